/*
 * bsf_oracle.c -- CPU restatement of the reference's multi-D B-spline PSF
 * Poisson-MLE Levenberg-Marquardt fitter (fittypes 6, 7, 8).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU fitter is compared
 * against.  The shipped library never links, loads or calls it; only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
 *
 * One source, two builds:
 *   default    "parity" semantics: float state with the double-precision
 *              sub-expressions the reference's C++ has under g++/glibc
 *              (derivative sums, pow(model,2), 1.0/U in the Cholesky, the
 *              0.8 lambda decay, double-literal bounds).  Symbols: orc_*.
 *   -DORC_F64  the same algorithm with every quantity in double: the "FP64
 *              oracle" used to state the error of fast paths.  Symbols:
 *              orc_*_f64; real-valued outputs are double.
 *
 * Reference files this follows (paths relative to the reference root,
 * VC = "MultiD_bspline/visual code"):
 *   VC/CPUmleFit_LM/CPUsplineLib.cpp:1289-1326  B-spline basis      (bsp)
 *   VC/CPUmleFit_LM/CPUsplineLib.cpp:1129-1197, 972-1126, 477-758
 *                                               weight tables       (delta_tables)
 *   VC/CPUmleFit_LM/CPUsplineLib.cpp:1199-1287, 761-970, 43-475
 *                                               gather-contract     (deriv_eval)
 *   VC/CPUmleFit_LM/CPUsplineLib.cpp:1363-1388  Cholesky            (cholesky)
 *   VC/CPUmleFit_LM/CPUsplineLib.cpp:1390-1416, 15-41
 *                                               triangular solves   (lu_eval)
 *   VC/CPUmleFit_LM/MatInvLib.h:17-77           Fisher inverse      (matinv)
 *   VC/CPUmleFit_LM/CPUgaussLib.cpp:172-223     initial estimates   (com2d, maxmin2d)
 *   VC/CPUmleFit_LM/CPUmleFit_LM.cpp:2411-2760  fittype 6 fit       (fit_one, ft==6)
 *   VC/CPUmleFit_LM/CPUmleFit_LM.cpp:2763-3158  fittype 7 fit       (fit_one, ft==7)
 *   VC/CPUmleFit_LM/CPUmleFit_LM.cpp:3163-3620  fittype 8 fit       (fit_one, ft==8)
 *   VC/CPUmleFit_LM/mexFunction.cpp:392-408     per-ROI dispatch    (orc_fit)
 *
 * Deliberate, documented departures (each is undefined behaviour in the
 * reference, so no defined result is changed):
 *   - lu_eval uses an 8-entry scratch vector for n=8 (the reference's 5D
 *     solve declares 7 entries, CPUsplineLib.cpp:17).
 *   - a derivative tap whose forward-difference partner lies past the end of
 *     the coefficient array reads 0 (its weight is exactly 0; the reference
 *     reads past the array).
 *   - fittype 8 with initP == NULL starts g2 at 0.9 (the reference reads an
 *     uninitialised value; 0.9 is its local g2 initialiser, CPUmleFit_LM.cpp:3203).
 *   - fittype 7 with initP == NULL uses the defaults nz/2, nw/2 (the
 *     reference dereferences the pointer).
 *   - the 5D z-derivative's irregular term (see deriv_eval) reads past the
 *     end of the reference's z table for blocks 24..31; the oracle reads the
 *     next table in declaration order (the w table).
 *
 * Build (oracle/Makefile): gcc -O2 -ffp-contract=off, no -march, so float
 * arithmetic is plain SSE2 with no contraction and no excess precision.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#ifdef ORC_F64
typedef double real;
#define R(x) (x)
#define SYM(n) n##_f64
#define LOGR(x) log(x)
#define FABSR(x) fabs(x)
#define FLOORR(x) floor(x)
#define SQRTR(x) sqrt(x)
#else
typedef float real;
#define R(x) (x##f)
#define SYM(n) n
#define LOGR(x) logf(x)
#define FABSR(x) fabsf(x)
#define FLOORR(x) floorf(x)
#define SQRTR(x) sqrtf(x)
#endif

#define PI_C R(3.141592)       /* definitions.h:15 */
#define TOL R(1e-6)            /* definitions.h:27 */
#define TOL1 R(1e-5)           /* definitions.h:28 */

/* Planted mistakes, for the tests that show the whole-fit pins catch a wrong
 * LM control (0 = none, the restatement as written).  Set only by tests,
 * never while another fit runs:
 *   1 limiter sign flipped (theta += step)     2 clamp fade dropped
 *   3 damping never raised on a rejected step  4 step limiter dropped
 *   5 lambda decay on accepted steps dropped   6 Hessian weight d/mu instead of d/mu^2
 *   7 x-derivative sign flipped in the model's dudt
 *   8 (not a mistake) the 5D contraction's two irregular derivative terms
 *     replaced by the regular ones, so that the fittype-8 LM control can be
 *     checked as an optimiser apart from the reference's derivative defects */
int SYM(orc_plant) = 0;

/* ------------------------------------------------------------------ A6 */
/* Centred cubic (deg 3) and quadratic (deg 2) B-spline basis,
 * CPUsplineLib.cpp:1289-1326.  The left quadratic branch has a double
 * literal (1.5*xi), kept here. */
real SYM(orc_bspline)(real xi, int deg)
{
    real x2, x3, b = 0;
    if (deg == 2) {
        x2 = xi * xi;
        if (xi >= R(0.5) && xi < R(1.5))
            b = R(1.125) - R(1.5) * xi + R(0.5) * x2;
        else if (xi >= R(-0.5) && xi < R(0.5))
            b = R(0.75) - x2;
        else if (xi >= R(-1.5) && xi < R(-0.5))
            b = (real)((double)R(1.125) + 1.5 * (double)xi + (double)(R(0.5) * x2));
    } else if (deg == 3) {
        const real c43 = R(4.0) / R(3.0), c23 = R(2.0) / R(3.0), c16 = R(1.0) / R(6.0);
        x2 = xi * xi;
        x3 = x2 * xi;
        if (xi >= R(1.0) && xi < R(2.0))
            b = c43 - R(2.0) * xi + x2 - c16 * x3;
        else if (xi >= R(0.0) && xi < R(1.0))
            b = c23 - x2 + R(0.5) * x3;
        else if (xi >= R(-1.0) && xi < R(0.0))
            b = c23 - x2 - R(0.5) * x3;
        else if (xi >= R(-2.0) && xi < R(-1.0))
            b = c43 + R(2.0) * xi + x2 + c16 * x3;
    }
    return b;
}

/* ------------------------------------------------------------------ A7 */
/* Tensor-product weight tables.
 *   frac[] in the reference's argument order: x (column), y (row), z, w, u.
 *   Term axes ("bits") run y, x, z, w, u (bit 0 = y = coefficient dim 0).
 *   tab[k*T + 2^D*blk + t], T = 4^D, k = 0: value, k = 1+a: derivative along
 *   term axis a (the reference's delta_dxf is a = 0, i.e. the row axis).
 *   blk's bit a holds (idx_a - 1); t's bit a picks tap k-idx (0) or
 *   k+idx-1 (1), with weights beta3(d+idx-1) / beta3(d-idx) and derivative
 *   weights beta2(d-1+idx-1/2) / beta2(d-1-idx+1/2).
 * Products run left to right in axis order y, x, z, w, u, as written in
 * CPUsplineLib.cpp:1129-1197 (3D), 972-1126 (4D), 477-758 (5D). */
void SYM(orc_delta_tables)(int D, const real *frac, real *tab)
{
    real B1[5][3], B2[5][3], dB1[5][3], dB2[5][3];
    const int T = 1 << (2 * D), nb = 1 << D;
    for (int a = 0; a < D; a++) {
        real d = (a == 0) ? frac[1] : (a == 1) ? frac[0] : frac[a];
        real dd = d - R(1.0);
        for (int idx = 1; idx <= 2; idx++) {
            B1[a][idx] = SYM(orc_bspline)(d + (real)idx - R(1.0), 3);
            B2[a][idx] = SYM(orc_bspline)(d - (real)idx, 3);
            dB1[a][idx] = SYM(orc_bspline)(dd + (real)idx - R(1.0) / R(2.0), 2);
            dB2[a][idx] = SYM(orc_bspline)(dd - (real)idx + R(1.0) / R(2.0), 2);
        }
    }
    for (int k = 0; k <= D; k++)
        for (int blk = 0; blk < nb; blk++)
            for (int t = 0; t < nb; t++) {
                real w = 0;
                for (int a = 0; a < D; a++) {
                    int idx = ((blk >> a) & 1) + 1, bit = (t >> a) & 1;
                    real fa;
                    if (k == a + 1)
                        fa = bit ? dB2[a][idx] : dB1[a][idx];
                    else
                        fa = bit ? B2[a][idx] : B1[a][idx];
                    w = (a == 0) ? fa : w * fa;
                }
                tab[k * T + nb * blk + t] = w;
            }
}

/* ------------------------------------------------------------------ A8 */
/* Per-pixel gather-contract: model value and derivatives.
 *   coord[]: x (column), y (row), z, w, u spline coordinates (1-based).
 *   nd[]:    nDatax, nDatay, nDataz, nDataw, nDatau (coefficient dims - 3).
 * Coefficient memory: dim 0 (extent nDatax+3) is indexed by the row tap,
 * dim 1 by the column tap, then z, w, u; column-major.
 * The value sum is float per block; every derivative block is summed in
 * double from (-1.0*c[tap] + c[tap+e_a]) * weight and added into a float
 * accumulator (CPUsplineLib.cpp:1199-1287, 761-970, 43-475).
 * dudt = [-N*d_col, -N*d_row, f, 1, N*d_z, N*d_w, N*d_u], model = bg + N*f. */
void SYM(orc_deriv_eval)(int D, const real *coord, const int *nd, const real *tab,
                         const float *c, const real *theta, real *dudt, real *model)
{
    const int T = 1 << (2 * D), nb = 1 << D;
    int k[5];
    long stride[5];
    long total;
    /* term axis a -> coordinate / clamp extent */
    for (int a = 0; a < D; a++) {
        real ci = (a == 0) ? coord[1] : (a == 1) ? coord[0] : coord[a];
        int n = (a == 0) ? nd[1] : (a == 1) ? nd[0] : nd[a];
        int kc = (int)FLOORR(ci);
        if (kc < 1) kc = 1;
        if (kc > n - 1) kc = n - 1;
        if (ci == (real)n) kc = n;
        k[a] = kc + 1;
    }
    stride[0] = 1;
    stride[1] = nd[0] + 3;
    stride[2] = stride[1] * (nd[1] + 3);
    for (int a = 3; a < 5; a++) stride[a] = stride[a - 1] * (nd[a - 1] + 3);
    total = stride[D - 1] * (nd[D - 1] + 3);

    real f = 0, df[5] = {0, 0, 0, 0, 0};
    for (int blk = 0; blk < nb; blk++) {
        long addr[32];
        for (int t = 0; t < nb; t++) {
            long ad = 0;
            for (int a = 0; a < D; a++) {
                int idx = ((blk >> a) & 1) + 1, bit = (t >> a) & 1;
                int tap = bit ? (k[a] + idx - 1) : (k[a] - idx);
                ad += (long)tap * stride[a];
            }
            addr[t] = ad;
        }
        real s = 0;
        for (int t = 0; t < nb; t++) {
            real p = c[addr[t]] * tab[nb * blk + t];
            s = (t == 0) ? p : s + p;
        }
        f += s;
        for (int a = 0; a < D; a++) {
            const real *wd = tab + (a + 1) * T + nb * blk;
            double sd = 0;
            for (int t = 0; t < nb; t++) {
                long a0 = addr[t];
                real w = wd[t];
                /* The reference's 5D contraction has two irregular terms, in
                 * both its CPU and GPU code:
                 *  z: term 27 takes weight [32*blk + 271] of the z table
                 *     (CPUsplineLib.cpp:313, GPUsplineLib.cuh:573); past the
                 *     table's end this reads the next table in declaration
                 *     order (the w table), which is how tab[] is laid out.
                 *  u: term 16 differences the u tap k-idx instead of
                 *     k+idx-1 (CPUsplineLib.cpp:420-421, GPUsplineLib.cuh:681). */
                if (D == 5 && a == 2 && t == 27 && SYM(orc_plant) != 8) w = wd[271];
                if (D == 5 && a == 4 && t == 16 && SYM(orc_plant) != 8) {
                    int idx = ((blk >> 4) & 1) + 1;
                    a0 -= (long)(2 * idx - 1) * stride[4];
                }
                long pa = a0 + stride[a];
                double cp = (pa < total) ? (double)c[pa] : 0.0;
                double term = (-1.0 * (double)c[a0] + cp) * (double)w;
                sd = (t == 0) ? term : sd + term;
            }
            df[a] = (real)((double)df[a] + sd);
        }
    }
    dudt[0] = R(-1.0) * theta[2] * df[1];
    dudt[1] = R(-1.0) * theta[2] * df[0];
    dudt[2] = f;
    dudt[3] = R(1.0);
    for (int a = 2; a < D; a++) dudt[2 + a] = theta[2] * df[a];
    *model = theta[3] + theta[2] * f;
}

/* ----------------------------------------------------------------- A10 */
/* Cholesky, CPUsplineLib.cpp:1363-1388: U row-major lower, L = U^T.
 * Fails (returns 1) on a negative pivot radicand; a zero pivot passes. */
int SYM(orc_cholesky)(const real *A, int n, real *L, real *U)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i + 1; j++) {
            real s = 0;
            for (int kk = 0; kk < j; kk++) s += U[i * n + kk] * U[j * n + kk];
            if (i == j) {
                if (A[i * n + i] - s >= 0) {
                    U[i * n + j] = SQRTR(A[i * n + i] - s);
                    L[j * n + i] = U[i * n + j];
                } else {
                    return 1;
                }
            } else {
                U[i * n + j] = (real)(1.0 / (double)U[j * n + j] * (double)(A[i * n + j] - s));
                L[j * n + i] = U[i * n + j];
            }
        }
    return 0;
}

/* Forward / backward substitution, CPUsplineLib.cpp:1390-1416 and 15-41. */
void SYM(orc_lu_eval)(const real *L, const real *U, const real *b, int n, real *x)
{
    real y[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        y[i] = b[i];
        for (int j = 0; j < i; j++) y[i] -= L[j * n + i] * y[j];
        y[i] /= L[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        x[i] = y[i];
        for (int j = i + 1; j < n; j++) x[i] -= U[j * n + i] * x[j];
        x[i] /= U[i * n + i];
    }
}

/* ----------------------------------------------------------------- A12 */
/* Unpivoted LU inverse; returns the diagonal.  MatInvLib.h:17-77.
 * M is overwritten with its LU factors, as in the reference. */
void SYM(orc_matinv)(real *M, real *Minv, real *Diag, int sz)
{
    int ii, jj, kk, num, b;
    real tmp1 = 0;
    real yy[25];
    for (jj = 0; jj < sz; jj++) {
        for (ii = 0; ii <= jj; ii++)
            if (ii > 0) {
                for (kk = 0; kk <= ii - 1; kk++) tmp1 += M[ii + kk * sz] * M[kk + jj * sz];
                M[ii + jj * sz] -= tmp1;
                tmp1 = 0;
            }
        for (ii = jj + 1; ii < sz; ii++)
            if (jj > 0) {
                for (kk = 0; kk <= jj - 1; kk++) tmp1 += M[ii + kk * sz] * M[kk + jj * sz];
                M[ii + jj * sz] = (1 / M[jj + jj * sz]) * (M[ii + jj * sz] - tmp1);
                tmp1 = 0;
            } else {
                M[ii + jj * sz] = (1 / M[jj + jj * sz]) * M[ii + jj * sz];
            }
    }
    tmp1 = 0;
    for (num = 0; num < sz; num++) {
        yy[0] = (num == 0) ? 1 : 0;
        for (ii = 1; ii < sz; ii++) {
            b = (ii == num) ? 1 : 0;
            for (jj = 0; jj <= ii - 1; jj++) tmp1 += M[ii + jj * sz] * yy[jj];
            yy[ii] = b - tmp1;
            tmp1 = 0;
        }
        Minv[sz - 1 + num * sz] = yy[sz - 1] / M[(sz - 1) + (sz - 1) * sz];
        for (ii = sz - 2; ii >= 0; ii--) {
            for (jj = ii + 1; jj < sz; jj++) tmp1 += M[ii + jj * sz] * Minv[jj + num * sz];
            Minv[ii + num * sz] = (1 / M[ii + ii * sz]) * (yy[ii] - tmp1);
            tmp1 = 0;
        }
    }
    if (Diag)
        for (ii = 0; ii < sz; ii++) Diag[ii] = Minv[ii * sz + ii];
}

/* ------------------------------------------------------------------ A5 */
/* Centre of mass (row centroid into *x), CPUgaussLib.cpp:172-192. */
void SYM(orc_com2d)(int sz, const float *data, real *x, real *y)
{
    real tx = 0, ty = 0, ts = 0;
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++) {
            real d = (real)data[sz * jj + ii];
            tx += d * (real)ii;
            ty += d * (real)jj;
            ts += d;
        }
    *x = tx / ts;
    *y = ty / ts;
}

/* Max pixel (floored at 0) and min pixel, CPUgaussLib.cpp:195-223. */
void SYM(orc_maxmin2d)(int sz, const float *data, real *maxN, real *minBG)
{
    real mx = R(0.0), mn = R(10e10);
    for (int kk = 0; kk < sz; kk++)
        for (int ll = 0; ll < sz; ll++) {
            real d = (real)data[kk * sz + ll];
            mx = (mx > d) ? mx : d;
            mn = (mn < d) ? mn : d;
        }
    *maxN = mx;
    *minBG = mn;
}

/* ------------------------------------------------------------- A1-A3, A9, A11 */
typedef struct {
    int ft, NV, D, sz;
    int nd[5];
    const float *c, *c3, *d;
    real off;
    real coord[5]; /* column, row, z, w, u ROI-level spline coordinates */
    real g2;
    real tab[6 * 1024];
    real tab3[4 * 64];
} fitctx;

/* ROI-level coordinates, fractional parts and weight tables for theta
 * (CPUmleFit_LM.cpp:2512-2521, 2875-2887, 3276-3292). */
static void set_coords(fitctx *x, const real *th)
{
    real frac[5];
    x->coord[0] = (real)((x->sz + 1) / 2) - th[0];
    x->coord[1] = (real)((x->sz + 1) / 2) - th[1];
    for (int a = 2; a < x->D; a++) x->coord[a] = th[2 + a];
    for (int a = 0; a < x->D; a++) frac[a] = x->coord[a] - FLOORR(x->coord[a]);
    SYM(orc_delta_tables)(x->D, frac, x->tab);
    if (x->ft == 8) {
        SYM(orc_delta_tables)(3, frac, x->tab3);
        x->g2 = th[7];
    }
}

/* Model value and dudt at pixel (row ii, column jj). */
static void pixel_model(const fitctx *x, const real *th, int ii, int jj, real *dudt, real *model)
{
    real cc[5];
    cc[0] = x->coord[0] + (real)jj + x->off;
    cc[1] = x->coord[1] + (real)ii + x->off;
    for (int a = 2; a < x->D; a++) cc[a] = x->coord[a];
    if (x->ft != 8) {
        SYM(orc_deriv_eval)(x->D, cc, x->nd, x->tab, x->c, th, dudt, model);
        if (SYM(orc_plant) == 7) dudt[0] = -dudt[0];
        return;
    }
    /* fittype 8: mu = (1-g2) mu3 + g2 mu5; only dudt[0..4] blended;
     * dudt[7] = mu5 - mu3 (CPUmleFit_LM.cpp:3312-3322) */
    real d3[5], m3, m5;
    real g2 = x->g2;
    SYM(orc_deriv_eval)(5, cc, x->nd, x->tab, x->c, th, dudt, &m5);
    SYM(orc_deriv_eval)(3, cc, x->nd, x->tab3, x->c3, th, d3, &m3);
    for (int l = 0; l < 5; l++) dudt[l] = dudt[l] * g2 + (R(1.0) - g2) * d3[l];
    *model = m3 * (R(1.0) - g2) + m5 * g2;
    dudt[7] = m5 - m3;
    if (SYM(orc_plant) == 7) dudt[0] = -dudt[0];
}

/* One pass over the ROI: Poisson deviance, gradient and expected-information
 * Hessian (CPUmleFit_LM.cpp:2524-2549 and twins).  first_ft8 selects the
 * float pow(model, 2.0f) of the fittype-8 first pass (CPUmleFit_LM.cpp:3338). */
static real eval_pass(const fitctx *x, const real *th, int first, real *J, real *H)
{
    const int NV = x->NV, sz = x->sz;
    real err = 0, dudt[8], model;
    memset(J, 0, sizeof(real) * NV);
    memset(H, 0, sizeof(real) * NV * NV);
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++) {
            pixel_model(x, th, ii, jj, dudt, &model);
            real data = (real)x->d[sz * jj + ii];
            if (x->ft == 7 && first) dudt[5] = dudt[4];
            if (data > 0)
                err = err + 2 * ((model - data) - data * LOGR(model / data));
            else {
                err = err + 2 * model;
                data = 0;
            }
            real t1 = 1 - data / model;
            for (int l = 0; l < NV; l++) J[l] += t1 * dudt[l];
            real t2;
            if (x->ft == 8 && first)
                t2 = data / (model * model);
            else
                t2 = (real)((double)data / ((double)model * (double)model));
            if (SYM(orc_plant) == 6) t2 = data / model;
            for (int l = 0; l < NV; l++)
                for (int m = l; m < NV; m++) {
                    H[l * NV + m] += t2 * dudt[l] * dudt[m];
                    H[m * NV + l] = H[l * NV + m];
                }
        }
    return err;
}

static real dmax(real a, double b) { return ((double)a > b) ? a : (real)b; }
static real dmin(real a, double b) { return ((double)a < b) ? a : (real)b; }

/* Whole fit of one ROI; writes column-major outputs with stride Nfits. */
static void fit_one(int ft, int n, const float *data, const float *coeff, const float *coeff3,
                    const int *nd, int sz, int iterations, const float *initP,
                    real *P, real *CRLB, real *LL, int Nfits, unsigned char *status)
{
    const int plant = SYM(orc_plant);
    static const float clamp6[5] = {2.0f, 2.0f, 10000, 20, 20};
    static const float clamp7[6] = {1.0f, 1.0f, 10000, 20, 2, 0.00001f};
    static const float fade7[6] = {0.5f, 0.5f, 0.4f, 0.5f, 0.7f, 0.8f};
    static const float clamp8[8] = {1.0f, 1.0f, 10000, 20, 2, 2, 2, 0.1f};
    fitctx *x = (fitctx *)malloc(sizeof(fitctx));
    const int NV = (ft == 6) ? 5 : (ft == 7) ? 6 : 8;
    real th[9], oth[9], sgn[8] = {0}, osgn[8] = {0}, clampv[8], oclamp[8], fade[8], upd[8] = {0};
    real J[8], H[64], L[64], U[64], M[64], Minv[64], Diag[8];
    real newErr, oldErr = (real)1e13, newLambda = 1, oldLambda = 1, sumUpdate = (real)1e10;
    int kk, turn = 1, info = 0, firstinfo = 0;

    x->ft = ft;
    x->NV = NV;
    x->D = (ft == 6) ? 3 : (ft == 7) ? 4 : 5;
    x->sz = sz;
    for (int a = 0; a < 5; a++) x->nd[a] = nd[a];
    x->c = coeff;
    x->c3 = coeff3;
    x->d = data + (long)sz * sz * n;
    x->off = (real)(nd[0] - sz) / R(2.0);
    x->g2 = R(0.9);

    for (int i = 0; i < NV; i++) {
        clampv[i] = (ft == 6) ? clamp6[i] : (ft == 7) ? clamp7[i] : clamp8[i];
        fade[i] = (ft == 7) ? fade7[i] : R(0.5);
        oclamp[i] = clampv[i];
    }
    for (int i = 0; i < 64; i++) { M[i] = 0; Minv[i] = 0; L[i] = 0; U[i] = 0; }

    /* initial values (CPUmleFit_LM.cpp:2478-2482, 2831-2845, 3228-3258) */
    {
        real Nmax;
        SYM(orc_com2d)(sz, x->d, &th[0], &th[1]);
        SYM(orc_maxmin2d)(sz, x->d, &Nmax, &th[3]);
        double t = (double)(((Nmax - th[3]) * 2) * PI_C) * 1.5 * 1.5;
        th[2] = (real)((0.0 > t) ? 0.0 : t);
        th[3] = dmax(th[3], 0.01);
        if (ft == 6) {
            th[4] = (real)nd[2] / R(2.0);
        } else if (ft == 7) {
            if (initP && initP[0] != 0) {
                th[4] = initP[0];
                th[5] = initP[1];
            } else {
                th[4] = (real)nd[2] / R(2.0);
                th[5] = (real)nd[3] / R(2.0);
            }
        } else {
            if (initP) {
                for (int i = 4; i < 8; i++) th[i] = initP[i - 4];
            } else {
                th[4] = (real)nd[2] / R(2.0);
                th[5] = (real)nd[3] / R(2.0);
                th[6] = (real)nd[4] / R(2.0);
                th[7] = R(0.9);
            }
        }
    }
    for (int i = 0; i < NV; i++) oth[i] = th[i];

    set_coords(x, th);
    newErr = eval_pass(x, th, 1, J, H);

    for (kk = 0; kk < iterations; kk++) {
        if (FABSR((newErr - oldErr) / newErr) < TOL || (ft == 8 && sumUpdate < TOL1))
            break;
        if ((double)newErr > 1.5 * (double)oldErr) {
            for (int i = 0; i < NV; i++) {
                sgn[i] = osgn[i];
                clampv[i] = oclamp[i];
                th[i] = oth[i];
            }
            newLambda = oldLambda;
            newErr = oldErr;
            if (plant != 3) newLambda = 10 * newLambda;
        } else if (newErr < oldErr && plant != 5) {
            if (newLambda > 1)
                newLambda = (real)((double)newLambda * 0.8);
            else if (newLambda < 1)
                newLambda = 1;
        }
        for (int i = 0; i < NV; i++) H[i * NV + i] = H[i * NV + i] * newLambda;
        info = SYM(orc_cholesky)(H, NV, L, U);
        if (info == 0) {
            SYM(orc_lu_eval)(L, U, J, NV, upd);
            if (ft != 6)
                for (int l = 0; l < NV; l++)
                    if (isnan(upd[l])) info = 1;
        }
        if (kk == 0) firstinfo = info;
        if (info != 0) {
            newLambda = 10 * newLambda;
            continue;
        }
        if (ft == 7) upd[5] = upd[5] * 10;
        for (int i = 0; i < NV; i++) {
            osgn[i] = sgn[i];
            oclamp[i] = clampv[i];
            oth[i] = th[i];
        }
        oldLambda = newLambda;
        oldErr = newErr;
        if (ft == 7 && (double)FABSR(upd[4]) < 0.005 && turn) {
            clampv[5] = (real)nd[3] / 10;
            oclamp[5] = (real)nd[3] / 10;
            turn = 0;
            clampv[4] = 2;
            oclamp[4] = 2;
        }
        if (ft == 8) sumUpdate = 0;
        for (int l = 0; l < NV; l++) {
            if (sgn[l] != 0 && plant != 2) {
                if ((sgn[l] == 1 && upd[l] < 0) || (sgn[l] == -1 && upd[l] > 0))
                    clampv[l] = (ft == 6) ? (real)((double)clampv[l] * 0.5) : clampv[l] * fade[l];
            }
            sgn[l] = (upd[l] > 0) ? 1 : -1;
            if (plant == 1 || plant == 4) {
                real st = (plant == 4) ? upd[l] : upd[l] / (1 + FABSR(upd[l] / clampv[l]));
                th[l] = (plant == 1) ? th[l] + st : th[l] - st;
                if (ft == 8) sumUpdate += FABSR(st);
            } else if (ft == 8) {
                upd[l] = upd[l] / (1 + FABSR(upd[l] / clampv[l]));
                th[l] = th[l] - upd[l];
                sumUpdate += FABSR(upd[l]);
            } else {
                th[l] = th[l] - upd[l] / (1 + FABSR(upd[l] / clampv[l]));
            }
        }
        /* bounds (CPUmleFit_LM.cpp:2631-2634, 3021-3026, 3449-3472) */
        if (ft == 8) {
            double lo = (double)(((real)sz - 1) / 2) - sz / 4.0;
            double hi = (double)(((real)sz - 1) / 2) + sz / 4.0;
            th[0] = dmax(th[0], lo);
            th[0] = dmin(th[0], hi);
            th[1] = dmax(th[1], lo);
            th[1] = dmin(th[1], hi);
        }
        th[2] = dmax(th[2], 1.0);
        th[3] = dmax(th[3], 0.01);
        th[4] = dmax(th[4], 0.0);
        if (ft == 6) {
            real zb = (real)(nd[2] - 4);
            th[4] = (th[4] < zb) ? th[4] : zb;
        } else {
            th[4] = dmin(th[4], (double)(real)nd[2] - 0.5);
        }
        if (ft == 7) {
            th[5] = dmax(th[5], 0.0);
            th[5] = dmin(th[5], (double)(real)nd[3] - 0.5);
        }
        if (ft == 8) {
            if (th[5] <= 1)
                th[5] = (real)(nd[3] - 2) + th[5];
            else if (th[5] > (real)(nd[3] - 1))
                th[5] = (2 + th[5]) - (real)nd[3];
            th[6] = (th[6] > 1) ? th[6] : 1;
            th[6] = dmin(th[6], (double)(real)nd[4] - 0.5);
            th[7] = (th[7] < 1) ? th[7] : 1;
            th[7] = (th[7] > 0) ? th[7] : 0;
        }
        set_coords(x, th);
        newErr = eval_pass(x, th, 0, J, H);
    }

    /* Fisher information, CRLB and log-likelihood (CPUmleFit_LM.cpp:2714-2754,
     * 3110-3150, 3566-3611).  Fittype 8 reuses the last pass's tables. */
    if (ft != 8) set_coords(x, th);
    real Div = 0;
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++) {
            real dudt[8], model;
            pixel_model(x, th, ii, jj, dudt, &model);
            real data = (real)x->d[sz * jj + ii];
            for (int a = 0; a < NV; a++)
                for (int b = a; b < NV; b++) {
                    M[a * NV + b] += dudt[b] * dudt[a] / model;
                    M[b * NV + a] = M[a * NV + b];
                }
            if (model > 0) {
                if (data > 0)
                    Div += data * LOGR(model) - model - data * LOGR(data) + data;
                else
                    Div += -model;
            }
        }
    SYM(orc_matinv)(M, Minv, Diag, NV);
    for (int a = 0; a < NV; a++) P[(long)Nfits * a + n] = th[a];
    P[(long)Nfits * NV + n] = (real)kk;
    for (int a = 0; a < NV; a++) CRLB[(long)Nfits * a + n] = Diag[a];
    LL[n] = Div;
    if (status) {
        /* the library's per-ROI flags (1 iteration limit, 2 last solve failed,
         * 4 a CRLB entry not finite positive, 8 a non-finite output, 16 the
         * first damped solve failed) */
        unsigned f = 0;
        if (kk >= iterations) f |= 1u;
        if (info != 0) f |= 2u;
        for (int a = 0; a < NV; a++) {
            if (!(Diag[a] > 0 && isfinite(Diag[a]))) f |= 4u;
            if (!isfinite(th[a])) f |= 8u;
        }
        if (!isfinite(Div)) f |= 8u;
        if (iterations > 0 && firstinfo != 0) f |= 16u;
        status[n] = (unsigned char)f;
    }
    free(x);
}

/* ---------------------------------------------------------------- A4 */
/* 4Pi multi-channel fit.  The reference calls GPUmleFit_LM_4Pi_bspline
 * (demo_4pi/bspline_based_4piloc_v4.m:196-206) but ships no source for it, so
 * this restatement is specified from that call site, the channel model of
 * demo_4pi/store/simBSplinePSF.m:38-53 (channel k evaluates the 4D (x,y,z,phi)
 * spline at w = phi/dphi + 1 shifted by the channel's phase offset) and the
 * shared-parameter layout of demo_4pi/store/calculate_CRLB_YL_shared.m:11-13,
 * 98-124 (each of the 6 base parameters x, y, N, bg, z, phi is one slot when
 * shared, nch consecutive slots otherwise; NV = 6 nch - 3 sum(shared) for
 * nch = 4).  The LM control is fittype 7's (the reference's 4D fitter,
 * CPUmleFit_LM.cpp:2763-3158), slot by slot, except that phase slots wrap with
 * period nw-2 (fittype 8's azimuth rule, CPUmleFit_LM.cpp:3460-3467) instead of
 * being clamped.  With one channel, zero shift, all parameters shared and no
 * wrap this is exactly the fittype-7 fit (tests pin that reduction).
 *   data     [sz, sz, Nfits, nch] column-major (channel slowest)
 *   shared   6 flags;  dT [6, nch, Nfits] per-channel x/y offsets (or NULL)
 *   initZ, initPhi  [Nfits] per-ROI starts (NULL: nz/2, nw/2)
 *   phaseShift [nch] in w units. */
static int slot_map(const int *shared, int nch, int slot[6][8])
{
    int n = 0;
    for (int p = 0; p < 6; p++) {
        for (int k = 0; k < nch; k++) slot[p][k] = shared[p] ? n : n + k;
        n += shared[p] ? 1 : nch;
    }
    return n;
}

int SYM(orc_nv_4pi)(const int *shared, int nch)
{
    int slot[6][8];
    if (nch < 1 || nch > 8) return -1;
    return slot_map(shared, nch, slot);
}

static real wrap_phase(real w, int nw)
{
    if (w <= 1)
        w = (real)(nw - 2) + w;
    else if (w > (real)(nw - 1))
        w = (2 + w) - (real)nw;
    return w;
}

typedef struct {
    fitctx ch[8];
    int nch, NV, slot[6][8];
    const float *dT;
    const float *shift;
    int n, Nfits;
} ctx4pi;

static void set_coords_4pi(ctx4pi *c, const real *th, real thk[8][6])
{
    for (int k = 0; k < c->nch; k++) {
        for (int p = 0; p < 6; p++) thk[k][p] = th[c->slot[p][k]];
        if (c->dT) {
            thk[k][0] += (real)c->dT[((long)c->n * c->nch + k) * 6 + 0];
            thk[k][1] += (real)c->dT[((long)c->n * c->nch + k) * 6 + 1];
        }
        thk[k][5] = wrap_phase(thk[k][5] + (real)c->shift[k], c->ch[k].nd[3]);
        set_coords(&c->ch[k], thk[k]);
    }
}

static real eval_pass_4pi(ctx4pi *c, const real *th, int first, real *J, real *H)
{
    const int NV = c->NV, sz = c->ch[0].sz;
    real thk[8][6], err = 0, dudt[8], model, dall[24];
    set_coords_4pi(c, th, thk);
    memset(J, 0, sizeof(real) * NV);
    memset(H, 0, sizeof(real) * NV * NV);
    for (int ii = 0; ii < sz; ii++)
        for (int jj = 0; jj < sz; jj++)
            for (int k = 0; k < c->nch; k++) {
                pixel_model(&c->ch[k], thk[k], ii, jj, dudt, &model);
                if (first) dudt[5] = dudt[4];
                for (int l = 0; l < NV; l++) dall[l] = 0;
                for (int p = 0; p < 6; p++) dall[c->slot[p][k]] = dudt[p];
                real data = (real)c->ch[k].d[sz * jj + ii];
                if (data > 0)
                    err = err + 2 * ((model - data) - data * LOGR(model / data));
                else {
                    err = err + 2 * model;
                    data = 0;
                }
                real t1 = 1 - data / model;
                for (int l = 0; l < NV; l++) J[l] += t1 * dall[l];
                real t2 = (real)((double)data / ((double)model * (double)model));
                for (int l = 0; l < NV; l++)
                    for (int m = l; m < NV; m++) {
                        H[l * NV + m] += t2 * dall[l] * dall[m];
                        H[m * NV + l] = H[l * NV + m];
                    }
            }
    return err;
}

static void fit_one_4pi(int n, const float *data, int sz, int Nfits, int nch, const int *shared, int iterations,
                        const float *coeff, const int *nd, const float *dT, const float *initZ,
                        const float *initPhi, const float *shift, real *P, real *CRLB, real *LL)
{
    static const float clamp7[6] = {1.0f, 1.0f, 10000, 20, 2, 0.00001f};
    static const float fade7[6] = {0.5f, 0.5f, 0.4f, 0.5f, 0.7f, 0.8f};
    ctx4pi *c = (ctx4pi *)malloc(sizeof(ctx4pi));
    real th[24], oth[24], sgn[24] = {0}, osgn[24] = {0}, clampv[24], oclamp[24], fade[24], upd[24] = {0};
    real J[24], H[576], L[576], U[576], M[576], Minv[576], Diag[24];
    real newErr, oldErr = (real)1e13, newLambda = 1, oldLambda = 1;
    int kk, turn = 1, base_of[24], info = 0;
    const int plant = SYM(orc_plant);
    c->nch = nch;
    c->NV = slot_map(shared, nch, c->slot);
    c->dT = dT;
    c->shift = shift;
    c->n = n;
    c->Nfits = Nfits;
    const int NV = c->NV;
    for (int k = 0; k < nch; k++) {
        fitctx *x = &c->ch[k];
        x->ft = 7;
        x->NV = 6;
        x->D = 4;
        x->sz = sz;
        for (int a = 0; a < 5; a++) x->nd[a] = nd[a];
        x->c = coeff;
        x->c3 = 0;
        x->d = data + ((long)k * Nfits + n) * sz * sz;
        x->off = (real)(nd[0] - sz) / R(2.0);
        x->g2 = 0;
    }
    for (int p = 0; p < 6; p++)
        for (int k = 0; k < nch; k++) base_of[c->slot[p][k]] = p;
    /* initial values: fittype 7's per channel; a shared slot takes the
     * channel mean (CPUmleFit_LM.cpp:2831-2845) */
    {
        real v[8][6];
        for (int k = 0; k < nch; k++) {
            real Nmax;
            SYM(orc_com2d)(sz, c->ch[k].d, &v[k][0], &v[k][1]);
            SYM(orc_maxmin2d)(sz, c->ch[k].d, &Nmax, &v[k][3]);
            double t = (double)(((Nmax - v[k][3]) * 2) * PI_C) * 1.5 * 1.5;
            v[k][2] = (real)((0.0 > t) ? 0.0 : t);
            v[k][3] = dmax(v[k][3], 0.01);
            v[k][4] = initZ ? (real)initZ[n] : (real)nd[2] / R(2.0);
            v[k][5] = initPhi ? (real)initPhi[n] : (real)nd[3] / R(2.0);
        }
        for (int p = 0; p < 6; p++) {
            if (shared[p]) {
                real s = 0;
                for (int k = 0; k < nch; k++) s += v[k][p];
                th[c->slot[p][0]] = s / (real)nch;
            } else {
                for (int k = 0; k < nch; k++) th[c->slot[p][k]] = v[k][p];
            }
        }
    }
    for (int i = 0; i < NV; i++) {
        clampv[i] = clamp7[base_of[i]];
        oclamp[i] = clampv[i];
        fade[i] = fade7[base_of[i]];
        oth[i] = th[i];
    }
    for (int i = 0; i < NV * NV; i++) { M[i] = 0; Minv[i] = 0; L[i] = 0; U[i] = 0; }

    newErr = eval_pass_4pi(c, th, 1, J, H);
    for (kk = 0; kk < iterations; kk++) {
        if (FABSR((newErr - oldErr) / newErr) < TOL) break;
        if ((double)newErr > 1.5 * (double)oldErr) {
            for (int i = 0; i < NV; i++) {
                sgn[i] = osgn[i];
                clampv[i] = oclamp[i];
                th[i] = oth[i];
            }
            newLambda = oldLambda;
            newErr = oldErr;
            if (plant != 3) newLambda = 10 * newLambda;
        } else if (newErr < oldErr && plant != 5) {
            if (newLambda > 1)
                newLambda = (real)((double)newLambda * 0.8);
            else if (newLambda < 1)
                newLambda = 1;
        }
        for (int i = 0; i < NV; i++) H[i * NV + i] = H[i * NV + i] * newLambda;
        info = SYM(orc_cholesky)(H, NV, L, U);
        if (info == 0) {
            SYM(orc_lu_eval)(L, U, J, NV, upd);
            for (int l = 0; l < NV; l++)
                if (isnan(upd[l])) info = 1;
        }
        if (info != 0) {
            newLambda = 10 * newLambda;
            continue;
        }
        for (int l = 0; l < NV; l++)
            if (base_of[l] == 5) upd[l] = upd[l] * 10;
        for (int i = 0; i < NV; i++) {
            osgn[i] = sgn[i];
            oclamp[i] = clampv[i];
            oth[i] = th[i];
        }
        oldLambda = newLambda;
        oldErr = newErr;
        if ((double)FABSR(upd[c->slot[4][0]]) < 0.005 && turn) {
            for (int l = 0; l < NV; l++) {
                if (base_of[l] == 5) {
                    clampv[l] = (real)nd[3] / 10;
                    oclamp[l] = (real)nd[3] / 10;
                }
                if (base_of[l] == 4) {
                    clampv[l] = 2;
                    oclamp[l] = 2;
                }
            }
            turn = 0;
        }
        for (int l = 0; l < NV; l++) {
            if (sgn[l] != 0 && plant != 2) {
                if ((sgn[l] == 1 && upd[l] < 0) || (sgn[l] == -1 && upd[l] > 0)) clampv[l] = clampv[l] * fade[l];
            }
            sgn[l] = (upd[l] > 0) ? 1 : -1;
            const real st = (plant == 4) ? upd[l] : upd[l] / (1 + FABSR(upd[l] / clampv[l]));
            th[l] = (plant == 1) ? th[l] + st : th[l] - st;
        }
        for (int l = 0; l < NV; l++) {
            switch (base_of[l]) {
            case 2: th[l] = dmax(th[l], 1.0); break;
            case 3: th[l] = dmax(th[l], 0.01); break;
            case 4:
                th[l] = dmax(th[l], 0.0);
                th[l] = dmin(th[l], (double)(real)nd[2] - 0.5);
                break;
            case 5: th[l] = wrap_phase(th[l], nd[3]); break;
            default: break;
            }
        }
        newErr = eval_pass_4pi(c, th, 0, J, H);
    }

    /* Fisher information, CRLB and log-likelihood over all channels */
    {
        real thk[8][6], dudt[8], model, dall[24], Div = 0;
        set_coords_4pi(c, th, thk);
        for (int ii = 0; ii < sz; ii++)
            for (int jj = 0; jj < sz; jj++)
                for (int k = 0; k < nch; k++) {
                    pixel_model(&c->ch[k], thk[k], ii, jj, dudt, &model);
                    for (int l = 0; l < NV; l++) dall[l] = 0;
                    for (int p = 0; p < 6; p++) dall[c->slot[p][k]] = dudt[p];
                    real data = (real)c->ch[k].d[sz * jj + ii];
                    for (int a = 0; a < NV; a++)
                        for (int b = a; b < NV; b++) {
                            M[a * NV + b] += dall[b] * dall[a] / model;
                            M[b * NV + a] = M[a * NV + b];
                        }
                    if (model > 0) {
                        if (data > 0)
                            Div += data * LOGR(model) - model - data * LOGR(data) + data;
                        else
                            Div += -model;
                    }
                }
        SYM(orc_matinv)(M, Minv, Diag, NV);
        for (int a = 0; a < NV; a++) P[(long)Nfits * a + n] = th[a];
        P[(long)Nfits * NV + n] = (real)kk;
        for (int a = 0; a < NV; a++) CRLB[(long)Nfits * a + n] = Diag[a];
        LL[n] = Div;
    }
    free(c);
}

int SYM(orc_fit_4pi)(const float *data, int sz, int Nfits, int nch, const int *shared, int iterations,
                     const float *coeff, const int *cdims, const float *dT, const float *initZ,
                     const float *initPhi, const float *phaseShift, real *P, real *CRLB, real *LL, int nthreads)
{
    int nd[5] = {cdims[0] - 3, cdims[1] - 3, cdims[2] - 3, cdims[3] - 3, 0};
    if (nch < 1 || nch > 8) return -1;
    if (SYM(orc_nv_4pi)(shared, nch) > 24) return -1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int n = 0; n < Nfits; n++)
        fit_one_4pi(n, data, sz, Nfits, nch, shared, iterations, coeff, nd, dT, initZ, initPhi, phaseShift, P,
                    CRLB, LL);
    return 0;
}

/* One deviance / gradient / Hessian pass (A9) at a given theta, for tests:
 * theta has NV entries; err, J[NV], H[NV*NV] and the model image
 * (mu[sz*jj+ii], column-major like the data) are returned. */
real SYM(orc_eval_pass)(int fittype, const float *roi, int sz, const float *coeff,
                        const int *cdims, const float *coeff3, const real *theta,
                        real *J, real *H, real *mu, int first)
{
    fitctx *x = (fitctx *)malloc(sizeof(fitctx));
    real th[9], err;
    x->ft = fittype;
    x->NV = (fittype == 6) ? 5 : (fittype == 7) ? 6 : 8;
    x->D = (fittype == 6) ? 3 : (fittype == 7) ? 4 : 5;
    x->sz = sz;
    for (int a = 0; a < 5; a++) x->nd[a] = (a < x->D) ? cdims[a] - 3 : 0;
    x->c = coeff;
    x->c3 = coeff3;
    x->d = roi;
    x->off = (real)(x->nd[0] - sz) / R(2.0);
    for (int i = 0; i < x->NV; i++) th[i] = theta[i];
    set_coords(x, th);
    err = eval_pass(x, th, first, J, H);
    if (mu)
        for (int ii = 0; ii < sz; ii++)
            for (int jj = 0; jj < sz; jj++) {
                real dudt[8];
                pixel_model(x, th, ii, jj, dudt, &mu[sz * jj + ii]);
            }
    free(x);
    return err;
}

/* Batch driver mirroring the MEX's per-ROI loop (mexFunction.cpp:392-408).
 *   cdims[]: coefficient array dims as stored (data grid + 3 per axis).
 *   nthreads > 1 fits ROIs in parallel (OpenMP); 1 = the reference's
 *   sequential loop.  Returns 0, or -1 for an unsupported fittype. */
int SYM(orc_fit_status)(int fittype, const float *data, int sz, int Nfits, const float *coeff,
                        const int *cdims, const float *coeff3, int iterations, const float *initP,
                        real *P, real *CRLB, real *LL, unsigned char *status, int nthreads)
{
    int nd[5] = {0, 0, 0, 0, 0};
    int D = (fittype == 6) ? 3 : (fittype == 7) ? 4 : (fittype == 8) ? 5 : 0;
    if (!D) return -1;
    for (int a = 0; a < D; a++) nd[a] = cdims[a] - 3;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int n = 0; n < Nfits; n++)
        fit_one(fittype, n, data, coeff, coeff3, nd, sz, iterations, initP, P, CRLB, LL, Nfits, status);
    return 0;
}

int SYM(orc_fit)(int fittype, const float *data, int sz, int Nfits, const float *coeff,
                 const int *cdims, const float *coeff3, int iterations, const float *initP,
                 real *P, real *CRLB, real *LL, int nthreads)
{
    int nd[5] = {0, 0, 0, 0, 0};
    int D = (fittype == 6) ? 3 : (fittype == 7) ? 4 : (fittype == 8) ? 5 : 0;
    if (!D) return -1;
    for (int a = 0; a < D; a++) nd[a] = cdims[a] - 3;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int n = 0; n < Nfits; n++)
        fit_one(fittype, n, data, coeff, coeff3, nd, sz, iterations, initP, P, CRLB, LL, Nfits, NULL);
    return 0;
}
