// bsf_session.cpp -- serving sessions (include/bsfit.h): device-resident
// models, a frame arena with HBM-budget admission and a pinned host mirror,
// and per-frame launches spread over kernel streams.
//
// Per frame:
//   host     inputs of every launch packed into the frame's pinned slot
//            (ROI descriptors, images, 4Pi starts / offsets);
//   s_in     one H2D copy of the slot's input section, event `in`;
//   ks[j]    the frame's launches (one per model / iterations / start, mixed
//            ROI sizes, largest first), each waiting on `in`;
//   s_out    waits for the launches, one D2H copy of the output section,
//            event `done`;
//   poll     scatters the outputs into the groups' host arrays.
// The arena is a ring: frames take slots in submission order and give them
// back in submission order once their results have been delivered, so the
// device and pinned slots of a frame are never reused while it runs.
//
// Graph mode (bsf_session_set_graphs): the arena is cut into equal slots
// taken in turn.  A frame's launches are matched to a plan
// (same models, iterations, starts and optional inputs, with ROI counts and
// sizes within the plan's capacities), whose inputs and outputs sit at fixed
// offsets of a slot; the plan's launches -- one graph branch each, reading
// their ROI counts from the slot's header -- are captured once per slot, so
// a frame is one H2D copy, one hipGraphLaunch and one D2H copy of the output
// section.  A frame that outgrows its plan gets a plan of twice its counts
// (and is captured anew); one that no slot can hold goes the per-call way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <vector>

#include "../../include/bsfit.h"
#include "bsf_kernels.h"

struct bsf_session;

namespace {

using bsf::fail;

constexpr size_t kAlign = 256;
constexpr int kMaxLaunchRois = 1 << 24;   // one grid of one-wave workgroups

size_t up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

int nv_of(int ft) { return ft == 6 ? 5 : ft == 7 ? 6 : ft == 8 ? 8 : -1; }
int dim_of(int ft) { return ft == 6 ? 3 : ft == 7 ? 4 : ft == 8 ? 5 : ft == BSF_FT_4PI ? 4 : -1; }

struct Model {
    bool live = false;
    int ft = 0, nv = 0, ndim = 0;
    int cdims[5] = {1, 1, 1, 1, 1};
    float *d_coeff = nullptr, *d_coeff3 = nullptr;
    size_t bytes = 0;
    int nch = 1;
    int shared[6] = {0, 0, 0, 0, 0, 0};
    float shift[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// One kernel launch of a frame: the groups sharing model, iterations and start.
constexpr int kMaxKernelStreams = 16;
constexpr int kMaxGroupStarts = 8;

struct Launch {
    int model = 0, iterations = 0, n = 0, Smax = 0, Smin = 0, nch = 1, nv = 0;
    float initP[kMaxGroupStarts][4] = {};   // start rows (multi-start: one launch per row)
    int nstarts = 1, strict = 0;
    int has_initP = 0, has_dT = 0, has_iz = 0, has_ip = 0;
    int wpc = 0;               // 4Pi: one wave per channel
    std::vector<int> groups;   // launch order (largest ROI size first)
    std::vector<int> start;    // first launch row of each group
    size_t o_desc = 0, o_data = 0, o_dT = 0, o_iz = 0, o_ip = 0;   // input section
    size_t o_end = 0;                                             // its end
    size_t o_P = 0, o_CR = 0, o_LL = 0;                           // output section
    double cost = 0;
    int stream = 0;
};

struct GroupOut {
    float *P, *CRLB, *LL;
    int count, nv, start, launch_n;
    size_t o_P, o_CR, o_LL;   // slot-relative byte offsets of the launch outputs
};

struct Frame {
    unsigned long long ticket = 0;
    size_t off = 0, bytes = 0;
    std::vector<GroupOut> outs;
    std::vector<std::pair<int, double>> work;   // (kernel stream, cost) of each launch
    std::vector<hipEvent_t> events;   // back() is `done`
    bool delivered = false;
};

// Graph mode: the arena splits into equal slots (at most kMaxGraphSlots),
// and a frame whose launches match a plan (same models, iterations, starts
// and optional inputs; counts and ROI sizes within the plan's capacities)
// replays that plan's graph, captured once per slot: the launches with
// device-side ROI counts, at fixed slot offsets.
constexpr int kMaxGraphSlots = 64;

struct Plan {
    std::vector<Launch> L;   // the keys; n = capacity (ROIs), Smax = largest ROI size; slot offsets
    int f64 = 0;
    size_t o_hdr = 0;        // per-launch ROI counts (int32), read by the kernels
    size_t o_data = 0;       // ROI data, packed launch after launch
    size_t data_cap = 0;     // bytes available to it
    size_t o_out = 0;        // output section: [o_out, slot end)
    hipGraphExec_t exec[kMaxGraphSlots] = {};
};

struct DeviceGuard {
    int old = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&old) != hipSuccess) old = -1;
        if (old != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (old >= 0) (void)hipSetDevice(old);
    }
};

}  // namespace

struct bsf_session {
    int device = 0, precision = BSF_PREC_MIX, nks = 2;
    double load[kMaxKernelStreams] = {};           // queued, undelivered launch cost per kernel stream
    size_t budget = 0, arena = 0, model_bytes = 0, lds_limit = 0;
    char *d_arena = nullptr, *h_arena = nullptr;
    hipStream_t s_in = nullptr, s_out = nullptr;
    hipStream_t ks[kMaxKernelStreams] = {};
    std::vector<Model> models;
    std::deque<Frame> inflight;
    std::vector<hipEvent_t> event_pool;
    unsigned long long next_ticket = 1;
    std::mutex mu;
    // per-submit scratch, kept to avoid reallocation on the steady-state path
    std::vector<Launch> launches;
    std::vector<int> order;
    int hetero = 0;          // bsf_session_set_hetero
    // graph mode (bsf_session_set_graphs)
    int graphs = 0;          // slots (0: per-call launches)
    size_t slot_bytes = 0;
    std::vector<Plan *> plans;
    std::vector<hipGraphExec_t> old_execs;   // of replaced plans, destroyed once nothing is in flight
    hipStream_t s_cap[kMaxKernelStreams + 1] = {};   // capture only: [0] origin, [1..] one per launch
    unsigned long long n_capture = 0, n_replay = 0, n_eager = 0;
};

namespace {

hipEvent_t take_event(bsf_session *s)
{
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Copies a finished frame's outputs from the pinned slot into the groups'
// host arrays (column-major, stride = group count).
void deliver(bsf_session *s, Frame &f)
{
    const char *base = s->h_arena + f.off;
    for (const GroupOut &g : f.outs) {
        const size_t cnt = (size_t)g.count, ln = (size_t)g.launch_n, st = (size_t)g.start;
        const float *P = (const float *)(base + g.o_P), *CR = (const float *)(base + g.o_CR);
        const float *LL = (const float *)(base + g.o_LL);
        for (int c = 0; c <= g.nv; c++) std::memcpy(g.P + c * cnt, P + c * ln + st, cnt * 4);
        for (int c = 0; c < g.nv; c++) std::memcpy(g.CRLB + c * cnt, CR + c * ln + st, cnt * 4);
        std::memcpy(g.LL, LL + st, cnt * 4);
    }
    for (hipEvent_t e : f.events) s->event_pool.push_back(e);
    f.events.clear();
    for (const auto &w : f.work) s->load[w.first] -= w.second;
    f.work.clear();
    f.delivered = true;
}

void retire(bsf_session *s)
{
    while (!s->inflight.empty() && s->inflight.front().delivered) s->inflight.pop_front();
}

// Waits for a frame, delivers it and returns its slots if it is the oldest.
int finish(bsf_session *s, Frame &f)
{
    if (f.delivered) return BSF_OK;
    hipError_t e = hipEventSynchronize(f.events.back());
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: frame %llu failed: %s", f.ticket, hipGetErrorString(e));
    deliver(s, f);
    retire(s);
    return BSF_OK;
}

// Ring allocation of `bytes` behind the newest in-flight frame; SIZE_MAX if
// there is no room now.
size_t ring_alloc(const bsf_session *s, size_t bytes)
{
    if (s->inflight.empty()) return bytes <= s->arena ? 0 : SIZE_MAX;
    const Frame &a = s->inflight.front(), &b = s->inflight.back();
    const size_t head = b.off + b.bytes, tail = a.off;
    if (b.off >= a.off) {   // not wrapped: free [head, arena) and [0, tail)
        if (s->arena - head >= bytes) return head;
        if (tail >= bytes) return 0;
        return SIZE_MAX;
    }
    return (tail - head >= bytes) ? head : SIZE_MAX;   // wrapped: free [head, tail)
}

// Waits for every frame in flight and delivers it (destroy, model release).
int drain(bsf_session *s)
{
    int rc = BSF_OK;
    while (!s->inflight.empty()) {
        Frame &f = s->inflight.front();
        if (f.delivered) {
            s->inflight.pop_front();
            continue;
        }
        int r = finish(s, f);   // delivers and pops on success
        if (r) {
            if (!rc) rc = r;
            for (hipEvent_t e : f.events) s->event_pool.push_back(e);
            for (const auto &w : f.work) s->load[w.first] -= w.second;
            s->inflight.pop_front();
        }
    }
    return rc;
}

int validate_model(const bsf_model_desc *m)
{
    if (!m) return fail(BSF_ERR_ARG, "bsf: NULL model description");
    const int D = dim_of(m->fittype);
    if (D < 0) return fail(BSF_ERR_FITTYPE, "bsf: model fittype %d is not 6, 7, 8 or BSF_FT_4PI", m->fittype);
    if (!m->coeff) return fail(BSF_ERR_ARG, "bsf: coeff must not be NULL");
    if (m->coeff_ndim != D)
        return fail(BSF_ERR_SHAPE, "bsf: fittype %d needs a %d-D coefficient array (got %d-D)", m->fittype, D,
                    m->coeff_ndim);
    for (int a = 0; a < D; a++)
        if (m->coeff_dims[a] < 5)
            return fail(BSF_ERR_SHAPE, "bsf: coefficient dim %d is %d; each axis needs >= 2 data samples (+3)", a,
                        m->coeff_dims[a]);
    if (m->fittype == 8 && !m->coeff3) return fail(BSF_ERR_ARG, "bsf: fittype 8 needs coeff3");
    if (m->fittype == BSF_FT_4PI) {
        if (m->nch < 1 || m->nch > 8) return fail(BSF_ERR_ARG, "bsf: 4Pi models need 1..8 channels (got %d)", m->nch);
        int nv = 0;
        for (int p = 0; p < 6; p++) nv += m->shared[p] ? 1 : m->nch;
        if (nv > 24) return fail(BSF_ERR_ARG, "bsf: 4Pi parameter count %d exceeds 24", nv);
    }
    return BSF_OK;
}

// Fills M from a validated description, checks the HBM budget and allocates
// the model's device buffers (nc / nc3 coefficient counts).
int alloc_model(bsf_session *s, const bsf_model_desc *m, Model &M, size_t &nc, size_t &nc3)
{
    M.ft = m->fittype;
    M.ndim = m->coeff_ndim;
    nc = 1;
    for (int a = 0; a < M.ndim; a++) {
        M.cdims[a] = m->coeff_dims[a];
        nc *= (size_t)m->coeff_dims[a];
    }
    nc3 = M.ft == 8 ? (size_t)M.cdims[0] * M.cdims[1] * M.cdims[2] : 0;
    if (nc * 4 >= ((size_t)1 << 31) || nc3 * 4 >= ((size_t)1 << 31))
        return fail(BSF_ERR_SHAPE, "bsf: coefficient arrays must be below 2 GiB");
    M.bytes = (nc * 4 + 255) / 256 * 256 + (nc3 * 4 + 255) / 256 * 256;
    if (s->model_bytes + M.bytes + s->arena > s->budget)
        return fail(BSF_ERR_NOMEM, "bsf: model of %zu B exceeds the HBM budget (%zu B used of %zu B)", M.bytes,
                    s->model_bytes + s->arena, s->budget);
    if (M.ft == BSF_FT_4PI) {
        M.nch = m->nch;
        M.nv = 0;
        for (int p = 0; p < 6; p++) {
            M.shared[p] = m->shared[p] ? 1 : 0;
            M.nv += M.shared[p] ? 1 : M.nch;
        }
        for (int k = 0; k < M.nch; k++) M.shift[k] = m->phaseShift[k];
    } else {
        M.nv = nv_of(M.ft);
    }
    hipError_t e = hipMalloc((void **)&M.d_coeff, nc * 4);
    if (e == hipSuccess && nc3) e = hipMalloc((void **)&M.d_coeff3, nc3 * 4);
    if (e != hipSuccess) {
        if (M.d_coeff) (void)hipFree(M.d_coeff);
        M.d_coeff = nullptr;
        return fail(BSF_ERR_NOMEM, "bsf: model allocation failed: %s", hipGetErrorString(e));
    }
    return BSF_OK;
}

// Publishes a model whose buffers were filled (e: the fill's status).
int add_model(bsf_session *s, Model &M, hipError_t e, int *model_id)
{
    if (e != hipSuccess) {
        (void)hipFree(M.d_coeff);
        if (M.d_coeff3) (void)hipFree(M.d_coeff3);
        return fail(BSF_ERR_DEVICE, "bsf: model upload failed: %s", hipGetErrorString(e));
    }
    M.live = true;
    s->model_bytes += M.bytes;
    s->models.push_back(M);
    *model_id = (int)s->models.size() - 1;
    return BSF_OK;
}

// ---------------------------------------------------------------- submission
// Launch checks of a frame's launch at its largest ROI size: LDS, and the 4Pi
// kernel choice.
int launch_lds(const bsf_session *s, Launch &l, int f64)
{
    const Model &M = s->models[l.model];
    // 4Pi: one wave per ROI.  The wave-per-channel kernel halves one
    // launch's latency on an idle GPU, but with frames overlapping it
    // gained nothing measurable in the streamed mix (p50 2.17 vs 2.25 ms,
    // p99 5.19 vs 4.88 ms) and saturated lower (3.15M vs 3.24M fits/s)
    l.wpc = M.ft == BSF_FT_4PI && bsf::use_wpc_4pi(0, f64, l.Smax, M.nch, M.nv, s->lds_limit);
    const size_t lds = M.ft == BSF_FT_4PI ? (l.wpc ? bsf::lds_bytes_4pi_wpc(f64, l.Smax, M.nch, M.nv)
                                                   : bsf::lds_bytes_4pi(f64, l.Smax, M.nch, M.nv))
                                          : bsf::lds_bytes(M.ft, f64, l.Smax);
    if (lds > s->lds_limit)
        return fail(BSF_ERR_ROI_SIZE, "bsf: ROI size %d of model %d needs %zu B of LDS (limit %zu B)", l.Smax,
                    l.model, lds, s->lds_limit);
    return BSF_OK;
}

// Orders a launch's groups largest ROI size first and numbers their rows;
// returns the launch's data floats.
size_t launch_rows(Launch &l, const bsf_frame *f)
{
    std::stable_sort(l.groups.begin(), l.groups.end(),
                     [&](int a, int b) { return f->groups[a].sz > f->groups[b].sz; });
    size_t nd = 0;
    l.start.clear();
    int row = 0;
    for (int gi : l.groups) {
        const bsf_roi_group &g = f->groups[gi];
        l.start.push_back(row);
        row += g.count;
        nd += (size_t)g.count * g.sz * g.sz * l.nch;
    }
    return nd;
}

// Packs launch l's inputs into the pinned slot h: ROI descriptors at
// at.o_desc (their offsets count from the float at desc_base), images at
// byte offset o_data, 4Pi starts / offsets at at.o_dT / o_iz / o_ip.
void pack_launch(const bsf_frame *f, const Launch &l, const Launch &at, char *h, size_t o_data, long long desc_base)
{
    bsf::RoiDesc *desc = (bsf::RoiDesc *)(h + at.o_desc);
    float *dst = (float *)(h + o_data);
    long long e0 = 0;
    int row = 0;
    for (size_t j = 0; j < l.groups.size(); j++) {
        const bsf_roi_group &g = f->groups[l.groups[j]];
        const size_t S2 = (size_t)g.sz * g.sz, cnt = (size_t)g.count;
        const size_t cs = g.chan_stride ? g.chan_stride : cnt * S2;
        for (int k = 0; k < l.nch; k++) std::memcpy(dst + e0 + k * cnt * S2, g.data + k * cs, cnt * S2 * 4);
        for (size_t i = 0; i < cnt; i++) {
            desc[row + i].off = desc_base + e0 + (long long)(i * S2);
            desc[row + i].cstride = (int)(cnt * S2);
            desc[row + i].S = g.sz;
        }
        const int nch = l.nch;
        if (l.has_dT) std::memcpy((float *)(h + at.o_dT) + (size_t)row * 6 * nch, g.dT, cnt * 6 * nch * 4);
        if (l.has_iz) std::memcpy((float *)(h + at.o_iz) + row, g.initZ, cnt * 4);
        if (l.has_ip) std::memcpy((float *)(h + at.o_ip) + row, g.initPhi, cnt * 4);
        e0 += (long long)(cnt * S2 * l.nch);
        row += g.count;
    }
}

// The fit launch (fittypes 6-8) of launch l's first start: slot base d, the
// kernels' data pointer `data`, ncount as for launch_one.
bsf::FitLaunch fit_launch(const bsf_session *s, const Launch &l, int f64, char *d, const float *data,
                          const int *ncount)
{
    const Model &M = s->models[l.model];
    bsf::FitLaunch K;
    std::memset(&K, 0, sizeof K);
    K.fittype = M.ft;
    K.f64 = f64;
    K.data = data;
    K.coeff = M.d_coeff;
    K.coeff3 = M.d_coeff3;
    K.P = (float *)(d + l.o_P);
    K.CRLB = (float *)(d + l.o_CR);
    K.LL = (float *)(d + l.o_LL);
    K.N = l.n;
    K.S = l.Smax;
    K.iterations = l.iterations;
    K.roi_offset = 0;
    K.count = l.n;
    K.rois = (const bsf::RoiDesc *)(d + l.o_desc);
    K.uniform = !ncount && l.Smin == l.Smax;
    K.ndim = M.ndim;
    for (int a = 0; a < 5; a++) K.cdims[a] = M.cdims[a];
    K.has_initP = l.has_initP;
    for (int q = 0; q < 4; q++) K.initP[q] = l.initP[0][q];
    K.ncount = ncount;
    return K;
}

// Queues launch l (all its starts) on ks: slot base d, the kernels' data
// pointer `data` (descriptor offsets count from it), outputs and tables at
// l's offsets; ncount: a device-side ROI count (l.n is then the capacity).
hipError_t launch_one(const bsf_session *s, const Launch &l, int f64, char *d, const float *data,
                      const int *ncount, hipStream_t ks)
{
    const Model &M = s->models[l.model];
    const bsf::RoiDesc *d_desc = (const bsf::RoiDesc *)(d + l.o_desc);
    float *dP = (float *)(d + l.o_P), *dCR = (float *)(d + l.o_CR), *dLL = (float *)(d + l.o_LL);
    if (M.ft == BSF_FT_4PI) {
        bsf::Fit4PiLaunch K;
        std::memset(&K, 0, sizeof K);
        K.f64 = f64;
        K.data = data;
        K.coeff = M.d_coeff;
        K.P = dP;
        K.CRLB = dCR;
        K.LL = dLL;
        K.N = l.n;
        K.S = l.Smax;
        K.iterations = l.iterations;
        K.nch = M.nch;
        K.roi_offset = 0;
        K.count = l.n;
        K.rois = d_desc;
        for (int a = 0; a < 4; a++) K.cdims[a] = M.cdims[a];
        for (int p = 0; p < 6; p++) K.shared[p] = M.shared[p];
        for (int k = 0; k < 8; k++) K.phaseShift[k] = M.shift[k];
        K.dT = l.has_dT ? (const float *)(d + l.o_dT) : nullptr;
        K.initZ = l.has_iz ? (const float *)(d + l.o_iz) : nullptr;
        K.initPhi = l.has_ip ? (const float *)(d + l.o_ip) : nullptr;
        K.wpc = l.wpc;
        K.ncount = ncount;
        return bsf::launch_fit_4pi(K, ks);
    }
    bsf::FitLaunch K = fit_launch(s, l, f64, d, data, ncount);
    // multi-start: one launch per start on this stream; after the first,
    // each keeps the results whose log-likelihood it beats
    hipError_t e = hipSuccess;
    for (int k = 0; k < l.nstarts && e == hipSuccess; k++) {
        for (int q = 0; q < 4; q++) K.initP[q] = l.initP[k][q];
        K.keep = k == 0 ? 0 : (l.strict ? 2 : 1);
        e = bsf::launch_fit(K, ks);
    }
    return e;
}

// Room for `bytes` in the arena behind the frames in flight (waiting for the
// oldest as needed, or BSF_BUSY with BSF_SUBMIT_NONBLOCK).
int admit(bsf_session *s, size_t bytes, int flags, size_t *off)
{
    while ((*off = ring_alloc(s, bytes)) == SIZE_MAX) {
        if (flags & BSF_SUBMIT_NONBLOCK)
            return fail(BSF_BUSY, "bsf: arena full (%d frames in flight)", (int)s->inflight.size());
        Frame &oldest = s->inflight.front();
        if (oldest.delivered) {
            retire(s);
            continue;
        }
        int rc = finish(s, oldest);
        if (rc) return rc;
    }
    return BSF_OK;
}

// Work queued for a frame whose submission failed may still read its slot:
// wait for it before the slot can be handed out again.
int abandon(bsf_session *s, Frame &F, hipError_t e)
{
    (void)hipStreamSynchronize(s->s_in);
    for (int j = 0; j < s->nks; j++) (void)hipStreamSynchronize(s->ks[j]);
    (void)hipStreamSynchronize(s->s_out);
    for (hipEvent_t ev : F.events) s->event_pool.push_back(ev);
    for (const auto &w : F.work) s->load[w.first] -= w.second;
    return fail(BSF_ERR_DEVICE, "bsf: frame submission failed: %s", hipGetErrorString(e));
}

void add_outputs(Frame &F, const bsf_frame *f, const Launch &l, const Launch &at)
{
    for (size_t j = 0; j < l.groups.size(); j++) {
        const bsf_roi_group &g = f->groups[l.groups[j]];
        GroupOut go;
        go.P = g.P;
        go.CRLB = g.CRLB;
        go.LL = g.LL;
        go.count = g.count;
        go.nv = l.nv;
        go.start = l.start[j];
        go.launch_n = at.n;
        go.o_P = at.o_P;
        go.o_CR = at.o_CR;
        go.o_LL = at.o_LL;
        F.outs.push_back(go);
    }
}

// One frame, launched call by call: inputs then outputs in a ring slot of the
// frame's size (graph mode: whole graph slots), one H2D copy, the launches
// spread over the kernel streams, one D2H copy.
int submit_eager(bsf_session *s, const bsf_frame *f, int flags, unsigned long long *ticket, std::vector<Launch> &L,
                 int f64)
{
    size_t o = 0;
    for (Launch &l : L) {
        const Model &M = s->models[l.model];
        int rc = launch_lds(s, l, f64);
        if (rc) return rc;
        const size_t nd = launch_rows(l, f);
        double cost = 0;
        for (int gi : l.groups) {
            const bsf_roi_group &g = f->groups[gi];
            cost += (double)g.count * g.sz * g.sz * l.nch * (double)(1 << (2 * (dim_of(M.ft) - 2)));
        }
        l.cost = cost * (l.iterations + 2) * l.nstarts;
        l.o_desc = o;
        o += up((size_t)l.n * sizeof(bsf::RoiDesc));
        l.o_data = o;
        o += up(nd * 4);
        l.o_dT = o;
        o += l.has_dT ? up((size_t)l.n * 6 * l.nch * 4) : 0;
        l.o_iz = o;
        o += l.has_iz ? up((size_t)l.n * 4) : 0;
        l.o_ip = o;
        o += l.has_ip ? up((size_t)l.n * 4) : 0;
        l.o_end = o;
    }
    const size_t in_bytes = o;
    for (Launch &l : L) {
        l.o_P = o;
        o += up((size_t)l.n * (l.nv + 1) * 4);
        l.o_CR = o;
        o += up((size_t)l.n * l.nv * 4);
        l.o_LL = o;
        o += up((size_t)l.n * 4);
    }
    const size_t total = o;
    // in graph mode a frame takes whole slots, so that slots stay aligned
    const size_t take = s->graphs ? (total + s->slot_bytes - 1) / s->slot_bytes * s->slot_bytes : total;
    if (take > s->arena)
        return fail(BSF_ERR_NOMEM, "bsf: frame needs %zu B, more than the whole arena (%zu B)", take, s->arena);

    size_t off;
    int rc = admit(s, take, flags, &off);
    if (rc) return rc;
    char *h = s->h_arena + off, *d = s->d_arena + off;
    // The longest launch first: its inputs are packed, copied in and the
    // launch queued before the other launches' inputs are packed (and copied
    // in one piece), so the frame's longest kernel starts while the rest of
    // the frame is still being packed.  A heterogeneous launch needs several
    // launches' inputs, so in that mode the whole input section is packed
    // and copied first.
    const bool piped = !s->hetero;
    if (!piped)
        for (Launch &l : L) pack_launch(f, l, l, h, l.o_data, 0);

    Frame F;
    F.off = off;
    F.bytes = take;
    hipEvent_t ev_in = nullptr;
    hipError_t e = hipSuccess;
    if (!piped) {
        ev_in = take_event(s);
        if (!ev_in) return fail(BSF_ERR_DEVICE, "bsf: hipEventCreate failed");
        F.events.push_back(ev_in);
        e = hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, s->s_in);
        if (e == hipSuccess) e = hipEventRecord(ev_in, s->s_in);
    }
    // longest launches first, each to the kernel stream with the least queued
    // work (launches of earlier frames not yet delivered included), so the
    // big launches of consecutive frames run side by side.  A launch's length
    // is its cost, a 4Pi launch's 1.6 times that: with the same work per ROI
    // its one wave per ROI takes the channels in turn (a lone frame's 4Pi
    // launch of ~400 ROIs ran 1.43 ms against 0.87 ms for its vortex-5D one).
    auto length = [&](const Launch &l) { return l.cost * (s->models[l.model].ft == BSF_FT_4PI ? 1.6 : 1.0); };
    std::vector<int> &ord = s->order;
    ord.resize(L.size());
    for (size_t i = 0; i < L.size(); i++) ord[i] = (int)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return length(L[a]) > length(L[b]); });
    // heterogeneous mode: the single-start launches of fittypes 6-8 go
    // together, up to kHeteroMax per launch (the first of them carries it)
    std::vector<std::vector<int>> het;
    if (s->hetero) {
        for (int i : ord) {
            const int ft = s->models[L[i].model].ft;
            if (ft < 6 || ft > 8 || L[i].nstarts != 1) continue;
            if (het.empty() || het.back().size() == (size_t)bsf::kHeteroMax) het.emplace_back();
            het.back().push_back(i);
        }
        for (auto &grp : het)
            for (size_t q = 1; q < grp.size(); q++) L[grp[0]].cost += L[grp[q]].cost, L[grp[q]].cost = -1;
    }
    for (int i : ord) {
        if (e != hipSuccess) break;
        Launch &l = L[i];
        if (l.cost < 0) continue;   // part of a heterogeneous launch
        if (piped && (i == ord[0] || (ord.size() > 1 && i == ord[1]))) {
            // the longest launch's inputs, then (at the second) everyone else's
            ev_in = take_event(s);
            if (!ev_in) {
                e = hipErrorOutOfMemory;
                break;
            }
            F.events.push_back(ev_in);
            if (i == ord[0]) {
                pack_launch(f, l, l, h, l.o_data, 0);
                e = hipMemcpyAsync(d + l.o_desc, h + l.o_desc, l.o_end - l.o_desc, hipMemcpyHostToDevice, s->s_in);
            } else {
                for (size_t q = 1; q < ord.size(); q++) pack_launch(f, L[ord[q]], L[ord[q]], h, L[ord[q]].o_data, 0);
                // everything but the first launch's inputs: before and after them
                const Launch &a = L[ord[0]];
                if (a.o_desc > 0) e = hipMemcpyAsync(d, h, a.o_desc, hipMemcpyHostToDevice, s->s_in);
                if (e == hipSuccess && a.o_end < in_bytes)
                    e = hipMemcpyAsync(d + a.o_end, h + a.o_end, in_bytes - a.o_end, hipMemcpyHostToDevice, s->s_in);
            }
            if (e == hipSuccess) e = hipEventRecord(ev_in, s->s_in);
            if (e != hipSuccess) break;
        }
        hipEvent_t ev_l = ev_in;
        int j = 0;
        for (int q = 1; q < s->nks; q++)
            if (s->load[q] < s->load[j]) j = q;
        s->load[j] += l.cost;
        F.work.emplace_back(j, l.cost);
        l.stream = j;
        hipStream_t ks = s->ks[j];
        e = hipStreamWaitEvent(ks, ev_l, 0);
        const std::vector<int> *grp = nullptr;
        for (const auto &x : het)
            if (x[0] == i && x.size() > 1) grp = &x;
        if (e == hipSuccess && grp) {
            bsf::FitLaunch K[bsf::kHeteroMax];
            for (size_t q = 0; q < grp->size(); q++) {
                const Launch &m = L[(*grp)[q]];
                K[q] = fit_launch(s, m, f64, d, (const float *)(d + m.o_data), nullptr);
            }
            e = bsf::launch_fit_hetero(K, (int)grp->size(), ks);
        } else if (e == hipSuccess) {
            e = launch_one(s, l, f64, d, (const float *)(d + l.o_data), nullptr, ks);
        }
        if (e != hipSuccess) break;
        hipEvent_t ev = take_event(s);
        if (!ev) {
            e = hipErrorOutOfMemory;
            break;
        }
        F.events.push_back(ev);
        e = hipEventRecord(ev, ks);
        if (e == hipSuccess) e = hipStreamWaitEvent(s->s_out, ev, 0);
    }
    hipEvent_t ev_done = nullptr;
    if (e == hipSuccess) {
        ev_done = take_event(s);
        if (!ev_done) e = hipErrorOutOfMemory;
    }
    if (e == hipSuccess) {
        F.events.push_back(ev_done);
        e = hipMemcpyAsync(h + in_bytes, d + in_bytes, total - in_bytes, hipMemcpyDeviceToHost, s->s_out);
        if (e == hipSuccess) e = hipEventRecord(ev_done, s->s_out);
    }
    if (e != hipSuccess) return abandon(s, F, e);
    for (Launch &l : L) add_outputs(F, f, l, l);
    F.ticket = s->next_ticket++;
    *ticket = F.ticket;
    s->inflight.push_back(std::move(F));
    return BSF_OK;
}

// ---------------------------------------------------------------- graph mode
bool same_key(const Launch &a, const Launch &b)
{
    return a.model == b.model && a.iterations == b.iterations && a.has_initP == b.has_initP &&
           a.nstarts == b.nstarts && a.strict == b.strict && std::memcmp(a.initP, b.initP, sizeof a.initP) == 0 &&
           a.has_dT == b.has_dT && a.has_iz == b.has_iz && a.has_ip == b.has_ip;
}

bool key_less(const Launch &a, const Launch &b)
{
    if (a.model != b.model) return a.model < b.model;
    if (a.iterations != b.iterations) return a.iterations < b.iterations;
    if (a.nstarts != b.nstarts) return a.nstarts < b.nstarts;
    if (a.strict != b.strict) return a.strict < b.strict;
    if (a.has_initP != b.has_initP) return a.has_initP < b.has_initP;
    if (a.has_dT != b.has_dT) return a.has_dT < b.has_dT;
    if (a.has_iz != b.has_iz) return a.has_iz < b.has_iz;
    if (a.has_ip != b.has_ip) return a.has_ip < b.has_ip;
    return std::memcmp(a.initP, b.initP, sizeof a.initP) < 0;
}

void drop_plan(bsf_session *s, Plan *P, bool now)
{
    for (hipGraphExec_t &x : P->exec)
        if (x) {
            if (now) (void)hipGraphExecDestroy(x);
            else s->old_execs.push_back(x);
            x = nullptr;
        }
    delete P;
}

void drop_plans(bsf_session *s)   // nothing may be in flight
{
    for (Plan *P : s->plans) drop_plan(s, P, true);
    s->plans.clear();
    for (hipGraphExec_t x : s->old_execs) (void)hipGraphExecDestroy(x);
    s->old_execs.clear();
}

// A plan for launches L (keys, counts n, sizes Smax) with room for twice
// their counts and their sizes or `prev`'s, whichever is larger; NULL if even
// the frame's own counts do not fit a slot with its data.
Plan *make_plan(const bsf_session *s, const std::vector<Launch> &L, int f64, const Plan *prev, size_t data_bytes)
{
    for (int grow = 2; grow >= 1; grow--) {
        Plan *P = new Plan;
        P->f64 = f64;
        P->L = L;
        size_t o = up(L.size() * 4);
        P->o_hdr = 0;
        bool ok = true;
        for (size_t i = 0; i < L.size(); i++) {
            Launch &l = P->L[i];
            int cap = L[i].n * grow;
            if (prev) cap = std::max(cap, prev->L[i].n);
            cap = std::max(cap, 64);
            l.n = (cap + 63) / 64 * 64;
            if (prev) l.Smax = std::max(l.Smax, prev->L[i].Smax);
            l.Smin = 0;
            l.groups.clear();
            l.start.clear();
            if (launch_lds(s, l, f64)) ok = false;
            l.o_desc = o;
            o += up((size_t)l.n * sizeof(bsf::RoiDesc));
            l.o_dT = o;
            o += l.has_dT ? up((size_t)l.n * 6 * l.nch * 4) : 0;
            l.o_iz = o;
            o += l.has_iz ? up((size_t)l.n * 4) : 0;
            l.o_ip = o;
            o += l.has_ip ? up((size_t)l.n * 4) : 0;
        }
        P->o_data = o;
        size_t out = 0;
        for (Launch &l : P->L)
            out += up((size_t)l.n * (l.nv + 1) * 4) + up((size_t)l.n * l.nv * 4) + up((size_t)l.n * 4);
        if (!ok || o + out + data_bytes > s->slot_bytes) {
            delete P;
            if (!ok) return nullptr;
            continue;
        }
        P->o_out = s->slot_bytes - out;
        P->data_cap = P->o_out - P->o_data;
        size_t q = P->o_out;
        for (Launch &l : P->L) {
            l.o_P = q;
            q += up((size_t)l.n * (l.nv + 1) * 4);
            l.o_CR = q;
            q += up((size_t)l.n * l.nv * 4);
            l.o_LL = q;
            q += up((size_t)l.n * 4);
        }
        return P;
    }
    return nullptr;
}

// Captures plan P's launches for slot `slot` (one branch per launch, all
// reading their ROI counts from the slot header) and instantiates the graph.
hipError_t capture(bsf_session *s, Plan &P, int slot)
{
    hipError_t e = hipSuccess;
    const int nb = (int)std::min<size_t>(P.L.size(), kMaxKernelStreams);
    for (int b = 0; b <= nb && e == hipSuccess; b++)
        if (!s->s_cap[b]) e = hipStreamCreateWithFlags(&s->s_cap[b], hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    std::vector<hipEvent_t> ev(nb + 1, nullptr);
    for (int b = 0; b <= nb && e == hipSuccess; b++) e = hipEventCreateWithFlags(&ev[b], hipEventDisableTiming);
    hipGraph_t g = nullptr;
    if (e == hipSuccess) e = hipStreamBeginCapture(s->s_cap[0], hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
        char *d = s->d_arena + (size_t)slot * s->slot_bytes;
        const int *cnt = (const int *)(d + P.o_hdr);
        hipError_t c = hipEventRecord(ev[0], s->s_cap[0]);
        for (int b = 0; b < nb && c == hipSuccess; b++) c = hipStreamWaitEvent(s->s_cap[1 + b], ev[0], 0);
        for (size_t i = 0; i < P.L.size() && c == hipSuccess; i++)
            c = launch_one(s, P.L[i], P.f64, d, (const float *)d, cnt + i, s->s_cap[1 + i % nb]);
        for (int b = 0; b < nb && c == hipSuccess; b++) {
            c = hipEventRecord(ev[1 + b], s->s_cap[1 + b]);
            if (c == hipSuccess) c = hipStreamWaitEvent(s->s_cap[0], ev[1 + b], 0);
        }
        e = hipStreamEndCapture(s->s_cap[0], &g);   // always ends the capture
        if (c != hipSuccess) e = c;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&P.exec[slot], g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e == hipSuccess) s->n_capture++;
    return e;
}

constexpr int kNoPlan = -1000;   // submit_graph: no plan fits, submit eagerly

// One frame as a graph replay: find (or make) the plan of its launches, pack
// its inputs at the plan's offsets of the next slot, one H2D copy, the slot's
// graph on a kernel stream, one D2H copy of the output section.
int submit_graph(bsf_session *s, const bsf_frame *f, int flags, unsigned long long *ticket, std::vector<Launch> &L,
                 int f64)
{
    std::sort(L.begin(), L.end(), key_less);
    size_t data_bytes = 0;
    std::vector<size_t> nd(L.size());
    for (size_t i = 0; i < L.size(); i++) {
        nd[i] = launch_rows(L[i], f);
        data_bytes += up(nd[i] * 4);
    }
    Plan *P = nullptr;
    size_t pi = 0;
    for (; pi < s->plans.size(); pi++) {
        Plan *Q = s->plans[pi];
        if (Q->f64 != f64 || Q->L.size() != L.size()) continue;
        bool same = true;
        for (size_t i = 0; i < L.size() && same; i++) same = same_key(Q->L[i], L[i]);
        if (same) {
            P = Q;
            break;
        }
    }
    bool fits = P && data_bytes <= P->data_cap;
    for (size_t i = 0; i < L.size() && fits; i++) fits = L[i].n <= P->L[i].n && L[i].Smax <= P->L[i].Smax;
    if (!fits) {
        Plan *Q = make_plan(s, L, f64, P, data_bytes);
        if (!Q) {
            bsf::clear_error();   // the eager path reports its own errors
            return kNoPlan;
        }
        if (P) {
            drop_plan(s, P, false);   // its graphs may be in flight
            s->plans[pi] = Q;
        } else {
            s->plans.push_back(Q);
        }
        P = Q;
    }

    size_t off;
    int rc = admit(s, s->slot_bytes, flags, &off);
    if (rc) return rc;
    const int slot = (int)(off / s->slot_bytes);
    if (!P->exec[slot]) {
        hipError_t e = capture(s, *P, slot);
        if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: graph capture failed: %s", hipGetErrorString(e));
    }
    char *h = s->h_arena + off, *d = s->d_arena + off;
    size_t od = P->o_data;
    for (size_t i = 0; i < L.size(); i++) {
        ((int *)(h + P->o_hdr))[i] = L[i].n;
        pack_launch(f, L[i], P->L[i], h, od, (long long)(od / 4));
        od += up(nd[i] * 4);
    }

    Frame F;
    F.off = off;
    F.bytes = s->slot_bytes;
    hipEvent_t ev_in = take_event(s), ev_k = take_event(s), ev_done = take_event(s);
    for (hipEvent_t x : {ev_in, ev_k, ev_done})
        if (x) F.events.push_back(x);
    hipError_t e = F.events.size() == 3 ? hipSuccess : hipErrorOutOfMemory;
    const hipStream_t gs = s->ks[slot % s->nks];
    if (e == hipSuccess) e = hipMemcpyAsync(d, h, od, hipMemcpyHostToDevice, s->s_in);
    if (e == hipSuccess) e = hipEventRecord(ev_in, s->s_in);
    if (e == hipSuccess) e = hipStreamWaitEvent(gs, ev_in, 0);
    if (e == hipSuccess) e = hipGraphLaunch(P->exec[slot], gs);
    if (e == hipSuccess) e = hipEventRecord(ev_k, gs);
    if (e == hipSuccess) e = hipStreamWaitEvent(s->s_out, ev_k, 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h + P->o_out, d + P->o_out, s->slot_bytes - P->o_out, hipMemcpyDeviceToHost, s->s_out);
    if (e == hipSuccess) e = hipEventRecord(ev_done, s->s_out);
    if (e != hipSuccess) return abandon(s, F, e);
    s->n_replay++;
    for (size_t i = 0; i < L.size(); i++) add_outputs(F, f, L[i], P->L[i]);
    F.ticket = s->next_ticket++;
    *ticket = F.ticket;
    s->inflight.push_back(std::move(F));
    return BSF_OK;
}

}  // namespace

extern "C" {

int bsf_session_create(const bsf_session_config *cfg, bsf_session **out)
{
    bsf::clear_error();
    if (!cfg || !out) return fail(BSF_ERR_ARG, "bsf: NULL session config or output");
    *out = nullptr;
    if (cfg->precision != BSF_PREC_MIX && cfg->precision != BSF_PREC_F64)
        return fail(BSF_ERR_ARG, "bsf: unknown precision mode %d", cfg->precision);
    const int nks = cfg->kernel_streams ? cfg->kernel_streams : 2;
    if (nks < 1 || nks > kMaxKernelStreams)
        return fail(BSF_ERR_ARG, "bsf: kernel_streams must be 1..%d (got %d)", kMaxKernelStreams, nks);
    const size_t budget = cfg->hbm_budget ? cfg->hbm_budget : ((size_t)2 << 30);
    const size_t arena = up(cfg->arena_bytes ? cfg->arena_bytes : ((size_t)256 << 20));
    if (arena > budget)
        return fail(BSF_ERR_NOMEM, "bsf: arena of %zu B exceeds the HBM budget of %zu B", arena, budget);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || cfg->device < 0 || cfg->device >= ndev)
        return fail(BSF_ERR_DEVICE, "bsf: no HIP device %d (%d visible)", cfg->device, e == hipSuccess ? ndev : 0);
    hipDeviceProp_t p;
    if ((e = hipGetDeviceProperties(&p, cfg->device)) != hipSuccess)
        return fail(BSF_ERR_DEVICE, "bsf: hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(BSF_ERR_DEVICE, "bsf: device %d is %s; this build targets gfx950 (MI355X)", cfg->device,
                    p.gcnArchName);
    DeviceGuard g(cfg->device);
    bsf_session *s = new bsf_session;
    s->device = cfg->device;
    s->precision = cfg->precision;
    s->nks = nks;
    s->budget = budget;
    s->arena = arena;
    s->lds_limit = p.sharedMemPerBlock;
    e = hipMalloc((void **)&s->d_arena, arena);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->h_arena, arena, hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->s_in, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->s_out, hipStreamNonBlocking);
    for (int j = 0; j < nks && e == hipSuccess; j++) e = hipStreamCreateWithFlags(&s->ks[j], hipStreamNonBlocking);
    if (e != hipSuccess) {
        bsf_session_destroy(s);
        return fail(e == hipErrorOutOfMemory ? BSF_ERR_NOMEM : BSF_ERR_DEVICE, "bsf: session set-up failed: %s",
                    hipGetErrorString(e));
    }
    // Touch every page of the pinned mirror now: a page's first CPU write
    // faults, and packing a frame into fresh pages runs at a quarter of the
    // speed, which a stream of frames cannot absorb once the ring advances
    // into them.
    std::memset(s->h_arena, 0, arena);
    *out = s;
    return BSF_OK;
}

int bsf_session_destroy(bsf_session *s)
{
    bsf::clear_error();
    if (!s) return BSF_OK;
    int rc = BSF_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        DeviceGuard g(s->device);
        rc = drain(s);
        for (Model &m : s->models) {
            if (m.d_coeff) (void)hipFree(m.d_coeff);
            if (m.d_coeff3) (void)hipFree(m.d_coeff3);
        }
        drop_plans(s);
        for (hipStream_t c : s->s_cap)
            if (c) (void)hipStreamDestroy(c);
        for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
        if (s->s_in) (void)hipStreamDestroy(s->s_in);
        if (s->s_out) (void)hipStreamDestroy(s->s_out);
        for (int j = 0; j < kMaxKernelStreams; j++)   // entries past nks are null
            if (s->ks[j]) (void)hipStreamDestroy(s->ks[j]);
        if (s->d_arena) (void)hipFree(s->d_arena);
        if (s->h_arena) (void)hipHostFree(s->h_arena);
    }
    delete s;
    return rc;
}

int bsf_model_register(bsf_session *s, const bsf_model_desc *m, int *model_id)
{
    bsf::clear_error();
    if (!s || !model_id) return fail(BSF_ERR_ARG, "bsf: NULL session or model id");
    int rc = validate_model(m);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    DeviceGuard g(s->device);
    Model M;
    size_t nc = 0, nc3 = 0;
    if ((rc = alloc_model(s, m, M, nc, nc3))) return rc;
    hipError_t e = hipMemcpy(M.d_coeff, m->coeff, nc * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && nc3) e = hipMemcpy(M.d_coeff3, m->coeff3, nc3 * 4, hipMemcpyHostToDevice);
    return add_model(s, M, e, model_id);
}

int bsf_model_register_samples(bsf_session *s, const bsf_model_samples *m, int *model_id)
{
    bsf::clear_error();
    if (!s || !m || !model_id) return fail(BSF_ERR_ARG, "bsf: NULL session, samples or model id");
    if (m->ndim < 1 || m->ndim > 5) return fail(BSF_ERR_SHAPE, "bsf: model samples need 3..5 dims (got %d)", m->ndim);
    bsf_model_desc d;
    std::memset(&d, 0, sizeof d);
    d.fittype = m->fittype;
    d.coeff = m->samples;
    d.coeff_ndim = m->ndim;
    for (int a = 0; a < m->ndim; a++) d.coeff_dims[a] = m->dims[a] + 3;
    d.coeff3 = m->samples3;
    d.nch = m->nch;
    for (int p = 0; p < 6; p++) d.shared[p] = m->shared[p];
    for (int k = 0; k < 8; k++) d.phaseShift[k] = m->phaseShift[k];
    int rc = validate_model(&d);
    if (rc) return rc;
    for (int a = 0; a < m->ndim; a++)
        if ((m->lambda && (!(m->lambda[a] >= 0.0) || std::isinf(m->lambda[a]))) ||
            (a < 3 && m->fittype == 8 && m->lambda3 && (!(m->lambda3[a] >= 0.0) || std::isinf(m->lambda3[a]))))
            return fail(BSF_ERR_ARG, "bsf: lambda must be finite and >= 0 (axis %d)", a);
    std::lock_guard<std::mutex> lk(s->mu);
    DeviceGuard g(s->device);
    Model M;
    size_t nc = 0, nc3 = 0;
    if ((rc = alloc_model(s, &d, M, nc, nc3))) return rc;
    // samples and the builder's scratch live only for the registration
    size_t n = 1;
    for (int a = 0; a < m->ndim; a++) n *= (size_t)m->dims[a];
    char *tmp = nullptr;
    const size_t xb = (n * 4 + 255) / 256 * 256;
    hipError_t e = hipMalloc((void **)&tmp, xb + bsf::coeff_work_bytes(m->dims, m->ndim));
    if (e == hipSuccess) {
        hipStream_t st = s->s_in;
        e = hipMemcpyAsync(tmp, m->samples, n * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = bsf::coeff_build((float *)tmp, m->dims, m->ndim, m->lambda, M.d_coeff, tmp + xb, st);
        if (e == hipSuccess && nc3) {
            const size_t n3 = (size_t)m->dims[0] * m->dims[1] * m->dims[2];
            if (e == hipSuccess) e = hipStreamSynchronize(st);   // tmp is reused
            if (e == hipSuccess) e = hipMemcpyAsync(tmp, m->samples3, n3 * 4, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = bsf::coeff_build((float *)tmp, m->dims, 3, m->lambda3, M.d_coeff3, tmp + xb, st);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        (void)hipFree(tmp);
    }
    return add_model(s, M, e, model_id);
}

int bsf_model_release(bsf_session *s, int model_id)
{
    bsf::clear_error();
    if (!s) return fail(BSF_ERR_ARG, "bsf: NULL session");
    std::lock_guard<std::mutex> lk(s->mu);
    if (model_id < 0 || model_id >= (int)s->models.size() || !s->models[model_id].live)
        return fail(BSF_ERR_ARG, "bsf: unknown model id %d", model_id);
    DeviceGuard g(s->device);
    int rc = drain(s);
    drop_plans(s);   // their graphs launch with the model's coefficients
    Model &M = s->models[model_id];
    (void)hipFree(M.d_coeff);
    if (M.d_coeff3) (void)hipFree(M.d_coeff3);
    M.d_coeff = M.d_coeff3 = nullptr;
    M.live = false;
    s->model_bytes -= M.bytes;
    return rc;
}

int bsf_session_set_precision(bsf_session *s, int precision)
{
    bsf::clear_error();
    if (!s) return fail(BSF_ERR_ARG, "bsf: NULL session");
    if (precision != BSF_PREC_MIX && precision != BSF_PREC_F64)
        return fail(BSF_ERR_ARG, "bsf: unknown precision mode %d", precision);
    std::lock_guard<std::mutex> lk(s->mu);
    s->precision = precision;
    return BSF_OK;
}

int bsf_session_set_graphs(bsf_session *s, int slots)
{
    bsf::clear_error();
    if (!s) return fail(BSF_ERR_ARG, "bsf: NULL session");
    if (slots < 0 || slots > kMaxGraphSlots)
        return fail(BSF_ERR_ARG, "bsf: graph slots must be 0..%d (got %d)", kMaxGraphSlots, slots);
    std::lock_guard<std::mutex> lk(s->mu);
    DeviceGuard g(s->device);
    int rc = drain(s);
    drop_plans(s);
    s->graphs = slots;
    s->slot_bytes = slots ? s->arena / slots / kAlign * kAlign : 0;
    return rc;
}

int bsf_session_set_hetero(bsf_session *s, int on)
{
    bsf::clear_error();
    if (!s) return fail(BSF_ERR_ARG, "bsf: NULL session");
    if (on != 0 && on != 1) return fail(BSF_ERR_ARG, "bsf: heterogeneous mode must be 0 or 1 (got %d)", on);
    std::lock_guard<std::mutex> lk(s->mu);
    s->hetero = on;
    return BSF_OK;
}

int bsf_session_graph_stats(bsf_session *s, unsigned long long *captures, unsigned long long *replays,
                            unsigned long long *eager)
{
    bsf::clear_error();
    if (!s) return fail(BSF_ERR_ARG, "bsf: NULL session");
    std::lock_guard<std::mutex> lk(s->mu);
    if (captures) *captures = s->n_capture;
    if (replays) *replays = s->n_replay;
    if (eager) *eager = s->n_eager;
    return BSF_OK;
}

int bsf_session_inflight(bsf_session *s, int *frames, size_t *arena_used)
{
    bsf::clear_error();
    if (!s) return fail(BSF_ERR_ARG, "bsf: NULL session");
    std::lock_guard<std::mutex> lk(s->mu);
    size_t used = 0;
    for (const Frame &f : s->inflight) used += f.bytes;
    if (frames) *frames = (int)s->inflight.size();
    if (arena_used) *arena_used = used;
    return BSF_OK;
}

int bsf_submit(bsf_session *s, const bsf_frame *f, int flags, unsigned long long *ticket)
{
    bsf::clear_error();
    if (!s || !f || !ticket) return fail(BSF_ERR_ARG, "bsf: NULL session, frame or ticket");
    if (f->ngroups < 0 || (f->ngroups > 0 && !f->groups)) return fail(BSF_ERR_ARG, "bsf: bad group list");
    if (flags & ~BSF_SUBMIT_NONBLOCK) return fail(BSF_ERR_ARG, "bsf: unknown submit flags 0x%x", flags);
    std::lock_guard<std::mutex> lk(s->mu);
    const int f64 = s->precision == BSF_PREC_F64;

    // ---- checks (the MEX's, per group) and grouping into launches
    std::vector<Launch> &L = s->launches;
    L.clear();
    for (int gi = 0; gi < f->ngroups; gi++) {
        const bsf_roi_group &g = f->groups[gi];
        if (g.model < 0 || g.model >= (int)s->models.size() || !s->models[g.model].live)
            return fail(BSF_ERR_ARG, "bsf: group %d: unknown model id %d", gi, g.model);
        const Model &M = s->models[g.model];
        if (g.sz < 1 || g.sz > BSF_MAX_ROI)
            return fail(BSF_ERR_ROI_SIZE, "bsf: group %d: ROI size must be between 1 and %d (got %d)", gi,
                        BSF_MAX_ROI, g.sz);
        if (g.count < 0 || g.iterations < 0) return fail(BSF_ERR_ARG, "bsf: group %d: negative count or iterations", gi);
        if (g.count == 0) continue;
        if (!g.data || !g.P || !g.CRLB || !g.LL)
            return fail(BSF_ERR_ARG, "bsf: group %d: NULL data or output pointer", gi);
        if (g.count > kMaxLaunchRois) return fail(BSF_ERR_ARG, "bsf: group %d: more than %d ROIs", gi, kMaxLaunchRois);
        if ((size_t)g.count * g.sz * g.sz >= ((size_t)1 << 31))
            return fail(BSF_ERR_ARG, "bsf: group %d: %d ROIs of size %d exceed 2^31 pixels", gi, g.count, g.sz);
        const int need = M.ft == 7 ? 2 : M.ft == 8 ? 4 : 0;
        const bool use_initP = g.initP && need > 0;
        if (use_initP && g.n_initP < need)
            return fail(BSF_ERR_ARG, "bsf: group %d: initP for fittype %d needs %d values (got %d)", gi, M.ft, need,
                        g.n_initP);
        const int nst = g.nstarts > 1 ? g.nstarts : 1;
        if (g.nstarts < 0 || nst > kMaxGroupStarts)
            return fail(BSF_ERR_ARG, "bsf: group %d: nstarts must be 0..%d (got %d)", gi, kMaxGroupStarts, g.nstarts);
        if (nst > 1 && !use_initP)
            return fail(BSF_ERR_ARG, "bsf: group %d: several starts need fittype 7 or 8 and initP rows", gi);
        if (M.ft == BSF_FT_4PI && g.chan_stride && g.chan_stride < (size_t)g.count * g.sz * g.sz)
            return fail(BSF_ERR_ARG, "bsf: group %d: channel stride %zu is below count*sz*sz", gi, g.chan_stride);
        Launch key;
        key.model = g.model;
        key.iterations = g.iterations;
        key.has_initP = use_initP ? 1 : 0;
        key.nstarts = nst;
        key.strict = nst > 1 && g.strict ? 1 : 0;
        for (int k = 0; k < nst && use_initP; k++)
            for (int i = 0; i < need; i++) key.initP[k][i] = g.initP[(size_t)k * g.n_initP + i];
        key.has_dT = M.ft == BSF_FT_4PI && g.dT;
        key.has_iz = M.ft == BSF_FT_4PI && g.initZ;
        key.has_ip = M.ft == BSF_FT_4PI && g.initPhi;
        Launch *hit = nullptr;
        for (Launch &l : L)
            if (same_key(l, key)) {
                hit = &l;
                break;
            }
        if (!hit) {
            key.nch = M.ft == BSF_FT_4PI ? M.nch : 1;
            key.nv = M.nv;
            L.push_back(key);
            hit = &L.back();
        }
        if ((long long)hit->n + g.count > kMaxLaunchRois)
            return fail(BSF_ERR_ARG, "bsf: more than %d ROIs of one model in a frame", kMaxLaunchRois);
        hit->groups.push_back(gi);
        hit->n += g.count;
        hit->Smax = std::max(hit->Smax, g.sz);
        hit->Smin = hit->Smin ? std::min(hit->Smin, g.sz) : g.sz;
    }

    if (L.empty()) {   // nothing to fit: complete at once
        *ticket = s->next_ticket++;
        return BSF_OK;
    }
    DeviceGuard dg(s->device);
    if (s->graphs) {
        int rc = submit_graph(s, f, flags, ticket, L, f64);
        if (rc != kNoPlan) return rc;
        s->n_eager++;   // no plan fits the slot: this frame goes the eager way, in a slot of its own
    }
    return submit_eager(s, f, flags, ticket, L, f64);
}

int bsf_poll(bsf_session *s, unsigned long long ticket, int wait)
{
    bsf::clear_error();
    if (!s) return fail(BSF_ERR_ARG, "bsf: NULL session");
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticket == 0 || ticket >= s->next_ticket) return fail(BSF_ERR_ARG, "bsf: unknown ticket %llu", ticket);
    for (Frame &f : s->inflight) {
        if (f.ticket != ticket) continue;
        if (f.delivered) return BSF_OK;
        DeviceGuard g(s->device);
        if (!wait) {
            hipError_t e = hipEventQuery(f.events.back());
            if (e == hipErrorNotReady) return BSF_PENDING;
            if (e != hipSuccess)
                return fail(BSF_ERR_DEVICE, "bsf: frame %llu failed: %s", ticket, hipGetErrorString(e));
        }
        return finish(s, f);
    }
    return BSF_OK;   // delivered and retired, or a frame with no ROIs
}

}  // extern "C"
