"""Serving sessions (bsf_session_* / bsf_submit / bsf_poll) through the C-ABI.

A frame packs the groups of one model into one launch with mixed ROI sizes
and runs a frame's launches on several streams; every ROI is still fitted by
its own wavefront from its own data, so each group's results must equal, bit
for bit, the same ROIs fitted alone by bsf_fit_host / bsf_fit_4pi_host (the
calls whose parity with the oracle the other GPU tests establish).  Also:
HBM-budget admission (BUSY / NOMEM), the arena ring wrapping under many
frames in flight, polling out of order, and argument errors.
"""
import numpy as np
import pytest

import bsfit
from bsfit import gen

pytestmark = pytest.mark.gpu

SIZES = (11, 13, 15, 17)
POOL = 40


def _eq(a, b):
    np.testing.assert_array_equal(a, b)   # NaN == NaN here, bit-level otherwise


@pytest.fixture(scope="module")
def pools():
    """Per model kind and ROI size: a seeded ROI pool and its standalone fits
    (MIX and F64)."""
    c3d = gen.model_coeffs("astig3d")
    c4d = gen.model_coeffs("color4d")
    c5, c53 = gen.model_coeffs("dipole5d")
    c4pi = gen.model_coeffs("4pi")
    out = {}
    for S in SIZES:
        w6 = gen.make_rois(6, POOL, S, 1000 + S, c3d)
        w7 = gen.make_rois(7, POOL, S, 2000 + S, c4d)
        w8 = gen.make_rois(8, POOL, S, 3000 + S, c5, c53)
        w4 = gen.make_rois_4pi(POOL, S, 4000 + S, c4pi, shared=(1, 1, 1, 0, 1, 1))
        for kind, w in (("ft6", w6), ("ft7", w7), ("ft8", w8), ("4pi", w4)):
            ref = {}
            for prec in (bsfit.PREC_MIX, bsfit.PREC_F64):
                if kind == "4pi":
                    ref[prec] = bsfit.fit_4pi(w.imstack, w.shared, w.iterations, w.coeff, w.phaseShift,
                                              initZ=w.initZ, initPhi=w.initPhi, precision=prec)
                else:
                    ref[prec] = bsfit.fit(w.fittype, w.imstack, w.iterations, w.coeff, w.initP, w.coeff3,
                                          precision=prec)
            out[kind, S] = (w, ref)
    return dict(c3d=c3d, c4d=c4d, c5=c5, c53=c53, c4pi=c4pi, pool=out)


def _register(s, pools):
    w4 = pools["pool"]["4pi", 11][0]
    return {
        "ft6": s.register_model(6, pools["c3d"]),
        "ft7": s.register_model(7, pools["c4d"]),
        "ft8": s.register_model(8, pools["c5"], pools["c53"]),
        "4pi": s.register_model(bsfit.FT_4PI, pools["c4pi"], nch=w4.nch, shared=w4.shared,
                                phaseShift=w4.phaseShift),
    }


def _group(models, kind, w, a, b):
    """Group of pool ROIs [a, b) -- slices of the pool, not copies."""
    if kind == "4pi":
        return bsfit.Group(models[kind], w.imstack[:, :, a:b, :], w.iterations,
                           initZ=w.initZ[a:b], initPhi=w.initPhi[a:b])
    return bsfit.Group(models[kind], w.imstack[:, :, a:b], w.iterations, initP=w.initP)


def _check_group(res, ref, a, b):
    P, CR, LL = res
    Pr, CRr, LLr = ref
    _eq(P, Pr[a:b])
    _eq(CR, CRr[a:b])
    _eq(LL, LLr[a:b])


@pytest.mark.parametrize("prec", [bsfit.PREC_MIX, bsfit.PREC_F64])
def test_frame_equals_standalone_fits(pools, prec):
    with bsfit.Session(arena_bytes=64 << 20, precision=prec) as s:
        models = _register(s, pools)
        groups, want = [], []
        for kind in ("ft6", "ft7", "ft8", "4pi"):
            for j, S in enumerate(SIZES):
                w, ref = pools["pool"][kind, S]
                a, b = 3 * j, 3 * j + 17 + 5 * j
                groups.append(_group(models, kind, w, a, b))
                want.append((ref[prec], a, b))
            # a second group of the same model and size in the same frame
            w, ref = pools["pool"][kind, 13]
            groups.append(_group(models, kind, w, 30, 40))
            want.append((ref[prec], 30, 40))
        t = s.submit(groups)
        res = s.poll(t, wait=True)
        assert len(res) == len(groups)
        for r, (ref, a, b) in zip(res, want):
            _check_group(r, ref, a, b)
        print("frame of %d groups (%d ROIs) equals standalone fits, precision %d"
              % (len(groups), sum(b - a for _, a, b in want), prec), flush=True)


def test_admission_busy_and_nomem(pools):
    ref6 = pools["pool"]["ft6", 13]
    w, ref = ref6
    with bsfit.Session(arena_bytes=64 << 20) as big:
        m = big.register_model(6, pools["c3d"])
        t = big.submit([bsfit.Group(m, w.imstack[:, :, 0:20], w.iterations)])
        nfr, frame_bytes = big.inflight()
        assert nfr == 1 and frame_bytes > 0
        _check_group(big.poll(t, wait=True)[0], ref[bsfit.PREC_MIX], 0, 20)
        assert big.inflight() == (0, 0)
    # an arena that holds one such frame but not two
    arena = frame_bytes + frame_bytes // 2
    with bsfit.Session(arena_bytes=arena, hbm_budget=arena + (8 << 20)) as s:
        m = s.register_model(6, pools["c3d"])
        g1 = [bsfit.Group(m, w.imstack[:, :, 0:20], w.iterations)]
        g2 = [bsfit.Group(m, w.imstack[:, :, 20:40], w.iterations)]
        t1 = s.submit(g1)
        # frame 1 keeps its slot until it is delivered: a non-blocking submit is refused
        assert s.submit(g2, block=False) is None
        assert s.inflight()[0] == 1
        # a blocking submit waits for frame 1, delivers it and takes its place
        t2 = s.submit(g2)
        assert t2 > t1
        _check_group(s.poll(t1)[0], ref[bsfit.PREC_MIX], 0, 20)
        _check_group(s.poll(t2, wait=True)[0], ref[bsfit.PREC_MIX], 20, 40)
        # a frame larger than the whole arena can never be admitted
        with pytest.raises(bsfit.BsfError) as e:
            s.submit([bsfit.Group(m, np.concatenate([w.imstack] * 3, axis=2), w.iterations)])
        assert e.value.code == 5
        # models count against the HBM budget
        with pytest.raises(bsfit.BsfError) as e:
            s.register_model(7, np.zeros((30, 30, 64, 1000), np.float32))   # 230 MB
        assert e.value.code == 5


def test_many_frames_ring_wraps_and_out_of_order_polls(pools):
    rng = np.random.default_rng(7)
    kinds = ("ft6", "ft7", "ft8", "4pi")
    with bsfit.Session(arena_bytes=3 << 20, kernel_streams=3) as s:
        models = _register(s, pools)
        pending = {}
        for fno in range(40):
            groups, want = [], []
            for kind in kinds:
                for S in SIZES:
                    if rng.random() < 0.35:
                        continue
                    w, ref = pools["pool"][kind, S]
                    n = int(rng.integers(0, 12))
                    a = int(rng.integers(0, POOL - n + 1))
                    groups.append(_group(models, kind, w, a, a + n))
                    want.append((ref[bsfit.PREC_MIX], a, a + n))
            t = s.submit(groups)
            pending[t] = want
            # collect a random subset now, the rest later
            for tt in list(pending):
                if rng.random() < 0.3:
                    r = s.poll(tt, wait=bool(rng.random() < 0.5))
                    if r is not None:
                        for rr, (ref, a, b) in zip(r, pending.pop(tt)):
                            _check_group(rr, ref, a, b)
        for tt in sorted(pending, key=lambda _: rng.random()):
            r = s.poll(tt, wait=True)
            for rr, (ref, a, b) in zip(r, pending.pop(tt)):
                _check_group(rr, ref, a, b)
        assert s.inflight() == (0, 0)


def test_session_argument_errors(pools):
    w, _ = pools["pool"]["ft6", 11]
    with bsfit.Session(arena_bytes=8 << 20) as s:
        m = s.register_model(6, pools["c3d"])
        with pytest.raises(bsfit.BsfError) as e:
            s.submit([bsfit.Group(m + 5, w.imstack, 10)])
        assert e.value.code == 1
        with pytest.raises(bsfit.BsfError) as e:
            s.register_model(8, pools["c5"])            # fittype 8 needs coeff3
        assert e.value.code == 1
        with pytest.raises(bsfit.BsfError) as e:
            s.register_model(7, pools["c3d"])           # rank mismatch
        assert e.value.code == 4
        with pytest.raises(bsfit.BsfError) as e:
            s.set_precision(3)
        assert e.value.code == 1
        for bad in (-1, 65):
            with pytest.raises(bsfit.BsfError) as e:
                s.set_graphs(bad)
            assert e.value.code == 1
        with pytest.raises(bsfit.BsfError) as e:        # several starts: fittypes 7 and 8 only
            s.submit([bsfit.Group(m, w.imstack, 10, starts=np.zeros((2, 2), np.float32))])
        assert e.value.code == 1
        m7 = s.register_model(7, gen.model_coeffs("color4d"))
        w7 = gen.config(1, N=4)
        with pytest.raises(bsfit.BsfError) as e:        # at most 8 starts per group
            s.submit([bsfit.Group(m7, w7.imstack, 10, starts=np.ones((9, 2), np.float32))])
        assert e.value.code == 1
        t = s.submit([])                                # nothing to fit: complete at once
        assert s.poll(t) == []
        with pytest.raises(bsfit.BsfError) as e:
            s.poll(t + 100)
        assert e.value.code == 1
        s.release_model(m)
        with pytest.raises(bsfit.BsfError) as e:
            s.submit([bsfit.Group(m, w.imstack, 10)])
        assert e.value.code == 1


def test_multistart_groups_equal_standalone_multistart():
    """A session group with several starts equals bsf_fit_multistart_host on
    the same ROIs bit for bit (fittype 7 strict, fittype 8 not), next to a
    single-start group of the same model in the same frame."""
    w7 = gen.config(1, N=90)
    st7 = np.array([[30.5, 1.5], [24.4, 2.5], [36.6, 1.0]], np.float32)
    w8 = gen.config(3, N=60, S=13)[0]
    c5, c3 = gen.model_coeffs("dipole5d")
    st8 = np.array([[10.5, 6.0, 4.5, 0.75], [10.5, 12.0, 4.5, 0.75]], np.float32)
    with bsfit.Session(arena_bytes=32 << 20) as s:
        m7 = s.register_model(7, w7.coeff)
        m8 = s.register_model(8, c5, c3)
        t = s.submit([bsfit.Group(m7, w7.imstack, 50, starts=st7, strict=True),
                      bsfit.Group(m7, w7.imstack, 50, initP=st7[0]),
                      bsfit.Group(m8, w8.imstack, 50, starts=st8)])
        r = s.poll(t, wait=True)
    want = [bsfit.fit_multistart(7, w7.imstack, 50, w7.coeff, st7, strict=True),
            bsfit.fit(7, w7.imstack, 50, w7.coeff, st7[0]),
            bsfit.fit_multistart(8, w8.imstack, 50, c5, st8, coeff3=c3)]
    for got, exp in zip(r, want):
        for a, b in zip(got, exp):
            np.testing.assert_array_equal(a, b)
    assert np.any(r[0][2] != r[1][2])          # the later starts changed some ROIs


def test_sixteen_stream_sessions_create_and_destroy():
    """Sessions with the largest number of kernel streams (16) release every
    stream: many create / destroy cycles, then a session still fits."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    for _ in range(24):
        with bsfit.Session(arena_bytes=8 << 20, kernel_streams=16) as s:
            pass
    c3d = gen.model_coeffs("astig3d")
    w = gen.make_rois(6, 32, 13, 99, c3d)
    ref = bsfit.fit(6, w.imstack, 50, c3d)
    with bsfit.Session(arena_bytes=8 << 20, kernel_streams=16) as s:
        m = s.register_model(6, c3d)
        t = s.submit([bsfit.Group(m, w.imstack, 50)])
        res = s.poll(t, wait=True)
    for a, b in zip(res[0], ref):
        _eq(a, b)
    assert hip.hipDeviceSynchronize() == 0
