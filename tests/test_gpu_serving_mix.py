"""The streaming serving mix (30 frames of the stated load) through one
session: each frame holds Poisson(2,000) ROIs drawn 40/20/20/20 from the 3D,
colour-4D, 4Pi (4 channels) and vortex-5D models with sizes uniform in
{11, 13, 15, 17}; frames are submitted without blocking, several in flight,
and polled in arrival order.

Asserted, with per-call launches, in graph mode (frames replaying
captured graphs of their launches) and with one heterogeneous launch for a
frame's fittype 6-8 groups: admission refuses nothing (the arena
holds the frames in flight at this load), and every group's results equal
the same ROIs fitted alone
(bsf_fit_host / bsf_fit_4pi_host, whose parity with the oracle the other GPU
tests establish) bit for bit -- each ROI is fitted by its own wavefront from
its own data whichever launch, stream and neighbours it gets."""
import collections

import numpy as np
import pytest

import bsfit
from bsfit import gen

pytestmark = pytest.mark.gpu

SIZES = (11, 13, 15, 17)
KINDS = ("ft6", "ft7", "4pi", "ft8")
MIX = (0.4, 0.2, 0.2, 0.2)
POOL = 400
FRAMES = 30
DEPTH = 6


@pytest.fixture(scope="module")
def serving_pools():
    seed = gen.SEED0 + 4100
    c3 = gen.model_coeffs("astig3d")
    c4 = gen.model_coeffs("color4d")
    c4pi = gen.model_coeffs("4pi")
    c5, c53 = gen.model_coeffs("dipole5d")
    pools = {}
    for S in SIZES:
        pools["ft6", S] = gen.make_rois(6, POOL, S, seed + 10 + S, c3)
        pools["ft7", S] = gen.make_rois(7, POOL, S, seed + 20 + S, c4)
        pools["4pi", S] = gen.make_rois_4pi(POOL, S, seed + 30 + S, c4pi)
        pools["ft8", S] = gen.make_rois(8, POOL, S, seed + 40 + S, c5, c53)
    ref = {}
    for (k, S), w in pools.items():
        if k == "4pi":
            ref[k, S] = bsfit.fit_4pi(w.imstack, w.shared, w.iterations, w.coeff, w.phaseShift,
                                      initZ=w.initZ, initPhi=w.initPhi)
        else:
            ref[k, S] = bsfit.fit(w.fittype, w.imstack, w.iterations, w.coeff, w.initP, w.coeff3)
    return dict(pools=pools, ref=ref, coeffs=(c3, c4, c4pi, c5, c53))


@pytest.mark.parametrize("graphs,hetero", [(False, False), (True, False), (False, True)],
                         ids=["per-call", "graphs", "hetero"])
def test_streaming_mix_equals_standalone_fits(serving_pools, graphs, hetero):
    pools, ref = serving_pools["pools"], serving_pools["ref"]
    c3, c4, c4pi, c5, c53 = serving_pools["coeffs"]
    rng = np.random.default_rng(2026)
    cells = [(k, S) for k in KINDS for S in SIZES]
    probs = np.array([MIX[KINDS.index(k)] / len(SIZES) for k, S in cells])
    frames = []
    for _ in range(FRAMES):
        n = int(rng.poisson(2000))
        groups, want = [], []
        for (k, S), c in zip(cells, rng.multinomial(n, probs)):
            c = int(min(c, POOL))
            if c == 0:
                continue
            a = int(rng.integers(0, POOL - c + 1))
            groups.append((k, S, a, a + c))
        frames.append(groups)
    nroi = sum(b - a for g in frames for _, _, a, b in g)
    print("case serving mix (%s): %d frames, %d ROIs (mean %.0f per frame)"
          % ("graphs" if graphs else "hetero" if hetero else "per-call", FRAMES, nroi, nroi / FRAMES), flush=True)
    w4 = pools["4pi", 11]
    with bsfit.Session(arena_bytes=256 << 20 if graphs else 96 << 20, kernel_streams=8) as s:
        if graphs:
            s.set_graphs(32)
        if hetero:
            s.set_hetero(True)
        models = {"ft6": s.register_model(6, c3), "ft7": s.register_model(7, c4),
                  "4pi": s.register_model(bsfit.FT_4PI, c4pi, nch=w4.nch, shared=w4.shared,
                                          phaseShift=w4.phaseShift),
                  "ft8": s.register_model(8, c5, c53)}

        def groups_of(spec):
            out = []
            for k, S, a, b in spec:
                w = pools[k, S]
                if k == "4pi":
                    out.append(bsfit.Group(models[k], w.imstack[:, :, a:b, :], w.iterations,
                                           initZ=w.initZ[a:b], initPhi=w.initPhi[a:b]))
                else:
                    out.append(bsfit.Group(models[k], w.imstack[:, :, a:b], w.iterations, initP=w.initP))
            return out

        pending = collections.deque()
        checked = 0
        for spec in frames:
            if len(pending) >= DEPTH:
                t, sp = pending.popleft()
                checked += _check_frame(s.poll(t, wait=True), sp, ref)
            t = s.submit(groups_of(spec), block=False)
            assert t is not None, "admission refused a frame with %d frames in flight" % len(pending)
            pending.append((t, spec))
        while pending:
            t, sp = pending.popleft()
            checked += _check_frame(s.poll(t, wait=True), sp, ref)
        assert s.inflight() == (0, 0)
        if graphs:
            cap, rep, eager = s.graph_stats()
            print("graphs captured %d, frames replayed %d, per-call %d" % (cap, rep, eager), flush=True)
            assert rep == FRAMES and eager == 0 and 1 <= cap <= 2 * 32
    assert checked == nroi
    print("all %d ROIs of %d frames equal their standalone fits" % (checked, FRAMES), flush=True)


def _check_frame(res, spec, ref):
    assert len(res) == len(spec)
    n = 0
    for (P, CR, LL), (k, S, a, b) in zip(res, spec):
        Pr, CRr, LLr = ref[k, S]
        np.testing.assert_array_equal(P, Pr[a:b])
        np.testing.assert_array_equal(CR, CRr[a:b])
        np.testing.assert_array_equal(LL, LLr[a:b])
        n += b - a
    return n


def test_graph_plans_grow_and_follow_the_frame(serving_pools):
    """Graph mode: a frame with three times the ROIs of the first outgrows
    its plan (a new plan, captured anew), a frame with other models gets a
    plan of its own, and a frame larger than a graph slot is launched call by
    call; every result equals the standalone fits."""
    pools, ref = serving_pools["pools"], serving_pools["ref"]
    c3, c4, c4pi, c5, c53 = serving_pools["coeffs"]
    with bsfit.Session(arena_bytes=24 << 20, kernel_streams=4) as s:
        s.set_graphs(8)
        m6, m8 = s.register_model(6, c3), s.register_model(8, c5, c53)
        m7 = s.register_model(7, c4)

        def frame(spec):
            out = []
            for m, k, S, a, b in spec:
                w = pools[k, S]
                out.append(bsfit.Group(m, w.imstack[:, :, a:b], w.iterations, initP=w.initP))
            return out

        specs = [
            [(m6, "ft6", 13, 0, 40), (m8, "ft8", 11, 0, 30)],                        # plan 1
            [(m6, "ft6", 11, 0, 50), (m8, "ft8", 11, 10, 40)],                       # replay
            [(m6, "ft6", 17, 0, 200), (m8, "ft8", 15, 0, 150)],                      # outgrows: plan 1'
            [(m7, "ft7", 15, 0, 60)],                                                # plan 2
            [(m6, "ft6", 11, 100, 160), (m8, "ft8", 11, 50, 70)],                    # replay 1'
            [(m6, "ft6", S, 0, POOL) for S in SIZES] + [(m8, "ft8", S, 0, POOL) for S in SIZES]
            + [(m7, "ft7", S, 0, POOL) for S in SIZES],                              # beyond a slot
        ]
        for i, spec in enumerate(specs):
            print("case graph plan frame %d: %d ROIs" % (i, sum(b - a for _, _, _, a, b in spec)), flush=True)
            res = s.poll(s.submit(frame(spec)), wait=True)
            _check_frame(res, [(k, S, a, b) for _, k, S, a, b in spec], ref)
        cap, rep, eager = s.graph_stats()
        print("graphs captured %d, frames replayed %d, per-call %d" % (cap, rep, eager), flush=True)
        assert (cap, rep, eager) == (3, 5, 1)
