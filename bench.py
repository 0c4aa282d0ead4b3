"""Throughput of the MI355X B-spline PSF fitter (one JSON line).

Headline workload (one "step" = one fit launch over the whole batch, inputs
resident in HBM): the vortex-5D orientation fit, fittype 8, 100,000 seeded
ROIs with sizes mixed uniformly in {11, 13, 15, 17} fitted in ONE launch
(bsf_fit_device_ex with a per-ROI size table), a 21x21x21x17x9 cubic
B-spline (coefficients 24x24x24x20x12) blended with a 21^3 averaged model,
50 LM iterations, FP32 evaluation + FP64 solve (MIX).  Synthetic data: ROIs
rendered from the spline models themselves plus Poisson noise
(gen.config(3)).

Beside it, in the same line: the fit kernel's roofline (algorithmic FP32
flops = sum over ROIs of the evaluations each fit ran x F_alg(S), over the
kernel time), per ROI size frac, the same workload host -> host (a serving
session: pack, copy in, one launch, copy out), per-fittype lines in both
precision modes (3D 10k x 13^2, colour-4D 100k x 15^2, 4Pi 200k x 4 x 13^2
with every parameter shared and with a background per channel, vortex-5D
100k mixed) and the CPU oracle on the host cores (1 core and all
allowed cores, up to 16) on a bounded sample of the same ROIs.

--model 3d / colour4d / 4pi / 4pi_bg / vortex5d (and --roi-size) run one of them as
the headline; --workload serve / coeff / frames run the streaming mix, the
coefficient builder and the camera-frame front-end.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gpu-based-b-spline-fitter_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

FP32_PEAK = 157.3e12     # MI355X FP32 vector, FLOP/s
METRIC = "ROI fits/sec per fittype (FP32+FP64-solve vs FP64); p50/p99 frame latency"
TRAFFIC_FILES = {"vortex5d": os.path.join(ROOT, "profiles", "fit_kernel_vortex5d_mixed_traffic.md"),
                 "colour4d": os.path.join(ROOT, "profiles", "fit_kernel_colour4d_traffic.md"),
                 "3d": os.path.join(ROOT, "profiles", "fit_kernel_3d_traffic.md"),
                 "4pi": os.path.join(ROOT, "profiles", "fit4pi_kernel_traffic.md")}
TRAFFIC_FILE = TRAFFIC_FILES["colour4d"]


def f_alg(ft, S, nch=1, NV=None):
    """Algorithmic flops of one ROI-evaluation (factorised contraction):
    2 P^2 T (1+K) + 32 S^2 (3+K) + S^2 (NV(NV+1) + 2 NV + 8), P = S+3,
    T = 4^(D-2), K = D-2; fittype 8 adds the 3D model and the blend."""
    D = {6: 3, 7: 4, 8: 5}[ft]
    NV = NV or {6: 5, 7: 6, 8: 8}[ft]
    P, K = S + 3, D - 2
    T = 4 ** K
    f = nch * (2 * P * P * T * (1 + K) + 32 * S * S * (3 + K) + S * S * (NV * (NV + 1) + 2 * NV + 8))
    if ft == 8:
        f += 2 * P * P * 4 * 2 + 2 * S * S * 16 * 4 + 20 * S * S
    return f


# bench.py --model: configuration of bsfit.gen.config, fittype, label, grid
FIT_MODELS = {
    "3d": (0, 6, "3D astigmatic", "27x27x61 B-spline"),
    "colour4d": (1, 7, "colour-4D", "27x27x61x3 B-spline"),
    "4pi": (2, 104, "4Pi-4D (4 channels)", "27x27x61x18 B-spline, phase offsets 0, pi/2, pi, 3pi/2"),
    "4pi_bg": (2, 104, "4Pi-4D (4 channels, a background per channel: 9 parameters)",
               "27x27x61x18 B-spline, phase offsets 0, pi/2, pi, 3pi/2"),
    "vortex5d": (3, 8, "vortex-5D", "21x21x21x17x9 B-spline + 21^3 averaged model"),
}
DEFAULT_N = {"3d": 10000, "colour4d": 100000, "4pi": 200000, "4pi_bg": 200000, "vortex5d": 100000}


def cpu_threads():
    """Host cores this process may use, capped at 16 (the CPU share of one
    GPU on the boxes this runs on)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(groups, threads, budget_s=10.0):
    """The oracle (C restatement of the reference CPU fitter) on a bounded
    sample of the same ROIs: an equal share of every size group (one oracle
    call per size, as the reference's one-size-per-call MEX), ROI-parallel
    over `threads` host threads.  Returns (fits/s, sample text)."""
    import numpy as np
    import oracle
    n = 8
    while True:
        t0 = time.perf_counter()
        done = 0
        for w in groups:
            k = min(n, w.N)
            sl = np.asfortranarray(w.imstack[:, :, :k])
            oracle.fit(w.fittype, sl, w.iterations, w.coeff, w.initP, w.coeff3, nthreads=threads)
            done += k
        t = time.perf_counter() - t0
        if t > budget_s / 4 or all(n >= w.N for w in groups):
            break
        n *= 4
    total = sum(w.N for w in groups)
    return done / t, "%d of the %d ROIs (%s), %.1f s, oracle/bsf_oracle.c parity build" % (
        done, total, ", ".join("%d x %d^2" % (min(n, w.N), w.S) for w in groups), t)


def traffic_bytes(path=None):
    """HBM bytes per launch of a kernel from its committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy;
    default: the fit kernel's), or None when no summary is present."""
    try:
        for line in open(path or TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def copy_calibration():
    """A plain device copy of 256 MiB: the reference point for scaling the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))


SERVE_SIZES = (11, 13, 15, 17)
SERVE_KINDS = ("ft6", "ft7", "4pi", "ft8")
SERVE_MIX = (0.4, 0.2, 0.2, 0.2)


def serve(args):
    """Streaming workload: frames arrive as a Poisson process (args.rate per
    second, seeded exponential gaps), each with Poisson(args.rois) ROIs drawn
    40/20/20/20 from the 3D / colour-4D / 4Pi / vortex-5D fitters and
    uniformly from ROI sizes 11..17.  Every model stays resident in one
    serving session; a frame is one bsf_submit (one launch per model, mixed
    sizes) and its latency runs from its arrival time to the moment its
    results are in host memory (bsf_poll).  A second, closed-loop phase keeps
    args.depth frames in flight to find the sustainable rate of the mix."""
    import collections
    import numpy as np
    import bsfit
    from bsfit import gen

    seed = gen.SEED0 + 4000
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    c3 = gen.model_coeffs("astig3d")
    c4 = gen.model_coeffs("color4d")
    c4pi = gen.model_coeffs("4pi")
    c5, c53 = gen.model_coeffs("dipole5d")
    pool_n = 768
    pools = {}
    for S in SERVE_SIZES:
        pools["ft6", S] = gen.make_rois(6, pool_n, S, seed + 10 + S, c3)
        pools["ft7", S] = gen.make_rois(7, pool_n, S, seed + 20 + S, c4)
        pools["4pi", S] = gen.make_rois_4pi(pool_n, S, seed + 30 + S, c4pi)
        pools["ft8", S] = gen.make_rois(8, pool_n, S, seed + 40 + S, c5, c53)
    prec = bsfit.PREC_F64 if args.precision == "f64" else bsfit.PREC_MIX
    sess = bsfit.Session(arena_bytes=256 << 20, kernel_streams=args.kernel_streams, precision=prec)
    if args.graphs:
        sess.set_graphs(args.graphs)
    if args.hetero:
        sess.set_hetero(True)
    w4 = pools["4pi", 11]
    models = {"ft6": sess.register_model(6, c3), "ft7": sess.register_model(7, c4),
              "4pi": sess.register_model(bsfit.FT_4PI, c4pi, nch=w4.nch, shared=w4.shared,
                                         phaseShift=w4.phaseShift),
              "ft8": sess.register_model(8, c5, c53)}
    cells = [(k, S) for k in SERVE_KINDS for S in SERVE_SIZES]
    probs = np.array([SERVE_MIX[SERVE_KINDS.index(k)] / len(SERVE_SIZES) for k, S in cells])

    def make_frame():
        n = int(rng.poisson(args.rois))
        groups = []
        for (k, S), c in zip(cells, rng.multinomial(n, probs)):
            c = int(min(c, pool_n))
            if c == 0:
                continue
            w = pools[k, S]
            a = int(rng.integers(0, pool_n - c + 1))
            if k == "4pi":
                groups.append(bsfit.Group(models[k], w.imstack[:, :, a:a + c, :], w.iterations,
                                          initZ=w.initZ[a:a + c], initPhi=w.initPhi[a:a + c]))
            else:
                groups.append(bsfit.Group(models[k], w.imstack[:, :, a:a + c], w.iterations, initP=w.initP))
        return groups, sum(int(np.shape(g.imstack)[2]) for g in groups)

    frames = [make_frame() for _ in range(args.frames)]
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, args.frames))
    print("serve: models, pools and %d frames ready in %.1f s (mean %.0f ROIs/frame)"
          % (args.frames, time.perf_counter() - t0, np.mean([f[1] for f in frames])), file=sys.stderr, flush=True)
    # warm-up with frames in flight, so that every stream (and its hardware
    # queue) and the event pool are set up before the timed phases
    warm = [sess.submit(g) for g, _ in frames[:4 * args.kernel_streams]]
    for t in warm:
        sess.poll(t, wait=True)

    # ---- open loop at the offered rate.  The driver's own garbage collector
    # is frozen for the timed phases: a full pass over the pre-built frames
    # stalls this loop for tens of milliseconds (a pause of the harness, not
    # of the session)
    import gc
    gc.collect()
    gc.freeze()
    gc.disable()
    lat = []
    pending = collections.deque()
    i = 0
    start = time.perf_counter()
    last = start
    while i < args.frames or pending:
        now = time.perf_counter() - start
        if i < args.frames and now >= arrivals[i]:
            pending.append((sess.submit(frames[i][0]), arrivals[i]))
            i += 1
            continue
        for item in list(pending):
            if sess.poll(item[0]) is not None:
                last = time.perf_counter()
                lat.append(last - start - item[1])
                pending.remove(item)
    span = last - start
    nroi = sum(f[1] for f in frames)
    lat_ms = np.array(lat) * 1e3
    slow = np.nonzero(lat_ms > 20.0)[0]
    if len(slow):
        print("serve: %d frames above 20 ms (completion order %d..%d)" % (len(slow), slow[0], slow[-1]),
              file=sys.stderr, flush=True)

    # ---- closed loop: args.depth frames in flight
    pending = collections.deque()
    cstart = time.perf_counter()
    for g, _ in frames:
        if len(pending) >= args.depth:
            sess.poll(pending.popleft(), wait=True)
        pending.append(sess.submit(g))
    while pending:
        sess.poll(pending.popleft(), wait=True)
    cspan = time.perf_counter() - cstart
    gc.enable()
    gc.unfreeze()
    gstats = sess.graph_stats() if args.graphs else None
    sess.close()
    out = {
        "metric": "serving: ROI fits/sec and frame latency, streamed 3D/4D/4Pi/5D mix",
        "value": nroi / span,
        "unit": "fits/s",
        "n_gpus": 1,
        "frames": args.frames,
        "higher_is_better": True,
        "dtype": "f32" if args.precision == "mix" else "f64",
        "data": "synthetic (seeded spline-rendered ROI pools + Poisson noise)",
        "config": {"workload": "frames at %g/s (Poisson arrivals), Poisson(%d) ROIs per frame, fitters "
                               "40/20/20/20 3D/colour-4D/4Pi(4 ch)/vortex-5D, S uniform in {11,13,15,17}, "
                               "50 LM iterations" % (args.rate, args.rois),
                   "kernel_streams": args.kernel_streams, "precision": args.precision,
                   "submission": (("graphs, %d slots" % args.graphs) if args.graphs else "per-call launches")
                   + (", heterogeneous fittype 6-8 launch" if args.hetero else ""),
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
        "offered_fits_per_s": nroi / float(arrivals[-1]),
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p90": float(np.percentile(lat_ms, 90)),
                       "p99": float(np.percentile(lat_ms, 99)), "max": float(lat_ms.max()),
                       "mean": float(lat_ms.mean())},
        "saturated": {"fits_per_s": nroi / cspan, "frames_per_s": args.frames / cspan, "depth": args.depth},
    }
    if gstats:
        out["graphs"] = {"captured": gstats[0], "replayed": gstats[1], "per_call": gstats[2]}
    print(json.dumps(out))


def coeff_bench(args):
    """Cubic B-spline coefficient build of the vortex-5D model (21x21x21x17x9
    samples, lambda 0 on every axis), samples and coefficients resident in
    HBM, one bsf_coeff_build_device per step.  Roofline: HBM, algorithmic
    bytes = samples read + coefficients written (FP32).  CPU baseline: the
    oracle's numpy FP64 restatement of the reference's bsarray."""
    import ctypes as C
    import numpy as np
    import torch
    import bsfit
    from bsfit import gen
    lib = bsfit.load()
    lib.bsf_coeff_work_bytes.argtypes = [C.c_void_p, C.c_int]
    lib.bsf_coeff_work_bytes.restype = C.c_size_t
    lib.bsf_coeff_build_device.argtypes = [C.c_void_p] * 2 + [C.c_int] + [C.c_void_p] * 4
    lib.bsf_coeff_build_device.restype = C.c_int
    x = np.asfortranarray(gen.psf_dipole5d().astype(np.float32))
    dims = np.array(x.shape, np.int32)
    nout = int(np.prod([d + 3 for d in x.shape]))
    dev = torch.device("cuda:0")
    d_x = torch.from_numpy(np.ascontiguousarray(x.ravel(order="F"))).to(dev)
    d_c = torch.empty(nout, dtype=torch.float32, device=dev)
    d_w = torch.empty(lib.bsf_coeff_work_bytes(dims.ctypes.data, x.ndim), dtype=torch.uint8, device=dev)
    st = torch.cuda.current_stream()

    def step():
        rc = lib.bsf_coeff_build_device(d_x.data_ptr(), dims.ctypes.data, x.ndim, None, d_c.data_ptr(),
                                        d_w.data_ptr(), st.cuda_stream)
        if rc:
            raise bsfit.BsfError(rc, lib.bsf_last_error().decode())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record(st)
    for _ in range(args.steps):
        step()
    b.record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ms = a.elapsed_time(b) / args.steps
    alg = x.size * 4 + nout * 4
    cpu = None
    if not args.no_cpu:
        from oracle import bsarray as B
        t1 = time.perf_counter()
        B.bsarray(x.astype(np.float64))
        tc = time.perf_counter() - t1
        cpu = {"value": tc * 1e3, "unit": "ms", "cores": 1, "kind": "port",
               "sample": "the whole build (oracle/bsarray.py, numpy FP64 restatement of bsarray)"}
    print(json.dumps({
        "metric": "cubic B-spline coefficient build time (bsarray, degree 3)", "value": ms, "unit": "ms",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
        "higher_is_better": False, "vs_baseline": None, "dtype": "f64", "data": "synthetic (vortex-5D PSF samples)",
        "config": {"workload": "vortex-5D model, 21x21x21x17x9 samples -> 24x24x24x20x12 coefficients, lambda 0"},
        "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": alg / (ms * 1e-3) / 8e12, "traffic": None,
                     "note": "algorithmic bytes = %d samples + %d coefficients, FP32; the build keeps an FP64 "
                             "complex work array and makes one pass per axis and pole" % (x.size, nout)},
        "cpu_baseline": cpu}))


def frames_bench(args):
    """Frame -> ROI front-end on device-resident camera frames: 32 frames of
    1024x1024 uint16 ADU, 400 emitters each (3D astigmatic model), the GUI's
    DoG mode (sigma 1, 7x7 kernel, first-column minimum subtracted), 13x13
    ROIs; one bsf_find_rois_device call over all frames per step.  Then the
    GUI's whole path on those frames with bsf_fit_frames_device: ROIs found
    and fitted with fittype 6 (50 iterations), the fit reading the ROI count
    on the device (no host round trip).
    Roofline: the detect kernel's FP32 filter arithmetic, 2 mr mc flop per
    pixel (the fused NMS, the halo recomputation and the compaction are not
    counted).  CPU baseline: the oracle's front-end (oracle/bsf_frontend.c)
    on a sample of the frames, one thread."""
    import ctypes as C
    import numpy as np
    import torch
    import bsfit
    from bsfit import frames as FR
    from bsfit import gen
    lib = FR._bind()
    nf, nr, nc, per, roi, cutoff = args.nframes, 1024, 1024, 400, 13, 15.0
    t0 = time.perf_counter()
    coeff = gen.model_coeffs("astig3d")
    frames, truth = gen.make_frames(nf, nr, nc, per, 20261203 + 5000, coeff=coeff)
    print("bench: generated %d frames %dx%d in %.1f s" % (nf, nr, nc, time.perf_counter() - t0), file=sys.stderr,
          flush=True)
    h = FR.peak_filter("dog", 1.0)
    dev = torch.device("cuda:0")
    d_fr = torch.from_numpy(np.ascontiguousarray(frames.ravel(order="F"))).to(dev)
    d_h = torch.from_numpy(np.ascontiguousarray(h.ravel(order="F"))).to(dev)
    cap = nf * 2 * per
    q = FR.make_request(d_fr.data_ptr(), FR.PIXEL_U16, frames.shape, roi, d_h.data_ptr(), h.shape, cutoff, 100.0,
                        0.5, True, False, 0.0, 1, cap)
    wb = lib.bsf_find_rois_work_bytes(C.byref(q))
    d_w = torch.empty(wb, dtype=torch.uint8, device=dev)
    d_st = torch.empty(roi * roi * cap, dtype=torch.float32, device=dev)
    d_pk = torch.empty(3 * cap, dtype=torch.int32, device=dev)
    d_n = torch.zeros(1, dtype=torch.int32, device=dev)
    res = FR.RoiResult(d_st.data_ptr(), d_pk.data_ptr(), None, 0)
    st = torch.cuda.current_stream()

    def detect():
        rc = lib.bsf_find_rois_device(C.byref(q), C.byref(res), d_n.data_ptr(), d_w.data_ptr(), wb, st.cuda_stream)
        if rc:
            raise bsfit.BsfError(rc, lib.bsf_last_error().decode())

    for _ in range(args.warmup):
        detect()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record(st)
    for _ in range(args.steps):
        detect()
    b.record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ms = a.elapsed_time(b) / args.steps
    n = int(d_n.item())
    # the GUI's whole path: frames -> ROIs -> fits (fittype 6, 50 iterations)
    NV = 5
    d_c = torch.from_numpy(np.ascontiguousarray(coeff.ravel(order="F"))).to(dev)

    d_P = torch.empty((NV + 1) * cap, dtype=torch.float32, device=dev)
    d_CR = torch.empty(NV * cap, dtype=torch.float32, device=dev)
    d_LL = torch.empty(cap, dtype=torch.float32, device=dev)
    fq = bsfit.api._request(6, roi, cap, 50, d_c.data_ptr(), coeff.shape, None, None, bsfit.PREC_MIX, None)
    fres = bsfit.api.Result(d_P.data_ptr(), d_CR.data_ptr(), d_LL.data_ptr())

    def end_to_end():
        # frames -> ROIs -> fits queued on the stream; the fit reads the ROI count on the device
        rc = lib.bsf_fit_frames_device(C.byref(q), C.byref(fq), C.byref(res), d_n.data_ptr(), C.byref(fres),
                                       d_w.data_ptr(), wb, st.cuda_stream)
        if rc:
            raise bsfit.BsfError(rc, lib.bsf_last_error().decode())

    end_to_end()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        end_to_end()
    torch.cuda.synchronize()
    e2e = (time.perf_counter() - t0) / args.steps
    flops = 2.0 * h.size * nr * nc * nf
    cpu = None
    if not args.no_cpu:
        import oracle
        k = max(1, min(nf, args.cpu_frames))
        t1 = time.perf_counter()
        co = oracle.find_rois(frames[:, :, :k], roi, h, cutoff=cutoff, offset=100.0, conversion=0.5,
                              subtract_min=True, capacity=k * 2 * per)
        tc = time.perf_counter() - t1
        cpu = {"value": k / tc, "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": "%d of the %d frames (oracle/bsf_frontend.c, %d ROIs), %.1f s" % (k, nf, co[3], tc)}
    print("bench: front-end %.3f ms for %d frames (%d ROIs); frames -> fits %.3f ms" % (ms, nf, n, e2e * 1e3),
          file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "camera frames/sec through the frame -> ROI front-end (1024x1024 uint16)", "value": nf / (ms * 1e-3),
        "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded spline-rendered emitters, Poisson noise, ADU = photons/0.5 + 100)",
        "config": {"workload": "%d frames of %dx%d uint16, %d emitters per frame, DoG sigma 1 (7x7), 13x13 ROIs"
                               % (nf, nr, nc, per), "rois_found": n,
                   "frames_to_fits_ms": e2e * 1e3, "frames_to_fits_per_s": n / e2e},
        "roofline": {"bound": "valu", "achieved": flops / (ms * 1e-3) / 1e12, "peak": FP32_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": flops / (ms * 1e-3) / FP32_PEAK,
                     "traffic": traffic_bytes(os.path.join(ROOT, "profiles", "detect_kernel_frames_traffic.md")),
                     "note": "FP32 vector-ALU bound (filter2: 2 x 49 flop per pixel against 2 B of raw pixel); "
                             "time = the whole front-end call (6 kernels), flops = the filter's only; traffic = "
                             "the detect kernel's HBM bytes per call (rocprofv3 FETCH_SIZE+WRITE_SIZE, "
                             "copy-calibrated; algorithmic 71,303,168 B)"},
        "cpu_baseline": cpu,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # five untimed launches bring the GPU to its steady clock (the first few
    # launches of a process run up to 20 % longer); ten timed ones are 80 ms
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=None, help="fit: ROIs (default: the workload's own count)")
    ap.add_argument("--precision", choices=["mix", "f64"], default="mix")
    ap.add_argument("--model", choices=sorted(FIT_MODELS), default="vortex5d",
                    help="fit: PSF model of the headline line (vortex5d, sizes mixed in one launch, by default)")
    ap.add_argument("--roi-size", type=int, default=None,
                    help="fit: one ROI size (default: the workload's own: 13 for 3d and 4pi, 15 for colour4d, "
                         "11-17 mixed for vortex5d)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="fit: only the headline (no per-size, host-to-host and per-fittype lines)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 256 MiB device copy (for scaling rocprofv3 byte counters)")
    ap.add_argument("--workload", choices=["fit", "serve", "coeff", "frames", "tiles"], default="fit",
                    help="fit: batch fits (default: vortex-5D, mixed sizes); serve: the streamed mix through a session; "
                         "coeff: the B-spline coefficient build of the 5D model; frames: camera frames -> ROIs "
                         "(-> fits); tiles: one vs two ROIs per wavefront (bsf_tune's timings) per fittype 6 / 7, "
                         "precision and ROI size")
    ap.add_argument("--nframes", type=int, default=32, help="frames: camera frames per step")
    ap.add_argument("--cpu-frames", type=int, default=4, help="frames: frames in the CPU baseline sample")
    ap.add_argument("--frames", type=int, default=3000, help="serve: frames")
    ap.add_argument("--rate", type=float, default=1000.0, help="serve: frame arrivals per second")
    ap.add_argument("--rois", type=int, default=2000, help="serve: mean ROIs per frame")
    ap.add_argument("--kernel-streams", type=int, default=16, help="serve: session kernel streams")
    ap.add_argument("--depth", type=int, default=16, help="serve: frames in flight in the closed-loop phase")
    ap.add_argument("--hetero", action="store_true",
                    help="serve: one heterogeneous launch for a frame's fittype 6-8 groups")
    ap.add_argument("--graphs", type=int, default=0, metavar="SLOTS",
                    help="serve: graph mode with SLOTS arena slots (frames replay captured graphs of their "
                         "launches); 0: per-call launches")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: this fitter is measured on one GPU; --gpus %d is not supported "
              "(ROIs are independent: N GPUs would each fit their own share of the ROIs with no exchange, "
              "one process per GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "coeff":
        coeff_bench(args)
        return
    if args.workload == "frames":
        frames_bench(args)
        return
    if args.workload == "tiles":
        tiles_bench(args)
        return
    if args.workload == "serve":
        # the session's kernel streams plus its two copy streams each need a
        # hardware queue of their own to run concurrently; HIP reads this
        # when it initialises (its default is 4)
        need = min(32, args.kernel_streams + 2)
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < need:
            os.environ["GPU_MAX_HW_QUEUES"] = str(need)
        serve(args)
        return
    # the host-to-host figure streams frames through a session with 4 kernel
    # streams and 2 copy streams: one hardware queue each (HIP's default is 4)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 6:
        os.environ["GPU_MAX_HW_QUEUES"] = "6"
    fit_bench(args)


# ------------------------------------------------------------------ tile sweep
def tiles_bench(args):
    """One vs two ROIs per wavefront: bsf_tune on 8,192 seeded ROIs repeated
    16 times (131,072 per launch), 50 iterations, `--steps` timed launches per
    tile; one JSON line per (fittype, precision, ROI size).  The tuned choice
    is reset afterwards, so the process keeps the measured defaults."""
    import numpy as np
    import torch
    import bsfit
    from bsfit import gen
    nu, rep = 8192, 16
    N = nu * rep
    for ft in (6, 7, 8):
        c3 = None
        if ft == 8:
            c, c3 = gen.model_coeffs("dipole5d")
        else:
            c = gen.model_coeffs("astig3d" if ft == 6 else "color4d")
        NV = bsfit.num_params(ft)
        up = lambda a: torch.from_numpy(np.asfortranarray(a, np.float32).ravel(order="F").copy()).cuda()
        dc = up(c)
        dc3 = up(c3) if c3 is not None else None
        sizes = (args.roi_size,) if args.roi_size else ((5, 7, 9, 11, 13, 15, 17, 21) if ft != 8 else (11, 13, 15, 17))
        for S in sizes:
            w = gen.make_rois(ft, nu, S, 5000 + S, c, c3)
            im = np.asfortranarray(w.imstack, np.float32).ravel(order="F")
            data = torch.from_numpy(np.tile(im, rep)).cuda()
            P = torch.empty((NV + 1) * N, device="cuda")
            CR = torch.empty(NV * N, device="cuda")
            LL = torch.empty(N, device="cuda")
            for prec in (bsfit.PREC_MIX, bsfit.PREC_F64):
                default = bsfit.fit_tile(ft, prec, S)
                t, ms = bsfit.tune(ft, data.data_ptr(), S, N, 50, dc.data_ptr(), c.shape, P.data_ptr(),
                                   CR.data_ptr(), LL.data_ptr(), initP=w.initP, reps=args.steps,
                                   coeff3_ptr=dc3.data_ptr() if dc3 is not None else None, precision=prec)
                bsfit.set_fit_tile(ft, prec, S, 0)
                print(json.dumps({"workload": "tiles", "fittype": ft, "precision": "f64" if prec else "mix",
                                  "roi_size": S, "rois": N, "ms_tile1": round(ms[0], 4), "ms_tile2": round(ms[1], 4),
                                  "speedup_tile2": round(ms[0] / ms[1], 3) if ms[1] else None, "tuned": t,
                                  "default": default}), flush=True)


# ------------------------------------------------------------------ fit bench
class FitCase:
    """One fit workload resident on the device, timed launch by launch."""

    def __init__(self, model, N=None, S=None, precision="mix"):
        import numpy as np
        import torch
        import bsfit
        from bsfit import gen
        self.model = model
        cfg, ft, label, grid = FIT_MODELS[model]
        self.ft, self.label, self.grid = ft, label, grid
        N = N or DEFAULT_N[model]
        dev = torch.device("cuda:0")
        up = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F"))).to(dev)
        self.prec = bsfit.PREC_F64 if precision == "f64" else bsfit.PREC_MIX
        self.precision = precision
        if ft == 104:
            # a seeded pool of 25k linked ROIs repeated to N (fits are independent per ROI)
            pool = min(N, 25000)
            w = gen.config(2, N=pool, S=S, bg_per_channel=model == "4pi_bg")
            reps = (N + pool - 1) // pool
            im = np.concatenate([w.imstack] * reps, axis=2)[:, :, :N, :]
            self.groups = [w]
            self.S_list = np.full(N, w.S)
            self.N = N
            self.nch = w.nch
            self.shared = w.shared
            self.NV = bsfit.num_params_4pi(w.shared, w.nch)
            self.d_data = up(im)
            self.d_iz = up(np.tile(w.initZ, reps)[:N])
            self.d_ip = up(np.tile(w.initPhi, reps)[:N])
            self.coeff, self.phaseShift, self.iterations = w.coeff, w.phaseShift, w.iterations
            self.d_coeff = up(w.coeff)
            self.sample = "25,000 seeded linked ROIs repeated to %d" % N if reps > 1 else "%d seeded ROIs" % N
        else:
            ws = gen.config(cfg, N=N, S=S)
            ws = ws if isinstance(ws, list) else [ws]
            ws = sorted(ws, key=lambda w: -w.S)           # largest ROIs first in the launch
            self.groups = ws
            self.S_list = np.concatenate([np.full(w.N, w.S) for w in ws])
            self.N = len(self.S_list)
            self.NV = {6: 5, 7: 6, 8: 8}[ft]
            self.d_data = up(np.concatenate([np.asarray(w.imstack).ravel(order="F") for w in ws]))
            w0 = ws[0]
            self.coeff, self.coeff3, self.initP, self.iterations = w0.coeff, w0.coeff3, w0.initP, w0.iterations
            self.d_coeff = up(w0.coeff)
            self.d_c3 = None if w0.coeff3 is None else up(w0.coeff3)
            self.mixed = len(ws) > 1
            self.d_rois = None
            if self.mixed:
                t = bsfit.roi_table(self.S_list)
                self.d_rois = torch.from_numpy(t.view(np.uint8).copy()).to(dev)
            self.sample = "%d seeded ROIs" % self.N
        self.Smax = int(self.S_list.max())
        self.d_P = torch.empty((self.NV + 1) * self.N, dtype=torch.float32, device=dev)
        self.d_CR = torch.empty(self.NV * self.N, dtype=torch.float32, device=dev)
        self.d_LL = torch.empty(self.N, dtype=torch.float32, device=dev)
        self.d_ev = torch.zeros(self.N, dtype=torch.int32, device=dev)
        self.stream = torch.cuda.current_stream()

    def step(self):
        import bsfit
        if self.ft == 104:
            bsfit.fit_4pi_device(self.d_data.data_ptr(), self.Smax, self.N, self.nch, self.shared, self.iterations,
                                 self.d_coeff.data_ptr(), self.coeff.shape, self.phaseShift, self.d_P.data_ptr(),
                                 self.d_CR.data_ptr(), self.d_LL.data_ptr(), initZ_ptr=self.d_iz.data_ptr(),
                                 initPhi_ptr=self.d_ip.data_ptr(), precision=self.prec,
                                 stream=self.stream.cuda_stream, evals_ptr=self.d_ev.data_ptr())
            return
        bsfit.fit_device(self.ft, self.d_data.data_ptr(), self.Smax, self.N, self.iterations,
                         self.d_coeff.data_ptr(), self.coeff.shape, self.d_P.data_ptr(), self.d_CR.data_ptr(),
                         self.d_LL.data_ptr(), initP=self.initP,
                         coeff3_ptr=None if self.d_c3 is None else self.d_c3.data_ptr(), precision=self.prec,
                         stream=self.stream.cuda_stream,
                         rois_ptr=None if self.d_rois is None else self.d_rois.data_ptr(),
                         evals_ptr=self.d_ev.data_ptr())

    def run(self, steps, warmup):
        """Time `steps` launches after `warmup` untimed ones: per-launch HIP
        events on the launch stream and the wall clock around all of them."""
        import numpy as np
        import torch
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        for a, b in ev:
            a.record(self.stream)
            self.step()
            b.record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        evals = self.d_ev.cpu().numpy().astype(np.float64)
        P = self.d_P.cpu().numpy().reshape(self.NV + 1, self.N)
        nch = self.nch if self.ft == 104 else 1
        ftb = 7 if self.ft == 104 else self.ft
        fl = {int(S): f_alg(ftb, int(S), nch=nch, NV=self.NV) for S in np.unique(self.S_list)}
        flops = float(sum(evals[self.S_list == S].sum() * f for S, f in fl.items()))
        return dict(ms_per_step=wall * 1e3 / steps, kernel_ms=kern_ms, fits_per_s=self.N / (wall / steps),
                    kernel_fits_per_s=self.N / (kern_ms * 1e-3), achieved=flops / (kern_ms * 1e-3),
                    frac=flops / (kern_ms * 1e-3) / FP32_PEAK, mean_evals=float(evals.mean()),
                    mean_iterations=float(P[self.NV].mean()), f_alg=fl)

    def per_size(self, steps, warmup):
        """frac of uniform launches over each size group of this workload."""
        out = {}
        if self.ft == 104 or len(self.groups) < 2:
            return out
        import numpy as np
        import torch
        import bsfit
        for w in self.groups:
            sel = np.nonzero(self.S_list == w.S)[0]
            a, b = int(sel[0]), int(sel[-1]) + 1
            off = int(sum(int(S) * int(S) for S in self.S_list[:a]))
            sub = FitCase.__new__(FitCase)
            sub.__dict__.update(self.__dict__)
            sub.N, sub.S_list, sub.Smax, sub.d_rois, sub.groups = b - a, self.S_list[a:b], int(w.S), None, [w]
            sub.d_data = self.d_data[off:off + (b - a) * w.S * w.S]
            sub.d_P = torch.empty((self.NV + 1) * sub.N, dtype=torch.float32, device=self.d_P.device)
            sub.d_CR = torch.empty(self.NV * sub.N, dtype=torch.float32, device=self.d_P.device)
            sub.d_LL = torch.empty(sub.N, dtype=torch.float32, device=self.d_P.device)
            sub.d_ev = torch.zeros(sub.N, dtype=torch.int32, device=self.d_P.device)
            r = sub.run(steps, warmup)
            out[str(w.S)] = {"rois": sub.N, "kernel_fits_per_s": r["kernel_fits_per_s"], "frac": r["frac"],
                             "f_alg": r["f_alg"][int(w.S)], "mean_evals": r["mean_evals"]}
        return out

    def host_to_host(self, steps, frames=8, kernel_streams=4):
        """The same ROIs host -> host through a serving session, pipelined as a
        streaming caller would: `frames` frames, each a share of every ROI
        size group (pack into the frame's pinned slot, one copy in, one launch
        of mixed sizes, one copy out, results scattered to the caller's
        arrays), all submitted before the first poll, so one frame's packing
        and copies overlap the launches of the others on `kernel_streams`
        streams (with the session's two copy streams, one hardware queue
        each: fit_bench raises HIP's default of four to six)."""
        import bsfit
        if self.ft == 104:
            return None
        need = sum(w.imstack.nbytes for w in self.groups) + self.N * (2 * self.NV + 2) * 4 + self.N * 16
        parts = []
        for i in range(frames):
            fr = []
            for w in self.groups:
                a, b = i * w.N // frames, (i + 1) * w.N // frames
                fr.append((w.imstack[:, :, a:b], w))
            parts.append(fr)
        with bsfit.Session(arena_bytes=int(need * 1.5) + (64 << 20), kernel_streams=kernel_streams,
                           precision=self.prec) as s:
            m = s.register_model(self.ft, self.coeff, self.coeff3)

            def step():
                tickets = [s.submit([bsfit.Group(m, im, w.iterations, initP=w.initP) for im, w in fr])
                           for fr in parts]
                for t in tickets:
                    s.poll(t, wait=True)
            for _ in range(2):
                step()
            t0 = time.perf_counter()
            for _ in range(steps):
                step()
            dt = (time.perf_counter() - t0) / steps
        return {"fits_per_s": self.N / dt, "ms_per_step": dt * 1e3, "steps": steps, "frames": frames,
                "kernel_streams": kernel_streams,
                "path": "bsf_submit x %d frames, then bsf_poll each: host arrays -> pinned arena -> H2D -> one "
                        "launch of mixed sizes per frame -> D2H -> host arrays" % frames}


def fit_bench(args):
    import torch
    model = args.model
    t0 = time.perf_counter()
    head = FitCase(model, N=args.n, S=args.roi_size, precision=args.precision)
    print("bench: %s workload ready (%d ROIs, sizes %s) in %.1f s" % (
        model, head.N, sorted(set(int(S) for S in head.S_list)), time.perf_counter() - t0), file=sys.stderr, flush=True)
    r = head.run(args.steps, args.warmup)
    print("bench: %.2f ms/step (kernel %.2f ms), mean iterations %.2f, mean evaluations %.2f, frac %.3f"
          % (r["ms_per_step"], r["kernel_ms"], r["mean_iterations"], r["mean_evals"], r["frac"]),
          file=sys.stderr, flush=True)
    extra = {}
    if not args.no_extras:
        extra["per_size"] = head.per_size(3, 2)
        h2h = head.host_to_host(3)
        if h2h:
            extra["host_to_host"] = h2h
        per = {}
        for m in ("3d", "colour4d", "4pi", "4pi_bg", "vortex5d"):
            per[m] = {}
            for prec in ("mix", "f64"):
                if m == model and prec == args.precision and args.n is None and args.roi_size is None:
                    rr = r
                    groups = head.groups
                else:
                    c = FitCase(m, precision=prec)
                    rr = c.run(3, 2)
                    groups = c.groups
                    del c
                    torch.cuda.empty_cache()
                per[m][prec] = {"fits_per_s": rr["fits_per_s"], "kernel_fits_per_s": rr["kernel_fits_per_s"],
                                "frac_fp32": rr["frac"], "mean_evals": rr["mean_evals"]}
                # the reference CPU fitter (restated by the oracle) on a bounded
                # sample of the same ROIs, for the 3D and colour-4D
                # configurations (the 4Pi fitter has no CPU path in the
                # reference; vortex-5D is the line's own cpu_baseline)
                if prec == "mix" and m in ("3d", "colour4d") and not args.no_cpu:
                    th = cpu_threads()
                    v_all, samp = cpu_baseline(groups, th, budget_s=4.0)
                    v_one, _ = cpu_baseline(groups, 1, budget_s=2.0)
                    per[m]["cpu"] = {"fits_per_s": v_all, "cores": th, "fits_per_s_1core": v_one, "kind": "port",
                                     "sample": samp}
            print("bench: %s MIX %.3g fits/s (frac %.3f), F64 %.3g fits/s" % (
                m, per[m]["mix"]["fits_per_s"], per[m]["mix"]["frac_fp32"], per[m]["f64"]["fits_per_s"]),
                file=sys.stderr, flush=True)
        extra["per_fittype"] = per
    cpu = None
    if not args.no_cpu and head.ft != 104:
        th = cpu_threads()
        v_all, samp_all = cpu_baseline(head.groups, th)
        v_one, samp_one = cpu_baseline(head.groups, 1)
        cpu = {"value": v_all, "unit": "fits/s", "cores": th, "kind": "port", "sample": samp_all,
               "value_1core": v_one, "sample_1core": samp_one}
    sizes = sorted(set(int(S) for S in head.S_list))
    size_txt = ("sizes mixed uniformly in {%s}, one launch" % ",".join(map(str, sizes))) if len(sizes) > 1 else \
        "%dx%d" % (sizes[0], sizes[0])
    nk = "%gk" % (head.N / 1000) if head.N % 1000 == 0 else str(head.N)
    traffic = traffic_bytes(TRAFFIC_FILES.get(model)) if (args.n is None and args.roi_size is None
                                                          and args.precision == "mix") else None
    out = {
        "metric": METRIC,
        "value": r["fits_per_s"],
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "mix" else "f64",
        "data": "synthetic (seeded spline-rendered ROIs + Poisson noise; %s)" % head.sample,
        "config": {"workload": "%s fit (fittype %s): %s ROIs, %s, %s, %d LM iterations, %s"
                               % (head.label, "4Pi" if head.ft == 104 else head.ft, nk, size_txt, head.grid,
                                  head.iterations,
                                  "FP32 evaluation + FP64 solve" if args.precision == "mix" else "all FP64"),
                   "fittype": head.ft, "rois": head.N, "roi_sizes": sizes, "iterations": head.iterations,
                   "precision": args.precision, "kernel_fits_per_s": r["kernel_fits_per_s"]},
        "roofline": {"bound": "valu", "achieved": r["achieved"] / 1e12, "peak": FP32_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": r["frac"], "traffic": traffic,
                     "note": "FP32 vector-ALU bound (no MFMA: not a dense contraction); peak = MI355X FP32 "
                             "157.3 TFLOP/s (counts packed FP32, which this build does not issue); achieved = "
                             "sum over ROIs of the evaluations each fit ran (kernel counter, mean %.2f) x "
                             "F_alg(S) (%s flop per ROI-evaluation) / kernel time (HIP events on the launch "
                             "stream); traffic = bytes per launch beyond L2 (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                             "copy-calibrated)%s" % (r["mean_evals"], ", ".join(
                                 "S=%d: %d" % (S, f) for S, f in sorted(r["f_alg"].items())),
                                 "; for vortex-5D these are the coefficient gathers that miss an XCD's 4 MB L2 "
                                 "(the 13.3 MB 5D array is re-gathered every evaluation), served by the Infinity "
                                 "Cache or HBM, against 93.6 MB of algorithmic bytes (ROI data, outputs, the "
                                 "coefficient arrays once)" if model == "vortex5d" else "")},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
