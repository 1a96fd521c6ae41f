"""Benchmark of the phase-retrieval hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-iters I]
                    [--no-cpu] [--no-batched] [--no-serving] [--requests R]

Headline workload (single reconstruction, 1024 x 1024 complex64): synthetic
frame (random star polygon, ~15 % fill, amplitude U[0.5, 1], phase
U[-pi, pi], Poisson 1e8 photons, support dilated 3 px), 1000 iterations
cycling 90 HIO / 10 ER (beta 0.8, numpy semantics), in-loop shrinkwrap after
every 20 iterations (sigma 3 -> 1.5 px, tau 0.2), FP64 modulus-error trace,
ROI copy into the savebuffer.  One step = one whole reconstruction (seed from
a device-resident guess, reset support, 1000 iterations, ROI copy); inputs
are resident in HBM before the timed region.  value = iterations per second.

Reported beside it in the same JSON line:
  roofline      dominant kernel of the headline workload: algorithmic bytes
                per launch / average launch time (HIP events around each
                launch on the library stream, one profiled reconstruction)
  single_sizes  one 512^2 and one 256^2 frame with the headline schedule:
                iterations/s of the resident launch (tile 4, the default,
                and tile 2) and of one launch per pass
  batched       64 trials of one 512^2 frame (RAAR beta 0.87 and DM beta 0.7,
                500 iterations each, aligned and averaged on the GPU), and 16
                trials of the headline frame and schedule in one plan
                (recons/s and the ratio to the single-reconstruction form),
                and uneven frame sizes (64 x 256^2, 16 x 512^2, 4 x 1024^2)
                in one merged launch per pass against per-plan launches
  serving       Poisson arrivals of mixed sizes {256: 40 %, 512: 30 %,
                1024: 20 %, 2048: 10 %}, 500 HIO/ER iterations with shrinkwrap
                each, continuous batching under an HBM budget: p50/p99
                latency and recons/s
  large_4096    one 4096^2 reconstruction (polygon ~8 %, amplitude U[0.3, 1],
                Poisson 1e9; 2000 iterations 90 HIO / 10 ER with shrinkwrap
                every 20): iterations/s, per-pass times and HBM fractions
  cpu_baseline  the restated reference numpy path on the host cores
                (bounded sample of the headline frame)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
B_COL, B_ROW = 20, 36      # algorithmic bytes per pixel of the column / row pass
B_SHRINK, B_ONEOFF = 32, 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-iters", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the partial-coherence / spectrum / PRTF / RFTF / RL timings")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--budget-gb", type=float, default=48.0)
    ap.add_argument("--kernel-probe", action="store_true",
                    help="only a 256 MiB device copy (HBM counter calibration) and the "
                         "iterations of one workload (--probe-frame), for rocprofv3 --pmc runs")
    ap.add_argument("--probe-frame", default="single_1024",
                    choices=("single_1024", "large_4096", "trials_512", "trials_1024"),
                    help="the workload of --kernel-probe: one 1024^2 or 4096^2 frame (headline "
                         "schedule), 64 trials of the 512^2 frame (RAAR beta 0.87), or 16 "
                         "trials of the 1024^2 frame (headline schedule)")
    return ap.parse_args()


def kernel_probe(ctx, which="single_1024"):
    """A copy of known size, then iterations of one workload: the whole
    1000-iteration headline schedule on the 1024^2 headline frame (one
    resident launch, the launch the single-frame bench times), 20 iterations
    of that schedule on the 4096^2 large frame or on 16 trials of the 1024^2
    frame, or 20 RAAR (beta 0.87) iterations on 64 trials of the 512^2 frame
    sharing one modulus and support."""
    import torch
    from als_speckle_amd import _capi
    from als_speckle_amd.device import Plan
    from als_speckle_amd import synth
    nbytes = 256 << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=ctx.device)
    dst = torch.empty_like(src)
    for _ in range(3):
        ctx.copy_bytes(dst, src)
    frame = synth.config_frame("single_1024" if which == "trials_1024" else which, 0)
    n = frame.modulus.shape[0]
    batch = {"trials_512": 64, "trials_1024": 16}.get(which, 1)
    shared = which != "trials_1024"
    plan = Plan(ctx, n, batch=batch, shared=shared)
    m = torch.from_numpy(frame.modulus).to(ctx.device)
    s = torch.from_numpy(frame.support).to(ctx.device)
    plan.set_modulus(m if shared else m.expand(batch, n, n).contiguous())
    plan.set_support(s if shared else s.expand(batch, n, n).contiguous())
    g = np.empty((batch, n, n), np.complex64)
    for t in range(batch):
        np.random.seed(1 + t)
        g[t] = (np.random.rand(n, n) + 1j * np.random.rand(n, n)).astype(np.complex64)
    plan.set_psi(torch.from_numpy(g if batch > 1 else g[0]).to(ctx.device))
    iters = 1000 if which == "single_1024" else 20
    if which == "trials_512":
        steps = [(_capi.RAAR, _capi.SEM_NUMPY, 0.87, 0.0, 0.0)] * iters
    else:
        steps, _ = headline_steps(iters)
    plan.iterate(steps)
    ctx.sync()
    print(json.dumps({"kernel_probe": True, "workload": which, "copy_bytes": nbytes, "copies": 3,
                      "iterations": iters, "n": n, "frames": batch,
                      "resident_launches": plan.resident_counts()[0]}), flush=True)


KERNEL_NAMES = {"column": "colpass", "row": "rowpass", "resident": "resident"}


def traffic_file(dom):
    """The committed PMC summary of the dominant kernel ("column" pass, "row"
    pass or "resident" launch) of the 1024^2 headline frame: HBM bytes per
    dispatch from FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel
    calibration."""
    return os.path.join(ROOT, "profiles", f"traffic_{KERNEL_NAMES[dom]}_1024.json")


def measured_traffic(dom):
    try:
        with open(traffic_file(dom)) as fh:
            return json.load(fh).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def oracle_final_errors(workload, run, key="final_error"):
    """The complex128 oracle's final modulus errors over its 8 random starts
    (tests/golden/final_error_<workload>.json, written offline by
    tests/golden/make_fixtures.py; read here as data)."""
    with open(os.path.join(ROOT, "tests", "golden", f"final_error_{workload}.json")) as fh:
        d = json.load(fh)["runs"][run]
    return np.array([r[key] for r in d["runs"]]), d["seeds"]


def compare_final(gpu, ref, label):
    """Distribution of GPU final errors against the oracle's (chaotic
    iterations: single trajectories part, error levels must agree)."""
    gpu = np.asarray(gpu, np.float64)
    med, omed = float(np.median(gpu)), float(np.median(ref))
    return {"what": label, "gpu_runs": int(gpu.size), "oracle_runs": int(ref.size),
            "gpu_min": float(gpu.min()), "gpu_median": med, "gpu_max": float(gpu.max()),
            "oracle_min": float(ref.min()), "oracle_median": omed, "oracle_max": float(ref.max()),
            "median_rel_diff": round(abs(med - omed) / omed, 5),
            "gpu_median_within_oracle_range": bool(ref.min() <= med <= ref.max())}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def headline_steps(iters):
    from als_speckle_amd import _capi
    from als_speckle_amd.phasing import expand_order, shrinkwrap_schedule
    algs = expand_order((("hio", 90), ("er", 10)), iters)
    sw = shrinkwrap_schedule(iters, every=20, sigma0=3.0, sigma1=1.5, tau=0.2)
    return [(_capi.ALG_CODES[x], _capi.SEM_NUMPY, 0.8, *sw.get(k, (0.0, 0.0)))
            for k, x in enumerate(algs)], len(sw)


def bench_next_rows(ctx):
    """Throughput of the paths around the loop, on the context's stream with
    HIP events: HIO with a partial-coherence ipsf and with a spectrum on the
    1024^2 headline frame, PRTF and stack RFTF of 64 estimates of 512^2, and
    Richardson-Lucy coherence iterations on the 1024^2 frame."""
    import torch
    from als_speckle_amd import _capi, synth
    from als_speckle_amd.device import Plan
    dev = ctx.device

    def timed(fn, reps):
        fn()
        ctx.sync()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(ctx.stream)
        for _ in range(reps):
            fn()
        e1.record(ctx.stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps          # seconds per call

    out = {}
    n = 1024
    frame = synth.config_frame("single_1024", 0)
    x = np.fft.fftfreq(n) * n
    psf = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (2 * 2.0 ** 2))
    ipsf = torch.from_numpy(np.fft.fft2(psf / psf.sum()).astype(np.complex64)).to(dev)
    rng = np.random.default_rng(3)
    g = (rng.random((n, n)) + 1j * rng.random((n, n))).astype(np.complex64)
    hio = [(_capi.HIO, _capi.SEM_NUMPY, 0.8, 0.0, 0.0)] * 100
    for name in ("ipsf", "spectrum"):
        p = Plan(ctx, n)
        p.set_modulus(torch.from_numpy(frame.modulus).to(dev))
        p.set_support(torch.from_numpy(frame.support).to(dev))
        p.set_psi(torch.from_numpy(g).to(dev))
        if name == "ipsf":
            p.set_ipsf(ipsf)
            model = 104 * n * n               # bytes per iteration (column A, row conv, column B, row pass)
        else:
            e = np.linspace(0.98, 1.02, 5)
            w = np.exp(-((e - 1.0) / 0.01) ** 2)
            p.set_spectrum((e / 1.0).astype(np.float32), (w / w.sum()).astype(np.float32))
            model = None
        dt = timed(lambda: p.iterate(hio, error_trace=False), 3) / len(hio)
        out[f"hio_{name}_1024"] = {"iterations_per_s": round(1.0 / dt, 1),
                                   "us_per_iteration": round(dt * 1e6, 2)}
        if model:
            out[f"hio_{name}_1024"]["model_hbm_frac"] = round(model / dt / HBM_PEAK, 4)
        p.close()
    fr = synth.config_frame("trials_512", 0)
    stack = torch.from_numpy((rng.random((64, 512, 512)) *
                              np.exp(2j * np.pi * rng.random((64, 512, 512)))).astype(np.complex64)).to(dev)
    goal = torch.from_numpy(fr.modulus.astype(np.float32)).to(dev)
    q = Plan(ctx, 512)
    dt = timed(lambda: q.prtf(stack), 5)
    out["prtf_64x512"] = {"ms_per_call": round(dt * 1e3, 3), "frames_per_s": round(64 / dt, 1)}
    dt = timed(lambda: q.rftf_stack(stack, goal), 5)
    out["rftf_stack_64x512"] = {"ms_per_call": round(dt * 1e3, 3), "frames_per_s": round(64 / dt, 1)}
    q.close()
    r = Plan(ctx, n)
    est = torch.from_numpy(g[:256, :256].copy()).to(dev)
    mod = torch.from_numpy(frame.modulus).to(dev)
    dt = timed(lambda: r.coherence_rl(est, mod, 20), 3) / 20
    out["richardson_lucy_1024"] = {"iterations_per_s": round(1.0 / dt, 1),
                                   "us_per_iteration": round(dt * 1e6, 2)}
    r.close()
    return out


def copy_bandwidth(ctx, nbytes=1 << 30, reps=10):
    """A plain device copy of 1 GiB (k_copy16), read + write bytes per second:
    what a streaming kernel reaches on this box, beside the 8 TB/s peak."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=ctx.device)
    dst = torch.empty_like(src)
    ctx.copy_bytes(dst, src)
    ctx.sync()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(ctx.stream)
    for _ in range(reps):
        ctx.copy_bytes(dst, src)
    e1.record(ctx.stream)
    e1.synchronize()
    dt = e0.elapsed_time(e1) / 1e3 / reps
    assert torch.equal(dst[:1024], src[:1024]) and torch.equal(dst[-1024:], src[-1024:])
    return {"bytes": nbytes, "tb_per_s": round(2 * nbytes / dt / 1e12, 3),
            "frac_of_peak": round(2 * nbytes / dt / HBM_PEAK, 4)}


def cpu_baseline(frame, iters):
    """Restated reference numpy path (oracle/ref_numpy.py) on a bounded sample."""
    from oracle.ref_numpy import RefNumpyPhasing
    from als_speckle_amd.phasing import expand_order
    r = RefNumpyPhasing(frame.modulus, frame.support, beta=0.8)
    r.seed(1)
    algs = expand_order((("hio", 90), ("er", 10)), iters)
    r.iteration("hio")  # warm-up
    t0 = time.perf_counter()
    r.iterate(algs)
    dt = time.perf_counter() - t0
    cores = len(os.sched_getaffinity(0))
    # the complex128 oracle on the same frame, a shorter sample
    from oracle import phasing_c128 as O
    psi = r.psi_in.astype(np.complex128)
    m, s = frame.modulus.astype(np.float64), r.support.astype(np.float64)
    n_or = max(5, iters // 15)
    t0 = time.perf_counter()
    for a in algs[:n_or]:
        psi, _, _ = O.iteration(psi, m, s, a, beta=0.8)
    dt_or = time.perf_counter() - t0
    return {"value": round(iters / dt, 3), "unit": "iterations/s", "cores": 1, "kind": "port",
            "sample": f"{iters} iterations (90 HIO/10 ER, beta 0.8) of the 1024x1024 frame, "
                      f"numpy {np.__version__} complex64 path with the reference's redundant "
                      f"FFT per _fft2 call, single-threaded ({cores} host cores visible)",
            "oracle_c128": {"value": round(n_or / dt_or, 3), "unit": "iterations/s", "cores": 1,
                            "sample": f"{n_or} iterations of the complex128 oracle (with its "
                                      f"FP64 modulus error) on the same frame"}}


def bench_single(a):
    """Single 1024^2 reconstructions on a dedicated stream, eager and as a
    replayed HIP graph of the whole reconstruction."""
    import torch
    from als_speckle_amd.device import Context, Plan
    from als_speckle_amd.masking import bounding_box
    from als_speckle_amd import _capi, synth

    n, iters = 1024, 1000
    frame = synth.config_frame("single_1024", 0)
    stream = torch.cuda.Stream()
    ctx = Context(0, stream=stream)
    dev = ctx.device
    m_d = torch.from_numpy(frame.modulus).to(dev)
    s_d = torch.from_numpy(frame.support).to(dev)
    np.random.seed(1)
    guess = (np.random.rand(n, n) + 1j * np.random.rand(n, n)).astype(np.complex64)
    g_d = torch.from_numpy(guess).to(dev)
    r0, r1, c0, c1 = (int(v) for v in bounding_box(frame.support, pad=10))
    savebuf = torch.zeros((a.steps + a.warmup + 2, r1 - r0, c1 - c0), dtype=torch.complex64,
                          device=dev)
    err = torch.empty((iters, 1), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    steps, nsw = headline_steps(iters)
    # the library's default for one 1024^2 frame (resident launches), and the
    # one-launch-per-pass form, also replayed as a captured HIP graph (graph
    # capture runs the per-pass form)
    plan = Plan(ctx, n, batch=1, shared=True)
    plan_pp = Plan(ctx, n, batch=1, shared=True, resident=2)
    for p in (plan, plan_pp):
        p.set_modulus(m_d)

    def reconstruction(slot, p=plan):
        p.set_support(s_d)
        p.set_psi(g_d, mult_support=True)
        p.iterate(steps, err_out=err)
        p.copy_roi(0, savebuf, slot, r0, c0, r1 - r0, c1 - c0)

    def timed(fn):
        ctx.sync()
        t0 = time.perf_counter()
        for k in range(a.steps):
            fn(a.warmup + k)
        ctx.sync()
        return time.perf_counter() - t0, float(err[-1, 0].item())

    with torch.cuda.stream(stream):
        for w in range(a.warmup):
            reconstruction(w)
            reconstruction(w, plan_pp)
        dt_eager, final_err = timed(reconstruction)
        dt_pp, pp_err = timed(lambda k: reconstruction(k, plan_pp))
        graph = plan_pp.capture(lambda: reconstruction(a.steps + a.warmup + 1, plan_pp))
        graph.launch()
        dt_graph, graph_err = timed(lambda k: graph.launch())
        graph.close()

        # per-launch timing of one more reconstruction of each form
        la0, st0 = plan.resident_counts()
        fu0, _ = plan.sw_counts()
        plan.profile(True)
        reconstruction(a.warmup + a.steps)
        nres, tres = plan.profile_read(2)
        plan.profile(False)
        la1, st1 = plan.resident_counts()
        fu1, _ = plan.sw_counts()
        plan_pp.profile(True)
        reconstruction(a.warmup + a.steps, plan_pp)
        ncol, tcol = plan_pp.profile_read(0)
        nrow, trow = plan_pp.profile_read(1)
        plan_pp.profile(False)
        res = None
        if nres and la1 > la0:
            # the resident launch's algorithmic bytes: the survey model's
            # bytes of the steps it ran and of the shrinkwraps inside it
            res_steps, res_sw = (st1 - st0) / (la1 - la0), (fu1 - fu0) / (la1 - la0)
            res = {"avg": tres / nres * 1e-3, "steps": res_steps, "shrinkwraps": res_sw,
                   "bytes": int(res_steps * (B_COL + B_ROW) * n * n + res_sw * B_SHRINK * n * n)}

        # equal final error: the same schedule from the oracle fixture's
        # random starts; the error of each final iterate is the error trace
        # entry of one more (ER, s * psi_in) step
        ref, seeds = oracle_final_errors("single_1024", "schedule")
        fin = []
        probe = [(_capi.ER, _capi.SEM_NUMPY, 0.8, 0.0, 0.0)]
        for sd in seeds:
            np.random.seed(sd)
            g = torch.from_numpy((np.random.rand(n, n) + 1j * np.random.rand(n, n))
                                 .astype(np.complex64)).to(dev)
            plan.set_support(s_d)
            plan.set_psi(g, mult_support=True)
            plan.iterate(steps, error_trace=False)
            fin.append(float(plan.iterate(probe)[0, 0].item()))
        vs_oracle = compare_final(fin, ref, "final modulus error over the oracle fixture's "
                                            "8 random starts, 1000-iteration headline schedule")
    avg = {"column": tcol / ncol * 1e-3, "row": trow / nrow * 1e-3}
    byts = {"column": B_COL * n * n, "row": B_ROW * n * n}
    forms = {"resident": dt_eager, "per_pass": dt_pp, "per_pass_graph": dt_graph}
    best = min(forms, key=forms.get)
    if best == "resident" and res is not None:
        # one launch runs the whole reconstruction: it is the dominant kernel
        avg["resident"], byts["resident"] = res["avg"], res["bytes"]
        dom = "resident"
    else:
        dom = max(("column", "row"), key=lambda k: avg[k])
    model = iters * (B_COL + B_ROW) * n * n + nsw * B_SHRINK * n * n + B_ONEOFF * n * n
    assert graph_err == final_err == pp_err, "the launch forms must give the same reconstruction"
    return dict(frame=frame, n=n, iters=iters, dt=forms[best], forms=forms, launch=best,
                final_err=final_err, avg=avg, byts=byts, dom=dom, model=model, res=res,
                vs_oracle=vs_oracle)


def bench_trials_512(ctx):
    """64 random-start trials of one 512^2 frame, RAAR then DM, 500 iterations."""
    import torch
    from als_speckle_amd import synth
    from als_speckle_amd.batch import phasing_batch
    n, T, iters = 512, 64, 500
    f = synth.config_frame("trials_512", 0)
    b = phasing_batch(T, gpu_info=ctx, autotune=True)
    b.load(f.modulus, f.support)
    log(f"trials_512 autotuned schedule: tile {b.tuning.tile}, {b.tuning.ept} values/thread, "
        f"persistent {b.tuning.persist} ({b.tuning.us_per_iter} us per batched HIO iteration)")
    g = np.empty((T, n, n), np.complex64)
    for t in range(T):
        np.random.seed(t)
        g[t] = (np.random.rand(n, n) + 1j * np.random.rand(n, n)).astype(np.complex64)
    g_d = torch.from_numpy(g).to(ctx.device)
    out = {"schedule": {"tile": b.tuning.tile, "values_per_thread": b.tuning.ept,
                        "persistent": b.tuning.persist, "autotune_table": b.tuning.table}}
    # DM ends with 20 ER iterations in the s * psi_out form: the DM iterate
    # grows at a solution (the reference's DM update doubles it inside the
    # support), so a projecting tail turns its trials into reconstructions
    # whose error can be compared (ER in the numpy form, s * psi_in, would
    # leave the iterate as it is)
    tail = 20
    for alg, beta, bpx, run in (("raar", 0.87, B_COL + B_ROW, "raar"),
                                ("dm", 0.7, 92, "dm_er_tail")):
        b.seed(guesses=g_d)
        b.iterate(10, order=((alg, 1),), beta=beta)           # warm-up
        ctx.sync()
        t0 = time.perf_counter()
        b.seed(guesses=g_d)
        b.iterate(iters, order=((alg, 1),), beta=beta)
        if alg == "dm":
            b.iterate(tail, order=(("er", 1),), beta=beta, semantics="opencl")
        aligned, avg = b.aligned_average()
        ctx.sync()
        dt = time.perf_counter() - t0
        tr = b.get_error_trace()
        # m and s are shared by the trials: charged once per batched
        # iteration; the ER tail (s * psi_out) moves 48 B/px
        shared_px = 8 if alg == "raar" else 12
        model = iters * n * n * ((bpx - shared_px) * T + shared_px)
        its = iters
        if alg == "dm":
            model += tail * n * n * ((48 - 8) * T + 8)
            its += tail
        ref, _ = oracle_final_errors("trials_512", run, key="trace_last")
        out[alg] = {"iterations": its, "recons_per_s": round(T / dt, 2),
                    "ms_per_batch": round(dt * 1e3, 2),
                    "us_per_batched_iteration": round(dt / its * 1e6, 2),
                    "model_hbm_frac": round(model / dt / HBM_PEAK, 4),
                    "final_error_mean": float(np.mean(tr[-1])),
                    "final_error_min": float(np.min(tr[-1])),
                    "final_error_vs_oracle": compare_final(
                        tr[-1], ref, f"modulus error entering the last of {its} iterations, "
                                     f"{T} GPU trials vs the oracle fixture's 8")}
        if alg == "dm":
            out[alg]["schedule"] = f"{iters} DM beta {beta} + {tail} ER (s * psi_out)"
    # the FFT + projection (column) pass and the update (row) pass of RAAR at
    # this batched shape, timed per launch; m and s are shared by the trials,
    # so the bytes are Phi r/w (16 B/px) + m once, and Phi, psi r/w (32 B/px)
    # + s once
    plan = b._plan
    b.seed(guesses=g_d)
    plan.profile(True)
    b.iterate(20, order=(("raar", 1),), beta=0.87)
    ctx.sync()
    nc, tc = plan.profile_read(0)
    nr, tr_ = plan.profile_read(1)
    plan.profile(False)
    col_us, row_us = tc / nc * 1e3, tr_ / nr * 1e3
    col_b = (16 * T + 4) * n * n
    row_b = (32 * T + 4) * n * n
    out["raar"].update({"column_pass_us": round(col_us, 2), "row_pass_us": round(row_us, 2),
                        "column_pass_hbm_frac": round(col_b / (col_us * 1e-6) / HBM_PEAK, 4),
                        "row_pass_hbm_frac": round(row_b / (row_us * 1e-6) / HBM_PEAK, 4)})
    return out


def bench_mixed_sizes(ctx, iters=100):
    """Uneven frame sizes in one launch: 64 frames of 256^2, 16 of 512^2 and
    4 of 1024^2 (the same pixel count per size), HIO beta 0.85 with the
    error trace, as one merged column-pass and row-pass launch per step for
    all three plans (spk_group_iterate), against the same three plans
    iterated by their own calls one after another on one stream and on three
    concurrent streams.  Same kernels' arithmetic in every form (the final
    iterates are compared bitwise)."""
    import torch
    from als_speckle_amd import _capi, synth
    from als_speckle_amd.device import Context, Group, Plan, group_shape
    mix = [(256, 64), (512, 16), (1024, 4)]
    dev = ctx.device
    steps = [(_capi.HIO, _capi.SEM_NUMPY, 0.85, 0.0, 0.0)] * iters
    data = {}
    for n, b in mix:
        f = synth.make_frame(n, 7, shape="polygon", fill=0.14, photons=1e7)
        rng = np.random.default_rng(n)
        g = (rng.random((b, n, n)) + 1j * rng.random((b, n, n))).astype(np.complex64)
        data[n] = (torch.from_numpy(np.stack([f.modulus] * b)).to(dev),
                   torch.from_numpy(np.stack([f.support] * b)).to(dev),
                   torch.from_numpy(g).to(dev))

    def make(c, n, b):
        tile, ept = group_shape(n)
        p = Plan(c, n, batch=b, shared=False, tile=tile, ept=ept, resident=2)
        m, s, g = data[n]
        p.set_modulus(m)
        p.set_support(s)
        p.set_psi(g)
        return p

    def reset(plans):
        for p in plans:
            p.set_support(data[p.n][1])
            p.set_psi(data[p.n][2])

    grp = Group(ctx)
    gplans = [make(ctx, n, b) for n, b in mix]
    streams = [torch.cuda.Stream(device=dev) for _ in mix]
    sctx = [Context(ctx.index, stream=st) for st in streams]
    splans = [make(c, n, b) for c, (n, b) in zip(sctx, mix)]
    forms = {
        "merged": lambda: grp.iterate(gplans, [steps] * len(mix)),
        "one_stream": lambda: [p.iterate(steps) for p in gplans],
        "three_streams": lambda: [p.iterate(steps) for p in splans],
    }
    out = {}
    finals = {}
    for name, fn in forms.items():
        plans = splans if name == "three_streams" else gplans
        best = None
        for rep in range(3):
            reset(plans)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        outs = [p.get_psi() for p in plans]
        torch.cuda.synchronize()          # the plans' streams wrote them
        finals[name] = [o.cpu().numpy() for o in outs]
        px = sum(n * n * b for n, b in mix)
        out[name] = {"us_per_iteration": round(best / iters * 1e6, 2),
                     "frame_iterations_per_s": round(sum(b for _, b in mix) * iters / best, 1),
                     "model_hbm_frac": round((B_COL + B_ROW) * px * iters / best / HBM_PEAK, 4)}
    out["bitwise_equal_to_merged"] = {
        name: all(np.array_equal(a, b) for a, b in zip(finals["merged"], finals[name]))
        for name in ("one_stream", "three_streams")}
    out["frames"] = {str(n): b for n, b in mix}
    out["iterations"] = iters
    out["launches_merged"] = grp.counts()[0]
    for p in gplans + splans:
        p.close()
    grp.close()
    for c in sctx:
        c.close()
    return out


def bench_batched_1024(ctx, frame, trials=16):
    """The headline frame and schedule, `trials` random starts in one plan."""
    import torch
    from als_speckle_amd.device import Plan
    n, iters = 1024, 1000
    steps, _ = headline_steps(iters)
    dev = ctx.device
    plan = Plan(ctx, n, batch=trials, shared=False)
    m = torch.from_numpy(np.stack([frame.modulus] * trials)).to(dev)
    s = torch.from_numpy(np.stack([frame.support] * trials)).to(dev)
    g = np.empty((trials, n, n), np.complex64)
    for t in range(trials):
        np.random.seed(1000 + t)
        g[t] = (np.random.rand(n, n) + 1j * np.random.rand(n, n)).astype(np.complex64)
    g = torch.from_numpy(g).to(dev)
    plan.set_modulus(m)
    err = torch.empty((iters, trials), dtype=torch.float64, device=dev)

    def run():
        plan.set_support(s)
        plan.set_psi(g, mult_support=True)
        plan.iterate(steps, err_out=err)

    run()
    ctx.sync()
    t0 = time.perf_counter()
    run()
    ctx.sync()
    dt = time.perf_counter() - t0
    fe = err[-1].cpu().numpy()
    # the passes of one more run, timed per launch (dispatch-recorded events):
    # the FFT + projection (column) pass and the update (row) pass at their
    # batched shape, against the 8 TB/s roofline (m and s per trial here, so
    # the model bytes are the bytes moved)
    plan.profile(True)
    run()
    ctx.sync()
    nc, tc = plan.profile_read(0)
    nr, tr = plan.profile_read(1)
    plan.profile(False)
    px = trials * n * n
    col_us, row_us = tc / nc * 1e3, tr / nr * 1e3
    return {"trials": trials, "recons_per_s": round(trials / dt, 3),
            "iterations_per_s": round(trials * iters / dt, 1),
            "final_error_mean": float(fe.mean()), "final_error_min": float(fe.min()),
            "tile": plan.tile, "values_per_thread": plan.ept,
            "column_pass_us": round(col_us, 2), "row_pass_us": round(row_us, 2),
            "column_pass_hbm_frac": round(B_COL * px / (col_us * 1e-6) / HBM_PEAK, 4),
            "row_pass_hbm_frac": round(B_ROW * px / (row_us * 1e-6) / HBM_PEAK, 4)}


def bench_single_sizes(ctx, reps=3):
    """One 512^2 and one 256^2 frame (polygon 15 %, Poisson 1e7) with the
    headline schedule: the resident launch in tile 4 (the default) and tile 2
    against one launch per pass."""
    import torch
    from als_speckle_amd import synth
    from als_speckle_amd.device import Plan
    steps, _ = headline_steps(1000)
    out = {}
    for n in (512, 256):
        f = synth.make_frame(n, 3, fill=0.15, photons=1e7)
        np.random.seed(1)
        g = torch.from_numpy((np.random.rand(n, n) + 1j * np.random.rand(n, n))
                             .astype(np.complex64)).to(ctx.device)
        m = torch.from_numpy(f.modulus).to(ctx.device)
        sp = torch.from_numpy(f.support).to(ctx.device)
        forms = {}
        for name, kw in (("resident", {}), ("resident_tile_2", {"tile": 2}),
                         ("per_pass", {"resident": 2})):
            p = Plan(ctx, n, **kw)
            p.set_modulus(m)
            err = torch.empty((len(steps), 1), dtype=torch.float64, device=ctx.device)

            def rec():
                p.set_support(sp)
                p.set_psi(g, mult_support=True)
                p.iterate(steps, err_out=err)

            rec()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                rec()
            ctx.sync()
            dt = (time.perf_counter() - t0) / reps
            p.check()
            forms[name] = round(len(steps) / dt, 1)
            p.close()
        out[str(n)] = {"iterations_per_s": forms}
    return out


def bench_large_4096(ctx):
    """The large-frame workload: one 4096^2 reconstruction of 2000 iterations."""
    import torch
    from als_speckle_amd import synth
    from als_speckle_amd.device import Plan
    from als_speckle_amd.phasing import expand_order, shrinkwrap_schedule
    from als_speckle_amd import _capi
    c = synth.CONFIGS["large_4096"]
    n, iters = c["n"], c["iterations"]
    frame = synth.config_frame("large_4096", 0)
    algs = expand_order(c["order"], iters)
    sw = shrinkwrap_schedule(iters, **c["shrinkwrap"])
    steps = [(_capi.ALG_CODES[x], _capi.SEM_NUMPY, c["beta"], *sw.get(k, (0.0, 0.0)))
             for k, x in enumerate(algs)]
    dev = ctx.device
    plan = Plan(ctx, n, batch=1, shared=True)
    s_d = torch.from_numpy(frame.support).to(dev)
    np.random.seed(1)
    g_d = torch.from_numpy((np.random.rand(n, n) + 1j * np.random.rand(n, n))
                           .astype(np.complex64)).to(dev)
    plan.set_modulus(torch.from_numpy(frame.modulus).to(dev))
    err = torch.empty((iters, 1), dtype=torch.float64, device=dev)

    def run():
        plan.set_support(s_d)
        plan.set_psi(g_d, mult_support=True)
        plan.iterate(steps, err_out=err)

    run()
    ctx.sync()
    t0 = time.perf_counter()
    run()
    ctx.sync()
    dt = time.perf_counter() - t0
    final_err = float(err[-1, 0].item())
    plan.profile(True)
    run()
    ctx.sync()
    nc, tc = plan.profile_read(0)
    nr, tr = plan.profile_read(1)
    plan.profile(False)
    plan.close()
    px = n * n
    col_us, row_us = tc / nc * 1e3, tr / nr * 1e3
    model = iters * (B_COL + B_ROW) * px + len(sw) * B_SHRINK * px + B_ONEOFF * px
    return {"n": n, "iterations": iters, "iterations_per_s": round(iters / dt, 1),
            "recons_per_s": round(1.0 / dt, 4), "final_modulus_error": final_err,
            "tile": plan.tile, "values_per_thread": plan.ept,
            "model_hbm_frac_whole_recon": round(model / dt / HBM_PEAK, 4),
            "column_pass_us": round(col_us, 2), "row_pass_us": round(row_us, 2),
            "column_pass_hbm_frac": round(B_COL * px / (col_us * 1e-6) / HBM_PEAK, 4),
            "row_pass_hbm_frac": round(B_ROW * px / (row_us * 1e-6) / HBM_PEAK, 4)}


def bench_serving(a):
    """Poisson arrivals of mixed frame sizes through the continuous-batching server."""
    import torch
    from als_speckle_amd import synth
    from als_speckle_amd.serve import Request, Server
    mix = {256: 0.4, 512: 0.3, 1024: 0.2, 2048: 0.1}
    rng = np.random.Generator(np.random.PCG64(7))
    dev = torch.device("cuda", 0)
    bank = {}
    for n in mix:
        bank[n] = []
        for k in range(2):
            f = synth.make_frame(n, 100 + 10 * k + n, shape="polygon",
                                 fill=float(rng.uniform(0.10, 0.15)), photons=1e7)
            bank[n].append((torch.from_numpy(f.modulus).to(dev),
                            torch.from_numpy(f.support).to(dev)))
    sizes = list(mix)
    probs = np.array([mix[n] for n in sizes])

    def make_requests(count, arrivals):
        reqs = []
        gen = torch.Generator(device=dev)
        gen.manual_seed(11)
        for i in range(count):
            n = int(sizes[rng.choice(len(sizes), p=probs)])
            m, s = bank[n][int(rng.integers(0, len(bank[n])))]
            re = torch.rand((n, n), device=dev, generator=gen)
            im = torch.rand((n, n), device=dev, generator=gen)
            g = torch.complex(re, im)
            reqs.append(Request(i, m, s, g, iterations=500, shrinkwrap=True,
                                arrival=float(arrivals[i])))
        return reqs

    budget = int(a.budget_gb * (1 << 30))

    def latencies(res, span):
        lat = np.array([r.latency for r in res]) * 1e3
        by = {}
        for n in sizes:
            ln = np.array([r.latency for r in res if r.n == n]) * 1e3
            if ln.size:
                by[str(n)] = {"count": int(ln.size),
                              "p50_ms": round(float(np.percentile(ln, 50)), 2),
                              "p99_ms": round(float(np.percentile(ln, 99)), 2)}
        fin = np.array([r.trace[-1] for r in res])
        return {"requests": len(res), "recons_per_s": round(len(res) / span, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2), "by_size": by,
                "final_error_median": float(np.median(fin))}

    # capacity of the budget-shared server: a saturated burst
    srv = Server(device=0, budget_bytes=budget, mix=mix)
    burst = make_requests(min(200, a.requests), np.zeros(a.requests))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    srv.run(burst)
    cap = len(burst) / (time.perf_counter() - t0)
    srv.close()
    log(f"serving capacity {cap:.1f} req/s (budget {a.budget_gb} GB shared by the sizes)")
    lam = 0.8 * cap
    arrivals = np.cumsum(rng.exponential(1.0 / lam, a.requests))
    reqs = make_requests(a.requests, arrivals)
    out = {"arrival_rate_per_s": round(lam, 2), "capacity_req_per_s": round(cap, 2),
           "hbm_budget_gb": a.budget_gb}
    # the same arrivals through the budget-shared form, the same with the
    # 256/512/1024 classes in merged launches, and the static form (one
    # fixed pool per size, sized from the budget by expected work)
    for form in ("shared_budget", "merged_sizes", "static_pools"):
        if form == "shared_budget":
            srv = Server(device=0, budget_bytes=budget, mix=mix)
        elif form == "merged_sizes":
            srv = Server(device=0, budget_bytes=budget, mix=mix, merge_sizes=True)
        else:
            srv = Server(device=0, slots=Server.size_slots(budget, mix))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = srv.run(reqs)
        span = time.perf_counter() - t0
        d = latencies(res, span)
        d["chunk_slots"] = {str(k): v for k, v in srv.chunk_slots.items()}
        d["peak_reserved_gb"] = round(srv.peak_bytes / (1 << 30), 2)
        d["chunks_allocated"], d["chunks_released"] = srv.chunks_allocated, srv.chunks_released
        srv.close()
        out[form] = d
        log(f"serving {form}: p50 {d['p50_ms']} ms, p99 {d['p99_ms']} ms, "
            f"{d['recons_per_s']} recons/s, peak {d['peak_reserved_gb']} GB")
    # headline fields: the budget-shared form
    for k in ("requests", "recons_per_s", "p50_ms", "p99_ms", "by_size", "final_error_median"):
        out[k] = out["shared_budget"][k]
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(the path runs on a single MI355X)",
              file=sys.stderr)
        sys.exit(2)
    from als_speckle_amd.device import Context
    ctx = Context(0)
    if a.kernel_probe:
        kernel_probe(ctx, a.probe_frame)
        return
    log("single 1024^2 reconstructions")
    s = bench_single(a)
    n, iters, dt = s["n"], s["iters"], s["dt"]
    its_per_s = iters * a.steps / dt
    plain_rps = a.steps / dt
    dom = s["dom"]
    d_avg, d_bytes = s["avg"][dom], s["byts"][dom]
    achieved = d_bytes / d_avg

    kname = KERNEL_NAMES[dom]
    traffic = measured_traffic(dom)
    col_bytes, col_us = B_COL * n * n, s["avg"]["column"]

    out = {
        "metric": "HIO iters/s at 1024^2 c64",
        "value": round(its_per_s, 1),
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": "single_1024: 1024x1024 complex64, 1000 iterations 90 HIO/10 ER "
                        "(beta 0.8, numpy semantics), shrinkwrap every 20 (sigma 3->1.5, "
                        "tau 0.2), FP64 error trace, polygon support 15% fill, Poisson 1e8",
            "n": n, "iterations": iters, "recons_per_s": round(plain_rps, 3),
            "launch": s["launch"],
            "iterations_per_s_by_form": {k: round(iters * a.steps / v, 1)
                                         for k, v in s["forms"].items()},
            "final_modulus_error": s["final_err"],
            "final_error_vs_oracle": s["vs_oracle"],
            "model_bytes_per_recon": s["model"],
            "model_hbm_frac_whole_recon": round(s["model"] * a.steps / dt / HBM_PEAK, 4),
            "column_pass_us": round(s["avg"]["column"] * 1e6, 3),
            "row_pass_us": round(s["avg"]["row"] * 1e6, 3),
            # the FFT + projection (column) pass of the one-launch-per-pass
            # form: 20 B/px per launch
            "fft_projection_pass": {"us": round(col_us * 1e6, 3),
                                    "achieved_gb_s": round(col_bytes / col_us / 1e9, 1),
                                    "frac": round(col_bytes / col_us / HBM_PEAK, 4)},
        },
        "roofline": {
            "bound": "hbm", "kernel": f"k_{kname}<1024>",
            "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": d_bytes,
            "avg_launch_us": round(d_avg * 1e6, 3),
        },
    }
    if s["res"] is not None:
        out["roofline"]["resident_launch"] = {
            "steps_per_launch": s["res"]["steps"], "shrinkwraps_per_launch": s["res"]["shrinkwraps"],
            "us_per_step": round(s["res"]["avg"] * 1e6 / s["res"]["steps"], 3)}
    out["hbm_copy_reference"] = copy_bandwidth(ctx)
    log("single 512^2 and 256^2 frames")
    out["single_sizes"] = bench_single_sizes(ctx)
    if not a.no_batched:
        log("batched trials 64 x 512^2 (RAAR, DM)")
        out["batched"] = {"trials_512": bench_trials_512(ctx)}
        log("batched 16 x 1024^2 headline schedule")
        b16 = bench_batched_1024(ctx, s["frame"])
        b16["speedup_vs_single"] = round(b16["recons_per_s"] / plain_rps, 3)
        out["batched"]["trials_1024"] = b16
        log("uneven frame sizes in one launch (256^2, 512^2, 1024^2)")
        out["batched"]["mixed_sizes"] = bench_mixed_sizes(ctx)
    if not a.no_large:
        log("large 4096^2 reconstruction")
        out["large_4096"] = bench_large_4096(ctx)
    if not a.no_next:
        log("partial coherence, spectrum, PRTF, RFTF, Richardson-Lucy")
        out["next_rows"] = bench_next_rows(ctx)
    if not a.no_serving:
        log(f"serving {a.requests} requests")
        out["serving"] = bench_serving(a)
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(s["frame"], a.cpu_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
