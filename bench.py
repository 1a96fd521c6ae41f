"""Benchmark of the phase-retrieval hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-iters I] [--no-cpu]

Workload (single reconstruction, 1024 x 1024 complex64): synthetic frame
(random star polygon, ~15 % fill, amplitude U[0.5, 1], phase U[-pi, pi],
Poisson 1e8 photons, support dilated 3 px), 1000 iterations cycling 90 HIO /
10 ER (beta 0.8, numpy semantics), in-loop shrinkwrap after every 20
iterations (sigma 3 -> 1.5 px, tau 0.2), FP64 modulus-error trace, ROI copy
into the savebuffer.  One step = one whole reconstruction (seed from a
device-resident guess, reset support, 1000 iterations, ROI copy); inputs are
resident in HBM before the timed region.  value = iterations per second.

Also reported: the dominant kernel's HBM roofline (algorithmic bytes per
launch / average launch time, timed with HIP events on the library stream
during one profiled reconstruction) and the restated reference numpy path
timed on the host cores (bounded sample).
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
    ap.add_argument("--cpu-iters", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


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
    return {"value": iters / dt, "unit": "iterations/s", "cores": 1, "kind": "port",
            "sample": f"{iters} iterations (90 HIO/10 ER, beta 0.8) of the 1024x1024 frame, "
                      f"numpy {np.__version__} complex64 path with the reference's redundant "
                      f"FFT per _fft2 call, single-threaded ({cores} host cores visible)"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(the path runs on a single MI355X; see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from als_speckle_amd import _capi, synth
    from als_speckle_amd.device import Context, Plan
    from als_speckle_amd.phasing import expand_order, shrinkwrap_schedule

    n, iters = 1024, 1000
    frame = synth.config_frame("single_1024", 0)
    ctx = Context(0)
    dev = ctx.device
    m_d = torch.from_numpy(frame.modulus).to(dev)
    s_d = torch.from_numpy(frame.support).to(dev)
    np.random.seed(1)
    guess = (np.random.rand(n, n) + 1j * np.random.rand(n, n)).astype(np.complex64)
    g_d = torch.from_numpy(guess).to(dev)
    r0, r1, c0, c1 = (int(v) for v in __import__("als_speckle_amd.masking", fromlist=["x"]).bounding_box(frame.support, pad=10))
    savebuf = torch.zeros((a.steps + a.warmup + 2, r1 - r0, c1 - c0), dtype=torch.complex64, device=dev)
    err = torch.empty((iters, 1), dtype=torch.float64, device=dev)

    algs = expand_order((("hio", 90), ("er", 10)), iters)
    sw = shrinkwrap_schedule(iters, every=20, sigma0=3.0, sigma1=1.5, tau=0.2)
    steps = [(_capi.ALG_CODES[x], _capi.SEM_NUMPY, 0.8, *sw.get(k, (0.0, 0.0)))
             for k, x in enumerate(algs)]

    plan = Plan(ctx, n, batch=1, shared=True)
    plan.set_modulus(m_d)

    def reconstruction(slot):
        plan.set_support(s_d)
        plan.set_psi(g_d, mult_support=True)
        plan.iterate(steps, err_out=err)
        plan.copy_roi(0, savebuf, slot, r0, c0, r1 - r0, c1 - c0)

    for w in range(a.warmup):
        reconstruction(w)
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        reconstruction(a.warmup + k)
    ctx.sync()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    its_per_s = iters * a.steps / dt
    final_err = float(err[-1, 0].item())

    # dominant kernel: one profiled reconstruction, HIP events around each pass
    plan.profile(True)
    reconstruction(a.warmup + a.steps)
    ncol, tcol = plan.profile_read(0)
    nrow, trow = plan.profile_read(1)
    plan.profile(False)
    avg_col, avg_row = tcol / ncol * 1e-3, trow / nrow * 1e-3     # seconds
    bytes_col, bytes_row = B_COL * n * n, B_ROW * n * n
    passes = {"column": (avg_col, bytes_col), "row": (avg_row, bytes_row)}
    dom = max(passes, key=lambda k: passes[k][0] * (ncol if k == "column" else nrow))
    d_avg, d_bytes = passes[dom]
    achieved = d_bytes / d_avg
    model_bytes = iters * (B_COL + B_ROW) * n * n + len(sw) * B_SHRINK * n * n + B_ONEOFF * n * n

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}pass_1024.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("bytes_per_launch")
        except Exception:
            traffic = None

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
            "n": n, "iterations": iters, "recons_per_s": round(a.steps / dt, 3),
            "final_modulus_error": final_err,
            "model_bytes_per_recon": model_bytes,
            "model_hbm_frac_whole_recon": round(model_bytes * a.steps / dt / HBM_PEAK, 4),
            "column_pass_us": round(avg_col * 1e6, 3), "row_pass_us": round(avg_row * 1e6, 3),
        },
        "roofline": {
            "bound": "hbm", "kernel": f"k_{'colpass' if dom == 'column' else 'rowpass'}<1024>",
            "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": d_bytes,
            "avg_launch_us": round(d_avg * 1e6, 3),
        },
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(frame, a.cpu_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
