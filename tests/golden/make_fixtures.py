"""Write the golden fixtures of the test suite (run from the repository root):

    python tests/golden/make_fixtures.py [--jobs J] [--only NAME ...]

Every expected value here comes from the complex128 oracle
(``oracle/phasing_c128.py``), the restatement of the reference's numpy
phasing path; the reference itself is Python 2 with pyopencl/pyfft and does
not run in this environment, so no fixture is produced by it.  Inputs come
from ``als_speckle_amd.synth`` (seeded) and the reference's own seeding rule
(``np.random.seed(val)``; ``rand + 1j rand``, ``phasing.py:1002-1006``).

Fixtures:

reference_vectors.json
    Data the reference's own tests hold: the ``bounding_box`` cases of
    ``test/test_mask.py:19-31`` and the circle/annulus areas of
    ``test/test_shape.py:31-51``.
iterates_small.npz
    Short free-running iterations of every algorithm (both ER/SF
    semantics, DM, and HIO with an in-loop shrinkwrap) on 32^2 and 64^2
    frames: inputs, the iterate after each step (complex64) and the error
    trace (float64).
final_error_<workload>.json
    The distribution over random starts of the final modulus error of whole
    reconstructions of the benchmark workloads, for "equal final error"
    comparisons (HIO is chaotic, so single trajectories of a complex64 and
    a complex128 implementation part after some tens of iterations; their
    error levels over seeds must agree):
      cpu_disk_256   200 HIO (beta 0.9) + 20 ER on the 256^2 disk frame
      single_1024    the headline schedule: 1000 iterations 90 HIO / 10 ER
                     (beta 0.8), shrinkwrap every 20 (sigma 3 -> 1.5, tau 0.2)
      trials_512     RAAR beta 0.87 x 500; DM beta 0.7 x 500 followed by 20
                     ER iterations in the s * psi_out form (a projecting
                     tail: the DM iterate grows at a solution, and ER in
                     the numpy form, s * psi_in, leaves it as it is)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SEEDS = tuple(range(1, 9))


def reference_vectors():
    """Inputs and expected outputs of the reference's helper tests."""
    return {
        "bounding_box": {
            "source": "test/test_mask.py:19-31",
            "shape": [256, 256], "box_rows": [18, 90], "box_cols": [45, 63],
            "cases": [{"pad": 0, "expect": [18, 90, 45, 63]},
                      {"pad": 14, "expect": [4, 104, 31, 77]}],
        },
        "circle_area": {
            "source": "test/test_shape.py:31-35",
            "note": "shape.circle(shape, radius, AA=False) (the disk of the reference's CPU "
                    "workload) sums to pi r^2 within a relative 0.001",
            "shape": [256, 256], "radius": 100, "rel_tol": 0.001,
        },
    }


def _steps_case(args):
    """Free-running oracle steps; returns (name, dict of arrays)."""
    from als_speckle_amd import synth
    from oracle import phasing_c128 as O
    n, alg, sem, beta, nsteps, sw = args
    f = synth.make_frame(n, 5 + n, shape="polygon", fill=0.15, photons=1e6)
    s = f.support.astype(np.float64)
    g = O.seed_guess(n, 3, s)
    psi = g.copy()
    out, tr, sups = [], [], []
    for k in range(nsteps):
        psi, _, e = O.iteration(psi, f.modulus, s, alg, beta, sem)
        if sw and k == sw[0]:
            s, _ = O.shrinkwrap_support(psi, s, sw[1], sw[2])
        out.append(psi.astype(np.complex64))
        tr.append(e)
        sups.append(s.astype(np.float32))
    name = f"n{n}_{alg}_{sem}" + ("_sw" if sw else "")
    d = {"modulus": f.modulus, "support": f.support, "guess": g.astype(np.complex64),
         "psi": np.stack(out), "trace": np.array(tr), "beta": np.array(beta),
         "supports": np.stack(sups)}
    if sw:
        d["sw"] = np.array(sw, np.float64)
    return name, d


def iterates_small(jobs):
    cases = []
    for n in (32, 64):
        for alg, sem in (("hio", "numpy"), ("er", "numpy"), ("er", "opencl"), ("raar", "numpy"),
                         ("sf", "numpy"), ("sf", "opencl"), ("dm", "numpy")):
            cases.append((n, alg, sem, 0.7 if alg == "dm" else 0.8, 3, None))
    cases.append((64, "hio", "numpy", 0.9, 3, (1, 2.0, 0.2)))
    with ProcessPoolExecutor(jobs) as ex:
        res = list(ex.map(_steps_case, cases))
    flat = {}
    for name, d in res:
        for k, v in d.items():
            flat[f"{name}/{k}"] = v
    return flat


def _final_run(args):
    """Final modulus error (and trace checkpoints) of one oracle reconstruction."""
    from als_speckle_amd import synth
    from oracle import phasing_c128 as O
    workload, seed, alg, beta, iters, tail = args
    c = synth.CONFIGS[workload]
    f = synth.config_frame(workload, 0)
    g = O.seed_guess(f.modulus.shape[0], seed, f.support)
    t0 = time.time()
    if workload == "trials_512":
        psi, tr, s = O.iterate(g, f.modulus, f.support, iters, ((alg, 1),), beta, "numpy")
        if tail:
            psi, tr2, s = O.iterate(psi, f.modulus, s, tail, (("er", 1),), beta, "opencl")
            tr = np.concatenate([tr, tr2])
    else:
        psi, tr, s = O.iterate(g, f.modulus, f.support, c["iterations"], c["order"], c["beta"],
                               "numpy", shrinkwrap=c.get("shrinkwrap"))
    # the error of the final iterate (the state after the last step)
    final = O.modulus_error(O.fft2(psi), f.modulus)
    return dict(seed=seed, final_error=final, trace_last=float(tr[-1]),
                trace_points={str(k): float(tr[k]) for k in (0, 9, 49, len(tr) // 2, len(tr) - 1)},
                support_fill=float(np.mean(s > 0)), seconds=round(time.time() - t0, 1))


def final_errors(workload, jobs):
    from als_speckle_amd import synth
    c = synth.CONFIGS[workload]
    runs = {}
    if workload == "trials_512":
        specs = {"raar": ("raar", 0.87, 500, 0), "dm_er_tail": ("dm", 0.7, 500, 20)}
    else:
        specs = {"schedule": (None, c["beta"], c["iterations"], 0)}
    for key, (alg, beta, iters, tail) in specs.items():
        with ProcessPoolExecutor(jobs) as ex:
            res = list(ex.map(_final_run, [(workload, sd, alg, beta, iters, tail) for sd in SEEDS]))
        fe = np.array([r["final_error"] for r in res])
        runs[key] = {"algorithm": alg, "beta": beta, "iterations": iters, "er_opencl_tail": tail,
                     "seeds": list(SEEDS), "runs": res,
                     "final_error": {"min": float(fe.min()), "max": float(fe.max()),
                                     "median": float(np.median(fe)), "mean": float(fe.mean())}}
    desc = {k: v for k, v in c.items() if k in ("n", "shape", "fill", "radius", "photons", "amp",
                                                 "order", "iterations", "beta", "shrinkwrap")}
    return {"workload": workload, "config": desc, "frame_seed": 0,
            "guess": "np.random.seed(seed); rand(n, n) + 1j rand(n, n) as complex64, x support",
            "final_error_definition": "sqrt(sum((|fft2 psi| - m)^2) / sum(m^2)) of the final "
                                      "iterate, complex128", "runs": runs}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--only", nargs="*", default=None)
    a = ap.parse_args(argv)
    want = set(a.only) if a.only else None

    def on(name):
        return want is None or name in want

    if on("reference_vectors"):
        with open(os.path.join(HERE, "reference_vectors.json"), "w") as fh:
            json.dump(reference_vectors(), fh, indent=1)
    if on("iterates_small"):
        np.savez_compressed(os.path.join(HERE, "iterates_small.npz"), **iterates_small(a.jobs))
    for wl in ("cpu_disk_256", "trials_512", "single_1024"):
        if on(wl):
            t0 = time.time()
            d = final_errors(wl, a.jobs)
            with open(os.path.join(HERE, f"final_error_{wl}.json"), "w") as fh:
                json.dump(d, fh, indent=1)
            print(f"{wl}: {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
