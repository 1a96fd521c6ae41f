"""The HIP path against the committed golden fixtures (tests/golden/).

1. iterates_small.npz: three free-running iterations of every algorithm and
   semantics (and HIO with an in-loop shrinkwrap) on 32^2 and 64^2 frames,
   from the fixture's start: iterates within 1e-5 relative L2 per step, error
   trace within 1e-5 relative, shrinkwrap masks equal except at pixels whose
   blurred value lies within 1e-5 of the threshold.

2. Equal final error.  HIO from a random start is chaotic: a complex64 and a
   complex128 trajectory from the same start part after some tens of
   iterations (the reference's own complex64 numpy path parts from the
   complex128 one the same way), so a single final state cannot be compared;
   the final error level can.  For the reference's CPU workload
   (256^2 disk, 200 HIO beta 0.9 + 20 ER), the 1024^2 headline schedule
   (1000 iterations with 50 shrinkwraps) and the batched-trials frame (512^2,
   RAAR beta 0.87 x 500; DM beta 0.7 x 500 + 20 ER in the s * psi_out form),
   the GPU's final modulus errors over the fixture's 8 random starts (and
   more for the trials) must have their median inside the range the
   complex128 oracle's 8 runs span.

3. Where the iteration is contractive the final states themselves agree: from
   each GPU HIO state of the CPU workload, 50 ER iterations in the
   s * psi_out form free-running on the GPU and in the oracle end within
   1e-3 relative L2 after global-phase alignment.
"""
import json
import os

import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd.device import Plan
from als_speckle_amd.phasing import expand_order, shrinkwrap_schedule
from oracle import phasing_c128 as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def fixture(name):
    with open(os.path.join(GOLD, f"final_error_{name}.json")) as fh:
        return json.load(fh)


def test_small_iterate_fixtures(gpu_ctx):
    z = np.load(os.path.join(GOLD, "iterates_small.npz"))
    names = sorted({k.split("/")[0] for k in z.files})
    for name in names:
        print(f"[golden {name}]", flush=True)
        g = lambda k: z[f"{name}/{k}"]
        n = g("modulus").shape[0]
        alg, sem = name.split("_")[1], name.split("_")[2]
        sw = g("sw") if f"{name}/sw" in z.files else None
        p = Plan(gpu_ctx, n)
        p.set_modulus(dev(g("modulus"), torch.float32))
        p.set_support(dev(g("support"), torch.float32))
        p.set_psi(dev(g("guess"), torch.complex64), mult_support=False)
        for k in range(g("psi").shape[0]):
            sig, tau = (float(sw[1]), float(sw[2])) if sw is not None and k == int(sw[0]) else (0.0, 0.0)
            e = p.iterate([(_capi.ALG_CODES[alg], _capi.SEM_CODES[sem], float(g("beta")), sig, tau)])
            e = float(e.cpu().numpy()[0, 0])
            got = p.get_psi().cpu().numpy()
            assert rel(got, g("psi")[k]) < 1e-5, (name, k, rel(got, g("psi")[k]))
            assert abs(e - g("trace")[k]) < 1e-5 * g("trace")[k], (name, k)
            sg = p.get_support().cpu().numpy()
            if sig > 0:
                prev = g("supports")[k - 1] if k else g("support")
                _, b = O.shrinkwrap_support(g("psi")[k].astype(np.complex128), prev, sig, tau)
                near = np.abs(b - tau * b.max()) < 1e-5 * b.max()
                assert np.all((sg == g("supports")[k]) | near), (name, k)
                # continue from the fixture's support where a near-threshold pixel differs
                p.set_support(dev(g("supports")[k], torch.float32))
            else:
                assert np.array_equal(sg, g("supports")[k]), (name, k)
        p.close()


def gpu_final_errors(ctx, workload, steps, seeds, frame):
    """Final modulus error of each seed's reconstruction (one plan, a frame per seed)."""
    n = frame.modulus.shape[0]
    B = len(seeds)
    p = Plan(ctx, n, batch=B, shared=False)
    p.set_modulus(dev(np.stack([frame.modulus] * B), torch.float32))
    p.set_support(dev(np.stack([frame.support] * B), torch.float32))
    g = np.stack([O.seed_guess(n, sd, np.ones((n, n))).astype(np.complex64) for sd in seeds])
    p.set_psi(dev(g, torch.complex64), mult_support=True)
    for k0 in range(0, len(steps), 100):           # error traces in chunks
        p.iterate(steps[k0:k0 + 100], error_trace=False)
    psi = p.get_psi().cpu().numpy()
    p.close()
    m = frame.modulus
    return np.array([O.modulus_error(O.fft2(psi[b].astype(np.complex128)), m) for b in range(B)]), psi


def within(fe, ref, label):
    lo, hi = ref["final_error"]["min"], ref["final_error"]["max"]
    med = float(np.median(fe))
    print(f"  {label}: gpu median {med:.5f} (range {fe.min():.5f}..{fe.max():.5f}); "
          f"oracle median {ref['final_error']['median']:.5f} range {lo:.5f}..{hi:.5f}", flush=True)
    assert lo <= med <= hi, label


def test_equal_final_error_cpu_workload_and_contractive_tail(gpu_ctx):
    print("[final error: cpu_disk_256]", flush=True)
    ref = fixture("cpu_disk_256")["runs"]["schedule"]
    c = synth.CONFIGS["cpu_disk_256"]
    f = synth.config_frame("cpu_disk_256", 0)
    steps = [(_capi.ALG_CODES[a], _capi.SEM_NUMPY, c["beta"], 0.0, 0.0)
             for a in expand_order(c["order"], c["iterations"])]
    fe, _ = gpu_final_errors(gpu_ctx, "cpu_disk_256", steps, ref["seeds"], f)
    within(fe, ref, "200 HIO + 20 ER")
    # the contractive tail: 50 ER (s * psi_out) from each GPU HIO state
    hio = steps[:200]
    er = [(_capi.ER, _capi.SEM_OPENCL, c["beta"], 0.0, 0.0)] * 50
    B = len(ref["seeds"])
    p = Plan(gpu_ctx, 256, batch=B, shared=True)
    p.set_modulus(dev(f.modulus, torch.float32))
    p.set_support(dev(f.support, torch.float32))
    g = np.stack([O.seed_guess(256, sd, np.ones((256, 256))).astype(np.complex64)
                  for sd in ref["seeds"]])
    p.set_psi(dev(g, torch.complex64), mult_support=True)
    p.iterate(hio, error_trace=False)
    start = p.get_psi().cpu().numpy()
    p.iterate(er, error_trace=False)
    end = p.get_psi().cpu().numpy()
    p.close()
    worst = 0.0
    for b in range(B):
        want, _, _ = O.iterate(start[b], f.modulus, f.support, 50, (("er", 1),), c["beta"], "opencl")
        r = rel(O.align_global_phase_closed(end[b]), O.align_global_phase_closed(want))
        worst = max(worst, r)
    print(f"  ER tail from the GPU HIO states: worst aligned rel {worst:.2e}", flush=True)
    assert worst < 1e-3


def test_equal_final_error_headline_schedule(gpu_ctx):
    print("[final error: single_1024 headline schedule]", flush=True)
    ref = fixture("single_1024")["runs"]["schedule"]
    c = synth.CONFIGS["single_1024"]
    f = synth.config_frame("single_1024", 0)
    sw = shrinkwrap_schedule(c["iterations"], **c["shrinkwrap"])
    steps = [(_capi.ALG_CODES[a], _capi.SEM_NUMPY, c["beta"], *sw.get(k, (0.0, 0.0)))
             for k, a in enumerate(expand_order(c["order"], c["iterations"]))]
    fe, _ = gpu_final_errors(gpu_ctx, "single_1024", steps, ref["seeds"], f)
    within(fe, ref, "1000 iterations, 50 shrinkwraps")


@pytest.mark.parametrize("run", ["raar", "dm_er_tail"])
def test_equal_final_error_batched_trials(gpu_ctx, run):
    print(f"[final error: trials_512 {run}]", flush=True)
    ref = fixture("trials_512")["runs"][run]
    f = synth.config_frame("trials_512", 0)
    alg, beta = ref["algorithm"], ref["beta"]
    steps = [(_capi.ALG_CODES[alg], _capi.SEM_NUMPY, beta, 0.0, 0.0)] * ref["iterations"]
    steps += [(_capi.ER, _capi.SEM_OPENCL, beta, 0.0, 0.0)] * ref["er_opencl_tail"]
    seeds = list(range(1, 33))            # the oracle's 8 seeds and 24 more
    fe, _ = gpu_final_errors(gpu_ctx, "trials_512", steps, seeds, f)
    within(fe, ref, f"{alg} beta {beta} x {ref['iterations']} + {ref['er_opencl_tail']} ER")
