"""The hot path on the frames of the headline workloads (synth.CONFIGS), against
the complex128 oracle:

- the 1024^2 single reconstruction (polygon ~15 % fill, Poisson 1e8, 90 HIO /
  10 ER, beta 0.8, numpy semantics, shrinkwrap every 20 with sigma 3 -> 1.5
  and tau 0.2): the first 100 iterations of that schedule in lockstep (the
  oracle restarted from the GPU iterate at every iteration), covering 90 HIO,
  10 ER and 5 shrinkwraps; per-iteration modulus error within 1e-5 relative,
  iterates within 1e-5 relative L2;
- ER on the same frame free-running for 30 iterations: trace within 1e-5,
  final iterate within 1e-3 relative L2 after global-phase alignment;
- the 4096^2 large frame (polygon ~8 %, amplitude U[0.3, 1], Poisson 1e9):
  HIO, HIO with the shrinkwrap, ER from the seeded start, one iteration each
  in lockstep (the dual Phi layout runs at this size).
- the same 4096^2 frame and a 2048^2 frame: one RAAR, SF (both semantics)
  and ER (s * psi_out) step each against the oracle, and three DM steps in
  one call (the row pass emits the next step's modified estimate; dual
  layout at 4096) bitwise equal to three single-step calls, each of those
  against the oracle;
- the batched-trials frame (512^2, polygon ~10 %, Poisson 1e7) in a plan of
  64 resident trials: RAAR beta 0.87 and DM beta 0.7 in lockstep on frames
  0, 31 and 63, then the global-phase alignment and trial mean of all 64
  final iterates against the oracle's;
- the 256^2 disk frame of the reference's CPU case: HIO beta 0.9 in lockstep
  for 20 iterations, then 20 ER iterations (s * psi_out form) free-running
  (trace and final iterate).

Shrinkwrap masks must agree except at pixels whose blurred value lies within
1e-5 of the threshold (FP32 blur vs FP64 oracle).
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd.device import Plan
from als_speckle_amd.phasing import expand_order, shrinkwrap_schedule
from oracle import phasing_c128 as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def seeded(n, s, val):
    """phasing.seed(val): U[0,1) + i U[0,1), times the support (phasing.py:1000-1031)."""
    np.random.seed(val)
    g = np.random.rand(n, n) + 1j * np.random.rand(n, n)
    return (g * s).astype(np.complex64)


def plan_for(ctx, f, psi):
    n = f.modulus.shape[0]
    p = Plan(ctx, n)
    p.set_modulus(dev(f.modulus, torch.float32))
    p.set_support(dev(f.support, torch.float32))
    p.set_psi(dev(psi, torch.complex64), mult_support=False)
    return p


def lockstep(p, f, psi, steps, beta=0.8):
    """steps: (alg, sigma, tau) per iteration; returns the worst trace and
    iterate errors against the oracle restarted from the GPU state."""
    cur = psi.astype(np.complex128)
    s = f.support.astype(np.float64)
    worst_t = worst_s = 0.0
    for k, (alg, sig, tau) in enumerate(steps):
        e_gpu = p.iterate([(_capi.ALG_CODES[alg], _capi.SEM_NUMPY, beta, sig, tau)]).cpu().numpy()[0, 0]
        nxt = p.get_psi().cpu().numpy()
        want, _, e = O.iteration(cur, f.modulus, s, alg, beta, "numpy")
        worst_t = max(worst_t, abs(e_gpu - e) / e)
        worst_s = max(worst_s, rel(nxt, want))
        cur = nxt.astype(np.complex128)
        if sig > 0:
            sg = p.get_support().cpu().numpy()
            so, b = O.shrinkwrap_support(cur, s, sig, tau)
            near = np.abs(b - tau * b.max()) < 1e-5 * b.max()
            assert np.all((sg == so) | near), f"shrinkwrap mask at iteration {k}"
            s = sg.astype(np.float64)
    return worst_t, worst_s


def test_single_1024_first_100_lockstep(gpu_ctx):
    print("[single_1024 lockstep 100]", flush=True)
    c = synth.CONFIGS["single_1024"]
    f = synth.config_frame("single_1024", seed=0)
    psi = seeded(1024, f.support, 1)
    algs = expand_order(c["order"], c["iterations"])[:100]
    sw = shrinkwrap_schedule(c["iterations"], **c["shrinkwrap"])
    steps = [(a, *sw.get(k, (0.0, 0.0))) for k, a in enumerate(algs)]
    assert sum(1 for s in steps if s[1] > 0) == 5 and algs.count("er") == 10
    p = plan_for(gpu_ctx, f, psi)
    wt, ws = lockstep(p, f, psi, steps)
    print(f"  worst trace rel {wt:.2e}, worst step rel {ws:.2e}")
    assert wt < TOL and ws < TOL


def test_single_1024_er_free_running(gpu_ctx):
    print("[single_1024 er free-running 30]", flush=True)
    f = synth.config_frame("single_1024", seed=0)
    psi = seeded(1024, f.support, 2)
    p = plan_for(gpu_ctx, f, psi)
    tr = p.iterate([(_capi.ER, _capi.SEM_NUMPY, 0.8, 0.0, 0.0)] * 30).cpu().numpy()[:, 0]
    got = p.get_psi().cpu().numpy()
    want, otr, _ = O.iterate(psi, f.modulus, f.support, 30, (("er", 1),), 0.8, "numpy")
    print(f"  worst trace rel {np.max(np.abs(tr - otr) / otr):.2e}")
    assert np.max(np.abs(tr - otr) / otr) < TOL
    assert rel(O.align_global_phase_closed(got), O.align_global_phase_closed(want)) < 1e-3


def test_large_4096_lockstep(gpu_ctx):
    print("[large_4096 lockstep 3]", flush=True)
    f = synth.config_frame("large_4096", seed=0)
    psi = seeded(4096, f.support, 3)
    p = plan_for(gpu_ctx, f, psi)
    wt, ws = lockstep(p, f, psi, [("hio", 0.0, 0.0), ("hio", 3.0, 0.2), ("er", 0.0, 0.0)])
    print(f"  worst trace rel {wt:.2e}, worst step rel {ws:.2e}")
    assert wt < TOL and ws < TOL


def test_trials_512_batched_raar_dm_align_mean(gpu_ctx):
    print("[trials_512 batch 64: raar/dm lockstep, align, mean]", flush=True)
    f = synth.config_frame("trials_512", seed=0)
    n, b = 512, 64
    psis = np.stack([seeded(n, f.support, 10 + t) for t in range(b)])
    p = Plan(gpu_ctx, n, batch=b, shared=True, tile=4, ept=8)
    p.set_modulus(dev(f.modulus, torch.float32))
    p.set_support(dev(f.support, torch.float32))
    p.set_psi(dev(psis, torch.complex64), mult_support=False)
    s = f.support.astype(np.float64)
    cur = psis
    for alg, beta in [("raar", 0.87), ("raar", 0.87), ("dm", 0.7), ("dm", 0.7)]:
        e = p.iterate([(_capi.ALG_CODES[alg], _capi.SEM_NUMPY, beta, 0.0, 0.0)]).cpu().numpy()[0]
        nxt = p.get_psi().cpu().numpy()
        for t in (0, 31, 63):
            want, _, eo = O.iteration(cur[t].astype(np.complex128), f.modulus, s, alg, beta,
                                      "numpy")
            assert abs(e[t] - eo) / eo < TOL, (alg, t)
            assert rel(nxt[t], want) < TOL, (alg, t)
        cur = nxt
    gal = gpu_ctx.align_global_phase(dev(cur, torch.complex64))
    mean = gpu_ctx.stack_mean(gal).cpu().numpy()
    want_al = O.align_global_phase_closed(cur.astype(np.complex128))
    assert rel(gal.cpu().numpy(), want_al) < TOL
    assert rel(mean, want_al.mean(axis=0)) < TOL
    p.close()


def test_cpu_disk_256_hio_then_er(gpu_ctx):
    print("[cpu_disk_256 hio lockstep 20, er free-running 20]", flush=True)
    f = synth.config_frame("cpu_disk_256", seed=0)
    psi = seeded(256, f.support, 4)
    p = plan_for(gpu_ctx, f, psi)
    wt, ws = lockstep(p, f, psi, [("hio", 0.0, 0.0)] * 20, beta=0.9)
    print(f"  worst trace rel {wt:.2e}, worst step rel {ws:.2e}")
    assert wt < TOL and ws < TOL
    start = p.get_psi().cpu().numpy()
    # ER in the reference's GPU-kernel form (s * psi_out, phasing_er.cl); the
    # numpy form (s * psi_in) leaves a supported iterate unchanged
    tr = p.iterate([(_capi.ER, _capi.SEM_OPENCL, 0.9, 0.0, 0.0)] * 20).cpu().numpy()[:, 0]
    got = p.get_psi().cpu().numpy()
    want, otr, _ = O.iterate(start, f.modulus, f.support, 20, (("er", 1),), 0.9, "opencl")
    assert otr[-1] < otr[0]
    print(f"  worst trace rel {np.max(np.abs(tr - otr) / otr):.2e}")
    assert np.max(np.abs(tr - otr) / otr) < TOL
    assert rel(O.align_global_phase_closed(got), O.align_global_phase_closed(want)) < 1e-3


@pytest.mark.parametrize("n", [2048, 4096])
def test_large_every_update_and_dm_emission(gpu_ctx, n):
    """4096: the large frame, dual layout; 2048: a frame of the same
    generator (polygon ~10 %, Poisson 1e8), single tiled layout (TT 2, E 16)."""
    print(f"[n={n} raar/sf/er one step each, dm 3 steps]", flush=True)
    if n == 4096:
        f = synth.config_frame("large_4096", seed=0)
    else:
        f = synth.make_frame(n, 3, shape="polygon", fill=0.10, photons=1e8)
    psi = seeded(n, f.support, 5)
    s = f.support.astype(np.float64)
    for alg, sem, beta in [("raar", "numpy", 0.87), ("sf", "numpy", 0.8), ("sf", "opencl", 0.8),
                           ("er", "opencl", 0.8)]:
        p = plan_for(gpu_ctx, f, psi)
        e = p.iterate([(_capi.ALG_CODES[alg], _capi.SEM_CODES[sem], beta, 0.0, 0.0)]).cpu().numpy()[0, 0]
        got = p.get_psi().cpu().numpy()
        p.close()
        want, _, eo = O.iteration(psi.astype(np.complex128), f.modulus, s, alg, beta, sem)
        print(f"  {alg}/{sem}: trace rel {abs(e - eo) / eo:.2e}, step rel {rel(got, want):.2e}")
        assert abs(e - eo) / eo < TOL and rel(got, want) < TOL, (alg, sem)
    dm = (_capi.DM, _capi.SEM_NUMPY, 0.7, 0.0, 0.0)
    p = plan_for(gpu_ctx, f, psi)
    tr_all = p.iterate([dm] * 3).cpu().numpy()[:, 0]
    end_all = p.get_psi().cpu().numpy()
    p.close()
    q = plan_for(gpu_ctx, f, psi)
    cur = psi.astype(np.complex128)
    for k in range(3):
        e = q.iterate([dm]).cpu().numpy()[0, 0]
        nxt = q.get_psi().cpu().numpy()
        want, _, eo = O.iteration(cur, f.modulus, s, "dm", 0.7, "numpy")
        print(f"  dm step {k}: trace rel {abs(e - eo) / eo:.2e}, step rel {rel(nxt, want):.2e}")
        assert rel(nxt, want) < TOL and abs(e - eo) / eo < TOL
        assert e == tr_all[k]
        cur = nxt.astype(np.complex128)
    assert np.array_equal(nxt, end_all)
    q.close()
