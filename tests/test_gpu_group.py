"""Frames of different sizes in one launch (spk_group_iterate) against each
plan's own spk_iterate call.

A merged launch runs, for every member's frame and band, the same per-band
code as the member's own pass kernel (the tile of the mixed-size shape), so
iterates, supports and error traces agree bitwise.  Covered: 256^2, 512^2
and 1024^2 plans of several frames with per-frame schedules (HIO, ER and SF
in both semantics, RAAR, idle frames, in-loop shrinkwraps fused and
stand-alone), members idle for a whole step, members whose steps run their
own launches (DM, a partially coherent projection) inside the same call,
one merged step against the oracle, and the argument checks.
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd._capi import SpkError
from als_speckle_amd.device import Context, Group, group_shape
from oracle import phasing_c128 as O

pytestmark = pytest.mark.gpu


def step(alg, beta=0.85, sem="numpy", sigma=0.0, tau=0.0):
    if alg == "noop":
        return (_capi.NOOP, 0, 0.0, 0.0, 0.0)
    return (_capi.ALG_CODES[alg], _capi.SEM_CODES[sem], beta, sigma, tau)


def member_data(n, frames, seed):
    fr = [synth.make_frame(n, seed + j, shape="polygon", fill=0.14, photons=1e7)
          for j in range(frames)]
    rng = np.random.default_rng(seed)
    g = (rng.random((frames, n, n)) + 1j * rng.random((frames, n, n))).astype(np.complex64)
    m = np.stack([f.modulus for f in fr]).astype(np.float32)
    s = np.stack([f.support for f in fr]).astype(np.float32)
    return m, s, g


def load(plan, data, dev):
    m, s, g = data
    b = plan.batch
    plan.set_modulus(torch.from_numpy(m[:b] if b > 1 else m[0]).to(dev))
    plan.set_support(torch.from_numpy(s[:b] if b > 1 else s[0]).to(dev))
    plan.set_psi(torch.from_numpy(g[:b] if b > 1 else g[0]).to(dev))


def schedule(nsteps, frames, seed, sw_every=None):
    """Per-frame tables: each frame its own mix of algorithms and idle steps,
    with shrinkwraps every sw_every steps (shifted per frame)."""
    rng = np.random.default_rng(seed)
    algs = [("hio", "numpy"), ("er", "numpy"), ("er", "opencl"), ("raar", "numpy"),
            ("sf", "numpy"), ("sf", "opencl")]
    tab = []
    for k in range(nsteps):
        for f in range(frames):
            if rng.random() < 0.1:
                tab.append(step("noop"))
                continue
            a, sem = algs[rng.integers(len(algs))]
            sig = tau = 0.0
            if sw_every and (k + f) % sw_every == sw_every - 1:
                sig, tau = (2.0 if k % 2 else 5.0), 0.2    # sigma 5: stand-alone
            tab.append(step(a, 0.85, sem, sig, tau))
    return tab


MEMBERS = [(256, 3), (512, 2), (1024, 1)]


def state(plan):
    return plan.get_psi().cpu().numpy(), plan.get_support().cpu().numpy()


def run_alone(ctx, n, frames, data, calls):
    tile, ept = group_shape(n)
    from als_speckle_amd.device import Plan
    p = Plan(ctx, n, batch=frames, shared=False, tile=tile, ept=ept, resident=2)
    load(p, data, ctx.device)
    errs = [p.iterate(t, per_frame=True).cpu().numpy() for t in calls]
    out = state(p) + (errs,)
    p.close()
    return out


def test_mixed_sizes_equal_their_own_calls(gpu_ctx):
    print("[group mixed sizes]", flush=True)
    ctx = gpu_ctx
    datas = [member_data(n, f, 100 + 10 * j) for j, (n, f) in enumerate(MEMBERS)]
    calls = [[schedule(23, f, 7 + j, sw_every=6), schedule(9, f, 70 + j, sw_every=4)]
             for j, (n, f) in enumerate(MEMBERS)]
    grp = Group(ctx)
    plans = [grp.plan(n, batch=f, shared=False) for n, f in MEMBERS]
    for p, d in zip(plans, datas):
        load(p, d, ctx.device)
    got_errs = [[] for _ in MEMBERS]
    for c in range(2):
        errs = grp.iterate(plans, [calls[j][c] for j in range(len(MEMBERS))],
                           per_frame=[True] * len(MEMBERS))
        for j, e in enumerate(errs):
            got_errs[j].append(e.cpu().numpy())
    launches, passes = grp.counts()
    assert launches >= 2 * 30 and passes > launches    # merged launches covered several plans
    for j, ((n, f), p) in enumerate(zip(MEMBERS, plans)):
        psi, sup = state(p)
        want = run_alone(ctx, n, f, datas[j], calls[j])
        assert np.array_equal(psi, want[0]), f"iterate n={n}"
        assert np.array_equal(sup, want[1]), f"support n={n}"
        for a, b in zip(got_errs[j], want[2]):
            assert np.array_equal(a, b, equal_nan=True), f"error trace n={n}"
        p.close()
    grp.close()


def gaussian_ipsf(n, width):
    x = np.fft.fftfreq(n) * n
    psf = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (2.0 * width ** 2))
    return np.fft.fft2(psf / psf.sum()).astype(np.complex64)


def test_members_with_dm_ipsf_and_idle_steps(gpu_ctx):
    """A DM member (its DM steps and their predecessors run with its own
    launches inside the call), a member with a partially coherent ipsf (all
    its steps run with its own launches), a member idle for whole steps, and
    shared step tables."""
    print("[group dm + ipsf + idle]", flush=True)
    ctx = gpu_ctx
    datas = [member_data(256, 2, 300), member_data(512, 1, 310), member_data(1024, 1, 320)]
    dm_tab = ([step("hio")] * 3 + [step("dm", 0.7)] * 3 + [step("er")] * 2 + [step("raar")] * 2)
    idle_tab = ([step("noop")] * 4 + [step("hio")] * 4 + [step("noop")] * 2)
    ipsf_tab = [step("hio", 0.9)] * 6 + [step("er")] * 4
    tabs = [dm_tab, idle_tab, ipsf_tab]
    ipsf = torch.from_numpy(gaussian_ipsf(1024, 1.5)).to(ctx.device)
    grp = Group(ctx)
    plans = [grp.plan(256, batch=2, shared=False), grp.plan(512, batch=1), grp.plan(1024)]
    for p, d in zip(plans, datas):
        load(p, d, ctx.device)
    plans[2].set_ipsf(ipsf)
    errs = grp.iterate(plans, tabs)
    from als_speckle_amd.device import Plan
    for j, (p, d, tab, e) in enumerate(zip(plans, datas, tabs, errs)):
        tile, ept = group_shape(p.n)
        q = Plan(ctx, p.n, batch=p.batch, shared=p.shared, tile=tile, ept=ept)
        load(q, d, ctx.device)
        if j == 2:
            q.set_ipsf(ipsf)
        want_e = q.iterate(tab)
        assert np.array_equal(p.get_psi().cpu().numpy(), q.get_psi().cpu().numpy())
        assert np.array_equal(e.cpu().numpy(), want_e.cpu().numpy(), equal_nan=True)
        q.close()
    for p in plans:
        p.close()
    grp.close()


def test_one_merged_step_matches_the_oracle(gpu_ctx):
    print("[group oracle step]", flush=True)
    ctx = gpu_ctx
    grp = Group(ctx)
    frames = {n: synth.make_frame(n, 50 + n // 256, shape="polygon", fill=0.15, photons=1e7)
              for n in (256, 512, 1024)}
    plans, starts = [], {}
    for n, f in frames.items():
        p = grp.plan(n)
        rng = np.random.default_rng(n)
        g = (rng.random((n, n)) + 1j * rng.random((n, n))).astype(np.complex64)
        starts[n] = g.astype(np.complex128) * f.support
        p.set_modulus(torch.from_numpy(f.modulus).to(ctx.device))
        p.set_support(torch.from_numpy(f.support).to(ctx.device))
        p.set_psi(torch.from_numpy(g).to(ctx.device))
        plans.append(p)
    errs = grp.iterate(plans, [[step("hio", 0.8)]] * 3)
    for p, e, (n, f) in zip(plans, errs, frames.items()):
        want, _, err = O.iteration(starts[n], f.modulus, f.support, "hio", beta=0.8)
        got = p.get_psi().cpu().numpy()
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5, n
        assert abs(e[0, 0].item() - err) / err < 1e-5, n
        p.close()
    grp.close()


def test_argument_checks(gpu_ctx):
    ctx = gpu_ctx
    grp = Group(ctx)
    from als_speckle_amd.device import Plan
    with pytest.raises(SpkError, match="256, 512 and 1024"):
        group_shape(2048)
    good = grp.plan(512)
    wrong = Plan(ctx, 512, tile=8, ept=8)             # not the 512 group tile
    f = synth.make_frame(512, 1, shape="disk", radius=60, photons=1e6)
    for p in (good, wrong):
        p.set_modulus(torch.from_numpy(f.modulus).to(ctx.device))
        p.set_support(torch.from_numpy(f.support).to(ctx.device))
        p.set_psi(torch.from_numpy(f.support.astype(np.complex64)).to(ctx.device))
    with pytest.raises(SpkError, match="mixed-size shape"):
        grp.iterate([good, wrong], [[step("hio")]] * 2)
    with pytest.raises(SpkError, match="twice"):
        grp.iterate([good, good], [[step("hio")]] * 2)
    other = Context(ctx.index)
    q = Group(other).plan(512)
    q.set_modulus(torch.from_numpy(f.modulus).to(ctx.device))
    q.set_support(torch.from_numpy(f.support).to(ctx.device))
    q.set_psi(torch.from_numpy(f.support.astype(np.complex64)).to(ctx.device))
    with pytest.raises(SpkError, match="context"):
        grp.iterate([good, q], [[step("hio")]] * 2)
    with pytest.raises(ValueError, match="one step table per plan"):
        grp.iterate([good], [[step("hio")] * 2, [step("hio")]])
    small = grp.plan(256)
    with pytest.raises(ValueError, match="same number of steps"):
        grp.iterate([good, small], [[step("hio")] * 2, [step("hio")]])
    before = good.get_psi().cpu().numpy()
    with pytest.raises(SpkError):
        grp.iterate([good], [[step("hio"), (99, 0, 0.8, 0.0, 0.0)]])   # bad entry
    assert np.array_equal(good.get_psi().cpu().numpy(), before)      # nothing launched
    for p in (good, wrong, q, small):
        p.close()
    grp.close()
    other.close()
