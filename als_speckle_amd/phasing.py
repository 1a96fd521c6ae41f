"""Drop-in for ``speckle.phasing`` (reference ``dhparks/als_speckle``).

The ``phasing`` class keeps the reference object API
(``speckle/phasing.py:33-1353``): ``load`` modulus / support / numtrials,
``seed``, ``iterate`` with ER/HIO/RAAR/SF/DM, ``get``, ``status`` and the
attributes ``savebuffer, support, r0, c0, rows, cols, numtrial, numtrials, N,
shape, compute_device, can_iterate``.  All iteration work runs in
``libspeckle_hip.so`` on the GPU; there is no CPU path (``force_cpu=True``
raises).

Additive keywords of ``iterate``: ``shrinkwrap`` (in-loop support refinement),
``semantics`` ('numpy' = the reference CPU formulas, the default; 'opencl' =
its GPU kernels, which differ for ER and SF), ``error_trace``.
"""
from __future__ import annotations

import sys
import warnings

import numpy as np
import torch

from . import _capi
from .device import Context, Plan, require_gpu, to_device
from .masking import bounding_box, find_all_objects

__all__ = ["phasing", "expand_order", "shrinkwrap_schedule", "DeviceArray",
           "align_global_phase", "refine_support", "prtf", "rftf", "gpu_prtf", "gpu_rftf",
           "center_of_mass_average"]


def expand_order(order, iterations):
    """Algorithm name per iteration, ``phasing.py:244-260``.

    ``order`` is an iterable of (name, count) entries, default
    (('hio', 99), ('er', 1)); the expanded list repeats cyclically.
    """
    assert isinstance(order, (type(None), list, tuple)), "algorithm order must be iterable"
    if order is None:
        order = (("hio", 99), ("er", 1))
    flat = []
    for entry in order:
        try:
            name, count = str(entry[0]), int(entry[1])
        except (IndexError, TypeError, ValueError):
            raise AssertionError("entry %s in iteration order is improperly formatted" % (entry,))
        flat.extend([name] * count)
    assert flat, "algorithm order is empty"
    for name in set(flat):
        assert name in _capi.ALG_CODES, "real space enforcement algorithm %s is unknown" % name
    k = len(flat)
    return [flat[i % k] for i in range(int(iterations))]


def shrinkwrap_schedule(iterations, every=20, sigma0=3.0, sigma1=1.5, tau=0.2):
    """{iteration index (0-based): (sigma, tau)} for an in-loop shrinkwrap.

    After iteration k (1-based) with k % every == 0; sigma goes linearly from
    sigma0 to sigma1 over the iterations // every events.
    """
    events = int(iterations) // int(every)
    out = {}
    for e in range(events):
        sig = sigma0 if events <= 1 else sigma0 + (sigma1 - sigma0) * e / (events - 1)
        out[(e + 1) * every - 1] = (float(sig), float(tau))
    return out


class DeviceArray:
    """Handle to a device-resident array of a reconstruction.

    Plays the role of the reference's ``arrayWrapper`` (gpu.py:3-15):
    ``phasing.get(handle)`` returns a host numpy copy.
    """

    def __init__(self, name, shape, dtype, fetch):
        self.name, self.shape, self.dtype = name, tuple(shape), np.dtype(dtype)
        self._fetch = fetch
        self.isgpu = True

    @property
    def size(self):
        return int(np.prod(self.shape))

    @property
    def ndim(self):
        return len(self.shape)

    def tensor(self) -> torch.Tensor:
        return self._fetch()

    def get(self) -> np.ndarray:
        return self._fetch().cpu().numpy()

    def __repr__(self):
        return f"DeviceArray({self.name}, shape={self.shape}, dtype={self.dtype})"


class phasing:
    """A reconstruction resident on one MI355X (``speckle/phasing.py:33-179``).

    Typical use is the reference's::

        r = phasing()
        r.load(modulus=m, support=s, numtrials=10)
        for t in range(10):
            r.seed()
            r.iterate(500)
        frames = r.get(r.savebuffer)
    """

    def __init__(self, force_cpu=False, force_np_fft=False, gpu_info=None):
        if force_cpu:
            raise RuntimeError(
                "als_speckle_amd has no CPU path: force_cpu=True is not supported")
        require_gpu()
        # gpu_info: None, a device index, or a Context (the reference passes
        # its OpenCL (context, device, queue, platform) tuple, gpu.py:399-421)
        if isinstance(gpu_info, Context):
            self.ctx = gpu_info
        else:
            self.ctx = Context(int(gpu_info) if gpu_info is not None else 0)
        self.use_gpu = True
        self.compute_device = "gpu"
        self.gpu_info = self.ctx

        # state flags, phasing.py:121-127
        self.modulus_state = 0
        self.support_state = 0
        self.ipsf_state = 0
        self.buffer_state = 0
        self.rl_state = 0
        self.spectrum_state = 0
        self.can_iterate = False
        self.numtrial = 0
        self.array_dtypes = ("float32", "complex64")

        self.N = None
        self.shape = None
        self.size = None
        self.modulus = None
        self.support = None
        self.psi_in = None
        self.numtrials = None
        self.r0 = self.c0 = self.rows = self.cols = None
        self.savebuffer = None
        self.reconstruction = None
        self.ipsf = None

        self._plan = None
        self._support_host = None   # host copy of the last loaded support (N x N)
        self._seeded = False
        self._trace = None           # device float64 trace of the last iterate()
        self._traces = []

    # ------------------------------------------------------------------
    def _make_plan(self):
        self._plan = Plan(self.ctx, self.N, batch=1, shared=True)
        if getattr(self, "_spectrum_host", None) is not None:
            self._plan.set_spectrum(*self._spectrum_host)
        plan = self._plan
        self.modulus = DeviceArray("modulus", self.shape, np.float32, lambda: self._modulus_dev)
        self.psi_in = DeviceArray("psi_in", self.shape, np.complex64, lambda: plan.get_psi(-1))

    def load(self, modulus=None, support=None, ipsf=None, spectrum=None, numtrials=None):
        """Load data, ``phasing.py:291-505``; keyword arguments only, any order.

        modulus: corner-centred N x N Fourier modulus (N a power of two).
        support: square array <= N x N, normalised to max 1; a smaller
            support is embedded top-left (phasing.py:466-485).
        numtrials: number of independent trials (savebuffer slots + 1).
        """
        types = (type(None), np.ndarray)
        assert isinstance(modulus, types), "modulus must be ndarray if supplied"
        assert isinstance(support, types), "support must be ndarray if supplied"
        assert isinstance(ipsf, types), "psf must be ndarray if supplied"
        if spectrum is not None:
            # phasing.py:430-448: rows (energies, weights); rescale factors are
            # the energies over the energy of the largest weight
            assert isinstance(spectrum, np.ndarray)
            assert spectrum.ndim == 2
            assert spectrum.shape[0] == 2
            energies, weights = spectrum[0], spectrum[1]
            center_e = energies[weights.argmax()]
            self._spectrum_host = ((energies / center_e).astype(np.float32),
                                   weights.astype(np.float32))
            self.N_spctrm = np.int32(energies.shape[0])
            self.spectrum_state = 2
            if self._plan is not None:
                self._plan.set_spectrum(*self._spectrum_host)
        if ipsf is not None:
            # phasing.py:414-425: transverse coherence, same size as the modulus
            assert ipsf.ndim == 2
            assert ipsf.shape[0] == ipsf.shape[1]
            ipsf = ipsf.astype(np.complex64)
            if self.ipsf_state == 2:
                assert ipsf.shape == self._ipsf_host.shape
            self._ipsf_host = ipsf
            self.ipsf_state = 1

        if modulus is not None:
            assert modulus.ndim == 2
            modulus = modulus.astype(np.float32)
            if self.modulus_state == 2:
                assert modulus.shape == self.shape
            else:
                assert modulus.shape[0] == modulus.shape[1]
                n = modulus.shape[0]
                assert n & (n - 1) == 0, "modulus size must be a power of 2"
                self.N, self.shape, self.size = n, modulus.shape, modulus.size
                self._make_plan()
            self._modulus_dev = to_device(modulus, self.ctx.device, torch.float32)
            self._plan.set_modulus(self._modulus_dev, -1)
            self.modulus_state = 2

        if self.modulus_state == 2 and self.ipsf_state > 0:
            # phasing.py:458-460
            assert self._ipsf_host.shape == self.shape
            if self.ipsf_state == 1:
                self._ipsf_dev = to_device(self._ipsf_host, self.ctx.device, torch.complex64)
                self._plan.set_ipsf(self._ipsf_dev)
                dev = self._ipsf_dev
                self.ipsf = DeviceArray("ipsf", self.shape, np.complex64, lambda: dev)
            self.ipsf_state = 2

        if support is not None:
            assert support.ndim == 2, "support is %s-dim" % support.ndim
            assert support.shape[0] == support.shape[1], \
                "support shape is %s %s, must be square" % support.shape
            support = support.astype(np.float32)
            support *= 1.0 / support.max()
            if self.support_state == 2:
                self.numtrial = 0
            self._support_host = support
            bounds = bounding_box(support, pad=10)
            self.r0, self.rows = int(bounds[0]), int(bounds[1] - bounds[0])
            self.c0, self.cols = int(bounds[2]), int(bounds[3] - bounds[2])
            self.support_state = 1

        if numtrials is not None:
            self.numtrials = int(numtrials)
            if self.buffer_state == 0:
                self.buffer_state = 1

        if self.modulus_state == 2 and self.support_state > 0:
            sup = self._support_host
            if sup.shape != self.shape:
                assert self.rows <= self.shape[0]
                assert self.cols <= self.shape[1]
                crop = sup[self.r0:self.r0 + self.rows, self.c0:self.c0 + self.cols]
                new = np.zeros(self.shape, np.float32)
                new[:self.rows, :self.cols] = crop
                bounds = bounding_box(new, force_to_square=True, pad=4)
                self.r0, self.rows = int(bounds[0]), int(bounds[1] - bounds[0])
                self.c0, self.cols = int(bounds[2]), int(bounds[3] - bounds[2])
                sup = new
                self._support_host = sup
            if self.support_state == 1:
                self._support_dev = to_device(sup, self.ctx.device, torch.float32)
                self._plan.set_support(self._support_dev, -1)
                plan = self._plan
                self.support = DeviceArray("support", self.shape, np.float32,
                                           lambda: plan.get_support(-1))
            self.support_state = 2
            if self.buffer_state > 0:
                shape = (self.numtrials + 1, self.rows, self.cols)
                if self.buffer_state == 1 or tuple(self.savebuffer.shape) != shape:
                    buf = torch.zeros(shape, dtype=torch.complex64, device=self.ctx.device)
                    self._savebuffer = buf
                    self.savebuffer = DeviceArray("savebuffer", shape, np.complex64,
                                                  lambda: self._savebuffer)
                    self.buffer_state = 2

        if self.modulus_state == 2 and self.support_state == 2 and self.buffer_state == 2:
            self.can_iterate = True

    # ------------------------------------------------------------------
    def seed(self, val=None, supplied=None):
        """Random start, ``phasing.py:990-1031``.

        ``np.random.seed(val)`` if given, then U[0,1) + iU[0,1) (real part
        drawn first) from numpy's global legacy generator, times the support.
        ``supplied`` embeds a guess top-left and rolls it onto the support by
        the argmax of the magnitude cross-correlation (conditioning.align_frames
        with use_mag_only, phasing.py:1008-1024).
        """
        assert self.N is not None, "cannot seed without N being set. load data first."
        assert self.support_state == 2, "load a support before seeding"
        if val is not None:
            np.random.seed(val)
        if supplied is None:
            guess = np.random.rand(self.N, self.N) + 1j * np.random.rand(self.N, self.N)
            guess = guess.astype(np.complex64)
            g = to_device(guess, self.ctx.device, torch.complex64)
        else:
            s = np.asarray(supplied)
            assert s.ndim == 2
            assert s.shape[0] <= self.N and s.shape[1] <= self.N
            if s.shape != (self.N, self.N):
                s2 = np.zeros(self.shape, np.complex64)
                s2[:s.shape[0], :s.shape[1]] = s.astype(np.complex64)
                s = s2
            g = to_device(s.astype(np.complex64), self.ctx.device, torch.complex64)
            g = self._align_to_support(g)
        self._plan.set_psi(g, -1, mult_support=True)
        self._seeded = True

    def _align_to_support(self, g: torch.Tensor) -> torch.Tensor:
        """Roll g so |g| best overlaps the support (device FFT cross-correlation).

        crosscorr.alignment_coordinates (crosscorr.py:526-568): the roll is
        minus the argmax of |IFFT(F|g| * conj(F support))| taken as a signed
        cyclic shift.
        """
        n = self.N
        mag = torch.abs(g).to(torch.complex64).contiguous()
        ref = self._support_dev.to(torch.complex64).contiguous()
        fa = self._plan.fft2(mag)
        fb = self._plan.fft2(ref)
        corr = self._plan.fft2((fa * torch.conj(fb)).contiguous(), inverse=True)
        k = int(torch.argmax(torch.abs(corr)).item())
        dr, dc = k // n, k % n
        if dr > n // 2:
            dr -= n
        if dc > n // 2:
            dc -= n
        return torch.roll(g, shifts=(-dr, -dc), dims=(0, 1)).contiguous()

    # ------------------------------------------------------------------
    def _debug_snapshots(self, iteration, alg, when):
        """``iterate(debug_on=...)`` dumps of ``_iteration``
        (``phasing.py:1193-1226``, ``1244-1246``), as FITS files in the
        working directory with the reference's names: before the iteration
        ``psi_fourier_1`` (fftshift of Psi = fft2(psi_in)), ``fourier_div``
        (the divisor) and ``psi_fourier_2`` (Psi after the modulus
        replacement), in polar components, and ``out`` (psi_out) in cartesian
        components; after a HIO iteration ``hio update`` (the new psi_in).
        The intermediates come from the library's stand-alone transform of
        the same iterate (the iteration itself runs the fused passes, which
        keep them on chip); with a spectral rescale only ``psi_fourier_1``
        and the update are written.  For DM the reference's second
        projection overwrites the same files; these hold the first."""
        from . import io as sio
        shift = np.fft.fftshift
        tag = "%s i%s" % (self.compute_device, iteration)
        if when == "after":
            if alg == "hio":
                sio.save("%s hio update.fits" % tag, self.get(self.psi_in), components="cartesian")
            return
        psi = self._plan.get_psi()
        Psi_d = self._plan.fft2(psi)
        Psi = Psi_d.cpu().numpy()
        sio.save("%s psi_fourier_1.fits" % tag, shift(Psi), components="polar")
        if getattr(self, "_spectrum_host", None) is not None:
            warnings.warn("debug_on: with a spectrum only psi_fourier_1 and the update are written")
            return
        ipsf = getattr(self, "_ipsf_host", None)
        if ipsf is not None:
            # div = sqrt(|ifft2(fft2(|Psi|^2) * ipsf)|)  (phasing.py:1175-1183)
            dev = self.ctx.device
            I = torch.from_numpy((np.abs(Psi) ** 2).astype(np.complex64)).to(dev)
            F = self._plan.fft2(I).cpu().numpy() * ipsf
            conv = self._plan.fft2(torch.from_numpy(F.astype(np.complex64)).to(dev), inverse=True)
            div = np.sqrt(np.abs(conv.cpu().numpy())).astype(np.float32)
        else:
            div = np.abs(Psi).astype(np.float32)
        sio.save("%s fourier_div.fits" % tag, shift(div.astype(np.complex64)), components="polar")
        m = self._modulus_dev.cpu().numpy()
        with np.errstate(divide="ignore", invalid="ignore"):
            Pp = np.where(div > 0, Psi * (m / np.where(div > 0, div, 1)), 0).astype(np.complex64)
        sio.save("%s psi_fourier_2.fits" % tag, shift(Pp), components="polar")
        out = self._plan.fft2(torch.from_numpy(Pp).to(self.ctx.device), inverse=True)
        sio.save("%s out.fits" % tag, out.cpu().numpy(), components="cartesian")

    def iterate(self, iterations, order=None, silent=True, beta=0.8, debug_on=None,
                shrinkwrap=None, semantics="numpy", error_trace=True):
        """Run iterations, ``phasing.py:181-289``.

        order: iterable of (name, count), names in hio/er/raar/sf/dm, cycled;
            default (('hio', 99), ('er', 1)).
        beta: feedback parameter for hio/raar/dm (default 0.8).
        silent: True, False or an int (progress every int iterations).
        shrinkwrap: None, True, or a dict(every=20, sigma0=3, sigma1=1.5,
            tau=0.2): support *= [G_sigma * |psi| > tau max] after every
            `every` iterations (build extension; the reference refines the
            support between rounds with refine_support).
        semantics: 'numpy' (reference CPU formulas) or 'opencl' (reference GPU
            kernels) for ER and SF.
        error_trace: record the modulus error of every iteration
            (``get_error_trace``).
        debug_on: iterations (0-based) at which to write the reference's
            FITS debugging snapshots (see ``_debug_snapshots``).
        At the end the ROI of the iterate is copied to savebuffer[numtrial].
        """
        if not (self.modulus_state == 2 and self.support_state == 2):
            print("cant iterate before loading support and modulus.")
            print("modulus state: " + str(self.modulus_state))
            print("support state: " + str(self.support_state))
            print("buffer  state: " + str(self.buffer_state))
            raise AssertionError("cant iterate before loading support and modulus")
        assert self._seeded, "seed() before iterate()"
        assert semantics in _capi.SEM_CODES, "semantics must be 'numpy' or 'opencl'"
        iterations = int(iterations)
        algs = expand_order(order, iterations)
        sw = {}
        if shrinkwrap:
            kw = dict(every=20, sigma0=3.0, sigma1=1.5, tau=0.2)
            if isinstance(shrinkwrap, dict):
                kw.update(shrinkwrap)
            sw = shrinkwrap_schedule(iterations, **kw)
        sem = _capi.SEM_CODES[semantics]
        steps = []
        for k, a in enumerate(algs):
            sig, tau = sw.get(k, (0.0, 0.0))
            steps.append((_capi.ALG_CODES[a], sem, beta, sig, tau))

        # progress printing and debug snapshots need host sync points;
        # otherwise one call
        every = None
        if silent is not True:
            every = 1 if (silent is False or not isinstance(silent, int)) else max(1, int(silent))
        debug = sorted({int(k) for k in (debug_on or ()) if 0 <= int(k) < iterations})
        cuts = set(debug) | {k + 1 for k in debug} | {0, iterations}
        if every is not None:
            cuts |= set(range(0, iterations, every))
        cuts = sorted(c for c in cuts if c <= iterations)
        traces = []
        resident_before = self._plan.resident_counts()[0]
        for start, stop in zip(cuts[:-1], cuts[1:]):
            if start in debug:
                self._debug_snapshots(start, algs[start], "before")
            traces.append(self._plan.iterate(steps[start:stop], error_trace=error_trace))
            if start in debug:
                self._debug_snapshots(start, algs[start], "after")
            if every is not None and stop % every == 0:
                self.ctx.sync()
                print("  iteration %s" % stop)
                sys.stdout.flush()
        if self._plan.resident_counts()[0] != resident_before:
            # resident launches wait at grid barriers: raise if one gave up
            # (the reference's iterate blocks on its kernels, gpu.py:349)
            self._plan.check()
        if error_trace and traces:
            self._trace = torch.cat([t[:, 0] for t in traces if t is not None])
        else:
            self._trace = None

        if self.buffer_state == 2:
            assert self.numtrial < self.numtrials + 1, \
                "numtrial %d exceeds the savebuffer (%d slots)" % (self.numtrial, self.numtrials + 1)
            self._plan.copy_roi(0, self._savebuffer, self.numtrial, self.r0, self.c0,
                                self.rows, self.cols)
        self.numtrial += 1

    def get_error_trace(self):
        """FP64 modulus error of each iteration of the last ``iterate`` call.

        E_k = sqrt(sum(|F psi_k| - m)^2 / sum m^2) for the iterate entering
        iteration k (build-defined; the reference computes none).
        """
        return None if self._trace is None else self._trace.cpu().numpy()

    # ------------------------------------------------------------------
    def get(self, something):
        """Host copy of a device array (``gpu.py:389-397``); other values pass through."""
        if isinstance(something, DeviceArray):
            return something.get()
        if isinstance(something, torch.Tensor):
            return something.detach().cpu().numpy()
        return something

    def status(self):
        """Print the data status, ``phasing.py:1033-1044``."""
        ro = {2: "complete", 1: "incomplete", 0: "nothing"}
        print("states:")
        print("  modulus %s %s" % (ro[self.modulus_state], self.shape))
        print("  support %s" % ro[self.support_state])
        print("  ipsf    %s" % ro[self.ipsf_state])
        print("  buffer  %s" % ro[self.buffer_state])
        print("  spectrm %s" % ro[self.spectrum_state])
        print("  device  %s" % self.compute_device)

    def optimize_coherence(self, best_estimate, modulus=None, force_cpu=False, load=True,
                           silent=False, iterations=500, method="richardson_lucy",
                           is_fourier=False, semantics="numpy"):
        """Estimate the coherence function from reconstructions,
        ``phasing.py:507-930``, on the GPU.

        best_estimate: 2-d or 3-d (frames) real-space reconstruction(s),
        embedded top-left in the modulus shape.  method:
          'richardson_lucy': ``iterations`` Richardson-Lucy updates from the
            start Gaussian; the ipsf is fft2 of the result (its real part with
            semantics 'numpy', the reference's CPU branch; complex with
            'opencl') over |ipsf[0, 0]|.  The reference's result depends only
            on the last frame (its running sum ``rl_sum`` is never read), so
            only the last frame is computed.  Loaded as the ipsf if ``load``;
            returns it centre-shifted, as the reference.
          'gaussian': per frame, scipy ``fmin`` (ftol 0.1, xtol 0.1, from
            (n/2, n/2)) of sum (sqrt|fft2(g * ac)| - m)^2 over the Gaussian's
            two coherence lengths, each objective evaluated on the GPU; the
            mean lengths' Gaussian is loaded as the ipsf if ``load``; returns
            the (2, frames) lengths.
        """
        assert method in ("richardson_lucy", "gaussian"), "unknown coherence method %s" % method
        if force_cpu:
            raise RuntimeError("there is no CPU path; optimize_coherence runs on the GPU")
        if modulus is None:
            assert self.modulus_state == 2, "load a modulus or pass one"
            m_dev = self._modulus_dev
        else:
            assert isinstance(modulus, np.ndarray) and modulus.ndim == 2
            assert modulus.shape[0] == modulus.shape[1]
            m_dev = to_device(modulus.astype(np.float32), self.ctx.device, torch.float32)
        n = int(m_dev.shape[0])
        if isinstance(best_estimate, DeviceArray):
            best_estimate = best_estimate.get()
        est = np.asarray(best_estimate)
        assert est.ndim in (2, 3), "best_estimate must be 2d or 3d"
        if est.ndim == 2:
            est = est[None]
        assert est.shape[1] <= n and est.shape[2] <= n
        plan = self._plan if (self._plan is not None and self.N == n) else Plan(self.ctx, n)
        if method == "richardson_lucy":
            frame = to_device(est[-1].astype(np.complex64), self.ctx.device, torch.complex64)
            ipsf = plan.coherence_rl(frame, m_dev, int(iterations), semantics, is_fourier)
            out = ipsf.cpu().numpy()
            if load:
                self.load(ipsf=out)
            return np.fft.fftshift(out)
        from scipy.optimize import fmin
        optcls = np.zeros((2, est.shape[0]), np.float32)
        for k, fr in enumerate(est):
            plan.gauss_prepare(to_device(fr.astype(np.complex64), self.ctx.device,
                                         torch.complex64), m_dev)
            optcls[:, k] = fmin(lambda c: plan.gauss_eval(c[0], c[1]), (n // 2, n // 2),
                                disp=False, ftol=.1, xtol=0.1)
        ave = np.average(optcls, axis=1)
        if not silent:
            print("recovered coherence lengths")
            print(ave)
        if load:
            rows, cols = np.indices((n, n)) - n // 2
            rows, cols = np.fft.fftshift(rows), np.fft.fftshift(cols)
            g = np.exp(-rows ** 2 / (2 * ave[0] ** 2)) * np.exp(-cols ** 2 / (2 * ave[1] ** 2))
            self.load(ipsf=g)
        return optcls


_default_ctx = None


def _context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def align_global_phase(data_in):
    """Remove the global phase of each frame, ``phasing.py:1355-1404``.

    Per frame the rotation exp(i p) minimising sum (Re sin p + Im cos p)^2,
    then a sign flip so that sum Re >= 0.  The reference finds p with
    ``fminbound``; here p is the closed-form minimiser (the smaller
    eigenvector of the Gram matrix of (Re, Im)), computed on the GPU with
    FP64 sums.  Returns a new array of the input's shape and dtype; the
    rotation is applied in complex64.
    """
    data = np.asarray(data_in)
    assert isinstance(data, np.ndarray), "data must be array"
    assert data.ndim in (2, 3), "data must be 2d or 3d"
    assert np.iscomplexobj(data), "data must be complex"
    ctx = _context()
    t = to_device(data.astype(np.complex64), ctx.device, torch.complex64)
    t3 = t.unsqueeze(0) if t.dim() == 2 else t
    ctx.align_global_phase(t3)
    return t.cpu().numpy().astype(data.dtype)


_post_plans = {}


def _post_plan(n):
    """A one-frame plan of size n, kept for the transforms of prtf / rftf."""
    if n < 16 or n > 4096 or n & (n - 1):
        raise ValueError("the transform size must be a power of 2 in [16, 4096] (is %d)" % n)
    p = _post_plans.get(n)
    if p is None:
        p = _post_plans[n] = Plan(_context(), n, batch=1, shared=True)
    return p


def refine_support(support, average_mag, blur=3, local_threshold=.2,
                   global_threshold=0, kill_weakest=False):
    """Shrinkwrap between rounds, ``phasing.py:1747-1815`` (Marchesini 2003).

    The blur of |average_mag| by the normalised, fftshifted Gaussian of
    ``shape.gaussian(support.shape, (blur, blur))`` runs on the GPU as the
    equivalent periodic separable convolution (FP64 sums of FP32 taps); it
    keeps the reference's odd-size kernel offset.  The connected parts of the
    support (values in [0.95, 1.05], ``masking.find_all_objects``) and the
    thresholds are host work, as in the reference:
    refined = sum over parts of [B*part > local_threshold * max(B*part)],
    times [B > global_threshold * max B], optionally without the weakest part.
    Returns (refined, blurred); refined has the support's dtype, blurred is
    float64 (computed in FP32 storage).  Unlike the reference, the caller's
    support array is not zeroed by the part search (``masking.py:428``).
    """
    assert isinstance(support, np.ndarray), "support must be array"
    assert support.ndim == 2, "support must be 2d"
    assert isinstance(average_mag, np.ndarray), "average_mag must be array"
    assert average_mag.ndim == 2, "average_mag must be 2d"
    assert isinstance(blur, (float, int)), "blur must be a number (is %s)" % type(blur)
    assert isinstance(global_threshold, (float, int)), \
        "global_threshold must be a number (is %s)" % type(global_threshold)
    assert isinstance(local_threshold, (float, int)), \
        "local_threshold must be a number (is %s)" % type(local_threshold)

    ctx = _context()
    mag = to_device(np.abs(average_mag).astype(np.float32), ctx.device, torch.float32)
    blurred = ctx.blur2d(mag, float(blur)).cpu().numpy().astype(np.float64)
    refined = np.zeros_like(support)
    global_passed = np.where(blurred > blurred.max() * global_threshold, 1, 0)
    parts = find_all_objects(support)
    part_sums = np.zeros(len(parts), float)
    for n, part in enumerate(parts):
        current = blurred * part
        refined += np.where(current > current.max() * local_threshold, 1, 0).astype(
            refined.dtype)
        if kill_weakest:
            part_sums[n] = np.sum(np.abs(average_mag) * part)
    refined *= global_passed.astype(refined.dtype)
    if kill_weakest:
        refined *= (1 - parts[part_sums.argmin()]).astype(refined.dtype)
    return refined, blurred


def prtf(estimates, N=None, silent=True, prtfq=False):
    """Phase-retrieval transfer function, ``phasing.py:1406-1461``.

    fftshift(|mean_f F_f / |F_f||) over the aligned estimates (frames, r, c),
    each zero-padded top-left into N x N (default max(r, c)) and transformed
    on the GPU; N must be a power of 2 in [16, 4096] (the reference's GPU
    variant has the same restriction).  The reference's CPU version adds the
    r x c transforms into an N x N sum, so it only runs when r == c == N; the
    padding is the same thing there.  The q-average (``prtfq``) needs
    ``wrapping.unwrap`` and is not provided.
    """
    assert isinstance(estimates, (np.ndarray, list, tuple)), "must be iterable"
    est = np.asarray(estimates)
    assert est.ndim == 3, "must be 3d"
    if prtfq:
        raise NotImplementedError("prtfq needs wrapping.unwrap, which is not provided")
    f, r, c = est.shape
    if N is None:
        N = max(r, c)
    plan = _post_plan(int(N))
    t = to_device(est.astype(np.complex64), plan.ctx.device, torch.complex64)
    return plan.prtf(t).cpu().numpy()


def rftf(estimate, goal_modulus, hot_pixels=False, ipsf=None, rftfq=False, scale=False):
    """Fourier-error transfer function, ``phasing.py:1521-1609``.

    |F| of the estimate zero-padded into the goal's shape (with an ipsf:
    sqrt(|ifft2(fft2(|F|) * ipsf)|), as the reference blurs the modulus) is
    compared with the goal modulus at the corner and the centre alignment
    (the closer one is used, ``phasing.py:1564-1573``); returns
    sqrt(1/(1+err)) with err = (|F| - g)^2 / g^2, centre-shifted, optionally
    rescaled from [1/sqrt(2), 1] to [0, 1].  The goal must be square with a
    power-of-2 side in [16, 4096].  ``hot_pixels`` (the conditioning module)
    and ``rftfq`` (``wrapping.unwrap``) are not provided.
    """
    assert isinstance(estimate, np.ndarray), "estimate must be ndarray"
    assert isinstance(goal_modulus, np.ndarray), "goal_modulus must be ndarray"
    assert estimate.shape <= goal_modulus.shape, "estimate must be smaller than goal_modulus"
    if hot_pixels or rftfq:
        raise NotImplementedError("hot_pixels and rftfq are not provided")
    n = goal_modulus.shape[0]
    assert goal_modulus.shape == (n, n), "goal_modulus must be square"
    if ipsf is not None:
        assert isinstance(ipsf, np.ndarray) and ipsf.ndim == 2
        assert ipsf.shape == goal_modulus.shape
    plan = _post_plan(n)
    dev = plan.ctx.device
    e = to_device(estimate.astype(np.complex64), dev, torch.complex64)
    g = to_device(goal_modulus.astype(np.float32), dev, torch.float32)
    ip = None if ipsf is None else to_device(ipsf.astype(np.complex64), dev, torch.complex64)
    return plan.rftf(e, g, scale=bool(scale), ipsf=ip).cpu().numpy()


def gpu_prtf(gpuinfo, estimates, N=None, prtfq=False):
    """``phasing.py:1463-1519``: the same as ``prtf`` (which already runs on the GPU)."""
    return prtf(estimates, N=N, prtfq=prtfq)


def gpu_rftf(gpuinfo, estimates, goal_modulus, ipsf=None, rftfq=False):
    """Stack RFTF, ``phasing.py:1611-1708`` with ``phasing_rftf_accumulate.cl:15-24``.

    Unlike ``rftf`` this averages the per-frame value
    1/sqrt(1 + (g - |F_f|)^2 / g^2) over the estimates (frames, r, c), with
    the goal taken at the corner (no alignment search), then centre-shifts.
    With an ipsf |F_f| becomes sqrt(|ifft2(fft2(|F_f|^2) * Re ipsf)|): the
    reference blurs the intensity here and uploads the ipsf as float32 (its
    real part).  ``gpuinfo`` is accepted for the signature and ignored;
    ``rftfq`` is not provided.
    """
    assert isinstance(estimates, np.ndarray) and estimates.ndim == 3
    assert isinstance(goal_modulus, np.ndarray) and goal_modulus.ndim == 2
    assert goal_modulus.shape[0] == goal_modulus.shape[1]
    if rftfq:
        raise NotImplementedError("rftfq is not provided")
    if ipsf is not None:
        assert isinstance(ipsf, np.ndarray) and ipsf.shape == goal_modulus.shape
    plan = _post_plan(goal_modulus.shape[0])
    dev = plan.ctx.device
    e = to_device(estimates.astype(np.complex64), dev, torch.complex64)
    g = to_device(goal_modulus.astype(np.float32), dev, torch.float32)
    ip = None if ipsf is None else to_device(np.real(ipsf).astype(np.complex64), dev,
                                              torch.complex64)
    return plan.rftf_stack(e, g, ipsf=ip).cpu().numpy()


def center_of_mass_average(imgs):
    """Sum of reconstructions aligned by the centre of mass of their
    magnitudes, ``phasing.py:1710-1745``, on the GPU.

    Each frame is rolled by minus the integer (truncated) offset of its
    centre of mass from frame 0's and the rolled frames are summed (the
    reference returns the sum, not the mean) in FP64; returns complex64 (the
    reference accumulates in the input dtype).  Unlike the reference, a 2-d
    input is not reshaped in place.
    """
    assert isinstance(imgs, np.ndarray)
    assert imgs.ndim in (2, 3)
    stack = imgs[None] if imgs.ndim == 2 else imgs
    ctx = _context()
    t = to_device(stack.astype(np.complex64), ctx.device, torch.complex64)
    return ctx.com_average(t).cpu().numpy()
