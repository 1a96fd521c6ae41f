"""The reference's CPU iteration, restated as written (complex64 storage).

TEST INFRASTRUCTURE ONLY -- the timed CPU baseline (``bench.py``
``cpu_baseline``, kind "port") and the "FP32 numpy" drift envelope in tests.

This mirrors ``phasing._iteration`` on its non-GPU, non-numexpr branch with
``np.fft`` as the FFT backend (``phasing.py:112-116``), keeping what costs
time there:

* ``_fft2`` computes a redundant ``ifft2`` on every call
  (``phasing.py:1077-1084``) and casts to complex64;
* the divisor is ``np.abs(psi)`` (``phasing.py:1177``), the replacement is
  ``m*pf/np.abs(fd)`` (``phasing.py:1217``);
* updates are the numpy formulas (``phasing.py:1231-1328``), followed by
  ``.astype(np.complex64)`` of psi_out and psi_in (``phasing.py:1339``,
  ``1348``).

With numpy >= 2 complex64 transforms run in FP32 (numpy 1.x promoted them to
FP64); this module uses whatever the installed numpy does.  It computes no
error metric, like the reference.
"""
from __future__ import annotations

import numpy as np


class RefNumpyPhasing:
    """State of one reconstruction on the reference's CPU path."""

    def __init__(self, modulus, support, beta=0.8):
        self.modulus = np.asarray(modulus, np.float32)
        sup = np.asarray(support, np.float32)
        self.support = sup * (1.0 / sup.max())          # phasing.py:389-390
        self.beta = beta
        self.psi_in = np.zeros(self.modulus.shape, np.complex64)
        self.psi_out = np.zeros(self.modulus.shape, np.complex64)
        self.psi_fourier = np.zeros(self.modulus.shape, np.complex64)

    def seed(self, val=None):
        """``phasing.py:1002-1031`` (CPU branch)."""
        if val is not None:
            np.random.seed(val)
        n = self.modulus.shape[0]
        supplied = np.random.rand(n, n) + 1j * np.random.rand(n, n)
        self.psi_in = supplied.astype(np.complex64)
        self.psi_in *= self.support

    @staticmethod
    def _fft2(data_in, inverse=False):
        """``phasing.py:1077-1085`` including the redundant transform."""
        if inverse:
            tmp = np.fft.ifft2(data_in)  # noqa: F841  (as written)
            data_out = np.fft.ifft2(data_in)
        else:
            tmp = np.fft.ifft2(data_in)  # noqa: F841  (as written)
            data_out = np.fft.fft2(data_in)
        return data_out.astype(np.complex64)

    def _fourier_constraint(self, data):
        """``phasing.py:1187-1228``."""
        self.psi_fourier = self._fft2(data)
        fd = np.abs(self.psi_fourier)
        m, pf = self.modulus, self.psi_fourier
        self.psi_fourier = m * pf / np.abs(fd)
        return self._fft2(self.psi_fourier, inverse=True)

    def iteration(self, alg):
        """``phasing.py:1331-1353``."""
        beta = self.beta
        support = self.support
        self.psi_out = self._fourier_constraint(self.psi_in).astype(np.complex64)
        psi_in, psi_out = self.psi_in, self.psi_out
        if alg == "hio":
            new = (1 - support) * (psi_in - beta * psi_out) + support * psi_out
        elif alg == "er":
            new = support * psi_in
        elif alg == "raar":
            new = (1 - support) * (beta * psi_in + (1 - 2 * beta) * psi_out) + support * psi_out
        elif alg == "sf":
            new = (1 - support) * (-1 * psi_in) + support * psi_out
        elif alg == "dm":
            gamma_m = 1.0 / beta
            gamma_s = -gamma_m
            tmp = support * psi_in - gamma_m * (1 - support) * psi_in
            tmp = self._fourier_constraint(tmp)
            t1 = 2 * psi_out - beta * tmp + beta * ((1 + gamma_s) * psi_out - gamma_s * psi_in)
            t2 = psi_in - beta * psi_out
            new = support * t1 + (1 - support) * t2
        else:
            raise AssertionError("real space enforcement algorithm %s is unknown" % alg)
        self.psi_in = new.astype(np.complex64)

    def iterate(self, algs):
        for a in algs:
            self.iteration(a)
        return self.psi_in
