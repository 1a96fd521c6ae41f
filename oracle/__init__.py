"""CPU oracle for the iterative phase-retrieval path (test infrastructure).

This package is the checker, not the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``als_speckle_amd`` never imports it and has
no CPU fallback.

Modules
-------
phasing_c128
    complex128 / FP64 restatement of the reference's numpy phasing path
    (``speckle/phasing.py``), plus the build-defined modulus-error metric and
    in-loop shrinkwrap.  Every function cites the reference lines it follows.
ref_numpy
    The reference CPU iteration restated as written (complex64 storage,
    ``np.fft`` including the redundant transform per ``_fft2`` call).  Used
    only as the timed CPU baseline and as the "FP32 numpy" drift envelope.

Pinning status of each oracle function is recorded in the module headers.
"""
