"""MI355X-native iterative phase retrieval behind the ``speckle.phasing`` API.

Modules
-------
phasing   drop-in ``phasing`` class and module functions
device    Context / Plan over torch ROCm buffers (C ABI calls)
masking   host-side mask helpers used by the phasing API
synth     seeded synthetic diffraction data for the workloads
_capi     ctypes binding of libspeckle_hip.so

Importing the package does not touch the GPU; constructing a ``phasing``
object or a ``Context`` does, and fails loudly without one.
"""
__version__ = "0.1.0"
