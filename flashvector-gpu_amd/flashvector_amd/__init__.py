"""MI355X-native query-scan path of FlashVector (exact flat L2/IP/cosine scan,
reduced-precision MFMA flat scan with exact rerank, IVF-PQ ADC list scan), all
fused with top-k selection, behind the reference's C ABI.

The HIP library is loaded lazily by flashvector_amd.lib on first use; the
pure-host modules (synthetic workloads, table construction) import without it.
"""
__version__ = "0.1.0"


def __getattr__(name):
    # `from flashvector_amd import FlashVectorGPU` without importing the
    # library-backed modules at package import.
    if name == "FlashVectorGPU":
        from .flashvector_gpu import FlashVectorGPU
        return FlashVectorGPU
    raise AttributeError(name)
