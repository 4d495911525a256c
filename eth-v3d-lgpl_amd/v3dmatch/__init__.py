"""Host package of the gfx950 SIFT putative matcher.

`SIFT_Matching` mirrors the C++ class V3D_CUDA::SIFT_Matching
(V3D/CUDA/v3d_cudamatching.h:15-67) over the C ABI of libv3dmatch.so (include/v3dm.h).
`synth` generates the synthetic SIFT-like descriptor sets used by tests and benchmarks.
"""
from ._native import V3DMError, lib_path, load_library  # noqa: F401
from .matching import (  # noqa: F401
    FALLBACK_BINOMIAL, FALLBACK_THRESHOLD, PREC_BF16, PREC_FP8, PREC_FP32, PREC_LEARNED, SIFT_Matching,
)
from . import synth  # noqa: F401
