"""MI355X-native per-keyframe depth pipeline of QuadtreeMapping.

quadtree -> multi-view cost volume -> hierarchical BP + winner extraction ->
tridiagonal smoothness upsampling -> Gaussian x Beta fusion, as hand-written
gfx950 HIP kernels behind the C ABI of libqm.so (include/qm.h).
"""
from ._lib import LIB_PATH, QmError, default_params, lib, qm_frame_desc, qm_params, release_scratch  # noqa: F401

__all__ = ["LIB_PATH", "QmError", "default_params", "lib", "qm_params", "qm_frame_desc", "release_scratch"]
