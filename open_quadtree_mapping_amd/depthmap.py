"""quadmap::Depthmap for Python callers (include/quadmap/depthmap.h:31-89).

Same method names and argument meaning as the reference class; the state
machine and the GPU pipeline run inside libqm.so (qm_create / qm_add_frame).
Images are host uint8 arrays, poses 3x4 float32 world->camera matrices.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, default_params, lib, qm_params

__all__ = ["Depthmap"]


class Depthmap:
    def __init__(self, width, height, fx, cx, fy, cy, remap_1=None, remap_2=None, semi2dense_ratio=1,
                 params: qm_params | None = None):
        # note the reference order fx, cx, fy, cy (src/depthmap.cpp:19-40)
        self.W, self.H = int(width), int(height)
        self._fx, self._fy, self._cx, self._cy = float(fx), float(fy), float(cx), float(cy)
        self.params = params or default_params()
        self._h = C.c_void_p()
        check(lib().qm_create(self.W, self.H, self._fx, self._fy, self._cx, self._cy, C.byref(self.params),
                              C.byref(self._h)))
        check(lib().qm_set_semi2dense_ratio(self._h, int(semi2dense_ratio)))
        if remap_1 is not None:
            mx = np.ascontiguousarray(remap_1, dtype=np.float32)
            my = np.ascontiguousarray(remap_2, dtype=np.float32)
            if mx.shape != (self.H, self.W) or my.shape != (self.H, self.W):
                raise ValueError("remap maps must be H x W float32 coordinate maps")
            check(lib().qm_set_remap(self._h, mx.ctypes.data_as(C.c_void_p), my.ctypes.data_as(C.c_void_p)))

    def add_frames(self, img_curr, T_curr_world) -> bool:
        img = np.ascontiguousarray(img_curr, dtype=np.uint8)
        if img.ndim != 2:
            raise ValueError("image must be a 2-D uint8 array")
        ih, iw = img.shape
        T = np.ascontiguousarray(T_curr_world, dtype=np.float32).reshape(12)
        has = C.c_int(0)
        check(lib().qm_add_frame(self._h, img.ctypes.data_as(C.c_void_p), iw, ih, iw, T.ctypes.data_as(C.c_void_p),
                                 C.byref(has)))
        return bool(has.value)

    def getDepthmap(self):
        out = np.empty((self.H, self.W), np.float32)
        check(lib().qm_get_depth(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def getDebugmap(self):
        out = np.empty((self.H, self.W), np.float32)
        check(lib().qm_get_debug(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def getReferenceImage(self):
        out = np.empty((self.H, self.W), np.uint8)
        check(lib().qm_get_reference(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def getT_world_ref(self):
        out = np.empty(12, np.float32)
        check(lib().qm_get_T_world_ref(self._h, out.ctypes.data_as(C.c_void_p)))
        return out.reshape(3, 4)

    def getPointCloud(self):
        """(n, 4) float32 rows (x, y, z, intensity) of the last result in the
        world frame (Publisher::publishPointCloud, src/publisher.cpp:129-164)."""
        out = np.empty((self.H * self.W, 4), np.float32)
        n = C.c_int()
        check(lib().qm_get_point_cloud(self._h, out.ctypes.data_as(C.c_void_p), C.byref(n)))
        return out[: n.value].copy()

    def set_semidense(self, on: bool):
        check(lib().qm_set_semidense(self._h, int(bool(on))))

    def set_prior_alternates(self, prior_cost: bool = False, prior_quadtree: bool = False):
        """The reference's prior-driven alternates (prior_to_cost,
        quadtree_depth_kernal) fed by the semi-dense module; before the first frame."""
        check(lib().qm_set_prior_alternates(self._h, int(bool(prior_cost)), int(bool(prior_quadtree))))

    def getFx(self):
        return self._fx

    def getFy(self):
        return self._fy

    def getCx(self):
        return self._cx

    def getCy(self):
        return self._cy

    def getWidth(self):
        return self.W

    def getHeight(self):
        return self.H

    def num_references(self):
        return lib().qm_num_references(self._h)

    def close(self):
        if self._h:
            lib().qm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
