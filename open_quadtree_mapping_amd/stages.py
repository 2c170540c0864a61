"""Stage-level calls of the per-keyframe depth pipeline on torch device tensors.

Each function is a thin wrapper over one C-ABI entry point of libqm.so
(include/qm.h); torch only provides the device memory and the stream. They
mirror the reference's stage functions:

  quadtree        quadtree_image_kernal        src/quadtree.cu:17-77
  cost_volume     image_to_cost + normalize    src/depth_extract.cu:135-244
  bp_extract      bp_extract                   src/bpextract.cu:37-395
  upsample        global_upsample              src/depth_upsample.cu:20-302
  fuse            fuse_transform/hole_filling/fuse_currentmap  src/depth_fusion.cu:48-213
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, default_params, device_empty, lib, qm_params

__all__ = ["quadtree", "cost_volume", "bp_extract", "upsample", "upsample_general", "fuse", "default_params"]


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _need(t, dtype, shape=None, name="tensor"):
    if t is None:
        return
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")


def quadtree(img: torch.Tensor) -> torch.Tensor:
    """uint8 HxW keyframe -> uint8 HxW quadtree level in {2, 3, 4}."""
    _need(img, torch.uint8, name="img")
    H, W = img.shape
    out = device_empty((H, W), torch.uint8, img.device)
    check(lib().qm_quadtree(_ptr(img), W, H, W, _ptr(out), _stream()))
    return out


def cost_volume(key, refs, T_ref_key, cam, levels, params: qm_params | None = None, want_count=False):
    """key, refs[4]: uint8 HxW; T_ref_key: (4, 3, 4) float32 host array; cam (fx, fy, cx, cy).
    Returns the (H/4, W/4, N) cost volume (and the int32 node count)."""
    p = params or default_params()
    _need(key, torch.uint8, name="key")
    H, W = key.shape
    if len(refs) != 4:
        raise ValueError("need four references")
    for r in refs:
        _need(r, torch.uint8, (H, W), name="ref")
    _need(levels, torch.uint8, (H, W), name="levels")
    T = np.ascontiguousarray(T_ref_key, dtype=np.float32).reshape(48)
    camv = np.ascontiguousarray(cam, dtype=np.float32).reshape(4)
    cost = device_empty((H // 4, W // 4, p.n_hyp), torch.float32, key.device)
    count = device_empty((H // 4, W // 4), torch.int32, key.device) if want_count else None
    refp = (C.c_void_p * 4)(*[r.data_ptr() for r in refs])
    check(lib().qm_cost_volume(_ptr(key), refp, W, H, W, T.ctypes.data_as(C.c_void_p),
                               camv.ctypes.data_as(C.c_void_p), C.byref(p), _ptr(levels), _ptr(cost),
                               _ptr(count), _stream()))
    return (cost, count) if want_count else cost


def bp_extract(cost, W, H, levels, params: qm_params | None = None, want_messages=False):
    """Returns (depth_coarse (H/4, W/4) float32, label int32 (-1 off-node)[, messages (4, H/4, W/4, N)])."""
    p = params or default_params()
    _need(cost, torch.float32, (H // 4, W // 4, p.n_hyp), name="cost")
    _need(levels, torch.uint8, (H, W), name="levels")
    dev = cost.device
    depth = device_empty((H // 4, W // 4), torch.float32, dev)
    label = device_empty((H // 4, W // 4), torch.int32, dev)
    msg = device_empty((4, H // 4, W // 4, p.n_hyp), torch.float32, dev) if want_messages else None
    check(lib().qm_bp_extract(_ptr(cost), W, H, _ptr(levels), C.byref(p), _ptr(depth), _ptr(label), _ptr(msg),
                              _stream()))
    return (depth, label, msg) if want_messages else (depth, label)


def upsample(sparse, img, params: qm_params | None = None):
    """Reference global upsample: sparse HxW float32, weights from the uint8 keyframe."""
    p = params or default_params()
    _need(img, torch.uint8, name="img")
    H, W = img.shape
    _need(sparse, torch.float32, (H, W), name="sparse")
    out = device_empty((H, W), torch.float32, img.device)
    check(lib().qm_upsample(_ptr(sparse), _ptr(img), W, H, W, C.byref(p), _ptr(out), _stream()))
    return out


def upsample_general(d, w_row, w_col, conf=None, lam=10.0, row_seg=None, col_seg=None, want_rows=False):
    """Generalised upsample (rhs_x = conf*d, rhs_n = conf); w_row H x (W+1), w_col (H+1) x W."""
    _need(d, torch.float32, name="d")
    H, W = d.shape
    _need(w_row, torch.float32, (H, W + 1), name="w_row")
    _need(w_col, torch.float32, (H + 1, W), name="w_col")
    _need(conf, torch.float32, (H, W), name="conf")
    out = device_empty((H, W), torch.float32, d.device)
    rx = device_empty(out.shape, out.dtype, out.device) if want_rows else None
    rn = device_empty(out.shape, out.dtype, out.device) if want_rows else None
    check(lib().qm_upsample_general(_ptr(d), _ptr(conf), _ptr(w_row), _ptr(w_col), W, H, float(lam),
                                    W if row_seg is None else int(row_seg), H if col_seg is None else int(col_seg),
                                    _ptr(out), _ptr(rx), _ptr(rn), _stream()))
    return (out, rx, rn) if want_rows else out


def fuse(seeds, depth_in, cam, T_last_to_cur, want_table=False):
    """seeds (H, W, 4) float32 is warped IN PLACE; returns (depth_out, seeds_out[, table])."""
    _need(depth_in, torch.float32, name="depth_in")
    H, W = depth_in.shape
    _need(seeds, torch.float32, (H, W, 4), name="seeds")
    camv = np.ascontiguousarray(cam, dtype=np.float32).reshape(4)
    T = np.ascontiguousarray(T_last_to_cur, dtype=np.float32).reshape(12)
    dout = device_empty((H, W), torch.float32, seeds.device)
    sout = device_empty((H, W, 4), torch.float32, seeds.device)
    table = device_empty((H, W), torch.int32, seeds.device) if want_table else None
    check(lib().qm_fuse(_ptr(seeds), _ptr(depth_in), W, H, camv.ctypes.data_as(C.c_void_p),
                        T.ctypes.data_as(C.c_void_p), _ptr(dout), _ptr(sout), _ptr(table), _stream()))
    return (dout, sout, table) if want_table else (dout, sout)


def point_cloud(depth, ref_img, cam, T_world_ref):
    """World-frame points (n, 4) = (x, y, z, intensity) of a depth map, in
    raster order, for pixels whose depth is not below 0.1 (qm_point_cloud)."""
    _need(depth, torch.float32, name="depth")
    H, W = depth.shape
    _need(ref_img, torch.uint8, (H, W), name="ref_img")
    camv = np.ascontiguousarray(cam, dtype=np.float32).reshape(4)
    T = np.ascontiguousarray(T_world_ref, dtype=np.float32).reshape(12)
    out = device_empty((H * W, 4), torch.float32, depth.device)
    n = C.c_int()
    check(lib().qm_point_cloud(_ptr(depth), _ptr(ref_img), W, H, camv.ctypes.data_as(C.c_void_p),
                               T.ctypes.data_as(C.c_void_p), _ptr(out), C.byref(n), _stream()))
    return out[: n.value]


def upsample_local(sparse, img):
    """The reference's 7x7 bilateral alternate to the global upsample (qm_upsample_local)."""
    _need(img, torch.uint8, name="img")
    H, W = img.shape
    _need(sparse, torch.float32, (H, W), name="sparse")
    out = device_empty((H, W), torch.float32, img.device)
    check(lib().qm_upsample_local(_ptr(sparse), _ptr(img), W, H, W, _ptr(out), _stream()))
    return out


def quadtree_prior(prior_depth, levels):
    """quadtree_depth_kernal: raises the levels (in place) where the prior
    inverse depth is flat over a block (qm_quadtree_prior)."""
    H, W = levels.shape
    _need(levels, torch.uint8, name="levels")
    _need(prior_depth, torch.float32, (H, W), name="prior_depth")
    check(lib().qm_quadtree_prior(_ptr(prior_depth), W, H, _ptr(levels), _stream()))
    return levels


def prior_to_cost(prior, levels, cost, params: qm_params | None = None, count=None):
    """prior_to_cost on a cost volume of cost_volume (in place): prior is
    (H, W, 2) float32 (inverse depth, variance) (qm_prior_to_cost)."""
    p = params or default_params()
    _need(levels, torch.uint8, name="levels")
    H, W = levels.shape
    _need(prior, torch.float32, (H, W, 2), name="prior")
    _need(cost, torch.float32, (H // 4, W // 4, p.n_hyp), name="cost")
    if count is not None:
        _need(count, torch.int32, (H // 4, W // 4), name="count")
    check(lib().qm_prior_to_cost(_ptr(prior), W, H, _ptr(levels), C.byref(p), _ptr(cost), _ptr(count), _stream()))
    return cost
