"""CPU oracle for the per-keyframe depth pipeline -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over oracle/libqm_oracle.so (built from oracle/qm_oracle.cpp by
oracle/Makefile). Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py may import this package; it is the checker, never the product path.

Each function follows a reference file (see qm_oracle.cpp's header):
  quadtree           src/quadtree.cu:17-77
  cost_volume        src/depth_extract.cu:135-244
  bp_extract         src/bpextract.cu:37-395
  weights/upsample   src/depth_upsample.cu:20-302
  fuse               src/depth_fusion.cu:13-213
  keyframe           src/seed_matrix.cu:448-597
  semi-dense seeds   src/lsd_semidense.cu (qm_semidense.cpp)
  prior alternates   src/quadtree.cu:79-154, src/depth_extract.cu:23-47 (qm_prior.cpp)
  local upsample     src/depth_upsample.cu:304-370
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# QM_ORACLE_LIB names another build of the oracle (used to check that the
# oracle's pins fail on a build with a deliberately planted mistake)
_LIB_PATH = os.environ.get("QM_ORACLE_LIB") or os.path.join(_HERE, "libqm_oracle.so")
_lib = None

P = C.c_void_p
I = C.c_int
F = C.c_float
D = C.c_double


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "all"], cwd=_HERE)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.qmo_quadtree.argtypes = [P, I, I, I, P]
        L.qmo_depth_table_f32.argtypes = [I, D, D, P]
        L.qmo_depth_table_f64.argtypes = [I, D, D, P]
        L.qmo_cost_f32.argtypes = [P, P, I, I, I, P, F, F, F, F, I, P, P, P, P, I]
        L.qmo_cost_f64.argtypes = [P, P, I, I, I, P, F, F, F, F, I, P, P, P, P, I]
        L.qmo_bp_extract_f32.argtypes = [P, I, I, P, I, F, F, P, D, D, P, P, P]
        L.qmo_bp_extract_f64.argtypes = [P, I, I, P, I, F, F, P, D, D, P, P, P]
        L.qmo_weights_f32.argtypes = [P, I, I, I, P, P]
        L.qmo_weights_f64.argtypes = [P, I, I, I, P, P]
        L.qmo_weight_of_delta_f32.argtypes = [I]
        L.qmo_weight_of_delta_f32.restype = F
        L.qmo_upsample_f32.argtypes = [I, I, P, P, P, P, D, I, I, P, P, P, I]
        L.qmo_upsample_f64.argtypes = [I, I, P, P, P, P, D, I, I, P, P, P, I]
        L.qmo_fuse_f32.argtypes = [I, I, F, F, F, F, P, P, P, P, P, P]
        L.qmo_fuse_f64.argtypes = [I, I, F, F, F, F, P, P, P, P, P, P]
        L.qmo_fuse_update_f32.argtypes = [I, I, P, P, P, P, P]
        L.qmo_fuse_update_f64.argtypes = [I, I, P, P, P, P, P]
        L.qmo_compose_sparse_f32.argtypes = [I, I, P, P, P, P]
        L.qmo_compose_sparse_f64.argtypes = [I, I, P, P, P, P]
        L.qmo_keyframe_f32.argtypes = [P, P, I, I, P, F, F, F, F, I, D, D, P, P, P, P, P, P, P, P, P, P, I, I]
        L.qmo_sd_gradient.argtypes = [P, I, I, I, P]
        L.qmo_sd_init.argtypes = [P, I, I, P]
        L.qmo_sd_step.argtypes = [P, P, P, P, P, I, I, I, F, F, F, F, P, P, I, P, P]
        L.qmo_sd_search_point.argtypes = [I, I, I, I, F, F, F, F, F, F, F, F, F, F, P, P, P, P, P, P]
        L.qmo_sd_search_point.restype = F
        L.qmo_sd_fill_holes.argtypes = [P, I, I, P]
        L.qmo_sd_tex_bilinear.argtypes = [P, I, I, F, F]
        L.qmo_sd_tex_bilinear.restype = F
        L.qmo_sd_regularize.argtypes = [P, I, I, I]
        L.qmo_sd_new_keyframe.argtypes = [P, I, I, F, F, F, F, P, P, P, P, P, P]
        L.qmo_quadtree_prior.argtypes = [P, I, I, P]
        L.qmo_prior_to_cost.argtypes = [P, I, I, P, I, D, D, P, P]
        L.qmo_set_threads.argtypes = [I]
        L.qmo_get_threads.restype = I
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _ptr_array(arrs):
    arr = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arr


def _dt(fp64):
    return np.float64 if fp64 else np.float32


def set_threads(n: int):
    lib().qmo_set_threads(int(n))


def get_threads() -> int:
    return lib().qmo_get_threads()


def quadtree(img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.uint8)
    H, W = img.shape
    out = np.zeros((H, W), np.uint8)
    lib().qmo_quadtree(_p(img), W, H, W, _p(out))
    return out


def depth_table(N: int, imin: float, imax: float, fp64: bool = False) -> np.ndarray:
    out = np.zeros(N, _dt(fp64))
    (lib().qmo_depth_table_f64 if fp64 else lib().qmo_depth_table_f32)(N, imin, imax, _p(out))
    return out


COST_REFERENCE, COST_FUSED = 0, 1  # projection arithmetic (the library's QM_COST_*)
UPSAMPLE_THOMAS, UPSAMPLE_SCAN = 0, 1  # line solver of the upsample (the library's QM_UPSAMPLE_*)


def cost_volume(key, refs, T, cam, N, imin, imax, levels, fp64=False, arith=COST_REFERENCE):
    """refs: 4 uint8 images; T: (4,3,4) float32 T_ref_key. Returns (cost HcxWcxN, count HcxWc).
    arith: COST_REFERENCE (the reference's operation order) or COST_FUSED
    (hoisted per-cell products, fused multiply-adds, one reciprocal)."""
    key = np.ascontiguousarray(key, dtype=np.uint8)
    refs = [np.ascontiguousarray(r, dtype=np.uint8) for r in refs]
    assert len(refs) == 4
    H, W = key.shape
    T = np.ascontiguousarray(T, dtype=np.float32).reshape(4, 12)
    table = depth_table(N, imin, imax, fp64)
    levels = np.ascontiguousarray(levels, dtype=np.uint8)
    cost = np.zeros((H // 4, W // 4, N), _dt(fp64))
    count = np.zeros((H // 4, W // 4), np.int32)
    fn = lib().qmo_cost_f64 if fp64 else lib().qmo_cost_f32
    rp = _ptr_array(refs)
    fn(_p(key), rp, W, H, W, _p(T), cam.fx, cam.fy, cam.cx, cam.cy, N, _p(table), _p(levels), _p(cost), _p(count),
       int(arith))
    return cost, count


def bp_extract(cost, W, H, levels, N, imin, imax, P1=0.003, P2=0.01, iters=(4, 10, 10, 10), fp64=False,
               want_messages=False):
    """Returns (depth_coarse HcxWc, label HcxWc int32 (-1: not a node)[, messages 4xHcxWcxN])."""
    dt = _dt(fp64)
    cost = np.ascontiguousarray(cost, dtype=dt)
    levels = np.ascontiguousarray(levels, dtype=np.uint8)
    Hc, Wc = H // 4, W // 4
    depth = np.zeros((Hc, Wc), dt)
    label = np.zeros((Hc, Wc), np.int32)
    it = np.array(iters, np.int32)
    msg = np.zeros((4, Hc, Wc, N), dt) if want_messages else None
    fn = lib().qmo_bp_extract_f64 if fp64 else lib().qmo_bp_extract_f32
    fn(_p(cost), W, H, _p(levels), N, P1, P2, _p(it), imin, imax, _p(depth), _p(label), _p(msg))
    return (depth, label, msg) if want_messages else (depth, label)


def weights(img, fp64=False):
    """Reference smoothness weights: (w_row H x (W+1), w_col (H+1) x W)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    H, W = img.shape
    dt = _dt(fp64)
    wr = np.zeros((H, W + 1), dt)
    wc = np.zeros((H + 1, W), dt)
    (lib().qmo_weights_f64 if fp64 else lib().qmo_weights_f32)(_p(img), W, H, W, _p(wr), _p(wc))
    return wr, wc


def weight_lut():
    return np.array([lib().qmo_weight_of_delta_f32(i) for i in range(256)], np.float32)


def upsample(d, w_row, w_col, conf=None, lam=10.0, row_seg=None, col_seg=None, fp64=False, want_rows=False,
             solver=UPSAMPLE_THOMAS):
    """The global upsample of src/depth_upsample.cu:20-302, generalised
    (confidences, explicit weights). solver UPSAMPLE_THOMAS: the reference's
    Thomas recurrences and order; UPSAMPLE_SCAN: the segmented-scan solve of the
    same systems (64 lanes per line, lines up to 64 * 24 steps; longer lines use
    Thomas)."""
    dt = _dt(fp64)
    d = np.ascontiguousarray(d, dtype=dt)
    H, W = d.shape
    w_row = np.ascontiguousarray(w_row, dtype=dt)
    w_col = np.ascontiguousarray(w_col, dtype=dt)
    conf = None if conf is None else np.ascontiguousarray(conf, dtype=dt)
    row_seg = W if row_seg is None else row_seg
    col_seg = H if col_seg is None else col_seg
    out = np.zeros((H, W), dt)
    rx = np.zeros((H, W), dt)
    rn = np.zeros((H, W), dt)
    fn = lib().qmo_upsample_f64 if fp64 else lib().qmo_upsample_f32
    fn(W, H, _p(d), _p(conf), _p(w_row), _p(w_col), lam, row_seg, col_seg, _p(out), _p(rx), _p(rn), int(solver))
    return (out, rx, rn) if want_rows else out


def ref_segments(W, H):
    """Chain split points of the reference upsample: 16*floor(W/16), 8*floor(H/8)."""
    return 16 * (W // 16), 8 * (H // 8)


def fuse(seeds, depth_in, T_last_to_cur, cam, fp64=False):
    """Returns (warped seeds (in-place copy), depth_out, seeds_out, table)."""
    dt = _dt(fp64)
    H, W = depth_in.shape
    s = np.array(seeds, dtype=dt, copy=True).reshape(H, W, 4)
    din = np.ascontiguousarray(depth_in, dtype=dt)
    T = np.ascontiguousarray(T_last_to_cur, dtype=np.float32).reshape(12)
    dout = np.zeros((H, W), dt)
    sout = np.zeros((H, W, 4), dt)
    table = np.zeros((H, W), np.int32)
    fn = lib().qmo_fuse_f64 if fp64 else lib().qmo_fuse_f32
    fn(W, H, cam.fx, cam.fy, cam.cx, cam.cy, _p(T), _p(s), _p(din), _p(dout), _p(sout), _p(table))
    return s, dout, sout, table


def fuse_update(table, seeds_warped, depth_in, fp64=False):
    """The per-pixel update alone, given the hole-filled table and the warped
    seeds. Returns (depth_out, seeds_out)."""
    dt = _dt(fp64)
    H, W = depth_in.shape
    tb = np.ascontiguousarray(table, dtype=np.int32)
    s = np.ascontiguousarray(seeds_warped, dtype=dt).reshape(H, W, 4)
    din = np.ascontiguousarray(depth_in, dtype=dt)
    dout = np.zeros((H, W), dt)
    sout = np.zeros((H, W, 4), dt)
    fn = lib().qmo_fuse_update_f64 if fp64 else lib().qmo_fuse_update_f32
    fn(W, H, _p(tb), _p(s), _p(din), _p(dout), _p(sout))
    return dout, sout


def compose_sparse(W, H, depth_coarse, label, prefill=None, fp64=False):
    dt = _dt(fp64)
    out = np.zeros((H, W), dt)
    dc = np.ascontiguousarray(depth_coarse, dtype=dt)
    lb = np.ascontiguousarray(label, dtype=np.int32)
    pf = None if prefill is None else np.ascontiguousarray(prefill, dtype=dt)
    (lib().qmo_compose_sparse_f64 if fp64 else lib().qmo_compose_sparse_f32)(W, H, _p(dc), _p(lb), _p(pf), _p(out))
    return out


def keyframe(key, refs, T, cam, N, imin, imax, seeds, T_last_to_cur, prefill=None, prior=None, prior_depth=None,
             arith=COST_REFERENCE, solver=UPSAMPLE_THOMAS):
    """Full FP32 keyframe. seeds (HxWx4 float32) is updated in place.
    Returns dict(depth, sparse, levels, label, stage_ms)."""
    key = np.ascontiguousarray(key, dtype=np.uint8)
    H, W = key.shape
    refs = [np.ascontiguousarray(r, dtype=np.uint8) for r in refs]
    T = np.ascontiguousarray(T, dtype=np.float32).reshape(4, 12)
    assert seeds.dtype == np.float32 and seeds.flags.c_contiguous
    Tl = np.ascontiguousarray(T_last_to_cur, dtype=np.float32).reshape(12)
    pf = None if prefill is None else np.ascontiguousarray(prefill, dtype=np.float32)
    pr = None if prior is None else np.ascontiguousarray(prior, dtype=np.float32)
    pd = None if prior_depth is None else np.ascontiguousarray(prior_depth, dtype=np.float32)
    depth = np.zeros((H, W), np.float32)
    sparse = np.zeros((H, W), np.float32)
    levels = np.zeros((H, W), np.uint8)
    label = np.zeros((H // 4, W // 4), np.int32)
    ms = np.zeros(5, np.float64)
    lib().qmo_keyframe_f32(_p(key), _ptr_array(refs), W, H, _p(T), cam.fx, cam.fy, cam.cx, cam.cy, N, imin, imax,
                           _p(pf), _p(pr), _p(pd), _p(seeds), _p(Tl), _p(depth), _p(sparse), _p(levels), _p(label),
                           _p(ms), int(arith), int(solver))
    return dict(depth=depth, sparse=sparse, levels=levels, label=label, stage_ms=ms)


def upsample_local(sparse, img):
    """local_upsample / depth_interpolate (src/depth_upsample.cu:304-370) in
    float32, accumulation in the reference's order (node column a outer, row b
    inner), exp as (float)exp((double)x)."""
    sparse = np.asarray(sparse, np.float32)
    img = np.asarray(img, np.uint8)
    H, W = img.shape
    f32 = np.float32
    yy, xx = np.mgrid[0:H, 0:W]
    me = img.astype(f32)
    acc = np.zeros((H, W), f32)
    wsum = np.zeros((H, W), f32)
    for a in range(7):
        for b in range(7):
            cx = (xx // 4 + a - 3) * 4
            cy = (yy // 4 + b - 3) * 4
            inb = (cx >= 0) & (cx < W) & (cy >= 0) & (cy < H)
            cxc, cyc = np.clip(cx, 0, W - 1), np.clip(cy, 0, H - 1)
            dep = np.where(inb, sparse[cyc, cxc], f32(0))
            use = inb & (dep.astype(np.float64) > 0.0)
            di = me - img[cyc, cxc].astype(f32)
            sx, sy = (xx - cx).astype(f32), (yy - cy).astype(f32)
            sw = sx * sx + sy * sy
            iw = di * di
            arg = (-sw / f32(100.0)) - iw / f32(50.0)
            w = np.exp(arg.astype(np.float64)).astype(f32)
            w = np.where(use, w, f32(0))
            dep = np.where(use, dep, f32(0))
            skip = inb & ~use
            acc = np.where(skip, acc, acc + w * dep)
            wsum = np.where(skip, wsum, wsum + w)
    with np.errstate(invalid="ignore", divide="ignore"):
        return acc / wsum


def quadtree_prior(prior_depth, levels):
    """quadtree_depth_kernal (src/quadtree.cu:79-154): returns max(levels,
    depth level) for an (H, W) prior depth map (qm_prior.cpp)."""
    out = np.array(levels, np.uint8, copy=True, order="C")
    H, W = out.shape
    lib().qmo_quadtree_prior(_p(np.ascontiguousarray(prior_depth, np.float32)), W, H, _p(out))
    return out


def prior_to_cost(prior, levels, cost, count, N, imin, imax):
    """prior_to_cost (src/depth_extract.cu:23-47) on a cost volume of
    cost_volume(), renormalised by add_num; returns new (cost, count)."""
    cost = np.array(cost, np.float32, copy=True, order="C")
    count = np.array(count, np.int32, copy=True, order="C")
    levels = np.ascontiguousarray(levels, np.uint8)
    H, W = levels.shape
    lib().qmo_prior_to_cost(_p(np.ascontiguousarray(prior, np.float32)), W, H, _p(levels), N, imin, imax,
                            _p(cost), _p(count))
    return cost, count


def point_cloud(depth, ref_img, cam, T_world_ref):
    """Publisher::publishPointCloud (src/publisher.cpp:129-164) in float32:
    raster order, pixels with depth not below 0.1 (NaN kept), rows
    (x, y, z, intensity) of T_world_ref * (normalize(K^-1 (x, y, 1)) * d)."""
    depth = np.asarray(depth, np.float32)
    keep = ~(depth.astype(np.float64) < 0.1)
    ys, xs = np.nonzero(keep)
    f32 = np.float32
    rx = (xs.astype(f32) - f32(cam.cx)) / f32(cam.fx)
    ry = (ys.astype(f32) - f32(cam.cy)) / f32(cam.fy)
    rz = np.ones_like(rx)
    inv = f32(1.0) / np.sqrt((rx * rx + ry * ry) + rz * rz)
    d = depth[ys, xs]
    p = [(rx * inv) * d, (ry * inv) * d, (rz * inv) * d]
    T = np.asarray(T_world_ref, f32).reshape(12)
    out = np.empty((len(xs), 4), f32)
    for r in range(3):
        out[:, r] = ((T[4 * r] * p[0] + T[4 * r + 1] * p[1]) + T[4 * r + 2] * p[2]) + T[4 * r + 3]
    out[:, 3] = np.asarray(ref_img)[ys, xs].astype(f32)
    return out


# ---------------------------------------------------------------------------
# Semi-dense keyframe seeds (src/seed_init.cu, src/lsd_semidense.cu,
# src/seed_matrix.cu:211-304,399-446): the sparse map's pre-fill.
# ---------------------------------------------------------------------------
def sd_gradient(img):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    H, W = img.shape
    g = np.zeros((H, W, 2), np.float32)
    lib().qmo_sd_gradient(_p(img), W, H, W, _p(g))
    return g


def sd_init(grad):
    H, W = grad.shape[:2]
    seeds = np.zeros((H, W, 8), np.float32)
    lib().qmo_sd_init(_p(np.ascontiguousarray(grad, np.float32)), W, H, _p(seeds))
    return seeds


def sd_step(seeds, key, key_grad, inc, inc_grad, cam, T_key_world, T_inc_world, switch_keyframe, want_prior=False):
    """One frame of the semi-dense module; seeds (H, W, 8) float32 is updated
    in place. Returns the sparse-map pre-fill (H, W) [and the (H, W, 2)
    (inverse depth, variance) prior of src/lsd_semidense.cu:241-294]."""
    assert seeds.dtype == np.float32 and seeds.flags.c_contiguous
    key = np.ascontiguousarray(key, np.uint8)
    inc = np.ascontiguousarray(inc, np.uint8)
    H, W = key.shape
    prefill = np.zeros((H, W), np.float32)
    prior = np.zeros((H, W, 2), np.float32) if want_prior else None
    Tk = np.ascontiguousarray(T_key_world, np.float32).reshape(12)
    Ti = np.ascontiguousarray(T_inc_world, np.float32).reshape(12)
    lib().qmo_sd_step(_p(seeds), _p(key), _p(np.ascontiguousarray(key_grad, np.float32)), _p(inc),
                      _p(np.ascontiguousarray(inc_grad, np.float32)), W, H, W, cam.fx, cam.fy, cam.cx, cam.cy,
                      _p(Tk), _p(Ti), int(bool(switch_keyframe)), _p(prefill), _p(prior))
    return (prefill, prior) if want_prior else prefill


def sd_search_point(x, y, epx, epy, gradient_max, my_gradient, min_idep, max_idep, cam, k2i, key, key_grad, inc,
                    inc_grad):
    """search_point (src/lsd_semidense.cu:529-753) of one key pixel. Returns
    (score or failure code -1/-2/-3, inverse depth, variance, epipolar length)."""
    key = np.ascontiguousarray(key, np.uint8)
    H, W = key.shape
    out = np.zeros(3, np.float32)
    err = lib().qmo_sd_search_point(int(x), int(y), W, H, epx, epy, gradient_max, my_gradient, min_idep, max_idep,
                                    cam.fx, cam.fy, cam.cx, cam.cy,
                                    _p(np.ascontiguousarray(k2i, np.float32).reshape(12)), _p(key),
                                    _p(np.ascontiguousarray(key_grad, np.float32)),
                                    _p(np.ascontiguousarray(inc, np.uint8)),
                                    _p(np.ascontiguousarray(inc_grad, np.float32)), _p(out))
    return float(err), float(out[0]), float(out[1]), float(out[2])


def sd_tex_bilinear(img, u, v):
    """The epipolar search's bilinear texture fetch at (u, v), texel centres at +0.5."""
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    return float(lib().qmo_sd_tex_bilinear(_p(img), W, H, float(u), float(v)))


def sd_fill_holes(seeds, inc_grad):
    """fill_holes (src/lsd_semidense.cu:345-387) on a copy of (H, W, 8) seeds."""
    s = np.array(seeds, np.float32, copy=True, order="C")
    H, W = s.shape[:2]
    lib().qmo_sd_fill_holes(_p(s), W, H, _p(np.ascontiguousarray(inc_grad, np.float32)))
    return s


def sd_regularize(seeds, remove_occlusions=False):
    """regularize (src/lsd_semidense.cu:296-344) on a copy of (H, W, 8) seeds."""
    s = np.array(seeds, np.float32, copy=True, order="C")
    H, W = s.shape[:2]
    lib().qmo_sd_regularize(_p(s), W, H, int(bool(remove_occlusions)))
    return s


def sd_new_keyframe(old_seeds, cam, T_old_to_new, key, key_grad, inc, inc_grad):
    """Propagation into a new keyframe (src/lsd_semidense.cu:64-181 as driven by
    src/seed_matrix.cu:247-304): returns the new (H, W, 8) seeds."""
    old = np.ascontiguousarray(old_seeds, np.float32)
    H, W = old.shape[:2]
    out = np.zeros_like(old)
    lib().qmo_sd_new_keyframe(_p(old), W, H, cam.fx, cam.fy, cam.cx, cam.cy,
                              _p(np.ascontiguousarray(T_old_to_new, np.float32).reshape(12)),
                              _p(np.ascontiguousarray(key, np.uint8)), _p(np.ascontiguousarray(key_grad, np.float32)),
                              _p(np.ascontiguousarray(inc, np.uint8)), _p(np.ascontiguousarray(inc_grad, np.float32)),
                              _p(out))
    return out


def axis_cos_baseline(T_a_world, T_b_world):
    """cos of the angle between the optical axes and the distance between the
    camera centres of two world->camera poses, in float32 as the reference."""
    A = _se3_inv(T_a_world).reshape(12)
    B = _se3_inv(T_b_world).reshape(12)
    cosv = (A[2] * B[2] + A[6] * B[6]) + A[10] * B[10]
    dx, dy, dz = A[3] - B[3], A[7] - B[7], A[11] - B[11]
    return np.float32(cosv), np.sqrt((dx * dx + dy * dy) + dz * dz, dtype=np.float32)


def need_switch_keyframe(T_key_world, T_inc_world):
    """src/seed_matrix.cu:200-209 (thresholds compared in double)."""
    c, b = axis_cos_baseline(T_key_world, T_inc_world)
    return bool(float(c) < 0.86 or float(b) > 0.5)


# ---------------------------------------------------------------------------
# Frame-sequence state machine (src/seed_matrix.cu:106-198, src/depthmap.cpp:42-55)
# without the semi-dense seeds (their pre-fill of the sparse map is zero here).
# ---------------------------------------------------------------------------
def _se3_inv(T):
    d = np.asarray(T, np.float32).reshape(12)
    r = np.zeros(12, np.float32)
    r[0], r[1], r[2], r[4], r[5], r[6], r[8], r[9], r[10] = d[0], d[4], d[8], d[1], d[5], d[9], d[2], d[6], d[10]
    r[3] = ((-d[0]) * d[3] - d[4] * d[7]) - d[8] * d[11]
    r[7] = ((-d[1]) * d[3] - d[5] * d[7]) - d[9] * d[11]
    r[11] = ((-d[2]) * d[3] - d[6] * d[7]) - d[10] * d[11]
    return r.reshape(3, 4)


def _se3_mul(A, B):
    a = np.asarray(A, np.float32).reshape(12)
    b = np.asarray(B, np.float32).reshape(12)
    r = np.zeros(12, np.float32)
    for row in range(3):
        o = 4 * row
        for col in range(3):
            r[o + col] = (a[o] * b[col] + a[o + 1] * b[4 + col]) + a[o + 2] * b[8 + col]
        r[o + 3] = ((a[o + 3] + a[o] * b[3]) + a[o + 1] * b[7]) + a[o + 2] * b[11]
    return r.reshape(3, 4)


class DepthmapSequence:
    """Oracle of quadmap::Depthmap over a frame sequence (FP32)."""

    def __init__(self, cam, N=64, imin=0.01, imax=2.0, ratio=1, min_refs=11, max_refs=60, ref_cos=0.95,
                 ref_baseline=0.03, semidense=True, prior_cost=False, prior_quadtree=False, arith=COST_REFERENCE):
        self.cam, self.N, self.imin, self.imax, self.ratio = cam, N, imin, imax, ratio
        self.min_refs, self.max_refs = min_refs, max_refs
        self.ref_cos, self.ref_baseline = np.float32(ref_cos), np.float32(ref_baseline)
        self.frame_index = 0
        self.initialized = False
        self.refs = []  # newest first: (image, T_ref_world)
        self.fuse_pose = np.eye(3, 4, dtype=np.float32)
        self.seeds = np.zeros((cam.H, cam.W, 4), np.float32)
        self.depth = self.debug = self.reference = None
        self.T_world_ref = np.eye(3, 4, dtype=np.float32)
        self.semidense = semidense
        self.prior_cost, self.prior_quadtree = prior_cost, prior_quadtree
        self.arith = arith
        self.prior = None
        self.sd_key = self.sd_key_grad = self.sd_key_T = self.sd_seeds = None
        self.prefill = None

    def _need_add_reference(self, T_income):
        if not self.refs:
            return True
        A = _se3_inv(self.refs[0][1]).reshape(12)
        B = _se3_inv(T_income).reshape(12)
        cosv = (A[2] * B[2] + A[6] * B[6]) + A[10] * B[10]
        dx, dy, dz = A[3] - B[3], A[7] - B[7], A[11] - B[11]
        base = np.sqrt((dx * dx + dy * dy) + dz * dz, dtype=np.float32)
        return bool(cosv < self.ref_cos or base > self.ref_baseline)

    def add_frames(self, img, T_curr_world):
        T_curr_world = np.asarray(T_curr_world, np.float32)
        self.frame_index += 1
        W, H = self.cam.W, self.cam.H
        if self.semidense:
            grad = sd_gradient(img)
            if not self.initialized:
                self.sd_key, self.sd_key_grad, self.sd_key_T = img.copy(), grad, T_curr_world.copy()
                self.sd_seeds = sd_init(grad)
            else:
                sw = need_switch_keyframe(self.sd_key_T, T_curr_world)
                out = sd_step(self.sd_seeds, self.sd_key, self.sd_key_grad, img, grad, self.cam,
                              self.sd_key_T, T_curr_world, sw, want_prior=self.prior_cost)
                self.prefill, self.prior = out if self.prior_cost else (out, None)
                if sw:
                    self.sd_key, self.sd_key_grad, self.sd_key_T = img.copy(), grad, T_curr_world.copy()
        if not self.initialized:
            self.initialized = True
            self.depth = np.zeros((H, W), np.float32)
            self.debug = np.zeros((H, W), np.float32)
            self.reference = img.copy()
            self.T_world_ref = _se3_inv(T_curr_world)
            return True
        if self.frame_index % self.ratio != 0:
            return False
        has_output = len(self.refs) >= self.min_refs
        if has_output:
            inv_inc = _se3_inv(T_curr_world)
            T = np.stack([_se3_mul(self.refs[r][1], inv_inc) for r in range(4)])
            T_last = _se3_mul(T_curr_world, _se3_inv(self.fuse_pose))
            res = keyframe(img, [self.refs[r][0] for r in range(4)], T, self.cam, self.N, self.imin, self.imax,
                           self.seeds, T_last, prefill=self.prefill if self.semidense else None,
                           prior=self.prior if self.semidense and self.prior_cost else None,
                           prior_depth=self.prefill if self.semidense and self.prior_quadtree else None,
                           arith=self.arith)
        if self._need_add_reference(T_curr_world):
            self.refs.insert(0, (img.copy(), T_curr_world.copy()))
            if len(self.refs) > self.max_refs:
                self.refs.pop()
        if has_output:
            self.fuse_pose = T_curr_world.copy()
            self.depth, self.debug, self.reference = res["depth"], res["sparse"], img.copy()
            self.T_world_ref = _se3_inv(T_curr_world)
        return has_output
