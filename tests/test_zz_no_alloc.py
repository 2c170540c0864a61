"""qm_fuse_noalloc (static device scratch, no allocation) against qm_fuse,
through the C ABI: bit-identical warped seeds, fused depth, new seeds and warp
table on consecutive calls whose sizes shrink, so the second call runs on the
scratch the first one left behind."""
import ctypes as C

import numpy as np
import pytest
import torch

from open_quadtree_mapping_amd._lib import check, lib
from tests._data import keyframe_set
from open_quadtree_mapping_amd.se3 import relative_pose

pytestmark = pytest.mark.gpu


def case(W, H, seed):
    rng = np.random.default_rng(seed)
    ks = keyframe_set(W, H, 3 if W == 376 else 5)
    seeds = np.zeros((H, W, 4), np.float32)
    seeds[..., 0] = rng.uniform(0.0, 30.0, (H, W))
    seeds[..., 1] = rng.uniform(0.0, 30.0, (H, W))
    seeds[..., 2] = ks["gt"] * rng.uniform(0.95, 1.05, (H, W))
    seeds[..., 3] = rng.uniform(0.001, 5.0, (H, W))
    seeds[rng.random((H, W)) < 0.05, 0] = 0.0
    depth = (ks["gt"] * rng.uniform(0.9, 1.1, (H, W))).astype(np.float32)
    depth[rng.random((H, W)) < 0.02] = np.nan
    depth[rng.random((H, W)) < 0.02] = -1.0
    T = relative_pose(ks["poses"][-1], ks["poses"][-2])
    c = ks["cam"]
    return seeds, depth, (c.fx, c.fy, c.cx, c.cy), T


def call(fn, seeds, depth, cam, T):
    H, W = depth.shape
    s = torch.from_numpy(seeds).cuda()
    d = torch.from_numpy(depth).cuda()
    # outputs start as a sentinel that no result takes (table entries are >= 0)
    dout = torch.full((H, W), -7.0, dtype=torch.float32, device="cuda")
    sout = torch.full((H, W, 4), -7.0, dtype=torch.float32, device="cuda")
    table = torch.full((H, W), -7, dtype=torch.int32, device="cuda")
    camv = np.ascontiguousarray(cam, np.float32)
    Tv = np.ascontiguousarray(T, np.float32).reshape(12)
    p = lambda t: C.c_void_p(t.data_ptr())
    check(fn(p(s), p(d), W, H, camv.ctypes.data_as(C.c_void_p), Tv.ctypes.data_as(C.c_void_p), p(dout), p(sout),
             p(table), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (s, dout, sout, table)]


def test_fuse_noalloc_matches_fuse_on_shrinking_calls():
    L = lib()
    for (W, H, seed) in [(752, 480, 41), (376, 240, 42), (752, 480, 43)]:
        print(f"fuse_noalloc {W}x{H} seed {seed}", flush=True)
        seeds, depth, cam, T = case(W, H, seed)
        want = call(L.qm_fuse, seeds, depth, cam, T)
        got = call(L.qm_fuse_noalloc, seeds, depth, cam, T)
        for name, g, w in zip(("warped seeds", "depth_out", "seeds_out", "table"), got, want):
            assert np.array_equal(g, w, equal_nan=True), f"{W}x{H}: {name} differs"
        assert (want[3] >= 0).all() and (want[1] != -7.0).all() and np.isfinite(want[1]).any()
