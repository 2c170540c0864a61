"""quadmap::Depthmap-level parity: a frame sequence through qm_create /
qm_add_frame (gates, reference deque, fusion state, downloads) against the
oracle's restatement of SeedMatrix::add_frames (src/seed_matrix.cu:106-198)."""
import numpy as np
import pytest

import oracle as O
from open_quadtree_mapping_amd import default_params
from open_quadtree_mapping_amd.depthmap import Depthmap
from open_quadtree_mapping_amd.scene import camera_for, lateral_trajectory, make_scene, render

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ratio", [1, 2])
def test_sequence_matches_oracle(ratio):
    W, H = 376, 240
    cam = camera_for(W, H)
    n = 18 if ratio == 1 else 30
    traj = lateral_trajectory(n, step=0.04, seed=11)
    sc = make_scene(11)
    imgs = [render(sc, cam, T)[0] for T in traj]
    p = default_params(idepth_min=0.02, idepth_max=2.0)
    dm = Depthmap(W, H, cam.fx, cam.cx, cam.fy, cam.cy, semi2dense_ratio=ratio, params=p)
    ref = O.DepthmapSequence(cam, 64, 0.02, 2.0, ratio=ratio)
    outputs = 0
    for k in range(n):
        print(f"depthmap ratio={ratio} frame {k}", flush=True)
        got = dm.add_frames(imgs[k], traj[k])
        want = ref.add_frames(imgs[k], traj[k])
        assert got == want, f"frame {k}: has_result {got} vs {want}"
        assert dm.num_references() == len(ref.refs)
        if got:
            assert np.array_equal(dm.getDebugmap(), ref.debug)
            assert np.array_equal(dm.getDepthmap(), ref.depth, equal_nan=True)
            assert np.array_equal(dm.getReferenceImage(), ref.reference)
            assert np.array_equal(dm.getT_world_ref(), ref.T_world_ref)
            outputs += 1
    assert outputs >= 3
    dm.close()


def test_remap_identity_shift_and_half_pixel():
    W, H = 64, 48
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (H, W)).astype(np.uint8)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    T = np.eye(3, 4, dtype=np.float32)
    outs = {}
    for name, (mx, my) in {"identity": (xx, yy), "shift": (xx + 3.0, yy - 2.0), "half": (xx + 0.5, yy)}.items():
        dm = Depthmap(W, H, 50.0, 32.0, 50.0, 24.0, mx, my, 1)
        assert dm.add_frames(img, T)  # the first frame always reports
        outs[name] = dm.getReferenceImage()
        dm.close()
    assert np.array_equal(outs["identity"], img)
    sh = np.zeros_like(img)
    sh[2:, : W - 3] = img[: H - 2, 3:]
    assert np.array_equal(outs["shift"], sh)
    a = img[:, :-1].astype(np.int64)
    b = img[:, 1:].astype(np.int64)
    half = (a * 16 * 32 * 32 + b * 16 * 32 * 32 + (1 << 14)) >> 15
    assert np.array_equal(outs["half"][:, :-1], half)
    assert np.array_equal(outs["half"][:, -1], (img[:, -1].astype(np.int64) * 16 * 32 * 32 + (1 << 14)) >> 15)


def remap_fixed_point(src, mx, my):
    """OpenCV's 8-bit INTER_LINEAR remap in fixed point: coordinates rounded to
    1/32 pixel (half to even), weights (32-a)(32-b)*32 ..., (sum + 2^14) >> 15,
    constant border 0."""
    SH, SW = src.shape
    ix = np.rint(mx.astype(np.float32) * np.float32(32)).astype(np.int64)
    iy = np.rint(my.astype(np.float32) * np.float32(32)).astype(np.int64)
    x0, y0, a, b = ix >> 5, iy >> 5, ix & 31, iy & 31

    def px(x, y):
        ok = (x >= 0) & (x < SW) & (y >= 0) & (y < SH)
        return np.where(ok, src[np.clip(y, 0, SH - 1), np.clip(x, 0, SW - 1)].astype(np.int64), 0)
    v = ((32 - a) * (32 - b) * 32 * px(x0, y0) + a * (32 - b) * 32 * px(x0 + 1, y0) +
         (32 - a) * b * 32 * px(x0, y0 + 1) + a * b * 32 * px(x0 + 1, y0 + 1))
    return np.clip((v + (1 << 14)) >> 15, 0, 255).astype(np.uint8)


def test_remap_downsample_from_larger_raw_image():
    # the ROS node's maps undistort and resize (downsample_factor): a W x H depth
    # map fed from a larger raw camera image
    W, H = 64, 48
    rng = np.random.default_rng(1)
    raw = rng.integers(0, 256, (2 * H + 3, 2 * W + 5)).astype(np.uint8)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    mx = xx * 2.03 + 0.37
    my = yy * 1.97 + 0.61
    dm = Depthmap(W, H, 50.0, 32.0, 50.0, 24.0, mx, my, 1)
    assert dm.add_frames(raw, np.eye(3, 4, dtype=np.float32))
    got = dm.getReferenceImage()
    dm.close()
    assert np.array_equal(got, remap_fixed_point(raw, mx, my))


def test_point_cloud_matches_oracle():
    """qm_point_cloud on a fused depth map (with -1 and NaN pixels) and the
    Depthmap-level point cloud of a sequence's last result, against the
    float32 restatement of Publisher::publishPointCloud, in raster order."""
    import torch
    from open_quadtree_mapping_amd import stages as S
    W, H = 376, 240
    cam = camera_for(W, H)
    rng = np.random.default_rng(4)
    depth = rng.uniform(0.05, 6.0, (H, W)).astype(np.float32)
    depth[rng.random((H, W)) < 0.2] = -1.0
    depth[rng.random((H, W)) < 0.01] = np.nan
    ref = rng.integers(0, 256, (H, W)).astype(np.uint8)
    T = np.array([[0.99, -0.1, 0.05, 0.3], [0.1, 0.98, 0.02, -0.2], [-0.05, 0.0, 0.99, 1.5]], np.float32)
    print("point cloud stage", flush=True)
    got = S.point_cloud(torch.from_numpy(depth).cuda(), torch.from_numpy(ref).cuda(), (cam.fx, cam.fy, cam.cx, cam.cy),
                        T).cpu().numpy()
    want = O.point_cloud(depth, ref, cam, T)
    assert got.shape == want.shape
    assert np.array_equal(got, want, equal_nan=True)
    # Depthmap level: 32 frames, so that the fused depth has become confident
    n = 32
    traj = lateral_trajectory(n, step=0.04, seed=12)
    sc = make_scene(12)
    imgs = [render(sc, cam, Tk)[0] for Tk in traj]
    dm = Depthmap(W, H, cam.fx, cam.cx, cam.fy, cam.cy, semi2dense_ratio=1,
                  params=default_params(idepth_min=0.02, idepth_max=2.0))
    last = None
    for k in range(n):
        if dm.add_frames(imgs[k], traj[k]) and k > 0:
            last = k
    assert last is not None
    pc = dm.getPointCloud()
    want = O.point_cloud(dm.getDepthmap(), dm.getReferenceImage(), cam, dm.getT_world_ref())
    assert len(pc) > 1000 and np.array_equal(pc, want, equal_nan=True)
    dm.close()
