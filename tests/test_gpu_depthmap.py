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
