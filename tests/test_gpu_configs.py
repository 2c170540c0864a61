"""The two plain-pipeline workloads of the scope table, run exactly as stated,
against the FP32 oracle.

Config 1: one keyframe at 376x240 (fx = fy = 183, cx = 188, cy = 120), a
10-frame trajectory of 4 cm lateral steps with U(-0.5, 0.5) deg jitter per
axis, 64 inverse-depth samples over 0.5-50 m, through the stage-level calls:
keyframe = frame 9 with the references of slots {8, 8, 7, 6, 5}
(src/depth_extract.cu:153-162), quadtree, cost volume, BP, upsample and the
first fusion (every output -1, src/depth_fusion.cu:207-210).

Config 2: 752x480 (fx = fy = 365, cx = 376, cy = 240), 5 cm per frame along
normalize(1, 0, 1), semi2dense_ratio = 3, through quadmap::Depthmap: 33
warm-up frames and 30 more give 10 depth maps (frames 35, 38, ..., 62), the
gates of src/seed_matrix.cu:144-155."""
import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd import default_params
from open_quadtree_mapping_amd import stages as S
from open_quadtree_mapping_amd.depthmap import Depthmap
from open_quadtree_mapping_amd.scene import Camera, camera_for, forward_lateral_trajectory, lateral_trajectory, \
    make_scene, render
from open_quadtree_mapping_amd.se3 import relative_pose

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("arith", [O.COST_REFERENCE, O.COST_FUSED])
def test_config1_stage_level_keyframe(arith):
    W, H = 376, 240
    cam = Camera(W, H, 183.0, 183.0, 188.0, 120.0)
    traj = lateral_trajectory(10, step=0.04, jitter_deg=0.5, seed=1)
    sc = make_scene(1)
    imgs = [render(sc, cam, T)[0] for T in traj]
    k = 9
    refs = [imgs[k - 1 - i] for i in range(4)]  # slots {8, 8, 7, 6, 5}: references 8, 7, 6, 5
    T = np.stack([relative_pose(traj[k - 1 - i], traj[k]) for i in range(4)])
    camt = (cam.fx, cam.fy, cam.cx, cam.cy)
    N, imin, imax = 64, 0.02, 2.0
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax, cost_arith=arith)
    print(f"config 1 keyframe 9, arith {arith}", flush=True)
    # quadtree
    lv = host(S.quadtree(dev(imgs[k])))
    assert np.array_equal(lv, O.quadtree(imgs[k]))
    # cost volume
    cost, count = S.cost_volume(dev(imgs[k]), [dev(r) for r in refs], T, camt, dev(lv), p, want_count=True)
    rc, rn = O.cost_volume(imgs[k], refs, T, cam, N, imin, imax, lv, arith=arith)
    assert np.array_equal(host(count), rn)
    assert np.array_equal(host(cost), rc)
    # BP and winner extraction
    dc, lab = S.bp_extract(dev(rc), W, H, dev(lv), p)
    rd, rl = O.bp_extract(rc, W, H, lv, N, imin, imax)
    assert np.array_equal(host(lab), rl)
    assert np.array_equal(host(dc), rd)
    assert (rl >= 0).sum() > 5000
    # upsample of the sparse map
    sparse = O.compose_sparse(W, H, rd, rl)
    dense = host(S.upsample(dev(sparse), dev(imgs[k])))
    wr, wc = O.weights(imgs[k])
    rs, cs = O.ref_segments(W, H)
    ref_dense = O.upsample(sparse, wr, wc, row_seg=rs, col_seg=cs)
    assert np.array_equal(dense, ref_dense, equal_nan=True)
    # the first fusion: empty seeds, so every output is -1 and the new seeds
    # carry the dense depth
    seeds = np.zeros((H, W, 4), np.float32)
    gs = dev(seeds)
    gd, gsn = S.fuse(gs, dev(ref_dense), camt, np.eye(3, 4, dtype=np.float32))
    _, od, osn, _ = O.fuse(seeds, ref_dense, np.eye(3, 4, dtype=np.float32), cam)
    god = host(gd)
    assert np.all(god == -1.0) and np.array_equal(god, od)
    assert np.array_equal(host(gsn), osn, equal_nan=True)
    # and the whole keyframe through the oracle's keyframe flow agrees with the stages
    seeds_kf = np.zeros((H, W, 4), np.float32)
    res = O.keyframe(imgs[k], refs, T, cam, N, imin, imax, seeds_kf, np.eye(3, 4, dtype=np.float32), arith=arith)
    assert np.array_equal(res["sparse"], sparse) and np.array_equal(res["label"], rl)
    assert np.array_equal(seeds_kf, osn, equal_nan=True)


def test_config2_ratio3_ten_maps():
    W, H = 752, 480
    cam = camera_for(W, H)
    assert (cam.fx, cam.fy, cam.cx, cam.cy) == (365.0, 365.0, 376.0, 240.0)
    n = 33 + 30
    traj = forward_lateral_trajectory(n, step=0.05, seed=41)
    sc = make_scene(41)
    imgs = [render(sc, cam, T)[0] for T in traj]
    p = default_params(idepth_min=0.02, idepth_max=2.0)
    dm = Depthmap(W, H, cam.fx, cam.cx, cam.fy, cam.cy, semi2dense_ratio=3, params=p)
    ref = O.DepthmapSequence(cam, 64, 0.02, 2.0, ratio=3)
    outputs = []
    for k in range(n):
        got = dm.add_frames(imgs[k], traj[k])
        want = ref.add_frames(imgs[k], traj[k])
        assert got == want, f"frame {k}: has_result {got} vs {want}"
        assert dm.num_references() == len(ref.refs)
        if got and k > 0:
            print(f"config 2 depth map at frame {k}", flush=True)
            assert np.array_equal(dm.getDebugmap(), ref.debug)
            assert np.array_equal(dm.getDepthmap(), ref.depth, equal_nan=True)
            assert np.array_equal(dm.getReferenceImage(), ref.reference)
            assert np.array_equal(dm.getT_world_ref(), ref.T_world_ref)
            outputs.append(k)
    assert outputs == list(range(35, 63, 3)) and len(outputs) == 10
    dm.close()
