"""End-to-end parity of the batched keyframe pipeline with the FP32 oracle.

Several camera streams of uneven resolution run through every stage in one
launch per stage (plain launches and hipGraph replay), three consecutive
keyframes each, so the fusion state carries over between keyframes.
"""
import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd import default_params
from open_quadtree_mapping_amd.batch import Batch
from open_quadtree_mapping_amd.scene import camera_for, forward_lateral_trajectory, make_scene, render
from open_quadtree_mapping_amd.se3 import relative_pose

pytestmark = pytest.mark.gpu
DEV = "cuda"


def stream_frames(W, H, seed, n=8):
    cam = camera_for(W, H)
    sc = make_scene(seed)
    traj = forward_lateral_trajectory(n, seed=seed)
    imgs = [render(sc, cam, T)[0] for T in traj]
    return cam, traj, imgs


def keyframe_inputs(cam, traj, imgs, k, dev_imgs):
    T = np.stack([relative_pose(traj[k - 1 - i], traj[k]) for i in range(4)])
    T_last = relative_pose(traj[k], traj[k - 1])  # previous keyframe -> this one
    fr = dict(key=dev_imgs[k], refs=[dev_imgs[k - 1 - i] for i in range(4)], T=T,
              cam=(cam.fx, cam.fy, cam.cx, cam.cy), T_last=T_last)
    return fr, T, T_last


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("N,imin,imax", [(64, 0.02, 2.0), (128, 0.02, 1.0 / 0.3)])
def test_batch_matches_oracle(graph, N, imin, imax):
    sizes = [(376, 240, 3), (752, 480, 5), (376, 240, 7)]
    streams = [stream_frames(W, H, s) for (W, H, s) in sizes]
    dev_imgs = [[torch.from_numpy(im).to(DEV) for im in imgs] for (_, _, imgs) in streams]
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    first_k = 4
    frames = [keyframe_inputs(cam, traj, imgs, first_k, dev_imgs[i])[0] for i, (cam, traj, imgs) in
              enumerate(streams)]
    b = Batch(frames, p, outputs=("depth", "dense", "sparse"))
    seeds = [np.zeros((H, W, 4), np.float32) for (W, H, _) in sizes]
    for k in range(first_k, first_k + 3):
        print(f"batch keyframe {k} graph={graph} N={N}", flush=True)
        ref = []
        for i, (cam, traj, imgs) in enumerate(streams):
            fr, T, T_last = keyframe_inputs(cam, traj, imgs, k, dev_imgs[i])
            if k > first_k:
                b.set_frame(i, fr)
            ref.append(O.keyframe(imgs[k], [imgs[k - 1 - j] for j in range(4)], T, cam, N, imin, imax, seeds[i],
                                  T_last))
        b.run(graph=graph)
        torch.cuda.synchronize()
        for i in range(len(sizes)):
            out = b.outputs[i]
            sp = out["sparse"].cpu().numpy()
            assert np.array_equal(sp, ref[i]["sparse"]), f"stream {i} sparse map differs"
            dn = out["dense"].cpu().numpy()
            dd = out["depth"].cpu().numpy()
            assert np.array_equal(dd, ref[i]["depth"], equal_nan=True), f"stream {i} fused depth differs"
            assert dn.shape == (sizes[i][1], sizes[i][0])
    b.close()
