"""Shared synthetic inputs for the tests (seeded, cached per process)."""
from functools import lru_cache

import numpy as np

from open_quadtree_mapping_amd.scene import Camera, camera_for, forward_lateral_trajectory, lateral_trajectory, \
    make_scene, render
from open_quadtree_mapping_amd.se3 import relative_pose


@lru_cache(maxsize=16)
def keyframe_set(W=376, H=240, seed=3, motion="forward", n=6):
    """Key = last frame, references = the four before it (newest first).
    Returns dict(key, refs, T (4,3,4), cam, gt, poses, frames)."""
    cam = camera_for(W, H)
    sc = make_scene(seed)
    traj = forward_lateral_trajectory(n, seed=seed) if motion == "forward" else lateral_trajectory(n, seed=seed)
    frames = [render(sc, cam, T) for T in traj]
    k = n - 1
    refs = [frames[k - 1 - i][0] for i in range(4)]
    T = np.stack([relative_pose(traj[k - 1 - i], traj[k]) for i in range(4)])
    return dict(key=frames[k][0], refs=refs, T=T, cam=cam, gt=frames[k][1], poses=traj, frames=frames)


def cam_tuple(cam: Camera):
    return (cam.fx, cam.fy, cam.cx, cam.cy)
