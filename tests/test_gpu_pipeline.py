"""End-to-end parity of the batched keyframe pipeline with the FP32 oracle.

Several camera streams of uneven resolution run through every stage in one
launch per stage (plain launches and hipGraph replay), three consecutive
keyframes each, so the fusion state carries over between keyframes.
"""
import json

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


@pytest.mark.parametrize("arith", [O.COST_REFERENCE, O.COST_FUSED])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("N,imin,imax", [(64, 0.02, 2.0), (128, 0.02, 1.0 / 0.3)])
def test_batch_matches_oracle(graph, N, imin, imax, arith):
    sizes = [(376, 240, 3), (752, 480, 5), (376, 240, 7)]
    streams = [stream_frames(W, H, s) for (W, H, s) in sizes]
    dev_imgs = [[torch.from_numpy(im).to(DEV) for im in imgs] for (_, _, imgs) in streams]
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax, cost_arith=arith)
    first_k = 4
    frames = [keyframe_inputs(cam, traj, imgs, first_k, dev_imgs[i])[0] for i, (cam, traj, imgs) in
              enumerate(streams)]
    b = Batch(frames, p, outputs=("depth", "dense", "sparse"))
    seeds = [np.zeros((H, W, 4), np.float32) for (W, H, _) in sizes]
    for k in range(first_k, first_k + 3):
        print(f"batch keyframe {k} graph={graph} N={N} arith={arith}", flush=True)
        ref = []
        for i, (cam, traj, imgs) in enumerate(streams):
            fr, T, T_last = keyframe_inputs(cam, traj, imgs, k, dev_imgs[i])
            if k > first_k:
                b.set_frame(i, fr)
            ref.append(O.keyframe(imgs[k], [imgs[k - 1 - j] for j in range(4)], T, cam, N, imin, imax, seeds[i],
                                  T_last, arith=arith))
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


def test_twelve_uneven_streams_config4():
    """Config 4: 12 streams of uneven size (4 x 376x240, 4 x 752x480,
    4 x 1280x720, intrinsics scaled from 752x480), 64 samples, every stage in
    one launch per stage, the keyframe pipeline replayed as a hipGraph; two
    consecutive keyframes per stream against the FP32 oracle."""
    N, imin, imax = 64, 0.02, 2.0
    kinds = [(376, 240, 11), (376, 240, 12), (752, 480, 13), (752, 480, 14), (1280, 720, 15), (1280, 720, 16)]
    rendered = {k: stream_frames(k[0], k[1], k[2], n=6) for k in kinds}
    sizes = [kinds[i % 2 + 2 * (i // 4)] for i in range(12)]  # 4 of each size, two scenes per size
    streams = [rendered[k] for k in sizes]
    dev_imgs = {k: [torch.from_numpy(im).to(DEV) for im in rendered[k][2]] for k in kinds}
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    frames = [keyframe_inputs(cam, traj, imgs, 4, dev_imgs[sizes[i]])[0] for i, (cam, traj, imgs) in
              enumerate(streams)]
    b = Batch(frames, p, outputs=("depth", "sparse"))
    seeds = [np.zeros((H, W, 4), np.float32) for (W, H, _) in sizes]
    for k in (4, 5):
        print(f"12 uneven streams, keyframe {k}", flush=True)
        refs = []
        for i, (cam, traj, imgs) in enumerate(streams):
            fr, T, T_last = keyframe_inputs(cam, traj, imgs, k, dev_imgs[sizes[i]])
            if k > 4:
                b.set_frame(i, fr)
            refs.append(O.keyframe(imgs[k], [imgs[k - 1 - j] for j in range(4)], T, cam, N, imin, imax, seeds[i],
                                   T_last))
        b.run(graph=True)
        torch.cuda.synchronize()
        for i in range(12):
            assert np.array_equal(b.outputs[i]["sparse"].cpu().numpy(), refs[i]["sparse"]), f"stream {i}"
            assert np.array_equal(b.outputs[i]["depth"].cpu().numpy(), refs[i]["depth"], equal_nan=True), f"stream {i}"
    b.close()


def test_every_launch_shape_gives_identical_results_and_autotune(tmp_path):
    """The autotuner's knobs change launch shapes only: every value of every
    knob reproduces the default run bit for bit; the autotuner then picks
    offered values and the batch still matches."""
    sizes = [(376, 240, 21), (752, 480, 22)]
    streams = [stream_frames(W, H, s, n=6) for (W, H, s) in sizes]
    dev_imgs = [[torch.from_numpy(im).to(DEV) for im in imgs] for (_, _, imgs) in streams]
    p = default_params(n_hyp=64, idepth_min=0.02, idepth_max=2.0)
    frames = [keyframe_inputs(cam, traj, imgs, 5, dev_imgs[i])[0] for i, (cam, traj, imgs) in enumerate(streams)]

    def run(b):
        b.reset_seeds()
        b.run(graph=True)
        torch.cuda.synchronize()
        return [(o["sparse"].cpu().numpy(), o["dense"].cpu().numpy()) for o in b.outputs]

    b = Batch(frames, p, outputs=("sparse", "dense"))
    want = run(b)
    offered = {"cost_waves": (2, 4, 8), "bp_rows": (4, 8, 16), "extract_rows": (4, 8, 16), "overlap": tuple(range(14)),
               "bp_pair": (0, 8, 16), "bp_seg": (8, 16, 32, 64), "bp_pair_levels": (1, 2, 3, 4)}
    defaults = {"cost_waves": 4, "bp_rows": 16, "extract_rows": 16, "overlap": 0, "bp_pair": 8, "bp_seg": 8,
                "bp_pair_levels": 2}
    assert b.tuning() == defaults
    for knob, values in offered.items():
        for v in values:
            print(f"launch shape {knob}={v}", flush=True)
            b.set_tuning(**{knob: v})
            got = run(b)
            for (gs, gd), (ws, wd) in zip(got, want):
                assert np.array_equal(gs, ws) and np.array_equal(gd, wd, equal_nan=True), f"{knob}={v}"
        b.set_tuning(**{knob: defaults[knob]})
    # the two sizes with shapes of their own (each stage then launches per size)
    assert b.groups() == [(376, 240), (752, 480)]
    for g0, g1 in [({"bp_pair": 0, "cost_waves": 2}, {"bp_pair": 16, "bp_seg": 16, "extract_rows": 4}),
                   ({"bp_pair": 8, "bp_seg": 64, "bp_pair_levels": 4}, {"bp_pair": 0, "bp_rows": 8})]:
        print(f"group shapes {g0} {g1}", flush=True)
        b.set_tuning(**defaults)
        b.set_group_tuning(0, **g0)
        b.set_group_tuning(1, **g1)
        assert all(b.group_tuning(0)[k] == v for k, v in g0.items())
        assert all(b.group_tuning(1)[k] == v for k, v in g1.items())
        got = run(b)
        for (gs, gd), (ws, wd) in zip(got, want):
            assert np.array_equal(gs, ws) and np.array_equal(gd, wd, equal_nan=True), f"{g0} {g1}"
    b.set_tuning(**defaults)
    cache = str(tmp_path / "shapes.json")
    res = b.autotune(reps=3, cache=cache)
    print(f"autotune: {res} groups {[b.group_tuning(g) for g in range(2)]}", flush=True)
    for knob, (v, ms) in res.items():
        assert v in offered[knob] and ms > 0
    assert b.tuning() == {k: v for k, (v, _) in res.items()}
    for g in range(2):
        assert all(v in offered[k] for k, v in b.group_tuning(g).items())
    got = run(b)
    for (gs, gd), (ws, wd) in zip(got, want):
        assert np.array_equal(gs, ws) and np.array_equal(gd, wd, equal_nan=True)
    # a second batch of the same composition takes the cached choices without timing
    db = json.load(open(cache))
    assert len(db) == 3
    b2 = Batch(frames, p, outputs=("sparse", "dense"))
    res2 = b2.autotune(reps=3, cache=cache)
    assert all(ms is None for _, ms in res2.values())
    assert b2.tuning() == b.tuning() and all(b2.group_tuning(g) == b.group_tuning(g) for g in range(2))
    b2.close()
    b.close()


@pytest.mark.parametrize("graph", [False, True])
def test_two_group_overlap_matches_one_stream_over_keyframes(graph):
    """The batch run as two frame groups on two streams (every overlap
    setting) gives the one-stream results over consecutive keyframes, so the
    fusion state carried between runs is the same; the autotuner's whole-run
    timings leave that state untouched."""
    sizes = [(376, 240, 31), (752, 480, 32), (376, 240, 33)]  # odd count: groups of 2 and 1
    streams = [stream_frames(W, H, s) for (W, H, s) in sizes]
    dev_imgs = [[torch.from_numpy(im).to(DEV) for im in imgs] for (_, _, imgs) in streams]
    p = default_params(n_hyp=64, idepth_min=0.02, idepth_max=2.0)

    def frames_at(k):
        return [keyframe_inputs(cam, traj, imgs, k, dev_imgs[i])[0] for i, (cam, traj, imgs) in enumerate(streams)]

    def sequence(b, autotune_after_first=False):
        b.reset_seeds()
        out = []
        for k in range(4, 7):
            for i, fr in enumerate(frames_at(k)):
                b.set_frame(i, fr)
            b.run(graph=graph)
            torch.cuda.synchronize()
            out.append([o["depth"].cpu().numpy() for o in b.outputs])
            if autotune_after_first and k == 4:
                b.autotune(reps=1)
        return out

    b = Batch(frames_at(4), p, outputs=("depth",))
    b.set_tuning(overlap=0)
    want = sequence(b)
    for ov in range(1, 14):
        print(f"overlap={ov} graph={graph}", flush=True)
        b.set_tuning(overlap=ov)
        got = sequence(b)
        for k, (g, w) in enumerate(zip(got, want)):
            for i in range(len(sizes)):
                assert np.array_equal(g[i], w[i], equal_nan=True), f"overlap {ov} keyframe {k} stream {i}"
    print(f"autotune between keyframes graph={graph}", flush=True)
    got = sequence(b, autotune_after_first=True)
    for k, (g, w) in enumerate(zip(got, want)):
        for i in range(len(sizes)):
            assert np.array_equal(g[i], w[i], equal_nan=True), f"after autotune: keyframe {k} stream {i}"
    b.close()
