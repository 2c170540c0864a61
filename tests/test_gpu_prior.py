"""The reference's prior-driven alternates on the GPU against the oracle:
quadtree_depth_kernal (src/quadtree.cu:79-154) and prior_to_cost
(src/depth_extract.cu:23-47) as stages, inside the batched pipeline (frames
with and without a prior in one launch, plain and graph replay), and behind
the Depthmap interface fed by the semi-dense module. All bit-exact."""
import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd import default_params
from open_quadtree_mapping_amd import stages as S
from open_quadtree_mapping_amd.batch import Batch
from open_quadtree_mapping_amd.depthmap import Depthmap
from open_quadtree_mapping_amd.scene import camera_for, forward_lateral_trajectory, lateral_trajectory, make_scene, \
    render
from open_quadtree_mapping_amd.se3 import relative_pose
from tests._data import cam_tuple, keyframe_set

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def prior_maps(gt, seed, frac=0.35):
    """(inverse depth, variance) prior from a ground-truth depth map: noisy,
    sparse, with NaN and non-positive entries; and a dense prior depth map
    with a few holes."""
    rng = np.random.default_rng(seed)
    H, W = gt.shape
    prior = np.zeros((H, W, 2), np.float32)
    m = rng.random((H, W)) < frac
    prior[m, 0] = (1.0 / gt[m]) * rng.uniform(0.97, 1.03, m.sum())
    prior[m, 1] = rng.uniform(0.005, 0.2, m.sum())
    prior[rng.random((H, W)) < 0.01] = [np.nan, 0.1]
    prior[rng.random((H, W)) < 0.01] = [0.5, 0.0]
    depth = gt.astype(np.float32).copy()
    depth[rng.random((H, W)) < 0.0005] = 0.0
    return prior, depth


@pytest.mark.parametrize("W,H,seed", [(376, 240, 3), (752, 480, 5), (100, 70, 8)])
def test_quadtree_prior_matches_oracle(W, H, seed):
    ks = keyframe_set(W, H, seed)
    print(f"quadtree prior {W}x{H}", flush=True)
    _, depth = prior_maps(ks["gt"], seed)
    levels = O.quadtree(ks["key"])
    want = O.quadtree_prior(depth, levels)
    got = host(S.quadtree_prior(dev(depth), dev(levels)))
    assert np.array_equal(got, want)
    assert (want > levels).any()
    zero = np.zeros((H, W), np.uint8)  # the full depth-level map
    assert np.array_equal(host(S.quadtree_prior(dev(depth), dev(zero))), O.quadtree_prior(depth, zero))


@pytest.mark.parametrize("W,H,seed,N,imin,imax", [(376, 240, 3, 64, 0.02, 2.0), (752, 480, 5, 64, 0.02, 2.0),
                                                  (376, 240, 4, 128, 0.02, 1.0 / 0.3)])
def test_prior_to_cost_matches_oracle(W, H, seed, N, imin, imax):
    ks = keyframe_set(W, H, seed)
    print(f"prior to cost {W}x{H} N={N}", flush=True)
    prior, _ = prior_maps(ks["gt"], seed)
    levels = O.quadtree(ks["key"])
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    cost, count = S.cost_volume(dev(ks["key"]), [dev(r) for r in ks["refs"]], ks["T"], cam_tuple(ks["cam"]),
                                dev(levels), p, want_count=True)
    S.prior_to_cost(dev(prior), dev(levels), cost, p, count)
    rc, rn = O.cost_volume(ks["key"], ks["refs"], ks["T"], ks["cam"], N, imin, imax, levels)
    wc, wn = O.prior_to_cost(prior, levels, rc, rn, N, imin, imax)
    assert np.array_equal(host(count), wn)
    got = host(cost)
    assert np.array_equal(got, wc), f"max abs diff {np.abs(got - wc).max()}"
    assert (wn > 64).sum() > 100


def _stream(W, H, seed, n=8):
    cam = camera_for(W, H)
    sc = make_scene(seed)
    traj = forward_lateral_trajectory(n, seed=seed)
    fr = [render(sc, cam, T) for T in traj]
    return cam, traj, [f[0] for f in fr], [f[1] for f in fr]


@pytest.mark.parametrize("graph", [False, True])
def test_batch_with_priors_matches_oracle(graph):
    sizes = [(376, 240, 3), (752, 480, 5), (376, 240, 7)]
    use = [(True, True), (False, False), (True, False)]  # (prior cost, prior quadtree) per stream
    streams = [_stream(W, H, s) for (W, H, s) in sizes]
    dimgs = [[dev(im) for im in st[2]] for st in streams]
    N, imin, imax = 64, 0.02, 2.0
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    seeds = [np.zeros((H, W, 4), np.float32) for (W, H, _) in sizes]
    b = None
    for k in range(4, 7):
        print(f"batch with priors keyframe {k} graph={graph}", flush=True)
        frames, ref = [], []
        for i, (cam, traj, imgs, gts) in enumerate(streams):
            T = np.stack([relative_pose(traj[k - 1 - j], traj[k]) for j in range(4)])
            T_last = relative_pose(traj[k], traj[k - 1])
            prior, pdepth = prior_maps(gts[k], 100 * i + k)
            fr = dict(key=dimgs[i][k], refs=[dimgs[i][k - 1 - j] for j in range(4)], T=T,
                      cam=(cam.fx, cam.fy, cam.cx, cam.cy), T_last=T_last)
            if use[i][0]:
                fr["prior"] = dev(prior)
            if use[i][1]:
                fr["prior_depth"] = dev(pdepth)
            frames.append(fr)
            ref.append(O.keyframe(imgs[k], [imgs[k - 1 - j] for j in range(4)], T, cam, N, imin, imax, seeds[i],
                                  T_last, prior=prior if use[i][0] else None,
                                  prior_depth=pdepth if use[i][1] else None))
        if b is None:
            b = Batch(frames, p, outputs=("depth", "dense", "sparse"))
        else:
            for i, fr in enumerate(frames):
                b.set_frame(i, fr)
        b.run(graph=graph)
        torch.cuda.synchronize()
        for i in range(len(sizes)):
            out = b.outputs[i]
            assert np.array_equal(b.read_buffer(i, Batch.BUF_LEVELS), ref[i]["levels"]), f"stream {i} levels"
            assert np.array_equal(out["sparse"].cpu().numpy(), ref[i]["sparse"]), f"stream {i} sparse map differs"
            assert np.array_equal(out["depth"].cpu().numpy(), ref[i]["depth"], equal_nan=True), f"stream {i} depth"
    b.close()


def test_depthmap_prior_alternates_match_oracle():
    W, H = 376, 240
    cam = camera_for(W, H)
    n = 18
    traj = lateral_trajectory(n, step=0.04, seed=11)
    sc = make_scene(11)
    imgs = [render(sc, cam, T)[0] for T in traj]
    p = default_params(idepth_min=0.02, idepth_max=2.0)
    dm = Depthmap(W, H, cam.fx, cam.cx, cam.fy, cam.cy, params=p)
    dm.set_prior_alternates(prior_cost=True, prior_quadtree=True)
    ref = O.DepthmapSequence(cam, 64, 0.02, 2.0, prior_cost=True, prior_quadtree=True)
    plain = O.DepthmapSequence(cam, 64, 0.02, 2.0)
    outputs = differ = 0
    for k in range(n):
        print(f"depthmap prior alternates frame {k}", flush=True)
        got = dm.add_frames(imgs[k], traj[k])
        want = ref.add_frames(imgs[k], traj[k])
        plain.add_frames(imgs[k], traj[k])
        assert got == want
        if got:
            assert np.array_equal(dm.getDebugmap(), ref.debug)
            assert np.array_equal(dm.getDepthmap(), ref.depth, equal_nan=True)
            outputs += 1
            differ += int(not np.array_equal(ref.debug, plain.debug))
    assert outputs >= 3 and differ >= 1  # the alternates change the result
    dm.close()
