"""The segmented-scan line solver of the upsample (QM_UPSAMPLE_SCAN) against
the oracle's restatement of it, bit for bit, through the C ABI.

The stage call and the batched pipeline both run it; frames whose lines are
longer than 64 x 24 steps fall back to the Thomas kernels line direction by
line direction, as the oracle does. The accuracy of the scan against FP64 and
its effect on the fused depth are gated on the CPU
(tests/test_upsample_scan.py).
"""
import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd import QmError, default_params
from open_quadtree_mapping_amd import stages as S
from open_quadtree_mapping_amd.batch import Batch
from tests._data import keyframe_set
from tests.test_gpu_pipeline import keyframe_inputs, stream_frames

pytestmark = pytest.mark.gpu
DEV = "cuda"
SCAN = O.UPSAMPLE_SCAN


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _random_case(W, H, seed, density=0.05):
    rng = np.random.default_rng(seed)
    # a smooth image with a few steps, so the couplings take both small and large values
    yy, xx = np.mgrid[0:H, 0:W]
    img = (100 + 60 * np.sin(xx / 37.0) * np.cos(yy / 23.0) + 80 * ((xx // 53 + yy // 41) % 2)
           + rng.integers(-3, 4, (H, W)))
    img = np.clip(img, 0, 255).astype(np.uint8)
    sparse = ((rng.random((H, W)) * 9 + 1) * (rng.random((H, W)) < density)).astype(np.float32)
    return img, sparse


# sizes: the headline rows (64 lanes x 12 steps) and columns (32 x 16), a
# split row chain (376 -> 368, 32 x 12), rows on 64 lanes x 24 steps (1280),
# odd sizes with short lanes and a column split (150 x 100 -> 96), rows longer
# than 1536 (Thomas rows, scan columns), columns longer than 1536 (scan rows,
# Thomas columns) and lambda above 2^100 (Thomas)
@pytest.mark.parametrize("W,H,lam", [(752, 480, 10.0), (376, 240, 10.0), (1280, 720, 10.0), (150, 100, 10.0),
                                     (752, 480, 1000.0), (150, 100, 0.5), (1600, 40, 10.0), (40, 1600, 10.0),
                                     (150, 100, 1e31)])
def test_stage_scan_matches_oracle(W, H, lam):
    img, sparse = _random_case(W, H, W + H)
    print(f"scan upsample {W}x{H} lambda {lam}", flush=True)
    wr, wc = O.weights(img)
    rs, cs = O.ref_segments(W, H)
    ref = O.upsample(sparse, wr, wc, lam=lam, row_seg=rs, col_seg=cs, solver=SCAN)
    p = default_params(upsample_solver=SCAN)
    p.lambda_ = lam
    got = host(S.upsample(dev(sparse), dev(img), p))
    assert np.array_equal(got, ref, equal_nan=True)
    # the Thomas mode of the same call stays the reference order
    thomas = O.upsample(sparse, wr, wc, lam=lam, row_seg=rs, col_seg=cs)
    p.upsample_solver = O.UPSAMPLE_THOMAS
    assert np.array_equal(host(S.upsample(dev(sparse), dev(img), p)), thomas, equal_nan=True)


def test_stage_scan_keyframe_sparse_map():
    ks = keyframe_set(752, 480, 5)
    W, H = 752, 480
    levels = O.quadtree(ks["key"])
    rc, _ = O.cost_volume(ks["key"], ks["refs"], ks["T"], ks["cam"], 64, 0.02, 2.0, levels)
    rd, rl = O.bp_extract(rc, W, H, levels, 64, 0.02, 2.0)
    sparse = O.compose_sparse(W, H, rd, rl)
    wr, wc = O.weights(ks["key"])
    rs, cs = O.ref_segments(W, H)
    ref = O.upsample(sparse, wr, wc, row_seg=rs, col_seg=cs, solver=SCAN)
    got = host(S.upsample(dev(sparse), dev(ks["key"]), default_params(upsample_solver=SCAN)))
    assert np.array_equal(got, ref, equal_nan=True)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_scan_matches_oracle(graph):
    """Uneven streams (376x240 with its split row chain, 752x480, 1280x720 on
    the 64-lane 24-step row kernel), three consecutive keyframes each, the
    fusion state carried over, in the fused cost arithmetic and the scan
    solver."""
    N, imin, imax = 64, 0.02, 2.0
    sizes = [(376, 240, 3), (752, 480, 5), (1280, 720, 9)]
    streams = [stream_frames(W, H, s) for (W, H, s) in sizes]
    dev_imgs = [[torch.from_numpy(im).to(DEV) for im in imgs] for (_, _, imgs) in streams]
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax, cost_arith=O.COST_FUSED, upsample_solver=SCAN)
    first_k = 4
    frames = [keyframe_inputs(cam, traj, imgs, first_k, dev_imgs[i])[0] for i, (cam, traj, imgs) in
              enumerate(streams)]
    b = Batch(frames, p, outputs=("depth", "dense", "sparse"))
    seeds = [np.zeros((H, W, 4), np.float32) for (W, H, _) in sizes]
    for k in range(first_k, first_k + 3):
        print(f"scan batch keyframe {k} graph={graph}", flush=True)
        ref = []
        for i, (cam, traj, imgs) in enumerate(streams):
            fr, T, T_last = keyframe_inputs(cam, traj, imgs, k, dev_imgs[i])
            if k > first_k:
                b.set_frame(i, fr)
            ref.append(O.keyframe(imgs[k], [imgs[k - 1 - j] for j in range(4)], T, cam, N, imin, imax, seeds[i],
                                  T_last, arith=O.COST_FUSED, solver=SCAN))
        b.run(graph=graph)
        torch.cuda.synchronize()
        for i in range(len(sizes)):
            out = b.outputs[i]
            assert np.array_equal(out["sparse"].cpu().numpy(), ref[i]["sparse"]), f"stream {i} sparse map differs"
            assert np.array_equal(out["depth"].cpu().numpy(), ref[i]["depth"], equal_nan=True), \
                f"stream {i} fused depth differs"
    b.close()


def test_batch_backward_refuses_scan():
    cam, traj, imgs = stream_frames(376, 240, 3)
    d = [torch.from_numpy(im).to(DEV) for im in imgs]
    fr = keyframe_inputs(cam, traj, imgs, 4, d)[0]
    b = Batch([fr], default_params(upsample_solver=SCAN), outputs=("depth",))
    b.run()
    with pytest.raises(QmError, match="Thomas"):
        b.backward([torch.zeros(240, 376, device=DEV)])
    b.close()
