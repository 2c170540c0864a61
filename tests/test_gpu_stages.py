"""Parity of every HIP stage with the CPU oracle (FP32), through the C ABI.

Bar: bit-exact. The HIP kernels follow the same pinned arithmetic as the
oracle: quadtree levels, cost volumes, BP messages, labels and depths,
upsampled maps and fusion outputs are compared for equality. A relative-error
check before some of them only makes a large mismatch fail with a clearer
message.
"""
import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd import default_params
from open_quadtree_mapping_amd import stages as S
from tests._data import cam_tuple, keyframe_set

pytestmark = pytest.mark.gpu
DEV = "cuda"
REL = 1e-4


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-12)
    d[both_nan] = 0.0
    return d


def constructed_images():
    rng = np.random.default_rng(11)
    out = {}
    out["constant"] = np.full((64, 96), 120, np.uint8)
    yy, xx = np.mgrid[0:64, 0:96]
    out["checker20"] = np.where((xx + yy) % 2 == 0, 100, 120).astype(np.uint8)
    out["random"] = rng.integers(0, 256, size=(80, 112), dtype=np.uint8)
    img = np.full((64, 96), 50, np.uint8)
    img[:, 52:] = 200  # step on a 4-pixel boundary, not 8 or 16
    out["step4"] = img
    # smooth ramp with flat patches: exercises levels 3 and 4 and near-threshold deviations
    ramp = (xx * 0.7 + yy * 0.3).astype(np.int32)
    ramp[8:40, 16:48] = 77
    ramp[40:56, 60:76] = ramp[40:56, 60:76] // 13 * 13
    out["ramp"] = np.clip(ramp, 0, 255).astype(np.uint8)
    # random piecewise-flat blocks with +-10 noise: right at the thresholds
    base = rng.integers(0, 200, size=(8, 12))
    blk = np.kron(base, np.ones((16, 16), np.int64))
    noise = rng.integers(-10, 11, size=blk.shape) * (rng.random(blk.shape) < 0.3)
    out["blocks"] = np.clip(blk + noise, 0, 255).astype(np.uint8)
    # partial tiles: width 376-like remainder (W % 16 = 8) and odd sizes
    out["partial"] = np.clip(np.kron(rng.integers(0, 255, size=(6, 6)), np.ones((13, 17))), 0, 255)[:70, :90].astype(
        np.uint8)
    return out


@pytest.mark.parametrize("name", sorted(constructed_images().keys()))
def test_quadtree_constructed(name):
    img = constructed_images()[name]
    print(f"quadtree constructed {name} {img.shape}", flush=True)
    got = host(S.quadtree(dev(img)))
    ref = O.quadtree(img)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("W,H,seed", [(376, 240, 3), (752, 480, 5), (1280, 720, 9)])
def test_quadtree_scene(W, H, seed):
    ks = keyframe_set(W, H, seed)
    print(f"quadtree scene {W}x{H} seed {seed}", flush=True)
    got = host(S.quadtree(dev(ks["key"])))
    assert np.array_equal(got, O.quadtree(ks["key"]))


ARITHS = [O.COST_REFERENCE, O.COST_FUSED]


@pytest.mark.parametrize("arith", ARITHS)
@pytest.mark.parametrize("W,H,seed,N,imin,imax", [(376, 240, 3, 64, 0.02, 2.0), (752, 480, 5, 64, 0.01, 2.0),
                                                  (376, 240, 4, 128, 0.02, 1.0 / 0.3)])
def test_cost_volume(W, H, seed, N, imin, imax, arith):
    ks = keyframe_set(W, H, seed)
    print(f"cost volume {W}x{H} N={N} arith={arith}", flush=True)
    levels = O.quadtree(ks["key"])
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax, cost_arith=arith)
    cost, count = S.cost_volume(dev(ks["key"]), [dev(r) for r in ks["refs"]], ks["T"], cam_tuple(ks["cam"]),
                                dev(levels), p, want_count=True)
    rc, rn = O.cost_volume(ks["key"], ks["refs"], ks["T"], ks["cam"], N, imin, imax, levels, arith=arith)
    assert np.array_equal(host(count), rn)
    got = host(cost)
    assert np.array_equal(got, rc), f"max abs diff {np.abs(got - rc).max()}"


@pytest.mark.parametrize("arith", ARITHS)
@pytest.mark.parametrize("N", [64, 128])
def test_cost_volume_general_transform(N, arith):
    # transforms whose rotation block is not a rotation (rows scaled by 2, so
    # |u2u| + |v2u| exceeds 1.75) take the kernel's clamped general tap path;
    # one such reference switches the whole launch to it
    W, H, seed = 376, 240, 6
    ks = keyframe_set(W, H, seed)
    T = ks["T"].copy()
    T[1, :2, :3] *= np.float32(2.0)
    T[3, :2, :3] *= np.float32(0.7)
    print(f"cost volume general transform N={N} arith={arith}", flush=True)
    levels = O.quadtree(ks["key"])
    p = default_params(n_hyp=N, idepth_min=0.02, idepth_max=2.0, cost_arith=arith)
    cost = S.cost_volume(dev(ks["key"]), [dev(r) for r in ks["refs"]], T, cam_tuple(ks["cam"]), dev(levels), p)
    rc, _ = O.cost_volume(ks["key"], ks["refs"], T, ks["cam"], N, 0.02, 2.0, levels, arith=arith)
    assert np.array_equal(host(cost), rc)


@pytest.mark.parametrize("W,H,seed,N", [(376, 240, 3, 64), (752, 480, 5, 64), (376, 240, 4, 128)])
def test_bp_extract(W, H, seed, N):
    ks = keyframe_set(W, H, seed)
    print(f"bp {W}x{H} N={N}", flush=True)
    imin, imax = 0.02, 2.0
    levels = O.quadtree(ks["key"])
    rc, _ = O.cost_volume(ks["key"], ks["refs"], ks["T"], ks["cam"], N, imin, imax, levels)
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    d, lab, msg = S.bp_extract(dev(rc), W, H, dev(levels), p, want_messages=True)
    rd, rl, rm = O.bp_extract(rc, W, H, levels, N, imin, imax, want_messages=True)
    gl = host(lab)
    assert np.array_equal(gl < 0, rl < 0), "node mask differs"
    frac = np.mean(gl != rl)
    assert frac <= 1e-3, f"label mismatch fraction {frac}"
    assert np.array_equal(host(msg), rm), "messages differ"
    assert np.array_equal(gl, rl)
    assert np.array_equal(host(d), rd, equal_nan=True), "depths differ"


# half-sweep counts per level (fine -> coarse) that leave odd last sweeps, a
# level with one sweep (no pair) or none (initialisation only), and pairs only
@pytest.mark.parametrize("iters", [(3, 5, 2, 1), (2, 2, 2, 2), (1, 3, 0, 2), (5, 1, 3, 4), (0, 0, 0, 0), (6, 7, 1, 0)])
def test_bp_schedules_match_oracle(iters):
    W, H, N = 376, 240, 64
    ks = keyframe_set(W, H, 7)
    print(f"bp schedule {iters}", flush=True)
    imin, imax = 0.02, 2.0
    levels = O.quadtree(ks["key"])
    rc, _ = O.cost_volume(ks["key"], ks["refs"], ks["T"], ks["cam"], N, imin, imax, levels)
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    for l in range(4):
        p.bp_iters[l] = iters[l]
    d, lab, msg = S.bp_extract(dev(rc), W, H, dev(levels), p, want_messages=True)
    rd, rl, rm = O.bp_extract(rc, W, H, levels, N, imin, imax, want_messages=True, iters=iters)
    assert np.array_equal(host(msg), rm), "messages differ"
    assert np.array_equal(host(lab), rl)
    assert np.array_equal(host(d), rd, equal_nan=True), "depths differ"


def test_bp_uniform_cost_gives_label_zero():
    # uniform data cost: all messages are 0 and the argmin tree keeps index 0
    W, H, N = 128, 96, 64
    levels = np.full((H, W), 2, np.uint8)
    cost = np.full((H // 4, W // 4, N), 0.25, np.float32)
    d, lab, msg = S.bp_extract(dev(cost), W, H, dev(levels), default_params(), want_messages=True)
    assert np.all(host(lab) == 0)
    assert np.all(host(msg) == 0)
    assert np.allclose(host(d), 100.0)


def crafted_cost(kind, H, W, N, rng):
    """Cost volumes that put the winner where the extraction has edge cases:
    the last sample (no parabola step), interior minima of random depth and
    width (the step), integer costs with many ties (the argmin tree's lower
    index), and flat runs of equal minima."""
    Hc, Wc = H // 4, W // 4
    d = np.arange(N, dtype=np.float64)
    if kind == "descending":
        c = (N - 1 - d)[None, None, :] * 0.01 + rng.uniform(0, 1e-3, (Hc, Wc, 1))
    elif kind == "vshape":
        m = rng.integers(0, N, (Hc, Wc, 1))
        s = rng.uniform(0.001, 0.05, (Hc, Wc, 1))
        c = np.abs(d[None, None, :] - m) * s + rng.uniform(0, 0.3, (Hc, Wc, 1))
    elif kind == "ties":
        c = rng.integers(0, 4, (Hc, Wc, N)) * 0.125
    else:  # plateau: a run of equal minima starting at a random sample
        c = np.full((Hc, Wc, N), 0.5)
        m = rng.integers(0, N - 4, (Hc, Wc))
        for k in range(4):
            np.put_along_axis(c, (m + k)[..., None], 0.1, axis=2)
    return c.astype(np.float32)


@pytest.mark.parametrize("kind", ["descending", "vshape", "ties", "plateau"])
def test_bp_extract_crafted_winners(kind):
    # BP and the winner extraction (argmin tree, parabola step, non-finite
    # step, depth) on costs chosen for their edge cases; every output bit-exact
    W, H, N = 128, 96, 64
    rng = np.random.default_rng(11)
    print(f"bp crafted {kind}", flush=True)
    cost = crafted_cost(kind, H, W, N, rng)
    levels = np.full((H, W), 2, np.uint8)
    imin, imax = 0.02, 2.0
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    d, lab, msg = S.bp_extract(dev(cost), W, H, dev(levels), p, want_messages=True)
    rd, rl, rm = O.bp_extract(cost, W, H, levels, N, imin, imax, want_messages=True)
    assert np.array_equal(host(msg), rm), "messages differ"
    assert np.array_equal(host(lab), rl), "labels differ"
    assert np.array_equal(host(d), rd, equal_nan=True), "depths differ"
    if kind == "descending":
        assert (rl == N - 1).mean() > 0.5  # the case is exercised


@pytest.mark.parametrize("W,H,seed", [(376, 240, 3), (752, 480, 5)])
def test_upsample_reference_mode(W, H, seed):
    ks = keyframe_set(W, H, seed)
    print(f"upsample {W}x{H}", flush=True)
    levels = O.quadtree(ks["key"])
    rc, _ = O.cost_volume(ks["key"], ks["refs"], ks["T"], ks["cam"], 64, 0.02, 2.0, levels)
    rd, rl = O.bp_extract(rc, W, H, levels, 64, 0.02, 2.0)
    sparse = O.compose_sparse(W, H, rd, rl)
    wr, wc = O.weights(ks["key"])
    rs, cs = O.ref_segments(W, H)
    ref = O.upsample(sparse, wr, wc, row_seg=rs, col_seg=cs)
    got = host(S.upsample(dev(sparse), dev(ks["key"])))
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.nanmax(rel_err(got, ref)) <= REL
    assert np.array_equal(got, ref)


# odd sizes take the unaligned load paths; segment splits inside a chunk of
# steps leave partial chunks on both sides of the split
@pytest.mark.parametrize("H,W,rs,cs", [(96, 136, 136, 96), (96, 136, 128, 88), (95, 137, 50, 7), (17, 19, 3, 0)])
def test_upsample_general(H, W, rs, cs):
    rng = np.random.default_rng(5)
    d = (rng.uniform(1, 10, (H, W)) * (rng.random((H, W)) < 0.3)).astype(np.float32)
    conf = rng.uniform(1e-3, 1.0, (H, W)).astype(np.float32)
    wr = np.exp(rng.uniform(np.log(1e-3), np.log(10), (H, W + 1))).astype(np.float32)
    wc = np.exp(rng.uniform(np.log(1e-3), np.log(10), (H + 1, W))).astype(np.float32)
    ref, rx, rn = O.upsample(d, wr, wc, conf=conf, row_seg=rs, col_seg=cs, want_rows=True)
    got, gx, gn = S.upsample_general(dev(d), dev(wr), dev(wc), conf=dev(conf), row_seg=rs, col_seg=cs,
                                     want_rows=True)
    assert np.array_equal(host(gx), rx)
    assert np.array_equal(host(gn), rn)
    assert np.array_equal(host(got), ref)


# right-hand sides below 2^-70 and above 2^70 and couplings that push the
# pivots above 2^20, beside ordinary values: the chains' chunks that meet them
# take the plain division, the others the unscaled one
def test_upsample_general_extreme_magnitudes():
    H, W = 40, 72
    rng = np.random.default_rng(9)
    d = (rng.uniform(1, 10, (H, W)) * (rng.random((H, W)) < 0.3)).astype(np.float32)
    pick = rng.random((H, W))
    d[pick < 0.02] = 1e-25
    d[(pick >= 0.02) & (pick < 0.04)] = 1e24
    d[(pick >= 0.04) & (pick < 0.05)] = 3e-38
    conf = rng.uniform(1e-3, 1.0, (H, W)).astype(np.float32)
    wr = np.exp(rng.uniform(np.log(1e-3), np.log(10), (H, W + 1))).astype(np.float32)
    wc = np.exp(rng.uniform(np.log(1e-3), np.log(10), (H + 1, W))).astype(np.float32)
    wr[rng.random((H, W + 1)) < 0.02] = 1e6
    wc[rng.random((H + 1, W)) < 0.02] = 1e6
    ref, rx, rn = O.upsample(d, wr, wc, conf=conf, want_rows=True)
    got, gx, gn = S.upsample_general(dev(d), dev(wr), dev(wc), conf=dev(conf), want_rows=True)
    assert np.array_equal(host(gx), rx, equal_nan=True)
    assert np.array_equal(host(gn), rn, equal_nan=True)
    assert np.array_equal(host(got), ref, equal_nan=True)


# the stage calls keep their scratch for the next call: repeated calls of one
# size reuse it (device bytes held stay flat) and give the same results
def test_stage_scratch_is_reused():
    from open_quadtree_mapping_amd._lib import lib
    H, W = 64, 96
    rng = np.random.default_rng(4)
    d = (rng.uniform(1, 10, (H, W)) * (rng.random((H, W)) < 0.3)).astype(np.float32)
    wr = np.ones((H, W + 1), np.float32)
    wc = np.ones((H + 1, W), np.float32)
    conf = np.ones((H, W), np.float32)
    ref = O.upsample(d, wr, wc, conf=conf)
    held = []
    for _ in range(3):
        got = S.upsample_general(dev(d), dev(wr), dev(wc), conf=dev(conf))
        torch.cuda.synchronize()
        assert np.array_equal(host(got), ref, equal_nan=True)
        held.append(lib().qm_device_bytes_in_use())
    assert held[1] == held[2], held
    # qm_release_scratch frees what the calls kept; the next call allocates afresh
    assert lib().qm_release_scratch() == 0
    assert lib().qm_device_bytes_in_use() < held[2]
    got = S.upsample_general(dev(d), dev(wr), dev(wc), conf=dev(conf))
    assert np.array_equal(host(got), ref, equal_nan=True)


def fusion_case(W, H, seed):
    rng = np.random.default_rng(seed)
    ks = keyframe_set(W, H, 3 if W == 376 else 5)
    cam = ks["cam"]
    seeds = np.zeros((H, W, 4), np.float32)
    seeds[..., 0] = rng.uniform(0.0, 30.0, (H, W))
    seeds[..., 1] = rng.uniform(0.0, 30.0, (H, W))
    seeds[..., 2] = ks["gt"] * rng.uniform(0.95, 1.05, (H, W))
    seeds[..., 3] = rng.uniform(0.001, 5.0, (H, W))
    seeds[rng.random((H, W)) < 0.05, 0] = 0.0  # bad points
    depth = (ks["gt"] * rng.uniform(0.9, 1.1, (H, W))).astype(np.float32)
    depth[rng.random((H, W)) < 0.02] = np.nan
    depth[rng.random((H, W)) < 0.02] = -1.0
    from open_quadtree_mapping_amd.se3 import relative_pose
    T = relative_pose(ks["poses"][-1], ks["poses"][-2])
    return seeds, depth, cam, T


@pytest.mark.parametrize("W,H,seed", [(376, 240, 1), (752, 480, 2)])
def test_fuse(W, H, seed):
    seeds, depth, cam, T = fusion_case(W, H, seed)
    print(f"fuse {W}x{H}", flush=True)
    rs_w, rd, rs, rt = O.fuse(seeds, depth, T, cam)
    gs = dev(seeds)
    gd, gsn, gt = S.fuse(gs, dev(depth), cam_tuple(cam), T, want_table=True)
    assert np.array_equal(host(gt), rt), "warp table differs"
    assert np.array_equal(host(gs), rs_w, equal_nan=True), "warped seeds differ"
    god = host(gd)
    assert np.array_equal(god == -1.0, rd == -1.0)
    assert np.nanmax(rel_err(god, rd)) <= REL
    assert np.array_equal(god, rd, equal_nan=True)
    assert np.array_equal(host(gsn), rs, equal_nan=True)


@pytest.mark.parametrize("W,H,seed", [(376, 240, 3), (752, 480, 5)])
def test_upsample_local_matches_oracle(W, H, seed):
    ks = keyframe_set(W, H, seed)
    print(f"local upsample {W}x{H}", flush=True)
    rng = np.random.default_rng(seed)
    sparse = np.zeros((H, W), np.float32)
    m = rng.random((H // 4, W // 4)) < 0.6
    sparse[0::4, 0::4][: H // 4, : W // 4][m] = rng.uniform(0.5, 10.0, m.sum()).astype(np.float32)
    got = host(S.upsample_local(dev(sparse), dev(ks["key"])))
    want = O.upsample_local(sparse, ks["key"])
    assert np.array_equal(got, want, equal_nan=True)
