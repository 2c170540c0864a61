"""Per-stage error of the HIP pipeline against the FP64 oracle.

Each stage is fed the same inputs on both sides (the FP32 outputs of the stage
before it, widened to double), so the figures isolate the stage's own
rounding; the last line chains the whole keyframe in FP64 from the images.
FP32 and FP64 round the texel coordinates of the cost volume and the BP
messages differently, so a small fraction of texel choices and labels flips;
the table reports how many, and the error of the rest. The test prints one
JSON line per resolution and asserts loose bounds on it.
"""
import json

import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd import default_params
from open_quadtree_mapping_amd import stages as S
from tests._data import cam_tuple, keyframe_set

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def stats(got, want, mask=None):
    """Relative error of got against want over entries finite in both."""
    g = np.asarray(got, np.float64).ravel()
    w = np.asarray(want, np.float64).ravel()
    ok = np.isfinite(g) & np.isfinite(w)
    if mask is not None:
        ok &= np.asarray(mask).ravel()
    nan_agree = float(np.mean(np.isnan(g) == np.isnan(w)))
    r = np.abs(g[ok] - w[ok]) / np.maximum(np.abs(w[ok]), 1e-30)
    if r.size == 0:
        return dict(n=0, nan_agree=nan_agree)
    return dict(n=int(r.size), p50=float(np.median(r)), p99=float(np.quantile(r, 0.99)), max=float(r.max()),
                frac_gt_1e4=float(np.mean(r > 1e-4)), nan_agree=nan_agree)


@pytest.mark.parametrize("W,H,seed", [(376, 240, 21), (752, 480, 22)])
def test_stage_errors_vs_fp64(W, H, seed):
    N, imin, imax = 64, 0.02, 2.0
    ks = keyframe_set(W, H, seed)
    cam = ks["cam"]
    print(f"fp64 stage errors {W}x{H} seed {seed}", flush=True)
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    levels = O.quadtree(ks["key"])
    report = {"size": f"{W}x{H}", "N": N}

    # quadtree: integer arithmetic, identical in any precision
    report["quadtree_mismatch"] = int(np.sum(host(S.quadtree(dev(ks["key"]))) != levels))

    # cost volume
    cost = host(S.cost_volume(dev(ks["key"]), [dev(r) for r in ks["refs"]], ks["T"], cam_tuple(cam), dev(levels), p))
    c64, _ = O.cost_volume(ks["key"], ks["refs"], ks["T"], cam, N, imin, imax, levels, fp64=True)
    report["cost"] = stats(cost, c64)

    # BP + extraction on the same (FP32) cost volume
    dco, lab = S.bp_extract(dev(cost), W, H, dev(levels), p)
    dco, lab = host(dco), host(lab)
    d64, l64 = O.bp_extract(cost.astype(np.float64), W, H, levels, N, imin, imax, fp64=True)
    node = l64 >= 0
    report["bp_label_flip_frac"] = float(np.mean(lab[node] != l64[node]))
    report["bp_depth"] = stats(dco, d64, mask=(lab == l64) & node)

    # upsample of the same sparse map (weights: FP32 LUT vs FP64 exp)
    sparse = O.compose_sparse(W, H, dco, lab)
    dense = host(S.upsample(dev(sparse), dev(ks["key"]), p))
    wr64, wc64 = O.weights(ks["key"], fp64=True)
    rs, cs = O.ref_segments(W, H)
    up64 = O.upsample(sparse.astype(np.float64), wr64, wc64, row_seg=rs, col_seg=cs, fp64=True)
    report["upsample"] = stats(dense, up64)

    # fusion: seeds brought to confidence by 16 static FP32 steps (the first
    # confident output of a static camera comes at step 16), then one moving step
    seeds1 = np.zeros((H, W, 4), np.float32)
    eye = np.eye(3, 4, dtype=np.float32)
    for _ in range(16):
        _, _, seeds1, _ = O.fuse(seeds1, dense, eye, cam)
    dense2 = (dense * np.float32(1.01)).astype(np.float32)
    T = np.ascontiguousarray(ks["T"][0], np.float32)
    sd = dev(seeds1.copy())
    dout, sout = S.fuse(sd, dev(dense2), cam_tuple(cam), T)
    dout, sout = host(dout), host(sout)
    _, do64, so64, _ = O.fuse(seeds1.astype(np.float64), dense2.astype(np.float64), T, cam, fp64=True)
    report["fuse_depth"] = stats(dout, do64, mask=(dout > 0) & (do64 > 0))
    report["fuse_seed_mu"] = stats(sout[..., 2], so64[..., 2])
    report["fuse_valid_agree"] = float(np.mean((dout > 0) == (do64 > 0)))

    # whole keyframe chained in FP64 from the images (dense depth before fusion)
    sp64 = O.compose_sparse(W, H, *O.bp_extract(c64, W, H, levels, N, imin, imax, fp64=True), fp64=True)
    e2e64 = O.upsample(sp64, wr64, wc64, row_seg=rs, col_seg=cs, fp64=True)
    report["end_to_end_dense"] = stats(dense, e2e64)

    print(json.dumps(report), flush=True)
    assert report["quadtree_mismatch"] == 0
    assert report["cost"]["frac_gt_1e4"] < 1e-3
    assert report["bp_label_flip_frac"] < 0.01
    assert report["bp_depth"]["max"] < 1e-5
    assert report["upsample"]["p99"] < 1e-5
    assert report["fuse_depth"]["n"] > 0 and report["fuse_depth"]["p99"] < 1e-4
    assert report["fuse_valid_agree"] > 0.99
    assert report["end_to_end_dense"]["frac_gt_1e4"] < 0.02
