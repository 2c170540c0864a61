"""The oracle's segmented-scan upsample solver (UPSAMPLE_SCAN) on the CPU.

The scan solves the same line systems as the reference's Thomas recurrences
(src/depth_upsample.cu:94-302) by composing each recurrence's maps per lane
and scanning them over 64 lanes. Pinned here independently of the oracle's
Thomas code:
  * in FP64 the scan and the Thomas order are the same algebra: they agree to
    rounding (1e-13 lambda) on every pixel, chain splits included;
  * in FP32 it solves (I + lambda L) x = r as accurately as the Thomas order:
    both are compared with a dense numpy solve of every row and column system;
  * lines longer than 64 x 24 steps take the Thomas solve, bit for bit (lines
    of up to 512 steps run on 32 lanes, longer ones on 64);
  * through the whole keyframe (fused cost arithmetic, BP, upsample, fusion)
    the fused depths of the two solvers differ by more than 1e-4 relative on
    at most 0.1% of the pixels, and the scan's are no further from the FP64
    chain than the Thomas order's.
"""
import numpy as np
import pytest

import oracle as O
from tests._data import keyframe_set

SCAN, THOMAS = O.UPSAMPLE_SCAN, O.UPSAMPLE_THOMAS


def _problem(W, H, seed, density=0.1):
    rng = np.random.default_rng(seed)
    d = ((rng.random((H, W)) * 9 + 1) * (rng.random((H, W)) < density)).astype(np.float32)
    wr = np.exp(rng.uniform(np.log(1e-3), 0.0, (H, W + 1))).astype(np.float32)
    wc = np.exp(rng.uniform(np.log(1e-3), 0.0, (H + 1, W))).astype(np.float32)
    wr[:, 0] = 0.0
    wc[0, :] = 0.0
    return d, wr, wc


def _dense_solve_lines(w, r, lam, seg):
    """Rows of r (shape L x n) each solved as (I + lam L(w)) x = r with the
    couplings w (L x (n + 1)) cut at seg, by numpy's dense solver in FP64."""
    L, n = r.shape
    out = np.zeros((L, n))
    for i in range(L):
        for s0, s1 in ((0, min(seg, n)), (min(seg, n), n)):
            if s1 <= s0:
                continue
            m = s1 - s0
            # w(k) and w(k + 1) on the diagonal, w(k + 1) off it inside the chain:
            # a cut coupling stays on the diagonal, as in the recurrences
            wl = w[i, s0:s1].astype(np.float64)
            wr_ = w[i, s0 + 1:s1 + 1].astype(np.float64)
            A = np.diag(1.0 + lam * (wl + wr_))
            A -= np.diag(lam * wr_[:-1], 1) + np.diag(lam * wr_[:-1], -1)
            out[i, s0:s1] = np.linalg.solve(A, r[i, s0:s1].astype(np.float64))
    return out


def _dense_upsample(d, wr, wc, lam, rs, cs):
    n = (d > 0).astype(np.float64)
    rx = _dense_solve_lines(wr, d.astype(np.float64), lam, rs)
    rn = _dense_solve_lines(wr, n, lam, rs)
    cx = _dense_solve_lines(wc.T, rx.T, lam, cs).T
    cn = _dense_solve_lines(wc.T, rn.T, lam, cs).T
    return cx / cn


@pytest.mark.parametrize("W,H,rs,cs,lam", [(150, 100, 150, 96, 10.0), (150, 100, 37, 61, 10.0), (96, 70, 96, 70, 0.5),
                                           (96, 70, 90, 64, 1000.0), (200, 40, 200, 40, 10.0)])
def test_fp64_scan_equals_fp64_thomas(W, H, rs, cs, lam):
    d, wr, wc = _problem(W, H, W * 7 + H)
    t = O.upsample(d, wr, wc, lam=lam, row_seg=rs, col_seg=cs, fp64=True)
    s = O.upsample(d, wr, wc, lam=lam, row_seg=rs, col_seg=cs, fp64=True, solver=SCAN)
    rel = np.abs(s - t) / np.abs(t)
    print(W, H, lam, "max", np.nanmax(rel))
    # FP64 rounding, amplified like the systems' condition (about 4 lambda)
    assert np.nanmax(rel) < 1e-13 * max(1.0, lam)


@pytest.mark.parametrize("W,H,rs,cs,lam", [(150, 100, 144, 96, 10.0), (150, 100, 37, 61, 10.0), (96, 70, 96, 70, 0.5),
                                           (96, 70, 90, 64, 1000.0)])
def test_fp32_scan_against_dense_fp64_solve(W, H, rs, cs, lam):
    d, wr, wc = _problem(W, H, W + 3 * H)
    ref = _dense_upsample(d, wr, wc, lam, rs, cs)
    t = O.upsample(d, wr, wc, lam=lam, row_seg=rs, col_seg=cs)
    s = O.upsample(d, wr, wc, lam=lam, row_seg=rs, col_seg=cs, solver=SCAN)
    et = np.abs(t - ref) / np.abs(ref)
    es = np.abs(s - ref) / np.abs(ref)
    print(W, H, lam, "thomas max", et.max(), "scan max", es.max(), "p99", np.percentile(es, 99))
    # as accurate as the reference order (the bound on the FP32 solve of these
    # systems grows with lambda: the couplings bring the pivots near 1 - 1/lambda)
    assert es.max() <= 4.0 * et.max() + 1e-6
    assert np.percentile(es, 99) <= 2.0 * np.percentile(et, 99) + 1e-7


def test_long_lines_take_the_thomas_solve():
    # rows of 1600 > 64 x 24 steps: Thomas; columns of 40: scan
    W, H = 1600, 40
    d, wr, wc = _problem(W, H, 5)
    rs, cs = O.ref_segments(W, H)
    _, rx_t, rn_t = O.upsample(d, wr, wc, row_seg=rs, col_seg=cs, want_rows=True)
    out_s, rx_s, rn_s = O.upsample(d, wr, wc, row_seg=rs, col_seg=cs, want_rows=True, solver=SCAN)
    assert np.array_equal(rx_s, rx_t) and np.array_equal(rn_s, rn_t)
    # and rows of exactly 1536 steps run the scan (they differ from Thomas somewhere)
    W = 1536
    d, wr, wc = _problem(W, H, 6)
    _, rx_t, _ = O.upsample(d, wr, wc, row_seg=W, col_seg=H, want_rows=True)
    _, rx_s, _ = O.upsample(d, wr, wc, row_seg=W, col_seg=H, want_rows=True, solver=SCAN)
    assert not np.array_equal(rx_s, rx_t)
    assert np.max(np.abs(rx_s - rx_t) / np.maximum(np.abs(rx_t), 1e-30)) < 1e-4


SCENES = [(376, 240, 3), (752, 480, 5), (752, 480, 6)]


@pytest.fixture(scope="module")
def pipeline_stats():
    out = []
    for (W, H, seed) in SCENES:
        ks = keyframe_set(W, H, seed)
        lv = O.quadtree(ks["key"])
        cost, _ = O.cost_volume(ks["key"], ks["refs"], ks["T"], ks["cam"], 64, 0.02, 2.0, lv, arith=O.COST_FUSED)
        dc, lb = O.bp_extract(cost, W, H, lv, 64, 0.02, 2.0)
        sparse = O.compose_sparse(W, H, dc, lb)
        wr, wc = O.weights(ks["key"])
        rs, cs = O.ref_segments(W, H)
        dense_t = O.upsample(sparse, wr, wc, row_seg=rs, col_seg=cs)
        dense_s = O.upsample(sparse, wr, wc, row_seg=rs, col_seg=cs, solver=SCAN)
        dense_64 = O.upsample(sparse, wr, wc, row_seg=rs, col_seg=cs, fp64=True)
        seeds = np.zeros((H, W, 4), np.float32)
        T0 = np.eye(3, 4, dtype=np.float32)
        # two keyframes through the whole chain: the second fuses into the first's seeds
        fused = {}
        for solver in (THOMAS, SCAN):
            sd = seeds.copy()
            O.keyframe(ks["key"], ks["refs"], ks["T"], ks["cam"], 64, 0.02, 2.0, sd, T0, arith=O.COST_FUSED,
                       solver=solver)
            r = O.keyframe(ks["key"], ks["refs"], ks["T"], ks["cam"], 64, 0.02, 2.0, sd, T0, arith=O.COST_FUSED,
                           solver=solver)
            fused[solver] = r["depth"]
        ok = np.isfinite(dense_64) & (dense_64 > 0)
        e_t = np.abs(dense_t[ok] - dense_64[ok]) / dense_64[ok]
        e_s = np.abs(dense_s[ok] - dense_64[ok]) / dense_64[ok]
        ft, fs = fused[THOMAS], fused[SCAN]
        valid_t, valid_s = ft > 0, fs > 0
        both = valid_t & valid_s
        far = int(np.count_nonzero(np.abs(fs[both] - ft[both]) > 1e-4 * np.abs(ft[both])))
        s = dict(scene=(W, H, seed), px=W * H, dense_t_max=float(e_t.max()), dense_s_max=float(e_s.max()),
                 dense_t_p99=float(np.percentile(e_t, 99)), dense_s_p99=float(np.percentile(e_s, 99)),
                 valid_flips=int(np.count_nonzero(valid_t != valid_s)), fused_far=far)
        print(s, flush=True)
        out.append(s)
    return out


def test_pipeline_dense_depth_as_accurate_as_thomas(pipeline_stats):
    for s in pipeline_stats:
        assert s["dense_s_max"] <= 4.0 * s["dense_t_max"] + 1e-6, s
        assert s["dense_s_p99"] <= 2.0 * s["dense_t_p99"] + 1e-7, s


def test_pipeline_fused_depth_within_tolerance(pipeline_stats):
    for s in pipeline_stats:
        assert s["valid_flips"] <= 1e-3 * s["px"], s
        assert s["fused_far"] <= 1e-3 * s["px"], s
