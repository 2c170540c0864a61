"""Pins of the CPU oracle that do not go through the oracle's own formulas.

Each stage is checked against an independent formulation: a simulation of the
reference kernel's shared-memory algorithm, closed forms, brute force on tiny
inputs, a dense FP64 linear solve, or invariants the mathematics fixes. The
reference repository ships no tests, golden vectors or fixtures for this path.
"""
import numpy as np
import pytest

import oracle as O
from open_quadtree_mapping_amd.scene import Camera


# ---------------------------------------------------------------------------
# Quadtree: simulate src/quadtree.cu:17-77 thread by thread, one 16x16 tile at
# a time, with barrier-separated phases (pyramid_num / approve in "shared memory").
# ---------------------------------------------------------------------------
def quadtree_simulated(img):
    H, W = img.shape
    out = np.zeros((H, W), np.uint8)
    for ty in range(0, H, 16):
        for tx in range(0, W, 16):
            inten = {}
            num = {}
            level = {}
            alive = set()
            for ly in range(16):
                for lx in range(16):
                    x, y = tx + lx, ty + ly
                    if x < W and y < H:
                        inten[(lx, ly)] = float(img[y, x])
                        num[(lx, ly)] = 1
                        level[(lx, ly)] = 0
                        alive.add((lx, ly))
            for i in range(1, 5):
                s = 1 << i
                rep = {p: (p[0] - p[0] % s, p[1] - p[1] % s) for p in alive}
                # phase 1: last nodes of level i-1 add into their level-i representative
                h = 1 << (i - 1)
                for p in sorted(alive):
                    if p[0] % h == 0 and p[1] % h == 0 and rep[p] != p:
                        inten[rep[p]] += inten[p]
                        num[rep[p]] += num[p]
                approve = {rep[p]: True for p in alive}
                # phase 2: count check, deviation test
                broke = set()
                for p in alive:
                    if num[rep[p]] != s * s:
                        broke.add(p)
                still = alive - broke
                for p in still:
                    avg = inten[rep[p]] / float(s * s)
                    if abs(inten_pixel(img, tx, ty, p) - avg) > 10.0:
                        approve[rep[p]] = False
                # phase 3
                nxt = set()
                for p in still:
                    if approve[rep[p]]:
                        level[p] = i
                        nxt.add(p)
                    else:
                        num[rep[p]] = 0
                alive = nxt
            for (lx, ly), lv in level.items():
                out[ty + ly, tx + lx] = max(lv, 2)
    return out


def inten_pixel(img, tx, ty, p):
    return float(img[ty + p[1], tx + p[0]])


def _qt_images():
    rng = np.random.default_rng(2)
    imgs = []
    for k in range(4):
        base = rng.integers(0, 230, size=(5, 6))
        blk = np.kron(base, np.ones((8, 8), np.int64))[: 37 + k, : 45 + 2 * k]
        noise = rng.integers(-12, 13, size=blk.shape) * (rng.random(blk.shape) < 0.15)
        imgs.append(np.clip(blk + noise, 0, 255).astype(np.uint8))
    imgs.append(rng.integers(0, 256, size=(20, 33)).astype(np.uint8))
    yy, xx = np.mgrid[0:32, 0:48]
    imgs.append(np.clip(xx // 3 + yy // 5 + 60, 0, 255).astype(np.uint8))
    return imgs


@pytest.mark.parametrize("k", range(6))
def test_quadtree_matches_kernel_simulation(k):
    img = _qt_images()[k]
    assert np.array_equal(O.quadtree(img), quadtree_simulated(img))


def test_quadtree_constructed_cases():
    assert np.all(O.quadtree(np.full((48, 64), 77, np.uint8)) == 4)
    yy, xx = np.mgrid[0:48, 0:64]
    assert np.all(O.quadtree(np.where((xx + yy) % 2 == 0, 90, 130).astype(np.uint8)) == 2)  # +-20
    assert np.all(O.quadtree(np.where((xx + yy) % 2 == 0, 100, 120).astype(np.uint8)) == 4)  # +-10: not > 10
    img = np.full((32, 64), 50, np.uint8)
    img[:, 52:] = 200
    lv = O.quadtree(img)
    assert np.all(lv[:, 48:56] == 2)   # 8x8 block straddles the step
    assert np.all(lv[:, 56:64] == 3)   # 16x16 block straddles it, 8x8 does not
    assert np.all(lv[:, 32:48] == 4)
    # W = 376: the last tile column is 8 wide, so at most level 3 there
    lv = O.quadtree(np.full((32, 376), 10, np.uint8))
    assert np.all(lv[:, 368:] == 3) and np.all(lv[:, :368] == 4)


# ---------------------------------------------------------------------------
# Cost volume closed forms.
# ---------------------------------------------------------------------------
CAM = Camera(96, 64, 60.0, 60.0, 48.0, 32.0)
EYE = np.tile(np.eye(3, 4, dtype=np.float32), (4, 1, 1))


def _nodes(levels, W, H):
    Hc, Wc = H // 4, W // 4
    m = np.zeros((Hc, Wc), bool)
    for cy in range(Hc):
        for cx in range(Wc):
            x, y = 4 * cx, 4 * cy
            qs = 1 << int(levels[y, x])
            m[cy, cx] = 0 < x < W - 1 and 0 < y < H - 1 and x % qs == 0 and y % qs == 0
    return m


ARITHS = [O.COST_REFERENCE, O.COST_FUSED]


@pytest.mark.parametrize("arith", ARITHS)
def test_cost_identical_images_identity_pose_is_zero(arith):
    rng = np.random.default_rng(1)
    key = rng.integers(0, 256, (64, 96)).astype(np.uint8)
    lv = O.quadtree(key)
    cost, cnt = O.cost_volume(key, [key] * 4, EYE, CAM, 64, 0.02, 2.0, lv, arith=arith)
    m = _nodes(lv, 96, 64)
    assert np.array_equal(cnt == 64, m) and np.all((cnt == 0) | (cnt == 64))
    assert np.all(cost[~m] == 0)
    # every projection of the same pixel lands inside (2 <= p < W-2) except near the border
    inner = m.copy()
    inner[:1, :] = inner[-1:, :] = False
    inner[:, :1] = inner[:, -1:] = False
    assert np.all(cost[inner] == 0)


@pytest.mark.parametrize("arith", ARITHS)
def test_cost_slot_multiset_is_0_0_1_2_3(arith):
    key = np.full((64, 96), 100, np.uint8)
    vals = [90, 130, 100, 40]
    refs = [np.full((64, 96), v, np.uint8) for v in vals]
    lv = O.quadtree(key)
    cost, _ = O.cost_volume(key, refs, EYE, CAM, 64, 0.02, 2.0, lv, arith=arith)
    sads = [9.0 * abs(100 - v) for v in vals]
    # normalize_the_cost: add_num = 64, one increment per sample thread of the node
    want = np.float32((sads[0] + sads[3] + sads[1] + sads[0] + sads[2]) / 5.0 / 255.0 / 64.0)
    cy, cx = 32 // 4, 48 // 4
    lvl = int(lv[32, 48])
    assert (48 % (1 << lvl)) == 0 and (32 % (1 << lvl)) == 0
    assert np.allclose(cost[cy, cx], want, rtol=1e-6, atol=0)


@pytest.mark.parametrize("arith", ARITHS)
def test_cost_out_of_view_is_100_over_64(arith):
    key = np.full((64, 96), 100, np.uint8)
    T = EYE.copy()
    T[:, 0, 3] = 1e4  # everything projects far outside
    lv = O.quadtree(key)
    cost, cnt = O.cost_volume(key, [key] * 4, T, CAM, 64, 0.02, 2.0, lv, arith=arith)
    m = cnt == 64
    assert m.any() and np.all(cost[m] == 1.5625)


@pytest.mark.parametrize("arith", ARITHS)
@pytest.mark.parametrize("tx", [0.25, -0.17, 0.4])
def test_cost_translation_minimum_at_true_disparity(arith, tx):
    # pure x translation: the centre pixel (on the optical axis) at depth z shows
    # in the reference at x + fx*tx/z; a textured reference shifted by that
    # disparity makes the SAD exactly 0 at the samples whose projection rounds
    # to it. The projection is taken in exact FP64 (x = 48 + 60 tx / z, z the
    # float sample depth); samples within 1e-4 px of a rounding boundary, where
    # FP32 roundings may legitimately go either way, are not asserted.
    W, H = 96, 64
    rng = np.random.default_rng(4)
    key = rng.integers(0, 256, (H, W)).astype(np.uint8)
    N, imin, imax = 64, 0.02, 2.0
    table = O.depth_table(N, imin, imax)
    tx32 = float(np.float32(tx))
    dstar = 20
    shift = int(np.floor(60.0 * tx32 / float(table[dstar]) + 0.5))
    ref = np.roll(key, shift, axis=1)
    T = EYE.copy()
    T[:, 0, 3] = tx
    lv = np.full((H, W), 2, np.uint8)
    cost, _ = O.cost_volume(key, [ref] * 4, T, CAM, N, imin, imax, lv, arith=arith)
    c = cost[32 // 4, 48 // 4]
    px = 48.0 + 60.0 * tx32 / table.astype(np.float64)
    frac = px + 0.5 - np.floor(px + 0.5)
    clear = (frac > 1e-4) & (frac < 1 - 1e-4)
    hit = np.floor(px + 0.5) == 48 + shift
    assert hit[dstar] and clear[dstar]
    assert np.all(c[hit & clear] == 0)
    assert np.all(c[~hit & clear] > 0)
    assert clear.sum() >= N - 2


# ---------------------------------------------------------------------------
# BP: independent FP64 formulation (dense O(N^2) min over the P1/P2 penalty
# matrix, vectorised sweeps) of the reference schedule, on tiny grids.
# ---------------------------------------------------------------------------
def bp_independent(cost, levels, W, H, N, P1=float(np.float32(0.003)), P2=float(np.float32(0.01)),
                   iters=(4, 10, 10, 10)):
    # P1, P2 are float constants in the reference (src/bpextract.cu:222-223)
    d = np.arange(N)
    V = np.where(d[:, None] == d[None, :], 0.0, np.where(np.abs(d[:, None] - d[None, :]) == 1, P1, P2))
    costs = [cost.astype(np.float64)]
    for _ in range(3):
        c = costs[-1]
        h0, w0 = c.shape[:2]
        h1, w1 = (h0 + 1) // 2, (w0 + 1) // 2
        pad = np.zeros((2 * h1, 2 * w1, N))
        pad[:h0, :w0] = c
        costs.append(pad[0::2, 0::2] + pad[1::2, 0::2] + pad[0::2, 1::2] + pad[1::2, 1::2])
    msg = None
    for l in range(3, -1, -1):
        h, w = costs[l].shape[:2]
        if msg is None:
            msg = np.zeros((4, h, w, N))
        else:
            msg = np.repeat(np.repeat(msg, 2, axis=1), 2, axis=2)[:, :h, :w]
        lev = l + 2
        for it in range(iters[l]):
            A = it % 2 == 0
            new = msg.copy()
            for y in range(h):
                for x in range(w):
                    if ((x + y) % 2 == 0) != A:
                        continue
                    bx = x // 2
                    if lev <= 4 and levels[min(y << lev, H - 1), min(bx << lev, W - 1)] > lev:
                        continue
                    inc = {}
                    if y > 0:
                        inc[0] = msg[1, y - 1, x]
                    if y < h - 1:
                        inc[1] = msg[0, y + 1, x]
                    if x > 0:
                        inc[2] = msg[3, y, x - 1]
                    if x < w - 1:
                        inc[3] = msg[2, y, x + 1]
                    for dr in range(4):
                        hh = costs[l][y, x] + sum(v for k, v in inc.items() if k != dr)
                        m = (hh[None, :] + V).min(axis=1)
                        new[dr, y, x] = m - m.mean()
            msg = new
    h, w = costs[0].shape[:2]
    labels = np.full((h, w), -1)
    for y in range(h):
        for x in range(w):
            qs = 1 << int(levels[4 * y, 4 * x])
            if (4 * x) % qs or (4 * y) % qs:
                continue
            b = costs[0][y, x].copy()
            if x > 0:
                b += msg[3, y, x - 1]
            if x < w - 1:
                b += msg[2, y, x + 1]
            if y > 0:
                b += msg[1, y - 1, x]
            if y < h - 1:
                b += msg[0, y + 1, x]
            labels[y, x] = int(np.argmin(b))
    return labels, msg


@pytest.mark.parametrize("seed,W,H", [(0, 64, 48), (1, 72, 40)])
def test_bp_matches_independent_fp64_formulation(seed, W, H):
    rng = np.random.default_rng(seed)
    N = 16
    cost = rng.random((H // 4, W // 4, N)) * 0.2
    cost[..., 5] -= 0.05  # a preferred label with noise around it
    levels = np.full((H, W), 2, np.uint8)
    levels[:16, :32] = 3  # some skipped cells at the finest level
    lab_ref, msg_ref = bp_independent(cost, levels, W, H, N)
    _, lab, msg = O.bp_extract(cost, W, H, levels, N, 0.01, 2.0, fp64=True, want_messages=True)
    assert np.allclose(msg, msg_ref, atol=1e-12, rtol=0)
    assert np.array_equal(lab, lab_ref)


def test_bp_messages_are_zero_mean_and_uniform_cost_is_label_zero():
    rng = np.random.default_rng(3)
    W, H, N = 96, 64, 64
    cost = rng.random((H // 4, W // 4, N))
    lv = np.full((H, W), 2, np.uint8)
    _, _, msg = O.bp_extract(cost, W, H, lv, N, 0.01, 2.0, fp64=True, want_messages=True)
    assert np.abs(msg.mean(axis=-1)).max() < 1e-13
    d, lab = O.bp_extract(np.full((H // 4, W // 4, N), 0.3, np.float32), W, H, lv, N, 0.01, 2.0)
    assert np.all(lab == 0) and np.allclose(d, 100.0)


def test_extract_parabola_vertex():
    # a cell far from everything: data cost = (d - 20.3)^2 sampled; with uniform
    # neighbours the messages are constant per cell, so the winner is 20 and the
    # sub-sample offset is the parabola vertex 0.3
    W, H, N = 16, 16, 64
    d = np.arange(N, dtype=np.float64)
    cost = np.tile(((d - 20.3) ** 2 * 1e-4)[None, None, :], (H // 4, W // 4, 1))
    lv = np.full((H, W), 4, np.uint8)
    dep, lab = O.bp_extract(cost, W, H, lv, N, 0.01, 2.0, fp64=True, iters=(0, 0, 0, 0))
    assert lab[0, 0] == 20
    step = (2.0 - 0.01) / 63.0
    assert np.isclose(dep[0, 0], 1.0 / (step * 20.3 + 0.01), rtol=1e-12)


# ---------------------------------------------------------------------------
# Upsample: dense FP64 solve of (I + lambda L) per line.
# ---------------------------------------------------------------------------
def line_matrix(w, lam, n, seg):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 1 + lam * (w[i] + w[i + 1])
        if i + 1 < n and i + 1 != seg:
            A[i, i + 1] = A[i + 1, i] = -lam * w[i + 1]
    return A


def upsample_dense(d, conf, wr, wc, lam, rs, cs):
    H, W = d.shape
    rx, rn = conf * d, conf
    X = np.zeros((H, W))
    Nn = np.zeros((H, W))
    for y in range(H):
        A = line_matrix(wr[y], lam, W, rs)
        X[y] = np.linalg.solve(A, rx[y])
        Nn[y] = np.linalg.solve(A, rn[y])
    OX = np.zeros((H, W))
    ON = np.zeros((H, W))
    for x in range(W):
        A = line_matrix(wc[:, x], lam, H, cs)
        OX[:, x] = np.linalg.solve(A, X[:, x])
        ON[:, x] = np.linalg.solve(A, Nn[:, x])
    return OX / ON


@pytest.mark.parametrize("seg", [False, True])
def test_upsample_equals_dense_solve(seg):
    rng = np.random.default_rng(7)
    H, W = 24, 40
    d = rng.uniform(1, 10, (H, W)) * (rng.random((H, W)) < 0.3)
    conf = rng.uniform(0.01, 1, (H, W))
    wr = np.exp(rng.uniform(-7, 2.3, (H, W + 1)))
    wc = np.exp(rng.uniform(-7, 2.3, (H + 1, W)))
    rs, cs = (32, 16) if seg else (W, H)
    want = upsample_dense(d, conf, wr, wc, 10.0, rs, cs)
    got64 = O.upsample(d, wr, wc, conf=conf, row_seg=rs, col_seg=cs, fp64=True)
    assert np.allclose(got64, want, rtol=1e-10, atol=0)
    got32 = O.upsample(d, wr, wc, conf=conf, row_seg=rs, col_seg=cs)
    assert np.max(np.abs(got32 - want) / np.abs(want)) < 1e-4


def test_upsample_constant_and_scale_invariance():
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (32, 48)).astype(np.uint8)
    wr, wc = O.weights(img, fp64=True)
    sp = np.where(rng.random((32, 48)) < 0.1, 3.5, 0.0)
    sp[0, 0] = 3.5
    out = O.upsample(sp, wr, wc, fp64=True)
    assert np.allclose(out, 3.5, rtol=1e-12)
    conf = rng.uniform(0.1, 1, (32, 48))
    d = rng.uniform(1, 5, (32, 48))
    a = O.upsample(d, wr, wc, conf=conf, fp64=True)
    b = O.upsample(d, wr, wc, conf=conf * 7.0, fp64=True)
    assert np.allclose(a, b, rtol=1e-12)


def test_weights_boundaries_follow_launch_grid():
    img = np.random.default_rng(9).integers(0, 256, (40, 56)).astype(np.uint8)  # 56 % 16 = 8, 40 % 16 = 8
    wr, wc = O.weights(img)
    assert np.all(wr[:, 0] == 0) and np.all(wr[:, 56] == 1)
    assert np.all(wc[0, :] == 0) and np.all(wc[40, :] == 1)
    img2 = np.zeros((32, 64), np.uint8)
    wr2, wc2 = O.weights(img2)
    assert np.all(wr2[:, 64] == 0) and np.all(wc2[32, :] == 0)
    lut = O.weight_lut()
    assert lut[0] == 1.0 and np.isclose(lut[20], np.exp(-1.0), rtol=1e-7)


# ---------------------------------------------------------------------------
# Fusion: Vogiatzis & Hernandez (2011) Gaussian x Beta update in FP64.
# ---------------------------------------------------------------------------
def vh_update(a, b, mu, s2, x, tau2):
    s = s2 + tau2
    m = (s2 * x + tau2 * mu) / s
    v = s2 * tau2 / s
    c1 = a / (a + b) * np.exp(-(x - mu) ** 2 / (2 * s)) / np.sqrt(2 * np.pi * s)
    c2 = b / (a + b) / 50.0
    c1, c2 = c1 / (c1 + c2), c2 / (c1 + c2)
    f = c1 * (a + 1) / (a + b + 1) + c2 * a / (a + b + 1)
    e = c1 * (a + 1) * (a + 2) / ((a + b + 1) * (a + b + 2)) + c2 * a * (a + 1) / ((a + b + 1) * (a + b + 2))
    mu_n = c1 * m + c2 * mu
    s2_n = c1 * (v + m * m) + c2 * (s2 + mu * mu) - mu_n ** 2
    a_n = (e - f) / (f - e / f)
    b_n = a_n * (1 - f) / f
    return a_n, b_n, mu_n, s2_n


CAMF = Camera(64, 48, 50.0, 50.0, 32.0, 24.0)


def test_fusion_first_map_is_all_invalid_and_matches_closed_form():
    rng = np.random.default_rng(0)
    d = rng.uniform(1, 10, (48, 64))
    seeds = np.zeros((48, 64, 4))
    _, out, sn, table = O.fuse(seeds, d, np.eye(3, 4, dtype=np.float32), CAMF, fp64=True)
    assert np.all(out == -1.0) and np.all(table == 0)
    tau2 = np.maximum(0.5, 0.2 * d) ** 2
    want = vh_update(10.0, 10.0, d, 2500.0, d, tau2)
    for k in range(4):
        assert np.allclose(sn[..., k], want[k], rtol=1e-12)


def test_fusion_static_recurrence_becomes_confident():
    # static camera, exact measurement every map: the inlier ratio a/(a+b) first
    # dips (the broad initial prior, sigma^2 = 2500, explains little), then rises;
    # the output stays -1 until the prior's ratio exceeds 0.6, at map index 15
    d = np.full((48, 64), 4.0)
    seeds = np.zeros((48, 64, 4))
    ratios, first_valid = [], None
    for k in range(30):
        _, out, seeds, _ = O.fuse(seeds, d, np.eye(3, 4, dtype=np.float32), CAMF, fp64=True)
        ratios.append(seeds[24, 32, 0] / (seeds[24, 32, 0] + seeds[24, 32, 1]))
        if first_valid is None and out[24, 32] > 0:
            first_valid = k
            assert np.isclose(out[24, 32], 4.0, rtol=1e-12)
    assert first_valid == 15
    assert ratios[13] <= 0.6 < ratios[14]
    assert np.all(np.diff(ratios[5:]) > 0) and np.all(np.diff(ratios[:4]) < 0)


def test_fusion_zbuffer_nearest_wins_and_reset_rule():
    seeds = np.zeros((48, 64, 4))
    seeds[..., :2] = (20.0, 5.0)
    seeds[..., 3] = 0.01
    seeds[..., 2] = 5.0
    T = np.eye(3, 4, dtype=np.float32)
    d = np.full((48, 64), 5.0)
    ws, out, sn, table = O.fuse(seeds, d, T, CAMF, fp64=True)
    idx = np.arange(48 * 64).reshape(48, 64)
    inner = table[2:-2, 2:-2]
    assert np.array_equal(inner, idx[2:-2, 2:-2])  # identity warp: every pixel is its own source
    # reset: a measurement far from the prior (outside tau^2 + sigma^2) restarts the seed
    d2 = d.copy()
    d2[10, 10] = 9.0
    _, _, sn2, _ = O.fuse(seeds, d2, T, CAMF, fp64=True)
    want = vh_update(10.0, 10.0, 9.0, 2500.0, 9.0, max(0.5, 1.8) ** 2)
    assert np.allclose(sn2[10, 10], want, rtol=1e-12)
    # two seeds warped onto one pixel: the camera moves +x by 0.1 m, so the seed
    # at depth 5 on x = 31 shifts one pixel left onto x = 30, where the seed at
    # depth 500 stays; the nearer one wins
    s3 = np.zeros((48, 64, 4))
    s3[..., :2] = (20.0, 5.0)
    s3[..., 3] = 0.01
    s3[20, 30] = (20.0, 5.0, 500.0, 0.01)  # far seed: stays on x = 30
    s3[20, 31] = (20.0, 5.0, 5.0, 0.01)
    T2 = np.eye(3, 4, dtype=np.float32)
    T2[0, 3] = -0.1  # camera moves +x: points shift left by fx*0.1/5 = 1 px
    ws3, _, _, tab3 = O.fuse(s3, np.full((48, 64), 5.0), T2, CAMF, fp64=True)
    assert tab3[20, 30] == 20 * 64 + 31


# ---------------------------------------------------------------------------
# Semi-dense keyframe seeds (oracle/qm_semidense.cpp)
# ---------------------------------------------------------------------------
def test_sd_gradient_closed_form():
    H, W = 24, 32
    yy, xx = np.mgrid[0:H, 0:W]
    img = (3 * xx + 5 * yy).astype(np.uint8)
    g = O.sd_gradient(img)
    assert np.all(g[1:-1, 1:-1, 0] == 3.0) and np.all(g[1:-1, 1:-1, 1] == 5.0)
    assert np.all(g[0] == 0) and np.all(g[-1] == 0) and np.all(g[:, 0] == 0) and np.all(g[:, -1] == 0)


def _textured(H, W, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, (H, W)).astype(np.uint8)


def _static_seeds(H, W, idepth, var, counter):
    s = np.zeros((H, W, 8), np.float32)
    s[3:H - 3, 3:W - 3, 0] = 1.0
    s[3:H - 3, 3:W - 3, 2] = counter
    s[3:H - 3, 3:W - 3, 4] = idepth
    s[3:H - 3, 3:W - 3, 5] = var
    s[3:H - 3, 3:W - 3, 6] = 1.0
    s[3:H - 3, 3:W - 3, 7] = 4.0
    return s


def test_sd_zero_motion_regularisation_closed_form():
    """Without motion the epipolar segment is empty, so the update leaves every
    seed alone; the regularisation of a constant field keeps the inverse depth
    exactly and sets the smoothed variance to 1 / sum 1/(var + 0.075^2 r^2)."""
    from open_quadtree_mapping_amd.scene import Camera
    H, W = 40, 48
    cam = Camera(W, H, 60.0, 60.0, W / 2, H / 2)
    img = _textured(H, W, 1)
    g = O.sd_gradient(img)
    s = _static_seeds(H, W, 0.5, 0.1, 10.0)
    before = s.copy()
    T = np.eye(3, 4, dtype=np.float32)
    O.sd_step(s, img, g, img, g, cam, T, T, False)
    inner = (slice(5, H - 5), slice(5, W - 5))
    assert np.array_equal(s[inner][..., [0, 1, 2, 4, 5]], before[inner][..., [0, 1, 2, 4, 5]])
    assert np.all(s[inner][..., 6] == np.float32(0.5))
    acc = np.float32(0.0)
    for j in range(-2, 3):
        for i in range(-2, 3):
            spatial = np.float32(np.float32(0.075) * np.float32(0.075)) * np.float32(i * i + j * j)
            acc = np.float32(acc + np.float32(np.float32(1.0) / np.float32(np.float32(0.1) + spatial)))
    assert np.all(s[inner][..., 7] == np.float32(np.float32(1.0) / acc))


def test_sd_projection_identity_pose():
    """Confident seeds (validity >= 20, smoothed variance <= 0.02) project into
    the sparse map at their own pixel under the identity pose, with depth =
    |K^-1 (x, y, 1)| / idepth, where the image gradient is at least 5."""
    from open_quadtree_mapping_amd.scene import Camera
    H, W = 40, 48
    cam = Camera(W, H, 60.0, 60.0, W / 2, H / 2)
    img = _textured(H, W, 2)
    g = O.sd_gradient(img)
    s = _static_seeds(H, W, 0.25, 0.001, 30.0)
    T = np.eye(3, 4, dtype=np.float32)
    pre = O.sd_step(s, img, g, img, g, cam, T, T, False)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    want = np.sqrt(((xx - cam.cx) / cam.fx) ** 2 + ((yy - cam.cy) / cam.fy) ** 2 + 1.0) / 0.25
    strong = (g[..., 0] ** 2 + g[..., 1] ** 2) >= 25.0
    inner = np.zeros((H, W), bool)
    inner[5:H - 5, 5:W - 5] = True
    m = strong & inner
    assert m.sum() > 100
    assert np.allclose(pre[m], want[m], rtol=1e-6)
    assert np.all(pre[~strong] == 0)


def test_sd_converges_to_rendered_depth():
    """The epipolar filter on a lateral sequence over the textured planes of the
    scene generator: after a few frames the pre-fill depths approach the
    rendered ground truth (a sign or index slip in the search geometry would
    leave them far off)."""
    from open_quadtree_mapping_amd.scene import camera_for, lateral_trajectory, make_scene, render
    W, H = 376, 240
    cam = camera_for(W, H)
    traj = lateral_trajectory(8, step=0.04, seed=11)
    sc = make_scene(11)
    fr = [render(sc, cam, T) for T in traj]
    seq = O.DepthmapSequence(cam, 64, 0.02, 2.0, ratio=1)
    med = []
    for k in range(8):
        seq.add_frames(fr[k][0], traj[k])
        if seq.prefill is not None:
            m = seq.prefill > 0
            med.append(float(np.median(np.abs(seq.prefill[m] - fr[k][1][m]) / fr[k][1][m])))
    assert med[-1] < 0.1 and med[-1] < med[0]


# ---------------------------------------------------------------------------
# Local (bilateral) upsample, src/depth_upsample.cu:304-370: window extent from
# a single node, the weight constants from two nodes in closed form (FP64),
# constant maps, and a scalar per-pixel loop in the reference's loop order.
# ---------------------------------------------------------------------------
def test_local_upsample_single_node_window_and_value():
    H, W = 48, 64
    sparse = np.zeros((H, W), np.float32)
    ni, nj = 7, 5  # node column / row
    sparse[nj * 4, ni * 4] = 3.25
    img = np.full((H, W), 90, np.uint8)
    out = O.upsample_local(sparse, img)
    yy, xx = np.mgrid[0:H, 0:W]
    reach = (np.abs(xx // 4 - ni) <= 3) & (np.abs(yy // 4 - nj) <= 3)
    assert np.all(np.isnan(out[~reach]))
    np.testing.assert_allclose(out[reach], 3.25, rtol=2e-7)


def test_local_upsample_two_node_weights_closed_form():
    H, W = 32, 32
    sparse = np.zeros((H, W), np.float32)
    sparse[8, 8], sparse[8, 16] = 2.0, 5.0
    img = np.full((H, W), 100, np.uint8)
    img[8, 16] = 107  # intensity step at the second node
    out = O.upsample_local(sparse, img)
    for (y, x) in [(8, 10), (9, 13), (11, 12), (6, 9)]:
        d1 = (x - 8) ** 2 + (y - 8) ** 2
        d2 = (x - 16) ** 2 + (y - 8) ** 2
        w1 = np.exp(-d1 / 100.0)
        w2 = np.exp(-d2 / 100.0 - 49.0 / 50.0)
        want = (2.0 * w1 + 5.0 * w2) / (w1 + w2)
        assert abs(out[y, x] - want) <= 2e-6 * want, (y, x)


def test_local_upsample_constant_and_nonpositive_skip():
    rng = np.random.default_rng(2)
    H, W = 40, 56
    img = rng.integers(0, 256, (H, W), dtype=np.uint8)
    sparse = np.zeros((H, W), np.float32)
    sparse[0::4, 0::4] = 4.5
    sparse[0::8, 0::8] = -1.0  # skipped like empty nodes
    out = O.upsample_local(sparse, img)
    np.testing.assert_allclose(out, 4.5, rtol=1e-6)


def _local_upsample_scalar(sparse, img):
    H, W = img.shape
    f32 = np.float32
    out = np.empty((H, W), f32)
    for y in range(H):
        for x in range(W):
            acc, ws = f32(0), f32(0)
            for a in range(7):
                for b in range(7):
                    cx, cy = (x // 4 + a - 3) * 4, (y // 4 + b - 3) * 4
                    if cx < 0 or cx >= W or cy < 0 or cy >= H:
                        w, d = f32(0), f32(0)
                    else:
                        d = sparse[cy, cx]
                        if float(d) <= 0.0:
                            continue
                        di = f32(img[y, x]) - f32(img[cy, cx])
                        sx, sy = f32(x - cx), f32(y - cy)
                        arg = -(sx * sx + sy * sy) / f32(100) - (di * di) / f32(50)
                        w = f32(np.exp(np.float64(arg)))
                    acc = f32(acc + w * d)
                    ws = f32(ws + w)
            with np.errstate(invalid="ignore", divide="ignore"):
                out[y, x] = acc / ws
    return out


def test_local_upsample_matches_scalar_loop():
    rng = np.random.default_rng(4)
    H, W = 20, 28
    img = rng.integers(0, 256, (H, W), dtype=np.uint8)
    sparse = np.zeros((H, W), np.float32)
    m = rng.random((H // 4, W // 4)) < 0.5
    sparse[0::4, 0::4][m] = rng.uniform(0.5, 9.0, m.sum()).astype(np.float32)
    assert np.array_equal(O.upsample_local(sparse, img), _local_upsample_scalar(sparse, img), equal_nan=True)


# ---------------------------------------------------------------------------
# Prior-driven alternates (oracle/qm_prior.cpp).
# quadtree_depth_kernal: simulated thread by thread (src/quadtree.cu:79-154),
# shared atomics applied in thread raster order, plus closed forms.
# ---------------------------------------------------------------------------
def quadtree_prior_simulated(prior_depth, levels):
    H, W = levels.shape
    out = levels.copy()
    f32 = np.float32
    for ty in range(0, H, 16):
        for tx in range(0, W, 16):
            threads = [(lx, ly) for ly in range(16) for lx in range(16) if tx + lx < W and ty + ly < H]
            inv, pin, num = {}, {}, {}
            all_valid = True
            for p in threads:
                d = f32(prior_depth[ty + p[1], tx + p[0]])
                valid = bool(d > 0.0)
                inv[p] = f32(f32(1.0) / d) if valid else f32(0.0)
                pin[p], num[p] = inv[p], 1
                all_valid &= valid
            if not all_valid:
                continue
            level = {p: 0 for p in threads}
            alive = set(threads)
            for i in range(1, 5):
                s, h = 1 << i, 1 << (i - 1)
                rep = {p: (p[0] - p[0] % s, p[1] - p[1] % s) for p in alive}
                for p in sorted(alive, key=lambda q: (q[1], q[0])):  # thread raster order
                    if p[0] % h == 0 and p[1] % h == 0 and rep[p] != p:
                        pin[rep[p]] = f32(pin[rep[p]] + pin[p])
                        num[rep[p]] = num.get(rep[p], 0) + num[p]
                approve = {rep[p]: True for p in alive}
                still = {p for p in alive if num.get(rep[p], 0) == s * s}
                for p in still:
                    avg = f32(pin[rep[p]] / f32(s * s))
                    if abs(f32(inv[p] - avg)) > f32(0.01):
                        approve[rep[p]] = False
                nxt = set()
                for p in still:
                    if approve[rep[p]]:
                        level[p] = i
                        nxt.add(p)
                    else:
                        num[rep[p]] = 0
                alive = nxt
            for (lx, ly), lv in level.items():
                out[ty + ly, tx + lx] = max(int(out[ty + ly, tx + lx]), lv)
    return out


def _prior_depth_maps():
    rng = np.random.default_rng(9)
    maps = []
    # piecewise-planar inverse depth with noise, some holes, ragged sizes
    for k, (H, W) in enumerate([(48, 64), (37, 45), (32, 50)]):
        yy, xx = np.mgrid[0:H, 0:W]
        inv = 0.5 + 0.002 * (xx // 8) + 0.0015 * (yy // 8) + (xx >= W // 2) * 0.05
        inv = inv + rng.normal(0, 0.003, (H, W)) * (rng.random((H, W)) < 0.2)
        d = (1.0 / inv).astype(np.float32)
        if k == 1:
            d[rng.random((H, W)) < 0.002] = 0.0
        maps.append(d)
    return maps


@pytest.mark.parametrize("k", range(3))
def test_quadtree_prior_matches_kernel_simulation(k):
    d = _prior_depth_maps()[k]
    H, W = d.shape
    rng = np.random.default_rng(k)
    levels = rng.integers(2, 5, (H, W)).astype(np.uint8)
    got = O.quadtree_prior(d, levels)
    assert np.array_equal(got, quadtree_prior_simulated(d, levels))
    assert np.all(got >= levels) and np.any(got > levels)
    # with zero image levels the full depth-level map (0..4) shows
    zero = np.zeros((H, W), np.uint8)
    got0 = O.quadtree_prior(d, zero)
    assert np.array_equal(got0, quadtree_prior_simulated(d, zero))
    assert len(np.unique(got0)) >= 3


def test_quadtree_prior_closed_forms():
    H, W = 32, 48
    base = np.full((H, W), 2, np.uint8)
    # constant prior: every complete 16x16 block is flat -> level 4
    assert np.all(O.quadtree_prior(np.full((H, W), 3.0, np.float32), base) == 4)
    # inverse depth ramp of slope 1/512 per pixel: a 2^i block deviates by
    # (2^i - 1)/1024 from its mean: <= 0.01 up to 8x8, 15/1024 > 0.01 at 16x16
    yy, xx = np.mgrid[0:H, 0:W]
    inv = (0.25 + xx / 512.0).astype(np.float32)
    assert np.all(O.quadtree_prior((1.0 / inv).astype(np.float32), base) == 3)
    # one pixel without prior (0, negative or NaN) keeps its whole tile unchanged
    for bad in (0.0, -1.0, np.nan):
        d = np.full((H, W), 3.0, np.float32)
        d[5, 20] = bad
        got = O.quadtree_prior(d, base)
        assert np.all(got[0:16, 16:32] == 2) and np.all(got[16:32, :] == 4) and np.all(got[0:16, 0:16] == 4)
    # ragged edge: blocks reaching past the image never complete
    d = np.full((20, 24), 3.0, np.float32)
    got = O.quadtree_prior(d, np.full((20, 24), 2, np.uint8))
    assert np.all(got[:16, :16] == 4) and np.all(got[16:, :] == 2) and np.all(got[:16, 16:] == 3)


# prior_to_cost: the reference's atomics replayed in thread order (prior pixels
# in raster order, then the image cost), add_num counted per sample thread as
# the 64 x 1 and 64 x 5 launches do, then normalize_the_cost; plus closed forms.
def _prior_to_cost_replay(prior, levels, cost_in, cnt_in, N, imin, imax):
    H, W = levels.shape
    Hc, Wc = H // 4, W // 4
    f32 = np.float32
    step = (imax - imin) / (N - 1)
    acc = np.zeros((Hc, Wc, N), f32)
    num = np.zeros((Hc, Wc), np.int64)
    for y in range(H):  # prior_to_cost: one block per pixel, one thread per sample
        for x in range(W):
            if x >= W - 3 or y >= H - 3 or x <= 2 or y <= 2:
                continue
            p, v = prior[y, x]
            if p <= 0 or v <= 0:
                continue
            for d in range(N):
                s = f32(step * d + imin)
                c = f32(0.1 * float(f32(s - p)) * float(f32(s - p)) / float(v))
                c = c if c < 0.5 else f32(0.5)
                acc[y // 4, x // 4, d] = f32(acc[y // 4, x // 4, d] + c)
                num[y // 4, x // 4] += 1
    out = np.zeros_like(cost_in)
    for cy in range(Hc):  # image_to_cost adds the raw image cost, then normalize_the_cost
        for cx in range(Wc):
            raw = cost_in[cy, cx] * f32(64.0)  # the cost oracle's volume is raw / 64
            a = acc[cy, cx] + raw if cnt_in[cy, cx] else acc[cy, cx]
            n = num[cy, cx] + (64 if cnt_in[cy, cx] else 0)
            out[cy, cx] = a / f32(n) if n > 0 else a
    return out, (num + cnt_in).astype(np.int32)


def _prior_case(H, W, seed):
    rng = np.random.default_rng(seed)
    key = rng.integers(0, 256, (H, W)).astype(np.uint8)
    lv = O.quadtree(key)
    T = np.tile(np.eye(3, 4, dtype=np.float32), (4, 1, 1))
    T[:, 0, 3] = [0.05, 0.05, -0.04, 0.03]
    cost, cnt = O.cost_volume(key, [np.roll(key, s, axis=1) for s in (1, 1, 2, -1)], T, CAM, 64, 0.02, 2.0, lv)
    prior = np.zeros((H, W, 2), np.float32)
    m = rng.random((H, W)) < 0.3
    prior[m, 0] = rng.uniform(0.02, 2.0, m.sum())
    prior[m, 1] = rng.uniform(0.005, 0.5, m.sum())
    prior[rng.random((H, W)) < 0.02] = [np.nan, 0.1]
    prior[rng.random((H, W)) < 0.02] = [0.5, -1.0]
    return prior, lv, cost, cnt


def test_prior_to_cost_matches_atomic_replay():
    prior, lv, cost, cnt = _prior_case(24, 32, 5)
    got_c, got_n = O.prior_to_cost(prior, lv, cost, cnt, 64, 0.02, 2.0)
    want_c, want_n = _prior_to_cost_replay(prior, lv, cost, cnt, 64, 0.02, 2.0)
    assert np.array_equal(got_n, want_n)
    assert np.array_equal(got_c, want_c)


def test_prior_to_cost_closed_forms():
    H, W = 32, 48
    levels = np.full((H, W), 4, np.uint8)  # nodes only at multiples of 16 ...
    levels[8, 12] = 2                      # ... and at cell (3, 2)
    Hc, Wc, N = H // 4, W // 4, 64
    cost = np.zeros((Hc, Wc, N), np.float32)
    cnt = np.zeros((Hc, Wc), np.int32)
    cost[2, 3] = np.linspace(0.0, 0.01, N, dtype=np.float32)
    cnt[2, 3] = 64
    prior = np.zeros((H, W, 2), np.float32)
    prior[9, 13] = [0.8, 0.05]      # one pixel in node cell (3, 2)
    prior[16:20, 20:24] = [0.3, 0.2]  # the full 4 x 4 block of non-node cell (5, 4)
    prior[1, 1] = [0.5, 0.1]        # border pixel: ignored
    prior[25, 30] = [np.nan, 0.1]   # NaN passes the "<= 0" tests and truncates to 0.5
    got, n = O.prior_to_cost(prior, levels, cost, cnt, N, 0.02, 2.0)
    s = (0.02 + (2.0 - 0.02) / (N - 1) * np.arange(N)).astype(np.float32).astype(np.float64)
    q = lambda p, v: np.minimum(0.1 * (s - p) ** 2 / v, 0.5)
    want = (q(np.float32(0.8), np.float32(0.05)) + 64.0 * cost[2, 3]) / 128.0
    np.testing.assert_allclose(got[2, 3], want, rtol=2e-6, atol=1e-9)
    np.testing.assert_allclose(got[4, 5], q(np.float32(0.3), np.float32(0.2)) / 64.0, rtol=2e-6)
    assert np.all(got[6, 7] == np.float32(0.5 / 64.0))
    assert (n[2, 3], n[4, 5], n[6, 7], n[0, 0], n[4, 4]) == (128, 1024, 64, 0, 0)
    untouched = np.ones((Hc, Wc), bool)
    untouched[2, 3] = untouched[4, 5] = untouched[6, 7] = False
    assert np.array_equal(got[untouched], cost[untouched])


def test_sd_prior_projection_identity_pose():
    """Under the identity pose seeds with validity >= 10 give the prior
    (s_idepth / |K^-1 (x, y, 1)|, max(|K^-1 (x, y, 1)|^2 s_var, 0.01)) at their
    own pixel where the gradient is at least 5, and (0, 0) elsewhere."""
    from open_quadtree_mapping_amd.scene import Camera
    H, W = 40, 48
    cam = Camera(W, H, 60.0, 60.0, W / 2, H / 2)
    img = _textured(H, W, 2)
    g = O.sd_gradient(img)
    s = _static_seeds(H, W, 0.25, 0.001, 12.0)
    s[..., 7] = np.where(np.arange(W)[None, :] < W // 2, 0.004, 0.03).astype(np.float32)
    T = np.eye(3, 4, dtype=np.float32)
    pre, prior = O.sd_step(s, img, g, img, g, cam, T, T, False, want_prior=True)
    assert np.all(pre == 0)  # validity 12 < 20: nothing in the depth pre-fill
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    r2 = ((xx - cam.cx) / cam.fx) ** 2 + ((yy - cam.cy) / cam.fy) ** 2 + 1.0
    strong = (g[..., 0] ** 2 + g[..., 1] ** 2) >= 25.0
    inner = np.zeros((H, W), bool)
    inner[5:H - 5, 5:W - 5] = True
    m = strong & inner
    assert m.sum() > 100
    np.testing.assert_allclose(prior[m, 0], (s[..., 6] / np.sqrt(r2))[m], rtol=1e-6)
    np.testing.assert_allclose(prior[m, 1], np.maximum(r2 * s[..., 7], 0.01)[m], rtol=2e-6)
    assert np.all(prior[~strong] == 0)


# ---------------------------------------------------------------------------
# Fusion hole filling (src/depth_fusion.cu:121-144), pinned by a source-centric
# simulation of the reference kernel: every source thread with 2 <= x <= W-2,
# 2 <= y <= H-2 and a non-empty entry CASes its entry into the empty cells of
# its 3x3 neighbourhood (y offset outer, x offset inner). The schedule
# simulated is "every thread reads its own entry, then the threads run their
# CAS loops one after another in raster order": one legal execution of the
# racy kernel. With the identity pose each good seed warps onto its own pixel,
# so the warp table before filling is known from the good/bad seed pattern.
# ---------------------------------------------------------------------------
def hole_fill_reference_kernel(table):
    H, W = table.shape
    snap = table.copy()
    out = table.copy()
    for y in range(H):
        for x in range(W):
            if x >= W - 1 or y >= H - 1 or x <= 1 or y <= 1:
                continue
            ti = snap[y, x]
            if ti == 0:
                continue
            for i in (-1, 0, 1):
                for j in (-1, 0, 1):
                    if out[y + i, x + j] == 0:
                        out[y + i, x + j] = ti
    return out


def _fusion_table(good, W, H):
    cam = Camera(W, H, 40.0, 40.0, W / 2.0, H / 2.0)
    seeds = np.zeros((H, W, 4), np.float32)
    seeds[..., 0] = np.where(good, 10.0, 0.0)   # a < 0.001: a bad point, not warped
    seeds[..., 1] = 1.0
    seeds[..., 2] = 5.0
    seeds[..., 3] = 1.0
    depth = np.full((H, W), 5.0, np.float32)
    _, _, _, table = O.fuse(seeds, depth, np.eye(3, 4, dtype=np.float32), cam)
    warped = np.where(good, np.arange(W * H, dtype=np.int32).reshape(H, W), 0).astype(np.int32)
    warped[0, 0] = 0  # source index 0 is the empty sentinel: pixel (0, 0) is never stored
    return table, warped


@pytest.mark.parametrize("seed,density", [(1, 0.5), (2, 0.2), (3, 0.8), (4, 0.05)])
def test_fusion_hole_filling_matches_reference_kernel_simulation(seed, density):
    W, H = 23, 17
    good = np.random.default_rng(seed).random((H, W)) < density
    good[0, 0] = True
    table, warped = _fusion_table(good, W, H)
    assert np.array_equal(table, hole_fill_reference_kernel(warped))


def test_fusion_hole_filling_constructed_cases():
    W, H = 12, 10
    # two sources compete for one hole: the raster-first source wins
    good = np.zeros((H, W), bool)
    good[4, 6] = good[5, 4] = True
    table, _ = _fusion_table(good, W, H)
    assert table[5, 5] == 4 * W + 6          # (6, 4) precedes (4, 5) in raster order
    assert table[4, 5] == 4 * W + 6 and table[6, 3] == 5 * W + 4
    # no cascade: cells two steps away stay empty
    assert table[7, 7] == 0 and table[2, 2] == 0
    # the source band is 2 <= x <= W-2, 2 <= y <= H-2: x = 1 and x = W-1 sources fill nothing
    good = np.zeros((H, W), bool)
    good[5, 1] = good[5, W - 1] = good[1, 6] = good[H - 1, 6] = True
    table, _ = _fusion_table(good, W, H)
    filled = table != 0
    assert filled.sum() == 4 and filled[5, 1] and filled[5, W - 1] and filled[1, 6] and filled[H - 1, 6]
    # x = 2 and x = W-2 sources do fill
    good = np.zeros((H, W), bool)
    good[5, 2] = good[5, W - 2] = True
    table, _ = _fusion_table(good, W, H)
    assert table[4, 1] == 5 * W + 2 and table[6, W - 1] == 5 * W + W - 2
    # the sentinel: pixel (0, 0)'s own seed stores index 0, i.e. nothing
    good = np.zeros((H, W), bool)
    good[0, 0] = True
    table, _ = _fusion_table(good, W, H)
    assert np.all(table == 0)


# ---------------------------------------------------------------------------
# Semi-dense epipolar search (src/lsd_semidense.cu:529-753): a pure x
# translation of a texture that varies along x only. A key pixel at inverse
# depth q shows in the income frame at u = x + fx tx q, so with an income image
# that is the key shifted by an integer disparity the search returns q to a
# small fraction of a pixel, and the variance is the reference's model in
# closed form: alpha = d(idepth)/du = 1 / (fx tx) and
# var = alpha^2 (0.05 [subpixel] + 0.25 gmax^2 / g^2 + 4^2 / g^2).
# ---------------------------------------------------------------------------
SD_W, SD_H = 160, 64
SD_CAM = Camera(SD_W, SD_H, 100.0, 100.0, 80.0, 32.0)


def _sd_signal(n, kind="smooth"):
    x = np.arange(n)
    if kind == "smooth":
        return np.round(128 + 50 * np.sin(x / 3.1) + 30 * np.sin(x / 7.3 + 1.0)).astype(np.int64)
    return np.where((x // 2) % 2 == 0, 90, 170).astype(np.int64)  # period 4: ambiguous


def _sd_pair(disp_px, kind="smooth"):
    sig = _sd_signal(SD_W + 64, kind)
    key = np.tile(sig[32:32 + SD_W], (SD_H, 1)).astype(np.uint8)
    inc = np.tile(sig[32 - disp_px:32 - disp_px + SD_W], (SD_H, 1)).astype(np.uint8)
    return key, inc


@pytest.mark.parametrize("idepth,tx", [(0.6, 0.1), (0.5, 0.08), (1.0, 0.05)])
def test_sd_search_recovers_translation_depth_and_closed_form_variance(idepth, tx):
    disp = 100.0 * float(np.float32(tx)) * idepth
    assert abs(disp - round(disp)) < 1e-6
    key, inc = _sd_pair(int(round(disp)))
    kg, ig = O.sd_gradient(key), O.sd_gradient(inc)
    k2i = np.eye(3, 4, dtype=np.float32)
    k2i[0, 3] = tx
    alpha = 1.0 / (100.0 * float(np.float32(tx)))
    checked = 0
    for x in range(20, 110, 3):
        g = float(kg[32, x, 0])
        if abs(g) < 5.0:
            continue
        for gmax in (abs(g), 2.0 * abs(g)):
            err, q, var, _ = O.sd_search_point(x, 32, 1.0, 0.0, gmax, abs(g), 0.01, 2.0, SD_CAM, k2i, key, kg, inc,
                                               ig)
            assert err >= 0.0, (x, err)
            assert abs(q - idepth) * 100.0 * tx <= 0.05, (x, q)      # within 0.05 px
            want = alpha * alpha * (0.05 + 0.25 * gmax * gmax / (g * g) + 16.0 / (g * g))
            assert var == pytest.approx(want, rel=2e-6), (x, var, want)
            checked += 1
    assert checked >= 30


def test_sd_bilinear_fetch_uses_8_bit_fraction_weights():
    # CUDA's linear filtering (Programming Guide, texture fetching appendix)
    # stores the weights in 9-bit fixed point with 8 fractional bits. On a ramp
    # T(i, j) = 4 i + 2 j the fetch at (i0 + 0.5 + f, j0 + 0.5 + g) is therefore
    # 4 (i0 + round(256 f) / 256) + 2 (j0 + round(256 g) / 256), not the exact
    # 4 (i0 + f) + 2 (j0 + g); clamped at the border.
    W, H = 40, 30
    yy, xx = np.mgrid[0:H, 0:W]
    img = (4 * xx + 2 * yy).astype(np.uint8)
    rng = np.random.default_rng(12)
    exact_differs = 0
    for _ in range(400):
        i0, j0 = int(rng.integers(0, W - 1)), int(rng.integers(0, H - 1))
        f, g = float(np.float32(rng.random())), float(np.float32(rng.random()))
        u, v = np.float32(i0 + 0.5 + f), np.float32(j0 + 0.5 + g)
        fu = float(u) - 0.5 - i0  # the fraction the fetch sees
        fv = float(v) - 0.5 - j0
        got = O.sd_tex_bilinear(img, u, v)
        want = 4 * (i0 + np.rint(256 * fu) / 256) + 2 * (j0 + np.rint(256 * fv) / 256)
        assert got == pytest.approx(want, abs=1e-5), (i0, j0, f, g)
        exact_differs += abs(got - (4 * (i0 + fu) + 2 * (j0 + fv))) > 1e-3
    assert exact_differs > 100
    # clamp addressing: beyond the last texel the fetch holds the edge value
    assert O.sd_tex_bilinear(img, W + 3.0, 5.5) == float(img[5, W - 1])
    assert O.sd_tex_bilinear(img, -2.0, 5.5) == float(img[5, 0])


def test_sd_search_failure_codes():
    k2i = np.eye(3, 4, dtype=np.float32)
    k2i[0, 3] = 0.1
    key, inc = _sd_pair(6)
    kg, ig = O.sd_gradient(key), O.sd_gradient(inc)
    g = float(kg[32, 20, 0])
    # the far end of the search range leaves the image: -1
    err, *_ = O.sd_search_point(0, 32, 1.0, 0.0, abs(g), abs(g), 0.01, 2.0, SD_CAM, k2i, key, kg, inc, ig)
    assert err == -1.0
    # no match anywhere (a black income image): -3
    err, *_ = O.sd_search_point(20, 32, 1.0, 0.0, abs(g), abs(g), 0.01, 2.0, SD_CAM, k2i, key, kg, 0 * inc, ig)
    assert err == -3.0
    # a period-4 texture: the second best match is 4 px away and as good: -2
    key, inc = _sd_pair(6, kind="periodic")
    kg, ig = O.sd_gradient(key), O.sd_gradient(inc)
    x = next(x for x in range(20, 40) if abs(kg[32, x, 0]) >= 5)
    g = float(kg[32, x, 0])
    err, *_ = O.sd_search_point(x, 32, 1.0, 0.0, abs(g), abs(g), 0.01, 2.0, SD_CAM, k2i, key, kg, inc, ig)
    assert err == -2.0


# ---------------------------------------------------------------------------
# Semi-dense hole filling (src/lsd_semidense.cu:345-387): an invalid interior
# pixel with enough valid neighbours in its 5x5 window (validity sum > 100, or
# > 24 when not blacklisted below -1) and an income gradient of at least 5
# becomes a seed with the inverse-variance weighted mean of the valid
# neighbours: idepth = sum(q / v) / sum(1 / v), variance = 1 / sum(1 / v),
# validity counter 0. Fills read a snapshot: no cascade.
# ---------------------------------------------------------------------------
def _sd_seeds(W, H):
    return np.zeros((H, W, 8), np.float32)


def _put_seed(s, x, y, q, var, counter, blacklist=0.0):
    s[y, x] = [1.0, blacklist, counter, 0.0, q, var, 1.0, 4.0]


def test_sd_fill_holes_inverse_variance_mean():
    W, H = 20, 16
    rng = np.random.default_rng(5)
    s = _sd_seeds(W, H)
    nb = [(8 + i, 7 + j) for j in range(-2, 3) for i in range(-2, 3) if (i, j) != (0, 0) and rng.random() < 0.6]
    qs = rng.uniform(0.2, 1.5, len(nb)).astype(np.float32)
    vs = rng.uniform(0.01, 0.5, len(nb)).astype(np.float32)
    for (x, y), q, v in zip(nb, qs, vs):
        _put_seed(s, x, y, q, v, 20.0)
    grad = np.full((H, W, 2), 4.0, np.float32)  # |g|^2 = 32 >= 25 everywhere
    out = O.sd_fill_holes(s, grad)
    w = 1.0 / vs.astype(np.float64)
    c = out[7, 8]
    assert c[0] == 1.0 and c[2] == 0.0 and c[1] == 0.0
    assert c[4] == pytest.approx(float((qs * w).sum() / w.sum()), rel=2e-6)
    assert c[5] == pytest.approx(float(1.0 / w.sum()), rel=2e-6)
    # valid seeds are left alone
    for (x, y) in nb:
        assert np.array_equal(out[y, x], s[y, x])


def test_sd_fill_holes_thresholds_and_snapshot():
    W, H = 24, 16
    grad = np.full((H, W, 2), 4.0, np.float32)
    # validity sum exactly 24 does not create, 25 does
    s = _sd_seeds(W, H)
    _put_seed(s, 6, 6, 0.5, 0.1, 24.0)
    out = O.sd_fill_holes(s, grad)
    assert out[6, 7, 0] == 0.0
    s[6, 6, 2] = 25.0
    out = O.sd_fill_holes(s, grad)
    assert out[6, 7, 0] == 1.0 and out[6, 7, 4] == pytest.approx(0.5, rel=1e-7)
    # a hole blacklisted below -1 needs a sum above 100
    s = _sd_seeds(W, H)
    _put_seed(s, 6, 6, 0.5, 0.1, 60.0)
    s[6, 7, 1] = -2.0
    out = O.sd_fill_holes(s, grad)
    assert out[6, 7, 0] == 0.0 and out[6, 8, 0] == 1.0
    s[6, 6, 2] = 101.0
    out = O.sd_fill_holes(s, grad)
    assert out[6, 7, 0] == 1.0
    # an income gradient below 5 creates nothing
    low = grad.copy()
    low[6, 7] = 3.0  # |g|^2 = 18
    out = O.sd_fill_holes(s, low)
    assert out[6, 7, 0] == 0.0 and out[6, 8, 0] == 1.0
    # snapshot: a fill two pixels beyond the window of the only source stays empty
    assert out[6, 9, 0] == 0.0 and out[6, 8, 0] == 1.0 and out[6, 4, 0] == 1.0


# ---------------------------------------------------------------------------
# Keyframe propagation (src/lsd_semidense.cu:64-181 via src/seed_matrix.cu:
# 247-304): seeds of a fronto-parallel plane at inverse depth 0.5 (a power of
# two, so the regularisation's weighted means of a constant field are exact)
# land where the motion puts them with the inverse depth it implies; of two
# seeds that land on one pixel the larger inverse depth wins.
# ---------------------------------------------------------------------------
def _sd_textured(W, H):
    yy, xx = np.mgrid[0:H, 0:W]
    img = (100 + 10 * ((xx + yy) % 4)).astype(np.uint8)  # |gx| = |gy| = 10 everywhere inside
    return img, O.sd_gradient(img)


def test_sd_propagation_identity_and_forward_motion():
    W, H = 48, 40
    cam = Camera(W, H, 40.0, 40.0, 24.0, 20.0)
    img, grad = _sd_textured(W, H)
    old = _sd_seeds(W, H)
    for y in range(8, 32):
        for x in range(8, 40):
            old[y, x] = [1.0, 0.0, 30.0, 0.0, 0.5, 0.01, 0.5, 0.01]
    out = O.sd_new_keyframe(old, cam, np.eye(3, 4, dtype=np.float32), img, grad, img, grad)
    inner = (slice(10, 30), slice(10, 38))
    assert np.all(out[inner][..., 0] == 1.0)
    assert np.all(out[inner][..., 4] == 0.5) and np.all(out[inner][..., 6] == 0.5)
    # forward by 0.5 m: depth 2 -> 1.5, so inverse depth 1 / 1.5 wherever a seed lands
    T = np.eye(3, 4, dtype=np.float32)
    T[2, 3] = -0.5
    out = O.sd_new_keyframe(old, cam, T, img, grad, img, grad)
    v = out[..., 0] == 1.0
    assert v.sum() > 300
    assert np.allclose(out[v][:, 4], 1.0 / 1.5, rtol=1e-6, atol=0)
    # the points spread out from the principal point by 2 / 1.5
    ys, xs = np.nonzero(v)
    assert xs.min() < 8 - 2 and xs.max() > 40 + 2


def test_sd_propagation_collision_larger_inverse_depth_wins():
    W, H = 48, 40
    cam = Camera(W, H, 40.0, 40.0, 24.0, 20.0)
    img, grad = _sd_textured(W, H)
    T = np.eye(3, 4, dtype=np.float32)
    T[0, 3] = 0.2  # fx tx = 8: inverse depth q moves 8 q pixels
    for qa, qb in [(0.5, 0.25), (0.25, 0.5)]:
        old = _sd_seeds(W, H)
        # A at x = 20 moves 8 qa, B at x = 20 + 8 (qa - qb) moves 8 qb: both land on x = 20 + 8 qa
        xa, xb, y = 20, 20 + int(8 * (qa - qb)), 20
        old[y, xa] = [1.0, 0.0, 30.0, 0.0, qa, 0.01, qa, 0.01]
        old[y, xb] = [1.0, 0.0, 30.0, 0.0, qb, 0.01, qb, 0.01]
        out = O.sd_new_keyframe(old, cam, T, img, grad, img, grad)
        tx_ = 20 + int(8 * qa)
        assert out[y, tx_, 0] == 1.0
        assert out[y, tx_, 4] == max(qa, qb)


# ---------------------------------------------------------------------------
# Point cloud (src/publisher.cpp:129-164): under the identity pose every point
# lies at distance d along its pixel's ray; a translation shifts every point
# by it; a rotation by 90 degrees about z maps (x, y, z) to (-y, x, z); pixels
# below 0.1 are left out, NaN kept, raster order.
# ---------------------------------------------------------------------------
def test_point_cloud_closed_forms():
    W, H = 24, 16
    cam = Camera(W, H, 20.0, 22.0, 11.5, 7.5)
    rng = np.random.default_rng(9)
    depth = rng.uniform(0.2, 9.0, (H, W)).astype(np.float32)
    depth[3, 4] = 0.05
    depth[5, 6] = -1.0
    depth[7, 8] = np.nan
    img = rng.integers(0, 256, (H, W)).astype(np.uint8)
    I = np.eye(3, 4, dtype=np.float32)
    pts = O.point_cloud(depth, img, cam, I)
    keep = ~(depth.astype(np.float64) < 0.1)
    assert len(pts) == keep.sum() == W * H - 2
    ys, xs = np.nonzero(keep)
    d = depth[ys, xs].astype(np.float64)
    fin = np.isfinite(d)
    assert np.allclose(np.linalg.norm(pts[fin, :3].astype(np.float64), axis=1), d[fin], rtol=1e-6)
    ray = np.stack([(xs - 11.5) / 20.0, (ys - 7.5) / 22.0, np.ones_like(d)], 1)
    ray /= np.linalg.norm(ray, axis=1, keepdims=True)
    assert np.allclose(pts[fin, :3], ray[fin] * d[fin, None], rtol=1e-5, atol=1e-6)
    assert np.isnan(pts[~fin, :3]).all()
    assert np.array_equal(pts[:, 3], img[ys, xs].astype(np.float32))
    T = I.copy()
    T[:, 3] = [0.5, -2.0, 3.0]
    pt = O.point_cloud(depth, img, cam, T)
    assert np.allclose(pt[fin, :3], pts[fin, :3] + T[:, 3], rtol=1e-6, atol=1e-6)
    R = np.zeros((3, 4), np.float32)
    R[0, 1], R[1, 0], R[2, 2] = -1.0, 1.0, 1.0
    pr = O.point_cloud(depth, img, cam, R)
    assert np.allclose(pr[fin, 0], -pts[fin, 1], rtol=1e-6, atol=1e-7)
    assert np.allclose(pr[fin, 1], pts[fin, 0], rtol=1e-6, atol=1e-7)
    assert np.allclose(pr[fin, 2], pts[fin, 2], rtol=1e-6, atol=1e-7)
