"""Hand-written adjoints (upsample solve, fusion update) against FP64 central
finite differences of the FP64 oracle (h = 1e-6), through torch.autograd."""
import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd import autograd as QA
from open_quadtree_mapping_amd.scene import Camera

pytestmark = pytest.mark.gpu
DEV = "cuda"
H_FD = 1e-6


def problem(H, W, seed):
    """Config-3 distributions: depths U(1,10) on 30% of pixels, confidences
    U(0,1], weights log-U(1e-3, 1e1)."""
    rng = np.random.default_rng(seed)
    d = (rng.uniform(1, 10, (H, W)) * (rng.random((H, W)) < 0.3)).astype(np.float64)
    conf = rng.uniform(0, 1, (H, W))
    conf = np.where(conf == 0, 1.0, conf)
    wr = np.exp(rng.uniform(np.log(1e-3), np.log(10), (H, W + 1)))
    wc = np.exp(rng.uniform(np.log(1e-3), np.log(10), (H + 1, W)))
    r = rng.standard_normal((H, W))
    return d, conf, wr, wc, r


def fd_upsample(arrays, which, idx, r, lam, row_seg=None, col_seg=None):
    def L(arrs):
        d, conf, wr, wc = arrs
        out = O.upsample(d, wr, wc, conf=conf, lam=lam, row_seg=row_seg, col_seg=col_seg, fp64=True)
        return float(np.sum(r * out))
    vals = []
    for sgn in (1, -1):
        a = [x.copy() for x in arrays]
        a[which][idx] += sgn * H_FD
        vals.append(L(a))
    return (vals[0] - vals[1]) / (2 * H_FD)


# Error of one gradient entry against its FP64 central difference, relative to
# max(|fd|, 1e-3 * max |gradient of that input|): entries whose derivative is
# tiny next to the input's others are held to an absolute bound at that floor.
def grad_err(g, fd, grads_of_input):
    scale = max(abs(fd), 1e-3 * float(np.abs(grads_of_input).max()), 1e-9)
    return abs(g - fd) / scale


def split_entries(H, W, row_seg, col_seg, rng, per=3):
    """Weight entries at and beside the chain splits and the line ends, for
    every split that is one: (which, index) pairs, which 2 = w_row, 3 = w_col."""
    out = []
    for c in sorted({0, 1, row_seg - 1, row_seg, row_seg + 1, W - 1, W}):
        if 0 <= c <= W:
            out += [(2, (int(rng.integers(0, H)), c)) for _ in range(per)]
    for r_ in sorted({0, 1, col_seg - 1, col_seg, col_seg + 1, H - 1, H}):
        if 0 <= r_ <= H:
            out += [(3, (r_, int(rng.integers(0, W)))) for _ in range(per)]
    for c in sorted({row_seg - 1, row_seg}):
        if 0 <= c < W:
            out += [(0, (int(rng.integers(0, H)), c)), (1, (int(rng.integers(0, H)), c))]
    return out


@pytest.mark.parametrize("H,W,row_seg,col_seg,n_entries", [(120, 188, None, None, 256), (480, 752, None, None, 64),
                                                           (240, 376, 368, 240, 128), (240, 376, 373, 229, 128),
                                                           (100, 150, 37, 61, 128)])
def test_upsample_adjoint_vs_fd(H, W, row_seg, col_seg, n_entries):
    print(f"upsample adjoint {W}x{H} splits {row_seg} {col_seg}", flush=True)
    d, conf, wr, wc, r = problem(H, W, 5)
    lam = 10.0
    td = torch.tensor(d, dtype=torch.float32, device=DEV, requires_grad=True)
    tc = torch.tensor(conf, dtype=torch.float32, device=DEV, requires_grad=True)
    twr = torch.tensor(wr, dtype=torch.float32, device=DEV, requires_grad=True)
    twc = torch.tensor(wc, dtype=torch.float32, device=DEV, requires_grad=True)
    out = QA.upsample(td, twr, twc, conf=tc, lam=lam, row_seg=row_seg, col_seg=col_seg)
    rs = W if row_seg is None else row_seg
    cs = H if col_seg is None else col_seg
    want = O.upsample(d.astype(np.float32), wr.astype(np.float32), wc.astype(np.float32),
                      conf=conf.astype(np.float32), lam=lam, row_seg=rs, col_seg=cs)
    assert np.array_equal(out.detach().cpu().numpy(), want, equal_nan=True)
    loss = (out * torch.tensor(r, dtype=torch.float32, device=DEV)).sum()
    loss.backward()
    grads = [td.grad.cpu().numpy(), tc.grad.cpu().numpy(), twr.grad.cpu().numpy(), twc.grad.cpu().numpy()]
    arrays = [d, conf, wr, wc]
    rng = np.random.default_rng(1)
    entries = []
    for k in range(n_entries):
        which = k % 4
        entries.append((which, tuple(int(rng.integers(0, s)) for s in arrays[which].shape)))
    if row_seg is not None:
        entries += split_entries(H, W, rs, cs, rng)
    errs = []
    for which, idx in entries:
        fd = fd_upsample(arrays, which, idx, r, lam, rs, cs)
        errs.append(grad_err(float(grads[which][idx]), fd, grads[which]))
    errs = np.array(errs)
    print(f"  {len(errs)} entries: max {errs.max():.2e} p99 {np.quantile(errs, 0.99):.2e}", flush=True)
    assert errs.max() < 2e-3, f"max {errs.max()} at entry {entries[int(errs.argmax())]}"


def test_upsample_lambda_gradient():
    H, W = 60, 94
    d, conf, wr, wc, r = problem(H, W, 9)
    lam = torch.tensor(10.0, dtype=torch.float32, device=DEV, requires_grad=True)
    out = QA.upsample(torch.tensor(d, dtype=torch.float32, device=DEV),
                      torch.tensor(wr, dtype=torch.float32, device=DEV),
                      torch.tensor(wc, dtype=torch.float32, device=DEV),
                      conf=torch.tensor(conf, dtype=torch.float32, device=DEV), lam=lam)
    (out * torch.tensor(r, dtype=torch.float32, device=DEV)).sum().backward()
    f = lambda l: float(np.sum(r * O.upsample(d, wr, wc, conf=conf, lam=l, fp64=True)))  # noqa: E731
    fd = (f(10.0 + 1e-6) - f(10.0 - 1e-6)) / 2e-6
    assert abs(lam.grad.item() - fd) <= 2e-3 * abs(fd)


def test_fuse_adjoint_vs_fd():
    H, W = 96, 128
    cam = Camera(W, H, 100.0, 100.0, 64.0, 48.0)
    rng = np.random.default_rng(3)
    seeds = np.zeros((H, W, 4))
    seeds[..., 0] = rng.uniform(5, 30, (H, W))
    seeds[..., 1] = rng.uniform(1, 10, (H, W))
    seeds[..., 2] = rng.uniform(2, 6, (H, W))
    seeds[..., 3] = rng.uniform(0.05, 2, (H, W))
    depth = seeds[..., 2] * rng.uniform(0.9, 1.1, (H, W))
    reset = rng.random((H, W)) < 0.1  # measurements far from the prior: the reset branch
    depth[reset] = rng.uniform(1, 8, (H, W))[reset]
    T = np.eye(3, 4, dtype=np.float32)
    T[0, 3] = 0.01
    gdep = rng.standard_normal((H, W))
    gse = rng.standard_normal((H, W, 4))
    td = torch.tensor(depth, dtype=torch.float32, device=DEV, requires_grad=True)
    ts = torch.tensor(seeds, dtype=torch.float32, device=DEV)
    dout, sout = QA.fuse(td, ts, (cam.fx, cam.fy, cam.cx, cam.cy), T)
    loss = (dout * torch.tensor(gdep, dtype=torch.float32, device=DEV)).sum() + \
        (sout * torch.tensor(gse, dtype=torch.float32, device=DEV)).sum()
    loss.backward()
    g = td.grad.cpu().numpy()

    def L(dep):
        _, o, s, _ = O.fuse(seeds, dep, T, cam, fp64=True)
        return float(np.sum(gdep * np.where(o == -1.0, 0.0, o)) + np.sum(gse * s))
    # the -1 outputs are constants; their gradient is 0 on both sides.
    # Excluded by rule: pixels where the central difference straddles a
    # discontinuity or kink of the update in the measured depth d -- the reset
    # test (d - mu)^2 > tau^2 + s2 changing sign within d +- h, or the kink of
    # tau = max(0.5, 0.2 d) at d = 2.5 within d +- h (prior = the warped seed
    # the hole-filled table names, held fixed).
    warped, _, _, table = O.fuse(seeds, depth, T, cam, fp64=True)
    flat = warped.reshape(-1, 4)

    def straddles(y, x):
        t = int(table[y, x])
        if t > 0:
            mu, s2 = flat[t, 2], flat[t, 3]
        else:
            mu, s2 = depth[y, x], 2500.0
        sgn = [np.sign((dd - mu) ** 2 - (max(0.5, 0.2 * dd) ** 2 + s2)) for dd in (depth[y, x] - H_FD,
                                                                                  depth[y, x] + H_FD)]
        return sgn[0] != sgn[1] or abs(depth[y, x] - 2.5) <= H_FD
    errs, excluded = [], 0
    for k in range(256):
        y, x = int(rng.integers(0, H)), int(rng.integers(0, W))
        if straddles(y, x):
            excluded += 1
            continue
        dp, dm = depth.copy(), depth.copy()
        dp[y, x] += H_FD
        dm[y, x] -= H_FD
        fd = (L(dp) - L(dm)) / (2 * H_FD)
        errs.append(grad_err(float(g[y, x]), fd, g))
    errs = np.array(errs)
    print(f"fuse adjoint: {len(errs)} entries checked, {excluded} excluded by rule, max {errs.max():.2e}", flush=True)
    assert excluded <= 8
    # bound 1e-2: the FP32 update itself loses up to ~2e-4 relative in a', b'
    # against FP64 on these seeds (the cancellation in (e - f) / (f - e / f)),
    # and its derivative through the same expression more
    assert errs.max() < 1e-2, f"max {errs.max()}"


def test_upsample_batch16_matches_oracle_and_fd():
    """Config 3: a batch of 16 upsampling problems at 752x480 through one
    launch per pass, with one smoothness weight per problem; forward against
    the FP32 oracle, adjoint against FP64 central differences on 256 entries."""
    B, H, W = 16, 480, 752
    probs = [problem(H, W, 100 + b) for b in range(B)]
    lams = np.linspace(5.0, 20.0, B)
    print("upsample batch of 16 at 752x480", flush=True)
    t = lambda k: torch.tensor(np.stack([p[k] for p in probs]), dtype=torch.float32, device=DEV,  # noqa: E731
                               requires_grad=True)
    td, tc, twr, twc = t(0), t(1), t(2), t(3)
    tl = torch.tensor(lams, dtype=torch.float32, device=DEV, requires_grad=True)
    out = QA.upsample(td, twr, twc, conf=tc, lam=tl)
    r = np.stack([p[4] for p in probs])
    (out * torch.tensor(r, dtype=torch.float32, device=DEV)).sum().backward()
    got = out.detach().cpu().numpy()
    for b in range(B):
        d, conf, wr, wc, _ = probs[b]
        want = O.upsample(d.astype(np.float32), wr.astype(np.float32), wc.astype(np.float32),
                          conf=conf.astype(np.float32), lam=float(lams[b]))
        assert np.array_equal(got[b], want, equal_nan=True), f"problem {b}"
    grads = [td.grad.cpu().numpy(), tc.grad.cpu().numpy(), twr.grad.cpu().numpy(), twc.grad.cpu().numpy()]
    rng = np.random.default_rng(2)
    errs = []
    for k in range(256):
        b, which = int(rng.integers(0, B)), k % 4
        arrays = list(probs[b][:4])
        idx = tuple(int(rng.integers(0, s)) for s in arrays[which].shape)
        fd = fd_upsample(arrays, which, idx, probs[b][4], float(lams[b]))
        errs.append(grad_err(float(grads[which][b][idx]), fd, grads[which][b]))
    errs = np.array(errs)
    print(f"batch of 16: 256 entries, max {errs.max():.2e}", flush=True)
    assert errs.max() < 2e-3, f"max {errs.max()}"
    gl = tl.grad.cpu().numpy()
    for b in (0, 7, 15):
        d, conf, wr, wc, rr = probs[b]
        f = lambda l: float(np.sum(rr * O.upsample(d, wr, wc, conf=conf, lam=l, fp64=True)))  # noqa: E731
        fd = (f(lams[b] + 1e-6) - f(lams[b] - 1e-6)) / 2e-6
        assert abs(gl[b] - fd) <= 2e-3 * abs(fd), f"problem {b}: {gl[b]} vs {fd}"


def test_differentiable_lambda_must_be_positive():
    H, W = 32, 48
    d, conf, wr, wc, _ = problem(H, W, 2)
    t = lambda a: torch.tensor(a, dtype=torch.float32, device=DEV)  # noqa: E731
    lam = torch.tensor(0.0, device=DEV, requires_grad=True)
    with pytest.raises(ValueError, match="> 0"):
        QA.upsample(t(d), t(wr), t(wc), conf=t(conf), lam=lam)
    # lambda = 0 without a gradient is a plain forward (A = I)
    out = QA.upsample(t(d), t(wr), t(wc), conf=t(conf), lam=0.0)
    assert torch.allclose(out, t(d), rtol=1e-6, atol=0)


def test_upsample_backward_on_kept_state_equals_recomputed():
    """The backward given the forward's kept row and column results gives the
    gradients of the backward that recomputes them, bit for bit (and the
    forward's output is the same with or without a state to keep)."""
    from open_quadtree_mapping_amd._lib import check, lib
    B, H, W = 3, 96, 136
    probs = [problem(H, W, 60 + b) for b in range(B)]
    t = lambda k: torch.tensor(np.stack([p[k] for p in probs]), dtype=torch.float32, device=DEV)  # noqa: E731
    d, conf, wr, wc, g = t(0), t(1), t(2), t(3), t(4)
    lams = [7.0, 10.0, 13.0]
    stream = torch.cuda.current_stream().cuda_stream
    state = torch.empty((B, 4, H, W), dtype=torch.float32, device=DEV)
    outs, grads = [], []
    for keep in (False, True):
        out = torch.empty_like(d)
        a = QA._problems(d, conf, wr, wc, lams, 128, 88, dense=out, state=state if keep else None)
        check(lib().qm_upsample_general_batch(a.ctypes.data, B, stream))
        gs = [torch.empty_like(d), torch.empty_like(d), torch.empty_like(wr), torch.empty_like(wc),
              torch.zeros(B, dtype=torch.float64, device=DEV)]
        a = QA._problems(d, conf, wr, wc, lams, 128, 88, g=g, grad_d=gs[0], grad_conf=gs[1], grad_w_row=gs[2],
                         grad_w_col=gs[3], grad_lambda=gs[4], state=state if keep else None)
        check(lib().qm_upsample_backward_batch(a.ctypes.data, B, stream))
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
        grads.append([x.cpu().numpy() for x in gs])
    assert np.array_equal(outs[0], outs[1], equal_nan=True)
    for k, (x, y) in enumerate(zip(*grads)):
        assert np.array_equal(x, y), f"gradient {k}"
