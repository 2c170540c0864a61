"""Batched plus differentiated pipeline (config 5 shape): many 752x480 streams
with 128 inverse-depth samples over 0.3-50 m, the full pipeline forward, and
the backward through the fusion-update and upsample adjoints with respect to
each stream's smoothness weight lambda, the upsample forward recomputed.
Checked against FP64 central differences of the FP64 oracle (the upsample and
the fusion update in double, on the run's own sparse map, warp table and
warped prior seeds)."""
import numpy as np
import pytest
import torch

import oracle as O
from open_quadtree_mapping_amd import default_params
from open_quadtree_mapping_amd import autograd as QA
from open_quadtree_mapping_amd.batch import Batch
from tests._data import cam_tuple, keyframe_set

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_streams", [32])
def test_batch_lambda_gradient_vs_fp64(n_streams):
    W, H, N = 752, 480, 128
    imin, imax = 1.0 / 50.0, 1.0 / 0.3
    sets = [keyframe_set(W, H, 40 + k) for k in range(4)]
    print(f"batched backward: {n_streams} streams {W}x{H} N={N}", flush=True)
    frames = []
    for i in range(n_streams):
        ks = sets[i % len(sets)]
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        frames.append(dict(key=dev(ks["key"]), refs=[dev(r) for r in ks["refs"]], T=ks["T"], cam=cam_tuple(ks["cam"])))
    p = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax)
    batch = Batch(frames, p, outputs=("depth", "dense", "sparse"))
    lams = torch.tensor(8.0 + 0.25 * np.arange(n_streams), dtype=torch.float64, device="cuda", requires_grad=True)
    # a static camera: 16 fusions bring the seeds to confidence, then the
    # differentiated run
    batch.set_lambdas(lams.detach().cpu().numpy())
    for _ in range(16):
        batch.run(graph=True)
    outs = QA.batch_depth(lams, batch, graph=True)
    rng = np.random.default_rng(5)
    g = [rng.standard_normal((H, W)) for _ in range(n_streams)]
    loss = sum((o.double() * torch.tensor(gi, device="cuda")).sum() for o, gi in zip(outs, g))
    loss.backward()
    grad = lams.grad.cpu().numpy()
    assert np.all(np.isfinite(grad))
    checked = 0
    for i in range(n_streams):
        ks = sets[i % len(sets)]
        sparse = batch.outputs[i]["sparse"].cpu().numpy()
        dense = batch.outputs[i]["dense"].cpu().numpy()
        depth = outs[i].detach().cpu().numpy()
        table = batch.read_buffer(i, Batch.BUF_TABLE)
        prior = batch.read_buffer(i, Batch.BUF_PRIOR)
        wr, wc = O.weights(ks["key"])
        rs, cs = O.ref_segments(W, H)
        lam = float(lams[i])
        # the forward of this stream with its own lambda, against the FP32 oracle
        assert np.array_equal(dense, O.upsample(sparse, wr, wc, lam=lam, row_seg=rs, col_seg=cs), equal_nan=True)
        d32, _ = O.fuse_update(table, prior, dense)
        assert np.array_equal(depth, d32, equal_nan=True)

        def f(lv):
            de = O.upsample(sparse.astype(np.float64), wr.astype(np.float64), wc.astype(np.float64), lam=lv,
                            row_seg=rs, col_seg=cs, fp64=True)
            do, _ = O.fuse_update(table, prior.astype(np.float64), de, fp64=True)
            return float(np.sum(np.where(do == -1.0, 0.0, do) * g[i]))
        fd = (f(lam + 1e-6) - f(lam - 1e-6)) / 2e-6
        assert abs(grad[i] - fd) <= 5e-3 * abs(fd) + 1e-6, f"stream {i}: {grad[i]} vs {fd}"
        checked += 1 if (depth > 0).mean() >= 0.05 else 0
    # every stream is asserted; most also carry a substantial confident area
    assert checked >= n_streams // 2
    batch.close()


def _small_batch(n=2, lam=None):
    W, H = 376, 240
    sets = [keyframe_set(W, H, 60 + k) for k in range(n)]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    frames = [dict(key=dev(ks["key"]), refs=[dev(r) for r in ks["refs"]], T=ks["T"], cam=cam_tuple(ks["cam"]))
              for ks in sets]
    kw = {} if lam is None else {"lambda": lam}
    p = default_params(n_hyp=64, idepth_min=0.02, idepth_max=2.0, **kw)
    return Batch(frames, p, outputs=("depth", "dense", "sparse")), sets, frames


def test_batch_backward_refuses_a_changed_batch():
    from open_quadtree_mapping_amd._lib import QmError
    batch, sets, frames = _small_batch()
    g = [torch.ones(240, 376, device="cuda") for _ in range(2)]
    with pytest.raises(QmError, match="no forward run"):
        batch.backward(g)
    batch.run()
    assert batch.runs() == 1
    batch.backward(g)  # the last run is intact
    batch.set_frame(0, frames[0])
    with pytest.raises(QmError, match="changed after its last run"):
        batch.backward(g)
    batch.run(graph=True)
    batch.backward(g)
    batch.autotune(reps=1)
    with pytest.raises(QmError, match="changed after its last run"):
        batch.backward(g)
    batch.run()
    batch.reset_seeds()
    with pytest.raises(QmError, match="changed after its last run"):
        batch.backward(g)
    # the autograd function notices a later run
    lams = torch.tensor([9.0, 11.0], dtype=torch.float64, device="cuda", requires_grad=True)
    outs = QA.batch_depth(lams, batch)
    batch.run()
    with pytest.raises(RuntimeError, match="ran again"):
        sum(o.sum() for o in outs).backward()
    for bad in ([0.0, 1.0], [1.0, -2.0], [float("nan"), 1.0]):
        with pytest.raises(ValueError, match="> 0"):
            batch.set_lambdas(bad)
    batch.close()


def test_batch_backward_on_kept_forward_equals_recomputed():
    """The first backward of a batch recomputes the upsample forward; later
    runs keep their column results for their backward, which then reuses them
    and the run's row results: the same run state gives the same gradient, bit
    for bit."""
    batch, sets, frames = _small_batch()
    rng = np.random.default_rng(4)
    g = [torch.tensor(rng.standard_normal((240, 376)).astype(np.float32), device="cuda") for _ in range(2)]
    got = []
    for graph in (False, False, True):
        batch.reset_seeds()
        for _ in range(3):
            batch.run(graph=graph)
        got.append(np.array(batch.backward(g)))
    assert np.all(np.isfinite(got[0]))
    assert np.array_equal(got[0], got[1]) and np.array_equal(got[0], got[2])
    batch.close()


@pytest.mark.parametrize("lam", [0.0, 1e-3])
def test_batch_lambda_gradient_at_small_lambda(lam):
    """params.lambda = 0 or 1e-3. The lambda gradient is formed as
    sum_e w_e dL/d(lambda w_e), with no division by lambda: at 0 it is finite
    (the upsample there returns sparse / indicator, NaN off the sparse pixels,
    and any lambda != 0 makes those finite, so no difference quotient across 0
    exists to compare with); at 1e-3 it matches the FP64 central difference."""
    batch, sets, frames = _small_batch(lam=lam)
    W, H = 376, 240
    for _ in range(16):
        batch.run()
    batch.run()
    rng = np.random.default_rng(8)
    g = [rng.standard_normal((H, W)).astype(np.float32) for _ in range(2)]
    grad = batch.backward([torch.tensor(gi, device="cuda") for gi in g])
    assert np.all(np.isfinite(grad))
    if lam == 0.0:
        batch.close()
        return
    for i in range(2):
        ks = sets[i]
        sparse = batch.outputs[i]["sparse"].cpu().numpy()
        table = batch.read_buffer(i, Batch.BUF_TABLE)
        prior = batch.read_buffer(i, Batch.BUF_PRIOR)
        wr, wc = O.weights(ks["key"])
        rs, cs = O.ref_segments(W, H)

        def f(lv):
            de = O.upsample(sparse.astype(np.float64), wr.astype(np.float64), wc.astype(np.float64), lam=lv,
                            row_seg=rs, col_seg=cs, fp64=True)
            do, _ = O.fuse_update(table, prior.astype(np.float64), de, fp64=True)
            return float(np.sum(np.where(do == -1.0, 0.0, do) * g[i].astype(np.float64)))
        fd = (f(lam + 1e-6) - f(lam - 1e-6)) / 2e-6
        assert abs(grad[i] - fd) <= 5e-3 * abs(fd) + 1e-6, f"stream {i}: {grad[i]} vs {fd}"
    batch.close()


def test_batch_fusion_state_replayed_and_kept():
    """A batch run writes neither the hole-filled warp table nor the warped
    prior seeds until a backward has asked for them; reading them replays
    both from the run's prior seeds and z-buffer (once: a second read returns
    the same, not a second warp). Replayed, and after a backward written by
    the run itself, they equal the oracle's warp and table for the run's prior
    seeds, and the run's depth and seeds equal the oracle's update."""
    W, H = 376, 240
    sets = [keyframe_set(W, H, 60 + k) for k in range(2)]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    c, s = np.cos(0.01), np.sin(0.01)
    T_last = np.array([[c, 0, s, 0.01], [0, 1, 0, 0.0], [-s, 0, c, 0.005]], np.float32)
    frames = [dict(key=dev(ks["key"]), refs=[dev(r) for r in ks["refs"]], T=ks["T"], cam=cam_tuple(ks["cam"]),
                   T_last=T_last) for ks in sets]
    p = default_params(n_hyp=64, idepth_min=0.02, idepth_max=2.0)
    batch = Batch(frames, p, outputs=("depth", "dense"))

    def check(kept):
        prior_in = [batch.read_buffer(i, Batch.BUF_SEEDS) for i in range(2)]
        batch.run(graph=True)
        torch.cuda.synchronize()
        for i, ks in enumerate(sets):
            dense = batch.outputs[i]["dense"].cpu().numpy()
            warped, dout, sout, table = O.fuse(prior_in[i], dense, T_last, ks["cam"])
            got_t = batch.read_buffer(i, Batch.BUF_TABLE)
            got_p = batch.read_buffer(i, Batch.BUF_PRIOR)
            assert np.array_equal(got_t, table), f"stream {i} table (kept={kept})"
            assert np.array_equal(got_p, warped, equal_nan=True), f"stream {i} warped prior (kept={kept})"
            # a second read neither warps again nor changes the table
            assert np.array_equal(batch.read_buffer(i, Batch.BUF_PRIOR), got_p, equal_nan=True)
            assert np.array_equal(batch.read_buffer(i, Batch.BUF_TABLE), got_t)
            assert np.array_equal(batch.outputs[i]["depth"].cpu().numpy(), dout, equal_nan=True)
            assert np.array_equal(batch.read_buffer(i, Batch.BUF_SEEDS), sout, equal_nan=True)

    for _ in range(4):
        batch.run(graph=True)
    check(kept=False)
    batch.backward([torch.ones(H, W, device="cuda") for _ in range(2)])
    check(kept=True)
    batch.close()
