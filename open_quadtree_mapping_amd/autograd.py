"""Differentiable upsample and fusion: torch.autograd.Functions over the C ABI.

UpsampleFunction: out = x_col / n_col of the generalised upsample
  (qm_upsample_general); backward = qm_upsample_backward, which runs the
  transposed (= same, A is symmetric) tridiagonal solves on the forward's row
  and column results (kept from the forward: 16 bytes per pixel). Gradients w.r.t. sparse depths, confidences, per-pixel smoothness
  weights, and the smoothness scale lambda when it is given as a tensor
  (dL/dlambda = sum_e w_e dL/d(lambda w_e), since A = I + lambda L(w), summed on
  the device in a fixed order; such a lambda must be > 0).
FuseFunction: the Gaussian x Beta update of one keyframe (qm_fuse) with
  backward qm_fuse_backward w.r.t. the dense depth input (prior seeds fixed).
BatchDepthFunction: one batched pipeline run (every stream through every
  stage, Batch) as a function of the per-stream smoothness weights; backward
  = qm_batch_backward (fusion adjoint, then the upsample adjoint with its
  forward recomputed) gives dL/dlambda per stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, device_empty, lib, qm_upsample_problem
from . import stages as S

__all__ = ["UpsampleFunction", "FuseFunction", "BatchDepthFunction", "upsample", "fuse", "batch_depth"]


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


_PROBLEM = np.dtype(qm_upsample_problem)


def _rows(t, B):
    """Device addresses of the B leading-dimension slices of t (0 for None)."""
    if t is None:
        return 0
    return t.data_ptr() + np.arange(B, dtype=np.uint64) * np.uint64(t.stride(0) * t.element_size())


def _problems(d, conf, w_row, w_col, lams, row_seg, col_seg, **ptrs):
    """qm_upsample_problem array for a (B, H, W) batch (B = 1 for 2-D inputs),
    as a numpy record array laid out as the C struct; ptrs: further pointer
    fields from (B, ...) tensors (or None)."""
    B, H, W = d.shape
    a = np.zeros(B, _PROBLEM)
    a["W"], a["H"] = W, H
    a["row_seg"] = W if row_seg is None else int(row_seg)
    a["col_seg"] = H if col_seg is None else int(col_seg)
    a["lambda_"] = np.asarray(lams, np.float64)
    a["d"], a["conf"], a["w_row"], a["w_col"] = _rows(d, B), _rows(conf, B), _rows(w_row, B), _rows(w_col, B)
    for k, t in ptrs.items():
        a[k] = _rows(t, B)
    return a


class UpsampleFunction(torch.autograd.Function):
    """d (H, W) or (B, H, W); conf like d or None; w_row (.., H, W+1);
    w_col (.., H+1, W); lam a number, a 0-d tensor or a (B,) tensor (one
    smoothness weight per problem). A batch runs as one launch per pass."""

    @staticmethod
    def forward(ctx, d, conf, w_row, w_col, lam, row_seg, col_seg):
        batched = d.dim() == 3
        d_ = d.detach().contiguous()
        w_row_, w_col_ = w_row.detach().contiguous(), w_col.detach().contiguous()
        conf_ = None if conf is None else conf.detach().contiguous()
        if not batched:
            d_, w_row_, w_col_ = d_[None], w_row_[None], w_col_[None]
            conf_ = None if conf_ is None else conf_[None]
        B = d_.shape[0]
        if torch.is_tensor(lam):
            lv = lam.detach().double().cpu().reshape(-1)
            lams = [float(lv[0])] * B if lv.numel() == 1 else [float(v) for v in lv]
        else:
            lams = [float(lam)] * B
        if len(lams) != B:
            raise ValueError("lam must be a scalar or hold one value per problem")
        if ctx.needs_input_grad[4] and not all(v > 0.0 for v in lams):
            raise ValueError("a differentiable lam must be > 0")
        out = device_empty(d_.shape, d_.dtype, d_.device)
        # the row and column pass results, kept for the backward
        state = device_empty((B, 4) + tuple(d_.shape[1:]), torch.float32, d_.device)
        probs = _problems(d_, conf_, w_row_, w_col_, lams, row_seg, col_seg, dense=out, state=state)
        check(lib().qm_upsample_general_batch(probs.ctypes.data, B, _stream()))
        ctx.save_for_backward(d_, w_row_, w_col_, state, *([conf_] if conf_ is not None else []))
        ctx.has_conf = conf_ is not None
        ctx.lams = lams
        ctx.lam_shape = tuple(lam.shape) if torch.is_tensor(lam) else None
        ctx.batched = batched
        ctx.segs = (row_seg, col_seg)
        return out if batched else out[0]

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        d, w_row, w_col, state = saved[:4]
        conf = saved[4] if ctx.has_conf else None
        B = d.shape[0]
        g = (g if ctx.batched else g[None]).contiguous()
        gd = device_empty(d.shape, d.dtype, d.device) if ctx.needs_input_grad[0] else None
        gc = device_empty(d.shape, d.dtype, d.device) if (conf is not None and ctx.needs_input_grad[1]) else None
        gwr = device_empty(w_row.shape, w_row.dtype, w_row.device) if ctx.needs_input_grad[2] else None
        gwc = device_empty(w_col.shape, w_col.dtype, w_col.device) if ctx.needs_input_grad[3] else None
        want_lam = ctx.lam_shape is not None and ctx.needs_input_grad[4]
        # dL/dlambda = sum_e w_e dL/d(lambda w_e), summed on the device
        per = device_empty((B,), torch.float64, d.device) if want_lam else None
        probs = _problems(d, conf, w_row, w_col, ctx.lams, *ctx.segs, g=g, grad_d=gd, grad_conf=gc, grad_w_row=gwr,
                          grad_w_col=gwc, grad_lambda=per, state=state)
        check(lib().qm_upsample_backward_batch(probs.ctypes.data, B, _stream()))
        glam = None
        if want_lam:
            one_per_problem = len(ctx.lam_shape) == 1 and ctx.lam_shape[0] == B
            glam = (per if one_per_problem else per.sum()).to(torch.float32).reshape(ctx.lam_shape)
        unb = (lambda t: t) if ctx.batched else (lambda t: None if t is None else t[0])
        return unb(gd), unb(gc), unb(gwr), unb(gwc), glam, None, None


def upsample(d, w_row, w_col, conf=None, lam=10.0, row_seg=None, col_seg=None):
    return UpsampleFunction.apply(d, conf, w_row, w_col, lam, row_seg, col_seg)


class FuseFunction(torch.autograd.Function):
    """Inputs: dense depth (H, W) [differentiable], previous seeds (H, W, 4)
    (not modified: a copy is warped), cam (fx, fy, cx, cy), T_last_to_cur (3x4).
    Outputs: fused depth (H, W), new seeds (H, W, 4)."""

    @staticmethod
    def forward(ctx, depth_in, seeds_prev, cam, T_last_to_cur):
        d = depth_in.detach().contiguous()
        warped = seeds_prev.detach().clone().contiguous()
        dout, sout, table = S.fuse(warped, d, cam, T_last_to_cur, want_table=True)
        ctx.save_for_backward(d, warped, table)
        return dout, sout

    @staticmethod
    def backward(ctx, g_depth, g_seeds):
        d, warped, table = ctx.saved_tensors
        H, W = d.shape
        grad = device_empty(d.shape, d.dtype, d.device)
        gd = None if g_depth is None else g_depth.contiguous()
        gs = None if g_seeds is None else g_seeds.contiguous()
        if gd is None and gs is None:
            return None, None, None, None
        check(lib().qm_fuse_backward(_p(warped), _p(table), _p(d), W, H, _p(gd), _p(gs), _p(grad), _stream()))
        return grad, None, None, None


def fuse(depth_in, seeds_prev, cam, T_last_to_cur=None):
    T = np.eye(3, 4, dtype=np.float32) if T_last_to_cur is None else T_last_to_cur
    return FuseFunction.apply(depth_in, seeds_prev, cam, T)


class BatchDepthFunction(torch.autograd.Function):
    """lams: (n,) tensor of per-stream smoothness weights; batch: a Batch with
    the "depth" output. Runs the batch once (graph replay when graph=True)
    and returns the n fused depth maps (copies)."""

    @staticmethod
    def forward(ctx, lams, batch, graph=False):
        batch.set_lambdas([float(v) for v in lams.detach().double().cpu()])
        batch.run(graph=graph)
        ctx.batch = batch
        ctx.run_id = batch.runs()
        ctx.dtype, ctx.device = lams.dtype, lams.device
        return tuple(o["depth"].clone() for o in batch.outputs)

    @staticmethod
    def backward(ctx, *grads):
        # the backward differentiates the state the batch holds now: it must
        # still be this forward's run (the library also refuses after a
        # set_frame / reset_seeds / autotune that followed the run)
        if ctx.batch.runs() != ctx.run_id:
            raise RuntimeError("batch_depth: the batch ran again after this forward; its backward needs the state "
                               "of the run it differentiates (call backward before the next run)")
        g = ctx.batch.backward(list(grads))
        return torch.tensor(g, dtype=ctx.dtype, device=ctx.device), None, None


def batch_depth(lams, batch, graph=False):
    return BatchDepthFunction.apply(lams, batch, graph)
