"""Differentiable upsample and fusion: torch.autograd.Functions over the C ABI.

UpsampleFunction: out = x_col / n_col of the generalised upsample
  (qm_upsample_general); backward = qm_upsample_backward, which recomputes the
  forward sweep and runs the transposed (= same, A is symmetric) tridiagonal
  solves. Gradients w.r.t. sparse depths, confidences, per-pixel smoothness
  weights, and the smoothness scale lambda when it is given as a tensor
  (dL/dlambda = sum_e w_e dL/dw_e / lambda, since A = I + lambda L(w)).
FuseFunction: the Gaussian x Beta update of one keyframe (qm_fuse) with
  backward qm_fuse_backward w.r.t. the dense depth input (prior seeds fixed).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib
from . import stages as S

__all__ = ["UpsampleFunction", "FuseFunction", "upsample", "fuse"]


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class UpsampleFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, conf, w_row, w_col, lam, row_seg, col_seg):
        lam_f = float(lam.detach().item()) if torch.is_tensor(lam) else float(lam)
        d_, w_row_, w_col_ = d.detach().contiguous(), w_row.detach().contiguous(), w_col.detach().contiguous()
        conf_ = None if conf is None else conf.detach().contiguous()
        out = S.upsample_general(d_, w_row_, w_col_, conf=conf_, lam=lam_f, row_seg=row_seg, col_seg=col_seg)
        ctx.save_for_backward(d_, w_row_, w_col_, *( [conf_] if conf_ is not None else []))
        ctx.has_conf = conf_ is not None
        ctx.lam = lam_f
        ctx.lam_is_tensor = torch.is_tensor(lam)
        ctx.segs = (row_seg, col_seg)
        return out

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        d, w_row, w_col = saved[:3]
        conf = saved[3] if ctx.has_conf else None
        H, W = d.shape
        g = g.contiguous()
        gd = torch.empty_like(d) if ctx.needs_input_grad[0] else None
        gc = torch.empty_like(d) if (conf is not None and ctx.needs_input_grad[1]) else None
        need_w = ctx.needs_input_grad[2] or ctx.needs_input_grad[3] or ctx.needs_input_grad[4]
        gwr = torch.empty_like(w_row) if need_w else None
        gwc = torch.empty_like(w_col) if need_w else None
        rs, cs = ctx.segs
        check(lib().qm_upsample_backward(_p(d), _p(conf), _p(w_row), _p(w_col), W, H, ctx.lam,
                                         W if rs is None else int(rs), H if cs is None else int(cs), _p(g),
                                         _p(gd), _p(gc), _p(gwr), _p(gwc), _stream()))
        glam = None
        if ctx.lam_is_tensor and ctx.needs_input_grad[4]:
            glam = ((w_row.double() * gwr.double()).sum() + (w_col.double() * gwc.double()).sum()) / ctx.lam
            glam = glam.to(torch.float32)
        return (gd, gc, gwr if ctx.needs_input_grad[2] else None, gwc if ctx.needs_input_grad[3] else None, glam,
                None, None)


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
        grad = torch.empty_like(d)
        gd = None if g_depth is None else g_depth.contiguous()
        gs = None if g_seeds is None else g_seeds.contiguous()
        if gd is None and gs is None:
            return None, None, None, None
        check(lib().qm_fuse_backward(_p(warped), _p(table), _p(d), W, H, _p(gd), _p(gs), _p(grad), _stream()))
        return grad, None, None, None


def fuse(depth_in, seeds_prev, cam, T_last_to_cur=None):
    T = np.eye(3, 4, dtype=np.float32) if T_last_to_cur is None else T_last_to_cur
    return FuseFunction.apply(depth_in, seeds_prev, cam, T)
