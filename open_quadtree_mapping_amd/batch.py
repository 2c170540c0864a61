"""Batched keyframes: many camera streams of uneven resolution through every
stage in one launch per stage and launch shape (optionally one hipGraph replay
per run); the frames of one size share launch shapes (a resolution group).

Thin owner of a qm_batch (include/qm.h). Inputs are torch uint8 device tensors
kept alive by this object; per-stream fusion seeds persist across runs.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import warnings

import numpy as np
import torch

from ._lib import STAGES, TUNE_KNOBS, check, default_params, lib, qm_frame_desc, qm_params


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            db = json.load(f)
        return db if isinstance(db, dict) else {}
    except (OSError, ValueError):
        return {}


def _store_cache(path: str, db: dict):
    d = os.path.dirname(os.path.abspath(path))
    try:
        os.makedirs(d, exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError as e:
        warnings.warn(f"launch-shape cache {path} not written: {e}")

__all__ = ["Batch", "make_desc"]


def make_desc(key, refs, T_ref_key, cam, T_last_to_cur=None, prefill=None, depth_out=None, dense_out=None,
              sparse_out=None, lam=None, prior=None, prior_depth=None) -> qm_frame_desc:
    H, W = key.shape
    d = qm_frame_desc()
    d.W, d.H, d.pitch = W, H, W
    for i, v in enumerate(cam):
        d.cam[i] = float(v)
    d.key = key.data_ptr()
    for r in range(4):
        d.refs[r] = refs[r].data_ptr()
    T = np.ascontiguousarray(T_ref_key, dtype=np.float32).reshape(48)
    for i in range(48):
        d.T_ref_key[i] = float(T[i])
    Tl = np.eye(3, 4, dtype=np.float32) if T_last_to_cur is None else np.asarray(T_last_to_cur, np.float32)
    for i, v in enumerate(Tl.reshape(12)):
        d.T_last_to_cur[i] = float(v)
    d.prefill = None if prefill is None else prefill.data_ptr()
    d.depth_out = None if depth_out is None else depth_out.data_ptr()
    d.dense_out = None if dense_out is None else dense_out.data_ptr()
    d.sparse_out = None if sparse_out is None else sparse_out.data_ptr()
    d.lambda_ = 0.0 if lam is None else float(lam)
    d.prior = None if prior is None else prior.data_ptr()
    d.prior_depth = None if prior_depth is None else prior_depth.data_ptr()
    return d


class Batch:
    """frames: list of dicts with key (HxW uint8 device), refs (4 of them),
    T (4,3,4 float32), cam (fx, fy, cx, cy), optional T_last, prefill,
    lambda (this stream's smoothness weight; default the params' lambda) and
    the prior-driven alternates' inputs prior (HxWx2 float32) and prior_depth.
    Output tensors depth / dense / sparse are allocated per stream when asked."""

    BUF_LEVELS, BUF_COST, BUF_DCOARSE, BUF_LABEL, BUF_DENSE, BUF_SEEDS, BUF_TABLE, BUF_PRIOR = range(8)

    def __init__(self, frames, params: qm_params | None = None, outputs=("depth",)):
        self.params = params or default_params()
        self.frames = list(frames)
        self.n = len(self.frames)
        self.outputs = []
        descs = (qm_frame_desc * self.n)()
        self._keep = []
        for i, fr in enumerate(self.frames):
            H, W = fr["key"].shape
            dev = fr["key"].device
            outs = {k: torch.empty((H, W), dtype=torch.float32, device=dev) for k in outputs}
            self.outputs.append(outs)
            descs[i] = make_desc(fr["key"], fr["refs"], fr["T"], fr["cam"], fr.get("T_last"), fr.get("prefill"),
                                 outs.get("depth"), outs.get("dense"), outs.get("sparse"), fr.get("lambda"),
                                 fr.get("prior"), fr.get("prior_depth"))
            self._keep.append(fr)
        self._descs = descs
        self._h = C.c_void_p()
        check(lib().qm_batch_create(descs, self.n, C.byref(self.params), C.byref(self._h)))

    def set_frame(self, i, fr):
        self.frames[i] = fr
        o = self.outputs[i]
        d = make_desc(fr["key"], fr["refs"], fr["T"], fr["cam"], fr.get("T_last"), fr.get("prefill"),
                      o.get("depth"), o.get("dense"), o.get("sparse"), fr.get("lambda"), fr.get("prior"),
                      fr.get("prior_depth"))
        self._descs[i] = d
        check(lib().qm_batch_set_frame(self._h, i, C.byref(d)))

    def set_lambdas(self, lams):
        """Per-stream smoothness weights for the next run (each finite and > 0;
        the library reads a value <= 0 as "use the params' lambda")."""
        lams = [float(v) for v in lams]
        if len(lams) != self.n:
            raise ValueError(f"need {self.n} lambdas, got {len(lams)}")
        bad = [v for v in lams if not (np.isfinite(v) and v > 0.0)]
        if bad:
            raise ValueError(f"lambdas must be finite and > 0, got {bad}")
        for i, v in enumerate(lams):
            fr = dict(self.frames[i])
            fr["lambda"] = float(v)
            self.set_frame(i, fr)

    def backward(self, g_depth_out, stream=None) -> np.ndarray:
        """dL/dlambda per stream of the last run, given dL/d(depth_out) per
        stream (float32 device tensors of the stream's size, or None)."""
        s = stream if stream is not None else torch.cuda.current_stream()
        gs = [None if g is None else g.contiguous() for g in g_depth_out]
        arr = (C.c_void_p * self.n)(*[None if g is None else g.data_ptr() for g in gs])
        out = (C.c_double * self.n)()
        check(lib().qm_batch_backward(self._h, arr, out, C.c_void_p(s.cuda_stream)))
        return np.array(out[:], dtype=np.float64)

    def runs(self) -> int:
        """Number of runs of this batch so far (qm_batch_runs)."""
        r = C.c_longlong()
        check(lib().qm_batch_runs(self._h, C.byref(r)))
        return r.value

    def run(self, graph: bool = False, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream()
        check(lib().qm_batch_run(self._h, 1 if graph else 0, C.c_void_p(s.cuda_stream)))

    def run_timed(self, stream=None):
        """One plain run with an event pair around every launch.
        Returns {stage: (total_ms, launches)}."""
        s = stream if stream is not None else torch.cuda.current_stream()
        ms = (C.c_double * len(STAGES))()
        cnt = (C.c_int * len(STAGES))()
        check(lib().qm_batch_run_timed(self._h, C.c_void_p(s.cuda_stream), ms, cnt))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(STAGES)}

    def buffer_ptr(self, i, which) -> int:
        p = C.c_void_p()
        check(lib().qm_batch_buffer(self._h, i, which, C.byref(p)))
        return p.value

    def autotune(self, reps: int = 5, stream=None, cache: str | None = None, refresh: bool = False,
                 seeds: tuple = ()) -> dict:
        """Pick the fastest launch shapes for this batch (qm_batch_autotune):
        per resolution group, then the overlap knob. Returns group 0's
        {knob: (value, ms)} (group_tuning gives the others').
        cache: a JSON file of earlier choices, keyed by the device, each
        group's size, frame count, hypotheses, cost arithmetic and sweep counts
        (shape knobs), and the batch's whole composition and upsample solver
        (overlap knob). When
        it holds every group's and the batch's entries they are applied without
        timing (ms None); else (or with refresh) the batch is tuned and its
        choices are written there. seeds: further read-only JSON files of the
        same form, consulted after cache (earlier measured choices that ship
        with an application)."""
        keys = self._cache_keys() if (cache or seeds) else None
        if keys is not None and not refresh:
            db = {}
            for path in reversed([cache, *seeds] if cache else list(seeds)):
                db.update(_load_cache(path))
            if all(k in db for k in keys[0]) and keys[1] in db:
                for g, k in enumerate(keys[0]):
                    self.set_group_tuning(g, **db[k])
                self.set_tuning(overlap=db[keys[1]]["overlap"])
                return {k: (v, None) for k, v in self.tuning().items()}
        s = stream if stream is not None else torch.cuda.current_stream()
        chosen = (C.c_int * len(TUNE_KNOBS))()
        ms = (C.c_double * len(TUNE_KNOBS))()
        check(lib().qm_batch_autotune(self._h, int(reps), C.c_void_p(s.cuda_stream), chosen, ms))
        if keys is not None and cache:
            db = _load_cache(cache)
            for g, k in enumerate(keys[0]):
                db[k] = self.group_tuning(g)
            db[keys[1]] = {"overlap": self.tuning()["overlap"]}
            _store_cache(cache, db)
        return {k: (chosen[i], ms[i]) for i, k in enumerate(TUNE_KNOBS)}

    def _cache_keys(self):
        # the architecture and CU count (the device name can be empty, e.g. under a profiler)
        props = torch.cuda.get_device_properties(torch.cuda.current_device())
        dev = f"{getattr(props, 'gcnArchName', '') or props.name}/{props.multi_processor_count}CU"
        p = self.params
        common = f"N{p.n_hyp}|arith{p.cost_arith}|iters{','.join(str(v) for v in p.bp_iters)}"
        sizes = [tuple(f["key"].shape[::-1]) for f in self.frames]
        groups = self.groups()
        per_group = [f"{dev}|{W}x{H}|frames{sizes.count((W, H))}|{common}" for (W, H) in groups]
        # the overlap knob also depends on the upsample's line solver (the scan
        # solver's passes are not latency-bound); Thomas batches keep their keys
        solver = f"|solver{p.upsample_solver}" if p.upsample_solver else ""
        whole = f"{dev}|batch:" + ";".join(f"{W}x{H}" for (W, H) in sizes) + f"|{common}{solver}"
        return per_group, whole

    def set_tuning(self, **knobs):
        """Set knobs for the batch (shape knobs for every resolution group)."""
        for k, v in knobs.items():
            check(lib().qm_batch_set_tuning(self._h, TUNE_KNOBS.index(k), int(v)))

    def tuning(self) -> dict:
        """The overlap knob and group 0's shape knobs."""
        out = {}
        for i, k in enumerate(TUNE_KNOBS):
            v = C.c_int()
            check(lib().qm_batch_get_tuning(self._h, i, C.byref(v)))
            out[k] = v.value
        return out

    def groups(self) -> list:
        """Resolution groups [(W, H), ...] in the order of their first frame."""
        cnt = C.c_int()
        check(lib().qm_batch_groups(self._h, C.byref(cnt), None, None, 0))
        W, H = (C.c_int * cnt.value)(), (C.c_int * cnt.value)()
        check(lib().qm_batch_groups(self._h, C.byref(cnt), W, H, cnt.value))
        return [(W[g], H[g]) for g in range(cnt.value)]

    def group_tuning(self, g: int) -> dict:
        """Shape knobs of resolution group g."""
        out = {}
        for i, k in enumerate(TUNE_KNOBS):
            if k == "overlap":
                continue
            v = C.c_int()
            check(lib().qm_batch_get_group_tuning(self._h, int(g), i, C.byref(v)))
            out[k] = v.value
        return out

    def set_group_tuning(self, g: int, **knobs):
        for k, v in knobs.items():
            check(lib().qm_batch_set_group_tuning(self._h, int(g), TUNE_KNOBS.index(k), int(v)))

    def read_buffer(self, i, which) -> np.ndarray:
        """Host copy of an internal buffer of stream i (BUF_* ids)."""
        H, W = self.frames[i]["key"].shape
        N, Hc, Wc = self.params.n_hyp, H // 4, W // 4
        shapes = {self.BUF_LEVELS: ((H, W), np.uint8), self.BUF_COST: ((Hc, Wc, N), np.float32),
                  self.BUF_DCOARSE: ((Hc, Wc), np.float32), self.BUF_LABEL: ((Hc, Wc), np.int32),
                  self.BUF_DENSE: ((H, W), np.float32), self.BUF_SEEDS: ((H, W, 4), np.float32),
                  self.BUF_TABLE: ((H, W), np.int32), self.BUF_PRIOR: ((H, W, 4), np.float32)}
        shape, dt = shapes[which]
        out = np.empty(shape, dt)
        check(lib().qm_batch_read_buffer(self._h, i, which, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def reset_seeds(self):
        check(lib().qm_batch_reset_seeds(self._h))

    def close(self):
        if self._h:
            lib().qm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
