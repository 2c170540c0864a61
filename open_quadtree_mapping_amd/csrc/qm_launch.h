// qm_launch.h -- host-side launchers of the pipeline stages. Each launcher runs
// its stage for n frames in one launch; hframes mirrors the device array.
#pragma once

#include <hip/hip_runtime.h>
#include "qm_device.h"
#include "../../include/qm.h"

namespace qm {

// Fill bm from per-frame block counts; returns the total number of blocks.
inline int make_block_map(BlockMap& bm, const int* counts, int n) {
  bm.n = n;
  int acc = 0;
  for (int i = 0; i < n; ++i) {
    bm.start[i] = acc;
    acc += counts[i];
  }
  for (int i = n; i <= kMaxFrames; ++i) bm.start[i] = acc;
  return acc;
}

struct BpLaunchParams {
  float p1, p2;
  double idepth_min, idepth_step;
};

// Launch shapes the autotuner chooses among (every choice gives identical
// results): waves per cost-volume block (each wave takes N / cost_waves
// samples of the block's 64 cells), cells (16-lane rows) per block of the
// single BP half-sweep and of the extraction, and the fused half-sweep pairs:
// off (0) or A cells per strip row (8 or 16: strips of 14 or 30 columns), rows
// per strip segment, and how many of the finest levels use them.
struct Tuning {
  int cost_waves = 4;    // 2, 4, 8
  int bp_rows = 16;      // 4, 8, 16
  int extract_rows = 16;  // 4, 8, 16
  int overlap = 0;       // 0 one stream; 1-7 two to eight frame groups on their own streams (run_pipeline)
  int bp_pair = 8;       // 0, 8, 16
  int bp_seg = 8;        // 8, 16, 32, 64
  int bp_pair_levels = 2;  // 1-4: levels 0 .. bp_pair_levels - 1
};

void launch_quadtree(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
// natural = write the volume in sample order (stage API) instead of the row layout
// arith = QM_COST_REFERENCE or QM_COST_FUSED (the projection arithmetic)
// pyramid: the cost kernel also writes levels 1-3 (no stand-alone pyramid needed)
void launch_cost(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const float* table, bool natural,
                 int arith, const Tuning& t, hipStream_t s, bool pyramid = false);
void launch_pyramid(const FrameDev* dframes, const FrameDev* hframes, int n, int N, hipStream_t s);
// one half-sweep of level l; first = the level's first sweep (also initialises
// cells' messages from the parent level, or zero at the coarsest level: every
// cell it does not update when copy_all, else only the cells the quadtree
// skips, which suffices when the level has a second half-sweep);
// active = 0 turns the first sweep into a pure initialisation.
// ext: this level-0 (non-first) sweep also extracts the winners of its colour's
// cells (every one of them, updated or skipped), for the sweep that ends BP.
// buf: the level's message buffer (read and written in place); pbuf: the
// parent level's (read by a first sweep).
void launch_bp_sweep(const FrameDev* dframes, const FrameDev* hframes, int n, int N, int l, bool A, bool first,
                     bool active, const BpLaunchParams& p, const Tuning& t, hipStream_t s, bool copy_all = true,
                     bool ext = false, int buf = 0, int pbuf = 0);
// Every half-sweep of the hierarchical BP, coarse to fine (iters[l] half-sweeps
// at level l; a level without any only initialises its messages), as single
// half-sweeps or fused pairs (t.bp_pair). ext_last: the last level-0 sweep
// extracts its colour's winners (the extraction stage then does the other
// colour's); want_msgs: level 0's final messages of both colours are read
// afterwards (copied out).
void launch_bp_levels(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const int iters[4],
                      const BpLaunchParams& p, const Tuning& t, hipStream_t s, bool ext_last, bool want_msgs);
// level l's half-sweeps run as fused pairs; the message buffer holding its
// final messages after its iters half-sweeps
bool bp_level_pairs(int l, int iters, const Tuning& t);
int bp_final_buf(int l, int iters, const Tuning& t);
// colour < 0: every cell; 0 / 1: only the cells with x + y even / odd; buf: level 0's final message buffer
void launch_extract(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const BpLaunchParams& p,
                    const Tuning& t, hipStream_t s, int colour = -1, int buf = 0);
void launch_copy_messages(const FrameDev* dframes, const FrameDev* hframes, int n, int N, float* out, hipStream_t s,
                          int buf = 0);
// a cost volume in sample order -> the row layout of the pipeline
void launch_cost_to_rows(const float* in, float* out, size_t ncell, int N, hipStream_t s);
// upsample: mode 0 = compose the sparse map from BP (+ pre-fill), weights from
// the keyframe; 1 = explicit sparse map, weights from the keyframe;
// 2 = generalised (d, conf, w_row, w_col given); lambda is FrameDev::lam.
// solver QM_UPSAMPLE_SCAN runs modes 0 and 1 by segmented scans (k_scan.hip)
// for the frames whose lines fit it; everything else runs the Thomas kernels.
void launch_upsample(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                     bool clear_zbuf, hipStream_t s, int solver = QM_UPSAMPLE_THOMAS);
void launch_fuse(const FrameDev* dframes, const FrameDev* hframes, int n, bool clear_zbuf, hipStream_t s);
void launch_upsample_row(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         hipStream_t s, int solver = QM_UPSAMPLE_THOMAS, bool clear_zbuf = false);
void launch_upsample_col(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         bool clear_zbuf, hipStream_t s, int solver = QM_UPSAMPLE_THOMAS);
// the scan kernels for the frames whose lines fit them (rows of up to 1536
// pixels, columns likewise); rest[i] = 1 marks a frame left to the Thomas
// kernels; returns how many there are
int launch_upsample_row_scan(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                             hipStream_t s, int* rest, bool clear_zbuf = false);
// true when the row scan takes every frame of the launch (then a row pass asked
// to reset the fusion z-buffer does it, and the column pass need not)
bool row_scan_covers(const FrameDev* hframes, int n);
int launch_upsample_col_scan(const FrameDev* dframes, const FrameDev* hframes, int n, const float* wlut,
                             bool clear_zbuf, hipStream_t s, int* rest);
// the reference's prior-driven alternates; frames without the prior get no work
void launch_quadtree_prior(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
void launch_prior_cost(const FrameDev* dframes, const FrameDev* hframes, int n, int N, bool natural,
                       double idepth_min, double idepth_step, hipStream_t s);
// the reference's 7x7 bilateral alternate to the global upsample
void launch_upsample_local(const float* sparse, const uint8_t* img, int W, int H, int pitch, float* dense,
                           hipStream_t s);
void launch_fill_zbuf(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
void launch_fuse_warp(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
void launch_fuse_update(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
void launch_fuse_replay(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
// adjoints: the generalised upsample forward recomputed (recompute; else its
// rowx planes and colxn are the forward's), then the column and row adjoint
// solves; fusion update adjoint w.r.t. its dense depth input
void launch_upsample_backward(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s,
                              bool recompute = true);
void launch_fuse_backward(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
// dL/dd and dL/dconf of frames with grad_d / grad_conf from their adj_rowv planes
void launch_adj_dconf(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
// pipeline backward: the keyframe upsample's inputs written as explicit arrays
// into each descriptor's g_d (when want_sparse) / g_wrow / g_wcol
void launch_compose_inputs(const FrameDev* dframes, const FrameDev* hframes, int n, const float* wlut,
                           hipStream_t s, bool want_sparse = true);

}  // namespace qm
