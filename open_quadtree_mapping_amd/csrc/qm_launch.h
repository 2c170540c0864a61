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
// samples of the block's 64 cells), and cells (16-lane rows) per block of the
// BP sweep and of the extraction.
struct Tuning {
  int cost_waves = 4;    // 2, 4, 8
  int bp_rows = 16;      // 4, 8, 16
  int extract_rows = 16;  // 4, 8, 16
  int overlap = 0;       // 0 one stream; 1-7 two to eight frame groups on their own streams (run_pipeline)
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
void launch_bp_sweep(const FrameDev* dframes, const FrameDev* hframes, int n, int N, int l, bool A, bool first,
                     bool active, const BpLaunchParams& p, const Tuning& t, hipStream_t s, bool copy_all = true,
                     bool ext = false);
// colour < 0: every cell; 0 / 1: only the cells with x + y even / odd
void launch_extract(const FrameDev* dframes, const FrameDev* hframes, int n, int N, const BpLaunchParams& p,
                    const Tuning& t, hipStream_t s, int colour = -1);
void launch_copy_messages(const FrameDev* dframes, const FrameDev* hframes, int n, int N, float* out, hipStream_t s);
// a cost volume in sample order -> the row layout of the pipeline
void launch_cost_to_rows(const float* in, float* out, size_t ncell, int N, hipStream_t s);
// upsample: mode 0 = compose the sparse map from BP (+ pre-fill), weights from
// the keyframe; 1 = explicit sparse map, weights from the keyframe;
// 2 = generalised (d, conf, w_row, w_col given); lambda is FrameDev::lam
void launch_upsample(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                     bool clear_zbuf, hipStream_t s);
void launch_fuse(const FrameDev* dframes, const FrameDev* hframes, int n, bool clear_zbuf, hipStream_t s);
void launch_upsample_row(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         hipStream_t s);
void launch_upsample_col(const FrameDev* dframes, const FrameDev* hframes, int n, int mode, const float* wlut,
                         bool clear_zbuf, hipStream_t s);
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
// adjoints: recompute the generalised upsample forward, then the column and row
// adjoint solves; fusion update adjoint w.r.t. its dense depth input
void launch_upsample_backward(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
void launch_fuse_backward(const FrameDev* dframes, const FrameDev* hframes, int n, hipStream_t s);
// pipeline backward: the keyframe upsample's inputs written as explicit arrays
// into each descriptor's g_d / g_wrow / g_wcol
void launch_compose_inputs(const FrameDev* dframes, const FrameDev* hframes, int n, const float* wlut,
                           hipStream_t s);

}  // namespace qm
