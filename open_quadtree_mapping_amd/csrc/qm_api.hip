// qm_api.hip -- C ABI of libqm: stage-level entry points and batched keyframes.
//
// Memory: a batch owns one device arena holding every per-stream buffer
// (quadtree levels, cost pyramid, messages of all four BP levels, Thomas
// coefficients, row results, dense map, fusion seeds (double-buffered) and the
// fusion z-buffer); nothing is allocated per keyframe. Stage-level calls are
// for integration and tests: they allocate their scratch, run on the given
// stream and synchronise it before returning.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/qm.h"
#include "qm_device.h"
#include "qm_launch.h"
#include "qm_internal.h"

using namespace qm;

namespace qm {

thread_local std::string g_err;
std::atomic<size_t> g_bytes{0};
std::atomic<size_t> g_cap{0};
std::mutex g_mu;
std::unordered_map<void*, size_t> g_sizes;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

void dev_free(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sizes.find(p);
    if (it != g_sizes.end()) {
      g_bytes -= it->second;
      g_sizes.erase(it);
    }
  }
  (void)hipFree(p);
}

// Scratch blocks of the stage-level calls, kept for the next call: a hipMalloc
// and hipFree pair costs far more than the kernels of one such call. A block
// enters the cache only after its call has synchronised its stream, so no
// kernel can still use it; the cache holds at most kCacheMax bytes, and it is
// emptied when an allocation is refused.
struct CachedBlock {
  int dev;
  size_t bytes;
  void* p;
};
std::mutex g_cache_mu;
std::vector<CachedBlock> g_cache;
size_t g_cache_bytes = 0;
constexpr size_t kCacheMax = size_t(1) << 30;  // 1 GiB over all devices

// frees each block with its own device current
void free_blocks(const std::vector<CachedBlock>& blocks) {
  if (blocks.empty()) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (const CachedBlock& b : blocks) {
    if (b.dev != cur) (void)hipSetDevice(b.dev);
    dev_free(b.p);
    if (b.dev != cur) (void)hipSetDevice(cur);
  }
}

void flush_scratch_cache() {
  std::vector<CachedBlock> blocks;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    blocks.swap(g_cache);
    g_cache_bytes = 0;
  }
  free_blocks(blocks);
}

static int dev_alloc_once(void** p, size_t bytes) {
  const size_t cap = g_cap.load();
  if (cap && g_bytes.load() + bytes > cap)
    return fail(QM_ERR_ALLOC, "device allocation of " + std::to_string(bytes) + " bytes refused by the cap");
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(QM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_sizes[*p] = bytes;
  g_bytes += bytes;
  return QM_OK;
}

int dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  int rc = dev_alloc_once(p, bytes);
  if (rc == QM_OK) return rc;
  bool had_cache;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    had_cache = !g_cache.empty();
  }
  if (!had_cache) return rc;
  (void)hipGetLastError();  // a failed hipMalloc leaves its error for the next check
  flush_scratch_cache();
  return dev_alloc_once(p, bytes);
}

// the smallest cached block of the current device holding bytes, at most
// twice that (else a new allocation)
int scratch_take(void** p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    int best = -1;
    for (int i = 0; i < (int)g_cache.size(); ++i) {
      const CachedBlock& b = g_cache[i];
      if (b.dev == dev && b.bytes >= bytes && b.bytes <= 2 * bytes && (best < 0 || b.bytes < g_cache[best].bytes))
        best = i;
    }
    if (best >= 0) {
      *p = g_cache[best].p;
      g_cache_bytes -= g_cache[best].bytes;
      g_cache[best] = g_cache.back();
      g_cache.pop_back();
      return QM_OK;
    }
  }
  return dev_alloc(p, bytes);
}

// keeps a block for reuse; when the cache would exceed its cap the oldest
// blocks are freed first (a block larger than the cap is freed at once)
void scratch_return(void* p) {
  size_t bytes = 0;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sizes.find(p);
    if (it != g_sizes.end()) bytes = it->second;
  }
  if (!bytes || bytes > kCacheMax) {
    dev_free(p);
    return;
  }
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::vector<CachedBlock> evicted;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    size_t k = 0;
    while (k < g_cache.size() && g_cache_bytes + bytes > kCacheMax) g_cache_bytes -= g_cache[k++].bytes;
    evicted.assign(g_cache.begin(), g_cache.begin() + k);
    g_cache.erase(g_cache.begin(), g_cache.begin() + k);
    g_cache.push_back({dev, bytes, p});
    g_cache_bytes += bytes;
  }
  free_blocks(evicted);
}

void init_geometry(FrameDev& F, int W, int H, int pitch) {
  std::memset(&F, 0, sizeof(F));
  F.W = W;
  F.H = H;
  F.pitch = pitch;
  F.Wc = W / 4;
  F.Hc = H / 4;
  F.lw[0] = F.Wc;
  F.lh[0] = F.Hc;
  for (int l = 1; l < 4; ++l) {
    F.lw[l] = (F.lw[l - 1] + 1) / 2;
    F.lh[l] = (F.lh[l - 1] + 1) / 2;
  }
  F.row_seg = 16 * (W / 16);
  F.col_seg = 8 * (H / 8);
  F.edge_r = (W % 16) ? 1 : 0;
  F.edge_c = (H % 16) ? 1 : 0;
  F.lam = 10.0;
}

}  // namespace qm

namespace {

// RAII scratch list for stage-level calls. done(): the call has synchronised
// its stream, so its blocks go back to the cache; a call that returns before
// that (an error) frees them (hipFree waits for the device).
struct Scratch {
  std::vector<void*> ptrs;
  bool synced = false;
  int get(void** p, size_t bytes) {
    int rc = scratch_take(p, bytes);
    if (rc == QM_OK) ptrs.push_back(*p);
    return rc;
  }
  void done() { synced = true; }
  ~Scratch() {
    for (void* p : ptrs) {
      if (synced)
        scratch_return(p);
      else
        dev_free(p);
    }
  }
};

int check_params(const qm_params* p) {
  if (!p) return fail(QM_ERR_INVALID, "params is NULL");
  if (p->n_hyp != 64 && p->n_hyp != 128) return fail(QM_ERR_INVALID, "n_hyp must be 64 or 128");
  if (p->cost_arith != QM_COST_REFERENCE && p->cost_arith != QM_COST_FUSED)
    return fail(QM_ERR_INVALID, "cost_arith must be QM_COST_REFERENCE or QM_COST_FUSED");
  if (p->upsample_solver != QM_UPSAMPLE_THOMAS && p->upsample_solver != QM_UPSAMPLE_SCAN)
    return fail(QM_ERR_INVALID, "upsample_solver must be QM_UPSAMPLE_THOMAS or QM_UPSAMPLE_SCAN");
  if (!(p->idepth_max > p->idepth_min) || !(p->idepth_min > 0.0))
    return fail(QM_ERR_INVALID, "need 0 < idepth_min < idepth_max");
  for (int i = 0; i < 4; ++i)
    if (p->bp_iters[i] < 0 || p->bp_iters[i] > 1000) return fail(QM_ERR_INVALID, "bp_iters out of range");
  if (!(p->lambda >= 0.0)) return fail(QM_ERR_INVALID, "lambda must be >= 0");
  return QM_OK;
}

int check_size(int W, int H, int pitch) {
  if (W < 16 || H < 16 || W > 16384 || H > 16384) return fail(QM_ERR_INVALID, "image size out of range [16, 16384]");
  if (pitch < W) return fail(QM_ERR_INVALID, "pitch < width");
  // texel offsets y * pitch + x are formed in 32 bits (24-bit multiply)
  if (pitch > 262144) return fail(QM_ERR_INVALID, "pitch above 262144 bytes");
  return QM_OK;
}

void host_depth_table(int N, double imin, double imax, float* out) {
  const double step = (imax - imin) / (double)(N - 1);
  for (int d = 0; d < N; ++d) out[d] = (float)(1.0 / (step * d + imin));
}

// w(dI) = expf(-(dI)^2 / 400) of src/depth_upsample.cu:72, with the argument
// rounded to float as the reference does (the division is in double).
void host_weight_lut(float* out) {
  for (int d = 0; d < 256; ++d) {
    const float arg = (float)(-(double)(d * d) / 400.0);
    out[d] = (float)std::exp((double)arg);
  }
}

BpLaunchParams bp_params(const qm_params* p) {
  BpLaunchParams b;
  b.p1 = p->p1;
  b.p2 = p->p2;
  b.idepth_min = p->idepth_min;
  b.idepth_step = (p->idepth_max - p->idepth_min) / (double)(p->n_hyp - 1);
  return b;
}

// The coarse-to-fine half-sweeps of the hierarchical BP (every cell's winner
// extracted afterwards).
void run_sweeps(const FrameDev* df, const FrameDev* hf, int n, const qm_params* p, const Tuning& t, hipStream_t s,
                bool want_msgs = false) {
  launch_bp_levels(df, hf, n, p->n_hyp, p->bp_iters, bp_params(p), t, s, false, want_msgs);
}

// launches of the BP half-sweep stage
int bp_launches(const int iters[4], const Tuning& t) {
  int k = 0;
  for (int l = 0; l < 4; ++l) {
    const int it = iters[l];
    k += it == 0 ? 1 : !bp_level_pairs(l, it, t) ? it : it / 2 + it % 2;
  }
  return k;
}

// The whole hierarchical BP: pyramid, coarse-to-fine half-sweeps, extraction;
// want_msgs: level 0's final messages are copied out afterwards (returns their buffer).
int run_bp(const FrameDev* df, const FrameDev* hf, int n, const qm_params* p, hipStream_t s, bool want_msgs) {
  const Tuning t;
  launch_pyramid(df, hf, n, p->n_hyp, s);
  run_sweeps(df, hf, n, p, t, s, want_msgs);
  const int buf = bp_final_buf(0, p->bp_iters[0], t);
  launch_extract(df, hf, n, p->n_hyp, bp_params(p), t, s, -1, buf);
  return buf;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// bytes of the per-frame pipeline state for N hypotheses
size_t frame_bytes(const FrameDev& F, int N) {
  const size_t px = (size_t)F.W * F.H;
  size_t b = align256(px);  // levels
  for (int l = 0; l < 4; ++l) {
    const size_t c = (size_t)F.lw[l] * F.lh[l];
    b += align256(c * N * 4) + 2 * align256(c * 4 * N * 4);  // cost, two message buffers
  }
  const size_t cc = (size_t)F.Wc * F.Hc;
  b += 2 * align256(cc * 4);               // dcoarse, label
  b += align256(px * 16);                  // fwd
  b += align256(px * 8);                   // rowx (X and N planes)
  b += align256(px * 4);                   // dense
  b += 2 * align256(px * 16);              // seeds x2
  b += 2 * align256(px * 8);               // z-buffers, one per run parity
  b += align256(px * 4);                   // hole-filled warp table
  return b;
}

struct Carver {
  char* p;
  template <typename T> T* take(size_t bytes) {
    T* r = reinterpret_cast<T*>(p);
    p += align256(bytes);
    return r;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
struct qm_batch {
  qm_params p;
  int n = 0;
  std::vector<FrameDev> hf[2];
  FrameDev* df[2] = {nullptr, nullptr};
  void* arena = nullptr;
  float* table = nullptr;  // depth samples, then the weight LUT (256)
  float* wlut = nullptr;
  std::vector<float4*> seeds[2];
  // the fusion z-buffers: a run of parity p scatters into zbuf[p] and its
  // column pass resets zbuf[p ^ 1] for the next run, so that the warp may
  // run anywhere before the update
  std::vector<unsigned long long*> zbuf[2];
  // the hole-filled warp table of each frame. A run writes it, and the warp
  // writes the warped prior seeds back, only once a backward has been asked
  // for (keep_fusion); before that the first backward, or a read of those
  // buffers, replays them from the run's prior seeds and z-buffer (once per
  // run: replayed_run)
  std::vector<int32_t*> table_buf;
  bool keep_fusion = false;
  long long replayed_run = -1;
  FrameDev* dfr = nullptr;  // descriptors of the replay
  // graph mode: the run's linear segments (see plan_segments), each captured
  // as a graph of its own, per parity; captured is set once both exist
  struct Segment {
    int f0, nf;               // frames [f0, f0 + nf)
    std::vector<int> stages;  // QM_STAGE_* in order
    int stream;               // 0: the caller's stream, k: side stream k - 1
    std::vector<int> waits;   // events (event ids) waited for before the segment
    int record;               // event id recorded after the segment, or -1
  };
  std::vector<Segment> segs;
  std::vector<hipGraphExec_t> seg_exec[2];
  bool captured[2] = {false, false};
  int parity = 0;
  long long runs = 0;
  // the state a backward differentiates (the last run's descriptors, warp
  // tables and warped prior seeds) is intact: set by a run, cleared by
  // set_frame, reset_seeds and autotune
  bool run_state_valid = false;
  // launch shapes: the frames of one size form a resolution group with shapes
  // of its own (gtune; a stage's launch covers consecutive frames whose groups
  // agree on its shape); tune holds the batch-wide overlap knob
  Tuning tune;
  std::vector<int> group_of;  // frame -> group
  std::vector<int> group_w, group_h;
  std::vector<Tuning> gtune;
  // extra streams and fork / mark / join events of the frame-group overlap
  static constexpr int kMaxGroups = 8;
  hipStream_t side[kMaxGroups - 1] = {};
  hipEvent_t ev_fork = nullptr, ev_warp = nullptr, ev_mark[kMaxGroups - 1] = {}, ev_join[kMaxGroups - 1] = {};
  bool side_ready = false;
  std::vector<hipEvent_t> events;
  // backward (per-stream lambda), allocated on the first qm_batch_backward
  void* bwd_arena = nullptr;
  FrameDev* dfb = nullptr;
  double* partial = nullptr;
  int partial_n = 0;
  struct Bwd {
    float *sparse, *wrow, *wcol, *gdense;
    float2 *colxn, *adjxn;
    double* lam_part;  // H + W per-line partials of dL/dlambda
  };
  std::vector<Bwd> bwd;
  // once the backward buffers exist, every run's column pass also keeps its
  // results (colxn) there: a backward of a run after run number colxn_since
  // reuses them and the run's row planes instead of recomputing the forward
  long long colxn_since = -1;
};

namespace {

// Optional timing: an event pair around each stage's launches (the BP
// half-sweeps of a run back to back under one pair, so that event overhead
// does not inflate their per-launch figure), attributed to a stage kind
// (QM_STAGE_*) with its launch count.
struct Timer {
  std::vector<hipEvent_t>* pool;
  std::vector<int> stage, launches;
  size_t used = 0;
  hipStream_t s;
  hipEvent_t next() {
    if (used == pool->size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      pool->push_back(e);
    }
    return (*pool)[used++];
  }
  void begin(int st, int n) {
    stage.push_back(st);
    launches.push_back(n);
    (void)hipEventRecord(next(), s);
  }
  void end() { (void)hipEventRecord(next(), s); }
};

struct Mark {
  Timer* t;
  Mark(Timer* tt, int st, int n = 1) : t(tt) {
    if (t) t->begin(st, n);
  }
  ~Mark() {
    if (t) t->end();
  }
};

bool any_prior(const qm_batch* b) {
  for (const FrameDev& F : b->hf[0])
    if (F.prior) return true;
  return false;
}

bool any_prior_depth(const qm_batch* b) {
  for (const FrameDev& F : b->hf[0])
    if (F.prior_depth) return true;
  return false;
}

// the shapes a stage's launch depends on
bool same_cost_shape(const Tuning& a, const Tuning& c) { return a.cost_waves == c.cost_waves; }
bool same_bp_shape(const Tuning& a, const Tuning& c) {
  return a.bp_rows == c.bp_rows && a.bp_pair == c.bp_pair && a.bp_seg == c.bp_seg &&
         a.bp_pair_levels == c.bp_pair_levels;
}
bool same_extract_shape(const Tuning& a, const Tuning& c) {
  return a.extract_rows == c.extract_rows && (a.bp_pair != 0) == (c.bp_pair != 0);
}

// fn(a, k, tuning) for each run [f0 + a, f0 + a + k) of consecutive frames
// whose resolution groups share the stage's shape (same) in the group shapes gt
template <typename Same, typename Fn>
void for_shape_runs(const std::vector<Tuning>& gt, const std::vector<int>& group_of, int f0, int n, Same same,
                    Fn fn) {
  for (int a = 0; a < n;) {
    const Tuning& ta = gt[group_of[f0 + a]];
    int k = 1;
    while (a + k < n && same(ta, gt[group_of[f0 + a + k]])) ++k;
    fn(a, k, ta);
    a += k;
  }
}
template <typename Same, typename Fn>
void for_shape_runs(const qm_batch* b, int f0, int n, Same same, Fn fn) {
  for_shape_runs(b->gtune, b->group_of, f0, n, same, fn);
}

// One stage (QM_STAGE_* numbering: 0 quadtree, 1 cost volume, 2 pyramid,
// 3 BP half-sweeps, 4 extraction, 5 / 6 upsample rows / columns, 7 / 8 fusion
// warp / update) for frames [f0, f0 + n) on stream s.
void enqueue_stage(qm_batch* b, int par, hipStream_t s, Timer* t, int f0, int n, int stage) {
  const FrameDev* df = b->df[par] + f0;
  const FrameDev* hf = b->hf[par].data() + f0;
  const int N = b->p.n_hyp;
  const qm_params* p = &b->p;
  const BpLaunchParams bp = bp_params(p);
  switch (stage) {
    case QM_STAGE_QUADTREE: {
      Mark m(t, QM_STAGE_QUADTREE);
      launch_quadtree(df, hf, n, s);
      bool prior_depth = false;
      for (int i = 0; i < n; ++i) prior_depth |= hf[i].prior_depth != nullptr;
      if (prior_depth) launch_quadtree_prior(df, hf, n, s);
      break;
    }
    case QM_STAGE_COST: {
      int runs = 0;
      for_shape_runs(b, f0, n, same_cost_shape, [&](int, int, const Tuning&) { ++runs; });
      Mark m(t, QM_STAGE_COST, runs);
      bool prior = false;
      for (int i = 0; i < n; ++i) prior |= hf[i].prior != nullptr;
      // without a prior the cost kernel builds the pyramid from its tiles;
      // the prior cost changes level 0 first, then the pyramid stage runs
      for_shape_runs(b, f0, n, same_cost_shape, [&](int a, int k, const Tuning& tu) {
        launch_cost(df + a, hf + a, k, N, b->table, false, p->cost_arith, tu, s, !prior);
      });
      if (prior) launch_prior_cost(df, hf, n, N, false, bp.idepth_min, bp.idepth_step, s);
      break;
    }
    case QM_STAGE_PYRAMID: {
      bool prior = false;
      for (int i = 0; i < n; ++i) prior |= hf[i].prior != nullptr;
      if (prior) {
        Mark m(t, QM_STAGE_PYRAMID);
        launch_pyramid(df, hf, n, N, s);
      }
      break;
    }
    case QM_STAGE_BP_SWEEP: {
      int launches = 0;
      for_shape_runs(b, f0, n, same_bp_shape,
                     [&](int, int, const Tuning& tu) { launches += bp_launches(p->bp_iters, tu); });
      Mark m(t, QM_STAGE_BP_SWEEP, launches);
      // the last level-0 sweep (when not the level's first) extracts its colour's cells
      for_shape_runs(b, f0, n, same_bp_shape, [&](int a, int k, const Tuning& tu) {
        launch_bp_levels(df + a, hf + a, k, N, p->bp_iters, bp, tu, s, true, false);
      });
      break;
    }
    case QM_STAGE_EXTRACT: {
      // the other colour's cells: colour 0 (x + y even) is the A set, swept at even it
      int runs = 0;
      for_shape_runs(b, f0, n, same_extract_shape, [&](int, int, const Tuning&) { ++runs; });
      Mark m(t, QM_STAGE_EXTRACT, runs);
      const int it0 = p->bp_iters[0];
      for_shape_runs(b, f0, n, same_extract_shape, [&](int a, int k, const Tuning& tu) {
        launch_extract(df + a, hf + a, k, N, bp, tu, s, it0 >= 2 ? ((it0 - 1) % 2 == 0 ? 1 : 0) : -1,
                       bp_final_buf(0, it0, tu));
      });
      break;
    }
    case QM_STAGE_UPSAMPLE_ROW: {
      Mark m(t, QM_STAGE_UPSAMPLE_ROW);
      // the row scan resets the next run's z-buffer in whole lines when it
      // takes every frame of the launch; else the column pass does
      launch_upsample_row(df, hf, n, 0, b->wlut, s, p->upsample_solver, true);
      break;
    }
    case QM_STAGE_UPSAMPLE_COL: {
      Mark m(t, QM_STAGE_UPSAMPLE_COL);
      const bool rows_reset = p->upsample_solver == QM_UPSAMPLE_SCAN && row_scan_covers(hf, n);
      launch_upsample_col(df, hf, n, 0, b->wlut, !rows_reset, s, p->upsample_solver);
      break;
    }
    case QM_STAGE_FUSE_WARP: {
      Mark m(t, QM_STAGE_FUSE_WARP);
      launch_fuse_warp(df, hf, n, s);
      break;
    }
    default: {
      Mark m(t, QM_STAGE_FUSE_UPDATE);
      launch_fuse_update(df, hf, n, s);
      break;
    }
  }
}

// All stages of the whole batch on one stream.
void run_stages(qm_batch* b, int par, hipStream_t s, Timer* t = nullptr) {
  for (int st = 0; st < QM_NUM_STAGES; ++st) enqueue_stage(b, par, s, t, 0, b->n, st);
}

// The overlap setting. With it on (and two frames or more) a run splits the
// batch into G frame groups, group 0 on s and the others on their own streams
// forked from s and joined back into it: the kernels of one group fill the GPU
// while another group's latency-bound ones (coarse BP levels, the upsample)
// run. Group g + 1 starts when group g has run the stage the setting names. Settings: 1-3 two groups, starting together,
// after the cost volume, after BP; 4 three and 5 four groups, each after the
// previous group's cost volume; 6 eight such groups; 7 four groups, each after
// the previous group's BP. Settings 8, 9 and 10 (two, three, four groups) run
// every group's cost volume on s, one after the other, and fork each group's
// later stages onto a stream of its own once its cost volume is done (the
// last group's stay on s): s then runs one long kernel while the other streams
// start theirs, where a string of short kernels on s (the coarse BP levels)
// holds back another stream's first kernel. Settings 11, 12 and 13 are 8, 9
// and 10 with the fusion warp of the whole batch run first, on the first side
// stream beside the cost volumes (a run's warp scatters into its own parity's
// z-buffer, which the previous run's column pass reset). Every setting
// computes the same results.
struct OverlapPlan {
  int groups, mark_after;
  bool costs_first, early_warp;
};
constexpr int kOverlapSettings = 14;
OverlapPlan overlap_plan(int ov, int n) {
  OverlapPlan p{1, -1, false, false};
  if (ov == 1) p = {2, -1, false, false};
  if (ov == 2) p = {2, 1, false, false};
  if (ov == 3) p = {2, 3, false, false};
  if (ov == 4) p = {3, 1, false, false};
  if (ov == 5) p = {4, 1, false, false};
  if (ov == 6) p = {8, 1, false, false};
  if (ov == 7) p = {4, 3, false, false};
  if (ov >= 8 && ov <= 13) p = {2 + (ov - 8) % 3, 1, true, ov >= 11};
  if (p.groups > n) p.groups = n;
  return p;
}

// the extra streams and the events, created once (outside any capture)
int ensure_side(qm_batch* b) {
  if (b->side_ready) return QM_OK;
  if (!b->ev_fork) QM_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  if (!b->ev_warp) QM_HIP(hipEventCreateWithFlags(&b->ev_warp, hipEventDisableTiming));
  for (int g = 0; g < qm_batch::kMaxGroups - 1; ++g) {
    if (!b->ev_mark[g]) QM_HIP(hipEventCreateWithFlags(&b->ev_mark[g], hipEventDisableTiming));
    if (!b->ev_join[g]) QM_HIP(hipEventCreateWithFlags(&b->ev_join[g], hipEventDisableTiming));
    if (!b->side[g]) QM_HIP(hipStreamCreateWithFlags(&b->side[g], hipStreamNonBlocking));
  }
  b->side_ready = true;
  return QM_OK;
}

void apply_desc(FrameDev& F, const qm_frame_desc& d, const qm_params& p) {
  F.lam = d.lambda > 0.0 ? d.lambda : p.lambda;
  F.fx = d.cam[0];
  F.fy = d.cam[1];
  F.cx = d.cam[2];
  F.cy = d.cam[3];
  F.key = d.key;
  for (int r = 0; r < 4; ++r) {
    F.ref[r] = d.refs[r];
    std::memcpy(F.T[r], d.T_ref_key[r], sizeof(float) * 12);
  }
  std::memcpy(F.T_fuse, d.T_last_to_cur, sizeof(float) * 12);
  F.prefill = d.prefill;
  F.prior = (const float2*)d.prior;
  F.prior_depth = d.prior_depth;
  F.depth_out = d.depth_out;
  F.dense_out = d.dense_out;
  F.sparse_out = d.sparse_out;
}

int check_desc(const qm_frame_desc& d) {
  int rc = check_size(d.W, d.H, d.pitch);
  if (rc) return rc;
  if (!d.key) return fail(QM_ERR_INVALID, "frame key image is NULL");
  for (int r = 0; r < 4; ++r)
    if (!d.refs[r]) return fail(QM_ERR_INVALID, "frame reference image is NULL");
  if (!(d.cam[0] != 0.0f) || !(d.cam[1] != 0.0f)) return fail(QM_ERR_INVALID, "focal length is zero");
  return QM_OK;
}

int upload_frames(qm_batch* b) {
  for (int par = 0; par < 2; ++par)
    QM_HIP(hipMemcpy(b->df[par], b->hf[par].data(), sizeof(FrameDev) * b->n, hipMemcpyHostToDevice));
  return QM_OK;
}

}  // namespace

extern "C" {

int qm_abi_version(void) { return QM_ABI_VERSION; }

const char* qm_last_error(void) { return g_err.c_str(); }

void qm_default_params(qm_params* p) {
  if (!p) return;
  p->n_hyp = 64;
  p->idepth_min = 0.01;
  p->idepth_max = 2.0;
  p->p1 = 0.003f;
  p->p2 = 0.01f;
  p->bp_iters[0] = 4;
  p->bp_iters[1] = 10;
  p->bp_iters[2] = 10;
  p->bp_iters[3] = 10;
  p->lambda = 10.0;
  p->min_refs = 11;
  p->max_refs = 60;
  p->ref_cos = 0.95f;
  p->ref_baseline = 0.03f;
  p->cost_arith = QM_COST_REFERENCE;
  p->upsample_solver = QM_UPSAMPLE_THOMAS;
}

void qm_set_alloc_cap(size_t bytes) { g_cap.store(bytes); }
int qm_release_scratch(void) {
  flush_scratch_cache();
  return QM_OK;
}
size_t qm_device_bytes_in_use(void) { return g_bytes.load(); }

int qm_quadtree(const uint8_t* img, int W, int H, int pitch, uint8_t* levels, void* stream) {
  int rc = check_size(W, H, pitch);
  if (rc) return rc;
  if (!img || !levels) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  FrameDev F;
  init_geometry(F, W, H, pitch);
  F.key = img;
  F.levels = levels;
  Scratch sc;
  void* df;
  if ((rc = sc.get(&df, sizeof(FrameDev)))) return rc;
  QM_HIP(hipMemcpyAsync(df, &F, sizeof(F), hipMemcpyHostToDevice, s));
  launch_quadtree((FrameDev*)df, &F, 1, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_cost_volume(const uint8_t* key, const uint8_t* const refs[4], int W, int H, int pitch,
                   const float T_ref_key[48], const float cam[4], const qm_params* p, const uint8_t* levels,
                   float* cost, int32_t* count, void* stream) {
  int rc = check_size(W, H, pitch);
  if (rc) return rc;
  if ((rc = check_params(p))) return rc;
  if (!key || !refs || !T_ref_key || !cam || !levels || !cost) return fail(QM_ERR_INVALID, "NULL pointer");
  for (int r = 0; r < 4; ++r)
    if (!refs[r]) return fail(QM_ERR_INVALID, "NULL reference");
  hipStream_t s = (hipStream_t)stream;
  FrameDev F;
  init_geometry(F, W, H, pitch);
  F.key = key;
  for (int r = 0; r < 4; ++r) {
    F.ref[r] = refs[r];
    std::memcpy(F.T[r], T_ref_key + 12 * r, sizeof(float) * 12);
  }
  F.fx = cam[0];
  F.fy = cam[1];
  F.cx = cam[2];
  F.cy = cam[3];
  F.levels = const_cast<uint8_t*>(levels);
  F.cost[0] = cost;
  F.count = count;
  std::vector<float> tab(p->n_hyp);
  host_depth_table(p->n_hyp, p->idepth_min, p->idepth_max, tab.data());
  Scratch sc;
  void *df, *dt;
  if ((rc = sc.get(&df, sizeof(FrameDev)))) return rc;
  if ((rc = sc.get(&dt, sizeof(float) * p->n_hyp))) return rc;
  QM_HIP(hipMemcpyAsync(df, &F, sizeof(F), hipMemcpyHostToDevice, s));
  QM_HIP(hipMemcpyAsync(dt, tab.data(), sizeof(float) * p->n_hyp, hipMemcpyHostToDevice, s));
  launch_cost((FrameDev*)df, &F, 1, p->n_hyp, (const float*)dt, true, p->cost_arith, Tuning(), s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_bp_extract(const float* cost, int W, int H, const uint8_t* levels, const qm_params* p, float* depth_coarse,
                  int32_t* label, float* messages, void* stream) {
  int rc = check_size(W, H, W);
  if (rc) return rc;
  if ((rc = check_params(p))) return rc;
  if (!cost || !levels || !depth_coarse || !label) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  const int N = p->n_hyp;
  FrameDev F;
  init_geometry(F, W, H, W);
  F.levels = const_cast<uint8_t*>(levels);
  F.dcoarse = depth_coarse;
  F.label = label;
  Scratch sc;
  void* df;
  if ((rc = sc.get(&df, sizeof(FrameDev)))) return rc;
  for (int l = 0; l < 4; ++l) {
    const size_t c = (size_t)F.lw[l] * F.lh[l];
    void* m;
    if ((rc = sc.get(&m, c * 4 * N * sizeof(float)))) return rc;
    F.msg[l] = (float*)m;
    if ((rc = sc.get(&m, c * 4 * N * sizeof(float)))) return rc;
    F.msg2[l] = (float*)m;
    void* cl;
    if ((rc = sc.get(&cl, c * N * sizeof(float)))) return rc;
    F.cost[l] = (float*)cl;
  }
  // the caller's volume is in sample order; the BP kernels read the row layout
  launch_cost_to_rows(cost, F.cost[0], (size_t)F.lw[0] * F.lh[0], N, s);
  QM_HIP(hipMemcpyAsync(df, &F, sizeof(F), hipMemcpyHostToDevice, s));
  const int buf = run_bp((FrameDev*)df, &F, 1, p, s, messages != nullptr);
  if (messages) launch_copy_messages((FrameDev*)df, &F, 1, N, messages, s, buf);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_upsample(const float* sparse, const uint8_t* img, int W, int H, int pitch, const qm_params* p, float* dense,
                void* stream) {
  int rc = check_size(W, H, pitch);
  if (rc) return rc;
  if ((rc = check_params(p))) return rc;
  if (!sparse || !img || !dense) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  FrameDev F;
  init_geometry(F, W, H, pitch);
  F.key = img;
  F.sparse_in = sparse;
  F.lam = p->lambda;
  F.dense = dense;
  float lut[256];
  host_weight_lut(lut);
  Scratch sc;
  void *df, *fwd, *rxn, *dl;
  if ((rc = sc.get(&df, sizeof(FrameDev)))) return rc;
  if ((rc = sc.get(&fwd, (size_t)W * H * 16))) return rc;
  if ((rc = sc.get(&rxn, (size_t)W * H * 8))) return rc;
  if ((rc = sc.get(&dl, sizeof(lut)))) return rc;
  F.fwd = (float4*)fwd;
  F.rowx = (float*)rxn;
  QM_HIP(hipMemcpyAsync(df, &F, sizeof(F), hipMemcpyHostToDevice, s));
  QM_HIP(hipMemcpyAsync(dl, lut, sizeof(lut), hipMemcpyHostToDevice, s));
  launch_upsample((FrameDev*)df, &F, 1, 1, (const float*)dl, false, s, p->upsample_solver);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_upsample_general_batch(const qm_upsample_problem* probs, int n, void* stream) {
  if (!probs || n <= 0 || n > kMaxFrames) return fail(QM_ERR_INVALID, "need 1..128 problems");
  int rc;
  for (int i = 0; i < n; ++i) {
    const qm_upsample_problem& q = probs[i];
    if ((rc = check_size(q.W, q.H, q.W))) return rc;
    if (!q.d || !q.w_row || !q.w_col || !q.dense) return fail(QM_ERR_INVALID, "NULL pointer");
    if ((q.row_x == nullptr) != (q.row_n == nullptr)) return fail(QM_ERR_INVALID, "row_x and row_n go together");
    if (q.row_seg < 0 || q.col_seg < 0) return fail(QM_ERR_INVALID, "negative segment");
    if (!(q.lambda >= 0.0)) return fail(QM_ERR_INVALID, "lambda must be >= 0");
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<FrameDev> hf(n);
  Scratch sc;
  void* df;
  if ((rc = sc.get(&df, sizeof(FrameDev) * n))) return rc;
  for (int i = 0; i < n; ++i) {
    const qm_upsample_problem& q = probs[i];
    FrameDev& F = hf[i];
    init_geometry(F, q.W, q.H, q.W);
    F.row_seg = q.row_seg;
    F.col_seg = q.col_seg;
    F.lam = q.lambda;
    F.g_d = q.d;
    F.g_conf = q.conf;
    F.g_wrow = q.w_row;
    F.g_wcol = q.w_col;
    F.row_x_out = q.row_x;
    F.row_n_out = q.row_n;
    F.dense = q.dense;
    const size_t px = (size_t)q.W * q.H;
    void* fwd;
    if ((rc = sc.get(&fwd, px * 16))) return rc;
    F.fwd = (float4*)fwd;
    if (q.state) {
      // the row pass's X and N planes, then the column pass's (x, n) pairs
      F.rowx = q.state;
      F.colxn = (float2*)(q.state + 2 * px);
    } else {
      void* rxn;
      if ((rc = sc.get(&rxn, px * 8))) return rc;
      F.rowx = (float*)rxn;
    }
  }
  QM_HIP(hipMemcpyAsync(df, hf.data(), sizeof(FrameDev) * n, hipMemcpyHostToDevice, s));
  launch_upsample((FrameDev*)df, hf.data(), n, 2, nullptr, false, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_upsample_backward_batch(const qm_upsample_problem* probs, int n, void* stream) {
  if (!probs || n <= 0 || n > kMaxFrames) return fail(QM_ERR_INVALID, "need 1..128 problems");
  int rc;
  for (int i = 0; i < n; ++i) {
    const qm_upsample_problem& q = probs[i];
    if ((rc = check_size(q.W, q.H, q.W))) return rc;
    if (!q.d || !q.w_row || !q.w_col || !q.g) return fail(QM_ERR_INVALID, "NULL pointer");
    if (q.row_seg < 0 || q.col_seg < 0) return fail(QM_ERR_INVALID, "negative segment");
    if (!(q.lambda >= 0.0)) return fail(QM_ERR_INVALID, "lambda must be >= 0");
    if ((q.state == nullptr) != (probs[0].state == nullptr))
      return fail(QM_ERR_INVALID, "every problem of a backward batch has a forward state, or none does");
  }
  const bool saved = probs[0].state != nullptr;
  hipStream_t s = (hipStream_t)stream;
  std::vector<FrameDev> hf(n);
  Scratch sc;
  void* df;
  if ((rc = sc.get(&df, sizeof(FrameDev) * n))) return rc;
  for (int i = 0; i < n; ++i) {
    const qm_upsample_problem& q = probs[i];
    FrameDev& F = hf[i];
    init_geometry(F, q.W, q.H, q.W);
    F.row_seg = q.row_seg;
    F.col_seg = q.col_seg;
    F.lam = q.lambda;
    F.g_d = q.d;
    F.g_conf = q.conf;
    F.g_wrow = q.w_row;
    F.g_wcol = q.w_col;
    F.g_out = q.g;
    F.grad_d = q.grad_d;
    F.grad_conf = q.grad_conf;
    F.grad_wrow = q.grad_w_row;
    if (q.grad_lambda) {
      void* lp;
      if ((rc = sc.get(&lp, sizeof(double) * (size_t)(q.W + q.H)))) return rc;
      F.lam_part = (double*)lp;
      F.grad_lambda = q.grad_lambda;
    }
    const size_t px = (size_t)q.W * q.H;
    void *fwd, *axn;
    if ((rc = sc.get(&fwd, px * 16))) return rc;
    if ((rc = sc.get(&axn, px * 8))) return rc;
    F.fwd = (float4*)fwd;
    F.adjxn = (float2*)axn;
    if (q.grad_d || q.grad_conf) {
      void* rv;  // the row adjoint's solution planes, for dL/dd and dL/dconf
      if ((rc = sc.get(&rv, px * 8))) return rc;
      F.adj_rowv = (float*)rv;
    }
    if (saved) {
      F.rowx = q.state;
      F.colxn = (float2*)(q.state + 2 * px);
    } else {
      void *rxn, *cxn, *dense;
      if ((rc = sc.get(&rxn, px * 8))) return rc;
      if ((rc = sc.get(&cxn, px * 8))) return rc;
      if ((rc = sc.get(&dense, px * 4))) return rc;
      F.rowx = (float*)rxn;
      F.colxn = (float2*)cxn;
      F.dense = (float*)dense;
    }
    F.grad_wcol = q.grad_w_col;
  }
  QM_HIP(hipMemcpyAsync(df, hf.data(), sizeof(FrameDev) * n, hipMemcpyHostToDevice, s));
  launch_upsample_backward((FrameDev*)df, hf.data(), n, s, !saved);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_upsample_general(const float* d, const float* conf, const float* w_row, const float* w_col, int W, int H,
                        double lambda, int row_seg, int col_seg, float* dense, float* row_x, float* row_n,
                        void* stream) {
  qm_upsample_problem q;
  std::memset(&q, 0, sizeof(q));
  q.W = W;
  q.H = H;
  q.row_seg = row_seg;
  q.col_seg = col_seg;
  q.lambda = lambda;
  q.d = d;
  q.conf = conf;
  q.w_row = w_row;
  q.w_col = w_col;
  q.dense = dense;
  q.row_x = row_x;
  q.row_n = row_n;
  return qm_upsample_general_batch(&q, 1, stream);
}

int qm_upsample_backward(const float* d, const float* conf, const float* w_row, const float* w_col, int W, int H,
                         double lambda, int row_seg, int col_seg, const float* g, float* grad_d, float* grad_conf,
                         float* grad_w_row, float* grad_w_col, void* stream) {
  qm_upsample_problem q;
  std::memset(&q, 0, sizeof(q));
  q.W = W;
  q.H = H;
  q.row_seg = row_seg;
  q.col_seg = col_seg;
  q.lambda = lambda;
  q.d = d;
  q.conf = conf;
  q.w_row = w_row;
  q.w_col = w_col;
  q.g = g;
  q.grad_d = grad_d;
  q.grad_conf = grad_conf;
  q.grad_w_row = grad_w_row;
  q.grad_w_col = grad_w_col;
  return qm_upsample_backward_batch(&q, 1, stream);
}

int qm_quadtree_prior(const float* prior_depth, int W, int H, uint8_t* levels, void* stream) {
  int rc = check_size(W, H, W);
  if (rc) return rc;
  if (!prior_depth || !levels) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  FrameDev F;
  init_geometry(F, W, H, W);
  F.prior_depth = prior_depth;
  F.levels = levels;
  Scratch sc;
  void* df;
  if ((rc = sc.get(&df, sizeof(FrameDev)))) return rc;
  QM_HIP(hipMemcpyAsync(df, &F, sizeof(F), hipMemcpyHostToDevice, s));
  launch_quadtree_prior((FrameDev*)df, &F, 1, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_prior_to_cost(const float* prior, int W, int H, const uint8_t* levels, const qm_params* p, float* cost,
                     int32_t* count, void* stream) {
  int rc = check_size(W, H, W);
  if (rc) return rc;
  if ((rc = check_params(p))) return rc;
  if (!prior || !levels || !cost) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  FrameDev F;
  init_geometry(F, W, H, W);
  F.prior = (const float2*)prior;
  F.levels = const_cast<uint8_t*>(levels);
  F.cost[0] = cost;
  F.count = count;
  Scratch sc;
  void* df;
  if ((rc = sc.get(&df, sizeof(FrameDev)))) return rc;
  QM_HIP(hipMemcpyAsync(df, &F, sizeof(F), hipMemcpyHostToDevice, s));
  const BpLaunchParams bp = bp_params(p);
  launch_prior_cost((FrameDev*)df, &F, 1, p->n_hyp, true, bp.idepth_min, bp.idepth_step, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_upsample_local(const float* sparse, const uint8_t* img, int W, int H, int pitch, float* dense, void* stream) {
  int rc = check_size(W, H, pitch);
  if (rc) return rc;
  if (!sparse || !img || !dense) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  launch_upsample_local(sparse, img, W, H, pitch, dense, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  return QM_OK;
}

int qm_point_cloud(const float* depth, const uint8_t* ref_img, int W, int H, const float cam[4],
                   const float T_world_ref[12], float* xyzi, int* count, void* stream) {
  int rc = check_size(W, H, W);
  if (rc) return rc;
  if (!depth || !ref_img || !cam || !T_world_ref || !xyzi || !count) return fail(QM_ERR_INVALID, "NULL pointer");
  return point_cloud(depth, ref_img, W, H, cam, T_world_ref, xyzi, count, (hipStream_t)stream);
}

int qm_sd_gradient(const uint8_t* img, int W, int H, int pitch, float* grad, void* stream) {
  int rc = check_size(W, H, pitch);
  if (rc) return rc;
  if (!img || !grad) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  sd_gradient(img, W, H, pitch, (float2*)grad, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  return QM_OK;
}

int qm_sd_init(const float* grad, int W, int H, float* seeds, void* stream) {
  int rc = check_size(W, H, W);
  if (rc) return rc;
  if (!grad || !seeds) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  sd_init((const float2*)grad, W, H, (float4*)seeds, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  return QM_OK;
}

int qm_sd_step(float* seeds, const uint8_t* key, const float* key_grad, const uint8_t* inc, const float* inc_grad,
               int W, int H, int pitch, const float cam[4], const float T_key_world[12],
               const float T_inc_world[12], int switch_keyframe, float* prefill, float* prior, void* stream) {
  int rc = check_size(W, H, pitch);
  if (rc) return rc;
  if (!seeds || !key || !key_grad || !inc || !inc_grad || !cam || !T_key_world || !T_inc_world || !prefill)
    return fail(QM_ERR_INVALID, "NULL pointer");
  if (!(cam[0] != 0.0f) || !(cam[1] != 0.0f)) return fail(QM_ERR_INVALID, "focal length is zero");
  hipStream_t s = (hipStream_t)stream;
  const size_t px = (size_t)W * H;
  Scratch sc;
  void *second, *u64, *info;
  if ((rc = sc.get(&second, px * 32))) return rc;
  if ((rc = sc.get(&u64, px * 8))) return rc;
  if ((rc = sc.get(&info, px * 16))) return rc;
  float4* bufs[2] = {(float4*)seeds, (float4*)second};
  sd_step(bufs, key, (const float2*)key_grad, inc, (const float2*)inc_grad, W, H, pitch, cam, T_key_world,
          T_inc_world, switch_keyframe != 0, prefill, (float2*)prior, (unsigned long long*)u64, (float4*)info, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_fuse_backward(const float* seeds_warped, const int32_t* table, const float* depth_in, int W, int H,
                     const float* g_depth_out, const float* g_seeds_out, float* grad_depth_in, void* stream) {
  int rc = check_size(W, H, W);
  if (rc) return rc;
  if (!seeds_warped || !table || !depth_in || !grad_depth_in) return fail(QM_ERR_INVALID, "NULL pointer");
  if (!g_depth_out && !g_seeds_out) return fail(QM_ERR_INVALID, "need at least one incoming gradient");
  hipStream_t s = (hipStream_t)stream;
  FrameDev F;
  init_geometry(F, W, H, W);
  F.seeds = (float4*)const_cast<float*>(seeds_warped);
  F.table_in = table;
  F.dense = const_cast<float*>(depth_in);
  F.g_depth_out = g_depth_out;
  F.g_seeds_out = (const float4*)g_seeds_out;
  F.grad_depth_in = grad_depth_in;
  Scratch sc;
  void* df;
  if ((rc = sc.get(&df, sizeof(FrameDev)))) return rc;
  QM_HIP(hipMemcpyAsync(df, &F, sizeof(F), hipMemcpyHostToDevice, s));
  launch_fuse_backward((FrameDev*)df, &F, 1, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

int qm_fuse(float* seeds, const float* depth_in, int W, int H, const float cam[4], const float T_last_to_cur[12],
            float* depth_out, float* seeds_out, int32_t* table, void* stream) {
  int rc = check_size(W, H, W);
  if (rc) return rc;
  if (!seeds || !depth_in || !cam || !T_last_to_cur || !seeds_out) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  FrameDev F;
  init_geometry(F, W, H, W);
  F.fx = cam[0];
  F.fy = cam[1];
  F.cx = cam[2];
  F.cy = cam[3];
  std::memcpy(F.T_fuse, T_last_to_cur, sizeof(float) * 12);
  F.seeds = (float4*)seeds;
  F.seeds_new = (float4*)seeds_out;
  F.dense = const_cast<float*>(depth_in);
  F.depth_out = depth_out;
  F.table_out = table;
  Scratch sc;
  void *df, *zb;
  if ((rc = sc.get(&df, sizeof(FrameDev)))) return rc;
  if ((rc = sc.get(&zb, (size_t)W * H * 8))) return rc;
  F.zbuf = (unsigned long long*)zb;
  QM_HIP(hipMemcpyAsync(df, &F, sizeof(F), hipMemcpyHostToDevice, s));
  launch_fuse((FrameDev*)df, &F, 1, true, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  sc.done();
  return QM_OK;
}

// ---------------------------------------------------------------------------
int qm_batch_create(const qm_frame_desc* frames, int n, const qm_params* p, qm_batch** out) {
  if (!out) return fail(QM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  if (!frames || n <= 0 || n > kMaxFrames) return fail(QM_ERR_INVALID, "need 1..128 frames");
  for (int i = 0; i < n; ++i)
    if ((rc = check_desc(frames[i]))) return rc;
  qm_batch* b = new qm_batch();
  b->p = *p;
  b->n = n;
  b->group_of.resize(n);
  for (int i = 0; i < n; ++i) {
    int g = 0;
    while (g < (int)b->group_w.size() && (b->group_w[g] != frames[i].W || b->group_h[g] != frames[i].H)) ++g;
    if (g == (int)b->group_w.size()) {
      b->group_w.push_back(frames[i].W);
      b->group_h.push_back(frames[i].H);
    }
    b->group_of[i] = g;
  }
  b->gtune.assign(b->group_w.size(), Tuning());
  const int N = p->n_hyp;
  std::vector<FrameDev> base(n);
  size_t total = align256(sizeof(float) * (N + 256));
  for (int i = 0; i < n; ++i) {
    init_geometry(base[i], frames[i].W, frames[i].H, frames[i].pitch);
    total += frame_bytes(base[i], N);
  }
  void* fd0 = nullptr;
  void* fd1 = nullptr;
  if ((rc = dev_alloc(&b->arena, total)) || (rc = dev_alloc(&fd0, sizeof(FrameDev) * n)) ||
      (rc = dev_alloc(&fd1, sizeof(FrameDev) * n))) {
    dev_free(b->arena);
    dev_free(fd0);
    dev_free(fd1);
    delete b;
    return rc;
  }
  b->df[0] = (FrameDev*)fd0;
  b->df[1] = (FrameDev*)fd1;
  Carver cv{(char*)b->arena};
  b->table = cv.take<float>(sizeof(float) * (N + 256));
  b->wlut = b->table + N;
  b->seeds[0].resize(n);
  b->seeds[1].resize(n);
  b->zbuf[0].resize(n);
  b->zbuf[1].resize(n);
  b->table_buf.resize(n);
  for (int i = 0; i < n; ++i) {
    FrameDev& F = base[i];
    const size_t px = (size_t)F.W * F.H;
    F.levels = cv.take<uint8_t>(px);
    for (int l = 0; l < 4; ++l) {
      const size_t c = (size_t)F.lw[l] * F.lh[l];
      F.cost[l] = cv.take<float>(c * N * 4);
      F.msg[l] = cv.take<float>(c * 4 * N * 4);
      F.msg2[l] = cv.take<float>(c * 4 * N * 4);
    }
    const size_t cc = (size_t)F.Wc * F.Hc;
    F.dcoarse = cv.take<float>(cc * 4);
    F.label = cv.take<int32_t>(cc * 4);
    F.fwd = cv.take<float4>(px * 16);
    F.rowx = cv.take<float>(px * 8);
    F.dense = cv.take<float>(px * 4);
    b->seeds[0][i] = cv.take<float4>(px * 16);
    b->seeds[1][i] = cv.take<float4>(px * 16);
    b->zbuf[0][i] = cv.take<unsigned long long>(px * 8);
    b->zbuf[1][i] = cv.take<unsigned long long>(px * 8);
    b->table_buf[i] = cv.take<int32_t>(px * 4);
    F.table_out = nullptr;
    F.warp_skip = 1;
    apply_desc(F, frames[i], *p);
  }
  for (int par = 0; par < 2; ++par) {
    b->hf[par] = base;
    for (int i = 0; i < n; ++i) {
      b->hf[par][i].seeds = b->seeds[par][i];
      b->hf[par][i].seeds_new = b->seeds[par ^ 1][i];
      b->hf[par][i].zbuf = b->zbuf[par][i];
      b->hf[par][i].zbuf_clear = b->zbuf[par ^ 1][i];
    }
  }
  std::vector<float> tabs(N + 256);
  host_depth_table(N, p->idepth_min, p->idepth_max, tabs.data());
  host_weight_lut(tabs.data() + N);
  hipError_t e = hipMemcpy(b->table, tabs.data(), sizeof(float) * (N + 256), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    for (int i = 0; i < n && e == hipSuccess; ++i) {
      const size_t px = (size_t)base[i].W * base[i].H;
      e = hipMemset(b->seeds[0][i], 0, px * 16);
      if (e == hipSuccess) e = hipMemset(b->seeds[1][i], 0, px * 16);
      // both z-buffers start empty (the first run's warp scatters into one)
      for (int q = 0; q < 2 && e == hipSuccess; ++q) e = hipMemset(b->zbuf[q][i], 0xFF, px * 8);
    }
  }
  if (e != hipSuccess || upload_frames(b) != QM_OK) {
    const std::string msg = e != hipSuccess ? std::string("setup: ") + hipGetErrorString(e) : g_err;
    qm_batch_destroy(b);
    return fail(QM_ERR_HIP, msg);
  }
  *out = b;
  return QM_OK;
}

static void drop_graphs(qm_batch* b);

int qm_batch_set_frame(qm_batch* b, int i, const qm_frame_desc* d) {
  if (!b || !d) return fail(QM_ERR_INVALID, "NULL pointer");
  if (i < 0 || i >= b->n) return fail(QM_ERR_INVALID, "frame index out of range");
  int rc = check_desc(*d);
  if (rc) return rc;
  const FrameDev& F0 = b->hf[0][i];
  if (d->W != F0.W || d->H != F0.H || d->pitch != F0.pitch)
    return fail(QM_ERR_INVALID, "frame size differs from the planned one");
  const bool had_prior = any_prior(b), had_prior_depth = any_prior_depth(b);
  for (int par = 0; par < 2; ++par) apply_desc(b->hf[par][i], *d, b->p);
  b->run_state_valid = false;
  // the alternates' launches are part of a captured graph only when some frame has the prior
  if (had_prior != any_prior(b) || had_prior_depth != any_prior_depth(b)) drop_graphs(b);
  return upload_frames(b);
}

// A run as a list of linear segments, each a run of consecutive stages of one
// frame group on one stream, with events between the streams. Plain runs
// launch the segments' kernels; graph runs capture each segment as a graph of
// its own (a linear graph runs on the stream it is launched on) and launch
// those, which replays faster than one graph holding every branch (a 32-map
// run 1.775 -> 1.750 ms).
enum { kEvFork = 0, kEvWarp = 1, kEvMark = 2, kEvJoin = 2 + qm_batch::kMaxGroups - 1 };
static hipEvent_t event_of(const qm_batch* b, int id) {
  if (id == kEvFork) return b->ev_fork;
  if (id == kEvWarp) return b->ev_warp;
  return id < kEvJoin ? b->ev_mark[id - kEvMark] : b->ev_join[id - kEvJoin];
}

static std::vector<qm_batch::Segment> plan_segments(const qm_batch* b) {
  using Seg = qm_batch::Segment;
  const OverlapPlan plan = overlap_plan(b->tune.overlap, b->n);
  std::vector<int> active;
  for (int u = 0; u < QM_NUM_STAGES; ++u)
    if (u != QM_STAGE_PYRAMID || any_prior(b)) active.push_back(u);
  std::vector<Seg> out;
  const int G = plan.groups;
  if (G < 2) {
    out.push_back(Seg{0, b->n, active, 0, {}, -1});
    return out;
  }
  int f0[qm_batch::kMaxGroups], nf[qm_batch::kMaxGroups];
  for (int g = 0, at = 0; g < G; ++g) {
    nf[g] = (b->n - at) / (G - g);
    f0[g] = at;
    at += nf[g];
  }
  if (plan.costs_first) {
    std::vector<int> heads, tails;
    for (int u : active) {
      if (u <= QM_STAGE_PYRAMID) heads.push_back(u);
      else if (!(plan.early_warp && u == QM_STAGE_FUSE_WARP)) tails.push_back(u);
    }
    if (plan.early_warp) out.push_back(Seg{0, b->n, {QM_STAGE_FUSE_WARP}, 1, {kEvFork}, kEvWarp});
    for (int g = 0; g < G; ++g) out.push_back(Seg{f0[g], nf[g], heads, 0, {}, g + 1 < G ? kEvMark + g : -1});
    for (int g = 0; g < G; ++g) {
      Seg t{f0[g], nf[g], tails, g + 1 < G ? g + 1 : 0, {}, g + 1 < G ? kEvJoin + g : -1};
      if (g + 1 < G) t.waits.push_back(kEvMark + g);
      if (plan.early_warp && g > 0) t.waits.push_back(kEvWarp);
      out.push_back(t);
    }
    return out;
  }
  int mark = -1;
  for (int u = 0; u < (int)active.size(); ++u)
    if (active[u] == plan.mark_after) mark = u;
  for (int g = 0; g < G; ++g) {
    std::vector<int> first_waits;
    if (g > 0) first_waits.push_back(mark >= 0 ? kEvMark + g - 1 : kEvFork);
    const int end_rec = g > 0 ? kEvJoin + g - 1 : -1;
    if (mark >= 0 && g + 1 < G) {
      out.push_back(Seg{f0[g], nf[g], std::vector<int>(active.begin(), active.begin() + mark + 1), g, first_waits,
                        kEvMark + g});
      out.push_back(Seg{f0[g], nf[g], std::vector<int>(active.begin() + mark + 1, active.end()), g, {}, end_rec});
    } else {
      out.push_back(Seg{f0[g], nf[g], active, g, first_waits, end_rec});
    }
  }
  return out;
}

// the segments' graphs of parity par, captured once (set-up calls drop them)
static int capture_segments(qm_batch* b, int par) {
  if (b->captured[par]) return QM_OK;
  // both parities share one plan: planned when neither is captured
  if (!b->captured[par ^ 1]) b->segs = plan_segments(b);
  if (b->segs.size() > 1) {
    const int rc = ensure_side(b);
    if (rc) return rc;
  }
  hipStream_t cs;
  QM_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipSuccess;
  std::vector<hipGraphExec_t>& ex = b->seg_exec[par];
  for (const qm_batch::Segment& sg : b->segs) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) break;
    for (int st : sg.stages) enqueue_stage(b, par, cs, nullptr, sg.f0, sg.nf, st);
    e = hipStreamEndCapture(cs, &g);
    size_t nodes = 0;
    if (e == hipSuccess) e = hipGraphGetNodes(g, nullptr, &nodes);
    if (e == hipSuccess && nodes > 0) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) break;
    ex.push_back(ge);  // null for a segment without kernels
  }
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) {
    for (hipGraphExec_t x : ex)
      if (x) (void)hipGraphExecDestroy(x);
    ex.clear();
    return fail(QM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  }
  b->captured[par] = true;
  return QM_OK;
}

// Launches the segments on their streams, as the captured graphs of parity
// par (graph) or as plain kernel launches: every event is recorded (in host
// order) before a wait on it is issued, and the side streams' segments go
// first whenever their events are recorded, so that their queues start at once.
static int launch_segments(qm_batch* b, int par, hipStream_t s, const std::vector<qm_batch::Segment>& segs,
                           bool graph) {
  const int ns = (int)segs.size();
  if (ns > 1) {
    const int rc = ensure_side(b);
    if (rc) return rc;
  }
  auto stream_of = [&](int k) { return k == 0 ? s : b->side[k - 1]; };
  bool recorded[kEvJoin + qm_batch::kMaxGroups] = {};
  std::vector<bool> done(ns, false);
  if (ns > 1) {
    QM_HIP(hipEventRecord(b->ev_fork, s));
    recorded[kEvFork] = true;
  }
  for (int left = ns; left > 0;) {
    int pick = -1;
    for (int pass = 0; pass < 2 && pick < 0; ++pass)
      for (int i = 0; i < ns && pick < 0; ++i) {
        if (done[i] || (pass == 0 && segs[i].stream == 0)) continue;
        // a stream's segments keep their list order
        bool earlier = false;
        for (int k = 0; k < i; ++k) earlier |= !done[k] && segs[k].stream == segs[i].stream;
        bool ready = !earlier;
        for (int w : segs[i].waits) ready &= recorded[w];
        if (ready) pick = i;
      }
    if (pick < 0) return fail(QM_ERR_STATE, "graph segments: no launchable segment");
    const qm_batch::Segment& sg = segs[pick];
    hipStream_t st = stream_of(sg.stream);
    for (int w : sg.waits) QM_HIP(hipStreamWaitEvent(st, event_of(b, w), 0));
    if (!graph) {
      for (int u : sg.stages) enqueue_stage(b, par, st, nullptr, sg.f0, sg.nf, u);
    } else if (b->seg_exec[par][pick]) {
      QM_HIP(hipGraphLaunch(b->seg_exec[par][pick], st));
    }
    if (sg.record >= 0) {
      QM_HIP(hipEventRecord(event_of(b, sg.record), st));
      recorded[sg.record] = true;
    }
    done[pick] = true;
    --left;
  }
  for (int i = 0; i < ns; ++i)
    if (segs[i].stream != 0 && segs[i].record >= kEvJoin) QM_HIP(hipStreamWaitEvent(s, event_of(b, segs[i].record), 0));
  return QM_OK;
}

// One pipeline run as plain launches (the overlap setting's segments).
int run_pipeline(qm_batch* b, int par, hipStream_t s) {
  return launch_segments(b, par, s, plan_segments(b), false);
}

int qm_batch_run(qm_batch* b, int mode, void* stream) {
  if (!b) return fail(QM_ERR_INVALID, "NULL batch");
  hipStream_t s = (hipStream_t)stream;
  const int par = b->parity;
  if (mode == 0) {
    int rc = run_pipeline(b, par, s);
    if (rc) return rc;
    QM_HIP(hipGetLastError());
  } else if (mode == 1) {
    int rc = capture_segments(b, par);
    if (!rc) rc = launch_segments(b, par, s, b->segs, true);
    if (rc) return rc;
  } else {
    return fail(QM_ERR_INVALID, "mode must be 0 or 1");
  }
  b->parity ^= 1;
  b->runs += 1;
  b->run_state_valid = true;
  return QM_OK;
}

int qm_batch_runs(qm_batch* b, long long* runs) {
  if (!b || !runs) return fail(QM_ERR_INVALID, "NULL pointer");
  *runs = b->runs;
  return QM_OK;
}

int qm_batch_run_timed(qm_batch* b, void* stream, double* stage_ms, int* stage_launches) {
  if (!b || !stage_ms || !stage_launches) return fail(QM_ERR_INVALID, "NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  Timer t;
  t.pool = &b->events;
  t.s = s;
  run_stages(b, b->parity, s, &t);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < QM_NUM_STAGES; ++i) {
    stage_ms[i] = 0.0;
    stage_launches[i] = 0;
  }
  for (size_t k = 0; k < t.stage.size(); ++k) {
    float ms = 0.0f;
    QM_HIP(hipEventElapsedTime(&ms, b->events[2 * k], b->events[2 * k + 1]));
    stage_ms[t.stage[k]] += ms;
    stage_launches[t.stage[k]] += t.launches[k];
  }
  b->parity ^= 1;
  b->runs += 1;
  b->run_state_valid = true;
  return QM_OK;
}

// The warp table and the warped prior seeds of the last run, for a run that
// did not write them (no backward asked for yet): replayed on stream s from
// its prior seeds and its z-buffer, both intact until the next run. Once per
// run; nothing to do for a run that wrote them or after the state changed.
static int replay_fusion(qm_batch* b, hipStream_t s) {
  if (b->keep_fusion || b->runs == 0 || !b->run_state_valid || b->replayed_run == b->runs) return QM_OK;
  const int n = b->n;
  if (!b->dfr) {
    void* p = nullptr;
    const int rc = dev_alloc(&p, sizeof(FrameDev) * n);
    if (rc) return rc;
    b->dfr = (FrameDev*)p;
  }
  std::vector<FrameDev> hr = b->hf[b->parity ^ 1];
  for (int i = 0; i < n; ++i) hr[i].table_out = b->table_buf[i];
  QM_HIP(hipMemcpyAsync(b->dfr, hr.data(), sizeof(FrameDev) * n, hipMemcpyHostToDevice, s));
  launch_fuse_replay(b->dfr, hr.data(), n, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));  // hr goes out of scope
  b->replayed_run = b->runs;
  return QM_OK;
}

int qm_batch_buffer(qm_batch* b, int i, int which, void** ptr) {
  if (!b || !ptr) return fail(QM_ERR_INVALID, "NULL pointer");
  if (i < 0 || i >= b->n) return fail(QM_ERR_INVALID, "frame index out of range");
  if (which == 6 || which == 7) {
    const int rc = replay_fusion(b, nullptr);
    if (rc) return rc;
    QM_HIP(hipStreamSynchronize(nullptr));
  }
  const FrameDev& F = b->hf[b->parity][i];
  switch (which) {
    case 0: *ptr = F.levels; break;
    case 1: *ptr = F.cost[0]; break;
    case 2: *ptr = F.dcoarse; break;
    case 3: *ptr = F.label; break;
    case 4: *ptr = F.dense; break;
    case 5: *ptr = F.seeds; break;
    case 6: *ptr = b->table_buf[i]; break;
    case 7: *ptr = b->hf[b->parity ^ 1][i].seeds; break;
    default: return fail(QM_ERR_INVALID, "unknown buffer id");
  }
  return QM_OK;
}

static int* tune_slot(Tuning& t, int knob) {
  switch (knob) {
    case QM_TUNE_COST_WAVES: return &t.cost_waves;
    case QM_TUNE_BP_ROWS: return &t.bp_rows;
    case QM_TUNE_EXTRACT_ROWS: return &t.extract_rows;
    case QM_TUNE_OVERLAP: return &t.overlap;
    case QM_TUNE_BP_PAIR: return &t.bp_pair;
    case QM_TUNE_BP_SEG: return &t.bp_seg;
    case QM_TUNE_BP_PAIR_LEVELS: return &t.bp_pair_levels;
    default: return nullptr;
  }
}

static void drop_graphs(qm_batch* b) {
  if (b->captured[0] || b->captured[1]) (void)hipDeviceSynchronize();  // no replay may still be running
  for (int par = 0; par < 2; ++par) {
    for (hipGraphExec_t e : b->seg_exec[par])
      if (e) (void)hipGraphExecDestroy(e);
    b->seg_exec[par].clear();
    b->captured[par] = false;
  }
}

static bool knob_value_ok(int knob, int value) {
  switch (knob) {
    case QM_TUNE_COST_WAVES: return value == 2 || value == 4 || value == 8;
    case QM_TUNE_OVERLAP: return value >= 0 && value < kOverlapSettings;
    case QM_TUNE_BP_PAIR: return value == 0 || value == 8 || value == 16;
    case QM_TUNE_BP_SEG: return value == 8 || value == 16 || value == 32 || value == 64;
    case QM_TUNE_BP_PAIR_LEVELS: return value >= 1 && value <= 4;
    default: return value == 4 || value == 8 || value == 16;  // bp_rows, extract_rows
  }
}

// set one knob: the overlap for the batch, a shape knob for group g (every
// group when g < 0); drops the captured graphs on a change
static int set_knob(qm_batch* b, int g, int knob, int value) {
  Tuning probe;
  if (!tune_slot(probe, knob)) return fail(QM_ERR_INVALID, "unknown tuning knob");
  if (!knob_value_ok(knob, value)) return fail(QM_ERR_INVALID, "value not offered for this knob");
  bool changed = false;
  if (knob == QM_TUNE_OVERLAP) {
    changed = b->tune.overlap != value;
    b->tune.overlap = value;
  } else {
    for (int i = 0; i < (int)b->gtune.size(); ++i) {
      if (g >= 0 && i != g) continue;
      int* slot = tune_slot(b->gtune[i], knob);
      changed |= *slot != value;
      *slot = value;
    }
  }
  if (changed) drop_graphs(b);
  return QM_OK;
}

int qm_batch_set_tuning(qm_batch* b, int knob, int value) {
  if (!b) return fail(QM_ERR_INVALID, "NULL batch");
  return set_knob(b, -1, knob, value);
}

int qm_batch_get_tuning(qm_batch* b, int knob, int* value) {
  if (!b || !value) return fail(QM_ERR_INVALID, "NULL pointer");
  int* slot = tune_slot(knob == QM_TUNE_OVERLAP ? b->tune : b->gtune[0], knob);
  if (!slot) return fail(QM_ERR_INVALID, "unknown tuning knob");
  *value = *slot;
  return QM_OK;
}

int qm_batch_groups(qm_batch* b, int* count, int* W, int* H, int cap) {
  if (!b || !count) return fail(QM_ERR_INVALID, "NULL pointer");
  const int G = (int)b->group_w.size();
  *count = G;
  for (int g = 0; g < G && g < cap; ++g) {
    if (W) W[g] = b->group_w[g];
    if (H) H[g] = b->group_h[g];
  }
  return QM_OK;
}

int qm_batch_set_group_tuning(qm_batch* b, int group, int knob, int value) {
  if (!b) return fail(QM_ERR_INVALID, "NULL batch");
  if (group < 0 || group >= (int)b->gtune.size()) return fail(QM_ERR_INVALID, "group index out of range");
  if (knob == QM_TUNE_OVERLAP) return fail(QM_ERR_INVALID, "the overlap knob is the batch's, not a group's");
  return set_knob(b, group, knob, value);
}

int qm_batch_get_group_tuning(qm_batch* b, int group, int knob, int* value) {
  if (!b || !value) return fail(QM_ERR_INVALID, "NULL pointer");
  if (group < 0 || group >= (int)b->gtune.size()) return fail(QM_ERR_INVALID, "group index out of range");
  if (knob == QM_TUNE_OVERLAP) return fail(QM_ERR_INVALID, "the overlap knob is the batch's, not a group's");
  int* slot = tune_slot(b->gtune[group], knob);
  if (!slot) return fail(QM_ERR_INVALID, "unknown tuning knob");
  *value = *slot;
  return QM_OK;
}

int qm_batch_autotune(qm_batch* b, int reps, void* stream, int* chosen, double* best_ms) {
  if (!b) return fail(QM_ERR_INVALID, "NULL batch");
  if (reps < 1 || reps > 1000) return fail(QM_ERR_INVALID, "reps must be in [1, 1000]");
  b->run_state_valid = false;  // its timing runs overwrite the last run's intermediate buffers
  hipStream_t s = (hipStream_t)stream;
  const int par = b->parity;
  const FrameDev* df = b->df[par];
  const FrameDev* hf = b->hf[par].data();
  const int n = b->n, N = b->p.n_hyp;
  const BpLaunchParams bp = bp_params(&b->p);
  // inputs of the timed stages: quadtree, cost and its pyramid (the seeds are
  // untouched; every shape gives the same results)
  launch_quadtree(df, hf, n, s);
  launch_cost(df, hf, n, N, b->table, false, b->p.cost_arith, Tuning(), s, true);
  hipEvent_t e0, e1;
  QM_HIP(hipEventCreate(&e0));
  QM_HIP(hipEventCreate(&e1));
  auto time = [&](const std::function<void(const Tuning&)>& stage, const Tuning& t, double* ms) -> int {
    stage(t);  // warm
    QM_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) stage(t);
    QM_HIP(hipEventRecord(e1, s));
    QM_HIP(hipEventSynchronize(e1));
    float f = 0.0f;
    QM_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = (double)f / reps;
    return QM_OK;
  };
  struct Knob {
    int id;
    std::vector<int> values;
    std::function<void(const Tuning&)> stage;
  };
  int rc = QM_OK;
  const int G = (int)b->gtune.size();
  // the shape knobs searched on the frames of the given runs of consecutive
  // frames (launches of a stage cover one run each); best_ms of the shape
  // knobs when times is given
  auto tune_runs = [&](const std::vector<std::pair<int, int>>& runs, Tuning best, double* times) -> Tuning {
    auto sweeps = [&](const Tuning& t) {
      for (const auto& r : runs) run_sweeps(df + r.first, hf + r.first, r.second, &b->p, t, s);
    };
    // the single half-sweep's shape is timed with the pairs off (it runs
    // alone there, at the levels without pairs, and for a level's odd last sweep)
    auto single_sweeps = [&](const Tuning& t) {
      Tuning u = t;
      u.bp_pair = 0;
      sweeps(u);
    };
    const Knob knobs[3] = {
        {QM_TUNE_COST_WAVES, {2, 4, 8}, [&](const Tuning& t) {
           for (const auto& r : runs)
             launch_cost(df + r.first, hf + r.first, r.second, N, b->table, false, b->p.cost_arith, t, s, true);
         }},
        {QM_TUNE_BP_ROWS, {4, 8, 16}, single_sweeps},
        {QM_TUNE_EXTRACT_ROWS, {4, 8, 16}, [&](const Tuning& t) {
           for (const auto& r : runs) launch_extract(df + r.first, hf + r.first, r.second, N, bp, t, s);
         }},
    };
    for (const Knob& k : knobs) {
      double best_t = 1e30;
      for (int v : k.values) {
        Tuning t = best;
        *tune_slot(t, k.id) = v;
        double ms = 0.0;
        if ((rc = time(k.stage, t, &ms))) return best;
        if (ms < best_t) {
          best_t = ms;
          *tune_slot(best, k.id) = v;
        }
      }
      if (times) times[k.id] = best_t;
    }
    // the fused pairs' three knobs depend on each other: every combination is
    // timed over the whole sweep schedule (with the pairs off, the other two
    // keep their values)
    double best_t = 1e30;
    const Tuning base = best;
    for (int pair : {0, 8, 16})
      for (int seg : {8, 16, 32, 64})
        for (int lv = 1; lv <= 4; ++lv) {
          if (pair == 0 && (seg != base.bp_seg || lv != base.bp_pair_levels)) continue;
          Tuning t = base;
          t.bp_pair = pair;
          t.bp_seg = seg;
          t.bp_pair_levels = lv;
          double ms = 0.0;
          if ((rc = time(sweeps, t, &ms))) return best;
          if (ms < best_t) {
            best_t = ms;
            best.bp_pair = pair;
            best.bp_seg = seg;
            best.bp_pair_levels = lv;
          }
        }
    if (times) times[QM_TUNE_BP_PAIR] = times[QM_TUNE_BP_SEG] = times[QM_TUNE_BP_PAIR_LEVELS] = best_t;
    return best;
  };
  // one shape for every frame (each stage in one launch), searched on the whole batch
  double common_ms[QM_NUM_TUNE] = {};
  const Tuning common = tune_runs({{0, n}}, b->gtune[0], common_ms);
  std::vector<Tuning> gbest(G, common);
  if (G > 1 && !rc) {
    // each resolution group on its own frames (its runs of consecutive frames)
    std::vector<Tuning> own(G);
    for (int g = 0; g < G && !rc; ++g) {
      std::vector<std::pair<int, int>> runs;
      for (int a = 0; a < n;) {
        if (b->group_of[a] != g) {
          ++a;
          continue;
        }
        int k = 1;
        while (a + k < n && b->group_of[a + k] == g) ++k;
        runs.push_back({a, k});
        a += k;
      }
      own[g] = tune_runs(runs, common, nullptr);
    }
    // per stage, the groups' own shapes (a launch per run of agreeing groups)
    // against the common one (one launch), both over the whole batch
    const std::vector<Tuning> one(G, common);
    auto cost_all = [&](const std::vector<Tuning>& gt) {
      for_shape_runs(gt, b->group_of, 0, n, same_cost_shape, [&](int a, int k, const Tuning& tu) {
        launch_cost(df + a, hf + a, k, N, b->table, false, b->p.cost_arith, tu, s, true);
      });
    };
    auto bp_all = [&](const std::vector<Tuning>& gt) {
      for_shape_runs(gt, b->group_of, 0, n, same_bp_shape,
                     [&](int a, int k, const Tuning& tu) { run_sweeps(df + a, hf + a, k, &b->p, tu, s); });
    };
    auto extract_all = [&](const std::vector<Tuning>& gt) {
      for_shape_runs(gt, b->group_of, 0, n, same_extract_shape,
                     [&](int a, int k, const Tuning& tu) { launch_extract(df + a, hf + a, k, N, bp, tu, s); });
    };
    auto faster = [&](const std::function<void(const std::vector<Tuning>&)>& stage, bool* split) -> int {
      double t_own = 0.0, t_one = 0.0;
      int r;
      if ((r = time([&](const Tuning&) { stage(own); }, common, &t_own))) return r;
      if ((r = time([&](const Tuning&) { stage(one); }, common, &t_one))) return r;
      *split = t_own < t_one;
      return QM_OK;
    };
    bool split_cost = false, split_bp = false, split_extract = false;
    if (!rc) rc = faster(cost_all, &split_cost);
    if (!rc) rc = faster(bp_all, &split_bp);
    if (!rc) rc = faster(extract_all, &split_extract);
    for (int g = 0; g < G && !rc; ++g) {
      Tuning& t = gbest[g];
      if (split_cost) t.cost_waves = own[g].cost_waves;
      if (split_bp) {
        t.bp_rows = own[g].bp_rows;
        t.bp_pair = own[g].bp_pair;
        t.bp_seg = own[g].bp_seg;
        t.bp_pair_levels = own[g].bp_pair_levels;
      }
      if (split_extract) t.extract_rows = own[g].extract_rows;
    }
  }
  if (best_ms)
    for (int k = 0; k < QM_NUM_TUNE; ++k)
      if (k != QM_TUNE_OVERLAP) best_ms[k] = common_ms[k];
  // the overlap knob, with the groups' chosen shapes: whole pipeline runs
  // (plain launches), which move the fusion state, so both seed buffers of
  // every frame are saved first and restored afterwards (the runs' depth /
  // dense / sparse outputs are not)
  int best_overlap = b->tune.overlap;
  if (!rc) {
    size_t total = 0;
    for (int i = 0; i < n; ++i) total += 2 * (size_t)hf[i].W * hf[i].H * sizeof(float4);
    void* snap = nullptr;
    rc = dev_alloc(&snap, total);
    if (!rc) {
      auto copy_seeds = [&](bool save) -> int {
        char* p = (char*)snap;
        for (int i = 0; i < n; ++i) {
          const size_t bytes = (size_t)hf[i].W * hf[i].H * sizeof(float4);
          for (int q = 0; q < 2; ++q, p += bytes)
            QM_HIP(hipMemcpyAsync(save ? (void*)p : (void*)b->seeds[q][i], save ? (void*)b->seeds[q][i] : (void*)p,
                                  bytes, hipMemcpyDeviceToDevice, s));
        }
        return QM_OK;
      };
      rc = copy_seeds(true);
      const Tuning keep = b->tune;
      const std::vector<Tuning> keep_g = b->gtune;
      b->gtune = gbest;
      double best_t = 1e30;
      for (int v = 0; v < kOverlapSettings && !rc; ++v) {
        Tuning t = keep;
        t.overlap = v;
        double ms = 0.0;
        rc = time([&](const Tuning& tt) { b->tune = tt; (void)run_pipeline(b, par, s); }, t, &ms);
        if (!rc && ms < best_t) {
          best_t = ms;
          best_overlap = v;
        }
      }
      b->tune = keep;
      b->gtune = keep_g;
      if (!rc) rc = copy_seeds(false);
      // the timing runs all used one parity and left its z-buffer filled:
      // both start empty again
      for (int i = 0; i < n && !rc; ++i)
        for (int q = 0; q < 2 && !rc; ++q) {
          const hipError_t e = hipMemsetAsync(b->zbuf[q][i], 0xFF, (size_t)hf[i].W * hf[i].H * 8, s);
          if (e != hipSuccess) rc = fail(QM_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        }
      if (!rc) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(QM_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
      }
      dev_free(snap);
      if (best_ms) best_ms[QM_TUNE_OVERLAP] = best_t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  QM_HIP(hipGetLastError());
  for (int g = 0; g < G; ++g)
    for (int k = 0; k < QM_NUM_TUNE; ++k)
      if (k != QM_TUNE_OVERLAP && (rc = set_knob(b, g, k, *tune_slot(gbest[g], k)))) return rc;
  if ((rc = set_knob(b, -1, QM_TUNE_OVERLAP, best_overlap))) return rc;
  if (chosen)
    for (int k = 0; k < QM_NUM_TUNE; ++k) chosen[k] = k == QM_TUNE_OVERLAP ? best_overlap : *tune_slot(gbest[0], k);
  return QM_OK;
}

int qm_batch_read_buffer(qm_batch* b, int i, int which, void* dst, size_t bytes) {
  void* src = nullptr;
  int rc = qm_batch_buffer(b, i, which, &src);
  if (rc) return rc;
  if (!dst) return fail(QM_ERR_INVALID, "dst is NULL");
  const FrameDev& F = b->hf[0][i];
  const size_t px = (size_t)F.W * F.H, cc = (size_t)F.Wc * F.Hc;
  const size_t sizes[8] = {px, cc * b->p.n_hyp * 4, cc * 4, cc * 4, px * 4, px * 16, px * 4, px * 16};
  if (bytes != sizes[which]) return fail(QM_ERR_INVALID, "byte count differs from the buffer size");
  QM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
  return QM_OK;
}

int qm_batch_reset_seeds(qm_batch* b) {
  if (!b) return fail(QM_ERR_INVALID, "NULL batch");
  b->run_state_valid = false;
  for (int i = 0; i < b->n; ++i) {
    const size_t px = (size_t)b->hf[0][i].W * b->hf[0][i].H;
    QM_HIP(hipMemset(b->seeds[0][i], 0, px * 16));
    QM_HIP(hipMemset(b->seeds[1][i], 0, px * 16));
  }
  return QM_OK;
}

int qm_batch_destroy(qm_batch* b) {
  if (!b) return QM_OK;
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  for (int g = 0; g < qm_batch::kMaxGroups - 1; ++g) {
    if (b->side[g]) {
      (void)hipStreamSynchronize(b->side[g]);
      (void)hipStreamDestroy(b->side[g]);
    }
    if (b->ev_mark[g]) (void)hipEventDestroy(b->ev_mark[g]);
    if (b->ev_join[g]) (void)hipEventDestroy(b->ev_join[g]);
  }
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_warp) (void)hipEventDestroy(b->ev_warp);
  drop_graphs(b);
  for (int par = 0; par < 2; ++par) dev_free(b->df[par]);
  dev_free(b->dfr);
  dev_free(b->arena);
  dev_free(b->bwd_arena);
  dev_free(b->dfb);
  dev_free(b->partial);
  delete b;
  return QM_OK;
}

int qm_batch_backward(qm_batch* b, const float* const* g_depth_out, double* grad_lambda, void* stream) {
  if (!b || !grad_lambda) return fail(QM_ERR_INVALID, "NULL pointer");
  if (b->runs == 0) return fail(QM_ERR_STATE, "no forward run to differentiate");
  if (b->p.upsample_solver != QM_UPSAMPLE_THOMAS)
    return fail(QM_ERR_INVALID, "the batch backward differentiates the Thomas upsample; create the batch with "
                                "upsample_solver = QM_UPSAMPLE_THOMAS");
  if (!b->run_state_valid)
    return fail(QM_ERR_STATE,
                "the batch changed after its last run (set_frame, reset_seeds or autotune); run it again before "
                "the backward");
  hipStream_t s = (hipStream_t)stream;
  const int n = b->n, last = b->parity ^ 1;
  int rc;
  if (!b->bwd_arena) {
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
      const FrameDev& F = b->hf[0][i];
      const size_t px = (size_t)F.W * F.H, er = (size_t)F.H * (F.W + 1), ec = (size_t)(F.H + 1) * F.W;
      total += 2 * align256(px * 4) + align256(er * 4) + align256(ec * 4) + 2 * align256(px * 8) +
               align256(sizeof(double) * (F.W + F.H));
    }
    void *arena = nullptr, *dfb = nullptr, *partial = nullptr;
    if ((rc = dev_alloc(&arena, total)) || (rc = dev_alloc(&dfb, sizeof(FrameDev) * n)) ||
        (rc = dev_alloc(&partial, sizeof(double) * n))) {
      dev_free(arena);
      dev_free(dfb);
      dev_free(partial);
      return rc;
    }
    b->bwd_arena = arena;
    b->dfb = (FrameDev*)dfb;
    b->partial = (double*)partial;  // dL/dlambda of each frame
    b->partial_n = n;
    Carver cv{(char*)arena};
    b->bwd.resize(n);
    for (int i = 0; i < n; ++i) {
      const FrameDev& F = b->hf[0][i];
      const size_t px = (size_t)F.W * F.H, er = (size_t)F.H * (F.W + 1), ec = (size_t)(F.H + 1) * F.W;
      qm_batch::Bwd& w = b->bwd[i];
      w.sparse = cv.take<float>(px * 4);
      w.gdense = cv.take<float>(px * 4);
      w.wrow = cv.take<float>(er * 4);
      w.wcol = cv.take<float>(ec * 4);
      w.colxn = cv.take<float2>(px * 8);
      w.adjxn = cv.take<float2>(px * 8);
      w.lam_part = cv.take<double>(sizeof(double) * (F.W + F.H));
    }
  }
  // the run's warp table and warped prior seeds, which the fusion adjoint reads
  if ((rc = replay_fusion(b, s))) return rc;
  std::vector<FrameDev> hb = b->hf[last];
  for (int i = 0; i < n; ++i) {
    FrameDev& F = hb[i];
    const qm_batch::Bwd& w = b->bwd[i];
    F.table_in = b->table_buf[i];
    F.table_out = nullptr;
    F.depth_out = nullptr;
    F.dense_out = nullptr;
    F.sparse_out = nullptr;
    F.g_depth_out = g_depth_out ? g_depth_out[i] : nullptr;
    F.g_seeds_out = nullptr;
    F.grad_depth_in = w.gdense;
    F.g_d = w.sparse;
    F.g_conf = nullptr;
    F.g_wrow = w.wrow;
    F.g_wcol = w.wcol;
    F.g_out = w.gdense;
    F.grad_d = nullptr;
    F.grad_conf = nullptr;
    F.grad_wrow = nullptr;  // only dL/dlambda is wanted: the adjoint sums w dL/d(lambda w) per line
    F.grad_wcol = nullptr;
    F.grad_w_unscaled = 1;
    F.lam_part = w.lam_part;
    F.grad_lambda = b->partial + i;
    F.colxn = w.colxn;
    F.adjxn = w.adjxn;
    F.row_x_out = nullptr;
    F.row_n_out = nullptr;
  }
  const bool kept = b->colxn_since >= 0 && b->runs > b->colxn_since;
  QM_HIP(hipMemcpyAsync(b->dfb, hb.data(), sizeof(FrameDev) * n, hipMemcpyHostToDevice, s));
  // fusion adjoint first: it reads the forward's dense depth, which an
  // upsample recomputation then rewrites with the same values
  launch_fuse_backward(b->dfb, hb.data(), n, s);
  // the sparse map is read only by a recomputed forward (the adjoint wants no dL/dd here)
  launch_compose_inputs(b->dfb, hb.data(), n, b->wlut, s, !kept);
  launch_upsample_backward(b->dfb, hb.data(), n, s, !kept);
  QM_HIP(hipGetLastError());
  QM_HIP(hipMemcpyAsync(grad_lambda, b->partial, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  QM_HIP(hipStreamSynchronize(s));
  if (b->colxn_since < 0) {
    // later runs keep their column results, their warp tables and their
    // warped prior seeds for their backward (a few MB of writes per run)
    for (int par = 0; par < 2; ++par)
      for (int i = 0; i < n; ++i) {
        b->hf[par][i].colxn = b->bwd[i].colxn;
        b->hf[par][i].table_out = b->table_buf[i];
        b->hf[par][i].warp_skip = 0;
      }
    if ((rc = upload_frames(b))) return rc;
    b->colxn_since = b->runs;
    b->keep_fusion = true;
  }
  return QM_OK;
}

}  // extern "C"
