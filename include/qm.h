/*
 * qm.h -- C ABI of the MI355X per-keyframe depth pipeline (libqm.so).
 *
 * Drop-in for the hot path of QuadtreeMapping (quadmap::Depthmap ->
 * SeedMatrix::extract_depth + fuse_output_depth). Plain C types only; every
 * device buffer is a HIP device pointer; `stream` is a hipStream_t passed as
 * void* (NULL = the null stream). Every call returns 0 on success or a negative
 * QM_ERR_* code; qm_last_error() gives a thread-local message.
 *
 * Layouts (all row-major, no padding unless a pitch argument says otherwise):
 *   image        uint8  H x W            (pitch in bytes, W <= pitch <= 262144)
 *   levels       uint8  H x W            quadtree level per pixel, in {2,3,4}
 *   cost         float  (H/4) x (W/4) x N
 *   depth_coarse float  (H/4) x (W/4)    BP winner depth per node cell, 0 elsewhere
 *   label        int32  (H/4) x (W/4)    BP winner index, -1 for non-node cells
 *   sparse/dense float  H x W
 *   seeds        float4 H x W            (a, b, mu, sigma^2) fusion state
 *   transforms   float[12] 3x4 row-major [R | t]
 */
#ifndef QM_H
#define QM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QM_OK 0
#define QM_ERR_INVALID (-1) /* bad argument (size, null pointer, unsupported N) */
#define QM_ERR_HIP (-2)     /* a HIP runtime call failed */
#define QM_ERR_ALLOC (-3)   /* device allocation failed or was refused */
#define QM_ERR_STATE (-4)   /* call out of order */

/* Pipeline constants. Defaults are the reference's compile-time values:
 * include/quadmap/stereo_parameter.cuh:5-7,41-42, include/quadmap/pixel_cost.cuh:4,
 * src/bpextract.cu:45,222-223, src/seed_matrix.cu:151,182,208. */
typedef struct qm_params {
  int n_hyp;            /* inverse-depth samples: 64 (DEPTH_NUM) or 128 */
  double idepth_min;    /* MIN_INV_DEPTH = 0.01 */
  double idepth_max;    /* MAX_INV_DEPTH = 2.0; sample step = (max-min)/(n_hyp-1) */
  float p1;             /* BP penalty for a +-1 label change, 0.003 */
  float p2;             /* BP penalty for any label change, 0.01 */
  int bp_iters[4];      /* half-sweeps per level, fine -> coarse: {4,10,10,10} */
  double lambda;        /* upsample smoothness, upsample_lambda = 10 */
  int min_refs;         /* references needed before a depth map is produced: 11 (> 10) */
  int max_refs;         /* reference deque length, KEYFRAME_NUM = 60 */
  float ref_cos;        /* add a reference if cos(z, z_ref) < 0.95 ... */
  float ref_baseline;   /* ... or the baseline exceeds 0.03 m */
  int cost_arith;       /* projection arithmetic of the cost volume, QM_COST_*: both give
                           the same algebra, and both are bit-identical to the oracle's
                           matching mode; default QM_COST_REFERENCE */
  int upsample_solver;  /* line solver of the global upsample, QM_UPSAMPLE_*: both solve
                           the same systems, and both are bit-identical to the oracle's
                           matching mode; default QM_UPSAMPLE_THOMAS */
} qm_params;

/* Cost-volume projection arithmetic (qm_params.cost_arith).
 * QM_COST_REFERENCE: the reference's operation order
 *   (src/depth_extract.cu:178-197): P = dir * depth, q = R P + t with every
 *   product rounded, px = (fx q.x) / q.z + cx by IEEE division, texel
 *   (int)(px + i u2u + j v2u + 0.5).
 * QM_COST_FUSED: the same projection with the per-cell products hoisted out of
 *   the sample loop and fused multiply-adds: a = (fx R0.dir, fy R1.dir, R2.dir),
 *   U = fma(depth, a, (fx t0, fy t1, t2)), r = 1 / U.z (correctly rounded),
 *   px + 0.5 = fma(U.x, r, cx + 0.5), texel floor(px + 0.5 + i u2u + j v2u).
 *   4 roundings instead of 12 on the way to each pixel; about a third fewer
 *   instructions per (cell, sample). */
#define QM_COST_REFERENCE 0
#define QM_COST_FUSED 1

/* Line solver of the global upsample (qm_params.upsample_solver), for the
 * row and column systems (I + lambda L(w)) x = r of src/depth_upsample.cu:94-302.
 * QM_UPSAMPLE_THOMAS: the reference's Thomas recurrences in its order, one
 *   serial chain per line.
 * QM_UPSAMPLE_SCAN: the same recurrences by segmented scans: 32 lanes per
 *   line (lines of up to 512 steps) or 64 lanes (up to 1536 steps) each own
 *   a run of steps; the maps of the three recurrences (a
 *   Moebius map for the coefficients, affine maps for the right-hand sides
 *   and the back substitution) are composed per lane and scanned over the
 *   lanes in double, and each lane then runs its steps in float from its
 *   incoming state. As accurate as the Thomas order against FP64; a
 *   different rounding. Lines longer than 1536 steps, and frames with lambda
 *   above 2^100, use the Thomas solve.
 *   Applies to the pipeline and qm_upsample; the differentiable entry points
 *   (qm_upsample_general*, qm_upsample_backward*, qm_batch_backward) use the
 *   Thomas solve and qm_batch_backward refuses a batch run with the scan. */
#define QM_UPSAMPLE_THOMAS 0
#define QM_UPSAMPLE_SCAN 1

/* Version of this ABI; bindings check it at load time so that a stale library
 * fails loudly instead of being called with the wrong argument list. */
#define QM_ABI_VERSION 13
int qm_abi_version(void);
const char* qm_last_error(void);
void qm_default_params(qm_params* p);
/* Fault injection for tests: refuse device allocations beyond `bytes` (0 = no cap). */
void qm_set_alloc_cap(size_t bytes);
/* Device bytes the library holds. This includes the scratch of the stage-level
 * calls below, which is kept for reuse by later calls: at most 1 GiB over all
 * devices, the oldest blocks freed first beyond that, everything freed when a
 * device allocation of the library is refused or qm_release_scratch is called. */
size_t qm_device_bytes_in_use(void);
/* Free the kept stage-level scratch now (for a caller whose own allocator ran
 * out of memory, or before teardown). Safe to call at any time between calls. */
int qm_release_scratch(void);

/* ---------------- stage level (device pointers) ---------------------------- */

/* Quadtree level assignment. Replaces quadtree_image_kernal
 * (src/quadtree.cu:17-77, launched at src/seed_matrix.cu:479). */
int qm_quadtree(const uint8_t* img, int W, int H, int pitch, uint8_t* levels, void* stream);

/* Multi-view patch-cost volume over n_hyp inverse-depth samples. refs[0..3] are
 * the four distinct references read by slots {0,0,1,2,3}; T_ref_key[4][12] is
 * the host array of key->reference transforms; cam = {fx, fy, cx, cy}.
 * Costs are the reference's normalised ones: (SAD sum / views / 255, or 100
 * where no reference sees the sample) / 64, the 64 being normalize_the_cost's
 * add_num, which every sample thread of a node increments once.
 * count (nullable) receives that add_num: 64 for node cells, 0 otherwise. Replaces
 * image_to_cost + normalize_the_cost (src/depth_extract.cu:135-244, launched at
 * src/seed_matrix.cu:508,523). */
int qm_cost_volume(const uint8_t* key, const uint8_t* const refs[4], int W, int H, int pitch,
                   const float T_ref_key[48], const float cam[4], const qm_params* p,
                   const uint8_t* levels, float* cost, int32_t* count, void* stream);

/* Hierarchical BP (cost pyramid, 4 levels, checkerboard min-sum) and winner
 * extraction. Replaces bp_extract (src/bpextract.cu:37-167) and its kernels
 * cost_distribute, bp, upsample, depth_extract (:169-395). messages (nullable)
 * receives the level-0 messages, 4 x (H/4) x (W/4) x N in the order up, down,
 * left, right (the reference's up/dm/lm/rm maps). */
int qm_bp_extract(const float* cost, int W, int H, const uint8_t* levels, const qm_params* p,
                  float* depth_coarse, int32_t* label, float* messages, void* stream);

/* Tridiagonal smoothness upsampling of a sparse depth map, weights from the
 * keyframe. Replaces global_upsample (src/depth_upsample.cu:20-302). */
int qm_upsample(const float* sparse, const uint8_t* img, int W, int H, int pitch, const qm_params* p,
                float* dense, void* stream);

/* Generalised upsampling used by the differentiable path: rhs_x = conf*d,
 * rhs_n = conf (conf == NULL: rhs_x = d, rhs_n = 1{d > 0} as the reference);
 * w_row is H x (W+1), w_col is (H+1) x W; the chains are split at row_seg and
 * col_seg (pass W and H for no split). row_x/row_n (nullable) receive the row
 * pass result. Same arithmetic as qm_upsample. */
int qm_upsample_general(const float* d, const float* conf, const float* w_row, const float* w_col, int W,
                        int H, double lambda, int row_seg, int col_seg, float* dense, float* row_x,
                        float* row_n, void* stream);

/* Adjoint of qm_upsample_general: given g = dL/d(dense), writes dL/dd,
 * dL/dconf (zeros when conf is NULL), dL/dw_row (H x (W+1)) and dL/dw_col
 * ((H+1) x W); any gradient output may be NULL. The forward passes are
 * recomputed, not stored; A = I + lambda L(w) is symmetric, so each adjoint is a
 * tridiagonal solve with the forward matrices (run in the forward passes' wave
 * pipeline). Pixels where the forward denominator n is 0 (output NaN) contribute
 * no gradient. The batched form can also return dL/dlambda (grad_lambda of a
 * problem). No counterpart in the reference (its upsample is forward-only). */
int qm_upsample_backward(const float* d, const float* conf, const float* w_row, const float* w_col, int W, int H,
                         double lambda, int row_seg, int col_seg, const float* g, float* grad_d, float* grad_conf,
                         float* grad_w_row, float* grad_w_col, void* stream);

/* Many generalised upsample problems (uneven sizes allowed) in one launch per
 * pass: forward as qm_upsample_general and adjoint as qm_upsample_backward,
 * per problem. Each problem names its own lambda, chain splits, inputs and
 * outputs; the backward fields (g, grad_*) are read only by the backward call,
 * the forward outputs (dense, row_x, row_n) only by the forward call. At most
 * 128 problems per call. No counterpart in the reference. */
typedef struct qm_upsample_problem {
  int W, H;
  int row_seg, col_seg;        /* chain splits; the reference's are 16*floor(W/16), 8*floor(H/8) */
  double lambda;
  const float* d;              /* H x W */
  const float* conf;           /* H x W or NULL (then rhs = d, indicator = d > 0) */
  const float* w_row;          /* H x (W+1) */
  const float* w_col;          /* (H+1) x W */
  float* dense;                /* forward output, H x W */
  float* row_x;                /* optional forward output of the row pass, H x W */
  float* row_n;
  const float* g;              /* backward input dL/d(dense), H x W */
  float* grad_d;               /* backward outputs, each may be NULL */
  float* grad_conf;
  float* grad_w_row;
  float* grad_w_col;
  double* grad_lambda;         /* backward output dL/dlambda: one double in device memory, or NULL */
  float* state;                /* optional, 4 x H x W floats in device memory: the forward keeps its row
                                  and column pass results there, and a backward given the state its
                                  forward wrote (same inputs) skips recomputing them; every problem of a
                                  backward batch has one, or none does */
} qm_upsample_problem;
int qm_upsample_general_batch(const qm_upsample_problem* problems, int n, void* stream);
int qm_upsample_backward_batch(const qm_upsample_problem* problems, int n, void* stream);

/* Semi-dense keyframe seeds: the LSD-style per-pixel inverse-depth filter that
 * pre-fills the sparse map every frame (src/lsd_semidense.cu, driven by
 * SeedMatrix::initial_keyframe / update_keyframe / create_new_keyframe,
 * src/seed_matrix.cu:211-304,399-446). Device buffers: grad W x H x 2 floats
 * (central differences, 0 on the border), seeds W x H x 8 floats (valid,
 * blacklist, validity, 0, idepth, variance, smoothed idepth, smoothed
 * variance). qm_sd_gradient replaces generate_gradient (src/seed_init.cu:13-42);
 * qm_sd_init replaces initialize_keyframe_kernel of the first frame
 * (src/lsd_semidense.cu:64-81); qm_sd_step runs one frame: the epipolar
 * update of the keyframe seeds from the income frame, hole filling,
 * regularisation and the projection into prefill (H x W, 0 where no seed
 * lands), then, when switch_keyframe is set, the propagation to the income
 * frame as the new keyframe (the caller then makes the income frame the
 * keyframe). Poses are world -> camera. */
int qm_sd_gradient(const uint8_t* img, int W, int H, int pitch, float* grad, void* stream);
int qm_sd_init(const float* grad, int W, int H, float* seeds, void* stream);
int qm_sd_step(float* seeds, const uint8_t* key, const float* key_grad, const uint8_t* inc, const float* inc_grad,
               int W, int H, int pitch, const float cam[4], const float T_key_world[12],
               const float T_inc_world[12], int switch_keyframe, float* prefill, float* prior, void* stream);
/* prior (nullable, H x W x 2) of qm_sd_step receives the (inverse depth,
 * variance) map of depth_project_kernel's float2 overload
 * (src/lsd_semidense.cu:241-294; seeds with validity >= 10, variance scaled
 * by (depth / z)^2 and held at >= 0.01, nearest depth per pixel, 0 elsewhere),
 * the input of the prior-driven alternates. The reference returns when both
 * values are finite (:290-291) and so never stores one; the map here keeps the
 * finite values, as that code intends. */

/* The reference's prior-driven alternates (not on its live path).
 * qm_quadtree_prior replaces quadtree_depth_kernal (src/quadtree.cu:79-154; no
 * launch in the reference): per 16 x 16 tile whose pixels all have a prior
 * depth > 0, a pixel's level becomes max(level, i) for the largest 2^i block
 * (i <= 4) whose four children qualified and whose inverse depths all lie
 * within 0.01 of the block mean. levels is read and written in place.
 * qm_prior_to_cost replaces prior_to_cost (src/depth_extract.cu:23-47, launch
 * commented out at src/seed_matrix.cu:492-499) on a cost volume from
 * qm_cost_volume (same levels and params): each cell gets the truncated
 * quadratic cost min(0.1 (s_d - p)^2 / v, 0.5) of every prior pixel (p, v) of
 * its 4 x 4 block (3 <= x <= W-4, 3 <= y <= H-4, not p <= 0, not v <= 0),
 * renormalised with the image cost as the reference's add_num does:
 * (sum + 64 cost) / (64 (P + node)). count (nullable) receives that add_num.
 * Device pointers; prior is H x W x 2 floats. */
int qm_quadtree_prior(const float* prior_depth, int W, int H, uint8_t* levels, void* stream);
int qm_prior_to_cost(const float* prior, int W, int H, const uint8_t* levels, const qm_params* p, float* cost,
                     int32_t* count, void* stream);

/* The reference's local alternate to the global upsample, local_upsample /
 * depth_interpolate (src/depth_upsample.cu:304-370; not on its live path):
 * each pixel averages the positive sparse depths of the 7 x 7 node grid (nodes
 * every 4 pixels) around it, weighted by exp(-(dx^2 + dy^2)/100 - dI^2/50);
 * NaN where no node contributes. Device pointers. */
int qm_upsample_local(const float* sparse, const uint8_t* img, int W, int H, int pitch, float* dense, void* stream);

/* World-frame point cloud of a depth map, replacing the CPU loop of
 * Publisher::publishPointCloud (src/publisher.cpp:129-164): in raster order,
 * every pixel whose depth d is not below 0.1 (NaN kept, as there) gives
 * (x, y, z, intensity) = (T_world_ref * (normalize(K^-1 (x, y, 1)) * d),
 * ref_img(x, y)). depth, ref_img, xyzi are device pointers; xyzi must hold
 * W * H * 4 floats; *count (host) receives the number of points. */
int qm_point_cloud(const float* depth, const uint8_t* ref_img, int W, int H, const float cam[4],
                   const float T_world_ref[12], float* xyzi, int* count, void* stream);

/* Gaussian x Beta fusion of a dense depth map into the seed state. seeds
 * (H x W float4) is warped in place by T_last_to_cur; the update goes to
 * seeds_out; depth_out gets mu' where the prior is confident, else -1.
 * table (nullable) receives the hole-filled warp table. Replaces fuse_transform,
 * hole_filling, fuse_currentmap (src/depth_fusion.cu:48-213) as driven by
 * SeedMatrix::fuse_output_depth (src/seed_matrix.cu:553-597). */
int qm_fuse(float* seeds, const float* depth_in, int W, int H, const float cam[4], const float T_last_to_cur[12],
            float* depth_out, float* seeds_out, int32_t* table, void* stream);

/* qm_fuse without any device allocation. No counterpart in the reference: it
 * is for callers that cannot allocate (tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, results and synchronisation as qm_fuse;
 * its scratch (descriptor and z-buffer) is one static device array of
 * 16 MiB + 4 KiB per GPU, held until the process ends, so W * H is at most
 * 2097152 pixels (QM_ERR_INVALID beyond). Calls on one GPU share that array:
 * calls that overlap on one GPU (other host threads, other streams) race. */
int qm_fuse_noalloc(float* seeds, const float* depth_in, int W, int H, const float cam[4],
                    const float T_last_to_cur[12], float* depth_out, float* seeds_out, int32_t* table, void* stream);

/* Adjoint of the per-pixel fusion update (src/depth_fusion.cu:146-213) with
 * respect to its dense depth input, for the warped seed state and hole-filled
 * table qm_fuse produced: grad = g_depth_out * d(depth_out)/d(depth) +
 * g_seeds_out . d(seeds_out)/d(depth) (either gradient may be NULL, not both).
 * The prior seed is held fixed. No counterpart in the reference. */
int qm_fuse_backward(const float* seeds_warped, const int32_t* table, const float* depth_in, int W, int H,
                     const float* g_depth_out, const float* g_seeds_out, float* grad_depth_in, void* stream);

/* ---------------- batched keyframes ---------------------------------------- */

/* One keyframe of one camera stream. Input pointers are device pointers owned by
 * the caller; NULL outputs are not written. */
typedef struct qm_frame_desc {
  int W, H, pitch;
  float cam[4];                /* fx, fy, cx, cy */
  const uint8_t* key;          /* keyframe image */
  const uint8_t* refs[4];      /* references for slots {0,0,1,2,3} */
  float T_ref_key[4][12];      /* key -> reference */
  float T_last_to_cur[12];     /* previous fused frame -> this keyframe */
  const float* prefill;        /* semi-dense pre-fill of the sparse map (H x W) or NULL */
  float* depth_out;            /* fused depth (H x W), -1 where not confident */
  float* dense_out;            /* upsampled depth before fusion (H x W) */
  float* sparse_out;           /* sparse map handed to the upsample (H x W) */
  double lambda;               /* this stream's upsample smoothness weight; <= 0: qm_params.lambda */
  /* the reference's prior-driven alternates (see qm_prior_to_cost and
   * qm_quadtree_prior), off when NULL: */
  const float* prior;          /* H x W x 2 (inverse depth, variance) prior added to the cost volume */
  const float* prior_depth;    /* H x W prior depth that can raise the quadtree level */
} qm_frame_desc;

typedef struct qm_batch qm_batch;

/* Plan a batch: allocates all per-stream state (fusion seeds persist across
 * runs), builds the per-stage work tables. */
int qm_batch_create(const qm_frame_desc* frames, int n, const qm_params* p, qm_batch** out);
/* Replace frame i's inputs/transforms (same size); takes effect on the next run. */
int qm_batch_set_frame(qm_batch* b, int i, const qm_frame_desc* frame);
/* Run every stage for every frame. mode 0: one launch per stage over all
 * frames; mode 1: the same captured once as a hipGraph and replayed. */
int qm_batch_run(qm_batch* b, int mode, void* stream);
/* Stage kinds for qm_batch_run_timed. */
#define QM_STAGE_QUADTREE 0
#define QM_STAGE_COST 1
#define QM_STAGE_PYRAMID 2
#define QM_STAGE_BP_SWEEP 3
#define QM_STAGE_EXTRACT 4
#define QM_STAGE_UPSAMPLE_ROW 5
#define QM_STAGE_UPSAMPLE_COL 6
#define QM_STAGE_FUSE_WARP 7
#define QM_STAGE_FUSE_UPDATE 8
#define QM_NUM_STAGES 9
/* One plain single-stream run with a hipEvent pair around each stage (around
 * every launch, except that the BP half-sweeps of the run share one pair);
 * returns per stage kind the summed time (ms) and the number of launches,
 * after synchronising the stream. For measurement only. */
int qm_batch_run_timed(qm_batch* b, void* stream, double stage_ms[QM_NUM_STAGES], int stage_launches[QM_NUM_STAGES]);
/* Launch-shape knobs of a batch; every value gives bit-identical results.
 * Changing one drops the batch's captured graphs (recaptured on the next
 * graph run). Defaults: 4, 16, 16, 0, 8, 8, 2. */
#define QM_TUNE_COST_WAVES 0   /* 2, 4 or 8 waves per cost-volume block of 64 cells */
#define QM_TUNE_BP_ROWS 1      /* 4, 8 or 16 cells (16-lane rows) per BP half-sweep block */
#define QM_TUNE_EXTRACT_ROWS 2 /* 4, 8 or 16 cells per winner-extraction block */
#define QM_TUNE_OVERLAP 3      /* 0: one stream; 1, 2, 3: the batch as two frame groups on two
                                  streams, the second starting with the first, after its cost
                                  volume, or after its BP; 4, 5, 6: three, four or eight
                                  groups, each starting after the previous group's cost volume;
                                  7: four groups, each after the previous group's BP; 8, 9, 10:
                                  two, three or four groups whose cost volumes run one after
                                  the other on the caller's stream, each group's later stages
                                  on a stream of its own; 11, 12, 13: 8, 9, 10 with the fusion
                                  warp of the whole batch run first, beside the cost volumes
                                  (results identical in every case) */
#define QM_TUNE_BP_PAIR 4      /* 0: one launch per BP half-sweep; 8 or 16: the half-sweeps of a
                                  level in fused pairs, one launch per pair, over strips of 14 or
                                  30 cell columns (the first half-sweep's messages pass to the
                                  second through LDS) */
#define QM_TUNE_BP_SEG 5       /* 8, 16, 32 or 64 cell rows per strip segment of a fused pair */
#define QM_TUNE_BP_PAIR_LEVELS 6 /* 1-4: the fused pairs run at levels 0 .. value - 1 (the coarser
                                    levels, with few cells, keep one launch per half-sweep) */
#define QM_NUM_TUNE 7
int qm_batch_set_tuning(qm_batch* b, int knob, int value);
int qm_batch_get_tuning(qm_batch* b, int knob, int* value);
/* Resolution groups: the frames of one size (W x H) share their launch shapes;
 * a stage's launch covers consecutive frames whose groups agree on its shapes.
 * qm_batch_groups gives the number of groups (in the order of their first
 * frame) and, up to cap, their sizes (W, H nullable). qm_batch_set_tuning sets
 * a shape knob for every group (get: group 0's); the group calls set and get
 * one group's (not the overlap knob, which is the batch's). */
int qm_batch_groups(qm_batch* b, int* count, int* W, int* H, int cap);
int qm_batch_set_group_tuning(qm_batch* b, int group, int knob, int value);
int qm_batch_get_group_tuning(qm_batch* b, int group, int knob, int* value);
/* Autotuner, on this batch's own frames: times every value of every shape
 * knob (reps launches each, after one warm-up; the BP knobs jointly, over the
 * whole coarse-to-fine sweep schedule) and keeps the fastest, first with one
 * shape for the whole batch, then (with two or more resolution groups) for
 * each group on its own frames; per stage, the groups' own shapes are kept
 * when the whole batch runs that stage faster with them (a launch per run of
 * agreeing groups) than with the common shape (one launch). Then the overlap
 * knob over whole runs (plain launches) with those shapes. It first
 * recomputes the quadtree,
 * cost volume and pyramid; the fusion state is not touched. chosen receives
 * group 0's shapes and the overlap; best_ms the common search's time per
 * launch set for each shape knob and the overlap's per run (nullable both);
 * qm_batch_get_group_tuning gives every group's. Synchronises the stream. */
int qm_batch_autotune(qm_batch* b, int reps, void* stream, int chosen[QM_NUM_TUNE], double best_ms[QM_NUM_TUNE]);
/* Backward of the last run with respect to each stream's lambda: given
 * g_depth_out[i] = dL/d(depth_out) of frame i (device pointer, H x W; NULL
 * array or entry = zero), writes grad_lambda[i] = dL/d(lambda_i) (host array).
 * Goes through the fusion update adjoint (prior seeds and warp table held
 * fixed) and the upsample adjoint (its forward recomputed, not stored), one
 * launch per stage over all frames, then synchronises the stream. Scratch is
 * allocated on the first call. QM_ERR_STATE before the first run and after a
 * qm_batch_set_frame, qm_batch_reset_seeds or qm_batch_autotune that followed
 * the last run (they change what that run's state the backward reads). A
 * later run replaces the state: a caller that keeps a run for a later
 * backward compares qm_batch_runs. No counterpart in the reference. */
int qm_batch_backward(qm_batch* b, const float* const* g_depth_out, double* grad_lambda, void* stream);
/* Number of runs (qm_batch_run, qm_batch_run_timed) since qm_batch_create. */
int qm_batch_runs(qm_batch* b, long long* runs);
/* Device pointer of an internal buffer of frame i, for inspection:
 * 0 levels, 1 cost (level 0), 2 depth_coarse, 3 label, 4 dense, 5 seeds (current state),
 * 6 hole-filled warp table of the last run (int32), 7 warped prior seeds of the last run.
 * A run writes 6 and 7 only once the batch has had a backward; before that
 * asking for either replays them from the last run's prior seeds and z-buffer
 * (a kernel on the null stream, then a synchronisation, once per run), which
 * is possible until the next run or a change of the batch.
 * The internal cost volume keeps each cell's row of N samples in the row layout:
 * sample d = j + 16 s at position (N/16) * j + s (j < 16). */
int qm_batch_buffer(qm_batch* b, int i, int which, void** ptr);
/* Copy that buffer (its full size, which must equal bytes) to dst, host or
 * device memory, synchronously. */
int qm_batch_read_buffer(qm_batch* b, int i, int which, void* dst, size_t bytes);
int qm_batch_reset_seeds(qm_batch* b);
int qm_batch_destroy(qm_batch* b);

/* ---------------- Depthmap-level state machine ----------------------------- */

typedef struct qm_ctx qm_ctx;

/* quadmap::Depthmap::Depthmap (src/depthmap.cpp:19-40). The reference
 * constructor's argument order is fx, cx, fy, cy; this one takes fx, fy, cx, cy
 * (the C++ wrapper include/quadmap/depthmap.h keeps the reference order).
 * p == NULL uses qm_default_params. */
int qm_create(int W, int H, float fx, float fy, float cx, float cy, const qm_params* p, qm_ctx** out);
int qm_destroy(qm_ctx* c);
/* SeedMatrix::set_semi2dense_ratio (include/quadmap/seed_matrix.cuh:209). */
int qm_set_semi2dense_ratio(qm_ctx* c, int ratio);
/* Semi-dense pre-fill of the sparse map (on by default, as the reference);
 * off leaves the pre-fill zero. */
int qm_set_semidense(qm_ctx* c, int on);
/* SeedMatrix::set_remap (src/seed_matrix.cu:75-80) plus the cv::remap(INTER_LINEAR,
 * BORDER_CONSTANT) of input_raw (:85), done on the GPU: map_x, map_y are host
 * H x W float coordinates into the raw camera image (the CV_32FC1 maps of
 * cv::initUndistortRectifyMap, src/depthmap_node.cpp:84-93, which may also
 * resize). NULL maps = identity (the raw image must then be W x H). */
int qm_set_remap(qm_ctx* c, const float* map_x, const float* map_y);
/* Depthmap::add_frames (src/depthmap.cpp:42-55 -> src/seed_matrix.cu:106-169):
 * host uint8 raw image of img_w x img_h with row pitch in bytes, T_curr_world
 * 3x4 (world->camera). *has_result = 1 when a new result is available (the
 * first frame always reports one, with zero maps). */
int qm_add_frame(qm_ctx* c, const uint8_t* img, int img_w, int img_h, int pitch, const float T_curr_world[12],
                 int* has_result);
/* Host copies (H x W) of the last result: getDepthmap (fused depth, -1 where not
 * confident), getDebugmap (sparse map handed to the upsample), getReferenceImage
 * (the remapped uint8 frame), getT_world_ref (include/quadmap/depthmap.h:49-75). */
int qm_get_depth(qm_ctx* c, float* out);
int qm_get_debug(qm_ctx* c, float* out);
int qm_get_reference(qm_ctx* c, uint8_t* out);
int qm_get_T_world_ref(qm_ctx* c, float out[12]);
/* The point cloud of the last result (as qm_point_cloud with the context's
 * depth map, reference image, intrinsics and T_world_ref), computed on the
 * GPU; xyzi is host memory for W * H * 4 floats. */
int qm_get_point_cloud(qm_ctx* c, float* xyzi, int* count);
/* Switch the prior-driven alternates on for a context's keyframes (before the
 * first frame; they need the semi-dense module): prior_cost adds the
 * semi-dense (inverse depth, variance) prior of the income frame to the cost
 * volume (qm_prior_to_cost), prior_quadtree raises quadtree levels from the
 * semi-dense pre-fill depth (qm_quadtree_prior). No counterpart in the
 * reference, whose launches of both are absent or commented out. */
int qm_set_prior_alternates(qm_ctx* c, int prior_cost, int prior_quadtree);
/* Number of reference frames held (the reference's framelist_host.size()). */
int qm_num_references(qm_ctx* c);

#ifdef __cplusplus
}
#endif

#endif /* QM_H */
