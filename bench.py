"""Benchmark: batched per-keyframe depth maps/s at 752x480 with 64 hypotheses.

One step = every stage (quadtree, cost volume, hierarchical BP + extraction,
tridiagonal upsample, fusion) for a batch of B camera streams of 752x480
keyframes (B = --streams, default 32), one launch per stage over all streams
of a frame group (the autotuner picks how many groups run side by side on
their own streams), replayed as one hipGraph. Inputs are resident in HBM
before timing starts; the fusion state carries over from step to step as for
consecutive keyframes.

Also measured, outside the timed region:
  * the plain per-frame GPU path (one keyframe at a time, per-stage launches,
    no graph) -> plain_maps_per_s and batched_over_plain;
  * per-stage kernel time from a single-stream run with whole-batch launches
    and an event pair around each stage (the BP half-sweeps of the run share
    one pair), turned into achieved HBM GB/s from the algorithmic bytes of
    each stage (roofline);
  * the FP32 OpenMP CPU oracle on a bounded sample (cpu_baseline).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--streams B]
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "depth maps/s @752×480×64 hyps, batched vs plain GPU; % HBM roofline"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--streams", type=int, default=32)
    ap.add_argument("--width", type=int, default=752)
    ap.add_argument("--height", type=int, default=480)
    ap.add_argument("--hyps", type=int, default=64)
    ap.add_argument("--scenes", type=int, default=8, help="distinct rendered scenes shared round-robin by streams")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tune-cache", default=os.environ.get(
        "QM_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "open_quadtree_mapping_amd",
                                      "launch_shapes.json")),
                    help="JSON file of autotuned launch shapes: cached choices for this device and batch are "
                         "reused, new ones added ('' for none)")
    ap.add_argument("--retune", action="store_true", help="autotune even when a cache holds this batch's shapes")
    ap.add_argument("--no-shipped-shapes", action="store_true",
                    help="do not consult the launch shapes measured on MI355X that ship with the package")
    ap.add_argument("--tune-reps", type=int, default=5)
    ap.add_argument("--overlap", type=int, default=None,
                    help="force the frame-group overlap setting (0-13) after autotuning, e.g. for a profiled run")
    ap.add_argument("--no-graph", action="store_true",
                    help="time the batched launches without hipGraph replay (for counter collection)")
    ap.add_argument("--workload", default="keyframes32",
                    choices=["keyframes32", "uneven12", "diff32", "upsample16"],
                    help="keyframes32: the headline workload (default). The others print one supplementary "
                         "line for another shape: uneven12 = 4 x 376x240 + 4 x 752x480 + 4 x 1280x720 "
                         "streams, N = 64, batched; diff32 = 32 streams at 752x480, N = 128 over 0.3-50 m, "
                         "batched forward plus the backward to each stream's lambda; upsample16 = 16 "
                         "upsample problems at 752x480, forward plus adjoint")
    ap.add_argument("--cost-arith", choices=["fused", "reference"], default="fused",
                    help="projection arithmetic of the cost volume (qm_params.cost_arith): fused (hoisted "
                         "products, FMA, one exact reciprocal) or the reference's operation order; both are "
                         "bit-identical to the oracle's matching mode")
    ap.add_argument("--upsample-solver", choices=["scan", "thomas"], default="scan",
                    help="line solver of the upsample (qm_params.upsample_solver): scan (segmented scans over "
                         "64 lanes per line) or thomas (the reference's serial recurrences); both are "
                         "bit-identical to the oracle's matching mode")
    ap.add_argument("--jitter-deg", type=float, default=0.0,
                    help="rotation jitter of the synthetic camera poses, U(-j, +j) degrees about each axis "
                         "(default 0: the keyframe-to-reference poses are pure translations)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB five times and exit: the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def bp_traffic(levels, W, H, N, iters=(4, 10, 10, 10), pairs=True):
    """Algorithmic bytes of the BP half-sweeps for one keyframe, following the
    skip rule of the kernels (quadtree lookup at (bx << lev, y << lev)) and
    their schedule (the extraction of the last level-0 sweep's colour included
    there, as in the pipeline).
    One launch per half-sweep (pairs=False): an update reads the cell's data
    row and 4 incoming messages and writes 4 outgoing ones, 36N bytes; a cell
    initialised from its parent at a level's first sweep (the cells the
    quadtree skips, or every cell not updated when the level has fewer than two
    half-sweeps) costs 16N read + 16N write (16N write only at the coarsest
    level).
    Fused pairs (pairs=True): the first half-sweep's updates read 20N (data +
    4 incoming, from the parent level in a level's first pair) and write their
    4 messages (16N) only in the level's last pair, and not at level 0; the
    second half-sweep's updates read their data row (4N; the incoming messages
    come from LDS) and write 16N; a skipped cell initialised at a level's first
    pair is written to both message buffers (16N read + 32N write, 32N write
    at the coarsest level); an odd last half-sweep runs alone (36N). The halo
    cells a pair recomputes are not algorithmic bytes: the HBM counters show
    them."""
    import numpy as np
    Wc, Hc = W // 4, H // 4
    lw, lh = [Wc], [Hc]
    for _ in range(3):
        lw.append((lw[-1] + 1) // 2)
        lh.append((lh[-1] + 1) // 2)
    total = 0
    updates = 0
    for l in range(4):
        w, h = lw[l], lh[l]
        lev = l + 2
        ys, xs = np.mgrid[0:h, 0:w]
        bx = xs >> 1
        if lev <= 4:
            X = np.minimum(bx << lev, W - 1)
            Y = np.minimum(ys << lev, H - 1)
            active = levels[Y, X] <= lev
        else:
            active = np.ones((h, w), bool)
        colA = (xs & 1) == (ys & 1)
        nA = int((active & colA).sum())
        nB = int((active & ~colA).sum())
        n_skip = int((~active).sum())
        cells = w * h
        it_l = iters[l]
        if it_l == 0:
            total += (16 if l == 3 else 32) * N * cells
            continue
        if not pairs or it_l < 2:
            for it in range(it_l):
                n_upd = nA if it % 2 == 0 else nB
                updates += n_upd
                total += 36 * N * n_upd
                if it == 0:
                    copied = (cells - n_upd) if it_l < 2 else n_skip
                    total += (16 if l == 3 else 32) * N * copied
            continue
        for it in range(0, it_l - 1, 2):
            last = it + 2 == it_l
            updates += nA + nB
            total += 20 * N * nA + 20 * N * nB
            if last and l > 0:
                total += 16 * N * nA
            if it == 0:
                total += (32 if l == 3 else 48) * N * n_skip
        if it_l % 2:
            updates += nA
            total += 36 * N * nA
    return total, updates


def bp_launches(iters=(4, 10, 10, 10), pairs=True):
    return sum(1 if it == 0 else (it // 2 + it % 2 if pairs and it >= 2 else it) for it in iters)


def stage_bytes(levels, W, H, N, pairs=True, reset_in_rows=True):
    import numpy as np
    C = (W // 4) * (H // 4)
    lw, lh = [W // 4], [H // 4]
    for _ in range(3):
        lw.append((lw[-1] + 1) // 2)
        lh.append((lh[-1] + 1) // 2)
    sub = levels[0:4 * (H // 4):4, 0:4 * (W // 4):4].astype(np.int64)
    ys, xs = np.mgrid[0:H // 4, 0:W // 4]
    qs = 1 << sub
    node = ((4 * xs) % qs == 0) & ((4 * ys) % qs == 0)
    # the last level-0 half-sweep (B cells, x + y odd, after 4 half-sweeps)
    # extracts its colour's winners; the extraction kernel does the A cells
    last = ((xs + ys) & 1) == 1
    nodes = int((node & ~last).sum())
    others = int((~node & ~last).sum())
    bp, upd = bp_traffic(levels, W, H, N, pairs=pairs)
    bp += 8 * int((node & last).sum())  # fused extraction: depth + label of the B nodes
    px = W * H
    return {
        "quadtree": 2 * px,
        # level 0, and levels 1-3 from the cost kernel's tiles (its fused pyramid)
        "cost": 5 * px + C + 4 * N * C + 4 * N * sum(a * b for a, b in zip(lw[1:], lh[1:])),
        "pyramid": 4 * N * sum(a * b for a, b in zip(lw, lh)),
        "bp_sweep": bp,
        "extract": nodes * (20 * N + 8) + others * 9,
        # image, labels + depths, row results; the next run's z-buffer reset
        # (8 bytes per pixel) in the row scan, else in the column pass
        "upsample_row": px + 8 * C + 8 * px + (8 * px if reset_in_rows else 0),
        "upsample_col": px + 8 * px + 4 * px + (0 if reset_in_rows else 8 * px),  # image, row results, dense
        # the batch's warp reads the prior seeds and scatters z-buffer keys (the
        # warped seeds are rebuilt by the update from the winner's key); the
        # update reads keys, the winner's seed and the dense depth and writes
        # the new seeds and the depth (the warp table only once a backward asks)
        "fuse_warp": 16 * px + 8 * px,
        "fuse_update": 8 * px + 16 * px + 4 * px + 16 * px + 4 * px,
    }, upd


FP32_PEAK_TFLOPS = 157.3
COST_FLOP_PER_SLOT = 40  # FP32 operations per (node cell, sample, slot) of the reference's cost kernel


def cost_flops(levels, W, H, N):
    """Algorithmic FP32 operations of the cost volume for one keyframe: the
    reference's image_to_cost (src/depth_extract.cu:175-219) does about 40 per
    (node cell, sample, slot) -- the projection through R|t and K (27) plus
    the patch-tap sums of the nine taps -- over 5 slots."""
    import numpy as np
    sub = levels[0:4 * (H // 4):4, 0:4 * (W // 4):4].astype(np.int64)
    ys, xs = np.mgrid[0:H // 4, 0:W // 4]
    qs = 1 << sub
    x, y = 4 * xs, 4 * ys
    node = (x % qs == 0) & (y % qs == 0) & (x > 0) & (y > 0) & (x < W - 1) & (y < H - 1)
    return int(node.sum()) * N * 5 * COST_FLOP_PER_SLOT


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_traffic(stage, W, H, N, B):
    """HBM bytes per launch of the roofline kernel for this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration recorded there); (None, None) when no summary covers it."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"pipeline_{W}x{H}x{N}_b{B}_hbm_counters.md")
    if not os.path.exists(path):
        return None, None
    text = open(path).read()
    m = re.search(r"```json\n(.*?)\n```", text, re.S)
    if not m:
        return None, None
    rec = json.loads(m.group(1))
    if rec.get("stage") != stage:
        return None, None
    return int(rec["hbm_bytes_per_launch"]), os.path.relpath(path, os.path.dirname(os.path.abspath(__file__)))


def measured_cost_valu(W, H, N, arith):
    """VALU instructions per (cell, sample) of the cost kernel in this
    arithmetic, from the committed rocprofv3 issue-counter summary; (None, None)
    when no summary covers it."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"pipeline_{W}x{H}x{N}_b8_issue_counters.md")
    if not os.path.exists(path):
        return None, None
    m = re.search(r"```json\n(.*?)\n```", open(path).read(), re.S)
    if not m:
        return None, None
    rec = json.loads(m.group(1))
    v = rec.get("valu_per_cell_sample", {}).get(arith)
    return v, os.path.relpath(path, os.path.dirname(os.path.abspath(__file__)))


def run_workload(args):
    """Supplementary workloads (--workload): one JSON line with the rate of one
    step, timed like the headline workload (warm-up steps, then args.steps
    steps between two device synchronisations)."""
    import numpy as np
    import torch

    from open_quadtree_mapping_amd import autograd as QA
    from open_quadtree_mapping_amd import default_params
    from open_quadtree_mapping_amd.batch import Batch
    from open_quadtree_mapping_amd.scene import camera_for, forward_lateral_trajectory, make_scene, render
    from open_quadtree_mapping_amd.se3 import relative_pose

    dev = torch.device("cuda:0")

    def keyframes(W, H, n, seed0):
        cam = camera_for(W, H)
        out = []
        for s in range(n):
            traj = forward_lateral_trajectory(6, seed=seed0 + s)
            sc = make_scene(seed0 + s)
            imgs = [torch.from_numpy(render(sc, cam, T)[0]).to(dev) for T in traj[1:]]
            T = np.stack([relative_pose(traj[4 - i], traj[5]) for i in range(4)])
            out.append(dict(key=imgs[4], refs=[imgs[3], imgs[2], imgs[1], imgs[0]], T=T,
                            cam=(cam.fx, cam.fy, cam.cx, cam.cy)))
        return out

    def shapes_of(batch):
        out = {f"{W}x{H}": batch.group_tuning(g) for g, (W, H) in enumerate(batch.groups())}
        out["overlap"] = batch.tuning()["overlap"]
        return out

    shapes = None

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    wl = args.workload
    arith = 1 if args.cost_arith == "fused" else 0
    if wl == "uneven12":
        frames = keyframes(376, 240, 4, 300) + keyframes(752, 480, 4, 310) + keyframes(1280, 720, 4, 320)
        batch = Batch(frames, default_params(n_hyp=64, idepth_min=0.02, idepth_max=2.0, cost_arith=arith,
                                             upsample_solver=1 if args.upsample_solver == "scan" else 0),
                      outputs=("depth",))
        batch.autotune(reps=args.tune_reps, cache=args.tune_cache or None, refresh=args.retune,
                               seeds=shipped_shapes(args))
        shapes = shapes_of(batch)
        sec = timed(lambda: batch.run(graph=True))
        n, unit, what = 12, "depth maps/s", "batched graph replay of 12 uneven streams, N = 64"
        batch.close()
    elif wl == "diff32":
        frames = [dict(f, key=f["key"].clone(), refs=[r.clone() for r in f["refs"]])
                  for f in keyframes(752, 480, 8, 100) for _ in range(4)]
        batch = Batch(frames, default_params(n_hyp=128, idepth_min=1.0 / 50.0, idepth_max=1.0 / 0.3, cost_arith=arith),
                      outputs=("depth",))
        batch.autotune(reps=args.tune_reps, cache=args.tune_cache or None, refresh=args.retune,
                               seeds=shipped_shapes(args))
        shapes = shapes_of(batch)
        g = [torch.randn(480, 752, device=dev) for _ in range(32)]

        def step():
            batch.run(graph=True)
            batch.backward(g)  # synchronises the stream
        sec = timed(step)
        n, unit, what = 32, "depth maps/s", "batched forward (graph) plus backward to 32 lambdas, N = 128"
        batch.close()
    else:  # upsample16
        B, H, W = 16, 480, 752
        gen = torch.Generator(device="cpu").manual_seed(3)
        d = torch.rand(B, H, W, generator=gen) * 9 + 1
        d = (d * (torch.rand(B, H, W, generator=gen) < 0.3)).to(dev)
        conf = (1.0 - torch.rand(B, H, W, generator=gen)).to(dev)
        wr = torch.exp(torch.empty(B, H, W + 1).uniform_(np.log(1e-3), np.log(10), generator=gen)).to(dev)
        wc = torch.exp(torch.empty(B, H + 1, W).uniform_(np.log(1e-3), np.log(10), generator=gen)).to(dev)
        lams = torch.full((B,), 10.0, device=dev)
        g = torch.randn(B, H, W, device=dev)
        leaves = [t.requires_grad_() for t in (d, conf, wr, wc, lams)]

        def step():
            for t in leaves:
                t.grad = None
            QA.upsample(d, wr, wc, conf=conf, lam=lams).backward(g)
        sec = timed(step)
        n, unit, what = B, "problems/s", "16 upsample problems, forward plus adjoint (all inputs and lambda)"
    cfg = {"workload": what}
    if wl != "upsample16":
        cfg["cost_arith"] = args.cost_arith
        # the batch backward differentiates the Thomas solve
        cfg["upsample_solver"] = args.upsample_solver if wl == "uneven12" else "thomas"
        cfg["launch_shapes"] = shapes
    print(json.dumps({"workload": wl, "value": round(n / sec, 2), "unit": unit, "ms_per_step": round(sec * 1e3, 4),
                      "steps": args.steps, "warmup": args.warmup, "dtype": "f32", "data": "synthetic",
                      "config": cfg}), flush=True)


def shipped_shapes(args):
    """The package's JSON of launch shapes autotuned on MI355X for the bench
    workloads (consulted after the user cache)."""
    if args.no_shipped_shapes:
        return ()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "open_quadtree_mapping_amd",
                        "launch_shapes_mi355x.json")
    return (path,) if os.path.exists(path) else ()


def main():
    args = parse()
    # the CPU baseline's OpenMP threads stay on neighbouring cores; libgomp reads
    # this when the oracle library loads, so it is set before anything imports it
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(streams shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.copy_mb:
        n = args.copy_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}))
        return

    if args.workload != "keyframes32":
        run_workload(args)
        return

    from open_quadtree_mapping_amd import default_params
    from open_quadtree_mapping_amd import stages as S
    from open_quadtree_mapping_amd.batch import Batch
    from open_quadtree_mapping_amd.scene import camera_for, forward_lateral_trajectory, make_scene, render
    from open_quadtree_mapping_amd.se3 import relative_pose

    W, H, N, B = args.width, args.height, args.hyps, args.streams
    dev = torch.device("cuda:0")
    cam = camera_for(W, H)
    imin, imax = 0.02, 2.0  # 0.5 - 50 m
    arith = 1 if args.cost_arith == "fused" else 0
    solver = 1 if args.upsample_solver == "scan" else 0
    params = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax, cost_arith=arith, upsample_solver=solver)

    t_setup = time.time()
    nsc = min(args.scenes, B)
    scenes = []
    for s in range(nsc):
        traj = forward_lateral_trajectory(6, seed=100 + s, jitter_deg=args.jitter_deg)
        sc = make_scene(100 + s)
        imgs = [render(sc, cam, T)[0] for T in traj]
        k = 5
        T = np.stack([relative_pose(traj[k - 1 - i], traj[k]) for i in range(4)])
        scenes.append((imgs, T, traj))
    print(f"[bench] rendered {nsc} scenes in {time.time() - t_setup:.1f}s", file=sys.stderr, flush=True)

    camt = (cam.fx, cam.fy, cam.cx, cam.cy)
    frames = []
    for i in range(B):
        imgs, T, traj = scenes[i % nsc]
        d = [torch.from_numpy(im).to(dev) for im in imgs[1:]]  # distinct device copies per stream
        frames.append(dict(key=d[4], refs=[d[3], d[2], d[1], d[0]], T=T, cam=camt))

    # ---- batched path (hipGraph) -------------------------------------------------
    batch = Batch(frames, params, outputs=("depth",))
    stream = torch.cuda.current_stream()
    tuned = None
    if not args.no_autotune:
        tuned = batch.autotune(reps=args.tune_reps, cache=args.tune_cache or None, refresh=args.retune,
                               seeds=shipped_shapes(args))
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    if args.overlap is not None:
        batch.set_tuning(overlap=args.overlap)
    use_graph = not args.no_graph
    for _ in range(args.warmup):
        batch.run(graph=use_graph)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        batch.run(graph=use_graph)
    e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall_s = t1 - t0
    ev_ms = e0.elapsed_time(e1)
    maps_per_s = B * args.steps / wall_s
    ms_per_step = ev_ms / args.steps
    print(f"[bench] batched: {maps_per_s:.1f} maps/s ({ms_per_step:.3f} ms/step of {B} maps)", file=sys.stderr,
          flush=True)

    # ---- per-stage kernel time (events around every launch) ----------------------
    reps = 3
    acc = {}
    for _ in range(reps):
        st = batch.run_timed()
        for k, (ms, cnt) in st.items():
            a = acc.setdefault(k, [0.0, 0])
            a[0] += ms
            a[1] += cnt
    levels_per_stream = [S.quadtree(f["key"]).cpu().numpy() for f in frames]
    pairs = batch.tuning()["bp_pair"] != 0
    stage_alg = {}
    bp_updates = 0
    for lv in levels_per_stream:
        sb, upd = stage_bytes(lv, W, H, N, pairs=pairs, reset_in_rows=solver == 1 and W <= 1536)
        bp_updates += upd
        for k, v in sb.items():
            stage_alg[k] = stage_alg.get(k, 0) + v
    stages = {}
    for k, (ms, cnt) in acc.items():
        per_run_ms = ms / reps
        launches = cnt // reps
        if launches == 0:
            continue  # a stage folded into another (the pyramid, into the cost volume)
        gbs = stage_alg[k] / (per_run_ms * 1e-3) / 1e9 if per_run_ms > 0 else 0.0
        stages[k] = {"ms_per_run": round(per_run_ms, 4), "launches": launches,
                     "avg_launch_us": round(1e3 * per_run_ms / max(launches, 1), 2),
                     "alg_bytes_per_run": int(stage_alg[k]), "achieved_GBs": round(gbs, 1),
                     "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    # the cost stage's FP32 work (BASELINE's FLOP fraction) and its time in the
    # other projection arithmetic, from a batch of the same frames
    flops = sum(cost_flops(lv, W, H, N) for lv in levels_per_stream)
    if "cost" in stages:
        c = stages["cost"]
        tflops = flops / (c["ms_per_run"] * 1e-3) / 1e12
        c.update({"alg_gflop_per_run": round(flops / 1e9, 3), "achieved_TFLOPs": round(tflops, 2),
                  "flop_frac": round(tflops / FP32_PEAK_TFLOPS, 4), "arith": args.cost_arith})
        valu, vsrc = measured_cost_valu(W, H, N, args.cost_arith)
        if valu is not None:
            c.update({"valu_per_cell_sample": valu, "valu_source": vsrc})
        other = default_params(n_hyp=N, idepth_min=imin, idepth_max=imax, cost_arith=1 - arith,
                               upsample_solver=solver)
        ob = Batch(frames, other, outputs=("depth",))
        ob.set_tuning(**batch.tuning())
        ob.run_timed()
        oms = sum(ob.run_timed()["cost"][0] for _ in range(reps)) / reps
        ob.close()
        c["other_arith_ms_per_run"] = round(oms, 4)
    if "bp_sweep" in stages and bp_updates:
        # bytes per cell update in units of N (36 for one launch per half-sweep)
        stages["bp_sweep"].update({"updates_per_run": bp_updates, "fused_pairs": pairs,
                                   "alg_bytes_per_update_N": round(stage_alg["bp_sweep"] / bp_updates / N, 2)})
        tr, _ = measured_traffic("bp_sweep", W, H, N, B)
        if tr is not None:
            stages["bp_sweep"]["hbm_bytes_per_update_N"] = round(
                tr * stages["bp_sweep"]["launches"] / bp_updates / N, 2)
    dom = max(stages, key=lambda k: stages[k]["ms_per_run"])
    d = stages[dom]
    traffic, traffic_src = measured_traffic(dom, W, H, N, B)
    roofline = {"bound": "hbm", "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["hbm_frac"], "traffic": traffic, "traffic_source": traffic_src, "kernel": dom,
                "alg_bytes_per_launch": int(d["alg_bytes_per_run"] / max(d["launches"], 1)),
                "avg_launch_us": d["avg_launch_us"]}
    print(f"[bench] stages: {json.dumps(stages)}", file=sys.stderr, flush=True)

    # ---- plain per-frame path ---------------------------------------------------
    plain = None
    if not args.no_plain:
        singles = [Batch([f], params, outputs=("depth",)) for f in frames]
        for s_ in singles:
            s_.run(graph=False)
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        for _ in range(args.plain_steps):
            for s_ in singles:
                s_.run(graph=False)
        torch.cuda.synchronize()
        tp1 = time.perf_counter()
        plain = B * args.plain_steps / (tp1 - tp0)
        for s_ in singles:
            s_.close()
        print(f"[bench] plain per-frame: {plain:.1f} maps/s", file=sys.stderr, flush=True)
    batch_tuning = batch.tuning()
    batch.close()

    # ---- CPU baseline: FP32 OpenMP oracle on a bounded sample -------------------
    # one warm-up keyframe, then keyframes until the time budget is spent (at
    # least 5); per-stage and end-to-end medians as BASELINE.md's protocol asks
    cpu = None
    if not args.no_cpu:
        import oracle as O
        O.lib()
        threads = O.get_threads()
        imgs, T, traj = scenes[0]
        seeds = np.zeros((H, W, 4), np.float32)
        eye = np.eye(3, 4, dtype=np.float32)
        kf = lambda: O.keyframe(imgs[5], [imgs[4], imgs[3], imgs[2], imgs[1]], T, cam, N, imin, imax, seeds, eye,
                                arith=arith, solver=solver)
        kf()  # warm-up
        per_stage, ends = [], []
        tc0 = time.perf_counter()
        while True:
            t_a = time.perf_counter()
            r = kf()
            ends.append(1e3 * (time.perf_counter() - t_a))
            per_stage.append(r["stage_ms"])
            if len(ends) >= 5 and (time.perf_counter() - tc0 > args.cpu_seconds or len(ends) >= 400):
                break
        tc = time.perf_counter() - tc0
        med = np.median(np.array(per_stage), axis=0)
        cpu = {"value": round(len(ends) / tc, 3), "unit": "depth maps/s", "cores": threads, "kind": "port",
               "sample": f"{len(ends)} keyframes of {W}x{H}x{N} after one warm-up, full pipeline, FP32 OpenMP "
                         f"oracle ({args.cost_arith} cost arithmetic, {args.upsample_solver} upsample), {tc:.1f}s",
               "cpu_model": cpu_model(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
               "median_ms": round(float(np.median(ends)), 3),
               "stage_median_ms": {k: round(float(v), 3) for k, v in
                                   zip(("quadtree", "cost", "bp", "upsample", "fusion"), med)}}
        print(f"[bench] cpu oracle: {cpu['value']} maps/s on {threads} threads ({cpu['cpu_model']}), "
              f"stage medians {cpu['stage_median_ms']}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(maps_per_s, 2),
        "unit": "depth maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded planar scenes (6 textured rectangles, fractal value noise) rendered to uint8 "
                "along a camera path of 5 cm per frame"
                + (f" with U(-{args.jitter_deg}, {args.jitter_deg}) deg rotation jitter" if args.jitter_deg > 0
                   else " (pure translations)"),
        "config": {"workload": f"{B} streams x {W}x{H} keyframes, {N} inverse-depth hypotheses, full pipeline "
                               "(quadtree, cost volume, HBP, upsample, fusion), "
                               + ("graph replays per step (one graph per stream segment)" if use_graph
                                  else "plain batched launches per step"),
                   "streams": B, "width": W, "height": H, "hypotheses": N, "cost_arith": args.cost_arith,
                   "upsample_solver": args.upsample_solver, "pose_jitter_deg": args.jitter_deg,
                   "launch_shapes": batch_tuning,
                   "launch_shapes_source": ("defaults" if tuned is None else
                                            "cache" if all(ms is None for _, ms in tuned.values()) else "autotune")},
        "plain_maps_per_s": None if plain is None else round(plain, 2),
        "batched_over_plain": None if plain is None else round(maps_per_s / plain, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stages,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
