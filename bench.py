"""Benchmark of the sparse-voxel rasterizer training step on one MI355X.

    python bench.py [--gpus 1] [--steps 100] [--warmup 10]

Workload (the 1M-voxel 1080p training configuration): an octree of 1,000,000
voxels (inner levels 9-17 grown on 8 random sphere surfaces, plus the
reference's 5-level outside shell), one 1920x1080 pinhole camera (fovx 50 deg)
on a radius-3 sphere, per-voxel FP32 corner densities and degree-3 SH.  One
step is the reference's training step through the drop-in API:
rasterize_voxels (preprocess, SH colour, binning, compositor) -> MSE loss ->
backward (compositor adjoint, SH backward) -> SparseAdam (sparse=True,
lr 0.025 density / 0.01 SH, betas (0.1, 0.99)).

Prints one JSON line: steps/s, forward Mpix/s, per-stage kernel times measured
with HIP events recorded on the launch stream inside the timed region, the
roofline of the dominant kernel and the oracle (CPU restatement) timed on the
host cores on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STAGES = ["preprocess", "scan", "binning", "render_fwd", "render_bwd", "sh_fwd", "sh_bwd", "adam", "grad_reduce"]
METRIC = "train steps/s (fwd+bwd+Adam) & fwd Mpix/s, 1M voxels 1080p; batched vs plain"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="B")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--fwd-iters", type=int, default=50)
    ap.add_argument("--condition-steps", type=int, default=200,
                    help="untimed training steps before the warm-up, so that the timed frames (whose traversed-entry "
                         "count E grows as the densities train) do not depend on --steps / --warmup")
    ap.add_argument("--kernel-iters", type=int, default=20,
                    help="frames per compositor in the back-to-back launch timing (the roofline's launch duration)")
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed oracle steps on all host threads in the CPU "
                                                             "baseline (median; 0: skip the baseline)")
    ap.add_argument("--cpu-single-steps", type=int, default=5, help="timed oracle steps on one thread (median)")
    ap.add_argument("--batch-config", default="D", help="multi-view workload for the batched-vs-plain figure "
                                                          "('none' to skip)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--large-config", default="E", help="large batched workload: all its views in one batch per "
                                                         "step, autotuned ('none' to skip)")
    ap.add_argument("--large-steps", type=int, default=50, help="timed batched training steps of the large workload")
    ap.add_argument("--large-plain-rounds", type=int, default=1,
                    help="timed rounds of the plain path (one drop-in step per view) on the large workload")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the bench, run 3 plain copy kernels (torch.neg) of this many MiB: the known-size "
                         "dispatch that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes")
    return ap.parse_args()


def load_profile(kernel, config, n_vox, width, height, entries):
    """The committed rocprofv3 counter summary of `kernel` on this workload
    (profiles/<kernel>_<config>_counters.txt: FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel of known
    size, and the SQ counter passes, over the bench command it names).  Returns (traffic bytes per dispatch,
    extra fields, note).  The traffic is used only when the summary's frame traversed the same number of
    tile entries as this one to within 1 %; otherwise it is None and the note says why (the per-entry figure is
    still reported)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{config}_counters.txt")
    if not os.path.exists(path):
        return None, {}, "no counter summary for this kernel and workload"
    fields = {}
    for line in open(path):
        if ":" in line and not line.startswith((" ", "#", "|")):
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    if fields.get("workload") != f"{config} {n_vox} {width}x{height}":
        return None, {}, "counter summary is for another workload"
    try:
        traffic = int(fields["traffic_bytes_per_dispatch"])
        e_prof = int(fields["entries_traversed_E"])
    except (KeyError, ValueError):
        return None, {}, "counter summary lacks traffic or E"
    extra = {"traffic_bytes_per_entry": traffic / e_prof, "profile_E": e_prof}
    for k in ("valu_busy_frac", "lds_bank_conflict_frac", "occupancy_waves_per_simd"):
        if k in fields:
            extra[k] = float(fields[k])
    if abs(e_prof - entries) > 0.01 * entries:
        return None, extra, f"counter frame traversed {e_prof} entries, this one {entries} (> 1 % apart)"
    return traffic, extra, None


def progress(msg):
    """One progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def collect_stages(L):
    out = {}
    for i, name in enumerate(STAGES):
        tot, cnt = ctypes.c_double(0), ctypes.c_int64(0)
        rc = L.svr_profile_collect(i, ctypes.byref(tot), ctypes.byref(cnt))
        if rc != 0:
            raise RuntimeError("svr_profile_collect failed")
        out[name] = (float(tot.value), int(cnt.value))
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, runs, single_runs):
    """The oracle's training step (preprocess, SH, binning, compositor fwd/bwd, SH bwd, Adam) on the host cores:
    one warm-up step, then the median of `runs` timed steps with every host thread (OpenMP), and the median of
    `single_runs` steps on one thread."""
    from oracle import oracle as O
    sc, cam = case["scene"], case["cam"]
    cam_pos = cam["c2w"][:3, 3]
    geos, sh0, shs = sc["geos"].copy(), sc["sh0"].copy(), sc["shs"].copy()
    state = {k: (np.zeros_like(v), np.zeros_like(v)) for k, v in (("g", geos), ("s0", sh0), ("s", shs))}
    target = case["target"]
    H, W = cam["H"], cam["W"]
    counter = [0]

    def one():
        counter[0] += 1
        step = counter[0]
        t0 = time.perf_counter()
        n_dup, bbox, qbits = O.preprocess(cam, sc["vox_centers"], sc["vox_lengths"])
        idx = np.nonzero(n_dup > 0)[0]
        rgbs = O.sh_fwd(3, idx, sc["vox_centers"], cam_pos, sh0, shs)
        _, _, vl, ranges = O.binning(cam, sc["octpath"], n_dup, bbox, qbits)
        color, T, nc, tl = O.render_fwd(cam, ranges, vl, bbox, sc["vox_centers"], sc["vox_lengths"], geos, rgbs)
        dcol = (2.0 / color.size) * (color - target)
        dg, dr, _ = O.render_bwd(cam, ranges, vl, bbox, sc["vox_centers"], sc["vox_lengths"], geos, rgbs, T, nc, tl,
                                 dcol.astype(np.float32), np.zeros((H, W), np.float32))
        d0, ds = O.sh_bwd(3, 16, idx, sc["vox_centers"], cam_pos, dr.astype(np.float32))
        for key, p, g, lr in (("g", geos, dg.astype(np.float32), 0.025), ("s0", sh0, d0, 0.01), ("s", shs, ds, 0.01)):
            m, v = state[key]
            O.adam_step(p, np.ascontiguousarray(g), m, v, step, lr, 0.1, 0.99, 1e-15, sparse=True)
        return time.perf_counter() - t0

    all_threads = O.num_threads()
    one()  # warm-up (page faults, thread pool)
    multi = sorted(one() for _ in range(runs))
    O.set_num_threads(1)
    try:  # the data is paged in and warm from the runs above: no second warm-up
        single = sorted(one() for _ in range(single_runs))
    finally:
        O.set_num_threads(all_threads)
    med = multi[len(multi) // 2]
    med1 = single[len(single) // 2]
    return dict(value=1.0 / med, unit="steps/s", cores=all_threads, kind="port",
                sample=f"full training steps of the same 1M-voxel 1920x1080 workload (the oracle's C++ "
                       f"restatement): median of {runs} steps on {all_threads} OpenMP threads and of {single_runs} "
                       f"steps on 1 thread, after one warm-up step, from the parameters the GPU run ended with",
                cpu_model=cpu_model(), sec_per_step_median=med, sec_per_step_range=[multi[0], multi[-1]],
                single_thread={"value": 1.0 / med1, "unit": "steps/s", "cores": 1, "sec_per_step_median": med1,
                               "sec_per_step_range": [single[0], single[-1]]})


def batch_vs_plain(args, dev):
    """Multi-view training on one octree: the batched path (all views of a step in one launch per stage, loss
    summed over views, one SparseAdam step) against the plain path (one drop-in training step per view), both
    in views per second."""
    import torch
    from sv_raster_amd import _C, batch, renderer, sparse_adam, synthetic as S
    name = args.batch_config
    sc = S.make_scene(name, seed=args.seed)
    sizes = S.config_views(name, args.seed)
    cams = [S.make_camera(W, H, seed=args.seed + 1000 + u) for u, (W, H) in enumerate(sizes)]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    targets = [t(S.make_target(c["H"], c["W"], args.seed + u)) for u, c in enumerate(cams)]
    settings = [renderer.RasterSettings(color_mode="sh", n_samp_per_vox=1, image_width=c["W"], image_height=c["H"],
                                        tanfovx=c["tanfovx"], tanfovy=c["tanfovy"], cx=c["cx"], cy=c["cy"],
                                        w2c_matrix=t(c["w2c"]), c2w_matrix=t(c["c2w"])) for c in cams]
    N = centers.shape[0]
    V = len(cams)

    def fresh():
        geos = t(sc["geos"]).requires_grad_()
        sh0 = t(sc["sh0"]).requires_grad_()
        shs = t(sc["shs"]).requires_grad_()
        opt = sparse_adam.SparseAdam([{"params": [geos], "lr": 0.025}, {"params": [sh0], "lr": 0.01},
                                      {"params": [shs], "lr": 0.01}], betas=(0.1, 0.99), eps=1e-15, sparse=True)
        return geos, sh0, shs, opt

    _C.order_table_cache.get(octpath)
    # plain: one training step per view
    geos, sh0, shs, opt = fresh()
    subdiv = torch.ones((N, 1), device=dev)

    def plain_round():
        for u in range(V):
            def vox_fn(idx, cam_pos, color_mode):
                rgbs = renderer.SH_eval.apply(3, idx, centers, cam_pos, None, sh0, shs)
                return {"geos": geos, "rgbs": rgbs, "subdiv_p": subdiv}
            color, _, _, _, _ = renderer.rasterize_voxels(settings[u], octpath, centers, lengths, vox_fn)
            loss = torch.nn.functional.mse_loss(color, targets[u])
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()

    plain_round()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        plain_round()
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t0) / args.batch_steps

    # batched: all views in one step, binning tile picked by the autotuner
    geos, sh0, shs, opt = fresh()
    views, tune = batch.autotune(settings, octpath, centers, lengths, geos, sh0, shs)

    def batch_step():
        colors, _ = batch.rasterize_voxels_batch(views, octpath, centers, lengths, geos, sh0, shs)
        loss = sum(torch.nn.functional.mse_loss(c, tg) for c, tg in zip(colors, targets))
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()

    batch_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        batch_step()
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - t0) / args.batch_steps
    # batched, captured in a HIP graph: one replay per step (device step counts, fixed entry capacity)
    geos, sh0, shs, opt = fresh()
    trainer = batch.GraphTrainStep(views, octpath, centers, lengths, geos, sh0, shs, opt, targets)
    trainer.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        trainer.step()
    torch.cuda.synchronize()
    graph_s = (time.perf_counter() - t0) / args.batch_steps
    overflow = trainer.overflowed()
    pix = sum(W * H for W, H in sizes)
    return {"workload": f"{name}: {N} voxels, {V} views of uneven resolution ({pix / 1e6:.2f} Mpix per round), "
                        "fwd+bwd+sparse Adam",
            "views": V, "rounds_timed": args.batch_steps,
            "plain_views_per_s": V / plain_s, "batched_views_per_s": V / batch_s,
            "batched_over_plain": plain_s / batch_s,
            "plain_ms_per_round": plain_s * 1e3, "batched_ms_per_step": batch_s * 1e3,
            "graph_views_per_s": V / graph_s, "graph_ms_per_step": graph_s * 1e3,
            "graph_over_plain": plain_s / graph_s, "graph_capacity": trainer.capacity,
            "graph_overflow": overflow,
            "tile_shape": views.tile_shape, "bwd_window": views.bwd_window, "bwd_sync": views.bwd_sync,
            "autotune": tune,
            "note": "plain = one drop-in training step (rasterize_voxels + SH_eval, MSE, SparseAdam) per view; "
                    "batched = rasterize_voxels_batch over all views, summed MSE, one SparseAdam step; "
                    "graph = the batched step replayed from a captured HIP graph (GraphTrainStep)"}


def batched_large(args, dev):
    """The large batched training workload: every view of a step in one batch (one launch per stage), binning
    tile and backward depth picked by the autotuner, eager fwd + bwd (summed MSE) + SparseAdam steps; next to
    it the plain path (one drop-in training step per view), both in views per second."""
    import torch
    from sv_raster_amd import _C, batch, renderer, sparse_adam, synthetic as S
    name = args.large_config
    t0 = time.perf_counter()
    sc = S.make_scene(name, seed=args.seed)
    gen_s = time.perf_counter() - t0
    sizes = S.config_views(name, args.seed)
    cams = [S.make_camera(W, H, seed=args.seed + 1000 + u) for u, (W, H) in enumerate(sizes)]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    geos = t(sc["geos"]).requires_grad_()
    sh0 = t(sc["sh0"]).requires_grad_()
    shs = t(sc["shs"]).requires_grad_()
    del sc
    targets = [t(S.make_target(c["H"], c["W"], args.seed + u)) for u, c in enumerate(cams)]
    settings = [renderer.RasterSettings(color_mode="sh", n_samp_per_vox=1, image_width=c["W"], image_height=c["H"],
                                        tanfovx=c["tanfovx"], tanfovy=c["tanfovy"], cx=c["cx"], cy=c["cy"],
                                        w2c_matrix=t(c["w2c"]), c2w_matrix=t(c["c2w"])) for c in cams]
    N, V = centers.shape[0], len(cams)
    opt = sparse_adam.SparseAdam([{"params": [geos], "lr": 0.025}, {"params": [sh0], "lr": 0.01},
                                  {"params": [shs], "lr": 0.01}], betas=(0.1, 0.99), eps=1e-15, sparse=True)
    _C.order_table_cache.get(octpath)
    progress(f"{name}: {N} voxels, {V} views; autotune")
    t0 = time.perf_counter()
    views, tune = batch.autotune(settings, octpath, centers, lengths, geos, sh0, shs)
    tune_s = time.perf_counter() - t0
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()

    def batch_step():
        colors, _ = batch.rasterize_voxels_batch(views, octpath, centers, lengths, geos, sh0, shs)
        loss = sum(torch.nn.functional.mse_loss(c, tg) for c, tg in zip(colors, targets))
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        return loss

    batch_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.large_steps):
        loss = batch_step()
        if (k + 1) % 10 == 0:
            progress(f"{name}: {k + 1} batched steps")
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - t0) / args.large_steps
    peak = torch.cuda.max_memory_allocated()
    with torch.no_grad():
        st = batch.batch_forward(views, 3, octpath, centers, lengths, geos, sh0, shs)
        R = st["R"]
        del st
    progress(f"{name}: {batch_s * 1e3:.1f} ms per batched step; plain path")
    subdiv = torch.ones((N, 1), device=dev)

    def plain_round():
        for u in range(V):
            def vox_fn(idx, cam_pos, color_mode):
                rgbs = renderer.SH_eval.apply(3, idx, centers, cam_pos, None, sh0, shs)
                return {"geos": geos, "rgbs": rgbs, "subdiv_p": subdiv}
            color, _, _, _, _ = renderer.rasterize_voxels(settings[u], octpath, centers, lengths, vox_fn)
            loss = torch.nn.functional.mse_loss(color, targets[u])
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()

    plain_s = None
    if args.large_plain_rounds > 0:
        plain_round()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.large_plain_rounds):
            plain_round()
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - t0) / args.large_plain_rounds
    pix = sum(W * H for W, H in sizes)
    return {"workload": f"{name}: {N} voxels (levels 9-21 with a 5-level outside shell), {V} views of 720p-1440p "
                        f"({pix / 1e6:.1f} Mpix per step) in one batch, fwd+bwd+sparse Adam",
            "views": V, "entries_R": R, "steps_timed": args.large_steps,
            "batched_ms_per_step": batch_s * 1e3, "batched_views_per_s": V / batch_s,
            "batched_mpix_per_s": pix / batch_s / 1e6,
            "plain_ms_per_round": None if plain_s is None else plain_s * 1e3,
            "plain_views_per_s": None if plain_s is None else V / plain_s,
            "batched_over_plain": None if plain_s is None else plain_s / batch_s,
            "peak_device_memory_gib": peak / 2**30, "loss_last": float(loss.item()),
            "tile_shape": views.tile_shape, "bwd_window": views.bwd_window, "bwd_sync": views.bwd_sync,
            "autotune": tune, "autotune_s": tune_s, "scene_generation_s": gen_s}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from sv_raster_amd import _C, renderer, sparse_adam, synthetic as S
    from sv_raster_amd import _lib

    L = _lib.load()
    dev = torch.device("cuda:0")
    cfg = S.CONFIGS[args.config]
    sc = S.make_scene(args.config, seed=args.seed)
    W, H = S.config_views(args.config, args.seed)[0]
    cam = S.make_camera(W, H, seed=args.seed)
    target_np = S.make_target(H, W, args.seed)
    case = dict(scene=sc, cam=cam, target=target_np)

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    geos = t(sc["geos"]).requires_grad_()
    sh0 = t(sc["sh0"]).requires_grad_()
    shs = t(sc["shs"]).requires_grad_()
    subdiv = torch.ones((centers.shape[0], 1), device=dev)
    w2c, c2w = t(cam["w2c"]), t(cam["c2w"])
    target = t(target_np)
    N = centers.shape[0]

    def vox_fn(idx, cam_pos, color_mode):
        rgbs = renderer.SH_eval.apply(3, idx, centers, cam_pos, None, sh0, shs)
        return {"geos": geos, "rgbs": rgbs, "subdiv_p": subdiv}

    settings = renderer.RasterSettings(color_mode="sh", n_samp_per_vox=1, image_width=W, image_height=H,
                                       tanfovx=cam["tanfovx"], tanfovy=cam["tanfovy"], cx=cam["cx"], cy=cam["cy"],
                                       w2c_matrix=w2c, c2w_matrix=c2w)
    opt = sparse_adam.SparseAdam([{"params": [geos], "lr": 0.025}, {"params": [sh0], "lr": 0.01},
                                  {"params": [shs], "lr": 0.01}], betas=(0.1, 0.99), eps=1e-15, sparse=True)

    def step():
        color, _, _, _, _ = renderer.rasterize_voxels(settings, octpath, centers, lengths, vox_fn)
        loss = torch.nn.functional.mse_loss(color, target)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        return loss

    # one-time: rank-order table of this octree (reused while the octree is unchanged)
    torch.cuda.synchronize()
    t_tab = time.perf_counter()
    _C.order_table_cache.get(octpath)
    torch.cuda.synchronize()
    order_table_ms = (time.perf_counter() - t_tab) * 1e3

    for _ in range(args.condition_steps):
        step()
    torch.cuda.synchronize()
    progress(f"{args.condition_steps} conditioning steps done")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    progress(f"{args.warmup} warm-up steps done")

    # snapshot of parameters and optimizer state, so that the instrumented pass below repeats the same steps
    params = (geos, sh0, shs)
    snap_p = [p.detach().clone() for p in params]
    snap_s = {id(p): {k: (v.clone() if torch.is_tensor(v) else v) for k, v in opt.state[p].items()} for p in params}

    # the timed steps (no instrumentation: a stage event costs the queue a few microseconds of idle time)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    progress(f"{args.steps} timed steps: {ms_per_step:.4f} ms per step")
    # the same steps again (parameters and optimizer state restored) with HIP events around every stage,
    # recorded on the launch stream: per-stage kernel times and the roofline of the dominant kernel
    with torch.no_grad():
        for p, q in zip(params, snap_p):
            p.copy_(q)
            for k, v in snap_s[id(p)].items():
                if torch.is_tensor(v):
                    opt.state[p][k].copy_(v)
                else:
                    opt.state[p][k] = v
    _lib.check(L.svr_profile_enable(1), "profile")
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        loss = step()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    stages = collect_stages(L)
    _lib.check(L.svr_profile_enable(0), "profile")
    ms_per_step_instrumented = (t3 - t2) * 1e3 / args.steps
    progress(f"{args.steps} instrumented steps: {ms_per_step_instrumented:.4f} ms per step")

    # forward only (the FPS path): preprocess + SH + binning + compositor
    with torch.no_grad():
        for _ in range(3):
            renderer.rasterize_voxels(settings, octpath, centers, lengths, vox_fn)
        torch.cuda.synchronize()
        f0 = time.perf_counter()
        for _ in range(args.fwd_iters):
            renderer.rasterize_voxels(settings, octpath, centers, lengths, vox_fn)
        torch.cuda.synchronize()
        fwd_ms = (time.perf_counter() - f0) * 1e3 / args.fwd_iters

    # algorithmic traffic of the compositor kernels on this frame
    n_dup, geom = _C.rasterize_preprocess(W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w,
                                          cam["near"], octpath, centers, lengths)
    with torch.no_grad():
        vis_idx = torch.where(n_dup > 0)[0]
        rgbs = _C.sh_compute(3, vis_idx, centers, c2w[:3, 3], sh0, shs)
        R, binning, img, color, _, _, T_img, _ = _C.rasterize_voxels(
            1, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, 0.0, False, False, False, False,
            octpath, centers, lengths, geos, rgbs, geom)
        ranges, tile_last, _ = _C.unpack_ImageState(W, H, img)
        E = int((tile_last.reshape(-1).long() - ranges.reshape(-1, 2)[:, 0].long()).sum().item())
        dcol = 2.0 * (color - target) / color.numel()
        dg, _, _ = _C.rasterize_voxels_backward(R, 1, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c,
                                               c2w, 0.0, octpath, centers, lengths, geos, rgbs, geom, binning, img,
                                               T_img, dcol, None, None, torch.zeros_like(T_img), 0.0, None, 0.0, 0.0,
                                               False, False, None, None)
        N_h = int((dg != 0).any(1).sum().item())
        N_v = int((n_dup > 0).sum().item())
        # launch durations of the compositors with the queue kept full (frames back to back with no host
        # wait in between), so that a stage's events bracket its kernels only: the duration the roofline uses,
        # and the one a kernel trace of this command reports for these kernels
        _lib.check(L.svr_profile_enable(1), "profile")
        for _ in range(args.kernel_iters):
            _C.rasterize_voxels(1, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, 0.0, False,
                                False, False, False, octpath, centers, lengths, geos, rgbs, geom)
        for _ in range(args.kernel_iters):
            _C.sh_compute(3, vis_idx, centers, c2w[:3, 3], sh0, shs)
        for _ in range(args.kernel_iters):
            _C.rasterize_voxels_backward(R, 1, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w,
                                         0.0, octpath, centers, lengths, geos, rgbs, geom, binning, img, T_img, dcol,
                                         None, None, torch.zeros_like(T_img), 0.0, None, 0.0, 0.0, False, False,
                                         None, None)
        torch.cuda.synchronize()
        kst = collect_stages(L)
        _lib.check(L.svr_profile_enable(0), "profile")
    kernel_ms = {k: kst[k][0] / kst[k][1] for k in ("render_fwd", "render_bwd", "sh_fwd")}
    P = W * H
    T_tiles = ((W + 15) // 16) * ((H + 15) // 16)
    bytes_fwd = 72 * E + 20 * P + 12 * T_tiles
    bytes_bwd = 72 * E + 24 * P + 12 * T_tiles + 48 * N_h

    stage_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items()}
    # algorithmic (compulsory) bytes per training step of each stage; Adam counts its 3 launches
    # (8 density + 48 SH floats per voxel): a 4-byte gradient test per element plus 24 bytes per touched one
    stage_bytes = {"preprocess": 36 * N, "scan": 8 * N, "sh_fwd": 216 * N_v,
                   "binning": 28 * N + 36 * R + 8 * T_tiles, "render_fwd": bytes_fwd, "render_bwd": bytes_bwd,
                   "grad_reduce": 48 * R + 48 * N_v, "sh_bwd": 228 * N_v, "adam": 4 * 56 * N + 24 * 56 * N_h}
    stage_ms_step = {k: v[0] / args.steps for k, v in stages.items()}
    stages_gbs = {k: stage_bytes[k] / (stage_ms_step[k] * 1e-3) / 1e9 if stage_ms_step[k] else None
                  for k in STAGES}
    step_bytes = sum(stage_bytes.values())
    dom = "render_bwd" if kernel_ms["render_bwd"] >= kernel_ms["render_fwd"] else "render_fwd"
    other = "render_fwd" if dom == "render_bwd" else "render_bwd"
    dom_bytes = bytes_bwd if dom == "render_bwd" else bytes_fwd
    achieved = dom_bytes / (kernel_ms[dom] * 1e-3) / 1e9
    traffic, prof_extra, traffic_note = load_profile(dom, args.config, N, W, H, E)

    batched = None
    if args.batch_config.lower() != "none":
        progress("multi-view figure")
        batched = batch_vs_plain(args, dev)

    large = None
    if args.large_config.lower() != "none":
        torch.cuda.empty_cache()
        progress("large batched workload")
        large = batched_large(args, dev)
        torch.cuda.empty_cache()

    cpu = None
    if args.cpu_steps > 0:
        progress("CPU baseline")
        # the oracle steps start from the parameters the GPU steps ended with (the same training state)
        with torch.no_grad():
            sc_cpu = dict(case["scene"], geos=geos.detach().cpu().numpy(), sh0=sh0.detach().cpu().numpy(),
                          shs=shs.detach().cpu().numpy())
        cpu = cpu_baseline(dict(case, scene=sc_cpu), args.cpu_steps, args.cpu_single_steps)

    result = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "condition_steps": args.condition_steps,
        "ms_per_step": ms_per_step,
        "ms_per_step_instrumented": ms_per_step_instrumented,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded octree, camera and target image; random-init parameters)",
        "config": {"workload": f"{args.config}: {N} voxels (levels {cfg['lv_min'] + cfg['outside']}-"
                               f"{cfg['lv_max'] + cfg['outside']}, outside shell {cfg['outside']} levels), "
                               f"{W}x{H}, plain single view, fwd+bwd+sparse Adam",
                   "voxels": N, "width": W, "height": H, "entries_R": R, "entries_traversed_E": E,
                   "voxels_hit": N_h, "views": 1, "path": "plain"},
        "fwd_ms": fwd_ms,
        "fwd_mpix_s": P / (fwd_ms * 1e-3) / 1e6,
        "loss_last": float(loss.item()),
        "order_table_ms_once_per_octree": order_table_ms,
        "stages_ms_per_launch": stage_ms,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom_bytes,
                     "avg_launch_ms": kernel_ms[dom],
                     "launches_timed": args.kernel_iters,
                     "entries_traversed_E": E, "voxels_hit_N_h": N_h,
                     "traffic_note": traffic_note, **prof_extra,
                     "other_compositor": other,
                     "other_compositor_ms": kernel_ms[other],
                     "other_compositor_gbs": (bytes_fwd if dom == "render_bwd" else bytes_bwd) /
                                             (kernel_ms[other] * 1e-3) / 1e9},
        "stages_gbs_algorithmic": stages_gbs,
        # the SH forward (fill, list marks, dense pass) timed like the compositors, back to back: inside the
        # training step it also pays for the write-back of dirty cache lines the binning left behind
        "sh_fwd_back_to_back": {"ms": kernel_ms["sh_fwd"],
                                "gbs_algorithmic": stage_bytes["sh_fwd"] / (kernel_ms["sh_fwd"] * 1e-3) / 1e9},
        "step_gbs_algorithmic": step_bytes / (ms_per_step * 1e-3) / 1e9,
        "tile_entry_pixel_pairs_per_s_fwd": E * 256 / (kernel_ms["render_fwd"] * 1e-3),
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_large": large,
    }
    print(json.dumps(result), flush=True)

    if args.calib_copy_mib > 0:
        a = torch.ones(args.calib_copy_mib * 262144, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
