"""Benchmark: screen-mode NL-PF at 1920x1080 x 16 spp, depth 4 (5 vertex levels), 15x15
cross-bilateral window, 16-stop colour ramp + hue shift 0.1 fused into every level.

One step = one whole NL-PF pass (levels 4..0 + film) over 165,888,000 synthetic path
vertices already resident in HBM.  Prints one compact JSON line (rank 0): the contract fields,
the FP16 and FP16 + learned-cache results beside their controls, and the other configurations'
headline figures.  --detail PATH writes every measurement (window sweep, cache variants, world
sweep) as JSON.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-fp16] [--no-fp64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "practical-stylized_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "filtered path-vertices/s @1080p×16spp, plain vs FP16+MLP; HBM % of peak; L2 err"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic flops per neighbour pair: |dp|^2 8 (3 sub, mul, 2 fma), normal term 8 (3 fma,
# fma), luminance 3 (sub, fma), exp 1, accumulate 7 (add, 3 fma)
FLOPS_PER_PAIR = 27.0
FP32_NONPACKED_TFLOPS = 78.6  # one FP32 FMA per lane per cycle (the build emits no packed FP32)
VALU_ISSUE_PEAK = 256 * 4 * 16 * 2.4e9  # lane-instructions/s: 256 CUs x 4 SIMD16 at 2.4 GHz
# the FP32 level kernel each window radius runs by default (the library's per-radius choice:
# column strips with 2 or 3 rows per lane, 64x12 tiles at radius 1, 64x16 tiles at radius 5)
LEVEL_KERNEL_FP32 = {1: "screen_level_kernel<float,float,1,4,3>", 2: "screen_strip_kernel<float,float,2,4,2>",
                     3: "screen_strip_kernel<float,float,3,4,2>", 4: "screen_strip_kernel<float,float,4,4,2>",
                     5: "screen_level_kernel<float,float,5,4,4>", 6: "screen_strip_kernel<float,float,6,4,3>",
                     7: "screen_strip_kernel<float,float,7,4,2>"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def style(N):
    rng = np.random.default_rng(2024)
    pos = np.sort(rng.random(16)).astype(np.float32)
    val = np.sort(rng.random(16)).astype(np.float32)
    return N.style_table([N.ramp(pos, val), N.hue_shift(0.1, inner=0)]), 1


def level_bytes(S, plane, L, rec_b):
    """Algorithmic HBM bytes of each level launch: records once, child output once, output once."""
    out = []
    for d in range(L):
        child = 12 if d + 1 < L else 0
        out.append(S * plane * (rec_b + child + 12))
    return out


def timed_runs(h, desc, rec, anc, film, steps, warmup):
    for _ in range(warmup):
        h.screen_run(desc, rec, anc, film, None)
    h.sync()
    lv_ms = np.zeros(desc.levels)
    t0 = time.perf_counter()
    for _ in range(steps):
        h.screen_run(desc, rec, anc, film, None)
        st = h.stats()
        lv_ms += np.array(st["level_ms"][: desc.levels])
    h.sync()
    wall = time.perf_counter() - t0
    return wall / steps * 1e3, lv_ms / steps


def materials_leg(h, N, desc, rec, film, steps, warmup):
    """The FP32 pass through nlpf_screen_run_materials: every vertex stylized by its own material
    (uniform over four: the headline ramp16+hue tree, rex, a flags-selected EnableIf between
    saturation and toon, and none) with its own style flags (uniform over 0 and styles 2-4).
    Level kernels and bytes as the one-tree pass, plus the 3-byte material and flags planes per
    vertex. Restores nothing: the caller sets its style table again."""
    W, H, S, L = desc.width, desc.height, desc.spp, desc.levels
    n = W * H * S * L
    rng = np.random.default_rng(2024)
    pos = np.sort(rng.random(16)).astype(np.float32)
    val = np.sort(rng.random(16)).astype(np.float32)
    table = N.style_table([N.ramp(pos, val), N.hue_shift(0.1, inner=0), N.rex(0.6), N.saturation(0.5),
                           N.toon(0.25), N.enable_if(2, inner=3, else_=4)])
    h.set_style(table, [1, 2, 5, -1])
    rng = np.random.default_rng(5)
    material = h.to_device(rng.integers(0, 4, n, dtype=np.uint16))
    flags = h.to_device(rng.choice(np.array([0, N.FLAG_STYLE2, N.FLAG_STYLE3, N.FLAG_STYLE4], np.uint8), n))
    for _ in range(warmup):
        h.screen_run_materials(desc, rec, None, material, flags, film)
    h.sync()
    lv_ms = np.zeros(L)
    t0 = time.perf_counter()
    for _ in range(steps):
        h.screen_run_materials(desc, rec, None, material, flags, film)
        lv_ms += np.array(h.stats()["level_ms"][:L])
    h.sync()
    wall = time.perf_counter() - t0
    material.free()
    flags.free()
    return {"ms_per_step": wall / steps * 1e3, "level_ms": (lv_ms / steps).round(3).tolist(),
            "materials": 4, "flags": [0, 2, 3, 4]}


def window_sweep(h, N, W, H, S, L, recs, steps=3, warmup=1):
    """Level-kernel rate for every window radius 1..7 on the same records: per-pair work falls
    with the square of the window while the bytes per vertex stay fixed, so the small windows
    show how close the kernel's memory path comes to the HBM roof. recs: {"fp32": (rec,
    None), "fp16": (rec, anchors)}."""
    film = h.alloc(W * H * 12)
    plane = W * H
    out = {}
    for name, (rec, anc) in recs.items():
        prec, rec_b = (N.FP32, 60) if name == "fp32" else (N.FP16, 30)
        rows = []
        for r in range(1, 8):
            desc = h.screen_desc(W, H, S, L, prec, r)
            _, lv = timed_runs(h, desc, rec, anc, film, steps, warmup)
            t = lv.sum() / 1e3
            gbs = sum(level_bytes(S, plane, L, rec_b)) / 1e9 / t
            tfl = plane * S * L * (2 * r + 1) ** 2 * FLOPS_PER_PAIR / t / 1e12
            rows.append({"radius": r, "window": f"{2 * r + 1}x{2 * r + 1}", "level_ms_sum": lv.sum(),
                         "vertices_per_s": plane * S * L / t, "hbm_gbs_algorithmic": gbs,
                         "hbm_frac": gbs / HBM_PEAK_GBS, "tflops": tfl,
                         "fp32_nonpacked_frac": tfl / FP32_NONPACKED_TFLOPS})
            log(f"  {name} radius {r}: {lv.sum():.2f} ms, {gbs:.0f} GB/s, {tfl:.1f} TFLOP/s")
        out[name] = rows
    film.free()
    return out


def cpu_baseline(h, rec, W, H, S, L, radius, sig, table, root, slices):
    """FP32 OpenMP restatement (oracle/) on the first `slices` spp slices of the same records."""
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    plane = W * H
    host = rec.download(np.float32, slices * L * 15 * plane).reshape(slices, L, 15, H, W)
    t0 = time.perf_counter()
    O.screen_run(host, W, H, slices, L, radius, sig, nodes=table, root=root, precision="f32", threads=threads)
    dt = time.perf_counter() - t0
    verts = plane * L * slices
    return {"value": verts / dt, "unit": "vertices/s", "cores": threads, "kind": "port",
            "sample": f"{slices} of {S} spp slices ({W}x{H}x{L} levels x {slices} = {verts} vertices), "
                      f"FP32 OpenMP restatement of the same filter, {threads} threads, {dt:.1f} s"}


def measured_traffic():
    """HBM bytes per launch of the level kernel from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "screen_level_fp32_1080p16spp_hbm_counters.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_valu():
    """VALU instructions and FP32 flops per neighbour pair of the level kernel from the committed
    rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "screen_level_fp32_1080p16spp_valu_counters.md")
    out = {}
    try:
        for line in open(path):
            for key in ("valu_instructions_per_pair", "fp32_flops_per_pair"):
                if line.startswith(key + ":"):
                    out[key] = float(line.split(":")[1])
    except OSError:
        pass
    return out or None


def mlp_leg(h, args, desc, rec, anc, film, film_ref, verts, S):
    """FP16 records + learned cache (2^16 training records per frame, batch 2^14, Adam lr 1e-2).
    The audit slice rotates over the spp slices from frame to frame; `--mlp-frames` untimed
    frames train a fresh cache first.  At each cache depth (the first is the headline), with the
    per-tile fallback at `--tile-threshold` and without fallback, the trained network and the
    two controls run through the same pass: "zero" cuts the paths at the cache depth (what the
    pass costs and gives without a network), "mean" puts the audit slice's mean filtered value
    everywhere (the best estimator that ignores the vertex).  Returns a list of legs."""
    legs = []
    for depth in args.cache_depths:
        log(f"FP16 + learned cache at depth {depth}")
        h.mlp_init(1)
        h.mlp_set_ema(args.ema)
        kw = dict(depth=depth, train_records=1 << 16, batch=1 << 14, lr=1e-2)
        frame = 0

        def frames(n, thr, est):
            nonlocal frame
            acc, st = {}, None
            h.sync()
            t0 = time.perf_counter()
            for _ in range(n):
                st = h.screen_run_mlp(desc, rec, anc, film, None, audit_slice=frame % S, seed=frame,
                                      tile_threshold=thr, estimator=est, **kw)
                frame += 1
                for k in ("audit_ms", "train_ms", "infer_ms", "fallback_ms", "filter_ms", "fallback_tiles"):
                    acc[k] = acc.get(k, 0.0) + st[k]
            h.sync()
            ms = (time.perf_counter() - t0) / max(n, 1) * 1e3
            return ms, {k: v / max(n, 1) for k, v in acc.items()}, st

        frames(args.mlp_frames, 1e30, "network")
        for thr in (args.tile_threshold, 1e30):
            for est in ("network", "zero", "mean"):
                frames(args.warmup, thr, est)
                ms, ph, st = frames(args.steps, thr, est)
                leg = {"estimator": est, "cache_depth": depth, "tile_threshold": thr if thr < 1e29 else None,
                       "value": verts / (ms / 1e3), "unit": "vertices/s", "ms_per_step": ms,
                       "phases_ms": {k[:-3]: ph[k] for k in ph if k.endswith("_ms")},
                       "fallback_tile_frac": ph["fallback_tiles"] / st["tiles"],
                       "audit_rel_l2": st["audit_rel_l2"], "cache_rel_l2": st["cache_rel_l2"],
                       "train_loss": st["train_loss"] if est == "network" else None}
                if film_ref is not None:
                    # film error over several frames (the cache keeps training, so it varies a little)
                    errs = []
                    for _ in range(max(args.quality_frames, 1)):
                        h.screen_run_mlp(desc, rec, anc, film, None, audit_slice=frame % S, seed=frame,
                                         tile_threshold=thr, estimator=est, **kw)
                        frame += 1
                        f = film.download(np.float32).astype(np.float64)
                        errs.append(float(np.linalg.norm(f - film_ref) / np.linalg.norm(film_ref)))
                    leg["l2_err_vs_fp64"] = float(np.mean(errs))
                    leg["l2_err_vs_fp64_max"] = float(np.max(errs))
                    leg["quality_frames"] = len(errs)
                log(f"  depth {depth} threshold {leg['tile_threshold']} {est}: {ms:.2f} ms/step, phases "
                    f"{np.round(list(leg['phases_ms'].values()), 2).tolist()}, fallback "
                    f"{leg['fallback_tile_frac']:.3f}, cache fit {st['cache_rel_l2']:.3f}, "
                    f"film L2 {leg.get('l2_err_vs_fp64')}")
                legs.append(leg)
    return legs


def mlp_summary(legs, args, ms32):
    """The bench line's compact view of the cache legs: the headline (first depth, fallback on),
    its controls, and the same at every other depth."""
    def pick(depth, thr, est):
        for g in legs:
            if g["cache_depth"] == depth and g["tile_threshold"] == thr and g["estimator"] == est:
                return g
        return None

    def brief(g):
        return None if g is None else {"ms": round(g["ms_per_step"], 3), "film_l2": g.get("l2_err_vs_fp64"),
                                       "fallback": round(g["fallback_tile_frac"], 4)}

    d0, thr = args.cache_depths[0], args.tile_threshold
    head = pick(d0, thr, "network")
    out = {"value": head["value"], "unit": "vertices/s", "ms_per_step": head["ms_per_step"],
           "vs_plain_fp32": ms32 / head["ms_per_step"], "film_l2_vs_fp64": head.get("l2_err_vs_fp64"),
           "cache_depth": d0, "tile_threshold": thr, "fallback_tile_frac": head["fallback_tile_frac"],
           "cache_rel_l2": head["cache_rel_l2"],
           "train": {"frames_before_timing": args.mlp_frames, "records_per_frame": 1 << 16, "batch": 1 << 14,
                     "lr": 1e-2, "ema_decay": args.ema}}
    for d in args.cache_depths:
        out[f"depth{d}"] = {f"{est}{'' if t is not None else '_nofallback'}": brief(pick(d, t, est))
                            for t in (thr, None) for est in ("network", "zero", "mean")}
    return out


def adjoint_leg(h, N):
    """Config 3: forward + adjoint at 960x540 x 8 spp, 4 levels, 7x7 window; 19 parameters
    (3 bandwidths + 16 ramp-stop values), L2 loss to a target rendered at perturbed parameters."""
    W, H, S, L, R = 960, 540, 8, 4, 3
    log("adjoint (config 3)")
    rng = np.random.default_rng(3)
    pos = np.sort(rng.random(16)).astype(np.float32)
    val = np.sort(rng.random(16)).astype(np.float32)
    table = N.style_table([N.ramp(pos, val), N.hue_shift(0.1, inner=0)])
    sig = np.array([0.05, 0.9, 0.5])
    desc = h.screen_desc(W, H, S, L, N.FP32, R)
    rec = h.alloc(W * H * S * L * 60)
    h.generate_screen_records(desc, rec, seed=3)
    film = h.alloc(W * H * 12)
    tt = table.copy()
    tt[0]["p"][16:32] = (tt[0]["p"][16:32] + rng.normal(0, 0.05, 16)).astype(np.float32)
    h.set_style(tt, [1])
    h.set_sigmas(*(sig * np.exp(rng.normal(0, 0.05, 3))))
    h.screen_run(desc, rec, None, film, None)
    target = film.download(np.float32)
    h.set_style(table, [1])
    h.set_sigmas(*sig)
    lv = h.alloc(S * L * 3 * W * H * 4)
    grad = h.alloc(19 * 4)
    dfilm = h.alloc(W * H * 12)
    fwd, bwd = [], []
    for it in range(4):
        h.sync()
        t0 = time.perf_counter()
        h.screen_run(desc, rec, None, film, lv)
        h.sync()
        t1 = time.perf_counter()
        f = film.download(np.float32)
        dfilm.upload((2 * (f - target)).astype(np.float32))
        h.sync()
        t2 = time.perf_counter()
        h.screen_backward(desc, rec, lv, dfilm, grad)
        h.sync()
        t3 = time.perf_counter()
        if it:
            fwd.append(t1 - t0)
            bwd.append(t3 - t2)
    g = grad.download(np.float32, 19)
    for b in (rec, film, lv, grad, dfilm):
        b.free()
    verts = W * H * S * L
    return {"workload": f"{W}x{H}x{S}spp, {L} levels, {2 * R + 1}x{2 * R + 1} window, ramp16+hue0.1, 19 parameters",
            "forward_ms": 1e3 * float(np.mean(fwd)), "backward_ms": 1e3 * float(np.mean(bwd)),
            "value": verts / float(np.mean(fwd) + np.mean(bwd)), "unit": "vertices/s (forward + adjoint)",
            "grad_finite": bool(np.isfinite(g).all())}


def large_leg(h, N):
    """Config 5: 3840x2160 x 64 spp, depth 6 (7 levels), 15x15 window, toon (6 levels) at depth 1
    wrapping the 16-stop ramp at depth 3; records generated and filtered in 16-spp chunks."""
    W, H, S, L, R, chunks = 3840, 2160, 16, 7, 7, 4
    rng = np.random.default_rng(2024)
    pos = np.sort(rng.random(16)).astype(np.float32)
    val = np.sort(rng.random(16)).astype(np.float32)
    table = N.style_table([N.ramp(pos, val, depth_mask=N.depth_mask(3)),
                           N.toon(1.0 / 6.0, inner=0, depth_mask=N.depth_mask(1))])
    h.set_style(table, [1])
    plane = W * H
    verts = plane * S * L
    out = {"workload": f"{W}x{H}x{S * chunks}spp in {chunks} chunks of {S} spp, {L} levels, 15x15 window, "
                       "toon(6) at depth 1 over ramp16 at depth 3"}
    for name, prec, rb in (("fp32", N.FP32, 60), ("fp16", N.FP16, 30)):
        log(f"config 5 {name}")
        desc = h.screen_desc(W, H, S, L, prec, R)
        rec = h.alloc(verts * rb)
        anc = h.alloc(S * L * ((W + 7) // 8) * ((H + 7) // 8) * 12) if prec == N.FP16 else None
        film = h.alloc(plane * 12)
        acc = np.zeros(plane * 3)
        total = 0.0
        for c in range(chunks):
            h.generate_screen_records(desc, rec, anc, seed=100 + c)
            h.sync()
            t0 = time.perf_counter()
            h.screen_run(desc, rec, anc, film, None)
            h.sync()
            total += time.perf_counter() - t0
            acc += film.download(np.float32)
        for b in (rec, anc, film):
            if b is not None:
                b.free()
        bytes_alg = chunks * S * plane * ((rb + 12 + 12) * (L - 1) + rb + 12)
        out[name] = {"ms_per_frame": total * 1e3, "value": chunks * verts / total, "unit": "vertices/s",
                     "hbm_gbs_algorithmic": bytes_alg / total / 1e9,
                     "hbm_frac": bytes_alg / total / 1e9 / HBM_PEAK_GBS,
                     "film_finite": bool(np.isfinite(acc).all())}
        log(f"  {name}: {total * 1e3:.1f} ms per 64-spp frame")
    return out


def world_cpu(host, W, H, B, L, cfg, nodes, roots, film, pts, bsdfs=None):
    """The FP64 world-mode oracle (OpenMP over reservoir slots; grid build, sorts and
    propagation single-threaded) on the whole frame's records, beside the GPU pass, and the
    GPU film's relative L2 error against it."""
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    ofilm = O.world_run(host["records"], host["flags"], host["new_flags"], host["material"], host["path_depth"],
                        W, H, B, L, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), cfg, nodes=nodes, roots=roots,
                        bsdfs=bsdfs)
    dt = time.perf_counter() - t0
    err = float(np.linalg.norm(film.astype(np.float64) - ofilm) / max(np.linalg.norm(ofilm), 1e-30))
    return {"value": pts / dt, "unit": "vertices/s", "cores": threads, "kind": "port",
            "sample": f"the whole frame ({pts} vertices), FP64 oracle, {threads} threads, {dt:.1f} s",
            "film_rel_l2_fp32_gpu_vs_fp64": err}


def world_leg(h_unused, N, cpu=True):
    """Config 1 in world mode (the reference's hash-grid NL-PF): 256x256 x 4 paths per pixel,
    max_depth 3, SMIS, no clusters, reuse cap above every neighbour count, 4-level toon at
    depth 1, records traced on the GPU from the procedural box scene."""
    W, H, B, L = 256, 256, 4, 4
    cfg = {"type": "pf", "max_depth": L - 1, "branch": B, "smis": True, "cluster": False, "reuse": [1 << 16]}
    nodes = N.style_table([N.toon(0.25, depth_mask=N.depth_mask(1))])
    roots = [0, 0, 0, 0, 0, 0, 0, -1]
    with N.Handle(cfg) as hw:
        hw.set_style(nodes, roots)
        desc = hw.world_desc(W, H, B, L)
        bufs = hw.alloc_world(desc)
        hw.generate_world_records(desc, bufs, seed=1)
        hw.world_run(desc, bufs)
        times = []
        for _ in range(3):
            film, _ = hw.world_run(desc, bufs)
            st = hw.stats()
            times.append(st["total_ms"] / 1e3)  # HIP events around the depth loop and the film
        pts = sum(st["points"][:L])
        nb = sum(st["neighbors"][:L])
        ms = 1e3 * float(np.median(times))
        out = {"workload": f"{W}x{H}x{B} paths, {L} levels, radius_scale default, SMIS, toon(4) at depth 1",
               "ms_per_frame": ms, "vertices": int(pts), "value": pts / (ms / 1e3), "unit": "vertices/s",
               "neighbours_per_vertex": nb / max(pts, 1)}
        host = hw.download_world(desc, bufs) if cpu else None
    # the same records under the reference's default "pf" settings: clustering on, reuse
    # [128, 64], default radius
    with N.Handle({"type": "pf", "max_depth": L - 1, "branch": B}) as hw:
        hw.set_style(nodes, roots)
        desc = hw.world_desc(W, H, B, L)
        bufs = hw.alloc_world(desc)
        hw.generate_world_records(desc, bufs, seed=1)
        hw.world_run(desc, bufs)
        times = []
        for _ in range(3):
            hw.world_run(desc, bufs)
            times.append(hw.stats()["total_ms"])
        st = hw.stats()
        ms_d = float(np.median(times))
        pts_d = sum(st["points"][:L])
        out["reference_defaults"] = {"config": "cluster, reuse [128, 64], radius_scale [0.1, 0.05], SMIS",
                                     "ms_per_frame": ms_d, "value": pts_d / (ms_d / 1e3), "unit": "vertices/s",
                                     "neighbours_per_vertex": sum(st["neighbors"][:L]) / max(pts_d, 1)}
    # the same scene with principled materials: metal walls and spheres (roughness 0 to 0.5,
    # as the reference scene's Gold, Hinge and Pot materials), a diffuse / metal floor; the
    # records are traced with the metal lobes and filtered by the glossy kernels
    table = N.bsdf_table([
        {"kind": N.BSDF_PRINCIPLED, "metallic": 0.5, "roughness": 0.4},
        {},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.3, "specular_tint": (1.0, 0.8, 0.6)},
        {},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.12},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.0},
        {"kind": N.BSDF_PRINCIPLED, "metallic": 1.0, "roughness": 0.5},
    ])
    gl = {}
    for name, c in (("default_radius", cfg), ("reference_defaults", {"type": "pf", "max_depth": L - 1, "branch": B})):
        with N.Handle(c) as hw:
            hw.set_style(nodes, roots)
            hw.set_bsdfs(table)
            desc = hw.world_desc(W, H, B, L)
            bufs = hw.alloc_world(desc)
            hw.generate_world_records(desc, bufs, seed=1)
            hw.world_run(desc, bufs)
            times = []
            for _ in range(3):
                gfilm, _ = hw.world_run(desc, bufs)
                times.append(hw.stats()["total_ms"])
            st = hw.stats()
            gms = float(np.median(times))
            gpts = sum(st["points"][:L])
            gl[name] = {"ms_per_frame": gms, "value": gpts / (gms / 1e3), "unit": "vertices/s",
                        "neighbours_per_vertex": sum(st["neighbors"][:L]) / max(gpts, 1)}
            if cpu and name == "default_radius":
                log("world mode, principled materials: FP64 oracle on the same frame (CPU)")
                ghost = hw.download_world(desc, bufs)
                gl[name]["cpu_oracle"] = world_cpu(ghost, W, H, B, L, c, nodes, roots, gfilm, gpts, bsdfs=table)
    gl["materials"] = "floor metallic 0.5 / roughness 0.4, back, right, spheres metal (roughness 0.3, 0.12, 0, 0.5)"
    out["principled"] = gl
    # clustering on and off at three radii under the reference's default reuse [128, 64]
    sweep = []
    for rs in (None, 2.0, 4.0):
        for cluster in (True, False):
            c = {"type": "pf", "max_depth": L - 1, "branch": B, "cluster": cluster}
            if rs is not None:
                c["radius_scale"] = [rs]
            with N.Handle(c) as hw:
                hw.set_style(nodes, roots)
                desc = hw.world_desc(W, H, B, L)
                bufs = hw.alloc_world(desc)
                hw.generate_world_records(desc, bufs, seed=1)
                hw.world_run(desc, bufs)
                times = []
                for _ in range(2):
                    hw.world_run(desc, bufs)
                    times.append(hw.stats()["total_ms"])
                st = hw.stats()
                sweep.append({"radius_scale": rs or "default", "cluster": cluster, "ms_per_frame": min(times),
                              "neighbours_per_vertex": sum(st["neighbors"][:L]) / max(sum(st["points"][:L]), 1),
                              "smis_pairs": sum(st["smis_pairs"][:L])})
    out["cluster_sweep"] = sweep
    if cpu:
        log("world mode: FP64 oracle on the same frame (CPU)")
        out["cpu_oracle"] = world_cpu(host, W, H, B, L, cfg, nodes, roots, film, pts)
        out["cpu_oracle"]["gpu_speedup"] = out["value"] / out["cpu_oracle"]["value"]
    # the same records with a 20x larger reuse radius: the filter itself dominates
    cfg2 = dict(cfg, radius_scale=[2.0])
    with N.Handle(cfg2) as hw:
        hw.set_style(nodes, roots)
        desc = hw.world_desc(W, H, B, L)
        bufs = hw.alloc_world(desc)
        hw.generate_world_records(desc, bufs, seed=1)
        hw.world_run(desc, bufs)
        times = []
        for _ in range(3):
            film, _ = hw.world_run(desc, bufs)
            st = hw.stats()
            times.append(st["total_ms"] / 1e3)
        pts = sum(st["points"][:L])
        nb = sum(st["neighbors"][:L])
        pairs = sum(st["smis_pairs"][:L])
        ms = 1e3 * float(np.median(times))
        host = hw.download_world(desc, bufs) if cpu else None
        # the O(K^2) SMIS denominators dominate: their pdf evaluations per second, and the VALU
        # issue they imply at 17 instructions per evaluation (inner loop of w_filter_wave_kernel
        # for records with equal shading and geometric normals, as every record of this scene
        # has; 25 otherwise), over the whole frame time
        out["radius_scale_2"] = {"ms_per_frame": ms, "value": pts / (ms / 1e3), "unit": "vertices/s",
                                 "neighbours_per_vertex": nb / max(pts, 1),
                                 "smis_pairs_per_frame": pairs, "smis_pairs_per_s": pairs / (ms / 1e3),
                                 "smis_valu_issue_frac": pairs * 17.0 / (ms / 1e3) / VALU_ISSUE_PEAK}
    if cpu:
        log("world mode, radius_scale 2: FP64 oracle on the same frame (CPU)")
        o = world_cpu(host, W, H, B, L, cfg2, nodes, roots, film, pts)
        o["gpu_speedup"] = out["radius_scale_2"]["value"] / o["value"]
        out["radius_scale_2"]["cpu_oracle"] = o
    return out


def pcie_upload_gbs(h, nbytes=1 << 30):
    """Host-to-device rate of nlpf_upload from pageable host memory (1 GiB)."""
    host = np.ones(nbytes // 4, np.float32)
    dev = h.alloc(nbytes)
    dev.upload(host)
    h.sync()
    t0 = time.perf_counter()
    dev.upload(host)
    h.sync()
    dt = time.perf_counter() - t0
    dev.free()
    return nbytes / dt / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--max-depth", type=int, default=4)
    ap.add_argument("--radius", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=16, help="spp slices timed by the CPU baseline")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--cache-depths", type=int, nargs="+", default=[3, 2],
                    help="cache depths measured (the first is the headline fp16_mlp entry)")
    ap.add_argument("--mlp-frames", type=int, default=32, help="untimed training frames before the timed ones")
    ap.add_argument("--tile-threshold", type=float, default=0.05)
    ap.add_argument("--ema", type=float, default=0.95, help="moving-average decay of the cache's inference weights")
    ap.add_argument("--quality-frames", type=int, default=8,
                    help="untimed frames after the timed ones whose film error vs FP64 is averaged")
    ap.add_argument("--no-extra", action="store_true", help="skip the config 3 / 5 / world-mode measurements")
    ap.add_argument("--no-sweep", action="store_true", help="skip the window-radius sweep of the level kernel")
    ap.add_argument("--world-only", action="store_true",
                    help="only the world-mode measurement (config 1), printed as its own JSON line")
    ap.add_argument("--detail", default=None, help="write every measurement to this JSON file")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one FP32 pass and one 1 GiB device copy, for rocprofv3 counter runs")
    ap.add_argument("--screen-cfg", default="{}",
                    help="JSON object merged into the screen-mode handle's config (e.g. screen_kernel)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "screen-mode NL-PF shards by spp slice with one film all-reduce, which is not built here.",
              file=sys.stderr)
        sys.exit(2)

    import nlpf as N

    if args.world_only:
        print(json.dumps({"world_config1": world_leg(None, N, cpu=not args.no_cpu)}), flush=True)
        return

    W, H, S, L, R = args.width, args.height, args.spp, args.max_depth + 1, args.radius
    plane = W * H
    sig = (0.05, 0.9, 0.5)
    table, root = style(N)
    h = N.Handle({"type": "pf", "mode": "screen", "max_depth": L - 1, "pixel_radius": R,
                  **json.loads(args.screen_cfg)})
    h.set_style(table, [root])
    h.set_sigmas(*sig)
    verts = plane * S * L

    log(f"generating {verts} FP32 records ({verts * 60 / 1e9:.1f} GB)")
    d32 = h.screen_desc(W, H, S, L, N.FP32, R)
    rec32 = h.alloc(verts * 60)
    h.generate_screen_records(d32, rec32, seed=1)
    film32 = h.alloc(plane * 12)

    if args.pmc_probe:
        h.screen_run(d32, rec32, None, film32, None)
        src = h.alloc(1 << 30)
        dst = h.alloc(1 << 30)
        h.copy(dst, src, 1 << 30)
        h.sync()
        log("pmc probe: one FP32 pass (levels 4..0 + film) and one 1 GiB device-to-device copy")
        h.close()
        return

    log("timing FP32")
    ms32, lv32 = timed_runs(h, d32, rec32, None, film32, args.steps, args.warmup)
    log(f"FP32: {ms32:.2f} ms/step, levels {np.round(lv32, 2).tolist()} ms")
    log("timing FP32 with per-vertex materials and flags")
    mats = materials_leg(h, N, d32, rec32, film32, args.steps, args.warmup)
    h.set_style(table, [root])
    log(f"FP32 per-vertex materials: {mats['ms_per_step']:.2f} ms/step")
    lb32 = level_bytes(S, plane, L, 60)
    achieved_gbs = sum(lb32) / 1e9 / (lv32.sum() / 1e3)
    pairs = verts * (2 * R + 1) ** 2
    tflops = pairs * FLOPS_PER_PAIR / (lv32.sum() / 1e3) / 1e12

    counters = measured_valu() if (W, H, S, L, R) == (1920, 1080, 16, 5, 7) else None
    pair_rate = pairs / (lv32.sum() / 1e3)
    result = {
        "metric": METRIC,
        "value": verts / (ms32 / 1e3),
        "unit": "vertices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"screen NL-PF {W}x{H}x{S}spp depth {L - 1} ({L} levels), "
                               f"{2 * R + 1}x{2 * R + 1} window, ramp16+hue0.1, FP32 records",
                   "vertices_per_step": verts},
        # the 15x15 gather is compute-bound (~75 flop per algorithmic byte): the binding roof
        # is FP32 VALU; the HBM figures stand beside it
        "roofline": {"bound": "valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP32_PEAK_TFLOPS,
                     "traffic": measured_traffic() if (W, H, S, L, R) == (1920, 1080, 16, 5, 7) else None,
                     "kernel": LEVEL_KERNEL_FP32[R],
                     "avg_launch_ms": float(lv32.mean()),
                     "algorithmic_flops_per_launch": pairs / L * FLOPS_PER_PAIR,
                     "algorithmic_bytes_per_launch": float(np.mean(lb32)),
                     "frac_of_nonpacked_fp32_peak": tflops / FP32_NONPACKED_TFLOPS,
                     "hbm_gbs": achieved_gbs, "hbm_frac": achieved_gbs / HBM_PEAK_GBS},
    }
    if counters:
        # VALU instructions per pair from the committed rocprofv3 counter run, at the measured pair rate
        result["roofline"]["valu_issue"] = {
            "valu_per_pair": counters.get("valu_instructions_per_pair"),
            "frac": pair_rate * counters.get("valu_instructions_per_pair", 0) / VALU_ISSUE_PEAK}
    result["screen_materials"] = {"ms_per_step": round(mats["ms_per_step"], 3),
                                  "vs_one_tree": round(mats["ms_per_step"] / ms32, 4)}
    detail = {"roofline": dict(result["roofline"]), "level_ms_fp32": lv32.tolist(), "screen_materials": mats}

    h2d = pcie_upload_gbs(h)
    rec_bytes = verts * 60
    detail["pcie_inclusive"] = {"h2d_gbs": h2d, "record_bytes": rec_bytes,
                                "value": verts / (ms32 / 1e3 + rec_bytes / (h2d * 1e9)), "unit": "vertices/s",
                                "note": "FP32 records uploaded from pageable host memory each frame"}
    log(f"host->device {h2d:.1f} GB/s")

    film_ref = None
    if not args.no_fp64:
        log("FP64 GPU reference pass (error oracle)")
        rec64 = h.alloc(verts * 120)
        # widen FP32 -> FP64 slice by slice through the host
        slice_n = L * 15 * plane
        for s in range(S):
            part = rec32.download(np.float32, slice_n, offset=s * slice_n * 4)
            rec64.upload(part.astype(np.float64), offset=s * slice_n * 8)
        d64 = h.screen_desc(W, H, S, L, N.FP64, R)
        film64 = h.alloc(plane * 24)
        t0 = time.perf_counter()
        h.screen_run(d64, rec64, None, film64, None)
        log(f"FP64 pass {time.perf_counter() - t0:.2f} s")
        film_ref = film64.download(np.float64)
        rec64.free()
        film32_host = film32.download(np.float32).astype(np.float64)
        result["l2_err_fp32_vs_fp64"] = float(np.linalg.norm(film32_host - film_ref) / np.linalg.norm(film_ref))

    if not args.no_fp16:
        log("FP16 records")
        d16 = h.screen_desc(W, H, S, L, N.FP16, R)
        rec16 = h.alloc(verts * 30)
        anc = h.alloc(S * L * ((W + 7) // 8) * ((H + 7) // 8) * 12)
        h.convert_records_fp16(d32, rec32, rec16, anc)
        film16 = h.alloc(plane * 12)
        ms16, lv16 = timed_runs(h, d16, rec16, anc, film16, args.steps, args.warmup)
        lb16 = level_bytes(S, plane, L, 30)
        fp16 = {"value": verts / (ms16 / 1e3), "ms_per_step": ms16, "hbm_gbs": sum(lb16) / 1e9 / (lv16.sum() / 1e3)}
        if film_ref is not None:
            f16 = film16.download(np.float32).astype(np.float64)
            fp16["film_l2_vs_fp64"] = float(np.linalg.norm(f16 - film_ref) / np.linalg.norm(film_ref))
        result["fp16"] = fp16
        if not args.no_mlp:
            legs = mlp_leg(h, args, d16, rec16, anc, film16, film_ref, verts, S)
            detail["fp16_mlp_legs"] = legs
            result["fp16_mlp"] = mlp_summary(legs, args, ms32)
        if not args.no_sweep:
            log("window sweep, radius 1..7")
            sweep = window_sweep(h, N, W, H, S, L, {"fp32": (rec32, None), "fp16": (rec16, anc)})
            detail["window_sweep"] = sweep
            result["window_sweep_ms"] = {name: [round(r["level_ms_sum"], 3) for r in rows]
                                         for name, rows in sweep.items()}
        rec16.free()
        anc.free()

    if not args.no_extra:
        rec32.free()
        adj = adjoint_leg(h, N)
        big = large_leg(h, N)
        wl = world_leg(h, N, cpu=not args.no_cpu)
        detail.update({"adjoint_config3": adj, "large_config5": big, "world_config1": wl})
        result["config3_adjoint"] = {"forward_ms": adj["forward_ms"], "adjoint_ms": adj["backward_ms"]}
        result["config5"] = {k: {"ms_per_frame": big[k]["ms_per_frame"], "value": big[k]["value"]}
                             for k in ("fp32", "fp16")}
        result["world_config1"] = {"ms_per_frame": wl["ms_per_frame"], "value": wl["value"],
                                   "reference_defaults_ms": wl["reference_defaults"]["ms_per_frame"],
                                   "radius_scale_2_ms": wl["radius_scale_2"]["ms_per_frame"]}
        if "cpu_oracle" in wl:
            result["world_config1"]["film_l2_vs_fp64_oracle"] = wl["cpu_oracle"]["film_rel_l2_fp32_gpu_vs_fp64"]
        pr = wl["principled"]
        result["world_config1"]["principled"] = {"ms": pr["default_radius"]["ms_per_frame"],
                                                 "reference_defaults_ms": pr["reference_defaults"]["ms_per_frame"]}
        if "cpu_oracle" in pr["default_radius"]:
            result["world_config1"]["principled"]["film_l2_vs_fp64_oracle"] = \
                pr["default_radius"]["cpu_oracle"]["film_rel_l2_fp32_gpu_vs_fp64"]
        h.set_style(table, [root])
        rec32 = h.alloc(verts * 60)
        h.generate_screen_records(d32, rec32, seed=1)

    if not args.no_cpu:
        log("CPU baseline (oracle FP32 restatement)")
        result["cpu_baseline"] = cpu_baseline(h, rec32, W, H, S, L, R, sig, table, root,
                                              max(1, min(args.cpu_slices, S)))
        result["cpu_baseline"]["gpu_speedup"] = result["value"] / result["cpu_baseline"]["value"]
    else:
        result["cpu_baseline"] = None
    h.close()
    if args.detail:
        detail["line"] = result
        with open(args.detail, "w") as f:
            json.dump(detail, f, indent=1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
