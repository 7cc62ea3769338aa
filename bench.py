#!/usr/bin/env python3
"""Benchmark of the closest-hit hot path on one MI355X.

Workload (default, --workload c1): BVH microbenchmark -- 1M random triangles
in the unit cube (centroids uniform, edge lengths log-normal(ln 0.003, 0.5),
random orientation, seed 7) in a binned-SAH BVH with leaves <= 4; 16M rays:
8M coherent primary rays of a 4096x2048 pinhole at (0.5, 0.5, -1) with a 60
degree vertical field of view, plus 8M incoherent rays (origins uniform in
the unit cube, cosine-hemisphere directions about uniform random normals).
One step = one rt_trace_closest launch over all 16M rays with inputs already
resident in HBM.  Prints one JSON line (rank 0).

The CPU baseline is the oracle's restatement of the reference's own
closest-hit path (Morton-pairing BVH with leaf size 1 and its unordered
stack traversal, cuda_libraries/BVH.cu:355-448) on a bounded sample of the
same rays, run on the host's cores.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s closest-hit; time-to-RMSE≤1e-2 vs FP64 oracle (cache vs plain)"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_c1(xyz, idx, rays_np, budget_s=12.0):
    """Reference closest-hit algorithm restated in C (oracle), timed on the
    host over growing samples until the budget is used."""
    import numpy as np
    import oracle

    nthreads = cpu_threads()
    t0 = time.time()
    ref = oracle.RefBvh(xyz, idx)
    build_s = time.time() - t0
    n_all = rays_np[0].shape[0]
    half = n_all // 2
    rng = np.random.default_rng(1)
    total_rays, total_s = 0, 0.0
    m = 4096
    while total_s < budget_s and m <= (1 << 20):
        sel = np.concatenate([rng.choice(half, m // 2, replace=False), half + rng.choice(n_all - half, m // 2, replace=False)])
        sub = [np.ascontiguousarray(a[sel]) for a in rays_np]
        t1 = time.time()
        ref.trace(sub, nthreads=nthreads)
        dt = time.time() - t1
        total_rays += m
        total_s += dt
        m *= 2
    return {
        "value": round(total_rays / total_s / 1e6, 4),
        "unit": "Mrays/s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"{total_rays} rays (half coherent, half incoherent, drawn from the same 16M) through the "
                   f"reference's Morton-pairing BVH restated in C, {total_s:.1f} s of tracing; tree build "
                   f"{build_s:.1f} s not included"),
    }


def run_c1(args):
    import numpy as np
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    dev = torch.device("cuda", 0)
    log("generating 1M-triangle soup (seed 7)")
    xyz, idx = rt.gen_soup(1 << 20, seed=7)
    log("generating 8M coherent + 8M incoherent rays")
    rays_np = rt.gen_rays(4096, 2048, eye=(0.5, 0.5, -1.0), vfov=60.0, n_incoherent=8 << 20, seed=7)
    n = rays_np[0].shape[0]
    tmax_np = np.full(n, np.inf, np.float32)
    ctx = rt.Context(0)
    t0 = time.time()
    bvh = rt.Bvh(ctx, xyz, idx, max_leaf=4, bins=32, width=args.width)
    inf = bvh.info()
    log(f"BVH: {inf.num_nodes} nodes, {inf.num_leaves} leaves, depth {inf.max_depth}, "
        f"build {inf.build_ms:.0f} ms (wall {time.time() - t0:.1f} s)")
    rays = [torch.from_numpy(a).to(dev) for a in rays_np]
    tmax = torch.from_numpy(tmax_np).to(dev)
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    t = torch.empty(n, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    s = stream.cuda_stream

    # Algorithmic bytes of one launch: rays in (o, d, tMax = 28 B) + hits out
    # (8 B) + every node (64 B) and triangle record (48 B) the traversal reads.
    nn, ntr = bvh.trace_stats(*rays, tmax=tmax)
    alg_bytes = n * 36 + 64 * nn + 48 * ntr
    log(f"per ray: {nn / n:.1f} node visits, {ntr / n:.2f} triangle tests")

    for _ in range(args.warmup):
        bvh.trace_closest(*rays, tmax=tmax, prim=prim, t=t, stream=s)
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bvh.trace_closest(*rays, tmax=tmax, prim=prim, t=t, stream=s)
        kms.append(bvh.last_kernel_ms())  # waits on this launch's end event
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    k_avg = sum(kms) / len(kms)
    mrays = n * args.steps / wall / 1e6
    hit_rate = float((prim >= 0).float().mean().item())

    # coherent / incoherent halves separately (same kernel, one launch each)
    half = 4096 * 2048
    split = {"coherent": None, "incoherent": None}
    for name, sl in (() if args.no_split else (("coherent", slice(0, half)), ("incoherent", slice(half, n)))):
        sub = [r[sl] for r in rays]
        tm = tmax[sl]
        bvh.trace_closest(*sub, tmax=tm, prim=prim[sl], t=t[sl], stream=s)
        torch.cuda.synchronize()
        ks = []
        for _ in range(max(3, args.steps // 2)):
            bvh.trace_closest(*sub, tmax=tm, prim=prim[sl], t=t[sl], stream=s)
            ks.append(bvh.last_kernel_ms())
        split[name] = round((sl.stop - sl.start) / (sum(ks) / len(ks) / 1e3) / 1e6, 1)

    achieved = alg_bytes / (k_avg / 1e3) / 1e9
    # HBM traffic and shader counters from the committed profiles of this
    # workload, used only while they describe it (same algorithmic bytes per
    # launch within 1 %: the same rays, tree and visit counts)
    def profile(name):
        try:
            d = json.load(open(os.path.join(ROOT, "profiles", name)))
        except (OSError, ValueError):
            return None
        ref = d.get("algorithmic_bytes_per_launch") or 0
        return d if ref > 0 and abs(alg_bytes - ref) <= 0.01 * ref else None
    pmc = profile(f"c1_trace_closest_bvh{args.width}_hbm_traffic.json")
    traffic = pmc.get("hbm_bytes_per_launch") if pmc else None
    sq = profile(f"c1_trace_closest_bvh{args.width}_sq_counters.json")
    valu = None if sq is None else {k: sq[k] for k in ("valu_busy", "lane_utilisation", "waiting_share",
                                                         "waves_per_simd")}
    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("BVH closest hit: 1M-triangle soup in the unit cube (seed 7), binned-SAH BVH leaves<=4, "
                         "16M rays = 8M coherent 4096x2048 pinhole + 8M incoherent"),
            "rays": n, "triangles": int(idx.shape[0]),
            "bvh_layout": ("four-wide, 8-bit quantized child boxes" if args.width == 4 else "two-wide, FP32 boxes"),
            "bvh_nodes": int(inf.num_nodes4 if args.width == 4 else inf.num_nodes),
            "bvh_depth": int(inf.depth4 if args.width == 4 else inf.max_depth),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "valu": valu,
            "note": ("achieved = 36 B per ray + 64 B per node visit + 48 B per triangle test (counted by the "
                     "kernel's counting variant) / the kernel's HIP-event time; traffic and valu come from the "
                     "committed FETCH_SIZE / WRITE_SIZE and SQ profiles of this workload (null when the run's "
                     "algorithmic bytes differ from theirs)"),
        },
        "kernel_ms": round(k_avg, 3),
        "mrays_coherent": split["coherent"],
        "mrays_incoherent": split["incoherent"],
        "node_visits_per_ray": round(nn / n, 2),
        "tri_tests_per_ray": round(ntr / n, 3),
        "hit_rate": round(hit_rate, 4),
    }
    if not args.no_cpu_baseline:
        log("CPU baseline (reference algorithm restated, bounded sample)")
        out["cpu_baseline"] = cpu_baseline_c1(xyz, idx, rays_np, budget_s=args.cpu_budget)
    bvh.close()
    ctx.close()
    return out


# Algorithmic bytes of the renderer's traversal kernels (per ray, before the
# tree): k_extend reads a queue slot (4 B) and the ray (origin + direction,
# 24 B) and writes the hit (id + t, 8 B); k_connect reads the shadow ray
# (origin, direction, tMax: 28 B) and its path slot (4 B), and an occluded
# ray clears its light bit (a 4 B read-modify-write).  A batch path also
# reads its scene's root (4 B).  Every node visit reads 64 B and every
# triangle test 48 B, counted by the kernels' counting variants.
# rt_render's default wave size (paths in flight), capped at a render's paths
DEFAULT_WAVE = 1 << 24


def measured_traffic(profile, kernel, alg_bytes, field="hbm_bytes_per_frame", wave=None):
    """HBM bytes per frame of `kernel` (or another field of its entry, such
    as its shader counters) from a committed FETCH_SIZE / WRITE_SIZE profile
    (copy-scaled), or None when there is none or it was taken on another
    workload (its algorithmic bytes per frame differ from this run's by more
    than 1 %, or its wave size, when given, is not this run's)."""
    if not profile:
        return None
    path = os.path.join(ROOT, "profiles", profile)
    try:
        d = json.load(open(path))
        k = d["kernels"][kernel]
    except (OSError, ValueError, KeyError):
        return None
    if wave is not None and d.get("path_capacity") != wave:
        return None
    ref = k.get("algorithmic_bytes_per_frame") or 0
    if ref <= 0 or abs(alg_bytes - ref) > 0.01 * ref or field not in k:
        return None
    v = k[field]
    return int(v) if field == "hbm_bytes_per_frame" else v


def traversal_roofline(count_stats, frames, ext_ms, con_ms, batch=False, note="", traffic_profile=None, wave=None):
    """The roofline object of a path-tracing line: the dominant traversal
    kernel's algorithmic bytes per frame (from a counting frame over the same
    samples) over its HIP-event time per frame (timed frames).  traffic: that
    kernel's HBM bytes per frame from traffic_profile (profiles/*.json) when
    it covers this workload, else None."""
    c = count_stats
    root = 4 if batch else 0
    ext_b = c.extension_rays * (36 + root) + 64 * c.ext_node_visits + 48 * c.ext_tri_tests
    con_b = (c.shadow_rays * (32 + root) + 8 * c.shadow_occluded + 64 * c.shadow_node_visits
             + 48 * c.shadow_tri_tests)
    ext_s, con_s = ext_ms / frames / 1e3, con_ms / frames / 1e3
    if con_s >= ext_s:
        name, b, t = "k_connect (any hit, shadow rays)", con_b, con_s
    else:
        name, b, t = "k_extend (closest hit, extension rays)", ext_b, ext_s
    achieved = b / t / 1e9
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(traffic_profile, name, b, wave=wave),
        "valu": measured_traffic(traffic_profile, name, b, field="valu", wave=wave),
        "kernel": name,
        "bytes_per_frame": int(b), "kernel_ms_per_frame": round(t * 1e3, 3),
        "per_ray": {
            "extension": {"node_visits": round(c.ext_node_visits / max(1, c.extension_rays), 3),
                          "tri_tests": round(c.ext_tri_tests / max(1, c.extension_rays), 3),
                          "kernel_ms_per_frame": round(ext_s * 1e3, 3),
                          "achieved_gbs": round(ext_b / ext_s / 1e9, 1) if ext_s > 0 else None},
            "shadow": {"node_visits": round(c.shadow_node_visits / max(1, c.shadow_rays), 3),
                       "tri_tests": round(c.shadow_tri_tests / max(1, c.shadow_rays), 3),
                       "occluded": round(c.shadow_occluded / max(1, c.shadow_rays), 4),
                       "kernel_ms_per_frame": round(con_s * 1e3, 3),
                       "achieved_gbs": round(con_b / con_s / 1e9, 1) if con_s > 0 else None},
        },
        "note": ("achieved = the kernel's algorithmic bytes per frame (36 B per extension ray or 32 B per shadow "
                 "ray + 8 B per occluded one, 64 B per node visit, 48 B per triangle test; visits and tests from "
                 "a counting frame over the same samples) / its summed HIP-event time per frame over the timed "
                 "frames (events on the launch stream around every launch); traffic = its HBM bytes per frame "
                 "from the FETCH_SIZE / WRITE_SIZE profile of the same workload (null without one)" + note),
    }


def run_render(args, which):
    """Path-tracing workloads: c0 = Cornell box 256^2 16 spp depth 8;
    c2 = procedural 3.4M-triangle scene 1920x1080 64 spp depth 16.  One step
    = one full frame (all samples of all pixels); value = samples/s."""
    import numpy as np
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    scenes = importlib.import_module("2023_render_engine_amd.scenes")
    if which == "c0":
        desc = scenes.cornell()
        label = "Cornell box (10 triangles + 2 spheres, 1 point light) 256x256, 16 spp, depth 8, seed 1"
    else:
        log("generating the procedural scene")
        desc = scenes.procedural()
        label = ("procedural: 1024^2 heightfield + 64 icospheres (5 subdivisions) = "
                 f"{desc.idx.shape[0]} triangles, 8 point lights + 1 sphere light, 1920x1080, 64 spp, depth 16")
    ctx = rt.Context(0)
    t0 = time.time()
    sc = rt.Scene(ctx, desc, bvh_on_device=args.bvh_device)
    commit_s = time.time() - t0
    log(f"scene committed in {commit_s:.1f} s ({'GPU' if args.bvh_device else 'host binned-SAH'} BVH build)")
    film = sc.new_film()
    for _ in range(args.warmup):
        sc.render(film)
    # counting frame over the same samples as the timed frames (untimed)
    cfilm = sc.new_film()
    sc.render(cfilm, count=True)
    cstats = sc.stats()
    del cfilm
    torch.cuda.synchronize()
    ext = shadow = 0
    ext_ms = con_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sc.render(film, profile=True)
        st = sc.stats()
        ext += st.extension_rays
        shadow += st.shadow_rays
        ext_ms += st.extend_ms
        con_ms += st.connect_ms
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    samples = desc.width * desc.height * desc.spp * args.steps
    # RMSE (the reference's computeRMSE) of the frame against an FP64
    # high-spp render of the same scene with disjoint samples (untimed)
    rmse = None
    if args.ref_spp > 0:
        ref = sc.new_film()
        for k in range(0, args.ref_spp, desc.spp):
            sc.render(ref, spp=desc.spp, spp_offset=(1 << 24) + k)
        torch.cuda.synchronize()
        ref_img = ref[0].view(-1, 3) / ref[1].double().clamp(min=1).unsqueeze(1)
        rmse = {"ref_spp": args.ref_spp, "frame_spp": desc.spp, "rmse": round(_rmse_dev(film[0], film[1], ref_img), 6)}
    out = {
        "metric": METRIC,
        "value": round(samples / wall / 1e6, 3),
        "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 film)", "data": "synthetic",
        "config": {"workload": label, "triangles": int(desc.idx.shape[0]),
                   "bvh_build": "GPU (rt_bvh_build_device)" if args.bvh_device else "host binned SAH",
                   "scene_commit_s": round(commit_s, 3)},
        "mrays_total": round((ext + shadow) / wall / 1e6, 1),
        "mrays_extension": round(ext / wall / 1e6, 1),
        "mrays_shadow": round(shadow / wall / 1e6, 1),
        "rays_per_sample": round((ext + shadow) / samples, 2),
        "rmse_vs_ref": rmse,
        "roofline": traversal_roofline(cstats, args.steps, ext_ms, con_ms,
                                       traffic_profile="c2_path_trace_procedural_hbm_traffic.json",
                                       wave=min(DEFAULT_WAVE, desc.width * desc.height * desc.spp)
                                       if which == "c2" else None),
    }
    if not args.no_cpu_baseline:
        import oracle
        os_ = oracle.OrcScene(desc)
        W, H = desc.width, desc.height
        # bounded sample: a centred window, growing until the budget is used
        spent, done = 0.0, 0
        w = 32
        while spent < args.cpu_budget and w <= min(W, H):
            x0, y0 = (W - w) // 2, (H - w) // 2
            t1 = time.time()
            os_.render(window=(x0, y0, x0 + w, y0 + w), nthreads=cpu_threads())
            spent += time.time() - t1
            done += w * w * desc.spp
            w *= 2
        out["cpu_baseline"] = {
            "value": round(done / spent / 1e6, 4), "unit": "Msamples/s", "cores": cpu_threads(), "kind": "port",
            "sample": f"{done} samples (centred square windows of the same frame, all {desc.spp} spp) with the "
                      f"oracle's restatement of the reference estimator, {spent:.1f} s",
        }
    sc.close()
    ctx.close()
    return out


def _rmse_dev(acc, nv, ref_img):
    """computeRMSE (camera/Film.cu:611-632) on the device, FP64: mean over
    pixels of the RGB channel sum of squared differences of mean radiance."""
    import torch
    img = acc.view(-1, 3) / nv.double().clamp(min=1).unsqueeze(1)
    return float(torch.sqrt(((img - ref_img) ** 2).sum(1).mean()))


def run_c4(args):
    """Time to RMSE <= target against a high-spp FP64 reference, plain path
    tracing vs path tracing with the online-trained radiance cache, on the
    procedural scene at 1920x1080.

    Frames are 4 spp each and accumulate into one film.  Plain: rt_render
    frames until the image's RMSE against the reference reaches the target.
    Cache (rt_render_cache_frame): the cache trains inside the frames it
    renders, with no separate training paths.  The first --c4-warm frames
    are plain path tracing (their samples count towards the image) whose
    paths all yield training samples, followed by --c4-warm-steps Adam steps
    each; after them, paths end in the cache at their first cache vertex two
    bounces in, except one in 2^--c4-harvest-shift, which goes full length
    and yields a training sample (a full-path target), and --c4-steps Adam
    steps follow each frame.  The timed region holds every frame and every
    training step (device-synchronised per frame); the RMSE evaluation
    between frames is not timed.  The reference is the plain renderer with
    FP64 accumulation at --c4-ref-spp, with sample indices disjoint from both
    runs."""
    import numpy as np
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    scenes = importlib.import_module("2023_render_engine_amd.scenes")
    W, H = args.c4_width, args.c4_height
    log(f"generating the procedural scene at {W}x{H}")
    desc = scenes.procedural(width=W, height=H, spp=4)
    label = (f"radiance cache on the procedural scene ({desc.idx.shape[0]} triangles, 8 point lights + 1 sphere "
             f"light, depth 16, RR after 3) at {W}x{H}: 4 spp frames, cache after {args.c4_cache_depth} bounces, "
             f"trained inside the frames ({args.c4_warm} plain warm-up frames x {args.c4_warm_steps} steps, then 1 "
             f"path in 2^{args.c4_harvest_shift} harvested and {args.c4_steps} steps per frame, batch 65536); "
             f"reference {args.c4_ref_spp} spp FP64")
    ctx = rt.Context(0)
    sc = rt.Scene(ctx, desc)
    ref = sc.new_film()
    t0 = time.time()
    for k in range(0, args.c4_ref_spp, 256):
        sc.render(ref, spp=256, spp_offset=(1 << 24) + k)
    torch.cuda.synchronize()
    log(f"reference {args.c4_ref_spp} spp in {time.time() - t0:.1f} s")
    ref_img = ref[0].view(-1, 3) / ref[1].double().clamp(min=1).unsqueeze(1)
    target = args.c4_target
    lo = np.minimum(desc.xyz.min(0), [0.38, 0.38, 0.98]) - 1e-3
    hi = np.maximum(desc.xyz.max(0), [0.62, 0.62, 1.22]) + 1e-3

    def new_cache():
        return rt.RadianceCache(ctx, tuple(lo), tuple(hi), levels=16, log2_table=19, batch=65536, seed=1)

    def frame_call(cache, film, frame, **kw):
        warm = frame < args.c4_warm
        return sc.render_cache_frame(film, cache, cache_depth=args.c4_cache_depth,
                                     harvest_shift=args.c4_harvest_shift,
                                     train_steps=args.c4_warm_steps if warm else args.c4_steps, plain=warm, spp=4,
                                     spp_offset=4 * frame, **kw)

    def run(use_cache):
        film = sc.new_film()
        cache = new_cache() if use_cache else None
        curve, spent, frame, reached = [], 0.0, 0, None
        losses, samples, ext_ms, con_ms, cache_frames = [], [], 0.0, 0.0, 0
        while spent < args.c4_budget:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if use_cache:
                st = frame_call(cache, film, frame, profile=True)
                losses.append((frame, float(st.loss)))
                samples.append(int(st.samples))
            else:
                sc.render(film, spp=4, spp_offset=4 * frame, profile=True)
            torch.cuda.synchronize()
            spent += time.perf_counter() - t1
            rs = sc.stats()
            if not use_cache or frame >= args.c4_warm:
                ext_ms += rs.extend_ms
                con_ms += rs.connect_ms
                cache_frames += 1
            frame += 1
            r = _rmse_dev(film[0], film[1], ref_img)
            curve.append((round(spent, 3), round(r, 5)))
            if r <= target and reached is None:
                reached = spent
                break
        res = {"time_s": None if reached is None else round(reached, 3), "frames": frame,
               "final_rmse": curve[-1][1] if curve else None, "timed_s": round(spent, 3),
               "ms_per_frame": round(spent * 1e3 / max(1, frame), 2),
               "curve": curve[:: max(1, len(curve) // 12)] + curve[-1:]}
        if losses:
            res["loss_readings"] = [(f, round(v, 4)) for f, v in losses]
            res["loss_first_last"] = [round(losses[0][1], 4), round(losses[-1][1], 4)]
            res["samples_per_frame"] = {"warm_up": samples[0], "cache": samples[-1]}
        return res, cache, film, (ext_ms, con_ms, cache_frames)

    log("plain path tracing")
    plain, _, _, _ = run(False)
    log(f"plain: {plain}")
    log("radiance cache trained inside the frames")
    cached, cache, film, (ext_ms, con_ms, nprof) = run(True)
    log(f"cache: {cached}")
    # counting frame of the cache method, over the last timed frame's samples (untimed, no training)
    cf = sc.new_film()
    sc.render_cache_frame(cf, cache, cache_depth=args.c4_cache_depth, harvest_shift=args.c4_harvest_shift,
                          train_steps=0, spp=4, spp_offset=4 * (cached["frames"] - 1), count=True)
    cst = sc.stats()
    cache.close()
    tp, tc = plain["time_s"], cached["time_s"]
    ratio = round(tp / tc, 3) if (tp and tc) else None
    out = {
        "metric": METRIC,
        "value": ratio,
        "unit": "x (time-to-RMSE, plain / cache)",
        "n_gpus": 1, "steps": cached["frames"], "warmup": args.c4_warm,
        "ms_per_step": cached["ms_per_frame"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 shading, bf16 MFMA cache, FP64 film", "data": "synthetic",
        "config": {"workload": label, "triangles": int(desc.idx.shape[0]), "target_rmse": target},
        "plain": plain, "cache": cached,
        "roofline": traversal_roofline(cst, max(1, nprof), ext_ms, con_ms,
                                       note="; the cache frames after the warm-up (their training steps are outside "
                                            "these kernels)"),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_cache(args, lo, hi)
    sc.close()
    ctx.close()
    return out


def cpu_baseline_cache(args, lo, hi):
    """The oracle's FP64 forward + backward of the same cache (16 levels,
    2^19 entries, 64-wide MLP) on random records inside the scene's box, one
    core: training samples per second (the GPU trains 65536 per step)."""
    import numpy as np
    import oracle
    cfg = oracle.cache_cfg(16, 19, 16, 1.38, lo, hi)
    rng = np.random.default_rng(4)
    mlp = (rng.random(oracle.CACHE_PARAMS, dtype=np.float32) - 0.5) * 0.2
    grid = (rng.random(16 * (1 << 19) * 2, dtype=np.float32) - 0.5) * 2e-4
    spent, done, n = 0.0, 0, 512
    while spent < min(args.cpu_budget, 10.0):
        rec = np.zeros((n, 12), np.float32)
        rec[:, :3] = lo + rng.random((n, 3)) * (np.asarray(hi) - lo)
        rec[:, 3:6] = [0, 0, 1]
        rec[:, 6:9] = [0, 0, 1]
        rec[:, 9] = 1
        tgt = rng.random((n, 3)).astype(np.float32)
        t1 = time.time()
        oracle.cache_loss_grad(cfg, mlp, grid, rec, tgt)
        spent += time.time() - t1
        done += n
    return {"value": round(done / spent, 1), "unit": "training samples/s", "cores": 1, "kind": "port",
            "sample": f"{done} random records in the scene's box through the oracle's FP64 cache forward and "
                      f"backward pass (no optimiser step), {spent:.1f} s"}


def run_c3(args):
    """Batched rendering: 32 procedural scenes (triangles log-uniform in
    [1e4, 5e5], resolutions uniform over {256^2, 512^2, 1024^2}, 32 spp,
    depth 12, scene i seeded 1000 + i) in one rt_batch launch sequence with
    shared queues, frames replayed from a captured hipGraph, wave size and
    grid from the autotuner.  One step = one frame of every scene at 32 spp.
    Also timed: the same frame as 32 separate rt_render calls with the same
    wave size and grid.  Per-scene
    RMSE is taken against a plain FP64 render of the same scene at
    --c3-ref-spp with disjoint samples (itself noisy: its own RMSE to the
    converged image is about sqrt(32 / ref_spp) of the frame's)."""
    import numpy as np
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    scenes = importlib.import_module("2023_render_engine_amd.scenes")
    log("generating 32 procedural scenes")
    descs = scenes.batch_suite(spp=32, max_depth=12)
    ntris = sum(int(d.idx.shape[0]) for d in descs)
    npix = sum(d.width * d.height for d in descs)
    ctx = rt.Context(0)
    t0 = time.time()
    scs = [rt.Scene(ctx, d, bvh_on_device=args.bvh_device) for d in descs]
    commit_s = time.time() - t0
    log(f"32 scenes committed in {commit_s:.1f} s ({ntris} triangles, {npix} pixels)")
    batch = rt.Batch(ctx, scs)
    seeds = [1000 + i for i in range(len(descs))]
    # 2 spp per trial, so that waves up to 2^24 paths (above the 11.9M pixels) are among the candidates
    best, entries = batch.autotune(spp=2, max_depth=12)
    log("autotune: " + ", ".join(f"P={e.path_capacity}/grid={e.grid}: {e.ms:.1f} ms" for e in entries))
    films = [s.new_film() for s in scs]
    for w in range(args.warmup):
        batch.render(films, spp=32, spp_offset=32 * w, max_depth=12, seeds=seeds, graph=True)
    torch.cuda.synchronize()
    for f in films:
        f[0].zero_()
        f[1].zero_()
    ext = shadow = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch.render(films, spp=32, spp_offset=32 * k, max_depth=12, seeds=seeds, graph=True)
        st = batch.stats()
        ext += st.extension_rays
        shadow += st.shadow_rays
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    samples = npix * 32 * args.steps
    # traversal kernel times: one eager frame with event pairs (graph replays
    # cannot hold them) and a counting frame, both over the first timed
    # frame's samples (untimed)
    pfilms = [s.new_film() for s in scs]
    batch.render(pfilms, spp=32, spp_offset=0, max_depth=12, seeds=seeds, profile=True)
    pst = batch.stats()
    batch.render(pfilms, spp=32, spp_offset=0, max_depth=12, seeds=seeds, count=True)
    cst = batch.stats()
    del pfilms
    torch.cuda.synchronize()
    eager_ms = pst.ms
    # the same frame rendered scene by scene
    seq = [s.new_film() for s in scs]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s, f, sd in zip(scs, seq, seeds):
        s.render(f, spp=32, spp_offset=0, max_depth=12, seed=sd, path_capacity=best.path_capacity, grid=best.grid)
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t1
    rmse = []
    if args.c3_ref_spp > 0:
        log(f"references at {args.c3_ref_spp} spp")
        for s, f, sd in zip(scs, films, seeds):
            ref = s.new_film()
            s.render(ref, spp=args.c3_ref_spp, spp_offset=1 << 24, max_depth=12, seed=sd)
            ri = ref[0].view(-1, 3) / ref[1].double().clamp(min=1).unsqueeze(1)
            rmse.append(_rmse_dev(f[0], f[1], ri))
    out = {
        "metric": METRIC,
        "value": round(samples / wall / 1e6, 3),
        "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 film)", "data": "synthetic",
        "config": {"workload": ("batched: 32 procedural scenes (triangles log-uniform 1e4-5e5, resolutions "
                                "{256,512,1024}^2), 32 spp, depth 12, one rt_batch launch sequence, graph-replayed "
                                "frames, autotuned wave size and grid"),
                   "triangles": ntris, "pixels": npix,
                   "autotuned": {"path_capacity": best.path_capacity, "grid": best.grid},
                   "bvh_build": "GPU (rt_bvh_build_device)" if args.bvh_device else "host binned SAH",
                   "scene_commit_s": round(commit_s, 3)},
        "mrays_total": round((ext + shadow) / wall / 1e6, 1),
        "mrays_extension": round(ext / wall / 1e6, 1),
        "sequential_msamples_s": round(npix * 32 / seq_s / 1e6, 3),
        "batch_speedup_vs_sequential": round(seq_s / (wall / args.steps), 3),
        "roofline": traversal_roofline(cst, 1, pst.extend_ms, pst.connect_ms, batch=True,
                                       note=f"; kernel times from one eager frame after the timed graph frames "
                                            f"(frame {eager_ms:.1f} ms eager vs {wall * 1e3 / args.steps:.1f} ms "
                                            f"replayed)"),
    }
    if rmse:
        out["rmse_vs_ref"] = {"ref_spp": args.c3_ref_spp, "frames": args.steps, "mean": round(float(np.mean(rmse)), 5),
                              "max": round(float(np.max(rmse)), 5)}
    if not args.no_cpu_baseline:
        import oracle
        spent, done = 0.0, 0
        for d in descs[:2]:
            os_ = oracle.OrcScene(d)
            w = 32
            x0, y0 = (d.width - w) // 2, (d.height - w) // 2
            t2 = time.time()
            os_.render(spp=32, window=(x0, y0, x0 + w, y0 + w), nthreads=cpu_threads())
            spent += time.time() - t2
            done += w * w * 32
        out["cpu_baseline"] = {
            "value": round(done / spent / 1e6, 4), "unit": "Msamples/s", "cores": cpu_threads(), "kind": "port",
            "sample": f"{done} samples: centred 32x32 windows of 2 of the 32 scenes at 32 spp with the oracle's "
                      f"restatement of the reference estimator, {spent:.1f} s"}
    batch.close()
    for s in scs:
        s.close()
    ctx.close()
    return out


def run_film(args):
    """Film output: a 1920x1080 FP64 film (synthetic HDR pixel sums, 64
    samples) tone-mapped and packed into a BMP pixel array per step
    (rt_film_tonemap: max pass + map pass).  value = Mpixels/s; the roofline
    prices one call at 51 algorithmic bytes per pixel (24 B of sums read by
    each pass, 3 B written)."""
    import numpy as np
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    w, h, spp = 1920, 1080, 64
    mode = args.tone
    rng = np.random.default_rng(7)
    sums_np = rng.random((w * h, 3)) * 8.0 * spp * rng.random((w * h, 1)) ** 2
    ctx = rt.Context(0)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(stream):
        sums = torch.from_numpy(sums_np.reshape(-1)).to(dev)
        out = torch.empty(rt.lib().rt_image_pixel_bytes(0, w, h), dtype=torch.uint8, device=dev)
        for _ in range(args.warmup):
            rt.tone_map(ctx, sums, w, h, spp, mode, "bmp", out=out)
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            rt.tone_map(ctx, sums, w, h, spp, mode, "bmp", out=out)
        e1.record(stream)
        stream.synchronize()
        wall = time.perf_counter() - t0
    call_ms = e0.elapsed_time(e1) / args.steps
    npix = w * h
    bytes_per_call = 51 * npix
    gbs = bytes_per_call / (call_ms * 1e-3) / 1e9
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "film_tonemap_1080p_hbm_traffic.json")
    if mode == "gammaClamp" and os.path.exists(pmc_path):
        try:
            traffic = json.load(open(pmc_path)).get("hbm_bytes_per_call")
        except (OSError, ValueError):
            traffic = None
    res = {
        "metric": "film tone mapping + BMP packing", "value": round(npix * args.steps / wall / 1e6, 1),
        "unit": "Mpixels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"1920x1080 FP64 film, {spp} samples, {mode}, 24-bit BMP pixel array"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbs / 8000.0, 4), "traffic": traffic,
                     "note": "one call = k_film_max + k_film_fold + k_film_map, 51 B/pixel, timed with HIP events on "
                             "its stream; traffic from the FETCH_SIZE / WRITE_SIZE profile of the same call"},
    }
    if not args.no_cpu_baseline:
        import oracle
        spent, done = 0.0, 0
        while spent < min(args.cpu_budget, 10.0):
            t1 = time.time()
            oracle.film_image(sums_np, w, h, spp, mode, "bmp")
            spent += time.time() - t1
            done += npix
        res["cpu_baseline"] = {"value": round(done / spent / 1e6, 2), "unit": "Mpixels/s", "cores": 1,
                               "kind": "port", "sample": f"{done // npix} whole 1920x1080 images with the oracle's "
                                                         f"restatement, {spent:.1f} s"}
    ctx.close()
    return res


def run_photon(args):
    """Photon map gather queries: 2^20 photons (the reference's
    N_PHOTON_SAMPLES, clustered on 16 blobs as photons gather on surfaces),
    one k-nearest query per pixel of a 1920x1080 frame, k = 100 within radius
    100 (MAX_NEIGHBOURS, PHOTON_SEARCH_RADIUS).  One step = one
    rt_photon_map_knn call over all queries; value = Mqueries/s."""
    import numpy as np
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    n, m, k, radius = 1 << 20, 1920 * 1080, 100, 100.0
    rng = np.random.default_rng(3)
    centres = rng.random((16, 3)) * 10
    x = (centres[rng.integers(0, 16, n)] + rng.normal(0, 0.2, (n, 3))).astype(np.float32)
    qn = (centres[rng.integers(0, 16, m)] + rng.normal(0, 0.2, (m, 3))).astype(np.float32)
    ctx = rt.Context(0)
    t0 = time.time()
    pm = rt.PhotonMap(ctx, x)
    build_s = time.time() - t0
    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(stream):
        q = torch.from_numpy(qn).to(dev)
        for _ in range(args.warmup):
            pm.nearest(q, k, radius)
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            cnt, idx, dist = pm.nearest(q, k, radius)
        e1.record(stream)
        stream.synchronize()
        wall = time.perf_counter() - t0
    call_ms = e0.elapsed_time(e1) / args.steps
    # compulsory bytes per query: 12 B point, 4 B count, 8 B per neighbour
    io = m * (12 + 4 + 8 * k)
    gbs = io / (call_ms * 1e-3) / 1e9
    res = {
        "metric": "photon map k-nearest queries", "value": round(m * args.steps / wall / 1e6, 2),
        "unit": "Mqueries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} photons in 16 clusters, {m} queries (one per 1920x1080 pixel), k={k}, "
                               f"radius {radius}", "build_s": round(build_s, 3)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbs / 8000.0, 4), "traffic": None,
                     "note": "compulsory I/O only (12 + 4 + 8k B per query); node fetches are cached gathers"},
    }
    if not args.no_cpu_baseline:
        import oracle
        tree = oracle.kd_build(x)[:2]
        spent, done, chunk = 0.0, 0, 4096
        while spent < min(args.cpu_budget, 10.0) and done + chunk <= m:
            t1 = time.time()
            oracle.knn(x, qn[done:done + chunk], k, radius, tree=tree, nthreads=cpu_threads())
            spent += time.time() - t1
            done += chunk
        res["cpu_baseline"] = {"value": round(done / spent / 1e6, 4), "unit": "Mqueries/s", "cores": cpu_threads(),
                               "kind": "port", "sample": f"{done} of the queries with the oracle's restatement of "
                                                         f"the kd-tree search, {spent:.1f} s"}
    pm.close()
    ctx.close()
    return res


def run_bvh_build(args):
    """The BVH built on the GPU (rt_bvh_build_device, the arrangement of the
    reference's BVH constructor, cuda_libraries/BVH.cu:314-341) over the 1M
    triangle soup of the closest-hit workload (seed 7), vertices and indices
    resident in HBM.  One step = one build (work space allocation, Morton
    sort, radix tree, boxes, two-wide and four-wide records); value =
    Mtriangles/s.  Beside it: closest-hit Mrays/s over the same 16M rays on
    the tree built here and on the host's binned-SAH tree, the quality the
    build trades for its speed."""
    import numpy as np
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    dev = torch.device("cuda", 0)
    xyz, idx = rt.gen_soup(1 << 20, seed=7)
    nt = idx.shape[0]
    ctx = rt.Context(0)
    dx = torch.from_numpy(xyz).to(dev)
    di = torch.from_numpy(idx.view(np.int32)).to(dev)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        rt.Bvh.from_device(ctx, dx, di).close()
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        b = rt.Bvh.from_device(ctx, dx, di)
        times.append(time.perf_counter() - t1)
        b.close()
    wall = time.perf_counter() - t0
    ms = sum(times) / len(times) * 1e3
    lb = rt.Bvh.from_device(ctx, dx, di)
    inf = lb.info()
    # compulsory bytes of one build: each triangle's three indices and nine
    # coordinates in (48 B), its 48 B record out, 64 B per two-wide and
    # four-wide node out
    alg = nt * (48 + 48) + 64 * (inf.num_nodes + inf.num_nodes4)
    gbs = alg / (ms * 1e-3) / 1e9

    rays_np = rt.gen_rays(4096, 2048, eye=(0.5, 0.5, -1.0), vfov=60.0, n_incoherent=8 << 20, seed=7)
    n = rays_np[0].shape[0]
    rays = [torch.from_numpy(a).to(dev) for a in rays_np]
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    t = torch.empty(n, dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream

    def trace_rate(bvh):
        for _ in range(2):
            bvh.trace_closest(*rays, prim=prim, t=t, stream=s)
        torch.cuda.synchronize()
        ks = []
        for _ in range(5):
            bvh.trace_closest(*rays, prim=prim, t=t, stream=s)
            ks.append(bvh.last_kernel_ms())
        nn, ntr = bvh.trace_stats(*rays)
        return round(n / (sum(ks) / len(ks) / 1e3) / 1e6, 1), round(nn / n, 2), round(ntr / n, 3)

    lb_rate, lb_nodes, lb_tris = trace_rate(lb)
    t1 = time.time()
    sah = rt.Bvh(ctx, xyz, idx, max_leaf=4, bins=32, width=4)
    sah_wall = time.time() - t1
    sinf = sah.info()
    sah_rate, sah_nodes, sah_tris = trace_rate(sah)
    res = {
        "metric": "BVH build on the GPU", "value": round(nt * args.steps / wall / 1e6, 2),
        "unit": "Mtriangles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "rt_bvh_build_device over the 1M-triangle soup (seed 7), leaves <= 4, four-wide",
                   "triangles": nt, "nodes2": int(inf.num_nodes), "nodes4": int(inf.num_nodes4),
                   "depth4": int(inf.depth4), "stack_cap": int(inf.stack_cap)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None,
                     "note": ("whole build, not one kernel: compulsory bytes (96 B per triangle + 64 B per node "
                              "record) / the build's wall time; the build is a chain of short launches and "
                              "per-level read-backs, so it is latency bound")},
        "build_ms": round(ms, 3),
        "sah_cost": round(inf.sah_cost, 2),
        "trace_mrays_device_tree": lb_rate,
        "node_visits_per_ray_device_tree": lb_nodes,
        "tri_tests_per_ray_device_tree": lb_tris,
        "host_sah_build_ms": round(sinf.build_ms, 1),
        "host_sah_build_wall_s": round(sah_wall, 2),
        "host_sah_cost": round(sinf.sah_cost, 2),
        "trace_mrays_host_sah_tree": sah_rate,
        "node_visits_per_ray_host_sah_tree": sah_nodes,
        "tri_tests_per_ray_host_sah_tree": sah_tris,
    }
    if not args.no_cpu_baseline:
        import oracle
        t1 = time.time()
        oracle.lbvh_build(xyz, idx)
        cs = time.time() - t1
        res["cpu_baseline"] = {"value": round(nt / cs / 1e6, 3), "unit": "Mtriangles/s", "cores": 1, "kind": "port",
                               "sample": f"one build of the same 1M triangles by the oracle's restatement "
                                         f"(oracle/lbvh.c), {cs:.1f} s"}
    lb.close()
    sah.close()
    ctx.close()
    return res


def run_photon_render(args):
    """RENDER_MODE PHOTON_MAPPING on the reference's own photon-mapping setup
    (integration/PathTracerIntegrator.cu:157-161, Scene.cu:912-967): the
    Cornell box at 256x256, 16 spp, depth 8; 2^20 photon paths
    (N_PHOTON_SAMPLES) traced once by rt_scene_trace_photons and put in a
    photon map; every path ends at its first scattering hit with the k = 100
    nearest photons within radius 100 (MAX_NEIGHBOURS, PHOTON_SEARCH_RADIUS).
    One step = one rt_render_photon frame; value = Msamples/s.  The photon
    pass and the map build are outside the timed region (the reference also
    builds them once per render)."""
    import numpy as np
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    scenes = importlib.import_module("2023_render_engine_amd.scenes")
    desc = scenes.cornell(width=256, height=256, spp=16, max_depth=8, seed=1)
    ctx = rt.Context(0)
    sc = rt.Scene(ctx, desc)
    t0 = time.time()
    ph = sc.trace_photons(1 << 20, seed=1)
    torch.cuda.synchronize()
    pm = rt.PhotonMap(ctx, ph["location"].cpu().numpy(), ph["shape"].cpu().numpy())
    setup_s = time.time() - t0
    nphot = int(ph["location"].shape[0])
    film = sc.new_film()
    k, radius = 100, 100.0
    for _ in range(args.warmup):
        sc.render_photon(film, pm, ph["radiance"], ph["direction"], k=k, radius=radius, spp=16)
    torch.cuda.synchronize()
    dev = torch.device("cuda", 0)
    s = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for i in range(args.steps):
        sc.render_photon(film, pm, ph["radiance"], ph["direction"], k=k, radius=radius, spp=16,
                         spp_offset=16 * (i + 1), stream=s.cuda_stream)
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    frame_ms = e0.elapsed_time(e1) / args.steps
    samples = 256 * 256 * 16
    # compulsory bytes of the frame's photon searches: per query 12 B point,
    # 4 B count and 8 B per neighbour (index, distance), one query per sample
    # that reaches a scattering surface (at most one per sample)
    io = samples * (12 + 4 + 8 * k)
    gbs = io / (frame_ms * 1e-3) / 1e9
    res = {
        "metric": "photon-mapping render (RENDER_MODE PHOTON_MAPPING)", "value": round(samples * args.steps / wall / 1e6, 3),
        "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (FP64 film, FP64 photon sums)", "data": "synthetic",
        "config": {"workload": "Cornell box 256x256, 16 spp, depth 8; 2^20 photon paths, k=100 within radius 100",
                   "photons": nphot, "photon_pass_and_map_build_s": round(setup_s, 3)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None,
                     "note": ("whole frame (HIP events on the launch stream): the photon searches' compulsory I/O "
                              "(12 + 4 + 8k B per sample) / the frame time; the k-nearest search is a dependent "
                              "walk per query, bound by latency")},
        "frame_ms": round(frame_ms, 3),
    }
    if not args.no_cpu_baseline:
        import oracle
        host = {kk: ph[kk].cpu().numpy() for kk in ("location", "shape", "radiance", "direction")}
        osc = oracle.OrcScene(desc)
        t1 = time.time()
        win = (112, 112, 144, 144)
        osc.render_photon(host, k=k, radius=radius, spp=16, window=win, nthreads=cpu_threads())
        cs = time.time() - t1
        n = 32 * 32 * 16
        res["cpu_baseline"] = {"value": round(n / cs / 1e6, 4), "unit": "Msamples/s", "cores": cpu_threads(),
                               "kind": "port", "sample": f"the centre 32x32 window at 16 spp ({n} samples) with the "
                                                         f"oracle's restatement and the same photons, {cs:.1f} s "
                                                         f"(kd-tree build included)"}
    pm.close()
    sc.close()
    ctx.close()
    return res


def run_photon_pass(args):
    """The photon pre-pass (generatePhotons + computePhotons,
    integration/photonDispersion.cu:26-192) at the reference's size: 2^20
    photon paths (N_PHOTON_SAMPLES, types.h:57), at most 30 intersections
    each (MAX_PHOTON_RECURSION_DEPTH), on the Cornell box (the scene the
    reference renders with its photon pass on).  One step = one
    rt_scene_trace_photons call (light selection, tracing, sort by path and
    bounce, output); value = photon paths per second."""
    import torch

    rt = importlib.import_module("2023_render_engine_amd")
    scenes = importlib.import_module("2023_render_engine_amd.scenes")
    n = 1 << 20
    desc = scenes.cornell()
    ctx = rt.Context(0)
    sc = rt.Scene(ctx, desc)
    for w in range(args.warmup):
        ph = sc.trace_photons(n, seed=1 + w)
    torch.cuda.synchronize()
    stored = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        ph = sc.trace_photons(n, seed=100 + k)
        stored += ph["location"].shape[0]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = stored / args.steps
    # compulsory bytes per call: 64 B written per stored photon (3 x 12 B of
    # float vectors, 24 B of FP64 radiance, 4 B shape), read back and written
    # once more by the sort's output pass, plus 4 B keys twice
    io = per * (64 * 3 + 16)
    gbs = io / (wall / args.steps) / 1e9
    res = {
        "metric": "photon pre-pass", "value": round(n * args.steps / wall / 1e6, 2), "unit": "Mpaths/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 radiance)",
        "data": "synthetic",
        "config": {"workload": f"Cornell box (10 triangles, 2 spheres, 1 point light), {n} photon paths, at most 30 "
                               f"intersections, Russian roulette after 3", "photons_stored": int(per)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "note": "compulsory photon I/O over the whole call (append, sort, output); the pass is "
                             "latency bound: one lane walks one photon path, with two host round trips per call"},
    }
    if not args.no_cpu_baseline:
        import oracle
        o = oracle.OrcScene(desc)
        spent, done, chunk = 0.0, 0, 20000
        while spent < min(args.cpu_budget, 10.0):
            t1 = time.time()
            o.trace_photons(chunk, seed=7 + done, cap=chunk * 32)
            spent += time.time() - t1
            done += chunk
        res["cpu_baseline"] = {"value": round(done / spent / 1e6, 4), "unit": "Mpaths/s", "cores": 1, "kind": "port",
                               "sample": f"{done} photon paths in batches of {chunk} with the oracle's restatement, "
                                         f"{spent:.1f} s (its light counts are per batch)"}
    sc.close()
    ctx.close()
    return res


def run_copy(args):
    """HBM calibration: a plain device-to-device copy of 1 GiB (torch), so the
    FETCH_SIZE / WRITE_SIZE counters of a kernel of known traffic can scale
    the traversal kernel's counters.  Not a benchmark of the renderer."""
    import torch
    n = 1 << 28  # float32 elements = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gbs = 2 * 4 * n * args.steps / wall / 1e9
    return {"metric": "HBM copy calibration", "value": round(gbs, 1), "unit": "GB/s (read + write)",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "device-to-device copy of 2^28 float32 (1 GiB read, 1 GiB written)"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1", choices=["c1", "c0", "c2", "c3", "c4", "copy", "film", "photon", "photon_pass",
                                                         "bvh_build", "photon_render"])
    ap.add_argument("--tone", default="gammaClamp", help="film: tone mapping (a ToneMappingType name)")
    ap.add_argument("--c3-ref-spp", type=int, default=128)
    ap.add_argument("--ref-spp", type=int, default=1024,
                    help="c0/c2: spp of the FP64 reference the frame's RMSE is taken against (0: skip)")
    ap.add_argument("--bvh-device", action="store_true",
                    help="c0/c2/c3: build the scenes' BVHs on the GPU (rt_scene_set_bvh) instead of the host's binned SAH")
    ap.add_argument("--width", type=int, default=4, choices=[2, 4],
                    help="c1: tree layout (4 = four-wide quantized nodes, 2 = two-wide FP32 nodes)")
    ap.add_argument("--c4-width", type=int, default=1920)
    ap.add_argument("--c4-height", type=int, default=1080)
    ap.add_argument("--c4-ref-spp", type=int, default=2048)
    ap.add_argument("--c4-target", type=float, default=1e-2)
    ap.add_argument("--c4-budget", type=float, default=20.0, help="c4: timed seconds per method at most")
    ap.add_argument("--c4-warm", type=int, default=1, help="c4: plain frames whose paths all train the cache")
    ap.add_argument("--c4-warm-steps", type=int, default=32, help="c4: Adam steps after each warm-up frame")
    ap.add_argument("--c4-steps", type=int, default=2, help="c4: Adam steps after each cache frame")
    ap.add_argument("--c4-harvest-shift", type=int, default=7,
                    help="c4: one path in 2^shift of a cache frame goes full length and trains the cache")
    ap.add_argument("--c4-cache-depth", type=int, default=2,
                    help="c4: bounces before a path ends in the cache (the workload's definition is 2)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-split", action="store_true",
                    help="c1: skip the separate coherent/incoherent launches (profiling runs)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(--gpus must be 1)"}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    if args.workload == "c1":
        out = run_c1(args)
    elif args.workload == "c3":
        out = run_c3(args)
    elif args.workload == "c4":
        out = run_c4(args)
    elif args.workload == "copy":
        out = run_copy(args)
    elif args.workload == "film":
        out = run_film(args)
    elif args.workload == "photon":
        out = run_photon(args)
    elif args.workload == "photon_pass":
        out = run_photon_pass(args)
    elif args.workload == "bvh_build":
        out = run_bvh_build(args)
    elif args.workload == "photon_render":
        out = run_photon_render(args)
    else:
        out = run_render(args, args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
