"""Benchmark: incoherent secondary rays/s on the 1.02M-triangle scene, one MI355X.

Workload (config 2): ground plane + 200 level-4 icospheres = 1,024,002
triangles; 1920x1080 pinhole camera.  The ray set is one cosine-weighted
bounce per primary hit (seeded), built once and kept resident in HBM.  One
step = one closest-hit trace of the whole incoherent ray batch (one kernel
launch of the BVH8 traversal).  Primary and shadow rays/s, the GPU build time
and, when the render pass is available, frame rates are reported alongside.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import math
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "snake-vk_amd"))

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trace_c2_incoherent_traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_launches(fn, reps, stream):
    """Average duration (ms) of `fn` launches measured with events on `stream`."""
    import torch
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    stream.synchronize()
    return sum(a.elapsed_time(b) for a, b in evs) / reps


def settle_clocks(fn, stream, ms=30.0, max_launches=400):
    """Launch `fn` until `ms` of GPU time has passed (at most `max_launches`), so
    the GPU's clocks have left the idle state they drop to while the host sets
    up; the first ~40 launches of a 0.26 ms trace otherwise run up to 8 %
    slower.  Untimed set-up, like the autotuner's launches."""
    import torch
    done, launches = 0.0, 0
    while done < ms and launches < max_launches:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(10):
            fn()
        b.record(stream)
        stream.synchronize()
        done += a.elapsed_time(b)
        launches += 10
    return launches


def cpu_baseline(world, rays_np, threads, gpu_hits=None):
    """The CPU restatement (oracle BVH8 traversal) on the same incoherent rays;
    with gpu_hits = (t, inst, prim) of the timed GPU trace, also how many of the
    rays got the same hit bit for bit."""
    import numpy as np

    from oracle import binding as O
    t0 = time.perf_counter()
    bvh = O.Bvh(world)
    build_s = time.perf_counter() - t0
    n = rays_np.shape[0]
    t0 = time.perf_counter()
    ct, ci, cp, *_ = bvh.trace(rays_np, threads=threads)
    dt = time.perf_counter() - t0
    rec = {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": f"all {n} incoherent secondary rays of the 1080p ray set, closest hit, CPU BVH8 "
                     f"restatement (oracle), {threads} threads; CPU LBVH->BVH8 build of the 1.02M "
                     f"triangles took {build_s:.2f} s"}
    match = None
    if gpu_hits is not None:
        gt, gi, gp = (np.ascontiguousarray(a) for a in gpu_hits)
        same = (gt.view(np.uint32) == ct.view(np.uint32)) & (gi.view(np.uint32) == ci) & (gp.view(np.uint32) == cp)
        match = {"rays": int(n), "equal": int(same.sum()), "frac": round(float(same.mean()), 6),
                 "what": "hit t (bits), instance id and primitive id of every ray of the timed GPU trace "
                         "against the CPU BVH8 traversal"}
    return rec, match


# (config, form); reference_compat: the reference's own estimator (ReSTIR direct light +
# BounceDiffuse's 8 + up to 16 rays per pixel) on the config 4 room
FRAME_RUNS = [(1, "pt"), (2, "pt"), (3, "restir"), (4, "pt"), (4, "cache"), (4, "cache_indirect"),
              (4, "reference_compat")]


# Algorithmic bytes per unit (the scope table's measurement figures, with this build's
# record sizes): closest-hit ray 56 + 80 N_node + 48 N_tri, shadow ray 40 + 80 N + 48 N,
# shaded path vertex 370; ReSTIR per pixel: candidates 2,790, temporal 476, spatial
# 3,164, shade 212 (profiles/restir_c3_1080p_pmc.md); cache query: 16 levels x 8
# corners x 2 x 2 B of hash-grid features = 512.
SHADE_BYTES = 370.0
RESTIR_BYTES = {"primary": 2790.0, "temporal": 476.0, "spatial": 3164.0, "shade": 212.0}
CACHE_QUERY_BYTES = 512.0


def visit_means(ctx, scene, w, h, stream):
    """Mean BVH8 nodes and triangles visited per ray (deterministic, counted by
    the traversal itself) for the configuration's camera rays, one cosine bounce
    from their hits and shadow rays from their hits to the emitters; returns the
    algorithmic bytes per ray of each kind."""
    import numpy as np
    import torch

    import raysets
    dev = torch.device("cuda:0")
    with torch.cuda.stream(stream):
        porg, pdr = raysets.primary_rays(scene.camera(0, w, h), w, h)
        po = torch.from_numpy(porg).to(dev)
        pd = torch.from_numpy(pdr).to(dev)
        pt, pinst, pprim, _, _, pnv, ptv = ctx.trace(po, pd, counted=True, stream=stream)
        world = torch.from_numpy(np.frombuffer(ctx.export(6).tobytes(), np.float32).reshape(-1, 12).copy()).to(dev)
        inst5 = scene.arrays()[2]
        off = torch.from_numpy(np.concatenate([[0], np.cumsum(inst5[:, 4] // 3)]).astype(np.int64)).to(dev)
        ext = float((world[:, :9].reshape(-1, 3).max(0).values - world[:, :9].reshape(-1, 3).min(0).values).max())
        so, sd = raysets.secondary_rays_torch(po, pd, pt, pinst, pprim, world, off, ext, seed=1)
        *_, snv, stv = ctx.trace(so, sd, counted=True, stream=stream)
        em = np.ctypeslib.as_array(scene.s.emissive_instances, (scene.s.n_emissive,))
        off_np = off.cpu().numpy()
        gids = np.concatenate([np.arange(off_np[i], off_np[i + 1]) for i in em]) if len(em) else np.zeros(0, np.int64)
        res = {"primary": 56.0 + 80.0 * float(pnv.double().mean()) + 48.0 * float(ptv.double().mean()),
               "extension": 56.0 + 80.0 * float(snv.double().mean()) + 48.0 * float(stv.double().mean())}
        if gids.size:
            et = world[torch.from_numpy(gids).to(dev), :9]
            ea = torch.linalg.cross(et[:, 3:6] - et[:, 0:3], et[:, 6:9] - et[:, 0:3]).norm(dim=1)
            sho, shd = raysets.shadow_rays_torch(po, pd, pt, et, ea / ea.sum(), seed=2, scale=ext)
            *_, hnv, htv = ctx.trace(sho, shd, any_hit=True, counted=True, stream=stream)
            res["shadow"] = 40.0 + 80.0 * float(hnv.double().mean()) + 48.0 * float(htv.double().mean())
        else:
            res["shadow"] = 40.0
        stream.synchronize()
    return res


def frame_algorithmic_bytes(st, B, mode, pixels):
    """Algorithmic bytes of one frame from its ray counts and per-ray bytes B."""
    import snkpt
    b = (st.camera_rays * B["primary"] + st.extension_rays * B["extension"] + st.shadow_rays * B["shadow"] +
         (st.camera_rays + st.extension_rays) * SHADE_BYTES)
    if mode == snkpt.MODE_RESTIR:
        b += pixels * sum(RESTIR_BYTES.values())
    if mode == snkpt.MODE_CACHE:
        b += (st.cache_queries + st.cache_train_samples) * CACHE_QUERY_BYTES
    return float(b)


def frame_rates(stream, frames=10, warmup=3, only=None):
    """Frames/s per configuration (1 spp frames of the render pass, event-timed
    on the launch stream).  Cache frames include the online training step
    (cache: L_o queried at path vertex 3; cache_indirect: NEE at the primary
    hit and the cache's indirect light there); reference_compat frames run
    the reference's biased BounceDiffuse estimator."""
    import torch

    import snkpt
    modes = {"pt": snkpt.MODE_PT, "restir": snkpt.MODE_RESTIR, "cache": snkpt.MODE_CACHE,
             "cache_indirect": snkpt.MODE_CACHE, "reference_compat": snkpt.MODE_REFERENCE_COMPAT}
    res = {}
    oracles = {}  # (config, camera frame) -> FP64 oracle image
    for cfg, label in FRAME_RUNS:
        if only is not None and f"c{cfg}_{label}" != only:
            continue
        mode = modes[label]
        scene = snkpt.SceneData(cfg)
        ctx = snkpt.Context(0)
        ctx.load_scene(scene)
        with torch.cuda.stream(stream):
            ctx.build_accel(stream)
            w, h = scene.s.width, scene.s.height
            cam = scene.camera(0, w, h)
            out = torch.zeros(h, w, 4, device="cuda")
            if mode == snkpt.MODE_CACHE:
                ctx.cache_reset(1, stream=stream)

            def settings(f):
                s = snkpt.default_settings(mode=mode, max_depth=scene.s.max_depth, rr_start=scene.s.rr_start,
                                           seed=scene.s.seed, frame=f, cache_query_depth=3)
                if label == "cache_indirect":
                    s.cache_query_depth, s.flags = 1, snkpt.FLAG_CACHE_INDIRECT
                if cfg == 3:
                    s.num_temporal_resamples, s.num_spatial_resamples = 1, 5
                return s

            def frame_cam(f):
                # config 3 renders its seeded 64-frame camera orbit (temporal reuse under motion)
                return scene.camera(f, w, h) if cfg == 3 else cam

            def frame(f):
                ctx.render(frame_cam(f), settings(f), out, w, h, stream=stream)
            # warm-up: at least `warmup` frames and ~30 ms of frames (the clocks leave
            # the idle state the scene set-up let them drop to): the second frame is
            # timed and the count rounded up to a power of two, so runs agree on it
            # (config 3's frames stay within its 64-frame orbit)
            frame(0)
            stream.synchronize()
            t1 = time.perf_counter()
            frame(1)
            stream.synchronize()
            f_ms = max(1e-3, (time.perf_counter() - t1) * 1e3)
            nw = max(warmup, min(1 << max(1, math.ceil(math.log2(30.0 / f_ms))), 64 - frames - 4))
            for f in range(2, nw):
                frame(f)
            evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                   for _ in range(frames)]
            for k, (a, b) in enumerate(evs):
                a.record(stream)
                frame(nw + k)
                b.record(stream)
            stream.synchronize()
            st = ctx.frame_stats()  # the last timed frame's ray counts (before the frames below)
            # the last timed frame's error against the converged FP64 oracle of its camera
            # (plain PT of the same depth: the cache's and the compat estimator's bias shows)
            last = nw + frames - 1
            okey = (cfg, last if cfg == 3 else 0)
            spp = 1024
            if okey not in oracles:
                acc = torch.zeros(h * w * 3, dtype=torch.float64, device="cuda")
                os_ = snkpt.default_settings(max_depth=scene.s.max_depth, rr_start=scene.s.rr_start,
                                             seed=scene.s.seed ^ 0x0AC1E)
                for k in range(spp // 64):
                    os_.frame = k
                    ctx.oracle_accumulate(frame_cam(last), os_, acc, w, h, 64, stream=stream)
                oracles[okey] = (acc / spp).view(h, w, 3)
            ref = oracles[okey]
            err = ((out[..., :3].double() - ref) ** 2).mean().sqrt()
            mae = (out[..., :3].double() - ref).abs().mean()
            frame_err = {"oracle_spp": spp, "frame_rel_rmse": round(float(err / ref.mean()), 4),
                         "frame_rel_mae": round(float(mae / ref.mean()), 4)}
            if mode != snkpt.MODE_PT:
                # a plain PT frame of the same camera beside it (1 spp, same depth)
                pt_out = torch.zeros_like(out)
                ps = snkpt.default_settings(max_depth=scene.s.max_depth, rr_start=scene.s.rr_start,
                                            seed=scene.s.seed, frame=last)
                ctx.render(frame_cam(last), ps, pt_out, w, h, stream=stream)
                perr = ((pt_out[..., :3].double() - ref) ** 2).mean().sqrt()
                frame_err["pt_frame_rel_rmse"] = round(float(perr / ref.mean()), 4)
                frame_err["pt_frame_rel_mae"] = round(float((pt_out[..., :3].double() - ref).abs().mean() / ref.mean()), 4)
            gms = None
            if label == "pt":
                # the same frame recorded into a HIP graph (no end-of-frame readback)
                # and replayed: the launch-bound part of a small frame goes away
                gs = settings(nw)
                gs.flags |= snkpt.FLAG_NO_STATS
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=stream):
                    ctx.render(cam, gs, out, w, h, stream=stream)
                for _ in range(2):
                    g.replay()
                gev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                       for _ in range(frames)]
                for a, b in gev:
                    a.record(stream)
                    g.replay()
                    b.record(stream)
                stream.synchronize()
                gms = sum(a.elapsed_time(b) for a, b in gev) / frames
                del g
        ms = sum(a.elapsed_time(b) for a, b in evs) / frames
        key = f"c{cfg}_{label}"
        res[key] = {"fps": round(1e3 / ms, 2), "ms_per_frame": round(ms, 3), "res": f"{w}x{h}",
                    "depth": scene.s.max_depth, "rays_per_frame": int(st.camera_rays + st.extension_rays +
                                                                      st.shadow_rays)}
        res[key]["mrays_s"] = round(res[key]["rays_per_frame"] / (ms * 1e-3) / 1e6, 1)
        if gms is not None:
            res[key]["graph_ms_per_frame"] = round(gms, 3)
        # compulsory bytes of a frame (BVH node/triangle reads excluded): 56 B per closest-hit ray
        # (ray 32 + path id 4 + hit 20), 40 B per shadow ray, 370 B per shaded path vertex
        # (hit, ray, path state, indices + normals, instance, material, inverse transform read;
        # next ray, path state, shadow ray written)
        hits = int(st.camera_rays + st.extension_rays)
        comp = (56.0 + 370.0) * hits + 40.0 * int(st.shadow_rays)
        res[key]["hbm_frac_compulsory"] = round(comp / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)
        # algorithmic bytes: rays x (56 + 80 N_node + 48 N_tri) with N counted on this
        # configuration's camera, bounce and shadow ray sets, + shaded vertices, ReSTIR, cache
        B = visit_means(ctx, scene, w, h, stream)
        alg = frame_algorithmic_bytes(st, B, mode, w * h)
        res[key]["bytes_per_ray"] = {k: round(v, 1) for k, v in B.items()}
        res[key]["algorithmic_bytes_per_frame"] = int(alg)
        res[key]["hbm_frac_algorithmic"] = round(alg / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)
        res[key]["error_vs_fp64_oracle"] = frame_err
        if mode == snkpt.MODE_RESTIR:
            res[key]["camera"] = "seeded 64-frame orbit" if cfg == 3 else "static"
            res[key]["history_reused_frac"] = round(int(st.restir_history_reused) / (w * h), 4)
        if mode == snkpt.MODE_CACHE:
            res[key]["cache_queries"] = int(st.cache_queries)
            res[key]["cache_train_samples"] = int(st.cache_train_samples)
        log(f"bench: {key}: {ms:.3f} ms/frame ({res[key]['mrays_s']} Mrays/s incl. shadow)")
        ctx.close()
    return res


def f_rows(stream, frames=10, warmup=3):
    """The rows beside the hot path, on one GPU: (1) frames whose primary
    vertices come from a G-buffer in the reference's formats (config 4 room,
    1080p, PT depth 8; the G-buffer is made on the GPU from traced camera rays:
    D32 clip depth, geometric normal RGBA16F, SNORM8 albedo, RGBA16F rma, R16
    flags) beside the same frames from camera rays; (2) the TAA resolve of a
    1080p RGBA32F frame; (3) short particle collision rays with contact records
    on the 1.02M-triangle scene.  Event-timed on the launch stream."""
    import numpy as np
    import torch

    import raysets
    import snkpt
    res = {}

    def timed(fn, n, w_up):
        for k in range(w_up):
            fn(k)
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
        for k, (a, b) in enumerate(evs):
            a.record(stream)
            fn(w_up + k)
            b.record(stream)
        stream.synchronize()
        return sum(a.elapsed_time(b) for a, b in evs) / n

    # (1) G-buffer input
    scene = snkpt.SceneData(4)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    w, h = scene.s.width, scene.s.height
    with torch.cuda.stream(stream):
        ctx.build_accel(stream)
        cam = scene.camera(0, w, h)
        porg, pdr = raysets.primary_rays(cam, w, h)
        po, pd = torch.from_numpy(porg).cuda(), torch.from_numpy(pdr).cuda()
        t, inst, prim, _, _ = ctx.trace(po, pd, stream=stream)
        stream.synchronize()
        world = torch.from_numpy(np.frombuffer(ctx.export(6).tobytes(), np.float32).reshape(-1, 12).copy()).cuda()
        _, _, inst5, _, mats = scene.arrays()
        off = torch.from_numpy(np.concatenate([[0], np.cumsum(inst5[:, 4] // 3)]).astype(np.int64)).cuda()
        hit = t >= 0
        ii = torch.where(hit, inst, torch.zeros_like(inst)).long()
        g = off[ii] + torch.where(hit, prim, torch.zeros_like(prim)).long()
        tri = world[g]
        v0, v1, v2 = tri[:, 0:3], tri[:, 3:6], tri[:, 6:9]
        nrm = torch.linalg.cross(v1 - v0, v2 - v0)
        nrm = nrm / nrm.norm(dim=1, keepdim=True).clamp_min(1e-30)
        d3 = pd[:, :3]
        nrm = torch.where(((nrm * d3).sum(1) > 0)[:, None], -nrm, nrm)
        p = po[:, :3] + d3 * t[:, None]
        pv = torch.tensor(np.array(cam.proj_view, np.float32).reshape(4, 4).T.copy()).cuda()
        clip = p @ pv[:, :3].T + pv[:, 3]
        depth = torch.where(hit, clip[:, 2] / clip[:, 3], torch.ones_like(t))
        m = torch.from_numpy(mats.view(np.float32).copy()).cuda()[inst5[:, 3].astype(np.int64)][ii]
        mflags = torch.from_numpy(mats[:, 12].astype(np.int32)).cuda()[torch.from_numpy(inst5[:, 3].astype(np.int64)).cuda()][ii]
        gb = {"depth": depth.reshape(h, w).contiguous(),
              "normal": torch.cat([nrm, torch.ones_like(t)[:, None]], 1).half().reshape(h, w, 4).contiguous(),
              "albedo": torch.cat([torch.round(m[:, 8:11].clamp(-1, 1) * 127.0), torch.full_like(t, 127.0)[:, None]],
                                  1).to(torch.int8).reshape(h, w, 4).contiguous(),
              "rma": torch.stack([m[:, 5], m[:, 6], torch.ones_like(t), torch.ones_like(t)], 1).half()
                          .reshape(h, w, 4).contiguous(),
              "mat_flags": torch.where(hit, mflags, torch.zeros_like(mflags)).to(torch.int16).reshape(h, w)
                                .contiguous()}
        out = torch.zeros(h, w, 4, device="cuda")

        def st(f):
            return snkpt.default_settings(max_depth=scene.s.max_depth, rr_start=scene.s.rr_start, seed=scene.s.seed,
                                          frame=f)
        gb_ms = timed(lambda f: ctx.render_gbuffer(cam, st(f), gb, out, w, h, stream=stream), frames, warmup)
        cam_ms = timed(lambda f: ctx.render(cam, st(f), out, w, h, stream=stream), frames, warmup)
        # (2) TAA resolve of the last frame (zero motion), ping-pong history inside the library
        cur = out.clone()
        vel = torch.zeros(h, w, 2, dtype=torch.float16, device="cuda")
        taa_out = torch.empty_like(cur)
        ctx.taa_resolve(cur, taa_out, vel, w, h, reset=True, stream=stream)
        taa_ms = timed(lambda f: ctx.taa_resolve(cur, taa_out, vel, w, h, stream=stream), 50, 5)
    ctx.close()
    # per pixel: current 16 B + history 16 B + motion 4 B + previous velocity 4 B read, output 16 B +
    # next history 16 B + next velocity 4 B written (neighbourhood reads served from LDS)
    taa_bytes = 76.0 * w * h
    res["gbuffer_input_c4_1080p"] = {
        "ms_per_frame": round(gb_ms, 3), "camera_rays_ms_per_frame": round(cam_ms, 3),
        "note": "PT depth 8, RR from 3; primary vertices from the G-buffer instead of traced camera rays"}
    res["taa_resolve_1080p"] = {"us": round(taa_ms * 1e3, 1), "bytes_per_pixel": 76,
                                "hbm_frac": round(taa_bytes / (taa_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
    # (3) contact rays on the 1.02M-triangle scene
    scene = snkpt.SceneData(2)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    n = 1 << 21
    rng = np.random.default_rng(7)
    pts = np.stack([rng.uniform(-50, 50, n), rng.uniform(0.0, 12.0, n), rng.uniform(-50, 50, n)], 1)
    dv = rng.normal(size=(n, 3))
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    r = rng.uniform(0.01, 0.08, n)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = pts - dv * r[:, None]
    rays[:, 4:7] = dv
    rays[:, 7] = rng.uniform(0.0, 30.0, n) / 120.0 + 2 * r
    with torch.cuda.stream(stream):
        ctx.build_accel(stream)
        o = torch.from_numpy(np.ascontiguousarray(rays[:, :4])).cuda()
        d = torch.from_numpy(np.ascontiguousarray(rays[:, 4:])).cuda()
        *_, cnv, ctv = ctx.trace(o, d, any_hit=True, counted=True, stream=stream)
        ct = ctx.trace_contacts(o, d, 1.0 / 120.0, stream=stream)
        stream.synchronize()
        c_ms = timed(lambda f: ctx.trace_contacts(o, d, 1.0 / 120.0, stream=stream), frames, warmup)
    ctx.close()
    nn_, nt_ = float(cnv.double().mean()), float(ctv.double().mean())
    # ray 32 B + nodes + triangles + contact record 60 B (t, normal, point, velocity, instance, primitive)
    cb = 32.0 + 80.0 * nn_ + 48.0 * nt_ + 60.0
    res["contacts_c2"] = {"rays": n, "mrays_s": round(n / (c_ms * 1e-3) / 1e6, 1), "us": round(c_ms * 1e3, 1),
                          "hit_fraction": round(float((ct[0] >= 0).float().mean()), 4),
                          "bytes_per_ray": round(cb, 1), "hbm_frac": round(n * cb / (c_ms * 1e-3) /
                                                                            (HBM_PEAK_GBS * 1e9), 4),
                          "note": "particle-style rays (origin behind a particle of radius 0.01-0.08, "
                                  "tmax = speed / 120 + 2 r), terminate on first hit"}
    log(f"bench: f rows: G-buffer frame {gb_ms:.3f} ms (camera rays {cam_ms:.3f}), TAA {taa_ms * 1e3:.1f} us, "
        f"contacts {res['contacts_c2']['mrays_s']} Mrays/s")
    return res


def batched_config5(stream, frames=10, warmup=2):
    """Config 5: 16 scenes (10k-2M triangles) x 64 views (128^2-1024^2) in one
    batch, depth 6, RR from bounce 3, 1 spp.  Reports frames/s of the whole
    batch eagerly, with the frame captured as a hipGraph, after the autotuner,
    and the same 64 views rendered one snkpt_render call at a time."""
    import torch

    import snkpt
    t0 = time.perf_counter()
    scenes, ctxs = [], []
    for v in range(16):
        sd = snkpt.SceneData(5, v)
        c = snkpt.Context(0)
        c.load_scene(sd)
        with torch.cuda.stream(stream):
            c.build_accel(stream)
        scenes.append(sd)
        ctxs.append(c)
    stream.synchronize()
    tris = [int(sd.s.n_idx // 3) for sd in scenes]
    log(f"bench: config 5: 16 scenes built ({min(tris)}..{max(tris)} tris) in {time.perf_counter() - t0:.1f} s")
    b = snkpt.Batch(0)
    for c in ctxs:
        b.add_scene(c)
    views, outs = [], []
    for view in range(64):
        var, w, h = snkpt.scene_view(5, view)
        cam = scenes[var].camera(view, w, h)
        out = torch.zeros(h, w, 4, device="cuda")
        b.add_view(var, cam, w, h, 0x5000 + view, out)
        views.append((var, w, h, cam))
        outs.append(out)
    torch.cuda.synchronize()  # the zeroed outputs before the batch's stream writes them
    pixels = sum(w * h for _, w, h, _ in views)
    depth, rr = 6, 3

    def settings(f):
        return snkpt.default_settings(mode=snkpt.MODE_PT, max_depth=depth, rr_start=rr, frame=f)

    def timed(fn):
        for f in range(warmup):
            fn(f)
        a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for f in range(frames):
            fn(warmup + f)
        e.record(stream)
        stream.synchronize()
        return a.elapsed_time(e) / frames

    eager_ms = timed(lambda f: b.render(settings(f), use_graph=False, stream=stream))
    st = b.frame_stats(stream)
    rays = int(st.camera_rays + st.extension_rays + st.shadow_rays)
    graph_ms = timed(lambda f: b.render(settings(f), use_graph=True, stream=stream))
    t0 = time.perf_counter()
    tuning = b.autotune(settings(100), stream=stream)
    tune_s = time.perf_counter() - t0
    tuned_ms = timed(lambda f: b.render(settings(f), use_graph=True, stream=stream))

    def one_by_one(f):
        for k, (var, w, h, cam) in enumerate(views):
            s = settings(f)
            s.seed = 0x5000 + k
            ctxs[var].render(cam, s, outs[k], w, h, stream=stream)

    single_ms = timed(one_by_one)
    # the last frame of the smallest, a middle and the largest view (equal to the
    # batch's frame, as its tests check) against a 256-spp FP64 oracle of its camera
    # (torch work on the oracle's stream: its sums are read after it has written them)
    view_err = {}
    by_size = sorted(range(len(views)), key=lambda k: views[k][1] * views[k][2])
    with torch.cuda.stream(stream):
        for k in (by_size[0], by_size[len(by_size) // 2], by_size[-1]):
            var, w, h, cam = views[k]
            acc = torch.zeros(h * w * 3, dtype=torch.float64, device="cuda")
            os_ = snkpt.default_settings(max_depth=depth, rr_start=rr, seed=0x0AC1E)
            for j in range(4):
                os_.frame = j
                ctxs[var].oracle_accumulate(cam, os_, acc, w, h, 64, stream=stream)
            ref = (acc / 256).view(h, w, 3)
            dv = outs[k][..., :3].double() - ref
            view_err[f"view {k} ({w}x{h})"] = {
                "frame_rel_rmse": round(float((dv ** 2).mean().sqrt() / ref.mean()), 4),
                "frame_rel_mae": round(float(dv.abs().mean() / ref.mean()), 4)}

    def cache_settings(f):
        return snkpt.default_settings(mode=snkpt.MODE_CACHE, max_depth=depth, rr_start=rr, frame=f,
                                      cache_query_depth=1, flags=snkpt.FLAG_CACHE_INDIRECT)

    for c in ctxs:
        c.cache_reset(1, stream=stream)
    cache_ms = timed(lambda f: b.render(cache_settings(f), use_graph=False, stream=stream))
    cst = b.frame_stats(stream)
    cache_graph_ms = timed(lambda f: b.render(cache_settings(f), use_graph=True, stream=stream))
    log(f"bench: config 5 ({pixels} px / frame): eager {eager_ms:.2f} ms, graph {graph_ms:.2f} ms, "
        f"autotuned graph {tuned_ms:.2f} ms, one view per call {single_ms:.2f} ms, cache form {cache_ms:.2f} ms "
        f"eager / {cache_graph_ms:.2f} ms graph")
    res = {"workload": "config 5: 16 seeded room scenes, 10k-2M triangles log-uniform, 64 views of 128^2-1024^2 "
                       "log-uniform, plain PT depth 6, RR from bounce 3, 1 spp",
           "triangles_min_max": [min(tris), max(tris)], "pixels_per_frame": pixels, "rays_per_frame": rays,
           "eager_ms": round(eager_ms, 3), "graph_ms": round(graph_ms, 3), "autotuned_graph_ms": round(tuned_ms, 3),
           "one_view_per_call_ms": round(single_ms, 3), "fps": round(1e3 / tuned_ms, 2),
           "cache_form_ms": round(cache_ms, 3), "cache_form_graph_ms": round(cache_graph_ms, 3),
           "cache_form_queries": int(cst.cache_queries),
           "cache_form_train_samples": int(cst.cache_train_samples),
           "cache_form": "NEE + indirect from each scene's own network at the primary hit, one training step per "
                         "scene per frame (65,536 training paths split by pixels); eager and as a replayed "
                         "hipGraph (autotuned launch shapes)",
           "mrays_s": round(rays / (tuned_ms * 1e-3) / 1e6, 1), "tuning": tuning.as_dict(),
           "autotune_s": round(tune_s, 2),
           "error_vs_fp64_oracle": {"oracle_spp": 256, "views": view_err}}
    b.close()
    for c in ctxs:
        c.close()
    return res


def time_to_rmse(stream, ref_spp=4096, max_frames=600, warm_frames=256):
    """1080p room (config 4): GPU time for the running mean of 1-spp frames to
    reach a relative RMSE (RMSE / the oracle's mean radiance) against the FP64
    oracle, plain PT against the MLP cache started cold: a freshly initialised
    network, every frame timed (render, training paths, optimiser steps), the
    error recorded from the first frame on.  PT's frames count equally; the
    cache's frames count (f + 1)^3 in the mean (SNKPT_FLAG_ACCUMULATE with
    accum_weight), so the still-learning network's early frames fade out
    instead of being discarded.  Thresholds 0.1, 0.05, 0.03 (one PT frame is at
    0.43).  The warm row (network trained for `warm_frames` frames before, not
    timed) is shown for context only and is not a result.  Frame times are GPU
    event times of the render calls; the error checks between frames are not
    timed."""
    import torch

    import snkpt
    scene = snkpt.SceneData(4)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    w, h = scene.s.width, scene.s.height
    depth, rr = scene.s.max_depth, scene.s.rr_start
    with torch.cuda.stream(stream):
        ctx.build_accel(stream)
        cam = scene.camera(0, w, h)
        acc = torch.zeros(h * w * 3, dtype=torch.float64, device="cuda")
        s = snkpt.default_settings(max_depth=depth, rr_start=rr, seed=scene.s.seed ^ 0x0AC1E)
        t0 = time.perf_counter()
        for k in range(ref_spp // 64):
            s.frame = k
            ctx.oracle_accumulate(cam, s, acc, w, h, 64, stream=stream)
        stream.synchronize()
        ref = (acc / ref_spp).view(h, w, 3)
        ref_mean = float(ref.mean())
        log(f"bench: FP64 oracle {ref_spp} spp in {time.perf_counter() - t0:.1f} s, mean {ref_mean:.5f}")
        out = torch.zeros(h, w, 4, device="cuda")
        ctx.render(cam, snkpt.default_settings(max_depth=depth), out, w, h, stream=stream)  # sizes the buffers
        stream.synchronize()
        rel_thr = [0.1, 0.05, 0.03]

        def race(mode, warm=0, weight=None, early_steps=None, **kw):
            if mode == snkpt.MODE_CACHE:
                ctx.cache_reset(1, stream=stream)
                for f in range(warm):
                    ctx.render(cam, snkpt.default_settings(mode=mode, max_depth=depth, rr_start=rr,
                                                           seed=scene.s.seed, frame=100000 + f, **kw),
                               out, w, h, stream=stream)
            t_ms, first_rmse, hit_rel, curve = 0.0, None, {}, []
            for f in range(max_frames):
                fl = snkpt.FLAG_ACCUMULATE | (snkpt.FLAG_RESET if f == 0 else 0)
                kw2 = dict(kw)
                fl |= kw2.pop("flags", 0)
                if early_steps and f < early_steps[0]:  # (frames, optimiser steps per frame) at the start
                    kw2["cache_train_steps"] = early_steps[1]
                st = snkpt.default_settings(mode=mode, max_depth=depth, rr_start=rr, seed=scene.s.seed, frame=f,
                                            flags=fl, accum_weight=weight(f) if weight else 1.0, **kw2)
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                ctx.render(cam, st, out, w, h, stream=stream)
                b.record(stream)
                stream.synchronize()
                t_ms += a.elapsed_time(b)
                e = float(torch.sqrt(((out[..., :3].double() - ref) ** 2).mean())) / ref_mean
                if f == 0:
                    first_rmse = e
                if ((f + 1) & f) == 0:
                    curve.append([f + 1, round(t_ms, 2), round(e, 5)])
                for thr in rel_thr:
                    if str(thr) not in hit_rel and e <= thr:
                        hit_rel[str(thr)] = round(t_ms, 2)
                if len(hit_rel) == len(rel_thr):
                    break
            return {"ms_to_rel_rmse": hit_rel, "frames": f + 1, "ms_per_frame": round(t_ms / (f + 1), 3),
                    "first_frame_rel_rmse": round(first_rmse, 4), "final_rel_rmse": round(e, 5),
                    "warm_frames_untimed": warm, "curve_frame_ms_relrmse": curve}

        pt = race(snkpt.MODE_PT)
        log(f"bench: PT time-to-RMSE rel {pt['ms_to_rel_rmse']}")
        cube = lambda f: float(f + 1) ** 3  # noqa: E731
        ind_mis = snkpt.FLAG_CACHE_INDIRECT | snkpt.FLAG_CACHE_MIS
        atq = snkpt.FLAG_CACHE_TRAIN_AT_QUERY
        forms = {"cache_q1_indirect_mis": dict(cache_query_depth=1, flags=ind_mis | atq, weight=cube, cache_ema=0.0,
                                               cache_lr=2.5e-3, cache_train_paths=32768, early_steps=(16, 2)),
                 "cache_q1_indirect_mis_1step": dict(cache_query_depth=1, flags=ind_mis | atq, weight=cube,
                                                     cache_ema=0.0, cache_lr=2.5e-3, cache_train_paths=32768),
                 "cache_q1_indirect_nee_only": dict(cache_query_depth=1, flags=snkpt.FLAG_CACHE_INDIRECT | atq,
                                                    weight=cube, cache_ema=0.0, cache_lr=2.5e-3,
                                                    cache_train_paths=32768),
                 "cache_q1_indirect_mis_all_vertices": dict(cache_query_depth=1, flags=ind_mis, weight=cube,
                                                            cache_ema=0.0, cache_lr=5e-3, cache_train_paths=32768),
                 "cache_q3_full": dict(cache_query_depth=3, weight=cube, cache_ema=0.0, cache_lr=5e-3,
                                       cache_train_paths=32768),
                 "warm_not_a_result_q1_indirect_mis": dict(cache_query_depth=1, flags=ind_mis | atq,
                                                           warm=warm_frames, cache_ema=0.0, cache_lr=2.5e-3,
                                                           cache_train_paths=32768)}
        caches = {}
        for name, kw in forms.items():
            caches[name] = race(snkpt.MODE_CACHE, **kw)
            log(f"bench: {name} time-to-RMSE rel {caches[name]['ms_to_rel_rmse']}")
    for name, c in caches.items():
        ratio = {}
        for k, v in pt["ms_to_rel_rmse"].items():
            t = c["ms_to_rel_rmse"].get(k)
            ratio[k] = round(v / t, 3) if t else None
        c["speedup_vs_pt"] = ratio
    ctx.close()
    return {"workload": "config 4: room at 1920x1080, 1 spp frames, depth 8, RR from bounce 3, cold start, every "
                        "frame timed (render, training paths, optimiser steps). cache_q1_indirect_mis: NEE + a traced "
                        "BRDF sample under the power heuristic at the primary hit, the cache answers its indirect "
                        "light; it trains on the primary vertices of 32,768 training paths (where it is asked), two "
                        "Adam steps (lr 2.5e-3) per frame for the first 16 frames, then one; _1step: one step per "
                        "frame throughout; _nee_only: NEE alone at the primary hit (no BRDF probe); _all_vertices: "
                        "training samples from every vertex of the training paths, lr 5e-3; cache_q3_full: the cache "
                        "answers "
                        "L_o at path vertex 3 (after 2 bounces); cache frames weighted (f + 1)^3 in the running mean "
                        "(the still-learning network's early frames fade out), PT frames equally",
            "oracle": {"spp": ref_spp, "mean_radiance": round(ref_mean, 6),
                       "rel_noise_est": round(pt["first_frame_rel_rmse"] / ref_spp ** 0.5, 5)},
            "pt": pt, **caches}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only-trace", action="store_true", help="time only the incoherent trace (profiling)")
    ap.add_argument("--no-frames", action="store_true", help="skip the per-config frame rates")
    ap.add_argument("--no-rmse", action="store_true", help="skip the time-to-RMSE race (PT vs cache)")
    ap.add_argument("--leg", default=None,
                    help="run one workload only and print its record (profiling): "
                         + ", ".join(f"c{c}_{l}" for c, l in FRAME_RUNS) + ", c5_batched, rmse, f_rows")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calib:
        import torch
        n = args.copy_calib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes": n * 4, "reps": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path is built and "
              "measured on one MI355X; how it would shard across GPUs is a design, not code)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import raysets
    import snkpt

    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    if args.leg:
        if args.leg == "c5_batched":
            rec = batched_config5(stream)
        elif args.leg == "rmse":
            rec = time_to_rmse(stream)
        elif args.leg == "f_rows":
            rec = f_rows(stream)
        else:
            rec = frame_rates(stream, frames=args.steps, warmup=args.warmup, only=args.leg)
            if not rec:
                print(f"unknown leg {args.leg}", file=sys.stderr)
                sys.exit(2)
        print(json.dumps({"leg": args.leg, "record": rec}), flush=True)
        return
    W, H = 1920, 1080
    log("bench: generating the 1.02M-triangle scene")
    scene = snkpt.SceneData(2)
    ctx = snkpt.Context(0)
    ctx.load_scene(scene)
    with torch.cuda.stream(stream):
        ctx.build_accel(stream)
        build_ms = [ctx.accel_info().build_ms]
        for _ in range(2):
            ctx.build_accel(stream)
            build_ms.append(ctx.accel_info().build_ms)
    info = ctx.accel_info()
    log(f"bench: BVH8 nodes={info.num_nodes} depth={info.depth} build={min(build_ms):.2f} ms")
    # dynamic scene: new instance transforms every frame (the reference rebuilds its
    # TLAS per frame), meshes unchanged; wall time of set_instances + build_accel
    _, _, inst5, m4, _ = scene.arrays()
    m4 = np.asarray(m4, np.float32).reshape(-1, 16)
    dyn_ms = []
    with torch.cuda.stream(stream):
        for k in range(4):
            moved = m4.copy()
            moved[1:, 12] += 0.01 * (k + 1)  # column-major: x translation of every transform but the ground's
            stream.synchronize()
            t0 = time.perf_counter()
            ctx.set_instances(inst5, moved)
            ctx.build_accel(stream)
            stream.synchronize()
            dyn_ms.append((time.perf_counter() - t0) * 1e3)
        ctx.set_instances(inst5, m4)
        ctx.build_accel(stream)
    info = ctx.accel_info()
    log(f"bench: transform update + rebuild {min(dyn_ms[1:]):.2f} ms wall")

    with torch.cuda.stream(stream):
        porg, pdr = raysets.primary_rays(scene.camera(0, W, H), W, H)
        po = torch.from_numpy(porg).to(dev)
        pd = torch.from_numpy(pdr).to(dev)
        pt, pinst, pprim, _, _ = ctx.trace(po, pd, stream=stream)
        world_np = np.frombuffer(ctx.export(6).tobytes(), np.float32).reshape(-1, 12)
        inst_np = scene.arrays()[2]
        off = torch.from_numpy(np.concatenate([[0], np.cumsum(inst_np[:, 4] // 3)]).astype(np.int64)).to(dev)
        wt = torch.from_numpy(world_np.copy()).to(dev)
        so, sd = raysets.secondary_rays_torch(po, pd, pt, pinst, pprim, wt, off, 100.0, seed=1)
        n = so.shape[0]
        t = torch.empty(n, dtype=torch.float32, device=dev)
        hi = torch.empty(n, dtype=torch.int32, device=dev)
        hp = torch.empty(n, dtype=torch.int32, device=dev)
        # visit counts on the identical ray set (bit-identical to the CPU BVH8 traversal)
        _, _, _, _, _, nv, tv = ctx.trace(so, sd, counted=True, stream=stream)
        stream.synchronize()
    n_node = float(nv.double().mean())
    n_tri = float(tv.double().mean())
    hit_frac = float((pt >= 0).float().mean())
    log(f"bench: {n} incoherent rays, mean nodes {n_node:.2f}, tris {n_tri:.2f}")

    step = lambda: ctx.trace_into(so, sd, t, hi, hp, stream=stream)
    settle_clocks(step, stream)
    for _ in range(args.warmup):
        step()
    stream.synchronize()
    torch.cuda.synchronize()
    # timed region: K steps back to back, HIP events on the launch stream at its two
    # ends (an event pair around every launch would add its own gap to each step);
    # the average launch duration includes the gaps between launches
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    secondary_hit = float((t >= 0).float().mean())

    bytes_per_ray = 56.0 + 80.0 * n_node + 48.0 * n_tri
    achieved = n * bytes_per_ray / (launch_ms * 1e-3) / 1e9
    traffic, prof = None, {}
    if os.path.exists(TRAFFIC_FILE):
        try:
            prof = json.load(open(TRAFFIC_FILE))
            traffic = prof.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    extra = {}
    if not args.only_trace:
        with torch.cuda.stream(stream):
            pt2 = torch.empty(po.shape[0], dtype=torch.float32, device=dev)
            prim_ms = time_launches(lambda: ctx.trace_into(po, pd, pt2, stream=stream), 10, stream)
            emit_tris = []
            pos, idx, inst5, m4, mats = scene.arrays()
            # emitter vertices from the world triangles of emissive instances
            em_inst = np.ctypeslib.as_array(scene.s.emissive_instances, (scene.s.n_emissive,))
            off_np = np.concatenate([[0], np.cumsum(inst5[:, 4] // 3)]).astype(np.int64)
            for i in em_inst:
                for g in range(off_np[i], off_np[i + 1]):
                    emit_tris.append(world_np[g, :9])
            et = torch.from_numpy(np.array(emit_tris, np.float32)).to(dev)
            ea = torch.linalg.cross(et[:, 3:6] - et[:, 0:3], et[:, 6:9] - et[:, 0:3]).norm(dim=1)
            shadow_o, shadow_d = raysets.shadow_rays_torch(po, pd, pt, et, ea / ea.sum(), seed=2, scale=100.0)
            ns = shadow_o.shape[0]
            st_ = torch.empty(ns, dtype=torch.float32, device=dev)
            sh_ms = time_launches(lambda: ctx.trace_into(shadow_o, shadow_d, st_, any_hit=True, stream=stream), 10,
                                  stream)
        extra["primary_mrays_s"] = round(po.shape[0] / (prim_ms * 1e-3) / 1e6, 2)
        extra["shadow_mrays_s"] = round(ns / (sh_ms * 1e-3) / 1e6, 2)
        extra["shadow_rays"] = ns
        # the same roofline for the primary (closest hit, 56 B + nodes + triangles)
        # and shadow (any hit, 40 B + nodes + triangles) ray sets, from their own
        # deterministic visit counts
        with torch.cuda.stream(stream):
            *_, pnv, ptv = ctx.trace(po, pd, counted=True, stream=stream)
            *_, snv, stv = ctx.trace(shadow_o, shadow_d, any_hit=True, counted=True, stream=stream)
            stream.synchronize()
        for key, nv_, tv_, base, ms_, cnt in (("primary", pnv, ptv, 56.0, prim_ms, po.shape[0]),
                                               ("shadow", snv, stv, 40.0, sh_ms, ns)):
            b = base + 80.0 * float(nv_.double().mean()) + 48.0 * float(tv_.double().mean())
            extra[f"{key}_bytes_per_ray"] = round(b, 1)
            extra[f"{key}_roofline_frac"] = round(cnt * b / (ms_ * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)
    extra.update({"incoherent_rays": n, "primary_hit_fraction": round(hit_frac, 4),
                  "secondary_hit_fraction": round(secondary_hit, 4), "mean_nodes_visited": round(n_node, 3),
                  "mean_tris_visited": round(n_tri, 3), "bvh8_nodes": info.num_nodes, "bvh8_depth": info.depth,
                  "gpu_build_ms": round(min(build_ms), 3), "trace_tuning": ctx.trace_tuning().as_dict(),
                  "transform_update_rebuild_ms_wall": round(min(dyn_ms[1:]), 3), "launch_ms": round(launch_ms, 4)})

    if not args.only_trace and not args.no_frames:
        extra["frames"] = frame_rates(stream)
    if not args.only_trace and not args.no_frames:
        extra["frames"]["c5_batched_pt"] = batched_config5(stream)
        extra["f_rows"] = f_rows(stream)
    if not args.only_trace and not args.no_rmse:
        extra["time_to_rmse"] = time_to_rmse(stream)

    cpu = hits_match = None
    if not args.no_cpu and not args.only_trace:
        threads = max(1, min(16, os.cpu_count() or 1))
        rays_np = torch.cat([so, sd], 1).cpu().numpy()
        log(f"bench: CPU baseline on {threads} threads")
        cpu, hits_match = cpu_baseline(world_np, rays_np, threads,
                                       (t.cpu().numpy(), hi.cpu().numpy(), hp.cpu().numpy()))
        log(f"bench: GPU hits equal to the CPU traversal: {hits_match['equal']} of {hits_match['rays']}")

    out = {
        "metric": "Mrays/s incoherent secondary (closest hit), 1.02M-triangle scene, 1080p, 1 GPU",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene and rays)",
        "config": {"workload": "config 2: ground + 200 level-4 icospheres (1,024,002 tris), 1920x1080, "
                               "one cosine bounce per primary hit", "rays_per_step": n, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_ray": round(bytes_per_ray, 1),
                     # compulsory bytes only (ray 32 + path id 4 + hit 20 per ray, no BVH reads)
                     "compulsory_frac": round(n * 56.0 / (launch_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                     # rocprofv3 counters of the same launch (profiles/): HBM bytes over this run's
                     # launch time, and the L2 / L1 hit rates
                     "measured_hbm_gbs": (round(traffic / (launch_ms * 1e-3) / 1e9, 1) if traffic else None),
                     "l2_hit_rate": prof.get("l2_hit_rate"), "l1_hit_rate": prof.get("l1_hit_rate"),
                     "note": "algorithmic bytes = 56 + 80*nodes + 48*tris per ray; the BVH (~70 MB) is "
                             "cache-resident, so frac > 1 would mean cache reuse"},
        "cpu_baseline": cpu,
        "hits_match": hits_match,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
