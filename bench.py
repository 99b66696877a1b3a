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


def cpu_baseline(world, rays_np, threads):
    """The CPU restatement (oracle BVH8 traversal) on the same incoherent rays."""
    from oracle import binding as O
    t0 = time.perf_counter()
    bvh = O.Bvh(world)
    build_s = time.perf_counter() - t0
    n = rays_np.shape[0]
    t0 = time.perf_counter()
    bvh.trace(rays_np, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"all {n} incoherent secondary rays of the 1080p ray set, closest hit, CPU BVH8 "
                      f"restatement (oracle), {threads} threads; CPU LBVH->BVH8 build of the 1.02M "
                      f"triangles took {build_s:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only-trace", action="store_true", help="time only the incoherent trace (profiling)")
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
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (every box has one MI355X; "
              "the multi-GPU sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import raysets
    import snkpt

    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
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
    for _ in range(args.warmup):
        step()
    stream.synchronize()
    torch.cuda.synchronize()
    # timed region: K steps, per-launch events on the launch stream
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = sum(a.elapsed_time(b) for a, b in evs) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    secondary_hit = float((t >= 0).float().mean())

    bytes_per_ray = 56.0 + 80.0 * n_node + 48.0 * n_tri
    achieved = n * bytes_per_ray / (launch_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            traffic = json.load(open(TRAFFIC_FILE)).get("hbm_bytes_per_launch")
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
    extra.update({"incoherent_rays": n, "primary_hit_fraction": round(hit_frac, 4),
                  "secondary_hit_fraction": round(secondary_hit, 4), "mean_nodes_visited": round(n_node, 3),
                  "mean_tris_visited": round(n_tri, 3), "bvh8_nodes": info.num_nodes, "bvh8_depth": info.depth,
                  "gpu_build_ms": round(min(build_ms), 3), "launch_ms": round(launch_ms, 4)})

    cpu = None
    if not args.no_cpu and not args.only_trace:
        threads = max(1, min(16, os.cpu_count() or 1))
        rays_np = torch.cat([so, sd], 1).cpu().numpy()
        log(f"bench: CPU baseline on {threads} threads")
        cpu = cpu_baseline(world_np, rays_np, threads)

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
                     "note": "algorithmic bytes = 56 + 80*nodes + 48*tris per ray; the BVH (~70 MB) is "
                             "cache-resident, so frac > 1 would mean cache reuse"},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
