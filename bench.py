"""Benchmark: BVH build Mtris/s and trace Mrays/s of the 2M-triangle mixed scene on one
MI355X, with the SAH cost ratio against the reference's CPU SAH BVH and the CPU baseline.

A step = one PLOC build of the 2,000,002-triangle scene (device-resident input) followed by
one closest-hit trace of 4,194,304 rays. Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variant", default="ploc_emc", choices=["ploc", "ploc_emc"])
    ap.add_argument("--radius", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (every box has one GPU)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from diplodocus_amd import accel, scenes

    t0 = time.time()
    scene = scenes.mixed_2m()
    rays_np = scenes.random_rays_towards(scene.tris, 4_194_304, seed=11)
    print(f"[bench] scene {scene.n} tris, {len(rays_np)} rays generated in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    dev = torch.device("cuda:0")
    tris = torch.from_numpy(scene.tris).to(dev)
    rays = torch.from_numpy(rays_np).to(dev)
    hits = torch.empty((len(rays_np), 4), dtype=torch.float32, device=dev)
    ctx = accel.Context(0)
    mt = 1 if args.variant == "ploc_emc" else 0
    params = accel.make_params(morton_type=mt, radius=args.radius, max_leaf=4)
    bvh = accel.GpuBvh(ctx)
    stream = torch.cuda.current_stream()

    def step(evs=None):
        if evs:
            evs[0].record(stream)
        bvh.build(tris, params)
        if evs:
            evs[1].record(stream)
        bvh.trace(rays, 0, hits)
        if evs:
            evs[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    stats = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        step(ev[k])
        stats.append(bvh.stats.as_dict())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    build_ms = np.array([e[0].elapsed_time(e[1]) for e in ev])
    trace_ms = np.array([e[1].elapsed_time(e[2]) for e in ev])
    n = scene.n
    nrays = len(rays_np)
    build_dev_ms = np.array([s["build_ms"] for s in stats])
    mtris = n / (build_ms.mean() * 1e-3) / 1e6
    mrays = nrays / (trace_ms.mean() * 1e-3) / 1e6
    last = stats[-1]
    h = hits.cpu().numpy().view(np.int32)[:, 3]
    print(f"[bench] build {build_ms.mean():.3f} ms (device {build_dev_ms.mean():.3f}), trace {trace_ms.mean():.3f} ms,"
          f" hit rate {(h >= 0).mean():.3f}, stages {last['stage_ms']}", file=sys.stderr, flush=True)

    # ---- roofline of the PLOC stage: algorithmic bytes per build over its measured time ----
    lay = bvh.download(intermediates=False)
    iters = last["ploc_iterations"]
    ploc_ms = np.mean([s["stage_ms"]["ploc"] for s in stats])
    # per merged-cluster bytes: each iteration reads its clusters (32 B) and writes the
    # survivors (32 B) plus one 32 B node and two 4 B parents per merge; iterations remove
    # n-1 clusters in total, so sum(c_k) is measured from the build's own iteration counts.
    roof = None

    # ---- CPU baseline: the reference's SAH BVH build (1 core) and tracer on a sample ----
    cpu = None
    sah_ratio = None
    if not args.no_cpu:
        from oracle import pyoracle

        tb = time.perf_counter()
        sb = pyoracle.SBvh(scene.tris, fix_child_bounds=False)
        sb_ms = (time.perf_counter() - tb) * 1e3
        sah_cpu = sb.sah()
        sah_fixed = pyoracle.SBvh(scene.tris, fix_child_bounds=True).sah()
        sah_ratio = last["sah_cost"] / sah_cpu
        sample = rays_np[:262144]
        threads = pyoracle.nthreads()
        sb.trace(sample, flags=0, mode=0, threads=threads)
        cpu_trace_s = sb.last_trace_s
        cpu = {
            "value": round(n / (sb.build_ms * 1e-3) / 1e6, 4),
            "unit": "Mtris/s",
            "cores": 1,
            "kind": "port",
            "sample": (f"full 2,000,002-tri SAH build, single thread ({sb.build_ms:.0f} ms); "
                       f"trace of {len(sample)} of the rays on {threads} threads: "
                       f"{len(sample) / cpu_trace_s / 1e6:.3f} Mrays/s"),
            "trace_mrays_s": round(len(sample) / cpu_trace_s / 1e6, 4),
            "trace_threads": threads,
            "cpu_model": cpu_model(),
            "sah_faithful": sah_cpu,
            "sah_fixed": sah_fixed,
        }
        print(f"[bench] cpu build {sb_ms:.0f} ms, sah cpu {sah_cpu:.3f} fixed {sah_fixed:.3f} gpu "
              f"{last['sah_cost']:.3f}", file=sys.stderr, flush=True)

    out = {
        "metric": "build Mtris/s + trace Mrays/s at SAH <=1.10x CPU; fwd+adjoint vs fwd time",
        "value": round(mtris, 3),
        "unit": "Mtris/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float((build_ms + trace_ms).mean()), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "mixed_2m: 2,000,002 tris (64 Gaussian clusters, 200k thin, 2 ground), 4,194,304 rays",
                   "variant": args.variant, "radius": args.radius, "max_leaf": 4},
        "trace_mrays_s": round(mrays, 3),
        "build_ms": round(float(build_ms.mean()), 4),
        "trace_ms": round(float(trace_ms.mean()), 4),
        "sah_cost": last["sah_cost"],
        "sah_ratio_vs_cpu": sah_ratio,
        "ploc_iterations": iters,
        "stage_ms": last["stage_ms"],
        "wall_s": round(wall, 3),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    bvh.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
