"""Benchmark: BVH build Mtris/s and trace Mrays/s on one MI355X, with the SAH cost ratio
against the reference's CPU SAH BVH and the CPU baseline timed on the same host.

A step = one PLOC build of the 2,000,002-triangle mixed scene (device-resident input)
followed by one closest-hit trace of its 4,194,304 rays. `value` is the build rate
(Mtris/s). Beside it: the same trace over the BVH's wide forms (4 and 8 children),
primary-ray trace rate of the 1024x1024 camera of the 69,668-triangle icosphere room,
traversal counters, the roofline of the PLOC iteration kernel (the build's dominant kernel)
and the CPU baseline. Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# AccelerationStructureType -> (Morton type, bounding volume), src/gpu/config/gpu_config_bridge.cc:27-53
VARIANTS = {"ploc": (0, 0), "ploc_emc": (1, 0), "ploc_sobb": (0, 1), "ploc_emc_sobb": (1, 1)}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


PMC_PROFILE = os.path.join(ROOT, "profiles", "build_trace_mixed2m_pmc.md")


def profile_value(key: str, variant: str, radius: int):
    """The figure `key: <variant> r<radius> = V` recorded in the counter profile (rocprofv3
    --pmc passes summarised by --pmc-summary), or None when it holds none for this variant and
    radius."""
    try:
        for line in open(PMC_PROFILE):
            if line.startswith(key + ":"):
                k, _, val = line.split(":", 1)[1].partition("=")
                if k.strip() == f"{variant} r{radius}":
                    return float(val.strip())
    except OSError:
        pass
    return None


def measured_traffic(variant: str, radius: int):
    """HBM bytes per merging PLOC-iteration launch for this workload, from the counter profile
    (None when it has no measurement of this variant and radius)."""
    v = profile_value("traffic_bytes_per_launch", variant, radius)
    return None if v is None else int(v)


def measured_build_traffic(variant: str, radius: int, build_ms: float):
    """HBM bytes of one whole build (every kernel between the build's first and last event),
    from the counter profile, with the rate they give over this run's build time; None when no
    measurement is recorded."""
    b = profile_value("build_traffic_bytes", variant, radius)
    if b is None:
        return None
    gbs = b / (build_ms * 1e-3) / 1e9
    return {"bytes": int(b), "gbs": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 5),
            "source": os.path.relpath(PMC_PROFILE, ROOT)}


def trace_roofline(variant: str, radius: int, nrays: int, trace_ms: float, touched_per_ray: float):
    """The trace's bytes and instruction issue. Bytes: the tree bytes a ray touches (64 per node
    visit, 48 per triangle test, from the traversal counters) and the bytes per ray the L2
    fetched from the fabric (counter profile), each as a rate over this run's trace time
    against HBM's 8 TB/s (the scene's ~220 MB of nodes and triangles also fit the 256 MB
    Infinity Cache, so fabric bytes are not all HBM bytes). Issue: VALU wave-instructions per
    CU cycle and the fraction of lanes active in them, from the same profile."""
    per_ray = profile_value("trace_bytes_per_ray", variant, radius)
    issue = profile_value("trace_valu_per_cu_cycle", variant, radius)
    lanes = profile_value("trace_valu_lane_util", variant, radius)
    hit = profile_value("trace_l2_hit", variant, radius)
    touched_gbs = touched_per_ray * nrays / (trace_ms * 1e-3) / 1e9
    out = {"kernel": "KTraceRefill<256, SOBB, false, false, 4> (closest hit)", "touched_bytes_per_ray": round(touched_per_ray, 1),
           "touched_gbs": round(touched_gbs, 2), "touched_frac": round(touched_gbs / HBM_PEAK_GBS, 5),
           "counter_bytes_per_ray": None if per_ray is None else round(per_ray, 1),
           "counter_gbs": None, "counter_frac": None, "l2_hit_ratio": hit,
           "valu_per_cu_cycle": issue, "valu_lane_util": lanes,
           "source": os.path.relpath(PMC_PROFILE, ROOT)}
    if per_ray is not None:
        g = per_ray * nrays / (trace_ms * 1e-3) / 1e9
        out["counter_gbs"] = round(g, 2)
        out["counter_frac"] = round(g / HBM_PEAK_GBS, 5)
    return out


def build_algorithmic_bytes(n: int, morton_type: int, digit_bits: int, iter_counts, kept_nodes: int) -> dict:
    """Algorithmic bytes of one build, counted from sizes (never from the profiler): scene
    bounds 36 n; codes 36 n read + key n + id 4 n + leaf box 24 n written; the radix sort's
    passes read and write (key, id) pairs plus one histogram read of the keys; PLOC iteration
    k with c_k clusters and m_k = c_k - c_k+1 merges: 28 c_k window read + 4 c_k neighbour
    write + 8 c_k mutual check + 32 m_k node write + 28 c_k compaction read + 28 c_k+1
    compaction write + 4 c_k scan; layout: the final nodes once (64 B each) and the ordered
    triangles (36 n gathered, 48 n written)."""
    kb = 8 if morton_type == 1 else 4
    bits = 60 if morton_type == 1 else 30
    passes = -(-bits // digit_bits)
    c = [int(x) for x in iter_counts] + [1]
    ploc = sum(28 * c[k] + 4 * c[k] + 8 * c[k] + 32 * (c[k] - c[k + 1]) + 28 * c[k] + 28 * c[k + 1] + 4 * c[k]
               for k in range(len(c) - 1))
    st = {"bounds": 36 * n, "codes": (36 + kb + 4 + 24) * n, "sort": passes * 2 * (kb + 4) * n + kb * n,
          "ploc": ploc, "layout": 64 * kept_nodes + (36 + 48) * n}
    st["total"] = sum(st.values())
    return st


def _kname(full: str) -> str:
    """Kernel name without return type, namespace and argument list (rocprofv3 reports the
    demangled signature, e.g. `void dpl::KCodes<1>(float const*, ...)`)."""
    full = full.strip()
    if full.startswith("void "):
        full = full[5:]
    if full.startswith("dpl::"):
        full = full[5:]
    depth = 0
    for i, ch in enumerate(full):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return full[:i].strip()
    return full.strip()


def pmc_summary(dirs, variant: str, radius: int, n_tris: int = 2_000_002, n_rays: int = 4_194_304) -> str:
    """Markdown summary of rocprofv3 --pmc passes over `bench.py --minimal --calibrate-copy`
    (counter_collection.csv files under `dirs`, one pass each). HBM-side bytes per dispatch
    come from the L2-to-fabric request counters by request size: reads = 32 RDREQ_32B +
    64 RDREQ_64B + 128 RDREQ_128B, writes = 32 (WRREQ - WRREQ_64B) + 64 WRREQ_64B. The
    build is the first run of dispatches from KBuildInit to KFinish; the trace is the first
    KTrace dispatch after it."""
    import csv
    import glob

    disp = {}
    for d in dirs:
        for f in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
            for row in csv.DictReader(open(f)):
                k = int(row["Dispatch_Id"])
                e = disp.setdefault(k, {"kernel": _kname(row["Kernel_Name"]), "c": {}})
                if e["kernel"] != _kname(row["Kernel_Name"]):
                    raise ValueError(f"dispatch {k} differs between passes: {e['kernel']} / {row['Kernel_Name']}")
                e["c"][row["Counter_Name"]] = e["c"].get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    order = [disp[k] for k in sorted(disp)]

    def rd(c):
        return 32 * c.get("TCC_EA0_RDREQ_32B_sum", 0) + 64 * c.get("TCC_EA0_RDREQ_64B_sum", 0) + \
            128 * c.get("TCC_EA0_RDREQ_128B_sum", 0)

    def wr(c):
        w, w64 = c.get("TCC_EA0_WRREQ_sum", 0), c.get("TCC_EA0_WRREQ_64B_sum", 0)
        return 32 * (w - w64) + 64 * w64

    # FETCH_SIZE / WRITE_SIZE (KiB per dispatch) scaled by what they read for the plain copies
    cp = [e for e in order if "copyBuffer" in e["kernel"]]
    f_scale = w_scale = None
    if cp and cp[-1]["c"].get("FETCH_SIZE", 0) > 0:
        f_scale = 268435456 / (cp[-1]["c"]["FETCH_SIZE"] * 1024)
    if cp and cp[-1]["c"].get("WRITE_SIZE", 0) > 0:
        w_scale = 268435456 / (cp[-1]["c"]["WRITE_SIZE"] * 1024)

    def fw(c):
        f = c.get("FETCH_SIZE", 0) * 1024 * f_scale if f_scale else 0.0
        w = c.get("WRITE_SIZE", 0) * 1024 * w_scale if w_scale else 0.0
        return f, w

    start = next(i for i, e in enumerate(order) if e["kernel"] == "KBuildInit")
    end = next(i for i in range(start, len(order)) if order[i]["kernel"] == "KFinish")
    build = order[start:end + 1]
    per = {}
    for e in build:
        t = per.setdefault(e["kernel"], [0, 0.0, 0.0, 0.0, 0.0])
        t[0] += 1
        t[1] += rd(e["c"])
        t[2] += wr(e["c"])
        f, w = fw(e["c"])
        t[3] += f
        t[4] += w
    tot_r = sum(v[1] for v in per.values())
    tot_w = sum(v[2] for v in per.values())
    out = [f"Dispatches of the first {n_tris:,}-triangle build ({variant} r{radius}): {len(build)}.", "",
           "| kernel | dispatches | read MB | written MB | read+written per dispatch (MB) | FETCH_SIZE MB | "
           "WRITE_SIZE MB |", "|---|---|---|---|---|---|---|"]
    for k, (cnt, r, w, f, fwr) in sorted(per.items(), key=lambda kv: -(kv[1][1] + kv[1][2])):
        out.append(f"| `{k}` | {cnt} | {r / 1e6:.2f} | {w / 1e6:.2f} | {(r + w) / cnt / 1e6:.3f} | {f / 1e6:.2f} | "
                   f"{fwr / 1e6:.2f} |")
    out += ["", f"Whole build: {tot_r / 1e6:.1f} MB read + {tot_w / 1e6:.1f} MB written = "
                f"{(tot_r + tot_w) / 1e6:.1f} MB by the request counters"]
    if f_scale and w_scale:
        tf, tw = sum(v[3] for v in per.values()), sum(v[4] for v in per.values())
        out[-1] += (f"; {tf / 1e6:.1f} MB fetched + {tw / 1e6:.1f} MB written = {(tf + tw) / 1e6:.1f} MB by "
                    f"FETCH_SIZE x {f_scale:.4f} and WRITE_SIZE x {w_scale:.4f}")
    out[-1] += "."
    out.append("")
    # multi-workgroup PLOC iterations: one fused KPlocIteration launch each, or a KPlocNeighbours
    # + KPlocMerge pair; pass-through launches (clusters already at the tail size) move almost
    # nothing and are not iterations
    fused = [e for e in build if e["kernel"].startswith("KPlocIteration") and rd(e["c"]) + wr(e["c"]) > 256 * 1024]
    pairs = [e for e in build if e["kernel"].startswith(("KPlocNeighbours", "KPlocMerge"))]
    n_iter = len(fused) + sum(1 for e in pairs if e["kernel"].startswith("KPlocMerge"))
    if n_iter:
        pl = sum(rd(e["c"]) + wr(e["c"]) for e in fused + pairs) / n_iter
        out.append(f"Merging PLOC iterations: {n_iter} ({len(fused)} fused launches, {len(pairs) // 2} two-launch), "
                   f"{pl / 1e6:.2f} MB per iteration.")
        out.append("")
        out.append(f"traffic_bytes_per_launch: {variant} r{radius} = {int(round(pl))}")
    out.append(f"build_traffic_bytes: {variant} r{radius} = {int(round(tot_r + tot_w))}")
    tr = next((e for e in order[end:] if e["kernel"].startswith("KTrace")), None)
    if tr is not None:
        c = tr["c"]
        tag = f"{variant} r{radius}"
        out += ["", f"## `{tr['kernel']}` ({n_rays:,} rays over the build's tree)", "",
                f"read {rd(c) / 1e6:.1f} MB + written {wr(c) / 1e6:.1f} MB = "
                f"{(rd(c) + wr(c)) / n_rays:.1f} B per ray beyond L2", "",
                f"trace_bytes_per_ray: {tag} = {(rd(c) + wr(c)) / n_rays:.1f}"]
        if c.get("TCC_HIT_sum", 0) + c.get("TCC_MISS_sum", 0) > 0:
            h, m = c["TCC_HIT_sum"], c["TCC_MISS_sum"]
            out += [f"trace_l2_hit: {tag} = {h / (h + m):.4f}", "",
                    f"L2 hit ratio {h / (h + m):.4f} ({h:.3e} hits, {m:.3e} misses)"]
        if c.get("SQ_ACTIVE_INST_VALU", 0) > 0:
            lanes = c["SQ_THREAD_CYCLES_VALU"] / (64 * c["SQ_ACTIVE_INST_VALU"])
            # GRBM_GUI_ACTIVE is summed over the 8 XCDs (32 CUs each): per-XCD busy cycles of the
            # dispatch = GRBM_GUI_ACTIVE / 8
            cycles = c["GRBM_GUI_ACTIVE"] / 8
            issue = c["SQ_INSTS_VALU"] / (256 * cycles)
            out += [f"trace_valu_lane_util: {tag} = {lanes:.4f}", f"trace_valu_per_cu_cycle: {tag} = {issue:.4f}", "",
                    f"VALU lanes active per VALU instruction {lanes:.3f}; VALU wave-instructions per CU cycle "
                    f"{issue:.3f} ({c['SQ_INSTS_VALU']:.4e} over 256 CUs x {cycles:.4e} cycles); "
                    f"SQ_INSTS_VMEM_RD {c['SQ_INSTS_VMEM_RD']:.4e}, SQ_WAVES {c['SQ_WAVES']:.0f}, "
                    f"SQ_WAVE_CYCLES {c['SQ_WAVE_CYCLES']:.4e}, SQ_BUSY_CYCLES {c['SQ_BUSY_CYCLES']:.4e}, "
                    f"GRBM_GUI_ACTIVE {c['GRBM_GUI_ACTIVE']:.4e}"]
        for key in ("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum"):
            if key in c:
                out.append(f"{key} {c[key]:.4e}")
    if cp:
        c = cp[-1]["c"]
        out += ["", f"Copy check: a 268,435,456-byte device copy reads {rd(c) / 268435456:.4f}x and writes "
                    f"{wr(c) / 268435456:.4f}x its size by the request counters"]
        if f_scale and w_scale:
            out[-1] += (f"; FETCH_SIZE reads {c['FETCH_SIZE']:.1f} KiB and WRITE_SIZE {c['WRITE_SIZE']:.1f} KiB for "
                        f"its 262,144 KiB (scales {f_scale:.4f}, {w_scale:.4f})")
        out[-1] += "."
    return "\n".join(out)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variant", default="ploc_emc_sobb", choices=list(VARIANTS))
    ap.add_argument("--no-sweep", action="store_true", help="skip the variant x radius sweep")
    ap.add_argument("--radius", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 4096-mesh batched build and trace")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run three plain 256 MiB device copies (a known-size kernel for HBM counter scaling)")
    ap.add_argument("--minimal", action="store_true",
                    help="only the timed steps, the traversal counters and the roofline build (for profiler runs)")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR",
                    help="summarise rocprofv3 --pmc output directories of a --minimal run and exit (no GPU)")
    args = ap.parse_args()
    if args.minimal:
        args.no_batch = args.no_sweep = args.no_cpu = True
    if args.pmc_summary:
        print(pmc_summary(args.pmc_summary, args.variant, args.radius))
        return 0
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
    log(f"scene {scene.n} tris, {len(rays_np)} rays generated in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    tris = torch.from_numpy(scene.tris).to(dev)
    rays = torch.from_numpy(rays_np).to(dev)
    hits = torch.empty((len(rays_np), 4), dtype=torch.float32, device=dev)
    ctx = accel.Context(0)
    mt, bv = VARIANTS[args.variant]
    params = accel.make_params(morton_type=mt, radius=args.radius, max_leaf=4, bv_type=bv)
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
    mtris = n / (build_ms.mean() * 1e-3) / 1e6
    mrays = nrays / (trace_ms.mean() * 1e-3) / 1e6
    last = stats[-1]
    h = hits.cpu().numpy().view(np.int32)[:, 3]
    log(f"build {build_ms.mean():.3f} ms (device {np.mean([s['build_ms'] for s in stats]):.3f}), trace "
        f"{trace_ms.mean():.3f} ms, hit rate {(h >= 0).mean():.3f}, iterations {last['ploc_iterations']}")

    # ---- the same build from host memory (the boundary's DPL_MEM_HOST): the PCIe upload is
    #      reported beside the device time, never in `value` ----
    host_builds = []
    hb = accel.GpuBvh(ctx)
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hb.build(scene.tris, params)
        torch.cuda.synchronize()
        host_builds.append((time.perf_counter() - t0) * 1e3)
    hstats = hb.stats.as_dict()
    hb.close()
    from_host = {"upload_ms": round(hstats["upload_ms"], 4), "device_build_ms": round(hstats["build_ms"], 4),
                 "wall_ms": round(float(np.median(host_builds)), 4),
                 "note": "dpl_bvh_build from pageable host memory (72 MB of triangles over PCIe), median of 3 wall "
                         "times around the call; value uses device-resident input"}
    log(f"build from host memory: upload {from_host['upload_ms']:.3f} ms + device {from_host['device_build_ms']:.3f} ms, "
        f"wall {from_host['wall_ms']:.3f} ms")

    # ---- the same trace with DPL_TRACE_COHERENT (rays sorted by where they cross the scene
    #      box before tracing, sort included; hits must equal the step's) ----
    from diplodocus_amd import _lib

    step_hits = hits.clone()
    bvh.trace(rays, _lib.DPL_TRACE_COHERENT, hits)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    for _ in range(5):
        bvh.trace(rays, _lib.DPL_TRACE_COHERENT, hits)
    c1.record(stream)
    torch.cuda.synchronize()
    coherent_ms = c0.elapsed_time(c1) / 5
    coherent_equal = bool(torch.equal(hits.view(torch.int32), step_hits.view(torch.int32)))
    log(f"trace with DPL_TRACE_COHERENT: {coherent_ms:.3f} ms (sort included), hits equal: {coherent_equal}")

    # SAH with axis-aligned child boxes (a SOBB tree's boxes refitted): the same cost formula
    # as the CPU SAH BVH's, for the ratio that gates tree quality
    sah_aabb = bvh.sah_aabb()

    # ---- traversal counters of the same trace (outside the timed region) ----
    _, visits, tests = bvh.trace_counters(rays)
    touched = (visits * 64 + tests * 48) / nrays

    wide = {}
    if not args.minimal:
        # ---- wide trees of the same BVH (4 and 8 children, 8-bit quantized child boxes): the
        #      conversion (SOBB trees refitted to axis-aligned boxes first) and the closest-hit
        #      trace of the same rays, whose hits must equal the two-child trace's ----

        two_child_hits = step_hits
        for width in (4, 8):
            bvh.widen(width)
            widen_ms = float(np.median([bvh.widen(width)["widen_ms"] for _ in range(3)]))
            bvh.trace(rays, _lib.DPL_TRACE_WIDE, hits)
            w0, w1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            w0.record(stream)
            for _ in range(5):
                bvh.trace(rays, _lib.DPL_TRACE_WIDE, hits)
            w1.record(stream)
            torch.cuda.synchronize()
            wide_ms = w0.elapsed_time(w1) / 5
            equal = bool(torch.equal(hits.view(torch.int32), two_child_hits.view(torch.int32)))
            _, wv, wt = bvh.trace_counters(rays, _lib.DPL_TRACE_WIDE)
            wide[f"width{width}"] = {"node_bytes": bvh.wide_info()["node_bytes"], "widen_ms": round(widen_ms, 4),
                                     "trace_ms": round(wide_ms, 4), "trace_mrays_s": round(nrays / (wide_ms * 1e-3) / 1e6, 3),
                                     "nodes_per_ray": round(wv / nrays, 2), "tris_per_ray": round(wt / nrays, 2),
                                     "hits_equal_two_child": equal}
            log(f"wide{width}: widen {widen_ms:.3f} ms, trace {wide_ms:.3f} ms = {nrays / wide_ms / 1e3:.1f} Mrays/s, "
                f"{wv / nrays:.1f} nodes and {wt / nrays:.1f} tris per ray, hits equal: {equal}")

    # ---- roofline of the PLOC iteration kernel: one profiled build (per-launch events) ----
    prof = accel.GpuBvh(ctx).build(tris, accel.make_params(morton_type=mt, radius=args.radius, max_leaf=4,
                                                             bv_type=bv, keep_intermediates=3))
    ps = prof.stats
    stage_ms = {k: round(v, 4) for k, v in ps.as_dict()["stage_ms"].items()}
    log(f"stages of a build with per-stage events: {stage_ms}")
    pd = prof.download(intermediates=True)
    alg = build_algorithmic_bytes(n, mt, 8, pd["iter_counts"], ps.inner_node_count)
    build_gbs = alg["total"] / (build_ms.mean() * 1e-3) / 1e9
    build_roof = {"algorithmic_bytes": alg, "build_ms": round(float(build_ms.mean()), 4),
                  "achieved_gbs": round(build_gbs, 2), "frac": round(build_gbs / HBM_PEAK_GBS, 5),
                  "measured": measured_build_traffic(args.variant, args.radius, float(build_ms.mean()))}
    log(f"whole build: {alg['total'] / 1e6:.1f} MB algorithmic in {build_ms.mean():.3f} ms = {build_gbs:.1f} GB/s")
    roof = None
    if ps.ploc_launches > 0 and ps.ploc_launch_ms > 0:
        per_launch_bytes = ps.ploc_bytes / ps.ploc_launches
        per_launch_ms = ps.ploc_launch_ms / ps.ploc_launches
        achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
        roof = {"kernel": "one PLOC iteration (NN search, merge, compaction): KPlocNeighbours<256> + "
                          "KPlocMerge<256> from ~98k clusters up, KPlocIteration<256, 4> below",
                "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(args.variant, args.radius),
                "iterations": ps.ploc_launches, "avg_iteration_us": round(per_launch_ms * 1e3, 3),
                "algorithmic_bytes_per_iteration": round(per_launch_bytes)}
        log(f"ploc iterations: {ps.ploc_launches} multi-workgroup, {per_launch_ms * 1e3:.2f} us avg, "
            f"{achieved:.1f} GB/s algorithmic")
    prof.close()

    prim_mrays = bounce_ms = None
    diff = None
    whitted = {}
    cs = cmats = None
    if not args.minimal:
        # ---- primary rays: 1024x1024 camera of the icosphere room (PLOC, Morton30, r=16) ----
        room = scenes.icosphere_room()
        cam_rays = torch.from_numpy(scenes.camera_rays(room.camera, 1024, 1024)).to(dev)
        cam_hits = torch.empty((cam_rays.shape[0], 4), dtype=torch.float32, device=dev)
        rb = accel.GpuBvh(ctx).build(torch.from_numpy(room.tris).to(dev), accel.make_params(0, 16, 4))
        for _ in range(3):
            rb.trace(cam_rays, 0, cam_hits)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        e0.record(stream)
        for _ in range(reps):
            rb.trace(cam_rays, 0, cam_hits)
        e1.record(stream)
        torch.cuda.synchronize()
        prim_ms = e0.elapsed_time(e1) / reps
        prim_mrays = cam_rays.shape[0] / (prim_ms * 1e-3) / 1e6
        _, pv, pt = rb.trace_counters(cam_rays)
        log(f"primary rays: {prim_ms:.3f} ms per 1M rays = {prim_mrays:.1f} Mrays/s, {pv / cam_rays.shape[0]:.1f} "
            f"nodes and {pt / cam_rays.shape[0]:.1f} tris per ray")
        rb.widen(4)
        for _ in range(3):
            rb.trace(cam_rays, _lib.DPL_TRACE_WIDE, cam_hits)
        e0.record(stream)
        for _ in range(reps):
            rb.trace(cam_rays, _lib.DPL_TRACE_WIDE, cam_hits)
        e1.record(stream)
        torch.cuda.synchronize()
        prim_wide_mrays = cam_rays.shape[0] / (e0.elapsed_time(e1) / reps * 1e-3) / 1e6
        wide["primary_width4_mrays_s"] = round(prim_wide_mrays, 3)
        log(f"primary rays, wide4: {prim_wide_mrays:.1f} Mrays/s")

        # ---- path tracing on the same room: 1024x1024 primary + 1 diffuse bounce, and the
        #      differentiable case (512x512, 4 spp, depth 4): forward and adjoint path replay ----
        from diplodocus_amd import render

        mat_id = torch.from_numpy(room.mat_id).to(dev)
        albedo = torch.from_numpy(room.albedo).to(dev)
        emission = torch.from_numpy(room.emission).to(dev)

        def timed(fn, reps):
            fn()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(reps):
                fn()
            b.record(stream)
            torch.cuda.synchronize()
            return a.elapsed_time(b) / reps

        rp_b = render.make_render_params(1024, 1024, 1, 2, 1, room.camera, len(room.albedo))
        img_b = torch.empty((1024, 1024, 3), dtype=torch.float32, device=dev)
        bounce_ms = timed(lambda: render.render_forward(rb, rp_b, mat_id, albedo, emission, img_b), 10)
        rp = render.make_render_params(512, 512, 4, 4, 1, room.camera, len(room.albedo))
        gen = torch.Generator().manual_seed(5)
        target_albedo = (albedo.cpu() + 0.05 * torch.randn(albedo.shape, generator=gen)).clamp(0, 1).to(dev)
        target = render.render_forward(rb, rp, mat_id, target_albedo, emission)
        img = torch.empty((512, 512, 3), dtype=torch.float32, device=dev)
        fwd_ms = timed(lambda: render.render_forward(rb, rp, mat_id, albedo, emission, img), 10)
        d_img = (2.0 * (img - target)).contiguous()
        adj_ms = timed(lambda: render.render_adjoint(rb, rp, mat_id, albedo, emission, d_img), 10)
        diff = {"workload": "icosphere_room 69,668 tris, 16 materials, 512x512 x 4 spp, depth 4, L2 loss vs albedo "
                            "+ N(0, 0.05) target; gradients for 16x3 albedo and emission",
                "forward_ms": round(fwd_ms, 4), "adjoint_ms": round(adj_ms, 4),
                "fwd_plus_adjoint_ms": round(fwd_ms + adj_ms, 4),
                "adjoint_over_forward": round(adj_ms / fwd_ms, 3)}
        log(f"path tracing: 1024x1024 + 1 bounce {bounce_ms:.3f} ms; 512x512x4 depth 4 forward {fwd_ms:.3f} ms, "
            f"adjoint {adj_ms:.3f} ms ({adj_ms / fwd_ms:.2f}x)")
        rb.close()

        # ---- Whitted ray tracing of the 36-triangle Cornell room (256x256, depth 3, 8 light
        #      samples per hit), the reference's CPU renderer's own case ----
        cs, cmats = scenes.cornell_whitted()
        cb = accel.GpuBvh(ctx).build(torch.from_numpy(cs.tris).to(dev), accel.make_params(0, 16, 4))
        c_tris, c_mat = torch.from_numpy(cs.tris).to(dev), torch.from_numpy(cs.mat_id).to(dev)
        c_mats = torch.from_numpy(cmats).to(dev)
        c_lights = torch.from_numpy(render.emissive_triangles(cs.mat_id, cmats)).to(dev)
        whitted = {}
        for side in (256, 1024):
            wp = render.make_whitted_params(side, side, 3, 8, cs.camera)
            wimg = torch.empty((side, side, 3), dtype=torch.float32, device=dev)
            ms = timed(lambda: render.render_whitted(cb, wp, c_tris, c_mat, c_mats, c_lights, image=wimg), 10)
            whitted[f"{side}x{side}_ms"] = round(ms, 4)
        cb.close()
        log(f"whitted cornell: 256x256 {whitted['256x256_ms']:.3f} ms, 1024x1024 {whitted['1024x1024_ms']:.3f} ms")

    # ---- batched builds: 4096 meshes (log-uniform 64..65536 tris) in one persistent
    #      PLOC+EMC launch under autotuned radius / block / digit width, then a
    #      hipGraph-captured trace of 256 rays per mesh ----
    batch = None
    if not args.no_batch:
        t0 = time.time()
        btris_np, boff = scenes.batched_meshes()
        brays_np, bids_np = scenes.batched_rays(btris_np, boff, 256)
        log(f"batch: {len(boff) - 1} meshes, {boff[-1]} tris generated in {time.time() - t0:.1f}s")
        btris = torch.from_numpy(btris_np).to(dev)
        del btris_np
        bparams, tune = accel.autotune(ctx, btris, accel.make_params(morton_type=1, radius=16, max_leaf=4),
                                       offsets=boff, sah_tolerance=0.05)
        gb = accel.GpuBvh(ctx)
        bms = []
        for _ in range(4):
            gb.build_batched(btris, boff, bparams)
            bms.append(gb.stats.build_ms)
        bms = bms[1:]
        brays = torch.from_numpy(brays_np).to(dev)
        bids = torch.from_numpy(bids_np).to(dev)
        bhits = torch.empty((len(bids_np), 4), dtype=torch.float32, device=dev)
        graph = accel.TraceGraph(gb, brays, bhits, ray_mesh=bids)
        graph_ms = timed(graph.launch, 10)
        eager_ms = timed(lambda: gb.trace_batched(brays, bids, 0, bhits), 10)
        # the same rays over the batch's wide form (4 children), also graph-captured
        two_child = bhits.clone()
        winfo = gb.widen(4)
        wgraph = accel.TraceGraph(gb, brays, bhits, ray_mesh=bids, flags=_lib.DPL_TRACE_WIDE)
        wgraph_ms = timed(wgraph.launch, 10)
        wide_equal = bool(torch.equal(bhits.view(torch.int32), two_child.view(torch.int32)))
        wgraph.close()
        nb = int(boff[-1])
        batch = {"workload": f"{len(boff) - 1} meshes, log-uniform 64..65536 tris (noise spheres / soups), {nb:,} tris; "
                             f"256 rays per mesh", "meshes": len(boff) - 1, "triangles": nb,
                 "autotuned": {"nn_search_radius": bparams.nn_search_radius, "block_size": bparams.block_size,
                               "sort_digit_bits": bparams.sort_digit_bits, "candidates": tune["candidates"]},
                 "build_ms": round(float(np.median(bms)), 4),
                 "build_mtris_s": round(nb / (float(np.median(bms)) * 1e-3) / 1e6, 2),
                 "sah_cost_mean": gb.stats.sah_cost,
                 "trace_rays": len(bids_np), "trace_graph_ms": round(graph_ms, 4),
                 "trace_eager_ms": round(eager_ms, 4),
                 "trace_mrays_s": round(len(bids_np) / (graph_ms * 1e-3) / 1e6, 2),
                 "widen4_ms": round(winfo["widen_ms"], 4), "trace_wide4_graph_ms": round(wgraph_ms, 4),
                 "trace_wide4_mrays_s": round(len(bids_np) / (wgraph_ms * 1e-3) / 1e6, 2),
                 "wide_hits_equal_two_child": wide_equal}
        log(f"batch: build {batch['build_ms']:.2f} ms ({batch['build_mtris_s']:.0f} Mtris/s, r="
            f"{bparams.nn_search_radius} block={bparams.block_size} digit={bparams.sort_digit_bits}), trace graph "
            f"{graph_ms:.3f} ms / eager {eager_ms:.3f} ms for {len(bids_np)} rays; wide4 graph {wgraph_ms:.3f} ms "
            f"(hits equal: {wide_equal})")
        graph.close()
        gb.close()
        del btris, brays, bids, bhits

    # ---- variant x radius sweep of the 2M scene (build time, SAH, trace rate) ----
    sweep = []
    if not args.no_sweep:
        sb_ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        g = accel.GpuBvh(ctx)
        for name, (vm, vb) in VARIANTS.items():
            for r in (8, 16, 32):
                p = accel.make_params(morton_type=vm, radius=r, max_leaf=4, bv_type=vb)
                g.build(tris, p)
                g.trace(rays, 0, hits)
                sb_ev[0].record(stream)
                g.build(tris, p)
                sb_ev[1].record(stream)
                g.trace(rays, 0, hits)
                sb_ev[2].record(stream)
                torch.cuda.synchronize()
                sweep.append({"variant": name, "radius": r, "build_ms": round(sb_ev[0].elapsed_time(sb_ev[1]), 4),
                              "trace_ms": round(sb_ev[1].elapsed_time(sb_ev[2]), 4),
                              "sah_cost": g.stats.sah_cost, "sah_cost_aabb": g.sah_aabb(),
                              "iterations": g.stats.ploc_iterations})
        g.close()
        log("sweep: " + ", ".join(f"{s['variant']}/r{s['radius']}: {s['build_ms']:.2f} ms sah {s['sah_cost']:.1f} "
                                  f"trace {s['trace_ms']:.1f} ms" for s in sweep))

    # ---- CPU baseline: the reference's SAH BVH build (1 core) and tracer on a sample ----
    cpu = None
    sah_ratio = None
    sah_ratio_fixed = None
    sah_ratio_own = None
    if not args.no_cpu:
        from oracle import pyoracle

        sb = pyoracle.SBvh(scene.tris, fix_child_bounds=False)
        sah_cpu = sb.sah()
        for s in sweep:
            s["sah_ratio_vs_cpu"] = round(s["sah_cost_aabb"] / sah_cpu, 4)
            s["sah_ratio_vs_cpu_own_bounds"] = round(s["sah_cost"] / sah_cpu, 4)
        sah_fixed = pyoracle.SBvh(scene.tris, fix_child_bounds=True).sah()
        sah_ratio = sah_aabb / sah_cpu
        sah_ratio_fixed = sah_aabb / sah_fixed
        sah_ratio_own = last["sah_cost"] / sah_cpu
        sample = rays_np[:131072]
        threads = pyoracle.nthreads()
        sb.trace(sample, flags=0, mode=0, threads=threads)
        cpu_trace_s = sb.last_trace_s
        sample1 = rays_np[:16384]
        sb.trace(sample1, flags=0, mode=0, threads=1)
        cpu_trace1_s = sb.last_trace_s
        t0 = time.perf_counter()
        pyoracle.whitted(pyoracle.Ploc(cs.tris, 0, 16, 4), cs.mat_id, cmats, 256, 256, 3, 8, cs.camera,
                         threads=threads)
        whitted_cpu_ms = (time.perf_counter() - t0) * 1e3
        whitted["cpu_256x256_ms"] = round(whitted_cpu_ms, 2)
        whitted["cpu_threads"] = threads
        cpu = {
            "value": round(n / (sb.build_ms * 1e-3) / 1e6, 4),
            "unit": "Mtris/s",
            "cores": 1,
            "kind": "port",
            "sample": (f"full 2,000,002-tri SAH build on 1 thread ({sb.build_ms:.0f} ms, restated "
                       f"src/acceleration/sbvh.cc); trace of the first {len(sample)} rays with the reference "
                       f"rules on {threads} threads: {len(sample) / cpu_trace_s / 1e6:.4f} Mrays/s"),
            "trace_mrays_s": round(len(sample) / cpu_trace_s / 1e6, 5),
            "trace_threads": threads,
            "trace_1core_mrays_s": round(len(sample1) / cpu_trace1_s / 1e6, 5),
            "trace_1core_sample": f"the first {len(sample1)} rays on 1 thread",
            "cpu_model": cpu_model(),
            "host_cores": os.cpu_count(),
            "sah_faithful": sah_cpu,
            "sah_fixed": sah_fixed,
        }
        log(f"cpu build {sb.build_ms:.0f} ms, sah cpu {sah_cpu:.3f} (fixed {sah_fixed:.3f}) gpu {sah_aabb:.3f} in "
            f"axis-aligned areas ({last['sah_cost']:.3f} in its own bounds); cpu trace {threads} threads "
            f"{len(sample) / cpu_trace_s / 1e6:.4f} Mrays/s, 1 thread {len(sample1) / cpu_trace1_s / 1e6:.4f} Mrays/s")

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
        "config": {"workload": "mixed_2m: 2,000,002 tris (64 Gaussian clusters, 200k thin 50:1, 2 ground), "
                               "4,194,304 rays from a bounding sphere toward the centre",
                   "variant": args.variant, "radius": args.radius, "max_leaf": 4},
        "trace_mrays_s": round(mrays, 3),
        "trace_wide4_mrays_s": wide["width4"]["trace_mrays_s"] if wide else None,
        "primary_mrays_s": round(prim_mrays, 3) if prim_mrays else None,
        "primary_workload": "icosphere_room 69,668 tris, 1024x1024 camera rays, PLOC Morton30 r=16",
        "primary_bounce_ms": round(bounce_ms, 4) if bounce_ms else None,
        "differentiable": diff,
        "whitted_cornell": dict(whitted, workload="cornell room 36 tris, Whitted depth 3, 8 area-light samples "
                                                  "per hit (the reference CPU renderer's case)"),
        "batched": batch,
        "build_ms": round(float(build_ms.mean()), 4),
        "trace_ms": round(float(trace_ms.mean()), 4),
        "sah_cost": last["sah_cost"],
        "sah_cost_aabb": sah_aabb,
        "sah_ratio_vs_cpu": sah_ratio,
        "sah_ratio_vs_cpu_fixed": sah_ratio_fixed,
        "sah_ratio_vs_cpu_own_bounds": sah_ratio_own,
        "sah_note": "sah_ratio_vs_cpu: both trees costed with axis-aligned child boxes (the SOBB tree's boxes "
                    "refitted); _own_bounds: the SOBB tree costed with its SOBB child areas",
        "ploc_iterations": last["ploc_iterations"],
        "stage_ms": stage_ms,
        "stage_ms_note": "one build with per-stage and per-iteration events (keep_intermediates bit 1); the timed "
                         "builds record no events between stages (each adds ~5 us of idle GPU time)",
        "trace_nodes_per_ray": round(visits / nrays, 2),
        "trace_tris_per_ray": round(tests / nrays, 2),
        "build_from_host": from_host,
        "trace_touched_bytes_per_ray": round(touched, 1),
        "trace_coherent": {"flag": "DPL_TRACE_COHERENT", "trace_ms": round(coherent_ms, 4),
                           "trace_mrays_s": round(nrays / (coherent_ms * 1e-3) / 1e6, 3),
                           "hits_equal_step": coherent_equal,
                           "note": "the same rays sorted by where they cross the scene box before tracing, "
                                   "sort included; the step's trace_ms keeps the given order"},
        "trace_roofline": trace_roofline(args.variant, args.radius, nrays, float(trace_ms.mean()), touched),
        "wide": dict(wide, workload="the step's BVH and rays; wide trees converted from it (SOBB trees refitted "
                                    "to axis-aligned boxes), 4 or 8 children, 8-bit quantized child boxes"),
        "wall_s": round(wall, 3),
        "variant_sweep": sweep,
        "roofline": roof,
        "build_roofline": build_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 268,435,456-byte device-to-device copies")
    bvh.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
