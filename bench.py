#!/usr/bin/env python3
"""Throughput benchmark of the VVC in-loop filter chain on one MI355X.

Workload (the headline line): 8 pictures of 3840x2160 4:2:0 10-bit, full chain
(LMCS inverse, deblocking incl. long luma filters, SAO, ALF with 4 luma APS
sets, chroma ALF and CC-ALF), synthetic seeded inputs, filtered in the plain
per-picture form (one kernel per stage).  One step = every picture filtered
once with its input and side information already resident in HBM.

The same JSON line also carries:
- "batched": the same 8 pictures in one persistent fused launch replayed from
  a HIP graph (tile shape from the autotuner), and the 64-picture mixed batch
  (16 each of 416x240, 832x480, 1920x1080, 3840x2160);
- "served": continuous batching of Poisson arrivals (120 pictures/s over 32
  streams of 832x480 / 1920x1080 / 3840x2160) with p50 / p99 latency;
- "roofline" of the dominant plain kernel and of the fused kernel, from HIP
  events recorded on the launch stream, with the HBM traffic and the vector
  ALU instruction count per dispatch from the rocprofv3 summaries in
  profiles/ ("valu": issue rate against 256 CUs x 2.4 GHz x one wave64
  instruction per CU-cycle, the bound the integer filters actually meet);
- "cpu_baseline": the CPU oracle timed on host cores.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
STAGES = ["prep_lmcs", "deblock_v", "deblock_h", "sao", "alf_luma", "alf_chroma_ccalf"]
# Algorithmic bytes per luma pixel of each plain stage kernel: the planes it
# transforms, read once and written once (u16 samples, 4:2:0).  The fused
# kernel does the whole chain: 6 bytes per luma pixel.
STAGE_BYTES_PER_PX = {"prep_lmcs": 6.0, "deblock_v": 6.0, "deblock_h": 6.0, "sao": 6.0, "alf_luma": 4.0,
                      "alf_chroma_ccalf": 4.0}
CHAIN_BYTES_PER_PX = 6.0
KERNEL_NAMES = {"prep_lmcs": "k_prep", "deblock_v": "k_dbk<true>", "deblock_h": "k_dbk<false>", "sao": "k_sao",
                "alf_luma": "k_alf_luma", "alf_chroma_ccalf": "k_alf_chroma"}


def load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(vv, n_pics, w, h, seconds):
    """The CPU oracle (C restatement of the reference filters) on host cores,
    one picture per thread, repeated until ~`seconds` of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import binding
    orc = binding.load(vv)
    frames = [vv.SynthFrame(1, w, h, seed=1000 + i) for i in range(n_pics)]
    threads = max(1, min(n_pics, os.cpu_count() or 1))
    arr = (C.POINTER(vv.Frame) * n_pics)(*[f.desc for f in frames])
    reps, wall, cpu = 0, 0.0, 0.0
    while cpu < seconds:
        for f in frames:
            f.reset()
        t0, c0 = time.perf_counter(), time.process_time()
        rc = orc.oracle_apply_in_loop_filters_mt(arr, n_pics, threads)
        wall += time.perf_counter() - t0
        cpu += time.process_time() - c0
        reps += 1
        assert rc == 0
    mpx = reps * n_pics * w * h / 1e6
    return {"value": round(mpx / wall, 3), "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "sample": f"{reps} x {n_pics} pictures {w}x{h} (the benchmark workload), oracle C restatement, "
                      f"one picture per thread, {cpu:.1f} s CPU over {wall:.1f} s wall"}


def plain_run(vv, ctx, jobs, steps, warmup, W, H):
    """Per-picture kernel sequences (each job's captured graph replayed in
    turn), timed over `steps` passes; then a shorter pass with HIP events
    between the stage kernels for the per-stage split."""
    sh = ctx.stream  # every kernel goes to this HIP stream
    for _ in range(warmup):
        for j in jobs:
            j.graph_launch(sh)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for j in jobs:
            j.graph_launch(sh)
    ctx.synchronize()
    total_ms = (time.perf_counter() - t0) * 1e3
    psteps = max(1, min(steps, 3))
    prof = C.c_void_p()
    vv.check(vv.lib.vvclf_prof_create(ctx.h, psteps * len(jobs), C.byref(prof)))
    ctx.synchronize()
    for _ in range(psteps):
        for j in jobs:
            vv.check(vv.lib.vvclf_job_run_prof(j.h, sh, prof))
    ctx.synchronize()
    ms = (vv.f32 * 8)()
    runs = vv.u32()
    vv.check(vv.lib.vvclf_prof_stage_ms(prof, ms, C.byref(runs)))
    vv.lib.vvclf_prof_destroy(prof)
    stage_ms = {s: ms[i] / runs.value for i, s in enumerate(STAGES)}
    return total_ms / steps, stage_ms


def batched_run(vv, ctx, jobs, reps, warmup=2):
    """Autotune, build the batch, time graph replays; returns dict."""
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    tune = vv.Tune()
    vv.check(vv.lib.vvclf_autotune(ctx.h, arr, len(jobs), C.byref(tune)))
    b = C.c_void_p()
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), C.byref(tune), C.byref(b)))
    ms = C.c_double()
    vv.check(vv.lib.vvclf_batch_time(b, ctx.stream, warmup, reps, 1, C.byref(ms)))
    vv.lib.vvclf_batch_destroy(b)
    px = sum(j.luma_pixels for j in jobs)
    return {"ms_per_launch": round(ms.value, 4), "luma_pixels_per_launch": px,
            "value": round(px / (ms.value / 1e3) / 1e6, 2), "unit": "Mpixel/s",
            "tile": f"{tune.tile_w}x{tune.tile_h}", "wgs_per_cu": tune.wgs_per_cu,
            "hbm_algorithmic_GBs": round(CHAIN_BYTES_PER_PX * px / (ms.value / 1e3) / 1e9, 2)}


SERVE_STAGES = ["admission", "host_tables_and_upload", "wait_for_launch", "device", "download"]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_instr.json")
# Issue ceiling of vector ALU work: 256 CUs x 2.4 GHz x one wave64 VALU
# instruction per CU per cycle (four SIMD16 units, four cycles per wave64 op).
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9


def profile_entry(path, kernel):
    """The entry of a per-kernel rocprofv3 summary in profiles/ whose name
    starts with `kernel` (template arguments may follow), or None."""
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    for k, v in d.items():
        name = k.replace("void ", "")
        if name.startswith(kernel) and (len(name) == len(kernel) or name[len(kernel)] in "<,"):
            return v
    return None


def traffic_for(kernel):
    """HBM bytes per dispatch of `kernel` on this workload (FETCH_SIZE + WRITE_SIZE, calibrated), from the
    rocprofv3 summary in profiles/, or None when not profiled."""
    v = profile_entry(TRAFFIC_FILE, kernel)
    return int(round(v["read_bytes"] + v["write_bytes"])) if v else None


def valu_for(kernel, avg_launch_ms):
    """The kernel's vector-ALU issue rate against the VALU ceiling: SQ_INSTS_VALU
    per dispatch (rocprofv3 summary in profiles/) over the live launch time."""
    v = profile_entry(VALU_FILE, kernel)
    if not v or not avg_launch_ms:
        return None
    rate = v["valu_wave_instr"] / (avg_launch_ms / 1e3)
    return {"wave_instr_per_launch": int(v["valu_wave_instr"]), "achieved_G_per_s": round(rate / 1e9, 1),
            "peak_G_per_s": round(VALU_PEAK_WAVE_INSTR_PER_S / 1e9, 1),
            "frac": round(rate / VALU_PEAK_WAVE_INSTR_PER_S, 4)}


def served_run(vv, ctx, seconds, rate, n_streams=32, seed=7):
    """Poisson arrivals at `rate` pictures/s over `n_streams` streams."""
    sizes = [(832, 480), (1920, 1080), (3840, 2160)]
    rnd = random.Random(seed)
    cfg = vv.ServerConfig()
    cfg.hbm_budget_bytes = 64 << 30
    cfg.max_inflight = 64
    cfg.max_batch = 16
    cfg.prep_threads = 6
    s = C.c_void_p()
    vv.check(vv.lib.vvclf_server_create(ctx.h, C.byref(cfg), C.byref(s)))
    streams = []
    for k in range(n_streams):
        w, h = sizes[rnd.randrange(3)]
        sid = C.c_uint32()
        vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(sid)))
        # two pictures per stream sharing the stream's APS / LMCS state
        frames = [vv.SynthFrame(4, w, h, seed=5000 + 10 * k + i, aps_seed=900 + k) for i in range(2)]
        streams.append((sid.value, frames))
    n_jobs = int(seconds * rate)
    tickets, busy = [], {}
    t_next = time.perf_counter()
    for i in range(n_jobs):
        t_next += rnd.expovariate(rate)
        while time.perf_counter() < t_next:
            time.sleep(max(0.0, min(0.001, t_next - time.perf_counter())))
        sid, frames = streams[rnd.randrange(n_streams)]
        f = frames[i % 2]
        prev = busy.get(id(f))
        if prev is not None:  # the frame's previous job must be done before reuse
            vv.lib.vvclf_server_wait(s, prev, 60000)
        f.reset()
        t = C.c_uint64()
        vv.check(vv.lib.vvclf_server_submit(s, sid, f.desc, C.byref(t)))
        busy[id(f)] = t.value
        tickets.append(t.value)
    vv.check(vv.lib.vvclf_server_drain(s, 120000))
    st = vv.ServerStats()
    vv.check(vv.lib.vvclf_server_get_stats(s, C.byref(st)))
    vv.lib.vvclf_server_destroy(s)
    return {"streams": n_streams, "offered_pictures_per_s": rate, "pictures": int(st.completed),
            "failed": int(st.failed), "launches": int(st.launches),
            "throughput_Mpixel_s": round(st.luma_pixels / st.wall_s / 1e6, 2) if st.wall_s else None,
            "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3), "max_ms": round(st.max_ms, 3),
            "admitted_bytes_peak": int(st.bytes_admitted_peak), "hbm_budget_bytes": int(cfg.hbm_budget_bytes),
            "stage_p50_ms": {k: round(st.stage_p50_ms[i], 3) for i, k in enumerate(SERVE_STAGES)},
            "stage_p99_ms": {k: round(st.stage_p99_ms[i], 3) for i, k in enumerate(SERVE_STAGES)},
            "latency": "submit -> filtered samples back in host memory (includes host table derivation, "
                       "H2D and D2H copies)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pictures", type=int, default=8)
    ap.add_argument("--width", type=int, default=3840)
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-mixed64", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-rate", type=float, default=120.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (pictures shard by stream "
              f"across GPUs with no collectives, so the per-GPU figure is the one reported)", file=sys.stderr)
        sys.exit(2)

    vv = load("vvclf_amd", os.path.join(ROOT, "oxideav-h266_amd", "__init__.py"))
    W, H, N = args.width, args.height, args.pictures
    log("generating workload")
    frames = [vv.SynthFrame(1, W, H, seed=1000 + i) for i in range(N)]
    ctx = vv.Context(0)
    jobs = [vv.Job(ctx, f.desc) for f in frames]
    log("plain form")
    ms_per_step, stage_ms = plain_run(vv, ctx, jobs, args.steps, args.warmup, W, H)
    px = N * W * H
    value = px / (ms_per_step / 1e3) / 1e6
    dom = max(stage_ms, key=stage_ms.get)
    dom_bytes = STAGE_BYTES_PER_PX[dom] * W * H
    achieved = dom_bytes / (stage_ms[dom] / 1e3) / 1e9
    out = {
        "metric": "filtered Mpixel/s 4:2:0 10-bit (batched/served vs plain); p99 latency; HBM GB/s",
        "value": round(value, 2),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16 samples, int32 arithmetic",
        "data": "synthetic (seeded gradient + noise samples, random CU trees / SAO / ALF / LMCS side info)",
        "config": {"workload": f"{N}x {W}x{H} 4:2:0 10-bit, full chain (LMCS inverse, deblock incl. long filters, "
                               f"SAO, ALF 4 luma APS + chroma ALF + CC-ALF); plain per-picture stage kernels",
                   "pictures_per_step": N, "luma_pixels_per_step": px, "form": "plain"},
        "chain_hbm_algorithmic_GBs": round(CHAIN_BYTES_PER_PX * px / (ms_per_step / 1e3) / 1e9, 2),
        "stage_ms_per_picture": {k: round(v, 5) for k, v in stage_ms.items()},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_for(KERNEL_NAMES[dom]),
                     "algorithmic_bytes_per_launch": dom_bytes,
                     "avg_launch_ms": round(stage_ms[dom], 5),
                     "valu": valu_for(KERNEL_NAMES[dom], stage_ms[dom])},
    }
    if not args.no_batched:
        log("batched form, same pictures")
        bt = batched_run(vv, ctx, jobs, reps=max(5, args.steps // 2))
        bt["speedup_vs_plain"] = round(bt["value"] / value, 2)
        bt["roofline"] = {"bound": "hbm", "kernel": "k_fused", "achieved": bt["hbm_algorithmic_GBs"],
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(bt["hbm_algorithmic_GBs"] / HBM_PEAK_GBS, 4),
                          "traffic": traffic_for("k_fused<%d, %d" % tuple(int(v) for v in bt["tile"].split("x"))),
                          "algorithmic_bytes_per_launch": CHAIN_BYTES_PER_PX * px,
                          "avg_launch_ms": bt["ms_per_launch"],
                          "valu": valu_for("k_fused<%d, %d" % tuple(int(v) for v in bt["tile"].split("x")),
                                           bt["ms_per_launch"])}
        out["batched"] = {"same_8x4k": bt}
    for j in jobs:
        j.close()
    jobs = []
    if not args.no_dropin:
        # the drop-in call end to end: host planes in, host planes out (PCIe
        # and host table derivation included), one picture at a time
        log("drop-in host-to-host")
        ts = []
        for rep in range(2):
            for f in frames:
                f.reset()
                t0 = time.perf_counter()
                vv.check(vv.lib.vvclf_apply_in_loop_filters(f.desc))
                ts.append(time.perf_counter() - t0)
        ts = ts[len(frames):]  # first pass warms the pools and staging buffers
        out["dropin_host_to_host"] = {"value": round(N * W * H / sum(ts) / 1e6, 2), "unit": "Mpixel/s",
                                      "ms_per_picture": round(1e3 * sum(ts) / len(ts), 3),
                                      "note": "vvclf_apply_in_loop_filters per picture: host table derivation, "
                                              "upload, plain kernels, download"}
    if not args.no_batched and not args.no_mixed64:
        log("batched form, 64 mixed pictures")
        sizes = [(416, 240)] * 16 + [(832, 480)] * 16 + [(1920, 1080)] * 16 + [(3840, 2160)] * 16
        mframes = [vv.SynthFrame(2, w, h, seed=2000 + i) for i, (w, h) in enumerate(sizes)]
        mjobs = [vv.Job(ctx, f.desc) for f in mframes]
        bm = batched_run(vv, ctx, mjobs, reps=5)
        # plain per-picture reference point on the same 64 pictures
        ms_plain, _ = plain_run(vv, ctx, mjobs, 3, 1, 0, 0)
        bm["plain_value"] = round(bm["luma_pixels_per_launch"] / (ms_plain / 1e3) / 1e6, 2)
        bm["speedup_vs_plain"] = round(bm["value"] / bm["plain_value"], 2)
        out["batched"]["mixed_64"] = bm
        for j in mjobs:
            j.close()
    if args.serve_seconds > 0:
        log("serving")
        out["served"] = served_run(vv, ctx, args.serve_seconds, args.serve_rate)
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(vv, N, W, H, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
