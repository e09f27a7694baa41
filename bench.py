#!/usr/bin/env python3
"""Throughput benchmark of the VVC in-loop filter chain on one MI355X.

Workload (the headline line): 8 pictures of 3840x2160 4:2:0 10-bit, full chain
(LMCS inverse, deblocking incl. long luma filters, SAO, ALF with 4 luma APS
sets, chroma ALF and CC-ALF), synthetic seeded inputs, filtered in the plain
per-picture form (one kernel per stage).  One step = every picture filtered
once with its input and side information already resident in HBM.

The same JSON line also carries:
- "batched": the same 8 pictures in one persistent fused launch replayed from
  a HIP graph (tile shape from the autotuner); the 64-picture mixed batch
  (16 each of 416x240, 832x480, 1920x1080, 3840x2160); and 4 pictures of
  7680x4320, ALF-heavy with a dense edge grid (every CTB with luma ALF and
  CC-ALF, 8 APS sets, minimum CU 4x8), each with its own roofline;
- "served": continuous batching of Poisson arrivals (120 pictures/s over 32
  streams of 832x480 / 1920x1080 / 3840x2160) with p50 / p99 latency, and a
  saturation run (arrivals far above the device's rate);
- "roofline" of the dominant plain kernel and of the fused kernel, from HIP
  events recorded on the launch stream, with the HBM traffic and the vector
  ALU instruction count per dispatch from the rocprofv3 summaries in
  profiles/ ("valu": issue rate against 256 CUs x 2.4 GHz x one wave64
  instruction per CU-cycle, the bound the integer filters actually meet);
- "cpu_baseline": the CPU oracle (the C restatement of the reference filters)
  timed on the host's cores, one thread and all threads, with the CPU model.

Before any leg is timed, its output is compared with the oracle's output on
the same inputs, bit for bit; a mismatch aborts the run.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import random
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
STAGES = ["prep_lmcs", "deblock_v", "deblock_h", "sao", "alf_luma", "alf_chroma_ccalf"]
# Algorithmic bytes per luma pixel of each plain stage kernel: the planes it
# transforms, read once and written once (u16 samples, 4:2:0).  The fused
# kernel does the whole chain: 6 bytes per luma pixel.
STAGE_BYTES_PER_PX = {"prep_lmcs": 6.0, "deblock_v": 6.0, "deblock_h": 6.0, "sao": 6.0, "alf_luma": 4.0,
                      "alf_chroma_ccalf": 4.0}
CHAIN_BYTES_PER_PX = 6.0
# Parallel graph branches of the plain form: independent pictures' kernel
# sequences overlap (one picture's memory-bound stages beside another's ALF).
PLAIN_LANES = 4
KERNEL_NAMES = {"prep_lmcs": "k_prep", "deblock_v": "k_dbk<true>", "deblock_h": "k_dbk<false>", "sao": "k_sao",
                "alf_luma": "k_alf_luma", "alf_chroma_ccalf": "k_alf_chroma"}
# Host CPU share of one GPU on the benchmark machines.
HOST_THREADS_CAP = 16


def load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(HOST_THREADS_CAP, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_oracle(vv):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import binding
    return binding.load(vv)


class Reference:
    """Oracle outputs of generated pictures, computed on host threads."""

    def __init__(self, vv, orc):
        self.vv, self.orc = vv, orc

    def outputs(self, specs):
        vv = self.vv
        frames = [vv.SynthFrame(cfg, w, h, seed=seed, **kw) for cfg, w, h, seed, kw in specs]
        arr = (C.POINTER(vv.Frame) * len(frames))(*[f.desc for f in frames])
        rc = self.orc.oracle_apply_in_loop_filters_mt(arr, len(frames), min(len(frames), host_threads()))
        if rc != 0:
            raise RuntimeError(f"oracle failed ({rc})")
        outs = [f.output_arrays() for f in frames]
        for f in frames:
            f.close()
        return outs


def same(ref, got, what):
    for name, a, b in zip(("luma", "cb", "cr"), ref, got):
        if not np.array_equal(a, b):
            raise SystemExit(f"bench: {what}: {name} differs from the oracle at {int((a != b).sum())} samples")


def verify_jobs(vv, jobs, frames, refs, what):
    for i, (j, f, r) in enumerate(zip(jobs, frames, refs)):
        j.download(C.byref(f.desc.contents.pic))
        same(r, f.output_arrays(), f"{what}, picture {i}")


def cpu_baseline(vv, orc, seconds):
    """The oracle on host cores: one thread on one 3840x2160 picture, and all
    threads with one picture per thread, each repeated to about `seconds` of
    CPU work."""
    W, H = 3840, 2160
    nt = host_threads()

    def timed(n_pics, threads):
        frames = [vv.SynthFrame(1, W, H, seed=1000 + i) for i in range(n_pics)]
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
        for f in frames:
            f.close()
        return reps, wall, cpu

    r1, w1, c1 = timed(1, 1)
    rn, wn, cn = timed(nt, nt)
    single = round(r1 * W * H / 1e6 / w1, 3)
    allc = round(rn * nt * W * H / 1e6 / wn, 3)
    return {"value": allc, "unit": "Mpixel/s", "cores": nt, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{rn} x {nt} pictures 3840x2160 of the benchmark workload (one per thread), oracle C "
                      f"restatement of the reference filters, {cn:.1f} s CPU over {wn:.1f} s wall",
            "single_thread": {"value": single, "unit": "Mpixel/s", "cores": 1,
                              "sample": f"{r1} x 1 picture 3840x2160, {c1:.1f} s CPU over {w1:.1f} s wall"}}


def plain_run(vv, ctx, jobs, steps, warmup, lanes=1):
    """Per-picture kernel sequences timed over `steps` passes: with lanes = 1
    each job's captured graph replayed in turn on one stream; with lanes > 1
    all jobs in one graph of that many parallel branches (independent
    pictures overlap).  Then a shorter sequential pass with HIP events between
    the stage kernels for the per-stage split."""
    sh = ctx.stream  # every kernel goes to this HIP stream
    g = vv.JobsGraph(ctx, jobs, lanes) if lanes > 1 else None

    def one_pass():
        if g:
            g.launch(sh)
        else:
            for j in jobs:
                j.graph_launch(sh)
    for _ in range(warmup):
        one_pass()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_pass()
    ctx.synchronize()
    total_ms = (time.perf_counter() - t0) * 1e3
    if g:
        g.close()
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


def batched_run(vv, ctx, jobs, frames, refs, what, reps, warmup=2):
    """Autotune, build the batch, check one graph replay against the oracle,
    then time graph replays; returns a dict."""
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    tune = vv.Tune()
    vv.check(vv.lib.vvclf_autotune(ctx.h, arr, len(jobs), C.byref(tune)))
    b = C.c_void_p()
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), C.byref(tune), C.byref(b)))
    vv.check(vv.lib.vvclf_batch_graph_launch(b, ctx.stream))
    ctx.synchronize()
    verify_jobs(vv, jobs, frames, refs, f"{what} (fused {tune.tile_w}x{tune.tile_h})")
    ms = C.c_double()
    vv.check(vv.lib.vvclf_batch_time(b, ctx.stream, warmup, reps, 1, C.byref(ms)))
    vv.lib.vvclf_batch_destroy(b)
    px = sum(j.luma_pixels for j in jobs)
    return {"ms_per_launch": round(ms.value, 4), "luma_pixels_per_launch": px,
            "value": round(px / (ms.value / 1e3) / 1e6, 2), "unit": "Mpixel/s",
            "tile": f"{tune.tile_w}x{tune.tile_h}", "wgs_per_cu": tune.wgs_per_cu,
            "hbm_algorithmic_GBs": round(CHAIN_BYTES_PER_PX * px / (ms.value / 1e3) / 1e9, 2),
            "checked_against_oracle": True}


SERVE_STAGES = ["admission", "host_tables_and_upload", "wait_for_launch", "device", "download"]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_instr.json")
# Issue ceiling of vector ALU work: 256 CUs x 2.4 GHz x one wave64 VALU
# instruction per CU per cycle (four SIMD16 units, four cycles per wave64 op).
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9


def profile_entry(path, workload, kernel):
    """The entry of a per-kernel rocprofv3 summary in profiles/ for
    `workload` whose name starts with `kernel` (template arguments may
    follow), or None."""
    try:
        with open(path) as fh:
            d = json.load(fh).get(workload, {})
    except (OSError, ValueError):
        return None
    for k, v in d.items():
        name = k.replace("void ", "")
        if name.startswith(kernel) and (len(name) == len(kernel) or name[len(kernel)] in "<,"):
            return v
    return None


def traffic_for(workload, kernel):
    """HBM bytes per dispatch of `kernel` on `workload` (FETCH_SIZE + WRITE_SIZE,
    calibrated), from the rocprofv3 summary in profiles/, or None when not
    profiled."""
    v = profile_entry(TRAFFIC_FILE, workload, kernel)
    return int(round(v["read_bytes"] + v["write_bytes"])) if v else None


def valu_for(workload, kernel, avg_launch_ms):
    """The kernel's vector-ALU issue rate against the VALU ceiling: SQ_INSTS_VALU
    per dispatch (rocprofv3 summary in profiles/) over the live launch time."""
    v = profile_entry(VALU_FILE, workload, kernel)
    if not v or not avg_launch_ms:
        return None
    rate = v["valu_wave_instr"] / (avg_launch_ms / 1e3)
    return {"wave_instr_per_launch": int(v["valu_wave_instr"]), "achieved_G_per_s": round(rate / 1e9, 1),
            "peak_G_per_s": round(VALU_PEAK_WAVE_INSTR_PER_S / 1e9, 1),
            "frac": round(rate / VALU_PEAK_WAVE_INSTR_PER_S, 4)}


def fused_roofline(bt, workload, px):
    kern = "k_fused<%d, %d" % tuple(int(v) for v in bt["tile"].split("x"))
    return {"bound": "hbm", "kernel": "k_fused", "achieved": bt["hbm_algorithmic_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(bt["hbm_algorithmic_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic_for(workload, kern), "algorithmic_bytes_per_launch": CHAIN_BYTES_PER_PX * px,
            "avg_launch_ms": bt["ms_per_launch"], "valu": valu_for(workload, kern, bt["ms_per_launch"])}


SERVE_SIZES = [(832, 480), (1920, 1080), (3840, 2160)]


def served_run(vv, ctx, ref, seconds, rate, n_streams=32, seed=7, max_jobs=None, verify=True, prep_threads=4):
    """Poisson arrivals at `rate` pictures/s over `n_streams` streams (two
    pictures per stream, reused); each picture's final output is compared with
    the oracle's at the end."""
    rnd = random.Random(seed)
    cfg = vv.ServerConfig()
    cfg.hbm_budget_bytes = 64 << 30
    cfg.max_inflight = 64
    cfg.max_batch = 16
    cfg.prep_threads = prep_threads
    s = C.c_void_p()
    vv.check(vv.lib.vvclf_server_create(ctx.h, C.byref(cfg), C.byref(s)))
    streams, specs = [], []
    for k in range(n_streams):
        w, h = SERVE_SIZES[rnd.randrange(3)]
        sid = C.c_uint32()
        vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(sid)))
        # two pictures per stream sharing the stream's APS / LMCS state
        sp = [(4, w, h, 5000 + 10 * k + i, {"aps_seed": 900 + k}) for i in range(2)]
        frames = [vv.SynthFrame(c, ww, hh, seed=sd, **kw) for c, ww, hh, sd, kw in sp]
        streams.append((sid.value, frames))
        specs.extend(zip(frames, sp))
    # warm-up: every frame once, so that the first-use costs (a stream's
    # resident LMCS / ALF tables, the prep threads' pinned staging buffers,
    # the device memory pool) fall outside the statistics
    for sid, frames in streams:
        for f in frames:
            vv.check(vv.lib.vvclf_server_submit(s, sid, f.desc, C.byref(C.c_uint64())))
    vv.check(vv.lib.vvclf_server_drain(s, 120000))
    vv.check(vv.lib.vvclf_server_reset_stats(s))
    n_jobs = int(seconds * rate) if max_jobs is None else max_jobs
    picks = [rnd.randrange(n_streams) for _ in range(n_jobs)]
    # A frame's last submission starts from its generated input, so that its
    # final output is checked against the oracle.  At an offered rate every
    # submission does; at saturation the earlier ones filter what the frame
    # holds (the same work on the same side information), since restoring a
    # 4K picture's planes on the one submitting thread would bound the rate.
    last = {}
    for i, k in enumerate(picks):
        last[id(streams[k][1][i % 2])] = i
    busy, used = {}, set()
    t_next = time.perf_counter()
    for i in range(n_jobs):
        if rate:
            t_next += rnd.expovariate(rate)
            while time.perf_counter() < t_next:
                time.sleep(max(0.0, min(0.001, t_next - time.perf_counter())))
        sid, frames = streams[picks[i]]
        f = frames[i % 2]
        prev = busy.get(id(f))
        if prev is not None:  # the frame's previous job must be done before reuse
            vv.lib.vvclf_server_wait(s, prev, 60000)
        if rate or last[id(f)] == i:
            f.reset()
        t = C.c_uint64()
        vv.check(vv.lib.vvclf_server_submit(s, sid, f.desc, C.byref(t)))
        busy[id(f)] = t.value
        used.add(id(f))
    vv.check(vv.lib.vvclf_server_drain(s, 120000))
    st = vv.ServerStats()
    vv.check(vv.lib.vvclf_server_get_stats(s, C.byref(st)))
    vv.lib.vvclf_server_destroy(s)
    if verify:
        todo = [(f, sp) for f, sp in specs if id(f) in used]
        refs = ref.outputs([sp for _, sp in todo])
        for (f, sp), r in zip(todo, refs):
            same(r, f.output_arrays(), f"served picture {sp}")
    return {"streams": n_streams, "offered_pictures_per_s": rate if rate else "saturation",
            "warmup_pictures": sum(len(fr) for _, fr in streams), "pictures": int(st.completed),
            "failed": int(st.failed), "launches": int(st.launches),
            "pictures_per_launch": round(st.completed / st.launches, 2) if st.launches else None,
            "throughput_Mpixel_s": round(st.luma_pixels / st.wall_s / 1e6, 2) if st.wall_s else None,
            "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3), "max_ms": round(st.max_ms, 3),
            "admitted_bytes_peak": int(st.bytes_admitted_peak), "hbm_budget_bytes": int(cfg.hbm_budget_bytes),
            "stage_p50_ms": {k: round(st.stage_p50_ms[i], 3) for i, k in enumerate(SERVE_STAGES)},
            "stage_p99_ms": {k: round(st.stage_p99_ms[i], 3) for i, k in enumerate(SERVE_STAGES)},
            "checked_against_oracle": bool(verify),
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
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-same8", action="store_true", help="skip the fused batch of the same 8 pictures")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (a second known-size kernel for the "
                         "HBM traffic counters) and exit")
    ap.add_argument("--no-mixed64", action="store_true")
    ap.add_argument("--no-8k", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-rate", type=float, default=120.0)
    ap.add_argument("--saturation-jobs", type=int, default=300)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (pictures shard by stream "
              f"across GPUs with no collectives, so the per-GPU figure is the one reported)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy > 0:
        # device-to-device copies of a known size through torch (HIP's copy
        # kernel), to scale FETCH_SIZE / WRITE_SIZE: nothing else runs
        import torch
        n = args.calibrate_copy << 20
        a = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)                          # HIP's device-to-device copy
            torch.bitwise_or(a, 0, out=b)       # an elementwise kernel: n bytes in, n out
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "passes": 4}), flush=True)
        return

    vv = load("vvclf_amd", os.path.join(ROOT, "oxideav-h266_amd", "__init__.py"))
    orc = load_oracle(vv)
    ref = Reference(vv, orc)
    W, H, N = args.width, args.height, args.pictures
    log("generating workload and its oracle outputs")
    specs = [(1, W, H, 1000 + i, {}) for i in range(N)]
    refs = ref.outputs(specs)
    frames = [vv.SynthFrame(c, w, h, seed=s, **kw) for c, w, h, s, kw in specs]
    ctx = vv.Context(0)
    jobs = [vv.Job(ctx, f.desc) for f in frames]
    log("plain form")
    for j in jobs:
        j.graph_launch(ctx.stream)
    ctx.synchronize()
    verify_jobs(vv, jobs, frames, refs, "plain form")
    g = vv.JobsGraph(ctx, jobs, PLAIN_LANES)
    g.launch(ctx.stream)
    ctx.synchronize()
    g.close()
    verify_jobs(vv, jobs, frames, refs, f"plain form on {PLAIN_LANES} graph lanes")
    ms_seq, _ = plain_run(vv, ctx, jobs, args.steps, args.warmup)
    ms_per_step, stage_ms = plain_run(vv, ctx, jobs, args.steps, args.warmup, lanes=PLAIN_LANES)
    px = N * W * H
    value = px / (ms_per_step / 1e3) / 1e6
    dom = max(stage_ms, key=stage_ms.get)
    dom_bytes = STAGE_BYTES_PER_PX[dom] * W * H
    achieved = dom_bytes / (stage_ms[dom] / 1e3) / 1e9
    wl = f"{N}x{W}x{H}"
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
                               f"SAO, ALF 4 luma APS + chroma ALF + CC-ALF); plain per-picture stage kernels, "
                               f"the pictures' sequences on {PLAIN_LANES} parallel graph branches",
                   "pictures_per_step": N, "luma_pixels_per_step": px, "form": "plain", "lanes": PLAIN_LANES},
        "checked_against_oracle": True,
        "plain_one_lane": {"ms_per_step": round(ms_seq, 4), "value": round(px / (ms_seq / 1e3) / 1e6, 2),
                           "unit": "Mpixel/s", "note": "each picture's graph replayed in turn on one stream"},
        "chain_hbm_algorithmic_GBs": round(CHAIN_BYTES_PER_PX * px / (ms_per_step / 1e3) / 1e9, 2),
        "stage_ms_per_picture": {k: round(v, 5) for k, v in stage_ms.items()},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_for(wl, KERNEL_NAMES[dom]),
                     "algorithmic_bytes_per_launch": dom_bytes,
                     "avg_launch_ms": round(stage_ms[dom], 5),
                     "valu": valu_for(wl, KERNEL_NAMES[dom], stage_ms[dom])},
    }
    out["batched"] = {}
    if not args.no_batched and not args.no_same8:
        log("batched form, same pictures")
        bt = batched_run(vv, ctx, jobs, frames, refs, "fused batch 8x4K", reps=max(5, args.steps // 2))
        bt["speedup_vs_plain"] = round(bt["value"] / value, 2)
        bt["roofline"] = fused_roofline(bt, wl, px)
        out["batched"]["same_8x4k"] = bt
    for j in jobs:
        j.close()
    jobs = []
    if not args.no_dropin:
        # the drop-in call end to end: host planes in, host planes out (PCIe
        # and host table derivation included), one picture at a time
        log("drop-in host-to-host")
        ts = []
        for rep in range(3):
            for f, r in zip(frames, refs):
                f.reset()
                t0 = time.perf_counter()
                vv.check(vv.lib.vvclf_apply_in_loop_filters(f.desc))
                ts.append(time.perf_counter() - t0)
                if rep == 0:
                    same(r, f.output_arrays(), "drop-in call")
        ts = ts[len(frames):]  # first pass warms the pools and staging buffers
        out["dropin_host_to_host"] = {"value": round(len(ts) * W * H / sum(ts) / 1e6, 2), "unit": "Mpixel/s",
                                      "ms_per_picture": round(1e3 * sum(ts) / len(ts), 3),
                                      "ms_per_picture_median": round(1e3 * sorted(ts)[len(ts) // 2], 3),
                                      "checked_against_oracle": True,
                                      "note": "vvclf_apply_in_loop_filters per picture: host table derivation, "
                                              "upload, plain kernels, download"}
    for f in frames:
        f.close()
    frames, refs = [], []
    if not args.no_batched and not args.no_mixed64:
        log("batched form, 64 mixed pictures")
        sizes = [(416, 240)] * 16 + [(832, 480)] * 16 + [(1920, 1080)] * 16 + [(3840, 2160)] * 16
        mspecs = [(2, w, h, 2000 + i, {}) for i, (w, h) in enumerate(sizes)]
        mrefs = ref.outputs(mspecs)
        mframes = [vv.SynthFrame(c, w, h, seed=s, **kw) for c, w, h, s, kw in mspecs]
        mjobs = [vv.Job(ctx, f.desc) for f in mframes]
        bm = batched_run(vv, ctx, mjobs, mframes, mrefs, "fused batch, 64 mixed", reps=5)
        # plain per-picture reference point on the same 64 pictures
        ms_plain, _ = plain_run(vv, ctx, mjobs, 3, 1, lanes=PLAIN_LANES)
        bm["plain_value"] = round(bm["luma_pixels_per_launch"] / (ms_plain / 1e3) / 1e6, 2)
        bm["speedup_vs_plain"] = round(bm["value"] / bm["plain_value"], 2)
        bm["roofline"] = fused_roofline(bm, "64 mixed", bm["luma_pixels_per_launch"])
        out["batched"]["mixed_64"] = bm
        for j in mjobs:
            j.close()
        for f in mframes:
            f.close()
        mrefs = None
    if not args.no_batched and not args.no_8k:
        log("batched form, 4x 7680x4320 ALF-heavy")
        especs = [(3, 7680, 4320, 6000 + i, {}) for i in range(4)]
        erefs = ref.outputs(especs)
        eframes = [vv.SynthFrame(c, w, h, seed=s, **kw) for c, w, h, s, kw in especs]
        ejobs = [vv.Job(ctx, f.desc) for f in eframes]
        be = batched_run(vv, ctx, ejobs, eframes, erefs, "fused batch, 4x 8K", reps=10)
        ms_plain, _ = plain_run(vv, ctx, ejobs, 3, 1, lanes=PLAIN_LANES)
        be["plain_value"] = round(be["luma_pixels_per_launch"] / (ms_plain / 1e3) / 1e6, 2)
        be["speedup_vs_plain"] = round(be["value"] / be["plain_value"], 2)
        be["roofline"] = fused_roofline(be, "4x7680x4320", be["luma_pixels_per_launch"])
        be["workload"] = ("4x 7680x4320 4:2:0 10-bit, every CTB luma ALF + CC-ALF, 8 luma APS sets, minimum CU 4x8, "
                          "bS 2 at p 0.5")
        out["batched"]["alf_heavy_4x8k"] = be
        for j in ejobs:
            j.close()
        for f in eframes:
            f.close()
        erefs = None
    if args.serve_seconds > 0:
        log("serving")
        out["served"] = served_run(vv, ctx, ref, args.serve_seconds, args.serve_rate)
        if args.saturation_jobs > 0:
            log("serving at saturation")
            out["served_saturation"] = served_run(vv, ctx, ref, 0, 0, max_jobs=args.saturation_jobs, seed=11)
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(vv, orc, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
