"""Benchmark: remap + pyramid-contrast fusion of one synthetic stack per step on one MI355X.

Workload (N=1 line, --config 2, the default): the streaming large stack, 143 frames of
7952x5304 RGB u16 (12-bit values scaled to 16 bit) from the seeded synthetic generator
(band-limited value-noise scene, depth ramp + three Gaussian bumps, per-frame defocus blur sigma
up to 6 px, similarity poses with shifts U(+-20) px, roll U(+-0.3) deg, scale 1+U(+-0.01)),
remapped into the common-overlap ROI and fused with the reference's full pyramid depth into
device-resident FP32 per-level accumulators.  --config 1 is the 30-frame u8 stack.
Frames are generated on the GPU and stay resident in HBM; the timed region is K steps of
{reset, N x add_frame_device (remap + pass 1), finish_device (pass 2 + collapse)}.

Output: one JSON line.  roofline: the longer of the two dominant launches (k_pass1, the level-0
k_pass2), its per frame-pixel algorithmic bytes (kernel_models) times the frame-pixels of one
launch over its mean duration from HIP events on the stack's stream in the timed region.
cpu_baseline: the C oracle (the reference's algorithm restated, its independent loops spread over
every core the job has) fusing the first two frames of the same workload as a full-size stack.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "focus-stacker_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK_HBM_GBS = 8000.0
METRIC = "fused MPix/s (remap+pyramid-contrast), %HBM roofline; p50/p99 stack latency"


def alg_bytes_per_frame_px(sample_bytes: int, n: int) -> float:
    """Two-pass streaming model: pass 1 = 3s + 14, pass 2 = 3s + 41, collapse (12*4/3 + 3s)/N."""
    s = sample_bytes
    return (3 * s + 14) + (3 * s + 41) + (16 + 3 * s) / n


def kernel_models(s: int) -> dict:
    """Per frame-pixel algorithmic bytes of the two dominant launches (s = sample bytes).
    k_pass1: the pass-1 term, read the source (3s), write w (4), RMW the norm sum (8) and the
    union support (2).  k_pass2_level0: the level-0 share of the pass-2 term, read the source
    (3s), w (4), S (4), U (1) and RMW the level-0 accumulator (24)."""
    return {"k_pass1": ("pass1", 3 * s + 14), "k_pass2_level0": ("pass2_level0", 3 * s + 33)}


# HBM bytes per launch, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs scaled by the same
# counters on a plain copy of known size (profiles/pmc_traffic_<workload>.md), keyed by
# (frames, width, height, samples, method, levels)
TRAFFIC = {
    (143, 7952, 5304, "u16", "pyramid", 0): {"k_pass1": 130.303e9, "k_pass2_level0": 160.602e9},
    (30, 7952, 5304, "u8", "pyramid", 0): {"k_pass1": 16.125e9, "k_pass2_level0": 20.356e9},
}
COPY_PROBE_BYTES = 256 << 20


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    k = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[k]


def job_cores() -> int:
    """Cores this job may use: OMP_NUM_THREADS when set (the measurement boxes set it to the
    host-core share of one GPU), else the CPUs in the process's affinity mask."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return n if n > 0 else len(os.sched_getaffinity(0))


def cpu_baseline(frames_np, poses, roi, method=0, max_levels=0):
    """The oracle (the C restatement of the reference's algorithm, its independent per-pixel and
    per-row loops spread over every core of the job, as the reference spreads row blocks over its
    --threads) on a bounded sample of the same workload: the first two frames fused as a stack over
    the whole ROI, the same per frame-pixel work as every frame of the full stack (remap, weights,
    norm sum, both pyramids, accumulation; the collapse once)."""
    from oracle import Oracle, make_pose
    o = Oracle()
    cores = job_cores()
    o.set_threads(cores)
    l, t, r, b = roi
    src, opo = frames_np[:2], [make_pose(**p) for p in poses[:2]]
    c0 = time.process_time()
    t0 = time.perf_counter()
    o.stack(src, opo, roi, method=method, max_levels=max_levels)
    dt = time.perf_counter() - t0
    cpu_s = time.process_time() - c0
    px = len(src) * (r - l) * (b - t)
    return {"value": round(px / dt / 1e6, 3), "unit": "MPix/s", "cores": cores, "kind": "port",
            "sample": f"oracle/focus_oracle.c fusing frames 0-1 of the workload as a 2-frame stack over the whole "
                      f"ROI ({r - l}x{b - t}) on {cores} threads: remap + weights + pyramids + accumulation + "
                      f"collapse, {dt:.1f} s wall, {cpu_s:.1f} s of CPU time"}


MIX_SIZES = [(640, 480), (1920, 1080), (4000, 3000), (6000, 4000)]
BIG_SIZE = (7952, 5304)


def _pool(synth, torch, dev, config_id, stack_id, frames, w, h, **kw):
    spec = synth.config_spec(config_id, stack_id=stack_id, frames=frames, width=w, height=h, **kw)
    st = synth.Stack(spec, device=dev)
    out = [st.frame(k).contiguous() for k in range(frames)]
    del st
    torch.cuda.synchronize()
    return out


def make_mix(fg, synth, torch, dev, n_stacks, seed=7, sizes=None, n_big=0):
    """The uneven stack mix: n_stacks stacks of frame counts U{4..40}, sizes drawn from MIX_SIZES,
    poses of the 1024x768 configuration scaled to the size (each stack its own seeded poses), then
    n_big stacks of 100-143 frames of 7952x5304.  Frames come from the seeded generator: per size one
    generated focus stack (40 frames; 143 of the 30-frame workload's generator for the large size),
    whose first n frames a stack of that size takes as its sources."""
    rng = np.random.default_rng(seed)
    sizes = sizes or MIX_SIZES
    pools = {}
    for i, (w, h) in enumerate(sizes):
        pools[(w, h)] = _pool(synth, torch, dev, 0, 900 + i, 40, w, h, shift=3.0 * w / 1024.0)
        print(f"bench: frame pool {w}x{h} ready", file=sys.stderr, flush=True)
    if n_big > 0:
        pools[BIG_SIZE] = _pool(synth, torch, dev, 1, 0, 143, *BIG_SIZE)
        print("bench: frame pool 7952x5304 ready", file=sys.stderr, flush=True)
    stacks = []
    for i in range(n_stacks + n_big):
        big = i >= n_stacks
        if big:
            n = int(rng.integers(100, 144))
            w, h = BIG_SIZE
            spec = synth.config_spec(1, stack_id=i, frames=n)
        else:
            n = int(rng.integers(4, 41))
            w, h = sizes[int(rng.integers(0, len(sizes)))]
            spec = synth.config_spec(0, stack_id=i, frames=n, width=w, height=h, shift=3.0 * w / 1024.0)
        poses = [fg.pose(p) for p in synth.make_poses(spec)]
        roi = fg.common_overlap_roi([(w, h)] * n, poses)
        stacks.append(dict(n=n, w=w, h=h, big=big, frames=[f.data_ptr() for f in pools[(w, h)][:n]], poses=poses,
                           roi=roi))
    return stacks, pools


def _roi_px(s):
    return (s["roi"][2] - s["roi"][0]) * (s["roi"][3] - s["roi"][1])


def _plain_stack(fg, ctx, s, out_ptr):
    """One request through the plain per-stack calls (the drop-in path): create, add, finish, destroy."""
    rw, rh = s["roi"][2] - s["roi"][0], s["roi"][3] - s["roi"][1]
    st = fg.Stack(ctx, s["n"], rw, rh, fg.SAMPLE_U8)
    desc = fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8)
    for f, p in zip(s["frames"], s["poses"]):
        st.add_frame_device(f, desc, p, s["roi"])
    if st.needs_pass2_frames():
        for f, p in zip(s["frames"], s["poses"]):
            st.pass2_frame_device(f, desc, p, s["roi"])
    st.finish_device(out_ptr)
    st.sync()
    st.close()


def run_batch(args):
    """The 64-stack uneven workload three ways, one after the other on the same requests: the plain
    per-stack calls (sequential), the server (all submitted at once, --slots in flight) and the batch
    engine (one persistent pass-1 launch and one launch per level per group): its first run (plans
    built, launches one by one), then, after a second run that captures the group schedules, a run
    that replays them."""
    import torch
    import focus_gpu as fg
    import synth
    dev = torch.device("cuda:0")
    sizes = [tuple(int(v) for v in x.split("x")) for x in args.sizes] if args.sizes else None
    stacks, pools = make_mix(fg, synth, torch, dev, args.stacks, sizes=sizes)
    ctx = fg.Context(0, memory_fraction=0.8)
    px = sum(s["n"] * _roi_px(s) for s in stacks)
    outs = {k: [torch.zeros((s["roi"][3] - s["roi"][1], s["roi"][2] - s["roi"][0], 3), dtype=torch.uint8, device=dev)
                for s in stacks] for k in ("plain", "server", "batch")}
    torch.cuda.synchronize()

    def reqs(key):
        return [dict(frames=s["frames"], desc=fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8), out=o.data_ptr(),
                     poses=s["poses"], roi=s["roi"]) for s, o in zip(stacks, outs[key])]

    paths = set(args.paths)
    t_plain = t_srv = t_first = t_batch = float("nan")
    st_srv, st_b = [], []
    print(f"bench: {len(stacks)} stacks, {px / 1e9:.2f} GPix", file=sys.stderr, flush=True)
    if "plain" in paths:
        for s, o in list(zip(stacks, outs["plain"]))[:4]:
            _plain_stack(fg, ctx, s, o.data_ptr())
        t0 = time.perf_counter()
        for s, o in zip(stacks, outs["plain"]):
            _plain_stack(fg, ctx, s, o.data_ptr())
        t_plain = time.perf_counter() - t0
        print(f"bench: plain {t_plain:.3f} s", file=sys.stderr, flush=True)
    if "server" in paths:
        with fg.Server(ctx, slots=args.slots, memory_fraction=0.8) as srv:
            for r in reqs("server")[:4]:
                srv.wait(srv.submit(r["frames"], r["desc"], r["out"], poses=r["poses"], roi=r["roi"]))
            t0 = time.perf_counter()
            tickets = [srv.submit(r["frames"], r["desc"], r["out"], poses=r["poses"], roi=r["roi"])
                       for r in reqs("server")]
            st_srv = [srv.wait(t) for t in tickets]
            t_srv = time.perf_counter() - t0
        print(f"bench: server {t_srv:.3f} s", file=sys.stderr, flush=True)
    if "batch" in paths:
        with fg.Batch(ctx, memory_fraction=0.8) as b:
            t0 = time.perf_counter()
            b.fuse(reqs("batch"))                       # first run: plans built, launches one by one
            t_first = time.perf_counter() - t0
            b.fuse(reqs("batch"))                       # second run: schedules captured
            t0 = time.perf_counter()
            st_b = b.fuse(reqs("batch"))                # replays the captured schedules
            t_batch = time.perf_counter() - t0
        print(f"bench: batch {t_batch:.3f} s (first run {t_first:.3f} s)", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ran = [k for k in ("plain", "server", "batch") if k in paths]
    same = all(all(bool(torch.equal(outs[ran[0]][i], outs[k][i])) for k in ran[1:]) for i in range(len(stacks)))
    lat_b = [x["total_ms"] for x in st_b] or [float("nan")]
    lat_s = [x["total_ms"] for x in st_srv] or [float("nan")]

    def rnd(v):
        return None if v != v else round(v, 3)
    size_txt = ", ".join(f"{w}x{h}" for w, h in (sizes or MIX_SIZES))
    line = {"metric": f"stacks/s, uneven stacks (4-40 frames of {size_txt} u8): batch vs server vs plain",
            "value": rnd(len(stacks) / t_batch), "unit": "stacks/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": rnd(t_batch * 1000.0), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": f"{len(stacks)} uneven stacks, remap + pyramid-contrast, full depth", "samples": "u8",
                       "slots": args.slots, "frame_pixels": px},
            "server_stacks_per_s": rnd(len(stacks) / t_srv), "plain_stacks_per_s": rnd(len(stacks) / t_plain),
            "batch_first_run_stacks_per_s": rnd(len(stacks) / t_first),
            "speedup_vs_server": rnd(t_srv / t_batch), "speedup_vs_plain": rnd(t_plain / t_batch),
            "fused_MPix_per_s": {"batch": rnd(px / t_batch / 1e6), "server": rnd(px / t_srv / 1e6),
                                 "plain": rnd(px / t_plain / 1e6)},
            "latency_ms": {"batch_p50": rnd(pct(lat_b, 50)), "batch_p99": rnd(pct(lat_b, 99)),
                           "server_p50": rnd(pct(lat_s, 50)), "server_p99": rnd(pct(lat_s, 99))},
            "paths": ran, "identical_outputs": same}
    print(json.dumps(line), flush=True)


def _arrivals(rng, rate, duration, min_requests, n_small, n_big, big_share):
    """Poisson arrival times over `duration` s (at least min_requests of them) and, per request, a
    template: with probability big_share one of the n_big full-size stacks, else one of the mix."""
    times, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration and len(times) >= min_requests:
            break
        times.append(t)
    kinds = [n_small + int(rng.integers(0, n_big)) if n_big and rng.random() < big_share else int(rng.integers(0, n_small))
             for _ in times]
    return times, kinds


def run_serve(args):
    """The serving workload: stack requests arrive as a Poisson process at each rate of --rates; each
    draws its frame count and size from the uneven mix (4-40 frames of 640x480 ... 6000x4000 u8), and
    5% are 100-143 frames of 7952x5304.  Admission by the device-memory budget (0.8 of free HBM).
    Each rate offers --duration s of arrivals (at least --min-requests), first to the server (slots
    in flight, continuous admission), then the same arrival schedule to the plain per-stack path
    (one request at a time, the drop-in call), whose outputs must equal the server's bit for bit.
    Latency is submit (arrival) to done; throughput is requests over first arrival to last done."""
    import torch
    import focus_gpu as fg
    import synth
    dev = torch.device("cuda:0")
    stacks, pools = make_mix(fg, synth, torch, dev, args.stacks, n_big=args.big_templates)
    n_small = sum(1 for s in stacks if not s["big"])
    n_big = len(stacks) - n_small
    torch.cuda.synchronize()
    ctx = fg.Context(0, memory_fraction=0.8)
    rng = np.random.default_rng(11)

    def out_for(s):
        rw, rh = s["roi"][2] - s["roi"][0], s["roi"][3] - s["roi"][1]
        return torch.empty((rh, rw, 3), dtype=torch.uint8, device=dev)

    def submit(srv, s, out):
        d = fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8)
        return srv.submit(s["frames"], d, out.data_ptr(), poses=s["poses"], roi=s["roi"])

    def stats(lat, kinds, t_end, n):
        small = [x for x, k in zip(lat, kinds) if not stacks[k]["big"]] or [float("nan")]
        big = [x for x, k in zip(lat, kinds) if stacks[k]["big"]] or [float("nan")]
        return {"stacks_per_s": round(n / t_end, 3), "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
                "small_p50_ms": round(pct(small, 50), 2), "small_p99_ms": round(pct(small, 99), 2),
                "big_p50_ms": round(pct(big, 50), 2) if big[0] == big[0] else None}

    rates = []
    same_all = True
    for rate in args.rates:
        times, kinds = _arrivals(rng, rate, args.duration, args.min_requests, n_small, n_big, 0.05)
        outs = [out_for(stacks[k]) for k in kinds]
        torch.cuda.synchronize()
        px = sum(stacks[k]["n"] * _roi_px(stacks[k]) for k in kinds)
        # server (a fresh one per rate, so that the plain path below gets the memory back): submissions
        # at the arrival times, then the waits
        with fg.Server(ctx, slots=args.slots, memory_fraction=0.8) as srv:
            # warm-up: one request of every template (first-use costs: pooled stacks, graph captures)
            warm = [(submit(srv, s, o), o) for s, o in ((s, out_for(s)) for s in stacks)]
            for t, _ in warm:
                srv.wait(t)
            del warm
            t0 = time.perf_counter()
            tickets = []
            for at, k, o in zip(times, kinds, outs):
                while time.perf_counter() - t0 < at:
                    time.sleep(min(0.0005, max(0.0, at - (time.perf_counter() - t0))))
                tickets.append(submit(srv, stacks[k], o))
            status = [srv.wait(t) for t in tickets]
        t_srv = max(at + st["total_ms"] / 1000.0 for at, st in zip(times, status))
        lat_srv = [st["total_ms"] for st in status]
        srv_line = stats(lat_srv, kinds, t_srv, len(times))
        print(f"bench: serve rate {rate}/s: {len(times)} requests ({sum(stacks[k]['big'] for k in kinds)} full-size),"
              f" server {srv_line}", file=sys.stderr, flush=True)
        # plain path: the same schedule, one request at a time in arrival order
        outs_p = [out_for(stacks[k]) for k in kinds]
        torch.cuda.synchronize()
        lat_p = []
        t0 = time.perf_counter()
        for at, k, o in zip(times, kinds, outs_p):
            while time.perf_counter() - t0 < at:
                time.sleep(min(0.0005, max(0.0, at - (time.perf_counter() - t0))))
            _plain_stack(fg, ctx, stacks[k], o.data_ptr())
            lat_p.append((time.perf_counter() - t0 - at) * 1000.0)
        t_plain = time.perf_counter() - t0
        plain_line = stats(lat_p, kinds, t_plain, len(times))
        same = all(bool(torch.equal(a, b)) for a, b in zip(outs, outs_p))
        same_all = same_all and same
        r = {"rate": rate, "requests": len(times), "full_size_requests": int(sum(stacks[k]["big"] for k in kinds)),
             "offered_s": round(times[-1], 2), "fused_MPix": round(px / 1e6, 1), "server": srv_line,
             "plain": plain_line, "identical_outputs": same}
        rates.append(r)
        print(f"bench: serve rate {rate}/s: plain {plain_line}, identical {same}", file=sys.stderr, flush=True)
        del outs, outs_p
        torch.cuda.empty_cache()
    top = rates[-1]
    line = {"metric": "served stacks/s with p50/p99 latency under Poisson arrivals "
                      "(uneven mix 640x480-6000x4000 u8 + 5% 100-143 x 7952x5304), server vs plain path",
            "value": top["server"]["stacks_per_s"], "unit": "stacks/s", "n_gpus": 1, "steps": len(rates), "warmup": 1,
            "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic",
            "config": {"workload": f"Poisson arrivals at {args.rates} stacks/s, {args.duration} s of arrivals per rate "
                                   f"(at least {args.min_requests} requests)",
                       "templates": {"mix": n_small, "full_size": n_big}, "full_size_share": 0.05, "slots": args.slots,
                       "memory_fraction": 0.8, "samples": "u8"},
            "rates": rates, "identical_outputs": same_all}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=None, help="override the frame count (quick runs)")
    ap.add_argument("--levels", type=int, default=0, help="0 = full depth (reference)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host", action="store_true", help="skip the host-fed (PCIe-inclusive) step")
    ap.add_argument("--profile-only", action="store_true", help="one profiled step (for rocprofv3 runs)")
    ap.add_argument("--config", type=int, choices=(1, 2), default=2,
                    help="2: 143 x 7952x5304 u16, three focal planes (the reported line); 1: 30 x 7952x5304 u8")
    ap.add_argument("--method", choices=("pyramid", "hybrid", "hardmask", "softmask"), default="pyramid")
    ap.add_argument("--workload", choices=("stack", "batch", "serve"), default="stack",
                    help="stack: one large stack per step (default); batch / serve: uneven stack mixes")
    ap.add_argument("--stacks", type=int, default=64)
    ap.add_argument("--paths", nargs="+", choices=("plain", "server", "batch"), default=["plain", "server", "batch"],
                    help="batch workload: which paths to time")
    ap.add_argument("--sizes", nargs="+", default=None,
                    help="batch workload: frame sizes WxH of the mix (default 640x480 1920x1080 4000x3000 6000x4000)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--rates", type=float, nargs="+", default=[0.5, 2.0, 5.0, 10.0, 20.0],
                    help="serve workload: Poisson arrival rates (stacks/s)")
    ap.add_argument("--duration", type=float, default=20.0, help="serve workload: seconds of arrivals per rate")
    ap.add_argument("--min-requests", type=int, default=30, help="serve workload: requests per rate at least")
    ap.add_argument("--big-templates", type=int, default=4,
                    help="serve workload: full-size (100-143 x 7952x5304) stacks the 5%% share draws from")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; a stack would shard across GPUs by frames with an "
              "all-reduce of the norm sum and the accumulator pyramid, which is not built "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "serve":
        return run_serve(args)

    import numpy as np
    import torch
    import focus_gpu as fg
    import synth

    dev = torch.device("cuda:0")
    spec = synth.config_spec(args.config)
    if args.frames:
        spec.frames = args.frames
    n = spec.frames
    u16 = spec.dtype == "u16"
    st_type = fg.SAMPLE_U16 if u16 else fg.SAMPLE_U8
    s_b = 2 if u16 else 1
    tname = "u16" if u16 else "u8"
    method = {"pyramid": fg.METHOD_PYRAMID, "hybrid": fg.METHOD_HYBRID, "hardmask": fg.METHOD_HARDMASK,
              "softmask": fg.METHOD_SOFTMASK}[args.method]
    print(f"bench: generating {n} x {spec.width}x{spec.height} {tname} frames on the GPU", file=sys.stderr, flush=True)
    gen = synth.Stack(spec, device=dev)
    frames = []
    for k in range(n):
        f = gen.frame(k)
        if u16:  # int32-coded samples to 16-bit storage
            f = (f - (f > 32767).to(torch.int32) * 65536).to(torch.int16)
        frames.append(f)
        if (k + 1) % 20 == 0:
            print(f"bench: {k + 1} frames", file=sys.stderr, flush=True)
    poses = gen.poses
    del gen
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    fposes = [fg.pose(p) for p in poses]
    roi = fg.common_overlap_roi([(spec.width, spec.height)] * n, fposes)
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    desc = fg.ImageDesc(spec.width, spec.height, 3, 0, st_type)
    ctx = fg.Context(0, memory_fraction=0.8)
    stack = fg.Stack(ctx, n, w, h, st_type, method=method, max_levels=args.levels)
    levels = len(stack.levels())
    out = torch.empty((h, w, 3), dtype=torch.int16 if u16 else torch.uint8, device=dev)
    ext = torch.cuda.ExternalStream(stack.stream())

    def step():
        stack.reset()
        for k in range(n):
            stack.add_frame_device(frames[k].data_ptr(), desc, fposes[k], roi)
        stack.finish_device(out.data_ptr())

    if args.profile_only:
        # a plain copy of known size first: the scale of FETCH_SIZE / WRITE_SIZE per byte moved
        cp_src = torch.rand(COPY_PROBE_BYTES // 4, device=dev)
        cp_dst = torch.empty_like(cp_src)
        torch.mul(cp_src, 2.0, out=cp_dst)
        torch.cuda.synchronize()
        step()
        stack.sync()
        print(json.dumps({"profile_step": "done", "roi": roi, "levels": levels,
                          "copy_probe_bytes": {"read": COPY_PROBE_BYTES, "write": COPY_PROBE_BYTES}}), flush=True)
        return

    print(f"bench: roi={roi} ({w}x{h}), {levels} levels, warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    stack.sync()
    torch.cuda.synchronize()

    # timed region: K steps; HIP events on the stack's stream bracket every step and every kernel
    # group the library launches (fg_stack_set_profiling), so per-launch durations come from the
    # same run as the wall-clock rate
    stack.set_profiling(True)
    evs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(ext)
        step()
        b.record(ext)
        evs.append((a, b))
    stack.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ph = stack.phase_times()
    stack.set_profiling(False)
    lat = [a.elapsed_time(b) for a, b in evs]
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    frame_px = n * w * h
    value = frame_px / (ms_per_step / 1000.0) / 1e6
    print(f"bench: {ms_per_step:.2f} ms/step, {value:.1f} MPix/s, phases {ph}", file=sys.stderr, flush=True)

    kern = {}
    for name, (phase, per_unit) in kernel_models(s_b).items():
        total_ms, launches = ph[phase]
        if launches == 0:
            continue
        avg_ms = total_ms / launches
        # frame-pixels per launch: the stack's frames over its launches of this kind per step (the
        # pyramid methods run one pass-1 and one level-0 launch per stack; the single-pass methods
        # one pass-1 launch per group of frames)
        units = n * w * h * args.steps // launches
        ach = per_unit * units / (avg_ms / 1000.0) / 1e9
        kern[name] = {"avg_launch_ms": round(avg_ms, 4), "launches": launches, "bytes_per_unit": per_unit,
                      "units_per_launch": units, "achieved": round(ach, 1), "frac": round(ach / PEAK_HBM_GBS, 4)}
    dom = max(kern, key=lambda k: kern[k]["avg_launch_ms"])
    pipe_bytes = alg_bytes_per_frame_px(s_b, n) * frame_px
    pipe_gbs = pipe_bytes / (ms_per_step / 1000.0) / 1e9
    cfg_name = f"{n} x {spec.width}x{spec.height} RGB {tname}"
    traffic = TRAFFIC.get((n, spec.width, spec.height, tname, args.method, args.levels), {}).get(dom)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MPix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{cfg_name} stack (synthetic, seeded), remap + {args.method}-contrast, {levels} levels",
                   "frames": n, "source": [spec.width, spec.height], "samples": tname, "roi": list(roi),
                   "levels": levels, "focal_planes": spec.planes,
                   "arithmetic": "fp32 pyramids, masks and accumulators; f64 remap coordinates and contrast sums",
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": round(pct(lat, 50), 3), "p99": round(pct(lat, 99), 3)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["achieved"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": kern[dom]["frac"], "traffic": traffic,
                     "avg_launch_ms": kern[dom]["avg_launch_ms"]},
        "kernels": kern,
        "pipeline": {"alg_bytes_per_frame_px": round(alg_bytes_per_frame_px(s_b, n), 3),
                     "achieved_GBs": round(pipe_gbs, 1), "frac": round(pipe_gbs / PEAK_HBM_GBS, 4)},
        "phases_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in ph.items()},
    }
    if not args.no_host:
        # the same stack fed from pageable host memory through fg_stack_add_frame: PCIe-inclusive
        # (uploads overlap pass 1 of earlier frames on the stack's upload stream)
        print("bench: host-fed steps", file=sys.stderr, flush=True)
        frames_host = [f.cpu().numpy() for f in frames]
        hsteps = 2
        stack.reset()
        for k in range(n):
            stack.add_frame(frames_host[k], fposes[k], roi)
        stack.finish_device(out.data_ptr())
        stack.sync()
        t0 = time.perf_counter()
        for _ in range(hsteps):
            stack.reset()
            for k in range(n):
                stack.add_frame(frames_host[k], fposes[k], roi)
            stack.finish_device(out.data_ptr())
        stack.sync()
        hms = (time.perf_counter() - t0) * 1000.0 / hsteps
        src_gb = n * frames_host[0].nbytes / 1e9
        line["pcie_inclusive"] = {"value": round(frame_px / (hms / 1000.0) / 1e6, 2), "unit": "MPix/s",
                                  "ms_per_step": round(hms, 3), "source_GB_per_step": round(src_gb, 3),
                                  "h2d_GBs": round(src_gb / (hms / 1000.0), 2),
                                  "note": "frames in pageable host memory, fg_stack_add_frame; not the value"}
        # the same from page-locked buffers (fg_host_alloc), as a decoder writing into them would
        pinned = []
        for fh in frames_host:
            pa = fg.PinnedArray(tuple(fh.shape), np.uint16 if u16 else np.uint8)
            pa.array[...] = fh.view(np.uint16) if u16 else fh
            pinned.append(pa)
        del frames_host
        t0 = time.perf_counter()
        for _ in range(hsteps):
            stack.reset()
            for k in range(n):
                stack.add_frame(pinned[k].array, fposes[k], roi)
            stack.finish_device(out.data_ptr())
        stack.sync()
        pms = (time.perf_counter() - t0) * 1000.0 / hsteps
        line["pcie_inclusive"]["pinned"] = {"value": round(frame_px / (pms / 1000.0) / 1e6, 2), "ms_per_step": round(pms, 3),
                                            "h2d_GBs": round(src_gb / (pms / 1000.0), 2)}
        for pa in pinned:
            pa.close()
        print(f"bench: host-fed {hms:.1f} ms/step, pinned {pms:.1f} ms/step", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        del frames[2:]
        torch.cuda.synchronize()
        print("bench: timing the CPU baseline (oracle) on a bounded sample", file=sys.stderr, flush=True)
        frames_np = [f.cpu().numpy().view(np.uint16) if u16 else f.cpu().numpy() for f in frames[:2]]
        omethod = {"pyramid": 0, "hybrid": 1, "hardmask": 2, "softmask": 3}[args.method]
        line["cpu_baseline"] = cpu_baseline(frames_np, poses, roi, omethod, args.levels)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
