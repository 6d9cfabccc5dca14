"""Autotuner of libfocus_gpu's launch shapes (fg_tuning) on the GPU it runs on.

Results do not depend on the launch shapes (every shape computes the same bits; the tuner checks it),
only the time does.  The tuner measures, one knob at a time (the others at their current best):

  batches (the 64-stack uneven mix of bench.py --workload batch, fused by fg_batch_fuse, timed on a
  replay of the captured schedules):
    batch_large_kpix            output size (thousands of pixels) from which a stack runs on its own
    batch_streams               streams of a batch: the groups on the first, stacks of their own on
                                the others
    batch_group_mib             MiB of stack buffers one group may hold (0 = the whole budget)
    batch_pass1_blocks_per_cu   persistent pass-1 blocks per CU (0 = as many as fit)
    batch_wide_level            mean level width from which a level's launch uses 64x16 tiles
  single stacks (fg_stack_* on device frames: 30 x 7952x5304, 20 x 1920x1080 and 10 x 640x480 u8,
  timed over the pass-2 phase):
    pass2_min_tiles             64x16 tiles while a level has this many, then 16x16, then one frame
                                per block with an ordered sum

It writes every measurement and the chosen values to focus-stacker_amd/tuning_mi355x.json and, with
--write, the chosen values as the defaults in focus-stacker_amd/csrc/fg_tuning.h (rebuild the library
afterwards).  Run it from the repository root on an MI355X host:  python focus-stacker_amd/autotune.py
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

CANDIDATES = {
    "batch_large_kpix": [0, 1000, 2500, 4000, 12000],
    "batch_streams": [2, 3, 4, 5, 6],
    "batch_group_mib": [0, 32768, 8192, 2048],
    "batch_pass1_blocks_per_cu": [0, 2, 3],
    "batch_wide_level": [64, 128, 256, 512],
    "pass2_min_tiles": [128, 256, 512, 1024, 2048],
}
SINGLE_STACKS = [(30, 7952, 5304), (20, 1920, 1080), (10, 640, 480)]


def _time_batch(fg, torch, ctx, reqs, outs, reps):
    with fg.Batch(ctx, memory_fraction=0.8) as b:
        b.fuse(reqs)                      # builds the plans
        b.fuse(reqs)                      # captures their schedules
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            b.fuse(reqs)
            best = min(best, time.perf_counter() - t0)
    torch.cuda.synchronize()
    return best, [o.clone() for o in outs]


def _time_single(fg, torch, ctx, frames, poses, roi, w, h, steps):
    rw, rh = roi[2] - roi[0], roi[3] - roi[1]
    st = fg.Stack(ctx, len(frames), rw, rh, fg.SAMPLE_U8)
    desc = fg.ImageDesc(w, h, 3, 0, fg.SAMPLE_U8)
    out = torch.zeros((rh, rw, 3), dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()

    def step():
        st.reset()
        for f, p in zip(frames, poses):
            st.add_frame_device(f.data_ptr(), desc, p, roi)
        st.finish_device(out.data_ptr())

    step()
    st.sync()
    st.set_profiling(True)
    for _ in range(steps):
        step()
    ph = st.phase_times()
    st.set_profiling(False)
    st.sync()
    st.close()
    return ph["pass2"][0] / steps, out.clone()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--write", action="store_true", help="write the chosen values into csrc/fg_tuning.h")
    ap.add_argument("--stacks", type=int, default=64)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()
    import torch
    import focus_gpu as fg
    import synth
    import bench
    torch.cuda.init()
    dev = torch.device("cuda:0")
    ctx = fg.Context(0, memory_fraction=0.8)
    best = fg.get_tuning()
    log = {"device": torch.cuda.get_device_name(0), "start": dict(best), "batch": {}, "single": {}}

    # ---- batches ----
    stacks, pools = bench.make_mix(fg, synth, torch, dev, args.stacks)
    outs = [torch.zeros((s["roi"][3] - s["roi"][1], s["roi"][2] - s["roi"][0], 3), dtype=torch.uint8, device=dev)
            for s in stacks]
    torch.cuda.synchronize()
    reqs = [dict(frames=s["frames"], desc=fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8), out=o.data_ptr(),
                 poses=s["poses"], roi=s["roi"]) for s, o in zip(stacks, outs)]
    ref = None
    for knob in ("batch_large_kpix", "batch_streams", "batch_group_mib", "batch_pass1_blocks_per_cu", "batch_wide_level"):
        res = {}
        for v in CANDIDATES[knob]:
            fg.set_tuning(**{**best, knob: v})
            t, got = _time_batch(fg, torch, ctx, reqs, outs, args.reps)
            if ref is None:
                ref = got
            same = all(bool(torch.equal(a, b)) for a, b in zip(ref, got))
            if not same:
                raise SystemExit(f"autotune: {knob}={v} changed the output bits")
            res[v] = round(t * 1000.0, 3)
            print(f"autotune: batch {knob}={v}: {res[v]} ms for {len(stacks)} stacks", flush=True)
        best[knob] = min(res, key=res.get)
        log["batch"][knob] = {str(k): v for k, v in res.items()}
    del pools, outs, ref
    torch.cuda.empty_cache()

    # ---- single stacks ----
    fg.set_tuning(**best)
    totals = {v: 0.0 for v in CANDIDATES["pass2_min_tiles"]}
    for n, w, h in SINGLE_STACKS:
        spec = synth.config_spec(0, stack_id=77, frames=n, width=w, height=h, shift=3.0 * w / 1024.0)
        gen = synth.Stack(spec, device=dev)
        frames = [gen.frame(k).contiguous() for k in range(n)]
        poses = [fg.pose(p) for p in gen.poses]
        del gen
        torch.cuda.synchronize()
        roi = fg.common_overlap_roi([(w, h)] * n, poses)
        ref = None
        res = {}
        for v in CANDIDATES["pass2_min_tiles"]:
            fg.set_tuning(**{**best, "pass2_min_tiles": v})
            t, got = _time_single(fg, torch, ctx, frames, poses, roi, w, h, 5)
            if ref is None:
                ref = got
            if not torch.equal(ref, got):
                raise SystemExit(f"autotune: pass2_min_tiles={v} changed the output bits")
            res[v] = round(t, 4)
            totals[v] += t
            print(f"autotune: single {n} x {w}x{h} pass2_min_tiles={v}: pass 2 {res[v]} ms", flush=True)
        log["single"][f"{n}x{w}x{h}"] = {str(k): v for k, v in res.items()}
        del frames
        torch.cuda.empty_cache()
    best["pass2_min_tiles"] = min(totals, key=totals.get)
    fg.set_tuning(**best)
    log["chosen"] = dict(best)
    with open(os.path.join(HERE, "tuning_mi355x.json"), "w") as f:
        json.dump(log, f, indent=1)
        f.write("\n")
    print(json.dumps({"autotune": "done", "chosen": best}), flush=True)
    if args.write:
        write_header(best)


def write_header(best):
    """The chosen values as the #define defaults of csrc/fg_tuning.h."""
    path = os.path.join(HERE, "csrc", "fg_tuning.h")
    text = open(path).read()
    for knob, macro in (("pass2_min_tiles", "FG_TUNE_PASS2_MIN_TILES"), ("batch_wide_level", "FG_TUNE_BATCH_WIDE_LEVEL"),
                        ("batch_pass1_blocks_per_cu", "FG_TUNE_BATCH_PASS1_BLOCKS_PER_CU"),
                        ("batch_group_mib", "FG_TUNE_BATCH_GROUP_MIB"), ("batch_streams", "FG_TUNE_BATCH_STREAMS"),
                        ("batch_large_kpix", "FG_TUNE_BATCH_LARGE_KPIX")):
        lines = text.split("\n")
        lines = [f"#define {macro} {int(best[knob])}" if ln.startswith(f"#define {macro} ") else ln for ln in lines]
        text = "\n".join(lines)
    with open(path, "w") as f:
        f.write(text)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--write-from-json":
        write_header(json.load(open(os.path.join(HERE, "tuning_mi355x.json")))["chosen"])
    else:
        main()
