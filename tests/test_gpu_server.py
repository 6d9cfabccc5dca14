"""GPU: the stack server (fg_server_*) against the plain per-stack calls.

Continuous batching must not change any result: every request computes exactly what the same
stack computes alone, so the outputs are compared bit for bit, with generous and with starved
memory budgets (admission then runs the stacks one at a time, or uncached).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import focus_gpu
    return focus_gpu


@pytest.fixture(scope="module")
def ctx(fg):
    c = fg.Context(0)
    yield c
    c.close()


# uneven stacks: (frames, width, height, method name, window)
MIX = [
    (3, 160, 120, "pyramid", 5),
    (5, 97, 71, "pyramid", 5),
    (2, 210, 90, "hybrid", 5),
    (4, 130, 100, "softmask", 3),
    (6, 64, 48, "pyramid", 7),
    (3, 150, 110, "hardmask", 5),
    (4, 160, 120, "pyramid", 5),   # same geometry as the first: reuses a pooled stack
]


def _method(fg, name):
    return {"pyramid": fg.METHOD_PYRAMID, "hybrid": fg.METHOD_HYBRID, "softmask": fg.METHOD_SOFTMASK,
            "hardmask": fg.METHOD_HARDMASK}[name]


@pytest.fixture(scope="module")
def stacks(fg, ctx):
    import synth
    out = []
    for i, (n, w, h, m, win) in enumerate(MIX):
        spec = synth.config_spec(0, stack_id=i, frames=n, width=w, height=h)
        st = synth.Stack(spec)
        frames = [ctx.upload(st.frame_numpy(k)) for k in range(n)]
        poses = [fg.pose(p) for p in st.poses]
        roi = fg.common_overlap_roi([(w, h)] * n, poses)
        out.append(dict(frames=frames, poses=poses, roi=roi, w=w, h=h, method=_method(fg, m), window=win))
    yield out
    for s in out:
        for f in s["frames"]:
            ctx.free(f)


def _alone(fg, ctx, s):
    rw, rh = s["roi"][2] - s["roi"][0], s["roi"][3] - s["roi"][1]
    st = fg.Stack(ctx, len(s["frames"]), rw, rh, fg.SAMPLE_U8, method=s["method"], window=s["window"])
    desc = fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8)
    for f, p in zip(s["frames"], s["poses"]):
        st.add_frame_device(f, desc, p, s["roi"])
    if st.needs_pass2_frames():
        for f, p in zip(s["frames"], s["poses"]):
            st.pass2_frame_device(f, desc, p, s["roi"])
    out = ctx.alloc(rw * rh * 3)
    st.finish_device(out)
    st.sync()
    st.close()
    res = ctx.download(out, (rh, rw, 3))
    ctx.free(out)
    return res


def _served(fg, ctx, stacks, **server_kw):
    outs, tickets = [], []
    with fg.Server(ctx, **server_kw) as srv:
        for s in stacks:
            rw, rh = s["roi"][2] - s["roi"][0], s["roi"][3] - s["roi"][1]
            o = ctx.alloc(rw * rh * 3)
            outs.append((o, (rh, rw, 3)))
            desc = fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8)
            tickets.append(srv.submit(s["frames"], desc, o, poses=s["poses"], roi=s["roi"], method=s["method"],
                                      window=s["window"]))
        status = [srv.wait(t) for t in tickets]
    res = [ctx.download(o, shp) for o, shp in outs]
    for o, _ in outs:
        ctx.free(o)
    return res, status


@pytest.mark.parametrize("server_kw", [dict(slots=3), dict(slots=4, pool_per_slot=-1), dict(slots=2, memory_fraction=1e-6)],
                         ids=["slots3", "nopool", "starved"])
def test_server_results_equal_alone(fg, ctx, stacks, server_kw):
    alone = [_alone(fg, ctx, s) for s in stacks]
    served, status = _served(fg, ctx, stacks, **server_kw)
    for i, (a, b, st) in enumerate(zip(alone, served, status)):
        assert st["state"] == fg.REQ_DONE, (i, st)
        assert st["total_ms"] >= st["run_ms"] >= 0.0
        assert np.array_equal(a, b), f"stack {i} differs in the mix ({server_kw})"


def test_server_rejects_bad_requests(fg, ctx, stacks):
    with fg.Server(ctx, slots=2) as srv:
        s = stacks[0]
        desc = fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8)
        with pytest.raises(fg.FgError):
            srv.submit(s["frames"], desc, 0, poses=s["poses"], roi=s["roi"])
        with pytest.raises(fg.FgError):
            srv.submit(s["frames"], desc, 1 << 20, poses=s["poses"], roi=s["roi"], window=4)
        with pytest.raises(fg.FgError):
            srv.wait(12345)


@pytest.fixture(scope="module")
def full_size(fg):
    """A 104-frame stack of 7952x5304 u8 (the 30-frame workload's generator and pose distribution):
    eight generated frames, taken in turn under 104 seeded poses, so that the frames stay few."""
    import torch
    import synth
    n, distinct = 104, 8
    spec = synth.config_spec(1, stack_id=41, frames=n)
    poses = [fg.pose(p) for p in synth.make_poses(spec)]
    gen = synth.Stack(synth.config_spec(1, stack_id=41, frames=distinct), device=torch.device("cuda:0"))
    pool = [gen.frame(k).contiguous() for k in range(distinct)]
    del gen
    torch.cuda.synchronize()
    w, h = spec.width, spec.height
    roi = fg.common_overlap_roi([(w, h)] * n, poses)
    yield dict(frames=[pool[k % distinct].data_ptr() for k in range(n)], poses=poses, roi=roi, w=w, h=h,
               method=fg.METHOD_PYRAMID, window=5, pool=pool)
    del pool
    torch.cuda.empty_cache()


@pytest.mark.parametrize("fraction", [0.8, 0.06], ids=["budget-0.8", "budget-0.06"])
def test_server_full_size_stack_beside_small_ones(fg, ctx, stacks, full_size, fraction):
    """A 104-frame 7952x5304 request admitted between small ones: with 0.8 of the free memory it runs
    cached beside them; with 0.06 (less than its 34 GB frame cache) it is admitted uncached, its
    second pass remapping the frames again.  Every result equals its stack fused alone."""
    mix = stacks[:3] + [full_size] + stacks[3:]
    print("server mix: full-size", full_size["w"], "x", full_size["h"], len(full_size["frames"]), "frames,",
          "budget", fraction, flush=True)
    served, status = _served(fg, ctx, mix, slots=4, memory_fraction=fraction)
    for i, (s, b, st) in enumerate(zip(mix, served, status)):
        assert st["state"] == fg.REQ_DONE, (i, st)
        a = _alone(fg, ctx, s)
        assert np.array_equal(a, b), f"request {i} ({s['w']}x{s['h']}, {len(s['frames'])} frames) differs in the mix"
