"""GPU: batched fusion (fg_batch_fuse) through the C ABI.  Every request's output must equal, bit for
bit, the same stack fused alone with the plain calls (fg_stack_add_frame_device ... finish): the batch
runs the same kernels on the same tiles, with sums in frame order.

Stacks are drawn like the 64-stack uneven workload (sizes 640x480, 1920x1080, 4000x3000 and 6000x4000,
poses of the 1024x768 configuration scaled to the size) with fewer stacks and frames, so that the
tests stay short; the frames come from the seeded generator, made on the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(640, 480), (1920, 1080), (4000, 3000), (6000, 4000)]


@pytest.fixture(scope="module")
def fg():
    import focus_gpu
    return focus_gpu


@pytest.fixture(scope="module")
def ctx(fg):
    c = fg.Context(0, memory_fraction=0.8)
    yield c
    c.close()


def _make(fg, stack_id, n, w, h, dtype="u8", lens=False):
    """Device frames (torch tensors) of one seeded stack, its poses and common-overlap ROI."""
    import torch
    import synth
    spec = synth.config_spec(0, stack_id=stack_id, frames=n, width=w, height=h, shift=3.0 * w / 1024.0,
                             dtype=dtype, lens=lens)
    st = synth.Stack(spec, device=torch.device("cuda:0"))
    frames = []
    for k in range(n):
        f = st.frame(k)
        if dtype == "u16":
            f = (f - (f > 32767).to(torch.int32) * 65536).to(torch.int16)
        frames.append(f.contiguous())
    poses = [fg.pose(p) for p in st.poses]
    roi = fg.common_overlap_roi([(w, h)] * n, poses)
    del st
    return dict(frames=frames, poses=poses, roi=roi, w=w, h=h, st=fg.SAMPLE_U16 if dtype == "u16" else fg.SAMPLE_U8)


def _out(fg, s):
    import torch
    rw, rh = s["roi"][2] - s["roi"][0], s["roi"][3] - s["roi"][1]
    return torch.zeros((rh, rw, 3), dtype=torch.int16 if s["st"] == fg.SAMPLE_U16 else torch.uint8, device="cuda:0")


def _request(fg, s, out, **kw):
    """The library runs on its own streams: torch's work on the frames and outputs (queued on torch's
    stream) has to be finished before they are handed over."""
    import torch
    torch.cuda.synchronize()
    return dict(frames=[f.data_ptr() for f in s["frames"]], desc=fg.ImageDesc(s["w"], s["h"], 3, 0, s["st"]),
                out=out.data_ptr(), poses=s["poses"], roi=s["roi"], **kw)


def _alone(fg, ctx, s, method=None, window=5, max_levels=0):
    import torch
    torch.cuda.synchronize()
    rw, rh = s["roi"][2] - s["roi"][0], s["roi"][3] - s["roi"][1]
    m = fg.METHOD_PYRAMID if method is None else method
    st = fg.Stack(ctx, len(s["frames"]), rw, rh, s["st"], method=m, window=window, max_levels=max_levels)
    d = fg.ImageDesc(s["w"], s["h"], 3, 0, s["st"])
    for f, p in zip(s["frames"], s["poses"]):
        st.add_frame_device(f.data_ptr(), d, p, s["roi"])
    out = st.finish()
    st.close()
    return out


def _host(t):
    import torch
    if t.dtype == torch.int16:
        return t.cpu().numpy().view(np.uint16)
    return t.cpu().numpy()


def _mix(fg, seed, count, max_frames):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w, h = SIZES[int(rng.integers(0, len(SIZES)))]
        n = int(rng.integers(4, max_frames + 1))
        out.append(_make(fg, 500 + seed * 100 + i, n, w, h))
    return out


def test_batch_of_uneven_stacks_equals_each_alone(ctx, fg):
    stacks = _mix(fg, 1, 8, 10)
    print("batch:", [(s["w"], s["h"], len(s["frames"])) for s in stacks], flush=True)
    outs = [_out(fg, s) for s in stacks]
    with fg.Batch(ctx) as b:
        status = b.fuse([_request(fg, s, o) for s, o in zip(stacks, outs)])
    assert all(st["state"] == fg.REQ_DONE for st in status)
    for i, (s, o) in enumerate(zip(stacks, outs)):
        ref = _alone(fg, ctx, s)
        got = _host(o)
        assert np.array_equal(got, ref), f"stack {i}: {int((got != ref).sum())} samples differ"


def test_batch_replays_its_graph_with_other_frames(ctx, fg):
    """The second run of a group geometry captures its schedule and the third replays it; only the
    frame descriptors change.  Frames reversed and mirrored give a different stack of the same shape."""
    import torch
    stacks = _mix(fg, 2, 5, 8)
    outs = [_out(fg, s) for s in stacks]
    with fg.Batch(ctx) as b:
        b.fuse([_request(fg, s, o) for s, o in zip(stacks, outs)])
        first = [_host(o).copy() for o in outs]
        flipped = []
        for s in stacks:
            t = dict(s)
            t["frames"] = [torch.flip(f, dims=[1]).contiguous() for f in reversed(s["frames"])]
            flipped.append(t)
        b.fuse([_request(fg, s, o) for s, o in zip(flipped, outs)])
        second = [_host(o).copy() for o in outs]
        b.fuse([_request(fg, s, o) for s, o in zip(stacks, outs)])
        third = [_host(o).copy() for o in outs]
    for i in range(len(stacks)):
        assert not np.array_equal(first[i], second[i])
        assert np.array_equal(second[i], _alone(fg, ctx, flipped[i])), f"replay {i} differs from the stack alone"
        assert np.array_equal(third[i], first[i])
        assert np.array_equal(first[i], _alone(fg, ctx, stacks[i]))


def test_batch_recaptures_when_lens_frames_come_and_go(ctx, fg):
    """A batch run's pass-1 kernel carries the lens-path remap only when the run has lens frames, so a
    captured schedule is captured again when that changes.  One geometry (same sizes, same ROI) runs
    with homography poses, then with the same poses plus a centre shift and radial terms (the lens
    path), and so on: runs 2-5 each flip the flag against the graph of the run before.  Every run
    must equal the stacks fused alone under the same poses."""
    stacks = _mix(fg, 3, 3, 6)
    lensed = []
    for i, s in enumerate(stacks):
        t = dict(s)
        t["poses"] = [fg.pose({f: getattr(p, f) for f, _ in fg.Pose._fields_}, center_shift_x=0.3 + 0.1 * k,
                              center_shift_y=-0.2, radial_a=2e-4 * (i + 1)) for k, p in enumerate(s["poses"])]
        lensed.append(t)
    outs = [_out(fg, s) for s in stacks]
    ref_h = [_alone(fg, ctx, s) for s in stacks]
    ref_l = [_alone(fg, ctx, s) for s in lensed]
    for i in range(len(stacks)):
        assert not np.array_equal(ref_h[i], ref_l[i])
    with fg.Batch(ctx) as b:
        for run, (group, refs) in enumerate([(stacks, ref_h), (lensed, ref_l), (stacks, ref_h), (lensed, ref_l),
                                             (stacks, ref_h)]):
            status = b.fuse([_request(fg, s, o) for s, o in zip(group, outs)])
            assert all(st["state"] == fg.REQ_DONE for st in status)
            for i, o in enumerate(outs):
                got = _host(o)
                assert np.array_equal(got, refs[i]), f"run {run}, stack {i}: {int((got != refs[i]).sum())} differ"


def test_batch_groups_methods_types_and_windows(ctx, fg):
    """A hybrid request (run through the plain calls), a 16-bit stack, a window-7 stack, a capped
    pyramid and lens poses, interleaved with 8-bit window-5 stacks: groups split where the class
    changes, and every output equals its stack alone."""
    reqs = [
        (_make(fg, 701, 4, 640, 480), {}),
        (_make(fg, 702, 5, 1920, 1080), dict(method=fg.METHOD_HYBRID)),
        (_make(fg, 703, 4, 640, 480, dtype="u16"), {}),
        (_make(fg, 704, 3, 1920, 1080, dtype="u16", lens=True), {}),
        (_make(fg, 705, 4, 640, 480), dict(window=7)),
        (_make(fg, 706, 6, 1920, 1080), dict(max_levels=6)),
        (_make(fg, 707, 4, 640, 480, lens=True), {}),
    ]
    outs = [_out(fg, s) for s, _ in reqs]
    with fg.Batch(ctx) as b:
        status = b.fuse([_request(fg, s, o, **kw) for (s, kw), o in zip(reqs, outs)])
    assert all(st["state"] == fg.REQ_DONE for st in status)
    for i, ((s, kw), o) in enumerate(zip(reqs, outs)):
        ref = _alone(fg, ctx, s, method=kw.get("method"), window=kw.get("window", 5),
                     max_levels=kw.get("max_levels", 0))
        assert np.array_equal(_host(o), ref), f"request {i} ({kw}) differs from the stack alone"


def test_batch_under_a_small_budget(ctx, fg):
    """A budget that holds about one large stack: many groups, plans evicted between them."""
    stacks = _mix(fg, 3, 6, 6)
    outs = [_out(fg, s) for s in stacks]
    with fg.Batch(ctx, memory_fraction=0.01, plan_cache=2) as b:
        status = b.fuse([_request(fg, s, o) for s, o in zip(stacks, outs)])
    assert all(st["state"] == fg.REQ_DONE for st in status)
    for s, o in zip(stacks, outs):
        assert np.array_equal(_host(o), _alone(fg, ctx, s))


def test_batch_of_tiny_stacks(ctx, fg):
    """Levels of one pixel, a single level (1x1), images smaller than the contrast window."""
    import torch
    rng = np.random.default_rng(9)
    cases = [(1, 1, 3), (2, 3, 2), (5, 5, 4), (17, 3, 3), (40, 150, 3)]
    stacks, outs = [], []
    for w, h, n in cases:
        frames = [torch.from_numpy(rng.integers(0, 256, (h, w, 3)).astype(np.uint8)).cuda() for _ in range(n)]
        s = dict(frames=frames, poses=None, roi=None, w=w, h=h, st=fg.SAMPLE_U8)
        stacks.append(s)
        outs.append(torch.zeros((h, w, 3), dtype=torch.uint8, device="cuda:0"))
    torch.cuda.synchronize()
    with fg.Batch(ctx) as b:
        b.fuse([dict(frames=[f.data_ptr() for f in s["frames"]], desc=fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8),
                     out=o.data_ptr()) for s, o in zip(stacks, outs)])
    for s, o in zip(stacks, outs):
        st = fg.Stack(ctx, len(s["frames"]), s["w"], s["h"])
        for f in s["frames"]:
            st.add_frame_device(f.data_ptr(), fg.ImageDesc(s["w"], s["h"], 3, 0, fg.SAMPLE_U8))
        ref = st.finish()
        st.close()
        assert np.array_equal(_host(o), ref), f"{s['w']}x{s['h']}"


def test_batch_rejects_a_bad_request_and_runs_the_rest(ctx, fg):
    stacks = _mix(fg, 4, 2, 4)
    outs = [_out(fg, s) for s in stacks]
    reqs = [_request(fg, s, o) for s, o in zip(stacks, outs)]
    bad = dict(reqs[0])
    bad["desc"] = fg.ImageDesc(stacks[0]["w"], stacks[0]["h"], 3, 1, stacks[0]["st"])  # alpha on a posed frame
    with fg.Batch(ctx) as b:
        with pytest.raises(fg.FgError):
            b.fuse([reqs[0], bad, reqs[1]])
    for s, o in zip(stacks, outs):
        assert np.array_equal(_host(o), _alone(fg, ctx, s))


@pytest.mark.parametrize("knobs", [dict(batch_large_kpix=0, batch_streams=1),
                                   dict(batch_large_kpix=1, batch_streams=4, batch_group_mib=64)],
                         ids=["all-grouped-one-stream", "all-alone-four-streams"])
def test_batch_outputs_do_not_depend_on_tuning(ctx, fg, knobs):
    """Every stack in groups on one stream, or every stack on its own over three side streams (with
    small group caps): the same bits as each stack fused alone."""
    stacks = _mix(fg, 5, 6, 6)
    outs = [_out(fg, s) for s in stacks]
    saved = fg.get_tuning()
    try:
        fg.set_tuning(**{**saved, **knobs})
        with fg.Batch(ctx) as b:
            status = b.fuse([_request(fg, s, o) for s, o in zip(stacks, outs)])
    finally:
        fg.set_tuning(**saved)
    assert all(st["state"] == fg.REQ_DONE for st in status)
    for i, (s, o) in enumerate(zip(stacks, outs)):
        assert np.array_equal(_host(o), _alone(fg, ctx, s)), f"stack {i} differs from the stack alone"
