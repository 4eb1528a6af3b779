"""HIP path vs the oracle (bit-exact), through the C ABI.

Every test here runs the product library on the GPU and compares it with the
oracle's C restatement on the same seeded inputs."""
import ctypes

import numpy as np
import pytest

C_int64 = ctypes.c_int64

from oracle.oracle import Oracle, MASK_ALL, MASK_16x16, SHAPES
import uavs3e_amd as ua
from uavs3e_amd import workloads as wl

pytestmark = pytest.mark.gpu

SATD_SLOTS = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 1), (1, 2)]


@pytest.fixture(scope="module", params=[8, 10])
def libs(request):
    bd = request.param
    return bd, ua.MeHip(bd), Oracle(bd)


def test_block_costs(libs):
    bd, hip, orc = libs
    rng = np.random.default_rng(10 + bd)
    S = 320
    a = rng.integers(0, 1 << bd, (300, S)).astype(hip.dt)
    b = rng.integers(0, 1 << bd, (300, S)).astype(hip.dt)
    for (w, h) in SHAPES + [(4, 4)]:
        oo = int(rng.integers(0, 100)) * S + int(rng.integers(0, 100))
        po = int(rng.integers(0, 100)) * S + int(rng.integers(0, 100))
        assert hip.sad(w, h, a, S, b, S, oo, po) == orc.sad(w, h, a, S, b, S, oo, po), (w, h)
        assert hip.had(w, h, a, S, b, S, oo, po) == orc.had(w, h, a, S, b, S, oo, po), (w, h)
        offs = [po + k for k in range(4)]
        x4 = hip.sad_xn(w, h, a, S, oo, b, S, offs)
        x3 = hip.sad_xn(w, h, a, S, oo, b, S, offs[:3])
        ref = [orc.sad(w, h, a, S, b, S, oo, o) for o in offs]
        assert list(x4) == ref and list(x3) == ref[:3], (w, h)
    for wl_, hl in SATD_SLOTS:
        for t in range(20):
            oo = int(rng.integers(0, 200)) * S + int(rng.integers(0, 200))
            po = int(rng.integers(0, 200)) * S + int(rng.integers(0, 200))
            assert hip.satd(wl_, hl, a, S, b, S, oo, po) == orc.satd(wl_, hl, a, S, b, S, oo, po)
    assert hip.error_count() == 0


def test_filters(libs):
    bd, hip, orc = libs
    rng = np.random.default_rng(20 + bd)
    S = 320
    src = rng.integers(0, 1 << bd, (200, S)).astype(hip.dt)
    maxv = (1 << bd) - 1
    off = 10 * S + 10
    for typ in range(4):
        n = 8 if typ in (0, 2) else 4
        for w in (4, 8, 16, 32, 64) + ((128,) if n == 8 else ()):
            for frac in range(1, 4 if n == 8 else 8):
                c = orc.luma_taps(frac) if n == 8 else orc.chroma_taps(frac)
                h = int(rng.choice([4, 8, 16]))
                assert np.array_equal(hip.ipflt(typ, src, S, off, w, h, c, maxv), orc.ipflt(typ, src, S, off, w, h, c, maxv))
    for typ in range(2):
        n = 8 if typ == 0 else 4
        for w in (4, 8, 16, 32, 64) + ((128,) if n == 8 else ()):
            for fx in range(1, 4 if n == 8 else 8):
                fy = int(rng.integers(1, 4 if n == 8 else 8))
                cx = orc.luma_taps(fx) if n == 8 else orc.chroma_taps(fx)
                cy = orc.luma_taps(fy) if n == 8 else orc.chroma_taps(fy)
                h = int(rng.choice([4, 8, 16]))
                assert np.array_equal(hip.ipflt_ext(typ, src, S, off, w, h, cx, cy, maxv),
                                      orc.ipflt_ext(typ, src, S, off, w, h, cx, cy, maxv))
    for w in (4, 8, 16, 32, 64, 128):
        s1 = rng.integers(0, 1 << bd, (8, w)).astype(hip.dt)
        s2 = rng.integers(0, 1 << bd, (8, w)).astype(hip.dt)
        assert np.array_equal(hip.pel_avrg(w, 8, s1, s2), orc.pel_avrg(w, 8, s1, s2))
        assert np.array_equal(hip.ipcpy(src, S, off, w, 8), src[10:18, 10:10 + w])
    assert hip.error_count() == 0


# widths 202, 73 and 75 leave 2, 1 and 3 samples past the last whole dword
# of a row: per-sample stores, and right-side padding from an unaligned column
@pytest.mark.parametrize("size", [(72, 40), (200, 136), (202, 40), (73, 41), (75, 36)])
def test_planes_device_and_table(libs, size):
    """Device reference planes and com_if_luma_frame through the per-call
    frame filters both equal the oracle's 16 padded planes."""
    bd, hip, orc = libs
    w, h = size
    host = ua.MeHost(bd)
    luma = np.random.default_rng(30 + bd + w).integers(0, 1 << bd, (h, w)).astype(hip.dt)
    want = orc.make_planes(luma)
    ref = hip.ref_create(luma)
    got = hip.ref_planes(ref, w, h)
    hip.ref_destroy(ref)
    for i in range(16):
        assert np.array_equal(got[i], want[i]), "device plane %d" % i
    t = hip.dispatch_table()
    via = host.planes_via_table(t, luma)
    for i in range(16):
        assert np.array_equal(via[i], want[i]), "table plane %d" % i
    assert hip.error_count() == 0


def _run_plan(hip, cur, refs, R, lam, mask, grid_out=None, **kw):
    h, w = cur.shape
    dref = [hip.ref_create(r) for r in refs]
    plan = hip.plan_create(w, h, R, lam, mask, len(refs), **kw)
    if grid_out is not None:
        grid_out.append(hip.plan_int_grid(plan))
    hip.plan_set_cur(plan, cur)
    hip.plan_run(plan, dref)
    res, best = hip.plan_results(plan, len(refs))
    hip.plan_destroy(plan)
    for r in dref:
        hip.ref_destroy(r)
    return res, best


def _compare(res, best, ores, obest):
    msg = ua.compare_results(res, best, ores, obest)
    assert msg is None, msg


CASES = [
    # (bd, w, h, R, mask, nref)
    (8, 384, 264, 8, MASK_ALL, 2),
    (8, 384, 264, 40, MASK_16x16, 1),
    (8, 136, 136, 128, 1 << SHAPES.index((128, 128)), 1),
    (10, 384, 264, 8, MASK_ALL, 2),
    (10, 200, 136, 24, MASK_16x16, 2),
    # a width that is not a multiple of 4: rows of the 8-bit planes are not
    # dword-aligned to each other (the sub-pel kernel's general path)
    (8, 202, 136, 8, MASK_ALL, 1),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "bd%d_%dx%d_R%d_m%x_n%d" % c)
def test_frame_plan_vs_oracle(case):
    bd, w, h, R, mask, nref = case
    base = wl.hd_allpu() if bd == 8 else wl.uhd10_allpu()
    work = base.scaled(w, h, R)
    work.ref_shifts_qpel = work.ref_shifts_qpel[:nref]
    cur, refs = work.frames()
    hip, orc = ua.MeHip(bd), Oracle(bd)
    lam = wl.lambda_mv(32)
    res, best = _run_plan(hip, cur, refs, R, lam, mask)
    ores, obest = orc.me_frame(cur, [orc.make_planes(r) for r in refs], R, lam, mask)
    _compare(res, best, ores, obest)
    assert res["valid"].sum() > 0


def _spoil_planes(hip, ref, w, h):
    """Overwrites the 16 padded planes of a device reference (not its luma)
    through the HIP runtime the library itself links, so that only a rebuild
    inside the frame can restore them."""
    plane, stride, pad = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
    fill = hip.lib.hipMemset
    fill.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    for i in range(16):
        assert hip.lib.me_ref_plane(ref, i, ctypes.byref(plane), ctypes.byref(stride), ctypes.byref(pad)) == 0
        item = np.dtype(hip.dt).itemsize
        base = plane.value - (pad.value * stride.value + pad.value) * item
        assert fill(base, 0x5A, stride.value * (h + 2 * pad.value) * item) == 0


@pytest.mark.parametrize("bd,nref,mask", [(8, 2, 1), (8, 3, 3), (10, 2, 1), (8, 2, 3), (8, 2, 2)])
def test_frame_plan_rebuilds_references(bd, nref, mask):
    """References named in the prepare mask are rebuilt inside the frame
    graph before the search, for any mask and on a capture as on a replay:
    spoiled planes of those references must not reach the results."""
    base = wl.hd_allpu() if bd == 8 else wl.uhd10_allpu()
    w, h, R = 256, 200, 12
    work = base.scaled(w, h, R)
    work.ref_shifts_qpel = (work.ref_shifts_qpel * 2)[:nref]
    cur, refs = work.frames()
    hip, orc = ua.MeHip(bd), Oracle(bd)
    lam = wl.lambda_mv(32)
    dref = [hip.ref_create(r) for r in refs]
    plan = hip.plan_create(w, h, R, lam, MASK_ALL, nref)
    hip.plan_set_cur(plan, cur)
    ores, obest = orc.me_frame(cur, [orc.make_planes(r) for r in refs], R, lam, MASK_ALL)
    for _ in range(2):                       # a capture, then a replay of the same graph
        for r in range(nref):
            if mask >> r & 1:
                _spoil_planes(hip, dref[r], w, h)
        hip.plan_run(plan, dref, mask)
        res, best = hip.plan_results(plan, nref)
        _compare(res, best, ores, obest)
    # the same handle twice
    hip.plan_run(plan, [dref[0]] * nref, 1)
    res, best = hip.plan_results(plan, nref)
    o2, b2 = orc.me_frame(cur, [orc.make_planes(refs[0])] * nref, R, lam, MASK_ALL)
    _compare(res, best, o2, b2)
    hip.plan_destroy(plan)
    for r in dref:
        hip.ref_destroy(r)
    assert hip.error_count() == 0


@pytest.mark.parametrize("bd", [8, 10])
def test_plan_graph_follows_other_entry_points(bd):
    """The other entry points of a plan point its picture descriptor at their
    own references; a later me_plan_run with the references of its cached
    graph must search those again.  And a reference destroyed and created
    anew (whose handle address may come back) must not replay a graph that
    holds the old one's buffers."""
    base = wl.hd_allpu() if bd == 8 else wl.uhd10_allpu()
    w, h, R = 200, 136, 8
    work = base.scaled(w, h, R)
    work.ref_shifts_qpel = work.ref_shifts_qpel[:2]
    cur, refs = work.frames()
    other = [np.ascontiguousarray(np.flipud(r)) for r in refs]
    hip, orc = ua.MeHip(bd), Oracle(bd)
    lam = wl.lambda_mv(32)
    mask = sum(1 << SHAPES.index(s) for s in [(8, 8), (16, 16), (32, 32)])
    want = orc.me_frame(cur, [orc.make_planes(r) for r in refs], R, lam, mask)
    want_b = orc.me_frame(cur, [orc.make_planes(r) for r in other], R, lam, mask)
    da = [hip.ref_create(r) for r in refs]
    db = [hip.ref_create(r) for r in other]
    plan = hip.plan_create(w, h, R, lam, mask, 2)
    hip.plan_set_cur(plan, cur)
    hip.plan_run(plan, da)
    _compare(*hip.plan_results(plan, 2), *want)
    npu = hip.lib.me_plan_num_pus(plan)
    tz = ua.TzCfg(8, 64, 1, 0, 0, 8, (C_int64 * 4)(7, 4, 0, 0))
    hip.plan_run_tz(plan, db, np.zeros((hip.lib.me_plan_num_ctus(plan), npu, 2, 2), np.int16), tz)
    hip.plan_run(plan, da)                     # the cached graph, descriptor back at da
    _compare(*hip.plan_results(plan, 2), *want)
    hip.plan_run(plan, db)
    _compare(*hip.plan_results(plan, 2), *want_b)
    for r in da:
        hip.ref_destroy(r)
    dc = [hip.ref_create(r) for r in refs]     # new buffers, possibly at da's addresses
    for _ in range(2):
        hip.plan_run(plan, dc, 1)              # rebuilds reference 0 inside the graph
        _compare(*hip.plan_results(plan, 2), *want)
    hip.plan_destroy(plan)
    for r in db + dc:
        hip.ref_destroy(r)
    assert hip.error_count() == 0


@pytest.mark.parametrize("bd,nref", [(8, 2), (10, 1)])
def test_frame_plan_high_contrast(bd, nref):
    """Independent uniform-noise frames: mean |difference| of about a third of
    the sample range, so a 128x128 SAD exceeds 2^20 (8-bit) and 2^22
    (10-bit) at every displacement -- the largest costs the keys must order."""
    rng = np.random.default_rng(7 + bd)
    w, h, R = 256, 256, 16
    dt = np.uint8 if bd == 8 else np.uint16
    cur = rng.integers(0, 1 << bd, (h, w)).astype(dt)
    refs = [rng.integers(0, 1 << bd, (h, w)).astype(dt) for _ in range(nref)]
    hip, orc = ua.MeHip(bd), Oracle(bd)
    lam = wl.lambda_mv(32)
    res, best = _run_plan(hip, cur, refs, R, lam, MASK_ALL)
    ores, obest = orc.me_frame(cur, [orc.make_planes(r) for r in refs], R, lam, MASK_ALL)
    _compare(res, best, ores, obest)
    big = ores["cost_int"][:, -1, 0] >> (16 - (bd - 8))   # 128x128 SADs of the four CTUs
    assert (big > (1 << (20 if bd == 8 else 22))).all()


def test_plan_tuning_variants_agree():
    """Band heights, waves per block, sub-pel position groups and the
    integer-search grid change only the schedule.  Persistent grids of a few
    workgroups split segments (CTU, reference, column tile) into 1-4 pieces
    at every band position, each piece storing its own key slot."""
    work = wl.hd_allpu().scaled(256, 200, 12)
    cur, refs = work.frames()
    hip = ua.MeHip(8)
    lam = wl.lambda_mv(32)
    base = _run_plan(hip, cur, refs, 12, lam, MASK_ALL, int_grid=1)
    for band, waves, cols, g, grid in ((1, 1, 0, 0, 1), (5, 2, 8, 2, 1), (25, 4, 12, 4, 0), (3, 4, 4, 8, 1),
                                       (11, 8, 0, 1, 0), (16, 8, 8, 8, 1), (1, 4, 0, 0, 7), (5, 2, 8, 2, 5),
                                       (3, 8, 4, 1, 13), (2, 4, 0, 4, 2), (4, 1, 12, 0, 9), (25, 8, 0, 0, 3)):
        got = []
        res, best = _run_plan(hip, cur, refs, 12, lam, MASK_ALL, grid_out=got, band_rows=band, waves=waves,
                              dx_cols=cols, sp_groups=g, int_grid=grid)
        groups, slots = got[0]
        if grid >= 2:   # persistent on `grid` workgroups
            assert groups == grid and 1 <= slots <= 4, (grid, groups, slots)
        elif grid == 1:
            assert slots == 0
        assert np.array_equal(res, base[0]) and np.array_equal(best, base[1]), (band, waves, cols, g, grid)


def test_host_driver_percall_vs_oracle():
    """The per-call drop-in path: the host driver calling HIP table entries."""
    work = wl.hd_16x16().scaled(144, 136, 3)
    cur, refs = work.frames()
    hip, orc, host = ua.MeHip(8), Oracle(8), ua.MeHost(8)
    t = hip.dispatch_table()
    planes = [host.planes_via_table(t, r) for r in refs]
    lam = wl.lambda_mv(32)
    res, best, ms = host.me_frame(t, cur, planes, 3, lam, MASK_16x16, 0, 1, threads=2)
    ores, obest = orc.me_frame(cur, [orc.make_planes(r) for r in refs], 3, lam, MASK_16x16, 0, 1)
    _compare(res[:1], best[:1], ores[:1], obest[:1])
    assert hip.error_count() == 0


@pytest.mark.parametrize("bd", [8, 10])
def test_bi_plan_and_mc_vs_oracle(bd):
    """Two reference lists (2 past, 2 future pictures): per-list final costs
    and best references, then motion compensation of every PU -- luma from
    the quarter-pel planes, 1/8-pel chroma, rounding bi-average -- bit-exact
    against the oracle."""
    work = wl.fuhd10_bi().scaled(264, 200, 6)
    if bd == 8:
        import dataclasses
        work = dataclasses.replace(work, bd=8, sigma_ref=3.0, sigma_cur=3.0)
    cur, refs = work.frames()
    chroma = work.ref_chroma()
    hip, orc = ua.MeHip(bd), Oracle(bd)
    lam = wl.lambda_mv(32)
    dref = [hip.ref_create(r) for r in refs]
    for d, (u, v) in zip(dref, chroma):
        hip.ref_load_chroma(d, u, v)
    plan = hip.plan_create(work.w, work.h, work.R, lam, MASK_ALL, 4, num_refs_l1=2)
    hip.plan_set_cur(plan, cur)
    hip.plan_run(plan, dref)
    hip.plan_mc(plan, dref)
    res, best = hip.plan_results(plan, 4)
    py, pu, pv = hip.plan_mc_download(plan)
    hip.plan_destroy(plan)
    for r in dref:
        hip.ref_destroy(r)
    planes = [orc.make_planes(r) for r in refs]
    ores, obest = orc.me_frame(cur, planes, work.R, lam, MASK_ALL, num_refs_l1=2)
    _compare(res, best, ores, obest)
    assert best.shape[-1] == 2 and (best[..., 0] < 2).all() and ((best[..., 1] >= 2) | (best[..., 1] < 0)).all()
    oy, ou, ov = orc.mc_frame(work.w, work.h, planes, [(orc.make_chroma(u), orc.make_chroma(v)) for u, v in chroma],
                              2, MASK_ALL, ores, obest)
    assert np.array_equal(py, oy), "luma predictions differ at %d samples" % int((py != oy).sum())
    assert np.array_equal(pu, ou) and np.array_equal(pv, ov), "chroma predictions differ"
    assert py.any() and pu.any()


def test_autotune_picks_fastest_and_keeps_results():
    """The autotuner times every schedule that fits and returns the fastest;
    the tuned plan's results equal the default plan's (and the oracle's)."""
    work = wl.hd_allpu().scaled(392, 264, 10)
    cur, refs = work.frames()
    hip, orc = ua.MeHip(8), Oracle(8)
    lam = wl.lambda_mv(32)
    dref = [hip.ref_create(r) for r in refs]
    pick, table = hip.autotune(work.w, work.h, work.R, lam, MASK_ALL, dref, cur, iters=2)
    assert all(t["int_search_ms"] > 0 and t["subpel_ms"] > 0 for t in table)
    it = [t for t in table if t["sp_groups"] == 0]   # integer-search schedules, default groups
    sp = [t for t in table if t["sp_groups"] > 0]    # then the sub-pel position groups
    assert len(it) >= 6 and sorted(t["sp_groups"] for t in sp) == [1, 2, 4, 8]
    fastest = min(it, key=lambda t: t["int_search_ms"])
    assert {k: pick[k] for k in ("band_rows", "waves", "dx_cols")} == {k: fastest[k] for k in ("band_rows", "waves", "dx_cols")}
    assert all({k: t[k] for k in ("band_rows", "waves", "dx_cols")} == {k: fastest[k] for k in ("band_rows", "waves", "dx_cols")}
               for t in sp)
    assert pick["sp_groups"] == min(sp, key=lambda t: t["subpel_ms"])["sp_groups"]
    for r in dref:
        hip.ref_destroy(r)
    res, best = _run_plan(hip, cur, refs, work.R, lam, MASK_ALL, **pick)
    ores, obest = orc.me_frame(cur, [orc.make_planes(r) for r in refs], work.R, lam, MASK_ALL)
    _compare(res, best, ores, obest)


TZ_MASK = sum(1 << SHAPES.index(s) for s in [(4, 8), (8, 16), (16, 8), (32, 4), (32, 32), (64, 128), (128, 128)])


@pytest.mark.parametrize("bd,mvr,cost_type", [(8, 0, 0), (8, 1, 1), (8, 2, 0), (8, 4, 1), (10, 0, 0), (10, 3, 1)])
def test_search_tz_vs_oracle(bd, mvr, cost_type):
    """me_search_tz (diamond, raster, refinement, sub-pel, final cost) for
    every PU from random MVPs, against the oracle's restatement (itself
    pinned by the compiled reference's me_search_tz golden vectors)."""
    base = wl.hd_allpu() if bd == 8 else wl.uhd10_allpu()
    work = base.scaled(200, 136, 8)
    work.ref_shifts_qpel = work.ref_shifts_qpel[:2]
    cur, refs = work.frames()
    hip, orc = ua.MeHip(bd), Oracle(bd)
    lam = wl.lambda_mv(32)
    n_ctu, nref = 4, 2
    npu = sum((128 // w) * (128 // h) for i, (w, h) in enumerate(SHAPES) if TZ_MASK >> i & 1)
    rng = np.random.default_rng(100 * bd + 10 * mvr + cost_type)
    mvp = rng.integers(-60, 61, (n_ctu, npu, nref, 2)).astype(np.int16)
    tz = ua.TzCfg(8, 64, 1, mvr, cost_type, 8, (C_int64 * 4)(7, 4, 0, 0))
    dref = [hip.ref_create(r) for r in refs]
    plan = hip.plan_create(work.w, work.h, work.R, lam, TZ_MASK, nref)
    hip.plan_set_cur(plan, cur)
    hip.plan_run_tz(plan, dref, mvp, tz)
    res, best = hip.plan_results(plan, nref)
    hip.plan_destroy(plan)
    for r in dref:
        hip.ref_destroy(r)
    planes = [orc.make_planes(r) for r in refs]
    pocs = [7, 4]
    checked = 0
    for ctu in range(n_ctu):
        cx, cy = (ctu % 2) * 128, (ctu // 2) * 128
        pu = 0
        for si, (w, h) in enumerate(SHAPES):
            if not TZ_MASK >> si & 1:
                continue
            for oy in range(0, 128, h):
                for ox in range(0, 128, w):
                    x, y = cx + ox, cy + oy
                    for r in range(nref):
                        g = res[ctu, pu, r]
                        if x + w > work.w or y + h > work.h:
                            assert g["valid"] == 0
                            continue
                        (mx, my, ix, iy), rc = orc.me_search_tz(cur, planes[r], x, y, w, h, r, nref, 8, pocs[r], 8, 64,
                                                                mvr, cost_type, 1, lam, tuple(mvp[ctu, pu, r]))
                        if rc == (1 << 64) - 1:
                            assert g["valid"] == 0
                            continue
                        extra = orc.lib.orc_refi_bits(nref, r) + mvr + (mvr < 4) + 2
                        got = (int(g["mv_sub"][0]), int(g["mv_sub"][1]), int(g["mv_int"][0]), int(g["mv_int"][1]),
                               int(g["cost_final"]), int(g["cost_sub"]) + lam * extra)
                        assert got == (mx, my, ix, iy, rc, orc.last_cost64), (x, y, w, h, r)
                        checked += 1
                    pu += 1
    assert checked > 1000


@pytest.mark.gpu
def test_bad_arguments_are_rejected():
    """Invalid configurations and arguments fail with an error (and leave
    the library usable): window out of range, no shapes, too many or no
    references, a list-1 count that leaves list 0 empty, a reference of
    another size, a server stream smaller than the window."""
    hip = ua.MeHip(8)
    lam = wl.lambda_mv(32)
    for args in [(256, 136, 256, lam, ua.MASK_ALL, 1),    # R > 255
                 (256, 136, 16, lam, 0, 1),               # no shape
                 (256, 136, 16, lam, ua.MASK_ALL, 0),     # no reference
                 (256, 136, 16, lam, ua.MASK_ALL, 5),     # more than 4
                 (64, 40, 61, lam, ua.MASK_ALL, 1)]:      # R > max(W, H) - 4
        with pytest.raises(ua.MeError):
            hip.plan_create(*args)
    with pytest.raises(ua.MeError):
        hip.plan_create(256, 136, 16, lam, ua.MASK_ALL, 2, num_refs_l1=2)
    work = wl.hd_allpu().scaled(256, 136, 8)
    work.ref_shifts_qpel = work.ref_shifts_qpel[:1]
    cur, refs = work.frames()
    other = hip.ref_create(np.zeros((136, 264), hip.dt))
    plan = hip.plan_create(256, 136, 8, lam, ua.MASK_16x16, 1)
    hip.plan_set_cur(plan, cur)
    with pytest.raises(ua.MeError):
        hip.plan_run(plan, [other])                        # reference of another size
    good = hip.ref_create(refs[0])
    hip.plan_run(plan, [good])                             # the plan still runs
    res, best = hip.plan_results(plan, 1)
    assert (res["valid"] == 1).sum() > 0
    hip.plan_destroy(plan)
    hip.ref_destroy(good)
    hip.ref_destroy(other)
    srv = ua.MeServer(hip, 32, lam, ua.MASK_16x16, 1)
    with pytest.raises(ua.MeError):
        srv.open_stream(24, 24)                            # max(W, H) - 4 < R
    srv.close()
    assert hip.error_count() == 0
