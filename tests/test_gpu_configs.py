"""The benchmark configurations' search parameters under GPU parity.

The frame plan at the full ranges the workloads are quoted on -- 10-bit with
four list-0 references at +-128 (the 2160p workload, every schedule shape the
autotuner may pick, column-tiled ones included), +-64 over all 29 shapes at
both bit depths (the 1080p workload), +-64 with two lists of two references
and motion compensation (the 8K bi-directional workload) -- on pictures of
one to four CTUs, compared bit-exactly with the oracle.  Also: residuals at
the sample range's bound, where the packed 16-bit Hadamard of the sub-pel
kernel has the least headroom, and the double-precision SATD normalisation
for every reachable sum."""
import ctypes
import math

import numpy as np
import pytest

import orc_par
import uavs3e_amd as ua
from oracle.oracle import Oracle, MASK_ALL, SHAPES
from uavs3e_amd import workloads as wl

pytestmark = pytest.mark.gpu

LAM = wl.lambda_mv(32)

# One shape of every integer-search node class family: 4-wide, 8x8, the
# lane-region shapes, the first sparse class, 64x64 and the whole CTU.
CLASS_MASK = sum(1 << SHAPES.index(s) for s in [(4, 8), (8, 8), (16, 16), (32, 16), (64, 64), (128, 128)])


def _run_plan(hip, cur, refs, R, mask, num_refs_l1=0, chroma=None, **kw):
    h, w = cur.shape
    dref = [hip.ref_create(r) for r in refs]
    if chroma is not None:
        for d, (u, v) in zip(dref, chroma):
            hip.ref_load_chroma(d, u, v)
    plan = hip.plan_create(w, h, R, LAM, mask, len(refs), num_refs_l1=num_refs_l1, **kw)
    hip.plan_set_cur(plan, cur)
    hip.plan_run(plan, dref)
    mc = None
    if chroma is not None:
        hip.plan_mc(plan, dref)
        mc = hip.plan_mc_download(plan)
    res, best = hip.plan_results(plan, len(refs))
    hip.plan_destroy(plan)
    for r in dref:
        hip.ref_destroy(r)
    return res, best, mc


def _compare(res, best, ores, obest):
    msg = ua.compare_results(res, best, ores, obest)
    assert msg is None, msg


@pytest.fixture(scope="module")
def uhd10_r128():
    """2160p 10-bit workload parameters (4 list-0 references at POC
    distances 1, 2, 4, 8 with cumulative motion, +-128) on a 200x136
    picture, and the oracle's results."""
    work = wl.uhd10_allpu().scaled(200, 136, 128)
    assert len(work.ref_shifts_qpel) == 4
    cur, refs = work.frames()
    orc = Oracle(10)
    print("oracle: 10-bit 200x136, 4 refs, R=128, 6 shapes", flush=True)
    ores, obest = orc_par.me_frame(orc, cur, [orc.make_planes(r) for r in refs], 128, LAM, CLASS_MASK)
    return cur, refs, ores, obest


# (band_rows, waves, dx_cols, sp_groups, int_grid): 0 = the library default
# (16 rows, 4 waves, all columns, one sub-pel position group, persistent grid
# where the frame fills the GPU); the autotuner picks 32 rows of 8 waves over
# 64-column tiles on the persistent grid at 2160p.  Grids of a few persistent
# workgroups split the segments of the top CTU row, whose first band tests
# the window (R = 128 reaches past -127) and whose others do not, so one
# workgroup runs both kinds and merges their keys.
@pytest.mark.parametrize("sched", [(0, 0, 0, 0, 1), (8, 2, 64, 2, 1), (4, 1, 32, 0, 1), (32, 4, 64, 8, 1),
                                   (5, 4, 128, 0, 1), (16, 8, 0, 4, 1), (16, 8, 0, 0, 5), (32, 8, 64, 0, 7),
                                   (8, 4, 0, 1, 3), (5, 2, 128, 0, 11)],
                         ids=lambda s: "b%d_w%d_c%d_g%d_grid%d" % s)
def test_10bit_4ref_r128(uhd10_r128, sched):
    cur, refs, ores, obest = uhd10_r128
    hip = ua.MeHip(10)
    band, waves, cols, g, grid = sched
    res, best, _ = _run_plan(hip, cur, refs, 128, CLASS_MASK, band_rows=band, waves=waves, dx_cols=cols,
                             sp_groups=g, int_grid=grid)
    _compare(res, best, ores, obest)
    # the window reaches the range's edge: some PUs end on a displacement of 128
    assert res["valid"].sum() > 100
    assert hip.error_count() == 0


@pytest.mark.parametrize("bd", [8, 10])
def test_r64_all_shapes(bd):
    """All 29 shapes at +-64 with 2 references (the 1080p workload's
    search at both bit depths), including edge CTUs."""
    base = wl.hd_allpu() if bd == 8 else wl.uhd10_allpu()
    work = base.scaled(200, 136, 64)
    work.ref_shifts_qpel = work.ref_shifts_qpel[:2]
    cur, refs = work.frames()
    hip, orc = ua.MeHip(bd), Oracle(bd)
    print("oracle: %d-bit 200x136, 2 refs, R=64, 29 shapes" % bd, flush=True)
    ores, obest = orc_par.me_frame(orc, cur, [orc.make_planes(r) for r in refs], 64, LAM, MASK_ALL)
    res, best, _ = _run_plan(hip, cur, refs, 64, MASK_ALL)
    _compare(res, best, ores, obest)
    assert res["valid"].sum() > 5000


def test_r64_two_lists_mc_10bit():
    """The 8K bi-directional workload's parameters (10-bit, 2 list-0 + 2
    list-1 references, brightness drift, +-64, all shapes) on a 200x136
    picture: per-list results and the motion-compensated luma and chroma."""
    work = wl.fuhd10_bi().scaled(200, 136, 64)
    cur, refs = work.frames()
    chroma = work.ref_chroma()
    hip, orc = ua.MeHip(10), Oracle(10)
    planes = [orc.make_planes(r) for r in refs]
    print("oracle: 10-bit 200x136, 2+2 refs, R=64, 29 shapes", flush=True)
    ores, obest = orc_par.me_frame(orc, cur, planes, 64, LAM, MASK_ALL, num_refs_l1=2)
    res, best, (py, pu, pv) = _run_plan(hip, cur, refs, 64, MASK_ALL, num_refs_l1=2, chroma=chroma)
    _compare(res, best, ores, obest)
    oy, ou, ov = orc.mc_frame(work.w, work.h, planes, [(orc.make_chroma(u), orc.make_chroma(v)) for u, v in chroma],
                              2, MASK_ALL, ores, obest)
    assert np.array_equal(py, oy), "luma predictions differ at %d samples" % int((py != oy).sum())
    assert np.array_equal(pu, ou) and np.array_equal(pv, ov), "chroma predictions differ"


def _saturated(bd, kind, w, h, seed):
    top = (1 << bd) - 1
    dt = ua.pel_dtype(bd)
    rng = np.random.default_rng(seed)
    if kind == "const":      # every residual at +max: the largest DC sums
        return np.full((h, w), top, dt), [np.zeros((h, w), dt)]
    if kind == "binary":     # residuals of +-max and 0 in random signs
        return (rng.integers(0, 2, (h, w)) * top).astype(dt), [(rng.integers(0, 2, (h, w)) * top).astype(dt)]
    # checkerboards of opposite phase: residual +-max alternating, the
    # largest high-frequency coefficient of every Hadamard tile
    yy, xx = np.mgrid[0:h, 0:w]
    cb = ((xx + yy) & 1).astype(np.int64) * top
    return cb.astype(dt), [(top - cb).astype(dt)]


@pytest.mark.parametrize("bd", [8, 10])
@pytest.mark.parametrize("kind", ["const", "binary", "checker"])
def test_saturated_residuals(bd, kind):
    """Residuals at the bound of the sample range reach the packed 16-bit
    Hadamard's largest intermediates (+-16320 at 8 bits, +-32736 for 10-bit
    tiles up to 8x8) and the 32-bit 10-bit halves; costs stay bit-exact."""
    cur, refs = _saturated(bd, kind, 200, 136, 5 + bd)
    hip, orc = ua.MeHip(bd), Oracle(bd)
    ores, obest = orc.me_frame(cur, [orc.make_planes(r) for r in refs], 8, LAM, MASK_ALL)
    res, best, _ = _run_plan(hip, cur, refs, 8, MASK_ALL)
    _compare(res, best, ores, obest)


@pytest.mark.parametrize("big", [0, 1], ids=["sqrt32", "sqrt128"])
def test_satd_norm_every_reachable_sum(big):
    """(int)(s / sqrt(c) * 2.0) of com_cost.c:467, 586, 649, 708 on the GPU
    for every s below 2^24, against IEEE double on the host.  Every
    reachable sum is below 2^24: a 16x8 tile has 128 coefficients of at
    most 128 * 1023 at 10 bits (16,760,832 in all), an 8x4 tile 32 of at
    most 32 * 1023."""
    hip = ua.MeHip(10)
    c = 128.0 if big else 32.0
    root = math.sqrt(c)
    assert root == float.fromhex("0x1.6a09e667f3bcdp+3" if big else "0x1.6a09e667f3bcdp+2")
    n = 1 << 24
    got = np.zeros(n, np.uint32)
    hip.check(hip.lib.me_hip_satd_norm(big, 0, n, got.ctypes.data_as(ctypes.c_void_p)), "me_hip_satd_norm")
    s = np.arange(n, dtype=np.float64)
    want = np.trunc(s / root * 2.0).astype(np.uint32)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, "first mismatch at s = %d: %d vs %d" % (bad[0], got[bad[0]], want[bad[0]])
