"""The lookahead cost estimator's inter part (loka_estimate_coding_cost,
lookahead.c:82-208): the oracle pinned against the compiled reference, and
the GPU path (me_plan_lookahead) against the oracle unit by unit.

The reference returns only the picture's mean of min(inter, intra) over the
units; on the golden pictures an inter prediction beats every intra mode in
every unit (tests/golden/gen_golden_loka.py), so the reference's value must
equal the oracle's inter costs summed.  A unit whose inter cost the oracle
gets too high or too low moves the sum."""
import ctypes
import importlib.util
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, RefLib, MASK_16x16, MASK_ALL, SHAPES
import uavs3e_amd as ua

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(bd):
    return np.load(os.path.join(GOLD, "loka_bd%d.npz" % bd), allow_pickle=False)


def _gen():
    spec = importlib.util.spec_from_file_location("gen_golden_loka", os.path.join(GOLD, "gen_golden_loka.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _oracle_case(o, cur, refs, ref_poc, case, params):
    pad, n0, n1 = (int(v) for v in case[:3])
    l0, l1 = [int(i) for i in case[3:3 + n0]], [int(i) for i in case[7:7 + n1]]
    gop, msr, cur_poc = (int(v) for v in params)
    return o.loka_inter_cost(cur, [o.make_planes(refs[i]) for i in l0], [o.make_planes(refs[i]) for i in l1],
                             cur_poc, [int(ref_poc[i]) for i in l0], [int(ref_poc[i]) for i in l1], gop, msr, pad)


@pytest.mark.parametrize("bd", [8, 10])
def test_loka_oracle_golden(bd):
    g, o = _golden(bd), Oracle(bd)
    for case, want in zip(g["cases"], g["cost"]):
        u = _oracle_case(o, g["cur"], g["refs"], g["ref_poc"], case, g["params"])
        assert float(u.astype(np.float64).sum()) / u.size / 256 == want, case


@pytest.mark.skipif(not RefLib.available(8), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("bd,seed", [(8, 1), (10, 2)])
def test_loka_oracle_live(bd, seed):
    m = _gen()
    cur, refs = m.make_pictures(bd, seed)
    o, ref = Oracle(bd), RefLib(bd, 0)
    want = m.reference_costs(ref, cur, refs)
    for (pad, l0, l1), w in zip(m.CASES, want):
        case = [pad, len(l0), len(l1)] + (l0 + [-1] * 4)[:4] + (l1 + [-1] * 4)[:4]
        u = _oracle_case(o, cur, refs, m.REF_POC, case, (m.GOP, m.MSR, m.CUR_POC))
        assert float(u.astype(np.float64).sum()) / u.size / 256 == w, case


def test_loka_oracle_special_cases():
    """References equal to the current picture cost nothing; with both lists
    the zero-MV average of cur + n and cur - n is exact, even where each
    single prediction costs more."""
    o = Oracle(8)
    rng = np.random.default_rng(3)
    cur = rng.integers(20, 236, (48, 64)).astype(np.uint8)
    pl = o.make_planes(cur)
    for pad in (0, 1):
        u = o.loka_inter_cost(cur, [pl], [], 4, [0], [], 16, 64, pad)
        # without padding the last unit column / row cannot take MV 0
        assert (u[:-1, :-1] == 0).all() if pad == 0 else (u == 0).all()
    n = rng.integers(-6, 7, cur.shape)
    a = (cur.astype(int) + n).astype(np.uint8)
    b = (cur.astype(int) - n).astype(np.uint8)
    u0 = o.loka_inter_cost(cur, [o.make_planes(a)], [], 4, [0], [], 16, 64, 1)
    ub = o.loka_inter_cost(cur, [o.make_planes(a)], [o.make_planes(b)], 4, [0], [8], 16, 64, 1)
    assert (u0 > 0).all() and (ub == 0).all()


def _gpu_case(hip, cur, refs_l0, refs_l1, pad, mask=MASK_16x16):
    h, w = cur.shape
    refs = list(refs_l0) + list(refs_l1)
    pocs = [4, 0, 2, 6][:len(refs_l0)] + [12, 16][:len(refs_l1)]
    dref = [hip.ref_create(r) for r in refs]
    plan = hip.plan_create(w, h, 16, 314169, mask, len(refs), num_refs_l1=len(refs_l1))
    hip.plan_set_cur(plan, cur)
    # mvr, cost_type and the adaptive raster range are the lookahead's own whatever is given
    tz = ua.TzCfg(16, 64, 1, 0, 1, 8, (ctypes.c_int64 * 4)(*(pocs + [0] * 4)[:4]))
    got = hip.plan_lookahead(plan, dref, tz, pad, w, h)
    hip.plan_destroy(plan)
    for r in dref:
        hip.ref_destroy(r)
    return got, pocs


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_loka_gpu_golden(bd):
    """The golden pictures: every case equals the oracle and, through it,
    the reference's picture cost."""
    g = _golden(bd)
    hip, o = ua.MeHip(bd), Oracle(bd)
    for case, want in zip(g["cases"], g["cost"]):
        pad, n0, n1 = (int(v) for v in case[:3])
        l0, l1 = [int(i) for i in case[3:3 + n0]], [int(i) for i in case[7:7 + n1]]
        if n0 + n1 > 4:
            continue
        refs = g["refs"]
        h, w = g["cur"].shape
        dref = [hip.ref_create(refs[i]) for i in l0 + l1]
        plan = hip.plan_create(w, h, 16, 314169, MASK_16x16, n0 + n1, num_refs_l1=n1)
        hip.plan_set_cur(plan, g["cur"])
        pocs = [int(g["ref_poc"][i]) for i in l0 + l1]
        gop, msr, cur_poc = (int(v) for v in g["params"])
        tz = ua.TzCfg(gop, msr, 0, 2, 0, cur_poc, (ctypes.c_int64 * 4)(*(pocs + [0] * 4)[:4]))
        got = hip.plan_lookahead(plan, dref, tz, pad, w, h)
        hip.plan_destroy(plan)
        for r in dref:
            hip.ref_destroy(r)
        want_u = _oracle_case(o, g["cur"], refs, g["ref_poc"], case, g["params"])
        assert np.array_equal(got, want_u), case
        assert float(got.astype(np.float64).sum()) / got.size / 256 == want, case
    assert hip.error_count() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("bd,w,h,n0,n1,pad", [(8, 200, 136, 2, 2, 1), (8, 200, 136, 1, 1, 0), (10, 136, 104, 3, 0, 1),
                                              (10, 136, 104, 2, 1, 0), (8, 264, 152, 4, 0, 0)])
def test_loka_gpu_vs_oracle(bd, w, h, n0, n1, pad):
    """Smooth moving content with noise (the searches travel and the raster
    stage runs), pictures whose sizes are not multiples of 16 or 128 (partial
    units are left out, partial CTUs hold whole units); other schedules of
    the plan (all shapes) give the same."""
    rng = np.random.default_rng(w + h + n0 + 7 * n1 + pad + bd)
    top = (1 << bd) - 1
    yy, xx = np.mgrid[0:h, 0:w]

    def pic(dx, dy, amp):
        base = 0.5 + 0.25 * np.sin((xx - dx) / 7.0) * np.cos((yy - dy) / 5.0) + 0.15 * np.sin((xx + yy - dx - dy) / 3.0)
        v = base * top + rng.normal(0, amp * top / 255.0, (h, w))
        return np.clip(np.rint(v), 0, top).astype(np.uint8 if bd == 8 else np.uint16)

    cur = pic(0, 0, 3)
    l0 = [pic(3 * k + 2, -2 * k - 1, 4) for k in range(n0)]
    l1 = [pic(-5 * k - 3, 4 * k + 6, 4) for k in range(n1)]
    hip, o = ua.MeHip(bd), Oracle(bd)
    got, pocs = _gpu_case(hip, cur, l0, l1, pad)
    want = o.loka_inter_cost(cur, [o.make_planes(r) for r in l0], [o.make_planes(r) for r in l1], 8,
                             pocs[:n0], pocs[n0:], 16, 64, pad)
    assert np.array_equal(got, want)
    assert (want < np.uint32(0xFFFFFFFF)).all()
    got_all, _ = _gpu_case(hip, cur, l0, l1, pad, MASK_ALL)
    assert np.array_equal(got_all, want)
    assert hip.error_count() == 0


@pytest.mark.gpu
def test_loka_rejects_bad_plans():
    hip = ua.MeHip(8)
    cur = np.zeros((64, 64), np.uint8)
    r = hip.ref_create(cur)
    plan = hip.plan_create(64, 64, 8, 314169, 1 << SHAPES.index((8, 8)), 1)   # no 16x16 shape
    hip.plan_set_cur(plan, cur)
    tz = ua.TzCfg(16, 64, 0, 2, 0, 8, (ctypes.c_int64 * 4)(0, 0, 0, 0))
    with pytest.raises(ua.MeError):
        hip.plan_lookahead(plan, [r], tz, 1, 64, 64)
    hip.plan_destroy(plan)
    plan = hip.plan_create(64, 64, 8, 314169, MASK_16x16, 1)
    hip.plan_set_cur(plan, cur)
    assert (hip.plan_lookahead(plan, [r], tz, 1, 64, 64) == 0).all()
    hip.plan_destroy(plan)
    hip.ref_destroy(r)
