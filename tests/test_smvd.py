"""Symmetric-MVD search (analyze_smvd, inter.c:1194-1255, with smvd_refine,
inter.c:1083-1191): the GPU path (me_plan_run_smvd) against the oracle PU by
PU, and the oracle against what symmetric motion fixes.

The oracle composes pinned parts -- com_mc_cu's bi-average
(tests/golden/mc_bd*.npz), com_had and SAD (tests/golden/ref_bd*.npz), the
MV-bits table -- with the reference's candidate loop.  The loop is pinned by
tests/golden/smvd_bd*.npz: the compiled reference's SMVD search (its own
smvd_refine after analyze_smvd's first two candidates, through a minimal
core_t) for every PU of a frame-plan layout, which the oracle and the GPU
are both compared with."""
import ctypes
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, SHAPES
import uavs3e_amd as ua

LAMBDA = 314169
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _pic(rng, h, w, bd, dx=0, dy=0, noise=1.0):
    top = (1 << bd) - 1
    yy, xx = np.mgrid[0:h, 0:w]
    b = 0.5 + 0.25 * np.sin((xx - dx) / 6.0) * np.cos((yy - dy) / 4.0) + 0.1 * np.sin((xx + yy - dx - dy) / 2.5)
    return np.clip(np.rint(b * top + rng.normal(0, noise * top / 255, (h, w))), 0, top).astype(
        np.uint8 if bd == 8 else np.uint16)


@pytest.mark.parametrize("cost_type", [0, 1])
def test_smvd_oracle_finds_symmetric_motion(cost_type):
    """List 0 shifted by +d and list 1 by -d (quarter pel, on the sub-pel
    planes of one picture): from MVP 0 and list 0's uni MV at +d the search
    returns (+d, -d) with the cost of an exact prediction (0 distortion)."""
    o = Oracle(8)
    rng = np.random.default_rng(4)
    h, w = 64, 96
    cur = _pic(rng, h, w, 8, noise=0)
    planes = o.make_planes(cur)
    x, y, pw, ph = 32, 16, 16, 16
    # the current block itself is the sub-pel plane block at MV 0 of either
    # list: any symmetric pair (+d, -d) of the same picture averages back to
    # it only for d = 0, so the search must end at (0, 0) / (0, 0)
    mvp = np.zeros((2, 2), np.int16)
    ms = np.array([[8, -4], [0, 0]], np.int16)
    mv, cost = o.analyze_smvd(cur, [planes, planes], 1, x, y, pw, ph, mvp, ms, 0, cost_type, LAMBDA)
    assert mv == [[0, 0], [0, 0]]
    off, tab = o.mvbits(w, h)
    assert cost == (LAMBDA * (2 * int(tab[off]) + 0 + 1) + 32768) >> 16


def _smvd_inputs(rng, n_ctu, npu, nref, mvr):
    g = 1 << mvr
    mvp = (rng.integers(-12, 13, (n_ctu, npu, nref, 2)) * g).astype(np.int16)
    ms = (rng.integers(-12, 13, (n_ctu, npu, nref, 2)) * g).astype(np.int16)
    pus = np.zeros((n_ctu, npu), ua.SMVD_PU)
    pus["active"] = (rng.random((n_ctu, npu)) < 0.9).astype(np.int8)
    return mvp, ms, pus


@pytest.mark.gpu
@pytest.mark.parametrize("bd,n0,n1,mvr,ct", [(8, 1, 1, 0, 0), (8, 2, 2, 0, 1), (10, 1, 2, 0, 0), (10, 2, 1, 2, 1),
                                             (8, 1, 1, 1, 0), (8, 1, 1, 3, 0)])
def test_smvd_gpu_vs_oracle(bd, n0, n1, mvr, ct):
    shapes = [(16, 16), (8, 16), (32, 8), (64, 64), (128, 128), (4, 8), (16, 8), (8, 4)]
    mask = sum(1 << SHAPES.index(s) for s in shapes)
    w, h = 200, 136
    rng = np.random.default_rng(bd * 100 + n0 * 10 + n1 + mvr + 7 * ct)
    cur = _pic(rng, h, w, bd)
    refs = [_pic(rng, h, w, bd, 2 * k + 1, -k) for k in range(n0)] + [_pic(rng, h, w, bd, -3 * k - 1, k + 1)
                                                                       for k in range(n1)]
    hip, o = ua.MeHip(bd), Oracle(bd)
    dref = [hip.ref_create(r) for r in refs]
    plan = hip.plan_create(w, h, 8, LAMBDA, mask, n0 + n1, num_refs_l1=n1)
    hip.plan_set_cur(plan, cur)
    n_ctu, npu = hip.lib.me_plan_num_ctus(plan), hip.lib.me_plan_num_pus(plan)
    mvp, ms, pus = _smvd_inputs(rng, n_ctu, npu, n0 + n1, mvr)
    pus0 = pus.copy()
    tz = ua.TzCfg(16, 64, 0, mvr, ct, 8, (ctypes.c_int64 * 4)(0, 0, 0, 0))
    hip.plan_run_smvd(plan, dref, mvp, ms, pus, tz)
    hip.plan_destroy(plan)
    for r in dref:
        hip.ref_destroy(r)
    planes = [o.make_planes(r) for r in refs]
    cols = (w + 127) // 128
    sel = [s for s in range(len(SHAPES)) if mask >> s & 1]
    checked = 0
    for ctu in range(n_ctu):
        cx, cy = (ctu % cols) * 128, (ctu // cols) * 128
        pu = 0
        for s in sel:
            pw, ph = SHAPES[s]
            for i in range((128 // pw) * (128 // ph)):
                x, y = cx + (i % (128 // pw)) * pw, cy + (i // (128 // pw)) * ph
                if pus0[ctu, pu]["active"] and x + pw <= w and y + ph <= h:
                    mv, cost = o.analyze_smvd(cur, planes, n0, x, y, pw, ph, mvp[ctu, pu].astype(int),
                                              ms[ctu, pu].astype(int), mvr, ct, LAMBDA)
                    g = pus[ctu, pu]
                    where = (ctu, s, i, x, y)
                    assert g["mv"].astype(int).tolist() == mv, where
                    assert int(g["cost"]) == cost, where
                    checked += 1
                elif not pus0[ctu, pu]["active"]:
                    assert pus[ctu, pu].tobytes() == pus0[ctu, pu].tobytes()
                pu += 1
    assert checked > 300
    assert hip.error_count() == 0


def _smvd_golden(bd):
    return np.load(os.path.join(GOLD, "smvd_bd%d.npz" % bd), allow_pickle=False)


def _pus(g):
    w, h, mask = int(g["w"]), int(g["h"]), int(g["mask"])
    cols = (w + 127) // 128
    n_ctu = cols * ((h + 127) // 128)
    out = []
    for ctu in range(n_ctu):
        cx, cy = (ctu % cols) * 128, (ctu // cols) * 128
        pu = 0
        for s in range(len(SHAPES)):
            if mask >> s & 1:
                pw, ph = SHAPES[s]
                for i in range((128 // pw) * (128 // ph)):
                    out.append((ctu, pu, cx + (i % (128 // pw)) * pw, cy + (i // (128 // pw)) * ph, pw, ph))
                    pu += 1
    return n_ctu, pu, out


@pytest.mark.parametrize("bd", [8, 10])
def test_smvd_oracle_golden(bd):
    """The oracle's analyze_smvd against the compiled reference's SMVD search
    (its own smvd_refine; tests/golden/smvd_bd*.npz, gen_golden_bi_loop.py)
    on a spread sample of each configuration's PUs: both MVs and mecost."""
    g, o = _smvd_golden(bd), Oracle(bd)
    cur = g["cur"]
    h, w = cur.shape
    refs = list(g["refs"])
    _, _, pus = _pus(g)
    checked = 0
    for k in range(3):
        n0, n1, mvr, ct = (int(v) for v in g["cfg_%d" % k])
        lrefs = refs[:n0] + refs[2:2 + n1]
        planes = [o.make_planes(r) for r in lrefs]
        act = g["active_%d" % k]
        cand = [p for p in pus if act[p[0], p[1]] and p[2] + p[4] <= w and p[3] + p[5] <= h]
        for ctu, pu, x, y, pw, ph in cand[::7]:
            mv, cost = o.analyze_smvd(cur, planes, n0, x, y, pw, ph, g["mvp_%d" % k][ctu, pu].astype(int),
                                      g["mv_scale_%d" % k][ctu, pu].astype(int), mvr, ct, LAMBDA)
            where = (k, ctu, pu, x, y, pw, ph)
            assert mv == g["out_mv_%d" % k][ctu, pu].astype(int).tolist(), where
            assert cost == int(g["out_cost_%d" % k][ctu, pu]), where
            checked += 1
    assert checked > 250


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_smvd_gpu_golden(bd):
    """me_plan_run_smvd on the fixture's inputs against the compiled
    reference's SMVD search directly: every active PU inside the picture."""
    g = _smvd_golden(bd)
    cur = g["cur"]
    h, w = cur.shape
    refs = list(g["refs"])
    n_ctu, npu, pus_l = _pus(g)
    hip = ua.MeHip(bd)
    checked = 0
    for k in range(3):
        n0, n1, mvr, ct = (int(v) for v in g["cfg_%d" % k])
        dref = [hip.ref_create(r) for r in refs[:n0] + refs[2:2 + n1]]
        plan = hip.plan_create(w, h, 8, LAMBDA, int(g["mask"]), n0 + n1, num_refs_l1=n1)
        hip.plan_set_cur(plan, cur)
        assert (hip.lib.me_plan_num_ctus(plan), hip.lib.me_plan_num_pus(plan)) == (n_ctu, npu)
        pus = np.zeros((n_ctu, npu), ua.SMVD_PU)
        pus["active"] = g["active_%d" % k]
        tz = ua.TzCfg(16, 64, 0, mvr, ct, 8, (ctypes.c_int64 * 4)(0, 0, 0, 0))
        hip.plan_run_smvd(plan, dref, g["mvp_%d" % k], g["mv_scale_%d" % k], pus, tz)
        hip.plan_destroy(plan)
        for r in dref:
            hip.ref_destroy(r)
        for ctu, pu, x, y, pw, ph in pus_l:
            if not g["active_%d" % k][ctu, pu] or x + pw > w or y + ph > h:
                continue
            where = (k, ctu, pu, x, y, pw, ph)
            assert pus[ctu, pu]["mv"].tolist() == g["out_mv_%d" % k][ctu, pu].tolist(), where
            assert int(pus[ctu, pu]["cost"]) == int(g["out_cost_%d" % k][ctu, pu]), where
            checked += 1
    assert checked > 1500
    assert hip.error_count() == 0
