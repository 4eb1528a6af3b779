"""Bi-prediction ME (analyze_bi, inter.c:972-1033): me_search_tz with bi = 1
in the oracle pinned against the compiled reference, and the GPU path
(me_plan_run_bi) against the oracle's analyze_bi PU by PU.

The oracle's analyze_bi composes pinned parts -- me_search_tz(bi = 1)
(golden vectors here), com_mc_cu luma (tests/golden/mc_bd*.npz), the MV-bits
table and tbl_refi_bits (tests/test_oracle.py) -- with the reference's loop
(two rounds, strict '<' against the running best, the searched list becomes
the fixed one, a round without a change ends it).  The loop is pinned by
tests/golden/bi_loop_bd*.npz: the compiled reference's own analyze_bi,
called through a minimal core_t and returning right after its search, for
every PU of a frame-plan layout; the oracle and the GPU are both compared
with it.  best_mecost is not an output of analyze_bi; it is the
me_search_tz(bi = 1) cost of the winner, pinned by bi_bd*.npz."""
import ctypes
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, RefLib, SHAPES
import uavs3e_amd as ua

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LAMBDA = 314169


def _golden(bd):
    return np.load(os.path.join(GOLD, "bi_bd%d.npz" % bd), allow_pickle=False)


def _oracle_bi_case(o, cur, planes, tgts, row):
    s, x, y, mvr, ct, lidx, nref, refi, mb, px, py, mx, my, off = (int(v) for v in row)
    w, h = SHAPES[s]
    tgt = tgts[off:off + w * h].reshape(h, w)
    m, c = o.me_search_tz(cur, planes, x, y, w, h, refi, nref, 0, 0, 16, 64, mvr, ct, 0, LAMBDA, (px, py),
                          org_bi=tgt, mv_start=(mx, my), bi_bits=mb + 1)
    return c, m[:2]


@pytest.mark.parametrize("bd", [8, 10])
def test_bi_search_golden(bd):
    g, o = _golden(bd), Oracle(bd)
    planes = o.make_planes(g["ref"])
    for row, c_want, mv_want in zip(g["cases"], g["cost"], g["mv"]):
        c, m = _oracle_bi_case(o, g["cur"], planes, g["targets"], row)
        assert c == int(c_want) and m == tuple(int(v) for v in mv_want), row.tolist()


@pytest.mark.skipif(not RefLib.available(8), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("bd", [8, 10])
def test_bi_search_live(bd):
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_golden_bi", os.path.join(GOLD, "gen_golden_bi.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    cur, refimg = m.make_pictures(bd, 5)
    cases, tgts = m.make_cases(bd, 6, cur)
    o, ref = Oracle(bd), RefLib(bd, 2)
    planes = o.make_planes(refimg)
    pic = ref.pic_create(refimg, 0)
    for row in cases[:60]:
        s, x, y, mvr, ct, lidx, nref, refi, mb, px, py, mx, my, off = (int(v) for v in row)
        w, h = SHAPES[s]
        c_ref, m_ref = ref.me_search_tz_bi(pic, tgts[off:off + w * h].reshape(h, w), x, y, refi, nref, lidx, mb,
                                           mvr, ct, LAMBDA, (px, py), (mx, my))
        assert _oracle_bi_case(o, cur, planes, tgts, row) == (c_ref, m_ref), row.tolist()
    ref.pic_free(pic)


def test_analyze_bi_properties():
    """A target that the fixed list predicts exactly (org == its
    prediction, so 2*org - pred == org) turns the first round into a
    uni-prediction search from the start MV with halved distortion: the
    chosen list-1 reference is the first cheapest of those searches; the
    outputs describe the last searched list's winner (mot_bits from its MV);
    and the best cost never exceeds the first round's cheapest search."""
    o = Oracle(8)
    rng = np.random.default_rng(9)
    h, w = 64, 96
    yy, xx = np.mgrid[0:h, 0:w]
    cur = np.clip((0.5 + 0.3 * np.sin(xx / 5.0) * np.cos(yy / 3.0)) * 255 + rng.normal(0, 3, (h, w)), 0,
                  255).astype(np.uint8)
    refs = [cur.copy(), np.roll(cur, (1, 2), axis=(0, 1)), np.roll(cur, (-2, 1), axis=(0, 1))]
    planes = [o.make_planes(r) for r in refs]
    x, y, pw, ph = 32, 16, 16, 16
    mvp = np.zeros((3, 2), np.int16)
    mv_scale = np.array([[0, 0], [4, -4], [8, 4]], np.int16)
    refi, mv, mot = [0, 0], [[0, 0], [0, 0]], [4, 6]
    first = []
    for rc in range(2):   # list 1 = refs 1, 2 searched against org itself
        m, c = o.me_search_tz(cur, planes[1 + rc], x, y, pw, ph, rc, 2, 0, 0, 16, 64, 0, 0, 0, LAMBDA,
                              (0, 0), org_bi=cur[y:y + ph, x:x + pw], mv_start=tuple(mv_scale[1 + rc]),
                              bi_bits=mot[0] + 1)
        first.append((c, m[:2]))
    best = o.analyze_bi(cur, planes, 1, x, y, pw, ph, mvp, mv_scale, refi, mv, 0, mot, 0, 0, LAMBDA)
    c1 = min(c for c, _ in first)
    k = [c for c, _ in first].index(c1)
    assert best <= c1
    if best == c1:   # no improvement in the second round: list 1 keeps the first round's winner
        assert refi[1] == k and tuple(mv[1]) == first[k][1]
    off, tab = o.mvbits(w, h)
    for l in (0, 1):
        r = [0, 1 + refi[1]][l]
        nl = [1, 2][l]
        want = o.lib.orc_refi_bits(nl, refi[l]) + int(tab[off + mv[l][0] - mvp[r][0]]) + int(tab[off + mv[l][1] - mvp[r][1]])
        if l == 1 or best < c1:
            assert mot[l] == want, (l, mot, want)


def _bi_inputs(rng, n_ctu, npu, n0, n1, mvr):
    g = 1 << mvr
    nref = n0 + n1
    mvp = (rng.integers(-24, 25, (n_ctu, npu, nref, 2)) * g).astype(np.int16)
    mv_scale = (rng.integers(-16, 17, (n_ctu, npu, nref, 2)) * g).astype(np.int16)
    pus = np.zeros((n_ctu, npu), ua.BI_PU)
    pus["refi"][..., 0] = rng.integers(0, n0, (n_ctu, npu))
    pus["refi"][..., 1] = rng.integers(0, n1, (n_ctu, npu))
    pus["lidx_ref"] = rng.integers(0, 2, (n_ctu, npu))
    pus["active"] = (rng.random((n_ctu, npu)) < 0.9).astype(np.int8)
    pus["mv"] = (rng.integers(-16, 17, (n_ctu, npu, 2, 2)) * g).astype(np.int16)
    pus["mot_bits"] = rng.integers(2, 30, (n_ctu, npu, 2))
    return mvp, mv_scale, pus


@pytest.mark.gpu
@pytest.mark.parametrize("bd,n0,n1,mvr,ct", [(8, 1, 1, 0, 0), (8, 2, 2, 0, 1), (10, 2, 1, 0, 0), (8, 1, 2, 2, 0),
                                             (10, 1, 1, 3, 1), (8, 2, 2, 1, 0)])
def test_bi_gpu_vs_oracle(bd, n0, n1, mvr, ct):
    """Every PU of a set of shapes (pictures with partial CTUs): refined
    references, MVs, mot_bits, cost and the updated mv_scale equal the
    oracle's analyze_bi."""
    shapes = [(16, 16), (8, 16), (32, 8), (64, 64), (128, 128), (4, 8)]
    mask = sum(1 << SHAPES.index(s) for s in shapes)
    w, h = 200, 136
    rng = np.random.default_rng(bd * 100 + n0 * 10 + n1 + mvr + 7 * ct)
    top = (1 << bd) - 1
    yy, xx = np.mgrid[0:h, 0:w]

    def pic(dx, dy):
        b = 0.5 + 0.25 * np.sin((xx - dx) / 6.0) * np.cos((yy - dy) / 4.0) + 0.1 * np.sin((xx + yy - dx) / 2.5)
        return np.clip(np.rint(b * top + rng.normal(0, top / 80, (h, w))), 0, top).astype(
            np.uint8 if bd == 8 else np.uint16)

    cur = pic(0, 0)
    refs = [pic(2 * k + 1, -k) for k in range(n0)] + [pic(-3 * k - 2, k + 1) for k in range(n1)]
    hip, o = ua.MeHip(bd), Oracle(bd)
    dref = [hip.ref_create(r) for r in refs]
    plan = hip.plan_create(w, h, 8, LAMBDA, mask, n0 + n1, num_refs_l1=n1)
    hip.plan_set_cur(plan, cur)
    n_ctu, npu = hip.lib.me_plan_num_ctus(plan), hip.lib.me_plan_num_pus(plan)
    mvp, mv_scale, pus = _bi_inputs(rng, n_ctu, npu, n0, n1, mvr)
    pus0, ms0 = pus.copy(), mv_scale.copy()
    tz = ua.TzCfg(16, 64, 0, mvr, ct, 8, (ctypes.c_int64 * 4)(0, 0, 0, 0))
    hip.plan_run_bi(plan, dref, mvp, mv_scale, pus, tz)
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
                q0 = pus0[ctu, pu]
                if q0["active"] and x + pw <= w and y + ph <= h:
                    refi, mv, mot = [int(v) for v in q0["refi"]], q0["mv"].astype(int).tolist(), \
                        [int(v) for v in q0["mot_bits"]]
                    ms = ms0[ctu, pu].astype(int).tolist()
                    best = o.analyze_bi(cur, planes, n0, x, y, pw, ph, mvp[ctu, pu].astype(int).tolist(), ms, refi,
                                        mv, int(q0["lidx_ref"]), mot, mvr, ct, LAMBDA)
                    g = pus[ctu, pu]
                    where = (ctu, s, i, x, y)
                    assert int(g["cost"]) == min(best, 0xFFFFFFFF), where
                    assert [int(v) for v in g["refi"]] == refi, where
                    assert g["mv"].astype(int).tolist() == [list(v) for v in mv], where
                    assert [int(v) for v in g["mot_bits"]] == mot, where
                    assert mv_scale[ctu, pu].astype(int).tolist() == [list(v) for v in ms], where
                    checked += 1
                elif not q0["active"]:
                    assert pus[ctu, pu].tobytes()[:-4] == q0.tobytes()[:-4]
                pu += 1
    assert checked > 300
    assert hip.error_count() == 0


@pytest.mark.gpu
def test_bi_and_smvd_reject_bad_plans():
    """Both searches need references in both lists, and PU reference
    indices inside their lists."""
    hip = ua.MeHip(8)
    w, h = 128, 64
    cur = np.zeros((h, w), np.uint8)
    refs = [hip.ref_create(cur) for _ in range(2)]
    tz = ua.TzCfg(16, 64, 0, 0, 0, 8, (ctypes.c_int64 * 4)(0, 0, 0, 0))
    plan = hip.plan_create(w, h, 8, LAMBDA, 1 << SHAPES.index((16, 16)), 2)   # one list only
    hip.plan_set_cur(plan, cur)
    n_ctu, npu = hip.lib.me_plan_num_ctus(plan), hip.lib.me_plan_num_pus(plan)
    mvp = np.zeros((n_ctu, npu, 2, 2), np.int16)
    ms = mvp.copy()
    with pytest.raises(ua.MeError):
        hip.plan_run_bi(plan, refs, mvp, ms, np.zeros((n_ctu, npu), ua.BI_PU), tz)
    with pytest.raises(ua.MeError):
        hip.plan_run_smvd(plan, refs, mvp, ms, np.zeros((n_ctu, npu), ua.SMVD_PU), tz)
    hip.plan_destroy(plan)
    plan = hip.plan_create(w, h, 8, LAMBDA, 1 << SHAPES.index((16, 16)), 2, num_refs_l1=1)
    hip.plan_set_cur(plan, cur)
    pus = np.zeros((n_ctu, npu), ua.BI_PU)
    pus["active"] = 1
    pus["refi"][0, 0, 1] = 1                               # list 1 has one reference
    with pytest.raises(ua.MeError):
        hip.plan_run_bi(plan, refs, mvp, ms, pus, tz)
    pus["refi"][0, 0, 1] = 0
    hip.plan_run_bi(plan, refs, mvp, ms, pus, tz)          # identical pictures: both MVs stay 0
    assert (pus["mv"] == 0).all()
    hip.plan_destroy(plan)
    for r in refs:
        hip.ref_destroy(r)


def _loop_golden(bd):
    return np.load(os.path.join(GOLD, "bi_loop_bd%d.npz" % bd), allow_pickle=False)


def _loop_pus(g):
    """(ctu, pu, x, y, w, h) of every PU of the fixture's plan layout."""
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


def _loop_refs(g, n0, n1):
    refs = list(g["refs"])
    return refs[:n0] + refs[2:2 + n1]


@pytest.mark.parametrize("bd", [8, 10])
def test_analyze_bi_loop_golden(bd):
    """The oracle's analyze_bi against the compiled reference's own
    analyze_bi (tests/golden/bi_loop_bd*.npz, gen_golden_bi_loop.py): the
    references, MVs and mot_bits after the search and every updated
    mv_scale, on a spread sample of each configuration's PUs."""
    g, o = _loop_golden(bd), Oracle(bd)
    cur = g["cur"]
    h, w = cur.shape
    _, _, pus = _loop_pus(g)
    checked = 0
    for k in range(3):
        n0, n1, mvr, ct = (int(v) for v in g["cfg_%d" % k])
        planes = [o.make_planes(r) for r in _loop_refs(g, n0, n1)]
        act = g["active_%d" % k]
        cand = [p for p in pus if act[p[0], p[1]] and p[2] + p[4] <= w and p[3] + p[5] <= h]
        for ctu, pu, x, y, pw, ph in cand[::7]:
            refi = [int(v) for v in g["refi_%d" % k][ctu, pu]]
            mv = g["mv_%d" % k][ctu, pu].astype(int).tolist()
            mot = [int(v) for v in g["mot_%d" % k][ctu, pu]]
            ms = g["mv_scale_%d" % k][ctu, pu].astype(int).tolist()
            o.analyze_bi(cur, planes, n0, x, y, pw, ph, g["mvp_%d" % k][ctu, pu].astype(int).tolist(), ms, refi, mv,
                         int(g["lidx_%d" % k][ctu, pu]), mot, mvr, ct, LAMBDA)
            where = (k, ctu, pu, x, y, pw, ph)
            assert refi == g["out_refi_%d" % k][ctu, pu].astype(int).tolist(), where
            assert [list(v) for v in mv] == g["out_mv_%d" % k][ctu, pu].astype(int).tolist(), where
            assert mot == g["out_mot_%d" % k][ctu, pu].astype(int).tolist(), where
            assert [list(v) for v in ms] == g["out_mv_scale_%d" % k][ctu, pu].astype(int).tolist(), where
            checked += 1
    assert checked > 250


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_bi_gpu_golden(bd):
    """me_plan_run_bi on the fixture's inputs against the compiled
    reference's analyze_bi directly: every active PU inside the picture."""
    g = _loop_golden(bd)
    cur = g["cur"]
    h, w = cur.shape
    n_ctu, npu, pus_l = _loop_pus(g)
    hip = ua.MeHip(bd)
    checked = 0
    for k in range(3):
        n0, n1, mvr, ct = (int(v) for v in g["cfg_%d" % k])
        dref = [hip.ref_create(r) for r in _loop_refs(g, n0, n1)]
        plan = hip.plan_create(w, h, 8, LAMBDA, int(g["mask"]), n0 + n1, num_refs_l1=n1)
        hip.plan_set_cur(plan, cur)
        assert (hip.lib.me_plan_num_ctus(plan), hip.lib.me_plan_num_pus(plan)) == (n_ctu, npu)
        pus = np.zeros((n_ctu, npu), ua.BI_PU)
        pus["refi"] = g["refi_%d" % k]
        pus["lidx_ref"] = g["lidx_%d" % k]
        pus["active"] = g["active_%d" % k]
        pus["mv"] = g["mv_%d" % k]
        pus["mot_bits"] = g["mot_%d" % k]
        ms = np.ascontiguousarray(g["mv_scale_%d" % k]).copy()
        tz = ua.TzCfg(16, 64, 0, mvr, ct, 8, (ctypes.c_int64 * 4)(0, 0, 0, 0))
        hip.plan_run_bi(plan, dref, g["mvp_%d" % k], ms, pus, tz)
        hip.plan_destroy(plan)
        for r in dref:
            hip.ref_destroy(r)
        for ctu, pu, x, y, pw, ph in pus_l:
            if not g["active_%d" % k][ctu, pu] or x + pw > w or y + ph > h:
                continue
            where = (k, ctu, pu, x, y, pw, ph)
            assert pus[ctu, pu]["refi"].tolist() == g["out_refi_%d" % k][ctu, pu].tolist(), where
            assert pus[ctu, pu]["mv"].tolist() == g["out_mv_%d" % k][ctu, pu].tolist(), where
            assert pus[ctu, pu]["mot_bits"].tolist() == g["out_mot_%d" % k][ctu, pu].tolist(), where
            assert ms[ctu, pu].tolist() == g["out_mv_scale_%d" % k][ctu, pu].tolist(), where
            checked += 1
    assert checked > 1500
    assert hip.error_count() == 0


@pytest.mark.skipif(not RefLib.available(8), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("bd", [8, 10])
def test_analyze_bi_and_smvd_live(bd):
    """Random single PUs of every shape, lists of 1-2 references, every AMVR
    precision and cost type: the oracle's analyze_bi and analyze_smvd equal
    the compiled reference (C+SSE+AVX2 table) on seeds the fixtures do not
    hold."""
    o, ref = Oracle(bd), RefLib(bd, 2)
    rng = np.random.default_rng(40 + bd)
    top = (1 << bd) - 1
    H, W = 96, 160
    yy, xx = np.mgrid[0:H, 0:W]

    def pic(dx, dy):
        b = 0.5 + 0.25 * np.sin((xx - dx) / 6.0) * np.cos((yy - dy) / 4.0) + 0.1 * np.sin((xx + yy) / 2.5)
        return np.clip(np.rint(b * top + rng.normal(0, top / 120, (H, W))), 0, top).astype(
            np.uint8 if bd == 8 else np.uint16)

    cur, refs = pic(0, 0), [pic(2 * k + 1, -k) for k in range(4)]
    planes = [o.make_planes(r) for r in refs]
    pics = [ref.pic_create(r, p) for r, p in zip(refs, (-1, -2, 1, 2))]
    done = 0
    while done < 40:
        w, h = SHAPES[int(rng.integers(len(SHAPES)))]
        if w > W or h > H:
            continue
        x, y = int(rng.integers(0, W - w + 1)), int(rng.integers(0, H - h + 1))
        n0, n1 = int(rng.integers(1, 3)), int(rng.integers(1, 3))
        mvr, ct, lr = int(rng.integers(0, 5)), int(rng.integers(0, 2)), int(rng.integers(0, 2))
        gq = 1 << mvr
        refi = [int(rng.integers(0, n0)), int(rng.integers(0, n1))]
        mv = (rng.integers(-16, 17, (2, 2)) * gq).astype(int)
        mot = [int(v) for v in rng.integers(2, 30, 2)]
        mvp4, ms4 = np.zeros((2, 4, 2), np.int16), np.zeros((2, 4, 2), np.int16)
        mvp4[0, :n0], mvp4[1, :n1] = rng.integers(-24, 25, (n0, 2)) * gq, rng.integers(-24, 25, (n1, 2)) * gq
        ms4[0, :n0], ms4[1, :n1] = rng.integers(-16, 17, (n0, 2)) * gq, rng.integers(-16, 17, (n1, 2)) * gq
        r_, m_, b_, s_ = ref.analyze_bi(cur, pics[:n0], pics[2:2 + n1], x, y, w, h, mvp4, ms4, refi, mv, lr, mot, mvr,
                                        ct, LAMBDA)
        ms = np.concatenate([ms4[0, :n0], ms4[1, :n1]]).astype(int).tolist()
        orefi, omv, omot = list(refi), mv.tolist(), list(mot)
        o.analyze_bi(cur, planes[:n0] + planes[2:2 + n1], n0, x, y, w, h,
                     np.concatenate([mvp4[0, :n0], mvp4[1, :n1]]).astype(int).tolist(), ms, orefi, omv, lr, omot,
                     mvr, ct, LAMBDA)
        where = (x, y, w, h, n0, n1, mvr, ct, lr)
        assert orefi == r_.tolist() and [list(v) for v in omv] == m_.tolist() and omot == b_.tolist(), where
        assert [list(v) for v in ms] == np.concatenate([s_[0, :n0], s_[1, :n1]]).tolist(), where
        mvp2 = (rng.integers(-12, 13, (2, 2)) * gq).astype(np.int16)
        ms0 = (rng.integers(-12, 13, 2) * gq).astype(np.int16)
        m_, c_ = ref.smvd(cur, pics[0], pics[2], x, y, w, h, mvp2, ms0, mvr, ct, LAMBDA)
        mv2, c2 = o.analyze_smvd(cur, [planes[0], planes[2]], 1, x, y, w, h, mvp2.astype(int),
                                 np.array([ms0, [0, 0]]).astype(int), mvr, ct, LAMBDA)
        assert (mv2, c2) == (m_.tolist(), c_), where
        done += 1
    for p in pics:
        ref.pic_free(p)
