"""Affine gradient motion search (affine_me_gradient, inter.c:1339-1471):
the oracle's restatement pinned against the compiled reference's own
function (golden vectors, tests/golden/affine_me_bd*.npz, and live on other
seeds), and the GPU path (me_affine_me_gradient) against the golden vectors
directly and against the oracle on a larger batch.

The search solves a 4x4 least-squares system in IEEE double per iteration;
the results are compared bit-exactly (costs, control points): the system's
sums are exact 64-bit integers and the solve follows the reference's
operation order without contraction, so no tolerance is needed.  The
fixtures include CUs on a flat block and on vertical stripes, whose systems
are singular (the solve divides by zero)."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, RefLib
import uavs3e_amd as ua

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LAMBDA = 314169


def _golden(bd):
    return np.load(os.path.join(GOLD, "affine_me_bd%d.npz" % bd), allow_pickle=False)


def _case(row):
    x, y, w, h, sw, sh, refi, nref, mvr, bi, mb = (int(v) for v in row[:11])
    mvp = np.array(row[11:15], int).reshape(2, 2)
    mv = np.array(row[15:19], int).reshape(2, 2)
    return x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, mvp, mv, int(row[19])


@pytest.mark.parametrize("bd", [8, 10])
def test_affine_me_oracle_golden(bd):
    g, o = _golden(bd), Oracle(bd)
    ref = g["ref"]
    H, W = ref.shape
    planes = o.make_planes(ref)
    for row, c_want, mv_want in zip(g["cases"], g["cost"], g["mv"]):
        x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, mvp, mv, off = _case(row)
        org = g["orgs"][off:off + w * h].reshape(h, w)
        c, m = o.affine_me_gradient(planes, W, H, org, x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, LAMBDA, mvp, mv)
        assert (c, np.array(m).reshape(-1).tolist()) == (int(c_want), mv_want.tolist()), row.tolist()


@pytest.mark.skipif(not RefLib.available(8), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("bd", [8, 10])
def test_affine_me_oracle_live(bd):
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_golden_affine_me", os.path.join(GOLD, "gen_golden_affine_me.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    cur, refimg = m.pictures(bd, 7)
    rows, orgs = m.cases(bd, 8, cur)
    o, R = Oracle(bd), RefLib(bd, 2)
    H, W = refimg.shape
    planes = o.make_planes(refimg)
    pic = R.pic_create(refimg, 0)
    for row in rows[:80]:
        x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, mvp, mv, off = _case(row)
        org = orgs[off:off + w * h].reshape(h, w)
        want = R.affine_me_gradient(pic, org, x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, LAMBDA, mvp, mv)
        got = o.affine_me_gradient(planes, W, H, org, x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, LAMBDA, mvp, mv)
        assert got == want, row.tolist()
    R.pic_free(pic)


def _cus(rows):
    cus = np.zeros(len(rows), ua.AFFINE_ME_CU)
    for i, row in enumerate(rows):
        x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, mvp, mv, off = _case(row)
        cus[i] = (x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, w, mvp, mv, off, 0)
    return cus


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_affine_me_gpu_golden(bd):
    """Every fixture case in one batch: costs and control points equal the
    compiled reference's."""
    g = _golden(bd)
    hip = ua.MeHip(bd)
    ref = hip.ref_create(g["ref"])
    cus = _cus(g["cases"])
    hip.affine_me_gradient(ref, g["orgs"], cus, LAMBDA)
    hip.ref_destroy(ref)
    for i, (c_want, mv_want) in enumerate(zip(g["cost"], g["mv"])):
        assert int(cus[i]["cost"]) == int(c_want), (i, g["cases"][i].tolist())
        assert cus[i]["mv"].reshape(-1).tolist() == mv_want.tolist(), (i, g["cases"][i].tolist())
    assert hip.error_count() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_affine_me_gpu_vs_oracle(bd):
    """A larger batch on other seeds (uni and bi, all CU sizes, sub-blocks,
    AMVR indices), the CUs' originals taken from the current picture itself
    (org_stride = picture width) for uni-prediction."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_golden_affine_me", os.path.join(GOLD, "gen_golden_affine_me.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    cur, refimg = m.pictures(bd, 17)
    rows, orgs = m.cases(bd, 18, cur)
    rows, orgs = np.concatenate([rows, m.cases(bd, 19, cur)[0]]), orgs
    H, W = cur.shape
    o, hip = Oracle(bd), ua.MeHip(bd)
    planes = o.make_planes(refimg)
    # uni CUs read the picture itself; bi CUs their targets, appended after it
    buf = [cur.reshape(-1)]
    off = cur.size
    cus = np.zeros(len(rows), ua.AFFINE_ME_CU)
    want = []
    top = (1 << bd) - 1
    shifted = np.roll(cur, (1, -1), axis=(0, 1)).astype(np.int32)
    for i, row in enumerate(rows):
        x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, mvp, mv, _ = _case(row)
        if bi:
            org = np.clip(2 * cur[y:y + h, x:x + w].astype(np.int32) - shifted[y:y + h, x:x + w], 0, top).astype(cur.dtype)
            cus[i] = (x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, w, mvp, mv, off, 0)
            buf.append(org.reshape(-1))
            off += w * h
        else:
            org = cur[y:y + h, x:x + w]
            cus[i] = (x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, W, mvp, mv, y * W + x, 0)
        want.append(o.affine_me_gradient(planes, W, H, org, x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, LAMBDA, mvp,
                                         mv))
    ref = hip.ref_create(refimg)
    hip.affine_me_gradient(ref, np.concatenate(buf), cus, LAMBDA)
    hip.ref_destroy(ref)
    for i, (c, mv) in enumerate(want):
        assert (int(cus[i]["cost"]), cus[i]["mv"].tolist()) == (c, mv), (i, rows[i].tolist())
    assert len(want) >= 300
    assert hip.error_count() == 0


@pytest.mark.gpu
def test_affine_me_rejects_bad_arguments():
    hip = ua.MeHip(8)
    pic = np.zeros((64, 64), np.uint8)
    ref = hip.ref_create(pic)
    base = (0, 0, 16, 16, 4, 4, 0, 1, 0, 0, 0, 64, [[0, 0], [0, 0]], [[0, 0], [0, 0]], 0, 0)
    for field, bad in (("w", 8), ("h", 24), ("sub_w", 2), ("mvr", 3), ("refi", 1), ("num_refp", 5), ("bi", 2),
                       ("org_stride", 8), ("org_offset", 64 * 64)):
        cus = np.zeros(1, ua.AFFINE_ME_CU)
        cus[0] = base
        cus[0][field] = bad
        with pytest.raises(ua.MeError):
            hip.affine_me_gradient(ref, pic.reshape(-1), cus, LAMBDA)
    cus = np.zeros(1, ua.AFFINE_ME_CU)
    cus[0] = base
    hip.affine_me_gradient(ref, pic.reshape(-1), cus, LAMBDA)   # flat: ends at once with the start points
    assert cus[0]["mv"].tolist() == [[0, 0], [0, 0]]
    hip.ref_destroy(ref)
