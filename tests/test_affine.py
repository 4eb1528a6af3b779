"""Affine luma motion compensation (com_mc_blk_affine_luma,
com_mc.c:202-250, and the luma of com_mc_cu_affine with three control
points): the oracle pinned against the compiled reference, and the GPU path
(me_affine_mc_luma) against the oracle."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, RefLib
import uavs3e_amd as ua

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(bd):
    return np.load(os.path.join(GOLD, "affine_bd%d.npz" % bd), allow_pickle=False)


@pytest.mark.parametrize("bd", [8, 10])
def test_affine_oracle_golden(bd):
    g, o = _golden(bd), Oracle(bd)
    pic = g["pic"]
    h, w = pic.shape
    planes = o.make_planes(pic)
    for c in g["cases"]:
        x, y, cw, ch, sw, sh = (int(v) for v in c[:6])
        got = o.affine_luma(planes, w, h, x, y, cw, ch, c[6:10].reshape(2, 2), sw, sh)
        off = int(c[10])
        assert np.array_equal(got.reshape(-1), g["preds"][off:off + cw * ch]), c.tolist()


@pytest.mark.parametrize("bd", [8, 10])
def test_affine3_oracle_golden(bd):
    """com_mc_cu_affine, luma and chroma, two and three control points."""
    g, o = _golden(bd), Oracle(bd)
    pic = g["pic"]
    h, w = pic.shape
    planes = o.make_planes(pic)
    u, v = pic[::2, ::2].copy(), pic[1::2, 1::2].copy()
    cuv = (o.make_chroma(u), o.make_chroma(v))
    for c in g["cases3"]:
        x, y, cw, ch, s8, ncp = (int(q) for q in c[:6])
        sub = 8 if s8 else 4
        cp = c[6:12].reshape(3, 2)[:ncp]
        got = o.affine_luma(planes, w, h, x, y, cw, ch, cp, sub, sub)
        off, coff, ca = int(c[12]), int(c[13]), cw * ch // 4
        assert np.array_equal(got.reshape(-1), g["preds3"][off:off + cw * ch]), c.tolist()
        gu, gv = o.affine_chroma(cuv, w, h, x, y, cw, ch, cp, sub)
        assert np.array_equal(gu.reshape(-1), g["preds3_u"][coff:coff + ca]), c.tolist()
        assert np.array_equal(gv.reshape(-1), g["preds3_v"][coff:coff + ca]), c.tolist()


@pytest.mark.skipif(not RefLib.available(8), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("bd", [8, 10])
def test_affine3_oracle_live(bd):
    # the reference's C table: its 10-bit AVX2 two-pass chroma filter differs
    # from its own C path (a finding recorded with the chroma MC fixtures)
    o, r = Oracle(bd), RefLib(bd, 0)
    rng = np.random.default_rng(bd + 5)
    h, w = 96, 160
    pic = rng.integers(0, 1 << bd, (h, w)).astype(o.dt)
    u = rng.integers(0, 1 << bd, (h // 2, w // 2)).astype(o.dt)
    v = rng.integers(0, 1 << bd, (h // 2, w // 2)).astype(o.dt)
    planes, rp = o.make_planes(pic), r.pic_create(pic, 0)
    r.pic_set_chroma(rp, u, v)
    cuv = (o.make_chroma(u), o.make_chroma(v))
    for k in range(40):
        cw, ch = (16, 32, 64, 128)[k % 4], (16, 32, 64, 128)[(k // 4) % 4]
        x, y = int(rng.integers(0, w)) // 4 * 4, int(rng.integers(0, h)) // 4 * 4
        s8, ncp = k % 2, 2 + (k // 2) % 2
        cp = rng.integers(-2000, 2000, (ncp, 2))
        sub = 8 if s8 else 4
        ry, ru, rv = r.affine_cu(rp, x, y, cw, ch, cp, s8)
        where = (k, x, y, cw, ch, s8, cp.tolist())
        assert np.array_equal(o.affine_luma(planes, w, h, x, y, cw, ch, cp, sub, sub), ry), where
        gu, gv = o.affine_chroma(cuv, w, h, x, y, cw, ch, cp, sub)
        assert np.array_equal(gu, ru) and np.array_equal(gv, rv), where
    r.pic_free(rp)


@pytest.mark.skipif(not RefLib.available(8), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("bd", [8, 10])
def test_affine_oracle_live(bd):
    o, r = Oracle(bd), RefLib(bd, 2)
    rng = np.random.default_rng(bd + 3)
    h, w = 96, 160
    pic = rng.integers(0, 1 << bd, (h, w)).astype(o.dt)
    planes, rp = o.make_planes(pic), r.pic_create(pic, 0)
    for k in range(60):
        cw, ch = (16, 32, 64, 128)[k % 4], (16, 32, 64, 128)[(k // 4) % 4]
        x, y = int(rng.integers(-16, w)) // 4 * 4, int(rng.integers(-16, h)) // 4 * 4
        sw, sh = (4, 8)[k % 2], (4, 8)[(k // 2) % 2]
        cp = rng.integers(-3000, 3000, (2, 2))
        assert np.array_equal(o.affine_luma(planes, w, h, x, y, cw, ch, cp, sw, sh),
                              r.affine_luma(rp, x, y, cw, ch, cp, sw, sh)), (k, x, y, cw, ch, sw, sh, cp.tolist())
    r.pic_free(rp)


def test_affine_oracle_special_cases():
    """Equal control points at a whole-sample MV copy the displaced block;
    at a quarter-sample MV every sub-block equals the 1/16 filter at phase 4;
    the 1/16 table's phase 0 is the identity and every phase sums to 64."""
    o = Oracle(8)
    rng = np.random.default_rng(2)
    h, w = 64, 96
    pic = rng.integers(0, 256, (h, w)).astype(np.uint8)
    planes = o.make_planes(pic)
    got = o.affine_luma(planes, w, h, 16, 16, 32, 16, [[32, -16], [32, -16]], 4, 4)
    assert np.array_equal(got, pic[15:31, 18:50])
    for f in range(16):
        t = o.luma_taps_hp(f)
        assert int(t.sum()) == 64
    assert o.luma_taps_hp(0).tolist() == [0, 0, 0, 64, 0, 0, 0, 0]
    got = o.affine_luma(planes, w, h, 16, 16, 16, 16, [[4, 0], [4, 0]], 8, 8)
    want = o.ipflt(0, planes[0], w + 2 * 144, (144 + 16) * (w + 2 * 144) + 144 + 16, 16, 16, o.luma_taps_hp(4), 255)
    assert np.array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_affine_gpu_vs_oracle(bd):
    """The golden cases plus random CUs with large control-point MVs (reaching
    the clipped picture border), every CU size, all four sub-block shapes."""
    g, o = _golden(bd), Oracle(bd)
    hip = ua.MeHip(bd)
    pic = g["pic"]
    h, w = pic.shape
    planes = o.make_planes(pic)
    rng = np.random.default_rng(bd)
    # (x, y, w, h, sub_w, sub_h, control points)
    rows = [[int(v) for v in c[:6]] + [c[6:10].reshape(2, 2)] for c in g["cases"]]
    rows += [[int(c[0]), int(c[1]), int(c[2]), int(c[3])] + [8 if c[4] else 4] * 2 + [c[6:12].reshape(3, 2)[:int(c[5])]]
             for c in g["cases3"]]
    for k in range(64):
        cw, ch = (16, 32, 64, 128)[k % 4], (16, 32, 64, 128)[(k // 4) % 4]
        rows.append([int(rng.integers(0, w)) // 4 * 4, int(rng.integers(0, h)) // 4 * 4, cw, ch, (4, 8)[k % 2],
                     (4, 8)[(k // 2) % 2], rng.integers(-6000, 6000, (2 + (k // 8) % 2, 2))])
    cus = np.zeros(len(rows), ua.AFFINE_CU)
    off = coff = 0
    for i, r in enumerate(rows):
        cus[i]["x"], cus[i]["y"], cus[i]["w"], cus[i]["h"], cus[i]["sub_w"], cus[i]["sub_h"] = r[:6]
        cus[i]["vertex_num"] = len(r[6])
        cus[i]["cpmv"][:len(r[6])] = r[6]
        cus[i]["pred_offset"] = off
        cus[i]["chroma_offset"] = coff
        off += r[2] * r[3]
        coff += r[2] * r[3] // 4
    pred = np.zeros(off + 7, hip.dt)
    pred[-7:] = 77
    u = rng.integers(0, 1 << bd, (h // 2, w // 2)).astype(hip.dt)
    v = rng.integers(0, 1 << bd, (h // 2, w // 2)).astype(hip.dt)
    ref = hip.ref_create(pic)
    hip.ref_load_chroma(ref, u, v)
    hip.affine_mc_luma(ref, cus, pred)
    sq = cus[cus["sub_w"] == cus["sub_h"]]          # chroma: square sub-blocks
    pu = np.zeros(coff + 5, hip.dt)
    pv = np.zeros(coff + 5, hip.dt)
    hip.affine_mc_chroma(ref, sq, pu, pv)
    hip.ref_destroy(ref)
    for c in cus:
        o_ = int(c["pred_offset"])
        want = o.affine_luma(planes, w, h, int(c["x"]), int(c["y"]), int(c["w"]), int(c["h"]),
                             c["cpmv"][:int(c["vertex_num"])], int(c["sub_w"]), int(c["sub_h"]))
        assert np.array_equal(pred[o_:o_ + c["w"] * c["h"]], want.reshape(-1)), c
    assert (pred[-7:] == 77).all()
    cuv = (o.make_chroma(u), o.make_chroma(v))
    for c in sq:
        co, ca = int(c["chroma_offset"]), int(c["w"] * c["h"] // 4)
        wu, wv = o.affine_chroma(cuv, w, h, int(c["x"]), int(c["y"]), int(c["w"]), int(c["h"]),
                                 c["cpmv"][:int(c["vertex_num"])], int(c["sub_w"]))
        assert np.array_equal(pu[co:co + ca], wu.reshape(-1)) and np.array_equal(pv[co:co + ca], wv.reshape(-1)), c
    assert len(sq) > 40
    assert hip.error_count() == 0


@pytest.mark.gpu
def test_affine_rejects_bad_arguments():
    """Unsupported CU / sub-block sizes, control-point counts, offsets past
    the buffer and chroma without chroma planes are errors; the reference
    stays usable."""
    hip = ua.MeHip(8)
    pic = np.random.default_rng(1).integers(0, 256, (64, 96)).astype(np.uint8)
    ref = hip.ref_create(pic)
    good = np.zeros(1, ua.AFFINE_CU)
    good[0]["w"] = good[0]["h"] = 16
    good[0]["sub_w"] = good[0]["sub_h"] = 4
    good[0]["vertex_num"] = 2
    pred = np.zeros(256, np.uint8)
    for field, value in (("w", 8), ("h", 24), ("sub_w", 2), ("vertex_num", 4), ("pred_offset", 1)):
        bad = good.copy()
        bad[0][field] = value
        with pytest.raises(ua.MeError):
            hip.affine_mc_luma(ref, bad, pred)
    pu, pv = np.zeros(64, np.uint8), np.zeros(64, np.uint8)
    with pytest.raises(ua.MeError):
        hip.affine_mc_chroma(ref, good, pu, pv)          # no chroma loaded
    hip.ref_load_chroma(ref, pic[::2, ::2].copy(), pic[1::2, 1::2].copy())
    bad = good.copy()
    bad[0]["sub_w"] = 8                                   # chroma needs square sub-blocks
    with pytest.raises(ua.MeError):
        hip.affine_mc_chroma(ref, bad, pu, pv)
    hip.affine_mc_luma(ref, good, pred)
    hip.affine_mc_chroma(ref, good, pu, pv)
    assert np.array_equal(pred.reshape(16, 16), pic[:16, :16])
    hip.ref_destroy(ref)
