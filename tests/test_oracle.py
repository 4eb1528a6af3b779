"""Pins the oracle (the C restatement the GPU path is checked against).

Three independent kinds of evidence, chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails at least one of them:
  1. golden vectors written by the compiled reference (tests/golden/ref_bd*.npz);
  2. a live cross-check against oracle/_ref where it was built;
  3. closed forms and brute force written differently from the oracle
     (Hadamard matrices instead of butterflies, numpy convolution instead of
     tap loops, the reference's procedural MV-bits loop instead of the closed
     form, exhaustive search instead of the raster scan).
"""
import math
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, RefLib, SHAPES, MASK_ALL, MASK_16x16, PAD

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SATD_SLOTS = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 1), (1, 2)]


def golden(bd):
    return np.load(os.path.join(GOLD, "ref_bd%d.npz" % bd), allow_pickle=False)


@pytest.fixture(scope="module", params=[8, 10])
def bd(request):
    return request.param


# --------------------------------------------------------------------------
# 1. golden vectors from the compiled reference
# --------------------------------------------------------------------------

def test_golden_block_costs(bd):
    g, o = golden(bd), Oracle(bd)
    A, B, S = g["blk_a"], g["blk_b"], g["blk_a"].shape[1]
    for k, (w, h, oo, po) in enumerate(g["sad_cases"].tolist()):
        assert o.sad(w, h, A, S, B, S, oo, po) == g["sad"][k]
        assert o.had(w, h, A, S, B, S, oo, po) == g["had"][k], (w, h)
        assert [o.sad(w, h, A, S, B, S, oo, po + i) for i in range(4)] == g["sad_x4"][k].tolist()
        assert [o.sad(w, h, A, S, B, S, oo, po + 2 * i) for i in range(3)] == g["sad_x3"][k].tolist()
    maxv = (1 << bd) - 1
    Z = np.zeros((32, 32), o.dt)
    M = np.full((32, 32), maxv, o.dt)
    for k, (wl, hl, oo, po) in enumerate(g["satd_cases"].tolist()):
        if oo == -1:
            v = o.satd(wl, hl, M, 32, Z, 32)
        elif oo == -2:
            v = o.satd(wl, hl, Z, 32, M, 32)
        else:
            v = o.satd(wl, hl, A, S, B, S, oo, po)
        assert v == g["satd"][k], (wl, hl, oo)


def test_golden_filters(bd):
    g, o = golden(bd), Oracle(bd)
    src = g["flt_src"]
    maxv = (1 << bd) - 1
    pos = 0
    for k, (ext, typ, w, h, fx, fy) in enumerate(g["flt_cases"].tolist()):
        n = g["flt_len"][k]
        want = g["flt_out"][pos:pos + n].reshape(h, w)
        pos += n
        nt = 8 if ((not ext and typ in (0, 2)) or (ext and typ == 0)) else 4
        tap = o.luma_taps if nt == 8 else o.chroma_taps
        got = (o.ipflt_ext(typ, src, 160, 8 * 160 + 8, w, h, tap(fx), tap(fy), maxv) if ext
               else o.ipflt(typ, src, 160, 8 * 160 + 8, w, h, tap(fx), maxv))
        assert np.array_equal(got, want), (ext, typ, w, fx, fy)
    for i, w in enumerate((4, 8, 16, 32, 64, 128)):
        got = o.pel_avrg(w, 8, g["avg_s1"][:, :w].copy(), g["avg_s2"][:, :w].copy())
        assert np.array_equal(got, g["avg"][i][:, :w])


def test_golden_rate_tables(bd):
    g, o = golden(bd), Oracle(bd)
    for (w, h) in ((1920, 1080), (3840, 2160)):
        off, tab = o.mvbits(w, h)
        assert np.array_equal(tab, g["mvbits_%dx%d" % (w, h)])
    assert o.lambda_mv(32) == 314169


def test_golden_subpel_planes(bd):
    g, o = golden(bd), Oracle(bd)
    luma = g["planes_luma"]
    h, w = luma.shape
    planes = o.make_planes(luma)
    E = 16
    assert np.array_equal(planes[:, PAD - E:PAD + h + E, PAD - E:PAD + w + E], g["planes_crop"])
    # outside the computed band every plane is edge replicated
    for i in range(16):
        p = planes[i]
        assert np.all(p[:PAD - 4] == p[PAD - 4]) and np.all(p[PAD + h + 4:] == p[PAD + h + 3])
        assert np.all(p[:, :PAD - 4] == p[:, PAD - 4:PAD - 3]) and np.all(p[:, PAD + w + 4:] == p[:, PAD + w + 3:PAD + w + 4])


def test_golden_me_search_tz(bd):
    """The reference's own search (diamond / raster / refine / sub-pel, every
    AMVR precision, both sub-pel cost types, adaptive raster range)."""
    g, o = golden(bd), Oracle(bd)
    cur, ref = g["srch_cur"], g["srch_ref"]
    planes = o.make_planes(ref)
    for k, c in enumerate(g["tz_cases"].tolist()):
        x, y, w, h, refi, nref, cpoc, gop, msr, mvr, ct, adapt, lam, mx, my = c
        mv, cost = o.me_search_tz(cur, planes, x, y, w, h, refi, nref, cpoc, 4, gop, msr, mvr, ct, adapt, lam, (mx, my))
        assert list(mv) == g["tz_mv"][k].tolist() and cost == g["tz_cost"][k], (k, c)


def test_golden_subpel(bd):
    g, o = golden(bd), Oracle(bd)
    cur, ref = g["srch_cur"], g["srch_ref"]
    planes = o.make_planes(ref)
    for k, c in enumerate(g["sp_cases"].tolist()):
        x, y, w, h, mvr, ct, lam, ix, iy, px, py = c
        mv, cost = o.subpel(cur, planes, x, y, w, h, mvr, ct, lam, (ix, iy), (px, py))
        assert list(mv) == g["sp_mv"][k].tolist() and cost == g["sp_cost"][k], (k, c)


def test_golden_full_search_frame(bd):
    g, o = golden(bd), Oracle(bd)
    res, best = o.me_frame(g["fr_cur"], [o.make_planes(r) for r in g["fr_refs"]], 8, 314169, MASK_ALL)
    assert np.array_equal(res, g["fr_res"]) and np.array_equal(best, g["fr_best"])


def test_golden_records_reference_simd_divergence():
    """The reference's AVX2 10-bit 2-D chroma filter differs from its own C
    path on full-range random input; 8-bit tables agree everywhere."""
    assert len(golden(8)["simd_mismatch_index"]) == 0
    g10 = golden(10)
    kinds = set(g10["simd_mismatch_kind"].tolist())
    assert kinds <= {"filter"}
    for k in g10["simd_mismatch_index"].tolist():
        ext, typ = g10["flt_cases"][k][:2]
        assert ext == 1 and typ == 1   # ipflt_ext[IPFILTER_EXT_4] only


# --------------------------------------------------------------------------
# 2. live cross-check against the compiled reference (where present)
# --------------------------------------------------------------------------

@pytest.mark.skipif(not RefLib.available(8), reason="oracle/_ref not built (reference tree absent)")
def test_live_reference_random_sweep(bd):
    o, r = Oracle(bd), RefLib(bd, 0)
    rng = np.random.default_rng(77 + bd)
    S = 256
    A = rng.integers(0, 1 << bd, (S, S)).astype(o.dt)
    B = rng.integers(0, 1 << bd, (S, S)).astype(o.dt)
    for _ in range(300):
        w, h = SHAPES[int(rng.integers(0, 29))]
        oo = int(rng.integers(0, S - h)) * S + int(rng.integers(0, S - w))
        po = int(rng.integers(0, S - h)) * S + int(rng.integers(0, S - w))
        assert o.had(w, h, A, S, B, S, oo, po) == r.had(w, h, A, S, B, S, oo, po)
        assert o.sad(w, h, A, S, B, S, oo, po) == r.sad(w, h, A, S, B, S, oo, po)
    for (w, h) in ((72, 40), (33, 17)):
        luma = rng.integers(0, 1 << bd, (h, w)).astype(o.dt)
        pic = r.pic_create(luma)
        assert np.array_equal(r.pic_planes(pic, w, h), o.make_planes(luma))
        r.pic_free(pic)


# --------------------------------------------------------------------------
# 3. closed forms and brute force
# --------------------------------------------------------------------------

def hadamard(n):
    h = np.array([[1]])
    while h.shape[0] < n:
        h = np.block([[h, h], [h, -h]])
    return h


def satd_matrix(diff):
    """Normalised SATD of one tile from explicit Hadamard matrices."""
    th, tw = diff.shape
    s = int(np.abs(hadamard(th) @ diff @ hadamard(tw).T).sum())
    if (tw, th) == (4, 4):
        return (s + 1) >> 1
    if (tw, th) == (8, 8):
        return (s + 2) >> 2
    return int(s / math.sqrt(128.0 if tw * th == 128 else 32.0) * 2.0)


def test_satd_matches_hadamard_matrices(bd):
    o = Oracle(bd)
    rng = np.random.default_rng(5 + bd)
    for wl, hl in SATD_SLOTS:
        tw, th = 4 << wl, 4 << hl
        for _ in range(50):
            a = rng.integers(0, 1 << bd, (th, tw)).astype(o.dt)
            b = rng.integers(0, 1 << bd, (th, tw)).astype(o.dt)
            want = satd_matrix(a.astype(np.int64) - b.astype(np.int64))
            assert o.satd(wl, hl, a, tw, b, tw) == want
    # com_had tiling: sum of tile SATDs, then >> (bd - 8)
    for (w, h) in SHAPES:
        a = rng.integers(0, 1 << bd, (h, w)).astype(o.dt)
        b = rng.integers(0, 1 << bd, (h, w)).astype(o.dt)
        if w == h:
            tw = th = 4 if w == 4 else 8
        elif w > h:
            tw, th = (8, 4) if h == 4 else (16, 8)
        else:
            tw, th = (4, 8) if w == 4 else (8, 16)
        d = a.astype(np.int64) - b.astype(np.int64)
        tot = sum(satd_matrix(d[y:y + th, x:x + tw]) for y in range(0, h, th) for x in range(0, w, tw))
        assert o.had(w, h, a, w, b, w) == tot >> (bd - 8), (w, h)


def test_sad_brute_force(bd):
    o = Oracle(bd)
    rng = np.random.default_rng(9)
    a = rng.integers(0, 1 << bd, (64, 140)).astype(o.dt)
    b = rng.integers(0, 1 << bd, (64, 140)).astype(o.dt)
    for (w, h) in SHAPES:
        if h > 64:
            continue
        assert o.sad(w, h, a, 140, b, 140) == int(np.abs(a[:h, :w].astype(np.int64) - b[:h, :w]).sum())


def test_filters_closed_forms(bd):
    o = Oracle(bd)
    maxv = (1 << bd) - 1
    rng = np.random.default_rng(3)
    src = rng.integers(0, 1 << bd, (48, 48)).astype(o.dt)
    const = np.full((48, 48), maxv // 3, o.dt)
    off = 8 * 48 + 8
    for frac in range(4):
        c = o.luma_taps(frac)
        assert int(c.sum()) == 64
        # constant in -> constant out, for every phase and both passes
        assert np.all(o.ipflt(0, const, 48, off, 16, 8, c, maxv) == maxv // 3)
        assert np.all(o.ipflt_ext(0, const, 48, off, 16, 8, c, o.luma_taps(3 - frac), maxv) == maxv // 3)
        # horizontal = numpy correlation of the 8 taps, rounded and clipped
        want = np.zeros((8, 16), np.int64)
        for k in range(8):
            want += c[k].astype(np.int64) * src[8:16, 8 - 3 + k:8 - 3 + k + 16]
        assert np.array_equal(o.ipflt(0, src, 48, off, 16, 8, c, maxv), np.clip((want + 32) >> 6, 0, maxv))
    # the integer phase is a copy
    assert np.array_equal(o.ipflt(2, src, 48, off, 16, 8, o.luma_taps(0), maxv), src[8:16, 8:24])
    for frac in range(8):
        assert int(o.chroma_taps(frac).sum()) == 64


def test_planes_agree_with_block_filters(bd):
    """Each quarter-pel plane sample inside the picture equals the per-block
    filter at that phase (com_mc.c:668-758 vs 461-549)."""
    o = Oracle(bd)
    maxv = (1 << bd) - 1
    luma = np.random.default_rng(11).integers(0, 1 << bd, (24, 32)).astype(o.dt)
    planes = o.make_planes(luma)
    st = planes.shape[2]
    base = planes[0].ravel()
    off = PAD * st + PAD
    for dy in range(4):
        for dx in range(4):
            if dx == 0 and dy == 0:
                continue
            if dy == 0:
                blk = o.ipflt(0, base, st, off, 32, 24, o.luma_taps(dx), maxv)
            elif dx == 0:
                blk = o.ipflt(2, base, st, off, 32, 24, o.luma_taps(dy), maxv)
            else:
                blk = o.ipflt_ext(0, base, st, off, 32, 24, o.luma_taps(dx), o.luma_taps(dy), maxv)
            assert np.array_equal(planes[dy * 4 + dx][PAD:PAD + 24, PAD:PAD + 32], blk), (dy, dx)


def test_mvbits_procedural_form():
    """The reference's procedure (inter.c:2036-2054): seed 1/2/3 bits, then the
    exp-Golomb loop overwrites entries with 2*exp_bits up to the offset."""
    o = Oracle(8)
    for (w, h) in ((64, 48), (1920, 1080)):
        off = (max(w, h) << 2) + 3
        tab = np.zeros(2 * off + 1, np.int64)
        tab[off] = 1
        tab[off - 1] = tab[off + 1] = 2
        tab[off - 2] = tab[off + 2] = 3
        e = 2
        while True:
            imax, imin = min((1 << e) - 1, off), (1 << (e - 1)) - 1
            for i in range(imin, imax):
                tab[off - i] = tab[off + i] = 2 * e
            if imax == off:
                break
            e += 1
        got_off, got = o.mvbits(w, h)
        assert got_off == off and np.array_equal(got, tab)


def test_full_search_is_global_raster_minimum(bd):
    """Exhaustive numpy search of the u64 cost; ties to the first raster
    position; window clipped like get_max_search_range."""
    o = Oracle(bd)
    rng = np.random.default_rng(21)
    W, H = 160, 136
    cur = rng.integers(0, 1 << bd, (H, W)).astype(o.dt)
    ref = np.roll(cur, (1, -2), (0, 1)).copy()
    ref[::7] = rng.integers(0, 1 << bd, ref[::7].shape)
    planes = o.make_planes(ref)
    lam = 314169
    R = 5
    res, _ = o.me_frame(cur, [planes], R, lam, MASK_16x16)
    off, tab = o.mvbits(W, H)
    shift = 16 - (bd - 8)
    T = max(W, H)
    P = planes[0].astype(np.int64)
    for ctu in range(res.shape[0]):
        cx, cy = (ctu % 2) * 128, (ctu // 2) * 128
        for pu in range(64):
            x, y = cx + (pu % 8) * 16, cy + (pu // 8) * 16
            if x + 16 > W or y + 16 > H:
                assert res[ctu, pu, 0]["valid"] == 0
                continue
            best = None
            for my in range(max(-R, -127 - y, -T + 4), min(R, H - 1 - y, T - 4) + 1):
                for mx in range(max(-R, -127 - x, -T + 4), min(R, W - 1 - x, T - 4) + 1):
                    blk = P[PAD + y + my:PAD + y + my + 16, PAD + x + mx:PAD + x + mx + 16]
                    sad = int(np.abs(cur[y:y + 16, x:x + 16].astype(np.int64) - blk).sum())
                    cost = (sad << shift) + (int(tab[off + 4 * mx]) + int(tab[off + 4 * my])) * lam
                    if best is None or cost < best[0]:
                        best = (cost, mx, my)
            r = res[ctu, pu, 0]
            assert (int(r["cost_int"]), int(r["mv_int"][0]), int(r["mv_int"][1])) == best


def test_final_cost_and_refi_bits():
    o = Oracle(8)
    table = [[0], [0], [1, 1], [1, 2, 2], [1, 2, 3, 3]]   # tbl_refi_bits rows 0..4 (tables.c:52-57)
    for n, row in enumerate(table):
        for i, v in enumerate(row):
            assert o.lib.orc_refi_bits(n, i) == v
    # (c + lambda*(refi_bits + mvr + (mvr<4) + 2) + 2^15) >> 16
    assert o.lib.orc_final_cost(10 << 16, 314169, 2, 1, 0) == (((10 << 16) + 314169 * 4) + 32768) >> 16


# --------------------------------------------------------------------------
# Motion compensation and reference lists
# --------------------------------------------------------------------------

MC_W, MC_H = 160, 96


def mc_golden(bd):
    return np.load(os.path.join(GOLD, "mc_bd%d.npz" % bd), allow_pickle=False)


def test_golden_mc_pu(bd):
    """orc_mc_pu against com_mc_cu of the compiled reference (C table):
    luma copies from the sub-pel planes, chroma 1/8-pel 4-tap filters,
    clipped border positions, bi-averaging, and both lists on one picture
    with one MV."""
    d = mc_golden(bd)
    o = Oracle(bd)
    planes = [o.make_planes(l) for l in d["luma"]]
    chroma = [(o.make_chroma(u), o.make_chroma(v)) for u, v in zip(d["u"], d["v"])]
    preds, k = d["preds"], 0
    kinds = set()
    for x, y, w, h, nl, r0, m0x, m0y, r1, m1x, m1y in d["cases"].tolist():
        ch = w >= 8 and h >= 8
        rs, mvs = [r0, r1][:nl], [(m0x, m0y), (m1x, m1y)][:nl]
        py, pu, pv = o.mc_pu(MC_W, MC_H, [planes[r] for r in rs], [chroma[r] for r in rs] if ch else None,
                             x, y, w, h, mvs)
        for a in ([py, pu, pv] if ch else [py]):
            n = a.size
            assert np.array_equal(a.ravel(), preds[k:k + n]), (x, y, w, h, nl, mvs)
            k += n
        kinds.add((nl, (m0x & 7) != 0 and (m0y & 7) != 0))
    assert k == preds.size and kinds == {(1, False), (1, True), (2, False), (2, True)}
    if bd == 10:
        assert len(d["simd_mismatch"]) > 0   # the reference's AVX2 chroma filter differs from its C path


def test_mc_frame_layout_and_lists(bd):
    """orc_mc_frame writes orc_mc_pu of each PU's best L0 / L1 reference at
    its offset in the frame layout, and orc_me_frame_lists picks the best
    reference per list with the list's own tbl_refi_bits row."""
    rng = np.random.default_rng(40 + bd)
    dt = np.uint8 if bd == 8 else np.uint16
    o = Oracle(bd)
    w, h = 136, 136
    lumas = [rng.integers(0, 1 << bd, (h, w)).astype(dt) for _ in range(4)]
    cur = np.clip(lumas[0].astype(int) + rng.integers(-3, 4, (h, w)), 0, (1 << bd) - 1).astype(dt)
    planes = [o.make_planes(l) for l in lumas]
    chroma = [(o.make_chroma(rng.integers(0, 1 << bd, (h // 2, w // 2)).astype(dt)),
               o.make_chroma(rng.integers(0, 1 << bd, (h // 2, w // 2)).astype(dt))) for _ in range(4)]
    mask = sum(1 << SHAPES.index(s) for s in [(4, 8), (8, 8), (16, 8), (32, 32)])
    lam = o.lambda_mv(32)
    res, best = o.me_frame(cur, planes, 3, lam, mask, num_refs_l1=2)
    res1, best1 = o.me_frame(cur, planes, 3, lam, mask)
    assert np.array_equal(res[["mv_sub", "cost_sub"]], res1[["mv_sub", "cost_sub"]])
    valid = res["valid"][..., 0] == 1
    for l, (r0, n) in enumerate([(0, 2), (2, 2)]):
        fc = np.stack([res["cost_final"][..., r0 + i] for i in range(n)], -1).astype(np.int64)
        for i in range(n):
            want = [o.lib.orc_final_cost(int(c), lam, n, i, 0) for c in res["cost_sub"][..., r0 + i][valid]]
            assert (fc[..., i][valid] == np.array(want)).all()
        assert (best[..., l][valid] == r0 + np.argmin(fc, -1)[valid]).all()   # argmin takes the first on ties
    py, pu, pv = o.mc_frame(w, h, planes, chroma, 2, mask, res, best)
    nsel, nsel_c = 4, 3
    n_ctu = res.shape[0]
    pu_i = 0
    for s in [SHAPES.index(t) for t in [(4, 8), (8, 8), (16, 8), (32, 32)]]:
        sw, sh = SHAPES[s]
        rank = sorted([SHAPES.index(t) for t in [(4, 8), (8, 8), (16, 8), (32, 32)]]).index(s)
        rank_c = rank - 1 if sw >= 8 else None
        for i, (oy, ox) in enumerate([(a, b) for a in range(0, 128, sh) for b in range(0, 128, sw)]):
            for ctu in range(n_ctu):
                cx, cy = (ctu % 2) * 128, (ctu // 2) * 128
                b = best[ctu, pu_i + i]
                off = ctu * nsel * 16384 + rank * 16384 + i * sw * sh
                got = py[off:off + sw * sh].reshape(sh, sw)
                if (b < 0).any():
                    assert not got.any()
                    continue
                mvs = [tuple(int(v) for v in res[ctu, pu_i + i, r]["mv_sub"]) for r in b]
                ey, eu, ev = o.mc_pu(w, h, [planes[r] for r in b], [chroma[r] for r in b] if sw >= 8 else None,
                                     cx + ox, cy + oy, sw, sh, mvs)
                assert np.array_equal(got, ey), (s, i, ctu)
                if sw >= 8:
                    co = ctu * nsel_c * 4096 + rank_c * 4096 + i * (sw // 2) * (sh // 2)
                    assert np.array_equal(pu[co:co + eu.size].reshape(eu.shape), eu)
                    assert np.array_equal(pv[co:co + ev.size].reshape(ev.shape), ev)
        pu_i += (128 // sw) * (128 // sh)


@pytest.mark.skipif(not RefLib.available(8), reason="oracle/_ref not built (reference tree absent)")
def test_live_reference_mc_sweep(bd):
    """orc_mc_pu against com_mc_cu of the compiled reference's C table on
    fresh random pictures, every shape, uni and bi."""
    rng = np.random.default_rng(77 + bd)
    dt = np.uint8 if bd == 8 else np.uint16
    o, ref = Oracle(bd), RefLib(bd, level=0)
    w, h = 200, 136
    lumas = [rng.integers(0, 1 << bd, (h, w)).astype(dt) for _ in range(2)]
    cs = [(rng.integers(0, 1 << bd, (h // 2, w // 2)).astype(dt), rng.integers(0, 1 << bd, (h // 2, w // 2)).astype(dt))
          for _ in range(2)]
    pics = []
    for i in range(2):
        p = ref.pic_create(lumas[i], poc=i)
        ref.pic_set_chroma(p, *cs[i])
        pics.append(p)
    planes = [o.make_planes(l) for l in lumas]
    chroma = [(o.make_chroma(u), o.make_chroma(v)) for u, v in cs]
    try:
        for pw, ph in SHAPES:
            for nl in (1, 2):
                x = int(rng.integers(0, max(1, (w - pw) // 4 + 1))) * 4
                y = int(rng.integers(0, max(1, (h - ph) // 4 + 1))) * 4
                mvs = [tuple(int(v) for v in rng.integers(-300, 300, 2)) for _ in range(2)]
                chf = pw >= 8 and ph >= 8
                got = o.mc_pu(w, h, planes[:nl], chroma[:nl] if chf else None, x, y, pw, ph, mvs[:nl])
                want = ref.mc_cu(x, y, w, h, pw, ph, pics[0], pics[1] if nl == 2 else None, mvs[0], mvs[1], chroma=chf)
                for a, b in zip(got, want):
                    assert (a is None and b is None) or np.array_equal(a, b), (pw, ph, nl)
    finally:
        for p in pics:
            ref.pic_free(p)


def test_orc_par_equals_whole_frame():
    """The threaded per-(shape, CTU) split of the oracle's frame search that
    the GPU configuration tests use returns exactly the whole-frame call."""
    import orc_par
    from uavs3e_amd import workloads as wl
    work = wl.fuhd10_bi().scaled(136, 72, 6)
    cur, refs = work.frames()
    orc = Oracle(10)
    planes = [orc.make_planes(r) for r in refs]
    lam = wl.lambda_mv(32)
    mask = MASK_ALL & ~(1 << SHAPES.index((8, 8)))
    for l1 in (0, 2):
        a = orc.me_frame(cur, planes, 6, lam, mask, num_refs_l1=l1)
        b = orc_par.me_frame(orc, cur, planes, 6, lam, mask, num_refs_l1=l1)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
