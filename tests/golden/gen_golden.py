"""Writes the golden vectors of the ME path from the compiled reference.

Run in the container where /root/reference exists, after oracle/build.py has
compiled it into oracle/_ref:

    python tests/golden/gen_golden.py

For each bit depth it writes tests/golden/ref_bd{8,10}.npz holding seeded
inputs and the reference's outputs for: SAD / SAD x3 / SAD x4 and the six
SATD slots through the dispatch table, com_had, every luma/chroma block
filter, the MV-bits table, the 15 quarter-pel planes of com_if_luma_frame,
me_search_tz (diamond/raster/refine/sub-pel) on synthetic pictures,
me_sub_pel_search from given integer MVs, and the full-search frame driver.
Table outputs are taken from the C table (uavs3e_funs_init_c) and checked
equal to the production C+SSE+AVX2 table, except where the reference's own
SIMD path differs from its C path (recorded under 'simd_mismatch_*').
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib, SHAPES, MASK_ALL, MASK_16x16, aligned_zeros  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
SATD_SLOTS = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 1), (1, 2)]


def smooth_picture(rng, h, w, bd):
    """Random low-frequency picture plus noise (so searches have real minima)."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    v = np.full((h, w), float(1 << (bd - 1)))
    for _ in range(4):
        a = rng.uniform(10, 40) * (1 if bd == 8 else 4)
        fx, fy = rng.uniform(-0.08, 0.08, 2)
        v += a * np.sin(fx * xx + fy * yy + rng.uniform(0, 6.28))
    v += rng.normal(0, 3 if bd == 8 else 12, (h, w))
    return np.clip(np.floor(v + 0.5), 0, (1 << bd) - 1)


def gen(bd):
    rng = np.random.default_rng(1000 + bd)
    dt = np.uint8 if bd == 8 else np.uint16
    maxv = (1 << bd) - 1
    ref_c = RefLib(bd, level=0)
    out = {}

    # ---- block costs -------------------------------------------------------
    S = 192
    A = rng.integers(0, maxv + 1, (S, S)).astype(dt)
    B = rng.integers(0, maxv + 1, (S, S)).astype(dt)
    out["blk_a"], out["blk_b"] = A, B
    cases = []   # (w, h, org_off, pred_off)
    for (w, h) in SHAPES + [(4, 4)]:
        for _ in range(3):
            cases.append((w, h, int(rng.integers(0, S - h)) * S + int(rng.integers(0, S - w)),
                          int(rng.integers(0, S - h - 1)) * S + int(rng.integers(0, S - w - 4))))
    cases = np.array(cases, np.int64)
    out["sad_cases"] = cases
    out["sad"] = np.array([ref_c.sad(w, h, A, S, B, S, o, p) for w, h, o, p in cases.tolist()], np.uint32)
    out["had"] = np.array([ref_c.had(w, h, A, S, B, S, o, p) for w, h, o, p in cases.tolist()], np.uint32)
    out["sad_x4"] = np.array([ref_c.sad_x4(w, h, A, S, B, S, [p + k for k in range(4)], org_off=o)
                              for w, h, o, p in cases.tolist()], np.uint32)
    out["sad_x3"] = np.array([ref_c.sad_x3(w, h, A, S, B, S, [p + 2 * k for k in range(3)], org_off=o)
                              for w, h, o, p in cases.tolist()], np.uint32)
    satd_cases = []
    for wl, hl in SATD_SLOTS:
        for _ in range(40):
            satd_cases.append((wl, hl, int(rng.integers(0, S - 16)) * S + int(rng.integers(0, S - 16)),
                               int(rng.integers(0, S - 16)) * S + int(rng.integers(0, S - 16))))
    # extreme residuals: maximal positive / negative differences
    for wl, hl in SATD_SLOTS:
        satd_cases.append((wl, hl, -1, -1))
        satd_cases.append((wl, hl, -2, -2))
    satd_cases = np.array(satd_cases, np.int64)
    Z = np.zeros((32, 32), dt)
    M = np.full((32, 32), maxv, dt)
    vals = []
    for wl, hl, o, p in satd_cases:
        if o == -1:
            vals.append(ref_c.satd(int(wl), int(hl), M, 32, Z, 32))
        elif o == -2:
            vals.append(ref_c.satd(int(wl), int(hl), Z, 32, M, 32))
        else:
            vals.append(ref_c.satd(int(wl), int(hl), A, S, B, S, int(o), int(p)))
    out["satd_cases"], out["satd"] = satd_cases, np.array(vals, np.uint32)

    # ---- block filters -------------------------------------------------------
    src = rng.integers(0, maxv + 1, (160, 160)).astype(dt)
    out["flt_src"] = src
    off = 8 * 160 + 8
    flt_cases, flt_out = [], []
    for typ in range(4):
        n = 8 if typ in (0, 2) else 4
        for w in (4, 8, 16, 32, 64) + ((128,) if n == 8 else ()):
            for frac in range(1, 4 if n == 8 else 8):
                h = [4, 8, 16][(w + frac) % 3]
                c = ref_c.luma_coef(frac) if n == 8 else ref_c.chroma_coef(frac)
                d = ref_c.ipflt(typ, src, 160, off, w, h, c, maxv)
                flt_cases.append((0, typ, w, h, frac, 0))
                flt_out.append(d.ravel())
    for typ in range(2):
        n = 8 if typ == 0 else 4
        for w in (4, 8, 16, 32, 64) + ((128,) if n == 8 else ()):
            for fx in range(1, 4 if n == 8 else 8):
                for fy in range(1, 4 if n == 8 else 8):
                    if (fx + fy + w) % 2:
                        continue
                    h = [4, 8, 16][(w + fx + fy) % 3]
                    cv = ref_c.luma_coef if n == 8 else ref_c.chroma_coef
                    d = ref_c.ipflt_ext(typ, src, 160, off, w, h, cv(fx), cv(fy), maxv)
                    flt_cases.append((1, typ, w, h, fx, fy))
                    flt_out.append(d.ravel())
    out["flt_cases"] = np.array(flt_cases, np.int64)
    out["flt_out"] = np.concatenate(flt_out).astype(dt)
    out["flt_len"] = np.array([len(x) for x in flt_out], np.int64)

    s1 = rng.integers(0, maxv + 1, (8, 128)).astype(dt)
    s2 = rng.integers(0, maxv + 1, (8, 128)).astype(dt)
    out["avg_s1"], out["avg_s2"] = s1, s2
    out["avg"] = np.stack([np.pad(ref_c.pel_avrg(w, 8, s1[:, :w].copy(), s2[:, :w].copy()), ((0, 0), (0, 128 - w)))
                           for w in (4, 8, 16, 32, 64, 128)])

    # ---- rate ----------------------------------------------------------------
    for (w, h) in ((1920, 1080), (3840, 2160)):
        o, t = ref_c.mvbits(w, h)
        out["mvbits_%dx%d" % (w, h)] = t

    # ---- frame planes --------------------------------------------------------
    pw, ph = 40, 24
    luma = rng.integers(0, maxv + 1, (ph, pw)).astype(dt)
    pic = ref_c.pic_create(luma)
    planes = ref_c.pic_planes(pic, pw, ph)
    ref_c.pic_free(pic)
    P, E = 144, 16
    out["planes_luma"] = luma
    out["planes_crop"] = planes[:, P - E:P + ph + E, P - E:P + pw + E].copy()

    # ---- searches on synthetic pictures --------------------------------------
    W, H = 192, 136
    cur = smooth_picture(rng, H, W, bd).astype(dt)
    refpic = np.roll(smooth_picture(rng, H, W, bd).astype(dt) // 2 + cur // 2, (3, -5), (0, 1)).astype(dt)
    out["srch_cur"], out["srch_ref"] = cur, refpic
    pic = ref_c.pic_create(refpic, 4)
    tz_cases, tz_out = [], []
    for i in range(160):
        w, h = SHAPES[int(rng.integers(0, len(SHAPES)))]
        if w > 64 or h > 64:
            w, h = 16, 16
        x = int(rng.integers(0, (W - w) // 4 + 1)) * 4
        y = int(rng.integers(0, (H - h) // 4 + 1)) * 4
        mvr = int(rng.choice([0, 0, 0, 1, 2, 3, 4]))
        q = 1 << mvr
        mvp = (int(rng.integers(-40, 41)) // q * q, int(rng.integers(-40, 41)) // q * q)
        cost_type = int(rng.integers(0, 2))
        num_refp = int(rng.integers(1, 5))
        refi = int(rng.integers(0, num_refp))
        cur_poc = 4 + refi
        adapt = int(rng.integers(0, 2))
        lam = int(rng.integers(100000, 700000))
        mv = np.zeros(2, np.int16)
        mvi = np.zeros(2, np.int16)
        mvp_a = np.array(mvp, np.int16)
        c = ref_c.lib.refh_me_search_tz(pic, cur.ctypes.data + (y * W + x) * cur.itemsize, W, x, y, w, h, refi,
                                        num_refp, cur_poc, 8, 64 << 3 if i % 2 else 384, mvr, cost_type, adapt, lam,
                                        mvp_a.ctypes.data, mv.ctypes.data, mvi.ctypes.data)
        tz_cases.append((x, y, w, h, refi, num_refp, cur_poc, 8, 64 << 3 if i % 2 else 384, mvr, cost_type, adapt,
                         lam, mvp[0], mvp[1]))
        tz_out.append((int(mv[0]), int(mv[1]), int(mvi[0]), int(mvi[1]), c))
    out["tz_cases"] = np.array(tz_cases, np.int64)
    out["tz_mv"] = np.array([t[:4] for t in tz_out], np.int64)
    out["tz_cost"] = np.array([t[4] for t in tz_out], np.uint64)

    sp_cases, sp_out = [], []
    for i in range(120):
        w, h = SHAPES[int(rng.integers(0, len(SHAPES)))]
        if w > 64 or h > 64:
            w, h = 8, 16
        x = int(rng.integers(0, (W - w) // 4 + 1)) * 4
        y = int(rng.integers(0, (H - h) // 4 + 1)) * 4
        mvr = int(rng.choice([0, 0, 1]))
        cost_type = int(rng.integers(0, 2))
        lam = int(rng.integers(100000, 700000))
        imv = (int(rng.integers(-20, 21)), int(rng.integers(-20, 21)))
        mv = np.array([imv[0] * 4, imv[1] * 4], np.int16)
        mvp = np.array([int(rng.integers(-20, 21)) * 2 // (1 << mvr) * (1 << mvr),
                        int(rng.integers(-20, 21)) * 2 // (1 << mvr) * (1 << mvr)], np.int16)
        c = ref_c.lib.refh_subpel(pic, cur.ctypes.data + (y * W + x) * cur.itemsize, W, x, y, w, h, mvr,
                                  cost_type, lam, mvp.ctypes.data, mv.ctypes.data)
        sp_cases.append((x, y, w, h, mvr, cost_type, lam, imv[0], imv[1], int(mvp[0]), int(mvp[1])))
        sp_out.append((int(mv[0]), int(mv[1]), c))
    out["sp_cases"] = np.array(sp_cases, np.int64)
    out["sp_mv"] = np.array([t[:2] for t in sp_out], np.int64)
    out["sp_cost"] = np.array([t[2] for t in sp_out], np.uint64)
    ref_c.pic_free(pic)

    # ---- full-search frame driver (the batched workloads' semantics) ---------
    lam = 314169
    fcur = smooth_picture(rng, 128, 128, bd).astype(dt)
    frefs = [np.roll(fcur, (2, -3), (0, 1)), smooth_picture(rng, 128, 128, bd).astype(dt)]
    pics = [ref_c.pic_create(r) for r in frefs]
    res, best, _ = ref_c.me_frame(fcur, pics, 8, lam, MASK_ALL)
    out["fr_cur"], out["fr_refs"] = fcur, np.stack(frefs)
    out["fr_res"], out["fr_best"] = res, best
    for p in pics:
        ref_c.pic_free(p)

    # ---- the production SIMD table must agree with the C table ---------------
    ref_s = RefLib(bd, level=2)
    mism = []
    for k, (w, h, o, p) in enumerate(cases.tolist()):
        if ref_s.sad(w, h, A, S, B, S, o, p) != out["sad"][k] or ref_s.had(w, h, A, S, B, S, o, p) != out["had"][k]:
            mism.append(("cost", k))
    pos = 0
    for k, (ext, typ, w, h, fx, fy) in enumerate(flt_cases):
        n = out["flt_len"][k]
        want = out["flt_out"][pos:pos + n]
        pos += n
        nt = 8 if ((not ext and typ in (0, 2)) or (ext and typ == 0)) else 4
        cv = ref_s.luma_coef if nt == 8 else ref_s.chroma_coef
        got = (ref_s.ipflt_ext(typ, src, 160, off, w, h, cv(fx), cv(fy), maxv) if ext
               else ref_s.ipflt(typ, src, 160, off, w, h, cv(fx), maxv)).ravel()
        if not np.array_equal(got, want):
            mism.append(("filter", k))
    out["simd_mismatch_kind"] = np.array([m[0] for m in mism], dtype="U8")
    out["simd_mismatch_index"] = np.array([m[1] for m in mism], np.int64)
    ref_s.set_level(0)

    path = os.path.join(OUT, "ref_bd%d.npz" % bd)
    np.savez_compressed(path, **out)
    print("wrote %s (%d KB); SIMD-vs-C mismatches: %s" % (path, os.path.getsize(path) // 1024, mism))


if __name__ == "__main__":
    if not RefLib.available(8):
        sys.exit("oracle/_ref is not built (needs the reference tree): python oracle/build.py")
    gen(8)
    gen(10)
