"""Writes golden vectors of the bi-prediction search loop (analyze_bi,
inter.c:972-1070, up to the end of its search) and of the symmetric-MVD
search (analyze_smvd, inter.c:1194-1255, with smvd_refine, inter.c:1083-1191)
from the compiled reference, for every PU of a frame-plan layout.

analyze_bi is the reference's own function, called through a minimal
core_t (oracle/ref_harness.c refh_analyze_bi); the SMVD search runs the
reference's smvd_refine after the two first candidates of analyze_smvd
(refh_smvd).  Run in the container where /root/reference exists, after
oracle/build.py has compiled it into oracle/_ref:

    python tests/golden/gen_golden_bi_loop.py

For each bit depth it writes tests/golden/bi_loop_bd{8,10}.npz and
smvd_bd{8,10}.npz.  Each holds the current picture and the references
(136x136; list 0 first), the plan's shape mask, and per configuration
k: cfg_k = (n0, n1, mvr, cost_type), the per-PU inputs in the frame API's
layout ([ctu][pu], references list 0 first) and the reference's outputs.
bi: mvp_k / mv_scale_k [ctu][pu][ref][2], pus_k (me_bi_pu_t fields refi,
lidx_ref, active, mv, mot_bits), out_refi_k, out_mv_k, out_mot_k,
out_mv_scale_k; an active PU inside the picture is checked.
smvd: mvp_k / mv_scale_k, active_k, out_mv_k [ctu][pu][2][2] and
out_cost_k."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib, SHAPES  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
W = H = 136
LAMBDA = 314169
SEL = [(16, 16), (8, 16), (32, 8), (64, 64), (128, 128), (4, 8)]
MASK = sum(1 << SHAPES.index(s) for s in SEL)
BI_CFGS = {8: [(1, 1, 0, 0), (2, 2, 0, 1), (1, 2, 2, 0)], 10: [(2, 1, 0, 0), (1, 1, 3, 1), (2, 2, 1, 0)]}
SMVD_CFGS = {8: [(1, 1, 0, 0), (2, 2, 0, 1), (1, 1, 3, 0)], 10: [(1, 2, 0, 0), (2, 1, 2, 1), (1, 1, 1, 1)]}


def picture(rng, bd, dx=0, dy=0):
    top = (1 << bd) - 1
    yy, xx = np.mgrid[0:H, 0:W]
    b = 0.5 + 0.25 * np.sin((xx - dx) / 6.0) * np.cos((yy - dy) / 4.0) + 0.1 * np.sin((xx + yy - dx - dy) / 2.5)
    return np.clip(np.rint(b * top + rng.normal(0, top / 255, (H, W))), 0, top).astype(
        np.uint8 if bd == 8 else np.uint16)


def layout():
    """[(ctu, pu, x, y, w, h)] of the plan's PUs in the frame API's order."""
    cols = (W + 127) // 128
    n_ctu = cols * ((H + 127) // 128)
    out = []
    for ctu in range(n_ctu):
        cx, cy = (ctu % cols) * 128, (ctu // cols) * 128
        pu = 0
        for s in range(len(SHAPES)):
            if not MASK >> s & 1:
                continue
            pw, ph = SHAPES[s]
            for i in range((128 // pw) * (128 // ph)):
                out.append((ctu, pu, cx + (i % (128 // pw)) * pw, cy + (i // (128 // pw)) * ph, pw, ph))
                pu += 1
    return n_ctu, pu, out


def bi_inputs(rng, n_ctu, npu, n0, n1, mvr):
    """The same input distribution as the GPU tests of me_plan_run_bi."""
    g = 1 << mvr
    nref = n0 + n1
    mvp = (rng.integers(-24, 25, (n_ctu, npu, nref, 2)) * g).astype(np.int16)
    mv_scale = (rng.integers(-16, 17, (n_ctu, npu, nref, 2)) * g).astype(np.int16)
    refi = np.stack([rng.integers(0, n0, (n_ctu, npu)), rng.integers(0, n1, (n_ctu, npu))], -1).astype(np.int8)
    lidx = rng.integers(0, 2, (n_ctu, npu)).astype(np.int8)
    active = (rng.random((n_ctu, npu)) < 0.9).astype(np.int8)
    mv = (rng.integers(-16, 17, (n_ctu, npu, 2, 2)) * g).astype(np.int16)
    mot = rng.integers(2, 30, (n_ctu, npu, 2)).astype(np.int32)
    return mvp, mv_scale, refi, lidx, active, mv, mot


def main():
    for bd in (8, 10):
        rng = np.random.default_rng(900 + bd)
        cur = picture(rng, bd)
        refs = [picture(rng, bd, 2 * k + 1, -k) for k in range(2)] + [picture(rng, bd, -3 * k - 1, k + 1)
                                                                     for k in range(2)]
        ref = RefLib(bd, 0)
        n_ctu, npu, pus = layout()
        bi = {"cur": cur, "refs": np.stack(refs), "mask": np.uint32(MASK), "w": W, "h": H}
        for k, (n0, n1, mvr, ct) in enumerate(BI_CFGS[bd]):
            # list 0: refs[0 .. n0), list 1: refs[2 .. 2 + n1)
            lrefs = refs[:n0] + refs[2:2 + n1]
            pics = [ref.pic_create(r, -1 - i if i < n0 else i - n0 + 1) for i, r in enumerate(lrefs)]
            mvp, ms, refi, lidx, active, mv, mot = bi_inputs(rng, n_ctu, npu, n0, n1, mvr)
            o_refi, o_mv, o_mot, o_ms = refi.copy(), mv.copy(), mot.copy(), ms.copy()
            for ctu, pu, x, y, pw, ph in pus:
                if not active[ctu, pu] or x + pw > W or y + ph > H:
                    continue
                m4 = np.zeros((2, 4, 2), np.int16)
                s4 = np.zeros((2, 4, 2), np.int16)
                m4[0, :n0], m4[1, :n1] = mvp[ctu, pu, :n0], mvp[ctu, pu, n0:]
                s4[0, :n0], s4[1, :n1] = ms[ctu, pu, :n0], ms[ctu, pu, n0:]
                r_, m_, b_, s_ = ref.analyze_bi(cur, pics[:n0], pics[n0:], x, y, pw, ph, m4, s4, refi[ctu, pu],
                                                mv[ctu, pu], int(lidx[ctu, pu]), mot[ctu, pu], mvr, ct, LAMBDA)
                o_refi[ctu, pu], o_mv[ctu, pu], o_mot[ctu, pu] = r_, m_, b_
                o_ms[ctu, pu, :n0], o_ms[ctu, pu, n0:] = s_[0, :n0], s_[1, :n1]
            for p in pics:
                ref.pic_free(p)
            bi.update({"cfg_%d" % k: np.array([n0, n1, mvr, ct], np.int32), "mvp_%d" % k: mvp,
                       "mv_scale_%d" % k: ms, "refi_%d" % k: refi, "lidx_%d" % k: lidx, "active_%d" % k: active,
                       "mv_%d" % k: mv, "mot_%d" % k: mot, "out_refi_%d" % k: o_refi, "out_mv_%d" % k: o_mv,
                       "out_mot_%d" % k: o_mot, "out_mv_scale_%d" % k: o_ms})
            changed = int((o_refi != refi).any(-1).sum() + (o_mv != mv).any((-1, -2)).sum())
            print("bd %d bi cfg %s: %d PUs with a changed reference or MV" % (bd, (n0, n1, mvr, ct), changed))
        np.savez_compressed(os.path.join(OUT, "bi_loop_bd%d.npz" % bd), **bi)

        sm = {"cur": cur, "refs": np.stack(refs), "mask": np.uint32(MASK), "w": W, "h": H}
        for k, (n0, n1, mvr, ct) in enumerate(SMVD_CFGS[bd]):
            lrefs = refs[:n0] + refs[2:2 + n1]
            p0, p1 = ref.pic_create(lrefs[0], -1), ref.pic_create(lrefs[n0], 1)
            g = 1 << mvr
            mvp = (rng.integers(-12, 13, (n_ctu, npu, n0 + n1, 2)) * g).astype(np.int16)
            ms = (rng.integers(-12, 13, (n_ctu, npu, n0 + n1, 2)) * g).astype(np.int16)
            active = (rng.random((n_ctu, npu)) < 0.9).astype(np.int8)
            o_mv = np.zeros((n_ctu, npu, 2, 2), np.int16)
            o_cost = np.zeros((n_ctu, npu), np.uint32)
            for ctu, pu, x, y, pw, ph in pus:
                if not active[ctu, pu] or x + pw > W or y + ph > H:
                    continue
                m_, c_ = ref.smvd(cur, p0, p1, x, y, pw, ph, np.stack([mvp[ctu, pu, 0], mvp[ctu, pu, n0]]),
                                  ms[ctu, pu, 0], mvr, ct, LAMBDA)
                o_mv[ctu, pu], o_cost[ctu, pu] = m_, c_
            ref.pic_free(p0)
            ref.pic_free(p1)
            sm.update({"cfg_%d" % k: np.array([n0, n1, mvr, ct], np.int32), "mvp_%d" % k: mvp, "mv_scale_%d" % k: ms,
                       "active_%d" % k: active, "out_mv_%d" % k: o_mv, "out_cost_%d" % k: o_cost})
            print("bd %d smvd cfg %s done" % (bd, (n0, n1, mvr, ct)))
        np.savez_compressed(os.path.join(OUT, "smvd_bd%d.npz" % bd), **sm)


if __name__ == "__main__":
    main()
