"""Writes the golden vectors of me_search_tz with bi = 1 (me.c:529-568, as
analyze_bi calls it, inter.c:1006-1011) from the compiled reference: random
PUs of every shape, bi-prediction targets 2*org - pred clipped
(create_bi_org, inter.c:55-68), MVPs and start MVs at every AMVR precision,
both sub-pel cost types, both lists, list sizes 1..4 and the other list's
mot_bits.

Run in the container where /root/reference exists, after oracle/build.py has
compiled it into oracle/_ref:

    python tests/golden/gen_golden_bi.py

For each bit depth it writes tests/golden/bi_bd{8,10}.npz: the current and
reference pictures (160x96), the cases (shape, x, y, mvr, cost type, list,
list size, refi, mot_bits, MVP, start MV, offset of the target), the targets
concatenated, and the reference's cost and MV per case."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib, SHAPES  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
W, H, NCASES, LAMBDA = 160, 96, 160, 314169


def make_pictures(bd, seed):
    rng = np.random.default_rng(seed)
    top = (1 << bd) - 1
    yy, xx = np.mgrid[0:H, 0:W]
    base = (0.5 + 0.3 * np.sin(xx / 6.0) * np.cos(yy / 4.0)) * top
    dt = np.uint8 if bd == 8 else np.uint16
    cur = np.clip(base + rng.normal(0, top / 60, (H, W)), 0, top).astype(dt)
    ref = np.clip(np.roll(base, (2, -3), axis=(0, 1)) + rng.normal(0, top / 60, (H, W)), 0, top).astype(dt)
    return cur, ref


def make_cases(bd, seed, cur):
    """(case rows, targets): a target is 2*org - (cur displaced by one
    sample), clipped, as if the other list predicted the PU one sample off."""
    rng = np.random.default_rng(seed)
    top = (1 << bd) - 1
    rows, tgts, off = [], [], 0
    shifted = np.roll(cur, (1, 1), axis=(0, 1)).astype(np.int32)
    while len(rows) < NCASES:
        s = int(rng.integers(len(SHAPES)))
        w, h = SHAPES[s]
        if w > W or h > H:
            continue
        x = int(rng.integers(0, (W - w) // 4 + 1)) * 4
        y = int(rng.integers(0, H - h + 1))
        mvr, ct, lidx = int(rng.integers(0, 5)), int(rng.integers(0, 2)), int(rng.integers(0, 2))
        nref = int(rng.integers(1, 5))
        refi, mb = int(rng.integers(0, nref)), int(rng.integers(0, 30))
        g = 1 << mvr
        mvp = [int(rng.integers(-40, 41)) * g, int(rng.integers(-40, 41)) * g]
        mv0 = [int(rng.integers(-30, 31)) // g * g, int(rng.integers(-30, 31)) // g * g]
        org = cur[y:y + h, x:x + w].astype(np.int32)
        tgt = np.clip(2 * org - shifted[y:y + h, x:x + w], 0, top).astype(cur.dtype)
        rows.append([s, x, y, mvr, ct, lidx, nref, refi, mb] + mvp + mv0 + [off])
        tgts.append(tgt.reshape(-1))
        off += w * h
    return np.array(rows, np.int32), np.concatenate(tgts)


def main():
    for bd in (8, 10):
        cur, refimg = make_pictures(bd, 70 + bd)
        cases, tgts = make_cases(bd, 80 + bd, cur)
        ref = RefLib(bd, 0)
        pic = ref.pic_create(refimg, 0)
        cost, mv = [], []
        for s, x, y, mvr, ct, lidx, nref, refi, mb, px, py, mx, my, off in cases.tolist():
            w, h = SHAPES[s]
            tgt = tgts[off:off + w * h].reshape(h, w)
            c, m = ref.me_search_tz_bi(pic, tgt, x, y, refi, nref, lidx, mb, mvr, ct, LAMBDA, (px, py), (mx, my))
            cost.append(c)
            mv.append(m)
        ref.pic_free(pic)
        np.savez_compressed(os.path.join(OUT, "bi_bd%d.npz" % bd), cur=cur, ref=refimg, cases=cases, targets=tgts,
                            cost=np.array(cost, np.uint64), mv=np.array(mv, np.int16))
        print("bd %d: %d cases, %d with an empty window" % (bd, len(cases), sum(c == (1 << 64) - 1 for c in cost)))


if __name__ == "__main__":
    main()
