"""Writes golden vectors of the affine gradient search, affine_me_gradient
(inter.c:1339-1471), from the compiled reference (oracle/ref_harness.c
refh_affine_me_gradient calls the reference's own static function).

Run in the container where /root/reference exists, after oracle/build.py has
compiled it into oracle/_ref:

    python tests/golden/gen_golden_affine_me.py

For each bit depth it writes tests/golden/affine_me_bd{8,10}.npz: the
current picture and the reference (224x160; the reference is the current
field zoomed, rotated and shifted, with a flat block and a block of
vertical stripes, where the least-squares system is singular), the cases
(x, y, w, h, sub_w, sub_h, refi, num_refp, mvr, bi, mot_bits_other, mvp
[2][2], start mv [2][2], offset into `orgs`), the CU originals concatenated
(the current picture's block for uni-prediction, a bi-prediction target
2*org - pred for bi), and the reference's cost and best control points."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
W, H, NCASES, LAMBDA = 224, 160, 160, 314169
FLAT = (16, 16, 48, 48)      # x, y, w, h of the flat block of the reference
STRIPES = (144, 96, 64, 48)  # the block of vertical stripes


def pictures(bd, seed):
    rng = np.random.default_rng(seed)
    top = (1 << bd) - 1
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    dt = np.uint8 if bd == 8 else np.uint16

    def field(zx, zy, rot, dx, dy):
        u = (xx - W / 2) * zx - (yy - H / 2) * rot + W / 2 + dx
        v = (yy - H / 2) * zy + (xx - W / 2) * rot + H / 2 + dy
        b = 0.5 + 0.25 * np.sin(u / 7.0) * np.cos(v / 5.0) + 0.12 * np.sin((u + v) / 3.1)
        return np.clip(np.rint(b * top + rng.normal(0, top / 200, (H, W))), 0, top).astype(dt)

    cur = field(1, 1, 0, 0, 0)
    ref = field(1.03, 0.98, 0.02, 1.3, -0.7)
    x, y, w, h = FLAT
    ref[y:y + h, x:x + w] = top // 3
    cur[y:y + h, x:x + w] = top // 3
    x, y, w, h = STRIPES
    stripes = ((np.arange(w) // 3) % 2 * (top // 2) + top // 4).astype(dt)
    ref[y:y + h, x:x + w] = stripes[None, :]
    return cur, ref


def cases(bd, seed, cur):
    rng = np.random.default_rng(seed)
    top = (1 << bd) - 1
    shifted = np.roll(cur, (1, -1), axis=(0, 1)).astype(np.int32)
    rows, orgs, off = [], [], 0
    while len(rows) < NCASES:
        k = len(rows)
        if k < 8:        # CUs inside the flat block and the stripes
            bx, by, bw, bh = FLAT if k < 4 else STRIPES
            w, h = 16 if k % 2 else 32, 16 if k % 4 < 2 else 32
            x, y = bx + (bw - w) // 2, by + (bh - h) // 2
        else:
            w, h = int(rng.choice([16, 32, 64, 128])), int(rng.choice([16, 32, 64, 128]))
            if w > W or h > H:
                continue
            x, y = int(rng.integers(0, W - w + 1)), int(rng.integers(0, H - h + 1))
        sub = int(rng.choice([4, 8]))
        mvr, bi = int(rng.integers(0, 3)), int(rng.integers(0, 2))
        g = 1 << [2, 4, 0][mvr]
        # flat / stripe CUs start at MV 0 so that they predict from the block
        scale = 0 if k < 8 else 1
        mvp = rng.integers(-40, 41, 4) * g * scale
        mv = rng.integers(-40, 41, 4) * g * scale
        nref = int(rng.integers(1, 5))
        refi, mb = int(rng.integers(0, nref)), int(rng.integers(0, 30))
        org = cur[y:y + h, x:x + w].astype(np.int32)
        if bi:
            org = np.clip(2 * org - shifted[y:y + h, x:x + w], 0, top)
        rows.append([x, y, w, h, sub, sub, refi, nref, mvr, bi, mb] + list(mvp) + list(mv) + [off])
        orgs.append(org.astype(cur.dtype).reshape(-1))
        off += w * h
    return np.array(rows, np.int64), np.concatenate(orgs)


def main():
    for bd in (8, 10):
        cur, refimg = pictures(bd, 60 + bd)
        rows, orgs = cases(bd, 61 + bd, cur)
        ref = RefLib(bd, 0)
        pic = ref.pic_create(refimg, 0)
        cost, mv = [], []
        for r in rows.tolist():
            x, y, w, h, sw, sh, refi, nref, mvr, bi, mb = r[:11]
            mvp, mv0, off = r[11:15], r[15:19], r[19]
            org = orgs[off:off + w * h].reshape(h, w)
            c, m = ref.affine_me_gradient(pic, org, x, y, w, h, sw, sh, refi, nref, mvr, bi, mb, LAMBDA, mvp, mv0)
            cost.append(c)
            mv.append(np.array(m).reshape(-1))
        ref.pic_free(pic)
        mv = np.array(mv, np.int32)
        moved = int((mv != rows[:, 15:19]).any(1).sum())
        np.savez_compressed(os.path.join(OUT, "affine_me_bd%d.npz" % bd), cur=cur, ref=refimg, cases=rows,
                            orgs=orgs, cost=np.array(cost, np.uint64), mv=mv)
        print("bd %d: %d cases, %d moved from their start" % (bd, len(rows), moved))


if __name__ == "__main__":
    main()
