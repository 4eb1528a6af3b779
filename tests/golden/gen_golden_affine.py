"""Writes the golden vectors of affine luma motion compensation from the
compiled reference: com_mc_blk_affine_luma (com_mc.c:202-250) for CUs of
every size from 16x16 to 128x128 at random positions (some reaching past the
picture), 4x4 / 8x4 / 4x8 / 8x8 sub-blocks and random control-point MVs in
1/16 pel (a fifth of them translational).

Run in the container where /root/reference exists, after oracle/build.py has
compiled it into oracle/_ref:

    python tests/golden/gen_golden_affine.py

For each bit depth it writes tests/golden/affine_bd{8,10}.npz: the reference
picture (160x96), the cases (x, y, cu_w, cu_h, sub_w, sub_h, the two CPMVs,
offset of the prediction) and the reference's predictions concatenated; and
cases3 / preds3 / preds3_u / preds3_v, com_mc_cu_affine (com_mc.c:365-385)
with two or three control points, luma and chroma (x, y, cu_w, cu_h, 8x8
sub-blocks or not, number of control points, three CPMVs, luma offset,
chroma offset); the chroma planes are two subsampled phases of the picture."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
W, H, NCASES = 160, 96, 64


def make_cases(seed):
    rng = np.random.default_rng(seed)
    rows, off = [], 0
    for k in range(NCASES):
        cw, ch = (16, 16, 32, 32, 64, 128)[rng.integers(0, 6)], (16, 16, 32, 64, 128)[rng.integers(0, 5)]
        x, y = int(rng.integers(0, W)) // 4 * 4, int(rng.integers(0, H)) // 4 * 4
        sw, sh = (4, 8)[rng.integers(0, 2)], (4, 8)[rng.integers(0, 2)]
        cp = rng.integers(-700, 700, (2, 2))
        if k % 5 == 0:
            cp[1] = cp[0]
        rows.append([x, y, cw, ch, sw, sh] + cp.reshape(-1).tolist() + [off])
        off += cw * ch
    return np.array(rows, np.int32)


def make_cases3(seed):
    """com_mc_cu_affine cases (luma and chroma): x, y, cu_w, cu_h, 8x8
    sub-blocks or not, control points (2 or 3), the three CPMVs (the third
    unused with 2), luma offset, chroma offset."""
    rng = np.random.default_rng(seed)
    rows, off, coff = [], 0, 0
    for k in range(NCASES // 2):
        cw, ch = (16, 16, 32, 32, 64, 128)[rng.integers(0, 6)], (16, 16, 32, 64, 128)[rng.integers(0, 5)]
        x, y = int(rng.integers(0, W)) // 4 * 4, int(rng.integers(0, H)) // 4 * 4
        rows.append([x, y, cw, ch, int(rng.integers(0, 2)), 2 + k % 2] + rng.integers(-700, 700, 6).tolist() +
                    [off, coff])
        off += cw * ch
        coff += cw * ch // 4
    return np.array(rows, np.int32)


def chroma_planes(pic):
    """The 4:2:0 chroma planes of the golden picture: two subsampled phases."""
    return pic[::2, ::2].copy(), pic[1::2, 1::2].copy()


def main():
    for bd in (8, 10):
        rng = np.random.default_rng(90 + bd)
        dt = np.uint8 if bd == 8 else np.uint16
        yy, xx = np.mgrid[0:H, 0:W]
        top = (1 << bd) - 1
        pic = np.clip((0.5 + 0.3 * np.sin(xx / 5.0) * np.cos(yy / 3.0)) * top +
                      rng.normal(0, top / 20, (H, W)), 0, top).astype(dt)
        cases = make_cases(100 + bd)
        ref = RefLib(bd, 0)
        rp = ref.pic_create(pic, 0)
        preds = [ref.affine_luma(rp, *[int(v) for v in c[:4]], c[6:10].reshape(2, 2), int(c[4]), int(c[5])).reshape(-1)
                 for c in cases]
        u, v = chroma_planes(pic)
        ref.pic_set_chroma(rp, u, v)
        cases3 = make_cases3(200 + bd)
        out3 = [ref.affine_cu(rp, *[int(q) for q in c[:4]], c[6:12].reshape(3, 2)[:int(c[5])], int(c[4]))
                for c in cases3]
        ref.pic_free(rp)
        np.savez_compressed(os.path.join(OUT, "affine_bd%d.npz" % bd), pic=pic, cases=cases,
                            preds=np.concatenate(preds), cases3=cases3,
                            preds3=np.concatenate([o[0].reshape(-1) for o in out3]),
                            preds3_u=np.concatenate([o[1].reshape(-1) for o in out3]),
                            preds3_v=np.concatenate([o[2].reshape(-1) for o in out3]))
        print("bd %d: %d cases, %d samples" % (bd, len(cases), sum(p.size for p in preds)))


if __name__ == "__main__":
    main()
