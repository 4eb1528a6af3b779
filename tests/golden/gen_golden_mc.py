"""Writes the golden vectors of block motion compensation from the compiled
reference: com_mc_cu (com_mc.c:140-200) -- luma quarter-pel copies from the
sub-pel planes, chroma 1/8-pel 4-tap filtering, L0/L1 bi-averaging -- for
PUs of every shape at random positions with MVs that reach the clipped
border area.

Run in the container where /root/reference exists, after oracle/build.py has
compiled it into oracle/_ref:

    python tests/golden/gen_golden_mc.py

For each bit depth it writes tests/golden/mc_bd{8,10}.npz: the luma and
4:2:0 chroma of 3 reference pictures (160x96), the cases (x, y, w, h,
number of lists, ref and MV per list) and the reference's predictions
through its C table, concatenated (luma, then u and v for PUs whose chroma
exists); simd_mismatch lists the cases where the production C+SSE+AVX2 table
gives other samples (the 10-bit AVX2 2-D chroma filter)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib, SHAPES  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
W, H, NREF, NCASES = 160, 96, 3, 96


def picture(rng, h, w, bd):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    v = np.full((h, w), float(1 << (bd - 1)))
    for _ in range(4):
        a = rng.uniform(10, 40) * (1 if bd == 8 else 4)
        fx, fy = rng.uniform(-0.2, 0.2, 2)
        v += a * np.sin(fx * xx + fy * yy + rng.uniform(0, 6.28))
    v += rng.normal(0, 6 if bd == 8 else 24, (h, w))
    return np.clip(np.floor(v + 0.5), 0, (1 << bd) - 1)


def cases(rng):
    out = []
    for k in range(NCASES):
        w, h = SHAPES[k % len(SHAPES)]
        x = int(rng.integers(0, max(1, (W - w) // 4 + 1))) * 4 if w <= W else 0
        y = int(rng.integers(0, max(1, (H - h) // 4 + 1))) * 4 if h <= H else 0
        nl = 1 + (k % 3 != 0)
        row = [x, y, w, h, nl]
        for l in range(2):
            r = int(rng.integers(0, NREF))
            if k % 4 == 1:     # far outside the picture: the clipped border band
                mv = rng.integers(-4 * (W + 200), 4 * (W + 200), 2)
            else:
                mv = rng.integers(-80, 80, 2)
            if k % 7 == 3 and l == 1:   # both lists on the same picture with the same MV
                r, mv = row[5], np.array(row[6:8])
            row += [r, int(mv[0]), int(mv[1])]
        out.append(row)
    return np.array(out, np.int32)


def gen(bd):
    rng = np.random.default_rng(3000 + bd)
    dt = np.uint8 if bd == 8 else np.uint16
    ref = RefLib(bd, level=0)
    luma = np.stack([picture(rng, H, W, bd) for _ in range(NREF)]).astype(dt)
    cu = np.stack([picture(rng, H // 2, W // 2, bd) for _ in range(NREF)]).astype(dt)
    cv = np.stack([picture(rng, H // 2, W // 2, bd) for _ in range(NREF)]).astype(dt)
    pics = []
    for r in range(NREF):
        p = ref.pic_create(luma[r], poc=r)
        ref.pic_set_chroma(p, cu[r], cv[r])
        pics.append(p)
    cs = cases(rng)
    preds = []
    simd_mismatch = []
    for k, (x, y, w, h, nl, r0, m0x, m0y, r1, m1x, m1y) in enumerate(cs.tolist()):
        chroma = w >= 8 and h >= 8
        outs = []
        for level in (0, 2):   # the C table is the specification; note where SIMD differs
            ref.set_level(level)
            outs.append(ref.mc_cu(x, y, W, H, w, h, pics[r0], pics[r1] if nl == 2 else None, (m0x, m0y),
                                  (m1x, m1y), chroma=chroma))
        py, pu, pv = outs[0]
        if any(a is not None and not np.array_equal(a, b) for a, b in zip(outs[0], outs[1])):
            simd_mismatch.append(k)
        preds.append(py.ravel())
        if chroma:
            preds += [pu.ravel(), pv.ravel()]
    for p in pics:
        ref.pic_free(p)
    np.savez_compressed(os.path.join(OUT, "mc_bd%d.npz" % bd), luma=luma, u=cu, v=cv, cases=cs,
                        preds=np.concatenate(preds).astype(dt), simd_mismatch=np.array(simd_mismatch, np.int32))
    print("mc_bd%d.npz: %d cases, C and SIMD tables differ on %d" % (bd, len(cs), len(simd_mismatch)))


if __name__ == "__main__":
    for bd in (8, 10):
        gen(bd)
