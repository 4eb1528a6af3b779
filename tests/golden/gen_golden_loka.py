"""Writes the golden vectors of the lookahead cost estimator from the
compiled reference: loka_estimate_coding_cost (lookahead.c:82-290), which
returns the picture's mean per-sample cost, the sum over whole 16x16 units
of the least inter or intra com_had divided by the number of samples.

The pictures are made so that an inter prediction beats every intra mode in
every unit -- noise pictures (intra cannot predict them) and references that
are the current picture plus small noise, displaced by offsets that every
unit's search window holds -- so that the reference's result is the
oracle's inter cost summed over the units.  One reference of list 1 is the
current picture minus list 0's first noise on the left half, so there the
bi-average is exact.

Run in the container where /root/reference exists, after oracle/build.py has
compiled it into oracle/_ref:

    python tests/golden/gen_golden_loka.py

For each bit depth it writes tests/golden/loka_bd{8,10}.npz: the current
picture (96x80), 6 references, the cases (is_padding, list-0 and list-1
reference indices, POCs) and the reference's result per case."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
W, H = 96, 80
GOP, MSR, CUR_POC = 16, 64, 8
# (is_padding, list-0 references, list-1 references); POC of reference i: REF_POC[i]
CASES = [
    (1, [0, 1], []),
    (1, [0, 1], [4, 5]),
    (1, [0], [4]),
    (0, [0, 1, 2, 3], []),
    (0, [0, 1], [4, 5]),
    (0, [0], [4]),
]
REF_POC = [4, 0, 2, 6, 12, 16]


def make_pictures(bd, seed):
    """The current picture and 6 references (see the module docstring)."""
    rng = np.random.default_rng(seed)
    S, top = 1 << (bd - 8), (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    cur = rng.integers(16 * S, 240 * S, (H, W)).astype(np.int32)

    def noisy(amp, shift=(0, 0)):
        b = np.roll(cur, shift, axis=(0, 1)) + rng.integers(-amp, amp + 1, cur.shape) * S
        return np.clip(b, 0, top)

    n = rng.integers(-3, 4, cur.shape) * S
    l1a = np.clip(cur - n, 0, top)
    l1a[:, W // 2:] = noisy(3)[:, W // 2:]
    refs = [np.clip(cur + n, 0, top),     # MV 0 except in the last unit column / row without padding
            noisy(1, (-1, -1)),           # MV (-1, -1)
            noisy(1, (0, -1)),            # MV (-1, 0): the top-right unit without padding
            noisy(1, (-1, 0)),            # MV (0, -1): the bottom-left unit without padding
            l1a,
            noisy(2, (1, -2))]
    return cur.astype(dt), [r.astype(dt) for r in refs]


def reference_costs(ref, cur, refs):
    out = []
    for pad, l0, l1 in CASES:
        p0 = [ref.pic_create(refs[i], REF_POC[i]) for i in l0]
        p1 = [ref.pic_create(refs[i], REF_POC[i]) for i in l1]
        out.append(ref.loka_cost(cur, p0, p1, CUR_POC, GOP, MSR, pad))
        for p in p0 + p1:
            ref.pic_free(p)
    return np.array(out, np.float64)


def main():
    for bd in (8, 10):
        cur, refs = make_pictures(bd, 40 + bd)
        ref = RefLib(bd, 2)
        costs = reference_costs(ref, cur, refs)
        cases = np.array([[pad, len(l0), len(l1)] + (l0 + [-1] * 4)[:4] + (l1 + [-1] * 4)[:4]
                          for pad, l0, l1 in CASES], np.int32)
        np.savez_compressed(os.path.join(OUT, "loka_bd%d.npz" % bd), cur=cur, refs=np.stack(refs), cases=cases,
                            ref_poc=np.array(REF_POC, np.int64), params=np.array([GOP, MSR, CUR_POC], np.int64),
                            cost=costs)
        print("bd %d: %s" % (bd, costs))


if __name__ == "__main__":
    main()
