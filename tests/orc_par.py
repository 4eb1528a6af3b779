"""The oracle's frame search split over threads (test helper, not a test).

orc_me_frame_lists searches every PU of the selected shapes for CTUs
[ctu0, ctu1) independently, so a frame can be computed as one call per
(shape, CTU) and reassembled in the frame API's PU order (shapes in table
order, each shape's PUs in raster order).  ctypes releases the GIL during
the C call, so the calls run in parallel; results are identical to one
whole-frame call, which test_orc_par_equals_whole_frame checks."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from oracle.oracle import SHAPES


def _workers():
    return max(1, min(16, os.cpu_count() or 1))


def me_frame(orc, cur, planes, R, lam, mask, num_refs_l1=0):
    h, w = cur.shape
    n_ctu = ((w + 127) // 128) * ((h + 127) // 128)
    shapes = [s for s in range(len(SHAPES)) if mask >> s & 1]
    per = {s: (128 // SHAPES[s][0]) * (128 // SHAPES[s][1]) for s in shapes}
    base, n = {}, 0
    for s in shapes:
        base[s] = n
        n += per[s]
    nref = len(planes)
    res = None
    best = np.full((n_ctu, n, 2) if num_refs_l1 else (n_ctu, n), -1, np.int8)

    def one(task):
        s, ctu = task
        r, b = orc.me_frame(cur, planes, R, lam, 1 << s, ctu0=ctu, ctu1=ctu + 1, num_refs_l1=num_refs_l1)
        return s, ctu, r[ctu], b[ctu]

    tasks = [(s, c) for s in shapes for c in range(n_ctu)]
    # the largest PUs first: their calls take longest
    tasks.sort(key=lambda t: -SHAPES[t[0]][0] * SHAPES[t[0]][1])
    with ThreadPoolExecutor(max_workers=_workers()) as ex:
        for s, ctu, r, b in ex.map(one, tasks):
            if res is None:
                res = np.zeros((n_ctu, n, nref), r.dtype)
            res[ctu, base[s]:base[s] + per[s]] = r
            best[ctu, base[s]:base[s] + per[s]] = b
    return res, best
