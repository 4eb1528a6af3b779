"""me_frame_search_noalloc (the frame search whose keys and tables live in
the library's fixed device array) against the plan API's me_plan_run on the
same inputs: bit-identical results on consecutive calls, the second with a
smaller picture, fewer shapes and fewer references, then the first again;
a frame that does not fit the array is rejected and the entry stays usable.
Device buffers come from the HIP runtime the library links, through the C
ABI only."""
import ctypes

import numpy as np
import pytest

import uavs3e_amd as ua
from oracle.oracle import SHAPES
from uavs3e_amd import workloads as wl

pytestmark = pytest.mark.gpu

LAM = wl.lambda_mv(32)


class DevBuf:
    def __init__(self, lib, nbytes):
        self.lib = lib
        self.ptr = ctypes.c_void_p()
        lib.hipMalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        assert lib.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert self.lib.hipMemcpy(self.ptr, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def download(self, arr):
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert self.lib.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), self.ptr, arr.nbytes, 2) == 0
        return arr

    def free(self):
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        assert self.lib.hipFree(self.ptr) == 0


def _case(bd, w, h, R, mask, nref, l1=0):
    base = wl.hd_allpu() if bd == 8 else wl.fuhd10_bi()
    work = base.scaled(w, h, R)
    shifts = (list(work.ref_shifts_qpel) * 2)[:nref]
    work.ref_shifts_qpel = shifts
    cur, refs = work.frames()
    return cur, refs, ua.FrameCfg(w, h, R, LAM, mask, nref, 0, 0, 0, l1)


def _noalloc(hip, cfg, cur, dref):
    n_ctu = ((cfg.pic_w + 127) // 128) * ((cfg.pic_h + 127) // 128)
    npu = sum((128 // sw) * (128 // sh) for s, (sw, sh) in enumerate(SHAPES) if cfg.shape_mask >> s & 1)
    nl = 2 if cfg.num_refs_l1 else 1
    res = np.zeros((n_ctu, npu, cfg.num_refs), ua.PU_RESULT)
    best = np.zeros((n_ctu, npu, 2) if nl == 2 else (n_ctu, npu), np.int8)
    d_cur = DevBuf(hip.lib, cur.nbytes + 64)
    d_res = DevBuf(hip.lib, res.nbytes)
    d_best = DevBuf(hip.lib, best.nbytes)
    d_cur.upload(cur)
    arr = (ctypes.c_void_p * len(dref))(*[r.value for r in dref])
    f = hip.lib.me_frame_search_noalloc
    f.argtypes = [ctypes.c_void_p] * 6
    hip.check(f(ctypes.byref(cfg), d_cur.ptr, arr, d_res.ptr, d_best.ptr, None), "me_frame_search_noalloc")
    d_res.download(res)
    d_best.download(best)
    for b in (d_cur, d_res, d_best):
        b.free()
    return res, best


def _plan(hip, cfg, cur, dref):
    plan = hip.plan_create(cfg.pic_w, cfg.pic_h, cfg.search_range, cfg.lambda_mv, cfg.shape_mask, cfg.num_refs,
                           num_refs_l1=cfg.num_refs_l1)
    hip.plan_set_cur(plan, cur)
    hip.plan_run(plan, dref)
    out = hip.plan_results(plan, cfg.num_refs)
    hip.plan_destroy(plan)
    return out


@pytest.mark.parametrize("bd", [8, 10])
def test_noalloc_equals_plan_run(bd):
    hip = ua.MeHip(bd)
    small = sum(1 << SHAPES.index(s) for s in [(8, 8), (16, 16), (64, 32)])
    cases = [_case(bd, 392, 264, 16, ua.MASK_ALL, 2 if bd == 8 else 4, 0 if bd == 8 else 2),
             _case(bd, 200, 136, 8, small, 1)]
    cases.append(cases[0])
    for cur, refs, cfg in cases:
        print("no-alloc: bd %d %dx%d R=%d refs=%d" % (bd, cfg.pic_w, cfg.pic_h, cfg.search_range, cfg.num_refs),
              flush=True)
        dref = [hip.ref_create(r) for r in refs]
        got = _noalloc(hip, cfg, cur, dref)
        want = _plan(hip, cfg, cur, dref)
        for r in dref:
            hip.ref_destroy(r)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        assert got[0]["valid"].sum() > 0
    assert hip.error_count() == 0


def test_noalloc_rejects_frames_beyond_the_array():
    """3840x2160 with all 29 shapes and 4 references needs 44 MB of keys."""
    hip = ua.MeHip(8)
    w, h = 3840, 2160
    zeros = np.zeros((h, w), np.uint8)
    dref = [hip.ref_create(zeros)]
    cfg = ua.FrameCfg(w, h, 8, LAM, ua.MASK_ALL, 4, 0, 0, 0, 0)
    f = hip.lib.me_frame_search_noalloc
    f.argtypes = [ctypes.c_void_p] * 6
    arr = (ctypes.c_void_p * 4)(*([dref[0].value] * 4))
    d = DevBuf(hip.lib, 256)   # never touched: the call fails before any launch
    rc = f(ctypes.byref(cfg), d.ptr, arr, d.ptr, d.ptr, None)
    assert rc != 0 and b"maximum" in hip.lib.me_hip_last_error()
    d.free()
    for r in dref:
        hip.ref_destroy(r)
    # still usable
    cur, refs, cfg = _case(8, 200, 136, 8, ua.MASK_16x16, 1)
    dref = [hip.ref_create(r) for r in refs]
    got = _noalloc(hip, cfg, cur, dref)
    want = _plan(hip, cfg, cur, dref)
    for r in dref:
        hip.ref_destroy(r)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
