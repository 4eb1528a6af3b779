"""uavs3e motion estimation on MI355X (gfx950): Python bindings.

The product is the C-ABI library ``lib/libuavs3e_me_hip{8,10}.so`` declared
in ``include/uavs3e_me_hip.h``.  This module loads it with ctypes and mirrors
its two layers for Python callers and tests:

* ``MeHip(bd)`` -- library state, the per-call entries (the hot-path slots of
  uavs3e's funs_handle_t, com_util.h:300-330), the dispatch table, device
  reference pictures and frame plans.
* ``MeHost(bd)`` -- host C helpers: the ME driver that calls costs only
  through a dispatch table (the per-call drop-in path) and the synthetic
  frame generator.

Loading fails loudly when a library is missing; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")

SHAPES = [(4, 8), (8, 4), (4, 16), (16, 4), (4, 32), (32, 4), (8, 8), (8, 16), (16, 8), (8, 32),
          (32, 8), (8, 64), (64, 8), (16, 16), (16, 32), (32, 16), (16, 64), (64, 16), (16, 128),
          (128, 16), (32, 32), (32, 64), (64, 32), (32, 128), (128, 32), (64, 64), (64, 128),
          (128, 64), (128, 128)]
MASK_ALL = (1 << 29) - 1
MASK_16x16 = 1 << SHAPES.index((16, 16))
PAD = 144

PU_RESULT = np.dtype([("mv_int", "<i2", 2), ("mv_sub", "<i2", 2), ("cost_int", "<u8"),
                      ("cost_sub", "<u8"), ("cost_final", "<u4"), ("valid", "<u4")])

# me_dispatch_t: 78 function pointers, slot order of the C struct.
DISPATCH_SLOTS = (
    [("ipcpy", i) for i in range(6)] +
    [("ipflt", t, i) for t in range(4) for i in range(6)] +
    [("ipflt_ext", t, i) for t in range(2) for i in range(6)] +
    [("ip_flt_y_hor",), ("ip_flt_y_ver",), ("ip_flt_y_ver_ext",)] +
    [("cost_sad", i) for i in range(6)] +
    [("cost_sad_x3", i) for i in range(6)] +
    [("cost_sad_x4", i) for i in range(6)] +
    [("cost_satd", a, b) for a in range(3) for b in range(3)] +
    [("pel_avrg", i) for i in range(6)])
DispatchTable = C.c_void_p * len(DISPATCH_SLOTS)


RESULT_FIELDS = ("mv_int", "mv_sub", "cost_int", "cost_sub", "cost_final")


def compare_results(res, best, ores, obest):
    """Compares two frame results ([ctu][pu][ref] PU_RESULT arrays and
    [ctu][pu] best-reference arrays).  Entries of PUs outside the picture
    (valid == 0) carry no result and are not compared.  Returns None when
    equal, else a message naming the first difference."""
    if not np.array_equal(res["valid"], ores["valid"]):
        return "valid flags differ at %d entries" % int((res["valid"] != ores["valid"]).sum())
    valid = ores["valid"] == 1
    for f in RESULT_FIELDS:
        a, b = res[f][valid], ores[f][valid]
        bad = np.nonzero(np.any(a != b, axis=-1) if a.ndim > 1 else a != b)[0]
        if len(bad):
            return "%s differs at %d of %d valid PUs, first %s vs %s" % (
                f, len(bad), int(valid.sum()), a[bad[0]], b[bad[0]])
    if not np.array_equal(best, obest):
        return "best reference differs at %d PUs" % int((best != obest).sum())
    return None


class FrameCfg(C.Structure):
    _fields_ = [("pic_w", C.c_int), ("pic_h", C.c_int), ("search_range", C.c_int), ("lambda_mv", C.c_uint32),
                ("shape_mask", C.c_uint32), ("num_refs", C.c_int), ("band_rows", C.c_int), ("waves", C.c_int),
                ("dx_cols", C.c_int), ("num_refs_l1", C.c_int), ("sp_groups", C.c_int), ("int_grid", C.c_int)]


def mc_layout(mask):
    """Elements per CTU of the luma and of each chroma prediction buffer of
    me_plan_mc, and the per-shape offsets {shape: (luma_off, chroma_off or
    None)} inside a CTU's block."""
    ly = lc = 0
    offs = {}
    for s, (w, h) in enumerate(SHAPES):
        if not mask >> s & 1:
            continue
        has_c = w >= 8 and h >= 8
        offs[s] = (ly, lc if has_c else None)
        ly += 16384
        lc += 4096 if has_c else 0
    return ly, lc, offs


class TuneResult(C.Structure):
    _fields_ = [("band_rows", C.c_int), ("waves", C.c_int), ("dx_cols", C.c_int), ("int_search_ms", C.c_double),
                ("sp_groups", C.c_int), ("subpel_ms", C.c_double), ("int_grid", C.c_int)]


class TzCfg(C.Structure):
    _fields_ = [("gop_size", C.c_int), ("max_search_range", C.c_int), ("adaptive_raster_range", C.c_int),
                ("mvr", C.c_int), ("cost_type", C.c_int), ("cur_poc", C.c_int64), ("ref_poc", C.c_int64 * 4)]


class ServerCfg(C.Structure):
    _fields_ = [("search_range", C.c_int), ("lambda_mv", C.c_uint32), ("shape_mask", C.c_uint32),
                ("num_refs", C.c_int), ("hbm_budget", C.c_uint64), ("max_batch_ctus", C.c_int),
                ("max_inflight", C.c_int), ("max_wait_us", C.c_int), ("band_rows", C.c_int), ("waves", C.c_int),
                ("dx_cols", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("batches", C.c_uint64), ("max_batch_seen", C.c_uint64),
                ("max_inflight_seen", C.c_uint64), ("mixed_batches", C.c_uint64), ("picture_bytes", C.c_uint64),
                ("admission_waits", C.c_uint64), ("max_launch_us", C.c_uint64), ("max_batch_us", C.c_uint64)]


# me_bi_pu_t
BI_PU = np.dtype([("refi", np.int8, 2), ("lidx_ref", np.int8), ("active", np.int8), ("mv", np.int16, (2, 2)),
                  ("mot_bits", np.int32, 2), ("cost", np.uint32)], align=True)


# me_affine_cu_t
AFFINE_CU = np.dtype([("x", np.int32), ("y", np.int32), ("w", np.int32), ("h", np.int32), ("sub_w", np.int32),
                      ("sub_h", np.int32), ("vertex_num", np.int32), ("reserved", np.int32),
                      ("cpmv", np.int32, (3, 2)), ("pred_offset", np.uint64), ("chroma_offset", np.uint64)],
                     align=True)

# me_affine_me_cu_t
AFFINE_ME_CU = np.dtype([("x", np.int32), ("y", np.int32), ("w", np.int32), ("h", np.int32), ("sub_w", np.int32),
                         ("sub_h", np.int32), ("refi", np.int32), ("num_refp", np.int32), ("mvr", np.int32),
                         ("bi", np.int32), ("mot_bits_other", np.int32), ("org_stride", np.int32),
                         ("mvp", np.int32, (2, 2)), ("mv", np.int32, (2, 2)), ("org_offset", np.uint64),
                         ("cost", np.uint64)], align=True)

# me_smvd_pu_t
SMVD_PU = np.dtype([("active", np.int8), ("pad", np.int8, 3), ("mv", np.int16, (2, 2)), ("cost", np.uint32)],
                   align=True)


class MeError(RuntimeError):
    pass


def pel_dtype(bd):
    return np.uint8 if bd == 8 else np.uint16


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _pat(a, off):
    return C.c_void_p(a.ctypes.data + int(off) * a.itemsize)


def hip_lib_path(bd):
    return os.path.join(LIB_DIR, "libuavs3e_me_hip%d.so" % bd)


def host_lib_path(bd):
    return os.path.join(LIB_DIR, "libme_host%d.so" % bd)


class MeHip:
    """The HIP library for one bit depth."""

    def __init__(self, bd=8, init=True):
        path = hip_lib_path(bd)
        if not os.path.exists(path):
            raise MeError("HIP library not built: %s (run __graft_entry__.build())" % path)
        self.bd = bd
        self.dt = pel_dtype(bd)
        L = self.lib = C.CDLL(path)
        L.me_hip_last_error.restype = C.c_char_p
        L.me_hip_error_count.restype = C.c_uint64
        for n in ("me_hip_sad", "me_hip_satd", "me_hip_had"):
            getattr(L, n).restype = C.c_uint32
        L.me_plan_run.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        for n in ("me_plan_num_ctus", "me_plan_num_pus", "me_plan_set_cur", "me_plan_results", "me_ref_load",
                  "me_ref_plane"):
            getattr(L, n).argtypes = None
        L.me_ref_load_chroma.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.me_plan_mc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.me_plan_mc_size.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.me_plan_mc_download.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.me_plan_run_tz.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.me_plan_lookahead.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.me_plan_run_bi.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
        L.me_plan_run_smvd.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]
        L.me_affine_mc_luma.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p]
        L.me_affine_mc_chroma.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64,
                                          C.c_void_p]
        L.me_plan_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                       C.c_void_p]
        L.me_plan_destroy.argtypes = [C.c_void_p]
        L.me_ref_destroy.argtypes = [C.c_void_p]
        L.me_server_create.argtypes = [C.c_void_p, C.c_void_p]
        L.me_server_open_stream.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.me_server_put_picture.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.c_int, C.c_int]
        L.me_server_release_picture.argtypes = [C.c_void_p, C.c_int, C.c_int64]
        L.me_server_submit.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.me_server_submit.restype = C.c_int64
        L.me_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p]
        L.me_server_now_us.restype = C.c_double
        L.me_server_stats.argtypes = [C.c_void_p, C.c_void_p]
        L.me_server_destroy.argtypes = [C.c_void_p]
        assert L.me_hip_bit_depth() == bd
        if init:
            self.check(L.me_hip_init(0), "me_hip_init")

    def check(self, rc, what):
        if rc != 0:
            raise MeError("%s failed: %s" % (what, self.lib.me_hip_last_error().decode()))

    def error_count(self):
        return self.lib.me_hip_error_count()

    # -- dispatch table ----------------------------------------------------
    def dispatch_table(self):
        t = DispatchTable()
        self.check(self.lib.me_funs_init_hip(C.byref(t)), "me_funs_init_hip")
        return t

    # -- per-call entries (numpy views; offsets are flat element indices) ---
    def sad(self, w, h, org, i_org, pred, i_pred, org_off=0, pred_off=0):
        return self.lib.me_hip_sad(w, _pat(org, org_off), i_org, _pat(pred, pred_off), i_pred, h)

    def sad_xn(self, w, h, org, i_org, org_off, pred, i_pred, offs):
        out = np.zeros(4, np.uint32)
        ps = [_pat(pred, o) for o in offs]
        if len(offs) == 3:
            self.lib.me_hip_sad_x3(w, _pat(org, org_off), i_org, *ps, i_pred, _p(out), h)
        else:
            self.lib.me_hip_sad_x4(w, _pat(org, org_off), i_org, *ps, i_pred, _p(out), h)
        return out[:len(offs)]

    def satd(self, wl, hl, org, i_org, cur, i_cur, org_off=0, cur_off=0):
        return self.lib.me_hip_satd(wl, hl, _pat(org, org_off), i_org, _pat(cur, cur_off), i_cur)

    def had(self, w, h, org, i_org, cur, i_cur, org_off=0, cur_off=0):
        return self.lib.me_hip_had(w, h, _pat(org, org_off), i_org, _pat(cur, cur_off), i_cur, self.bd)

    def ipflt(self, typ, src, i_src, src_off, w, h, coef, max_val):
        dst = np.zeros((h, w), self.dt)
        coef = np.ascontiguousarray(coef, np.int8)
        self.lib.me_hip_ipflt(typ, _pat(src, src_off), i_src, _p(dst), w, w, h, _p(coef), max_val)
        return dst

    def ipflt_ext(self, typ, src, i_src, src_off, w, h, cx, cy, max_val):
        dst = np.zeros((h, w), self.dt)
        cx = np.ascontiguousarray(cx, np.int8)
        cy = np.ascontiguousarray(cy, np.int8)
        self.lib.me_hip_ipflt_ext(typ, _pat(src, src_off), i_src, _p(dst), w, w, h, _p(cx), _p(cy), max_val)
        return dst

    def ipcpy(self, src, i_src, src_off, w, h):
        dst = np.zeros((h, w), self.dt)
        self.lib.me_hip_ipcpy(_pat(src, src_off), i_src, _p(dst), w, w, h)
        return dst

    def pel_avrg(self, w, h, s1, s2):
        dst = np.zeros((h, w), self.dt)
        s1 = np.ascontiguousarray(s1, self.dt)
        s2 = np.ascontiguousarray(s2, self.dt)
        self.lib.me_hip_pel_avrg(w, _p(dst), w, _p(s1), _p(s2), h)
        return dst

    # -- device references ---------------------------------------------------
    def ref_create(self, luma):
        luma = np.ascontiguousarray(luma, self.dt)
        h, w = luma.shape
        r = C.c_void_p()
        self.check(self.lib.me_ref_create(C.byref(r), w, h), "me_ref_create")
        self.check(self.lib.me_ref_load(r, _p(luma), w, 0, None), "me_ref_load")
        return r

    def ref_load_chroma(self, ref, u, v):
        u = np.ascontiguousarray(u, self.dt)
        v = np.ascontiguousarray(v, self.dt)
        self.check(self.lib.me_ref_load_chroma(ref, _p(u), _p(v), u.shape[1], 0, None), "me_ref_load_chroma")

    def ref_planes(self, ref, w, h, pad=PAD):
        """Download the 16 padded planes of a device reference: (16, H+2pad, W+2pad)."""
        planes = np.zeros((16, h + 2 * pad, w + 2 * pad), self.dt)
        for i in range(16):
            self.check(self.lib.me_ref_download_plane(ref, i, _p(planes[i])), "me_ref_download_plane")
        return planes

    def ref_destroy(self, ref):
        self.lib.me_ref_destroy(ref)

    # -- plans ---------------------------------------------------------------
    def plan_create(self, w, h, R, lambda_mv, mask, num_refs, band_rows=0, waves=0, dx_cols=0, num_refs_l1=0,
                    sp_groups=0, int_grid=0):
        cfg = FrameCfg(w, h, R, lambda_mv, mask, num_refs, band_rows, waves, dx_cols, num_refs_l1, sp_groups,
                       int_grid)
        p = C.c_void_p()
        self.check(self.lib.me_plan_create(C.byref(p), C.byref(cfg)), "me_plan_create")
        self._nl = getattr(self, "_nl", {})
        self._nl[p.value] = 2 if num_refs_l1 else 1
        return p

    def autotune(self, w, h, R, lambda_mv, mask, refs, cur, num_refs_l1=0, iters=3):
        """me_plan_autotune: ({band_rows, waves, dx_cols, int_grid, sp_groups}
        of the fastest schedule, [timings of every candidate])."""
        cfg = FrameCfg(w, h, R, lambda_mv, mask, len(refs), 0, 0, 0, num_refs_l1, 0)
        arr = (C.c_void_p * len(refs))(*[r.value for r in refs])
        cur = np.ascontiguousarray(cur, self.dt)
        log = (TuneResult * 64)()
        n = C.c_int()
        self.check(self.lib.me_plan_autotune(C.byref(cfg), arr, _p(cur), cur.shape[1], iters, log, 64, C.byref(n)),
                   "me_plan_autotune")
        table = [{"band_rows": t.band_rows, "waves": t.waves, "dx_cols": t.dx_cols, "int_grid": t.int_grid,
                  "int_search_ms": t.int_search_ms, "sp_groups": t.sp_groups, "subpel_ms": t.subpel_ms}
                 for t in log[:n.value]]
        return {"band_rows": cfg.band_rows, "waves": cfg.waves, "dx_cols": cfg.dx_cols, "int_grid": cfg.int_grid,
                "sp_groups": cfg.sp_groups}, table

    def plan_set_cur(self, plan, cur):
        cur = np.ascontiguousarray(cur, self.dt)
        self.check(self.lib.me_plan_set_cur(plan, _p(cur), cur.shape[1], 0, None), "me_plan_set_cur")

    def plan_run(self, plan, refs, prepare_mask=0, stream=None):
        arr = (C.c_void_p * len(refs))(*[r.value for r in refs])
        self.check(self.lib.me_plan_run(plan, arr, prepare_mask, stream), "me_plan_run")

    def plan_profile(self, plan, refs, prepare_mask=0, iters=3):
        """Average ms per stage: [prepare, fill, integer search, sub-pel]."""
        arr = (C.c_void_p * len(refs))(*[r.value for r in refs])
        ms = (C.c_double * 4)()
        self.check(self.lib.me_plan_profile(plan, arr, prepare_mask, iters, ms), "me_plan_profile")
        return list(ms)

    def plan_int_grid(self, plan):
        """(workgroups, key slots per column tile) of the plan's integer
        search; 0 slots: one workgroup per band with atomic keys."""
        g, k = C.c_int(), C.c_int()
        self.check(self.lib.me_plan_int_grid(plan, C.byref(g), C.byref(k)), "me_plan_int_grid")
        return g.value, k.value

    def plan_results(self, plan, num_refs):
        """(res [ctu][pu][ref], best [ctu][pu] or [ctu][pu][list] for two lists)."""
        n_ctu = self.lib.me_plan_num_ctus(plan)
        npu = self.lib.me_plan_num_pus(plan)
        nl = getattr(self, "_nl", {}).get(plan.value, 1)
        res = np.zeros((n_ctu, npu, num_refs), PU_RESULT)
        best = np.zeros((n_ctu, npu, 2) if nl == 2 else (n_ctu, npu), np.int8)
        self.check(self.lib.me_plan_results(plan, _p(res), _p(best)), "me_plan_results")
        return res, best

    def plan_best(self, plan, out, best=None, stream=None):
        """me_plan_best into out (PU_BEST [ctu][pu] or [ctu][pu][list]) and,
        optionally, best (int8, the same shape)."""
        assert out.dtype == PU_BEST and out.flags.c_contiguous
        f = self.lib.me_plan_best
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        self.check(f(plan, _p(out), _p(best) if best is not None else None, stream), "me_plan_best")

    def plan_set_cur_raw(self, plan, ptr, stride, stream=None):
        """me_plan_set_cur from a host address (e.g. pinned memory)."""
        f = self.lib.me_plan_set_cur
        self.check(f(plan, C.c_void_p(ptr), C.c_int(stride), C.c_int(0), stream), "me_plan_set_cur")

    def plan_run_tz(self, plan, refs, mvp, tz, mvi_prev=None, stream=None):
        """me_plan_run_tz: mvp int16 [ctu][pu][ref][2] quarter-pel MVPs; tz a
        TzCfg.  Results via plan_results."""
        arr = (C.c_void_p * len(refs))(*[r.value for r in refs])
        mvp = np.ascontiguousarray(mvp, np.int16)
        mvi = np.ascontiguousarray(mvi_prev, np.int16) if mvi_prev is not None else None
        self.check(self.lib.me_plan_run_tz(plan, arr, _p(mvp), _p(mvi) if mvi is not None else None, C.byref(tz),
                                           stream), "me_plan_run_tz")

    def plan_run_bi(self, plan, refs, mvp, mv_scale, pus, tz, stream=None):
        """me_plan_run_bi: mvp, mv_scale int16 [ctu][pu][ref][2] (mv_scale
        updated in place), pus a BI_PU array [ctu][pu] (updated in place)."""
        arr = (C.c_void_p * len(refs))(*[r.value for r in refs])
        mvp = np.ascontiguousarray(mvp, np.int16)
        assert mv_scale.dtype == np.int16 and mv_scale.flags.c_contiguous
        assert pus.dtype == BI_PU and pus.flags.c_contiguous and BI_PU.itemsize == 24
        self.check(self.lib.me_plan_run_bi(plan, arr, _p(mvp), _p(mv_scale), _p(pus), C.byref(tz), stream),
                   "me_plan_run_bi")

    def affine_mc_luma(self, ref, cus, pred):
        """me_affine_mc_luma: cus an AFFINE_CU array, pred a flat array of
        samples (updated in place at each CU's pred_offset)."""
        assert cus.dtype == AFFINE_CU and AFFINE_CU.itemsize == 72 and pred.dtype == self.dt
        cus = np.ascontiguousarray(cus)
        assert pred.flags.c_contiguous
        self.check(self.lib.me_affine_mc_luma(ref, _p(cus), len(cus), _p(pred), pred.size, None),
                   "me_affine_mc_luma")

    def affine_mc_chroma(self, ref, cus, pred_u, pred_v):
        """me_affine_mc_chroma: pred_u / pred_v flat sample arrays updated in
        place at each CU's chroma_offset."""
        assert cus.dtype == AFFINE_CU and pred_u.dtype == self.dt and pred_v.dtype == self.dt
        assert pred_u.size == pred_v.size and pred_u.flags.c_contiguous and pred_v.flags.c_contiguous
        cus = np.ascontiguousarray(cus)
        self.check(self.lib.me_affine_mc_chroma(ref, _p(cus), len(cus), _p(pred_u), _p(pred_v), pred_u.size, None),
                   "me_affine_mc_chroma")

    def affine_me_gradient(self, ref, org, cus, lambda_mv):
        """me_affine_me_gradient: cus an AFFINE_ME_CU array (mv and cost
        updated in place), org a flat array of the CUs' originals."""
        assert cus.dtype == AFFINE_ME_CU and AFFINE_ME_CU.itemsize == 96 and cus.flags.c_contiguous
        org = np.ascontiguousarray(org, self.dt)
        f = self.lib.me_affine_me_gradient
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int, C.c_uint32, C.c_void_p]
        self.check(f(ref, _p(org), org.size, _p(cus), len(cus), lambda_mv, None), "me_affine_me_gradient")

    def plan_run_smvd(self, plan, refs, mvp, mv_scale, pus, tz, stream=None):
        """me_plan_run_smvd: mvp, mv_scale int16 [ctu][pu][ref][2]; pus an
        SMVD_PU array [ctu][pu] (active in, mv / cost out)."""
        arr = (C.c_void_p * len(refs))(*[r.value for r in refs])
        mvp = np.ascontiguousarray(mvp, np.int16)
        mv_scale = np.ascontiguousarray(mv_scale, np.int16)
        assert pus.dtype == SMVD_PU and pus.flags.c_contiguous and SMVD_PU.itemsize == 16
        self.check(self.lib.me_plan_run_smvd(plan, arr, _p(mvp), _p(mv_scale), _p(pus), C.byref(tz), stream),
                   "me_plan_run_smvd")

    def plan_lookahead(self, plan, refs, tz, is_padding, pic_w, pic_h, stream=None):
        """me_plan_lookahead: the lookahead's inter cost of each whole 16x16
        unit, as uint32 [pic_h // 16][pic_w // 16]."""
        arr = (C.c_void_p * len(refs))(*[r.value for r in refs])
        out = np.zeros((pic_h // 16, pic_w // 16), np.uint32)
        self.check(self.lib.me_plan_lookahead(plan, arr, C.byref(tz), int(is_padding), _p(out), stream),
                   "me_plan_lookahead")
        return out

    def plan_mc(self, plan, refs, stream=None):
        """me_plan_mc into the plan's own device buffers."""
        arr = (C.c_void_p * len(refs))(*[r.value for r in refs])
        self.check(self.lib.me_plan_mc(plan, arr, None, None, None, stream), "me_plan_mc")

    def plan_mc_download(self, plan):
        """Flat (pred_y, pred_u, pred_v) of the last me_plan_mc."""
        ly, lc = C.c_uint64(), C.c_uint64()
        self.check(self.lib.me_plan_mc_size(plan, C.byref(ly), C.byref(lc)), "me_plan_mc_size")
        py = np.zeros(ly.value, self.dt)
        pu = np.zeros(max(1, lc.value), self.dt)
        pv = np.zeros(max(1, lc.value), self.dt)
        self.check(self.lib.me_plan_mc_download(plan, _p(py), _p(pu), _p(pv)), "me_plan_mc_download")
        return py, pu, pv

    def plan_destroy(self, plan):
        getattr(self, "_nl", {}).pop(plan.value, None)
        self.lib.me_plan_destroy(plan)


PU_BEST = np.dtype([("mv", "<i2", 2), ("cost", "<u4")])


class Pinned:
    """A numpy array over page-locked host memory from the library's own HIP
    runtime (me_pinned_alloc), for uploads and result downloads."""

    def __init__(self, hip, shape, dtype):
        self.hip = hip
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        self.ptr = C.c_void_p()
        hip.lib.me_pinned_alloc.argtypes = [C.c_void_p, C.c_uint64]
        hip.check(hip.lib.me_pinned_alloc(C.byref(self.ptr), max(1, n)), "me_pinned_alloc")
        buf = (C.c_ubyte * max(1, n)).from_address(self.ptr.value)
        self.array = np.frombuffer(buf, np.uint8, n).view(dtype).reshape(shape)

    def free(self):
        if self.ptr:
            self.hip.lib.me_pinned_free.argtypes = [C.c_void_p]
            self.hip.lib.me_pinned_free(self.ptr)
            self.ptr = None
            self.array = None


def server_picture_bytes(w, h, bd):
    """Device bytes one resident picture takes in the server's HBM budget:
    16 padded planes, the unpadded luma (+64 B read slack) and three s16
    filter intermediates over (w+16) x (h+16)."""
    pel = 1 if bd == 8 else 2
    return 16 * (w + 2 * PAD) * (h + 2 * PAD) * pel + w * h * pel + 64 + 3 * (w + 16) * (h + 16) * 2


class MeServer:
    """The request server of the HIP library (continuous batching of per-CTU
    ME requests from several streams).  Result buffers handed to submit() are
    kept alive here until their ticket is waited for."""

    def __init__(self, hip, R, lambda_mv, mask, num_refs, hbm_budget=64 << 30, max_batch_ctus=64,
                 max_inflight=3, max_wait_us=200, band_rows=0, waves=0, dx_cols=0):
        self.hip = hip
        self.lib = hip.lib
        self.num_refs = num_refs
        self.npu = sum((128 // w) * (128 // h) for s, (w, h) in enumerate(SHAPES) if mask >> s & 1)
        cfg = ServerCfg(R, lambda_mv, mask, num_refs, hbm_budget, max_batch_ctus, max_inflight, max_wait_us,
                        band_rows, waves, dx_cols)
        self.srv = C.c_void_p()
        hip.check(self.lib.me_server_create(C.byref(self.srv), C.byref(cfg)), "me_server_create")
        self._bufs = {}

    def open_stream(self, w, h):
        sid = C.c_int()
        self.hip.check(self.lib.me_server_open_stream(self.srv, w, h, C.byref(sid)), "me_server_open_stream")
        return sid.value

    def put_picture(self, sid, poc, luma, timeout_ms=-1):
        """True when resident, False when the HBM budget did not free up in time."""
        luma = np.ascontiguousarray(luma, self.hip.dt)
        rc = self.lib.me_server_put_picture(self.srv, sid, poc, _p(luma), luma.shape[1], timeout_ms)
        if rc < 0:
            self.hip.check(rc, "me_server_put_picture")
        return rc == 0

    def release_picture(self, sid, poc):
        self.hip.check(self.lib.me_server_release_picture(self.srv, sid, poc), "me_server_release_picture")

    def submit(self, sid, cur_poc, ctu, ref_pocs):
        res = np.zeros((self.npu, self.num_refs), PU_RESULT)
        best = np.zeros(self.npu, np.int8)
        refs = (C.c_int64 * self.num_refs)(*ref_pocs)
        t = self.lib.me_server_submit(self.srv, sid, cur_poc, ctu, refs, _p(res), _p(best))
        if t <= 0:
            raise MeError("me_server_submit failed: %s" % self.lib.me_hip_last_error().decode())
        self._bufs[t] = (res, best)
        return t

    def wait(self, ticket, timeout_ms=-1):
        """(res, best, done_us) of a ticket, or None on timeout."""
        done = C.c_double()
        rc = self.lib.me_server_wait(self.srv, ticket, timeout_ms, C.byref(done))
        if rc == 1:
            return None
        res, best = self._bufs.pop(ticket)
        if rc != 0:
            raise MeError("request %d failed" % ticket)
        return res, best, done.value

    def now_us(self):
        return self.lib.me_server_now_us()

    def stats(self):
        st = ServerStats()
        self.hip.check(self.lib.me_server_stats(self.srv, C.byref(st)), "me_server_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.srv:
            self.lib.me_server_destroy(self.srv)
            self.srv = None


class MeHost:
    """Host C helpers: synthetic frames and the table-driven ME driver."""

    def __init__(self, bd=8):
        path = host_lib_path(bd)
        if not os.path.exists(path):
            raise MeError("host library not built: %s" % path)
        self.bd = bd
        self.dt = pel_dtype(bd)
        L = self.lib = C.CDLL(path)
        L.me_synth_frame.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, C.c_void_p]
        L.me_synth_key.restype = C.c_uint64
        L.me_synth_key.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int]
        L.me_synth_uniform.restype = C.c_double
        L.me_synth_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.me_synth_sin.restype = C.c_double
        L.me_synth_sin.argtypes = [C.c_double]
        L.me_synth_ln.restype = C.c_double
        L.me_synth_ln.argtypes = [C.c_double]
        L.me_synth_gauss.restype = C.c_double
        L.me_synth_gauss.argtypes = [C.c_uint64, C.c_uint64]
        L.me_synth_region_offset.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_void_p]
        L.me_host_frame.restype = C.c_double
        L.me_host_frame.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                    C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
        L.me_host_if_luma_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.me_host_serve_load.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                         C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]

    def synth(self, seed, stream, frame, w, h, sx=0.0, sy=0.0, region_qpel=0, sigma=0.0, noise_purpose=3,
              offset=0.0):
        out = np.zeros((h, w), self.dt)
        self.lib.me_synth_frame(seed, stream, frame, w, h, self.bd, sx, sy, region_qpel, sigma, noise_purpose,
                                offset, _p(out))
        return out

    def key(self, seed, stream, frame, purpose):
        return self.lib.me_synth_key(seed, stream, frame, purpose)

    def uniform(self, key, i):
        return self.lib.me_synth_uniform(key, i)

    def region_offset(self, seed, stream, frame, w, region_qpel, x, y):
        out = np.zeros(2, np.int32)
        self.lib.me_synth_region_offset(seed, stream, frame, w, region_qpel, x, y, _p(out))
        return tuple(int(v) for v in out)

    def serve_load(self, cfg, sizes, seeds, frames, rate, arrival_seed=1, keep_results=False):
        """Drives a request server (created from the ServerCfg `cfg`) with
        one uploader/submitter/collector thread set per stream (host C,
        csrc/serve_load.c).  Returns (latency_us per request, info dict,
        server stats dict, (res, best) or None)."""
        wh = np.array([v for s in sizes for v in s], np.int32)
        sd = np.array(seeds, np.uint64)
        n = sum(((w + 127) // 128) * ((h + 127) // 128) for w, h in sizes) * (frames - 1)
        npu = sum((128 // sw) * (128 // sh) for s, (sw, sh) in enumerate(SHAPES) if cfg.shape_mask >> s & 1)
        lat = np.zeros(n, np.float64)
        info = np.zeros(3, np.float64)
        st = ServerStats()
        res = np.zeros((n, npu, cfg.num_refs), PU_RESULT) if keep_results else None
        best = np.zeros((n, npu), np.int8) if keep_results else None
        rc = self.lib.me_host_serve_load(C.byref(cfg), len(sizes), _p(wh), _p(sd), frames, rate, arrival_seed,
                                         _p(lat), _p(res) if keep_results else None,
                                         _p(best) if keep_results else None, _p(info), C.byref(st))
        if rc:
            raise MeError("me_host_serve_load failed")
        stats = {f: getattr(st, f) for f, _ in ServerStats._fields_}
        return lat, {"wall_us": info[0], "requests": int(info[1]), "synth_s": info[2]}, stats, \
            ((res, best) if keep_results else None)

    def planes_via_table(self, table, luma, pad=PAD):
        """com_if_luma_frame through a dispatch table: (16, H+2pad, W+2pad)."""
        h, w = luma.shape
        stride = w + 2 * pad
        planes = np.zeros((16, h + 2 * pad, stride), self.dt)
        planes[0, pad:pad + h, pad:pad + w] = luma
        ptrs = (C.c_void_p * 16)(*[planes[i].ctypes.data + (pad * stride + pad) * planes.itemsize for i in range(16)])
        rc = self.lib.me_host_if_luma_frame(C.byref(table), ptrs, stride, w, h, pad, self.bd)
        if rc:
            raise MeError("me_host_if_luma_frame failed")
        return planes

    def me_frame(self, table, cur, ref_planes_list, R, lambda_mv, mask, ctu0=0, ctu1=None, threads=1,
                 ctu_ms=None, pad=PAD):
        """Full search + sub-pel through the table for CTUs [ctu0, ctu1)."""
        h, w = cur.shape
        n_ctu = ((w + 127) // 128) * ((h + 127) // 128)
        ctu1 = n_ctu if ctu1 is None else ctu1
        npu = sum((128 // sw) * (128 // sh) for s, (sw, sh) in enumerate(SHAPES) if mask >> s & 1)
        nref = len(ref_planes_list)
        stride = w + 2 * pad
        ptrs = (C.c_void_p * (16 * nref))()
        for r, pl in enumerate(ref_planes_list):
            for i in range(16):
                ptrs[16 * r + i] = pl[i].ctypes.data + (pad * stride + pad) * pl.itemsize
        res = np.zeros((n_ctu, npu, nref), PU_RESULT)
        best = np.full((n_ctu, npu), -1, np.int8)
        cur = np.ascontiguousarray(cur, self.dt)
        ms = self.lib.me_host_frame(C.byref(table), w, h, self.bd, _p(cur), w, nref, ptrs, stride, R, lambda_mv, mask,
                                    ctu0, ctu1, threads, _p(res), _p(best),
                                    _p(ctu_ms) if ctu_ms is not None else None)
        return res, best, ms
