"""ctypes wrappers for the test-side checkers (TEST INFRASTRUCTURE ONLY).

``Oracle(bd)`` loads the C restatement ``oracle/lib/liboracle{bd}.so``;
``RefLib(bd)`` loads the compiled reference ``oracle/_ref/libuavs3e_ref{bd}.so``
(present only where the reference tree was available at build time).
Both expose the same numpy-level helpers so that tests can compare them.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

u8p = C.POINTER(C.c_uint8)
PU_RES_DTYPE = np.dtype([("mv_int", "<i2", 2), ("mv_sub", "<i2", 2), ("cost_int", "<u8"),
                         ("cost_sub", "<u8"), ("cost_final", "<u4"), ("valid", "<u4")])

# The 29 AVS3 PU shapes in the order used by every frame-level API.
SHAPES = [(4, 8), (8, 4), (4, 16), (16, 4), (4, 32), (32, 4), (8, 8), (8, 16), (16, 8), (8, 32),
          (32, 8), (8, 64), (64, 8), (16, 16), (16, 32), (32, 16), (16, 64), (64, 16), (16, 128),
          (128, 16), (32, 32), (32, 64), (64, 32), (32, 128), (128, 32), (64, 64), (64, 128),
          (128, 64), (128, 128)]
MASK_ALL = (1 << 29) - 1
MASK_16x16 = 1 << SHAPES.index((16, 16))
PAD = 144  # PIC_PAD_SIZE_L (com_define.h:293)


def pel_dtype(bd):
    return np.uint8 if bd == 8 else np.uint16


def aligned_zeros(shape, dtype, align=64):
    """Zeroed C-contiguous array whose data pointer is `align`-byte aligned.
    The reference's SIMD kernels use aligned vector stores on prediction
    buffers, which the encoder declares ALIGNED_32 (com_mc.c:142)."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    raw = np.zeros(n + align, np.uint8)
    off = (-raw.ctypes.data) % align
    return raw[off:off + n].view(dtype).reshape(shape)


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def ptr_at(a, off):
    """Pointer to element `off` (flat index) of a C-contiguous array."""
    return C.c_void_p(a.ctypes.data + int(off) * a.itemsize)


class _Lib:
    def __init__(self, path, bd):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.bd = bd
        self.dt = pel_dtype(bd)


class OrcCtx(C.Structure):
    """orc_ctx_t of me_oracle.c."""
    _fields_ = [("pic_w", C.c_int), ("pic_h", C.c_int), ("bd", C.c_int), ("org", C.c_void_p), ("i_org", C.c_int),
                ("planes", C.c_void_p), ("i_ref", C.c_int), ("tab_mvbits", C.c_void_p), ("tab_off", C.c_int),
                ("lambda_mv", C.c_uint32), ("is_padding", C.c_int), ("mvr", C.c_int), ("cost_type", C.c_int),
                ("bi", C.c_int)]


class OrcTz(C.Structure):
    """orc_tz_t of me_oracle.c."""
    _fields_ = [("gop_size", C.c_int), ("max_search_range", C.c_int), ("adaptive_raster_range", C.c_int),
                ("ptr", C.c_longlong), ("ref_ptr", C.c_longlong), ("num_refp", C.c_int), ("mv_ipel", C.c_int16 * 2),
                ("bi_bits", C.c_int)]


class Oracle(_Lib):
    def __init__(self, bd=8):
        super().__init__(os.path.join(HERE, "lib", "liboracle%d.so" % bd), bd)
        L = self.lib
        L.orc_sad.restype = C.c_uint32
        L.orc_satd.restype = C.c_uint32
        L.orc_had.restype = C.c_uint32
        L.orc_lambda_mv.restype = C.c_uint32
        L.orc_luma_taps.restype = C.c_void_p
        L.orc_chroma_taps.restype = C.c_void_p
        L.orc_luma_taps_hp.restype = C.c_void_p
        L.orc_full_search.restype = C.c_uint64
        L.orc_subpel.restype = C.c_uint64
        L.orc_ipel_cost.restype = C.c_uint64
        L.orc_final_cost.restype = C.c_uint32
        L.orc_final_cost.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int]
        L.orc_me_search_tz.restype = C.c_uint64
        L.orc_subpel_ptr.restype = C.c_void_p
        assert L.orc_bit_depth() == bd
        assert L.orc_pu_res_size() == PU_RES_DTYPE.itemsize
        assert L.orc_ctx_size() == C.sizeof(OrcCtx) and L.orc_tz_size() == C.sizeof(OrcTz)

    def search_ctx(self, cur, x, y, planes, lambda_mv, mvr=0, cost_type=0, is_padding=1, pad=PAD, org_bi=None):
        """An orc_ctx_t for the PU at (x, y) of `cur` against padded `planes`
        (with org_bi, an h x w bi-prediction target, a bi = 1 search of it).
        Returns (ctx, keepalive)."""
        h, w = cur.shape
        stride = w + 2 * pad
        ptrs = (C.c_void_p * 16)(*[planes[i].ctypes.data + (pad * stride + pad) * planes.itemsize for i in range(16)])
        off, tab = self.mvbits(w, h)
        if org_bi is None:
            org, i_org = cur.ctypes.data + (y * w + x) * cur.itemsize, w
        else:
            org_bi = np.ascontiguousarray(org_bi, self.dt)
            org, i_org = org_bi.ctypes.data, org_bi.shape[1]
        c = OrcCtx(w, h, self.bd, org, i_org, C.cast(ptrs, C.c_void_p).value,
                   stride, tab.ctypes.data + off, off, lambda_mv, is_padding, mvr, cost_type, int(org_bi is not None))
        return c, (ptrs, tab, cur, planes, org_bi)

    def me_search_tz(self, cur, planes, x, y, w, h, refi, num_refp, cur_poc, ref_poc, gop, msr, mvr, cost_type,
                     adaptive, lambda_mv, mvp, is_padding=1, org_bi=None, mv_start=(0, 0), bi_bits=0):
        """me_search_tz (me.c:529-568); with org_bi the bi = 1 search from
        mv_start (quarter pel) with the other list's mot_bits + 1 = bi_bits."""
        c, keep = self.search_ctx(cur, x, y, planes, lambda_mv, mvr, cost_type, is_padding, org_bi=org_bi)
        t = OrcTz(gop, msr, adaptive, cur_poc, ref_poc, num_refp, (C.c_int16 * 2)(0, 0), bi_bits)
        mvp_a = np.array(mvp, np.int16)
        mv = np.array(mv_start, np.int16)
        mvi = np.zeros(2, np.int16)
        c64 = C.c_uint64()
        r = self.lib.orc_me_search_tz(C.byref(c), C.byref(t), x, y, w, h, refi, ptr(mvp_a), ptr(mv), ptr(mvi),
                                      C.byref(c64))
        self.last_cost64 = c64.value
        return (int(mv[0]), int(mv[1]), int(mvi[0]), int(mvi[1])), r

    def subpel(self, cur, planes, x, y, w, h, mvr, cost_type, lambda_mv, imv, mvp):
        c, keep = self.search_ctx(cur, x, y, planes, lambda_mv, mvr, cost_type)
        mvp_a = np.array(mvp, np.int16)
        mv = np.array([imv[0] * 4, imv[1] * 4], np.int16)
        r = self.lib.orc_subpel(C.byref(c), ptr(mvp_a), x, y, w, h, ptr(mv))
        return (int(mv[0]), int(mv[1])), r

    def analyze_bi(self, cur, planes, n0, x, y, w, h, mvp, mv_scale, refi, mv, lidx_ref, mot_bits, mvr, cost_type,
                   lambda_mv):
        """The search part of analyze_bi (inter.c:972-1033) for the PU at
        (x, y): planes per reference (list 0 first, n0 of them); mvp and
        mv_scale [nref][2] quarter pel (mv_scale updated in place as
        me_search_tz updates pi->mv_scale); refi [2] (index in each list),
        mv [2][2], mot_bits [2] updated in place as cur_info->refi / mv and
        pi->mot_bits.  Returns best_mecost (u64)."""
        pic_h, pic_w = cur.shape
        nref = len(planes)
        nl, base = [n0, nref - n0], [0, n0]
        top = (1 << self.bd) - 1
        off, tab = self.mvbits(pic_w, pic_h)
        org = cur[y:y + h, x:x + w].astype(np.int32)
        best = (1 << 64) - 1
        fixed = lidx_ref                       # the list whose prediction makes the target
        for _ in range(2):
            # com_mc_cu of the fixed list alone, then create_bi_org (inter.c:998-999)
            pred, _, _ = self.mc_pu(pic_w, pic_h, [planes[base[fixed] + refi[fixed]]], None, x, y, w, h,
                                    [tuple(mv[fixed])])
            org_bi = np.clip(2 * org - pred.astype(np.int32), 0, top).astype(self.dt)
            s = 1 - fixed                      # SWAP (inter.c:1001-1002)
            changed = False
            for rc in range(nl[s]):
                r = base[s] + rc
                m, c = self.me_search_tz(cur, planes[r], x, y, w, h, rc, nl[s], 0, 0, 16, 64, mvr, cost_type, 0,
                                         lambda_mv, tuple(mvp[r]), org_bi=org_bi, mv_start=tuple(mv_scale[r]),
                                         bi_bits=int(mot_bits[fixed]) + 1)
                mv_scale[r] = m[:2]
                if c < best:
                    best, changed = c, True
                    refi[s] = rc
                    mv[s] = m[:2]
                    mot_bits[s] = (self.lib.orc_refi_bits(nl[s], rc) + int(tab[off + ((m[0] - mvp[r][0]) >> mvr)]) +
                                   int(tab[off + ((m[1] - mvp[r][1]) >> mvr)]))
            if not changed:
                break
            fixed = s
        return best

    def affine_luma(self, planes, pic_w, pic_h, x, y, cu_w, cu_h, cpmv, sub_w, sub_h, pad=PAD):
        """Affine luma MC from padded planes (com_mc_blk_affine_luma,
        com_mc.c:202-250, or com_mc_blk_affine's luma with three points);
        cpmv [[x0, y0], [x1, y1]] or with [x2, y2] too, 1/16 pel.  Returns
        the (cu_h, cu_w) block."""
        stride = pic_w + 2 * pad
        cpa = np.asarray(cpmv).reshape(-1, 2)
        cp_num = len(cpa)
        cp = (C.c_int * 6)(*[int(v) for v in np.concatenate([cpa.reshape(-1), np.zeros(6 - cpa.size, int)])])
        out = np.zeros((cu_h, cu_w), self.dt)
        self.lib.orc_affine_luma(x, y, pic_w, pic_h, cu_w, cu_h, cp, cp_num,
                                 C.c_void_p(planes[0].ctypes.data + (pad * stride + pad) * planes.itemsize), stride,
                                 ptr(out), sub_w, sub_h, self.bd)
        return out

    def affine_chroma(self, chroma_uv, pic_w, pic_h, x, y, cu_w, cu_h, cpmv, sub, cpad=PAD // 2):
        """The chroma of com_mc_blk_affine (com_mc.c:290-360) from padded
        chroma planes (make_chroma); square sub-blocks sub = 4 or 8.
        Returns (u, v) blocks of (cu_h/2, cu_w/2)."""
        cs = pic_w // 2 + 2 * cpad
        cpa = np.asarray(cpmv).reshape(-1, 2)
        cp = (C.c_int * 6)(*[int(v) for v in np.concatenate([cpa.reshape(-1), np.zeros(6 - cpa.size, int)])])
        u = np.zeros((cu_h // 2, cu_w // 2), self.dt)
        v = np.zeros((cu_h // 2, cu_w // 2), self.dt)
        off = (cpad * cs + cpad) * self.dt().itemsize
        self.lib.orc_affine_chroma(x, y, pic_w, pic_h, cu_w, cu_h, cp, len(cpa), sub,
                                   C.c_void_p(chroma_uv[0].ctypes.data + off), C.c_void_p(chroma_uv[1].ctypes.data + off),
                                   cs, ptr(u), ptr(v), self.bd)
        return u, v

    def affine_me_gradient(self, planes, pic_w, pic_h, org, x, y, cu_w, cu_h, sub_w, sub_h, refi, num_refp, mvr, bi,
                           mot_bits_other, lambda_mv, mvp, mv, pad=PAD):
        """affine_me_gradient (inter.c:1339-1471) from padded planes: org the
        CU's (cu_h, cu_w) original (or bi target); mvp, mv [[x0, y0], [x1,
        y1]] in 1/16 pel.  Returns (cost without the MV bits, best mv)."""
        stride = pic_w + 2 * pad
        org = np.ascontiguousarray(org, self.dt)
        mvp_a = (C.c_int * 4)(*[int(v) for v in np.asarray(mvp).reshape(-1)])
        mv_a = (C.c_int * 4)(*[int(v) for v in np.asarray(mv).reshape(-1)])
        f = self.lib.orc_affine_me_gradient
        f.restype = C.c_uint64
        c = f(x, y, pic_w, pic_h, cu_w, cu_h, sub_w, sub_h,
              C.c_void_p(planes[0].ctypes.data + (pad * stride + pad) * planes.itemsize), stride, ptr(org),
              org.shape[1], refi, num_refp, mvr, bi, mot_bits_other, C.c_uint32(lambda_mv), mvp_a, mv_a, self.bd)
        return int(c), [[mv_a[0], mv_a[1]], [mv_a[2], mv_a[3]]]

    def luma_taps_hp(self, frac):
        return np.ctypeslib.as_array(C.cast(self.lib.orc_luma_taps_hp(frac), C.POINTER(C.c_int8)), (8,)).copy()

    def analyze_smvd(self, cur, planes, n0, x, y, w, h, mvp, mv_scale, mvr, cost_type, lambda_mv):
        """The search part of analyze_smvd (inter.c:1194-1255) with
        smvd_refine (inter.c:1083-1191) for the PU at (x, y): planes per
        reference (list 0 first, n0 of them; reference 0 of each list is
        used), mvp / mv_scale [nref][2] quarter pel.  Returns (mv [2][2],
        mecost)."""
        pic_h, pic_w = cur.shape
        off, tab = self.mvbits(pic_w, pic_h)
        org = np.ascontiguousarray(cur[y:y + h, x:x + w])
        p0, p1 = [int(v) for v in mvp[0]], [int(v) for v in mvp[n0]]
        M = 0xFFFFFFFF

        def bits(v):   # tab_mvbits[v >> mvr] (GET_MVBITS_X, util.h:41)
            return int(tab[off + max(-off, min(off, v >> mvr))])

        def mv_cost(b):   # MV_COST (inter.c:41), u32 arithmetic
            return (((lambda_mv * b) & M) + 32768 & M) >> 16

        def s16(v):
            return (v + 32768) % 65536 - 32768

        def cost(m0, m1, had):   # get_smvd_cost of com_mc_cu's bi-average (inter.c:1071-1081)
            pred, _, _ = self.mc_pu(pic_w, pic_h, [planes[0], planes[n0]], None, x, y, w, h, [tuple(m0), tuple(m1)])
            if had:
                return self.had(w, h, org, w, pred, w)
            return self.sad(w, h, org, w, pred, w) >> (self.bd - 8)

        had = cost_type == 0
        mv = [list(p0), list(p1)]
        best = (cost(mv[0], mv[1], had) + mv_cost(bits(0) * 2 + mvr + (mvr < 4))) & M
        m0 = [int(v) for v in mv_scale[0]]
        d = [m0[0] - p0[0], m0[1] - p0[1]]
        m1 = [s16(p1[0] - d[0]), s16(p1[1] - d[1])]
        cb = (cost(m0, m1, had) + mv_cost(bits(d[0]) + bits(d[1]) + mvr + (mvr < 4))) & M
        if cb < best:
            best, mv = cb, [m0, m1]
        cross = [(0, 1), (1, 0), (0, -1), (-1, 0)]
        diamond = [(0, 2), (1, 1), (2, 0), (1, -1), (0, -2), (-1, -1), (-2, 0), (-1, 1)]
        for pattern, rounds in ((2, 8), (0, 1)):
            offs, de, rnd, mask = (cross, 3, 4, 3) if pattern == 0 else (diamond, 7, 8, 7)
            ds = 0
            for _ in range(rounds):
                c = list(mv[0])
                best_dir = -1
                for idx in range(ds, de + 1):
                    dr = (idx + rnd) & mask
                    c0 = [s16(c[0] + s16(offs[dr][0] << mvr)), s16(c[1] + s16(offs[dr][1] << mvr))]
                    dd = [c0[0] - p0[0], c0[1] - p0[1]]
                    c1 = [s16(p1[0] - dd[0]), s16(p1[1] - dd[1])]
                    t = (cost(c0, c1, had) + mv_cost(mvr + (mvr < 4) + bits(dd[0]) + bits(dd[1]))) & M
                    if t < best:
                        best, mv, best_dir = t, [c0, c1], dr
                if best_dir < 0:
                    break
                step = 2 - (best_dir & 1) if pattern == 2 else 1
                ds, de = best_dir - step, best_dir + step
        if not had:
            dd = [mv[0][0] - p0[0], mv[0][1] - p0[1]]
            best = (cost(mv[0], mv[1], True) + mv_cost(bits(dd[0]) + bits(dd[1]) + mvr + (mvr < 4))) & M
        return mv, best

    def loka_inter_cost(self, cur, planes_l0, planes_l1, cur_poc, poc_l0, poc_l1, gop, msr, is_padding, pad=PAD):
        """The inter part of loka_estimate_coding_cost (lookahead.c:82-208)
        for every whole 16x16 unit of `cur`, uint32 [h // 16][w // 16]:
        me_search_tz from MVP 0 with curr_mvr 2 and the lookahead's lambda
        (lookahead.c:85-110, 150-152), com_mc_blk_luma at the integer MV
        (153), com_had (155), the first cheapest per list (156-170), then
        with both lists the pel_avrg of the two kept predictions (172-179)
        and of both lists' first reference at MV 0 (181-207)."""
        h, w = cur.shape
        stride = w + 2 * pad
        lam = int(np.floor(65536.0 * np.sqrt(1.43631 * 2.0 ** ((32 - 16.0) / 4.0))))
        lists = [(planes_l0, poc_l0), (planes_l1, poc_l1)]
        is_bi = len(planes_l0) > 0 and len(planes_l1) > 0
        out = np.zeros((h // 16, w // 16), np.uint32)
        org_all = np.ascontiguousarray(cur)

        def block(planes, bx, by):
            return np.ascontiguousarray(planes[0][pad + by:pad + by + 16, pad + bx:pad + bx + 16])

        for uy in range(h // 16):
            for ux in range(w // 16):
                x, y = ux * 16, uy * 16
                org = np.ascontiguousarray(org_all[y:y + 16, x:x + 16])
                best = (1 << 64) - 1
                kept = [None, None]
                for lidx, (pl, pocs) in enumerate(lists):
                    best_l = (1 << 64) - 1
                    for refi, planes in enumerate(pl):
                        mv, _ = self.me_search_tz(cur, planes, x, y, 16, 16, refi, len(pl), cur_poc, pocs[refi],
                                                  gop, msr, 2, 0, 0, lam, (0, 0), is_padding)
                        pred = block(planes, x + (mv[0] >> 2), y + (mv[1] >> 2))
                        c = self.had(16, 16, org, 16, pred, 16)
                        if c < best_l:
                            best_l = c
                            best = min(best, c)
                            kept[lidx] = pred
                if is_bi:
                    c = self.had(16, 16, org, 16, self.pel_avrg(16, 16, kept[0], kept[1]), 16)
                    best = min(best, c)
                    z = self.pel_avrg(16, 16, block(planes_l0[0], x, y), block(planes_l1[0], x, y))
                    best = min(best, self.had(16, 16, org, 16, z, 16))
                out[uy, ux] = min(best, 0xFFFFFFFF)
        return out

    # -- block costs --------------------------------------------------------
    def sad(self, w, h, org, i_org, pred, i_pred, org_off=0, pred_off=0):
        return self.lib.orc_sad(w, ptr_at(org, org_off), i_org, ptr_at(pred, pred_off), i_pred, h)

    def satd(self, wl, hl, org, s_org, cur, s_cur, org_off=0, cur_off=0):
        return self.lib.orc_satd(wl, hl, ptr_at(org, org_off), s_org, ptr_at(cur, cur_off), s_cur)

    def had(self, w, h, org, s_org, cur, s_cur, org_off=0, cur_off=0):
        return self.lib.orc_had(w, h, ptr_at(org, org_off), s_org, ptr_at(cur, cur_off), s_cur, self.bd)

    # -- interpolation -----------------------------------------------------
    def luma_taps(self, frac):
        return np.ctypeslib.as_array(C.cast(self.lib.orc_luma_taps(frac), C.POINTER(C.c_int8)), (8,)).copy()

    def chroma_taps(self, frac):
        return np.ctypeslib.as_array(C.cast(self.lib.orc_chroma_taps(frac), C.POINTER(C.c_int8)), (4,)).copy()

    def ipflt(self, typ, src, i_src, src_off, w, h, coef, max_val):
        dst = aligned_zeros((h, w), self.dt)
        coef = np.ascontiguousarray(coef, np.int8)
        self.lib.orc_ipflt(typ, ptr_at(src, src_off), i_src, ptr(dst), w, w, h, ptr(coef), max_val)
        return dst

    def ipflt_ext(self, typ, src, i_src, src_off, w, h, cx, cy, max_val):
        dst = aligned_zeros((h, w), self.dt)
        cx = np.ascontiguousarray(cx, np.int8)
        cy = np.ascontiguousarray(cy, np.int8)
        self.lib.orc_ipflt_ext(typ, ptr_at(src, src_off), i_src, ptr(dst), w, w, h, ptr(cx), ptr(cy), max_val)
        return dst

    def pel_avrg(self, w, h, s1, s2):
        dst = aligned_zeros((h, w), self.dt)
        self.lib.orc_pel_avrg(w, ptr(dst), w, ptr(s1), ptr(s2), h)
        return dst

    def pad_plane(self, buf, stride, w, h, pad, ext):
        self.lib.orc_pad_plane(ptr_at(buf, pad * stride + pad), stride, w, h, pad, ext)

    def make_planes(self, luma, pad=PAD):
        """Padded integer plane + 15 sub-pel planes, as (16, H+2pad, W+2pad)."""
        h, w = luma.shape
        stride = w + 2 * pad
        planes = np.zeros((16, h + 2 * pad, stride), self.dt)
        planes[0, pad:pad + h, pad:pad + w] = luma
        self.pad_plane(planes[0], stride, w, h, pad, 0)
        ptrs = (C.c_void_p * 16)(*[C.c_void_p(planes[i].ctypes.data + (pad * stride + pad) * planes.itemsize)
                                    for i in range(16)])
        self.lib.orc_if_luma_frame(ptrs, stride, w, h, pad, self.bd)
        return planes

    # -- rate --------------------------------------------------------------
    def mvbits(self, pic_w, pic_h):
        off = self.lib.orc_mvbits_offset(pic_w, pic_h)
        tab = np.zeros(2 * off + 1, np.uint8)
        self.lib.orc_mvbits_fill(ptr(tab), off)
        return off, tab

    def lambda_mv(self, qp):
        return self.lib.orc_lambda_mv(qp)

    # -- frame -------------------------------------------------------------
    @staticmethod
    def _plane_ptrs(ref_planes_list, w, pad):
        stride = w + 2 * pad
        ptrs = (C.c_void_p * (16 * len(ref_planes_list)))()
        for r, planes in enumerate(ref_planes_list):
            for i in range(16):
                ptrs[16 * r + i] = planes[i].ctypes.data + (pad * stride + pad) * planes.itemsize
        return ptrs, stride

    def me_frame(self, cur, ref_planes_list, R, lambda_mv, mask, ctu0=0, ctu1=None, cost_type=0, pad=PAD,
                 num_refs_l1=0):
        """Full search + sub-pel for CTUs [ctu0, ctu1); returns (res, best_ref)
        with res shaped (n_ctu, npu, nref) of PU_RES_DTYPE and best_ref
        (n_ctu, npu), or (n_ctu, npu, 2) when the last num_refs_l1 references
        form list 1."""
        h, w = cur.shape
        n_ctu = ((w + 127) // 128) * ((h + 127) // 128)
        ctu1 = n_ctu if ctu1 is None else ctu1
        npu = self.lib.orc_pus_per_ctu(C.c_uint32(mask))
        nref = len(ref_planes_list)
        ptrs, stride = self._plane_ptrs(ref_planes_list, w, pad)
        res = np.zeros((n_ctu, npu, nref), PU_RES_DTYPE)
        best = np.full((n_ctu, npu, 2) if num_refs_l1 else (n_ctu, npu), -1, np.int8)
        cur = np.ascontiguousarray(cur)
        self.lib.orc_me_frame_lists(w, h, self.bd, ptr(cur), w, nref, num_refs_l1, ptrs, stride, R,
                                    C.c_uint32(lambda_mv), cost_type, C.c_uint32(mask), ctu0, ctu1, ptr(res),
                                    ptr(best))
        return res, best

    def make_chroma(self, plane, pad=PAD // 2):
        """A chroma plane edge-padded by `pad` (PIC_PAD_SIZE_C, com_img_padding
        with ext 0): (H/2 + 2pad, W/2 + 2pad)."""
        ch, cw = plane.shape
        out = np.zeros((ch + 2 * pad, cw + 2 * pad), self.dt)
        out[pad:pad + ch, pad:pad + cw] = plane
        self.lib.orc_pad_plane(ptr_at(out, pad * out.shape[1] + pad), out.shape[1], cw, ch, pad, 0)
        return out

    def mc_pu(self, pic_w, pic_h, planes_l, chroma_l, x, y, w, h, mvs, pad=PAD, cpad=PAD // 2):
        """com_mc_cu of one PU from len(planes_l) lists (luma planes per list,
        (u, v) padded chroma per list or None, quarter-pel MV per list);
        returns (pred_y, pred_u, pred_v) (chroma None without chroma)."""
        nl = len(planes_l)
        stride = pic_w + 2 * pad
        keep = []
        pl = (C.c_void_p * nl)()
        for l, planes in enumerate(planes_l):
            a = (C.c_void_p * 16)(*[planes[i].ctypes.data + (pad * stride + pad) * planes.itemsize for i in range(16)])
            keep.append(a)
            pl[l] = C.cast(a, C.c_void_p).value
        do_c = chroma_l is not None
        cs = pic_w // 2 + 2 * cpad
        ul = (C.c_void_p * nl)(*[chroma_l[l][0].ctypes.data + (cpad * cs + cpad) * self.dt().itemsize
                                 for l in range(nl)]) if do_c else None
        vl = (C.c_void_p * nl)(*[chroma_l[l][1].ctypes.data + (cpad * cs + cpad) * self.dt().itemsize
                                 for l in range(nl)]) if do_c else None
        mv = np.array(mvs, np.int16).reshape(-1)
        py = np.zeros((h, w), self.dt)
        pu = np.zeros((h // 2, w // 2), self.dt)
        pv = np.zeros((h // 2, w // 2), self.dt)
        self.lib.orc_mc_pu(pic_w, pic_h, self.bd, nl, pl, stride, ul, vl, cs, x, y, w, h, ptr(mv), int(do_c),
                           ptr(py), ptr(pu), ptr(pv))
        return py, (pu if do_c else None), (pv if do_c else None)

    def mc_frame(self, pic_w, pic_h, ref_planes_list, chroma_list, num_refs_l1, mask, res, best, pad=PAD,
                 cpad=PAD // 2):
        """Predictions of every PU from ME results (orc_mc_frame layout):
        returns flat (pred_y, pred_u, pred_v).  chroma_list: padded (u, v)
        per reference."""
        n_ctu = res.shape[0]
        nsel = sum(1 for s in range(29) if mask >> s & 1)
        nsel_c = sum(1 for s in range(29) if mask >> s & 1 and SHAPES[s][0] >= 8 and SHAPES[s][1] >= 8)
        ptrs, stride = self._plane_ptrs(ref_planes_list, pic_w, pad)
        cs = pic_w // 2 + 2 * cpad
        isz = np.dtype(self.dt).itemsize
        ul = (C.c_void_p * len(chroma_list))(*[u.ctypes.data + (cpad * cs + cpad) * isz for u, v in chroma_list])
        vl = (C.c_void_p * len(chroma_list))(*[v.ctypes.data + (cpad * cs + cpad) * isz for u, v in chroma_list])
        py = np.zeros(n_ctu * nsel * 16384, self.dt)
        pu = np.zeros(max(1, n_ctu * nsel_c * 4096), self.dt)
        pv = np.zeros(max(1, n_ctu * nsel_c * 4096), self.dt)
        res = np.ascontiguousarray(res)
        best = np.ascontiguousarray(best)
        self.lib.orc_mc_frame(pic_w, pic_h, self.bd, res.shape[2], num_refs_l1, ptrs, stride, ul, vl, cs,
                              C.c_uint32(mask), 0, n_ctu, ptr(res), ptr(best), ptr(py), ptr(pu), ptr(pv))
        return py, pu, pv


class RefLib(_Lib):
    """The compiled reference (oracle/_ref).  level: 0 = C table, 2 = C+SSE+AVX2."""

    @staticmethod
    def available(bd=8):
        return os.path.exists(os.path.join(HERE, "_ref", "libuavs3e_ref%d.so" % bd))

    def __init__(self, bd=8, level=0):
        super().__init__(os.path.join(HERE, "_ref", "libuavs3e_ref%d.so" % bd), bd)
        L = self.lib
        for n in ("refh_sad", "refh_satd", "refh_had"):
            getattr(L, n).restype = C.c_uint32
        L.refh_pic_create.restype = C.c_void_p
        L.refh_pic_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_longlong]
        L.refh_pic_plane.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.refh_pic_stride.argtypes = [C.c_void_p]
        L.refh_pic_free.argtypes = [C.c_void_p]
        L.refh_luma_coef.restype = C.c_void_p
        L.refh_chroma_coef.restype = C.c_void_p
        L.refh_me_search_tz_bi.restype = C.c_uint64
        L.refh_me_search_tz_bi.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32,
                                           C.c_void_p, C.c_void_p]
        L.refh_analyze_bi.restype = None
        L.refh_analyze_bi.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_uint32]
        L.refh_smvd.restype = C.c_uint32
        L.refh_smvd.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_void_p]
        L.refh_loka_cost.restype = C.c_double
        L.refh_loka_cost.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_longlong, C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_int, C.c_int, C.c_int, C.c_int]
        L.refh_me_search_tz.restype = C.c_uint64
        L.refh_me_search_tz.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.refh_subpel.restype = C.c_uint64
        L.refh_subpel.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_uint32, C.c_void_p, C.c_void_p]
        L.refh_me_frame.restype = C.c_double
        L.refh_me_frame.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                    C.c_uint32, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_void_p]
        L.refh_tz_frame.restype = C.c_double
        L.refh_tz_frame.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                    C.c_longlong, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                    C.c_void_p]
        L.refh_set_pinning.argtypes = [C.c_int]
        self.set_level(level)

    def set_level(self, level):
        assert self.lib.refh_init(level) == self.bd
        self.level = level

    def sad(self, w, h, org, i_org, pred, i_pred, org_off=0, pred_off=0):
        return self.lib.refh_sad(w, ptr_at(org, org_off), i_org, ptr_at(pred, pred_off), i_pred, h)

    def sad_x4(self, w, h, org, i_org, pred, i_pred, offs, org_off=0):
        out = np.zeros(4, np.uint32)
        self.lib.refh_sad_x4(w, ptr_at(org, org_off), i_org, *[ptr_at(pred, o) for o in offs], i_pred, ptr(out), h)
        return out

    def sad_x3(self, w, h, org, i_org, pred, i_pred, offs, org_off=0):
        out = np.zeros(3, np.uint32)
        self.lib.refh_sad_x3(w, ptr_at(org, org_off), i_org, *[ptr_at(pred, o) for o in offs], i_pred, ptr(out), h)
        return out

    def luma_coef(self, frac):
        return np.ctypeslib.as_array(C.cast(self.lib.refh_luma_coef(frac), C.POINTER(C.c_int8)), (8,)).copy()

    def chroma_coef(self, frac):
        return np.ctypeslib.as_array(C.cast(self.lib.refh_chroma_coef(frac), C.POINTER(C.c_int8)), (4,)).copy()

    def satd(self, wl, hl, org, s_org, cur, s_cur, org_off=0, cur_off=0):
        return self.lib.refh_satd(wl, hl, ptr_at(org, org_off), s_org, ptr_at(cur, cur_off), s_cur)

    def had(self, w, h, org, s_org, cur, s_cur, org_off=0, cur_off=0):
        return self.lib.refh_had(w, h, ptr_at(org, org_off), s_org, ptr_at(cur, cur_off), s_cur, self.bd)

    def ipflt(self, typ, src, i_src, src_off, w, h, coef, max_val):
        dst = aligned_zeros((h, w), self.dt)
        coef = np.ascontiguousarray(coef, np.int8)
        self.lib.refh_ipflt(typ, ptr_at(src, src_off), i_src, ptr(dst), w, w, h, ptr(coef), max_val)
        return dst

    def ipflt_ext(self, typ, src, i_src, src_off, w, h, cx, cy, max_val):
        dst = aligned_zeros((h, w), self.dt)
        cx = np.ascontiguousarray(cx, np.int8)
        cy = np.ascontiguousarray(cy, np.int8)
        self.lib.refh_ipflt_ext(typ, ptr_at(src, src_off), i_src, ptr(dst), w, w, h, ptr(cx), ptr(cy), max_val)
        return dst

    def pel_avrg(self, w, h, s1, s2):
        dst = aligned_zeros((h, w), self.dt)
        a1 = aligned_zeros((h, w), self.dt); a1[...] = s1
        a2 = aligned_zeros((h, w), self.dt); a2[...] = s2
        self.lib.refh_pel_avrg(w, ptr(dst), w, ptr(a1), ptr(a2), h)
        return dst

    def mvbits(self, pic_w, pic_h):
        cap = ((max(pic_w, pic_h) << 2) + 3) * 2 + 1
        tab = np.zeros(cap, np.uint8)
        off = self.lib.refh_mvbits(pic_w, pic_h, ptr(tab), cap)
        return off, tab[:2 * off + 1]

    def pic_create(self, luma, poc=0):
        luma = np.ascontiguousarray(luma, self.dt)
        h, w = luma.shape
        return self.lib.refh_pic_create(ptr(luma), w, h, poc)

    def pic_planes(self, pic, w, h, pad=PAD):
        stride = w + 2 * pad
        assert self.lib.refh_pic_stride(pic) == stride
        planes = np.zeros((16, h + 2 * pad, stride), self.dt)
        for i in range(16):
            self.lib.refh_pic_plane(pic, i, ptr(planes[i]))
        return planes

    def pic_free(self, pic):
        self.lib.refh_pic_free(pic)

    def pic_set_chroma(self, pic, u, v):
        u = np.ascontiguousarray(u, self.dt)
        v = np.ascontiguousarray(v, self.dt)
        self.lib.refh_pic_set_chroma(C.c_void_p(pic), ptr(u), ptr(v))

    def mc_cu(self, x, y, pic_w, pic_h, w, h, p0, p1, mv0, mv1, chroma=True):
        """com_mc_cu of one PU (p0 / p1: L0 / L1 pictures or None)."""
        py = np.zeros((h, w), self.dt)
        pu = np.zeros((h // 2, w // 2), self.dt)
        pv = np.zeros((h // 2, w // 2), self.dt)
        m0 = np.array(mv0 if mv0 is not None else (0, 0), np.int16)
        m1 = np.array(mv1 if mv1 is not None else (0, 0), np.int16)
        self.lib.refh_mc_cu(x, y, pic_w, pic_h, w, h, C.c_void_p(p0) if p0 else None, C.c_void_p(p1) if p1 else None,
                            ptr(m0), ptr(m1), 0 if chroma else 1, ptr(py), ptr(pu), ptr(pv))
        return py, (pu if chroma else None), (pv if chroma else None)

    def affine_cu_luma(self, pic, x, y, cu_w, cu_h, cpmv, sub8):
        """Luma of com_mc_cu_affine (com_mc.c:365-385) of the compiled
        reference, one list, two or three control points; sub8 selects 8x8
        sub-blocks (affine_subblk_size_idx)."""
        cpa = np.asarray(cpmv, np.int32).reshape(-1, 2)
        cp = np.zeros(8, np.int32)
        cp[:cpa.size] = cpa.reshape(-1)
        out = np.zeros((cu_h, cu_w), self.dt)
        self.lib.refh_affine_cu(C.c_void_p(pic), x, y, cu_w, cu_h, ptr(cp), len(cpa), int(sub8), ptr(out), None, None)
        return out

    def affine_cu(self, pic, x, y, cu_w, cu_h, cpmv, sub8):
        """com_mc_cu_affine of the compiled reference: (luma, u, v)."""
        cpa = np.asarray(cpmv, np.int32).reshape(-1, 2)
        cp = np.zeros(8, np.int32)
        cp[:cpa.size] = cpa.reshape(-1)
        out = np.zeros((cu_h, cu_w), self.dt)
        u = np.zeros((cu_h // 2, cu_w // 2), self.dt)
        v = np.zeros((cu_h // 2, cu_w // 2), self.dt)
        self.lib.refh_affine_cu(C.c_void_p(pic), x, y, cu_w, cu_h, ptr(cp), len(cpa), int(sub8), ptr(out), ptr(u),
                                ptr(v))
        return out, u, v

    def affine_luma(self, pic, x, y, cu_w, cu_h, cpmv, sub_w, sub_h):
        """com_mc_blk_affine_luma of the compiled reference."""
        cp = np.ascontiguousarray(np.asarray(cpmv, np.int32).reshape(-1))
        out = np.zeros((cu_h, cu_w), self.dt)
        self.lib.refh_affine_luma(C.c_void_p(pic), x, y, cu_w, cu_h, ptr(cp), sub_w, sub_h, ptr(out))
        return out

    def me_search_tz_bi(self, pic, org_bi, x, y, refi, num_refp, lidx, mot_bits_other, mvr, cost_type, lambda_mv,
                        mvp, mv):
        """me_search_tz(bi = 1) of the compiled reference; returns (cost, mv)."""
        org_bi = np.ascontiguousarray(org_bi, self.dt)
        h, w = org_bi.shape
        mvp_a = np.array(mvp, np.int16)
        mv_a = np.array(mv, np.int16)
        c = self.lib.refh_me_search_tz_bi(pic, ptr(org_bi), w, x, y, w, h, refi, num_refp, lidx, mot_bits_other,
                                          mvr, cost_type, lambda_mv, ptr(mvp_a), ptr(mv_a))
        return c, (int(mv_a[0]), int(mv_a[1]))

    def analyze_bi(self, cur, pics_l0, pics_l1, x, y, w, h, mvp, mv_scale, refi, mv, lidx_ref, mot_bits, mvr,
                   cost_type, lambda_mv):
        """The compiled reference's own analyze_bi (inter.c:972-1070) up to
        the end of its search, through a minimal core_t: mvp / mv_scale
        [2][4][2] (per list and reference index), refi [2], mv [2][2],
        mot_bits [2].  Returns (refi, mv, mot_bits, mv_scale) after the
        search, as cur_info->refi / mv, pi->mot_bits and pi->mv_scale."""
        cur = np.ascontiguousarray(cur, self.dt)
        ph, pw = cur.shape
        a0 = (C.c_void_p * max(1, len(pics_l0)))(*pics_l0)
        a1 = (C.c_void_p * max(1, len(pics_l1)))(*pics_l1)
        mvp_a = np.ascontiguousarray(mvp, np.int16).reshape(2, 4, 2)
        ms = np.ascontiguousarray(mv_scale, np.int16).reshape(2, 4, 2).copy()
        ri = np.array(refi, np.int8)
        mv_a = np.array(mv, np.int16).reshape(2, 2).copy()
        mb = np.array(mot_bits, np.int32)
        self.lib.refh_analyze_bi(ptr(cur), pw, ph, a0, len(pics_l0), a1, len(pics_l1), x, y, w, h, ptr(mvp_a), ptr(ms),
                                 ptr(ri), ptr(mv_a), lidx_ref, ptr(mb), mvr, cost_type, C.c_uint32(lambda_mv))
        return ri, mv_a, mb, ms

    def smvd(self, cur, p0, p1, x, y, w, h, mvp, mv_scale_l0, mvr, cost_type, lambda_mv):
        """analyze_smvd's search (inter.c:1194-1255) with the reference's own
        smvd_refine: mvp [2][2] (reference 0 of each list), mv_scale_l0
        [2].  Returns (mv [2][2], mecost)."""
        cur = np.ascontiguousarray(cur, self.dt)
        ph, pw = cur.shape
        mvp_a = np.ascontiguousarray(mvp, np.int16).reshape(2, 2)
        ms = np.array(mv_scale_l0, np.int16)
        out = np.zeros((2, 2), np.int16)
        c = self.lib.refh_smvd(ptr(cur), pw, ph, C.c_void_p(p0), C.c_void_p(p1), x, y, w, h, ptr(mvp_a), ptr(ms), mvr,
                               cost_type, C.c_uint32(lambda_mv), ptr(out))
        return out, int(c)

    def affine_me_gradient(self, pic, org, x, y, cu_w, cu_h, sub_w, sub_h, refi, num_refp, mvr, bi, mot_bits_other,
                           lambda_mv, mvp, mv):
        """The compiled reference's affine_me_gradient (inter.c:1339-1471);
        returns (cost without the MV bits, best mv [[x0, y0], [x1, y1]])."""
        org = np.ascontiguousarray(org, self.dt)
        mvp_a = (C.c_int * 4)(*[int(v) for v in np.asarray(mvp).reshape(-1)])
        mv_a = (C.c_int * 4)(*[int(v) for v in np.asarray(mv).reshape(-1)])
        f = self.lib.refh_affine_me_gradient
        f.restype = C.c_uint64
        f.argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 12 + [C.c_uint32, C.c_void_p, C.c_void_p]
        c = f(pic, ptr(org), org.shape[1], x, y, cu_w, cu_h, sub_w, sub_h, refi, num_refp, mvr, bi, mot_bits_other,
              lambda_mv, mvp_a, mv_a)
        return int(c), [[mv_a[0], mv_a[1]], [mv_a[2], mv_a[3]]]

    def loka_cost(self, cur, pics_l0, pics_l1, cur_poc, gop, msr, is_padding):
        """loka_estimate_coding_cost (lookahead.c:82-290) of the compiled
        reference: the picture's summed least inter / intra unit cost."""
        h, w = cur.shape
        cur = np.ascontiguousarray(cur, self.dt)
        a0 = (C.c_void_p * 4)(*pics_l0)
        a1 = (C.c_void_p * 4)(*pics_l1)
        return self.lib.refh_loka_cost(ptr(cur), w, h, cur_poc, a0, len(pics_l0), a1, len(pics_l1), gop, msr,
                                       is_padding)

    def set_pinning(self, on):
        """Pin the frame drivers' worker threads one per CPU of the process's
        affinity set."""
        self.lib.refh_set_pinning(1 if on else 0)

    def tz_frame(self, cur, pics, gop_size, max_search_range, cur_poc, lambda_mv, mask, ctu0=0, ctu1=None, threads=1):
        """me_search_tz from MVP (0,0) for every PU of CTUs [ctu0, ctu1) of
        the compiled reference: (res, best, ms); res carries mv_int, mv_sub,
        cost_final (the u32 me_search_tz returns) and valid."""
        h, w = cur.shape
        n_ctu = ((w + 127) // 128) * ((h + 127) // 128)
        ctu1 = n_ctu if ctu1 is None else ctu1
        npu = sum((128 // sw) * (128 // sh) for s, (sw, sh) in enumerate(SHAPES) if mask >> s & 1)
        nref = len(pics)
        arr = (C.c_void_p * nref)(*pics)
        res = np.zeros((n_ctu, npu, nref), PU_RES_DTYPE)
        best = np.full((n_ctu, npu), -1, np.int8)
        cur = np.ascontiguousarray(cur, self.dt)
        ms = self.lib.refh_tz_frame(w, h, ptr(cur), w, nref, arr, gop_size, max_search_range, cur_poc, lambda_mv, mask,
                                    ctu0, ctu1, threads, ptr(res), ptr(best))
        return res, best, ms

    def me_frame(self, cur, pics, R, lambda_mv, mask, ctu0=0, ctu1=None, cost_type=0, threads=1, ctu_ms=None):
        h, w = cur.shape
        n_ctu = ((w + 127) // 128) * ((h + 127) // 128)
        ctu1 = n_ctu if ctu1 is None else ctu1
        npu = sum((128 // sw) * (128 // sh) for s, (sw, sh) in enumerate(SHAPES) if mask >> s & 1)
        nref = len(pics)
        arr = (C.c_void_p * nref)(*pics)
        res = np.zeros((n_ctu, npu, nref), PU_RES_DTYPE)
        best = np.full((n_ctu, npu), -1, np.int8)
        cur = np.ascontiguousarray(cur, self.dt)
        ms = self.lib.refh_me_frame(w, h, ptr(cur), w, nref, arr, R, lambda_mv, cost_type, mask, ctu0, ctu1,
                                    threads, ptr(res), ptr(best), ptr(ctu_ms) if ctu_ms is not None else None)
        return res, best, ms
