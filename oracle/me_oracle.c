/*
 * me_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of uavs3e's motion-estimation path.  It is the
 * checker for the HIP library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and never as the thing measured or
 * shipped.  It is compiled once per bit depth (-DBIT_DEPTH=8 / 10), exactly as
 * the reference fixes `pel` at compile time (inc/com_api.h:46-52).
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The code is written from the stated semantics, not transcribed:
 *   - SATD uses one generic separable Walsh-Hadamard for all six tile shapes
 *     (the reference unrolls six different butterfly networks);
 *   - the MV-bits table is filled from its closed form;
 *   - the frame sub-pel planes are computed per plane from the tap formulas.
 * How each function is pinned (golden vectors from the compiled reference,
 * invariants, brute force) is listed in tests/test_oracle_*.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifndef BIT_DEPTH
#define BIT_DEPTH 8
#endif
#if BIT_DEPTH == 8
typedef uint8_t pel;
#else
typedef uint16_t pel;
#endif
typedef int8_t s8;
typedef int16_t s16;
typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t s64;

#define EXPORT __attribute__((visibility("default")))
#define CLIP3(lo, hi, v) ((v) < (lo) ? (lo) : ((v) > (hi) ? (hi) : (v)))

EXPORT int orc_bit_depth(void) { return BIT_DEPTH; }
EXPORT int orc_pel_size(void) { return (int)sizeof(pel); }

/* ------------------------------------------------------------------------ */
/* Block costs                                                               */
/* ------------------------------------------------------------------------ */

/* com_cost.c:40-58 (DEFINE_SAD): sum over h rows and w columns of |org-pred|. */
EXPORT u32 orc_sad(int w, const pel *org, int i_org, const pel *pred, int i_pred, int h)
{
    u32 s = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int d = (int)org[y * i_org + x] - (int)pred[y * i_pred + x];
            s += (u32)(d < 0 ? -d : d);
        }
    return s;
}

/* com_cost.c:60-85 (DEFINE_SAD_X3) */
EXPORT void orc_sad_x3(int w, const pel *org, int i_org, const pel *p0, const pel *p1,
                       const pel *p2, int i_pred, u32 sad[3], int h)
{
    sad[0] = orc_sad(w, org, i_org, p0, i_pred, h);
    sad[1] = orc_sad(w, org, i_org, p1, i_pred, h);
    sad[2] = orc_sad(w, org, i_org, p2, i_pred, h);
}

/* com_cost.c:87-115 (DEFINE_SAD_X4) */
EXPORT void orc_sad_x4(int w, const pel *org, int i_org, const pel *p0, const pel *p1,
                       const pel *p2, const pel *p3, int i_pred, u32 sad[4], int h)
{
    sad[0] = orc_sad(w, org, i_org, p0, i_pred, h);
    sad[1] = orc_sad(w, org, i_org, p1, i_pred, h);
    sad[2] = orc_sad(w, org, i_org, p2, i_pred, h);
    sad[3] = orc_sad(w, org, i_org, p3, i_pred, h);
}

/* In-place Sylvester fast Walsh-Hadamard transform of n values spaced by
 * `step`.  Any signed row permutation of the Hadamard matrix gives the same
 * sum of absolute coefficients, which is all SATD uses. */
static void fwht(int *v, int n, int step)
{
    for (int half = 1; half < n; half <<= 1)
        for (int i = 0; i < n; i += half << 1)
            for (int j = i; j < i + half; j++) {
                int a = v[j * step], b = v[(j + half) * step];
                v[j * step] = a + b;
                v[(j + half) * step] = a - b;
            }
}

/* Sum of |coef| of the 2-D WHT of the w x h residual org-cur. */
static int had_abs_sum(int w, int h, const pel *org, int s_org, const pel *cur, int s_cur)
{
    int m[16 * 16];
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            m[y * w + x] = (int)org[y * s_org + x] - (int)cur[y * s_cur + x];
    for (int y = 0; y < h; y++) fwht(m + y * w, w, 1);
    for (int x = 0; x < w; x++) fwht(m + x, h, w);
    int s = 0;
    for (int i = 0; i < w * h; i++) s += m[i] < 0 ? -m[i] : m[i];
    return s;
}

/* sqrt(32.0) and sqrt(128.0), set by com_dct_coef_create (com_tables.c:276-277). */
static double tbl_sqrt(int i) { return i == 0 ? sqrt(32.0) : sqrt(128.0); }

/* cost_satd[log2w-2][log2h-2] (com_cost.c:1025-1030).  Normalisation per shape:
 *   4x4 (s+1)>>1 (com_cost.c:259); 8x8 (s+2)>>2 (:340);
 *   16x8 / 8x16 (int)(s / sqrt(128) * 2.0) (:467, :586);
 *   8x4 / 4x8  (int)(s / sqrt(32)  * 2.0) (:649, :708).
 * Returns 0xFFFFFFFF for a slot the reference leaves empty. */
EXPORT u32 orc_satd(int wl, int hl, const pel *org, int s_org, const pel *cur, int s_cur)
{
    int w = 4 << wl, h = 4 << hl;
    if (wl == 0 && hl == 0) return (u32)((had_abs_sum(4, 4, org, s_org, cur, s_cur) + 1) >> 1);
    if (wl == 1 && hl == 1) return (u32)((had_abs_sum(8, 8, org, s_org, cur, s_cur) + 2) >> 2);
    if ((wl == 2 && hl == 1) || (wl == 1 && hl == 2))
        return (u32)(int)(had_abs_sum(w, h, org, s_org, cur, s_cur) / tbl_sqrt(1) * 2.0);
    if ((wl == 1 && hl == 0) || (wl == 0 && hl == 1))
        return (u32)(int)(had_abs_sum(w, h, org, s_org, cur, s_cur) / tbl_sqrt(0) * 2.0);
    return 0xFFFFFFFFu;
}

/* com_had (com_cost.c:712-812): tile the w x h block into SATD kernels, sum the
 * per-tile (already normalised) values, then shift right by bd-8. */
EXPORT u32 orc_had(int w, int h, const pel *org, int s_org, const pel *cur, int s_cur, int bd)
{
    int tw, th, wl, hl;
    if (w == h) {
        if (w == 4) { tw = 4; th = 4; }
        else { tw = 8; th = 8; }
    } else if (w > h) {
        if (h == 4) { tw = 8; th = 4; }      /* row of 8x4 tiles */
        else { tw = 16; th = 8; }            /* grid of 16x8 tiles */
    } else {
        if (w == 4) { tw = 4; th = 8; }      /* column of 4x8 tiles */
        else { tw = 8; th = 16; }            /* grid of 8x16 tiles */
    }
    wl = tw == 4 ? 0 : tw == 8 ? 1 : 2;
    hl = th == 4 ? 0 : th == 8 ? 1 : 2;
    int sum = 0;
    for (int y = 0; y < h; y += th)
        for (int x = 0; x < w; x += tw)
            sum += (int)orc_satd(wl, hl, org + y * s_org + x, s_org, cur + y * s_cur + x, s_cur);
    return (u32)(sum >> (bd - 8));
}

/* ------------------------------------------------------------------------ */
/* Interpolation                                                             */
/* ------------------------------------------------------------------------ */

/* Quarter-pel luma taps (com_tables.c:149-154) and 1/8-pel chroma taps
 * (com_tables.c:191-200). */
static const s8 LUMA_TAPS[4][8] = {
    {0, 0, 0, 64, 0, 0, 0, 0},
    {-1, 4, -10, 57, 19, -7, 3, -1},
    {-1, 4, -11, 40, 40, -11, 4, -1},
    {-1, 3, -7, 19, 57, -10, 4, -1}};
static const s8 CHROMA_TAPS[8][4] = {
    {0, 64, 0, 0}, {-4, 62, 6, 0}, {-6, 56, 15, -1}, {-5, 47, 25, -3},
    {-4, 36, 36, -4}, {-3, 25, 47, -5}, {-1, 15, 56, -6}, {0, 6, 62, -4}};

EXPORT const s8 *orc_luma_taps(int frac) { return LUMA_TAPS[frac & 3]; }

/* The 1/16-pel luma filter of affine motion compensation
 * (com_tbl_mc_l_coeff_hp, com_tables.c:130-147; the standard's table). */
static const s8 LUMA_TAPS_HP[16][8] = {
    {0, 0, 0, 64, 0, 0, 0, 0},        {0, 1, -3, 63, 4, -2, 1, 0},      {-1, 2, -5, 62, 8, -3, 1, 0},
    {-1, 3, -8, 60, 13, -4, 1, 0},    {-1, 4, -10, 58, 17, -5, 1, 0},   {-1, 4, -11, 52, 26, -8, 3, -1},
    {-1, 3, -9, 47, 31, -10, 4, -1},  {-1, 4, -11, 45, 34, -10, 4, -1}, {-1, 4, -11, 40, 40, -11, 4, -1},
    {-1, 4, -10, 34, 45, -11, 4, -1}, {-1, 4, -10, 31, 47, -9, 3, -1},  {-1, 3, -8, 26, 52, -11, 4, -1},
    {0, 1, -5, 17, 58, -10, 4, -1},   {0, 1, -4, 13, 60, -8, 3, -1},    {0, 1, -3, 8, 62, -5, 2, -1},
    {0, 1, -2, 4, 63, -3, 1, 0}};

EXPORT const s8 *orc_luma_taps_hp(int frac) { return LUMA_TAPS_HP[frac & 15]; }
EXPORT const s8 *orc_chroma_taps(int frac) { return CHROMA_TAPS[frac & 7]; }

/* n-tap dot product centred so that tap index (n/2-1) hits src[0]. */
static int taps(const pel *src, int step, const s8 *c, int n)
{
    int s = 0;
    for (int k = 0; k < n; k++) s += (int)src[(k - (n / 2 - 1)) * step] * c[k];
    return s;
}
static int taps16(const s16 *src, int step, const s8 *c, int n)
{
    int s = 0;
    for (int k = 0; k < n; k++) s += (int)src[(k - (n / 2 - 1)) * step] * c[k];
    return s;
}

/* ipcpy (com_mc.c:649-657) */
EXPORT void orc_ipcpy(const pel *src, int i_src, pel *dst, int i_dst, int w, int h)
{
    for (int y = 0; y < h; y++) memcpy(dst + y * i_dst, src + y * i_src, sizeof(pel) * (size_t)w);
}

/* ipflt[IPFILTER_H_8/H_4/V_8/V_4] (com_mc.c:461-491 luma, 555-585 chroma):
 * one-direction n-tap, clip((sum + 32) >> 6).  type: 0=H8 1=H4 2=V8 3=V4
 * (com_define.h:251-257). */
EXPORT void orc_ipflt(int type, const pel *src, int i_src, pel *dst, int i_dst, int w, int h,
                      const s8 *coef, int max_val)
{
    int n = (type == 0 || type == 2) ? 8 : 4;
    int vert = type >= 2;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int s = taps(src + y * i_src + x, vert ? i_src : 1, coef, n);
            int v = (s + 32) >> 6;
            dst[y * i_dst + x] = (pel)CLIP3(0, max_val, v);
        }
}

/* ipflt_ext[IPFILTER_EXT_8/EXT_4] (com_mc.c:493-549 luma, 587-643 chroma):
 * horizontal pass into s16, then vertical.  The bit depth is inferred from
 * max_val == 255 (com_mc.c:503, 597): 8-bit keeps the H sum unshifted and
 * shifts the V sum by 12; otherwise H is (s+2)>>2 and V shifts by 10. */
EXPORT void orc_ipflt_ext(int type, const pel *src, int i_src, pel *dst, int i_dst, int w, int h,
                          const s8 *coef_x, const s8 *coef_y, int max_val)
{
    int n = type == 0 ? 8 : 4;
    int above = n / 2 - 1, below = n / 2;
    int shift1 = max_val == 255 ? 0 : 2;
    int shift2 = max_val == 255 ? 12 : 10;
    int add1 = (1 << shift1) >> 1, add2 = 1 << (shift2 - 1);
    int rows = h + above + below;
    s16 *tmp = (s16 *)malloc(sizeof(s16) * (size_t)rows * (size_t)w);
    for (int r = 0; r < rows; r++)
        for (int x = 0; x < w; x++) {
            int s = taps(src + (r - above) * i_src + x, 1, coef_x, n);
            tmp[r * w + x] = (s16)((s + add1) >> shift1);
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int s = taps16(tmp + (y + above) * w + x, w, coef_y, n);
            int v = (s + add2) >> shift2;
            dst[y * i_dst + x] = (pel)CLIP3(0, max_val, v);
        }
    free(tmp);
}

/* pel_avrg (com_util.c:3088-3105): sources packed at stride w. */
EXPORT void orc_pel_avrg(int w, pel *dst, int i_dst, const pel *s1, const pel *s2, int h)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            dst[y * i_dst + x] = (pel)(((int)s1[y * w + x] + (int)s2[y * w + x] + 1) >> 1);
}

/* com_img_padding for one plane (com_util.c:105-165): the band
 * [-ext, w+ext) x [-ext, h+ext) is kept and replicated outward to fill the
 * pad; left/right first, then whole rows up/down. `p` points at (0,0). */
EXPORT void orc_pad_plane(pel *p, int stride, int w, int h, int pad, int ext)
{
    int e = pad - ext;
    pel *a = p - ext * stride - ext;
    int bw = w + 2 * ext, bh = h + 2 * ext;
    for (int y = 0; y < bh; y++) {
        pel *row = a + y * stride;
        for (int j = 1; j <= e; j++) { row[-j] = row[0]; row[bw - 1 + j] = row[bw - 1]; }
    }
    for (int j = 1; j <= e; j++) {
        memcpy(a - j * stride - e, a - e, sizeof(pel) * (size_t)stride);
        memcpy(a + (bh - 1 + j) * stride - e, a + (bh - 1) * stride - e, sizeof(pel) * (size_t)stride);
    }
}

/* The 15 quarter-pel luma planes of com_if_luma_frame (com_mc.c:760-800) with
 * the C filters com_if_filter_hor_8 (668-702), com_if_filter_ver_8 (704-729)
 * and uavs3e_if_ver_luma_frame_ext (731-758).
 *   planes[dy*4+dx] points at (0,0) of each padded plane (stride, pad);
 *   planes[0] is the integer plane, which must already be edge-padded.
 * Horizontal planes are evaluated over [-8, w+8) x [-8, h+8), the others over
 * [-4, w+4) x [-4, h+4); then every sub-pel plane is re-padded from its
 * [-4, w+4) band (com_img_padding(.., 1, 4)). */
EXPORT void orc_if_luma_frame(pel *const planes[16], int stride, int w, int h, int pad, int bd)
{
    const pel *src = planes[0];
    int maxv = (1 << bd) - 1;
    int sh_tmp = bd - 8, add_tmp = (1 << sh_tmp) >> 1;
    int tw = w + 16, th = h + 16;
    /* s16 horizontal intermediates for dx = 1..3 over [-8, w+8) x [-8, h+8) */
    s16 *tmp = (s16 *)malloc(sizeof(s16) * 3 * (size_t)tw * (size_t)th);
    for (int dx = 1; dx < 4; dx++) {
        pel *dst = planes[dx];
        s16 *t = tmp + (size_t)(dx - 1) * tw * th;
        for (int y = -8; y < h + 8; y++)
            for (int x = -8; x < w + 8; x++) {
                int s = taps(src + y * stride + x, 1, LUMA_TAPS[dx], 8);
                t[(y + 8) * tw + (x + 8)] = (s16)((s + add_tmp) >> sh_tmp);
                int v = (s + 32) >> 6;
                dst[y * stride + x] = (pel)CLIP3(0, maxv, v);
            }
    }
    for (int dy = 1; dy < 4; dy++) {
        pel *dst = planes[dy * 4];
        for (int y = -4; y < h + 4; y++)
            for (int x = -4; x < w + 4; x++) {
                int v = (taps(src + y * stride + x, stride, LUMA_TAPS[dy], 8) + 32) >> 6;
                dst[y * stride + x] = (pel)CLIP3(0, maxv, v);
            }
    }
    int sh2 = 20 - bd, add2 = 1 << (sh2 - 1);
    for (int dx = 1; dx < 4; dx++) {
        const s16 *t = tmp + (size_t)(dx - 1) * tw * th;
        for (int dy = 1; dy < 4; dy++) {
            pel *dst = planes[dy * 4 + dx];
            for (int y = -4; y < h + 4; y++)
                for (int x = -4; x < w + 4; x++) {
                    int s = taps16(t + (y + 8) * tw + (x + 8), tw, LUMA_TAPS[dy], 8);
                    int v = (s + add2) >> sh2;
                    dst[y * stride + x] = (pel)CLIP3(0, maxv, v);
                }
        }
    }
    free(tmp);
    for (int i = 1; i < 16; i++) orc_pad_plane(planes[i], stride, w, h, pad, 4);
}

/* ------------------------------------------------------------------------ */
/* Rate term                                                                 */
/* ------------------------------------------------------------------------ */

/* tab_mvbits_offset = (max(W,H) << 2) + 3 (inter.c:2034). */
EXPORT int orc_mvbits_offset(int pic_w, int pic_h) { return ((pic_w > pic_h ? pic_w : pic_h) << 2) + 3; }

/* inter_search_create (inter.c:2029-2058), from its closed form:
 *   tab[0] = 1;  tab[v] = 2*(floor(log2(|v|+1)) + 1) for 0 < |v| < off;
 *   tab[+-off] = 0 (never written; the allocation is zeroed, com_util.c:52).
 * The seeded 2/3 bits for |v| = 1, 2 (inter.c:2038-2039) are overwritten by
 * the loop with 4, which the closed form gives directly.
 * `base` has 2*off+1 entries; entry off is mv 0. */
EXPORT void orc_mvbits_fill(u8 *base, int off)
{
    base[off] = 1;
    base[0] = base[2 * off] = 0;
    for (int v = 1; v < off; v++) {
        int l = 0;
        while ((2 << l) <= v + 1) l++;   /* l = floor(log2(v+1)) */
        base[off + v] = base[off - v] = (u8)(2 * (l + 1));
    }
}

/* lambda_mv = floor(65536 * sqrt(1.43631 * 2^((qp-16)/4))) (uavs3e.c:834-851,
 * lookahead.c:85-107). */
EXPORT u32 orc_lambda_mv(int qp)
{
    double lambda = 1.43631 * pow(2.0, (qp - 16.0) / 4.0);
    return (u32)floor(65536.0 * sqrt(lambda));
}

/* tbl_refi_bits (tables.c:52-70): refi_bits[n][i] = i+1 for i < n-1, and n-1
 * for the last index; n <= 1 gives 0. */
EXPORT int orc_refi_bits(int num_refp, int refi)
{
    if (num_refp <= 1) return 0;
    return refi < num_refp - 1 ? refi + 1 : num_refp - 1;
}

/* ------------------------------------------------------------------------ */
/* Search                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
    int pic_w, pic_h, bd;
    const pel *org; int i_org;      /* PU top-left in the current picture */
    const pel *const *planes;       /* 16 padded planes at (0,0); [0] integer */
    int i_ref;                      /* plane stride in pels */
    const u8 *tab_mvbits;           /* centred: tab_mvbits[0] is mv 0 */
    int tab_off;
    u32 lambda_mv;
    int is_padding;
    int mvr;                        /* AMVR index; the batched configs use 0 */
    int cost_type;                  /* 0 = SATD, 1 = SAD then SATD at winner */
    int bi;                         /* me_search_tz's bi: org is the bi-prediction target, costs halved */
} orc_ctx_t;

EXPORT int orc_ctx_size(void) { return (int)sizeof(orc_ctx_t); }

/* get_max_search_range (me.c:502-527) with com_ipel_range_rounding (me.c:48-55).
 * Writes the integer-pel MV window; returns 0 when it is empty. */
EXPORT int orc_window(const orc_ctx_t *c, const s16 mvp[2], int x, int y, int w, int h,
                      s16 min_mv[2], s16 max_mv[2])
{
    int T = (c->tab_off << c->mvr) >> 2;
    int px = mvp[0] >> 2, py = mvp[1] >> 2;
    int lo_x, lo_y, hi_x, hi_y;
    if (c->is_padding) {
        lo_x = -128 + 1 - x; lo_y = -128 + 1 - y;
        hi_x = c->pic_w - 1 - x; hi_y = c->pic_h - 1 - y;
    } else {
        lo_x = -x; lo_y = -y;
        hi_x = c->pic_w - 1 - x - w; hi_y = c->pic_h - 1 - y - h;
    }
    if (px - T + 4 > lo_x) lo_x = px - T + 4;
    if (py - T + 4 > lo_y) lo_y = py - T + 4;
    if (px + T - 4 < hi_x) hi_x = px + T - 4;
    if (py + T - 4 < hi_y) hi_y = py + T - 4;
    min_mv[0] = (s16)lo_x; min_mv[1] = (s16)lo_y;
    max_mv[0] = (s16)hi_x; max_mv[1] = (s16)hi_y;
    if (c->mvr > 2) {
        int sh = c->mvr - 2;
        s16 *v[4] = {&min_mv[0], &min_mv[1], &max_mv[0], &max_mv[1]};
        for (int i = 0; i < 4; i++)
            *v[i] = (s16)(*v[i] >= 0 ? ((*v[i] >> sh) << sh) : -(((-*v[i]) >> sh) << sh));
    }
    return !(min_mv[0] > max_mv[0] || min_mv[1] > max_mv[1]);
}

/* Rate of an integer-pel candidate: GET_MVBITS_IPEL_X/Y (util.h:43-44) with
 * the table re-centred on the MVP (me.c:113-114). */
static inline u32 ipel_bits(const orc_ctx_t *c, const s16 mvp[2], int mx, int my)
{
    const u8 *tx = c->tab_mvbits - (mvp[0] >> c->mvr);
    const u8 *ty = c->tab_mvbits - (mvp[1] >> c->mvr);
    return (u32)tx[(mx << 2) >> c->mvr] + (u32)ty[(my << 2) >> c->mvr];
}
/* Rate of a quarter-pel candidate: GET_MVBITS_X/Y (util.h:41-42). */
static inline u32 qpel_bits(const orc_ctx_t *c, const s16 mvp[2], int mx, int my)
{
    const u8 *tx = c->tab_mvbits - (mvp[0] >> c->mvr);
    const u8 *ty = c->tab_mvbits - (mvp[1] >> c->mvr);
    return (u32)tx[mx >> c->mvr] + (u32)ty[my >> c->mvr];
}

static inline const pel *ipel_ptr(const orc_ctx_t *c, int x, int y, int mx, int my)
{
    return c->planes[0] + (y + my) * c->i_ref + (x + mx);
}

/* Integer-pel cost (me.c:151-187, 250-251):
 *   (sad << (16 - (bd-8))) + (bits_x + bits_y) * lambda_mv, as u64. */
EXPORT u64 orc_ipel_cost(const orc_ctx_t *c, const s16 mvp[2], int x, int y, int w, int h, int mx, int my)
{
    u32 sad = orc_sad(w, c->org, c->i_org, ipel_ptr(c, x, y, mx, my), c->i_ref, h);
    return ((u64)sad << (16 - (c->bd - 8))) + (u64)(ipel_bits(c, mvp, mx, my) * c->lambda_mv);
}

/* Full search: the search_raster loop (me.c:150-190) with step 1 over the
 * window [centre-R, centre+R] clipped to get_max_search_range, where the
 * centre is the rounded and clipped MVP (me.c:540-541).  Scan is y-major then
 * x, and a candidate replaces the best only when strictly cheaper, so ties go
 * to the first candidate in raster order.  Returns UINT64_MAX with mv = 0
 * when the window is empty (me.c:531-533). */
EXPORT u64 orc_full_search(const orc_ctx_t *c, const s16 mvp[2], int x, int y, int w, int h, int R,
                           s16 mv_out[2])
{
    s16 lo[2], hi[2];
    mv_out[0] = mv_out[1] = 0;
    if (!orc_window(c, mvp, x, y, w, h, lo, hi)) return UINT64_MAX;
    int cx = CLIP3(lo[0], hi[0], (mvp[0] + 2) >> 2);
    int cy = CLIP3(lo[1], hi[1], (mvp[1] + 2) >> 2);
    int x0 = CLIP3(lo[0], hi[0], cx - R), x1 = CLIP3(lo[0], hi[0], cx + R);
    int y0 = CLIP3(lo[1], hi[1], cy - R), y1 = CLIP3(lo[1], hi[1], cy + R);
    u64 best = UINT64_MAX;
    for (int my = y0; my <= y1; my++)
        for (int mx = x0; mx <= x1; mx++) {
            u64 cost = orc_ipel_cost(c, mvp, x, y, w, h, mx, my);
            if (cost < best) { best = cost; mv_out[0] = (s16)mx; mv_out[1] = (s16)my; }
        }
    return best;
}

/* com_mc_blk_luma_pointer (com_mc.c:89-101): quarter-pel position -> pointer
 * into plane [qy&3][qx&3] at the integer position clipped to
 * [-MAX_CU_SIZE-4, pic+4]. */
EXPORT const pel *orc_subpel_ptr(const orc_ctx_t *c, int qx, int qy)
{
    int dx = qx & 3, dy = qy & 3;
    int ix = CLIP3(-128 - 4, c->pic_w + 4, qx >> 2);
    int iy = CLIP3(-128 - 4, c->pic_h + 4, qy >> 2);
    return c->planes[dy * 4 + dx] + iy * c->i_ref + ix;
}

static u32 subpel_dist(const orc_ctx_t *c, int w, int h, const pel *p)
{
    if (c->cost_type == 0) return orc_had(w, h, c->org, c->i_org, p, c->i_ref, c->bd);
    return orc_sad(w, c->org, c->i_org, p, c->i_ref, h);
}

/* me_sub_pel_search (me.c:366-500) for bi = 0.  `mv` holds the integer-pel
 * result in quarter-pel units on entry and the refined MV on return.
 *   centre: distortion at the integer MV read from the integer plane
 *           without clipping (me.c:400);
 *   8 half-pel points in the order of me.c:415-417;
 *   if mvr == 0, 8 quarter-pel points around the best half-pel point in the
 *   order of me.c:460-462;
 *   cost = (dist << shift) + bits * lambda, strict '<' keeps the first;
 *   cost_type 1 recomputes SATD at the winner (me.c:490-497).
 * The SATD caches (me.c:392-407, 445-455) only replay values of the same
 * positions, so they are not modelled. */
EXPORT u64 orc_subpel(const orc_ctx_t *c, const s16 mvp[2], int x, int y, int w, int h, s16 mv[2])
{
    static const int HP[8][2] = {{-2, 0}, {-2, 2}, {0, 2}, {2, 2}, {2, 0}, {2, -2}, {0, -2}, {-2, -2}};
    static const int QP[8][2] = {{-1, 0}, {0, 1}, {1, 0}, {0, -1}, {-1, 1}, {1, 1}, {-1, -1}, {1, -1}};
    int shift = 16 - c->bi - (c->cost_type == 0 ? 0 : (c->bd - 8));
    int cx = mv[0], cy = mv[1];
    u32 d = subpel_dist(c, w, h, ipel_ptr(c, x, y, cx >> 2, cy >> 2));
    u64 best = (u64)(c->lambda_mv * qpel_bits(c, mvp, cx, cy)) + ((u64)d << shift);
    for (int i = 0; i < 8; i++) {
        int mx = cx + HP[i][0], my = cy + HP[i][1];
        d = subpel_dist(c, w, h, orc_subpel_ptr(c, mx + (x << 2), my + (y << 2)));
        u64 cost = (u64)(c->lambda_mv * qpel_bits(c, mvp, mx, my)) + ((u64)d << shift);
        if (cost < best) { best = cost; mv[0] = (s16)mx; mv[1] = (s16)my; }
    }
    if (c->mvr == 0) {
        cx = mv[0]; cy = mv[1];
        for (int i = 0; i < 8; i++) {
            int mx = cx + QP[i][0], my = cy + QP[i][1];
            d = subpel_dist(c, w, h, orc_subpel_ptr(c, mx + (x << 2), my + (y << 2)));
            u64 cost = (u64)(c->lambda_mv * qpel_bits(c, mvp, mx, my)) + ((u64)d << shift);
            if (cost < best) { best = cost; mv[0] = (s16)mx; mv[1] = (s16)my; }
        }
    }
    if (c->cost_type) {
        d = orc_had(w, h, c->org, c->i_org, orc_subpel_ptr(c, mv[0] + (x << 2), mv[1] + (y << 2)), c->i_ref, c->bd);
        best = (u64)(c->lambda_mv * qpel_bits(c, mvp, mv[0], mv[1])) + ((u64)d << (16 - c->bi));
    }
    return best;
}

/* Final uni-prediction cost of me_search_tz (me.c:560-567):
 *   c + lambda * (refi_bits[num_refp][refi] + mvr + (mvr < 4) + 2),
 *   returned as (u32)((c + 2^15) >> 16). */
EXPORT u32 orc_final_cost(u64 cost, u32 lambda_mv, int num_refp, int refi, int mvr)
{
    int extra = orc_refi_bits(num_refp, refi) + mvr + (mvr < 4) + 2;
    cost += (u64)(lambda_mv * (u32)extra);
    return (u32)((cost + (1u << 15)) >> 16);
}

/* ------------------------------------------------------------------------ */
/* The reference ME search (diamond -> raster -> refine -> sub-pel)          */
/* ------------------------------------------------------------------------ */

typedef struct {
    int gop_size;
    int max_search_range;
    int adaptive_raster_range;
    long long ptr, ref_ptr;     /* POC of the current and the reference picture */
    int num_refp;
    s16 mv_ipel[2];             /* previous integer result for this list/ref (adaptive range) */
    int bi_bits;                /* bi: mot_bits of the other list + 1 (me.c:561) */
} orc_tz_t;

EXPORT int orc_tz_size(void) { return (int)sizeof(orc_tz_t); }

/* get_raster_range (me.c:65-100) */
static void raster_range(const orc_ctx_t *c, const orc_tz_t *t, const s16 lo[2], const s16 hi[2],
                         const s16 centre[2], int step, int out[2][2])
{
    int mvr = c->mvr;
    long long dl = t->ptr - t->ref_ptr;
    int dist = (int)(dl < 0 ? -dl : dl);
    int msr = t->max_search_range >> (3 - mvr > 0 ? 3 - mvr : 0);
    int rxy = (msr * dist + (t->gop_size >> 1)) / t->gop_size;
    rxy = CLIP3(msr >> 2, msr, rxy);
    int rx = rxy, ry = rxy;
    if (mvr && t->adaptive_raster_range) {
        rx = abs(t->mv_ipel[0] - centre[0]) * 4;
        ry = abs(t->mv_ipel[1] - centre[1]) * 2;
        if (rx < (msr >> 2)) rx = msr >> 2;
        if (ry < (msr >> 2)) ry = msr >> 2;
        if (rx < step * 8) rx = step * 8;
        if (ry < step * 4) ry = step * 4;
        if (rx > rxy) rx = rxy;
        if (ry > rxy) ry = rxy;
    }
    if (mvr > 2) { int sh = mvr - 2; rx = (rx >> sh) << sh; ry = (ry >> sh) << sh; }
    out[0][0] = CLIP3(lo[0], hi[0], centre[0] - (s16)rx);
    out[1][0] = CLIP3(lo[0], hi[0], centre[0] + (s16)rx);
    out[0][1] = CLIP3(lo[1], hi[1], centre[1] - (s16)ry);
    out[1][1] = CLIP3(lo[1], hi[1], centre[1] + (s16)ry);
}

/* A candidate of a running search: the reference compares
 * tmp_best = best - cost_mvy (as s64) against sad_cost + cost_mvx, which is
 * best > cost exactly while best is finite (me.c:158, 169-172, 189). */
static int try_ipel(const orc_ctx_t *c, const s16 mvp[2], int x, int y, int w, int h, int shift,
                    int mx, int my, u64 *best, s16 mv[2])
{
    u32 sad = orc_sad(w, c->org, c->i_org, ipel_ptr(c, x, y, mx, my), c->i_ref, h);
    u64 cost = ((u64)sad << shift) + (u64)(ipel_bits(c, mvp, mx, my) * c->lambda_mv);
    if (cost < *best) { *best = cost; mv[0] = (s16)mx; mv[1] = (s16)my; return 1; }
    return 0;
}

/* search_raster (me.c:103-197).  Note: the row loop's s64 'tmp_best' makes a
 * row unable to improve while *best is UINT64_MAX (UINT64_MAX - cost_mvy is
 * negative as s64); the reference only calls it with a finite best. */
static void search_raster(const orc_ctx_t *c, const orc_tz_t *t, const s16 lo[2], const s16 hi[2],
                          int x, int y, int w, int h, int refi, const s16 mvp[2], s16 mv[2], u64 *best)
{
    static const int step_base[5] = {5, 5, 5, 6, 8};
    static const int step_min[5] = {1, 1, 1, 2, 4};
    int shift = 16 - (c->bd - 8);
    int mn = w < h ? w : h;
    int step = (step_base[c->mvr] > (mn >> 1) ? step_base[c->mvr] : (mn >> 1)) * (refi + 1);
    int rng[2][2];
    raster_range(c, t, lo, hi, mv, step, rng);
    for (int round = 0; step >= step_min[c->mvr]; round++) {
        int x0, y0, x1, y1;
        if (round == 0) { x0 = rng[0][0]; y0 = rng[0][1]; x1 = rng[1][0]; y1 = rng[1][1]; }
        else {
            x0 = mv[0] - step > lo[0] ? mv[0] - step : lo[0];
            y0 = mv[1] - step > lo[1] ? mv[1] - step : lo[1];
            x1 = mv[0] + step < hi[0] ? mv[0] + step : hi[0];
            y1 = mv[1] + step < hi[1] ? mv[1] + step : hi[1];
        }
        for (int my = y0; my <= y1; my += step) {
            if (*best == UINT64_MAX) continue;
            for (int mx = x0; mx <= x1; mx += step) try_ipel(c, mvp, x, y, w, h, shift, mx, my, best, mv);
        }
        step >>= 1;
        if (c->mvr > 2) { int sh = c->mvr - 2; step = (step >> sh) << sh; }
    }
}

/* search_diamond (me.c:199-355) for bi = 0. */
static int search_diamond(const orc_ctx_t *c, const s16 lo[2], const s16 hi[2], int x, int y, int w, int h,
                          const s16 mvp[2], s16 mv[2], u64 *best, int max_step)
{
    static const int DIA[2][16][2] = {
        {{-2, 0}, {-1, 1}, {0, 2}, {1, 1}, {2, 0}, {1, -1}, {0, -2}, {-1, -1}},
        {{-4, 0}, {-3, 1}, {-2, 2}, {-1, 3}, {0, 4}, {1, 3}, {2, 2}, {3, 1},
         {4, 0}, {3, -1}, {2, -2}, {1, -3}, {0, -4}, {-1, -3}, {-2, -2}, {-3, -1}}};
    int best_step = 0;
    int shift = 16 - c->bi - (c->bd - 8);
    int core = c->bi ? 5 : 2, cstep = 1;
    int cx = mv[0], cy = mv[1];
    if (c->bi) {   /* the core square spans the distance to the rounded MVP, 2..5 (6 above integer AMVR) */
        if (c->mvr > 2) core++;
        int dx = abs(((mvp[0] + 2) >> 2) - cx), dy = abs(((mvp[1] + 2) >> 2) - cy);
        int d = (dx > dy ? dx : dy) >> (c->mvr > 2 ? c->mvr - 2 : 0);
        core = CLIP3(2, core, d);
    }
    if (c->mvr > 2) { core <<= c->mvr - 2; cstep <<= c->mvr - 2; }
    int x0 = cx - core, y0 = cy - core, x1 = cx + core, y1 = cy + core;
    while (x0 <= lo[0]) x0 += cstep;
    while (y0 <= lo[1]) y0 += cstep;
    while (x1 >= hi[0]) x1 -= cstep;
    while (y1 >= hi[1]) y1 -= cstep;
    if (try_ipel(c, mvp, x, y, w, h, shift, cx, cy, best, mv)) best_step = 0;
    if (max_step) {
        for (int my = y0; my <= y1; my += cstep)
            for (int mx = x0; mx <= x1; mx += cstep) try_ipel(c, mvp, x, y, w, h, shift, mx, my, best, mv);
    }
    if (mv[0] != cx || mv[1] != cy) best_step = 2;
    if (c->bi) return best_step;   /* no multilayer diamond (me.c:317-319) */
    int scale = c->mvr - 2 > 0 ? c->mvr - 2 : 0;
    for (int step = 4, round = 0; step < max_step; step <<= 1, round++) {
        int ss = (step << scale) >> 1;
        int num = round < 2 ? 8 : 16, add = round < 1 ? 2 : 1, idx = round < 2 ? 0 : 1;
        for (int i = 0; i < num; i += add) {
            int mx = cx + (s16)(ss >> idx) * DIA[idx][i][0];
            int my = cy + (s16)(ss >> idx) * DIA[idx][i][1];
            if (mx <= hi[0] && mx >= lo[0] && my <= hi[1] && my >= lo[1])
                if (try_ipel(c, mvp, x, y, w, h, shift, mx, my, best, mv)) best_step = step;
        }
    }
    return best_step;
}

/* me_search_tz (me.c:529-568).  Returns the u64 that the reference
 * returns: UINT64_MAX for an empty window (mv_out untouched), otherwise the
 * rounded u32 cost.  mv_out is in quarter-pel; with bi = 0 the search starts
 * at the MVP and mv_ipel_out receives the integer-pel result (me.c:553);
 * with bi = 1 it starts at mv_out's value on entry, runs the core square
 * only and leaves mv_ipel_out alone; *cost64 is the unrounded cost. */
EXPORT u64 orc_me_search_tz(const orc_ctx_t *c, orc_tz_t *t, int x, int y, int w, int h, int refi,
                            const s16 mvp[2], s16 mv_out[2], s16 mv_ipel_out[2], u64 *cost64)
{
    s16 lo[2], hi[2];
    if (!orc_window(c, mvp, x, y, w, h, lo, hi)) { *cost64 = UINT64_MAX; return UINT64_MAX; }
    s16 mv[2];
    const s16 *start = c->bi ? mv_out : mvp;
    mv[0] = (s16)CLIP3(lo[0], hi[0], (start[0] + 2) >> 2);
    mv[1] = (s16)CLIP3(lo[1], hi[1], (start[1] + 2) >> 2);
    u64 best = UINT64_MAX;
    int bs = search_diamond(c, lo, hi, x, y, w, h, mvp, mv, &best, 32);
    if (!c->bi && (abs(mvp[0] - (mv[0] << 2)) > 2 || abs(mvp[1] - (mv[1] << 2)) > 2)) {
        if (bs > 5) search_raster(c, t, lo, hi, x, y, w, h, refi, mvp, mv, &best);
        search_diamond(c, lo, hi, x, y, w, h, mvp, mv, &best, 8);
    }
    if (!c->bi) {
        t->mv_ipel[0] = mv[0]; t->mv_ipel[1] = mv[1];
        mv_ipel_out[0] = mv[0]; mv_ipel_out[1] = mv[1];
    }
    mv[0] = (s16)(mv[0] * 4); mv[1] = (s16)(mv[1] * 4);
    if (c->mvr < 2) best = orc_subpel(c, mvp, x, y, w, h, mv);
    mv_out[0] = mv[0]; mv_out[1] = mv[1];
    int extra = orc_refi_bits(t->num_refp, refi) + c->mvr + (c->mvr < 4) + (c->bi ? t->bi_bits : 2);
    best += (u64)(c->lambda_mv * (u32)extra);
    *cost64 = best;
    return (u32)((best + (1u << 15)) >> 16);
}

/* ------------------------------------------------------------------------ */
/* Frame driver for the batched workloads                                    */
/* ------------------------------------------------------------------------ */

/* The 29 AVS3 PU shapes (w x h) allowed by com_check_split_mode
 * (com_util.c:1338-1420) with the encoder defaults max_part_ratio 8 and
 * min_cu 4 (uavs3e.c:1590-1596), 4x4 excluded, in a fixed order. */
static const int SHAPES[29][2] = {
    {4, 8}, {8, 4}, {4, 16}, {16, 4}, {4, 32}, {32, 4}, {8, 8}, {8, 16}, {16, 8}, {8, 32},
    {32, 8}, {8, 64}, {64, 8}, {16, 16}, {16, 32}, {32, 16}, {16, 64}, {64, 16}, {16, 128},
    {128, 16}, {32, 32}, {32, 64}, {64, 32}, {32, 128}, {128, 32}, {64, 64}, {64, 128},
    {128, 64}, {128, 128}};

EXPORT int orc_shape_w(int s) { return SHAPES[s][0]; }
EXPORT int orc_shape_h(int s) { return SHAPES[s][1]; }

/* Number of PUs per 128x128 CTU for a shape mask; PUs are enumerated shape
 * by shape (table order), each shape tiled in raster order over the CTU. */
EXPORT int orc_pus_per_ctu(u32 mask)
{
    int n = 0;
    for (int s = 0; s < 29; s++)
        if (mask >> s & 1) n += (128 / SHAPES[s][0]) * (128 / SHAPES[s][1]);
    return n;
}

typedef struct {
    s16 mv_int[2];       /* integer-pel full-search result */
    s16 mv_sub[2];       /* quarter-pel refined result */
    u64 cost_int;        /* u64 integer-pel cost, UINT64_MAX if not searched */
    u64 cost_sub;        /* u64 sub-pel cost */
    u32 cost_final;      /* rounded uni cost including ref/MVR bits */
    u32 valid;           /* 1 if the PU lies inside the picture and was searched */
} orc_pu_res_t;

EXPORT int orc_pu_res_size(void) { return (int)sizeof(orc_pu_res_t); }

/* Full search + sub-pel refinement for every PU of CTUs [ctu0, ctu1) and every
 * reference.  MVP = (0,0) for every PU; PUs not fully inside the picture are
 * skipped (valid = 0).  Results are indexed [ctu][pu][ref].
 * References [0, num_refs - num_refs_l1) form list 0, the rest list 1; the
 * final cost of reference r uses tbl_refi_bits of its list's size and its
 * index in the list (me.c:560-561 with num_refp of that list).
 * best_ref[(ctu*npu + pu)*nl + l] (nl = 1 without list 1, else 2) is the
 * reference of list l with the smallest final u32 cost, ties to the lower
 * index (inter.c:901-904), or -1 for a skipped PU.
 *   ref_planes: num_refs x 16 plane pointers, each at (0,0) of a padded plane. */
EXPORT void orc_me_frame_lists(int pic_w, int pic_h, int bd, const pel *cur, int i_cur,
                               int num_refs, int num_refs_l1, const pel *const *ref_planes, int i_ref,
                               int R, u32 lambda_mv, int cost_type, u32 shape_mask,
                               int ctu0, int ctu1, orc_pu_res_t *res, s8 *best_ref)
{
    int off = orc_mvbits_offset(pic_w, pic_h);
    u8 *tab = (u8 *)malloc((size_t)(2 * off + 1));
    orc_mvbits_fill(tab, off);
    int ctu_cols = (pic_w + 127) / 128;
    int npu = orc_pus_per_ctu(shape_mask);
    const int n0 = num_refs - num_refs_l1, nl = num_refs_l1 > 0 ? 2 : 1;
    s16 mvp[2] = {0, 0};
    for (int ctu = ctu0; ctu < ctu1; ctu++) {
        int cx = (ctu % ctu_cols) * 128, cy = (ctu / ctu_cols) * 128;
        int pu = 0;
        for (int s = 0; s < 29; s++) {
            if (!(shape_mask >> s & 1)) continue;
            int w = SHAPES[s][0], h = SHAPES[s][1];
            for (int oy = 0; oy < 128; oy += h)
                for (int ox = 0; ox < 128; ox += w, pu++) {
                    int x = cx + ox, y = cy + oy;
                    size_t base = ((size_t)ctu * npu + pu);
                    int inside = x + w <= pic_w && y + h <= pic_h;
                    u32 best_cost[2] = {0xFFFFFFFFu, 0xFFFFFFFFu};
                    s8 bref[2] = {-1, -1};
                    for (int r = 0; r < num_refs; r++) {
                        orc_pu_res_t *o = &res[base * num_refs + r];
                        memset(o, 0, sizeof(*o));
                        o->cost_int = o->cost_sub = UINT64_MAX;
                        o->cost_final = 0xFFFFFFFFu;
                        if (!inside) continue;
                        orc_ctx_t c;
                        c.pic_w = pic_w; c.pic_h = pic_h; c.bd = bd;
                        c.org = cur + y * i_cur + x; c.i_org = i_cur;
                        c.planes = ref_planes + 16 * r; c.i_ref = i_ref;
                        c.tab_mvbits = tab + off; c.tab_off = off;
                        c.lambda_mv = lambda_mv; c.is_padding = 1; c.mvr = 0; c.cost_type = cost_type;
                        c.bi = 0;
                        s16 mv[2];
                        u64 ci = orc_full_search(&c, mvp, x, y, w, h, R, mv);
                        o->mv_int[0] = mv[0]; o->mv_int[1] = mv[1];
                        o->cost_int = ci;
                        if (ci == UINT64_MAX) continue;
                        mv[0] = (s16)(mv[0] * 4); mv[1] = (s16)(mv[1] * 4);
                        u64 cs = orc_subpel(&c, mvp, x, y, w, h, mv);
                        o->mv_sub[0] = mv[0]; o->mv_sub[1] = mv[1];
                        o->cost_sub = cs;
                        const int l = r < n0 ? 0 : 1;
                        o->cost_final = l == 0 ? orc_final_cost(cs, lambda_mv, n0, r, 0)
                                               : orc_final_cost(cs, lambda_mv, num_refs_l1, r - n0, 0);
                        o->valid = 1;
                        if (o->cost_final < best_cost[l]) { best_cost[l] = o->cost_final; bref[l] = (s8)r; }
                    }
                    for (int l = 0; l < nl; l++) best_ref[base * nl + l] = bref[l];
                }
        }
    }
    free(tab);
}

/* The single-list form: every reference in list 0. */
EXPORT void orc_me_frame(int pic_w, int pic_h, int bd, const pel *cur, int i_cur,
                         int num_refs, const pel *const *ref_planes, int i_ref,
                         int R, u32 lambda_mv, int cost_type, u32 shape_mask,
                         int ctu0, int ctu1, orc_pu_res_t *res, s8 *best_ref)
{
    orc_me_frame_lists(pic_w, pic_h, bd, cur, i_cur, num_refs, 0, ref_planes, i_ref, R, lambda_mv, cost_type,
                       shape_mask, ctu0, ctu1, res, best_ref);
}

/* ------------------------------------------------------------------------ */
/* Motion compensation                                                       */
/* ------------------------------------------------------------------------ */

/* com_mc_blk_luma with hp_flag 0 (com_mc.c:41-87): quarter-pel position ->
 * copy of w x h samples from plane [qy&3][qx&3] at the integer position
 * clipped to [-MAX_CU_SIZE-4, pic+4] (max_posx = pic_w + 4, com_mc.c:144). */
static void mc_luma(const pel *const *planes, int i_ref, int pic_w, int pic_h, int qx, int qy, int w, int h,
                    pel *dst, int i_dst)
{
    int ix = CLIP3(-128 - 4, pic_w + 4, qx >> 2);
    int iy = CLIP3(-128 - 4, pic_h + 4, qy >> 2);
    orc_ipcpy(planes[(qy & 3) * 4 + (qx & 3)] + iy * i_ref + ix, i_ref, dst, i_dst, w, h);
}

/* com_mc_blk_chroma with hp_flag 0 (com_mc.c:103-138): the luma quarter-pel
 * position is a 1/8-pel chroma position; the integer part is clipped to
 * [-(MAX_CU_SIZE/2)-2, (pic+4)/2] and the 4-tap filters of com_tbl_mc_c_coeff
 * run on the padded chroma plane (copy, H_4, V_4 or EXT_4). */
static void mc_chroma(const pel *plane, int i_c, int pic_w, int pic_h, int qx, int qy, int w, int h,
                      pel *dst, int i_dst, int maxv)
{
    int dx = qx & 7, dy = qy & 7;
    int ix = CLIP3(-64 - 2, (pic_w + 4) >> 1, qx >> 3);
    int iy = CLIP3(-64 - 2, (pic_h + 4) >> 1, qy >> 3);
    const pel *src = plane + iy * i_c + ix;
    if (dx == 0 && dy == 0) orc_ipcpy(src, i_c, dst, i_dst, w, h);
    else if (dy == 0) orc_ipflt(1, src, i_c, dst, i_dst, w, h, CHROMA_TAPS[dx], maxv);
    else if (dx == 0) orc_ipflt(3, src, i_c, dst, i_dst, w, h, CHROMA_TAPS[dy], maxv);
    else orc_ipflt_ext(1, src, i_c, dst, i_dst, w, h, CHROMA_TAPS[dx], CHROMA_TAPS[dy], maxv);
}

/* com_mc_cu (com_mc.c:140-200) for one PU at (x, y), w x h, from nl lists:
 * list l predicts from luma planes planes_l[l] (16 plane pointers at (0,0),
 * stride i_ref) and chroma planes u_l[l], v_l[l] (padded, at (0,0), stride
 * i_c) at quarter-pel MV mv[l]; two lists are averaged with pel_avrg,
 * (a + b + 1) >> 1 (com_util.c:3088-3105).  When both lists name the same
 * picture with the same MV the reference keeps the L0 prediction, which the
 * average reproduces ((a + a + 1) >> 1 = a).  Chroma (4:2:0, (w/2) x (h/2))
 * only when do_chroma.  Outputs are packed (stride w, resp. w/2). */
EXPORT void orc_mc_pu(int pic_w, int pic_h, int bd, int nl, const pel *const *const *planes_l, int i_ref,
                      const pel *const *u_l, const pel *const *v_l, int i_c, int x, int y, int w, int h,
                      const s16 *mv, int do_chroma, pel *pred_y, pel *pred_u, pel *pred_v)
{
    const int maxv = (1 << bd) - 1;
    pel *t0 = (pel *)malloc(sizeof(pel) * 128 * 128), *t1 = (pel *)malloc(sizeof(pel) * 128 * 128);
    for (int comp = 0; comp < (do_chroma ? 3 : 1); comp++) {
        int cw = comp ? w / 2 : w, ch = comp ? h / 2 : h;
        pel *out = comp == 0 ? pred_y : (comp == 1 ? pred_u : pred_v);
        for (int l = 0; l < nl; l++) {
            int qx = (x << 2) + mv[2 * l], qy = (y << 2) + mv[2 * l + 1];
            pel *dst = l == 0 ? t0 : t1;
            if (comp == 0) mc_luma(planes_l[l], i_ref, pic_w, pic_h, qx, qy, cw, ch, dst, cw);
            else mc_chroma(comp == 1 ? u_l[l] : v_l[l], i_c, pic_w, pic_h, qx, qy, cw, ch, dst, cw, maxv);
        }
        if (nl == 2) orc_pel_avrg(cw, out, cw, t0, t1, ch);
        else memcpy(out, t0, sizeof(pel) * (size_t)cw * ch);
    }
    free(t0);
    free(t1);
}

/* Prediction of every PU of CTUs [ctu0, ctu1) from the ME results: orc_mc_pu
 * with the best reference of each list (best_ref, as orc_me_frame_lists
 * writes it) and its quarter-pel MV.  Chroma is predicted for PUs of width
 * and height >= 8 (the chroma slots widx - 1 of com_mc_cu exist only there).
 * Output layout, per CTU: the selected shapes in table order, each shape's
 * PUs in raster order, each PU's w x h samples row by row; luma takes
 * 16384 samples per shape and chroma (u and v each) 4096 per eligible
 * shape.  Skipped PUs (a best_ref < 0) are written as zeros.
 *   ref_planes: num_refs x 16 luma planes at (0,0), stride i_ref
 *   ref_u, ref_v: num_refs chroma planes at (0,0), padded, stride i_c */
EXPORT void orc_mc_frame(int pic_w, int pic_h, int bd, int num_refs, int num_refs_l1, const pel *const *ref_planes,
                         int i_ref, const pel *const *ref_u, const pel *const *ref_v, int i_c, u32 shape_mask,
                         int ctu0, int ctu1, const orc_pu_res_t *res, const s8 *best_ref,
                         pel *pred_y, pel *pred_u, pel *pred_v)
{
    const int nl = num_refs_l1 > 0 ? 2 : 1;
    int ctu_cols = (pic_w + 127) / 128;
    int npu = orc_pus_per_ctu(shape_mask);
    int nsel = 0, nsel_c = 0;
    for (int s = 0; s < 29; s++)
        if (shape_mask >> s & 1) { nsel++; if (SHAPES[s][0] >= 8 && SHAPES[s][1] >= 8) nsel_c++; }
    for (int ctu = ctu0; ctu < ctu1; ctu++) {
        int cx = (ctu % ctu_cols) * 128, cy = (ctu / ctu_cols) * 128;
        int pu = 0, rank = 0, rank_c = 0;
        for (int s = 0; s < 29; s++) {
            if (!(shape_mask >> s & 1)) continue;
            int w = SHAPES[s][0], h = SHAPES[s][1];
            int has_c = w >= 8 && h >= 8;
            pel *py = pred_y + (size_t)ctu * nsel * 16384 + (size_t)rank * 16384;
            pel *pu_ = pred_u + (size_t)ctu * nsel_c * 4096 + (size_t)rank_c * 4096;
            pel *pv_ = pred_v + (size_t)ctu * nsel_c * 4096 + (size_t)rank_c * 4096;
            int i = 0;
            for (int oy = 0; oy < 128; oy += h)
                for (int ox = 0; ox < 128; ox += w, pu++, i++) {
                    size_t base = ((size_t)ctu * npu + pu);
                    pel *oy_ = py + (size_t)i * w * h;
                    pel *ou = has_c ? pu_ + (size_t)i * (w / 2) * (h / 2) : NULL;
                    pel *ov = has_c ? pv_ + (size_t)i * (w / 2) * (h / 2) : NULL;
                    int ok = 1;
                    for (int l = 0; l < nl; l++) ok &= best_ref[base * nl + l] >= 0;
                    if (!ok) {
                        memset(oy_, 0, sizeof(pel) * (size_t)w * h);
                        if (has_c) {
                            memset(ou, 0, sizeof(pel) * (size_t)(w / 2) * (h / 2));
                            memset(ov, 0, sizeof(pel) * (size_t)(w / 2) * (h / 2));
                        }
                        continue;
                    }
                    const pel *const *pl[2];
                    const pel *ul[2], *vl[2];
                    s16 mv[4];
                    for (int l = 0; l < nl; l++) {
                        int r = best_ref[base * nl + l];
                        pl[l] = ref_planes + 16 * r;
                        ul[l] = ref_u ? ref_u[r] : NULL;
                        vl[l] = ref_v ? ref_v[r] : NULL;
                        mv[2 * l] = res[base * num_refs + r].mv_sub[0];
                        mv[2 * l + 1] = res[base * num_refs + r].mv_sub[1];
                    }
                    orc_mc_pu(pic_w, pic_h, bd, nl, pl, i_ref, ul, vl, i_c, cx + ox, cy + oy, w, h, mv,
                              has_c && ref_u && ref_v, oy_, ou, ov);
                }
            rank++;
            rank_c += has_c;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Affine motion compensation, luma                                          */
/* ------------------------------------------------------------------------ */

static int log2i(int v) { int l = 0; while ((1 << l) < v) l++; return l; }

/* Rounds a value with 7 fraction bits to an integer, half away from zero
 * (com_mv_rounding_s32 with right shift 7, left shift 0, com_util.h:189-194). */
static int round7(int v) { return v >= 0 ? (v + 64) >> 7 : -((-v + 64) >> 7); }

/* com_mc_blk_affine_luma (com_mc.c:202-250) and the luma part of
 * com_mc_blk_affine (com_mc.c:252-333): 4-parameter model from two
 * control-point MVs, 6-parameter from three (cp_num; cpmv in 1/16 pel); each
 * sub_w x sub_h sub-block takes the MV at its centre (the top-left sub-block
 * the MV at the CU's top-left corner, the top-right one at its top-right
 * corner, with three points the bottom-left one at its bottom-left corner),
 * rounded to 1/16 pel
 * and clipped to 18 bits, and is predicted by com_mc_blk_luma with hp_flag:
 * position clipped to [-132, pic + 4] (com_mc.c:61-62), then an integer
 * copy, a one-direction 8-tap filter or the two-pass filter with the 1/16
 * table.  (com_mc.c:65's test for the half-pel planes can never hold, so
 * the filters always run.)  planes0: the padded integer plane at (0,0). */
EXPORT void orc_affine_luma(int x, int y, int pic_w, int pic_h, int cu_w, int cu_h, const int cpmv[3][2],
                            int cp_num, const pel *plane0, int i_ref, pel *pred, int sub_w, int sub_h, int bd)
{
    const int lw = log2i(cu_w), lh = log2i(cu_h);
    const int base_x = cpmv[0][0] * 128, base_y = cpmv[0][1] * 128;
    const int dhx = ((cpmv[1][0] - cpmv[0][0]) * 128) >> lw, dhy = ((cpmv[1][1] - cpmv[0][1]) * 128) >> lw;
    /* the 6-parameter model takes the vertical gradient from the bottom-left
     * control point (com_mc.c:273-279) */
    const int dvx = cp_num == 3 ? ((cpmv[2][0] - cpmv[0][0]) * 128) >> lh : -dhy;
    const int dvy = cp_num == 3 ? ((cpmv[2][1] - cpmv[0][1]) * 128) >> lh : dhx;
    const int max_val = (1 << bd) - 1;
    for (int h = 0; h < cu_h; h += sub_h)
        for (int w = 0; w < cu_w; w += sub_w) {
            int px = w + sub_w / 2, py = h + sub_h / 2;
            if (w == 0 && h == 0) { px = 0; py = 0; }
            else if (w + sub_w == cu_w && h == 0) { px = cu_w; py = 0; }
            else if (w == 0 && h + sub_h == cu_h && cp_num == 3) { px = 0; py = cu_h; }
            int mx = round7(base_x + dhx * px + dvx * py), my = round7(base_y + dhy * px + dvy * py);
            mx = CLIP3(-131072, 131071, mx);
            my = CLIP3(-131072, 131071, my);
            const int gx = ((x + w) << 4) + mx, gy = ((y + h) << 4) + my;
            const int fx = gx & 15, fy = gy & 15;
            const int ix = CLIP3(-128 - 4, pic_w + 4, gx >> 4), iy = CLIP3(-128 - 4, pic_h + 4, gy >> 4);
            const pel *src = plane0 + iy * i_ref + ix;
            pel *dst = pred + h * cu_w + w;
            if (fx == 0 && fy == 0) orc_ipcpy(src, i_ref, dst, cu_w, sub_w, sub_h);
            else if (fy == 0) orc_ipflt(0, src, i_ref, dst, cu_w, sub_w, sub_h, LUMA_TAPS_HP[fx], max_val);
            else if (fx == 0) orc_ipflt(2, src, i_ref, dst, cu_w, sub_w, sub_h, LUMA_TAPS_HP[fy], max_val);
            else orc_ipflt_ext(0, src, i_ref, dst, cu_w, sub_w, sub_h, LUMA_TAPS_HP[fx], LUMA_TAPS_HP[fy], max_val);
        }
}

/* The 1/32-pel chroma filter of affine MC (com_tbl_mc_c_coeff_hp,
 * com_tables.c:156-189; the standard's table). */
static const s8 CHROMA_TAPS_HP[32][4] = {
    {0, 64, 0, 0},    {-1, 63, 2, 0},   {-2, 62, 4, 0},   {-2, 60, 7, -1},  {-2, 58, 10, -2}, {-3, 57, 12, -2},
    {-4, 56, 14, -2}, {-4, 55, 15, -2}, {-4, 54, 16, -2}, {-5, 53, 18, -2}, {-6, 52, 20, -2}, {-6, 49, 24, -3},
    {-6, 46, 28, -4}, {-5, 44, 29, -4}, {-4, 42, 30, -4}, {-4, 39, 33, -4}, {-4, 36, 36, -4}, {-4, 33, 39, -4},
    {-4, 30, 42, -4}, {-4, 29, 44, -5}, {-4, 28, 46, -6}, {-3, 24, 49, -6}, {-2, 20, 52, -6}, {-2, 18, 53, -5},
    {-2, 16, 54, -4}, {-2, 15, 55, -4}, {-2, 14, 56, -4}, {-2, 12, 57, -3}, {-2, 10, 58, -2}, {-1, 7, 60, -2},
    {0, 4, 62, -2},   {0, 2, 63, -1}};

/* The chroma of com_mc_blk_affine (com_mc.c:290-360) for square sub-blocks
 * (sub = 4 or 8): the luma sub-block MVs as above, kept per 4x4 luma unit
 * (mv_save); each 8x8 luma block's 4x4 chroma uses (sum of its four units'
 * MVs + 2) >> 2 as a 1/32-pel chroma position, clipped to [-66, (pic+4)/2],
 * and com_mc_blk_chroma's copy / 4-tap / two-pass filters with the 1/32
 * table.  u0 / v0: padded chroma planes at (0,0), stride i_c. */
EXPORT void orc_affine_chroma(int x, int y, int pic_w, int pic_h, int cu_w, int cu_h, const int cpmv[3][2],
                              int cp_num, int sub, const pel *u0, const pel *v0, int i_c, pel *pred_u, pel *pred_v,
                              int bd)
{
    const int lw = log2i(cu_w), lh = log2i(cu_h);
    const int dhx = ((cpmv[1][0] - cpmv[0][0]) * 128) >> lw, dhy = ((cpmv[1][1] - cpmv[0][1]) * 128) >> lw;
    const int dvx = cp_num == 3 ? ((cpmv[2][0] - cpmv[0][0]) * 128) >> lh : -dhy;
    const int dvy = cp_num == 3 ? ((cpmv[2][1] - cpmv[0][1]) * 128) >> lh : dhx;
    static int save[32][32][2];
    for (int h = 0; h < cu_h; h += sub)
        for (int w = 0; w < cu_w; w += sub) {
            int px = w + sub / 2, py = h + sub / 2;
            if (w == 0 && h == 0) { px = 0; py = 0; }
            else if (w + sub == cu_w && h == 0) { px = cu_w; py = 0; }
            else if (w == 0 && h + sub == cu_h && cp_num == 3) { px = 0; py = cu_h; }
            int mx = CLIP3(-131072, 131071, round7(cpmv[0][0] * 128 + dhx * px + dvx * py));
            int my = CLIP3(-131072, 131071, round7(cpmv[0][1] * 128 + dhy * px + dvy * py));
            for (int j = 0; j < sub / 4; j++)
                for (int i = 0; i < sub / 4; i++) {
                    save[w / 4 + i][h / 4 + j][0] = mx;
                    save[w / 4 + i][h / 4 + j][1] = my;
                }
        }
    const int max_val = (1 << bd) - 1;
    for (int h = 0; h < cu_h; h += 8)
        for (int w = 0; w < cu_w; w += 8) {
            const int a = w / 4, b = h / 4;
            const int mx = (save[a][b][0] + save[a + 1][b][0] + save[a][b + 1][0] + save[a + 1][b + 1][0] + 2) >> 2;
            const int my = (save[a][b][1] + save[a + 1][b][1] + save[a][b + 1][1] + save[a + 1][b + 1][1] + 2) >> 2;
            const int gx = ((x + w) << 4) + mx, gy = ((y + h) << 4) + my;
            const int fx = gx & 31, fy = gy & 31;
            const int ix = CLIP3(-66, (pic_w + 4) >> 1, gx >> 5), iy = CLIP3(-66, (pic_h + 4) >> 1, gy >> 5);
            for (int pl = 0; pl < 2; pl++) {
                const pel *src = (pl ? v0 : u0) + iy * i_c + ix;
                pel *dst = (pl ? pred_v : pred_u) + (h / 2) * (cu_w / 2) + w / 2;
                if (fx == 0 && fy == 0) orc_ipcpy(src, i_c, dst, cu_w / 2, 4, 4);
                else if (fy == 0) orc_ipflt(1, src, i_c, dst, cu_w / 2, 4, 4, CHROMA_TAPS_HP[fx], max_val);
                else if (fx == 0) orc_ipflt(3, src, i_c, dst, cu_w / 2, 4, 4, CHROMA_TAPS_HP[fy], max_val);
                else orc_ipflt_ext(1, src, i_c, dst, cu_w / 2, 4, 4, CHROMA_TAPS_HP[fx], CHROMA_TAPS_HP[fy], max_val);
            }
        }
}

/* ------------------------------------------------------------------------ */
/* Affine gradient ME (affine_me_gradient, inter.c:1339-1471)                */
/* ------------------------------------------------------------------------ */

static int ilog2(int v) { int l = 0; while ((2 << l) <= v) l++; return l; }

/* Control-point MV bits (affine_mv_bits, inter.c:1317-1337): the reference
 * index bits, then per control point 2 * (floor(log2(|dx|+1)) +
 * floor(log2(|dy|+1))) + 2 + (dx != 0) + (dy != 0) with d = MVD >> shift. */
static int affine_bits(const int mv[2][2], const int mvp[2][2], int num_refp, int refi, int shift)
{
    int bits = orc_refi_bits(num_refp, refi);
    for (int v = 0; v < 2; v++) {
        const int dx = abs((mv[v][0] - mvp[v][0]) >> shift), dy = abs((mv[v][1] - mvp[v][1]) >> shift);
        bits += 2 * (ilog2(dx + 1) + ilog2(dy + 1)) + 2 + (dx != 0) + (dy != 0);
    }
    return bits;
}

/* MV_COST (inter.c:41): u32 arithmetic. */
static u32 mv_cost32(u32 lambda_mv, int bits) { return (lambda_mv * (u32)bits + 32768u) >> 16; }

/* Sobel gradients of the prediction (affine_sobel_flt_hor/ver,
 * com_util.c:2966-3018): the 3x3 kernels at every interior sample; a
 * border sample takes the value of the nearest interior one. */
static int sobel(const pel *p, int w, int h, int j, int k, int vertical)
{
    j = CLIP3(1, h - 2, j);
    k = CLIP3(1, w - 2, k);
    const pel *c = p + j * w + k;
    if (!vertical) return (c[-w + 1] - c[-w - 1]) + 2 * (c[1] - c[-1]) + (c[w + 1] - c[w - 1]);
    return (c[w - 1] - c[-w - 1]) + 2 * (c[w] - c[-w]) + (c[w + 1] - c[-w + 1]);
}

/* The 4-parameter least-squares system of one iteration
 * (affine_coef_computer, com_util.c:3020-3049): for every sample, with the
 * gradients (gx, gy), its coordinates (k, j) and the residual e,
 * g = (gx, k gx + j gy, gy, j gx - k gy); A[c][r] += g_c g_r and
 * b[c] += 8 g_c e, summed exactly in 64 bits. */
static void affine_system(const pel *org, int s_org, const pel *pred, int w, int h, s64 A[4][4], s64 b[4])
{
    memset(A, 0, sizeof(s64) * 16);
    memset(b, 0, sizeof(s64) * 4);
    for (int j = 0; j < h; j++)
        for (int k = 0; k < w; k++) {
            const int gx = sobel(pred, w, h, j, k, 0), gy = sobel(pred, w, h, j, k, 1);
            const int g[4] = {gx, k * gx + j * gy, gy, j * gx - k * gy};
            const s64 e = (s64)((int)org[j * s_org + k] - (int)pred[j * w + k]);
            for (int c = 0; c < 4; c++) {
                for (int r = 0; r < 4; r++) A[c][r] += (s64)g[c] * g[r];
                b[c] += (s64)g[c] * e * 8;
            }
        }
}

/* solve_equal (inter.c:1280-1315) of the 4x4 system [A | b]: Gaussian
 * elimination with partial pivoting in IEEE double, in the reference's
 * exact operation order (the result is defined by its roundings): rows 1..4
 * of m hold the equations, row 0 is the swap buffer; for each column c
 * the first row of largest |m[.][c]| at or below row c + 1 becomes the
 * pivot, and every element right of the column is reduced as
 * m[j][k] - m[i][k] * m[j][c] / m[i][c]; back substitution sums the known
 * terms from left to right. */
static void solve4(const s64 A[4][4], const s64 b[4], double x[4])
{
    double m[5][5];
    memset(m, 0, sizeof(m));
    for (int r = 0; r < 4; r++) {
        for (int c = 0; c < 4; c++) m[r + 1][c] = (double)A[r][c];
        m[r + 1][4] = (double)b[r];
    }
    const int n = 4;
    for (int i = 1; i < n; i++) {
        int piv = i;
        double big = fabs(m[i][i - 1]);
        for (int j = i + 1; j <= n; j++)
            if (fabs(m[j][i - 1]) > big) { big = fabs(m[j][i - 1]); piv = j; }
        if (piv != i)
            for (int c = 0; c <= n; c++) { m[0][c] = m[i][c]; m[i][c] = m[piv][c]; m[piv][c] = m[0][c]; }
        for (int j = i + 1; j <= n; j++)
            for (int k = i; k <= n; k++) m[j][k] = m[j][k] - m[i][k] * m[j][i - 1] / m[i][i - 1];
    }
    x[n - 1] = m[n][n] / m[n][n - 1];
    for (int i = n - 2; i >= 0; i--) {
        double t = 0;
        for (int j = i + 1; j < n; j++) t += m[i + 1][j] * x[j];
        x[i] = (m[i + 1][n] - t) / m[i + 1][i];
    }
}

/* (s16)(v * scale + (v >= 0 ? 0.5 : -0.5)) as the reference computes it
 * (inter.c:1413-1422). */
static s16 round_delta(double v, double scale) { return (s16)(v * scale + (v >= 0 ? 0.5 : -0.5)); }

/* affine_me_gradient (inter.c:1339-1471) with two control points for the
 * cu_w x cu_h CU at (x, y): start from mv; per iteration (7, or 5 with bi)
 * solve the gradient system of the current prediction for the change of
 * the 4-parameter model, turn it into control-point deltas at the AMVR
 * precision (Tab_Affine_AMVR: mvr 0 -> 1/4, 1 -> 1, 2 -> 1/16 pel), stop
 * when both deltas are zero, else move and clip the control points
 * (18-bit CPMVs, a clipped maximum rounded down to the precision) and keep
 * them when their cost MV_COST(bits) + (com_had >> bi) is strictly lower.
 * mv receives the best control points; returns that cost without its MV
 * bits.  plane0: the padded integer plane of the reference at (0, 0);
 * org / s_org: the CU's original (or bi target). */
EXPORT u64 orc_affine_me_gradient(int x, int y, int pic_w, int pic_h, int cu_w, int cu_h, int sub_w, int sub_h,
                                  const pel *plane0, int i_ref, const pel *org, int s_org, int refi, int num_refp,
                                  int mvr, int bi, int mot_bits_other, u32 lambda_mv, const int mvp[2][2],
                                  int mv[2][2], int bd)
{
    static pel pred[128 * 128];
    const int shift = mvr == 0 ? 2 : (mvr == 1 ? 4 : 0);
    const int iters = bi ? 5 : 7;
    const int CPMV_LO = -131072, CPMV_HI = 131071;
    int cp[3][2] = {{mv[0][0], mv[0][1]}, {mv[1][0], mv[1][1]}, {0, 0}};
    orc_affine_luma(x, y, pic_w, pic_h, cu_w, cu_h, cp, 2, plane0, i_ref, pred, sub_w, sub_h, bd);
    int best_bits = affine_bits((const int(*)[2])cp, mvp, num_refp, refi, shift) + (bi ? mot_bits_other : 0);
    u64 best = (u64)mv_cost32(lambda_mv, best_bits) + (orc_had(cu_w, cu_h, org, s_org, pred, cu_w, bd) >> bi);
    for (int it = 0; it < iters; it++) {
        s64 A[4][4], b[4];
        double p[4];
        affine_system(org, s_org, pred, cu_w, cu_h, A, b);
        solve4((const s64(*)[4])A, b, p);
        const double d[4] = {p[0], p[2], p[1] * cu_w + p[0], -p[3] * cu_w + p[2]};   /* v0x, v0y, v1x, v1y */
        s16 dv[4];
        for (int i = 0; i < 4; i++) {
            if (shift == 0) {
                dv[i] = round_delta(d[i], 16);
            } else {
                dv[i] = (s16)(round_delta(d[i], 4) * 4);                   /* quarter pel -> 1/16 pel */
                dv[i] = (s16)(((dv[i] + (1 << (shift - 1))) >> shift) << shift);
            }
        }
        if (!dv[0] && !dv[1] && !dv[2] && !dv[3]) break;
        for (int v = 0; v < 2; v++)
            for (int c = 0; c < 2; c++) {
                int t = CLIP3(CPMV_LO, CPMV_HI, cp[v][c] + dv[2 * v + c]);
                if (t == CPMV_HI) t = (t >> shift) << shift;
                cp[v][c] = t;
            }
        orc_affine_luma(x, y, pic_w, pic_h, cu_w, cu_h, cp, 2, plane0, i_ref, pred, sub_w, sub_h, bd);
        const int bits = affine_bits((const int(*)[2])cp, mvp, num_refp, refi, shift) + (bi ? mot_bits_other : 0);
        const u64 cost = (u64)mv_cost32(lambda_mv, bits) + (orc_had(cu_w, cu_h, org, s_org, pred, cu_w, bd) >> bi);
        if (cost < best) {
            best = cost;
            best_bits = bits;
            for (int v = 0; v < 2; v++) { mv[v][0] = cp[v][0]; mv[v][1] = cp[v][1]; }
        }
    }
    return best - mv_cost32(lambda_mv, best_bits);
}
