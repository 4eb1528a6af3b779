// Motion compensation of every PU of a set of CTUs (gfx950): com_mc_cu
// (com_mc.c:140-200) with the best reference of each list and its
// quarter-pel MV, for all selected PU shapes at once.
//
//   luma     com_mc_blk_luma, hp_flag 0 (com_mc.c:41-87): a copy from plane
//            [qy&3][qx&3] at the integer position clipped to
//            [-MAX_CU_SIZE-4, pic+4]
//   chroma   com_mc_blk_chroma, hp_flag 0 (com_mc.c:103-138): the luma
//            quarter-pel position read as a 1/8-pel chroma position, integer
//            part clipped to [-(MAX_CU_SIZE/2)-2, (pic+4)/2], then copy /
//            H_4 / V_4 / EXT_4 with com_tbl_mc_c_coeff (com_mc.c:555-657)
//   bi       pel_avrg, (a + b + 1) >> 1 (com_util.c:3088-3105)
//
// The device keeps chroma unpadded; every read clamps its coordinates into
// the plane, which is what the reference's edge-replicated pad of 72
// samples holds at every position these reads can reach (the clipped
// origin is >= -66, filter taps reach 1 sample before and 2 after, and a
// 64-sample block starting at (pic+4)/2 ends 68 samples past the edge).
#include "frame.h"
#include "affine.h"

#include <string.h>
#include <algorithm>

struct McArgs {
    JobSet js;
    const me_pu_result_t *res;
    const s8 *best;
    int npu, nref, nl;
    int pw, ph, base, per_ctu;
    size_t luma_off, chroma_off;   // this shape's offset inside a CTU's block
    size_t luma_ctu, chroma_ctu;   // elements per CTU
    int chroma;                    // 1 if this shape has chroma
    pel *py, *pu, *pv;
};

// Four luma samples from an arbitrary address (aligned dword loads and
// v_alignbyte; may read 3 bytes past the last, inside the padded plane).
__device__ __forceinline__ void load4(const pel *p, int o[4])
{
    const uintptr_t ad = (uintptr_t)p;
    const u32 *q = (const u32 *)(ad & ~(uintptr_t)3);
    const u32 sh = (u32)(ad & 3);
#if BIT_DEPTH == 8
    const u32 w = __builtin_amdgcn_alignbyte(q[1], q[0], sh);
#pragma unroll
    for (int i = 0; i < 4; i++) o[i] = (int)((w >> (8 * i)) & 0xFF);
#else
    const u32 w0 = __builtin_amdgcn_alignbyte(q[1], q[0], sh), w1 = __builtin_amdgcn_alignbyte(q[2], q[1], sh);
    o[0] = (int)(w0 & 0xFFFF);
    o[1] = (int)(w0 >> 16);
    o[2] = (int)(w1 & 0xFFFF);
    o[3] = (int)(w1 >> 16);
#endif
}

// grid (16, jobs): 4096 threads per (CTU, shape); thread t writes luma
// samples 4t..4t+3 of the shape's 16384 (chroma: k_mc_chroma).
__global__ __launch_bounds__(256) void k_mc(McArgs a)
{
    const int job = blockIdx.y;
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int cx0 = (ctu % P.ctu_cols) * 128, cy0 = (ctu / P.ctu_cols) * 128;
    const int cols = 128 / a.pw;
    const size_t pu0 = (size_t)job * a.npu + a.base;

    // luma
    {
        const int area4 = a.pw * a.ph / 4;
        const int i = t / area4, q = t - i * area4;
        const int row = q / (a.pw / 4), c4 = (q - row * (a.pw / 4)) * 4;
        const int x = cx0 + (i % cols) * a.pw, y = cy0 + (i / cols) * a.ph;
        const s8 *bp = a.best + (pu0 + i) * a.nl;
        int acc[4] = {0, 0, 0, 0};
        bool ok = true;
        for (int l = 0; l < a.nl; l++) ok &= bp[l] >= 0;
        if (ok) {
            for (int l = 0; l < a.nl; l++) {
                const int r = bp[l];
                const me_pu_result_t &o = a.res[(pu0 + i) * a.nref + r];
                const int qx = (x << 2) + o.mv_sub[0], qy = (y << 2) + o.mv_sub[1];
                const int ix = me_clip3(-ME_MAX_CU - 4, P.pic_w + 4, qx >> 2);
                const int iy = me_clip3(-ME_MAX_CU - 4, P.pic_h + 4, qy >> 2);
                const pel *src = P.planes[r] + (size_t)((qy & 3) * 4 + (qx & 3)) * P.plane_elems +
                                 (size_t)(iy + row + ME_PAD) * P.ref_stride + (ix + c4 + ME_PAD);
                int v[4];
                load4(src, v);
#pragma unroll
                for (int k = 0; k < 4; k++) acc[k] = l ? (acc[k] + v[k] + 1) >> 1 : v[k];
            }
        }
        pel *dst = a.py + (size_t)job * a.luma_ctu + a.luma_off + (size_t)i * a.pw * a.ph +
                   (size_t)row * a.pw + c4;
#if BIT_DEPTH == 8
        *(u32 *)dst = (u32)acc[0] | ((u32)acc[1] << 8) | ((u32)acc[2] << 16) | ((u32)acc[3] << 24);
#else
        *(uint2 *)dst = make_uint2((u32)acc[0] | ((u32)acc[1] << 16), (u32)acc[2] | ((u32)acc[3] << 16));
#endif
    }

}

// Four vertically adjacent chroma samples (rows row0 .. row0+3, column col)
// of a PU's (w/2) x (h/2) block predicted from plane `pl` (cw x chh,
// unpadded) at luma quarter-pel position (qx, qy), com_mc_blk_chroma: the
// position read as 1/8-pel, its integer part clipped, then copy / H_4 / V_4
// / 2-D, the 2-D filter's horizontal pass made once per input row for all
// four outputs (each output is the reference's own sum).  The 7 x 4
// input window (rows row0-1 .. row0+5, columns col-1 .. col+2) is read with
// dword loads away from the plane's edges, sample by sample with clamped
// coordinates near them.
__device__ __forceinline__ void mc_chroma_col4(const pel *pl, int cw, int chh, int pic_w, int pic_h, int qx, int qy,
                                               int col, int row0, int out[4])
{
    const int fx = qx & 7, fy = qy & 7;
    const int ix = me_clip3(-(ME_MAX_CU >> 1) - 2, (pic_w + 4) >> 1, qx >> 3) + col;
    const int iy = me_clip3(-(ME_MAX_CU >> 1) - 2, (pic_h + 4) >> 1, qy >> 3) + row0;
    const int maxv = (1 << BIT_DEPTH) - 1;
    const int x0 = ix - 1, y0 = iy - 1;
    int W[7][4];
    // load4 may read 2 samples past the fourth: keep that inside the row
    if (x0 >= 0 && x0 + 5 <= cw - 1 && y0 >= 0 && y0 + 6 <= chh - 1) {
#pragma unroll
        for (int j = 0; j < 7; j++) load4(pl + (size_t)(y0 + j) * cw + x0, W[j]);
    } else {
#pragma unroll
        for (int j = 0; j < 7; j++) {
            const pel *rp = pl + (size_t)me_clip3(0, chh - 1, y0 + j) * cw;
#pragma unroll
            for (int k = 0; k < 4; k++) W[j][k] = (int)rp[me_clip3(0, cw - 1, x0 + k)];
        }
    }
    if (fx == 0 && fy == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) out[r] = W[r + 1][1];
    } else if (fy == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int t = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) t += ME_CHROMA_TAPS[fx][k] * W[r + 1][k];
            out[r] = me_clip3(0, maxv, (t + 32) >> 6);
        }
    } else if (fx == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int t = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) t += ME_CHROMA_TAPS[fy][j] * W[r + j][1];
            out[r] = me_clip3(0, maxv, (t + 32) >> 6);
        }
    } else {
        // 2-D: horizontal pass stored as s16, then vertical (com_mc.c:587-643;
        // 8-bit: shift1 0, shift2 12; 10-bit: shift1 2, shift2 10)
        const int shift1 = BIT_DEPTH == 8 ? 0 : 2, shift2 = BIT_DEPTH == 8 ? 12 : 10;
        const int add1 = (1 << shift1) >> 1, add2 = 1 << (shift2 - 1);
        int h[7];
#pragma unroll
        for (int j = 0; j < 7; j++) {
            int t = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) t += ME_CHROMA_TAPS[fx][k] * W[j][k];
            h[j] = (int)(s16)((t + add1) >> shift1);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int v = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) v += ME_CHROMA_TAPS[fy][j] * h[r + j];
            out[r] = me_clip3(0, maxv, (v + add2) >> shift2);
        }
    }
}

// grid (4, jobs): 1024 threads per (CTU, shape with chroma); thread t
// writes a column of 4 samples of u and of v: PU i = t / (cw * ch / 4), then
// row group and column inside the PU's cw x ch chroma block.
__global__ __launch_bounds__(256) void k_mc_chroma(McArgs a)
{
    const int job = blockIdx.y;
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int cx0 = (ctu % P.ctu_cols) * 128, cy0 = (ctu / P.ctu_cols) * 128;
    const int cols = 128 / a.pw;
    const size_t pu0 = (size_t)job * a.npu + a.base;
    const int cw = a.pw / 2, chh = a.ph / 2, area = cw * chh;
    const int per_pu = area / 4;
    const int i = t / per_pu, q = t - i * per_pu;
    const int rg = q / cw, col = q - rg * cw, row0 = rg * 4;
    const int x = cx0 + (i % cols) * a.pw, y = cy0 + (i / cols) * a.ph;
    const s8 *bp = a.best + (pu0 + i) * a.nl;
    const int pcw = P.pic_w / 2, pch = P.pic_h / 2;
    int u[4] = {0, 0, 0, 0}, v[4] = {0, 0, 0, 0};
    bool ok = true;
    for (int l = 0; l < a.nl; l++) ok &= bp[l] >= 0;
    if (ok) {
        for (int l = 0; l < a.nl; l++) {
            const int r = bp[l];
            const me_pu_result_t &o = a.res[(pu0 + i) * a.nref + r];
            const int qx = (x << 2) + o.mv_sub[0], qy = (y << 2) + o.mv_sub[1];
            const pel *pu = P.chroma[r], *pv = pu + (size_t)pcw * pch;
            int su[4], sv[4];
            mc_chroma_col4(pu, pcw, pch, P.pic_w, P.pic_h, qx, qy, col, row0, su);
            mc_chroma_col4(pv, pcw, pch, P.pic_w, P.pic_h, qx, qy, col, row0, sv);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                u[k] = l ? (u[k] + su[k] + 1) >> 1 : su[k];
                v[k] = l ? (v[k] + sv[k] + 1) >> 1 : sv[k];
            }
        }
    }
    const size_t o = (size_t)job * a.chroma_ctu + a.chroma_off + (size_t)i * area + (size_t)row0 * cw + col;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        a.pu[o + (size_t)k * cw] = (pel)u[k];
        a.pv[o + (size_t)k * cw] = (pel)v[k];
    }
}

void me_mc_sizes(const MeSearch *s, size_t *luma_per_ctu, size_t *chroma_per_ctu)
{
    size_t ly = 0, lc = 0;
    for (int k = 0; k < ME_NSHAPES; k++) {
        if (s->base[k] < 0) continue;
        ly += 16384;
        if (ME_SHAPES_HOST[k][0] >= 8 && ME_SHAPES_HOST[k][1] >= 8) lc += 4096;
    }
    *luma_per_ctu = ly;
    *chroma_per_ctu = lc;
}

int me_mc_enqueue(const MeSearch *s, const JobSet &js, const me_pu_result_t *res, const s8 *best, pel *pred_y,
                  pel *pred_u, pel *pred_v, hipStream_t st)
{
    McArgs a;
    memset(&a, 0, sizeof(a));
    a.js = js;
    a.res = res;
    a.best = best;
    a.npu = s->npu;
    a.nref = s->nref;
    a.nl = s->nl;
    me_mc_sizes(s, &a.luma_ctu, &a.chroma_ctu);
    a.py = pred_y;
    a.pu = pred_u;
    a.pv = pred_v;
    const bool with_chroma = pred_u && pred_v && a.chroma_ctu > 0;
    size_t lo = 0, co = 0;
    for (int k = 0; k < ME_NSHAPES; k++) {
        if (s->base[k] < 0) continue;
        a.pw = ME_SHAPES_HOST[k][0];
        a.ph = ME_SHAPES_HOST[k][1];
        a.base = s->base[k];
        a.per_ctu = (128 / a.pw) * (128 / a.ph);
        a.luma_off = lo;
        a.chroma_off = co;
        const bool has_c = a.pw >= 8 && a.ph >= 8;
        a.chroma = with_chroma && has_c;
        hipLaunchKernelGGL(k_mc, dim3(16, js.n_jobs), dim3(256), 0, st, a);
        ME_CHECK(hipGetLastError());
        if (a.chroma) {
            hipLaunchKernelGGL(k_mc_chroma, dim3(4, js.n_jobs), dim3(256), 0, st, a);
            ME_CHECK(hipGetLastError());
        }
        lo += 16384;
        co += has_c ? 4096 : 0;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Affine motion compensation, luma (com_mc_blk_affine_luma, com_mc.c:202-250)
// ---------------------------------------------------------------------------

// The 1/32-pel chroma interpolation filter (com_tbl_mc_c_coeff_hp,
// com_tables.c:156-189): the standard's table.
__constant__ static const s8 AFF_CTAPS[32][4] = {
    {0, 64, 0, 0},    {-1, 63, 2, 0},   {-2, 62, 4, 0},   {-2, 60, 7, -1},  {-2, 58, 10, -2}, {-3, 57, 12, -2},
    {-4, 56, 14, -2}, {-4, 55, 15, -2}, {-4, 54, 16, -2}, {-5, 53, 18, -2}, {-6, 52, 20, -2}, {-6, 49, 24, -3},
    {-6, 46, 28, -4}, {-5, 44, 29, -4}, {-4, 42, 30, -4}, {-4, 39, 33, -4}, {-4, 36, 36, -4}, {-4, 33, 39, -4},
    {-4, 30, 42, -4}, {-4, 29, 44, -5}, {-4, 28, 46, -6}, {-3, 24, 49, -6}, {-2, 20, 52, -6}, {-2, 18, 53, -5},
    {-2, 16, 54, -4}, {-2, 15, 55, -4}, {-2, 14, 56, -4}, {-2, 12, 57, -3}, {-2, 10, 58, -2}, {-1, 7, 60, -2},
    {0, 4, 62, -2},   {0, 2, 63, -1}};

// The sub-block MV (1/16 pel, 18 bits) of the sub-block at (w0, h0) of CU c
// (com_mc.c:224-241, 290-310).
__device__ __forceinline__ void aff_sub_mv(const me_affine_cu_t &c, int w0, int h0, int &mx, int &my)
{
    const int lw = __builtin_ctz((unsigned)c.w), lh = __builtin_ctz((unsigned)c.h);
    const int dhx = ((c.cpmv[1][0] - c.cpmv[0][0]) * 128) >> lw, dhy = ((c.cpmv[1][1] - c.cpmv[0][1]) * 128) >> lw;
    const bool three = c.vertex_num == 3;
    const int dvx = three ? ((c.cpmv[2][0] - c.cpmv[0][0]) * 128) >> lh : -dhy;
    const int dvy = three ? ((c.cpmv[2][1] - c.cpmv[0][1]) * 128) >> lh : dhx;
    int px = w0 + c.sub_w / 2, py = h0 + c.sub_h / 2;
    if (w0 == 0 && h0 == 0) { px = 0; py = 0; }
    else if (w0 + c.sub_w == c.w && h0 == 0) { px = c.w; py = 0; }
    else if (w0 == 0 && h0 + c.sub_h == c.h && three) { px = 0; py = c.h; }
    mx = me_clip3(-131072, 131071, aff_round7(c.cpmv[0][0] * 128 + dhx * px + dvx * py));
    my = me_clip3(-131072, 131071, aff_round7(c.cpmv[0][1] * 128 + dhy * px + dvy * py));
}

// One thread per chroma sample of u and v; grid (blocks for the batch's
// largest CU, n): block b of CU c covers its chroma samples [256 b, 256 b + 256).
__global__ __launch_bounds__(256) void k_affine_chroma(const me_affine_cu_t *cus, const pel *chroma, int pic_w,
                                                       int pic_h, pel *pred_u, pel *pred_v)
{
    const me_affine_cu_t c = cus[blockIdx.y];
    const int cwid = c.w >> 1, k = blockIdx.x * 256 + threadIdx.x;
    if (k >= cwid * (c.h >> 1)) return;
    const int cx = k % cwid, cy = k / cwid;
    const int w8 = (cx >> 2) << 3, h8 = (cy >> 2) << 3;   // the 8x8 luma block
    // its four 4x4 luma units' MVs (mv_save, com_mc.c:312-325), averaged (com_mc.c:346-347)
    int sx = 0, sy = 0;
    for (int u = 0; u < 4; u++) {
        const int lx = w8 + (u & 1) * 4, ly = h8 + (u >> 1) * 4;
        int mx, my;
        aff_sub_mv(c, lx - lx % c.sub_w, ly - ly % c.sub_h, mx, my);
        sx += mx;
        sy += my;
    }
    const int gx = ((c.x + w8) << 4) + ((sx + 2) >> 2), gy = ((c.y + h8) << 4) + ((sy + 2) >> 2);
    const int fx = gx & 31, fy = gy & 31;
    const int ix = me_clip3(-(ME_MAX_CU >> 1) - 2, (pic_w + 4) >> 1, gx >> 5) + (cx & 3);
    const int iy = me_clip3(-(ME_MAX_CU >> 1) - 2, (pic_h + 4) >> 1, gy >> 5) + (cy & 3);
    const int cw = pic_w >> 1, chh = pic_h >> 1, maxv = (1 << BIT_DEPTH) - 1;
    for (int pl = 0; pl < 2; pl++) {
        const pel *P = chroma + (size_t)pl * cw * chh;
        auto S = [&](int xx, int yy) -> int {
            return (int)P[(size_t)me_clip3(0, chh - 1, yy) * cw + me_clip3(0, cw - 1, xx)];
        };
        int v;
        if (fx == 0 && fy == 0) {
            v = S(ix, iy);
        } else if (fy == 0 || fx == 0) {
            int t = 0;
            for (int j = 0; j < 4; j++)
                t += (fy == 0 ? AFF_CTAPS[fx][j] * S(ix - 1 + j, iy) : AFF_CTAPS[fy][j] * S(ix, iy - 1 + j));
            v = me_clip3(0, maxv, (t + 32) >> 6);
        } else {
            const int shift1 = BIT_DEPTH == 8 ? 0 : 2, shift2 = BIT_DEPTH == 8 ? 12 : 10;
            const int add1 = (1 << shift1) >> 1, add2 = 1 << (shift2 - 1);
            int t2 = 0;
            for (int j = 0; j < 4; j++) {
                int t = 0;
                for (int i = 0; i < 4; i++) t += AFF_CTAPS[fx][i] * S(ix - 1 + i, iy - 1 + j);
                t2 += AFF_CTAPS[fy][j] * (int)(s16)((t + add1) >> shift1);
            }
            v = me_clip3(0, maxv, (t2 + add2) >> shift2);
        }
        (pl ? pred_v : pred_u)[c.chroma_offset + k] = (pel)v;
    }
}

// One output sample per thread; grid (blocks for the batch's largest CU, n):
// block b of CU c covers its samples [256 b, 256 b + 256).
__global__ __launch_bounds__(256) void k_affine_luma(const me_affine_cu_t *cus, const pel *plane0, int stride,
                                                     int pic_w, int pic_h, pel *pred)
{
    const me_affine_cu_t c = cus[blockIdx.y];
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= c.w * c.h) return;
    pred[c.pred_offset + k] = (pel)aff_luma_sample(c.cpmv, c.vertex_num, c.x, c.y, c.w, c.h, c.sub_w, c.sub_h,
                                                   k % c.w, k / c.w, plane0, stride, pic_w, pic_h);
}

ME_EXPORT int me_affine_mc_luma(me_ref_t r, const me_affine_cu_t *cus, int n, me_pel *pred, uint64_t pred_elems,
                                void *stream)
{
    if (!r || !cus || !pred || n < 0) {
        me_set_error("me_affine_mc_luma: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    for (int i = 0; i < n; i++) {
        const me_affine_cu_t &c = cus[i];
        const bool pow2 = c.w >= 16 && c.w <= 128 && c.h >= 16 && c.h <= 128 && !(c.w & (c.w - 1)) && !(c.h & (c.h - 1));
        const bool sub = (c.sub_w == 4 || c.sub_w == 8) && (c.sub_h == 4 || c.sub_h == 8);
        const bool cps = c.vertex_num == 2 || c.vertex_num == 3;
        if (!pow2 || !sub || !cps || c.pred_offset + (uint64_t)c.w * c.h > pred_elems) {
            me_set_error("me_affine_mc_luma: CU size, sub-block size, control points or prediction offset out of range",
                         hipErrorInvalidValue, __FILE__, __LINE__);
            return -1;
        }
    }
    if (n == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (r->aff_cus_cap < n) {
        (void)hipFree(r->aff_cus);
        r->aff_cus = nullptr;
        r->aff_cus_cap = 0;
        ME_CHECK(hipMalloc((void **)&r->aff_cus, (size_t)n * sizeof(me_affine_cu_t)));
        r->aff_cus_cap = n;
    }
    if (r->aff_pred_cap < pred_elems) {
        (void)hipFree(r->aff_pred);
        r->aff_pred = nullptr;
        r->aff_pred_cap = 0;
        ME_CHECK(hipMalloc((void **)&r->aff_pred, pred_elems * sizeof(pel)));
        r->aff_pred_cap = pred_elems;
    }
    ME_CHECK(hipMemcpyAsync(r->aff_cus, cus, (size_t)n * sizeof(me_affine_cu_t), hipMemcpyHostToDevice, st));
    ME_CHECK(hipMemcpyAsync(r->aff_pred, pred, pred_elems * sizeof(pel), hipMemcpyHostToDevice, st));
    int max_area = 0;
    for (int i = 0; i < n; i++) max_area = std::max(max_area, cus[i].w * cus[i].h);
    hipLaunchKernelGGL(k_affine_luma, dim3((max_area + 255) / 256, n), dim3(256), 0, st, r->aff_cus,
                       r->planes + (size_t)ME_PAD * r->stride + ME_PAD, r->stride, r->w, r->h, r->aff_pred);
    ME_CHECK(hipGetLastError());
    ME_CHECK(hipMemcpyAsync(pred, r->aff_pred, pred_elems * sizeof(pel), hipMemcpyDeviceToHost, st));
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

ME_EXPORT int me_affine_mc_chroma(me_ref_t r, const me_affine_cu_t *cus, int n, me_pel *pred_u, me_pel *pred_v,
                                  uint64_t chroma_elems, void *stream)
{
    if (!r || !cus || !pred_u || !pred_v || n < 0) {
        me_set_error("me_affine_mc_chroma: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (!r->chroma) {
        me_set_error("me_affine_mc_chroma: the reference has no chroma (me_ref_load_chroma)", hipErrorInvalidValue,
                     __FILE__, __LINE__);
        return -1;
    }
    for (int i = 0; i < n; i++) {
        const me_affine_cu_t &c = cus[i];
        const bool pow2 = c.w >= 16 && c.w <= 128 && c.h >= 16 && c.h <= 128 && !(c.w & (c.w - 1)) && !(c.h & (c.h - 1));
        const bool sub = (c.sub_w == 4 || c.sub_w == 8) && c.sub_h == c.sub_w;
        const bool cps = c.vertex_num == 2 || c.vertex_num == 3;
        if (!pow2 || !sub || !cps || c.chroma_offset + (uint64_t)(c.w / 2) * (c.h / 2) > chroma_elems) {
            me_set_error("me_affine_mc_chroma: CU size, sub-block size (4x4 or 8x8), control points or offset out of "
                         "range", hipErrorInvalidValue, __FILE__, __LINE__);
            return -1;
        }
    }
    if (n == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (r->aff_cus_cap < n) {
        (void)hipFree(r->aff_cus);
        r->aff_cus = nullptr;
        r->aff_cus_cap = 0;
        ME_CHECK(hipMalloc((void **)&r->aff_cus, (size_t)n * sizeof(me_affine_cu_t)));
        r->aff_cus_cap = n;
    }
    if (r->aff_pred_cap < 2 * chroma_elems) {
        (void)hipFree(r->aff_pred);
        r->aff_pred = nullptr;
        r->aff_pred_cap = 0;
        ME_CHECK(hipMalloc((void **)&r->aff_pred, 2 * chroma_elems * sizeof(pel)));
        r->aff_pred_cap = 2 * chroma_elems;
    }
    pel *du = r->aff_pred, *dv = r->aff_pred + chroma_elems;
    ME_CHECK(hipMemcpyAsync(r->aff_cus, cus, (size_t)n * sizeof(me_affine_cu_t), hipMemcpyHostToDevice, st));
    ME_CHECK(hipMemcpyAsync(du, pred_u, chroma_elems * sizeof(pel), hipMemcpyHostToDevice, st));
    ME_CHECK(hipMemcpyAsync(dv, pred_v, chroma_elems * sizeof(pel), hipMemcpyHostToDevice, st));
    int max_area = 0;
    for (int i = 0; i < n; i++) max_area = std::max(max_area, (cus[i].w / 2) * (cus[i].h / 2));
    hipLaunchKernelGGL(k_affine_chroma, dim3((max_area + 255) / 256, n), dim3(256), 0, st, r->aff_cus, r->chroma, r->w,
                       r->h, du, dv);
    ME_CHECK(hipGetLastError());
    ME_CHECK(hipMemcpyAsync(pred_u, du, chroma_elems * sizeof(pel), hipMemcpyDeviceToHost, st));
    ME_CHECK(hipMemcpyAsync(pred_v, dv, chroma_elems * sizeof(pel), hipMemcpyDeviceToHost, st));
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}
