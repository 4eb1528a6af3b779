// Affine luma prediction shared by affine motion compensation (mc.hip) and
// the affine gradient search (affine_me.hip).
#pragma once
#include "common.h"

// The 1/16-pel luma interpolation filter (com_tbl_mc_l_coeff_hp,
// com_tables.c:130-147): the standard's table.
__constant__ static const s8 AFF_TAPS[16][8] = {
    {0, 0, 0, 64, 0, 0, 0, 0},        {0, 1, -3, 63, 4, -2, 1, 0},      {-1, 2, -5, 62, 8, -3, 1, 0},
    {-1, 3, -8, 60, 13, -4, 1, 0},    {-1, 4, -10, 58, 17, -5, 1, 0},   {-1, 4, -11, 52, 26, -8, 3, -1},
    {-1, 3, -9, 47, 31, -10, 4, -1},  {-1, 4, -11, 45, 34, -10, 4, -1}, {-1, 4, -11, 40, 40, -11, 4, -1},
    {-1, 4, -10, 34, 45, -11, 4, -1}, {-1, 4, -10, 31, 47, -9, 3, -1},  {-1, 3, -8, 26, 52, -11, 4, -1},
    {0, 1, -5, 17, 58, -10, 4, -1},   {0, 1, -4, 13, 60, -8, 3, -1},    {0, 1, -3, 8, 62, -5, 2, -1},
    {0, 1, -2, 4, 63, -3, 1, 0}};

__device__ __forceinline__ int aff_round7(int v) { return v >= 0 ? (v + 64) >> 7 : -((-v + 64) >> 7); }

// The luma sample (sx, sy) of the affine prediction of the w x h CU at
// (x, y) (com_mc_blk_affine_luma, com_mc.c:202-250, and the luma of
// com_mc_blk_affine, com_mc.c:252-333): its sub_w x sub_h sub-block takes
// the model's MV at the sub-block centre (the top-left and top-right
// sub-blocks the CU's corners, with three control points also the
// bottom-left one), rounded to 1/16 pel and clipped to 18 bits
// (com_mc.c:224-241); the integer position is clipped to [-132, pic + 4]
// and the 1/16 phase picks a copy, a one-direction 8-tap filter or the
// two-pass filter with the ipflt_ext shifts (com_mc.c:461-657).  The
// reference's test for the half-sample planes (com_mc.c:65) never holds, so
// the filters always run.  plane0: the padded integer plane at (0, 0).
__device__ __forceinline__ int aff_luma_sample(const int32_t (*cpmv)[2], int vertex_num, int x, int y, int w, int h,
                                               int sub_w, int sub_h, int sx, int sy, const pel *plane0, int stride,
                                               int pic_w, int pic_h)
{
    const int w0 = sx - sx % sub_w, h0 = sy - sy % sub_h;
    const int lw = __builtin_ctz((unsigned)w), lh = __builtin_ctz((unsigned)h);
    const int dhx = ((cpmv[1][0] - cpmv[0][0]) * 128) >> lw, dhy = ((cpmv[1][1] - cpmv[0][1]) * 128) >> lw;
    const bool three = vertex_num == 3;
    const int dvx = three ? ((cpmv[2][0] - cpmv[0][0]) * 128) >> lh : -dhy;
    const int dvy = three ? ((cpmv[2][1] - cpmv[0][1]) * 128) >> lh : dhx;
    int px = w0 + sub_w / 2, py = h0 + sub_h / 2;
    if (w0 == 0 && h0 == 0) { px = 0; py = 0; }
    else if (w0 + sub_w == w && h0 == 0) { px = w; py = 0; }
    else if (w0 == 0 && h0 + sub_h == h && three) { px = 0; py = h; }
    int mx = aff_round7(cpmv[0][0] * 128 + dhx * px + dvx * py);
    int my = aff_round7(cpmv[0][1] * 128 + dhy * px + dvy * py);
    mx = me_clip3(-131072, 131071, mx);
    my = me_clip3(-131072, 131071, my);
    const int gx = ((x + w0) << 4) + mx, gy = ((y + h0) << 4) + my;
    const int fx = gx & 15, fy = gy & 15;
    const int ix = me_clip3(-ME_MAX_CU - 4, pic_w + 4, gx >> 4), iy = me_clip3(-ME_MAX_CU - 4, pic_h + 4, gy >> 4);
    const pel *src = plane0 + (ptrdiff_t)(iy + sy - h0) * stride + (ix + sx - w0);
    const int maxv = (1 << BIT_DEPTH) - 1;
    if (fx == 0 && fy == 0) return src[0];
    if (fy == 0 || fx == 0) {
        const int step = fy == 0 ? 1 : stride, f = fy == 0 ? fx : fy;
        int t = 0;
        for (int j = 0; j < 8; j++) t += (int)src[(j - 3) * step] * AFF_TAPS[f][j];
        return me_clip3(0, maxv, (t + 32) >> 6);
    }
    const int shift1 = BIT_DEPTH == 8 ? 0 : 2, shift2 = BIT_DEPTH == 8 ? 12 : 10;
    const int add1 = (1 << shift1) >> 1, add2 = 1 << (shift2 - 1);
    int t = 0;
    for (int j = 0; j < 8; j++) {
        const pel *r = src + (ptrdiff_t)(j - 3) * stride;
        int u = 0;
        for (int i = 0; i < 8; i++) u += (int)r[i - 3] * AFF_TAPS[fx][i];
        t += (int)(s16)((u + add1) >> shift1) * AFF_TAPS[fy][j];
    }
    return me_clip3(0, maxv, (t + add2) >> shift2);
}
