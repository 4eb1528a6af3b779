// Affine gradient motion search (affine_me_gradient, inter.c:1339-1471) for
// a batch of CUs on gfx950.
//
// One workgroup of 256 lanes per CU.  Per iteration (7, or 5 for the
// bi-prediction form) the workgroup
//   1. predicts the CU from its two control points into LDS (every lane a
//      stride of samples, com_mc_blk_affine_luma, shared with affine MC);
//   2. accumulates the 4-parameter least-squares system of the Sobel
//      gradients and the residual (affine_sobel_flt_hor/ver and
//      affine_coef_computer, com_util.c:2966-3049): 14 distinct 64-bit sums,
//      exact integers, so the lane / wave order of the reduction cannot
//      change them;
//   3. on one lane, solves the system in IEEE double exactly as solve_equal
//      does (inter.c:1280-1315; FP contraction off, the reference's
//      operation order) and turns the solution into control-point deltas at
//      the AMVR precision (inter.c:1405-1431);
//   4. stops on zero deltas, else moves the control points, predicts again
//      and keeps them when MV_COST(bits) + (com_had >> bi) is strictly lower.
// com_had is one lane per 8x8 / 16x8 / 8x16 tile (at most 256 tiles for a
// 128x128 CU), summed across the workgroup.
#include "frame.h"
#include "affine.h"

#include <string.h>

// The reference's double -> s16 conversions (inter.c:1413-1422) compile to
// the 32-bit truncating convert of x86-64 (cvttsd2si), which gives INT_MIN
// for NaN and for values outside the int32 range; the low 16 bits are kept.
// gfx950's convert saturates instead, so the out-of-range case is explicit.
__device__ __forceinline__ s16 cvt_s16_like_x86(double v)
{
    const int i = (v > -2147483649.0 && v < 2147483648.0) ? (int)v : (int)0x80000000;
    return (s16)i;
}

__device__ __forceinline__ s16 round_delta(double v, double scale)
{
#pragma clang fp contract(off)
    return cvt_s16_like_x86(v * scale + (v >= 0 ? 0.5 : -0.5));
}

// affine_mv_bits (inter.c:1317-1337).
__device__ __forceinline__ int affine_bits(const int (*mv)[2], const int32_t (*mvp)[2], int num_refp, int refi,
                                           int shift)
{
    int bits = num_refp <= 1 ? 0 : (refi < num_refp - 1 ? refi + 1 : num_refp - 1);   // tbl_refi_bits
    for (int v = 0; v < 2; v++) {
        const int dx = abs((mv[v][0] - mvp[v][0]) >> shift), dy = abs((mv[v][1] - mvp[v][1]) >> shift);
        bits += 2 * ((31 - __clz(dx + 1)) + (31 - __clz(dy + 1))) + 2 + (dx != 0) + (dy != 0);
    }
    return bits;
}

__device__ __forceinline__ u32 mv_cost32(u32 lambda, int bits) { return (lambda * (u32)bits + 32768u) >> 16; }

// solve_equal (inter.c:1280-1315) for the 4 parameters: m rows 1..4 hold the
// equations [A | b], row 0 is the swap buffer.
__device__ void solve4(double (*m)[5], double *x)
{
#pragma clang fp contract(off)
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

// The Sobel gradient at (j, k) of the w x h prediction p: every border
// sample takes the nearest interior sample's value (com_util.c:2966-3018).
__device__ __forceinline__ void sobel2(const pel *p, int w, int h, int j, int k, int &gx, int &gy)
{
    j = me_clip3(1, h - 2, j);
    k = me_clip3(1, w - 2, k);
    const pel *c = p + j * w + k;
    const int a = c[-w - 1], b = c[-w], d = c[-w + 1], e = c[-1], f = c[1], g = c[w - 1], hh = c[w], i = c[w + 1];
    gx = (d - a) + 2 * (f - e) + (i - g);
    gy = (g - a) + 2 * (hh - b) + (i - d);
}

__device__ __forceinline__ s64 wave_sum64(s64 v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ u32 wave_sum32(u32 v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct AffMeArgs {
    me_affine_me_cu_t *cus;
    const pel *org;        // originals of the batch (org_offset / org_stride per CU)
    const pel *plane0;     // padded integer plane of the reference at (0, 0)
    int stride, pic_w, pic_h;
    u32 lambda;
};

// The CU's prediction from control points cp into LDS.
__device__ __forceinline__ void predict(const me_affine_me_cu_t &c, const int (*cp)[2], const AffMeArgs &a, pel *pred)
{
    const int32_t cpm[3][2] = {{cp[0][0], cp[0][1]}, {cp[1][0], cp[1][1]}, {0, 0}};
    for (int k = threadIdx.x; k < c.w * c.h; k += 256)
        pred[k] = (pel)aff_luma_sample(cpm, 2, c.x, c.y, c.w, c.h, c.sub_w, c.sub_h, k % c.w, k / c.w, a.plane0, a.stride,
                                       a.pic_w, a.pic_h);
}

// com_had (com_cost.c:712-812) of the CU's original against pred, one lane
// per tile; the sum is valid on lane 0 of the workgroup.
__device__ __forceinline__ u32 had_cu(const me_affine_me_cu_t &c, const pel *org, const pel *pred, u32 *part)
{
    int tt, tw, th;
    had_tiling(c.w, c.h, &tt, &tw, &th);
    const int tpr = c.w / tw, nt = tpr * (c.h / th);
    u32 v = 0;
    if ((int)threadIdx.x < nt) {
        const int tx = (threadIdx.x % tpr) * tw, ty = (threadIdx.x / tpr) * th;
        v = (u32)satd_any(tt, org + (size_t)ty * c.org_stride + tx, c.org_stride, pred + ty * c.w + tx, c.w);
    }
    v = wave_sum32(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    const u32 s = part[0] + part[1] + part[2] + part[3];
    __syncthreads();
    return s >> (BIT_DEPTH - 8);
}

__global__ __launch_bounds__(256) void k_affine_me(AffMeArgs a)
{
    __shared__ pel pred[128 * 128];
    __shared__ s64 sums[4][14];
    __shared__ u32 part[4];
    __shared__ int s_cp[2][2];
    __shared__ int s_stop;
    const me_affine_me_cu_t c = a.cus[blockIdx.x];
    const pel *org = a.org + c.org_offset;
    const int w = c.w, h = c.h, bi = c.bi;
    const int shift = c.mvr == 0 ? 2 : (c.mvr == 1 ? 4 : 0);   // Tab_Affine_AMVR
    const int extra = bi ? c.mot_bits_other : 0;
    int cp[2][2] = {{c.mv[0][0], c.mv[0][1]}, {c.mv[1][0], c.mv[1][1]}};
    predict(c, cp, a, pred);
    __syncthreads();
    int best_bits = affine_bits(cp, c.mvp, c.num_refp, c.refi, shift) + extra;
    u64 best = (u64)mv_cost32(a.lambda, best_bits) + (had_cu(c, org, pred, part) >> bi);
    int best_cp[2][2] = {{cp[0][0], cp[0][1]}, {cp[1][0], cp[1][1]}};
    const int iters = bi ? 5 : 7;   // AF_ITER_BI, AF_ITER_UNI (com_define.h:198-199)
    for (int it = 0; it < iters; it++) {
        // the least-squares system: A[c][r] (10 distinct, symmetric) and b[c]
        s64 acc[14];
        for (int q = 0; q < 14; q++) acc[q] = 0;
        for (int k = threadIdx.x; k < w * h; k += 256) {
            const int j = k / w, kk = k % w;
            int gx, gy;
            sobel2(pred, w, h, j, kk, gx, gy);
            const int g[4] = {gx, kk * gx + j * gy, gy, j * gx - kk * gy};
            const s64 e = (s64)((int)org[(size_t)j * c.org_stride + kk] - (int)pred[k]);
            int q = 0;
            for (int r = 0; r < 4; r++)
                for (int s = r; s < 4; s++) acc[q++] += (s64)g[r] * g[s];
            for (int r = 0; r < 4; r++) acc[10 + r] += (s64)g[r] * e * 8;
        }
        for (int q = 0; q < 14; q++) {
            const s64 v = wave_sum64(acc[q]);
            if ((threadIdx.x & 63) == 0) sums[threadIdx.x >> 6][q] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s64 t[14];
            for (int q = 0; q < 14; q++) t[q] = sums[0][q] + sums[1][q] + sums[2][q] + sums[3][q];
            double m[5][5];
            for (int i = 0; i < 5; i++) m[0][i] = 0;
            int q = 0;
            for (int r = 0; r < 4; r++)
                for (int s = r; s < 4; s++, q++) m[r + 1][s] = m[s + 1][r] = (double)t[q];
            for (int r = 0; r < 4; r++) m[r + 1][4] = (double)t[10 + r];
            double p[4];
            solve4(m, p);
            double d[4];
            {
#pragma clang fp contract(off)
                d[0] = p[0];
                d[1] = p[2];
                d[2] = p[1] * w + p[0];
                d[3] = -p[3] * w + p[2];
            }
            s16 dv[4];
            for (int i = 0; i < 4; i++) {
                if (shift == 0) {
                    dv[i] = round_delta(d[i], 16);
                } else {
                    dv[i] = (s16)(round_delta(d[i], 4) * 4);   // quarter pel -> 1/16 pel
                    dv[i] = (s16)(((dv[i] + (1 << (shift - 1))) >> shift) << shift);
                }
            }
            s_stop = !dv[0] && !dv[1] && !dv[2] && !dv[3];
            for (int v = 0; v < 2; v++)
                for (int k = 0; k < 2; k++) {
                    int u = me_clip3(-131072, 131071, cp[v][k] + dv[2 * v + k]);
                    if (u == 131071) u = (u >> shift) << shift;
                    s_cp[v][k] = u;
                }
        }
        __syncthreads();
        if (s_stop) break;
        for (int v = 0; v < 2; v++)
            for (int k = 0; k < 2; k++) cp[v][k] = s_cp[v][k];
        predict(c, cp, a, pred);
        __syncthreads();
        const int bits = affine_bits(cp, c.mvp, c.num_refp, c.refi, shift) + extra;
        const u64 cost = (u64)mv_cost32(a.lambda, bits) + (had_cu(c, org, pred, part) >> bi);
        if (cost < best) {   // uniform: every lane holds the same sums
            best = cost;
            best_bits = bits;
            for (int v = 0; v < 2; v++)
                for (int k = 0; k < 2; k++) best_cp[v][k] = cp[v][k];
        }
    }
    if (threadIdx.x == 0) {
        me_affine_me_cu_t &out = a.cus[blockIdx.x];
        for (int v = 0; v < 2; v++)
            for (int k = 0; k < 2; k++) out.mv[v][k] = best_cp[v][k];
        out.cost = best - mv_cost32(a.lambda, best_bits);
    }
}

ME_EXPORT int me_affine_me_gradient(me_ref_t r, const me_pel *org, uint64_t org_elems, me_affine_me_cu_t *cus, int n,
                                    uint32_t lambda_mv, void *stream)
{
    if (!r || !org || !cus || n < 0) {
        me_set_error("me_affine_me_gradient: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (n == 0) return 0;
    auto pow2 = [](int v) { return v == 16 || v == 32 || v == 64 || v == 128; };
    for (int i = 0; i < n; i++) {
        const me_affine_me_cu_t &c = cus[i];
        const bool ok = pow2(c.w) && pow2(c.h) && (c.sub_w == 4 || c.sub_w == 8) && (c.sub_h == 4 || c.sub_h == 8) &&
                        c.mvr >= 0 && c.mvr <= 2 && (c.bi == 0 || c.bi == 1) && c.num_refp >= 1 &&
                        c.num_refp <= ME_MAX_REFS && c.refi >= 0 && c.refi < c.num_refp && c.org_stride >= c.w &&
                        c.org_offset + (uint64_t)(c.h - 1) * c.org_stride + c.w <= org_elems;
        if (!ok) {
            me_set_error("me_affine_me_gradient: CU size, sub-block size, AMVR index, reference or original out of range",
                         hipErrorInvalidValue, __FILE__, __LINE__);
            return -1;
        }
    }
    hipStream_t st = (hipStream_t)stream;
    me_affine_me_cu_t *d_cus = nullptr;
    pel *d_org = nullptr;
    ME_CHECK(hipMalloc((void **)&d_cus, (size_t)n * sizeof(me_affine_me_cu_t)));
    hipError_t e = hipMalloc((void **)&d_org, (size_t)org_elems * sizeof(pel));
    if (e == hipSuccess) e = hipMemcpyAsync(d_cus, cus, (size_t)n * sizeof(me_affine_me_cu_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_org, org, (size_t)org_elems * sizeof(pel), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        AffMeArgs a;
        a.cus = d_cus;
        a.org = d_org;
        a.plane0 = r->planes + (size_t)ME_PAD * r->stride + ME_PAD;
        a.stride = r->stride;
        a.pic_w = r->w;
        a.pic_h = r->h;
        a.lambda = lambda_mv;
        hipLaunchKernelGGL(k_affine_me, dim3(n), dim3(256), 0, st, a);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(cus, d_cus, (size_t)n * sizeof(me_affine_me_cu_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_cus);
    (void)hipFree(d_org);
    if (e != hipSuccess) {
        me_set_error("me_affine_me_gradient", e, __FILE__, __LINE__);
        return -1;
    }
    return 0;
}
