// Sub-pel refinement of every PU of a set of CTUs (gfx950):
// me_sub_pel_search (me.c:366-500) with cost_type 0 (SATD), bi 0, AMVR 0,
// then the final cost of me_search_tz (me.c:560-567).
//
// This file is compiled once per group of com_had tile types (ME_SP_PART =
// 0: 4x4, 8x8, 8x4, 4x8 tiles; 1: 16x8; 2: 8x16) so that the instances build
// in parallel.
#include "had.h"

#ifndef ME_SP_PART
#define ME_SP_PART 0
#endif

// Sub-pel search of one shape.  A unit is one (CTU, PU of this shape,
// reference); a.lpu lanes share a unit: one lane per com_had tile
// (com_cost.c:719-811), two for 16x8 / 8x16 tiles (sums and differences).
// The 29 shapes have 1..64 tiles, or 256 (128x128 in 8x8 tiles).  Per search
// position the lanes' tile SATDs are summed across the unit's lanes, so
// every lane of a unit holds the same costs and makes the same (strict '<',
// reference order) decisions.  A unit of more than 64 lanes is a whole
// workgroup.
template <int TT>
__device__ __forceinline__ void subpel_body(const SubArgs &a)
{
    // grid (per_ctu * lpu / blockDim, jobs, refs): a workgroup works on one
    // (job, reference), its lanes on PUs i of that CTU, a.lpu lanes per PU
    const int LPU = a.lpu, lg = __builtin_ctz((unsigned)LPU);   // LPU: a power of two
    const int li = threadIdx.x & (LPU - 1);
    const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> lg);
    const int job = blockIdx.y, r = blockIdx.z;
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    const int cols = 128 / a.pw;
    const int x = (ctu % P.ctu_cols) * 128 + (i % cols) * a.pw, y = (ctu / P.ctu_cols) * 128 + (i / cols) * a.ph;
    const size_t slot = ((size_t)job * a.npu + a.base + i) * a.nref + r;
    const bool inside = x + a.pw <= P.pic_w && y + a.ph <= P.pic_h;
    const u64 key = a.keys[slot];
    me_pu_result_t out;
    out.mv_int[0] = out.mv_int[1] = out.mv_sub[0] = out.mv_sub[1] = 0;
    out.cost_int = out.cost_sub = ~0ull;
    out.cost_final = 0xFFFFFFFFu;
    out.valid = 0;
    if (!inside || key == ~0ull) {
        if (li == 0) a.res[slot] = out;
        return;
    }
    const u64 k = key - ((u64)KEY_BIAS_HI << 32);
    const int span = 2 * a.R + 1;
    const int idx = (int)(k & ((1ull << a.IB) - 1));
    const int dx = idx % span - a.R, dy = idx / span - a.R;
    out.mv_int[0] = (s16)dx;
    out.mv_int[1] = (s16)dy;
    out.cost_int = k >> a.IB;

    constexpr bool SPLIT = TT == 3 || TT == 5;
    const pel *org = P.cur + (size_t)y * P.pic_w + x;
    const int t = SPLIT ? li >> 1 : li;
    const int tx = (t % a.tpr) * a.tw, ty = (t / a.tpr) * a.th;
    const int toff_o = ty * P.pic_w + tx, toff_p = ty * P.ref_stride + tx;
    const int red_w = LPU < 64 ? LPU : 64;
    auto had = [&](const pel *pred) -> u32 {
        int v;
        if (SPLIT) {
            // the two lanes of a tile hold its sum and difference halves; the
            // tile is normalised (com_cost.c:467, 586) before tiles are summed
            const int hs = had_half<TT>(org + toff_o, P.pic_w, pred + toff_p, P.ref_stride, li & 1);
            const int s = hs + __builtin_amdgcn_mov_dpp(hs, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
            const double SQRT128 = 0x1.6a09e667f3bcdp+3;
            v = (li & 1) ? 0 : (int)((double)s / SQRT128 * 2.0);
        } else {
            v = satd_t<TT>(org + toff_o, P.pic_w, pred + toff_p, P.ref_stride);
        }
        v = group_sum(v, red_w);
        if (LPU > 64) {
            __shared__ int red[4];
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
            __syncthreads();
            v = 0;
            for (int k = 0; k < (LPU >> 6); k++) v += red[k];
            __syncthreads();
        }
        return (u32)(v >> (BIT_DEPTH - 8));   // com_had: sum >> (bd - 8)
    };
    const u8 *tb = a.tab;
    const int cmx = dx * 4, cmy = dy * 4;
    // 17 positions, one SATD call site: 0 = the centre, read from the
    // integer plane without clipping (me.c:400); 1..8 the half-pel points
    // around it (me.c:415-417); 9..16 the quarter-pel points around the best
    // half-pel position (me.c:460-462).  Strict '<' in this order.
    u64 best = 0;
    int bx = cmx, by = cmy, hx = cmx, hy = cmy;
#pragma unroll 1
    for (int p = 0; p < 17; p++) {
        int mx = cmx, my = cmy;
        const pel *pred;
        if (p == 0) {
            pred = plane_ptr(P, r, 0, x + dx, y + dy);
        } else {
            if (p == 9) { hx = bx; hy = by; }
            if (p < 9) {
                mx = cmx + 2 * pat(ME_SUBPEL_HALF_X, p - 1);
                my = cmy + 2 * pat(ME_SUBPEL_HALF_Y, p - 1);
            } else {
                mx = hx + pat(ME_SUBPEL_QUARTER_X, p - 9);
                my = hy + pat(ME_SUBPEL_QUARTER_Y, p - 9);
            }
            pred = qpel_ptr(P, r, mx + (x << 2), my + (y << 2));
        }
        const u64 c = (u64)(a.lambda * (u32)(tb[mx] + tb[my])) + ((u64)had(pred) << 16);
        if (p == 0 || c < best) { best = c; bx = mx; by = my; }
    }
    if (li == 0) {
        out.mv_sub[0] = (s16)bx;
        out.mv_sub[1] = (s16)by;
        out.cost_sub = best;
        // tbl_refi_bits[n][i] of the reference's list: n its size, i its index
        // in it (tables.c:52-70); AMVR 0 adds 0 + 1 + 2 (me.c:560-561)
        const int n = r < a.n0 ? a.n0 : a.nref - a.n0, ri = r < a.n0 ? r : r - a.n0;
        const int rb = n <= 1 ? 0 : (ri < n - 1 ? ri + 1 : n - 1);
        const u64 cf = best + (u64)(a.lambda * (u32)(rb + 3));
        out.cost_final = (u32)((cf + (1u << 15)) >> 16);
        out.valid = 1;
        a.res[slot] = out;
    }
}

template <int TT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_subpel(SubArgs a)
{
    subpel_body<TT>(a);
}

#if ME_SP_PART == 0
subpel_fn me_subpel_kernel_small(int tt)
{
    switch (tt) {
    case 0: return k_subpel<0>;
    case 1: return k_subpel<1>;
    case 2: return k_subpel<2>;
    case 4: return k_subpel<4>;
    default: return nullptr;
    }
}
#elif ME_SP_PART == 1
subpel_fn me_subpel_kernel_16x8(int tt) { return tt == 3 ? k_subpel<3> : nullptr; }
#else
subpel_fn me_subpel_kernel_8x16(int tt) { return tt == 5 ? k_subpel<5> : nullptr; }
#endif
