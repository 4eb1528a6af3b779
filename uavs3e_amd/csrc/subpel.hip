// Sub-pel refinement of every PU of a set of CTUs (gfx950):
// me_sub_pel_search (me.c:366-500) with cost_type 0 (SATD), bi 0, AMVR 0,
// then the final cost of me_search_tz (me.c:560-567).
//
// One launch covers every searched shape: a workgroup looks up its shape
// and runs the instance of that shape's com_had tile type.
#include "had.h"

#include <string.h>

// Sub-pel search of one shape.  A unit is one (CTU, PU of this shape,
// reference).  Its lanes form S.G position groups of S.lpu lanes: in a group
// one lane per com_had tile (com_cost.c:719-811), two for 16x8 / 8x16 tiles
// (sums and differences); the groups evaluate different search positions at
// the same time.  The 29 shapes have 1..64 tiles, or 256 (128x128 in 8x8
// tiles).  Per position the tile SATDs are summed across the group's lanes
// (through LDS when a group spans several waves), and the groups' best keys
// (cost, position index) are merged through LDS, so every lane of a unit
// makes the same decisions.  Taking the smallest (cost, index) pair is the
// reference's strict '<' scan in its evaluation order.
template <int TT>
__device__ __forceinline__ void subpel_body(const SubArgs &a, const SubShape &S, int blk, int job, int r)
{
    __shared__ u64 gk[256];   // per (unit, group): best key of a phase
    __shared__ int red[4];    // per wave: SATD partial sums of groups > 64 lanes
    // the shape's blocks (per_ctu * lpu * G / 256 per job and reference): a
    // workgroup works on one (job, reference), its units on PUs i of that CTU
    const int LPU = S.lpu, lg = __builtin_ctz((unsigned)LPU);   // powers of two
    const int G = S.G, gg = __builtin_ctz((unsigned)G);
    const int tid = threadIdx.x;
    const int li = tid & (LPU - 1), g = (tid >> lg) & (G - 1), ul = tid >> (lg + gg);
    const int i = (int)((blk * 256 + tid) >> (lg + gg));
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    const int cols = 128 / S.pw;
    const int x = (ctu % P.ctu_cols) * 128 + (i % cols) * S.pw, y = (ctu / P.ctu_cols) * 128 + (i / cols) * S.ph;
    const size_t slot = ((size_t)job * a.npu + S.base + i) * a.nref + r;
    const bool inside = x + S.pw <= P.pic_w && y + S.ph <= P.pic_h;
    const u64 key = key_merged(a.kl, a.keys, slot, job, r, a.nref);
    const bool lead = li == 0 && g == 0;
    me_pu_result_t out;
    out.mv_int[0] = out.mv_int[1] = out.mv_sub[0] = out.mv_sub[1] = 0;
    out.cost_int = out.cost_sub = ~0ull;
    out.cost_final = 0xFFFFFFFFu;
    out.valid = 0;
    // Units without a search result take no sample reads; they still pass
    // every barrier of the workgroup.
    const bool active = inside && key < KEY_INIT;
    const bool sync = G > 1 || LPU > 64;
    if (!active) {
        if (lead) a.res[slot] = out;
        if (!sync) return;
    }
    const u64 k = active ? key - ((u64)KEY_BIAS_HI << 32) : 0;
    const int span = 2 * a.R + 1;
    const int idx = (int)(k & ((1ull << a.IB) - 1));
    const int dx = idx % span - a.R, dy = idx / span - a.R;
    out.mv_int[0] = (s16)dx;
    out.mv_int[1] = (s16)dy;
    out.cost_int = k >> a.IB;

    constexpr bool SPLIT = TT == 3 || TT == 5;
    const pel *org = P.cur + (size_t)y * P.pic_w + x;
    const int t = SPLIT ? li >> 1 : li;
    const int tx = (t % S.tpr) * S.tw, ty = (t / S.tpr) * S.th;
    const int toff_o = ty * P.pic_w + tx, toff_p = ty * P.ref_stride + tx;
    const int red_w = LPU < 64 ? LPU : 64;
    // the lane's current tile as 16-bit pairs, made once for all 17 positions
    s16x2 opairs[org_pairs_n<TT>()];
    if (active) org_pairs<TT>(org + toff_o, P.pic_w, li & 1, opairs);
    // com_had of the unit at pred, in every lane of the group; run = 0 skips
    // the sample reads (the sum is then 0) but not the barriers
    auto had = [&](const pel *pred, bool run) -> u32 {
        int v = 0;
        if (run) {
            if (SPLIT) {
                // the two lanes of a tile hold its sum and difference halves;
                // the tile is normalised (com_cost.c:467, 586) before tiles
                // are summed
                const int hs = had_pre<TT>(opairs, pred + toff_p, P.ref_stride, li & 1);
                const int s = hs + __builtin_amdgcn_mov_dpp(hs, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
                v = (li & 1) ? 0 : me_satd_norm128(s);
            } else {
                v = satd_pre<TT>(opairs, pred + toff_p, P.ref_stride);
            }
        }
        v = group_sum(v, red_w);
        if (LPU > 64) {
            if ((tid & 63) == 0) red[tid >> 6] = v;
            __syncthreads();
            const int w0 = (tid >> 6) & ~((LPU >> 6) - 1);   // first wave of the group
            v = 0;
            for (int q = 0; q < (LPU >> 6); q++) v += red[w0 + q];
            __syncthreads();
        }
        return (u32)(v >> (BIT_DEPTH - 8));   // com_had: sum >> (bd - 8)
    };
    const u8 *tb = a.tab;
    const int cmx = dx * 4, cmy = dy * 4;
    // Position p of the 17: 0 = the centre, read from the integer plane
    // without clipping (me.c:400); 1..8 the half-pel points around it
    // (me.c:415-417); 9..16 the quarter-pel points around the best of 0..8
    // (me.c:460-462).  A key is (cost << 5) | p.
    auto phase = [&](int p0, int n, int hx, int hy) -> u64 {
        u64 best = ~0ull;
        for (int s0 = 0; s0 < n; s0 += G) {
            const bool valid = s0 + g < n;
            const int p = p0 + (valid ? s0 + g : 0);
            int mx = cmx, my = cmy;
            const pel *pred;
            if (p == 0) {
                pred = plane_ptr(P, r, 0, x + dx, y + dy);
            } else {
                if (p < 9) {
                    mx = cmx + 2 * pat(ME_SUBPEL_HALF_X, p - 1);
                    my = cmy + 2 * pat(ME_SUBPEL_HALF_Y, p - 1);
                } else {
                    mx = hx + pat(ME_SUBPEL_QUARTER_X, p - 9);
                    my = hy + pat(ME_SUBPEL_QUARTER_Y, p - 9);
                }
                pred = qpel_ptr(P, r, mx + (x << 2), my + (y << 2));
            }
            const u32 h = had(pred, active && valid);
            const u64 c = (u64)(a.lambda * (u32)(tb[mx] + tb[my])) + ((u64)h << 16);
            const u64 kk = (c << 5) | (u64)p;
            if (valid && kk < best) best = kk;
        }
        if (G > 1) {
            if (li == 0) gk[ul * G + g] = best;
            __syncthreads();
            best = gk[ul * G];
            for (int q = 1; q < G; q++) best = gk[ul * G + q] < best ? gk[ul * G + q] : best;
            __syncthreads();
        }
        return best;
    };
    auto pos = [&](int p, int hx, int hy, int &mx, int &my) {
        mx = cmx;
        my = cmy;
        if (p >= 1 && p < 9) {
            mx = cmx + 2 * pat(ME_SUBPEL_HALF_X, p - 1);
            my = cmy + 2 * pat(ME_SUBPEL_HALF_Y, p - 1);
        } else if (p >= 9) {
            mx = hx + pat(ME_SUBPEL_QUARTER_X, p - 9);
            my = hy + pat(ME_SUBPEL_QUARTER_Y, p - 9);
        }
    };
    const u64 k1 = phase(0, 9, cmx, cmy);
    int hx, hy;
    pos((int)(k1 & 31), cmx, cmy, hx, hy);
    const u64 k2 = phase(9, 8, hx, hy);
    const u64 kb = k2 < k1 ? k2 : k1;
    int bx, by;
    pos((int)(kb & 31), hx, hy, bx, by);
    const u64 best = kb >> 5;
    if (lead && active) {
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

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_subpel(SubArgs a)
{
    if (a.js.n_jobs == 0) return;   // a warm-up launch (me_subpel_preload)
    // workgroup w = (block, job, reference), block fastest, in XCD order: the
    // blocks of one (job, reference) read the same plane neighbourhoods
    const unsigned nx = gridDim.x, ny = gridDim.y;
    const unsigned w = me_xcd_order(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
    const int bxi = (int)(w % nx), job = (int)((w / nx) % ny), r = (int)(w / (nx * ny));
    // the shape and its parameters are uniform across the workgroup
    const int si = __builtin_amdgcn_readfirstlane(a.blk_shape[bxi]);
    const SubShape S = a.shp[si];
    const int blk = bxi - S.blk0;
    switch (S.tt) {
    case 0: subpel_body<0>(a, S, blk, job, r); break;
    case 1: subpel_body<1>(a, S, blk, job, r); break;
    case 2: subpel_body<2>(a, S, blk, job, r); break;
    case 3: subpel_body<3>(a, S, blk, job, r); break;
    case 4: subpel_body<4>(a, S, blk, job, r); break;
    default: subpel_body<5>(a, S, blk, job, r); break;
    }
}

int me_subpel_launch(const SubArgs &a, int blocks, int jobs, int refs, hipStream_t st)
{
    hipLaunchKernelGGL(k_subpel, dim3(blocks, jobs, refs), dim3(256), 0, st, a);
    ME_CHECK(hipGetLastError());
    return 0;
}

int me_subpel_preload(hipStream_t st)
{
    // An empty launch: besides loading the code it makes the runtime set up
    // the kernel's private (spill) memory for st's queue, which otherwise
    // happens at the first real launch and stalls it for milliseconds.
    SubArgs a;
    memset(&a, 0, sizeof(a));
    hipLaunchKernelGGL(k_subpel, dim3(1), dim3(256), 0, st, a);
    ME_CHECK(hipGetLastError());
    return 0;
}
