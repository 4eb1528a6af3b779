// The reference encoder's own integer + sub-pel search, me_search_tz
// (me.c:529-568), for every PU of a set of CTUs (gfx950): uni-prediction
// (bi = 0) from a per-PU MVP, any AMVR precision and both sub-pel cost types.
//
//   window        get_max_search_range + com_ipel_range_rounding   me.c:48-55, 502-527
//   diamond       centre, core square, multilayer diamond rounds   me.c:199-355
//   raster        get_raster_range, rounds of halving steps        me.c:65-197
//   refine        diamond with max step 8                          me.c:546-551
//   sub-pel       me_sub_pel_search (AMVR 0: half + quarter,       me.c:366-500
//                 AMVR 1: half; cost type 1: SAD, then SATD)
//   final cost    + lambda*(refi_bits+mvr+(mvr<4)+2), (c+2^15)>>16 me.c:560-567
//
// One group of lanes per (PU, reference): 4 lanes for the smallest shapes up
// to a whole wave from 32x32 on (area / 16, a power of two).  The search
// decisions are made on values every lane of the group holds (control flow
// is uniform within a group; groups of one wave may diverge); each
// candidate's SAD or com_had is computed by the group's lanes together and
// summed across them with DPP / swizzle steps that stay inside the group.
// The running best is replaced only by a strictly cheaper candidate in the
// reference's evaluation order, exactly as the sequential loops do.
#include "had.h"

#include <string.h>
#include <algorithm>

// Steps of a block walk whose loads are issued together (TZ_U), and k_tz's
// waves per SIMD (TZ_WPE): 2 and 3 measured fastest (4 steps at 2 waves per
// SIMD: 6.8 against 5.6 ms per 1080p all-PU frame).
#ifndef TZ_U
#define TZ_U 2
#endif
#ifndef TZ_WPE
#define TZ_WPE 3
#endif

// compiled once per group of com_had tile types, like subpel.hip (ME_TZ_PART
// 0: 4x4, 8x8, 8x4, 4x8 tiles and the launcher; 1: 16x8; 2: 8x16)
#ifndef ME_TZ_PART
#define ME_TZ_PART 0
#endif

// One shape of a grouped search launch: its geometry, lanes per (PU,
// reference) and the first workgroup (grid x) of its range.
struct TzShape {
    int pw, ph, base, per_ctu, lpu, orgbi_off, blk0;
};

struct TzArgs {
    JobSet js;
    const s16 *mvp;          // [job][npu][nref][2], quarter pel
    const s16 *mvi_prev;     // previous integer MV per (job, pu, ref) for the adaptive raster range, or null
    me_pu_result_t *res;
    const u8 *tab;           // centred MV-bits table
    int tab_off;
    u32 lambda;
    int npu, nref, n0;       // n0: list-0 size
    int lpu;                 // lanes per (PU, reference): a power of two, 4..64
    int pw, ph, base, per_ctu, tt, tw, th;
    int gop, msr, adaptive, mvr, cost_type;
    int pad_win;             // inter_search_t.is_padding: windows may leave the picture
    long long cur_poc, ref_poc[ME_MAX_REFS];
    // bi = 1 (analyze_bi): per PU, the list searched is the one not fixed in
    // its BiState, the target is its bi-prediction original and the search
    // starts from (and writes back) mv_scale
    int bi;
    const pel *orgbi;        // [job][selected shape][16384]
    int orgbi_off, nsel;     // this shape's rank * 16384; selected shapes
    s16 *mv_scale;           // [job][npu][nref][2], quarter pel
    const struct BiState *bst;   // [job][npu]
    const me_bi_pu_t *bpu;       // [job][npu]
    // k_tz: the shapes of the launch (one com_had tile type), largest PUs
    // first; workgroup x belongs to the last shape whose blk0 <= x
    TzShape shp[ME_NSHAPES];
    int nshp;
};

// analyze_bi's running state of one PU (inter.c:986-1033)
struct BiState {
    u64 best;                // best_mecost
    int fixed;               // the list whose prediction makes the target
    int done;                // no change in the last round, or not a bi PU
};

// Diamond patterns of search_diamond (me.c:330-333): 8 points of radius 2,
// and the 16-point ring of radius 4 in the same rotation.
__constant__ static const s8 TZ_DIA8[8][2] = {{-2, 0}, {-1, 1}, {0, 2}, {1, 1}, {2, 0}, {1, -1}, {0, -2}, {-1, -1}};
__constant__ static const s8 TZ_DIA16[16][2] = {{-4, 0}, {-3, 1}, {-2, 2}, {-1, 3}, {0, 4}, {1, 3}, {2, 2}, {3, 1},
                                                {4, 0}, {3, -1}, {2, -2}, {1, -3}, {0, -4}, {-1, -3}, {-2, -2}, {-3, -1}};


typedef const __attribute__((address_space(1))) u32 gu32_t;

// SAD of a w x h block by the lpu lanes of a group; every lane gets the sum.
// A lane takes 4 horizontally adjacent samples at a time: the original's
// are aligned (PU x and the picture width are multiples of 4), the
// prediction's are realigned from two aligned dwords with v_alignbyte.
__device__ __forceinline__ u32 wave_sad(const pel *org, int so, const pel *pred, int sp, int w, int h, int lpu)
{
    const int li = threadIdx.x & (lpu - 1), q = w >> 2, lq = __builtin_ctz((unsigned)q);
    u32 s = 0;
    for (int i = li; i < q * h; i += lpu) {
        const int xx = (i & (q - 1)) << 2, yy = i >> lq;
        const pel *o = org + yy * so + xx, *p = pred + yy * sp + xx;
        const uintptr_t ad = (uintptr_t)p;
        const u32 *pq = (const u32 *)(ad & ~(uintptr_t)3);
        const u32 sh = (u32)(ad & 3);
#if BIT_DEPTH == 8
        const u32 pw = __builtin_amdgcn_alignbyte(pq[1], pq[0], sh);
        s = __builtin_amdgcn_sad_u8(pw, *(const u32 *)o, s);
#else
        const u32 p0 = __builtin_amdgcn_alignbyte(pq[1], pq[0], sh), p1 = __builtin_amdgcn_alignbyte(pq[2], pq[1], sh);
        const uint2 ow = *(const uint2 *)o;
        s = __builtin_amdgcn_sad_u16(p0, ow.x, __builtin_amdgcn_sad_u16(p1, ow.y, s));
#endif
    }
    return (u32)group_sum((int)s, lpu);
}

// A 64-bit address every lane of the wave holds, moved to scalar registers
// (the search's picture and plane bases are the workgroup's own).
template <typename T> __device__ __forceinline__ T *uniform_ptr(T *p)
{
    const u64 v = (u64)(uintptr_t)p;
    return (T *)(uintptr_t)(((u64)(u32)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
                            (u32)__builtin_amdgcn_readfirstlane((int)(u32)v));
}
typedef const __attribute__((address_space(1))) unsigned char gbyte_t;
__device__ __forceinline__ u32 ld32(gbyte_t *base, u32 off) { return *(const gu32_t *)(base + off); }
// NW consecutive dwords in one load instruction (dword-aligned)
typedef u32 u32x2u __attribute__((ext_vector_type(2), aligned(4)));
typedef u32 u32x3u __attribute__((ext_vector_type(3), aligned(4)));
template <int NW> __device__ __forceinline__ void ldn(gbyte_t *base, u32 off, u32 *d)
{
    if (NW == 1) {
        d[0] = ld32(base, off);
    } else if (NW == 2) {
        const u32x2u v = *(const __attribute__((address_space(1))) u32x2u *)(base + off);
        d[0] = v.x;
        d[1] = v.y;
    } else {
        const u32x3u v = *(const __attribute__((address_space(1))) u32x3u *)(base + off);
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
    }
}

// SADs of NB candidate blocks against the same original, by the lpu lanes of
// a group: one walk over the block's samples that accumulates every
// candidate at each step, so that the candidates' loads are in flight
// together (a search evaluated one candidate at a time waits for each
// candidate's loads in turn).  The original is at byte offset ooff from ob,
// candidate c at poff[c] from pb: both bases are the workgroup's (scalar
// registers), so each load takes a 32-bit lane offset instead of a 64-bit
// address.  out[c] is in every lane of the group.
template <int NB, bool UNI>
__device__ __forceinline__ void wave_sad_n_off(const pel *ob_, u32 ooff, int so, const pel *pb_, const u32 *poff, int sp,
                                           int w, int h, int lpu, u32 *out)
{
    // UNI: the plane base is the workgroup's; else (bi = 1, the lanes' PUs may
    // search different lists) each lane keeps its own 64-bit base
    gbyte_t *ob = (gbyte_t *)uniform_ptr(ob_), *pb = (gbyte_t *)(UNI ? uniform_ptr(pb_) : pb_);
    const int li = threadIdx.x & (lpu - 1), q = w >> 2, lq = __builtin_ctz((unsigned)q);
    u32 s[NB];
#pragma unroll
    for (int c = 0; c < NB; c++) s[c] = 0;
    // TZ_U steps of the walk at a time: every load of the chunk is issued
    // before the first is used (the scheduler would otherwise wait for each
    // candidate's pair in turn).  A lane's step count, q * h / lpu, is the
    // same for every lane; steps past it repeat the chunk's first reads and
    // are not summed.
    constexpr int NW = BIT_DEPTH == 8 ? 2 : 3, OW = BIT_DEPTH == 8 ? 1 : 2;
    const int total = q * h;
    for (int i0 = li; i0 < total; i0 += TZ_U * lpu) {
        u32 qd[TZ_U][NB][NW], sh[TZ_U][NB], ow[TZ_U][OW];
#pragma unroll
        for (int t = 0; t < TZ_U; t++) {
            const int i = i0 + t * lpu < total ? i0 + t * lpu : i0;
            const int xx = (i & (q - 1)) << 2, yy = i >> lq;
            const u32 oo = ooff + (u32)((yy * so + xx) * (int)sizeof(pel)), po = (u32)((yy * sp + xx) * (int)sizeof(pel));
#pragma unroll
            for (int c = 0; c < NB; c++) {
                const u32 ad = poff[c] + po;
                sh[t][c] = ad & 3u;
                ldn<NW>(pb, ad & ~3u, qd[t][c]);
            }
            ldn<OW>(ob, oo, ow[t]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < TZ_U; t++) {
            if (t > 0 && i0 + t * lpu >= total) break;
#pragma unroll
            for (int c = 0; c < NB; c++) {
#if BIT_DEPTH == 8
                s[c] = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(qd[t][c][1], qd[t][c][0], sh[t][c]), ow[t][0], s[c]);
#else
                const u32 p0 = __builtin_amdgcn_alignbyte(qd[t][c][1], qd[t][c][0], sh[t][c]);
                const u32 p1 = __builtin_amdgcn_alignbyte(qd[t][c][2], qd[t][c][1], sh[t][c]);
                s[c] = __builtin_amdgcn_sad_u16(p0, ow[t][0], __builtin_amdgcn_sad_u16(p1, ow[t][1], s[c]));
#endif
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NB; c++) out[c] = (u32)group_sum((int)s[c], lpu);
}

// Candidates per batch of the integer searches: 4 keeps every instance of
// k_tz at 3 waves per SIMD (at most 168 VGPRs, amdgpu_waves_per_eu(3));
// each candidate in flight costs registers.
template <int TT> struct TzBatch { static constexpr int NB = 4; };

// One normalised com_had tile of type TT (com_cost.c:259, 340, 467, 586,
// 649, 708), done by one lane.
template <int TT>
__device__ __forceinline__ int satd_tile_fn(const pel *org, int so, const pel *cur, int sc)
{
    if (TT == 3 || TT == 5) {
        return me_satd_norm128(had_half<TT>(org, so, cur, sc, 0) + had_half<TT>(org, so, cur, sc, 1));
    }
    return satd_t<TT>(org, so, cur, sc);
}

// com_had (com_cost.c:712-812): the PU's tiles spread over the lanes; the
// tile sum is shifted by bd - 8.  One out-of-line copy (it is called from
// four places of the sub-pel search).
template <int TT>
__device__ __noinline__ u32 wave_had(int lpu, int pw, int ph, const pel *org, int so, const pel *pred, int sp)
{
    // tile size of the type (had_tiling)
    constexpr int tw = TT == 0 ? 4 : (TT == 1 ? 8 : (TT == 2 ? 8 : (TT == 3 ? 16 : (TT == 4 ? 4 : 8))));
    constexpr int th = TT == 0 ? 4 : (TT == 1 ? 8 : (TT == 2 ? 4 : (TT == 3 ? 8 : (TT == 4 ? 8 : 16))));
    const int li = threadIdx.x & (lpu - 1), tpr = pw / tw, T = tpr * (ph / th);
    int s = 0;
    if (TT == 3 || TT == 5) {
        // a 16x8 / 8x16 tile on a lane pair: the 16-point transform is the
        // 8-point transform of the half sums (even lane) and of the half
        // differences (odd lane); the pair's sum is normalised before the
        // tiles are summed (com_cost.c:467, 586).  lpu is even, so both
        // lanes of a pair run the same iterations.
        for (int t2 = li; t2 < 2 * T; t2 += lpu) {
            const int t = t2 >> 1, tx = (t % tpr) * tw, ty = (t / tpr) * th;
            const int hs = had_half<TT>(org + ty * so + tx, so, pred + ty * sp + tx, sp, t2 & 1);
            const int v = hs + __builtin_amdgcn_mov_dpp(hs, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
            s += (t2 & 1) ? 0 : me_satd_norm128(v);
        }
    } else {
        for (int t = li; t < T; t += lpu) {
            const int tx = (t % tpr) * tw, ty = (t / tpr) * th;
            s += satd_tile_fn<TT>(org + ty * so + tx, so, pred + ty * sp + tx, sp);
        }
    }
    return (u32)(group_sum(s, lpu) >> (BIT_DEPTH - 8));
}

struct TzUnit {
    const TzArgs *a;
    const PicDesc *P;
    int r, x, y, w, h, lpu;
    const pel *org;
    int so;                  // org stride
    const pel *ob;           // the workgroup's original picture (or bi targets) and ...
    u32 ooff;                // ... org's byte offset from it
    const pel *pb;           // reference r's integer plane set
    int mvpx, mvpy;
    int lo[2], hi[2];
    u64 best;
    int mv[2];
    int shift;

    __device__ __forceinline__ u32 ipel_bits(int mx, int my) const
    {
        const int m = a->mvr;
        return (u32)a->tab[((mx << 2) >> m) - (mvpx >> m)] + (u32)a->tab[((my << 2) >> m) - (mvpy >> m)];
    }
    __device__ __forceinline__ u32 qpel_bits(int mx, int my) const
    {
        const int m = a->mvr;
        return (u32)a->tab[(mx >> m) - (mvpx >> m)] + (u32)a->tab[(my >> m) - (mvpy >> m)];
    }
    // The costs of up to NB integer candidates (mx[c], my[c]), c < n,
    // computed together; a candidate's cost is that of me.c:158-189.
    // Entries c >= n read the first candidate's block and are not used.
    template <int NB>
    __device__ __forceinline__ void eval_ipel(const int *mx, const int *my, int n, u64 *cost) const
    {
        u32 sad[NB], poff[NB];
#pragma unroll
        for (int c = 0; c < NB; c++) {
            const int k = c < n ? c : 0;
            poff[c] = (u32)(((y + my[k] + ME_PAD) * P->ref_stride + (x + mx[k] + ME_PAD)) * (int)sizeof(pel));
        }
        if (!a->bi) wave_sad_n_off<NB, true>(ob, ooff, so, pb, poff, P->ref_stride, w, h, lpu, sad);
        else wave_sad_n_off<NB, false>(ob, ooff, so, pb, poff, P->ref_stride, w, h, lpu, sad);
#pragma unroll
        for (int c = 0; c < NB; c++) {
            const int k = c < n ? c : 0;
            cost[c] = ((u64)sad[c] << shift) + (u64)(ipel_bits(mx[k], my[k]) * a->lambda);
        }
    }
};

// Scans a batch of evaluated candidates in the reference's order: a
// strictly cheaper valid one becomes the best (me.c:158-189).  Returns the
// tag of the batch's last improvement, or `prev` without one.
template <int NB>
__device__ __forceinline__ int scan_batch(TzUnit &u, const int *mx, const int *my, const u64 *cost, const bool *ok,
                                          const int *tag, int prev)
{
    int t = prev;
#pragma unroll
    for (int c = 0; c < NB; c++)
        if (ok[c] && cost[c] < u.best) {
            u.best = cost[c];
            u.mv[0] = mx[c];
            u.mv[1] = my[c];
            t = tag[c];
        }
    return t;
}

// search_diamond (me.c:199-355); returns the step of the last improvement
// (0: centre, 2: core square, else the diamond step).  With bi the core
// square spans the distance to the rounded MVP (2..5, 6 above integer AMVR)
// and the multilayer diamond is skipped (me.c:221-231, 317-319).
// Every candidate of the call lies around the entry MV, fixed before any is
// evaluated: the centre, the core square in raster order, then each
// multilayer round's points in table order (those outside the window are
// not candidates).  They are evaluated NB at a time and scanned in that
// order, which keeps the reference's first strict minimum.
template <int NB>
__device__ __noinline__ int tz_diamond(TzUnit &ur, int max_step)
{
    // work on a copy that stays in registers (the caller's unit is in
    // private memory, as it is passed by reference); the result goes back
    TzUnit u = ur;
    const int mvr = u.a->mvr;
    int core = 2, cstep = 1;
    const int cx = u.mv[0], cy = u.mv[1];
    if (u.a->bi) {
        const int dx = abs(((u.mvpx + 2) >> 2) - cx), dy = abs(((u.mvpy + 2) >> 2) - cy);
        core = me_clip3(2, mvr > 2 ? 6 : 5, (dx > dy ? dx : dy) >> (mvr > 2 ? mvr - 2 : 0));
    }
    if (mvr > 2) { core <<= mvr - 2; cstep <<= mvr - 2; }
    int x0 = cx - core, y0 = cy - core, x1 = cx + core, y1 = cy + core;
    while (x0 <= u.lo[0]) x0 += cstep;
    while (y0 <= u.lo[1]) y0 += cstep;
    while (x1 >= u.hi[0]) x1 -= cstep;
    while (y1 >= u.hi[1]) y1 -= cstep;
    const int nx = max_step && x1 >= x0 ? (x1 - x0) / cstep + 1 : 0;
    const int ny = max_step && y1 >= y0 ? (y1 - y0) / cstep + 1 : 0;
    const int ncore = nx * ny;
    // multilayer diamonds: 4 points at step 4, 8 at step 8, 16 beyond
    // (me.c:330-333), scaled by the AMVR precision
    int nml = 0;
    if (!u.a->bi)
        for (int step = 4, round = 0; step < max_step; step <<= 1, round++) nml += round == 0 ? 4 : (round == 1 ? 8 : 16);
    const int scale = mvr - 2 > 0 ? mvr - 2 : 0;
    const int total = 1 + ncore + nml;
    int best_step = 0;
    for (int b0 = 0; b0 < total; b0 += NB) {
        int mx[NB], my[NB], tag[NB];
        bool ok[NB];
#pragma unroll
        for (int c = 0; c < NB; c++) {
            int k = b0 + c;
            mx[c] = cx;
            my[c] = cy;
            tag[c] = 0;
            ok[c] = k < total;
            if (k == 0 || k >= total) continue;
            k -= 1;
            if (k < ncore) {
                const int qy = k / nx;
                mx[c] = x0 + (k - qy * nx) * cstep;
                my[c] = y0 + qy * cstep;
                tag[c] = 2;
                continue;
            }
            k -= ncore;
            const int round = k < 4 ? 0 : (k < 12 ? 1 : 2 + (k - 12) / 16);
            const int i = round == 0 ? 2 * k : (round == 1 ? k - 4 : (k - 12) % 16);
            const int step = 4 << round, big = round < 2 ? 0 : 1;
            const int ss = (step << scale) >> 1;
            const int dx = big ? TZ_DIA16[i][0] : TZ_DIA8[i][0], dy = big ? TZ_DIA16[i][1] : TZ_DIA8[i][1];
            const int px = cx + (s16)(ss >> big) * dx, py = cy + (s16)(ss >> big) * dy;
            tag[c] = step;
            ok[c] = px <= u.hi[0] && px >= u.lo[0] && py <= u.hi[1] && py >= u.lo[1];
            if (ok[c]) { mx[c] = px; my[c] = py; }
        }
        u64 cost[NB];
        u.eval_ipel<NB>(mx, my, NB, cost);
        best_step = scan_batch<NB>(u, mx, my, cost, ok, tag, best_step);
    }
    ur.best = u.best;
    ur.mv[0] = u.mv[0];
    ur.mv[1] = u.mv[1];
    return best_step;
}

// search_raster (me.c:103-197) with get_raster_range (me.c:65-100).  Each
// round's grid is fixed by its start, so it is evaluated NB points at a
// time and scanned in raster order; a row is skipped while no candidate has
// a cost yet (the reference's s64 tmp_best never improves then, me.c:158).
template <int NB>
__device__ __noinline__ void tz_raster(TzUnit &ur, int refi, int mvi_prev_x, int mvi_prev_y)
{
    TzUnit u = ur;   // a register copy, as in tz_diamond
    const TzArgs &a = *u.a;
    const int mvr = a.mvr;
    const int step_base = mvr == 3 ? 6 : (mvr == 4 ? 8 : 5), step_min = mvr == 3 ? 2 : (mvr == 4 ? 4 : 1);
    const int mn = u.w < u.h ? u.w : u.h;
    int step = (step_base > (mn >> 1) ? step_base : (mn >> 1)) * (refi + 1);
    // get_raster_range
    long long dl = a.cur_poc - a.ref_poc[u.r];
    const int dist = (int)(dl < 0 ? -dl : dl);
    const int msr = a.msr >> (3 - mvr > 0 ? 3 - mvr : 0);
    int rxy = (msr * dist + (a.gop >> 1)) / a.gop;
    rxy = me_clip3(msr >> 2, msr, rxy);
    int rx = rxy, ry = rxy;
    if (mvr && a.adaptive) {
        rx = abs(mvi_prev_x - u.mv[0]) * 4;
        ry = abs(mvi_prev_y - u.mv[1]) * 2;
        if (rx < (msr >> 2)) rx = msr >> 2;
        if (ry < (msr >> 2)) ry = msr >> 2;
        if (rx < step * 8) rx = step * 8;
        if (ry < step * 4) ry = step * 4;
        if (rx > rxy) rx = rxy;
        if (ry > rxy) ry = rxy;
    }
    if (mvr > 2) { const int sh = mvr - 2; rx = (rx >> sh) << sh; ry = (ry >> sh) << sh; }
    const int r0x = me_clip3(u.lo[0], u.hi[0], u.mv[0] - (s16)rx), r1x = me_clip3(u.lo[0], u.hi[0], u.mv[0] + (s16)rx);
    const int r0y = me_clip3(u.lo[1], u.hi[1], u.mv[1] - (s16)ry), r1y = me_clip3(u.lo[1], u.hi[1], u.mv[1] + (s16)ry);
    for (int round = 0; step >= step_min; round++) {
        int x0, y0, x1, y1;
        if (round == 0) { x0 = r0x; y0 = r0y; x1 = r1x; y1 = r1y; }
        else {
            x0 = u.mv[0] - step > u.lo[0] ? u.mv[0] - step : u.lo[0];
            y0 = u.mv[1] - step > u.lo[1] ? u.mv[1] - step : u.lo[1];
            x1 = u.mv[0] + step < u.hi[0] ? u.mv[0] + step : u.hi[0];
            y1 = u.mv[1] + step < u.hi[1] ? u.mv[1] + step : u.hi[1];
        }
        const int nx = x1 >= x0 ? (x1 - x0) / step + 1 : 0, ny = y1 >= y0 ? (y1 - y0) / step + 1 : 0;
        const int total = nx * ny;
        bool skip_row = false;
        for (int b0 = 0; b0 < total; b0 += NB) {
            int mx[NB], my[NB];
#pragma unroll
            for (int c = 0; c < NB; c++) {
                const int k = b0 + c < total ? b0 + c : 0, qy = k / nx;
                mx[c] = x0 + (k - qy * nx) * step;
                my[c] = y0 + qy * step;
            }
            u64 cost[NB];
            u.eval_ipel<NB>(mx, my, NB, cost);
#pragma unroll
            for (int c = 0; c < NB; c++) {
                const int k = b0 + c;
                if (k >= total) continue;
                if (k % nx == 0) skip_row = u.best == ~0ull;
                if (!skip_row && cost[c] < u.best) {
                    u.best = cost[c];
                    u.mv[0] = mx[c];
                    u.mv[1] = my[c];
                }
            }
        }
        step >>= 1;
        if (mvr > 2) { const int sh = mvr - 2; step = (step >> sh) << sh; }
    }
    ur.best = u.best;
    ur.mv[0] = u.mv[0];
    ur.mv[1] = u.mv[1];
}

// me_sub_pel_search (me.c:366-500); u.mv holds the quarter-pel
// integer result on entry and the refined MV on return.
//
// With SATD costs (cost_type 0) a PU's lanes are split into G position
// groups of Lg lanes, Lg the lanes its com_had tiles take (one per tile,
// two per 16x8 / 8x16 tile): a group evaluates one candidate position, so G
// positions of a phase (the centre and the 8 half-pel points, then the 8
// quarter-pel points) are evaluated at a time and their costs then scanned
// in the reference's order with its strict '<' (with one lane per tile and
// lpu = area/16 lanes, three of four lanes would otherwise idle through every
// SATD).  SAD costs (cost_type 1) take every lane per candidate.
template <int TT>
__device__ __noinline__ u64 tz_subpel(TzUnit &ur)
{
    TzUnit u = ur;   // a register copy, as in tz_diamond
    const TzArgs &a = *u.a;
    const int shift = 16 - a.bi - (a.cost_type == 0 ? 0 : (BIT_DEPTH - 8));
    const int cx = u.mv[0], cy = u.mv[1];
    if (a.cost_type == 0) {
        constexpr int tw = TT == 0 ? 4 : (TT == 1 ? 8 : (TT == 2 ? 8 : (TT == 3 ? 16 : (TT == 4 ? 4 : 8))));
        constexpr int th = TT == 0 ? 4 : (TT == 1 ? 8 : (TT == 2 ? 4 : (TT == 3 ? 8 : (TT == 4 ? 8 : 16))));
        const int T = (u.w / tw) * (u.h / th), L = (TT == 3 || TT == 5) ? 2 * T : T;
        int G = L < u.lpu ? u.lpu / L : 1;
        if (G > 4) G = 4;
        const int Lg = G > 1 ? L : u.lpu;                         // lanes per position group
        const int lane = threadIdx.x & 63, base = lane & ~(u.lpu - 1);
        int g = (lane - base) / Lg;
        if (g >= G) g = 0;                                        // spare lanes repeat group 0's position
        // position k of a phase: the half phase's k = 0 is the centre (the
        // integer plane, no clip, me.c:400) and k = 1..8 its half-pel points;
        // the quarter phase's k = 0..7 are the quarter-pel points around
        // (ox, oy)
        auto pos = [&](bool half, int ox, int oy, int k, int &mx, int &my) {
            if (half) {
                mx = ox + (k ? 2 * pat(ME_SUBPEL_HALF_X, k - 1) : 0);
                my = oy + (k ? 2 * pat(ME_SUBPEL_HALF_Y, k - 1) : 0);
            } else {
                mx = ox + pat(ME_SUBPEL_QUARTER_X, k);
                my = oy + pat(ME_SUBPEL_QUARTER_Y, k);
            }
        };
        // Where every lane of a group holds at most one tile (the PU's tiles
        // fit its lanes), the lane's tile pairs of the original are made once
        // for all 17 positions and each position reads only its prediction
        // rows (had_pre / satd_pre); otherwise wave_had walks the tiles.
        constexpr bool SPLIT = TT == 3 || TT == 5;
        const bool pre = L <= Lg;
        const int li = lane - base - g * Lg;                      // lane within its group
        const bool tile = li < L;
        const int t = SPLIT ? li >> 1 : li, tpr = u.w / tw;
        const int tx = (t % tpr) * tw, ty = (t / tpr) * th;
        s16x2 opairs[org_pairs_n<TT>()];
        if (pre && tile) org_pairs<TT>(u.org + ty * u.so + tx, u.so, li & 1, opairs);
        auto had = [&](const pel *ptr) -> u32 {
            if (!pre) return wave_had<TT>(Lg, u.w, u.h, u.org, u.so, ptr, u.P->ref_stride);
            int v = 0;
            if (SPLIT) {
                // the two lanes of a tile hold its sum and difference halves;
                // the tile is normalised (com_cost.c:467, 586) before tiles
                // are summed
                const int hs = tile ? had_pre<TT>(opairs, ptr + ty * u.P->ref_stride + tx, u.P->ref_stride, li & 1) : 0;
                const int sv = hs + __builtin_amdgcn_mov_dpp(hs, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
                v = tile && !(li & 1) ? me_satd_norm128(sv) : 0;
            } else if (tile) {
                v = satd_pre<TT>(opairs, ptr + ty * u.P->ref_stride + tx, u.P->ref_stride);
            }
            return (u32)(group_sum(v, Lg) >> (BIT_DEPTH - 8));
        };
        auto phase = [&](bool half, int ox, int oy, u64 &best, int &bx, int &by) {
            const int n = half ? 9 : 8;
            for (int b0 = 0; b0 < n; b0 += G) {
                const int k = b0 + g < n ? b0 + g : b0;
                int mx, my;
                pos(half, ox, oy, k, mx, my);
                const pel *ptr = half && k == 0 ? plane_ptr(*u.P, u.r, 0, u.x + (mx >> 2), u.y + (my >> 2))
                                                : qpel_ptr(*u.P, u.r, mx + (u.x << 2), my + (u.y << 2));
                const u32 d = had(ptr);
                for (int q = 0; q < G && b0 + q < n; q++) {
                    const u32 dq = (u32)__shfl((int)d, base + q * Lg, 64);
                    pos(half, ox, oy, b0 + q, mx, my);
                    const u64 c = (u64)(a.lambda * u.qpel_bits(mx, my)) + ((u64)dq << shift);
                    if (c < best) { best = c; bx = mx; by = my; }
                }
            }
        };
        u64 best = ~0ull;
        int bx = cx, by = cy;
        phase(true, cx, cy, best, bx, by);
        if (a.mvr == 0) phase(false, bx, by, best, bx, by);
        ur.mv[0] = bx;
        ur.mv[1] = by;
        return best;
    }
    auto dist = [&](const pel *p) -> u32 { return wave_sad(u.org, u.so, p, u.P->ref_stride, u.w, u.h, u.lpu); };
    u32 d = dist(plane_ptr(*u.P, u.r, 0, u.x + (cx >> 2), u.y + (cy >> 2)));
    u64 best = (u64)(a.lambda * u.qpel_bits(cx, cy)) + ((u64)d << shift);
    for (int p = 0; p < 8; p++) {
        const int mx = cx + 2 * pat(ME_SUBPEL_HALF_X, p), my = cy + 2 * pat(ME_SUBPEL_HALF_Y, p);
        d = dist(qpel_ptr(*u.P, u.r, mx + (u.x << 2), my + (u.y << 2)));
        const u64 c = (u64)(a.lambda * u.qpel_bits(mx, my)) + ((u64)d << shift);
        if (c < best) { best = c; u.mv[0] = mx; u.mv[1] = my; }
    }
    if (a.mvr == 0) {
        const int hx = u.mv[0], hy = u.mv[1];
        for (int p = 0; p < 8; p++) {
            const int mx = hx + pat(ME_SUBPEL_QUARTER_X, p), my = hy + pat(ME_SUBPEL_QUARTER_Y, p);
            d = dist(qpel_ptr(*u.P, u.r, mx + (u.x << 2), my + (u.y << 2)));
            const u64 c = (u64)(a.lambda * u.qpel_bits(mx, my)) + ((u64)d << shift);
            if (c < best) { best = c; u.mv[0] = mx; u.mv[1] = my; }
        }
    }
    // cost_type 1: the final cost is com_had's at the chosen position
    d = wave_had<TT>(u.lpu, u.w, u.h, u.org, u.so,
                     qpel_ptr(*u.P, u.r, u.mv[0] + (u.x << 2), u.mv[1] + (u.y << 2)), u.P->ref_stride);
    best = (u64)(a.lambda * u.qpel_bits(u.mv[0], u.mv[1])) + ((u64)d << (16 - a.bi));
    ur.mv[0] = u.mv[0];
    ur.mv[1] = u.mv[1];
    return best;
}

// grid (the shapes' workgroups, jobs, refs): each shape of the launch (one
// com_had tile type TT) has ceil(per_ctu / (4 * 64 / lpu)) workgroups of 4
// waves of 64 / lpu (PU, reference) groups.  All shapes of a tile type run
// in one launch, the largest PUs first, so that their long sequential
// searches overlap with the many short ones of the small PUs.
// The argument block is read where the launch put it, in the kernarg
// segment (the first explicit argument sits at offset 0): a search unit keeps
// a pointer to it, and taking the address of the parameter itself would make
// every lane copy the whole block into private memory first.
#define TZ_ARGS(T) (*(const T *)__builtin_amdgcn_kernarg_segment_ptr())

template <int TT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TZ_WPE))) void k_tz(TzArgs)
{
    const TzArgs &a = TZ_ARGS(TzArgs);
    int k = 0;
    while (k + 1 < a.nshp && a.shp[k + 1].blk0 <= (int)blockIdx.x) k++;
    const TzShape &S = a.shp[k];
    const int lpu = S.lpu, pw = S.pw, ph = S.ph, base = S.base;
    const int lg = __builtin_ctz((unsigned)lpu), upw = 64 >> lg;
    const int li = threadIdx.x & (lpu - 1), wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = (((int)blockIdx.x - S.blk0) * 4 + wv) * upw + ((threadIdx.x & 63) >> lg);
    if (i >= S.per_ctu) return;
    const int job = blockIdx.y;
    const int n0 = a.n0, n1 = a.nref - a.n0;
    int r = blockIdx.z, blist = 0;
    if (a.bi) {   // the searched list's reference blockIdx.z
        const BiState &bs = a.bst[(size_t)job * a.npu + base + i];
        if (bs.done) return;
        blist = 1 - bs.fixed;
        if (r >= (blist ? n1 : n0)) return;
        r += blist ? n0 : 0;
    }
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    const int cols = 128 / pw;
    TzUnit u;
    u.a = &a;
    u.P = &P;
    u.r = r;
    u.w = pw;
    u.h = ph;
    u.lpu = lpu;
    u.x = (ctu % P.ctu_cols) * 128 + (i % cols) * pw;
    u.y = (ctu / P.ctu_cols) * 128 + (i / cols) * ph;
    const size_t slot = ((size_t)job * a.npu + base + i) * a.nref + r;
    me_pu_result_t out;
    out.mv_int[0] = out.mv_int[1] = out.mv_sub[0] = out.mv_sub[1] = 0;
    out.cost_int = out.cost_sub = ~0ull;
    out.cost_final = 0xFFFFFFFFu;
    out.valid = 0;
    u.mvpx = a.mvp ? a.mvp[slot * 2] : 0;
    u.mvpy = a.mvp ? a.mvp[slot * 2 + 1] : 0;
    // get_max_search_range (me.c:502-527)
    const int mvr = a.mvr;
    const int T = (a.tab_off << mvr) >> 2;
    const int px = u.mvpx >> 2, py = u.mvpy >> 2;
    int lox, loy, hix, hiy;
    if (a.pad_win) {
        lox = -128 + 1 - u.x; loy = -128 + 1 - u.y; hix = P.pic_w - 1 - u.x; hiy = P.pic_h - 1 - u.y;
    } else {
        lox = -u.x; loy = -u.y; hix = P.pic_w - 1 - u.x - u.w; hiy = P.pic_h - 1 - u.y - u.h;
    }
    if (px - T + 4 > lox) lox = px - T + 4;
    if (py - T + 4 > loy) loy = py - T + 4;
    if (px + T - 4 < hix) hix = px + T - 4;
    if (py + T - 4 < hiy) hiy = py + T - 4;
    int v[4] = {(s16)lox, (s16)loy, (s16)hix, (s16)hiy};
    if (mvr > 2) {   // com_ipel_range_rounding (me.c:48-55)
        const int sh = mvr - 2;
        for (int k = 0; k < 4; k++) v[k] = (s16)(v[k] >= 0 ? ((v[k] >> sh) << sh) : -(((-v[k]) >> sh) << sh));
    }
    u.lo[0] = v[0]; u.lo[1] = v[1]; u.hi[0] = v[2]; u.hi[1] = v[3];
    const bool inside = u.x + u.w <= P.pic_w && u.y + u.h <= P.pic_h;
    if (!inside || u.lo[0] > u.hi[0] || u.lo[1] > u.hi[1]) {
        if (li == 0) a.res[slot] = out;
        return;
    }
    if (a.bi) {
        u.org = a.orgbi + ((size_t)job * a.nsel * 16384 + S.orgbi_off) + (size_t)i * pw * ph;
        u.so = pw;
        u.ob = a.orgbi;
    } else {
        u.org = P.cur + (size_t)u.y * P.pic_w + u.x;
        u.so = P.pic_w;
        u.ob = P.cur;
    }
    u.ooff = (u32)((u.org - u.ob) * (int)sizeof(pel));
    u.pb = P.planes[r];
    u.shift = 16 - a.bi - (BIT_DEPTH - 8);
    u.best = ~0ull;
    const int sx = a.bi ? a.mv_scale[slot * 2] : u.mvpx, sy = a.bi ? a.mv_scale[slot * 2 + 1] : u.mvpy;
    u.mv[0] = me_clip3(u.lo[0], u.hi[0], (sx + 2) >> 2);
    u.mv[1] = me_clip3(u.lo[1], u.hi[1], (sy + 2) >> 2);
    const int refi = r < a.n0 ? r : r - a.n0, nlist = r < a.n0 ? a.n0 : a.nref - a.n0;
    constexpr int NB = TzBatch<TT>::NB;
    const int bs = tz_diamond<NB>(u, 32);
    if (!a.bi && (abs(u.mvpx - (u.mv[0] << 2)) > 2 || abs(u.mvpy - (u.mv[1] << 2)) > 2)) {
        if (bs > 5) {
            const int pvx = a.mvi_prev ? a.mvi_prev[slot * 2] : 0, pvy = a.mvi_prev ? a.mvi_prev[slot * 2 + 1] : 0;
            tz_raster<NB>(u, refi, pvx, pvy);
        }
        tz_diamond<NB>(u, 8);
    }
    out.mv_int[0] = (s16)u.mv[0];
    out.mv_int[1] = (s16)u.mv[1];
    out.cost_int = u.best;
    u.mv[0] *= 4;
    u.mv[1] *= 4;
    u64 best = u.best;
    if (mvr < 2) best = tz_subpel<TT>(u);
    out.mv_sub[0] = (s16)u.mv[0];
    out.mv_sub[1] = (s16)u.mv[1];
    out.cost_sub = best;
    const int rb = nlist <= 1 ? 0 : (refi < nlist - 1 ? refi + 1 : nlist - 1);   // tbl_refi_bits
    // bi: the fixed list's mot_bits + 1 (me.c:560-561)
    const int last = a.bi ? a.bpu[(size_t)job * a.npu + base + i].mot_bits[1 - blist] + 1 : 2;
    const u64 cf = best + (u64)(a.lambda * (u32)(rb + mvr + (mvr < 4) + last));
    out.cost_final = (u32)((cf + (1u << 15)) >> 16);
    out.valid = 1;
    if (li == 0) {
        a.res[slot] = out;
        if (a.bi) {
            a.mv_scale[slot * 2] = (s16)u.mv[0];
            a.mv_scale[slot * 2 + 1] = (s16)u.mv[1];
        }
    }
}

// ---------------------------------------------------------------------------
// analyze_smvd (inter.c:1194-1243) with smvd_refine (inter.c:1083-1191):
// symmetric MVD search of list 0 / list 1 reference 0 for every PU.
// ---------------------------------------------------------------------------

// smvd_refine's patterns (inter.c:1097-1098): cross, then diamond
__constant__ static const s8 SMVD_CROSS[4][2] = {{0, 1}, {1, 0}, {0, -1}, {-1, 0}};
__constant__ static const s8 SMVD_DIAMOND[8][2] = {{0, 2}, {1, 1}, {2, 0}, {1, -1}, {0, -2}, {-1, -1}, {-2, 0}, {-1, 1}};

// LDS tile of one lane: the bi-average of a com_had tile (up to 16x8)
#define SMVD_TILE 136

struct SmvdUnit {
    const TzArgs *a;
    const PicDesc *P;
    int x, y;
    const pel *org;
    pel *tile;               // this lane's LDS tile
    int mvp[2][2];
    __device__ u32 mvbits(int v) const   // tab_mvbits[v >> mvr], the table held to its range
    {
        const int i = v >> a->mvr;
        return a->tab[i < -a->tab_off ? -a->tab_off : (i > a->tab_off ? a->tab_off : i)];
    }
    __device__ u32 mv_cost(u32 bits) const { return (a->lambda * bits + 32768u) >> 16; }   // MV_COST (inter.c:41)
};

// get_smvd_cost (inter.c:1071-1081) of the bi-average of list 0 ref 0 at m0
// and list 1 ref 0 at m1 (com_mc_cu, com_mc.c:140-200: the rounding
// average of two quarter-pel plane blocks); `had` forces com_had.  The
// aligned groups of `lanes` lanes (a.lpu, or fewer for position groups) sum
// their own candidates.
template <int TT>
__device__ u32 smvd_cost(const SmvdUnit &u, int m0x, int m0y, int m1x, int m1y, bool had, int lanes)
{
    const TzArgs &a = *u.a;
    const PicDesc &P = *u.P;
    const pel *p0 = qpel_ptr(P, 0, (u.x << 2) + m0x, (u.y << 2) + m0y);
    const pel *p1 = qpel_ptr(P, a.n0, (u.x << 2) + m1x, (u.y << 2) + m1y);
    const int li = threadIdx.x & (lanes - 1);
    int s = 0;
    if (!had) {   // SAD >> (bd - 8)
        const int q = a.pw >> 2;
        for (int i = li; i < q * a.ph; i += lanes) {
            const int xx = (i % q) * 4, yy = i / q;
            int v0[4], v1[4];
            load_row<4>(p0 + yy * P.ref_stride + xx, v0);
            load_row<4>(p1 + yy * P.ref_stride + xx, v1);
            const pel *o = u.org + yy * P.pic_w + xx;
            for (int k = 0; k < 4; k++) s += abs((int)o[k] - ((v0[k] + v1[k] + 1) >> 1));
        }
        return (u32)(group_sum(s, lanes) >> (BIT_DEPTH - 8));
    }
    // the tile type's tile size (had_tiling)
    constexpr int TW = TT == 0 ? 4 : (TT == 1 ? 8 : (TT == 2 ? 8 : (TT == 3 ? 16 : (TT == 4 ? 4 : 8))));
    constexpr int TH = TT == 0 ? 4 : (TT == 1 ? 8 : (TT == 2 ? 4 : (TT == 3 ? 8 : (TT == 4 ? 8 : 16))));
    constexpr int NW = TW * (int)sizeof(pel) / 4;   // dwords per tile row
    // (a + b + 1) >> 1 of every sample of two dwords: (a | b) - ((a ^ b) >> 1)
    // per sample, the shifted-in bit of the next sample masked off
    constexpr u32 HM = sizeof(pel) == 1 ? 0x7F7F7F7Fu : 0x7FFF7FFFu;
    const int tpr = a.pw / TW, T = tpr * (a.ph / TH);
    for (int t = li; t < T; t += lanes) {
        const int tx = (t % tpr) * TW, ty = (t / tpr) * TH;
        // the tile's average, stored as whole dwords (com_had reads dwords)
#pragma unroll
        for (int yy = 0; yy < TH; yy++) {
            u32 *d = (u32 *)(u.tile + yy * TW);
            u32 w0[NW], w1[NW];
            load_words<NW>(p0 + (ty + yy) * P.ref_stride + tx, w0);
            load_words<NW>(p1 + (ty + yy) * P.ref_stride + tx, w1);
#pragma unroll
            for (int k = 0; k < NW; k++) d[k] = (w0[k] | w1[k]) - (((w0[k] ^ w1[k]) >> 1) & HM);
        }
        s += satd_tile_fn<TT>(u.org + ty * P.pic_w + tx, P.pic_w, u.tile, TW);
    }
    return (u32)(group_sum(s, lanes) >> (BIT_DEPTH - 8));
}

// smvd_refine (inter.c:1083-1191) for the patterns analyze_smvd uses:
// diamond (pattern 2) and cross (pattern 0), offsets scaled by the AMVR shift.
// A round's points all lie around the round's centre, so with com_had costs
// (one lane per tile, area/16 lanes per PU) the PU's lanes form up to four
// position groups that evaluate several points at a time; the costs are then
// scanned in the reference's order with its strict '<'.
template <int TT>
__device__ u32 smvd_refine(const SmvdUnit &u, int mv[2][2], u32 cost, int pattern, int rounds)
{
    const TzArgs &a = *u.a;
    const int mvr = a.mvr;
    const u32 ref_mvr_bits = (u32)(mvr + (mvr < 4));   // tbl_refi_bits[0][0] = 0
    const bool had = a.cost_type == 0;
    const int T = (a.pw / a.tw) * (a.ph / a.th);
    int G = had && T < a.lpu ? a.lpu / T : 1;
    if (G > 4) G = 4;
    const int Lg = G > 1 ? T : a.lpu;
    const int lane = threadIdx.x & 63, base = lane & ~(a.lpu - 1);
    int g = (lane - base) / Lg;
    if (g >= G) g = 0;   // spare lanes repeat group 0's point
    int ds = 0, de = pattern == 0 ? 3 : 7;
    const int rnd = pattern == 0 ? 4 : 8, mask = pattern == 0 ? 3 : 7;
    for (int round = 0; round < rounds; round++) {
        const int cx = mv[0][0], cy = mv[0][1];
        // point idx of the round: its direction, both lists' MVs and MV bits
        auto point = [&](int idx, int &dir, int &c0x, int &c0y, int &c1x, int &c1y, u32 &bits) {
            dir = (idx + rnd) & mask;
            const int ox = pattern == 0 ? SMVD_CROSS[dir][0] : SMVD_DIAMOND[dir][0];
            const int oy = pattern == 0 ? SMVD_CROSS[dir][1] : SMVD_DIAMOND[dir][1];
            c0x = (s16)(cx + (s16)(ox << mvr));
            c0y = (s16)(cy + (s16)(oy << mvr));
            const int dx = c0x - u.mvp[0][0], dy = c0y - u.mvp[0][1];
            c1x = (s16)(u.mvp[1][0] - dx);
            c1y = (s16)(u.mvp[1][1] - dy);
            bits = ref_mvr_bits + u.mvbits(dx) + u.mvbits(dy);
        };
        int best_dir = -1;
        for (int b0 = ds; b0 <= de; b0 += G) {
            const int idx = b0 + g <= de ? b0 + g : b0;
            int dir, c0x, c0y, c1x, c1y;
            u32 bits;
            point(idx, dir, c0x, c0y, c1x, c1y, bits);
            const u32 d = smvd_cost<TT>(u, c0x, c0y, c1x, c1y, had, Lg);
            for (int q = 0; q < G && b0 + q <= de; q++) {
                point(b0 + q, dir, c0x, c0y, c1x, c1y, bits);
                const u32 t = (u32)__shfl((int)d, base + q * Lg, 64) + u.mv_cost(bits);
                if (t < cost) {
                    cost = t;
                    mv[0][0] = c0x; mv[0][1] = c0y; mv[1][0] = c1x; mv[1][1] = c1y;
                    best_dir = dir;
                }
            }
        }
        if (best_dir < 0) break;
        const int step = pattern == 2 ? 2 - (best_dir & 1) : 1;
        ds = best_dir - step;
        de = best_dir + step;
    }
    return cost;
}

// grid (ceil(per_ctu / (2 * 64 / lpu)), jobs), 2 waves of 64 / lpu PU groups.
// out[job][npu]: .active selects the PU; mv and cost are written.
template <int TT>
__global__ __launch_bounds__(128) void k_smvd(TzArgs, me_smvd_pu_t *out)
{
    const TzArgs &a = TZ_ARGS(TzArgs);
    __shared__ __align__(16) pel tiles[128 * SMVD_TILE];
    const int lg = __builtin_ctz((unsigned)a.lpu), upw = 64 >> lg;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = (blockIdx.x * 2 + wv) * upw + ((threadIdx.x & 63) >> lg);
    if (i >= a.per_ctu) return;
    const int job = blockIdx.y;
    const size_t pk = (size_t)job * a.npu + a.base + i;
    if (!out[pk].active) return;
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    const int cols = 128 / a.pw;
    SmvdUnit u;
    u.a = &a;
    u.P = &P;
    u.x = (ctu % P.ctu_cols) * 128 + (i % cols) * a.pw;
    u.y = (ctu / P.ctu_cols) * 128 + (i / cols) * a.ph;
    if (u.x + a.pw > P.pic_w || u.y + a.ph > P.pic_h) return;
    u.org = P.cur + (size_t)u.y * P.pic_w + u.x;
    u.tile = tiles + threadIdx.x * SMVD_TILE;
    const size_t s0 = pk * a.nref, s1 = pk * a.nref + a.n0;   // list 0 / list 1 reference 0
    u.mvp[0][0] = a.mvp[s0 * 2]; u.mvp[0][1] = a.mvp[s0 * 2 + 1];
    u.mvp[1][0] = a.mvp[s1 * 2]; u.mvp[1][1] = a.mvp[s1 * 2 + 1];
    const int mvr = a.mvr;
    const bool had = a.cost_type == 0;
    // MV = MVP in both lists (inter.c:1219-1225)
    int mv[2][2] = {{u.mvp[0][0], u.mvp[0][1]}, {u.mvp[1][0], u.mvp[1][1]}};
    u32 cost = smvd_cost<TT>(u, mv[0][0], mv[0][1], mv[1][0], mv[1][1], had, a.lpu) +
               u.mv_cost(u.mvbits(0) * 2 + mvr + (mvr < 4));
    // list 0's uni MV, mirrored into list 1 (inter.c:1227-1243)
    {
        const int m0x = a.mv_scale[s0 * 2], m0y = a.mv_scale[s0 * 2 + 1];
        const int dx = m0x - u.mvp[0][0], dy = m0y - u.mvp[0][1];
        const int m1x = (s16)(u.mvp[1][0] - dx), m1y = (s16)(u.mvp[1][1] - dy);
        const u32 cb = smvd_cost<TT>(u, m0x, m0y, m1x, m1y, had, a.lpu) +
                       u.mv_cost(u.mvbits(dx) + u.mvbits(dy) + mvr + (mvr < 4));
        if (cb < cost) {
            cost = cb;
            mv[0][0] = m0x; mv[0][1] = m0y; mv[1][0] = m1x; mv[1][1] = m1y;
        }
    }
    cost = smvd_refine<TT>(u, mv, cost, 2, 8);
    cost = smvd_refine<TT>(u, mv, cost, 0, 1);
    if (!had) {   // SAD search: the winner's com_had (inter.c:1248-1255)
        const int dx = mv[0][0] - u.mvp[0][0], dy = mv[0][1] - u.mvp[0][1];
        cost = smvd_cost<TT>(u, mv[0][0], mv[0][1], mv[1][0], mv[1][1], true, a.lpu) +
               u.mv_cost(u.mvbits(dx) + u.mvbits(dy) + mvr + (mvr < 4));
    }
    if ((threadIdx.x & (a.lpu - 1)) == 0) {
        for (int l = 0; l < 2; l++) {
            out[pk].mv[l][0] = (s16)mv[l][0];
            out[pk].mv[l][1] = (s16)mv[l][1];
        }
        out[pk].cost = cost;
    }
}

typedef void (*smvd_fn)(TzArgs, me_smvd_pu_t *);
#if ME_TZ_PART == 0
smvd_fn me_smvd_kernel_small(int tt)
{
    switch (tt) {
    case 0: return k_smvd<0>;
    case 1: return k_smvd<1>;
    case 2: return k_smvd<2>;
    case 4: return k_smvd<4>;
    default: return nullptr;
    }
}
#elif ME_TZ_PART == 1
smvd_fn me_smvd_kernel_16x8(int tt) { return tt == 3 ? k_smvd<3> : nullptr; }
#else
smvd_fn me_smvd_kernel_8x16(int tt) { return tt == 5 ? k_smvd<5> : nullptr; }
#endif

typedef void (*tz_fn)(TzArgs);
#if ME_TZ_PART == 0
tz_fn me_tz_kernel_small(int tt)
{
    switch (tt) {
    case 0: return k_tz<0>;
    case 1: return k_tz<1>;
    case 2: return k_tz<2>;
    case 4: return k_tz<4>;
    default: return nullptr;
    }
}
#elif ME_TZ_PART == 1
tz_fn me_tz_kernel_16x8(int tt) { return tt == 3 ? k_tz<3> : nullptr; }
#else
tz_fn me_tz_kernel_8x16(int tt) { return tt == 5 ? k_tz<5> : nullptr; }
#endif

#if ME_TZ_PART == 0
tz_fn me_tz_kernel_16x8(int tt);
tz_fn me_tz_kernel_8x16(int tt);

// The lookahead's inter cost of each 16x16 unit (loka_estimate_coding_cost,
// lookahead.c:130-208) from the searched MVs: per list, the com_had of the
// unit against each reference at its MV (com_mc_blk_luma at an integer MV is
// the integer plane at x + mv, y + mv, com_mc.c:41-86), the list's first
// cheapest kept; then the average (pel_avrg, com_util.c:3088-3105) of the
// two lists' kept predictions, and of both lists' first reference at MV 0.
// out[unit] is the cheapest (u32 max when nothing was searched).  A group of
// 4 lanes per unit, one 8x8 com_had tile per lane; the averaged tiles go
// through LDS.
#if ME_TZ_PART == 0
struct LokaArgs {
    const PicDesc *pic;
    const me_pu_result_t *res;
    u32 *out;
    int ucols, units, npu, nref, n0, base16;
};

__device__ __forceinline__ u32 loka_had(const pel *org, int so, const pel *pred, int sp)
{
    const int li = threadIdx.x & 3, tx = (li & 1) * 8, ty = (li >> 1) * 8;
    const int s = satd_t<1>(org + ty * so + tx, so, pred + ty * sp + tx, sp);
    return (u32)(group_sum(s, 4) >> (BIT_DEPTH - 8));
}

// this lane's 8x8 tile of (p + q + 1) >> 1 into t (stride 16: the group's
// block).  Stored as whole dwords: com_had reads the tile back as dwords
// (load_row), and 16-bit stores would not be ordered before those loads.
__device__ __forceinline__ void loka_avg_tile(pel *t, const pel *p, const pel *q, int sp, int sq)
{
    constexpr int PW = 4 / (int)sizeof(pel);   // samples per dword
    const int li = threadIdx.x & 3, tx = (li & 1) * 8, ty = (li >> 1) * 8;
    for (int y = 0; y < 8; y++) {
        u32 *d = (u32 *)(t + (ty + y) * 16 + tx);
        const pel *a = p + (ty + y) * sp + tx, *b = q + (ty + y) * sq + tx;
        for (int k = 0; k < 8 / PW; k++) {
            u32 v = 0;
            for (int j = 0; j < PW; j++)
                v |= (u32)((a[k * PW + j] + b[k * PW + j] + 1) >> 1) << (8 * (int)sizeof(pel) * j);
            d[k] = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_loka(LokaArgs a)
{
    // 64 groups' blocks, + 8 samples of slack: tile rows are read as dwords
    __shared__ __align__(16) pel avg[64 * 256 + 8];
    const int g = threadIdx.x >> 2, unit = blockIdx.x * 64 + g;
    if (unit >= a.units) return;
    const PicDesc &P = *a.pic;
    const int x = (unit % a.ucols) * 16, y = (unit / a.ucols) * 16;
    const int ctu = (y >> 7) * P.ctu_cols + (x >> 7), i = ((y & 127) >> 4) * 8 + ((x & 127) >> 4);
    const size_t slot = ((size_t)ctu * a.npu + a.base16 + i) * a.nref;
    const pel *org = P.cur + (size_t)y * P.pic_w + x;
    u32 best = 0xFFFFFFFFu;
    const pel *kept[2] = {nullptr, nullptr};
    for (int l = 0; l < 2; l++) {
        const int r0 = l ? a.n0 : 0, r1 = l ? a.nref : a.n0;
        u32 best_l = 0xFFFFFFFFu;
        for (int r = r0; r < r1; r++) {
            const me_pu_result_t &q = a.res[slot + r];
            if (!q.valid) continue;
            const pel *pred = plane_ptr(P, r, 0, x + (q.mv_sub[0] >> 2), y + (q.mv_sub[1] >> 2));
            const u32 c = loka_had(org, P.pic_w, pred, P.ref_stride);
            if (c < best_l) {
                best_l = c;
                best = c < best ? c : best;
                kept[l] = pred;
            }
        }
    }
    if (a.n0 > 0 && a.nref > a.n0) {
        pel *t = avg + g * 256;
        if (kept[0] && kept[1]) {
            loka_avg_tile(t, kept[0], kept[1], P.ref_stride, P.ref_stride);
            const u32 c = loka_had(org, P.pic_w, t, 16);
            best = c < best ? c : best;
        }
        loka_avg_tile(t, plane_ptr(P, 0, 0, x, y), plane_ptr(P, a.n0, 0, x, y), P.ref_stride, P.ref_stride);
        const u32 c = loka_had(org, P.pic_w, t, 16);
        best = c < best ? c : best;
    }
    if ((threadIdx.x & 3) == 0) a.out[unit] = best;
}

int me_loka_enqueue(const MeSearch *s, const PicDesc *pic, const me_pu_result_t *res, int pic_w, int pic_h, u32 *out,
                    hipStream_t st)
{
    LokaArgs a;
    memset(&a, 0, sizeof(a));
    a.pic = pic;
    a.res = res;
    a.out = out;
    a.ucols = pic_w / 16;
    a.units = a.ucols * (pic_h / 16);
    a.npu = s->npu;
    a.nref = s->nref;
    a.n0 = s->nref - s->nref_l1;
    a.base16 = s->base[S16x16];
    if (a.units <= 0) return 0;
    hipLaunchKernelGGL(k_loka, dim3((a.units + 63) / 64), dim3(256), 0, st, a);
    ME_CHECK(hipGetLastError());
    return 0;
}
#endif

// ---------------------------------------------------------------------------
// analyze_bi (inter.c:972-1033): target, search of the other list, choice
// ---------------------------------------------------------------------------

#if ME_TZ_PART == 0
size_t me_bi_state_bytes(size_t n) { return n * sizeof(BiState); }

__global__ void k_bi_init(const me_bi_pu_t *pu, BiState *st, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    st[k].best = ~0ull;
    st[k].fixed = pu[k].lidx_ref;
    st[k].done = pu[k].active ? 0 : 1;
}

// create_bi_org (inter.c:55-68) of one shape's PUs from the fixed list's
// prediction (com_mc_blk_luma at its quarter-pel MV: the plane [qy&3][qx&3]
// at the clipped integer position, com_mc.c:41-86, 140-165).  Block (256):
// min(256, area) threads per PU, 256 / that PUs per block, one sample per
// thread and step.  grid (ceil(per_ctu / PUs per block), jobs).
__global__ __launch_bounds__(256) void k_bi_org(TzArgs a)
{
    const int area = a.pw * a.ph, tpp = area < 256 ? area : 256;
    const int i = blockIdx.x * (256 / tpp) + threadIdx.x / tpp, job = blockIdx.y;
    if (i >= a.per_ctu) return;
    const size_t pk = (size_t)job * a.npu + a.base + i;
    if (a.bst[pk].done) return;
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    const int cols = 128 / a.pw;
    const int x = (ctu % P.ctu_cols) * 128 + (i % cols) * a.pw, y = (ctu / P.ctu_cols) * 128 + (i / cols) * a.ph;
    if (x + a.pw > P.pic_w || y + a.ph > P.pic_h) return;
    const me_bi_pu_t &q = a.bpu[pk];
    const int f = a.bst[pk].fixed, r = (f ? a.n0 : 0) + q.refi[f];
    const pel *pred = qpel_ptr(P, r, (x << 2) + q.mv[f][0], (y << 2) + q.mv[f][1]);
    const pel *org = P.cur + (size_t)y * P.pic_w + x;
    pel *dst = (pel *)a.orgbi + ((size_t)job * a.nsel * 16384 + a.orgbi_off) + (size_t)i * a.pw * a.ph;
    const int maxv = (1 << BIT_DEPTH) - 1;
    for (int k = threadIdx.x % tpp; k < area; k += tpp) {
        const int yy = k / a.pw, xx = k - yy * a.pw;
        const int v = 2 * (int)org[yy * P.pic_w + xx] - (int)pred[yy * P.ref_stride + xx];
        dst[k] = (pel)me_clip3(0, maxv, v);
    }
}

// The reference loop over the searched list (inter.c:1006-1021): a strictly
// cheaper me_search_tz cost takes the reference, its MV and get_mv_bits.
__global__ void k_bi_select(TzArgs a, me_bi_pu_t *pu, BiState *st, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || st[k].done) return;
    const int s = 1 - st[k].fixed, nl = s ? a.nref - a.n0 : a.n0, r0 = s ? a.n0 : 0;
    bool changed = false;
    for (int rc = 0; rc < nl; rc++) {
        const size_t slot = (size_t)k * a.nref + r0 + rc;
        const me_pu_result_t &q = a.res[slot];
        const u64 c = q.valid ? (u64)q.cost_final : ~0ull;
        if (c < st[k].best) {
            st[k].best = c;
            changed = true;
            const int mx = a.mv_scale[slot * 2], my = a.mv_scale[slot * 2 + 1];
            const int px = a.mvp[slot * 2], py = a.mvp[slot * 2 + 1];
            pu[k].refi[s] = (s8)rc;
            pu[k].mv[s][0] = (s16)mx;
            pu[k].mv[s][1] = (s16)my;
            const int rb = nl <= 1 ? 0 : (rc < nl - 1 ? rc + 1 : nl - 1);
            pu[k].mot_bits[s] = rb + a.tab[(mx - px) >> a.mvr] + a.tab[(my - py) >> a.mvr];
        }
    }
    if (changed) st[k].fixed = s;
    else st[k].done = 1;
    pu[k].cost = st[k].best > 0xFFFFFFFFull ? 0xFFFFFFFFu : (u32)st[k].best;
}
#endif

int me_tz_enqueue(const MeSearch *s, const JobSet &js, const int16_t *mvp, const int16_t *mvi_prev,
                  const me_tz_cfg_t *tz, me_pu_result_t *res, hipStream_t st, int only_shape, int pad_win,
                  u32 lambda, const BiBufs *bi)
{
    TzArgs a;
    memset(&a, 0, sizeof(a));
    a.js = js;
    a.mvp = mvp;
    a.mvi_prev = mvi_prev;
    a.res = res;
    a.tab = s->d_tab_base + s->tab_off;
    a.tab_off = s->tab_off;
    a.lambda = lambda ? lambda : s->lambda;
    a.pad_win = pad_win;
    a.npu = s->npu;
    a.nref = s->nref;
    a.n0 = s->nref - s->nref_l1;
    a.gop = tz->gop_size;
    a.msr = tz->max_search_range;
    a.adaptive = tz->adaptive_raster_range;
    a.mvr = tz->mvr;
    a.cost_type = tz->cost_type;
    a.cur_poc = tz->cur_poc;
    for (int r = 0; r < ME_MAX_REFS; r++) a.ref_poc[r] = tz->ref_poc[r];
    if (bi) {
        a.bi = 1;
        a.orgbi = bi->orgbi;
        a.mv_scale = bi->mv_scale;
        a.bst = bi->state;
        a.bpu = bi->pus;
        for (int k = 0; k < ME_NSHAPES; k++) a.nsel += s->base[k] >= 0;
    }
    // the selected shapes' ranks (bi targets are laid out in shape order)
    int rank_of[ME_NSHAPES], rank = 0;
    for (int k = 0; k < ME_NSHAPES; k++) rank_of[k] = s->base[k] >= 0 ? rank++ : -1;
    const int nz = bi ? std::max(a.n0, a.nref - a.n0) : s->nref;
    // one launch per com_had tile type, its shapes largest PUs first.  The
    // launches write disjoint results: with more than one, each runs on a
    // side stream of its own, forked from st and joined back into it, so that
    // the long searches of one launch's largest PUs overlap the others.
    int tt_of[ME_NSHAPES], nlaunch = 0;
    bool has_tt[6] = {false, false, false, false, false, false};
    for (int k = 0; k < ME_NSHAPES; k++) {
        int tw, th;
        had_tiling(ME_SHAPES_HOST[k][0], ME_SHAPES_HOST[k][1], &tt_of[k], &tw, &th);
        if (s->base[k] >= 0 && (only_shape < 0 || k == only_shape)) has_tt[tt_of[k]] = true;
    }
    for (int tt = 0; tt < 6; tt++) nlaunch += has_tt[tt];
    const bool side = nlaunch > 1 && s->tz_side[0];
    if (side) ME_CHECK(hipEventRecord(s->tz_ev[6], st));
    for (int tt = 0; tt < 6; tt++) {
        int order[ME_NSHAPES], no = 0;
        for (int k = 0; k < ME_NSHAPES; k++) {
            if (s->base[k] < 0 || (only_shape >= 0 && k != only_shape)) continue;
            if (tt_of[k] == tt) order[no++] = k;
        }
        if (!no) continue;
        const hipStream_t ls = side ? s->tz_side[tt] : st;
        if (side) ME_CHECK(hipStreamWaitEvent(ls, s->tz_ev[6], 0));
        std::stable_sort(order, order + no, [](int p, int q) {
            return ME_SHAPES_HOST[p][0] * ME_SHAPES_HOST[p][1] > ME_SHAPES_HOST[q][0] * ME_SHAPES_HOST[q][1];
        });
        tz_fn fn = tt == 3 ? me_tz_kernel_16x8(tt) : (tt == 5 ? me_tz_kernel_8x16(tt) : me_tz_kernel_small(tt));
        if (!fn) { me_set_error("no search kernel for shape", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
        a.nshp = no;
        int blocks = 0;
        for (int j = 0; j < no; j++) {
            const int k = order[j];
            TzShape &q = a.shp[j];
            q.pw = ME_SHAPES_HOST[k][0];
            q.ph = ME_SHAPES_HOST[k][1];
            q.base = s->base[k];
            q.per_ctu = (128 / q.pw) * (128 / q.ph);
            const int area = q.pw * q.ph;
            q.lpu = area / 16 < 4 ? 4 : (area / 16 > 64 ? 64 : area / 16);
            q.orgbi_off = 16384 * rank_of[k];
            q.blk0 = blocks;
            const int upb = 4 * (64 / q.lpu);
            blocks += (q.per_ctu + upb - 1) / upb;
            if (bi) {   // the shape's bi-prediction targets (k_bi_org reads the per-shape fields)
                a.pw = q.pw;
                a.ph = q.ph;
                a.base = q.base;
                a.per_ctu = q.per_ctu;
                a.orgbi_off = q.orgbi_off;
                const int per_block = 256 / std::min(256, area);
                hipLaunchKernelGGL(k_bi_org, dim3((a.per_ctu + per_block - 1) / per_block, js.n_jobs), dim3(256), 0,
                                   ls, a);
            }
        }
        hipLaunchKernelGGL(fn, dim3(blocks, js.n_jobs, nz), dim3(256), 0, ls, a);
        ME_CHECK(hipGetLastError());
        if (side) {
            ME_CHECK(hipEventRecord(s->tz_ev[tt], ls));
            ME_CHECK(hipStreamWaitEvent(st, s->tz_ev[tt], 0));
        }
    }
    if (bi) {
        const int n = js.n_jobs * s->npu;
        hipLaunchKernelGGL(k_bi_select, dim3((n + 255) / 256), dim3(256), 0, st, a, bi->pus, bi->state, n);
        ME_CHECK(hipGetLastError());
    }
    return 0;
}

smvd_fn me_smvd_kernel_16x8(int tt);
smvd_fn me_smvd_kernel_8x16(int tt);

int me_smvd_enqueue(const MeSearch *s, const JobSet &js, const int16_t *mvp, const int16_t *mv_scale,
                    const me_tz_cfg_t *tz, me_smvd_pu_t *out, hipStream_t st)
{
    TzArgs a;
    memset(&a, 0, sizeof(a));
    a.js = js;
    a.mvp = mvp;
    a.mv_scale = (s16 *)mv_scale;
    a.tab = s->d_tab_base + s->tab_off;
    a.tab_off = s->tab_off;
    a.lambda = s->lambda;
    a.npu = s->npu;
    a.nref = s->nref;
    a.n0 = s->nref - s->nref_l1;
    a.mvr = tz->mvr;
    a.cost_type = tz->cost_type;
    // one launch per shape, writing its own PUs' results: the launches run
    // side by side on the search's side streams (forked from st, joined back)
    int nshape = 0;
    for (int k = 0; k < ME_NSHAPES; k++) nshape += s->base[k] >= 0;
    const bool side = nshape > 1 && s->tz_side[0];
    if (side) ME_CHECK(hipEventRecord(s->tz_ev[6], st));
    for (int i = 0; side && i < 6; i++) ME_CHECK(hipStreamWaitEvent(s->tz_side[i], s->tz_ev[6], 0));
    int launch = 0;
    for (int k = 0; k < ME_NSHAPES; k++) {
        if (s->base[k] < 0) continue;
        const hipStream_t ls = side ? s->tz_side[launch++ % 6] : st;
        a.pw = ME_SHAPES_HOST[k][0];
        a.ph = ME_SHAPES_HOST[k][1];
        a.base = s->base[k];
        a.per_ctu = (128 / a.pw) * (128 / a.ph);
        had_tiling(a.pw, a.ph, &a.tt, &a.tw, &a.th);
        const int area = a.pw * a.ph;
        a.lpu = area / 16 < 4 ? 4 : (area / 16 > 64 ? 64 : area / 16);
        const int upb = 2 * (64 / a.lpu);
        smvd_fn fn = a.tt == 3 ? me_smvd_kernel_16x8(a.tt)
                               : (a.tt == 5 ? me_smvd_kernel_8x16(a.tt) : me_smvd_kernel_small(a.tt));
        if (!fn) { me_set_error("no SMVD kernel for shape", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
        hipLaunchKernelGGL(fn, dim3((a.per_ctu + upb - 1) / upb, js.n_jobs), dim3(128), 0, ls, a, out);
        ME_CHECK(hipGetLastError());
    }
    for (int i = 0; side && i < 6; i++) {
        ME_CHECK(hipEventRecord(s->tz_ev[i], s->tz_side[i]));
        ME_CHECK(hipStreamWaitEvent(st, s->tz_ev[i], 0));
    }
    return 0;
}

int me_bi_enqueue(const MeSearch *s, const JobSet &js, const int16_t *mvp, const me_tz_cfg_t *tz,
                  me_pu_result_t *res, const BiBufs &bi, hipStream_t st)
{
    const int n = js.n_jobs * s->npu;
    hipLaunchKernelGGL(k_bi_init, dim3((n + 255) / 256), dim3(256), 0, st, bi.pus, bi.state, n);
    ME_CHECK(hipGetLastError());
    for (int round = 0; round < 2; round++)
        if (me_tz_enqueue(s, js, mvp, nullptr, tz, res, st, -1, 1, 0, &bi)) return -1;
    return 0;
}
#endif
