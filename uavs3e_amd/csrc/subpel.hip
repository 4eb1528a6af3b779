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

#if BIT_DEPTH == 8
// ---------------------------------------------------------------------------
// 8 bits on the matrix cores.  A com_had tile of n = TW x TH samples (n = 32,
// 64, 128) is the n-point Sylvester Hadamard transform H_n of its residual in
// row-major order (H_TW (x) H_TH is H_n with the index y * TW + x), and its
// cost is the sum of |coef|.  H_n's 32 x 32 blocks are +-H_32 (block (m, s)
// has the sign of popcount(m & s)), so the coefficients 32m..32m+31 of a set
// of 32 tiles are sum_s +-H_32 B_s, B_s the tiles' samples 32s..32s+31: one
// v_mfma_i32_32x32x32_i8 per (m, s).  The samples enter as int8 x - 128, the
// current block's transform T(org - 128) is made once per PU, and each
// position's chain starts from 0x8000 - T(org - 128), so that it ends at
// 0x8000 + T(pred - org) (every coefficient within +-n * 255 <= 32640): one
// v_sad_u16 against 0x8000 per coefficient sums |coef|.
//
// Operand layout of v_mfma_i32_32x32x32_i8 (gfx950): A and B lane l hold row
// (column) l % 32 at K = 16 (l / 32) + q in byte q; D lane l holds column
// l % 32, rows 8 (r / 4) + 4 (l / 32) + r % 4 in register r.  Only the
// column of D matters here (a tile's coefficients are summed in any order).
// ---------------------------------------------------------------------------
// Waves per SIMD the kernel is compiled for: at 3 (168 registers) the
// few spills are in the set-up, while the latency of a position's loads and
// product chains needs the third wave to hide it.
#ifndef MX_WPE
#define MX_WPE 3
#endif

typedef int mx_v4 __attribute__((ext_vector_type(4)));
typedef int mx_v16 __attribute__((ext_vector_type(16)));

// Row lane % 32 of H_32 (or of -H_32) as the A operand, entries +-1.
__device__ __forceinline__ mx_v4 mx_h32(int lane, bool neg)
{
    const u32 i = (u32)lane & 31u, k0 = 16u * ((u32)lane >> 5);
    mx_v4 r;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        u32 x = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const u32 odd = ((u32)__builtin_popcount(i & (k0 + 4u * d + b)) & 1u) ^ (neg ? 1u : 0u);
            x |= (odd ? 0xFFu : 0x01u) << (8 * b);
        }
        r[d] = (int)x;
    }
    return r;
}

// The B operand of K-step s for half h (= lane / 32) of a TW-wide tile with
// its top-left sample at p, row stride sp: samples 32s + 16h .. + 15 in
// row-major order, each minus 128 (x ^ 0x80).  Rows are read as aligned
// dwords (up to 3 bytes past a row's end: padded planes, current-frame
// slack) and aligned with v_alignbyte.
template <int TW>
__device__ __forceinline__ mx_v4 mx_operand(const pel *p, int sp, int s, int h)
{
    constexpr int RH = 16 / TW, NW = TW / 4;
    const pel *r0 = p + ((32 * s + 16 * h) / TW) * sp;
    u32 w[4];
#pragma unroll
    for (int q = 0; q < RH; q++) {
        const uintptr_t ad = (uintptr_t)(r0 + q * sp);
        const gu32p a4 = (gu32p)(ad & ~(uintptr_t)3);
        const u32 sh = (u32)(ad & 3);
        u32 d[NW + 1];
#pragma unroll
        for (int i = 0; i <= NW; i++) d[i] = a4[i];
#pragma unroll
        for (int i = 0; i < NW; i++) w[q * NW + i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh) ^ 0x80808080u;
    }
    return (mx_v4){(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
}

// A position's rows, fetched one position ahead of their use: the aligned
// dwords of K-step s, row q of batch b, and the low address bits of each
// batch's tile (row r starts (sh0 + r * sp) & 3 bytes into its first dword).
template <int TW, int KS, int NB>
struct MxRaw {
    u32 d[NB][KS][16 / TW][TW / 4 + 1];
    u32 sh0[NB];
};

template <int TW, int KS, int NB>
__device__ __forceinline__ void mx_fetch(MxRaw<TW, KS, NB> &R, const pel *pred, const int *toff, int sp, int hf,
                                         bool active)
{
    constexpr int RH = 16 / TW, NW = TW / 4;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const pel *p = pred + toff[b];
        R.sh0[b] = (u32)(uintptr_t)p & 3u;
#pragma unroll
        for (int s = 0; s < KS; s++)
#pragma unroll
            for (int q = 0; q < RH; q++)
#pragma unroll
                for (int i = 0; i <= NW; i++) R.d[b][s][q][i] = 0;
        if (active) {
#pragma unroll
            for (int s = 0; s < KS; s++)
#pragma unroll
                for (int q = 0; q < RH; q++) {
                    const uintptr_t ad = (uintptr_t)(p + ((32 * s + 16 * hf) / TW + q) * sp);
                    const gu32p a4 = (gu32p)(ad & ~(uintptr_t)3);
#pragma unroll
                    for (int i = 0; i <= NW; i++) R.d[b][s][q][i] = a4[i];
                }
        }
    }
}

// The B operand of K-step s of batch b from fetched rows (as mx_operand).
template <int TW, int KS, int NB>
__device__ __forceinline__ mx_v4 mx_operand_raw(const MxRaw<TW, KS, NB> &R, int b, int s, int sp, int hf)
{
    constexpr int RH = 16 / TW, NW = TW / 4;
    u32 w[4];
#pragma unroll
    for (int q = 0; q < RH; q++) {
        const u32 sh = (R.sh0[b] + (u32)(((32 * s + 16 * hf) / TW + q) * sp)) & 3u;
#pragma unroll
        for (int i = 0; i < NW; i++)
            w[q * NW + i] = __builtin_amdgcn_alignbyte(R.d[b][s][q][i + 1], R.d[b][s][q][i], sh) ^ 0x80808080u;
    }
    return (mx_v4){(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
}

// The MV-bits table me_search_init fills (GET_MVBITS, inter.c:2029-2058) in
// its closed form: 1 for 0, else 2 (floor(log2(|v| + 1)) + 1) -- computed
// rather than loaded, so that no load issued after a position's prefetch has
// to be waited for (loads complete in order).
__device__ __forceinline__ u32 mx_mvbits(int v)
{
    const int a = v < 0 ? -v : v;
    return a == 0 ? 1u : 2u * (u32)(32 - __clz(a + 1));
}

// Coefficient block m of the tiles whose K-step operands are B[0..KS): the
// chain of +-H_32 products started from acc.
template <int KS>
__device__ __forceinline__ mx_v16 mx_block(int m, const mx_v4 *B, mx_v4 Ap, mx_v4 An, mx_v16 acc)
{
#pragma unroll
    for (int s = 0; s < KS; s++)
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8((__builtin_popcount(m & s) & 1) ? An : Ap, B[s], acc, 0, 0, 0);
    return acc;
}

// Sub-pel search of one shape on the matrix cores.  A wave works on NB
// batches of 32 tiles (one per D column, lanes j and j + 32 hold its two
// K halves); a PU's tiles are consecutive: Tn <= 32 tiles -> 32 / Tn PUs per
// wave, Tn = 64 -> two waves per PU, Tn = 256 (128x128) -> the whole
// workgroup with NB = 2.  Every lane runs every position (the matrix
// instructions need the whole wave); lanes of PUs without a search result
// read nothing and write nothing.
template <int TT, int NB>
__device__ __forceinline__ void subpel_mx(const SubArgs &a, const SubShape &S, int blk, int job, int r)
{
    // the whole com_had tile (had_tiling): 8x8, 8x4, 16x8, 4x8, 8x16
    constexpr int TW = TT == 3 ? 16 : (TT == 4 ? 4 : 8), TH = TT == 2 ? 4 : (TT == 5 ? 16 : 8);
    constexpr int N = TW * TH, KS = N / 32;
    __shared__ int red[4];   // per wave: SATD of its tiles of a PU spanning waves
    const int Tn = S.lpu;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, jc = lane & 31, hf = lane >> 5;
    const int U = 128 * NB / Tn;                       // PUs per workgroup
    const int tw0 = wave * 32 * NB + jc;               // the lane's tile of batch 0 in the workgroup
    const int uw = tw0 / Tn;                           // its PU in the workgroup (the same for every batch)
    const int i = blk * U + uw;
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    const int cols = 128 / S.pw;
    const int x = (ctu % P.ctu_cols) * 128 + (i % cols) * S.pw, y = (ctu / P.ctu_cols) * 128 + (i / cols) * S.ph;
    const size_t slot = ((size_t)job * a.npu + S.base + i) * a.nref + r;
    const bool inside = x + S.pw <= P.pic_w && y + S.ph <= P.pic_h;
    const u64 key = key_merged(a.kl, a.keys, slot, job, r, a.nref);
    const bool active = inside && key < KEY_INIT;
    const bool lead = tw0 % Tn == 0 && hf == 0;
    const u64 k = active ? key - ((u64)KEY_BIAS_HI << 32) : 0;
    const int span = 2 * a.R + 1;
    const int idx = (int)(k & ((1ull << a.IB) - 1));
    const int dx = idx % span - a.R, dy = idx / span - a.R;

    const mx_v4 Ap = mx_h32(lane, false), An = mx_h32(lane, true);
    // per batch: the lane's tile offsets, and 0x8000 - T(org - 128) per block
    int toff_p[NB];
    mx_v16 init[NB][KS];
    const pel *org = P.cur + (size_t)y * P.pic_w + x;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int t = (tw0 + 32 * b) % Tn;
        const int tx = (t % S.tpr) * TW, ty = (t / S.tpr) * TH;
        toff_p[b] = ty * P.ref_stride + tx;
        mx_v4 B[KS];
#pragma unroll
        for (int s = 0; s < KS; s++) B[s] = active ? mx_operand<TW>(org + ty * P.pic_w + tx, P.pic_w, s, hf) : (mx_v4)0;
#pragma unroll
        for (int m = 0; m < KS; m++) init[b][m] = (mx_v16)0x8000 - mx_block<KS>(m, B, Ap, An, (mx_v16)0);
    }
    // com_had of the PU from its fetched rows (every lane of the PU gets the sum)
    typedef MxRaw<TW, KS, NB> Raw;
    auto had = [&](const Raw &R) -> u32 {
        int v = 0;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            mx_v4 B[KS];
#pragma unroll
            for (int s = 0; s < KS; s++) B[s] = mx_operand_raw<TW, KS, NB>(R, b, s, P.ref_stride, hf);
            u32 sa = 0;
#pragma unroll
            for (int m = 0; m < KS; m++) {
                const mx_v16 d = mx_block<KS>(m, B, Ap, An, init[b][m]);
#pragma unroll
                for (int q = 0; q < 16; q++) sa = __builtin_amdgcn_sad_u16((u32)d[q], 0x8000u, sa);
                // one block's products live at a time (registers for occupancy)
                __builtin_amdgcn_sched_barrier(0);
            }
            {   // plus the tile's other K half (lane ^ 32), without LDS
                const auto w = __builtin_amdgcn_permlane32_swap(sa, sa, false, false);
                sa = w[0] + w[1];
            }
            const int st = (int)sa;
            v += TT == 1 ? (st + 2) >> 2 : (TT == 2 || TT == 4 ? me_satd_norm32(st) : me_satd_norm128(st));
        }
        if (Tn <= 32) return (u32)group_sum(v, Tn);
        v = group_sum(v, 32);
        if (lane == 0) red[wave] = v;
        __syncthreads();
        const int wpu = Tn / (32 * NB);   // waves per PU
        const int w0 = wave & ~(wpu - 1);
        v = 0;
        for (int q = 0; q < wpu; q++) v += red[w0 + q];
        __syncthreads();
        return (u32)v;
    };
    const u8 *tb = a.tab;
    const int cmx = dx * 4, cmy = dy * 4;
    // positions as in subpel_body: 0 the integer centre, 1..8 half-pel,
    // 9..16 quarter-pel around the best of 0..8; key (cost << 5) | p
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
    auto pred_at = [&](int p, int hx, int hy) -> const pel * {
        int mx, my;
        pos(p, hx, hy, mx, my);
        return p == 0 ? plane_ptr(P, r, 0, x + dx, y + dy) : qpel_ptr(P, r, mx + (x << 2), my + (y << 2));
    };
    // a phase's positions are known in advance: the rows of position p + 1
    // are fetched before position p's transforms
    auto phase = [&](int p0, int n, int hx, int hy) -> u64 {
        u64 best = ~0ull;
        Raw R;
        mx_fetch<TW, KS, NB>(R, pred_at(p0, hx, hy), toff_p, P.ref_stride, hf, active);
        for (int p = p0; p < p0 + n; p++) {
            Raw Rn;
            const bool more = p + 1 < p0 + n;
            if (more) mx_fetch<TW, KS, NB>(Rn, pred_at(p + 1, hx, hy), toff_p, P.ref_stride, hf, active);
            int mx, my;
            pos(p, hx, hy, mx, my);
            const u32 hv = had(R);
            const u64 c = (u64)(a.lambda * (mx_mvbits(mx) + mx_mvbits(my))) + ((u64)hv << 16);
            const u64 kk = (c << 5) | (u64)p;
            if (kk < best) best = kk;
            if (more) R = Rn;
        }
        return best;
    };
    const u64 k1 = phase(0, 9, cmx, cmy);
    int hx, hy;
    pos((int)(k1 & 31), cmx, cmy, hx, hy);
    const u64 k2 = phase(9, 8, hx, hy);
    const u64 kb = k2 < k1 ? k2 : k1;
    int bx, by;
    pos((int)(kb & 31), hx, hy, bx, by);
    if (lead) {
        me_pu_result_t out;
        out.mv_int[0] = out.mv_int[1] = out.mv_sub[0] = out.mv_sub[1] = 0;
        out.cost_int = out.cost_sub = ~0ull;
        out.cost_final = 0xFFFFFFFFu;
        out.valid = 0;
        if (active) {
            const u64 best = kb >> 5;
            out.mv_int[0] = (s16)dx;
            out.mv_int[1] = (s16)dy;
            out.cost_int = k >> a.IB;
            out.mv_sub[0] = (s16)bx;
            out.mv_sub[1] = (s16)by;
            out.cost_sub = best;
            const int n = r < a.n0 ? a.n0 : a.nref - a.n0, ri = r < a.n0 ? r : r - a.n0;
            const int rb = n <= 1 ? 0 : (ri < n - 1 ? ri + 1 : n - 1);
            const u64 cf = best + (u64)(a.lambda * (u32)(rb + 3));
            out.cost_final = (u32)((cf + (1u << 15)) >> 16);
            out.valid = 1;
        }
        a.res[slot] = out;
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MX_WPE))) void k_subpel_mx(SubArgs a)
{
    if (a.js.n_jobs == 0) return;   // a warm-up launch (me_subpel_preload)
    const unsigned nx = gridDim.x, ny = gridDim.y;
    const unsigned w = me_xcd_order(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
    const int bxi = (int)(w % nx), job = (int)((w / nx) % ny), r = (int)(w / (nx * ny));
    const int si = __builtin_amdgcn_readfirstlane(a.blk_shape[bxi]);
    const SubShape S = a.shp[si];
    const int blk = bxi - S.blk0;
    switch (S.tt) {
    case 1:
        if (S.G == 2) subpel_mx<1, 2>(a, S, blk, job, r);
        else subpel_mx<1, 1>(a, S, blk, job, r);
        break;
    case 2: subpel_mx<2, 1>(a, S, blk, job, r); break;
    case 3: subpel_mx<3, 1>(a, S, blk, job, r); break;
    case 4: subpel_mx<4, 1>(a, S, blk, job, r); break;
    default: subpel_mx<5, 1>(a, S, blk, job, r); break;
    }
}
#endif

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
#if BIT_DEPTH == 8
    if (a.mx) {
        hipLaunchKernelGGL(k_subpel_mx, dim3(blocks, jobs, refs), dim3(256), 0, st, a);
        ME_CHECK(hipGetLastError());
        return 0;
    }
#endif
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
#if BIT_DEPTH == 8
    hipLaunchKernelGGL(k_subpel_mx, dim3(1), dim3(256), 0, st, a);
    ME_CHECK(hipGetLastError());
#endif
    return 0;
}
