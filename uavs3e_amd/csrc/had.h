// Device helpers of the sub-pel search shared by subpel.hip and tz.hip:
// search patterns, lane-group sums, plane addressing (com_mc_blk_luma_pointer)
// and the com_had tiles as in-register Walsh-Hadamard transforms.
#pragma once
#include "frame.h"

// Search patterns of me_sub_pel_search in evaluation order, packed 2 bits
// per point (field q = offset + 1, in half-pel resp. quarter-pel steps):
//   half-pel    {-2,0} {-2,2} {0,2} {2,2} {2,0} {2,-2} {0,-2} {-2,-2}  me.c:415-417
//   quarter-pel {-1,0} {0,1} {1,0} {0,-1} {-1,1} {1,1} {-1,-1} {1,-1}  me.c:460-462
#define ME_SUBPEL_HALF_X 0x1A90u
#define ME_SUBPEL_HALF_Y 0x01A9u
#define ME_SUBPEL_QUARTER_X 0x8864u
#define ME_SUBPEL_QUARTER_Y 0x0A19u
__device__ __forceinline__ int pat(u32 packed, int q) { return (int)((packed >> (2 * q)) & 3u) - 1; }

// Sum over aligned groups of w lanes (w a power of two, w <= 64), result in
// every lane of the group; w is uniform.  Quad permutes and row mirrors run
// as DPP moves, the 32-lane step as ds_swizzle.
__device__ __forceinline__ int group_sum(int v, int w)
{
    if (w >= 2) v += __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
    if (w >= 4) v += __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
    if (w >= 8) v += __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false);   // row_half_mirror
    if (w >= 16) v += __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false);  // row_mirror
    if (w >= 32) v += __builtin_amdgcn_ds_swizzle(v, 0x401F);               // lane ^ 16
    if (w >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

__device__ __forceinline__ const pel *plane_ptr(const PicDesc &P, int ref, int idx, int x, int y)
{
    return P.planes[ref] + (size_t)idx * P.plane_elems + (size_t)(y + ME_PAD) * P.ref_stride + (x + ME_PAD);
}

// com_mc_blk_luma_pointer (com_mc.c:89-101).
__device__ __forceinline__ const pel *qpel_ptr(const PicDesc &P, int ref, int qx, int qy)
{
    int ix = me_clip3(-ME_MAX_CU - 4, P.pic_w + 4, qx >> 2);
    int iy = me_clip3(-ME_MAX_CU - 4, P.pic_h + 4, qy >> 2);
    return plane_ptr(P, ref, (qy & 3) * 4 + (qx & 3), ix, iy);
}

// N consecutive samples from an arbitrary address as int, read with aligned
// dword loads and v_alignbyte (may read up to 3 bytes past the last sample,
// which stays inside the padded planes and the slack of the current frame).
template <int N>
__device__ __forceinline__ void load_row(const pel *p, int *out)
{
    const uintptr_t ad = (uintptr_t)p;
    const u32 *q = (const u32 *)(ad & ~(uintptr_t)3);
    const u32 sh = (u32)(ad & 3);
    constexpr int NW = (N * (int)sizeof(pel) + 3) / 4;
    u32 d[NW + 1];
#pragma unroll
    for (int i = 0; i <= NW; i++) d[i] = q[i];
    u32 w[NW];
#pragma unroll
    for (int i = 0; i < NW; i++) w[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
#pragma unroll
    for (int i = 0; i < N; i++) {
#if BIT_DEPTH == 8
        out[i] = (int)((w[i >> 2] >> (8 * (i & 3))) & 0xFF);
#else
        out[i] = (int)((w[i >> 1] >> (16 * (i & 1))) & 0xFFFF);
#endif
    }
}

template <int N>
__device__ __forceinline__ void fwht(int *v, int step)
{
#pragma unroll
    for (int hs = 1; hs < N; hs <<= 1)
#pragma unroll
        for (int i = 0; i < N; i += hs << 1)
#pragma unroll
            for (int j = i; j < i + hs; j++) {
                int p = v[j * step], q = v[(j + hs) * step];
                v[j * step] = p + q;
                v[(j + hs) * step] = p - q;
            }
}

// Tiles on packed 16-bit lanes, two samples per operation.  The last
// butterfly is never formed: for the last stage's pairs (a, b),
// |a + b| + |a - b| = 2 max(|a|, |b|), and a Walsh-Hadamard transform's
// stages commute, so the two samples of a dword take that last stage.  Every
// other intermediate of a transform of N samples is at most N/2 times the
// largest residual: 8 bits, any tile and the 16x8 / 8x16 halves (residuals
// within +-510): at most 16320; 10 bits, tiles up to 8x8 (residuals within
// +-1023): at most 32736 -- both inside a signed 16-bit lane.  The 10-bit
// 16x8 / 8x16 halves (residuals to +-2046) take four stages on 16-bit lanes
// (at most 32736) and their fifth on 32-bit lanes (had_pk_sum_half).
typedef short s16x2 __attribute__((ext_vector_type(2)));

// NW dwords of the samples starting at p (any alignment).
template <int NW>
__device__ __forceinline__ void load_words(const pel *p, u32 *w)
{
    const uintptr_t ad = (uintptr_t)p;
    const u32 *q = (const u32 *)(ad & ~(uintptr_t)3);
    const u32 sh = (u32)(ad & 3);
    u32 d[NW + 1];
#pragma unroll
    for (int i = 0; i <= NW; i++) d[i] = q[i];
#pragma unroll
    for (int i = 0; i < NW; i++) w[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
}

// One row of TW samples (TW = 4, 8) as TW/2 pairs of 16-bit lanes.  8 bits:
// v[2j] = (p[4j], p[4j+2]), v[2j+1] = (p[4j+1], p[4j+3]); 10 bits: v[j] =
// (p[2j], p[2j+1]), straight from the samples' dwords.
#if BIT_DEPTH == 8
// 8 bits: the pairs of dword j straight from the aligned dwords d[j],
// d[j + 1] around it, for a row starting sh bytes into d[0]: v_perm picks
// bytes sh, sh + 2 (s0) and sh + 1, sh + 3 (s1) of {d[j + 1]:d[j]} and
// zeroes the high bytes, so one instruction aligns and widens a pair.
__device__ __forceinline__ u32 perm_sel(u32 sh) { return sh * 0x10001u + 0x0C020C00u; }
template <int NW>
__device__ __forceinline__ void pairs_from(const u32 *d, u32 s0, s16x2 *v)
{
    const u32 s1 = s0 + 0x10001u;
#pragma unroll
    for (int j = 0; j < NW; j++) {
        v[2 * j] = __builtin_bit_cast(s16x2, __builtin_amdgcn_perm(d[j + 1], d[j], s0));
        v[2 * j + 1] = __builtin_bit_cast(s16x2, __builtin_amdgcn_perm(d[j + 1], d[j], s1));
    }
}
#endif

template <int TW>
__device__ __forceinline__ void pairs_row(const pel *p, s16x2 *v)
{
#if BIT_DEPTH == 8
    const uintptr_t ad = (uintptr_t)p;
    const u32 *q = (const u32 *)(ad & ~(uintptr_t)3);
    u32 d[TW / 4 + 1];
#pragma unroll
    for (int i = 0; i <= TW / 4; i++) d[i] = q[i];
    pairs_from<TW / 4>(d, perm_sel((u32)(ad & 3)), v);
#else
    u32 w[TW / 2];
    load_words<TW / 2>(p, w);
#pragma unroll
    for (int j = 0; j < TW / 2; j++) v[j] = __builtin_bit_cast(s16x2, w[j]);
#endif
}

// The same from a dword-aligned global address q plus a byte shift sh (the
// rows of a block at a stride of whole dwords share sh): one address add
// per row and global (not flat) loads.
typedef const __attribute__((address_space(1))) u32 *gu32p;
template <int TW>
__device__ __forceinline__ void pairs_row_al(gu32p q, u32 sh, s16x2 *v)
{
    constexpr int NW = TW * (int)sizeof(pel) / 4;
    u32 d[NW + 1];
#pragma unroll
    for (int i = 0; i <= NW; i++) d[i] = q[i];
#if BIT_DEPTH == 8
    pairs_from<NW>(d, perm_sel(sh), v);
#else
#pragma unroll
    for (int j = 0; j < NW; j++) v[j] = __builtin_bit_cast(s16x2, __builtin_amdgcn_alignbyte(d[j + 1], d[j], sh));
#endif
}

// The residual org - cur of one row as pairs.
template <int TW>
__device__ __forceinline__ void res_pairs(const pel *org, const pel *cur, s16x2 *v)
{
    s16x2 o[TW / 2], c[TW / 2];
    pairs_row<TW>(org, o);
    pairs_row<TW>(cur, c);
#pragma unroll
    for (int j = 0; j < TW / 2; j++) v[j] = o[j] - c[j];
}

template <int N>
__device__ __forceinline__ void fwht_pk(s16x2 *v, int step)
{
#pragma unroll
    for (int hs = 1; hs < N; hs <<= 1)
#pragma unroll
        for (int i = 0; i < N; i += hs << 1)
#pragma unroll
            for (int j = i; j < i + hs; j++) {
                const s16x2 p = v[j * step], q = v[(j + hs) * step];
                v[j * step] = p + q;
                v[(j + hs) * step] = p - q;
            }
}

// Sum of |coef| of the TW x TH transform whose rows are given as pairs.
// The pair stage reads the low half of every pair with a bias of 0x8000
// (mod 2^16).  The low halves of the pairs form a transform of their own
// (only the pair stage mixes the halves), and input 0 enters every output of
// a Walsh-Hadamard transform with sign +1, so 0x8000 added to the low half
// of v[0] before the transform reaches the low half of every pair.  PRE:
// the caller's pairs already carry it (org_pairs puts it into the current
// block's first pair, see pk_bias); otherwise it is added here.
template <int TW, int TH, bool PRE = false>
__device__ __forceinline__ int had_pk_sum(s16x2 *v)
{
    constexpr int K = TW / 2;
    if (!PRE) v[0] += (s16x2){(short)-32768, 0};
#pragma unroll
    for (int y = 0; y < TH; y++) fwht_pk<K>(v + y * K, 1);
#pragma unroll
    for (int k = 0; k < K; k++) fwht_pk<TH>(v + k, K);
    u32 s = 0;
#if BIT_DEPTH == 8
    // The pair stage formed (every |lo +- hi| <= 32640): t = (lo + hi, lo -
    // hi), each half offset by the bias, in one packed multiply-add, and
    // |t.lo| + |t.hi| by v_sad_u16 against 0x8000.
    const s16x2 pm = {1, -1};
#pragma unroll
    for (int i = 0; i < K * TH; i++) {
        const s16x2 t = v[i].yy * pm + v[i].xx;
        s = __builtin_amdgcn_sad_u16(__builtin_bit_cast(u32, t), 0x80008000u, s);
    }
#else
    // 10 bits: lo + hi could leave 16 bits.  With L = lo + 0x8000 and both
    // halves within +-32736, |lo + hi| + |lo - hi| (= 2 max(|lo|, |hi|)) is
    // |L - (0x8000 - hi)| + |L - (0x8000 + hi)| on unsigned 16-bit lanes:
    // one v_sad_u16 of (L, L) against (0x8000 - hi, 0x8000 + hi).
    const s16x2 mp = {-1, 1}, bias = {(short)-32768, (short)-32768};
#pragma unroll
    for (int i = 0; i < K * TH; i++) {
        const s16x2 b = v[i].yy * mp + bias;
        s = __builtin_amdgcn_sad_u16(__builtin_bit_cast(u32, v[i].xx), __builtin_bit_cast(u32, b), s);
    }
#endif
    return (int)s;
}

// had_pk_sum<8, 8> of the half of a 16x8 / 8x16 tile.  At 10 bits the half
// residual reaches +-2046, so the column transform's last stage (rows y and
// y + 4, at most 65472) and the pair stage run on 32-bit lanes: for the
// pairs P, Q of rows y and y + 4, the four coefficients of their last two
// stages sum to 2 max(|p.lo + q.lo|, |p.hi + q.hi|) +
// 2 max(|p.lo - q.lo|, |p.hi - q.hi|).
template <bool PRE = false>
__device__ __forceinline__ int had_pk_sum_half(s16x2 *v)
{
#if BIT_DEPTH == 8
    return had_pk_sum<8, 8, PRE>(v);
#else
#pragma unroll
    for (int y = 0; y < 8; y++) fwht_pk<4>(v + y * 4, 1);
    // the column stages within rows 0-3 and within rows 4-7
#pragma unroll
    for (int k = 0; k < 4; k++) {
        fwht_pk<4>(v + k, 4);
        fwht_pk<4>(v + k + 16, 4);
    }
    int s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const s16x2 p = v[i], q = v[i + 16];   // rows y and y + 4 of column pair k
        const int plo = p.x, phi = p.y, qlo = q.x, qhi = q.y;
        const int a = plo + qlo, b = phi + qhi, c = plo - qlo, d = phi - qhi;
        s += __builtin_elementwise_max(__builtin_elementwise_max(a, -a), __builtin_elementwise_max(b, -b)) +
             __builtin_elementwise_max(__builtin_elementwise_max(c, -c), __builtin_elementwise_max(d, -d));
    }
    return 2 * s;
#endif
}

// Sum of |coef| of the 2-D Walsh-Hadamard transform of a TW x TH residual.
template <int TW, int TH>
__device__ __forceinline__ int had_abs(const pel *org, int so, const pel *cur, int sc)
{
    s16x2 v[TH * TW / 2];
#pragma unroll
    for (int y = 0; y < TH; y++) res_pairs<TW>(org + y * so, cur + y * sc, v + y * (TW / 2));
    return had_pk_sum<TW, TH>(v);
}

// A 16-point WHT is the 8-point WHT of the half sums and of the half
// differences, so a 16x8 (TT 3) or 8x16 (TT 5) tile is two 8x8 transforms:
// half 0 of the sums, half 1 of the differences of its two 8x8 halves.
// Returns that half's sum of |coef|.
template <int TT>
__device__ __forceinline__ int had_half(const pel *org, int so, const pel *cur, int sc, int half)
{
    const int sg = half ? -1 : 1;
    const int o2 = TT == 3 ? 8 : 8 * so, c2 = TT == 3 ? 8 : 8 * sc;
    s16x2 v[32];
    const s16x2 sgv = (s16x2)(short)sg;
#pragma unroll
    for (int y = 0; y < 8; y++) {
        s16x2 b[4];
        res_pairs<8>(org + y * so, cur + y * sc, v + y * 4);
        res_pairs<8>(org + y * so + o2, cur + y * sc + c2, b);
#pragma unroll
        for (int k = 0; k < 4; k++) v[y * 4 + k] += b[k] * sgv;
    }
    return had_pk_sum_half(v);
}

// One normalised cost_satd tile (com_cost.c:259, 340, 467, 586, 649, 708)
// for the tile types done by one lane (0, 1, 2, 4).
template <int TT>
__device__ __forceinline__ int satd_t(const pel *org, int so, const pel *cur, int sc)
{
    if (TT == 0) return (had_abs<4, 4>(org, so, cur, sc) + 1) >> 1;
    if (TT == 1) return (had_abs<8, 8>(org, so, cur, sc) + 2) >> 2;
    if (TT == 2) return me_satd_norm32(had_abs<8, 4>(org, so, cur, sc));
    return me_satd_norm32(had_abs<4, 8>(org, so, cur, sc));
}

// The same costs with the current block's pairs made once (a sub-pel search
// compares one current block with 17 predictions): ORG_PAIRS<TT> pairs per
// lane, from org_pairs<TT>, then satd_pre<TT> / had_pre<TT> per prediction.
template <int TT> struct tile_dims { static constexpr int W = TT == 0 || TT == 4 ? 4 : 8, H = TT == 0 || TT == 2 ? 4 : 8; };
template <int TT> constexpr int org_pairs_n() { return TT == 3 || TT == 5 ? 32 : tile_dims<TT>::W * tile_dims<TT>::H / 2; }
// whether the pair stage of TT's tiles takes the low-half bias (the 10-bit
// 16x8 / 8x16 halves end on 32-bit lanes without it)
template <int TT> constexpr bool pk_bias() { return BIT_DEPTH == 8 || !(TT == 3 || TT == 5); }

// For TT 3/5 the pairs of the tile's half sum (half 0) or difference (half 1)
// of its two 8x8 halves (at most 2046 at 10 bits).
template <int TT>
__device__ __forceinline__ void org_pairs(const pel *org, int so, int half, s16x2 *o)
{
    if (TT == 3 || TT == 5) {
        const s16x2 sgv = (s16x2)(short)(half ? -1 : 1);
        const int o2 = TT == 3 ? 8 : 8 * so;
#pragma unroll
        for (int y = 0; y < 8; y++) {
            s16x2 b[4];
            pairs_row<8>(org + y * so, o + y * 4);
            pairs_row<8>(org + y * so + o2, b);
#pragma unroll
            for (int k = 0; k < 4; k++) o[y * 4 + k] += b[k] * sgv;
        }
    } else {
        constexpr int W = tile_dims<TT>::W, H = tile_dims<TT>::H;
#pragma unroll
        for (int y = 0; y < H; y++) pairs_row<W>(org + y * so, o + y * (W / 2));
    }
    if (pk_bias<TT>()) o[0] += (s16x2){(short)-32768, 0};
}

template <int TT>
__device__ __forceinline__ int had_pre(const s16x2 *o, const pel *cur, int sc, int half)
{
    s16x2 v[org_pairs_n<TT>()];
    if ((sc * (int)sizeof(pel)) % 4 == 0) {
        // rows a whole number of dwords apart: one alignment for all rows
        const uintptr_t ad = (uintptr_t)cur;
        const u32 sh = (u32)(ad & 3);
        const gu32p q0 = (gu32p)(ad & ~(uintptr_t)3);
        const int sw = sc * (int)sizeof(pel) / 4;
        if (TT == 3 || TT == 5) {
            const s16x2 sgv = (s16x2)(short)(half ? -1 : 1);
            const int c2 = TT == 3 ? 8 * (int)sizeof(pel) / 4 : 8 * sw;
#pragma unroll
            for (int y = 0; y < 8; y++) {
                s16x2 a[4], b[4];
                pairs_row_al<8>(q0 + y * sw, sh, a);
                pairs_row_al<8>(q0 + y * sw + c2, sh, b);
#pragma unroll
                for (int k = 0; k < 4; k++) v[y * 4 + k] = o[y * 4 + k] - (a[k] + b[k] * sgv);
            }
            return had_pk_sum_half<pk_bias<TT>()>(v);
        } else {
            constexpr int W = tile_dims<TT>::W, H = tile_dims<TT>::H;
#pragma unroll
            for (int y = 0; y < H; y++) {
                s16x2 c[W / 2];
                pairs_row_al<W>(q0 + y * sw, sh, c);
#pragma unroll
                for (int k = 0; k < W / 2; k++) v[y * (W / 2) + k] = o[y * (W / 2) + k] - c[k];
            }
            return had_pk_sum<W, H, true>(v);
        }
    }
    if (TT == 3 || TT == 5) {
        const s16x2 sgv = (s16x2)(short)(half ? -1 : 1);
        const int c2 = TT == 3 ? 8 : 8 * sc;
#pragma unroll
        for (int y = 0; y < 8; y++) {
            s16x2 a[4], b[4];
            pairs_row<8>(cur + y * sc, a);
            pairs_row<8>(cur + y * sc + c2, b);
#pragma unroll
            for (int k = 0; k < 4; k++) v[y * 4 + k] = o[y * 4 + k] - (a[k] + b[k] * sgv);
        }
        return had_pk_sum_half<pk_bias<TT>()>(v);
    } else {
        constexpr int W = tile_dims<TT>::W, H = tile_dims<TT>::H;
#pragma unroll
        for (int y = 0; y < H; y++) {
            s16x2 c[W / 2];
            pairs_row<W>(cur + y * sc, c);
#pragma unroll
            for (int k = 0; k < W / 2; k++) v[y * (W / 2) + k] = o[y * (W / 2) + k] - c[k];
        }
        return had_pk_sum<W, H, true>(v);
    }
}

// satd_t (TT 0, 1, 2, 4) from the precomputed pairs.
template <int TT>
__device__ __forceinline__ int satd_pre(const s16x2 *o, const pel *cur, int sc)
{
    const int s = had_pre<TT>(o, cur, sc, 0);
    if (TT == 0) return (s + 1) >> 1;
    if (TT == 1) return (s + 2) >> 2;
    return me_satd_norm32(s);
}

