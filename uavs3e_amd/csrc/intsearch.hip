// Integer full search of every PU of a set of CTUs (gfx950).
//
//   integer cost  (sad << (16-(bd-8))) + (bits_x+bits_y)*lambda   me.c:119,151-187
//   window        get_max_search_range (is_padding = 1)           me.c:502-527
//   full search   raster y-major, first strict minimum wins        me.c:150-190
//
// This file is compiled once per workgroup size (ME_IS_WAVES = 1, 2, 4, 8) so
// that the instances build in parallel.
#include "frame.h"

#include <string.h>

#ifndef ME_IS_WAVES
#define ME_IS_WAVES 4
#endif
// One pass over the displacements updates all 58 node classes of a lane.
// (Two passes, 33 small and then 25 large classes with half the live node
// states, measured slower: the leaves' SAD instructions then run twice.)

// A step of the search covers four displacements dxb .. dxb+3 of one row
// (NKEY).  Values of NK displacements of one node: at 8 bits all four are
// packed as the u16 lanes of one 64-bit value (bits 0-15 d0 ... 48-63 d3) --
// a sum of at most 256 samples stays below 2^16, so a sum of two values is
// one 64-bit add (v_lshl_add_u64) that never carries between lanes; at 10
// bits one displacement per u32 (a step then holds four of them).
#define NKEY 4
#if BIT_DEPTH == 8
#define NK 4
struct Val {
    u64 v;
    __device__ __forceinline__ u32 lo() const { return (u32)v; }
    __device__ __forceinline__ u32 hi() const { return (u32)(v >> 32); }
    __device__ __forceinline__ u32 get(int k) const
    {
        const u32 w = k < 2 ? lo() : hi();
        return (k & 1) ? (w >> 16) : (w & 0xFFFFu);
    }
};
__device__ __forceinline__ Val mkval(u32 lo, u32 hi) { return {((u64)hi << 32) | lo}; }
__device__ __forceinline__ Val operator+(Val a, Val b) { return {a.v + b.v}; }
__device__ __forceinline__ Val vsel(bool c, Val a, Val b) { return c ? a : b; }
#else
#define NK 1
struct Val {
    u32 v;
    __device__ __forceinline__ u32 get(int) const { return v; }
};
__device__ __forceinline__ Val operator+(Val a, Val b) { return {a.v + b.v}; }
__device__ __forceinline__ Val vsel(bool c, Val a, Val b) { return c ? a : b; }
#endif

// Uniform key parts of the step's four displacements.  A displacement
// outside the block's range gets a high word that can never win.
// The key parts are vector values rather than arrays: an array member let
// the compiler keep them in private memory and reload them as pairs.
typedef u32 u32xk __attribute__((ext_vector_type(NKEY)));
struct KeyCtx {
    u32xk khi, klo;
    int dx[NKEY];
    int dy, sh;
};


typedef u32 u32x2 __attribute__((ext_vector_type(2)));
typedef u32 u32x2a __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ bool node_valid(int px, int py, int dx, int dy, int wlim, int hlim)
{
    // is_padding = 1 window of get_max_search_range (me.c:508-512):
    // -127 - x <= dx <= W-1-x and the same for y (the +-T limit is uniform).
    return (u32)(px + dx + 127) <= (u32)wlim && (u32)(py + dy + 127) <= (u32)hlim;
}

// One candidate: key = ((sad << sh) + khi) : klo, kept as the bits of a
// double; st = min(st, key).  SH0: sh is known to be 0 (every configuration
// except 8-bit with R > 127), so the add can take a 16-bit lane directly.
template <bool EDGE, bool SH0>
__device__ __forceinline__ void upd1(double &st, u32 sad, const KeyCtx &c, int k, int px, int py, int wlim, int hlim)
{
    u32 hi = (SH0 ? sad : (sad << c.sh)) + c.khi[k];
    if (EDGE) {
        if (!node_valid(px, py, c.dx[k], c.dy, wlim, hlim)) hi |= KEY_INVALID_HI;
    }
#if BIT_DEPTH == 8
    // a two-word vector: measured fastest at 8 bits
    u32x2 key = {c.klo[k], hi};
    st = me_vmin64(st, __builtin_bit_cast(double, key));
#else
    // a 64-bit or: keeps the 10-bit kernel free of private-memory traffic
    st = me_vmin64(st, __builtin_bit_cast(double, ((u64)hi << 32) | c.klo[k]));
#endif
}

// Cross-lane partner values.  Lane = ry*8 + rx (16x16 regions of the CTU).
//   xor 1, xor 2 : DPP quad permutes        (x pairs, x quads)
//   x octet      : DPP row_half_mirror       (the other x quad of a row of 8)
//   xor 8        : DPP row_ror:8             (y pairs)
//   xor 16       : ds_swizzle xor mode       (y quads)
//   xor 32       : v_permlane32_swap         (the other half of the CTU; only
//                  pair sums over the halves are needed, see hsum2)
// Every source lane of these four patterns exists, so bound_ctrl (zero for a
// missing source) never applies; with it the compiler folds each move into
// the add that consumes it (v_add_u32_dpp).
template <int CTRL> __device__ __forceinline__ u32 dppx(u32 v)
{
    return (u32)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ u32 px1(u32 v) { return dppx<0xB1>(v); }
__device__ __forceinline__ u32 px2(u32 v) { return dppx<0x4E>(v); }
__device__ __forceinline__ u32 px4(u32 v) { return dppx<0x141>(v); }
__device__ __forceinline__ u32 py1(u32 v) { return dppx<0x128>(v); }
__device__ __forceinline__ u32 py2(u32 v) { return (u32)__builtin_amdgcn_ds_swizzle((int)v, 0x401F); }
// Pair sums over the wave's halves for two values at once: lanes 0..31 get
// a + a[lane + 32], lanes 32..63 get b + b[lane - 32].  v_permlane32_swap
// exchanges lanes 32..63 of its first operand with lanes 0..31 of its
// second, so afterwards the lower halves of the two results hold a's two
// halves and the upper halves b's.
__device__ __forceinline__ u32 hsum2(u32 a, u32 b)
{
    const auto w = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    return w[0] + w[1];
}
// A value plus its x- or y-partner's: adds on 32-bit halves, which take
// the partner's value as a DPP operand.
#if BIT_DEPTH == 8
__device__ __forceinline__ Val vaddx1(Val a) { return mkval(a.lo() + px1(a.lo()), a.hi() + px1(a.hi())); }
__device__ __forceinline__ Val vaddy1(Val a) { return mkval(a.lo() + py1(a.lo()), a.hi() + py1(a.hi())); }
#else
__device__ __forceinline__ Val vaddx1(Val a) { return {a.v + px1(a.v)}; }
__device__ __forceinline__ Val vaddy1(Val a) { return {a.v + py1(a.v)}; }
#endif

// Region-level composition from the 16 leaves L[r][c] (4x4 blocks, row r,
// column c of the lane's 16x16 region).
struct Region {
    Val h8[4][2], v8[2][4], q[2][2], h16[4], v16[4], h16x8[2], v8x16[2], s16;
    __device__ __forceinline__ void build(const Val L[4][4])
    {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c2 = 0; c2 < 2; c2++) h8[r][c2] = L[r][2 * c2] + L[r][2 * c2 + 1];
#pragma unroll
        for (int r2 = 0; r2 < 2; r2++)
#pragma unroll
            for (int c = 0; c < 4; c++) v8[r2][c] = L[2 * r2][c] + L[2 * r2 + 1][c];
#pragma unroll
        for (int r2 = 0; r2 < 2; r2++)
#pragma unroll
            for (int c2 = 0; c2 < 2; c2++) q[r2][c2] = h8[2 * r2][c2] + h8[2 * r2 + 1][c2];
#pragma unroll
        for (int r = 0; r < 4; r++) h16[r] = h8[r][0] + h8[r][1];
#pragma unroll
        for (int c = 0; c < 4; c++) v16[c] = v8[0][c] + v8[1][c];
#pragma unroll
        for (int i = 0; i < 2; i++) { h16x8[i] = q[i][0] + q[i][1]; v8x16[i] = q[0][i] + q[1][i]; }
        s16 = h16x8[0] + h16x8[1];
    }
};

// Pass S: the 33 PUs inside the lane's 16x16 region.
//   st[0..7] 4x8 (c, r2)   st[8..15] 8x4 (c2, r)  st[16..19] 4x16 (c)
//   st[20..23] 16x4 (r)    st[24..27] 8x8 (c2,r2) st[28..29] 8x16 (c2)
//   st[30..31] 16x8 (r2)   st[32] 16x16
// The region's values hold displacements kb .. kb+NK-1 of the step.
#define N_SMALL 33
template <bool EDGE, bool SH0>
__device__ __forceinline__ void pass_small(const Region &g, double *st, const KeyCtx &c, int kb, int px, int py, int wlim,
                                           int hlim)
{
#pragma unroll
    for (int k = 0; k < NK; k++) {
#define U(I, V, X_, Y_) upd1<EDGE, SH0>(st[I], (V).get(k), c, kb + k, X_, Y_, wlim, hlim)
#pragma unroll
        for (int cc = 0; cc < 4; cc++)
#pragma unroll
            for (int r2 = 0; r2 < 2; r2++) U(cc * 2 + r2, g.v8[r2][cc], px + 4 * cc, py + 8 * r2);
#pragma unroll
        for (int c2 = 0; c2 < 2; c2++)
#pragma unroll
            for (int r = 0; r < 4; r++) U(8 + c2 * 4 + r, g.h8[r][c2], px + 8 * c2, py + 4 * r);
#pragma unroll
        for (int cc = 0; cc < 4; cc++) U(16 + cc, g.v16[cc], px + 4 * cc, py);
#pragma unroll
        for (int r = 0; r < 4; r++) U(20 + r, g.h16[r], px, py + 4 * r);
#pragma unroll
        for (int c2 = 0; c2 < 2; c2++)
#pragma unroll
            for (int r2 = 0; r2 < 2; r2++) U(24 + c2 * 2 + r2, g.q[r2][c2], px + 8 * c2, py + 8 * r2);
#pragma unroll
        for (int c2 = 0; c2 < 2; c2++) U(28 + c2, g.v8x16[c2], px + 8 * c2, py);
#pragma unroll
        for (int r2 = 0; r2 < 2; r2++) U(30 + r2, g.h16x8[r2], px, py + 8 * r2);
        U(32, g.s16, px, py);
#undef U
    }
}

// Key parts of one displacement per lane (k chosen by lane bits).
struct KeyV {
    u32 hi, lo;
    int dx;
};

template <bool EDGE, bool SH0>
__device__ __forceinline__ void updv(double &st, u32 sad, const KeyV &v, const KeyCtx &c, int px, int py, int wlim,
                                     int hlim)
{
    u32 hi = (SH0 ? sad : (sad << c.sh)) + v.hi;
    if (EDGE) {
        if (!node_valid(px, py, v.dx, c.dy, wlim, hlim)) hi |= KEY_INVALID_HI;
    }
    u32x2 key = {v.lo, hi};
    st = me_vmin64(st, __builtin_bit_cast(double, key));
}

__device__ __forceinline__ KeyV key_sel(bool b, const KeyV &one, const KeyV &zero)
{
    return {b ? one.hi : zero.hi, b ? one.lo : zero.lo, b ? one.dx : zero.dx};
}

__device__ __forceinline__ KeyV key_k(const KeyCtx &c, int k) { return {c.khi[k], c.klo[k], c.dx[k]}; }

// The values of one node for the step's four displacements (Quad), and two
// of them, displacements 2p and 2p+1 (Pair), as the sparse classes take
// them.  At 8 bits a Quad is a Val and a Pair one of its words (u16 lanes);
// at 10 bits a Quad is four u32 and a Pair two.
#if BIT_DEPTH == 8
typedef Val Quad;
typedef u32 Pair;
__device__ __forceinline__ Pair quad_pair(const Quad &q, bool p) { return p ? q.hi() : q.lo(); }
__device__ __forceinline__ u32 pair_elem(Pair v, int j) { return j ? (v >> 16) : (v & 0xFFFFu); }
template <typename F> __device__ __forceinline__ Pair pair_map(Pair v, F f) { return f(v); }
#else
// Four named members, not an array: a choice between two of them by a lane
// bit (quad_pair) would otherwise be folded into a variable array index,
// which puts the array in private memory.
struct Quad {
    u32 v0, v1, v2, v3;
    __device__ __forceinline__ void set(int k, u32 x)   // k a compile-time constant
    {
        if (k == 0) v0 = x;
        else if (k == 1) v1 = x;
        else if (k == 2) v2 = x;
        else v3 = x;
    }
};
struct Pair {
    u32 a, b;
};
__device__ __forceinline__ Pair quad_pair(const Quad &q, bool p) { return {p ? q.v2 : q.v0, p ? q.v3 : q.v1}; }
__device__ __forceinline__ u32 pair_elem(Pair v, int j) { return j ? v.b : v.a; }
template <typename F> __device__ __forceinline__ Pair pair_map(Pair v, F f) { return {f(v.a), f(v.b)}; }
#endif

// The 19 sparse classes (nodes of 32x16 and larger that are not dense pair
// slots) with the four displacements spread over the lanes that share a
// node: a sum over a lane pair hands each lane of the pair the sums of
// different displacements (a reduce-scatter), so each lane updates each
// class for one or two displacements instead of four.  Lane bits
// x0 = (rx ^ rx>>2) & 1, x1 = (rx>>1 ^ rx>>2) & 1, y0 = ry&1, y1 = ry>>1&1:
// the x bits are taken relative to x2 = rx>>2 because the 128-wide sums pair
// lane rx with lane 7-rx (DPP row_half_mirror), which then holds the same
// displacement.  The displacement of a value:
//   pair by x0 (k = 2x0, 2x0+1): 32x16, 64x8        pair by y0: 16x32, 8x64
//   k = 2x0 + y0: 32x32 and every class built from it (64x32 ... 128x128)
//   k = 2x0 + x1: 64x16, 128x16                       k = 2y0 + y1: 16x64, 16x128
// The lanes holding different displacements of a node are merged before
// publishing (merge_sparse).  A 128-tall node spans both halves of the wave,
// so after its sum over the halves each half holds it whole: the four such
// classes live in two node states, one class per half.
template <bool EDGE, bool SH0>
__device__ __forceinline__ void sparse_scatter(const Quad &s, const Quad *t32x8, const Quad *t8x32, double *st,
                                               const KeyCtx &c, int rx, int ry, int cx, int cy, int wlim, int hlim)
{
    const int X = rx * 16, Y = ry * 16, px = cx + X, py = cy + Y;
    const int x2 = cx + (X & ~31), x4 = cx + (X & ~63), y2 = cy + (Y & ~31), y4 = cy + (Y & ~63);
    const bool x0 = (rx ^ (rx >> 2)) & 1, x1 = ((rx >> 1) ^ (rx >> 2)) & 1, y0 = ry & 1, y1 = (ry >> 1) & 1;
    const KeyV kxa = key_sel(x0, key_k(c, 2), key_k(c, 0)), kxb = key_sel(x0, key_k(c, 3), key_k(c, 1));
    const KeyV kya = key_sel(y0, key_k(c, 2), key_k(c, 0)), kyb = key_sel(y0, key_k(c, 3), key_k(c, 1));
    const KeyV kxy = key_sel(y0, kxb, kxa), kxx = key_sel(x1, kxb, kxa), kyy = key_sel(y1, kyb, kya);
    const auto fx1 = [](u32 v) { return px1(v); };
    const auto fy1 = [](u32 v) { return py1(v); };
    const auto fx2 = [](u32 v) { return px2(v); };
    const auto fy2 = [](u32 v) { return py2(v); };
    // x pairs: lane x0 keeps displacements 2x0, 2x0+1 and receives them from
    // its partner, which selects the pair by its own (opposite) x0
    const Pair sx = pair_map(quad_pair(s, !x0), fx1), kx = quad_pair(s, x0);
    const u32 t32x16a = pair_elem(kx, 0) + pair_elem(sx, 0), t32x16b = pair_elem(kx, 1) + pair_elem(sx, 1);
    const Pair sy = pair_map(quad_pair(s, !y0), fy1), ky = quad_pair(s, y0);
    const u32 t16x32a = pair_elem(ky, 0) + pair_elem(sy, 0), t16x32b = pair_elem(ky, 1) + pair_elem(sy, 1);
    updv<EDGE, SH0>(st[6], t32x16a, kxa, c, x2, py, wlim, hlim);
    updv<EDGE, SH0>(st[6], t32x16b, kxb, c, x2, py, wlim, hlim);
    updv<EDGE, SH0>(st[7], t16x32a, kya, c, px, y2, wlim, hlim);
    updv<EDGE, SH0>(st[7], t16x32b, kyb, c, px, y2, wlim, hlim);
    // 32x32 and up: lane y0 keeps displacement 2x0 + y0
    const u32 t32x32 = (y0 ? t32x16b : t32x16a) + py1(y0 ? t32x16a : t32x16b);
    const u32 t64x32 = t32x32 + px2(t32x32), t32x64 = t32x32 + py2(t32x32);
    const u32 t64x64 = t64x32 + py2(t64x32);
    const u32 t128x32 = t64x32 + px4(t64x32), t128x64 = t64x64 + px4(t64x64);
    // The 128-tall classes: one sum over the wave's halves serves two of
    // them, one per half, which then share a node state (st[21]: 32x128 in
    // the lower half, 64x128 in the upper; st[19]: 128x128 / 16x128).
    const bool upper = ry >= 4;
    const u32 t32x128_64x128 = hsum2(t32x64, t64x64);
    updv<EDGE, SH0>(st[8], t32x32, kxy, c, x2, y2, wlim, hlim);
    updv<EDGE, SH0>(st[15], t64x32, kxy, c, x4, y2, wlim, hlim);
    updv<EDGE, SH0>(st[16], t32x64, kxy, c, x2, y4, wlim, hlim);
    updv<EDGE, SH0>(st[17], t64x64, kxy, c, x4, y4, wlim, hlim);
    updv<EDGE, SH0>(st[20], t128x32, kxy, c, cx, y2, wlim, hlim);
    updv<EDGE, SH0>(st[21], t32x128_64x128, kxy, c, upper ? x4 : x2, cy, wlim, hlim);
    updv<EDGE, SH0>(st[22], t128x64, kxy, c, cx, y4, wlim, hlim);
    // 64x16, 128x16: lane x1 keeps displacement 2x0 + x1
    const u32 t64x16 = (x1 ? t32x16b : t32x16a) + px2(x1 ? t32x16a : t32x16b);
    const u32 t128x16 = t64x16 + px4(t64x16);
    updv<EDGE, SH0>(st[11], t64x16, kxx, c, x4, py, wlim, hlim);
    updv<EDGE, SH0>(st[18], t128x16, kxx, c, cx, py, wlim, hlim);
    // 16x64, 16x128: lane y1 keeps displacement 2y0 + y1
    const u32 t16x64 = (y1 ? t16x32b : t16x32a) + py2(y1 ? t16x32a : t16x32b);
    updv<EDGE, SH0>(st[14], t16x64, kyy, c, px, y4, wlim, hlim);
    // 128x128 (lower half, displacement 2x0 + y0) and 16x128 (upper half,
    // 2y0 + y1)
    const u32 t128x128_16x128 = hsum2(t128x64, t16x64);
    updv<EDGE, SH0>(st[19], t128x128_16x128, key_sel(upper, kyy, kxy), c, upper ? px : cx, cy, wlim, hlim);
    // 64x8 (from 32x8, equal on both lanes of an x pair) and 8x64
#pragma unroll
    for (int r2 = 0; r2 < 2; r2++) {
        const Pair v = quad_pair(t32x8[r2], x0), w = pair_map(v, fx2);
        updv<EDGE, SH0>(st[9 + r2], pair_elem(v, 0) + pair_elem(w, 0), kxa, c, x4, py + 8 * r2, wlim, hlim);
        updv<EDGE, SH0>(st[9 + r2], pair_elem(v, 1) + pair_elem(w, 1), kxb, c, x4, py + 8 * r2, wlim, hlim);
    }
#pragma unroll
    for (int c2 = 0; c2 < 2; c2++) {
        const Pair v = quad_pair(t8x32[c2], y0), w = pair_map(v, fy2);
        updv<EDGE, SH0>(st[12 + c2], pair_elem(v, 0) + pair_elem(w, 0), kya, c, px + 8 * c2, y4, wlim, hlim);
        updv<EDGE, SH0>(st[12 + c2], pair_elem(v, 1) + pair_elem(w, 1), kyb, c, px + 8 * c2, y4, wlim, hlim);
    }
}

__device__ __forceinline__ double kmin(double a, double b) { return me_vmin64(a, b); }
__device__ __forceinline__ double dpp64_x1(double v)
{
    const u64 u = __builtin_bit_cast(u64, v);
    return __builtin_bit_cast(double, ((u64)px1((u32)(u >> 32)) << 32) | px1((u32)u));
}
__device__ __forceinline__ double dpp64_x2(double v)
{
    const u64 u = __builtin_bit_cast(u64, v);
    return __builtin_bit_cast(double, ((u64)px2((u32)(u >> 32)) << 32) | px2((u32)u));
}
__device__ __forceinline__ double dpp64_y1(double v)
{
    const u64 u = __builtin_bit_cast(u64, v);
    return __builtin_bit_cast(double, ((u64)py1((u32)(u >> 32)) << 32) | py1((u32)u));
}
__device__ __forceinline__ double dpp64_y2(double v)
{
    const u64 u = __builtin_bit_cast(u64, v);
    return __builtin_bit_cast(double, ((u64)py2((u32)(u >> 32)) << 32) | py2((u32)u));
}

// Merges the lanes that hold different displacements of one sparse node, so
// that the node's publishing lane holds the minimum over all four.
__device__ __forceinline__ void merge_sparse(double *st)
{
    st[6] = kmin(st[6], dpp64_x1(st[6]));
    st[9] = kmin(st[9], dpp64_x1(st[9]));
    st[10] = kmin(st[10], dpp64_x1(st[10]));
    st[7] = kmin(st[7], dpp64_y1(st[7]));
    st[12] = kmin(st[12], dpp64_y1(st[12]));
    st[13] = kmin(st[13], dpp64_y1(st[13]));
    const int xy[] = {8, 15, 16, 17, 20, 21, 22};   // 21: 32x128 and 64x128
#pragma unroll
    for (int i = 0; i < 7; i++) {
        st[xy[i]] = kmin(st[xy[i]], dpp64_x1(st[xy[i]]));
        st[xy[i]] = kmin(st[xy[i]], dpp64_y1(st[xy[i]]));
    }
    st[11] = kmin(st[11], dpp64_x1(st[11]));
    st[11] = kmin(st[11], dpp64_x2(st[11]));
    st[18] = kmin(st[18], dpp64_x1(st[18]));
    st[18] = kmin(st[18], dpp64_x2(st[18]));
    st[14] = kmin(st[14], dpp64_y1(st[14]));
    st[14] = kmin(st[14], dpp64_y2(st[14]));
    // st[19]: 128x128 in the lower half (partners x1, y1), 16x128 in the
    // upper half (y1, y2)
    st[19] = kmin(st[19], dpp64_y1(st[19]));
    const double o19x = dpp64_x1(st[19]), o19y = dpp64_y2(st[19]);
    st[19] = kmin(st[19], (threadIdx.x & 63) >= 32 ? o19y : o19x);
}

// Pass L: the 609 PUs larger than a 16x16 region.
// Pairs of regions (x pairs, y pairs) hold 14 nodes between two lanes; each
// lane keeps half of them (3 "dense" slots per direction, chosen by the lane's
// rank in the pair).  The 19 sparser classes are reduce-scattered over the
// lanes of their group (sparse_scatter); one lane per group publishes.
//   st[0..2] x-pair slots: 32x4 rows {0,2}, {1,3} by rank; 32x8 row = rank
//   st[3..5] y-pair slots: 4x32 cols {0,2}, {1,3}; 8x32 col = rank
//   st[6] 32x16  st[7] 16x32  st[8] 32x32  st[9..10] 64x8 (r2)  st[11] 64x16
//   st[12..13] 8x64 (c2)  st[14] 16x64  st[15] 64x32  st[16] 32x64  st[17] 64x64
//   st[18] 128x16  st[19] 128x128 (lanes 0..31) / 16x128 (lanes 32..63)
//   st[20] 128x32  st[21] 32x128 (lanes 0..31) / 64x128 (lanes 32..63)
//   st[22] 128x64
#define N_LARGE 23
// The dense slots for the region's displacements kb .. kb+NK-1; returns the
// 32x8 and 8x32 sums that the sparse classes build on.
template <bool EDGE, bool SH0>
__device__ __forceinline__ void pass_dense(const Region &g, double *st, const KeyCtx &c, int kb, int rx, int ry, int cx,
                                           int cy, int wlim, int hlim, Val *t32x8, Val *t8x32)
{
    const int X = rx * 16, Y = ry * 16, px = cx + X, py = cy + Y;
    const bool xr = rx & 1, yr = ry & 1;
    const int x2 = cx + (X & ~31), y2 = cy + (Y & ~31);
    Val t32x4[4], t4x32[4];
#pragma unroll
    for (int r = 0; r < 4; r++) t32x4[r] = vaddx1(g.h16[r]);
#pragma unroll
    for (int r2 = 0; r2 < 2; r2++) t32x8[r2] = vaddx1(g.h16x8[r2]);
#pragma unroll
    for (int cc = 0; cc < 4; cc++) t4x32[cc] = vaddy1(g.v16[cc]);
#pragma unroll
    for (int c2 = 0; c2 < 2; c2++) t8x32[c2] = vaddy1(g.v8x16[c2]);
    const Val d0 = vsel(xr, t32x4[2], t32x4[0]), d1 = vsel(xr, t32x4[3], t32x4[1]), d2 = vsel(xr, t32x8[1], t32x8[0]);
    const Val d3 = vsel(yr, t4x32[2], t4x32[0]), d4 = vsel(yr, t4x32[3], t4x32[1]), d5 = vsel(yr, t8x32[1], t8x32[0]);
#pragma unroll
    for (int k = 0; k < NK; k++) {
        upd1<EDGE, SH0>(st[0], d0.get(k), c, kb + k, x2, py + (xr ? 8 : 0), wlim, hlim);
        upd1<EDGE, SH0>(st[1], d1.get(k), c, kb + k, x2, py + (xr ? 12 : 4), wlim, hlim);
        upd1<EDGE, SH0>(st[2], d2.get(k), c, kb + k, x2, py + (xr ? 8 : 0), wlim, hlim);
        upd1<EDGE, SH0>(st[3], d3.get(k), c, kb + k, px + (yr ? 8 : 0), y2, wlim, hlim);
        upd1<EDGE, SH0>(st[4], d4.get(k), c, kb + k, px + (yr ? 12 : 4), y2, wlim, hlim);
        upd1<EDGE, SH0>(st[5], d5.get(k), c, kb + k, px + (yr ? 8 : 0), y2, wlim, hlim);
    }
}

__device__ __forceinline__ void atomic_key_min(u64 *p, double st)
{
    u64 v = __builtin_bit_cast(u64, st);
    if (v != KEY_INIT) atomicMin((unsigned long long *)p, (unsigned long long)v);
}

// PU index inside the CTU of shape s at CTU-local (x, y).
__device__ __forceinline__ int pu_index(const IntArgs &a, int s, int w, int h, int x, int y)
{
    return a.base[s] + (y / h) * (128 / w) + x / w;
}

// The LDS window is stored with every block of 16 rows shifted right by
// ((row >> 4) & 3) dwords.  Lanes read rows 16 apart at columns 16 samples
// apart, so without the shift a half-wave's 32 dword reads fall into only 8
// of the 32 banks; with it they fall into 32 distinct banks (8-bit).
__device__ __forceinline__ int wskew(int row) { return ((row >> 4) & 3) * (4 / (int)sizeof(pel)); }

// Leaves: the 16 4x4-block SADs of the lane's 16x16 region, NK displacements
// at a time, as L[r][c].
#if BIT_DEPTH == 8
// v_qsad_pk_u16_u8 gives the SADs of one 4-sample row at the four
// displacements dxb..dxb+3 in one instruction (4 x u16).
__device__ __forceinline__ void leaves(const pel *win, int wcols, int wrow0, int c0, const pel *curl, int crow0, int ccol0,
                                       Val L[4][4])
{
    u64 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        // each overlapping dword pair is read on its own: a 64-bit operand
        // needs an even register pair, so pairs (d1,d2) and (d3,d4) of one
        // five-dword read would each cost a register copy
        const u32 *rp = (const u32 *)(win + (wrow0 + r) * wcols + wskew(wrow0 + r) + c0);
        const uint4 cw = *(const uint4 *)(curl + (crow0 + r) * 128 + ccol0);
        const u32 cc[4] = {cw.x, cw.y, cw.z, cw.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const u32x2a p = *(const u32x2a *)(rp + j);
            acc[r >> 2][j] = __builtin_amdgcn_qsad_pk_u16_u8(((u64)p.y << 32) | p.x, cc[j], acc[r >> 2][j]);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) L[r][c] = {acc[r][c]};
}
#else
// v_sad_u16 on sample pairs.  The step's four displacements dxb + k share
// each window row read (10 dwords, samples 0..19 from the step's column):
// k = 0 and 2 take dword-aligned pairs, k = 1 and 3 pairs realigned with
// v_alignbyte (nine realigned dwords serve both).  L[k] receives the 16
// leaves of displacement k.
__device__ __forceinline__ void leaves4(const pel *win, int wcols, int wrow0, int c0, const pel *curl, int crow0,
                                        int ccol0, Val L[NKEY][4][4])
{
#pragma unroll
    for (int k = 0; k < NKEY; k++)
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) L[k][r][c].v = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const u32 *rp = (const u32 *)(win + (wrow0 + r) * wcols + wskew(wrow0 + r) + c0);
        u32 d[10], o[9];
#pragma unroll
        for (int j = 0; j < 10; j++) d[j] = rp[j];
#pragma unroll
        for (int m = 0; m < 9; m++) o[m] = __builtin_amdgcn_alignbyte(d[m + 1], d[m], 2);   // samples 2m+1, 2m+2
        const uint4 c01 = *(const uint4 *)(curl + (crow0 + r) * 128 + ccol0);
        const uint4 c23 = *(const uint4 *)(curl + (crow0 + r) * 128 + ccol0 + 8);
        const u32 cc[8] = {c01.x, c01.y, c01.z, c01.w, c23.x, c23.y, c23.z, c23.w};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            // leaf column c: current samples 4c .. 4c+3, window samples 4c+k ..
            u32 *l = &L[0][r >> 2][c].v;
            l[0] = __builtin_amdgcn_sad_u16(d[2 * c], cc[2 * c], __builtin_amdgcn_sad_u16(d[2 * c + 1], cc[2 * c + 1], l[0]));
            u32 &l1 = L[1][r >> 2][c].v, &l2 = L[2][r >> 2][c].v, &l3 = L[3][r >> 2][c].v;
            l1 = __builtin_amdgcn_sad_u16(o[2 * c], cc[2 * c], __builtin_amdgcn_sad_u16(o[2 * c + 1], cc[2 * c + 1], l1));
            l2 = __builtin_amdgcn_sad_u16(d[2 * c + 1], cc[2 * c], __builtin_amdgcn_sad_u16(d[2 * c + 2], cc[2 * c + 1], l2));
            l3 = __builtin_amdgcn_sad_u16(o[2 * c + 1], cc[2 * c], __builtin_amdgcn_sad_u16(o[2 * c + 2], cc[2 * c + 1], l3));
        }
        // keep the compiler from hoisting the LDS reads of all 16 rows ahead
        // of their use, which alone exceeds the registers: a block of 4 rows
        // is summed before the next block is read
        if ((r & 3) == 3) {
#pragma unroll
            for (int k = 0; k < NKEY; k++) {
                Val *q = L[k][r >> 2];
                asm volatile("" : "+v"(q[0].v), "+v"(q[1].v), "+v"(q[2].v), "+v"(q[3].v)::"memory");
            }
        }
    }
}
#endif

struct Block {
    int cx, cy, dy0, dxs, dxe, ref;
};

// The rate row is read-only for the whole launch: reading it through the
// constant address space keeps its loads scalar even where the workgroup's
// key stores of an earlier segment precede them in the loop.  The row holds
// ME_RATE_PAD entries past its end, so that four consecutive entries can be
// read from any step of four displacements, and the step after the last.
typedef const __attribute__((address_space(4))) u32 *const_u32p;
typedef u32 u32x4a __attribute__((ext_vector_type(4), aligned(4)));
typedef const __attribute__((address_space(4))) u32x4a *const_u32x4p;
__device__ __forceinline__ u32 rate_at(const IntArgs &a, int i) { return ((const_u32p)a.rate)[i]; }
__device__ __forceinline__ u32x4a rate4_at(const IntArgs &a, int i)
{
    return *(const_u32x4p)((const_u32p)a.rate + i);
}

// One pass over the band's displacements: PASS 3 = every PU class, 2 =
// 16x16 only.  Updates the node states in st.
//
// The band's (row, step) pairs -- rows inside [lo, hi], steps of four
// displacements from the first that reaches lo -- are split in row-major
// order into WAVES contiguous ranges of equal length, one per wave.  A full
// band of band_rows = k * WAVES rows gives each wave k whole rows; a short
// band (a segment's last, 2R+1 being odd) gives each wave a part of a row
// instead of leaving all but one wave idle.  Keys carry their raster index,
// so the order in which a node's candidates are met does not matter.
template <int PASS, int WAVES, bool EDGE, bool SH0, int NS>
__device__ __forceinline__ void run_pass(const IntArgs &a, const PicDesc &P, const Block &b, const pel *win, const pel *curl,
                                         double *st, int tid)
{
    // the wave index is uniform across the wave: keep it (and everything the
    // displacement loop derives from it) in scalar registers
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int rx = lane & 7, ry = lane >> 3, X = rx * 16, Y = ry * 16;
    const int wlim = P.pic_w + 126, hlim = P.pic_h + 126;
    const int span = 2 * a.R + 1;
    const int r0 = a.lo - b.dy0 > 0 ? a.lo - b.dy0 : 0;
    const int r1 = a.hi - b.dy0 + 1 < a.band_rows ? a.hi - b.dy0 + 1 : a.band_rows;
    const int s0 = a.lo - 3 - b.dxs > 0 ? (a.lo - 3 - b.dxs + 3) >> 2 : 0;   // first step with dxb + 3 >= lo
    const int S = ((b.dxe - b.dxs) >> 2) + 1 - s0;                          // steps per row
    const int total = r1 > r0 && S > 0 ? (r1 - r0) * S : 0;
    int p = total * wave / WAVES;
    const int pe = total * (wave + 1) / WAVES;
    while (p < pe) {
        const int rr = p / S, sb = p - rr * S;
        const int n = S - sb < pe - p ? S - sb : pe - p;
        p += n;
        const int row = r0 + rr;
        const int dy = b.dy0 + row;
        const int dxs = b.dxs + 4 * (s0 + sb), dxl = dxs + 4 * (n - 1);   // the chunk's first and last step
        const int dyi = dy + a.R;
        const u32 ry_rate = rate_at(a, dyi);
        // The key parts of the step's four displacements.  A displacement
        // outside the range gets a high word above every valid key (never
        // NaN).  They are made at the end of the previous step, with one
        // scalar load of the four rates: waiting for a scalar load also waits
        // for every LDS read in flight, and at a step's end none is.
        auto make_keys = [&](int dxb, const u32x4a &rx4) {
            KeyCtx kc;
            kc.dy = dy;
            kc.sh = a.sh;
#pragma unroll
            for (int k = 0; k < NKEY; k++) {
                const int dx = dxb + k, dxi = dx + a.R;
                kc.dx[k] = dx;
                const u64 K = ((u64)(rx4[k] + ry_rate) << a.IB) | (u64)(dyi * span + dxi);
                const bool out = dx < a.lo || dx > b.dxe;
                kc.khi[k] = __builtin_amdgcn_readfirstlane(out ? KEY_OUT_HI : (u32)(K >> 32) + KEY_BIAS_HI);
                kc.klo[k] = __builtin_amdgcn_readfirstlane(out ? 0u : (u32)K);
            }
            return kc;
        };
        KeyCtx kc = make_keys(dxs, rate4_at(a, dxs + a.R));
        for (int dxb = dxs; dxb <= dxl; dxb += 4) {
            // always true from s0 on; the branch keeps the step's search in a
            // block of its own, which the scheduler orders better
            if (dxb + 3 >= a.lo) {
#if BIT_DEPTH == 8
                Val L[4][4];
                leaves(win, a.wpitch, Y + row, X + (dxb - b.dxs), curl, Y, X, L);
                Region g;
                g.build(L);
                if (PASS == 3) {
                    pass_small<EDGE, SH0>(g, st, kc, 0, b.cx + X, b.cy + Y, wlim, hlim);
                    Val t32x8[2], t8x32[2];
                    pass_dense<EDGE, SH0>(g, st + N_SMALL, kc, 0, rx, ry, b.cx, b.cy, wlim, hlim, t32x8, t8x32);
                    sparse_scatter<EDGE, SH0>(g.s16, t32x8, t8x32, st + N_SMALL, kc, rx, ry, b.cx, b.cy, wlim, hlim);
                } else {
#pragma unroll
                    for (int k = 0; k < NK; k++) upd1<EDGE, SH0>(st[0], g.s16.get(k), kc, k, b.cx + X, b.cy + Y, wlim, hlim);
                }
#else
                // the four displacements' leaves from one pass over the window
                // rows; then each displacement's region is composed and compared
                // on its own, and the sparse classes take all four together
                Val L[NKEY][4][4];
                leaves4(win, a.wpitch, Y + row, X + (dxb - b.dxs), curl, Y, X, L);
                Quad s16, t32x8[2], t8x32[2];
#pragma unroll
                for (int k = 0; k < NKEY; k++) {
                    Region g;
                    g.build(L[k]);
                    if (PASS == 3) {
                        pass_small<EDGE, SH0>(g, st, kc, k, b.cx + X, b.cy + Y, wlim, hlim);
                        Val a8[2], b8[2];
                        pass_dense<EDGE, SH0>(g, st + N_SMALL, kc, k, rx, ry, b.cx, b.cy, wlim, hlim, a8, b8);
                        s16.set(k, g.s16.v);
                        t32x8[0].set(k, a8[0].v);
                        t32x8[1].set(k, a8[1].v);
                        t8x32[0].set(k, b8[0].v);
                        t8x32[1].set(k, b8[1].v);
                    } else {
                        upd1<EDGE, SH0>(st[0], g.s16.v, kc, k, b.cx + X, b.cy + Y, wlim, hlim);
                    }
                }
                if (PASS == 3) sparse_scatter<EDGE, SH0>(s16, t32x8, t8x32, st + N_SMALL, kc, rx, ry, b.cx, b.cy, wlim, hlim);
#endif
            }
            // the next step's rates are loaded only here (see make_keys)
            __builtin_amdgcn_sched_barrier(0);
            kc = make_keys(dxb + 4, rate4_at(a, dxb + 4 + a.R));
        }
    }
}

// A node's key into the slot of this piece (kb), as the minimum with the key
// there when `merge` is set, or merged into the one key array with atomicMin
// (atomic layouts, kb there).
#define PUB(S, W, H, XX, YY, ST)                                                    \
    do {                                                                            \
        if (a.base[S] >= 0) {                                                       \
            u64 *kp_ = kb + (size_t)pu_index(a, S, W, H, XX, YY) * a.nref;          \
            const u64 k_ = __builtin_bit_cast(u64, ST);                             \
            if (atomic) atomic_key_min(kp_, ST);                                    \
            else if (merge) { const u64 o_ = *kp_; *kp_ = k_ < o_ ? k_ : o_; }      \
            else *kp_ = k_;                                                         \
        }                                                                           \
    } while (0)

__device__ __forceinline__ void publish_small(const IntArgs &a, u64 *kb, bool atomic, bool merge, const double *st, int X,
                                              int Y)
{
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int r2 = 0; r2 < 2; r2++) PUB(S4x8, 4, 8, X + 4 * c, Y + 8 * r2, st[c * 2 + r2]);
#pragma unroll
    for (int c2 = 0; c2 < 2; c2++)
#pragma unroll
        for (int r = 0; r < 4; r++) PUB(S8x4, 8, 4, X + 8 * c2, Y + 4 * r, st[8 + c2 * 4 + r]);
#pragma unroll
    for (int c = 0; c < 4; c++) PUB(S4x16, 4, 16, X + 4 * c, Y, st[16 + c]);
#pragma unroll
    for (int r = 0; r < 4; r++) PUB(S16x4, 16, 4, X, Y + 4 * r, st[20 + r]);
#pragma unroll
    for (int c2 = 0; c2 < 2; c2++)
#pragma unroll
        for (int r2 = 0; r2 < 2; r2++) PUB(S8x8, 8, 8, X + 8 * c2, Y + 8 * r2, st[24 + c2 * 2 + r2]);
#pragma unroll
    for (int c2 = 0; c2 < 2; c2++) PUB(S8x16, 8, 16, X + 8 * c2, Y, st[28 + c2]);
#pragma unroll
    for (int r2 = 0; r2 < 2; r2++) PUB(S16x8, 16, 8, X, Y + 8 * r2, st[30 + r2]);
    PUB(S16x16, 16, 16, X, Y, st[32]);
}

__device__ __forceinline__ void publish_large(const IntArgs &a, u64 *kb, bool atomic, bool merge, const double *st, int rx,
                                              int ry)
{
    const int X = rx * 16, Y = ry * 16, X2 = X & ~31, Y2 = Y & ~31, X4 = X & ~63, Y4 = Y & ~63;
    const bool xr = rx & 1, yr = ry & 1;
    // dense slots: every lane owns distinct nodes
    PUB(S32x4, 32, 4, X2, Y + (xr ? 8 : 0), st[0]);
    PUB(S32x4, 32, 4, X2, Y + (xr ? 12 : 4), st[1]);
    PUB(S32x8, 32, 8, X2, Y + (xr ? 8 : 0), st[2]);
    PUB(S4x32, 4, 32, X + (yr ? 8 : 0), Y2, st[3]);
    PUB(S4x32, 4, 32, X + (yr ? 12 : 4), Y2, st[4]);
    PUB(S8x32, 8, 32, X + (yr ? 8 : 0), Y2, st[5]);
    // sparse classes: the group's top-left lane publishes
    const bool ex1 = !(rx & 1), ex2 = !(rx & 3), ex3 = rx == 0;
    const bool ey1 = !(ry & 1), ey2 = !(ry & 3), ey3 = ry == 0;
    if (ex1) PUB(S32x16, 32, 16, X, Y, st[6]);
    if (ey1) PUB(S16x32, 16, 32, X, Y, st[7]);
    if (ex1 && ey1) PUB(S32x32, 32, 32, X, Y, st[8]);
    if (ex2) {
        PUB(S64x8, 64, 8, X, Y, st[9]);
        PUB(S64x8, 64, 8, X, Y + 8, st[10]);
        PUB(S64x16, 64, 16, X, Y, st[11]);
    }
    if (ey2) {
        PUB(S8x64, 8, 64, X, Y, st[12]);
        PUB(S8x64, 8, 64, X + 8, Y, st[13]);
        PUB(S16x64, 16, 64, X, Y, st[14]);
    }
    if (ex2 && ey1) PUB(S64x32, 64, 32, X, Y, st[15]);
    if (ex1 && ey2) PUB(S32x64, 32, 64, X, Y, st[16]);
    if (ex2 && ey2) PUB(S64x64, 64, 64, X, Y, st[17]);
    if (ex3) PUB(S128x16, 128, 16, X, Y, st[18]);
    // the 128-tall classes kept in the upper half publish from its first row
    // (ry == 4; pu_index rounds Y down to the node)
    const bool eyh = ry == 4;
    if (eyh) PUB(S16x128, 16, 128, X, Y, st[19]);
    if (ex3 && ey1) PUB(S128x32, 128, 32, X, Y, st[20]);
    if (ex1 && ey3) PUB(S32x128, 32, 128, X, Y, st[21]);
    if (ex3 && ey2) PUB(S128x64, 128, 64, X, Y, st[22]);
    if (ex2 && eyh) PUB(S64x128, 64, 128, X, Y, st[21]);
    if (ex3 && ey3) PUB(S128x128, 128, 128, X, Y, st[19]);
}

// The waves of a block hold node states for different displacement rows of
// the same nodes: merge them in LDS (the window is no longer needed), so that
// one wave publishes per block.  Returns whether this wave publishes.
template <int WAVES, int NS>
__device__ __forceinline__ bool merge_waves(double *st, unsigned char *smem)
{
    if (WAVES == 1) return true;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    u64 *red = (u64 *)smem;   // [NS][64]
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int i = 0; i < NS; i++) red[i * 64 + lane] = __builtin_bit_cast(u64, st[i]);
    }
    __syncthreads();
    if (wave != 0) {
#pragma unroll
        for (int i = 0; i < NS; i++)
            __hip_atomic_fetch_min(red + i * 64 + lane, __builtin_bit_cast(u64, st[i]), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    if (wave != 0) return false;
#pragma unroll
    for (int i = 0; i < NS; i++) st[i] = __builtin_bit_cast(double, red[i * 64 + lane]);
    return true;
}

// Stages the reference window of band b in LDS: rows cy + dy0 ..
// cy + dy0 + 128 + band_rows - 2, columns cx + dxs .. cx + dxs + wcols - 1,
// a dword (4 / sizeof(pel) samples) per thread and item.  Samples are
// clamped to the padded plane (it is edge-replicated, so clamping is exact);
// away from the plane's edges a dword is realigned from two aligned loads.
// A thread issues the loads of SB items before it stores any of them: the
// planes are read through global (not generic) pointers, which cannot alias
// the LDS, so the loads stay in flight together instead of one memory round
// trip per item.
typedef const __attribute__((address_space(1))) pel *gpel_p;
typedef const __attribute__((address_space(1))) u32 *gu32_p;
#define SB 8
__device__ __forceinline__ void stage_window(const IntArgs &a, const PicDesc &P, const Block &b, pel *win, int nthr,
                                             int tid)
{
    constexpr int SPD = 4 / (int)sizeof(pel);
    const int wrows = 128 + a.band_rows - 1;
    const gpel_p ref0 = (gpel_p)P.planes[b.ref] + (size_t)ME_PAD * P.ref_stride + ME_PAD;
    const int xmin = -ME_PAD, xmax = P.pic_w + ME_PAD - 1, ymin = -ME_PAD, ymax = P.pic_h + ME_PAD - 1;
    const int wq = a.wcols / SPD, dq = nthr % wq, dr = nthr / wq;
    int r = tid / wq, q = tid - r * wq;
    while (r < wrows) {
        // the loads of the batch from straight-line code: every item reads
        // an address inside the padded plane (items past the window repeat
        // its last row; items off the plane's sides read a clamped position
        // and are redone sample by sample below)
        u64 w[SB];
        u32 sh[SB];
        int o[SB];
        bool side = false;
#pragma unroll
        for (int j = 0; j < SB; j++) {
            const int x0 = b.cx + b.dxs + q * SPD;
            const int rc = r < wrows ? r : wrows - 1;
            const int y = me_clip3(ymin, ymax, b.cy + b.dy0 + rc);
            const int xc = me_clip3(xmin, xmax - SPD + 1, x0);
            const uintptr_t ad = (uintptr_t)(ref0 + (ptrdiff_t)y * P.ref_stride + xc);
            const gu32_p qa = (gu32_p)(ad & ~(uintptr_t)3);
            w[j] = ((u64)qa[1] << 32) | qa[0];
            sh[j] = (u32)(ad & 3);
            o[j] = r < wrows ? r * a.wpitch + wskew(r) + q * SPD : -1;
            side |= r < wrows && xc != x0;
            r += dr;
            q += dq;
            if (q >= wq) { q -= wq; r++; }
        }
#pragma unroll
        for (int j = 0; j < SB; j++)
            if (o[j] >= 0) *(u32 *)(win + o[j]) = __builtin_amdgcn_alignbyte((u32)(w[j] >> 32), (u32)w[j], sh[j]);
        if (side) {
            // the batch's items off the plane's sides (rare: windows reaching
            // more than the padding past the picture), sample by sample
#pragma unroll
            for (int j = 0; j < SB; j++) {
                if (o[j] < 0) continue;
                const int rj = o[j] / a.wpitch, qj = (o[j] - rj * a.wpitch - wskew(rj)) / SPD;
                const int x0 = b.cx + b.dxs + qj * SPD;
                if (x0 >= xmin && x0 + SPD - 1 <= xmax) continue;
                const gpel_p src = ref0 + (ptrdiff_t)me_clip3(ymin, ymax, b.cy + b.dy0 + rj) * P.ref_stride;
                u32 v = 0;
#pragma unroll
                for (int k = 0; k < SPD; k++) v |= (u32)src[me_clip3(xmin, xmax, x0 + k)] << (8 * (int)sizeof(pel) * k);
                *(u32 *)(win + o[j]) = v;
            }
        }
    }
}

// Stages the current CTU (clamped to the picture) in LDS, 128 x 128 samples,
// SB loads in flight per thread as above.
__device__ __forceinline__ void stage_cur(const PicDesc &P, const Block &b, pel *curl, int nthr, int tid)
{
    constexpr int SPD = 4 / (int)sizeof(pel);
    constexpr int CQ = 128 / SPD;   // dwords per CTU row
    const int q = tid % CQ, rstep = nthr / CQ;
    const bool inner = b.cx + 127 <= P.pic_w - 1;
    const gpel_p cur = (gpel_p)P.cur;
    for (int r0 = tid / CQ; r0 < 128; r0 += SB * rstep) {
        u32 v[SB];
#pragma unroll
        for (int j = 0; j < SB; j++) {
            const int r = r0 + j * rstep;
            v[j] = 0;
            if (r < 128) {
                const int y = me_clip3(0, P.pic_h - 1, b.cy + r), x0 = b.cx + q * SPD;
                const gpel_p src = cur + (size_t)y * P.pic_w;
                if (inner) {
                    v[j] = *(gu32_p)(src + x0);
                } else {
#pragma unroll
                    for (int k = 0; k < SPD; k++)
                        v[j] |= (u32)src[me_clip3(0, P.pic_w - 1, x0 + k)] << (8 * (int)sizeof(pel) * k);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < SB; j++) {
            const int r = r0 + j * rstep;
            if (r < 128) *(u32 *)(curl + r * 128 + q * SPD) = v[j];
        }
    }
}

// The bands [band0, band1) of one segment (job, reference, column tile):
// per band the window is staged (with the current CTU at the first band) and
// every wave covers the whole CTU (lane = 16x16 region) for every WAVES-th
// displacement row, updating the node states st of all classes.
template <int PASS, int WAVES, bool EDGE, bool SH0, int NS>
__device__ __forceinline__ void segment_bands(const IntArgs &a, const PicDesc &P, const Block seg, int band0, int band1,
                                              unsigned char *smem, double *st)
{
    const int wrows = 128 + a.band_rows - 1, nthr = 64 * WAVES;
    pel *win = (pel *)smem;
    pel *curl = win + (size_t)wrows * a.wpitch;
    for (int band = band0; band < band1; band++) {
        Block b = seg;
        b.dy0 = -a.R + band * a.band_rows;
        __syncthreads();   // every wave is done with the previous window (or merge area)
        // The staging addresses and the lane's window offsets repeat from
        // band to band; an opaque thread index keeps them from being hoisted
        // out of the band and segment loops into registers held through the
        // whole kernel: each band computes them anew.
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        stage_window(a, P, b, win, nthr, tid);
        if (band == band0) stage_cur(P, b, curl, nthr, tid);
        __syncthreads();
        run_pass<PASS, WAVES, EDGE, SH0, NS>(a, P, b, win, curl, st, tid);
    }
}

// The bands [band0, band1) of one segment in this workgroup, then one
// publication of the node states.  merge: this workgroup has already stored
// keys of the segment into its slot (non-atomic layouts), so each key is
// stored as the minimum with the one there (the same lane wrote it).
template <int MODE, int WAVES, int NS, bool EDGE, bool SH0>
__device__ __forceinline__ void search_segment(const IntArgs &a, const PicDesc &P, const Block b, int band0, int band1,
                                               unsigned char *smem, u64 *kb, bool atomic, bool merge)
{
    const int lane = threadIdx.x & 63, rx = lane & 7, ry = lane >> 3;
    constexpr int PASS = MODE == 1 ? 2 : 3;
    double st[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) st[i] = __builtin_bit_cast(double, KEY_INIT);
    segment_bands<PASS, WAVES, EDGE, SH0, NS>(a, P, b, band0, band1, smem, st);
    if (merge_waves<WAVES, NS>(st, smem)) {
        // The publishing addresses do not change from segment to segment; kept
        // opaque here, they are computed when used rather than hoisted out of
        // the workgroup's segment loop into registers held for the whole kernel.
        int px = rx, py = ry;
        asm volatile("" : "+v"(px), "+v"(py));
        if (MODE == 1) {
            PUB(S16x16, 16, 16, px * 16, py * 16, st[0]);
        } else {
            merge_sparse(st + N_SMALL);
            publish_small(a, kb, atomic, merge, st, px * 16, py * 16);
            publish_large(a, kb, atomic, merge, st + N_SMALL, px, py);
        }
    }
}

// Values every lane of a wave holds, marked uniform again inside a function
// that is not inlined (its arguments arrive in vector registers).
__device__ __forceinline__ u32 uni(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ u64 uni64(u64 v) { return ((u64)uni((u32)(v >> 32)) << 32) | uni((u32)v); }

typedef const __attribute__((address_space(4))) IntArgs *IntArgsC;

// One segment's bands [band0, band1) in this workgroup, with one
// publication.  Kept out of line: the search needs the whole register
// budget, and the workgroup's segment loop around it would otherwise share
// that budget (its invariants hoisted and held through the search).  The
// launch arguments are read from the kernel's argument segment through ap.
// A node's window (me.c:508-512) is -127-px <= dx <= W-1-px and the same for
// y.  The upper bounds never need a test: for a node inside the picture (the
// only ones whose keys are used) a displacement past W-1-px puts the whole
// block into the edge-replicated padding, so its SAD equals the SAD at
// dx = W-1-px (whose left column is the last picture column and the rest
// padding) while its MV bits are no smaller (bits grow with |dx|,
// dx > W-1-px >= 0) and its raster index is larger: its key can never be the
// minimum.  The same holds below the picture and in the corner.  The lower
// bounds do need it, where they bind: only for R >= 128, in every band of a
// segment at the left of the picture whose column tile reaches past -127,
// and at the top of the picture in the bands that reach past -127 (a
// segment's first bands: dy grows with the band).  Bands without the test and
// with sh == 0 (every 10-bit configuration and 8-bit with R <= 127) take the
// fast path (EDGE false, SH0 true); the others test each node's window and
// shift the SAD.  merge: see search_segment.
template <int MODE, int WAVES, int NS, bool EDGE, bool SH0>
__device__ __attribute__((noinline)) void search_piece(IntArgsC ap, u32 job, u32 ref, u32 bx, u32 band0, u32 band1,
                                                       u64 kb_addr, u32 flags)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const IntArgs &a = *(const IntArgs *)(IntArgsC)(uintptr_t)uni64((u64)(uintptr_t)ap);
    job = uni(job);
    int ctu;
    const PicDesc &P = *job_decode(a.js, (int)job, ctu);
    Block b;
    b.ref = (int)uni(ref);
    b.cx = (ctu % P.ctu_cols) * 128;
    b.cy = (ctu / P.ctu_cols) * 128;
    b.dy0 = 0;
    b.dxs = -a.R + (int)uni(bx) * a.dx_cols;
    b.dxe = b.dxs + a.dx_cols - 1 < a.hi ? b.dxs + a.dx_cols - 1 : a.hi;
    flags = uni(flags);
    search_segment<MODE, WAVES, NS, EDGE, SH0>(a, P, b, (int)uni(band0), (int)uni(band1), smem,
                                              (u64 *)(uintptr_t)uni64(kb_addr), flags & 1, (flags >> 1) & 1);
}

// One workgroup per unit (atomic layouts, groups == units): the search of
// one band inline, as a kernel of its own -- no segment loop around it and
// no call.
template <int MODE, int WAVES>
__device__ __forceinline__ void search_unit(const IntArgs &a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NS = MODE == 1 ? 1 : N_SMALL + N_LARGE;
    // units in XCD order: the bands of a segment read overlapping windows
    const u32 nb = (u32)a.kl.nbands, u = me_xcd_order(blockIdx.x, gridDim.x);
    const u32 x = u / nb;
    const int band = (int)(u - x * nb);
    const int bx = (int)(x % (u32)a.kl.nbx);
    const u32 jr = x / (u32)a.kl.nbx;
    const int job = (int)(jr / (u32)a.nref);
    int ctu;
    const PicDesc &P = *job_decode(a.js, job, ctu);
    Block b;
    b.ref = (int)(jr % (u32)a.nref);
    b.cx = (ctu % P.ctu_cols) * 128;
    b.cy = (ctu / P.ctu_cols) * 128;
    b.dy0 = 0;
    b.dxs = -a.R + bx * a.dx_cols;
    b.dxe = b.dxs + a.dx_cols - 1 < a.hi ? b.dxs + a.dx_cols - 1 : a.hi;
    u64 *kb = a.keys + (size_t)job * a.npu * a.nref + b.ref;
    const bool edge = b.cx + b.dxs < -127 || b.cy - a.R + band * a.band_rows < -127;
    if (!edge && a.sh == 0) search_segment<MODE, WAVES, NS, false, true>(a, P, b, band, band + 1, smem, kb, true, false);
    else search_segment<MODE, WAVES, NS, true, false>(a, P, b, band, band + 1, smem, kb, true, false);
}

// Workgroup g of the partition (KeyLayout) with 64*WAVES threads works
// through its units segment by segment (PERSIST), or takes one unit
// (atomic layouts).  A segment with both window-testing and fast bands runs
// them as two searches, the second merging its keys into the first's.
template <int MODE, int WAVES, bool PERSIST>
// The register budget is 2 waves per SIMD: 2 workgroups of 4 waves, or one
// of 8 (the 10-bit windows at R = 128 leave room for only one workgroup's
// LDS per CU).
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_int_search(IntArgs a)
{
    if (a.js.n_jobs == 0) return;   // a warm-up launch
    if (!PERSIST) {
        search_unit<MODE, WAVES>(a);
        return;
    }
    const IntArgsC ap = (IntArgsC)__builtin_amdgcn_kernarg_segment_ptr();   // a is the only argument
    constexpr int NS = MODE == 1 ? 1 : N_SMALL + N_LARGE;
    const KeyLayout &kl = a.kl;
    // workgroups in XCD order: neighbouring workgroups share a segment
    const u32 g = me_xcd_order(blockIdx.x, gridDim.x), nb = (u32)kl.nbands;
    const u32 ub = (u32)((u64)kl.units * g / kl.groups), ue = (u32)((u64)kl.units * (g + 1) / kl.groups);
    for (u32 u = ub; u < ue;) {
        const u32 x = u / nb;
        const u32 band0 = u - x * nb;
        const u32 left = ue - u;
        const u32 band1 = left < nb - band0 ? band0 + left : nb;
        const u32 bx = x % (u32)kl.nbx;
        const u32 jr = x / (u32)kl.nbx;
        const u32 job = jr / (u32)a.nref, ref = jr % (u32)a.nref;
        int ctu;
        const PicDesc &P = *job_decode(a.js, (int)job, ctu);
        const int cx = (ctu % P.ctu_cols) * 128, cy = (ctu / P.ctu_cols) * 128;
        // this workgroup's piece of the segment
        const u64 kba = (u64)(uintptr_t)(a.keys + (size_t)job * a.npu * a.nref + ref +
                                         (size_t)(bx * kl.nslots + (int)(g - key_owner(kl, (u64)x * nb))) * kl.slot_elems);
        if (cx - a.R + (int)bx * a.dx_cols < -127 || a.sh != 0) {
            search_piece<MODE, WAVES, NS, true, false>(ap, job, ref, bx, band0, band1, kba, 0);
        } else {
            const int need = a.R - 127 - cy;   // window-testing bands: -R + band * band_rows < -127 - cy
            u32 split = need <= 0 ? 0u : (u32)((need + a.band_rows - 1) / a.band_rows);
            split = split < band0 ? band0 : (split > band1 ? band1 : split);
            if (split > band0) search_piece<MODE, WAVES, NS, true, false>(ap, job, ref, bx, band0, split, kba, 0);
            if (band1 > split)
                search_piece<MODE, WAVES, NS, false, true>(ap, job, ref, bx, split, band1, kba, split > band0 ? 2u : 0u);
        }
        u += band1 - band0;
    }
}
#undef PUB

#define ME_IS_CAT2(a, b) a##b
#define ME_IS_CAT(a, b) ME_IS_CAT2(a, b)

int ME_IS_CAT(me_int_search_launch_w, ME_IS_WAVES)(int mode, const IntArgs &a, size_t lds, hipStream_t st)
{
    const dim3 grid(a.kl.groups ? a.kl.groups : 1), block(64 * ME_IS_WAVES);
    const bool persist = a.kl.nslots > 0;
    const void *fn = mode == 1 ? (persist ? (const void *)k_int_search<1, ME_IS_WAVES, true>
                                          : (const void *)k_int_search<1, ME_IS_WAVES, false>)
                               : (persist ? (const void *)k_int_search<0, ME_IS_WAVES, true>
                                          : (const void *)k_int_search<0, ME_IS_WAVES, false>);
    if (lds > 65536) ME_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (mode == 1 && persist) hipLaunchKernelGGL((k_int_search<1, ME_IS_WAVES, true>), grid, block, lds, st, a);
    else if (mode == 1) hipLaunchKernelGGL((k_int_search<1, ME_IS_WAVES, false>), grid, block, lds, st, a);
    else if (persist) hipLaunchKernelGGL((k_int_search<0, ME_IS_WAVES, true>), grid, block, lds, st, a);
    else hipLaunchKernelGGL((k_int_search<0, ME_IS_WAVES, false>), grid, block, lds, st, a);
    ME_CHECK(hipGetLastError());
    return 0;
}

int ME_IS_CAT(me_int_search_preload_w, ME_IS_WAVES)(int mode, size_t lds, hipStream_t st)
{
    // an empty launch (no jobs): the first launch of a kernel in a process
    // costs milliseconds, which a request should not wait for
    IntArgs a;
    memset(&a, 0, sizeof(a));
    return ME_IS_CAT(me_int_search_launch_w, ME_IS_WAVES)(mode, a, lds, st);
}
