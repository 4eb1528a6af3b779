// Batched frame motion estimation for gfx950.
//
// Path per frame (one hipGraph):
//   [optional] pad + 15 quarter-pel planes of a just-reconstructed reference
//   k_fill_keys       -> key buffer := all ones
//   k_int_search      -> integer full search of every PU, every reference
//   k_subpel          -> me_sub_pel_search + final cost + reference choice
//
// Reference semantics (file:line in uavs3e):
//   integer cost  (sad << (16-(bd-8))) + (bits_x+bits_y)*lambda   me.c:119,151-187
//   window        get_max_search_range (is_padding = 1)           me.c:502-527
//   full search   raster y-major, first strict minimum wins        me.c:150-190
//   sub-pel       me_sub_pel_search, cost_type 0 (SATD)            me.c:366-500
//   final cost    + lambda*(refi_bits+mvr+(mvr<4)+2), (c+2^15)>>16 me.c:560-567
#include "common.h"
#include "satd.h"
#include "../../include/uavs3e_me_hip.h"

#include <string.h>
#include <vector>

#define ME_MAX_REFS 4

// ---------------------------------------------------------------------------
// Reference pictures: padded integer plane + 15 quarter-pel planes
// ---------------------------------------------------------------------------

struct me_ref_s {
    int w, h, stride, rows;
    size_t plane_elems;      // stride * rows
    pel *planes;             // 16 planes, plane i at planes + i * plane_elems
    pel *luma;               // unpadded w x h copy (source for re-preparation)
    s16 *tmp;                // 3 horizontal s16 planes over [-8, w+8) x [-8, h+8)
};

// Edge padding of the integer plane (com_img_padding(img, .., 0),
// com_util.c:105-165): every padded sample copies the nearest picture sample.
__global__ void k_pad_int(const pel *__restrict__ luma, int w, int h, pel *__restrict__ plane0, int stride, int rows)
{
    int X = blockIdx.x * blockDim.x + threadIdx.x;
    int Y = blockIdx.y;
    if (X >= stride || Y >= rows) return;
    int x = me_clip3(0, w - 1, X - ME_PAD);
    int y = me_clip3(0, h - 1, Y - ME_PAD);
    plane0[(size_t)Y * stride + X] = luma[(size_t)y * w + x];
}

// Horizontal s16 intermediates of com_if_filter_hor_8 (com_mc.c:668-702):
// t = sum(taps[dx] * P0), stored as (t + add) >> (bd - 8), over
// [-8, w+8) x [-8, h+8) for dx = 1..3.
__global__ void k_htmp(const pel *__restrict__ p0, int stride, int w, int h, s16 *__restrict__ tmp)
{
    int tw = w + 16, th = h + 16;
    int xi = blockIdx.x * blockDim.x + threadIdx.x;
    int yi = blockIdx.y;
    int dx = blockIdx.z + 1;
    if (xi >= tw || yi >= th) return;
    const pel *s = p0 + (ptrdiff_t)(yi - 8) * stride + (xi - 8);
    int t = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) t += (int)s[k - 3] * ME_LUMA_TAPS[dx][k];
    const int sh = BIT_DEPTH - 8, add = (1 << sh) >> 1;
    tmp[(size_t)(dx - 1) * tw * th + (size_t)yi * tw + xi] = (s16)((t + add) >> sh);
}

// The 15 sub-pel planes of com_if_luma_frame (com_mc.c:760-800) including
// the final com_img_padding(plane, 1, 4): every padded sample equals the
// filter output at its position clamped into the [-4, w+4) x [-4, h+4) band.
//   dy == 0: clip((H8(P0) + 32) >> 6)              com_mc.c:668-702
//   dx == 0: clip((V8(P0) + 32) >> 6)              com_mc.c:704-729
//   else:    clip((V8(tmp) + 2^(19-bd)) >> (20-bd)) com_mc.c:731-758
__global__ void k_planes(pel *__restrict__ planes, size_t plane_elems, int stride, int rows, int w, int h,
                         const s16 *__restrict__ tmp)
{
    int X = blockIdx.x * blockDim.x + threadIdx.x;
    int Y = blockIdx.y;
    int idx = blockIdx.z + 1;
    if (X >= stride || Y >= rows) return;
    int dx = idx & 3, dy = idx >> 2;
    int x = me_clip3(-4, w + 3, X - ME_PAD);
    int y = me_clip3(-4, h + 3, Y - ME_PAD);
    const int maxv = (1 << BIT_DEPTH) - 1;
    const pel *p0 = planes + (size_t)ME_PAD * stride + ME_PAD;
    int v;
    if (dy == 0) {
        const pel *s = p0 + (ptrdiff_t)y * stride + x;
        int t = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) t += (int)s[k - 3] * ME_LUMA_TAPS[dx][k];
        v = (t + 32) >> 6;
    } else if (dx == 0) {
        const pel *s = p0 + (ptrdiff_t)y * stride + x;
        int t = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) t += (int)s[(ptrdiff_t)(k - 3) * stride] * ME_LUMA_TAPS[dy][k];
        v = (t + 32) >> 6;
    } else {
        int tw = w + 16, th = h + 16;
        const s16 *s = tmp + (size_t)(dx - 1) * tw * th + (size_t)(y + 8) * tw + (x + 8);
        int t = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) t += (int)s[(ptrdiff_t)(k - 3) * tw] * ME_LUMA_TAPS[dy][k];
        const int sh2 = 20 - BIT_DEPTH, add2 = 1 << (sh2 - 1);
        v = (t + add2) >> sh2;
    }
    planes[(size_t)idx * plane_elems + (size_t)Y * stride + X] = (pel)me_clip3(0, maxv, v);
}

static int ref_prepare_async(me_ref_s *r, hipStream_t st)
{
    dim3 b(256);
    hipLaunchKernelGGL(k_pad_int, dim3((r->stride + 255) / 256, r->rows), b, 0, st, r->luma, r->w, r->h,
                       r->planes, r->stride, r->rows);
    hipLaunchKernelGGL(k_htmp, dim3((r->w + 16 + 255) / 256, r->h + 16, 3), b, 0, st,
                       r->planes + (size_t)ME_PAD * r->stride + ME_PAD, r->stride, r->w, r->h, r->tmp);
    hipLaunchKernelGGL(k_planes, dim3((r->stride + 255) / 256, r->rows, 15), b, 0, st, r->planes, r->plane_elems,
                       r->stride, r->rows, r->w, r->h, r->tmp);
    ME_CHECK(hipGetLastError());
    return 0;
}

ME_EXPORT int me_ref_create(me_ref_t *out, int pic_w, int pic_h)
{
    if (me_hip_init(-1)) return -1;
    if (!out || pic_w <= 0 || pic_h <= 0) { me_set_error("me_ref_create: bad size", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    me_ref_s *r = (me_ref_s *)calloc(1, sizeof(me_ref_s));
    r->w = pic_w; r->h = pic_h;
    r->stride = pic_w + 2 * ME_PAD;
    r->rows = pic_h + 2 * ME_PAD;
    r->plane_elems = (size_t)r->stride * r->rows;
    hipError_t e = hipMalloc((void **)&r->planes, 16 * r->plane_elems * sizeof(pel));
    if (e == hipSuccess) e = hipMalloc((void **)&r->luma, (size_t)pic_w * pic_h * sizeof(pel));
    if (e == hipSuccess) e = hipMalloc((void **)&r->tmp, 3 * (size_t)(pic_w + 16) * (pic_h + 16) * sizeof(s16));
    if (e != hipSuccess) {
        me_set_error("me_ref_create: hipMalloc", e, __FILE__, __LINE__);
        hipFree(r->planes); hipFree(r->luma); hipFree(r->tmp); free(r);
        return -1;
    }
    *out = r;
    return 0;
}

ME_EXPORT int me_ref_load(me_ref_t r, const me_pel *luma, int stride, int on_device, void *stream)
{
    if (!r || !luma || stride < r->w) { me_set_error("me_ref_load: bad argument", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    hipStream_t st = (hipStream_t)stream;
    ME_CHECK(hipMemcpy2DAsync(r->luma, r->w * sizeof(pel), luma, stride * sizeof(pel), r->w * sizeof(pel), r->h,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    if (ref_prepare_async(r, st)) return -1;
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

ME_EXPORT int me_ref_plane(me_ref_t r, int idx, me_pel **plane, int *stride, int *pad)
{
    if (!r || idx < 0 || idx > 15) return -1;
    if (plane) *plane = r->planes + (size_t)idx * r->plane_elems + (size_t)ME_PAD * r->stride + ME_PAD;
    if (stride) *stride = r->stride;
    if (pad) *pad = ME_PAD;
    return 0;
}

ME_EXPORT int me_ref_download_plane(me_ref_t r, int idx, me_pel *out)
{
    if (!r || !out || idx < 0 || idx > 15) { me_set_error("me_ref_download_plane: bad argument", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    ME_CHECK(hipMemcpy(out, r->planes + (size_t)idx * r->plane_elems, r->plane_elems * sizeof(pel), hipMemcpyDeviceToHost));
    return 0;
}

ME_EXPORT void me_ref_destroy(me_ref_t r)
{
    if (!r) return;
    hipFree(r->planes); hipFree(r->luma); hipFree(r->tmp);
    free(r);
}

// ---------------------------------------------------------------------------
// Shapes and PU numbering
// ---------------------------------------------------------------------------

ME_EXPORT int me_shape_w(int s) { return s >= 0 && s < ME_NSHAPES ? ME_SHAPES_HOST[s][0] : 0; }
ME_EXPORT int me_shape_h(int s) { return s >= 0 && s < ME_NSHAPES ? ME_SHAPES_HOST[s][1] : 0; }
ME_EXPORT int me_pus_per_ctu(uint32_t mask)
{
    int n = 0;
    for (int s = 0; s < ME_NSHAPES; s++)
        if (mask >> s & 1) n += (128 / ME_SHAPES_HOST[s][0]) * (128 / ME_SHAPES_HOST[s][1]);
    return n;
}

enum {
    S4x8 = 0, S8x4, S4x16, S16x4, S4x32, S32x4, S8x8, S8x16, S16x8, S8x32, S32x8, S8x64, S64x8, S16x16,
    S16x32, S32x16, S16x64, S64x16, S16x128, S128x16, S32x32, S32x64, S64x32, S32x128, S128x32, S64x64,
    S64x128, S128x64, S128x128
};

// ---------------------------------------------------------------------------
// Integer full search
// ---------------------------------------------------------------------------

// Keys: a candidate's cost c (u64, < 2^41) and its raster index i in the
// (2R+1)^2 window are packed as key = (c << IB) | i, so the smallest key is
// the smallest cost with the first raster position on ties -- the reference's
// strict '<' scan (me.c:169-172, 187).  With cost = (sad << shift) + rate and
// IB >= 32 - shift, key = sad * 2^(shift+IB) + K(d), K(d) = (rate << IB) | i,
// so the low word of a key is K's low word and its high word is
// (sad << (shift+IB-32)) + K_hi(d): one shift-add per candidate.  KEY_BIAS
// lifts every high word above 2^28 so that keys are normal doubles.
// Valid high words stay below 2^29 (sad < 2^24 for a 128x128 PU at 10 bits,
// K_hi < 2^15).  A node outside the search window sets bit 30; a lane whose
// displacement lies past the block's range (the tail of the last group of
// four) gets KEY_OUT_HI plus its SAD.  Both stay below 0x7FF00000, so every
// key is a finite positive double that v_min_f64 orders like the u64 key.
#define KEY_BIAS_HI 0x10000000u
#define KEY_INVALID_HI 0x40000000u
#define KEY_OUT_HI 0x60000000u
#define KEY_INIT 0x7FF0000000000000ull

struct IntArgs {
    const pel *cur;          // current picture, stride pic_w
    const pel *ref0[ME_MAX_REFS];  // integer plane of each reference at (0,0)
    int ref_stride;
    int pic_w, pic_h, ctu_cols;
    int R;                   // window half size
    int lo, hi;              // uniform displacement clip: [max(-R,-T+4), min(R,T-4)]
    int IB, sh;              // index bits; sad shift into the high word
    const u32 *rate;         // [2R+1]: bits(4*(i-R)) * lambda
    u64 *keys;               // [ctu][npu][nref]
    int npu, nref;
    int band_rows, wcols;    // displacement rows per block; LDS window row length
    int dx_cols, nbx;        // displacement columns per block; blocks along x
    int wpitch;              // LDS row pitch of the window (wcols + skew room)
    int base[ME_NSHAPES];    // first PU index of each shape inside a CTU (-1: not selected)
};

// Values of NK displacements of one node: at 8 bits four displacements are
// packed as u16 lanes (lo = d0,d1; hi = d2,d3) -- a sum of at most 256
// samples stays below 2^16, so plain 32-bit adds never carry between lanes;
// at 10 bits one displacement per u32.
#if BIT_DEPTH == 8
#define NK 4
struct Val {
    u32 lo, hi;
    __device__ __forceinline__ u32 get(int k) const
    {
        const u32 w = k < 2 ? lo : hi;
        return (k & 1) ? (w >> 16) : (w & 0xFFFFu);
    }
};
__device__ __forceinline__ Val operator+(Val a, Val b) { return {a.lo + b.lo, a.hi + b.hi}; }
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

// Uniform key parts of the NK displacements being processed.  A displacement
// outside the block's range gets a high word that can never win.
struct KeyCtx {
    u32 khi[NK], klo[NK];
    int dx[NK];
    int dy, sh;
};

typedef u32 u32x2 __attribute__((ext_vector_type(2)));

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
    u32x2 key = {c.klo[k], hi};
    st = me_vmin64(st, __builtin_bit_cast(double, key));
}

// Cross-lane partner values.  Lane = ry*8 + rx (16x16 regions of the CTU).
//   xor 1, xor 2 : DPP quad permutes        (x pairs, x quads)
//   x octet      : DPP row_half_mirror       (the other x quad of a row of 8)
//   xor 8        : DPP row_ror:8             (y pairs)
//   xor 16       : ds_swizzle xor mode       (y quads)
//   xor 32       : ds_bpermute               (the other half of the CTU)
__device__ __forceinline__ u32 px1(u32 v) { return (u32)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false); }
__device__ __forceinline__ u32 px2(u32 v) { return (u32)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false); }
__device__ __forceinline__ u32 px4(u32 v) { return (u32)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false); }
__device__ __forceinline__ u32 py1(u32 v) { return (u32)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false); }
__device__ __forceinline__ u32 py2(u32 v) { return (u32)__builtin_amdgcn_ds_swizzle((int)v, 0x401F); }
__device__ __forceinline__ u32 py4(u32 v) { return (u32)__shfl_xor((int)v, 32, 64); }
#if BIT_DEPTH == 8
__device__ __forceinline__ Val vpx1(Val a) { return {px1(a.lo), px1(a.hi)}; }
__device__ __forceinline__ Val vpy1(Val a) { return {py1(a.lo), py1(a.hi)}; }
#else
__device__ __forceinline__ Val vpx1(Val a) { return {px1(a.v)}; }
__device__ __forceinline__ Val vpy1(Val a) { return {py1(a.v)}; }
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
#define N_SMALL 33
template <bool EDGE, bool SH0>
__device__ __forceinline__ void pass_small(const Region &g, double *st, const KeyCtx &c, int px, int py, int wlim, int hlim)
{
#pragma unroll
    for (int k = 0; k < NK; k++) {
#define U(I, V, X_, Y_) upd1<EDGE, SH0>(st[I], (V).get(k), c, k, X_, Y_, wlim, hlim)
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

// Pass L: the 609 PUs larger than a 16x16 region.
// Pairs of regions (x pairs, y pairs) hold 14 nodes between two lanes; each
// lane keeps half of them (3 "dense" slots per direction, chosen by the lane's
// rank in the pair).  The 19 sparser classes are summed on every lane of their
// group and kept by all of them; one lane per group publishes.
//   st[0..2] x-pair slots: 32x4 rows {0,2}, {1,3} by rank; 32x8 row = rank
//   st[3..5] y-pair slots: 4x32 cols {0,2}, {1,3}; 8x32 col = rank
//   st[6] 32x16  st[7] 16x32  st[8] 32x32  st[9..10] 64x8 (r2)  st[11] 64x16
//   st[12..13] 8x64 (c2)  st[14] 16x64  st[15] 64x32  st[16] 32x64  st[17] 64x64
//   st[18] 128x16  st[19] 16x128  st[20] 128x32  st[21] 32x128  st[22] 128x64
//   st[23] 64x128  st[24] 128x128
#define N_LARGE 25
template <bool EDGE, bool SH0>
__device__ __forceinline__ void pass_large(const Region &g, double *st, const KeyCtx &c, int rx, int ry, int cx, int cy,
                                           int wlim, int hlim)
{
    const int X = rx * 16, Y = ry * 16, px = cx + X, py = cy + Y;
    const bool xr = rx & 1, yr = ry & 1;
    const int x2 = cx + (X & ~31), x4 = cx + (X & ~63), y2 = cy + (Y & ~31), y4 = cy + (Y & ~63);
    // dense x pairs
    Val t32x4[4], t32x8[2], t4x32[4], t8x32[2];
#pragma unroll
    for (int r = 0; r < 4; r++) t32x4[r] = g.h16[r] + vpx1(g.h16[r]);
#pragma unroll
    for (int r2 = 0; r2 < 2; r2++) t32x8[r2] = g.h16x8[r2] + vpx1(g.h16x8[r2]);
#pragma unroll
    for (int cc = 0; cc < 4; cc++) t4x32[cc] = g.v16[cc] + vpy1(g.v16[cc]);
#pragma unroll
    for (int c2 = 0; c2 < 2; c2++) t8x32[c2] = g.v8x16[c2] + vpy1(g.v8x16[c2]);
    const Val d0 = vsel(xr, t32x4[2], t32x4[0]), d1 = vsel(xr, t32x4[3], t32x4[1]), d2 = vsel(xr, t32x8[1], t32x8[0]);
    const Val d3 = vsel(yr, t4x32[2], t4x32[0]), d4 = vsel(yr, t4x32[3], t4x32[1]), d5 = vsel(yr, t8x32[1], t8x32[0]);
    // sparse classes are summed one displacement at a time (sums exceed 16 bits)
#pragma unroll
    for (int k = 0; k < NK; k++) {
        upd1<EDGE, SH0>(st[0], d0.get(k), c, k, x2, py + (xr ? 8 : 0), wlim, hlim);
        upd1<EDGE, SH0>(st[1], d1.get(k), c, k, x2, py + (xr ? 12 : 4), wlim, hlim);
        upd1<EDGE, SH0>(st[2], d2.get(k), c, k, x2, py + (xr ? 8 : 0), wlim, hlim);
        upd1<EDGE, SH0>(st[3], d3.get(k), c, k, px + (yr ? 8 : 0), y2, wlim, hlim);
        upd1<EDGE, SH0>(st[4], d4.get(k), c, k, px + (yr ? 12 : 4), y2, wlim, hlim);
        upd1<EDGE, SH0>(st[5], d5.get(k), c, k, px + (yr ? 8 : 0), y2, wlim, hlim);
        const u32 s = g.s16.get(k);
        const u32 t32x16 = s + px1(s), t16x32 = s + py1(s);
        const u32 t32x32 = t32x16 + py1(t32x16);
        const u32 a8 = t32x8[0].get(k), b8 = t32x8[1].get(k), c8 = t8x32[0].get(k), d8 = t8x32[1].get(k);
        const u32 t64x8a = a8 + px2(a8), t64x8b = b8 + px2(b8);
        const u32 t8x64a = c8 + py2(c8), t8x64b = d8 + py2(d8);
        const u32 t64x16 = t32x16 + px2(t32x16), t16x64 = t16x32 + py2(t16x32);
        const u32 t64x32 = t32x32 + px2(t32x32), t32x64 = t32x32 + py2(t32x32);
        const u32 t64x64 = t64x32 + py2(t64x32);
        const u32 t128x16 = t64x16 + px4(t64x16), t16x128 = t16x64 + py4(t16x64);
        const u32 t128x32 = t64x32 + px4(t64x32), t32x128 = t32x64 + py4(t32x64);
        const u32 t128x64 = t64x64 + px4(t64x64), t64x128 = t64x64 + py4(t64x64);
        const u32 t128x128 = t128x64 + py4(t128x64);
        upd1<EDGE, SH0>(st[6], t32x16, c, k, x2, py, wlim, hlim);
        upd1<EDGE, SH0>(st[7], t16x32, c, k, px, y2, wlim, hlim);
        upd1<EDGE, SH0>(st[8], t32x32, c, k, x2, y2, wlim, hlim);
        upd1<EDGE, SH0>(st[9], t64x8a, c, k, x4, py, wlim, hlim);
        upd1<EDGE, SH0>(st[10], t64x8b, c, k, x4, py + 8, wlim, hlim);
        upd1<EDGE, SH0>(st[11], t64x16, c, k, x4, py, wlim, hlim);
        upd1<EDGE, SH0>(st[12], t8x64a, c, k, px, y4, wlim, hlim);
        upd1<EDGE, SH0>(st[13], t8x64b, c, k, px + 8, y4, wlim, hlim);
        upd1<EDGE, SH0>(st[14], t16x64, c, k, px, y4, wlim, hlim);
        upd1<EDGE, SH0>(st[15], t64x32, c, k, x4, y2, wlim, hlim);
        upd1<EDGE, SH0>(st[16], t32x64, c, k, x2, y4, wlim, hlim);
        upd1<EDGE, SH0>(st[17], t64x64, c, k, x4, y4, wlim, hlim);
        upd1<EDGE, SH0>(st[18], t128x16, c, k, cx, py, wlim, hlim);
        upd1<EDGE, SH0>(st[19], t16x128, c, k, px, cy, wlim, hlim);
        upd1<EDGE, SH0>(st[20], t128x32, c, k, cx, y2, wlim, hlim);
        upd1<EDGE, SH0>(st[21], t32x128, c, k, x2, cy, wlim, hlim);
        upd1<EDGE, SH0>(st[22], t128x64, c, k, cx, y4, wlim, hlim);
        upd1<EDGE, SH0>(st[23], t64x128, c, k, x4, cy, wlim, hlim);
        upd1<EDGE, SH0>(st[24], t128x128, c, k, cx, cy, wlim, hlim);
    }
}

template <bool EDGE, bool SH0>
__device__ __forceinline__ void pass_16(const Region &g, double *st, const KeyCtx &c, int px, int py, int wlim, int hlim)
{
#pragma unroll
    for (int k = 0; k < NK; k++) upd1<EDGE, SH0>(st[0], g.s16.get(k), c, k, px, py, wlim, hlim);
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
                                       int k0, Val L[4][4])
{
    (void)k0;
    u64 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const u32 *rp = (const u32 *)(win + (wrow0 + r) * wcols + wskew(wrow0 + r) + c0);
        u32 d[5];
#pragma unroll
        for (int j = 0; j < 5; j++) d[j] = rp[j];
        const uint4 cw = *(const uint4 *)(curl + (crow0 + r) * 128 + ccol0);
        const u32 cc[4] = {cw.x, cw.y, cw.z, cw.w};
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[r >> 2][j] = __builtin_amdgcn_qsad_pk_u16_u8(((u64)d[j + 1] << 32) | d[j], cc[j], acc[r >> 2][j]);
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) L[r][c] = {(u32)acc[r][c], (u32)(acc[r][c] >> 32)};
}
#else
// v_sad_u16 on sample pairs; odd displacements realign the pairs with
// v_alignbyte.  One displacement (dxb + k0) per call.
__device__ __forceinline__ void leaves(const pel *win, int wcols, int wrow0, int c0, const pel *curl, int crow0, int ccol0,
                                       int k0, Val L[4][4])
{
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) L[r][c].v = 0;
#pragma unroll 4
    for (int r = 0; r < 16; r++) {
        const u32 *rp = (const u32 *)(win + (wrow0 + r) * wcols + wskew(wrow0 + r) + c0 + (k0 & ~1));
        u32 d[9];
#pragma unroll
        for (int j = 0; j < 9; j++) d[j] = rp[j];
        const uint4 c01 = *(const uint4 *)(curl + (crow0 + r) * 128 + ccol0);
        const uint4 c23 = *(const uint4 *)(curl + (crow0 + r) * 128 + ccol0 + 8);
        const u32 cc[8] = {c01.x, c01.y, c01.z, c01.w, c23.x, c23.y, c23.z, c23.w};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            u32 p0, p1;
            if (k0 & 1) {
                p0 = __builtin_amdgcn_alignbyte(d[2 * c + 1], d[2 * c], 2);
                p1 = __builtin_amdgcn_alignbyte(d[2 * c + 2], d[2 * c + 1], 2);
            } else {
                p0 = d[2 * c];
                p1 = d[2 * c + 1];
            }
            L[r >> 2][c].v = __builtin_amdgcn_sad_u16(p0, cc[2 * c], __builtin_amdgcn_sad_u16(p1, cc[2 * c + 1], L[r >> 2][c].v));
        }
    }
}
#endif

struct Block {
    int cx, cy, dy0, dxs, dxe, ref;
};

// One pass over the block's displacements: PASS 0 = small PUs, 1 = large
// PUs, 2 = 16x16 only.  Returns the node states in st.
template <int PASS, int WAVES, bool EDGE, bool SH0, int NS>
__device__ __forceinline__ void run_pass(const IntArgs &a, const Block &b, const pel *win, const pel *curl, double *st)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rx = lane & 7, ry = lane >> 3, X = rx * 16, Y = ry * 16;
    const int wlim = a.pic_w + 126, hlim = a.pic_h + 126;
    const int span = 2 * a.R + 1;
#pragma unroll
    for (int i = 0; i < NS; i++) st[i] = __builtin_bit_cast(double, KEY_INIT);
    for (int row = wave; row < a.band_rows; row += WAVES) {
        const int dy = b.dy0 + row;
        if (dy < a.lo || dy > a.hi) continue;
        const int dyi = dy + a.R;
        const u32 ry_rate = a.rate[dyi];
        for (int dxb = b.dxs; dxb <= b.dxe; dxb += 4) {
            if (dxb + 3 < a.lo) continue;
#pragma unroll
            for (int k0 = 0; k0 < 4; k0 += NK) {
                if (NK == 1 && (dxb + k0 < a.lo || dxb + k0 > b.dxe)) continue;
                KeyCtx kc;
                kc.dy = dy;
                kc.sh = a.sh;
#pragma unroll
                for (int k = 0; k < NK; k++) {
                    const int dx = dxb + k0 + k;
                    kc.dx[k] = dx;
                    if (dx < a.lo || dx > b.dxe) {
                        kc.khi[k] = KEY_OUT_HI;   // above every valid key, never NaN
                        kc.klo[k] = 0;
                    } else {
                        const int dxi = dx + a.R;
                        const u64 K = ((u64)(a.rate[dxi] + ry_rate) << a.IB) | (u64)(dyi * span + dxi);
                        kc.khi[k] = (u32)(K >> 32) + KEY_BIAS_HI;
                        kc.klo[k] = (u32)K;
                    }
                }
                Val L[4][4];
                leaves(win, a.wpitch, Y + row, X + (dxb - b.dxs), curl, Y, X, k0, L);
                Region g;
                g.build(L);
                if (PASS == 0) pass_small<EDGE, SH0>(g, st, kc, b.cx + X, b.cy + Y, wlim, hlim);
                else if (PASS == 1) pass_large<EDGE, SH0>(g, st, kc, rx, ry, b.cx, b.cy, wlim, hlim);
                else pass_16<EDGE, SH0>(g, st, kc, b.cx + X, b.cy + Y, wlim, hlim);
            }
        }
    }
}

#define PUB(S, W, H, XX, YY, ST)                                                                    \
    do {                                                                                            \
        if (a.base[S] >= 0) atomic_key_min(kb + (size_t)pu_index(a, S, W, H, XX, YY) * a.nref, ST); \
    } while (0)

__device__ __forceinline__ void publish_small(const IntArgs &a, u64 *kb, const double *st, int X, int Y)
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

__device__ __forceinline__ void publish_large(const IntArgs &a, u64 *kb, const double *st, int rx, int ry)
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
    if (ey3) PUB(S16x128, 16, 128, X, Y, st[19]);
    if (ex3 && ey1) PUB(S128x32, 128, 32, X, Y, st[20]);
    if (ex1 && ey3) PUB(S32x128, 32, 128, X, Y, st[21]);
    if (ex3 && ey2) PUB(S128x64, 128, 64, X, Y, st[22]);
    if (ex2 && ey3) PUB(S64x128, 64, 128, X, Y, st[23]);
    if (ex3 && ey3) PUB(S128x128, 128, 128, X, Y, st[24]);
}

// One block = (displacement tile, CTU, reference); 64*WAVES threads.  The
// reference window of the tile and the current CTU are staged in LDS once;
// then each wave covers the whole CTU (lane = 16x16 region) for every
// WAVES-th displacement row of the tile, in two passes (small PUs, large
// PUs) so that each pass's node states fit in registers.
template <int MODE, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_int_search(IntArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    Block b;
    const int band = blockIdx.x / a.nbx, bandx = blockIdx.x - band * a.nbx, ctu = blockIdx.y;
    b.ref = blockIdx.z;
    b.cx = (ctu % a.ctu_cols) * 128;
    b.cy = (ctu / a.ctu_cols) * 128;
    b.dy0 = -a.R + band * a.band_rows;
    b.dxs = -a.R + bandx * a.dx_cols;
    b.dxe = b.dxs + a.dx_cols - 1 < a.hi ? b.dxs + a.dx_cols - 1 : a.hi;
    const int wrows = 128 + a.band_rows - 1;
    pel *win = (pel *)smem;
    pel *curl = win + (size_t)wrows * a.wpitch;
    const int tid = threadIdx.x, nthr = 64 * WAVES;

    // Stage the window: reference samples (cx + dxs + c, cy + dy0 + r), clamped
    // to the padded plane (it is edge-replicated, so clamping is exact), and
    // the current CTU clamped to the picture.
    const pel *ref0 = a.ref0[b.ref];
    const int xmin = -ME_PAD, xmax = a.pic_w + ME_PAD - 1, ymin = -ME_PAD, ymax = a.pic_h + ME_PAD - 1;
    for (int i = tid; i < wrows * a.wcols; i += nthr) {
        int r = i / a.wcols, c = i - r * a.wcols;
        int x = me_clip3(xmin, xmax, b.cx + b.dxs + c), y = me_clip3(ymin, ymax, b.cy + b.dy0 + r);
        win[r * a.wpitch + wskew(r) + c] = ref0[(ptrdiff_t)y * a.ref_stride + x];
    }
    for (int i = tid; i < 128 * 128; i += nthr) {
        int r = i >> 7, c = i & 127;
        int x = me_clip3(0, a.pic_w - 1, b.cx + c), y = me_clip3(0, a.pic_h - 1, b.cy + r);
        curl[i] = a.cur[(size_t)y * a.pic_w + x];
    }
    __syncthreads();

    const int lane = tid & 63;
    const int rx = lane & 7, ry = lane >> 3;
    const bool edge = !(b.cx + 124 + a.R <= a.pic_w - 1 && b.cy + 124 + a.R <= a.pic_h - 1 && a.R <= 127);
    u64 *kb = a.keys + (size_t)ctu * a.npu * a.nref + b.ref;
    // Interior blocks with sh == 0 (the common case) take the fast path;
    // the others test each node's window and shift the SAD.
    const bool fast = !edge && a.sh == 0;
    if (MODE == 1) {
        double st[1];
        if (fast) run_pass<2, WAVES, false, true, 1>(a, b, win, curl, st);
        else run_pass<2, WAVES, true, false, 1>(a, b, win, curl, st);
        PUB(S16x16, 16, 16, rx * 16, ry * 16, st[0]);
        return;
    }
    {
        double st[N_SMALL];
        if (fast) run_pass<0, WAVES, false, true, N_SMALL>(a, b, win, curl, st);
        else run_pass<0, WAVES, true, false, N_SMALL>(a, b, win, curl, st);
        publish_small(a, kb, st, rx * 16, ry * 16);
    }
    {
        double st[N_LARGE];
        if (fast) run_pass<1, WAVES, false, true, N_LARGE>(a, b, win, curl, st);
        else run_pass<1, WAVES, true, false, N_LARGE>(a, b, win, curl, st);
        publish_large(a, kb, st, rx, ry);
    }
}
#undef PUB

__global__ void k_fill_keys(u64 *keys, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = ~0ull;
}

// ---------------------------------------------------------------------------
// Sub-pel refinement (me_sub_pel_search, me.c:366-500, cost_type 0, bi 0)
// ---------------------------------------------------------------------------

struct SubArgs {
    const pel *cur;
    const pel *planes[ME_MAX_REFS];  // plane 0 base (padded buffer start) of each reference
    size_t plane_elems;
    int stride;
    int pic_w, pic_h, ctu_cols;
    int R, IB;
    u32 lambda;
    const u8 *tab;              // MV-bits table, centred (tab[0] is mv 0)
    const u64 *keys;
    me_pu_result_t *res;
    s8 *best;
    int npu, nref, n_ctu;
    // the shape handled by one launch
    int pw, ph, tw, th, tpr, base, per_ctu;
};

__device__ __forceinline__ const pel *plane_ptr(const SubArgs &a, int ref, int idx, int x, int y)
{
    return a.planes[ref] + (size_t)idx * a.plane_elems + (size_t)(y + ME_PAD) * a.stride + (x + ME_PAD);
}

// com_mc_blk_luma_pointer (com_mc.c:89-101).
__device__ __forceinline__ const pel *qpel_ptr(const SubArgs &a, int ref, int qx, int qy)
{
    int ix = me_clip3(-ME_MAX_CU - 4, a.pic_w + 4, qx >> 2);
    int iy = me_clip3(-ME_MAX_CU - 4, a.pic_h + 4, qy >> 2);
    return plane_ptr(a, ref, (qy & 3) * 4 + (qx & 3), ix, iy);
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

// Sum of |coef| of the 2-D Walsh-Hadamard transform of a TW x TH residual.
template <int TW, int TH>
__device__ __forceinline__ int had_abs(const pel *org, int so, const pel *cur, int sc)
{
    int m[TW * TH];
#pragma unroll
    for (int y = 0; y < TH; y++) {
        int o[TW], c[TW];
        load_row<TW>(org + y * so, o);
        load_row<TW>(cur + y * sc, c);
#pragma unroll
        for (int x = 0; x < TW; x++) m[y * TW + x] = o[x] - c[x];
    }
#pragma unroll
    for (int y = 0; y < TH; y++) fwht<TW>(m + y * TW, 1);
#pragma unroll
    for (int x = 0; x < TW; x++) fwht<TH>(m + x, TW);
    int s = 0;
#pragma unroll
    for (int i = 0; i < TW * TH; i++) s += abs(m[i]);
    return s;
}

// One normalised cost_satd tile (com_cost.c:259, 340, 467, 586, 649, 708).
// A 16-point WHT is the 8-point WHT of the half sums and of the half
// differences, so a 16x8 (8x16) tile is two 8x8 transforms.
template <int TT>
__device__ __forceinline__ int satd_t(const pel *org, int so, const pel *cur, int sc)
{
    const double SQRT32 = 0x1.6a09e667f3bcdp+2, SQRT128 = 0x1.6a09e667f3bcdp+3;
    if (TT == 0) return (had_abs<4, 4>(org, so, cur, sc) + 1) >> 1;
    if (TT == 1) return (had_abs<8, 8>(org, so, cur, sc) + 2) >> 2;
    if (TT == 2) return (int)((double)had_abs<8, 4>(org, so, cur, sc) / SQRT32 * 2.0);
    if (TT == 4) return (int)((double)had_abs<4, 8>(org, so, cur, sc) / SQRT32 * 2.0);
    // 16x8 (TT 3) and 8x16 (TT 5): rows of the two 8x8 halves are combined
    // into sums and differences as they are read
    int mS[64], mD[64];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        int o1[8], c1[8], o2[8], c2[8];
        if (TT == 3) {
            load_row<8>(org + y * so, o1);
            load_row<8>(cur + y * sc, c1);
            load_row<8>(org + y * so + 8, o2);
            load_row<8>(cur + y * sc + 8, c2);
        } else {
            load_row<8>(org + y * so, o1);
            load_row<8>(cur + y * sc, c1);
            load_row<8>(org + (y + 8) * so, o2);
            load_row<8>(cur + (y + 8) * sc, c2);
        }
#pragma unroll
        for (int x = 0; x < 8; x++) {
            const int l = o1[x] - c1[x], r = o2[x] - c2[x];
            mS[y * 8 + x] = l + r;
            mD[y * 8 + x] = l - r;
        }
    }
    int s = 0;
#pragma unroll
    for (int y = 0; y < 8; y++) { fwht<8>(mS + y * 8, 1); fwht<8>(mD + y * 8, 1); }
#pragma unroll
    for (int x = 0; x < 8; x++) { fwht<8>(mS + x, 8); fwht<8>(mD + x, 8); }
#pragma unroll
    for (int i = 0; i < 64; i++) s += abs(mS[i]) + abs(mD[i]);
    return (int)((double)s / SQRT128 * 2.0);
}

// Sub-pel search of one shape.  A unit is one (CTU, PU of this shape,
// reference); LPU lanes share a unit and each takes T/LPU of its com_had
// tiles (com_cost.c:719-811).  Per search position the lanes' tile SATDs are
// summed across the unit's lanes, so every lane of a unit holds the same
// costs and makes the same (strict '<', reference order) decisions.
template <int TT, int T>
__device__ __forceinline__ void subpel_body(const SubArgs &a)
{
    // lanes per unit: one per tile up to a wave; the 256-tile 128x128 PU
    // takes a whole 256-thread block
    constexpr int LPU = T < 64 ? T : (T < 256 ? 64 : 256), TPL = T / LPU, UPB = 256 / LPU;
    const int li = threadIdx.x % LPU;
    const long unit = (long)blockIdx.x * UPB + threadIdx.x / LPU;
    const long nunits = (long)a.n_ctu * a.per_ctu * a.nref;
    if (unit >= nunits) return;
    const int r = (int)(unit % a.nref);
    const long cp = unit / a.nref;
    const int i = (int)(cp % a.per_ctu), ctu = (int)(cp / a.per_ctu);
    const int cols = 128 / a.pw;
    const int x = (ctu % a.ctu_cols) * 128 + (i % cols) * a.pw, y = (ctu / a.ctu_cols) * 128 + (i / cols) * a.ph;
    const size_t slot = ((size_t)ctu * a.npu + a.base + i) * a.nref + r;
    const bool inside = x + a.pw <= a.pic_w && y + a.ph <= a.pic_h;
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

    const pel *org = a.cur + (size_t)y * a.pic_w + x;
    int toff_o[TPL], toff_p[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        const int t = li + LPU * j;
        const int tx = (t % a.tpr) * a.tw, ty = (t / a.tpr) * a.th;
        toff_o[j] = ty * a.pic_w + tx;
        toff_p[j] = ty * a.stride + tx;
    }
    auto had = [&](const pel *pred) -> u32 {
        int v = 0;
#pragma unroll 1
        for (int j = 0; j < TPL; j++) v += satd_t<TT>(org + toff_o[j], a.pic_w, pred + toff_p[j], a.stride);
#pragma unroll
        for (int m = 1; m < (LPU < 64 ? LPU : 64); m <<= 1) v += __shfl_xor(v, m, 64);
        if (LPU > 64) {
            __shared__ int red[4];
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
            __syncthreads();
            v = red[0] + red[1] + red[2] + red[3];
            __syncthreads();
        }
        return (u32)(v >> (BIT_DEPTH - 8));   // com_had: sum >> (bd - 8)
    };
    const u8 *tb = a.tab;
    const int cmx = dx * 4, cmy = dy * 4;
    // centre: the integer plane, not clipped (me.c:400)
    u64 best = (u64)(a.lambda * (u32)(tb[cmx] + tb[cmy])) + ((u64)had(plane_ptr(a, r, 0, x + dx, y + dy)) << 16);
    int bx = cmx, by = cmy;
    const int HP[8][2] = {{-2, 0}, {-2, 2}, {0, 2}, {2, 2}, {2, 0}, {2, -2}, {0, -2}, {-2, -2}};
    const int QP[8][2] = {{-1, 0}, {0, 1}, {1, 0}, {0, -1}, {-1, 1}, {1, 1}, {-1, -1}, {1, -1}};
#pragma unroll 1
    for (int p = 0; p < 8; p++) {
        const int mx = cmx + HP[p][0], my = cmy + HP[p][1];
        const u64 c = (u64)(a.lambda * (u32)(tb[mx] + tb[my])) + ((u64)had(qpel_ptr(a, r, mx + (x << 2), my + (y << 2))) << 16);
        if (c < best) { best = c; bx = mx; by = my; }
    }
    const int hx = bx, hy = by;
#pragma unroll 1
    for (int p = 0; p < 8; p++) {
        const int mx = hx + QP[p][0], my = hy + QP[p][1];
        const u64 c = (u64)(a.lambda * (u32)(tb[mx] + tb[my])) + ((u64)had(qpel_ptr(a, r, mx + (x << 2), my + (y << 2))) << 16);
        if (c < best) { best = c; bx = mx; by = my; }
    }
    if (li == 0) {
        out.mv_sub[0] = (s16)bx;
        out.mv_sub[1] = (s16)by;
        out.cost_sub = best;
        // tbl_refi_bits[n][r] (tables.c:52-70); AMVR 0 adds 0 + 1 + 2 (me.c:560-561)
        const int rb = a.nref <= 1 ? 0 : (r < a.nref - 1 ? r + 1 : a.nref - 1);
        const u64 cf = best + (u64)(a.lambda * (u32)(rb + 3));
        out.cost_final = (u32)((cf + (1u << 15)) >> 16);
        out.valid = 1;
        a.res[slot] = out;
    }
}

// 8x8-class tiles fit 4 waves per SIMD; the 16x8 / 8x16 tiles (two 8x8
// halves) get the register budget of 2.
template <int TT, int T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_subpel(SubArgs a)
{
    subpel_body<TT, T>(a);
}
template <int TT, int T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_subpel_wide(SubArgs a)
{
    subpel_body<TT, T>(a);
}

// Reference choice per PU: smallest final cost, lower index on ties
// (inter.c:901-904); -1 for a PU that was not searched.
__global__ void k_select(const me_pu_result_t *res, s8 *best, int n, int nref)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u32 bc = 0xFFFFFFFFu;
    int br = -1;
    for (int r = 0; r < nref; r++) {
        const me_pu_result_t &o = res[(size_t)i * nref + r];
        if (o.valid && o.cost_final < bc) { bc = o.cost_final; br = r; }
    }
    best[i] = (s8)br;
}

typedef void (*subpel_fn)(SubArgs);

// Kernel instance for a shape's (tile type, tile count).
static subpel_fn subpel_kernel(int tt, int T)
{
#define K(TT_, T_) if (tt == TT_ && T == T_) return k_subpel<TT_, T_>
    K(4, 1); K(4, 2); K(4, 4);
    K(2, 1); K(2, 2); K(2, 4);
    K(1, 1); K(1, 4); K(1, 16); K(1, 64); K(1, 256);
#undef K
#define K(TT_, T_) if (tt == TT_ && T == T_) return k_subpel_wide<TT_, T_>
    K(5, 1); K(5, 2); K(5, 4); K(5, 8); K(5, 16); K(5, 32); K(5, 64);
    K(3, 1); K(3, 2); K(3, 4); K(3, 8); K(3, 16); K(3, 32); K(3, 64);
#undef K
    return nullptr;
}

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------

struct me_plan_s {
    me_frame_cfg_t cfg;
    int ctu_cols, ctu_rows, n_ctu, npu, nref;
    int IB, sh, lo, hi, span;
    int mode, waves, band_rows, nbands, wcols, wpitch, dx_cols, nbx;
    size_t lds;
    int base[ME_NSHAPES];
    u32 *d_rate;
    u8 *d_tab_base;
    int tab_off;
    int *d_pu_tab;
    pel *d_cur;
    u64 *d_keys;
    me_pu_result_t *d_res;
    s8 *d_best;
    hipStream_t own;
    hipGraphExec_t exec;
    me_ref_t graph_refs[ME_MAX_REFS];
    u32 graph_mask;
    hipStream_t graph_stream;
};

// Closed form of the MV-bits table of inter_search_create (inter.c:2029-2058).
static void mvbits_fill(u8 *base, int off)
{
    base[off] = 1;
    base[0] = base[2 * off] = 0;
    for (int v = 1; v < off; v++) {
        int l = 0;
        while ((2 << l) <= v + 1) l++;
        base[off + v] = base[off - v] = (u8)(2 * (l + 1));
    }
}

ME_EXPORT int me_plan_create(me_plan_t *out, const me_frame_cfg_t *cfg)
{
    if (me_hip_init(-1)) return -1;
    if (!out || !cfg || cfg->pic_w <= 0 || cfg->pic_h <= 0 || cfg->num_refs < 1 || cfg->num_refs > ME_MAX_REFS ||
        cfg->search_range < 0 || cfg->search_range > 255 || !(cfg->shape_mask & ((1u << ME_NSHAPES) - 1))) {
        me_set_error("me_plan_create: bad configuration", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    me_plan_s *p = (me_plan_s *)calloc(1, sizeof(me_plan_s));
    p->cfg = *cfg;
    p->ctu_cols = (cfg->pic_w + 127) / 128;
    p->ctu_rows = (cfg->pic_h + 127) / 128;
    p->n_ctu = p->ctu_cols * p->ctu_rows;
    p->nref = cfg->num_refs;
    const int R = cfg->search_range;
    p->span = 2 * R + 1;
    int need = 0;
    while ((1 << need) < p->span * p->span) need++;
    const int shift = 16 - (BIT_DEPTH - 8);
    p->IB = need > 32 - shift ? need : 32 - shift;
    p->sh = shift + p->IB - 32;
    // Uniform part of the window: |mv| <= T - 4 with T = max(W,H) (me.c:504-512).
    const int T = cfg->pic_w > cfg->pic_h ? cfg->pic_w : cfg->pic_h;
    p->lo = -R > -T + 4 ? -R : -T + 4;
    p->hi = R < T - 4 ? R : T - 4;
    const uint32_t m = cfg->shape_mask & ((1u << ME_NSHAPES) - 1);
    p->mode = m == (1u << S16x16) ? 1 : 0;
    int n = 0;
    for (int s = 0; s < ME_NSHAPES; s++) {
        if (m >> s & 1) { p->base[s] = n; n += (128 / ME_SHAPES_HOST[s][0]) * (128 / ME_SHAPES_HOST[s][1]); }
        else p->base[s] = -1;
    }
    p->npu = n;
    p->waves = cfg->waves > 0 ? (cfg->waves > 4 ? 4 : cfg->waves) : 4;
    p->band_rows = cfg->band_rows > 0 ? cfg->band_rows : 16;
    if (p->band_rows > p->span) p->band_rows = p->span;
    p->nbands = (p->span + p->band_rows - 1) / p->band_rows;
    // Columns: multiples of 4 displacements (one v_qsad_pk_u16_u8 step).
    const int full_cols = (p->span + 3) & ~3;
    p->dx_cols = cfg->dx_cols > 0 ? ((cfg->dx_cols + 3) & ~3) : full_cols;
    if (p->dx_cols > full_cols) p->dx_cols = full_cols;
    p->nbx = (p->span + p->dx_cols - 1) / p->dx_cols;
    // A lane reads up to 20 samples from column X + (dxb - dxs) <= 112 + dx_cols - 4.
    p->wcols = ((128 + p->dx_cols + 16 + 15) / 16) * 16;
    p->wpitch = p->wcols + 16;
    p->lds = (size_t)(128 + p->band_rows - 1) * p->wpitch * sizeof(pel) + 128 * 128 * sizeof(pel);

    // rate[i] = bits(4*(i-R)) * lambda (GET_MVBITS_IPEL, util.h:43-44)
    p->tab_off = ((cfg->pic_w > cfg->pic_h ? cfg->pic_w : cfg->pic_h) << 2) + 3;
    std::vector<u8> tab(2 * p->tab_off + 1);
    mvbits_fill(tab.data(), p->tab_off);
    std::vector<u32> rate(p->span);
    for (int i = 0; i < p->span; i++) rate[i] = (u32)tab[p->tab_off + 4 * (i - R)] * cfg->lambda_mv;
    std::vector<int> putab(3 * n);
    int k = 0;
    for (int s = 0; s < ME_NSHAPES; s++) {
        if (!(m >> s & 1)) continue;
        int w = ME_SHAPES_HOST[s][0], h = ME_SHAPES_HOST[s][1];
        for (int y = 0; y < 128; y += h)
            for (int x = 0; x < 128; x += w) { putab[3 * k] = s; putab[3 * k + 1] = x; putab[3 * k + 2] = y; k++; }
    }
    const size_t nkeys = (size_t)p->n_ctu * p->npu * p->nref;
    hipError_t e = hipSuccess;
#define ALLOC(ptr, bytes) if (e == hipSuccess) e = hipMalloc((void **)&(ptr), (bytes))
    ALLOC(p->d_rate, p->span * sizeof(u32));
    ALLOC(p->d_tab_base, tab.size());
    ALLOC(p->d_pu_tab, putab.size() * sizeof(int));
    ALLOC(p->d_cur, (size_t)cfg->pic_w * cfg->pic_h * sizeof(pel) + 64);   // slack for dword row reads
    ALLOC(p->d_keys, nkeys * sizeof(u64));
    ALLOC(p->d_res, nkeys * sizeof(me_pu_result_t));
    ALLOC(p->d_best, (size_t)p->n_ctu * p->npu);
#undef ALLOC
    if (e == hipSuccess) e = hipStreamCreate(&p->own);
    if (e == hipSuccess) e = hipMemcpy(p->d_rate, rate.data(), p->span * sizeof(u32), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_tab_base, tab.data(), tab.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_pu_tab, putab.data(), putab.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        me_set_error("me_plan_create", e, __FILE__, __LINE__);
        me_plan_destroy(p);
        return -1;
    }
    *out = p;
    return 0;
}

ME_EXPORT int me_plan_set_cur(me_plan_t p, const me_pel *cur, int stride, int on_device, void *stream)
{
    if (!p || !cur || stride < p->cfg.pic_w) { me_set_error("me_plan_set_cur: bad argument", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    hipStream_t st = stream ? (hipStream_t)stream : p->own;
    ME_CHECK(hipMemcpy2DAsync(p->d_cur, p->cfg.pic_w * sizeof(pel), cur, stride * sizeof(pel), p->cfg.pic_w * sizeof(pel),
                              p->cfg.pic_h, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

// Enqueues the frame; ev (optional, 5 events) brackets the four stages.
static int plan_enqueue_stages(me_plan_s *p, me_ref_t const *refs, uint32_t prepare_mask, hipStream_t st, hipEvent_t *ev)
{
    if (ev) ME_CHECK(hipEventRecord(ev[0], st));
    for (int r = 0; r < p->nref; r++)
        if (prepare_mask >> r & 1)
            if (ref_prepare_async(refs[r], st)) return -1;
    if (ev) ME_CHECK(hipEventRecord(ev[1], st));
    const size_t nkeys = (size_t)p->n_ctu * p->npu * p->nref;
    hipLaunchKernelGGL(k_fill_keys, dim3((unsigned)((nkeys + 255) / 256)), dim3(256), 0, st, p->d_keys, nkeys);
    if (ev) ME_CHECK(hipEventRecord(ev[2], st));

    IntArgs a;
    memset(&a, 0, sizeof(a));
    a.cur = p->d_cur;
    for (int r = 0; r < p->nref; r++) a.ref0[r] = refs[r]->planes + (size_t)ME_PAD * refs[r]->stride + ME_PAD;
    a.ref_stride = refs[0]->stride;
    a.pic_w = p->cfg.pic_w; a.pic_h = p->cfg.pic_h; a.ctu_cols = p->ctu_cols;
    a.R = p->cfg.search_range; a.lo = p->lo; a.hi = p->hi; a.IB = p->IB; a.sh = p->sh;
    a.rate = p->d_rate; a.keys = p->d_keys; a.npu = p->npu; a.nref = p->nref;
    a.band_rows = p->band_rows; a.wcols = p->wcols; a.dx_cols = p->dx_cols; a.nbx = p->nbx; a.wpitch = p->wpitch;
    memcpy(a.base, p->base, sizeof(a.base));
    dim3 grid(p->nbands * p->nbx, p->n_ctu, p->nref);
    if (p->lds > 65536) {
        const void *fn[6] = {(const void *)k_int_search<0, 1>, (const void *)k_int_search<0, 2>, (const void *)k_int_search<0, 4>,
                             (const void *)k_int_search<1, 1>, (const void *)k_int_search<1, 2>, (const void *)k_int_search<1, 4>};
        for (int i = 0; i < 6; i++) ME_CHECK(hipFuncSetAttribute(fn[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds));
    }
    if (p->mode == 0) {
        switch (p->waves) {
        case 1: hipLaunchKernelGGL((k_int_search<0, 1>), grid, dim3(64), p->lds, st, a); break;
        case 2: hipLaunchKernelGGL((k_int_search<0, 2>), grid, dim3(128), p->lds, st, a); break;
        default: hipLaunchKernelGGL((k_int_search<0, 4>), grid, dim3(256), p->lds, st, a); break;
        }
    } else {
        switch (p->waves) {
        case 1: hipLaunchKernelGGL((k_int_search<1, 1>), grid, dim3(64), p->lds, st, a); break;
        case 2: hipLaunchKernelGGL((k_int_search<1, 2>), grid, dim3(128), p->lds, st, a); break;
        default: hipLaunchKernelGGL((k_int_search<1, 4>), grid, dim3(256), p->lds, st, a); break;
        }
    }

    SubArgs b;
    memset(&b, 0, sizeof(b));
    b.cur = p->d_cur;
    for (int r = 0; r < p->nref; r++) b.planes[r] = refs[r]->planes;
    b.plane_elems = refs[0]->plane_elems;
    b.stride = refs[0]->stride;
    b.pic_w = p->cfg.pic_w; b.pic_h = p->cfg.pic_h; b.ctu_cols = p->ctu_cols;
    b.R = p->cfg.search_range; b.IB = p->IB; b.lambda = p->cfg.lambda_mv;
    b.tab = p->d_tab_base + p->tab_off;
    b.keys = p->d_keys; b.res = p->d_res; b.best = p->d_best;
    b.npu = p->npu; b.nref = p->nref; b.n_ctu = p->n_ctu;
    if (ev) ME_CHECK(hipEventRecord(ev[3], st));
    for (int s = 0; s < ME_NSHAPES; s++) {
        if (p->base[s] < 0) continue;
        const int w = ME_SHAPES_HOST[s][0], h = ME_SHAPES_HOST[s][1];
        int tt, tw, th;
        had_tiling(w, h, &tt, &tw, &th);
        const int T = (w / tw) * (h / th);
        subpel_fn fn = subpel_kernel(tt, T);
        if (!fn) { me_set_error("no sub-pel kernel for shape", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
        b.pw = w; b.ph = h; b.tw = tw; b.th = th; b.tpr = w / tw; b.base = p->base[s];
        b.per_ctu = (128 / w) * (128 / h);
        const int lpu = T < 64 ? T : (T < 256 ? 64 : 256), upb = 256 / lpu;
        const long units = (long)p->n_ctu * b.per_ctu * p->nref;
        hipLaunchKernelGGL(fn, dim3((unsigned)((units + upb - 1) / upb)), dim3(256), 0, st, b);
    }
    const int npus = p->n_ctu * p->npu;
    hipLaunchKernelGGL(k_select, dim3((npus + 255) / 256), dim3(256), 0, st, p->d_res, p->d_best, npus, p->nref);
    if (ev) ME_CHECK(hipEventRecord(ev[4], st));
    return 0;
}

static int plan_enqueue(me_plan_s *p, me_ref_t const *refs, uint32_t prepare_mask, hipStream_t st)
{
    return plan_enqueue_stages(p, refs, prepare_mask, st, nullptr);
}

ME_EXPORT int me_plan_run(me_plan_t p, me_ref_t const *refs, uint32_t prepare_mask, void *stream)
{
    if (!p || !refs) { me_set_error("me_plan_run: bad argument", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    for (int r = 0; r < p->nref; r++)
        if (!refs[r] || refs[r]->w != p->cfg.pic_w || refs[r]->h != p->cfg.pic_h) {
            me_set_error("me_plan_run: reference size differs from the plan", hipErrorInvalidValue, __FILE__, __LINE__);
            return -1;
        }
    if (p->lds > 160 * 1024) { me_set_error("me_plan_run: LDS window too large", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    hipStream_t st = stream ? (hipStream_t)stream : p->own;
    bool same = p->exec && p->graph_mask == prepare_mask && p->graph_stream == st;
    for (int r = 0; same && r < p->nref; r++) same = p->graph_refs[r] == refs[r];
    if (!same) {
        if (p->exec) { hipGraphExecDestroy(p->exec); p->exec = nullptr; }
        hipGraph_t g;
        ME_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = plan_enqueue(p, refs, prepare_mask, st);
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) return -1;
        if (e != hipSuccess) { me_set_error("hipStreamEndCapture", e, __FILE__, __LINE__); return -1; }
        e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) { me_set_error("hipGraphInstantiate", e, __FILE__, __LINE__); p->exec = nullptr; return -1; }
        for (int r = 0; r < p->nref; r++) p->graph_refs[r] = refs[r];
        p->graph_mask = prepare_mask;
        p->graph_stream = st;
    }
    ME_CHECK(hipGraphLaunch(p->exec, st));
    return 0;
}

// Stage timing: runs `iters` frames stage by stage (no graph) with HIP
// events on the plan's stream between the stages and returns the average
// milliseconds of: [0] reference preparation (prepare_mask), [1] key fill,
// [2] integer search, [3] sub-pel + reference choice.
ME_EXPORT int me_plan_profile(me_plan_t p, me_ref_t const *refs, uint32_t prepare_mask, int iters, double *ms)
{
    if (!p || !refs || !ms || iters < 1) return -1;
    hipStream_t st = p->own;
    hipEvent_t ev[5];
    for (int i = 0; i < 5; i++) ME_CHECK(hipEventCreate(&ev[i]));
    double acc[4] = {0, 0, 0, 0};
    int rc = 0;
    for (int it = 0; it < iters && !rc; it++) {
        rc = plan_enqueue_stages(p, refs, prepare_mask, st, ev);
        if (!rc) {
            ME_CHECK(hipEventSynchronize(ev[4]));
            for (int k = 0; k < 4; k++) {
                float t = 0;
                ME_CHECK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
                acc[k] += t;
            }
        }
    }
    for (int i = 0; i < 5; i++) hipEventDestroy(ev[i]);
    for (int k = 0; k < 4; k++) ms[k] = acc[k] / iters;
    return rc;
}

ME_EXPORT int me_plan_results(me_plan_t p, me_pu_result_t *res, int8_t *best_ref)
{
    if (!p) return -1;
    hipStream_t st = p->graph_stream ? p->graph_stream : p->own;
    ME_CHECK(hipStreamSynchronize(st));
    const size_t n = (size_t)p->n_ctu * p->npu;
    if (res) ME_CHECK(hipMemcpy(res, p->d_res, n * p->nref * sizeof(me_pu_result_t), hipMemcpyDeviceToHost));
    if (best_ref) ME_CHECK(hipMemcpy(best_ref, p->d_best, n, hipMemcpyDeviceToHost));
    return 0;
}

ME_EXPORT int me_plan_num_ctus(me_plan_t p) { return p ? p->n_ctu : 0; }
ME_EXPORT int me_plan_num_pus(me_plan_t p) { return p ? p->npu : 0; }

ME_EXPORT void me_plan_destroy(me_plan_t p)
{
    if (!p) return;
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->own) hipStreamDestroy(p->own);
    hipFree(p->d_rate); hipFree(p->d_tab_base); hipFree(p->d_pu_tab); hipFree(p->d_cur);
    hipFree(p->d_keys); hipFree(p->d_res); hipFree(p->d_best);
    free(p);
}
