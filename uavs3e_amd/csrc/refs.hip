// Device reference pictures for batched ME (gfx950): edge padding and the
// 15 quarter-pel planes of com_if_luma_frame, plus the PU shape API.
#include "frame.h"

#include <atomic>

#include <algorithm>

#include <string.h>

// ---------------------------------------------------------------------------
// Reference pictures: padded integer plane + 15 quarter-pel planes
// ---------------------------------------------------------------------------


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

// N consecutive samples of type T from any sample-aligned address, read as
// dwords and realigned with v_alignbyte (reads at most 3 bytes past the last
// sample: inside the padded planes, whose rows and allocations run further).
template <typename T, int N>
__device__ __forceinline__ void load_samples(const T *p, int *out)
{
    constexpr int NB = N * (int)sizeof(T), NW = (NB + 3) / 4;
    const uintptr_t ad = (uintptr_t)p;
    const u32 *q = (const u32 *)(ad & ~(uintptr_t)3);
    const u32 sh = (u32)(ad & 3);
    u32 d[NW + 1], w[NW];
#pragma unroll
    for (int i = 0; i <= NW; i++) d[i] = q[i];
#pragma unroll
    for (int i = 0; i < NW; i++) w[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
#pragma unroll
    for (int i = 0; i < N; i++) {
        if (sizeof(T) == 1) out[i] = (int)((w[i >> 2] >> (8 * (i & 3))) & 0xFFu);
        else out[i] = (int)(T)((w[i >> 1] >> (16 * (i & 1))) & 0xFFFFu);
    }
}

// 4 consecutive clipped samples at dst: one dword (8 bits) or two (10 bits)
// when dst is aligned to the 4 samples' size, else sample by sample.
__device__ __forceinline__ void store4(pel *dst, const int *v, int maxv, bool aligned)
{
    if (aligned) {
        if (sizeof(pel) == 1) {
            u32 d = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) d |= (u32)me_clip3(0, maxv, v[j]) << (8 * j);
            *(u32 *)dst = d;
        } else {
            uint2 d;
            d.x = (u32)me_clip3(0, maxv, v[0]) | (u32)me_clip3(0, maxv, v[1]) << 16;
            d.y = (u32)me_clip3(0, maxv, v[2]) | (u32)me_clip3(0, maxv, v[3]) << 16;
            *(uint2 *)dst = d;
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) dst[j] = (pel)me_clip3(0, maxv, v[j]);
}

// One padded sample of plane idx at (X, Y) (plane coordinates).
__device__ __forceinline__ int plane_sample(const pel *p0, int stride, int w, int h, const s16 *tmp, int idx, int X,
                                            int Y)
{
    const int dx = idx & 3, dy = idx >> 2;
    const int x = me_clip3(-4, w + 3, X - ME_PAD), y = me_clip3(-4, h + 3, Y - ME_PAD);
    int t = 0;
    if (dy == 0) {
        const pel *s = p0 + (ptrdiff_t)y * stride + x;
#pragma unroll
        for (int k = 0; k < 8; k++) t += (int)s[k - 3] * ME_LUMA_TAPS[dx][k];
        return (t + 32) >> 6;
    }
    if (dx == 0) {
        const pel *s = p0 + (ptrdiff_t)y * stride + x;
#pragma unroll
        for (int k = 0; k < 8; k++) t += (int)s[(ptrdiff_t)(k - 3) * stride] * ME_LUMA_TAPS[dy][k];
        return (t + 32) >> 6;
    }
    const int tw = w + 16, th = h + 16;
    const s16 *s = tmp + (size_t)(dx - 1) * tw * th + (size_t)(y + 8) * tw + (x + 8);
#pragma unroll
    for (int k = 0; k < 8; k++) t += (int)s[(ptrdiff_t)(k - 3) * tw] * ME_LUMA_TAPS[dy][k];
    const int sh2 = 20 - BIT_DEPTH;
    return (t + (1 << (sh2 - 1))) >> sh2;
}

// The 15 sub-pel planes of com_if_luma_frame (com_mc.c:760-800) including
// the final com_img_padding(plane, 1, 4): every padded sample equals the
// filter output at its position clamped into the [-4, w+4) x [-4, h+4) band.
//   dy == 0: clip((H8(P0) + 32) >> 6)              com_mc.c:668-702
//   dx == 0: clip((V8(P0) + 32) >> 6)              com_mc.c:704-729
//   else:    clip((V8(tmp) + 2^(19-bd)) >> (20-bd)) com_mc.c:731-758
// A thread writes 4 consecutive samples of a row; where none of them is
// clamped, the 8-tap inputs are read as whole rows of samples.
// Launched over the band [-4, w+4) x [-4, h+4) only (k_pad_planes copies the
// band's edges outwards); a thread's 4 samples may run past the band's
// right edge, where they get the clamped values the copy would give.
__global__ void k_planes(pel *__restrict__ planes, size_t plane_elems, int stride, int rows, int w, int h,
                         const s16 *__restrict__ tmp)
{
    const int X0 = ME_PAD - 4 + 4 * (blockIdx.x * blockDim.x + threadIdx.x);
    const int Y = ME_PAD - 4 + blockIdx.y;
    const int idx = blockIdx.z + 1;
    if (X0 >= ME_PAD + w + 4 || Y >= rows) return;
    const int dx = idx & 3, dy = idx >> 2;
    const int maxv = (1 << BIT_DEPTH) - 1;
    const pel *p0 = planes + (size_t)ME_PAD * stride + ME_PAD;
    pel *dst = planes + (size_t)idx * plane_elems + (size_t)Y * stride + X0;
    const int x0 = X0 - ME_PAD, y = me_clip3(-4, h + 3, Y - ME_PAD);
    int v[4];
    if (x0 >= -4 && x0 + 3 <= w + 3 && X0 + 3 < stride) {
        int acc[4] = {0, 0, 0, 0};
        if (dy == 0) {
            int s[11];
            load_samples<pel, 11>(p0 + (ptrdiff_t)y * stride + x0 - 3, s);
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int k = 0; k < 8; k++) acc[j] += s[j + k] * ME_LUMA_TAPS[dx][k];
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = (acc[j] + 32) >> 6;
        } else if (dx == 0) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                int s[4];
                load_samples<pel, 4>(p0 + (ptrdiff_t)(y + k - 3) * stride + x0, s);
#pragma unroll
                for (int j = 0; j < 4; j++) acc[j] += s[j] * ME_LUMA_TAPS[dy][k];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = (acc[j] + 32) >> 6;
        } else {
            const int tw = w + 16, th = h + 16;
            const s16 *t0 = tmp + (size_t)(dx - 1) * tw * th + (size_t)(y + 8) * tw + (x0 + 8);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                int s[4];
                load_samples<s16, 4>(t0 + (ptrdiff_t)(k - 3) * tw, s);
#pragma unroll
                for (int j = 0; j < 4; j++) acc[j] += s[j] * ME_LUMA_TAPS[dy][k];
            }
            const int sh2 = 20 - BIT_DEPTH;
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = (acc[j] + (1 << (sh2 - 1))) >> sh2;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = X0 + j < stride ? plane_sample(p0, stride, w, h, tmp, idx, X0 + j, Y) : 0;
    }
    if (X0 + 3 < stride) store4(dst, v, maxv, (stride & 3) == 0 && (plane_elems & 3) == 0);
    else
        for (int j = 0; j < 4; j++)
            if (X0 + j < stride) dst[j] = (pel)me_clip3(0, maxv, v[j]);
}

// com_img_padding(plane, 1, 4) of the 15 sub-pel planes: every sample
// outside the band [-4, w+4) x [-4, h+4) takes the band sample at its
// clamped position.  sides = 0: grid.y over the rows above and below the
// band (whole rows); sides = 1: grid.y over the band's rows (the columns
// left and right of it); 4 samples per thread; grid.z: the plane.  Reads
// stay inside the band, writes outside it.
__global__ void k_pad_planes(pel *__restrict__ planes, size_t plane_elems, int stride, int rows, int w, int h,
                             int sides)
{
    const int bx0 = ME_PAD - 4, bx1 = ME_PAD + w + 3, by0 = ME_PAD - 4, by1 = ME_PAD + h + 3;
    const int ntop = by0, nbot = rows - 1 - by1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int maxv = (1 << BIT_DEPTH) - 1;
    const bool rows_aligned = (stride & 3) == 0 && (plane_elems & 3) == 0;
    pel *pl = planes + (size_t)(blockIdx.z + 1) * plane_elems;
    int Y = blockIdx.y + (sides ? ntop + nbot : 0), X, end;
    const pel *src;
    if (Y < ntop + nbot) {   // 4 samples of a row outside the band
        Y = Y < ntop ? Y : by1 + 1 + (Y - ntop);
        X = 4 * t;
        end = stride;
        src = pl + (size_t)me_clip3(by0, by1, Y) * stride;
    } else {                 // 4 samples left or right of the band in one of its rows
        Y = by0 + (Y - ntop - nbot);
        const int gl = (bx0 + 3) / 4, gr = (stride - 1 - bx1 + 3) / 4;
        if (t >= gl + gr) return;
        X = t < gl ? 4 * t : bx1 + 1 + 4 * (t - gl);
        end = t < gl ? bx0 : stride;
        src = pl + (size_t)Y * stride;
    }
    if (X >= end) return;
    pel *dst = pl + (size_t)Y * stride + X;
    int v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = src[me_clip3(bx0, bx1, X + j)];
    if (X + 3 < end) store4(dst, v, maxv, rows_aligned && (X & 3) == 0);
    else
        for (int j = 0; X + j < end; j++) dst[j] = (pel)v[j];
}

int me_ref_prepare_async(me_ref_s *r, hipStream_t st)
{
    dim3 b(256);
    hipLaunchKernelGGL(k_pad_int, dim3((r->stride + 255) / 256, r->rows), b, 0, st, r->luma, r->w, r->h,
                       r->planes, r->stride, r->rows);
    hipLaunchKernelGGL(k_htmp, dim3((r->w + 16 + 255) / 256, r->h + 16, 3), b, 0, st,
                       r->planes + (size_t)ME_PAD * r->stride + ME_PAD, r->stride, r->w, r->h, r->tmp);
    hipLaunchKernelGGL(k_planes, dim3((r->w + 8 + 1023) / 1024, r->h + 8, 15), b, 0, st, r->planes, r->plane_elems,
                       r->stride, r->rows, r->w, r->h, r->tmp);
    const int pad_rows = r->rows - (r->h + 8);
    const int side_threads = (ME_PAD - 4 + 3) / 4 + (r->stride - (ME_PAD + r->w + 4) + 3) / 4;
    hipLaunchKernelGGL(k_pad_planes, dim3((r->stride + 3) / 4 / 256 + 1, pad_rows, 15), b, 0, st, r->planes,
                       r->plane_elems, r->stride, r->rows, r->w, r->h, 0);
    hipLaunchKernelGGL(k_pad_planes, dim3((side_threads + 255) / 256, r->h + 8, 15), b, 0, st, r->planes,
                       r->plane_elems, r->stride, r->rows, r->w, r->h, 1);
    ME_CHECK(hipGetLastError());
    return 0;
}

ME_EXPORT int me_ref_create(me_ref_t *out, int pic_w, int pic_h)
{
    if (me_hip_init(-1)) return -1;
    if (!out || pic_w <= 0 || pic_h <= 0) { me_set_error("me_ref_create: bad size", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    static std::atomic<u64> next_id{1};
    me_ref_s *r = (me_ref_s *)calloc(1, sizeof(me_ref_s));
    r->id = next_id.fetch_add(1);
    r->w = pic_w; r->h = pic_h;
    r->stride = pic_w + 2 * ME_PAD;
    r->rows = pic_h + 2 * ME_PAD;
    r->plane_elems = (size_t)r->stride * r->rows;
    hipError_t e = hipMalloc((void **)&r->planes, 16 * r->plane_elems * sizeof(pel));
    // +64 B: the unpadded luma doubles as a current picture, read with dword row loads
    if (e == hipSuccess) e = hipMalloc((void **)&r->luma, (size_t)pic_w * pic_h * sizeof(pel) + 64);
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
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (stride == r->w)   // one copy: a 2-D copy from pageable memory goes row by row
        ME_CHECK(hipMemcpyAsync(r->luma, luma, (size_t)r->w * r->h * sizeof(pel), kind, st));
    else
        ME_CHECK(hipMemcpy2DAsync(r->luma, r->w * sizeof(pel), luma, stride * sizeof(pel), r->w * sizeof(pel), r->h,
                                  kind, st));
    if (me_ref_prepare_async(r, st)) return -1;
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

ME_EXPORT int me_ref_load_chroma(me_ref_t r, const me_pel *u, const me_pel *v, int stride, int on_device, void *stream)
{
    const int cw = r ? r->w / 2 : 0, chh = r ? r->h / 2 : 0;
    if (!r || !u || !v || stride < cw || cw < 1 || chh < 1) {
        me_set_error("me_ref_load_chroma: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (!r->chroma) ME_CHECK(hipMalloc((void **)&r->chroma, 2 * (size_t)cw * chh * sizeof(pel)));
    hipStream_t st = (hipStream_t)stream;
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const me_pel *src[2] = {u, v};
    for (int c = 0; c < 2; c++) {
        pel *dst = r->chroma + (size_t)c * cw * chh;
        if (stride == cw) ME_CHECK(hipMemcpyAsync(dst, src[c], (size_t)cw * chh * sizeof(pel), kind, st));
        else ME_CHECK(hipMemcpy2DAsync(dst, cw * sizeof(pel), src[c], stride * sizeof(pel), cw * sizeof(pel), chh, kind, st));
    }
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
    (void)hipFree(r->planes);
    (void)hipFree(r->luma);
    (void)hipFree(r->tmp);
    (void)hipFree(r->chroma);
    (void)hipFree(r->aff_pred);
    (void)hipFree(r->aff_cus);
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
