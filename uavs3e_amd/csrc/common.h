// Shared definitions for the uavs3e ME HIP library (gfx950).
// The library is compiled once per sample bit depth, like the reference
// (inc/com_api.h:46-52): BIT_DEPTH=8 gives pel = uint8_t, 10 gives uint16_t.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#ifndef BIT_DEPTH
#define BIT_DEPTH 8
#endif

#if BIT_DEPTH == 8
typedef uint8_t pel;
#else
typedef uint16_t pel;
#endif
typedef int8_t s8;
typedef int16_t s16;
typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t s64;

#define ME_EXPORT extern "C" __attribute__((visibility("default")))

#define ME_PAD 144          // PIC_PAD_SIZE_L (com_define.h:293)
#define ME_MAX_CU 128       // MAX_CU_SIZE (com_define.h:268)
#define ME_NSHAPES 29

// Error reporting: every HIP call is checked; the library records the last
// error string and returns a negative code from its C entry points.
void me_set_error(const char *what, hipError_t e, const char *file, int line);
// The device selected by me_hip_init (threads the library starts select it).
int me_hip_device(void);
#define ME_CHECK(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            me_set_error(#call, e_, __FILE__, __LINE__);                      \
            return -1;                                                        \
        }                                                                     \
    } while (0)

// The 29 PU shapes (w, h) of the frame-level API, in their fixed order.
static const int ME_SHAPES_HOST[ME_NSHAPES][2] = {
    {4, 8}, {8, 4}, {4, 16}, {16, 4}, {4, 32}, {32, 4}, {8, 8}, {8, 16}, {16, 8}, {8, 32},
    {32, 8}, {8, 64}, {64, 8}, {16, 16}, {16, 32}, {32, 16}, {16, 64}, {64, 16}, {16, 128},
    {128, 16}, {32, 32}, {32, 64}, {64, 32}, {32, 128}, {128, 32}, {64, 64}, {64, 128},
    {128, 64}, {128, 128}};

// Quarter-pel luma taps (com_tables.c:149-154) and 1/8-pel chroma taps
// (com_tables.c:191-200).
__constant__ static const s8 ME_LUMA_TAPS[4][8] = {
    {0, 0, 0, 64, 0, 0, 0, 0},
    {-1, 4, -10, 57, 19, -7, 3, -1},
    {-1, 4, -11, 40, 40, -11, 4, -1},
    {-1, 3, -7, 19, 57, -10, 4, -1}};

__constant__ static const s8 ME_CHROMA_TAPS[8][4] = {
    {0, 64, 0, 0}, {-4, 62, 6, 0}, {-6, 56, 15, -1}, {-5, 47, 25, -3},
    {-4, 36, 36, -4}, {-3, 25, 47, -5}, {-1, 15, 56, -6}, {0, 6, 62, -4}};

// The normalisation of the 8x4 / 4x8 (32 coefficients) and 16x8 / 8x16 (128
// coefficients) cost_satd tiles: (int)(s / com_tbl_sqrt[i] * 2.0) in IEEE
// double (com_cost.c:467, 586, 649, 708; com_tbl_sqrt = {sqrt(32.0),
// sqrt(128.0)}, com_tables.c:276-277).  sqrt(32) = 4 sqrt(2) and sqrt(128) =
// 8 sqrt(2) are the correctly rounded sqrt(2) times a power of two.  Every
// kernel normalises through these two functions.  They multiply by the
// double nearest 2 / com_tbl_sqrt[i] instead of dividing: for every s in
// [0, 2^27) -- a tile's sum of |coef| stays below 2^24 at 10 bits -- both
// give the same integer (checked for every such s on the host, and again in
// the CPU tests over [0, 2^24]), and the multiply is one instruction where
// the division is a dozen.
__host__ __device__ __forceinline__ int me_satd_norm32(int s) { return (int)((double)s * 0x1.6a09e667f3bccp-2); }
__host__ __device__ __forceinline__ int me_satd_norm128(int s) { return (int)((double)s * 0x1.6a09e667f3bccp-3); }

__device__ __forceinline__ int me_clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }

// Workgroup order for L2 locality.  A launch's workgroups are dealt to the
// GPU's 8 XCDs (each with its own L2) in turn by linear index, so
// consecutive workgroups -- which here work on the same CTU and reference
// and read the same planes -- would fill 8 L2s with the same lines.  This
// bijection of [0, n) gives the workgroups of XCD x (linear index = x mod 8)
// one contiguous run of work indices instead.  Results never depend on it.
#define ME_XCDS 8
__device__ __forceinline__ unsigned me_xcd_order(unsigned lin, unsigned n)
{
    const unsigned q = n / ME_XCDS, rem = n % ME_XCDS, x = lin % ME_XCDS, slot = lin / ME_XCDS;
    return x < rem ? x * (q + 1) + slot : rem * (q + 1) + (x - rem) * q + slot;
}

// One wave64 min of two u64 keys held as the bit patterns of positive normal
// doubles: for such values the IEEE order equals the unsigned integer order,
// and v_min_f64 does it in one instruction (the compiler's fmin would first
// canonicalise both operands).
__device__ __forceinline__ double me_vmin64(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
