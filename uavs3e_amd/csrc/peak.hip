// Issue-rate microbenchmarks of the VALU instructions the search kernels are
// built from, on gfx950 (measurement support for bench.py; not part of the
// ME library's C ABI).
//
// Every lane runs 8 independent dependency chains of one instruction kind,
// 16 instructions per loop iteration, so that the issue rate and not the
// latency of a chain is measured; the grid holds 8 waves per SIMD.  The
// instructions are written as inline asm so that exactly that instruction
// is issued, and every chain's result is stored at the end.
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PEAK_EXPORT extern "C" __attribute__((visibility("default")))

typedef uint32_t u32;
typedef uint64_t u64;

enum { K_ADD_U32 = 0, K_SAD_U8, K_QSAD_PK, K_SAD_U16, K_MIN_F64, K_ADD_DPP, K_LSHL_ADD_U64, K_ADD_SDWA, K_NKINDS };

template <int KIND>
__global__ __launch_bounds__(256) void k_peak(u32 seed, int iters, u32 *out)
{
    const u32 t = blockIdx.x * 256 + threadIdx.x;
    u32 a[8];
    u64 q[8];
    double d[8];
    const u32 b = seed ^ (t * 2654435761u), c = b * 40503u + 1u;
    const u64 qb = ((u64)c << 32) | b;
    const double db = __builtin_bit_cast(double, (u64)0x4000000000000000ull | (qb & 0xFFFFFFFFFFFull));
#pragma unroll
    for (int j = 0; j < 8; j++) {
        a[j] = b + j;
        q[j] = qb + j;
        d[j] = __builtin_bit_cast(double, (u64)0x3FF0000000000000ull | ((qb + j) & 0xFFFFFFFFFFFull));
    }
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int k = j & 7;
            if (KIND == K_ADD_U32) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
            if (KIND == K_SAD_U8) asm volatile("v_sad_u8 %0, %1, %2, %0" : "+v"(a[k]) : "v"(b), "v"(c));
            if (KIND == K_QSAD_PK) asm volatile("v_qsad_pk_u16_u8 %0, %1, %2, %0" : "+v"(q[k]) : "v"(qb), "v"(c));
            if (KIND == K_SAD_U16) asm volatile("v_sad_u16 %0, %1, %2, %0" : "+v"(a[k]) : "v"(b), "v"(c));
            if (KIND == K_MIN_F64) asm volatile("v_min_f64 %0, %0, %1" : "+v"(d[k]) : "v"(db));
            if (KIND == K_ADD_DPP)
                asm volatile("v_add_u32_dpp %0, %1, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1"
                             : "+v"(a[k]) : "v"(b));
            if (KIND == K_LSHL_ADD_U64) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(q[k]) : "v"(qb));
            if (KIND == K_ADD_SDWA)
                asm volatile("v_add_u32_sdwa %0, %1, %0 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_1 src1_sel:DWORD"
                             : "+v"(a[k]) : "v"(b));
        }
    }
    u32 r = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) r ^= a[j] ^ (u32)q[j] ^ (u32)(q[j] >> 32) ^ (u32)__builtin_bit_cast(u64, d[j]);
    out[t] = r;
}

// Runs kind `kind` (0 v_add_u32, 1 v_sad_u8, 2 v_qsad_pk_u16_u8, 3
// v_sad_u16, 4 v_min_f64, 5 v_add_u32 with a DPP operand, 6 v_lshl_add_u64,
// 7 v_add_u32 of a 16-bit half (SDWA)) on `blocks`
// workgroups of 256 lanes, iters x 16 instructions per lane, and returns the
// kernel's time (HIP events) and the lane-instructions it issued.
PEAK_EXPORT int me_peak_run(int kind, int iters, int blocks, double *ms, double *lane_instrs)
{
    if (kind < 0 || kind >= K_NKINDS || iters < 1 || blocks < 1 || !ms || !lane_instrs) return -1;
    u32 *out = nullptr;
    if (hipMalloc((void **)&out, (size_t)blocks * 256 * sizeof(u32)) != hipSuccess) return -1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
    void (*fn)(u32, int, u32 *) = nullptr;
    switch (kind) {
    case K_ADD_U32: fn = k_peak<K_ADD_U32>; break;
    case K_SAD_U8: fn = k_peak<K_SAD_U8>; break;
    case K_QSAD_PK: fn = k_peak<K_QSAD_PK>; break;
    case K_SAD_U16: fn = k_peak<K_SAD_U16>; break;
    case K_MIN_F64: fn = k_peak<K_MIN_F64>; break;
    case K_ADD_DPP: fn = k_peak<K_ADD_DPP>; break;
    case K_LSHL_ADD_U64: fn = k_peak<K_LSHL_ADD_U64>; break;
    default: fn = k_peak<K_ADD_SDWA>; break;
    }
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(256), 0, 0, 7u, 4, out);   // warm-up
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(256), 0, 0, 7u, iters, out);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return -1;
    *ms = t;
    *lane_instrs = (double)blocks * 256.0 * iters * 16.0;
    return 0;
}
