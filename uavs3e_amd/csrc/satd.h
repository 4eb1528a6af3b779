// SATD tiles and the com_had tiling rules, shared by the per-call entries and
// the batched sub-pel kernel.
#pragma once
#include "common.h"

// Sum of |coef| of the 2-D Walsh-Hadamard transform of a TW x TH residual,
// then the per-shape normalisation of cost_satd (com_cost.c:259, 340, 467,
// 586, 649, 708).
template <int TW, int TH>
__device__ __forceinline__ int satd_tile(const pel *org, int s_org, const pel *cur, int s_cur)
{
    int m[TW * TH];
#pragma unroll
    for (int y = 0; y < TH; y++)
#pragma unroll
        for (int x = 0; x < TW; x++) m[y * TW + x] = (int)org[y * s_org + x] - (int)cur[y * s_cur + x];
#pragma unroll
    for (int y = 0; y < TH; y++)
#pragma unroll
        for (int hs = 1; hs < TW; hs <<= 1)
#pragma unroll
            for (int i = 0; i < TW; i += hs << 1)
#pragma unroll
                for (int j = i; j < i + hs; j++) {
                    int p = m[y * TW + j], q = m[y * TW + j + hs];
                    m[y * TW + j] = p + q;
                    m[y * TW + j + hs] = p - q;
                }
#pragma unroll
    for (int x = 0; x < TW; x++)
#pragma unroll
        for (int hs = 1; hs < TH; hs <<= 1)
#pragma unroll
            for (int i = 0; i < TH; i += hs << 1)
#pragma unroll
                for (int j = i; j < i + hs; j++) {
                    int p = m[j * TW + x], q = m[(j + hs) * TW + x];
                    m[j * TW + x] = p + q;
                    m[(j + hs) * TW + x] = p - q;
                }
    int s = 0;
#pragma unroll
    for (int i = 0; i < TW * TH; i++) s += abs(m[i]);
    if (TW == 4 && TH == 4) return (s + 1) >> 1;
    if (TW == 8 && TH == 8) return (s + 2) >> 2;
    if (TW * TH == 128) return me_satd_norm128(s);
    return me_satd_norm32(s);
}

__device__ __forceinline__ int satd_any(int tt, const pel *org, int s_org, const pel *cur, int s_cur)
{
    switch (tt) {
    case 0: return satd_tile<4, 4>(org, s_org, cur, s_cur);
    case 1: return satd_tile<8, 8>(org, s_org, cur, s_cur);
    case 2: return satd_tile<8, 4>(org, s_org, cur, s_cur);
    case 3: return satd_tile<16, 8>(org, s_org, cur, s_cur);
    case 4: return satd_tile<4, 8>(org, s_org, cur, s_cur);
    default: return satd_tile<8, 16>(org, s_org, cur, s_cur);
    }
}

// com_had tiling rules (com_cost.c:719-811) -> tile type, tile w, tile h.
__host__ __device__ inline void had_tiling(int w, int h, int *tt, int *tw, int *th)
{
    if (w == h) {
        if (w == 4) { *tt = 0; *tw = 4; *th = 4; } else { *tt = 1; *tw = 8; *th = 8; }
    } else if (w > h) {
        if (h == 4) { *tt = 2; *tw = 8; *th = 4; } else { *tt = 3; *tw = 16; *th = 8; }
    } else {
        if (w == 4) { *tt = 4; *tw = 4; *th = 8; } else { *tt = 5; *tw = 8; *th = 16; }
    }
}

