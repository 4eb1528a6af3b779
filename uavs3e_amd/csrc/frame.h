// Internal declarations shared by the batched ME translation units
// (refs.hip, intsearch.hip, subpel.hip, frame.hip, server.hip).
#pragma once
#include "common.h"
#include "satd.h"
#include "../../include/uavs3e_me_hip.h"

#define ME_MAX_REFS 4
// Entries past the end of the integer-pel rate row (see IntArgs::rate).
#define ME_RATE_PAD 8

// A reference picture resident in device memory: the edge-padded integer
// plane + 15 quarter-pel planes (com_pic_t / com_subpel_t, com_util.c:62-91,
// 251-284), plus the sources from which they are rebuilt.
struct me_ref_s {
    u64 id;                  // unique per me_ref_create in this process (never reused)
    int w, h, stride, rows;
    size_t plane_elems;      // stride * rows
    pel *planes;             // 16 planes, plane i at planes + i * plane_elems
    pel *luma;               // unpadded w x h copy (source for re-preparation)
    s16 *tmp;                // 3 horizontal s16 planes over [-8, w+8) x [-8, h+8)
    pel *chroma;             // optional 4:2:0 u then v, (w/2) x (h/2) each, unpadded
    // affine motion compensation buffers (grown on demand)
    pel *aff_pred;
    size_t aff_pred_cap;
    me_affine_cu_t *aff_cus;
    int aff_cus_cap;
};

// Pads the integer plane of r from r->luma and rebuilds its 15 sub-pel
// planes on stream st (com_img_padding + com_if_luma_frame).
int me_ref_prepare_async(me_ref_s *r, hipStream_t st);

enum {
    S4x8 = 0, S8x4, S4x16, S16x4, S4x32, S32x4, S8x8, S8x16, S16x8, S8x32, S32x8, S8x64, S64x8, S16x16,
    S16x32, S32x16, S16x64, S64x16, S16x128, S128x16, S32x32, S32x64, S64x32, S32x128, S128x32, S64x64,
    S64x128, S128x64, S128x128
};

// One current picture and the references its PUs are searched against.
struct PicDesc {
    const pel *cur;                  // current picture, stride pic_w (+64 B of read slack)
    const pel *planes[ME_MAX_REFS];  // plane-0 base (padded buffer start) of each reference
    size_t plane_elems;              // samples per plane (stride * rows)
    int ref_stride;                  // pic_w + 2 * ME_PAD
    int pic_w, pic_h, ctu_cols;
    const pel *chroma[ME_MAX_REFS];  // u plane of each reference (v follows), or null
};

// The CTUs a launch works on, always indexed by uniform (workgroup-level)
// values so that descriptor reads are scalar loads.  jobs == nullptr: job j
// is CTU j of pics[0] (a whole frame).  Otherwise jobs[j] = (picture index
// << 16) | CTU index.  pics is device memory.  Keys and results are indexed
// by job.
struct JobSet {
    const PicDesc *pics;
    const u32 *jobs;
    int n_jobs;
};

__device__ __forceinline__ const PicDesc *job_decode(const JobSet &js, int job, int &ctu)
{
    if (js.jobs) {
        const u32 j = js.jobs[job];
        ctu = (int)(j & 0xFFFFu);
        return js.pics + (j >> 16);
    }
    ctu = job;
    return js.pics;
}

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

// How the integer search's work is spread over workgroups and where each
// workgroup leaves its keys.  A unit is (job, reference, column tile bx,
// band of band_rows displacement rows), numbered band fastest:
//   u = ((job * nref + ref) * nbx + bx) * nbands + band.
// Workgroup g of `groups` takes the contiguous units
// [units * g / groups, units * (g + 1) / groups).  The units of one
// (job, reference, bx) -- a segment -- are therefore split into pieces over
// consecutive workgroups, and a workgroup keeps a segment's node states in
// registers across its bands.
//   nslots == 0: every unit is a workgroup of its own (groups == units) and
//     publishes with a 64-bit atomicMin into one key array (filled with ones
//     first).
//   nslots > 0 (persistent): piece p of a segment stores its keys, with
//     plain stores, into slot bx * nslots + p, each slot a [job][npu][nref]
//     array of slot_elems keys; the piece count of a segment follows from
//     the partition (key_pieces), so a reader takes the minimum over exactly
//     the slots written and nothing needs filling.
// A segment with no candidate in a piece stores KEY_INIT there; a key
// >= KEY_INIT after the merge means "no candidate".
struct KeyLayout {
    u32 units, groups;
    int nbands, nbx, nslots;
    size_t slot_elems;
};

// The workgroup whose range holds unit u.
__host__ __device__ __forceinline__ u32 key_owner(const KeyLayout &k, u64 u)
{
    return (u32)(((u + 1) * k.groups - 1) / k.units);
}

// Pieces of segment x (x = (job * nref + ref) * nbx + bx).
__host__ __device__ __forceinline__ int key_pieces(const KeyLayout &k, u64 x)
{
    const u64 u0 = x * (u64)k.nbands;
    return (int)(key_owner(k, u0 + k.nbands - 1) - key_owner(k, u0)) + 1;
}

// The merged integer-search key of (job, PU, reference); idx is its index in
// a [job][npu][nref] array.
__device__ __forceinline__ u64 key_merged(const KeyLayout &k, const u64 *keys, size_t idx, int job, int ref, int nref)
{
    if (k.nslots == 0) return keys[idx];
    u64 m = ~0ull;
    for (int bx = 0; bx < k.nbx; bx++) {
        const int np = key_pieces(k, ((u64)job * nref + ref) * k.nbx + bx);
        for (int p = 0; p < np; p++) {
            const u64 v = keys[(size_t)(bx * k.nslots + p) * k.slot_elems + idx];
            m = v < m ? v : m;
        }
    }
    return m;
}

struct IntArgs {
    JobSet js;
    int R;                   // window half size
    int lo, hi;              // uniform displacement clip: [max(-R,-T+4), min(R,T-4)]
    int IB, sh;              // index bits; sad shift into the high word
    const u32 *rate;         // [2R+1 + ME_RATE_PAD]: bits(4*(i-R)) * lambda, then padding
    u64 *keys;               // see KeyLayout
    KeyLayout kl;
    int npu, nref;
    int band_rows, wcols;    // displacement rows per band; LDS window row length
    int dx_cols;             // displacement columns per column tile
    int wpitch;              // LDS row pitch of the window (wcols + skew room)
    int base[ME_NSHAPES];    // first PU index of each shape inside a CTU (-1: not selected)
};

// Launches k_int_search<mode, waves> (waves 1, 2, 4 or 8; mode 1 = 16x16 PUs
// only) on a.kl.groups workgroups with `lds` bytes of LDS.
int me_int_search_launch(int mode, int waves, const IntArgs &a, size_t lds, hipStream_t st);
int me_int_search_launch_w1(int mode, const IntArgs &a, size_t lds, hipStream_t st);
int me_int_search_launch_w2(int mode, const IntArgs &a, size_t lds, hipStream_t st);
int me_int_search_launch_w4(int mode, const IntArgs &a, size_t lds, hipStream_t st);
int me_int_search_launch_w8(int mode, const IntArgs &a, size_t lds, hipStream_t st);

// ---------------------------------------------------------------------------
// Sub-pel refinement (me_sub_pel_search, me.c:366-500, cost_type 0, bi 0)
// ---------------------------------------------------------------------------

// One PU shape of the sub-pel launch: its com_had tiling, the lanes per PU
// and the first workgroup (per job and reference) of its range.
struct SubShape {
    int pw, ph, tw, th, tpr;    // PU and tile size, tiles per row
    int base, per_ctu;          // PU index base in the CTU, PUs per CTU
    int lpu;                    // lanes per PU position group: tiles, 2 x tiles for 16x8 / 8x16 tiles
    int G;                      // position groups per PU (lpu * G <= 256)
    int tt;                     // com_had tile type (had_tiling)
    int blk0;                   // first workgroup of the shape
    int pad;
};

struct SubArgs {
    JobSet js;
    int n0;                     // list-0 size (references [n0, nref) form list 1)
    int R, IB;
    u32 lambda;
    const u8 *tab;              // MV-bits table, centred (tab[0] is mv 0)
    const u64 *keys;            // integer-search keys (KeyLayout kl)
    KeyLayout kl;
    me_pu_result_t *res;
    int npu, nref;
    const SubShape *shp;        // device table of the searched shapes
    const u8 *blk_shape;        // workgroup -> index into shp
    int mx;                     // 8 bits: the matrix-core kernel (its shape layout: lpu = tiles per PU, G = 32-tile batches per wave)
};

// Every shape in one launch: grid (blocks, jobs, refs) of 256 lanes.
int me_subpel_launch(const SubArgs &a, int blocks, int jobs, int refs, hipStream_t st);
// Empty launches on st (no jobs; the kernels return at once): a kernel's
// first launch in a process, and its first on a queue when it has private
// memory, otherwise stall for milliseconds.
int me_subpel_preload(hipStream_t st);
int me_int_search_preload_w1(int mode, size_t lds, hipStream_t st);
int me_int_search_preload_w2(int mode, size_t lds, hipStream_t st);
int me_int_search_preload_w4(int mode, size_t lds, hipStream_t st);
int me_int_search_preload_w8(int mode, size_t lds, hipStream_t st);

// ---------------------------------------------------------------------------
// One search configuration: window geometry, key layout, schedule and the
// device rate tables, shared by frame plans and the request server.
// ---------------------------------------------------------------------------

struct MeSearch {
    int R, span, IB, sh, lo, hi;
    int npu, nref, nref_l1, nl, mode;   // nl: lists (1, or 2 when nref_l1 > 0)
    int waves, band_rows, nbands, wcols, wpitch, dx_cols, nbx;
    size_t lds;
    int base[ME_NSHAPES];
    u32 lambda;
    u32 *d_rate;             // [2R+1] integer-pel rate per displacement component
    u8 *d_tab_base;          // MV-bits table, tab_off samples each side of 0
    int tab_off;
    SubShape *d_shp;         // sub-pel shape table of the searched shapes
    u8 *d_blk_shape;         // sub-pel workgroup -> shape
    int sub_blocks;          // sub-pel workgroups per (job, reference)
    int sub_mx;              // the sub-pel launch runs k_subpel_mx (8 bits, one position at a time)
    // me_search_tz runs one launch per com_had tile type; the launches write
    // disjoint results, so they run side by side on these streams (default
    // priority), forked from the caller's stream by tz_ev[6] and joined back
    // by tz_ev[0..5]
    hipStream_t tz_side[6];
    hipEvent_t tz_ev[7];
};

// max_dim = max(pic_w, pic_h) of the pictures searched (the |mv| <= T - 4
// limit and the MV-bits table size of inter_search_create).  band_rows,
// waves and dx_cols are schedule choices (0 = default); sp_groups is the
// number of sub-pel search positions a PU evaluates at the same time (1, 2,
// 4 or 8: 1 does the least work, more cut the latency of small batches).
// Results do not depend on them.
int me_search_init(MeSearch *s, int R, u32 lambda, u32 shape_mask, int nref, int nref_l1, int max_dim, int band_rows,
                   int waves, int dx_cols, int sp_groups);
void me_search_free(MeSearch *s);
// Makes st ready for the search's launches without delay: loads the
// kernels' code and creates the stream's device queue (first launches on a
// new stream otherwise pay for both), then waits for st.
int me_search_warmup(const MeSearch *s, hipStream_t st);
// The integer search's partition and key slots for n_jobs jobs.  grid 0:
// with enough units to fill the device twice over, as many workgroups as the
// device holds at once (persistent); grid n >= 2: n persistent workgroups
// when that leaves some with several units; otherwise, and for grid 1, one
// workgroup per unit (atomic keys).  A persistent partition whose segments
// would take more than 4 pieces also falls back to atomic keys.  Returns the
// keys the buffer must hold.
size_t me_search_keys(const MeSearch *s, int n_jobs, int grid, KeyLayout *kl);
// Enqueues key initialisation, integer search, sub-pel refinement and the
// reference choice for the jobs of js.  keys hold kl (me_search_keys for
// js.n_jobs); res is [job][npu][nref], best [job][npu][nl].  ev, if given,
// receives 4 events recorded before the key fill, before the integer search,
// before sub-pel and at the end.
// The same in three parts: key fill (atomic layouts only), integer search,
// sub-pel + choice.
int me_search_fill(const MeSearch *s, const JobSet &js, const KeyLayout &kl, u64 *keys, hipStream_t st);
int me_search_int(const MeSearch *s, const JobSet &js, const KeyLayout &kl, u64 *keys, hipStream_t st);
int me_search_finish(const MeSearch *s, const JobSet &js, const KeyLayout &kl, const u64 *keys, me_pu_result_t *res,
                     s8 *best, hipStream_t st);
int me_search_enqueue(const MeSearch *s, const JobSet &js, const KeyLayout &kl, u64 *keys, me_pu_result_t *res,
                      s8 *best, hipStream_t st, hipEvent_t *ev);
// The MV-bits table of inter_search_create (inter.c:2029-2058) in closed
// form: tab[off + v] for |v| < off, 0 at +-off.
void me_mvbits_fill(u8 *base, int off);

// ---------------------------------------------------------------------------
// Motion compensation of every PU from a search's results
// ---------------------------------------------------------------------------

// Enqueues the predictions (layout of me_plan_mc) for the jobs of js from
// res / best of me_search_enqueue; the pictures' chroma pointers must be set
// when pred_u / pred_v are given.
int me_mc_enqueue(const MeSearch *s, const JobSet &js, const me_pu_result_t *res, const s8 *best, pel *pred_y,
                  pel *pred_u, pel *pred_v, hipStream_t st);
// Elements per CTU of the luma and of each chroma prediction buffer.
void me_mc_sizes(const MeSearch *s, size_t *luma_per_ctu, size_t *chroma_per_ctu);

// ---------------------------------------------------------------------------
// me_search_tz of every PU from per-PU MVPs
// ---------------------------------------------------------------------------

// Enqueues the reference's search (me.c:529-568) for the jobs of js: mvp
// (null: all zero) and mvi_prev (optional) are [job][npu][nref][2] device
// arrays; results go to res [job][npu][nref].  only_shape >= 0 searches that
// shape alone; pad_win is inter_search_t.is_padding; lambda 0 takes the
// search's own.
struct BiState;
struct BiBufs {
    me_bi_pu_t *pus;         // [job][npu]
    BiState *state;          // [job][npu]
    s16 *mv_scale;           // [job][npu][nref][2]
    pel *orgbi;              // [job][selected shape][16384]
};
int me_tz_enqueue(const MeSearch *s, const JobSet &js, const int16_t *mvp, const int16_t *mvi_prev,
                  const me_tz_cfg_t *tz, me_pu_result_t *res, hipStream_t st, int only_shape = -1, int pad_win = 1,
                  u32 lambda = 0, const BiBufs *bi = nullptr);
// analyze_bi's search (two rounds of target, search and choice) for the jobs
// of js; bi.pus holds the inputs and receives the results.
int me_bi_enqueue(const MeSearch *s, const JobSet &js, const int16_t *mvp, const me_tz_cfg_t *tz,
                  me_pu_result_t *res, const BiBufs &bi, hipStream_t st);
size_t me_bi_state_bytes(size_t n);
// analyze_smvd's search for the jobs of js; out.active selects PUs.
int me_smvd_enqueue(const MeSearch *s, const JobSet &js, const int16_t *mvp, const int16_t *mv_scale,
                    const me_tz_cfg_t *tz, me_smvd_pu_t *out, hipStream_t st);
// The lookahead's inter cost of every whole 16x16 unit of the picture, in
// raster order, from the 16x16 results of me_tz_enqueue (device out).
int me_loka_enqueue(const MeSearch *s, const PicDesc *pic, const me_pu_result_t *res, int pic_w, int pic_h, u32 *out,
                    hipStream_t st);
