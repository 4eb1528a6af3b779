/*
 * uavs3e_me_hip.h -- C ABI of the MI355X (gfx950) motion-estimation library.
 *
 * The library is built once per sample bit depth, exactly as uavs3e fixes
 * `pel` at compile time (inc/com_api.h:46-52):
 *   libuavs3e_me_hip8.so  (BIT_DEPTH 8,  me_pel = uint8_t)
 *   libuavs3e_me_hip10.so (BIT_DEPTH 10, me_pel = uint16_t)
 * Include this header with the same BIT_DEPTH as the library you link.
 *
 * Two layers:
 *   1. Per-call drop-in entries.  me_funs_init_hip() overwrites the hot-path
 *      slots of a dispatch table laid out like uavs3e's funs_handle_t
 *      (inc/com_util.h:278-341), the way uavs3e_funs_init_avx2() does
 *      (src/avx2/avx2.c:41).  Each entry takes host pointers and strides in
 *      pels, runs on the GPU and returns only when its outputs are written,
 *      like the reference's C/SIMD kernels.
 *   2. Batched frame ME (no counterpart in the reference): device-resident
 *      reference pictures with their 15 quarter-pel planes, and a per-frame
 *      plan that runs integer full search + sub-pel refinement + reference
 *      selection for every PU of a frame in one hipGraph launch.
 *
 * All functions returning int give 0 on success and a negative value on
 * failure; me_hip_last_error() describes the failure.
 */
#ifndef UAVS3E_ME_HIP_H
#define UAVS3E_ME_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef BIT_DEPTH
#define BIT_DEPTH 8
#endif
#if BIT_DEPTH == 8
typedef uint8_t me_pel;
#else
typedef uint16_t me_pel;
#endif

/* ------------------------------------------------------------------------ */
/* Library state                                                            */
/* ------------------------------------------------------------------------ */

/* Bit depth the library was built for (8 or 10). */
int me_hip_bit_depth(void);

/* Selects HIP device `device` for this process and checks that it can run
 * the library's code objects.  Called implicitly by every entry point. */
int me_hip_init(int device);

/* Description of the most recent failure in this thread ("" if none). */
const char *me_hip_last_error(void);

/* Number of per-call entries that failed since load. */
uint64_t me_hip_error_count(void);

/* The per-call entries have no error channel (the reference slots return
 * void/u32, com_util.h:300-330).  By default a failing entry prints the HIP
 * error and aborts the process.  With enable=1 it instead calls the slot that
 * me_funs_init_hip() replaced and counts the failure. */
void me_hip_set_fallback(int enable);

/* ------------------------------------------------------------------------ */
/* 1. Per-call drop-in: the hot-path slots of funs_handle_t                 */
/* ------------------------------------------------------------------------ */

/* Same slots, signatures and index conventions as funs_handle_t:
 *   widx = log2(width) - 2 (com_tables.c:75-105), CU_SIZE_NUM = 6;
 *   ipflt[IPFILTER_H_8, H_4, V_8, V_4], ipflt_ext[IPFILTER_EXT_8, EXT_4]
 *   (com_define.h:251-262); cost_satd[log2w-2][log2h-2] with the six slots
 *   [0][0] [1][0] [0][1] [1][1] [2][1] [1][2] filled (com_cost.c:1025-1030). */
typedef struct me_dispatch_t {
    /* com_util.h:300 */
    void (*ipcpy[6])(const me_pel *src, int i_src, me_pel *dst, int i_dst, int width, int height);
    /* com_util.h:301 */
    void (*ipflt[4][6])(const me_pel *src, int i_src, me_pel *dst, int i_dst, int width, int height,
                        const int8_t *coeff, int max_val);
    /* com_util.h:302 */
    void (*ipflt_ext[2][6])(const me_pel *src, int i_src, me_pel *dst, int i_dst, int width, int height,
                            const int8_t *coef_x, const int8_t *coef_y, int max_val);
    /* com_util.h:304 */
    void (*ip_flt_y_hor)(const me_pel *src, int i_src, me_pel *dst[3], int i_dst, int16_t *dst_tmp[3],
                         int i_dst_tmp, int width, int height, int8_t (*coeff)[8], int bit_depth);
    /* com_util.h:305 */
    void (*ip_flt_y_ver)(const me_pel *src, int i_src, me_pel *dst[3], int i_dst, int width, int height,
                         int8_t (*coeff)[8], int bit_depth);
    /* com_util.h:306 */
    void (*ip_flt_y_ver_ext)(const int16_t *src, int i_src, me_pel *dst[3], int i_dst, int width, int height,
                             int8_t (*coeff)[8], int bit_depth);
    /* com_util.h:316 */
    uint32_t (*cost_sad[6])(me_pel *p_org, int i_org, me_pel *p_pred, int i_pred, int height);
    /* com_util.h:317 */
    void (*cost_sad_x3[6])(me_pel *p_org, int i_org, me_pel *pred0, me_pel *pred1, me_pel *pred2, int i_pred,
                           uint32_t sad[3], int height);
    /* com_util.h:318 */
    void (*cost_sad_x4[6])(me_pel *p_org, int i_org, me_pel *pred0, me_pel *pred1, me_pel *pred2, me_pel *pred3,
                           int i_pred, uint32_t sad[4], int height);
    /* com_util.h:322 */
    uint32_t (*cost_satd[3][3])(me_pel *p_org, int i_org, me_pel *p_pred, int i_pred);
    /* com_util.h:330 */
    void (*pel_avrg[6])(me_pel *dst, int i_dst, me_pel *src1, me_pel *src2, int height);
} me_dispatch_t;

/* Overwrites every slot above with its HIP entry, remembering the previous
 * contents for the optional fallback.  Replaces, for these slots,
 * uavs3e_funs_init_sse()/uavs3e_funs_init_avx2() (sse.c:213, avx2.c:41). */
int me_funs_init_hip(me_dispatch_t *table);

/* The entries themselves, for callers that bind them one by one. */
uint32_t me_hip_sad(int width, me_pel *p_org, int i_org, me_pel *p_pred, int i_pred, int height);        /* com_cost.c:40-58 */
void me_hip_sad_x3(int width, me_pel *p_org, int i_org, me_pel *p0, me_pel *p1, me_pel *p2, int i_pred,
                   uint32_t sad[3], int height);                                                          /* com_cost.c:60-85 */
void me_hip_sad_x4(int width, me_pel *p_org, int i_org, me_pel *p0, me_pel *p1, me_pel *p2, me_pel *p3,
                   int i_pred, uint32_t sad[4], int height);                                              /* com_cost.c:87-115 */
uint32_t me_hip_satd(int log2w_m2, int log2h_m2, me_pel *p_org, int i_org, me_pel *p_pred, int i_pred);  /* com_cost.c:176-710 */
uint32_t me_hip_had(int w, int h, me_pel *org, int s_org, me_pel *cur, int s_cur, int bit_depth);        /* com_had, com_cost.c:712-812 */
void me_hip_ipcpy(const me_pel *src, int i_src, me_pel *dst, int i_dst, int width, int height);           /* com_mc.c:649-657 */
void me_hip_ipflt(int type, const me_pel *src, int i_src, me_pel *dst, int i_dst, int width, int height,
                  const int8_t *coeff, int max_val);                                                      /* com_mc.c:461-491, 555-585 */
void me_hip_ipflt_ext(int type, const me_pel *src, int i_src, me_pel *dst, int i_dst, int width, int height,
                      const int8_t *coef_x, const int8_t *coef_y, int max_val);                           /* com_mc.c:493-549, 587-643 */
void me_hip_pel_avrg(int width, me_pel *dst, int i_dst, me_pel *src1, me_pel *src2, int height);         /* com_util.c:3088-3105 */
void me_hip_ip_flt_y_hor(const me_pel *src, int i_src, me_pel *dst[3], int i_dst, int16_t *dst_tmp[3], int i_dst_tmp,
                         int width, int height, int8_t (*coeff)[8], int bit_depth);                       /* com_mc.c:668-702 */
void me_hip_ip_flt_y_ver(const me_pel *src, int i_src, me_pel *dst[3], int i_dst, int width, int height,
                         int8_t (*coeff)[8], int bit_depth);                                              /* com_mc.c:704-729 */
void me_hip_ip_flt_y_ver_ext(const int16_t *src, int i_src, me_pel *dst[3], int i_dst, int width, int height,
                             int8_t (*coeff)[8], int bit_depth);                                          /* com_mc.c:731-758 */

/* Self-check of the path's one floating-point step (no counterpart in the
 * reference): out[i] = (int)((s0 + i) / sqrt(c) * 2.0) for i < n, evaluated
 * on the GPU by the function every SATD kernel of this library normalises
 * with -- c = 32 (big = 0: the 8x4 / 4x8 tiles, com_cost.c:649, 708) or 128
 * (big = 1: 16x8 / 8x16, com_cost.c:467, 586).  out: host memory. */
int me_hip_satd_norm(int big, uint32_t s0, uint32_t n, uint32_t *out);

/* ------------------------------------------------------------------------ */
/* 2. Batched frame ME (no counterpart in the reference)                    */
/* ------------------------------------------------------------------------ */

/* One result per (PU, reference). */
typedef struct me_pu_result_t {
    int16_t mv_int[2];     /* integer-pel full-search MV (me.c:553 mv_ipel) */
    int16_t mv_sub[2];     /* quarter-pel refined MV (me_sub_pel_search, me.c:366-500) */
    uint64_t cost_int;     /* u64 integer-pel cost (me.c:151-187); UINT64_MAX if not searched */
    uint64_t cost_sub;     /* u64 sub-pel cost (me.c:409-487) */
    uint32_t cost_final;   /* (u32)((cost_sub + lambda*extra_bits + 2^15) >> 16) (me.c:560-567) */
    uint32_t valid;        /* 1 if the PU lies inside the picture and its window is not empty */
} me_pu_result_t;

/* A reference picture resident in device memory: the edge-padded integer
 * luma plane and its 15 quarter-pel planes, laid out like the reference's
 * com_pic_t / com_subpel_t (com_util.c:62-91, 251-284): plane index
 * dy*4+dx, pad ME 144 luma samples (PIC_PAD_SIZE_L), stride pic_w + 288. */
typedef struct me_ref_s *me_ref_t;

int me_ref_create(me_ref_t *out, int pic_w, int pic_h);
/* Loads luma (pic_h rows of pic_w samples at `stride`) from host memory, or
 * from device memory when on_device != 0, pads it (com_img_padding, com_util.c:
 * 105-165) and builds the 15 planes (com_if_luma_frame, com_mc.c:760-800).
 * stream: a hipStream_t or NULL. */
int me_ref_load(me_ref_t ref, const me_pel *luma, int stride, int on_device, void *stream);
/* Loads the 4:2:0 chroma planes (pic_h/2 rows of pic_w/2 samples each, at
 * `stride`) used by me_plan_mc; reads outside them see the nearest edge
 * sample, as in the reference's padded chroma planes (pad 72, com_util.c:
 * 148-165). */
int me_ref_load_chroma(me_ref_t ref, const me_pel *u, const me_pel *v, int stride, int on_device, void *stream);
/* Device address of plane `idx` at sample (0,0); *stride and *pad receive the
 * layout.  Valid until me_ref_destroy. */
int me_ref_plane(me_ref_t ref, int idx, me_pel **plane, int *stride, int *pad);
/* Copies the whole padded plane `idx` ((pic_h+288) rows of pic_w+288
 * samples) to host memory. */
int me_ref_download_plane(me_ref_t ref, int idx, me_pel *out);
void me_ref_destroy(me_ref_t ref);

/* Frame-level configuration (the batched workloads use MVP = (0,0) for every
 * PU and AMVR precision 0). */
typedef struct me_frame_cfg_t {
    int pic_w, pic_h;
    int search_range;      /* R: integer full search over [-R, R]^2 clipped like me.c:502-527 */
    uint32_t lambda_mv;    /* floor(65536*sqrt(lambda)) (uavs3e.c:851) */
    uint32_t shape_mask;   /* bit s selects shape s of the 29-shape list below */
    int num_refs;          /* references searched per PU (tbl_refi_bits row) */
    int band_rows;         /* tuning: displacement rows per workgroup, 0 = default */
    int waves;             /* tuning: waves per workgroup (1, 2, 4 or 8), 0 = default */
    int dx_cols;           /* tuning: displacement columns per workgroup (multiple of 4), 0 = all */
    int num_refs_l1;       /* 0: every reference is in list 0.  n > 0: refs[num_refs-n ..] form list 1
                              (bi-directional frames); the final cost of a reference then uses
                              tbl_refi_bits of its own list (tables.c:52-70) and the best reference is
                              chosen per list */
    int sp_groups;         /* tuning: sub-pel search positions of a PU evaluated side by side by one
                              workgroup (1, 2, 4 or 8), 0 = 1 */
    int int_grid;          /* tuning: integer-search grid.  0: persistent workgroups (as many as the
                              GPU holds at once, each sweeping a contiguous run of displacement bands
                              and writing each key once) where the frame has the work to fill the GPU
                              twice over; 1: one workgroup per band, keys merged with atomics; n >= 2:
                              n persistent workgroups */
} me_frame_cfg_t;

/* The 29 PU shapes, in mask-bit order: 4x8 8x4 4x16 16x4 4x32 32x4 8x8 8x16
 * 16x8 8x32 32x8 8x64 64x8 16x16 16x32 32x16 16x64 64x16 16x128 128x16 32x32
 * 32x64 64x32 32x128 128x32 64x64 64x128 128x64 128x128.  PUs of a CTU are
 * numbered shape by shape in this order, each shape tiled in raster order
 * over the 128x128 CTU; CTUs are numbered in raster order. */
int me_shape_w(int s);
int me_shape_h(int s);
int me_pus_per_ctu(uint32_t shape_mask);

typedef struct me_plan_s *me_plan_t;

int me_plan_create(me_plan_t *out, const me_frame_cfg_t *cfg);
/* Current picture, from host or device memory. */
int me_plan_set_cur(me_plan_t plan, const me_pel *cur, int stride, int on_device, void *stream);
/* Enqueues ME of the whole frame against refs[0..num_refs-1] on `stream`
 * as one graph launch.  Bit r of prepare_mask re-pads and re-interpolates
 * reference r from its device luma inside the same graph (a reference that
 * was just reconstructed).  Results stay on the device until me_plan_results. */
int me_plan_run(me_plan_t plan, me_ref_t const *refs, uint32_t prepare_mask, void *stream);
/* Autotuner of the search schedule (no counterpart in the reference).
 * Times cfg's frame against refs, with the current picture `cur` (host
 * memory, `stride` samples per row), `iters` runs per candidate (HIP events
 * on the plan's stream): first the integer search for every workgroup size
 * (1, 2, 4, 8 waves) x displacement rows per workgroup (4..32) x
 * displacement columns per workgroup (all, 64, 32) whose LDS window fits
 * (persistent grid), then the fastest of those with one workgroup per band
 * (int_grid 1), then, with the faster, the sub-pel search for every
 * position group count (1, 2, 4, 8).  Writes the fastest into
 * cfg->band_rows / waves / dx_cols / int_grid / sp_groups.  The schedule
 * never changes results.  log (optional) receives up to log_cap timings, one
 * per candidate. */
/* The integer search's grid of a plan (no counterpart in the reference):
 * *groups workgroups and *slots key slots per displacement-column tile (0:
 * one workgroup per band, keys merged with atomics). */
int me_plan_int_grid(me_plan_t plan, int *groups, int *slots);
typedef struct me_tune_result_t {
    int band_rows, waves, dx_cols;
    double int_search_ms;
    int sp_groups;
    double subpel_ms;
    int int_grid;
} me_tune_result_t;
int me_plan_autotune(me_frame_cfg_t *cfg, me_ref_t const *refs, const me_pel *cur, int stride, int iters,
                     me_tune_result_t *log, int log_cap, int *n_log);

/* The reference encoder's own search instead of the full search:
 * me_search_tz (me.c:529-568) -- diamond, conditional raster and refinement
 * around each PU's MVP, then me_sub_pel_search -- for every PU of the frame
 * and every reference, uni-prediction.  mvp: quarter-pel MVPs
 * [(ctu*npu + pu)*num_refs + r][2] (host memory); mv_ipel_prev: the previous
 * integer MV of each (PU, reference) for the adaptive raster range
 * (inter_search_t.mv_ipel, me.c:553), or NULL for zeros.  Results and the
 * per-list best reference are read with me_plan_results (cost_int: the
 * integer-pel cost; cost_sub: the cost after sub-pel, before the extra
 * bits; cost_final: the u32 me_search_tz returns).  Runs on `stream` (NULL:
 * the plan's stream), outside the plan's graph. */
typedef struct me_tz_cfg_t {
    int gop_size;               /* get_raster_range (me.c:65-100) */
    int max_search_range;       /* inter_search_t.max_search_range */
    int adaptive_raster_range;
    int mvr;                    /* AMVR index 0..4 */
    int cost_type;              /* sub-pel: 0 = SATD, 1 = SAD, then SATD at the winner */
    int64_t cur_poc;
    int64_t ref_poc[4];         /* POC of each reference */
} me_tz_cfg_t;
int me_plan_run_tz(me_plan_t plan, me_ref_t const *refs, const int16_t *mvp, const int16_t *mv_ipel_prev,
                   const me_tz_cfg_t *tz, void *stream);

/* Affine motion compensation of luma for a batch of CUs against one
 * reference: com_mc_blk_affine_luma (com_mc.c:202-250) with two control
 * points, the luma part of com_mc_blk_affine (com_mc.c:252-333, through
 * com_mc_cu_affine, com_mc.c:365-385) with two or three (vertex_num).  The
 * model comes from the control-point MVs (CPMV, 1/16 pel: top-left,
 * top-right, bottom-left); every sub_w x sub_h sub-block (4 or 8) takes the
 * MV at its centre (the top-left, top-right and, with three control points,
 * bottom-left sub-blocks the MVs at the CU's corners), rounded to 1/16 pel
 * and clipped to 18 bits, and is interpolated from the integer plane with
 * the 1/16-pel 8-tap filters (com_tables.c:130-147) as com_mc_blk_luma does
 * with hp_flag (position clipped to [-132, pic + 4]).  CU sizes 16..128
 * (powers of two).  pred (host memory, pred_elems samples) receives each
 * CU's w x h prediction row by row at its pred_offset; samples outside
 * every CU's block are left as they were.  Blocking.  The reference holds
 * the device staging buffers of both affine calls: calls on one reference
 * must not overlap. */
typedef struct me_affine_cu_t {
    int32_t x, y, w, h;
    int32_t sub_w, sub_h;
    int32_t vertex_num;       /* 2 or 3 control points */
    int32_t reserved;
    int32_t cpmv[3][2];
    uint64_t pred_offset;
    uint64_t chroma_offset;   /* me_affine_mc_chroma: where the (w/2) x (h/2) u and v blocks go */
} me_affine_cu_t;
int me_affine_mc_luma(me_ref_t ref, const me_affine_cu_t *cus, int n, me_pel *pred, uint64_t pred_elems,
                      void *stream);
/* The chroma of com_mc_blk_affine (com_mc.c:335-360) for the same CUs
 * (square 4x4 or 8x8 sub-blocks, as com_mc_cu_affine uses): each 8x8 luma
 * block's 4x4 chroma takes the rounded average of its four 4x4 luma units'
 * sub-block MVs, read as a 1/32-pel chroma position and interpolated with
 * the 4-tap 1/32-pel filters (com_tables.c:156-189) as com_mc_blk_chroma
 * does with hp_flag (position clipped to [-66, (pic + 4) / 2]; reads outside
 * the chroma planes see the nearest edge sample, as in the reference's
 * padded planes).  Needs me_ref_load_chroma.  pred_u / pred_v (host memory,
 * chroma_elems samples each) receive each CU's blocks at chroma_offset. */
int me_affine_mc_chroma(me_ref_t ref, const me_affine_cu_t *cus, int n, me_pel *pred_u, me_pel *pred_v,
                        uint64_t chroma_elems, void *stream);

/* Affine gradient motion search, affine_me_gradient (inter.c:1339-1471),
 * for a batch of CUs against one reference: from each CU's start control
 * points (two, the 4-parameter model), up to 7 iterations (5 with bi) of --
 * the affine luma prediction (as me_affine_mc_luma), the Sobel gradients of
 * the prediction and the 4-parameter least-squares system with the residual
 * (affine_sobel_flt_hor/ver, affine_coef_computer, com_util.c:2966-3049),
 * solved in IEEE double as solve_equal does (inter.c:1280-1315), the
 * control-point deltas rounded to the AMVR precision; zero deltas end the
 * search; moved control points (clipped to 18 bits) are kept when
 * MV_COST(affine_mv_bits (+ the other list's mot_bits with bi)) + (com_had
 * >> bi) is strictly lower.  Per CU: mv receives the best control points
 * and cost the best cost without its MV bits (the function's return value).
 * org (host memory, org_elems samples) holds each CU's w x h original at
 * org_offset with org_stride samples per row: the current picture at the
 * CU for the uni-prediction search, the bi-prediction target of
 * analyze_affine_bi (inter.c:1760-1790) with bi = 1.  CU sizes 16..128
 * (powers of two), sub-blocks 4 or 8, mvr 0..2 (Tab_Affine_AMVR: 1/4, 1,
 * 1/16 pel), num_refp 1..4.  Blocking. */
typedef struct me_affine_me_cu_t {
    int32_t x, y, w, h;
    int32_t sub_w, sub_h;
    int32_t refi, num_refp;     /* tbl_refi_bits of the reference */
    int32_t mvr;                /* affine AMVR index 0..2 */
    int32_t bi;                 /* 1: the bi-prediction form */
    int32_t mot_bits_other;     /* bi: the other list's mot_bits */
    int32_t org_stride;
    int32_t mvp[2][2];          /* control-point predictors, 1/16 pel */
    int32_t mv[2][2];           /* in: start control points; out: the best */
    uint64_t org_offset;
    uint64_t cost;              /* out */
} me_affine_me_cu_t;
int me_affine_me_gradient(me_ref_t ref, const me_pel *org, uint64_t org_elems, me_affine_me_cu_t *cus, int n,
                          uint32_t lambda_mv, void *stream);

/* Bi-prediction ME of analyze_bi (inter.c:972-1033), the search part, for
 * every PU of the plan that has references in both lists: starting from the
 * uni-prediction results, up to two rounds of -- the prediction of the fixed
 * list (com_mc_cu luma, com_mc.c:140-165) makes the target 2*org - pred
 * clipped (create_bi_org, inter.c:55-68); every reference of the other list
 * is searched against it by me_search_tz with bi = 1 (core square around
 * the previous MV, halved distortion, the fixed list's mot_bits in the
 * final cost, me.c:199-355, 366-500, 529-568); a strictly cheaper cost than
 * the best so far takes that list's reference, MV and mot_bits (get_mv_bits,
 * inter.c:50-53); the searched list becomes the fixed one, and a round
 * without a change ends the search.  The RDO after the search
 * (inter.c:1035-1070) is the caller's.
 * Per PU (pus[ctu*npu + pu]): refi, mv, lidx_ref and mot_bits on entry are
 * analyze_bi's refi_L0L1, mv_L0L1, lidx_ref and pi->mot_bits; on return refi,
 * mv and mot_bits are cur_info->refi / mv and pi->mot_bits after the search
 * and cost is best_mecost (u32 max when nothing was searched).  active = 0
 * skips a PU.  mvp and mv_scale are [(ctu*npu + pu)*num_refs + r][2] quarter
 * pel (pi->mvp_scale / mv_scale, list 0 first); mv_scale is updated as the
 * reference updates it.  tz: mvr and cost_type.  Blocking (results on the
 * host when the call returns); the plan's results are overwritten. */
typedef struct me_bi_pu_t {
    int8_t refi[2];
    int8_t lidx_ref;
    int8_t active;
    int16_t mv[2][2];
    int32_t mot_bits[2];
    uint32_t cost;
} me_bi_pu_t;
int me_plan_run_bi(me_plan_t plan, me_ref_t const *refs, const int16_t *mvp, int16_t *mv_scale, me_bi_pu_t *pus,
                   const me_tz_cfg_t *tz, void *stream);

/* Symmetric-MVD search of analyze_smvd (inter.c:1194-1255) with
 * smvd_refine (inter.c:1083-1191) for every PU of the plan: list 0 and
 * list 1 reference 0, list 1's MVD the negation of list 0's.  Candidates:
 * both MVPs; list 0's uni MV mirrored; then smvd_refine's diamond (8 rounds)
 * and cross (1 round) around list 0's MV, offsets scaled by the AMVR shift;
 * each costs get_smvd_cost of the bi-average prediction (com_mc_cu) --
 * com_had, or SAD >> (bd - 8) with cost_type 1, then com_had of the winner --
 * plus MV_COST of the MVD bits.  pus[ctu*npu + pu]: active selects the PU;
 * mv receives the two MVs (quarter pel) and cost the u32 mecost.  mvp and
 * mv_scale as in me_plan_run_bi (reference 0 of each list is read).  tz:
 * mvr and cost_type.  The RDO after the search (inter.c:1257-1275) is the
 * caller's.  Blocking. */
typedef struct me_smvd_pu_t {
    int8_t active;
    int8_t pad[3];
    int16_t mv[2][2];
    uint32_t cost;
} me_smvd_pu_t;
int me_plan_run_smvd(me_plan_t plan, me_ref_t const *refs, const int16_t *mvp, const int16_t *mv_scale,
                     me_smvd_pu_t *pus, const me_tz_cfg_t *tz, void *stream);

/* The inter part of the lookahead's cost estimator, loka_estimate_coding_cost
 * (lookahead.c:82-208), for every whole 16x16 unit of the plan's current
 * picture: each reference (the plan's list-0 references, then list 1) is
 * searched by me_search_tz from MVP 0 at integer precision (curr_mvr 2, no
 * adaptive raster range, lambda_mv floor(65536*sqrt(1.43631*2^4)) = 314169,
 * lookahead.c:85-110; gop_size, max_search_range and the POCs from tz, its
 * mvr / cost_type / adaptive_raster_range ignored); the unit's cost is the
 * least com_had of (a) each reference's prediction at its MV, (b) with both
 * lists, the average of each list's first cheapest prediction and (c) the
 * average of both lists' first reference at MV 0 (lookahead.c:130-208).
 * is_padding: inter_search_init's flag (inter.c:2015-2027) -- 1 lets windows
 * reach 127 samples outside the picture (the picture-encoding pinter,
 * uavs3e.c:185), 0 keeps every block inside it (the lookahead thread's,
 * uavs3e.c:1326).  The intra part of the estimator (lookahead.c:210-231) is
 * the caller's: the picture's cost is the sum over units of the minimum of
 * this and the intra cost.  unit_cost: host array of (pic_w/16)*(pic_h/16)
 * u32 in raster order, written before the call returns.  The plan needs the
 * 16x16 shape; its 16x16 results are overwritten. */
int me_plan_lookahead(me_plan_t plan, me_ref_t const *refs, const me_tz_cfg_t *tz, int is_padding,
                      uint32_t *unit_cost, void *stream);

/* The frame search of me_plan_create + me_plan_run + me_plan_results in one
 * call that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The picture
 * descriptor, the rate tables and the per-(CTU, PU, reference) keys live in
 * a fixed 8 MiB device array of the library (one per GPU, held until the
 * process ends); a frame whose keys and tables need more (8 bytes per PU and
 * reference, e.g. 1080p with all 29 shapes and 2 references takes 5.9 MB)
 * is rejected with an error.  cfg as for me_plan_create (band_rows, waves
 * and dx_cols choose the schedule); cur: device memory, pic_h rows of pic_w
 * samples at stride pic_w followed by at least 64 readable bytes; refs: the
 * cfg->num_refs references, already loaded; res / best_ref: device memory
 * laid out as me_plan_results writes them.  Enqueued on `stream` (NULL: the
 * null stream); the results are complete when the stream reaches this point.
 * Every call on a GPU uses the same array: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialised by the caller. */
int me_frame_search_noalloc(const me_frame_cfg_t *cfg, const me_pel *cur, me_ref_t const *refs, me_pu_result_t *res,
                            int8_t *best_ref, void *stream);

/* Waits for the last run and copies results to host memory:
 * res[(ctu*npu + pu)*num_refs + r], best_ref[(ctu*npu + pu)*L + l] with L = 1
 * (single list) or 2 (num_refs_l1 > 0): the reference index (into refs) of
 * list l's best reference, -1 = PU skipped. */
int me_plan_results(me_plan_t plan, me_pu_result_t *res, int8_t *best_ref);
/* The results as an encoder consumes them, 8 bytes per (PU, list) instead
 * of 32 per (PU, reference): the best reference's quarter-pel MV and u32
 * final cost (the value me_search_tz returns, me.c:560-567), cost
 * 0xFFFFFFFF for a skipped PU; best_ref (optional) as in me_plan_results.
 * out[(ctu*npu + pu)*L + l] in host memory (pinned memory from
 * me_pinned_alloc makes the copy a single DMA).  Enqueued on `stream` (NULL:
 * the plan's) after the last run; returns when the results are on the host. */
typedef struct me_pu_best_t {
    int16_t mv[2];
    uint32_t cost;
} me_pu_best_t;
int me_plan_best(me_plan_t plan, me_pu_best_t *out, int8_t *best_ref, void *stream);
/* Page-locked host memory for uploads and result downloads. */
int me_pinned_alloc(void **out, uint64_t bytes);
void me_pinned_free(void *p);
/* Motion compensation of every PU from the last run's results: com_mc_cu
 * (com_mc.c:140-200) with the best reference of each list and its
 * quarter-pel MV -- luma copied from the quarter-pel planes, 4:2:0 chroma
 * by the 1/8-pel 4-tap filters of com_mc_blk_chroma (com_mc.c:103-138), and
 * with two lists the rounding average (a + b + 1) >> 1 of pel_avrg
 * (com_util.c:3088-3105).  Chroma is predicted for shapes of width and
 * height >= 8 (the chroma slots widx - 1 of com_mc_cu exist only there) and
 * needs me_ref_load_chroma on every reference.
 * Layout per CTU: the selected shapes in table order, each shape's PUs in
 * raster order, each PU's w x h samples row by row; luma takes 16384
 * samples per shape, u and v 4096 per chroma shape; skipped PUs are zero.
 * pred_* are device buffers of me_plan_mc_size elements, or all NULL to use
 * buffers owned by the plan (read back with me_plan_mc_download).
 * Enqueued on `stream` (NULL: the plan's stream) after the ME of the frame. */
int me_plan_mc(me_plan_t plan, me_ref_t const *refs, me_pel *pred_y, me_pel *pred_u, me_pel *pred_v, void *stream);
int me_plan_mc_size(me_plan_t plan, uint64_t *luma_elems, uint64_t *chroma_elems);
/* Waits for the plan's stream and copies the plan-owned predictions to host. */
int me_plan_mc_download(me_plan_t plan, me_pel *pred_y, me_pel *pred_u, me_pel *pred_v);
/* Runs `iters` frames stage by stage (no graph) on the plan's own stream and
 * writes the average milliseconds of each stage to ms[0..3]: reference
 * preparation, key initialisation, integer search, sub-pel + reference choice. */
int me_plan_profile(me_plan_t plan, me_ref_t const *refs, uint32_t prepare_mask, int iters, double ms[4]);
int me_plan_num_ctus(me_plan_t plan);
int me_plan_num_pus(me_plan_t plan);   /* per CTU */
void me_plan_destroy(me_plan_t plan);

/* ------------------------------------------------------------------------ */
/* 3. Request server (no counterpart in the reference)                      */
/* ------------------------------------------------------------------------ */

/* Continuous batching of per-CTU ME requests from several concurrent
 * streams of different resolutions.  Pictures stay resident on the device
 * (integer plane + 15 quarter-pel planes each) and are admitted against an
 * HBM budget; requests are gathered into batches of up to max_batch_ctus
 * CTUs from any stream, up to max_inflight batches run on the device at
 * once on their own streams, and a request that arrives while batches run
 * joins the next batch as soon as a batch slot frees up.  Every request's
 * result is the same as for the frame API (and the same whatever else is in
 * its batch).  All entry points are thread-safe. */
typedef struct me_server_cfg_t {
    int search_range;       /* R of the integer full search (the same for every stream) */
    uint32_t lambda_mv;     /* floor(65536*sqrt(lambda)) (uavs3e.c:851) */
    uint32_t shape_mask;    /* PU shapes searched in each CTU (29-shape list) */
    int num_refs;           /* references per request, 1..4 */
    uint64_t hbm_budget;    /* device bytes pictures may hold in total */
    int max_batch_ctus;     /* CTUs per batch, 1..65535 */
    int max_inflight;       /* batches on the device at once, 1..8 */
    int max_wait_us;        /* while other batches run, a partial batch waits at most this long to fill */
    int band_rows, waves, dx_cols;  /* search schedule, as in me_frame_cfg_t (0 = default) */
} me_server_cfg_t;

typedef struct me_server_stats_t {
    uint64_t requests_done;
    uint64_t batches;
    uint64_t max_batch_seen;      /* largest batch launched */
    uint64_t max_inflight_seen;   /* most batches on the device at once */
    uint64_t mixed_batches;       /* batches holding CTUs of more than one stream */
    uint64_t picture_bytes;       /* device bytes held by pictures now (live + pooled) */
    uint64_t admission_waits;     /* me_server_put_picture calls that had to wait for budget */
    uint64_t max_launch_us;       /* longest enqueue of one batch by the dispatcher (us) */
    uint64_t max_batch_us;        /* longest time from a batch's enqueue to its results on the host (us) */
} me_server_stats_t;

typedef struct me_server_s *me_server_t;

int me_server_create(me_server_t *out, const me_server_cfg_t *cfg);
/* Registers a stream of pic_w x pic_h pictures; *stream_id receives its id. */
int me_server_open_stream(me_server_t srv, int pic_w, int pic_h, int *stream_id);
/* Makes picture `poc` of a stream resident: uploads its luma (host memory,
 * `stride` samples per row) and builds its padded planes.  Waits while the
 * HBM budget is exhausted, up to timeout_ms (-1: no limit).  Returns 0 when
 * resident, 1 when the budget did not free up in time, negative on error. */
int me_server_put_picture(me_server_t srv, int stream_id, int64_t poc, const me_pel *luma, int stride, int timeout_ms);
/* No further requests will name this picture; its memory returns to the
 * budget once the requests already submitted with it have completed. */
int me_server_release_picture(me_server_t srv, int stream_id, int64_t poc);
/* Submits ME of CTU `ctu` (raster index) of picture cur_poc against the
 * pictures ref_pocs[0..num_refs-1] of the same stream.  On completion
 * res[pu * num_refs + r] and best_ref[pu] (host memory that must stay valid
 * until then; PU numbering as in the frame API) are written.  Returns a
 * ticket (> 0) or a negative value on error. */
int64_t me_server_submit(me_server_t srv, int stream_id, int64_t cur_poc, int ctu, const int64_t *ref_pocs,
                         me_pu_result_t *res, int8_t *best_ref);
/* Waits for a ticket up to timeout_ms (-1: no limit).  Returns 0 when done
 * (the ticket is then retired), 1 on timeout.  done_us (optional) receives
 * the completion time on the me_server_now_us clock. */
int me_server_wait(me_server_t srv, int64_t ticket, int timeout_ms, double *done_us);
/* Monotonic clock of the server in microseconds (CLOCK_MONOTONIC). */
double me_server_now_us(void);
int me_server_stats(me_server_t srv, me_server_stats_t *stats);
/* Waits for every submitted request, then frees the server and its pictures. */
void me_server_destroy(me_server_t srv);

#ifdef __cplusplus
}
#endif
#endif
