/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Thin C entry points around the *compiled reference* (uavs3e built from its
 * own sources by oracle/ref.mk into oracle/_ref/).  Used to write golden
 * vectors, to cross-check the oracle restatement, and as the timed CPU
 * baseline ("reference AVX2 path") of bench.py.  Nothing in the product
 * links it.
 *
 * me.c and inter.c are compiled as part of this translation unit (they are
 * #included from the reference tree, not copied) so that their static
 * functions can be called: get_max_search_range and me_sub_pel_search
 * (me.c:366-527), analyze_bi (inter.c:972-1070), smvd_refine
 * (inter.c:1083-1191) and the helpers analyze_smvd uses.
 */
#define _GNU_SOURCE
#include "define.h"
#include "me.c"
#include "inter.c"
#include <pthread.h>
#include <sched.h>
#include <time.h>

#define RH_EXPORT __attribute__((visibility("default")))

static int g_level = -1;

/* Fill the dispatch table the way uavs3e_create does (uavs3e.c:1332-1342):
 * C first, then SSE, then AVX2 when the CPU has it.  level 0 = C only. */
RH_EXPORT int refh_init(int level)
{
    uavs3e_funs_init_c();
    if (level >= 1) uavs3e_funs_init_sse();
    if (level >= 2 && uavs3e_simd_avx_level(NULL) >= 2) uavs3e_funs_init_avx2();
    com_dct_coef_create();
    g_level = level;
    return BIT_DEPTH;
}

RH_EXPORT int refh_bit_depth(void) { return BIT_DEPTH; }

static int wlog(int w) { return CONV_LOG2(w) - 2; }

RH_EXPORT u32 refh_sad(int w, pel *org, int i_org, pel *pred, int i_pred, int h)
{ return uavs3e_funs_handle.cost_sad[wlog(w)](org, i_org, pred, i_pred, h); }

RH_EXPORT void refh_sad_x3(int w, pel *org, int i_org, pel *p0, pel *p1, pel *p2, int i_pred, u32 *sad, int h)
{ uavs3e_funs_handle.cost_sad_x3[wlog(w)](org, i_org, p0, p1, p2, i_pred, sad, h); }

RH_EXPORT void refh_sad_x4(int w, pel *org, int i_org, pel *p0, pel *p1, pel *p2, pel *p3, int i_pred, u32 *sad, int h)
{ uavs3e_funs_handle.cost_sad_x4[wlog(w)](org, i_org, p0, p1, p2, p3, i_pred, sad, h); }

RH_EXPORT u32 refh_satd(int wl, int hl, pel *org, int s_org, pel *cur, int s_cur)
{ return uavs3e_funs_handle.cost_satd[wl][hl](org, s_org, cur, s_cur); }

RH_EXPORT u32 refh_had(int w, int h, pel *org, int s_org, pel *cur, int s_cur, int bd)
{ return com_had(w, h, org, s_org, cur, s_cur, bd); }

RH_EXPORT void refh_ipcpy(const pel *src, int i_src, pel *dst, int i_dst, int w, int h)
{ uavs3e_funs_handle.ipcpy[wlog(w)](src, i_src, dst, i_dst, w, h); }

RH_EXPORT void refh_ipflt(int type, const pel *src, int i_src, pel *dst, int i_dst, int w, int h, const s8 *coef, int max_val)
{ uavs3e_funs_handle.ipflt[type][wlog(w)](src, i_src, dst, i_dst, w, h, coef, max_val); }

RH_EXPORT void refh_ipflt_ext(int type, const pel *src, int i_src, pel *dst, int i_dst, int w, int h,
                              const s8 *cx, const s8 *cy, int max_val)
{ uavs3e_funs_handle.ipflt_ext[type][wlog(w)](src, i_src, dst, i_dst, w, h, cx, cy, max_val); }

RH_EXPORT void refh_pel_avrg(int w, pel *dst, int i_dst, pel *s1, pel *s2, int h)
{ uavs3e_funs_handle.pel_avrg[wlog(w)](dst, i_dst, s1, s2, h); }

RH_EXPORT const s8 *refh_luma_coef(int frac) { return com_tbl_mc_l_coeff[frac & 3]; }
RH_EXPORT const s8 *refh_chroma_coef(int frac) { return com_tbl_mc_c_coeff[frac & 7]; }

/* MV-bits table from inter_search_create (inter.c:2029-2058); copies
 * 2*off+1 entries into out (out[off] is mv 0) and returns off. */
RH_EXPORT int refh_mvbits(int pic_w, int pic_h, u8 *out, int cap)
{
    com_info_t *info = (com_info_t *)calloc(1, sizeof(com_info_t));
    u8 *tab = NULL;
    info->pic_width = pic_w;
    info->pic_height = pic_h;
    int off = inter_search_create(&tab, info);
    if (2 * off + 1 <= cap) memcpy(out, tab - off, (size_t)(2 * off + 1));
    inter_search_free(tab, off);
    free(info);
    return off;
}

/* ---------------------------------------------------------------------- */
/* Reference pictures with sub-pel planes                                   */
/* ---------------------------------------------------------------------- */

typedef struct {
    com_pic_t *pic;
    int w, h;
} refh_pic_t;

/* Build a reference picture the way the encoder does after reconstruction
 * (uavs3e.c:1086-1092): copy the luma samples, edge-pad (com_img_padding ..,0),
 * allocate the sub-pel images (com_subpel_create) and run com_if_luma_frame. */
RH_EXPORT void *refh_pic_create(const pel *luma, int w, int h, long long poc)
{
    int err = 0;
    refh_pic_t *r = (refh_pic_t *)calloc(1, sizeof(refh_pic_t));
    r->w = w; r->h = h;
    r->pic = com_pic_create(w, h, PIC_PAD_SIZE_L, PIC_PAD_SIZE_C, &err);
    for (int y = 0; y < h; y++) memcpy(r->pic->y + y * r->pic->stride_luma, luma + y * w, sizeof(pel) * w);
    com_img_padding(r->pic->img, 1, 0);
    r->pic->subpel = com_subpel_create(w, h, PIC_PAD_SIZE_L, PIC_PAD_SIZE_C, &err);
    r->pic->subpel->imgs[0][0] = r->pic->img;
    r->pic->ptr = poc;
    r->pic->img->ptr = poc;   /* what the lookahead reads (lookahead.c:143) */
    int ts = w + 16;
    s16 *tmp[3];
    /* The SIMD frame filters write whole vectors past the end of a row (the
     * encoder carves the three buffers out of one allocation, uavs3e.c:122,
     * 167-169), so give each buffer slack rows. */
    for (int i = 0; i < 3; i++) tmp[i] = (s16 *)calloc((size_t)ts * (h + 16) + 8 * (size_t)ts + 256, sizeof(s16));
    com_if_luma_frame(r->pic->subpel->imgs, tmp, BIT_DEPTH);
    for (int i = 0; i < 3; i++) free(tmp[i]);
    return r;
}

RH_EXPORT int refh_pic_stride(void *p) { return ((refh_pic_t *)p)->pic->stride_luma; }
RH_EXPORT int refh_pic_pad(void *p) { return ((refh_pic_t *)p)->pic->padsize_luma; }

/* Copy one whole padded plane (plane index dy*4+dx) into out. */
RH_EXPORT void refh_pic_plane(void *p, int idx, pel *out)
{
    refh_pic_t *r = (refh_pic_t *)p;
    com_img_t *img = r->pic->subpel->imgs[idx >> 2][idx & 3];
    int s = img->stride[0] / (int)sizeof(pel), pad = img->pad[0];
    const pel *base = (const pel *)img->planes[0] - pad * s - pad;
    memcpy(out, base, sizeof(pel) * (size_t)s * (r->h + 2 * pad));
}

/* Sets the picture's 4:2:0 chroma planes ((w/2) x (h/2) samples each, rows
 * packed) and edge-pads them as the encoder pads a reconstructed picture
 * (com_img_padding(img, 3, 0), uavs3e.c:1087; the luma plane is padded
 * again, which leaves it unchanged). */
RH_EXPORT void refh_pic_set_chroma(void *p, const pel *u, const pel *v)
{
    refh_pic_t *r = (refh_pic_t *)p;
    int cw = r->w / 2, chh = r->h / 2;
    for (int y = 0; y < chh; y++) {
        memcpy(r->pic->u + y * r->pic->stride_chroma, u + y * cw, sizeof(pel) * cw);
        memcpy(r->pic->v + y * r->pic->stride_chroma, v + y * cw, sizeof(pel) * cw);
    }
    com_img_padding(r->pic->img, 3, 0);
}

/* com_mc_cu (com_mc.c:140-200) for one PU at (x, y), w x h: p0 / p1 are the
 * L0 / L1 reference pictures (either may be NULL for uni-prediction), mv0 /
 * mv1 their quarter-pel MVs.  pred_y receives w x h luma samples (stride w),
 * pred_u / pred_v (w/2) x (h/2) chroma samples (stride w/2); channel 0 =
 * luma and chroma, 1 = luma only. */
RH_EXPORT void refh_mc_cu(int x, int y, int pic_w, int pic_h, int w, int h, void *p0, void *p1, const s16 *mv0,
                          const s16 *mv1, int channel, pel *pred_y, pel *pred_u, pel *pred_v)
{
    com_ref_pic_t refp[1][REFP_NUM];
    memset(refp, 0, sizeof(refp));
    s8 refi[REFP_NUM] = {-1, -1};
    s16 mv[REFP_NUM][MV_D] = {{0, 0}, {0, 0}};
    if (p0) { refp[0][REFP_0].pic = ((refh_pic_t *)p0)->pic; refp[0][REFP_0].ptr = refp[0][REFP_0].pic->ptr; refi[REFP_0] = 0; mv[REFP_0][0] = mv0[0]; mv[REFP_0][1] = mv0[1]; }
    if (p1) { refp[0][REFP_1].pic = ((refh_pic_t *)p1)->pic; refp[0][REFP_1].ptr = refp[0][REFP_1].pic->ptr; refi[REFP_1] = 0; mv[REFP_1][0] = mv1[0]; mv[REFP_1][1] = mv1[1]; }
    static __thread pel buf[N_C][MAX_CU_DIM] __attribute__((aligned(32)));
    com_mc_cu(x, y, pic_w, pic_h, w, h, refi, mv, refp, buf, w, channel ? CHANNEL_L : CHANNEL_LC, BIT_DEPTH);
    for (int i = 0; i < h; i++) memcpy(pred_y + i * w, buf[Y_C] + i * w, sizeof(pel) * w);
    if (!channel)
        for (int i = 0; i < h / 2; i++) {
            memcpy(pred_u + i * (w / 2), buf[U_C] + i * (w / 2), sizeof(pel) * (w / 2));
            memcpy(pred_v + i * (w / 2), buf[V_C] + i * (w / 2), sizeof(pel) * (w / 2));
        }
}

RH_EXPORT void refh_pic_free(void *p)
{
    refh_pic_t *r = (refh_pic_t *)p;
    r->pic->subpel->imgs[0][0] = NULL;
    com_subpel_free(r->pic->subpel);
    r->pic->subpel = NULL;
    com_pic_destroy(r->pic);
    free(r);
}

/* ---------------------------------------------------------------------- */
/* Search                                                                   */
/* ---------------------------------------------------------------------- */

typedef struct {
    inter_search_t pi;
    u8 *tab;
} refh_search_t;

static void search_setup(refh_search_t *s, int pic_w, int pic_h, u32 lambda_mv, int cost_type, int is_padding)
{
    com_info_t *info = (com_info_t *)calloc(1, sizeof(com_info_t));
    info->pic_width = pic_w;
    info->pic_height = pic_h;
    info->bit_depth_internal = BIT_DEPTH;
    info->gop_size = 16;
    memset(&s->pi, 0, sizeof(s->pi));
    int off = inter_search_create(&s->tab, info);
    info->tab_mvbits = s->tab;
    inter_search_init(&s->pi, info, is_padding);
    s->pi.tab_mvbits_offset = off;
    s->pi.lambda_mv = lambda_mv;
    s->pi.subpel_cost_type = cost_type;
    s->pi.curr_mvr = 0;
    free(info);
}

static void search_free(refh_search_t *s) { inter_search_free(s->tab, s->pi.tab_mvbits_offset); }

/* me_search_tz (me.c:529-568) exactly as the encoder calls it for uni
 * prediction (inter.c:896), with the SATD caches cleared first as
 * analyze_inter_cu does per CU (inter.c:1907-1908). */
RH_EXPORT u64 refh_me_search_tz(void *ref, const pel *org, int i_org, int x, int y, int w, int h,
                                int refi, int num_refp, long long cur_poc, int gop_size, int max_search_range,
                                int mvr, int cost_type, int adaptive_raster_range, u32 lambda_mv,
                                const s16 *mvp, s16 *mv_out, s16 *mv_ipel_out)
{
    refh_pic_t *r = (refh_pic_t *)ref;
    refh_search_t s;
    search_setup(&s, r->w, r->h, lambda_mv, cost_type, 1);
    inter_search_t *pi = &s.pi;
    pi->gop_size = gop_size;
    pi->max_search_range = max_search_range;
    pi->adaptive_raster_range = adaptive_raster_range;
    pi->ptr = cur_poc;
    pi->ref_pic = r->pic;
    pi->num_refp = (u8)num_refp;
    pi->curr_mvr = (u8)mvr;
    pi->org = (pel *)org;
    pi->i_org = i_org;
    s16 mvp_l[2] = {mvp[0], mvp[1]}, mv[2] = {0, 0};
    u64 c = me_search_tz(pi, x, y, w, h, r->w, r->h, (s8)refi, 0, mvp_l, mv, 0);
    mv_out[0] = mv[0]; mv_out[1] = mv[1];
    mv_ipel_out[0] = pi->mv_ipel[0][refi][0];
    mv_ipel_out[1] = pi->mv_ipel[0][refi][1];
    search_free(&s);
    return c;
}

/* me_search_tz with bi = 1 as analyze_bi calls it (inter.c:1006-1011): org
 * is the bi-prediction target (create_bi_org, inter.c:55-68) with stride
 * i_org, the search starts at mv (quarter pel, updated in place) for list
 * lidx, and the other list's mot_bits enter the final cost. */
RH_EXPORT u64 refh_me_search_tz_bi(void *ref, const pel *org, int i_org, int x, int y, int w, int h,
                                   int refi, int num_refp, int lidx, int mot_bits_other, int mvr, int cost_type,
                                   u32 lambda_mv, const s16 *mvp, s16 *mv)
{
    refh_pic_t *r = (refh_pic_t *)ref;
    refh_search_t s;
    search_setup(&s, r->w, r->h, lambda_mv, cost_type, 1);
    inter_search_t *pi = &s.pi;
    pi->ref_pic = r->pic;
    pi->num_refp = (u8)num_refp;
    pi->curr_mvr = (u8)mvr;
    pi->org = (pel *)org;
    pi->i_org = i_org;
    pi->adaptive_raster_range = 0;
    pi->mot_bits[1 - lidx] = mot_bits_other;
    s16 mvp_l[2] = {mvp[0], mvp[1]};
    u64 c = me_search_tz(pi, x, y, w, h, r->w, r->h, (s8)refi, lidx, mvp_l, mv, 1);
    search_free(&s);
    return c;
}

/* me_sub_pel_search (me.c:366-500) from a given integer MV (quarter-pel
 * units in mv, refined on return), bi = 0, caches cleared. */
RH_EXPORT u64 refh_subpel(void *ref, const pel *org, int i_org, int x, int y, int w, int h,
                          int mvr, int cost_type, u32 lambda_mv, const s16 *mvp, s16 *mv)
{
    refh_pic_t *r = (refh_pic_t *)ref;
    refh_search_t s;
    search_setup(&s, r->w, r->h, lambda_mv, cost_type, 1);
    s.pi.ref_pic = r->pic;
    s.pi.curr_mvr = (u8)mvr;
    s.pi.org = (pel *)org;
    s.pi.i_org = i_org;
    s16 mvp_l[2] = {mvp[0], mvp[1]}, m[2] = {mv[0], mv[1]};
    u64 c = me_sub_pel_search(&s.pi, x, y, w, h, 0, 0, mvp_l, m, 0);
    mv[0] = m[0]; mv[1] = m[1];
    search_free(&s);
    return c;
}

/* get_max_search_range (me.c:502-527); returns ok and writes the window. */
RH_EXPORT int refh_window(int pic_w, int pic_h, int x, int y, int w, int h, int mvr, int is_padding,
                          const s16 *mvp, s16 *min_mv, s16 *max_mv)
{
    refh_search_t s;
    search_setup(&s, pic_w, pic_h, 0, 0, is_padding);
    s.pi.curr_mvr = (u8)mvr;
    int ok = get_max_search_range(&s.pi, mvp, x, y, w, h, pic_w, pic_h);
    min_mv[0] = s.pi.min_mv[0]; min_mv[1] = s.pi.min_mv[1];
    max_mv[0] = s.pi.max_mv[0]; max_mv[1] = s.pi.max_mv[1];
    search_free(&s);
    return ok;
}

/* ---------------------------------------------------------------------- */
/* Full-search frame driver (the timed "reference AVX2 path")             */
/* ---------------------------------------------------------------------- */

static const int SHAPES[29][2] = {
    {4, 8}, {8, 4}, {4, 16}, {16, 4}, {4, 32}, {32, 4}, {8, 8}, {8, 16}, {16, 8}, {8, 32},
    {32, 8}, {8, 64}, {64, 8}, {16, 16}, {16, 32}, {32, 16}, {16, 64}, {64, 16}, {16, 128},
    {128, 16}, {32, 32}, {32, 64}, {64, 32}, {32, 128}, {128, 32}, {64, 64}, {64, 128},
    {128, 64}, {128, 128}};

typedef struct {
    s16 mv_int[2];
    s16 mv_sub[2];
    u64 cost_int;
    u64 cost_sub;
    u32 cost_final;
    u32 valid;
} refh_pu_res_t;

/* Step-1 raster over the window with the table's SAD x4 / x3 / x1 grouping
 * of search_raster (me.c:150-188); strict '<' keeps the first minimum. */
static u64 full_search(inter_search_t *pi, int x, int y, int w, int h, int R, s16 mv[2])
{
    s16 mvp[2] = {0, 0};
    mv[0] = mv[1] = 0;
    if (!get_max_search_range(pi, mvp, x, y, w, h, pi->ref_pic->width_luma, pi->ref_pic->height_luma))
        return COM_UINT64_MAX;
    int cx = COM_CLIP3(pi->min_mv[0], pi->max_mv[0], 0), cy = COM_CLIP3(pi->min_mv[1], pi->max_mv[1], 0);
    int x0 = COM_CLIP3(pi->min_mv[0], pi->max_mv[0], cx - R), x1 = COM_CLIP3(pi->min_mv[0], pi->max_mv[0], cx + R);
    int y0 = COM_CLIP3(pi->min_mv[1], pi->max_mv[1], cy - R), y1 = COM_CLIP3(pi->min_mv[1], pi->max_mv[1], cy + R);
    int i_ref = pi->ref_pic->stride_luma;
    pel *ref = pi->ref_pic->y + y * i_ref + x;
    u8 *tab = pi->tab_mvbits;
    u32 lambda = pi->lambda_mv;
    int shift = 16 - (pi->bit_depth - 8);
    int widx = CONV_LOG2(w) - 2;
    u64 best = COM_UINT64_MAX;
    for (int my = y0; my <= y1; my++) {
        pel *p = ref + my * i_ref;
        u64 cmy = (u64)(tab[my << 2] * lambda);
        int mx = x0;
        u32 sad[4];
        for (; mx + 3 <= x1; mx += 4) {
            uavs3e_funs_handle.cost_sad_x4[widx](pi->org, pi->i_org, p + mx, p + mx + 1, p + mx + 2, p + mx + 3, i_ref, sad, h);
            for (int k = 0; k < 4; k++) {
                u64 c = cmy + (u64)(tab[(mx + k) << 2] * lambda) + ((u64)sad[k] << shift);
                if (c < best) { best = c; mv[0] = (s16)(mx + k); mv[1] = (s16)my; }
            }
        }
        if (mx + 2 <= x1) {
            uavs3e_funs_handle.cost_sad_x3[widx](pi->org, pi->i_org, p + mx, p + mx + 1, p + mx + 2, i_ref, sad, h);
            for (int k = 0; k < 3; k++) {
                u64 c = cmy + (u64)(tab[(mx + k) << 2] * lambda) + ((u64)sad[k] << shift);
                if (c < best) { best = c; mv[0] = (s16)(mx + k); mv[1] = (s16)my; }
            }
            mx += 3;
        }
        for (; mx <= x1; mx++) {
            u64 c = cmy + (u64)(tab[mx << 2] * lambda) +
                    ((u64)uavs3e_funs_handle.cost_sad[widx](pi->org, pi->i_org, p + mx, i_ref, h) << shift);
            if (c < best) { best = c; mv[0] = (s16)mx; mv[1] = (s16)my; }
        }
    }
    return best;
}

typedef struct {
    int pic_w, pic_h;
    const pel *cur; int i_cur;
    int num_refs; void **refs;
    int R; u32 lambda_mv; int cost_type; u32 mask;
    int ctu0, ctu1, npu;
    refh_pu_res_t *res; s8 *best_ref;
    double *ctu_ms;
    int next;                 /* shared work counter */
    int next_worker;          /* worker index for CPU pinning */
    int tz;                   /* 0: full search + sub-pel; 1: me_search_tz from MVP (0,0) */
    int gop_size, max_search_range;
    long long cur_poc;
    pthread_mutex_t lock;
} frame_job_t;

/* Workers are pinned one per CPU of the process's affinity set, in order,
 * when g_pin is set (refh_set_pinning). */
static int g_pin = 0;
RH_EXPORT void refh_set_pinning(int on) { g_pin = on; }

static void pin_worker(int idx)
{
    cpu_set_t allowed, one;
    if (sched_getaffinity(0, sizeof(allowed), &allowed)) return;
    int n = CPU_COUNT(&allowed);
    if (n <= 0) return;
    int k = idx % n, seen = 0;
    for (int c = 0; c < CPU_SETSIZE; c++)
        if (CPU_ISSET(c, &allowed) && seen++ == k) {
            CPU_ZERO(&one);
            CPU_SET(c, &one);
            pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
            return;
        }
}

static void run_ctu(frame_job_t *j, refh_search_t *s, int ctu)
{
    inter_search_t *pi = &s->pi;
    int cols = (j->pic_w + 127) / 128;
    int cx = (ctu % cols) * 128, cy = (ctu / cols) * 128;
    int pu = 0;
    for (int sh = 0; sh < 29; sh++) {
        if (!(j->mask >> sh & 1)) continue;
        int w = SHAPES[sh][0], h = SHAPES[sh][1];
        for (int oy = 0; oy < 128; oy += h)
            for (int ox = 0; ox < 128; ox += w, pu++) {
                int x = cx + ox, y = cy + oy;
                size_t base = (size_t)ctu * j->npu + pu;
                int inside = x + w <= j->pic_w && y + h <= j->pic_h;
                u32 bc = 0xFFFFFFFFu;
                s8 br = -1;
                for (int r = 0; r < j->num_refs; r++) {
                    refh_pu_res_t *o = &j->res[base * j->num_refs + r];
                    memset(o, 0, sizeof(*o));
                    o->cost_int = o->cost_sub = COM_UINT64_MAX;
                    o->cost_final = 0xFFFFFFFFu;
                    if (!inside) continue;
                    pi->ref_pic = ((refh_pic_t *)j->refs[r])->pic;
                    pi->org = (pel *)j->cur + y * j->i_cur + x;
                    pi->i_org = j->i_cur;
                    memset(pi->hpel_satd, 0, sizeof(pi->hpel_satd));
                    memset(pi->qpel_satd, 0, sizeof(pi->qpel_satd));
                    memset(pi->hpel_start_mv, 0, sizeof(pi->hpel_start_mv));
                    memset(pi->qpel_start_mv, 0, sizeof(pi->qpel_start_mv));
                    s16 mv[2], mvp[2] = {0, 0};
                    if (j->tz) {
                        /* me_search_tz as the encoder calls it for uni-prediction (inter.c:896) */
                        pi->gop_size = j->gop_size;
                        pi->max_search_range = j->max_search_range;
                        pi->adaptive_raster_range = 0;
                        pi->ptr = j->cur_poc;
                        pi->num_refp = (u8)j->num_refs;
                        pi->curr_mvr = 0;
                        u64 c = me_search_tz(pi, x, y, w, h, j->pic_w, j->pic_h, (s8)r, 0, mvp, mv, 0);
                        if (c == COM_UINT64_MAX) continue;
                        o->mv_int[0] = pi->mv_ipel[0][r][0];
                        o->mv_int[1] = pi->mv_ipel[0][r][1];
                        o->mv_sub[0] = mv[0]; o->mv_sub[1] = mv[1];
                        o->cost_final = (u32)c;
                        o->valid = 1;
                        if (o->cost_final < bc) { bc = o->cost_final; br = (s8)r; }
                        continue;
                    }
                    u64 ci = full_search(pi, x, y, w, h, j->R, mv);
                    o->mv_int[0] = mv[0]; o->mv_int[1] = mv[1];
                    o->cost_int = ci;
                    if (ci == COM_UINT64_MAX) continue;
                    mv[0] <<= 2; mv[1] <<= 2;
                    u64 cs = me_sub_pel_search(pi, x, y, w, h, (s8)r, 0, mvp, mv, 0);
                    o->mv_sub[0] = mv[0]; o->mv_sub[1] = mv[1];
                    o->cost_sub = cs;
                    int extra = tbl_refi_bits[j->num_refs][r] + 0 + 1 + 2;
                    u64 cf = cs + j->lambda_mv * extra;
                    o->cost_final = (u32)((cf + (1 << 15)) >> 16);
                    o->valid = 1;
                    if (o->cost_final < bc) { bc = o->cost_final; br = (s8)r; }
                }
                j->best_ref[base] = br;
            }
    }
}

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void *frame_worker(void *arg)
{
    frame_job_t *j = (frame_job_t *)arg;
    pthread_mutex_lock(&j->lock);
    const int widx = j->next_worker++;
    pthread_mutex_unlock(&j->lock);
    if (g_pin) pin_worker(widx);
    refh_search_t s;
    search_setup(&s, j->pic_w, j->pic_h, j->lambda_mv, j->cost_type, 1);
    for (;;) {
        pthread_mutex_lock(&j->lock);
        int ctu = j->ctu0 + j->next++;
        pthread_mutex_unlock(&j->lock);
        if (ctu >= j->ctu1) break;
        double t0 = now_ms();
        run_ctu(j, &s, ctu);
        if (j->ctu_ms) j->ctu_ms[ctu - j->ctu0] = now_ms() - t0;
    }
    search_free(&s);
    return NULL;
}

static double run_frame_job(frame_job_t *jp, int threads);

/* Full search + me_sub_pel_search for CTUs [ctu0, ctu1) over `threads`
 * pthreads (dynamic CTU distribution).  Results use the oracle's layout.
 * ctu_ms (optional) receives each CTU's wall time.  Returns total ms. */
RH_EXPORT double refh_me_frame(int pic_w, int pic_h, const pel *cur, int i_cur, int num_refs, void **refs,
                               int R, u32 lambda_mv, int cost_type, u32 mask, int ctu0, int ctu1,
                               int threads, refh_pu_res_t *res, s8 *best_ref, double *ctu_ms)
{
    frame_job_t j;
    memset(&j, 0, sizeof(j));
    j.pic_w = pic_w; j.pic_h = pic_h; j.cur = cur; j.i_cur = i_cur;
    j.num_refs = num_refs; j.refs = refs; j.R = R; j.lambda_mv = lambda_mv;
    j.cost_type = cost_type; j.mask = mask; j.ctu0 = ctu0; j.ctu1 = ctu1;
    j.res = res; j.best_ref = best_ref; j.ctu_ms = ctu_ms;
    return run_frame_job(&j, threads);
}

/* me_search_tz (me.c:529-568) from MVP (0,0), AMVR 0, SATD sub-pel, for
 * every PU of CTUs [ctu0, ctu1) and every reference (their POCs from
 * refh_pic_create; the current picture's is cur_poc), over `threads`
 * pthreads.  res receives mv_int (pi->mv_ipel), mv_sub and cost_final (the
 * u32 me_search_tz returns); best_ref the first cheapest reference.
 * Returns total ms. */
RH_EXPORT double refh_tz_frame(int pic_w, int pic_h, const pel *cur, int i_cur, int num_refs, void **refs,
                               int gop_size, int max_search_range, long long cur_poc, u32 lambda_mv, u32 mask,
                               int ctu0, int ctu1, int threads, refh_pu_res_t *res, s8 *best_ref)
{
    frame_job_t j;
    memset(&j, 0, sizeof(j));
    j.pic_w = pic_w; j.pic_h = pic_h; j.cur = cur; j.i_cur = i_cur;
    j.num_refs = num_refs; j.refs = refs; j.lambda_mv = lambda_mv;
    j.mask = mask; j.ctu0 = ctu0; j.ctu1 = ctu1;
    j.res = res; j.best_ref = best_ref;
    j.tz = 1;
    j.gop_size = gop_size;
    j.max_search_range = max_search_range;
    j.cur_poc = cur_poc;
    return run_frame_job(&j, threads);
}

static double run_frame_job(frame_job_t *jp, int threads)
{
    frame_job_t j = *jp;
    int n = 0;
    for (int sh = 0; sh < 29; sh++)
        if (j.mask >> sh & 1) n += (128 / SHAPES[sh][0]) * (128 / SHAPES[sh][1]);
    j.npu = n;
    pthread_mutex_init(&j.lock, NULL);
    if (threads < 1) threads = 1;
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    double t0 = now_ms();
    for (int i = 0; i < threads; i++) pthread_create(&tid[i], NULL, frame_worker, &j);
    for (int i = 0; i < threads; i++) pthread_join(tid[i], NULL);
    double t = now_ms() - t0;
    free(tid);
    pthread_mutex_destroy(&j.lock);
    return t;
}

/* ---------------------------------------------------------------------- */
/* Lookahead                                                                */
/* ---------------------------------------------------------------------- */

/* loka_estimate_coding_cost (lookahead.c:82-290) of `cur` against pictures
 * made by refh_pic_create (l0[0..n0), l1[0..n1)), with an inter_search_t set
 * up by inter_search_init(.., is_padding) (inter.c:2015-2027) as the encoder
 * does (uavs3e.c:185, 1326); no icost / chroma / dqp outputs.  Returns the
 * picture cost: the sum over whole 16x16 units of the least inter or intra
 * com_had. */
RH_EXPORT double refh_loka_cost(const pel *cur, int w, int h, long long cur_poc, void **l0, int n0, void **l1,
                                int n1, int gop_size, int max_search_range, int is_padding)
{
    int err = 0;
    com_pic_t *org = com_pic_create(w, h, PIC_PAD_SIZE_L, PIC_PAD_SIZE_C, &err);
    for (int y = 0; y < h; y++) memcpy(org->y + y * org->stride_luma, cur + y * w, sizeof(pel) * w);
    org->img->ptr = cur_poc;
    com_img_t *ref_l0[4], *ref_l1[4];
    for (int i = 0; i < n0; i++) ref_l0[i] = ((refh_pic_t *)l0[i])->pic->img;
    for (int i = 0; i < n1; i++) ref_l1[i] = ((refh_pic_t *)l1[i])->pic->img;
    refh_search_t s;
    search_setup(&s, w, h, 0, 0, is_padding);
    s.pi.gop_size = gop_size;
    s.pi.max_search_range = max_search_range;
    int num_ref[2] = {n0, n1};
    double c = loka_estimate_coding_cost(&s.pi, org->img, ref_l0, ref_l1, num_ref, BIT_DEPTH, NULL, NULL, NULL);
    search_free(&s);
    com_pic_destroy(org);
    return c;
}

/* ---------------------------------------------------------------------- */
/* Affine motion compensation                                               */
/* ---------------------------------------------------------------------- */

/* com_mc_blk_affine_luma (com_mc.c:202-250) of a cu_w x cu_h CU at (x, y)
 * from the two control-point MVs (1/16 pel) into pred (stride cu_w). */
RH_EXPORT void refh_affine_luma(void *ref, int x, int y, int cu_w, int cu_h, const int *cpmv, int sub_w, int sub_h,
                                pel *pred)
{
    refh_pic_t *r = (refh_pic_t *)ref;
    CPMV mv[VER_NUM][MV_D];
    memset(mv, 0, sizeof(mv));
    mv[0][0] = cpmv[0]; mv[0][1] = cpmv[1];
    mv[1][0] = cpmv[2]; mv[1][1] = cpmv[3];
    ALIGNED_32(pel buf[MAX_CU_DIM]);
    com_mc_blk_affine_luma(x, y, r->w, r->h, cu_w, cu_h, mv, r->pic, buf, sub_w, sub_h, BIT_DEPTH);
    memcpy(pred, buf, sizeof(pel) * (size_t)cu_w * cu_h);
}

/* com_mc_cu_affine (com_mc.c:365-385) for list 0 alone, with two or three
 * control points (1/16 pel) and 4x4 or 8x8 sub-blocks
 * (affine_subblk_size_idx): luma into pred, chroma into pred_u / pred_v
 * (either may be NULL). */
RH_EXPORT void refh_affine_cu(void *ref, int x, int y, int cu_w, int cu_h, const int *cpmv, int vertex_num,
                              int sub8, pel *pred, pel *pred_u, pel *pred_v)
{
    refh_pic_t *r = (refh_pic_t *)ref;
    static com_ref_pic_t refp[MAX_REFS][REFP_NUM];
    memset(refp, 0, sizeof(refp));
    refp[0][REFP_0].pic = r->pic;
    s8 refi[REFP_NUM] = {0, REFI_INVALID};
    CPMV mv[REFP_NUM][VER_NUM][MV_D];
    memset(mv, 0, sizeof(mv));
    for (int i = 0; i < vertex_num; i++) {
        mv[REFP_0][i][MV_X] = cpmv[2 * i];
        mv[REFP_0][i][MV_Y] = cpmv[2 * i + 1];
    }
    com_pic_header_t sh;
    memset(&sh, 0, sizeof(sh));
    sh.affine_subblk_size_idx = sub8;
    static pel buf[N_C][MAX_CU_DIM];
    com_mc_cu_affine(x, y, r->w, r->h, cu_w, cu_h, refi, mv, refp, buf, vertex_num, &sh, BIT_DEPTH);
    memcpy(pred, buf[Y_C], sizeof(pel) * (size_t)cu_w * cu_h);
    if (pred_u) memcpy(pred_u, buf[U_C], sizeof(pel) * (size_t)(cu_w / 2) * (cu_h / 2));
    if (pred_v) memcpy(pred_v, buf[V_C], sizeof(pel) * (size_t)(cu_w / 2) * (cu_h / 2));
}

/* ---------------------------------------------------------------------- */
/* analyze_bi's search and analyze_smvd's search with a minimal core_t      */
/* ---------------------------------------------------------------------- */

/* The CU-level encoder state both searches read: the current picture, the
 * reference lists, the CU geometry and an inter_search_t set up as
 * analyze_inter_cu leaves it for one CU (inter_search_init, is_padding 1). */
typedef struct {
    core_t *core;
    com_info_t info;
    com_pic_t *org;
    com_ref_pic_t refp[MAX_REFS][REFP_NUM];
    int num_refp[REFP_NUM];
    u8 *tab;
} refh_cu_t;

static void cu_setup(refh_cu_t *c, const pel *cur, int pic_w, int pic_h, void **l0, int n0, void **l1, int n1,
                     int x, int y, int w, int h, int mvr, int cost_type, u32 lambda_mv)
{
    int err = 0;
    memset(c, 0, sizeof(*c));
    c->info.pic_width = pic_w;
    c->info.pic_height = pic_h;
    c->info.bit_depth_internal = BIT_DEPTH;
    c->info.gop_size = 16;
    c->info.i_scu = (pic_w + MIN_CU_SIZE - 1) >> MIN_CU_LOG2;
    /* analyze_bi returns right after its search when the search cost
     * exceeds inter_satd * satd_threshold (inter.c:1036-1038): with
     * inter_satd 0 that is always, so inter_rdcost is never reached */
    c->info.rmv_inter_candi_by_satd = 1;
    c->org = com_pic_create(pic_w, pic_h, PIC_PAD_SIZE_L, PIC_PAD_SIZE_C, &err);
    for (int i = 0; i < pic_h; i++) memcpy(c->org->y + i * c->org->stride_luma, cur + i * pic_w, sizeof(pel) * pic_w);
    for (int i = 0; i < n0; i++) {
        c->refp[i][REFP_0].pic = ((refh_pic_t *)l0[i])->pic;
        c->refp[i][REFP_0].ptr = c->refp[i][REFP_0].pic->ptr;
    }
    for (int i = 0; i < n1; i++) {
        c->refp[i][REFP_1].pic = ((refh_pic_t *)l1[i])->pic;
        c->refp[i][REFP_1].ptr = c->refp[i][REFP_1].pic->ptr;
    }
    c->num_refp[REFP_0] = n0;
    c->num_refp[REFP_1] = n1;
    /* core_t holds 32-byte aligned buffers (the prediction of com_mode_t)
     * that the SIMD kernels store to with aligned stores */
    void *mem = NULL;
    if (posix_memalign(&mem, 64, sizeof(core_t))) abort();
    memset(mem, 0, sizeof(core_t));
    core_t *core = c->core = (core_t *)mem;
    core->info = &c->info;
    core->pic_org = c->org;
    core->refp = c->refp;
    core->num_refp = c->num_refp;
    core->cu_pix_x = x;
    core->cu_pix_y = y;
    core->cu_width = w;
    core->cu_height = h;
    core->cu_width_log2 = CONV_LOG2(w);
    core->cu_height_log2 = CONV_LOG2(h);
    core->cu_scup_in_pic = (x >> MIN_CU_LOG2) + (y >> MIN_CU_LOG2) * c->info.i_scu;
    core->inter_satd = 0;
    core->satd_threshold = 1.0;
    core->mod_info_curr.pb_part = SIZE_2Nx2N;
    core->mod_info_curr.mvr_idx = (u8)mvr;
    inter_search_t *pi = &core->pinter;
    int off = inter_search_create(&c->tab, &c->info);
    c->info.tab_mvbits = c->tab;
    inter_search_init(pi, &c->info, 1);
    pi->tab_mvbits_offset = off;
    pi->lambda_mv = lambda_mv;
    pi->subpel_cost_type = cost_type;
    pi->curr_mvr = (u8)mvr;
}

static void cu_free(refh_cu_t *c)
{
    inter_search_free(c->tab, c->core->pinter.tab_mvbits_offset);
    free(c->core);
    com_pic_destroy(c->org);
}

/* analyze_bi (inter.c:972-1070) of the reference itself, up to the end of
 * its search: the CU at (x, y), w x h, of `cur` with the list-0 pictures
 * l0[0..n0) and list-1 pictures l1[0..n1) (refh_pic_create).  In:
 * mvp[l][r][2] and mv_scale[l][r][2] (pi->mvp_scale / mv_scale, quarter
 * pel, r < 4), refi[2] and mv[2][2] (analyze_bi's refi_L0L1 / mv_L0L1),
 * lidx_ref, mot_bits[2] (pi->mot_bits).  Out: refi, mv (cur_info->refi /
 * mv), mot_bits (pi->mot_bits) and mv_scale as the search leaves them. */
RH_EXPORT void refh_analyze_bi(const pel *cur, int pic_w, int pic_h, void **l0, int n0, void **l1, int n1, int x,
                               int y, int w, int h, const s16 *mvp, s16 *mv_scale, s8 *refi, s16 *mv, int lidx_ref,
                               int *mot_bits, int mvr, int cost_type, u32 lambda_mv)
{
    refh_cu_t c;
    cu_setup(&c, cur, pic_w, pic_h, l0, n0, l1, n1, x, y, w, h, mvr, cost_type, lambda_mv);
    inter_search_t *pi = &c.core->pinter;
    for (int l = 0; l < REFP_NUM; l++)
        for (int r = 0; r < 4; r++)
            for (int k = 0; k < MV_D; k++) {
                pi->mvp_scale[l][r][k] = mvp[(l * 4 + r) * 2 + k];
                pi->mv_scale[l][r][k] = mv_scale[(l * 4 + r) * 2 + k];
            }
    pi->mot_bits[0] = mot_bits[0];
    pi->mot_bits[1] = mot_bits[1];
    s16 mv_l0l1[REFP_NUM][MV_D] = {{mv[0], mv[1]}, {mv[2], mv[3]}};
    s8 refi_l0l1[REFP_NUM] = {refi[0], refi[1]};
    analyze_bi(c.core, NULL, mv_l0l1, refi_l0l1, (s8)lidx_ref);
    com_mode_t *ci = &c.core->mod_info_curr;
    refi[0] = ci->refi[REFP_0];
    refi[1] = ci->refi[REFP_1];
    for (int k = 0; k < 4; k++) mv[k] = ci->mv[k >> 1][k & 1];
    mot_bits[0] = pi->mot_bits[0];
    mot_bits[1] = pi->mot_bits[1];
    for (int l = 0; l < REFP_NUM; l++)
        for (int r = 0; r < 4; r++)
            for (int k = 0; k < MV_D; k++) mv_scale[(l * 4 + r) * 2 + k] = pi->mv_scale[l][r][k];
    cu_free(&c);
}

/* analyze_smvd's search (inter.c:1194-1255) for the CU at (x, y), w x h:
 * reference 0 of each list; the first two candidates (the MVP pair, and
 * list 0's uni MV with list 1 mirrored about its MVP), then the reference's
 * own smvd_refine twice (diamond, 8 rounds; cross, 1 round; offsets shifted
 * by the AMVR index) and, for cost type 1, the com_had of the winner, all
 * through the reference's com_mc_cu, get_smvd_cost and MV-bits functions in
 * analyze_smvd's order.  In: mvp[l][2], mv_scale_l0[2] (pi->mvp_scale[l][0],
 * pi->mv_scale[0][0]).  Out: mv[2][2]; returns mecost. */
RH_EXPORT u32 refh_smvd(const pel *cur, int pic_w, int pic_h, void *p0, void *p1, int x, int y, int w, int h,
                        const s16 *mvp, const s16 *mv_scale_l0, int mvr, int cost_type, u32 lambda_mv, s16 *mv_out)
{
    void *l0[1] = {p0}, *l1[1] = {p1};
    refh_cu_t c;
    cu_setup(&c, cur, pic_w, pic_h, l0, 1, l1, 1, x, y, w, h, mvr, cost_type, lambda_mv);
    core_t *core = c.core;
    inter_search_t *pi = &core->pinter;
    com_mode_t *ci = &core->mod_info_curr;
    const int log2w = core->cu_width_log2, log2h = core->cu_height_log2, widx = log2w - MIN_CU_LOG2;
    pel *org = c.org->y + x + y * c.org->stride_luma;
    init_inter_data(core);
    ci->smvd_flag = 1;
    ci->refi[REFP_0] = 0;
    ci->refi[REFP_1] = 0;
    s16 m[REFP_NUM][MV_D], p[REFP_NUM][MV_D], cand[REFP_NUM][MV_D], d[MV_D];
    for (int l = 0; l < REFP_NUM; l++) {
        p[l][MV_X] = m[l][MV_X] = mvp[2 * l];
        p[l][MV_Y] = m[l][MV_Y] = mvp[2 * l + 1];
    }
    /* candidate 1: both MVPs, MVD 0 */
    com_mc_cu(x, y, pic_w, pic_h, w, h, ci->refi, m, core->refp, ci->pred, w, CHANNEL_L, BIT_DEPTH);
    u32 best = get_smvd_cost(cost_type, w, h, widx, org, c.org->stride_luma, ci->pred[0], w, BIT_DEPTH) +
               MV_COST(get_mv_bits_with_mvr(pi, 0, 0, (u8)mvr));
    /* candidate 2: list 0's uni MV, list 1 mirrored about its MVP */
    cand[REFP_0][MV_X] = mv_scale_l0[0];
    cand[REFP_0][MV_Y] = mv_scale_l0[1];
    d[MV_X] = cand[REFP_0][MV_X] - p[REFP_0][MV_X];
    d[MV_Y] = cand[REFP_0][MV_Y] - p[REFP_0][MV_Y];
    cand[REFP_1][MV_X] = p[REFP_1][MV_X] - d[MV_X];
    cand[REFP_1][MV_Y] = p[REFP_1][MV_Y] - d[MV_Y];
    com_mc_cu(x, y, pic_w, pic_h, w, h, ci->refi, cand, core->refp, ci->pred, w, CHANNEL_L, BIT_DEPTH);
    u32 c2 = get_smvd_cost(cost_type, w, h, widx, org, c.org->stride_luma, ci->pred[0], w, BIT_DEPTH) +
             MV_COST(get_mv_bits_with_mvr(pi, d[MV_X], d[MV_Y], (u8)mvr));
    if (c2 < best) {
        best = c2;
        memcpy(m, cand, sizeof(m));
    }
    best = smvd_refine(core, x, y, log2w, log2h, m, p, ci->refi, 0, 1, best, 2, 8, ci->mvr_idx);
    best = smvd_refine(core, x, y, log2w, log2h, m, p, ci->refi, 0, 1, best, 0, 1, ci->mvr_idx);
    if (cost_type) {
        com_mc_cu(x, y, pic_w, pic_h, w, h, ci->refi, m, core->refp, ci->pred, w, CHANNEL_L, BIT_DEPTH);
        d[MV_X] = m[REFP_0][MV_X] - p[REFP_0][MV_X];
        d[MV_Y] = m[REFP_0][MV_Y] - p[REFP_0][MV_Y];
        best = com_had(w, h, org, c.org->stride_luma, ci->pred[0], w, BIT_DEPTH) +
               MV_COST(get_mv_bits_with_mvr(pi, d[MV_X], d[MV_Y], (u8)mvr));
    }
    for (int k = 0; k < 4; k++) mv_out[k] = m[k >> 1][k & 1];
    cu_free(&c);
    return best;
}

/* affine_me_gradient (inter.c:1339-1471) of the reference itself for the
 * cu_w x cu_h CU at (x, y) against `ref`: org (stride i_org) is the CU's
 * original (the current picture at the CU, or analyze_affine_bi's target
 * for bi), mvp / mv the two control-point MVs (1/16 pel) as CPMV[2][2];
 * mv receives the best control points.  refi / num_refp index
 * tbl_refi_bits, mvr is the affine AMVR index (0..2), bi selects the
 * bi-prediction form (5 iterations, distortion halved, mot_bits_other
 * added to the bits).  Returns the cost without its MV bits, as the
 * function does. */
RH_EXPORT u64 refh_affine_me_gradient(void *ref, const pel *org, int i_org, int x, int y, int cu_w, int cu_h,
                                      int sub_w, int sub_h, int refi, int num_refp, int mvr, int bi,
                                      int mot_bits_other, u32 lambda_mv, const int *mvp, int *mv)
{
    refh_pic_t *r = (refh_pic_t *)ref;
    refh_search_t s;
    search_setup(&s, r->w, r->h, lambda_mv, 0, 1);
    inter_search_t *pi = &s.pi;
    pi->ref_pic = r->pic;
    pi->org = (pel *)org;
    pi->i_org = i_org;
    pi->num_refp = (u8)num_refp;
    pi->curr_mvr = (u8)mvr;
    const int lidx = 0;
    pi->mot_bits[1 - lidx] = mot_bits_other;
    CPMV p[VER_NUM][MV_D], m[VER_NUM][MV_D];
    memset(p, 0, sizeof(p));
    memset(m, 0, sizeof(m));
    for (int v = 0; v < 2; v++) {
        p[v][MV_X] = mvp[2 * v]; p[v][MV_Y] = mvp[2 * v + 1];
        m[v][MV_X] = mv[2 * v];  m[v][MV_Y] = mv[2 * v + 1];
    }
    s8 ri = (s8)refi;
    u64 c = affine_me_gradient(pi, x, y, CONV_LOG2(cu_w), CONV_LOG2(cu_h), &ri, lidx, p, m, bi, sub_w, sub_h);
    for (int v = 0; v < 2; v++) {
        mv[2 * v] = m[v][MV_X];
        mv[2 * v + 1] = m[v][MV_Y];
    }
    search_free(&s);
    return c;
}
