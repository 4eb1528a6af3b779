/*
 * me_host.c -- host-side ME driver that reaches every block cost and filter
 * only through a dispatch table (me_dispatch_t, the hot-path slots of
 * uavs3e's funs_handle_t).  With the table filled by me_funs_init_hip() this
 * is the "per-call HIP drop-in" path: the search control runs on the host,
 * like uavs3e's me.c, and every SAD/SATD/filter is one synchronous GPU call.
 *
 * It plays the role of the reference's callers on a machine where the
 * reference itself is not present:
 *   me_host_if_luma_frame  = com_if_luma_frame (com_mc.c:760-800)
 *   me_host_full_search    = search_raster's step-1 loop (me.c:150-190)
 *   me_host_subpel         = me_sub_pel_search, cost_type 0 (me.c:366-500)
 *   me_host_had            = com_had (com_cost.c:712-812)
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "uavs3e_me_hip.h"

typedef me_pel pel;
typedef int8_t s8;
typedef int16_t s16;
typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;

#define EXPORT __attribute__((visibility("default")))
#define CLIP3(lo, hi, v) ((v) < (lo) ? (lo) : ((v) > (hi) ? (hi) : (v)))

static const s8 LUMA_TAPS[4][8] = {
    {0, 0, 0, 64, 0, 0, 0, 0},
    {-1, 4, -10, 57, 19, -7, 3, -1},
    {-1, 4, -11, 40, 40, -11, 4, -1},
    {-1, 3, -7, 19, 57, -10, 4, -1}};

static const int SHAPES[29][2] = {
    {4, 8}, {8, 4}, {4, 16}, {16, 4}, {4, 32}, {32, 4}, {8, 8}, {8, 16}, {16, 8}, {8, 32},
    {32, 8}, {8, 64}, {64, 8}, {16, 16}, {16, 32}, {32, 16}, {16, 64}, {64, 16}, {16, 128},
    {128, 16}, {32, 32}, {32, 64}, {64, 32}, {32, 128}, {128, 32}, {64, 64}, {64, 128},
    {128, 64}, {128, 128}};

static int wlog(int w) { int l = 0; while ((4 << l) < w) l++; return l; }

/* picture_padding (com_util.c:105-147) of one plane from its
 * [-ext, w+ext) x [-ext, h+ext) band. */
EXPORT void me_host_pad_plane(pel *p, int stride, int w, int h, int pad, int ext)
{
    int e = pad - ext;
    pel *a = p - ext * stride - ext;
    int bw = w + 2 * ext, bh = h + 2 * ext;
    for (int y = 0; y < bh; y++) {
        pel *row = a + y * stride;
        for (int j = 1; j <= e; j++) { row[-j] = row[0]; row[bw - 1 + j] = row[bw - 1]; }
    }
    for (int j = 1; j <= e; j++) {
        memcpy(a - j * stride - e, a - e, sizeof(pel) * (size_t)stride);
        memcpy(a + (bh - 1 + j) * stride - e, a + (bh - 1) * stride - e, sizeof(pel) * (size_t)stride);
    }
}

/* com_if_luma_frame (com_mc.c:760-800) through the table's frame filters.
 * planes[dy*4+dx] point at (0,0) of padded planes; plane 0 must hold the
 * picture (it is edge-padded here). */
EXPORT int me_host_if_luma_frame(const me_dispatch_t *t, pel *const planes[16], int stride, int w, int h, int pad, int bd)
{
    me_host_pad_plane(planes[0], stride, w, h, pad, 0);
    int ts = w + 16;
    s16 *tmp = (s16 *)malloc(sizeof(s16) * 3 * (size_t)ts * (h + 16));
    if (!tmp) return -1;
    s16 *tb[3] = {tmp, tmp + (size_t)ts * (h + 16), tmp + 2 * (size_t)ts * (h + 16)};
    s8 coefs[3][8];
    memcpy(coefs, LUMA_TAPS[1], sizeof(coefs));
    pel *d[3];
    for (int i = 0; i < 3; i++) d[i] = planes[i + 1] - 8 * stride - 8;
    t->ip_flt_y_hor(planes[0] - 8 * stride - 8, stride, d, stride, tb, ts, w + 16, h + 16, coefs, bd);
    for (int i = 0; i < 3; i++) d[i] = planes[(i + 1) * 4] - 4 * stride - 4;
    t->ip_flt_y_ver(planes[0] - 4 * stride - 4, stride, d, stride, w + 8, h + 8, coefs, bd);
    for (int dx = 1; dx < 4; dx++) {
        for (int i = 0; i < 3; i++) d[i] = planes[(i + 1) * 4 + dx] - 4 * stride - 4;
        t->ip_flt_y_ver_ext(tb[dx - 1] + 4 * ts + 4, ts, d, stride, w + 8, h + 8, coefs, bd);
    }
    free(tmp);
    for (int i = 1; i < 16; i++) me_host_pad_plane(planes[i], stride, w, h, pad, 4);
    return 0;
}

/* com_had (com_cost.c:712-812) through cost_satd. */
EXPORT u32 me_host_had(const me_dispatch_t *t, int w, int h, pel *org, int s_org, pel *cur, int s_cur, int bd)
{
    int tw, th, wl, hl;
    if (w == h) { tw = th = (w == 4 ? 4 : 8); }
    else if (w > h) { if (h == 4) { tw = 8; th = 4; } else { tw = 16; th = 8; } }
    else { if (w == 4) { tw = 4; th = 8; } else { tw = 8; th = 16; } }
    wl = wlog(tw); hl = wlog(th);
    int sum = 0;
    for (int y = 0; y < h; y += th)
        for (int x = 0; x < w; x += tw)
            sum += (int)t->cost_satd[wl][hl](org + y * s_org + x, s_org, cur + y * s_cur + x, s_cur);
    return (u32)(sum >> (bd - 8));
}

typedef struct {
    const me_dispatch_t *t;
    int w, h, bd;
    const u8 *tab;      /* centred MV-bits table */
    int T;
    u32 lambda;
} host_ctx_t;

/* Integer full search (MVP 0, AMVR 0): window [-R,R]^2 clipped as in
 * get_max_search_range (me.c:502-527, is_padding = 1); per row the
 * SAD x4 / x3 / x1 grouping of search_raster (me.c:162-188); first strict
 * minimum in raster order. */
static u64 full_search(const host_ctx_t *c, pel *org, int i_org, pel *ref0, int i_ref, int x, int y, int pw, int ph,
                       int R, s16 mv[2])
{
    int lox = -127 - x > -c->T + 4 ? -127 - x : -c->T + 4, hix = c->w - 1 - x < c->T - 4 ? c->w - 1 - x : c->T - 4;
    int loy = -127 - y > -c->T + 4 ? -127 - y : -c->T + 4, hiy = c->h - 1 - y < c->T - 4 ? c->h - 1 - y : c->T - 4;
    mv[0] = mv[1] = 0;
    if (lox > hix || loy > hiy) return UINT64_MAX;
    int cx = CLIP3(lox, hix, 0), cy = CLIP3(loy, hiy, 0);
    int x0 = CLIP3(lox, hix, cx - R), x1 = CLIP3(lox, hix, cx + R);
    int y0 = CLIP3(loy, hiy, cy - R), y1 = CLIP3(loy, hiy, cy + R);
    int widx = wlog(pw), shift = 16 - (c->bd - 8);
    u64 best = UINT64_MAX;
    for (int my = y0; my <= y1; my++) {
        pel *p = ref0 + (y + my) * i_ref + x;
        u64 cmy = (u64)(c->tab[my << 2] * c->lambda);
        int mx = x0;
        u32 sad[4];
        for (; mx + 3 <= x1; mx += 4) {
            c->t->cost_sad_x4[widx](org, i_org, p + mx, p + mx + 1, p + mx + 2, p + mx + 3, i_ref, sad, ph);
            for (int k = 0; k < 4; k++) {
                u64 v = cmy + (u64)(c->tab[(mx + k) << 2] * c->lambda) + ((u64)sad[k] << shift);
                if (v < best) { best = v; mv[0] = (s16)(mx + k); mv[1] = (s16)my; }
            }
        }
        if (mx + 2 <= x1) {
            c->t->cost_sad_x3[widx](org, i_org, p + mx, p + mx + 1, p + mx + 2, i_ref, sad, ph);
            for (int k = 0; k < 3; k++) {
                u64 v = cmy + (u64)(c->tab[(mx + k) << 2] * c->lambda) + ((u64)sad[k] << shift);
                if (v < best) { best = v; mv[0] = (s16)(mx + k); mv[1] = (s16)my; }
            }
            mx += 3;
        }
        for (; mx <= x1; mx++) {
            u64 v = cmy + (u64)(c->tab[mx << 2] * c->lambda) +
                    ((u64)c->t->cost_sad[widx](org, i_org, p + mx, i_ref, ph) << shift);
            if (v < best) { best = v; mv[0] = (s16)mx; mv[1] = (s16)my; }
        }
    }
    return best;
}

/* me_sub_pel_search (me.c:366-500), cost_type 0, bi 0, AMVR 0, MVP 0. */
static u64 subpel(const host_ctx_t *c, pel *org, int i_org, pel *const *planes, int i_ref, int x, int y, int pw, int ph,
                  s16 mv[2])
{
    static const int HP[8][2] = {{-2, 0}, {-2, 2}, {0, 2}, {2, 2}, {2, 0}, {2, -2}, {0, -2}, {-2, -2}};
    static const int QP[8][2] = {{-1, 0}, {0, 1}, {1, 0}, {0, -1}, {-1, 1}, {1, 1}, {-1, -1}, {1, -1}};
    int cx = mv[0], cy = mv[1];
    /* centre: the integer plane, not clipped (me.c:400) */
    u32 d = me_host_had(c->t, pw, ph, org, i_org, planes[0] + (y + (cy >> 2)) * i_ref + x + (cx >> 2), i_ref, c->bd);
    u64 best = (u64)(c->lambda * (u32)(c->tab[cx] + c->tab[cy])) + ((u64)d << 16);
    for (int pass = 0; pass < 2; pass++) {
        const int (*pat)[2] = pass ? QP : HP;
        int bx = mv[0], by = mv[1];
        for (int i = 0; i < 8; i++) {
            int mx = bx + pat[i][0], my = by + pat[i][1];
            int qx = mx + (x << 2), qy = my + (y << 2);
            int ix = CLIP3(-132, c->w + 4, qx >> 2), iy = CLIP3(-132, c->h + 4, qy >> 2);
            pel *q = planes[(qy & 3) * 4 + (qx & 3)] + iy * i_ref + ix;
            d = me_host_had(c->t, pw, ph, org, i_org, q, i_ref, c->bd);
            u64 v = (u64)(c->lambda * (u32)(c->tab[mx] + c->tab[my])) + ((u64)d << 16);
            if (v < best) { best = v; mv[0] = (s16)mx; mv[1] = (s16)my; }
        }
    }
    return best;
}

typedef struct {
    host_ctx_t c;
    pel *cur; int i_cur;
    int nref; pel *const *planes; int i_ref;
    int R; u32 mask; int npu;
    int ctu0, ctu1, next;
    me_pu_result_t *res; int8_t *best;
    double *ctu_ms;
    pthread_mutex_t mu;
} frame_job_t;

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void run_ctu(frame_job_t *j, int ctu)
{
    const host_ctx_t *c = &j->c;
    int cols = (c->w + 127) / 128;
    int cx = (ctu % cols) * 128, cy = (ctu / cols) * 128, pu = 0;
    for (int s = 0; s < 29; s++) {
        if (!(j->mask >> s & 1)) continue;
        int pw = SHAPES[s][0], ph = SHAPES[s][1];
        for (int oy = 0; oy < 128; oy += ph)
            for (int ox = 0; ox < 128; ox += pw, pu++) {
                int x = cx + ox, y = cy + oy;
                size_t base = (size_t)ctu * j->npu + pu;
                int inside = x + pw <= c->w && y + ph <= c->h;
                u32 bc = 0xFFFFFFFFu;
                int br = -1;
                for (int r = 0; r < j->nref; r++) {
                    me_pu_result_t *o = &j->res[base * j->nref + r];
                    memset(o, 0, sizeof(*o));
                    o->cost_int = o->cost_sub = UINT64_MAX;
                    o->cost_final = 0xFFFFFFFFu;
                    if (!inside) continue;
                    pel *const *pl = j->planes + 16 * r;
                    pel *org = j->cur + y * j->i_cur + x;
                    s16 mv[2];
                    u64 ci = full_search(c, org, j->i_cur, pl[0], j->i_ref, x, y, pw, ph, j->R, mv);
                    o->mv_int[0] = mv[0]; o->mv_int[1] = mv[1];
                    o->cost_int = ci;
                    if (ci == UINT64_MAX) continue;
                    mv[0] = (s16)(mv[0] * 4); mv[1] = (s16)(mv[1] * 4);
                    u64 cs = subpel(c, org, j->i_cur, pl, j->i_ref, x, y, pw, ph, mv);
                    o->mv_sub[0] = mv[0]; o->mv_sub[1] = mv[1];
                    o->cost_sub = cs;
                    int rb = j->nref <= 1 ? 0 : (r < j->nref - 1 ? r + 1 : j->nref - 1);
                    u64 cf = cs + (u64)(c->lambda * (u32)(rb + 3));
                    o->cost_final = (u32)((cf + (1u << 15)) >> 16);
                    o->valid = 1;
                    if (o->cost_final < bc) { bc = o->cost_final; br = r; }
                }
                j->best[base] = (int8_t)br;
            }
    }
}

static void *worker(void *arg)
{
    frame_job_t *j = (frame_job_t *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int ctu = j->ctu0 + j->next++;
        pthread_mutex_unlock(&j->mu);
        if (ctu >= j->ctu1) break;
        double t0 = now_ms();
        run_ctu(j, ctu);
        if (j->ctu_ms) j->ctu_ms[ctu - j->ctu0] = now_ms() - t0;
    }
    return NULL;
}

/* Full search + sub-pel for CTUs [ctu0, ctu1) on `threads` host threads, all
 * costs through `t`.  Layout of res/best as in me_plan_results.  Returns the
 * wall time in ms. */
EXPORT double me_host_frame(const me_dispatch_t *t, int w, int h, int bd, pel *cur, int i_cur, int nref,
                            pel *const *ref_planes, int i_ref, int R, u32 lambda, u32 mask, int ctu0, int ctu1,
                            int threads, me_pu_result_t *res, int8_t *best, double *ctu_ms)
{
    frame_job_t j;
    memset(&j, 0, sizeof(j));
    int off = ((w > h ? w : h) << 2) + 3;
    u8 *tab = (u8 *)malloc((size_t)(2 * off + 1));
    tab[off] = 1;
    tab[0] = tab[2 * off] = 0;
    for (int v = 1; v < off; v++) {
        int l = 0;
        while ((2 << l) <= v + 1) l++;
        tab[off + v] = tab[off - v] = (u8)(2 * (l + 1));
    }
    j.c.t = t; j.c.w = w; j.c.h = h; j.c.bd = bd; j.c.tab = tab + off; j.c.T = w > h ? w : h; j.c.lambda = lambda;
    j.cur = cur; j.i_cur = i_cur; j.nref = nref; j.planes = ref_planes; j.i_ref = i_ref;
    j.R = R; j.mask = mask; j.ctu0 = ctu0; j.ctu1 = ctu1; j.res = res; j.best = best; j.ctu_ms = ctu_ms;
    for (int s = 0; s < 29; s++)
        if (mask >> s & 1) j.npu += (128 / SHAPES[s][0]) * (128 / SHAPES[s][1]);
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    double t0 = now_ms();
    for (int i = 0; i < threads; i++) pthread_create(&tid[i], NULL, worker, &j);
    for (int i = 0; i < threads; i++) pthread_join(tid[i], NULL);
    double ms = now_ms() - t0;
    free(tid);
    free(tab);
    pthread_mutex_destroy(&j.mu);
    return ms;
}
