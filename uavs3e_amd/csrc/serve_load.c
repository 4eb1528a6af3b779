/*
 * serve_load.c -- encoder-side load for the request server.
 *
 * Each stream is driven by two threads, as an encoder would be: an uploader
 * that makes the stream's pictures resident at most one frame ahead of the
 * frame being searched (me_server_put_picture; frames 0-2 before the clock), and a submitter that issues
 * the per-CTU ME requests of frame f against frame f-1 at precomputed
 * Poisson arrival times and releases frame f-1 once frame f is submitted.
 * A collector per stream retires the tickets.  A request's latency is its
 * completion time (taken by the server when its batch's results are on the
 * host) minus its scheduled arrival time, so a late submission counts
 * against the server, not for it.
 *
 * Frames are synthesised up front (me_synth_frame): stream i uses seed
 * seeds[i], global motion (+13, -6) quarter pel per frame, per-64x64 region
 * offsets in [-16, 16] quarter pel, noise sigma 3 (x4 at 10 bits).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/uavs3e_me_hip.h"

#define EXPORT __attribute__((visibility("default")))

void me_synth_frame(uint64_t seed, int stream, int frame, int w, int h, int bd, double sx, double sy, int region_qpel,
                    double sigma, int noise_purpose, double offset, void *out);

typedef struct {
    me_server_t srv;
    int sid, w, h, n_ctu, frames, npu;
    me_pel **pics;
    const double *arrive;          /* [frames-1][n_ctu] scheduled arrival, server clock */
    double *lat;                   /* [frames-1][n_ctu] completion - arrival, us */
    int64_t *tickets;
    me_pu_result_t *res;
    int8_t *best;
    pthread_mutex_t mu;
    pthread_cond_t cv;
    int resident;                  /* highest frame made resident */
    int searching;                 /* frame whose requests are being submitted */
    int submitted;                 /* requests submitted so far */
    int rc;
    pthread_barrier_t *go;         /* all threads running; then the schedule is final */
} stream_t;

/* Every load thread waits here twice: once when it runs, once more after
 * the main thread has fixed the clock, so that thread start-up (first-touch
 * of new stacks, symbol binding) never delays the first requests. */
static void wait_go(stream_t *s)
{
    pthread_barrier_wait(s->go);
    pthread_barrier_wait(s->go);
}

static void sleep_until(double t_us)
{
    for (;;) {
        double now = me_server_now_us(), d = t_us - now;
        if (d <= 0) return;
        if (d > 300) {
            struct timespec ts = {0, (long)((d - 150) * 1000)};
            nanosleep(&ts, NULL);
        }
    }
}

static void *uploader(void *arg)
{
    stream_t *s = (stream_t *)arg;
    wait_go(s);
    for (int f = s->resident + 1; f < s->frames; f++) {
        pthread_mutex_lock(&s->mu);
        while (f > s->searching + 1 && !s->rc) pthread_cond_wait(&s->cv, &s->mu);
        pthread_mutex_unlock(&s->mu);
        int rc = me_server_put_picture(s->srv, s->sid, f, s->pics[f], s->w, -1);
        pthread_mutex_lock(&s->mu);
        if (rc) s->rc = -1;
        s->resident = f;
        pthread_cond_broadcast(&s->cv);
        pthread_mutex_unlock(&s->mu);
        if (rc) return NULL;
    }
    return NULL;
}

static void *submitter(void *arg)
{
    stream_t *s = (stream_t *)arg;
    wait_go(s);
    for (int f = 1; f < s->frames; f++) {
        pthread_mutex_lock(&s->mu);
        s->searching = f;
        pthread_cond_broadcast(&s->cv);
        while (s->resident < f && !s->rc) pthread_cond_wait(&s->cv, &s->mu);
        int bad = s->rc;
        pthread_mutex_unlock(&s->mu);
        if (bad) break;
        const int64_t ref = f - 1;
        for (int c = 0; c < s->n_ctu; c++) {
            const int k = (f - 1) * s->n_ctu + c;
            sleep_until(s->arrive[k]);
            int64_t t = me_server_submit(s->srv, s->sid, f, c, &ref, s->res + (size_t)k * s->npu,
                                         s->best + (size_t)k * s->npu);
            pthread_mutex_lock(&s->mu);
            if (t <= 0) s->rc = -1;
            s->tickets[k] = t;
            s->submitted = k + 1;
            pthread_cond_broadcast(&s->cv);
            pthread_mutex_unlock(&s->mu);
            if (t <= 0) return NULL;
        }
        if (me_server_release_picture(s->srv, s->sid, f - 1)) { s->rc = -1; break; }
    }
    if (!s->rc) me_server_release_picture(s->srv, s->sid, s->frames - 1);
    pthread_mutex_lock(&s->mu);
    s->searching = s->frames;
    pthread_cond_broadcast(&s->cv);
    pthread_mutex_unlock(&s->mu);
    return NULL;
}

static void *collector(void *arg)
{
    stream_t *s = (stream_t *)arg;
    wait_go(s);
    const int n = (s->frames - 1) * s->n_ctu;
    for (int k = 0; k < n; k++) {
        pthread_mutex_lock(&s->mu);
        while (s->submitted <= k && !s->rc) pthread_cond_wait(&s->cv, &s->mu);
        int bad = s->rc && s->submitted <= k;
        pthread_mutex_unlock(&s->mu);
        if (bad) return NULL;
        double done = 0;
        if (me_server_wait(s->srv, s->tickets[k], -1, &done)) { s->rc = -1; return NULL; }
        s->lat[k] = done - s->arrive[k];
    }
    return NULL;
}

typedef struct {
    stream_t *S;
    const uint64_t *seeds;
    int nstreams, frames, bd;
    int next;                      /* next (stream, frame) task, under mu */
    pthread_mutex_t mu;
} gen_t;

static void *synth_worker(void *arg)
{
    gen_t *g = (gen_t *)arg;
    const double sigma = g->bd == 8 ? 3.0 : 12.0;
    for (;;) {
        pthread_mutex_lock(&g->mu);
        const int t = g->next++;
        pthread_mutex_unlock(&g->mu);
        if (t >= g->nstreams * g->frames) return NULL;
        const int i = t / g->frames, f = t % g->frames;
        stream_t *s = &g->S[i];
        me_synth_frame(g->seeds[i], i, f, s->w, s->h, g->bd, f * 13 / 4.0, f * -6 / 4.0, 16, sigma, 4, 0.0, s->pics[f]);
    }
}

/* Runs the serving load: nstreams streams of wh[2i] x wh[2i+1] pictures,
 * `frames` frames each, requests arriving as one Poisson process of
 * `rate` CTUs/s split over the streams in proportion to their CTUs per
 * frame (so all streams advance at the same frame rate).
 *   lat_us   out: per request latency, stream by stream, frame 1.. by CTU
 *   res,best out (optional): per request results, same order, npu PUs each
 *   info     out: [0] wall time from the first arrival to the last
 *            completion (us), [1] requests, [2] frame synthesis time (s)
 *   stats    out (optional): the server's statistics at the end
 * Returns 0, or -1 if any request failed. */
EXPORT int me_host_serve_load(const me_server_cfg_t *cfg, int nstreams, const int *wh, const uint64_t *seeds,
                              int frames, double rate, uint64_t arrival_seed, double *lat_us, me_pu_result_t *res_out,
                              int8_t *best_out, double *info, me_server_stats_t *stats)
{
    if (nstreams < 1 || frames < 2 || rate <= 0) return -1;
    me_server_t srv;
    if (me_server_create(&srv, cfg)) return -1;
    int npu = me_pus_per_ctu(cfg->shape_mask);
    const int bd = me_hip_bit_depth();
    stream_t *S = (stream_t *)calloc(nstreams, sizeof(stream_t));
    int total_ctu = 0, total_req = 0, rc = 0;
    for (int i = 0; i < nstreams; i++) {
        stream_t *s = &S[i];
        s->srv = srv;
        s->w = wh[2 * i];
        s->h = wh[2 * i + 1];
        s->n_ctu = ((s->w + 127) / 128) * ((s->h + 127) / 128);
        s->frames = frames;
        s->npu = npu;
        s->resident = -1;
        s->searching = 0;
        pthread_mutex_init(&s->mu, NULL);
        pthread_cond_init(&s->cv, NULL);
        if (me_server_open_stream(srv, s->w, s->h, &s->sid)) rc = -1;
        total_ctu += s->n_ctu;
        total_req += s->n_ctu * (frames - 1);
    }
    /* frames, synthesised in parallel (16 threads over (stream, frame) tasks) */
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < nstreams; i++) {
        S[i].pics = (me_pel **)calloc(frames, sizeof(me_pel *));
        for (int f = 0; f < frames; f++) S[i].pics[f] = (me_pel *)malloc((size_t)S[i].w * S[i].h * sizeof(me_pel));
    }
    {
        gen_t g;
        g.S = S;
        g.seeds = seeds;
        g.nstreams = nstreams;
        g.frames = frames;
        g.bd = bd;
        g.next = 0;
        pthread_mutex_init(&g.mu, NULL);
        pthread_t th[16];
        for (int t = 0; t < 16; t++) pthread_create(&th[t], NULL, synth_worker, &g);
        for (int t = 0; t < 16; t++) pthread_join(th[t], NULL);
        pthread_mutex_destroy(&g.mu);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (info) info[2] = (t1.tv_sec - t0.tv_sec) + (t1.tv_nsec - t0.tv_nsec) * 1e-9;

    /* the first three pictures of every stream are resident before the clock
     * starts: an encoder has two before its first ME request, and the third
     * completes the steady-state set (reference, current, next) that the
     * server's picture pool then recycles, so no picture allocation falls
     * inside the measured run */
    for (int i = 0; i < nstreams && !rc; i++)
        for (int f = 0; f < 3 && f < frames && !rc; f++) {
            if (me_server_put_picture(srv, S[i].sid, f, S[i].pics[f], S[i].w, -1)) rc = -1;
            S[i].resident = f;
        }
    /* arrival schedule: per stream a Poisson process at the stream's share
     * of the rate (its CTUs per frame over all streams'); the superposition
     * is a Poisson process at the total rate.  Offsets from the start, which
     * is fixed once every load thread runs. */
    double start = 0.0;
    uint64_t x = arrival_seed * 0x9E3779B97F4A7C15ull + 1;
    double **arr = (double **)calloc(nstreams, sizeof(double *));
    for (int i = 0; i < nstreams; i++) {
        const int n = S[i].n_ctu * (frames - 1);
        arr[i] = (double *)malloc(n * sizeof(double));
        const double rate_i = rate * S[i].n_ctu / total_ctu;
        double t = 0.0;
        for (int k = 0; k < n; k++) {
            x ^= x << 13; x ^= x >> 7; x ^= x << 17;
            const double u = ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            t += -log(u) / rate_i * 1e6;
            arr[i][k] = t;
        }
        S[i].arrive = arr[i];
    }
    size_t off = 0;
    for (int i = 0; i < nstreams; i++) {
        const int n = S[i].n_ctu * (frames - 1);
        S[i].lat = lat_us + off;
        S[i].tickets = (int64_t *)calloc(n, sizeof(int64_t));
        S[i].res = res_out ? res_out + off * npu : (me_pu_result_t *)malloc((size_t)n * npu * sizeof(me_pu_result_t));
        S[i].best = best_out ? best_out + off * npu : (int8_t *)malloc((size_t)n * npu);
        off += n;
    }
    pthread_t *th = (pthread_t *)calloc(3 * nstreams, sizeof(pthread_t));
    if (!rc) {
        pthread_barrier_t go;
        pthread_barrier_init(&go, NULL, 3 * nstreams + 1);
        for (int i = 0; i < nstreams; i++) {
            S[i].go = &go;
            pthread_create(&th[3 * i], NULL, uploader, &S[i]);
            pthread_create(&th[3 * i + 1], NULL, submitter, &S[i]);
            pthread_create(&th[3 * i + 2], NULL, collector, &S[i]);
        }
        pthread_barrier_wait(&go);
        start = me_server_now_us() + 2000.0;
        for (int i = 0; i < nstreams; i++) {
            const int n = S[i].n_ctu * (frames - 1);
            for (int k = 0; k < n; k++) arr[i][k] += start;
        }
        pthread_barrier_wait(&go);
        for (int i = 0; i < 3 * nstreams; i++) pthread_join(th[i], NULL);
        pthread_barrier_destroy(&go);
    }
    double last = start;
    for (int i = 0; i < nstreams; i++) {
        if (S[i].rc) rc = -1;
        const int n = S[i].n_ctu * (frames - 1);
        for (int k = 0; k < n; k++)
            if (S[i].arrive[k] + S[i].lat[k] > last) last = S[i].arrive[k] + S[i].lat[k];
    }
    if (info) {
        info[0] = last - start;
        info[1] = total_req;
    }
    if (stats) me_server_stats(srv, stats);
    me_server_destroy(srv);
    for (int i = 0; i < nstreams; i++) {
        for (int f = 0; f < frames; f++) free(S[i].pics[f]);
        free(S[i].pics);
        free(arr[i]);
        free(S[i].tickets);
        if (!res_out) free(S[i].res);
        if (!best_out) free(S[i].best);
        pthread_mutex_destroy(&S[i].mu);
        pthread_cond_destroy(&S[i].cv);
    }
    free(arr);
    free(th);
    free(S);
    return rc;
}
