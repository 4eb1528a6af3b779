// Per-call drop-in entries: the hot-path slots of uavs3e's funs_handle_t
// (inc/com_util.h:300-330) as synchronous HIP calls on host buffers.
//
// Each entry stages exactly the samples the reference kernel reads (block
// plus filter halo) into a per-thread pinned buffer, copies them to the GPU
// in one transfer, runs one kernel, and copies the result back before
// returning -- the reference slots are synchronous and keep no pointer.
// Encoder threads call the table concurrently (uavs3e.c:1060, 1501), so all
// state is per thread: a HIP stream, a pinned staging buffer and a device
// buffer.
#include "common.h"
#include "satd.h"
#include "../../include/uavs3e_me_hip.h"

#include <atomic>
#include <mutex>
#include <string.h>
#include <vector>

// ---------------------------------------------------------------------------
// Library state
// ---------------------------------------------------------------------------

static thread_local char t_err[512];
static std::atomic<uint64_t> g_errors{0};
static std::atomic<int> g_fallback{0};
static std::mutex g_init_mu;
static int g_init_rc = 1;   // 1 = not tried
static int g_device = 0;
// Slots replaced by me_funs_init_hip (the fallback of a failing entry).  The
// reference has one process-wide table (com_util.c:40); a slot that already
// holds a HIP entry when the table is initialised again keeps the previous
// non-HIP slot recorded here, so a fallback never calls back into HIP.
static me_dispatch_t g_prev;

void me_set_error(const char *what, hipError_t e, const char *file, int line)
{
    snprintf(t_err, sizeof(t_err), "%s: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

ME_EXPORT const char *me_hip_last_error(void) { return t_err; }
int me_hip_device(void) { return g_device; }
ME_EXPORT int me_hip_bit_depth(void) { return BIT_DEPTH; }
ME_EXPORT uint64_t me_hip_error_count(void) { return g_errors.load(); }
ME_EXPORT void me_hip_set_fallback(int enable) { g_fallback.store(enable ? 1 : 0); }

__global__ void k_probe(int *p) { if (p) *p = BIT_DEPTH; }

ME_EXPORT int me_hip_init(int device)
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_init_rc == 0) return 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        me_set_error("me_hip_init: no HIP device", e == hipSuccess ? hipErrorNoDevice : e, __FILE__, __LINE__);
        g_init_rc = -1;
        return -1;
    }
    g_device = device < 0 ? 0 : device;
    e = hipSetDevice(g_device);
    if (e == hipSuccess) {
        hipFuncAttributes attr;
        e = hipFuncGetAttributes(&attr, (const void *)k_probe);
    }
    if (e != hipSuccess) {
        me_set_error("me_hip_init: device unusable or no gfx950 code object", e, __FILE__, __LINE__);
        g_init_rc = -1;
        return -1;
    }
    g_init_rc = 0;
    return 0;
}

// ---------------------------------------------------------------------------
// Per-thread staging context
// ---------------------------------------------------------------------------

// A thread's stream and staging buffers.  Contexts are never freed: a thread
// that ends hands its context back to a pool (no HIP call, since thread-exit
// destructors can run after the HIP runtime has been torn down at process
// exit), and the next thread that calls an entry takes it over.  The pool
// holds at most as many contexts as threads ever called entries at once.
struct CallCtx {
    hipStream_t st = nullptr;
    unsigned char *h = nullptr, *d = nullptr;
    size_t cap = 0;
    bool ok = false;
};
static std::mutex g_ctx_mu;
static std::vector<CallCtx *> g_ctx_pool;

struct CtxSlot {
    CallCtx *c = nullptr;
    CallCtx &get()
    {
        if (!c) {
            std::lock_guard<std::mutex> lk(g_ctx_mu);
            if (!g_ctx_pool.empty()) { c = g_ctx_pool.back(); g_ctx_pool.pop_back(); }
            else c = new CallCtx();
        }
        return *c;
    }
    ~CtxSlot()
    {
        if (!c) return;
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        g_ctx_pool.push_back(c);
    }
};
static thread_local CtxSlot t_slot;
static CallCtx &ctx() { return t_slot.get(); }

// Ensures this thread has a stream and at least `bytes` of pinned + device
// staging.  Returns 0 on success.
static int ctx_reserve(size_t bytes)
{
    CallCtx &c = ctx();
    if (!c.ok) {
        if (me_hip_init(-1)) return -1;
        ME_CHECK(hipSetDevice(g_device));
        ME_CHECK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
        c.ok = true;
    }
    if (bytes > c.cap) {
        size_t n = c.cap ? c.cap : (1 << 20);
        while (n < bytes) n <<= 1;
        if (c.h) { (void)hipHostFree(c.h); c.h = nullptr; }
        if (c.d) { (void)hipFree(c.d); c.d = nullptr; }
        c.cap = 0;
        ME_CHECK(hipHostMalloc((void **)&c.h, n, hipHostMallocDefault));
        ME_CHECK(hipMalloc((void **)&c.d, n));
        c.cap = n;
    }
    return 0;
}

// A staging layout: regions are appended at 256-byte aligned offsets.
struct Stage {
    size_t used = 0;
    size_t add(size_t bytes)
    {
        size_t off = used;
        used = (used + bytes + 255) & ~(size_t)255;
        return off;
    }
};

// Copy `rows` rows of `cols` elements (element size es) starting at src with
// a stride of `stride` elements into dst, packed.
static void gather(void *dst, const void *src, int rows, int cols, int stride, int es)
{
    const unsigned char *s = (const unsigned char *)src;
    unsigned char *d = (unsigned char *)dst;
    for (int y = 0; y < rows; y++) memcpy(d + (size_t)y * cols * es, s + (ptrdiff_t)y * stride * es, (size_t)cols * es);
}
static void scatter(void *dst, const void *src, int rows, int cols, int stride, int es)
{
    const unsigned char *s = (const unsigned char *)src;
    unsigned char *d = (unsigned char *)dst;
    for (int y = 0; y < rows; y++) memcpy(d + (ptrdiff_t)y * stride * es, s + (size_t)y * cols * es, (size_t)cols * es);
}

static int ctx_upload(size_t bytes)
{
    ME_CHECK(hipMemcpyAsync(ctx().d, ctx().h, bytes, hipMemcpyHostToDevice, ctx().st));
    return 0;
}
static int ctx_download(size_t off, size_t bytes)
{
    ME_CHECK(hipMemcpyAsync(ctx().h + off, ctx().d + off, bytes, hipMemcpyDeviceToHost, ctx().st));
    ME_CHECK(hipStreamSynchronize(ctx().st));
    return 0;
}

// Failure policy of the entries (they have no error channel): abort with the
// HIP error, or -- when enabled -- run the replaced slot and count.
static bool entry_failed(const char *entry, bool has_prev)
{
    g_errors.fetch_add(1);
    if (g_fallback.load() && has_prev) return true;
    fprintf(stderr, "uavs3e_me_hip: %s failed: %s\n", entry, t_err);
    abort();
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// SAD of n predictions against one original, packed at stride w.
__global__ __launch_bounds__(256) void k_sad_n(const pel *org, const pel *preds, size_t pred_step, int w, int h,
                                               int n, u32 *out)
{
    __shared__ u32 part[4][4];
    const int tid = threadIdx.x;
    u32 acc[4] = {0, 0, 0, 0};
    for (int i = tid; i < w * h; i += 256) {
        const int o = org[i];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < n) acc[k] += (u32)abs(o - (int)preds[k * pred_step + i]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        u32 v = acc[k];
        for (int m = 32; m; m >>= 1) v += (u32)__shfl_xor((int)v, m, 64);
        if ((tid & 63) == 0) part[tid >> 6][k] = v;
    }
    __syncthreads();
    if (tid < n) out[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

// One SATD tile.
__global__ void k_satd(int tt, const pel *org, int s_org, const pel *cur, int s_cur, u32 *out)
{
    if (threadIdx.x == 0) out[0] = (u32)satd_any(tt, org, s_org, cur, s_cur);
}

// com_had: one thread per tile, block sum, then >> (bd - 8).
__global__ __launch_bounds__(256) void k_had(int w, int h, const pel *org, int s_org, const pel *cur, int s_cur,
                                             int bd, u32 *out)
{
    __shared__ int tot;
    int tt, tw, th;
    had_tiling(w, h, &tt, &tw, &th);
    const int tpr = w / tw, n = tpr * (h / th);
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int x = (t % tpr) * tw, y = (t / tpr) * th;
        atomicAdd(&tot, satd_any(tt, org + y * s_org + x, s_org, cur + y * s_cur + x, s_cur));
    }
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (u32)(tot >> (bd - 8));
}

struct Taps8 { s8 c[8]; };

// ipflt: src points at (0,0) of a staged region whose row stride is s_src;
// n-tap, vertical or horizontal, clip((sum + 32) >> 6).
__global__ void k_ipflt(const pel *src, int s_src, pel *dst, int w, int h, int n, int vert, Taps8 cf, int max_val)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const pel *s = src + (ptrdiff_t)y * s_src + x;
    const int step = vert ? s_src : 1, c0 = n / 2 - 1;
    int t = 0;
    for (int k = 0; k < n; k++) t += (int)s[(k - c0) * step] * cf.c[k];
    dst[y * w + x] = (pel)me_clip3(0, max_val, (t + 32) >> 6);
}

// ipflt_ext: horizontal pass stored as s16 ((t + add1) >> shift1), then
// vertical ((t + add2) >> shift2); 8-bit is recognised by max_val == 255
// (com_mc.c:503, 597).
__global__ void k_ipflt_ext(const pel *src, int s_src, pel *dst, int w, int h, int n, Taps8 cx, Taps8 cy, int max_val)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const int shift1 = max_val == 255 ? 0 : 2, shift2 = max_val == 255 ? 12 : 10;
    const int add1 = (1 << shift1) >> 1, add2 = 1 << (shift2 - 1);
    const int c0 = n / 2 - 1;
    int t = 0;
    for (int j = 0; j < n; j++) {
        const pel *s = src + (ptrdiff_t)(y + j - c0) * s_src + x;
        int u = 0;
        for (int k = 0; k < n; k++) u += (int)s[k - c0] * cx.c[k];
        t += (int)(s16)((u + add1) >> shift1) * cy.c[j];
    }
    dst[y * w + x] = (pel)me_clip3(0, max_val, (t + add2) >> shift2);
}

__global__ void k_avrg(const pel *a, const pel *b, pel *dst, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (pel)(((int)a[i] + (int)b[i] + 1) >> 1);
}

// Frame filters of com_if_luma_frame (com_mc.c:668-758) on staged regions.
struct Coef3 { s8 c[3][8]; };

__global__ void k_flt_hor(const pel *src, int s_src, pel *d, s16 *t, int w, int h, Coef3 cf, int bd)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, p = blockIdx.z;
    if (x >= w || y >= h) return;
    const pel *s = src + (ptrdiff_t)y * s_src + x;
    int v = 0;
    for (int k = 0; k < 8; k++) v += (int)s[k - 3] * cf.c[p][k];
    const int sh = bd - 8, add = (1 << sh) >> 1, maxv = (1 << bd) - 1;
    size_t o = (size_t)p * w * h + (size_t)y * w + x;
    t[o] = (s16)((v + add) >> sh);
    d[o] = (pel)me_clip3(0, maxv, (v + 32) >> 6);
}

__global__ void k_flt_ver(const pel *src, int s_src, pel *d, int w, int h, Coef3 cf, int bd)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, p = blockIdx.z;
    if (x >= w || y >= h) return;
    const pel *s = src + (ptrdiff_t)y * s_src + x;
    int v = 0;
    for (int k = 0; k < 8; k++) v += (int)s[(ptrdiff_t)(k - 3) * s_src] * cf.c[p][k];
    d[(size_t)p * w * h + (size_t)y * w + x] = (pel)me_clip3(0, (1 << bd) - 1, (v + 32) >> 6);
}

__global__ void k_flt_ver_ext(const s16 *src, int s_src, pel *d, int w, int h, Coef3 cf, int bd)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, p = blockIdx.z;
    if (x >= w || y >= h) return;
    const s16 *s = src + (ptrdiff_t)y * s_src + x;
    int v = 0;
    for (int k = 0; k < 8; k++) v += (int)s[(ptrdiff_t)(k - 3) * s_src] * cf.c[p][k];
    const int sh = 20 - bd, add = 1 << (sh - 1);
    d[(size_t)p * w * h + (size_t)y * w + x] = (pel)me_clip3(0, (1 << bd) - 1, (v + add) >> sh);
}

// ---------------------------------------------------------------------------
// Entries
// ---------------------------------------------------------------------------

static int log2w(int w) { int l = 0; while ((4 << l) < w) l++; return l; }

static int sad_n(int w, const pel *org, int i_org, pel *const *p, int n, int i_pred, int h, u32 *out)
{
    const size_t blk = (size_t)w * h * sizeof(pel);
    Stage sg;
    size_t o_org = sg.add(blk);
    size_t o_pred = sg.add(blk * 4);
    size_t o_out = sg.add(16);
    if (ctx_reserve(sg.used)) return -1;
    gather(ctx().h + o_org, org, h, w, i_org, sizeof(pel));
    for (int k = 0; k < n; k++) gather(ctx().h + o_pred + k * blk, p[k], h, w, i_pred, sizeof(pel));
    if (ctx_upload(o_out)) return -1;
    hipLaunchKernelGGL(k_sad_n, dim3(1), dim3(256), 0, ctx().st, (const pel *)(ctx().d + o_org),
                       (const pel *)(ctx().d + o_pred), (size_t)w * h, w, h, n, (u32 *)(ctx().d + o_out));
    ME_CHECK(hipGetLastError());
    if (ctx_download(o_out, n * sizeof(u32))) return -1;
    memcpy(out, ctx().h + o_out, n * sizeof(u32));
    return 0;
}

ME_EXPORT uint32_t me_hip_sad(int w, me_pel *org, int i_org, me_pel *pred, int i_pred, int h)
{
    u32 v = 0;
    pel *p[1] = {pred};
    if (sad_n(w, org, i_org, p, 1, i_pred, h, &v) && entry_failed("cost_sad", g_prev.cost_sad[log2w(w)]))
        return g_prev.cost_sad[log2w(w)](org, i_org, pred, i_pred, h);
    return v;
}

ME_EXPORT void me_hip_sad_x3(int w, me_pel *org, int i_org, me_pel *p0, me_pel *p1, me_pel *p2, int i_pred, uint32_t sad[3], int h)
{
    pel *p[3] = {p0, p1, p2};
    if (sad_n(w, org, i_org, p, 3, i_pred, h, sad) && entry_failed("cost_sad_x3", g_prev.cost_sad_x3[log2w(w)]))
        g_prev.cost_sad_x3[log2w(w)](org, i_org, p0, p1, p2, i_pred, sad, h);
}

ME_EXPORT void me_hip_sad_x4(int w, me_pel *org, int i_org, me_pel *p0, me_pel *p1, me_pel *p2, me_pel *p3, int i_pred,
                             uint32_t sad[4], int h)
{
    pel *p[4] = {p0, p1, p2, p3};
    if (sad_n(w, org, i_org, p, 4, i_pred, h, sad) && entry_failed("cost_sad_x4", g_prev.cost_sad_x4[log2w(w)]))
        g_prev.cost_sad_x4[log2w(w)](org, i_org, p0, p1, p2, p3, i_pred, sad, h);
}

static int satd_call(int wl, int hl, const pel *org, int i_org, const pel *cur, int i_cur, u32 *out)
{
    const int tw = 4 << wl, th = 4 << hl;
    int tt;
    if (wl == 0 && hl == 0) tt = 0;
    else if (wl == 1 && hl == 1) tt = 1;
    else if (wl == 1 && hl == 0) tt = 2;
    else if (wl == 2 && hl == 1) tt = 3;
    else if (wl == 0 && hl == 1) tt = 4;
    else if (wl == 1 && hl == 2) tt = 5;
    else { me_set_error("cost_satd: empty slot", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    const size_t blk = (size_t)tw * th * sizeof(pel);
    Stage sg;
    size_t oo = sg.add(blk), oc = sg.add(blk), ou = sg.add(16);
    if (ctx_reserve(sg.used)) return -1;
    gather(ctx().h + oo, org, th, tw, i_org, sizeof(pel));
    gather(ctx().h + oc, cur, th, tw, i_cur, sizeof(pel));
    if (ctx_upload(ou)) return -1;
    hipLaunchKernelGGL(k_satd, dim3(1), dim3(64), 0, ctx().st, tt, (const pel *)(ctx().d + oo), tw,
                       (const pel *)(ctx().d + oc), tw, (u32 *)(ctx().d + ou));
    ME_CHECK(hipGetLastError());
    if (ctx_download(ou, 4)) return -1;
    memcpy(out, ctx().h + ou, 4);
    return 0;
}

ME_EXPORT uint32_t me_hip_satd(int wl, int hl, me_pel *org, int i_org, me_pel *cur, int i_cur)
{
    u32 v = 0;
    if (satd_call(wl, hl, org, i_org, cur, i_cur, &v) && entry_failed("cost_satd", wl < 3 && hl < 3 && g_prev.cost_satd[wl][hl]))
        return g_prev.cost_satd[wl][hl](org, i_org, cur, i_cur);
    return v;
}

ME_EXPORT uint32_t me_hip_had(int w, int h, me_pel *org, int s_org, me_pel *cur, int s_cur, int bd)
{
    const size_t blk = (size_t)w * h * sizeof(pel);
    Stage sg;
    size_t oo = sg.add(blk), oc = sg.add(blk), ou = sg.add(16);
    u32 v = 0;
    int rc = ctx_reserve(sg.used);
    if (!rc) {
        gather(ctx().h + oo, org, h, w, s_org, sizeof(pel));
        gather(ctx().h + oc, cur, h, w, s_cur, sizeof(pel));
        rc = ctx_upload(ou);
    }
    if (!rc) {
        hipLaunchKernelGGL(k_had, dim3(1), dim3(256), 0, ctx().st, w, h, (const pel *)(ctx().d + oo), w,
                           (const pel *)(ctx().d + oc), w, bd, (u32 *)(ctx().d + ou));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { me_set_error("k_had", e, __FILE__, __LINE__); rc = -1; }
    }
    if (!rc) rc = ctx_download(ou, 4);
    if (rc) { entry_failed("com_had", false); return 0; }
    memcpy(&v, ctx().h + ou, 4);
    return v;
}

static int ipcpy_call(const pel *src, int i_src, pel *dst, int i_dst, int w, int h)
{
    const size_t blk = (size_t)w * h * sizeof(pel);
    if (ctx_reserve(blk)) return -1;
    gather(ctx().h, src, h, w, i_src, sizeof(pel));
    if (ctx_upload(blk)) return -1;
    // the copy itself is the kernel-free device round trip of the block
    if (ctx_download(0, blk)) return -1;
    scatter(dst, ctx().h, h, w, i_dst, sizeof(pel));
    return 0;
}

ME_EXPORT void me_hip_ipcpy(const me_pel *src, int i_src, me_pel *dst, int i_dst, int w, int h)
{
    if (ipcpy_call(src, i_src, dst, i_dst, w, h) && entry_failed("ipcpy", g_prev.ipcpy[log2w(w)]))
        g_prev.ipcpy[log2w(w)](src, i_src, dst, i_dst, w, h);
}

// Footprint read by the reference filter of each type (com_mc.c:424-643):
// 8-tap: -3..+4 around each sample, 4-tap: -1..+2, only along the filtered
// direction(s).
static int ipflt_call(int type, int ext, const pel *src, int i_src, pel *dst, int i_dst, int w, int h,
                      const s8 *cx, const s8 *cy, int max_val)
{
    const int n = ext ? (type == 0 ? 8 : 4) : ((type == 0 || type == 2) ? 8 : 4);
    const int hor = ext || type == 0 || type == 1, ver = ext || type == 2 || type == 3;
    const int lo = n / 2 - 1, hi = n / 2;
    const int x0 = hor ? -lo : 0, y0 = ver ? -lo : 0;
    const int rw = w + (hor ? lo + hi : 0), rh = h + (ver ? lo + hi : 0);
    Stage sg;
    size_t os = sg.add((size_t)rw * rh * sizeof(pel)), od = sg.add((size_t)w * h * sizeof(pel));
    if (ctx_reserve(sg.used)) return -1;
    gather(ctx().h + os, src + (ptrdiff_t)y0 * i_src + x0, rh, rw, i_src, sizeof(pel));
    if (ctx_upload(od)) return -1;
    Taps8 a, b;
    memset(&a, 0, sizeof(a));
    memset(&b, 0, sizeof(b));
    for (int k = 0; k < n; k++) { a.c[k] = cx[k]; if (cy) b.c[k] = cy[k]; }
    const pel *s0 = (const pel *)(ctx().d + os) + (ptrdiff_t)(-y0) * rw + (-x0);
    dim3 g((w + 63) / 64, h);
    if (ext) hipLaunchKernelGGL(k_ipflt_ext, g, dim3(64), 0, ctx().st, s0, rw, (pel *)(ctx().d + od), w, h, n, a, b, max_val);
    else hipLaunchKernelGGL(k_ipflt, g, dim3(64), 0, ctx().st, s0, rw, (pel *)(ctx().d + od), w, h, n, (int)ver, a, max_val);
    ME_CHECK(hipGetLastError());
    if (ctx_download(od, (size_t)w * h * sizeof(pel))) return -1;
    scatter(dst, ctx().h + od, h, w, i_dst, sizeof(pel));
    return 0;
}

ME_EXPORT void me_hip_ipflt(int type, const me_pel *src, int i_src, me_pel *dst, int i_dst, int w, int h,
                            const int8_t *coeff, int max_val)
{
    if (ipflt_call(type, 0, src, i_src, dst, i_dst, w, h, coeff, nullptr, max_val) &&
        entry_failed("ipflt", type >= 0 && type < 4 && g_prev.ipflt[type][log2w(w)]))
        g_prev.ipflt[type][log2w(w)](src, i_src, dst, i_dst, w, h, coeff, max_val);
}

ME_EXPORT void me_hip_ipflt_ext(int type, const me_pel *src, int i_src, me_pel *dst, int i_dst, int w, int h,
                                const int8_t *coef_x, const int8_t *coef_y, int max_val)
{
    if (ipflt_call(type, 1, src, i_src, dst, i_dst, w, h, coef_x, coef_y, max_val) &&
        entry_failed("ipflt_ext", type >= 0 && type < 2 && g_prev.ipflt_ext[type][log2w(w)]))
        g_prev.ipflt_ext[type][log2w(w)](src, i_src, dst, i_dst, w, h, coef_x, coef_y, max_val);
}

static int avrg_call(int w, pel *dst, int i_dst, const pel *s1, const pel *s2, int h)
{
    const size_t blk = (size_t)w * h * sizeof(pel);
    Stage sg;
    size_t oa = sg.add(blk), ob = sg.add(blk), od = sg.add(blk);
    if (ctx_reserve(sg.used)) return -1;
    memcpy(ctx().h + oa, s1, blk);
    memcpy(ctx().h + ob, s2, blk);
    if (ctx_upload(od)) return -1;
    hipLaunchKernelGGL(k_avrg, dim3((w * h + 255) / 256), dim3(256), 0, ctx().st, (const pel *)(ctx().d + oa),
                       (const pel *)(ctx().d + ob), (pel *)(ctx().d + od), w * h);
    ME_CHECK(hipGetLastError());
    if (ctx_download(od, blk)) return -1;
    scatter(dst, ctx().h + od, h, w, i_dst, sizeof(pel));
    return 0;
}

ME_EXPORT void me_hip_pel_avrg(int w, me_pel *dst, int i_dst, me_pel *s1, me_pel *s2, int h)
{
    if (avrg_call(w, dst, i_dst, s1, s2, h) && entry_failed("pel_avrg", g_prev.pel_avrg[log2w(w)]))
        g_prev.pel_avrg[log2w(w)](dst, i_dst, s1, s2, h);
}

static Coef3 coef3(int8_t (*coeff)[8])
{
    Coef3 c;
    for (int p = 0; p < 3; p++)
        for (int k = 0; k < 8; k++) c.c[p][k] = coeff[p][k];
    return c;
}

static int flt_hor_call(const pel *src, int i_src, pel *dst[3], int i_dst, s16 *dt[3], int i_dt, int w, int h,
                        int8_t (*coeff)[8], int bd)
{
    const int rw = w + 7;
    Stage sg;
    size_t os = sg.add((size_t)rw * h * sizeof(pel));
    size_t od = sg.add((size_t)3 * w * h * sizeof(pel));
    size_t ot = sg.add((size_t)3 * w * h * sizeof(s16));
    if (ctx_reserve(sg.used)) return -1;
    gather(ctx().h + os, src - 3, h, rw, i_src, sizeof(pel));
    if (ctx_upload(od)) return -1;
    hipLaunchKernelGGL(k_flt_hor, dim3((w + 255) / 256, h, 3), dim3(256), 0, ctx().st,
                       (const pel *)(ctx().d + os) + 3, rw, (pel *)(ctx().d + od), (s16 *)(ctx().d + ot), w, h,
                       coef3(coeff), bd);
    ME_CHECK(hipGetLastError());
    if (ctx_download(od, sg.used - od)) return -1;
    for (int p = 0; p < 3; p++) {
        scatter(dst[p], ctx().h + od + (size_t)p * w * h * sizeof(pel), h, w, i_dst, sizeof(pel));
        scatter(dt[p], ctx().h + ot + (size_t)p * w * h * sizeof(s16), h, w, i_dt, sizeof(s16));
    }
    return 0;
}

ME_EXPORT void me_hip_ip_flt_y_hor(const me_pel *src, int i_src, me_pel *dst[3], int i_dst, int16_t *dst_tmp[3],
                                   int i_dst_tmp, int w, int h, int8_t (*coeff)[8], int bd)
{
    if (flt_hor_call(src, i_src, dst, i_dst, dst_tmp, i_dst_tmp, w, h, coeff, bd) &&
        entry_failed("ip_flt_y_hor", g_prev.ip_flt_y_hor))
        g_prev.ip_flt_y_hor(src, i_src, dst, i_dst, dst_tmp, i_dst_tmp, w, h, coeff, bd);
}

static int flt_ver_call(const void *src, int es, int i_src, pel *dst[3], int i_dst, int w, int h,
                        int8_t (*coeff)[8], int bd, bool ext)
{
    const int rh = h + 7;
    Stage sg;
    size_t os = sg.add((size_t)w * rh * es);
    size_t od = sg.add((size_t)3 * w * h * sizeof(pel));
    if (ctx_reserve(sg.used)) return -1;
    gather(ctx().h + os, (const unsigned char *)src - (ptrdiff_t)3 * i_src * es, rh, w, i_src, es);
    if (ctx_upload(od)) return -1;
    if (ext)
        hipLaunchKernelGGL(k_flt_ver_ext, dim3((w + 255) / 256, h, 3), dim3(256), 0, ctx().st,
                           (const s16 *)(ctx().d + os) + 3 * w, w, (pel *)(ctx().d + od), w, h, coef3(coeff), bd);
    else
        hipLaunchKernelGGL(k_flt_ver, dim3((w + 255) / 256, h, 3), dim3(256), 0, ctx().st,
                           (const pel *)(ctx().d + os) + 3 * w, w, (pel *)(ctx().d + od), w, h, coef3(coeff), bd);
    ME_CHECK(hipGetLastError());
    if (ctx_download(od, (size_t)3 * w * h * sizeof(pel))) return -1;
    for (int p = 0; p < 3; p++) scatter(dst[p], ctx().h + od + (size_t)p * w * h * sizeof(pel), h, w, i_dst, sizeof(pel));
    return 0;
}

ME_EXPORT void me_hip_ip_flt_y_ver(const me_pel *src, int i_src, me_pel *dst[3], int i_dst, int w, int h,
                                   int8_t (*coeff)[8], int bd)
{
    if (flt_ver_call(src, sizeof(pel), i_src, dst, i_dst, w, h, coeff, bd, false) &&
        entry_failed("ip_flt_y_ver", g_prev.ip_flt_y_ver))
        g_prev.ip_flt_y_ver(src, i_src, dst, i_dst, w, h, coeff, bd);
}

ME_EXPORT void me_hip_ip_flt_y_ver_ext(const int16_t *src, int i_src, me_pel *dst[3], int i_dst, int w, int h,
                                       int8_t (*coeff)[8], int bd)
{
    if (flt_ver_call(src, sizeof(s16), i_src, dst, i_dst, w, h, coeff, bd, true) &&
        entry_failed("ip_flt_y_ver_ext", g_prev.ip_flt_y_ver_ext))
        g_prev.ip_flt_y_ver_ext(src, i_src, dst, i_dst, w, h, coeff, bd);
}

// ---------------------------------------------------------------------------
// Self-check of the one floating-point step
// ---------------------------------------------------------------------------

__global__ void k_satd_norm(int big, u32 s0, u32 n, u32 *out)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = (int)(s0 + i);
    out[i] = (u32)(big ? me_satd_norm128(s) : me_satd_norm32(s));
}

ME_EXPORT int me_hip_satd_norm(int big, uint32_t s0, uint32_t n, uint32_t *out)
{
    if (!out || n == 0 || (uint64_t)s0 + n > 0x80000000ull) {
        me_set_error("me_hip_satd_norm: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (me_hip_init(-1)) return -1;
    u32 *d = nullptr;
    ME_CHECK(hipMalloc((void **)&d, (size_t)n * sizeof(u32)));
    hipLaunchKernelGGL(k_satd_norm, dim3((n + 255) / 256), dim3(256), 0, 0, big, s0, n, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, d, (size_t)n * sizeof(u32), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) { me_set_error("me_hip_satd_norm", e, __FILE__, __LINE__); return -1; }
    return 0;
}

// Width-specialised wrappers: the reference slots of SAD and pel_avrg carry
// the block width in their table index only.
template <int W> static u32 sad_w(me_pel *o, int io, me_pel *p, int ip, int h) { return me_hip_sad(W, o, io, p, ip, h); }
template <int W> static void sad3_w(me_pel *o, int io, me_pel *a, me_pel *b, me_pel *c, int ip, u32 s[3], int h)
{ me_hip_sad_x3(W, o, io, a, b, c, ip, s, h); }
template <int W> static void sad4_w(me_pel *o, int io, me_pel *a, me_pel *b, me_pel *c, me_pel *d, int ip, u32 s[4], int h)
{ me_hip_sad_x4(W, o, io, a, b, c, d, ip, s, h); }
template <int W> static void avrg_w(me_pel *d, int id, me_pel *a, me_pel *b, int h) { me_hip_pel_avrg(W, d, id, a, b, h); }
template <int T> static void flt_w(const me_pel *s, int is, me_pel *d, int id, int w, int h, const int8_t *c, int mv)
{ me_hip_ipflt(T, s, is, d, id, w, h, c, mv); }
template <int T> static void fltx_w(const me_pel *s, int is, me_pel *d, int id, int w, int h, const int8_t *a, const int8_t *b, int mv)
{ me_hip_ipflt_ext(T, s, is, d, id, w, h, a, b, mv); }
template <int WL, int HL> static u32 satd_w(me_pel *o, int io, me_pel *p, int ip) { return me_hip_satd(WL, HL, o, io, p, ip); }

static void fill_hip_slots(me_dispatch_t *t);

ME_EXPORT int me_funs_init_hip(me_dispatch_t *t)
{
    if (!t) return -1;
    if (me_hip_init(-1)) return -1;
    me_dispatch_t hip;
    memset(&hip, 0, sizeof(hip));
    fill_hip_slots(&hip);
    {
        // Slot by slot: remember what the table held unless it is already ours.
        std::lock_guard<std::mutex> lk(g_init_mu);
        void **slot = (void **)t, **ours = (void **)&hip, **prev = (void **)&g_prev;
        for (size_t i = 0; i < sizeof(me_dispatch_t) / sizeof(void *); i++)
            if (ours[i] && slot[i] != ours[i]) prev[i] = slot[i];
    }
    fill_hip_slots(t);
    return 0;
}

// Writes every HIP entry into its slot of t.
static void fill_hip_slots(me_dispatch_t *t)
{
    t->cost_sad[0] = sad_w<4>; t->cost_sad[1] = sad_w<8>; t->cost_sad[2] = sad_w<16>;
    t->cost_sad[3] = sad_w<32>; t->cost_sad[4] = sad_w<64>; t->cost_sad[5] = sad_w<128>;
    t->cost_sad_x3[0] = sad3_w<4>; t->cost_sad_x3[1] = sad3_w<8>; t->cost_sad_x3[2] = sad3_w<16>;
    t->cost_sad_x3[3] = sad3_w<32>; t->cost_sad_x3[4] = sad3_w<64>; t->cost_sad_x3[5] = sad3_w<128>;
    t->cost_sad_x4[0] = sad4_w<4>; t->cost_sad_x4[1] = sad4_w<8>; t->cost_sad_x4[2] = sad4_w<16>;
    t->cost_sad_x4[3] = sad4_w<32>; t->cost_sad_x4[4] = sad4_w<64>; t->cost_sad_x4[5] = sad4_w<128>;
    t->pel_avrg[0] = avrg_w<4>; t->pel_avrg[1] = avrg_w<8>; t->pel_avrg[2] = avrg_w<16>;
    t->pel_avrg[3] = avrg_w<32>; t->pel_avrg[4] = avrg_w<64>; t->pel_avrg[5] = avrg_w<128>;
    for (int i = 0; i < 6; i++) {
        t->ipcpy[i] = me_hip_ipcpy;
        t->ipflt[0][i] = flt_w<0>; t->ipflt[1][i] = flt_w<1>; t->ipflt[2][i] = flt_w<2>; t->ipflt[3][i] = flt_w<3>;
        t->ipflt_ext[0][i] = fltx_w<0>; t->ipflt_ext[1][i] = fltx_w<1>;
    }
    t->ip_flt_y_hor = me_hip_ip_flt_y_hor;
    t->ip_flt_y_ver = me_hip_ip_flt_y_ver;
    t->ip_flt_y_ver_ext = me_hip_ip_flt_y_ver_ext;
    t->cost_satd[0][0] = satd_w<0, 0>; t->cost_satd[1][0] = satd_w<1, 0>; t->cost_satd[0][1] = satd_w<0, 1>;
    t->cost_satd[1][1] = satd_w<1, 1>; t->cost_satd[2][1] = satd_w<2, 1>; t->cost_satd[1][2] = satd_w<1, 2>;
}
