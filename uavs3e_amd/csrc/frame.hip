// Batched frame motion estimation for gfx950: search configurations and
// frame plans.
//
// Path per frame (one hipGraph):
//   [optional] pad + 15 quarter-pel planes of a just-reconstructed reference
//   k_fill_keys       -> key buffer := all ones
//   k_int_search      -> integer full search of every PU, every reference
//   k_subpel          -> me_sub_pel_search + final cost of every (PU, reference)
//   k_select          -> reference choice per PU
//
// Reference semantics (file:line in uavs3e):
//   integer cost  (sad << (16-(bd-8))) + (bits_x+bits_y)*lambda   me.c:119,151-187
//   window        get_max_search_range (is_padding = 1)           me.c:502-527
//   full search   raster y-major, first strict minimum wins        me.c:150-190
//   sub-pel       me_sub_pel_search, cost_type 0 (SATD)            me.c:366-500
//   final cost    + lambda*(refi_bits+mvr+(mvr<4)+2), (c+2^15)>>16 me.c:560-567
//   ref choice    smallest final cost, lower index on ties         inter.c:901-904
#include "frame.h"

#include <math.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

__global__ void k_fill_keys(u64 *keys, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = ~0ull;
}

// Reference choice per PU and list: smallest final cost, lower index on
// ties (inter.c:901-904); -1 for a PU that was not searched.  References
// [0, n0) form list 0, [n0, nref) list 1; best[i * nl + l].
__global__ void k_select(const me_pu_result_t *res, s8 *best, int n, int nref, int n0, int nl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int l = 0; l < nl; l++) {
        u32 bc = 0xFFFFFFFFu;
        int br = -1;
        for (int r = l ? n0 : 0; r < (l ? nref : n0); r++) {
            const me_pu_result_t &o = res[(size_t)i * nref + r];
            if (o.valid && o.cost_final < bc) { bc = o.cost_final; br = r; }
        }
        best[(size_t)i * nl + l] = (s8)br;
    }
}

int me_int_search_launch(int mode, int waves, const IntArgs &a, size_t lds, hipStream_t st)
{
    switch (waves) {
    case 1: return me_int_search_launch_w1(mode, a, lds, st);
    case 2: return me_int_search_launch_w2(mode, a, lds, st);
    case 8: return me_int_search_launch_w8(mode, a, lds, st);
    default: return me_int_search_launch_w4(mode, a, lds, st);
    }
}


// ---------------------------------------------------------------------------
// Search configurations
// ---------------------------------------------------------------------------

// Closed form of the MV-bits table of inter_search_create (inter.c:2029-2058):
// tab[0] = 1, tab[+-v] = 2 * (floor(log2(v + 1)) + 1) for 0 < v < off (the
// loop overwrites the seeded +-1, +-2 entries), tab[+-off] = 0.
void me_mvbits_fill(u8 *base, int off)
{
    base[off] = 1;
    base[0] = base[2 * off] = 0;
    for (int v = 1; v < off; v++) {
        int l = 0;
        while ((2 << l) <= v + 1) l++;
        base[off + v] = base[off - v] = (u8)(2 * (l + 1));
    }
}

// The host part of a search configuration: window geometry, key layout,
// schedule, and the sub-pel shape table (shp[0 .. *n_shp), in launch order).
// Allocates nothing.
static int search_layout(MeSearch *p, int R, u32 lambda, u32 shape_mask, int nref, int nref_l1, int max_dim,
                         int band_rows, int waves, int dx_cols, int sp_groups, SubShape shp[ME_NSHAPES], int *n_shp)
{
    memset(p, 0, sizeof(*p));
    const uint32_t m = shape_mask & ((1u << ME_NSHAPES) - 1);
    if (R < 0 || R > 255 || nref < 1 || nref > ME_MAX_REFS || nref_l1 < 0 || nref_l1 >= nref || !m || max_dim < 1) {
        me_set_error("me_search_init: bad configuration", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    p->R = R;
    p->lambda = lambda;
    p->nref = nref;
    p->nref_l1 = nref_l1;
    p->nl = nref_l1 > 0 ? 2 : 1;
    p->span = 2 * R + 1;
    int need = 0;
    while ((1 << need) < p->span * p->span) need++;
    const int shift = 16 - (BIT_DEPTH - 8);
    p->IB = need > 32 - shift ? need : 32 - shift;
    p->sh = shift + p->IB - 32;
    // Uniform part of the window: |mv| <= T - 4 with T = max(W,H) (me.c:504-512).
    const int T = max_dim;
    p->lo = -R > -T + 4 ? -R : -T + 4;
    p->hi = R < T - 4 ? R : T - 4;
    p->mode = m == (1u << S16x16) ? 1 : 0;
    int n = 0;
    for (int s = 0; s < ME_NSHAPES; s++) {
        if (m >> s & 1) { p->base[s] = n; n += (128 / ME_SHAPES_HOST[s][0]) * (128 / ME_SHAPES_HOST[s][1]); }
        else p->base[s] = -1;
    }
    p->npu = n;
    p->waves = waves > 0 ? (waves >= 8 ? 8 : waves >= 4 ? 4 : (waves >= 2 ? 2 : 1)) : 4;
    p->band_rows = band_rows > 0 ? band_rows : 16;
    if (p->band_rows > p->span) p->band_rows = p->span;
    p->nbands = (p->span + p->band_rows - 1) / p->band_rows;
    // Columns: multiples of 4 displacements (one v_qsad_pk_u16_u8 step).
    const int full_cols = (p->span + 3) & ~3;
    p->dx_cols = dx_cols > 0 ? ((dx_cols + 3) & ~3) : full_cols;
    if (p->dx_cols > full_cols) p->dx_cols = full_cols;
    p->nbx = (p->span + p->dx_cols - 1) / p->dx_cols;
    // A lane reads up to 20 samples from column X + (dxb - dxs) <= 112 + dx_cols - 4.
    p->wcols = ((128 + p->dx_cols + 16 + 15) / 16) * 16;
    p->wpitch = p->wcols + 16;
    p->lds = (size_t)(128 + p->band_rows - 1) * p->wpitch * sizeof(pel) + 128 * 128 * sizeof(pel);
    // the same LDS later merges the waves' node states: at most 58 keys per lane
    if (p->lds < 58 * 64 * sizeof(u64)) p->lds = 58 * 64 * sizeof(u64);
    if (p->lds > 160 * 1024) {
        me_set_error("me_search_init: LDS window too large (reduce band_rows or dx_cols)", hipErrorInvalidValue,
                     __FILE__, __LINE__);
        return -1;
    }
    p->tab_off = (max_dim << 2) + 3;
    // Sub-pel shape table: every searched shape's workgroups in one launch,
    // the shapes with the most samples per PU first (their workgroups run
    // longest, so they should not trail at the end of the launch).
    const int G = sp_groups >= 8 ? 8 : sp_groups >= 4 ? 4 : sp_groups >= 2 ? 2 : 1;
    // 8 bits, one position at a time: the matrix-core kernel (no 4x4 tiles:
    // the smallest shapes are 8x4 / 4x8)
    const bool mx = BIT_DEPTH == 8 && G == 1;
    int order[ME_NSHAPES], no = 0;
    for (int s = 0; s < ME_NSHAPES; s++)
        if (p->base[s] >= 0) order[no++] = s;
    std::stable_sort(order, order + no, [](int a, int b) {
        return ME_SHAPES_HOST[a][0] * ME_SHAPES_HOST[a][1] > ME_SHAPES_HOST[b][0] * ME_SHAPES_HOST[b][1];
    });
    int blocks_total = 0;
    for (int k = 0; k < no; k++) {
        const int s = order[k];
        SubShape &q = shp[k];
        memset(&q, 0, sizeof(q));
        q.pw = ME_SHAPES_HOST[s][0];
        q.ph = ME_SHAPES_HOST[s][1];
        had_tiling(q.pw, q.ph, &q.tt, &q.tw, &q.th);
        const int Tn = (q.pw / q.tw) * (q.ph / q.th);
        if (Tn > 64 && Tn != 256) {
            me_set_error("me_search_init: no sub-pel tiling for a shape", hipErrorInvalidValue, __FILE__, __LINE__);
            return -1;
        }
        q.tpr = q.pw / q.tw;
        q.base = p->base[s];
        q.per_ctu = (128 / q.pw) * (128 / q.ph);
        q.blk0 = blocks_total;
        if (mx) {
            // k_subpel_mx: one D column per tile, 128 tiles per workgroup
            // (256 for 128x128: two 32-tile batches per wave); the tiles of
            // a CTU's PUs of one shape fill 1, 2 or 4 workgroups
            q.lpu = Tn;
            q.G = Tn > 128 ? 2 : 1;
            blocks_total += q.per_ctu * Tn / (128 * q.G);
            continue;
        }
        q.lpu = q.tt == 3 || q.tt == 5 ? 2 * Tn : Tn;   // 1..128, or 256 for 128x128
        // up to sp_groups search positions of a PU at a time, at most
        // 256 lanes per PU; lanes per (CTU, reference): a multiple of 256
        q.G = 256 / q.lpu < G ? 256 / q.lpu : G;
        blocks_total += q.per_ctu * q.lpu * q.G / 256;
    }
    p->sub_blocks = blocks_total;
    p->sub_mx = mx ? 1 : 0;
    *n_shp = no;
    return 0;
}

// Workgroup -> shape of the sub-pel launch: shape k covers the workgroups
// [shp[k].blk0, shp[k + 1].blk0).
static void blk_shape_fill(const SubShape *shp, int n_shp, int sub_blocks, u8 *out)
{
    for (int k = 0; k < n_shp; k++) {
        const int end = k + 1 < n_shp ? shp[k + 1].blk0 : sub_blocks;
        for (int b = shp[k].blk0; b < end; b++) out[b] = (u8)k;
    }
}

int me_search_init(MeSearch *p, int R, u32 lambda, u32 shape_mask, int nref, int nref_l1, int max_dim, int band_rows,
                   int waves, int dx_cols, int sp_groups)
{
    SubShape shp[ME_NSHAPES];
    int n_shp = 0;
    if (search_layout(p, R, lambda, shape_mask, nref, nref_l1, max_dim, band_rows, waves, dx_cols, sp_groups, shp,
                      &n_shp))
        return -1;
    // rate[i] = bits(4*(i-R)) * lambda (GET_MVBITS_IPEL, util.h:43-44)
    std::vector<u8> tab(2 * p->tab_off + 1);
    me_mvbits_fill(tab.data(), p->tab_off);
    std::vector<u32> rate(p->span + ME_RATE_PAD, 0u);
    for (int i = 0; i < p->span; i++) rate[i] = (u32)tab[p->tab_off + 4 * (i - R)] * lambda;
    std::vector<u8> blk_shape(p->sub_blocks);
    blk_shape_fill(shp, n_shp, p->sub_blocks, blk_shape.data());
    hipError_t e = hipMalloc((void **)&p->d_rate, rate.size() * sizeof(u32));
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_tab_base, tab.size());
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_shp, n_shp * sizeof(SubShape));
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_blk_shape, blk_shape.size());
    if (e == hipSuccess) e = hipMemcpy(p->d_rate, rate.data(), rate.size() * sizeof(u32), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_tab_base, tab.data(), tab.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_shp, shp, n_shp * sizeof(SubShape), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_blk_shape, blk_shape.data(), blk_shape.size(), hipMemcpyHostToDevice);
    for (int i = 0; i < 6 && e == hipSuccess; i++) e = hipStreamCreateWithFlags(&p->tz_side[i], hipStreamNonBlocking);
    for (int i = 0; i < 7 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&p->tz_ev[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        me_set_error("me_search_init", e, __FILE__, __LINE__);
        me_search_free(p);
        return -1;
    }
    return 0;
}

void me_search_free(MeSearch *p)
{
    if (p->d_rate) (void)hipFree(p->d_rate);
    if (p->d_tab_base) (void)hipFree(p->d_tab_base);
    if (p->d_shp) (void)hipFree(p->d_shp);
    if (p->d_blk_shape) (void)hipFree(p->d_blk_shape);
    p->d_rate = nullptr;
    p->d_tab_base = nullptr;
    p->d_shp = nullptr;
    p->d_blk_shape = nullptr;
    for (int i = 0; i < 6; i++) {
        if (p->tz_side[i]) (void)hipStreamDestroy(p->tz_side[i]);
        p->tz_side[i] = nullptr;
    }
    for (int i = 0; i < 7; i++) {
        if (p->tz_ev[i]) (void)hipEventDestroy(p->tz_ev[i]);
        p->tz_ev[i] = nullptr;
    }
}

int me_search_warmup(const MeSearch *p, hipStream_t st)
{
    int rc = p->waves == 1   ? me_int_search_preload_w1(p->mode, p->lds, st)
             : p->waves == 2 ? me_int_search_preload_w2(p->mode, p->lds, st)
             : p->waves == 8 ? me_int_search_preload_w8(p->mode, p->lds, st)
                             : me_int_search_preload_w4(p->mode, p->lds, st);
    if (rc || me_subpel_preload(st)) return -1;
    // n = 0: no access
    hipLaunchKernelGGL(k_fill_keys, dim3(1), dim3(256), 0, st, (u64 *)nullptr, (size_t)0);
    hipLaunchKernelGGL(k_select, dim3(1), dim3(256), 0, st, (const me_pu_result_t *)nullptr, (s8 *)nullptr, 0, 1, 1, 1);
    ME_CHECK(hipGetLastError());
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

// Workgroups of the integer search the current device holds at once: the
// kernels are built for 2 waves per SIMD (4 SIMDs per CU), and a CU's LDS
// holds floor(LDS per CU / lds) workgroups.
static int resident_groups(const MeSearch *p)
{
    int dev = 0, cus = 0, lds_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        lds_cu <= 0)
        lds_cu = 160 * 1024;
    int per_cu = 8 / p->waves;
    const int by_lds = (int)((size_t)lds_cu / p->lds);
    if (by_lds < per_cu) per_cu = by_lds;
    if (per_cu < 1) per_cu = 1;
    return cus * per_cu;
}

size_t me_search_keys(const MeSearch *p, int n_jobs, int grid, KeyLayout *kl)
{
    memset(kl, 0, sizeof(*kl));
    kl->nbands = p->nbands;
    kl->nbx = p->nbx;
    kl->slot_elems = (size_t)n_jobs * p->npu * p->nref;
    const u64 units = (u64)n_jobs * p->nref * p->nbx * p->nbands;
    kl->units = (u32)units;
    kl->groups = (u32)units;
    // grid 0: as many workgroups as the device holds, when there are units
    // to fill it twice over; n >= 2: n workgroups whenever that leaves some
    // with several units; 1: one workgroup per unit
    const int cap = grid == 0 ? resident_groups(p) : grid >= 2 ? grid : 0;
    const bool enough = grid == 0 ? units >= 2 * (u64)cap : units > (u64)cap;
    if (cap > 0 && enough && units < 0x7FFFFFFFull) {
        KeyLayout t = *kl;
        t.groups = (u32)cap;
        int most = 1;
        const u64 nseg = (u64)n_jobs * p->nref * p->nbx;
        for (u64 x = 0; x < nseg; x++) {
            const int np = key_pieces(t, x);
            if (np > most) most = np;
        }
        // a persistent launch writes most * nbx slots; beyond 4 pieces the
        // atomic layout moves fewer bytes
        if (most <= 4) {
            t.nslots = most;
            *kl = t;
            return (size_t)p->nbx * most * kl->slot_elems;
        }
    }
    return kl->slot_elems;
}

int me_search_fill(const MeSearch *p, const JobSet &js, const KeyLayout &kl, u64 *keys, hipStream_t st)
{
    (void)p;
    if (kl.nslots) return 0;   // persistent: every slot read is written by its piece
    const size_t nkeys = kl.slot_elems;
    hipLaunchKernelGGL(k_fill_keys, dim3((unsigned)((nkeys + 255) / 256)), dim3(256), 0, st, keys, nkeys);
    ME_CHECK(hipGetLastError());
    return 0;
}

int me_search_int(const MeSearch *p, const JobSet &js, const KeyLayout &kl, u64 *keys, hipStream_t st)
{
    IntArgs a;
    memset(&a, 0, sizeof(a));
    a.js = js;
    a.R = p->R; a.lo = p->lo; a.hi = p->hi; a.IB = p->IB; a.sh = p->sh;
    a.rate = p->d_rate; a.keys = keys; a.kl = kl; a.npu = p->npu; a.nref = p->nref;
    a.band_rows = p->band_rows; a.wcols = p->wcols; a.dx_cols = p->dx_cols; a.wpitch = p->wpitch;
    memcpy(a.base, p->base, sizeof(a.base));
    return me_int_search_launch(p->mode, p->waves, a, p->lds, st);
}

int me_search_finish(const MeSearch *p, const JobSet &js, const KeyLayout &kl, const u64 *keys, me_pu_result_t *res,
                     s8 *best, hipStream_t st)
{
    SubArgs b;
    memset(&b, 0, sizeof(b));
    b.js = js;
    b.n0 = p->nref - p->nref_l1;
    b.R = p->R; b.IB = p->IB; b.lambda = p->lambda;
    b.tab = p->d_tab_base + p->tab_off;
    b.keys = keys; b.kl = kl; b.res = res;
    b.npu = p->npu; b.nref = p->nref;
    b.shp = p->d_shp;
    b.blk_shape = p->d_blk_shape;
    b.mx = p->sub_mx;
    if (me_subpel_launch(b, p->sub_blocks, js.n_jobs, p->nref, st)) return -1;
    const int npus = js.n_jobs * p->npu;
    hipLaunchKernelGGL(k_select, dim3((npus + 255) / 256), dim3(256), 0, st, res, best, npus, p->nref,
                       p->nref - p->nref_l1, p->nl);
    ME_CHECK(hipGetLastError());
    return 0;
}

int me_search_enqueue(const MeSearch *p, const JobSet &js, const KeyLayout &kl, u64 *keys, me_pu_result_t *res,
                      s8 *best, hipStream_t st, hipEvent_t *ev)
{
    if (js.n_jobs <= 0) return 0;
    if (ev) ME_CHECK(hipEventRecord(ev[0], st));
    if (me_search_fill(p, js, kl, keys, st)) return -1;
    if (ev) ME_CHECK(hipEventRecord(ev[1], st));
    if (me_search_int(p, js, kl, keys, st)) return -1;
    if (ev) ME_CHECK(hipEventRecord(ev[2], st));
    if (me_search_finish(p, js, kl, keys, res, best, st)) return -1;
    if (ev) ME_CHECK(hipEventRecord(ev[3], st));
    return 0;
}

// ---------------------------------------------------------------------------
// Frame plans
// ---------------------------------------------------------------------------

struct me_plan_s {
    me_frame_cfg_t cfg;
    int ctu_cols, ctu_rows, n_ctu;
    MeSearch s;
    PicDesc pic;          // host copy of *d_pic
    PicDesc *d_pic;       // the frame's picture descriptor (current + references)
    pel *d_cur;
    pel *d_pred;          // plan-owned predictions of me_plan_mc (y, then u, then v), or null
    s16 *d_mvp;           // me_plan_run_tz inputs: MVPs, then previous integer MVs
    u64 *d_keys;
    KeyLayout kl;         // the integer search's partition and key slots for the whole frame
    me_pu_result_t *d_res;
    s8 *d_best;
    u32 *d_loka;          // lookahead unit costs (allocated on first use)
    me_bi_pu_t *d_bipu;   // analyze_bi per-PU inputs / results (allocated on first use)
    void *d_bist;         // analyze_bi per-PU state
    pel *d_orgbi;         // analyze_bi targets, [ctu][selected shape][16384]
    me_smvd_pu_t *d_smvd; // analyze_smvd per-PU results (allocated on first use)
    me_pu_best_t *d_bestc; // me_plan_best output (allocated on first use)
    hipStream_t own;
    hipGraphExec_t exec;
    u64 graph_refs[ME_MAX_REFS];   // me_ref_s::id of the references the graph was captured with
    u32 graph_mask;
    hipStream_t graph_stream;
};

ME_EXPORT int me_plan_create(me_plan_t *out, const me_frame_cfg_t *cfg)
{
    if (me_hip_init(-1)) return -1;
    if (!out || !cfg || cfg->pic_w <= 0 || cfg->pic_h <= 0 || cfg->pic_w > 128 * 0xFFFF) {
        me_set_error("me_plan_create: bad configuration", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    me_plan_s *p = (me_plan_s *)calloc(1, sizeof(me_plan_s));
    p->cfg = *cfg;
    p->ctu_cols = (cfg->pic_w + 127) / 128;
    p->ctu_rows = (cfg->pic_h + 127) / 128;
    p->n_ctu = p->ctu_cols * p->ctu_rows;
    const int T = cfg->pic_w > cfg->pic_h ? cfg->pic_w : cfg->pic_h;
    if (cfg->search_range > T - 4) {   // the window and the MV-bits table (4*T+3 each side) hold |mv| <= T-4
        me_set_error("me_plan_create: search range exceeds max(pic_w, pic_h) - 4", hipErrorInvalidValue, __FILE__,
                     __LINE__);
        free(p);
        return -1;
    }
    if (me_search_init(&p->s, cfg->search_range, cfg->lambda_mv, cfg->shape_mask, cfg->num_refs, cfg->num_refs_l1, T,
                       cfg->band_rows, cfg->waves, cfg->dx_cols, cfg->sp_groups)) {
        free(p);
        return -1;
    }
    const size_t nkeys = (size_t)p->n_ctu * p->s.npu * p->s.nref;
    const size_t nslot_keys = me_search_keys(&p->s, p->n_ctu, cfg->int_grid, &p->kl);
    hipError_t e = hipSuccess;
#define ALLOC(ptr, bytes) if (e == hipSuccess) e = hipMalloc((void **)&(ptr), (bytes))
    ALLOC(p->d_cur, (size_t)cfg->pic_w * cfg->pic_h * sizeof(pel) + 64);   // slack for dword row reads
    ALLOC(p->d_pic, sizeof(PicDesc));
    ALLOC(p->d_keys, nslot_keys * sizeof(u64));
    ALLOC(p->d_res, nkeys * sizeof(me_pu_result_t));
    ALLOC(p->d_best, (size_t)p->n_ctu * p->s.npu * p->s.nl);
#undef ALLOC
    if (e == hipSuccess) e = hipStreamCreate(&p->own);
    if (e != hipSuccess) {
        me_set_error("me_plan_create", e, __FILE__, __LINE__);
        me_plan_destroy(p);
        return -1;
    }
    *out = p;
    return 0;
}

ME_EXPORT int me_plan_int_grid(me_plan_t p, int *groups, int *slots)
{
    if (!p) { me_set_error("me_plan_int_grid: bad argument", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    if (groups) *groups = (int)p->kl.groups;
    if (slots) *slots = p->kl.nslots;
    return 0;
}

ME_EXPORT int me_plan_set_cur(me_plan_t p, const me_pel *cur, int stride, int on_device, void *stream)
{
    if (!p || !cur || stride < p->cfg.pic_w) { me_set_error("me_plan_set_cur: bad argument", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    hipStream_t st = stream ? (hipStream_t)stream : p->own;
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (stride == p->cfg.pic_w)   // one copy: a 2-D copy from pageable memory goes row by row
        ME_CHECK(hipMemcpyAsync(p->d_cur, cur, (size_t)p->cfg.pic_w * p->cfg.pic_h * sizeof(pel), kind, st));
    else
        ME_CHECK(hipMemcpy2DAsync(p->d_cur, p->cfg.pic_w * sizeof(pel), cur, stride * sizeof(pel),
                                  p->cfg.pic_w * sizeof(pel), p->cfg.pic_h, kind, st));
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}


// Points the plan's device picture descriptor at refs.  The descriptor is
// read by every launch of the frame, so it is rewritten only between frames
// (after the stream has drained) and only when the references change.
static int plan_set_pic(me_plan_s *p, me_ref_t const *refs, hipStream_t st)
{
    PicDesc d;
    memset(&d, 0, sizeof(d));
    d.cur = p->d_cur;
    for (int r = 0; r < p->s.nref; r++) d.planes[r] = refs[r]->planes;
    d.plane_elems = refs[0]->plane_elems;
    d.ref_stride = refs[0]->stride;
    d.pic_w = p->cfg.pic_w;
    d.pic_h = p->cfg.pic_h;
    d.ctu_cols = p->ctu_cols;
    for (int r = 0; r < p->s.nref; r++) d.chroma[r] = refs[r]->chroma;
    if (memcmp(&d, &p->pic, sizeof(d)) == 0) return 0;
    ME_CHECK(hipStreamSynchronize(st));
    if (p->graph_stream && p->graph_stream != st) ME_CHECK(hipStreamSynchronize(p->graph_stream));
    ME_CHECK(hipMemcpy(p->d_pic, &d, sizeof(d), hipMemcpyHostToDevice));
    p->pic = d;
    return 0;
}

// Enqueues the frame; ev (optional, 5 events) brackets the four stages.
static int plan_enqueue_stages(me_plan_s *p, me_ref_t const *refs, uint32_t prepare_mask, hipStream_t st, hipEvent_t *ev)
{
    if (ev) ME_CHECK(hipEventRecord(ev[0], st));
    for (int r = 0; r < p->s.nref; r++)
        if (prepare_mask >> r & 1)
            if (me_ref_prepare_async(refs[r], st)) return -1;
    JobSet js;
    js.pics = p->d_pic;
    js.jobs = nullptr;
    js.n_jobs = p->n_ctu;
    return me_search_enqueue(&p->s, js, p->kl, p->d_keys, p->d_res, p->d_best, st, ev ? ev + 1 : nullptr);
}

static int plan_enqueue(me_plan_t p, me_ref_t const *refs, uint32_t prepare_mask, hipStream_t st)
{
    return plan_enqueue_stages(p, refs, prepare_mask, st, nullptr);
}

ME_EXPORT int me_plan_run(me_plan_t p, me_ref_t const *refs, uint32_t prepare_mask, void *stream)
{
    if (!p || !refs) { me_set_error("me_plan_run: bad argument", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    for (int r = 0; r < p->s.nref; r++)
        if (!refs[r] || refs[r]->w != p->cfg.pic_w || refs[r]->h != p->cfg.pic_h) {
            me_set_error("me_plan_run: reference size differs from the plan", hipErrorInvalidValue, __FILE__, __LINE__);
            return -1;
        }
    hipStream_t st = stream ? (hipStream_t)stream : p->own;
    // The other entry points point the shared descriptor at their own
    // references, so it is set for every run (a no-op when unchanged).
    if (plan_set_pic(p, refs, st)) return -1;
    // The graph's preparation kernels hold the references' buffers, so the
    // cache is keyed on their ids: a destroyed reference whose handle address
    // comes back from a later me_ref_create does not match.
    bool same = p->exec && p->graph_mask == prepare_mask && p->graph_stream == st;
    for (int r = 0; same && r < p->s.nref; r++) same = p->graph_refs[r] == refs[r]->id;
    if (!same) {
        if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
        hipGraph_t g = nullptr;
        ME_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = plan_enqueue(p, refs, prepare_mask, st);
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (!rc) me_set_error("hipStreamEndCapture", e, __FILE__, __LINE__);
            return -1;
        }
        e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { me_set_error("hipGraphInstantiate", e, __FILE__, __LINE__); p->exec = nullptr; return -1; }
        for (int r = 0; r < p->s.nref; r++) p->graph_refs[r] = refs[r]->id;
        p->graph_mask = prepare_mask;
        p->graph_stream = st;
    }
    ME_CHECK(hipGraphLaunch(p->exec, st));
    return 0;
}

// Stage timing: runs `iters` frames stage by stage (no graph) with HIP
// events on the plan's stream between the stages and returns the average
// milliseconds of: [0] reference preparation (prepare_mask), [1] key fill,
// [2] integer search, [3] sub-pel + reference choice.
ME_EXPORT int me_plan_profile(me_plan_t p, me_ref_t const *refs, uint32_t prepare_mask, int iters, double *ms)
{
    if (!p || !refs || !ms || iters < 1) return -1;
    hipStream_t st = p->own;
    for (int r = 0; r < p->s.nref; r++)
        if (!refs[r] || refs[r]->w != p->cfg.pic_w || refs[r]->h != p->cfg.pic_h) return -1;
    if (plan_set_pic(p, refs, st)) return -1;
    hipEvent_t ev[5];
    for (int i = 0; i < 5; i++) ME_CHECK(hipEventCreate(&ev[i]));
    double acc[4] = {0, 0, 0, 0};
    int rc = 0;
    for (int it = 0; it < iters && !rc; it++) {
        rc = plan_enqueue_stages(p, refs, prepare_mask, st, ev);
        if (!rc) {
            ME_CHECK(hipEventSynchronize(ev[4]));
            for (int k = 0; k < 4; k++) {
                float t = 0;
                ME_CHECK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
                acc[k] += t;
            }
        }
    }
    for (int i = 0; i < 5; i++) (void)hipEventDestroy(ev[i]);
    for (int k = 0; k < 4; k++) ms[k] = acc[k] / iters;
    return rc;
}

ME_EXPORT int me_plan_results(me_plan_t p, me_pu_result_t *res, int8_t *best_ref)
{
    if (!p) return -1;
    hipStream_t st = p->graph_stream ? p->graph_stream : p->own;
    ME_CHECK(hipStreamSynchronize(st));
    const size_t n = (size_t)p->n_ctu * p->s.npu;
    if (res) ME_CHECK(hipMemcpy(res, p->d_res, n * p->s.nref * sizeof(me_pu_result_t), hipMemcpyDeviceToHost));
    if (best_ref) ME_CHECK(hipMemcpy(best_ref, p->d_best, n * p->s.nl, hipMemcpyDeviceToHost));
    return 0;
}

// The best reference of every (PU, list) as an encoder consumes it: its
// quarter-pel MV and u32 final cost (cost 0xFFFFFFFF for a skipped PU).
__global__ void k_best(const me_pu_result_t *res, const s8 *best, me_pu_best_t *out, int n, int nref, int nl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * nl) return;
    const int r = best[i];
    me_pu_best_t o;
    if (r >= 0) {
        const me_pu_result_t &q = res[(size_t)(i / nl) * nref + r];
        o.mv[0] = q.mv_sub[0];
        o.mv[1] = q.mv_sub[1];
        o.cost = q.cost_final;
    } else {
        o.mv[0] = o.mv[1] = 0;
        o.cost = 0xFFFFFFFFu;
    }
    out[i] = o;
}

ME_EXPORT int me_plan_best(me_plan_t p, me_pu_best_t *out, int8_t *best_ref, void *stream)
{
    if (!p || !out) { me_set_error("me_plan_best: bad argument", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    hipStream_t st = stream ? (hipStream_t)stream : (p->graph_stream ? p->graph_stream : p->own);
    const int n = p->n_ctu * p->s.npu;
    if (!p->d_bestc) ME_CHECK(hipMalloc((void **)&p->d_bestc, (size_t)n * p->s.nl * sizeof(me_pu_best_t)));
    hipLaunchKernelGGL(k_best, dim3((n * p->s.nl + 255) / 256), dim3(256), 0, st, p->d_res, p->d_best, p->d_bestc, n,
                       p->s.nref, p->s.nl);
    ME_CHECK(hipGetLastError());
    ME_CHECK(hipMemcpyAsync(out, p->d_bestc, (size_t)n * p->s.nl * sizeof(me_pu_best_t), hipMemcpyDeviceToHost, st));
    if (best_ref) ME_CHECK(hipMemcpyAsync(best_ref, p->d_best, (size_t)n * p->s.nl, hipMemcpyDeviceToHost, st));
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

ME_EXPORT int me_pinned_alloc(void **out, uint64_t bytes)
{
    if (!out) return -1;
    if (me_hip_init(-1)) return -1;
    ME_CHECK(hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault));
    return 0;
}

ME_EXPORT void me_pinned_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

ME_EXPORT int me_plan_autotune(me_frame_cfg_t *cfg, me_ref_t const *refs, const me_pel *cur, int stride, int iters,
                               me_tune_result_t *log, int log_cap, int *n_log)
{
    if (!cfg || !refs || !cur || iters < 1) {
        me_set_error("me_plan_autotune: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    const int span = 2 * cfg->search_range + 1;
    const int full = (span + 3) & ~3;
    const int bands[] = {4, 8, 16, 32}, waves[] = {1, 2, 4, 8}, cols[] = {0, 64, 32};
    double best = 1e30;
    me_frame_cfg_t c = *cfg, pick = *cfg;
    int n = 0;
    // the integer search's time: fill (atomic layouts only) + search
    auto time_int = [&](const me_frame_cfg_t &cc, double *ms) -> int {
        me_plan_t p;
        if (me_plan_create(&p, &cc)) return 1;   // LDS window too large
        int rc = me_plan_set_cur(p, cur, stride, 0, nullptr);
        if (!rc) rc = me_plan_profile(p, refs, 0, 1, ms);          // warm-up
        if (!rc) rc = me_plan_profile(p, refs, 0, iters, ms);
        me_plan_destroy(p);
        return rc ? -1 : 0;
    };
    c.int_grid = 0;
    for (int w : waves)
        for (int b : bands)
            for (int dc : cols) {
                if (dc && dc >= full) continue;   // same as "all columns"
                if (b > span) continue;
                c.band_rows = b;
                c.waves = w;
                c.dx_cols = dc;
                double ms[4] = {0, 0, 0, 0};
                const int rc = time_int(c, ms);
                if (rc > 0) continue;
                if (rc < 0) return -1;
                const double t = ms[1] + ms[2];
                if (log && n < log_cap) log[n] = {b, w, dc, t, c.sp_groups, ms[3], c.int_grid};
                n++;
                if (t < best) { best = t; pick = c; }
            }
    if (best < 1e30) {   // the fastest schedule with one workgroup per band
        c = pick;
        c.int_grid = 1;
        double ms[4] = {0, 0, 0, 0};
        const int rc = time_int(c, ms);
        if (rc < 0) return -1;
        if (rc == 0) {
            const double t = ms[1] + ms[2];
            if (log && n < log_cap) log[n] = {c.band_rows, c.waves, c.dx_cols, t, c.sp_groups, ms[3], 1};
            n++;
            if (t < best) { best = t; pick = c; }
        }
    }
    if (best >= 1e30) {
        if (n_log) *n_log = n < log_cap ? n : log_cap;
        me_set_error("me_plan_autotune: no schedule fits", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    // sub-pel position groups, with the integer schedule just picked
    c = pick;
    double best_sp = 1e30;
    for (int g : {1, 2, 4, 8}) {
        c.sp_groups = g;
        me_plan_t p;
        if (me_plan_create(&p, &c)) continue;
        double ms[4] = {0, 0, 0, 0};
        int rc = me_plan_set_cur(p, cur, stride, 0, nullptr);
        if (!rc) rc = me_plan_profile(p, refs, 0, 1, ms);          // warm-up
        if (!rc) rc = me_plan_profile(p, refs, 0, iters, ms);
        me_plan_destroy(p);
        if (rc) return -1;
        if (log && n < log_cap) log[n] = {c.band_rows, c.waves, c.dx_cols, ms[1] + ms[2], g, ms[3], c.int_grid};
        n++;
        if (ms[3] < best_sp) { best_sp = ms[3]; pick.sp_groups = g; }
    }
    if (n_log) *n_log = n < log_cap ? n : log_cap;
    cfg->band_rows = pick.band_rows;
    cfg->waves = pick.waves;
    cfg->dx_cols = pick.dx_cols;
    cfg->int_grid = pick.int_grid;
    cfg->sp_groups = pick.sp_groups;
    return 0;
}

ME_EXPORT int me_plan_run_tz(me_plan_t p, me_ref_t const *refs, const int16_t *mvp, const int16_t *mv_ipel_prev,
                             const me_tz_cfg_t *tz, void *stream)
{
    if (!p || !refs || !mvp || !tz || tz->gop_size < 1 || tz->mvr < 0 || tz->mvr > 4 || tz->cost_type < 0 ||
        tz->cost_type > 1) {
        me_set_error("me_plan_run_tz: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (p->cfg.pic_w % 4) {   // the search reads the original 4 aligned samples at a time
        me_set_error("me_plan_run_tz: picture width must be a multiple of 4", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    for (int r = 0; r < p->s.nref; r++)
        if (!refs[r] || refs[r]->w != p->cfg.pic_w || refs[r]->h != p->cfg.pic_h) {
            me_set_error("me_plan_run_tz: reference size differs from the plan", hipErrorInvalidValue, __FILE__, __LINE__);
            return -1;
        }
    hipStream_t st = stream ? (hipStream_t)stream : (p->graph_stream ? p->graph_stream : p->own);
    const size_t n = (size_t)p->n_ctu * p->s.npu * p->s.nref * 2;
    if (!p->d_mvp) ME_CHECK(hipMalloc((void **)&p->d_mvp, 2 * n * sizeof(s16)));
    if (plan_set_pic(p, refs, st)) return -1;
    ME_CHECK(hipMemcpyAsync(p->d_mvp, mvp, n * sizeof(s16), hipMemcpyHostToDevice, st));
    if (mv_ipel_prev) ME_CHECK(hipMemcpyAsync(p->d_mvp + n, mv_ipel_prev, n * sizeof(s16), hipMemcpyHostToDevice, st));
    JobSet js;
    js.pics = p->d_pic;
    js.jobs = nullptr;
    js.n_jobs = p->n_ctu;
    if (me_tz_enqueue(&p->s, js, p->d_mvp, mv_ipel_prev ? p->d_mvp + n : nullptr, tz, p->d_res, st)) return -1;
    const int npus = p->n_ctu * p->s.npu;
    hipLaunchKernelGGL(k_select, dim3((npus + 255) / 256), dim3(256), 0, st, p->d_res, p->d_best, npus, p->s.nref,
                       p->s.nref - p->s.nref_l1, p->s.nl);
    ME_CHECK(hipGetLastError());
    // the host arrays may be reused once the call returns
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

// analyze_bi's search (inter.c:972-1033) for every PU with both lists.
ME_EXPORT int me_plan_run_bi(me_plan_t p, me_ref_t const *refs, const int16_t *mvp, int16_t *mv_scale,
                             me_bi_pu_t *pus, const me_tz_cfg_t *tz, void *stream)
{
    if (!p || !refs || !mvp || !mv_scale || !pus || !tz || tz->mvr < 0 || tz->mvr > 4 || tz->cost_type < 0 ||
        tz->cost_type > 1) {
        me_set_error("me_plan_run_bi: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (p->s.nref_l1 < 1 || p->s.nref_l1 >= p->s.nref || p->cfg.pic_w % 4) {
        me_set_error("me_plan_run_bi: the plan needs references in both lists and a width that is a multiple of 4",
                     hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    const int n0 = p->s.nref - p->s.nref_l1;
    const size_t npus = (size_t)p->n_ctu * p->s.npu;
    for (size_t k = 0; k < npus; k++) {
        const me_bi_pu_t &q = pus[k];
        if (q.active && (q.lidx_ref < 0 || q.lidx_ref > 1 || q.refi[0] < 0 || q.refi[0] >= n0 || q.refi[1] < 0 ||
                         q.refi[1] >= p->s.nref_l1)) {
            me_set_error("me_plan_run_bi: PU reference index or list out of range", hipErrorInvalidValue, __FILE__,
                         __LINE__);
            return -1;
        }
    }
    for (int r = 0; r < p->s.nref; r++)
        if (!refs[r] || refs[r]->w != p->cfg.pic_w || refs[r]->h != p->cfg.pic_h) {
            me_set_error("me_plan_run_bi: reference size differs from the plan", hipErrorInvalidValue, __FILE__,
                         __LINE__);
            return -1;
        }
    hipStream_t st = stream ? (hipStream_t)stream : (p->graph_stream ? p->graph_stream : p->own);
    const size_t n = npus * p->s.nref * 2;
    int nsel = 0;
    for (int k = 0; k < ME_NSHAPES; k++) nsel += p->s.base[k] >= 0;
    if (!p->d_mvp) ME_CHECK(hipMalloc((void **)&p->d_mvp, 2 * n * sizeof(s16)));
    if (!p->d_bipu) ME_CHECK(hipMalloc((void **)&p->d_bipu, npus * sizeof(me_bi_pu_t)));
    if (!p->d_bist) ME_CHECK(hipMalloc(&p->d_bist, me_bi_state_bytes(npus)));
    // +64 B: com_had reads tile rows as dwords, up to 3 bytes past a row's end
    if (!p->d_orgbi) ME_CHECK(hipMalloc((void **)&p->d_orgbi, (size_t)p->n_ctu * nsel * 16384 * sizeof(pel) + 64));
    if (plan_set_pic(p, refs, st)) return -1;
    ME_CHECK(hipMemcpyAsync(p->d_mvp, mvp, n * sizeof(s16), hipMemcpyHostToDevice, st));
    ME_CHECK(hipMemcpyAsync(p->d_mvp + n, mv_scale, n * sizeof(s16), hipMemcpyHostToDevice, st));
    ME_CHECK(hipMemcpyAsync(p->d_bipu, pus, npus * sizeof(me_bi_pu_t), hipMemcpyHostToDevice, st));
    JobSet js;
    js.pics = p->d_pic;
    js.jobs = nullptr;
    js.n_jobs = p->n_ctu;
    BiBufs bb;
    bb.pus = p->d_bipu;
    bb.state = (BiState *)p->d_bist;
    bb.mv_scale = p->d_mvp + n;
    bb.orgbi = p->d_orgbi;
    if (me_bi_enqueue(&p->s, js, p->d_mvp, tz, p->d_res, bb, st)) return -1;
    ME_CHECK(hipMemcpyAsync(pus, p->d_bipu, npus * sizeof(me_bi_pu_t), hipMemcpyDeviceToHost, st));
    ME_CHECK(hipMemcpyAsync(mv_scale, p->d_mvp + n, n * sizeof(s16), hipMemcpyDeviceToHost, st));
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

// analyze_smvd's search (inter.c:1083-1255) for every PU.
ME_EXPORT int me_plan_run_smvd(me_plan_t p, me_ref_t const *refs, const int16_t *mvp, const int16_t *mv_scale,
                               me_smvd_pu_t *pus, const me_tz_cfg_t *tz, void *stream)
{
    if (!p || !refs || !mvp || !mv_scale || !pus || !tz || tz->mvr < 0 || tz->mvr > 4 || tz->cost_type < 0 ||
        tz->cost_type > 1) {
        me_set_error("me_plan_run_smvd: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (p->s.nref_l1 < 1 || p->s.nref_l1 >= p->s.nref) {
        me_set_error("me_plan_run_smvd: the plan needs references in both lists", hipErrorInvalidValue, __FILE__,
                     __LINE__);
        return -1;
    }
    for (int r = 0; r < p->s.nref; r++)
        if (!refs[r] || refs[r]->w != p->cfg.pic_w || refs[r]->h != p->cfg.pic_h) {
            me_set_error("me_plan_run_smvd: reference size differs from the plan", hipErrorInvalidValue, __FILE__,
                         __LINE__);
            return -1;
        }
    hipStream_t st = stream ? (hipStream_t)stream : (p->graph_stream ? p->graph_stream : p->own);
    const size_t npus = (size_t)p->n_ctu * p->s.npu;
    const size_t n = npus * p->s.nref * 2;
    if (!p->d_mvp) ME_CHECK(hipMalloc((void **)&p->d_mvp, 2 * n * sizeof(s16)));
    if (!p->d_smvd) ME_CHECK(hipMalloc((void **)&p->d_smvd, npus * sizeof(me_smvd_pu_t)));
    if (plan_set_pic(p, refs, st)) return -1;
    ME_CHECK(hipMemcpyAsync(p->d_mvp, mvp, n * sizeof(s16), hipMemcpyHostToDevice, st));
    ME_CHECK(hipMemcpyAsync(p->d_mvp + n, mv_scale, n * sizeof(s16), hipMemcpyHostToDevice, st));
    ME_CHECK(hipMemcpyAsync(p->d_smvd, pus, npus * sizeof(me_smvd_pu_t), hipMemcpyHostToDevice, st));
    JobSet js;
    js.pics = p->d_pic;
    js.jobs = nullptr;
    js.n_jobs = p->n_ctu;
    if (me_smvd_enqueue(&p->s, js, p->d_mvp, p->d_mvp + n, tz, p->d_smvd, st)) return -1;
    ME_CHECK(hipMemcpyAsync(pus, p->d_smvd, npus * sizeof(me_smvd_pu_t), hipMemcpyDeviceToHost, st));
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

// loka_estimate_coding_cost (lookahead.c:82-208), the inter part: every
// whole 16x16 unit searched by me_search_tz from MVP 0 at integer precision
// with the lookahead's lambda (lookahead.c:85-110), then its com_had costs.
ME_EXPORT int me_plan_lookahead(me_plan_t p, me_ref_t const *refs, const me_tz_cfg_t *tz, int is_padding,
                                uint32_t *unit_cost, void *stream)
{
    if (!p || !refs || !tz || !unit_cost || tz->gop_size < 1) {
        me_set_error("me_plan_lookahead: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (p->s.base[S16x16] < 0 || p->cfg.pic_w % 4 || p->cfg.pic_w < 32 || p->cfg.pic_h < 32) {
        me_set_error("me_plan_lookahead: the plan needs the 16x16 shape and a picture of at least 32x32 with a "
                     "width that is a multiple of 4", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    for (int r = 0; r < p->s.nref; r++)
        if (!refs[r] || refs[r]->w != p->cfg.pic_w || refs[r]->h != p->cfg.pic_h) {
            me_set_error("me_plan_lookahead: reference size differs from the plan", hipErrorInvalidValue, __FILE__,
                         __LINE__);
            return -1;
        }
    hipStream_t st = stream ? (hipStream_t)stream : (p->graph_stream ? p->graph_stream : p->own);
    const int units = (p->cfg.pic_w / 16) * (p->cfg.pic_h / 16);
    if (!p->d_loka) ME_CHECK(hipMalloc((void **)&p->d_loka, (size_t)units * sizeof(u32)));
    if (plan_set_pic(p, refs, st)) return -1;
    me_tz_cfg_t t = *tz;
    t.mvr = 2;                       // curr_mvr = 2: integer MVs, no sub-pel
    t.adaptive_raster_range = 0;
    t.cost_type = 0;
    const u32 lambda = (u32)floor(65536.0 * sqrt(1.43631 * pow(2.0, (32 - 16.0) / 4.0)));
    JobSet js;
    js.pics = p->d_pic;
    js.jobs = nullptr;
    js.n_jobs = p->n_ctu;
    if (me_tz_enqueue(&p->s, js, nullptr, nullptr, &t, p->d_res, st, S16x16, is_padding ? 1 : 0, lambda)) return -1;
    if (me_loka_enqueue(&p->s, p->d_pic, p->d_res, p->cfg.pic_w, p->cfg.pic_h, p->d_loka, st)) return -1;
    ME_CHECK(hipMemcpyAsync(unit_cost, p->d_loka, (size_t)units * sizeof(u32), hipMemcpyDeviceToHost, st));
    ME_CHECK(hipStreamSynchronize(st));
    return 0;
}

ME_EXPORT int me_plan_mc_size(me_plan_t p, uint64_t *luma_elems, uint64_t *chroma_elems)
{
    if (!p) return -1;
    size_t ly, lc;
    me_mc_sizes(&p->s, &ly, &lc);
    if (luma_elems) *luma_elems = (uint64_t)ly * p->n_ctu;
    if (chroma_elems) *chroma_elems = (uint64_t)lc * p->n_ctu;
    return 0;
}

ME_EXPORT int me_plan_mc(me_plan_t p, me_ref_t const *refs, me_pel *pred_y, me_pel *pred_u, me_pel *pred_v,
                         void *stream)
{
    if (!p || !refs || (!pred_y && (pred_u || pred_v))) {
        me_set_error("me_plan_mc: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    size_t ly, lc;
    me_mc_sizes(&p->s, &ly, &lc);
    ly *= p->n_ctu;
    lc *= p->n_ctu;
    for (int r = 0; r < p->s.nref; r++)
        if (!refs[r] || refs[r]->w != p->cfg.pic_w || refs[r]->h != p->cfg.pic_h || (lc && !refs[r]->chroma)) {
            me_set_error("me_plan_mc: reference missing, of another size, or without chroma", hipErrorInvalidValue,
                         __FILE__, __LINE__);
            return -1;
        }
    hipStream_t st = stream ? (hipStream_t)stream : (p->graph_stream ? p->graph_stream : p->own);
    if (!pred_y) {
        if (!p->d_pred) ME_CHECK(hipMalloc((void **)&p->d_pred, (ly + 2 * lc) * sizeof(pel)));
        pred_y = p->d_pred;
        pred_u = lc ? p->d_pred + ly : nullptr;
        pred_v = lc ? p->d_pred + ly + lc : nullptr;
    }
    if (plan_set_pic(p, refs, st)) return -1;
    JobSet js;
    js.pics = p->d_pic;
    js.jobs = nullptr;
    js.n_jobs = p->n_ctu;
    return me_mc_enqueue(&p->s, js, p->d_res, p->d_best, pred_y, pred_u, pred_v, st);
}

ME_EXPORT int me_plan_mc_download(me_plan_t p, me_pel *pred_y, me_pel *pred_u, me_pel *pred_v)
{
    if (!p || !p->d_pred) { me_set_error("me_plan_mc_download: no plan-owned predictions", hipErrorInvalidValue, __FILE__, __LINE__); return -1; }
    size_t ly, lc;
    me_mc_sizes(&p->s, &ly, &lc);
    ly *= p->n_ctu;
    lc *= p->n_ctu;
    ME_CHECK(hipStreamSynchronize(p->graph_stream ? p->graph_stream : p->own));
    ME_CHECK(hipStreamSynchronize(p->own));
    if (pred_y) ME_CHECK(hipMemcpy(pred_y, p->d_pred, ly * sizeof(pel), hipMemcpyDeviceToHost));
    if (pred_u && lc) ME_CHECK(hipMemcpy(pred_u, p->d_pred + ly, lc * sizeof(pel), hipMemcpyDeviceToHost));
    if (pred_v && lc) ME_CHECK(hipMemcpy(pred_v, p->d_pred + ly + lc, lc * sizeof(pel), hipMemcpyDeviceToHost));
    return 0;
}

ME_EXPORT int me_plan_num_ctus(me_plan_t p) { return p ? p->n_ctu : 0; }
ME_EXPORT int me_plan_num_pus(me_plan_t p) { return p ? p->s.npu : 0; }

ME_EXPORT void me_plan_destroy(me_plan_t p)
{
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->own) (void)hipStreamDestroy(p->own);
    me_search_free(&p->s);
    (void)hipFree(p->d_loka);
    (void)hipFree(p->d_bipu);
    (void)hipFree(p->d_bist);
    (void)hipFree(p->d_orgbi);
    (void)hipFree(p->d_smvd);
    (void)hipFree(p->d_bestc);
    (void)hipFree(p->d_cur);
    (void)hipFree(p->d_pic);
    (void)hipFree(p->d_pred);
    (void)hipFree(p->d_mvp);
    (void)hipFree(p->d_keys);
    (void)hipFree(p->d_res);
    (void)hipFree(p->d_best);
    free(p);
}

// ---------------------------------------------------------------------------
// Frame search without allocation
// ---------------------------------------------------------------------------

// Everything me_frame_search_noalloc keeps on the device -- the picture
// descriptor, the integer-pel rate row, the MV-bits table, the sub-pel shape
// tables and the key of every (CTU, PU, reference) -- lives in this array.
// Each GPU holds its own copy until the process ends; host code reaches it
// only through the per-device address from hipGetSymbolAddress.
#define ME_NOALLOC_SCRATCH (8u << 20)
static __device__ __align__(256) unsigned char scratch[ME_NOALLOC_SCRATCH];

#define ME_NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[ME_NOALLOC_MAX_DEVICES];

// The scratch array's address on the current device, looked up once per
// device (with that device current).
static unsigned char *scratch_address(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= ME_NOALLOC_MAX_DEVICES) {
        me_set_error("me_frame_search_noalloc: no current device", hipErrorInvalidDevice, __FILE__, __LINE__);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        unsigned char *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            me_set_error("me_frame_search_noalloc: hipGetSymbolAddress", e, __FILE__, __LINE__);
            return nullptr;
        }
        g_scratch_dev[dev] = ptr;
    }
    return g_scratch_dev[dev];
}

// Device tables of a search, filled in place (the values me_search_init
// uploads): the MV-bits table in its closed form, the rate row, the sub-pel
// shape table, workgroup -> shape, and the picture descriptor.
struct TableArgs {
    SubShape shp[ME_NSHAPES];
    PicDesc pic;
    int n_shp, sub_blocks, span, R, tab_off;
    u32 lambda;
    u32 *rate;
    u8 *tab_base;
    SubShape *d_shp;
    u8 *blk_shape;
    PicDesc *d_pic;
};

__device__ __forceinline__ u32 mvbits_at(int v, int off)
{
    const int a = v < 0 ? -v : v;
    if (a == 0) return 1;
    if (a >= off) return 0;
    return 2u * (u32)(31 - __clz(a + 1) + 1);   // 2 * (floor(log2(|v| + 1)) + 1)
}

__global__ void k_search_tables(TableArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= 2 * a.tab_off) a.tab_base[i] = (u8)mvbits_at(i - a.tab_off, a.tab_off);
    if (i < a.span) a.rate[i] = mvbits_at(4 * (i - a.R), a.tab_off) * a.lambda;
    else if (i < a.span + ME_RATE_PAD) a.rate[i] = 0;
    if (i < a.n_shp) a.d_shp[i] = a.shp[i];
    if (i < a.sub_blocks) {
        int k = 0;
        while (k + 1 < a.n_shp && a.shp[k + 1].blk0 <= i) k++;
        a.blk_shape[i] = (u8)k;
    }
    if (i == 0) *a.d_pic = a.pic;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

ME_EXPORT int me_frame_search_noalloc(const me_frame_cfg_t *cfg, const me_pel *cur, me_ref_t const *refs,
                                      me_pu_result_t *res, int8_t *best_ref, void *stream)
{
    if (!cfg || !cur || !refs || !res || !best_ref || cfg->pic_w <= 0 || cfg->pic_h <= 0 ||
        cfg->pic_w > 128 * 0xFFFF) {
        me_set_error("me_frame_search_noalloc: bad argument", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    if (me_hip_init(-1)) return -1;
    const int T = cfg->pic_w > cfg->pic_h ? cfg->pic_w : cfg->pic_h;
    if (cfg->search_range > T - 4) {
        me_set_error("me_frame_search_noalloc: search range exceeds max(pic_w, pic_h) - 4", hipErrorInvalidValue,
                     __FILE__, __LINE__);
        return -1;
    }
    MeSearch s;
    SubShape shp[ME_NSHAPES];
    int n_shp = 0;
    if (search_layout(&s, cfg->search_range, cfg->lambda_mv, cfg->shape_mask, cfg->num_refs, cfg->num_refs_l1, T,
                      cfg->band_rows, cfg->waves, cfg->dx_cols, cfg->sp_groups, shp, &n_shp))
        return -1;
    for (int r = 0; r < s.nref; r++)
        if (!refs[r] || refs[r]->w != cfg->pic_w || refs[r]->h != cfg->pic_h) {
            me_set_error("me_frame_search_noalloc: reference missing or of another size", hipErrorInvalidValue,
                         __FILE__, __LINE__);
            return -1;
        }
    const int ctu_cols = (cfg->pic_w + 127) / 128;
    const int n_ctu = ctu_cols * ((cfg->pic_h + 127) / 128);
    // scratch layout: descriptor, rate row, MV-bits table, shape table,
    // workgroup -> shape, keys
    const size_t o_rate = align256(sizeof(PicDesc));
    const size_t o_tab = o_rate + align256((size_t)(s.span + ME_RATE_PAD) * sizeof(u32));
    const size_t o_shp = o_tab + align256((size_t)2 * s.tab_off + 1);
    const size_t o_blk = o_shp + align256(sizeof(SubShape) * ME_NSHAPES);
    const size_t o_keys = o_blk + align256((size_t)s.sub_blocks);
    const size_t need = o_keys + (size_t)n_ctu * s.npu * s.nref * sizeof(u64);
    if (need > ME_NOALLOC_SCRATCH) {
        char msg[160];
        snprintf(msg, sizeof(msg), "me_frame_search_noalloc: the frame needs %zu bytes of scratch, the maximum is %u",
                 need, ME_NOALLOC_SCRATCH);
        me_set_error(msg, hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    unsigned char *base = scratch_address();
    if (!base) return -1;
    s.d_rate = (u32 *)(base + o_rate);
    s.d_tab_base = base + o_tab;
    s.d_shp = (SubShape *)(base + o_shp);
    s.d_blk_shape = base + o_blk;
    u64 *keys = (u64 *)(base + o_keys);
    hipStream_t st = (hipStream_t)stream;

    TableArgs a;
    memset(&a, 0, sizeof(a));
    memcpy(a.shp, shp, sizeof(shp));
    a.pic.cur = cur;
    for (int r = 0; r < s.nref; r++) {
        a.pic.planes[r] = refs[r]->planes;
        a.pic.chroma[r] = refs[r]->chroma;
    }
    a.pic.plane_elems = refs[0]->plane_elems;
    a.pic.ref_stride = refs[0]->stride;
    a.pic.pic_w = cfg->pic_w;
    a.pic.pic_h = cfg->pic_h;
    a.pic.ctu_cols = ctu_cols;
    a.n_shp = n_shp;
    a.sub_blocks = s.sub_blocks;
    a.span = s.span;
    a.R = s.R;
    a.tab_off = s.tab_off;
    a.lambda = s.lambda;
    a.rate = s.d_rate;
    a.tab_base = s.d_tab_base;
    a.d_shp = s.d_shp;
    a.blk_shape = s.d_blk_shape;
    a.d_pic = (PicDesc *)base;
    int n = 2 * s.tab_off + 1;
    if (s.span + ME_RATE_PAD > n) n = s.span + ME_RATE_PAD;
    if (s.sub_blocks > n) n = s.sub_blocks;
    hipLaunchKernelGGL(k_search_tables, dim3((n + 255) / 256), dim3(256), 0, st, a);
    ME_CHECK(hipGetLastError());
    JobSet js;
    js.pics = (const PicDesc *)base;
    js.jobs = nullptr;
    js.n_jobs = n_ctu;
    KeyLayout kl;
    me_search_keys(&s, n_ctu, 1, &kl);   // one key per (CTU, PU, reference), merged with atomics
    return me_search_enqueue(&s, js, kl, keys, res, best_ref, st, nullptr);
}
