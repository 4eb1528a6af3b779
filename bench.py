"""Benchmark: batched frame motion estimation, 1080p 8-bit, 2 references,
all 29 PU shapes, full search +-64 then quarter-pel refinement (the
1080p-all-PU workload).  Prints one JSON line.

A step is one frame: re-pad + re-interpolate the newest reference (each frame
of an encoder contributes one new reference picture) and run integer full
search, sub-pel refinement and reference choice for every PU of the frame
against both references, as one hipGraph launch.  Inputs are resident in HBM
before timing starts.

Beside the batched number the line carries:
  per_call     the 1080p 16x16 workload (+-32, 1 reference) on the same frames:
               the per-call drop-in path (host driver, every SAD/SATD/filter a
               synchronous HIP call) on whole CTUs, the batched plan and the
               reference CPU path, each in CTUs/s
  end_to_end   upload of the current picture + the frame + compact results
               into pinned host memory, per frame
  cpu_baseline the reference uavs3e kernels (C+SSE+AVX2 table, compiled from
               the reference sources into oracle/_ref) on the host's cores, on
               a bounded sample of CTUs, else the oracle restatement
  roofline     the integer-search kernel against HBM bandwidth, plus its VALU
               fraction
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_OPS = 256 * 64 * 2.4e9     # lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz


def pct(a, q):
    return float(np.percentile(np.asarray(a, dtype=np.float64), q))


def host_cpu():
    """(model name, logical CPUs, CPUs this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return model, os.cpu_count() or 1, allowed


def cpu_threads():
    """Worker threads of the CPU baselines: one per CPU this process may use,
    at most 16 (the CPU share of one GPU on the GPU box)."""
    return max(1, min(host_cpu()[2], 16))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="hd_allpu")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-runs", type=int, default=3, help="passes of the CPU baseline over the frame")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="after the timed region, run a plain byte copy of this many bytes each way "
                         "(three launches, torch bitwise_not) to scale HBM counters against")
    ap.add_argument("--no-percall", action="store_true", help="skip the 1080p-16x16 leg (batched vs per-call vs CPU)")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end leg (upload, frame, results)")
    ap.add_argument("--valu-peak", action="store_true",
                    help="also measure the issue rate of the search's VALU instructions (microbenchmarks)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-tz", action="store_true", help="skip the me_search_tz leg")
    ap.add_argument("--no-lookahead", action="store_true", help="skip the lookahead leg (its me_search_tz launches too)")
    ap.add_argument("--serve-frames", type=int, default=32, help="frames per stream in the serving leg")
    ap.add_argument("--serve-rate", type=float, default=40000.0, help="offered load of the serving leg, CTUs/s")
    ap.add_argument("--no-tune", action="store_true", help="use the default schedule instead of the autotuner's")
    ap.add_argument("--tune-iters", type=int, default=3)
    ap.add_argument("--tune-log", default="", help="write the autotuner's timing table to this JSON file")
    ap.add_argument("--band-rows", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--dx-cols", type=int, default=0)
    ap.add_argument("--sp-groups", type=int, default=0, help="sub-pel position groups (0: 1, or the autotuner's)")
    ap.add_argument("--int-grid", type=int, default=0,
                    help="integer-search grid: 0 persistent where the frame fills the GPU, 1 one workgroup per band")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the ME path runs on a single MI355X "
              "(sharding across GPUs is designed, not built). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import uavs3e_amd as ua
    from uavs3e_amd import workloads as wl

    work = wl.ALL[args.workload]()
    bd = work.bd
    t0 = time.time()
    cur, refs = work.frames()
    gen_s = time.time() - t0
    print("[bench] %s frames generated in %.1fs" % (work.name, gen_s), file=sys.stderr, flush=True)

    hip = ua.MeHip(bd)
    dref = [hip.ref_create(r) for r in refs]
    bi = work.num_refs_l1 > 0
    if bi:   # bi-directional frames: ME of both lists, then MC (luma + 4:2:0 chroma, bi-average)
        for d, (u, v) in zip(dref, work.ref_chroma()):
            hip.ref_load_chroma(d, u, v)
    # integer-search schedule: the autotuner's pick unless given on the command line
    tune = None
    if not (args.band_rows or args.waves or args.dx_cols) and not args.no_tune:
        t0 = time.perf_counter()
        pick, table = hip.autotune(work.w, work.h, work.R, work.lam, work.mask, dref, cur, work.num_refs_l1,
                                   iters=args.tune_iters)
        # the integer-search candidates come first (sp_groups 0: the
        # default), then the sub-pel position groups with the picked schedule
        it = [t for t in table if t["sp_groups"] == 0]
        sp = [t for t in table if t["sp_groups"] > 0]
        tune = {"pick": pick, "candidates": len(table), "seconds": time.perf_counter() - t0,
                "fastest_ms": min(t["int_search_ms"] for t in it),
                "default_ms": next((t["int_search_ms"] for t in it
                                    if t["band_rows"] == 16 and t["waves"] == 4 and t["dx_cols"] == 0
                                    and t["int_grid"] == 0), None),
                "per_band_grid_ms": next((t["int_search_ms"] for t in it if t["int_grid"] == 1), None),
                "subpel_ms_by_groups": {str(t["sp_groups"]): t["subpel_ms"] for t in sp}}
        args.band_rows, args.waves, args.dx_cols = pick["band_rows"], pick["waves"], pick["dx_cols"]
        args.sp_groups, args.int_grid = pick["sp_groups"], pick["int_grid"]
        print("[bench] autotune: %s (%.3f ms; default %s ms) over %d schedules in %.1fs" % (
            pick, tune["fastest_ms"], tune["default_ms"], len(table), tune["seconds"]), file=sys.stderr, flush=True)
        if args.tune_log:
            with open(args.tune_log, "w") as f:
                json.dump({"workload": work.name, "pick": pick, "table": table}, f, indent=1)
    plan = hip.plan_create(work.w, work.h, work.R, work.lam, work.mask, work.num_refs,
                           band_rows=args.band_rows, waves=args.waves, dx_cols=args.dx_cols,
                           num_refs_l1=work.num_refs_l1, sp_groups=args.sp_groups, int_grid=args.int_grid)
    hip.plan_set_cur(plan, cur)
    n_ctu = hip.lib.me_plan_num_ctus(plan)
    npu = hip.lib.me_plan_num_pus(plan)

    def step():
        hip.plan_run(plan, dref, prep)
        if bi:
            hip.plan_mc(plan, dref)

    # ---- batched: one graph launch per frame --------------------------------
    prep = 1   # the newest reference is rebuilt inside every step
    for _ in range(args.warmup):
        step()
    hip.plan_results(plan, work.num_refs)
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        hip.lib.me_plan_results(plan, None, None)   # waits for this frame
        lat.append((time.perf_counter() - s0) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_step = elapsed * 1e3 / args.steps
    ctus_per_s = n_ctu * args.steps / elapsed
    print("[bench] batched: %.3f ms/frame, %.0f CTUs/s" % (ms_step, ctus_per_s), file=sys.stderr, flush=True)

    # candidate evaluations per frame (integer full search, all valid PUs and refs)
    res, best = hip.plan_results(plan, work.num_refs)

    # ---- end to end: current picture up, frame, compact results down --------
    e2e = None
    if not args.no_e2e:
        nl = 2 if bi else 1
        shape_b = (n_ctu, npu, nl) if bi else (n_ctu, npu)
        pin_cur = ua.Pinned(hip, cur.shape, cur.dtype)
        pin_cur.array[...] = cur
        pin_out = ua.Pinned(hip, shape_b, ua.PU_BEST)
        pin_best = ua.Pinned(hip, shape_b, np.int8)

        def e2e_step():
            hip.plan_set_cur_raw(plan, pin_cur.ptr.value, work.w)
            step()
            hip.plan_best(plan, pin_out.array, pin_best.array)

        for _ in range(2):
            e2e_step()
        e_lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s0 = time.perf_counter()
            e2e_step()
            e_lat.append((time.perf_counter() - s0) * 1e3)
        e_el = time.perf_counter() - t0
        # the compact results are the full results' best references
        bsel = np.where(best >= 0, best, 0).astype(np.int64)
        if bi:
            full = np.take_along_axis(res, bsel, axis=2)
        else:
            full = np.take_along_axis(res, bsel[..., None], axis=2)[..., 0]
        ok = bool(np.array_equal(pin_best.array, best) and
                  np.array_equal(np.where(best >= 0, full["cost_final"], 0xFFFFFFFF), pin_out.array["cost"]) and
                  np.array_equal(np.where((best >= 0)[..., None], full["mv_sub"], 0), pin_out.array["mv"]))
        e2e = {"ms_per_frame": e_el * 1e3 / args.steps, "ctus_per_s": n_ctu * args.steps / e_el,
               "p50_ms": pct(e_lat, 50), "p99_ms": pct(e_lat, 99),
               "upload_bytes": int(cur.nbytes), "download_bytes": int(pin_out.array.nbytes + pin_best.array.nbytes),
               "compact_equals_full_results": ok,
               "covers": "per frame, serially: the current picture from pinned host memory to HBM (me_plan_set_cur), "
                         "the frame's graph (reference rebuild + search), then per (PU, list) the best reference's "
                         "MV and final cost (8 B) and index (1 B) into pinned host memory (me_plan_best); "
                         "the headline value keeps the inputs resident and the results on the device"}
        for b_ in (pin_cur, pin_out, pin_best):
            b_.free()
        print("[bench] end to end: %.3f ms/frame (%.0f CTUs/s)" % (e2e["ms_per_frame"], e2e["ctus_per_s"]),
              file=sys.stderr, flush=True)
    valid = res["valid"] == 1
    geo = []
    for s in range(29):
        if work.mask >> s & 1:
            sw, sh = ua.SHAPES[s]
            geo += [(ox, oy) for oy in range(0, 128, sh) for ox in range(0, 128, sw)]
    geo = np.array(geo)
    cols = (work.w + 127) // 128
    T = max(work.w, work.h)
    ctu_xy = np.array([((c % cols) * 128, (c // cols) * 128) for c in range(n_ctu)])
    px = ctu_xy[:, None, 0] + geo[None, :, 0]
    py = ctu_xy[:, None, 1] + geo[None, :, 1]
    nx = np.minimum(np.minimum(work.w - 1 - px, T - 4), work.R) - np.maximum(np.maximum(-127 - px, -T + 4), -work.R) + 1
    ny = np.minimum(np.minimum(work.h - 1 - py, T - 4), work.R) - np.maximum(np.maximum(-127 - py, -T + 4), -work.R) + 1
    cands = int((np.clip(nx, 0, None) * np.clip(ny, 0, None) * valid[:, :, 0]).sum()) * work.num_refs
    shapes = [s for s in range(29) if work.mask >> s & 1]
    cand_per_s = cands * args.steps / elapsed

    # ---- stage timing with HIP events on the plan's stream -------------------
    hip.plan_profile(plan, dref, prep, iters=2)   # warm-up of the eager (non-graph) launches
    stage = hip.plan_profile(plan, dref, prep, iters=10)
    mc_ms = None
    if bi:
        t0 = time.perf_counter()
        for _ in range(3):
            hip.plan_mc(plan, dref)
        hip.lib.me_plan_results(plan, None, None)
        mc_ms = (time.perf_counter() - t0) * 1e3 / 3
    k_ms = stage[2]
    pel = 1 if bd == 8 else 2
    alg_bytes = work.w * work.h * pel * (1 + work.num_refs) + n_ctu * npu * work.num_refs * 8
    achieved_gbs = alg_bytes / (k_ms * 1e-3) / 1e9
    # algorithmic VALU work of the integer search: every candidate SAD built
    # from 4x4-block SADs plus one compare per PU candidate
    blocks4 = (work.w // 4) * (work.h // 4)
    span = 2 * work.R + 1
    # abs-diffs per lane-op of the widest SAD instruction at the bit depth:
    # v_sad_u8 (and v_qsad_pk_u16_u8 at a quarter of its rate) 4, v_sad_u16 2
    ad_per_op = 4 if bd == 8 else 2
    valu_ops = blocks4 * span * span * work.num_refs * 16 / ad_per_op + cands
    valu_frac = valu_ops / (k_ms * 1e-3) / VALU_PEAK_OPS
    valu_ops4 = blocks4 * span * span * work.num_refs * 16 / 4 + cands
    # HBM bytes per launch of the integer search, from the rocprofv3 counter
    # runs summarised in profiles/ (measured with the schedule they name)
    # (the entry of the same grid, the same schedule if there is one)
    traffic, traffic_sched = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % work.name)
    if os.path.exists(tpath):
        with open(tpath) as f:
            entries = json.load(f).get("entries", [])
        sched = {"band_rows": args.band_rows, "waves": args.waves, "dx_cols": args.dx_cols, "int_grid": args.int_grid}
        same = [e for e in entries if e["schedule"] == sched] or \
               [e for e in entries if e["schedule"]["int_grid"] == args.int_grid]
        if same:
            traffic, traffic_sched = same[0]["bytes_per_launch"], same[0]["schedule"]

    # ---- the reference's own search (me_search_tz) on the GPU ----------------
    search_tz = None
    if not args.no_tz and not bi:
        import ctypes
        mvp = np.zeros((n_ctu, npu, work.num_refs, 2), np.int16)   # MVP (0,0), as the full search
        tz = ua.TzCfg(8, work.R, 0, 0, 0, 8, (ctypes.c_int64 * 4)(*([7, 6, 4, 0][:work.num_refs] +
                                                                       [0] * (4 - work.num_refs))))
        hip.plan_run_tz(plan, dref, mvp, tz)     # warm-up
        t0 = time.perf_counter()
        for _ in range(3):
            hip.plan_run_tz(plan, dref, mvp, tz)
        tz_ms = (time.perf_counter() - t0) * 1e3 / 3
        tres, _ = hip.plan_results(plan, work.num_refs)
        tz_cpu = None
        from oracle.oracle import RefLib
        if RefLib.available(bd) and not args.no_cpu:
            rl = RefLib(bd, level=2)
            rl.set_pinning(1)
            tpocs = [7, 6, 4, 0][:work.num_refs]
            tpics = [rl.pic_create(r, p_) for r, p_ in zip(refs, tpocs)]
            nthr = cpu_threads()
            _, _, t1 = rl.tz_frame(cur, tpics, 8, work.R, 8, work.lam, work.mask, 0, 1, threads=1)
            n_s = int(min(n_ctu, max(nthr, 10e3 * nthr / max(t1, 1e-3))))
            cres_t, cbest_t, cms_t = rl.tz_frame(cur, tpics, 8, work.R, 8, work.lam, work.mask, 0, n_s, threads=nthr)
            for q in tpics:
                rl.pic_free(q)
            g_ = tres[:n_s]
            c_ = cres_t[:n_s]
            v_ = c_["valid"] == 1
            same = bool(np.array_equal(g_["valid"], c_["valid"]) and
                        np.array_equal(g_["mv_sub"][v_], c_["mv_sub"][v_]) and
                        np.array_equal(g_["mv_int"][v_], c_["mv_int"][v_]) and
                        np.array_equal(g_["cost_final"][v_], c_["cost_final"][v_]))
            tz_cpu = {"value": n_s / (cms_t * 1e-3), "unit": "CTUs/s", "cores": nthr, "kind": "reference",
                      "sample": "CTUs 0..%d of %s: the reference's me_search_tz (C+SSE4.2+AVX2 table), same "
                                "MVP / GOP / POCs / range, %d pinned threads, %.2f s; bit-exact with the GPU: %s"
                                % (n_s - 1, work.name, nthr, cms_t * 1e-3, "yes" if same else "NO")}
            print("[bench] me_search_tz CPU: %.1f CTUs/s" % tz_cpu["value"], file=sys.stderr, flush=True)
        search_tz = {"ms_per_frame": tz_ms, "ctus_per_s": n_ctu / (tz_ms * 1e-3), "cpu_reference": tz_cpu,
                     "note": "me_search_tz (diamond, raster, refinement, sub-pel) for every PU and reference, "
                             "MVP (0,0), GOP 8, max search range %d, AMVR 0, SATD sub-pel; one wave per "
                             "(PU, reference); includes the MVP upload" % work.R,
                     "pus_searched": int((tres["valid"] == 1).sum())}
        print("[bench] me_search_tz: %.2f ms/frame" % tz_ms, file=sys.stderr, flush=True)

    # ---- bi-prediction ME (analyze_bi's search) on a two-list frame ---------
    bi_search = None
    if not args.no_tz and bi:
        import ctypes
        n0 = work.num_refs - work.num_refs_l1
        mvp = np.zeros((n_ctu, npu, work.num_refs, 2), np.int16)
        tz = ua.TzCfg(16, work.R, 0, 0, 0, 8, (ctypes.c_int64 * 4)(*([7, 6, 9, 10][:work.num_refs] +
                                                                        [0] * (4 - work.num_refs))))
        hip.plan_run_tz(plan, dref, mvp, tz)     # the uni-prediction searches analyze_bi starts from
        ures, ubest = hip.plan_results(plan, work.num_refs)
        pus = np.zeros((n_ctu, npu), ua.BI_PU)
        ok = (ubest[..., 0] >= 0) & (ubest[..., 1] >= 0)
        pus["active"] = ok
        idx = np.arange(n_ctu)[:, None], np.arange(npu)[None, :]

        def mvbits(v):   # the MV-bits table of inter_search_create in closed form
            a = np.abs(v.astype(np.int64))
            return np.where(a == 0, 1, 2 * (np.floor(np.log2(a + 1)).astype(np.int64) + 1))

        costs = []
        for l in range(2):
            r = np.where(ok, ubest[..., l], 0).astype(np.int64)
            nl = n0 if l == 0 else work.num_refs_l1
            pus["refi"][..., l] = np.where(ok, r - (n0 if l else 0), 0)
            m = ures["mv_sub"][idx[0], idx[1], r]
            pus["mv"][..., l, :] = m
            ri = r - (n0 if l else 0)
            rb = np.where(nl <= 1, 0, np.where(ri < nl - 1, ri + 1, nl - 1))
            pus["mot_bits"][..., l] = rb + mvbits(m[..., 0]) + mvbits(m[..., 1])
            costs.append(ures["cost_final"][idx[0], idx[1], r])
        pus["lidx_ref"] = (costs[1] < costs[0]).astype(np.int8)
        ms0 = np.ascontiguousarray(ures["mv_sub"])
        pus0 = pus.copy()
        ms = ms0.copy()
        hip.plan_run_bi(plan, dref, mvp, ms, pus, tz)   # warm-up
        t_bi = []
        for _ in range(3):
            ms, pus = ms0.copy(), pus0.copy()
            t0 = time.perf_counter()
            hip.plan_run_bi(plan, dref, mvp, ms, pus, tz)
            t_bi.append((time.perf_counter() - t0) * 1e3)
        b_ms = float(np.median(t_bi))
        bi_search = {"ms_per_frame": b_ms, "pus": int(ok.sum()), "pus_per_s": int(ok.sum()) / (b_ms * 1e-3),
                     "note": "me_plan_run_bi: analyze_bi's search (bi-prediction target, me_search_tz bi = 1 "
                             "over the other list's references, choice; up to two rounds) for every PU with "
                             "both lists, from the frame's uni-prediction me_search_tz results; includes the "
                             "host<->device copies of the per-PU state (%.0f MB)" % (
                                 (pus.nbytes + 2 * ms.nbytes) / 1e6)}
        print("[bench] bi search: %.2f ms/frame (%d PUs)" % (b_ms, int(ok.sum())), file=sys.stderr, flush=True)
        # symmetric MVD (analyze_smvd) over the same PUs
        spus = np.zeros((n_ctu, npu), ua.SMVD_PU)
        spus["active"] = ok
        hip.plan_run_smvd(plan, dref, mvp, ms0, spus, tz)   # warm-up
        t_s = []
        for _ in range(3):
            t0 = time.perf_counter()
            hip.plan_run_smvd(plan, dref, mvp, ms0, spus, tz)
            t_s.append((time.perf_counter() - t0) * 1e3)
        s_ms = float(np.median(t_s))
        bi_search["smvd_ms_per_frame"] = s_ms
        bi_search["smvd_note"] = ("me_plan_run_smvd: analyze_smvd's search (MVP pair, mirrored uni MV, diamond x8 "
                                  "and cross refinement of bi-average com_had costs) for the same PUs; includes "
                                  "the host<->device copies")
        print("[bench] smvd search: %.2f ms/frame" % s_ms, file=sys.stderr, flush=True)

    # ---- affine luma motion compensation --------------------------------------
    affine = None
    if not args.no_tz:
        # the picture tiled by 32x32 CUs with 8x8 sub-blocks (4x4 in every other
        # CU), control-point MVs of a small zoom / rotation around the global motion
        arng = np.random.default_rng(12)
        xs, ys = np.meshgrid(np.arange(0, work.w - 31, 32), np.arange(0, work.h - 31, 32))
        nc = xs.size
        cus = np.zeros(nc, ua.AFFINE_CU)
        cus["x"], cus["y"], cus["w"], cus["h"] = xs.reshape(-1), ys.reshape(-1), 32, 32
        cus["sub_w"] = cus["sub_h"] = np.where(np.arange(nc) % 2 == 0, 8, 4)
        cus["vertex_num"] = 2
        cus["cpmv"][:, 0, :] = arng.integers(-64, 65, (nc, 2))
        cus["cpmv"][:, 1, :] = cus["cpmv"][:, 0, :] + arng.integers(-12, 13, (nc, 2))
        cus["pred_offset"] = np.arange(nc, dtype=np.uint64) * 1024
        apred = np.zeros(nc * 1024, hip.dt)
        hip.affine_mc_luma(dref[0], cus, apred)   # warm-up
        t_a = []
        for _ in range(5):
            t0 = time.perf_counter()
            hip.affine_mc_luma(dref[0], cus, apred)
            t_a.append((time.perf_counter() - t0) * 1e3)
        a_ms = float(np.median(t_a))
        affine = {"ms_per_frame": a_ms, "cus": int(nc), "samples_per_s": nc * 1024 / (a_ms * 1e-3),
                  "note": "me_affine_mc_luma: every 32x32 CU of the picture, 8x8 / 4x4 sub-blocks, 1/16-pel "
                          "8-tap interpolation; includes the host<->device copies of the predictions"}
        print("[bench] affine MC: %.3f ms/frame (%d CUs)" % (a_ms, nc), file=sys.stderr, flush=True)

    # ---- lookahead cost estimator (inter part) -------------------------------
    lookahead = None
    if not args.no_tz and not args.no_lookahead and hip.lib.me_plan_num_pus(plan) > 0:
        import ctypes
        n0 = work.num_refs - work.num_refs_l1
        pocs = [4, 0, 2, 6][:n0] + [12, 16][:work.num_refs_l1]
        # random access: GOP 16, SEARCH_RANGE_IPEL_RA 384 (com_define.h:137), padded references
        ltz = ua.TzCfg(16, 384, 0, 2, 0, 8, (ctypes.c_int64 * 4)(*(pocs + [0] * 4)[:4]))
        if work.mask >> 13 & 1:   # the 16x16 shape
            lcost = hip.plan_lookahead(plan, dref, ltz, 1, work.w, work.h)   # warm-up
            t0 = time.perf_counter()
            for _ in range(5):
                lcost = hip.plan_lookahead(plan, dref, ltz, 1, work.w, work.h)
            l_ms = (time.perf_counter() - t0) * 1e3 / 5
            lookahead = {"ms_per_frame": l_ms, "units_per_s": lcost.size / (l_ms * 1e-3), "units": int(lcost.size),
                         "note": "me_plan_lookahead: me_search_tz of every whole 16x16 unit from MVP 0 at "
                                 "integer precision against every reference, then com_had of each "
                                 "prediction (and the bi averages with two lists); includes the copy of "
                                 "the unit costs to the host"}
            from oracle.oracle import RefLib
            if RefLib.available(bd) and not args.no_cpu:
                rl = RefLib(bd, level=2)
                p0 = [rl.pic_create(refs[i], pocs[i]) for i in range(n0)]
                p1 = [rl.pic_create(refs[i], pocs[i]) for i in range(n0, work.num_refs)]
                t_all, t_intra = [], []
                for _ in range(3):
                    t0 = time.perf_counter()
                    rl.loka_cost(cur, p0, p1, 8, 16, 384, 1)
                    t_all.append((time.perf_counter() - t0) * 1e3)
                    t0 = time.perf_counter()
                    rl.loka_cost(cur, [], [], 8, 16, 384, 1)
                    t_intra.append((time.perf_counter() - t0) * 1e3)
                for q in p0 + p1:
                    rl.pic_free(q)
                c_ms = float(np.median(t_all))
                i_ms = float(np.median(t_intra))
                lookahead["cpu_reference_ms"] = c_ms
                lookahead["cpu_reference_inter_ms"] = c_ms - i_ms
                lookahead["speedup_inter_vs_cpu"] = (c_ms - i_ms) / l_ms
                lookahead["cpu_note"] = ("loka_estimate_coding_cost of the compiled reference (C+SSE4.2+AVX2), "
                                         "one thread as the encoder runs it, median of 3; its inter part alone is "
                                         "the call with the references minus the same call with none (intra only, "
                                         "%.1f ms)" % i_ms)
            print("[bench] lookahead: %.3f ms/frame (%d units)%s" % (
                l_ms, lcost.size, "; reference CPU inter part %.1f ms" % lookahead["cpu_reference_inter_ms"]
                if "cpu_reference_inter_ms" in lookahead else ""), file=sys.stderr, flush=True)

    # ---- the 1080p 16x16 workload: batched, per-call and CPU on the same frames
    # (1920x1080 8-bit, 16x16 PUs, full search +-32, 1 reference, sub-pel)
    per_call = None
    if not args.no_percall and bd == 8:
        from oracle.oracle import RefLib
        w1 = wl.hd_16x16()
        cur1, refs1 = w1.frames()
        d1 = [hip.ref_create(r) for r in refs1]
        plan1 = hip.plan_create(w1.w, w1.h, w1.R, w1.lam, w1.mask, 1)
        hip.plan_set_cur(plan1, cur1)
        n1 = hip.lib.me_plan_num_ctus(plan1)
        for _ in range(args.warmup):
            hip.plan_run(plan1, d1, 1)
        hip.lib.me_plan_results(plan1, None, None)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            hip.plan_run(plan1, d1, 1)
            hip.lib.me_plan_results(plan1, None, None)
        b1 = n1 * args.steps / (time.perf_counter() - t0)
        r1, bst1 = hip.plan_results(plan1, 1)
        hip.plan_destroy(plan1)
        for r in d1:
            hip.ref_destroy(r)
        # per-call: the host driver calling every SAD / SATD / filter through
        # the HIP table, on whole CTUs of the same frame (the middle row)
        host = ua.MeHost(8)
        table = hip.dispatch_table()
        planes1 = [host.planes_via_table(table, r) for r in refs1]
        cols1 = (w1.w + 127) // 128
        mid1 = ((w1.h + 127) // 128 // 2) * cols1 + cols1 // 2
        pc = {}
        for nthr in (1, 4):
            t0 = time.perf_counter()
            pres, pbest, _ = host.me_frame(table, cur1, planes1, w1.R, w1.lam, w1.mask, mid1, mid1 + nthr,
                                           threads=nthr)
            el = time.perf_counter() - t0
            same = ua.compare_results(pres[mid1:mid1 + nthr], pbest[mid1:mid1 + nthr], r1[mid1:mid1 + nthr],
                                      bst1[mid1:mid1 + nthr]) is None
            pc[nthr] = {"ctus_per_s": nthr / el, "seconds": el, "bit_exact_with_batched": same}
            print("[bench] per-call, %d thread(s): %.3f CTUs/s" % (nthr, nthr / el), file=sys.stderr, flush=True)
        c1 = None
        if RefLib.available(8) and not args.no_cpu:
            rl = RefLib(8, level=2)
            rl.set_pinning(1)
            nthr = cpu_threads()
            pics1 = [rl.pic_create(r) for r in refs1]
            runs = []
            for _ in range(3):
                cres1, cbest1, cms1 = rl.me_frame(cur1, pics1, w1.R, w1.lam, w1.mask, 0, n1, threads=nthr)
                runs.append(cms1)
            for q in pics1:
                rl.pic_free(q)
            c1 = {"ctus_per_s": n1 / (float(np.median(runs)) * 1e-3), "cores": nthr, "kind": "reference",
                  "bit_exact_with_batched": ua.compare_results(r1, bst1, cres1, cbest1) is None}
            print("[bench] 16x16 workload CPU: %.0f CTUs/s" % c1["ctus_per_s"], file=sys.stderr, flush=True)
        best_pc = max(v["ctus_per_s"] for v in pc.values())
        per_call = {"value": pc[1]["ctus_per_s"], "unit": "CTUs/s", "workload": w1.name,
                    "threads_1": pc[1], "threads_4": pc[4],
                    "batched_ctus_per_s": b1,
                    "speedup_batched_vs_per_call": b1 / best_pc,
                    "cpu_reference": c1,
                    "speedup_batched_vs_cpu": (b1 / c1["ctus_per_s"]) if c1 else None,
                    "hip_errors": int(hip.error_count()),
                    "sample": "whole CTUs of the 1080p 16x16 workload (+-32, 1 reference): the per-call table "
                              "(every SAD / SATD a synchronous HIP call from the host driver) on the middle "
                              "CTU(s), with 1 and 4 host threads; batched = the frame plan on the same frame, "
                              "same run; speedup against the faster per-call run"}

    # ---- CPU baseline leg -----------------------------------------------------
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        from oracle.oracle import RefLib, Oracle
        if RefLib.available(bd):
            # the reference kernels on a bounded sample: the whole frame when a
            # pass fits in about 10 s on the host's threads, else interior CTUs
            rl = RefLib(bd, level=2)
            rl.set_pinning(1)
            pics = [rl.pic_create(r) for r in refs]
            rows = (work.h + 127) // 128
            mid = (rows // 2) * cols
            _, _, t1 = rl.me_frame(cur, pics, work.R, work.lam, work.mask, mid, mid + 1, threads=1)
            n_s = int(min(n_ctu, max(threads, 10e3 * threads / max(t1, 1e-3))))
            c0 = 0 if n_s == n_ctu else max(0, min(n_ctu - n_s, mid - n_s // 2))
            c1 = c0 + n_s
            runs = []
            for _ in range(args.cpu_runs if n_s == n_ctu else 1):
                cres, cbest, cms = rl.me_frame(cur, pics, work.R, work.lam, work.mask, c0, c1, threads=threads)
                runs.append(cms)
            for pc_ in pics:
                rl.pic_free(pc_)
            cms = float(np.median(runs))
            kind = "reference"
            what = ("uavs3e kernels (C+SSE4.2+AVX2 dispatch table) compiled from the reference sources, "
                    "median of %d runs" % len(runs))
        else:
            # the scalar C oracle on one interior CTU
            rows = (work.h + 127) // 128
            c0 = (rows // 2) * cols + cols // 2
            c1 = c0 + 1
            orc = Oracle(bd)
            t0 = time.perf_counter()
            cres, cbest = orc.me_frame(cur, [orc.make_planes(r) for r in refs], work.R, work.lam, work.mask,
                                       c0, c1)
            cms = (time.perf_counter() - t0) * 1e3
            threads = 1
            kind = "port"
            what = "the scalar C oracle"
        if bi:   # the reference harness has one list: compare the per-reference search results
            cres = cres.copy()
            cres["cost_final"] = res["cost_final"][:cres.shape[0]]
            diff = ua.compare_results(res[c0:c1], best[c0:c1], cres[c0:c1], best[c0:c1])
        else:
            diff = ua.compare_results(res[c0:c1], best[c0:c1], cres[c0:c1], cbest[c0:c1])
        model, logical, allowed = host_cpu()
        cpu = {"value": (c1 - c0) / (cms * 1e-3), "unit": "CTUs/s", "cores": threads, "kind": kind,
               "cpu_model": model, "host_logical_cpus": logical, "process_cpus": allowed,
               "pinned": kind == "reference",
               "sample": "CTUs %d..%d of %s, %s, %d threads, %.2f s per pass; results bit-exact with the GPU: %s"
                         % (c0, c1 - 1, work.name, what + (", one thread pinned per CPU" if kind == "reference"
                                                            else ""), threads, cms * 1e-3,
                            "yes" if diff is None else "NO (%s)" % diff)}
        print("[bench] cpu baseline: %.2f CTUs/s (%s)" % (cpu["value"], kind), file=sys.stderr, flush=True)

    # ---- serving leg: 6 concurrent streams, Poisson per-CTU requests ---------
    serving = None
    if not args.no_serve:
        sizes = [(1280, 720)] * 3 + [(1920, 1080)] * 2 + [(3840, 2160)]
        seeds = [0x40 + i for i in range(len(sizes))]
        # 4 displacement rows per workgroup: short per-batch latency at light load
        cfg = ua.ServerCfg(32, wl.lambda_mv(32), ua.MASK_16x16, 1, 64 << 30, 128, 4, 100, 4, 4, 0)
        host8 = ua.MeHost(8)
        # one unmeasured run of the same load first: the first server run in
        # a process stalls once for several milliseconds after its first batch
        host8.serve_load(cfg, sizes, seeds, args.serve_frames, args.serve_rate, arrival_seed=8)
        slat, info, st, _ = host8.serve_load(cfg, sizes, seeds, args.serve_frames, args.serve_rate, arrival_seed=7)
        lat_ms = slat / 1e3
        serving = {"streams": "3x1280x720 + 2x1920x1080 + 1x3840x2160, 8-bit, %d frames each" % args.serve_frames,
                   "job": "one CTU: 16x16 PUs, full search +-32, 1 reference (the previous frame), QP 32, sub-pel",
                   "warmup": "one unmeasured run of the same load first (other arrival seed)",
                   "offered_ctus_per_s": args.serve_rate, "requests": info["requests"],
                   "throughput_ctus_per_s": info["requests"] / (info["wall_us"] * 1e-6),
                   "latency_ms": {"p50": pct(lat_ms, 50), "p99": pct(lat_ms, 99), "max": float(lat_ms.max())},
                   "server": {"hbm_budget_gb": 64, "max_batch_ctus": 128, "max_inflight": 4, "max_wait_us": 100,
                              "band_rows": 4,
                              "batches": st["batches"], "mean_batch": info["requests"] / max(1, st["batches"]),
                              "mixed_batches": st["mixed_batches"], "max_inflight_seen": st["max_inflight_seen"],
                              "max_launch_us": st["max_launch_us"], "max_batch_us": st["max_batch_us"],
                              "peak_picture_bytes_end": st["picture_bytes"]},
                   "frame_synthesis_s": info["synth_s"]}
        print("[bench] serving: p50 %.3f ms p99 %.3f ms, %.0f CTUs/s" % (
            serving["latency_ms"]["p50"], serving["latency_ms"]["p99"], serving["throughput_ctus_per_s"]),
            file=sys.stderr, flush=True)

    if args.calib_copy:
        src = torch.zeros(args.calib_copy, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.bitwise_not(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    hip.plan_destroy(plan)
    for r in dref:
        hip.ref_destroy(r)

    valu_peak = None
    if args.valu_peak:
        import ctypes
        from uavs3e_amd import build as ub
        pk = ctypes.CDLL(ub.peak_lib())
        pk.me_peak_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        names = ["v_add_u32", "v_sad_u8", "v_qsad_pk_u16_u8", "v_sad_u16", "v_min_f64", "v_add_u32_dpp",
                 "v_lshl_add_u64", "v_add_u32_sdwa"]
        valu_peak = {"note": "lane-instructions/s of one instruction kind: 8 independent chains per lane, 2048 "
                             "workgroups of 256 lanes (8 waves per SIMD), HIP events; per CU and clock = "
                             "value / (256 CUs x clock)"}
        for k, nm in enumerate(names):
            ms_ = ctypes.c_double()
            li_ = ctypes.c_double()
            if pk.me_peak_run(k, 2048, 2048, ctypes.byref(ms_), ctypes.byref(li_)) != 0:
                continue
            valu_peak[nm] = li_.value / (ms_.value * 1e-3)
        print("[bench] VALU peak: %s" % ", ".join("%s %.1f T/s" % (k, v / 1e12) for k, v in valu_peak.items()
                                                  if k != "note"), file=sys.stderr, flush=True)

    line = {
        "metric": "1080p ME CTUs/s batched vs per-call vs AVX2, bit-exact MVs; p50/p99 CTU latency",
        "value": ctus_per_s,
        "unit": "CTUs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8" if bd == 8 else "u16",
        "data": "synthetic (seeded sinusoid field + noise, global + per-region motion)",
        "config": {"workload": work.name, "pic": "%dx%d" % (work.w, work.h), "bit_depth": bd,
                   "refs": work.num_refs, "search_range": work.R, "pu_shapes": len(shapes),
                   "pus_per_frame": int(valid[:, :, 0].sum()), "ctus_per_frame": n_ctu,
                   "step": "rebuild newest reference planes + full search + sub-pel + ref choice, one hipGraph" +
                           (" + MC of every PU (luma, 4:2:0 chroma, bi-average)" if bi else ""),
                   "lists": "2 L0 + 2 L1" if bi else "L0"},
        "autotune": tune,
        "schedule": {"band_rows": args.band_rows, "waves": args.waves, "dx_cols": args.dx_cols,
                     "int_grid": args.int_grid, "sp_groups": args.sp_groups},
        "candidate_evals_per_s": cand_per_s,
        "ctu_latency_ms": ({"p50": serving["latency_ms"]["p50"], "p99": serving["latency_ms"]["p99"],
                            "note": "serving leg: per-CTU request arrival to result on the host"}
                           if serving else None),
        "frame_latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99),
                             "note": "batched: every CTU of a frame completes with the frame"},
        "stage_ms": {"prepare_ref": stage[0], "fill": stage[1], "int_search": stage[2], "subpel": stage[3],
                     "mc": mc_ms},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic, "traffic_schedule": traffic_sched,
                     "kernel": "k_int_search", "algorithmic_bytes": alg_bytes, "kernel_ms": k_ms,
                     "valu": {"ops": valu_ops, "frac": valu_frac, "peak_lane_ops_per_s": VALU_PEAK_OPS,
                              "abs_diffs_per_lane_op": ad_per_op,
                              "frac_at_4_abs_diffs_per_op": valu_ops4 / (k_ms * 1e-3) / VALU_PEAK_OPS,
                              "ops_note": "every candidate SAD composed from 4x4-block SADs (16 abs-diffs per "
                                          "block and displacement, abs_diffs_per_lane_op per lane-op: 4 for 8-bit "
                                          "samples with v_sad_u8, 2 for 10-bit with v_sad_u16) plus one compare "
                                          "per PU candidate"}},
        "per_call": per_call,
        "end_to_end": e2e,
        "serving": serving,
        "search_tz": search_tz,
        "lookahead": lookahead,
        "bi_search": bi_search,
        "affine_mc": affine,
        "cpu_baseline": cpu,
        "valu_peak": valu_peak,
        "speedup_vs_cpu": (ctus_per_s / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
