"""The request server (continuous batching across streams) through the C ABI.

Each request's result must be bit-identical whether it runs alone or mixed
with other streams' requests in shared batches, and equal to the frame API's
result for the same CTU (the frame API is itself checked against the oracle
in test_gpu_parity.py; here one picture set is also checked against the
oracle directly)."""
import random
import threading

import numpy as np
import pytest

import uavs3e_amd as ua
from uavs3e_amd import workloads as wl
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

R = 12
NREF = 2
MASK = sum(1 << ua.SHAPES.index(s) for s in [(8, 8), (16, 8), (8, 32), (16, 16), (32, 32), (64, 16), (128, 128)])
SIZES = [(264, 200), (392, 136), (200, 264)]   # three streams of different resolutions
FRAMES = 4                                     # pocs 0..3; requests for pocs 2, 3 against (poc-1, poc-2)


def stream_frames(i, w, h):
    host = ua.MeHost(8)
    frames = []
    for f in range(FRAMES):
        frames.append(host.synth(0x40 + i, i, f, w, h, f * 13 / 4.0, f * -6 / 4.0, 16, 3.0, 4, 0.0))
    return frames


@pytest.fixture(scope="module")
def streams():
    return [stream_frames(i, w, h) for i, (w, h) in enumerate(SIZES)]


def n_ctus(w, h):
    return ((w + 127) // 128) * ((h + 127) // 128)


def requests_of(i):
    w, h = SIZES[i]
    return [(i, poc, ctu) for poc in range(2, FRAMES) for ctu in range(n_ctus(w, h))]


def run_server(hip, streams, order, threads, **cfg):
    """Runs the requests `order` ((stream, poc, ctu) tuples) through one
    server, submitted by `threads` threads; returns {(stream, poc, ctu):
    (res, best)} and the server statistics."""
    lam = wl.lambda_mv(32)
    srv = ua.MeServer(hip, R, lam, MASK, NREF, **cfg)
    try:
        sids = {}
        for i in sorted({o[0] for o in order}):
            sids[i] = srv.open_stream(*SIZES[i])
            for poc, fr in enumerate(streams[i]):
                assert srv.put_picture(sids[i], poc, fr)
        out = {}
        lock = threading.Lock()

        def worker(part):
            tickets = [(o, srv.submit(sids[o[0]], o[1], o[2], [o[1] - 1, o[1] - 2])) for o in part]
            for o, t in tickets:
                res, best, _ = srv.wait(t)
                with lock:
                    out[o] = (res, best)

        parts = [order[k::threads] for k in range(threads)]
        ts = [threading.Thread(target=worker, args=(p,)) for p in parts]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        return out, srv.stats()
    finally:
        srv.close()


def test_server_alone_vs_mixed_vs_frame_api(streams):
    hip = ua.MeHip(8)
    # alone: one stream per server, one request in flight at a time
    alone = {}
    for i in range(len(SIZES)):
        got, st = run_server(hip, streams, requests_of(i), 1, max_batch_ctus=1, max_inflight=1, max_wait_us=0)
        assert st["mixed_batches"] == 0
        alone.update(got)
    # mixed: every stream in one server, shuffled, three submitting threads,
    # batches of up to 5 CTUs that wait for company, three batches in flight
    order = [o for i in range(len(SIZES)) for o in requests_of(i)]
    random.Random(5).shuffle(order)
    mixed, st = run_server(hip, streams, order, 3, max_batch_ctus=5, max_inflight=3, max_wait_us=20000)
    assert st["requests_done"] == len(order)
    assert st["mixed_batches"] > 0, st
    assert st["max_batch_seen"] > 1, st
    assert sorted(mixed) == sorted(alone)
    for k in alone:
        assert np.array_equal(alone[k][0], mixed[k][0]) and np.array_equal(alone[k][1], mixed[k][1]), k
    # the frame API on the same (current, references) gives the same CTU results
    lam = wl.lambda_mv(32)
    for i, (w, h) in enumerate(SIZES):
        fr = streams[i]
        for poc in range(2, FRAMES):
            dref = [hip.ref_create(fr[poc - 1]), hip.ref_create(fr[poc - 2])]
            plan = hip.plan_create(w, h, R, lam, MASK, NREF)
            hip.plan_set_cur(plan, fr[poc])
            hip.plan_run(plan, dref)
            res, best = hip.plan_results(plan, NREF)
            hip.plan_destroy(plan)
            for r in dref:
                hip.ref_destroy(r)
            for ctu in range(n_ctus(w, h)):
                sres, sbest = mixed[(i, poc, ctu)]
                msg = ua.compare_results(sres[None], sbest[None], res[ctu:ctu + 1], best[ctu:ctu + 1])
                assert msg is None, (i, poc, ctu, msg)


def test_server_vs_oracle(streams):
    """One stream's requests against the oracle directly."""
    hip, orc = ua.MeHip(8), Oracle(8)
    i, poc = 1, 3
    got, _ = run_server(hip, streams, [(i, poc, c) for c in range(n_ctus(*SIZES[i]))], 2, max_batch_ctus=2,
                        max_inflight=2, max_wait_us=0)
    fr = streams[i]
    ores, obest = orc.me_frame(fr[poc], [orc.make_planes(fr[poc - 1]), orc.make_planes(fr[poc - 2])], R,
                               wl.lambda_mv(32), MASK)
    for ctu in range(n_ctus(*SIZES[i])):
        res, best = got[(i, poc, ctu)]
        msg = ua.compare_results(res[None], best[None], ores[ctu:ctu + 1], obest[ctu:ctu + 1])
        assert msg is None, (ctu, msg)


def test_server_hbm_admission_and_reuse():
    """Pictures are admitted against the HBM budget: a put that does not fit
    waits (here: times out) until a released picture's memory frees up, and a
    released picture's buffers are reused by the next picture of its size."""
    hip = ua.MeHip(8)
    w, h = 256, 136
    host = ua.MeHost(8)
    frames = [host.synth(9, 0, f, w, h, 0.0, 0.0, 16, 3.0, 4, 0.0) for f in range(4)]
    srv = ua.MeServer(hip, R, wl.lambda_mv(32), ua.MASK_16x16, 1, hbm_budget=1)
    per_pic = None
    try:
        sid = srv.open_stream(w, h)
        assert not srv.put_picture(sid, 0, frames[0], timeout_ms=0)   # budget of 1 byte: nothing fits
    finally:
        srv.close()
    from uavs3e_amd import server_picture_bytes
    per_pic = server_picture_bytes(w, h, 8)
    srv = ua.MeServer(hip, R, wl.lambda_mv(32), ua.MASK_16x16, 1, hbm_budget=2 * per_pic)
    try:
        sid = srv.open_stream(w, h)
        assert srv.put_picture(sid, 0, frames[0])
        assert srv.put_picture(sid, 1, frames[1])
        assert srv.stats()["picture_bytes"] == 2 * per_pic
        assert not srv.put_picture(sid, 2, frames[2], timeout_ms=50)   # over budget
        t = srv.submit(sid, 1, 0, [0])
        done = []

        def put_late():
            done.append(srv.put_picture(sid, 2, frames[2]))   # waits for the budget

        th = threading.Thread(target=put_late)
        th.start()
        th.join(timeout=0.3)
        assert th.is_alive() and not done     # still waiting: nothing was released
        srv.release_picture(sid, 0)           # named by the pending request until it completes
        res, best, _ = srv.wait(t)
        assert (best >= 0).any()
        th.join(timeout=30)
        assert not th.is_alive() and done == [True]
        st = srv.stats()
        assert st["picture_bytes"] == 2 * per_pic   # picture 0's buffers were reused
        assert st["admission_waits"] >= 2
        t = srv.submit(sid, 2, 1, [1])
        assert srv.wait(t) is not None
    finally:
        srv.close()


def test_serve_load_matches_frame_api():
    """The encoder-side load driver (uploader / submitter / collector threads
    per stream, Poisson arrivals, pictures released as frames complete): every
    request completes and equals the frame API's result for its CTU."""
    hip, host = ua.MeHip(8), ua.MeHost(8)
    sizes = [(264, 200), (136, 136), (392, 136)]
    seeds = [0x40, 0x41, 0x42]
    frames, Rs = 4, 16
    lam = wl.lambda_mv(32)
    cfg = ua.ServerCfg(Rs, lam, ua.MASK_16x16, 1, 1 << 30, 4, 3, 100, 0, 0, 0)
    lat, info, st, (res, best) = host.serve_load(cfg, sizes, seeds, frames, 3000.0, arrival_seed=3,
                                                 keep_results=True)
    assert info["requests"] == len(lat) == sum(n_ctus(w, h) for w, h in sizes) * (frames - 1)
    assert (lat > 0).all() and st["requests_done"] == info["requests"]
    k = 0
    for i, (w, h) in enumerate(sizes):
        fr = [host.synth(seeds[i], i, f, w, h, f * 13 / 4.0, f * -6 / 4.0, 16, 3.0, 4, 0.0) for f in range(frames)]
        for f in range(1, frames):
            dref = [hip.ref_create(fr[f - 1])]
            plan = hip.plan_create(w, h, Rs, lam, ua.MASK_16x16, 1)
            hip.plan_set_cur(plan, fr[f])
            hip.plan_run(plan, dref)
            pres, pbest = hip.plan_results(plan, 1)
            hip.plan_destroy(plan)
            hip.ref_destroy(dref[0])
            n = n_ctus(w, h)
            msg = ua.compare_results(res[k:k + n], best[k:k + n], pres, pbest)
            assert msg is None, (i, f, msg)
            k += n
