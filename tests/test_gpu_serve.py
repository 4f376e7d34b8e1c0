"""The request server (genfire_server_*): continuous batching of reconstruction
jobs on one GPU.

Bar: every request's reconstruction, errK and errR equal those of the same job
run alone through genfire_job_* (bit-exact: batching shares launches, not
arithmetic); admission respects the HBM budget and the slot count; invalid
requests fail at submission with the reference's message."""
import numpy as np
import pytest

from genfire_amd import _lib

pytestmark = pytest.mark.gpu


def _phantom(dim, P, seed, single_axis):
    rng = np.random.default_rng(seed)
    k = np.arange(dim) - dim // 2
    vol = np.zeros((dim, dim, dim), np.float32)
    for _ in range(max(4, dim ** 3 // 800)):
        c = rng.uniform(-dim / 3, dim / 3, 3)
        g = [np.exp(-((k - q) ** 2) / 3.0) for q in c]
        vol += rng.uniform(0.5, 1.0) * (g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :])
    if single_axis:
        ang = np.stack([np.zeros(P), np.linspace(-70, 70, P), np.zeros(P)], 1)
    else:
        ang = np.stack([rng.uniform(-5, 5, P), np.linspace(-70, 70, P), rng.uniform(-5, 5, P)], 1)
    sup = (k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2) <= (0.42 * dim) ** 2
    return vol, ang, sup


def _params(iters, dr, monitor, prec=_lib.PREC_FP32):
    p = _lib.default_params()
    p.griddingMethod = 2   # DFT gridding
    p.numIterations = iters
    p.oversamplingRatio = 2.0
    p.interpolationCutoffDistance = 0.3
    p.constraintEnforcementMode = 3 if dr else 1
    p.algorithm = _lib.ALG_DR if dr else _lib.ALG_GENFIRE
    p.monitor_R = 1 if monitor else 0
    p.monitorR_loopLength = 5
    p.precision = prec
    for i in range(3):
        p.vector3[i] = (0.0, 0.0, 1.0)[i]
    return p


# (dim, P, iterations, DR, monitor, seed): dim 64 -> 128^3 half-spectrum path
# (three same-shape jobs batch together), dim 32 -> 64^3 and dim 48 -> 96^3 on
# the full-spectrum path
MIX = [
    (64, 21, 30, True, True, 1),
    (64, 17, 45, True, False, 2),
    (64, 25, 20, False, True, 3),
    (32, 13, 25, True, True, 4),
    (48, 15, 15, False, False, 5),
    (32, 9, 35, False, True, 6),
]


@pytest.fixture(scope="module")
def mix_inputs(ctx):
    out = []
    for dim, P, iters, dr, mon, seed in MIX:
        vol, ang, sup = _phantom(dim, P, seed, single_axis=seed % 2 == 0)
        projs = np.stack([ctx.project(vol, *a) for a in ang], 2).astype(np.float32)
        out.append((projs, ang, sup, _params(iters, dr, mon)))
    return out


def _alone(ctx, projs, ang, sup, p):
    d1, d2, P = projs.shape
    job = _lib.Job(ctx, p, d1, d2, P)
    job.set_inputs(projs, ang, sup)
    job.grid()
    job.reconstruct()
    return job.reconstruction(), job.errK(), job.errR()[0]


def test_results_same_alone_and_in_mix(ctx, mix_inputs):
    print("[serve] mix of 6 requests, 4 slots", flush=True)
    alone = [_alone(ctx, *m) for m in mix_inputs]
    print("[serve] alone runs done", flush=True)
    cfg = _lib.default_server_config()
    cfg.max_active = 4
    cfg.lookahead = 2
    srv = _lib.Server(ctx, cfg)
    ids = [srv.submit(p, projs, ang, sup) for projs, ang, sup, p in mix_inputs]
    srv.drain()
    for rid, (rec0, ek0, er0), m in zip(ids, alone, MIX):
        assert srv.status(rid) == _lib.REQ_DONE, m
        rec, ek, er, lat = srv.result(rid)
        print(f"[serve] {m}: latency {lat['total']:.1f} ms", flush=True)
        np.testing.assert_array_equal(rec, rec0)
        np.testing.assert_array_equal(ek, ek0)
        np.testing.assert_array_equal(er, er0)
    st = srv.stats()
    print(f"[serve] stats {st}", flush=True)
    assert st["completed"] == len(MIX) and st["failed"] == 0
    assert st["max_active_seen"] <= 4
    # the three 128^3 jobs iterate together at least part of the time
    assert st["batched_jobs"] > st["batched_launches"]
    # steady-state iterations replay the group's captured graph
    assert 0 < st["graph_replays"] < st["batched_launches"], st
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_direct_launches_equal_graph_replays(ctx, mix_inputs, monkeypatch):
    """GENFIRE_GRAPHS=0 launches every scheduler iteration directly; the results
    equal those of the jobs alone (which replay graphs), bit for bit."""
    print("[serve] graphs off", flush=True)
    sub = mix_inputs[:3] + mix_inputs[3:4]
    alone = [_alone(ctx, *m) for m in sub]
    monkeypatch.setenv("GENFIRE_GRAPHS", "0")
    cfg = _lib.default_server_config()
    cfg.max_active = 4
    srv = _lib.Server(ctx, cfg)
    monkeypatch.delenv("GENFIRE_GRAPHS")
    ids = [srv.submit(p, projs, ang, sup) for projs, ang, sup, p in sub]
    srv.drain()
    for rid, (rec0, ek0, er0) in zip(ids, alone):
        assert srv.status(rid) == _lib.REQ_DONE
        rec, ek, er, _ = srv.result(rid)
        np.testing.assert_array_equal(rec, rec0)
        np.testing.assert_array_equal(ek, ek0)
        np.testing.assert_array_equal(er, er0)
    st = srv.stats()
    assert st["graph_replays"] == 0 and st["batched_jobs"] > st["batched_launches"], st
    srv.close()


def test_batched_jobs_with_different_boxes(ctx):
    """Two 128^3 jobs whose support boxes differ (Dim 64 at ratio 2, Dim 48 at
    ratio 8/3) share the batched launches; each result equals its job alone."""
    print("[serve] one grid shape, two box sizes", flush=True)
    reqs = []
    for dim, ratio, P, seed in [(64, 2.0, 15, 11), (48, 128 / 48, 13, 12)]:
        vol, ang, sup = _phantom(dim, P, seed, single_axis=False)
        projs = np.stack([ctx.project(vol, *a) for a in ang], 2).astype(np.float32)
        p = _params(25, True, True)
        p.oversamplingRatio = ratio
        reqs.append((projs, ang, sup, p))
    alone = [_alone(ctx, *r) for r in reqs]
    cfg = _lib.default_server_config()
    cfg.max_active = 2
    srv = _lib.Server(ctx, cfg)
    ids = [srv.submit(p, projs, ang, sup) for projs, ang, sup, p in reqs]
    srv.drain()
    for rid, (rec0, ek0, er0) in zip(ids, alone):
        assert srv.status(rid) == _lib.REQ_DONE
        rec, ek, er, _ = srv.result(rid)
        np.testing.assert_array_equal(rec, rec0)
        np.testing.assert_array_equal(ek, ek0)
        np.testing.assert_array_equal(er, er0)
    st = srv.stats()
    assert st["batched_jobs"] > st["batched_launches"], st
    srv.close()


def test_budget_admits_one_at_a_time(ctx, mix_inputs):
    print("[serve] budget for one job", flush=True)
    cfg = _lib.default_server_config()
    cfg.max_active = 8
    cfg.hbm_budget_bytes = 200_000_000   # fits one 128^3 job (about 150 MB estimated), not two
    srv = _lib.Server(ctx, cfg)
    ids = [srv.submit(m[3], *m[:3]) for m in mix_inputs[:3]]
    for rid in ids:
        srv.wait(rid)
    st = srv.stats()
    assert st["completed"] == 3
    assert st["max_active_seen"] == 1, st
    assert st["peak_footprint_bytes"] <= cfg.hbm_budget_bytes
    srv.close()


def test_invalid_request_fails_at_submission(ctx, mix_inputs):
    print("[serve] invalid request", flush=True)
    srv = _lib.Server(ctx)
    projs, ang, sup, p = mix_inputs[0]
    p.constraintEnforcementMode = 4
    with pytest.raises(_lib.GenfireError, match="constraintEnforcementMode value 4 not understood"):
        srv.submit(p, projs, ang, sup)
    p.constraintEnforcementMode = 3
    st = srv.stats()
    assert st["submitted"] == 0
    srv.close()


def test_failed_iteration_fails_its_group(ctx, mix_inputs, monkeypatch):
    """An error inside the scheduler's iteration of a group (forced from the
    host: GENFIRE_SERVER_FAIL_AT makes its third group iteration raise after the
    launches are queued) fails every request of that group with the error: the
    waits return, nothing stays ITERATING, and the server goes on serving."""
    print("[serve] injected scheduler failure", flush=True)
    monkeypatch.setenv("GENFIRE_SERVER_FAIL_AT", "3")
    cfg = _lib.default_server_config()
    cfg.max_active = 4
    srv = _lib.Server(ctx, cfg)
    monkeypatch.delenv("GENFIRE_SERVER_FAIL_AT")
    projs, ang, sup, p = mix_inputs[0]
    ids = [srv.submit(p, projs, ang, sup) for _ in range(2)]
    srv.drain()
    states = [srv.status(i) for i in ids]
    assert _lib.REQ_FAILED in states, states
    assert all(s in (_lib.REQ_FAILED, _lib.REQ_DONE) for s in states), states
    with pytest.raises(_lib.GenfireError, match="failure injected"):
        srv.result(ids[states.index(_lib.REQ_FAILED)])
    st = srv.stats()
    assert st["failed"] >= 1 and st["failed"] + st["completed"] == 2, st
    # the server keeps serving: a later request completes and equals the job alone
    rid = srv.submit(p, projs, ang, sup)
    srv.wait(rid)
    assert srv.status(rid) == _lib.REQ_DONE
    rec, ek, er, _ = srv.result(rid)
    rec0, ek0, er0 = _alone(ctx, projs, ang, sup, p)
    np.testing.assert_array_equal(rec, rec0)
    np.testing.assert_array_equal(ek, ek0)
    srv.close()
