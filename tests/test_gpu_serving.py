"""The continuous-batching query server: every answer is the same alone and
in a mix of concurrent requests (bit-exact), matches the oracle, and the
server batches requests of one stage into one call."""
import numpy as np
import pytest

import oracle
import shell_gpu as sg
from common import expand_oracle, inner_oracle, ntt_primes

pytestmark = pytest.mark.gpu

LOG_N, L, LEVELS, C = 10, 2, 3, 2


@pytest.fixture(scope="module")
def ctx():
    c = sg.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def setup(ctx):
    qs = ntt_primes(LOG_N, L)
    rns = oracle.Rns(qs, LOG_N)
    pids = ctx.primes(qs, LOG_N)
    log_bs = [10] * L
    k, _ = rns.gadget_dims(log_bs)
    keys = [(rns.uniform(100 + j, k), rns.uniform(200 + j, k)) for j in range(LEVELS)]
    dkeys = [(ctx.array(a), ctx.array(b)) for a, b in keys]
    K = 1 << LEVELS
    dbs = [rns.uniform(300, K * C), rns.uniform(301, K * (C + 1))]
    dbx = []
    for db in dbs:
        x = ctx.empty(db.shape)
        ctx.pir_db_export(ctx.array(db), db.shape[0], x, pids)
        dbx.append(x)
    queries = [rns.uniform(400 + g, 2) for g in range(6)]
    return dict(rns=rns, pids=pids, log_bs=log_bs, keys=keys, dkeys=dkeys, dbs=dbs, dbx=dbx,
                queries=queries)


def _server(ctx, s, max_active, budget=0, cohort=0):
    srv = sg.PirServer(ctx, s["pids"], LOG_N, s["log_bs"], s["dkeys"], max_active, budget, cohort)
    ids = [srv.add_db(s["dbx"][0], C), srv.add_db(s["dbx"][1], C + 1)]
    srv.start()
    return srv, ids


def test_alone_equals_mixed_and_oracle(ctx, setup):
    s = setup
    srv, dbs = _server(ctx, s, max_active=4)
    which = [0, 1, 0, 1, 1, 0]
    tickets = [srv.submit(dbs[w], q) for w, q in zip(which, s["queries"])]
    mixed = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    assert st["capacity"] == 4
    assert st["batched_items"] > st["stage_runs"]  # stages ran for several requests at once
    alone = []
    for w, q in zip(which, s["queries"]):
        alone.append(srv.wait(srv.submit(dbs[w], q))[0])
    srv.close()
    rns = s["rns"]
    for g, (a, m) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a, m), g
    for g in (0, 1):
        e = expand_oracle(rns, s["queries"][g], LEVELS, s["keys"], s["log_bs"])
        cols = C + which[g]
        exp = inner_oracle(rns, e, s["dbs"][which[g]], cols)
        assert np.array_equal(mixed[g].reshape(exp.shape), exp), g


def test_hbm_budget_limits_admission(ctx, setup):
    s = setup
    n = 1 << LOG_N
    per_request = ((1 << LEVELS) + C + 1) * 2 * L * n * 8
    srv, dbs = _server(ctx, s, max_active=4, budget=per_request)
    assert srv.stats()["capacity"] == 1
    tickets = [srv.submit(dbs[0], q) for q in s["queries"][:3]]
    outs = [srv.wait(t)[0] for t in tickets]
    lat, wall = srv.load_test(dbs[0], np.stack(s["queries"][:2]), 4, rate=2000.0)
    srv.close()
    assert len(outs) == 3 and all(o.size == C * 2 * L * n for o in outs)
    assert np.all(lat > 0) and wall > 0


def test_cohort_admission_same_answers(ctx, setup):
    """Newcomers admitted in cohorts of three while others run: every answer
    equals the one computed alone, under a Poisson load as well."""
    s = setup
    srv, dbs = _server(ctx, s, max_active=6, cohort=3)
    alone = [srv.wait(srv.submit(dbs[0], q))[0] for q in s["queries"]]
    first = [srv.submit(dbs[0], q) for q in s["queries"][:2]]
    later = [srv.submit(dbs[0], q) for q in s["queries"]]
    got_first = [srv.wait(t)[0] for t in first]
    got_later = [srv.wait(t)[0] for t in later]
    lat, wall = srv.load_test(dbs[0], np.stack(s["queries"]), 12, rate=500.0)
    srv.close()
    for g, m in enumerate(got_first):
        assert np.array_equal(alone[g], m), g
    for g, m in enumerate(got_later):
        assert np.array_equal(alone[g], m), g
    assert np.all(lat > 0) and wall > 0
