"""GPU tests of the serving loop (hdg_server_*): results identical alone
and in a mix of concurrent requests, HBM-budget admission, error isolation.
"""
import collections

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QD = {1: 2, 2: 3, 3: 6, 4: 6}


def _req(hdg, nx, ny, k, seed, want_trace=False, precision=0, ka=3.0, kb=5.0):
    rng = np.random.default_rng(seed)
    return hdg.make_request(nx=nx, ny=ny, k=k, jitter=0.2, seed=seed, src_kind=hdg._lib.SRC_SINE_SERIES,
                            kappa=(ka, kb), coef=rng.standard_normal(16), want_trace=want_trace,
                            precision=precision)


def test_server_alone_equals_mix_and_context(hdg):
    """A request gives the same trace solution, bit for bit, run alone in a
    one-slot server, in the middle of 12 concurrent requests of other sizes,
    orders and precisions in a 4-slot server, and through a plain Context."""
    x = _req(hdg, 48, 40, 2, 9, want_trace=True)
    with hdg.Server(slots=1) as s1:
        r1 = s1.wait(s1.submit(x))
        assert r1["status"] == 0
        t_alone = s1.trace(r1["id"])
    others = [(64, 64, 1), (33, 17, 3), (96, 80, 2), (20, 100, 4), (128, 128, 2), (50, 50, 1),
              (40, 72, 3), (81, 81, 2), (16, 16, 4), (100, 60, 1), (64, 32, 3)]
    with hdg.Server(slots=4) as s4:
        ids = []
        for i, (nx, ny, k) in enumerate(others):
            ids.append(s4.submit(_req(hdg, nx, ny, k, 100 + i, precision=i % 2)))
            if i == 5:
                xid = s4.submit(x)
        res = [s4.wait(i) for i in ids]
        rx = s4.wait(xid)
        t_mix = s4.trace(xid)
        st = s4.stats()
    assert all(r["status"] == 0 for r in res) and rx["status"] == 0
    assert st["completed"] == len(others) + 1 and st["failed"] == 0
    assert np.array_equal(t_alone, t_mix)
    assert rx["iters"] == r1["iters"]
    # the same problem through the Context API
    m = hdg.rectangle_mesh((48, 40), jitter=0.2, seed=9)
    ctx = hdg.Context(0)
    ctx.set_mesh(m)
    ctx.setup(2, quad_degree=3, rtol=1e-12)
    rng = np.random.default_rng(9)
    ctx.set_source(kind=hdg._lib.SRC_SINE_SERIES, kappa=(3.0, 5.0), coef=rng.standard_normal(16))
    ctx.set_dirichlet(m.boundary_dofs(3), np.zeros(len(m.boundary_dofs(3))))
    ctx.run()
    t_ctx = ctx.trace()
    ctx.close()
    assert np.array_equal(t_alone, t_ctx)
    # overlap happened: some requests started before others finished
    iv = sorted((r["t_start"], r["t_done"]) for r in res + [rx])
    assert any(iv[i + 1][0] < iv[i][1] for i in range(len(iv) - 1))


def test_server_budget_admission(hdg):
    """With a budget below two footprints, requests run one at a time."""
    reqs = [_req(hdg, 96, 96, 2, 200 + i) for i in range(5)]
    foot = max(hdg.footprint(r) for r in reqs)
    budget = int(1.5 * foot)
    with hdg.Server(slots=4, hbm_budget=budget) as s:
        ids = [s.submit(r) for r in reqs]
        res = [s.wait(i) for i in ids]
        st = s.stats()
    assert all(r["status"] == 0 for r in res)
    iv = sorted((r["t_start"], r["t_done"]) for r in res)
    assert all(iv[i + 1][0] >= iv[i][1] for i in range(len(iv) - 1))
    assert st["peak_reserved"] <= budget * 1.1
    assert all(r["device_bytes"] <= foot for r in res)


def test_server_arena_equals_pool(hdg):
    """A server with a budget carves its contexts' buffers out of one device
    allocation of that size (freed pieces are reused by the next set-ups of
    other sizes); the results are bitwise those of contexts on the device
    pool, including when a slot gives its memory back under budget pressure."""
    specs = [(40, 40, 1), (96, 64, 3), (17, 33, 2), (128, 128, 1), (64, 64, 4), (50, 20, 2),
             (72, 72, 3), (24, 90, 1), (110, 110, 2), (31, 31, 4), (80, 40, 3), (60, 60, 2),
             (128, 96, 1), (45, 45, 3), (20, 20, 2), (100, 50, 4)]
    reqs = [_req(hdg, nx, ny, k, 900 + i, want_trace=True) for i, (nx, ny, k) in enumerate(specs)]
    with hdg.Server(slots=1) as s1:
        ref = []
        for r in reqs:
            res = s1.wait(s1.submit(r))
            assert res["status"] == 0
            ref.append((res, s1.trace(res["id"])))
    foot = max(hdg.footprint(r) for r in reqs)
    for budget in (int(3.5 * foot), int(1.2 * foot)):
        with hdg.Server(slots=3, hbm_budget=budget) as s:
            ids = [s.submit(r) for r in reqs]
            got = [s.wait(i) for i in ids]
            tr = [s.trace(i) for i, g in zip(ids, got)]
            st = s.stats()
        assert st["failed"] == 0 and st["peak_reserved"] <= budget
        for (ra, ta), rb, tb in zip(ref, got, tr):
            assert rb["status"] == 0 and rb["iters"] == ra["iters"]
            assert np.array_equal(ta, tb)


def test_server_rejects_and_isolates_failures(hdg):
    with hdg.Server(slots=2, hbm_budget=64 << 20) as s:
        with pytest.raises(hdg.HDGError, match="HBM_BUDGET"):
            s.submit(_req(hdg, 512, 512, 3, 1))
        with pytest.raises(hdg.HDGError, match="UNSUPPORTED"):
            s.submit(hdg.make_request(nx=8, ny=8, k=5))
        # k=3 at quad_degree 4 (the reference default k+1): singular local matrix
        bad = s.submit(hdg.make_request(nx=8, ny=8, k=3, quad_degree=4))
        good = s.submit(hdg.make_request(nx=16, ny=16, k=1, uex_kind=0))
        rb, rg = s.wait(bad), s.wait(good)
    assert rb["status"] == 4
    assert rg["status"] == 0 and 0 < rg["errornorm"] < 1e-3


def test_footprint_matches_device_use(hdg):
    """The admission estimate covers what a context really allocates."""
    for (n, k, prec) in [(64, 1, 0), (128, 2, 0), (100, 3, 1), (40, 4, 0)]:
        r = _req(hdg, n, n, k, 3, precision=prec)
        with hdg.Server(slots=1) as s:
            res = s.wait(s.submit(r))
        assert res["status"] == 0
        assert res["device_bytes"] <= hdg.footprint(r) <= 1.5 * res["device_bytes"] + (16 << 20)


def test_server_batching_equals_alone(hdg):
    """With continuous batching, queued requests that share options are
    solved together in grouped launches; each result is bitwise the one the
    request gets alone, and a request whose mesh folds over (negative
    Jacobian) fails alone inside its batch."""
    specs = [(48, 40, 2), (64, 64, 1), (33, 17, 3), (96, 80, 2), (20, 100, 1), (50, 50, 1),
             (40, 72, 3), (81, 81, 2), (16, 16, 1), (100, 60, 1), (64, 32, 3), (24, 24, 2)]
    reqs = [_req(hdg, nx, ny, k, 700 + i, want_trace=True) for i, (nx, ny, k) in enumerate(specs)]
    alone = []
    with hdg.Server(slots=1) as s1:
        for r in reqs:
            res = s1.wait(s1.submit(r))
            assert res["status"] == 0
            alone.append((res, s1.trace(res["id"])))
    folded = _req(hdg, 16, 16, 1, 799)
    folded.jitter = 0.95
    with hdg.Server(slots=2, batch=16) as sb:
        # two large requests hold both slots while the rest queue up
        busy = [sb.submit(_req(hdg, 256, 256, 3, 790 + i)) for i in range(2)]
        ids = [sb.submit(r) for r in reqs]
        bad = sb.submit(folded)
        res = [sb.wait(i) for i in ids]
        tr = [sb.trace(i) for i, r in zip(ids, res)]
        rbad = sb.wait(bad)
        assert all(sb.wait(i)["status"] == 0 for i in busy)
    for (ra, ta), rb, tb in zip(alone, res, tr):
        assert rb["status"] == 0 and rb["iters"] == ra["iters"]
        assert rb["true_relres"] == ra["true_relres"]
        assert np.array_equal(ta, tb)
    assert rbad["status"] == 2  # HDG_ERR_NEG_JACOBIAN
    # several requests shared one batch (same slot, same start)
    together = collections.Counter((r["slot"], r["t_start"]) for r in res + [rbad])
    assert max(together.values()) >= 3
