"""GPU tests of batches (hdg_batch_*): requests of one order and options
are packed into one mesh and share every launch; each instance must give,
bit for bit, what it gives solved alone (a one-slot server runs hdg_run on
the request by itself)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _req(hdg, nx, ny, k, seed, kind=None, precision=0, precond=0, uex=-1):
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(1, 8, size=2)
    if kind is None:
        kind = hdg._lib.SRC_SINE_SERIES if seed % 2 else hdg._lib.SRC_KAPPA
    return hdg.make_request(nx=nx, ny=ny, k=k, jitter=0.2, seed=seed, src_kind=kind, kappa=(a, b),
                            coef=rng.standard_normal(16), precision=precision, precond=precond,
                            uex_kind=uex if kind == hdg._lib.SRC_KAPPA else -1, want_trace=True)


def _alone(hdg, reqs):
    out = []
    with hdg.Server(slots=1) as s:
        for r in reqs:
            res = s.wait(s.submit(r))
            assert res["status"] == 0, res
            out.append((res, s.trace(res["id"])))
    return out


def _check_same(hdg, reqs, expect_groups):
    alone = _alone(hdg, reqs)
    with hdg.Batch(reqs) as b:
        assert b.info()["groups"] == expect_groups
        res = b.run()
        traces = [b.trace(i) for i in range(len(reqs))]
        # a second run on the resident batch repeats the same bits
        res2 = b.run()
        traces2 = [b.trace(i) for i in range(len(reqs))]
    for i, ((ra, ta), rb, tb, rb2, tb2) in enumerate(zip(alone, res, traces, res2, traces2)):
        print(f"instance {i}: n_dofs {rb['n_dofs']} iters {rb['iters']} alone {ra['iters']}", flush=True)
        assert rb["status"] == 0
        assert rb["n_dofs"] == ra["n_dofs"] and rb["n_cells"] == ra["n_cells"]
        assert rb["iters"] == ra["iters"]
        assert rb["relres"] == ra["relres"] and rb["true_relres"] == ra["true_relres"]
        assert rb["errornorm"] == ra["errornorm"]
        assert np.array_equal(ta, tb)
        assert np.array_equal(tb, tb2) and rb2["iters"] == rb["iters"]
        assert rb["true_relres"] < 1e-10
    return res


def test_batch_multigrid_orders_equal_alone(hdg):
    """Uneven sizes (square, rectangular, odd, tiny) and three orders: three
    groups; every instance bitwise equal to its solo solve."""
    shapes = [(16, 16, 1), (33, 17, 2), (5, 4, 3), (40, 72, 1), (2, 3, 2), (64, 64, 3), (17, 9, 1),
              (96, 80, 2), (48, 48, 3), (1, 1, 1), (100, 60, 2), (24, 31, 3)]
    reqs = [_req(hdg, nx, ny, k, 300 + i, uex=0) for i, (nx, ny, k) in enumerate(shapes)]
    res = _check_same(hdg, reqs, expect_groups=3)
    # manufactured solution instances report their L2 error
    for r, q in zip(res, reqs):
        if q.uex_kind == 0:
            assert 0 < r["errornorm"] < 1e-2


def test_batch_mixed_precision_and_block_jacobi(hdg):
    """Mixed-precision multigrid and block-Jacobi groups beside FP64 ones."""
    reqs = [_req(hdg, 32, 32, 3, 401, precision=1), _req(hdg, 57, 23, 3, 402, precision=1),
            _req(hdg, 20, 20, 1, 403, precond=1), _req(hdg, 12, 30, 2, 404, precond=1),
            _req(hdg, 44, 44, 3, 405), _req(hdg, 8, 8, 4, 406), _req(hdg, 30, 14, 4, 407)]
    _check_same(hdg, reqs, expect_groups=5)


def test_batch_many_instances(hdg):
    """40 instances with n log-uniform in [16, 128] and k in {1, 2, 3}, as
    the batched workload draws them (smaller)."""
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(40):
        n = int(round(np.exp(rng.uniform(np.log(16), np.log(128)))))
        reqs.append(_req(hdg, n, n, int(rng.integers(1, 4)), 500 + i, kind=hdg._lib.SRC_KAPPA, uex=0))
    _check_same(hdg, reqs, expect_groups=len({q.opt.k for q in reqs}))


def test_batch_rejects_bad_requests(hdg):
    with pytest.raises(hdg.HDGError, match="k=5"):
        hdg.Batch([_req(hdg, 8, 8, 1, 1), hdg.make_request(nx=8, ny=8, k=5)])
    bad = _req(hdg, 8, 8, 1, 2)
    bad.src.kind = hdg._lib.SRC_VALUES
    with pytest.raises(hdg.HDGError, match="source kind"):
        hdg.Batch([bad])
    with pytest.raises(hdg.HDGError, match="budget"):
        hdg.Batch([_req(hdg, 256, 256, 3, 3)], hbm_budget=16 << 20)
