"""Batched uneven solves (one persistent, graph-captured launch) against the
CPU oracle: per system, pivots bit-exact with the reference's row-major LU
(src/lapack/getrf.rs:46-120) and backward error of X <= n * eps of the
arithmetic; graph replays reproduce the eager result bit for bit."""
import numpy as np
import pytest

import lair_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_batch(count, seed, dtype, nmin=1, nmax=256):
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, count).astype(np.int32)
    rs = rng.choice(np.array([1, 2, 4], np.int32), count)
    As = [(rng.standard_normal((n, n)) + n * np.eye(n)).astype(dtype) for n in ns]
    Bs = [rng.standard_normal((n, r)).astype(dtype) for n, r in zip(ns, rs)]
    return ns, rs, As, Bs


def run_batch(ctx, ns, rs, As, Bs, dtype, replays=1, autotune=True):
    a = np.concatenate([A.ravel() for A in As]).reshape(-1, 1)
    b = np.concatenate([B.ravel() for B in Bs]).reshape(-1, 1)
    da = L.DeviceArray.from_host(a)
    db = L.DeviceArray.from_host(b)
    dp = L.DeviceArray(int(ns.sum()), 1, np.int64)
    di = L.DeviceArray(len(ns), 1, np.int32)
    plan = L.BatchedPlan(ctx, ns, rs, da, db, dp, di, precision=0 if dtype == np.float64 else 1,
                         autotune=autotune)
    outs = []
    for _ in range(replays):
        da.upload(a)
        db.upload(b)
        plan.run()
        ctx.synchronize()
        outs.append((da.download().ravel(), db.download().ravel(), dp.download().ravel(),
                     di.download().ravel()))
    cfg = plan.config()
    plan.close()
    return outs, cfg


def split(ns, rs, lu, x, piv):
    res = []
    oa = ob = op = 0
    for n, r in zip(ns, rs):
        res.append((lu[oa:oa + n * n].reshape(n, n), x[ob:ob + n * r].reshape(n, r), piv[op:op + n]))
        oa += n * n
        ob += n * r
        op += n
    return res


@pytest.fixture(scope="module")
def ctx():
    c = L.Context()
    yield c
    c.close()


# Without the autotuner the plan keeps its first candidate: 16-column panels
# with every system updated in global memory; the autotuned plans of these
# batches stage the smaller systems in LDS.
@pytest.mark.parametrize("autotune", [True, False])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_matches_oracle(ctx, dtype, autotune):
    ns, rs, As, Bs = make_batch(300, 11 if dtype == np.float64 else 12, dtype)
    outs, cfg = run_batch(ctx, ns, rs, As, Bs, dtype, autotune=autotune)
    if not autotune:
        assert cfg["nb"] == 16 and cfg["lds_elems"] == 0
    lu, x, piv, info = outs[0]
    eps = np.finfo(dtype).eps
    for i, (lui, xi, pi) in enumerate(split(ns, rs, lu, x, piv)):
        n = ns[i]
        ref = As[i].copy()
        if dtype == np.float64:
            piv_o, sing_o = O.getrf(ref)
        else:
            piv_o, sing_o = O.sgetrf(ref)
        assert info[i] == (-1 if sing_o is None else sing_o)
        np.testing.assert_array_equal(pi, piv_o, err_msg=f"system {i} n={n}")
        a64 = As[i].astype(np.float64)
        r = Bs[i].astype(np.float64) - a64 @ xi.astype(np.float64)
        be = np.abs(r).max(0) / (np.abs(a64).sum(1).max() * np.abs(xi).max(0))
        assert be.max() <= n * eps, (i, n, be.max())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_extreme_shapes(ctx, dtype):
    """The limits of the batched path: n in {1, 2, 255, 256} with up to 16
    right-hand sides (at n = 256 and 16 of them, Y fills its whole LDS region
    in the back substitution), mixed with mid sizes."""
    rng = np.random.default_rng(41 if dtype == np.float64 else 42)
    ns = np.array([256, 256, 255, 1, 2, 256, 100, 1, 255, 64, 2, 17], np.int32)
    rs = np.array([16, 1, 16, 16, 16, 8, 16, 1, 3, 16, 5, 16], np.int32)
    As = [(rng.standard_normal((n, n)) + n * np.eye(n)).astype(dtype) for n in ns]
    Bs = [rng.standard_normal((n, r)).astype(dtype) for n, r in zip(ns, rs)]
    eps = np.finfo(dtype).eps
    for autotune in (True, False):
        outs, _ = run_batch(ctx, ns, rs, As, Bs, dtype, autotune=autotune)
        lu, x, piv, info = outs[0]
        assert (info == -1).all()
        for i, (lui, xi, pi) in enumerate(split(ns, rs, lu, x, piv)):
            ref = As[i].copy()
            piv_o, _ = O.getrf(ref) if dtype == np.float64 else O.sgetrf(ref)
            np.testing.assert_array_equal(pi, piv_o, err_msg=f"system {i} n={ns[i]}")
            a64 = As[i].astype(np.float64)
            r = Bs[i].astype(np.float64) - a64 @ xi.astype(np.float64)
            be = np.abs(r).max(0) / (np.abs(a64).sum(1).max() * np.abs(xi).max(0))
            assert be.max() <= ns[i] * eps, (autotune, i, ns[i], rs[i], be.max())


def test_batched_graph_replays_are_identical(ctx):
    ns, rs, As, Bs = make_batch(200, 21, np.float64, 1, 200)
    outs, _ = run_batch(ctx, ns, rs, As, Bs, np.float64, replays=3)
    for o in outs[1:]:
        for u, v in zip(outs[0], o):
            assert np.array_equal(u, v)


def test_batched_singular_system_reports_last_zero_pivot(ctx):
    ns, rs, As, Bs = make_batch(20, 31, np.float64, 20, 60)
    As[5][:, 3] = 0.0
    As[5][:, 17] = 0.0
    outs, _ = run_batch(ctx, ns, rs, As, Bs, np.float64)
    info = outs[0][3]
    ref = As[5].copy()
    _, sing_o = O.getrf(ref)
    assert sing_o == 17 and info[5] == 17
    assert all(info[i] == -1 for i in range(len(ns)) if i != 5)


def test_batched_c4_full_size(ctx):
    """20,000 systems, n ~ U{8..256}, A = N(0,1) + n I, nrhs in {1,2,4}: every
    system's backward error, and pivots of a sample against the oracle."""
    count = 20000
    rng = np.random.default_rng(2024)
    ns = rng.integers(8, 257, count).astype(np.int32)
    rs = rng.choice(np.array([1, 2, 4], np.int32), count)
    As = [rng.standard_normal((n, n)) + n * np.eye(n) for n in ns]
    Bs = [rng.standard_normal((n, r)) for n, r in zip(ns, rs)]
    outs, cfg = run_batch(ctx, ns, rs, As, Bs, np.float64)
    lu, x, piv, info = outs[0]
    assert (info == -1).all()
    eps = np.finfo(np.float64).eps
    parts = split(ns, rs, lu, x, piv)
    for i, (lui, xi, pi) in enumerate(parts):
        r = Bs[i] - As[i] @ xi
        be = np.abs(r).max(0) / (np.abs(As[i]).sum(1).max() * np.abs(xi).max(0))
        assert be.max() <= ns[i] * eps, (i, ns[i], be.max())
    for i in range(0, count, 200):
        ref = As[i].copy()
        piv_o, _ = O.getrf(ref)
        np.testing.assert_array_equal(parts[i][2], piv_o)
