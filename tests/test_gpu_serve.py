"""Serving loop on the GPU (run with -m gpu): hpl_serve_* through the C ABI.

Bars:
  * every request's refined solution has backward error <= 1e-14 n and lies
    within 1e-10 (relative, inf-norm) of the FP64 oracle solution
    (gefa/gesl restated in FP64) -- host matrices of orders that need identity
    padding, several right-hand sides, FP8 trailing update;
  * the same request gives the same result alone and in the mix: a set of
    requests (device-generated systems, 20% of them reusing a resident factor)
    is run one at a time, then all at once with overlapping execution; the
    solutions must be bitwise identical (tolerance 0: the block size and
    padding depend on n alone and every kernel is deterministic);
  * admission by budget: a system larger than the factor budget is refused.
"""
import numpy as np
import pytest

from hpcc_fpga_amd import HplError, Server
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GB = 10 ** 9


def _say(msg):
    print(msg, flush=True)


def _fp64_solution(A, b):
    F = O.gefa_nopvt_f64(A.astype(np.float64))
    return O.gesl_nopvt_f64(F, b)


@pytest.fixture(scope="module")
def server():
    s = Server(slots=4, hbm_budget=6 * GB, prec="fp8", max_n=2048, max_nrhs=16)
    yield s
    s.close()


@pytest.mark.parametrize("n,nrhs,seed", [(100, 1, 1), (333, 4, 2), (777, 16, 3), (1024, 2, 4)])
def test_request_matches_fp64_oracle(server, n, nrhs, seed):
    _say(f"serve host system n={n} nrhs={nrhs}")
    A, _, _, _ = O.generate_sized(n, seed, 1, 1.2)
    rng = np.random.default_rng(seed)
    Bm = rng.uniform(-1, 1, (n, nrhs))
    t = server.submit(Bm, a=A)
    X, resp = server.wait(t)
    X = X.reshape(n, nrhs)
    assert resp["status"] == 0 and resp["backward_error"] <= 1e-14 * n
    assert resp["npad"] % resp["block"] == 0 and resp["npad"] >= n
    for j in range(nrhs):
        x_ref = _fp64_solution(A, Bm[:, j])
        fwd = np.max(np.abs(X[:, j] - x_ref)) / np.max(np.abs(x_ref))
        assert fwd <= 1e-10, (j, fwd)
        assert O.backward_error(A, X[:, j], Bm[:, j]) <= 1e-14 * n


def test_device_generated_system_is_the_sized_generator(server):
    n = 500
    _say(f"serve generated n={n}")
    rng = np.random.default_rng(0)
    b = rng.uniform(-1, 1, n)
    x, resp = server.wait(server.submit(b, seed=123))
    A, _, _, _ = O.generate_sized(n, 123, 0)
    x_ref = _fp64_solution(A, b)
    assert np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)) <= 1e-10
    assert O.backward_error(A, x, b) <= 1e-14 * n


def test_resident_factor_reuse_and_release(server):
    n = 640
    _say("serve reuse")
    rng = np.random.default_rng(1)
    x0, r0 = server.wait(server.submit(rng.uniform(-1, 1, n), seed=77, keep=True))
    fid = r0["factor"]
    assert fid != 0 and r0["reused"] == 0
    assert server.stats()["resident_factors"] >= 1
    B2 = rng.uniform(-1, 1, (n, 5))
    X, r1 = server.wait(server.submit(B2, factor=fid))
    assert r1["reused"] == 1 and r1["factor"] == fid
    A, _, _, _ = O.generate_sized(n, 77, 0)
    for j in range(5):
        assert O.backward_error(A, X[:, j], B2[:, j]) <= 1e-14 * n
    server.release(fid)
    with pytest.raises(HplError):
        server.submit(B2, factor=fid)


def test_budget_refusal():
    _say("serve budget")
    # 0.1 GB: the slot workspace (~10 MB at n <= 4096) plus a 90 MB factor
    # arena -- an order-4096 factor (~136 MB) can never fit, an order-512 one can
    with Server(slots=1, hbm_budget=int(0.1 * GB), prec="fp8", max_n=4096, max_nrhs=1) as s:
        with pytest.raises(HplError) as e:
            s.submit(np.ones(4096), seed=1)
        assert e.value.status == 5  # HPL_ERR_BUDGET
        x, r = s.wait(s.submit(np.ones(512), seed=1))  # a small one still fits
        assert r["status"] == 0


def _workload(count, seed):
    """Orders log-uniform in [64, 2048], nrhs in 1..16, every fifth request
    reusing the factor of an earlier kept one."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(64), np.log(2048)))))
        nrhs = int(rng.integers(1, 17))
        reqs.append({"n": n, "nrhs": nrhs, "seed": 1000 + i, "reuse_of": None,
                     "b": rng.uniform(-1, 1, (n, nrhs))})
    for i in range(4, count, 5):  # 20% reuse: same order as the kept request i-3
        j = i - 3
        reqs[i]["n"] = reqs[j]["n"]
        reqs[i]["b"] = rng.uniform(-1, 1, (reqs[j]["n"], reqs[i]["nrhs"]))
        reqs[i]["reuse_of"] = j
    return reqs


def test_same_result_alone_and_in_the_mix():
    reqs = _workload(30, 5)
    keepers = {r["reuse_of"] for r in reqs if r["reuse_of"] is not None}
    with Server(slots=4, hbm_budget=8 * GB, prec="fp8", max_n=2048, max_nrhs=16) as s:
        # alone: one request at a time
        _say("serve alone")
        alone, fids = {}, {}
        for i, r in enumerate(reqs):
            if r["reuse_of"] is None:
                t = s.submit(r["b"], seed=r["seed"], keep=i in keepers)
            else:
                t = s.submit(r["b"], factor=fids[r["reuse_of"]])
            x, resp = s.wait(t)
            assert resp["backward_error"] <= 1e-14 * r["n"]
            alone[i] = x
            if i in keepers:
                fids[i] = resp["factor"]
        for f in fids.values():
            s.release(f)
        # mix: everything submitted at once (fresh systems first so that the
        # reuses find their factors), waited for afterwards
        _say("serve mix")
        fids = {}
        tickets = {}
        for i, r in enumerate(reqs):
            if r["reuse_of"] is None:
                tickets[i] = s.submit(r["b"], seed=r["seed"], keep=i in keepers)
        mixed = {}
        for i in sorted(keepers):
            mixed[i], resp = s.wait(tickets.pop(i))
            fids[i] = resp["factor"]
        for i, r in enumerate(reqs):
            if r["reuse_of"] is not None:
                tickets[i] = s.submit(r["b"], factor=fids[r["reuse_of"]])
        for i, t in tickets.items():
            x, resp = s.wait(t)
            mixed[i] = x
        st = s.stats()
        _say(f"  max in flight {st['max_in_flight']}")
        assert st["max_in_flight"] >= 2  # requests overlapped on the device
        assert sorted(mixed) == list(range(len(reqs)))
        for i in mixed:
            assert np.array_equal(alone[i], mixed[i]), i
