"""CPU checks that pin the oracle (no GPU).

Each oracle function is checked against something other than itself:
  generate_ref      - known answer derived from the reference README's emulator
                      run (LINPACK/README.md:168-172: ||b||inf = resid /
                      (residn * n * eps) = 35.88236618041992 at n=32, seed 0),
                      an independent numpy MT19937 restatement (bit-exact),
                      and the invariants of test_host_reference_implementations.cpp:30-53;
  generate_sized    - independent numpy splitmix64 restatement (bit-exact),
                      dominance margin, b = M 1 in FP64;
  generate_batch_system, batch_orders
                    - the same splitmix64 restatement, identity pad, entry
                      ranges, b = M 1 (x = 1 in an FP64 solve; b = A 1 fails);
  gefa_nopvt_omp    - bit-identical to gefa_nopvt for 1, 3 and 8 threads;
  gefa_nopvt        - decoded factor reproduces M (numpy FP64), equals scipy's
                      LU (no row exchange happens for these column-dominant
                      matrices), brute force on 2x2/3x3, committed golden vector;
  gesl_nopvt        - the reference's own dtrsm_ref (oracle/_ref) and scipy
                      solve_triangular on the decoded factor;
  *_f64             - numpy.linalg.solve, the reference's FGMRES (gmres_ref);
  top/left/inner    - triangular-solve identities and a plain matmul;
  gefa_blocked      - bit-identical to gefa_nopvt;
  gefa_pvt/gesl_pvt - LAPACK getrf (scipy): same pivots (first index on exact
                      ties, as isamax), same L and U after the layout
                      conversion, P M = L U; brute force 3x3 with a known
                      pivot; the double-precision multiplier on an input where
                      float evaluation differs; numpy.linalg.solve; the
                      reference's own ReferenceSolveWithPivoting check
                      (test_host_reference_implementations.cpp:82-89);
  residn_ref/norms  - direct numpy evaluation of the reference formulas.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# --------------------------------------------------------------------------- generators
def test_ref_generator_readme_known_answer():
    # LINPACK/README.md:168-172 (n=32, B=16 emulator run, rank 0 -> seed 0)
    resid, residn, eps = 5.960464477539062e-07, 0.004354506590071576, 1.1920928955078125e-07
    normb_readme = resid / (residn * 32 * eps)
    _, b, _, normb = O.generate_ref(32, seed=0)
    assert normb == pytest.approx(normb_readme, rel=1e-12)
    assert float(np.max(b)) == normb


def _numpy_ref_generator(n, seed):
    """Independent restatement: numpy's MT19937 (init_genrand seeding, as
    std::mt19937(seed)) + libstdc++ generate_canonical<double,53> with two
    32-bit draws: u = (g1 + g2 * 2^32) / 2^64."""
    rs = np.random.RandomState(seed)
    g = rs.randint(0, 2 ** 32, size=2 * n * n, dtype=np.uint64).astype(np.float64)
    u = (g[0::2] + g[1::2] * 4294967296.0) / 18446744073709551616.0
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    A = u.astype(np.float32).reshape(n, n)
    for r in range(n):
        A[r, r] = 0.0
        s = np.float32(0.0)
        for v in A[r]:
            s = np.float32(s + v)
        A[r, r] = s
    b = np.zeros(n, np.float32)
    for j in range(n):
        s = np.float32(0.0)
        for v in A[:, j]:
            s = np.float32(s + v)
        b[j] = s
    return A, b


@pytest.mark.parametrize("n,seed", [(16, 0), (33, 1), (64, 5)])
def test_ref_generator_matches_independent_mt19937(n, seed):
    A, b, _, _ = O.generate_ref(n, seed=seed)
    A2, b2 = _numpy_ref_generator(n, seed)
    assert np.array_equal(A.view(np.uint32), A2.view(np.uint32))
    assert np.array_equal(b.view(np.uint32), b2.view(np.uint32))


def test_ref_generator_invariants():
    # test_host_reference_implementations.cpp:30-53
    A, b, norma, normb = O.generate_ref(128, seed=0, diag_dominant=False)
    assert np.all((A > 0) & (A < 1))
    A, b, norma, normb = O.generate_ref(128, seed=0)
    off = A - np.diag(np.diag(A))
    assert np.all((off[~np.eye(128, dtype=bool)] > 0) & (off[~np.eye(128, dtype=bool)] < 1))
    assert np.allclose(np.diag(A), off.sum(axis=1, dtype=np.float64), rtol=1e-5)
    assert norma == float(np.max(A))
    x = np.linalg.solve(A.T.astype(np.float64), b.astype(np.float64))
    assert np.max(np.abs(x - 1.0)) < 1e-5


def _splitmix_uniform(seed, stream, k):
    M = np.uint64(0xFFFFFFFFFFFFFFFF)

    def mix(z):
        z = np.asarray(z, np.uint64)
        with np.errstate(over="ignore"):
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return (z ^ (z >> np.uint64(31))) & M

    key = mix(np.uint64(seed) ^ (np.uint64(stream) << np.uint64(56)))
    with np.errstate(over="ignore"):
        z = mix(key + (np.asarray(k, np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
    return (z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def test_sized_generator_matches_independent_splitmix():
    n = 96
    A, b, b64, xt = O.generate_sized(n, 42, 0)
    U = _splitmix_uniform(42, 0, np.arange(n * n)).reshape(n, n) - np.float32(0.5)
    U[np.arange(n), np.arange(n)] += np.float32(n)
    assert np.array_equal(A.view(np.uint32), U.view(np.uint32))
    assert O.hash_uniform(42, 0, 5) == float(_splitmix_uniform(42, 0, [5])[0])
    # b = M 1 in FP64 (sum order differs from numpy: tolerance)
    assert np.allclose(b64, A.T.astype(np.float64).sum(axis=1), rtol=1e-14, atol=1e-12)
    assert np.array_equal(xt, np.ones(n))
    # dominance margin: |A_ii| >= n - 0.5 > sum_{j!=i} |A_ij| <= 0.5 (n-1)
    off = np.abs(A.T.astype(np.float64)).sum(axis=1) - np.abs(np.diag(A))
    assert np.all(np.abs(np.diag(A)) > off)


def test_sized_generator_mode1():
    n, d = 80, 1.1
    A, b, b64, xt = O.generate_sized(n, 7, 1, d)
    M = A.T.astype(np.float64)
    offs = np.abs(M).sum(axis=1) - np.abs(np.diag(M))
    assert np.allclose(np.diag(M), d * offs, rtol=1e-6)
    assert np.all((xt >= -1) & (xt < 1))
    assert np.allclose(b64, M @ xt, rtol=1e-13, atol=1e-12)
    assert np.array_equal(b, b64.astype(np.float32))
    xs = _splitmix_uniform(7, 1, np.arange(n)).astype(np.float32)
    assert np.array_equal(xt, (np.float32(2) * xs - np.float32(1)).astype(np.float64))


def test_batch_system_generator():
    """orc_generate_batch_system against an independent splitmix64
    restatement (stream 2 + s of the seed, entries U[-0.5,0.5) with n added on
    the diagonal), the identity pad and a zero right-hand side there, and the
    orientation of b: b = M 1 with M = A^T, so the FP64 solve gives x = 1
    (b = A 1 would not: A is not symmetric)."""
    seed = 7
    for s, n in [(0, 33), (5, 64), (11, 100)]:
        A, b, npad = O.generate_batch_system(n, seed, s)
        assert npad == (n + 31) // 32 * 32 and A.shape == (npad, npad)
        U = _splitmix_uniform(seed, 2 + s, np.arange(n * n)).reshape(n, n) - np.float32(0.5)
        U[np.arange(n), np.arange(n)] += np.float32(n)
        assert np.array_equal(A[:n, :n].view(np.uint32), U.view(np.uint32))
        off = A[:n, :n].copy()
        off[np.arange(n), np.arange(n)] = 0
        assert np.all((off >= -0.5) & (off < 0.5))
        assert np.all((np.diag(A)[:n] >= n - 0.5) & (np.diag(A)[:n] < n + 0.5))
        pad = A.copy()
        pad[:n, :n] = 0
        assert np.array_equal(pad, np.diag(np.r_[np.zeros(n), np.ones(npad - n)]).astype(np.float32))
        assert np.all(b[n:] == 0.0)
        M = A.T.astype(np.float64)
        x = np.linalg.solve(M, b)
        assert np.max(np.abs(x[:n] - 1.0)) <= 1e-13 and np.all(x[n:] == 0.0)
        assert np.max(np.abs(np.linalg.solve(A.astype(np.float64), b)[:n] - 1.0)) > 1e-6


def test_batch_orders_log_uniform():
    """batch_orders: n = round(exp(log lo + u (log hi - log lo))) with u the
    counter-based uniform of stream 1, inside [lo, hi]."""
    o = O.batch_orders(4096, 32, 1024, 7)
    u = _splitmix_uniform(7, 1, np.arange(4096)).astype(np.float64)
    assert np.array_equal(o, np.rint(np.exp(np.log(32) + u * (np.log(1024) - np.log(32)))))
    assert o.min() >= 32 and o.max() <= 1024
    # log-uniform: the median order is near sqrt(32 * 1024) = 181
    assert 160 <= np.median(o) <= 200


def test_gefa_nopvt_omp_bit_identical():
    """The parallel restatement (the OpenMP CPU timing) spreads the independent
    columns of each step over threads; every element receives the same
    operations in the same order, so it equals the sequential oracle bit for
    bit, for any thread count."""
    A, _, _, _ = O.generate_sized(300, 3, 0)
    F = O.gefa_nopvt(A)
    for th in (1, 3, 8):
        assert np.array_equal(O.gefa_nopvt_omp(A, th).view(np.uint32), F.view(np.uint32))


# --------------------------------------------------------------------------- factor / solve
def test_gefa_brute_force_2x2():
    # M = [[a, c], [b, d]] stored row-major as A = M^T = [[a, b], [c, d]]
    a, b_, c, d = 4.0, 1.0, 2.0, 5.0
    A = np.array([[a, b_], [c, d]], np.float32)
    F = O.gefa_nopvt(A)
    l = b_ / a
    u22 = d - l * c
    assert F[0, 0] == np.float32(-1.0 / a)
    assert F[0, 1] == pytest.approx(-l)          # -l stored below M's diagonal
    assert F[1, 0] == c                          # u_12
    assert F[1, 1] == pytest.approx(-1.0 / u22)


@pytest.mark.parametrize("n", [3, 17, 64])
def test_gefa_decoded_reproduces_matrix(n):
    A, _, _, _ = O.generate_ref(n, seed=n)
    F = O.gefa_nopvt_f64(A.astype(np.float64))
    L, U = O.decode_factor(F)
    M = A.T.astype(np.float64)
    assert np.allclose(L @ U, M, rtol=1e-13, atol=1e-12 * np.abs(M).max())
    P, Ls, Us = sla.lu(M)
    assert np.array_equal(P, np.eye(n))  # column dominance: no row exchange
    assert np.allclose(Ls, L, rtol=1e-12, atol=1e-14)
    assert np.allclose(Us, U, rtol=1e-12, atol=1e-12)
    F32 = O.gefa_nopvt(A)
    assert np.max(np.abs(F32 - F)) <= 1e-5 * np.max(np.abs(F))


def test_gesl_against_reference_dtrsm_and_scipy():
    n = 48
    A, b, _, normb = O.generate_ref(n, seed=2)
    F = O.gefa_nopvt(A)
    x = O.gesl_nopvt(F, b)
    L, U = O.decode_factor(F)
    y = sla.solve_triangular(L, b.astype(np.float64), lower=True, unit_diagonal=True)
    xs = sla.solve_triangular(U, y, lower=False)
    assert np.max(np.abs(x - xs)) < 1e-5
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    # LAPACK-form factor, column-major (= our row-major transpose layout)
    LU = np.ascontiguousarray((np.tril(L, -1) + U).T)  # col-major storage of L\U
    r = b.astype(np.float64).copy()
    R.dtrsm_ref(b"L", b"L", b"N", b"U", n, 1, 1.0, LU, n, r, n)
    R.dtrsm_ref(b"L", b"U", b"N", b"N", n, 1, 1.0, LU, n, r, n)
    assert np.max(np.abs(r - xs)) < 1e-10
    _, residn = O.residn_ref(x, normb)
    assert residn < 1.0


def test_f64_solve_against_numpy_and_reference_gmres():
    n = 64
    A, b, b64, xt = O.generate_sized(n, 7, 1, 1.1)
    M = A.T.astype(np.float64)
    F = O.gefa_nopvt_f64(A.astype(np.float64))
    x = O.gesl_nopvt_f64(F, b64)
    assert np.allclose(x, np.linalg.solve(M, b64), rtol=1e-12, atol=1e-13)
    assert np.allclose(x, xt, rtol=1e-10, atol=1e-11)
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    L, U = O.decode_factor(O.gefa_nopvt(A))  # FP32 factor as preconditioner
    LU = np.ascontiguousarray((np.tril(L, -1) + U).T)
    Acm = np.ascontiguousarray(A.astype(np.float64))  # col-major M
    xg = np.zeros(n)
    R.gmres_ref(n, Acm, n, xg, b64.copy(), LU, n, 50, 1, 1e-14)
    assert np.max(np.abs(xg - x)) < 1e-11


def test_block_closed_forms_against_triangular_identities():
    B = 32
    D, _, _, _ = O.generate_ref(B, seed=11)
    lu = O.gefa_nopvt(D)
    L, U = O.decode_factor(lu.astype(np.float64))
    rng = np.random.default_rng(0)
    X = rng.random((B, B)).astype(np.float32)
    top = O.top_update(X, lu)       # U^T top = -X  (rows of M's L-panel)
    assert np.allclose(U.T @ top.astype(np.float64), -X, rtol=1e-4, atol=1e-5)
    left = O.left_update(X, lu)     # L left^T = X^T  (M's U-panel)
    assert np.allclose(L @ left.T.astype(np.float64), X.T, rtol=1e-4, atol=1e-5)
    Cm = rng.random((B, B)).astype(np.float32)
    inner = O.inner_update(Cm, left, top)
    ref = Cm.astype(np.float64) + left.astype(np.float64) @ top.astype(np.float64)
    bound = B * np.finfo(np.float32).eps * (np.abs(Cm) + np.abs(left) @ np.abs(top))
    assert np.all(np.abs(inner - ref) <= bound + 1e-7)


@pytest.mark.parametrize("n,B", [(64, 16), (96, 32), (128, 64)])
def test_blocked_equals_unblocked(n, B):
    A, _, _, _ = O.generate_ref(n, seed=1)
    assert np.array_equal(O.gefa_blocked(A, B).view(np.uint32), O.gefa_nopvt(A).view(np.uint32))


def test_metrics():
    x = np.array([1.0, 1.5, 0.75], np.float32)
    resid, residn = O.residn_ref(x, 2.0)
    assert resid == 0.5
    assert residn == pytest.approx(0.5 / (3 * 2.0 * np.finfo(np.float32).eps))
    A, b, _, _ = O.generate_ref(20, seed=4)
    xs = np.linspace(-1, 1, 20)
    rn, an, xn, bn = O.norms(A, xs, b.astype(np.float64))
    M = A.T.astype(np.float64)
    assert rn == pytest.approx(np.max(np.abs(b - M @ xs)), rel=1e-12)
    assert an == pytest.approx(np.max(np.abs(M).sum(axis=1)), rel=1e-12)
    assert xn == 1.0 and bn == pytest.approx(float(np.max(np.abs(b))))


# --------------------------------------------------------------------------- golden vectors
def test_golden_vectors():
    """Committed fixtures (tests/golden/make_golden.py): the oracle must keep
    reproducing them bit for bit."""
    path = os.path.join(GOLDEN, "linpack_golden.npz")
    g = np.load(path, allow_pickle=False)
    A, b, _, normb = O.generate_ref(32, seed=0)
    assert np.array_equal(A, g["ref32_A"]) and np.array_equal(b, g["ref32_b"])
    F = O.gefa_nopvt(A)
    assert np.array_equal(F, g["ref32_factor"])
    assert np.array_equal(O.gesl_nopvt(F, b), g["ref32_x"])
    As, bs, _, _ = O.generate_sized(64, 42, 0)
    assert np.array_equal(As, g["sized64_A"]) and np.array_equal(bs, g["sized64_b"])
    assert float(g["readme_normb"]) == normb


# --------------------------------------------------------------------------- mixed precision
def _bf16_bits(v):
    """bfloat16 rounding by the usual bit trick (ties to even on the kept bits)."""
    u = np.float32(v).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return float(np.uint32(r).view(np.float32))


def _e4m3_table():
    vals = []
    for e in range(16):
        for m in range(8):
            if e == 15 and m == 7:
                continue  # NaN code
            vals.append(m * 2.0 ** -9 if e == 0 else (1 + m / 8) * 2.0 ** (e - 7))
    return np.array(vals)


def test_low_precision_rounding_against_independent_definitions():
    """round_lp: FP16 against numpy's float16 cast, BF16 against the bit
    trick, FP8 E4M3 against brute-force nearest value of the code table (ties
    to the even code)."""
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(4000) * np.exp2(rng.integers(-20, 12, 4000))).astype(np.float32)
    assert all(O.round_lp(v, "fp16") == float(np.float16(v)) for v in x)
    assert all(O.round_lp(v, "bf16") == _bf16_bits(v) for v in x)
    vals = _e4m3_table()
    xs = rng.uniform(-448, 448, 4000) * np.exp2(rng.integers(-12, 1, 4000))
    # exact midpoints between neighbouring codes exercise ties-to-even
    mids = (vals[1:] + vals[:-1]) / 2
    for v in np.concatenate([xs, mids, -mids[::7]]):
        d = np.abs(vals - abs(v))
        cand = np.flatnonzero(d == d.min())
        code = cand[0] if len(cand) == 1 else [c for c in cand if c % 2 == 0][0]
        assert O.round_lp(v, "fp8") == np.copysign(vals[code], v), v


def test_fp8_block_scale_is_smallest_power_of_two():
    for amax in [448.0, 449.0, 1.0, 1e-30, 3.5e5, 0.017]:
        s = O.fp8_scale(amax)
        assert s == 2.0 ** round(np.log2(s))
        assert amax <= 448 * s and amax > 448 * s / 2
    assert O.fp8_scale(0.0) == 1.0


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp8"])
def test_lp_blocked_factor_against_numpy_restatement(prec):
    """One trailing update (n = 2B): the oracle's low-precision blocked
    factorisation equals a numpy restatement of quantise-and-update."""
    n, B = 64, 32
    A, _, _, _ = O.generate_ref(n, seed=5)
    got = O.gefa_blocked_lp(A, B, prec)
    a = A.copy()
    a[:B, :B] = O.gefa_nopvt(a[:B, :B])
    a[:B, B:] = O.top_update(a[:B, B:], a[:B, :B])
    a[B:, :B] = O.left_update(a[B:, :B], a[:B, :B])
    T = a[:B, B:].T.astype(np.float64)  # vectors = columns of the top strip
    L = a[B:, :B].astype(np.float64)    # vectors = rows of the left strip

    def q(V):
        if prec == "fp16":
            return V.astype(np.float16).astype(np.float64)
        if prec == "bf16":
            return np.vectorize(_bf16_bits)(V)
        out = np.empty_like(V)
        for g in range(0, V.shape[0], 32):
            blk = V[g:g + 32]
            amax = np.max(np.abs(blk))
            e = int(np.ceil(np.log2(amax / 448)))
            s = 2.0 ** e
            out[g:g + 32] = s * np.vectorize(lambda v: O.round_lp(v, "fp8"))(blk / s)
        return out

    upd = (a[B:, B:].astype(np.float64) + q(L) @ q(T).T).astype(np.float32)
    a[B:, B:] = O.gefa_nopvt(upd)
    assert np.array_equal(got[:B], a[:B]) and np.array_equal(got[B:, :B], a[B:, :B])
    assert np.max(np.abs(got[B:, B:] - a[B:, B:])) <= 1e-6 * np.max(np.abs(a[B:, B:]))


def test_lp_factor_error_ordered_by_precision():
    """Coarser update formats move the factor further from the FP32 one."""
    A, _, _, _ = O.generate_ref(256, seed=2)
    ref = O.gefa_blocked(A, 64)
    err = {p: np.max(np.abs(O.gefa_blocked_lp(A, 64, p) - ref)) for p in ("fp16", "bf16", "fp8")}
    assert 0 < err["fp16"] < err["bf16"] < err["fp8"]


# --------------------------------------------------------------------------- partial pivoting
def _lapack_lu_from_pvt(F, ipvt):
    """(L, U) of P M = L U from a gefa_pvt factor (negated multipliers)."""
    G = O.pvt_to_lapack(F, ipvt).T.astype(np.float64)  # M layout
    L = np.eye(G.shape[0]) - np.tril(G, -1)
    U = np.triu(G)
    return L, U


@pytest.mark.parametrize("kind", ["uniform01", "normal"])
def test_gefa_pvt_against_lapack_getrf(kind):
    n = 96
    rng = np.random.default_rng(7)
    A = (rng.random((n, n)) if kind == "uniform01" else rng.standard_normal((n, n))).astype(
        np.float32)
    F, ipvt = O.gefa_pvt(A)
    M = A.T.astype(np.float64)
    lu, piv = sla.lu_factor(M)
    assert np.array_equal(ipvt, piv)
    L, U = _lapack_lu_from_pvt(F, ipvt)
    assert np.allclose(L, np.tril(lu, -1) + np.eye(n), atol=2e-5)
    assert np.allclose(U, np.triu(lu), atol=2e-5 * np.abs(lu).max())
    perm = O.pivot_permutation(ipvt)
    assert np.allclose(L @ U, M[perm], atol=1e-5 * np.abs(M).max())
    assert np.array_equal(O.lapack_to_pvt(O.pvt_to_lapack(F, ipvt), ipvt), F)


def test_gefa_pvt_ties_take_first_index():
    # small integers give exact ties in |M(i,k)| at many steps; LAPACK's isamax
    # and gefa_ref's strict '<' both keep the first
    rng = np.random.default_rng(5)
    n = 40
    A = rng.integers(-2, 3, size=(n, n)).astype(np.float32)
    A[np.arange(n), np.arange(n)] += 0.5
    F, ipvt = O.gefa_pvt(A)
    _, piv = sla.lu_factor(A.T.astype(np.float64))
    assert np.array_equal(ipvt, piv)
    # column 0 of M is A[0]: the pivot is the first of its largest magnitudes
    col = np.abs(A[0].astype(np.float64))
    assert ipvt[0] == int(np.flatnonzero(col == col.max())[0])


def test_gefa_pvt_brute_force_3x3():
    # M rows: [1 2 3], [4 5 6], [7 8 10]; column 0 pivots on row 2 (|7|)
    M = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 10]], np.float64)
    A = np.ascontiguousarray(M.T.astype(np.float32))
    F, ipvt = O.gefa_pvt(A)
    assert ipvt[0] == 2
    # by hand: after exchanging rows 0 and 2, l = (4/7, 1/7); the 2x2 remainder
    # [[5-32/7, 6-40/7], [2-8/7, 3-10/7]] = [[3/7, 2/7], [6/7, 11/7]] pivots on 6/7
    assert ipvt[1] == 2 and ipvt[2] == 2
    L, U = _lapack_lu_from_pvt(F, ipvt)
    P = M[O.pivot_permutation(ipvt)]
    assert np.allclose(L @ U, P, atol=1e-5)
    assert U[0, 0] == 7.0 and U[1, 1] == pytest.approx(6 / 7) and F[0, 0] == 7.0
    # gefa_ref's layout: step 0 stored -(4/7, 1/7) in rows 1, 2 before step 1
    # exchanged rows 1 and 2 in columns >= 1 only
    assert F[0, 1] == pytest.approx(-4 / 7) and F[0, 2] == pytest.approx(-1 / 7)


def test_gefa_pvt_multiplier_in_double():
    # gefa_ref: a[k*lda+i] *= -1.0 / a[k*lda+k] -- the double literal makes
    # the reciprocal and the product double, rounded to float once.  Find an
    # input where evaluating it in float would round differently.
    rng = np.random.default_rng(0)
    for _ in range(100000):
        akk, aik = rng.random(2).astype(np.float32) + np.float32(0.5)
        if aik >= akk:
            continue
        dbl = np.float32(np.float64(aik) * (-1.0 / np.float64(akk)))
        flt = np.float32(aik * (np.float32(-1.0) / akk))
        if dbl != flt:
            break
    else:
        pytest.fail("no distinguishing input found")
    A = np.array([[akk, aik], [0.25, 1.0]], np.float32)  # M = [[akk, .25], [aik, 1]]
    F, ipvt = O.gefa_pvt(A)
    assert ipvt[0] == 0
    assert F[0, 1] == dbl and F[0, 1] != flt


def test_gesl_pvt_against_numpy_and_reference_check():
    n = 128
    A, b, _, normb = O.generate_ref(n, seed=0, diag_dominant=False)
    F, ipvt = O.gefa_pvt(A)
    x = O.gesl_pvt(F, b, ipvt)
    xs = np.linalg.solve(A.T.astype(np.float64), b.astype(np.float64))
    assert np.max(np.abs(x - xs)) < 1e-3
    _, residn = O.residn_ref(x, normb)  # ReferenceSolveWithPivoting: validateOutput
    assert residn < 1.0
    # a zero leading pivot: the unpivoted factorisation divides by zero, the
    # pivoted one exchanges it away
    A0 = A.copy()
    A0[0, 0] = 0.0
    b0 = A0.T.astype(np.float64).sum(axis=1).astype(np.float32)
    with np.errstate(all="ignore"):
        xn = O.gesl_nopvt(O.gefa_nopvt(A0), b0)
    assert not np.all(np.isfinite(xn))
    F0, ipvt0 = O.gefa_pvt(A0)
    assert ipvt0[0] != 0
    assert O.hpl_scaled_residual(A0, O.gesl_pvt(F0, b0, ipvt0).astype(np.float64),
                                 b0.astype(np.float64)) < 1.0


def test_ref_generator_f64_matches_float_draws():
    """The DATA_TYPE=double generator draws the same mt19937 sequence; the
    float build rounds each draw, the double build keeps it."""
    n = 48
    A64, b64 = O.generate_ref_f64(n, seed=3)
    A32, b32, _, _ = O.generate_ref(n, seed=3)
    off = ~np.eye(n, dtype=bool)
    assert np.array_equal(A64.astype(np.float32)[off], A32[off])
    assert np.allclose(np.diag(A64), A64.sum(axis=1) - np.diag(A64), rtol=1e-14)
    assert np.allclose(b64, A64.sum(axis=0), rtol=1e-14)
    x = np.linalg.solve(A64.T, b64)
    assert np.max(np.abs(x - 1.0)) < 1e-13
