"""Mixed-precision trailing update + FP64 iterative refinement (run with -m gpu).

Bars:
  * BF16/FP16 factor vs the oracle's FP32 gefa_ref_nopvt factor: relative
    max difference <= 2e-2 (BF16, unit roundoff 2^-8) / 4e-3 (FP16, 2^-11) --
    a wrong MFMA operand layout gives O(1) differences;
  * BF16/FP16/FP8 factor vs the oracle's emulation of the same quantised
    update (gefa_blocked_lp): the RMS distance to the emulation is at most a
    quarter of the emulation's own RMS distance to the FP32 factor, i.e. the
    GPU's departure from FP32 is the modelled quantisation (rounding mode,
    FP8 code and block scale) and not something else;
  * refinement: normwise backward error <= 1e-14 * n (the target of the
    mixed-precision configuration) and the refined x within 1e-10 (relative,
    inf-norm) of the FP64 oracle solution (gefa_ref_nopvt/gesl_ref_nopvt
    restated in FP64, cross-checked against numpy and the reference's FGMRES
    in the CPU tests).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,tol", [("bf16", 2e-2), ("fp16", 4e-3)])
def test_lowprec_factor_close_to_fp32(ctx, prec, tol):
    n, B = 1024, 128
    print(f"lowprec factor {prec}", flush=True)
    A, b, _, _ = O.generate_sized(n, 42, 0)
    ctx.keep_original(False)
    ctx.upload(A, b)
    ctx.gefa_dev(B, prec)
    F, _ = ctx.download(matrix=True, solution=False)
    F_ref = O.gefa_nopvt(A)
    err = np.max(np.abs(F - F_ref)) / np.max(np.abs(F_ref))
    assert err <= tol, err


@pytest.mark.parametrize("prec", ["bf16"])
def test_strip_solve_by_inverse(ctx, prec):
    """The BF16 factorisation solves the strips as products with the
    diagonal block's full inverse, split into high and low BF16 halves (about
    2^-17 relative).  Step 0's strips see no earlier update, so they equal the
    oracle's first-panel strips to that accuracy (FP32 recurrence: ~1e-7); a
    wrong operand layout or a dropped half gives 1e-3 or worse.  The 32 x 32
    recurrence (HPL_STRIP_SOLVE=recurrence) must meet the same bar."""
    import os

    from hpcc_fpga_amd import Context

    n, B = 1024, 256
    print(f"strip solve by inverse {prec}", flush=True)
    A, _, _, _ = O.generate_sized(n, 5, 0)
    F_ref = O.gefa_nopvt(A)
    ref = np.concatenate([F_ref[:B, B:].ravel(), F_ref[B:, :B].ravel()])
    scale = np.max(np.abs(ref))
    errs = {}
    for mode in ("inverse", "recurrence"):
        if mode == "recurrence":
            os.environ["HPL_STRIP_SOLVE"] = "recurrence"
        try:
            c = ctx if mode == "inverse" else Context(0)
        finally:
            os.environ.pop("HPL_STRIP_SOLVE", None)
        try:
            c.keep_original(False)
            c.upload(A)
            c.gefa_dev(B, prec)
            F, _ = c.download(matrix=True, solution=False)
        finally:
            if c is not ctx:
                c.close()
        got = np.concatenate([F[:B, B:].ravel(), F[B:, :B].ravel()])
        errs[mode] = float(np.max(np.abs(got - ref)) / scale)
    print(f"  first-panel strips vs oracle: {errs}", flush=True)
    assert errs["inverse"] <= 5e-5, errs
    assert errs["recurrence"] <= 5e-6, errs


@pytest.mark.parametrize("prec,n,B", [("bf16", 512, 64), ("fp16", 512, 64), ("fp8", 512, 64),
                                      ("fp8", 768, 128), ("bf16", 1024, 256), ("fp16", 1536, 256)])
def test_lowprec_factor_matches_quantised_emulation(ctx, prec, n, B):
    print(f"quantised emulation {prec} n={n} B={B}", flush=True)
    A, _, _, _ = O.generate_ref(n, seed=9)  # positive entries: coherent, large updates
    ctx.keep_original(False)
    ctx.upload(A)
    ctx.gefa_dev(B, prec)
    F, _ = ctx.download(matrix=True, solution=False)
    emu = O.gefa_blocked_lp(A, B, prec).astype(np.float64)
    f32 = O.gefa_blocked(A, B).astype(np.float64)
    rms = lambda d: float(np.sqrt(np.mean(d * d)))  # noqa: E731
    d_gpu = rms(F - emu)
    d_model = rms(emu - f32)
    print(f"  rms(gpu-emu)={d_gpu:.3e} rms(emu-fp32)={d_model:.3e}", flush=True)
    assert d_model > 0
    assert d_gpu <= 0.25 * d_model


@pytest.mark.parametrize("prec,d", [("bf16", 2.0), ("bf16", 1.1), ("fp16", 1.1), ("fp8", 2.0),
                                    ("fp8", 1.1)])
def test_refinement_reaches_fp64_accuracy(ctx, prec, d):
    n, B = 2048, 256
    print(f"refinement {prec} d={d}", flush=True)
    ctx.keep_original(True)
    ctx.generate(n, 1, 7, d)
    res = ctx.solve_ir(B, prec, max_iter=30)
    assert res["converged"] == 1
    assert res["backward_error"] <= 1e-14 * n
    x = ctx.download_solution64()
    A, b32, b64, xt = O.generate_sized(n, 7, 1, d)
    F = O.gefa_nopvt_f64(A.astype(np.float64))
    x_ref = O.gesl_nopvt_f64(F, b64)
    fwd = np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref))
    assert fwd <= 1e-10, fwd
    # independent residual on the CPU
    be = O.backward_error(A, x, b64)
    assert be <= 1e-14 * n, be
    ctx.keep_original(False)


@pytest.mark.parametrize("prec,nrhs", [("bf16", 1), ("fp8", 3), ("fp8", 16), ("fp32", 5)])
def test_further_right_hand_sides(ctx, prec, nrhs):
    """hpl_solve_rhs64: new right-hand sides with the resident factor, each
    refined to 1e-14 n and within 1e-10 of the FP64 oracle solution."""
    n, B = 1024, 128
    print(f"further rhs {prec} nrhs={nrhs}", flush=True)
    ctx.keep_original(True)
    ctx.generate(n, 1, 5, 1.5)
    ctx.solve_ir(B, prec, max_iter=30)
    rng = np.random.default_rng(nrhs)
    Bm = rng.uniform(-1, 1, (n, nrhs))
    X, res = ctx.solve_rhs64(Bm if nrhs > 1 else Bm[:, 0], max_iter=30)
    X = X.reshape(n, nrhs)
    A, _, _, _ = O.generate_sized(n, 5, 1, 1.5)
    F = O.gefa_nopvt_f64(A.astype(np.float64))
    for j in range(nrhs):
        assert res[j]["converged"] == 1 and res[j]["backward_error"] <= 1e-14 * n
        x_ref = O.gesl_nopvt_f64(F, Bm[:, j])
        fwd = np.max(np.abs(X[:, j] - x_ref)) / np.max(np.abs(x_ref))
        assert fwd <= 1e-10, (j, fwd)
        assert O.backward_error(A, X[:, j], Bm[:, j]) <= 1e-14 * n
    ctx.keep_original(False)


def test_refinement_fp32_factor(ctx):
    """FP32 trailing update through the same refinement loop."""
    n, B = 1024, 128
    ctx.keep_original(True)
    ctx.generate(n, 1, 3, 1.5)
    res = ctx.solve_ir(B, "fp32", max_iter=10)
    assert res["converged"] == 1 and res["iterations"] <= 3
    ctx.keep_original(False)


@pytest.mark.parametrize("prec,d", [("bf16", 1.1), ("fp8", 1.01)])
def test_gmres_refinement_matches_fp64_and_reference_gmres(ctx, prec, d):
    """hpl_solve_gmres reaches 1e-14 n; its x agrees with the FP64 oracle
    solution and with the reference's own gmres_ref (LINPACK/src/host/gmres.c,
    compiled unchanged into oracle/_ref) run with an FP64 LU preconditioner."""
    n, B = 1024, 128
    print(f"gmres {prec} d={d}", flush=True)
    ctx.keep_original(True)
    ctx.generate(n, 1, 11, d)
    res = ctx.solve_gmres(B, prec, restart=20, max_iter=10)
    assert res["converged"] == 1 and res["backward_error"] <= 1e-14 * n
    x = ctx.download_solution64()
    A, _, b64, _ = O.generate_sized(n, 11, 1, d)
    F = O.gefa_nopvt_f64(A.astype(np.float64))
    x_ref = O.gesl_nopvt_f64(F, b64)
    assert np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)) <= 1e-10
    assert O.backward_error(A, x, b64) <= 1e-14 * n
    R = O.ref_lib()
    if R is not None:
        L, U = O.decode_factor(F)
        LU = np.ascontiguousarray((np.tril(L, -1) + U).T)
        xg = np.zeros(n)
        R.gmres_ref(n, np.ascontiguousarray(A.astype(np.float64)), n, xg, b64.copy(), LU, n, 20,
                    10, 1e-14)
        assert np.max(np.abs(x - xg)) / np.max(np.abs(xg)) <= 1e-10
    ctx.keep_original(False)


def test_gmres_needs_fewer_corrections_than_classic_with_fp8(ctx):
    """Reference generator (positive entries, coherent trailing updates) with
    an FP8 factor: both refinements reach 1e-14 n; GMRES takes at most as many
    outer corrections (matrix passes through the residual) as the classic one."""
    n, B = 2048, 256
    print("gmres vs classic fp8", flush=True)
    A, b, _, _ = O.generate_ref(n, seed=3)
    ctx.keep_original(True)
    ctx.upload(A)
    ctx.upload_rhs64(b.astype(np.float64))
    classic = ctx.solve_ir(B, "fp8", max_iter=60, check=False)
    gm = ctx.solve_gmres(B, "fp8", restart=20, max_iter=10, check=False)
    print(f"  classic: {classic['iterations']} it, berr {classic['backward_error']:.2e}; "
          f"gmres: {gm['iterations']} Arnoldi steps, berr {gm['backward_error']:.2e}", flush=True)
    assert gm["converged"] == 1 and gm["backward_error"] <= 1e-14 * n
    x = ctx.download_solution64()
    assert O.backward_error(A, x, b.astype(np.float64)) <= 1e-14 * n
    ctx.keep_original(False)
