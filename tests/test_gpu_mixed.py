"""Mixed-precision trailing update + FP64 iterative refinement (run with -m gpu).

Bars:
  * BF16/FP16 factor vs the oracle's FP32 gefa_ref_nopvt factor: relative
    max difference <= 2e-2 (BF16, unit roundoff 2^-8) / 4e-3 (FP16, 2^-11) --
    a wrong MFMA operand layout gives O(1) differences;
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


@pytest.mark.parametrize("prec,d", [("bf16", 2.0), ("bf16", 1.1), ("fp16", 1.1)])
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


def test_refinement_fp32_factor(ctx):
    """FP32 trailing update through the same refinement loop."""
    n, B = 1024, 128
    ctx.keep_original(True)
    ctx.generate(n, 1, 3, 1.5)
    res = ctx.solve_ir(B, "fp32", max_iter=10)
    assert res["converged"] == 1 and res["iterations"] <= 3
    ctx.keep_original(False)
