"""GPU parity of the FP64 path -- the reference built with DATA_TYPE=double
(run with -m gpu).

Oracle: gefa_ref_nopvt / gesl_ref_nopvt in double (oracle.gefa_nopvt_f64 /
gesl_nopvt_f64, LINPACK/src/host/linpack_data.cpp:166-229 with HOST_DATA_TYPE
double) on the DP build's own generator (oracle.generate_ref_f64).

Bars:
  * the first panel -- stored rows 0..63 and columns 0..63 -- is bit-identical
    to gefa_ref_nopvt: the diagonal block follows its operation order, and the
    strips follow the closed forms of top_update / left_update, which give
    the same operations on those elements (test_kernel_communication.cpp
    :624-632,776-789);
  * the whole factor within 1e-13 of max |F| (MFMA summation order);
  * the substitution with a given factor bit-identical to gesl_ref_nopvt (same
    order, contraction off);
  * the solution: the reference's check residn < 1 with the double epsilon,
    HPL scaled residual < 1, |x - 1| <= 1e-12.
"""
import numpy as np
import pytest

from hpcc_fpga_amd import Context, HplError
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float64).eps)


def _say(msg):
    print(msg, flush=True)


def test_block_bit_exact(ctx):
    _say("f64 single block n=64")
    A, b = O.generate_ref_f64(64, seed=0)
    F = A.copy()
    ctx.gefa_f64(F)
    assert np.array_equal(F.view(np.uint64), O.gefa_nopvt_f64(A).view(np.uint64))


@pytest.mark.parametrize("n", [128, 256, 640, 1024, 2048])
def test_factor_and_solution(ctx, n):
    _say(f"f64 factor n={n}")
    A, b = O.generate_ref_f64(n, seed=0)
    F = A.copy()
    ctx.gefa_f64(F)
    F_ref = O.gefa_nopvt_f64(A)
    assert np.array_equal(F[:64].view(np.uint64), F_ref[:64].view(np.uint64))
    assert np.array_equal(F[:, :64].view(np.uint64), F_ref[:, :64].view(np.uint64))
    err = np.max(np.abs(F - F_ref)) / np.max(np.abs(F_ref))
    assert err <= 1e-13, err
    # substitution with the GPU factor: bit-identical to gesl_ref_nopvt's order
    x = b.copy()
    ctx.gesl_f64(F, x)
    assert np.array_equal(x.view(np.uint64), O.gesl_nopvt_f64(F, b).view(np.uint64))
    # the reference's validation with the double epsilon
    normb = float(np.max(b))
    residn = np.max(np.abs(x - 1.0)) / (n * normb * EPS)
    assert residn < 1.0, residn
    assert np.max(np.abs(x - 1.0)) <= 1e-12
    M = A.T
    hpl = np.max(np.abs(M @ x - b)) / (EPS * np.max(np.abs(M).sum(axis=1)) * np.max(np.abs(x)) * n)
    assert hpl < 1.0, hpl


def test_device_resident_sized(ctx):
    n = 4096
    _say(f"f64 device-resident n={n}")
    c = Context(0)
    try:
        c.generate_f64(n, 0, 42)
        A32, _, b64, _ = O.generate_sized(n, 42, 0)
        for rep in range(2):
            if rep:
                c.restore_f64()
            c.gefa_dev_f64()
            c.gesl_dev_f64()
            rn, an, xn, bn = c.residual_f64()
            assert rn / (EPS * an * xn * n) < 1.0
        _, x = c.download_f64(matrix=False, solution=True)
        assert np.max(np.abs(x - 1.0)) <= 1e-12
        # the device system is the FP32 generator's, widened
        a, _ = c.download_f64(matrix=True, solution=False)
        c.restore_f64()
        a0, _ = c.download_f64(matrix=True, solution=False)
        assert np.array_equal(a0, A32.astype(np.float64))
        del a, b64
    finally:
        c.close()


def test_f64_errors(ctx):
    _say("f64 errors")
    with pytest.raises(HplError) as e:
        ctx.gefa_f64(np.zeros((96, 96)))  # n not a multiple of 64
    assert e.value.status == 1
    with pytest.raises(HplError) as e:
        ctx.gefa_f64(np.zeros((64, 64)))  # zero pivot
    assert e.value.status == 4
    c = Context(0)
    try:
        with pytest.raises(HplError) as e:
            c.gesl_dev_f64()
        assert e.value.status == 6
    finally:
        c.close()
