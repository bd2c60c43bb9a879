"""Batched solves of uneven order (run with -m gpu).

Bars: the device generator is bit-identical to the oracle's; every system's
refined solution reaches backward error <= 1e-14 * n and matches the FP64
oracle (gefa/gesl restated in FP64) to 1e-10 relative; each system's result
is bit-identical whether it is solved alone or in the mix, across graph
replays with other device work in between, and whether the factorisation runs
one workgroup per system or blocked (one launch per step over the bucket).
"""
import numpy as np
import pytest

from hpcc_fpga_amd import Context
from hpcc_fpga_amd.linpack import Batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _fp64_solution(A, b, n):
    F = O.gefa_nopvt_f64(A[:n, :n].astype(np.float64))
    return O.gesl_nopvt_f64(F, b[:n])


def test_batch_generator_and_solutions(ctx):
    orders = np.array([33, 32, 100, 257, 500, 64, 1, 5, 992, 130], np.int32)
    print("batch mixed orders", flush=True)
    bt = Batch(ctx, orders)
    bt.generate(7)
    for s in (0, 3, 8):
        a, b = bt.get_system(s)
        A_ref, b_ref, _ = O.generate_batch_system(int(orders[s]), 7, s)
        assert np.array_equal(a.view(np.uint32), A_ref.view(np.uint32))
        assert np.array_equal(b, b_ref)
    st = bt.solve("fp16", max_iter=20)
    assert st["failed"] == 0 and st["zero_pivots"] == 0
    for s, n in enumerate(orders):
        x, it, be = bt.get(s)
        assert be <= 1e-14 * n
        A_ref, b_ref, _ = O.generate_batch_system(int(n), 7, s)
        xr = _fp64_solution(A_ref, b_ref, int(n))
        assert np.max(np.abs(x - xr)) <= 1e-10 * max(1.0, np.max(np.abs(xr)))
    bt.close()


def test_alone_equals_in_mix_and_replays(ctx):
    orders = O.batch_orders(256, 32, 1024, 7)
    print("batch determinism", flush=True)
    mix = Batch(ctx, orders)
    mix.generate(7)
    st1 = mix.solve("fp16")
    assert st1["failed"] == 0
    picks = [0, 17, 100, 255]
    first = {s: mix.get(s)[0] for s in picks}
    # other device work between replays, then replay twice
    tmp = Context(0)
    tmp.generate(2048, 0, 1)
    tmp.gefa_dev(256)
    tmp.close()
    for _ in range(2):
        mix.solve("fp16")
        for s in picks:
            assert np.array_equal(mix.get(s)[0], first[s])
    # alone: a batch holding only system s, with the same matrix
    for s in picks:
        a, b = mix.get_system(s)
        n = int(orders[s])
        one = Batch(ctx, [n])
        one.set_system(0, a[:n, :n], b[:n])
        one.solve("fp16")
        assert np.array_equal(one.get(0)[0], first[s])
        one.close()
    mix.close()


def test_autotune_and_fp32(ctx):
    orders = O.batch_orders(64, 32, 480, 3)
    bt = Batch(ctx, orders)
    bt.generate(3)
    th = bt.autotune("fp16")
    # the workgroup size, negated for the blocked factorisation
    assert all(t in (0, 128, 256, 512, -128, -256, -512) for t in th)
    st = bt.solve("fp16")
    assert st["failed"] == 0
    st32 = bt.solve("fp32")
    assert st32["failed"] == 0 and st32["max_iterations"] <= 3
    bt.close()


@pytest.mark.parametrize("prec,hi", [("fp16", 1024), ("fp32", 480)])
def test_blocked_equals_persistent(ctx, prec, hi):
    """The blocked factorisation applies the persistent kernel's operations in
    the same order: every system's refined solution, iteration count and
    backward error are bit-identical between the two modes."""
    orders = np.concatenate([O.batch_orders(120, 32, hi, 5), [hi, 33, 1, 64]]).astype(np.int32)
    print(f"batch blocked vs persistent {prec} n<={hi}", flush=True)
    bt = Batch(ctx, orders)
    bt.generate(5)
    bt.set_blocked(False)
    st0 = bt.solve(prec)
    ref = [bt.get(s) for s in range(len(orders))]
    bt.set_blocked(True)
    st1 = bt.solve(prec)
    assert st0["failed"] == 0 and st1["failed"] == 0
    assert st1["zero_pivots"] == st0["zero_pivots"] == 0
    for s in range(len(orders)):
        x, it, be = bt.get(s)
        assert np.array_equal(x, ref[s][0]), s
        assert it == ref[s][1] and be == ref[s][2]
    # and again after a replay of the blocked graph
    bt.solve(prec)
    for s in (0, len(orders) - 1):
        assert np.array_equal(bt.get(s)[0], ref[s][0])
    bt.close()
