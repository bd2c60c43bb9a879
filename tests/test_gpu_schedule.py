"""GPU: the pipelined GEFA schedule gives the grouped schedule's factor.

Both schedules apply every step's update to each element exactly once, in
step order, through the same GEMM kernels; they only split the steps into
launches differently (windows of whole steps, so the K splits fall on block
boundaries and every MFMA sees the same operand pairs).  Each element is
therefore formed by the same floating-point operations in the same order,
and the factors must agree bit for bit.  The grouped schedule is the one the
earlier parity tests pinned against the oracle; the cases below cover group
counts that do not divide the number of blocks, more slices than steps per
group, the 16-bit update and the FP64 path.

The schedule, update depth, slice count and tail are read when a context is
created (HPL_SCHEDULE, HPL_UPDATE_DEPTH, HPL_BULK_SLICES, HPL_F64_DEPTH,
HPL_TAIL_M, HPL_TAIL_DEPTH).
"""
import os

import numpy as np
import pytest

from hpcc_fpga_amd import Context

pytestmark = pytest.mark.gpu

KEYS = ("HPL_SCHEDULE", "HPL_UPDATE_DEPTH", "HPL_BULK_SLICES", "HPL_F64_DEPTH", "HPL_TAIL_M",
        "HPL_TAIL_DEPTH", "HPL_HEAD_M", "HPL_HEAD_DEPTH")


def _context(env):
    saved = {k: os.environ.get(k) for k in KEYS}
    for k in KEYS:
        os.environ.pop(k, None)
    os.environ.update(env)
    try:
        return Context(0)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _factor(env, n, B, prec="fp32", seed=5):
    print(f"schedule {env} n={n} B={B} {prec}", flush=True)
    c = _context(env)
    try:
        c.generate(n, 0, seed)
        c.gefa_dev(B, prec)
        a, _ = c.download(matrix=True, solution=False)
        return a
    finally:
        c.close()


@pytest.mark.parametrize("n,B,depth,slices,tail", [
    (2048, 128, 3, 0, None),     # 16 blocks: groups of 3, the last one short
    (1792, 256, 3, 0, None),     # 7 blocks
    (2560, 128, 4, 3, None),     # fewer slices than steps per group
    (1024, 64, 2, 5, None),      # more slices than steps per group
    (4096, 256, 5, 0, None),
    (4096, 256, 4, 0, None),       # the default depth, slices and tail
    (4096, 256, 3, 0, (0, 0)),     # no tail
    (4096, 256, 3, 0, (1536, 1)),  # groups of 3, then single steps
    (3072, 128, 4, 0, (1280, 2)),  # a group straddles the tail threshold
])
def test_pipelined_equals_grouped_fp32(n, B, depth, slices, tail):
    g = _factor({"HPL_SCHEDULE": "grouped", "HPL_UPDATE_DEPTH": str(depth)}, n, B)
    env = {"HPL_SCHEDULE": "pipelined", "HPL_UPDATE_DEPTH": str(depth),
           "HPL_BULK_SLICES": str(slices)}
    # None: the default tail (single steps once the order left is at most n/2)
    if tail:
        env.update({"HPL_TAIL_M": str(tail[0]), "HPL_TAIL_DEPTH": str(tail[1])})
    p = _factor(env, n, B)
    assert np.array_equal(g.view(np.uint32), p.view(np.uint32))


@pytest.mark.parametrize("prec,n,depth,extra", [
    ("bf16", 3072, 5, {}),  # the default head (groups of 6 above 5n/8) and tail
    ("fp8", 3072, 5, {}),
    # two head groups of 6 blocks, then the tail's groups of 2: 16 blocks
    ("bf16", 4096, 4, {"HPL_HEAD_M": "2048", "HPL_HEAD_DEPTH": "6"}),
])
def test_pipelined_equals_grouped_low_precision(prec, n, depth, extra):
    B = 256
    g = _factor({"HPL_SCHEDULE": "grouped", "HPL_UPDATE_DEPTH": str(depth)}, n, B, prec)
    env = {"HPL_SCHEDULE": "pipelined", "HPL_UPDATE_DEPTH": str(depth)}
    env.update(extra)
    p = _factor(env, n, B, prec)
    assert np.array_equal(g.view(np.uint32), p.view(np.uint32))


def test_pipelined_equals_grouped_fp64():
    n = 1600  # 25 blocks of 64: groups of 6, the last one short
    out = []
    for env in ({"HPL_SCHEDULE": "grouped", "HPL_F64_DEPTH": "6"},
                {"HPL_SCHEDULE": "pipelined", "HPL_F64_DEPTH": "6"},
                {"HPL_SCHEDULE": "pipelined", "HPL_F64_DEPTH": "6", "HPL_TAIL_M": "640",
                 "HPL_TAIL_DEPTH": "2"},
                {"HPL_SCHEDULE": "pipelined"}):  # the default depth and tail
        print(f"schedule {env} FP64 n={n}", flush=True)
        c = _context(env)
        try:
            c.generate_f64(n, 0, 3)
            c.gefa_dev_f64()
            a, _ = c.download_f64(matrix=True, solution=False)
            out.append(a)
        finally:
            c.close()
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))
    assert np.array_equal(out[0].view(np.uint64), out[2].view(np.uint64))
    assert np.array_equal(out[0].view(np.uint64), out[3].view(np.uint64))


def test_kernel_timing_counts_overlap():
    """hpl_kernel_stats / hpl_kernel_busy: the union of the update launches'
    intervals is positive and at most their summed durations (equal when no
    two launches overlap, which at this size depends on timing)."""
    n, B = 4096, 256
    c = _context({"HPL_SCHEDULE": "pipelined"})
    try:
        c.generate(n, 0, 7)
        c.set_profiling(True)
        c.gefa_dev(B)
        launches, summed, flops = c.kernel_stats()
        busy = c.kernel_busy()
        c.set_profiling(False)
    finally:
        c.close()
    print(f"launches {launches} summed {summed:.6f} s busy {busy:.6f} s", flush=True)
    assert launches > 0 and flops > 0
    assert 0.0 < busy <= summed * (1 + 1e-5) + 1e-6
