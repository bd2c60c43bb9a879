"""Host-side model of every memory access batch_solve_kernel forms (CPU).

The batched kernel (hpcc_fpga_amd/csrc/batch.hip) works on one padded system
per workgroup with its strips in dynamic LDS.  This model replays, for every
bucket the host builds (api_batch.hip: padded orders up to 128 / 256 / 512 /
1024, maxm = the bucket's largest npad - 32, at least 32) and every padded
order in the bucket, the index arithmetic of each phase -- the diagonal
block D, the claim slot, the strip copies Left / TopT, the MFMA operand reads
of the trailing tiles with TPW tiles in flight per wave (the committed 4 and
the removed 8), the C tile loads and stores with the clamped tile index, and
the solve's FP32 vector x32 (over TopT), the FP64 vector xsh (over Left) and
the reduction slots red (over D) -- and checks each against its buffer:
LDS offsets below batch_lds_bytes(maxm) <= 160 KiB, regions that must not
overlap while both are live, and global offsets inside the system's
npad x npad matrix.  For both TPW values every access is in bounds.
"""
import numpy as np
import pytest

LDT = 40              # padded strip row (elements), batch.hip
D_BYTES = 32 * 33 * 4
LDS_LIMIT = 160 * 1024
HIS = (128, 256, 512, 1024)


def lds_bytes(maxm, f16):
    return D_BYTES + 2 * maxm * LDT * (2 if f16 else 4)


def acc_row(r, lane):
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)


def check_system(npad, maxm, f16, NT, TPW):
    es = 2 if f16 else 4
    total = lds_bytes(maxm, f16)
    left0, top0 = D_BYTES, D_BYTES + maxm * LDT * es
    lane = np.arange(64)
    lr, lh = lane & 31, lane >> 5
    # claim = D + 32 (one int): inside D, beyond the reduction slots red[0..NT/64)
    claim = 32 * 4
    assert claim + 4 <= D_BYTES and claim >= 8 * (NT // 64)
    # D[r * 33 + c], r, c < 32
    assert (31 * 33 + 31) * 4 + 4 <= D_BYTES
    for k0 in range(0, npad, 32):
        r0 = k0 + 32
        m = npad - r0
        if m <= 0:
            continue
        assert m <= maxm
        # strip copies: TopT[c * LDT + i], Left[r * LDT + i], c, r < m, i < 32
        assert top0 + ((m - 1) * LDT + 31) * es + es <= total
        assert left0 + ((m - 1) * LDT + 31) * es + es <= top0
        tiles = m // 32
        ntile = tiles * tiles
        waves = NT // 64
        for wave in range(waves):
            for tt0 in range(wave * TPW, ntile, waves * TPW):
                tt = np.minimum(tt0 + np.arange(TPW), ntile - 1)   # clamped tiles
                ti, tj = tt // tiles, tt % tiles
                # MFMA operand rows and k offsets
                row_l = ti[:, None] * 32 + lr[None, :]
                row_t = tj[:, None] * 32 + lr[None, :]
                assert row_l.max() < m and row_t.max() < m
                if f16:
                    kmax = 1 * 16 + 8 * lh.max() + 7         # ks < 2, 8 halves per read
                else:
                    kmax = 2 * 15 + lh.max()                 # kp < 16
                assert kmax < 32 < LDT
                assert left0 + (row_l.max() * LDT + kmax) * es + es <= top0
                assert top0 + (row_t.max() * LDT + kmax) * es + es <= total
                # C tiles in the system: rows/cols r0 + t*32 + (0..31)
                rows = r0 + ti[:, None, None] * 32 + acc_row(np.arange(16)[None, :, None],
                                                             lane[None, None, :])
                cols = r0 + tj[:, None] * 32 + lr[None, :]
                assert rows.min() >= 0 and rows.max() < npad
                assert cols.max() < npad
    # solve: x32 (npad floats) over TopT, xsh (npad doubles) over Left
    assert top0 + 4 * npad <= total
    assert left0 + 8 * npad <= top0
    # red[w] (doubles, w < NT/64) over D, below the claim slot
    assert 8 * (NT // 64) <= claim
    return total


def buckets(orders):
    out = []
    npads = [(int(n) + 31) // 32 * 32 for n in orders]
    for k, hi in enumerate(HIS):
        lo = HIS[k - 1] if k else 0
        mine = [p for p in npads if lo < p <= hi]
        if mine:
            out.append((max(32, max(mine) - 32), sorted(set(mine))))
    return out


def batch_orders(count, lo, hi, seed):
    from oracle import oracle as O

    return O.batch_orders(count, lo, hi, seed)


@pytest.mark.parametrize("TPW", [4, 8])
@pytest.mark.parametrize("NT", [128, 256, 512])
def test_every_access_in_bounds(TPW, NT):
    workloads = [
        np.array([33, 32, 100, 257, 500, 64, 1, 5, 992, 130]),  # test_gpu_batch mixed orders
        batch_orders(64, 32, 480, 3),                           # the autotune test
        batch_orders(256, 32, 1024, 7),
        batch_orders(4096, 32, 1024, 7),                        # the batched configuration
    ]
    for orders in workloads:
        for maxm, npads in buckets(orders):
            for f16 in (True, False):
                if lds_bytes(maxm, f16) > LDS_LIMIT:
                    continue  # refused on the host (FP32 strips of large orders)
                for npad in npads:
                    assert check_system(npad, maxm, f16, NT, TPW) <= LDS_LIMIT


def test_largest_bucket_fits_lds_exactly():
    # npad 1024: maxm 992 -> 162944 bytes of the 163840 a workgroup may hold
    assert lds_bytes(992, True) == 162944 <= LDS_LIMIT
    assert lds_bytes(992, False) > LDS_LIMIT
