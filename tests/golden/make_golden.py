"""Writes tests/golden/linpack_golden.npz from the oracle.

Fixtures: the reference generator at n=32, seed 0 (the configuration of the
reference README's emulator run, LINPACK/README.md:130-136,160-233) with its
gefa_ref_nopvt factor and gesl_ref_nopvt solution; the sized generator at n=64,
seed 42; and the README-derived ||b||inf = resid / (residn * n * eps).
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

A, b, _, normb = O.generate_ref(32, seed=0)
F = O.gefa_nopvt(A)
x = O.gesl_nopvt(F, b)
As, bs, _, _ = O.generate_sized(64, 42, 0)
readme = 5.960464477539062e-07 / (0.004354506590071576 * 32 * 1.1920928955078125e-07)
assert abs(readme - normb) < 1e-9 * normb
np.savez_compressed(os.path.join(HERE, "linpack_golden.npz"), ref32_A=A, ref32_b=b,
                    ref32_factor=F, ref32_x=x, sized64_A=As, sized64_b=bs,
                    readme_normb=np.float64(normb))
print("wrote linpack_golden.npz")
