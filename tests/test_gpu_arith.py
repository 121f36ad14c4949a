"""The short correctly rounded square root and reciprocal of the QEF's Givens rotation
(lvn_device.hpp: sqrt_ge1 for arguments >= 1, recip_1_2 for divisors in [1, 2)) against
the compiler's IEEE sqrtf and division, bit for bit, on the GPU: every float of [1, 4)
and of [1, 2) respectively, a stride-97 sample of every binade up to FLT_MAX, and +inf
(tests/cpp/arith_check.hip; needs a GPU).
"""
import ctypes as C
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libarith_check.so")


def test_short_sqrt_and_reciprocal_round_like_ieee():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    out = (C.c_ulonglong * 4)()
    assert lib.lvn_arith_check(out) == 0
    bad_sqrt, n_sqrt, bad_recip, n_recip = list(out)
    print(f"sqrt: {bad_sqrt} of {n_sqrt} differ; reciprocal: {bad_recip} of {n_recip} differ", flush=True)
    assert n_sqrt > 27_000_000 and n_recip == 1 << 23
    assert bad_sqrt == 0 and bad_recip == 0
