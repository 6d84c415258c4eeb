"""The reference's BLAS build restated on the host (oracle/blas_path.py:
the C restatement's permutes and O2N maps, numpy's dgemm): within 1e-13 of
the oracle on random contractions of rank 1..6 and on the ring and ladder
block shapes, with one and with several BLAS threads.  It is the CPU
baseline of the reference's production path."""
import os
import sys

import numpy as np
import pytest

from cases import dims_of, random_case, rel_fro
from checker import Oracle

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import blas_path  # noqa: E402

o = Oracle()


@pytest.mark.parametrize("threads", [1, 4])
def test_blas_path_random_vs_oracle(threads):
    rng = np.random.default_rng(17)
    with blas_path.blas_threads(threads):
        for _ in range(150):
            dl, ll, rl, ext = random_case(rng, max_ext=7)
            ptrn, e = o.get_contraction_ptrn(len(dl), len(ll), len(rl), dl + ll + rl)
            assert e == 0
            ld, rd, dd = dims_of(ll, ext), dims_of(rl, ext), dims_of(dl, ext)
            L = rng.standard_normal(int(np.prod(ld)))
            R = rng.standard_normal(int(np.prod(rd)))
            want, e = o.contract(ptrn, L, ld, R, rd, dd)
            got = blas_path.contract(o, ptrn, L, ld, R, rd, dd)
            assert rel_fro(got, want) <= 1e-13


def test_blas_path_ring_block_pair():
    # R(a,i,b,j) = W(k,b,c,j) T(a,i,c,k) at 30/15 segments (M = N = K = 450)
    ptrn, e = o.get_contraction_ptrn(4, 4, 4, [1, 2, 3, 4, 5, 3, 6, 4, 1, 2, 6, 5])
    ld, rd, dd = [15, 30, 30, 15], [30, 15, 30, 15], [30, 15, 30, 15]
    rng = np.random.default_rng(2)
    L = rng.uniform(-0.5, 0.5, int(np.prod(ld)))
    R = rng.standard_normal(int(np.prod(rd))) * 1e-2
    want, _ = o.contract(ptrn, L, ld, R, rd, dd)
    assert rel_fro(blas_path.contract(o, ptrn, L, ld, R, rd, dd), want) <= 1e-13
