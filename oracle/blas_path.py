"""The reference's production contraction path on the host: permutes plus
a host BLAS dgemm('T','N').  Test infrastructure only (a CPU baseline and a
checker), never the product.

tensor_block_contract_ (src/sip/tensor_algebra/tensor_dil_omp.F90:662-906)
in a BLAS build: determine_index_permutations (:799-859) gives the O2N maps
lo2n / ro2n / do2n and the transpose flags; the operands are permuted
contracted-first with tensor_block_copy_ when their flag is set (:730-741);
the destination is zeroed (:752) and dgemm('T','N', lld, lrd, lcd, 1, ltp,
lcd, rtp, lcd, 1, dtp, lld) runs (:762); the matrix result is permuted into
the destination with do2n over the extents ddims(do2n(:)) (:782-785).

Here the maps and the permutes are the C restatement's (oracle.c, passed in
as an Oracle object) and the dgemm is numpy's, i.e. the OpenBLAS numpy
links; its thread count is set with threadpoolctl.  Summation order is
OpenBLAS's, so results agree with the oracle to rounding (1e-13 relative),
not bit for bit.
"""
import numpy as np


def contract(orc, ptrn, l, ldims, r, rdims, ddims):
    """D = L (x)_ptrn R for rank >= 1 operands and destination (the
    partial-contraction branch of tensor_block_contract_, :757-766)."""
    p = orc.perms(ptrn, ldims, rdims, ddims)
    M, N, K = int(p["M"]), int(p["N"]), int(p["K"])
    ltp = orc.copy(list(ldims), [1] + list(p["lo2n"]), l)[0] if p["ltransp"] else l
    rtp = orc.copy(list(rdims), [1] + list(p["ro2n"]), r)[0] if p["rtransp"] else r
    A = np.asarray(ltp).reshape((K, M), order="F")
    B = np.asarray(rtp).reshape((K, N), order="F")
    dtp = np.dot(A.T, B).ravel(order="F")  # dgemm('T','N') into a zeroed dtp
    if not p["dtransp"]:
        return dtp
    do2n = list(p["do2n"])
    src_dims = [int(ddims[k - 1]) for k in do2n]
    return orc.copy(src_dims, [1] + do2n, dtp)[0]


def blas_threads(n):
    """Context manager limiting the host BLAS to n threads."""
    from threadpoolctl import threadpool_limits
    return threadpool_limits(limits=int(n), user_api="blas")
