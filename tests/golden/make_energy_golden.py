"""Writes tests/golden/reference_energy_den.npz from the reference's own
energy-denominator loops.

Run in the build container, where oracle/_ref/libenergyden_ref.so has been
compiled from the reference's energy_denominator_rhf.F (make -C oracle ref):

    python tests/golden/make_energy_golden.py

Each case calls one of do_rhf_den2 (energy_denominator_rhf.F:241-274),
do_rhf_den4 (:310-352) or do_rhf_den6 (:399-444) on a block of seeded normal
values with a Fock matrix whose diagonal holds non-dyadic orbital energies
(occupied in (-2, -0.3), virtual in (0.1, 3)) and whose off-diagonal
elements are nonzero, so the fixture pins which elements are read and the
left-to-right order of the sum.  Stored per case: the block extents, the
sign of each index (+1 occupied, -1 virtual), the orbital energies each
index reads, the input block and the reference's output (column-major).
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libenergyden_ref.so")
OUT = os.path.join(HERE, "reference_energy_den.npz")

NOCC, NBAS = 9, 30  # orbitals 1..9 occupied, 10..30 virtual


def fock_matrix(rng):
    f = rng.uniform(-0.05, 0.05, (NBAS, NBAS))
    f = 0.5 * (f + f.T)
    diag = np.concatenate([rng.uniform(-2.0, -0.3, NOCC), rng.uniform(0.1, 3.0, NBAS - NOCC)])
    np.fill_diagonal(f, diag)
    return np.asfortranarray(f)


def call(lib, rank, dims, offs, x, fock):
    """do_rhf_den<rank> on x (column-major, modified in place): index d runs
    1..dims[d] and reads the Fock diagonal at index + offs[d]."""
    i = lambda v: ctypes.byref(ctypes.c_int(int(v)))
    args = []
    for d in range(rank):
        args += [i(1), i(dims[d]), i(offs[d])]
    xp = x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    fp = fock.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    fn = {2: lib.do_rhf_den2_, 4: lib.do_rhf_den4_, 6: lib.do_rhf_den6_}[rank]
    fn(*args, xp, i(1), i(NBAS), i(1), i(NBAS), fp)


def main():
    # lazy binding: the file's driver routine calls into the SIP runtime,
    # which is absent here; only the do_rhf_den* loops are called
    lib = ctypes.CDLL(LIB, mode=os.RTLD_LAZY)
    rng = np.random.default_rng(20261203)
    fock = fock_matrix(rng)
    diag = np.diag(fock).copy()
    # (rank, extents, first orbital of each index); the reference's layouts:
    # den2 (a, i), den4 (a, i, b, j), den6 (a, i, b, j, c, k): even positions
    # virtual, odd positions occupied
    cases = [
        (2, [7, 4], [12, 3]),
        (4, [5, 3, 6, 4], [10, 1, 18, 5]),
        (4, [1, 9, 21, 1], [10, 1, 10, 9]),
        (6, [4, 3, 5, 2, 3, 4], [11, 2, 15, 7, 25, 1]),
        (6, [2, 5, 3, 4, 6, 2], [20, 1, 10, 3, 12, 8]),
    ]
    arrays, meta = {}, []
    for n, (rank, dims, first) in enumerate(cases):
        signs = [-1 if d % 2 == 0 else 1 for d in range(rank)]
        for d in range(rank):  # virtual indices read virtual orbitals, occupied occupied
            lo, hi = first[d], first[d] + dims[d] - 1
            assert (lo > NOCC) == (signs[d] < 0) and hi <= NBAS and (signs[d] < 0 or hi <= NOCC)
        offs = [f - 1 for f in first]
        x = rng.standard_normal(int(np.prod(dims)))
        y = x.copy()
        call(lib, rank, dims, offs, y, fock)
        arrays["c%d_dims" % n] = np.array(dims, dtype=np.int32)
        arrays["c%d_signs" % n] = np.array(signs, dtype=np.int32)
        for d in range(rank):
            arrays["c%d_eps%d" % (n, d)] = diag[first[d] - 1:first[d] - 1 + dims[d]]
        arrays["c%d_x" % n] = x
        arrays["c%d_out" % n] = y
        meta.append({"routine": "do_rhf_den%d" % rank, "dims": dims, "first_orbital": first})
    arrays["meta"] = np.array(json.dumps({
        "generator": "tests/golden/make_energy_golden.py",
        "reference": "src/sip/super_instructions/qm/qm-generic/energy_denominator_rhf.F",
        "build": "amdflang -O2 -ffp-contract=off (oracle/Makefile, target ref)",
        "numpy_seed": 20261203, "cases": meta}))
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, len(cases), "cases")


if __name__ == "__main__":
    main()
