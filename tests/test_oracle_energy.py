"""Oracle energy routines against closed forms (CPU).

energy_denominator (energy_denominator_rhf.F:399-444) and fcontract
(tensor_dil_omp.F90:910-938) are pinned by exact cases: small-integer and
dyadic energies make every denominator exact in any summation order, so the
index-to-sign mapping and the division are checked exactly.  The
left-to-right order of the sums is pinned by the reference's own loops:
tests/golden/reference_energy_den.npz holds do_rhf_den2/4/6 outputs on
non-dyadic energies (written by tests/golden/make_energy_golden.py from the
reference build), which the oracle must reproduce bit for bit while two
plausible other orders do not.
"""
import json
import os

import numpy as np
import pytest

from checker import Oracle

o = Oracle()
HERE = os.path.dirname(os.path.abspath(__file__))
DEN_GOLDEN = os.path.join(HERE, "golden", "reference_energy_den.npz")


def den_cases():
    z = np.load(DEN_GOLDEN)
    meta = json.loads(str(z["meta"]))
    out = []
    for n, c in enumerate(meta["cases"]):
        dims = [int(d) for d in z["c%d_dims" % n]]
        out.append((c["routine"], dims, [int(s) for s in z["c%d_signs" % n]],
                    [z["c%d_eps%d" % (n, d)] for d in range(len(dims))],
                    z["c%d_x" % n], z["c%d_out" % n]))
    return out


@pytest.mark.parametrize("case", den_cases(), ids=lambda c: "%s-%s" % (c[0], "x".join(map(str, c[1]))))
def test_denominator_matches_the_reference_loops(case):
    """Bit-exact against do_rhf_den2/4/6; for rank 4 and 6 the fixture tells
    the reference's order ((b + d + f) - a - c - e, left to right) from
    (b + d + f) - (a + c + e) and from index order (-a + b - c + d ...)."""
    routine, dims, signs, eps, x, want = case
    got = o.energy_denominator(dims, x.copy(), eps, signs)
    assert np.array_equal(got, want)
    if len(dims) < 4:
        return
    grids = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    e = [eps[d][grids[d]] for d in range(len(dims))]
    occ = sum(e[d] for d in range(len(dims)) if signs[d] > 0)
    vir = sum(e[d] for d in range(len(dims)) if signs[d] < 0)
    grouped = x / (occ - vir).ravel(order="F")
    inorder = np.zeros(dims)
    for d in range(len(dims)):
        inorder = inorder + signs[d] * e[d]
    inorder = x / inorder.ravel(order="F")
    assert not np.array_equal(grouped, want)
    assert not np.array_equal(inorder, want)


def test_denominator_rank6_exact():
    # (a,i,b,j,c,k) layout of do_rhf_den6: occupied at positions 2,4,6
    dims = [2, 3, 2, 2, 1, 2]
    signs = [-1, 1, -1, 1, -1, 1]
    eps = [np.array([0.5, 1.5]), np.array([-1.0, -2.0, -0.25]), np.array([2.0, 0.75]),
           np.array([-0.5, -1.5]), np.array([3.0]), np.array([-4.0, -0.125])]
    n = int(np.prod(dims))
    rng = np.random.default_rng(0)
    q = rng.integers(-9, 10, n).astype(np.float64)
    den = np.zeros(dims, order="F")
    for idx in np.ndindex(*dims):
        occ = sum(eps[d][idx[d]] for d in range(6) if signs[d] > 0)
        vir = sum(eps[d][idx[d]] for d in range(6) if signs[d] < 0)
        den[idx] = occ - vir
    x = q * den.ravel(order="F")  # exact: dyadic values, small magnitudes
    got = o.energy_denominator(dims, x, eps, signs)
    assert np.array_equal(got, q)


def test_denominator_rank2_and_4_signs():
    eps_v, eps_o = np.array([1.0, 2.0, 4.0]), np.array([-0.5, -8.0])
    x = np.ones(6)
    got = o.energy_denominator([3, 2], x, [eps_v, eps_o], [-1, 1])
    want = np.array([[1.0 / (eo - ev) for ev in eps_v] for eo in eps_o]).ravel()
    assert np.array_equal(got, want)
    got4 = o.energy_denominator([3, 2, 3, 2], np.ones(36), [eps_v, eps_o, eps_v, eps_o],
                                [-1, 1, -1, 1])
    assert got4[0] == 1.0 / (eps_o[0] + eps_o[0] - eps_v[0] - eps_v[0])
    assert got4[-1] == 1.0 / (eps_o[1] + eps_o[1] - eps_v[2] - eps_v[2])


def test_fcontract_closed_form():
    n = 1000
    k = np.arange(1, n + 1, dtype=np.float64)
    assert o.fcontract(k, k) == n * (n + 1) * (2 * n + 1) / 6
    assert o.fcontract(k, np.ones(n), d=5.0) == 5.0 + n * (n + 1) / 2
