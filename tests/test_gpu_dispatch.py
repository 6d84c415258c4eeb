"""GPU tests of the worker dispatch: the call sequences a SIAL program's
pardo / do loops issue, replayed through sipgpu_worker_* in the
interpreter's own terms (block selectors pushed in the compiler's order,
index ids of the SIAL index table, extents from the segment tables).

Known answers of the reference's test suite, exact (integer data):
  contraction_small_test  test_basic_sial.cpp:695-770 (contraction_small_test.sialx)
  transpose4d_tmp         test_basic_sial.cpp:1285-1327 (transpose4d_tmp.sialx,
                          test/test_transpose4d_op.F)
  contract_to_scalar      test_basic_sial.cpp:1037-1084 (contract_to_scalar.sialx)
The fills follow fill_block_cyclic.f / fill_block_sequential.f.  A ring
iteration through the worker (accumulating contractions, a permute of the
result and an energy-like contraction to a scalar that must see every queued
write) is checked against the oracle's sequence within 1e-12 (half of its
contractions run split over 13 moduli).
"""
import numpy as np
import pytest

import golden_io as g
from cases import rel_fro
from checker import Oracle

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu
o = Oracle()
I, J, K, L = 1, 2, 3, 4  # index ids (slots of the SIAL index table)


def _worker(budget=64 << 20):
    c = sg.Cache(budget)
    srv = sg.Server(c, max_batch=64)
    return c, srv, sg.Worker(c, srv)


def _close(c, srv, w):
    w.destroy()
    srv.destroy()
    c.destroy()


def test_contraction_small_test_replay():
    """do k / do j: execute fill_block_cyclic b[j,k] 1.0; do i / do l:
    allocate c[i,l]; execute fill_block_cyclic a[i,j,k,l] 1.0; print a;
    c[i,l] = a[i,j,k,l] * b[j,k]; print c.  One 15-segment each."""
    c, srv, w = _worker()
    A, B, C = 1 << 48, 2 << 48, 3 << 48  # block keys of a, b, c
    a = g.cyclic(15 ** 4, 1)
    b = g.cyclic(15 ** 2, 1)
    w.write(B, [15, 15], b)
    w.write(A, [15] * 4, a)
    assert np.array_equal(w.read(A, a.size), a)  # print a
    w.push(A, [I, J, K, L], [15] * 4)  # left operand pushed first
    w.push(B, [J, K], [15, 15])        # right operand on top
    w.contract(C, [I, L], [15, 15])
    assert w.depth() == 0  # all stacks empty
    got = w.read(C, 225)   # print c: waits for the queued contraction
    want = np.einsum("ijkl,jk->il", a.reshape((15,) * 4, order="F").astype(np.int64),
                     b.reshape((15, 15), order="F").astype(np.int64))
    assert np.array_equal(got, want.ravel(order="F").astype(np.float64))
    st = w.stats()
    assert st["contractions"] == 1 and st["writes"] == 2
    _close(c, srv, w)


def test_transpose4d_tmp_replay():
    """b[k,j,i,l] = a[i,j,k,l] with j, l over the occupied segment of
    extent 1 and i, k over the virtual segment of extent 5 (moaindex segments
    {1, 5, 4}); a filled by fill_block_sequential from 53.0.  The compiler's
    permutation vector is checked as the interpreter checks it."""
    c, srv, w = _worker()
    A, B = 1 << 48, 2 << 48
    a = g.sequential(25, 53.0)
    w.write(A, [5, 1, 5, 1], a)
    w.push(A, [I, J, K, L], [5, 1, 5, 1])  # right-hand side pushed first
    w.push(B, [K, J, I, L], [5, 1, 5, 1])  # left-hand side on top
    w.permute(compiler_perm=[2, 1, 0, 3, -1, -1])
    got = w.read(B, 25)
    Am = a.reshape((5, 1, 5, 1), order="F")
    Bm = np.empty((5, 1, 5, 1))
    for i in range(5):
        for k in range(5):
            Bm[k, 0, i, 0] = Am[i, 0, k, 0]  # test_transpose4d_op.F: b(k,j,i,l) = a(i,j,k,l)
    assert np.array_equal(got, Bm.ravel(order="F"))
    # a compiler vector that differs is refused, as the interpreter's check fails
    w.push(A, [I, J, K, L], [5, 1, 5, 1])
    w.push(B, [K, J, I, L], [5, 1, 5, 1])
    with pytest.raises(sg.SipGpuError) as ei:
        w.permute(compiler_perm=[0, 1, 2, 3, -1, -1])
    assert ei.value.code == -103
    assert w.depth() == 0
    _close(c, srv, w)


def test_contract_to_scalar_replay():
    """x = a[i,j] * b[i,j] with 8 x 8 blocks, a cyclic from 1.0 and b from
    5.0: the scalar the interpreter pushes onto its expression stack."""
    c, srv, w = _worker()
    A, B = 1 << 48, 2 << 48
    a, b = g.cyclic(64, 1), g.cyclic(64, 5)
    w.write(A, [8, 8], a)
    w.write(B, [8, 8], b)
    w.push(A, [I, J], [8, 8])
    w.push(B, [I, J], [8, 8])
    x = w.contract_to_scalar()
    assert x == float(np.dot(a.astype(np.int64), b.astype(np.int64)))
    _close(c, srv, w)


def test_ring_iteration_through_the_worker():
    """R(a,i,b,j) += W(k,b,c,j) T(a,i,c,k) over every block pair in SIP order
    (pushed and queued without waiting), then P[b,j,a,i] = R[a,i,b,j] for
    every block, then e = sum over blocks of P . P as contractions to a
    scalar, which must see every queued write.  Split precision on half of
    the contractions (13 moduli)."""
    w_ = wl.ring(o_segs=(5, 4), v_segs=(10, 7))
    D, Lb, Rb, ptrn = w_["D"], w_["L"], w_["R"], w_["ptrn"]
    A_, I_, B_, J_, K_, C_ = 1, 2, 3, 4, 5, 6  # ids of a, i, b, j, k, c
    keys = {}
    c, srv, w = _worker(256 << 20)
    for tag, arr in ((1, Lb), (2, Rb)):
        for blk in arr.block_ids():
            keys[(tag, blk)] = (tag << 48) | len(keys)
            w.write(keys[(tag, blk)], [int(x) for x in arr.dims(blk)], arr.block(blk))
    for blk in D.block_ids():
        keys[(3, blk)] = (3 << 48) | len(keys)
        keys[(4, blk)] = (4 << 48) | len(keys)
    want = {}
    for n, (db, lb, rb) in enumerate(w_["entries"]):
        w.push(keys[(1, lb)], [K_, B_, C_, J_], Lb.dims(lb))
        w.push(keys[(2, rb)], [A_, I_, C_, K_], Rb.dims(rb))
        w.contract(keys[(3, db)], [A_, I_, B_, J_], D.dims(db), accumulate=True,
                   moduli=13 if n % 2 else 0)
        tmp, e = o.contract(ptrn, Lb.block(lb), Lb.dims(lb), Rb.block(rb), Rb.dims(rb),
                            D.dims(db))
        assert e == 0
        want[db] = want.get(db, 0.0) + tmp
    energy = 0.0
    for db in D.block_ids():
        dd = [int(x) for x in D.dims(db)]
        w.push(keys[(3, db)], [A_, I_, B_, J_], dd)                  # rhs
        w.push(keys[(4, db)], [B_, J_, A_, I_], [dd[2], dd[3], dd[0], dd[1]])  # lhs
        w.permute()
        pw, _ = o.copy(dd, [1, 3, 4, 1, 2], want[db])
        energy += float(np.dot(pw, pw))
    got_e = 0.0
    for db in D.block_ids():
        dd = [int(x) for x in D.dims(db)]
        pd = [dd[2], dd[3], dd[0], dd[1]]
        w.push(keys[(4, db)], [B_, J_, A_, I_], pd)
        w.push(keys[(4, db)], [B_, J_, A_, I_], pd)
        got_e += w.contract_to_scalar()
    assert abs(got_e - energy) <= 1e-11 * abs(energy)
    w.flush(writeback=True)
    worst = max(rel_fro(w.read(keys[(3, db)], want[db].size), want[db])
                for db in D.block_ids())
    assert worst <= 1e-12
    st = w.stats()
    assert st["contractions"] == len(w_["entries"]) and st["queued"] == 0
    _close(c, srv, w)


def test_worker_errors():
    """An empty selector stack, a label pattern get_contraction_ptrn_
    rejects (its ierr in the message), an illegal transpose, and extents
    that do not fit the pattern."""
    c, srv, w = _worker()
    with pytest.raises(sg.SipGpuError) as ei:
        w.contract(9, [I, L], [3, 3])
    assert ei.value.code == -107
    w.write(1, [3, 4], np.ones(12))
    w.write(2, [4, 3], np.ones(12))
    w.push(1, [I, K], [3, 4])
    w.push(2, [J, L], [4, 3])  # no label shared, K and J appear once
    with pytest.raises(sg.SipGpuError) as ei:
        w.contract(9, [I, L], [3, 3])
    assert ei.value.code == -101 and "ierr" in str(ei.value)
    w.push(1, [I, K], [3, 4])
    w.push(3, [I, J], [3, 4])
    with pytest.raises(sg.SipGpuError) as ei:
        w.permute()
    assert ei.value.code == -103
    w.push(1, [I, K], [3, 4])
    w.push(2, [K, J], [5, 3])  # contracted extents differ
    with pytest.raises(sg.SipGpuError) as ei:
        w.contract(9, [I, J], [3, 3])
    assert ei.value.code == -101
    assert w.depth() == 0
    _close(c, srv, w)
