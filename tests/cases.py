"""Random and fixed contraction cases shared by the tests.

A case is described the way the SIP interpreter describes it: integer index
labels for the destination, left and right blocks (the `aces_pattern` of
interpreter.cpp:1229-1236) plus one extent per label.
"""
import numpy as np


def random_case(rng, max_rank=6, max_ext=6, allow_scalar=False):
    """Draw (dlab, llab, rlab, ext) with ranks <= max_rank."""
    while True:
        nc = int(rng.integers(0, 4))
        nl = int(rng.integers(0, 4))
        nr = int(rng.integers(0, 4))
        lr, rr, dr = nl + nc, nr + nc, nl + nr
        if max(lr, rr, dr) > max_rank:
            continue
        if not allow_scalar and (lr == 0 or rr == 0 or dr == 0):
            continue
        break
    labels = list(rng.permutation(np.arange(1, 1 + nc + nl + nr)) + 10)
    c, l, r = labels[:nc], labels[nc:nc + nl], labels[nc + nl:]
    llab = list(rng.permutation(c + l)) if lr else []
    rlab = list(rng.permutation(c + r)) if rr else []
    dlab = list(rng.permutation(l + r)) if dr else []
    ext = {lab: int(rng.integers(1, max_ext + 1)) for lab in labels}
    return [int(x) for x in dlab], [int(x) for x in llab], [int(x) for x in rlab], ext


def dims_of(labels, ext):
    return [ext[x] for x in labels]


def brute_force(dlab, llab, rlab, ext, l, r):
    """Long-double einsum of the contraction (column-major blocks)."""
    letters = {}
    for lab in list(dlab) + list(llab) + list(rlab):
        letters.setdefault(lab, chr(ord("a") + len(letters)))
    L = np.asarray(l, dtype=np.longdouble).reshape(dims_of(llab, ext) or [1], order="F")
    R = np.asarray(r, dtype=np.longdouble).reshape(dims_of(rlab, ext) or [1], order="F")
    if not llab:
        L = L.reshape(())
    if not rlab:
        R = R.reshape(())
    spec = "%s,%s->%s" % ("".join(letters[x] for x in llab), "".join(letters[x] for x in rlab),
                          "".join(letters[x] for x in dlab))
    D = np.einsum(spec, L, R)
    return np.asarray(D).reshape(-1, order="F")


def rel_fro(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
