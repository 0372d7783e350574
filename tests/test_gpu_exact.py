"""Gray-code exact density of states on the GPU (wl_exact_dos) against the
oracle: every count equal to the transfer-matrix counts (and to brute force
where N <= 26), the total 2^N, the symmetry g(E) = g(-E) of bipartite
lattices, and the closed forms of the ferromagnet's two ground states and
first excited level."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import oracle_disorder
from oracle_lib import oracle, p
from qec_pf_wang_landau_amd import engine

pytestmark = pytest.mark.gpu


def oracle_counts(J, X, Y, boundary):
    N = X * Y
    out = []
    for s in range(J.size // (2 * N)):
        lng = np.zeros(4 * N + 1)
        cnt = np.zeros(4 * N + 1, np.uint64)
        rc = oracle().wlo_exact_dos_transfer(p(np.ascontiguousarray(J[s * 2 * N:(s + 1) * 2 * N])), X, Y, boundary,
                                             p(lng), p(cnt))
        assert rc == 0
        out.append(cnt)
    return np.array(out)


@pytest.mark.parametrize("X,Y,boundary,prob", [(3, 3, 0, 0.2), (4, 4, 0, 0.1), (4, 4, 1, 0.1), (5, 5, 0, 0.15),
                                               (3, 5, 1, 0.3), (6, 4, 1, 0.1), (5, 6, 0, 0.1), (6, 6, 0, 0.04)])
def test_counts_equal_transfer_matrix(cuda, X, Y, boundary, prob):
    S = 3
    J = oracle_disorder(41, S, X, Y, prob, "c")
    got = engine.exact_counts_gpu(J, X, Y, boundary)
    ref = oracle_counts(J, X, Y, boundary)
    assert np.array_equal(got, ref)
    assert (got.sum(axis=1) == np.uint64(1) << np.uint64(X * Y)).all()


def test_counts_equal_brute_force(cuda):
    X, Y = 4, 4
    J = oracle_disorder(3, 2, X, Y, 0.2, "v")
    got = engine.exact_counts_gpu(J, X, Y, 0)
    for s in range(2):
        cnt = np.zeros(4 * X * Y + 1, np.uint64)
        assert oracle().wlo_exact_dos_bruteforce(p(np.ascontiguousarray(J[s * 32:(s + 1) * 32])), X, Y, 0, p(cnt)) == 0
        assert np.array_equal(got[s], cnt)


def test_ferromagnet_closed_forms(cuda):
    """J = +1 everywhere, 6x6 periodic: g(-2N) = 2, g(-2N + 8) = 2N (one
    flipped spin breaks four bonds), g symmetric under E -> -E."""
    X = Y = 6
    N = X * Y
    J = np.ones(2 * N, np.int8)
    g = engine.exact_counts_gpu(J, X, Y, 0)[0]
    assert g[0] == 2 and g[8] == 2 * N and g[4] == 0
    assert np.array_equal(g, g[::-1])
    assert int(g.sum()) == 2 ** N


def test_refuses_unsupported_shapes(cuda):
    with pytest.raises(Exception, match="periodic lattices need"):
        engine.exact_counts_gpu(np.ones(2 * 2 * 8, np.int8), 2, 8, 0)
    with pytest.raises(Exception, match="X Y <= 40"):
        engine.exact_counts_gpu(np.ones(2 * 7 * 7, np.int8), 7, 7, 0)
