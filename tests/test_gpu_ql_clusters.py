"""The GPU's two QL eigensolvers on spectra with clusters of eigenvalues at
zero (free fermions, U = V = 0, on the twisted 4x2 lattice with t' = 0.13).

There the couplings next to the zero eigenvalues stall at rounding noise
(~1e-17 against a matrix norm of ~30), so a split test relative to the two
neighbouring diagonals alone (both ~1e-15) never fires and the iteration
gives up; the solvers also split below eps ||T||. Checked against numpy's
eigvalsh of the same matrices:

* the batched exact solver (k_herm_eig) on the k = (0, 0) blocks of sectors
  (3, 5) and (4, 4), dims 392 and 628;
* the Ritz-value QL (k_ritz_ql, the tree-order FTLM path) on the real
  symmetric tridiagonal form of the first of them, with the weights (squared
  first eigenvector components) summing to one.
"""
import numpy as np
import pytest
import scipy.linalg

import oracle_lib

pytestmark = pytest.mark.gpu

MODEL = dict(lx=4, ly=2, tp=0.13, phix=0.3, phiy=0.2, u=0.0, v=0.0)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def eng():
    import ftlm_extended_hubbard_development_amd as pkg

    e = pkg.Engine(**MODEL)
    yield e
    e.close()


def dense(orc, nup, ndn):
    A = orc.block(oracle_lib.model(**MODEL), nup, ndn, 0, 0).dense()
    return 0.5 * (A + A.conj().T)


@pytest.mark.parametrize("nup,ndn", [(3, 5), (4, 4)])
def test_exact_blocks_zero_clusters(eng, orc, nup, ndn):
    eng.sector(nup, ndn)
    ev = eng.exact_blocks([(0, 0)])[0]
    ref = np.linalg.eigvalsh(dense(orc, nup, ndn))
    zeros = int((np.abs(ref) < 1e-10).sum())
    err = np.max(np.abs(ev - ref))
    say(f"exact ({nup},{ndn}) k=(0,0): dim {ev.size}, {zeros} zero eigenvalues, max err {err:.2e}")
    assert ev.size == ref.size and zeros >= 8
    assert err < 1e-12 * max(1.0, np.abs(ref).max())


def test_ritz_ql_zero_cluster(eng, orc):
    A = dense(orc, 3, 5)
    H = scipy.linalg.hessenberg(A)
    d = H.diagonal().real.copy()
    e = np.abs(H.diagonal(-1))  # a diagonal unitary makes the couplings real
    n = d.size
    alpha = d.reshape(n, 1)
    beta = np.append(e, 0.0).reshape(n, 1)
    eig, w = eng.ritz(alpha, beta, np.array([n], np.int32), 1, method=1)  # 1: QL (0: Jacobi)
    ref = np.linalg.eigvalsh(A)
    err = np.max(np.abs(eig[0] - ref))
    say(f"Ritz QL on the {n}-row tridiagonal form: max err {err:.2e}, weight sum {w[0].sum():.15f}")
    assert err < 1e-12 * max(1.0, np.abs(ref).max())
    assert abs(w[0].sum() - 1.0) < 1e-12
