"""The 4x4 U=7 sweep's largest sector at full size, through properties that
do not depend on a CPU run (the oracle cannot build a 10.35M-row block in a
test's time):

* sector (8, 8): full dimension C(16,8)^2 = 165,636,900 and the k-block
  dimensions sum to it (the reference prints active_k_dim_sum == full_dim);
  block k = (0, 0) has 10,353,252 rows;
* H is Hermitian through the GPU apply: <x, H y> = <H x, y> for random x, y
  (relative 1e-12; a wrong sign, phase or index breaks it at O(1));
* the hop-list and on-the-fly applies agree bit for bit (both equal the
  reference's apply on every smaller block of the parity tests);
* FTLM on the block: Ritz values ascending per sample, weights summing to 1,
  and tree-order alpha/beta equal to the serial (reference-order) ones over
  the first 20 steps (before any Ritz value converges) within dim * eps, the
  worst-case rounding of the reference's own row-order sums at this size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def eng():
    import ftlm_extended_hubbard_development_amd as pkg

    e = pkg.Engine(lx=4, ly=4, u=7.0)
    yield e
    e.close()


def test_sector_dimensions(eng):
    full, parents = eng.sector(8, 8)
    assert full == 12870 ** 2 == 165636900
    dims = [eng.block(mx, my) for my in range(4) for mx in range(4)]
    say(f"(8,8): parents {parents}, k-block dims {dims}")
    assert sum(dims) == full
    assert eng.block(0, 0) == 10353252


def test_hermitian_and_kernels_agree(eng):
    eng.sector(8, 8)
    dim = eng.block(0, 0)
    rng = np.random.default_rng(11)
    xy = rng.standard_normal((dim, 2)) + 1j * rng.standard_normal((dim, 2))
    h = eng.apply(xy)  # hop-list path, both columns in one launch
    a = np.vdot(xy[:, 0], h[:, 1])
    b = np.vdot(h[:, 0], xy[:, 1])
    say(f"<x,Hy> = {a:.12e}, <Hx,y> = {b:.12e}")
    assert abs(a - b) <= 1e-12 * abs(a)
    eng.set_apply_config(2, 0)  # on the fly
    try:
        h2 = eng.apply(xy)
    finally:
        eng.set_apply_config(0, -1)
    assert np.array_equal(h, h2)


def test_ftlm_largest_block(eng):
    import ftlm_extended_hubbard_development_amd as pkg

    eng.sector(8, 8)
    eng.block(0, 0)
    R, steps = 2, 20
    bs = pkg.block_seed(12345, 8, 8, 0, 0)
    at, bt, mt = eng.lanczos(R, steps, bseed=bs, reduce=pkg.REDUCE_TREE)
    aser, bser, mser = eng.lanczos(R, steps, bseed=bs, reduce=pkg.REDUCE_SERIAL)
    assert (mt == steps).all() and (mser == steps).all()
    # the reference's row-order sum of dim terms is itself only accurate to
    # dim * eps relative (worst case): that, not 1e-12, bounds tree vs serial
    tol = 10353252 * 2.3e-16
    da = np.max(np.abs(at - aser) / np.abs(aser))
    db = np.max(np.abs(bt[:steps - 1] - bser[:steps - 1]) / np.abs(bser[:steps - 1]))
    say(f"tree vs serial over {steps} steps: alpha {da:.2e}, beta {db:.2e} (bound {tol:.1e})")
    assert da <= tol and db <= tol
    e, w = eng.ftlm_block(R, steps, bs, reduce=pkg.REDUCE_TREE)
    assert e.size == R * steps and abs(w.sum() - 1.0) < 1e-12
    for r in range(R):
        assert np.all(np.diff(e[r * steps:(r + 1) * steps]) >= 0)
    say(f"lowest Ritz values {e[0]:.10f} {e[steps]:.10f}, weight sum {w.sum():.15f}")
