"""GPU parity of the engine (through the C ABI) against the CPU oracle.

The oracle is bit-identical to the reference on every case checked by
tests/test_oracle_vs_reference.py and the committed golden fixtures, so a
bit-exact match with it is a bit-exact match with the reference. Tolerances
are stated where the GPU legitimately rounds differently (tree reductions,
the Householder eigensolver that replaces LAPACK zheev).
"""
import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu

MODELS = {
    "2x2": dict(lx=2, ly=2, tp=0.2, u=4.0, v=1.0),
    "4x2": dict(lx=4, ly=2, u=8.0),
    "3x3tw": dict(lx=3, ly=3, tp=-0.25, u=6.0, v=1.0, phix=0.3, phiy=0.7),
    "4x4": dict(lx=4, ly=4, u=7.0),
    "3x4tw": dict(lx=3, ly=4, tp=-0.25, u=6.0, v=1.0, phix=0.25, phiy=0.75),
    "2x3": dict(lx=2, ly=3, tp=0.3, u=5.0, v=0.5, phix=0.25, phiy=0.5),
    "1x5": dict(lx=1, ly=5, tx=0.7, u=3.0, v=0.4, phiy=0.2),
}

BLOCKS = [
    ("2x2", 1, 1, 0, 0), ("2x2", 2, 1, 1, 0), ("2x2", 2, 2, 1, 1), ("2x2", 4, 3, 0, 1),
    ("4x2", 4, 4, 0, 0), ("4x2", 3, 4, 1, 1), ("4x2", 2, 2, 2, 0), ("4x2", 4, 4, 2, 1),
    ("3x3tw", 3, 3, 0, 0), ("3x3tw", 4, 3, 1, 2), ("3x3tw", 3, 0, 1, 1),
    ("4x4", 2, 2, 0, 0), ("4x4", 3, 2, 1, 3), ("4x4", 2, 2, 2, 2), ("4x4", 4, 4, 2, 0),
    ("3x4tw", 3, 3, 1, 2), ("2x3", 3, 2, 1, 1), ("1x5", 2, 2, 0, 3),
]


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pkg():
    import ftlm_extended_hubbard_development_amd as p

    p.load_library()
    return p


_engines = {}


def engine(pkg, name):
    if name not in _engines:
        _engines[name] = pkg.Engine(**MODELS[name])
    return _engines[name]


_oracle_blocks = {}


def both(pkg, orc, name, nup, ndn, mx, my):
    say(f"case {name} ({nup},{ndn},{mx},{my})")
    eng = engine(pkg, name)
    if (eng.nup, eng.ndn) != (nup, ndn):
        eng.sector(nup, ndn)
    dim = eng.block(mx, my)
    key = (name, nup, ndn, mx, my)
    if key not in _oracle_blocks:
        _oracle_blocks[key] = orc.block(oracle_lib.model(**MODELS[name]), nup, ndn, mx, my)
    blk = _oracle_blocks[key]
    assert dim == blk.dim
    return eng, blk


@pytest.mark.parametrize("case", BLOCKS, ids=lambda c: "-".join(map(str, c)))
def test_block_basis_bitexact(pkg, orc, case):
    eng, blk = both(pkg, orc, *case)
    if blk.dim == 0:
        return
    up, dn, deg, dnf = eng.reps()
    oup, odn, odeg, odnf = blk.reps()
    assert np.array_equal(up, oup) and np.array_equal(dn, odn)
    assert np.array_equal(deg, odeg)
    assert np.array_equal(dnf.view(np.uint64), odnf.view(np.uint64))


@pytest.mark.parametrize("case", BLOCKS, ids=lambda c: "-".join(map(str, c)))
def test_block_csr_bitexact(pkg, orc, case):
    eng, blk = both(pkg, orc, *case)
    if blk.dim == 0:
        return
    rp, col, val, diag = eng.csr(all_dirs=True)
    orp, ocol, oval, odiag = blk.csr()
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(val, oval)  # numeric equality (+0 == -0)
    assert np.array_equal(diag, odiag)


@pytest.mark.parametrize("R", [1, 3, 5, 32, 64, 128, 130])
@pytest.mark.parametrize("case", [("4x2", 4, 4, 0, 0), ("3x3tw", 4, 3, 1, 2), ("4x4", 3, 2, 1, 3)],
                         ids=lambda c: "-".join(map(str, c)))
def test_apply_bitexact(pkg, orc, case, R):
    eng, blk = both(pkg, orc, *case)
    rng = np.random.default_rng(R)
    x = rng.standard_normal((blk.dim, R)) + 1j * rng.standard_normal((blk.dim, R))
    y = eng.apply(x, kernel=0)
    yo = blk.apply(x)
    assert np.array_equal(y, yo)
    if R == 1:
        assert np.array_equal(eng.apply(x, kernel=1), yo)


@pytest.mark.parametrize("path,W", [(1, 0), (1, 8), (1, 16), (1, 64), (2, 0), (2, 8), (2, 16), (2, 32)])
@pytest.mark.parametrize("R", [1, 5, 33, 64, 128, 130])
@pytest.mark.parametrize("case", [("4x2", 4, 4, 0, 0), ("3x3tw", 4, 3, 1, 2), ("4x4", 3, 2, 1, 3),
                                  ("4x4", 4, 4, 2, 0)], ids=lambda c: "-".join(map(str, c)))
def test_apply_configs_bitexact(pkg, orc, case, R, path, W):
    """Hop-list and on-the-fly kernels, untiled and column-tiled: the same
    bits as the reference's apply for every configuration."""
    eng, blk = both(pkg, orc, *case)
    rng = np.random.default_rng(R + W)
    x = rng.standard_normal((blk.dim, R)) + 1j * rng.standard_normal((blk.dim, R))
    eng.set_apply_config(path, W)
    try:
        assert np.array_equal(eng.apply(x), blk.apply(x))
    finally:
        eng.set_apply_config(0, -1)


def test_autotune_and_configured_lanczos(pkg, orc):
    """The autotuner picks a candidate; tree-order Lanczos through every
    apply configuration matches the oracle before ghost onset."""
    case = ("4x4", 3, 2, 1, 3)
    eng, blk = both(pkg, orc, *case)
    bs = pkg.block_seed(12345, *case[1:])
    R = 32
    best, ms, cands = eng.autotune(R)
    assert ms > 0 and len(cands) >= 3
    try:
        for path, W in ((1, 0), (1, 8), (2, 0), (2, 8)):
            eng.set_apply_config(path, W)
            a, b, m = eng.lanczos(R, 40, bseed=bs, reduce=pkg.REDUCE_TREE)
            for r in range(0, R, 7):
                oa, ob = orc.lanczos(blk, 40, bs, r)
                assert np.allclose(a[:30, r], oa[:30], rtol=1e-12, atol=1e-12)
                assert np.allclose(b[:30, r], ob[:30], rtol=1e-12, atol=0)
    finally:
        eng.set_apply_config(0, -1)


@pytest.mark.parametrize("path", [1, 2])
def test_serial_bitexact_both_paths(pkg, orc, path):
    case = ("4x4", 3, 2, 1, 3)
    eng, blk = both(pkg, orc, *case)
    bs = pkg.block_seed(12345, *case[1:])
    eng.set_apply_config(path, -1)
    try:
        for pv in (False, True):
            e, w = eng.ftlm_block(4, 60, bs, reduce=pkg.REDUCE_SERIAL, per_vector=pv)
            oe, ow = orc.ftlm_block(blk, 4, 60, bs)
            assert np.array_equal(e, oe) and np.array_equal(w, ow)
    finally:
        eng.set_apply_config(0, -1)


@pytest.mark.parametrize("case", [("4x2", 4, 4, 0, 0), ("3x3tw", 4, 3, 1, 2), ("4x4", 3, 2, 1, 3),
                                  ("2x2", 2, 2, 1, 1)], ids=lambda c: "-".join(map(str, c)))
def test_start_vectors_bitexact(pkg, orc, case):
    eng, blk = both(pkg, orc, *case)
    bs = pkg.block_seed(12345, *case[1:])
    for s in (0, 1, 7):
        assert np.array_equal(eng.start_vector(bs, s), orc.start_vector(bs, s, blk.dim))


@pytest.mark.parametrize("per_vector", [False, True])
@pytest.mark.parametrize("case,R,steps", [(("4x2", 4, 4, 0, 0), 8, 120), (("3x3tw", 4, 3, 1, 2), 5, 80),
                                          (("4x4", 3, 2, 1, 3), 3, 80), (("2x2", 2, 2, 1, 1), 3, 40)],
                         ids=lambda c: "-".join(map(str, c)) if isinstance(c, tuple) else str(c))
def test_lanczos_serial_bitexact(pkg, orc, case, R, steps, per_vector):
    """Serial reductions reproduce the reference's alpha/beta bit for bit."""
    eng, blk = both(pkg, orc, *case)
    bs = pkg.block_seed(12345, *case[1:])
    a, b, m = eng.lanczos(R, steps, bseed=bs, reduce=pkg.REDUCE_SERIAL, per_vector=per_vector)
    for r in range(R):
        oa, ob = orc.lanczos(blk, steps, bs, r)
        assert m[r] == len(oa)
        assert np.array_equal(a[: m[r], r], oa)
        assert np.array_equal(b[: m[r] - 1, r], ob)


@pytest.mark.parametrize("case,R,steps", [(("4x2", 4, 4, 0, 0), 8, 120), (("3x3tw", 4, 3, 1, 2), 5, 80),
                                          (("4x4", 3, 2, 1, 3), 3, 80)],
                         ids=lambda c: "-".join(map(str, c)) if isinstance(c, tuple) else str(c))
def test_lanczos_tree_matches_before_ghosts(pkg, orc, case, R, steps):
    """Tree reductions round differently; before loss of orthogonality sets
    in (the first 30 steps) alpha/beta agree to 1e-12 relative."""
    eng, blk = both(pkg, orc, *case)
    bs = pkg.block_seed(12345, *case[1:])
    a, b, m = eng.lanczos(R, steps, bseed=bs, reduce=pkg.REDUCE_TREE)
    for r in range(R):
        oa, ob = orc.lanczos(blk, steps, bs, r)
        k = min(30, len(ob))
        assert np.allclose(a[:k, r], oa[:k], rtol=1e-12, atol=1e-12)
        assert np.allclose(b[:k, r], ob[:k], rtol=1e-12, atol=0)


@pytest.mark.parametrize("n", [1, 2, 5, 40, 80, 120])
def test_ritz_jacobi_bitexact(pkg, orc, n):
    """Device Jacobi == the reference's jacobi_diagonalize + sort (oracle)."""
    eng = engine(pkg, "2x2")
    rng = np.random.default_rng(n)
    R = 3
    al = rng.standard_normal((n, R)) * 3.0
    be = np.abs(rng.standard_normal((n, R))) + 0.1
    m = np.full(R, n, np.int32)
    e, w = eng.ritz(al, be, m, 7, method=0)
    for r in range(R):
        oe, ow = orc.ritz(al[:, r], be[: n - 1, r], 7)
        assert np.array_equal(e[r, :n], oe)
        assert np.array_equal(w[r, :n], ow)


@pytest.mark.parametrize("n", [1, 2, 5, 40, 80, 120])
def test_ritz_ql_vs_numpy(pkg, n):
    """QL Ritz pairs (tree-order runs) vs numpy eigh of the tridiagonal:
    eigenvalues to 1e-13 * ||T||, weights summed over each cluster of equal
    eigenvalues to 1e-12 (individual weights of degenerate pairs are not
    unique)."""
    eng = engine(pkg, "2x2")
    rng = np.random.default_rng(100 + n)
    R = 4
    al = rng.standard_normal((n, R)) * 3.0
    be = np.abs(rng.standard_normal((n, R))) + 0.1
    m = np.full(R, n, np.int32)
    e, w = eng.ritz(al, be, m, 1, method=1)
    for r in range(R):
        T = np.diag(al[:, r]) + np.diag(be[: n - 1, r], 1) + np.diag(be[: n - 1, r], -1)
        ev, V = np.linalg.eigh(T)
        scale = np.abs(ev).max()
        assert np.max(np.abs(e[r, :n] - ev)) < 1e-13 * scale
        assert np.max(np.abs(w[r, :n] - V[0] ** 2)) < 1e-12
        assert abs(w[r, :n].sum() - 1.0) < 1e-13


@pytest.mark.parametrize("case,R,steps", [(("4x2", 4, 4, 0, 0), 32, 120), (("3x3tw", 4, 3, 1, 2), 8, 80),
                                          (("4x4", 3, 2, 1, 3), 4, 80)],
                         ids=lambda c: "-".join(map(str, c)) if isinstance(c, tuple) else str(c))
def test_ftlm_block_serial_bitexact(pkg, orc, case, R, steps):
    eng, blk = both(pkg, orc, *case)
    bs = pkg.block_seed(12345, *case[1:])
    e, w = eng.ftlm_block(R, steps, bs, reduce=pkg.REDUCE_SERIAL)
    oe, ow = orc.ftlm_block(blk, R, steps, bs)
    assert np.array_equal(e, oe) and np.array_equal(w, ow)


def _block_n(eig, w, pref, N, beta, mu):
    sh = eig - mu * N
    m = sh.min()
    z = np.sum(pref * w * np.exp(-beta * (sh - m)))
    return z, m


@pytest.mark.parametrize("case,R,steps,beta", [(("4x2", 4, 4, 0, 0), 32, 120, 20.0),
                                               (("3x3tw", 4, 3, 1, 2), 8, 80, 10.0),
                                               (("4x4", 3, 2, 1, 3), 8, 80, 12.5)],
                         ids=lambda c: "-".join(map(str, c)) if isinstance(c, tuple) else str(c))
def test_ftlm_block_tree_thermo(pkg, orc, case, R, steps, beta):
    """Tree reductions: the block's partition-function contribution agrees
    with the reference's to 1e-10 relative over a mu grid (alpha/beta
    themselves diverge after ghost states appear)."""
    eng, blk = both(pkg, orc, *case)
    bs = pkg.block_seed(12345, *case[1:])
    e, w = eng.ftlm_block(R, steps, bs, reduce=pkg.REDUCE_TREE)
    oe, ow = orc.ftlm_block(blk, R, steps, bs)
    assert abs(w.sum() - 1.0) < 1e-12
    N = case[1] + case[2]
    for mu in np.linspace(-3, 6, 10):
        z, m = _block_n(e, w, blk.dim, N, beta, mu)
        zo, mo = _block_n(oe, ow, blk.dim, N, beta, mu)
        assert abs(m - mo) < 1e-10
        assert abs(z - zo) <= 1e-10 * abs(zo)


@pytest.mark.parametrize("name,nup,ndn", [("2x2", 2, 1), ("4x2", 2, 3), ("4x2", 4, 4), ("3x3tw", 2, 1),
                                          ("2x3", 3, 3), ("1x5", 2, 2)])
def test_exact_blocks(pkg, orc, name, nup, ndn):
    """Batched device eigensolver vs numpy eigvalsh of the reference's
    Hermitised dense block (the reference calls zheev)."""
    eng = engine(pkg, name)
    g = MODELS[name]
    eng.sector(nup, ndn)
    ks = [(mx, my) for my in range(g["ly"]) for mx in range(g["lx"])]
    ks = [k for k in ks if eng.block(*k) > 0]
    say(f"exact {name} ({nup},{ndn}) blocks={len(ks)}")
    eigs = eng.exact_blocks(ks)
    for (mx, my), ev in zip(ks, eigs):
        blk = orc.block(oracle_lib.model(**g), nup, ndn, mx, my)
        H = blk.dense()
        Hh = 0.5 * (H + H.conj().T)
        ref = np.linalg.eigvalsh(Hh)
        scale = max(1.0, np.abs(ref).max())
        assert len(ev) == blk.dim
        assert np.max(np.abs(ev - ref)) < 1e-12 * scale
        if blk.dim == 1:
            assert ev[0] == orc.exact_block(blk)[0]


@pytest.mark.parametrize("R", [257, 300, 512])
@pytest.mark.parametrize("reduce", ["serial", "tree"])
def test_ftlm_block_beyond_one_batch(pkg, orc, R, reduce):
    """More samples than one batched launch holds (256 columns): the block
    runs in batches of samples and gives the reference's spectrum, bit for
    bit with serial reductions, with weights divided by the full R."""
    case = ("4x2", 3, 4, 1, 1)
    eng, blk = both(pkg, orc, *case)
    bs = pkg.block_seed(12345, *case[1:])
    e, w = eng.ftlm_block(R, 40, bs, reduce=pkg.REDUCE_SERIAL if reduce == "serial" else pkg.REDUCE_TREE)
    oe, ow = orc.ftlm_block(blk, R, 40, bs)
    assert len(e) == len(oe)
    if reduce == "serial":
        assert np.array_equal(e, oe) and np.array_equal(w, ow)
    else:
        assert abs(w.sum() - 1.0) < 1e-12
        for mu in np.linspace(-3, 10, 7):
            z, m = _block_n(e, w, blk.dim, 7, 20.0, mu)
            zo, mo = _block_n(oe, ow, blk.dim, 7, 20.0, mu)
            assert abs(m - mo) < 1e-10 and abs(z - zo) <= 1e-10 * abs(zo)


def test_batched_calls_reject_more_than_256_columns(pkg, orc):
    """The single-launch entry points hold at most 256 columns and say so
    (FTLM_ERR_LIMIT) instead of leaving columns unwritten; tile widths are
    bounded per apply path."""
    eng, blk = both(pkg, orc, "4x2", 3, 4, 1, 1)
    x = np.zeros((blk.dim, 257), np.complex128)
    with pytest.raises(pkg.FtlmError, match="error -5"):
        eng.apply(x)
    with pytest.raises(pkg.FtlmError, match="error -5"):
        eng.lanczos(257, 10, bseed=1)
    with pytest.raises(pkg.FtlmError, match="error -5"):
        eng.prepare(300, 1)
    with pytest.raises(pkg.FtlmError, match="error -5"):
        eng.set_apply_config(2, 256)
    with pytest.raises(pkg.FtlmError, match="error -5"):
        eng.set_apply_config(1, 512)
    eng.set_apply_config(1, 256)
    eng.set_apply_config(0, -1)
    # per-vector runs take any R up to 1024
    a, b, m = eng.lanczos(257, 5, bseed=1, reduce=pkg.REDUCE_SERIAL, per_vector=True)
    assert np.all(m == 5)
