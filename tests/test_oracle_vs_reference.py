"""The oracle against the reference compiled from its own sources
(oracle/_ref, built where the reference is present): bit-equality on blocks,
applies, seeded FTLM spectra, Jacobi and the thermodynamic accumulation over
models and blocks beyond the committed fixtures. Skipped where the harness
was not built."""
import numpy as np
import pytest

import oracle_lib

MODELS = [
    dict(lx=2, ly=2, tp=0.2, u=4.0, v=1.0),
    dict(lx=4, ly=2, u=8.0),
    dict(lx=3, ly=2, tx=0.9, ty=1.1, tp=-0.3, u=5.0, v=0.7, phix=0.25, phiy=0.6),
    dict(lx=3, ly=3, tp=-0.25, u=6.0, v=1.0, phix=0.75, phiy=0.25),
    dict(lx=1, ly=6, ty=0.8, u=3.0, v=0.2, phiy=0.4),
    dict(lx=5, ly=2, tp=0.15, u=4.5, v=0.3, phix=0.5, phiy=0.0),
]


@pytest.mark.parametrize("mi", range(len(MODELS)))
def test_blocks_and_spectra(ref, orc, mi):
    md = MODELS[mi]
    m = oracle_lib.model(**md)
    L = md["lx"] * md["ly"]
    rng = np.random.default_rng(mi)
    for _ in range(4):
        nup, ndn = (int(t) for t in rng.integers(0, L + 1, 2))
        mx, my = int(rng.integers(0, md["lx"])), int(rng.integers(0, md["ly"]))
        a, b = orc.block(m, nup, ndn, mx, my), ref.block(m, nup, ndn, mx, my)
        assert (a.dim, a.nnz, a.full_dim, a.parents) == (b.dim, b.nnz, b.full_dim, b.parents)
        if a.dim == 0:
            continue
        for x, y in zip(a.reps(), b.reps()):
            assert np.array_equal(x, y)
        for x, y in zip(a.csr(), b.csr()):
            assert np.array_equal(x, y)
        v = rng.standard_normal((a.dim, 3)) + 1j * rng.standard_normal((a.dim, 3))
        assert np.array_equal(a.apply(v), b.apply(v))
        bs = orc.block_seed(777, nup, ndn, mx, my)
        assert bs == ref.block_seed(777, nup, ndn, mx, my)
        if a.dim > 1:
            ea, wa = orc.ftlm_block(a, 3, 30, bs, threads=2)
            eb, wb = ref.ftlm_block(b, 3, 30, bs, threads=2)
            assert np.array_equal(ea, eb) and np.array_equal(wa, wb)


def test_jacobi_random(ref, orc):
    rng = np.random.default_rng(9)
    for n in (2, 9, 33):
        A = rng.standard_normal((n, n))
        A = A + A.T
        e1, v1 = ref.jacobi(A)
        import ctypes as C

        dp = C.POINTER(C.c_double)
        e2 = np.zeros(n)
        v2 = np.zeros(n)
        orc.lib.orc_jacobi(n, np.ascontiguousarray(A).copy().ctypes.data_as(dp), e2.ctypes.data_as(dp),
                           v2.ctypes.data_as(dp))
        assert np.array_equal(e1, e2) and np.array_equal(v1, v2)


def test_accumulate_random(ref, orc):
    rng = np.random.default_rng(4)
    mu = ref.linspace(-3.0, 4.0, 57)
    sr = dict(emin=np.full(57, np.inf), z=np.zeros(57), n1=np.zeros(57), n2=np.zeros(57))
    so = {k: v.copy() for k, v in sr.items()}
    for blk in range(20):
        n = int(rng.integers(1, 500))
        e = rng.standard_normal(n) * 5
        w = rng.random(n) if blk % 2 else None
        N = int(rng.integers(0, 10))
        ref.accumulate(e, w, 3.0, N, 2.857142857142857, mu, sr)
        orc.accumulate(e, w, 3.0, N, 2.857142857142857, mu, so)
    for k in sr:
        assert np.array_equal(sr[k], so[k])


def test_metadata_text(ref):
    m = oracle_lib.model(lx=4, ly=4, tp=-0.25, u=7.0, v=0.1, phix=0.3, phiy=0.7)
    txt = ref.metadata_text(m, 128, 80, 256, 12345)
    assert txt.startswith("lx=4\nly=4\ntx=1\nty=1\ntp=-0.25\n") and txt.endswith("seed=12345\n")
