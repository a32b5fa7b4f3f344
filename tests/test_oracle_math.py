"""Pins of the oracle by mathematics that does not depend on the reference:
brute-force Fock-space spectra, free fermions, the atomic limit, Burnside
block dimensions, Toeplitz tridiagonal closed forms, the mt19937_64 known
answer, and numpy eigvalsh. Each is chosen so that a plausible mistake (a
dropped term, a wrong sign or phase, a transposed operand) fails it.
"""
import itertools
import math

import numpy as np
import pytest

import oracle_lib


def sites_xy(lx, s):
    return s % lx, s // lx


def neighbours(lx, ly):
    out = []
    for s in range(lx * ly):
        x, y = sites_xy(lx, s)
        at = lambda a, b: ((b % ly) * lx + (a % lx))
        out.append([at(x + 1, y), at(x - 1, y), at(x, y + 1), at(x, y - 1),
                    at(x + 1, y + 1), at(x - 1, y + 1), at(x + 1, y - 1), at(x - 1, y - 1)])
    return out


def amplitudes(m):
    """-t_dir * exp(i theta_dir), directions as in the model definition."""
    ax = 2 * math.pi * m["phix"] / m["lx"]
    ay = 2 * math.pi * m["phiy"] / m["ly"]
    ang = [ax, -ax, ay, -ay, ax + ay, -ax + ay, ax - ay, -ax - ay]
    t = [m["tx"], m["tx"], m["ty"], m["ty"], m["tp"], m["tp"], m["tp"], m["tp"]]
    return [-t[d] * complex(math.cos(ang[d]), math.sin(ang[d])) for d in range(8)]


def full(m):
    d = dict(tx=1.0, ty=1.0, tp=0.0, phix=1.0, phiy=1.0, u=8.0, v=0.0)
    d.update(m)
    return d


def brute_sector_spectrum(m, nup, ndn):
    """Eigenvalues of H on the full (nup, ndn) configuration space, built
    from second quantisation directly: |up, dn> = prod up c^dag (ascending
    sites) prod dn c^dag (ascending sites) |0>, hopping c^dag_to c_from with
    amplitude -t e^{i theta} per direction, U doublons, V density bonds."""
    L = m["lx"] * m["ly"]
    nb = neighbours(m["lx"], m["ly"])
    amp = amplitudes(m)
    ups = [sum(1 << s for s in c) for c in itertools.combinations(range(L), nup)]
    dns = [sum(1 << s for s in c) for c in itertools.combinations(range(L), ndn)]
    states = [(u, d) for u in ups for d in dns]
    index = {s: i for i, s in enumerate(states)}
    H = np.zeros((len(states), len(states)), complex)
    bonds = set()
    for s in range(L):
        bonds.add((s, nb[s][0]))
        bonds.add((s, nb[s][2]))
    for i, (u, d) in enumerate(states):
        dbl = bin(u & d).count("1")
        nn = 0
        for a, b in bonds:
            na = ((u >> a) & 1) + ((d >> a) & 1)
            nbb = ((u >> b) & 1) + ((d >> b) & 1)
            nn += na * nbb
        H[i, i] += m["u"] * dbl + m["v"] * nn
        for spin in (0, 1):
            bits = u if spin == 0 else d
            for frm in range(L):
                if not (bits >> frm) & 1:
                    continue
                for dr in range(8):
                    to = nb[frm][dr]
                    if (bits >> to) & 1:
                        continue
                    lo, hi = min(frm, to), max(frm, to)
                    between = bin(bits & (((1 << max(hi - lo - 1, 0)) - 1) << (lo + 1))).count("1")
                    sign = -1.0 if between % 2 else 1.0
                    nbits = bits ^ (1 << frm) ^ (1 << to)
                    tgt = (nbits, d) if spin == 0 else (u, nbits)
                    # row = source, column = target (the block convention)
                    H[i, index[tgt]] += amp[dr] * sign
    assert np.allclose(H, H.conj().T, atol=1e-13)
    return np.linalg.eigvalsh(H)


def union_of_blocks(orc, m, nup, ndn, exact=True):
    ev = []
    dims = []
    for my in range(m["ly"]):
        for mx in range(m["lx"]):
            b = orc.block(oracle_lib.model(**m), nup, ndn, mx, my)
            dims.append(b.dim)
            if b.dim:
                ev.append(orc.exact_block(b))
    return np.sort(np.concatenate(ev)) if ev else np.zeros(0), dims


MODELS = [
    full(dict(lx=2, ly=2, tp=0.2, u=4.0, v=1.0)),
    full(dict(lx=3, ly=2, tp=-0.3, u=5.0, v=0.7, phix=0.25, phiy=0.6)),
    full(dict(lx=2, ly=3, tx=0.8, ty=1.3, tp=0.1, u=3.0, v=-0.4, phix=0.1, phiy=0.9)),
    full(dict(lx=4, ly=2, u=8.0)),
    full(dict(lx=1, ly=4, ty=1.1, u=2.0, v=0.5, phiy=0.3)),
]


@pytest.mark.parametrize("m", MODELS, ids=lambda m: f"{m['lx']}x{m['ly']}")
def test_blocks_reproduce_fock_space_spectrum(orc, m):
    """Union over momenta of the block spectra == brute-force sector
    spectrum (pins representatives, orbit signs, Bloch phases, norms and
    coefficients together)."""
    L = m["lx"] * m["ly"]
    for nup, ndn in [(1, 0), (1, 1), (2, 1), (L // 2, L // 2), (L - 1, 2)]:
        if nup > L or ndn > L:
            continue
        ev, dims = union_of_blocks(orc, m, nup, ndn)
        assert sum(dims) == math.comb(L, nup) * math.comb(L, ndn)
        ref = brute_sector_spectrum(m, nup, ndn)
        assert np.max(np.abs(ev - ref)) < 1e-10, (nup, ndn)


def test_free_fermions(orc):
    """U = V = 0: many-body energies are sums of single-particle energies of
    the twisted hopping matrix."""
    m = full(dict(lx=3, ly=2, tp=0.25, u=0.0, v=0.0, phix=0.3, phiy=0.2))
    L = 6
    nb = neighbours(3, 2)
    amp = amplitudes(m)
    h = np.zeros((L, L), complex)
    for frm in range(L):
        for d in range(8):
            h[frm, nb[frm][d]] += amp[d]
    eps = np.linalg.eigvalsh(h)
    for nup, ndn in [(1, 1), (2, 1), (3, 2)]:
        e_up = [sum(c) for c in itertools.combinations(eps, nup)]
        e_dn = [sum(c) for c in itertools.combinations(eps, ndn)]
        want = np.sort([a + b for a in e_up for b in e_dn])
        ev, _ = union_of_blocks(orc, m, nup, ndn)
        assert np.max(np.abs(ev - want)) < 1e-10


def test_atomic_limit(orc):
    """t = 0: the spectrum is the multiset of diagonal energies."""
    m = full(dict(lx=3, ly=3, tx=0.0, ty=0.0, tp=0.0, u=5.0, v=1.5))
    L = 9
    for nup, ndn in [(2, 3), (4, 4)]:
        ev, _ = union_of_blocks(orc, m, nup, ndn)
        want = []
        for cu in itertools.combinations(range(L), nup):
            for cd in itertools.combinations(range(L), ndn):
                u = sum(1 << s for s in cu)
                d = sum(1 << s for s in cd)
                want.append(orc.diag_energy(oracle_lib.model(**m), u, d))
        assert np.array_equal(ev, np.sort(want))


def burnside_dims(lx, ly, nup, ndn):
    """dim_k = (1/L) sum_T conj(chi_k(T)) Tr_up(T) Tr_dn(T), with
    Tr(T) = sum over T-invariant configurations of the translation sign."""
    L = lx * ly

    def shift(bits, dx, dy):
        moved = [((s // lx + dy) % ly) * lx + (s % lx + dx) % lx for s in range(L) if (bits >> s) & 1]
        inv = sum(1 for i in range(len(moved)) for j in range(i + 1, len(moved)) if moved[i] > moved[j])
        return sum(1 << s for s in moved), (-1) ** inv

    def trace(n, dx, dy):
        t = 0
        for c in itertools.combinations(range(L), n):
            b = sum(1 << s for s in c)
            b2, sg = shift(b, dx, dy)
            if b2 == b:
                t += sg
        return t

    dims = {}
    for my in range(ly):
        for mx in range(lx):
            acc = 0j
            for dy in range(ly):
                for dx in range(lx):
                    chi = np.exp(2j * math.pi * (mx * dx / lx + my * dy / ly))
                    acc += np.conj(chi) * trace(nup, dx, dy) * trace(ndn, dx, dy)
            dims[(mx, my)] = int(round((acc / L).real))
    return dims


@pytest.mark.parametrize("lx,ly,nup,ndn", [(2, 2, 1, 1), (4, 2, 3, 3), (3, 3, 2, 2), (4, 2, 4, 2), (2, 3, 3, 2),
                                           (4, 4, 2, 2)])
def test_burnside_block_dimensions(orc, lx, ly, nup, ndn):
    want = burnside_dims(lx, ly, nup, ndn)
    m = oracle_lib.model(lx=lx, ly=ly)
    for (mx, my), d in want.items():
        assert orc.block(m, nup, ndn, mx, my).dim == d


@pytest.mark.parametrize("m", MODELS[:3], ids=lambda m: f"{m['lx']}x{m['ly']}")
def test_blocks_hermitian(orc, m):
    L = m["lx"] * m["ly"]
    for my in range(m["ly"]):
        for mx in range(m["lx"]):
            b = orc.block(oracle_lib.model(**m), L // 2, L // 2 - 1, mx, my)
            if b.dim:
                H = b.dense()
                assert np.max(np.abs(H - H.conj().T)) < 1e-14


def test_full_krylov_recovers_spectrum(orc):
    """With m = dim Lanczos steps the Ritz values equal the exact spectrum
    of a small block (no degeneracies lost to the start vector's support)."""
    m = oracle_lib.model(lx=3, ly=2, tp=-0.3, u=5.0, v=0.7, phix=0.25, phiy=0.6)
    b = orc.block(m, 2, 1, 1, 0)
    exact = orc.exact_block(b)
    e, w = orc.ftlm_block(b, 1, b.dim, orc.block_seed(12345, 2, 1, 1, 0))
    # every Ritz value sits on an eigenvalue, and the extremal ones are found
    assert max(np.min(np.abs(exact - r)) for r in e) < 1e-8
    assert abs(e.min() - exact.min()) < 1e-10 and abs(e.max() - exact.max()) < 1e-10
    assert abs(w.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("n", [2, 5, 17, 40])
def test_toeplitz_tridiagonal(orc, n):
    """Jacobi Ritz values/weights of tridiag(b, a, b):
    a + 2b cos(j pi/(n+1)), weights (2/(n+1)) sin^2(j pi/(n+1))."""
    a, b = 0.7, -1.3
    e, w = orc.ritz(np.full(n, a), np.full(n - 1, b), 1)
    j = np.arange(1, n + 1)
    ev = a + 2 * b * np.cos(j * np.pi / (n + 1))
    wt = 2.0 / (n + 1) * np.sin(j * np.pi / (n + 1)) ** 2
    order = np.argsort(ev)
    assert np.max(np.abs(e - ev[order])) < 1e-12
    assert np.max(np.abs(w - wt[order])) < 1e-12


def test_weights_sum_per_sample(orc):
    b = orc.block(oracle_lib.model(lx=4, ly=2, u=8.0), 3, 3, 1, 0)
    R = 6
    e, w = orc.ftlm_block(b, R, 50, orc.block_seed(12345, 3, 3, 1, 0))
    for r in range(R):
        assert abs(w[r * 50:(r + 1) * 50].sum() - 1.0 / R) < 1e-14


def test_mt19937_64_known_answer(orc):
    """C++11 [rand.predef]: the 10000th output of a default-constructed
    mt19937_64 (seed 5489) is 9981545732273789042."""
    import ctypes as C

    f = orc.lib.orc_mt_nth
    f.restype = C.c_ulonglong
    f.argtypes = [C.c_ulonglong, C.c_longlong]
    assert f(5489, 10000) == 9981545732273789042


def test_start_vector_is_unit_gaussian(orc):
    x = orc.start_vector(orc.block_seed(12345, 4, 4, 0, 0), 3, 20000)
    assert abs(np.linalg.norm(x) - 1.0) < 1e-14
    z = x * np.sqrt(2 * 20000)  # each real component ~ N(0, 1/(2n)): rescaled to unit variance
    assert abs(np.mean(z.real)) < 0.05 and abs(np.std(z.real) - 1.0) < 0.05


@pytest.mark.parametrize("n", [1, 2, 3, 10, 64, 150])
def test_hermitian_eigensolver_vs_numpy(orc, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    assert np.max(np.abs(orc.herm_eigvals(A) - np.linalg.eigvalsh(A))) < 1e-12 * max(1, np.abs(A).max()) * n


def test_thermo_closed_form(orc):
    """Streaming log-sum-exp accumulation == direct evaluation of Z, <N>,
    <N^2> for a few blocks of known spectra (including a later block with a
    lower minimum, which exercises the rescale)."""
    mu = orc.linspace(-2.0, 3.0, 11)
    beta = 3.7
    blocks = [(np.array([0.5, 1.5, 2.0]), None, 1.0, 1), (np.array([-1.0, 0.2]), np.array([0.3, 0.7]), 4.0, 2),
              (np.array([-4.0, -3.5, 0.0]), None, 1.0, 3)]
    st = dict(emin=np.full(11, np.inf), z=np.zeros(11), n1=np.zeros(11), n2=np.zeros(11))
    for e, w, pref, N in blocks:
        orc.accumulate(e, w, pref, N, beta, mu, st)
    for k, m_ in enumerate(mu):
        Z = N1 = N2 = 0.0
        for e, w, pref, N in blocks:
            ww = np.ones_like(e) if w is None else w
            t = pref * ww * np.exp(-beta * (e - m_ * N))
            Z += t.sum()
            N1 += N * t.sum()
            N2 += N * N * t.sum()
        got_logz = math.log(st["z"][k]) - beta * st["emin"][k]
        assert abs(got_logz - math.log(Z)) < 1e-12
        assert abs(st["n1"][k] / st["z"][k] - N1 / Z) < 1e-12
        assert abs(st["n2"][k] / st["z"][k] - N2 / Z) < 1e-11


@pytest.mark.parametrize("nup,ndn", [(3, 5), (4, 4)])
def test_hermitian_eigensolver_zero_clusters(orc, nup, ndn):
    """Free fermions (U = V = 0) on the twisted 4x2 lattice: the k = (0, 0)
    blocks of these sectors (dims 392 and 628) hold clusters of eigenvalues
    at zero, where a QL split test relative to the neighbouring diagonals
    alone never fires (the couplings stall at rounding noise). The solver
    must converge and agree with numpy."""
    m = oracle_lib.model(lx=4, ly=2, tp=0.13, phix=0.3, phiy=0.2, u=0.0, v=0.0)
    blk = orc.block(m, nup, ndn, 0, 0)
    A = blk.dense()
    A = 0.5 * (A + A.conj().T)
    ref = np.linalg.eigvalsh(A)
    assert (np.abs(ref) < 1e-10).sum() >= 8
    w = orc.herm_eigvals(A)
    assert np.max(np.abs(w - ref)) < 1e-12 * max(1.0, np.abs(ref).max())
