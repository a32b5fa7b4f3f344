"""Device thermodynamic sums (ftlm_thermo) against the oracle's streaming
accumulation (accumulate_block_thermo, main.cpp:1172-1219, bit-identical to
the reference).

emin is the same minimum of E - mu N in both (floating-point subtraction is
monotone), so it must match exactly; z, n1, n2 differ only by summation
order and the reference's rescale products: 1e-12 relative. The finalised
n, compressibility and log Z are then within the engine's 1e-10 target.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def random_spectra(seed, nblk=60):
    rng = np.random.default_rng(seed)
    specs = []
    for b in range(nblk):
        particles = int(rng.integers(0, 19))
        if rng.random() < 0.4:  # exact block: eigenvalues only
            n = int(rng.integers(1, 40))
            e = np.sort(rng.uniform(-8, 30, n)) + 3.0 * particles
            specs.append((e, None, particles, 1.0))
        else:  # FTLM block: R samples of m Ritz values, weights summing to 1/R
            R, m = int(rng.integers(1, 6)), int(rng.integers(1, 25))
            dim = int(rng.integers(m, 5000))
            e = np.concatenate([np.sort(rng.uniform(-8, 30, m)) for _ in range(R)]) + 3.0 * particles
            w = rng.random(R * m)
            w = w.reshape(R, m) / w.reshape(R, m).sum(axis=1, keepdims=True) / R
            specs.append((e, w.ravel(), particles, float(dim)))
    return specs


@pytest.fixture(scope="module")
def eng():
    import ftlm_extended_hubbard_development_amd as pkg

    e = pkg.Engine(lx=3, ly=3)
    yield e
    e.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_thermo_matches_streaming_accumulation(eng, orc, seed):
    specs = random_spectra(seed)
    betas = [0.7, 4.0, 20.0]
    mu = np.linspace(-5.0, 15.0, 41)
    emin, z, n1, n2 = eng.thermo(specs, betas, mu)
    for i, beta in enumerate(betas):
        st = {k: np.zeros(mu.size) for k in ("z", "n1", "n2")}
        st["emin"] = np.full(mu.size, np.inf)
        for e, w, part, pref in specs:
            orc.accumulate(e, w, pref, part, beta, mu, st)
        assert np.array_equal(emin[i], st["emin"])
        for name, got in (("z", z[i]), ("n1", n1[i]), ("n2", n2[i])):
            err = np.max(np.abs(got - st[name]) / np.maximum(np.abs(st[name]), 1e-300))
            say(f"beta={beta} {name}: max rel err {err:.2e}")
            assert err <= 1e-12


def test_thermo_many_values(eng, orc):
    """More values than one workgroup holds, uneven tail."""
    rng = np.random.default_rng(5)
    specs = [(rng.uniform(-10, 10, 70001), np.full(70001, 1.0 / 70001), 7, 123.0),
             (rng.uniform(-12, 10, 3), None, 6, 1.0)]
    mu = np.array([-1.0, 0.5, 3.0])
    emin, z, n1, n2 = eng.thermo(specs, [2.0], mu)
    st = {k: np.zeros(mu.size) for k in ("z", "n1", "n2")}
    st["emin"] = np.full(mu.size, np.inf)
    for e, w, part, pref in specs:
        orc.accumulate(e, w, pref, part, 2.0, mu, st)
    assert np.array_equal(emin[0], st["emin"])
    assert np.allclose(z[0], st["z"], rtol=1e-12, atol=0)
    assert np.allclose(n2[0], st["n2"], rtol=1e-12, atol=0)
