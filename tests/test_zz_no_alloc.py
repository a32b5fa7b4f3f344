"""The allocation-free entry point ftlm_thermo_noalloc against ftlm_thermo.

Device inputs and outputs are torch tensors on the engine's GPU; the entry
point's scratch is the library's fixed per-GPU array. Two calls in a row,
the second with a smaller input, must each give ftlm_thermo's sums for the
same spectra bit for bit (the same kernels and grid; the running minima are
exact on both paths); a call beyond the scratch size is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spectra(rng, nblk):
    out = []
    for _ in range(nblk):
        n = int(rng.integers(1, 200))
        e = np.sort(rng.standard_normal(n) * 4.0 - 10.0)
        if rng.random() < 0.3:  # an exact block
            out.append((e, None, int(rng.integers(0, 33)), 1.0))
        else:
            w = rng.random(n)
            w /= w.sum()
            out.append((e, w, int(rng.integers(0, 33)), float(rng.integers(300, 10 ** 6))))
    return out


def per_value(spec):
    e = np.concatenate([s[0] for s in spec])
    w = np.concatenate([(np.ones(len(s[0])) if s[1] is None else s[1]) * s[3] for s in spec])
    n = np.concatenate([np.full(len(s[0]), s[2], np.uint8) for s in spec])
    return e, w, n


def test_thermo_noalloc_matches_thermo():
    import torch

    import ftlm_extended_hubbard_development_amd as pkg

    eng = pkg.Engine(lx=4, ly=4, u=7.0)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(21)
    betas = np.array([2.857142857142857, 12.5])
    mu = np.linspace(-3.0, 4.0, 281)
    d_betas = torch.tensor(betas, device=dev)
    d_mu = torch.tensor(mu, device=dev)
    for nblk in (900, 37):  # second call smaller: stale scratch must not leak into it
        print(f"case thermo_noalloc nblk={nblk}", flush=True)
        spec = spectra(rng, nblk)
        emin, z, n1, n2 = eng.thermo(spec, betas, mu)
        e, w, n = per_value(spec)
        d_e = torch.tensor(e, device=dev)
        d_w = torch.tensor(w, device=dev)
        d_n = torch.tensor(n, device=dev)
        d_out = torch.full((len(betas), len(mu), 4), float("nan"), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        rc = eng.lib.ftlm_thermo_noalloc(eng.h, len(e), d_e.data_ptr(), d_w.data_ptr(), d_n.data_ptr(), len(betas),
                                         d_betas.data_ptr(), len(mu), d_mu.data_ptr(), d_out.data_ptr(), None)
        assert rc == 0, eng.lib.ftlm_last_error()
        out = d_out.cpu().numpy()
        assert np.array_equal(out[..., 0], emin)
        assert np.array_equal(out[..., 1], z) and np.array_equal(out[..., 2], n1) and np.array_equal(out[..., 3], n2)
    # beyond the fixed scratch: refused with FTLM_ERR_LIMIT, nothing launched
    big = 4096 * 3000  # 3000 * 562 * 3 partials > 4 Mi doubles
    rc = eng.lib.ftlm_thermo_noalloc(eng.h, big, d_e.data_ptr(), d_w.data_ptr(), d_n.data_ptr(), len(betas),
                                     d_betas.data_ptr(), len(mu), d_mu.data_ptr(), d_out.data_ptr(), None)
    assert rc == -5 and b"32 MiB" in eng.lib.ftlm_last_error()
    eng.close()
