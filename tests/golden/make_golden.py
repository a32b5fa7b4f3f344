"""Generate the golden fixtures under tests/golden/ from the reference.

Runs only where the reference harness oracle/_ref/libftlm_ref.so was built
(from /root/reference/src/main.cpp by oracle/Makefile). Every number written
here comes from the reference's own functions (build_lattice,
build_sector_basis, build_momentum_block, apply_hamiltonian, run_ftlm_block,
jacobi_diagonalize, accumulate_block_thermo, write_results, the checkpoint
writer) except the eigenvalues of exactly diagonalised blocks: the reference
calls LAPACK zheev('N', 'U') there (main.cpp:888-910), which the harness does
not link, so this script calls the same LAPACK routine itself -- zheev of
Intel MKL (the LAPACK the reference was built against when it was measured,
loaded from /opt/conda/lib/libmkl_rt.so through ctypes, sequential layer) --
on the reference's own Hermitised dense block (main.cpp:862-886).

    python tests/golden/make_golden.py            # writes the fixtures
    python tests/golden/make_golden.py zheev csv  # only the exact-block ones
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib  # noqa: E402

ref = oracle_lib.reference()
if ref is None:
    sys.exit("reference harness not built (make -C oracle with /root/reference present)")


def hermitised(blk):
    """main.cpp:875-886: the diagonal's imaginary part dropped, (i, j) and
    conj(j, i) averaged (0.5 * (a + conj(b)) rounds like numpy's expression)."""
    H = blk.dense()
    A = 0.5 * (H + H.conj().T)
    A[np.diag_indices_from(A)] = H.diagonal().real
    return A


ZHEEV_SOURCE = "LAPACK zheev('N','U') of Intel MKL (libmkl_rt, sequential)"
_mkl = None


def zheev(A):
    """Eigenvalues (ascending) by LAPACK zheev with jobz='N', uplo='U', the
    call of main.cpp:888-910 (workspace query, then the solve)."""
    global _mkl
    import ctypes as C

    if _mkl is None:
        os.environ["MKL_THREADING_LAYER"] = "SEQUENTIAL"
        _mkl = C.CDLL("/opt/conda/lib/libmkl_rt.so")
    n = A.shape[0]
    a = np.asfortranarray(A.astype(np.complex128))
    w = np.zeros(n)
    rwork = np.zeros(max(1, 3 * n - 2))
    info = C.c_int(0)
    jobz, uplo, nn = C.c_char(b"N"), C.c_char(b"U"), C.c_int(n)
    dp = C.POINTER(C.c_double)
    q = np.zeros(2)
    lwork = C.c_int(-1)
    _mkl.zheev_(C.byref(jobz), C.byref(uplo), C.byref(nn), a.ctypes.data_as(dp), C.byref(nn), w.ctypes.data_as(dp),
                q.ctypes.data_as(dp), C.byref(lwork), rwork.ctypes.data_as(dp), C.byref(info))
    assert info.value == 0
    lw = max(1, int(q[0]))
    work = np.zeros(2 * lw)
    lwork = C.c_int(lw)
    _mkl.zheev_(C.byref(jobz), C.byref(uplo), C.byref(nn), a.ctypes.data_as(dp), C.byref(nn), w.ctypes.data_as(dp),
                work.ctypes.data_as(dp), C.byref(lwork), rwork.ctypes.data_as(dp), C.byref(info))
    assert info.value == 0
    return w


def exact_eigs(blk):
    if blk.dim == 1:
        rp, col, val, diag = blk.csr()
        ev = diag[0]
        for e in range(rp[0], rp[1]):
            if col[e] == 0:
                ev += val[e].real
        return np.array([ev])
    return zheev(hermitised(blk))


def sweep(model, *, beta_list, mu_min, mu_max, mu_count, samples, steps, threshold=256, seed=12345):
    """Block loop of main.cpp:1519-1728 on reference functions; returns the
    per-beta (z, n1, n2, emin) arrays and the mu grid."""
    m = oracle_lib.model(**model)
    L = model["lx"] * model["ly"]
    mu = ref.linspace(mu_min, mu_max, mu_count)
    stored = []
    for nup in range(L + 1):
        for ndn in range(L + 1):
            for my in range(model["ly"]):
                for mx in range(model["lx"]):
                    blk = ref.block(m, nup, ndn, mx, my)
                    if blk.dim == 0:
                        continue
                    if blk.dim <= threshold:
                        stored.append((exact_eigs(blk), None, 1.0, nup + ndn))
                    else:
                        bs = ref.block_seed(seed, nup, ndn, mx, my)
                        e, w = ref.ftlm_block(blk, samples, steps, bs)
                        stored.append((e, w, float(blk.dim), nup + ndn))
    out = []
    for beta in beta_list:
        st = dict(emin=np.full(mu_count, np.inf), z=np.zeros(mu_count), n1=np.zeros(mu_count),
                  n2=np.zeros(mu_count))
        for e, w, pref, N in stored:
            ref.accumulate(e, w, pref, N, beta, mu, st)
        out.append(st)
    return mu, out, L


def write_csv(path, betas, mu, states, L):
    """finalisation of main.cpp:1770-1788, written by the reference's
    write_results / write_multi_beta_results."""
    cols = {k: [] for k in ("n", "cc", "kappa", "z", "logz")}
    for beta, st in zip(betas, states):
        z, n1, n2, emin = st["z"], st["n1"], st["n2"], st["emin"]
        logz = np.array([math.log(a) for a in z]) - beta * emin
        mean = n1 / z
        mean2 = n2 / z
        cc = (mean2 - mean * mean) / float(L)
        cols["n"].append(mean / float(L))
        cols["cc"].append(np.maximum(0.0, cc))
        cols["kappa"].append(beta * np.maximum(0.0, cc))
        cols["z"].append(z)
        cols["logz"].append(logz)
    import ctypes as C

    dp = C.POINTER(C.c_double)
    arr = {k: np.ascontiguousarray(np.concatenate(v)) for k, v in cols.items()}
    b = np.ascontiguousarray(betas, np.float64)
    mu = np.ascontiguousarray(mu)
    rc = ref.lib.ref_write_results(path.encode(), len(betas), b.ctypes.data_as(dp), len(mu), mu.ctypes.data_as(dp),
                                   arr["n"].ctypes.data_as(dp), arr["cc"].ctypes.data_as(dp),
                                   arr["kappa"].ctypes.data_as(dp), arr["z"].ctypes.data_as(dp),
                                   arr["logz"].ctypes.data_as(dp))
    assert rc == 0


CSV_CASES = {
    # name: (model, sweep kwargs)
    "cfg1_2x2": (dict(lx=2, ly=2, tx=1.0, ty=1.0, tp=0.2, phix=1.0, phiy=1.0, u=4.0, v=1.0),
                 dict(beta_list=[1.5], mu_min=-2.0, mu_max=6.0, mu_count=33, samples=5, steps=40)),
    "cfg1_2x2_ftlm_only": (dict(lx=2, ly=2, tx=1.0, ty=1.0, tp=0.2, phix=1.0, phiy=1.0, u=4.0, v=1.0),
                           dict(beta_list=[1.5], mu_min=-2.0, mu_max=6.0, mu_count=33, samples=5, steps=40,
                                threshold=0)),
    "cfg2_4x2": (dict(lx=4, ly=2, u=8.0, v=0.0, tp=0.0),
                 dict(beta_list=[20.0], mu_min=-5.0, mu_max=15.0, mu_count=81, samples=32, steps=120)),
    "4x2_ftlm_only_r4_s40": (dict(lx=4, ly=2, u=8.0),
                             dict(beta_list=[20.0], mu_min=-5.0, mu_max=15.0, mu_count=81, samples=4, steps=40,
                                  threshold=0)),
    "4x2_multibeta_r8_s60": (dict(lx=4, ly=2, u=8.0),
                             dict(beta_list=[20.0, 5.0], mu_min=-5.0, mu_max=15.0, mu_count=41, samples=8,
                                  steps=60)),
    "3x3_twist_r4_s60": (dict(lx=3, ly=3, tp=-0.25, u=6.0, v=1.0, phix=0.25, phiy=0.5),
                         dict(beta_list=[10.0], mu_min=-5.0, mu_max=15.0, mu_count=41, samples=4, steps=60)),
    # every block exact: the ED result that bounds the cfg2 FTLM error
    "ed_4x2_beta20": (dict(lx=4, ly=2, u=8.0, v=0.0, tp=0.0),
                      dict(beta_list=[20.0], mu_min=-5.0, mu_max=15.0, mu_count=81, samples=32, steps=120,
                           threshold=1000000000)),
    # the 3x4 model of the twist-averaged scan (t' = -0.25, U = 6, V = 1,
    # beta = 10) at two twists of its 4x4 grid, few samples
    "cfg4_3x4_tw_0.25_0.75_r4_s40": (dict(lx=3, ly=4, tp=-0.25, u=6.0, v=1.0, phix=0.25, phiy=0.75),
                                     dict(beta_list=[10.0], mu_min=-2.0, mu_max=8.0, mu_count=21, samples=4,
                                          steps=40)),
    "cfg4_3x4_tw_0.5_0_r4_s40": (dict(lx=3, ly=4, tp=-0.25, u=6.0, v=1.0, phix=0.5, phiy=0.0),
                                 dict(beta_list=[10.0], mu_min=-2.0, mu_max=8.0, mu_count=21, samples=4,
                                      steps=40)),
    # every block exact on the twisted 3x3 lattice (blocks up to 1,764 rows)
    "ed_3x3_twist_beta10": (dict(lx=3, ly=3, tp=-0.25, u=6.0, v=1.0, phix=0.25, phiy=0.5),
                            dict(beta_list=[10.0], mu_min=-5.0, mu_max=15.0, mu_count=41, samples=4, steps=60,
                                 threshold=1000000000)),
}


def command_line(model, kw):
    """The ftlm_n_vs_mu flags reproducing one CSV case."""
    a = []
    for k in ("lx", "ly", "tx", "ty", "tp", "phix", "phiy", "u", "v"):
        if k in model:
            a += [f"--{k}", repr(model[k])]
    if len(kw["beta_list"]) > 1:
        a += ["--beta-list", ",".join(repr(b) for b in kw["beta_list"])]
    else:
        a += ["--beta", repr(kw["beta_list"][0])]
    a += ["--mu-min", repr(kw["mu_min"]), "--mu-max", repr(kw["mu_max"]), "--mu-count", str(kw["mu_count"]),
          "--samples", str(kw["samples"]), "--lanczos-steps", str(kw["steps"])]
    if "threshold" in kw:
        a += ["--exact-block-threshold", str(kw["threshold"])]
    return a


def csv_fixtures(only=None):
    os.makedirs(os.path.join(HERE, "csv"), exist_ok=True)
    meta = {}
    cases_path = os.path.join(HERE, "csv", "cases.json")
    if only and os.path.exists(cases_path):
        with open(cases_path) as f:
            meta = json.load(f)
    for name, (model, kw) in CSV_CASES.items():
        if only and name not in only:
            continue
        print("csv", name, flush=True)
        mu, states, L = sweep(model, **kw)
        write_csv(os.path.join(HERE, "csv", name + ".csv"), kw["beta_list"], mu, states, L)
        meta[name] = {"args": command_line(model, kw),
                      "exact_blocks_from": ZHEEV_SOURCE if kw.get("threshold", 256) > 0 else "none"}
    with open(os.path.join(HERE, "csv", "cases.json"), "w") as f:
        json.dump(meta, f, indent=1)


BLOCK_CASES = [
    (dict(lx=2, ly=2, tp=0.2, u=4.0, v=1.0), [(1, 1, 0, 0), (2, 1, 1, 0), (2, 2, 1, 1), (4, 3, 0, 1)]),
    (dict(lx=4, ly=2, u=8.0), [(4, 4, 0, 0), (3, 4, 1, 1), (2, 3, 2, 0), (4, 4, 2, 1)]),
    (dict(lx=3, ly=3, tp=-0.25, u=6.0, v=1.0, phix=0.3, phiy=0.7), [(3, 3, 0, 0), (4, 3, 1, 2), (3, 0, 1, 1)]),
    (dict(lx=4, ly=4, u=7.0), [(2, 2, 0, 0), (3, 2, 1, 3), (2, 2, 2, 2), (4, 4, 2, 0)]),
    (dict(lx=3, ly=4, tp=-0.25, u=6.0, v=1.0, phix=0.25, phiy=0.75), [(3, 3, 1, 2), (2, 2, 0, 0)]),
    (dict(lx=2, ly=3, tp=0.3, u=5.0, v=0.5, phix=0.25, phiy=0.5), [(3, 2, 1, 1)]),
    (dict(lx=1, ly=5, tx=0.7, u=3.0, v=0.4, phiy=0.2), [(2, 2, 0, 3)]),
]

FTLM_CASES = [  # (model index, block, R, steps)
    (1, (4, 4, 0, 0), 32, 120),
    (1, (3, 4, 1, 1), 8, 80),
    (2, (4, 3, 1, 2), 8, 80),
    (3, (3, 2, 1, 3), 4, 80),
    (3, (4, 4, 2, 0), 2, 80),
    (4, (3, 3, 1, 2), 4, 80),
    (0, (2, 2, 1, 1), 5, 40),
]


def block_fixtures():
    data = {}
    index = []
    for mi, (model, blocks) in enumerate(BLOCK_CASES):
        m = oracle_lib.model(**model)
        for (nup, ndn, mx, my) in blocks:
            key = f"m{mi}_{nup}_{ndn}_{mx}_{my}"
            print("block", key, flush=True)
            blk = ref.block(m, nup, ndn, mx, my)
            up, dn, deg, dnf = blk.reps()
            rp, col, val, diag = blk.csr()
            rng = np.random.default_rng(7)
            x = rng.standard_normal(blk.dim) + 1j * rng.standard_normal(blk.dim)
            hx = blk.apply(x)
            if blk.dim <= 20000:
                data[key + "_up"] = up
                data[key + "_dn"] = dn
                data[key + "_deg"] = deg
                data[key + "_dnf"] = dnf
                data[key + "_hx"] = hx
                data[key + "_diag"] = diag
            else:  # large blocks: order-sensitive checksums (x is regenerated from the seed)
                w = np.arange(1, blk.dim + 1, dtype=np.float64)
                data[key + "_sums"] = np.array([
                    float((up.astype(np.float64) * w).sum()), float((dn.astype(np.float64) * w).sum()),
                    float((deg * w).sum()), float((dnf * w).sum()), float((diag * w).sum()),
                    float((hx.real * w).sum()), float((hx.imag * w).sum())])
            if blk.nnz <= 60000:
                data[key + "_rowptr"] = rp
                data[key + "_col"] = col
                data[key + "_val"] = val
            else:
                data[key + "_csrsum"] = np.array([rp.sum(), col.astype(np.int64).sum(), val.sum().real, val.sum().imag,
                                                  np.abs(val).sum()])
            index.append(dict(key=key, model=model, block=[nup, ndn, mx, my], dim=int(blk.dim), nnz=int(blk.nnz),
                              full_dim=int(blk.full_dim), parents=int(blk.parents)))
    for mi, (nup, ndn, mx, my), R, steps in FTLM_CASES:
        model = BLOCK_CASES[mi][0]
        key = f"ftlm_m{mi}_{nup}_{ndn}_{mx}_{my}_R{R}_s{steps}"
        print("ftlm", key, flush=True)
        blk = ref.block(oracle_lib.model(**model), nup, ndn, mx, my)
        bs = ref.block_seed(12345, nup, ndn, mx, my)
        e, w = ref.ftlm_block(blk, R, steps, bs)
        data[key + "_eig"] = e
        data[key + "_w"] = w
        index.append(dict(key=key, model=model, block=[nup, ndn, mx, my], R=R, steps=steps, seed=12345,
                          block_seed=str(bs)))
    # classical Jacobi on random symmetric tridiagonal / dense matrices
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 7, 20, 45):
        A = rng.standard_normal((n, n))
        A = A + A.T
        e, v0 = ref.jacobi(A)
        data[f"jacobi_{n}_A"] = A
        data[f"jacobi_{n}_eig"] = e
        data[f"jacobi_{n}_v0"] = v0
    np.savez_compressed(os.path.join(HERE, "blocks.npz"), **data)
    with open(os.path.join(HERE, "blocks.json"), "w") as f:
        json.dump(index, f, indent=1)


ZHEEV_CASES = [  # (model, blocks): exact-block eigenvalues at every size the engine's solver meets
    (dict(lx=2, ly=2, tp=0.2, u=4.0, v=1.0), [(1, 1, 0, 0), (2, 1, 1, 0), (2, 2, 1, 1)]),
    (dict(lx=4, ly=2, u=8.0), [(4, 4, 0, 0), (3, 4, 1, 1), (2, 3, 2, 0), (1, 1, 0, 0)]),
    (dict(lx=3, ly=3, tp=-0.25, u=6.0, v=1.0, phix=0.25, phiy=0.5), [(4, 5, 0, 0), (4, 4, 1, 2), (3, 3, 2, 1),
                                                                     (2, 1, 1, 1)]),
    (dict(lx=4, ly=4, u=7.0), [(1, 2, 0, 0), (2, 1, 3, 1), (15, 15, 0, 0)]),
    (dict(lx=3, ly=4, tp=-0.25, u=6.0, v=1.0, phix=0.25, phiy=0.75), [(2, 1, 1, 2), (11, 10, 0, 3)]),
    (dict(lx=4, ly=2, u=0.0, v=0.0, phix=0.25, phiy=0.5), [(4, 4, 1, 0)]),
]


def zheev_fixtures():
    """Eigenvalues of exact blocks by the reference's LAPACK call (zheev on
    its Hermitised dense block), keyed zheev_m<model>_<nup>_<ndn>_<mx>_<my>."""
    data = {}
    index = []
    for mi, (model, blocks) in enumerate(ZHEEV_CASES):
        m = oracle_lib.model(**model)
        for (nup, ndn, mx, my) in blocks:
            blk = ref.block(m, nup, ndn, mx, my)
            key = f"zheev_m{mi}_{nup}_{ndn}_{mx}_{my}"
            print(key, blk.dim, flush=True)
            data[key] = exact_eigs(blk)
            index.append(dict(key=key, model=model, block=[nup, ndn, mx, my], dim=int(blk.dim),
                              source=ZHEEV_SOURCE))
    np.savez_compressed(os.path.join(HERE, "zheev.npz"), **data)
    with open(os.path.join(HERE, "zheev.json"), "w") as f:
        json.dump(index, f, indent=1)


def primitive_fixtures():
    out = {"lattices": {}, "twist": [], "mix": [], "block_seed": [], "hop_sign": [], "diag": [], "translate": [],
           "combos": {}}
    for lx, ly in [(2, 2), (4, 2), (3, 3), (1, 3), (3, 1), (2, 3), (4, 4), (1, 1), (5, 1), (2, 1)]:
        nb, bonds = ref.lattice(lx, ly)
        out["lattices"][f"{lx}x{ly}"] = {"nb": nb.tolist(), "bonds": bonds.tolist()}
    for md in [dict(lx=2, ly=2, phix=1.0, phiy=1.0), dict(lx=3, ly=4, phix=0.25, phiy=0.75),
               dict(lx=4, ly=4, phix=0.0, phiy=0.5), dict(lx=5, ly=3, phix=0.3, phiy=0.7)]:
        z = ref.twist_phases(oracle_lib.model(**md))
        out["twist"].append({"model": md, "re": z.real.tolist(), "im": z.imag.tolist()})
    for v in [0, 1, 12345, 2**63, 2**64 - 1, 0x9e3779b97f4a7c15]:
        out["mix"].append([str(v), str(ref.mix_u64(v))])
    for args in [(12345, 0, 0, 0, 0), (12345, 4, 4, 0, 0), (7, 8, 8, 3, 3), (2**64 - 1, 16, 0, 1, 2)]:
        out["block_seed"].append([str(a) for a in args] + [str(ref.block_seed(*args))])
    rng = np.random.default_rng(3)
    for _ in range(40):
        bits = int(rng.integers(0, 2**40))
        a, b = (int(t) for t in rng.integers(0, 40, 2))
        out["hop_sign"].append([str(bits), a, b, ref.hop_sign(bits, a, b)])
    for md in [dict(lx=4, ly=4, u=7.0, v=1.3), dict(lx=2, ly=2, u=4.0, v=1.0), dict(lx=3, ly=2, u=2.5, v=-0.5)]:
        L = md["lx"] * md["ly"]
        for _ in range(10):
            up = int(rng.integers(0, 2**L))
            dn = int(rng.integers(0, 2**L))
            out["diag"].append({"model": md, "up": up, "dn": dn, "e": ref.diag_energy(oracle_lib.model(**md), up, dn)})
    for lx, ly in [(4, 4), (3, 4), (2, 3), (4, 2), (5, 3)]:
        L = lx * ly
        for _ in range(10):
            up = int(rng.integers(0, 2**L))
            dn = int(rng.integers(0, 2**L))
            dx = int(rng.integers(0, lx))
            dy = int(rng.integers(0, ly))
            out["translate"].append([lx, ly, up, dn, dx, dy, *ref.translate(up, dn, dx, dy, lx, ly)])
    for L, n in [(4, 2), (8, 4), (16, 8), (12, 6), (9, 0), (5, 5)]:
        c = ref.combos(L, n)
        out["combos"][f"{L}_{n}"] = {"count": len(c), "first": [int(t) for t in c[:5]], "last": [int(t) for t in c[-5:]],
                                     "sum": str(int(np.sum(c.astype(object))))}
    with open(os.path.join(HERE, "primitives.json"), "w") as f:
        json.dump(out, f, indent=0)


def checkpoint_fixture():
    """A checkpoint written by the reference's own writer (cfg1 blocks)."""
    path = os.path.join(HERE, "ckpt_2x2.bin")
    if os.path.exists(path):
        os.remove(path)
    model = dict(lx=2, ly=2, tp=0.2, u=4.0, v=1.0)
    m = oracle_lib.model(**model)
    import ctypes as C

    dp = C.POINTER(C.c_double)
    n = 0
    for nup in range(5):
        for ndn in range(5):
            for my in range(2):
                for mx in range(2):
                    blk = ref.block(m, nup, ndn, mx, my)
                    if blk.dim == 0:
                        continue
                    if (nup + ndn) % 2 == 0:
                        e = np.ascontiguousarray(exact_eigs(blk))
                        w = np.zeros(0)
                        pref, ex = 1.0, 1
                    else:  # FTLM records, R = 3, 10 steps
                        e, w = ref.ftlm_block(blk, 3, 10, ref.block_seed(12345, nup, ndn, mx, my))
                        e = np.ascontiguousarray(e)
                        w = np.ascontiguousarray(w)
                        pref, ex = float(blk.dim), 0
                    rc = ref.lib.ref_checkpoint_append(path.encode(), nup, ndn, mx, my, nup + ndn, blk.dim, pref, ex,
                                                       len(e), e.ctypes.data_as(dp), len(w),
                                                       (w if len(w) else np.zeros(1)).ctypes.data_as(dp))
                    assert rc == 0
                    n += 1
    meta = ref.metadata_text(m, 5, 40, 256, 12345)
    with open(path + ".meta", "w") as f:
        f.write(meta)
    print("checkpoint records", n)


if __name__ == "__main__":
    which = sys.argv[1:] or ["primitives", "blocks", "checkpoint", "zheev", "csv"]
    if "primitives" in which:
        primitive_fixtures()
    if "blocks" in which:
        block_fixtures()
    if "checkpoint" in which:
        checkpoint_fixture()
    if "zheev" in which:
        zheev_fixtures()
    if "csv" in which:
        csv_fixtures()
    extra = [w for w in which if w in CSV_CASES]
    if extra:
        csv_fixtures(extra)
