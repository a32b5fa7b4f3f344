"""The CPU oracle against the golden fixtures written from the reference.

Every comparison is bit-exact unless a tolerance is written next to it.
Runs anywhere the oracle is built (no reference sources, no GPU needed).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
ORACLE_CLI = os.path.join(os.path.dirname(HERE), "oracle", "ftlm_oracle_cli")

with open(os.path.join(GOLD, "primitives.json")) as f:
    PRIM = json.load(f)
with open(os.path.join(GOLD, "blocks.json")) as f:
    INDEX = json.load(f)
BLOCKS = np.load(os.path.join(GOLD, "blocks.npz"))


def test_lattices(orc):
    for name, d in PRIM["lattices"].items():
        lx, ly = (int(t) for t in name.split("x"))
        nb, bonds = orc.lattice(lx, ly)
        assert nb.tolist() == d["nb"], name
        assert bonds.tolist() == d["bonds"], name


def test_twist_phases(orc):
    for t in PRIM["twist"]:
        z = orc.twist_phases(oracle_lib.model(**t["model"]))
        assert z.real.tolist() == t["re"] and z.imag.tolist() == t["im"]


def test_seeds(orc):
    for v, m in PRIM["mix"]:
        assert orc.mix_u64(int(v)) == int(m)
    for row in PRIM["block_seed"]:
        args = [int(a) for a in row[:5]]
        assert orc.block_seed(*args) == int(row[5])


def test_hop_sign_diag_translate(orc):
    for bits, a, b, s in PRIM["hop_sign"]:
        assert orc.hop_sign(int(bits), a, b) == s
    for d in PRIM["diag"]:
        assert orc.diag_energy(oracle_lib.model(**d["model"]), d["up"], d["dn"]) == d["e"]
    for lx, ly, up, dn, dx, dy, u2, d2, sg in PRIM["translate"]:
        assert orc.shift_state(up, dn, dy * lx + dx, lx, ly) == (u2, d2, sg)


def test_combos(orc):
    for key, d in PRIM["combos"].items():
        L, n = (int(t) for t in key.split("_"))
        c = orc.combos(L, n)
        assert len(c) == d["count"]
        assert [int(t) for t in c[:5]] == d["first"] and [int(t) for t in c[-5:]] == d["last"]
        assert str(int(np.sum(c.astype(object)))) == d["sum"]


_blocks = {}


def oracle_block(orc, model, blk):
    key = (json.dumps(model, sort_keys=True), tuple(blk))
    if key not in _blocks:
        _blocks[key] = orc.block(oracle_lib.model(**model), *blk)
    return _blocks[key]


@pytest.mark.parametrize("entry", [e for e in INDEX if "R" not in e], ids=lambda e: e["key"])
def test_blocks(orc, entry):
    b = oracle_block(orc, entry["model"], entry["block"])
    k = entry["key"]
    assert (b.dim, b.nnz, b.full_dim, b.parents) == (entry["dim"], entry["nnz"], entry["full_dim"], entry["parents"])
    up, dn, deg, dnf = b.reps()
    rp, col, val, diag = b.csr()
    rng = np.random.default_rng(7)
    x = rng.standard_normal(b.dim) + 1j * rng.standard_normal(b.dim)
    hx = b.apply(x)
    if k + "_up" in BLOCKS:
        assert np.array_equal(up, BLOCKS[k + "_up"]) and np.array_equal(dn, BLOCKS[k + "_dn"])
        assert np.array_equal(deg, BLOCKS[k + "_deg"])
        assert np.array_equal(dnf.view(np.uint64), BLOCKS[k + "_dnf"].view(np.uint64))
        assert np.array_equal(diag, BLOCKS[k + "_diag"])
        assert np.array_equal(hx, BLOCKS[k + "_hx"])
    else:
        w = np.arange(1, b.dim + 1, dtype=np.float64)
        sums = np.array([float((up.astype(np.float64) * w).sum()), float((dn.astype(np.float64) * w).sum()),
                         float((deg * w).sum()), float((dnf * w).sum()), float((diag * w).sum()),
                         float((hx.real * w).sum()), float((hx.imag * w).sum())])
        assert np.array_equal(sums, BLOCKS[k + "_sums"])
    if k + "_col" in BLOCKS:
        assert np.array_equal(rp, BLOCKS[k + "_rowptr"]) and np.array_equal(col, BLOCKS[k + "_col"])
        assert np.array_equal(val, BLOCKS[k + "_val"])


@pytest.mark.parametrize("entry", [e for e in INDEX if "R" in e], ids=lambda e: e["key"])
def test_ftlm_spectra(orc, entry):
    """run_ftlm_block: Ritz values and weights bit-identical to the reference."""
    b = oracle_block(orc, entry["model"], entry["block"])
    bs = orc.block_seed(entry["seed"], *entry["block"])
    assert bs == int(entry["block_seed"])
    e, w = orc.ftlm_block(b, entry["R"], entry["steps"], bs)
    assert np.array_equal(e, BLOCKS[entry["key"] + "_eig"])
    assert np.array_equal(w, BLOCKS[entry["key"] + "_w"])


@pytest.mark.parametrize("n", [1, 2, 3, 7, 20, 45])
def test_jacobi(orc, n):
    import ctypes as C

    A = np.ascontiguousarray(BLOCKS[f"jacobi_{n}_A"])
    e = np.zeros(n)
    v = np.zeros(n)
    dp = C.POINTER(C.c_double)
    orc.lib.orc_jacobi(n, A.copy().ctypes.data_as(dp), e.ctypes.data_as(dp), v.ctypes.data_as(dp))
    assert np.array_equal(e, BLOCKS[f"jacobi_{n}_eig"]) and np.array_equal(v, BLOCKS[f"jacobi_{n}_v0"])


with open(os.path.join(GOLD, "csv", "cases.json")) as f:
    CASES = json.load(f)

# oracle end-to-end runs cheap enough for the CPU suite
CPU_CASES = ["cfg1_2x2", "cfg1_2x2_ftlm_only", "4x2_ftlm_only_r4_s40", "3x3_twist_r4_s60"]


@pytest.mark.parametrize("name", CPU_CASES)
def test_oracle_cli_vs_golden_csv(tmp_path, name):
    """End-to-end n(mu): FTLM-only runs are byte-identical (no exact block);
    runs with exact blocks agree to 1e-12 (the oracle's Householder+QL vs
    the golden's eigvalsh)."""
    out = str(tmp_path / "o.csv")
    p = subprocess.run([ORACLE_CLI] + CASES[name]["args"] + ["--output", out, "--threads", "4"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    gold = os.path.join(GOLD, "csv", name + ".csv")
    if CASES[name]["exact_blocks_from"] == "none":
        assert open(out, "rb").read() == open(gold, "rb").read()
    else:
        a = np.genfromtxt(out, delimiter=",", names=True)
        g = np.genfromtxt(gold, delimiter=",", names=True)
        assert np.max(np.abs(a["n"] - g["n"])) <= 1e-12
        assert np.max(np.abs(a["compressibility"] - g["compressibility"])) <= 1e-11
