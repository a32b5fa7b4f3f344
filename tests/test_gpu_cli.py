"""End-to-end parity of the drop-in executables on the GPU.

The golden CSVs under tests/golden/csv were produced by the reference's own
block loop functions (tests/golden/make_golden.py); exact blocks there use
numpy eigvalsh in place of zheev. Tolerances (relative, written per column in compare()):
  n                      1e-10 of |n| (the engine's target), no absolute floor
                         beyond 1e-300
  charge_correlation,
  compressibility        1e-9 of the value plus the FP64 cancellation floor
                         of <N^2> - <N>^2 (64 eps <N>^2 / L)
  log_partition          1e-11 relative
With serial reductions and FTLM on every block (--exact-block-threshold 0),
the Ritz values are bit-identical to the reference's and the host
accumulation uses the reference's order, so the CSV must match byte for byte.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "ftlm_extended_hubbard_development_amd", "ftlm_n_vs_mu")
ED_EXE = os.path.join(ROOT, "ftlm_extended_hubbard_development_amd", "ftlm_ed_n_vs_mu")

with open(os.path.join(GOLD, "csv", "cases.json")) as f:
    CASES = json.load(f)


def run(exe, args, cwd):
    print("run", os.path.basename(exe), " ".join(args), flush=True)
    p = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr + p.stdout[-2000:]
    return p.stdout


def load(path):
    return np.genfromtxt(path, delimiter=",", names=True)


def sites_of(args):
    """lattice sites L = lx * ly of a command line (reference defaults 2x2)"""
    lx = int(args[args.index("--lx") + 1]) if "--lx" in args else 2
    ly = int(args[args.index("--ly") + 1]) if "--ly" in args else 2
    return lx * ly


def beta_of(args):
    """the single beta of a command line (multi-beta files carry a column)"""
    return float(args[args.index("--beta") + 1]) if "--beta" in args else 2.0


def compare(out, gold, sites, beta=None):
    """Relative bounds per column (n: 1e-10 of |n|; charge correlation and
    compressibility: 1e-9 of the value, plus the cancellation floor of
    <N^2> - <N>^2 computed in FP64, 64 eps <N>^2 / L; log_partition: 1e-11).
    Only a denormal-scale floor (1e-300) is added to the n bound."""
    a, b = load(out), load(gold)
    assert a.dtype.names == b.dtype.names
    assert np.array_equal(a["mu"], b["mu"])
    if "beta" in a.dtype.names:
        assert np.array_equal(a["beta"], b["beta"])
    n_ok = np.abs(a["n"] - b["n"]) <= 1e-10 * np.abs(b["n"]) + 1e-300
    assert np.all(n_ok), (np.max(np.abs(a["n"] - b["n"]) / np.maximum(np.abs(b["n"]), 1e-300)), b["n"][~n_ok][:5])
    L = float(sites)
    assert "beta" in b.dtype.names or beta is not None
    cancel = 64 * np.finfo(float).eps * (b["n"] * L) ** 2 / L
    for col in ("charge_correlation", "compressibility"):
        scale = 1.0 if col == "charge_correlation" else (b["beta"] if "beta" in b.dtype.names else beta)
        bound = 1e-9 * np.abs(b[col]) + scale * cancel + 1e-300
        ok = np.abs(a[col] - b[col]) <= bound
        assert np.all(ok), (col, np.abs(a[col] - b[col])[~ok][:5], b[col][~ok][:5])
    assert np.all(np.abs(a["log_partition"] - b["log_partition"]) <= 1e-11 * np.abs(b["log_partition"]) + 1e-12)


@pytest.mark.parametrize("name", sorted(CASES))
def test_cli_matches_reference_sweep(tmp_path, name):
    out = str(tmp_path / "out.csv")
    stdout = run(EXE, CASES[name]["args"] + ["--output", out], str(tmp_path))
    assert "Wrote " + out in stdout
    compare(out, os.path.join(GOLD, "csv", name + ".csv"), sites_of(CASES[name]["args"]), beta_of(CASES[name]["args"]))


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n]["exact_blocks_from"] == "none"])
def test_cli_serial_byte_identical(tmp_path, name):
    out = str(tmp_path / "out.csv")
    run(EXE, CASES[name]["args"] + ["--output", out, "--reduction", "serial"], str(tmp_path))
    with open(out, "rb") as f1, open(os.path.join(GOLD, "csv", name + ".csv"), "rb") as f2:
        assert f1.read() == f2.read()


def test_cli_per_vector_matches(tmp_path):
    name = "4x2_ftlm_only_r4_s40"
    out = str(tmp_path / "out.csv")
    run(EXE, CASES[name]["args"] + ["--output", out, "--per-vector", "--reduction", "serial"], str(tmp_path))
    with open(out, "rb") as f1, open(os.path.join(GOLD, "csv", name + ".csv"), "rb") as f2:
        assert f1.read() == f2.read()


@pytest.mark.parametrize("reduction", ["serial", "tree"])
def test_resume_from_reference_checkpoint(tmp_path, reduction):
    """A checkpoint written by the reference's writer resumes: every block
    loads, nothing is recomputed, and the result is the reference
    accumulation of those records: bit for bit with serial reductions (the
    reference's streaming host sums), within 1e-12 with tree reductions (one
    device launch for all blocks, ftlm_thermo)."""
    ck = str(tmp_path / "ck.bin")
    shutil.copy(os.path.join(GOLD, "ckpt_2x2.bin"), ck)
    shutil.copy(os.path.join(GOLD, "ckpt_2x2.bin.meta"), ck + ".meta")
    out = str(tmp_path / "out.csv")
    args = ["--lx", "2", "--ly", "2", "--tp", "0.2", "--u", "4", "--v", "1", "--beta", "1.5", "--samples", "5",
            "--lanczos-steps", "40", "--mu-min", "-2", "--mu-max", "6", "--mu-count", "33", "--checkpoint", ck,
            "--output", out, "--reduction", reduction]
    before = open(ck, "rb").read()
    stdout = run(EXE, args, str(tmp_path))
    assert "loaded_blocks=84" in stdout
    assert open(ck, "rb").read() == before  # nothing appended
    # expected: oracle accumulation of the same records in block order
    import struct

    data = open(ck, "rb").read()
    pos = 8
    recs = {}
    while pos < len(data):
        assert data[pos:pos + 8] == b"BLKREC1\n"
        nup, ndn, mx, my, part = struct.unpack_from("<5i", data, pos + 8)
        dim, pref = struct.unpack_from("<Qd", data, pos + 28)
        ex = data[pos + 44]
        ne, nw = struct.unpack_from("<QQ", data, pos + 45)
        e = np.frombuffer(data, np.float64, ne, pos + 61)
        w = np.frombuffer(data, np.float64, nw, pos + 61 + 8 * ne)
        recs[(nup, ndn, mx, my)] = (e.copy(), w.copy(), pref, part)
        pos += 61 + 8 * (ne + nw)
    orc = oracle_lib.oracle()
    mu = orc.linspace(-2.0, 6.0, 33)
    st = dict(emin=np.full(33, np.inf), z=np.zeros(33), n1=np.zeros(33), n2=np.zeros(33))
    for nup in range(5):
        for ndn in range(5):
            for my in range(2):
                for mx in range(2):
                    if (nup, ndn, mx, my) in recs:
                        e, w, pref, part = recs[(nup, ndn, mx, my)]
                        orc.accumulate(e, w if len(w) else None, pref, part, 1.5, mu, st)
    n = st["n1"] / st["z"] / 4.0
    got = load(out)
    if reduction == "serial":
        assert np.array_equal(got["n"], np.array([float("%.15g" % v) for v in n]))
    else:
        assert np.max(np.abs(got["n"] - n)) <= 1e-12


def test_checkpoint_roundtrip_and_resume(tmp_path):
    """Checkpointed run == plain run; a resume after cutting the file mid
    record recomputes the missing blocks and gives the same bytes."""
    base = ["--lx", "4", "--ly", "2", "--u", "8", "--beta", "20", "--samples", "4", "--lanczos-steps", "30",
            "--mu-min", "-5", "--mu-max", "15", "--mu-count", "21", "--reduction", "serial"]
    plain = str(tmp_path / "plain.csv")
    run(EXE, base + ["--output", plain], str(tmp_path))
    ck = str(tmp_path / "ck.bin")
    o1 = str(tmp_path / "o1.csv")
    run(EXE, base + ["--checkpoint", ck, "--output", o1], str(tmp_path))
    assert open(o1, "rb").read() == open(plain, "rb").read()
    full = open(ck, "rb").read()
    cut = int(len(full) * 0.6)
    with open(ck, "wb") as f:
        f.write(full[:cut])
    o2 = str(tmp_path / "o2.csv")
    run(EXE, base + ["--checkpoint", ck, "--output", o2], str(tmp_path))
    assert open(o2, "rb").read() == open(plain, "rb").read()
    # a second resume reads a well-formed file (the torn tail was cut)
    o3 = str(tmp_path / "o3.csv")
    out3 = run(EXE, base + ["--checkpoint", ck, "--output", o3], str(tmp_path))
    assert open(o3, "rb").read() == open(plain, "rb").read()
    assert "loaded_blocks=620" in out3


def test_ed_executable(tmp_path):
    name = "cfg1_2x2"
    out = str(tmp_path / "ed.csv")
    run(ED_EXE, CASES[name]["args"] + ["--output", out], str(tmp_path))
    compare(out, os.path.join(GOLD, "csv", name + ".csv"), sites_of(CASES[name]["args"]), beta_of(CASES[name]["args"]))


def test_ed_4x2_bounds_ftlm(tmp_path):
    """The 4x2 exact diagonalisation bounds the FTLM error of cfg2 like the
    reference's own pair does (max |dn| 6.2e-3 there)."""
    if "ed_4x2_beta20" not in CASES:
        pytest.skip("no ED golden")
    out = str(tmp_path / "ed.csv")
    run(ED_EXE, CASES["ed_4x2_beta20"]["args"] + ["--output", out], str(tmp_path))
    compare(out, os.path.join(GOLD, "csv", "ed_4x2_beta20.csv"), 8, 20.0)
    ftlm = load(os.path.join(GOLD, "csv", "cfg2_4x2.csv"))
    ed = load(out)
    assert np.max(np.abs(ftlm["n"] - ed["n"])) < 1e-2


@pytest.mark.parametrize("pack", [["--pack-max-dim", "0"], ["--pack-max-dim", "100"],
                                  ["--pack-width", "8", "--pack-wg", "1"], ["--pack-budget-gb", "0.002"]],
                         ids=["unpacked", "mixed", "w8", "many-packs"])
@pytest.mark.parametrize("name", ["cfg2_4x2", "4x2_multibeta_r8_s60"])
def test_cli_pack_configurations(tmp_path, name, pack):
    """Packed, unpacked and mixed block execution give the reference's
    output: byte for byte in serial mode when every block is FTLM, within the
    stated tolerances otherwise."""
    if name not in CASES:
        pytest.skip("no golden")
    out = str(tmp_path / "out.csv")
    gold = os.path.join(GOLD, "csv", name + ".csv")
    if CASES[name]["exact_blocks_from"] == "none":
        run(EXE, CASES[name]["args"] + ["--output", out, "--reduction", "serial"] + pack, str(tmp_path))
        assert open(out, "rb").read() == open(gold, "rb").read()
    out2 = str(tmp_path / "out2.csv")
    run(EXE, CASES[name]["args"] + ["--output", out2] + pack, str(tmp_path))
    compare(out2, gold, sites_of(CASES[name]["args"]), beta_of(CASES[name]["args"]))


TW_ARGS = ["--lx", "3", "--ly", "3", "--tp", "-0.25", "--u", "6", "--v", "1", "--beta", "4", "--samples", "4",
           "--lanczos-steps", "30", "--mu-min", "-2", "--mu-max", "8", "--mu-count", "11",
           "--exact-block-threshold", "0"]


def test_cli_twist_grid_equals_separate_runs(tmp_path):
    """--twist-grid 2 runs the four twists {0, 1/2}^2 in one process with
    their blocks packed together; each per-twist CSV equals the reference
    run at that twist (the oracle's CLI), byte for byte in serial mode."""
    orc = oracle_lib.oracle()
    out = str(tmp_path / "tw.csv")
    stdout = run(EXE, TW_ARGS + ["--twist-grid", "2", "--reduction", "serial", "--output", out], str(tmp_path))
    assert "twist grid 2x2 (4 twists)" in stdout
    for px, py in [(0, 0), (0.5, 0), (0, 0.5), (0.5, 0.5)]:
        got = str(tmp_path / f"tw.phix{px:g}_phiy{py:g}.csv")
        assert os.path.exists(got), got
        ref = str(tmp_path / f"ref_{px}_{py}.csv")
        rc = orc.main(TW_ARGS + ["--phix", str(px), "--phiy", str(py), "--output", ref])
        assert rc == 0
        assert open(got, "rb").read() == open(ref, "rb").read(), (px, py)
    # tree reductions (QL Ritz in the packs): within the stated tolerances
    out2 = str(tmp_path / "tt.csv")
    run(EXE, TW_ARGS + ["--twist-grid", "2", "--output", out2], str(tmp_path))
    for px, py in [(0, 0), (0.5, 0), (0, 0.5), (0.5, 0.5)]:
        compare(str(tmp_path / f"tt.phix{px:g}_phiy{py:g}.csv"), str(tmp_path / f"ref_{px}_{py}.csv"), 9, 4.0)


def _raw_thermo(e, w, pref, N, beta, mu):
    sh = e - mu * N
    m = sh.min()
    wt = pref * w * np.exp(-beta * (sh - m)) if w is not None else np.exp(-beta * (sh - m))
    return wt.sum(), N * wt.sum(), N * N * wt.sum(), m


TR_ARGS = CASES["cfg1_2x2"]["args"]  # 2x2, beta 1.5, 5 samples, 40 steps, 33 mu points
TR_MODEL = dict(lx=2, ly=2, tp=0.2, u=4.0, v=1.0)


@pytest.mark.parametrize("threshold", ["0", "256"])
def test_cli_trace_partition_csv(tmp_path, threshold):
    """--trace-partition-csv: per block and mu, the used (FTLM or exact) and
    the exact partition sums (main.cpp:1505-1663), against the oracle's
    spectra and numpy's eigenvalues of the reference's dense blocks."""
    orc = oracle_lib.oracle()
    tr = str(tmp_path / "trace.csv")
    out = str(tmp_path / "out.csv")
    run(EXE, TR_ARGS + ["--exact-block-threshold", threshold, "--reduction", "serial", "--trace-partition-csv", tr,
                        "--output", out], str(tmp_path))
    rows = list(__import__("csv").DictReader(open(tr)))
    m = oracle_lib.model(**TR_MODEL)
    mu = orc.linspace(-2.0, 6.0, 33)
    k = 0
    for nup in range(5):
        for ndn in range(5):
            for my in range(2):
                for mx in range(2):
                    blk = orc.block(m, nup, ndn, mx, my)
                    if blk.dim == 0:
                        continue
                    ed = np.linalg.eigvalsh(blk.dense())
                    ftlm = blk.dim > int(threshold)
                    if ftlm:
                        e, w = orc.ftlm_block(blk, 5, 40, orc.block_seed(12345, nup, ndn, mx, my))
                    for mv in mu:
                        r = rows[k]
                        k += 1
                        assert (int(r["nup"]), int(r["ndown"]), int(r["mx"]), int(r["my"])) == (nup, ndn, mx, my)
                        assert int(r["basis_dim"]) == blk.dim and r["method"] == ("ftlm" if ftlm else "exact")
                        N = nup + ndn
                        zu = _raw_thermo(e, w, blk.dim, N, 1.5, mv) if ftlm else _raw_thermo(ed, None, 1, N, 1.5, mv)
                        ze = _raw_thermo(ed, None, 1, N, 1.5, mv)
                        assert abs(float(r["z_used"]) - zu[0]) <= 1e-10 * abs(zu[0])
                        assert abs(float(r["z_ed"]) - ze[0]) <= 1e-10 * abs(ze[0])
                        assert abs(float(r["n_used_total"]) - zu[1]) <= 1e-10 * abs(zu[1]) + 1e-300
                        assert abs(float(r["min_ftlm"]) - zu[3]) <= 1e-12 * (1 + abs(zu[3]))
                        assert abs(float(r["min_ed"]) - ze[3]) <= 1e-12 * (1 + abs(ze[3]))
                        sw = w.sum() if ftlm else 1.0
                        assert abs(float(r["sum_sample_weights"]) - sw) <= 1e-13
    assert k == len(rows)
    compare(out, os.path.join(GOLD, "csv", "cfg1_2x2_ftlm_only.csv" if threshold == "0" else "cfg1_2x2.csv"), 4, 1.5)


def test_cli_debug_block(tmp_path):
    """--debug-block-*: the selected FTLM block's comparison with its exact
    spectrum (stdout lines and CSV of main.cpp:1233-1267, 1627-1663)."""
    orc = oracle_lib.oracle()
    dbg = str(tmp_path / "dbg.csv")
    out = str(tmp_path / "out.csv")
    stdout = run(EXE, TR_ARGS + ["--exact-block-threshold", "0", "--reduction", "serial", "--debug-block-nup", "2",
                                 "--debug-block-ndown", "1", "--debug-block-mx", "1", "--debug-block-my", "0",
                                 "--debug-block-csv", dbg, "--output", out], str(tmp_path))
    blk = orc.block(oracle_lib.model(**TR_MODEL), 2, 1, 1, 0)
    assert f"Selected block debug Nup=2 Ndown=1 mx=1 my=0 basis_dim={blk.dim} method=ftlm" in stdout
    assert stdout.count("  debug mu=") == 33
    rows = list(__import__("csv").DictReader(open(dbg)))
    assert len(rows) == 33
    e, w = orc.ftlm_block(blk, 5, 40, orc.block_seed(12345, 2, 1, 1, 0))
    ed = np.linalg.eigvalsh(blk.dense())
    for r, mv in zip(rows, orc.linspace(-2.0, 6.0, 33)):
        zf = _raw_thermo(e, w, blk.dim, 3, 1.5, mv)
        ze = _raw_thermo(ed, None, 1, 3, 1.5, mv)
        assert abs(float(r["z_ftlm"]) - zf[0]) <= 1e-10 * zf[0] and abs(float(r["z_ed"]) - ze[0]) <= 1e-10 * ze[0]
        assert abs(float(r["n_diff"]) - (zf[1] - ze[1])) <= 1e-9 * max(zf[1], ze[1])
    compare(out, os.path.join(GOLD, "csv", "cfg1_2x2_ftlm_only.csv"), 4, 1.5)
    # a debug target that is not a block fails like the reference
    p = subprocess.run([EXE] + TR_ARGS + ["--debug-block-nup", "9", "--debug-block-ndown", "1", "--debug-block-mx",
                                          "0", "--debug-block-my", "0", "--debug-block-csv", dbg,
                                          "--output", out], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 1 and "Selected block debug target was not found" in p.stderr


def test_ed_3x3_blocks_up_to_1764(tmp_path):
    """The ED executable on the twisted 3x3 lattice, where blocks reach 1,764
    rows (ed_main.cpp:74-113 diagonalises every block), against the
    reference's block loop with LAPACK zheev on every block."""
    import time

    name = "ed_3x3_twist_beta10"
    out = str(tmp_path / "ed.csv")
    t0 = time.time()
    run(ED_EXE, CASES[name]["args"] + ["--output", out], str(tmp_path))
    print(f"ED 3x3 (868 blocks, dims up to 1764): {time.time() - t0:.2f} s wall", flush=True)
    compare(out, os.path.join(GOLD, "csv", name + ".csv"), 9, 10.0)


CFG4_TWISTS = {"cfg4_3x4_tw_0.25_0.75_r4_s40": (0.25, 0.75), "cfg4_3x4_tw_0.5_0_r4_s40": (0.5, 0.0)}


@pytest.mark.parametrize("reduction", ["tree", "serial"])
def test_cfg4_model_twist_list(tmp_path, reduction):
    """The twisted 3x4 model (t' = -0.25, U = 6, V = 1, beta = 10) at two
    twists of its 4x4 grid in one --twist-list run (blocks of both twists
    packed together, exact blocks queued across twists): each twist's CSV
    against the reference's block loop run at that twist alone."""
    names = sorted(CFG4_TWISTS)
    if any(n not in CASES for n in names):
        pytest.skip("3x4 goldens not generated")
    args = list(CASES[names[0]]["args"])
    for flag in ("--phix", "--phiy"):
        i = args.index(flag)
        del args[i:i + 2]
    tl = ",".join(f"{CFG4_TWISTS[n][0]}:{CFG4_TWISTS[n][1]}" for n in names)
    out = str(tmp_path / "c4.csv")
    stdout = run(EXE, args + ["--twist-list", tl, "--reduction", reduction, "--output", out], str(tmp_path))
    assert "twist list (2 twists)" in stdout
    for t, n in enumerate(names):
        compare(str(tmp_path / f"c4.twist{t}.csv"), os.path.join(GOLD, "csv", n + ".csv"), 12, 10.0)


@pytest.mark.parametrize("reduction", ["serial", "tree"])
def test_partial_sums_resume(tmp_path, reduction):
    """--partial-sums with --stop-after: a run stopped between sectors and
    started again until done gives the uninterrupted run's CSV -- byte for
    byte with serial reductions (the streaming sums continue from the stored
    hexadecimal state), within the stated tolerances with tree reductions
    (device sums folded per sector). Each run here stops after one sector."""
    base = CASES["cfg1_2x2_ftlm_only"]["args"] + ["--beta-list", "1.5,4.0", "--reduction", reduction]
    plain = str(tmp_path / "plain.csv")
    run(EXE, base + ["--output", plain], str(tmp_path))
    ps = str(tmp_path / "run.sums")
    out = str(tmp_path / "out.csv")
    runs = 0
    while True:
        stdout = run(EXE, base + ["--output", out, "--partial-sums", ps, "--stop-after", "1e-9"], str(tmp_path))
        runs += 1
        assert runs < 200
        if "Wrote " + out in stdout:
            break
        assert "Stopped before sector" in stdout and not os.path.exists(out)
    assert runs == 25  # one sector per run
    assert "partial sums: 84 blocks summed over 25 run(s)" in stdout
    if reduction == "serial":
        assert open(out, "rb").read() == open(plain, "rb").read()
    compare(out, plain, 4)
    # a partial-sum file of another run is refused
    p = subprocess.run([EXE] + base + ["--seed", "7", "--output", out, "--partial-sums", ps], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 1 and "Partial-sum file does not match current run" in p.stderr
