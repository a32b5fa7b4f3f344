"""The drop-in host layer (libftlm_host.so: flags, checkpoint, thermodynamic
accumulation, CSV) on the CPU, against the reference's formats: the golden
checkpoint written by the reference's own writer, the golden CSVs written by
its write_results, and the oracle's accumulation (bit-identical to the
reference's, tests/test_oracle_golden.py). No GPU is involved."""
import ctypes as C
import os
import shutil
import struct

import numpy as np
import pytest

import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
HOST = os.path.join(ROOT, "ftlm_extended_hubbard_development_amd", "libftlm_host.so")

dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST):
        pytest.fail("libftlm_host.so not built")
    L = C.CDLL(HOST)
    L.ftlmh_last_error.restype = C.c_char_p
    L.ftlmh_parse.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
    L.ftlmh_metadata.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_char_p, C.c_int]
    L.ftlmh_checkpoint_count.restype = C.c_longlong
    L.ftlmh_checkpoint_count.argtypes = [C.c_char_p, C.POINTER(C.c_ulonglong)]
    L.ftlmh_checkpoint_get.restype = C.c_longlong
    L.ftlmh_checkpoint_get.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp,
                                       C.POINTER(C.c_longlong), dp, C.POINTER(C.c_ulonglong), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]
    L.ftlmh_checkpoint_resume.restype = C.c_longlong
    L.ftlmh_checkpoint_resume.argtypes = [C.c_char_p]
    L.ftlmh_checkpoint_append.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong,
                                          C.c_double, C.c_int, C.c_longlong, dp, C.c_longlong, dp]
    L.ftlmh_accumulate.argtypes = [dp, dp, C.c_longlong, C.c_double, C.c_int, C.c_double, dp, C.c_int, dp, dp, dp,
                                   dp, C.c_int]
    L.ftlmh_write_results.argtypes = [C.c_char_p, C.c_int, dp, C.c_int, dp, dp, dp, dp, dp, dp]
    L.ftlmh_partial_write.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_longlong,
                                      C.c_int, C.c_int, C.c_int, dp]
    L.ftlmh_partial_read.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_int), dp,
                                     C.POINTER(C.c_longlong), dp, C.c_longlong]
    L.ftlmh_merge.argtypes = [C.c_int, C.c_double, dp, dp]
    return L


def argv(args):
    a = [b"ftlm_n_vs_mu"] + [x.encode() for x in args]
    return len(a), (C.c_char_p * len(a))(*a)


# messages of parse_args (main.cpp:257-368) for bad command lines
BAD = [
    (["--lx", "0"], "Lattice dimensions must be positive."),
    (["--lx", "9", "--ly", "8"], "This implementation supports at most 63 sites."),
    (["--beta", "-1"], "All beta values must be positive."),
    (["--beta-list", "1,,2"], "Empty entry in comma-separated value list: 1,,2"),
    (["--samples", "0"], "--samples and --lanczos-steps must be positive."),
    (["--threads", "-2"], "--threads must be non-negative."),
    (["--mu-count", "1"], "--mu-count must be at least 2."),
    (["--debug-block-nup", "1"], "Selected-block debug requires --debug-block-nup/--debug-block-ndown/"
                                 "--debug-block-mx/--debug-block-my together."),
    (["--u"], "Missing value for --u"),
    (["--u", "abc"], "Failed to parse value: abc"),
    (["--u", "1.5x"], "Failed to parse value: 1.5x"),
    (["--frobnicate"], "Unknown argument: --frobnicate"),
]


@pytest.mark.parametrize("args,msg", BAD, ids=[" ".join(a) for a, _ in BAD])
def test_parse_errors(host, args, msg):
    n, a = argv(args)
    assert host.ftlmh_parse(n, a) == 1
    assert host.ftlmh_last_error().decode() == msg


# the engine's additive flags for resumable runs
BAD_ENGINE = [
    (["--stop-after", "10"], "--stop-after needs --partial-sums or --checkpoint to resume from."),
    (["--stop-after", "-1", "--partial-sums", "p"], "--stop-after must be non-negative."),
    (["--partial-sums", "p", "--checkpoint", "c"], "--partial-sums and --checkpoint are exclusive."),
]


@pytest.mark.parametrize("args,msg", BAD_ENGINE, ids=[" ".join(a) for a, _ in BAD_ENGINE])
def test_parse_errors_engine_flags(host, args, msg):
    n, a = argv(args)
    assert host.ftlmh_parse(n, a) == 1
    assert host.ftlmh_last_error().decode() == msg


def test_parse_errors_match_reference(host, ref):
    for args, _ in BAD:
        rc, err = ref.parse_args(args)
        n, a = argv(args)
        assert host.ftlmh_parse(n, a) == 1 and rc == 1
        assert host.ftlmh_last_error().decode() == err


def test_parse_ok_and_help(host):
    n, a = argv(["--lx", "4", "--ly", "4", "--beta-list", "2.857142857142857,12.5", "--max-sector-dim",
                 "165636900", "--reduction", "serial", "--per-vector", "--device", "0"])
    assert host.ftlmh_parse(n, a) == 0


def test_metadata_matches_reference_file(host):
    """.meta text (main.cpp:370-387) for the golden checkpoint's run."""
    args = ["--lx", "2", "--ly", "2", "--tp", "0.2", "--u", "4", "--v", "1", "--samples", "5", "--lanczos-steps",
            "40"]
    n, a = argv(args)
    buf = C.create_string_buffer(4096)
    k = host.ftlmh_metadata(n, a, buf, 4096)
    with open(os.path.join(GOLD, "ckpt_2x2.bin.meta")) as f:
        assert buf.value[:k].decode() == f.read()


def records(path):
    data = open(path, "rb").read()
    assert data[:8] == b"FTLMCP1\n"
    pos = 8
    out = []
    while pos + 8 <= len(data) and data[pos:pos + 8] == b"BLKREC1\n":
        if pos + 61 > len(data):
            break
        nup, ndn, mx, my, part = struct.unpack_from("<5i", data, pos + 8)
        dim, pref = struct.unpack_from("<Qd", data, pos + 28)
        ex = data[pos + 44]
        ne, nw = struct.unpack_from("<QQ", data, pos + 45)
        end = pos + 61 + 8 * (ne + nw)
        if end > len(data):
            break
        e = np.frombuffer(data, np.float64, ne, pos + 61).copy()
        w = np.frombuffer(data, np.float64, nw, pos + 61 + 8 * ne).copy()
        out.append(((nup, ndn, mx, my), part, dim, pref, ex, e, w, pos, end))
        pos = end
    return out


def test_read_reference_checkpoint(host):
    path = os.path.join(GOLD, "ckpt_2x2.bin")
    valid = C.c_ulonglong(0)
    recs = records(path)
    assert host.ftlmh_checkpoint_count(path.encode(), C.byref(valid)) == len(recs) == 84
    assert valid.value == os.path.getsize(path)
    for key, part, dim, pref, ex, e, w, _, _ in recs[::7]:
        eb = np.zeros(max(len(e), 1))
        wb = np.zeros(max(len(w), 1))
        nw = C.c_longlong(0)
        pf = C.c_double(0)
        dd = C.c_ulonglong(0)
        pp = C.c_int(0)
        xx = C.c_int(0)
        ne = host.ftlmh_checkpoint_get(path.encode(), *key, eb.ctypes.data_as(dp), wb.ctypes.data_as(dp),
                                       C.byref(nw), C.byref(pf), C.byref(dd), C.byref(pp), C.byref(xx))
        assert ne == len(e) and nw.value == len(w)
        assert np.array_equal(eb[:ne], e) and np.array_equal(wb[: nw.value], w)
        assert (pf.value, dd.value, pp.value, xx.value) == (pref, dim, part, ex)


def test_append_is_byte_identical(host, tmp_path):
    """Re-writing the reference-written records with the host writer gives
    the same bytes (header, field order, padding-free little endian)."""
    src = os.path.join(GOLD, "ckpt_2x2.bin")
    out = str(tmp_path / "ck.bin")
    for key, part, dim, pref, ex, e, w, _, _ in records(src):
        e = np.ascontiguousarray(e)
        w = np.ascontiguousarray(w) if len(w) else np.zeros(1)
        assert host.ftlmh_checkpoint_append(out.encode(), *key, part, dim, pref, ex, len(e), e.ctypes.data_as(dp),
                                            len(w) if ex == 0 else 0, w.ctypes.data_as(dp)) == 0
    assert open(out, "rb").read() == open(src, "rb").read()


@pytest.mark.parametrize("frac", [0.13, 0.5, 0.77, 0.999])
def test_torn_tail_is_cut_on_resume(host, tmp_path, frac):
    """A file cut mid-record loads every complete record, and resuming cuts
    the torn tail so that later appends stay readable (the reference appends
    after the torn bytes, and its second resume then reads garbage)."""
    src = os.path.join(GOLD, "ckpt_2x2.bin")
    recs = records(src)
    data = open(src, "rb").read()
    cut = int(len(data) * frac)
    path = str(tmp_path / "ck.bin")
    with open(path, "wb") as f:
        f.write(data[:cut])
    complete = [r for r in recs if r[8] <= cut]
    assert host.ftlmh_checkpoint_resume(path.encode()) == len(complete)
    assert os.path.getsize(path) == (complete[-1][8] if complete else 8)
    # append the missing records, then a second resume sees all of them
    for key, part, dim, pref, ex, e, w, _, end in recs[len(complete):]:
        e = np.ascontiguousarray(e)
        w = np.ascontiguousarray(w) if len(w) else np.zeros(1)
        host.ftlmh_checkpoint_append(path.encode(), *key, part, dim, pref, ex, len(e), e.ctypes.data_as(dp),
                                     len(w) if ex == 0 else 0, w.ctypes.data_as(dp))
    assert host.ftlmh_checkpoint_resume(path.encode()) == len(recs)
    assert open(path, "rb").read() == data


def test_bad_magic_is_an_error(host, tmp_path):
    p = str(tmp_path / "x.bin")
    with open(p, "wb") as f:
        f.write(b"NOTFTLM\n" + b"\0" * 16)
    assert host.ftlmh_checkpoint_count(p.encode(), None) == -1
    assert host.ftlmh_last_error().decode() == "Checkpoint has an unrecognized format: " + p


@pytest.mark.parametrize("threads", [1, 4])
def test_accumulate_bit_identical(host, orc, threads):
    """accumulate_block_thermo order: bit-identical to the oracle (itself
    bit-identical to the reference) with any host thread count."""
    rng = np.random.default_rng(5)
    mu = orc.linspace(-5.0, 15.0, 81)
    s_h = [np.full(81, np.inf), np.zeros(81), np.zeros(81), np.zeros(81)]
    s_o = dict(emin=np.full(81, np.inf), z=np.zeros(81), n1=np.zeros(81), n2=np.zeros(81))
    for blk in range(30):
        n = int(rng.integers(1, 4000))
        e = np.ascontiguousarray(rng.standard_normal(n) * 6 + rng.integers(-5, 5))
        w = np.ascontiguousarray(rng.random(n)) if blk % 3 else None
        N = int(rng.integers(0, 16))
        pref = float(rng.integers(1, 1000)) if w is not None else 1.0
        host.ftlmh_accumulate(e.ctypes.data_as(dp), None if w is None else w.ctypes.data_as(dp), n, pref, N, 12.5,
                              mu.ctypes.data_as(dp), 81, *[a.ctypes.data_as(dp) for a in s_h], threads)
        orc.accumulate(e, w, pref, N, 12.5, mu, s_o)
    for a, k in zip(s_h, ["emin", "z", "n1", "n2"]):
        assert np.array_equal(a, s_o[k])


def test_write_results_bytes(host, tmp_path):
    """CSV writer (main.cpp:1409-1455): re-writing the values parsed from a
    reference-written CSV reproduces its bytes (single- and multi-beta)."""
    for name in ["cfg1_2x2", "4x2_multibeta_r8_s60"]:
        gold = os.path.join(GOLD, "csv", name + ".csv")
        t = np.genfromtxt(gold, delimiter=",", names=True)
        multi = "beta" in t.dtype.names
        betas = np.unique(t["beta"]) if multi else np.array([1.0])
        if multi:
            betas = np.array(list(dict.fromkeys(t["beta"].tolist())))
        nmu = len(t) // len(betas)
        mu = np.ascontiguousarray(t["mu"][:nmu])
        cols = [np.ascontiguousarray(t[c]) for c in
                ("n", "charge_correlation", "compressibility", "partition_like", "log_partition")]
        out = str(tmp_path / (name + ".csv"))
        b = np.ascontiguousarray(betas)
        assert host.ftlmh_write_results(out.encode(), len(betas) if multi else 1, b.ctypes.data_as(dp), nmu,
                                        mu.ctypes.data_as(dp), *[c.ctypes.data_as(dp) for c in cols]) == 0
        assert open(out, "rb").read() == open(gold, "rb").read()


def test_partial_sum_file_roundtrip(host, tmp_path):
    """--partial-sums files keep every double exactly (hexadecimal), the
    infinite minima of empty streams included, and the run identity."""
    path = str(tmp_path / "p.sums").encode()
    rng = np.random.default_rng(5)
    nt, nb, nmu = 2, 3, 7
    sums = rng.standard_normal(nt * nb * 4 * nmu) * 10.0 ** rng.integers(-300, 300, nt * nb * 4 * nmu)
    sums[0] = np.inf
    sums[5] = 0.0
    meta = b"lx=4\nly=4\nbetas=2.857142857142857,12.5\n"
    assert host.ftlmh_partial_write(path, meta, 5, 11, 3, 1234.5678, 999, nt, nb, nmu, sums.ctypes.data_as(dp)) == 0
    buf = C.create_string_buffer(4096)
    head = (C.c_int * 6)()
    el = C.c_double(0)
    nblk = C.c_longlong(0)
    back = np.zeros_like(sums)
    assert host.ftlmh_partial_read(path, buf, 4096, head, C.byref(el), C.byref(nblk), back.ctypes.data_as(dp),
                                   back.size) == 1
    assert buf.value == meta and list(head) == [5, 11, 3, nt, nb, nmu]
    assert el.value == 1234.5678 and nblk.value == 999
    assert np.array_equal(back.view(np.uint64), sums.view(np.uint64))
    assert host.ftlmh_partial_read(str(tmp_path / "absent").encode(), buf, 4096, head, C.byref(el), C.byref(nblk),
                                   back.ctypes.data_as(dp), back.size) == 0
    (tmp_path / "bad.sums").write_text("FTLM_PARTIAL_SUMS 1\nmeta_lines 1\nx\nnext 1\n")
    assert host.ftlmh_partial_read(str(tmp_path / "bad.sums").encode(), buf, 4096, head, C.byref(el),
                                   C.byref(nblk), back.ctypes.data_as(dp), back.size) == -1
    assert b"Malformed partial-sum file" in host.ftlmh_last_error()


def test_merge_equals_one_stream(host, orc):
    """Folding the sums of a second set of blocks into those of a first
    gives the sums of one stream over all of them (the reference's streaming
    accumulation, main.cpp:1172-1219) to rounding; an empty side is the
    identity."""
    rng = np.random.default_rng(9)
    mu = orc.linspace(-3.0, 4.0, 41)
    beta = 12.5
    blocks = []
    for _ in range(12):
        n = int(rng.integers(1, 50))
        e = np.sort(rng.standard_normal(n) * 5 - 20)
        w = rng.random(n)
        w /= w.sum()
        blocks.append((e, w, float(rng.integers(1, 500)), int(rng.integers(0, 17))))

    def stream(bl):
        st = dict(emin=np.full(41, np.inf), z=np.zeros(41), n1=np.zeros(41), n2=np.zeros(41))
        for e, w, pref, N in bl:
            orc.accumulate(e, w, pref, N, beta, mu, st)
        return np.concatenate([st["emin"], st["z"], st["n1"], st["n2"]])

    whole = stream(blocks)
    run = stream(blocks[:5])
    part = stream(blocks[5:])
    host.ftlmh_merge(41, beta, run.ctypes.data_as(dp), part.ctypes.data_as(dp))
    assert np.array_equal(run[:41], whole[:41])
    assert np.allclose(run[41:], whole[41:], rtol=1e-13, atol=0)
    empty = np.concatenate([np.full(41, np.inf), np.zeros(123)])
    before = part.copy()
    host.ftlmh_merge(41, beta, part.ctypes.data_as(dp), empty.ctypes.data_as(dp))
    assert np.array_equal(part, before)
    host.ftlmh_merge(41, beta, empty.ctypes.data_as(dp), before.ctypes.data_as(dp))
    assert np.array_equal(empty, before)
