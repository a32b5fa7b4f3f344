"""Command-line behaviour of the drop-in executables that does not need a GPU:
help, flag errors (the reference's messages and exit code 1), checkpoint
.meta guard and flag conflicts, all decided before the engine is created;
and, without a GPU, a loud failure instead of any CPU fallback."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ftlm_extended_hubbard_development_amd", "ftlm_n_vs_mu")
ED = os.path.join(ROOT, "ftlm_extended_hubbard_development_amd", "ftlm_ed_n_vs_mu")


def run(exe, args, cwd):
    return subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_help(tmp_path):
    for exe in (EXE, ED):
        p = run(exe, ["--help"], str(tmp_path))
        assert p.returncode == 0
        assert p.stdout.startswith("Usage: ")
        for flag in ["--lx N", "--beta-list LIST", "--exact-block-threshold N", "--checkpoint PATH",
                     "--trace-partition-csv PATH", "--debug-block-csv PATH", "--output PATH"]:
            assert flag in p.stdout


@pytest.mark.parametrize("args,msg", [
    (["--lx", "9", "--ly", "8"], "Error: This implementation supports at most 63 sites."),
    (["--bogus"], "Error: Unknown argument: --bogus"),
    (["--mu-count", "1"], "Error: --mu-count must be at least 2."),
    (["--vector-placement", "2"], "Error: --vector-placement must be 0 or 3 to 16."),
    (["--checkpoint", "ck.bin", "--trace-partition-csv", "t.csv"],
     "Error: --checkpoint is not currently compatible with trace/debug block output."),
])
def test_errors_exit_1(tmp_path, args, msg):
    p = run(EXE, args, str(tmp_path))
    assert p.returncode == 1
    assert p.stderr.strip() == msg


def test_meta_mismatch(tmp_path):
    ck = tmp_path / "ck.bin"
    (tmp_path / "ck.bin.meta").write_text("lx=2\nly=2\n")
    p = run(EXE, ["--checkpoint", str(ck)], str(tmp_path))
    assert p.returncode == 1
    assert p.stderr.strip() == f"Error: Checkpoint metadata does not match current run: {ck}.meta"


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible here")
def test_no_gpu_fails_loudly(tmp_path):
    for exe in (EXE, ED):
        p = run(exe, ["--output", str(tmp_path / "o.csv")], str(tmp_path))
        assert p.returncode == 1
        assert "GPU engine unavailable" in p.stderr
        assert not (tmp_path / "o.csv").exists()
