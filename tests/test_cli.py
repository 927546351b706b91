"""bzip2gpu CLI paths that exit before any device work: help, version and
unknown options (reference: src/utils/argsparce.cpp:6-101)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bzip2gpu_amd", "bzip2gpu")


def _run(*args):
    if not os.path.exists(BIN):
        pytest.skip("bzip2gpu is not built")
    return subprocess.run([BIN, *args], capture_output=True, timeout=60)


def test_help_lists_bzip2_flags():
    r = _run("-h")
    assert r.returncode == 0
    text = r.stdout.decode()
    for flag in ("--decompress", "--compress", "--keep", "--force", "--test", "--stdout", "--quiet",
                 "--verbose", "--license", "--version", "--small", "--fast", "--best"):
        assert flag in text


@pytest.mark.parametrize("flag", ["-V", "-L", "--version", "--license"])
def test_version(flag):
    r = _run(flag)
    assert r.returncode == 0
    assert b"bzip2gpu_amd" in r.stdout


def test_unknown_option_fails():
    r = _run("-x")
    assert r.returncode == 1
    assert b"UNKNOWN OPTION: -x" in r.stdout
