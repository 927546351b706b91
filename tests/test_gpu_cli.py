"""bzip2gpu CLI on the device: compression output equals the library's,
-d/-c/-t/stdin paths, bunzip/bzcat program names, and bzip2's exit codes
(2 for corrupt input, 1 for a missing file)."""
import bz2
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bzip2gpu_amd", "bzip2gpu")


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _run(args, data=None, prog=BIN):
    return subprocess.run([prog, *args], input=data, capture_output=True, timeout=120)


@pytest.fixture()
def payload():
    import bzip2gpu_amd.datagen as dg
    return dg.mix(2_500_000, 7).tobytes()


def test_compress_file_then_decompress(dev, tmp_path, payload):
    import bzip2gpu_amd as bz
    src = tmp_path / "data.bin"
    src.write_bytes(payload)
    _case("cli compress file -v")
    r = _run(["-v", "-6", str(src)])
    assert r.returncode == 0, r.stderr
    assert b"saved" in r.stderr
    comp = (tmp_path / "data.bin.bz2").read_bytes()
    assert comp == bz.compress(payload, 6)
    assert src.exists()  # inputs are kept
    os.remove(src)
    _case("cli decompress file")
    r = _run(["-d", str(tmp_path / "data.bin.bz2")])
    assert r.returncode == 0, r.stderr
    assert src.read_bytes() == payload
    _case("cli decompress to stdout")
    r = _run(["-dc", str(tmp_path / "data.bin.bz2")])
    assert r.returncode == 0 and r.stdout == payload


def test_stdin_stdout_and_names(dev, tmp_path, payload):
    _case("cli stdin compress")
    r = _run([], data=payload)
    assert r.returncode == 0
    import bzip2gpu_amd as bz
    assert bz.decompress(r.stdout) == payload
    _case("cli stdin decompress")
    r2 = _run(["-d"], data=r.stdout)
    assert r2.returncode == 0 and r2.stdout == payload
    stock = bz2.compress(payload, 9)
    (tmp_path / "s.tbz2").write_bytes(stock)
    _case("cli .tbz2 name")
    assert _run(["-d", str(tmp_path / "s.tbz2")]).returncode == 0
    assert (tmp_path / "s.tar").read_bytes() == payload
    (tmp_path / "odd").write_bytes(stock)
    _case("cli unknown suffix")
    assert _run(["-d", "-q", str(tmp_path / "odd")]).returncode == 0
    assert (tmp_path / "odd.out").read_bytes() == payload
    for name, args in (("bunzip2gpu", ["-c"]), ("bzcatgpu", [])):
        link = tmp_path / name
        os.symlink(BIN, link)
        _case(f"cli program name {name}")
        r3 = _run([*args, str(tmp_path / "odd")], prog=str(link))
        assert r3.returncode == 0 and r3.stdout == payload


def test_test_mode_and_exit_codes(dev, tmp_path, payload):
    good = bz2.compress(payload, 9)
    (tmp_path / "good.bz2").write_bytes(good)
    bad = bytearray(good)
    bad[len(bad) // 2] ^= 0x10
    (tmp_path / "bad.bz2").write_bytes(bytes(bad))
    _case("cli -t good")
    r = _run(["-tv", str(tmp_path / "good.bz2")])
    assert r.returncode == 0 and b"ok" in r.stderr
    assert not (tmp_path / "good").exists()
    _case("cli -t corrupt")
    r = _run(["-t", str(tmp_path / "good.bz2"), str(tmp_path / "bad.bz2")])
    assert r.returncode == 2
    _case("cli missing file")
    r = _run(["-d", str(tmp_path / "missing.bz2")])
    assert r.returncode == 1
    _case("cli not bzip2")
    (tmp_path / "plain.bz2").write_bytes(b"hello")
    assert _run(["-d", str(tmp_path / "plain.bz2")]).returncode == 2
