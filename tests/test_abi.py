"""The C-ABI library loads and exports every function include/bzg.h
declares (no GPU calls).  CPU only."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "bzg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bzg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import bzip2gpu_amd
    so = bzip2gpu_amd.library_path()
    if not os.path.exists(so):
        pytest.fail("libbzg.so not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(so)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_bounds_without_gpu():
    import bzip2gpu_amd
    assert "gfx950" in bzip2gpu_amd.version()
    assert bzip2gpu_amd.compress_bound(0) >= 14
    assert bzip2gpu_amd.compress_bound(1 << 20) > (1 << 20)
    assert bzip2gpu_amd.lib().bzg_rle1_bound(4) >= 5


def test_compress_fails_loudly_without_device():
    import torch
    import bzip2gpu_amd
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(bzip2gpu_amd.BzgError):
        bzip2gpu_amd.compress(b"hello", 9)


def test_drop_in_headers_parse():
    # the drop-in C++ headers keep the reference's names
    txt = open(os.path.join(ROOT, "include", "compression.h")).read()
    for name in ["rle1_compress", "rle2_compress", "make_symbols_table", "fmtf", "huffman_build_trees",
                 "fbwt", "bzip2_gpu_compress", "huffman_encode"]:
        assert re.search(r"\b%s\s*\(" % name, txt), name
    io = open(os.path.join(ROOT, "include", "io.h")).read()
    assert "BZFileInputStream" in io and "fetch" in io
