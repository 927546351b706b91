"""The reference's own test cases, driven through the drop-in C++ API.

tests/refapi/refapi_test (built by __graft_entry__.build()) regenerates the
inputs of the reference's gtests (srand(2137); tests/compression/rle1.cu,
fbwt.cu, fmtf.cu, rle2.cu, huffman.cu and tests/io.cpp) and calls
rle1_compress, fbwt, make_symbols_table, fmtf, rle2_compress,
huffman_build_trees, huffman_encode, BZFileInputStream and
bzip2_gpu_compress exactly as those tests do, with every output compared
with the CPU oracle.  The binary runs once per session; each of its cases is
reported as a test of its own.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "refapi", "refapi_test")

CASES = [
    "io.file_input_stream",
    "io.read_error_partway",
    "compress.rle1_normal",
    "compress.rle1_incompressible",
    "compress.rle1_long_runs",
    "compress.fbwt_normal",
    "compress.fbwt_small_alphabet",
    "compress.fbwt_single_character",
    "compress.mtf_symbol_table_generation",
    "compress.fmtf_normal",
    "compress.fmtf_small_alphabet",
    "compress.rle2_normal",
    "compress.rle2_incompressible",
    "compress.rle2_long_runs",
    "compress.huffman_count_frequencies",
    "compress.huffman_tree_builder_stability",
    "compress.huffman_encoding",
    "compress.huffman_encode_rejects_out_of_range",
    "compress.stream_empty",
    "compress.stream_via_ring",
    "compress.stream_via_ring_segmented",
]


def test_refapi_binary_lists_every_case():
    """CPU: the binary is built and holds exactly the cases listed here."""
    if not os.path.exists(BIN):
        pytest.fail("tests/refapi/refapi_test is not built; run __graft_entry__.build()")
    out = subprocess.run([BIN, "--list"], capture_output=True, text=True, timeout=60, check=True).stdout
    assert out.split() == CASES


@pytest.fixture(scope="module")
def refapi_results(dev):
    if not os.path.exists(BIN):
        pytest.fail("tests/refapi/refapi_test is not built; run __graft_entry__.build()")
    print("running", BIN, flush=True)
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=600)
    res = {}
    for line in p.stdout.splitlines():
        if line.startswith("[ OK   ] "):
            res[line[9:].strip()] = (True, "")
        elif line.startswith("[ FAIL ] "):
            name, _, msg = line[9:].partition(": ")
            res[name.strip()] = (False, msg)
    return res, p.returncode, p.stdout[-2000:] + p.stderr[-2000:]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_reference_case_through_drop_in(refapi_results, case):
    res, rc, tail = refapi_results
    assert case in res, f"case did not finish (exit {rc}):\n{tail}"
    ok, msg = res[case]
    assert ok, msg
