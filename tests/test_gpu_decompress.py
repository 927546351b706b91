"""Device decompression: byte-identical to stock libbz2 (Python's bz2) on
stock streams, and the exact inverse of this library's compressor on its own
streams, including the reference-format blocks stock bzip2 refuses (RLE1
longer than 100000*level) and runs longer than 255 bytes."""
import bz2
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _same(a: bytes, b: bytes, what: str) -> None:
    if a == b:
        return
    n = min(len(a), len(b))
    first = next((i for i in range(n) if a[i] != b[i]), n)
    raise AssertionError(f"{what}: lengths {len(a)} vs {len(b)}, first difference at byte {first}")


def _inputs():
    import bzip2gpu_amd.datagen as dg
    rng = np.random.default_rng(23)
    runs = b"".join(bytes([int(rng.integers(0, 256))]) * int(rng.integers(1, 700)) for _ in range(3000))
    return [
        ("one", b"x"),
        ("hello", b"hello world!\n"),
        ("four_equal", b"aaaa"),
        ("run_259_260_261", b"q" * 259 + b"r" * 260 + b"s" * 261 + b"t" * 4 + b"u" * 5),
        ("zeros900k", b"\x00" * 900000),
        ("period2", b"ab" * 300000),
        ("random_899k", rng.integers(0, 256, 899000, dtype=np.uint8).tobytes()),
        ("runs", runs),
        ("text1_1MiB", dg.workload_text1().tobytes()),
        ("mix_3MiB", dg.mix(3 << 20, 5).tobytes()),
        ("periodic_p7", dg.periodic(950000, 8, 7, 0.001).tobytes()),
        ("periodic_p1000", dg.periodic(700000, 9, 1000, 0.0).tobytes()),
    ]


def test_empty_inputs(dev):
    import bzip2gpu_amd as bz
    _case("decompress empty")
    assert bz.decompress(b"") == b""
    _case("decompress empty stream")
    assert bz.decompress(bz2.compress(b"")) == b""


@pytest.mark.parametrize("level", [1, 9])
def test_stock_streams(dev, level):
    """Streams written by stock libbz2 decode to bz2.decompress's bytes."""
    import bzip2gpu_amd as bz
    for name, data in _inputs():
        c = bz2.compress(data, level)
        _case(f"stock level={level} {name} raw={len(data)} bz2={len(c)}")
        _same(bz.decompress(c), bz2.decompress(c), f"stock {name}")


def test_own_streams_round_trip(dev):
    """decompress(compress(x)) == x for this library's streams: the
    reference's fixed 900000-byte cut at levels 1 and 9 (blocks stock bzip2
    refuses at level 1) and stock-sized blocks."""
    import bzip2gpu_amd as bz
    for name, data in _inputs():
        for level, block in ((9, 900000), (1, 900000), (3, 300000)):
            c = bz.compress(data, level, block)
            _case(f"own level={level} block={block} {name} raw={len(data)} bz2={len(c)}")
            _same(bz.decompress(c), data, f"own {name} level {level}")


def test_multi_stream_and_trailing_bytes(dev):
    import bzip2gpu_amd as bz
    a, b = b"first stream " * 1000, bytes(range(256)) * 50
    ca, cb = bz2.compress(a, 9), bz.compress(b, 5, 500000)
    _case("two streams")
    assert bz.decompress(ca + cb) == a + b
    _case("three streams incl. an empty one")
    assert bz.decompress(ca + bz2.compress(b"") + ca) == a + a
    for tail in (b"x", b"\x00\x00\x00\x00garbage", b"BZ!!"):
        _case(f"trailing {tail!r}")
        assert bz.decompress(ca + tail) == bz2.decompress(ca + tail) == a
    for tail in (b"B", b"BZ", b"BZh", b"BZh9"):
        _case(f"truncated next stream {tail!r}")
        with pytest.raises(bz.BzgDataError):
            bz.decompress(ca + tail)
        with pytest.raises((EOFError, ValueError)):
            bz2.decompress(ca + tail)


def test_corrupt_and_truncated(dev):
    """Bit flips and truncation: where stock bzip2 rejects the stream, so does
    the device decoder (CRC, header and code checks); where stock accepts it
    (padding bits), both give the same bytes."""
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = dg.mix(1 << 20, 8).tobytes()
    c = bz2.compress(data, 9)
    rng = np.random.default_rng(3)
    positions = [4, 5, 9, 12, 20, 40, 100, 300] + [int(x) for x in rng.integers(40, len(c), 24)] + [len(c) - 1]
    for pos in positions:
        for bit in (0, 5):
            bad = bytearray(c)
            bad[pos] ^= 1 << bit
            bad = bytes(bad)
            _case(f"flip byte {pos} bit {bit}")
            try:
                want = bz2.decompress(bad)
            except (OSError, EOFError, ValueError):
                want = None
            if want is None:
                with pytest.raises(bz.BzgDataError):
                    bz.decompress(bad)
            else:
                assert bz.decompress(bad) == want
    for cut in (3, 10, 60, len(c) // 2, len(c) - 11, len(c) - 1):
        _case(f"truncate at {cut}")
        with pytest.raises(bz.BzgDataError):
            bz.decompress(c[:cut])
    _case("not bzip2")
    with pytest.raises(bz.BzgDataError):
        bz.decompress(b"PK\x03\x04 not a bzip2 file")


def test_symbol_slot_retry(dev):
    """Symbol slots are sized from the distance to the next magic candidate;
    a block that overflows them (a false candidate inside it) is decoded
    again with full slots.  BZG_DEC_SLAB_MAX shrinks every slot to force
    that path."""
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = dg.mix(2 << 20, 4).tobytes()
    c = bz2.compress(data, 9)
    os.environ["BZG_DEC_SLAB_MAX"] = "1000"
    try:
        _case("forced slot overflow")
        _same(bz.decompress(c), data, "retry path")
    finally:
        del os.environ["BZG_DEC_SLAB_MAX"]


def test_device_input_large(dev):
    """A 64 MiB mix stream from the batched compressor, decoded from device
    memory; the output is copied back and compared."""
    import torch

    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = dg.mix(64 << 20, 2)
    c = bz.compress(data.tobytes(), 9)
    d_in = torch.from_numpy(np.frombuffer(c, dtype=np.uint8).copy()).to(dev)
    _case(f"device decompress 64 MiB mix bz2={len(c)}")
    ptr, n = bz.decompress_device(d_in)
    assert n == data.size
    ms, nb, ns = bz.last_decode_times()
    assert nb == (data.size + 899999) // 900000 and ns == 1
    out = bz.decompress(c)
    _same(out, data.tobytes(), "64 MiB mix")


def test_many_streams_small_tables(dev):
    """More blocks than two per CU (here 700 streams, 1400 candidates): the
    decoder takes its 9-bit lookup tables (codes of 10-20 bits through the
    length limits).  Stock streams and this library's, levels 1-9."""
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    rng = np.random.default_rng(31)
    parts, streams = [], []
    for i in range(700):
        p = dg.mix(int(rng.integers(200, 6000)), 300 + i).tobytes()
        lv = int(rng.integers(1, 10))
        parts.append(p)
        streams.append(bz2.compress(p, lv) if i % 2 else bz.compress(p, lv))
    _case("decompress 700 concatenated streams")
    _same(bz.decompress(b"".join(streams)), b"".join(parts), "700 streams")
