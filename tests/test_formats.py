"""On-disk formats around the path (include/minicore/coreset/coreset.h:94-136,267-338).

The byte layouts are restated here with struct from the reference's write/read
code and checked both ways against libmcg's streams: what the library writes,
decompressed, must be exactly those bytes, and bytes built here (gzip or plain,
as the reference's gzFile and FILE* writers produce) must read back.  These
entry points are host-only, so the CPU suite runs them; the pyminicore sampler
round trip, which needs sensitivities from the GPU, is marked gpu.
"""
import gzip
import struct

import numpy as np
import pytest

from minicore_amd import _lib


def sampler_bytes(probs, seed, weights):
    """coreset.h:282-293: u64 n, u64 seed, f32 probs[n], u32 flag (137 | 0), f32 weights[n]."""
    n = len(probs)
    b = struct.pack("<QQ", n, seed) + np.asarray(probs, "<f4").tobytes()
    b += struct.pack("<I", 137 if weights is not None else 0)
    if weights is not None:
        b += np.asarray(weights, "<f4").tobytes()
    return b


def index_coreset_bytes(idx, w):
    """coreset.h:113-123: u64 m, IT indices[m], FT weights[m]."""
    return struct.pack("<Q", len(idx)) + idx.tobytes() + w.tobytes()


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


@pytest.mark.parametrize("weighted", [False, True])
def test_sampler_write_layout(lib, tmp_path, weighted):
    rng = np.random.default_rng(5)
    p = rng.random(37).astype(np.float32)
    w = rng.random(37).astype(np.float32) if weighted else None
    path = tmp_path / "s.gz"
    _lib.sampler_write(path, p, 0xDEADBEEF12345, w)
    with open(path, "rb") as f:
        assert f.read(2) == b"\x1f\x8b"  # a gzip stream
    assert gzip.decompress(path.read_bytes()) == sampler_bytes(p, 0xDEADBEEF12345, w)


@pytest.mark.parametrize("compressed", [True, False])
@pytest.mark.parametrize("weighted", [False, True])
def test_sampler_read_reference_bytes(lib, tmp_path, compressed, weighted):
    rng = np.random.default_rng(6)
    p = rng.random(50).astype(np.float32)
    w = rng.random(50).astype(np.float32) if weighted else None
    raw = sampler_bytes(p, 13, w)
    path = tmp_path / "s.bin"
    path.write_bytes(gzip.compress(raw) if compressed else raw)
    got_p, seed, got_w = _lib.sampler_read(path)
    assert seed == 13
    assert np.array_equal(got_p, p)
    if weighted:
        assert np.array_equal(got_w, w)
    else:
        assert got_w is None


def test_sampler_any_nonzero_flag_means_weights(lib, tmp_path):
    # the reference asserts flag == 137 only in debug builds and reads weights for any nonzero flag
    p = np.arange(4, dtype=np.float32)
    raw = struct.pack("<QQ", 4, 1) + p.tobytes() + struct.pack("<I", 5) + (p + 1).tobytes()
    path = tmp_path / "s.gz"
    path.write_bytes(gzip.compress(raw))
    _, _, w = _lib.sampler_read(path)
    assert np.array_equal(w, p + 1)


def test_sampler_empty(lib, tmp_path):
    path = tmp_path / "e.gz"
    _lib.sampler_write(path, np.zeros(0, np.float32), 7)
    p, seed, w = _lib.sampler_read(path)
    assert len(p) == 0 and seed == 7 and w is None
    assert gzip.decompress(path.read_bytes()) == sampler_bytes([], 7, None)


def test_sampler_truncated_and_missing(lib, tmp_path):
    p = np.ones(10, np.float32)
    raw = sampler_bytes(p, 3, p)
    path = tmp_path / "t.gz"
    path.write_bytes(gzip.compress(raw[:-7]))
    with pytest.raises(_lib.MCGError) as e:
        _lib.sampler_read(path)
    assert e.value.code == -6 and "read" in str(e.value)
    with pytest.raises(_lib.MCGError) as e:
        _lib.sampler_read(tmp_path / "absent.gz")
    assert e.value.code == -6
    with pytest.raises(_lib.MCGError) as e:
        _lib.sampler_write(tmp_path / "no_dir" / "x.gz", p, 1)
    assert e.value.code == -6


@pytest.mark.parametrize("it,ft", [(np.uint32, np.float32), (np.uint64, np.float64), (np.uint16, np.float32),
                                   (np.uint64, np.float32)])
def test_index_coreset_roundtrip_and_layout(lib, tmp_path, it, ft):
    rng = np.random.default_rng(7)
    idx = rng.integers(0, 60000, 23).astype(it)
    w = rng.random(23).astype(ft)
    path = tmp_path / "ic.gz"
    _lib.index_coreset_write(path, idx, w, it, ft)
    assert gzip.decompress(path.read_bytes()) == index_coreset_bytes(idx, w)
    gi, gw = _lib.index_coreset_read(path, it, ft)
    assert gi.dtype == it and gw.dtype == ft
    assert np.array_equal(gi, idx) and np.array_equal(gw, w)
    # plain (FILE*-written) stream of the same layout
    path.write_bytes(index_coreset_bytes(idx, w))
    gi, gw = _lib.index_coreset_read(path, it, ft)
    assert np.array_equal(gi, idx) and np.array_equal(gw, w)


def test_index_coreset_bad_dtype_and_truncation(lib, tmp_path):
    path = tmp_path / "ic.gz"
    with pytest.raises(TypeError):
        _lib.index_coreset_write(path, np.zeros(2, np.uint32), np.zeros(2, np.float16), np.uint32, np.float16)
    with pytest.raises(_lib.MCGError) as e:
        _lib.index_coreset_write(path, np.zeros(2, np.int32), np.zeros(2, np.float32), np.int32, np.float32)
    assert e.value.code == -1
    raw = index_coreset_bytes(np.arange(5, dtype=np.uint32), np.ones(5, np.float32))
    path.write_bytes(gzip.compress(raw[:-1]))
    with pytest.raises(_lib.MCGError) as e:
        _lib.index_coreset_read(path)
    assert e.value.code == -6


@pytest.mark.gpu
def test_pyminicore_sampler_write_read_sample(tmp_path):
    """make_sampler -> write -> CoresetSampler(path): the same probabilities, and the
    same draws for the same seed (the stream carries probs, seed and weights)."""
    import minicore_amd as mc
    rng = np.random.default_rng(8)
    n, k = 2000, 7
    costs = rng.random(n) * 3
    asn = rng.integers(0, k, n).astype(np.uint32)
    w = (rng.random(n) + .5).astype(np.float32)
    for weights in (None, w):
        cs = mc.CoresetSampler()
        cs.make_sampler(k, costs, asn, weights=weights, seed=99, sens=4)
        path = str(tmp_path / "cs.gz")
        cs.write(path)
        cs2 = mc.CoresetSampler(path)
        assert np.array_equal(cs.get_probs(), cs2.get_probs())
        a = cs.sample(300, seed=17)
        b = cs2.sample(300, seed=17)
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])
        p, seed, ww = _lib.sampler_read(path)
        assert seed == 99
        assert (ww is None) == (weights is None)
    fl = mc.CoresetSampler()
    fl.make_sampler(k, costs, asn, sens=1)
    with pytest.raises(RuntimeError, match="FL"):
        fl.write(str(tmp_path / "fl.gz"))
