"""Binary system files in the reference's fixture format
(benchmarks/matrix_analysis/ilu_fill.cpp:20-44): layout, round trip and
rejection of malformed files (host code only)."""
import numpy as np
import pytest

import owtk


def c0_csr():
    s = owtk.generate_split(64, 32, 3, 0, ordering=owtk.ORDER_RCM)
    o, c, v = s.csr()
    return s.nodes, o, c, v


def test_layout_matches_reference_format(tmp_path):
    n, o, c, v = c0_csr()
    path = tmp_path / "system.bin"
    owtk.write_block_csr(path, n, 3, o, c, v)
    raw = path.read_bytes()
    header = np.frombuffer(raw[:24], np.uint64)
    assert header.tolist() == [n, 3, o[-1]]
    pos = 24
    offs = np.frombuffer(raw[pos:pos + 8 * (n + 1)], np.uint64)
    pos += 8 * (n + 1)
    cols = np.frombuffer(raw[pos:pos + 8 * o[-1]], np.uint64)
    pos += 8 * o[-1]
    vals = np.frombuffer(raw[pos:], np.float64)
    assert np.array_equal(offs, o) and np.array_equal(cols, c) and np.array_equal(vals, v)


def test_round_trip(tmp_path):
    n, o, c, v = c0_csr()
    path = tmp_path / "system.bin"
    owtk.write_block_csr(path, n, 3, o, c, v)
    n2, b2, o2, c2, v2 = owtk.read_block_csr(path)
    assert (n2, b2) == (n, 3)
    assert np.array_equal(o2, o) and np.array_equal(c2, c) and np.array_equal(v2, v)


def test_numpy_written_file_reads(tmp_path):
    # written independently of the library, as the reference's tools do
    n, o, c, v = c0_csr()
    path = tmp_path / "np.bin"
    with open(path, "wb") as f:
        np.array([n, 3, o[-1]], np.uint64).tofile(f)
        o.astype(np.uint64).tofile(f)
        c.astype(np.uint64).tofile(f)
        v.astype(np.float64).tofile(f)
    n2, b2, o2, c2, v2 = owtk.read_block_csr(path)
    assert np.array_equal(o2, o) and np.array_equal(v2, v)


@pytest.mark.parametrize("damage", ["truncate", "header", "column", "offsets"])
def test_malformed_files_rejected(tmp_path, damage):
    n, o, c, v = c0_csr()
    path = tmp_path / "bad.bin"
    owtk.write_block_csr(path, n, 3, o, c, v)
    raw = bytearray(path.read_bytes())
    if damage == "truncate":
        raw = raw[:-8]
    elif damage == "header":
        raw[16:24] = np.array([o[-1] + 5], np.uint64).tobytes()
    elif damage == "column":
        pos = 24 + 8 * (n + 1)
        raw[pos:pos + 8] = np.array([n + 7], np.uint64).tobytes()
    else:
        raw[24 + 8:24 + 16] = np.array([10 ** 9], np.uint64).tobytes()
    path.write_bytes(bytes(raw))
    with pytest.raises(owtk.InvalidArgumentError):
        owtk.read_block_csr(path)
