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


# --------------------------------------------------------------- state files --
# The cases of the reference's own check of read_state
# (benchmarks/test_specwave_limon.py:38-66), on files written here the same way.
def _state(path, ids, values, width=8, ns=1, nd=2):
    import struct
    with open(path, "wb") as f:
        f.write(b"SWSTATE1" + struct.pack("=5Q", width, len(ids), ns, nd, 2))
        np.asarray(ids, dtype="=u8").tofile(f)
        np.asarray(values, dtype=f"=f{width}").tofile(f)
    return path


def test_state_reordered_owned_spectrum(tmp_path):
    first = _state(tmp_path / "rank0.bin", [1], [3, 4])
    second = _state(tmp_path / "rank1.bin", [0], [1, 2])
    assert np.array_equal(owtk.read_state([first, second], 1, 2), [[1, 2], [3, 4]])


def test_state_float_spectrum(tmp_path):
    single = tmp_path / "single.bin"
    import struct
    with open(single, "wb") as f:
        f.write(b"SWSTATE1" + struct.pack("=5Q", 4, 2, 1, 2, 2))
        np.asarray([0, 1], dtype="=u8").tofile(f)
        np.asarray([1, 2, 3, 4], dtype="=f4").tofile(f)
    assert np.array_equal(owtk.read_state([single], 1, 2), [[1, 2], [3, 4]])


@pytest.mark.parametrize("case", ["missing rank", "duplicate global IDs", "wrong dimensions", "nan", "infinity",
                                  "negative", "truncated-header", "truncated-data", "extra-data", "wrong-magic"])
def test_state_rejected(tmp_path, case):
    first = _state(tmp_path / "rank0.bin", [1], [3, 4])
    second = _state(tmp_path / "rank1.bin", [0], [1, 2])
    if case == "missing rank":
        files, dims = [first], (1, 2)
    elif case == "duplicate global IDs":
        files, dims = [first, first], (1, 2)
    elif case == "wrong dimensions":
        files, dims = [first, second], (2, 2)
    elif case in ("nan", "infinity", "negative"):
        bad = {"nan": [1, np.nan], "infinity": [1, np.inf], "negative": [1, -1]}[case]
        files, dims = [first, _state(tmp_path / "bad.bin", [0], bad)], (1, 2)
    else:
        data = {"truncated-header": b"SWSTATE1", "truncated-data": first.read_bytes()[:-1],
                "extra-data": first.read_bytes() + b"X", "wrong-magic": b"invalid!"}[case]
        path = tmp_path / "bad.bin"
        path.write_bytes(data)
        files, dims = [path], (1, 2)
    with pytest.raises(owtk.OwtkError):
        owtk.read_state(files, *dims)
