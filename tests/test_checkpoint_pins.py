"""Pins of the checkpoint CRC-64 and the variables file (the oracle's and
the product host's restatements of crc64.c:153-256, norms.c:440-458 and
checkpoint_save.c:412-468).

- The published check value of this CRC (reflected, ECMA-182 polynomial,
  initial value and final XOR all ones: CRC-64/XZ): CRC("123456789") =
  0x995DC9BBDF1939FA.
- A bit-at-a-time shift register in Python, written from the polynomial
  alone, equals both table-driven restatements on random buffers.
- Combining: CRC(A||B) = combine(CRC(A), CRC(B), |B|) for random splits,
  including empty pieces.
- The "vars" CRC of the oracle equals the bitwise CRC of the byte stream
  built in numpy: elements in order, each with phi, pi, cx, cy, cz (the
  reference's variable order) as stored doubles.
- variables.bin of the oracle, parsed by a reader written from the format:
  the header, element names "<patch>_<child digits>" in element order, np,
  and per variable its list index and np little-endian doubles.
"""
import os

import numpy as np
import pytest

from nmesh_amd import checkpoint as ck
from oracle.orc import OMesh, lib

POLY = 0xC96C5795D7870F42          # 0x42F0E1EBA9EA3693 bit-reversed


def crc_bitwise(data, crc=0):
    """shift register, one bit at a time (no table)"""
    crc ^= 0xFFFFFFFFFFFFFFFF
    for b in data:
        crc ^= b
        for _ in range(8):
            crc = (crc >> 1) ^ POLY if crc & 1 else crc >> 1
    return crc ^ 0xFFFFFFFFFFFFFFFF


def orc_crc(data, crc=0):
    return int(lib().orc_crc64(crc, data, len(data)))


def test_check_value():
    assert crc_bitwise(b"123456789") == 0x995DC9BBDF1939FA
    assert orc_crc(b"123456789") == 0x995DC9BBDF1939FA
    assert ck.crc64(b"123456789") == 0x995DC9BBDF1939FA
    assert orc_crc(b"") == 0 and ck.crc64(b"") == 0


def test_table_crc_equals_bitwise_and_continues():
    rng = np.random.default_rng(1)
    for n in (1, 7, 8, 9, 64, 333):
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        ref = crc_bitwise(data)
        assert orc_crc(data) == ref and ck.crc64(data) == ref
        # continuing from a previous CRC is the CRC of the concatenation
        k = n // 3
        assert ck.crc64(data[k:], ck.crc64(data[:k])) == ref
        assert orc_crc(data[k:], orc_crc(data[:k])) == ref


def test_combine():
    rng = np.random.default_rng(2)
    for _ in range(20):
        a = rng.integers(0, 256, int(rng.integers(0, 300)), dtype=np.uint8).tobytes()
        b = rng.integers(0, 256, int(rng.integers(0, 300)), dtype=np.uint8).tobytes()
        ref = crc_bitwise(a + b)
        assert ck.crc64_combine(ck.crc64(a), ck.crc64(b), len(b)) == ref
        assert int(lib().orc_crc64_combine(orc_crc(a), orc_crc(b), len(b))) == ref


def stream(om, u):
    """the vars byte stream: per element phi, pi, cx, cy, cz"""
    u = u.reshape(5, -1)
    parts = []
    for e in range(om.nelem):
        info = om.elm_info(e)
        n, off = int(info[2]), int(info[3])
        for v in (4, 0, 1, 2, 3):
            parts.append(u[v, off:off + n ** 3].astype("<f8").tobytes())
    return b"".join(parts)


@pytest.mark.parametrize("kind,luni,n", [("box", 1, 3), ("shell", 0, 2)])
def test_oracle_vars_crc_is_the_crc_of_the_element_variable_stream(kind, luni, n):
    om = OMesh(kind, luni, n, dout=2.0, rin=2.0, rout=6.0)
    u = np.random.default_rng(3).standard_normal(5 * om.nnodes)
    assert om.vars_crc(u) == crc_bitwise(stream(om, u))


def parse_variables_bin(path):
    """a reader written from the format of checkpoint_write_vl"""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    assert line() == "number of variables and variable list in this file:"
    nv = int(line())
    names = [line() for _ in range(nv)]
    assert line() == "" and line() == ""
    assert line() == "variable-list-data:" and line() == ""
    elms = []
    while True:
        s = line()
        if s == "EndOfFile.":
            break
        assert s == "{"
        name = line()
        np_ = int(line())
        vals = {}
        for _ in range(nv):
            vli = int(line())
            vals[names[vli]] = np.frombuffer(data[pos:pos + 8 * np_], dtype="<f8")
            pos += 8 * np_
            assert data[pos:pos + 1] == b"\n"
            pos += 1
        assert line() == "}" and line() == ""
        elms.append((name, np_, vals))
    assert pos == len(data)
    return names, elms


def test_oracle_variables_bin_format(tmp_path):
    om = OMesh("box", 1, 3, dout=2.0)
    u = np.random.default_rng(4).standard_normal(5 * om.nnodes)
    path = os.path.join(tmp_path, "variables.bin")
    om.write_variables_bin(u, path)
    names, elms = parse_variables_bin(path)
    assert names == ["scalarwave1_phi", "scalarwave1_pi", "scalarwave1_cx", "scalarwave1_cy",
                     "scalarwave1_cz"]
    # one uniform refinement of patch 0: children 0..7 in Morton order
    assert [e[0] for e in elms] == ["0_%d" % c for c in range(8)]
    uu = u.reshape(5, -1)
    slot = {"scalarwave1_pi": 0, "scalarwave1_cx": 1, "scalarwave1_cy": 2, "scalarwave1_cz": 3,
            "scalarwave1_phi": 4}
    for e, (name, np_, vals) in enumerate(elms):
        off = int(om.elm_info(e)[3])
        assert np_ == 27
        for k, v in vals.items():
            assert np.array_equal(v, uu[slot[k], off:off + np_])
