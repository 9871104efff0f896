"""SciDAC single-file output (cmx_field_write_scidac / cmx_field_read_scidac), host only.

QIO, LIME and QDP++ are not in the reference tree and no SciDAC file is among its fixtures, so the
writer is checked against the published format by an independent reader written here with `struct`
and numpy: the LIME record sequence and flags, the private XML records, the binary data decoded as
big-endian numbers in lexicographic site order with QDP++'s element order ([sink spin][source spin]
[sink colour][source colour] for a propagator), and the checksum recomputed with zlib's CRC-32 and the
SciDAC rotations (index mod 29, mod 31).  Byte compatibility with the QIO library itself is unpinned.
"""
import re
import struct
import zlib

import numpy as np
import pytest

from chroma_amd import lib as cmx

L = (4, 4, 2, 6)
V = int(np.prod(L))


def lex_to_host(L):
    """host index (parity * V/2 + lex / 2) of every lexicographic site (x fastest)"""
    V = int(np.prod(L))
    lex = np.arange(V)
    x = lex % L[0]
    y = (lex // L[0]) % L[1]
    z = (lex // (L[0] * L[1])) % L[2]
    t = lex // (L[0] * L[1] * L[2])
    par = (x + y + z + t) % 2
    return par * (V // 2) + lex // 2


def parse_lime(buf):
    recs, pos = [], 0
    while pos < len(buf):
        magic, version, flags, length = struct.unpack(">IHHQ", buf[pos:pos + 16])
        assert magic == 0x456789AB and version == 1
        rtype = buf[pos + 16:pos + 144].rstrip(b"\0").decode()
        data = buf[pos + 144:pos + 144 + length]
        padded = (length + 7) // 8 * 8
        assert buf[pos + 144 + length:pos + 144 + padded] == b"\0" * (padded - length)
        recs.append((rtype, bool(flags & 0x8000), bool(flags & 0x4000), data))
        pos += 144 + padded
    assert pos == len(buf)
    return recs


def tag(xml, name):
    m = re.search(f"<{name}>(.*?)</{name}>", xml)
    return m.group(1) if m else None


def random_field(kind, seed):
    rng = np.random.default_rng(seed)
    shape = (V, 12, 12) if kind == cmx.SCIDAC_PROPAGATOR else (V, 4, 3)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("kind", [cmx.SCIDAC_PROPAGATOR, cmx.SCIDAC_FERMION])
@pytest.mark.parametrize("prec", [cmx.DOUBLE, cmx.SINGLE])
def test_scidac_file_layout(tmp_path, kind, prec):
    f = random_field(kind, 3 + kind + 7 * prec)
    path = tmp_path / "field.lime"
    cmx.field_write_scidac(path, L, f, kind=kind, prec=prec, file_xml="<propagator>file</propagator>",
                           record_xml="<Propagator>record</Propagator>")
    recs = parse_lime(path.read_bytes())
    assert [(r[0], r[1], r[2]) for r in recs] == [
        ("scidac-private-file-xml", True, False), ("scidac-file-xml", False, True),
        ("scidac-private-record-xml", True, False), ("scidac-record-xml", False, False),
        ("scidac-binary-data", False, False), ("scidac-checksum", False, True)]
    pf, fx, pr, rx, data, ck = [r[3] for r in recs]
    pf, pr, ck = pf.decode(), pr.decode(), ck.decode()
    assert fx == b"<propagator>file</propagator>" and rx == b"<Propagator>record</Propagator>"
    assert tag(pf, "dims").split() == [str(v) for v in L]
    assert tag(pf, "spacetime") == "4" and tag(pf, "volfmt") == "0"
    name = "DiracPropagator" if kind == cmx.SCIDAC_PROPAGATOR else "DiracFermion"
    p = "D" if prec == cmx.DOUBLE else "F"
    assert tag(pr, "datatype") == f"QDP_{p}3_{name}"
    assert tag(pr, "precision") == p and tag(pr, "colors") == "3" and tag(pr, "spins") == "4"
    assert tag(pr, "recordtype") == "0" and tag(pr, "datacount") == "1"
    ncomp = 144 if kind == cmx.SCIDAC_PROPAGATOR else 12
    real = 8 if prec == cmx.DOUBLE else 4
    site_bytes = ncomp * 2 * real
    assert tag(pr, "typesize") == str(site_bytes) and len(data) == V * site_bytes
    # the binary data: big-endian, lexicographic sites, QDP++ element order
    dt = ">c16" if prec == cmx.DOUBLE else ">c8"
    arr = np.frombuffer(data, dt).astype(np.complex128)
    host = f[lex_to_host(L)]
    if kind == cmx.SCIDAC_PROPAGATOR:
        # host [s1*3+c1][s2*3+c2] -> file [s1][s2][c1][c2]
        want = host.reshape(V, 4, 3, 4, 3).transpose(0, 1, 3, 2, 4).reshape(-1)
    else:
        want = host.reshape(-1)
    if prec == cmx.SINGLE:
        want = want.astype(np.complex64).astype(np.complex128)
    assert np.array_equal(arr, want)
    # the checksum
    suma = sumb = 0
    for s in range(V):
        c = zlib.crc32(data[s * site_bytes:(s + 1) * site_bytes]) & 0xFFFFFFFF
        ra, rb = s % 29, s % 31
        suma ^= ((c << ra) | (c >> (32 - ra))) & 0xFFFFFFFF if ra else c
        sumb ^= ((c << rb) | (c >> (32 - rb))) & 0xFFFFFFFF if rb else c
    assert int(tag(ck, "suma"), 16) == suma and int(tag(ck, "sumb"), 16) == sumb


@pytest.mark.parametrize("kind", [cmx.SCIDAC_PROPAGATOR, cmx.SCIDAC_FERMION])
def test_scidac_round_trip(tmp_path, kind):
    f = random_field(kind, 11 + kind)
    for prec in (cmx.DOUBLE, cmx.SINGLE):
        path = tmp_path / f"rt{prec}.lime"
        cmx.field_write_scidac(path, L, f, kind=kind, prec=prec, file_xml="<a/>", record_xml="<b>x</b>")
        g, fx, rx = cmx.field_read_scidac(path, L, kind=kind)
        assert fx == "<a/>" and rx == "<b>x</b>"
        if prec == cmx.DOUBLE:
            assert np.array_equal(g, f)
        else:
            assert np.array_equal(g, f.astype(np.complex64).astype(np.complex128))


def test_scidac_reader_rejects(tmp_path):
    f = random_field(cmx.SCIDAC_PROPAGATOR, 21)
    path = tmp_path / "p.lime"
    cmx.field_write_scidac(path, L, f)
    with pytest.raises(cmx.CmxError):  # other lattice
        cmx.field_read_scidac(path, (4, 4, 4, 6))
    with pytest.raises(cmx.CmxError):  # a propagator read as a fermion
        cmx.field_read_scidac(path, L, kind=cmx.SCIDAC_FERMION)
    buf = bytearray(path.read_bytes())
    recs = parse_lime(bytes(buf))
    off = bytes(buf).index(recs[4][3][:64]) + 1000
    buf[off] ^= 0x01  # one bit of the binary data
    bad = tmp_path / "bad.lime"
    bad.write_bytes(bytes(buf))
    with pytest.raises(cmx.CmxError, match="checksum"):
        cmx.field_read_scidac(bad, L)
    with pytest.raises(cmx.CmxError):
        cmx.field_write_scidac(tmp_path / "x.lime", (3, 4, 4, 4), np.zeros((192, 12, 12), complex))
