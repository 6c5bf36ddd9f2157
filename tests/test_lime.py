"""LIME / SciDAC / ILDG file I/O (CPU).

An independent Python reader of the LIME record layout (144-byte big-endian
header: magic 0x456789ab, version, MB/ME flags, length, type; payload padded to
8 bytes) and of the SciDAC checksum (zlib crc32 of each site's stored bytes,
rotated by rank mod 29 / 31 and XOR-ed; reference dml.c:41-53) checks what the
library writes; files written here by Python in the ILDG layout of the
reference's reader (io.c:202-347) are read back by the library."""
import struct
import zlib

import numpy as np
import pytest

import cvc_amd


def records(path):
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        magic, ver, flags, n = struct.unpack(">IHHQ", data[pos:pos + 16])
        assert magic == 0x456789AB and ver == 1
        typ = data[pos + 16:pos + 144].split(b"\0")[0].decode()
        payload = data[pos + 144:pos + 144 + n]
        pad = (8 - n % 8) % 8
        assert data[pos + 144 + n:pos + 144 + n + pad] == b"\0" * pad
        out.append((typ, bool(flags & 0x8000), bool(flags & 0x4000), payload))
        pos += 144 + n + pad
    assert pos == len(data)
    return out


def scidac(payload, site_bytes, ranks):
    a = b = 0
    for k, r in enumerate(ranks):
        w = zlib.crc32(payload[k * site_bytes:(k + 1) * site_bytes]) & 0xFFFFFFFF
        ra, rb = r % 29, r % 31
        a ^= ((w << ra) | (w >> (32 - ra))) & 0xFFFFFFFF if ra else w
        b ^= ((w << rb) | (w >> (32 - rb))) & 0xFFFFFFFF if rb else w
    return a, b


@pytest.mark.parametrize("prec", [64, 32])
def test_contraction_file(tmp_path, prec):
    dims = (4, 2, 3, 2)
    V = int(np.prod(dims))
    rng = np.random.default_rng(prec)
    s = rng.standard_normal((V, 16, 2))
    path = tmp_path / "cvc2_v_x.0007"
    ck = cvc_amd.lime_write_contraction(path, s, dims, prec=prec, nconf=7, source=0)
    recs = records(path)
    assert [r[0] for r in recs] == ["cvc_contraction-format", "scidac-binary-data", "scidac-checksum"]
    assert [(r[1], r[2]) for r in recs] == [(True, True), (True, False), (False, True)]
    xml = recs[0][3].decode()
    for tag, val in (("precision", prec), ("components", 16), ("lx", 2), ("ly", 3), ("lz", 2), ("lt", 4),
                     ("nconf", 7), ("source", 0)):
        assert f"<{tag}>{val}</{tag}>" in xml
    fmt = ">f8" if prec == 64 else ">f4"
    raw = np.frombuffer(recs[1][3], dtype=fmt).reshape(V, 16, 2)
    want = s if prec == 64 else s.astype(np.float32).astype(np.float64)
    assert np.array_equal(raw.astype(np.float64), want)
    site = 16 * 2 * (8 if prec == 64 else 4)
    assert scidac(recs[1][3], site, range(V)) == ck
    assert f"<suma>{ck[0]:08x}</suma>" in recs[2][3].decode() and f"<sumb>{ck[1]:08x}</sumb>" in recs[2][3].decode()
    back, ck2 = cvc_amd.lime_read_contraction(path, dims)
    assert ck2 == ck and np.array_equal(back, want)


def test_contraction_checksum_mismatch(tmp_path):
    dims = (2, 2, 2, 2)
    s = np.ones((16, 16, 2))
    path = tmp_path / "c"
    cvc_amd.lime_write_contraction(path, s, dims)
    raw = bytearray(path.read_bytes())
    hdr = raw.find(b"scidac-binary-data") - 16
    raw[hdr + 144 + 3] ^= 0x40                     # flip a bit of the first value
    path.write_bytes(bytes(raw))
    with pytest.raises(cvc_amd.CvchipError):
        cvc_amd.lime_read_contraction(path, dims)
    with pytest.raises(cvc_amd.CvchipError):
        cvc_amd.lime_read_contraction(path, (2, 2, 2, 4))   # wrong volume


def write_ildg_python(path, dims, g, prec=64):
    """An ILDG file in the layout read by io.c:202-347, written without the library."""
    T, LX, LY, LZ = dims
    G = g.reshape(T, LX, LY, LZ, 4, 18)
    site = np.concatenate([G[..., 1, :], G[..., 2, :], G[..., 3, :], G[..., 0, :]], -1)   # x, y, z, t
    site = site.transpose(0, 3, 2, 1, 4)                                                  # t, z, y, x
    payload = site.astype(">f8" if prec == 64 else ">f4").tobytes()
    a, b = scidac(payload, 72 * (prec // 8), range(T * LX * LY * LZ))
    xml = b"<?xml version=\"1.0\"?><ildgFormat><precision>%d</precision></ildgFormat>" % prec
    ck = (b"<?xml version=\"1.0\" encoding=\"UTF-8\"?><scidacChecksum><version>1.0</version>"
          b"<suma>%08x</suma><sumb>%08x</sumb></scidacChecksum>" % (a, b))
    with open(path, "wb") as f:
        for typ, flags, data in (("ildg-format", 0x8000, xml), ("ildg-binary-data", 0, payload),
                                 ("scidac-checksum", 0x4000, ck)):
            f.write(struct.pack(">IHHQ", 0x456789AB, 1, flags, len(data)) + typ.encode().ljust(128, b"\0"))
            f.write(data + b"\0" * ((8 - len(data) % 8) % 8))
    return a, b


@pytest.mark.parametrize("prec", [64, 32])
def test_gauge_file(tmp_path, prec):
    dims = (4, 2, 3, 2)
    g = cvc_amd.gauge_random(dims, 3, 0.3)
    path = tmp_path / "conf.0003"
    ck = write_ildg_python(path, dims, g, prec)
    back, ck2 = cvc_amd.lime_read_gauge(path, dims)
    want = g if prec == 64 else g.astype(np.float32).astype(np.float64)
    assert ck2 == ck and np.array_equal(back, want)
    # the library's writer produces a file the same reader returns unchanged, same checksum
    path2 = tmp_path / "conf2"
    ck3 = cvc_amd.lime_write_gauge(path2, dims, g, prec)
    assert ck3 == ck
    back2, _ = cvc_amd.lime_read_gauge(path2, dims)
    assert np.array_equal(back2, want)
    assert [r[0] for r in records(path2)] == ["ildg-format", "ildg-binary-data", "scidac-checksum"]


def spinor_file_order(s, dims):
    """s[V, 24] (sites t, x, y, z, z fastest) -> the file's site order t, z, y, x (x fastest)."""
    T, LX, LY, LZ = dims
    return s.reshape(T, LX, LY, LZ, 24).transpose(0, 3, 2, 1, 4).reshape(-1, 24)


@pytest.mark.parametrize("prec", [64, 32])
def test_propagator_file(tmp_path, prec):
    """write_propagator + a second flavour appended (propagator_io.c:125-133, 522-647): record
    sequence and flags, etmc XML, big-endian data in t z y x order, checksums over the file
    order; read_lime_spinor returns record 0 (up) and 1 (down)."""
    dims = (4, 2, 3, 2)
    V = int(np.prod(dims))
    rng = np.random.default_rng(prec + 1)
    up, dn = rng.standard_normal((V, 24)), rng.standard_normal((V, 24))
    path = tmp_path / "prop.0007.04.05.inverted"
    ck_up = cvc_amd.lime_write_propagator(path, dims, up, prec)
    ck_dn = cvc_amd.lime_write_spinor(path, dims, dn, append=True, prec=prec)
    recs = records(path)
    assert [r[0] for r in recs] == ["etmc-propagator-format", "scidac-binary-data", "scidac-checksum",
                                    "scidac-binary-data", "scidac-checksum"]
    assert [(r[1], r[2]) for r in recs] == [(True, False), (False, True), (True, False), (False, True),
                                            (True, False)]
    xml = recs[0][3].decode()
    for tag, val in (("field", "diracFermion"), ("precision", prec), ("flavours", 1), ("lx", 2), ("ly", 3),
                     ("lz", 2), ("lt", 4)):
        assert f"<{tag}>{val}</{tag}>" in xml
    fmt = ">f8" if prec == 64 else ">f4"
    cast = (lambda a: a) if prec == 64 else (lambda a: a.astype(np.float32).astype(np.float64))
    for data_rec, ck_rec, s, ck in ((recs[1], recs[2], up, ck_up), (recs[3], recs[4], dn, ck_dn)):
        raw = np.frombuffer(data_rec[3], dtype=fmt).reshape(V, 24).astype(np.float64)
        assert np.array_equal(raw, spinor_file_order(cast(s), dims))
        assert scidac(data_rec[3], 24 * prec // 8, range(V)) == ck
        txt = ck_rec[3].decode()
        assert f"<suma>{ck[0]:#010x}</suma>" in txt and f"<sumb>{ck[1]:#010x}</sumb>" in txt
    back_up, c0 = cvc_amd.lime_read_spinor(path, dims, 0)
    back_dn, c1 = cvc_amd.lime_read_spinor(path, dims, 1)
    assert c0 == ck_up and c1 == ck_dn
    assert np.array_equal(back_up, cast(up)) and np.array_equal(back_dn, cast(dn))


def test_spinor_file_written_by_python(tmp_path):
    """A spinor file in the layout of read_binary_spinor_data (propagator_io.c:344-405) written
    without the library (64-bit up record, 32-bit down record) reads back exactly."""
    dims = (2, 3, 2, 4)
    V = int(np.prod(dims))
    rng = np.random.default_rng(5)
    up, dn = rng.standard_normal((V, 24)), rng.standard_normal((V, 24))
    with open(tmp_path / "p", "wb") as f:
        for typ, flags, data in (("etmc-propagator-format", 0x8000, b"<etmcFormat/>"),
                                 ("scidac-binary-data", 0x4000, spinor_file_order(up, dims).astype(">f8").tobytes()),
                                 ("scidac-binary-data", 0xC000, spinor_file_order(dn, dims).astype(">f4").tobytes())):
            f.write(struct.pack(">IHHQ", 0x456789AB, 1, flags, len(data)) + typ.encode().ljust(128, b"\0"))
            f.write(data + b"\0" * ((8 - len(data) % 8) % 8))
    a, _ = cvc_amd.lime_read_spinor(tmp_path / "p", dims, 0)
    b, _ = cvc_amd.lime_read_spinor(tmp_path / "p", dims, 1)
    assert np.array_equal(a, up)
    assert np.array_equal(b, dn.astype(np.float32).astype(np.float64))


def test_spinor_cmi_fallback(tmp_path):
    """A file that is not LIME is read as CMI (read_cmi, propagator_io.c:842-892): host-order
    floats, sites x, y, z, t (t fastest), times 2 kappa."""
    dims = (4, 2, 3, 2)
    T, LX, LY, LZ = dims
    V = int(np.prod(dims))
    rng = np.random.default_rng(9)
    s = rng.standard_normal((V, 24)).astype(np.float32)
    cmi = s.reshape(T, LX, LY, LZ, 24).transpose(1, 2, 3, 0, 4)
    (tmp_path / "c").write_bytes(np.ascontiguousarray(cmi).astype("=f4").tobytes())
    kappa = 0.15
    back, ck = cvc_amd.lime_read_spinor(tmp_path / "c", dims, 0, kappa=kappa)
    assert np.array_equal(back, 2.0 * kappa * s.astype(np.float64))
    (tmp_path / "short").write_bytes(b"\0" * (V * 96 - 4))
    with pytest.raises(cvc_amd.CvchipError):
        cvc_amd.lime_read_spinor(tmp_path / "short", dims, 0)


def test_spinor_errors(tmp_path):
    dims = (2, 2, 2, 2)
    s = np.arange(16 * 24, dtype=np.float64).reshape(16, 24)
    path = tmp_path / "p"
    cvc_amd.lime_write_propagator(path, dims, s)
    with pytest.raises(cvc_amd.CvchipError):
        cvc_amd.lime_read_spinor(path, (2, 2, 2, 4), 0)          # wrong volume
    with pytest.raises(cvc_amd.CvchipError):
        cvc_amd.lime_read_spinor(path, dims, 1)                  # no record 1, and not a CMI file
    raw = bytearray(path.read_bytes())
    hdr = raw.find(b"scidac-binary-data") - 16
    raw[hdr + 144 + 9] ^= 0x01                                   # a bit of the second value
    path.write_bytes(bytes(raw))
    with pytest.raises(cvc_amd.CvchipError):
        cvc_amd.lime_read_spinor(path, dims, 0)


@pytest.mark.parametrize("length", [2 ** 64 - 64, 2 ** 63, 10 ** 6])
def test_malformed_record_length_is_rejected(tmp_path, length):
    """A record header whose length runs past the end of the file (also one
    near 2^64, where offset + length wraps around) makes the reader return an
    error instead of allocating that length."""
    dims = (2, 2, 2, 2)
    path = tmp_path / "c"
    cvc_amd.lime_write_contraction(path, np.ones((16, 16, 2)), dims)
    raw = bytearray(path.read_bytes())
    hdr = raw.find(b"scidac-binary-data") - 16
    raw[hdr + 8:hdr + 16] = struct.pack(">Q", length)
    path.write_bytes(bytes(raw))
    with pytest.raises(cvc_amd.CvchipError):
        cvc_amd.lime_read_contraction(path, dims)
