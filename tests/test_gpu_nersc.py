"""NERSC gauge files (NerscIO.h) and gauge statistics (avgPlaquette, linkTrace).  The file layout,
byte order and checksum are restated here with numpy, independently of the library: files the
library writes must hold numpy's big-endian bytes of the links, and files numpy writes (IEEE32
little-endian, three or two rows) must read back to numpy's values."""
import numpy as np
import pytest

import grid_amd as gx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


def np_stats(U):
    """avgPlaquette and linkTrace in numpy: U[t, z, y, x, mu, a, b]; direction mu shifts axis 3 - mu."""
    ax = [3, 2, 1, 0]
    V = np.prod(U.shape[:4])
    s = 0.0
    for mu in range(4):
        for nu in range(mu):
            Umu, Unu = U[..., mu, :, :], U[..., nu, :, :]
            Unu_xmu = np.roll(Unu, -1, axis=ax[mu])
            Umu_xnu = np.roll(Umu, -1, axis=ax[nu])
            P = Umu @ Unu_xmu @ np.conj(np.swapaxes(Umu_xnu, -1, -2)) @ np.conj(np.swapaxes(Unu, -1, -2))
            s += np.trace(P, axis1=-2, axis2=-1).real.sum()
    tr = np.trace(U, axis1=-2, axis2=-1).real.sum()
    return s / V / 6 / 3, tr / V / 12


def np_reconstruct(U2):
    """third row = conj(row0 x row1)"""
    r2 = np.conj(np.cross(U2[..., 0, :], U2[..., 1, :]))
    return np.concatenate([U2, r2[..., None, :]], axis=-2)


def header_text(dims, datatype, fp, csum, plaq, trace):
    lines = ['BEGIN_HEADER', 'HDR_VERSION = 1.0', f'DATATYPE = {datatype}', 'STORAGE_FORMAT = 1.0']
    lines += [f'DIMENSION_{d + 1} = {dims[d]}' for d in range(4)]
    lines += [f'LINK_TRACE = {trace:.10g}', f'PLAQUETTE = {plaq:.10g}']
    lines += [f'BOUNDARY_{d + 1} = PERIODIC' for d in range(4)]
    lines += [f'CHECKSUM = {csum:x}', 'ENSEMBLE_ID = test', 'ENSEMBLE_LABEL = numpy', 'SEQUENCE_NUMBER = 7',
              f'FLOATING_POINT = {fp}', 'END_HEADER']
    return ('\n'.join(lines) + '\n').encode()


def split_header(raw):
    i = raw.index(b'END_HEADER\n') + len(b'END_HEADER\n')
    return raw[:i].decode(), raw[i:]


def test_gauge_statistics(ctx):
    dims = (4, 6, 2, 8)                       # distinct extents: a swapped axis would show
    U = gx.random_gauge(dims, 501, 0.4)
    p, t = gx.gauge_statistics(ctx, U, dims)
    rp, rt = np_stats(U)
    assert abs(p - rp) < 1e-13 and abs(t - rt) < 1e-13, (p, rp, t, rt)
    unit = np.zeros_like(U)
    unit[..., :, :] = np.eye(3)
    assert gx.gauge_statistics(ctx, unit, dims) == (1.0, 1.0)


@pytest.mark.parametrize('two_row', [False, True])
def test_nersc_write_read(ctx, tmp_path, two_row):
    """The writer's data bytes are numpy's '>f8' bytes of the (first two rows of the) links, its
    CHECKSUM is numpy's 32-bit word sum of the host-order doubles, and reading the file returns the
    links (bitwise for three rows; the third row rebuilt to rounding for two)."""
    dims = (4, 4, 6, 4)
    U = gx.random_gauge(dims, 502, 0.5)
    path = tmp_path / 'cfg.nersc'
    gx.nersc_write(ctx, str(path), U, dims, two_row=two_row, ensemble_label='LBL', sequence_number=3)
    head, data = split_header(path.read_bytes())
    stored = U[..., :2, :] if two_row else U
    assert data == np.ascontiguousarray(stored).astype('>c16').tobytes()
    keys = dict(l.split('=', 1) for l in head.splitlines() if '=' in l)
    keys = {k.strip(): v.strip() for k, v in keys.items()}
    assert keys['DATATYPE'] == ('4D_SU3_GAUGE' if two_row else '4D_SU3_GAUGE_3x3')
    assert keys['FLOATING_POINT'] == 'IEEE64BIG'
    words = np.ascontiguousarray(stored).astype('<c16').view('<u4')
    assert int(keys['CHECKSUM'], 16) == int(words.sum(dtype=np.uint64) & 0xffffffff)
    rp, rt = np_stats(U)
    assert abs(float(keys['PLAQUETTE']) - rp) < 1e-9 and abs(float(keys['LINK_TRACE']) - rt) < 1e-9
    V, info = gx.nersc_read(ctx, str(path), dims)
    if two_row:
        assert np.max(np.abs(V - U)) < 1e-14
    else:
        assert np.array_equal(V, U)
    assert info['ok'] and info['two_row'] == int(two_row) and info['dims'] == dims
    assert info['checksum_data'] == info['checksum_header']
    assert info['ensemble_label'] == 'LBL' and info['sequence_number'] == 3


@pytest.mark.parametrize('two_row', [False, True])
def test_nersc_reads_numpy_ieee32(ctx, tmp_path, two_row):
    """A file numpy writes in IEEE32 (little-endian), with numpy's checksum and statistics, reads
    back to numpy's single-precision links."""
    dims = (4, 4, 4, 6)
    U = gx.random_gauge(dims, 503, 0.3)
    stored = (U[..., :2, :] if two_row else U).astype(np.complex64)
    full = np_reconstruct(stored.astype(np.complex128)) if two_row else stored.astype(np.complex128)
    p, t = np_stats(full)
    csum = int(np.ascontiguousarray(stored).view('<u4').sum(dtype=np.uint64) & 0xffffffff)
    dt = '4D_SU3_GAUGE' if two_row else '4D_SU3_GAUGE_3x3'
    path = tmp_path / 'cfg32.nersc'
    path.write_bytes(header_text(dims, dt, 'IEEE32', csum, p, t) + np.ascontiguousarray(stored).astype('<c8').tobytes())
    V, info = gx.nersc_read(ctx, str(path), dims)
    assert info['ok'] and info['floating_point'] == 'IEEE32' and info['ensemble_label'] == 'numpy'
    if two_row:
        assert np.max(np.abs(V - full)) < 1e-14
    else:
        assert np.array_equal(V, full)


def test_nersc_rejects_bad_files(ctx, tmp_path):
    """A flipped data byte fails the checksum; an edited PLAQUETTE fails the plaquette check; a file
    of another lattice, or a missing file, is refused."""
    dims = (4, 4, 4, 4)
    U = gx.random_gauge(dims, 504, 0.5)
    good = tmp_path / 'good.nersc'
    gx.nersc_write(ctx, str(good), U, dims)
    raw = bytearray(good.read_bytes())
    head, data = split_header(bytes(raw))
    bad = tmp_path / 'bad_sum.nersc'
    raw[len(head) + 100] ^= 0x10
    bad.write_bytes(bytes(raw))
    with pytest.raises(gx.GxError, match='checksum'):
        gx.nersc_read(ctx, str(bad), dims)
    _, info = gx.nersc_read(ctx, str(bad), dims, check=False)
    assert not info['ok'] and info['checksum_data'] != info['checksum_header']
    lines = head.splitlines()
    lines = [('PLAQUETTE = 0.123' if l.startswith('PLAQUETTE') else l) for l in lines]
    plaq = tmp_path / 'bad_plaq.nersc'
    plaq.write_bytes(('\n'.join(lines) + '\n').encode() + data)
    with pytest.raises(gx.GxError, match='plaquette'):
        gx.nersc_read(ctx, str(plaq), dims)
    with pytest.raises(gx.GxError):
        gx.nersc_read(ctx, str(good), (4, 4, 4, 8))
    with pytest.raises(gx.GxError):
        gx.nersc_read(ctx, str(tmp_path / 'missing.nersc'), dims)


# ---------------- ILDG (LIME) files ----------------
import struct
import zlib

LIME_MAGIC = 0x456789ab


def lime_records(path):
    """Independent LIME parser: [(type, mb, me, data)] from the published record layout."""
    out = []
    with open(path, 'rb') as f:
        while True:
            h = f.read(144)
            if not h:
                return out
            magic, version, flags, length = struct.unpack('>IHHQ', h[:16])
            assert magic == LIME_MAGIC and version == 1
            typ = h[16:144].split(b'\0')[0].decode()
            data = f.read(length)
            pad = (length + 7) // 8 * 8 - length
            assert f.read(pad) == b'\0' * pad
            out.append((typ, bool(flags & 0x8000), bool(flags & 0x4000), data))


def lime_write(path, records):
    with open(path, 'wb') as f:
        for typ, mb, me, data in records:
            flags = (0x8000 if mb else 0) | (0x4000 if me else 0)
            f.write(struct.pack('>IHHQ', LIME_MAGIC, 1, flags, len(data)) + typ.encode().ljust(128, b'\0'))
            f.write(data + b'\0' * ((len(data) + 7) // 8 * 8 - len(data)))


def scidac_sums(binary, nsite):
    """SciDAC checksums: CRC-32 (zlib) of each site's stored bytes, rotated by site % 29 / % 31, XORed."""
    per = len(binary) // nsite
    a = b = 0
    rot = lambda x, r: ((x << r) | (x >> (32 - r))) & 0xffffffff if r else x
    for s in range(nsite):
        c = zlib.crc32(binary[s * per:(s + 1) * per]) & 0xffffffff
        a ^= rot(c, s % 29)
        b ^= rot(c, s % 31)
    return a, b


@pytest.mark.parametrize('precision', [64, 32])
def test_ildg_write_layout_and_round_trip(ctx, tmp_path, precision):
    """The written file, parsed by an independent LIME reader: the record sequence and flags of
    Grid's writer (file message, then field message), ildg-format of this lattice, the binary
    record equal to numpy's big-endian bytes of the links, SciDAC checksums equal to zlib's CRC-32
    recipe; reading it back gives the field (exactly for 64 bits) and the recomputed metadata."""
    dims = (4, 6, 2, 4)
    U = gx.random_gauge(dims, 601, 0.4)
    path = str(tmp_path / f'cfg{precision}.ildg')
    gx.ildg_write(ctx, path, U, dims, precision=precision, lfn='lfn://test/cfg', description='test field')
    recs = lime_records(path)
    assert [r[0] for r in recs] == ['scidac-private-file-xml', 'scidac-file-xml', 'scidac-private-record-xml',
                                    'scidac-record-xml', 'ildg-format', 'ildg-data-lfn', 'ildg-binary-data',
                                    'scidac-checksum']
    assert [(r[1], r[2]) for r in recs] == [(True, False), (False, True), (True, False)] + [(False, False)] * 4 + \
        [(False, True)]
    rec = {r[0]: r[3] for r in recs}
    fmt = rec['ildg-format'].decode()
    for tag, v in (('field', 'su3gauge'), ('precision', str(precision)), ('lx', '4'), ('ly', '6'), ('lz', '2'),
                   ('lt', '4')):
        assert f'<{tag}>{v}</{tag}>' in fmt
    assert rec['ildg-data-lfn'] == b'lfn://test/cfg'
    want = U.astype('>c16' if precision == 64 else '>c8').tobytes()
    assert rec['ildg-binary-data'] == want
    a, b = scidac_sums(want, int(np.prod(dims)))
    ck = rec['scidac-checksum'].decode()
    assert f'<suma>{a:x}</suma>' in ck and f'<sumb>{b:x}</sumb>' in ck
    V, info = gx.ildg_read(ctx, path, dims)
    if precision == 64:
        assert np.array_equal(V, U)
    else:
        assert np.array_equal(V, U.astype(np.complex64).astype(np.complex128))
    assert info['precision'] == precision and info['has_plaquette'] and info['lfn'] == 'lfn://test/cfg'
    assert (info['scidac_checksuma'], info['scidac_checksumb']) == (a, b)
    assert abs(info['plaquette'] - info['plaquette_file']) < 1e-6


def test_ildg_reads_foreign_file_and_rejects_damage(ctx, tmp_path):
    """A file assembled here from the published formats (records in another order, Grid's
    FieldNormMetaData and grid-format metadata, no USQCD record) reads back to the numpy field;
    a flipped byte, a missing checksum record or another lattice is refused."""
    dims = (4, 4, 4, 2)
    U = gx.random_gauge(dims, 611, 0.3)
    binary = U.astype('>c16').tobytes()
    a, b = scidac_sums(binary, int(np.prod(dims)))
    P, T = gx.gauge_statistics(ctx, U, dims)
    fmt = ('<?xml version="1.0"?><ildgFormat><version>1.0</version><field>su3gauge</field>'
           '<precision>64</precision><lx>4</lx><ly>4</ly><lz>4</lz><lt>2</lt></ildgFormat>').encode()
    grid = (f'<?xml version="1.0"?><FieldMetaData><nd>4</nd><link_trace>{T!r}</link_trace>'
            f'<plaquette>{P!r}</plaquette><floating_point>IEEE64BIG</floating_point></FieldMetaData>').encode()
    norm = f'<?xml version="1.0"?><FieldNormMetaData><norm2>{float(np.sum(np.abs(U) ** 2))!r}</norm2></FieldNormMetaData>'
    ck = f'<?xml version="1.0"?><scidacChecksum><version>1.0</version><suma>{a:x}</suma><sumb>{b:x}</sumb></scidacChecksum>'
    recs = [('grid-format', True, False, grid), ('FieldNormMetaData', False, False, norm.encode()),
            ('ildg-format', False, False, fmt), ('ildg-data-lfn', False, False, b'foreign'),
            ('ildg-binary-data', False, False, binary), ('scidac-checksum', False, True, ck.encode())]
    good = str(tmp_path / 'foreign.ildg')
    lime_write(good, recs)
    V, info = gx.ildg_read(ctx, good, dims)
    assert np.array_equal(V, U) and info['has_norm'] and info['has_plaquette'] and info['lfn'] == 'foreign'
    bad = bytearray(binary)
    bad[1000] ^= 0x10
    for name, rs, d in (('flipped', recs[:4] + [('ildg-binary-data', False, False, bytes(bad))] + recs[5:], dims),
                        ('no-checksum', recs[:5], dims), ('lattice', recs, (4, 4, 2, 4))):
        p = str(tmp_path / f'{name}.ildg')
        lime_write(p, rs)
        with pytest.raises(gx.GxError):
            gx.ildg_read(ctx, p, d)
