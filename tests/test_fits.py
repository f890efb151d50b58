"""FITS reader and writer of libcsx.so (csim_fits.cpp's load_mat / save_mat
conventions), checked against an independent numpy parse of the FITS standard's
layout and against the reference's own data files where they are present."""
import os

import numpy as np
import pytest

from csim_amd import fits

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = "/root/reference"


def parse(path):
    """Header dict and data bytes of a primary-HDU FITS file, read per the standard
    (80-byte cards in 2880-byte blocks, big-endian data)."""
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    hdr, off = {}, 0
    while True:
        card = raw[off:off + 80].decode("ascii")
        off += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            hdr[key] = card[10:].split("/")[0].strip()
    off = (off + 2879) // 2880 * 2880
    return hdr, raw[off:]


def write_std(path, data_rows_first, bitpix, naxis, bscale=None, bzero=None):
    """An independent FITS writer (numpy, big-endian), data given in file order."""
    cards = ["SIMPLE  = %20s" % "T", "BITPIX  = %20d" % bitpix, "NAXIS   = %20d" % len(naxis)]
    cards += ["NAXIS%-3d= %20d" % (i + 1, n) for i, n in enumerate(naxis)]
    if bscale is not None:
        cards += ["BSCALE  = %20s" % repr(float(bscale)), "BZERO   = %20s" % repr(float(bzero))]
    cards += ["END"]
    h = "".join(c.ljust(80) for c in cards)
    h = h.ljust((len(h) + 2879) // 2880 * 2880)
    dt = {8: ">u1", 16: ">i2", 32: ">i4", -32: ">f4", -64: ">f8"}[bitpix]
    d = np.ascontiguousarray(data_rows_first, dtype=dt).tobytes()
    d = d.ljust((len(d) + 2879) // 2880 * 2880, b"\0")
    open(path, "wb").write(h.encode("ascii") + d)


def test_write_matches_save_mat_layout(tmp_path):
    rng = np.random.default_rng(7)
    m = rng.standard_normal((5, 3))          # 5 rows, 3 columns
    p = str(tmp_path / "m.fits")
    fits.write_fits(p, m)
    hdr, data = parse(p)
    assert hdr["SIMPLE"] == "T" and int(hdr["BITPIX"]) == -64 and int(hdr["NAXIS"]) == 2
    assert int(hdr["NAXIS1"]) == 3 and int(hdr["NAXIS2"]) == 5        # NAXIS1 = columns
    got = np.frombuffer(data[:m.size * 8], dtype=">f8").reshape(5, 3)   # rows one after another
    assert np.array_equal(got, m)
    assert not any(data[m.size * 8:])                                   # zero padding


def test_round_trip_and_vector(tmp_path):
    rng = np.random.default_rng(8)
    for shape in ((1, 1), (4, 9), (37, 2), (360, 361)):
        m = rng.standard_normal(shape) * 10.0 ** rng.integers(-12, 12, size=shape)
        p = str(tmp_path / f"r{shape[0]}x{shape[1]}.fits")
        fits.write_fits(p, m)
        assert np.array_equal(fits.read_fits(p), m)
    v = np.arange(6.0)
    fits.write_fits(str(tmp_path / "v.fits"), v)            # a vector saves as n x 1
    assert np.array_equal(fits.read_fits(str(tmp_path / "v.fits")), v[:, None])


def test_spaces_become_underscores(tmp_path):
    fits.write_fits(str(tmp_path / "a b.fits"), np.eye(2))
    assert os.path.exists(str(tmp_path / "a_b.fits"))


@pytest.mark.parametrize("bitpix", [8, 16, 32, -32, -64])
def test_read_integer_and_float_images(tmp_path, bitpix):
    rng = np.random.default_rng(bitpix + 100)
    rows, cols, slices = 6, 4, 3
    if bitpix == 8:
        f = rng.integers(0, 256, size=(slices, rows, cols))
    elif bitpix > 0:
        f = rng.integers(-2 ** (bitpix - 1), 2 ** (bitpix - 1), size=(slices, rows, cols))
    else:
        f = rng.standard_normal((slices, rows, cols)).astype(np.float32 if bitpix == -32 else np.float64)
    p = str(tmp_path / "c.fits")
    write_std(p, f, bitpix, (cols, rows, slices), bscale=0.5, bzero=-3.0)
    got = fits.read_fits(p)
    assert got.shape == (rows, cols, slices)
    want = -3.0 + 0.5 * np.asarray(f, dtype=np.float64)
    assert np.array_equal(got, want.transpose(1, 2, 0))
    write_std(p, f[0], bitpix, (cols, rows))
    assert np.array_equal(fits.read_fits(p), np.asarray(f[0], dtype=np.float64))


def test_complex_pairs(tmp_path):
    rng = np.random.default_rng(9)
    z = rng.standard_normal((7, 5)) + 1j * rng.standard_normal((7, 5))
    p = str(tmp_path / "z.fits")
    fits.save_mat(p, z, "reIm")
    assert os.path.exists(str(tmp_path / "z_re.fits")) and os.path.exists(str(tmp_path / "z_im.fits"))
    assert np.array_equal(fits.load_mat(p, complex_pair=True), z)
    fits.save_mat(p, z, "amPh")
    am, ph = fits.read_fits(str(tmp_path / "z_am.fits")), fits.read_fits(str(tmp_path / "z_ph.fits"))
    assert np.array_equal(am, np.abs(z)) and np.array_equal(ph, np.angle(z))
    with pytest.raises(ValueError):
        fits.save_mat(p, z, "polar")


def test_reader_errors_are_reported(tmp_path):
    from csim_amd import CsxError
    with pytest.raises(CsxError):
        fits.read_fits(str(tmp_path / "missing.fits"))
    p = str(tmp_path / "short.fits")
    open(p, "wb").write(b"SIMPLE  =                    T".ljust(80))
    with pytest.raises(CsxError):
        fits.read_fits(p)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data files absent")
def test_reads_reference_files_like_the_golden_fixture():
    fx = np.load(os.path.join(GOLDEN, "mono_kat.npz"))
    base = os.path.join(REF, "donutPiaaFullCircleExample", "mono_2_6")
    M = fits.load_mat(os.path.join(base, "responseData", "responseData_s0_p0.fits"), complex_pair=True)
    assert np.array_equal(M, fx["M"])
    best = fits.read_fits(os.path.join(base, "optTest", "bestOptVector.fits"))
    assert np.array_equal(best.ravel(), fx["best"])
