"""FITS reader/writer (als_speckle_amd/io.py) on the CPU.

The reference's own test suite has no FITS fixtures (its phasing examples
read data files that are not in the repository), so the reader is pinned by
byte strings built here from the FITS standard's rules (independent of the
writer) and by write/read round trips.
"""
import gzip
import os

import numpy as np
import pytest

from als_speckle_amd import io as sio


def fits_bytes(cards, data_be):
    """A FITS HDU: 80-char cards, END, blank padding to 2880; data padded."""
    head = "".join(c.ljust(80) for c in cards + ["END"]).encode("ascii")
    head += b" " * (-len(head) % 2880)
    return head + data_be + b"\x00" * (-len(data_be) % 2880)


def test_reads_a_hand_built_primary_hdu(tmp_path):
    a = np.arange(12, dtype=np.int16).reshape(3, 4) - 5
    cards = ["SIMPLE  =                    T", "BITPIX  =                   16",
             "NAXIS   =                    2", "NAXIS1  =                    4",
             "NAXIS2  =                    3", "BZERO   =                 10.0",
             "OBJECT  = 'it''s a test'       / comment"]
    path = tmp_path / "h.fits"
    path.write_bytes(fits_bytes(cards, a.astype(">i2").tobytes()))
    got = sio.openfits(str(path))
    assert got.shape == (3, 4)                 # NAXIS1 is the fastest axis
    assert np.array_equal(got, a + 10.0)       # BZERO applied
    h = sio.read_header(str(path))
    assert h["OBJECT"] == "it's a test" and h["NAXIS1"] == 4 and h["SIMPLE"] is True


def test_image_extension_is_reoriented_like_labview(tmp_path):
    prim = ["SIMPLE  =                    T", "BITPIX  =                    8",
            "NAXIS   =                    0", "EXTEND  =                    T"]
    img = np.arange(6, dtype=np.float32).reshape(2, 3)
    ext = ["XTENSION= 'IMAGE   '", "BITPIX  =                  -32",
           "NAXIS   =                    2", "NAXIS1  =                    3",
           "NAXIS2  =                    2", "PCOUNT  =                    0",
           "GCOUNT  =                    1"]
    path = tmp_path / "lv.fits"
    path.write_bytes(fits_bytes(prim, b"") + fits_bytes(ext, img.astype(">f4").tobytes()))
    assert np.array_equal(sio.openfits(str(path)), np.rot90(np.fliplr(img)))
    assert np.array_equal(sio.openfits(str(path), orientImageHDU=False), img)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int16, np.int32, np.uint8])
def test_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(1)
    a = (rng.random((5, 7, 9)) * 100).astype(dtype)
    path = str(tmp_path / "r.fits")
    sio.writefits(path, a, {"EXPOSURE": 0.25, "DETECTOR": "ccd"})
    assert os.path.getsize(path) % 2880 == 0
    b = sio.openfits(path)
    assert b.dtype == np.dtype(dtype) and np.array_equal(a, b)
    h = sio.read_header(path)
    assert h["EXPOSURE"] == 0.25 and h["DETECTOR"] == "ccd" and h["NAXIS3"] == 5
    with open(path, "rb") as fh:
        assert fh.read(30) == b"SIMPLE  =                    T"
    with pytest.raises(IOError):
        sio.writefits(path, a)                 # no overwrite by default
    sio.writefits(path, a[0], overwrite=True)
    assert sio.openfits(path).shape == (7, 9)


def test_gzip_and_open_dispatch_and_components(tmp_path):
    rng = np.random.default_rng(2)
    z = (rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))).astype(np.complex64)
    base = str(tmp_path / "recon.fits")
    sio.save(base, z, components="polar", do_zip=True)
    mag = sio.open(str(tmp_path / "recon_mag.fits.gz"))
    ph = sio.open(str(tmp_path / "recon_phase.fits.gz"))
    assert np.allclose(mag * np.exp(1j * ph), z, atol=1e-6)
    sio.save_fits(str(tmp_path / "c.fits"), z, components="cartesian")
    re = sio.open(str(tmp_path / "c_real.fits"))
    im = sio.open(str(tmp_path / "c_imag.fits"))
    assert np.array_equal(re + 1j * im, z)
    np.savetxt(str(tmp_path / "t.txt"), np.eye(3), delimiter="\t")
    assert np.array_equal(sio.open(str(tmp_path / "t.txt")), np.eye(3))
    with pytest.raises(NotImplementedError):
        sio.open(str(tmp_path / "x.png"))
    raw = gzip.decompress(open(str(tmp_path / "recon_mag.fits.gz"), "rb").read())
    assert raw[:6] == b"SIMPLE"
