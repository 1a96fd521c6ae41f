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
    with pytest.raises(NotImplementedError):      # pickles are not opened
        sio.open(str(tmp_path / "x.pck"))
    raw = gzip.decompress(open(str(tmp_path / "recon_mag.fits.gz"), "rb").read())
    assert raw[:6] == b"SIMPLE"


# --- images (io.py:480-500) and ds9 region masks (io.py:1309-1474) ---------

def test_openimage_greyscale_and_colour(tmp_path):
    from PIL import Image
    g = (np.arange(48 * 40).reshape(48, 40) % 251).astype(np.uint8)
    Image.fromarray(g, "L").save(tmp_path / "g.png")
    got = sio.open(str(tmp_path / "g.png"))
    assert got.dtype == np.uint8 and np.array_equal(got, g)
    rgb = np.stack([g, g[::-1], (g * 3) % 255], -1).astype(np.uint8)
    Image.fromarray(rgb, "RGB").save(tmp_path / "c.bmp")
    want = np.asarray(Image.fromarray(rgb, "RGB").convert("L"))
    assert np.array_equal(sio.open(str(tmp_path / "c.bmp")), want)
    with pytest.raises(NotImplementedError):
        sio.open(str(tmp_path / "x.pck"))


def region_file(tmp_path, shapes, fits_shape=(64, 80), name=None):
    f = tmp_path / "frame.fits"
    sio.writefits(str(f), np.zeros(fits_shape, np.float32), overwrite=True)
    path = tmp_path / (name or "r.reg")
    lines = ["# Region file format: DS9 version 4.1", "# Filename: %s" % f,
             'global color=green dashlist=8 3 width=1', "image"] + shapes
    path.write_text("\n".join(lines) + "\n")
    return str(path)


def test_ds9_shapes_follow_the_reference_rules(tmp_path):
    from PIL import Image, ImageDraw
    rows, cols = np.indices((64, 80), float)
    # circle(x, y, r): 1-based position; pixels with distance < r from (y-1, x-1)
    m = sio.open(region_file(tmp_path, ["circle(20,30,5)"]))
    assert m.shape == (64, 80)
    assert np.array_equal(m, np.hypot(rows - 29, cols - 19) < 5)
    # box(x, y, w, h, angle): centre (y-1, x-1) truncated; w spans rows, h columns
    # (the reference passes (w, h) as (row, column) lengths), half lengths floored
    m = sio.open(region_file(tmp_path, ["box(40,20,10,6,0)"]))
    want = np.zeros((64, 80), int)
    want[14:24, 36:42] = 1
    assert np.array_equal(m, want)
    # polygon: the first vertex is shifted by 2 px, the others by 1 (as the
    # reference subtracts 1 twice from its first two numbers); PIL fill+outline
    m = sio.open(region_file(tmp_path, ["polygon(10,10,40,12,30,50,12,40)"]))
    img = Image.new("L", (80, 64), 0)
    ImageDraw.Draw(img).polygon([(8, 8), (39, 11), (29, 49), (11, 39)], outline=1, fill=1)
    assert np.array_equal(m, np.array(img))
    # ellipse(x, y, a, b, angle): the stretched-circle rule; area ~ pi a b
    m = sio.open(region_file(tmp_path, ["ellipse(40,32,16,8,0)"]))
    assert abs(m.sum() - np.pi * 16 * 8) < 0.05 * np.pi * 16 * 8
    assert np.array_equal(m, ((rows - 31) ** 2 + (2 * (cols - 39)) ** 2) < 16 ** 2)
    r90 = sio.open(region_file(tmp_path, ["ellipse(40,32,16,8,90)"]))
    assert abs(int(r90.sum()) - int(m.sum())) <= 4
    # a quarter turn swaps the axes: 12 px along the columns is inside only
    # the turned ellipse, 12 px along the rows only the unturned one
    assert r90[31, 39 + 12] == 1 and m[31, 39 + 12] == 0
    assert r90[31 + 12, 39] == 0 and m[31 + 12, 39] == 1
    # annulus(x, y, r1, r2): ring between the two radii
    m = sio.open(region_file(tmp_path, ["annulus(50,30,3,6)"]))
    d = np.hypot(rows - 29, cols - 49)
    assert np.array_equal(m, (d < 6) & ~(d < 3))


def test_ds9_combination_rules_and_size(tmp_path):
    shapes = ["circle(20,20,8)", "circle(30,20,8)"]          # overlapping discs
    path = region_file(tmp_path, shapes)
    rows, cols = np.indices((64, 80), float)
    a = np.hypot(rows - 19, cols - 19) < 8
    b = np.hypot(rows - 19, cols - 29) < 8
    assert np.array_equal(sio.open_ds9_mask(path), a | b)
    assert np.array_equal(sio.open_ds9_mask(path, remove_intersections=True), a ^ b)
    numbered = sio.open_ds9_mask(path, individual_regions=True)
    assert np.array_equal(numbered, 1 * a + 2 * b)
    # the FITS file named in the region file is missing: force_reg_size, else 2048^2
    missing = tmp_path / "m.reg"
    missing.write_text("# Filename: /nonexistent/frame.fits\ncircle(5,5,2)\n")
    assert sio.open(str(missing), force_reg_size=32).shape == (32, 32)
    assert sio.open_ds9_mask(str(missing), force_reg_size=(16, 24)).shape == (16, 24)
    assert sio.open_ds9_mask(str(missing)).shape == (2048, 2048)
