"""Input and output of moduli, supports and reconstructions without
pyfits/astropy (``speckle/io.py``: ``open`` 35-91, ``save`` 93-175,
``openfits`` 384-436, ``openimage`` 480-500, ``writefits`` 502-551,
``save_fits`` 714-745, ``_labview_to_andor`` 626-645, ``open_ds9_mask``
1309-1474): FITS, greyscale images through PIL, text arrays, ds9 regions.

Implements the FITS standard's primary HDU and IMAGE extensions: 2880-byte
blocks of 80-character header cards, big-endian data with
BITPIX 8/16/32/64/-32/-64, BSCALE/BZERO, gzip-compressed files.  As in the
reference, ``openfits`` returns the first IMAGE extension if the file has one
(LabView files), reoriented by ``rot90(fliplr(.))`` per frame, and the
primary data otherwise (Andor/pyfits files).
"""
from __future__ import annotations

import builtins
import gzip
import os

import numpy as np

__all__ = ["open", "save", "openfits", "writefits", "save_fits", "read_header",
           "set_overwrite", "openimage", "open_ds9_mask", "get_fits_dimensions"]

BLOCK = 2880
CARD = 80
_BITPIX = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}

overwrite_default = False


def set_overwrite(val):
    """Default overwrite behaviour of the FITS writers (``io.py:19-30``)."""
    global overwrite_default
    if isinstance(val, bool):
        overwrite_default = val


def _read_bytes(filename):
    with builtins.open(filename, "rb") as fh:
        raw = fh.read()
    if raw[:2] == b"\x1f\x8b":
        raw = gzip.decompress(raw)
    return raw


def _parse_value(text):
    text = text.strip()
    if not text:
        return None
    if text.startswith("'"):
        # string value: quotes doubled inside, trailing blanks not significant
        end, out = 1, []
        while end < len(text):
            c = text[end]
            if c == "'":
                if end + 1 < len(text) and text[end + 1] == "'":
                    out.append("'")
                    end += 2
                    continue
                break
            out.append(c)
            end += 1
        return "".join(out).rstrip()
    text = text.split("/")[0].strip()
    if text == "T":
        return True
    if text == "F":
        return False
    try:
        return int(text)
    except ValueError:
        pass
    try:
        return float(text.replace("D", "E"))
    except ValueError:
        return text


def _parse_header(raw, pos):
    """Cards from `pos` up to END; returns (ordered dict, data offset)."""
    header = {}
    while True:
        if pos + BLOCK > len(raw):
            raise IOError("truncated FITS header")
        block = raw[pos:pos + BLOCK].decode("ascii", errors="replace")
        pos += BLOCK
        for k in range(0, BLOCK, CARD):
            card = block[k:k + CARD]
            key = card[:8].strip()
            if key == "END":
                return header, pos
            if card[8:10] == "= " and key:
                header[key] = _parse_value(card[10:])
            elif key and key not in ("COMMENT", "HISTORY"):
                header.setdefault(key, None)


def _hdus(raw):
    """(header, data array or None) for every HDU of the file."""
    out, pos = [], 0
    while pos < len(raw):
        if not raw[pos:pos + BLOCK].strip(b" \x00"):
            break
        header, pos = _parse_header(raw, pos)
        naxis = int(header.get("NAXIS", 0))
        bitpix = int(header["BITPIX"])
        dims = [int(header["NAXIS%d" % (i + 1)]) for i in range(naxis)]
        count = int(np.prod(dims)) if naxis else 0
        gcount = int(header.get("GCOUNT", 1) or 1)
        pcount = int(header.get("PCOUNT", 0) or 0)
        nbytes = abs(bitpix) // 8 * gcount * (pcount + count)
        data = None
        # the primary HDU holds an image; extensions only when XTENSION = IMAGE
        is_image = not out or header.get("XTENSION") == "IMAGE"
        if naxis and count and is_image:
            if bitpix not in _BITPIX:
                raise IOError("unsupported BITPIX %s" % bitpix)
            arr = np.frombuffer(raw, dtype=_BITPIX[bitpix], count=count, offset=pos)
            arr = arr.reshape(dims[::-1]).astype(_BITPIX[bitpix][1:])
            bscale = header.get("BSCALE", 1.0)
            bzero = header.get("BZERO", 0.0)
            if bscale not in (None, 1, 1.0) or bzero not in (None, 0, 0.0):
                arr = arr * float(bscale if bscale is not None else 1.0) + \
                    float(bzero if bzero is not None else 0.0)
            data = arr
        out.append((header, data))
        pos += (nbytes + BLOCK - 1) // BLOCK * BLOCK
    return out


def read_header(filename, card=0):
    """Header of HDU `card` as a dict (``open_fits_header``, ``io.py:747-755``)."""
    return _hdus(_read_bytes(filename))[card][0]


def _labview_to_andor(img):
    """``io.py:626-645``: rotate 90 degrees counter-clockwise after a
    left-right flip, per frame."""
    assert img.ndim in (2, 3), "_labview_to_andor: Image is neither 2D nor 3D."
    if img.ndim == 3:
        return np.stack([np.rot90(np.fliplr(f)) for f in img])
    return np.rot90(np.fliplr(img))


def openfits(filename, quiet=True, orientImageHDU=True):
    """Data of a FITS file (``io.py:384-436``).

    The first IMAGE extension if there is one (reoriented with
    ``_labview_to_andor`` when ``orientImageHDU``), else the primary data.
    """
    assert os.path.exists(filename), "unknown file %s" % filename
    hdus = _hdus(_read_bytes(filename))
    for k, (header, data) in enumerate(hdus):
        if k > 0 and header.get("XTENSION") == "IMAGE":
            if not quiet:
                print("openfits: %s data located in ImageHDU" % filename)
            return _labview_to_andor(data) if orientImageHDU else data
    if not quiet:
        print("openfits: %s data located in PrimaryHDU" % filename)
    return hdus[0][1]


def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, (int, np.integer)):
        v = "%20d" % int(value)
    elif isinstance(value, (float, np.floating)):
        v = "%20s" % repr(float(value)).upper()
    else:
        s = str(value).replace("'", "''")
        v = "'%-8s'" % s
    text = "%-8s= %s" % (key[:8].upper(), v)
    if comment:
        text += " / " + comment
    return text[:CARD].ljust(CARD)


def writefits(filename, img, header_items=None, overwrite=None):
    """Write `img` as the primary HDU of a FITS file (``io.py:502-551``).

    Real integer and floating arrays up to 999 dimensions; complex arrays
    must be split into components first (``save_fits``).  header_items: extra
    cards {key: value}.
    """
    if header_items is None:
        header_items = {}
    assert isinstance(header_items, dict), "header_items must be a dictionary"
    if overwrite is None:
        overwrite = overwrite_default
    assert isinstance(overwrite, bool), "overwrite must be True/False"
    if os.path.exists(filename) and not overwrite:
        raise IOError("%s exists and overwrite is False" % filename)
    img = np.asarray(img)
    assert not np.iscomplexobj(img), "complex data: save components with save_fits"
    kind = {np.dtype(np.uint8): 8, np.dtype(np.int16): 16, np.dtype(np.int32): 32,
            np.dtype(np.int64): 64, np.dtype(np.float32): -32, np.dtype(np.float64): -64}
    if img.dtype == np.bool_:
        img = img.astype(np.uint8)
    if img.dtype not in kind:
        img = img.astype(np.float64 if img.dtype.kind == "f" else np.int64)
    bitpix = kind[img.dtype]
    cards = [_card("SIMPLE", True), _card("BITPIX", bitpix), _card("NAXIS", img.ndim)]
    for i, d in enumerate(img.shape[::-1]):
        cards.append(_card("NAXIS%d" % (i + 1), d))
    reserved = {"SIMPLE", "BITPIX", "NAXIS", "END", "EXTEND"}
    for key, value in header_items.items():
        if key.upper() in reserved or key.upper().startswith("NAXIS"):
            continue
        cards.append(_card(key, value))
    cards.append("END".ljust(CARD))
    head = "".join(cards).encode("ascii")
    head += b" " * ((-len(head)) % BLOCK)
    data = np.ascontiguousarray(img).astype(_BITPIX[bitpix]).tobytes()
    data += b"\x00" * ((-len(data)) % BLOCK)
    with builtins.open(filename, "wb") as fh:
        fh.write(head)
        fh.write(data)


_save_maps = {"mag": np.abs, "phase": np.angle, "real": np.real, "imag": np.imag}


def _process_components(components):
    """``io.py:1271-1307``: names, plus 'polar', 'cartesian' and 'all'."""
    if isinstance(components, str):
        components = [components]
    out = []
    for c in components:
        if c == "polar":
            out += ["mag", "phase"]
        elif c == "cartesian":
            out += ["real", "imag"]
        elif c == "all":
            out += ["mag", "phase", "real", "imag"]
        else:
            assert c in _save_maps, "components %s not recognized" % c
            out.append(c)
    return out


def save_fits(filename, img, header=None, components=None, overwrite=None,
              append_component=True):
    """Save components of an array as FITS files (``io.py:714-745``):
    'mag', 'phase', 'real', 'imag', 'polar', 'cartesian' (default 'mag');
    each as <name>_<component>.fits unless append_component is False."""
    if header is None:
        header = {}
    if components is None:
        components = ("mag",)
    for ext in ("fits", "jpg", "gif", "png", "csv", "jpeg", "bmp"):
        filename = filename.replace("." + ext, "")
    for c in _process_components(components):
        name = "%s_%s.fits" % (filename, c) if append_component else "%s.fits" % filename
        writefits(name, _save_maps[c](np.asarray(img)), dict(header), overwrite)


def get_fits_dimensions(filename, card=0):
    """Dimensions of a FITS image from its header, slowest axis first
    (``io.py:863-880``)."""
    hdr = read_header(filename, card)
    return tuple(int(hdr["NAXIS%d" % k]) for k in range(int(hdr["NAXIS"]), 0, -1))


def openimage(filename):
    """A greyscale image as an integer array (``io.py:480-500``: PIL
    ``convert("L")``; colour images are converted to luminance)."""
    from PIL import Image
    with Image.open(filename) as im:
        return np.asarray(im.convert("L")).copy()


# -- ds9 region masks (io.py:1309-1474 with shape.py:138-420) ----------------
# The shape rules restate the reference's helpers with AA=False, including
# how it reads each region's numbers: ds9 positions are 1-based and become
# 0-based; a box's width and height are passed as (rows, columns) lengths;
# polygon vertices are shifted by one pixel more than the other shapes; an
# annulus uses its first two radii.

def _radial_mask(size, center, radius):
    """``shape.circle(size, radius, center, AA=False)``: r < radius."""
    rows, cols = np.indices(size, float)
    return np.where(np.hypot(rows - center[0], cols - center[1]) < radius, 1, 0)


def _rect(size, lengths, center):
    """``shape.rect`` (``shape.py:138-215``): integer lengths and centre,
    half lengths by floor division; where the box leaves the array the
    reference's clipping is kept as written (a lower edge below 0 is clipped
    to 0; an upper edge past the end replaces the other bound)."""
    size = (int(size[0]), int(size[1]))
    lengths = (int(lengths[0]), int(lengths[1]))
    center = (int(center[0]), int(center[1]))
    assert lengths[0] > 0 and lengths[1] > 0, "lengths must be > 0"
    assert center[0] > 0 and center[1] > 0, "center must be > 0"
    r_min, r_max = center[0] - lengths[0] // 2, center[0] + lengths[0] // 2
    c_min, c_max = center[1] - lengths[1] // 2, center[1] + lengths[1] // 2
    if r_min < 0:
        r_min = 0
    if c_min < 0:
        c_min = 0
    if r_max > size[0]:
        r_min = size[0]
    if c_max > size[1]:
        r_max = size[1]
    temp = np.zeros(size, int)
    temp[r_min:r_max, c_min:c_max] = 1
    return temp


def _ellipse(size, axes, center, angle):
    """``shape.ellipse(..., AA=False)`` (``shape.py:330-400``) with the
    rotated indices of ``shape._make_indices`` (``shape.py:532-556``)."""
    rows, cols = np.indices(size).astype(float)
    rows -= center[0]
    cols -= center[1]
    if angle != 0:
        a = angle * np.pi / 180.0
        rows, cols = rows * np.cos(a) + cols * np.sin(a), cols * np.cos(a) - rows * np.sin(a)
    ratio = float(axes[1]) / float(axes[0])
    if ratio >= 1:
        rows = rows * (float(axes[1]) / axes[0])
        radius = axes[1]
    else:
        cols = cols * (1.0 / ratio)
        radius = axes[0]
    return np.where(np.hypot(rows, cols) < radius, 1, 0)


def _polygon(desc, dim):
    """``_draw_polygon`` (``io.py:1309-1326``): PIL fill and outline of the
    (x, y) vertices."""
    from PIL import Image, ImageDraw
    ys, xs = dim
    verts = [(desc[2 * v], desc[2 * v + 1]) for v in range(len(desc) // 2)]
    img = Image.new("L", (xs, ys), 0)
    ImageDraw.Draw(img).polygon(verts, outline=1, fill=1)
    return np.array(img).astype(int)


_REGION_SHAPES = ("box", "circle", "polygon", "ellipse", "annulus")


def _parse_region_file(filename, force_reg_size):
    import re
    file_exp = re.compile(r"# Filename: ((?:\/[\w\.\-\ ]+)+)(?:(\[\w\])?)")
    shapes, dim = [], None
    with builtins.open(filename, "r") as fh:
        for line in fh:
            if file_exp.match(line):
                spl = file_exp.split(line)
                afile = spl[1]
                card = spl[2][1:-1] if len(spl) >= 3 and spl[2] else 0
                try:
                    dim = get_fits_dimensions(afile, card)
                except (IOError, OSError, KeyError, ValueError):
                    dim = None
                continue
            if line.split("(")[0] in _REGION_SHAPES:
                shapes.append(line.rstrip("\n"))
    if dim is None:
        # the FITS file named in the region file is unreadable (or no file is
        # named: the reference then fails) -- force_reg_size, else 2048^2
        if force_reg_size is None:
            dim = (2048, 2048)
        elif isinstance(force_reg_size, (tuple, list)):
            dim = (int(force_reg_size[0]), int(force_reg_size[1]))
        else:
            dim = (int(force_reg_size), int(force_reg_size))
    if len(dim) == 3:
        dim = dim[1:]
    return shapes, tuple(int(d) for d in dim)


def open_ds9_mask(filename, individual_regions=False, remove_intersections=False,
                  force_reg_size=None):
    """A mask from a ds9 region file (``io.py:1328-1474``): box, circle,
    polygon, ellipse and annulus regions, on an array the size of the FITS
    file the region file names (else ``force_reg_size`` or 2048 x 2048).

    Returns a 0/1 mask of the union of the regions (with
    ``remove_intersections``, of the pixels covered an odd number of times);
    with ``individual_regions`` the mask times the sum of the region numbers
    (1, 2, ... in file order) covering each pixel."""
    assert isinstance(individual_regions, (bool, int)), \
        "individual_regions must be boolean-evaluable"
    assert isinstance(remove_intersections, (bool, int)), \
        "remove_intersections must be boolean-evaluable"
    assert isinstance(force_reg_size, (type(None), int, tuple, list)), \
        "force_size must be None, int, or iterable"
    if isinstance(force_reg_size, (tuple, list)):
        assert len(force_reg_size) == 2, "force_reg_size must be of length 2"
    shapes, dim = _parse_region_file(filename, force_reg_size)
    data = np.zeros(dim)
    binarydata = np.zeros(dim)
    for count, sdesc in enumerate(shapes, start=1):
        kind = sdesc.split("(")[0]
        desc = [float(v) for v in sdesc[sdesc.find("(") + 1:sdesc.find(")")].split(",")]
        desc[0] -= 1
        desc[1] -= 1
        if kind == "circle":
            obj = _radial_mask(dim, (desc[1], desc[0]), desc[2])
        elif kind == "box":
            center = (desc[1], desc[0])
            obj = _rect(dim, (desc[2], desc[3]), center)
            if len(desc) == 7:
                obj = _rect(dim, (desc[4], desc[5]), center) - obj
        elif kind == "polygon":
            obj = _polygon([v - 1 for v in desc], dim)
        elif kind == "ellipse":
            angle = -1 * desc[-1]
            center = (desc[1], desc[0])
            obj = _ellipse(dim, (desc[2], desc[3]), center, angle)
            if len(desc) == 7:
                obj = _ellipse(dim, (desc[4], desc[5]), center, angle) - obj
        else:  # annulus
            center = (desc[1], desc[0])
            assert center[0] > 0 and center[1] > 0, "center must be > 0"
            r = (desc[2], desc[3])
            obj = _radial_mask(dim, center, max(r)) - _radial_mask(dim, center, min(r))
        data += count * obj
        binarydata += obj
    if remove_intersections:
        mask = np.where((binarydata % 2) == 0, 0, 1)
    else:
        mask = np.where(binarydata >= 1, 1, 0)
    return mask * data if individual_regions else mask


def open(filename, quiet=True, orientImageHDU=True, convert_to="float", delimiter="\t",
         force_reg_size=None):
    """Open a data file by extension (``io.py:35-91``): FITS (.fits, .fit and
    gzip .gz/.zip), images (.jpg, .jpeg, .gif, .png, .bmp: greyscale),
    text arrays (.txt, .csv) and ds9 regions (.reg: a mask).  Pickles
    (.pck) are not opened: loading one runs code from the file."""
    assert len(filename.split(".")) >= 2, "filename appears to have no extension"
    ext = filename.split(".")[-1].lower()
    if ext in ("fits", "fit", "gz", "zip"):
        return openfits(filename, quiet=quiet, orientImageHDU=orientImageHDU)
    if ext in ("jpg", "jpeg", "gif", "png", "bmp"):
        return openimage(filename)
    if ext in ("txt", "csv"):
        return np.loadtxt(filename, delimiter=delimiter if ext == "txt" else ",").astype(
            convert_to)
    if ext == "reg":
        return open_ds9_mask(filename, force_reg_size=force_reg_size)
    raise NotImplementedError("opening .%s files is not provided" % ext)


def save(filename, data, header=None, components=None, overwrite=None, do_zip=False,
         append_component=True, **kwargs):
    """Save by extension (``io.py:93-175``): FITS only; ``do_zip`` gzips each
    written file in place of the original (the reference runs gzip)."""
    assert len(filename.split(".")) >= 2, "filename appears to have no extension"
    ext = filename.split(".")[-1].lower()
    if ext not in ("fits", "fit"):
        raise NotImplementedError("saving .%s files is not provided (FITS only)" % ext)
    if components is None:
        components = ("mag",)        # the reference's default (io.py:733-734)
    base = filename[: -len(ext) - 1]
    save_fits(base + ".fits", data, header, components, overwrite, append_component)
    if do_zip:
        for c in _process_components(components):
            name = "%s_%s.fits" % (base, c) if append_component else "%s.fits" % base
            with builtins.open(name, "rb") as fh, gzip.open(name + ".gz", "wb") as gz:
                gz.write(fh.read())
            os.remove(name)
