"""Light curves from FITS files (TESS / Kepler SPOC light-curve products) without astropy.

The reference reads them with astropy.io.fits (test_multi_gpu_consistency.py:14-31,
test_memory_leak.py:36-39, test1.py:88-94): TIME, PDCSAP_FLUX and PDCSAP_FLUX_ERR of the
first extension, NaN rows dropped, flux normalised by its median. This module parses the
FITS structure directly (FITS Standard 4.0): headers of 2880-byte blocks of 80-character
cards, and a BINTABLE extension whose rows are NAXIS1 big-endian bytes described by the
TFORMn / TTYPEn / TSCALn / TZEROn keywords.
"""
from __future__ import annotations

import numpy as np

BLOCK = 2880
CARD = 80

# binary-table TFORM codes of fixed-width numeric columns -> big-endian numpy types
_CODES = {"L": "i1", "B": "u1", "I": ">i2", "J": ">i4", "K": ">i8", "E": ">f4", "D": ">f8",
          "C": ">c8", "M": ">c16", "A": "S1"}


def _value(text):
    """Value of one header card's value field (string, logical, integer or float)."""
    v = text.split("/", 1)[0].strip() if not text.strip().startswith("'") else text.strip()
    if v.startswith("'"):
        end = 1
        out = []
        while end < len(v):          # '' is an escaped quote
            ch = v[end]
            if ch == "'":
                if end + 1 < len(v) and v[end + 1] == "'":
                    out.append("'")
                    end += 2
                    continue
                break
            out.append(ch)
            end += 1
        return "".join(out).rstrip()
    if v in ("T", "F"):
        return v == "T"
    if v == "":
        return None
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace("D", "E"))
    except ValueError:
        return v


def _read_header(buf, pos):
    """(header dict, position of the data after the header)."""
    hdr = {}
    while True:
        if pos + BLOCK > len(buf):
            raise ValueError("truncated FITS header")
        block = buf[pos:pos + BLOCK]
        pos += BLOCK
        for i in range(0, BLOCK, CARD):
            card = block[i:i + CARD].decode("ascii", errors="replace")
            key = card[:8].strip()
            if key == "END":
                return hdr, pos
            if card[8:10] == "= " and key:
                hdr[key] = _value(card[10:])


def _data_bytes(hdr):
    naxis = int(hdr.get("NAXIS", 0))
    if naxis == 0:
        return 0
    n = 1
    for k in range(1, naxis + 1):
        n *= int(hdr["NAXIS%d" % k])
    bitpix = abs(int(hdr["BITPIX"]))
    size = n * bitpix // 8 + int(hdr.get("PCOUNT", 0))
    return (size + BLOCK - 1) // BLOCK * BLOCK


def hdus(path):
    """[(header, data bytes)] of every HDU of the file."""
    with open(path, "rb") as f:
        buf = f.read()
    out = []
    pos = 0
    while pos < len(buf):
        if not buf[pos:pos + 8].strip():
            break
        hdr, pos = _read_header(buf, pos)
        nbytes = _data_bytes(hdr)
        out.append((hdr, buf[pos:pos + nbytes]))
        pos += nbytes
    return out


def _tform(tform):
    """(repeat count, code) of a binary-table TFORM value such as 'D', '1E', '16A'."""
    s = tform.strip()
    i = 0
    while i < len(s) and s[i].isdigit():
        i += 1
    r = int(s[:i]) if i else 1
    return r, s[i]


def table(hdr, data):
    """Columns of a BINTABLE HDU as a dict of numpy arrays (native byte order, TSCAL and
    TZERO applied, float columns as float64)."""
    if hdr.get("XTENSION") != "BINTABLE":
        raise ValueError("not a binary table: %r" % hdr.get("XTENSION"))
    rowlen, nrows = int(hdr["NAXIS1"]), int(hdr["NAXIS2"])
    names, formats, offsets = [], [], []
    off = 0
    for k in range(1, int(hdr["TFIELDS"]) + 1):
        r, code = _tform(hdr["TFORM%d" % k])
        name = str(hdr.get("TTYPE%d" % k, "COL%d" % k))
        if code == "X":                        # bit array: ceil(r / 8) bytes
            dt, width = ("u1", (r + 7) // 8), (r + 7) // 8
        elif code == "A":
            dt, width = ("S%d" % r), r
        elif code in _CODES:
            base = np.dtype(_CODES[code])
            dt = (base, r) if r != 1 else base
            width = base.itemsize * r
        else:
            raise ValueError("unsupported TFORM %r" % hdr["TFORM%d" % k])
        names.append(name)
        formats.append(dt)
        offsets.append(off)
        off += width
    if off != rowlen:
        raise ValueError("row width %d differs from NAXIS1 %d" % (off, rowlen))
    rows = np.frombuffer(data, dtype=np.dtype({"names": names, "formats": formats,
                                               "offsets": offsets, "itemsize": rowlen}),
                         count=nrows)
    cols = {}
    for k, name in enumerate(names, start=1):
        col = rows[name]
        scale = hdr.get("TSCAL%d" % k)
        zero = hdr.get("TZERO%d" % k)
        if col.dtype.kind in "iuf" and (scale is not None or zero is not None):
            col = col.astype(np.float64) * (1.0 if scale is None else float(scale)) + (
                0.0 if zero is None else float(zero))
        elif col.dtype.kind == "f":
            col = col.astype(np.float64)
        elif col.dtype.kind in "iu":
            col = col.astype(col.dtype.newbyteorder("="))
        cols[name] = col
    return cols


def read_lightcurve(path, flux="PDCSAP_FLUX", flux_err="PDCSAP_FLUX_ERR", normalize=True):
    """(t, y, dy, header) of a SPOC light-curve file: the first extension's TIME, `flux`
    and `flux_err` columns, rows with a NaN in any of them dropped, and (normalize) flux
    and error divided by the median flux, as the reference's scripts prepare it
    (test_multi_gpu_consistency.py:14-31). header merges the primary and table headers."""
    h = hdus(path)
    if len(h) < 2:
        raise ValueError("no table extension in %s" % path)
    cols = table(*h[1])
    t, y, dy = cols["TIME"], cols[flux], cols[flux_err]
    keep = ~np.isnan(t) & ~np.isnan(y) & ~np.isnan(dy)
    t, y, dy = t[keep], y[keep], dy[keep]
    if normalize:
        med = np.median(y)
        y = y / med
        dy = dy / med
    header = dict(h[0][0])
    header.update(h[1][0])
    return t, y, dy, header
