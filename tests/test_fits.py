"""FITS reading of the TESS light-curve file the reference's scripts load
(test_multi_gpu_consistency.py:14-31; the file is copied under tests/golden/), pinned by
the invariants of the format and of a SPOC 2-min light curve: the row width equals the sum
of the column widths, cadence numbers are consecutive, times are increasing, inside
[TSTART, TSTOP] and 2 minutes apart, and the primary-header stellar parameters read as
numbers."""
import os

import numpy as np
import pytest

from gtls_amd import fits

FILE = os.path.join(os.path.dirname(__file__), "golden",
                    "tess2022302161335-s0058-0000000021132157-0247-s_lc.fits")


def test_header_cards():
    assert fits._value("                   T / conforms") is True
    assert fits._value("                 100 / width") == 100
    assert fits._value("   -1.5E+02 / x") == -150.0
    assert fits._value(" 1.0D-3") == 1e-3
    assert fits._value(" 'TIC 21132157'      / target") == "TIC 21132157"
    assert fits._value(" 'O''Neil  '") == "O'Neil"
    assert fits._value("                      / empty") is None


def test_tess_table_structure():
    h = fits.hdus(FILE)
    assert [x[0].get("XTENSION") for x in h] == [None, "BINTABLE", "IMAGE"]
    hdr, data = h[1]
    assert hdr["NAXIS1"] == 100 and hdr["NAXIS2"] == 19962 and hdr["TFIELDS"] == 20
    assert len(data) >= hdr["NAXIS1"] * hdr["NAXIS2"]
    cols = fits.table(hdr, data)
    assert list(cols)[:10] == ["TIME", "TIMECORR", "CADENCENO", "SAP_FLUX", "SAP_FLUX_ERR",
                               "SAP_BKG", "SAP_BKG_ERR", "PDCSAP_FLUX", "PDCSAP_FLUX_ERR",
                               "QUALITY"]
    assert all(len(c) == 19962 for c in cols.values())
    assert np.all(np.diff(cols["CADENCENO"]) == 1)
    t = cols["TIME"]
    ok = ~np.isnan(t)
    assert np.all(np.diff(t[ok]) > 0)
    assert hdr["TSTART"] - 0.01 <= t[ok].min() and t[ok].max() <= hdr["TSTOP"] + 0.01
    assert np.median(np.diff(t[ok])) * 1440 == pytest.approx(2.0, abs=1e-3)
    # flux and its error: positive where defined, error well below the flux
    f, e = cols["PDCSAP_FLUX"], cols["PDCSAP_FLUX_ERR"]
    good = ~np.isnan(f) & ~np.isnan(e)
    assert good.sum() == 18774
    assert np.all(f[good] > 0) and np.all(e[good] > 0) and np.median(e[good] / f[good]) < 0.01
    assert h[0][0]["OBJECT"] == "TIC 21132157"
    assert isinstance(h[0][0]["RADIUS"], float) and isinstance(h[0][0]["LOGG"], float)


def test_read_lightcurve_cleans_and_normalises():
    t, y, dy, hdr = fits.read_lightcurve(FILE)
    assert len(t) == len(y) == len(dy) == 18774
    assert np.median(y) == 1.0
    assert not np.any(np.isnan(t) | np.isnan(y) | np.isnan(dy))
    assert 0.98 < y.min() and y.max() < 1.02
    assert hdr["OBJECT"] == "TIC 21132157" and hdr["XTENSION"] == "BINTABLE"
