"""CPU ORACLE (test infrastructure only) of the history buffer updates.

numpy restatement of histFileMod.F90 for fields kept at the resolution of their
source (the branch of hist_update_hbuf_field_1d / _2d that does not map to the
gridcell, :1612-1763 and :2035-2169), hfields_normalize (:2302-2358) and
hfields_zero (:2361-2381).  Arrays are (num2d, n1d), point fastest as the
reference's hbuf(k, j); `active` is a boolean array over the points or None.
Each element is an independent sequence of IEEE operations, so elementwise
numpy gives the reference's values.  Pinned by tests/test_oracle_history.py
against closed forms (means of exactly representable series, maxima, minima,
missing and inactive points).
"""
import numpy as np

SPVAL = 1.e36


def hist_update(hbuf, nacs, field, active, avgflag):
    """One step of hist_update_hbuf for one field (in place)."""
    valid = np.ones(field.shape, dtype=bool) if active is None else np.broadcast_to(active, field.shape).copy()
    good = valid & (field != SPVAL)
    first = nacs == 0
    if avgflag == "I":
        hbuf[...] = np.where(valid, np.where(field != SPVAL, field, SPVAL), SPVAL)
        nacs[...] = 1
    elif avgflag in ("A", "S"):
        h = np.where(good & first, 0.0, hbuf)
        h = np.where(good, h + field, h)
        h = np.where(~good & first, SPVAL, h)
        nacs[...] = np.where(good, nacs + 1, nacs)
        hbuf[...] = h
    elif avgflag in ("X", "M"):
        start = -1.e50 if avgflag == "X" else 1.e50
        h = np.where(good & first, start, hbuf)
        if avgflag == "X":
            h = np.where(good, np.where(field > h, field, h), h)
        else:
            h = np.where(good, np.where(field < h, field, h), h)
        h = np.where(~good & first, SPVAL, h)
        hbuf[...] = h
        nacs[...] = 1
    else:
        raise ValueError(avgflag)


def hist_normalize(hbuf, nacs, avgflag):
    """hfields_normalize: 'A' buffers divided by float(nacs) where nacs /= 0."""
    if avgflag == "A":
        nz = nacs != 0
        hbuf[nz] = hbuf[nz] / nacs[nz].astype(np.float32).astype(np.float64)


def hist_zero(hbuf, nacs):
    hbuf[...] = 0.0
    nacs[...] = 0
