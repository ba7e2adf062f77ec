"""Comparison helpers shared by the parity tests."""
import numpy as np

NS = 12   # nlevsno


def relerr(a, b):
    """Elementwise |a-b| / max(|b|, 1e-12 * max|b|): relative error with a
    floor at 1e-12 of the field's scale so exact zeros compare sanely."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.nanmax(np.abs(b)) if b.size else 0.0
    floor = max(scale * 1e-12, 1e-300)
    with np.errstate(invalid="ignore"):
        return np.abs(a - b) / np.maximum(np.abs(b), floor)


def max_relerr(a, b, mask=None):
    e = relerr(a, b)
    if mask is not None:
        e = e[mask]
    if e.size == 0:
        return 0.0
    both_nan = np.isnan(np.asarray(a, dtype=np.float64)) & np.isnan(np.asarray(b, dtype=np.float64))
    if mask is not None:
        both_nan = both_nan[mask]
    e = np.where(both_nan, 0.0, e)
    return float(np.nanmax(np.where(np.isnan(e), np.inf, e)))


def active_sno_mask(st):
    """Mask of [level, column] entries of SNO arrays that the step defines
    (snow layers snl+1..0 and ground layers)."""
    snl = st.col["snl"]
    nl = st.col["t_soisno"].shape[0]
    j = np.arange(nl)[:, None] - NS + 1          # level index of each row
    return j >= (snl[None, :] + 1)


def soil_mask(st, nlev=20):
    nb = st.col["nbedrock"]
    j = np.arange(st.col["watsat"].shape[0])[:, None] + 1
    return j <= nb[None, :]


STATE_FIELDS = ["h2osoi_liq", "h2osoi_ice", "t_soisno"]


def compare_states(g, o, tol, fields=None, report=None):
    """Max relative error per field over the entries the step defines."""
    out = {}
    msk_sno = active_sno_mask(o)
    msk_soi = soil_mask(o)
    for name in fields or STATE_FIELDS:
        a, b = g.col[name], o.col[name]
        if a.ndim == 1:
            m = None
        elif a.shape[0] == msk_sno.shape[0]:
            m = msk_sno
        else:
            m = np.zeros(a.shape, bool)
            m[:msk_soi.shape[0]] = msk_soi
        out[name] = max_relerr(a, b, m)
    if report is not None:
        report.update(out)
    bad = {k: v for k, v in out.items() if not v <= tol}
    assert not bad, f"relative error above {tol}: {bad}"
    return out
