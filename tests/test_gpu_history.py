"""History buffers on the GPU against the oracle, bit for bit: the kernels on
planted series (every flag, missing values, inactive points, level strides,
more fields than one launch takes), and a history tape over full land steps
written to netCDF."""
import ctypes as C

import numpy as np
import pytest

from ctsm_amd import abi
from ctsm_amd import synthetic as S
from ctsm_amd.state import DeviceState
from oracle import history as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SPVAL = H.SPVAL


def test_hist_kernels_bitwise(ctx):
    rng = np.random.default_rng(21)
    flags = "ASIXM"
    nf = 30                                        # more than one launch's batch
    specs = []
    for i in range(nf):
        n1d = int(rng.integers(1, 3000))
        num2d = int(rng.integers(1, 4))
        ld = n1d + int(rng.integers(0, 7)) if num2d > 1 else n1d
        specs.append((n1d, num2d, ld, flags[i % 5], i % 3 == 0))
    print(f"case: history kernels, {nf} fields, 6 steps", flush=True)
    dev, host, desc = [], [], []
    for n1d, num2d, ld, flag, use_active in specs:
        src = ctx.alloc(8 * ld * num2d)
        hb = ctx.alloc(8 * n1d * num2d)
        na = ctx.alloc(4 * n1d * num2d)
        act = None
        act_h = None
        if use_active:
            act_h = (rng.random(n1d) > 0.2).astype(np.int32)
            act = ctx.alloc(4 * n1d)
            ctx.h2d(act, act_h)
        dev.append((src, hb, na, act))
        host.append([np.zeros((num2d, n1d)), np.zeros((num2d, n1d), dtype=np.int32), act_h])
        desc.append(abi.HistField(src, act, hb, na, ld, n1d, num2d, ord(flag), 0))
    arr = (abi.HistField * nf)(*desc)
    try:
        ctx.hist_zero(arr)
        for step in range(6):
            for (n1d, num2d, ld, flag, _), (src, _, _, _), hs in zip(specs, dev, host):
                x = rng.normal(280.0, 30.0, (num2d, ld))
                x[rng.random((num2d, ld)) < 0.1] = SPVAL
                ctx.h2d(src, np.ascontiguousarray(x))
                H.hist_update(hs[0], hs[1], x[:, :n1d], None if hs[2] is None else hs[2] != 0, flag)
            ctx.hist_update(arr)
            ctx.sync()
        ctx.hist_normalize(arr)
        for (n1d, num2d, ld, flag, _), (src, hb, na, _), hs in zip(specs, dev, host):
            H.hist_normalize(hs[0], hs[1], flag)
            g = np.empty((num2d, n1d))
            n = np.empty((num2d, n1d), dtype=np.int32)
            ctx.d2h(g, hb)
            ctx.d2h(n, na)
            assert np.array_equal(g, hs[0]), flag
            assert np.array_equal(n, hs[1]), flag
        ctx.hist_zero(arr)
        g = np.empty((specs[0][1], specs[0][0]))
        ctx.d2h(g, dev[0][1])
        assert np.all(g == 0.0)
    finally:
        for t in dev:
            for p in t:
                if p is not None:
                    ctx.free(p)


def test_history_tape_over_land_steps(tmp_path):
    from scipy.io import netcdf_file
    from ctsm_amd.history import HistoryTape, DEFAULT_FIELDS, device_field
    from ctsm_amd.lib import Context
    c = Context()
    c.set_canopy_params(S.canopy_params(), S.pft_table())
    st = S.canopy_workload("canopy_uneven", ncol=256, seed=0xC75D0041)
    print(f"case: history tape over 4 full land steps, 256 columns / {st.dims.npatch} patches", flush=True)
    ds = DeviceState(st, c)
    tape = HistoryTape(c, ds)
    hydro = O.hydro_params()
    d = st.dims
    ref = {}
    try:
        for step in range(4):
            c.land_step_full_dev(ds, hydro=hydro, water_table=True)
            tape.update()
            h = ds.to_host()
            for name, group, key, levels, flag, _, _ in DEFAULT_FIELDS:
                arr = {"c": h.col, "p": h.pch, "g": h.can, "s": h.sf, "y": h.hy}
                _, _, kind = device_field(ds, group, key)
                from ctsm_amd.state import CANOPY_ALIAS, SOILFLUX_ALIAS, SURFWATER_ALIAS
                g_, k_ = group, key
                for _ in range(3):
                    if g_ == "g" and k_ in CANOPY_ALIAS:
                        g_, k_ = CANOPY_ALIAS[k_]
                    elif g_ == "s" and k_ in SOILFLUX_ALIAS:
                        g_, k_ = SOILFLUX_ALIAS[k_]
                    elif g_ == "y" and k_ in SURFWATER_ALIAS:
                        g_, k_ = SURFWATER_ALIAS[k_]
                x = arr[g_][k_]
                if levels is None:
                    x = x[None, :]
                else:
                    lk, first, count = levels
                    num2d = {"levgrnd": d.nlevgrnd, "levsoi": d.nlevsoi}[count]
                    off = d.nlevsno + first - 1 if lk == "sno" else first - 1
                    x = x[off:off + num2d]
                patch = kind in ("p", "soip", "lyr")
                act = (h.can["active"] != 0) if patch else None
                if name not in ref:
                    ref[name] = [np.zeros(x.shape), np.zeros(x.shape, dtype=np.int32)]
                H.hist_update(ref[name][0], ref[name][1], x, act, flag)
        bufs = tape.buffers()
        for name, (hb, na) in bufs.items():
            assert np.array_equal(np.atleast_2d(hb), ref[name][0]), name
            assert np.array_equal(np.atleast_2d(na), ref[name][1]), name
        path = tmp_path / "hist.nc"
        tape.write(path, time=4 * 1800.0)
        with netcdf_file(str(path), "r", mmap=False) as f:
            for name, (hb, na) in ref.items():
                H.hist_normalize(hb, na, "A")
                v = f.variables[name][0]
                assert np.array_equal(np.atleast_2d(v), hb), name
            assert f.variables["TSOI"].units.decode() == "K"
        assert tape.nsteps == 0
        b = tape.buffers()
        assert all(np.all(v[1] == 0) for v in b.values())
    finally:
        tape.free()
        ds.free()
        c.close()


def test_hist_rejects_bad_descriptors(ctx):
    from ctsm_amd.lib import CTSM_E_ARG, CtsmError
    print("case: history descriptors with a bad flag / a short level stride", flush=True)
    hb, na, src = ctx.alloc(8 * 20), ctx.alloc(4 * 20), ctx.alloc(8 * 20)
    try:
        for d in (abi.HistField(src, None, hb, na, 10, 10, 2, ord("L"), 0),
                  abi.HistField(src, None, hb, na, 5, 10, 2, ord("A"), 0),
                  abi.HistField(None, None, hb, na, 10, 10, 2, ord("A"), 0)):
            arr = (abi.HistField * 1)(d)
            with pytest.raises(CtsmError) as ei:
                ctx.hist_update(arr)
            assert ei.value.code == CTSM_E_ARG
    finally:
        for p in (hb, na, src):
            ctx.free(p)
