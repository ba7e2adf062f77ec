"""Checkpoint and resume of device-resident state on the GPU: a run resumed
from a checkpoint continues bit for bit as the uninterrupted run; a damaged
or mismatched file is refused (CTSM_E_IO)."""
import numpy as np
import pytest

from ctsm_amd import restart
from ctsm_amd import synthetic as S
from ctsm_amd.lib import CTSM_E_IO, CtsmError
from ctsm_amd.state import DeviceState
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rctx():
    from ctsm_amd.lib import Context
    c = Context()
    c.set_canopy_params(S.canopy_params(), S.pft_table())
    yield c
    c.close()


def _offset(raw, name):
    import struct
    (n,) = struct.unpack("<I", bytes(raw[8:12]))
    for i in range(n):
        e = bytes(raw[16 + i * 88:16 + (i + 1) * 88])
        if e[:64].split(b"\0", 1)[0].decode() == name:
            return struct.unpack("<Q", e[72:80])[0]
    raise KeyError(name)


def _same(a, b):
    for grp in ("col", "pch", "can", "sf", "hy"):
        x, y = getattr(a, grp), getattr(b, grp)
        assert set(x) == set(y), grp
        for k in x:
            np.testing.assert_array_equal(x[k], y[k], err_msg=f"{grp}.{k}")


def test_resume_is_bitwise(rctx, tmp_path):
    st = S.canopy_workload("canopy_uneven", ncol=384, seed=0xC75D0051)
    print(f"case: checkpoint after 2 full land steps, resume 2 more; 384 columns / {st.dims.npatch} patches",
          flush=True)
    hydro = O.hydro_params()
    path = tmp_path / "state.ckpt"
    ds = DeviceState(st, rctx)
    try:
        for _ in range(2):
            rctx.land_step_full_dev(ds, hydro=hydro, water_table=True)
        restart.checkpoint(rctx, ds, path, extra={"step": 2})
        for _ in range(2):
            rctx.land_step_full_dev(ds, hydro=hydro, water_table=True)
        rctx.sync()
        a = ds.to_host()
    finally:
        ds.free()
    ds2, extra = restart.restore(rctx, path)
    try:
        assert extra == {"step": 2}
        for _ in range(2):
            rctx.land_step_full_dev(ds2, hydro=hydro, water_table=True)
        rctx.sync()
        b = ds2.to_host()
    finally:
        ds2.free()
    _same(a, b)


def test_damaged_or_mismatched_checkpoint_is_refused(rctx, tmp_path):
    st = S.canopy_workload("canopy", ncol=64, seed=0xC75D0052)
    print("case: checkpoint integrity checks, 64 columns", flush=True)
    path = tmp_path / "state.ckpt"
    ds = DeviceState(st, rctx)
    try:
        restart.checkpoint(rctx, ds, path)
        # one flipped byte inside the soil temperatures
        raw = bytearray(path.read_bytes())
        raw[_offset(raw, "c/t_soisno") + 1000] ^= 0x01
        bad = tmp_path / "bad.ckpt"
        bad.write_bytes(bytes(raw))
        with pytest.raises(CtsmError) as ei:
            restart.restore(rctx, bad)
        assert ei.value.code == CTSM_E_IO
        # a field of another size, and a name the file does not hold
        ptr = ds.col["t_soisno"]
        nbytes = 8 * (st.dims.nlevsno + st.dims.nlevgrnd) * st.dims.ncol
        with pytest.raises(CtsmError) as ei:
            rctx.checkpoint_read(path, [("c/t_soisno", ptr, nbytes - 8)])
        assert ei.value.code == CTSM_E_IO
        with pytest.raises(CtsmError) as ei:
            rctx.checkpoint_read(path, [("c/no_such_field", ptr, nbytes)])
        assert ei.value.code == CTSM_E_IO
        # the intact file still reads
        rctx.checkpoint_read(path, [("c/t_soisno", ptr, nbytes)])
    finally:
        ds.free()


def test_checkpoint_of_raw_ranges(rctx, tmp_path):
    """Ranges larger than one staging chunk and ranges whose size is a multiple
    of the 4 KiB alignment (the last one ending the file) come back intact."""
    rng = np.random.default_rng(17)
    sizes = [40 << 20, 12345, 4096, 3 << 20]
    print(f"case: checkpoint of raw device ranges {sizes}", flush=True)
    host = [rng.integers(0, 256, n, dtype=np.uint8) for n in sizes]
    dev = [rctx.alloc(n) for n in sizes]
    back = [rctx.alloc(n) for n in sizes]
    try:
        for p, h in zip(dev, host):
            rctx.h2d(p, h)
        path = tmp_path / "raw.ckpt"
        rctx.checkpoint_write(path, [(f"r{i}", p, n) for i, (p, n) in enumerate(zip(dev, sizes))])
        rctx.checkpoint_read(path, [(f"r{i}", p, n) for i, (p, n) in enumerate(zip(back, sizes))])
        for p, h in zip(back, host):
            g = np.empty_like(h)
            rctx.d2h(g, p)
            assert np.array_equal(g, h)
    finally:
        for p in dev + back:
            rctx.free(p)
