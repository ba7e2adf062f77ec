"""ctsm_canopyfluxes_dev_noalloc on the GPU, through the C-ABI: the same
results as ctsm_canopyfluxes_dev, bit for bit, on two calls in a row (the
second on a smaller batch, so the fixed scratch holds the first call's
leftovers), no device memory taken by the call, and a batch too large for
the fixed scratch refused with CTSM_E_SCRATCH before anything runs.

Each test prints its case before the GPU work starts.
"""
import ctypes as C

import numpy as np
import pytest

from ctsm_amd import synthetic as S
from ctsm_amd.lib import CTSM_E_SCRATCH, NOALLOC_SCRATCH_BYTES, CtsmError
from ctsm_amd.state import DeviceState

pytestmark = pytest.mark.gpu


def _free_device_bytes():
    """hipMemGetInfo's free bytes on the current device."""
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _run(ctx, st, noalloc, filt=None):
    ds = DeviceState(st, ctx)
    fdev = None
    try:
        if filt is not None:
            fdev = ctx.alloc(filt.nbytes)
            ctx.h2d(fdev, filt)
        ctx.sync()
        before = _free_device_bytes()
        n = None if filt is None else len(filt)
        if noalloc:
            ctx.canopyfluxes_dev_noalloc(ds, filterp_dev=fdev, n=n)
        else:
            ctx.canopyfluxes_dev(ds, filterp_dev=fdev, n=n)
        ctx.sync()
        after = _free_device_bytes()
        return ds.to_host(), before - after
    finally:
        if fdev is not None:
            ctx.free(fdev)
        ds.free()


def _same(a, b):
    for grp in ("col", "pch", "can"):
        for k, v in getattr(b, grp).items():
            np.testing.assert_array_equal(getattr(a, grp)[k], v, err_msg=f"{grp}.{k}")


def test_noalloc_equals_dev_on_two_calls():
    from ctsm_amd.lib import Context
    ctx = Context()
    try:
        ctx.set_canopy_params(S.canopy_params(), S.pft_table())
        # first call: 4,096 uneven columns, every patch; second: 700 columns
        # behind a descending filter of every third patch (smaller than the
        # first call, so the scratch still holds the first call's state)
        for ncol, step in ((4096, None), (700, 3)):
            st = S.canopy_workload("canopy_uneven", ncol=ncol)
            filt = None if step is None else np.arange(st.dims.npatch, 0, -step, dtype=np.int32)
            print(f"case: no-allocation CanopyFluxes, {ncol} columns, "
                  f"{st.dims.npatch if filt is None else len(filt)} patches", flush=True)
            # the allocating entry first: the runtime sizes its own spill
            # scratch for the kernel on its first dispatch
            g_dev, _ = _run(ctx, st.copy(), noalloc=False, filt=filt)
            g_na, taken = _run(ctx, st.copy(), noalloc=True, filt=filt)
            _same(g_na, g_dev)
            assert taken == 0, f"the no-allocation entry took {taken} bytes of device memory"
    finally:
        ctx.close()


def test_noalloc_refuses_a_batch_beyond_its_scratch():
    from ctsm_amd.lib import Context
    ctx = Context()
    try:
        ctx.set_canopy_params(S.canopy_params(), S.pft_table())
        st = S.canopy_workload("canopy", ncol=8)
        ds = DeviceState(st, ctx)
        n = NOALLOC_SCRATCH_BYTES // 136 + 1              # one patch past what the scratch holds
        # a filter of n valid entries (patch 1 throughout), so the call's
        # arguments are in bounds whatever it does with them
        filt = np.ones(n, dtype=np.int32)
        fdev = ctx.alloc(filt.nbytes)
        try:
            ctx.h2d(fdev, filt)
            print(f"case: no-allocation CanopyFluxes refused for {n} patches", flush=True)
            with pytest.raises(CtsmError) as e:
                ctx.canopyfluxes_dev_noalloc(ds, filterp_dev=fdev, n=n)
            assert e.value.code == CTSM_E_SCRATCH
            assert e.value.index * 1024 >= NOALLOC_SCRATCH_BYTES
            # nothing ran: the context still works and the state is unchanged
            ctx.sync()
            np.testing.assert_array_equal(ds.to_host().can["t_veg"], st.can["t_veg"])
        finally:
            ctx.free(fdev)
            ds.free()
    finally:
        ctx.close()
