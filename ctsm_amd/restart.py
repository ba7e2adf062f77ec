"""Checkpoint and resume of device-resident state.

In the drop-in path the Fortran host keeps writing its restart files
(restFileMod.F90:74-156) from its own arrays.  A device-resident run (a
batched or served ensemble) holds its state in HBM instead; `checkpoint`
streams every array of a DeviceState to one file through the library
(ctsm_checkpoint_write: pinned double-buffered copies, per-array checksums),
with the dimensions and array shapes in a small JSON record, and `restore`
rebuilds the DeviceState from such a file (ctsm_checkpoint_read).  A resumed
run continues bit for bit as the uninterrupted one.
"""
import json
import struct
from types import SimpleNamespace

import numpy as np

from .state import DeviceState, Dims

_GROUPS = (("c", "col"), ("p", "pch"), ("g", "can"), ("s", "sf"), ("y", "hy"))
_META = "meta.json"
_ENTRY = 64 + 3 * 8


def _fields(ds):
    out = []
    for g, attr in _GROUPS:
        for k, ptr in getattr(ds, attr).items():
            shape, dt = ds._shapes[(g, k)]
            out.append((f"{g}/{k}", ptr, int(np.prod(shape)) * np.dtype(dt).itemsize, list(shape), np.dtype(dt).str))
    return out


def checkpoint(ctx, ds, path, extra=None):
    """Write every array of `ds` (after all enqueued work) to `path`.
    `extra` (JSON-serialisable) is stored with the dimensions, e.g. a step count."""
    fields = _fields(ds)
    d = ds.dims
    meta = {"dims": {"ncol": d.ncol, "npatch": d.npatch, "ngrid": d.ngrid, "nlevsno": d.nlevsno,
                     "nlevsoi": d.nlevsoi, "nlevgrnd": d.nlevgrnd},
            "arrays": {n: [shape, dt] for n, _, _, shape, dt in fields}, "extra": extra}
    blob = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8).copy()
    mptr = ctx.alloc(max(blob.nbytes, 8))
    try:
        ctx.h2d(mptr, blob)
        ctx.checkpoint_write(path, [(_META, mptr, blob.nbytes)] + [(n, p, b) for n, p, b, _, _ in fields])
    finally:
        ctx.free(mptr)


def read_meta(path):
    """The JSON record of a checkpoint file (read directly from the file)."""
    with open(path, "rb") as f:
        head = f.read(16)
        if head[:8] != b"CTSMCKP1":
            raise ValueError(f"{path}: not a checkpoint file")
        (n,) = struct.unpack("<I", head[8:12])
        tab = f.read(n * _ENTRY)
        for i in range(n):
            e = tab[i * _ENTRY:(i + 1) * _ENTRY]
            name = e[:64].split(b"\0", 1)[0].decode()
            nbytes, off, _ = struct.unpack("<QQQ", e[64:])
            if name == _META:
                f.seek(off)
                return json.loads(f.read(nbytes).decode())
    raise ValueError(f"{path}: no {_META} record")


def restore(ctx, path):
    """A new DeviceState holding the state saved in `path`, and the saved `extra`."""
    meta = read_meta(path)
    dm = meta["dims"]
    dims = Dims(dm["ncol"], dm["npatch"], dm["nlevsno"], dm["nlevsoi"], dm["nlevgrnd"], dm["ngrid"])
    host = SimpleNamespace(dims=dims, col={}, pch={}, can={}, sf={}, hy={})
    attr = dict(_GROUPS)
    for name, (shape, dt) in meta["arrays"].items():
        g, k = name.split("/", 1)
        getattr(host, attr[g])[k] = np.zeros(shape, dtype=np.dtype(dt))
    ds = DeviceState(host, ctx)
    try:
        ctx.checkpoint_read(path, [(n, p, b) for n, p, b, _, _ in _fields(ds)])
    except Exception:
        ds.free()
        raise
    return ds, meta.get("extra")
