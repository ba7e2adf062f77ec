"""The checkpoint file format, on the CPU: the writer and reader of
ctsm_amd/csrc/checkpoint.hip built by g++ against a host-only stand-in of the
HIP calls (tests/host_hip), the file parsed independently here -- magic, field
table, 4 KiB-aligned ranges, FNV-1a checksums over 64-bit words -- with fields
larger than one 32 MiB staging chunk, empty fields and a range ending exactly on
the alignment; a flipped byte or a wrong size is refused (CTSM_E_IO) with the
destination untouched; a write that is killed or fails partway leaves the
previous file whole.  Every case runs on two builds of the tool: -O1, and
-O1 with AddressSanitizer and UndefinedBehaviorSanitizer, which aborts the
run (and fails the case) on an out-of-bounds access, a use after free, a leak
or undefined arithmetic."""
import os
import resource
import shutil
import signal
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIZES = [100, (70 << 20) + 5, 0, 8, 12345, 3 << 20]     # the last ends on the 4 KiB alignment
CTSM_E_IO = 7


SANITIZE = ["-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


@pytest.fixture(scope="module", params=["plain", "asan"])
def tool(request, tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("ckpt_" + request.param)
    exe = d / "ckpt"
    extra = SANITIZE if request.param == "asan" else []
    subprocess.run(["g++", "-O1", "-std=c++17", *extra, "-I", os.path.join(ROOT, "tests", "host_hip"),
                    "-I", os.path.join(ROOT, "include"), "-x", "c++",
                    os.path.join(ROOT, "ctsm_amd", "csrc", "checkpoint.hip"), "-x", "c++",
                    os.path.join(ROOT, "tests", "host_hip", "ckpt_main.cpp"), "-o", str(exe)], check=True)
    return exe, d


def _run(exe, mode, path, sizes):
    out = subprocess.run([str(exe), mode, str(path)] + [str(s) for s in sizes], capture_output=True, text=True,
                         check=True).stdout.split()
    return int(out[-1])


def _fnv1a_words(b):
    h = 0xcbf29ce484222325
    m = (1 << 64) - 1
    n8 = len(b) // 8
    for w in np.frombuffer(b[:8 * n8], dtype="<u8").tolist():
        h = ((h ^ w) * 0x100000001b3) & m
    for x in b[8 * n8:]:
        h = ((h ^ x) * 0x100000001b3) & m
    return h


def _pattern(i, n):
    k = np.arange(n, dtype=np.uint64)
    return ((k * 131 + i * 7) & 0xff).astype(np.uint8).tobytes()


def test_file_layout_and_checksums(tool):
    exe, d = tool
    path = d / "a.ckpt"
    assert _run(exe, "w", path, SIZES) == 0
    raw = path.read_bytes()
    assert raw[:8] == b"CTSMCKP1"
    (n,) = struct.unpack("<I", raw[8:12])
    assert n == len(SIZES)
    ends = []
    for i in range(n):
        e = raw[16 + 88 * i:16 + 88 * (i + 1)]
        name = e[:64].split(b"\0", 1)[0].decode()
        size, off, csum = struct.unpack("<QQQ", e[64:])
        assert name == f"field{i}" and size == SIZES[i] and off % 4096 == 0
        data = raw[off:off + size]
        assert data == _pattern(i, size), name
        assert csum == _fnv1a_words(data), name
        ends.append(off + size)
    assert len(raw) >= max(ends)
    assert _run(exe, "r", path, SIZES) == 0


def test_damaged_or_mismatched_file_is_refused(tool):
    exe, d = tool
    path = d / "b.ckpt"
    assert _run(exe, "w", path, SIZES) == 0
    raw = bytearray(path.read_bytes())
    e = raw[16 + 88 * 5:16 + 88 * 6]
    size, off, _ = struct.unpack("<QQQ", bytes(e[64:]))
    raw[off + size - 1] ^= 0x40                       # the last byte of the range that ends the file
    bad = d / "bad.ckpt"
    bad.write_bytes(bytes(raw))
    assert _run(exe, "r", bad, SIZES) == CTSM_E_IO
    assert _run(exe, "r", path, SIZES[:-1] + [SIZES[-1] - 8]) == CTSM_E_IO


def _limited(limit, ignore_xfsz):
    """preexec_fn: files this process writes stop at `limit` bytes; past it
    the write is killed by SIGXFSZ (a crash mid-write) or, with the signal
    ignored, fails with EFBIG (a full disk)."""
    def f():
        resource.setrlimit(resource.RLIMIT_FSIZE, (limit, limit))
        signal.signal(signal.SIGXFSZ, signal.SIG_IGN if ignore_xfsz else signal.SIG_DFL)
    return f


@pytest.mark.parametrize("ignore_xfsz", [False, True], ids=["killed", "write-fails"])
def test_interrupted_write_keeps_the_previous_file(tool, ignore_xfsz):
    exe, d = tool
    path = d / f"keep_{int(ignore_xfsz)}.ckpt"
    assert _run(exe, "w", path, SIZES) == 0
    before = path.read_bytes()
    big = [80 << 20, 3 << 20]                       # a larger checkpoint, cut off at 16 MiB
    r = subprocess.run([str(exe), "w", str(path)] + [str(s) for s in big], capture_output=True, text=True,
                       preexec_fn=_limited(16 << 20, ignore_xfsz))
    if ignore_xfsz:
        assert r.returncode == 0 and int(r.stdout.split()[-1]) == CTSM_E_IO, r.stdout + r.stderr
        assert not os.path.exists(str(path) + ".tmp")     # the failed write cleans up after itself
    else:
        assert r.returncode == -signal.SIGXFSZ, (r.returncode, r.stderr)
    assert path.read_bytes() == before
    assert _run(exe, "r", path, SIZES) == 0


def test_refused_file_leaves_the_destination_untouched(tool):
    exe, d = tool
    path = d / "c.ckpt"
    assert _run(exe, "w", path, SIZES) == 0
    raw = bytearray(path.read_bytes())
    e = raw[16 + 88 * 5:16 + 88 * 6]                 # damage the last field: every earlier one checks out
    size, off, _ = struct.unpack("<QQQ", bytes(e[64:]))
    raw[off + size // 2] ^= 0x01
    bad = d / "c_bad.ckpt"
    bad.write_bytes(bytes(raw))
    # exit 0 and CTSM_E_IO: ckpt_main exits 8 if any destination byte changed
    assert _run(exe, "r", bad, SIZES) == CTSM_E_IO
