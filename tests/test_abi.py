"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/aarbf.h declares, its struct layout matches the bindings, and
the host-only importer honours the importParticles.h format."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import aarbf
from aarbf import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("aarbf.h",):
        for line in open(os.path.join(ROOT, "include", h)):
            if line.lstrip().startswith(("*", "/*", "#")):
                continue
            m = re.match(r"^\s*(?:const\s+)?(?:int|void|uint32_t|char)\s*\*?\s*(aarbf_[a-z0-9_]+)\s*\(", line)
            if m:
                names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_declared()) <= set(_lib.SIGNATURES)


def test_struct_layout_matches_header():
    lib = _lib.load()
    buf = (C.c_uint32 * 32)()
    m = lib.aarbf_abi_layout(buf, 32)
    assert list(buf[:m]) == _lib._expected_layout()


def test_status_strings():
    lib = _lib.load()
    assert lib.aarbf_status_string(0) == b"ok"
    assert lib.aarbf_status_string(5) == b"malformed particle file"


def test_context_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(aarbf.AarbfError) as e:
        aarbf.Context(0)
    assert e.value.status == 7  # AARBF_ERR_NO_DEVICE


def _export(path, steps):
    lib = _lib.load()
    names = (C.c_char_p * len(steps))(*[n.encode() for n, _ in steps])
    arrs = [np.ascontiguousarray(a, np.float32) for _, a in steps]
    ptrs = (C.POINTER(C.c_float) * len(steps))(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrs])
    counts = (C.c_uint64 * len(steps))(*[a.shape[0] for a in arrs])
    return lib.aarbf_export_particles(path.encode(), len(steps), names, ptrs, counts)


def _import(path):
    lib = _lib.load()
    h = C.c_void_p()
    st = lib.aarbf_import_particles(path.encode(), C.byref(h))
    if st:
        return st, None
    out = []
    for i in range(lib.aarbf_timesteps_count(h)):
        p = C.POINTER(C.c_float)()
        n = C.c_uint64()
        nm = C.c_char_p()
        assert lib.aarbf_timesteps_get(h, i, C.byref(p), C.byref(n), C.byref(nm)) == 0
        a = np.ctypeslib.as_array(p, shape=(n.value * 4,)).reshape(-1, 4).copy() if n.value else np.zeros((0, 4))
        out.append((nm.value.decode(), a))
    lib.aarbf_timesteps_destroy(h)
    return 0, out


def test_particle_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    steps = [("0", rng.normal(size=(100, 4))), ("10", rng.normal(size=(37, 4))), ("20", np.zeros((0, 4)))]
    path = str(tmp_path / "set.particles")
    assert _export(path, steps) == 0
    # the on-disk layout: header of names, then path_<name> = u64 count + float4s
    assert open(path).read().splitlines() == ["0", "10", "20"]
    raw = open(path + "_10", "rb").read()
    assert int.from_bytes(raw[:8], "little") == 37 and len(raw) == 8 + 37 * 16
    st, got = _import(path)
    assert st == 0
    for (n0, a0), (n1, a1) in zip(steps, got):
        assert n0 == n1 and np.array_equal(np.float32(a0), a1)


def test_particle_file_errors(tmp_path):
    path = str(tmp_path / "bad.particles")
    assert _import(str(tmp_path / "missing"))[0] == 4  # AARBF_ERR_IO
    open(path, "w").write("1\n")
    assert _import(path)[0] == 4  # timestep file missing
    with open(path + "_1", "wb") as f:
        f.write((10).to_bytes(8, "little") + b"\0" * 16 * 9)  # one particle short
    assert _import(path)[0] == 5  # AARBF_ERR_FORMAT
