"""Builds copies of a C source with one planted mistake, so a test can show
that its check fails on that mistake (a check that a plausible slip passes
pins nothing).  Only used by the CPU tests; the copies are compiled into the
test's temporary directory with the flags of the real build."""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_CFLAGS = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fvisibility=hidden", "-w"]


def _plant(text: str, old: str, new: str, what: str) -> str:
    n = text.count(old)
    if n != 1:
        raise AssertionError(f"{what}: the text to replace occurs {n} times (expected once): {old!r}")
    return text.replace(old, new)


def oracle_mutant(tmp_path, old: str, new: str, name: str = "mut") -> str:
    """oracle.c with `old` replaced by `new` (exactly one occurrence),
    built like oracle/Makefile builds liboracle.so; returns the library path."""
    d = os.path.join(str(tmp_path), name)
    os.makedirs(os.path.join(d, "oracle"), exist_ok=True)
    inc = os.path.join(d, "include")
    if not os.path.exists(inc):
        os.symlink(os.path.join(ROOT, "include"), inc)
    src = open(os.path.join(ROOT, "oracle", "oracle.c")).read()
    open(os.path.join(d, "oracle", "oracle.c"), "w").write(_plant(src, old, new, "oracle.c"))
    shutil.copy(os.path.join(ROOT, "oracle", "adjoint.c"), os.path.join(d, "oracle", "adjoint.c"))
    out = os.path.join(d, "liboracle_mut.so")
    subprocess.run(["gcc", *ORACLE_CFLAGS, "-shared", "-o", out, os.path.join(d, "oracle", "oracle.c"),
                    os.path.join(d, "oracle", "adjoint.c"), "-lm", "-lpthread"], check=True)
    return out


def host_mutant(tmp_path, source: str, old: str, new: str, name: str = "mut") -> str:
    """One host-only C++ file of the library (e.g. csrc/camera.cpp) with a
    planted mistake, built alone into a shared library; returns its path."""
    d = os.path.join(str(tmp_path), name)
    os.makedirs(d, exist_ok=True)
    path = os.path.join(ROOT, "attribute-aware-rbfs_amd", "csrc", source)
    dst = os.path.join(d, source)
    open(dst, "w").write(_plant(open(path).read(), old, new, source))
    out = os.path.join(d, "lib" + os.path.splitext(source)[0] + "_mut.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-shared", "-I",
                    os.path.join(ROOT, "include"), "-o", out, dst], check=True)
    return out
