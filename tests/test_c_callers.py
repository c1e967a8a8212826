"""Compiled C / C++ callers of the drop-in boundary.

CPU: the reference's header and this build's header compile together in
one translation unit (either order, C and C++, -Wall -Wextra -Werror), and
a C++ program written like the reference's test/versioninfo.cpp links
libtopotoolbox_amd.so, calls has_topotoolbox() (no GPU call) and prints the
reference's version macros -- built once against this build's header and,
where the reference is present, against the reference's own topotoolbox.h
(a caller relinked with no source change).

GPU: a C99 caller (tests/c/dropin_caller.c) calls fillsinks, reconstruct
and excesstopography_fsm2d on host buffers; its outputs are compared with
the oracle.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, assert_bitwise

INCLUDE = os.path.join(ROOT, "include")
CDIR = os.path.join(ROOT, "tests", "c")
REF_INCLUDE = "/root/reference/include"


def _libdir():
    from libtopotoolbox_amd.build import OUT_DIR, LIB_PATH

    if not os.path.exists(LIB_PATH):
        pytest.fail(f"{LIB_PATH} is missing: build the library first")
    return OUT_DIR


def _run(cmd, **kw):
    res = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert res.returncode == 0, f"{' '.join(cmd)}\n{res.stdout}\n{res.stderr}"
    return res.stdout


def _link_flags():
    d = _libdir()
    return [f"-L{d}", "-ltopotoolbox_amd", f"-Wl,-rpath,{d}"]


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_INCLUDE, "topotoolbox.h")),
                    reason="reference headers absent")
@pytest.mark.parametrize("lang", ["c", "c++"])
@pytest.mark.parametrize("order", ["ref_first", "amd_first"])
def test_headers_compile_together(tmp_path, lang, order):
    inc = ['#include "topotoolbox.h"', '#include "topotoolbox_amd.h"']
    if order == "amd_first":
        inc.reverse()
    src = tmp_path / ("both.c" if lang == "c" else "both.cpp")
    body = "\n".join(inc) + """
int probe(float *o, float *d, uint8_t *b, ptrdiff_t *dims) {
  fillsinks(o, d, b, dims);
  return has_topotoolbox() + TOPOTOOLBOX_VERSION_MAJOR;
}
"""
    if lang == "c++":
        body = 'extern "C" {\n' + "\n".join(inc) + "\n}\n" + body.split("\n", 2)[2]
    src.write_text(body)
    cc = "gcc" if lang == "c" else "g++"
    std = "-std=c99" if lang == "c" else "-std=c++17"
    _run([cc, std, "-Wall", "-Wextra", "-Werror", "-I", REF_INCLUDE, "-I", INCLUDE, "-c",
          str(src), "-o", str(tmp_path / "both.o")])


@pytest.mark.parametrize("header", ["amd", "reference"])
def test_versioninfo_links_and_runs(tmp_path, header):
    if header == "reference" and not os.path.isfile(os.path.join(REF_INCLUDE, "topotoolbox.h")):
        pytest.skip("reference headers absent")
    exe = tmp_path / "versioninfo"
    flags = ["-I", REF_INCLUDE, "-DUSE_REFERENCE_HEADER"] if header == "reference" else \
        ["-I", INCLUDE]
    _run(["g++", "-std=c++17", "-Wall", "-Werror", *flags, os.path.join(CDIR, "versioninfo.cpp"),
          "-o", str(exe), *_link_flags()])
    out = _run([str(exe)])
    assert "topotoolbox v3.0.0" in out
    assert "openmp v0" in out


@pytest.mark.gpu
def test_c_dropin_caller(tmp_path, oracle_lib):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    exe = tmp_path / "dropin_caller"
    _run(["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-I", INCLUDE,
          os.path.join(CDIR, "dropin_caller.c"), "-o", str(exe), *_link_flags()])
    ny, nx = 301, 413
    rng = np.random.default_rng(77)
    dem = oracle_lib.grid_random(77, (ny, nx))
    dem[rng.uniform(size=dem.shape) < 0.01] = np.nan
    bc = np.zeros(dem.shape, np.uint8)
    bc[0, :] = bc[-1, :] = 1
    bc[:, 0] = bc[:, -1] = 1
    slopes = rng.uniform(0.1, 0.9, size=dem.shape).astype(np.float32)
    pin, pout = str(tmp_path / "in"), str(tmp_path / "out")
    dem.tofile(pin + ".dem")
    bc.tofile(pin + ".bc")
    slopes.tofile(pin + ".slopes")
    print(f"dropin_caller {nx} x {ny}", flush=True)
    out = _run([str(exe), str(nx), str(ny), pin, pout])
    assert "ok" in out
    load = lambda ext: np.fromfile(pout + "." + ext, np.float32).reshape(ny, nx)  # noqa: E731
    assert_bitwise(load("fill"), oracle_lib.fillsinks_pf(dem, bc), "C caller fillsinks")
    marker = (dem - np.float32(10.0)).astype(np.float32)
    want_r, _, converged = oracle_lib.reconstruct(marker, dem)
    assert converged
    assert_bitwise(load("recon"), want_r, "C caller reconstruct")
    u64, _ = oracle_lib.fsm2d_f64(dem, slopes, 30.0)
    u = load("excess")
    ok = ~np.isnan(u64)
    np.testing.assert_array_equal(np.isnan(u), ~ok)
    err = np.abs(u[ok] - u64[ok]) / np.maximum(np.abs(u64[ok]), 1.0)
    assert err.max() <= 1e-4
