"""CPU-side checks of the C-ABI library: it loads without a GPU and exports
every entry point declared in include/liggghts_amd.h; the Python binding
covers all of them; and the product path refuses to run without the library
(no CPU fallback)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "liggghts_amd.h")
LIB = os.path.join(ROOT, "liggghts-public_amd", "libdem_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dem_[a-z_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("dem_create", "dem_set_particles", "dem_set_materials", "dem_run",
                 "dem_get_state", "dem_get_pairs", "dem_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (make -C liggghts-public_amd/csrc)"
    lib = ctypes.CDLL(LIB)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (dem_[a-z_]+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_python_binding_covers_the_header():
    import pydem
    assert sorted(pydem.EXPORTED) == declared_functions()


def test_create_without_gpu_reports_an_error_not_a_crash():
    """On a machine without a GPU dem_create must fail with DEM_ERR_NODEV."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import ctypes, pydem\n"
        "lib = pydem.load_library()\n"
        "p = pydem.DemParams(); p.dt = 1e-5; p.skin = 1e-3; p.cdf = 1.0\n"
        "for d in range(3): p.boxhi[d] = 1.0\n"
        "h = ctypes.c_void_p()\n"
        "rc = lib.dem_create(ctypes.byref(p), ctypes.byref(h))\n"
        "print(rc)\n" % os.path.join(ROOT, "liggghts-public_amd"))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    assert int(res.stdout.strip().splitlines()[-1]) == -6


def test_missing_library_fails_loudly(tmp_path):
    import pydem
    with pytest.raises(RuntimeError):
        pydem._lib_saved = pydem._lib
        try:
            pydem._lib = None
            pydem.load_library(str(tmp_path / "nope.so"))
        finally:
            pydem._lib = pydem._lib_saved
