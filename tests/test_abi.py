"""The C-ABI library loads and exports every symbol include/cpgpu.h declares
(no compute: runs without a GPU)."""
import ctypes as C
import os
import re

import numpy as np

import carpios_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "cpgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cpgpu_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ca.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(ca.EXPORTS) == syms


def test_invalid_arguments_are_rejected_without_a_device():
    L = ca.lib()
    assert L.cpgpu_create(C.c_int(0), None) == ca.EINVAL
    assert L.cpgpu_poisson_destroy(None) == ca.EINVAL
    c = np.zeros(4)
    assert L.cpgpu_grid_uniform_axis(0, 1, C.c_double(0), C.c_double(1),
                                     ca._p(c), ca._p(c)) == ca.EINVAL


def test_uniform_axis_matches_grid_construction():
    import oracle as O
    for n, gl, lo, hi in ((128, 1, 0.0, 1.0), (512, 2, -0.5, 0.5), (7, 3, 2.0, 5.5)):
        c, s = ca.uniform_axis(n, gl, lo, hi)
        a = O.Axis(n, gl, lo, hi)
        assert np.array_equal(c, a.c) and np.array_equal(s, a.s)


def test_packed_boundary_values_are_checked_on_the_host():
    # the library reads 2*(ny+2gl) + 2*(nx+2gl) values: anything else is
    # refused before it reaches the C side
    nx, ny, gl = 6, 4, 2
    full = ca.pack_bc_values(nx, ny, gl, [1.0, 2.0, 3.0, 4.0])
    assert full.shape == (2 * (ny + 2 * gl) + 2 * (nx + 2 * gl),)
    assert np.array_equal(ca.pack_bc_values(nx, ny, gl, full), full)
    import pytest
    for bad in (full[:-1], np.concatenate([full, [0.0]]), full.reshape(2, -1),
                [1.0, 2.0, 3.0]):
        with pytest.raises(ValueError):
            ca.pack_bc_values(nx, ny, gl, bad)


def test_cpp_bindings_compile():
    # the C++ host side a maintainer adds to the reference: the structured
    # helper needs only the C-ABI header; the algebra solvers' counterparts
    # compile against the reference's own lib/algebra where that tree is
    import shutil
    import subprocess
    import pytest
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    inc = os.path.join(ROOT, "include")
    r = subprocess.run([gxx, "-std=c++11", "-fsyntax-only", "-Wall", "-I" + inc,
                        os.path.join(ROOT, "tests", "cpp", "caller.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ref = "/root/reference/lib"
    if os.path.isdir(os.path.join(ref, "algebra")):
        r = subprocess.run([gxx, "-std=c++11", "-fsyntax-only", "-w", "-I" + ref, "-I" + inc,
                            os.path.join(ROOT, "tests", "cpp", "algebra_shim.cpp")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
