"""The scratch-free solve (cpgpu_poisson2d_solve_static) against the
allocating one-shot solve (cpgpu_poisson2d_solve) through the C-ABI: two
calls in a row on one GPU, the second on a smaller grid, each bitwise equal
to the allocating entry point (Jacobi-PCG is FP64 and deterministic), and
the size limit refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

import carpios_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu


def case(nx, ny, types, vals, seed):
    og = O.StructGrid(nx, ny, 1, 0.0, 1.3, 0.0, 1.0)
    src = og.field()
    og.interior(src)[:] = np.random.default_rng(seed).standard_normal((nx, ny))
    phi0 = og.field()
    og.interior(phi0)[:] = np.random.default_rng(seed + 1).standard_normal((nx, ny))
    desc = ca.uniform_grid_desc(nx, ny, 1, types, vals, xlim=(0.0, 1.3))
    return og, src, phi0, desc


def allocating(ctx, desc, src, phi0, method, max_iter, tol, norm):
    phi = phi0.copy()
    mi, t, hl = C.c_int(max_iter), C.c_double(tol), C.c_int(0)
    code = ca.lib().cpgpu_poisson2d_solve(ctx.h, C.byref(desc), ca._p(src), C.c_double(0.0),
                                          ca._p(phi), method, C.byref(mi), C.byref(t),
                                          norm, None, 0, C.byref(hl))
    return code, mi.value, t.value, phi


def static(ctx, desc, src, phi0, method, max_iter, tol, norm):
    # device arrays of the caller: the interior vectors of two problems
    # created for the purpose (the entry point itself allocates nothing)
    hs, hp = ca.Poisson(ctx, desc), ca.Poisson(ctx, desc)
    hs.set_phi(src)
    hp.set_phi(phi0)
    d_src = ca.lib().cpgpu_poisson_device_vec(hs.h, 0)
    d_phi = ca.lib().cpgpu_poisson_device_vec(hp.h, 0)
    mi, t = C.c_int(max_iter), C.c_double(tol)
    code = ca.lib().cpgpu_poisson2d_solve_static(ctx.h, C.byref(desc), C.c_void_p(d_src),
                                                 C.c_double(0.0), C.c_void_p(d_phi), method,
                                                 C.byref(mi), C.byref(t), norm)
    phi = hp.get_phi()
    hs.close()
    hp.close()
    return code, mi.value, t.value, phi


@pytest.mark.parametrize("order", ["large_then_small"])
def test_static_solve_matches_the_allocating_solve(ctx, order):
    runs = [((200, 150), [1, 2, 1, 2], [0.5, -0.25, 1.0, 0.0], 3, ca.JPCG, ca.NORM_REF),
            ((48, 40), [2, 1, 2, 2], [0.0, 2.0, -1.0, 0.5], 7, ca.CG, ca.NORM_TRUE)]
    for (nx, ny), types, vals, seed, method, norm in runs:
        og, src, phi0, desc = case(nx, ny, types, vals, seed)
        a = allocating(ctx, desc, src, phi0, method, 20000, 1e-10, norm)
        s = static(ctx, desc, src, phi0, method, 20000, 1e-10, norm)
        print(f"static solve {nx}x{ny} method {method}: allocating {a[:3]}, static {s[:3]}",
              flush=True)
        assert a[0] == s[0] == ca.OK
        assert a[1] == s[1] and a[2] == s[2]
        assert np.array_equal(og.interior(a[3]), og.interior(s[3]))


def test_static_solve_refuses_grids_beyond_its_scratch(ctx):
    og, src, phi0, desc = case(4097, 8, [1, 1, 1, 1], None, 1)
    mi, t = C.c_int(10), C.c_double(1e-8)
    code = ca.lib().cpgpu_poisson2d_solve_static(ctx.h, C.byref(desc), None, C.c_double(1.0),
                                                 C.c_void_p(1), ca.JPCG, C.byref(mi),
                                                 C.byref(t), ca.NORM_TRUE)
    assert code == ca.EINVAL
    assert b"exceed the static scratch" in ca.lib().cpgpu_last_error(ctx.h)
    og, src, phi0, desc = case(600, 600, [1, 1, 1, 1], None, 1)
    code = ca.lib().cpgpu_poisson2d_solve_static(ctx.h, C.byref(desc), None, C.c_double(1.0),
                                                 C.c_void_p(1), ca.JPCG, C.byref(mi),
                                                 C.byref(t), ca.NORM_TRUE)
    assert code == ca.EINVAL
