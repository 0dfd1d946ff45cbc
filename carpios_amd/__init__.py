"""carpios_amd -- MI355X-native structured-grid projection hot path of Carpios.

The compute path is libcpgpu.so (hand-written HIP for gfx950 behind the C-ABI
of include/cpgpu.h).  This module is a thin ctypes view of that ABI for
Python callers (tests, bench); it has no compute of its own and raises if
the library is missing -- there is no CPU fallback.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpgpu.so")

# status codes / constants mirrored from include/cpgpu.h
OK, MAXITER, BREAKDOWN_RHO, BREAKDOWN_OMEGA = 0, 1, 2, 3
EINVAL, EHIP, EUNSUPPORTED, ENOMEM, EZERODIAG = -1, -2, -3, -4, -10
BC_DIRICHLET, BC_NEUMANN = 1, 2
CG, JPCG, BICGSTAB, JACOBI, MIXED = 0, 1, 2, 3, 4
NORM_REF, NORM_TRUE = 0, 1
OPT_PROJECT_MEAN = 1

MIXED_RESTART = 5
FCG_UNET = 6
METHODS = {"CG": CG, "JPCG": JPCG, "Dia_BiCGSTAB": BICGSTAB,
           "Jacobi": JACOBI, "MIXED": MIXED, "MIXED_RESTART": MIXED_RESTART,
           "FCG_UNET": FCG_UNET}
UNET_WARM, UNET_PRECOND = 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Grid2D(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("gl", C.c_int),
                ("cx", _dp), ("sx", _dp), ("cy", _dp), ("sy", _dp),
                ("beta", C.c_double), ("bc_type", C.c_int * 4),
                ("bc_val", _dp)]


class CpgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cpgpu error {code}: {msg}")
        self.code = code


_lib = None

# every symbol include/cpgpu.h declares
EXPORTS = [
    "cpgpu_grid_uniform_axis", "cpgpu_create", "cpgpu_destroy", "cpgpu_last_error", "cpgpu_synchronize",
    "cpgpu_stream", "cpgpu_poisson_create", "cpgpu_poisson_destroy",
    "cpgpu_poisson_set_source", "cpgpu_poisson_set_source_device",
    "cpgpu_poisson_set_phi", "cpgpu_poisson_get_phi", "cpgpu_poisson_solve",
    "cpgpu_poisson_set_mixed", "cpgpu_poisson_apply",
    "cpgpu_poisson_residual", "cpgpu_poisson_device_vec",
    "cpgpu_poisson_get_vec", "cpgpu_poisson_set_source_blobs",
    "cpgpu_blob_params", "cpgpu_poisson_fill_phi",
    "cpgpu_poisson2d_solve", "cpgpu_timer_enable", "cpgpu_timer_read",
    "cpgpu_ns_create", "cpgpu_ns_destroy", "cpgpu_ns_set_field",
    "cpgpu_ns_get_field", "cpgpu_ns_step", "cpgpu_ns_time",
    "cpgpu_ns_projection", "cpgpu_poisson_set_variant",
    "cpgpu_poisson_autotune", "cpgpu_poisson_kernel_times",
    "cpgpu_poisson_set_batch", "cpgpu_unet_create", "cpgpu_unet_destroy",
    "cpgpu_unet_num_params", "cpgpu_unet_get_params", "cpgpu_unet_set_params",
    "cpgpu_unet_forward", "cpgpu_unet_backward", "cpgpu_unet_adam",
    "cpgpu_unet_set_sample", "cpgpu_unet_num_samples", "cpgpu_unet_train",
    "cpgpu_unet_eval", "cpgpu_poisson_set_unet", "cpgpu_unet_tensor",
    "cpgpu_ns_set_time", "cpgpu_ns_set_solver", "cpgpu_poisson_set_fcg",
    "cpgpu_poisson_error_norms", "cpgpu_ns_error_norms", "cpgpu_batch_create",
    "cpgpu_batch_destroy", "cpgpu_batch_set_tiles", "cpgpu_batch_set_graphs",
    "cpgpu_batch_set_persistent",
    "cpgpu_batch_solve", "cpgpu_batch_autotune", "cpgpu_batch_draw",
    "cpgpu_ns_field_change", "cpgpu_unet_set_conv_shape", "cpgpu_unet_set_fusion", "cpgpu_advection_create",
    "cpgpu_advection_destroy", "cpgpu_advection_set_field", "cpgpu_advection_get_field",
    "cpgpu_advection_set_scheme", "cpgpu_advection_initial", "cpgpu_advection_step",
    "cpgpu_advection_time", "cpgpu_unet_clear_samples",
    "cpgpu_unet_set_walls", "cpgpu_copy_bandwidth", "cpgpu_csr_create",
    "cpgpu_csr_destroy", "cpgpu_csr_mult", "cpgpu_csr_solve",
    "cpgpu_poisson_set_alpha", "cpgpu_poisson_set_bc_values", "cpgpu_ns_set_kim",
    "cpgpu_ns_kim_iters", "cpgpu_ns_set_bc_values", "cpgpu_poisson2d_solve_static",
    "cpgpu_grid_axis_sizes",
]


def lib():
    """Load libcpgpu.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        L.cpgpu_last_error.restype = C.c_char_p
        L.cpgpu_stream.restype = C.c_void_p
        L.cpgpu_poisson_device_vec.restype = C.c_void_p
        L.cpgpu_unet_num_params.restype = C.c_longlong
        L.cpgpu_ns_projection.restype = C.c_void_p
        _lib = L
    return _lib


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


class Context:
    """One device, one stream (cpgpu_ctx)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        self._children = weakref.WeakSet()
        rc = lib().cpgpu_create(C.c_int(device), C.byref(self.h))
        if rc != OK:
            raise CpgpuError(rc, "cpgpu_create failed (no GPU?)")

    def adopt(self, obj):
        """Objects created on this context are closed before it."""
        self._children.add(obj)

    def check(self, rc, allow=(OK,)):
        if rc not in allow:
            msg = lib().cpgpu_last_error(self.h)
            raise CpgpuError(rc, msg.decode() if msg else "")
        return rc

    def sync(self):
        self.check(lib().cpgpu_synchronize(self.h))

    def copy_bandwidth(self, nbytes=1 << 30, reps=20):
        """(GB/s, us per copy) of a plain device copy of nbytes."""
        gbs, us = C.c_double(0), C.c_double(0)
        self.check(lib().cpgpu_copy_bandwidth(self.h, C.c_size_t(nbytes), C.c_int(reps),
                                              C.byref(gbs), C.byref(us)))
        return gbs.value, us.value

    def timer_enable(self, family, on=True):
        self.check(lib().cpgpu_timer_enable(self.h, family, int(on)))

    def timer_read(self, family):
        n, ms = C.c_int(0), C.c_double(0)
        self.check(lib().cpgpu_timer_read(self.h, family, C.byref(n),
                                          C.byref(ms)))
        return n.value, ms.value

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.close()
            lib().cpgpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_grid_desc(cx, sx, cy, sy, nx, ny, gl, bc_type, bc_val=None,
                   beta=1.0):
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (cx, sx, cy, sy)]
    g = Grid2D()
    g.nx, g.ny, g.gl = nx, ny, gl
    g.cx, g.sx, g.cy, g.sy = [_p(a) for a in keep]
    g.beta = beta
    for k in range(4):
        g.bc_type[k] = int(bc_type[k])
    if bc_val is not None:
        # 4 per-segment scalars or the packed per-ghost-line array: the
        # library reads 2*(nx+2gl) + 2*(ny+2gl) values
        bv = pack_bc_values(nx, ny, gl, bc_val)
        keep.append(bv)
        g.bc_val = _p(bv)
    g._keep = keep
    return g


def uniform_axis(n, gl, vmin, vmax):
    """(centres, sizes) of a uniform structure::Grid_ axis, ghosts included."""
    c = np.zeros(n + 2 * gl)
    s = np.zeros(n + 2 * gl)
    rc = lib().cpgpu_grid_uniform_axis(C.c_int(n), C.c_int(gl),
                                       C.c_double(vmin), C.c_double(vmax),
                                       _p(c), _p(s))
    if rc != OK:
        raise CpgpuError(rc, "bad axis")
    return c, s


def axis_from_sizes(sizes, gl, vmin=0.0):
    """(centres, sizes) of an axis of given interior cell sizes, ghosts
    included (structure::Grid_ from size arrays, s_grid.hpp:397-434)."""
    sizes = np.ascontiguousarray(sizes, dtype=np.float64)
    n = len(sizes)
    c = np.zeros(n + 2 * gl)
    s = np.zeros(n + 2 * gl)
    rc = lib().cpgpu_grid_axis_sizes(C.c_int(n), C.c_int(gl), C.c_double(vmin), _p(sizes),
                                     _p(c), _p(s))
    if rc != OK:
        raise CpgpuError(rc, "bad axis sizes")
    return c, s


def sized_grid_desc(xsizes, ysizes, gl, bc_type, bc_val=None, beta=1.0, x0=0.0, y0=0.0):
    """Descriptor of a grid of given cell sizes along x and y."""
    cx, sx = axis_from_sizes(xsizes, gl, x0)
    cy, sy = axis_from_sizes(ysizes, gl, y0)
    return make_grid_desc(cx, sx, cy, sy, len(xsizes), len(ysizes), gl, bc_type, bc_val, beta)


def uniform_grid_desc(nx, ny, gl, bc_type, bc_val=None, beta=1.0,
                      xlim=(0.0, 1.0), ylim=(0.0, 1.0)):
    """Descriptor of a uniform grid on [xlim] x [ylim]."""
    cx, sx = uniform_axis(nx, gl, *xlim)
    cy, sy = uniform_axis(ny, gl, *ylim)
    return make_grid_desc(cx, sx, cy, sy, nx, ny, gl, bc_type, bc_val, beta)


class Poisson:
    """Device-resident Poisson problem (cpgpu_poisson)."""

    def __init__(self, ctx, desc):
        self.ctx, self.desc = ctx, desc
        self.nx, self.ny, self.gl = desc.nx, desc.ny, desc.gl
        self.h = C.c_void_p()
        ctx.check(lib().cpgpu_poisson_create(ctx.h, C.byref(desc),
                                             C.byref(self.h)))
        ctx.adopt(self)

    @property
    def shape(self):
        return (self.nx + 2 * self.gl, self.ny + 2 * self.gl)

    def set_source(self, src_field=None, uniform=0.0):
        f = None if src_field is None else np.ascontiguousarray(
            src_field, dtype=np.float64)
        self.ctx.check(lib().cpgpu_poisson_set_source(self.h, _p(f),
                                                      C.c_double(uniform)))

    def set_phi(self, phi_field):
        f = np.ascontiguousarray(phi_field, dtype=np.float64)
        self.ctx.check(lib().cpgpu_poisson_set_phi(self.h, _p(f)))

    def get_phi(self, out=None):
        out = np.zeros(self.shape) if out is None else out
        self.ctx.check(lib().cpgpu_poisson_get_phi(self.h, _p(out)))
        return out

    def set_alpha(self, alpha):
        """Helmholtz rows (Poisson_::set_uniform_alpha, StencilHelmholtz_)."""
        self.ctx.check(lib().cpgpu_poisson_set_alpha(self.h, C.c_double(alpha)))

    def set_bc_values(self, vals):
        arr = pack_bc_values(self.nx, self.ny, self.gl, vals)
        self.ctx.check(lib().cpgpu_poisson_set_bc_values(self.h, _p(arr)))

    def apply(self, x_field):
        x = np.ascontiguousarray(x_field, dtype=np.float64)
        y = np.zeros(self.shape)
        self.ctx.check(lib().cpgpu_poisson_apply(self.h, _p(x), _p(y)))
        return y

    def residual(self):
        r = C.c_double(0)
        self.ctx.check(lib().cpgpu_poisson_residual(self.h, C.byref(r)))
        return r.value

    def error_norms(self, exact):
        """(Error1, Error2, Errori) of phi against a ghost-padded exact
        field (Operation_::Error1/2/i, s_operation.hpp:869-908)."""
        exact = np.ascontiguousarray(exact, dtype=np.float64)
        assert exact.shape == self.shape
        out = np.zeros(3)
        self.ctx.check(lib().cpgpu_poisson_error_norms(self.h, _p(exact), _p(out)))
        return out

    def set_mixed(self, inner_tol=1e-4, max_outer=40):
        self.ctx.check(lib().cpgpu_poisson_set_mixed(
            self.h, C.c_double(inner_tol), C.c_int(max_outer)))

    def solve(self, method="JPCG", max_iter=100000, tol=1e-10, norm=NORM_REF,
              opts=0, hist_cap=0):
        """Returns (code, iters, tol_out, hist)."""
        mi, t = C.c_int(max_iter), C.c_double(tol)
        hist = np.zeros(max(hist_cap, 1))
        hl = C.c_int(0)
        code = lib().cpgpu_poisson_solve(
            self.h, C.c_int(METHODS[method]), C.byref(mi), C.byref(t),
            C.c_int(norm), C.c_int(opts), _p(hist) if hist_cap else None,
            C.c_int(hist_cap), C.byref(hl))
        self.ctx.check(code, allow=(OK, MAXITER, BREAKDOWN_RHO,
                                    BREAKDOWN_OMEGA))
        return code, mi.value, t.value, hist[:min(hl.value, hist_cap)].copy()

    def close(self):
        if self.h and self.ctx.h and getattr(self, "_owned", True):
            lib().cpgpu_poisson_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _get_vec(self, which):
    out = np.zeros(self.shape)
    self.ctx.check(lib().cpgpu_poisson_get_vec(self.h, C.c_int(which), _p(out)))
    return out


Poisson.get_vec = _get_vec


def _set_source_blobs(self, seed, nblobs=64, wmin=0.01, wmax=0.1):
    self.ctx.check(lib().cpgpu_poisson_set_source_blobs(
        self.h, C.c_ulonglong(seed), C.c_int(nblobs), C.c_double(wmin),
        C.c_double(wmax)))


Poisson.set_source_blobs = _set_source_blobs


def _fill_phi(self, value=0.0):
    self.ctx.check(lib().cpgpu_poisson_fill_phi(self.h, C.c_double(value)))


Poisson.fill_phi = _fill_phi


def blob_params(seed, nblobs=64, wmin=0.01, wmax=0.1):
    out = np.zeros(4 * nblobs)
    rc = lib().cpgpu_blob_params(C.c_ulonglong(seed), C.c_int(nblobs),
                                 C.c_double(wmin), C.c_double(wmax), _p(out))
    if rc != OK:
        raise CpgpuError(rc, "bad blob parameters")
    return out.reshape(nblobs, 4)


# ---------------------------------------------------------------- NS step
LIMITERS = {"VanLeer": 0, "superbee": 1, "WAHYD": 2, "QUICK": 3, "CUI": 4,
            "upwind2": 5}
FIELDS = {"u": 0, "v": 1, "uf": 2, "vf": 3, "us": 4, "vs": 5, "usf": 6,
          "vsf": 7, "p": 8, "phi": 9, "div": 10}


class NSDesc(C.Structure):
    _fields_ = [("grid", Grid2D), ("bcu_type", C.c_int * 4), ("bcu_val", _dp),
                ("bcv_type", C.c_int * 4), ("bcv_val", _dp),
                ("rho", C.c_double), ("mu", C.c_double), ("cfl", C.c_double),
                ("dnum", C.c_double), ("dcoef", C.c_double),
                ("limiter", C.c_int), ("solver", C.c_int),
                ("max_iter", C.c_int), ("tol", C.c_double),
                ("norm_mode", C.c_int), ("dt", C.c_double)]


class AdvectionDesc(C.Structure):
    _fields_ = [("grid", Grid2D), ("bcu_type", C.c_int * 4), ("bcu_val", _dp),
                ("bcv_type", C.c_int * 4), ("bcv_val", _dp), ("scheme", C.c_int),
                ("dt", C.c_double)]


# Advection_::set_advection_scheme names (s_advection.hpp:140-182)
ADV_SCHEMES = {"VanLeer": 0, "superbee": 1, "WAHYD": 2, "QUICK": 3, "CUI": 4, "upwind2": 5,
               "upwind1": 10, "center": 11, "center4": 12}


def pack_bc_values(nx, ny, gl, values):
    """None, 4 per-segment scalars, or an already packed array."""
    NX, NY = nx + 2 * gl, ny + 2 * gl
    if values is None:
        return np.zeros(2 * NY + 2 * NX)
    if np.ndim(values) == 1 and len(values) == 4:
        return np.concatenate([np.full(NY, values[0]), np.full(NY, values[1]),
                               np.full(NX, values[2]), np.full(NX, values[3])])
    arr = np.ascontiguousarray(values, dtype=np.float64)
    # the library reads exactly 2*(ny+2gl) + 2*(nx+2gl) values
    if arr.ndim != 1 or arr.size != 2 * NY + 2 * NX:
        raise ValueError(f"boundary values: expected 4 scalars or a 1-D array of "
                         f"{2 * NY + 2 * NX} per-ghost-line values, got shape {arr.shape}")
    return arr


class NS:
    """Device-resident NS_explicit_ projection stepper (cpgpu_ns)."""

    def __init__(self, ctx, nx, ny, gl=2, xlim=(0.0, 1.0), ylim=(0.0, 1.0),
                 bcu=((1, 1, 1, 1), None), bcv=((1, 1, 1, 1), None), rho=1.0,
                 mu=1.0, cfl=0.5, dnum=1e10, dcoef=1.0, limiter="VanLeer",
                 solver="JPCG", max_iter=100000, tol=1e-10, norm=NORM_REF,
                 dt=1e-3, xsizes=None, ysizes=None):
        self.ctx, self.nx, self.ny, self.gl = ctx, nx, ny, gl
        if xsizes is None and ysizes is None:
            g = uniform_grid_desc(nx, ny, gl, [2, 2, 2, 2], xlim=xlim, ylim=ylim)
        else:
            # cell sizes from arrays (starting at xlim[0], ylim[0])
            xs = np.full(nx, (xlim[1] - xlim[0]) / nx) if xsizes is None else xsizes
            ys = np.full(ny, (ylim[1] - ylim[0]) / ny) if ysizes is None else ysizes
            g = sized_grid_desc(xs, ys, gl, [2, 2, 2, 2], x0=xlim[0], y0=ylim[0])
        d = NSDesc()
        d.grid = g
        self._keep = [g, g._keep]
        for name, (types, vals) in (("bcu", bcu), ("bcv", bcv)):
            arr = pack_bc_values(nx, ny, gl, vals)
            self._keep.append(arr)
            t = getattr(d, name + "_type")
            for k in range(4):
                t[k] = int(types[k])
            setattr(d, name + "_val", _p(arr))
        d.rho, d.mu, d.cfl, d.dnum, d.dcoef = rho, mu, cfl, dnum, dcoef
        d.limiter = LIMITERS[limiter]
        d.solver = METHODS[solver]
        d.max_iter, d.tol, d.norm_mode, d.dt = max_iter, tol, norm, dt
        self.desc = d
        self.h = C.c_void_p()
        ctx.check(lib().cpgpu_ns_create(ctx.h, C.byref(d), C.byref(self.h)))
        ctx.adopt(self)

    @property
    def shape(self):
        return (self.nx + 2 * self.gl, self.ny + 2 * self.gl)

    def set_field(self, name, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == self.shape
        self.ctx.check(lib().cpgpu_ns_set_field(self.h, FIELDS[name], _p(f)))

    def get_field(self, name):
        out = np.zeros(self.shape)
        self.ctx.check(lib().cpgpu_ns_get_field(self.h, FIELDS[name], _p(out)))
        return out

    def error_norms(self, name, exact):
        """(Error1, Error2, Errori) of field `name` against a ghost-padded
        exact field (Operation_::Error1/2/i, s_operation.hpp:869-908)."""
        exact = np.ascontiguousarray(exact, dtype=np.float64)
        assert exact.shape == self.shape
        out = np.zeros(3)
        self.ctx.check(lib().cpgpu_ns_error_norms(self.h, FIELDS[name], _p(exact),
                                                  _p(out)))
        return out

    def field_change(self, name):
        """Error1/2/i of field `name` against its value at the previous call
        (-1s on the first), the steady-state test of EventStopScalarError_
        (s_event.hpp:520-566)."""
        out = np.zeros(3)
        self.ctx.check(lib().cpgpu_ns_field_change(self.h, FIELDS[name], _p(out)))
        return out

    def step(self, n=1):
        it = np.zeros(max(n, 1), dtype=np.int32)
        res = np.zeros(max(n, 1))
        code = lib().cpgpu_ns_step(self.h, C.c_int(n), _p(it, _ip), _p(res))
        self.ctx.check(code, allow=(OK, MAXITER))
        return it[:n], res[:n]

    def time(self):
        t, dt = C.c_double(0), C.c_double(0)
        self.ctx.check(lib().cpgpu_ns_time(self.h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def set_time(self, t, dt):
        self.ctx.check(lib().cpgpu_ns_set_time(self.h, C.c_double(t), C.c_double(dt)))

    def set_kim(self, enable=True, helm="Jacobi", helm_max_iter=1000,
                helm_tol=1e-4):
        """NS_kim_ stepping (s_ns_kim.hpp:130-287): fixed dt, upwind
        advection, a Helmholtz predictor per component solved by `helm`
        (Solver_Jacobi_ and the Solver_ limits by default)."""
        self.ctx.check(lib().cpgpu_ns_set_kim(self.h, int(bool(enable)),
                                              METHODS[helm], int(helm_max_iter),
                                              C.c_double(helm_tol)))

    def kim_iters(self):
        it = np.zeros(2, dtype=np.int32)
        self.ctx.check(lib().cpgpu_ns_kim_iters(self.h, _p(it, _ip)))
        return int(it[0]), int(it[1])

    def set_bc_values(self, comp, vals):
        """New boundary values of velocity component comp ("u"/"v" or 0/1):
        4 scalars or the packed per-ghost-line array."""
        c = {"u": 0, "v": 1}.get(comp, comp)
        arr = pack_bc_values(self.nx, self.ny, self.gl, vals)
        self.ctx.check(lib().cpgpu_ns_set_bc_values(self.h, int(c), _p(arr)))

    def set_solver(self, solver, max_iter=100000, tol=1e-10, norm=NORM_REF):
        self.ctx.check(lib().cpgpu_ns_set_solver(self.h, METHODS[solver],
                                                 int(max_iter), C.c_double(tol),
                                                 int(norm)))

    def projection(self):
        """The step's projection problem (all-Neumann Poisson) as a Poisson
        view that the stepper owns: b and phi of the last step, U-Net
        attachment, solver options."""
        P = Poisson.__new__(Poisson)
        P.ctx, P.desc = self.ctx, None
        P.nx, P.ny, P.gl = self.nx, self.ny, self.gl
        P.h = C.c_void_p(lib().cpgpu_ns_projection(self.h))
        P._owned = False
        P._ns = self
        return P

    STATE = ("u", "v", "uf", "vf", "p", "phi")

    def save(self):
        """Host copy of the step state (fields, t, dt)."""
        return {"fields": {k: self.get_field(k) for k in self.STATE},
                "time": self.time()}

    def restore(self, state):
        for k, f in state["fields"].items():
            self.set_field(k, f)
        self.set_time(*state["time"])

    def close(self):
        if self.h and self.ctx.h:
            lib().cpgpu_ns_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _set_variant(self, v64, v32):
    self.ctx.check(lib().cpgpu_poisson_set_variant(self.h, int(v64), int(v32)))


def _autotune(self, reps=10):
    v64, v32 = C.c_int(0), C.c_int(0)
    t64, t32 = C.c_float(0), C.c_float(0)
    self.ctx.check(lib().cpgpu_poisson_autotune(
        self.h, int(reps), C.byref(v64), C.byref(v32), C.byref(t64),
        C.byref(t32)))
    return {"fp64": (v64.value, t64.value), "fp32": (v32.value, t32.value)}


Poisson.set_variant = _set_variant
Poisson.autotune = _autotune


def _kernel_times(self, precision=32, reps=50):
    d, u = C.c_double(0), C.c_double(0)
    self.ctx.check(lib().cpgpu_poisson_kernel_times(
        self.h, int(precision), int(reps), C.byref(d), C.byref(u)))
    return d.value, u.value


def _set_batch(self, iters=16, use_graphs=True):
    """use_graphs: False, True (replayed Krylov batches) or 2 (each
    flexible-CG iteration replayed as one graph as well)."""
    g = 2 if (not isinstance(use_graphs, bool) and use_graphs == 2) else int(bool(use_graphs))
    self.ctx.check(lib().cpgpu_poisson_set_batch(self.h, int(iters), g))


Poisson.kernel_times = _kernel_times
Poisson.set_batch = _set_batch


# ------------------------------------------------------------ batched solves
def draw_problems(seed, n, lo=6, hi=10):
    """Random problem set (cpgpu_batch_draw): nx, ny, bc_type (n x 4),
    blob seeds."""
    nx = np.zeros(n, dtype=np.int32)
    ny = np.zeros(n, dtype=np.int32)
    bc = np.zeros((n, 4), dtype=np.int32)
    bs = np.zeros(n, dtype=np.uint64)
    code = lib().cpgpu_batch_draw(C.c_ulonglong(seed), C.c_int(n), C.c_int(lo),
                                  C.c_int(hi), _p(nx, _ip), _p(ny, _ip), _p(bc, _ip),
                                  bs.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    if code != OK:
        raise CpgpuError(code, "cpgpu_batch_draw")
    return nx, ny, bc, bs


def random_problems(ctx, seed=2, n=256, lo=6, hi=10, nblobs=64, wmin=0.01, wmax=0.1):
    """The drawn problem set on [0,1]^2 grids (gl 1, boundary values 0) with
    zero-meaned Gaussian-blob sources; phi = 0."""
    nx, ny, bc, bs = draw_problems(seed, n, lo, hi)
    probs = []
    for p in range(n):
        P = Poisson(ctx, uniform_grid_desc(int(nx[p]), int(ny[p]), 1, [int(t) for t in bc[p]]))
        P.set_source_blobs(int(bs[p]), nblobs=nblobs, wmin=wmin, wmax=wmax)
        P.fill_phi(0.0)
        probs.append(P)
    return probs


EZERODIAG = -10



class Advection:
    """Device-resident Advection_ scalar driver (cpgpu_advection): phi's
    boundary (types, values), the centre velocities' boundaries, the scheme
    name and the Time_ dt."""

    def __init__(self, ctx, nx, ny, gl=2, xlim=(0.0, 1.0), ylim=(0.0, 1.0),
                 bc_phi=((1, 1, 1, 1), None), bcu=((2, 2, 2, 2), None),
                 bcv=((2, 2, 2, 2), None), scheme="upwind1", dt=1e-3):
        self.ctx, self.nx, self.ny, self.gl = ctx, nx, ny, gl
        types, vals = bc_phi
        pv = pack_bc_values(nx, ny, gl, vals)
        g = uniform_grid_desc(nx, ny, gl, list(types), pv, xlim=xlim, ylim=ylim)
        d = AdvectionDesc()
        d.grid = g
        self._keep = [g, g._keep, pv]
        for name, (t_, v_) in (("bcu", bcu), ("bcv", bcv)):
            arr = pack_bc_values(nx, ny, gl, v_)
            self._keep.append(arr)
            t = getattr(d, name + "_type")
            for k in range(4):
                t[k] = int(t_[k])
            setattr(d, name + "_val", _p(arr))
        d.scheme = ADV_SCHEMES[scheme]
        d.dt = dt
        self.desc = d
        self.h = C.c_void_p()
        ctx.check(lib().cpgpu_advection_create(ctx.h, C.byref(d), C.byref(self.h)))
        ctx.adopt(self)

    @property
    def shape(self):
        return (self.nx + 2 * self.gl, self.ny + 2 * self.gl)

    def set_field(self, name, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == self.shape
        self.ctx.check(lib().cpgpu_advection_set_field(self.h, FIELDS[name], _p(f)))

    def get_field(self, name):
        out = np.zeros(self.shape)
        self.ctx.check(lib().cpgpu_advection_get_field(self.h, FIELDS[name], _p(out)))
        return out

    def set_scheme(self, name):
        self.ctx.check(lib().cpgpu_advection_set_scheme(self.h, ADV_SCHEMES[name]))

    def initial(self):
        self.ctx.check(lib().cpgpu_advection_initial(self.h))

    def step(self, n=1):
        self.ctx.check(lib().cpgpu_advection_step(self.h, C.c_int(n)))

    def time(self):
        t, k = C.c_double(0), C.c_longlong(0)
        self.ctx.check(lib().cpgpu_advection_time(self.h, C.byref(t), C.byref(k)))
        return t.value, k.value

    def close(self):
        if self.h:
            lib().cpgpu_advection_destroy(self.h)
            self.h = C.c_void_p()

class CSR:
    """A general CSR matrix on the device (cpgpu_csr): the algebra solvers'
    MatrixSCR_ (lib/algebra/matrix_SparCompRow.hpp) with the free-function
    solvers of lib/algebra/solver_matrix.hpp on it."""

    def __init__(self, ctx, row_ptr, col, val):
        self.ctx = ctx
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.row_ptr) - 1
        self.h = C.c_void_p()
        ctx.check(lib().cpgpu_csr_create(ctx.h, C.c_int(self.n), _p(self.row_ptr, _ip),
                                         _p(self.col, _ip), _p(self.val), C.byref(self.h)))
        ctx.adopt(self)

    def mult(self, x):
        """y = A x (MatrixSCR_::operator*, rows summed in CSR order)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        self.ctx.check(lib().cpgpu_csr_mult(self.h, _p(x), _p(y)))
        return y

    def solve(self, method, x, b, max_iter=1000, tol=1e-4, norm=NORM_REF, hist_cap=100000):
        """The reference's int F(A, x, b, max_iter, tol, lresid): returns
        (code, x, max_iter, tol, residual history); x is not modified."""
        x = np.array(x, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        mi, t, hl = C.c_int(max_iter), C.c_double(tol), C.c_int(0)
        hist = np.zeros(max(hist_cap, 1))
        code = lib().cpgpu_csr_solve(self.h, C.c_int(METHODS[method]), _p(x), _p(b),
                                     C.byref(mi), C.byref(t), C.c_int(norm), _p(hist),
                                     C.c_int(hist_cap), C.byref(hl))
        self.ctx.check(code, allow=(OK, MAXITER, 2, 3))
        return code, x, mi.value, t.value, hist[:min(hl.value, hist_cap)]

    def close(self):
        if self.h and self.ctx.h:
            lib().cpgpu_csr_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched solves over Poisson problems of uneven sizes (cpgpu_batch)."""

    def __init__(self, ctx, problems):
        self.ctx, self.problems = ctx, list(problems)
        self.n = len(self.problems)
        arr = (C.c_void_p * self.n)(*[P.h for P in self.problems])
        self.h = C.c_void_p()
        ctx.check(lib().cpgpu_batch_create(ctx.h, C.c_int(self.n), arr, C.byref(self.h)))
        ctx.adopt(self)

    def set_tiles(self, tpb):
        self.ctx.check(lib().cpgpu_batch_set_tiles(self.h, C.c_int(tpb)))

    def set_graphs(self, on):
        self.ctx.check(lib().cpgpu_batch_set_graphs(self.h, C.c_int(int(bool(on)))))

    def set_persistent(self, on):
        """Every batched pass as one persistent cooperative launch."""
        self.ctx.check(lib().cpgpu_batch_set_persistent(self.h, C.c_int(int(bool(on)))))

    def autotune(self, method="JPCG", reps=2):
        best = C.c_int(0)
        ms = np.zeros(4, dtype=np.float32)
        self.ctx.check(lib().cpgpu_batch_autotune(self.h, C.c_int(METHODS[method]),
                                                  C.c_int(reps), C.byref(best),
                                                  ms.ctypes.data_as(C.POINTER(C.c_float))))
        return best.value, ms

    def solve(self, method="JPCG", max_iter=100000, tol=1e-8):
        """Returns (code, iters, resid, codes, latency_ms) per problem."""
        it = np.zeros(self.n, dtype=np.int32)
        res = np.zeros(self.n)
        codes = np.zeros(self.n, dtype=np.int32)
        lat = np.zeros(self.n)
        code = lib().cpgpu_batch_solve(self.h, C.c_int(METHODS[method]), C.c_int(max_iter),
                                       C.c_double(tol), _p(it, _ip), _p(res),
                                       _p(codes, _ip), _p(lat))
        self.ctx.check(code, allow=(OK, MAXITER))
        return code, it, res, codes, lat

    def close(self):
        if self.h and self.ctx.h:
            lib().cpgpu_batch_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- learned preconditioner
_fp = C.POINTER(C.c_float)


class UNet:
    """BF16 U-Net preconditioner (cpgpu_unet): 4 levels, 16-32-64-128
    channels, stride-4 pooling, FP32 master parameters, Adam."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(lib().cpgpu_unet_create(ctx.h, C.c_ulonglong(seed),
                                          C.byref(self.h)))
        ctx.adopt(self)
        self.num_params = int(lib().cpgpu_unet_num_params(self.h))

    def get_params(self, grads=False):
        out = np.zeros(self.num_params, dtype=np.float32)
        self.ctx.check(lib().cpgpu_unet_get_params(self.h, _p(out, _fp),
                                                   int(bool(grads))))
        return out

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        assert a.size == self.num_params
        self.ctx.check(lib().cpgpu_unet_set_params(self.h, _p(a, _fp)))

    def set_walls(self, bc_type=None):
        """Wall types (x-, x+, y-, y+) of the raw forward/backward input."""
        arr = None if bc_type is None else (C.c_int * 4)(*[int(t) for t in bc_type])
        self.ctx.check(lib().cpgpu_unet_set_walls(self.h, arr))

    def set_conv_shape(self, rows=4, persist=4):
        self.ctx.check(lib().cpgpu_unet_set_conv_shape(self.h, C.c_int(rows), C.c_int(persist)))

    def set_fusion(self, level0=True):
        """Run an applied network's finest level in the two fused kernels."""
        self.ctx.check(lib().cpgpu_unet_set_fusion(self.h, C.c_int(int(bool(level0)))))

    def forward(self, x):
        """Raw network output for an (H, W) array (no input scaling)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros_like(x)
        H, W = x.shape
        self.ctx.check(lib().cpgpu_unet_forward(self.h, H, W, _p(x, _fp),
                                                _p(y, _fp)))
        return y

    def backward(self, dy):
        """Gradients of sum(dy * y) for the last forward()."""
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        g = np.zeros(self.num_params, dtype=np.float32)
        self.ctx.check(lib().cpgpu_unet_backward(self.h, _p(dy, _fp),
                                                 _p(g, _fp)))
        return g

    def adam(self, lr):
        self.ctx.check(lib().cpgpu_unet_adam(self.h, C.c_float(lr)))

    TENSORS = ["x", "a0", "C0", "d0a", "d0", "p0", "a1", "C1", "d1a", "d1",
               "p1", "a2", "C2", "d2a", "d2", "p2", "a3", "b3", "y"]

    def tensor(self, name, grad=False):
        """Internal activation (or its gradient) of the last forward
        (backward) as an (H, W, C) float32 array."""
        tid = self.TENSORS.index(name)
        dims = (C.c_int * 3)()
        self.ctx.check(lib().cpgpu_unet_tensor(self.h, tid, int(grad), None, dims))
        out = np.zeros((dims[0], dims[1], dims[2]), dtype=np.float32)
        self.ctx.check(lib().cpgpu_unet_tensor(self.h, tid, int(grad),
                                               _p(out, _fp), dims))
        return out

    def set_sample(self, P, slot, e, smooth=0, omega=0.8):
        """Training sample: error e (nx, ny) of problem P, after `smooth`
        weighted-Jacobi sweeps; r = A e on device."""
        e = np.ascontiguousarray(e, dtype=np.float64)
        assert e.shape == (P.nx, P.ny)
        self.ctx.check(lib().cpgpu_unet_set_sample(self.h, P.h, int(slot),
                                                   _p(e), int(smooth),
                                                   C.c_double(omega)))

    @property
    def num_samples(self):
        return int(lib().cpgpu_unet_num_samples(self.h))

    def clear_samples(self):
        self.ctx.check(lib().cpgpu_unet_clear_samples(self.h))

    def train(self, P, steps, batch=4, lr=1e-3, seed=0):
        losses = np.zeros(max(steps, 1), dtype=np.float32)
        self.ctx.check(lib().cpgpu_unet_train(
            self.h, P.h, int(steps), int(batch), C.c_float(lr),
            C.c_ulonglong(seed), _p(losses, _fp)))
        return losses[:steps]

    def eval(self, P):
        out = np.zeros(self.num_samples, dtype=np.float32)
        self.ctx.check(lib().cpgpu_unet_eval(self.h, P.h, _p(out, _fp)))
        return out

    def close(self):
        if self.h and self.ctx.h:
            lib().cpgpu_unet_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _set_unet(self, unet, flags=UNET_PRECOND):
    self.ctx.check(lib().cpgpu_poisson_set_unet(
        self.h, unet.h if unet is not None else None, int(flags)))
    self._unet = unet  # keep it alive while attached


Poisson.set_unet = _set_unet


def _set_fcg(self, pre=2, post=2, omega=0.8):
    self.ctx.check(lib().cpgpu_poisson_set_fcg(self.h, int(pre), int(post),
                                               C.c_double(omega)))


Poisson.set_fcg = _set_fcg
