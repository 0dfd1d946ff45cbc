"""ctypes view of the CPU oracle (oracle/liboracle.so) and, where built, of the
reference's own lib/algebra solvers (oracle/_ref/libcarpios_ref.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module as the checker.  The product package
``carpios_amd`` never imports it.
"""
import ctypes as C  # noqa: N812
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libcarpios_ref.so")

BC1, BC2 = 1, 2
NORM_REF, NORM_TRUE = 0, 1
LIM = {"VanLeer": 0, "superbee": 1, "WAHYD": 2, "QUICK": 3, "CUI": 4,
       "upwind2": 5}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("gl", C.c_int),
                ("cx", _dp), ("sx", _dp), ("cy", _dp), ("sy", _dp)]


class CSR(C.Structure):
    _fields_ = [("n", C.c_int), ("row_ptr", _ip), ("col", _ip), ("val", _dp)]


class BC(C.Structure):
    _fields_ = [("type", C.c_int * 4), ("val", _dp)]


class NSParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("cfl", C.c_double),
                ("dnum", C.c_double), ("dcoef", C.c_double),
                ("limiter", C.c_int), ("max_iter", C.c_int),
                ("tol", C.c_double), ("norm_mode", C.c_int)]


class NSState(C.Structure):
    _fields_ = [(n, _dp) for n in ("u", "v", "uf", "vf", "us", "vs", "usf",
                                   "vsf", "p", "phi", "div")] + \
               [("dt", C.c_double), ("t", C.c_double), ("kim_src", _dp)]


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_dot.restype = C.c_double
        L.orc_nrm2_ref.restype = C.c_double
        L.orc_nrm2_true.restype = C.c_double
        L.orc_tvd_cell.restype = C.c_double
        L.orc_diffusion_cell.restype = C.c_double
        L.orc_upwind1_cell.restype = C.c_double
        L.orc_max_magnitude.restype = C.c_double
        L.orc_center_cell.restype = C.c_double
        L.orc_center4_cell.restype = C.c_double
        L.orc_adv_cell.restype = C.c_double
        L.orc_advection_step.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_void_p, C.c_int,
                                         C.c_double]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref not built (needs the reference tree)")
        L = C.CDLL(REF_PATH)
        L.ref_dot.restype = C.c_double
        L.ref_nrm2.restype = C.c_double
        _ref = L
    return _ref


class Axis:
    """Per-axis centre and size arrays (ghosts included)."""

    def __init__(self, n, gl, vmin, vmax):
        self.n, self.gl = n, gl
        self.c = np.zeros(n + 2 * gl)
        self.s = np.zeros(n + 2 * gl)
        lib().orc_axis_uniform(C.c_int(n), C.c_int(gl), C.c_double(vmin),
                               C.c_double(vmax), _ptr(self.c), _ptr(self.s))


class SizedAxis:
    """Axis of given interior cell sizes (Grid_::_set_non_uniform_grid)."""

    def __init__(self, sizes, gl, vmin):
        sizes = np.ascontiguousarray(sizes, dtype=np.float64)
        self.n, self.gl = len(sizes), gl
        self.c = np.zeros(self.n + 2 * gl)
        self.s = np.zeros(self.n + 2 * gl)
        lib().orc_axis_sizes(C.c_int(self.n), C.c_int(gl), C.c_double(vmin),
                             _ptr(sizes), _ptr(self.c), _ptr(self.s))


class StructGrid:
    """Structured grid as lib/structure/s_grid.hpp builds it: uniform on
    [xmin, xmax] x [ymin, ymax], or from interior cell sizes (xsizes,
    ysizes) starting at (xmin, ymin)."""

    def __init__(self, nx, ny, gl=1, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0,
                 xsizes=None, ysizes=None):
        self.nx, self.ny, self.gl = nx, ny, gl
        self.ax = Axis(nx, gl, xmin, xmax) if xsizes is None else SizedAxis(xsizes, gl, xmin)
        self.ay = Axis(ny, gl, ymin, ymax) if ysizes is None else SizedAxis(ysizes, gl, ymin)
        self.c = Grid(nx, ny, gl, _ptr(self.ax.c), _ptr(self.ax.s),
                      _ptr(self.ay.c), _ptr(self.ay.s))

    @property
    def shape(self):
        return (self.nx + 2 * self.gl, self.ny + 2 * self.gl)

    def field(self):
        return np.zeros(self.shape)

    def interior(self, f):
        g = self.gl
        return f[g:g + self.nx, g:g + self.ny]

    def centres(self):
        """Interior cell-centre coordinates as (X, Y) arrays of shape (nx, ny)."""
        g = self.gl
        return np.meshgrid(self.ax.c[g:g + self.nx], self.ay.c[g:g + self.ny],
                           indexing="ij")


def make_bc(grid, types, values=None):
    """types: 4 ints (x-, x+, y-, y+); values: None, 4 scalars, or the packed
    [seg0 | seg1 | seg2 | seg3] array of per-line values."""
    NX, NY = grid.shape
    if values is None:
        arr = np.zeros(2 * NY + 2 * NX)
    elif np.ndim(values) == 1 and len(values) == 4:
        arr = np.concatenate([np.full(NY, values[0]), np.full(NY, values[1]),
                              np.full(NX, values[2]), np.full(NX, values[3])])
    else:
        arr = np.ascontiguousarray(values, dtype=np.float64)
    bc = BC()
    for k in range(4):
        bc.type[k] = int(types[k])
    bc.val = _ptr(arr)
    bc._keep = arr
    return bc


class Matrix:
    """Assembled CSR (i-fastest unknowns) plus its right-hand side."""

    def __init__(self, grid, bc, beta=1.0, src_field=None, src_uniform=0.0,
                 alpha=None):
        n = grid.nx * grid.ny
        nnz_max = 5 * n
        self.n = n
        self.row_ptr = np.zeros(n + 1, dtype=np.int32)
        self.col = np.zeros(nnz_max, dtype=np.int32)
        self.val = np.zeros(nnz_max)
        self.b = np.zeros(n)
        sp = None
        if src_field is not None:
            self._src = np.ascontiguousarray(src_field, dtype=np.float64)
            sp = _ptr(self._src)
        # alpha: the StencilHelmholtz_ rows of Poisson_::set_uniform_alpha
        head = ((lib().orc_poisson_assemble,) if alpha is None else
                (lib().orc_helmholtz_assemble, C.c_double(alpha)))
        nnz = head[0](
            C.byref(grid.c), *head[1:], C.c_double(beta), sp,
            C.c_double(src_uniform),
            C.byref(bc), _ptr(self.row_ptr, _ip), _ptr(self.col, _ip),
            _ptr(self.val), _ptr(self.b))
        if nnz < 0:
            raise ValueError("assembly failed")
        self.nnz = nnz
        self.col = self.col[:nnz].copy()
        self.val = self.val[:nnz].copy()
        self.c = CSR(n, _ptr(self.row_ptr, _ip), _ptr(self.col, _ip),
                     _ptr(self.val))

    def dense(self):
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            for k in range(self.row_ptr[i], self.row_ptr[i + 1]):
                A[i, self.col[k]] = self.val[k]
        return A

    def mult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        lib().orc_csr_mult(C.byref(self.c), _ptr(x), _ptr(y))
        return y


_SOLVERS = {"CG": "orc_cg", "JPCG": "orc_jpcg",
            "Dia_BiCGSTAB": "orc_dia_bicgstab", "Jacobi": "orc_jacobi"}


def solve(method, A, b, x0=None, max_iter=1000, tol=1e-10, norm=NORM_REF):
    """Returns (code, x, iters, tol_out, hist)."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    mi = C.c_int(max_iter)
    t = C.c_double(tol)
    hist = np.zeros(max_iter + 2)
    hl = C.c_int(0)
    fn = getattr(lib(), _SOLVERS[method])
    code = fn(C.byref(A.c), _ptr(x), _ptr(b), C.byref(mi), C.byref(t),
              C.c_int(norm), _ptr(hist), C.c_int(len(hist)), C.byref(hl))
    return code, x, mi.value, t.value, hist[:min(hl.value, len(hist))].copy()


REF_METHODS = {"CG": 0, "JPCG": 1, "Dia_BiCGSTAB": 2, "Jacobi": 3,
               "Solver_Jacobi_": 4, "IC_CG": 5, "IC_CGS": 6, "IC_BiCGSTAB": 7}


def ref_solve(method, A, b, x0=None, max_iter=1000, tol=1e-10):
    """The reference's own solver on the same CSR (oracle/_ref)."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    mi = C.c_int(max_iter)
    t = C.c_double(tol)
    hist = np.zeros(max_iter + 2)
    hl = C.c_int(0)
    code = ref().ref_solve(C.c_int(REF_METHODS[method]), C.c_int(A.n),
                           _ptr(A.row_ptr, _ip), _ptr(A.col, _ip),
                           _ptr(A.val), _ptr(x), _ptr(b), C.byref(mi),
                           C.byref(t), _ptr(hist), C.c_int(len(hist)),
                           C.byref(hl))
    return code, x, mi.value, t.value, hist[:min(hl.value, len(hist))].copy()


def jpcg_mixed(A, b, x0=None, tol=1e-10, inner_tol=1e-4, max_outer=50,
               max_inner=100000, project_mean=False):
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    outer, inner, res = C.c_int(0), C.c_int(0), C.c_double(0)
    code = lib().orc_jpcg_mixed(C.byref(A.c), _ptr(x), _ptr(b),
                                C.c_double(tol), C.c_double(inner_tol),
                                C.c_int(max_outer), C.c_int(max_inner),
                                C.c_int(int(project_mean)), C.byref(outer),
                                C.byref(inner), C.byref(res))
    return code, x, outer.value, inner.value, res.value


def field_to_x(grid, f):
    x = np.zeros(grid.nx * grid.ny)
    f = np.ascontiguousarray(f, dtype=np.float64)
    lib().orc_field_to_x(C.byref(grid.c), _ptr(f), _ptr(x))
    return x


def x_to_field(grid, x, f=None):
    f = grid.field() if f is None else f
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_x_to_field(C.byref(grid.c), _ptr(x), _ptr(f))
    return f


def apply_bc(grid, f, bc):
    assert f.flags.c_contiguous and f.dtype == np.float64
    lib().orc_apply_bc(C.byref(grid.c), _ptr(f), C.byref(bc))
    return f


class NS:
    """NS_explicit_ state driven by the oracle (s_ns_explicit.hpp:118-155)."""

    def __init__(self, grid, bcu, bcv, rho=1.0, mu=1.0, cfl=0.5, dnum=1e10,
                 dcoef=1.0, limiter="VanLeer", max_iter=100000, tol=1e-10,
                 norm=NORM_REF, dt=1e-3):
        self.g = grid
        self.bcu, self.bcv = bcu, bcv
        self.prm = NSParams(rho, mu, cfl, dnum, dcoef, LIM[limiter], max_iter,
                            tol, norm)
        self.f = {n: grid.field() for n in ("u", "v", "uf", "vf", "us", "vs",
                                            "usf", "vsf", "p", "phi", "div")}
        self.st = NSState(*[_ptr(self.f[n]) for n in
                            ("u", "v", "uf", "vf", "us", "vs", "usf", "vsf",
                             "p", "phi", "div")], dt, 0.0)
        bcphi = make_bc(grid, [BC2] * 4)
        self.A = Matrix(grid, bcphi, beta=1.0)
        self.iters = []

    @property
    def dt(self):
        return self.st.dt

    def step(self):
        it, res = C.c_int(0), C.c_double(0)
        code = lib().orc_ns_step(C.byref(self.g.c), C.byref(self.A.c),
                                 C.byref(self.bcu), C.byref(self.bcv),
                                 C.byref(self.prm), C.byref(self.st),
                                 C.byref(it), C.byref(res))
        self.iters.append(it.value)
        return code, it.value, res.value


class KimParams(C.Structure):
    _fields_ = [("method", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double)]


SOLVE = {"CG": 0, "JPCG": 1, "Dia_BiCGSTAB": 2, "Jacobi": 3}


class NSKim(NS):
    """NS_kim_ state driven by the oracle (s_ns_kim.hpp:130-287): fixed dt,
    upwind1 advection, a Helmholtz predictor per component (Solver_Jacobi_
    with the Solver_ defaults unless told otherwise) and the projection."""

    def __init__(self, grid, bcu, bcv, helm="Jacobi", helm_max_iter=1000,
                 helm_tol=1e-4, **kw):
        super().__init__(grid, bcu, bcv, **kw)
        self.kp = KimParams(SOLVE[helm], helm_max_iter, helm_tol)
        # the last step's predictor sources of u and v (ghost-padded)
        self.kim_src = np.zeros((2,) + grid.shape)
        self.st.kim_src = _ptr(self.kim_src)
        self.nstep = 0
        self.helm_iters = []
        self._build()

    def _build(self):
        # the predictor Poisson_ has its own empty BoundaryIndex
        # (s_ns_kim.hpp:105, s_equation.hpp:104): BC2 value 0 on every wall
        beta = self.dt * 0.5 / self.prm.rho
        walls = _plant == PLANT["kim_pred_walls"]
        bp = make_bc(self.g, [BC2] * 4)
        self.Hu = Matrix(self.g, self.bcu if walls else bp, beta=beta, alpha=1.0)
        self.Hv = Matrix(self.g, self.bcv if walls else bp, beta=beta, alpha=1.0)

    def step(self):
        it = (C.c_int * 3)()
        res = C.c_double(0)
        code = lib().orc_ns_kim_step(
            C.byref(self.g.c), C.byref(self.A.c), C.byref(self.Hu.c),
            C.byref(self.Hv.c), C.byref(self.bcu), C.byref(self.bcv),
            C.byref(self.prm), C.byref(self.kp), self.nstep, C.byref(self.st),
            it, C.byref(res))
        self.nstep += 1
        self.iters.append(it[0])
        self.helm_iters.append((it[1], it[2]))
        return code, it[0], res.value


PLANT = {"none": 0, "adv_sign": 1, "dif_rho": 2, "face_upwind": 3,
         "neumann_sign": 4, "kim_pred_walls": 5, "kim_ab2_ghosts": 6,
         "adv_two_buffer": 7}
_plant = 0


def set_plant(name):
    """Switch on one deliberate mistake (ORC_PLANT_*) so that a test can show
    its pin fails on a wrong restatement; "none" restores the oracle."""
    global _plant
    _plant = PLANT[name]
    lib().orc_set_plant(C.c_int(_plant))


def set_threads(n):
    """Threads of the CPU baseline; 1 (the default) keeps every routine in
    the reference's serial order, as the parity checks need."""
    lib().orc_set_threads(C.c_int(int(n)))


# Advection_::set_advection_scheme names (s_advection.hpp:140-182): the TVD
# family by limiter id, and the three plain schemes
ADV = dict(LIM, upwind1=10, center=11, center4=12)


def adv_field(grid, uf, vf, phi, scheme):
    """The advection term of every cell, evaluated on phi as it is."""
    out = np.zeros((grid.nx, grid.ny))
    L = lib()
    for i in range(grid.nx):
        for j in range(grid.ny):
            out[i, j] = L.orc_adv_cell(C.byref(grid.c), _ptr(uf), _ptr(vf), _ptr(phi), i, j,
                                       ADV[scheme])
    return out


class Advection:
    """Advection_ state driven by the oracle (s_advection.hpp:20-217): centre
    and face velocities, phi and its boundary, the scheme and dt."""

    def __init__(self, grid, bc_phi, scheme="upwind1", dt=1e-3):
        self.g, self.bc, self.dt = grid, bc_phi, dt
        self.scheme = scheme
        self.f = {n: grid.field() for n in ("u", "v", "uf", "vf", "phi")}

    def initial(self, bcu, bcv):
        """Advection_::initial (s_advection.hpp:103-135): the centre
        velocities' ghosts, then the faces by InterpolateC2F."""
        apply_bc(self.g, self.f["u"], bcu)
        apply_bc(self.g, self.f["v"], bcv)
        c2f(self.g, self.f["u"], self.f["v"], self.f["uf"], self.f["vf"])

    def step(self, n=1):
        for _ in range(n):
            lib().orc_advection_step(C.byref(self.g.c), _ptr(self.f["uf"]), _ptr(self.f["vf"]),
                                     _ptr(self.f["phi"]), C.byref(self.bc), ADV[self.scheme],
                                     self.dt)


def upwind1_field(grid, uf, vf, phi):
    out = np.zeros((grid.nx, grid.ny))
    L = lib()
    for i in range(grid.nx):
        for j in range(grid.ny):
            out[i, j] = L.orc_upwind1_cell(C.byref(grid.c), _ptr(uf), _ptr(vf),
                                           _ptr(phi), i, j)
    return out


def tvd_field(grid, uf, vf, phi, limiter="VanLeer"):
    out = np.zeros((grid.nx, grid.ny))
    L = lib()
    for i in range(grid.nx):
        for j in range(grid.ny):
            out[i, j] = L.orc_tvd_cell(C.byref(grid.c), _ptr(uf), _ptr(vf),
                                       _ptr(phi), i, j, LIM[limiter])
    return out


def diffusion_field(grid, mu, phi):
    out = np.zeros((grid.nx, grid.ny))
    L = lib()
    for i in range(grid.nx):
        for j in range(grid.ny):
            out[i, j] = L.orc_diffusion_cell(C.byref(grid.c), C.c_double(mu),
                                             _ptr(phi), i, j)
    return out


def c2f(grid, uc, vc, uf=None, vf=None):
    uf = grid.field() if uf is None else uf
    vf = grid.field() if vf is None else vf
    lib().orc_interp_c2f(C.byref(grid.c), _ptr(uc), _ptr(vc), _ptr(uf),
                         _ptr(vf))
    return uf, vf


def div(grid, uf, vf):
    d = grid.field()
    lib().orc_div(C.byref(grid.c), _ptr(uf), _ptr(vf), _ptr(d))
    return d


def grad(grid, phi):
    gx, gy = grid.field(), grid.field()
    lib().orc_grad(C.byref(grid.c), _ptr(phi), _ptr(gx), _ptr(gy))
    return gx, gy


def error_norms(grid, exa, val):
    """(Error1, Error2, Errori) of val against exa, s_operation.hpp:869-908."""
    out = np.zeros(3)
    exa = np.ascontiguousarray(exa, dtype=np.float64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    lib().orc_error_norms(C.byref(grid.c), _ptr(exa), _ptr(val), _ptr(out))
    return out


def f2c(grid, gx, gy):
    cx, cy = grid.field(), grid.field()
    lib().orc_interp_f2c(C.byref(grid.c), _ptr(gx), _ptr(gy), _ptr(cx),
                         _ptr(cy))
    return cx, cy
