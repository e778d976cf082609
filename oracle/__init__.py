"""CPU restatement of the reference's acoustic RK4 hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product path (solar_amd)
never imports or calls it.  The C source (solar_oracle.c) cites the reference
file:line it follows for every routine.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_D = ctypes.POINTER(ctypes.c_double)
_lib = None


def build():
    r = subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, ctypes.c_int, ctypes.c_double]
    lib.orc_destroy.argtypes = [ctypes.c_void_p]
    v = ctypes.c_void_p
    for name, nargs in [("orc_compute_rhs", 2), ("orc_stage_rhs", 2), ("orc_filter_exp", 2), ("orc_sponge", 2),
                        ("orc_scalar_ghosts", 3), ("orc_vector_ghosts", 4), ("orc_gradient", 4),
                        ("orc_divergence", 4), ("orc_ddx", 4), ("orc_ddy", 4), ("orc_filter_xi", 4),
                        ("orc_filter_eta", 4), ("orc_filter_r", 2), ("orc_interp_x", 2), ("orc_interp_y", 2)]:
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = [v] + [_D] * nargs
    lib.orc_bc_rhs.argtypes = [v, _D]
    lib.orc_set_c2_field.argtypes = [v, _D]
    lib.orc_ddr.argtypes = [v, _D, _D, ctypes.c_long]
    lib.orc_rk4.argtypes = [v, _D, ctypes.c_int, ctypes.c_double, _D]
    lib.orc_convert.argtypes = [v, ctypes.c_int, _D, _D, _D, _D]
    lib.orc_sponge_sigma.restype = ctypes.c_double
    lib.orc_sponge_sigma.argtypes = [v, ctypes.c_int]
    lib.orc_get_tables.argtypes = [v, _D, _D, _D, _D, _D, _D]
    lib.orc_rk_coeffs.argtypes = [_D, _D]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_D)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """The reference's hot path on the CPU for one 6-panel grid."""

    def __init__(self, n, nr, gr, c2, rho0, g0, sponge_len, a_sponge=0.1):
        self.lib = _load()
        self.n, self.nr = n, nr
        self.shape = (6, 5, nr, n, n)
        self._keep = [_c(gr), _c(c2), _c(rho0), _c(g0)]
        self.h = self.lib.orc_create(n, nr, *[_p(a) for a in self._keep], int(sponge_len), float(a_sponge))
        if not self.h:
            raise ValueError("oracle: unsupported sizes")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def _out(self, shape=None):
        return np.zeros(shape or self.shape)

    def set_c2_field(self, c2):
        """c2(i,j,k) per panel, shape (6, nr, n, n); None restores the radial profile."""
        if c2 is None:
            self.lib.orc_set_c2_field(self.h, None)
        else:
            self.lib.orc_set_c2_field(self.h, _p(_c(c2)))

    def compute_rhs(self, q):
        out = self._out()
        self.lib.orc_compute_rhs(self.h, _p(_c(q)), _p(out))
        return out

    def stage_rhs(self, q):
        """COMPUTE_RHS, BC_RHS, SPONGE as one stage applies them."""
        out = self._out()
        self.lib.orc_stage_rhs(self.h, _p(_c(q)), _p(out))
        return out

    def bc_rhs(self, rhs):
        rhs = _c(rhs).copy()
        self.lib.orc_bc_rhs(self.h, _p(rhs))
        return rhs

    def sponge(self, q, rhs):
        rhs = _c(rhs).copy()
        self.lib.orc_sponge(self.h, _p(_c(q)), _p(rhs))
        return rhs

    def filter_exp(self, q):
        out = self._out()
        self.lib.orc_filter_exp(self.h, _p(_c(q)), _p(out))
        return out

    def rk4(self, q, nsteps, dt, time=0.0):
        q = _c(q).copy()
        t = ctypes.c_double(time)
        self.lib.orc_rk4(self.h, _p(q), int(nsteps), float(dt), ctypes.byref(t))
        return q, t.value

    def _gshapes(self):
        return (6, 2, self.nr, self.n, 3), (6, 2, self.nr, 3, self.n)

    def scalar_ghosts(self, f):
        sx, sy = self._gshapes()
        gx, gy = np.zeros(sx), np.zeros(sy)
        self.lib.orc_scalar_ghosts(self.h, _p(_c(f)), _p(gx), _p(gy))
        return gx, gy

    def vector_ghosts(self, vxi, veta):
        sx, sy = self._gshapes()
        gx, gy = np.zeros(sx), np.zeros(sy)
        self.lib.orc_vector_ghosts(self.h, _p(_c(vxi)), _p(_c(veta)), _p(gx), _p(gy))
        return gx, gy

    def gradient(self, p):
        outs = [np.zeros((6, self.nr, self.n, self.n)) for _ in range(3)]
        self.lib.orc_gradient(self.h, _p(_c(p)), *[_p(o) for o in outs])
        return outs

    def divergence(self, vxi, veta, vr):
        out = np.zeros((6, self.nr, self.n, self.n))
        self.lib.orc_divergence(self.h, _p(_c(vxi)), _p(_c(veta)), _p(_c(vr)), _p(out))
        return out

    def ddr(self, f):
        f = _c(f)
        ncol = f.size // self.nr
        out = np.zeros_like(f)
        self.lib.orc_ddr(self.h, _p(f), _p(out), ncol)
        return out

    def ddx(self, f, fp, fn):
        out = np.zeros((self.nr, self.n, self.n))
        self.lib.orc_ddx(self.h, _p(_c(f)), _p(out), _p(_c(fp)), _p(_c(fn)))
        return out

    def ddy(self, f, fp, fn):
        out = np.zeros((self.nr, self.n, self.n))
        self.lib.orc_ddy(self.h, _p(_c(f)), _p(out), _p(_c(fp)), _p(_c(fn)))
        return out

    def filter_xi(self, g, gp, gn):
        out = np.zeros((self.nr, self.n, self.n))
        self.lib.orc_filter_xi(self.h, _p(_c(g)), _p(out), _p(_c(gp)), _p(_c(gn)))
        return out

    def filter_eta(self, g, gp, gn):
        out = np.zeros((self.nr, self.n, self.n))
        self.lib.orc_filter_eta(self.h, _p(_c(g)), _p(out), _p(_c(gp)), _p(_c(gn)))
        return out

    def filter_r(self, g):
        out = np.zeros((self.nr, self.n, self.n))
        self.lib.orc_filter_r(self.h, _p(_c(g)), _p(out))
        return out

    def interp_x(self, strip):
        out = np.zeros((self.nr, 3, self.n))
        self.lib.orc_interp_x(self.h, _p(_c(strip)), _p(out))
        return out

    def interp_y(self, strip):
        out = np.zeros((self.nr, self.n, 3))
        self.lib.orc_interp_y(self.h, _p(_c(strip)), _p(out))
        return out

    def convert(self, panel, vxi, veta):
        th, ph = np.zeros_like(_c(vxi)), np.zeros_like(_c(vxi))
        self.lib.orc_convert(self.h, int(panel), _p(_c(vxi)), _p(_c(veta)), _p(th), _p(ph))
        return th, ph

    def sponge_sigma(self, gk):
        return self.lib.orc_sponge_sigma(self.h, int(gk))

    def tables(self):
        n, nr = self.n, self.nr
        out = dict(scal=np.zeros(4), n1=np.zeros(8 * n), r1=np.zeros(3 * nr), n2=np.zeros(4 * n * n),
                   interp=np.zeros(27 * n), ghost=np.zeros(24))
        self.lib.orc_get_tables(self.h, *[_p(out[k]) for k in ("scal", "n1", "r1", "n2", "interp", "ghost")])
        return out


def rk_coeffs():
    lib = _load()
    a, b = np.zeros(5), np.zeros(5)
    lib.orc_rk_coeffs(_p(a), _p(b))
    return a, b
