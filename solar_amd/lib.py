"""ctypes binding of libsolar_amd.so (the C ABI of include/solar_amd.h)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SOLAR_AMD_LIB") or os.path.join(_HERE, "libsolar_amd.so")

RHS_BC = 1
RHS_SPONGE = 2
# header:68-76
ALPHA_RK4 = (0.0, -4.801594388477952e-01, -1.404247195200064e+00, -2.016477077503357e+00,
             -1.056444269767207e+00)
BETA_RK4 = (1.028639988104959e-01, 7.408540575766301e-01, 7.426530946683799e-01,
            4.694937902357905e-01, 1.881733382887932e-01)

_D = ctypes.POINTER(ctypes.c_double)
_lib = None

# (name, restype, argtypes) of every entry point declared in include/solar_amd.h
_IP = ctypes.POINTER(ctypes.c_int)

EXPORTS = [
    ("solar_init", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, _D, _D, _D, _D,
                                  ctypes.c_int, ctypes.c_double]),
    ("solar_init_ex", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, _D, _D, _D, _D,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    ("solar_state_bits", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    ("solar_finalize", ctypes.c_int, [ctypes.c_void_p]),
    ("solar_set_state", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("solar_get_state", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("solar_set_state_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("solar_get_state_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("solar_rk4", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _D]),
    ("solar_rk4_async", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    ("solar_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("solar_compute_rhs", ctypes.c_int, [ctypes.c_void_p, _D, _D, ctypes.c_int]),
    ("solar_compute_rhs_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_void_p]),
    ("solar_filter_exp", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("solar_scalar_ghosts", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D]),
    ("solar_vector_ghosts", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _D]),
    ("solar_check_finite", ctypes.c_int, [ctypes.c_void_p]),
    ("solar_norm_vr", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("solar_kernel_times", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("solar_device_bytes", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]),
    ("solar_algorithmic_bytes", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("solar_set_c2_field", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("solar_set_adjoint_options", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("solar_set_adjoint_cache", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("solar_adjoint_recomputed_steps", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    ("solar_set_receivers", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    ("solar_record", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _D]),
    ("solar_adj_rhs", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("solar_adj_filter", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("solar_adj_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _D, _D]),
    ("solar_misfit_gradient", ctypes.c_int, [ctypes.c_void_p, _D, ctypes.c_int, ctypes.c_double, _D, ctypes.c_int,
                                             _D, _D, _D, ctypes.POINTER(ctypes.c_int)]),
    ("solar_write_restart", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_int]),
    ("solar_read_restart", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _D, _IP]),
    ("solar_write_visual", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
    ("solar_write_radial_trace", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("solar_restart_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int]),
    ("solar_restart_read", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _D, _IP]),
    ("solar_visual_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int,
                                          ctypes.c_int]),
    ("solar_radial_trace_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _D]),
    ("solar_grid_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
    ("solar_output_radius", ctypes.c_int, [ctypes.c_int, _D, ctypes.c_double, _IP]),
    ("solar_norm", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]),
    ("solar_ens_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _IP, _IP, _D, _D, _D, _D,
                                        _IP, _D, _D]),
    ("solar_ens_finalize", ctypes.c_int, [ctypes.c_void_p]),
    ("solar_ens_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("solar_ens_tiles", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _IP]),
    ("solar_ens_set_tiles", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _IP]),
    ("solar_ens_set_state", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    ("solar_ens_get_state", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    ("solar_ens_rk4", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("solar_ens_rk4_async", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("solar_ens_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("solar_ens_info", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), _IP,
                                      ctypes.POINTER(ctypes.c_size_t)]),
    ("solar_adjoint_times", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("solar_tangent_linear", ctypes.c_int, [ctypes.c_void_p, _D, ctypes.c_int, ctypes.c_double, _D, _D]),
    ("solar_last_error", ctypes.c_char_p, []),
    ("solar_host_tables", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, ctypes.c_int, ctypes.c_double,
                                         _D, _D, _D, _D, _D, _D, _D, _D]),
    ("solar_background", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]),
    ("solar_panel_points", ctypes.c_int, [ctypes.c_int, _D]),
    ("solar_generate_ic", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_uint64, _D, _D,
                                         _D]),
]


class SolarError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SolarError(f"{path} not found: build it with `python -m solar_amd.build`")
    lib = ctypes.CDLL(path)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_D)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _check(rc, what):
    if rc != 0:
        msg = load_library().solar_last_error().decode()
        raise SolarError(f"{what} failed ({rc}): {msg}")


def background(kind, nr):
    """Radial grid and background profiles (gr, c2, rho0, g0); kind 0 uniform, 1 stratified."""
    lib = load_library()
    out = [np.zeros(nr) for _ in range(4)]
    _check(lib.solar_background(kind, nr, *[_p(a) for a in out]), "solar_background")
    return tuple(out)


def panel_points(n):
    """Unit position vectors of all grid points, shape (6, n, n, 3) indexed [p, j, i]."""
    lib = load_library()
    xyz = np.zeros((6, n, n, 3))
    _check(lib.solar_panel_points(n, _p(xyz)), "solar_panel_points")
    return xyz


def generate_ic(kind, n, nr, gr, c2, seed, params=None):
    """Seeded initial state q of shape (6, 5, nr, n, n) and a 4-vector of IC info."""
    lib = load_library()
    q = np.zeros((6, 5, nr, n, n))
    info = np.zeros(4)
    pp = None if params is None else _f64(params)
    _check(lib.solar_generate_ic(kind, n, nr, _p(_f64(gr)), _p(_f64(c2)), ctypes.c_uint64(seed),
                                 None if pp is None else _p(pp), _p(q), _p(info)), "solar_generate_ic")
    return q, info


def host_tables(n, nr, gr, c2, rho0, g0, sponge_len, a_sponge):
    """Host tables exactly as the device uses them (no GPU needed)."""
    lib = load_library()
    out = dict(scal=np.zeros(4), n1=np.zeros(8 * n), r1=np.zeros(3 * nr), n2=np.zeros(4 * n * n),
               interp=np.zeros(27 * n), ghost=np.zeros(24), sigma=np.zeros(nr), drho0dr=np.zeros(nr))
    _check(lib.solar_host_tables(n, nr, _p(_f64(gr)), _p(_f64(c2)), _p(_f64(rho0)), _p(_f64(g0)), sponge_len,
                                 a_sponge, *[_p(out[k]) for k in
                                             ("scal", "n1", "r1", "n2", "interp", "ghost", "sigma", "drho0dr")]),
           "solar_host_tables")
    return out


def _dir(d):
    return None if d is None else str(d).encode()


def restart_write(dirname, q, cue, time, step):
    """WRITE_RESTART files Restart/restart.<ddd>.<zone> for a host state (6, 5, nr, n, n)."""
    q = _f64(q)
    _check(load_library().solar_restart_write(_dir(dirname), q.shape[-1], q.shape[2], _p(q), int(cue), float(time),
                                              int(step)), "solar_restart_write")


def restart_read(dirname, n, nr, cue):
    """Inverse of restart_write: (q of shape (6, 5, nr, n, n), time, step)."""
    q = np.empty((6, 5, nr, n, n))
    t, st = ctypes.c_double(), ctypes.c_int()
    _check(load_library().solar_restart_read(_dir(dirname), n, nr, _p(q), int(cue), ctypes.byref(t),
                                             ctypes.byref(st)), "solar_restart_read")
    return q, t.value, st.value


def visual_write(dirname, q, cue, o_rad):
    """WRITE_VISUAL files Visual/<ddd>/vis_<rank> (o_rad 1-based)."""
    q = _f64(q)
    _check(load_library().solar_visual_write(_dir(dirname), q.shape[-1], q.shape[2], _p(q), int(cue), int(o_rad)),
           "solar_visual_write")


def radial_trace_write(dirname, q):
    """Appends one WRITE_RADIAL_TRACE record to trace_<zone> for every zone."""
    q = _f64(q)
    _check(load_library().solar_radial_trace_write(_dir(dirname), q.shape[-1], q.shape[2], _p(q)),
           "solar_radial_trace_write")


def grid_write(dirname, n, nr):
    """WRITE_GRID files Visual/grid/grid_<rank>."""
    _check(load_library().solar_grid_write(_dir(dirname), int(n), int(nr)), "solar_grid_write")


def output_radius(gr, obsheight):
    """o_rad (1-based) of SOLAR_DATA for an observation height."""
    gr = _f64(gr)
    k = ctypes.c_int()
    _check(load_library().solar_output_radius(gr.size, _p(gr), float(obsheight), ctypes.byref(k)),
           "solar_output_radius")
    return k.value


class Solver:
    """The RK4 hot path on one GPU: state q(n,n,nr,5,6) resident in HBM.

    Mirrors the reference's boundary: construct (the SETUP_* calls), set_state
    (GENERATE_IC filling /RKWORK/ q), rk4 (call RK4, cube.f90:74), get_state.
    Arrays are numpy float64 of shape (6, 5, nr, n, n) (= Fortran q(n,n,nr,5,6)).
    """

    def __init__(self, n, nr, gr, c2, rho0, g0, sponge_len, a_sponge=0.1, state_bits=64):
        self.lib = load_library()
        self.n, self.nr = n, nr
        self.shape = (6, 5, nr, n, n)
        self.state_bits = state_bits
        h = ctypes.c_void_p()
        _check(self.lib.solar_init_ex(ctypes.byref(h), n, nr, _p(_f64(gr)), _p(_f64(c2)), _p(_f64(rho0)),
                                      _p(_f64(g0)), int(sponge_len), float(a_sponge), int(state_bits)),
               "solar_init")
        self.h = h
        self.time = 0.0

    def close(self):
        if self.h:
            self.lib.solar_finalize(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        a = _f64(a)
        if a.size != int(np.prod(self.shape)):
            raise SolarError(f"array of {a.size} elements, expected shape {self.shape}")
        return a

    def set_state(self, q):
        _check(self.lib.solar_set_state(self.h, _p(self._arr(q))), "solar_set_state")

    def get_state(self):
        q = np.empty(self.shape)
        _check(self.lib.solar_get_state(self.h, _p(q)), "solar_get_state")
        return q

    def set_state_device(self, ptr):
        _check(self.lib.solar_set_state_device(self.h, ctypes.c_void_p(ptr)), "solar_set_state_device")

    def get_state_device(self, ptr):
        _check(self.lib.solar_get_state_device(self.h, ctypes.c_void_p(ptr)), "solar_get_state_device")

    def rk4(self, nsteps, dt):
        t = ctypes.c_double(self.time)
        _check(self.lib.solar_rk4(self.h, int(nsteps), float(dt), ctypes.byref(t)), "solar_rk4")
        self.time = t.value
        return self.time

    def rk4_async(self, nsteps, dt):
        _check(self.lib.solar_rk4_async(self.h, int(nsteps), float(dt)), "solar_rk4_async")

    def sync(self):
        _check(self.lib.solar_sync(self.h), "solar_sync")

    def compute_rhs(self, q, flags=0):
        out = np.empty(self.shape)
        _check(self.lib.solar_compute_rhs(self.h, _p(self._arr(q)), _p(out), int(flags)), "solar_compute_rhs")
        return out

    def compute_rhs_noalloc(self, q_ptr, rhs_ptr, flags=0, stream=None):
        """solar_compute_rhs_noalloc on device pointers (ints), enqueued on
        `stream` (a hipStream_t as int, None = the handle's stream) without
        waiting; allocates no device memory."""
        _check(self.lib.solar_compute_rhs_noalloc(self.h, ctypes.c_void_p(q_ptr), ctypes.c_void_p(rhs_ptr),
                                                  int(flags), ctypes.c_void_p(stream) if stream else None),
               "solar_compute_rhs_noalloc")

    def filter_exp(self, q):
        out = np.empty(self.shape)
        _check(self.lib.solar_filter_exp(self.h, _p(self._arr(q)), _p(out)), "solar_filter_exp")
        return out

    def _gshapes(self):
        n, nr = self.n, self.nr
        return (6, 2, nr, n, 3), (6, 2, nr, 3, n)

    def scalar_ghosts(self, f):
        f = _f64(f)
        sx, sy = self._gshapes()
        gx, gy = np.empty(sx), np.empty(sy)
        _check(self.lib.solar_scalar_ghosts(self.h, _p(f), _p(gx), _p(gy)), "solar_scalar_ghosts")
        return gx, gy

    def vector_ghosts(self, vxi, veta):
        sx, sy = self._gshapes()
        gx, gy = np.empty(sx), np.empty(sy)
        _check(self.lib.solar_vector_ghosts(self.h, _p(_f64(vxi)), _p(_f64(veta)), _p(gx), _p(gy)),
               "solar_vector_ghosts")
        return gx, gy

    def check_finite(self):
        rc = self.lib.solar_check_finite(self.h)
        if rc == -3:
            return False
        _check(rc, "solar_check_finite")
        return True

    def norm_vr(self):
        v = ctypes.c_double()
        _check(self.lib.solar_norm_vr(self.h, ctypes.byref(v)), "solar_norm_vr")
        return v.value

    def norm(self, field, reference_formula=False):
        """cube.f90:77 diagnostic of field 0..4 (auxiliary.f90:100-125 when reference_formula)."""
        v = ctypes.c_double()
        _check(self.lib.solar_norm(self.h, int(field), int(bool(reference_formula)), ctypes.byref(v)), "solar_norm")
        return v.value

    def write_restart(self, dirname, cue, time, step):
        _check(self.lib.solar_write_restart(self.h, _dir(dirname), int(cue), float(time), int(step)),
               "solar_write_restart")

    def read_restart(self, dirname, cue):
        t, st = ctypes.c_double(), ctypes.c_int()
        _check(self.lib.solar_read_restart(self.h, _dir(dirname), int(cue), ctypes.byref(t), ctypes.byref(st)),
               "solar_read_restart")
        return t.value, st.value

    def write_visual(self, dirname, cue, o_rad):
        _check(self.lib.solar_write_visual(self.h, _dir(dirname), int(cue), int(o_rad)), "solar_write_visual")

    def write_radial_trace(self, dirname):
        _check(self.lib.solar_write_radial_trace(self.h, _dir(dirname)), "solar_write_radial_trace")

    def kernel_times(self):
        out = np.zeros(5)
        _check(self.lib.solar_kernel_times(self.h, _p(out)), "solar_kernel_times")
        return dict(stage_ms=out[0], stage_gather_ms=out[1], filter_ms=out[2], filter_gather_ms=out[3],
                    step_ms=out[4])

    def algorithmic_bytes(self):
        """Algorithmic HBM bytes per grid point of each launch of a step:
        (stage 1, ..., stage 5, filter)."""
        out = np.zeros(6)
        _check(self.lib.solar_algorithmic_bytes(self.h, _p(out)), "solar_algorithmic_bytes")
        return tuple(float(x) for x in out)

    def device_bytes(self):
        v = ctypes.c_size_t()
        _check(self.lib.solar_device_bytes(self.h, ctypes.byref(v)), "solar_device_bytes")
        return v.value

    # ---------------------------------------------------------- sensitivities

    def set_c2_field(self, c2):
        """Full c^2 field, shape (6, nr, n, n) (= Fortran c2(n,n,nr,6)); None restores the radial profile."""
        if c2 is None:
            _check(self.lib.solar_set_c2_field(self.h, None), "solar_set_c2_field")
            return
        c2 = _f64(c2)
        if c2.size != 6 * self.nr * self.n * self.n:
            raise SolarError(f"c2 field of {c2.size} elements, expected shape {(6, self.nr, self.n, self.n)}")
        _check(self.lib.solar_set_c2_field(self.h, _p(c2)), "solar_set_c2_field")

    def set_adjoint_options(self, div_bits=64):
        """Storage of the stage divergences kept by misfit_gradient (64 exact, 32 half the memory)."""
        _check(self.lib.solar_set_adjoint_options(self.h, int(div_bits)), "solar_set_adjoint_options")

    def set_adjoint_cache(self, enable=True):
        _check(self.lib.solar_set_adjoint_cache(self.h, int(bool(enable))), "solar_set_adjoint_cache")

    @property
    def recomputed_steps(self):
        v = ctypes.c_int()
        _check(self.lib.solar_adjoint_recomputed_steps(self.h, ctypes.byref(v)), "solar_adjoint_recomputed_steps")
        return v.value

    def set_receivers(self, pkji):
        """Receivers of p' at (panel, k, j, i), 0-based; array of shape (nrec, 4)."""
        a = np.ascontiguousarray(pkji, dtype=np.int32).reshape(-1, 4)
        self.nrec = a.shape[0]
        _check(self.lib.solar_set_receivers(self.h, a.shape[0], a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
               "solar_set_receivers")

    def record(self, nsteps, dt):
        """nsteps RK4 steps, p' at the receivers after each: array (nsteps, nrec)."""
        out = np.zeros((int(nsteps), getattr(self, "nrec", 0)))
        _check(self.lib.solar_record(self.h, int(nsteps), float(dt), _p(out)), "solar_record")
        return out

    def tangent_linear(self, q0, nsteps, dt, dc2):
        """d(traces)/d(c^2) . dc2, shape (nsteps, nrec)."""
        dc2 = _f64(dc2)
        if dc2.size != 6 * self.nr * self.n * self.n:
            raise SolarError(f"dc2 of {dc2.size} values, expected shape {(6, self.nr, self.n, self.n)}")
        out = np.zeros((int(nsteps), getattr(self, "nrec", 0)))
        _check(self.lib.solar_tangent_linear(self.h, _p(self._arr(q0)), int(nsteps), float(dt), _p(dc2), _p(out)),
               "solar_tangent_linear")
        return out

    def adjoint_times(self):
        """Phases of the last misfit_gradient (s): workspace, forward, recompute, adjoint."""
        out = np.zeros(4)
        _check(self.lib.solar_adjoint_times(self.h, _p(out)), "solar_adjoint_times")
        return dict(workspace_s=out[0], forward_s=out[1], recompute_s=out[2], adjoint_s=out[3])

    def adj_rhs(self, w):
        out = np.empty(self.shape)
        _check(self.lib.solar_adj_rhs(self.h, _p(self._arr(w)), _p(out)), "solar_adj_rhs")
        return out

    def adj_filter(self, o):
        out = np.empty(self.shape)
        _check(self.lib.solar_adj_filter(self.h, _p(self._arr(o)), _p(out)), "solar_adj_filter")
        return out

    def adj_step(self, lam, dt):
        out = np.empty(self.shape)
        _check(self.lib.solar_adj_step(self.h, float(dt), _p(self._arr(lam)), _p(out)), "solar_adj_step")
        return out

    def misfit_gradient(self, q0, nsteps, dt, data, checkpoint_every=0):
        """(J, dJ/dc2 of shape (6, nr, n, n), modelled traces (nsteps, nrec)).

        checkpoint_every <= 0 picks the longest segment that fits in device
        memory; self.segment_used holds the length used."""
        data = _f64(data)
        nrec = getattr(self, "nrec", 0)
        if data.size != int(nsteps) * nrec:
            raise SolarError(f"data of {data.size} values, expected ({nsteps}, {nrec})")
        J = ctypes.c_double()
        seg = ctypes.c_int()
        grad = np.zeros((6, self.nr, self.n, self.n))
        seis = np.zeros((int(nsteps), nrec))
        _check(self.lib.solar_misfit_gradient(self.h, _p(self._arr(q0)), int(nsteps), float(dt), _p(data),
                                              int(checkpoint_every), ctypes.cast(ctypes.byref(J), _D), _p(grad), _p(seis),
                                              ctypes.byref(seg)),
               "solar_misfit_gradient")
        self.segment_used = seg.value
        return J.value, grad, seis


class Ensemble:
    """Independent simulations with uneven grids stepped together.

    members: sequence of dicts with n, nr, gr, c2, rho0, g0, sponge_len,
    a_sponge, dt.  States are numpy float64 of shape (6, 5, nr, n, n)."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = [dict(m) for m in members]
        k = len(self.members)
        ia = lambda key: (ctypes.c_int * k)(*[int(m[key]) for m in self.members])  # noqa: E731
        cat = lambda key: _f64(np.concatenate([_f64(m[key]) for m in self.members]))  # noqa: E731
        self._keep = [cat(x) for x in ("gr", "c2", "rho0", "g0")]
        asp = _f64([m.get("a_sponge", 0.1) for m in self.members])
        dts = _f64([m["dt"] for m in self.members])
        h = ctypes.c_void_p()
        _check(self.lib.solar_ens_create(ctypes.byref(h), k, ia("n"), ia("nr"), *[_p(a) for a in self._keep],
                                         ia("sponge_len"), _p(asp), _p(dts)), "solar_ens_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.solar_ens_finalize(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, m):
        d = self.members[m]
        return (6, 5, d["nr"], d["n"], d["n"])

    def autotune(self, reps=3):
        _check(self.lib.solar_ens_autotune(self.h, int(reps)), "solar_ens_autotune")

    def tiles(self, m):
        out = (ctypes.c_int * 6)()
        _check(self.lib.solar_ens_tiles(self.h, int(m), out), "solar_ens_tiles")
        return tuple(out)

    def set_tiles(self, m, six):
        arr = (ctypes.c_int * 6)(*[int(v) for v in six])
        _check(self.lib.solar_ens_set_tiles(self.h, int(m), arr), "solar_ens_set_tiles")

    def set_state(self, m, q):
        q = _f64(q)
        if q.size != int(np.prod(self._shape(m))):
            raise SolarError(f"member {m}: array of {q.size} elements, expected {self._shape(m)}")
        _check(self.lib.solar_ens_set_state(self.h, int(m), _p(q)), "solar_ens_set_state")

    def get_state(self, m):
        q = np.empty(self._shape(m))
        _check(self.lib.solar_ens_get_state(self.h, int(m), _p(q)), "solar_ens_get_state")
        return q

    def rk4(self, nsteps):
        _check(self.lib.solar_ens_rk4(self.h, int(nsteps)), "solar_ens_rk4")

    def rk4_async(self, nsteps):
        _check(self.lib.solar_ens_rk4_async(self.h, int(nsteps)), "solar_ens_rk4_async")

    def sync(self):
        _check(self.lib.solar_ens_sync(self.h), "solar_ens_sync")

    def info(self):
        pts, nb = ctypes.c_size_t(), ctypes.c_size_t()
        nl = ctypes.c_int()
        _check(self.lib.solar_ens_info(self.h, ctypes.byref(pts), ctypes.byref(nl), ctypes.byref(nb)),
               "solar_ens_info")
        return dict(points=pts.value, launches_per_step=nl.value, device_bytes=nb.value)
