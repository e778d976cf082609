"""solar_compute_rhs_noalloc: the right-hand side on device arrays with the
panel-exchange ghost strips in the library's static device scratch (no
allocation in the call).  Checked through the C ABI against the allocating
entry point solar_compute_rhs, bit for bit, on two calls in a row (the
second on a smaller grid, sharing the same scratch), and for the size limit.

The device buffers come from the HIP runtime the library itself links
(/opt/rocm), called through ctypes: torch's wheel carries a second copy of the
HIP runtime, and two runtimes in one process need not agree on the devices.
"""
import ctypes
import os

import numpy as np
import pytest

from solar_amd import RHS_BC, RHS_SPONGE, SolarError, Solver, background, generate_ic, load_library

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    """The HIP runtime libsolar_amd.so is linked against (loaded with it)."""
    load_library()
    maps = open("/proc/self/maps").read().split()
    paths = sorted({p for p in maps if os.path.basename(p).startswith("libamdhip64.so") and "/torch/" not in p})
    hip = ctypes.CDLL(paths[0] if paths else "libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class _Dev:
    """A device buffer holding a copy of a host array."""

    def __init__(self, hip, host):
        self.hip, self.nbytes = hip, host.nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, _H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _case(n, nr, seed, bits=64):
    gr, c2, rho0, g0 = background(1, nr)
    q, _ = generate_ic(1, n, nr, gr, c2, seed, [1e-3, 1e-6, 6.0])
    q = q + 1e-4 * np.random.default_rng(seed).standard_normal(q.shape)
    return Solver(n, nr, gr, c2, rho0, g0, max(1, round(0.08 * nr)), 0.1, state_bits=bits), q


def test_noalloc_rhs_equals_compute_rhs_on_two_calls():
    hip = _hip()
    flags = RHS_BC | RHS_SPONGE
    # first call on a 6x48x48x40 grid, then on a smaller 6x16x16x24 grid
    for n, nr, seed in ((48, 40, 3), (16, 24, 4)):
        print(f"[case] noalloc rhs n={n} nr={nr} flags={flags}", flush=True)
        s, q = _case(n, nr, seed)
        with s:
            want = s.compute_rhs(q, flags)
            qd, out = _Dev(hip, q), _Dev(hip, np.full_like(q, np.nan))
            s.compute_rhs_noalloc(qd.ptr.value, out.ptr.value, flags)
            s.sync()
            got = out.get(q)
            qd.free()
            out.free()
        ndiff = int(np.count_nonzero(got != want))
        print(f"[noalloc] n={n} nr={nr}: differing elements {ndiff} of {got.size}", flush=True)
        assert ndiff == 0


def test_noalloc_rhs_fp32_state_and_caller_stream():
    hip = _hip()
    s, q = _case(32, 28, 5, bits=32)
    with s:
        want = s.compute_rhs(q, RHS_BC)
        q32 = q.astype(np.float32)
        qd, out = _Dev(hip, q32), _Dev(hip, np.full_like(q32, np.nan))
        st = ctypes.c_void_p()
        assert hip.hipStreamCreate(ctypes.byref(st)) == 0
        s.compute_rhs_noalloc(qd.ptr.value, out.ptr.value, RHS_BC, stream=st.value)
        assert hip.hipStreamSynchronize(st) == 0
        got = out.get(q32).astype(np.float64)
        hip.hipStreamDestroy(st)
        qd.free()
        out.free()
    assert np.array_equal(got, want)


def test_noalloc_rhs_rejects_grids_beyond_the_scratch():
    # n * nr = 352 * 336 = 118272 > 116508: four ghost sets need more than 128 MiB
    hip = _hip()
    gr, c2, rho0, g0 = background(1, 336)
    with Solver(352, 336, gr, c2, rho0, g0, 27, 0.1, state_bits=32) as s:
        # real buffers of the right size, so nothing would fault if the check let the call through
        size = int(np.prod(s.shape)) * 4
        bufs = [ctypes.c_void_p(), ctypes.c_void_p()]
        for b in bufs:
            assert hip.hipMalloc(ctypes.byref(b), size) == 0
        try:
            with pytest.raises(SolarError, match="scratch"):
                s.compute_rhs_noalloc(bufs[0].value, bufs[1].value, 0)
            s.sync()
        finally:
            hip.hipDeviceSynchronize()
            for b in bufs:
                hip.hipFree(b)
