"""MI355X (gfx950) implementation of TopologicalCYMSolver's real-time lattice
update loop, driven through the C ABI of libtcym.so (include/tcym.h).

`Solver` owns one device context.  Its methods carry the reference's names
(Dynamics::Update -> dynamics_update, ChernSimonsNumber::NCsDot -> ncs_dot,
...) and take/return host arrays in the reference's AoS layout
(U[s_U*(mu + 3*site)], E[a + n_E*(mu + 3*site)], site = x + Nx*(y + Ny*z)).
There is no CPU fallback: without the built library or a gfx950 device every
call raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TCYM_LIBRARY") or os.path.join(PKG, "libtcym.so")

OK, E_ARG, E_HIP, E_NOCONV, E_NONFINITE, E_UNSUPPORTED, E_IO = 0, -1, -2, -3, -4, -5, -6
FP64, FP32 = 0, 1
FORM_AUTO, FORM_LINK, FORM_PLAQUETTE = 0, 1, 2
U_HOT, U_COOL, U_OLD, U_NEW, U_MID, U_AUX = range(6)
E_HOT, E_COOL, E_MID, E_XI, E_AUX = range(5)

# every exported symbol of include/tcym.h
EXPORTS = [
    "tcym_create", "tcym_destroy", "tcym_last_error", "tcym_version",
    "tcym_set_links", "tcym_get_links", "tcym_set_efield", "tcym_get_efield",
    "tcym_copy_links", "tcym_copy_efield", "tcym_zero_efield", "tcym_unit_links",
    "tcym_random_links", "tcym_random_efield",
    "tcym_leapfrog", "tcym_update_efield", "tcym_update_links", "tcym_langevin", "tcym_langevin_xi", "tcym_evolve_ncs",
    "tcym_noise", "tcym_flow", "tcym_cool", "tcym_cooling_start", "tcym_cooling_update", "tcym_reunitarize",
    "tcym_edotb", "tcym_ncsdot", "tcym_ncsdot_noalloc", "tcym_edotb_map", "tcym_bulk", "tcym_hard_scales", "tcym_gauss", "tcym_unitarity",
    "tcym_gauss_restore", "tcym_geodesic", "tcym_delta_ncs_cooling", "tcym_vacuum_ncs",
    "tcym_block", "tcym_level", "tcym_cool_blocked", "tcym_cool_continue", "tcym_continue_cooling", "tcym_thermalize",
    "tcym_quasi_particles",
    "tcym_set_update_form", "tcym_timing", "tcym_kernel_time", "tcym_timing_reset", "tcym_sync", "tcym_device_bytes",
    "tcym_get_params",
]
# every exported symbol of include/tcym_io.h
IO_EXPORTS = [
    "tcym_io_last_error", "tcym_write_text_links", "tcym_write_text_efield", "tcym_read_text_links",
    "tcym_read_text_efield", "tcym_save_text", "tcym_load_text", "tcym_write_checkpoint",
    "tcym_read_checkpoint", "tcym_save_checkpoint", "tcym_load_checkpoint",
]
# every exported symbol of include/tcym_batch.h
BATCH_EXPORTS = [
    "tcym_batch_create", "tcym_batch_destroy", "tcym_batch_last_error", "tcym_batch_sync",
    "tcym_batch_device_bytes", "tcym_batch_set_links", "tcym_batch_get_links", "tcym_batch_set_efield",
    "tcym_batch_get_efield", "tcym_batch_initial_condition", "tcym_batch_set_step", "tcym_batch_get_step",
    "tcym_batch_leapfrog", "tcym_batch_langevin", "tcym_batch_flow", "tcym_batch_evolve",
    "tcym_batch_cooling_start", "tcym_batch_cooling_update", "tcym_batch_ncsdot", "tcym_batch_bulk",
    "tcym_batch_gauss", "tcym_batch_unitarity", "tcym_batch_tune", "tcym_batch_set_schedule", "tcym_batch_timing",
    "tcym_batch_kernel_time", "tcym_batch_timing_reset",
]


class TcymError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("tcym error %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int), ("precision", ctypes.c_int), ("N", ctypes.c_int * 3),
                ("a", ctypes.c_double * 3), ("aScale", ctypes.c_double), ("device", ctypes.c_int)]


class BatchParams(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int), ("precision", ctypes.c_int), ("count", ctypes.c_int),
                ("N", ctypes.POINTER(ctypes.c_int)), ("a", ctypes.c_double * 3), ("aScale", ctypes.c_double),
                ("device", ctypes.c_int)]


class CheckpointInfo(ctypes.Structure):
    """tcym_checkpoint_info (include/tcym_io.h)."""
    _fields_ = [("nc", ctypes.c_int), ("precision", ctypes.c_int), ("N", ctypes.c_int * 3),
                ("a", ctypes.c_double * 3), ("aScale", ctypes.c_double), ("step", ctypes.c_uint64),
                ("seed", ctypes.c_uint64), ("time", ctypes.c_double), ("acc", ctypes.c_double * 8),
                ("has_efield", ctypes.c_int)]

    def as_dict(self):
        return {"nc": self.nc, "precision": self.precision, "N": tuple(self.N), "a": tuple(self.a),
                "aScale": self.aScale, "step": self.step, "seed": self.seed, "time": self.time,
                "acc": tuple(self.acc), "has_efield": bool(self.has_efield)}


_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def load_library(path=LIB_PATH):
    """Load libtcym.so and declare its signatures.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libtcym.so is not built (%s); run python -m topologicalcymsolver_amd.build" % path)
    L = ctypes.CDLL(path)
    V, I, D, P = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Params)
    U32, U64, LL = ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_longlong)
    sig = {
        "tcym_create": [P, ctypes.POINTER(V)], "tcym_destroy": [V],
        "tcym_set_links": [V, I, _dp], "tcym_get_links": [V, I, _dp],
        "tcym_set_efield": [V, I, _dp], "tcym_get_efield": [V, I, _dp],
        "tcym_copy_links": [V, I, I], "tcym_copy_efield": [V, I, I], "tcym_zero_efield": [V, I],
        "tcym_unit_links": [V, I], "tcym_random_links": [V, I, U64, D], "tcym_random_efield": [V, I, U64, D],
        "tcym_leapfrog": [V, I, I, I, D], "tcym_update_efield": [V, I, I, D], "tcym_update_links": [V, I, I, D],
        "tcym_langevin": [V, I, I, I, D, D, D, U64, U32], "tcym_langevin_xi": [V, I, I, I, D, D],
        "tcym_evolve_ncs": [V, I, I, I, I, D, D, D, U64, U32, _dp],
        "tcym_noise": [V, I, U64, U32, D], "tcym_flow": [V, I, I, I, D, I, _dp],
        "tcym_cool": [V, I, I, I, D], "tcym_cooling_start": [V, I, I, D],
        "tcym_cooling_update": [V, I, I, D, I, _dp], "tcym_reunitarize": [V, I],
        "tcym_edotb": [V, I, I, I, _dp], "tcym_ncsdot": [V, I, I, _dp], "tcym_ncsdot_noalloc": [V, I, I, V], "tcym_edotb_map": [V, I, I, _dp],
        "tcym_bulk": [V, I, I, _dp], "tcym_gauss": [V, I, I, _dp, _dp], "tcym_unitarity": [V, I, _dp, _dp],
        "tcym_hard_scales": [V, I, _dp],
        "tcym_gauss_restore": [V, I, I, I, D, ctypes.POINTER(I), _dp],
        "tcym_geodesic": [V, I, I, I, I, D], "tcym_delta_ncs_cooling": [V, I, I, _dp],
        "tcym_vacuum_ncs": [V, I, _dp], "tcym_timing": [V, I], "tcym_set_update_form": [V, I],
        "tcym_block": [V, I, I], "tcym_level": [V, I, I, ctypes.POINTER(I), ctypes.POINTER(I)],
        "tcym_cool_blocked": [V, I, I, I, D, I, I, ctypes.POINTER(I)],
        "tcym_cool_continue": [V, I, I, I, D, I, I, ctypes.POINTER(I)],
        "tcym_continue_cooling": [V, I, I, I, I, I, D, D, I, I, ctypes.POINTER(I), _dp, _dp],
        "tcym_kernel_time": [V, ctypes.c_char_p, _dp, LL], "tcym_timing_reset": [V], "tcym_sync": [V],
        "tcym_device_bytes": [V, LL],
        "tcym_batch_create": [ctypes.POINTER(BatchParams), ctypes.POINTER(V)], "tcym_batch_destroy": [V],
        "tcym_batch_sync": [V], "tcym_batch_device_bytes": [V, LL],
        "tcym_batch_set_links": [V, I, I, _dp], "tcym_batch_get_links": [V, I, I, _dp],
        "tcym_batch_set_efield": [V, I, I, _dp], "tcym_batch_get_efield": [V, I, I, _dp],
        "tcym_batch_initial_condition": [V, U64, D, D, I, D, ctypes.POINTER(I)],
        "tcym_batch_set_step": [V, U32], "tcym_batch_get_step": [V, ctypes.POINTER(U32)],
        "tcym_batch_leapfrog": [V, I, D], "tcym_batch_langevin": [V, I, D, D, D, U64],
        "tcym_batch_flow": [V, I, I, D], "tcym_batch_evolve": [V, I, I, I, D, D, D, U64, _dp],
        "tcym_batch_cooling_start": [V, I, D], "tcym_batch_cooling_update": [V, I, D, _dp],
        "tcym_batch_ncsdot": [V, I, I, _dp], "tcym_batch_bulk": [V, I, I, _dp],
        "tcym_batch_gauss": [V, I, I, _dp], "tcym_batch_unitarity": [V, I, _dp],
        "tcym_batch_tune": [V, I, D, I, ctypes.POINTER(I), _dp], "tcym_batch_timing": [V, I],
        "tcym_batch_set_schedule": [V, I, I, I, I],
        "tcym_batch_kernel_time": [V, ctypes.c_char_p, _dp, LL], "tcym_batch_timing_reset": [V],
        "tcym_get_params": [V, P],
        "tcym_thermalize": [V, I, I, I, D, D, D, U64, I, D, _dp, I, ctypes.POINTER(I)],
        "tcym_quasi_particles": [V, I, I, D, D, U64],
        "tcym_write_text_links": [I, ctypes.POINTER(I), _dp, ctypes.c_char_p, I, I],
        "tcym_write_text_efield": [I, ctypes.POINTER(I), _dp, ctypes.c_char_p, I, I],
        "tcym_read_text_links": [I, ctypes.POINTER(I), _dp, ctypes.c_char_p, I],
        "tcym_read_text_efield": [I, ctypes.POINTER(I), _dp, ctypes.c_char_p, I],
        "tcym_save_text": [V, I, I, ctypes.c_char_p, ctypes.c_char_p, I, I],
        "tcym_load_text": [V, I, I, ctypes.c_char_p, ctypes.c_char_p, I],
        "tcym_write_checkpoint": [ctypes.c_char_p, ctypes.POINTER(CheckpointInfo), _dp, _dp],
        "tcym_read_checkpoint": [ctypes.c_char_p, ctypes.POINTER(CheckpointInfo), _dp, _dp],
        "tcym_save_checkpoint": [V, I, I, ctypes.c_char_p, ctypes.POINTER(CheckpointInfo)],
        "tcym_load_checkpoint": [V, I, I, ctypes.c_char_p, ctypes.POINTER(CheckpointInfo)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.tcym_last_error.argtypes = [V]
    L.tcym_last_error.restype = ctypes.c_char_p
    L.tcym_version.argtypes = []
    L.tcym_version.restype = ctypes.c_char_p
    L.tcym_batch_last_error.argtypes = [V]
    L.tcym_batch_last_error.restype = ctypes.c_char_p
    L.tcym_io_last_error.argtypes = []
    L.tcym_io_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


# ---- configuration files (include/tcym_io.h); host arrays, no device needed ----
def _io_ck(r):
    if r != OK:
        raise TcymError(r, load_library().tcym_io_last_error().decode())


def _record_size(nc, kind):
    if kind not in ("links", "efield"):
        raise ValueError("kind must be 'links' or 'efield'")
    return (4 if nc == 2 else 18) if kind == "links" else (3 if nc == 2 else 8)


def write_text(kind, nc, N, A, path, digits=15, nthreads=0):
    """IO::SaveConfiguration's text file of links or E (IO/SaveConfiguration.cpp:15-51);
    digits 15 is the reference's OUTPUT_PRECISION, 17 reads back bit-exact."""
    L = load_library()
    n = (ctypes.c_int * 3)(*N)
    A = _arr(A, 3 * N[0] * N[1] * N[2] * _record_size(nc, kind))
    fn = L.tcym_write_text_links if kind == "links" else L.tcym_write_text_efield
    _io_ck(fn(nc, n, A.ctypes.data_as(_dp), os.fsencode(path), digits, nthreads))


def read_text(kind, nc, N, path, nthreads=0):
    """IO::LoadConfiguration of one text file (IO/LoadConfiguration.cpp:46-122)."""
    L = load_library()
    n = (ctypes.c_int * 3)(*N)
    A = np.empty(3 * N[0] * N[1] * N[2] * _record_size(nc, kind))
    fn = L.tcym_read_text_links if kind == "links" else L.tcym_read_text_efield
    _io_ck(fn(nc, n, A.ctypes.data_as(_dp), os.fsencode(path), nthreads))
    return A


def _info(nc=2, precision=FP64, N=(1, 1, 1), a=(1.0, 1.0, 1.0), aScale=1.0, step=0, seed=0, time=0.0, acc=()):
    acc = tuple(acc) + (0.0,) * (8 - len(acc))
    return CheckpointInfo(nc, precision, (ctypes.c_int * 3)(*N), (ctypes.c_double * 3)(*a), aScale, step, seed,
                          time, (ctypes.c_double * 8)(*acc))


def write_checkpoint(path, U, E=None, **info):
    """Binary checkpoint of host arrays; info: nc, precision, N, a, aScale, step, seed, time, acc."""
    ci = _info(**info)
    su, ne = (4, 3) if ci.nc == 2 else (18, 8)
    V = ci.N[0] * ci.N[1] * ci.N[2]
    U = _arr(U, 3 * V * su)
    E = None if E is None else _arr(E, 3 * V * ne)
    _io_ck(load_library().tcym_write_checkpoint(os.fsencode(path), ctypes.byref(ci), U.ctypes.data_as(_dp),
                                                None if E is None else E.ctypes.data_as(_dp)))


def read_checkpoint(path, arrays=True):
    """(info dict, U, E) of a binary checkpoint; E is None for a links-only file,
    U and E are None when arrays is False."""
    L = load_library()
    ci = CheckpointInfo()
    _io_ck(L.tcym_read_checkpoint(os.fsencode(path), ctypes.byref(ci), None, None))
    if not arrays:
        return ci.as_dict(), None, None
    su, ne = (4, 3) if ci.nc == 2 else (18, 8)
    V = ci.N[0] * ci.N[1] * ci.N[2]
    U = np.empty(3 * V * su)
    E = np.empty(3 * V * ne) if ci.has_efield else None
    _io_ck(L.tcym_read_checkpoint(os.fsencode(path), ctypes.byref(ci), U.ctypes.data_as(_dp),
                                  None if E is None else E.ctypes.data_as(_dp)))
    return ci.as_dict(), U, E


def _arr(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


class Solver:
    """One lattice context on one GPU (Lattice::Init + the GLinks/EFields singletons)."""

    def __init__(self, nc, N, precision=FP64, a=(1.0, 1.0, 1.0), aScale=1.0, device=0):
        self.L = load_library()
        if isinstance(N, int):
            N = (N, N, N)
        self.nc, self.N, self.precision = nc, tuple(N), precision
        self.V = N[0] * N[1] * N[2]
        self.su = 4 if nc == 2 else 18
        self.ne = 3 if nc == 2 else 8
        p = Params(nc, precision, (ctypes.c_int * 3)(*N), (ctypes.c_double * 3)(*a), aScale, device)
        h = ctypes.c_void_p()
        r = self.L.tcym_create(ctypes.byref(p), ctypes.byref(h))
        if r != OK:
            raise TcymError(r, "tcym_create failed (group %d, N=%s, device %d)" % (nc, N, device))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.tcym_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, r):
        if r != OK:
            raise TcymError(r, self.L.tcym_last_error(self.h).decode())
        return r

    # ---- fields ----
    def set_links(self, U, slot=U_HOT):
        U = _arr(U, 3 * self.V * self.su)
        self._ck(self.L.tcym_set_links(self.h, slot, U.ctypes.data_as(_dp)))

    def get_links(self, slot=U_HOT):
        n = self.level(u=slot)[1]
        U = np.empty(3 * n[0] * n[1] * n[2] * self.su)
        self._ck(self.L.tcym_get_links(self.h, slot, U.ctypes.data_as(_dp)))
        return U

    def set_efield(self, E, slot=E_HOT):
        E = _arr(E, 3 * self.V * self.ne)
        self._ck(self.L.tcym_set_efield(self.h, slot, E.ctypes.data_as(_dp)))

    def get_efield(self, slot=E_HOT):
        n = self.level(e=slot)[1]
        E = np.empty(3 * n[0] * n[1] * n[2] * self.ne)
        self._ck(self.L.tcym_get_efield(self.h, slot, E.ctypes.data_as(_dp)))
        return E

    def copy_links(self, dst, src):
        self._ck(self.L.tcym_copy_links(self.h, dst, src))

    def copy_efield(self, dst, src):
        self._ck(self.L.tcym_copy_efield(self.h, dst, src))

    def zero_efield(self, slot=E_HOT):
        self._ck(self.L.tcym_zero_efield(self.h, slot))

    def unit_links(self, slot=U_HOT):
        self._ck(self.L.tcym_unit_links(self.h, slot))

    def random_links(self, seed, sigma, slot=U_HOT):
        self._ck(self.L.tcym_random_links(self.h, slot, seed, sigma))

    def random_efield(self, seed, sigma, slot=E_HOT):
        self._ck(self.L.tcym_random_efield(self.h, slot, seed, sigma))

    # ---- updates (reference names) ----
    def dynamics_update(self, nsteps=1, dt=0.025, u=U_HOT, e=E_HOT):
        """Dynamics::Update (DYNAMICS/Dynamics.cpp:204-218), dTau = dt."""
        self._ck(self.L.tcym_leapfrog(self.h, u, e, nsteps, dt))

    def update_electric_fields(self, dt=0.025, u=U_HOT, e=E_HOT):
        self._ck(self.L.tcym_update_efield(self.h, u, e, dt))

    def update_gauge_links(self, dt=0.025, u=U_HOT, e=E_HOT):
        self._ck(self.L.tcym_update_links(self.h, u, e, dt))

    def langevin_update(self, nsteps=1, dt=0.001, damping=1.0, T=0.5, seed=0, step0=0, u=U_HOT, e=E_HOT):
        """LangevinDynamics::Update (DYNAMICS/LangevinDynamics.cpp:320-347)."""
        self._ck(self.L.tcym_langevin(self.h, u, e, nsteps, dt, damping, T, seed, step0))

    def evolve_ncs(self, mode, nsteps, dt, damping=0.0, T=0.0, seed=0, step0=0, u=U_HOT, e=E_HOT):
        """nsteps leapfrog (mode 0) or Langevin (mode 1) steps; returns the hot-trajectory
        N_CS integral sum dt/2 [NCsDot(U_n, E_n+1/2) + NCsDot(U_n+1, E_n+1/2)]
        (ChernSimonsNumber::DeltaNCsRealTime, TOPOLOGY/ChernSimonsDerivative.cpp:8)."""
        d = ctypes.c_double(0.0)
        self._ck(self.L.tcym_evolve_ncs(self.h, mode, u, e, nsteps, dt, damping, T, seed, step0, ctypes.byref(d)))
        return d.value

    def langevin_update_fields(self, dt=0.001, damping=1.0, xi=E_XI, u=U_HOT, e=E_HOT):
        """LangevinDynamics::UpdateFields with LangevinVariables::Xi (DYNAMICS/LangevinDynamics.cpp:300-311)."""
        self._ck(self.L.tcym_langevin_xi(self.h, u, e, xi, dt, damping))

    def random_gaussian_noise(self, seed, step, stddev, xi=E_XI):
        self._ck(self.L.tcym_noise(self.h, xi, seed, step, stddev))

    def gradient_flow(self, nsteps=1, dtau=0.125, calibrate=False, u=U_COOL, e=E_COOL):
        """GradientFlow::StandardUpdate / CalibrationUpdate (GRADIENTFLOW/GradientFlow.cpp:150-188)."""
        d = ctypes.c_double(0.0)
        self._ck(self.L.tcym_flow(self.h, u, e, nsteps, dtau, 1 if calibrate else 0, ctypes.byref(d)))
        return d.value

    def cool_n_save(self, nsteps, dtau, u_hot=U_HOT, u_cold=U_NEW):
        """Cooling::CoolNSave (GRADIENTFLOW/Cooling.cpp:156-179)."""
        self._ck(self.L.tcym_cool(self.h, u_hot, u_cold, nsteps, dtau))

    def cooling_method_start(self, nsteps, dtau, u_hot=U_HOT):
        self._ck(self.L.tcym_cooling_start(self.h, u_hot, nsteps, dtau))

    def cooling_method_update(self, nsteps, dtau, measure=True, u_hot=U_HOT):
        """ChernSimonsNumber::CoolingMethod::Update (TOPOLOGY/CoolingMethod.cpp:143-165); returns GetDeltaNCs(1)."""
        d = ctypes.c_double(0.0)
        self._ck(self.L.tcym_cooling_update(self.h, u_hot, nsteps, dtau, 1 if measure else 0, ctypes.byref(d)))
        return d.value

    def block(self, u, e=-1):
        """Block(GaugeLinks**) / Block(ElectricFields**) (LATTICE/GaugeLinks.cpp:125-151)."""
        self._ck(self.L.tcym_block(self.h, u, e))

    def level(self, u=-1, e=-1):
        """(blocking level, extents) of link slot u or field slot e."""
        lv = ctypes.c_int()
        n = (ctypes.c_int * 3)()
        self._ck(self.L.tcym_level(self.h, u, e, ctypes.byref(lv), n))
        return lv.value, tuple(n)

    def cool_blocked(self, nsteps, dtau, block_frequency, max_blocking, u_hot=U_HOT, u_cold=U_NEW):
        """Cooling::CoolNSave with blocking (GRADIENTFLOW/Cooling.cpp:69-179); returns the blockings done."""
        nb = ctypes.c_int()
        self._ck(self.L.tcym_cool_blocked(self.h, u_hot, u_cold, nsteps, dtau, block_frequency, max_blocking,
                                          ctypes.byref(nb)))
        return nb.value

    def continue_cooling(self, step0, nsteps, dtau, ctau0=0.0, block_frequency=0, max_blocking=0, nblocked=0,
                         u_src=U_NEW, u_work=U_COOL, e_work=E_COOL):
        """Cooling::ContinueCooling in calibration mode (GRADIENTFLOW/Cooling.cpp:181-195);
        returns (DeltaNCsCooling, blockings, log rows [t, MAGNETIC, ELECTRIC, DeltaNCs, VacEst])."""
        nb = ctypes.c_int(nblocked)
        d = ctypes.c_double()
        log = np.empty((nsteps - step0 + 1, 5))
        self._ck(self.L.tcym_continue_cooling(self.h, u_src, u_work, e_work, step0, nsteps, dtau, ctau0,
                                              block_frequency, max_blocking, ctypes.byref(nb), ctypes.byref(d),
                                              log.ctypes.data_as(_dp)))
        return d.value, nb.value, log

    def thermalize(self, cycles, max_time, dt=0.0375, beta=2.0, seed=0, maxiter=10000, tol=1e-14, u=U_HOT, e=E_HOT):
        """ThermalDynamics::Thermalize (DYNAMICS/ThermalDynamics.cpp:260-316); returns the
        ThermalEnergy rows {time, T00, TXX, TYY, TZZ, ELECTRIC, MAGNETIC} as an (n, 7) array."""
        cap = cycles * (int(max_time / dt) // 20 + 2)
        log = np.zeros(7 * max(cap, 1))
        n = ctypes.c_int(0)
        self._ck(self.L.tcym_thermalize(self.h, u, e, cycles, max_time, dt, beta, seed, maxiter, tol,
                                        log.ctypes.data_as(_dp), cap, ctypes.byref(n)))
        assert n.value <= cap
        return log[:7 * n.value].reshape(n.value, 7)

    def quasi_particles(self, Qs=1.0, n0=1.0, seed=0, u=U_HOT, e=E_HOT):
        """InitialConditions::SetQuasiParticles before its Gauss-law restoration
        (INITIALCONDITIONS/SetQuasiParticles.cpp:24-174)."""
        self._ck(self.L.tcym_quasi_particles(self.h, u, e, Qs, n0, seed))

    def reunitarize(self, u=U_HOT):
        self._ck(self.L.tcym_reunitarize(self.h, u))

    # ---- observables ----
    def edotb(self, improved=True, u=U_HOT, e=E_HOT):
        d = ctypes.c_double()
        self._ck(self.L.tcym_edotb(self.h, u, e, 1 if improved else 0, ctypes.byref(d)))
        return d.value

    def ncs_dot(self, u=U_HOT, e=E_HOT):
        """ChernSimonsNumber::NCsDot (TOPOLOGY/ChernSimonsDerivative.cpp:31-49)."""
        d = ctypes.c_double()
        self._ck(self.L.tcym_ncsdot(self.h, u, e, ctypes.byref(d)))
        return d.value

    def ncs_dot_noalloc(self, out_dev, u=U_HOT, e=E_HOT):
        """tcym_ncsdot_noalloc: NCsDot written to the device address out_dev (an
        int), asynchronously on the context's stream; call sync() before reading it."""
        self._ck(self.L.tcym_ncsdot_noalloc(self.h, u, e, ctypes.c_void_p(out_dev)))

    def edotb_map(self, u=U_HOT, e=E_HOT):
        m = np.empty(self.V)
        self._ck(self.L.tcym_edotb_map(self.h, u, e, m.ctypes.data_as(_dp)))
        return m

    def bulk_update(self, u=U_HOT, e=E_HOT):
        """Observables::Bulk::Update: {ESqr[3], BSqr[3]} (OBSERVABLES/BulkObservables.cpp:133-143)."""
        o = np.empty(6)
        self._ck(self.L.tcym_bulk(self.h, u, e, o.ctypes.data_as(_dp)))
        return o

    def hard_scales(self, u=U_HOT):
        """Observables::HardScales::Update: H[mu] (OBSERVABLES/HardScales.cpp:40-104)."""
        o = np.empty(3)
        self._ck(self.L.tcym_hard_scales(self.h, u, o.ctypes.data_as(_dp)))
        return o

    def energy(self, u=U_HOT, e=E_HOT):
        """Bulk::T00 for a = aScale = 1 (OBSERVABLES/BulkObservables.cpp:29-31)."""
        return 0.5 * float(self.bulk_update(u, e).sum())

    def gauss_violation(self, u=U_HOT, e=E_HOT):
        m, r = ctypes.c_double(), ctypes.c_double()
        self._ck(self.L.tcym_gauss(self.h, u, e, ctypes.byref(m), ctypes.byref(r)))
        return m.value, r.value

    def unitarity_violation(self, u=U_HOT):
        m, r = ctypes.c_double(), ctypes.c_double()
        self._ck(self.L.tcym_unitarity(self.h, u, ctypes.byref(m), ctypes.byref(r)))
        return m.value, r.value

    def gauss_law_restore(self, maxiter=10000, tol=1e-14, u=U_HOT, e=E_HOT, check=True):
        """GaussLawRestoration::Restore (GAUSSLAWFIXING/GaussRestore.cpp:131-167)."""
        it, fm = ctypes.c_int(), ctypes.c_double()
        r = self.L.tcym_gauss_restore(self.h, u, e, maxiter, tol, ctypes.byref(it), ctypes.byref(fm))
        if r != OK and (check or r != E_NOCONV):
            self._ck(r)
        return it.value, fm.value

    def geodesic_interpolation(self, c, uold=U_OLD, unew=U_NEW, umid=U_MID, emid=E_MID):
        self._ck(self.L.tcym_geodesic(self.h, uold, unew, umid, emid, c))

    def delta_ncs_cooling(self, uold=U_OLD, unew=U_NEW):
        d = ctypes.c_double()
        self._ck(self.L.tcym_delta_ncs_cooling(self.h, uold, unew, ctypes.byref(d)))
        return d.value

    def vacuum_ncs(self, u=U_HOT):
        d = ctypes.c_double()
        self._ck(self.L.tcym_vacuum_ncs(self.h, u, ctypes.byref(d)))
        return d.value

    # ---- measurement ----
    # ---- configuration files (include/tcym_io.h) ----
    def _io(self, r):
        if r != OK:
            raise TcymError(r, self.L.tcym_io_last_error().decode())

    def save_configuration(self, ufile, efile, u=U_HOT, e=E_HOT, digits=15, nthreads=0):
        """IO::SaveConfiguration (IO/SaveConfiguration.cpp:3-52); either file may be None."""
        self._io(self.L.tcym_save_text(self.h, u, e, None if ufile is None else os.fsencode(ufile),
                                       None if efile is None else os.fsencode(efile), digits, nthreads))

    def load_configuration(self, ufile, efile, u=U_HOT, e=E_HOT, nthreads=0):
        """IO::LoadConfiguration (IO/LoadConfiguration.cpp:3-136); either file may be None."""
        self._io(self.L.tcym_load_text(self.h, u, e, None if ufile is None else os.fsencode(ufile),
                                       None if efile is None else os.fsencode(efile), nthreads))

    def save_checkpoint(self, path, u=U_HOT, e=E_HOT, step=0, seed=0, time=0.0, acc=()):
        ci = _info(step=step, seed=seed, time=time, acc=acc)
        self._io(self.L.tcym_save_checkpoint(self.h, u, e, os.fsencode(path), ctypes.byref(ci)))

    def load_checkpoint(self, path, u=U_HOT, e=E_HOT):
        """Loads a checkpoint into the slots; returns its header as a dict."""
        ci = CheckpointInfo()
        self._io(self.L.tcym_load_checkpoint(self.h, u, e, os.fsencode(path), ctypes.byref(ci)))
        return ci.as_dict()

    def params(self):
        p = Params()
        self._ck(self.L.tcym_get_params(self.h, ctypes.byref(p)))
        return {"nc": p.nc, "precision": p.precision, "N": tuple(p.N), "a": tuple(p.a), "aScale": p.aScale,
                "device": p.device}

    def set_update_form(self, form):
        """FORM_AUTO: plaquettes shared among their links where the lattice
        allows it and fills the device; FORM_LINK: the staples of every link
        apart; FORM_PLAQUETTE: plaquettes wherever the lattice allows it."""
        self._ck(self.L.tcym_set_update_form(self.h, form))

    def timing(self, enable=True):
        self._ck(self.L.tcym_timing(self.h, 1 if enable else 0))

    def timing_reset(self):
        self._ck(self.L.tcym_timing_reset(self.h))

    def kernel_time(self, name):
        t, n = ctypes.c_double(), ctypes.c_longlong()
        self._ck(self.L.tcym_kernel_time(self.h, name.encode(), ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def sync(self):
        self._ck(self.L.tcym_sync(self.h))

    def device_bytes(self):
        b = ctypes.c_longlong()
        self._ck(self.L.tcym_device_bytes(self.h, ctypes.byref(b)))
        return b.value


class Batch:
    """Many independent lattices of uneven sizes in one device allocation
    (include/tcym_batch.h): every pass runs over all of them in one persistent
    kernel, and every lattice computes bit-for-bit what a Solver of the same
    size, precision and seed computes.  The reference runs such ensembles as
    separate MPI ranks with seed + rank (SU2Solver.cpp:161-164)."""

    def __init__(self, nc, sizes, precision=FP32, a=(1.0, 1.0, 1.0), aScale=1.0, device=0):
        self.L = load_library()
        sizes = [(n, n, n) if isinstance(n, int) else tuple(n) for n in sizes]
        self.nc, self.sizes, self.precision = nc, sizes, precision
        self.count = len(sizes)
        self.su = 4 if nc == 2 else 18
        self.ne = 3 if nc == 2 else 8
        self.V = [n[0] * n[1] * n[2] for n in sizes]
        flat = (ctypes.c_int * (3 * self.count))(*[k for n in sizes for k in n])
        self._flat = flat
        p = BatchParams(nc, precision, self.count, ctypes.cast(flat, ctypes.POINTER(ctypes.c_int)),
                        (ctypes.c_double * 3)(*a), aScale, device)
        h = ctypes.c_void_p()
        r = self.L.tcym_batch_create(ctypes.byref(p), ctypes.byref(h))
        if r != OK:
            raise TcymError(r, "tcym_batch_create failed (group %d, %d lattices, device %d)" % (nc, self.count, device))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.tcym_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, r):
        if r != OK:
            raise TcymError(r, self.L.tcym_batch_last_error(self.h).decode())
        return r

    # ---- fields ----
    def set_links(self, e, U, slot=U_HOT):
        U = _arr(U, 3 * self.V[e] * self.su)
        self._ck(self.L.tcym_batch_set_links(self.h, slot, e, U.ctypes.data_as(_dp)))

    def get_links(self, e, slot=U_HOT):
        U = np.empty(3 * self.V[e] * self.su)
        self._ck(self.L.tcym_batch_get_links(self.h, slot, e, U.ctypes.data_as(_dp)))
        return U

    def set_efield(self, e, E, slot=E_HOT):
        E = _arr(E, 3 * self.V[e] * self.ne)
        self._ck(self.L.tcym_batch_set_efield(self.h, slot, e, E.ctypes.data_as(_dp)))

    def get_efield(self, e, slot=E_HOT):
        E = np.empty(3 * self.V[e] * self.ne)
        self._ck(self.L.tcym_batch_get_efield(self.h, slot, e, E.ctypes.data_as(_dp)))
        return E

    def initial_condition(self, seed0, sigma_u, sigma_e, maxiter=10000, tol=1e-14):
        """Lattice e: random links/E with seed0 + e, then GaussRestore (FP64)."""
        it = (ctypes.c_int * self.count)()
        self._ck(self.L.tcym_batch_initial_condition(self.h, seed0, sigma_u, sigma_e, maxiter, tol, it))
        return list(it)

    def set_step(self, step):
        self._ck(self.L.tcym_batch_set_step(self.h, step))

    def get_step(self):
        s = ctypes.c_uint32()
        self._ck(self.L.tcym_batch_get_step(self.h, ctypes.byref(s)))
        return s.value

    # ---- updates ----
    def leapfrog(self, nsteps, dt):
        self._ck(self.L.tcym_batch_leapfrog(self.h, nsteps, dt))

    def langevin(self, nsteps, dt, damping, T, seed0):
        self._ck(self.L.tcym_batch_langevin(self.h, nsteps, dt, damping, T, seed0))

    def flow(self, nsteps, dtau, u=U_HOT):
        self._ck(self.L.tcym_batch_flow(self.h, u, nsteps, dtau))

    def evolve(self, mode, nseg, steps, dt, damping=0.0, T=0.0, seed0=0):
        """Graph-replayed segments of `steps` updates + NCsDot; returns NCsDot [count, nseg]."""
        out = np.empty(self.count * nseg)
        self._ck(self.L.tcym_batch_evolve(self.h, mode, nseg, steps, dt, damping, T, seed0, out.ctypes.data_as(_dp)))
        return out.reshape(self.count, nseg)

    def cooling_start(self, nsteps, dtau):
        self._ck(self.L.tcym_batch_cooling_start(self.h, nsteps, dtau))

    def cooling_update(self, nsteps, dtau):
        out = np.empty(self.count)
        self._ck(self.L.tcym_batch_cooling_update(self.h, nsteps, dtau, out.ctypes.data_as(_dp)))
        return out

    # ---- measurements (one entry per lattice) ----
    def ncs_dot(self, u=U_HOT, e=E_HOT):
        out = np.empty(self.count)
        self._ck(self.L.tcym_batch_ncsdot(self.h, u, e, out.ctypes.data_as(_dp)))
        return out

    def bulk(self, u=U_HOT, e=E_HOT):
        out = np.empty(self.count * 6)
        self._ck(self.L.tcym_batch_bulk(self.h, u, e, out.ctypes.data_as(_dp)))
        return out.reshape(self.count, 6)

    def gauss(self, u=U_HOT, e=E_HOT):
        out = np.empty(self.count * 2)
        self._ck(self.L.tcym_batch_gauss(self.h, u, e, out.ctypes.data_as(_dp)))
        return out.reshape(self.count, 2)

    def unitarity(self, u=U_HOT):
        out = np.empty(self.count * 2)
        self._ck(self.L.tcym_batch_unitarity(self.h, u, out.ctypes.data_as(_dp)))
        return out.reshape(self.count, 2)

    def tune(self, mode=1, dt=0.05, reps=3):
        """Autotune the update kernels; returns ((shape, blocks per CU, XCD split), us per step)."""
        ch = (ctypes.c_int * 3)()
        us = ctypes.c_double()
        self._ck(self.L.tcym_batch_tune(self.h, mode, dt, reps, ch, ctypes.byref(us)))
        return tuple(ch), us.value

    def set_schedule(self, mode, shape, blocks_per_cu=4, xcd_split=True):
        """The schedule of one pass (mode 0/1: leapfrog and Langevin, 3: flow) set
        by hand: shape -1, 0, 4..6 (per-link form) or 7..9 (plaquette form, units
        of 4, 8, 16 planes)."""
        self._ck(self.L.tcym_batch_set_schedule(self.h, mode, shape, blocks_per_cu, 1 if xcd_split else 0))

    # ---- timing ----
    def timing(self, enable=True):
        self._ck(self.L.tcym_batch_timing(self.h, 1 if enable else 0))

    def timing_reset(self):
        self._ck(self.L.tcym_batch_timing_reset(self.h))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._ck(self.L.tcym_batch_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def sync(self):
        self._ck(self.L.tcym_batch_sync(self.h))
