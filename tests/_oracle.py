"""ctypes binding of the oracle (CPU FP64 restatement in oracle/).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it, always as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "libtcym_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


class OrParams(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int), ("N", ctypes.c_int * 3), ("a", ctypes.c_double * 3),
                ("aScale", ctypes.c_double), ("nthreads", ctypes.c_int)]


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        L = ctypes.CDLL(ORACLE_LIB)
        P = ctypes.POINTER(OrParams)
        d, i, u32, u64 = ctypes.c_double, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64
        sig = {
            "or_mul": (None, [i, _dp, _dp, _dp]), "or_mul_nd": (None, [i, _dp, _dp, _dp]),
            "or_mul_dn": (None, [i, _dp, _dp, _dp]), "or_proj": (None, [i, d, _dp, _dp]),
            "or_expi": (None, [i, d, _dp, _dp]), "or_logi": (None, [i, d, _dp, _dp]),
            "or_adj_act": (None, [i, _dp, _dp, _dp]), "or_inv_adj_act": (None, [i, _dp, _dp, _dp]),
            "or_unitarity_norm": (d, [i, _dp]), "or_reunitarize": (None, [i, _dp]),
            "or_philox4x32": (None, [ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
            "or_link_normals": (None, [u64, u32, u32, u32, i, _dp]),
            "or_random_links": (None, [P, u64, d, _dp]), "or_random_efield": (None, [P, u64, d, _dp]),
            "or_noise_field": (None, [P, u64, u32, d, _dp]),
            "or_leapfrog": (None, [P, _dp, _dp, i, d]),
            "or_update_efield": (None, [P, _dp, _dp, d]), "or_update_links": (None, [P, _dp, _dp, d]),
            "or_langevin_xi": (None, [P, _dp, _dp, _dp, d, d]),
            "or_langevin": (None, [P, _dp, _dp, i, d, d, d, u64, u32]),
            "or_flow_force": (None, [P, _dp, _dp]), "or_flow_links": (None, [P, _dp, _dp, d]),
            "or_flow": (None, [P, _dp, _dp, i, d]), "or_flow_calibration": (d, [P, _dp, _dp, i, d]),
            "or_edotb": (d, [P, _dp, _dp]), "or_edotb_unimproved": (d, [P, _dp, _dp]),
            "or_edotb_map": (None, [P, _dp, _dp, _dp]), "or_ncsdot": (d, [P, _dp, _dp]),
            "or_bulk": (None, [P, _dp, _dp, _dp]),
            "or_hard_scales": (None, [P, _dp, _dp]),
            "or_block_links": (None, [P, _dp, _dp]),
            "or_gauss": (None, [P, _dp, _dp, _dp, _dp]), "or_gauss_field": (None, [P, _dp, _dp, _dp]),
            "or_unitarity": (None, [P, _dp, _dp, _dp]),
            "or_gauss_restore": (i, [P, _dp, _dp, i, d, _dp]),
            "or_geodesic": (None, [P, _dp, _dp, d, _dp, _dp]),
            "or_delta_ncs_cooling": (d, [P, _dp, _dp, _dp, _dp]),
            "or_vacuum_ncs": (d, [P, _dp]), "or_reunitarize_field": (None, [P, _dp]),
            "or_thermal_efield": (None, [P, u64, u32, d, _dp]),
            "or_thermalize": (i, [P, _dp, _dp, i, d, d, d, u64, i, d, _dp, i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def su(nc):
    return 4 if nc == 2 else 18


def ne(nc):
    return 3 if nc == 2 else 8


class Lattice:
    """One lattice of the oracle: U, E in the reference's AoS layout."""

    def __init__(self, nc, N, a=(1.0, 1.0, 1.0), aScale=1.0, nthreads=0):
        if isinstance(N, int):
            N = (N, N, N)
        self.nc, self.N = nc, tuple(N)
        self.V = N[0] * N[1] * N[2]
        self.p = OrParams(nc, (ctypes.c_int * 3)(*N), (ctypes.c_double * 3)(*a), aScale, nthreads)
        self.U = np.zeros(3 * self.V * su(nc))
        self.E = np.zeros(3 * self.V * ne(nc))
        self.set_identity()

    # ---- set-up ----
    def set_identity(self):
        u = self.U.reshape(-1, su(self.nc))
        u[:] = 0
        if self.nc == 2:
            u[:, 3] = 1
        else:
            u[:, 0] = u[:, 8] = u[:, 16] = 1

    def random(self, seed, sigma_u, sigma_e, restore=True, maxiter=10000, tol=1e-14):
        L = lib()
        L.or_random_links(ctypes.byref(self.p), seed, sigma_u, _ptr(self.U))
        L.or_random_efield(ctypes.byref(self.p), seed, sigma_e, _ptr(self.E))
        if restore:
            return self.gauss_restore(maxiter, tol)
        return None

    def copy(self):
        c = Lattice.__new__(Lattice)
        c.nc, c.N, c.V, c.p = self.nc, self.N, self.V, self.p
        c.U, c.E = self.U.copy(), self.E.copy()
        return c

    # ---- sweeps ----
    def leapfrog(self, n, dt):
        lib().or_leapfrog(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), n, dt)

    def update_efield(self, dt):
        """Dynamics::UpdateElectricFields alone."""
        lib().or_update_efield(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), dt)

    def update_links(self, dt):
        """Dynamics::UpdateGaugeLinks alone."""
        lib().or_update_links(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), dt)

    def langevin(self, n, dt, damping, T, seed, step0=0):
        lib().or_langevin(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), n, dt, damping, T, seed, step0)

    def langevin_xi(self, Xi, dt, damping):
        lib().or_langevin_xi(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), _ptr(Xi), dt, damping)

    def noise(self, seed, step, stddev):
        Xi = np.zeros_like(self.E)
        lib().or_noise_field(ctypes.byref(self.p), seed, step, stddev, _ptr(Xi))
        return Xi

    def thermalize(self, cycles, max_time, dt, beta, seed, maxiter=10000, tol=1e-14):
        """ThermalDynamics::Thermalize: returns the ThermalEnergy rows (n x 7)."""
        cap = cycles * (int(max_time / dt) // 20 + 2)
        log = np.zeros(7 * max(cap, 1))
        n = lib().or_thermalize(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), cycles, max_time, dt, beta, seed,
                                maxiter, tol, _ptr(log), cap)
        assert n <= cap
        return log[:7 * n].reshape(n, 7)

    def flow(self, n, dtau):
        lib().or_flow(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), n, dtau)

    def flow_calibration(self, n, dtau):
        return lib().or_flow_calibration(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), n, dtau)

    def flow_force(self):
        lib().or_flow_force(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E))

    # ---- observables ----
    def edotb(self, U=None, E=None):
        U = self.U if U is None else U
        E = self.E if E is None else E
        return lib().or_edotb(ctypes.byref(self.p), _ptr(U), _ptr(E))

    def edotb_unimproved(self):
        return lib().or_edotb_unimproved(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E))

    def edotb_map(self):
        m = np.zeros(self.V)
        lib().or_edotb_map(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), _ptr(m))
        return m

    def ncsdot(self, U=None, E=None):
        U = self.U if U is None else U
        E = self.E if E is None else E
        return lib().or_ncsdot(ctypes.byref(self.p), _ptr(U), _ptr(E))

    def bulk(self):
        o = np.zeros(6)
        lib().or_bulk(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), _ptr(o))
        return o

    def blocked(self):
        """Block(GaugeLinks**): a new Lattice of half the extent and twice the spacing
        with the blocked links (E zero, as Block(ElectricFields**))."""
        a = tuple(2.0 * x for x in self.p.a)
        c = Lattice(self.nc, tuple(n // 2 for n in self.N), a=a, aScale=self.p.aScale, nthreads=self.p.nthreads)
        lib().or_block_links(ctypes.byref(self.p), _ptr(self.U), _ptr(c.U))
        return c

    def hard_scales(self):
        o = np.zeros(3)
        lib().or_hard_scales(ctypes.byref(self.p), _ptr(self.U), _ptr(o))
        return o

    def energy(self):
        """T00 = (ESqr + BSqr summed over directions)/2 (a = aScale = 1)."""
        return 0.5 * float(self.bulk().sum())

    def gauss(self):
        m, r = ctypes.c_double(), ctypes.c_double()
        lib().or_gauss(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), ctypes.byref(m), ctypes.byref(r))
        return m.value, r.value

    def gauss_field(self):
        G = np.zeros(self.V * ne(self.nc))
        lib().or_gauss_field(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), _ptr(G))
        return G

    def unitarity(self, U=None):
        U = self.U if U is None else U
        m, r = ctypes.c_double(), ctypes.c_double()
        lib().or_unitarity(ctypes.byref(self.p), _ptr(U), ctypes.byref(m), ctypes.byref(r))
        return m.value, r.value

    def gauss_restore(self, maxiter=10000, tol=1e-14):
        fm = ctypes.c_double()
        it = lib().or_gauss_restore(ctypes.byref(self.p), _ptr(self.U), _ptr(self.E), maxiter, tol, ctypes.byref(fm))
        return it, fm.value

    def geodesic(self, Uold, Unew, c):
        Um, Em = np.zeros_like(Uold), np.zeros_like(self.E)
        lib().or_geodesic(ctypes.byref(self.p), _ptr(Uold), _ptr(Unew), c, _ptr(Um), _ptr(Em))
        return Um, Em

    def delta_ncs_cooling(self, Uold, Unew):
        Um, Em = np.zeros_like(Uold), np.zeros_like(self.E)
        return lib().or_delta_ncs_cooling(ctypes.byref(self.p), _ptr(Uold), _ptr(Unew), _ptr(Um), _ptr(Em))

    def vacuum_ncs(self, U=None):
        U = self.U if U is None else U
        return lib().or_vacuum_ncs(ctypes.byref(self.p), _ptr(U))

    def reunitarize(self):
        lib().or_reunitarize_field(ctypes.byref(self.p), _ptr(self.U))


# ---- matrix-level helpers ----
def expi(nc, c, alpha):
    out = np.zeros(su(nc))
    lib().or_expi(nc, c, _ptr(np.ascontiguousarray(alpha, dtype=np.float64)), _ptr(out))
    return out


def logi(nc, c, U):
    out = np.zeros(ne(nc))
    lib().or_logi(nc, c, _ptr(np.ascontiguousarray(U, dtype=np.float64)), _ptr(out))
    return out


def mul(nc, A, B, kind=""):
    out = np.zeros(su(nc))
    fn = {"": lib().or_mul, "nd": lib().or_mul_nd, "dn": lib().or_mul_dn}[kind]
    fn(nc, _ptr(np.ascontiguousarray(A)), _ptr(np.ascontiguousarray(B)), _ptr(out))
    return out


def proj(nc, c, M):
    out = np.zeros(ne(nc))
    lib().or_proj(nc, c, _ptr(np.ascontiguousarray(M)), _ptr(out))
    return out


def adj_act(nc, V, E, inverse=False):
    out = np.zeros(ne(nc))
    fn = lib().or_inv_adj_act if inverse else lib().or_adj_act
    fn(nc, _ptr(np.ascontiguousarray(V)), _ptr(np.ascontiguousarray(E, dtype=np.float64)), _ptr(out))
    return out


def unitarity_norm(nc, U):
    return lib().or_unitarity_norm(nc, _ptr(np.ascontiguousarray(U)))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32(c, k, o)
    return list(o)


def link_normals(seed, stream, step, link, n):
    z = np.zeros(n)
    lib().or_link_normals(seed, stream, step, link, n, _ptr(z))
    return z


# ---- numpy matrix conversions (independent representation for checks) ----
PAULI = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], complex)]


def gellmann():
    l = np.zeros((8, 3, 3), complex)
    l[0][0, 1] = l[0][1, 0] = 1
    l[1][0, 1], l[1][1, 0] = -1j, 1j
    l[2][0, 0], l[2][1, 1] = 1, -1
    l[3][0, 2] = l[3][2, 0] = 1
    l[4][0, 2], l[4][2, 0] = -1j, 1j
    l[5][1, 2] = l[5][2, 1] = 1
    l[6][1, 2], l[6][2, 1] = -1j, 1j
    l[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return l


def generators(nc):
    return [s / 2 for s in PAULI] if nc == 2 else list(gellmann() / 2)


def to_matrix(nc, u):
    """link storage -> complex matrix"""
    u = np.asarray(u)
    if nc == 2:
        return u[3] * np.eye(2) + 1j * sum(u[a] * PAULI[a] for a in range(3))
    z = u[0::2] + 1j * u[1::2]
    return z.reshape(3, 3, order="F")


def from_matrix(nc, M):
    if nc == 2:
        u3 = np.real(np.trace(M)) / 2
        ua = [np.real(np.trace(M @ s) / (2j)) for s in PAULI]
        return np.array(ua + [u3])
    z = M.reshape(-1, order="F")
    out = np.zeros(18)
    out[0::2], out[1::2] = z.real, z.imag
    return out


def alg_matrix(nc, alpha):
    return sum(alpha[a] * t for a, t in enumerate(generators(nc)))


def quasi_particles(lat, Qs, n0, seed):
    """InitialConditions::SetQuasiParticles before its Gauss-law restoration
    (INITIALCONDITIONS/SetQuasiParticles.cpp:24-174, Minkowski metric), in
    numpy: lattice momenta (LATTICE/Momenta.cpp:4-8), the polarisation vectors
    of MISC/MinkowskiWaveVectors.cpp:6-58, the occupancy n0/sqrt(p^2/Qs^2 + 0.1)
    below Qs (SetQuasiParticles.cpp:27-39), complex normals from Philox stream 4
    (counter colour + NE p; the reference draws RandomNumberGenerator::ComplexGauss),
    the unnormalised inverse FFT (FFTW_BACKWARD; numpy's ifftn times V), and
    U = exp(-i A^a t^a), E from twice the real parts (SetQuasiParticles.cpp:119-165).
    Sets lat.U and lat.E and returns (Xi1, Xi2, cDp) for checks."""
    N = lat.N
    V, nE = lat.V, ne(lat.nc)
    a = np.array(list(lat.p.a))
    ac = float(np.prod(a))
    kz, ky, kx = np.meshgrid(np.arange(N[2]), np.arange(N[1]), np.arange(N[0]), indexing="ij")
    k = (kx, ky, kz)
    cDp = np.array([-1j * (1.0 - np.exp(-1j * 2.0 * np.pi * k[m] / N[m])) / a[m] for m in range(3)])
    psq = (np.abs(cDp) ** 2).sum(axis=0)
    om = np.sqrt(psq)
    bv = np.array([1.0 / np.sqrt(3.0), 1.0 / np.sqrt(2.0), 1.0 / (np.sqrt(2.0) * np.sqrt(3.0))])
    nz = psq > 0
    safe = np.where(nz, psq, 1.0)
    bdp = sum(bv[m] * cDp[m] for m in range(3))
    nrm = np.sqrt(np.abs((bv ** 2).sum() - (bdp * np.conj(bdp)).real / safe))
    den = nrm * np.sqrt(2.0 * np.where(nz, om, 1.0))
    Xi1 = np.array([np.where(nz, (bv[m] - bdp * np.conj(cDp[m]) / safe) / den, 0.0) for m in range(3)])
    Xi2 = np.array([np.where(nz, (cDp[(m + 1) % 3] * np.conj(Xi1[(m + 2) % 3])
                                  - cDp[(m + 2) % 3] * np.conj(Xi1[(m + 1) % 3])) / np.where(nz, om, 1.0), 0.0)
                    for m in range(3)])
    pv2 = lat.p.aScale ** 2 * psq
    f = np.where(pv2 < Qs * Qs, n0 / np.sqrt(pv2 / (Qs * Qs) + 0.1), 0.0)
    amp = np.sqrt(f) * np.sqrt(V * ac)
    L = lib()
    z = np.zeros(4)
    c1 = np.zeros((V, nE), complex)
    c2 = np.zeros((V, nE), complex)
    for p in range(V):
        for c in range(nE):
            L.or_link_normals(seed, 4, 0, c + nE * p, 4, _ptr(z))
            c1[p, c] = complex(z[0], z[1]) / np.sqrt(2.0)
            c2[p, c] = complex(z[2], z[3]) / np.sqrt(2.0)
    shp = (N[2], N[1], N[0])
    lat.E[:] = 0.0
    U = lat.U.reshape(V, 3, -1)
    E = lat.E.reshape(V, 3, nE)
    for m in range(3):
        x1, x2 = Xi1[m].reshape(-1), Xi2[m].reshape(-1)
        Am = amp.reshape(-1)[:, None] * (c1 * x1[:, None] + c2 * x2[:, None])
        Em = 1j * om.reshape(-1)[:, None] * Am
        XA = np.fft.ifftn(Am.T.reshape((nE,) + shp), axes=(1, 2, 3)).reshape(nE, V).T * V
        XE = np.fft.ifftn(Em.T.reshape((nE,) + shp), axes=(1, 2, 3)).reshape(nE, V).T * V
        Al = 2.0 * a[m] * XA.real / (V * ac)
        E[:, m, :] = 2.0 * (ac / a[m]) * XE.real / (V * ac)
        for site in range(V):
            L.or_expi(lat.nc, -1.0, _ptr(np.ascontiguousarray(Al[site])), _ptr(U[site, m]))
    return Xi1, Xi2, cDp