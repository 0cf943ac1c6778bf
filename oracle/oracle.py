"""ctypes wrapper of the FP64 CPU oracle (oracle/sph_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product library never
uses it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

F_X, F_V, F_A, F_RHO, F_P, F_DRHO, F_RHO0, F_TYPE, F_ID = range(9)
_FIELD = {"x": F_X, "v": F_V, "a": F_A, "rho": F_RHO, "P": F_P, "Drho": F_DRHO,
          "rho0": F_RHO0, "type": F_TYPE, "id": F_ID}


class orc_params(ctypes.Structure):
    _fields_ = [("m", ctypes.c_double), ("rho0", ctypes.c_double), ("c", ctypes.c_double),
                ("alpha", ctypes.c_double), ("hs", ctypes.c_double), ("mu", ctypes.c_double),
                ("nu", ctypes.c_double), ("g", ctypes.c_double * 3), ("dt", ctypes.c_double),
                ("coeff", ctypes.c_double)]


_L = None
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int64)
_pp = ctypes.POINTER(orc_params)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_create": (ctypes.c_void_p, [ctypes.c_double, _pd, _pd]),
            "orc_create_periodic": (ctypes.c_void_p, [ctypes.c_double, _pd, _pd]),
            "orc_destroy": (None, [ctypes.c_void_p]),
            "orc_resize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
            "orc_n": (ctypes.c_int64, [ctypes.c_void_p]),
            "orc_dim": (ctypes.c_int, [ctypes.c_void_p]),
            "orc_field": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_int]),
            "orc_key_geometry": (None, [ctypes.c_void_p, _pi, _pi, _pi, _pi,
                                        ctypes.POINTER(ctypes.c_int)]),
            "orc_find_key": (ctypes.c_int64, [ctypes.c_void_p, _pd]),
            "orc_create_cell_list": (ctypes.c_int64, [ctypes.c_void_p]),
            "orc_neighbour_pairs": (ctypes.c_int64, [ctypes.c_void_p, _pi, ctypes.c_int64]),
            "orc_neighbour_counts": (ctypes.c_int64, [ctypes.c_void_p, _pi]),
            "orc_matrix_entries": (ctypes.c_int64, [ctypes.c_void_p, _pi, ctypes.c_int64]),
            "orc_probe_sum": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _pd,
                                             ctypes.c_int64, _pd]),
            "orc_brute_pairs": (ctypes.c_int64, [_pd, ctypes.c_int64, ctypes.c_double, _pi,
                                                 ctypes.c_int64]),
            "orc_density_sum": (ctypes.c_int64, [ctypes.c_void_p, _pp, ctypes.c_int]),
            "orc_density_sum_rho0": (None, [ctypes.c_void_p, _pp, ctypes.c_int]),
            "orc_internal_force": (None, [ctypes.c_void_p, _pp]),
            "orc_pressure_force_rho0": (None, [ctypes.c_void_p, _pp]),
            "orc_balance_of_mass": (ctypes.c_int64, [ctypes.c_void_p, _pp]),
            "orc_reset_rho": (None, [ctypes.c_void_p]),
            "orc_reset_a": (None, [ctypes.c_void_p]),
            "orc_eos": (None, [ctypes.c_void_p, _pp, ctypes.c_int]),
            "orc_eos_continuity": (None, [ctypes.c_void_p, _pp]),
            "orc_accelerate": (None, [ctypes.c_void_p, _pp]),
            "orc_move": (None, [ctypes.c_void_p, _pp]),
            "orc_drift": (None, [ctypes.c_void_p, _pp]),
            "orc_kick": (None, [ctypes.c_void_p, _pp]),
            "orc_step_collapse_dry": (ctypes.c_int64, [ctypes.c_void_p, _pp, ctypes.c_int]),
            "orc_neighbour_digest": (ctypes.c_int64, [ctypes.c_void_p, _pi,
                                                      ctypes.POINTER(ctypes.c_uint64)]),
            "orc_mix64": (ctypes.c_uint64, [ctypes.c_uint64]),
            "orc_particle_sum_lj": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double, _pd]),
            "orc_energy": (ctypes.c_double, [ctypes.c_void_p, _pp, ctypes.c_int]),
            "orc_momentum": (None, [ctypes.c_void_p, _pp, _pd]),
            "orc_step_wcsph": (ctypes.c_int64, [ctypes.c_void_p, _pp, ctypes.c_int, ctypes.c_int]),
            "orc_num_threads": (ctypes.c_int, []),
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_wendland2": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "orc_rDwendland2": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "orc_wendland3": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "orc_rDwendland3": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "orc_Dwendland2": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "orc_Dwendland3": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _L = L
    return _L


def params(**kw):
    p = orc_params()
    for k in ("m", "rho0", "c", "alpha", "hs", "mu", "nu", "dt"):
        setattr(p, k, float(kw.get(k, 0.0)))
    p.coeff = float(kw.get("coeff", 1.0))
    g = kw.get("g", (0.0, 0.0, 0.0))
    for i in range(3):
        p.g[i] = float(g[i])
    return p


class OracleSystem:
    """The reference's ParticleSystem restated in C (FP64, CPU).  With
    periodic=True, a system periodic along every axis of the box with extent
    (not in the reference)."""

    def __init__(self, h, lo, hi, periodic=False):
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        make = lib().orc_create_periodic if periodic else lib().orc_create
        self._s = make(float(h), lo.ctypes.data_as(_pd), hi.ctypes.data_as(_pd))
        self.periodic = periodic
        if not self._s:
            raise ValueError("invalid oracle system")
        self.h = h

    def __del__(self):
        if getattr(self, "_s", None):
            lib().orc_destroy(self._s)
            self._s = None

    @property
    def n(self):
        return lib().orc_n(self._s)

    @property
    def dim(self):
        return lib().orc_dim(self._s)

    def load(self, x, v=None, rho=None, P=None, type=None):
        x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
        n = len(x)
        lib().orc_resize(self._s, n)
        self.field("x")[:] = x
        if v is not None:
            self.field("v")[:] = v
        if rho is not None:
            self.field("rho")[:] = rho
        if P is not None:
            self.field("P")[:] = P
        if type is not None:
            self.field("type")[:] = type

    def field(self, name):
        """Writable numpy view of a field (reference order)."""
        n = self.n
        ptr = lib().orc_field(self._s, _FIELD[name])
        if name == "id":
            buf = (ctypes.c_int64 * max(n, 1)).from_address(ptr)
            return np.frombuffer(buf, dtype=np.int64, count=n)
        comps = 3 if name in ("x", "v", "a") else 1
        buf = (ctypes.c_double * max(n * comps, 1)).from_address(ptr)
        a = np.frombuffer(buf, dtype=np.float64, count=n * comps)
        return a.reshape(n, 3) if comps == 3 else a

    def key_geometry(self):
        ph = (ctypes.c_int64 * 3)()
        lim = (ctypes.c_int64 * 3)()
        km = ctypes.c_int64()
        diff = (ctypes.c_int64 * 27)()
        nd = ctypes.c_int()
        lib().orc_key_geometry(self._s, ph, lim, ctypes.byref(km), diff, ctypes.byref(nd))
        return tuple(ph), tuple(lim), km.value, list(diff)[: nd.value]

    def find_key(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return lib().orc_find_key(self._s, x.ctypes.data_as(_pd))

    def create_cell_list(self):
        return lib().orc_create_cell_list(self._s)

    def pairs(self):
        """Accepted pairs (p, q) in reference visiting order."""
        total = lib().orc_neighbour_pairs(self._s, None, 0)
        out = np.empty((max(total, 1), 2), dtype=np.int64)
        lib().orc_neighbour_pairs(self._s, out.ctypes.data_as(_pi), total)
        return out[:total]

    def matrix_entries(self):
        """The (p, q) visits of assemble_matrix, p == q included."""
        total = lib().orc_matrix_entries(self._s, None, 0)
        out = np.empty((max(total, 1), 2), dtype=np.int64)
        lib().orc_matrix_entries(self._s, out.ctypes.data_as(_pi), total)
        return out[:total]

    def neighbour_counts(self):
        c = np.empty(max(self.n, 1), dtype=np.int64)
        tot = lib().orc_neighbour_counts(self._s, c.ctypes.data_as(_pi))
        return c[: self.n], tot

    def probe_sum(self, x, m, fluid_only=False):
        """SmoothedParticles.sum at probe points x (np, 3): (np, 4) array of
        sum (phi m) W and sum (phi m) v_k W (src/core.jl:240-258)."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, 3))
        out = np.zeros((x.shape[0], 4), dtype=np.float64)
        rc = lib().orc_probe_sum(self._s, float(m), 1 if fluid_only else 0,
                                 x.ctypes.data_as(_pd), x.shape[0], out.ctypes.data_as(_pd))
        if rc != 0:
            raise RuntimeError("orc_probe_sum: no current cell list or periodic system")
        return out

    def particle_sum_lj(self, m, dr_wall, E_wall, eps=1e-16):
        """SmoothedParticles.sum(sys, LJ_potential, p) for every particle p
        (src/core.jl:271-291, examples/collapse_symplectic.jl:144-151)."""
        out = np.zeros(max(self.n, 1))
        rc = lib().orc_particle_sum_lj(self._s, float(m), float(dr_wall), float(E_wall),
                                       float(eps), out.ctypes.data_as(_pd))
        if rc != 0:
            raise RuntimeError("orc_particle_sum_lj: no current cell list or periodic system")
        return out[: self.n]

    def density_sum(self, self_term=True, **kw):
        p = params(**kw)
        return lib().orc_density_sum(self._s, ctypes.byref(p), 1 if self_term else 0)

    def density_sum_rho0(self, self_term=True, **kw):
        p = params(**kw)
        lib().orc_density_sum_rho0(self._s, ctypes.byref(p), 1 if self_term else 0)

    def internal_force(self, **kw):
        p = params(**kw)
        lib().orc_internal_force(self._s, ctypes.byref(p))

    def pressure_force_rho0(self, **kw):
        p = params(**kw)
        lib().orc_pressure_force_rho0(self._s, ctypes.byref(p))

    def balance_of_mass(self, **kw):
        p = params(**kw)
        return lib().orc_balance_of_mass(self._s, ctypes.byref(p))

    def reset_rho(self):
        lib().orc_reset_rho(self._s)

    def reset_a(self):
        lib().orc_reset_a(self._s)

    def eos(self, use_field=False, **kw):
        p = params(**kw)
        lib().orc_eos(self._s, ctypes.byref(p), 1 if use_field else 0)

    def eos_continuity(self, **kw):
        p = params(**kw)
        lib().orc_eos_continuity(self._s, ctypes.byref(p))

    def accelerate(self, **kw):
        p = params(**kw)
        lib().orc_accelerate(self._s, ctypes.byref(p))

    def move(self, **kw):
        p = params(**kw)
        lib().orc_move(self._s, ctypes.byref(p))

    def drift(self, **kw):
        p = params(**kw)
        lib().orc_drift(self._s, ctypes.byref(p))

    def kick(self, **kw):
        p = params(**kw)
        lib().orc_kick(self._s, ctypes.byref(p))

    def step_collapse_dry(self, nsteps, **kw):
        """nsteps iterations of the time loop of examples/collapse_dry.jl:202-211;
        returns the balance_of_mass! pairs summed over them."""
        p = params(**kw)
        return lib().orc_step_collapse_dry(self._s, ctypes.byref(p), int(nsteps))

    def neighbour_digest(self):
        """(counts, hashes): per particle the number of neighbours and the sum
        modulo 2^64 of mix64(q + 1) over them."""
        n = self.n
        c = np.empty(max(n, 1), dtype=np.int64)
        hsh = np.empty(max(n, 1), dtype=np.uint64)
        tot = lib().orc_neighbour_digest(self._s, c.ctypes.data_as(_pi),
                                         hsh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if tot < 0:
            raise RuntimeError("no current cell list")
        return c[:n], hsh[:n]

    def energy(self, kind="wcsph", **kw):
        p = params(**kw)
        k = {"wcsph": 0, "collision": 1, "wcsph_rho0": 2}[kind]
        return lib().orc_energy(self._s, ctypes.byref(p), k)

    def momentum(self, **kw):
        p = params(**kw)
        out = (ctypes.c_double * 3)()
        lib().orc_momentum(self._s, ctypes.byref(p), out)
        return np.array(list(out))

    def step_wcsph(self, nsteps, use_rho0=False, **kw):
        p = params(**kw)
        return lib().orc_step_wcsph(self._s, ctypes.byref(p), int(nsteps), 1 if use_rho0 else 0)


def brute_pairs(x, h):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    total = lib().orc_brute_pairs(x.ctypes.data_as(_pd), len(x), float(h), None, 0)
    out = np.empty((max(total, 1), 2), dtype=np.int64)
    lib().orc_brute_pairs(x.ctypes.data_as(_pd), len(x), float(h), out.ctypes.data_as(_pi), total)
    return out[:total]


def mix64(z):
    return lib().orc_mix64(int(z))


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()
