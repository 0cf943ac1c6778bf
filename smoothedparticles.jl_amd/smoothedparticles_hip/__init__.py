"""smoothedparticles_hip -- host-side mirror of SmoothedParticles.jl's
hot-path API over the gfx950 library libsph_hip (C ABI, include/sph.h).

Reference call surface and its counterpart here:

    ParticleSystem(T, domain, h)       ParticleSystem(domain, h, precision)
    generate_particles!/push!          sys.add_particles(x, v, rho, P, type)
    ParticleField(sys, :rho)           ParticleField(sys, "rho")
    create_cell_list!(sys)             create_cell_list(sys)
    apply!(sys, op!; self, params...)  apply(sys, op, self=..., **params)
    apply_binary!(sys, op!)            apply_binary(sys, op, **params)
    apply_unary!(sys, op!)             apply_unary(sys, op, **params)
    sum(p -> energy(p), sys.particles) energy(sys, **params)

``Batch`` (include/sph_batch.h, no counterpart in the reference) holds many
independent systems, periodic ones included, and steps them together.

Operators cannot be arbitrary closures across the C ABI; the built-in ones
are registered by the reference name they replace, per reference program
(``PROGRAMS``; ``OPERATORS`` holds the names that are unambiguous).  A Python
callable passed to apply/apply_binary/apply_unary runs through the host
fallback (``host_apply``): the device finds the pairs, the closure runs on
the host over downloaded fields.  Every registered operator runs in the HIP
library; if the library is missing, every call raises.
"""
import ctypes

import numpy as np

from . import _native as N
from ._native import SPHError, lib, check
from .kernels import wendland2, rDwendland2, wendland3, rDwendland3  # noqa: F401

__all__ = [
    "Box", "Rectangle", "ParticleSystem", "ParticleField", "create_cell_list",
    "apply", "apply_binary", "apply_unary", "energy", "momentum", "neighbour_pairs",
    "step", "OPERATORS", "SPHError", "find_rho", "find_rho0", "internal_force",
    "internal_force_rho0", "balance_of_mass", "reset_rho", "reset_a", "find_pressure",
    "find_pressure_rho0", "find_pressure_continuity", "accelerate", "move",
    "wendland2", "rDwendland2", "wendland3", "rDwendland3", "FLUID", "WALL", "Batch",
    "probe_sums", "assemble_vector", "assemble_matrix", "drift", "kick", "PROGRAMS",
    "program_operators", "host_apply", "Particle", "neighbour_digest", "cell_keys",
    "visit_order", "particle_sums_lj_wall", "collapse_dry_init", "Snapshot",
]

FLUID = 0.0
WALL = 1.0

_dptr = ctypes.POINTER(ctypes.c_double)


def _as_ptr(a):
    return a.ctypes.data_as(_dptr)


class Box:
    """Box(x1_min, x2_min, x3_min, x1_max, x2_max, x3_max), src/geometry.jl:15-22."""

    def __init__(self, x1_min, x2_min, x3_min, x1_max, x2_max, x3_max):
        self.lo = np.array([x1_min, x2_min, x3_min], dtype=np.float64)
        self.hi = np.array([x1_max, x2_max, x3_max], dtype=np.float64)


def Rectangle(x1_min, x2_min, x1_max, x2_max):
    """Rectangle = Box with x3 in [0, 0] (src/geometry.jl:41-43)."""
    return Box(x1_min, x2_min, 0.0, x1_max, x2_max, 0.0)


class Operator:
    """A registered operator: its library id, whether it is binary, and the
    parameters the reference function fixes in its body (for example the
    0.5*dt half step of collapse_dry.jl's move!), which override the call's
    keyword parameters."""

    def __init__(self, name, op_id, binary, fixed=None):
        self.name = name
        self.id = op_id
        self.binary = binary
        self.fixed = dict(fixed or {})

    def __repr__(self):
        return f"<operator {self.name}>"


find_rho = Operator("find_rho!", N.SPH_OP_DENSITY_SUM, True)
find_rho0 = Operator("find_rho0!", N.SPH_OP_DENSITY_SUM_RHO0, True)
internal_force = Operator("internal_force!", N.SPH_OP_INTERNAL_FORCE, True)
internal_force_rho0 = Operator("internal_force! (rho0)", N.SPH_OP_PRESSURE_FORCE_RHO0, True)
balance_of_mass = Operator("balance_of_mass!", N.SPH_OP_BALANCE_OF_MASS, True)
reset_rho = Operator("reset_rho!", N.SPH_OP_RESET_RHO, False)
reset_a = Operator("reset_a!", N.SPH_OP_RESET_A, False)
find_pressure = Operator("find_pressure!", N.SPH_OP_FIND_PRESSURE, False)
find_pressure_rho0 = Operator("find_pressure! (rho0 field)", N.SPH_OP_FIND_PRESSURE_RHO0, False)
find_pressure_continuity = Operator(
    "find_pressure! (continuity)", N.SPH_OP_FIND_PRESSURE_CONTINUITY, False)
accelerate = Operator("accelerate!", N.SPH_OP_ACCELERATE, False)
move = Operator("move!", N.SPH_OP_MOVE, False)
drift = Operator("move! (all particles, a kept)", N.SPH_OP_DRIFT, False)
kick = Operator("accelerate! (all particles, no gravity)", N.SPH_OP_KICK, False)

# The operators of each reference program, by the name the program gives
# them.  Several names (find_pressure!, internal_force!, move!,
# accelerate!) have different bodies in different programs, so a bare name
# is resolved only where it is unambiguous; otherwise the program is named
# ("collapse_dry:move!") or the table below is used directly.
PROGRAMS = {
    "collapse_dry": {  # examples/collapse_dry.jl
        "balance_of_mass!": balance_of_mass,                         # :112-115
        "find_pressure!": find_pressure_continuity,                  # :123-127
        "internal_force!": internal_force,                           # :135-141
        "move!": Operator("move!", N.SPH_OP_MOVE, False, {"coeff": 0.5}),           # :148-153
        "accelerate!": Operator("accelerate!", N.SPH_OP_ACCELERATE, False, {"coeff": 0.5}),  # :155-159
    },
    "test_collision_2d": {  # tests/test_collision_2d.jl
        "find_rho!": find_rho,                                       # :63-65
        "find_rho0!": find_rho0,                                     # :67-69
        "find_pressure!": find_pressure_rho0,                        # :71-73
        "internal_force!": internal_force_rho0,                      # :75-78
        "reset_a!": reset_a,                                         # :80-82
        "reset_rho!": reset_rho,                                     # :84-86
        "move!": Operator("move!", N.SPH_OP_DRIFT, False, {"coeff": 1.0}),          # :88-90
        "accelerate!": Operator("accelerate!", N.SPH_OP_KICK, False, {"coeff": 0.5}),  # :92-94
    },
}


def _bare_names():
    """Names that mean the same operator in every program that defines
    them, plus the library's own unambiguous names."""
    by_name = {}
    for ops in PROGRAMS.values():
        for name, op in ops.items():
            by_name.setdefault(name, set()).add((op.id, tuple(sorted(op.fixed.items()))))
    table = {name: next(op for ops in PROGRAMS.values() for n, op in ops.items() if n == name)
             for name, ids in by_name.items() if len(ids) == 1}
    table.update({
        "internal_force_rho0!": internal_force_rho0,
        "find_pressure_rho0!": find_pressure_rho0,
        "find_pressure_continuity!": find_pressure_continuity,
        "find_pressure_constant!": find_pressure,   # P = c^2 (rho - rho0), collapse_dry.jl:126
    })
    return table


# Unambiguous name -> registered operator.
OPERATORS = _bare_names()
AMBIGUOUS = sorted({n for ops in PROGRAMS.values() for n in ops} - set(OPERATORS))


def program_operators(program):
    """The operators of one reference program, by its own names."""
    if program not in PROGRAMS:
        raise KeyError(f"unknown program {program!r}; known: {sorted(PROGRAMS)}")
    return dict(PROGRAMS[program])


_PARAM_NAMES = ("m", "rho0", "c", "alpha", "hs", "mu", "nu", "g", "dt", "coeff")


def make_params(**kw):
    """Build an sph_params struct from keyword parameters (missing = 0,
    coeff defaults to 1)."""
    unknown = set(kw) - set(_PARAM_NAMES)
    if unknown:
        raise TypeError(f"unknown parameters: {sorted(unknown)}")
    p = N.sph_params()
    for k in ("m", "rho0", "c", "alpha", "hs", "mu", "nu", "dt"):
        setattr(p, k, float(kw.get(k, 0.0)))
    p.coeff = float(kw.get("coeff", 1.0))
    g = kw.get("g", (0.0, 0.0, 0.0))
    for i in range(3):
        p.g[i] = float(g[i])
    return p


class ParticleSystem:
    """ParticleSystem(domain, h, precision) -- src/structs.jl:43-92.

    ``domain`` is a Box (only its bounding box matters, as in the
    reference); ``h`` is the neighbour radius and cell size; ``precision``
    is "fp64" (plain FP64, the drop-in) or "mixed" (FP32 pair arithmetic on
    relative coordinates with FP64 accumulation, FP64 state).
    """

    def __init__(self, domain, h, precision="fp64"):
        if not (h > 0.0):
            raise AssertionError("invalid ParticleSystem declaration! (h must be a positive float)")
        mode = {"fp64": N.SPH_FP64, "mixed": N.SPH_MIXED}[precision]
        self.h = float(h)
        self.domain = domain
        self.precision = precision
        self._h = ctypes.c_void_p()
        check(lib().sph_create(ctypes.byref(self._h), self.h, _as_ptr(domain.lo),
                               _as_ptr(domain.hi), mode))

    def close(self):
        if self._h:
            lib().sph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        n = ctypes.c_int64()
        check(lib().sph_count(self._h, ctypes.byref(n)))
        return n.value

    @property
    def n(self):
        return len(self)

    def add_particles(self, x, v=None, rho=None, P=None, type=None):
        """push! n particles (generate_particles!, src/grids.jl:253-258)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]

        def opt(a, shape):
            if a is None:
                return None
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape))
            return a

        v = opt(v, (n, 3))
        rho = opt(rho, (n,))
        P = opt(P, (n,))
        t = opt(type, (n,))
        keep = [x, v, rho, P, t]  # noqa: F841 (lifetime during the call)
        check(lib().sph_append(self._h, n, _as_ptr(x),
                               _as_ptr(v) if v is not None else None,
                               _as_ptr(rho) if rho is not None else None,
                               _as_ptr(P) if P is not None else None,
                               _as_ptr(t) if t is not None else None))

    def snapshot(self, *names):
        """Start an asynchronous download of the named fields (Snapshot)."""
        return Snapshot(lambda ids, k, out: lib().sph_snapshot_begin(self._h, ids, k, out), names)

    def generate_lattice(self, spec):
        """Device-side generate_particles! of a lattice (include/sph_setup.h,
        sph_generate_lattice); returns the number of particles added."""
        n = ctypes.c_int64()
        check(lib().sph_generate_lattice(self._h, ctypes.byref(spec), ctypes.byref(n)))
        return n.value

    def get(self, name):
        n = len(self)
        shape = (n, 3) if name in N.VECTOR_FIELDS else (n,)
        out = np.empty(shape, dtype=np.float64)
        check(lib().sph_get_field(self._h, N.FIELDS[name], _as_ptr(out), n))
        return out

    def set(self, name, values):
        n = len(self)
        shape = (n, 3) if name in N.VECTOR_FIELDS else (n,)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), shape))
        check(lib().sph_set_field(self._h, N.FIELDS[name], _as_ptr(a), n))

    def key_geometry(self):
        """(key_phase, key_lim, key_max, key_diff) as the reference computes them."""
        ph = (ctypes.c_int64 * 3)()
        lim = (ctypes.c_int64 * 3)()
        km = ctypes.c_int64()
        diff = (ctypes.c_int64 * 27)()
        nd = ctypes.c_int()
        check(lib().sph_key_geometry(self._h, ph, lim, ctypes.byref(km), diff, ctypes.byref(nd)))
        return tuple(ph), tuple(lim), km.value, list(diff)[: nd.value]

    def set_tile(self, T):
        check(lib().sph_set_tile(self._h, int(T)))

    def autotune(self, **params):
        best = ctypes.c_int()
        p = make_params(**params)
        check(lib().sph_autotune(self._h, ctypes.byref(p), ctypes.byref(best)))
        return best.value

    def set_pair_masks(self, enable):
        """Reuse the density pass's accepted-candidate masks in later pair
        passes on the same cell list (default on)."""
        check(lib().sph_set_option(self._h, N.SPH_OPTION_PAIR_MASKS, 1 if enable else 0))

    def set_kept_lists(self, enable):
        """Record the candidates that survive each tile's cull in the density
        pass and gather only those in the passes that read its masks
        (default on)."""
        check(lib().sph_set_option(self._h, N.SPH_OPTION_KEPT_LISTS, 1 if enable else 0))

    def set_graph(self, enable):
        """Replay the fused steps of a step() call from a hipGraph (default
        off; same bits)."""
        check(lib().sph_set_option(self._h, N.SPH_OPTION_GRAPH, 1 if enable else 0))

    def step_pairs(self):
        """(density-pass pairs, force-pass pairs) of the last step() call."""
        d = ctypes.c_int64()
        f = ctypes.c_int64()
        check(lib().sph_step_pairs(self._h, ctypes.byref(d), ctypes.byref(f)))
        return d.value, f.value

    def step_tests(self):
        """Distance tests of the density passes of the last step() call."""
        t = ctypes.c_int64()
        check(lib().sph_step_tests(self._h, ctypes.byref(t)))
        return t.value

    def candidates(self):
        """(candidates over all targets, over fluid targets) of one pass."""
        a = ctypes.c_int64()
        f = ctypes.c_int64()
        check(lib().sph_candidates(self._h, ctypes.byref(a), ctypes.byref(f)))
        return a.value, f.value

    def profile(self, enable=True):
        check(lib().sph_profile(self._h, 1 if enable else 0))

    def profile_read(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(lib().sph_profile_read(self._h, kernel, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class Snapshot:
    """An asynchronous download of fields (sph_snapshot_begin): started on
    the device at creation, in the order of the steps; wait() returns the
    arrays (reference order) once the copy is done.  Close it (or use it as
    a context manager) before the system is destroyed."""

    def __init__(self, begin, names):
        self.names = list(names)
        for nm in self.names:
            if nm not in N.FIELDS:
                raise KeyError(f"Variable {nm} does not exist!")
        ids = (ctypes.c_int * len(self.names))(*[N.FIELDS[nm] for nm in self.names])
        self._q = ctypes.c_void_p()
        check(begin(ids, len(self.names), ctypes.byref(self._q)))

    def wait(self):
        n = ctypes.c_int64()
        check(lib().sph_snapshot_wait(self._q, ctypes.byref(n)))
        out = {}
        for nm in self.names:
            ptr = _dptr()
            check(lib().sph_snapshot_field(self._q, N.FIELDS[nm], ctypes.byref(ptr)))
            k = 3 if nm in N.VECTOR_FIELDS else 1
            a = (np.ctypeslib.as_array(ptr, shape=(n.value * k,)).copy() if n.value
                 else np.empty(0))
            out[nm] = a.reshape(n.value, 3) if k == 3 else a
        return out

    def close(self):
        if self._q:
            lib().sph_snapshot_free(self._q)
            self._q = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParticleField:
    """ParticleField(sys, name) -- src/structs.jl:118-125: the n-th element
    is field `name` of the n-th particle (reference order)."""

    def __init__(self, sys, name):
        if name not in N.FIELDS:
            raise KeyError(f"Variable {name} does not exist!")
        self.sys = sys
        self.name = name

    def __len__(self):
        return len(self.sys)

    def __array__(self, dtype=None, copy=None):
        a = self.sys.get(self.name)
        return a if dtype is None else a.astype(dtype)

    def __getitem__(self, i):
        return self.sys.get(self.name)[i]

    def __setitem__(self, i, val):
        a = self.sys.get(self.name)
        a[i] = val
        self.sys.set(self.name, a)


def create_cell_list(sys):
    """create_cell_list!(sys) -- src/core.jl:51-90.  Returns the new count."""
    n = ctypes.c_int64()
    check(lib().sph_build_cells(sys._h, ctypes.byref(n)))
    return n.value


def _resolve(op):
    """A registered operator for an Operator, a name or "program:name";
    None for a Python callable (it runs through the host fallback)."""
    if isinstance(op, Operator):
        return op
    if isinstance(op, str):
        if ":" in op:
            prog, name = op.split(":", 1)
            ops = program_operators(prog)
            if name not in ops:
                raise KeyError(f"{prog} defines no operator {name!r}")
            return ops[name]
        if op in OPERATORS:
            return OPERATORS[op]
        if op in AMBIGUOUS:
            progs = sorted(p for p, ops in PROGRAMS.items() if op in ops)
            raise KeyError(
                f"{op!r} names different operators in {progs}; write "
                f"'{progs[0]}:{op}' or use program_operators()")
        raise KeyError(f"{op!r} is not a registered operator")
    if callable(op):
        return None
    raise TypeError(f"{op!r} is neither an operator nor a callable")


def _params(op, params):
    merged = dict(params)
    merged.update(op.fixed)
    return make_params(**merged)


def apply_binary(sys, op, **params):
    """apply_binary!(sys, op) -- src/core.jl:125-129.  A Python callable
    op(p, q, r) runs through the host fallback (host_apply)."""
    reg = _resolve(op)
    if reg is None:
        return host_apply(sys, op, binary=True)
    if not reg.binary:
        raise TypeError(f"{reg.name} is unary")
    p = _params(reg, params)
    check(lib().sph_apply_binary(sys._h, reg.id, ctypes.byref(p), 0))


def apply_unary(sys, op, **params):
    """apply_unary!(sys, op) -- src/core.jl:138-142.  A Python callable
    op(p) runs through the host fallback (host_apply)."""
    reg = _resolve(op)
    if reg is None:
        return host_apply(sys, op, binary=False)
    if reg.binary:
        raise TypeError(f"{reg.name} is binary")
    p = _params(reg, params)
    check(lib().sph_apply_unary(sys._h, reg.id, ctypes.byref(p)))


def apply(sys, op, self=False, **params):
    """apply!(sys, op; self=false, parameters...) -- src/core.jl:151-161.
    A Python callable is dispatched as the reference does
    (hasmethod(op, (T, T, Float64)), :152): three positional parameters make
    it binary, one unary; it runs through the host fallback."""
    reg = _resolve(op)
    if reg is None:
        return host_apply(sys, op, self=self, **params)
    p = _params(reg, params)
    check(lib().sph_apply(sys._h, reg.id, ctypes.byref(p), 1 if self else 0))


class Particle:
    """One particle of the host fallback: attribute access to its fields in
    the downloaded arrays (x, v, a -- also Dv --, rho, P, Drho, rho0,
    type).  Vector fields read as copies and are written whole, as the
    reference's immutable RealVector fields are (p.x += ... rebinds)."""

    __slots__ = ("_f", "_i")
    _ALIAS = {"Dv": "a"}

    def __init__(self, fields, i):
        object.__setattr__(self, "_f", fields)
        object.__setattr__(self, "_i", i)

    def __getattr__(self, name):
        f = self._f[self._ALIAS.get(name, name)]
        v = f[self._i]
        return v.copy() if f.ndim == 2 else float(v)

    def __setattr__(self, name, value):
        self._f[self._ALIAS.get(name, name)][self._i] = value

    def __eq__(self, other):  # identity, as == on a mutable struct
        return isinstance(other, Particle) and other._f is self._f and other._i == self._i

    def __hash__(self):
        return hash(self._i)


def _arity(func):
    import inspect

    try:
        sig = inspect.signature(func)
    except (TypeError, ValueError):
        return None
    pos = [p for p in sig.parameters.values()
           if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD) and p.default is p.empty]
    return len(pos)


def visit_order(sys):
    """The accepted pairs of the current cell list in the order
    _apply_binary! visits them (src/core.jl:94-112): by target p, then by
    the key_diff entry of q's cell (di outermost, :70-82 of
    src/structs.jl), then by descending q (the order add_index! keeps,
    src/core.jl:26-41).  Pairs and cell keys come from the device; the
    order is a sort on the host."""
    pairs = neighbour_pairs(sys)
    if len(pairs) == 0:
        return pairs
    n = len(sys)
    keys = np.empty(n, dtype=np.float64)
    check(lib().sph_get_field(sys._h, N.SPH_FIELD_KEY, _as_ptr(keys), n))
    keys = keys.astype(np.int64)
    diff = np.asarray(sys.key_geometry()[3], dtype=np.int64)
    delta = keys[pairs[:, 1]] - keys[pairs[:, 0]]
    # position of each pair's key difference in key_diff; a difference that
    # key_diff holds more than once (aliased grids) is visited once per entry,
    # and the device then lists the pair once per visit
    order_of = {}
    for i, d in enumerate(diff):
        order_of.setdefault(int(d), []).append(i)
    rank = np.empty(len(pairs), dtype=np.int64)
    idx = np.lexsort((-pairs[:, 1], delta, pairs[:, 0]))
    pairs, delta = pairs[idx], delta[idx]
    start = 0
    while start < len(pairs):
        end = start + 1
        while (end < len(pairs) and pairs[end, 0] == pairs[start, 0]
               and pairs[end, 1] == pairs[start, 1]):
            end += 1
        slots = order_of[int(delta[start])]
        rank[start:end] = slots[: end - start]
        start = end
    idx = np.lexsort((-pairs[:, 1], rank, pairs[:, 0]))
    return pairs[idx]


def host_apply(sys, func, self=False, binary=None, **params):
    """The host fallback for an operator that is not registered (a Python
    callable): download every field, run func over the particles (unary,
    func(p)) or over the accepted pairs of the current cell list in the
    reference's visiting order (binary, func(p, q, r)), then upload every
    field.  With self=True a binary func also gets func(p, p, 0.0) for each
    p after the pair pass (src/core.jl:155-157).  Keyword parameters are
    passed on to func, as apply! forwards them (:154).  Only the fields the
    closure changed are uploaded.  Correct but slow: the pairs are found on
    the device, the closure runs in Python."""
    if binary is None:
        k = _arity(func)
        if k not in (1, 3):
            raise TypeError("an operator takes (p) or (p, q, r)")
        binary = k == 3
    n = len(sys)
    fields = {k: sys.get(k) for k in N.FIELDS}
    before = {k: a.copy() for k, a in fields.items()}
    parts = [Particle(fields, i) for i in range(n)]
    call = (lambda *a: func(*a, **params)) if params else func
    if binary:
        x = fields["x"]
        pairs = visit_order(sys)
        for p, q in pairs:
            d = x[p] - x[q]
            r = float(np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
            call(parts[p], parts[q], r)
        if self:
            for i in range(n):
                call(parts[i], parts[i], 0.0)
    else:
        for i in range(n):
            call(parts[i])
    # upload what the closure changed (an unchanged x keeps the cell list)
    for k, a in fields.items():
        if not np.array_equal(a, before[k], equal_nan=True):
            sys.set(k, a)


def _reduce(sys, what, nout, **params):
    p = make_params(**params)
    out = (ctypes.c_double * 3)()
    check(lib().sph_reduce(sys._h, what, ctypes.byref(p), out))
    return list(out)[:nout]


def energy(sys, kind="wcsph", **params):
    """Total energy: kind "wcsph" = examples/collapse_dry.jl:166-171,
    kind "collision" = tests/test_collision_2d.jl:96-100, kind "wcsph_rho0" =
    the WCSPH energy with the per-particle reference density rho0_p."""
    what = {"wcsph": N.SPH_REDUCE_ENERGY, "collision": N.SPH_REDUCE_ENERGY_COLLISION,
            "wcsph_rho0": N.SPH_REDUCE_ENERGY_RHO0}[kind]
    return _reduce(sys, what, 1, **params)[0]


def momentum(sys, **params):
    """Sum of m v over all particles."""
    return np.array(_reduce(sys, N.SPH_REDUCE_MOMENTUM, 3, **params))


def neighbour_digest(sys):
    """(counts, hashes) of every particle's neighbour set under the current
    cell list: the number of accepted q and the sum modulo 2^64 of
    mix64(q + 1) over them (include/sph.h, sph_neighbour_digest)."""
    n = len(sys)
    c = np.empty(max(n, 1), dtype=np.int64)
    hsh = np.empty(max(n, 1), dtype=np.uint64)
    check(lib().sph_neighbour_digest(sys._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     hsh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n))
    return c[:n], hsh[:n]


def cell_keys(sys):
    """find_key of each particle at the last cell-list build (1-based)."""
    n = len(sys)
    keys = np.empty(n, dtype=np.float64)
    check(lib().sph_get_field(sys._h, N.SPH_FIELD_KEY, _as_ptr(keys), n))
    return keys.astype(np.int64)


def neighbour_pairs(sys):
    """All accepted ordered pairs (p, q) of the current cell list as
    reference indices, shape (npairs, 2), unspecified order."""
    n = ctypes.c_int64()
    check(lib().sph_neighbour_pairs(sys._h, None, 0, ctypes.byref(n)))
    total = n.value
    out = np.empty((max(total, 1), 2), dtype=np.int64)
    check(lib().sph_neighbour_pairs(
        sys._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), total, ctypes.byref(n)))
    return out[:total]


def probe_sums(sys, x, fluid_only=False, **params):
    """SmoothedParticles.sum(sys, func, x) (src/core.jl:240-258) at probe
    points x (shape (np, 3) or (3,)) for the interpolation closures of
    examples/cavity_flow.jl:166-172.  Returns shape (np, 4): column 0 is the
    sum of (phi*m)*W(h,r) (the Shepard denominator gamma), columns 1..3 the
    sums of ((phi*m)*v_k)*W(h,r), with phi = Float64(type == FLUID) when
    fluid_only, else 1, and m = params["m"].  Needs a current cell list."""
    xs = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, 3))
    out = np.zeros((xs.shape[0], 4), dtype=np.float64)
    p = make_params(**params)
    flags = N.SPH_PROBE_FLUID_ONLY if fluid_only else 0
    check(lib().sph_probe_sum(sys._h, ctypes.byref(p), flags, xs.ctypes.data_as(_dptr),
                              xs.shape[0], out.ctypes.data_as(_dptr)))
    return out


def particle_sums_lj_wall(sys, m, dr_wall, E_wall, eps=1e-16):
    """SmoothedParticles.sum(sys, LJ_potential, p) (src/core.jl:271-291) for
    every particle p, with LJ_potential of
    examples/collapse_symplectic.jl:144-151: the wall-potential energy of
    each particle (reference order).  q == p is visited too, as in the
    reference.  Needs a current cell list."""
    n = len(sys)
    out = np.empty(max(n, 1), dtype=np.float64)
    fp = np.array([m, dr_wall, E_wall, eps], dtype=np.float64)
    check(lib().sph_particle_sum(sys._h, N.SPH_SUM_LJ_WALL, _as_ptr(fp), 4, _as_ptr(out), n))
    return out[:n]


def _fields(sys):
    return {k: sys.get(k) for k in N.FIELDS}


def assemble_vector(sys, func):
    """assemble_vector(sys, func) (src/core.jl:175-182): v_i = func(p_i).
    `func` is vectorised: it receives a dict of the field arrays (reference
    order) and returns the n values."""
    v = np.asarray(func(_fields(sys)), dtype=np.float64).reshape(-1)
    if v.shape[0] != len(sys):
        raise ValueError(f"func returned {v.shape[0]} values for {len(sys)} particles")
    return v


def assemble_matrix(sys, func):
    """assemble_matrix(sys, func) (src/core.jl:196-225): the sparse matrix
    with A_ij = sum of func(p_i, p_j, r_ij) over the visits of the cell loop,
    which unlike apply_binary! includes q == p (r = 0).  The device finds the
    pairs (the current cell list's neighbour pairs); `func` is vectorised: it
    receives the field dict, the row and column index arrays and the FP64
    distances r = sqrt((d1*d1 + d2*d2) + d3*d3), and returns the values.
    Entries visited more than once (aliased small grids) are summed, as
    SparseArrays.sparse does.  Returns a scipy.sparse CSR matrix."""
    import scipy.sparse

    n = len(sys)
    pairs = neighbour_pairs(sys)
    # A particle meets itself once per key_diff entry equal to 0.
    self_visits = sum(1 for d in sys.key_geometry()[3] if d == 0)
    diag = np.repeat(np.arange(n, dtype=np.int64), self_visits)
    I = np.concatenate([pairs[:, 0], diag])
    J = np.concatenate([pairs[:, 1], diag])
    x = sys.get("x")
    d = x[I] - x[J]
    r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    V = np.asarray(func(_fields(sys), I, J, r), dtype=np.float64).reshape(-1)
    if V.shape[0] != I.shape[0]:
        raise ValueError("func must return one value per (i, j, r) entry")
    return scipy.sparse.csr_matrix((V, (I, J)), shape=(n, n))


_SCHEMES = {"wcsph_se": N.SPH_SCHEME_WCSPH_SE, "wcsph_se_rho0": N.SPH_SCHEME_WCSPH_SE_RHO0,
            "collapse_dry": N.SPH_SCHEME_COLLAPSE_DRY}


def step(sys, nsteps, scheme="wcsph_se", **params):
    """nsteps fused device-resident steps.  scheme "wcsph_se" (density sum,
    EOS with the constant rho0, internal force, symplectic Euler kick and
    drift), "wcsph_se_rho0" (the same with the per-particle rho0 field), or
    "collapse_dry" (the time loop of examples/collapse_dry.jl:202-211:
    half-step kick and drift, rebuild, balance_of_mass!, continuity EOS,
    half drift, rebuild, internal force, half kick; call collapse_dry_init
    first).  Returns the accepted ordered pairs summed over the first pair
    pass of each step (density sum, or balance_of_mass!)."""
    sid = _SCHEMES[scheme]
    p = make_params(**params)
    pairs = ctypes.c_int64()
    check(lib().sph_step(sys._h, sid, ctypes.byref(p), int(nsteps), ctypes.byref(pairs)))
    return pairs.value


def collapse_dry_init(sys, **params):
    """create_cell_list!(sys); apply!(sys, internal_force!): the set-up of
    the time loop of examples/collapse_dry.jl:198-200."""
    create_cell_list(sys)
    apply(sys, internal_force, **params)


def init_reference_density(sys, **params):
    """create_cell_list!(sys); apply!(sys, find_rho0!, self = true): the
    per-particle reference density of tests/test_collision_2d.jl:123-124."""
    create_cell_list(sys)
    sys.set("rho0", 0.0)
    apply(sys, find_rho0, self=True, **params)




class Batch:
    """Many independent particle systems stepped together (include/sph_batch.h).

    Each system keeps its own cell geometry, parameters and reference
    particle order and evolves exactly as it would alone; systems join with
    ``add`` and leave with ``retire`` between calls.  Every system of a batch
    has the same dimension.
    """

    def __init__(self, precision="mixed"):
        mode = {"fp64": N.SPH_FP64, "mixed": N.SPH_MIXED}[precision]
        self.precision = precision
        self._h = ctypes.c_void_p()
        check(lib().sph_batch_create(ctypes.byref(self._h), mode))

    def close(self):
        if self._h:
            lib().sph_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, x, h, bbox_lo, bbox_hi, periodic=False, v=None, rho=None, P=None,
            type=None, **params):
        """Add a system; returns its id.  ``params`` are those of
        make_params (m, rho0, c, alpha, hs, mu, nu, g, dt)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]

        def opt(a, shape):
            if a is None:
                return None
            return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape))

        v, rho, P, t = opt(v, (n, 3)), opt(rho, (n,)), opt(P, (n,)), opt(type, (n,))
        d = N.sph_system_desc()
        d.h = float(h)
        for i in range(3):
            d.bbox_min[i] = float(bbox_lo[i])
            d.bbox_max[i] = float(bbox_hi[i])
        d.periodic = 1 if periodic else 0
        d.params = make_params(**params)
        sid = ctypes.c_int()
        check(lib().sph_batch_add(self._h, ctypes.byref(d), n, _as_ptr(x),
                                  _as_ptr(v) if v is not None else None,
                                  _as_ptr(rho) if rho is not None else None,
                                  _as_ptr(P) if P is not None else None,
                                  _as_ptr(t) if t is not None else None, ctypes.byref(sid)))
        return sid.value

    def snapshot(self, sid, *names):
        """Start an asynchronous download of fields of system sid (Snapshot)."""
        return Snapshot(lambda ids, k, out: lib().sph_batch_snapshot_begin(self._h, int(sid), ids,
                                                                           k, out), names)

    def generate_lattice(self, sid, spec):
        """Device-side generate_particles! of a lattice into system sid (the
        last system holding particles); returns the number added."""
        n = ctypes.c_int64()
        check(lib().sph_batch_generate_lattice(self._h, int(sid), ctypes.byref(spec),
                                               ctypes.byref(n)))
        return n.value

    def add_state(self, st, periodic=False):
        """Add a workload state of ``states`` (dam break, Taylor-Green)."""
        return self.add(st.x, st.R, st.bbox_lo, st.bbox_hi, periodic=periodic, v=st.v,
                        rho=st.rho, P=st.P, type=st.type, **st.params())

    def retire(self, sid):
        check(lib().sph_batch_retire(self._h, int(sid)))

    def size(self):
        """(live systems, their particles)."""
        ns = ctypes.c_int()
        n = ctypes.c_int64()
        check(lib().sph_batch_size(self._h, ctypes.byref(ns), ctypes.byref(n)))
        return ns.value, n.value

    def count(self, sid):
        n = ctypes.c_int64()
        check(lib().sph_batch_count(self._h, int(sid), ctypes.byref(n)))
        return n.value

    def get(self, sid, name):
        n = self.count(sid)
        out = np.empty((n, 3) if name in N.VECTOR_FIELDS else (n,), dtype=np.float64)
        check(lib().sph_batch_get_field(self._h, int(sid), N.FIELDS[name], _as_ptr(out), n))
        return out

    def set(self, sid, name, values):
        n = self.count(sid)
        shape = (n, 3) if name in N.VECTOR_FIELDS else (n,)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), shape))
        check(lib().sph_batch_set_field(self._h, int(sid), N.FIELDS[name], _as_ptr(a), n))

    def reduce(self, sid, what):
        out = (ctypes.c_double * 3)()
        check(lib().sph_batch_reduce(self._h, int(sid), what, out))
        return list(out)

    def energy(self, sid, kind="wcsph_rho0"):
        what = {"wcsph": N.SPH_REDUCE_ENERGY, "collision": N.SPH_REDUCE_ENERGY_COLLISION,
                "wcsph_rho0": N.SPH_REDUCE_ENERGY_RHO0}[kind]
        return self.reduce(sid, what)[0]

    def momentum(self, sid):
        return np.array(self.reduce(sid, N.SPH_REDUCE_MOMENTUM))

    def build_cells(self):
        check(lib().sph_batch_build_cells(self._h))

    def probe_sums(self, sid, x, fluid_only=False):
        """probe_sums of one system (SmoothedParticles.sum at points,
        src/core.jl:240-258) with the mass of its own parameters."""
        xs = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, 3))
        out = np.zeros((xs.shape[0], 4), dtype=np.float64)
        flags = N.SPH_PROBE_FLUID_ONLY if fluid_only else 0
        check(lib().sph_batch_probe_sum(self._h, int(sid), flags, xs.ctypes.data_as(_dptr),
                                        xs.shape[0], out.ctypes.data_as(_dptr)))
        return out

    def neighbour_digest(self, sid):
        """neighbour_digest of one system under the current cell lists."""
        n = self.count(sid)
        c = np.empty(max(n, 1), dtype=np.int64)
        hsh = np.empty(max(n, 1), dtype=np.uint64)
        check(lib().sph_batch_neighbour_digest(
            self._h, int(sid), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            hsh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n))
        return c[:n], hsh[:n]

    def neighbour_pairs(self, sid):
        """Accepted ordered pairs of one system under the current cell lists
        (reference indices of that system), unspecified order."""
        n = ctypes.c_int64()
        check(lib().sph_batch_neighbour_pairs(self._h, int(sid), None, 0, ctypes.byref(n)))
        total = n.value
        out = np.empty((max(total, 1), 2), dtype=np.int64)
        check(lib().sph_batch_neighbour_pairs(
            self._h, int(sid), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), total,
            ctypes.byref(n)))
        return out[:total]

    def init_rho0(self):
        """find_rho0! (self = true) for the systems added since the last call."""
        check(lib().sph_batch_init_rho0(self._h))

    def step(self, nsteps, scheme="wcsph_se_rho0"):
        """nsteps fused steps of every system; returns the density-pass pairs."""
        pairs = ctypes.c_int64()
        check(lib().sph_batch_step(self._h, _SCHEMES[scheme], int(nsteps), ctypes.byref(pairs)))
        return pairs.value

    def step_pairs(self):
        d = ctypes.c_int64()
        f = ctypes.c_int64()
        check(lib().sph_step_pairs(self._h, ctypes.byref(d), ctypes.byref(f)))
        return d.value, f.value

    def set_tile(self, T):
        check(lib().sph_batch_set_tile(self._h, int(T)))

    def set_pair_masks(self, enable):
        """Reuse the density pass's accepted-candidate masks (default on); a
        batch handle is a handle of sph.h, so sph_set_option applies."""
        check(lib().sph_set_option(self._h, N.SPH_OPTION_PAIR_MASKS, 1 if enable else 0))

    def set_kept_lists(self, enable):
        """Gather only the candidates each tile's cull kept in the passes that
        read the masks (default on)."""
        check(lib().sph_set_option(self._h, N.SPH_OPTION_KEPT_LISTS, 1 if enable else 0))

    def set_graph(self, enable):
        """Replay the fused steps of a step() call from a hipGraph (default
        off; same bits)."""
        check(lib().sph_set_option(self._h, N.SPH_OPTION_GRAPH, 1 if enable else 0))

    def autotune(self):
        best = ctypes.c_int()
        check(lib().sph_batch_autotune(self._h, ctypes.byref(best)))
        return best.value

    def footprint(self):
        """Device bytes held by the batch."""
        b = ctypes.c_int64()
        check(lib().sph_batch_footprint(self._h, ctypes.byref(b)))
        return b.value

    def profile(self, enable=True):
        check(lib().sph_profile(self._h, 1 if enable else 0))

    def profile_read(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(lib().sph_profile_read(self._h, kernel, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
