"""ctypes wrapper of the FP64 CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke``
and by the benchmark's CPU-baseline leg, as the checker of the HIP path.
The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

INFLOW, OUTFLOW, WALL, PERIODIC, FACE_STATES, FAR_FIELD = range(6)
IMIN, IMAX, JMIN, JMAX = range(4)
FIRST_ORDER, MINMOD, VAN_LEER, VAN_ALBADA, MONOTONIZED_CENTRAL, UNLIMITED_CENTRAL = range(6)
PER_VARIABLE, SHARED, SHARED_SCALED = range(3)
HLLC, RUSANOV, HLL = range(3)
FORWARD_EULER, SSPRK2, SSPRK3 = range(3)
RELAX_SWEEPS, RELAX_JACOBI, RELAX_RED_BLACK, RELAX_RED_BLACK_LINEAR = range(4)

BC_KINDS = {"inflow": INFLOW, "outflow": OUTFLOW, "wall": WALL, "periodic": PERIODIC,
            "faceStates": FACE_STATES, "farField": FAR_FIELD}
LIMITERS = {"firstOrder": FIRST_ORDER, "minmod": MINMOD, "vanLeer": VAN_LEER,
            "vanAlbada": VAN_ALBADA, "monotonizedCentral": MONOTONIZED_CENTRAL,
            "unlimitedCentral": UNLIMITED_CENTRAL}
RECONSTRUCTIONS = {"perVariable": PER_VARIABLE, "shared": SHARED, "sharedScaled": SHARED_SCALED}
FLUXES = {"hllc": HLLC, "rusanov": RUSANOV, "hll": HLL}
INTEGRATORS = {"forwardEuler": FORWARD_EULER, "ssprk2": SSPRK2, "ssprk3": SSPRK3}
RELAXATIONS = {"sweeps": RELAX_SWEEPS, "jacobi": RELAX_JACOBI, "redBlack": RELAX_RED_BLACK,
               "redBlackLinear": RELAX_RED_BLACK_LINEAR}


class OracleConfig(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("x", C.POINTER(C.c_double)),
                ("y", C.POINTER(C.c_double)), ("axisymmetric", C.c_int), ("gamma", C.c_double),
                ("flux", C.c_int), ("limiter", C.c_int), ("reconstruction", C.c_int),
                ("integrator", C.c_int), ("cfl", C.c_double), ("local_stepping", C.c_int),
                ("bc_kind", C.c_int * 4), ("bc_state", C.c_double * 16),
                ("source", C.POINTER(C.c_double))]


class OracleError(RuntimeError):
    pass


class NonPhysicalStateError(OracleError):
    pass


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.argtypes = [C.POINTER(OracleConfig), C.POINTER(C.c_void_p)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_compute_metrics.argtypes = [C.c_int, C.c_int, dp, dp, C.c_int] + [dp] * 10
        for name in ("orc_set_primitives", "orc_set_conserved"):
            getattr(L, name).argtypes = [C.c_void_p, dp]
        L.orc_get_conserved.argtypes = [C.c_void_p, dp]
        L.orc_get_clock.argtypes = [C.c_void_p, dp, C.POINTER(C.c_longlong)]
        L.orc_set_face_states.argtypes = [C.c_void_p, C.c_int, dp, C.c_int]
        L.orc_set_face_fluxes.argtypes = [C.c_void_p, C.c_int, dp, C.c_int]
        L.orc_iterate.argtypes = [C.c_void_p, C.c_int]
        L.orc_advance_to.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_stable_time_step.argtypes = [C.c_void_p]
        L.orc_stable_time_step.restype = C.c_double
        L.orc_time_steps.argtypes = [C.c_void_p, dp]
        L.orc_residual_vector.argtypes = [C.c_void_p, dp]
        L.orc_density_residual_norm.argtypes = [C.c_void_p]
        L.orc_density_residual_norm.restype = C.c_double
        L.orc_boundary_fluxes.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_implicit_correction.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, dp]
        L.orc_iterate_implicit.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_forebody_pressure_drag.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_forebody_pressure_drag.restype = C.c_double
        L.orc_face_frame_flux.argtypes = [C.c_int, dp, dp, C.c_double, dp]
        L.orc_face_flux_2d.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_limiter.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_limiter.restype = C.c_double
        L.orc_neighbour_term_linear.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                                C.c_double, dp]
        L.orc_neighbour_term_linear.restype = None
        L.orc_reconstruct.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_int, C.c_double, dp]
        L.orc_far_field_state.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc):
    if rc == 0:
        return
    msg = lib().orc_last_error().decode()
    if rc == 2:
        raise NonPhysicalStateError(msg)
    raise OracleError(msg)


def compute_metrics(grid, axisymmetric):
    ni, nj = grid.ni, grid.nj
    out = dict(inx=np.zeros((ni + 1) * nj), iny=np.zeros((ni + 1) * nj), iarea=np.zeros((ni + 1) * nj),
               jnx=np.zeros(ni * (nj + 1)), jny=np.zeros(ni * (nj + 1)), jarea=np.zeros(ni * (nj + 1)),
               volume=np.zeros(ni * nj), planar=np.zeros(ni * nj), cx=np.zeros(ni * nj), cy=np.zeros(ni * nj))
    x = np.ascontiguousarray(grid.x, dtype=np.float64)
    y = np.ascontiguousarray(grid.y, dtype=np.float64)
    _check(lib().orc_compute_metrics(ni, nj, _p(x), _p(y), int(axisymmetric),
                                     *[_p(out[k]) for k in ("inx", "iny", "iarea", "jnx", "jny", "jarea",
                                                             "volume", "planar", "cx", "cy")]))
    return out


class Euler2D:
    """FP64 CPU engine with the reference's method set (euler2d.ts:290)."""

    def __init__(self, grid, *, geometry="axisymmetric", gamma=1.4, flux="hllc", limiter="minmod",
                 reconstruction="perVariable", time_integrator="ssprk3", cfl=0.5,
                 time_stepping="local", boundaries=None, source=None):
        self.grid = grid
        self.ni, self.nj = grid.ni, grid.nj
        self.gamma = gamma
        self._x = np.ascontiguousarray(grid.x, dtype=np.float64)
        self._y = np.ascontiguousarray(grid.y, dtype=np.float64)
        cfg = OracleConfig()
        cfg.ni, cfg.nj = grid.ni, grid.nj
        cfg.x, cfg.y = _p(self._x), _p(self._y)
        cfg.axisymmetric = 1 if geometry == "axisymmetric" else 0
        cfg.gamma = gamma
        cfg.flux = FLUXES[flux]
        cfg.limiter = LIMITERS[limiter]
        cfg.reconstruction = RECONSTRUCTIONS[reconstruction]
        cfg.integrator = INTEGRATORS[time_integrator]
        cfg.cfl = cfl
        cfg.local_stepping = 1 if time_stepping == "local" else 0
        for k, side in enumerate(("iMin", "iMax", "jMin", "jMax")):
            bc = boundaries[side]
            cfg.bc_kind[k] = BC_KINDS[bc["kind"]]
            st = bc.get("state", (0.0, 0.0, 0.0, 0.0))
            for q in range(4):
                cfg.bc_state[4 * k + q] = st[q]
        self._src = None
        if source is not None:
            self._src = np.ascontiguousarray(source, dtype=np.float64)
            cfg.source = _p(self._src)
        h = C.c_void_p()
        _check(lib().orc_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.metrics = compute_metrics(grid, geometry == "axisymmetric")
        for k, side in enumerate(("iMin", "iMax", "jMin", "jMax")):
            bc = boundaries[side]
            if bc["kind"] == "faceStates" and "states" in bc:
                self.set_face_states(side, bc["states"])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    @property
    def n(self):
        return self.ni * self.nj

    def set_primitives(self, prim):
        """prim: (n, 4) rho, u, v, p per cell; resets time and steps."""
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(prim, dtype=np.float64), (self.n, 4)))
        _check(lib().orc_set_primitives(self._h, _p(a)))

    def set_initial_condition(self, field):
        """field(x, y) -> (rho, u, v, p) at the cell centroids."""
        cx, cy = self.metrics["cx"], self.metrics["cy"]
        self.set_primitives(np.array([field(a, b) for a, b in zip(cx, cy)]))

    def set_conserved(self, q):
        a = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        _check(lib().orc_set_conserved(self._h, _p(a)))

    def conserved(self):
        q = np.zeros(4 * self.n)
        lib().orc_get_conserved(self._h, _p(q))
        return q.reshape(self.n, 4)

    def primitives(self):
        q = self.conserved()
        rho = q[:, 0]
        u, v = q[:, 1] / rho, q[:, 2] / rho
        p = (self.gamma - 1) * (q[:, 3] - 0.5 * (q[:, 1] ** 2 + q[:, 2] ** 2) / rho)
        return np.stack([rho, u, v, p], axis=1)

    @property
    def time(self):
        t = C.c_double()
        lib().orc_get_clock(self._h, C.byref(t), None)
        return t.value

    @property
    def steps(self):
        s = C.c_longlong()
        lib().orc_get_clock(self._h, None, C.byref(s))
        return s.value

    def set_face_states(self, side, states):
        a = np.ascontiguousarray(states, dtype=np.float64).reshape(-1)
        _check(lib().orc_set_face_states(self._h, ("iMin", "iMax", "jMin", "jMax").index(side), _p(a), len(a) // 4))

    def set_face_fluxes(self, side, fluxes):
        k = ("iMin", "iMax", "jMin", "jMax").index(side)
        if fluxes is None:
            _check(lib().orc_set_face_fluxes(self._h, k, None, 0))
            return
        a = np.ascontiguousarray(fluxes, dtype=np.float64).reshape(-1)
        _check(lib().orc_set_face_fluxes(self._h, k, _p(a), len(a)))

    def iterate(self, count):
        _check(lib().orc_iterate(self._h, int(count)))

    def advance_to(self, t_end, max_steps=1_000_000):
        _check(lib().orc_advance_to(self._h, float(t_end), int(max_steps)))

    def stable_time_step(self):
        return lib().orc_stable_time_step(self._h)

    def time_steps(self):
        dt = np.zeros(self.n)
        lib().orc_time_steps(self._h, _p(dt))
        return dt

    def residual_vector(self):
        r = np.zeros(4 * self.n)
        lib().orc_residual_vector(self._h, _p(r))
        return r.reshape(self.n, 4)

    def density_residual_norm(self):
        return lib().orc_density_residual_norm(self._h)

    def boundary_fluxes(self, side):
        k = ("iMin", "iMax", "jMin", "jMax").index(side)
        m = self.nj if k < 2 else self.ni
        out = np.zeros(9 * m)
        _check(lib().orc_boundary_fluxes(self._h, k, _p(out)))
        return out.reshape(m, 9)

    def forebody_drag(self, p_inf, q_inf, base_radius, side="jMin"):
        f = np.ascontiguousarray(self.boundary_fluxes(side))
        return lib().orc_forebody_pressure_drag(_p(f), f.shape[0], p_inf, q_inf, base_radius)

    def implicit_correction(self, cfl, omega=1.5, relaxation="sweeps", sweeps=1):
        dq = np.zeros(4 * self.n)
        _check(lib().orc_implicit_correction(self._h, cfl, omega, RELAXATIONS[relaxation], sweeps, _p(dq)))
        return dq.reshape(self.n, 4)

    def iterate_implicit(self, count, cfl, omega=1.5, relaxation="sweeps", sweeps=1):
        _check(lib().orc_iterate_implicit(self._h, int(count), cfl, omega, RELAXATIONS[relaxation], sweeps))


def face_frame_flux(flux, left, right, gamma):
    out = np.zeros(4)
    lib().orc_face_frame_flux(FLUXES[flux], _p(np.asarray(left, float)), _p(np.asarray(right, float)), gamma, _p(out))
    return out


def face_flux_2d(flux, left, right, nx, ny, gamma):
    out = np.zeros(4)
    lib().orc_face_flux_2d(FLUXES[flux], _p(np.asarray(left, float)), _p(np.asarray(right, float)), nx, ny, gamma, _p(out))
    return out


def limiter(name, a, b):
    return lib().orc_limiter(LIMITERS[name], a, b)


def reconstruct(reconstruction, limiter_name, prev, centre, nxt, side, gamma):
    out = np.zeros(4)
    lib().orc_reconstruct(RECONSTRUCTIONS[reconstruction], LIMITERS[limiter_name],
                          _p(np.asarray(prev, float)), _p(np.asarray(centre, float)),
                          _p(np.asarray(nxt, float)), side, gamma, _p(out))
    return out


def far_field_state(inside, free, nx, ny, gamma):
    out = np.zeros(4)
    lib().orc_far_field_state(_p(np.asarray(inside, float)), _p(np.asarray(free, float)), nx, ny, gamma, _p(out))
    return out


def neighbour_term_linear(q, d, gamma, nx, ny, area, lam_a, omega):
    """The linearized neighbour term (euler2d-two-float.wgsl:574-598)."""
    out = np.zeros(4)
    lib().orc_neighbour_term_linear(_p(np.asarray(q, float)), _p(np.asarray(d, float)), gamma, nx, ny, area, lam_a,
                                    omega, _p(out))
    return out
