"""Explicit Runge-Kutta time integration around OceanoOperator.perform_stage_hydro.

Mirrors TimeIntegrator::ExplicitRungeKuttaIntegrator (ExplicitRungeKuttaIntegrator.h:
77-298): Butcher tableaux of TimeSteppingOceano::ExplicitRungeKutta
(time_stepping.h:137-255), dt-scaled rows, one perform_stage_hydro call per stage with
factor = a[stage] (b at the last stage) and the stage inputs {r_h, r_q, tracer, velocity}.
The reference copies four vectors into `vec_ri` per stage
(ExplicitRungeKuttaIntegrator.h:207-210, 255-258) so that the stage may overwrite
next_ri while reading it; here two r registers alternate instead (stage s reads one and
writes the other), which gives the same values without the copies.
"""
import math

import numpy as np


def _tableau(name):
    if name == "stage_1_order_1":  # FORWARD_EULER
        return [[1.0]], [1.0], [0.0]
    if name == "stage_2_order_2":  # SSP_SECOND_ORDER
        return [[1.0]], [0.5, 0.5], [0.0, 1.0]
    if name == "stage_3_order_3":  # SSP_THIRD_ORDER
        return [[1.0], [0.25, 0.25]], [1 / 6, 1 / 6, 2 / 3], [0.0, 1.0, 0.5]
    if name == "stage_4_order_4":  # RK_CLASSIC_FOURTH_ORDER
        return ([[0.5], [0.0, 0.5], [0.0, 0.0, 1.0]], [1 / 6, 1 / 3, 1 / 3, 1 / 6],
                [0.0, 0.5, 0.5, 1.0])
    if name == "stage_3_order_2":  # THREE_STAGE_SECOND_ORDER (Giraldo)
        chi = 2.0 - math.sqrt(2.0)
        a32 = 0.5
        return ([[chi], [1.0 - a32, a32]], [0.5 - chi / 4, 0.5 - chi / 4, chi / 2],
                [0.0, chi, 1.0])
    if name == "lsrk_5_order_4":
        return lsrk54_butcher()
    raise ValueError("Unimplemented explicit Runge-Kutta method: %s" % name)


def lsrk54_butcher():
    """Carpenter & Kennedy (1994) five-stage fourth-order 2N-storage scheme, written as a
    Butcher tableau (rows a[s] build the input of stage s+1)."""
    A = [0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
         -3550918686646 / 2091501179385, -1275806237668 / 842570457699]
    B = [1432997174477 / 9575080441755, 5161836677717 / 13612068292357,
         1720146321549 / 2090206949498, 3134564353537 / 4481467310338,
         2277821191437 / 14882151754819]
    Cc = [0.0, 1432997174477 / 9575080441755, 2526269341429 / 6820363962896,
          2006345519317 / 3224310063776, 2802321613138 / 2924317926251]
    S = 5

    def coeff(j, i):  # weight of dt*k_j in u_i (stages 0-based, j <= i)
        s = 0.0
        for l in range(j, i + 1):
            p = B[l]
            for m in range(j + 1, l + 1):
                p *= A[m]
            s += p
        return s

    a = [[coeff(j, i) for j in range(i + 1)] for i in range(S - 1)]
    b = [coeff(j, S - 1) for j in range(S)]
    return a, b, Cc


class ExplicitRungeKuttaIntegrator:
    def __init__(self, scheme="stage_4_order_4"):
        self.ai, self.bi, self.ci = _tableau(scheme)
        self.scheme = scheme

    def n_stages(self):
        return len(self.bi)

    def reinit(self, op):
        """Auxiliary registers: r_h, r_q and n_stages+1 k vectors per field."""
        S = self.n_stages()
        self.r_h = [op.initialize_vector(0), op.initialize_vector(0)]
        self.r_q = [op.initialize_vector(1), op.initialize_vector(1)]
        self.ki_h = [op.initialize_vector(0) for _ in range(S + 1)]
        self.ki_q = [op.initialize_vector(1) for _ in range(S + 1)]

    def perform_time_step(self, op, current_time, time_step, solution_height,
                          solution_discharge, postprocess_velocity=None):
        S = self.n_stages()
        b = [v * time_step for v in self.bi]
        a = [[v * time_step for v in row] for row in self.ai]
        for stage in range(S):
            last = stage == S - 1
            if stage == 0:
                ri = [solution_height, solution_discharge, None, postprocess_velocity]
            else:
                ri = [self.r_h[stage % 2], self.r_q[stage % 2], None, postprocess_velocity]
            nxt = (stage + 1) % 2
            op.perform_stage_hydro(stage, current_time + self.ci[stage] * time_step,
                                   b if last else a[stage], ri, self.ki_h, self.ki_q,
                                   solution_height, solution_discharge, self.r_h[nxt],
                                   self.r_q[nxt])


def ark_tableau(name="stage_3_order_2"):
    """Additive (IMEX) RK of AdditiveRungeKuttaIntegrator.h:149-175: explicit
    THREE_STAGE_SECOND_ORDER (time_stepping.h:223-245) with the implicit TRAPEZOIDAL_BDF2
    (time_stepping.h:340-379, Giraldo et al. 2012).  Returns (a, b, c, a_tilde, d_tilde):
    a, a_tilde the strictly lower rows feeding stages 1.., d_tilde the diagonal implicit
    coefficients of stages 0..S-2 (the last stage has no implicit solve)."""
    if name != "stage_3_order_2":
        raise ValueError("unknown additive RK scheme " + name)
    chi, a32 = 2.0 - np.sqrt(2.0), 0.5
    a = [[chi], [1.0 - a32, a32]]
    b = [0.5 - chi / 4.0, 0.5 - chi / 4.0, chi / 2.0]
    c = [0.0, chi, 1.0]
    at = [[chi / 2.0], [1.0 / (2.0 * np.sqrt(2.0)), 1.0 / (2.0 * np.sqrt(2.0))]]
    dt_ = [chi / 2.0, 1.0 - 1.0 / np.sqrt(2.0)]
    return a, b, c, at, dt_


class AdditiveRungeKuttaIntegrator:
    """IMEX time step of AdditiveRungeKuttaIntegrator::perform_time_step
    (AdditiveRungeKuttaIntegrator.h:207-339): the stiff Manning friction implicit, the rest
    explicit, one perform_stage_hydro_ark call per stage with factor = a[stage]
    (b at the last stage), factor_tilde = a~[stage] (b at the last stage) and the implicit
    diagonal factor_matrix = d~[stage] dt.  The reference reads d~[S-1] at the last stage,
    one past the two coefficients TRAPEZOIDAL_BDF2 defines; the last stage has no implicit
    solve, so 0 is passed there.  Two r registers alternate as in the explicit integrator."""

    def __init__(self, scheme="stage_3_order_2"):
        self.ai, self.bi, self.ci, self.ati, self.dti = ark_tableau(scheme)
        self.scheme = scheme

    def n_stages(self):
        return len(self.bi)

    def reinit(self, op):
        """r registers and vec_ki: n_stages+1 height and 2 n_stages discharge vectors
        (AdditiveRungeKuttaIntegrator.h:158-182)."""
        S = self.n_stages()
        self.r_h = [op.initialize_vector(0), op.initialize_vector(0)]
        self.r_q = [op.initialize_vector(1), op.initialize_vector(1)]
        self.ki_h = [op.initialize_vector(0) for _ in range(S + 1)]
        self.ki_q = [op.initialize_vector(1) for _ in range(2 * S)]

    def perform_time_step(self, op, current_time, time_step, solution_height,
                          solution_discharge, postprocess_velocity=None):
        S = self.n_stages()
        b = [v * time_step for v in self.bi]
        for stage in range(S):
            last = stage == S - 1
            fac = b if last else [v * time_step for v in self.ai[stage]]
            fat = b if last else [v * time_step for v in self.ati[stage]]
            fm = self.dti[stage] * time_step if stage < len(self.dti) else 0.0
            if stage == 0:
                ri = [solution_height, solution_discharge, None, postprocess_velocity]
            else:
                ri = [self.r_h[stage % 2], self.r_q[stage % 2], None, postprocess_velocity]
            nxt = (stage + 1) % 2
            op.perform_stage_hydro_ark(stage, current_time + self.ci[stage] * time_step, fac,
                                       fat, fm, ri, self.ki_h, self.ki_q, solution_height,
                                       solution_discharge, self.r_h[nxt], self.r_q[nxt])


def butcher_rows(scheme, dt):
    """(factor rows per stage, c) for a scheme: what perform_stage_hydro receives."""
    ai, bi, ci = _tableau(scheme)
    S = len(bi)
    rows = [np.array(ai[s]) * dt for s in range(S - 1)] + [np.array(bi) * dt]
    return rows, ci


class DeviceIntegrator:
    """Device-resident time loop of the library (oc_rk_*): the S stages of a scheme with
    integrator-owned k / r registers, optionally followed by the velocity update, launched
    eagerly (step) or captured once into a HIP graph and replayed (capture / replay).
    ExplicitRungeKuttaIntegrator.h:179-298 driven as in main.cc:1546-1577."""

    def __init__(self, op, scheme="stage_4_order_4"):
        import ctypes as C
        from . import _capi
        ai, bi, _ = _tableau(scheme)
        rows = [v for s in range(len(bi) - 1) for v in ai[s]] + list(bi)
        arr = (C.c_double * len(rows))(*rows)
        h = C.c_void_p()
        _capi.check(_capi.lib().oc_rk_create(op._h, len(bi), arr, C.byref(h)), "oc_rk_create")
        self._h, self.op, self.n_stages = h, op, len(bi)
        self._lib = _capi.lib()  # kept for __del__, which may run at interpreter shutdown

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.oc_rk_destroy(self._h)
            self._h = None

    @staticmethod
    def _p(t):
        return None if t is None else t.data_ptr()

    def set_time(self, t):
        """time of the next step (stage s runs at t + c_s dt; every step advances it)"""
        from . import _capi
        _capi.check(_capi.lib().oc_rk_set_time(self._h, float(t)), "oc_rk_set_time")

    def time(self):
        import ctypes as C
        from . import _capi
        t = C.c_double()
        _capi.check(_capi.lib().oc_rk_time(self._h, C.byref(t)), "oc_rk_time")
        return t.value

    def step(self, dt, w_h, w_q, velocity=None):
        from . import _capi
        from .operator import _stream
        _capi.check(_capi.lib().oc_rk_step(self._h, dt, self._p(w_h), self._p(w_q),
                                           self._p(velocity), _stream()), "oc_rk_step")

    def capture(self, dt, w_h, w_q, velocity=None, n_steps=1):
        """Record n_steps steps into one graph on torch's current stream (must not be the
        null stream); each replay runs them all."""
        from . import _capi
        from .operator import _stream
        _capi.check(_capi.lib().oc_rk_capture_steps(self._h, n_steps, dt, self._p(w_h),
                                                    self._p(w_q), self._p(velocity),
                                                    _stream()), "oc_rk_capture")

    def replay(self, n=1):
        from . import _capi
        from .operator import _stream
        _capi.check(_capi.lib().oc_rk_replay(self._h, n, _stream()), "oc_rk_replay")

    def run_to_steady(self, w_h, w_q, dt, tol=1e-10, max_steps=200000, chunk=500,
                      stream=None):
        """Replays the captured-step graph until one step changes the state by less than
        tol in relative L2 (the C3 criterion ||w^{n+1} - w^n|| / ||w^n||, zeta and q
        together), checked every `chunk` steps.  Returns (steps, last relative change)."""
        import torch
        side = stream or torch.cuda.Stream()
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            self.capture(dt, w_h, w_q)
        side.synchronize()
        steps, change = 0, float("inf")
        while steps < max_steps:
            with torch.cuda.stream(side):
                self.replay(chunk - 1)
                ph, pq = w_h.clone(), w_q.clone()
                self.replay(1)
                d = torch.sqrt(((w_h - ph) ** 2).sum() + ((w_q - pq) ** 2).sum())
                n = torch.sqrt((ph ** 2).sum() + (pq ** 2).sum())
                change = float(d / n)
            steps += chunk
            if not change == change or change < tol:  # NaN or converged
                break
        side.synchronize()
        return steps, change


class EnsembleIntegrator:
    """The device-resident RK loop over several operators advanced together (oc_rk_*_multi):
    one launch per stage for all of them (the batched ensemble's degree groups, one operator
    per degree over the union of that degree's instances).  step / capture take lists of
    solution vectors, one pair per operator; replay runs the captured graph."""

    def __init__(self, ops, scheme="stage_4_order_4"):
        import ctypes as C
        from . import _capi
        ai, bi, _ = _tableau(scheme)
        rows = [v for s in range(len(bi) - 1) for v in ai[s]] + list(bi)
        arr = (C.c_double * len(rows))(*rows)
        handles = (C.c_void_p * len(ops))(*[op._h.value for op in ops])
        h = C.c_void_p()
        _capi.check(_capi.lib().oc_rk_create_multi(len(ops), handles, len(bi), arr,
                                                   C.byref(h)), "oc_rk_create_multi")
        self._h, self.ops, self.n_stages = h, list(ops), len(bi)
        self._lib = _capi.lib()  # kept for __del__, which may run at interpreter shutdown

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.oc_rk_destroy(self._h)
            self._h = None

    @staticmethod
    def _ptrs(ts):
        import ctypes as C
        return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    def set_time(self, t):
        from . import _capi
        _capi.check(_capi.lib().oc_rk_set_time(self._h, float(t)), "oc_rk_set_time")

    def _dts(self, dt):
        import ctypes as C
        d = list(dt) if hasattr(dt, "__len__") else [dt] * len(self.ops)
        return (C.c_double * len(d))(*d)

    def step(self, dt, ws_h, ws_q):
        """dt: one step for all operators or one per operator"""
        from . import _capi
        from .operator import _stream
        _capi.check(_capi.lib().oc_rk_step_multi(self._h, self._dts(dt), self._ptrs(ws_h),
                                                 self._ptrs(ws_q), _stream()),
                    "oc_rk_step_multi")

    def capture(self, dt, ws_h, ws_q, n_steps=1):
        from . import _capi
        from .operator import _stream
        _capi.check(_capi.lib().oc_rk_capture_multi(self._h, n_steps, self._dts(dt),
                                                    self._ptrs(ws_h), self._ptrs(ws_q),
                                                    _stream()), "oc_rk_capture_multi")

    def replay(self, n=1):
        from . import _capi
        from .operator import _stream
        _capi.check(_capi.lib().oc_rk_replay(self._h, n, _stream()), "oc_rk_replay")
