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


def butcher_rows(scheme, dt):
    """(factor rows per stage, c) for a scheme: what perform_stage_hydro receives."""
    ai, bi, ci = _tableau(scheme)
    S = len(bi)
    rows = [np.array(ai[s]) * dt for s in range(S - 1)] + [np.array(bi) * dt]
    return rows, ci
