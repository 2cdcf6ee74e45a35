"""OceanoOperator: the host-side mirror of SpaceDiscretization::OceanoOperator.

Same member names, argument meaning and error behaviour as the reference class
(OceanDG.h:157-380): construction + reinit set the operator up on a mesh,
initialize_data_at_quadrature tabulates the problem data at the quadrature points
(OceanDG.h:498-569), perform_stage_hydro runs one explicit Runge-Kutta stage
(OceanDG.h:182-194, 1782-1912) and apply evaluates M^-1 L (OceanDG.h:178-180).
Vectors are torch tensors on the GPU in the operator's storage order (see
include/oceano_dg.h); every call goes through liboceano_dg.so, which raises on failure.
"""
import ctypes as C

import numpy as np
import torch

from . import _capi
from ._capi import check, dptr, lib

BC_KINDS = {
    "wall": _capi.OC_BC_WALL,
    "supercritical_inflow": _capi.OC_BC_SUPERCRITICAL_INFLOW,
    "supercritical_outflow": _capi.OC_BC_SUPERCRITICAL_OUTFLOW,
    "height_inflow": _capi.OC_BC_HEIGHT_INFLOW,
    "discharge_inflow": _capi.OC_BC_DISCHARGE_INFLOW,
    "absorbing_outflow": _capi.OC_BC_ABSORBING_OUTFLOW,
}

STORES = {
    "cell_0": _capi.OC_STORE_CELL_GAUSS,
    "cell_1": _capi.OC_STORE_CELL_LOBATTO,
    "cell_2": _capi.OC_STORE_CELL_COLLOC,
    "face": _capi.OC_STORE_FACE,
    "boundary": _capi.OC_STORE_BOUNDARY,
}


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class OceanoOperator:
    """GPU operator for one mesh and one polynomial degree (1..4)."""

    def __init__(self, mesh, degree, g=9.81, threshold_wetdry=0.1, viscosity=0.0,
                 max_iteration_height=15, precision="fp64", boundary_conditions=None,
                 device="cuda"):
        self.mesh = mesh
        self.degree = degree
        self.device = torch.device(device)
        self.precision = precision
        bcs = dict(boundary_conditions or {})
        ids = np.array(list(bcs.keys()), np.int32)
        kinds = np.array([BC_KINDS[k] if isinstance(k, str) else int(k) for k in bcs.values()],
                         np.int32)
        self._keep = (mesh.geom, mesh.inner, mesh.boundary, ids, kinds)
        desc = _capi.MeshDesc(mesh.n_cells, dptr(mesh.geom), mesh.inner.shape[0],
                              _capi.iptr(mesh.inner), mesh.boundary.shape[0],
                              _capi.iptr(mesh.boundary))
        prm = _capi.Params(degree, g, threshold_wetdry, viscosity, max_iteration_height,
                           {"fp64": _capi.OC_PRECISION_FP64,
                            "mixed": _capi.OC_PRECISION_MIXED}[precision],
                           len(ids), _capi.iptr(ids), _capi.iptr(kinds))
        self.g = g
        self.viscosity = viscosity
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().oc_dg_create(C.byref(desc), C.byref(prm), C.byref(h)), "oc_dg_create")
        self._h = h
        sizes = np.zeros(7, np.int64)
        check(lib().oc_dg_sizes(self._h, sizes.ctypes.data_as(_capi.PI64)), "oc_dg_sizes")
        (self.n_cells, self.n_cells_padded, self.dofs_per_cell, self.n_q_points_1d,
         self.n_lobatto_1d, self.n_inner, self.n_boundary) = (int(v) for v in sizes)

    # ---------------------------------------------------------------- set-up -------
    def quadrature_points(self, store):
        s = STORES[store]
        n = np.zeros(1, np.int64)
        check(lib().oc_dg_n_points(self._h, s, n.ctypes.data_as(_capi.PI64)), "oc_dg_n_points")
        xy = np.zeros((int(n[0]), 2))
        check(lib().oc_dg_quadrature_points(self._h, s, dptr(xy)), "oc_dg_quadrature_points")
        return xy

    # values per point of each store (OC_STORE_* in oceano_dg.h)
    VALUES_PER_POINT = {"cell_0": 5, "cell_1": 1, "cell_2": 2, "face": 1, "boundary": 1}

    def n_points(self, store):
        n = np.zeros(1, np.int64)
        check(lib().oc_dg_n_points(self._h, STORES[store], n.ctypes.data_as(_capi.PI64)),
              "oc_dg_n_points")
        return int(n[0])

    def set_static_data(self, store, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        expect = self.n_points(store) * self.VALUES_PER_POINT[store]
        if v.size != expect:
            raise ValueError("%s: %d values given, the store holds %d points x %d values"
                             % (store, v.size, self.n_points(store),
                                self.VALUES_PER_POINT[store]))
        check(lib().oc_dg_set_static_data(self._h, STORES[store], dptr(v)),
              "oc_dg_set_static_data")

    def initialize_data_at_quadrature(self, problem_data):
        """problem_data(xy[n,2]) -> [n,5] = z_b, friction, tau_x, tau_y, coriolis."""
        data = {}
        data["cell_0"] = problem_data(self.quadrature_points("cell_0"))
        data["cell_1"] = problem_data(self.quadrature_points("cell_1"))[:, 0]
        data["cell_2"] = problem_data(self.quadrature_points("cell_2"))[:, :2]
        data["face"] = problem_data(self.quadrature_points("face"))[:, 0]
        data["boundary"] = problem_data(self.quadrature_points("boundary"))[:, 0]
        for k, v in data.items():
            self.set_static_data(k, v)
        return data

    def set_boundary_values(self, values):
        """time-independent boundary data [n_boundary, nq, 3] (used at every time)"""
        v = np.ascontiguousarray(values, dtype=np.float64)
        expect = self.n_boundary * self.n_q_points_1d * 3
        if v.size != expect:
            raise ValueError("boundary values: %d given, expected n_boundary x nq x 3 = %d"
                             % (v.size, expect))
        check(lib().oc_dg_set_boundary_values(self._h, dptr(v)), "oc_dg_set_boundary_values")

    def set_boundary_table(self, times, values):
        """boundary data at the given times, [n_times, n_boundary, nq, 3]: every call with a
        time uses the entry at that time (OceanDG.h:1798-1803 sets the boundary Functions'
        time each stage); see oc_dg_set_boundary_table for the order a replayed time loop
        reads"""
        t = np.ascontiguousarray(times, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(values, dtype=np.float64)
        expect = t.size * self.n_boundary * self.n_q_points_1d * 3
        if v.size != expect:
            raise ValueError("boundary table: %d values given, expected n_times x n_boundary "
                             "x nq x 3 = %d" % (v.size, expect))
        check(lib().oc_dg_set_boundary_table(self._h, t.size, dptr(t), dptr(v)),
              "oc_dg_set_boundary_table")

    # --------------------------------------------------------------- vectors -------
    def vector_length(self, n_components):
        n = np.zeros(1, np.int64)
        check(lib().oc_dg_vector_length(self._h, n_components, n.ctypes.data_as(_capi.PI64)),
              "oc_dg_vector_length")
        return int(n[0])

    def initialize_vector(self, variable):
        """variable 0 = free surface (1 component), 1 = discharge / velocity (2)."""
        nc = 1 if variable == 0 else 2
        return torch.zeros(self.vector_length(nc), dtype=torch.float64, device=self.device)

    def pack(self, cell_major, n_components):
        src = torch.as_tensor(cell_major, dtype=torch.float64).to(self.device).contiguous()
        dst = torch.empty(self.vector_length(n_components), dtype=torch.float64,
                          device=self.device)
        check(lib().oc_dg_pack(self._h, n_components, _vp(src), _vp(dst), _stream()),
              "oc_dg_pack")
        return dst

    def unpack(self, op_vec, n_components):
        dst = torch.empty(self.n_cells * n_components * self.dofs_per_cell, dtype=torch.float64,
                          device=self.device)
        check(lib().oc_dg_unpack(self._h, n_components, _vp(op_vec), _vp(dst), _stream()),
              "oc_dg_unpack")
        return dst

    def project_collocation(self, values, n_components):
        """Nodal coefficients (cell-major, host) from values at the cell_2 points."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        out = np.zeros(self.n_cells * n_components * self.dofs_per_cell)
        check(lib().oc_dg_project_collocation(self._h, n_components, dptr(v), dptr(out)),
              "oc_dg_project_collocation")
        return out

    def project_hydro(self, function, zb_cell2):
        """OceanDG.h:2242-2352: function(xy) -> [n,3] (free surface, u, v); returns the
        operator-order vectors (zeta, q) with q = (u, v) * depth(zeta, z_b)."""
        xy = self.quadrature_points("cell_2")
        w = function(xy)
        zeta = w[:, 0]
        h = np.maximum(zeta + zb_cell2, 1e-12)
        qv = np.stack([w[:, 1] * h, w[:, 2] * h], axis=1)
        zh = self.project_collocation(zeta, 1)
        zq = self.project_collocation(qv, 2)
        return self.pack(zh, 1), self.pack(zq, 2)

    # ------------------------------------------------------------- the stage -------
    def perform_stage_hydro(self, cur_stage, cur_time, factor_residual, current_ri,
                            vec_ki_height, vec_ki_discharge, solution_height,
                            solution_discharge, next_ri_height, next_ri_discharge,
                            write_scratch=False):
        """OceanDG.h:182-194.  current_ri = [zeta_r, q_r, tracer, velocity]."""
        n_ki = len(vec_ki_height)
        if len(vec_ki_discharge) != n_ki:
            raise ValueError("vec_ki_height and vec_ki_discharge differ in length")
        fac = np.ascontiguousarray(np.asarray(factor_residual, np.float64)[:cur_stage + 1])
        kh = (C.c_void_p * n_ki)(*[k.data_ptr() for k in vec_ki_height])
        kq = (C.c_void_p * n_ki)(*[k.data_ptr() for k in vec_ki_discharge])
        vel = current_ri[-1] if len(current_ri) > 2 else None
        check(lib().oc_dg_perform_stage_hydro(
            self._h, cur_stage, cur_time, dptr(fac), cur_stage + 1, _vp(current_ri[0]),
            _vp(current_ri[1]), _vp(vel), kh, kq, n_ki, _vp(solution_height),
            _vp(solution_discharge), _vp(next_ri_height), _vp(next_ri_discharge),
            int(write_scratch), _stream()), "oc_dg_perform_stage_hydro")

    def perform_stage_hydro_ark(self, cur_stage, cur_time, factor_residual,
                                factor_tilde_residual, factor_matrix, current_ri,
                                vec_ki_height, vec_ki_discharge, solution_height,
                                solution_discharge, next_ri_height, next_ri_discharge):
        """The additive (IMEX) stage, OceanDG.h:196-208 with body 1936-2119.
        factor_matrix is the operator member the integrator sets before every stage
        (d~_stage dt, AdditiveRungeKuttaIntegrator.h:239, 291); vec_ki_height holds
        n_stages + 1 vectors, vec_ki_discharge 2 n_stages (explicit and implicit parts)."""
        n_h, n_q = len(vec_ki_height), len(vec_ki_discharge)
        fac = np.ascontiguousarray(np.asarray(factor_residual, np.float64)[:cur_stage + 1])
        fat = np.ascontiguousarray(np.asarray(factor_tilde_residual,
                                              np.float64)[:cur_stage + 1])
        kh = (C.c_void_p * n_h)(*[k.data_ptr() if k is not None else None
                                  for k in vec_ki_height])
        kq = (C.c_void_p * n_q)(*[k.data_ptr() if k is not None else None
                                  for k in vec_ki_discharge])
        vel = current_ri[-1] if len(current_ri) > 2 else None
        check(lib().oc_dg_perform_stage_hydro_ark(
            self._h, cur_stage, cur_time, dptr(fac), dptr(fat), cur_stage + 1,
            float(factor_matrix), _vp(current_ri[0]), _vp(current_ri[1]), _vp(vel), kh, n_h,
            kq, n_q, _vp(solution_height), _vp(solution_discharge), _vp(next_ri_height),
            _vp(next_ri_discharge), _stream()), "oc_dg_perform_stage_hydro_ark")

    def apply(self, current_time, src, dst):
        """dst = M^-1 L(src); src/dst = [zeta, q] (+ velocity in src when viscous)."""
        vel = src[2] if len(src) > 2 else None
        check(lib().oc_dg_apply(self._h, current_time, _vp(src[0]), _vp(src[1]), _vp(vel),
                                _vp(dst[0]), _vp(dst[1]), _stream()), "oc_dg_apply")

    def rhs(self, src, dst, current_time=0.0):
        """dst = L(src) without inverse mass (the two MatrixFree::loop calls)."""
        vel = src[2] if len(src) > 2 else None
        check(lib().oc_dg_rhs(self._h, current_time, _vp(src[0]), _vp(src[1]), _vp(vel),
                              _vp(dst[0]), _vp(dst[1]), _stream()), "oc_dg_rhs")

    def compute_errors(self, zeta, q, exact, time):
        """OceanDG.h:2382-2459 per field: {depth, discharge, hu, hv} L2 and Linf errors of
        operator-order vectors against exact(xy, t) -> [n,3] (free surface, u, v)."""
        zh = np.ascontiguousarray(self.unpack(zeta, 1).cpu().numpy())
        qh = np.ascontiguousarray(self.unpack(q, 2).cpu().numpy())
        ex = np.ascontiguousarray(exact(self.quadrature_points("cell_0"), time))
        err = np.zeros(7)
        check(lib().oc_dg_compute_errors(self._h, dptr(zh), dptr(qh), dptr(ex), dptr(err)),
              "oc_dg_compute_errors")
        return {"l2_depth": err[0], "l2_discharge": err[1], "l2_hu": err[2], "l2_hv": err[3],
                "linf_depth": err[4], "linf_hu": err[5], "linf_hv": err[6]}

    def velocity(self, zeta, q, vel=None):
        """evaluate_velocity_field (OceanDG.h:2600-2662) into vel (operator order, 2 comps)."""
        vel = self.initialize_vector(1) if vel is None else vel
        check(lib().oc_dg_velocity(self._h, _vp(zeta), _vp(q), _vp(vel), _stream()),
              "oc_dg_velocity")
        return vel

    def transport_speed(self, zeta, q):
        """compute_cell_transport_speed (OceanDG.h:2506-2580); synchronises."""
        v = C.c_double()
        check(lib().oc_dg_transport_speed(self._h, _vp(zeta), _vp(q), C.byref(v), _stream()),
              "oc_dg_transport_speed")
        return v.value

    def volume(self, zeta):
        """check_mass cell integral (OceanDG.h:2138-2184): integral of the depth."""
        v = C.c_double()
        check(lib().oc_dg_volume(self._h, _vp(zeta), C.byref(v), _stream()), "oc_dg_volume")
        return v.value

    def sum(self, vec, n_components):
        """sum of the active entries of an operator-order vector; synchronises."""
        v = C.c_double()
        check(lib().oc_dg_sum(self._h, _vp(vec), n_components, C.byref(v), _stream()),
              "oc_dg_sum")
        return v.value

    # ------------------------------------------------------ learned closure -------
    def set_subgrid_stats(self, stats, active=None):
        """[n_cells, 3] sub-grid (mean, variance, skewness) of the bed, cell-major;
        active [n_cells] (None = all): the cells a closure network corrects"""
        v = np.ascontiguousarray(stats, dtype=np.float64)
        a = None if active is None else np.ascontiguousarray(active, dtype=np.int32)
        check(lib().oc_dg_set_subgrid_stats(self._h, dptr(v),
                                            None if a is None else _capi.iptr(a)),
              "oc_dg_set_subgrid_stats")

    def set_closure(self, network):
        """attach a ClosureNetwork (None detaches); the operator keeps a reference.  A
        network whose output gain is 0 is the plain model: nothing is attached."""
        if network is not None and network.output_gain == 0:
            network = None
        self._closure = network
        check(lib().oc_dg_set_closure(self._h, network.handle if network else None),
              "oc_dg_set_closure")

    def closure_features(self, zeta, q):
        """network inputs [n_cells, 3 nd + 3] (float32, device, cell-major rows)"""
        x = torch.empty((self.n_cells, 3 * self.dofs_per_cell + 3), dtype=torch.float32,
                        device=self.device)
        check(lib().oc_dg_closure_features(self._h, _vp(zeta), _vp(q), _vp(x), _stream()),
              "oc_dg_closure_features")
        return x

    def closure_source(self, zeta, q):
        """the attached network's correction of a state (operator order, 2 components)"""
        tau = self.initialize_vector(1)
        check(lib().oc_dg_closure_source(self._h, _vp(zeta), _vp(q), _vp(tau), _stream()),
              "oc_dg_closure_source")
        return tau

    def count_nonfinite(self, vec):
        n = np.zeros(1, np.int64)
        check(lib().oc_dg_count_nonfinite(self._h, _vp(vec), vec.numel(),
                                          n.ctypes.data_as(_capi.PI64), _stream()),
              "oc_dg_count_nonfinite")
        return int(n[0])

    KERNELS = ("stage", "height_rhs", "discharge", "rk_update", "height_jacobi")

    def set_timing(self, enable):
        check(lib().oc_dg_set_timing(self._h, int(enable)), "oc_dg_set_timing")

    def timing_report(self):
        """{kernel: (total ms, launches)} since the last report (events on the stream)."""
        ms = np.zeros(5)
        cnt = np.zeros(5, np.int64)
        check(lib().oc_dg_timing_report(self._h, dptr(ms), cnt.ctypes.data_as(_capi.PI64)),
              "oc_dg_timing_report")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.KERNELS)}

    def last_sweeps(self):
        n = np.zeros(1, np.int64)
        check(lib().oc_dg_last_sweeps(self._h, n.ctypes.data_as(_capi.PI64), _stream()),
              "oc_dg_last_sweeps")
        return int(n[0])

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().oc_dg_destroy(self._h)
                self._h = None
        except Exception:
            pass
