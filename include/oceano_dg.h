/* oceano_dg.h -- C ABI of the MI355X shallow-water DG stage operator (liboceano_dg.so).
 *
 * Plain pointers and sizes only: no HIP, torch or deal.II types cross this boundary.
 * Device buffers are hipMalloc'ed by the caller; `stream` is a hipStream_t passed as void*
 * (NULL = the null stream).  Every call returns an int status (OC_OK = 0); the message of
 * the last failure on the calling thread is oc_last_error().  No exception crosses the ABI.
 * A handle is not thread-safe and belongs to the GPU that was current at oc_dg_create.
 *
 * Vector storage.  The operator owns the storage order of its DoF vectors, as deal.II's
 * MatrixFree::initialize_dof_vector does (OceanDG.h:498-504): cells are grouped in
 * batches of OC_CELL_BATCH consecutive cells and, inside a batch, the entry of
 * (component, node) for the batch's lane-th cell sits at
 *   ((batch * n_components + component) * dofs_per_cell + node) * OC_CELL_BATCH + lane.
 * `node` is lexicographic over the (p+1)^2 Gauss-Lobatto nodes (x fastest), the
 * FE_DGQ(p) numbering.  oc_dg_pack / oc_dg_unpack convert from / to the cell-major
 * order (cell, component, node) used by deal.II's FESystem(FE_DGQ(p), c) DoF numbering.
 * The time integrator only copies and combines whole vectors, so it never depends on it.
 */
#ifndef OCEANO_DG_H
#define OCEANO_DG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OC_CELL_BATCH 64

/* status codes */
#define OC_OK 0
#define OC_ERR_ARGUMENT 1      /* bad argument (null pointer, size, degree, stage index)    */
#define OC_ERR_MESH 2          /* inconsistent topology (face not covered, 2:1 violated)   */
#define OC_ERR_BOUNDARY 3      /* boundary id without a condition: OceanDG.h:1199-1203     */
#define OC_ERR_HIP 4           /* a HIP runtime call failed                                 */
#define OC_ERR_NONFINITE 5     /* non-finite entry found by oc_dg_count_nonfinite callers   */
#define OC_ERR_UNSUPPORTED 6   /* configuration the kernels do not implement                */

const char *oc_last_error(void);

/* Library tuning options read by later oc_dg_create calls (no reference counterpart: the
 * reference leaves cell batching to deal.II's MatrixFree).
 *   "waves_per_block" 1 (default), 2 or 4: 64-cell work items per workgroup of the
 *                 unpaired one-pass stage kernel (with 2 or 4, the height and discharge
 *                 items of 1 or 2 batches, each on an LDS slice of its own size).  Results
 *                 do not depend on it.
 *   "paired_items" -1 (default: 1 at degree <= 3, 0 at degree 4), 0 or 1: the
 *                 inviscid one-pass stage runs each batch's height and discharge items as
 *                 one two-wave workgroup that splits the faces between them (each face
 *                 point's fluxes formed once per cell, the other equation's integrals
 *                 handed over through LDS).  On hanging faces the face integrals are summed
 *                 in another order, so results change at rounding level.
 *   "slot_chunk"  cells are grouped by face signature (which faces are the coarse side of
 *                 two subfaces or on the boundary) within chunks of this many 64-cell
 *                 batches of the caller's order; 0 (default) = one XCD's share of the
 *                 batches for OC_PRECISION_FP64 and the caller's order for the mixed
 *                 path; < 0 = keep the caller's order.  Results do not depend on it.
 *   "xcd_swizzle" 1 (default): workgroups are dealt so that each of the 8 XCDs runs a
 *                 contiguous range of batches (neighbour traces stay in its L2); 0: the
 *                 hardware's round-robin order.  Results do not depend on it. */
int oc_set_option(const char *name, int64_t value);

/* ------------------------------------------------------------------------------------
 * Boundary-condition kinds.  Replaces ICBC::BcBase's six boundary maps
 * (IcbcBase.h:65-77): each boundary id carries exactly one kind (IcbcBase.h:146-260).
 * ---------------------------------------------------------------------------------- */
#define OC_BC_WALL 1
#define OC_BC_SUPERCRITICAL_INFLOW 2
#define OC_BC_SUPERCRITICAL_OUTFLOW 3
#define OC_BC_HEIGHT_INFLOW 4
#define OC_BC_DISCHARGE_INFLOW 5
#define OC_BC_ABSORBING_OUTFLOW 6

/* ------------------------------------------------------------------------------------
 * Mesh builder (setup path).  Replaces GridIn::read_msh + Triangulation::refine_global
 * + the MatrixFree face setup the reference reaches through main.cc:617-663 and
 * OceanDG.h:438-470, for the axis-aligned quadrilateral meshes of the test cases.
 * ---------------------------------------------------------------------------------- */
typedef struct oc_mesh oc_mesh;

/* nx x ny cells on [x0,x1]x[y0,y1]; boundary_id = {x0 side, x1 side, y0 side, y1 side}
 * (deal.II face numbers 0..3); a periodic direction ignores its two ids. */
int oc_mesh_create_rectangle(double x0, double x1, double y0, double y1, int32_t nx,
                             int32_t ny, int32_t periodic_x, int32_t periodic_y,
                             const int32_t boundary_id[4], oc_mesh **mesh);
/* Gmsh 4.1 ASCII file (GridIn::read_msh, main.cc:621-637) whose quadrilaterals tile a
 * uniform axis-aligned grid, e.g. the reference decks' shallowWaterVortex_20x10cells.msh;
 * boundary ids from the physical tags of the boundary curves, one per side. */
int oc_mesh_read_msh(const char *path, int32_t periodic_x, int32_t periodic_y,
                     oc_mesh **mesh);
/* split every active cell into four (Triangulation::refine_global) */
int oc_mesh_refine_global(oc_mesh *mesh, int32_t times);
/* split round(fraction * n_active) active cells chosen by a Fisher-Yates shuffle driven
 * by std::mt19937_64(seed) with index = engine() % (i + 1); fails if 2:1 is violated */
int oc_mesh_refine_fraction(oc_mesh *mesh, double fraction, uint64_t seed);
/* info = {n_cells, n_inner_faces, n_boundary_faces, n_hanging_subfaces} */
int oc_mesh_info(const oc_mesh *mesh, int64_t info[4]);
/* cell_geom [n_cells][4] = x0, y0, hx, hy; cell_level [n_cells];
 * inner [n_inner][6] = cell_m, face_m, subface_m, cell_p, face_p, subface_p
 *   (subface = -1 for a whole face, 0/1 = lower/upper half of a coarse face);
 * boundary [n_bdry][3] = cell, face, boundary_id.  Any pointer may be NULL. */
int oc_mesh_export(const oc_mesh *mesh, double *cell_geom, int32_t *cell_level,
                   int32_t *inner, int32_t *boundary);
void oc_mesh_destroy(oc_mesh *mesh);

/* ------------------------------------------------------------------------------------
 * Stage operator.  Replaces SpaceDiscretization::OceanoOperator (OceanDG.h:157-380).
 * ---------------------------------------------------------------------------------- */
typedef struct oc_mlp oc_mlp; /* oceano_mlp.h */
typedef struct oc_dg oc_dg;

typedef struct
{
  int32_t n_cells;
  const double *cell_geom;   /* [n_cells][4] x0, y0, hx, hy (axis-aligned cells)         */
  int32_t n_inner;
  const int32_t *inner;      /* [n_inner][6] as oc_mesh_export                            */
  int32_t n_boundary;
  const int32_t *boundary;   /* [n_boundary][3] as oc_mesh_export                         */
} oc_mesh_desc;

#define OC_PRECISION_FP64 0  /* plain FP64: the reference arithmetic                       */
#define OC_PRECISION_MIXED 1 /* FP32 cell/face integrals, FP64 state, depth and bathymetry */

typedef struct
{
  int32_t degree;               /* fe_degree, 1..4 (main.cc:213)                           */
  double g;                     /* "g" (ShallowWater.h:294)                                */
  double threshold_wetdry;      /* "Threshold_for_wetdry", cu (ShallowWater.h:295-296)     */
  double viscosity;             /* "horizontal_viscosity" (ViscosityCoefficient.h:68)      */
  int32_t max_iteration_height; /* 15 in the reference (main.cc:226)                       */
  int32_t precision;            /* OC_PRECISION_*                                          */
  int32_t n_boundary_kinds;     /* entries of the two arrays below                         */
  const int32_t *boundary_ids;  /* boundary ids that carry a condition                     */
  const int32_t *boundary_kinds;/* OC_BC_* per id (IcbcBase.h:146-260)                     */
} oc_params;

/* Replaces OceanoOperator::OceanoOperator + reinit (OceanDG.h:415-470). */
int oc_dg_create(const oc_mesh_desc *mesh, const oc_params *params, oc_dg **op);
int oc_dg_destroy(oc_dg *op);

/* Static data at quadrature points, the five CellDataStorage stores of
 * OceanoOperator::initialize_data_at_quadrature (OceanDG.h:498-569):
 *   OC_STORE_CELL_GAUSS   [n_cells][nq*nq]   points; 5 values each: z_b, friction,
 *                                             tau_x, tau_y, coriolis  (problem_data)
 *   OC_STORE_CELL_LOBATTO [n_cells][ng*ng]   points; 1 value: z_b
 *   OC_STORE_CELL_COLLOC  [n_cells][n1*n1]   points; 2 values: z_b, friction
 *   OC_STORE_FACE         [n_inner][ng][2]   points x -+ 1e-12 n (minus, plus side); 1 value
 *   OC_STORE_BOUNDARY     [n_bdry][nq]       points; 1 value: z_b
 * with nq = floor(1.5p)+1, ng = p+2, n1 = p+1 and points lexicographic (x fastest).
 * oc_dg_quadrature_points writes 2 doubles (x, y) per point to host memory; the caller
 * evaluates its functions there and hands the values to oc_dg_set_static_data (host).
 * A store with no points (no boundary faces, no inner faces) accepts null pointers. */
#define OC_STORE_CELL_GAUSS 0
#define OC_STORE_CELL_LOBATTO 1
#define OC_STORE_CELL_COLLOC 2
#define OC_STORE_FACE 3
#define OC_STORE_BOUNDARY 4
int oc_dg_n_points(const oc_dg *op, int32_t store, int64_t *n_points);
int oc_dg_quadrature_points(const oc_dg *op, int32_t store, double *xy);
int oc_dg_set_static_data(oc_dg *op, int32_t store, const double *values);

/* Boundary data at the boundary quadrature points (OC_STORE_BOUNDARY points),
 * [n_bdry][nq][3] per time: the values of the boundary-condition Functions, which the
 * reference evaluates per quadrature point (OceanDG.h:1146-1184) after setting their time to
 * the stage time (OceanDG.h:1798-1803; apply: 1700-1705).  Component 0 = free surface,
 * 1..2 = velocity (supercritical inflow, absorbing outflow); component 1 = the discharge
 * magnitude for OC_BC_DISCHARGE_INFLOW.  Host memory; both calls synchronise the device
 * (a stage still running may read the previous data).
 *   oc_dg_set_boundary_values  time-independent data: one entry used at every time.
 *   oc_dg_set_boundary_table   values at n_times times (finite, non-decreasing; repeated
 *                              times allowed), [n_times][n_bdry][nq][3].  Every call that
 *                              takes a `time` (stage, additive stage, apply, rhs) uses the
 *                              entry whose time equals it to within 1e-12 max(1, |t|), the
 *                              first such entry, and fails with OC_ERR_ARGUMENT when there is
 *                              none.  A time loop replayed from a graph (oc_rk_replay) reads
 *                              entries e0 + r S + s for step r and stage s, e0 the entry at
 *                              the loop's current time: tabulate the stage times
 *                              t + r dt + c_s dt in that order.
 * Operators whose boundary faces are all walls or supercritical outflows read no data, and
 * their stages accept any time.  The reference sets the time of the inflow Functions only
 * (supercritical, height, discharge inflow); absorbing-outflow data stay at the time they
 * had, so a caller that reproduces it tabulates them at that fixed time. */
int oc_dg_set_boundary_values(oc_dg *op, const double *values);
int oc_dg_set_boundary_table(oc_dg *op, int32_t n_times, const double *times,
                             const double *values);

/* Geometry queries (host memory): cells [n_cells][4], degree and sizes {n_cells,
 * n_cells_padded, dofs_per_cell, nq, ng, n_inner, n_bdry}. */
int oc_dg_sizes(const oc_dg *op, int64_t sizes[7]);
/* number of doubles of an operator vector with n_components components (padded) */
int oc_dg_vector_length(const oc_dg *op, int32_t n_components, int64_t *length);
/* device-to-device conversions between cell-major order and the operator order */
int oc_dg_pack(oc_dg *op, int32_t n_components, const double *cell_major, double *op_order,
               void *stream);
int oc_dg_unpack(oc_dg *op, int32_t n_components, const double *op_order, double *cell_major,
                 void *stream);

/* One explicit Runge-Kutta stage of the hydrodynamic operator.  Replaces
 * OceanoOperator::perform_stage_hydro (OceanDG.h:182-194, body 1782-1912) as called by
 * ExplicitRungeKuttaIntegrator::perform_time_step (ExplicitRungeKuttaIntegrator.h:212-271).
 *   factor[0..stage]   dt-scaled Butcher row (b for the last stage), n_factor = stage+1
 *   ri_h, ri_q, ri_vel current_ri = {zeta_r, q_r, (tracer), velocity}
 *   ki_h[0..n_ki-1]    vec_ki_height: [0] scratch (front), [stage+1] <- L_zeta (not inverted)
 *   ki_q[0..n_ki-1]    vec_ki_discharge: [0] scratch (L_q), [stage+1] <- M^-1 L_q
 *   w_h, w_q           solution vectors; updated in place at the last stage
 *   r_h, r_q           next_ri vectors; written at the other stages
 *   time               the stage time (current_time + c_s dt): selects the boundary data
 *                      (oc_dg_set_boundary_table), OC_ERR_ARGUMENT if none is tabulated
 * ki_h / ki_q are host arrays of n_ki device pointers; n_ki = n_stages + 1 and the last
 * stage is stage == n_ki - 2 (OceanDG.h:1827).  The scratch entries ki_h[0], ki_q[0] are
 * written only if `write_scratch` is nonzero (the time integrator never reads them).
 * ri_vel is read only when the viscosity is nonzero (it multiplies the only use).
 * OC_PRECISION_MIXED: the k vectors ki_h[1..], ki_q[1..] are operator-internal and hold
 * FP32 values in the first half of each vector (the FP32 integrals' values, stored
 * exactly); every combination with them stays FP64.  The scratch entries stay FP64. */
int oc_dg_perform_stage_hydro(oc_dg *op, uint32_t stage, double time, const double *factor,
                              uint32_t n_factor, const double *ri_h, const double *ri_q,
                              const double *ri_vel, double *const *ki_h, double *const *ki_q,
                              uint32_t n_ki, double *w_h, double *w_q, double *r_h,
                              double *r_q, int32_t write_scratch, void *stream);

/* oc_dg_perform_stage_hydro for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments, results and errors as oc_dg_perform_stage_hydro; it calls no
 * allocating routine.  The stage kernels' per-batch scratch (Jacobi sweep counts, the
 * mixed path's deferred depth-solve masks) lives in one static device array per GPU, sized
 * for OC_NOALLOC_MAX_BATCHES batches of OC_CELL_BATCH cells; an operator with more
 * batches gets OC_ERR_ARGUMENT.  Every call on a GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads, other operators) race.  Every entry of the
 * array a stage reads is written earlier in the same stage, so nothing needs clearing.
 * oc_dg_last_sweeps does not see these calls (their counts stay in the shared array), and
 * per-kernel timing (oc_dg_set_timing) records nothing for them. */
#define OC_NOALLOC_MAX_BATCHES 16384
int oc_dg_perform_stage_hydro_noalloc(oc_dg *op, uint32_t stage, double time,
                                      const double *factor, uint32_t n_factor,
                                      const double *ri_h, const double *ri_q,
                                      const double *ri_vel, double *const *ki_h,
                                      double *const *ki_q, uint32_t n_ki, double *w_h,
                                      double *w_q, double *r_h, double *r_q,
                                      int32_t write_scratch, void *stream);

/* The explicit stage of n_ops operators in one launch (the batched ensemble: instances
 * grouped into one operator per degree, every group advanced by a single grid).  No
 * reference counterpart (oceano runs one instance per process).  Entry g of each array is
 * operator g's argument of oc_dg_perform_stage_hydro (its time, its dt-scaled factor row,
 * its vectors); n_factor, n_ki and stage are shared.  Results equal n_ops separate
 * oc_dg_perform_stage_hydro calls.  The
 * operators must share precision and device, be inviscid, have no closure attached and
 * write outputs distinct from their inputs (OC_ERR_UNSUPPORTED otherwise: call the
 * single-operator entry per operator).  At degree 4 in OC_PRECISION_MIXED the FP64 depth
 * solve of near-floor cells follows as a small launch per such operator. */
#define OC_MULTI_MAX_OPERATORS 4
int oc_dg_perform_stage_hydro_multi(int32_t n_ops, oc_dg *const *ops, uint32_t stage,
                                    const double *time, const double *const *factor,
                                    uint32_t n_factor,
                                    const double *const *ri_h, const double *const *ri_q,
                                    double *const *const *ki_h, double *const *const *ki_q,
                                    uint32_t n_ki, double *const *w_h, double *const *w_q,
                                    double *const *r_h, double *const *r_q, void *stream);

/* One additive (IMEX) Runge-Kutta stage: the stiff Manning friction is implicit, the rest
 * of the operator explicit.  Replaces the TIMEINTEGRATOR_ADDITIVERUNGEKUTTA overload of
 * OceanoOperator::perform_stage_hydro (OceanDG.h:196-208, body 1936-2119) as called by
 * AdditiveRungeKuttaIntegrator::perform_time_step (AdditiveRungeKuttaIntegrator.h:239-316).
 *   factor[0..stage]        dt-scaled explicit row a_i (b_i at the last stage)
 *   factor_tilde[0..stage]  dt-scaled implicit row a~_i (ignored at the last stage)
 *   factor_matrix           d~_stage dt: the implicit diagonal (pde_operator.factor_matrix)
 *   ki_h[0..n_ki_h-1]       n_ki_h = n_stages + 1: [stage+1] <- L_zeta
 *   ki_q[0..n_ki_q-1]       n_ki_q = 2 n_stages: [2 stage+1] <- explicit L_q (without the
 *                           friction source before the last stage), [2 stage+2] <- the
 *                           friction source tested against the basis (not at the last stage)
 *   r_h, r_q                next_ri (written before the last stage; must differ from ri_h,
 *                           ri_q: the stage reads the input after writing r_h)
 * The last stage updates w_h, w_q in place and writes r_q = M^-1 (b . kex + b . kim).
 * `time` selects the boundary data as in oc_dg_perform_stage_hydro.  ri_vel is read only
 * when the viscosity is nonzero.  OC_PRECISION_MIXED: ki_h[1..]
 * and ki_q[1..] hold FP32 values in their first half as in the explicit stage (the explicit
 * and stiff momentum terms are FP32 integrals, so the storage is exact); the update that
 * combines them is FP64. */
int oc_dg_perform_stage_hydro_ark(oc_dg *op, uint32_t stage, double time,
                                  const double *factor, const double *factor_tilde,
                                  uint32_t n_factor, double factor_matrix, const double *ri_h,
                                  const double *ri_q, const double *ri_vel,
                                  double *const *ki_h, uint32_t n_ki_h, double *const *ki_q,
                                  uint32_t n_ki_q, double *w_h, double *w_q, double *r_h,
                                  double *r_q, void *stream);

/* Learned sub-grid closure (no counterpart in the reference, which has no closure; the
 * correction is a momentum source, the slot of ShallowWater.h:425-429 used at
 * OceanDG.h:734-736).  Degrees 1 and 2.
 *   oc_dg_set_subgrid_stats  host [n_cells][3]: mean, variance, skewness of the sub-grid
 *                            bed per cell (caller's cell order); `active` (host
 *                            [n_cells], null = all) selects the cells the network corrects,
 *                            the others keep the plain model; also allocates the
 *                            correction buffer
 *   oc_dg_set_closure        attach a network (oceano_mlp.h) with 3 nd + 3 inputs (modal
 *                            coefficients of zeta, hu, hv in the orthonormal tensor
 *                            Legendre basis of the cell, then the three statistics) and
 *                            2 nd outputs (modal coefficients of the correction of hu, hv);
 *                            null detaches.  The handle is not owned and must outlive its
 *                            use.  With a network attached every explicit stage
 *                            (oc_dg_perform_stage_hydro, oc_rk_*) adds the correction of
 *                            its input state to M^-1 L_q; apply, rhs and the additive
 *                            stage do not.  A cell whose output is not finite gets no
 *                            correction (the plain sub-grid model).
 *   oc_dg_closure_features   the network inputs of every cell, device [n_cells][3 nd + 3]
 *                            FP32, caller's cell order (training data)
 *   oc_dg_closure_source     the correction of a state, device vector in operator order
 *                            (2 components) */
int oc_dg_set_subgrid_stats(oc_dg *op, const double *stats, const int32_t *active);
int oc_dg_set_closure(oc_dg *op, const oc_mlp *mlp);
int oc_dg_closure_features(oc_dg *op, const double *zeta, const double *q, float *features,
                           void *stream);
int oc_dg_closure_source(oc_dg *op, const double *zeta, const double *q, double *tau,
                         void *stream);

/* dst = M^-1 L(src): step-67's OceanoOperator::apply (declared OceanDG.h:178-180; its
 * body in the reference is dead code).  Both fields use the exact tensor-product inverse
 * of the mass matrix (affine cells); L is the spatial operator of perform_stage_hydro with
 * the boundary data at `time` (OceanDG.h:1700-1705). */
int oc_dg_apply(oc_dg *op, double time, const double *src_h, const double *src_q,
                const double *src_vel, double *dst_h, double *dst_q, void *stream);

/* L_zeta and L_q without inverse mass (the two MatrixFree::loop calls, OceanDG.h:1805-1823),
 * boundary data at `time` */
int oc_dg_rhs(oc_dg *op, double time, const double *src_h, const double *src_q,
              const double *src_vel, double *dst_h, double *dst_q, void *stream);

/* Nodal coefficients from values at the OC_STORE_CELL_COLLOC points (Gauss(p+1)^2 per
 * cell): transform_from_q_points_to_basis of deal.II's CellwiseInverseMassMatrix, as used
 * by OceanoOperator::project_hydro (OceanDG.h:2242-2352) and evaluate_velocity_field
 * (OceanDG.h:2600-2662).  values [n_cells][n1*n1][n_comp] -> coeffs [n_cells][n_comp][n1*n1]
 * (cell-major).  Host memory, no GPU work. */
int oc_dg_project_collocation(const oc_dg *op, int32_t n_comp, const double *values,
                              double *coeffs);

/* Errors against an exact solution: compute_errors_hydro (OceanDG.h:2382-2459) with the
 * fields reported separately.  zeta [n_cells][n1^2] and q [n_cells][2][n1^2] cell-major,
 * exact [n_cells][nq^2][3] = (free surface, u, v) at the OC_STORE_CELL_GAUSS points.
 * err[0] = L2 of the depth error, err[1] = L2 of the discharge error (hu, hv together,
 * as the reference), err[2], err[3] = L2 of hu, hv, err[4..6] = Linf of depth, hu, hv.
 * Depths use the z_b of OC_STORE_CELL_GAUSS.  Host memory, no GPU work. */
int oc_dg_compute_errors(const oc_dg *op, const double *zeta, const double *q,
                         const double *exact, double err[7]);

/* ------------------------------------------------------------------------------------
 * Work around the stage that the time loop needs (main.cc:1546-1577), on the GPU.
 *
 * Velocity field for the viscous term and output: evaluate_velocity_field
 * (OceanDG.h:2600-2662) -- velocity at the Gauss(p+1) points (z_b of
 * OC_STORE_CELL_COLLOC), collocated back to the nodal basis.  vel: operator order, 2
 * components.  Asynchronous on `stream`. */
int oc_dg_velocity(oc_dg *op, const double *zeta, const double *q, double *vel, void *stream);

/* Largest transport speed over the mesh: compute_cell_transport_speed
 * (OceanDG.h:2506-2580), which the time loop turns into the step every five steps
 * (main.cc:1550-1555).  Synchronises the stream; *speed on the host. */
int oc_dg_transport_speed(oc_dg *op, const double *zeta, const double *q, double *speed,
                          void *stream);

/* Mass-conservation check (OceanDG.h:2138-2184): *volume = integral of the depth
 * max(zeta + z_b, 1e-12) at the GL(p+2) points (z_b of OC_STORE_CELL_LOBATTO); *sum =
 * sum of all entries of an operator-order vector (the boundary part of the check is the
 * sum of ki_h[s+1] weighted by the step's b_s dt: inner faces and cell terms of L_zeta sum
 * to zero).  Both synchronise the stream. */
int oc_dg_volume(oc_dg *op, const double *zeta, double *volume, void *stream);
int oc_dg_sum(oc_dg *op, const double *x, int32_t n_components, double *sum, void *stream);

/* ------------------------------------------------------------------------------------
 * Device-resident explicit RK time loop (ExplicitRungeKuttaIntegrator.h:179-298 driven as
 * in main.cc:1546-1577).  The integrator owns S+1 k registers per field and two r
 * registers that alternate between stages (the reference copies r into vec_ri instead;
 * the results are the same).  rows: S rows packed one after the other, row s holding
 * s+1 coefficients -- the Butcher row a[s+1][0..s] for s < S-1 and b for s = S-1
 * (unscaled; the step multiplies by dt).  A step is the S stages, then, when `velocity`
 * is given (viscosity != 0), the velocity update of the new state.  The loop keeps its
 * time t (oc_rk_set_time, initially 0): stage s of a step runs at t + c_s dt with c_s the
 * row sum of a[s] (ExplicitRungeKuttaIntegrator.h:212-271), and every executed step
 * advances t by dt.  oc_rk_capture records one step with fixed arguments into a HIP graph
 * (on a non-null stream) without running it; oc_rk_replay launches it n times (n times
 * the recorded steps).  With
 * time-dependent boundary data (oc_dg_set_boundary_table) the recorded stages read the
 * table through a device-side entry counter that oc_rk_replay sets from t and the graph
 * advances; the replay checks on the host that the table lists the n steps' stage times.
 * A graph holds the operator's static data, boundary table and closure as they were at the
 * capture: oc_rk_replay fails with OC_ERR_ARGUMENT after any oc_dg_set_static_data,
 * oc_dg_set_boundary_values / _table, oc_dg_set_subgrid_stats or oc_dg_set_closure call on
 * the operator (capture again). */
typedef struct oc_rk oc_rk;
int oc_rk_create(oc_dg *op, int32_t n_stages, const double *rows, oc_rk **out);
void oc_rk_destroy(oc_rk *rk);
int oc_rk_set_time(oc_rk *rk, double t);
int oc_rk_time(const oc_rk *rk, double *t);
int oc_rk_step(oc_rk *rk, double dt, double *w_h, double *w_q, double *velocity, void *stream);
int oc_rk_capture(oc_rk *rk, double dt, double *w_h, double *w_q, double *velocity,
                  void *stream);
/* n_steps consecutive steps in one graph (one launch per replay runs them all: the host
 * launch and the gap between graph launches are paid once per n_steps steps) */
int oc_rk_capture_steps(oc_rk *rk, int32_t n_steps, double dt, double *w_h, double *w_q,
                        double *velocity, void *stream);
int oc_rk_replay(oc_rk *rk, int32_t n, void *stream);
/* A time loop over n_ops operators advanced together, one launch per stage
 * (oc_dg_perform_stage_hydro_multi; the batched ensemble).  Registers and time per
 * operator; dt, w_h, w_q are arrays of n_ops entries (each operator keeps its own step, e.g.
 * the CFL step of its degree).  oc_rk_replay replays a captured multi-operator graph;
 * oc_rk_registers reports the first operator's single-loop registers, not these, and
 * oc_rk_time the first operator's time.  Time-dependent boundary tables are not supported
 * in a multi-operator graph. */
int oc_rk_create_multi(int32_t n_ops, oc_dg *const *ops, int32_t n_stages, const double *rows,
                       oc_rk **out);
int oc_rk_step_multi(oc_rk *rk, const double *dt, double *const *w_h, double *const *w_q,
                     void *stream);
int oc_rk_capture_multi(oc_rk *rk, int32_t n_steps, const double *dt, double *const *w_h,
                        double *const *w_q, void *stream);
/* device pointers of the integrator's registers (arrays of S+1 and 2 entries; any may be
 * null).  On an OC_PRECISION_MIXED operator the k registers hold FP32 numbers (as every
 * mixed k vector does), and with paired work items at degree <= 3 the r registers hold the
 * intermediate stage inputs in centred FP32 form: the values as floats in the first half of
 * each register (zeta minus its cell's DoF-0 value, q as is), then, in the zeta register,
 * each cell's DoF-0 value as a double (the FP32 stage reads nothing else of them). */
int oc_rk_registers(oc_rk *rk, double **ki_h, double **ki_q, double **r_h, double **r_q);

/* Count non-finite entries of a device vector of `n` doubles into *count (host).  The
 * reference has no such check (main.cc:1527-1529 only mentions blow-up); hosts use it
 * to turn a NaN state into OC_ERR_NONFINITE. Synchronises the stream. */
int oc_dg_count_nonfinite(oc_dg *op, const double *x, int64_t n, int64_t *count, void *stream);

/* Sum of Jacobi sweeps executed by the depth-mass solve in the last stage call, over all
 * cells (the k loop of OceanDG.h:1443).  Synchronises the stream. */
int oc_dg_last_sweeps(oc_dg *op, int64_t *sweeps, void *stream);

/* Per-kernel timing with HIP events recorded on the launch stream around each kernel of
 * oc_dg_perform_stage_hydro: 0 = the one-pass stage kernel; on the split path taken when
 * outputs alias inputs, 1 = L_zeta, 2 = L_q + inverse mass, 3 = discharge RK update,
 * 4 = front + depth-mass Jacobi (also the mixed path's FP64 solve of cells near the depth
 * floor, which follows kernel 0).  Enabling creates a pool of events up front.
 * oc_dg_timing_report synchronises, returns the summed milliseconds
 * and launch counts per kernel since the last report, and clears them.  The reference
 * times the same two parts with TimerOutput sections "rk_stage hydro - integrals L_h" and
 * "rk_stage hydro - inv mass + vec upd" (OceanDG.h:1796, 1829). */
int oc_dg_set_timing(oc_dg *op, int32_t enable);
int oc_dg_timing_report(oc_dg *op, double ms[5], int64_t count[5]);

/* ------------------------------------------------------------------------------------
 * Test-case functions (ICBC layer).  Host memory; evaluate at n points xy[n][2].
 * ---------------------------------------------------------------------------------- */
/* Travelling cos^4 vortex of Icbc_ShallowWaterVortex.h:139-226 (regularity p = 2),
 * parametrised: prm = {h0, u_inf, v_inf, amplitude (h0 - h_min), radius, xc, yc, g,
 * period_x (0 = none), period_y}; out[n][3] = depth, u, v at time t. */
int oc_icbc_vortex(const double prm[10], double t, int64_t n, const double *xy, double *out);
/* Channel bed with a parabolic bump b(x) = 0.2 - 0.05 (x-10)^2 on 8 < x < 12 plus seeded
 * Gaussian sub-cell bumps; returns problem_data rows out[n][5] = {z_b = -b, manning, 0,0,0}.
 * prm = {manning, n_bumps(-1 = draw from seed), amplitude_min, amplitude_max, width_min,
 * width_max, x_min, x_max}. */
int oc_icbc_channel(const double prm[8], uint64_t seed, int64_t n, const double *xy,
                    double *out);

#ifdef __cplusplus
}
#endif

#endif
