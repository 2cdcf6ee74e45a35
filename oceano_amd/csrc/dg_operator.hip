// dg_operator.hip -- host side of liboceano_dg.so: operator set-up, static data, launches.
//
// Mirrors SpaceDiscretization::OceanoOperator (OceanDG.h:157-380): create = constructor +
// reinit + initialize_vector (438-504); set_static_data = initialize_data_at_quadrature
// (498-569) with the function evaluation done by the caller at the exported points;
// perform_stage_hydro / apply as declared at OceanDG.h:178-194.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "dg_internal.hpp"
#include "oceano_dg.h"
#include "mlp_core.hpp"
#include "rules.hpp"

namespace oceano {

thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

constexpr int mlp_inputs_max = mlp::IN;

static int bad_argument(const char *fn)
{
  set_error(std::string(fn) + ": null pointer or invalid size");
  return OC_ERR_ARGUMENT;
}

namespace {

DegreeTables make_tables(int p)
{
  DegreeTables t;
  std::memset(&t, 0, sizeof(t));
  const int n1 = p + 1, nq = nq_of(p), ng = p + 2;
  const Rule1D nodes = gauss_lobatto(n1), rq = gauss(nq), rg = gauss_lobatto(ng), rn = gauss(n1);
  for (int a = 0; a < nq; ++a)
    {
      t.wq[a] = rq.w[a];
      for (int i = 0; i < n1; ++i)
        {
          t.Sq[a][i] = lagrange_value(nodes.x, i, rq.x[a]);
          t.Dq[a][i] = lagrange_slope(nodes.x, i, rq.x[a]);
        }
    }
  for (int a = 0; a < ng; ++a)
    {
      t.wg[a] = rg.w[a];
      for (int i = 0; i < n1; ++i)
        {
          t.Sg[a][i] = lagrange_value(nodes.x, i, rg.x[a]);
          for (int s = 0; s < 2; ++s)
            t.Sgs[s][a][i] = lagrange_value(nodes.x, i, 0.5 * (rg.x[a] + s));
        }
    }
  std::vector<double> S(n1 * n1);
  for (int a = 0; a < n1; ++a)
    {
      t.wn[a] = rn.w[a];
      for (int i = 0; i < n1; ++i)
        S[a * n1 + i] = lagrange_value(nodes.x, i, rn.x[a]);
    }
  for (int a = 0; a < n1; ++a)
    for (int i = 0; i < n1; ++i)
      t.Sn[a][i] = S[a * n1 + i];
  for (int s = 0; s < 2; ++s)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n1; ++k)
        t.Psub[s][j][k] = lagrange_value(nodes.x, k, 0.5 * (nodes.x[j] + s));
  const std::vector<double> Si = small_inverse(S, n1);  // Si[i][a]
  for (int i = 0; i < n1; ++i)
    for (int a = 0; a < n1; ++a)
      t.Ti[i][a] = Si[i * n1 + a];
  for (int i = 0; i < n1; ++i)
    {
      double s = 0;
      for (int a = 0; a < ng; ++a)
        s += t.Sg[a][i] * t.wg[a];
      t.lumped[i] = s;
    }
  double smax = 0.0;
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j)
      {
        t.ilumped2[i][j] = 1.0 / (t.lumped[i] * t.lumped[j]);
        double m = 0.0;
        for (int a = 0; a < ng; ++a)
          m += t.Sg[a][i] * t.Sg[a][j] * t.wg[a];
        t.M1[i][j] = m;
      }
  for (int a = 0; a < ng; ++a)
    for (int i = 0; i < n1; ++i)
      smax = std::max(smax, std::fabs(t.Sg[a][i]));
  t.lam = smax * smax;
  return t;
}

Basis<float> to_float(const DegreeTables &t)
{
  Basis<float> f;
  const double *src = reinterpret_cast<const double *>(&t);
  float *dst = reinterpret_cast<float *>(&f);
  static_assert(sizeof(DegreeTables) == 2 * sizeof(Basis<float>), "table layout");
  for (size_t i = 0; i < sizeof(DegreeTables) / sizeof(double); ++i)
    dst[i] = float(src[i]);
  return f;
}

// K linear Jacobi sweeps on a wet unit cell from a zero increment:
// J = sum_{k<K} (I - D^-1 M)^k D^-1 with M the GL(p+2) mass (exact) and D its row sums
std::vector<float> jacobi_matrix(int p, int K)
{
  const DegreeTables t = make_tables(p);
  const int n1 = p + 1, nd = n1 * n1, ng = p + 2;
  std::vector<double> M1(n1 * n1, 0.0), M(nd * nd), D(nd), A(nd * nd), J(nd * nd, 0.0),
    P(nd * nd, 0.0);
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j)
      for (int a = 0; a < ng; ++a)
        M1[i * n1 + j] += t.Sg[a][i] * t.Sg[a][j] * t.wg[a];
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j)
      M[i * nd + j] = M1[(i % n1) * n1 + (j % n1)] * M1[(i / n1) * n1 + (j / n1)];
  for (int i = 0; i < nd; ++i)
    D[i] = t.lumped[i % n1] * t.lumped[i / n1];
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j)
      A[i * nd + j] = (i == j ? 1.0 : 0.0) - M[i * nd + j] / D[i];
  // P = D^-1, then J = sum_k A^k D^-1 by Horner: J = D^-1 + A (D^-1 + A (...))
  for (int i = 0; i < nd; ++i)
    P[i * nd + i] = 1.0 / D[i];
  for (int k = 0; k < K; ++k)
    {
      std::vector<double> N(nd * nd, 0.0);
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j)
          {
            double s = (i == j) ? 1.0 / D[i] : 0.0;
            if (k > 0)
              for (int m = 0; m < nd; ++m)
                s += A[i * nd + m] * J[m * nd + j];
            N[i * nd + j] = s;
          }
      J = N;
    }
  std::vector<float> out(nd * nd);
  for (int i = 0; i < nd * nd; ++i)
    out[i] = float(J[i]);
  return out;
}

// The momentum inverse mass of the unit cell as one matrix: the basis change of
// CellwiseInverseMassMatrix (OceanDG.h:1541-1553) from GL(p+1) nodal values to Gauss(p+1)
// collocation, the inverse quadrature weights and back, is B1 (x) B1 with
// B1[i][j] = sum_a Ti[i][a] Ti[j][a] / wn[a]; on a cell of area |K|, M^-1 = (B1 (x) B1) / |K|.
// Row-major [nd][nd], node index i + n1 j (x fastest).
std::vector<float> inverse_mass_matrix(int p)
{
  const DegreeTables t = make_tables(p);
  const int n1 = p + 1, nd = n1 * n1;
  std::vector<double> B(n1 * n1, 0.0);
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j)
      for (int a = 0; a < n1; ++a)
        B[i * n1 + j] += t.Ti[i][a] * t.Ti[j][a] / t.wn[a];
  std::vector<float> out(nd * nd);
  for (int r = 0; r < nd; ++r)
    for (int c = 0; c < nd; ++c)
      out[r * nd + c] = float(B[(r % n1) * n1 + (c % n1)] * B[(r / n1) * n1 + (c / n1)]);
  return out;
}

int closure_dispatch(oc_dg *op, const OpArgs &O, const double *zeta, const double *q,
                     float *features, hipStream_t s)
{
  switch (op->P)
    {
      case 1:
        return closure_d1(op, O, zeta, q, features, s);
      case 2:
        return closure_d2(op, O, zeta, q, features, s);
      case 3:
        return closure_d3(op, O, zeta, q, features, s);
      default:
        return closure_d4(op, O, zeta, q, features, s);
    }
}

// orthonormal Legendre polynomial of degree n on [0, 1] at t
double legendre01(int n, double t)
{
  const double x = 2.0 * t - 1.0;
  double a = 1.0, b = x;
  if (n == 0)
    return 1.0;
  for (int k = 1; k < n; ++k)
    {
      const double c = ((2 * k + 1) * x * b - k * a) / (k + 1);
      a = b;
      b = c;
    }
  return std::sqrt(2.0 * n + 1.0) * b;
}

bool g_tables_ready[64] = {};

// Per-batch scratch of the stage kernels for oc_dg_perform_stage_hydro_noalloc: deferred
// depth-solve masks (one 64-bit word per batch), then Jacobi sweep counts (one 32-bit word
// per batch).  Each GPU holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device (g_scratch).
constexpr int NOALLOC_WORDS = OC_NOALLOC_MAX_BATCHES + OC_NOALLOC_MAX_BATCHES / 2;
static __device__ __align__(256) unsigned long long scratch[NOALLOC_WORDS];
unsigned long long *g_scratch[64] = {};

// once per device, with that device current (called by oc_dg_create)
int ensure_scratch(int dev)
{
  if (g_scratch[dev])
    return OC_OK;
  void *ptr = nullptr;
  HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch[dev] = static_cast<unsigned long long *>(ptr);
  return OC_OK;
}

int ensure_tables()
{
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64)
    {
      set_error("device index out of range");
      return OC_ERR_HIP;
    }
  if (g_tables_ready[dev])
    return OC_OK;
  DegreeTables t[4];
  Basis<float> tf[4];
  for (int p = 1; p <= 4; ++p)
    {
      t[p - 1] = make_tables(p);
      tf[p - 1] = to_float(t[p - 1]);
    }
  int rc;
  if ((rc = upload_tables_d1(t, tf)) || (rc = upload_tables_d2(t, tf)) ||
      (rc = upload_tables_d3(t, tf)) || (rc = upload_tables_d4(t, tf)) ||
      (rc = upload_tables_multi(t, tf)))
    return rc;
  g_tables_ready[dev] = true;
  return OC_OK;
}

template <typename T>
int dalloc(T **p, size_t n)
{
  *p = nullptr;
  if (n == 0)
    return OC_OK;
  HIP_TRY(hipMalloc((void **)p, n * sizeof(T)));
  return OC_OK;
}

template <typename T>
int upload(T **p, const std::vector<T> &h)
{
  if (*p)
    {
      (void)hipFree(*p);
      *p = nullptr;
    }
  if (h.empty())
    return OC_OK;
  int rc = dalloc(p, h.size());
  if (rc)
    return rc;
  HIP_TRY(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return OC_OK;
}

__global__ void k_pack(int n_cells, int ncomp, int nd, const int *__restrict__ order,
                       const double *__restrict__ in, double *__restrict__ out, int unpack)
{
  const size_t n_batches = ((size_t)n_cells + CB - 1) / CB;
  const size_t total = n_batches * CB * ncomp * nd;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x)
    {
      // t enumerates the operator order: ((batch*ncomp + comp)*nd + k)*CB + lane
      const int lane = t % CB;
      const size_t r = t / CB;
      const int k = r % nd;
      const size_t r2 = r / nd;
      const int comp = r2 % ncomp;
      const size_t batch = r2 / ncomp;
      const size_t slot = batch * CB + lane;
      if (slot >= (size_t)n_cells)
        {
          if (!unpack)
            out[t] = 0.0;  // padding lanes
          continue;
        }
      const size_t cell = (size_t)order[slot];
      const size_t cm = (cell * ncomp + comp) * nd + k;
      if (unpack)
        out[cm] = in[t];
      else
        out[t] = in[cm];
    }
}

__global__ void k_zero_count(unsigned long long *p) { *p = 0ull; }

// per batch: the sum of the active lanes' entries of an operator-order vector
__global__ void k_batch_sum(int n_cells, int per_lane, const double *__restrict__ x,
                            double *__restrict__ out_blk)
{
  const int batch = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  if (batch * CB + lane < n_cells)
    for (int k = 0; k < per_lane; ++k)
      s += x[((size_t)batch * per_lane + k) * CB + lane];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    s += __shfl_xor(s, off);
  if (lane == 0)
    out_blk[batch] = s;
}

__global__ void k_count_nonfinite(const double *__restrict__ x, size_t n,
                                  unsigned long long *count)
{
  unsigned long long local = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    local += !isfinite(x[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    local += __shfl_xor(local, off);
  if ((threadIdx.x & 63) == 0 && local)
    atomicAdd(count, local);
}

}  // namespace

static hipEvent_t pool_event(oc_dg *op)
{
  if (!op->event_pool.empty())
    {
      hipEvent_t e = op->event_pool.back();
      op->event_pool.pop_back();
      return e;
    }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess)
    return nullptr;
  return e;
}

int timed_begin(oc_dg *op, hipStream_t s, hipEvent_t *start)
{
  *start = nullptr;
  if (!op->timing)
    return OC_OK;
  *start = pool_event(op);
  if (*start)
    (void)hipEventRecord(*start, s);
  return OC_OK;
}

void timed_end(oc_dg *op, hipStream_t s, int kernel, hipEvent_t start)
{
  if (!op->timing || !start)
    return;
  hipEvent_t stop = pool_event(op);
  if (!stop)
    return;
  (void)hipEventRecord(stop, s);
  op->events.push_back({kernel, {start, stop}});
}

}  // namespace oceano

using namespace oceano;

namespace {

void free_all(oc_dg *op)
{
  for (auto &e : op->events)
    {
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
  for (auto e : op->event_pool)
    (void)hipEventDestroy(e);
  auto f = [](void *p) {
    if (p)
      (void)hipFree(p);
  };
  f(op->d_hxy);
  f(op->d_order);
  f(op->d_nbr);
  f(op->d_fid);
  f(op->d_finfo);
  f(op->d_bkind);
  for (auto p : op->d_cell0)
    f(p);
  f(op->d_cell1);
  f(op->d_cell2);
  f(op->d_cell2n);
  f(op->d_cstats);
  f(op->d_cvi);
  f(op->d_cv);
  f(op->d_closure);
  f(op->d_face);
  f(op->d_bz);
  f(op->d_btable);
  f(op->d_sweeps_blk);
  f(op->d_jmat);
  f(op->d_kinv);
  f(op->d_defer);
  f(op->d_post);
  f(op->d_count);
  f(op->d_multi_args);
}

// face point on local face f of cell geometry g at face parameter t in [0,1]
void face_point(const double *g, int f, double t, double &x, double &y)
{
  switch (f)
    {
      case 0:
        x = g[0];
        y = g[1] + t * g[3];
        break;
      case 1:
        x = g[0] + g[2];
        y = g[1] + t * g[3];
        break;
      case 2:
        x = g[0] + t * g[2];
        y = g[1];
        break;
      default:
        x = g[0] + t * g[2];
        y = g[1] + g[3];
        break;
    }
}

bool uniform(const std::vector<double> &v, double &val)
{
  if (v.empty())
    {
      val = 0.0;
      return true;
    }
  for (size_t i = 1; i < v.size(); ++i)
    if (std::memcmp(&v[i], &v[0], sizeof(double)) != 0)
      return false;
  val = v[0];
  return true;
}

// Cell slots.  A wave evaluates the second subface of face f whenever one of its 64 cells
// is the coarse side there, and the boundary branch whenever one of them has a boundary
// face f; cells are therefore grouped by their face signature (per face: conforming or
// fine side / coarse side / boundary), keeping the caller's order inside each group, so
// that almost every batch of 64 has one signature and runs one face path.
// Sorting is done within chunks of consecutive cells of the caller's order: a chunk of one
// XCD's share of the batches keeps each XCD's neighbour traces in its own L2.
// Default (0): FP64 sorts within one XCD's share (fewer divergent face passes pay off at one
// wave per SIMD); the mixed path keeps the caller's order (its two waves per SIMD hide the
// divergence and it gains more from neighbour locality) -- measured on the C2 vortex.
int g_xcd_swizzle = 1;  // batch -> workgroup order: each XCD gets a contiguous batch range
int g_waves_per_block = 1;  // work items (waves of 64 cells) per workgroup of the stage
// inviscid one-pass stage: a batch's two items in one workgroup, faces split between them;
// < 0: by degree -- paired at p <= 3 (measured, paired /
// unpaired time: hanging 97k-cell vortex at p = 3, mixed 52.0 / 57.7 us, FP64 100.9 / 104.9
// us per stage; the conforming ensemble groups' steps, FP64 0.87, 0.96, 1.08, 1.24 and mixed
// 0.86, 0.90, 0.96, 1.13 for p = 1..4: at p = 4 (one or two waves per SIMD) the pair's
// barrier costs more than the shared face work saves)
int g_paired_items = -1;
int g_slot_chunk = 0;  // batches per chunk; 0: by precision as above; < 0: the caller's order

int cell_slots(const oc_mesh_desc *mesh, int precision, std::vector<int32_t> &order,
               std::vector<int32_t> &pos)
{
  const int N = mesh->n_cells;
  std::vector<int> code((size_t)N * 4, 0);  // 0 conforming / fine side, 1 coarse, 2 boundary
  auto mark = [&](int c, int f, int k) {
    if (c < 0 || c >= N || f < 0 || f > 3)
      return false;
    code[4 * (size_t)c + f] = std::max(code[4 * (size_t)c + f], k);
    return true;
  };
  for (int i = 0; i < mesh->n_inner; ++i)
    {
      const int *F = &mesh->inner[6 * i];
      // a subface index marks that side as the coarse one (its face is split in two)
      if (!mark(F[0], F[1], F[2] >= 0 ? 1 : 0) || !mark(F[3], F[4], F[5] >= 0 ? 1 : 0))
        {
          set_error("inner face " + std::to_string(i) + " is malformed");
          return OC_ERR_MESH;
        }
    }
  for (int b = 0; b < mesh->n_boundary; ++b)
    if (!mark(mesh->boundary[3 * b], mesh->boundary[3 * b + 1], 2))
      {
        set_error("boundary face " + std::to_string(b) + " is malformed");
        return OC_ERR_MESH;
      }
  std::vector<int> sig(N);
  for (int c = 0; c < N; ++c)
    sig[c] = code[4 * c] + 3 * (code[4 * c + 1] + 3 * (code[4 * c + 2] + 3 * code[4 * c + 3]));
  order.resize(N);
  for (int c = 0; c < N; ++c)
    order[c] = c;
  const int nb = (N + CB - 1) / CB;
  const bool sort = g_slot_chunk > 0 || (g_slot_chunk == 0 && precision == OC_PRECISION_FP64);
  const long chunk =
    !sort ? 0 : (long)CB * (g_slot_chunk > 0 ? g_slot_chunk : (nb + NXCD - 1) / NXCD);
  for (long c0 = 0; chunk > 0 && c0 < N; c0 += chunk)
    std::stable_sort(order.begin() + c0, order.begin() + std::min<long>(N, c0 + chunk),
                     [&](int a, int b) { return sig[a] < sig[b]; });
  pos.resize(N);
  for (int i = 0; i < N; ++i)
    pos[order[i]] = i;
  return OC_OK;
}

}  // namespace

namespace oceano {

long find_boundary_time(const oc_dg *op, double time)
{
  if (op->bt_any)
    return 0;
  const double tol = 1e-12 * std::max(1.0, std::fabs(time));
  for (size_t k = 0; k < op->bt_times.size(); ++k)
    if (std::fabs(op->bt_times[k] - time) <= tol)
      return (long)k;
  return -1;
}

int boundary_entry(oc_dg *op, double time, uint32_t stage, OpArgs &O, const char *fn)
{
  if (!op->needs_bdata)
    return OC_OK;
  if (op->capture_tick)
    {
      O.btick = op->capture_tick;
      O.bstage = (int)stage;
      return OC_OK;
    }
  const long k = find_boundary_time(op, time);
  if (k < 0)
    {
      char buf[64];
      std::snprintf(buf, sizeof(buf), "%.17g", time);
      set_error(std::string(fn) + ": no boundary data tabulated at time " + buf +
                " (oc_dg_set_boundary_table)");
      return OC_ERR_ARGUMENT;
    }
  O.bvals = op->d_btable + (size_t)k * (size_t)O.bstride;
  return OC_OK;
}

}  // namespace oceano

extern "C" {

const char *oc_last_error(void) { return oceano::g_last_error.c_str(); }

int oc_set_option(const char *name, int64_t value)
{
  if (!name)
    return bad_argument(__func__);
  if (std::strcmp(name, "slot_chunk") == 0)
    {
      g_slot_chunk = (int)value;
      return OC_OK;
    }
  if (std::strcmp(name, "waves_per_block") == 0)
    {
      if (value != 1 && value != 2 && value != 4)
        {
          set_error("oc_set_option: waves_per_block must be 1, 2 or 4");
          return OC_ERR_ARGUMENT;
        }
      g_waves_per_block = (int)value;
      return OC_OK;
    }
  if (std::strcmp(name, "xcd_swizzle") == 0)
    {
      g_xcd_swizzle = value != 0;
      return OC_OK;
    }
  if (std::strcmp(name, "paired_items") == 0)
    {
      if (value < -1 || value > 1)
        {
          set_error("oc_set_option: paired_items must be -1, 0 or 1");
          return OC_ERR_ARGUMENT;
        }
      g_paired_items = (int)value;
      return OC_OK;
    }
  set_error(std::string("oc_set_option: unknown option ") + name);
  return OC_ERR_ARGUMENT;
}

int oc_dg_create(const oc_mesh_desc *mesh, const oc_params *prm, oc_dg **out)
{
  if (!mesh || !prm || !out || mesh->n_cells <= 0 || !mesh->cell_geom ||
      (mesh->n_inner > 0 && !mesh->inner) || (mesh->n_boundary > 0 && !mesh->boundary))
    {
      set_error("oc_dg_create: null or empty argument");
      return OC_ERR_ARGUMENT;
    }
  if (prm->degree < 1 || prm->degree > 4)
    {
      set_error("oc_dg_create: degree must be in 1..4");
      return OC_ERR_ARGUMENT;
    }
  if (prm->precision != OC_PRECISION_FP64 && prm->precision != OC_PRECISION_MIXED)
    {
      set_error("oc_dg_create: unknown precision");
      return OC_ERR_ARGUMENT;
    }
  if (prm->max_iteration_height < 1)
    {
      set_error("oc_dg_create: max_iteration_height must be >= 1");
      return OC_ERR_ARGUMENT;
    }
  int rc = OC_OK;
  oc_dg *op = new oc_dg;
  op->xcd_swizzle = g_xcd_swizzle;
  op->waves_per_block = g_waves_per_block;
  op->paired_items = g_paired_items >= 0 ? g_paired_items : prm->degree <= 3;
  op->P = prm->degree;
  op->n1 = op->P + 1;
  op->nd = op->n1 * op->n1;
  op->nq = nq_of(op->P);
  op->ng = op->P + 2;
  op->N = mesh->n_cells;
  op->NB = (op->N + CB - 1) / CB;
  op->precision = prm->precision;
  op->g = prm->g;
  const double cu = prm->threshold_wetdry;
  op->cu4 = cu * cu * cu * cu;
  op->nu = prm->viscosity;
  op->max_it = prm->max_iteration_height;
  op->geom.assign(mesh->cell_geom, mesh->cell_geom + 4 * (size_t)op->N);
  for (int i = 0; i < prm->n_boundary_kinds; ++i)
    {
      const int kind = prm->boundary_kinds[i];
      if (kind < OC_BC_WALL || kind > OC_BC_ABSORBING_OUTFLOW)
        {
          delete op;
          set_error("oc_dg_create: unknown boundary kind");
          return OC_ERR_ARGUMENT;
        }
      if (op->bc_of_id.count(prm->boundary_ids[i]))
        {
          delete op;
          set_error("You already set the boundary with id " +
                    std::to_string(prm->boundary_ids[i]) + " to another type of boundary");
          return OC_ERR_BOUNDARY;
        }
      op->bc_of_id[prm->boundary_ids[i]] = kind;
    }
  const int N = op->N, NB = op->NB;
  if (const int rc2 = cell_slots(mesh, op->precision, op->order, op->pos))
    {
      delete op;
      return rc2;
    }
  // topology -> per-cell face tables (cell c sits in slot pos[c])
  std::vector<int> finfo((size_t)NB * 4 * CB, -1), nbr((size_t)NB * 4 * 2 * CB, 0),
    fid((size_t)NB * 4 * 2 * CB, 0);
  std::vector<int> slots((size_t)N * 4, 0);
  auto at4 = [&](int c, int f) {
    const int i = op->pos[c];
    return ((size_t)(i / CB) * 4 + f) * CB + i % CB;
  };
  auto at8 = [&](int c, int f, int s) {
    const int i = op->pos[c];
    return (((size_t)(i / CB) * 4 + f) * 2 + s) * CB + i % CB;
  };
  auto fail = [&](const std::string &m, int code) {
    delete op;
    set_error(m);
    return code;
  };
  auto set_kind = [&](int c, int f, int kind) -> bool {
    int &k = finfo[at4(c, f)];
    const int base = k < 0 ? -1 : (k & 3);
    if (base >= 0 && base != (kind & 3))
      return false;
    k = kind;
    return true;
  };
  op->inner_in.assign(mesh->inner, mesh->inner + 6 * (size_t)mesh->n_inner);
  op->inner.resize(6 * (size_t)mesh->n_inner);
  op->swapped.assign(mesh->n_inner, 0);
  for (int i = 0; i < mesh->n_inner; ++i)
    {
      int cm = mesh->inner[6 * i], fm = mesh->inner[6 * i + 1], sm = mesh->inner[6 * i + 2];
      int cp = mesh->inner[6 * i + 3], fp = mesh->inner[6 * i + 4], sp = mesh->inner[6 * i + 5];
      if (cm < 0 || cm >= N || cp < 0 || cp >= N || fm < 0 || fm > 3 || fp < 0 || fp > 3 ||
          sm < -1 || sm > 1 || sp < -1 || sp > 1 || (sm >= 0 && sp >= 0))
        return fail("inner face " + std::to_string(i) + " is malformed", OC_ERR_MESH);
      if ((fm & 1) == 0)
        {
          std::swap(cm, cp);
          std::swap(fm, fp);
          std::swap(sm, sp);
          op->swapped[i] = 1;
        }
      if (fp != fm - 1)
        return fail("inner face " + std::to_string(i) + " joins non-opposite faces", OC_ERR_MESH);
      int *row = &op->inner[6 * i];
      row[0] = cm;
      row[1] = fm;
      row[2] = sm;
      row[3] = cp;
      row[4] = fp;
      row[5] = sp;
      bool ok = true;
      if (sm < 0 && sp < 0)
        {
          ok = set_kind(cm, fm, FK_CONF) && set_kind(cp, fp, FK_CONF);
          nbr[at8(cm, fm, 0)] = op->pos[cp];
          fid[at8(cm, fm, 0)] = i;
          nbr[at8(cp, fp, 0)] = op->pos[cm];
          fid[at8(cp, fp, 0)] = i;
          slots[4 * cm + fm] += 1;
          slots[4 * cp + fp] += 1;
        }
      else if (sm >= 0)
        {
          ok = set_kind(cm, fm, FK_COARSE) && set_kind(cp, fp, FK_FINE | (sm << 2));
          nbr[at8(cm, fm, sm)] = op->pos[cp];
          fid[at8(cm, fm, sm)] = i;
          nbr[at8(cp, fp, 0)] = op->pos[cm];
          fid[at8(cp, fp, 0)] = i;
          slots[4 * cm + fm] += 1 << (4 + sm);
          slots[4 * cp + fp] += 1;
        }
      else
        {
          ok = set_kind(cm, fm, FK_FINE | (sp << 2)) && set_kind(cp, fp, FK_COARSE);
          nbr[at8(cm, fm, 0)] = op->pos[cp];
          fid[at8(cm, fm, 0)] = i;
          nbr[at8(cp, fp, sp)] = op->pos[cm];
          fid[at8(cp, fp, sp)] = i;
          slots[4 * cm + fm] += 1;
          slots[4 * cp + fp] += 1 << (4 + sp);
        }
      if (!ok)
        return fail("cell face used by faces of different kinds (inner face " +
                      std::to_string(i) + ")",
                    OC_ERR_MESH);
    }
  op->bdry.assign(mesh->boundary, mesh->boundary + 3 * (size_t)mesh->n_boundary);
  std::vector<int> bkind(mesh->n_boundary);
  for (int b = 0; b < mesh->n_boundary; ++b)
    {
      const int c = mesh->boundary[3 * b], f = mesh->boundary[3 * b + 1],
                id = mesh->boundary[3 * b + 2];
      if (c < 0 || c >= N || f < 0 || f > 3)
        return fail("boundary face " + std::to_string(b) + " is malformed", OC_ERR_MESH);
      auto it = op->bc_of_id.find(id);
      if (it == op->bc_of_id.end())
        return fail("Unknown boundary id " + std::to_string(id) +
                      ", did you set a boundary condition for this part of the domain boundary?",
                    OC_ERR_BOUNDARY);
      bkind[b] = it->second;
      if (it->second != OC_BC_WALL && it->second != OC_BC_SUPERCRITICAL_OUTFLOW)
        op->needs_bdata = true;
      if (!set_kind(c, f, FK_BDRY))
        return fail("boundary face " + std::to_string(b) + " overlaps an inner face", OC_ERR_MESH);
      fid[at8(c, f, 0)] = b;
      slots[4 * c + f] += 1;
    }
  for (int c = 0; c < N; ++c)
    for (int f = 0; f < 4; ++f)
      {
        const int k = finfo[at4(c, f)];
        const int s = slots[4 * c + f];
        const bool ok = (k >= 0) && (((k & 3) == FK_COARSE) ? (s == 0x30) : (s == 1));
        if (!ok)
          return fail("face " + std::to_string(f) + " of cell " + std::to_string(c) +
                        " is not covered exactly once",
                      OC_ERR_MESH);
      }
  for (auto &k : finfo)
    if (k < 0)
      k = FK_CONF;  // padding lanes
  std::vector<double> hxy((size_t)NB * 2 * CB, 1.0);
  for (int c = 0; c < N; ++c)
    {
      const int i = op->pos[c];
      hxy[((size_t)(i / CB) * 2 + 0) * CB + i % CB] = op->geom[4 * c + 2];
      hxy[((size_t)(i / CB) * 2 + 1) * CB + i % CB] = op->geom[4 * c + 3];
    }
  if ((rc = hipGetDevice(&op->device) == hipSuccess ? OC_OK : OC_ERR_HIP) ||
      op->device < 0 || op->device >= 64)
    {
      set_error("oc_dg_create: no usable current device");
      delete op;
      return OC_ERR_HIP;
    }
  if ((rc = ensure_tables()) || (rc = ensure_scratch(op->device)))
    {
      delete op;
      return rc;
    }
  if ((rc = upload(&op->d_hxy, hxy)) || (rc = upload(&op->d_order, op->order)) ||
      (rc = upload(&op->d_nbr, nbr)) ||
      (rc = upload(&op->d_fid, fid)) || (rc = upload(&op->d_finfo, finfo)) ||
      (rc = upload(&op->d_bkind, bkind)) || (rc = dalloc(&op->d_sweeps_blk, (size_t)NB)) ||
      (rc = upload(&op->d_jmat, jacobi_matrix(op->P, op->max_it))) ||
      (rc = upload(&op->d_kinv, inverse_mass_matrix(op->P))) ||
      (rc = dalloc(&op->d_defer, (size_t)NB)) || (rc = dalloc(&op->d_post, (size_t)NB)) ||
      (rc = dalloc(&op->d_count, 1)) || (rc = dalloc(&op->d_multi_args, 1)))
    {
      free_all(op);
      delete op;
      return rc;
    }
  if (mesh->n_boundary > 0)
    {
      std::vector<double> bv((size_t)mesh->n_boundary * op->nq * 3, 0.0);
      if ((rc = upload(&op->d_btable, bv)))
        {
          free_all(op);
          delete op;
          return rc;
        }
    }
  *out = op;
  return OC_OK;
}

int oc_dg_destroy(oc_dg *op)
{
  if (!op)
    return bad_argument(__func__);
  free_all(op);
  delete op;
  return OC_OK;
}

int oc_dg_sizes(const oc_dg *op, int64_t sizes[7])
{
  if (!op || !sizes)
    return bad_argument(__func__);
  sizes[0] = op->N;
  sizes[1] = (int64_t)op->NB * CB;
  sizes[2] = op->nd;
  sizes[3] = op->nq;
  sizes[4] = op->ng;
  sizes[5] = (int64_t)op->inner.size() / 6;
  sizes[6] = (int64_t)op->bdry.size() / 3;
  return OC_OK;
}

int oc_dg_vector_length(const oc_dg *op, int32_t n_components, int64_t *length)
{
  if (!op || !length || n_components < 1)
    return bad_argument(__func__);
  *length = (int64_t)op->NB * CB * n_components * op->nd;
  return OC_OK;
}

int oc_dg_n_points(const oc_dg *op, int32_t store, int64_t *n)
{
  if (!op || !n)
    return bad_argument(__func__);
  const int64_t ni = (int64_t)op->inner.size() / 6, nb = (int64_t)op->bdry.size() / 3;
  switch (store)
    {
      case OC_STORE_CELL_GAUSS:
        *n = (int64_t)op->N * op->nq * op->nq;
        return OC_OK;
      case OC_STORE_CELL_LOBATTO:
        *n = (int64_t)op->N * op->ng * op->ng;
        return OC_OK;
      case OC_STORE_CELL_COLLOC:
        *n = (int64_t)op->N * op->n1 * op->n1;
        return OC_OK;
      case OC_STORE_FACE:
        *n = ni * op->ng * 2;
        return OC_OK;
      case OC_STORE_BOUNDARY:
        *n = nb * op->nq;
        return OC_OK;
      default:
        set_error("unknown store");
        return OC_ERR_ARGUMENT;
    }
}

int oc_dg_quadrature_points(const oc_dg *op, int32_t store, double *xy)
{
  int64_t npts = 0;
  if (const int rc = oc_dg_n_points(op, store, &npts))
    return rc;
  if (npts == 0)
    return OC_OK;  // empty store (no boundary / inner faces): xy may be null
  if (!xy)
    return bad_argument(__func__);
  const int N = op->N;
  auto cell_rule = [&](const Rule1D &r) {
    const int m = (int)r.x.size();
    size_t o = 0;
    for (int c = 0; c < N; ++c)
      {
        const double *g = &op->geom[4 * c];
        for (int b = 0; b < m; ++b)
          for (int a = 0; a < m; ++a)
            {
              xy[o++] = g[0] + r.x[a] * g[2];
              xy[o++] = g[1] + r.x[b] * g[3];
            }
      }
  };
  switch (store)
    {
      case OC_STORE_CELL_GAUSS:
        cell_rule(gauss(op->nq));
        return OC_OK;
      case OC_STORE_CELL_LOBATTO:
        cell_rule(gauss_lobatto(op->ng));
        return OC_OK;
      case OC_STORE_CELL_COLLOC:
        cell_rule(gauss(op->n1));
        return OC_OK;
      case OC_STORE_FACE:
        {
          const Rule1D r = gauss_lobatto(op->ng);
          const size_t ni = op->inner_in.size() / 6;
          size_t o = 0;
          for (size_t i = 0; i < ni; ++i)
            {
              const int32_t *F = &op->inner_in[6 * i];
              const double nx = F[1] == 0 ? -1.0 : F[1] == 1 ? 1.0 : 0.0;
              const double ny = F[1] == 2 ? -1.0 : F[1] == 3 ? 1.0 : 0.0;
              for (int a = 0; a < op->ng; ++a)
                {
                  double x, y;
                  const double tm = F[2] >= 0 ? 0.5 * (r.x[a] + F[2]) : r.x[a];
                  face_point(&op->geom[4 * F[0]], F[1], tm, x, y);
                  xy[o++] = x - 1e-12 * nx;
                  xy[o++] = y - 1e-12 * ny;
                  const double tp = F[5] >= 0 ? 0.5 * (r.x[a] + F[5]) : r.x[a];
                  face_point(&op->geom[4 * F[3]], F[4], tp, x, y);
                  xy[o++] = x + 1e-12 * nx;
                  xy[o++] = y + 1e-12 * ny;
                }
            }
          return OC_OK;
        }
      case OC_STORE_BOUNDARY:
        {
          const Rule1D r = gauss(op->nq);
          size_t o = 0;
          for (size_t b = 0; b < op->bdry.size() / 3; ++b)
            for (int a = 0; a < op->nq; ++a)
              {
                double x, y;
                face_point(&op->geom[4 * op->bdry[3 * b]], op->bdry[3 * b + 1], r.x[a], x, y);
                xy[o++] = x;
                xy[o++] = y;
              }
          return OC_OK;
        }
      default:
        set_error("unknown store");
        return OC_ERR_ARGUMENT;
    }
}

int oc_dg_set_static_data(oc_dg *op, int32_t store, const double *values)
{
  int64_t npts = 0;
  if (const int rc = oc_dg_n_points(op, store, &npts))
    return rc;
  // the stores are reallocated (or turned into constants): a graph recorded before holds
  // the old pointers and values
  op->generation++;
  if (npts == 0)
    return OC_OK;  // empty store: nothing is read from it
  if (!values)
    return bad_argument(__func__);
  const int N = op->N, NB = op->NB;
  int rc = OC_OK;
  auto interleave = [&](const std::vector<double> &cm, int npts, double **dptr, double &cval) {
    double u;
    if (uniform(cm, u))
      {
        if (*dptr)
          (void)hipFree(*dptr);
        *dptr = nullptr;
        cval = u;
        return OC_OK;
      }
    std::vector<double> il((size_t)NB * npts * CB, 0.0);
    for (int c = 0; c < N; ++c)
      for (int q = 0; q < npts; ++q)
        {
          const int i = op->pos[c];
          il[((size_t)(i / CB) * npts + q) * CB + i % CB] = cm[(size_t)c * npts + q];
        }
    cval = 0.0;
    return upload(dptr, il);
  };
  switch (store)
    {
      case OC_STORE_CELL_GAUSS:
        {
          const int np = op->nq * op->nq;
          op->zb_gauss_host.resize((size_t)N * np);
          for (size_t i = 0; i < op->zb_gauss_host.size(); ++i)
            op->zb_gauss_host[i] = values[i * 5];
          for (int f = 0; f < 5 && !rc; ++f)
            {
              std::vector<double> cm((size_t)N * np);
              for (size_t i = 0; i < cm.size(); ++i)
                cm[i] = values[i * 5 + f];
              rc = interleave(cm, np, &op->d_cell0[f], op->c_cell0[f]);
            }
          return rc;
        }
      case OC_STORE_CELL_LOBATTO:
        {
          const int np = op->ng * op->ng;
          std::vector<double> cm(values, values + (size_t)N * np);
          return interleave(cm, np, &op->d_cell1, op->c_cell1);
        }
      case OC_STORE_CELL_COLLOC:
        {
          op->cell2_host.assign(values, values + (size_t)N * op->nd * 2);
          std::vector<double> cm((size_t)N * op->nd), cn((size_t)N * op->nd);
          for (size_t i = 0; i < cm.size(); ++i)
            {
              cm[i] = values[2 * i];
              cn[i] = values[2 * i + 1];
            }
          if ((rc = interleave(cm, op->nd, &op->d_cell2, op->c_cell2)))
            return rc;
          return interleave(cn, op->nd, &op->d_cell2n, op->c_cell2n);
        }
      case OC_STORE_FACE:
        {
          const size_t ni = op->inner.size() / 6;
          std::vector<double> v(values, values + ni * op->ng * 2);
          for (size_t i = 0; i < ni; ++i)
            if (op->swapped[i])
              for (int a = 0; a < op->ng; ++a)
                std::swap(v[(i * op->ng + a) * 2], v[(i * op->ng + a) * 2 + 1]);
          double u;
          if (uniform(v, u))
            {
              if (op->d_face)
                (void)hipFree(op->d_face);
              op->d_face = nullptr;
              op->c_face = u;
              return OC_OK;
            }
          op->c_face = 0.0;
          return upload(&op->d_face, v);
        }
      case OC_STORE_BOUNDARY:
        {
          std::vector<double> v(values, values + (op->bdry.size() / 3) * op->nq);
          double u;
          if (uniform(v, u))
            {
              if (op->d_bz)
                (void)hipFree(op->d_bz);
              op->d_bz = nullptr;
              op->c_bz = u;
              return OC_OK;
            }
          op->c_bz = 0.0;
          return upload(&op->d_bz, v);
        }
      default:
        set_error("unknown store");
        return OC_ERR_ARGUMENT;
    }
}


static int upload_boundary(oc_dg *op, size_t n_times, const double *values)
{
  const size_t n = (op->bdry.size() / 3) * op->nq * 3;
  // a stage still running on another stream may read the current table
  HIP_TRY(hipDeviceSynchronize());
  std::vector<double> v(values, values + n * n_times);
  if (const int rc = upload(&op->d_btable, v))
    return rc;
  op->generation++;
  return OC_OK;
}

int oc_dg_set_boundary_values(oc_dg *op, const double *values)
{
  if (!op)
    return bad_argument(__func__);
  const size_t n = (op->bdry.size() / 3) * op->nq * 3;
  if (n == 0)
    return OC_OK;
  if (!values)
    return bad_argument(__func__);
  if (const int rc = upload_boundary(op, 1, values))
    return rc;
  op->bt_times.clear();
  op->bt_any = true;
  return OC_OK;
}

int oc_dg_set_boundary_table(oc_dg *op, int32_t n_times, const double *times,
                             const double *values)
{
  if (!op || n_times < 1 || !times)
    return bad_argument(__func__);
  // non-decreasing up to the matching tolerance: stage times computed as t0 + r dt + c dt
  // may step back by a rounding error where one step's last stage meets the next's first
  for (int32_t k = 0; k < n_times; ++k)
    if (!std::isfinite(times[k]) ||
        (k > 0 && times[k] < times[k - 1] - 1e-12 * std::max(1.0, std::fabs(times[k]))))
      {
        set_error("oc_dg_set_boundary_table: times must be finite and non-decreasing");
        return OC_ERR_ARGUMENT;
      }
  const size_t n = (op->bdry.size() / 3) * op->nq * 3;
  if (n > 0)
    {
      if (!values)
        return bad_argument(__func__);
      if (const int rc = upload_boundary(op, (size_t)n_times, values))
        return rc;
    }
  op->bt_times.assign(times, times + n_times);
  op->bt_any = false;
  return OC_OK;
}

int oc_dg_pack(oc_dg *op, int32_t n_components, const double *cell_major, double *op_order,
               void *stream)
{
  if (!op || n_components < 1 || !cell_major || !op_order)
    return bad_argument(__func__);
  const size_t total = (size_t)op->NB * CB * n_components * op->nd;
  hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, op->N,
                     n_components, op->nd, op->d_order, cell_major, op_order, 0);
  HIP_TRY(hipGetLastError());
  return OC_OK;
}

int oc_dg_unpack(oc_dg *op, int32_t n_components, const double *op_order, double *cell_major,
                 void *stream)
{
  if (!op || n_components < 1 || !cell_major || !op_order)
    return bad_argument(__func__);
  const size_t total = (size_t)op->NB * CB * n_components * op->nd;
  hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, op->N,
                     n_components, op->nd, op->d_order, op_order, cell_major, 1);
  HIP_TRY(hipGetLastError());
  return OC_OK;
}

// the stage with its per-batch scratch (sweep counts, deferral masks) at `sweeps` / `defer`
// (null: the operator's own buffers)
static int stage_impl(oc_dg *op, uint32_t stage, double time, const double *factor,
                      uint32_t n_factor, const double *ri_h, const double *ri_q,
                      const double *ri_vel, double *const *ki_h, double *const *ki_q,
                      uint32_t n_ki, double *w_h, double *w_q, double *r_h, double *r_q,
                      int32_t write_scratch, void *stream, unsigned int *sweeps,
                      unsigned long long *defer)
{
  if (!op || !factor || !ri_h || !ri_q || !ki_h || !ki_q || !w_h || !w_q)
    {
      set_error("oc_dg_perform_stage_hydro: null argument");
      return OC_ERR_ARGUMENT;
    }
  if (n_ki < 2 || n_ki > MAXKI || stage + 2 > n_ki || n_factor != stage + 1)
    {
      set_error("oc_dg_perform_stage_hydro: stage / n_ki / n_factor inconsistent");
      return OC_ERR_ARGUMENT;
    }
  const bool last = stage == n_ki - 2;
  if (!last && (!r_h || !r_q))
    {
      set_error("oc_dg_perform_stage_hydro: next_ri vectors required before the last stage");
      return OC_ERR_ARGUMENT;
    }
  if (op->nu != 0.0 && !ri_vel)
    {
      set_error("oc_dg_perform_stage_hydro: velocity vector required when viscosity != 0");
      return OC_ERR_ARGUMENT;
    }
  for (uint32_t j = 0; j <= stage + 1; ++j)
    if (!ki_h[j] || !ki_q[j])
      {
        if (j == 0 && !write_scratch)
          continue;
        set_error("oc_dg_perform_stage_hydro: null ki vector");
        return OC_ERR_ARGUMENT;
      }
  OpArgs O = op->args();
  if (sweeps)
    O.sweeps = sweeps;
  if (defer)
    O.defer = defer;
  if (const int rc = boundary_entry(op, time, stage, O, "oc_dg_perform_stage_hydro"))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  if (op->closure_mlp)
    {
      // the learned correction of this stage's input, added after the inverse mass
      if (const int rc = closure_dispatch(op, O, ri_h, ri_q, nullptr, s))
        return rc;
      O.closure = op->d_closure;
    }
  switch (op->P)
    {
      case 1:
        return stage_d1(op, O, stage, factor, ri_h, ri_q, ri_vel, ki_h, ki_q, n_ki, w_h,
                               w_q, r_h, r_q, write_scratch, s);
      case 2:
        return stage_d2(op, O, stage, factor, ri_h, ri_q, ri_vel, ki_h, ki_q, n_ki, w_h,
                               w_q, r_h, r_q, write_scratch, s);
      case 3:
        return stage_d3(op, O, stage, factor, ri_h, ri_q, ri_vel, ki_h, ki_q, n_ki, w_h,
                               w_q, r_h, r_q, write_scratch, s);
      default:
        return stage_d4(op, O, stage, factor, ri_h, ri_q, ri_vel, ki_h, ki_q, n_ki, w_h,
                               w_q, r_h, r_q, write_scratch, s);
    }
}

int oc_dg_perform_stage_hydro(oc_dg *op, uint32_t stage, double time, const double *factor,
                              uint32_t n_factor, const double *ri_h, const double *ri_q,
                              const double *ri_vel, double *const *ki_h, double *const *ki_q,
                              uint32_t n_ki, double *w_h, double *w_q, double *r_h,
                              double *r_q, int32_t write_scratch, void *stream)
{
  return stage_impl(op, stage, time, factor, n_factor, ri_h, ri_q, ri_vel, ki_h, ki_q, n_ki,
                    w_h, w_q, r_h, r_q, write_scratch, stream, nullptr, nullptr);
}

int oc_dg_perform_stage_hydro_noalloc(oc_dg *op, uint32_t stage, double time,
                                      const double *factor, uint32_t n_factor,
                                      const double *ri_h, const double *ri_q,
                                      const double *ri_vel, double *const *ki_h,
                                      double *const *ki_q, uint32_t n_ki, double *w_h,
                                      double *w_q, double *r_h, double *r_q,
                                      int32_t write_scratch, void *stream)
{
  if (!op)
    {
      set_error("oc_dg_perform_stage_hydro_noalloc: null operator");
      return OC_ERR_ARGUMENT;
    }
  if (op->NB > OC_NOALLOC_MAX_BATCHES)
    {
      set_error("oc_dg_perform_stage_hydro_noalloc: " + std::to_string(op->NB) +
                " cell batches exceed the scratch of " +
                std::to_string(OC_NOALLOC_MAX_BATCHES) + " batches (" +
                std::to_string((long long)OC_NOALLOC_MAX_BATCHES * CB) + " cells)");
      return OC_ERR_ARGUMENT;
    }
  unsigned long long *base = g_scratch[op->device];
  if (!base)
    {
      set_error("oc_dg_perform_stage_hydro_noalloc: scratch address not resolved");
      return OC_ERR_HIP;
    }
  // per-kernel timing would create events: not recorded by this entry point
  const int timing = op->timing;
  op->timing = 0;
  const int rc = stage_impl(op, stage, time, factor, n_factor, ri_h, ri_q, ri_vel, ki_h, ki_q,
                            n_ki, w_h, w_q, r_h, r_q, write_scratch, stream,
                            reinterpret_cast<unsigned int *>(base + OC_NOALLOC_MAX_BATCHES),
                            base);
  op->timing = timing;
  return rc;
}

static int rhs_or_apply(oc_dg *op, bool invert, double time, const double *src_h,
                        const double *src_q, const double *src_vel, double *dst_h,
                        double *dst_q, void *stream, const char *fn)
{
  if (!op || !src_h || !src_q || !dst_h || !dst_q)
    return bad_argument(fn);
  if (op->nu != 0.0 && !src_vel)
    {
      set_error("velocity vector required when viscosity != 0");
      return OC_ERR_ARGUMENT;
    }
  OpArgs O = op->args();
  if (const int rc = boundary_entry(op, time, 0, O, fn))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  switch (op->P)
    {
      case 1:
        return rhs_apply_d1(op, O, invert, src_h, src_q, src_vel, dst_h, dst_q, s);
      case 2:
        return rhs_apply_d2(op, O, invert, src_h, src_q, src_vel, dst_h, dst_q, s);
      case 3:
        return rhs_apply_d3(op, O, invert, src_h, src_q, src_vel, dst_h, dst_q, s);
      default:
        return rhs_apply_d4(op, O, invert, src_h, src_q, src_vel, dst_h, dst_q, s);
    }
}

int oc_dg_perform_stage_hydro_multi(int32_t n_ops, oc_dg *const *ops, uint32_t stage,
                                    const double *time, const double *const *factor,
                                    uint32_t n_factor,
                                    const double *const *ri_h, const double *const *ri_q,
                                    double *const *const *ki_h, double *const *const *ki_q,
                                    uint32_t n_ki, double *const *w_h, double *const *w_q,
                                    double *const *r_h, double *const *r_q, void *stream)
{
  const char *fn = "oc_dg_perform_stage_hydro_multi";
  if (n_ops < 1 || n_ops > OC_MULTI_MAX_OPERATORS || !ops || !time || !factor || !ri_h || !ri_q ||
      !ki_h || !ki_q || !w_h || !w_q || !r_h || !r_q)
    {
      set_error(std::string(fn) + ": null argument or operator count outside 1.." +
                std::to_string(OC_MULTI_MAX_OPERATORS));
      return OC_ERR_ARGUMENT;
    }
  if (n_ki < 2 || n_ki > MAXKI || stage + 2 > n_ki || n_factor != stage + 1)
    {
      set_error(std::string(fn) + ": stage / n_ki / n_factor inconsistent");
      return OC_ERR_ARGUMENT;
    }
  const bool last = stage == n_ki - 2;
  OpArgs O[OC_MULTI_MAX_OPERATORS];
  for (int g = 0; g < n_ops; ++g)
    {
      oc_dg *op = ops[g];
      if (!op || !factor[g] || !ri_h[g] || !ri_q[g] || !ki_h[g] || !ki_q[g] || !w_h[g] ||
          !w_q[g] || (!last && (!r_h[g] || !r_q[g])))
        {
          set_error(std::string(fn) + ": null operator or vector");
          return OC_ERR_ARGUMENT;
        }
      for (uint32_t j = 1; j <= stage + 1; ++j)
        if (!ki_h[g][j] || !ki_q[g][j])
          {
            set_error(std::string(fn) + ": null ki vector");
            return OC_ERR_ARGUMENT;
          }
      double *hd = last ? w_h[g] : r_h[g], *qd = last ? w_q[g] : r_q[g];
      if (op->precision != ops[0]->precision || op->nu != 0.0 || op->closure_mlp ||
          op->device != ops[0]->device || hd == ri_h[g] || qd == ri_q[g] || hd == ri_q[g] ||
          qd == ri_h[g])
        {
          set_error(std::string(fn) + ": the operators must share precision and device, be "
                    "inviscid without a closure, and write outputs distinct from the inputs "
                    "(call oc_dg_perform_stage_hydro per operator otherwise)");
          return OC_ERR_UNSUPPORTED;
        }
      O[g] = op->args();
      if (const int rc = boundary_entry(op, time[g], stage, O[g], fn))
        return rc;
    }
  return stage_multi(n_ops, ops, O, stage, factor, ri_h, ri_q, ki_h, ki_q, n_ki, w_h, w_q,
                     r_h, r_q, (hipStream_t)stream);
}

int oc_dg_perform_stage_hydro_ark(oc_dg *op, uint32_t stage, double time,
                                  const double *factor, const double *factor_tilde,
                                  uint32_t n_factor, double factor_matrix, const double *ri_h,
                                  const double *ri_q, const double *ri_vel,
                                  double *const *ki_h, uint32_t n_ki_h, double *const *ki_q,
                                  uint32_t n_ki_q, double *w_h, double *w_q, double *r_h,
                                  double *r_q, void *stream)
{
  if (!op || !factor || !ri_h || !ri_q || !ki_h || !ki_q || !w_h || !w_q || !r_q)
    {
      set_error("oc_dg_perform_stage_hydro_ark: null argument");
      return OC_ERR_ARGUMENT;
    }
  const uint32_t S = n_ki_h - 1;
  if (n_ki_h < 2 || n_ki_q != 2 * S || stage >= S || n_factor != stage + 1 || S + 1 > MAXKI)
    {
      set_error("oc_dg_perform_stage_hydro_ark: stage / n_ki_h / n_ki_q / n_factor "
                "inconsistent (n_ki_q = 2 (n_ki_h - 1))");
      return OC_ERR_ARGUMENT;
    }
  const bool last = stage == S - 1;
  if ((!last && (!r_h || !factor_tilde || r_h == ri_h || r_q == ri_q)) ||
      (op->nu != 0.0 && !ri_vel))
    {
      set_error("oc_dg_perform_stage_hydro_ark: r_h and factor_tilde are required before the "
                "last stage, r distinct from ri, the velocity when viscosity != 0");
      return OC_ERR_ARGUMENT;
    }
  for (uint32_t j = 1; j < n_ki_q; ++j)  // [0] (the front) is not written here
    if (!ki_q[j] || (j <= S && !ki_h[j]))
      {
        set_error("oc_dg_perform_stage_hydro_ark: null ki vector");
        return OC_ERR_ARGUMENT;
      }
  OpArgs O = op->args();
  if (const int rc = boundary_entry(op, time, stage, O, "oc_dg_perform_stage_hydro_ark"))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  switch (op->P)
    {
      case 1:
        return stage_ark_d1(op, O, stage, S, factor, factor_tilde, factor_matrix, ri_h, ri_q,
                            ri_vel, ki_h, ki_q, w_h, w_q, r_h, r_q, s);
      case 2:
        return stage_ark_d2(op, O, stage, S, factor, factor_tilde, factor_matrix, ri_h, ri_q,
                            ri_vel, ki_h, ki_q, w_h, w_q, r_h, r_q, s);
      case 3:
        return stage_ark_d3(op, O, stage, S, factor, factor_tilde, factor_matrix, ri_h, ri_q,
                            ri_vel, ki_h, ki_q, w_h, w_q, r_h, r_q, s);
      default:
        return stage_ark_d4(op, O, stage, S, factor, factor_tilde, factor_matrix, ri_h, ri_q,
                            ri_vel, ki_h, ki_q, w_h, w_q, r_h, r_q, s);
    }
}

int oc_dg_apply(oc_dg *op, double time, const double *src_h, const double *src_q,
                const double *src_vel, double *dst_h, double *dst_q, void *stream)
{
  return rhs_or_apply(op, true, time, src_h, src_q, src_vel, dst_h, dst_q, stream,
                      "oc_dg_apply");
}

int oc_dg_rhs(oc_dg *op, double time, const double *src_h, const double *src_q,
              const double *src_vel, double *dst_h, double *dst_q, void *stream)
{
  return rhs_or_apply(op, false, time, src_h, src_q, src_vel, dst_h, dst_q, stream,
                      "oc_dg_rhs");
}

int oc_dg_project_collocation(const oc_dg *op, int32_t n_comp, const double *values,
                              double *coeffs)
{
  if (!op || n_comp < 1 || !values || !coeffs)
    return bad_argument(__func__);
  const int n1 = op->n1, nd = op->nd;
  const Rule1D nodes = gauss_lobatto(n1), rn = gauss(n1);
  std::vector<double> S(nd);
  for (int a = 0; a < n1; ++a)
    for (int i = 0; i < n1; ++i)
      S[a * n1 + i] = lagrange_value(nodes.x, i, rn.x[a]);
  const std::vector<double> Ti = small_inverse(S, n1);  // Ti[i][a]
  std::vector<double> T(nd);
  for (int c = 0; c < op->N; ++c)
    for (int d = 0; d < n_comp; ++d)
      {
        const double *u = values + (size_t)c * nd * n_comp;
        double *U = coeffs + ((size_t)c * n_comp + d) * nd;
        for (int b = 0; b < n1; ++b)
          for (int i = 0; i < n1; ++i)
            {
              double s = 0;
              for (int a = 0; a < n1; ++a)
                s += Ti[i * n1 + a] * u[(a + n1 * b) * n_comp + d];
              T[i + n1 * b] = s;
            }
        for (int j = 0; j < n1; ++j)
          for (int i = 0; i < n1; ++i)
            {
              double s = 0;
              for (int b = 0; b < n1; ++b)
                s += Ti[j * n1 + b] * T[i + n1 * b];
              U[i + n1 * j] = s;
            }
      }
  return OC_OK;
}

int oc_dg_compute_errors(const oc_dg *op, const double *zeta, const double *q,
                         const double *exact, double err[7])
{
  if (!op || !zeta || !q || !exact || !err)
    return bad_argument(__func__);
  const int n1 = op->n1, nd = op->nd, m = op->nq, np = m * m;
  const Rule1D nodes = gauss_lobatto(n1), rq = gauss(m);
  std::vector<double> S(m * n1);
  for (int a = 0; a < m; ++a)
    for (int i = 0; i < n1; ++i)
      S[a * n1 + i] = lagrange_value(nodes.x, i, rq.x[a]);
  auto eval = [&](const double *U, double *out) {
    std::vector<double> T(m * n1);
    for (int j = 0; j < n1; ++j)
      for (int a = 0; a < m; ++a)
        {
          double s = 0;
          for (int i = 0; i < n1; ++i)
            s += S[a * n1 + i] * U[i + n1 * j];
          T[a + m * j] = s;
        }
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a)
        {
          double s = 0;
          for (int j = 0; j < n1; ++j)
            s += S[b * n1 + j] * T[a + m * j];
          out[a + m * b] = s;
        }
  };
  std::vector<double> z(np), qx(np), qy(np);
  double e2[3] = {0, 0, 0}, einf[3] = {0, 0, 0};
  for (int c = 0; c < op->N; ++c)
    {
      const double hx = op->geom[4 * c + 2], hy = op->geom[4 * c + 3];
      eval(zeta + (size_t)c * nd, z.data());
      eval(q + (size_t)c * 2 * nd, qx.data());
      eval(q + (size_t)c * 2 * nd + nd, qy.data());
      for (int b = 0; b < m; ++b)
        for (int a = 0; a < m; ++a)
          {
            const int qp = a + m * b;
            const double JxW = rq.w[a] * rq.w[b] * hx * hy;
            const double zb =
              op->zb_gauss_host.empty() ? 0.0 : op->zb_gauss_host[(size_t)c * np + qp];
            const double *ex = exact + ((size_t)c * np + qp) * 3;
            const double hex = std::max(ex[0] + zb, 1e-12);
            const double d0 = hex - std::max(z[qp] + zb, 1e-12);
            const double d1 = ex[1] * hex - qx[qp];
            const double d2 = ex[2] * hex - qy[qp];
            e2[0] += d0 * d0 * JxW;
            e2[1] += d1 * d1 * JxW;
            e2[2] += d2 * d2 * JxW;
            einf[0] = std::max(einf[0], std::fabs(d0));
            einf[1] = std::max(einf[1], std::fabs(d1));
            einf[2] = std::max(einf[2], std::fabs(d2));
          }
    }
  err[0] = std::sqrt(e2[0]);
  err[1] = std::sqrt(e2[1] + e2[2]);
  err[2] = std::sqrt(e2[1]);
  err[3] = std::sqrt(e2[2]);
  err[4] = einf[0];
  err[5] = einf[1];
  err[6] = einf[2];
  return OC_OK;
}

int oc_dg_count_nonfinite(oc_dg *op, const double *x, int64_t n, int64_t *count, void *stream)
{
  if (!op || !x || n < 0 || !count)
    return bad_argument(__func__);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, op->d_count);
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  if (n > 0)
    hipLaunchKernelGGL(k_count_nonfinite, dim3(std::max(blocks, 1)), dim3(256), 0, s, x,
                       (size_t)n, op->d_count);
  unsigned long long h = 0;
  HIP_TRY(hipMemcpyAsync(&h, op->d_count, sizeof(h), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *count = (int64_t)h;
  return OC_OK;
}

int oc_dg_set_timing(oc_dg *op, int32_t enable)
{
  if (!op)
    return bad_argument(__func__);
  op->timing = enable != 0;
  // events are created here, not between the launches they time
  while (op->timing && op->event_pool.size() < 4096)
    {
      hipEvent_t e = nullptr;
      HIP_TRY(hipEventCreate(&e));
      op->event_pool.push_back(e);
    }
  return OC_OK;
}

int oc_dg_timing_report(oc_dg *op, double *ms, int64_t *count)
{
  if (!op || !ms || !count)
    return bad_argument(__func__);
  for (int k = 0; k < TK_COUNT; ++k)
    {
      ms[k] = 0.0;
      count[k] = 0;
    }
  for (auto &e : op->events)
    {
      HIP_TRY(hipEventSynchronize(e.second.second));
      float t = 0.f;
      HIP_TRY(hipEventElapsedTime(&t, e.second.first, e.second.second));
      ms[e.first] += t;
      count[e.first] += 1;
      op->event_pool.push_back(e.second.first);
      op->event_pool.push_back(e.second.second);
    }
  op->events.clear();
  return OC_OK;
}

static int post_dispatch(oc_dg *op, int what, const double *zeta, const double *q, double *out,
                         hipStream_t s)
{
  const OpArgs O = op->args();
  switch (op->P)
    {
      case 1:
        return post_d1(op, O, what, zeta, q, out, s);
      case 2:
        return post_d2(op, O, what, zeta, q, out, s);
      case 3:
        return post_d3(op, O, what, zeta, q, out, s);
      default:
        return post_d4(op, O, what, zeta, q, out, s);
    }
}

// per-batch results -> host, reduced in batch order
static int fetch_batches(oc_dg *op, hipStream_t s, std::vector<double> &h)
{
  h.resize(op->NB);
  HIP_TRY(hipMemcpyAsync(h.data(), op->d_post, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                         s));
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

int oc_dg_velocity(oc_dg *op, const double *zeta, const double *q, double *vel, void *stream)
{
  if (!op || !zeta || !q || !vel)
    return bad_argument(__func__);
  return post_dispatch(op, POST_VELOCITY, zeta, q, vel, (hipStream_t)stream);
}

int oc_dg_transport_speed(oc_dg *op, const double *zeta, const double *q, double *speed,
                          void *stream)
{
  if (!op || !zeta || !q || !speed)
    return bad_argument(__func__);
  hipStream_t s = (hipStream_t)stream;
  if (const int rc = post_dispatch(op, POST_SPEED, zeta, q, op->d_post, s))
    return rc;
  std::vector<double> h;
  if (const int rc = fetch_batches(op, s, h))
    return rc;
  double m = 0.0;
  for (double v : h)
    m = std::max(m, v);
  *speed = m;
  return OC_OK;
}

int oc_dg_volume(oc_dg *op, const double *zeta, double *volume, void *stream)
{
  if (!op || !zeta || !volume)
    return bad_argument(__func__);
  hipStream_t s = (hipStream_t)stream;
  if (const int rc = post_dispatch(op, POST_VOLUME, zeta, nullptr, op->d_post, s))
    return rc;
  std::vector<double> h;
  if (const int rc = fetch_batches(op, s, h))
    return rc;
  double sum = 0.0;
  for (double v : h)
    sum += v;
  *volume = sum;
  return OC_OK;
}

int oc_dg_set_subgrid_stats(oc_dg *op, const double *stats, const int32_t *active)
{
  if (!op || (!stats && op->N > 0))
    return bad_argument(__func__);
  op->generation++;
  if (3 * op->nd + 3 > mlp_inputs_max)
    {
      set_error("oc_dg_set_subgrid_stats: the closure supports degrees 1 and 2");
      return OC_ERR_ARGUMENT;
    }
  const int n1 = op->n1, nd = op->nd, NB = op->NB;
  // [batch][4][lane]: mean, variance, skewness, active flag
  std::vector<double> il((size_t)NB * 4 * CB, 0.0);
  for (int c = 0; c < op->N; ++c)
    {
      const int i = op->pos[c];
      for (int k = 0; k < 3; ++k)
        il[((size_t)(i / CB) * 4 + k) * CB + i % CB] = stats[(size_t)c * 3 + k];
      il[((size_t)(i / CB) * 4 + 3) * CB + i % CB] = (!active || active[c]) ? 1.0 : 0.0;
    }
  int rc = upload(&op->d_cstats, il);
  if (rc)
    return rc;
  // V[k][m] = psi_m(node_k) for k = i + n1 j, m = a + n1 b; Vi = V^-1
  const Rule1D gl = gauss_lobatto(n1);
  std::vector<double> V1((size_t)n1 * n1);
  for (int i = 0; i < n1; ++i)
    for (int a = 0; a < n1; ++a)
      V1[i * n1 + a] = legendre01(a, gl.x[i]);
  const std::vector<double> Vi1 = small_inverse(V1, n1);  // [a][i]
  std::vector<double> V((size_t)nd * nd), Vi((size_t)nd * nd);
  for (int j = 0; j < n1; ++j)
    for (int i = 0; i < n1; ++i)
      for (int b = 0; b < n1; ++b)
        for (int a = 0; a < n1; ++a)
          {
            const int k = i + n1 * j, m = a + n1 * b;
            V[(size_t)k * nd + m] = V1[j * n1 + b] * V1[i * n1 + a];
            Vi[(size_t)m * nd + k] = Vi1[a * n1 + i] * Vi1[b * n1 + j];
          }
  if ((rc = upload(&op->d_cv, V)) || (rc = upload(&op->d_cvi, Vi)))
    return rc;
  return upload(&op->d_closure, std::vector<double>((size_t)NB * CB * 2 * nd, 0.0));
}

int oc_dg_set_closure(oc_dg *op, const oc_mlp *mlp)
{
  if (!op)
    return bad_argument(__func__);
  op->generation++;
  if (!mlp)
    {
      op->closure_mlp = nullptr;
      return OC_OK;
    }
  if (!op->d_cstats)
    {
      set_error("oc_dg_set_closure: oc_dg_set_subgrid_stats first");
      return OC_ERR_ARGUMENT;
    }
  int32_t n_in = 0, n_out = 0;
  oc_mlp_dims(mlp, &n_in, &n_out);
  if (n_in != 3 * op->nd + 3 || n_out != 2 * op->nd)
    {
      set_error("oc_dg_set_closure: the network needs 3 nd + 3 inputs and 2 nd outputs");
      return OC_ERR_ARGUMENT;
    }
  op->closure_mlp = mlp;
  return OC_OK;
}

int oc_dg_closure_features(oc_dg *op, const double *zeta, const double *q, float *features,
                           void *stream)
{
  if (!op || !zeta || !q || !features)
    return bad_argument(__func__);
  if (!op->d_cstats)
    {
      set_error("oc_dg_closure_features: oc_dg_set_subgrid_stats first");
      return OC_ERR_ARGUMENT;
    }
  return closure_dispatch(op, op->args(), zeta, q, features, (hipStream_t)stream);
}

int oc_dg_closure_source(oc_dg *op, const double *zeta, const double *q, double *tau,
                         void *stream)
{
  if (!op || !zeta || !q || !tau)
    return bad_argument(__func__);
  if (!op->closure_mlp)
    {
      set_error("oc_dg_closure_source: no closure network attached");
      return OC_ERR_ARGUMENT;
    }
  hipStream_t s = (hipStream_t)stream;
  if (const int rc = closure_dispatch(op, op->args(), zeta, q, nullptr, s))
    return rc;
  HIP_TRY(hipMemcpyAsync(tau, op->d_closure, (size_t)op->NB * CB * 2 * op->nd * sizeof(double),
                         hipMemcpyDeviceToDevice, s));
  return OC_OK;
}

int oc_dg_sum(oc_dg *op, const double *x, int32_t n_components, double *sum, void *stream)
{
  if (!op || !x || !sum || n_components < 1)
    return bad_argument(__func__);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_batch_sum, dim3(op->NB), dim3(CB), 0, s, op->N, n_components * op->nd, x,
                     op->d_post);
  HIP_TRY(hipGetLastError());
  std::vector<double> h;
  if (const int rc = fetch_batches(op, s, h))
    return rc;
  double acc = 0.0;
  for (double v : h)
    acc += v;
  *sum = acc;
  return OC_OK;
}

int oc_dg_last_sweeps(oc_dg *op, int64_t *sweeps, void *stream)
{
  if (!op || !sweeps)
    return bad_argument(__func__);
  hipStream_t s = (hipStream_t)stream;
  std::vector<unsigned int> h(op->NB);
  HIP_TRY(hipMemcpyAsync(h.data(), op->d_sweeps_blk, h.size() * sizeof(unsigned int),
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  int64_t t = 0;
  for (unsigned int v : h)
    t += v;
  *sweeps = t;
  return OC_OK;
}

}  // extern "C"
