// dg_kernels.hpp -- gfx950 kernels of one explicit RK stage of the shallow-water DG operator.
//
// Mapping: one thread per cell, one wave64 per batch of 64 consecutive cells.  Vectors
// and static data are stored lane-batched ([batch][...][lane]), so every load/store of a
// cell's own data is a fully coalesced 512-byte wave access; the 1-D basis tables live in
// __constant__ memory and are read with scalar loads (uniform across the wave).
// Tensor-product contractions are sum-factorised row by row (one quadrature row b at a
// time), which keeps the per-thread live set at O((p+1)^2) values.
//
// Faces are evaluated cell-centrically: each cell evaluates the flux on each of its
// faces (both traces, identical arithmetic and argument order on both sides, so both
// cells see bit-identical flux values) and integrates it against its own test functions.
// There is no face buffer and no atomic; accumulation order is fixed.
#pragma once

#include "dg_physics.hpp"
#include "oceano_dg.h"

namespace oceano {

constexpr int CB = OC_CELL_BATCH;
constexpr int NXCD = 8;  // MI355X: accelerator complex dies, each with its own L2
constexpr int MAXN1 = 5, MAXNQ = 7, MAXNG = 6, MAXKI = 9;

template <typename R>
struct Basis
{
  R Sq[MAXNQ][MAXN1];      // l_i at Gauss(nq) points
  R Dq[MAXNQ][MAXN1];      // l_i' at Gauss(nq) points
  R wq[MAXNQ];
  R Sg[MAXNG][MAXN1];      // l_i at Gauss-Lobatto(ng) points
  R Sgs[2][MAXNG][MAXN1];  // l_i at (t + s)/2, t in GL(ng): coarse side of subface s
  R wg[MAXNG];
  R Ti[MAXN1][MAXN1];      // inverse of [l_i(Gauss(n1)_a)]_{a,i}, stored [i][a]
  R wn[MAXN1];             // Gauss(n1) weights
  R lumped[MAXN1];         // sum_a Sg[a][i] wg[a]
  R ilumped2[MAXN1][MAXN1];  // 1 / (lumped[i] lumped[j]): inverse lumped mass / area
  R M1[MAXN1][MAXN1];        // sum_a Sg[a][i] Sg[a][j] wg[a]: 1-D GL(ng) mass (exact)
  R Sn[MAXN1][MAXN1];        // l_i at Gauss(n1) points (velocity projection)
  R Psub[2][MAXN1][MAXN1];   // l_k((x_j + s)/2), x_j the nodes: restriction to half s
  R lam;                     // max_{a,i} Sg[a][i]^2: bound of |phi_ij| at the GL(ng)^2 points
};
using DegreeTables = Basis<double>;

enum FaceKind
{
  FK_CONF = 0,    // same-level neighbour
  FK_COARSE = 1,  // this cell is the coarse side of two subfaces
  FK_FINE = 2,    // this cell is the fine side; neighbour coarse (subface index in bits 2)
  FK_BDRY = 3
};

struct Field
{
  const double *ptr;  // stored values, or nullptr when the field is uniform
  double cval;        // the uniform value
};

struct OpArgs
{
  int n_cells;
  int n_batches;
  const double *hxy;    // [batch][2][lane]   hx, hy
  const int *nbr;       // [batch][4][2][lane] neighbour cell per face / subface
  const int *fid;       // [batch][4][2][lane] inner face id or boundary face id
  const int *finfo;     // [batch][4][lane]    kind | subface << 2
  Field cell0[5];       // [batch][nq^2][lane] z_b, manning, tau_x, tau_y, coriolis
  Field cell1;          // [batch][ng^2][lane] z_b at GL points
  Field cell2;          // [batch][n1^2][lane] z_b at Gauss(n1) points
  Field cell2n;         // [batch][n1^2][lane] friction coefficient at Gauss(n1) points
  Field facez;          // [face][ng][2]
  Field bdryz;          // [bface][nq]
  // boundary data [bface][nq][3] of this stage: the table entry itself, or (btick != null,
  // a captured time loop) the table base, the entry being btick[0] + bstage
  const double *bvals;
  const int *btick;
  int bstage;
  long long bstride;    // doubles per table entry
  const int *bkind;     // [bface]
  double g, cu4, nu;
  int has_friction, has_wind, has_coriolis;
  int max_it;
  int xcd_swizzle;
  // mixed precision: closed form of max_it linear Jacobi sweeps on a wet unit cell,
  // [nd][nd] row-major (delta = J r / area), and the wet-cell safety factor
  const float *jmat;
  // mixed precision: the momentum inverse mass of the unit cell, [nd][nd] row-major
  // (M^-1 = K / area on an affine cell)
  const float *kinv;
  double wet_margin;
  unsigned long long *defer;  // mixed path: per batch, lanes whose depth solve is deferred
  unsigned int *sweeps;       // per batch: Jacobi sweeps of the last stage (sum over lanes)
  const double *closure;  // [batch][2][nd][lane] M^-1-form momentum source, or nullptr
};

struct StageArgs
{
  const double *src_h, *src_q, *src_v;
  double *dst;               // L / M^-1 L output (ki[stage+1] or apply / rhs output)
  double *scratch;           // optional ki[0] output
  const double *kprev[MAXKI];// ki[1..stage] for the RK combination
  double factor[MAXKI];
  int stage;                 // number of previous k vectors = stage
  // FP32 paired stage inside the device time loop: rk_dst receives the next stage input in
  // its centred FP32 form instead of FP64 (c32_values, c32_centres)
  int out32;
  const double *wsol;        // solution (w_h or w_q)
  double *rk_dst;            // r or w
  unsigned int *sweeps_blk;  // per-batch sum of Jacobi sweeps (height kernels)
};

// the arguments of one multi-operator stage launch (dg_stage_multi.hip)
constexpr int MAXG = OC_MULTI_MAX_OPERATORS;
struct MultiArgs
{
  int n_groups;
  int start[MAXG + 1];  // first workgroup of each group; start[n_groups] = grid size
  int degree[MAXG];
  OpArgs O[MAXG];
  StageArgs H[MAXG], Q[MAXG];
};

__host__ __device__ constexpr int nq_of(int p) { return (3 * p) / 2 + 1; }

template <int N1>
__host__ __device__ constexpr int face_dof(int f, int k)
{
  return f == 0 ? N1 * k : f == 1 ? (N1 - 1) + N1 * k : f == 2 ? k : k + N1 * (N1 - 1);
}

// block -> batch so that each XCD (blocks are dealt round-robin over 8 XCDs) works on a
// contiguous range of batches, keeping neighbour traces in its own L2
__device__ __forceinline__ int logical_batch(int b, int nb, int swz)
{
  if (!swz)
    return b;
  const int q = nb / NXCD, r = nb % NXCD, x = b % NXCD, i = b / NXCD;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

__device__ __forceinline__ size_t vix(int batch, int ncomp, int comp, int nd, int k, int lane)
{
  return (((size_t)batch * ncomp + comp) * nd + k) * CB + lane;
}

// Keeps the scheduler from hoisting the loads of a later phase (the next face's
// neighbour traces) above the current one, which would keep them all live at once.
__device__ __forceinline__ void phase_barrier()
{
  __builtin_amdgcn_sched_barrier(0);
  asm volatile("" ::: "memory");
}

// Placed after a group of independent loads, before their first use: the machine scheduler
// otherwise pairs each load with its use (in FP32 code with its conversion, to save
// registers), so the wave waits out one memory latency per load instead of one per group.
__device__ __forceinline__ void issue_barrier() { __builtin_amdgcn_sched_barrier(0); }

// index of the calling wave in its workgroup, as a wave-uniform (scalar) value: the
// compiler cannot tell that threadIdx.x / 64 is the same on every lane, and a branch on it
// would be compiled as divergent (both sides' live ranges overlapping, values in VGPRs)
__device__ __forceinline__ int wave_in_block()
{
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / CB));
}

// The centred FP32 form of an intermediate stage input (the FP32 stage's integrals read it
// only so): in the storage of the FP64 vector, the values as floats at the float offset
// vix(...) (the first half of the storage; zeta minus its cell's DoF-0 value, q as they
// are), and for zeta each cell's centre, its DoF-0 value in FP64, after them.  The
// integrals then see the same FP32 numbers as when they convert the FP64 vector.
__device__ __forceinline__ const float *c32_values(const double *v)
{
  return reinterpret_cast<const float *>(v);
}
__device__ __forceinline__ const double *c32_centres(const double *v, int n_batches, int nd)
{
  return v + (size_t)n_batches * nd * CB / 2;
}

// boundary data of the running stage (see OpArgs::bvals)
__device__ __forceinline__ const double *stage_bvals(const OpArgs &O)
{
  return O.btick ? O.bvals + (size_t)(O.btick[0] + O.bstage) * (size_t)O.bstride : O.bvals;
}

__device__ __forceinline__ double fld(const Field &f, size_t i)
{
  return f.ptr ? f.ptr[i] : f.cval;
}

// A field copied into VGPRs ahead of a hot loop.  Its pointer and value are wave-uniform
// kernel arguments; the stage kernels hold more uniform values than fit in the SGPRs, and
// the compiler then parks the argument block in VGPR lanes and restores a whole group of
// 16 SGPRs (one v_readlane each) at every read of a field inside a loop.  Copied into
// VGPRs once, a read is a select (or the load) and nothing more.
template <typename R>
struct FieldV
{
  const double *ptr;
  R cval;
  bool stored;
};

template <typename R>
__device__ __forceinline__ FieldV<R> in_vgprs(const Field &f)
{
  FieldV<R> v{f.ptr, R(f.cval), f.ptr != nullptr};
  asm volatile("" : "+v"(v.ptr), "+v"(v.cval));
  return v;
}

template <typename R>
__device__ __forceinline__ R fld(const FieldV<R> &f, size_t i)
{
  return f.stored ? R(f.ptr[i]) : f.cval;
}

}  // namespace oceano
