// dg_internal.hpp -- operator state shared by the host translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "dg_kernels.hpp"
#include "oceano_dg.h"
#include "oceano_mlp.h"

namespace oceano {
void set_error(const std::string &msg);
}

struct oc_dg
{
  using OpArgs = oceano::OpArgs;
  int device = 0;  // the GPU current at oc_dg_create
  int P, n1, nd, nq, ng;
  int N, NB;
  int precision;
  double g, cu4, nu;
  int max_it;
  std::vector<double> geom;
  std::vector<int32_t> inner;     // canonical (minus side owns face 1 or 3)
  std::vector<char> swapped;      // inner face listed from its plus side by the caller
  std::vector<int32_t> inner_in;  // as given
  std::vector<int32_t> bdry;
  std::map<int, int> bc_of_id;
  // cell slots: the kernels see cell order[i] in slot i (batch i / 64, lane i % 64);
  // pos is the inverse.  Every host-facing array stays in the caller's cell order.
  std::vector<int32_t> order, pos;
  // device
  double *d_hxy = nullptr;
  int *d_order = nullptr;
  int *d_nbr = nullptr, *d_fid = nullptr, *d_finfo = nullptr, *d_bkind = nullptr;
  double *d_cell0[5] = {};
  double c_cell0[5] = {};
  double *d_cell1 = nullptr;
  double c_cell1 = 0;
  double *d_cell2 = nullptr;
  double c_cell2 = 0;
  double *d_cell2n = nullptr;
  double c_cell2n = 0;
  double *d_face = nullptr;
  double c_face = 0;
  double *d_bz = nullptr;
  double c_bz = 0;
  // boundary data, one entry [n_bdry][nq][3] per tabulated time (oc_dg_set_boundary_table);
  // bt_any: a single entry valid at every time (oc_dg_set_boundary_values)
  double *d_btable = nullptr;
  std::vector<double> bt_times;
  bool bt_any = true;
  bool needs_bdata = false;     // some boundary face reads boundary data
  const int *capture_tick = nullptr;  // set while a time loop records a graph
  // bumped by every setter that replaces device data a captured graph may hold
  uint64_t generation = 0;
  unsigned int *d_sweeps_blk = nullptr;
  float *d_jmat = nullptr;  // mixed precision: closed-form Jacobi matrix (unit cell)
  float *d_kinv = nullptr;  // mixed precision: momentum inverse mass (unit cell)
  unsigned long long *d_defer = nullptr;  // mixed precision: per batch, deferred lanes
  double *d_post = nullptr; // per-batch results of the reductions [NB]
  // arguments of a multi-operator stage launch this operator leads (written by a kernel)
  oceano::MultiArgs *d_multi_args = nullptr;
  double wet_margin = 10.0;
  // learned closure (oc_dg_set_closure): network (not owned), per-cell sub-grid statistics
  // [batch][3][lane], modal<->nodal tables [nd][nd], correction buffer [batch][2][nd][lane]
  const oc_mlp *closure_mlp = nullptr;
  double *d_cstats = nullptr, *d_cvi = nullptr, *d_cv = nullptr, *d_closure = nullptr;
  unsigned long long *d_count = nullptr;
  std::vector<double> cell2_host;
  std::vector<double> zb_gauss_host;  // z_b at the OC_STORE_CELL_GAUSS points
  int xcd_swizzle = 1;
  int waves_per_block = 1;  // one-pass stage kernel: 64-cell items per workgroup (1, 2, 4)
  int paired_items = 0;     // inviscid one-pass stage: k_stage_pair (faces split per batch)
  // set by the device time loop (oc_rk) around each stage: the stage input is in its centred
  // FP32 form (rk_in32), the next stage input is to be written so (rk_out32); honoured only
  // by the FP32 paired stage at p <= 3, whose integrals read the input only as FP32 numbers
  int rk_in32 = 0, rk_out32 = 0;
  // optional per-kernel event timing: kernel id -> recorded (start, stop) pairs
  int timing = 0;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> events;
  std::vector<hipEvent_t> event_pool;

  OpArgs args() const
  {
    OpArgs o;
    o.n_cells = N;
    o.n_batches = NB;
    o.hxy = d_hxy;
    o.nbr = d_nbr;
    o.fid = d_fid;
    o.finfo = d_finfo;
    for (int f = 0; f < 5; ++f)
      o.cell0[f] = {d_cell0[f], c_cell0[f]};
    o.cell1 = {d_cell1, c_cell1};
    o.cell2 = {d_cell2, c_cell2};
    o.cell2n = {d_cell2n, c_cell2n};
    o.facez = {d_face, c_face};
    o.bdryz = {d_bz, c_bz};
    o.bvals = d_btable;
    o.btick = nullptr;
    o.bstage = 0;
    o.bstride = (long long)(bdry.size() / 3) * nq * 3;
    o.bkind = d_bkind;
    o.g = g;
    o.cu4 = cu4;
    o.nu = nu;
    o.has_friction = (d_cell0[1] != nullptr) || (c_cell0[1] != 0.0);
    o.has_wind = (d_cell0[2] != nullptr) || (c_cell0[2] != 0.0) || (d_cell0[3] != nullptr) ||
                 (c_cell0[3] != 0.0);
    o.has_coriolis = (d_cell0[4] != nullptr) || (c_cell0[4] != 0.0);
    o.max_it = max_it;
    o.xcd_swizzle = xcd_swizzle;
    o.jmat = d_jmat;
    o.kinv = d_kinv;
    o.wet_margin = wet_margin;
    o.defer = d_defer;
    o.sweeps = d_sweeps_blk;
    o.closure = nullptr;  // set per stage when a closure network is attached
    return o;
  }
};


namespace oceano {

#define HIP_TRY(call)                                                                       \
  do                                                                                        \
    {                                                                                       \
      hipError_t e_ = (call);                                                               \
      if (e_ != hipSuccess)                                                                 \
        {                                                                                   \
          set_error(std::string(#call) + ": " + hipGetErrorString(e_));                     \
          return OC_ERR_HIP;                                                                \
        }                                                                                   \
    }                                                                                       \
  while (0)

// per-degree translation units (dg_stage_p<P>.hip)
#define OC_DECLARE_DEGREE(P)                                                                \
  int upload_tables_d##P(const DegreeTables *t, const Basis<float> *tf);                  \
  int stage_d##P(oc_dg *op, const OpArgs &O, uint32_t stage, const double *factor,          \
                 const double *ri_h, const double *ri_q, const double *ri_v,                \
                 double *const *ki_h, double *const *ki_q, uint32_t n_ki, double *w_h,      \
                 double *w_q, double *r_h, double *r_q, int write_scratch, hipStream_t s);  \
  int rhs_apply_d##P(oc_dg *op, const OpArgs &O, bool invert, const double *src_h,          \
                     const double *src_q, const double *src_v, double *dst_h,              \
                     double *dst_q, hipStream_t s);                                        \
  int post_d##P(oc_dg *op, const OpArgs &O, int what, const double *zeta, const double *q,  \
                double *out, hipStream_t s);                                               \
  int closure_d##P(oc_dg *op, const OpArgs &O, const double *zeta, const double *q,         \
                   float *features, hipStream_t s);                                        \
  int stage_ark_d##P(oc_dg *op, const OpArgs &O, uint32_t stage, uint32_t n_stages,         \
                     const double *factor, const double *factor_tilde, double fm,          \
                     const double *ri_h, const double *ri_q, const double *ri_v,           \
                     double *const *ki_h, double *const *ki_q, double *w_h, double *w_q,   \
                     double *r_h, double *r_q, hipStream_t s);
// per-degree post-processing kernels (post_dP)
enum PostKernel
{
  POST_VELOCITY = 0,  // out: velocity field, operator order, 2 components
  POST_SPEED = 1,     // out: per-batch maximum transport speed [NB]
  POST_VOLUME = 2     // out: per-batch integral of the depth [NB]
};
// kernels timed by oc_dg_set_timing
enum TimedKernel
{
  TK_STAGE = 0,         // one-pass stage (height and discharge work in one grid)
  TK_HEIGHT_RHS = 1,    // split path: L_zeta
  TK_DISCHARGE = 2,     // split path: L_q + inverse mass
  TK_RK_UPDATE = 3,     // split path: discharge RK combination
  TK_HEIGHT_JACOBI = 4, // split path: front + depth solve
  TK_COUNT = 5
};
int timed_begin(oc_dg *op, hipStream_t s, hipEvent_t *start);
// points O at the boundary data of a stage at `time` (or, while a time loop records a graph,
// at the table indexed by the loop's device tick + stage); OC_ERR_ARGUMENT when the table
// holds no entry at that time
int boundary_entry(oc_dg *op, double time, uint32_t stage, OpArgs &O, const char *fn);
// the entry of the boundary table at `time`, -1 if none (always 0 for time-independent data)
long find_boundary_time(const oc_dg *op, double time);
void timed_end(oc_dg *op, hipStream_t s, int kernel, hipEvent_t start);

OC_DECLARE_DEGREE(1)
OC_DECLARE_DEGREE(2)
OC_DECLARE_DEGREE(3)
OC_DECLARE_DEGREE(4)

// one launch for the stage of several operators (dg_stage_multi.hip)
int upload_tables_multi(const DegreeTables *t, const Basis<float> *tf);
int stage_multi(int n, oc_dg *const *ops, const OpArgs *O, uint32_t stage,
                const double *const *factor, const double *const *ri_h,
                const double *const *ri_q,
                double *const *const *ki_h, double *const *const *ki_q, uint32_t n_ki,
                double *const *w_h, double *const *w_q, double *const *r_h,
                double *const *r_q, hipStream_t s);

}  // namespace oceano
