// dg_stage_multi.hip -- one launch for the explicit stage of several operators (the batched
// ensemble of many uneven instances).
//
// The instances of an ensemble are grouped into one operator per degree (a disjoint union
// of their meshes); this kernel advances every group in a single grid: workgroup b belongs
// to the group whose range of workgroups holds it and runs that group's work item (batch,
// kind) at that group's degree, exactly as the per-degree one-pass kernel k_stage does
// (k_stage_multi_pair: workgroup b runs a batch's two items paired, as k_stage_pair does,
// when every operator pairs its items).
// Each group's range starts at a multiple of the XCD count, so the batch -> XCD dealing of
// the per-degree kernel holds inside every group.  The kernel's register and LDS budget is
// that of its largest degree.  No reference counterpart: oceano runs one instance per
// process (OceanDG.h:462-463 batches cells, not instances).
#include "dg_internal.hpp"

namespace oceano {
static __constant__ Basis<double> c_tab[4];
static __constant__ Basis<float> c_tabf[4];
}

#include "dg_launch.hpp"

namespace oceano {

namespace {

// The arguments of all groups are written to device memory by a one-thread kernel ahead of
// the stage kernel (in stream order, so a captured graph replays both): the stage kernel
// reads the group it runs through a pointer, with uniform (scalar) loads.  Indexing the
// by-value argument block by a run-time group index instead makes the compiler copy the
// whole block to per-lane scratch memory.
__global__ void k_put_multi_args(MultiArgs A, MultiArgs *dst)
{
  if (threadIdx.x == 0)
    *dst = A;
}

template <int P, typename R>
__device__ __forceinline__ void multi_item(const OpArgs &O, const StageArgs &H,
                                           const StageArgs &Q, int w, R *lds)
{
  if (item_kind(w))
    discharge_block<P, false, 0, R>(O, Q, w >> 1, lds);
  else
    height_stage_block<P, R>(O, H, w >> 1, reinterpret_cast<float *>(lds));
}

template <typename R>
constexpr int multi_min_waves()
{
  return stage_min_waves<4, false, R>();
}

template <typename R>
__global__ __launch_bounds__(CB, (multi_min_waves<R>())) void k_stage_multi(
  const MultiArgs *__restrict__ A)
{
  __shared__ R lds[discharge_lds_words<4, false>()];
  const int b = blockIdx.x;
  int g = 0;
  while (g + 1 < A->n_groups && b >= A->start[g + 1])
    ++g;
  const OpArgs &O = A->O[g];
  const int n_items = 2 * O.n_batches;
  const int local = b - A->start[g];
  if (local >= n_items)  // padding of the group's range to a multiple of the XCD count
    return;
  const int w = logical_batch(local, n_items, O.xcd_swizzle);
  switch (A->degree[g])
    {
      case 1:
        multi_item<1, R>(O, A->H[g], A->Q[g], w, lds);
        break;
      case 2:
        multi_item<2, R>(O, A->H[g], A->Q[g], w, lds);
        break;
      case 3:
        multi_item<3, R>(O, A->H[g], A->Q[g], w, lds);
        break;
      default:
        multi_item<4, R>(O, A->H[g], A->Q[g], w, lds);
        break;
    }
}

// paired items (k_stage_pair): workgroup = one batch of a group, one wave its height item,
// the other its discharge item (pair_height_wave); the LDS is laid out as k_stage_pair's at degree 4 (the height
// slice of degree 4 is the largest, so every degree's height slice fits in front of the
// discharge slice)
template <int P, typename R>
__device__ __forceinline__ void multi_pair(const OpArgs &O, const StageArgs &H,
                                           const StageArgs &Q, int batch, R *lds)
{
  static_assert(pair_height_words<P, R>() <= pair_height_words<4, R>() &&
                  discharge_lds_words<P, false>() <= discharge_lds_words<4, false>(),
                "paired LDS slices of a lower degree exceed degree 4's");
  R *xch = lds + pair_xch_offset<P, R>();
  R *dlds = lds + pair_height_words<4, R>();
  if (pair_height_wave())
    height_stage_block<P, R, true>(O, H, batch, reinterpret_cast<float *>(lds), xch, dlds);
  else
    discharge_block<P, false, 0, R, true>(O, Q, batch, dlds, xch);
}

template <typename R>
__global__ __launch_bounds__(2 * CB, (multi_min_waves<R>())) void k_stage_multi_pair(
  const MultiArgs *__restrict__ A)
{
  __shared__ __align__(16) R lds[pair_lds_words<4, R>()];
  const int b = blockIdx.x;
  int g = 0;
  while (g + 1 < A->n_groups && b >= A->start[g + 1])
    ++g;
  const OpArgs &O = A->O[g];
  const int local = b - A->start[g];
  if (local >= O.n_batches)  // padding of the group's range to a multiple of the XCD count
    return;
  const int batch = logical_batch(local, O.n_batches, O.xcd_swizzle);
  switch (A->degree[g])
    {
      case 1:
        multi_pair<1, R>(O, A->H[g], A->Q[g], batch, lds);
        break;
      case 2:
        multi_pair<2, R>(O, A->H[g], A->Q[g], batch, lds);
        break;
      case 3:
        multi_pair<3, R>(O, A->H[g], A->Q[g], batch, lds);
        break;
      default:
        multi_pair<4, R>(O, A->H[g], A->Q[g], batch, lds);
        break;
    }
}

}  // namespace

int upload_tables_multi(const DegreeTables *t, const Basis<float> *tf)
{
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tab), t, sizeof(DegreeTables) * 4));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_tabf), tf, sizeof(Basis<float>) * 4));
  return OC_OK;
}

// the stage of n operators (same precision, inviscid, one-pass: outputs distinct from
// inputs) in one launch; O[g] carry each operator's boundary entry; the mixed path's FP64
// depth solve of near-floor cells at degree 4 follows as its own small launch per group
int stage_multi(int n, oc_dg *const *ops, const OpArgs *O, uint32_t stage,
                const double *const *factor, const double *const *ri_h,
                const double *const *ri_q,
                double *const *const *ki_h, double *const *const *ki_q, uint32_t n_ki,
                double *const *w_h, double *const *w_q, double *const *r_h,
                double *const *r_q, hipStream_t s)
{
  MultiArgs A{};
  A.n_groups = n;
  // paired items when every operator asks for them (one workgroup per batch), else one
  // workgroup per work item
  bool paired = true;
  for (int g = 0; g < n; ++g)
    paired = paired && ops[g]->paired_items;
  int start = 0;
  for (int g = 0; g < n; ++g)
    {
      A.start[g] = start;
      A.degree[g] = ops[g]->P;
      A.O[g] = O[g];
      stage_args(O[g], stage, factor[g], ri_h[g], ri_q[g], nullptr, ki_h[g], ki_q[g], n_ki,
                 w_h[g], w_q[g], r_h[g], r_q[g], 0, A.H[g], A.Q[g]);
      const int blocks = paired ? ops[g]->NB : 2 * ops[g]->NB;
      start += (blocks + NXCD - 1) / NXCD * NXCD;
    }
  A.start[n] = start;
  const bool mixed = ops[0]->precision == OC_PRECISION_MIXED;
  MultiArgs *dA = ops[0]->d_multi_args;
  hipLaunchKernelGGL(k_put_multi_args, dim3(1), dim3(CB), 0, s, A, dA);
  HIP_TRY(hipGetLastError());
  if (paired && mixed)
    hipLaunchKernelGGL(k_stage_multi_pair<float>, dim3(start), dim3(2 * CB), 0, s, dA);
  else if (paired)
    hipLaunchKernelGGL(k_stage_multi_pair<double>, dim3(start), dim3(2 * CB), 0, s, dA);
  else if (mixed)
    hipLaunchKernelGGL(k_stage_multi<float>, dim3(start), dim3(CB), 0, s, dA);
  else
    hipLaunchKernelGGL(k_stage_multi<double>, dim3(start), dim3(CB), 0, s, dA);
  HIP_TRY(hipGetLastError());
  if (mixed)
    for (int g = 0; g < n; ++g)
      if (defer_depth_solve(ops[g]->P))
        {
          StageArgs D = A.H[g];
          for (uint32_t j = 0; j <= stage; ++j)
            D.kprev[j] = ki_h[g][j + 1];
          const dim3 dgrid((ops[g]->NB + DEFER_BATCHES - 1) / DEFER_BATCHES);
          hipLaunchKernelGGL((k_depth_deferred<4>), dgrid, dim3(CB), 0, s, O[g], D);
          HIP_TRY(hipGetLastError());
        }
  return OC_OK;
}

}  // namespace oceano
