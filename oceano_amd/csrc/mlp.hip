// mlp.hip -- training and inference of the closure network (BF16 MFMA, FP32 accumulation).
//
// Training step = k_mlp_grad (forward of 16-sample tiles, MSE loss, backward; each block
// accumulates its weight gradients in LDS and writes one partial) + k_mlp_reduce (sum of
// the partials in block order: deterministic) + k_mlp_adam (Adam in FP32 on the master
// parameters, then the BF16 copies the products read).
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "mlp_core.hpp"
#include "oceano_dg.h"
#include "oceano_mlp.h"

namespace oceano {
void set_error(const std::string &msg);
}

using namespace oceano::mlp;
using oceano::set_error;

struct oc_mlp
{
  int n_in = 0, n_out = 0;
  float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *norm = nullptr;
  float *partial = nullptr, *lossp = nullptr;
  bf16 *wt = nullptr, *wb = nullptr;
  int64_t step = 0;
  int gmax = 256;
};

namespace {

#define MLP_TRY(x)                                                   \
  do                                                                 \
    {                                                                \
      const hipError_t e_ = (x);                                     \
      if (e_ != hipSuccess)                                          \
        {                                                            \
          set_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
          return OC_ERR_HIP;                                         \
        }                                                            \
    }                                                                \
  while (0)

constexpr int WB0 = W2, NWB = B4 - W2;  // backward weights W2..W4 (row-major, with bias gaps)

__global__ __launch_bounds__(64) void k_mlp_grad(const float *__restrict__ p,
                                                 const bf16 *__restrict__ wt_g,
                                                 const bf16 *__restrict__ wb_g,
                                                 const float *__restrict__ norm, int64_t n,
                                                 int n_in, int n_out,
                                                 const float *__restrict__ X,
                                                 const float *__restrict__ T, float gscale,
                                                 float *__restrict__ partial,
                                                 float *__restrict__ lossp)
{
  __shared__ __attribute__((aligned(16))) bf16 s_wt[NT];
  __shared__ __attribute__((aligned(16))) bf16 s_wb[NWB];
  __shared__ float s_g[NPARAM];
  __shared__ __attribute__((aligned(16))) bf16 s_x[16 * IN];
  __shared__ __attribute__((aligned(16))) bf16 s_h[3][16 * HID];
  __shared__ float s_z[3 * 16 * HID];
  __shared__ float s_y[16 * OUT], s_t[16 * OUT];
  __shared__ __attribute__((aligned(16))) bf16 s_d[2][16 * HID];
  __shared__ float s_df[2][16 * HID];
  __shared__ float s_norm[NNORM];
  __shared__ float s_loss[64];
  const int l = threadIdx.x, r = l & 15, kq = (l >> 4) * 4, k8 = (l >> 4) * 8;
  for (int i = l; i < NT; i += 64)
    s_wt[i] = wt_g[i];
  for (int i = l; i < NWB; i += 64)
    s_wb[i] = wb_g[WB0 + i];
  for (int i = l; i < NPARAM; i += 64)
    s_g[i] = 0.f;
  for (int i = l; i < NNORM; i += 64)
    s_norm[i] = norm[i];
  float lossacc = 0.f;
  __syncthreads();
  const int64_t n_chunks = (n + 15) / 16;
  for (int64_t c = blockIdx.x; c < n_chunks; c += gridDim.x)
    {
      const int64_t base = c * 16;
      for (int e = l; e < 16 * IN; e += 64)
        {
          const int s = e / IN, k = e % IN;
          const int64_t row = base + s;
          float x = 0.f;
          if (row < n && k < n_in)
            x = (X[row * n_in + k] - s_norm[k]) * s_norm[IN + k];
          s_x[e] = to_bf16(x);
        }
      for (int e = l; e < 16 * OUT; e += 64)
        {
          const int s = e / OUT, o = e % OUT;
          const int64_t row = base + s;
          s_t[e] = (row < n && o < n_out)
                     ? (T[row * n_out + o] - s_norm[2 * IN + OUT + o]) / s_norm[2 * IN + o]
                     : 0.f;
        }
      __syncthreads();
      forward_tile(s_x, s_wt, p, s_h[0], s_h[1], s_h[2], s_z, s_y);
      // dL/dy of the mean squared error over the n x n_out valid outputs
      for (int e = l; e < 16 * OUT; e += 64)
        {
          const int s = e / OUT, o = e % OUT;
          const bool valid = base + s < n && o < n_out;
          const float d = valid ? s_y[e] - s_t[e] : 0.f;
          lossacc += d * d;
          const float g = d * gscale;
          s_df[0][e] = g;
          s_d[0][e] = to_bf16(g);
        }
      __syncthreads();
      // output layer: db4, dW4 = H2^T dY, dH2 = dY W4^T
      if (l < OUT)
        {
          float sb = 0.f;
          for (int s = 0; s < 16; ++s)
            sb += s_df[0][s * OUT + l];
          s_g[B4 + l] += sb;
        }
#pragma unroll
      for (int rt = 0; rt < HID / 16; ++rt)
#pragma unroll
        for (int ct = 0; ct < OUT / 16; ++ct)
          {
            bf16x4 a, b;
#pragma unroll
            for (int j = 0; j < 4; ++j)
              {
                a[j] = (short)s_h[2][(kq + j) * HID + 16 * rt + r];
                b[j] = (short)s_d[0][(kq + j) * OUT + 16 * ct + r];
              }
            const f32x4 acc = mfma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int j = 0; j < 4; ++j)
              s_g[W4 + (16 * rt + kq + j) * OUT + 16 * ct + r] += acc[j];
          }
#pragma unroll
      for (int ct = 0; ct < HID / 16; ++ct)
        {
          f32x4 acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kk = 0; kk < OUT / 32; ++kk)
            acc = mfma32(load8(s_d[0] + r * OUT + 32 * kk + k8),
                         load8(s_wb + (W4 - WB0) + (16 * ct + r) * OUT + 32 * kk + k8), acc);
#pragma unroll
          for (int j = 0; j < 4; ++j)
            {
              const int s = kq + j, u = 16 * ct + r;
              const float dz = acc[j] * gelu_prime(s_z[2 * 16 * HID + s * HID + u]);
              s_df[1][s * HID + u] = dz;
              s_d[1][s * HID + u] = to_bf16(dz);
            }
        }
      __syncthreads();
      // hidden layers 3 and 2: db, dW = H_in^T dZ, dZ_in = (dZ W^T) * gelu'(Z_in)
#pragma unroll
      for (int layer = 2; layer >= 1; --layer)
        {
          const int cur = layer == 2 ? 1 : 0, nxt = 1 - cur;
          const int Wo = layer == 2 ? W3 : W2, Bo = layer == 2 ? B3 : B2;
          const bf16 *Hin = s_h[layer - 1];
          {
            float sb = 0.f;
            for (int s = 0; s < 16; ++s)
              sb += s_df[cur][s * HID + l];
            s_g[Bo + l] += sb;
          }
#pragma unroll
          for (int rt = 0; rt < HID / 16; ++rt)
#pragma unroll
            for (int ct = 0; ct < HID / 16; ++ct)
              {
                bf16x4 a, b;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                  {
                    a[j] = (short)Hin[(kq + j) * HID + 16 * rt + r];
                    b[j] = (short)s_d[cur][(kq + j) * HID + 16 * ct + r];
                  }
                const f32x4 acc = mfma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int j = 0; j < 4; ++j)
                  s_g[Wo + (16 * rt + kq + j) * HID + 16 * ct + r] += acc[j];
              }
#pragma unroll
          for (int ct = 0; ct < HID / 16; ++ct)
            {
              f32x4 acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
              for (int kk = 0; kk < HID / 32; ++kk)
                acc = mfma32(load8(s_d[cur] + r * HID + 32 * kk + k8),
                             load8(s_wb + (Wo - WB0) + (16 * ct + r) * HID + 32 * kk + k8),
                             acc);
#pragma unroll
              for (int j = 0; j < 4; ++j)
                {
                  const int s = kq + j, u = 16 * ct + r;
                  const float dz =
                    acc[j] * gelu_prime(s_z[(layer - 1) * 16 * HID + s * HID + u]);
                  s_df[nxt][s * HID + u] = dz;
                  s_d[nxt][s * HID + u] = to_bf16(dz);
                }
            }
          __syncthreads();
        }
      // first layer (dZ1 in buffer 1): db1, dW1 = X^T dZ1
      {
        float sb = 0.f;
        for (int s = 0; s < 16; ++s)
          sb += s_df[1][s * HID + l];
        s_g[B1 + l] += sb;
      }
#pragma unroll
      for (int rt = 0; rt < IN / 16; ++rt)
#pragma unroll
        for (int ct = 0; ct < HID / 16; ++ct)
          {
            bf16x4 a, b;
#pragma unroll
            for (int j = 0; j < 4; ++j)
              {
                a[j] = (short)s_x[(kq + j) * IN + 16 * rt + r];
                b[j] = (short)s_d[1][(kq + j) * HID + 16 * ct + r];
              }
            const f32x4 acc = mfma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int j = 0; j < 4; ++j)
              s_g[W1 + (16 * rt + kq + j) * HID + 16 * ct + r] += acc[j];
          }
      __syncthreads();
    }
  for (int i = l; i < NPARAM; i += 64)
    partial[(size_t)blockIdx.x * NPARAM + i] = s_g[i];
  s_loss[l] = lossacc;
  __syncthreads();
  if (l == 0)
    {
      float s = 0.f;
      for (int i = 0; i < 64; ++i)
        s += s_loss[i];
      lossp[blockIdx.x] = s;
    }
}

__global__ void k_mlp_reduce(const float *__restrict__ partial, int G, float *__restrict__ grad)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM)
    return;
  float s = 0.f;
  for (int b = 0; b < G; ++b)
    s += partial[(size_t)b * NPARAM + i];
  grad[i] = s;
}

__device__ inline int transposed_index(int i)
{
  if (i >= W1 && i < B1)
    return T1 + ((i - W1) % HID) * IN + (i - W1) / HID;
  if (i >= W2 && i < B2)
    return T2 + ((i - W2) % HID) * HID + (i - W2) / HID;
  if (i >= W3 && i < B3)
    return T3 + ((i - W3) % HID) * HID + (i - W3) / HID;
  if (i >= W4 && i < B4)
    return T4 + ((i - W4) % OUT) * HID + (i - W4) / OUT;
  return -1;
}

__global__ void k_mlp_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ grad, bf16 *__restrict__ wt,
                           bf16 *__restrict__ wb, float lr_t, float beta1, float beta2,
                           float eps)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM)
    return;
  const float g = grad[i];
  const float mi = beta1 * m[i] + (1.f - beta1) * g;
  const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float w = p[i] - lr_t * mi / (sqrtf(vi) + eps);
  p[i] = w;
  wb[i] = to_bf16(w);
  const int t = transposed_index(i);
  if (t >= 0)
    wt[t] = to_bf16(w);
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ p,
                                                     const bf16 *__restrict__ wt_g,
                                                     const float *__restrict__ norm,
                                                     int64_t n, int n_in, int n_out,
                                                     const float *__restrict__ X,
                                                     float *__restrict__ Y)
{
  __shared__ __attribute__((aligned(16))) bf16 s_wt[NT];
  __shared__ __attribute__((aligned(16))) bf16 s_x[4][16 * IN];
  __shared__ __attribute__((aligned(16))) bf16 s_h[4][3][16 * HID];
  __shared__ float s_y[4][16 * OUT];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int i = threadIdx.x; i < NT; i += 256)
    s_wt[i] = wt_g[i];
  __syncthreads();
  for (int tile = 0; tile < 4; ++tile)
    {
      const int64_t base = (int64_t)blockIdx.x * 256 + (int64_t)(w * 4 + tile) * 16;
      for (int e = l; e < 16 * IN; e += 64)
        {
          const int s = e / IN, k = e % IN;
          const int64_t row = base + s;
          float x = 0.f;
          if (row < n && k < n_in)
            x = (X[row * n_in + k] - norm[k]) * norm[IN + k];
          s_x[w][e] = to_bf16(x);
        }
      __syncthreads();
      forward_tile(s_x[w], s_wt, p, s_h[w][0], s_h[w][1], s_h[w][2], nullptr, s_y[w]);
      for (int e = l; e < 16 * OUT; e += 64)
        {
          const int s = e / OUT, o = e % OUT;
          const int64_t row = base + s;
          if (row < n && o < n_out)
            Y[row * n_out + o] = output_value(s_y[w][e], norm, o);
        }
      __syncthreads();
    }
}

void host_copies(const std::vector<float> &p, std::vector<bf16> &wt, std::vector<bf16> &wb)
{
  wt.assign(NT, 0);
  wb.assign(NPARAM, 0);
  for (int i = 0; i < NPARAM; ++i)
    wb[i] = to_bf16(p[i]);
  auto put = [&](int W, int K, int N, int T) {
    for (int k = 0; k < K; ++k)
      for (int n = 0; n < N; ++n)
        wt[T + n * K + k] = to_bf16(p[W + k * N + n]);
  };
  put(W1, IN, HID, T1);
  put(W2, HID, HID, T2);
  put(W3, HID, HID, T3);
  put(W4, HID, OUT, T4);
}

int upload_params(oc_mlp *m, const std::vector<float> &p)
{
  std::vector<bf16> wt, wb;
  host_copies(p, wt, wb);
  MLP_TRY(hipMemcpy(m->p, p.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(m->wt, wt.data(), NT * sizeof(bf16), hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(m->wb, wb.data(), NPARAM * sizeof(bf16), hipMemcpyHostToDevice));
  return OC_OK;
}

std::vector<float> default_norm()
{
  std::vector<float> nrm(NNORM, 0.f);
  for (int i = 0; i < IN; ++i)
    nrm[IN + i] = 1.f;
  for (int o = 0; o < OUT; ++o)
    {
      nrm[2 * IN + o] = 1.f;
      nrm[NLO + o] = -INFINITY;
      nrm[NHI + o] = INFINITY;
    }
  return nrm;
}

void release(oc_mlp *m)
{
  for (float *q : {m->p, m->m, m->v, m->grad, m->norm, m->partial, m->lossp})
    (void)hipFree(q);
  (void)hipFree(m->wt);
  (void)hipFree(m->wb);
}

int grad_launch(oc_mlp *m, int64_t n, const float *x, const float *t, hipStream_t s, int *G)
{
  const int64_t chunks = (n + 15) / 16;
  *G = (int)std::min<int64_t>(chunks, m->gmax);
  const float gscale = 2.0f / (float)((double)n * (double)m->n_out);
  hipLaunchKernelGGL(k_mlp_grad, dim3(*G), dim3(64), 0, s, m->p, m->wt, m->wb, m->norm, n,
                     m->n_in, m->n_out, x, t, gscale, m->partial, m->lossp);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0, s, m->partial, *G,
                     m->grad);
  MLP_TRY(hipGetLastError());
  return OC_OK;
}

}  // namespace

extern "C" {

int oc_mlp_create(int32_t n_in, int32_t n_out, uint64_t seed, oc_mlp **out)
{
  if (!out || n_in < 1 || n_in > IN || n_out < 1 || n_out > OUT)
    {
      set_error("oc_mlp_create: n_in must be in 1..32, n_out in 1..32");
      return OC_ERR_ARGUMENT;
    }
  oc_mlp *m = new oc_mlp;
  m->n_in = n_in;
  m->n_out = n_out;
  bool ok = true;
  for (float **q : {&m->p, &m->m, &m->v, &m->grad})
    ok = ok && hipMalloc(q, NPARAM * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->norm, NNORM * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->partial, (size_t)m->gmax * NPARAM * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->lossp, m->gmax * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&m->wt, NT * sizeof(bf16)) == hipSuccess;
  ok = ok && hipMalloc(&m->wb, NPARAM * sizeof(bf16)) == hipSuccess;
  ok = ok && hipMemset(m->m, 0, NPARAM * sizeof(float)) == hipSuccess &&
       hipMemset(m->v, 0, NPARAM * sizeof(float)) == hipSuccess;
  if (!ok)
    {
      release(m);
      delete m;
      set_error("oc_mlp_create: device allocation failed");
      return OC_ERR_HIP;
    }
  // Glorot-uniform weights on the used block of each layer, zero pads and biases
  std::mt19937_64 eng(seed);
  auto u01 = [&]() { return double(eng() >> 11) * (1.0 / 9007199254740992.0); };
  std::vector<float> p(NPARAM, 0.f);
  auto fill = [&](int W, int K, int N, int k_used, int n_used) {
    const double lim = std::sqrt(6.0 / (k_used + n_used));
    for (int k = 0; k < k_used; ++k)
      for (int n = 0; n < n_used; ++n)
        p[W + k * N + n] = float((2.0 * u01() - 1.0) * lim);
  };
  fill(W1, IN, HID, n_in, HID);
  fill(W2, HID, HID, HID, HID);
  fill(W3, HID, HID, HID, HID);
  fill(W4, HID, OUT, HID, n_out);
  std::vector<float> nrm = default_norm();
  int rc = upload_params(m, p);
  if (!rc && hipMemcpy(m->norm, nrm.data(), NNORM * sizeof(float), hipMemcpyHostToDevice) !=
               hipSuccess)
    rc = OC_ERR_HIP;
  if (rc)
    {
      release(m);
      delete m;
      return rc;
    }
  *out = m;
  return OC_OK;
}

void oc_mlp_destroy(oc_mlp *m)
{
  if (!m)
    return;
  release(m);
  delete m;
}

int32_t oc_mlp_n_params(void) { return NPARAM; }

int oc_mlp_get_params(oc_mlp *m, float *params)
{
  if (!m || !params)
    return OC_ERR_ARGUMENT;
  MLP_TRY(hipMemcpy(params, m->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return OC_OK;
}

int oc_mlp_set_params(oc_mlp *m, const float *params)
{
  if (!m || !params)
    {
      set_error("oc_mlp_set_params: null argument");
      return OC_ERR_ARGUMENT;
    }
  std::vector<float> p(params, params + NPARAM);
  // pads of the unused inputs / outputs stay zero
  for (int k = m->n_in; k < IN; ++k)
    for (int n = 0; n < HID; ++n)
      p[W1 + k * HID + n] = 0.f;
  for (int k = 0; k < HID; ++k)
    for (int o = m->n_out; o < OUT; ++o)
      p[W4 + k * OUT + o] = 0.f;
  for (int o = m->n_out; o < OUT; ++o)
    p[B4 + o] = 0.f;
  MLP_TRY(hipMemset(m->m, 0, NPARAM * sizeof(float)));
  MLP_TRY(hipMemset(m->v, 0, NPARAM * sizeof(float)));
  m->step = 0;
  return upload_params(m, p);
}

int oc_mlp_set_normalization(oc_mlp *m, const float *x_mean, const float *x_std,
                             const float *y_shift, const float *y_scale)
{
  if (!m || !x_mean || !x_std || !y_shift || !y_scale)
    {
      set_error("oc_mlp_set_normalization: null argument");
      return OC_ERR_ARGUMENT;
    }
  std::vector<float> nrm = default_norm();
  // keep the output range of an earlier oc_mlp_set_output_range
  MLP_TRY(hipMemcpy(nrm.data() + NLO, m->norm + NLO, 2 * OUT * sizeof(float),
                    hipMemcpyDeviceToHost));
  for (int i = 0; i < m->n_in; ++i)
    {
      if (!(x_std[i] > 0.f))
        {
          set_error("oc_mlp_set_normalization: x_std must be positive");
          return OC_ERR_ARGUMENT;
        }
      nrm[i] = x_mean[i];
      nrm[IN + i] = 1.f / x_std[i];
    }
  for (int o = 0; o < m->n_out; ++o)
    {
      if (!(y_scale[o] > 0.f))
        {
          set_error("oc_mlp_set_normalization: y_scale must be positive");
          return OC_ERR_ARGUMENT;
        }
      nrm[2 * IN + o] = y_scale[o];
      nrm[2 * IN + OUT + o] = y_shift[o];
    }
  MLP_TRY(hipMemcpy(m->norm, nrm.data(), NNORM * sizeof(float), hipMemcpyHostToDevice));
  return OC_OK;
}

int oc_mlp_set_output_range(oc_mlp *m, const float *lo, const float *hi)
{
  if (!m)
    {
      set_error("oc_mlp_set_output_range: null argument");
      return OC_ERR_ARGUMENT;
    }
  std::vector<float> r(2 * OUT);
  for (int o = 0; o < OUT; ++o)
    {
      r[o] = (lo && o < m->n_out) ? lo[o] : -INFINITY;
      r[OUT + o] = (hi && o < m->n_out) ? hi[o] : INFINITY;
      if (!(r[o] <= r[OUT + o]))
        {
          set_error("oc_mlp_set_output_range: lo > hi");
          return OC_ERR_ARGUMENT;
        }
    }
  MLP_TRY(hipMemcpy(m->norm + NLO, r.data(), 2 * OUT * sizeof(float), hipMemcpyHostToDevice));
  return OC_OK;
}

int oc_mlp_forward(oc_mlp *m, int64_t n, const float *x, float *y, void *stream)
{
  if (!m || n < 0 || (n > 0 && (!x || !y)))
    {
      set_error("oc_mlp_forward: null argument");
      return OC_ERR_ARGUMENT;
    }
  if (n == 0)
    return OC_OK;
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, m->p, m->wt, m->norm, n, m->n_in, m->n_out, x, y);
  MLP_TRY(hipGetLastError());
  return OC_OK;
}

int oc_mlp_gradients(oc_mlp *m, int64_t n, const float *x, const float *t, float *grad,
                     double *loss, void *stream)
{
  if (!m || n < 1 || !x || !t)
    {
      set_error("oc_mlp_gradients: null argument or empty batch");
      return OC_ERR_ARGUMENT;
    }
  hipStream_t s = (hipStream_t)stream;
  int G = 0;
  int rc = grad_launch(m, n, x, t, s, &G);
  if (rc)
    return rc;
  if (grad)
    {
      MLP_TRY(hipMemcpyAsync(grad, m->grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s));
    }
  if (loss)
    {
      std::vector<float> lp(G);
      MLP_TRY(hipMemcpyAsync(lp.data(), m->lossp, G * sizeof(float), hipMemcpyDeviceToHost, s));
      MLP_TRY(hipStreamSynchronize(s));
      double acc = 0.0;
      for (int b = 0; b < G; ++b)
        acc += lp[b];
      *loss = acc / ((double)n * m->n_out);
    }
  else if (grad)
    MLP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

int oc_mlp_adam(oc_mlp *m, float lr, float beta1, float beta2, float eps, void *stream)
{
  if (!m || !(lr >= 0.f))
    {
      set_error("oc_mlp_adam: null argument or negative learning rate");
      return OC_ERR_ARGUMENT;
    }
  m->step += 1;
  const double t = (double)m->step;
  const float lr_t =
    float(lr * std::sqrt(1.0 - std::pow((double)beta2, t)) / (1.0 - std::pow((double)beta1, t)));
  hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     m->p, m->m, m->v, m->grad, m->wt, m->wb, lr_t, beta1, beta2, eps);
  MLP_TRY(hipGetLastError());
  return OC_OK;
}

int oc_mlp_train_step(oc_mlp *m, int64_t n, const float *x, const float *t, float lr,
                      double *loss, void *stream)
{
  int rc = oc_mlp_gradients(m, n, x, t, nullptr, loss, stream);
  if (rc)
    return rc;
  return oc_mlp_adam(m, lr, 0.9f, 0.999f, 1e-8f, stream);
}

int oc_mlp_dims(const oc_mlp *m, int32_t *n_in, int32_t *n_out)
{
  if (!m)
    return OC_ERR_ARGUMENT;
  if (n_in)
    *n_in = m->n_in;
  if (n_out)
    *n_out = m->n_out;
  return OC_OK;
}

// device pointers for the in-stage inference (library internal)
void oc_mlp_device_state(const oc_mlp *m, const float **p, const uint16_t **wt,
                         const float **norm)
{
  *p = m->p;
  *wt = m->wt;
  *norm = m->norm;
}

}  // extern "C"
