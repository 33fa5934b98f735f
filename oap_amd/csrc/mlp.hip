// mlp.hip -- the learned-start network: a multilayer perceptron with tanh
// hidden layers and a linear output, BF16 operands and FP32 accumulation on
// the matrix cores, FP32 master weights, trained with Adam by hand-written
// forward and backward kernels.
//
// It is the batched counterpart of oap's neural routines (forward
// propagation = dotProduct + activation, back propagation = transposed
// weights x errors (.) f', nabla W = errors x inputs^T, the update
// W <- W - lr * nabla; oapNeuralHost/oapGenericNeuralApi.h:195-354), with a
// whole batch as one matrix product per layer:
//   Z_l = H_{l-1} W_l^T + b_l,  H_l = tanh(Z_l) (linear for the last layer)
//   loss = 1/(2B) sum_b sum_j w_j (Y_bj - T_bj)^2
//   dY = w (.) (Y - T) / B,  dW_l = dZ_l^T H_{l-1},  db_l = sum_b dZ_l,
//   dZ_{l-1} = (dZ_l W_l) (.) (1 - H_{l-1}^2).
// Every product whose inner dimension is at least 16 runs on
// v_mfma_f32_16x16x16_bf16 (64 x 64 output tile per 256-thread workgroup,
// 32 x 32 per wave); the 3-wide input layer uses FP32 FMAs.  BF16 values are
// FP32 values rounded to nearest-even (bf16_rne).
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace oaph;

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Epilogues of the GEMM: C = acc (+ bias[n]) then
enum Epi { EPI_STORE = 0, EPI_BIAS_TANH = 1, EPI_BIAS = 2, EPI_DTANH = 3 };

constexpr int kTM = 64, kTN = 64, kTK = 16;

// C[m][n] = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak] (bf16),
// B(k, n) = B[k*sbk + n*sbn] (bf16).  EPI_BIAS_TANH / EPI_BIAS add bias[n];
// EPI_BIAS_TANH stores tanh; EPI_DTANH multiplies by (1 - H[m][n]^2).
// C (fp32, ldc) and, when Cb != nullptr, its bf16 copy (ldc) are written.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A,
                                                   int sam, int sak, const uint16_t* __restrict__ B,
                                                   int sbk, int sbn, float* __restrict__ C,
                                                   uint16_t* __restrict__ Cb, int ldc,
                                                   const float* __restrict__ bias,
                                                   const float* __restrict__ H, int ldh) {
  __shared__ uint16_t As[kTM][kTK + 4];
  __shared__ uint16_t Bs[kTK][kTN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += kTK) {
    // stage 64 x 16 of A and 16 x 64 of B (4 elements per thread each)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int am = idx / kTK, ak = idx % kTK;
      const int gm = m0 + am, gk = k0 + ak;
      As[am][ak] = (gm < M && gk < K) ? A[(size_t)gm * sam + (size_t)gk * sak] : (uint16_t)0;
      const int bk = idx / kTN, bn = idx % kTN;
      const int gk2 = k0 + bk, gn = n0 + bn;
      Bs[bk][bn] = (gk2 < K && gn < N) ? B[(size_t)gk2 * sbk + (size_t)gn * sbn] : (uint16_t)0;
    }
    __syncthreads();
    const int r = lane & 15, kq = (lane >> 4) * 4;
    s4 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int row = wm + 16 * i + r;
      a[i] = s4{(short)As[row][kq], (short)As[row][kq + 1], (short)As[row][kq + 2],
                (short)As[row][kq + 3]};
      const int col = wn + 16 * i + r;
      b[i] = s4{(short)Bs[kq][col], (short)Bs[kq + 1][col], (short)Bs[kq + 2][col],
                (short)Bs[kq + 3][col]};
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // lane holds rows 4*(lane/16) + 0..3 of column lane % 16 of each 16 x 16 tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + q;
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m >= M || n >= N) continue;
        float v = acc[i][j][q];
        if (EPI == EPI_BIAS_TANH || EPI == EPI_BIAS) v += bias[n];
        if (EPI == EPI_BIAS_TANH) v = tanhf(v);
        if (EPI == EPI_DTANH) {
          const float h = H[(size_t)m * ldh + n];
          v *= 1.f - h * h;
        }
        C[(size_t)m * ldc + n] = v;
        if (Cb) Cb[(size_t)m * ldc + n] = bf16_rne(v);
      }
}

// First layer (inner dimension d_in, small): Z = X W^T + b, H = tanh(Z)
__global__ __launch_bounds__(256) void k_input_layer(int B, int din, int dout, const float* __restrict__ X,
                                                     const uint16_t* __restrict__ Wb,
                                                     const float* __restrict__ bias,
                                                     float* __restrict__ H, uint16_t* __restrict__ Hb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * dout) return;
  const int b = i / dout, o = i % dout;
  float z = bias[o];
  for (int k = 0; k < din; ++k)
    z += bf16_to_f(bf16_rne(X[(size_t)b * din + k])) * bf16_to_f(Wb[(size_t)o * din + k]);
  const float h = tanhf(z);
  H[i] = h;
  Hb[i] = bf16_rne(h);
}

// dY = w (.) (Y - T) / B (fp32 and bf16), partial loss per block
__global__ __launch_bounds__(256) void k_output_grad(int B, int dout, const float* __restrict__ Y,
                                                     const float* __restrict__ T,
                                                     const float* __restrict__ w, float* __restrict__ dY,
                                                     uint16_t* __restrict__ dYb, float* __restrict__ loss_part) {
  __shared__ float red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < B * dout) {
    const int o = i % dout;
    const float d = Y[i] - T[i];
    const float g = w[o] * d / (float)B;
    dY[i] = g;
    dYb[i] = bf16_rne(g);
    l = 0.5f * w[o] * d * d / (float)B;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// db[o] = sum_b dZ[b][o] (one thread per output, fixed order)
__global__ __launch_bounds__(256) void k_colsum(int B, int n, const float* __restrict__ dZ,
                                                float* __restrict__ db) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += dZ[(size_t)b * n + o];
  db[o] = s;
}

// dW1[o][k] = sum_b dZ1[b][o] X[b][k] (bf16 operands, fp32 sums)
__global__ __launch_bounds__(256) void k_input_grad(int B, int din, int dout, const uint16_t* __restrict__ dZb,
                                                    const float* __restrict__ X, float* __restrict__ dW) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dout * din) return;
  const int o = i / din, k = i % din;
  float s = 0.f;
  for (int b = 0; b < B; ++b)
    s += bf16_to_f(dZb[(size_t)b * dout + o]) * bf16_to_f(bf16_rne(X[(size_t)b * din + k]));
  dW[i] = s;
}

__global__ __launch_bounds__(256) void k_adam(size_t n, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr,
                                              float b1, float b2, float c1, float c2, float eps) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_to_bf16(size_t n, const float* __restrict__ src,
                                                 uint16_t* __restrict__ dst) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = bf16_rne(src[i]);
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct oaph_mlp {
  std::vector<uint32_t> dims;  // d_0 .. d_L
  size_t nparams = 0;
  std::vector<size_t> w_off, b_off;  // offsets in the flat parameter vector
  uint32_t max_batch = 0;
  hipStream_t stream = nullptr;
  DevBuf params, grads, adam_m, adam_v, wbf;  // fp32 params/grads/moments, bf16 weights
  DevBuf acts, actsb, dz, dzb;                 // per-layer activations and deltas
  DevBuf X, T, outw, loss_part, idx;           // batch staging
  DevBuf data_x, data_t;                       // resident training set
  uint32_t data_n = 0;
  long long step = 0;
  std::vector<size_t> act_off;  // offsets of layer activations (B x d_l) in acts
};

namespace {

oaph_status refresh_bf16(oaph_mlp* m) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(m->nparams)), dim3(256), 0, m->stream, m->nparams,
                     m->params.as<float>(), m->wbf.as<uint16_t>());
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

template <int EPI>
oaph_status gemm(hipStream_t s, int M, int N, int K, const uint16_t* A, int sam, int sak,
                 const uint16_t* B, int sbk, int sbn, float* C, uint16_t* Cb, int ldc,
                 const float* bias, const float* H, int ldh) {
  dim3 grid((N + kTN - 1) / kTN, (M + kTM - 1) / kTM);
  hipLaunchKernelGGL(k_gemm_bf16<EPI>, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn,
                     C, Cb, ldc, bias, H, ldh);
  OAPH_HIP(hipGetLastError());
  return OAPH_OK;
}

// forward of a batch already staged in m->X (B x d_0 fp32)
oaph_status forward(oaph_mlp* m, uint32_t B) {
  const int L = (int)m->dims.size() - 1;
  const float* P = m->params.as<float>();
  const uint16_t* Wb = m->wbf.as<uint16_t>();
  float* act = m->acts.as<float>();
  uint16_t* actb = m->actsb.as<uint16_t>();
  hipLaunchKernelGGL(k_input_layer, dim3(blocks((size_t)B * m->dims[1])), dim3(256), 0, m->stream,
                     (int)B, (int)m->dims[0], (int)m->dims[1], m->X.as<float>(), Wb + m->w_off[0],
                     P + m->b_off[0], act + m->act_off[1], actb + m->act_off[1]);
  OAPH_HIP(hipGetLastError());
  for (int l = 1; l < L; ++l) {
    const int din = (int)m->dims[l], dout = (int)m->dims[l + 1];
    const uint16_t* Hin = actb + m->act_off[l];
    float* Z = act + m->act_off[l + 1];
    uint16_t* Zb = actb + m->act_off[l + 1];
    // Z = H W^T: A(b, k) = H[b*din + k], B(k, o) = W[o*din + k]
    if (l + 1 < L)
      OAPH_TRY(gemm<EPI_BIAS_TANH>(m->stream, (int)B, dout, din, Hin, din, 1, Wb + m->w_off[l], 1,
                                   din, Z, Zb, dout, P + m->b_off[l], nullptr, 0));
    else
      OAPH_TRY(gemm<EPI_BIAS>(m->stream, (int)B, dout, din, Hin, din, 1, Wb + m->w_off[l], 1, din,
                              Z, Zb, dout, P + m->b_off[l], nullptr, 0));
  }
  return OAPH_OK;
}

// gradients of the loss for the batch in m->X / m->T into m->grads; the
// batch loss (sum of the per-block partials) into *loss when non-null
oaph_status backward(oaph_mlp* m, uint32_t B, float* loss) {
  const int L = (int)m->dims.size() - 1;
  const uint16_t* Wb = m->wbf.as<uint16_t>();
  float* act = m->acts.as<float>();
  uint16_t* actb = m->actsb.as<uint16_t>();
  float* G = m->grads.as<float>();
  float* dz = m->dz.as<float>();
  uint16_t* dzb = m->dzb.as<uint16_t>();
  const int dL = (int)m->dims[L];
  const size_t nout = (size_t)B * dL;
  const unsigned lb = blocks(nout);
  hipLaunchKernelGGL(k_output_grad, dim3(lb), dim3(256), 0, m->stream, (int)B, dL,
                     act + m->act_off[L], m->T.as<float>(), m->outw.as<float>(), dz, dzb,
                     m->loss_part.as<float>());
  OAPH_HIP(hipGetLastError());
  // dz / dzb hold dZ_l (B x d_l) for the layer being processed
  for (int l = L; l >= 2; --l) {
    const int dout = (int)m->dims[l], din = (int)m->dims[l - 1];
    // dW_{l-1} (dout x din) = dZ^T H_{l-1}: A(o, b) = dZ[b*dout + o], B(b, k) = H[b*din + k]
    OAPH_TRY(gemm<EPI_STORE>(m->stream, dout, din, (int)B, dzb, 1, dout, actb + m->act_off[l - 1],
                             din, 1, G + m->w_off[l - 1], nullptr, din, nullptr, nullptr, 0));
    hipLaunchKernelGGL(k_colsum, dim3(blocks(dout)), dim3(256), 0, m->stream, (int)B, dout, dz,
                       G + m->b_off[l - 1]);
    OAPH_HIP(hipGetLastError());
    // dZ_{l-1} = (dZ W) (.) (1 - H^2): A(b, o) = dZ[b*dout + o], B(o, k) = W[o*din + k];
    // written to the second half of the delta buffers, then swapped in
    float* dz2 = dz + (size_t)m->max_batch * 512;
    uint16_t* dzb2 = dzb + (size_t)m->max_batch * 512;
    OAPH_TRY(gemm<EPI_DTANH>(m->stream, (int)B, din, dout, dzb, dout, 1, Wb + m->w_off[l - 1], din,
                             1, dz2, dzb2, din, nullptr, act + m->act_off[l - 1], din));
    OAPH_HIP(hipMemcpyAsync(dz, dz2, (size_t)B * din * sizeof(float), hipMemcpyDeviceToDevice,
                            m->stream));
    OAPH_HIP(hipMemcpyAsync(dzb, dzb2, (size_t)B * din * sizeof(uint16_t),
                            hipMemcpyDeviceToDevice, m->stream));
  }
  // first layer: dW_0 = dZ_1^T X, db_0
  const int d1 = (int)m->dims[1], d0 = (int)m->dims[0];
  hipLaunchKernelGGL(k_input_grad, dim3(blocks((size_t)d1 * d0)), dim3(256), 0, m->stream, (int)B,
                     d0, d1, dzb, m->X.as<float>(), G + m->w_off[0]);
  hipLaunchKernelGGL(k_colsum, dim3(blocks(d1)), dim3(256), 0, m->stream, (int)B, d1, dz,
                     G + m->b_off[0]);
  OAPH_HIP(hipGetLastError());
  if (loss) {
    std::vector<float> part(lb);
    OAPH_HIP(hipMemcpyAsync(part.data(), m->loss_part.ptr, lb * sizeof(float),
                            hipMemcpyDeviceToHost, m->stream));
    OAPH_HIP(hipStreamSynchronize(m->stream));
    double s = 0;
    for (float p : part) s += p;
    *loss = (float)s;
  }
  return OAPH_OK;
}

oaph_status stage(oaph_mlp* m, const float* X, const float* T, uint32_t B, hipMemcpyKind kind) {
  OAPH_HIP(hipMemcpyAsync(m->X.ptr, X, (size_t)B * m->dims[0] * sizeof(float), kind, m->stream));
  if (T)
    OAPH_HIP(hipMemcpyAsync(m->T.ptr, T, (size_t)B * m->dims.back() * sizeof(float), kind,
                            m->stream));
  return OAPH_OK;
}

oaph_status adam(oaph_mlp* m, float lr) {
  ++m->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(m->nparams)), dim3(256), 0, m->stream, m->nparams,
                     m->params.as<float>(), m->grads.as<float>(), m->adam_m.as<float>(),
                     m->adam_v.as<float>(), lr, b1, b2, c1, c2, 1e-8f);
  OAPH_HIP(hipGetLastError());
  return refresh_bf16(m);
}

__global__ void k_gather(uint32_t B, const uint32_t* __restrict__ idx, uint32_t dx, uint32_t dt,
                         const float* __restrict__ X, const float* __restrict__ T,
                         float* __restrict__ Xo, float* __restrict__ To) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const uint32_t s = idx[i];
  for (uint32_t k = 0; k < dx; ++k) Xo[(size_t)i * dx + k] = X[(size_t)s * dx + k];
  for (uint32_t k = 0; k < dt; ++k) To[(size_t)i * dt + k] = T[(size_t)s * dt + k];
}

}  // namespace

extern "C" {

// dims: d_0 .. d_L (L >= 2 layers of weights), hidden widths <= 512
oaph_status oaph_mlp_create(oaph_mlp** out, const uint32_t* dims, uint32_t ndims,
                            uint32_t max_batch) {
  OAPH_CHECK(out && dims && ndims >= 3 && max_batch > 0, OAPH_ERR_ARG);
  for (uint32_t i = 0; i < ndims; ++i) OAPH_CHECK(dims[i] >= 1 && dims[i] <= 512, OAPH_ERR_UNSUPPORTED);
  OAPH_TRY(ensure_init());
  oaph_mlp* m = new oaph_mlp();
  m->dims.assign(dims, dims + ndims);
  m->max_batch = max_batch;
  size_t off = 0;
  for (uint32_t l = 0; l + 1 < ndims; ++l) {
    m->w_off.push_back(off);
    off += (size_t)dims[l + 1] * dims[l];
    m->b_off.push_back(off);
    off += dims[l + 1];
  }
  m->nparams = off;
  size_t aoff = 0;
  m->act_off.assign(ndims, 0);
  for (uint32_t l = 1; l < ndims; ++l) {
    m->act_off[l] = aoff;
    aoff += (size_t)max_batch * dims[l];
  }
  oaph_status st = OAPH_OK;
  auto R = [&](DevBuf& b, size_t bytes) {
    if (st == OAPH_OK) st = b.reserve(bytes);
  };
  R(m->params, off * 4);
  R(m->grads, off * 4);
  R(m->adam_m, off * 4);
  R(m->adam_v, off * 4);
  R(m->wbf, off * 2);
  R(m->acts, aoff * 4);
  R(m->actsb, aoff * 2);
  R(m->dz, (size_t)max_batch * 1024 * 4);
  R(m->dzb, (size_t)max_batch * 1024 * 2);
  R(m->X, (size_t)max_batch * dims[0] * 4);
  R(m->T, (size_t)max_batch * dims[ndims - 1] * 4);
  R(m->outw, (size_t)dims[ndims - 1] * 4);
  R(m->loss_part, (size_t)((max_batch * dims[ndims - 1] + 255) / 256 + 1) * 4);
  R(m->idx, (size_t)max_batch * 4);
  if (st == OAPH_OK && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    st = OAPH_ERR_HIP;
  if (st == OAPH_OK) {
    if (hipMemset(m->params.ptr, 0, off * 4) != hipSuccess ||
        hipMemset(m->adam_m.ptr, 0, off * 4) != hipSuccess ||
        hipMemset(m->adam_v.ptr, 0, off * 4) != hipSuccess)
      st = OAPH_ERR_HIP;
    std::vector<float> ones(dims[ndims - 1], 1.f);
    if (st == OAPH_OK &&
        hipMemcpy(m->outw.ptr, ones.data(), ones.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      st = OAPH_ERR_HIP;
  }
  if (st != OAPH_OK) {
    for (DevBuf* b : {&m->params, &m->grads, &m->adam_m, &m->adam_v, &m->wbf, &m->acts,
                      &m->actsb, &m->dz, &m->dzb, &m->X, &m->T, &m->outw, &m->loss_part, &m->idx})
      b->release();
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return st;
  }
  *out = m;
  return OAPH_OK;
}

oaph_status oaph_mlp_destroy(oaph_mlp* m) {
  if (!m) return OAPH_OK;
  (void)hipStreamSynchronize(m->stream);
  for (DevBuf* b : {&m->params, &m->grads, &m->adam_m, &m->adam_v, &m->wbf, &m->acts, &m->actsb,
                    &m->dz, &m->dzb, &m->X, &m->T, &m->outw, &m->loss_part, &m->idx, &m->data_x,
                    &m->data_t})
    b->release();
  (void)hipStreamDestroy(m->stream);
  delete m;
  return OAPH_OK;
}

oaph_status oaph_mlp_num_params(oaph_mlp* m, size_t* n) {
  OAPH_CHECK(m && n, OAPH_ERR_ARG);
  *n = m->nparams;
  return OAPH_OK;
}

// flat FP32 parameters: for each layer W (d_{l+1} x d_l, row-major) then b;
// resets the optimiser state
oaph_status oaph_mlp_set_params(oaph_mlp* m, const float* p) {
  OAPH_CHECK(m && p, OAPH_ERR_ARG);
  OAPH_HIP(hipStreamSynchronize(m->stream));
  OAPH_HIP(hipMemcpy(m->params.ptr, p, m->nparams * 4, hipMemcpyHostToDevice));
  OAPH_HIP(hipMemset(m->adam_m.ptr, 0, m->nparams * 4));
  OAPH_HIP(hipMemset(m->adam_v.ptr, 0, m->nparams * 4));
  m->step = 0;
  OAPH_TRY(refresh_bf16(m));
  OAPH_HIP(hipStreamSynchronize(m->stream));
  return OAPH_OK;
}

oaph_status oaph_mlp_get_params(oaph_mlp* m, float* p) {
  OAPH_CHECK(m && p, OAPH_ERR_ARG);
  OAPH_HIP(hipStreamSynchronize(m->stream));
  OAPH_HIP(hipMemcpy(p, m->params.ptr, m->nparams * 4, hipMemcpyDeviceToHost));
  return OAPH_OK;
}

// per-output loss weights w_j (default 1)
oaph_status oaph_mlp_set_output_weights(oaph_mlp* m, const float* w) {
  OAPH_CHECK(m && w, OAPH_ERR_ARG);
  OAPH_HIP(hipStreamSynchronize(m->stream));
  OAPH_HIP(hipMemcpy(m->outw.ptr, w, m->dims.back() * 4, hipMemcpyHostToDevice));
  return OAPH_OK;
}

// Y (n x d_L) = network(X (n x d_0)), host arrays, any n (split in batches)
oaph_status oaph_mlp_predict(oaph_mlp* m, const float* X, uint32_t n, float* Y) {
  OAPH_CHECK(m && X && Y, OAPH_ERR_ARG);
  const int L = (int)m->dims.size() - 1;
  for (uint32_t b0 = 0; b0 < n; b0 += m->max_batch) {
    const uint32_t B = std::min(m->max_batch, n - b0);
    OAPH_TRY(stage(m, X + (size_t)b0 * m->dims[0], nullptr, B, hipMemcpyHostToDevice));
    OAPH_TRY(forward(m, B));
    OAPH_HIP(hipMemcpyAsync(Y + (size_t)b0 * m->dims[L], m->acts.as<float>() + m->act_off[L],
                            (size_t)B * m->dims[L] * 4, hipMemcpyDeviceToHost, m->stream));
    OAPH_HIP(hipStreamSynchronize(m->stream));
  }
  return OAPH_OK;
}

// gradients (flat, like the parameters) and loss of one batch of host data
oaph_status oaph_mlp_gradients(oaph_mlp* m, const float* X, const float* T, uint32_t B,
                               float* grads, float* loss) {
  OAPH_CHECK(m && X && T && grads && B >= 1 && B <= m->max_batch, OAPH_ERR_ARG);
  OAPH_TRY(stage(m, X, T, B, hipMemcpyHostToDevice));
  OAPH_TRY(forward(m, B));
  float l = 0;
  OAPH_TRY(backward(m, B, &l));
  OAPH_HIP(hipMemcpyAsync(grads, m->grads.ptr, m->nparams * 4, hipMemcpyDeviceToHost, m->stream));
  OAPH_HIP(hipStreamSynchronize(m->stream));
  if (loss) *loss = l;
  return OAPH_OK;
}

// resident training set (host arrays copied once)
oaph_status oaph_mlp_set_dataset(oaph_mlp* m, const float* X, const float* T, uint32_t n) {
  OAPH_CHECK(m && X && T && n > 0, OAPH_ERR_ARG);
  OAPH_TRY(m->data_x.reserve((size_t)n * m->dims[0] * 4));
  OAPH_TRY(m->data_t.reserve((size_t)n * m->dims.back() * 4));
  OAPH_HIP(hipMemcpy(m->data_x.ptr, X, (size_t)n * m->dims[0] * 4, hipMemcpyHostToDevice));
  OAPH_HIP(hipMemcpy(m->data_t.ptr, T, (size_t)n * m->dims.back() * 4, hipMemcpyHostToDevice));
  m->data_n = n;
  return OAPH_OK;
}

// `epochs` passes of Adam over the resident set in shuffled batches of B
// (SplitMix64-seeded Fisher-Yates per epoch); losses[e] = mean batch loss
oaph_status oaph_mlp_train(oaph_mlp* m, uint32_t B, uint32_t epochs, float lr, uint64_t seed,
                           float* losses) {
  OAPH_CHECK(m && B >= 1 && B <= m->max_batch, OAPH_ERR_ARG);
  OAPH_CHECK(m->data_n >= B, OAPH_ERR_STATE);
  std::vector<uint32_t> perm(m->data_n);
  for (uint32_t i = 0; i < m->data_n; ++i) perm[i] = i;
  uint64_t state = seed;
  auto next = [&]() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  };
  DevBuf didx;
  OAPH_TRY(didx.reserve((size_t)m->data_n * 4));
  for (uint32_t e = 0; e < epochs; ++e) {
    for (uint32_t i = m->data_n - 1; i > 0; --i) std::swap(perm[i], perm[next() % (i + 1)]);
    OAPH_HIP(hipMemcpyAsync(didx.ptr, perm.data(), (size_t)m->data_n * 4, hipMemcpyHostToDevice,
                            m->stream));
    double esum = 0;
    uint32_t nb = 0;
    for (uint32_t b0 = 0; b0 + B <= m->data_n; b0 += B, ++nb) {
      hipLaunchKernelGGL(k_gather, dim3(blocks(B)), dim3(256), 0, m->stream, B,
                         didx.as<uint32_t>() + b0, m->dims[0], m->dims.back(),
                         m->data_x.as<float>(), m->data_t.as<float>(), m->X.as<float>(),
                         m->T.as<float>());
      OAPH_HIP(hipGetLastError());
      OAPH_TRY(forward(m, B));
      float l = 0;
      OAPH_TRY(backward(m, B, losses ? &l : nullptr));
      esum += l;
      OAPH_TRY(adam(m, lr));
    }
    if (losses) losses[e] = (float)(esum / nb);
  }
  OAPH_HIP(hipStreamSynchronize(m->stream));
  didx.release();
  return OAPH_OK;
}

}  // extern "C"
