// Learned branching-score surrogate: an MLP  in -> h1 -> h2 -> 1  (ReLU)
// trained and evaluated with BF16 MFMA kernels.
//
// Every matrix product of the forward and backward passes is one launch of
// mlp_gemm_kernel: 64 x 64 output tiles, four wavefronts of 32 x 32, each a
// 2 x 2 grid of v_mfma_f32_16x16x16bf16_1k blocks. Operands are read from the
// FP32 buffers (master weights, activations, gradients) and rounded to BF16
// (round to nearest even) as they are packed into the instruction's four-
// element operands; accumulation is FP32; the epilogue adds the bias and
// applies ReLU in the forward pass, or the ReLU derivative in the backward
// pass. Bias gradients are column sums, the loss is 0.5 mean (y - t)^2, the
// optimiser Adam on the FP32 master weights. The sizes are small (batch
// <= 1024, widths 256): the kernels aim at correctness and at keeping the
// whole training loop on the device, not at the MFMA peak.

#include "internal.hpp"

#include "sovereign_surrogate.h"

#include <cmath>
#include <cstdint>
#include <vector>

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int kTileMN = 64;

__device__ __forceinline__ short to_bf16(float f) {
  const unsigned u = __float_as_uint(f);
  return static_cast<short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

enum Epi { kPlain = 0, kBiasRelu = 1, kBias = 2, kReluMask = 3 };

// C(m, n) = epi(sum_k A(m, k) B(k, n)), m < M, n < N, k < K, with
//   A(m, k) = TA ? A[k * lda + m] : A[m * lda + k],
//   B(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n],
// C row-major (ldc). kBiasRelu: max(acc + bias[n], 0); kBias: acc + bias[n];
// kReluMask: acc where H[m * ldc + n] > 0, else 0.
// Instruction layout (16 x 16 x 16): operand X(i, k) sits on lane i + 16 (k / 4),
// element k % 4; Y(k, j) on lane j + 16 (k / 4); result (i, j) on lane
// j + 16 (i / 4), register i % 4.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                       const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                       int ldc, const float* __restrict__ bias,
                                                       const float* __restrict__ H) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m0 = blockIdx.y * kTileMN + (w & 1) * 32, n0 = blockIdx.x * kTileMN + (w >> 1) * 32;
  f32x4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int m = m0 + 16 * i + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int k = k0 + 4 * lk + v;
        float x = 0.f;
        if (m < M && k < K) x = TA ? A[static_cast<long long>(k) * lda + m] : A[static_cast<long long>(m) * lda + k];
        a[i][v] = to_bf16(x);
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + 16 * j + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int k = k0 + 4 * lk + v;
        float x = 0.f;
        if (n < N && k < K) x = TB ? B[static_cast<long long>(n) * ldb + k] : B[static_cast<long long>(k) * ldb + n];
        b[j][v] = to_bf16(x);
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + 16 * j + li;
      if (n >= N) continue;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int m = m0 + 16 * i + 4 * lk + g;
        if (m >= M) continue;
        float v = acc[i][j][g];
        const long long o = static_cast<long long>(m) * ldc + n;
        if (EPI == kBiasRelu) v = fmaxf(v + bias[n], 0.f);
        if (EPI == kBias) v += bias[n];
        if (EPI == kReluMask) v = H[o] > 0.f ? v : 0.f;
        C[o] = v;
      }
    }
}

// out[n] = sum_{m < M} X[m * ld + n] (in order of m: deterministic).
__global__ void col_sum_kernel(int M, int N, const float* __restrict__ X, int ld, float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += X[static_cast<long long>(m) * ld + n];
  out[n] = s;
}

// dy = (y - t) / B; loss_acc += 0.5 mean (y - t)^2 (one workgroup).
__global__ __launch_bounds__(256) void loss_kernel(int B, const float* __restrict__ y, const float* __restrict__ t,
                                                   float* __restrict__ dy, float* __restrict__ loss_acc) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < B; b += 256) {
    const float d = y[b] - t[b];
    dy[b] = d / static_cast<float>(B);
    s += 0.5 * static_cast<double>(d) * d;
  }
  s = block_reduce<0>(s, sh);
  if (threadIdx.x == 0) loss_acc[0] += static_cast<float>(s / B);
}

__global__ void adam_kernel(long long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// dst rows r < count: x_dst[r, :] = x_src[idx[r], :], t_dst[r] = t_src[idx[r]].
__global__ void gather_kernel(int count, int F, const int* __restrict__ idx, const float* __restrict__ xs,
                              const float* __restrict__ ts, float* __restrict__ xd, float* __restrict__ td) {
  const int r = blockIdx.x;
  if (r >= count) return;
  const long long src = idx[r];
  for (int f = threadIdx.x; f < F; f += blockDim.x) xd[static_cast<long long>(r) * F + f] = xs[src * F + f];
  if (threadIdx.x == 0) td[r] = ts[src];
}

__global__ void zero_kernel(int n, float* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

template <bool TA, bool TB, int EPI>
hipError_t gemm(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                const float* bias, const float* H, hipStream_t s) {
  const dim3 grid((N + kTileMN - 1) / kTileMN, (M + kTileMN - 1) / kTileMN);
  mlp_gemm_kernel<TA, TB, EPI><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, H);
  return hipGetLastError();
}

// splitmix64: the same initial weights for a seed on every host.
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (static_cast<double>(splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

}  // namespace

struct sov_mlp {
  int in = 0, h1 = 0, h2 = 0;
  long long np = 0;
  long long oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  int cap = 0;  // batch capacity of the activation buffers
  float *X = nullptr, *T = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr, *dH1 = nullptr,
        *dH2 = nullptr, *loss = nullptr;
  int* idx = nullptr;
  long long adam_t = 0;
  std::vector<void*> owned;
  ~sov_mlp() {
    for (void* q : owned) (void)hipFree(q);
  }
  template <typename U>
  U* alloc(std::size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<std::size_t>(n, 1) * sizeof(U)) != hipSuccess) return nullptr;
    owned.push_back(q);
    return static_cast<U*>(q);
  }
  int reserve(int batch) {
    if (batch <= cap) return SOV_OK;
    const std::size_t b = static_cast<std::size_t>(batch);
    X = alloc<float>(b * in);
    T = alloc<float>(b);
    H1 = alloc<float>(b * h1);
    H2 = alloc<float>(b * h2);
    Y = alloc<float>(b);
    dY = alloc<float>(b);
    dH1 = alloc<float>(b * h1);
    dH2 = alloc<float>(b * h2);
    idx = alloc<int>(b);
    if (!X || !T || !H1 || !H2 || !Y || !dY || !dH1 || !dH2 || !idx) return SOV_ERR_NOMEM;
    cap = batch;
    return SOV_OK;
  }
  // forward pass of the batch in X (device)
  int forward(int B, hipStream_t s) {
    SOV_HIP_TRY((gemm<false, false, kBiasRelu>(B, h1, in, X, in, p + oW1, h1, H1, h1, p + ob1, nullptr, s)));
    SOV_HIP_TRY((gemm<false, false, kBiasRelu>(B, h2, h1, H1, h1, p + oW2, h2, H2, h2, p + ob2, nullptr, s)));
    SOV_HIP_TRY((gemm<false, false, kBias>(B, 1, h2, H2, h2, p + oW3, 1, Y, 1, p + ob3, nullptr, s)));
    return SOV_OK;
  }
  // loss and gradients into g for the batch in (X, T)
  int backward(int B, hipStream_t s) {
    loss_kernel<<<1, 256, 0, s>>>(B, Y, T, dY, loss);
    SOV_LAUNCH_CHECK();
    SOV_HIP_TRY((gemm<true, false, kPlain>(h2, 1, B, H2, h2, dY, 1, g + oW3, 1, nullptr, nullptr, s)));
    col_sum_kernel<<<1, 64, 0, s>>>(B, 1, dY, 1, g + ob3);
    SOV_HIP_TRY((gemm<false, true, kReluMask>(B, h2, 1, dY, 1, p + oW3, 1, dH2, h2, nullptr, H2, s)));
    SOV_HIP_TRY((gemm<true, false, kPlain>(h1, h2, B, H1, h1, dH2, h2, g + oW2, h2, nullptr, nullptr, s)));
    col_sum_kernel<<<(h2 + 255) / 256, 256, 0, s>>>(B, h2, dH2, h2, g + ob2);
    SOV_HIP_TRY((gemm<false, true, kReluMask>(B, h1, h2, dH2, h2, p + oW2, h2, dH1, h1, nullptr, H1, s)));
    SOV_HIP_TRY((gemm<true, false, kPlain>(in, h1, B, X, in, dH1, h1, g + oW1, h1, nullptr, nullptr, s)));
    col_sum_kernel<<<(h1 + 255) / 256, 256, 0, s>>>(B, h1, dH1, h1, g + ob1);
    SOV_LAUNCH_CHECK();
    return SOV_OK;
  }
};

extern "C" {

int sov_mlp_create(const sov_mlp_shape* shape, uint64_t seed, sov_mlp** out) {
  if (!shape || !out || shape->in <= 0 || shape->hidden1 <= 0 || shape->hidden2 <= 0) return SOV_ERR_ARG;
  *out = nullptr;
  if (!sov::device_ready()) return SOV_ERR_NODEV;
  auto* n = new sov_mlp;
  n->in = shape->in;
  n->h1 = shape->hidden1;
  n->h2 = shape->hidden2;
  n->oW1 = 0;
  n->ob1 = n->oW1 + static_cast<long long>(n->in) * n->h1;
  n->oW2 = n->ob1 + n->h1;
  n->ob2 = n->oW2 + static_cast<long long>(n->h1) * n->h2;
  n->oW3 = n->ob2 + n->h2;
  n->ob3 = n->oW3 + n->h2;
  n->np = n->ob3 + 1;
  n->p = n->alloc<float>(n->np);
  n->g = n->alloc<float>(n->np);
  n->m1 = n->alloc<float>(n->np);
  n->m2 = n->alloc<float>(n->np);
  n->loss = n->alloc<float>(1);
  if (!n->p || !n->g || !n->m1 || !n->m2 || !n->loss) {
    delete n;
    return SOV_ERR_NOMEM;
  }
  // Glorot-uniform weights and biases: U(-r, r), r = sqrt(6 / (fan_in + fan_out)).
  std::vector<float> h(static_cast<std::size_t>(n->np), 0.f);
  uint64_t st = seed;
  auto fill = [&](long long off, long long cnt, int fan_in, int fan_out) {
    const double r = std::sqrt(6.0 / (fan_in + fan_out));
    for (long long k = 0; k < cnt; ++k) h[off + k] = static_cast<float>(r * (2.0 * uniform01(st) - 1.0));
  };
  fill(n->oW1, static_cast<long long>(n->in) * n->h1, n->in, n->h1);
  fill(n->ob1, n->h1, n->in, n->h1);
  fill(n->oW2, static_cast<long long>(n->h1) * n->h2, n->h1, n->h2);
  fill(n->ob2, n->h2, n->h1, n->h2);
  fill(n->oW3, n->h2, n->h2, 1);
  fill(n->ob3, 1, n->h2, 1);
  const int rc = sov_mlp_set_params(n, h.data());
  if (rc != SOV_OK) {
    delete n;
    return rc;
  }
  *out = n;
  return SOV_OK;
}

int64_t sov_mlp_param_count(const sov_mlp* net) { return net ? net->np : -1; }

int sov_mlp_get_params(const sov_mlp* net, float* params) {
  if (!net || !params) return SOV_ERR_ARG;
  SOV_HIP_TRY(hipMemcpyAsync(params, net->p, net->np * sizeof(float), hipMemcpyDeviceToHost, sov::stream()));
  SOV_HIP_TRY(hipStreamSynchronize(sov::stream()));
  return SOV_OK;
}

int sov_mlp_set_params(sov_mlp* net, const float* params) {
  if (!net || !params) return SOV_ERR_ARG;
  hipStream_t s = sov::stream();
  SOV_HIP_TRY(hipMemcpyAsync(net->p, params, net->np * sizeof(float), hipMemcpyHostToDevice, s));
  const unsigned blocks = static_cast<unsigned>((net->np + 255) / 256);
  zero_kernel<<<blocks, 256, 0, s>>>(static_cast<int>(net->np), net->m1);
  zero_kernel<<<blocks, 256, 0, s>>>(static_cast<int>(net->np), net->m2);
  SOV_LAUNCH_CHECK();
  net->adam_t = 0;
  SOV_HIP_TRY(hipStreamSynchronize(s));
  return SOV_OK;
}

int sov_mlp_forward(sov_mlp* net, const float* x, int32_t batch, float* y) {
  if (!net || !x || !y || batch < 0) return SOV_ERR_ARG;
  if (batch == 0) return SOV_OK;
  int rc = net->reserve(batch);
  if (rc != SOV_OK) return rc;
  hipStream_t s = sov::stream();
  SOV_HIP_TRY(hipMemcpyAsync(net->X, x, static_cast<std::size_t>(batch) * net->in * sizeof(float),
                             hipMemcpyHostToDevice, s));
  if ((rc = net->forward(batch, s)) != SOV_OK) return rc;
  SOV_HIP_TRY(hipMemcpyAsync(y, net->Y, static_cast<std::size_t>(batch) * sizeof(float), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  sov::count_operation();
  return SOV_OK;
}

int sov_mlp_gradients(sov_mlp* net, const float* x, const float* t, int32_t batch, float* grad, float* loss) {
  if (!net || !x || !t || !grad || batch <= 0) return SOV_ERR_ARG;
  int rc = net->reserve(batch);
  if (rc != SOV_OK) return rc;
  hipStream_t s = sov::stream();
  const std::size_t b = static_cast<std::size_t>(batch);
  SOV_HIP_TRY(hipMemcpyAsync(net->X, x, b * net->in * sizeof(float), hipMemcpyHostToDevice, s));
  SOV_HIP_TRY(hipMemcpyAsync(net->T, t, b * sizeof(float), hipMemcpyHostToDevice, s));
  zero_kernel<<<1, 64, 0, s>>>(1, net->loss);
  SOV_LAUNCH_CHECK();
  if ((rc = net->forward(batch, s)) != SOV_OK) return rc;
  if ((rc = net->backward(batch, s)) != SOV_OK) return rc;
  SOV_HIP_TRY(hipMemcpyAsync(grad, net->g, net->np * sizeof(float), hipMemcpyDeviceToHost, s));
  float l = 0.f;
  SOV_HIP_TRY(hipMemcpyAsync(&l, net->loss, sizeof(float), hipMemcpyDeviceToHost, s));
  SOV_HIP_TRY(hipStreamSynchronize(s));
  if (loss) *loss = l;
  sov::count_operation();
  return SOV_OK;
}

int sov_mlp_train(sov_mlp* net, const float* x, const float* t, int64_t n, int32_t batch, int32_t epochs,
                  float learning_rate, uint64_t seed, float* loss_per_epoch) {
  if (!net || !x || !t || n <= 0 || batch <= 0 || epochs < 0 || !(learning_rate > 0.f)) return SOV_ERR_ARG;
  const int B = static_cast<int>(std::min<int64_t>(batch, n));
  int rc = net->reserve(B);
  if (rc != SOV_OK) return rc;
  hipStream_t s = sov::stream();
  float* xs = nullptr;
  float* ts = nullptr;
  SOV_HIP_TRY(hipMalloc(&xs, static_cast<std::size_t>(n) * net->in * sizeof(float)));
  if (hipMalloc(&ts, static_cast<std::size_t>(n) * sizeof(float)) != hipSuccess) {
    (void)hipFree(xs);
    return SOV_ERR_NOMEM;
  }
  struct Free {
    float* a;
    float* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } fr{xs, ts};
  SOV_HIP_TRY(hipMemcpyAsync(xs, x, static_cast<std::size_t>(n) * net->in * sizeof(float), hipMemcpyHostToDevice, s));
  SOV_HIP_TRY(hipMemcpyAsync(ts, t, static_cast<std::size_t>(n) * sizeof(float), hipMemcpyHostToDevice, s));
  std::vector<int> perm(static_cast<std::size_t>(n));
  uint64_t st = seed;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const unsigned pblocks = static_cast<unsigned>((net->np + 255) / 256);
  for (int e = 0; e < epochs; ++e) {
    for (int64_t i = 0; i < n; ++i) perm[i] = static_cast<int>(i);
    for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(st) % static_cast<uint64_t>(i + 1)]);
    zero_kernel<<<1, 64, 0, s>>>(1, net->loss);
    int batches = 0;
    for (int64_t b0 = 0; b0 < n; b0 += B) {
      const int cnt = static_cast<int>(std::min<int64_t>(B, n - b0));
      SOV_HIP_TRY(hipMemcpyAsync(net->idx, perm.data() + b0, cnt * sizeof(int), hipMemcpyHostToDevice, s));
      gather_kernel<<<cnt, 64, 0, s>>>(cnt, net->in, net->idx, xs, ts, net->X, net->T);
      SOV_LAUNCH_CHECK();
      if ((rc = net->forward(cnt, s)) != SOV_OK) return rc;
      if ((rc = net->backward(cnt, s)) != SOV_OK) return rc;
      ++net->adam_t;
      const float c1 = 1.f - std::pow(b1, static_cast<float>(net->adam_t));
      const float c2 = 1.f - std::pow(b2, static_cast<float>(net->adam_t));
      adam_kernel<<<pblocks, 256, 0, s>>>(net->np, net->p, net->g, net->m1, net->m2, learning_rate, b1, b2, eps, c1,
                                           c2);
      SOV_LAUNCH_CHECK();
      ++batches;
      // perm is pageable host memory that the next epoch reshuffles: the
      // upload must have completed before that happens.
      SOV_HIP_TRY(hipStreamSynchronize(s));
    }
    if (loss_per_epoch) {
      float l = 0.f;
      SOV_HIP_TRY(hipMemcpyAsync(&l, net->loss, sizeof(float), hipMemcpyDeviceToHost, s));
      SOV_HIP_TRY(hipStreamSynchronize(s));
      loss_per_epoch[e] = l / static_cast<float>(batches);
    }
  }
  SOV_HIP_TRY(hipStreamSynchronize(s));
  sov::count_operation();
  return SOV_OK;
}

int sov_mlp_free(sov_mlp* net) {
  delete net;
  return SOV_OK;
}

}  // extern "C"
