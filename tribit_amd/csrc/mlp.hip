// mlp.hip -- the scheduler's cost model: a 3-layer perceptron (F -> 64 -> 64
// -> 1, ReLU), trained in FP32 by the kernels below (forward, backward, Adam)
// and evaluated in BF16 (weights and activations rounded to bfloat16,
// products accumulated in FP32) when the scheduler scores candidates.
//
// No counterpart in the reference. Sizes are tiny (a batch of a few thousand
// samples, 64 hidden units), so the kernels are plain: one thread per output
// element, fixed loop order, no atomics -- a training run is deterministic.
//
// Parameter layout (FP32, one flat array):
//   W1[F][64] b1[64] W2[64][64] b2[64] W3[64] b3[1]
// Loss: L = 1/(2B) * sum_b (y_b - t_b)^2.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace tb {

constexpr int kHidden = 64;
constexpr unsigned kMlpBlock = 256;

struct MlpShape {
  int f;
  size_t w1, b1, w2, b2, w3, b3, total;
};

__host__ __device__ static MlpShape mlp_shape(int f) {
  MlpShape s;
  s.f = f;
  s.w1 = 0;
  s.b1 = s.w1 + (size_t)f * kHidden;
  s.w2 = s.b1 + kHidden;
  s.b2 = s.w2 + (size_t)kHidden * kHidden;
  s.w3 = s.b2 + kHidden;
  s.b3 = s.w3 + kHidden;
  s.total = s.b3 + 1;
  return s;
}

// Y[b][n] = act(sum_k X[b][k] W[k][n] + bias[n])
__global__ void k_lin_fwd(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ bias,
                          float* __restrict__ Y, int B, int K, int N, int relu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * N) return;
  const int b = i / N, n = i % N;
  float acc = bias[n];
  for (int k = 0; k < K; ++k) acc = fmaf(X[(size_t)b * K + k], W[(size_t)k * N + n], acc);
  Y[i] = (relu && acc < 0.f) ? 0.f : acc;
}

// dy[b] = (y[b] - t[b]) / B ; loss terms
__global__ void k_mse_grad(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy,
                           float* __restrict__ sq, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float e = y[b] - t[b];
  dy[b] = e / (float)B;
  sq[b] = e * e;
}

// dX[b][k] = (sum_n dY[b][n] W[k][n]) * (Xact[b][k] > 0)   (Xact: the ReLU output feeding the layer)
__global__ void k_lin_bwd_in(const float* __restrict__ dY, const float* __restrict__ W,
                             const float* __restrict__ Xact, float* __restrict__ dX, int B, int K, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * K) return;
  const int b = i / K, k = i % K;
  float acc = 0.f;
  for (int n = 0; n < N; ++n) acc = fmaf(dY[(size_t)b * N + n], W[(size_t)k * N + n], acc);
  dX[i] = Xact[i] > 0.f ? acc : 0.f;
}

// dW[k][n] = sum_b X[b][k] dY[b][n];  dbias[n] = sum_b dY[b][n]  (thread k == K computes dbias)
__global__ void k_lin_bwd_w(const float* __restrict__ X, const float* __restrict__ dY, float* __restrict__ dW,
                            float* __restrict__ dbias, int B, int K, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (K + 1) * N) return;
  const int k = i / N, n = i % N;
  float acc = 0.f;
  if (k < K) {
    for (int b = 0; b < B; ++b) acc = fmaf(X[(size_t)b * K + k], dY[(size_t)b * N + n], acc);
    dW[(size_t)k * N + n] = acc;
  } else {
    for (int b = 0; b < B; ++b) acc += dY[(size_t)b * N + n];
    dbias[n] = acc;
  }
}

// Adam (Kingma & Ba 2015), bias-corrected step size lr_t.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr_t, float beta1, float beta2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = beta1 * m1[i] + (1.f - beta1) * gi;
  const float b = beta2 * m2[i] + (1.f - beta2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr_t * a / (sqrtf(b) + eps);
}

// FP32 -> BF16 with round-to-nearest-even on the bit pattern (NaN kept quiet).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__global__ void k_zero_f32(float* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

__global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ q, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = f32_to_bf16(p[i]);
}

// BF16 inference: one thread per sample; weights (bf16) staged in LDS; inputs
// and hidden activations rounded to bf16; FP32 accumulation.
constexpr int kMaxFeat = 32;
__global__ __launch_bounds__(kMlpBlock) void k_mlp_infer_bf16(const float* __restrict__ X, int B, int F,
                                                               const uint16_t* __restrict__ P,
                                                               float* __restrict__ out) {
  __shared__ uint16_t sP[kMaxFeat * kHidden + kHidden + kHidden * kHidden + kHidden + kHidden + 1];
  const MlpShape s = mlp_shape(F);
  for (size_t i = threadIdx.x; i < s.total; i += blockDim.x) sP[i] = P[i];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float x[kMaxFeat];
  for (int k = 0; k < F; ++k) x[k] = bf16_to_f32(f32_to_bf16(X[(size_t)b * F + k]));
  float h1[kHidden];
  for (int n = 0; n < kHidden; ++n) {
    float acc = bf16_to_f32(sP[s.b1 + n]);
    for (int k = 0; k < F; ++k) acc = fmaf(x[k], bf16_to_f32(sP[s.w1 + (size_t)k * kHidden + n]), acc);
    h1[n] = bf16_to_f32(f32_to_bf16(acc > 0.f ? acc : 0.f));
  }
  float y = bf16_to_f32(sP[s.b3]);
  for (int n = 0; n < kHidden; ++n) {
    float acc = bf16_to_f32(sP[s.b2 + n]);
    for (int k = 0; k < kHidden; ++k) acc = fmaf(h1[k], bf16_to_f32(sP[s.w2 + (size_t)k * kHidden + n]), acc);
    const float h2 = bf16_to_f32(f32_to_bf16(acc > 0.f ? acc : 0.f));
    y = fmaf(h2, bf16_to_f32(sP[s.w3 + n]), y);
  }
  out[b] = y;
}

}  // namespace tb

// ---------------------------------------------------------------- model object
struct tribit_mlp {
  tribit_ctx* ctx = nullptr;
  tb::MlpShape shape{};
  float* p = nullptr;   // parameters (device)
  float* m1 = nullptr;  // Adam moments
  float* m2 = nullptr;
  uint16_t* pb = nullptr;  // bf16 copy for inference (refreshed after training / set)
  uint64_t step = 0;
};

namespace tb {

static unsigned nblk(size_t n) { return (unsigned)((n + kMlpBlock - 1) / kMlpBlock); }

void mlp_refresh_bf16(tribit_mlp* m) {
  k_to_bf16<<<nblk(m->shape.total), kMlpBlock, 0, m->ctx->stream>>>(m->p, m->pb, (int)m->shape.total);
}

tribit_mlp* mlp_create(tribit_ctx* ctx, int nfeat, uint64_t seed) {
  if (nfeat < 1 || nfeat > kMaxFeat) fail(TRIBIT_ERR_ARG, "feature count must be in [1, 32]");
  tribit_mlp* m = new tribit_mlp;
  m->ctx = ctx;
  m->shape = mlp_shape(nfeat);
  const size_t T = m->shape.total;
  try {
    m->p = ctx->alloc_n<float>(T);
    m->m1 = ctx->alloc_n<float>(T);
    m->m2 = ctx->alloc_n<float>(T);
    m->pb = ctx->alloc_n<uint16_t>(T);
  } catch (...) {
    ctx->free(m->p);
    ctx->free(m->m1);
    ctx->free(m->m2);
    ctx->free(m->pb);
    delete m;
    throw;
  }
  // He-uniform initialisation from a SplitMix64 stream; biases zero
  std::vector<float> h(T, 0.f);
  uint64_t z = seed;
  auto uni = [&z]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](size_t off, size_t cnt, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (size_t i = 0; i < cnt; ++i) h[off + i] = (float)((2.0 * uni() - 1.0) * a);
  };
  const MlpShape& s = m->shape;
  fill(s.w1, (size_t)nfeat * kHidden, nfeat);
  fill(s.w2, (size_t)kHidden * kHidden, kHidden);
  fill(s.w3, kHidden, kHidden);
  TB_HIP(hipMemcpyAsync(m->p, h.data(), 4 * T, hipMemcpyHostToDevice, ctx->stream));
  k_zero_f32<<<nblk(T), kMlpBlock, 0, ctx->stream>>>(m->m1, T);
  k_zero_f32<<<nblk(T), kMlpBlock, 0, ctx->stream>>>(m->m2, T);
  mlp_refresh_bf16(m);
  ctx->sync();
  return m;
}

void mlp_free(tribit_mlp* m) {
  if (!m) return;
  if (m->ctx) {
    m->ctx->free(m->p);
    m->ctx->free(m->m1);
    m->ctx->free(m->m2);
    m->ctx->free(m->pb);
  }
  delete m;
}

// Forward + backward of one batch (device X[B][F], t[B]); gradients into g
// (device, parameter layout); returns the loss.
static double mlp_fwd_bwd(tribit_mlp* m, const float* X, const float* t, int B, float* g) {
  tribit_ctx* ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const MlpShape& s = m->shape;
  const int F = s.f, H = kHidden;
  DevBuf<float> h1(ctx, (size_t)B * H), h2(ctx, (size_t)B * H), y(ctx, B), dy(ctx, B), sq(ctx, B);
  DevBuf<float> d2(ctx, (size_t)B * H), d1(ctx, (size_t)B * H);
  k_lin_fwd<<<nblk((size_t)B * H), kMlpBlock, 0, st>>>(X, m->p + s.w1, m->p + s.b1, h1.p, B, F, H, 1);
  k_lin_fwd<<<nblk((size_t)B * H), kMlpBlock, 0, st>>>(h1.p, m->p + s.w2, m->p + s.b2, h2.p, B, H, H, 1);
  k_lin_fwd<<<nblk(B), kMlpBlock, 0, st>>>(h2.p, m->p + s.w3, m->p + s.b3, y.p, B, H, 1, 0);
  k_mse_grad<<<nblk(B), kMlpBlock, 0, st>>>(y.p, t, dy.p, sq.p, B);
  // layer 3
  k_lin_bwd_w<<<nblk((size_t)(H + 1)), kMlpBlock, 0, st>>>(h2.p, dy.p, g + s.w3, g + s.b3, B, H, 1);
  k_lin_bwd_in<<<nblk((size_t)B * H), kMlpBlock, 0, st>>>(dy.p, m->p + s.w3, h2.p, d2.p, B, H, 1);
  // layer 2
  k_lin_bwd_w<<<nblk((size_t)(H + 1) * H), kMlpBlock, 0, st>>>(h1.p, d2.p, g + s.w2, g + s.b2, B, H, H);
  k_lin_bwd_in<<<nblk((size_t)B * H), kMlpBlock, 0, st>>>(d2.p, m->p + s.w2, h1.p, d1.p, B, H, H);
  // layer 1
  k_lin_bwd_w<<<nblk((size_t)(F + 1) * H), kMlpBlock, 0, st>>>(X, d1.p, g + s.w1, g + s.b1, B, F, H);
  TB_HIP(hipGetLastError());
  std::vector<float> hs(B);
  TB_HIP(hipMemcpyAsync(hs.data(), sq.p, 4 * (size_t)B, hipMemcpyDeviceToHost, st));
  ctx->sync();
  double l = 0.0;
  for (int b = 0; b < B; ++b) l += hs[b];
  return l / (2.0 * B);
}

double mlp_grad(tribit_mlp* m, const float* hX, const float* ht, int B, float* hgrad) {
  tribit_ctx* ctx = m->ctx;
  const int F = m->shape.f;
  DevBuf<float> X(ctx, (size_t)B * F), t(ctx, B), g(ctx, m->shape.total);
  TB_HIP(hipMemcpyAsync(X.p, hX, 4 * (size_t)B * F, hipMemcpyHostToDevice, ctx->stream));
  TB_HIP(hipMemcpyAsync(t.p, ht, 4 * (size_t)B, hipMemcpyHostToDevice, ctx->stream));
  const double l = mlp_fwd_bwd(m, X.p, t.p, B, g.p);
  TB_HIP(hipMemcpyAsync(hgrad, g.p, 4 * m->shape.total, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  return l;
}

// Mini-batch Adam over host samples (X[n][F], t[n]); batches are taken in a
// fixed order (a SplitMix64 permutation per epoch). Returns the last epoch's
// mean loss.
double mlp_train(tribit_mlp* m, const float* hX, const float* ht, uint64_t n, int epochs, int batch, float lr,
                 uint64_t seed) {
  tribit_ctx* ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const int F = m->shape.f;
  const size_t T = m->shape.total;
  if (n == 0 || batch < 1 || n > 0xFFFFFFFFull) fail(TRIBIT_ERR_ARG, "empty training set or batch");
  DevBuf<float> g(ctx, T);
  std::vector<uint32_t> perm(n);
  for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
  DevBuf<float> bx(ctx, (size_t)batch * F), bt(ctx, batch);
  std::vector<float> hx((size_t)batch * F), hb(batch);
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  double last = 0.0;
  uint64_t z = seed;
  for (int ep = 0; ep < epochs; ++ep) {
    for (uint64_t i = n - 1; i > 0; --i) {  // Fisher-Yates with SplitMix64
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      x ^= x >> 31;
      std::swap(perm[i], perm[x % (i + 1)]);
    }
    double sum = 0.0;
    uint64_t nb = 0;
    for (uint64_t b0 = 0; b0 < n; b0 += batch) {
      const int B = (int)std::min<uint64_t>(batch, n - b0);
      for (int j = 0; j < B; ++j) {
        memcpy(&hx[(size_t)j * F], hX + (size_t)perm[b0 + j] * F, 4 * F);
        hb[j] = ht[perm[b0 + j]];
      }
      TB_HIP(hipMemcpyAsync(bx.p, hx.data(), 4 * (size_t)B * F, hipMemcpyHostToDevice, st));
      TB_HIP(hipMemcpyAsync(bt.p, hb.data(), 4 * (size_t)B, hipMemcpyHostToDevice, st));
      sum += mlp_fwd_bwd(m, bx.p, bt.p, B, g.p);
      ++nb;
      ++m->step;
      const double c1 = 1.0 - std::pow((double)beta1, (double)m->step);
      const double c2 = 1.0 - std::pow((double)beta2, (double)m->step);
      const float lr_t = (float)(lr * std::sqrt(c2) / c1);
      k_adam<<<nblk(T), kMlpBlock, 0, st>>>(m->p, g.p, m->m1, m->m2, (int)T, lr_t, beta1, beta2, eps);
    }
    last = sum / (double)nb;
  }
  mlp_refresh_bf16(m);
  ctx->sync();
  return last;
}

size_t mlp_num_params(const tribit_mlp* m) { return m->shape.total; }
tribit_ctx* mlp_context(const tribit_mlp* m) { return m->ctx; }
int mlp_num_features(const tribit_mlp* m) { return m->shape.f; }

void mlp_get(tribit_mlp* m, float* h) {
  TB_HIP(hipMemcpyAsync(h, m->p, 4 * m->shape.total, hipMemcpyDeviceToHost, m->ctx->stream));
  m->ctx->sync();
}

void mlp_set(tribit_mlp* m, const float* h) {
  TB_HIP(hipMemcpyAsync(m->p, h, 4 * m->shape.total, hipMemcpyHostToDevice, m->ctx->stream));
  k_zero_f32<<<nblk(m->shape.total), kMlpBlock, 0, m->ctx->stream>>>(m->m1, m->shape.total);
  k_zero_f32<<<nblk(m->shape.total), kMlpBlock, 0, m->ctx->stream>>>(m->m2, m->shape.total);
  m->step = 0;
  mlp_refresh_bf16(m);
  m->ctx->sync();
}

// BF16 inference over device samples.
void mlp_infer_dev(tribit_mlp* m, const float* dX, uint64_t n, float* dout) {
  if (!n) return;
  k_mlp_infer_bf16<<<nblk(n), kMlpBlock, 0, m->ctx->stream>>>(dX, (int)n, m->shape.f, m->pb, dout);
  TB_HIP(hipGetLastError());
}

void mlp_infer(tribit_mlp* m, const float* hX, uint64_t n, float* hout) {
  tribit_ctx* ctx = m->ctx;
  DevBuf<float> X(ctx, n * m->shape.f + 1), out(ctx, n + 1);
  TB_HIP(hipMemcpyAsync(X.p, hX, 4 * n * m->shape.f, hipMemcpyHostToDevice, ctx->stream));
  mlp_infer_dev(m, X.p, n, out.p);
  TB_HIP(hipMemcpyAsync(hout, out.p, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
}

}  // namespace tb
