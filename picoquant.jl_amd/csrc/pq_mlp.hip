// Learned cost model kernels and the contraction-shape timer.
//
// pq_time_step: time one pairwise contraction shape (index groups + batch)
//   with HIP events on the context stream -- the measurements the cost
//   model is trained on and the autotuner's probe.
// pq_mlp_*: a 3-layer perceptron (in -> H -> H -> 1, ReLU, FP32) trained
//   with hand-written forward / backward / Adam kernels; it predicts
//   log(kernel microseconds) from a contraction's shape features.
//
// Parameter layout (one flat FP32 vector):
//   W1[in][H] | b1[H] | W2[H][H] | b2[H] | w3[H] | b3
// Minibatch activations live in device buffers sized for the largest batch.
#include "pq_rt.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

using namespace pqrt;

// ---------------------------------------------------------------- timer
extern "C" int pq_time_step(pq_ctx* c, int dtype, int nm, int nn, int nk, const int64_t* modes, int batch,
                            int kclass, int tile, int reps, double* us) {
  if (!c || !us || nm < 0 || nn < 0 || nk < 0 || (nm + nn + nk > 0 && !modes) || batch < 1 || batch > 65535 ||
      reps < 1 || reps > 1000 || (dtype != PQ_C128 && dtype != PQ_C64))
    return fail(c, PQ_ERR_ARG, "bad pq_time_step arguments");
  StepSpec s;
  int64_t q = 0;
  auto take = [&](int n) {
    std::vector<Mode> v;
    for (int k = 0; k < n; ++k, ++q) v.push_back({modes[3 * q], modes[3 * q + 1], modes[3 * q + 2]});
    return v;
  };
  s.M = merge_modes(take(nm));
  s.N = merge_modes(take(nn));
  s.K = merge_modes(take(nk));
  for (auto* v : {&s.M, &s.N, &s.K})
    for (auto& m : *v)
      if (m.ext < 1 || m.s0 < 0 || m.s1 < 0) return fail(c, PQ_ERR_ARG, "bad mode");
  auto span = [](const std::vector<Mode>& ms, bool first) {
    int64_t x = 0;
    for (auto& m : ms) x += (m.ext - 1) * (first ? m.s0 : m.s1);
    return x;
  };
  const int64_t ea = span(s.M, true) + span(s.K, true) + 1;
  const int64_t eb = span(s.N, true) + span(s.K, false) + 1;
  const int64_t ec = span(s.M, false) + span(s.N, false) + 1;
  const size_t es = dtype == PQ_C128 ? 16 : 8;
  const double total = (double)(ea + eb + ec) * batch * es;
  if (ea >= (1ll << 31) || eb >= (1ll << 31) || ec >= (1ll << 31) || total > 12e9)
    return fail(c, PQ_ERR_UNSUPPORTED, "shape too large to time");
  s.dtype = dtype;
  s.kc = kclass;
  s.tile = tile;
  s.L = single_batch();
  if (batch > 1) {
    s.L.nmodes = 1;
    s.L.total = batch;
    s.L.ext[0] = make_fdiv((uint32_t)batch);
    s.L.sA[0] = ea;
    s.L.sB[0] = eb;
    s.L.sC[0] = ec;
  }
  int r = resolve_step(c, s, es);
  if (r) return r;
  size_t ga, gb, gc, gs = 0;
  void* pa = dev_alloc(c, (size_t)ea * batch * es, &ga);
  void* pb = dev_alloc(c, (size_t)eb * batch * es, &gb);
  void* pc = dev_alloc(c, (size_t)ec * batch * es, &gc);
  size_t sb = step_scratch_bytes(s, es);
  void* ps = sb ? dev_alloc(c, sb, &gs) : nullptr;
  if (!pa || !pb || !pc || (sb && !ps)) {
    if (pa) dev_release(c, pa, ga);
    if (pb) dev_release(c, pb, gb);
    if (pc) dev_release(c, pc, gc);
    if (ps) dev_release(c, ps, gs);
    return fail(c, PQ_ERR_OOM, "timer allocation failed");
  }
  hipMemsetAsync(pa, 0, (size_t)ea * batch * es, c->stream);
  hipMemsetAsync(pb, 0, (size_t)eb * batch * es, c->stream);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  std::vector<float> t;
  r = launch_step(c, pa, pb, pc, s, ps);  // warm-up
  for (int i = 0; i < reps && r == PQ_OK; ++i) {
    hipEventRecord(e0, c->stream);
    r = launch_step(c, pa, pb, pc, s, ps);
    hipEventRecord(e1, c->stream);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    t.push_back(ms);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  dev_release(c, pa, ga);
  dev_release(c, pb, gb);
  dev_release(c, pc, gc);
  if (ps) dev_release(c, ps, gs);
  if (r) return r;
  std::sort(t.begin(), t.end());
  *us = 1000.0 * t[t.size() / 2];
  return PQ_OK;
}

// ------------------------------------------------------------------ MLP
namespace {

// H[b][j] = act(sum_i X[b][i] W[i][j] + bias[j]); one lane per (b, j).
__global__ void k_dense_fwd(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ bias,
                            float* __restrict__ H, int n, int in, int out, int relu) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * out) return;
  int b = t / out, j = t % out;
  float acc = bias[j];
  for (int i = 0; i < in; ++i) acc = fmaf(X[b * in + i], W[i * out + j], acc);
  H[t] = relu ? fmaxf(acc, 0.f) : acc;
}

// dY[b] = 2 (Y[b] - y[b]) / n ; loss partials per b.
__global__ void k_mse_grad(const float* __restrict__ Y, const float* __restrict__ y, float* __restrict__ dY,
                           float* __restrict__ sq, int n) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  float d = Y[b] - y[b];
  dY[b] = 2.f * d / (float)n;
  sq[b] = d * d;
}

// Weight gradient: G[i][j] = sum_b X[b][i] D[b][j]; bias gradient gb[j] = sum_b D[b][j].
__global__ void k_dense_wgrad(const float* __restrict__ X, const float* __restrict__ D, float* __restrict__ G,
                              float* __restrict__ gb, int n, int in, int out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < in * out) {
    int i = t / out, j = t % out;
    float acc = 0.f;
    for (int b = 0; b < n; ++b) acc = fmaf(X[b * in + i], D[b * out + j], acc);
    G[t] = acc;
  } else if (t < in * out + out) {
    int j = t - in * out;
    float acc = 0.f;
    for (int b = 0; b < n; ++b) acc += D[b * out + j];
    gb[j] = acc;
  }
}

// Input gradient through a ReLU layer: DX[b][i] = (sum_j D[b][j] W[i][j]) * (Hin[b][i] > 0).
__global__ void k_dense_bgrad(const float* __restrict__ D, const float* __restrict__ W, const float* __restrict__ Hin,
                              float* __restrict__ DX, int n, int in, int out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * in) return;
  int b = t / in, i = t % in;
  float acc = 0.f;
  for (int j = 0; j < out; ++j) acc = fmaf(D[b * out + j], W[i * out + j], acc);
  DX[t] = Hin[t] > 0.f ? acc : 0.f;
}

// Adam step on every parameter.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int np, float lr, float b1, float b2, float c1, float c2, float eps) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  float gi = g[t];
  float mi = b1 * m[t] + (1.f - b1) * gi;
  float vi = b2 * v[t] + (1.f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

struct pq_mlp {
  int in = 0, H = 0, np = 0, cap = 0;
  float* p = nullptr;   // parameters
  float* g = nullptr;   // gradient
  float* m = nullptr;   // Adam moments
  float* v = nullptr;
  float *X = nullptr, *y = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr;
  float *dY = nullptr, *dH2 = nullptr, *dH1 = nullptr, *sq = nullptr;
  int64_t adam_t = 0;
  // offsets into p / g
  int oW1, ob1, oW2, ob2, ow3, ob3;
};

namespace {

int mlp_alloc(pq_ctx* c, pq_mlp* m, int cap) {
  if (cap <= m->cap) return PQ_OK;
  for (float** q : {&m->X, &m->y, &m->H1, &m->H2, &m->Y, &m->dY, &m->dH2, &m->dH1, &m->sq})
    if (*q) {
      hipFree(*q);
      *q = nullptr;
    }
  const size_t f = sizeof(float);
  bool ok = hipMalloc((void**)&m->X, (size_t)cap * m->in * f) == hipSuccess &&
            hipMalloc((void**)&m->y, (size_t)cap * f) == hipSuccess &&
            hipMalloc((void**)&m->H1, (size_t)cap * m->H * f) == hipSuccess &&
            hipMalloc((void**)&m->H2, (size_t)cap * m->H * f) == hipSuccess &&
            hipMalloc((void**)&m->Y, (size_t)cap * f) == hipSuccess &&
            hipMalloc((void**)&m->dY, (size_t)cap * f) == hipSuccess &&
            hipMalloc((void**)&m->dH2, (size_t)cap * m->H * f) == hipSuccess &&
            hipMalloc((void**)&m->dH1, (size_t)cap * m->H * f) == hipSuccess &&
            hipMalloc((void**)&m->sq, (size_t)cap * f) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    return fail(c, PQ_ERR_OOM, "mlp buffers");
  }
  m->cap = cap;
  return PQ_OK;
}

void forward(pq_ctx* c, pq_mlp* m, int n) {
  hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((int64_t)n * m->H)), dim3(256), 0, c->stream, m->X, m->p + m->oW1,
                     m->p + m->ob1, m->H1, n, m->in, m->H, 1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((int64_t)n * m->H)), dim3(256), 0, c->stream, m->H1, m->p + m->oW2,
                     m->p + m->ob2, m->H2, n, m->H, m->H, 1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(n)), dim3(256), 0, c->stream, m->H2, m->p + m->ow3, m->p + m->ob3,
                     m->Y, n, m->H, 1, 0);
}

void backward(pq_ctx* c, pq_mlp* m, int n) {
  hipLaunchKernelGGL(k_mse_grad, dim3(blocks(n)), dim3(256), 0, c->stream, m->Y, m->y, m->dY, m->sq, n);
  // layer 3: w3[H][1]
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks(m->H + 1)), dim3(256), 0, c->stream, m->H2, m->dY, m->g + m->ow3,
                     m->g + m->ob3, n, m->H, 1);
  hipLaunchKernelGGL(k_dense_bgrad, dim3(blocks((int64_t)n * m->H)), dim3(256), 0, c->stream, m->dY, m->p + m->ow3,
                     m->H2, m->dH2, n, m->H, 1);
  // layer 2
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((int64_t)m->H * m->H + m->H)), dim3(256), 0, c->stream, m->H1,
                     m->dH2, m->g + m->oW2, m->g + m->ob2, n, m->H, m->H);
  hipLaunchKernelGGL(k_dense_bgrad, dim3(blocks((int64_t)n * m->H)), dim3(256), 0, c->stream, m->dH2, m->p + m->oW2,
                     m->H1, m->dH1, n, m->H, m->H);
  // layer 1
  hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((int64_t)m->in * m->H + m->H)), dim3(256), 0, c->stream, m->X,
                     m->dH1, m->g + m->oW1, m->g + m->ob1, n, m->in, m->H);
}

}  // namespace

extern "C" {

int pq_mlp_create(pq_ctx* c, int in_dim, int hidden, uint64_t seed, pq_mlp** out) {
  if (!c || !out || in_dim < 1 || in_dim > 256 || hidden < 1 || hidden > 1024) return PQ_ERR_ARG;
  pq_mlp* m = new pq_mlp;
  m->in = in_dim;
  m->H = hidden;
  m->oW1 = 0;
  m->ob1 = m->oW1 + in_dim * hidden;
  m->oW2 = m->ob1 + hidden;
  m->ob2 = m->oW2 + hidden * hidden;
  m->ow3 = m->ob2 + hidden;
  m->ob3 = m->ow3 + hidden;
  m->np = m->ob3 + 1;
  std::vector<float> init(m->np, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_in, int count) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan_in));
    for (int i = 0; i < count; ++i) init[off + i] = nd(rng);
  };
  fill(m->oW1, in_dim, in_dim * hidden);
  fill(m->oW2, hidden, hidden * hidden);
  fill(m->ow3, hidden, hidden);
  const size_t bytes = (size_t)m->np * sizeof(float);
  bool ok = hipMalloc((void**)&m->p, bytes) == hipSuccess && hipMalloc((void**)&m->g, bytes) == hipSuccess &&
            hipMalloc((void**)&m->m, bytes) == hipSuccess && hipMalloc((void**)&m->v, bytes) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    delete m;
    return fail(c, PQ_ERR_OOM, "mlp parameters");
  }
  hipMemcpy(m->p, init.data(), bytes, hipMemcpyHostToDevice);
  std::vector<float> z(m->np, 0.f);
  hipMemcpy(m->m, z.data(), bytes, hipMemcpyHostToDevice);
  hipMemcpy(m->v, z.data(), bytes, hipMemcpyHostToDevice);
  hipMemcpy(m->g, z.data(), bytes, hipMemcpyHostToDevice);
  *out = m;
  return PQ_OK;
}

int pq_mlp_destroy(pq_mlp* m) {
  if (!m) return PQ_ERR_ARG;
  for (float* q : {m->p, m->g, m->m, m->v, m->X, m->y, m->H1, m->H2, m->Y, m->dY, m->dH2, m->dH1, m->sq})
    if (q) hipFree(q);
  delete m;
  return PQ_OK;
}

int pq_mlp_num_params(const pq_mlp* m) { return m ? m->np : -1; }

int pq_mlp_get_params(pq_ctx* c, pq_mlp* m, float* host) {
  if (!c || !m || !host) return PQ_ERR_ARG;
  hipStreamSynchronize(c->stream);
  hipError_t e = hipMemcpy(host, m->p, (size_t)m->np * sizeof(float), hipMemcpyDeviceToHost);
  return e == hipSuccess ? PQ_OK : hipfail(c, e, "mlp params");
}

int pq_mlp_set_params(pq_ctx* c, pq_mlp* m, const float* host) {
  if (!c || !m || !host) return PQ_ERR_ARG;
  hipStreamSynchronize(c->stream);
  hipError_t e = hipMemcpy(m->p, host, (size_t)m->np * sizeof(float), hipMemcpyHostToDevice);
  m->adam_t = 0;
  return e == hipSuccess ? PQ_OK : hipfail(c, e, "mlp params");
}

int pq_mlp_predict(pq_ctx* c, pq_mlp* m, const float* X, int n, float* out) {
  if (!c || !m || !X || !out || n < 1) return PQ_ERR_ARG;
  int r = mlp_alloc(c, m, n);
  if (r) return r;
  hipMemcpyAsync(m->X, X, (size_t)n * m->in * sizeof(float), hipMemcpyHostToDevice, c->stream);
  forward(c, m, n);
  hipMemcpyAsync(out, m->Y, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? PQ_OK : hipfail(c, e, "mlp predict");
}

int pq_mlp_loss_grad(pq_ctx* c, pq_mlp* m, const float* X, const float* y, int n, float* loss, float* grad) {
  if (!c || !m || !X || !y || n < 1) return PQ_ERR_ARG;
  int r = mlp_alloc(c, m, n);
  if (r) return r;
  hipMemcpyAsync(m->X, X, (size_t)n * m->in * sizeof(float), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(m->y, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream);
  forward(c, m, n);
  backward(c, m, n);
  std::vector<float> sq(n);
  hipMemcpyAsync(sq.data(), m->sq, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (grad) hipMemcpyAsync(grad, m->g, (size_t)m->np * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hipfail(c, e, "mlp loss/grad");
  double s = 0;
  for (float v : sq) s += v;
  if (loss) *loss = (float)(s / n);
  return PQ_OK;
}

int pq_mlp_train(pq_ctx* c, pq_mlp* m, const float* X, const float* y, int n, int epochs, int batch, float lr,
                 uint64_t seed, float* final_loss) {
  if (!c || !m || !X || !y || n < 1 || epochs < 0 || batch < 1) return PQ_ERR_ARG;
  batch = std::min(batch, n);
  int r = mlp_alloc(c, m, std::max(batch, 1));
  if (r) return r;
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::mt19937_64 rng(seed);
  std::vector<float> xb((size_t)batch * m->in), yb(batch);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double last = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(idx.begin(), idx.end(), rng);
    double tot = 0;
    int cnt = 0;
    for (int s0 = 0; s0 + batch <= n; s0 += batch) {
      for (int b = 0; b < batch; ++b) {
        const int i = idx[s0 + b];
        memcpy(&xb[(size_t)b * m->in], X + (size_t)i * m->in, m->in * sizeof(float));
        yb[b] = y[i];
      }
      hipMemcpyAsync(m->X, xb.data(), xb.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
      hipMemcpyAsync(m->y, yb.data(), yb.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
      forward(c, m, batch);
      backward(c, m, batch);
      m->adam_t++;
      const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
      hipLaunchKernelGGL(k_adam, dim3(blocks(m->np)), dim3(256), 0, c->stream, m->p, m->g, m->m, m->v, m->np, lr,
                         b1, b2, c1, c2, eps);
      if (ep == epochs - 1) {
        std::vector<float> sq(batch);
        hipMemcpyAsync(sq.data(), m->sq, (size_t)batch * sizeof(float), hipMemcpyDeviceToHost, c->stream);
        hipStreamSynchronize(c->stream);
        for (float v : sq) tot += v;
        cnt += batch;
      }
    }
    if (cnt) last = tot / cnt;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hipfail(c, e, "mlp train");
  if (final_loss) *final_loss = (float)last;
  return PQ_OK;
}

}  // extern "C"
