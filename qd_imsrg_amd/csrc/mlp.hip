// Learned warm start: a small FP32 multilayer perceptron, trained and
// evaluated by the device kernels below, that maps oscillator-frequency
// features to the coefficients of a PCA basis of converged Magnus Omegas;
// and the expansion of such coefficients into a device-resident operator.
// No counterpart in the reference: s_evolve (src/magnus_IMSRG.f90:194-248)
// always starts from Omega = 0.
//
// Network: layer l maps A_{l-1} (batch x in_l, row-major) to
// Z_l = A_{l-1} W_l^T + b_l, A_l = tanh(Z_l) for hidden layers and A_L = Z_L
// at the output.  Parameters are one flat FP32 vector, per layer W_l
// (out_l x in_l, row-major) then b_l.  Loss = mean over batch and outputs of
// (A_L - T)^2.  Training: full-batch Adam (beta1 0.9, beta2 0.999, eps 1e-8).
// Every reduction runs in a fixed order (no atomics): results are
// deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qd_internal.h"

using namespace qd;

namespace {

__global__ void k_dense_fwd(const float *__restrict__ W, const float *__restrict__ b, const float *__restrict__ A,
                            float *__restrict__ Z, int B, int nin, int nout, int act) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * nout) return;
  const int r = idx / nout, o = idx % nout;
  const float *w = W + (size_t)o * nin, *a = A + (size_t)r * nin;
  float s = b[o];
  for (int i = 0; i < nin; i++) s = fmaf(w[i], a[i], s);
  Z[idx] = act ? tanhf(s) : s;
}

// G = dLoss/dY = 2 (Y - T) / n and loss = mean (Y - T)^2, one work-group
__global__ void __launch_bounds__(256) k_mse_grad(const float *__restrict__ Y, const float *__restrict__ T,
                                                  float *__restrict__ G, int n, float *__restrict__ loss) {
  __shared__ float red[256];
  const float sc = 2.0f / (float)n;
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) {
    const float d = Y[i] - T[i];
    G[i] = sc * d;
    s = fmaf(d, d, s);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] / (float)n;
}

// dW[o][i] = sum_r G[r][o] A[r][i], db[o] = sum_r G[r][o]
__global__ void k_dense_bwd_w(const float *__restrict__ G, const float *__restrict__ A, float *__restrict__ dW,
                              float *__restrict__ db, int B, int nin, int nout) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nout * (nin + 1)) return;
  const int o = idx / (nin + 1), i = idx % (nin + 1);
  float s = 0.0f;
  if (i < nin)
    for (int r = 0; r < B; r++) s = fmaf(G[(size_t)r * nout + o], A[(size_t)r * nin + i], s);
  else
    for (int r = 0; r < B; r++) s += G[(size_t)r * nout + o];
  if (i < nin) dW[(size_t)o * nin + i] = s;
  else db[o] = s;
}

// Gp[r][i] = (sum_o G[r][o] W[o][i]) * (1 - A[r][i]^2), A = tanh output feeding this layer
__global__ void k_dense_bwd_x(const float *__restrict__ G, const float *__restrict__ W, const float *__restrict__ A,
                              float *__restrict__ Gp, int B, int nin, int nout) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * nin) return;
  const int r = idx / nin, i = idx % nin;
  float s = 0.0f;
  for (int o = 0; o < nout; o++) s = fmaf(G[(size_t)r * nout + o], W[(size_t)o * nin + i], s);
  const float a = A[idx];
  Gp[idx] = s * (1.0f - a * a);
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float c1, float c2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// out[j] = mean[j] + sum_r coef[r] vecs[r][j]
__global__ void k_basis_expand(const double *__restrict__ mean, const double *__restrict__ vecs,
                               const double *__restrict__ coef, int rank, int64_t n, double *__restrict__ out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double s = mean[j];
    for (int r = 0; r < rank; r++) s = fma(coef[r], vecs[(size_t)r * n + j], s);
    out[j] = s;
  }
}

int blocks(int64_t n) { return (int)((n + 255) / 256); }

}  // namespace

struct qd_mlp {
  qd_context *ctx = nullptr;
  hipStream_t s = nullptr;
  std::vector<int> width;              // nin, hidden..., nout
  std::vector<int64_t> woff, boff;     // per layer offsets into the flat vector
  int64_t np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  int64_t adam_t = 0;
  int cap = 0;                         // batch capacity of the activation buffers
  std::vector<float *> act;            // per layer output (batch x width[l]), act[0] = input
  float *grad_a = nullptr, *grad_b = nullptr, *target = nullptr, *loss = nullptr;
  int nlayer() const { return (int)width.size() - 1; }
};

struct qd_basis {
  qd_context *ctx = nullptr;
  int rank = 0;
  int64_t n = 0;
  double *mean = nullptr, *vecs = nullptr, *coef = nullptr;
};

#define HIPM(expr)                                                                              \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return set_error(QD_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

namespace {

int ensure_batch(qd_mlp *m, int B) {
  if (B <= m->cap) return QD_OK;
  for (float *a : m->act) hipFree(a);
  hipFree(m->grad_a); hipFree(m->grad_b); hipFree(m->target);
  m->act.assign(m->width.size(), nullptr);
  int wmax = 0;
  for (int w : m->width) wmax = std::max(wmax, w);
  for (size_t l = 0; l < m->width.size(); l++) HIPM(hipMalloc(&m->act[l], sizeof(float) * (size_t)B * m->width[l]));
  HIPM(hipMalloc(&m->grad_a, sizeof(float) * (size_t)B * wmax));
  HIPM(hipMalloc(&m->grad_b, sizeof(float) * (size_t)B * wmax));
  HIPM(hipMalloc(&m->target, sizeof(float) * (size_t)B * m->width.back()));
  m->cap = B;
  return QD_OK;
}

void forward(qd_mlp *m, int B) {
  const int L = m->nlayer();
  for (int l = 0; l < L; l++) {
    const int nin = m->width[l], nout = m->width[l + 1];
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((int64_t)B * nout)), dim3(256), 0, m->s, m->p + m->woff[l],
                       m->p + m->boff[l], m->act[l], m->act[l + 1], B, nin, nout, l + 1 < L ? 1 : 0);
  }
}

// loss into *loss_dst (device) and the gradient of every parameter into m->g
// (forward must have run)
void backward(qd_mlp *m, int B, float *loss_dst) {
  const int L = m->nlayer();
  float *G = m->grad_a, *Gp = m->grad_b;
  hipLaunchKernelGGL(k_mse_grad, dim3(1), dim3(256), 0, m->s, m->act[L], m->target, G, B * m->width[L], loss_dst);
  for (int l = L - 1; l >= 0; l--) {
    const int nin = m->width[l], nout = m->width[l + 1];
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((int64_t)nout * (nin + 1))), dim3(256), 0, m->s, G, m->act[l],
                       m->g + m->woff[l], m->g + m->boff[l], B, nin, nout);
    if (l > 0) {
      hipLaunchKernelGGL(k_dense_bwd_x, dim3(blocks((int64_t)B * nin)), dim3(256), 0, m->s, G, m->p + m->woff[l],
                         m->act[l], Gp, B, nin, nout);
      std::swap(G, Gp);
    }
  }
}

int upload_batch(qd_mlp *m, const float *x, const float *t, int B) {
  int rc = ensure_batch(m, B);
  if (rc) return rc;
  HIPM(hipMemcpyAsync(m->act[0], x, sizeof(float) * (size_t)B * m->width[0], hipMemcpyHostToDevice, m->s));
  if (t) HIPM(hipMemcpyAsync(m->target, t, sizeof(float) * (size_t)B * m->width.back(), hipMemcpyHostToDevice, m->s));
  return QD_OK;
}

}  // namespace

extern "C" {

int qd_mlp_create(qd_context *ctx, int nin, int nhidden, int nlayers, int nout, uint64_t seed, qd_mlp **out) {
  if (!ctx || !out) return set_error(QD_E_ARG, "null argument");
  *out = nullptr;
  if (nin < 1 || nhidden < 1 || nlayers < 1 || nout < 1) return set_error(QD_E_ARG, "layer sizes must be positive");
  qd_mlp *m = new qd_mlp();
  m->ctx = ctx;
  m->s = context_stream(ctx);
  m->width.push_back(nin);
  for (int l = 0; l < nlayers; l++) m->width.push_back(nhidden);
  m->width.push_back(nout);
  std::vector<float> h;
  uint64_t st = seed;
  auto uni = [&]() {  // splitmix64 -> uniform [0, 1)
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < m->nlayer(); l++) {
    const int a = m->width[l], b = m->width[l + 1];
    const double lim = std::sqrt(6.0 / (a + b));  // Glorot uniform
    m->woff.push_back((int64_t)h.size());
    for (int64_t k = 0; k < (int64_t)a * b; k++) h.push_back((float)((2.0 * uni() - 1.0) * lim));
    m->boff.push_back((int64_t)h.size());
    for (int k = 0; k < b; k++) h.push_back(0.0f);
  }
  m->np = (int64_t)h.size();
  auto bail = [&](int rc) { qd_mlp_destroy(m); return rc; };
  for (float **q : {&m->p, &m->g, &m->m, &m->v})
    if (hipMalloc(q, sizeof(float) * m->np) != hipSuccess) return bail(set_error(QD_E_HIP, "hipMalloc failed"));
  if (hipMalloc(&m->loss, sizeof(float) * 2) != hipSuccess) return bail(set_error(QD_E_HIP, "hipMalloc failed"));
  if (hipMemcpy(m->p, h.data(), sizeof(float) * m->np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->m, 0, sizeof(float) * m->np) != hipSuccess || hipMemset(m->v, 0, sizeof(float) * m->np) != hipSuccess ||
      hipMemset(m->g, 0, sizeof(float) * m->np) != hipSuccess)
    return bail(set_error(QD_E_HIP, "parameter upload failed"));
  *out = m;
  return QD_OK;
}

int qd_mlp_destroy(qd_mlp *m) {
  if (!m) return QD_OK;
  if (m->s) hipStreamSynchronize(m->s);
  for (float *a : m->act) hipFree(a);
  for (float *q : {m->p, m->g, m->m, m->v, m->grad_a, m->grad_b, m->target, m->loss}) hipFree(q);
  delete m;
  return QD_OK;
}

int64_t qd_mlp_num_params(const qd_mlp *m) { return m ? m->np : set_error(QD_E_ARG, "null network"); }

int qd_mlp_get_params(qd_mlp *m, float *p, int64_t n) {
  if (!m || !p || n != m->np) return set_error(QD_E_ARG, "need %lld parameters", (long long)(m ? m->np : 0));
  HIPM(hipMemcpyAsync(p, m->p, sizeof(float) * n, hipMemcpyDeviceToHost, m->s));
  HIPM(hipStreamSynchronize(m->s));
  return QD_OK;
}

int qd_mlp_set_params(qd_mlp *m, const float *p, int64_t n) {
  if (!m || !p || n != m->np) return set_error(QD_E_ARG, "need %lld parameters", (long long)(m ? m->np : 0));
  HIPM(hipMemcpyAsync(m->p, p, sizeof(float) * n, hipMemcpyHostToDevice, m->s));
  HIPM(hipMemsetAsync(m->m, 0, sizeof(float) * n, m->s));
  HIPM(hipMemsetAsync(m->v, 0, sizeof(float) * n, m->s));
  m->adam_t = 0;
  HIPM(hipStreamSynchronize(m->s));
  return QD_OK;
}

int qd_mlp_forward(qd_mlp *m, const float *x, int batch, float *y) {
  if (!m || !x || !y || batch < 1) return set_error(QD_E_ARG, "null argument or empty batch");
  int rc = upload_batch(m, x, nullptr, batch);
  if (rc) return rc;
  forward(m, batch);
  HIPM(hipGetLastError());
  HIPM(hipMemcpyAsync(y, m->act.back(), sizeof(float) * (size_t)batch * m->width.back(), hipMemcpyDeviceToHost, m->s));
  HIPM(hipStreamSynchronize(m->s));
  return QD_OK;
}

int qd_mlp_loss_grad(qd_mlp *m, const float *x, const float *t, int batch, float *loss, float *grad) {
  if (!m || !x || !t || batch < 1) return set_error(QD_E_ARG, "null argument or empty batch");
  int rc = upload_batch(m, x, t, batch);
  if (rc) return rc;
  forward(m, batch);
  backward(m, batch, m->loss);
  HIPM(hipGetLastError());
  if (loss) HIPM(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->s));
  if (grad) HIPM(hipMemcpyAsync(grad, m->g, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->s));
  HIPM(hipStreamSynchronize(m->s));
  return QD_OK;
}

int qd_mlp_train(qd_mlp *m, const float *x, const float *t, int batch, int steps, float lr, float *loss_trace) {
  if (!m || !x || !t || batch < 1 || steps < 0) return set_error(QD_E_ARG, "bad argument");
  int rc = upload_batch(m, x, t, batch);
  if (rc) return rc;
  float *trace = nullptr;
  if (loss_trace && steps > 0) HIPM(hipMalloc(&trace, sizeof(float) * steps));
  for (int k = 0; k < steps; k++) {
    forward(m, batch);
    backward(m, batch, trace ? trace + k : m->loss);
    m->adam_t++;
    const float c1 = 1.0f - std::pow(0.9f, (float)m->adam_t), c2 = 1.0f - std::pow(0.999f, (float)m->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(m->np)), dim3(256), 0, m->s, m->p, m->g, m->m, m->v, m->np, lr, c1, c2);
  }
  HIPM(hipGetLastError());
  if (trace) {
    HIPM(hipMemcpyAsync(loss_trace, trace, sizeof(float) * steps, hipMemcpyDeviceToHost, m->s));
    HIPM(hipStreamSynchronize(m->s));
    hipFree(trace);
  }
  HIPM(hipStreamSynchronize(m->s));
  return QD_OK;
}

int qd_basis_create(qd_context *ctx, int rank, const double *mean, const double *vecs, qd_basis **out) {
  if (!ctx || !mean || (rank > 0 && !vecs) || rank < 0 || !out) return set_error(QD_E_ARG, "bad argument");
  *out = nullptr;
  qd_basis *b = new qd_basis();
  b->ctx = ctx;
  b->rank = rank;
  b->n = context_neq(ctx);
  const size_t n = (size_t)b->n;
  auto bail = [&](const char *what) {
    qd_basis_destroy(b);
    return set_error(QD_E_HIP, "%s failed", what);
  };
  if (hipMalloc(&b->mean, sizeof(double) * n) != hipSuccess) return bail("hipMalloc");
  if (hipMalloc(&b->vecs, sizeof(double) * n * std::max(rank, 1)) != hipSuccess) return bail("hipMalloc");
  if (hipMalloc(&b->coef, sizeof(double) * std::max(rank, 1)) != hipSuccess) return bail("hipMalloc");
  if (hipMemcpy(b->mean, mean, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) return bail("upload");
  if (rank && hipMemcpy(b->vecs, vecs, sizeof(double) * n * rank, hipMemcpyHostToDevice) != hipSuccess)
    return bail("upload");
  *out = b;
  return QD_OK;
}

int qd_basis_destroy(qd_basis *b) {
  if (!b) return QD_OK;
  hipStreamSynchronize(context_stream(b->ctx));
  hipFree(b->mean);
  hipFree(b->vecs);
  hipFree(b->coef);
  delete b;
  return QD_OK;
}

int qd_basis_expand(qd_basis *b, const double *coef, qd_op *out) {
  if (!b || !out || (b->rank && !coef)) return set_error(QD_E_ARG, "null argument");
  if (op_context(out) != b->ctx) return set_error(QD_E_ARG, "operator belongs to another context");
  hipStream_t s = context_stream(b->ctx);
  if (b->rank) HIPM(hipMemcpyAsync(b->coef, coef, sizeof(double) * b->rank, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_basis_expand, dim3(2048), dim3(256), 0, s, b->mean, b->vecs, b->coef, b->rank, b->n,
                     op_data(out));
  HIPM(hipGetLastError());
  HIPM(hipStreamSynchronize(s));
  return QD_OK;
}

}  // extern "C"
