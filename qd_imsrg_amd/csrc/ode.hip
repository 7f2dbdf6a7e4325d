// The ODE path: the Shampine-Gordon variable-order, variable-step Adams
// PECE integrator the reference uses for dH/ds = dGam(H)
// (src/adams_ode.f90: ode :7, de :218, step :562, intrp :1187), restated
// from the published algorithm with the solution, the modified divided
// differences phi(:, 1..16) and every work vector resident on the device.
// The scalar recurrences (psi, alpha, beta, sig, v, w, g; order and step
// selection) run on the host exactly as written in the algorithm; every
// vector operation is one fused element-wise kernel or a deterministic
// two-pass reduction.  Element-wise kernels do not contract a*b+c into FMAs,
// so each element is rounded as in the reference's statement order.
#include <cmath>

#include "ode.h"

namespace qd {

namespace {

constexpr int OPART = 1024;  // fixed partition of every reduction
const double GSTR[13] = {0.5, 0.0833, 0.0417, 0.0264, 0.0188, 0.0143, 0.0114,
                         0.00936, 0.00789, 0.00679, 0.00592, 0.00524, 0.00468};

struct PredictArgs {
  double g[13], beta[12], h;
  int k, nsp1, nornd;
};

__device__ __forceinline__ double *col(double *phi, int64_t n, int i) { return phi + (int64_t)(i - 1) * n; }

// block 2 of STEP: phi -> phi*, predicted solution p, predicted differences
__global__ void __launch_bounds__(256) k_predict(double *__restrict__ phi, const double *__restrict__ y,
                                                 double *__restrict__ p, int64_t n, PredictArgs a) {
#pragma clang fp contract(off)
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n; l += (int64_t)gridDim.x * 256) {
    for (int i = a.nsp1; i <= a.k; i++) col(phi, n, i)[l] = a.beta[i - 1] * col(phi, n, i)[l];
    col(phi, n, a.k + 2)[l] = col(phi, n, a.k + 1)[l];
    col(phi, n, a.k + 1)[l] = 0.0;
    double pl = 0.0;
    for (int j = 1; j <= a.k; j++) {
      const int i = a.k + 1 - j;
      pl = pl + a.g[i - 1] * col(phi, n, i)[l];
      col(phi, n, i)[l] = col(phi, n, i)[l] + col(phi, n, i + 1)[l];
    }
    if (a.nornd) {
      p[l] = y[l] + a.h * pl;
    } else {
      const double tau = a.h * pl - col(phi, n, 15)[l];
      const double pv = y[l] + tau;
      p[l] = pv;
      col(phi, n, 16)[l] = (pv - y[l]) - tau;
    }
  }
}

// sums over a fixed partition: up to three weighted sums of squares per launch
//   mode 0: s0 = sum (a / wt)^2
//   mode 1: t = yp - phi1; s0 = sum (t / wt)^2, s1 = sum ((phik + t) / wt)^2,
//           s2 = sum ((phikm1 + t) / wt)^2 (s1, s2 only where the column is given)
__global__ void __launch_bounds__(256) k_wsum(int mode, const double *__restrict__ a, const double *__restrict__ yp,
                                              const double *__restrict__ phi1, const double *__restrict__ phik,
                                              const double *__restrict__ phikm1, const double *__restrict__ wt,
                                              int64_t n, double *__restrict__ part) {
#pragma clang fp contract(off)
  const int64_t chunk = (n + OPART - 1) / OPART;
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int64_t l = lo + threadIdx.x; l < hi; l += 256) {
    const double r = 1.0 / wt[l];
    if (mode == 0) {
      const double x = a[l] * r;
      s0 = s0 + x * x;
    } else {
      const double t = yp[l] - phi1[l];
      const double x0 = t * r;
      s0 = s0 + x0 * x0;
      if (phik) { const double x1 = (phik[l] + t) * r; s1 = s1 + x1 * x1; }
      if (phikm1) { const double x2 = (phikm1[l] + t) * r; s2 = s2 + x2 * x2; }
    }
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[threadIdx.x * OPART + blockIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_wsum_final(const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double red[256];
  for (int q = 0; q < 3; q++) {
    double s = 0.0;
    for (int i = threadIdx.x; i < OPART; i += 256) s += part[q * OPART + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_weights(const double *__restrict__ yy, double *__restrict__ wt, int64_t n,
                                                 double releps, double abseps) {
#pragma clang fp contract(off)
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n; l += (int64_t)gridDim.x * 256)
    wt[l] = releps * fabs(yy[l]) + abseps;
}

// start of an integration: phi1 = yp, phi2 = 0 (and phi15 = 0 when rounding
// control is on)
__global__ void __launch_bounds__(256) k_start(double *__restrict__ phi, const double *__restrict__ yp, int64_t n,
                                               int zero15) {
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n; l += (int64_t)gridDim.x * 256) {
    col(phi, n, 1)[l] = yp[l];
    col(phi, n, 2)[l] = 0.0;
    if (zero15) col(phi, n, 15)[l] = 0.0;
  }
}

// block 3 of STEP: restore phi after a failed step
__global__ void __launch_bounds__(256) k_restore(double *__restrict__ phi, int64_t n, int k, PredictArgs a) {
#pragma clang fp contract(off)
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n; l += (int64_t)gridDim.x * 256)
    for (int i = 1; i <= k; i++)
      col(phi, n, i)[l] = (1.0 / a.beta[i - 1]) * (col(phi, n, i)[l] - col(phi, n, i + 1)[l]);
}

// block 4 of STEP: corrected solution
__global__ void __launch_bounds__(256) k_correct(double *__restrict__ phi, const double *__restrict__ p,
                                                 const double *__restrict__ yp, double *__restrict__ y, int64_t n,
                                                 double temp1, int nornd) {
#pragma clang fp contract(off)
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n; l += (int64_t)gridDim.x * 256) {
    if (nornd) {
      y[l] = p[l] + temp1 * (yp[l] - col(phi, n, 1)[l]);
    } else {
      const double rho = temp1 * (yp[l] - col(phi, n, 1)[l]) - col(phi, n, 16)[l];
      const double yv = p[l] + rho;
      y[l] = yv;
      col(phi, n, 15)[l] = (yv - p[l]) - rho;
    }
  }
}

// block 4 of STEP: differences for the next step
__global__ void __launch_bounds__(256) k_update(double *__restrict__ phi, const double *__restrict__ yp, int64_t n,
                                                int k) {
#pragma clang fp contract(off)
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n; l += (int64_t)gridDim.x * 256) {
    const double d = yp[l] - col(phi, n, 1)[l];
    col(phi, n, k + 1)[l] = d;
    col(phi, n, k + 2)[l] = d - col(phi, n, k + 2)[l];
    for (int i = 1; i <= k; i++) col(phi, n, i)[l] = col(phi, n, i)[l] + d;
  }
}

struct IntrpArgs {
  double g[13], rho[13], hi;
  int ki;
};

// INTRP: yout = y + hi sum_i g_i phi_i, ypout = sum_i rho_i phi_i (i = ki..1)
__global__ void __launch_bounds__(256) k_intrp(const double *__restrict__ phi, const double *__restrict__ yy,
                                               double *__restrict__ yout, double *__restrict__ ypout, int64_t n,
                                               IntrpArgs a) {
#pragma clang fp contract(off)
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n; l += (int64_t)gridDim.x * 256) {
    double s = 0.0, sp = 0.0;
    for (int j = 1; j <= a.ki; j++) {
      const int i = a.ki + 1 - j;
      const double ph = phi[(int64_t)(i - 1) * n + l];
      s = s + a.g[i - 1] * ph;
      sp = sp + a.rho[i - 1] * ph;
    }
    yout[l] = yy[l] + a.hi * s;
    if (ypout) ypout[l] = sp;
  }
}

__global__ void __launch_bounds__(256) k_axpy_nc(const double *__restrict__ x, const double *__restrict__ d,
                                                 double *__restrict__ z, int64_t n, double h) {
#pragma clang fp contract(off)
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n; l += (int64_t)gridDim.x * 256)
    z[l] = x[l] + h * d[l];
}

unsigned grid_of(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

}  // namespace

#define HIPO(expr)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return set_error(QD_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));  \
  } while (0)

int AdamsDevice::init(hipStream_t stream, int64_t neq) {
  s = stream;
  n = neq;
  const size_t b = sizeof(double) * (size_t)n;
  HIPO(hipMalloc(&phi, 16 * b));
  HIPO(hipMemsetAsync(phi, 0, 16 * b, s));
  for (double **q : {&yy, &p, &yp, &wt})
    HIPO(hipMalloc(q, b));
  HIPO(hipMalloc(&part, sizeof(double) * 3 * OPART));
  HIPO(hipMalloc(&scal, sizeof(double) * 4));
  HIPO(hipHostMalloc((void **)&hscal, sizeof(double) * 4, hipHostMallocDefault));
  return QD_OK;
}

void AdamsDevice::release() {
  if (s) hipStreamSynchronize(s);
  for (double *q : {phi, yy, p, yp, wt, part, scal}) hipFree(q);
  if (hscal) hipHostFree(hscal);
  phi = yy = p = yp = wt = part = scal = hscal = nullptr;
}

int AdamsDevice::sums(int mode, const double *a, const double *phik, const double *phikm1, double out[3]) {
  hipLaunchKernelGGL(k_wsum, dim3(OPART), dim3(256), 0, s, mode, a, yp, phi, phik, phikm1, wt, n, part);
  hipLaunchKernelGGL(k_wsum_final, dim3(1), dim3(256), 0, s, part, scal);
  HIPO(hipGetLastError());
  HIPO(hipMemcpyAsync(hscal, scal, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
  HIPO(hipStreamSynchronize(s));
  for (int q = 0; q < 3; q++) out[q] = hscal[q];
  return QD_OK;
}

int AdamsDevice::rhs(const OdeRhs &f, const double *y, double *out) {
  nfe++;
  return f(x, y, out);
}

// STEP (src/adams_ode.f90:562-1185)
int AdamsDevice::step(const OdeRhs &f, double *eps, bool *crash) {
  const double u = 2.220446049250313e-16, twou = 2.0 * u, fouru = 2.0 * twou;
  int rc;
  double sm[3];
  *crash = true;
  if (std::fabs(h) < fouru * std::fabs(x)) {
    h = std::copysign(fouru * std::fabs(x), h);
    return QD_OK;
  }
  const double p5eps = 0.5 * *eps;
  if ((rc = sums(0, yy, nullptr, nullptr, sm))) return rc;
  const double rnd = twou * std::sqrt(sm[0]);
  if (p5eps < rnd) {
    *eps = 2.0 * rnd * (1.0 + fouru);
    return QD_OK;
  }
  *crash = false;
  g[0] = 1.0;
  g[1] = 0.5;
  sig[0] = 1.0;
  if (start) {
    if ((rc = rhs(f, yy, yp))) return rc;
    if ((rc = sums(0, yp, nullptr, nullptr, sm))) return rc;
    const double total = std::sqrt(sm[0]);
    double absh = std::fabs(h);
    if (*eps < 16.0 * total * h * h) absh = 0.25 * std::sqrt(*eps / total);
    h = std::copysign(std::max(absh, fouru * std::fabs(x)), h);
    hold = 0.0;
    k = 1;
    kold = 0;
    start = false;
    phase1 = true;
    nornd = true;
    if (p5eps <= 100.0 * rnd) nornd = false;
    hipLaunchKernelGGL(k_start, dim3(grid_of(n)), dim3(256), 0, s, phi, yp, n, nornd ? 0 : 1);
    HIPO(hipGetLastError());
  }
  int ifail = 0;
  int knew;
  double absh, erk, erkm1 = 0.0, erkm2 = 0.0, xold;
  PredictArgs pa{};
  while (true) {
    // block 1: coefficients
    const int kp1 = k + 1, kp2 = k + 2, km1 = k - 1, km2 = k - 2;
    if (h != hold) ns = 0;
    if (ns <= kold) ns++;
    const int nsp1 = ns + 1;
    if (k >= ns) {
      beta[ns - 1] = 1.0;
      alpha[ns - 1] = 1.0 / ns;
      double temp1 = h * ns;
      sig[nsp1 - 1] = 1.0;
      for (int i = nsp1; i <= k; i++) {
        const double temp2 = psi[i - 2];
        psi[i - 2] = temp1;
        beta[i - 1] = beta[i - 2] * psi[i - 2] / temp2;
        temp1 = temp2 + h;
        alpha[i - 1] = h / temp1;
        sig[i] = i * alpha[i - 1] * sig[i - 1];
      }
      psi[k - 1] = temp1;
      if (ns <= 1) {
        for (int iq = 1; iq <= k; iq++) {
          v[iq - 1] = 1.0 / (double)(iq * (iq + 1));
          w[iq - 1] = v[iq - 1];
        }
      } else {
        if (k > kold) {
          v[k - 1] = 1.0 / (double)(k * kp1);
          for (int j = 1; j <= ns - 2; j++) {
            const int i = k - j;
            v[i - 1] -= alpha[j] * v[i];
          }
        }
        const double temp5 = alpha[ns - 1];
        for (int iq = 1; iq <= kp1 - ns; iq++) {
          v[iq - 1] -= temp5 * v[iq];
          w[iq - 1] = v[iq - 1];
        }
        g[nsp1 - 1] = w[0];
      }
      for (int i = ns + 2; i <= kp1; i++) {
        const double temp6 = alpha[i - 2];
        for (int iq = 1; iq <= kp2 - i; iq++) w[iq - 1] -= temp6 * w[iq];
        g[i - 1] = w[0];
      }
    }
    // block 2: predict, evaluate, error estimates
    for (int i = 0; i < 13; i++) pa.g[i] = g[i];
    for (int i = 0; i < 12; i++) pa.beta[i] = beta[i];
    pa.h = h;
    pa.k = k;
    pa.nsp1 = nsp1;
    pa.nornd = nornd ? 1 : 0;
    hipLaunchKernelGGL(k_predict, dim3(grid_of(n)), dim3(256), 0, s, phi, yy, p, n, pa);
    HIPO(hipGetLastError());
    xold = x;
    x = x + h;
    absh = std::fabs(h);
    if ((rc = rhs(f, p, yp))) return rc;
    if ((rc = sums(1, nullptr, km2 >= 0 ? phi + (int64_t)(k - 1) * n : nullptr,
                   km2 > 0 ? phi + (int64_t)(km1 - 1) * n : nullptr, sm)))
      return rc;
    erkm2 = erkm1 = 0.0;
    if (km2 > 0) erkm2 = absh * sig[km1 - 1] * GSTR[km2 - 1] * std::sqrt(sm[2]);
    if (km2 >= 0) erkm1 = absh * sig[k - 1] * GSTR[km1 - 1] * std::sqrt(sm[1]);
    const double serk = std::sqrt(sm[0]);
    const double err = absh * serk * (g[k - 1] - g[kp1 - 1]);
    erk = absh * serk * sig[kp1 - 1] * GSTR[k - 1];
    knew = k;
    if (km2 > 0) {
      if (std::max(erkm1, erkm2) <= erk) knew = km1;
    } else if (km2 == 0) {
      if (erkm1 <= 0.5 * erk) knew = km1;
    }
    if (err <= *eps) break;
    // block 3: failed step
    phase1 = false;
    x = xold;
    hipLaunchKernelGGL(k_restore, dim3(grid_of(n)), dim3(256), 0, s, phi, n, k, pa);
    HIPO(hipGetLastError());
    for (int i = 2; i <= k; i++) psi[i - 2] = psi[i - 1] - h;
    ifail++;
    double temp2 = 0.5;
    if (ifail > 3 && p5eps < 0.25 * erk) temp2 = std::sqrt(p5eps / erk);
    if (ifail >= 3) knew = 1;
    h = temp2 * h;
    k = knew;
    if (std::fabs(h) < fouru * std::fabs(x)) {
      *crash = true;
      h = std::copysign(fouru * std::fabs(x), h);
      *eps = *eps + *eps;
      return QD_OK;
    }
  }
  // block 4: successful step
  const int kp1 = k + 1, km1 = k - 1;
  kold = k;
  hold = h;
  hipLaunchKernelGGL(k_correct, dim3(grid_of(n)), dim3(256), 0, s, phi, p, yp, yy, n, h * g[kp1 - 1],
                     nornd ? 1 : 0);
  HIPO(hipGetLastError());
  if ((rc = rhs(f, yy, yp))) return rc;
  hipLaunchKernelGGL(k_update, dim3(grid_of(n)), dim3(256), 0, s, phi, yp, n, k);
  HIPO(hipGetLastError());
  double erkp1 = 0.0;
  if (knew == km1 || k == 12) phase1 = false;
  int action;  // 1 raise, -1 lower, 0 keep
  if (phase1) {
    action = 1;
  } else if (knew == km1) {
    action = -1;
  } else if (kp1 > ns) {
    action = 0;
  } else {
    if ((rc = sums(0, phi + (int64_t)(kp1 + 1 - 1) * n, nullptr, nullptr, sm))) return rc;
    erkp1 = absh * GSTR[kp1 - 1] * std::sqrt(sm[0]);
    if (k > 1) {
      if (erkm1 <= std::min(erk, erkp1)) action = -1;
      else if (erkp1 >= erk || k == 12) action = 0;
      else action = 1;
    } else {
      action = erkp1 >= 0.5 * erk ? 0 : 1;
    }
  }
  if (action == 1) {
    k = kp1;
    erk = erkp1;
  } else if (action == -1) {
    k = km1;
    erk = erkm1;
  }
  double hnew = h + h;
  if (!phase1 && p5eps < erk * std::ldexp(1.0, k + 1)) {
    hnew = h;
    if (p5eps < erk) {
      const double r = std::pow(p5eps / erk, 1.0 / (double)(k + 1));
      hnew = absh * std::max(0.5, std::min(0.9, r));
      hnew = std::copysign(std::max(hnew, fouru * std::fabs(x)), h);
    }
  }
  h = hnew;
  nsteps++;
  return QD_OK;
}

// INTRP (src/adams_ode.f90:1187-1312): solution at xout into yout
int AdamsDevice::intrp(double xout, double *yout) {
  IntrpArgs a{};
  const double hi = xout - x;
  const int ki = kold + 1, kip1 = ki + 1;
  double ww[13];
  a.g[0] = a.rho[0] = 1.0;
  for (int i = 1; i <= ki; i++) ww[i - 1] = 1.0 / (double)i;
  double term = 0.0;
  for (int j = 2; j <= ki; j++) {
    const double psijm1 = psi[j - 2];
    const double gamma = (hi + term) / psijm1, eta = hi / psijm1;
    for (int i = 1; i <= kip1 - j; i++) ww[i - 1] = gamma * ww[i - 1] - eta * ww[i];
    a.g[j - 1] = ww[0];
    a.rho[j - 1] = gamma * a.rho[j - 2];
    term = psijm1;
  }
  a.hi = hi;
  a.ki = ki;
  hipLaunchKernelGGL(k_intrp, dim3(grid_of(n)), dim3(256), 0, s, phi, yy, yout, nullptr, n, a);
  HIPO(hipGetLastError());
  return QD_OK;
}

// DE (src/adams_ode.f90:218-560) behind ODE (:7-216): integrate from *t to
// tout; y (device, n) holds the initial value on a (re)start and receives
// the solution at the returned *t.
int AdamsDevice::ode(const OdeRhs &f, double *y, double *t, double tout, double *relerr, double *abserr,
                     int *iflag) {
  const double fouru = 4.0 * 2.220446049250313e-16;
  if (n < 1 || *t == tout || *relerr < 0.0 || *abserr < 0.0) { *iflag = 6; return QD_OK; }
  double eps = std::max(*relerr, *abserr);
  if (eps <= 0.0 || *iflag == 0) { *iflag = 6; return QD_OK; }
  const int isn = *iflag > 0 ? 1 : -1;
  int fl = std::abs(*iflag);
  if (fl != 1 && (*t != told || fl < 2 || fl > 5)) { *iflag = 6; return QD_OK; }
  const double del = tout - *t, absdel = std::fabs(del);
  const double tend = isn < 0 ? tout : *t + 10.0 * del;
  int nostep = 0, kle4 = 0;
  bool stiff = false;
  const double releps = *relerr / eps, abseps = *abserr / eps;
  int rc;
  if (fl == 1 || isnold < 0 || delsgn * del <= 0.0) {
    start = true;
    x = *t;
    HIPO(hipMemcpyAsync(yy, y, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    delsgn = del >= 0.0 ? 1.0 : -1.0;
    h = std::copysign(std::max(std::fabs(tout - x), fouru * std::fabs(x)), tout - x);
  }
  while (true) {
    if (std::fabs(x - *t) >= absdel) {
      if ((rc = intrp(tout, y))) return rc;
      *iflag = 2;
      *t = tout;
      told = *t;
      isnold = isn;
      return QD_OK;
    }
    if (isn < 0 && std::fabs(tout - x) < fouru * std::fabs(x)) {
      const double hh = tout - x;
      if ((rc = rhs(f, yy, yp))) return rc;
      hipLaunchKernelGGL(k_axpy_nc, dim3(grid_of(n)), dim3(256), 0, s, yy, yp, y, n, hh);
      HIPO(hipGetLastError());
      *iflag = 2;
      *t = tout;
      told = *t;
      isnold = isn;
      return QD_OK;
    }
    if (nostep >= maxnum) {
      *iflag = isn * (stiff ? 5 : 4);
      HIPO(hipMemcpyAsync(y, yy, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      *t = x;
      told = *t;
      isnold = 1;
      return QD_OK;
    }
    h = std::copysign(std::min(std::fabs(h), std::fabs(tend - x)), h);
    hipLaunchKernelGGL(k_weights, dim3(grid_of(n)), dim3(256), 0, s, yy, wt, n, releps, abseps);
    HIPO(hipGetLastError());
    bool crash;
    if ((rc = step(f, &eps, &crash))) return rc;
    if (crash) {
      *iflag = isn * 3;
      *relerr = eps * releps;
      *abserr = eps * abseps;
      HIPO(hipMemcpyAsync(y, yy, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      *t = x;
      told = *t;
      isnold = 1;
      return QD_OK;
    }
    nostep++;
    kle4++;
    if (kold > 4) kle4 = 0;
    if (kle4 >= 50) stiff = true;
  }
}

}  // namespace qd
