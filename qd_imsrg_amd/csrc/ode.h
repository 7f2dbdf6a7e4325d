// Device-resident Shampine-Gordon Adams PECE integrator (the ODE path).
#pragma once
#include <functional>

#include "qd_internal.h"

namespace qd {

// right-hand side: yp = f(x, y), both device vectors of the integrator's length
using OdeRhs = std::function<int(double x, const double *y, double *yp)>;

struct AdamsDevice {
  hipStream_t s = nullptr;
  int64_t n = 0;
  // phi(:, 1..16) column-major (column i at phi + (i-1) n); yy = solution at x
  double *phi = nullptr, *yy = nullptr, *p = nullptr, *yp = nullptr, *wt = nullptr;
  double *part = nullptr, *scal = nullptr, *hscal = nullptr;
  double psi[12] = {}, alpha[12] = {}, beta[12] = {}, sig[13] = {}, v[12] = {}, w[12] = {}, g[13] = {};
  double x = 0, h = 0, hold = 0, told = 0, delsgn = 0;
  int k = 0, kold = 0, ns = 0, isnold = 0;
  bool start = false, phase1 = false, nornd = false;
  int64_t nfe = 0;
  int nsteps = 0;
  int maxnum = 500;  // steps per call before iflag 4 (src/adams_ode.f90:376)

  int init(hipStream_t stream, int64_t neq);
  void release();
  // ode(f, neqn, y, ..., t, tout, relerr, abserr, iflag, ...) of the reference
  int ode(const OdeRhs &f, double *y, double *t, double tout, double *relerr, double *abserr, int *iflag);

 private:
  int step(const OdeRhs &f, double *eps, bool *crash);
  int intrp(double xout, double *yout);
  int sums(int mode, const double *a, const double *phik, const double *phikm1, double out[3]);
  int rhs(const OdeRhs &f, const double *y, double *out);
};

}  // namespace qd
