// Input side of the flow: analytic oscillator Coulomb integrals and the
// Hartree-Fock H(0) in the reference's flat layout (host FP64).
//
// v_int restates src/ME_general.f90:2205-2276 (Anisimovas & Matulis, J. Phys.:
// Condens. Matter 10, 601 (1998)); binomials are exact integers and the sums
// run in long double.  The HF driver follows construct_two_particle_HF_BASIS
// (src/HFQDmod.f90:9-254) with the interaction built from v_int the way
// get_ints (src/ME_general.f90:746-798) does, instead of reading ME.bin.
// Eigenvector phases: largest-magnitude component positive.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include <omp.h>

#include "hf_device.h"
#include "qd_internal.h"

namespace qd {
namespace {

typedef long double ld;

ld binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  ld r = 1;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return std::round(r);
}
ld fact(int n) {
  ld r = 1;
  for (int i = 2; i <= n; i++) r *= i;
  return r;
}
// Gamma(x/2) for integer x >= 1
ld gamma_half(int twice) {
  if (twice % 2 == 0) return fact(twice / 2 - 1);
  const int xc = (twice - 1) / 2;
  return fact(2 * xc) / (std::pow((ld)4, xc) * fact(xc)) * 1.772453850905516027298167483341145182L;
}

// The inner sum cancels heavily for large g; it is accumulated in binary128
// (exact integer binomials and factorials up to ~2^113) and memoized per tuple.
typedef __float128 qf;
qf qbinom(int n, int k) {
  if (k < 0 || k > n) return 0;
  qf r = 1;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return r;
}
qf qfact(int n) {
  qf r = 1;
  for (int i = 2; i <= n; i++) r *= i;
  return r;
}
// Gamma(x/2) / sqrt(pi)^[x odd] as an exact binary128 rational value
qf qgamma_half_rational(int twice) {
  if (twice % 2 == 0) return qfact(twice / 2 - 1);
  const int xc = (twice - 1) / 2;
  qf p4 = 1;
  for (int i = 0; i < xc; i++) p4 *= 4;
  return qfact(2 * xc) / (p4 * qfact(xc));
}

// I(g1, g2, g3, g4): the innermost triple sum of v_int, in binary128
ld inner_value(int g1, int g2, int g3, int g4) {
  const int G = g1 + g2 + g3 + g4;
  // lam = 2(l1+l2) is even: Gamma(1+lam/2) is an integer and
  // Gamma((G-lam+1)/2) carries sqrt(pi) exactly when G is even
  qf s = 0;
  for (int l1 = 0; l1 <= g1; l1++)
    for (int l2 = 0; l2 <= g2; l2++)
      for (int l3 = 0; l3 <= g3; l3++) {
        const int l4 = l1 + l2 - l3;
        if (l4 < 0 || l4 > g4) continue;
        const int lam = l1 + l2 + l3 + l4;
        const qf term = qbinom(g1, l1) * qbinom(g2, l2) * qbinom(g3, l3) * qbinom(g4, l4) *
                        qgamma_half_rational(2 + lam) * qgamma_half_rational(G - lam + 1);
        if ((g2 + g3 - l2 - l3) & 1) s -= term; else s += term;
      }
  ld r = (ld)s;
  if ((G & 1) == 0) r *= 1.772453850905516027298167483341145182L;
  return r;
}

uint64_t inner_key(int g1, int g2, int g3, int g4) {
  return ((uint64_t)g1 << 48) | ((uint64_t)g2 << 32) | ((uint64_t)g3 << 16) | (uint64_t)g4;
}

// read-only table of I values, keys sorted (filled before a parallel build)
struct InnerTable {
  std::vector<uint64_t> keys;
  std::vector<ld> vals;
  const ld *find(uint64_t k) const {
    auto it = std::lower_bound(keys.begin(), keys.end(), k);
    return (it != keys.end() && *it == k) ? &vals[it - keys.begin()] : nullptr;
  }
};

struct InnerCache {
  std::unordered_map<uint64_t, ld> memo;
  ld get(int g1, int g2, int g3, int g4) {
    const uint64_t key = inner_key(g1, g2, g3, g4);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    const ld r = inner_value(g1, g2, g3, g4);
    memo.emplace(key, r);
    return r;
  }
};

InnerCache &cache() {
  static thread_local InnerCache c;
  return c;
}

// The quadruple sum over (j1, j2, j3, j4) of the reference's v_int
// (src/ME_general.f90:2205-2396) depends on j1, j4 only through j1 + j4 (in
// g1, g4 and the sign) and on j2, j3 only through j2 + j3, so it is summed as
// sum_{J14, J23} A(J14) B(J23) 2^{-(G+1)/2} I(g1, g2, g3, g4) with
// A(J) = sum_{j1+j4=J} (-1)^J binom(n1+|m1|, n1-j1) binom(n4+|m4|, n4-j4) / (j1! j4!)
// and B likewise: (n1+n4+1)(n2+n3+1) terms instead of (n1+1)(n2+1)(n3+1)(n4+1).
// Keys (g1, g2, g3, g4) the sum below visits for this quadruple
void inner_keys(const int *qi, const int *qj, const int *qk, const int *ql, std::vector<uint64_t> &out) {
  if (qi[2] != ql[2] || qk[2] != qj[2]) return;
  const int ai = std::abs(qi[1]), aj = std::abs(qj[1]), ak = std::abs(qk[1]), al = std::abs(ql[1]);
  for (int J14 = 0; J14 <= qi[0] + ql[0]; J14++)
    for (int J23 = 0; J23 <= qj[0] + qk[0]; J23++)
      out.push_back(inner_key(J14 + (ai + qi[1]) / 2 + (al - ql[1]) / 2, J23 + (aj + qj[1]) / 2 + (ak - qk[1]) / 2,
                              J23 + (aj - qj[1]) / 2 + (ak + qk[1]) / 2, J14 + (ai - qi[1]) / 2 + (al + ql[1]) / 2));
}

double v_int_impl(const int *qi, const int *qj, const int *qk, const int *ql, double hw,
                  const InnerTable *tab = nullptr) {
  if (qi[2] != ql[2] || qk[2] != qj[2]) return 0.0;
  const int ai = std::abs(qi[1]), aj = std::abs(qj[1]), ak = std::abs(qk[1]), al = std::abs(ql[1]);
  ld A[64], B[64];
  const int n14 = qi[0] + ql[0], n23 = qj[0] + qk[0];
  for (int J = 0; J <= n14; J++) A[J] = 0;
  for (int J = 0; J <= n23; J++) B[J] = 0;
  for (int j1 = 0; j1 <= qi[0]; j1++)
    for (int j4 = 0; j4 <= ql[0]; j4++)
      A[j1 + j4] += (((j1 + j4) & 1) ? -1 : 1) / (fact(j1) * fact(j4)) * binom(qi[0] + ai, qi[0] - j1) *
                    binom(ql[0] + al, ql[0] - j4);
  for (int j2 = 0; j2 <= qj[0]; j2++)
    for (int j3 = 0; j3 <= qk[0]; j3++)
      B[j2 + j3] += (((j2 + j3) & 1) ? -1 : 1) / (fact(j2) * fact(j3)) * binom(qj[0] + aj, qj[0] - j2) *
                    binom(qk[0] + ak, qk[0] - j3);
  ld sm = 0;
  for (int J14 = 0; J14 <= n14; J14++) {
    const int g1 = J14 + (ai + qi[1]) / 2 + (al - ql[1]) / 2;
    const int g4 = J14 + (ai - qi[1]) / 2 + (al + ql[1]) / 2;
    for (int J23 = 0; J23 <= n23; J23++) {
      const int g2 = J23 + (aj + qj[1]) / 2 + (ak - qk[1]) / 2;
      const int g3 = J23 + (aj - qj[1]) / 2 + (ak + qk[1]) / 2;
      const int G = g1 + g2 + g3 + g4;
      const ld *hit = tab ? tab->find(inner_key(g1, g2, g3, g4)) : nullptr;
      sm += A[J14] * B[J23] * std::pow((ld)0.5, (ld)(G + 1) * 0.5L) * (hit ? *hit : cache().get(g1, g2, g3, g4));
    }
  }
  const ld nrm = fact(qi[0]) / fact(qi[0] + ai) * fact(qj[0]) / fact(qj[0] + aj) * fact(qk[0]) /
                 fact(qk[0] + ak) * fact(ql[0]) / fact(ql[0] + al);
  return (double)(std::sqrt((ld)hw) * sm * std::sqrt(nrm));
}

}  // namespace

// cyclic Jacobi eigensolver for a small symmetric matrix (row-major a[n*n]);
// eigenvalues ascending, eigenvectors in the columns of v
void jacobi(int n, std::vector<double> a, std::vector<double> &w, std::vector<double> &v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double th = (aqq - app) / (2 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - s * akq;
          a[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - s * aqk;
          a[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = c * vkp - s * vkq;
          v[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
  std::vector<double> vs((size_t)n * n);
  w.resize(n);
  for (int k = 0; k < n; k++) {
    w[k] = a[(size_t)ord[k] * n + ord[k]];
    int big = 0;
    for (int i = 0; i < n; i++)
      if (std::fabs(v[(size_t)i * n + ord[k]]) > std::fabs(v[(size_t)big * n + ord[k]])) big = i;
    const double sg = v[(size_t)big * n + ord[k]] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; i++) vs[(size_t)i * n + k] = sg * v[(size_t)i * n + ord[k]];
  }
  v.swap(vs);
}

namespace {

// <ab|V|cd> at hw = 1 for m_a + m_b = m_c + m_d: for each (a, b, c) the d run
// over the orbitals with m_d = m_a + m_b - m_c (compressed, N^3 x ~R/2)
struct Spatial {
  int N = 0, R = 0, mmax = 0;
  std::vector<int> n, m;
  std::vector<std::vector<int>> by_m;   // orbitals with m = mm - mmax
  std::vector<int> pos_in_m;            // position of an orbital in its m list
  std::vector<int64_t> start;           // N^3
  std::vector<double> tab;
  std::vector<uint8_t> present;         // table read from ME.bin: element given in the file
  double get(int a, int b, int c, int d) const {
    const int md = m[a] + m[b] - m[c];
    if (md != m[d]) return 0.0;
    return tab[start[((size_t)a * N + b) * N + c] + pos_in_m[d]];
  }
  // the exchange element <ab|V|dc> as construct_two_particle_HF_BASIS reads
  // it from ME.bin (src/HFQDmod.f90:163-166): if it is (near) zero there,
  // <ba|V|cd> is taken instead; for a complete table the two are equal
  double get_exch(int a, int b, int c, int d) const {
    const double x = get(a, b, d, c);
    if (present.empty() || std::fabs(x) >= 1e-10) return x;
    return get(b, a, c, d);
  }
};

// orbitals (shell, then m ascending), per-m lists and the compressed offsets
// of the (a, b, c) runs; the table is zero-filled
void spatial_layout(int nshell, Spatial &S) {
  S.R = nshell;
  for (int R = 0; R < nshell; R++)
    for (int m = -R; m <= R; m += 2) { S.n.push_back((R - std::abs(m)) / 2); S.m.push_back(m); }
  const int N = S.N = (int)S.n.size();
  S.mmax = nshell - 1;
  S.by_m.assign(2 * S.mmax + 1, {});
  S.pos_in_m.assign(N, 0);
  for (int a = 0; a < N; a++) {
    S.pos_in_m[a] = (int)S.by_m[S.m[a] + S.mmax].size();
    S.by_m[S.m[a] + S.mmax].push_back(a);
  }
  S.start.assign((size_t)N * N * N, 0);
  int64_t k = 0;
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++)
      for (int c = 0; c < N; c++) {
        S.start[((size_t)a * N + b) * N + c] = k;
        const int md = S.m[a] + S.m[b] - S.m[c];
        if (md >= -S.mmax && md <= S.mmax) k += (int64_t)S.by_m[md + S.mmax].size();
      }
  S.tab.assign(k, 0.0);
}

void build_spatial(int nshell, Spatial &S) {
  spatial_layout(nshell, S);
  const int N = S.N;
  // <ab|V|cd> is invariant under (a,b,c,d) -> (b,a,d,c) (particle exchange),
  // -> (c,d,a,b) (real and symmetric) and under m -> -m on all four orbitals
  // (time reversal); each orbit of this group of 8 is evaluated once, at its
  // lexicographically smallest member, and copied to the others.
  std::vector<int> rev(N);
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++)
      if (S.n[b] == S.n[a] && S.m[b] == -S.m[a]) rev[a] = b;
  auto canon = [&](int a, int b, int c, int d) {
    std::array<int, 4> best{a, b, c, d};
    const std::array<int, 4> base[2] = {{a, b, c, d}, {rev[a], rev[b], rev[c], rev[d]}};
    for (const auto &t : base) {
      const std::array<int, 4> im[4] = {{t[0], t[1], t[2], t[3]}, {t[1], t[0], t[3], t[2]},
                                        {t[2], t[3], t[0], t[1]}, {t[3], t[2], t[1], t[0]}};
      for (const auto &x : im) best = std::min(best, x);
    }
    return best;
  };
  auto at = [&](int a, int b, int c, int d) -> double & {
    return S.tab[S.start[((size_t)a * N + b) * N + c] + S.pos_in_m[d]];
  };
  // every inner sum I(g1..g4) the owned elements need, computed once (in
  // parallel) before the elements themselves
  InnerTable itab;
  {
    std::vector<std::vector<uint64_t>> per(omp_get_max_threads());
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int a = 0; a < N; a++)
      for (int b = 0; b < N; b++)
        for (int c = 0; c < N; c++) {
          const int md = S.m[a] + S.m[b] - S.m[c];
          if (md < -S.mmax || md > S.mmax) continue;
          auto &out = per[omp_get_thread_num()];
          for (int d : S.by_m[md + S.mmax]) {
            if (canon(a, b, c, d) != std::array<int, 4>{a, b, c, d}) continue;
            const int qa[3] = {S.n[a], S.m[a], 1}, qb[3] = {S.n[b], S.m[b], 1};
            const int qc[3] = {S.n[c], S.m[c], 1}, qd[3] = {S.n[d], S.m[d], 1};
            inner_keys(qa, qb, qd, qc, out);
          }
          if (out.size() > (1u << 20)) {  // keep the per-thread lists short
            std::sort(out.begin(), out.end());
            out.erase(std::unique(out.begin(), out.end()), out.end());
          }
        }
    for (auto &v : per) {
      itab.keys.insert(itab.keys.end(), v.begin(), v.end());
      std::vector<uint64_t>().swap(v);
    }
    std::sort(itab.keys.begin(), itab.keys.end());
    itab.keys.erase(std::unique(itab.keys.begin(), itab.keys.end()), itab.keys.end());
    itab.vals.resize(itab.keys.size());
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t i = 0; i < itab.keys.size(); i++) {
      const uint64_t k = itab.keys[i];
      itab.vals[i] = inner_value((int)(k >> 48), (int)((k >> 32) & 0xffff), (int)((k >> 16) & 0xffff), (int)(k & 0xffff));
    }
  }
  for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int a = 0; a < N; a++)
      for (int b = 0; b < N; b++)
        for (int c = 0; c < N; c++) {
          const int md = S.m[a] + S.m[b] - S.m[c];
          if (md < -S.mmax || md > S.mmax) continue;
          for (int d : S.by_m[md + S.mmax]) {
            const std::array<int, 4> cn = canon(a, b, c, d);
            const bool own = cn == std::array<int, 4>{a, b, c, d};
            if (pass == 0 && own) {
              // <ab|V|cd> = v_int(a, b, d, c) with equal spins
              const int qa[3] = {S.n[a], S.m[a], 1}, qb[3] = {S.n[b], S.m[b], 1};
              const int qc[3] = {S.n[c], S.m[c], 1}, qd[3] = {S.n[d], S.m[d], 1};
              at(a, b, c, d) = v_int_impl(qa, qb, qd, qc, 1.0, &itab);
            } else if (pass == 1 && !own) {
              at(a, b, c, d) = at(cn[0], cn[1], cn[2], cn[3]);
            }
          }
        }
  }
}

// one table per shell count per process (hw enters as sqrt(hw), applied on use)
const Spatial &spatial_table(int nshell) {
  static std::mutex mu;
  static std::map<int, Spatial> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(nshell);
  if (it == cache.end()) {
    Spatial S;
    build_spatial(nshell, S);
    it = cache.emplace(nshell, std::move(S)).first;
  }
  return it->second;
}

}  // namespace
}  // namespace qd

using namespace qd;

extern "C" {

double qd_v_int(const int *qi, const int *qj, const int *qk, const int *ql, double hw) {
  return v_int_impl(qi, qj, qk, ql, hw);
}

}  // extern "C"

namespace qd {
namespace {
// construct_two_particle_HF_BASIS + HF_E2 on a Coulomb table S (<ab|V|cd> at
// hw = 1 over the spatial orbitals of nshell shells)
// dctx/dout: the Fock builds and the two-body transform run on the device and
// H(0) is written into the device operator dout (h0 unused)
int hf_build(int nbody, int nshell, double hw, const Spatial &S, double *h0, int64_t neq, double *eig_out,
             int *iters, qd_context *dctx = nullptr, qd_op *dout = nullptr) {
  Channels ch;
  std::string why = build_channels(nbody, nshell, ch);
  if (!why.empty()) return set_error(QD_E_LAYOUT, "%s", why.c_str());
  if (!h0 && !dout) return set_error(QD_E_ARG, "h0 is null");
  if (neq != ch.neq) return set_error(QD_E_ARG, "h0 holds %lld elements, the layout needs %lld", (long long)neq,
                                      (long long)ch.neq);
  const int M = ch.M, n = ch.n;
  const bool trace = std::getenv("QD_HF_TRACE") != nullptr;
  auto now = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
  double t_mark = now();
  const double shw = std::sqrt(hw);
  // spin-orbital -> spatial orbital (nm_index order), spin
  std::vector<int> sp(M), spin(M);
  {
    int k = 0;
    for (int R = 0; R < nshell; R++)
      for (int m = -R; m <= R; m += 2) {
        const int a = R * (R + 1) / 2 + (R + m) / 2;  // nm_index - 1
        for (int s = -1; s <= 1; s += 2) { sp[k] = a; spin[k] = s; k++; }
      }
  }
  auto asym = [&](int p, int q, int r, int s) -> double {  // <pq|V|rs>_A
    double v = 0;
    if (spin[p] == spin[r] && spin[q] == spin[s]) v += S.get(sp[p], sp[q], sp[r], sp[s]);
    if (spin[p] == spin[s] && spin[q] == spin[r]) v -= S.get_exch(sp[p], sp[q], sp[r], sp[s]);
    return shw * v;
  };
  // single-particle channels
  std::vector<std::vector<int>> chan(ch.nch);
  for (int i = 0; i < M; i++) chan[ch.spch[i]].push_back(i);
  std::vector<double> Tm(M), C((size_t)M * M, 0.0), eig(M, 0.0), F((size_t)M * M, 0.0);
  for (int i = 0; i < M; i++) {
    const int R = 2 * S.n[sp[i]] + std::abs(S.m[sp[i]]);
    Tm[i] = hw * (R + 1);
    C[(size_t)i * M + i] = 1.0;  // C[ho*M + hf]
  }
  HfDevice *dev = nullptr;
  struct DevGuard {
    HfDevice *&d;
    ~DevGuard() { hf_device_destroy(d); }
  } guard{dev};
  if (dctx) {
    const CoulombView cv{S.N, S.tab.data(), (int64_t)S.tab.size(), S.start.data(), S.pos_in_m.data(), S.m.data(),
                         !S.present.empty()};
    int rc = hf_device_create(dctx, cv, sp, spin, shw, &dev);
    if (rc) return rc;
  }
  // Roothaan iteration as in the reference; if it has not converged after
  // 100 iterations (it can settle into a 2-cycle, e.g. N=20, hw=0.5, R>=10,
  // where the reference's loop never ends) restart from the oscillator basis
  // with the density mixed half/half with the previous one, which has the
  // same fixed point.
  double crit = 1.0;
  int it = 0, total = 0;
  double mix = 0.0;
  std::vector<double> rho_old;
  while (crit > 1e-6) {
    if (it == 100 && mix == 0.0) {
      mix = 0.5;
      it = 0;
      rho_old.clear();
      std::fill(C.begin(), C.end(), 0.0);
      for (int i = 0; i < M; i++) C[(size_t)i * M + i] = 1.0;
      std::fill(eig.begin(), eig.end(), 0.0);
    }
    if (it >= 2000) break;
    it++;
    total++;
    // rho_lj = sum_{h<n} C[l,h] C[j,h] (block-diagonal)
    std::vector<double> rho((size_t)M * M, 0.0);
    for (auto &cs : chan)
      for (int l : cs)
        for (int j : cs) {
          double s = 0;
          for (int h = 0; h < n; h++) s += C[(size_t)l * M + h] * C[(size_t)j * M + h];
          rho[(size_t)l * M + j] = s;
        }
    if (mix > 0.0) {
      if (!rho_old.empty())
        for (size_t e = 0; e < rho.size(); e++) rho[e] = (1.0 - mix) * rho[e] + mix * rho_old[e];
      rho_old = rho;
    }
    std::fill(F.begin(), F.end(), 0.0);
    if (dev) {
      int rc = hf_device_fock(dev, rho, Tm, F);
      if (rc) return rc;
    } else {
#pragma omp parallel for schedule(dynamic, 1)
      for (int cc = 0; cc < (int)chan.size(); cc++)
        for (int i : chan[cc])
          for (int k : chan[cc]) {
            double v = 0;
            for (auto &cj : chan)
              for (int j : cj)
                for (int l : cj) {
                  const double r = rho[(size_t)l * M + j];
                  if (r != 0.0) v += r * asym(j, i, l, k);
                }
            F[(size_t)i * M + k] = (i == k ? Tm[i] : 0.0) + v;
          }
    }
    std::vector<double> eold = eig;
    std::fill(C.begin(), C.end(), 0.0);
    for (auto &cs : chan) {
      const int d = (int)cs.size();
      std::vector<double> a((size_t)d * d), w, v;
      for (int x = 0; x < d; x++)
        for (int y = 0; y < d; y++) a[(size_t)x * d + y] = F[(size_t)cs[x] * M + cs[y]];
      jacobi(d, a, w, v);
      for (int k = 0; k < d; k++) {
        eig[cs[k]] = w[k];
        for (int x = 0; x < d; x++) C[(size_t)cs[x] * M + cs[k]] = v[(size_t)x * d + k];
      }
    }
    crit = 0;
    for (int i = 0; i < n; i++) crit += std::fabs(eold[i] - eig[i]);
    if (trace) {
      double e = 0;
      for (int i = 0; i < n; i++) e += eig[i];
      fprintf(stderr, "hf iter %d crit %.3e sum_eocc %.10f\n", it, crit, e);
    }
  }
  if (trace) fprintf(stderr, "hf: %d SCF iterations %.2f s\n", total, now() - t_mark);
  t_mark = now();
  // F_HF = C^T F C (F from the last Fock build)
  std::vector<double> FC((size_t)M * M, 0.0), FHF((size_t)M * M, 0.0);
  for (int i = 0; i < M; i++)
    for (int b = 0; b < M; b++) {
      double s = 0;
      for (int k = 0; k < M; k++) s += F[(size_t)i * M + k] * C[(size_t)k * M + b];
      FC[(size_t)i * M + b] = s;
    }
  for (int a = 0; a < M; a++)
    for (int b = 0; b < M; b++) {
      double s = 0;
      for (int i = 0; i < M; i++) s += C[(size_t)i * M + a] * FC[(size_t)i * M + b];
      FHF[(size_t)a * M + b] = s;
    }
  // pack: E0 set below, one-body parts column-major
  std::vector<double> hbuf;
  if (dev) hbuf.assign((size_t)ch.off_2b, 0.0);  // one-body region, uploaded below
  double *hv = dev ? hbuf.data() : h0;
  if (!dev) std::memset(h0, 0, sizeof(double) * neq);
  const int mp = ch.mp;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) hv[ch.off_fhh + i + (int64_t)j * n] = FHF[(size_t)i * M + j];
  for (int j = 0; j < mp; j++)
    for (int i = 0; i < mp; i++) hv[ch.off_fpp + i + (int64_t)j * mp] = FHF[(size_t)(n + i) * M + n + j];
  for (int j = 0; j < n; j++)
    for (int i = 0; i < mp; i++) hv[ch.off_fph + i + (int64_t)j * mp] = FHF[(size_t)(n + i) * M + j];
  // two-body: per block HFV = T V_HO T^T over the block's pairs (hh | ph | pp).
  // T[x][z] = C[b1,a1] C[b2,a2] - C[b2,a1] C[b1,a2] for HF pair x = (a1,a2) and
  // oscillator pair z = (b1,b2): non-zero only for b1, b2 in the channels of
  // a1, a2 (either order), so each row is built from those channels alone
  auto t_rows = [&](int q) {
    const int nh = ch.nh[q], nb = ch.nb[q], np = ch.np[q], nt = nh + nb + np;
    const uint32_t *P = &ch.pairs[ch.hh_start[q]];  // hh, ph, pp are contiguous
    std::vector<std::vector<std::pair<int, double>>> T(nt);
    auto comb = [&](int i, int j) {  // combined (hh | ph | pp) index of pair i < j in block q
      const uint32_t ia = ch.iarr[(size_t)i * M - (size_t)i * (i + 1) / 2 + (j - i - 1)];
      const int t = ia_type(ia), k = ia_idx(ia);
      return t == T_HH ? k : (t == T_PH ? nh + k : nh + nb + k);
    };
    for (int x = 0; x < nt; x++) {
      const int a1 = lo16(P[x]), a2 = hi16(P[x]);
      const auto &c1 = chan[ch.spch[a1]], &c2 = chan[ch.spch[a2]];
      for (int b1 : c1)
        for (int b2 : c2) {
          if (b1 == b2) continue;
          const int lo = std::min(b1, b2), hi = std::max(b1, b2);
          if (ch.spch[b1] == ch.spch[b2] && b1 > b2) continue;  // same channel: each unordered pair once
          const double t = C[(size_t)lo * M + a1] * C[(size_t)hi * M + a2] - C[(size_t)hi * M + a1] * C[(size_t)lo * M + a2];
          if (t != 0.0) T[x].push_back({comb(lo, hi), t});
        }
    }
    return T;
  };
  if (dev) {
    hipStream_t st = context_stream(dctx);
    if (hipMemcpyAsync(op_data(dout), hbuf.data(), sizeof(double) * hbuf.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemsetAsync(op_data(dout) + ch.off_2b, 0, sizeof(double) * (size_t)(neq - ch.off_2b), st) != hipSuccess)
      return set_error(QD_E_HIP, "upload of the one-body part failed");
    std::vector<std::vector<std::vector<std::pair<int, double>>>> rows(ch.nblock);
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < ch.nblock; q++) rows[q] = t_rows(q);
    SparseRows Tr;
    Tr.row_start.push_back(0);
    for (int q = 0; q < ch.nblock; q++)
      for (const auto &r : rows[q]) {
        for (const auto &zt : r) {
          Tr.col.push_back(zt.first);
          Tr.val.push_back(zt.second);
        }
        Tr.row_start.push_back((int)Tr.col.size());
      }
    int rc;
    if ((rc = hf_device_transform(dev, Tr, dout))) return rc;
  } else {
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < ch.nblock; q++) {
      const int nh = ch.nh[q], nb = ch.nb[q], np = ch.np[q], nt = nh + nb + np;
      if (nt == 0) continue;
      const uint32_t *P = &ch.pairs[ch.hh_start[q]];
      std::vector<double> V((size_t)nt * nt), TV((size_t)nt * nt, 0.0), H((size_t)nt * nt, 0.0);
      const auto T = t_rows(q);
      for (int x = 0; x < nt; x++)
        for (int y = 0; y < nt; y++) V[(size_t)x * nt + y] = asym(lo16(P[x]), hi16(P[x]), lo16(P[y]), hi16(P[y]));
      for (int x = 0; x < nt; x++)
        for (const auto &zt : T[x]) {
          const double *v = &V[(size_t)zt.first * nt];
          double *tv = &TV[(size_t)x * nt];
          for (int y = 0; y < nt; y++) tv[y] += zt.second * v[y];
        }
      for (int x = 0; x < nt; x++)
        for (int y = 0; y < nt; y++) {
          double sum = 0;
          for (const auto &zt : T[y]) sum += TV[(size_t)x * nt + zt.first] * zt.second;
          H[(size_t)x * nt + y] = sum;
        }
      const int64_t *o = &ch.off[(size_t)q * 6];
      auto put = [&](int sub, int r0, int rows, int c0, int cols) {
        for (int c = 0; c < cols; c++)
          for (int r = 0; r < rows; r++) h0[o[sub] + r + (int64_t)c * rows] = H[(size_t)(r0 + r) * nt + c0 + c];
      };
      const int h = 0, x = nh, p = nh + nb;
      put(HHHH, h, nh, h, nh);
      put(PPPP, p, np, p, np);
      put(PPHH, p, np, h, nh);
      put(PHPH, x, nb, x, nb);
      put(PPPH, p, np, x, nb);
      put(PHHH, x, nb, h, nh);
    }
  }
  if (trace) fprintf(stderr, "hf: basis transform %.2f s\n", now() - t_mark);
  double eh = 0;
  for (int i = 0; i < n; i++) eh += eig[i];
  if (dev) {
    int rc = hf_device_e0(dctx, eh, dout);
    if (rc) return rc;
  } else {
    double vsum = 0;  // sum over blocks of the Vhhhh diagonal, in block order
    for (int q = 0; q < ch.nblock; q++)
      for (int j = 0; j < ch.nh[q]; j++) vsum += h0[ch.off[(size_t)q * 6 + HHHH] + j + (int64_t)j * ch.nh[q]];
    h0[0] = eh - vsum;  // HF_E2 (src/HFQDmod.f90:292-323)
  }
  if (eig_out) std::memcpy(eig_out, eig.data(), sizeof(double) * M);
  if (iters) *iters = total;
  if (crit > 1e-6) return set_error(QD_E_NUMERIC, "Hartree-Fock did not converge (crit %g after %d iterations)", crit, total);
  return QD_OK;
}

// nm_index (src/ME_general.f90:2140-2149) less one: shell R = 2n + |m|
int nm_orbital(int n, int m) {
  const int R = 2 * n + std::abs(m);
  return R * (R + 1) / 2 + (R + m) / 2;
}

// ME.bin as construct_two_particle_HF_BASIS reads it (src/HFQDmod.f90:97-129):
// a stream of records n1,m1,n2,m2,n3,m3,n4,m4 (int32) and Vdir (real64),
// Vdir = <ab|V|cd> at hw = 1 with a = (n1,m1) .. d = (n4,m4); records with an
// orbital beyond nshell shells are skipped; each record sets <ab|V|cd> and
// <cd|ab> (the file's later records win).  Records that break m_a + m_b =
// m_c + m_d are never read by the HF and are dropped.
int load_me_bin(const char *path, int nshell, Spatial &S, int64_t *nrec) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return set_error(QD_E_ARG, "cannot open %s", path);
  spatial_layout(nshell, S);
  S.present.assign(S.tab.size(), 0);
  const int N = S.N;
  auto slot = [&](int a, int b, int c, int d) -> int64_t {
    if (S.m[a] + S.m[b] != S.m[c] + S.m[d]) return -1;
    return S.start[((size_t)a * N + b) * N + c] + S.pos_in_m[d];
  };
  int64_t n = 0;
  for (;;) {
    unsigned char rec[40];
    const size_t got = std::fread(rec, 1, sizeof rec, f);
    if (got == 0) break;
    if (got != sizeof rec) { std::fclose(f); return set_error(QD_E_ARG, "%s: truncated record %lld ('failfile')", path, (long long)n + 1); }
    int32_t q[8];
    double v;
    std::memcpy(q, rec, 32);
    std::memcpy(&v, rec + 32, 8);
    n++;
    bool skip = false;
    for (int k = 0; k < 4; k++)
      if (q[2 * k] < 0 || 2 * q[2 * k] + std::abs(q[2 * k + 1]) > nshell - 1) skip = true;
    if (skip) continue;
    const int a = nm_orbital(q[0], q[1]), b = nm_orbital(q[2], q[3]), c = nm_orbital(q[4], q[5]),
              d = nm_orbital(q[6], q[7]);
    const int64_t s1 = slot(a, b, c, d), s2 = slot(c, d, a, b);
    if (s1 < 0) continue;
    S.tab[s1] = v; S.present[s1] = 1;
    S.tab[s2] = v; S.present[s2] = 1;
  }
  std::fclose(f);
  if (nrec) *nrec = n;
  return QD_OK;
}
}  // namespace
}  // namespace qd

extern "C" {

int qd_hf_hamiltonian(int nbody, int nshell, double hw, double *h0, int64_t neq, double *eig_out, int *iters) {
  if (nshell < 1 || nshell > 40) return set_error(QD_E_LAYOUT, "nshell out of range");
  const bool trace = std::getenv("QD_HF_TRACE") != nullptr;
  auto t0 = std::chrono::steady_clock::now();
  const Spatial &S = spatial_table(nshell);
  if (trace)
    fprintf(stderr, "hf: Coulomb table (R=%d) %.2f s\n", nshell,
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  return hf_build(nbody, nshell, hw, S, h0, neq, eig_out, iters);
}

int qd_hf_hamiltonian_device(qd_context *ctx, double hw, qd_op *h0, double *eig_out, int *iters) {
  if (!ctx || !h0) return set_error(QD_E_ARG, "null argument");
  if (op_context(h0) != ctx) return set_error(QD_E_ARG, "operator belongs to another context");
  if (op_herm(h0) != 1) return set_error(QD_E_ARG, "H(0) needs a hermitian operator");
  const Channels &ch = context_channels(ctx);
  const bool trace = std::getenv("QD_HF_TRACE") != nullptr;
  auto t0 = std::chrono::steady_clock::now();
  const Spatial &S = spatial_table(ch.R);
  if (trace)
    fprintf(stderr, "hf: Coulomb table (R=%d) %.2f s\n", ch.R,
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  return hf_build(ch.n, ch.R, hw, S, nullptr, ch.neq, eig_out, iters, ctx, h0);
}

int qd_hf_hamiltonian_me(int nbody, int nshell, double hw, const char *me_path, double *h0, int64_t neq,
                         double *eig_out, int *iters) {
  if (!me_path) return set_error(QD_E_ARG, "null path");
  if (nshell < 1 || nshell > 40) return set_error(QD_E_LAYOUT, "nshell out of range");
  Spatial S;
  int rc = load_me_bin(me_path, nshell, S, nullptr);
  if (rc) return rc;
  return hf_build(nbody, nshell, hw, S, h0, neq, eig_out, iters);
}

int qd_me_bin_write(int nshell, const char *me_path, int64_t *nrec) {
  if (!me_path) return set_error(QD_E_ARG, "null path");
  if (nshell < 1 || nshell > 40) return set_error(QD_E_LAYOUT, "nshell out of range");
  const Spatial &S = spatial_table(nshell);
  FILE *f = std::fopen(me_path, "wb");
  if (!f) return set_error(QD_E_ARG, "cannot create %s", me_path);
  // one record per (a, b, c, d) with m_a + m_b = m_c + m_d and (a, b) <= (c, d)
  // (the loader stores each record at <ab|cd> and <cd|ab>)
  int64_t n = 0;
  const int N = S.N;
  std::vector<unsigned char> buf;
  buf.reserve(40 * 4096);
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++)
      for (int c = 0; c < N; c++) {
        const int md = S.m[a] + S.m[b] - S.m[c];
        if (md < -S.mmax || md > S.mmax) continue;
        for (int d : S.by_m[md + S.mmax]) {
          if (std::make_pair(a, b) > std::make_pair(c, d)) continue;
          const int32_t q[8] = {S.n[a], S.m[a], S.n[b], S.m[b], S.n[c], S.m[c], S.n[d], S.m[d]};
          const double v = S.get(a, b, c, d);
          unsigned char rec[40];
          std::memcpy(rec, q, 32);
          std::memcpy(rec + 32, &v, 8);
          buf.insert(buf.end(), rec, rec + 40);
          n++;
          if (buf.size() >= 40 * 4096) {
            if (std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) { std::fclose(f); return set_error(QD_E_ARG, "write failed: %s", me_path); }
            buf.clear();
          }
        }
      }
  if (!buf.empty() && std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
    std::fclose(f);
    return set_error(QD_E_ARG, "write failed: %s", me_path);
  }
  if (std::fclose(f) != 0) return set_error(QD_E_ARG, "write failed: %s", me_path);
  if (nrec) *nrec = n;
  return QD_OK;
}

}  // extern "C"
