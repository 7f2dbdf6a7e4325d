// Consumers of the decoupled H(s): the TDA spectrum of the per-step side
// output (src/operators.f90:7-124) and the EOM-IM-SRG excited states
// (src/EOM_IMSRG.f90:10-47, 132-280, 551-1080; src/commutators.f90:1162-1260).
//
// TDA: the matrix <ai|H|bj> - E0 = d_ij f_ab - d_ab f_ji + <aj|V|ib> over the
// (particle a < cut, hole i) pairs of one (M_L, M_S) transfer is gathered on
// the device, one thread per element, and diagonalised on the host.
//
// EOM: the ladder operator Q of one (M_L, M_S) transfer has a ph part Q_ai
// and a pphh part Q_abij; the EOM matrix applied to Q is the ph and pphh part
// of the IM-SRG(2) commutator [H, Q] (HQ_comm_1b / HQ_comm_2b).  On the device
// Q is scattered into dense arrays (ph: (M-N) x N column-major like fph; pphh:
// fermionic pair index of a < b times that of i < j), then one wave per ph
// output and one thread per pphh output form the commutator.  Every sum the
// reference takes over all states is kept where f enters (f need not conserve
// the channel); where V enters, the sum is restricted to the states or pairs
// that V's (M_L, M_S) blocks can connect -- the others are exact zeros of
// v_elem (src/ME_general.f90:1610-1612).  The eigenvalues come from a Lanczos
// iteration with full reorthogonalisation on device vectors in place of
// ARPACK's implicitly restarted Lanczos (dsaupd 'SA', src/EOM_IMSRG.f90:232-
// 277): the smallest algebraic eigenvalues, converged to a relative residual
// of 1e-10.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "qd_kernels.h"

namespace qd {
namespace {

#define HIPX(expr)                                                                            \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return set_error(QD_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---- TDA ---------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_tda_matrix(DevTab t, const double *__restrict__ v, int herm,
                                                    const int2 *__restrict__ lab, int d, double *__restrict__ m) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)d * d) return;
  const int I = (int)(e % d), J = (int)(e / d);
  const int i = lab[I].x, a = lab[I].y, j = lab[J].x, b = lab[J].y;
  double x = velem(t, v, herm, a, j, i, b);
  if (i == j) x += felem(t, v, herm, a, b);
  if (a == b) x -= felem(t, v, herm, j, i);
  m[e] = x;
}

// build_TDAmat (src/operators.f90:36-52, 67-90): (hole i, particle a) with
// a below the cut (1-based: a = N+1 .. cutshell) and m_l(a) - m_l(i) = ML,
// m_s(a) - m_s(i) = MS, hole outer
std::vector<int2> tda_labels(const Channels &ch, int ml, int ms, int cut) {
  std::vector<int2> lab;
  for (int i = 0; i < ch.n; i++)
    for (int a = ch.n; a < std::min(cut, ch.M); a++)
      if (ch.ml[a] - ch.ml[i] == ml && ch.ms[a] - ch.ms[i] == ms) lab.push_back(make_int2(i, a));
  return lab;
}

// ---- EOM ---------------------------------------------------------------------
struct EomDev {  // device view of one EOM space
  int n, mp, np2, nh2;            // holes, particles, particle pairs, hole pairs
  int sps, tps;                   // ph and pphh components
  const int2 *ph;                 // (a, i) per ph component
  const int4 *pp;                 // (a, b, i, j) per pphh component, a < b, i < j
  double *Qph, *Qpp;              // dense Q: (mp x n) and (np2 x nh2), column-major
};

__device__ __forceinline__ int fidx(int x, int y, int cnt) {  // x < y, 0-based
  return x * cnt - x * (x + 1) / 2 + (y - x - 1);
}
// Qf_elem (src/ME_general.f90:1838-1868): the ph part, zero elsewhere
__device__ __forceinline__ double qf(const EomDev &q, int a, int i) {
  return (a >= q.n && i < q.n) ? q.Qph[(a - q.n) + (int64_t)i * q.mp] : 0.0;
}
// Q_elem (src/ME_general.f90:1170-1225): the pphh part with pair signs
__device__ __forceinline__ double qe(const EomDev &q, int a, int b, int c, int d) {
  if (a == b || c == d || a < q.n || b < q.n || c >= q.n || d >= q.n) return 0.0;
  double pre = 1.0;
  if (a > b) { int x = a; a = b; b = x; pre = -pre; }
  if (c > d) { int x = c; c = d; d = x; pre = -pre; }
  return pre * q.Qpp[fidx(a - q.n, b - q.n, q.mp) + (int64_t)fidx(c, d, q.n) * q.np2];
}
// block of the pair (x, y), x != y
__device__ __forceinline__ int pair_block(const DevTab &t, int x, int y) {
  const int lo = x < y ? x : y, hi = x < y ? y : x;
  return ia_blk(t.iarr[pidx(lo, hi, t.M)]);
}
// single-particle channel with (m_l, m_s), or -1
__device__ __forceinline__ int chan_of(const DevTab &t, int ml, int ms) {
  if (ml < -t.ml0 || ml > t.ml0 || (ms != 1 && ms != -1)) return -1;
  return t.ch_lookup[(ml + t.ml0) * 2 + (ms > 0)];
}

__global__ void __launch_bounds__(256) k_eom_scatter(EomDev q, const double *__restrict__ x) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < q.sps) {
    const int2 p = q.ph[e];
    q.Qph[(p.x - q.n) + (int64_t)p.y * q.mp] = x[e];
  } else if (e < (int64_t)q.sps + q.tps) {
    const int4 p = q.pp[e - q.sps];
    q.Qpp[fidx(p.x - q.n, p.y - q.n, q.mp) + (int64_t)fidx(p.z, p.w, q.n) * q.np2] = x[e];
  }
}
__global__ void __launch_bounds__(256) k_fill0(double *p, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) p[e] = 0.0;
}

// HQ_comm_1b(a, i, H, Q) (src/commutators.f90:1162-1208): one wave per output
__global__ void __launch_bounds__(256) k_eom_1b(DevTab t, const double *__restrict__ hv, int hh, EomDev q,
                                                double *__restrict__ y) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= q.sps) return;
  const int a = q.ph[w].x, i = q.ph[w].y, n = q.n, M = t.M;
  double s = 0.0;
  // 11 -> 1
  for (int b = n + lane; b < M; b += 64) s += felem(t, hv, hh, a, b) * qf(q, b, i);
  for (int j = lane; j < n; j += 64) s -= qf(q, a, j) * felem(t, hv, hh, j, i);
  // 12 -> 1 and 21 -> 1
  for (int e = lane; e < q.mp * n; e += 64) {
    const int b = n + e % q.mp, j = e / q.mp;
    s += felem(t, hv, hh, j, b) * qe(q, a, b, i, j) + qf(q, b, j) * velem(t, hv, hh, a, j, i, b);
  }
  // 22 -> 1: sum_{b<c} sum_j Q(c,b,i,j) V(a,j,c,b), (b, c) the pp pairs of (a, j)'s block
  for (int j = 0; j < n; j++) {
    const int blk = pair_block(t, a, j), np = t.np[blk], s0 = t.pp_start[blk];
    for (int k = lane; k < np; k += 64) {
      const uint32_t pr = t.pairs[s0 + k];
      const int b = lo16(pr), c = hi16(pr);
      s += qe(q, c, b, i, j) * velem(t, hv, hh, a, j, c, b);
    }
  }
  // - sum_b sum_{k<j} Q(a,b,k,j) V(k,j,i,b), (k, j) the hh pairs of (i, b)'s block
  for (int b = n; b < M; b++) {
    const int blk = pair_block(t, i, b), nh = t.nh[blk], s0 = t.hh_start[blk];
    for (int k = lane; k < nh; k += 64) {
      const uint32_t pr = t.pairs[s0 + k];
      const int kk = lo16(pr), jj = hi16(pr);
      s -= qe(q, a, b, kk, jj) * velem(t, hv, hh, kk, jj, i, b);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) y[w] = s;
}

// HQ_comm_2b(a, b, i, j, H, Q) (src/commutators.f90:1211-1260): one thread per output
__global__ void __launch_bounds__(256) k_eom_2b(DevTab t, const double *__restrict__ hv, int hh, EomDev q,
                                                double *__restrict__ y) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= q.tps) return;
  const int4 p = q.pp[e];
  const int a = p.x, b = p.y, i = p.z, j = p.w, n = q.n, M = t.M;
  double s = 0.0;
  // 12 -> 2
  for (int c = n; c < M; c++)
    s += felem(t, hv, hh, a, c) * qe(q, c, b, i, j) - felem(t, hv, hh, b, c) * qe(q, c, a, i, j);
  {  // Qf(c,i) V(a,b,c,j) - Qf(c,j) V(a,b,c,i): c in the channel that conserves (M_L, M_S)
    const int chj = chan_of(t, t.ml[a] + t.ml[b] - t.ml[j], t.ms[a] + t.ms[b] - t.ms[j]);
    if (chj >= 0)
      for (int k = t.ch_start[chj]; k < t.ch_start[chj + 1]; k++) {
        const int c = t.ch_states[k];
        if (c >= n) s += qf(q, c, i) * velem(t, hv, hh, a, b, c, j);
      }
    const int chi = chan_of(t, t.ml[a] + t.ml[b] - t.ml[i], t.ms[a] + t.ms[b] - t.ms[i]);
    if (chi >= 0)
      for (int k = t.ch_start[chi]; k < t.ch_start[chi + 1]; k++) {
        const int c = t.ch_states[k];
        if (c >= n) s -= qf(q, c, j) * velem(t, hv, hh, a, b, c, i);
      }
  }
  for (int k = 0; k < n; k++) {
    s += -qe(q, a, b, k, j) * felem(t, hv, hh, k, i) + qe(q, a, b, k, i) * felem(t, hv, hh, k, j);
    s += -velem(t, hv, hh, b, k, j, i) * qf(q, a, k) + velem(t, hv, hh, a, k, j, i) * qf(q, b, k);
  }
  // 22 -> 2 ladders: the pp pairs of (a, b)'s block, the hh pairs of (i, j)'s block
  {
    const int blk = pair_block(t, a, b), np = t.np[blk], s0 = t.pp_start[blk];
    for (int k = 0; k < np; k++) {
      const uint32_t pr = t.pairs[s0 + k];
      s += velem(t, hv, hh, a, b, lo16(pr), hi16(pr)) * qe(q, lo16(pr), hi16(pr), i, j);
    }
  }
  {
    const int blk = pair_block(t, i, j), nh = t.nh[blk], s0 = t.hh_start[blk];
    for (int k = 0; k < nh; k++) {
      const uint32_t pr = t.pairs[s0 + k];
      s += velem(t, hv, hh, lo16(pr), hi16(pr), i, j) * qe(q, a, b, lo16(pr), hi16(pr));
    }
  }
  // 22 -> 2 ph terms: for V(x,k,c,z) the particle c lies in the channel
  // m(c) = m(x) + m(k) - m(z)
  for (int k = 0; k < n; k++) {
    const int xs[4] = {b, a, b, a}, zs[4] = {i, i, j, j}, us[4] = {a, b, a, b}, vs[4] = {j, j, i, i};
    const double sg[4] = {-1.0, 1.0, 1.0, -1.0};
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int x = xs[r], z = zs[r];
      const int ch = chan_of(t, t.ml[x] + t.ml[k] - t.ml[z], t.ms[x] + t.ms[k] - t.ms[z]);
      if (ch < 0) continue;
      for (int m = t.ch_start[ch]; m < t.ch_start[ch + 1]; m++) {
        const int c = t.ch_states[m];
        if (c < n) continue;
        s += sg[r] * velem(t, hv, hh, x, k, c, z) * qe(q, us[r], c, k, vs[r]);
      }
    }
  }
  y[q.sps + e] = s;
}

// c[j] = V(:, j) . w over the first k columns (one work-group per column)
__global__ void __launch_bounds__(256) k_coldots(const double *__restrict__ V, int64_t N, const double *__restrict__ w,
                                                 double *__restrict__ c) {
  const int j = blockIdx.x;
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < N; e += 256) s += V[e + (int64_t)j * N] * w[e];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) c[j] = red[0];
}
// w -= V(:, 0..k-1) c
__global__ void __launch_bounds__(256) k_subtract(const double *__restrict__ V, int64_t N, int k,
                                                  const double *__restrict__ c, double *__restrict__ w) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= N) return;
  double s = w[e];
  for (int j = 0; j < k; j++) s -= V[e + (int64_t)j * N] * c[j];
  w[e] = s;
}
__global__ void __launch_bounds__(256) k_scale_to(const double *__restrict__ w, int64_t N, double f,
                                                  double *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < N) out[e] = w[e] * f;
}

// The EOM space of one (M_L, M_S) transfer, in the reference's order
// (unwrap, src/EOM_IMSRG.f90:815-884): ph components with the particle
// outer, then pphh components (a < b particles, i < j holes), a outer
struct EomSpace {
  std::vector<int2> ph;
  std::vector<int4> pp;
};
EomSpace eom_space(const Channels &ch, int ml, int ms) {
  EomSpace s;
  const int n = ch.n, M = ch.M;
  for (int a = n; a < M; a++)
    for (int i = 0; i < n; i++)
      if (ch.ml[a] - ch.ml[i] == ml && ch.ms[a] - ch.ms[i] == ms) s.ph.push_back(make_int2(a, i));
  for (int a = n; a < M; a++)
    for (int b = a + 1; b < M; b++)
      for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
          if (ch.ml[a] + ch.ml[b] - ch.ml[i] - ch.ml[j] == ml && ch.ms[a] + ch.ms[b] - ch.ms[i] - ch.ms[j] == ms)
            s.pp.push_back(make_int4(a, b, i, j));
  return s;
}

// ---- particle-attached (mode 1) and hole-removed (mode 2) EOM ---------------
// LANCZOS_1p_ATTACHED / LANCZOS_1h_REMOVED with matvec_1p_attached /
// matvec_1h_removed (src/EOM_IMSRG.f90:283-548, 662-812) and HQ_comm_1p,
// HQ_comm_2p1h, HQ_comm_1h, HQ_comm_2h1p (src/commutators.f90:1263-1425).
// The ladder operator has one-index components X1 (mode 1: particles a with
// (m_l, m_s) = (ML, MS); mode 2: holes i with -(m_l, m_s) = (ML, MS)) and
// three-index components X3 (mode 1: a < b particles, hole i, m_a + m_b - m_i
// = M; mode 2: i < j holes, particle a, -m_i - m_j + m_a = M), first index
// outer.  On the device X1 is a dense vector over the states and X3 a dense
// array over (pair of the two like indices, third index); get_X3 is
// antisymmetric in its first two indices (src/ME_general.f90:1243-1280).
struct Eom13Dev {
  int mode, n, mp;
  int s1, s3;                    // X1 and X3 components
  const int *l1;                 // state of each X1 component
  const int4 *l3;                // (x, y, z, 0) of each X3 component
  double *X1, *X3;               // dense: X1[state], X3[pair * third + k]
};
__device__ __forceinline__ double gx1(const Eom13Dev &q, int a) { return q.X1[a]; }
__device__ __forceinline__ double gx3(const Eom13Dev &q, int x, int y, int z) {
  if (x == y) return 0.0;
  double pre = 1.0;
  if (x > y) { int w = x; x = y; y = w; pre = -pre; }
  if (q.mode == 1) {  // x, y particles, z hole
    if (x < q.n || z >= q.n) return 0.0;
    return pre * q.X3[(int64_t)fidx(x - q.n, y - q.n, q.mp) * q.n + z];
  }
  if (y >= q.n || z < q.n) return 0.0;  // x, y holes, z particle
  return pre * q.X3[(int64_t)fidx(x, y, q.n) * q.mp + (z - q.n)];
}
__global__ void __launch_bounds__(256) k_eom13_scatter(Eom13Dev q, const double *__restrict__ x) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < q.s1) {
    q.X1[q.l1[e]] = x[e];
  } else if (e < (int64_t)q.s1 + q.s3) {
    const int4 p = q.l3[e - q.s1];
    const int64_t k = q.mode == 1 ? (int64_t)fidx(p.x - q.n, p.y - q.n, q.mp) * q.n + p.z
                                  : (int64_t)fidx(p.x, p.y, q.n) * q.mp + (p.z - q.n);
    q.X3[k] = x[e];
  }
}
// one thread per output component
__global__ void __launch_bounds__(256) k_eom13(DevTab t, const double *__restrict__ hv, int hh, Eom13Dev q,
                                               double *__restrict__ y) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)q.s1 + q.s3) return;
  const int n = q.n, M = t.M;
  double s = 0.0;
  if (q.mode == 1 && e < q.s1) {  // HQ_comm_1p(a)
    const int a = q.l1[e];
    for (int b = n; b < M; b++) s += felem(t, hv, hh, a, b) * gx1(q, b);
    for (int b = n; b < M; b++)
      for (int i = 0; i < n; i++) s += felem(t, hv, hh, i, b) * gx3(q, a, b, i);
    for (int b = n; b < M; b++)
      for (int c = b + 1; c < M; c++)
        for (int i = 0; i < n; i++) s += velem(t, hv, hh, a, i, b, c) * gx3(q, b, c, i);
  } else if (q.mode == 1) {  // HQ_comm_2p1h(a, b, i)
    const int4 p = q.l3[e - q.s1];
    const int a = p.x, b = p.y, i = p.z;
    for (int c = n; c < M; c++)
      s += velem(t, hv, hh, a, b, c, i) * gx1(q, c) + felem(t, hv, hh, a, c) * gx3(q, c, b, i) -
           felem(t, hv, hh, b, c) * gx3(q, c, a, i);
    for (int j = 0; j < n; j++) s -= felem(t, hv, hh, j, i) * gx3(q, a, b, j);
    for (int c = n; c < M; c++)
      for (int d = c + 1; d < M; d++) s += velem(t, hv, hh, a, b, c, d) * gx3(q, c, d, i);
    for (int c = n; c < M; c++)
      for (int j = 0; j < n; j++)
        s += -velem(t, hv, hh, a, j, c, i) * gx3(q, c, b, j) + velem(t, hv, hh, b, j, c, i) * gx3(q, c, a, j);
  } else if (e < q.s1) {  // HQ_comm_1h(i)
    const int i = q.l1[e];
    for (int j = 0; j < n; j++) s -= felem(t, hv, hh, j, i) * gx1(q, j);
    for (int b = n; b < M; b++)
      for (int j = 0; j < n; j++) s += felem(t, hv, hh, j, b) * gx3(q, i, j, b);
    for (int j = 0; j < n; j++)
      for (int k = j + 1; k < n; k++)
        for (int b = n; b < M; b++) s += velem(t, hv, hh, j, k, b, i) * gx3(q, j, k, b);
  } else {  // HQ_comm_2h1p(i, j, a)
    const int4 p = q.l3[e - q.s1];
    const int i = p.x, j = p.y, a = p.z;
    for (int k = 0; k < n; k++)
      s += velem(t, hv, hh, a, k, i, j) * gx1(q, k) - felem(t, hv, hh, k, j) * gx3(q, i, k, a) +
           felem(t, hv, hh, k, i) * gx3(q, j, k, a);
    for (int b = n; b < M; b++) s += felem(t, hv, hh, a, b) * gx3(q, i, j, b);
    for (int k = 0; k < n; k++)
      for (int l = k + 1; l < n; l++) s += velem(t, hv, hh, k, l, i, j) * gx3(q, k, l, a);
    for (int b = n; b < M; b++)
      for (int k = 0; k < n; k++)
        s += -velem(t, hv, hh, a, k, b, i) * gx3(q, k, j, b) + velem(t, hv, hh, a, k, b, j) * gx3(q, k, i, b);
  }
  y[e] = s;
}

struct Eom13Space {
  std::vector<int> l1;
  std::vector<int4> l3;
};
Eom13Space eom13_space(const Channels &ch, int mode, int ml, int ms) {
  Eom13Space s;
  const int n = ch.n, M = ch.M;
  if (mode == 1) {
    for (int a = n; a < M; a++)
      if (ch.ml[a] == ml && ch.ms[a] == ms) s.l1.push_back(a);
    for (int a = n; a < M; a++)
      for (int b = a + 1; b < M; b++)
        for (int i = 0; i < n; i++)
          if (ch.ml[a] + ch.ml[b] - ch.ml[i] == ml && ch.ms[a] + ch.ms[b] - ch.ms[i] == ms)
            s.l3.push_back(make_int4(a, b, i, 0));
  } else {
    for (int i = 0; i < n; i++)
      if (-ch.ml[i] == ml && -ch.ms[i] == ms) s.l1.push_back(i);
    for (int a = 0; a < n; a++)
      for (int b = a + 1; b < n; b++)
        for (int i = n; i < M; i++)
          if (-ch.ml[a] - ch.ml[b] + ch.ml[i] == ml && -ch.ms[a] - ch.ms[b] + ch.ms[i] == ms)
            s.l3.push_back(make_int4(a, b, i, 0));
  }
  return s;
}

// device buffers of one EOM solve (freed by the destructor)
struct EomRun {
  int mode = 0;  // 0 excited states, 1 particle attached, 2 hole removed
  Eom13Dev q13{};
  EomDev q{};
  int2 *ph = nullptr;
  int4 *pp = nullptr;
  double *Qph = nullptr, *Qpp = nullptr, *x = nullptr, *y = nullptr;
  std::vector<void *> mem;
  ~EomRun() {
    for (void *p : mem) (void)hipFree(p);
  }
  int alloc(void **p, size_t bytes) {
    HIPX(hipMalloc(p, std::max<size_t>(bytes, 16)));
    mem.push_back(*p);
    return QD_OK;
  }
  int init(const Channels &ch, const EomSpace &s) {
    q.n = ch.n;
    q.mp = ch.mp;
    q.np2 = ch.mp * (ch.mp - 1) / 2;
    q.nh2 = ch.n * (ch.n - 1) / 2;
    q.sps = (int)s.ph.size();
    q.tps = (int)s.pp.size();
    int rc;
    if ((rc = alloc((void **)&ph, sizeof(int2) * s.ph.size())) || (rc = alloc((void **)&pp, sizeof(int4) * s.pp.size())) ||
        (rc = alloc((void **)&Qph, sizeof(double) * (size_t)q.mp * q.n)) ||
        (rc = alloc((void **)&Qpp, sizeof(double) * (size_t)q.np2 * std::max(q.nh2, 1))) ||
        (rc = alloc((void **)&x, sizeof(double) * (size_t)(q.sps + q.tps))) ||
        (rc = alloc((void **)&y, sizeof(double) * (size_t)(q.sps + q.tps))))
      return rc;
    if (!s.ph.empty()) HIPX(hipMemcpy(ph, s.ph.data(), sizeof(int2) * s.ph.size(), hipMemcpyHostToDevice));
    if (!s.pp.empty()) HIPX(hipMemcpy(pp, s.pp.data(), sizeof(int4) * s.pp.size(), hipMemcpyHostToDevice));
    q.ph = ph;
    q.pp = pp;
    q.Qph = Qph;
    q.Qpp = Qpp;
    return QD_OK;
  }
  int init13(const Channels &ch, int md, const Eom13Space &s) {
    mode = md;
    q13.mode = md;
    q13.n = ch.n;
    q13.mp = ch.mp;
    q13.s1 = (int)s.l1.size();
    q13.s3 = (int)s.l3.size();
    const int64_t x3 = md == 1 ? (int64_t)ch.mp * (ch.mp - 1) / 2 * ch.n : (int64_t)ch.n * (ch.n - 1) / 2 * ch.mp;
    int *l1 = nullptr;
    int4 *l3 = nullptr;
    int rc;
    if ((rc = alloc((void **)&l1, sizeof(int) * s.l1.size())) || (rc = alloc((void **)&l3, sizeof(int4) * s.l3.size())) ||
        (rc = alloc((void **)&Qph, sizeof(double) * (size_t)ch.M)) ||
        (rc = alloc((void **)&Qpp, sizeof(double) * (size_t)std::max<int64_t>(x3, 1))) ||
        (rc = alloc((void **)&x, sizeof(double) * (s.l1.size() + s.l3.size()))) ||
        (rc = alloc((void **)&y, sizeof(double) * (s.l1.size() + s.l3.size()))))
      return rc;
    if (!s.l1.empty()) HIPX(hipMemcpy(l1, s.l1.data(), sizeof(int) * s.l1.size(), hipMemcpyHostToDevice));
    if (!s.l3.empty()) HIPX(hipMemcpy(l3, s.l3.data(), sizeof(int4) * s.l3.size(), hipMemcpyHostToDevice));
    q13.l1 = l1;
    q13.l3 = l3;
    q13.X1 = Qph;
    q13.X3 = Qpp;
    x3size = std::max<int64_t>(x3, 1);
    return QD_OK;
  }
  int64_t x3size = 0;
  int64_t size() const { return mode == 0 ? (int64_t)q.sps + q.tps : (int64_t)q13.s1 + q13.s3; }
  int64_t first() const { return mode == 0 ? q.sps : q13.s1; }  // one-index (ph / X1) components
  // y = A x for device vectors (x may be this->x)
  int matvec(hipStream_t st, const DevTab &t, const double *hv, int hherm, const double *xin, double *yout) {
    if (mode != 0) {
      const int64_t N = size();
      hipLaunchKernelGGL(k_fill0, dim3(64), dim3(256), 0, st, q13.X1, (int64_t)t.M);
      hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, q13.X3, x3size);
      if (N > 0) {
        hipLaunchKernelGGL(k_eom13_scatter, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, q13, xin);
        hipLaunchKernelGGL(k_eom13, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, t, hv, hherm, q13, yout);
      }
      HIPX(hipGetLastError());
      return QD_OK;
    }
    const int64_t N = (int64_t)q.sps + q.tps;
    hipLaunchKernelGGL(k_fill0, dim3(256), dim3(256), 0, st, q.Qph, (int64_t)q.mp * q.n);
    hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, q.Qpp, (int64_t)q.np2 * q.nh2);
    if (N > 0) hipLaunchKernelGGL(k_eom_scatter, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, q, xin);
    if (q.sps > 0) hipLaunchKernelGGL(k_eom_1b, dim3((unsigned)((q.sps + 3) / 4)), dim3(256), 0, st, t, hv, hherm, q, yout);
    if (q.tps > 0) hipLaunchKernelGGL(k_eom_2b, dim3((unsigned)((q.tps + 255) / 256)), dim3(256), 0, st, t, hv, hherm, q, yout);
    HIPX(hipGetLastError());
    return QD_OK;
  }
};

}  // namespace
}  // namespace qd

using namespace qd;

extern "C" {

int qd_tda_spectrum(qd_context *ctx, const qd_op *h, int ml, int ms, int cut, double *eig, int cap, int *dim) {
  if (!ctx || !h || !dim) return set_error(QD_E_ARG, "null argument");
  const Channels &ch = context_channels(ctx);
  const std::vector<int2> lab = tda_labels(ch, ml, ms, cut);
  const int d = (int)lab.size();
  *dim = d;
  if (d == 0) return QD_OK;
  if (!eig || cap < d) return set_error(QD_E_ARG, "TDA block of dimension %d, eig holds %d", d, cap);
  hipStream_t st = context_stream(ctx);
  int2 *dl = nullptr;
  double *dm = nullptr;
  HIPX(hipMalloc(&dl, sizeof(int2) * d));
  if (hipMalloc(&dm, sizeof(double) * (size_t)d * d) != hipSuccess) {
    (void)hipFree(dl);
    return set_error(QD_E_HIP, "hipMalloc of the TDA matrix (%d x %d) failed", d, d);
  }
  std::vector<double> m((size_t)d * d);
  hipError_t e = hipMemcpyAsync(dl, lab.data(), sizeof(int2) * d, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_tda_matrix, dim3((unsigned)(((int64_t)d * d + 255) / 256)), dim3(256), 0, st,
                       context_devtab(ctx), op_cdata(h), op_herm(h), dl, d, dm);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(m.data(), dm, sizeof(double) * m.size(), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(dl);
  (void)hipFree(dm);
  if (e != hipSuccess) return set_error(QD_E_HIP, "TDA matrix: %s", hipGetErrorString(e));
  // dsyev('V','U', ...) reads the upper triangle (src/ME_general.f90:576-590)
  std::vector<double> a((size_t)d * d), w, v;
  for (int r = 0; r < d; r++)
    for (int c = 0; c < d; c++) {
      const int lo = std::min(r, c), hi = std::max(r, c);
      a[(size_t)r * d + c] = m[lo + (size_t)hi * d];  // matrix(lo, hi), column-major
    }
  jacobi(d, a, w, v);
  std::memcpy(eig, w.data(), sizeof(double) * d);
  return QD_OK;
}

int qd_eom_dim(const qd_context *ctx, int ml, int ms, int64_t *n_ph, int64_t *n_pphh) {
  if (!ctx) return set_error(QD_E_ARG, "null context");
  const EomSpace s = eom_space(context_channels(ctx), ml, ms);
  if (n_ph) *n_ph = (int64_t)s.ph.size();
  if (n_pphh) *n_pphh = (int64_t)s.pp.size();
  return QD_OK;
}

int qd_eom_matvec(qd_context *ctx, const qd_op *h, int ml, int ms, const double *x, double *y, int64_t n) {
  if (!ctx || !h || !x || !y) return set_error(QD_E_ARG, "null argument");
  const Channels &ch = context_channels(ctx);
  const EomSpace s = eom_space(ch, ml, ms);
  const int64_t N = (int64_t)s.ph.size() + (int64_t)s.pp.size();
  if (n != N) return set_error(QD_E_ARG, "EOM space has %lld components, got %lld", (long long)N, (long long)n);
  EomRun r;
  int rc;
  if ((rc = r.init(ch, s))) return rc;
  hipStream_t st = context_stream(ctx);
  HIPX(hipMemcpyAsync(r.x, x, sizeof(double) * N, hipMemcpyHostToDevice, st));
  if ((rc = r.matvec(st, context_devtab(ctx), op_cdata(h), op_herm(h), r.x, r.y))) return rc;
  HIPX(hipMemcpyAsync(y, r.y, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  HIPX(hipStreamSynchronize(st));
  return QD_OK;
}

}  // extern "C"

namespace {
// Lanczos for the nev smallest eigenvalues of the EOM matrix of mode 0/1/2
int eom_solve(qd_context *ctx, const qd_op *h, int mode, int ml, int ms, int nev, double *eig, double *w1p1h,
              int *iters) {
  if (!ctx || !h || !eig || nev < 1) return set_error(QD_E_ARG, "null argument or nev < 1");
  const Channels &ch = context_channels(ctx);
  EomRun r;
  int rc;
  if (mode == 0) rc = r.init(ch, eom_space(ch, ml, ms));
  else rc = r.init13(ch, mode, eom13_space(ch, mode, ml, ms));
  if (rc) return rc;
  const int64_t N = r.size();
  if (N < nev) return set_error(QD_E_ARG, "EOM space of dimension %lld below nev = %d", (long long)N, nev);
  hipStream_t st = context_stream(ctx);
  const DevTab &t = context_devtab(ctx);
  const int kmax = (int)std::min<int64_t>(N, std::max(20 * nev, 200));
  double *V = nullptr, *w = nullptr, *c = nullptr;
  if ((rc = r.alloc((void **)&V, sizeof(double) * (size_t)N * (kmax + 1))) ||
      (rc = r.alloc((void **)&w, sizeof(double) * (size_t)N)) || (rc = r.alloc((void **)&c, sizeof(double) * (kmax + 1))))
    return rc;
  const unsigned gN = (unsigned)((N + 255) / 256);
  // start vector: seeded pseudo-random, normalised
  {
    std::vector<double> v0((size_t)N);
    std::mt19937_64 gen(20261203);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    double nr = 0.0;
    for (auto &x : v0) { x = u(gen); nr += x * x; }
    nr = std::sqrt(nr);
    for (auto &x : v0) x /= nr;
    HIPX(hipMemcpyAsync(V, v0.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
  }
  std::vector<double> alpha, beta, cv(kmax + 1), ritz, wts;
  std::vector<double> evec;
  int k = 0;
  bool done = false;
  while (!done && k < kmax) {
    double *vk = V + (size_t)k * N;
    if ((rc = r.matvec(st, t, op_cdata(h), op_herm(h), vk, w))) return rc;
    // full reorthogonalisation against v_0..v_k, twice (classical Gram-Schmidt)
    for (int pass = 0; pass < 2; pass++) {
      hipLaunchKernelGGL(k_coldots, dim3(k + 1), dim3(256), 0, st, V, N, w, c);
      if (pass == 0) {
        HIPX(hipMemcpyAsync(cv.data(), c, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        alpha.push_back(cv[k]);
      }
      hipLaunchKernelGGL(k_subtract, dim3(gN), dim3(256), 0, st, V, N, k + 1, c, w);
    }
    hipLaunchKernelGGL(k_coldots, dim3(1), dim3(256), 0, st, w, N, w, c);
    double b2 = 0.0;
    HIPX(hipMemcpyAsync(&b2, c, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    const double b = std::sqrt(std::max(b2, 0.0));
    k++;
    // Ritz values of the k x k tridiagonal matrix
    const bool check = k >= nev && (k % 5 == 0 || k == kmax || k == N || b < 1e-13);
    if (check) {
      std::vector<double> T((size_t)k * k, 0.0);
      for (int i = 0; i < k; i++) {
        T[(size_t)i * k + i] = alpha[i];
        if (i + 1 < k) T[(size_t)i * k + i + 1] = T[(size_t)(i + 1) * k + i] = beta[i];
      }
      std::vector<double> th, S;
      jacobi(k, T, th, S);
      bool conv = true;
      for (int i = 0; i < nev; i++)
        if (std::fabs(b * S[(size_t)(k - 1) * k + i]) > 1e-10 * std::max(1.0, std::fabs(th[i]))) conv = false;
      if (conv || k == N || k == kmax || b < 1e-13) {
        done = true;
        ritz.assign(th.begin(), th.begin() + nev);
        evec = S;
      }
    }
    if (!done) {
      beta.push_back(b);
      hipLaunchKernelGGL(k_scale_to, dim3(gN), dim3(256), 0, st, w, N, 1.0 / b, V + (size_t)k * N);
    }
  }
  if (!done) return set_error(QD_E_NUMERIC, "EOM Lanczos did not converge in %d iterations", kmax);
  for (int i = 0; i < nev; i++) eig[i] = ritz[i];
  if (w1p1h) {  // ph weight of each Ritz vector: sum over its ph components squared
    const int sps = (int)r.first();
    std::vector<double> Vph((size_t)sps * k);
    for (int j = 0; j < k && sps > 0; j++)
      HIPX(hipMemcpyAsync(Vph.data() + (size_t)j * sps, V + (size_t)j * N, sizeof(double) * sps, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    for (int i = 0; i < nev; i++) {
      double sw = 0.0;
      for (int mu = 0; mu < sps; mu++) {
        double yv = 0.0;
        for (int j = 0; j < k; j++) yv += Vph[(size_t)j * sps + mu] * evec[(size_t)j * k + i];
        sw += yv * yv;
      }
      w1p1h[i] = sw;
    }
  }
  if (iters) *iters = k;
  return QD_OK;
}
}  // namespace

extern "C" {

int qd_eom_excited_states(qd_context *ctx, const qd_op *h, int ml, int ms, int nev, double *eig, double *w1p1h,
                          int *iters) {
  return eom_solve(ctx, h, 0, ml, ms, nev, eig, w1p1h, iters);
}

int qd_eom_attached_removed(qd_context *ctx, const qd_op *h, int mode, int ml, int ms, int nev, double *eig,
                            double *w1, int *iters) {
  if (mode != 1 && mode != 2) return set_error(QD_E_ARG, "mode must be 1 (particle attached) or 2 (hole removed)");
  return eom_solve(ctx, h, mode, ml, ms, nev, eig, w1, iters);
}

int qd_eom13_dim(const qd_context *ctx, int mode, int ml, int ms, int64_t *n1, int64_t *n3) {
  if (!ctx || (mode != 1 && mode != 2)) return set_error(QD_E_ARG, "null context or bad mode");
  const Eom13Space s = eom13_space(context_channels(ctx), mode, ml, ms);
  if (n1) *n1 = (int64_t)s.l1.size();
  if (n3) *n3 = (int64_t)s.l3.size();
  return QD_OK;
}

int qd_eom13_matvec(qd_context *ctx, const qd_op *h, int mode, int ml, int ms, const double *x, double *y, int64_t n) {
  if (!ctx || !h || !x || !y || (mode != 1 && mode != 2)) return set_error(QD_E_ARG, "null argument or bad mode");
  const Channels &ch = context_channels(ctx);
  EomRun r;
  int rc;
  if ((rc = r.init13(ch, mode, eom13_space(ch, mode, ml, ms)))) return rc;
  const int64_t N = r.size();
  if (n != N) return set_error(QD_E_ARG, "EOM space has %lld components, got %lld", (long long)N, (long long)n);
  hipStream_t st = context_stream(ctx);
  HIPX(hipMemcpyAsync(r.x, x, sizeof(double) * N, hipMemcpyHostToDevice, st));
  if ((rc = r.matvec(st, context_devtab(ctx), op_cdata(h), op_herm(h), r.x, r.y))) return rc;
  HIPX(hipMemcpyAsync(y, r.y, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  HIPX(hipStreamSynchronize(st));
  return QD_OK;
}

}  // extern "C"
