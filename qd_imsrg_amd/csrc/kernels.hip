// Gather / reduction kernels of the IM-SRG(2) commutator (gfx950, FP64).
//
// Every kernel reads operators in the reference's flat vectorize layout.  The
// element accessor velem() restates v_elem (src/ME_general.f90:1557-1680):
// pair sorting with sign, the i_array lookup, zero across blocks, and the
// herm factor for the three sub-blocks that are stored transposed.
#include "qd_kernels.h"

namespace qd {

__device__ __forceinline__ int64_t pidx(int i, int j, int M) {
  return (int64_t)i * M - (int64_t)i * (i + 1) / 2 + (j - i - 1);
}

__device__ __forceinline__ double velem(const DevTab &t, const double *__restrict__ v, int herm, int a1,
                                        int a2, int a3, int a4) {
  if (a1 == a2 || a3 == a4) return 0.0;
  double pre = 1.0;
  if (a1 > a2) { int x = a1; a1 = a2; a2 = x; pre = -pre; }
  if (a3 > a4) { int x = a3; a3 = a4; a4 = x; pre = -pre; }
  const uint32_t I = t.iarr[pidx(a1, a2, t.M)], J = t.iarr[pidx(a3, a4, t.M)];
  const int q = ia_blk(I);
  if (q != ia_blk(J)) return 0.0;
  const int w1 = ia_idx(I), w2 = ia_idx(J);
  const int64_t *o = t.off + 6 * q;
  const int nh = t.nh[q], np = t.np[q], nb = t.nb[q];
  switch (ia_type(I) * 4 + ia_type(J)) {
    case T_HH * 4 + T_HH: return pre * v[o[HHHH] + w1 + (int64_t)w2 * nh];
    case T_HH * 4 + T_PH: return pre * herm * v[o[PHHH] + w2 + (int64_t)w1 * nb];
    case T_HH * 4 + T_PP: return pre * herm * v[o[PPHH] + w2 + (int64_t)w1 * np];
    case T_PH * 4 + T_HH: return pre * v[o[PHHH] + w1 + (int64_t)w2 * nb];
    case T_PH * 4 + T_PH: return pre * v[o[PHPH] + w1 + (int64_t)w2 * nb];
    case T_PH * 4 + T_PP: return pre * herm * v[o[PPPH] + w2 + (int64_t)w1 * np];
    case T_PP * 4 + T_HH: return pre * v[o[PPHH] + w1 + (int64_t)w2 * np];
    case T_PP * 4 + T_PH: return pre * v[o[PPPH] + w1 + (int64_t)w2 * np];
    default: return pre * v[o[PPPP] + w1 + (int64_t)w2 * np];
  }
}

// f_elem (src/ME_general.f90:1780-1813)
__device__ __forceinline__ double felem(const DevTab &t, const double *__restrict__ v, int herm, int a, int b) {
  const int n = t.n;
  if (a < n) return (b < n) ? v[t.off_fhh + a + (int64_t)b * n] : herm * v[t.off_fph + (b - n) + (int64_t)a * t.mp];
  return (b < n) ? v[t.off_fph + (a - n) + (int64_t)b * t.mp] : v[t.off_fpp + (a - n) + (int64_t)(b - n) * t.mp];
}

// ---------------------------------------------------------------------------
// commutator_122 (src/commutators.f90:355-676) for one output element with
// indices in storage order.  For index position k the reference sums over
// holes t with f(x_k,t) and over particles t with f(t,x_k) (times herm of
// the one-body operand on the bra side / -1 and -herm on the ket side); only
// t in the (m_l, m_s) channel of x_k can give a non-zero two-body element.
__device__ double c122_half(const DevTab &t, const double *__restrict__ fa, int ha,
                            const double *__restrict__ vb, int hb, const int x[4]) {
  double r = 0.0;
  for (int k = 0; k < 4; k++) {
    const int xk = x[k];
    const int c = t.spch[xk];
    const int s0 = t.ch_start[c], s1 = t.ch_start[c + 1];
    for (int s = s0; s < s1; s++) {
      const int tt = t.ch_states[s];
      double coef;
      if (k < 2) coef = (tt < t.n) ? felem(t, fa, ha, xk, tt) : ha * felem(t, fa, ha, tt, xk);
      else coef = (tt < t.n) ? -ha * felem(t, fa, ha, xk, tt) : -felem(t, fa, ha, tt, xk);
      int y0 = x[0], y1 = x[1], y2 = x[2], y3 = x[3];
      if (k == 0) y0 = tt; else if (k == 1) y1 = tt; else if (k == 2) y2 = tt; else y3 = tt;
      r += coef * velem(t, vb, hb, y0, y1, y2, y3);
    }
  }
  return r;
}

// Pandya difference D[(a,b),(c,d)] = hh * P[(a,b),(c,d)] - eh * P[(c,d),(a,b)]
// with P = ETACC * HCC^T per CC block: XCC - YCC of construct_cc_ints
// (src/ME_general.f90:948-949), since YCC = eh * HCC ETACC^T = eh * P^T.
__device__ __forceinline__ double pandya_d(const DevTab &t, const double *__restrict__ P, double hh,
                                           double eh, int a, int b, int c, int d) {
  const int2 m1 = t.cc_map[a * t.M + b], m2 = t.cc_map[c * t.M + d];
  const int q = m1.x;
  if (t.cc_nb[q] == 0) return 0.0;
  const int64_t r = t.cc_r[q];
  const double *B = P + t.cc_poff[q];
  return hh * B[m1.y + m2.y * r] - eh * B[m2.y + m1.y * r];
}

__global__ void __launch_bounds__(256) k_twobody(DevTab t, Ops o, double *__restrict__ out,
                                                 const Seg2 *__restrict__ segs,
                                                 const Tile2 *__restrict__ tiles, int flags,
                                                 const double *__restrict__ w1, const double *__restrict__ w2,
                                                 const double *__restrict__ w3, const double *__restrict__ P,
                                                 double hh, double eh) {
  const Tile2 tl = tiles[blockIdx.x];
  const Seg2 sg = segs[tl.seg];
  const int e = tl.start + threadIdx.x;
  if (e >= sg.size) return;
  const int I = e % sg.rows, J = e / sg.rows;
  const uint32_t bra = t.pairs[sg.bra0 + I], ket = t.pairs[sg.ket0 + J];
  const int x[4] = {lo16(bra), hi16(bra), lo16(ket), hi16(ket)};
  const int64_t idx = sg.off + e;
  double acc = (flags & TB_ACCUM) ? out[idx] : 0.0;
  if (flags & TB_122)
    acc += c122_half(t, o.v1, o.h1, o.v2, o.h2, x) - c122_half(t, o.v2, o.h2, o.v1, o.h1, x);
  if (flags & TB_LADDER) {
    // transposed / signed combinations of the products of matrix_mult222
    // (src/commutators.f90:999-1127) that cannot be accumulated by the GEMM
    const double h1 = o.h1, h2 = o.h2;
    const int64_t tr = sg.off + J + (int64_t)I * sg.rows;  // (J,I) of a square sub-block
    switch (sg.sub) {
      case HHHH:
      case PPPP:
        acc += h1 * w1[tr] - h2 * w1[idx];
        acc += h1 * h2 * w3[tr] - w3[idx];
        break;
      case PHHH: acc += h1 * w2[idx] - h2 * w1[idx]; break;
      case PPPH: acc += -h2 * w1[idx] + h1 * w2[idx]; break;
      case PHPH:
        acc += h1 * w2[tr] - h2 * w2[idx];
        acc += h1 * w1[tr] - h2 * w1[idx];
        break;
      default: break;
    }
  }
  if (flags & TB_PANDYA) {
    // src/commutators.f90:746-751; storage-order indices carry the ph sign
    const int x1 = x[0], x2 = x[1], x3 = x[2], x4 = x[3];
    acc += pandya_d(t, P, hh, eh, x3, x1, x2, x4) - pandya_d(t, P, hh, eh, x3, x2, x1, x4) +
           pandya_d(t, P, hh, eh, x4, x2, x1, x3) - pandya_d(t, P, hh, eh, x4, x1, x2, x3);
  }
  out[idx] = acc;
}

void launch_twobody(hipStream_t s, const DevTab &t, const Ops &o, double *out, const Seg2 *segs,
                    const Tile2 *tiles, int ntiles, int flags, const double *w1, const double *w2,
                    const double *w3, const double *P, double hh, double eh) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_twobody, dim3(ntiles), dim3(256), 0, s, t, o, out, segs, tiles, flags, w1, w2, w3, P,
                     hh, eh);
}

// ---------------------------------------------------------------------------
// One-body part: commutator_111 (src/commutators.f90:137-195), _121 (:73-134)
// and the partial traces of _221 (:278-348).  One wave per output element of
// fhh, fpp, fph (contiguous in the flat layout from offset 1); the lanes split
// every summation and a fixed xor-shuffle tree reduces them (deterministic).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ double c121_lane(const DevTab &t, const Ops &o, int x, int y, int lane) {
  const int n = t.n;
  const int dl = t.ml[y] - t.ml[x], ds = t.ms[y] - t.ms[x];
  double r = 0.0;
  for (int i = lane; i < n; i += 64) {
    // particles a with (ml_a - ml_i, ms_a - ms_i) = +(dl,ds) or -(dl,ds)
    for (int sgn = 0; sgn < 2; sgn++) {
      if (sgn == 1 && dl == 0 && ds == 0) break;
      const int mla = t.ml[i] + (sgn ? -dl : dl), msa = t.ms[i] + (sgn ? -ds : ds);
      if (mla < -t.ml0 || mla > t.ml0 || (msa != 1 && msa != -1)) continue;
      const int c = t.ch_lookup[(mla + t.ml0) * 2 + (msa > 0)];
      if (c < 0) continue;
      for (int s = t.ch_start[c]; s < t.ch_start[c + 1]; s++) {
        const int a = t.ch_states[s];
        if (a < n) continue;
        const double f1 = o.v1[t.off_fph + (a - n) + (int64_t)i * t.mp];
        const double f2 = o.v2[t.off_fph + (a - n) + (int64_t)i * t.mp];
        r += f1 * (o.h1 * velem(t, o.v2, o.h2, a, x, i, y) - velem(t, o.v2, o.h2, i, x, a, y));
        r -= f2 * (o.h2 * velem(t, o.v1, o.h1, a, x, i, y) - velem(t, o.v1, o.h1, i, x, a, y));
      }
    }
  }
  return r;
}

__global__ void __launch_bounds__(256) k_onebody(DevTab t, Ops o, double *__restrict__ out, int flags,
                                                 const double *__restrict__ w1, const double *__restrict__ w2) {
  const int n = t.n, mp = t.mp;
  const int64_t nhh = (int64_t)n * n, npp = (int64_t)mp * mp, nph = (int64_t)mp * n;
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= nhh + npp + nph) return;  // whole wave exits together
  int kind, i, j, x, y;
  if (e < nhh) { kind = 0; i = (int)(e % n); j = (int)(e / n); x = i; y = j; }
  else if (e < nhh + npp) { kind = 1; int64_t f = e - nhh; i = (int)(f % mp); j = (int)(f / mp); x = n + i; y = n + j; }
  else { kind = 2; int64_t f = e - nhh - npp; i = (int)(f % mp); j = (int)(f / mp); x = n + i; y = j; }
  const double *A = o.v1, *B = o.v2;
  const double h1 = o.h1, h2 = o.h2;
  const int64_t ohh = t.off_fhh, opp = t.off_fpp, oph = t.off_fph;
#define FHH(v, a, b) (v)[ohh + (a) + (int64_t)(b) * n]
#define FPP(v, a, b) (v)[opp + (a) + (int64_t)(b) * mp]
#define FPH(v, a, b) (v)[oph + (a) + (int64_t)(b) * mp]
  const int64_t dst = 1 + e;
  double acc = (flags & OB_ACCUM) ? out[dst] : 0.0;
  if (flags & OB_111) {
    double q1 = 0, q1t = 0, q2 = 0, q2t = 0, q3 = 0, q4 = 0;
    if (kind == 0) {
      for (int k = lane; k < n; k += 64) { q1 += FHH(A, i, k) * FHH(B, k, j); q1t += FHH(A, j, k) * FHH(B, k, i); }
      for (int k = lane; k < mp; k += 64) { q2 += FPH(B, k, i) * FPH(A, k, j); q2t += FPH(B, k, j) * FPH(A, k, i); }
    } else if (kind == 1) {
      for (int k = lane; k < mp; k += 64) { q1 += FPP(B, i, k) * FPP(A, k, j); q1t += FPP(B, j, k) * FPP(A, k, i); }
      for (int k = lane; k < n; k += 64) { q2 += FPH(A, i, k) * FPH(B, j, k); q2t += FPH(A, j, k) * FPH(B, i, k); }
    } else {
      for (int k = lane; k < n; k += 64) { q1 += FPH(A, i, k) * FHH(B, k, j); q2 += FPH(B, i, k) * FHH(A, k, j); }
      for (int k = lane; k < mp; k += 64) { q3 += FPP(A, i, k) * FPH(B, k, j); q4 += FPP(B, i, k) * FPH(A, k, j); }
    }
    q1 = wave_sum(q1); q1t = wave_sum(q1t); q2 = wave_sum(q2); q2t = wave_sum(q2t);
    q3 = wave_sum(q3); q4 = wave_sum(q4);
    double val;
    if (kind == 0) val = q1 - h2 * q2 + h1 * (q2t - h2 * q1t);
    else if (kind == 1) val = h2 * q2 + h1 * (h2 * q1t - q2t) - q1;
    else val = q1 - q2 + q3 - q4;
    acc = (flags & OB_ACCUM) ? acc + val : val;
  }
#undef FHH
#undef FPP
#undef FPH
  if (flags & OB_121) acc += wave_sum(c121_lane(t, o, x, y, lane));
  if ((flags & OB_221) && t.spch[x] == t.spch[y]) {
    // partial traces of the 221 intermediates (w1, w2 have herm = +1)
    double r = 0.0;
    for (int a = lane; a < t.M; a += 64) {
      if (a < n) {
        if (kind == 2) r += h1 * velem(t, w2, 1, a, x, a, y) - h2 * velem(t, w1, 1, a, x, a, y);
        else r += h1 * velem(t, w1, 1, a, y, a, x) - h2 * velem(t, w1, 1, a, x, a, y);
      } else {
        if (kind == 1) r += -h1 * velem(t, w1, 1, a, y, a, x) + h2 * velem(t, w1, 1, a, x, a, y);
        else if (kind == 0) r += -h1 * velem(t, w2, 1, a, y, a, x) + h2 * velem(t, w2, 1, a, x, a, y);
        else r += -h1 * velem(t, w2, 1, a, x, a, y) + h2 * velem(t, w1, 1, a, x, a, y);
      }
    }
    acc += wave_sum(r);
  }
  if (lane == 0) out[dst] = acc;
}

void launch_onebody(hipStream_t s, const DevTab &t, const Ops &o, double *out, int flags, const double *w1,
                    const double *w2) {
  const int64_t tot = (int64_t)t.n * t.n + (int64_t)t.mp * t.mp + (int64_t)t.mp * t.n;
  hipLaunchKernelGGL(k_onebody, dim3((unsigned)((tot + 3) / 4)), dim3(256), 0, s, t, o, out, flags, w1, w2);
}

// ---------------------------------------------------------------------------
// deterministic block reduction helper (256 threads)
__device__ __forceinline__ double block_sum256(double v) {
  __shared__ double sh[256];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// commutator_110 + commutator_220 (src/commutators.f90:12-70): (h1 - h2) *
// (sum fph1 .* fph2 + sum over blocks Vpphh1 .* Vpphh2).  Work-group 0 takes
// the fph product, work-group g >= 1 the pphh segment g-1; k_zero_final adds
// the partials in a fixed order.
__global__ void __launch_bounds__(256) k_zerobody_part(DevTab t, Ops o, const Seg2 *__restrict__ segs, int which,
                                                       double *part) {
  double a = 0.0;
  if (blockIdx.x == 0) {
    if (which & 1) {
      const int64_t m = (int64_t)t.mp * t.n;
      for (int64_t e = threadIdx.x; e < m; e += 256) a += o.v1[t.off_fph + e] * o.v2[t.off_fph + e];
    }
  } else if (which & 2) {
    const Seg2 sg = segs[blockIdx.x - 1];
    for (int e = threadIdx.x; e < sg.size; e += 256) a += o.v1[sg.off + e] * o.v2[sg.off + e];
  }
  a = block_sum256(a);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}
__global__ void __launch_bounds__(256) k_zero_final(Ops o, const double *__restrict__ part, int np, double *dst) {
  double b = 0.0;
  for (int k = 1 + threadIdx.x; k < np; k += 256) b += part[k];
  b = block_sum256(b);
  if (threadIdx.x == 0) *dst = (double)(o.h1 - o.h2) * part[0] + (double)(o.h1 - o.h2) * b;
}

void launch_zerobody(hipStream_t s, const DevTab &t, const Ops &o, const Seg2 *pphh_segs, int npphh,
                     double *part, double *dst, int which) {
  hipLaunchKernelGGL(k_zerobody_part, dim3(1 + npphh), dim3(256), 0, s, t, o, pphh_segs, which, part);
  hipLaunchKernelGGL(k_zero_final, dim3(1), dim3(256), 0, s, o, part, 1 + npphh, dst);
}

// ---------------------------------------------------------------------------
// calc_cc (src/commutators.f90:1135-1159): HCC(r,c) = <j k|H|i l>,
// (i,j) = qnab(r), (k,l) = qnhp(c)
__global__ void __launch_bounds__(256) k_calc_cc(DevTab t, const double *__restrict__ v, int herm,
                                                 double *__restrict__ cc, const Seg2 *__restrict__ segs,
                                                 const Tile2 *__restrict__ tiles) {
  const Tile2 tl = tiles[blockIdx.x];
  const Seg2 sg = segs[tl.seg];
  const int e = tl.start + threadIdx.x;
  if (e >= sg.size) return;
  const int a = e % sg.rows, c = e / sg.rows;
  const uint32_t ab = t.cc_ab[sg.bra0 + a], hp = t.cc_hp[sg.ket0 + c];
  cc[sg.off + e] = velem(t, v, herm, hi16(ab), lo16(hp), lo16(ab), hi16(hp));
}

void launch_calc_cc(hipStream_t s, const DevTab &t, const double *v, int herm, double *cc, const Seg2 *cc_segs,
                    const Tile2 *cc_tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_calc_cc, dim3(ntiles), dim3(256), 0, s, t, v, herm, cc, cc_segs, cc_tiles);
}

// ---------------------------------------------------------------------------
// build_white (src/IMSRG_tools.f90:377-434)
__global__ void __launch_bounds__(256) k_white_ph(DevTab t, const double *__restrict__ rec, int hr,
                                                  double *__restrict__ gen) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)t.mp * t.n) return;
  const int ai = (int)(e % t.mp), i = (int)(e / t.mp), a = ai + t.n;
  const double den = rec[t.off_fpp + ai + (int64_t)ai * t.mp] - rec[t.off_fhh + i + (int64_t)i * t.n] -
                     velem(t, rec, hr, a, i, a, i);
  gen[t.off_fph + e] = rec[t.off_fph + e] / den;
}

__global__ void __launch_bounds__(256) k_white_pphh(DevTab t, const double *__restrict__ rec, int hr,
                                                    double *__restrict__ gen, const Seg2 *__restrict__ segs,
                                                    const Tile2 *__restrict__ tiles) {
  const Tile2 tl = tiles[blockIdx.x];
  const Seg2 sg = segs[tl.seg];
  const int e = tl.start + threadIdx.x;
  if (e >= sg.size) return;
  const int I = e % sg.rows, J = e / sg.rows;
  const uint32_t pp = t.pairs[sg.bra0 + I], hh = t.pairs[sg.ket0 + J];
  const int p1 = lo16(pp), p2 = hi16(pp), q1 = lo16(hh), q2 = hi16(hh);
  const int n = t.n, mp = t.mp;
  double efs = rec[t.off_fpp + (p1 - n) + (int64_t)(p1 - n) * mp];
  efs = efs + rec[t.off_fpp + (p2 - n) + (int64_t)(p2 - n) * mp];
  efs = efs - rec[t.off_fhh + q1 + (int64_t)q1 * n];
  efs = efs - rec[t.off_fhh + q2 + (int64_t)q2 * n];
  const double aph = velem(t, rec, hr, p1, p2, p1, p2) + velem(t, rec, hr, q1, q2, q1, q2) -
                     velem(t, rec, hr, p1, q1, p1, q1) - velem(t, rec, hr, p1, q2, p1, q2) -
                     velem(t, rec, hr, p2, q1, p2, q1) - velem(t, rec, hr, p2, q2, p2, q2);
  gen[sg.off + e] = rec[sg.off + e] / (efs + aph);
}

void launch_white(hipStream_t s, const DevTab &t, const double *rec, int hrec, double *gen, const Seg2 *pphh_segs,
                  const Tile2 *pphh_tiles, int ntiles) {
  const int64_t m = (int64_t)t.mp * t.n;
  if (m > 0) hipLaunchKernelGGL(k_white_ph, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, t, rec, hrec, gen);
  if (ntiles > 0)
    hipLaunchKernelGGL(k_white_pphh, dim3(ntiles), dim3(256), 0, s, t, rec, hrec, gen, pphh_segs, pphh_tiles);
}

// ---------------------------------------------------------------------------
// add_arrays: z = s1*x + s2*y
__global__ void __launch_bounds__(256) k_axpby(int64_t n, double s1, const double *__restrict__ x, double s2,
                                               const double *__restrict__ y, double *__restrict__ z) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    z[e] = s1 * x[e] + s2 * y[e];
}

void launch_axpby(hipStream_t s, int64_t n, double s1, const double *x, double s2, const double *y, double *z) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)blocks), dim3(256), 0, s, n, s1, x, s2, y, z);
}

// sum of squares over [1, n): 1024 fixed partials then one final block
constexpr int NPART = 1024;
__global__ void __launch_bounds__(256) k_sumsq_part(const double *__restrict__ v, int64_t n, double *part) {
  const int64_t chunk = (n - 1 + NPART - 1) / NPART;
  const int64_t b = 1 + (int64_t)blockIdx.x * chunk, e = b + chunk < n ? b + chunk : n;
  double s = 0.0;
  for (int64_t k = b + threadIdx.x; k < e; k += 256) s += v[k] * v[k];
  s = block_sum256(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void __launch_bounds__(256) k_sum_final(const double *__restrict__ part, int np, double *dst) {
  double s = 0.0;
  for (int k = threadIdx.x; k < np; k += 256) s += part[k];
  s = block_sum256(s);
  if (threadIdx.x == 0) *dst = s;
}

void launch_norm2(hipStream_t s, const double *v, int64_t neq, double *partials, double *dst) {
  hipLaunchKernelGGL(k_sumsq_part, dim3(NPART), dim3(256), 0, s, v, neq, partials);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partials, NPART, dst);
}

// MBPT2 (src/HFQDmod.f90:325-362): sum Vpphh^2 / (f_h1 + f_h2 - f_p1 - f_p2)
__global__ void __launch_bounds__(256) k_mbpt2_part(DevTab t, const double *__restrict__ v,
                                                    const Seg2 *__restrict__ segs,
                                                    const Tile2 *__restrict__ tiles, double *part) {
  const Tile2 tl = tiles[blockIdx.x];
  const Seg2 sg = segs[tl.seg];
  const int e = tl.start + threadIdx.x;
  double s = 0.0;
  if (e < sg.size) {
    const int I = e % sg.rows, J = e / sg.rows;
    const uint32_t pp = t.pairs[sg.bra0 + I], hh = t.pairs[sg.ket0 + J];
    const int n = t.n, mp = t.mp;
    const int p1 = lo16(pp) - n, p2 = hi16(pp) - n, q1 = lo16(hh), q2 = hi16(hh);
    const double f12 = v[t.off_fpp + p1 + (int64_t)p1 * mp] + v[t.off_fpp + p2 + (int64_t)p2 * mp];
    const double den = v[t.off_fhh + q1 + (int64_t)q1 * n] + v[t.off_fhh + q2 + (int64_t)q2 * n] - f12;
    const double x = v[sg.off + e];
    s = x * x / den;
  }
  s = block_sum256(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

void launch_mbpt2(hipStream_t s, const DevTab &t, const double *v, const Seg2 *pphh_segs, const Tile2 *pphh_tiles,
                  int ntiles, double *partials, double *dst) {
  if (ntiles > 0)
    hipLaunchKernelGGL(k_mbpt2_part, dim3(ntiles), dim3(256), 0, s, t, v, pphh_segs, pphh_tiles, partials);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partials, ntiles, dst);
}

__global__ void k_add_scalar(double *dst, const double *a, const double *b) { *dst = *a + *b; }
void launch_scale_add_scalar(hipStream_t s, double *dst, const double *a, const double *b) {
  hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(1), 0, s, dst, a, b);
}

}  // namespace qd
