// Gather / reduction kernels of the IM-SRG(2) commutator (gfx950, FP64).
//
// Every kernel reads operators in the reference's flat vectorize layout.  The
// element accessor velem() restates v_elem (src/ME_general.f90:1557-1680):
// pair sorting with sign, the i_array lookup, zero across blocks, and the
// herm factor for the three sub-blocks that are stored transposed.
#include "qd_kernels.h"

namespace qd {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// In place D = hh * P - eh * P^T on every r x r CC product, one 32x32 tile
// pair (ti <= tj) per work-group through LDS.
__global__ void __launch_bounds__(256) k_pandya_combine(DevTab t, double *__restrict__ P,
                                                        const TilePair *__restrict__ tp, double hh, double eh) {
  __shared__ double A[32][33], B[32][33];
  const TilePair p = tp[blockIdx.x];
  const int r = t.cc_r[p.blk];
  double *base = P + t.cc_poff[p.blk];
  const int i0 = p.ti * 32, j0 = p.tj * 32;
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int li = e & 31, lj = e >> 5;
    A[li][lj] = (i0 + li < r && j0 + lj < r) ? base[(i0 + li) + (int64_t)(j0 + lj) * r] : 0.0;
    B[li][lj] = (j0 + li < r && i0 + lj < r) ? base[(j0 + li) + (int64_t)(i0 + lj) * r] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int li = e & 31, lj = e >> 5;
    if (i0 + li < r && j0 + lj < r) base[(i0 + li) + (int64_t)(j0 + lj) * r] = hh * A[li][lj] - eh * B[lj][li];
    if (p.ti != p.tj && j0 + li < r && i0 + lj < r)
      base[(j0 + li) + (int64_t)(i0 + lj) * r] = hh * B[li][lj] - eh * A[lj][li];
  }
}

void launch_pandya_combine(hipStream_t s, const DevTab &t, double *P, const TilePair *tp, int ntp, double hh,
                           double eh) {
  if (ntp <= 0) return;
  hipLaunchKernelGGL(k_pandya_combine, dim3(ntp), dim3(256), 0, s, t, P, tp, hh, eh);
}

// ---------------------------------------------------------------------------
// One-body part: commutator_111 (src/commutators.f90:137-195), _121 (:73-134)
// and the partial traces of _221 (:278-348) for the output elements of fhh,
// fpp, fph (contiguous in the flat layout from offset 1).  All reductions
// have a fixed order (deterministic).

// commutator_121 through the cross-coupled matrices: with
// CC_k[(p,q),(i,a)] = <q i|V_k|p a> (calc_cc), v(a,x,i,y) = h * CC[(x,y),(i,a)]
// and v(i,x,a,y) = CC[(y,x),(i,a)], so the 121 sum over (i,a) of
// src/commutators.f90:83-132 becomes u_k[(p,q)] = sum_(i,a) CC_k[(p,q),(i,a)]
// f_other(a,i): one wave per ordered pair (p,q), gathering V.
__global__ void __launch_bounds__(256) k_cc_matvec(DevTab t, const double *__restrict__ v, int herm,
                                                   const double *__restrict__ f, double *__restrict__ u) {
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);  // one wave per ordered pair
  const int lane = threadIdx.x & 63;
  if (e >= (int64_t)t.M * t.M) return;
  const int p = (int)(e / t.M), q = (int)(e % t.M);
  const int blk = t.cc_map[e].x;
  const int nb = t.cc_nb[blk];
  const uint32_t *hp = t.cc_hp + t.cc_hp_start[blk];
  double s = 0.0;
  for (int c = lane; c < nb; c += 64) {
    const int i = lo16(hp[c]), a = hi16(hp[c]);
    s += velem(t, v, herm, q, i, p, a) * f[t.off_fph + (a - t.n) + (int64_t)i * t.mp];
  }
  s = wave_sum(s);
  if (lane == 0) u[e] = s;
}

void launch_cc_matvec(hipStream_t s, const DevTab &t, const double *v, int herm, const double *f, double *u) {
  const int64_t m = (int64_t)t.M * t.M;
  hipLaunchKernelGGL(k_cc_matvec, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, s, t, v, herm, f, u);
}

// The same mat-vec from a calc_cc-filled CC set (element <q i|V|p a> at
// device row (p,q), column (i,a)).  A work-group takes 32 consecutive device
// rows; its 8 column groups stride the columns (coalesced over the rows, f(a,i)
// uniform across them) and their partial sums are added in a fixed order.
__global__ void __launch_bounds__(256) k_cc_gemv(DevTab t, const double *__restrict__ cc,
                                                 const double *__restrict__ f, double *__restrict__ u) {
  __shared__ double part[8][33];
  const int rl = threadIdx.x & 31, cg = threadIdx.x >> 5;
  const int g = blockIdx.x * 32 + rl;
  double s = 0.0;
  if (g < t.M * t.M) {
    const int K = t.cc_rowblk[g];  // CC block of device row g
    const int d = g - t.cc_ab_start[K], r = t.cc_r[K], nb = t.cc_nb[K];
    const double *col = cc + t.cc_off[K] + d;
    const uint32_t *hp = t.cc_hp + t.cc_hp_start[K];
#pragma unroll 4
    for (int c = cg; c < nb; c += 8) {
      const int i = lo16(hp[c]), a = hi16(hp[c]);
      s += col[(int64_t)c * r] * f[t.off_fph + (a - t.n) + (int64_t)i * t.mp];
    }
  }
  part[cg][rl] = s;
  __syncthreads();
  if (cg == 0 && g < t.M * t.M) {
    double v = part[0][rl];
    for (int k = 1; k < 8; k++) v += part[k][rl];
    const uint32_t pq = t.cc_ab[g];
    u[(int64_t)lo16(pq) * t.M + hi16(pq)] = v;
  }
}

// The same mat-vec without a CC set: each CC element is gathered from V
// through the calc_cc table (the element k_calc_cc_map would write), in the
// work split and summation order of k_cc_gemv.  Saves the CC set's write and
// re-read, and replaces k_cc_matvec's per-element pair resolution (several
// dependent loads) by one table load.
__global__ void __launch_bounds__(256) k_cc_gemv_map(DevTab t, const double *__restrict__ v, int herm,
                                                     const double *__restrict__ f, double *__restrict__ u) {
  __shared__ double part[8][33];
  const int rl = threadIdx.x & 31, cg = threadIdx.x >> 5;
  const int g = blockIdx.x * 32 + rl;
  double s = 0.0;
  if (g < t.M * t.M) {
    const int K = t.cc_rowblk[g];  // CC block of device row g
    const int d = g - t.cc_ab_start[K], r = t.cc_r[K], nb = t.cc_nb[K];
    const uint32_t *col = t.cc_gmap + t.cc_off[K] + d;
    const uint32_t *hp = t.cc_hp + t.cc_hp_start[K];
#pragma unroll 4
    for (int c = cg; c < nb; c += 8) {
      const uint32_t m = col[(int64_t)c * r], k = m & CC_ZERO;
      double x = 0.0;
      if (k != CC_ZERO) {
        x = v[k];
        if ((m >> 30) & 1u) x = herm * x;
        if (m >> 31) x = -x;
      }
      const int i = lo16(hp[c]), a = hi16(hp[c]);
      s += x * f[t.off_fph + (a - t.n) + (int64_t)i * t.mp];
    }
  }
  part[cg][rl] = s;
  __syncthreads();
  if (cg == 0 && g < t.M * t.M) {
    double w = part[0][rl];
    for (int k = 1; k < 8; k++) w += part[k][rl];
    const uint32_t pq = t.cc_ab[g];
    u[(int64_t)lo16(pq) * t.M + hi16(pq)] = w;
  }
}

// Output element e of fhh, fpp, fph (flat offset 1 + e): kind 0 hh, 1 pp, 2 ph;
// (i, j) within its block, (x, y) as states
__device__ __forceinline__ void ob_decode(const DevTab &t, int64_t e, int &kind, int &i, int &j, int &x, int &y) {
  const int n = t.n, mp = t.mp;
  const int64_t nhh = (int64_t)n * n, npp = (int64_t)mp * mp;
  if (e < nhh) { kind = 0; i = (int)(e % n); j = (int)(e / n); x = i; y = j; }
  else if (e < nhh + npp) { kind = 1; int64_t f = e - nhh; i = (int)(f % mp); j = (int)(f / mp); x = n + i; y = n + j; }
  else { kind = 2; int64_t f = e - nhh - npp; i = (int)(f % mp); j = (int)(f / mp); x = n + i; y = j; }
}

// commutator_111 (src/commutators.f90:137-195) from the one-body products that
// the FP64 product list of the one-body pass leaves in the one-body region of w1,
// w2, w3 (P0 = A_hh B_hh, Q0 = B_ph^T A_ph, P1 = B_pp A_pp, Q1 = A_ph B_ph^T,
// R2 = A_ph B_hh - B_ph A_hh + A_pp B_ph - B_pp A_ph in three parts, for
// A = r1, B = r2), and
// commutator_121 (:73-134) from the CC mat-vecs u1, u2: one thread per output.
__global__ void __launch_bounds__(256) k_onebody(DevTab t, Ops o, double *__restrict__ out, int flags,
                                                 const double *__restrict__ w1, const double *__restrict__ w2,
                                                 const double *__restrict__ w3, const double *__restrict__ u1,
                                                 const double *__restrict__ u2) {
  const int n = t.n, mp = t.mp;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * n + (int64_t)mp * mp + (int64_t)mp * n) return;
  int kind, i, j, x, y;
  ob_decode(t, e, kind, i, j, x, y);
  const double h1 = o.h1, h2 = o.h2;
  const int64_t dst = 1 + e;
  double acc = (flags & OB_ACCUM) ? out[dst] : 0.0;
  if (flags & OB_111) {
    double val;
    if (kind == 0) {
      const double *P = w1 + t.off_fhh, *Q = w2 + t.off_fhh;
      const double q1 = P[i + (int64_t)j * n], q1t = P[j + (int64_t)i * n];
      const double q2 = Q[i + (int64_t)j * n], q2t = Q[j + (int64_t)i * n];
      val = q1 - h2 * q2 + h1 * (q2t - h2 * q1t);
    } else if (kind == 1) {
      const double *P = w1 + t.off_fpp, *Q = w2 + t.off_fpp;
      const double q1 = P[i + (int64_t)j * mp], q1t = P[j + (int64_t)i * mp];
      const double q2 = Q[i + (int64_t)j * mp], q2t = Q[j + (int64_t)i * mp];
      val = h2 * q2 + h1 * (h2 * q1t - q2t) - q1;
    } else {
      const int64_t k = t.off_fph + i + (int64_t)j * mp;
      val = w1[k] + w2[k] - w3[k];
    }
    acc = (flags & OB_ACCUM) ? acc + val : val;
  }
  if (flags & OB_121) {
    // u2 = CC(r2) f(r1), u1 = CC(r1) f(r2)
    const int64_t xy = (int64_t)x * t.M + y, yx = (int64_t)y * t.M + x;
    acc += (double)(o.h1 * o.h2) * u2[xy] - u2[yx] - ((double)(o.h1 * o.h2) * u1[xy] - u1[yx]);
  }
  out[dst] = acc;
}

// partial traces of commutator_221 (src/commutators.f90:278-348) for the
// outputs whose two states share a single-particle channel (the others are
// zero): one wave per listed output, added to out
__global__ void __launch_bounds__(256) k_onebody221(DevTab t, Ops o, double *__restrict__ out,
                                                    const int *__restrict__ list, int nlist,
                                                    const double *__restrict__ w1, const double *__restrict__ w2) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= nlist) return;  // whole wave exits together
  const int64_t e = list[k];
  int kind, i, j, x, y;
  ob_decode(t, e, kind, i, j, x, y);
  const int n = t.n;
  const double h1 = o.h1, h2 = o.h2;
  // w1, w2 have herm = +1
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
  r = wave_sum(r);
  if (lane == 0) out[1 + e] += r;
}

// The same traces from the code table: per lane and state a, one coalesced
// table load and two gathers, with the arrays and coefficients chosen by
// selects (no branch splits the loads).
__global__ void __launch_bounds__(256) k_onebody221_map(DevTab t, Ops o, double *__restrict__ out,
                                                        const int *__restrict__ list, int nlist,
                                                        const uint2 *__restrict__ g, const double *__restrict__ w1,
                                                        const double *__restrict__ w2) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= nlist) return;  // whole wave exits together
  const int64_t e = list[k];
  int kind, i, j, x, y;
  ob_decode(t, e, kind, i, j, x, y);
  const int n = t.n, M = t.M;
  const double h1 = o.h1, h2 = o.h2;
  const uint2 *gk = g + (int64_t)k * M;
  double r = 0.0;
#pragma unroll 4
  for (int a = lane; a < M; a += 64) {
    const uint2 m = gk[a];  // codes of <a x|w|a y>, <a y|w|a x>
    const uint32_t ka = m.x & CC_ZERO, kb = m.y & CC_ZERO;
    const bool hole = a < n;
    // hole:     kind 2: h1 w2(axay) - h2 w1(axay); else h1 w1(ayax) - h2 w1(axay)
    // particle: kind 1: -h1 w1(ayax) + h2 w1(axay); kind 0: -h1 w2(ayax) + h2 w2(axay);
    //           kind 2: -h1 w2(axay) + h2 w1(axay)
    const bool first_ax = kind == 2;
    const double *P1 = (kind == 2 || (!hole && kind == 0)) ? w2 : w1;
    const double *P2 = (!hole && kind == 0) ? w2 : w1;
    const uint32_t c1 = first_ax ? m.x : m.y, k1 = first_ax ? ka : kb;
    const double v1 = P1[k1 == CC_ZERO ? 0 : k1], v2 = P2[ka == CC_ZERO ? 0 : ka];
    const double x1 = k1 == CC_ZERO ? 0.0 : ((c1 >> 31) ? -v1 : v1);
    const double x2 = ka == CC_ZERO ? 0.0 : ((m.x >> 31) ? -v2 : v2);
    r += hole ? h1 * x1 - h2 * x2 : -h1 * x1 + h2 * x2;
  }
  r = wave_sum(r);
  if (lane == 0) out[1 + e] += r;
}

void launch_onebody(hipStream_t s, const DevTab &t, const Ops &o, double *out, int flags, const double *w1,
                    const double *w2, const double *w3, double *u1, double *u2, const double *cc1,
                    const double *cc2, const int *list221, int n221, const uint32_t *g221) {
  if (flags & OB_NOMV) {  // u1, u2 filled by an earlier OB_MV call
  } else if ((flags & OB_121) && cc1 && cc2) {
    const int m = t.M * t.M;
    hipLaunchKernelGGL(k_cc_gemv, dim3((m + 31) / 32), dim3(256), 0, s, t, cc2, o.v1, u2);
    hipLaunchKernelGGL(k_cc_gemv, dim3((m + 31) / 32), dim3(256), 0, s, t, cc1, o.v2, u1);
  } else if ((flags & OB_121) && t.cc_gmap) {
    const int m = t.M * t.M;
    hipLaunchKernelGGL(k_cc_gemv_map, dim3((m + 31) / 32), dim3(256), 0, s, t, o.v2, o.h2, o.v1, u2);
    hipLaunchKernelGGL(k_cc_gemv_map, dim3((m + 31) / 32), dim3(256), 0, s, t, o.v1, o.h1, o.v2, u1);
  } else if (flags & OB_121) {
    launch_cc_matvec(s, t, o.v2, o.h2, o.v1, u2);
    launch_cc_matvec(s, t, o.v1, o.h1, o.v2, u1);
  }
  if (flags & OB_MV) return;
  const int64_t tot = (int64_t)t.n * t.n + (int64_t)t.mp * t.mp + (int64_t)t.mp * t.n;
  if ((flags & (OB_111 | OB_121)) || !(flags & OB_ACCUM))
    hipLaunchKernelGGL(k_onebody, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, t, o, out, flags, w1, w2, w3,
                       u1, u2);
  if ((flags & OB_221) && n221 > 0) {
    if (g221)
      hipLaunchKernelGGL(k_onebody221_map, dim3((n221 + 3) / 4), dim3(256), 0, s, t, o, out, list221, n221,
                         (const uint2 *)g221, w1, w2);
    else
      hipLaunchKernelGGL(k_onebody221, dim3((n221 + 3) / 4), dim3(256), 0, s, t, o, out, list221, n221, w1, w2);
  }
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
__global__ void __launch_bounds__(256) k_zerobody_part(DevTab t, Ops o, const Seg2 *__restrict__ segs,
                                                       const Tile2 *__restrict__ tiles, int which, double *part) {
  double a = 0.0;
  if (blockIdx.x == 0) {
    if (which & 1) {
      const int64_t m = (int64_t)t.mp * t.n;
      for (int64_t e = threadIdx.x; e < m; e += 256) a += o.v1[t.off_fph + e] * o.v2[t.off_fph + e];
    }
  } else if (which & 2) {  // one tile of 256 elements of a Vpphh sub-block
    const Tile2 tl = tiles[blockIdx.x - 1];
    const Seg2 sg = segs[tl.seg];
    const int e = tl.start + threadIdx.x;
    if (e < sg.size) a = o.v1[sg.off + e] * o.v2[sg.off + e];
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

void launch_zerobody(hipStream_t s, const DevTab &t, const Ops &o, const Seg2 *pphh_segs, const Tile2 *pphh_tiles,
                     int ntiles, double *part, double *dst, int which) {
  hipLaunchKernelGGL(k_zerobody_part, dim3(1 + ntiles), dim3(256), 0, s, t, o, pphh_segs, pphh_tiles, which, part);
  hipLaunchKernelGGL(k_zero_final, dim3(1), dim3(256), 0, s, o, part, 1 + ntiles, dst);
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

// the same gather from the table: one coalesced index load, one V load and
// one store per element (k_calc_cc resolves the pair tables per element)
__global__ void __launch_bounds__(256) k_calc_cc_map(const uint32_t *__restrict__ map, const double *__restrict__ v,
                                                     int herm, double *__restrict__ cc, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const uint32_t m = map[e], k = m & CC_ZERO;
  double x = 0.0;
  if (k != CC_ZERO) {
    x = v[k];
    if ((m >> 30) & 1u) x = herm * x;
    if (m >> 31) x = -x;
  }
  cc[e] = x;
}

void launch_calc_cc(hipStream_t s, const DevTab &t, const double *v, int herm, double *cc, const Seg2 *cc_segs,
                    const Tile2 *cc_tiles, int ntiles) {
  if (ntiles <= 0) return;
  if (t.cc_gmap) {
    hipLaunchKernelGGL(k_calc_cc_map, dim3((unsigned)((t.cc_size + 255) / 256)), dim3(256), 0, s, t.cc_gmap, v, herm,
                       cc, t.cc_size);
    return;
  }
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

__global__ void k_zero_scan_init(uint32_t *nz) { *nz = 0u; }

__global__ void __launch_bounds__(256) k_zero_scan(const double *__restrict__ v, const ZTile *__restrict__ tiles,
                                                   uint32_t *nz) {
  const ZTile t = tiles[blockIdx.x];
  bool any = false;
  for (int k = threadIdx.x; k < t.len; k += 256) any = any || v[t.off + k] != 0.0;
  if (__any(any) && (threadIdx.x & 63) == 0) atomicOr(nz, 1u << t.bit);
}

__global__ void __launch_bounds__(256) k_zero_chunks(Slots sl, const ZChunk *__restrict__ chunks) {
  const ZChunk c = chunks[blockIdx.x];
  double *__restrict__ p = sl.p[c.slot] + c.off;
  for (int k = threadIdx.x; k < c.len; k += 256) p[k] = 0.0;
}

void launch_zero_chunks(hipStream_t s, const Slots &slots, const ZChunk *chunks, int n) {
  if (n > 0) hipLaunchKernelGGL(k_zero_chunks, dim3(n), dim3(256), 0, s, slots, chunks);
}

void launch_zero_scan(hipStream_t s, const double *v, const ZTile *tiles, int ntiles, uint32_t *nz) {
  hipLaunchKernelGGL(k_zero_scan_init, dim3(1), dim3(1), 0, s, nz);
  if (ntiles > 0) hipLaunchKernelGGL(k_zero_scan, dim3(ntiles), dim3(256), 0, s, v, tiles, nz);
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

// One BCH term's bookkeeping in one pass (src/magnus_IMSRG.f90:300-316):
// ||HS||^2 before the update and ||AD||^2 over [1, n) with the partition and
// summation order of k_sumsq_part (the same sums to the bit), and HS += cof AD
// over [0, n)
__global__ void __launch_bounds__(256) k_bch_update(double *__restrict__ hs, const double *__restrict__ ad,
                                                    double cof, int64_t n, double *part) {
  const int64_t chunk = (n - 1 + NPART - 1) / NPART;
  const int64_t b = 1 + (int64_t)blockIdx.x * chunk, e = b + chunk < n ? b + chunk : n;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t k = b + threadIdx.x; k < e; k += 256) {
    const double x = hs[k], y = ad[k];
    s1 += x * x;
    s2 += y * y;
    hs[k] = x + cof * y;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) hs[0] = hs[0] + cof * ad[0];
  s1 = block_sum256(s1);
  s2 = block_sum256(s2);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1;
    part[NPART + blockIdx.x] = s2;
  }
}

void launch_bch_update(hipStream_t s, double *hs, const double *ad, double cof, int64_t neq, double *partials,
                       double *dst) {
  hipLaunchKernelGGL(k_bch_update, dim3(NPART), dim3(256), 0, s, hs, ad, cof, neq, partials);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partials, NPART, dst);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partials + NPART, NPART, dst + 1);
}

// dot product over [1, n) (E0 excluded, as in mat_2_norm), same partition
__global__ void __launch_bounds__(256) k_dot_part(const double *__restrict__ a, const double *__restrict__ b,
                                                  int64_t n, double *part) {
  const int64_t chunk = (n - 1 + NPART - 1) / NPART;
  const int64_t lo = 1 + (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  double s = 0.0;
  for (int64_t k = lo + threadIdx.x; k < hi; k += 256) s += a[k] * b[k];
  s = block_sum256(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

void launch_dot(hipStream_t s, const double *a, const double *b, int64_t neq, double *partials, double *dst) {
  hipLaunchKernelGGL(k_dot_part, dim3(NPART), dim3(256), 0, s, a, b, neq, partials);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partials, NPART, dst);
}

// z = sum_i c_i x_i over up to LINCOMB_MAX vectors, one pass
__global__ void __launch_bounds__(256) k_lincomb(int64_t n, LinComb lc, double *__restrict__ z) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    double acc = 0.0;
    for (int i = 0; i < lc.k; i++) acc = fma(lc.c[i], lc.x[i][e], acc);
    z[e] = acc;
  }
}

void launch_lincomb(hipStream_t s, int64_t n, const LinComb &lc, double *z) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_lincomb, dim3((unsigned)blocks), dim3(256), 0, s, n, lc, z);
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
