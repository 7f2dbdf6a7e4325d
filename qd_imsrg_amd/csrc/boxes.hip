// Two-body part of the IM-SRG(2) commutator over class-pair boxes (gfx950,
// FP64): commutator_122, the ladder combinations of matrix_mult222 and the
// Pandya term of commutator_222, one write per stored output.
//
// Every substitution of the 1b x 2b commutator keeps the replaced state in
// its single-particle (m_l, m_s) channel, so the stored outputs
// <x1 x2|V|x3 x4> with x1, x2 in channels (A1, A2) and x3, x4 in (A3, A4)
// only need V on the ordered grid A1 x A2 x A3 x A4.  On that grid
// [f, V](x1..x4) = sum_t f(x1,t) V(t,x2,x3,x4) + f(x2,t) V(x1,t,x3,x4)
//                  - herm(f) (f(x3,t) V(x1,x2,t,x4) + f(x4,t) V(x1,x2,x3,t))
// (src/commutators.f90:355-676, which sums the same terms split by hole and
// particle substitutes) is four dense contractions with |c| x |c| channel
// blocks of f.  A work-group stages the grids of both operands and the
// channel blocks in LDS, contracts, and writes the stored elements of the
// grid with the sign of their canonical order.
#include "qd_kernels.h"

namespace qd {

__global__ void __launch_bounds__(256) k_fclass(DevTab t, Ops o, double *__restrict__ F) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= t.fcls_size) return;
  int lo = 0, hi = t.nch - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.fcls_off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  const int c0 = t.ch_start[lo], sz = t.ch_start[lo + 1] - c0, le = e - t.fcls_off[lo];
  const int x = t.ch_states[c0 + le / sz], y = t.ch_states[c0 + le % sz];
  F[e] = felem(t, o.v1, o.h1, x, y);
  F[t.fcls_size + e] = felem(t, o.v2, o.h2, x, y);
}

void launch_fclass(hipStream_t s, const DevTab &t, const Ops &o, double *F) {
  if (t.fcls_size > 0) hipLaunchKernelGGL(k_fclass, dim3((t.fcls_size + 255) / 256), dim3(256), 0, s, t, o, F);
}

// last box b in [lo, ng) with pre[b] <= e
__device__ __forceinline__ int box_of(const int *pre, int lo, int ng, int e) {
  int hi = ng - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// stored sub-block (enum Sub) of (row pair type, column pair type), row >= column
__device__ __forceinline__ int sub_type(int ta, int tb) {
  return tb == 0 ? (ta == 0 ? HHHH : (ta == 1 ? PHHH : PPHH)) : (tb == 1 ? (ta == 1 ? PHPH : PPPH) : PPPP);
}

// floor(a / d) for 0 <= a < 2^22 and 1 <= d < 2^12 from rd = 1.0f / d: the
// product's relative error (< 2^-23) times (a + 0.5) / d stays below the
// distance 0.5 / d of (a + 0.5) / d from the nearest integer
__device__ __forceinline__ int qdiv(int a, float rd) { return (int)(((float)a + 0.5f) * rd); }

// NT threads per task, grids of up to VC elements, RC row grid entries;
// STAGE: the D staging path (TB_DSTAGE) is compiled in; SPLIT: paired boxes
// write their transposed outputs in a second phase (below); LDIR: the square
// sub-blocks' ladder terms are formed here from the products (TB_LDIRECT)
template <int NT, int VC, int RC, bool STAGE, bool SPLIT, bool LDIR>
__global__ void __launch_bounds__(NT) k_box2b(DevTab t, Ops o, double *__restrict__ out,
                                               const BTask *__restrict__ tasks, const int *__restrict__ blist,
                                               int flags, const double *__restrict__ w1,
                                               const double *__restrict__ w2, const double *__restrict__ w3,
                                               const double *__restrict__ L, const double *__restrict__ F,
                                               const double *__restrict__ D, double dsign) {
  __shared__ double2 gg[VC];              // ordered grids of (V1, V2): [row][ldc]
  __shared__ double2 fl[BX_FCAP / 2];          // channel blocks of (f1, f2): A3, A4, then A1, A2 per box
  __shared__ int pre[BX_GMAX + 1];             // element offsets of the boxes
  __shared__ int4 bx[BX_GMAX];                 // class pair of the rows, rows in the task, grid offset, F offset
  __shared__ int4 bs[BX_GMAX];                 // |A1|, |A2|, A1 == A2, first row entry in rent
  __shared__ uint32_t rent[RC];           // grid entries of the task's rows, box after box
  __shared__ uint32_t cent[BX_CMAX];           // grid entries of the task's columns
  __shared__ int64_t soff[3][3];               // stored sub-block of (row type, column type), row >= column
  __shared__ float2 brc[BX_GMAX];              // 1 / rows in the task, 1 / |A2|
  __shared__ int64_t pb[BX_GMAX][4];           // Pandya: offset in D of each term's stored sub-block (-1: none)
  __shared__ int pl[BX_GMAX][4];               // Pandya: its leading dimension | stored transposed << 30
  __shared__ int2 bfo[BX_GMAX];                // offsets of the box's A1 and A2 blocks in F
  __shared__ int fsz_s;                        // f entries staged (A3, A4 and every box's A1, A2)
  __shared__ int bgo[BX_GMAX];                 // first row grid entry of the box in cp_grid
  __shared__ int bpr[BX_GMAX];                 // the box also writes the transposed box's outputs
  __shared__ int rtot_s;                       // row grid entries of the task
  __shared__ int anyp_s;                       // some box of the task is paired
  const BTask tk = tasks[blockIdx.x];
  if (tk.block < 0) return;
  // ZB bits of r1's nonzero sub-block types (all set outside the BCH terms)
  const uint32_t nz1 = (uint32_t)flags >> TB_NZ_SHIFT;
  const int q = tk.block, mode = tk.mode, ng = tk.ng;
  const int nh = t.nh[q], nb = t.nb[q], np = t.np[q];
  const int tstart[3] = {0, nh, nh + nb}, tcount[3] = {nh, nb, np};
  const int A3 = t.cp_a[tk.gcol], A4 = t.cp_b[tk.gcol];
  const int s3 = t.ch_start[A3 + 1] - t.ch_start[A3], s4 = t.ch_start[A4 + 1] - t.ch_start[A4];
  const int nc = mode == 1 ? tk.cnt : s3 * s4, ldc = nc | 1;
  const int c_off = mode == 1 ? tk.lo : 0, r_off = mode == 2 ? tk.lo : 0;
  const uint32_t *cgrid = t.cp_grid + t.cp_grid_off[tk.gcol] + c_off;
  const bool do122 = (flags & TB_122) != 0, first = mode != 2;
  const bool dopd = first && (flags & TB_PANDYA) && D;
  if (threadIdx.x < 9) {
    const int tA = threadIdx.x / 3, tB = threadIdx.x % 3;
    const int sub = tB == 0 ? (tA == 0 ? HHHH : (tA == 1 ? PHHH : PPHH)) : (tB == 1 ? (tA == 1 ? PHPH : PPPH) : PPPP);
    soff[tA][tB] = tA >= tB ? t.off[6 * q + sub] : 0;
  }
  for (int k = threadIdx.x; k < nc; k += NT) cent[k] = cgrid[k];
  if (threadIdx.x < 64) {
    // wave 0, one lane per box: the box's entries of the table built with
    // the task lists (one load), then its offsets by exclusive prefix sums
    // over the lanes (ng <= BX_GMAX <= 64)
    static_assert(BX_GMAX <= 64, "one lane of wave 0 per box");
    const int b = threadIdx.x;
    int g = 0, s1 = 1, s2 = 1, f1 = 0, f2 = 0, go = 0, same12 = 0;
    int pr = 0;
    if (b < ng) {
      const int4 i0 = t.box_info[2 * (tk.first + b)], i1 = t.box_info[2 * (tk.first + b) + 1];
      g = i0.x;
      pr = (g & BOX_PAIRED) != 0;
      g &= ~BOX_PAIRED;
      go = i0.y;
      s1 = i0.z;
      s2 = i0.w;
      f1 = i1.x;
      f2 = i1.y;
      same12 = i1.z;
    }
    const int nr = b < ng ? (mode == 2 ? tk.cnt : s1 * s2) : 0;
    const int fn = b < ng ? s1 * s1 + s2 * s2 : 0;
    int e = nr * nc, gz = nr * ldc, fz = fn, rz = nr;  // inclusive sums
    for (int d = 1; d < 64; d <<= 1) {
      const int e2 = __shfl_up(e, d, 64), g2 = __shfl_up(gz, d, 64), f2s = __shfl_up(fz, d, 64),
                r2 = __shfl_up(rz, d, 64);
      if (b >= d) {
        e += e2;
        gz += g2;
        fz += f2s;
        rz += r2;
      }
    }
    const int q34 = s3 * s3 + s4 * s4;
    if (b < ng) {
      bx[b] = make_int4(g, nr, gz - nr * ldc, q34 + fz - fn);
      bs[b] = make_int4(s1, s2, same12, rz - nr);
      brc[b] = make_float2(1.0f / (float)nr, 1.0f / (float)s2);
      pre[b] = e - nr * nc;
      bfo[b] = make_int2(f1, f2);
      bgo[b] = go + r_off;
      bpr[b] = pr;
    }
    const bool anyp = __any(pr);
    if (b == ng - 1) {
      pre[ng] = e;
      fsz_s = q34 + fz;
      rtot_s = rz;
      anyp_s = anyp;
    }
  } else if (dopd && (int)threadIdx.x - 64 < ng) {
    // wave 1, one lane per box: the bases of its four Pandya terms,
    // D(x3,x1|x2,x4), D(x3,x2|x1,x4), D(x4,x2|x1,x3), D(x4,x1|x2,x3), from
    // the table built with the task lists (box_pand_terms in capi.hip)
    const int b = threadIdx.x - 64;
    const int64_t e = ((int64_t)(dsign == 0.0 ? 0 : 1) * t.nblist + tk.first + b) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      pb[b][k] = t.box_pb[e + k];
      pl[b][k] = t.box_pl[e + k];
    }
  }
  __syncthreads();
  {
    // row grid entries of all boxes in one flat pass
    const int rtot = rtot_s;
    for (int k = threadIdx.x; k < rtot; k += NT) {
      int lo = 0, hi = ng - 1;  // last box whose rows start at or before k
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bs[mid].w <= k) lo = mid;
        else hi = mid - 1;
      }
      rent[k] = t.cp_grid[bgo[lo] + (k - bs[lo].w)];
    }
  }
  __syncthreads();
  const int total = pre[ng], fsz = fsz_s;
  const float rs4 = 1.0f / (float)s4;
  // value of Pandya term k at (row a, column c) of its sub-block, the
  // sub-block at src (global D or its staged copy in LDS)
  auto dterm = [&](const double *__restrict__ src, int b, int k, int a, int c) -> double {
    const int w = pl[b][k], ld = w & 0x3fffffff;
    const bool tr = (w >> 30) != 0;
    return (tr ? dsign : 1.0) * src[tr ? c + (int64_t)a * ld : a + (int64_t)c * ld];
  };
  // Pandya term of whole boxes with D in run panels: every term's sub-block
  // is contiguous, so it is staged in LDS with coalesced loads (two terms at
  // a time, in the grid's LDS before the grids), and each thread keeps the
  // sum for its grid elements e = tid + k NT, in order, in registers
  constexpr int KQ = (VC + NT - 1) / NT;
  const bool stage_d = STAGE && dopd && mode == 0 && dsign != 0.0 && (flags & TB_DSTAGE);
  // One register queue rq serves both the staged Pandya sums (entry k: element
  // tid + k NT, consumed from the front) and the values kept for the second
  // write phase (appended at the back as the front is consumed)
  constexpr int KV = STAGE || SPLIT ? KQ : 1;
  double rq[KV];
#pragma unroll
  for (int k = 0; k < KV; k++) rq[k] = 0.0;
  if (stage_d) {
    double *__restrict__ gd = (double *)gg;  // [2][total]
    for (int h = 0; h < 2; h++) {
#pragma unroll
      for (int w = 0; w < 2; w++) {
        const int k = 2 * h + w;
        for (int j0 = threadIdx.x, b = 0; j0 < total; j0 += 4 * NT) {
          double v[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int j = min(j0 + NT * u, total - 1);
            if (b + 1 < ng && pre[b + 1] <= j) b = box_of(pre, b + 1, ng, j);
            const int64_t base = pb[b][k];  // no D block: the zero at D[p_size]
            v[u] = D[base < 0 ? t.p_size : base + (j - pre[b])];
          }
#pragma unroll
          for (int u = 0; u < 4; u++)
            if (j0 + NT * u < total) gd[w * total + j0 + NT * u] = v[u];
        }
      }
      __syncthreads();
      int b = 0;
#pragma unroll
      for (int k = 0; k < (STAGE ? KV : 1); k++) {
        const int e = threadIdx.x + k * NT;
        if (e < total) {
          if (b + 1 < ng && pre[b + 1] <= e) b = box_of(pre, b + 1, ng, e);
          const int le = e - pre[b], c = qdiv(le, brc[b].x), r = le - c * bx[b].y;
          const int s1 = bs[b].x, s2 = bs[b].y;
          const int i1 = qdiv(r, brc[b].y), i2 = r - i1 * s2, i3 = qdiv(c, rs4), i4 = c - i3 * s4;
          const double *__restrict__ g0 = gd + pre[b], *__restrict__ g1 = gd + total + pre[b];
          if (h == 0)
            rq[k] = dterm(g0, b, 0, i3 * s1 + i1, i2 * s4 + i4) - dterm(g1, b, 1, i3 * s2 + i2, i1 * s4 + i4);
          else
            rq[k] = (rq[k] + dterm(g0, b, 2, i4 * s2 + i2, i1 * s3 + i3)) - dterm(g1, b, 3, i4 * s1 + i1, i2 * s3 + i3);
        }
      }
      __syncthreads();
    }
  }
  if (do122) {
    // channel blocks of f1 and f2 (A3, A4, then A1, A2 of every box) in one
    // flat pass, two entries per thread at a time with both loads issued
    // before the LDS stores
    const int fs = t.fcls_size, q3 = s3 * s3, q34 = q3 + s4 * s4, ftot = fsz;
    const int f3 = t.fcls_off[A3], f4 = t.fcls_off[A4];
    for (int k0 = threadIdx.x; k0 < ftot; k0 += 2 * NT) {
      double y1[2], y2[2];
      int kk[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int k = min(k0 + NT * u, ftot - 1);
        int src;
        if (k < q34) src = k < q3 ? f3 + k : f4 + (k - q3);
        else {
          int lo = 0, hi = ng - 1;  // last box whose blocks start at or before k
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (bx[mid].w <= k) lo = mid;
            else hi = mid - 1;
          }
          const int j = k - bx[lo].w, q1 = bs[lo].x * bs[lo].x;
          src = j < q1 ? bfo[lo].x + j : bfo[lo].y + (j - q1);
        }
        y1[u] = F[src];
        y2[u] = F[fs + src];
        kk[u] = k;
      }
#pragma unroll
      for (int u = 0; u < 2; u++)
        if (k0 + NT * u < ftot) fl[kk[u]] = make_double2(y1[u], y2[u]);
    }
    // ordered grids: V(x1,x2,x3,x4) = sign * (herm if stored transposed) *
    // stored element of the canonical pairs, 0 for x1 == x2 or x3 == x4.
    // Four elements per thread and pass: all their loads are issued
    // branch-free (past the end the last element is read again; an invalid
    // entry reads element 0) before the first LDS store, so the round trips
    // overlap; validity and sign are applied at the store.
    for (int e0 = threadIdx.x, b = 0; e0 < total; e0 += 4 * NT) {
      double x1[4], x2[4], g1[4], g2[4];
      int dst[4];
      bool ok[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = min(e0 + NT * u, total - 1);
        if (b + 1 < ng && pre[b + 1] <= e) b = box_of(pre, b + 1, ng, e);
        const int4 B = bx[b];
        const int le = e - pre[b], c = qdiv(le, brc[b].x), r = le - c * B.y;
        const uint32_t er = rent[bs[b].w + r], ec = cent[c];
        // stored element of (I, J), or of (J, I) times herm when the row type
        // is below the column type (branch-free over the pair types)
        const int tyr = (int)((er >> 18) & 3u), tyc = (int)((ec >> 18) & 3u);
        const int Il = (int)(er & 0xffffu) - tstart[tyr], Jl = (int)(ec & 0xffffu) - tstart[tyc];
        const bool tr = tyr < tyc;
        const int tA = tr ? tyc : tyr, tB = tr ? tyr : tyc;
        ok[u] = (er >> 17) & (ec >> 17) & 1u;
        const int64_t a = ok[u] ? soff[tA][tB] + (tr ? Jl : Il) + (int64_t)(tr ? Il : Jl) * tcount[tA] : 0;
        const bool neg = ((er ^ ec) >> 16) & 1u;
        // r1's sub-block types that are zero in every block are not read
        const bool z1 = !((nz1 >> sub_type(tA, tB)) & 1u);
        x1[u] = o.v1[z1 ? 0 : a];
        x2[u] = o.v2[a];
        g1[u] = z1 ? 0.0 : (neg ? -1.0 : 1.0) * (tr ? (double)o.h1 : 1.0);
        g2[u] = (neg ? -1.0 : 1.0) * (tr ? (double)o.h2 : 1.0);
        dst[u] = B.z + r * ldc + c;
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (e0 + NT * u < total)
          gg[dst[u]] = ok[u] ? make_double2(g1[u] * x1[u], g2[u] * x2[u]) : make_double2(0.0, 0.0);
    }
    __syncthreads();
  }
  const double h1 = o.h1, h2 = o.h2, hz = -h1 * h2;  // hz: herm of [r1, r2]
  const bool same34 = A3 == A4;
  // stored sub-block of (row type, column type), row type >= column type
  auto sub_of = [](int ta, int tb) { return sub_type(ta, tb); };
  // ladder inputs (signed combinations of the products of matrix_mult222,
  // src/commutators.f90:999-1127, and the summed products for Vpphh,
  // Vphhh, Vppph): xa = w1(I,J), xb = the summed product at (I,J) (L for
  // PPHH and PPPH, whose sums come in two halves; w3 for PHHH), xc =
  // w2(I,J) for PHHH / PPPH; the square sub-blocks read their whole ladder
  // term from L (without LDIR).  A value the sub-block does not use is read
  // at element 0 and ignored.  One set per stored output.
  const bool lad = first && (flags & TB_LADDER);
  // LDIR: the square sub-blocks (HHHH, PHPH, PPPP) read A = w1 and B = w3
  // (w2 for PHPH) at (I, J) and at (J, I) and form L(I,J) = U(J,I) + W(I,J),
  // U = h1 A + c B, W = -h2 A - d B (c = h1 h2, d = 1; PHPH c = h1, d = h2),
  // as k_ladder_sq does; both orientations come from the same four values.
  // A product of a sub-block type that is zero in r1 (ZB bits in flags >>
  // TB_NZ_SHIFT) is not read.
  auto ld = [&](int sub, int64_t idx, double &xa, double &xb, double &xc) {
    const bool sq = sub == HHHH || sub == PPPP || sub == PHPH;
    const bool usec = sub == PHHH || sub == PPPH;
    xa = (sq ? L : w1)[sub == PPHH ? 0 : idx];
    xb = (sub == PHHH ? w3 : L)[sq ? 0 : idx];
    xc = w2[usec ? idx : 0];
  };
  auto finish = [&](int sub, double acc, double xa, double xb, double xc) {
    if (lad) {
      switch (sub) {
        case HHHH:
        case PPPP:
        case PHPH: acc += xa; break;
        case PPHH: acc += xb; break;
        case PHHH: acc += h1 * xc - h2 * xa + xb; break;
        case PPPH: acc += -h2 * xa + h1 * xc + xb; break;
        default: break;
      }
    }
    return acc;
  };
  // the bra slabs of a split box add to what the ket slabs wrote
  const bool rd = first ? ((flags & TB_ACCUM) || (mode == 1 && do122)) : (flags & TB_ACCUM) != 0;
  // Whole boxes with paired boxes among them write in two phases: the outputs
  // <I|Z|J> of every grid element here, with consecutive threads on
  // consecutive rows (consecutive stored rows of a column: few cache lines
  // per wave); the paired outputs <J|Z|I> = h_Z (...) after all grid elements
  // are contracted, from the values kept in LDS, with consecutive threads on
  // consecutive columns, so that their stored rows are consecutive too.
  // Each thread keeps its values in registers (rq, in order) until
  // the grids in LDS are no longer read.
  // (The narrow form only: in the wide form, one 16-wave task per CU, the
  // extra barriers cost more than the scattered writes: N=42 R=20 two-body
  // class 2.30 -> 2.39 ms, while N=20 R=16 gains 0.481 -> 0.461 ms.)
  const bool split = SPLIT && mode == 0 && anyp_s;
  for (int e = threadIdx.x, b = 0; e < total; e += NT) {
    // this element's staged Pandya sum; the queues move up by one
    const double pdv = rq[0];
#pragma unroll
    for (int k = 0; k + 1 < KV; k++) rq[k] = rq[k + 1];
    if (b + 1 < ng && pre[b + 1] <= e) b = box_of(pre, b + 1, ng, e);
    const int4 B = bx[b];
    const int le = e - pre[b], c = qdiv(le, brc[b].x), r = le - c * B.y;
    const int rr = r_off + r, cc = c_off + c;
    const uint32_t er = rent[bs[b].w + r], ec = cent[c];
    if (!((er >> 17) & (ec >> 17) & 1u)) continue;
    const int tyr = (int)((er >> 18) & 3u), tyc = (int)((ec >> 18) & 3u);
    const bool sr = (er >> 16) & 1u, sc = (ec >> 16) & 1u;
    const int4 S = bs[b];
    // one orientation of a same-channel pair
    if ((S.z && sr) || (same34 && sc)) continue;
    // stored: <I|Z|J> with row type >= column type; in a paired box also
    // <J|Z|I> (column type >= row type), here unless written in the second phase
    const bool st1 = tyr >= tyc, st2 = bpr[b] && tyc >= tyr && !split;
    if (!st1 && !(bpr[b] && tyc >= tyr)) continue;
    const int s1 = S.x, s2 = S.y;
    const int i1 = qdiv(rr, brc[b].y), i2 = rr - i1 * s2, i3 = qdiv(cc, rs4), i4 = cc - i3 * s4;
    // stored positions (sub-blocks of the flat layout); 0 where not stored
    const int Il = (int)(er & 0xffffu) - tstart[tyr], Jl = (int)(ec & 0xffffu) - tstart[tyc];
    const int sub1 = st1 ? sub_of(tyr, tyc) : HHHH, sub2 = st2 ? sub_of(tyc, tyr) : HHHH;
    const int64_t idx1 = st1 ? soff[tyr][tyc] + Il + (int64_t)Jl * tcount[tyr] : 0;
    const int64_t idx2 = st2 ? soff[tyc][tyr] + Jl + (int64_t)Il * tcount[tyc] : 0;
    // global loads first, so that their latency overlaps the LDS contraction.
    // Every load is issued unconditionally and nothing reads its value before
    // the combination: loads under divergent branches share registers across
    // the paths, and the compiler then drains all outstanding loads between
    // them.  A term without a D block reads the zero at D[p_size].
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
    if (dopd && !stage_d) {
      // column slabs and whole column-major D: each term read from global D
      // (a term without a D block reads the zero at D[p_size])
      auto dg = [&](int k, int a, int c) -> double {
        const int64_t base = pb[b][k];
        const int w = pl[b][k], ld = w & 0x3fffffff;
        const bool tr = (w >> 30) != 0;
        const int64_t at = base < 0 ? t.p_size : base + (tr ? c + (int64_t)a * ld : a + (int64_t)c * ld);
        return (tr ? dsign : 1.0) * D[at];
      };
      d0 = dg(0, i3 * s1 + i1, i2 * s4 + i4);
      d1 = dg(1, i3 * s2 + i2, i1 * s4 + i4);
      d2 = dg(2, i4 * s2 + i2, i1 * s3 + i3);
      d3 = dg(3, i4 * s1 + i1, i2 * s3 + i3);
    }
    double xa1 = 0.0, xb1 = 0.0, xc1 = 0.0, xa2 = 0.0, xb2 = 0.0, xc2 = 0.0;
    const bool sqd = LDIR && tyr == tyc;  // st1 holds, and st2 in a paired box
    bool za = false, zb = false;
    if (lad) {
      if constexpr (LDIR) {
        if (sqd) {
          za = tyr == 1 && !((nz1 >> PPPH) & 1u);  // w1.phph = r2.ppph^T r1.ppph
          zb = !((nz1 >> (tyr == 0 ? HHHH : (tyr == 1 ? PHHH : PPPP))) & 1u);
        }
        const int64_t it = sqd ? soff[tyr][tyr] + Jl + (int64_t)Il * tcount[tyr] : 0;  // (J, I)
        const double *Bq = tyr == 1 ? w2 : w3;
        xa1 = sqd ? w1[za ? 0 : idx1] : 0.0;
        xb1 = sqd ? Bq[zb ? 0 : idx1] : 0.0;
        xa2 = sqd ? w1[za ? 0 : it] : 0.0;
        xb2 = sqd ? Bq[zb ? 0 : it] : 0.0;
        if (!sqd) ld(sub1, idx1, xa1, xb1, xc1);
        if (!sqd && !split) ld(sub2, idx2, xa2, xb2, xc2);
      } else {
        // a square sub-block's L is (anti)symmetric, L(J,I) = h_Z L(I,J):
        // its paired output needs no second read
        ld(sub1, idx1, xa1, xb1, xc1);
        if (!split) ld(tyr == tyc ? HHHH : sub2, tyr == tyc ? 0 : idx2, xa2, xb2, xc2);
      }
    }
    const double o1 = rd ? out[idx1] : 0.0;
    const double o2 = rd && !split ? out[idx2] : 0.0;
    double val = 0.0;
    if (do122) {
      const double2 *G = gg + B.z;
      if (first) {
        const double2 *fa = fl + B.w, *fb = fl + B.w + s1 * s1;
        double acc = 0.0;
        for (int k = 0; k < s1; k++) {
          const double2 f = fa[i1 * s1 + k], g = G[(k * s2 + i2) * ldc + c];
          acc += f.x * g.y - f.y * g.x;
        }
        for (int k = 0; k < s2; k++) {
          const double2 f = fb[i2 * s2 + k], g = G[(i1 * s2 + k) * ldc + c];
          acc += f.x * g.y - f.y * g.x;
        }
        val += acc;
      }
      if (mode != 1) {
        const double2 *fc = fl, *fd = fl + s3 * s3;
        double k2 = 0.0, k1 = 0.0;
        for (int k = 0; k < s3; k++) {
          const double2 f = fc[i3 * s3 + k], g = G[r * ldc + k * s4 + i4];
          k2 += f.x * g.y;
          k1 += f.y * g.x;
        }
        for (int k = 0; k < s4; k++) {
          const double2 f = fd[i4 * s4 + k], g = G[r * ldc + i3 * s4 + k];
          k2 += f.x * g.y;
          k1 += f.y * g.x;
        }
        val += h2 * k1 - h1 * k2;
      }
    }
    if (stage_d) val += pdv;
    else if (dopd) val += d0 - d1 + d2 - d3;
    if (sr != sc) val = -val;
    const bool sq = tyr == tyc;
    if (lad && sq) {
      if constexpr (LDIR) {
        const double a_ij = za ? 0.0 : xa1, b_ij = zb ? 0.0 : xb1, a_ji = za ? 0.0 : xa2, b_ji = zb ? 0.0 : xb2;
        const double cq = tyr == 1 ? h1 : h1 * h2, dq = tyr == 1 ? h2 : 1.0;
        xa1 = (h1 * a_ji + cq * b_ji) + (-h2 * a_ij - dq * b_ij);
        xa2 = (h1 * a_ij + cq * b_ij) + (-h2 * a_ji - dq * b_ji);
      } else {
        xa2 = hz * xa1;
      }
    }
    // the second phase adds this value times h_Z; a square sub-block's
    // ladder term is folded in (h_Z = +-1)
    rq[KV - 1] = lad && sq ? val + hz * xa2 : val;
    if (!first) {  // ket slabs run before the main pass: the first writers of their outputs
      if (st1) out[idx1] = o1 + val;
      if (st2) out[idx2] = o2 + hz * val;
      continue;
    }
    if (st1) out[idx1] = finish(sub1, o1 + val, xa1, xb1, xc1);
    if (st2) out[idx2] = finish(sub2, o2 + hz * val, xa2, xb2, xc2);
  }
  if constexpr (!SPLIT) return;
  if (!split) return;
  // second phase: the values of this thread's grid elements (iterations j of
  // the loop above, at rq[KV - n + j]) into LDS at their element positions,
  // then the paired outputs with consecutive threads on consecutive columns
  __syncthreads();  // every contraction has read the grids
  {
    double *__restrict__ gv = (double *)gg;
    const int n = total > (int)threadIdx.x ? (total - 1 - (int)threadIdx.x) / NT + 1 : 0;
#pragma unroll
    for (int k = 0; k < KV; k++) {
      const int j = k - (KV - n);
      if (j >= 0) gv[threadIdx.x + j * NT] = rq[k];
    }
  }
  __syncthreads();
  const double *__restrict__ gv = (const double *)gg;
  const float rnc = 1.0f / (float)nc;
  // four elements per thread and pass, all their loads in flight together
  // (an element without a paired output reads element 0 and writes nothing)
  constexpr int U2 = 4;
  for (int e0 = threadIdx.x, b = 0; e0 < total; e0 += U2 * NT) {
    double xa[U2], xb[U2], xc[U2], o2[U2], v[U2];
    int64_t ix[U2];
    int sb[U2];
    bool ok[U2];
#pragma unroll
    for (int u = 0; u < U2; u++) {
      const int e = min(e0 + NT * u, total - 1);
      if (b + 1 < ng && pre[b + 1] <= e) b = box_of(pre, b + 1, ng, e);
      const int nr = bx[b].y, le = e - pre[b], r = qdiv(le, rnc), c = le - r * nc;
      const uint32_t er = rent[bs[b].w + r], ec = cent[c];
      const int tyr = (int)((er >> 18) & 3u), tyc = (int)((ec >> 18) & 3u);
      const bool sr = (er >> 16) & 1u, sc = (ec >> 16) & 1u;
      ok[u] = e0 + NT * u < total && bpr[b] && ((er >> 17) & (ec >> 17) & 1u) && !(bs[b].z && sr) &&
              !(same34 && sc) && tyc >= tyr;
      const int Il = (int)(er & 0xffffu) - tstart[tyr], Jl = (int)(ec & 0xffffu) - tstart[tyc];
      sb[u] = ok[u] ? sub_of(tyc, tyr) : HHHH;
      ix[u] = ok[u] ? soff[tyc][tyr] + Jl + (int64_t)Il * tcount[tyc] : 0;
      xa[u] = xb[u] = xc[u] = 0.0;
      const bool folded = tyc == tyr;  // its ladder term is in the value
      if (lad) ld(folded ? HHHH : sb[u], folded ? 0 : ix[u], xa[u], xb[u], xc[u]);
      if (folded) xa[u] = 0.0;
      o2[u] = rd ? out[ix[u]] : 0.0;
      v[u] = gv[ok[u] ? pre[b] + c * nr + r : 0];
    }
#pragma unroll
    for (int u = 0; u < U2; u++)
      if (ok[u]) out[ix[u]] = finish(sb[u], o2[u] + hz * v[u], xa[u], xb[u], xc[u]);
  }
}

void launch_box2b(hipStream_t s, const DevTab &t, const Ops &o, double *out, const BTask *tasks, int ntasks,
                  const int *blist, int flags, const double *w1, const double *w2, const double *w3,
                  const double *L, const double *F, const double *D, double dsign, bool wide) {
  if (ntasks <= 0) return;
#define QD_BOX2B(NT_, VC_, RC_, ST_, SP_, LD_)                                                                  \
  hipLaunchKernelGGL((k_box2b<NT_, VC_, RC_, ST_, SP_, LD_>), dim3(ntasks), dim3(NT_), 0, s, t, o, out, tasks, blist, \
                     flags, w1, w2, w3, L, F, D, dsign)
#define QD_BOX2B_L(NT_, VC_, RC_, ST_, SP_)                     \
  do {                                                           \
    if (ldir) QD_BOX2B(NT_, VC_, RC_, ST_, SP_, true);           \
    else QD_BOX2B(NT_, VC_, RC_, ST_, SP_, false);               \
  } while (0)
  const bool stage = (flags & TB_DSTAGE) && D && dsign != 0.0;
  const bool ldir = (flags & TB_LDIRECT) != 0;
  if (wide) {
    if (stage) QD_BOX2B_L(BX_WIDE_THREADS, BX_WIDE_VCAP, BX_WIDE_RCAP, true, false);
    else QD_BOX2B_L(BX_WIDE_THREADS, BX_WIDE_VCAP, BX_WIDE_RCAP, false, false);
  } else {
    if (stage) QD_BOX2B_L(BX_THREADS, BX_VCAP, BX_RCAP, true, true);
    else QD_BOX2B_L(BX_THREADS, BX_VCAP, BX_RCAP, false, true);
  }
#undef QD_BOX2B_L
#undef QD_BOX2B
}

__global__ void __launch_bounds__(256) k_sum_halves(const double *__restrict__ w3, const double *__restrict__ w4,
                                                    double *__restrict__ L, const Seg2 *__restrict__ segs,
                                                    const Tile2 *__restrict__ tiles) {
  const Tile2 tl = tiles[blockIdx.x];
  const Seg2 sg = segs[tl.seg];
  const int e = tl.start + threadIdx.x;
  if (e >= sg.size) return;
  const int64_t k = sg.off + e;
  L[k] = w3[k] + w4[k];
}

void launch_sum_halves(hipStream_t s, const double *w3, const double *w4, double *L, const Seg2 *segs,
                       const Tile2 *tiles, int ntiles) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_sum_halves, dim3(ntiles), dim3(256), 0, s, w3, w4, L, segs, tiles);
}

// The square sub-blocks' ladder terms need the products at (J, I): one
// work-group per pair of 32x32 tiles (ij, ji).  With U = h1 A + c B and
// W = -h2 A - d B, L = U^T + W: only U goes through LDS (both tiles), W stays
// in the registers of the thread that loaded it, so the box pass reads one
// coalesced value per element instead of two scattered transposed ones.
// nz1: the nonzero sub-block types of r1 (ZB bits); a product of a zero type
// is known to be zero (the BCH terms) and is not read: w3 of HHHH and PPPP
// (r1 hhhh, pppp), w1 and w2 of PHPH (r1 ppph, phhh)
__global__ void __launch_bounds__(256) k_ladder_sq(DevTab t, const double *__restrict__ w1,
                                                   const double *__restrict__ w2, const double *__restrict__ w3,
                                                   double *__restrict__ L, const TilePair *__restrict__ tp, double h1,
                                                   double h2, uint32_t nz1) {
  __shared__ double Uij[32][33], Uji[32][33];
  const TilePair p = tp[blockIdx.x];
  const int q = p.blk / 3, kind = p.blk % 3;
  const int n = kind == 0 ? t.nh[q] : (kind == 1 ? t.nb[q] : t.np[q]);
  const int64_t off = t.off[6 * q + (kind == 0 ? HHHH : (kind == 1 ? PHPH : PPPP))];
  const double *__restrict__ A = w1 + off;
  const double *__restrict__ B = (kind == 1 ? w2 : w3) + off;
  double *__restrict__ Lq = L + off;
  const double c = kind == 1 ? h1 : h1 * h2, d = kind == 1 ? h2 : 1.0;
  const bool ua = kind != 1 || ((nz1 >> PPPH) & 1u);
  const bool ub = (nz1 >> (kind == 0 ? HHHH : (kind == 1 ? PHHH : PPPP))) & 1u;
  const int i0 = p.ti * 32, j0 = p.tj * 32;
  // all sixteen loads of a thread in flight together (unconditional, at an
  // in-bounds address; element 0 outside the block), then the LDS stores
  double ra[4][4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int e = threadIdx.x + 256 * u, li = e & 31, lj = e >> 5;
    const bool in1 = i0 + li < n && j0 + lj < n, in2 = j0 + li < n && i0 + lj < n;
    const int64_t a1 = in1 ? (i0 + li) + (int64_t)(j0 + lj) * n : 0;
    const int64_t a2 = in2 ? (j0 + li) + (int64_t)(i0 + lj) * n : 0;
    ra[u][0] = ua ? A[a1] : 0.0;
    ra[u][1] = ub ? B[a1] : 0.0;
    ra[u][2] = ua ? A[a2] : 0.0;
    ra[u][3] = ub ? B[a2] : 0.0;
  }
  double wij[4], wji[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int e = threadIdx.x + 256 * u, li = e & 31, lj = e >> 5;
    Uij[li][lj] = h1 * ra[u][0] + c * ra[u][1];
    Uji[li][lj] = h1 * ra[u][2] + c * ra[u][3];
    wij[u] = -h2 * ra[u][0] - d * ra[u][1];
    wji[u] = -h2 * ra[u][2] - d * ra[u][3];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int e = threadIdx.x + 256 * u, li = e & 31, lj = e >> 5;
    if (i0 + li < n && j0 + lj < n) Lq[(i0 + li) + (int64_t)(j0 + lj) * n] = Uji[lj][li] + wij[u];
    if (p.ti != p.tj && j0 + li < n && i0 + lj < n) Lq[(j0 + li) + (int64_t)(i0 + lj) * n] = Uij[lj][li] + wji[u];
  }
}

void launch_ladder_sq(hipStream_t s, const DevTab &t, const double *w1, const double *w2, const double *w3,
                      double *L, const TilePair *tp, int ntp, double h1, double h2, uint32_t nz1) {
  if (ntp <= 0) return;
  hipLaunchKernelGGL(k_ladder_sq, dim3(ntp), dim3(256), 0, s, t, w1, w2, w3, L, tp, h1, h2, nz1);
}

}  // namespace qd
