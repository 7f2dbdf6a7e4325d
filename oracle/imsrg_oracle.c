/*
 * CPU oracle for the IM-SRG(2) commutator / BCH hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the timed CPU baseline.  The product (libqdimsrg.so) never links or calls it.
 *
 * A plain-C FP64 restatement of the reference's algorithm, each function
 * citing the Fortran it follows.  Operators are flat vectors in the
 * reference's ``vectorize`` order (src/ME_general.f90:1871-1989) with an
 * integer hermiticity flag; cross-coupled (CC) matrices are flat per-block
 * column-major r x nb arrays with a real hermiticity flag.  Indices are
 * 0-based; holes are states 0..n-1 (convertF, src/ME_general.f90:1527-1536).
 *
 * Summation domains: the reference sums over every single-particle index and
 * relies on v_elem returning exactly 0 across (M_L, M_S) blocks
 * (src/ME_general.f90:1610-1612).  Here those sums are restricted to the
 * indices whose terms can be non-zero; skipped terms are exact zeros, so the
 * restriction changes no value.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

typedef struct {
  i64 n, M, mp, nblock, neq;
  const i64 *nh, *np, *nb, *off;           /* off[q*6 + s], s in vectorize order */
  const i64 *qnhh, *qnpp, *qnph;           /* concatenated (i,j) pair lists      */
  const i64 *hh_start, *pp_start, *ph_start;
  const i64 *iarr;                         /* [pair*3] = (type 1..3, block, idx) */
  i64 ncc, cc_size;
  const i64 *cc_r, *cc_nb, *cc_off, *cc_ab, *cc_hp, *cc_ab_start, *cc_hp_start;
  const i64 *cc_map;                       /* [(M*i+j)*2] = (block, row)         */
  const i64 *ml, *ms;
  i64 off_fhh, off_fpp, off_fph;
} tab_t;

enum { HHHH = 0, PPPP, PPHH, PHPH, PPPH, PHHH };

#define FHH(v, a, b) (v)[t->off_fhh + (a) + (b) * t->n]
#define FPP(v, a, b) (v)[t->off_fpp + (a) + (b) * t->mp]
#define FPH(v, a, b) (v)[t->off_fph + (a) + (b) * t->mp]

static inline i64 pidx(const tab_t *t, i64 i, i64 j) {
  return i * t->M - i * (i + 1) / 2 + (j - i - 1);
}

/* v_elem, src/ME_general.f90:1557-1680 */
double o_velem(const tab_t *t, const double *v, int herm, i64 a1, i64 a2, i64 a3, i64 a4) {
  if (a1 == a2 || a3 == a4) return 0.0;
  int pre = 1;
  i64 x;
  if (a1 > a2) { x = a1; a1 = a2; a2 = x; pre = -1; }
  if (a3 > a4) { x = a3; a3 = a4; a4 = x; pre = -pre; }
  const i64 *I = t->iarr + 3 * pidx(t, a1, a2);
  const i64 *J = t->iarr + 3 * pidx(t, a3, a4);
  if (I[1] != J[1]) return 0.0;
  i64 q = I[1], w1 = I[2], w2 = J[2];
  const i64 *o = t->off + 6 * q;
  i64 nh = t->nh[q], np = t->np[q], nb = t->nb[q];
  switch (I[0] * 4 + J[0]) {
    case 1 * 4 + 1: return v[o[HHHH] + w1 + w2 * nh] * pre;
    case 1 * 4 + 2: return v[o[PHHH] + w2 + w1 * nb] * pre * herm;
    case 1 * 4 + 3: return v[o[PPHH] + w2 + w1 * np] * pre * herm;
    case 2 * 4 + 1: return v[o[PHHH] + w1 + w2 * nb] * pre;
    case 2 * 4 + 2: return v[o[PHPH] + w1 + w2 * nb] * pre;
    case 2 * 4 + 3: return v[o[PPPH] + w2 + w1 * np] * pre * herm;
    case 3 * 4 + 1: return v[o[PPHH] + w1 + w2 * np] * pre;
    case 3 * 4 + 2: return v[o[PPPH] + w1 + w2 * np] * pre;
    default:        return v[o[PPPP] + w1 + w2 * np] * pre;
  }
}

/* f_elem, src/ME_general.f90:1780-1813 */
double o_felem(const tab_t *t, const double *v, int herm, i64 a, i64 b) {
  i64 n = t->n;
  if (a < n) return (b < n) ? FHH(v, a, b) : FPH(v, b - n, a) * herm;
  return (b < n) ? FPH(v, a - n, b) : FPP(v, a - n, b - n);
}

static inline int same_ch(const tab_t *t, i64 a, i64 b) {
  return t->ml[a] == t->ml[b] && t->ms[a] == t->ms[b];
}

/* column-major C(m,n) = alpha*op(A)*op(B) + beta*C, the dgemm contract.
 * Every element is the plain sum over l = 0..k-1 in order (no fused
 * multiply-add: ISO C, contraction off), computed for 8 x 4 elements at a
 * time so that the compiler vectorises across elements. */
#define GEMM_MI 8
#define GEMM_NJ 4
typedef double v4d __attribute__((vector_size(32)));
static void gemm(char ta, char tb, i64 m, i64 n, i64 k, double alpha, const double *A, i64 lda,
                 const double *B, i64 ldb, double beta, double *C, i64 ldc) {
  if (m <= 0 || n <= 0) return;
  const i64 kk = k > 0 ? k : 1;
  const i64 mp = (m + GEMM_MI - 1) / GEMM_MI * GEMM_MI;  /* rows padded with zeros */
  /* a: panels of GEMM_MI rows, a[(i/MI)*MI*k + l*MI + i%MI] = op(A)(i,l) */
  double *a = (double *)calloc((size_t)(mp * kk), sizeof(double));
  double *b = (double *)malloc(sizeof(double) * (size_t)(n * kk));  /* b[j*k + l] = op(B)(l,j) */
  for (i64 l = 0; l < k; l++)
    for (i64 i = 0; i < m; i++)
      a[(i / GEMM_MI) * GEMM_MI * kk + l * GEMM_MI + i % GEMM_MI] = (ta == 'N') ? A[i + l * lda] : A[l + i * lda];
  for (i64 j = 0; j < n; j++)
    for (i64 l = 0; l < k; l++) b[j * k + l] = (tb == 'N') ? B[l + j * ldb] : B[j + l * ldb];
  for (i64 j0 = 0; j0 < n; j0 += GEMM_NJ) {
    const i64 nj = n - j0 < GEMM_NJ ? n - j0 : GEMM_NJ;
    const double *b0 = b + j0 * k, *b1 = b0 + (nj > 1 ? k : 0), *b2 = b0 + (nj > 2 ? 2 * k : 0),
                 *b3 = b0 + (nj > 3 ? 3 * k : 0);
    for (i64 i0 = 0; i0 < m; i0 += GEMM_MI) {
      /* eight rows as two 4-wide vectors per column: accumulators stay in
       * registers; separate multiply and add, as the scalar sum does */
      v4d s0a = {0}, s0b = {0}, s1a = {0}, s1b = {0}, s2a = {0}, s2b = {0}, s3a = {0}, s3b = {0};
      const double *ap = a + i0 * kk;
      for (i64 l = 0; l < k; l++) {
        v4d lo, hi;
        memcpy(&lo, ap + l * GEMM_MI, sizeof lo);
        memcpy(&hi, ap + l * GEMM_MI + 4, sizeof hi);
        const v4d x0 = {b0[l], b0[l], b0[l], b0[l]}, x1 = {b1[l], b1[l], b1[l], b1[l]};
        const v4d x2 = {b2[l], b2[l], b2[l], b2[l]}, x3 = {b3[l], b3[l], b3[l], b3[l]};
        v4d p;
        p = lo * x0; s0a = s0a + p; p = hi * x0; s0b = s0b + p;
        p = lo * x1; s1a = s1a + p; p = hi * x1; s1b = s1b + p;
        p = lo * x2; s2a = s2a + p; p = hi * x2; s2b = s2b + p;
        p = lo * x3; s3a = s3a + p; p = hi * x3; s3b = s3b + p;
      }
      double s0[GEMM_MI], s1[GEMM_MI], s2[GEMM_MI], s3[GEMM_MI];
      memcpy(s0, &s0a, 32); memcpy(s0 + 4, &s0b, 32);
      memcpy(s1, &s1a, 32); memcpy(s1 + 4, &s1b, 32);
      memcpy(s2, &s2a, 32); memcpy(s2 + 4, &s2b, 32);
      memcpy(s3, &s3a, 32); memcpy(s3 + 4, &s3b, 32);
      const double *sv[GEMM_NJ] = {s0, s1, s2, s3};
      for (i64 jj = 0; jj < nj; jj++)
        for (i64 ii = 0; ii < GEMM_MI && i0 + ii < m; ii++) {
          double *cij = &C[(i0 + ii) + (j0 + jj) * ldc];
          const double c = (beta == 0.0) ? 0.0 : beta * *cij;
          *cij = c + alpha * sv[jj][ii];
        }
    }
  }
  free(a);
  free(b);
}

/* commutator_110, src/commutators.f90:12-33 */
double o_comm110(const tab_t *t, const double *v1, int h1, const double *v2, int h2) {
  double c = 0.0;
  for (i64 i = 0; i < t->mp; i++) {
    double q = 0.0;
    for (i64 k = 0; k < t->n; k++) q += FPH(v1, i, k) * FPH(v2, i, k);
    c += q;
  }
  return (h1 - h2) * c;
}

/* commutator_220, src/commutators.f90:36-70 */
double o_comm220(const tab_t *t, const double *v1, int h1, const double *v2, int h2) {
  double c = 0.0;
  for (i64 q = 0; q < t->nblock; q++) {
    i64 np = t->np[q], nh = t->nh[q];
    if (np * nh == 0) continue;
    const double *a = v1 + t->off[6 * q + PPHH], *b = v2 + t->off[6 * q + PPHH];
    for (i64 i = 0; i < np; i++) {
      double s = 0.0;
      for (i64 k = 0; k < nh; k++) s += a[i + k * np] * b[i + k * np];
      c += s * (h1 - h2);
    }
  }
  return c;
}

/* commutator_111, src/commutators.f90:137-195 (overwrites the one-body part) */
void o_comm111(const tab_t *t, const double *v1, int h1, const double *v2, int h2, double *v3) {
  i64 n = t->n, m = t->mp;
  double *Q1 = malloc(sizeof(double) * (size_t)(m * m + 1));
  double *Q2 = malloc(sizeof(double) * (size_t)(m * m + 1));
  double *Q3 = malloc(sizeof(double) * (size_t)(m * n + 1));
  double *Q4 = malloc(sizeof(double) * (size_t)(m * n + 1));
  /* fhh */
  gemm('N', 'N', n, n, n, 1.0, &FHH(v1, 0, 0), n, &FHH(v2, 0, 0), n, 0.0, Q1, n);
  gemm('T', 'N', n, n, m, 1.0, &FPH(v2, 0, 0), m, &FPH(v1, 0, 0), m, 0.0, Q2, n);
  for (i64 j = 0; j < n; j++)
    for (i64 i = 0; i < n; i++)
      FHH(v3, i, j) = Q1[i + j * n] - h2 * Q2[i + j * n] + h1 * (Q2[j + i * n] - h2 * Q1[j + i * n]);
  /* fpp */
  gemm('N', 'N', m, m, m, 1.0, &FPP(v2, 0, 0), m, &FPP(v1, 0, 0), m, 0.0, Q1, m);
  gemm('N', 'T', m, m, n, 1.0, &FPH(v1, 0, 0), m, &FPH(v2, 0, 0), m, 0.0, Q2, m);
  for (i64 j = 0; j < m; j++)
    for (i64 i = 0; i < m; i++)
      FPP(v3, i, j) = h2 * Q2[i + j * m] + h1 * (h2 * Q1[j + i * m] - Q2[j + i * m]) - Q1[i + j * m];
  /* fph */
  gemm('N', 'N', m, n, n, 1.0, &FPH(v1, 0, 0), m, &FHH(v2, 0, 0), n, 0.0, Q1, m);
  gemm('N', 'N', m, n, n, 1.0, &FPH(v2, 0, 0), m, &FHH(v1, 0, 0), n, 0.0, Q2, m);
  gemm('N', 'N', m, n, m, 1.0, &FPP(v1, 0, 0), m, &FPH(v2, 0, 0), m, 0.0, Q3, m);
  gemm('N', 'N', m, n, m, 1.0, &FPP(v2, 0, 0), m, &FPH(v1, 0, 0), m, 0.0, Q4, m);
  for (i64 j = 0; j < n; j++)
    for (i64 i = 0; i < m; i++)
      FPH(v3, i, j) = Q1[i + j * m] - Q2[i + j * m] + Q3[i + j * m] - Q4[i + j * m];
  free(Q1); free(Q2); free(Q3); free(Q4);
}

/* commutator_121, src/commutators.f90:73-134 (adds to the one-body part).
 * The pair (i,a) contributes only when (a,x) and (i,y) or (i,x) and (a,y)
 * share a block, i.e. ml_a - ml_i = +-(ml_y - ml_x) (same for m_s). */
static double c121_elem(const tab_t *t, const double *v1, int h1, const double *v2, int h2,
                        i64 x, i64 y) {
  i64 n = t->n;
  double r = 0.0;
  i64 dl = t->ml[y] - t->ml[x], ds = t->ms[y] - t->ms[x];
  for (i64 i = 0; i < n; i++)
    for (i64 a = n; a < t->M; a++) {
      i64 el = t->ml[a] - t->ml[i], es = t->ms[a] - t->ms[i];
      if (!((el == dl && es == ds) || (el == -dl && es == -ds))) continue;
      r = r + FPH(v1, a - n, i) * (h1 * o_velem(t, v2, h2, a, x, i, y) - o_velem(t, v2, h2, i, x, a, y))
            - FPH(v2, a - n, i) * (h2 * o_velem(t, v1, h1, a, x, i, y) - o_velem(t, v1, h1, i, x, a, y));
    }
  return r;
}

void o_comm121(const tab_t *t, const double *v1, int h1, const double *v2, int h2, double *v3) {
  i64 n = t->n, m = t->mp;
#pragma omp parallel for schedule(dynamic)
  for (i64 h2i = 0; h2i < n; h2i++) {
    for (i64 h1i = 0; h1i < n; h1i++) FHH(v3, h1i, h2i) += c121_elem(t, v1, h1, v2, h2, h1i, h2i);
    for (i64 p1 = 0; p1 < m; p1++) FPH(v3, p1, h2i) += c121_elem(t, v1, h1, v2, h2, p1 + n, h2i);
  }
#pragma omp parallel for schedule(dynamic)
  for (i64 p2 = 0; p2 < m; p2++)
    for (i64 p1 = 0; p1 < m; p1++) FPP(v3, p1, p2) += c121_elem(t, v1, h1, v2, h2, p1 + n, p2 + n);
}

/* commutator_221, src/commutators.f90:198-352 (adds to the one-body part).
 * w1, w2 are scratch operators with herm = +1 (allocate_everything). */
void o_comm221(const tab_t *t, const double *v1, int h1, const double *v2, int h2, double *v3) {
  i64 n = t->n, m = t->mp, M = t->M;
  double *w1 = calloc((size_t)t->neq, sizeof(double));
  double *w2 = calloc((size_t)t->neq, sizeof(double));
#pragma omp parallel for schedule(dynamic)
  for (i64 w = 0; w < t->nblock; w++) {
    i64 nh = t->nh[w], np = t->np[w], nb = t->nb[w];
    const i64 *o = t->off + 6 * w;
    if (np * nh)
      gemm('N', 'T', np, np, nh, 1.0, v1 + o[PPHH], np, v2 + o[PPHH], np, 0.0, w1 + o[PPPP], np);
    if (np * nb)
      gemm('T', 'N', nb, nb, np, 1.0, v2 + o[PPPH], np, v1 + o[PPPH], np, 0.0, w1 + o[PHPH], nb);
    if (nh * np)
      gemm('T', 'N', nh, nh, np, 1.0, v2 + o[PPHH], np, v1 + o[PPHH], np, 0.0, w1 + o[HHHH], nh);
    if (nh * nb)
      gemm('N', 'T', nb, nb, nh, 1.0, v1 + o[PHHH], nb, v2 + o[PHHH], nb, 0.0, w2 + o[PHPH], nb);
    if (nh * nb * np) {
      gemm('T', 'N', nb, nh, np, 1.0, v2 + o[PPPH], np, v1 + o[PPHH], np, 0.0, w1 + o[PHHH], nb);
      gemm('T', 'N', nb, nh, np, 1.0, v1 + o[PPPH], np, v2 + o[PPHH], np, 0.0, w2 + o[PHHH], nb);
      gemm('N', 'T', np, nb, nh, 1.0, v1 + o[PPHH], np, v2 + o[PHHH], nb, 0.0, w1 + o[PPPH], np);
      gemm('N', 'T', np, nb, nh, 1.0, v2 + o[PPHH], np, v1 + o[PHHH], nb, 0.0, w2 + o[PPPH], np);
    }
  }
  /* fpp: only same-channel (p,p2) pairs can be non-zero */
#pragma omp parallel for schedule(dynamic)
  for (i64 p = 0; p < m; p++)
    for (i64 p2 = 0; p2 < m; p2++) {
      if (!same_ch(t, p + n, p2 + n)) continue;
      double r = FPP(v3, p, p2);
      for (i64 i = 0; i < n; i++)
        r = r + h1 * o_velem(t, w1, 1, i, p2 + n, i, p + n) - h2 * o_velem(t, w1, 1, i, p + n, i, p2 + n);
      for (i64 a = n; a < M; a++)
        r = r - h1 * o_velem(t, w1, 1, a, p2 + n, a, p + n) + h2 * o_velem(t, w1, 1, a, p + n, a, p2 + n);
      FPP(v3, p, p2) = r;
    }
  /* fhh */
  for (i64 h = 0; h < n; h++)
    for (i64 hb = 0; hb < n; hb++) {
      if (!same_ch(t, h, hb)) continue;
      double r = FHH(v3, h, hb);
      for (i64 i = 0; i < n; i++)
        r = r + h1 * o_velem(t, w1, 1, i, hb, i, h) - h2 * o_velem(t, w1, 1, i, h, i, hb);
      for (i64 a = n; a < M; a++)
        r = r - h1 * o_velem(t, w2, 1, a, hb, a, h) + h2 * o_velem(t, w2, 1, a, h, a, hb);
      FHH(v3, h, hb) = r;
    }
  /* fph */
#pragma omp parallel for schedule(dynamic)
  for (i64 p = 0; p < m; p++)
    for (i64 h = 0; h < n; h++) {
      if (!same_ch(t, p + n, h)) continue;
      double r = FPH(v3, p, h);
      for (i64 i = 0; i < n; i++)
        r = r + h1 * o_velem(t, w2, 1, i, p + n, i, h) - h2 * o_velem(t, w1, 1, i, p + n, i, h);
      for (i64 a = n; a < M; a++)
        r = r - h1 * o_velem(t, w2, 1, a, p + n, a, h) + h2 * o_velem(t, w1, 1, a, p + n, a, h);
      FPH(v3, p, h) = r;
    }
  free(w1);
  free(w2);
}

/* One element of commutator_122, src/commutators.f90:355-676, for the bra
 * (x1,x2) and ket (x3,x4) given in the order the reference feeds v_elem
 * (particle first for ph pairs, src/commutators.f90:466-471).  f1/f2 = f_elem
 * of r1/r2 with the reference's herm factors: the reference writes
 *   A(x1,t)V(t x2|x3 x4) + A(x2,t)V(x1 t|x3 x4)
 *   - h_A [A(x3,t)V(x1 x2|t x4) + A(x4,t)V(x1 x2|x3 t)]
 * for hole t (fhh / fph(x-n,t) rows) and
 *   h_A [A(t,x1)V(t x2|..) + A(t,x2)V(x1 t|..)] - [A(t,x3)V(..|t x4) + A(t,x4)V(..|x3 t)]
 * for particle t (fph(t,x) / fpp(t,x) columns), minus the same with r1 <-> r2.
 * In f_elem language the hole terms read A(x,t) and the particle terms read
 * A(t,x) with t a particle, i.e. h_A f_elem(x,t) for a hole x. */
static inline double fcol(const tab_t *t, const double *v, i64 tt, i64 x) {
  /* reference reads fph(t-n, x) for hole x and fpp(t-n, x-n) for particle x */
  i64 n = t->n;
  return (x < n) ? FPH(v, tt - n, x) : FPP(v, tt - n, x - n);
}
static inline double frow(const tab_t *t, const double *v, i64 x, i64 tt) {
  /* reference reads fhh(x,t) for hole x and fph(x-n,t) for particle x */
  i64 n = t->n;
  return (x < n) ? FHH(v, x, tt) : FPH(v, x - n, tt);
}

/* The states t that can contribute to an element: those in the (m_l, m_s)
 * channel of one of x1..x4 (f conserves the channel; every other t has all
 * four terms zero).  chan[k] lists the states of a channel in ascending
 * order, cstart/clen give each state's run.  The candidates are visited in
 * ascending order, as a loop over all t would visit them. */
typedef struct {
  i64 *chan, *cstart, *clen;
} chan_t;

static chan_t make_chan(const tab_t *t) {
  chan_t c;
  i64 M = t->M;
  c.chan = malloc(sizeof(i64) * (size_t)M);
  c.cstart = malloc(sizeof(i64) * (size_t)M);
  c.clen = malloc(sizeof(i64) * (size_t)M);
  i64 k = 0;
  char *done = calloc((size_t)M, 1);
  for (i64 a = 0; a < M; a++) {
    if (done[a]) continue;
    i64 s0 = k;
    for (i64 b = a; b < M; b++)
      if (t->ml[b] == t->ml[a] && t->ms[b] == t->ms[a]) { c.chan[k++] = b; done[b] = 1; }
    for (i64 j = s0; j < k; j++) { c.cstart[c.chan[j]] = s0; c.clen[c.chan[j]] = k - s0; }
  }
  free(done);
  return c;
}

static void free_chan(chan_t c) { free(c.chan); free(c.cstart); free(c.clen); }

/* sorted, distinct candidates t for the element (x1, x2, x3, x4); returns the count */
static int candidates(const chan_t *c, i64 x1, i64 x2, i64 x3, i64 x4, i64 *out) {
  const i64 xs[4] = {x1, x2, x3, x4};
  int n = 0;
  for (int q = 0; q < 4; q++) {
    int seen = 0;
    for (int r = 0; r < q; r++)
      if (c->cstart[xs[r]] == c->cstart[xs[q]]) seen = 1;
    if (seen) continue;
    for (i64 j = 0; j < c->clen[xs[q]]; j++) out[n++] = c->chan[c->cstart[xs[q]] + j];
  }
  for (int i = 1; i < n; i++) {  /* insertion sort: a few dozen entries */
    i64 v = out[i];
    int j = i - 1;
    while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; j--; }
    out[j + 1] = v;
  }
  return n;
}

static double c122_half(const tab_t *t, const chan_t *c, const double *fa, int ha, const double *vb, int hb,
                        i64 x1, i64 x2, i64 x3, i64 x4) {
  i64 n = t->n;
  i64 cand[4 * 64];
  int nc = candidates(c, x1, x2, x3, x4, cand);
  double r = 0.0;
  /* hole sums: t in the (m_l, m_s) channel of the replaced index */
  for (int k = 0; k < nc && cand[k] < n; k++) {
    i64 tt = cand[k];
    double s = 0.0;
    if (same_ch(t, tt, x1)) s += frow(t, fa, x1, tt) * o_velem(t, vb, hb, tt, x2, x3, x4);
    if (same_ch(t, tt, x2)) s += frow(t, fa, x2, tt) * o_velem(t, vb, hb, x1, tt, x3, x4);
    double u = 0.0;
    if (same_ch(t, tt, x3)) u += frow(t, fa, x3, tt) * o_velem(t, vb, hb, x1, x2, tt, x4);
    if (same_ch(t, tt, x4)) u += frow(t, fa, x4, tt) * o_velem(t, vb, hb, x1, x2, x3, tt);
    r += s - ha * u;
  }
  for (int k = 0; k < nc; k++) {
    i64 tt = cand[k];
    if (tt < n) continue;
    double s = 0.0;
    if (same_ch(t, tt, x1)) s += fcol(t, fa, tt, x1) * o_velem(t, vb, hb, tt, x2, x3, x4);
    if (same_ch(t, tt, x2)) s += fcol(t, fa, tt, x2) * o_velem(t, vb, hb, x1, tt, x3, x4);
    double u = 0.0;
    if (same_ch(t, tt, x3)) u += fcol(t, fa, tt, x3) * o_velem(t, vb, hb, x1, x2, tt, x4);
    if (same_ch(t, tt, x4)) u += fcol(t, fa, tt, x4) * o_velem(t, vb, hb, x1, x2, x3, tt);
    r += ha * s - u;
  }
  return r;
}

static inline double c122_elem(const tab_t *t, const chan_t *c, const double *v1, int h1, const double *v2,
                               int h2, i64 x1, i64 x2, i64 x3, i64 x4) {
  return c122_half(t, c, v1, h1, v2, h2, x1, x2, x3, x4) - c122_half(t, c, v2, h2, v1, h1, x1, x2, x3, x4);
}

void o_comm122(const tab_t *t, const double *v1, int h1, const double *v2, int h2, double *v3) {
  /* blocks in turn, the columns of each sub-block in parallel (the largest
   * blocks hold most of the elements) */
  chan_t cs = make_chan(t);
  const chan_t *c = &cs;
  for (i64 q = 0; q < t->nblock; q++) {
    i64 nh = t->nh[q], np = t->np[q], nb = t->nb[q];
    const i64 *o = t->off + 6 * q;
    const i64 *HH = t->qnhh + 2 * t->hh_start[q], *PP = t->qnpp + 2 * t->pp_start[q],
              *PH = t->qnph + 2 * t->ph_start[q];
#pragma omp parallel for schedule(dynamic)
    for (i64 J = 0; J < nh; J++)
      for (i64 I = 0; I < nh; I++)
        v3[o[HHHH] + I + J * nh] = c122_elem(t, c, v1, h1, v2, h2, HH[2 * I], HH[2 * I + 1], HH[2 * J], HH[2 * J + 1]);
#pragma omp parallel for schedule(dynamic)
    for (i64 J = 0; J < np; J++)
      for (i64 I = 0; I < np; I++)
        v3[o[PPPP] + I + J * np] = c122_elem(t, c, v1, h1, v2, h2, PP[2 * I], PP[2 * I + 1], PP[2 * J], PP[2 * J + 1]);
#pragma omp parallel for schedule(dynamic)
    for (i64 J = 0; J < nh; J++)
      for (i64 I = 0; I < np; I++)
        v3[o[PPHH] + I + J * np] = c122_elem(t, c, v1, h1, v2, h2, PP[2 * I], PP[2 * I + 1], HH[2 * J], HH[2 * J + 1]);
    /* ph pairs: particle first, pre = -1 (src/commutators.f90:466-471) */
#pragma omp parallel for schedule(dynamic)
    for (i64 J = 0; J < nh; J++)
      for (i64 I = 0; I < nb; I++)
        v3[o[PHHH] + I + J * nb] = -c122_elem(t, c, v1, h1, v2, h2, PH[2 * I + 1], PH[2 * I], HH[2 * J], HH[2 * J + 1]);
#pragma omp parallel for schedule(dynamic)
    for (i64 J = 0; J < nb; J++)
      for (i64 I = 0; I < np; I++)
        v3[o[PPPH] + I + J * np] = -c122_elem(t, c, v1, h1, v2, h2, PP[2 * I], PP[2 * I + 1], PH[2 * J + 1], PH[2 * J]);
#pragma omp parallel for schedule(dynamic)
    for (i64 J = 0; J < nb; J++)
      for (i64 I = 0; I < nb; I++)
        v3[o[PHPH] + I + J * nb] = c122_elem(t, c, v1, h1, v2, h2, PH[2 * I + 1], PH[2 * I], PH[2 * J + 1], PH[2 * J]);
  }
  free_chan(cs);
}

/* construct_cc_ints, src/ME_general.f90:923-952:
 *   XCC = HCC%herm * ETACC HCC^T,  YCC = ETACC%herm * HCC ETACC^T   (r x r) */
static void cc_ints(const tab_t *t, const double *hcc, double hh, const double *ecc, double eh,
                    double **X, double **Y) {
  for (i64 q = 0; q < t->ncc; q++) {
    i64 r = t->cc_r[q], nb = t->cc_nb[q];
    X[q] = calloc((size_t)(r * r + 1), sizeof(double));
    Y[q] = calloc((size_t)(r * r + 1), sizeof(double));
  }
#pragma omp parallel for schedule(dynamic)
  for (i64 q = 0; q < t->ncc; q++) {
    i64 r = t->cc_r[q], nb = t->cc_nb[q];
    if (r == 0 || nb == 0) continue;
    gemm('N', 'T', r, r, nb, hh, ecc + t->cc_off[q], r, hcc + t->cc_off[q], r, 0.0, X[q], r);
    gemm('N', 'T', r, r, nb, eh, hcc + t->cc_off[q], r, ecc + t->cc_off[q], r, 0.0, Y[q], r);
  }
}

/* Pandya ph term for output <x1 x2|x3 x4>, src/commutators.f90:716-751 */
static inline double pandya(const tab_t *t, double **X, double **Y, i64 x1, i64 x2, i64 x3, i64 x4) {
  i64 M = t->M;
#define CQ(a, b) t->cc_map[2 * (M * (a) + (b))]
#define CR(a, b) t->cc_map[2 * (M * (a) + (b)) + 1]
#define XE(Z, a, b, c, d) Z[CQ(a, b)][CR(a, b) + CR(c, d) * t->cc_r[CQ(a, b)]]
  double x = XE(X, x3, x1, x2, x4) - XE(X, x3, x2, x1, x4) + XE(X, x4, x2, x1, x3) - XE(X, x4, x1, x2, x3);
  x = x - XE(Y, x4, x2, x1, x3) + XE(Y, x4, x1, x2, x3) - XE(Y, x3, x1, x2, x4) + XE(Y, x3, x2, x1, x4);
  return x;
#undef CQ
#undef CR
#undef XE
}

/* commutator_222 + matrix_mult222, src/commutators.f90:679-1132 (adds to 2b).
 * hcc = CC of r2 (with its own real herm), ecc = CC of r1. */
void o_comm222(const tab_t *t, const double *v1, int h1, const double *v2, int h2, double *v3,
               const double *hcc, double hh, const double *ecc, double eh) {
  double **X = malloc(sizeof(double *) * (size_t)t->ncc), **Y = malloc(sizeof(double *) * (size_t)t->ncc);
  cc_ints(t, hcc, hh, ecc, eh, X, Y);
#pragma omp parallel for schedule(dynamic)
  for (i64 q = 0; q < t->nblock; q++) {
    i64 nh = t->nh[q], np = t->np[q], nb = t->nb[q];
    const i64 *o = t->off + 6 * q;
    i64 mx = np > nh ? np : nh;
    mx = mx > nb ? mx : nb;
    double *w = calloc((size_t)(mx * mx + 1), sizeof(double));
    const double *A, *B;
    double *C;
#define V1(s) (v1 + o[s])
#define V2(s) (v2 + o[s])
#define V3(s) (v3 + o[s])
    if (nh * np > 0) {
      gemm('T', 'N', nh, nh, np, 1.0, V2(PPHH), np, V1(PPHH), np, 0.0, w, nh);
      C = V3(HHHH);
      for (i64 j = 0; j < nh; j++)
        for (i64 i = 0; i < nh; i++) C[i + j * nh] += h1 * w[j + i * nh] - h2 * w[i + j * nh];
      gemm('N', 'T', np, np, nh, 1.0, V1(PPHH), np, V2(PPHH), np, 0.0, w, np);
      C = V3(PPPP);
      for (i64 j = 0; j < np; j++)
        for (i64 i = 0; i < np; i++) C[i + j * np] += h1 * w[j + i * np] - h2 * w[i + j * np];
      C = V3(PPHH);
      gemm('N', 'N', np, nh, np, 1.0, V1(PPPP), np, V2(PPHH), np, 0.0, w, np);
      for (i64 e = 0; e < np * nh; e++) C[e] += w[e];
      gemm('N', 'N', np, nh, np, 1.0, V2(PPPP), np, V1(PPHH), np, 0.0, w, np);
      for (i64 e = 0; e < np * nh; e++) C[e] -= w[e];
      gemm('N', 'N', np, nh, nh, 1.0, V1(PPHH), np, V2(HHHH), nh, 0.0, w, np);
      for (i64 e = 0; e < np * nh; e++) C[e] -= w[e];
      gemm('N', 'N', np, nh, nh, 1.0, V2(PPHH), np, V1(HHHH), nh, 0.0, w, np);
      for (i64 e = 0; e < np * nh; e++) C[e] += w[e];
    }
    if (nh > 0) {
      gemm('N', 'N', nh, nh, nh, 1.0, V1(HHHH), nh, V2(HHHH), nh, 0.0, w, nh);
      C = V3(HHHH);
      for (i64 j = 0; j < nh; j++)
        for (i64 i = 0; i < nh; i++) C[i + j * nh] += h1 * h2 * w[j + i * nh] - w[i + j * nh];
    }
    if (np > 0) {
      gemm('N', 'N', np, np, np, 1.0, V2(PPPP), np, V1(PPPP), np, 0.0, w, np);
      C = V3(PPPP);
      for (i64 j = 0; j < np; j++)
        for (i64 i = 0; i < np; i++) C[i + j * np] += h1 * h2 * w[j + i * np] - w[i + j * np];
    }
    if (nb * np * nh > 0) {
      C = V3(PHHH);
      gemm('T', 'N', nb, nh, np, 1.0, V1(PPPH), np, V2(PPHH), np, 0.0, w, nb);
      for (i64 e = 0; e < nb * nh; e++) C[e] += h1 * w[e];
      gemm('T', 'N', nb, nh, np, 1.0, V2(PPPH), np, V1(PPHH), np, 0.0, w, nb);
      for (i64 e = 0; e < nb * nh; e++) C[e] -= h2 * w[e];
      C = V3(PPPH);
      gemm('N', 'T', np, nb, nh, 1.0, V1(PPHH), np, V2(PHHH), nb, 0.0, w, np);
      for (i64 e = 0; e < np * nb; e++) C[e] -= h2 * w[e];
      gemm('N', 'T', np, nb, nh, 1.0, V2(PPHH), np, V1(PHHH), nb, 0.0, w, np);
      for (i64 e = 0; e < np * nb; e++) C[e] += h1 * w[e];
    }
    if (nb * nh > 0) {
      C = V3(PHHH);
      gemm('N', 'N', nb, nh, nh, 1.0, V1(PHHH), nb, V2(HHHH), nh, 0.0, w, nb);
      for (i64 e = 0; e < nb * nh; e++) C[e] -= w[e];
      gemm('N', 'N', nb, nh, nh, 1.0, V2(PHHH), nb, V1(HHHH), nh, 0.0, w, nb);
      for (i64 e = 0; e < nb * nh; e++) C[e] += w[e];
      C = V3(PHPH);
      gemm('N', 'T', nb, nb, nh, 1.0, V1(PHHH), nb, V2(PHHH), nb, 0.0, w, nb);
      for (i64 j = 0; j < nb; j++)
        for (i64 i = 0; i < nb; i++) C[i + j * nb] += h1 * w[j + i * nb] - h2 * w[i + j * nb];
    }
    if (nb * np > 0) {
      C = V3(PPPH);
      gemm('N', 'N', np, nb, np, 1.0, V1(PPPP), np, V2(PPPH), np, 0.0, w, np);
      for (i64 e = 0; e < np * nb; e++) C[e] += w[e];
      gemm('N', 'N', np, nb, np, 1.0, V2(PPPP), np, V1(PPPH), np, 0.0, w, np);
      for (i64 e = 0; e < np * nb; e++) C[e] -= w[e];
      C = V3(PHPH);
      gemm('T', 'N', nb, nb, np, 1.0, V2(PPPH), np, V1(PPPH), np, 0.0, w, nb);
      for (i64 j = 0; j < nb; j++)
        for (i64 i = 0; i < nb; i++) C[i + j * nb] += h1 * w[j + i * nb] - h2 * w[i + j * nb];
    }
    free(w);
    /* Pandya (particle-hole) part: ph-type bras/kets are fed particle first
     * and multiplied by pre = -1 (src/commutators.f90:846-851, 892-897) */
    const i64 *HH = t->qnhh + 2 * t->hh_start[q], *PP = t->qnpp + 2 * t->pp_start[q],
              *PH = t->qnph + 2 * t->ph_start[q];
    for (i64 J = 0; J < nh; J++)
      for (i64 I = 0; I < nh; I++)
        V3(HHHH)[I + J * nh] += pandya(t, X, Y, HH[2 * I], HH[2 * I + 1], HH[2 * J], HH[2 * J + 1]);
    for (i64 J = 0; J < np; J++)
      for (i64 I = 0; I < np; I++)
        V3(PPPP)[I + J * np] += pandya(t, X, Y, PP[2 * I], PP[2 * I + 1], PP[2 * J], PP[2 * J + 1]);
    for (i64 J = 0; J < nh; J++)
      for (i64 I = 0; I < np; I++)
        V3(PPHH)[I + J * np] += pandya(t, X, Y, PP[2 * I], PP[2 * I + 1], HH[2 * J], HH[2 * J + 1]);
    for (i64 J = 0; J < nb; J++)
      for (i64 I = 0; I < np; I++)
        V3(PPPH)[I + J * np] += -pandya(t, X, Y, PP[2 * I], PP[2 * I + 1], PH[2 * J + 1], PH[2 * J]);
    for (i64 J = 0; J < nh; J++)
      for (i64 I = 0; I < nb; I++)
        V3(PHHH)[I + J * nb] += -pandya(t, X, Y, PH[2 * I + 1], PH[2 * I], HH[2 * J], HH[2 * J + 1]);
    for (i64 J = 0; J < nb; J++)
      for (i64 I = 0; I < nb; I++)
        V3(PHPH)[I + J * nb] += pandya(t, X, Y, PH[2 * I + 1], PH[2 * I], PH[2 * J + 1], PH[2 * J]);
#undef V1
#undef V2
#undef V3
  }
  for (i64 q = 0; q < t->ncc; q++) { free(X[q]); free(Y[q]); }
  free(X);
  free(Y);
}

/* calc_cc, src/commutators.f90:1135-1159: HCC(r,c) = <j k|H|i l> */
void o_calc_cc(const tab_t *t, const double *v, int h, double *cc) {
#pragma omp parallel for schedule(dynamic)
  for (i64 q = 0; q < t->ncc; q++) {
    i64 r = t->cc_r[q], nb = t->cc_nb[q];
    const i64 *AB = t->cc_ab + 2 * t->cc_ab_start[q], *HP = t->cc_hp + 2 * t->cc_hp_start[q];
    for (i64 c = 0; c < nb; c++)
      for (i64 a = 0; a < r; a++)
        cc[t->cc_off[q] + a + c * r] = o_velem(t, v, h, AB[2 * a + 1], HP[2 * c], AB[2 * a], HP[2 * c + 1]);
  }
}

/* build_white, src/IMSRG_tools.f90:377-434: writes gen fph and Vpphh only */
void o_build_white(const tab_t *t, double *gen, const double *rec, int hr) {
  i64 n = t->n, M = t->M;
  for (i64 a = n; a < M; a++)
    for (i64 i = 0; i < n; i++)
      FPH(gen, a - n, i) = FPH(rec, a - n, i) /
                           (FPP(rec, a - n, a - n) - FHH(rec, i, i) - o_velem(t, rec, hr, a, i, a, i));
  for (i64 q = 0; q < t->nblock; q++) {
    i64 nh = t->nh[q], np = t->np[q];
    if (np * nh == 0) continue;
    const i64 *HH = t->qnhh + 2 * t->hh_start[q], *PP = t->qnpp + 2 * t->pp_start[q];
    for (i64 i = 0; i < np; i++)
      for (i64 j = 0; j < nh; j++) {
        i64 p1 = PP[2 * i], p2 = PP[2 * i + 1], q1 = HH[2 * j], q2 = HH[2 * j + 1];
        double efs = FPP(rec, p1 - n, p1 - n);
        efs = efs + FPP(rec, p2 - n, p2 - n);
        efs = efs - FHH(rec, q1, q1);
        efs = efs - FHH(rec, q2, q2);
        double aph = o_velem(t, rec, hr, p1, p2, p1, p2) + o_velem(t, rec, hr, q1, q2, q1, q2) -
                     o_velem(t, rec, hr, p1, q1, p1, q1) - o_velem(t, rec, hr, p1, q2, p1, q2) -
                     o_velem(t, rec, hr, p2, q1, p2, q1) - o_velem(t, rec, hr, p2, q2, p2, q2);
        gen[t->off[6 * q + PPHH] + i + j * np] = rec[t->off[6 * q + PPHH] + i + j * np] / (efs + aph);
      }
  }
}

/* MBPT2, src/HFQDmod.f90:325-362 */
double o_mbpt2(const tab_t *t, const double *v) {
  i64 n = t->n;
  double sm = 0.0;
  for (i64 q = 0; q < t->nblock; q++) {
    i64 nh = t->nh[q], np = t->np[q];
    const i64 *HH = t->qnhh + 2 * t->hh_start[q], *PP = t->qnpp + 2 * t->pp_start[q];
    const double *V = v + t->off[6 * q + PPHH];
    for (i64 I = 0; I < np; I++) {
      double f12 = FPP(v, PP[2 * I] - n, PP[2 * I] - n) + FPP(v, PP[2 * I + 1] - n, PP[2 * I + 1] - n);
      for (i64 J = 0; J < nh; J++) {
        double denom = FHH(v, HH[2 * J], HH[2 * J]) + FHH(v, HH[2 * J + 1], HH[2 * J + 1]) - f12;
        sm = sm + V[I + J * np] * V[I + J * np] / denom;
      }
    }
  }
  return sm;
}

/* mat_2_norm, src/ME_general.f90:2112-2138 (excludes E0) */
double o_norm(const tab_t *t, const double *v) {
  double od = 0.0;
  for (i64 e = 1; e < t->neq; e++) od += v[e] * v[e];
  return sqrt(od);
}

/* add_arrays, src/ME_general.f90:592-618: r3 = s1*r1 + s2*r2 */
void o_add(const tab_t *t, const double *r1, double s1, const double *r2, double s2, double *r3) {
  for (i64 e = 0; e < t->neq; e++) r3[e] = s1 * r1[e] + s2 * r2[e];
}

/* Full [r1, r2] in the call order of BCH_EXPAND (src/magnus_IMSRG.f90:297-311)
 * and dGam (src/gs_IMSRG.f90:256-268).  Returns the zero-body part, also
 * stored in r3[0]. */
double o_commutator(const tab_t *t, const double *v1, int h1, const double *v2, int h2, double *v3,
                    const double *hcc, double hh, const double *ecc, double eh) {
  memset(v3, 0, sizeof(double) * (size_t)t->neq);
  double e0 = o_comm110(t, v1, h1, v2, h2) + o_comm220(t, v1, h1, v2, h2);
  v3[0] = e0;
  o_comm111(t, v1, h1, v2, h2, v3);
  o_comm121(t, v1, h1, v2, h2, v3);
  o_comm221(t, v1, h1, v2, h2, v3);
  o_comm122(t, v1, h1, v2, h2, v3);
  o_comm222(t, v1, h1, v2, h2, v3, hcc, hh, ecc, eh);
  return e0;
}

/* BCH_EXPAND, src/magnus_IMSRG.f90:253-318.  hs <- sum_k cof(k) ad_G^k H with
 * the reference's literal coefficients and truncation rule.  Scratch: int1,
 * int2, ad (neq each).  Returns the number of nested commutators taken. */
static const double BCH_COF[11] = {1.0, 0.5, 0.166666666666666666, 0.04166666666666666,
                                   0.0083333333333333333, .001388888888888, 1.984126984e-4,
                                   2.48015873e-5, 2.755731922e-6, 2.755731922e-7, 2.505210839e-8};
/* REAL(4) literal 1e-5 widened to REAL(8) (src/magnus_IMSRG.f90:259) */
#define BCH_CONV 9.999999747378752e-06

int o_bch_expand(const tab_t *t, double *hs, const double *g, int hg, const double *h,
                 const double *hcc, double hh, const double *occ, double oh,
                 double *int1, double *int2, double *ad) {
  size_t bytes = sizeof(double) * (size_t)t->neq;
  memcpy(hs, h, bytes);
  memcpy(int2, hs, bytes);
  int k;
  for (k = 2; k <= 12; k++) {
    memcpy(int1, hs, bytes);
    memcpy(ad, int2, bytes);
    /* INT2 = [G, AD] with INT1/INT2/AD herm = +1 */
    o_commutator(t, g, hg, ad, 1, int2, hcc, hh, occ, oh);
    o_add(t, int1, 1.0, int2, BCH_COF[k - 2], hs);
    if (o_norm(t, int2) * BCH_COF[k - 2] / o_norm(t, int1) < BCH_CONV) break;
  }
  return k > 12 ? 11 : k - 1;
}

/* ---- EOM-IM-SRG matrix-vector product (excited states) --------------------
 * matvec_prod (src/EOM_IMSRG.f90:551-660) with HQ_comm_1b / HQ_comm_2b
 * (src/commutators.f90:1162-1260) restated loop for loop: the ladder operator
 * Q of transfer (ml, ms) has ph part Qph (mp x n, column-major like fph) and
 * pphh part Qpp (fermionic index of a < b, times that of i < j); x and y are
 * the reference's wrapped vectors (unwrap / rewrap, src/EOM_IMSRG.f90:815-
 * 884, 1011-1080): ph components with the particle outer, then pphh
 * components (a < b, i < j), a outer.  Returns the vector length. */
static i64 fidx0(i64 x, i64 y, i64 cnt) { return x * cnt - x * (x + 1) / 2 + (y - x - 1); }

typedef struct {
  i64 n, mp, np2;
  const double *ph, *pp;
} eomq_t;

static double o_qf(const eomq_t *q, i64 a, i64 b) {
  if (a < q->n || b >= q->n) return 0.0;
  return q->ph[(a - q->n) + b * q->mp];
}
static double o_qe(const eomq_t *q, i64 a, i64 b, i64 c, i64 d) {
  i64 n = q->n;
  if (a == b || c == d || a < n || b < n || c >= n || d >= n) return 0.0;
  double pre = 1.0;
  if (a > b) { i64 x = a; a = b; b = x; pre = -pre; }
  if (c > d) { i64 x = c; c = d; d = x; pre = -pre; }
  return pre * q->pp[fidx0(a - n, b - n, q->mp) + fidx0(c, d, n) * q->np2];
}

static double hq_comm_1b(const tab_t *t, const double *h, int hh, const eomq_t *q, i64 a, i64 i) {
  i64 N = t->n, M = t->M;
  double sm = 0.0;
  for (i64 b = N; b < M; b++) sm += o_felem(t, h, hh, a, b) * o_qf(q, b, i);
  for (i64 j = 0; j < N; j++) sm -= o_qf(q, a, j) * o_felem(t, h, hh, j, i);
  for (i64 b = N; b < M; b++)
    for (i64 j = 0; j < N; j++) {
      sm += o_felem(t, h, hh, j, b) * o_qe(q, a, b, i, j);
      sm += o_qf(q, b, j) * o_velem(t, h, hh, a, j, i, b);
    }
  for (i64 b = N; b < M; b++)
    for (i64 c = b + 1; c < M; c++)
      for (i64 j = 0; j < N; j++) sm += o_qe(q, c, b, i, j) * o_velem(t, h, hh, a, j, c, b);
  for (i64 b = N; b < M; b++)
    for (i64 k = 0; k < N; k++)
      for (i64 j = k + 1; j < N; j++) sm -= o_qe(q, a, b, k, j) * o_velem(t, h, hh, k, j, i, b);
  return sm;
}

static double hq_comm_2b(const tab_t *t, const double *h, int hh, const eomq_t *q, i64 a, i64 b, i64 i, i64 j) {
  i64 N = t->n, M = t->M;
  double sm = 0.0;
  for (i64 c = N; c < M; c++) {
    sm += o_felem(t, h, hh, a, c) * o_qe(q, c, b, i, j) - o_felem(t, h, hh, b, c) * o_qe(q, c, a, i, j);
    sm += o_qf(q, c, i) * o_velem(t, h, hh, a, b, c, j) - o_qf(q, c, j) * o_velem(t, h, hh, a, b, c, i);
  }
  for (i64 k = 0; k < N; k++) {
    sm += -o_qe(q, a, b, k, j) * o_felem(t, h, hh, k, i) + o_qe(q, a, b, k, i) * o_felem(t, h, hh, k, j);
    sm += -o_velem(t, h, hh, b, k, j, i) * o_qf(q, a, k) + o_velem(t, h, hh, a, k, j, i) * o_qf(q, b, k);
  }
  for (i64 c = N; c < M; c++)
    for (i64 d = c + 1; d < M; d++) sm += o_velem(t, h, hh, a, b, c, d) * o_qe(q, c, d, i, j);
  for (i64 k = 0; k < N; k++)
    for (i64 l = k + 1; l < N; l++) sm += o_velem(t, h, hh, k, l, i, j) * o_qe(q, a, b, k, l);
  for (i64 k = 0; k < N; k++)
    for (i64 c = N; c < M; c++) {
      sm -= o_velem(t, h, hh, b, k, c, i) * o_qe(q, a, c, k, j);
      sm += o_velem(t, h, hh, a, k, c, i) * o_qe(q, b, c, k, j);
      sm += o_velem(t, h, hh, b, k, c, j) * o_qe(q, a, c, k, i);
      sm -= o_velem(t, h, hh, a, k, c, j) * o_qe(q, b, c, k, i);
    }
  return sm;
}

i64 o_eom_matvec(const tab_t *t, const double *h, int hh, int ml, int ms, const double *x, double *y) {
  i64 N = t->n, M = t->M, mp = t->mp;
  i64 np2 = mp * (mp - 1) / 2, nh2 = N * (N - 1) / 2;
  double *ph = calloc((size_t)(mp * N + 1), sizeof(double));
  double *pp = calloc((size_t)(np2 * (nh2 > 0 ? nh2 : 1) + 1), sizeof(double));
  eomq_t q = {N, mp, np2, ph, pp};
  i64 g = 0;
  /* unwrap */
  for (i64 a = N; a < M; a++)
    for (i64 i = 0; i < N; i++)
      if (t->ml[a] - t->ml[i] == ml && t->ms[a] - t->ms[i] == ms) {
        if (x) ph[(a - N) + i * mp] = x[g];
        g++;
      }
  i64 sps = g;
  for (i64 a = N; a < M; a++)
    for (i64 b = a + 1; b < M; b++)
      for (i64 i = 0; i < N; i++)
        for (i64 j = i + 1; j < N; j++)
          if (t->ml[a] + t->ml[b] - t->ml[i] - t->ml[j] == ml && t->ms[a] + t->ms[b] - t->ms[i] - t->ms[j] == ms) {
            if (x) pp[fidx0(a - N, b - N, mp) + fidx0(i, j, N) * np2] = x[g];
            g++;
          }
  if (x && y) {
    /* the products, then rewrap in the same order */
    i64 e = 0;
    for (i64 a = N; a < M; a++)
      for (i64 i = 0; i < N; i++)
        if (t->ml[a] - t->ml[i] == ml && t->ms[a] - t->ms[i] == ms) y[e++] = hq_comm_1b(t, h, hh, &q, a, i);
    i64 *lst = malloc(sizeof(i64) * 4 * (size_t)(g - sps + 1));
    i64 k = 0;
    for (i64 a = N; a < M; a++)
      for (i64 b = a + 1; b < M; b++)
        for (i64 i = 0; i < N; i++)
          for (i64 j = i + 1; j < N; j++)
            if (t->ml[a] + t->ml[b] - t->ml[i] - t->ml[j] == ml && t->ms[a] + t->ms[b] - t->ms[i] - t->ms[j] == ms) {
              lst[4 * k] = a; lst[4 * k + 1] = b; lst[4 * k + 2] = i; lst[4 * k + 3] = j;
              k++;
            }
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 r = 0; r < k; r++) y[sps + r] = hq_comm_2b(t, h, hh, &q, lst[4 * r], lst[4 * r + 1], lst[4 * r + 2], lst[4 * r + 3]);
    free(lst);
  }
  free(ph);
  free(pp);
  return g;
}

/* ---- EOM-IM-SRG for one particle attached (mode 1) or removed (mode 2) -----
 * matvec_1p_attached / matvec_1h_removed (src/EOM_IMSRG.f90:662-812) with
 * HQ_comm_1p, HQ_comm_2p1h, HQ_comm_1h, HQ_comm_2h1p (src/commutators.f90:
 * 1263-1425) on the reference's vec layout (get_X1 / get_X3, src/ME_general.
 * f90:1229-1280): mode 1: X1 of particle a at a - n, X3(a<b particles, hole
 * i) at mp + fidx(a-n, b-n, mp) n + i; mode 2: X1 of hole i at i, X3(a<b
 * holes, particle i) at fidx(a, b, n) mp + i.  x and y are the wrapped target
 * components (unwrap_1p / unwrap_1h order).  Returns their count. */
typedef struct { i64 n, mp, mode; double *v; } vec13_t;

static double gx1(const vec13_t *q, i64 a) { return a >= q->n ? q->v[a - q->n] : q->v[a]; }
static i64 x3pos(const vec13_t *q, i64 ax, i64 bx, i64 i, double *ph) {
  i64 a = ax, b = bx;
  *ph = 1.0;
  if (ax > bx) { a = bx; b = ax; *ph = -1.0; }
  if (i >= q->n) return fidx0(a, b, q->n) * q->mp + i;            /* third index a particle */
  return q->mp + fidx0(a - q->n, b - q->n, q->mp) * q->n + i;     /* third index a hole */
}
static double gx3(const vec13_t *q, i64 ax, i64 bx, i64 i) {
  if (ax == bx) return 0.0;
  double ph;
  i64 k = x3pos(q, ax, bx, i, &ph);
  return q->v[k] * ph;
}

static double hq_1p(const tab_t *t, const double *h, int hh, const vec13_t *q, i64 a) {
  i64 N = t->n, M = t->M;
  double sm = 0.0;
  for (i64 b = N; b < M; b++) sm += o_felem(t, h, hh, a, b) * gx1(q, b);
  for (i64 b = N; b < M; b++)
    for (i64 i = 0; i < N; i++) sm += o_felem(t, h, hh, i, b) * gx3(q, a, b, i);
  for (i64 b = N; b < M; b++)
    for (i64 c = b + 1; c < M; c++)
      for (i64 i = 0; i < N; i++) sm += o_velem(t, h, hh, a, i, b, c) * gx3(q, b, c, i);
  return sm;
}
static double hq_1h(const tab_t *t, const double *h, int hh, const vec13_t *q, i64 i) {
  i64 N = t->n, M = t->M;
  double sm = 0.0;
  for (i64 j = 0; j < N; j++) sm -= o_felem(t, h, hh, j, i) * gx1(q, j);
  for (i64 b = N; b < M; b++)
    for (i64 j = 0; j < N; j++) sm += o_felem(t, h, hh, j, b) * gx3(q, i, j, b);
  for (i64 j = 0; j < N; j++)
    for (i64 k = j + 1; k < N; k++)
      for (i64 b = N; b < M; b++) sm += o_velem(t, h, hh, j, k, b, i) * gx3(q, j, k, b);
  return sm;
}
static double hq_2p1h(const tab_t *t, const double *h, int hh, const vec13_t *q, i64 a, i64 b, i64 i) {
  i64 N = t->n, M = t->M;
  double sm = 0.0;
  for (i64 c = N; c < M; c++)
    sm += o_velem(t, h, hh, a, b, c, i) * gx1(q, c) + o_felem(t, h, hh, a, c) * gx3(q, c, b, i) -
          o_felem(t, h, hh, b, c) * gx3(q, c, a, i);
  for (i64 j = 0; j < N; j++) sm -= o_felem(t, h, hh, j, i) * gx3(q, a, b, j);
  for (i64 c = N; c < M; c++)
    for (i64 d = c + 1; d < M; d++) sm += o_velem(t, h, hh, a, b, c, d) * gx3(q, c, d, i);
  for (i64 c = N; c < M; c++)
    for (i64 j = 0; j < N; j++)
      sm = sm - o_velem(t, h, hh, a, j, c, i) * gx3(q, c, b, j) + o_velem(t, h, hh, b, j, c, i) * gx3(q, c, a, j);
  return sm;
}
static double hq_2h1p(const tab_t *t, const double *h, int hh, const vec13_t *q, i64 i, i64 j, i64 a) {
  i64 N = t->n, M = t->M;
  double sm = 0.0;
  for (i64 k = 0; k < N; k++)
    sm += o_velem(t, h, hh, a, k, i, j) * gx1(q, k) - o_felem(t, h, hh, k, j) * gx3(q, i, k, a) +
          o_felem(t, h, hh, k, i) * gx3(q, j, k, a);
  for (i64 b = N; b < M; b++) sm += o_felem(t, h, hh, a, b) * gx3(q, i, j, b);
  for (i64 k = 0; k < N; k++)
    for (i64 l = k + 1; l < N; l++) sm += o_velem(t, h, hh, k, l, i, j) * gx3(q, k, l, a);
  for (i64 b = N; b < M; b++)
    for (i64 k = 0; k < N; k++)
      sm = sm - o_velem(t, h, hh, a, k, b, i) * gx3(q, k, j, b) + o_velem(t, h, hh, a, k, b, j) * gx3(q, k, i, b);
  return sm;
}

i64 o_eom13_matvec(const tab_t *t, const double *h, int hh, int mode, int ml, int ms, const double *x, double *y) {
  i64 N = t->n, M = t->M, mp = t->mp;
  i64 dm = mode == 1 ? mp + N * mp * (mp - 1) / 2 : N + mp * N * (N - 1) / 2;
  double *vin = calloc((size_t)dm + 1, sizeof(double)), *vout = calloc((size_t)dm + 1, sizeof(double));
  vec13_t qi = {N, mp, mode, vin};
  /* unwrap: one-index components, then (x < y, z) */
  i64 g = 0;
  i64 (*lst)[3] = malloc(sizeof(i64[3]) * (size_t)(dm + 1));
  i64 nl = 0;
  if (mode == 1) {
    for (i64 a = N; a < M; a++)
      if (t->ml[a] == ml && t->ms[a] == ms) { lst[nl][0] = a; lst[nl][1] = -1; nl++; }
  } else {
    for (i64 i = 0; i < N; i++)
      if (-t->ml[i] == ml && -t->ms[i] == ms) { lst[nl][0] = i; lst[nl][1] = -1; nl++; }
  }
  i64 n1 = nl;
  if (mode == 1) {
    for (i64 a = N; a < M; a++)
      for (i64 b = a + 1; b < M; b++)
        for (i64 i = 0; i < N; i++)
          if (t->ml[a] + t->ml[b] - t->ml[i] == ml && t->ms[a] + t->ms[b] - t->ms[i] == ms) {
            lst[nl][0] = a; lst[nl][1] = b; lst[nl][2] = i; nl++;
          }
  } else {
    for (i64 a = 0; a < N; a++)
      for (i64 b = a + 1; b < N; b++)
        for (i64 i = N; i < M; i++)
          if (-t->ml[a] - t->ml[b] + t->ml[i] == ml && -t->ms[a] - t->ms[b] + t->ms[i] == ms) {
            lst[nl][0] = a; lst[nl][1] = b; lst[nl][2] = i; nl++;
          }
  }
  g = nl;
  if (x && y) {
    for (i64 k = 0; k < nl; k++) {
      if (k < n1) {
        i64 a = lst[k][0];
        vin[a >= N ? a - N : a] = x[k];
      } else {
        double ph;
        i64 pos = x3pos(&qi, lst[k][0], lst[k][1], lst[k][2], &ph);
        vin[pos] = x[k] * ph;
      }
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (i64 k = 0; k < nl; k++) {
      if (k < n1)
        y[k] = mode == 1 ? hq_1p(t, h, hh, &qi, lst[k][0]) : hq_1h(t, h, hh, &qi, lst[k][0]);
      else
        y[k] = mode == 1 ? hq_2p1h(t, h, hh, &qi, lst[k][0], lst[k][1], lst[k][2])
                         : hq_2h1p(t, h, hh, &qi, lst[k][0], lst[k][1], lst[k][2]);
    }
  }
  free(lst);
  free(vin);
  free(vout);
  return g;
}
