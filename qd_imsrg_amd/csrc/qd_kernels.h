// Kernel launchers and the work-list records they consume.
#pragma once
#include "qd_internal.h"

namespace qd {

// ---- two-body element tiles: one record per run of <= 256 consecutive
// elements of one stored sub-block
struct Seg2 {
  int64_t off;      // flat offset of the sub-block
  int rows, size;   // rows (ld) and rows*cols
  int block, sub;
  int bra0, ket0;   // start of the bra / ket pair lists in DevTab::pairs
};
struct Tile2 { int seg, start; };

constexpr int NXCD = 8;                     // MI355X: work-group i runs on XCD i % 8
struct TilePair { int blk, ti, tj; };       // 32x32 tiles (ti, tj) and (tj, ti) of one CC block

// ---- grouped FP64 GEMM work lists
// Operand slots resolved at launch time (the same lists serve every call).
enum Slot { S_R1 = 0, S_R2, S_OUT, S_W1, S_W2, S_W3, S_W4, S_CC1, S_CC2, S_P, NSLOT };
struct Slots { double *p[NSLOT]; };
// bits of the sub-block types of an operator: the six two-body sub-blocks
// (enum Sub), then the one-body fhh, fpp, fph
enum { ZB_FHH = 6, ZB_FPP = 7, ZB_FPH = 8, ZB_ALL = (1 << 9) - 1 };
struct GemmSeg {   // one product op(A)*op(B) of a problem, alpha = +-1
  int64_t a_off, b_off;
  int a_slot, b_slot, lda, ldb, ta, tb, K;
  double alpha;
};
struct GemmProb {  // C(M,N) (+)= sum over segments
  int64_t c_off;
  int c_slot, ldc, M, N, beta1, seg0, nseg, shape;
};
struct GemmTile { int prob, tm, tn; };
// a range of a workspace slot to zero: len doubles from off
struct ZChunk { int slot, len; int64_t off; };
void launch_zero_chunks(hipStream_t s, const Slots &slots, const ZChunk *chunks, int n);
constexpr int GEMM_TK = 32;
// tile shapes of the grouped GEMM (rows x cols of C per work-group)
// SH_64x16K: a 64 x 16 tile whose four waves split every K chunk and add
// their partial tiles in a fixed order at the end (for N <= 16 problems).
// SH_128x128: four waves of 64 x 64 (K chunks of 16), half the operand
// traffic per flop of 64 x 64 (for large problems)
enum GemmShape { SH_64x64 = 0, SH_128x64, SH_32x32, SH_128x16, SH_64x16, SH_16x64, SH_16x128, SH_64x16K, SH_128x128,
                 NSHAPE };
constexpr int SHAPE_TM[NSHAPE] = {64, 128, 32, 128, 64, 16, 16, 64, 128};
constexpr int SHAPE_TN[NSHAPE] = {64, 64, 32, 16, 16, 64, 128, 16, 128};
// arithmetic of the grouped GEMM: FP64, FP32, or BF16 mantissa slices (the
// value is the slice count)
enum GemmPrec { GEMM_FP64 = 0, GEMM_FP32 = 1, GEMM_BF16X2 = 2, GEMM_BF16X3 = 3, GEMM_BF16X4 = 4 };

struct Ops {  // two operators and their hermiticity
  const double *v1, *v2;
  int h1, h2;
};

// ---- device helpers shared by the kernel sources
// v_elem (src/ME_general.f90:1557-1680): pair sorting with sign, the i_array
// lookup, zero across blocks, and the herm factor for the three sub-blocks
// that are stored transposed
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

// flags of the fused two-body kernel
// TB_DSTAGE: whole boxes stage the Pandya sub-blocks of D (run panels) in LDS
// TB_LDIRECT: the square sub-blocks' ladder terms are formed from the products
// w1 and w3 (w2 for PHPH) at (I, J) and (J, I) instead of read from L.  The
// ZB bits of r1's nonzero sub-block types go in flags >> TB_NZ_SHIFT: r1's
// elements and products of a zero type are not read
enum { TB_122 = 1, TB_LADDER = 2, TB_PANDYA = 4, TB_ACCUM = 8, TB_DSTAGE = 32, TB_LDIRECT = 64 };
constexpr int TB_NZ_SHIFT = 16;
// flags of the one-body kernel; OB_MV: only the 121 mat-vecs into u1, u2
// (they read the operands and CC sets only), OB_NOMV: u1, u2 are filled
enum { OB_111 = 1, OB_121 = 2, OB_221 = 4, OB_ACCUM = 8, OB_MV = 16, OB_NOMV = 32 };

// Small products: 16 x 16 tiles (tm, tn in units of 16), FP64 FMA.
void launch_gemm_small(hipStream_t s, const GemmProb *probs, const GemmSeg *segs, const GemmTile *tiles,
                       int ntiles, const Slots &slots);
// Shapes that can share one launch (similar LDS and register budgets).
constexpr bool gemm_mixable(int sh) {
  return sh == SH_64x64 || sh == SH_32x32;
}
constexpr int SH_MIX = NSHAPE;  // launch_gemm: tiles of any mixable shape, each its problem's
// All tiles must have the given shape (or, for SH_MIX, a mixable one).
// grid <= 0: one work-group per tile; else a persistent grid striding the
// list; tiles with prob < 0 are skipped.
void launch_gemm(hipStream_t s, const GemmProb *probs, const GemmSeg *segs, const GemmTile *tiles,
                 int ntiles, const Slots &slots, int grid, int shape, int prec);

// ---- class-pair boxes of the two-body pass.  A box = the stored outputs
// whose bra lies in one class pair (rows) and whose ket in another (columns).
// mode 0: whole boxes, the class pairs blist[first .. first+ng-1] against
// column pair gcol; mode 1: one box, columns lo .. lo+cnt-1 of its grid, with
// the bra-side 1b x 2b terms (plus ladder and Pandya terms), added to what the
// mode-2 tasks wrote; mode 2: one box, rows lo .. lo+cnt-1, writing the
// ket-side 1b x 2b terms (their own launch, before the others).
struct BTask { int block, gcol, first, ng, mode, lo, cnt; };
// blist entry flag: the box (rows a, columns b) also writes the outputs of the
// box (rows b, columns a) as h_Z times its own 1b x 2b and Pandya parts
constexpr int BOX_PAIRED = 1 << 30;
// 512 threads per task: LDS per work-group stays under 80 KB (two per CU,
// 16 waves, as many as the kernel's registers allow), so tasks hold up to
// 3600 grid elements (no box of R = 14 is split); channels of up to 10
// states (R = 20 shells) fit the channel-block area
constexpr int BX_THREADS = 512;
constexpr int BX_VCAP = 3600;  // grid elements per operand staged in LDS
constexpr int BX_FCAP = 1200;  // doubles of one-body channel blocks staged in LDS
constexpr int BX_GMAX = 48;    // boxes per task (at most 64: one wave's lanes at set-up)
constexpr int BX_RCAP = 384;   // row grid entries per task (staged in LDS)
constexpr int BX_CMAX = 128;   // column grid entries per task (a class pair of <= 11 x 11 states)
// The wide form: 1024 threads and 148 KB of LDS per task (one per CU, 16
// waves), grids of up to 8000 elements, for channel structures where many
// boxes exceed 3600 elements (R = 20: a third of the grid elements), which
// the narrow form splits into slabs
constexpr int BX_WIDE_THREADS = 1024;
constexpr int BX_WIDE_VCAP = 8000;
constexpr int BX_WIDE_RCAP = 768;
// one-body matrices of both operands blocked by channel: F[0, fcls_size) of
// r1, F[fcls_size, 2 fcls_size) of r2
void launch_fclass(hipStream_t s, const DevTab &t, const Ops &o, double *F);
// all stored two-body outputs of the tasks: 1b x 2b (TB_122), ladder
// combinations of the products w1..w3 and, for the summed products of Vpphh
// and Vppph and (without TB_LDIRECT) the square sub-blocks, of L
// (TB_LADDER; launch_sum_halves and launch_ladder_sq have filled L), Pandya term from D
// (TB_PANDYA); TB_ACCUM adds to out (mode 1 tasks add to what the mode 2 tasks wrote).  wide: the
// tasks were built for the wide form (BX_WIDE_*).  dsign != 0: D is in run
// panels with only the sub-blocks of runs (a, b), a_start <= b_start, stored
// (launch_pandya_fused), a lower element (x, y) is read as dsign * D(y, x);
// dsign = 0: D is whole and column-major
// Paired boxes (BOX_PAIRED in blist, the bt1p and bt2p task lists) need operators that
// obey their herm flags (square sub-blocks and fhh, fpp (anti)symmetric) and,
// with the Pandya term, D built from calc_cc of those operators.
void launch_box2b(hipStream_t s, const DevTab &t, const Ops &o, double *out, const BTask *tasks, int ntasks,
                  const int *blist, int flags, const double *w1, const double *w2, const double *w3,
                  const double *L, const double *F, const double *D, double dsign, bool wide);
// L = w3 + w4 over the given sub-blocks (the two halves of the summed
// products of Vpphh and Vppph)
void launch_sum_halves(hipStream_t s, const double *w3, const double *w4, double *L, const Seg2 *segs,
                       const Tile2 *tiles, int ntiles);
// Ladder terms of the square sub-blocks (HHHH, PHPH, PPPP; they need the
// products at (J, I) as well as (I, J)) into L, over 32x32 tile pairs (ti <=
// tj; TilePair::blk = 3 * block + 0 HHHH / 1 PHPH / 2 PPPP):
// HHHH, PPPP: L = (h1 w1^T - h2 w1) + (h1 h2 w3^T - w3);
// PHPH:       L = (h1 w2^T - h2 w2) + (h1 w1^T - h2 w1)
// nz1: nonzero sub-block types of r1 (ZB bits; ZB_ALL outside the BCH terms)
void launch_ladder_sq(hipStream_t s, const DevTab &t, const double *w1, const double *w2, const double *w3,
                      double *L, const TilePair *tp, int ntp, double h1, double h2, uint32_t nz1);
// cc1 / cc2: calc_cc of r1 / r2 when the caller holds them (then the CC
// mat-vecs of commutator_121 read them instead of gathering V), else null
// OB_111 needs the one-body products of the one-body product list in the
// one-body region of w1, w2, w3; list221: the n221 outputs whose states share
// a channel (221 traces)
// g221: per listed output and state a, the codes (flat index | sign << 31,
// CC_ZERO) of <a x|w|a y> and <a y|w|a x>, or null (the traces then resolve
// the pair tables per element)
void launch_onebody(hipStream_t s, const DevTab &t, const Ops &o, double *out, int flags,
                    const double *w1, const double *w2, const double *w3, double *u1, double *u2,
                    const double *cc1, const double *cc2, const int *list221, int n221, const uint32_t *g221);
// zero-body part (commutator_110 + commutator_220) into *dst (device scalar)
void launch_zerobody(hipStream_t s, const DevTab &t, const Ops &o, const Seg2 *pphh_segs, const Tile2 *pphh_tiles,
                     int ntiles, double *part, double *dst, int which /*1: 110, 2: 220, 3: both*/);
void launch_calc_cc(hipStream_t s, const DevTab &t, const double *v, int herm, double *cc,
                    const Seg2 *cc_segs, const Tile2 *cc_tiles, int ntiles);
void launch_white(hipStream_t s, const DevTab &t, const double *rec, int hrec, double *gen,
                  const Seg2 *pphh_segs, const Tile2 *pphh_tiles, int ntiles);
void launch_pandya_combine(hipStream_t s, const DevTab &t, double *P, const TilePair *tp, int ntp, double hh,
                           double eh);
// D = hh * E H^T - eh * (E H^T)^T per CC block straight from the CC sets
// (E = ETACC, H = HCC), over 64x64 tile pairs (ti <= tj).  D is written in
// run panels (each run of device rows whose first states share a channel, of
// length l starting at row L, is a column-major l x r panel at L r), and only
// the sub-blocks of runs (a, b) with a_start <= b_start (the others are
// -hh * eh times their transpose).
// E H^T in FP64 or in BF16 slices (prec); pandya_fused_prec tells which
// precisions have the fused pass (FP32 keeps product GEMM + combine).  The
// BF16 pass first cuts both CC sets into slice planes (CcSlices): per CC
// block, units of 8 consecutive k of one row, zero past K, unit (k / 8, x)
// of slice s at planes + s * len + off[q] + 8 ((k / 8) r + x).
constexpr int CS_UNITS_WG = 256;  // slicing work-group: 256 units of one block
struct CcSlices {
  unsigned short *planes;  // E slices 0..3, then H slices 0..3, len elements each
  int64_t len;
  const int64_t *off;      // per CC block
  const int2 *units;       // (block, first unit) per slicing work-group
  int nunits;
};
void launch_pandya_fused(hipStream_t s, const DevTab &t, const double *etacc, const double *hcc, double *D,
                         const TilePair *tp, int ntp, double hh, double eh, int prec, const CcSlices &cs);
bool pandya_fused_prec(int prec);
void launch_axpby(hipStream_t s, int64_t n, double s1, const double *x, double s2, const double *y, double *z);
// nonzero sub-block types of an operator: *nz = OR over the tiles of (bit of
// the tile's type if any of its elements is nonzero)
struct ZTile { int64_t off; int len, bit; };
void launch_zero_scan(hipStream_t s, const double *v, const ZTile *tiles, int ntiles, uint32_t *nz);
// sum of squares of v[1..neq) -> *dst, via partials (deterministic)
void launch_norm2(hipStream_t s, const double *v, int64_t neq, double *partials, double *dst);
// BCH term: dst[0] = ||hs||^2 before and dst[1] = ||ad||^2 over [1, neq)
// (as launch_norm2), then hs += cof * ad over [0, neq); partials: 2048 doubles
void launch_bch_update(hipStream_t s, double *hs, const double *ad, double cof, int64_t neq, double *partials,
                       double *dst);
// sum of a[k] b[k] over [1, neq) -> *dst (deterministic)
void launch_dot(hipStream_t s, const double *a, const double *b, int64_t neq, double *partials, double *dst);
constexpr int LINCOMB_MAX = 32;
struct LinComb {  // z = sum_i c[i] x[i]
  int k;
  double c[LINCOMB_MAX];
  const double *x[LINCOMB_MAX];
};
void launch_lincomb(hipStream_t s, int64_t n, const LinComb &lc, double *z);
void launch_mbpt2(hipStream_t s, const DevTab &t, const double *v, const Seg2 *pphh_segs,
                  const Tile2 *pphh_tiles, int ntiles, double *partials, double *dst);
void launch_scale_add_scalar(hipStream_t s, double *dst, const double *a, const double *b);

}  // namespace qd
