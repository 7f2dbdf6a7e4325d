// Grouped FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// One launch runs every (problem, output tile) of a work list; the problems are
// the per-(M_L, M_S)-block products of matrix_mult222 (src/commutators.f90:
// 999-1127), commutator_221 (:236-272) and construct_cc_ints
// (src/ME_general.f90:948-949).  A problem sums up to four products op(A)op(B)
// (alpha = +-1, applied while staging A), so several reference dgemm calls
// that add into one sub-block become one pass.
//
// Each problem carries its own tile shape (chosen per size bucket by the
// autotuner in capi.hip); every shape is its own kernel instantiation (own
// register budget), launched once over that shape's part of the tile list.
// The launches are persistent: a fixed grid strides over the tiles, which are
// sorted by decreasing work.  Work-group = 256 threads = 4 waves; K is staged through
// LDS in chunks of 32 (8 MFMA k-steps; 16 for the FP64 mixed launch and the
// 128 x 128 tiles).  MFMA operand layout: A(16x4) lane l
// holds A[l%16][l/16], B(4x16) lane l holds B[l/16][l%16], D(16x16) lane l
// holds D[4*v + l/16][l%16] in component v (FP64: one row per group).
#include "qd_kernels.h"

namespace qd {

typedef double d4 __attribute__((ext_vector_type(4)));

// Walk of K in chunks of KC over all segments of a problem, with the next
// chunk's global loads issued into registers (load) before the current
// chunk's MFMAs (compute) and written to LDS (store) after the next barrier.
// compute(kv) gets the count of valid k of the chunk (the rest is zero fill).
// load only issues the loads: anything computed from the loaded values (the
// segment's alpha, a conversion, the slicing) belongs in store, which gets
// the staged segment's ta, tb and alpha.  Arithmetic on them inside load
// would make the loads complete before the current chunk's MFMAs start.
template <int KC, class Load, class Store, class Compute>
__device__ __forceinline__ void chunk_walk(const GemmProb &pr, const GemmSeg *__restrict__ segs, Load load,
                                           Store store, Compute compute) {
  if (pr.nseg <= 0) return;
  int si = 0, k0 = 0;
  GemmSeg sg = segs[pr.seg0];
  load(sg, 0);
  int cta = sg.ta, ctb = sg.tb;
  double calpha = sg.alpha;
  while (true) {
    const int kv = sg.K - k0 < KC ? sg.K - k0 : KC;
    int nsi = si, nk0 = k0 + KC;
    GemmSeg nsg = sg;
    if (nk0 >= sg.K) {
      nsi = si + 1;
      nk0 = 0;
      if (nsi < pr.nseg) nsg = segs[pr.seg0 + nsi];
    }
    const bool more = nsi < pr.nseg;
    __syncthreads();
    store(cta, ctb, calpha);
    __syncthreads();
    if (more) {
      load(nsg, nk0);
      cta = nsg.ta;
      ctb = nsg.tb;
      calpha = nsg.alpha;
    }
    compute(kv);
    if (!more) break;
    si = nsi;
    k0 = nk0;
    sg = nsg;
  }
}

// shape table: TM x TN tile, WM x WN waves, each wave FM x FN MFMA tiles
template <int TM, int TN, int WM, int WN, int KC = GEMM_TK>
__device__ __forceinline__ void gemm_tile(const GemmProb &pr, const GemmSeg *__restrict__ segs, const Slots &slots,
                                          int tm, int tn, double *__restrict__ lds) {
  constexpr int FM = TM / WM / 16, FN = TN / WN / 16;
  double *As = lds;                       // [KC][TM + 1]
  double *Bs = lds + KC * (TM + 1);  // [KC][TN + 1]
  const int row0 = tm * TM, col0 = tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int lr = lane & 15, lk = lane >> 4;
  const int m0 = wm * (TM / WM), n0 = wn * (TN / WN);
  d4 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; a++)
#pragma unroll
    for (int b = 0; b < FN; b++) acc[a][b] = d4{0, 0, 0, 0};

  constexpr int NA = (TM * KC + 255) / 256, NB = (TN * KC + 255) / 256;
  double ra[NA], rb[NB];
  auto load = [&](const GemmSeg &sg, int k0) {
    const double *__restrict__ A = slots.p[sg.a_slot] + sg.a_off;
    const double *__restrict__ B = slots.p[sg.b_slot] + sg.b_off;
#pragma unroll
    for (int r = 0; r < NA; r++) {
      const int idx = tid + 256 * r;
      int i, k;
      if (!sg.ta) { i = idx % TM; k = idx / TM; }        // columns of A contiguous in i
      else { k = idx % KC; i = idx / KC; }     // A stored transposed
      const int gi = row0 + i, gk = k0 + k;
      double v = 0.0;
      if (idx < TM * KC && gi < pr.M && gk < sg.K)
        v = sg.ta ? A[gk + (int64_t)gi * sg.lda] : A[gi + (int64_t)gk * sg.lda];
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
      const int idx = tid + 256 * r;
      int j, k;
      if (!sg.tb) { k = idx % KC; j = idx / KC; }  // columns of B contiguous in k
      else { j = idx % TN; k = idx / TN; }                   // B stored transposed
      const int gj = col0 + j, gk = k0 + k;
      double v = 0.0;
      if (idx < TN * KC && gj < pr.N && gk < sg.K)
        v = sg.tb ? B[gj + (int64_t)gk * sg.ldb] : B[gk + (int64_t)gj * sg.ldb];
      rb[r] = v;
    }
  };
  auto store = [&](int ta, int tb, double alpha) {
#pragma unroll
    for (int r = 0; r < NA; r++) {
      const int idx = tid + 256 * r;
      if (idx < TM * KC) {
        const int i = ta ? idx / KC : idx % TM, k = ta ? idx % KC : idx / TM;
        As[k * (TM + 1) + i] = alpha * ra[r];
      }
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
      const int idx = tid + 256 * r;
      if (idx < TN * KC) {
        const int j = tb ? idx % TN : idx / KC, k = tb ? idx / TN : idx % KC;
        Bs[k * (TN + 1) + j] = rb[r];
      }
    }
  };
  auto compute = [&](int kv) {
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      if (kk >= kv) break;  // zero fill past the segment's K adds nothing
      double a[FM], b[FN];
#pragma unroll
      for (int x = 0; x < FM; x++) a[x] = As[(kk + lk) * (TM + 1) + m0 + x * 16 + lr];
#pragma unroll
      for (int y = 0; y < FN; y++) b[y] = Bs[(kk + lk) * (TN + 1) + n0 + y * 16 + lr];
#pragma unroll
      for (int x = 0; x < FM; x++)
#pragma unroll
        for (int y = 0; y < FN; y++) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[y], a[x], acc[x][y], 0, 0, 0);
    }
  };
  chunk_walk<KC>(pr, segs, load, store, compute);
  // acc holds the tile transposed (B^T A^T): lane l, component v is row
  // l % 16, column 4 v + l / 16, so each store covers four columns of 16
  // consecutive rows of the column-major C (whole 128-byte lines).  With
  // beta = 1 the four old values of a 16 x 16 block are loaded before its
  // first store, one global round trip per block instead of one per element.
  double *__restrict__ C = slots.p[pr.c_slot] + pr.c_off;
#pragma unroll
  for (int x = 0; x < FM; x++)
#pragma unroll
    for (int y = 0; y < FN; y++) {
      const int r = row0 + m0 + x * 16 + lr;
      double old[4];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = col0 + n0 + y * 16 + 4 * v + lk;
        old[v] = 0.0;
        if (pr.beta1 && r < pr.M && c < pr.N) old[v] = C[r + (int64_t)c * pr.ldc];
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = col0 + n0 + y * 16 + 4 * v + lk;
        if (r < pr.M && c < pr.N) C[r + (int64_t)c * pr.ldc] = pr.beta1 ? old[v] + acc[x][y][v] : acc[x][y][v];
      }
    }
}

// 64 x 16 tile with split K: the chunk's 8 k-steps of 4 go to the four waves
// two each (wave w takes k-steps w and w + 4), every wave accumulating the whole
// 64 x 16 tile (four 16 x 16 MFMA blocks); the partial tiles are added through
// LDS in wave order.  For outputs with N <= 16 this keeps four MFMAs per
// k-step per wave without the padding of a wider tile.
__device__ __forceinline__ void gemm_tile_splitk(const GemmProb &pr, const GemmSeg *__restrict__ segs,
                                                 const Slots &slots, int tm, int tn, double *__restrict__ lds) {
  constexpr int TM = 64, TN = 16;
  double *As = lds;                       // [GEMM_TK][TM + 1]
  double *Bs = lds + GEMM_TK * (TM + 1);  // [GEMM_TK][TN + 1]
  const int row0 = tm * TM, col0 = tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc[4];
#pragma unroll
  for (int x = 0; x < 4; x++) acc[x] = d4{0, 0, 0, 0};
  constexpr int NA = (TM * GEMM_TK + 255) / 256, NB = (TN * GEMM_TK + 255) / 256;
  double ra[NA], rb[NB];
  auto load = [&](const GemmSeg &sg, int k0) {
    const double *__restrict__ A = slots.p[sg.a_slot] + sg.a_off;
    const double *__restrict__ B = slots.p[sg.b_slot] + sg.b_off;
#pragma unroll
    for (int r = 0; r < NA; r++) {
      const int idx = tid + 256 * r;
      int i, k;
      if (!sg.ta) { i = idx % TM; k = idx / TM; }
      else { k = idx % GEMM_TK; i = idx / GEMM_TK; }
      const int gi = row0 + i, gk = k0 + k;
      double v = 0.0;
      if (idx < TM * GEMM_TK && gi < pr.M && gk < sg.K)
        v = sg.ta ? A[gk + (int64_t)gi * sg.lda] : A[gi + (int64_t)gk * sg.lda];
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
      const int idx = tid + 256 * r;
      int j, k;
      if (!sg.tb) { k = idx % GEMM_TK; j = idx / GEMM_TK; }
      else { j = idx % TN; k = idx / TN; }
      const int gj = col0 + j, gk = k0 + k;
      double v = 0.0;
      if (idx < TN * GEMM_TK && gj < pr.N && gk < sg.K)
        v = sg.tb ? B[gj + (int64_t)gk * sg.ldb] : B[gk + (int64_t)gj * sg.ldb];
      rb[r] = v;
    }
  };
  auto store = [&](int ta, int tb, double alpha) {
#pragma unroll
    for (int r = 0; r < NA; r++) {
      const int idx = tid + 256 * r;
      if (idx < TM * GEMM_TK) {
        const int i = ta ? idx / GEMM_TK : idx % TM, k = ta ? idx % GEMM_TK : idx / TM;
        As[k * (TM + 1) + i] = alpha * ra[r];
      }
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
      const int idx = tid + 256 * r;
      if (idx < TN * GEMM_TK) {
        const int j = tb ? idx % TN : idx / GEMM_TK, k = tb ? idx / TN : idx % GEMM_TK;
        Bs[k * (TN + 1) + j] = rb[r];
      }
    }
  };
  auto compute = [&](int kv) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int kk = 4 * (wave + 4 * h);
      if (kk >= kv) break;  // zero fill past the segment's K adds nothing
      const double b = Bs[(kk + lk) * (TN + 1) + lr];
#pragma unroll
      for (int x = 0; x < 4; x++) {
        const double a = As[(kk + lk) * (TM + 1) + x * 16 + lr];
        acc[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[x], 0, 0, 0);
      }
    }
  };
  chunk_walk<GEMM_TK>(pr, segs, load, store, compute);
  // partial tiles of waves 1..3 through LDS (16 doubles per lane), added by wave 0
  __syncthreads();
  if (wave > 0) {
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int v = 0; v < 4; v++) lds[((wave - 1) * 16 + x * 4 + v) * 64 + lane] = acc[x][v];
  }
  __syncthreads();
  if (wave == 0) {
    double *__restrict__ C = slots.p[pr.c_slot] + pr.c_off;
    const int c = col0 + lr;
    double old[4][4];  // with beta = 1, every old value is loaded before the first store
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = row0 + x * 16 + 4 * v + lk;
        old[x][v] = 0.0;
        if (pr.beta1 && r < pr.M && c < pr.N) old[x][v] = C[r + (int64_t)c * pr.ldc];
      }
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        double sum = acc[x][v];
#pragma unroll
        for (int w = 0; w < 3; w++) sum += lds[(w * 16 + x * 4 + v) * 64 + lane];
        const int r = row0 + x * 16 + 4 * v + lk;
        if (r < pr.M && c < pr.N) C[r + (int64_t)c * pr.ldc] = pr.beta1 ? old[x][v] + sum : sum;
      }
  }
}

// ---- split-precision tiles: FP32, or the mantissa cut into S BF16 slices
//
// x = x_0 + x_1 + ... + x_{S-1} + r, each x_s a BF16 number taken by
// truncation from what the previous slices left (|r| < 2^(1-8S) |x|).  The
// product keeps the slice pairs (i, j) with i + j < S, smallest first, all on
// v_mfma_f32_16x16x32_bf16 (BF16 products are exact in FP32) into one FP32
// accumulator per K stage of 32; each stage's FP32 sum is then added into an
// FP64 accumulator, and the output update (beta) stays FP64.  FP32 mode rounds
// the operands to FP32 and uses v_mfma_f32_16x16x4_f32 with the same staging.
// FP32 MFMA output layout: lane l, component v holds D[4*(l/16) + v][l%16].
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));
constexpr int SPLIT_TK = 32;                    // K per staging round
constexpr int SPLIT_LDW = (SPLIT_TK + 8) / 2;   // 32-bit words per staged row (16-byte aligned, padded)

template <int S>
__device__ __forceinline__ void bf16_slices(double v, unsigned (&h)[S]) {
  double r = v;
#pragma unroll
  for (int s = 0; s < S; s++) {
    const unsigned u = __float_as_uint((float)r) & 0xffff0000u;
    h[s] = u >> 16;
    if (s + 1 < S) r -= (double)__uint_as_float(u);
  }
}

__device__ __forceinline__ double load_a(const GemmSeg &sg, const double *__restrict__ A, const GemmProb &pr, int gi,
                                         int gk) {
  if (gi >= pr.M || gk >= sg.K) return 0.0;
  return sg.ta ? A[gk + (int64_t)gi * sg.lda] : A[gi + (int64_t)gk * sg.lda];
}
__device__ __forceinline__ double load_b(const GemmSeg &sg, const double *__restrict__ B, const GemmProb &pr, int gj,
                                         int gk) {
  if (gj >= pr.N || gk >= sg.K) return 0.0;
  return sg.tb ? B[gj + (int64_t)gk * sg.ldb] : B[gk + (int64_t)gj * sg.ldb];
}

// C tile update from an FP64 accumulator in the FP32 MFMA output layout
template <int FM, int FN>
__device__ __forceinline__ void store_f32_layout(const GemmProb &pr, const Slots &slots, int r0, int c0, int lr, int lk,
                                                 const d4 (&acc)[FM][FN]) {
  double *__restrict__ C = slots.p[pr.c_slot] + pr.c_off;
#pragma unroll
  for (int x = 0; x < FM; x++) {
    double old[FN][4];  // with beta = 1, the row block's old values are loaded before its first store
#pragma unroll
    for (int y = 0; y < FN; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = r0 + x * 16 + 4 * lk + v, c = c0 + y * 16 + lr;
        old[y][v] = 0.0;
        if (pr.beta1 && r < pr.M && c < pr.N) old[y][v] = C[r + (int64_t)c * pr.ldc];
      }
#pragma unroll
    for (int y = 0; y < FN; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = r0 + x * 16 + 4 * lk + v, c = c0 + y * 16 + lr;
        if (r < pr.M && c < pr.N) C[r + (int64_t)c * pr.ldc] = pr.beta1 ? old[y][v] + acc[x][y][v] : acc[x][y][v];
      }
  }
}

template <int TM, int TN, int WM, int WN, int S>
__device__ __forceinline__ void gemm_tile_bf16(const GemmProb &pr, const GemmSeg *__restrict__ segs,
                                               const Slots &slots, int tm, int tn, unsigned *__restrict__ lds) {
  constexpr int FM = TM / WM / 16, FN = TN / WN / 16, KP = SPLIT_TK / 2;
  constexpr int NA = (TM * KP + 255) / 256, NB = (TN * KP + 255) / 256;
  unsigned *As = lds;                        // [S][TM][SPLIT_LDW]: slice s of A(i, k), k pairs packed
  unsigned *Bs = lds + S * TM * SPLIT_LDW;   // [S][TN][SPLIT_LDW]: slice s of B(k, j)
  const int row0 = tm * TM, col0 = tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int lr = lane & 15, lk = lane >> 4;
  const int m0 = wm * (TM / WM), n0 = wn * (TN / WN);
  d4 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; a++)
#pragma unroll
    for (int b = 0; b < FN; b++) acc[a][b] = d4{0, 0, 0, 0};
  double ra[NA][2], rb[NB][2];  // pairs of consecutive k
  auto load = [&](const GemmSeg &sg, int k0) {
    const double *__restrict__ A = slots.p[sg.a_slot] + sg.a_off;
    const double *__restrict__ B = slots.p[sg.b_slot] + sg.b_off;
#pragma unroll
    for (int r = 0; r < NA; r++) {
      const int idx = tid + 256 * r;
      const int i = sg.ta ? idx / KP : idx % TM, kp = sg.ta ? idx % KP : idx / TM;
      const bool in = idx < TM * KP;
      ra[r][0] = in ? load_a(sg, A, pr, row0 + i, k0 + 2 * kp) : 0.0;
      ra[r][1] = in ? load_a(sg, A, pr, row0 + i, k0 + 2 * kp + 1) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
      const int idx = tid + 256 * r;
      const int j = sg.tb ? idx % TN : idx / KP, kp = sg.tb ? idx / TN : idx % KP;
      const bool in = idx < TN * KP;
      rb[r][0] = in ? load_b(sg, B, pr, col0 + j, k0 + 2 * kp) : 0.0;
      rb[r][1] = in ? load_b(sg, B, pr, col0 + j, k0 + 2 * kp + 1) : 0.0;
    }
  };
  auto store = [&](int ta, int tb, double alpha) {
#pragma unroll
    for (int r = 0; r < NA; r++) {
      const int idx = tid + 256 * r;
      if (idx < TM * KP) {
        const int i = ta ? idx / KP : idx % TM, kp = ta ? idx % KP : idx / TM;
        unsigned h0[S], h1[S];
        bf16_slices<S>(alpha * ra[r][0], h0);
        bf16_slices<S>(alpha * ra[r][1], h1);
#pragma unroll
        for (int s = 0; s < S; s++) As[(s * TM + i) * SPLIT_LDW + kp] = h0[s] | (h1[s] << 16);
      }
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
      const int idx = tid + 256 * r;
      if (idx < TN * KP) {
        const int j = tb ? idx % TN : idx / KP, kp = tb ? idx / TN : idx % KP;
        unsigned h0[S], h1[S];
        bf16_slices<S>(rb[r][0], h0);
        bf16_slices<S>(rb[r][1], h1);
#pragma unroll
        for (int s = 0; s < S; s++) Bs[(s * TN + j) * SPLIT_LDW + kp] = h0[s] | (h1[s] << 16);
      }
    }
  };
  auto compute = [&](int) {
    bf8 a[S][FM], b[S][FN];
#pragma unroll
    for (int s = 0; s < S; s++) {
#pragma unroll
      for (int x = 0; x < FM; x++)
        a[s][x] = __builtin_bit_cast(bf8, *(const u4 *)(As + (s * TM + m0 + x * 16 + lr) * SPLIT_LDW + 4 * lk));
#pragma unroll
      for (int y = 0; y < FN; y++)
        b[s][y] = __builtin_bit_cast(bf8, *(const u4 *)(Bs + (s * TN + n0 + y * 16 + lr) * SPLIT_LDW + 4 * lk));
    }
#pragma unroll
    for (int x = 0; x < FM; x++)
#pragma unroll
      for (int y = 0; y < FN; y++) {
        f4 c = f4{0, 0, 0, 0};
#pragma unroll
        for (int d = S - 1; d >= 0; d--)
#pragma unroll
          for (int i = 0; i <= d; i++) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][x], b[d - i][y], c, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) acc[x][y][v] += (double)c[v];
      }
  };
  chunk_walk<SPLIT_TK>(pr, segs, load, store, compute);
  store_f32_layout<FM, FN>(pr, slots, row0 + m0, col0 + n0, lr, lk, acc);
}

template <int TM, int TN, int WM, int WN>
__device__ __forceinline__ void gemm_tile_f32(const GemmProb &pr, const GemmSeg *__restrict__ segs, const Slots &slots,
                                              int tm, int tn, float *__restrict__ lds) {
  constexpr int FM = TM / WM / 16, FN = TN / WN / 16;
  constexpr int NA = (TM * SPLIT_TK + 255) / 256, NB = (TN * SPLIT_TK + 255) / 256;
  float *As = lds;                          // [SPLIT_TK][TM + 1]
  float *Bs = lds + SPLIT_TK * (TM + 1);    // [SPLIT_TK][TN + 1]
  const int row0 = tm * TM, col0 = tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int lr = lane & 15, lk = lane >> 4;
  const int m0 = wm * (TM / WM), n0 = wn * (TN / WN);
  d4 acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; a++)
#pragma unroll
    for (int b = 0; b < FN; b++) acc[a][b] = d4{0, 0, 0, 0};
  double ra[NA], rb[NB];
  auto load = [&](const GemmSeg &sg, int k0) {
    const double *__restrict__ A = slots.p[sg.a_slot] + sg.a_off;
    const double *__restrict__ B = slots.p[sg.b_slot] + sg.b_off;
#pragma unroll
    for (int r = 0; r < NA; r++) {
      const int idx = tid + 256 * r;
      const int i = sg.ta ? idx / SPLIT_TK : idx % TM, k = sg.ta ? idx % SPLIT_TK : idx / TM;
      ra[r] = idx < TM * SPLIT_TK ? load_a(sg, A, pr, row0 + i, k0 + k) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
      const int idx = tid + 256 * r;
      const int j = sg.tb ? idx % TN : idx / SPLIT_TK, k = sg.tb ? idx / TN : idx % SPLIT_TK;
      rb[r] = idx < TN * SPLIT_TK ? load_b(sg, B, pr, col0 + j, k0 + k) : 0.0;
    }
  };
  auto store = [&](int ta, int tb, double alpha) {
#pragma unroll
    for (int r = 0; r < NA; r++) {
      const int idx = tid + 256 * r;
      if (idx < TM * SPLIT_TK) {
        const int i = ta ? idx / SPLIT_TK : idx % TM, k = ta ? idx % SPLIT_TK : idx / TM;
        As[k * (TM + 1) + i] = (float)(alpha * ra[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
      const int idx = tid + 256 * r;
      if (idx < TN * SPLIT_TK) {
        const int j = tb ? idx % TN : idx / SPLIT_TK, k = tb ? idx / TN : idx % SPLIT_TK;
        Bs[k * (TN + 1) + j] = (float)rb[r];
      }
    }
  };
  auto compute = [&](int) {
    f4 c[FM][FN];
#pragma unroll
    for (int x = 0; x < FM; x++)
#pragma unroll
      for (int y = 0; y < FN; y++) c[x][y] = f4{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < SPLIT_TK; kk += 4) {
      float a[FM], b[FN];
#pragma unroll
      for (int x = 0; x < FM; x++) a[x] = As[(kk + lk) * (TM + 1) + m0 + x * 16 + lr];
#pragma unroll
      for (int y = 0; y < FN; y++) b[y] = Bs[(kk + lk) * (TN + 1) + n0 + y * 16 + lr];
#pragma unroll
      for (int x = 0; x < FM; x++)
#pragma unroll
        for (int y = 0; y < FN; y++) c[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[x], b[y], c[x][y], 0, 0, 0);
    }
#pragma unroll
    for (int x = 0; x < FM; x++)
#pragma unroll
      for (int y = 0; y < FN; y++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[x][y][v] += (double)c[x][y][v];
  };
  chunk_walk<SPLIT_TK>(pr, segs, load, store, compute);
  store_f32_layout<FM, FN>(pr, slots, row0 + m0, col0 + n0, lr, lk, acc);
}

// LDS bytes of one work-group for shape (TM, TN) and precision P
constexpr int gemm_lds_bytes(int TM, int TN, int P) {
  return P == GEMM_FP64  ? GEMM_TK * (TM + 1 + TN + 1) * 8
         : P == GEMM_FP32 ? SPLIT_TK * (TM + 1 + TN + 1) * 4
                          : P * (TM + TN) * SPLIT_LDW * 4;
}

template <int SH, int P>
constexpr int tile_lds_bytes() {
  constexpr int TM = SHAPE_TM[SH], TN = SHAPE_TN[SH];
  return (SH == SH_64x16K && P == GEMM_FP64)
             ? (gemm_lds_bytes(TM, TN, P) > 3 * 16 * 64 * 8 ? gemm_lds_bytes(TM, TN, P) : 3 * 16 * 64 * 8)
         : SH == SH_128x128 ? (P == GEMM_FP64 ? 16 * (TM + 1 + TN + 1) * 8 : gemm_lds_bytes(64, 64, P))
                            : gemm_lds_bytes(TM, TN, P);
}

// one output tile of shape SH in precision P
template <int SH, int P, int KC = GEMM_TK>
__device__ __forceinline__ void run_tile(const GemmProb &pr, const GemmSeg *__restrict__ segs, const Slots &slots,
                                         const GemmTile &tl, unsigned char *lds) {
  constexpr int TM = SHAPE_TM[SH], TN = SHAPE_TN[SH];
  constexpr int WM = TN <= 16 ? 4 : (TM <= 16 ? 1 : 2), WN = 4 / WM;
  constexpr bool BIG = SH == SH_128x128;  // FP64: K chunks of 16; split precisions: four 64 x 64 sub-tiles
  if constexpr (SH == SH_64x16K && P == GEMM_FP64) gemm_tile_splitk(pr, segs, slots, tl.tm, tl.tn, (double *)lds);
  else if constexpr (BIG && P == GEMM_FP64) gemm_tile<128, 128, 2, 2, 16>(pr, segs, slots, tl.tm, tl.tn, (double *)lds);
  else if constexpr (BIG) {
    for (int q = 0; q < 4; q++) {
      if (q) __syncthreads();
      if constexpr (P == GEMM_FP32)
        gemm_tile_f32<64, 64, 2, 2>(pr, segs, slots, 2 * tl.tm + (q & 1), 2 * tl.tn + (q >> 1), (float *)lds);
      else
        gemm_tile_bf16<64, 64, 2, 2, P>(pr, segs, slots, 2 * tl.tm + (q & 1), 2 * tl.tn + (q >> 1), (unsigned *)lds);
    }
  }
  else if constexpr (P == GEMM_FP64) gemm_tile<TM, TN, WM, WN, KC>(pr, segs, slots, tl.tm, tl.tn, (double *)lds);
  else if constexpr (P == GEMM_FP32) gemm_tile_f32<TM, TN, WM, WN>(pr, segs, slots, tl.tm, tl.tn, (float *)lds);
  else gemm_tile_bf16<TM, TN, WM, WN, P>(pr, segs, slots, tl.tm, tl.tn, (unsigned *)lds);
}

template <int SH, int P>
__global__ void __launch_bounds__(256) k_gemm(const GemmProb *__restrict__ probs, const GemmSeg *__restrict__ segs,
                                              const GemmTile *__restrict__ tiles, int ntiles, Slots slots) {
  __shared__ __align__(16) unsigned char lds[tile_lds_bytes<SH, P>()];
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const GemmTile tl = tiles[w];
    if (tl.prob < 0) continue;  // padding of a short XCD queue (uniform over the work-group)
    const GemmProb pr = probs[tl.prob];
    if (pr.nseg < 0) continue;  // a problem left out of this list variant (uniform)
    run_tile<SH, P>(pr, segs, slots, tl, lds);
    __syncthreads();  // the next tile reuses the LDS staging area
  }
}

// The mixed launch: tiles of the shapes gemm_mixable() admits (64 x 64 and
// 32 x 32), each taking its problem's shape, in one persistent grid, so that
// a list whose problems want either shape still runs as one launch.  (Wider
// mixes raise the register count past four waves per SIMD.)
template <int P>
constexpr int mix_lds_bytes() {
  constexpr int a = tile_lds_bytes<SH_64x64, P>(), b = tile_lds_bytes<SH_32x32, P>();
  return a > b ? a : b;
}

template <int P, int KC = GEMM_TK>
__device__ __forceinline__ void gemm_mix_body(const GemmProb *__restrict__ probs, const GemmSeg *__restrict__ segs,
                                              const GemmTile *__restrict__ tiles, int ntiles, const Slots &slots) {
  __shared__ __align__(16) unsigned char lds[KC == GEMM_TK ? mix_lds_bytes<P>() : KC * (64 + 1 + 64 + 1) * 8];
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const GemmTile tl = tiles[w];
    if (tl.prob < 0) continue;  // padding of a short XCD queue (uniform over the work-group)
    const GemmProb pr = probs[tl.prob];
    if (pr.nseg < 0) continue;  // a problem left out of this list variant (uniform)
    if (pr.shape == SH_32x32) run_tile<SH_32x32, P, KC>(pr, segs, slots, tl, lds);
    else run_tile<SH_64x64, P, KC>(pr, segs, slots, tl, lds);
    __syncthreads();  // the next tile reuses the LDS staging area
  }
}
// four waves per SIMD (FP64, FP32, BF16x2 fit 128 registers without spills)
template <int P>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_gemm_mix(const GemmProb *__restrict__ probs, const GemmSeg *__restrict__ segs, const GemmTile *__restrict__ tiles,
           int ntiles, Slots slots) {
  gemm_mix_body<P>(probs, segs, tiles, ntiles, slots);
}
// FP64 in K chunks of 16 (16.6 KB of LDS, 78 registers): six waves per SIMD
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_gemm_mix_f64(const GemmProb *__restrict__ probs, const GemmSeg *__restrict__ segs, const GemmTile *__restrict__ tiles,
               int ntiles, Slots slots) {
  gemm_mix_body<GEMM_FP64, 16>(probs, segs, tiles, ntiles, slots);
}
// BF16x3, BF16x4: more slices in flight, left to the register allocator
template <int P>
__global__ void __launch_bounds__(256) k_gemm_mix_wide(const GemmProb *__restrict__ probs,
                                                       const GemmSeg *__restrict__ segs,
                                                       const GemmTile *__restrict__ tiles, int ntiles, Slots slots) {
  gemm_mix_body<P>(probs, segs, tiles, ntiles, slots);
}

template <int P>
static void launch_gemm_p(hipStream_t s, const GemmProb *probs, const GemmSeg *segs, const GemmTile *tiles,
                          int ntiles, const Slots &slots, int grid, int shape) {
  switch (shape) {
    case SH_64x64: hipLaunchKernelGGL((k_gemm<SH_64x64, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    case SH_128x64: hipLaunchKernelGGL((k_gemm<SH_128x64, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    case SH_32x32: hipLaunchKernelGGL((k_gemm<SH_32x32, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    case SH_128x16: hipLaunchKernelGGL((k_gemm<SH_128x16, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    case SH_64x16: hipLaunchKernelGGL((k_gemm<SH_64x16, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    case SH_16x64: hipLaunchKernelGGL((k_gemm<SH_16x64, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    case SH_16x128: hipLaunchKernelGGL((k_gemm<SH_16x128, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    case SH_64x16K: hipLaunchKernelGGL((k_gemm<SH_64x16K, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    case SH_128x128: hipLaunchKernelGGL((k_gemm<SH_128x128, P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
    default: break;
  }
}

void launch_gemm(hipStream_t s, const GemmProb *probs, const GemmSeg *segs, const GemmTile *tiles, int ntiles,
                 const Slots &slots, int grid, int shape, int prec) {
  if (ntiles <= 0) return;
  if (grid <= 0 || grid > ntiles) grid = ntiles;
  else if (grid >= NXCD) grid -= grid % NXCD;  // keeps tile slot t on XCD t % 8
  if (shape == SH_MIX) {
    switch (prec) {
#define QD_MIX(K, P) hipLaunchKernelGGL((K<P>), dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots)
      case GEMM_FP64: hipLaunchKernelGGL(k_gemm_mix_f64, dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots); break;
      case GEMM_FP32: QD_MIX(k_gemm_mix, GEMM_FP32); break;
      case GEMM_BF16X2: QD_MIX(k_gemm_mix, GEMM_BF16X2); break;
      case GEMM_BF16X3: QD_MIX(k_gemm_mix_wide, GEMM_BF16X3); break;
      case GEMM_BF16X4: QD_MIX(k_gemm_mix_wide, GEMM_BF16X4); break;
#undef QD_MIX
      default: break;
    }
    return;
  }
  switch (prec) {
    case GEMM_FP64: launch_gemm_p<GEMM_FP64>(s, probs, segs, tiles, ntiles, slots, grid, shape); break;
    case GEMM_FP32: launch_gemm_p<GEMM_FP32>(s, probs, segs, tiles, ntiles, slots, grid, shape); break;
    case GEMM_BF16X2: launch_gemm_p<GEMM_BF16X2>(s, probs, segs, tiles, ntiles, slots, grid, shape); break;
    case GEMM_BF16X3: launch_gemm_p<GEMM_BF16X3>(s, probs, segs, tiles, ntiles, slots, grid, shape); break;
    case GEMM_BF16X4: launch_gemm_p<GEMM_BF16X4>(s, probs, segs, tiles, ntiles, slots, grid, shape); break;
    default: break;
  }
}

// ---- Pandya product and combine in one pass (FP64)
//
// construct_cc_ints (src/ME_general.f90:923-952) needs D = XCC - YCC =
// h_H P - h_eta P^T with P = ETACC HCC^T per CC block.  One work-group takes
// a pair of 64x64 tiles (ti <= tj) of a block: it computes P_ij = E_i H_j^T
// and P_ji = E_j H_i^T (K staged through LDS in chunks of 8, the next chunk
// prefetched into registers, k ascending as in the grouped GEMM, so P is the
// same to the bit), then writes D_ij = h_H P_ij - h_eta P_ji^T, the
// transposed tile exchanged through LDS.  The product matrix P is never
// written.
//
// Layout of D ("run panels").  The device rows of a CC block come in runs,
// the rows whose first states share a channel; the Pandya term of one output
// box reads a dense (run x run) sub-block of D per term.  Each run of rows is
// stored as its own column-major panel (run length) x r at (run start) * r,
// so the sub-block of runs (a, b) is contiguous: element (x, y) at
// a_start * r + b_start * a_len + (x - a_start) + (y - b_start) * a_len.
// D^T = -h_H h_eta D, so only the sub-blocks with a_start <= b_start are
// written (the two-body pass reads a lower one as that sign times the
// transposed upper one), which halves D's write: the tile pair (ti < tj)
// writes all of D_ij and, from D_ji = -h_H h_eta D_ij^T, the few elements of
// the lower tile that lie in a diagonal sub-block (a run crossing the tile
// boundary); a diagonal tile writes its elements with a_start <= b_start.
// Four waves per SIMD (125 registers, 33 KB of LDS for the epilogue tile);
// chunks of 16 needed 160 registers and ran three.
//
// VLK: the 16 x 16 blocks of a1, a2 are transposed P blocks with lane l,
// component v at row l % 16 and column 4 (l / 16) + v (BF16 MFMA layout),
// else column 4 v + l / 16 (FP64 MFMA layout).  lds: at least 64 x 65 doubles.
template <bool VLK>
__device__ __forceinline__ void store_d_tile(const DevTab &t, int q, double *__restrict__ Dq, int r, int i0, int j0,
                                             bool diag, const d4 (&a1)[2][2], const d4 (&a2)[2][2],
                                             double *__restrict__ Tt, double hh, double eh) {
  constexpr int LDE = 65;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int m0 = (wave & 1) * 32, n0 = (wave >> 1) * 32;
  auto col_of = [&](int y, int v) { return n0 + y * 16 + (VLK ? 4 * lk + v : 4 * v + lk); };
  const uint32_t *__restrict__ runs = t.cc_run + t.cc_ab_start[q];
  // run start | length << 16 of the tile's rows and columns held by this lane
  uint32_t rrow[2], rcol[2][4];
#pragma unroll
  for (int x = 0; x < 2; x++) rrow[x] = i0 + m0 + x * 16 + lr < r ? runs[i0 + m0 + x * 16 + lr] : 0u;
#pragma unroll
  for (int y = 0; y < 2; y++)
#pragma unroll
    for (int v = 0; v < 4; v++) rcol[y][v] = j0 + col_of(y, v) < r ? runs[j0 + col_of(y, v)] : 0u;
  __syncthreads();
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) Tt[(m0 + x * 16 + lr) * LDE + col_of(y, v)] = diag ? a1[x][y][v] : a2[x][y][v];
  __syncthreads();
  double dv[2][2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int lx = m0 + x * 16 + lr, ly = col_of(y, v);
        dv[x][y][v] = hh * a1[x][y][v] - eh * Tt[ly * LDE + lx];
        const int row = i0 + lx, col = j0 + ly;
        const int L = (int)(rrow[x] & 0xffffu), l = (int)(rrow[x] >> 16);
        if (row < r && col < r && L <= (int)(rcol[y][v] & 0xffffu))
          Dq[(int64_t)L * r + (int64_t)col * l + (row - L)] = dv[x][y][v];
      }
  // rows of tile tj in a run that starts inside tile ti: the lower tile's
  // elements of diagonal sub-blocks, D(j0 + a, i0 + b) = -hh eh D_ij(b, a)
  if (diag || j0 >= r || (int)(runs[j0] & 0xffffu) > i0 + 63) return;
  __syncthreads();
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) Tt[(m0 + x * 16 + lr) * LDE + col_of(y, v)] = dv[x][y][v];
  __syncthreads();
  const double ds = -hh * eh;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int a = m0 + x * 16 + lr, b = col_of(y, v), row = j0 + a, col = i0 + b;
        if (row < r && col < r) {
          const uint32_t ru = runs[row];
          const int L = (int)(ru & 0xffffu), l = (int)(ru >> 16);
          if (L == (int)(runs[col] & 0xffffu)) Dq[(int64_t)L * r + (int64_t)col * l + (row - L)] = ds * Tt[b * LDE + a];
        }
      }
}

// D_ij of one tile pair from the CC sets in FP64 (layout and epilogue above)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_pandya_fused(DevTab t, const double *__restrict__ E,
                                                      const double *__restrict__ H, double *__restrict__ D,
                                                      const TilePair *__restrict__ tp, double hh, double eh) {
  // panel rows 80 doubles apart: a ds_read_b64 of 16 rows x 2 k-steps per
  // half-wave hits 32 distinct bank pairs; the transposed tile of the
  // epilogue keeps rows of 65 and sets the LDS size
  constexpr int T = 64, TK = 8, LD = T + 16, LDE = T + 1, NU = T * TK / 256;
  constexpr int NLDS = 4 * TK * LD > T * LDE ? 4 * TK * LD : T * LDE;
  __shared__ double lds[NLDS];  // E_i, E_j, H_i, H_j as [TK][LD]; later one [T][LDE] tile
  const TilePair pr = tp[blockIdx.x];
  const int q = pr.blk, r = t.cc_r[q], K = t.cc_nb[q];
  const double *__restrict__ Eq = E + t.cc_off[q];
  const double *__restrict__ Hq = H + t.cc_off[q];
  double *__restrict__ Dq = D + t.cc_poff[q];
  const bool diag = pr.ti == pr.tj;
  const int i0 = pr.ti * T, j0 = pr.tj * T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int m0 = (wave & 1) * 32, n0 = (wave >> 1) * 32;
  // a diagonal pair (ti == tj) stages only the i panels and reads H_j from H_i
  double *Ei = lds, *Ej = lds + TK * LD, *Hi = lds + 2 * TK * LD, *Hj = diag ? Hi : lds + 3 * TK * LD;
  d4 a1[2][2], a2[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) a1[x][y] = a2[x][y] = d4{0, 0, 0, 0};
  // staging: element (row x, k) of a panel is M[x + k r] (column-major r x K)
  double rg[4][NU];
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NU; u++) {
      const int idx = tid + 256 * u, x = idx % T, k = idx / T, gk = k0 + k;
      const bool kin = gk < K;
      rg[0][u] = (kin && i0 + x < r) ? Eq[(i0 + x) + (int64_t)gk * r] : 0.0;
      rg[1][u] = (!diag && kin && j0 + x < r) ? Eq[(j0 + x) + (int64_t)gk * r] : 0.0;
      rg[2][u] = (kin && i0 + x < r) ? Hq[(i0 + x) + (int64_t)gk * r] : 0.0;
      rg[3][u] = (!diag && kin && j0 + x < r) ? Hq[(j0 + x) + (int64_t)gk * r] : 0.0;
    }
  };
  load(0);
  for (int k0 = 0; k0 < K; k0 += TK) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NU; u++) {
      const int idx = tid + 256 * u, x = idx % T, k = idx / T;
      Ei[k * LD + x] = rg[0][u];
      Hi[k * LD + x] = rg[2][u];
      if (!diag) {
        Ej[k * LD + x] = rg[1][u];
        Hj[k * LD + x] = rg[3][u];
      }
    }
    __syncthreads();
    if (k0 + TK < K) load(k0 + TK);
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      if (k0 + kk >= K) break;  // zero fill past K adds nothing (14% of the k-steps at N=20 R=14)
      double ei[2], hj[2], ej[2], hi[2];
#pragma unroll
      for (int x = 0; x < 2; x++) {
        ei[x] = Ei[(kk + lk) * LD + m0 + x * 16 + lr];
        hj[x] = Hj[(kk + lk) * LD + n0 + x * 16 + lr];
        ej[x] = Ej[(kk + lk) * LD + m0 + x * 16 + lr];
        hi[x] = Hi[(kk + lk) * LD + n0 + x * 16 + lr];
      }
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
          a1[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(hj[y], ei[x], a1[x][y], 0, 0, 0);
          if (!diag) a2[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(hi[y], ej[x], a2[x][y], 0, 0, 0);
        }
    }
  }
  // epilogue; the MFMAs take H first, so they hold the tiles transposed:
  // lane l, component v is row l % 16, column 4 v + l / 16 of the P tile,
  // and each store covers four columns of 16 consecutive rows of D
  store_d_tile<false>(t, q, Dq, r, i0, j0, diag, a1, a2, lds, hh, eh);
}

// The same fused pass with P = E H^T in BF16 slices (split precision S,
// numerics of gemm_tile_bf16: K stages of 32 from k = 0, each stage's slice
// products summed in FP32 smallest first, stage sums added into FP64).
// Both CC sets are first cut into slice planes by k_cc_slices (one pass over
// each set), so the tiles, which read every E and H row panel once per tile
// pair it belongs to, stage ready slices: one 16-byte load per slice, row and
// 8 k, no slicing arithmetic.  The two products of a tile pair run one after
// the other, each with its own K walk over two staged panels (E rows, H
// rows); P_ji stays in registers until the epilogue.  The MFMAs take the H
// slices first, so they hold each 16 x 16 block transposed: lane l,
// component v is row l % 16, column 4 (l / 16) + v.
template <int S>
__global__ void __launch_bounds__(256) k_cc_slices(DevTab t, const double *__restrict__ cc,
                                                   unsigned short *__restrict__ planes, int64_t len,
                                                   const int64_t *__restrict__ off, const int2 *__restrict__ units) {
  const int2 w = units[blockIdx.x];
  const int q = w.x, r = t.cc_r[q], K = t.cc_nb[q];
  const int64_t u = (int64_t)w.y + threadIdx.x;
  if (u >= (int64_t)((K + 7) / 8) * r) return;
  const int kb = (int)(u / r), x = (int)(u % r);
  const double *__restrict__ src = cc + t.cc_off[q] + x;
  double v[8];
#pragma unroll
  for (int e = 0; e < 8; e++) v[e] = 8 * kb + e < K ? src[(int64_t)(8 * kb + e) * r] : 0.0;
  unsigned wv[S][4];
#pragma unroll
  for (int e = 0; e < 4; e++) {
    unsigned h0[S], h1[S];
    bf16_slices<S>(v[2 * e], h0);
    bf16_slices<S>(v[2 * e + 1], h1);
#pragma unroll
    for (int sl = 0; sl < S; sl++) wv[sl][e] = h0[sl] | (h1[sl] << 16);
  }
#pragma unroll
  for (int sl = 0; sl < S; sl++)
    *(u4 *)(planes + sl * len + off[q] + 8 * u) = u4{wv[sl][0], wv[sl][1], wv[sl][2], wv[sl][3]};
}

template <int S>
__device__ __forceinline__ void pandya_split_product(const unsigned short *__restrict__ Ep,
                                                     const unsigned short *__restrict__ Hp, int64_t len, int r, int K,
                                                     int e0, int h0, d4 (&acc)[2][2], unsigned *__restrict__ lds) {
  constexpr int T = 64;
  unsigned *Es = lds;                       // [S][T][SPLIT_LDW]: slice s of E(e0 + x, k), k pairs packed
  unsigned *Hs = lds + S * T * SPLIT_LDW;   // [S][T][SPLIT_LDW]: slice s of H(h0 + x, k)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int m0 = (wave & 1) * 32, n0 = (wave >> 1) * 32;
  const int x = tid & 63, j = tid >> 6, nkb = (K + 7) / 8;  // staging: row x, k unit j of the stage
  u4 re[S], rh[S];
  auto load = [&](int k0) {
    const int kb = k0 / 8 + j;
    const bool ein = kb < nkb && e0 + x < r, hin = kb < nkb && h0 + x < r;
#pragma unroll
    for (int sl = 0; sl < S; sl++) {
      re[sl] = ein ? *(const u4 *)(Ep + sl * len + 8 * ((int64_t)kb * r + e0 + x)) : u4{0, 0, 0, 0};
      rh[sl] = hin ? *(const u4 *)(Hp + sl * len + 8 * ((int64_t)kb * r + h0 + x)) : u4{0, 0, 0, 0};
    }
  };
  load(0);
  for (int k0 = 0; k0 < K; k0 += SPLIT_TK) {
    __syncthreads();
#pragma unroll
    for (int sl = 0; sl < S; sl++) {
      *(u4 *)(Es + (sl * T + x) * SPLIT_LDW + 4 * j) = re[sl];
      *(u4 *)(Hs + (sl * T + x) * SPLIT_LDW + 4 * j) = rh[sl];
    }
    __syncthreads();
    if (k0 + SPLIT_TK < K) load(k0 + SPLIT_TK);
    bf8 a[S][2], b[S][2];
#pragma unroll
    for (int sl = 0; sl < S; sl++)
#pragma unroll
      for (int y = 0; y < 2; y++) {
        a[sl][y] = __builtin_bit_cast(bf8, *(const u4 *)(Es + (sl * T + m0 + y * 16 + lr) * SPLIT_LDW + 4 * lk));
        b[sl][y] = __builtin_bit_cast(bf8, *(const u4 *)(Hs + (sl * T + n0 + y * 16 + lr) * SPLIT_LDW + 4 * lk));
      }
#pragma unroll
    for (int xx = 0; xx < 2; xx++)
#pragma unroll
      for (int y = 0; y < 2; y++) {
        f4 c = f4{0, 0, 0, 0};
#pragma unroll
        for (int d = S - 1; d >= 0; d--)
#pragma unroll
          for (int i = 0; i <= d; i++) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[d - i][y], a[i][xx], c, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) acc[xx][y][v] += (double)c[v];
      }
  }
}

template <int S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S <= 3 ? 3 : 2)))
k_pandya_fused_bf16(DevTab t, const unsigned short *__restrict__ planes,
                                                           int64_t len, const int64_t *__restrict__ off,
                                                           double *__restrict__ D, const TilePair *__restrict__ tp,
                                                           double hh, double eh) {
  constexpr int T = 64, LDE = T + 1;
  constexpr int NSTAGE = 2 * S * T * SPLIT_LDW * 4, NEPI = T * LDE * 8;
  __shared__ __align__(16) unsigned char lds[NSTAGE > NEPI ? NSTAGE : NEPI];
  const TilePair pr = tp[blockIdx.x];
  const int q = pr.blk, r = t.cc_r[q], K = t.cc_nb[q];
  const unsigned short *Ep = planes + off[q], *Hp = planes + 4 * len + off[q];
  double *__restrict__ Dq = D + t.cc_poff[q];
  const bool diag = pr.ti == pr.tj;
  const int i0 = pr.ti * T, j0 = pr.tj * T;
  d4 a1[2][2], a2[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) a1[x][y] = a2[x][y] = d4{0, 0, 0, 0};
  if (!diag) pandya_split_product<S>(Ep, Hp, len, r, K, j0, i0, a2, (unsigned *)lds);  // P_ji = E_j H_i^T
  pandya_split_product<S>(Ep, Hp, len, r, K, i0, j0, a1, (unsigned *)lds);             // P_ij = E_i H_j^T
  // epilogue of k_pandya_fused: D_ij = hh P_ij - eh P_ji^T, P_ji through LDS
  store_d_tile<true>(t, q, Dq, r, i0, j0, diag, a1, a2, (double *)lds, hh, eh);
}

template <int S>
static void launch_pandya_bf16(hipStream_t s, const DevTab &t, const double *etacc, const double *hcc, double *D,
                               const TilePair *tp, int ntp, double hh, double eh, const CcSlices &cs) {
  if (cs.nunits > 0) {
    hipLaunchKernelGGL(k_cc_slices<S>, dim3(cs.nunits), dim3(CS_UNITS_WG), 0, s, t, etacc, cs.planes, cs.len, cs.off,
                       cs.units);
    hipLaunchKernelGGL(k_cc_slices<S>, dim3(cs.nunits), dim3(CS_UNITS_WG), 0, s, t, hcc, cs.planes + 4 * cs.len,
                       cs.len, cs.off, cs.units);
  }
  hipLaunchKernelGGL(k_pandya_fused_bf16<S>, dim3(ntp), dim3(256), 0, s, t, cs.planes, cs.len, cs.off, D, tp, hh, eh);
}

void launch_pandya_fused(hipStream_t s, const DevTab &t, const double *etacc, const double *hcc, double *D,
                         const TilePair *tp, int ntp, double hh, double eh, int prec, const CcSlices &cs) {
  if (ntp <= 0) return;
  switch (prec) {
    case GEMM_FP64: hipLaunchKernelGGL(k_pandya_fused, dim3(ntp), dim3(256), 0, s, t, etacc, hcc, D, tp, hh, eh); break;
    case GEMM_BF16X2: launch_pandya_bf16<2>(s, t, etacc, hcc, D, tp, ntp, hh, eh, cs); break;
    case GEMM_BF16X3: launch_pandya_bf16<3>(s, t, etacc, hcc, D, tp, ntp, hh, eh, cs); break;
    case GEMM_BF16X4: launch_pandya_bf16<4>(s, t, etacc, hcc, D, tp, ntp, hh, eh, cs); break;
    default: break;
  }
}

bool pandya_fused_prec(int prec) { return prec == GEMM_FP64 || prec >= GEMM_BF16X2; }

// Small products (the one-body blocks of commutator_111, K <= a few hundred):
// one 16 x 16 output tile per work-group and one output per thread, K staged
// through LDS in chunks of 192 (one global round trip for K <= 192), plain
// FP64 FMA in a fixed order.  Many small work-groups keep the latency of the K
// walk short where the MFMA tiles of k_gemm would leave most of the chip idle.
constexpr int SMALL_KC = 192;
__global__ void __launch_bounds__(256) k_gemm_small(const GemmProb *__restrict__ probs,
                                                    const GemmSeg *__restrict__ segs,
                                                    const GemmTile *__restrict__ tiles, Slots slots) {
  __shared__ double As[16][SMALL_KC + 1], Bs[SMALL_KC][17];
  const GemmTile tl = tiles[blockIdx.x];
  const GemmProb p = probs[tl.prob];
  const int ti = threadIdx.x & 15, tj = threadIdx.x >> 4;
  const int i0 = tl.tm * 16, j0 = tl.tn * 16;
  double acc = 0.0;
  for (int s = 0; s < p.nseg; s++) {
    const GemmSeg sg = segs[p.seg0 + s];
    const double *__restrict__ A = slots.p[sg.a_slot] + sg.a_off;
    const double *__restrict__ B = slots.p[sg.b_slot] + sg.b_off;
    double part = 0.0;
    for (int k0 = 0; k0 < sg.K; k0 += SMALL_KC) {
      for (int idx = threadIdx.x; idx < 16 * SMALL_KC; idx += 256) {
        // the contiguous index of each operand runs fastest across the lanes
        const int i = sg.ta ? idx / SMALL_KC : idx % 16, k = sg.ta ? idx % SMALL_KC : idx / 16;
        const int gi = i0 + i, gk = k0 + k;
        double v = 0.0;
        if (gi < p.M && gk < sg.K) v = sg.ta ? A[gk + (int64_t)gi * sg.lda] : A[gi + (int64_t)gk * sg.lda];
        As[i][k] = v;
        const int j = sg.tb ? idx % 16 : idx / SMALL_KC, kb = sg.tb ? idx / 16 : idx % SMALL_KC;
        const int gj = j0 + j, gkb = k0 + kb;
        double w = 0.0;
        if (gj < p.N && gkb < sg.K) w = sg.tb ? B[gj + (int64_t)gkb * sg.ldb] : B[gkb + (int64_t)gj * sg.ldb];
        Bs[kb][j] = w;
      }
      __syncthreads();
      const int kn = sg.K - k0 < SMALL_KC ? sg.K - k0 : SMALL_KC;
      for (int k = 0; k < kn; k++) part += As[ti][k] * Bs[k][tj];
      __syncthreads();
    }
    acc += sg.alpha * part;
  }
  const int gi = i0 + ti, gj = j0 + tj;
  if (gi < p.M && gj < p.N) {
    double *C = slots.p[p.c_slot] + p.c_off + gi + (int64_t)gj * p.ldc;
    *C = p.beta1 ? *C + acc : acc;
  }
}

void launch_gemm_small(hipStream_t s, const GemmProb *probs, const GemmSeg *segs, const GemmTile *tiles,
                       int ntiles, const Slots &slots) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_gemm_small, dim3(ntiles), dim3(256), 0, s, probs, segs, tiles, slots);
}

}  // namespace qd
