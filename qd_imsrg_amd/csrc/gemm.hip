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
// autotuner in capi.hip); the kernel dispatches on it at run time.  The launch
// is persistent: a fixed grid strides over the tile list, which is sorted by
// decreasing work.  Work-group = 256 threads = 4 waves; K is staged through
// LDS in chunks of 16 (4 MFMA k-steps).  MFMA operand layout: A(16x4) lane l
// holds A[l%16][l/16], B(4x16) lane l holds B[l/16][l%16], D(16x16) lane l
// holds D[4*v + l/16][l%16] in component v (FP64: one row per group).
#include "qd_kernels.h"

namespace qd {

typedef double d4 __attribute__((ext_vector_type(4)));

// shape table: TM x TN tile, WM x WN waves, each wave FM x FN MFMA tiles
template <int TM, int TN, int WM, int WN>
__device__ __forceinline__ void gemm_tile(const GemmProb &pr, const GemmSeg *__restrict__ segs, const Slots &slots,
                                          int tm, int tn, double *__restrict__ lds) {
  constexpr int FM = TM / WM / 16, FN = TN / WN / 16;
  double *As = lds;                       // [GEMM_TK][TM + 1]
  double *Bs = lds + GEMM_TK * (TM + 1);  // [GEMM_TK][TN + 1]
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

  for (int si = 0; si < pr.nseg; si++) {
    const GemmSeg sg = segs[pr.seg0 + si];
    const double *__restrict__ A = slots.p[sg.a_slot] + sg.a_off;
    const double *__restrict__ B = slots.p[sg.b_slot] + sg.b_off;
    for (int k0 = 0; k0 < sg.K; k0 += GEMM_TK) {
#pragma unroll
      for (int r = 0; r < (TM * GEMM_TK + 255) / 256; r++) {
        const int idx = tid + 256 * r;
        if (idx < TM * GEMM_TK) {
          int i, k;
          if (!sg.ta) { i = idx % TM; k = idx / TM; }  // columns of A contiguous in i
          else { k = idx % GEMM_TK; i = idx / GEMM_TK; }  // A stored transposed
          const int gi = row0 + i, gk = k0 + k;
          double v = 0.0;
          if (gi < pr.M && gk < sg.K) v = sg.ta ? A[gk + (int64_t)gi * sg.lda] : A[gi + (int64_t)gk * sg.lda];
          As[k * (TM + 1) + i] = sg.alpha * v;
        }
      }
#pragma unroll
      for (int r = 0; r < (TN * GEMM_TK + 255) / 256; r++) {
        const int idx = tid + 256 * r;
        if (idx < TN * GEMM_TK) {
          int j, k;
          if (!sg.tb) { k = idx % GEMM_TK; j = idx / GEMM_TK; }  // columns of B contiguous in k
          else { j = idx % TN; k = idx / TN; }                   // B stored transposed
          const int gj = col0 + j, gk = k0 + k;
          double v = 0.0;
          if (gj < pr.N && gk < sg.K) v = sg.tb ? B[gj + (int64_t)gk * sg.ldb] : B[gk + (int64_t)gj * sg.ldb];
          Bs[k * (TN + 1) + j] = v;
        }
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < GEMM_TK; kk += 4) {
        double a[FM], b[FN];
#pragma unroll
        for (int x = 0; x < FM; x++) a[x] = As[(kk + lk) * (TM + 1) + m0 + x * 16 + lr];
#pragma unroll
        for (int y = 0; y < FN; y++) b[y] = Bs[(kk + lk) * (TN + 1) + n0 + y * 16 + lr];
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
          for (int y = 0; y < FN; y++) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
      }
      __syncthreads();
    }
  }
  double *__restrict__ C = slots.p[pr.c_slot] + pr.c_off;
#pragma unroll
  for (int x = 0; x < FM; x++)
#pragma unroll
    for (int y = 0; y < FN; y++) {
      const int c = col0 + n0 + y * 16 + lr;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = row0 + m0 + x * 16 + 4 * v + lk;
        if (r < pr.M && c < pr.N) {
          double *p = C + r + (int64_t)c * pr.ldc;
          *p = pr.beta1 ? *p + acc[x][y][v] : acc[x][y][v];
        }
      }
    }
}

__global__ void __launch_bounds__(256) k_gemm(const GemmProb *__restrict__ probs, const GemmSeg *__restrict__ segs,
                                              const GemmTile *__restrict__ tiles, int ntiles, Slots slots) {
  __shared__ double lds[GEMM_TK * (128 + 1) + GEMM_TK * (64 + 1) + 64];
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const GemmTile tl = tiles[w];
    const GemmProb pr = probs[tl.prob];
    switch (pr.shape) {
      case SH_64x64: gemm_tile<64, 64, 2, 2>(pr, segs, slots, tl.tm, tl.tn, lds); break;
      case SH_128x64: gemm_tile<128, 64, 2, 2>(pr, segs, slots, tl.tm, tl.tn, lds); break;
      case SH_32x32: gemm_tile<32, 32, 2, 2>(pr, segs, slots, tl.tm, tl.tn, lds); break;
      case SH_128x16: gemm_tile<128, 16, 4, 1>(pr, segs, slots, tl.tm, tl.tn, lds); break;
      case SH_64x16: gemm_tile<64, 16, 4, 1>(pr, segs, slots, tl.tm, tl.tn, lds); break;
      case SH_16x64: gemm_tile<16, 64, 1, 4>(pr, segs, slots, tl.tm, tl.tn, lds); break;
      default: gemm_tile<16, 128, 1, 4>(pr, segs, slots, tl.tm, tl.tn, lds); break;
    }
    __syncthreads();  // the next tile reuses the LDS staging area
  }
}

void launch_gemm(hipStream_t s, const GemmProb *probs, const GemmSeg *segs, const GemmTile *tiles, int ntiles,
                 const Slots &slots, int grid) {
  if (ntiles <= 0) return;
  if (grid <= 0 || grid > ntiles) grid = ntiles;
  hipLaunchKernelGGL(k_gemm, dim3(grid), dim3(256), 0, s, probs, segs, tiles, ntiles, slots);
}

}  // namespace qd
