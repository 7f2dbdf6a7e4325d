// Grouped FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// One launch runs every (problem, 64x64 output tile) of a work list; the
// problems are the per-(M_L, M_S)-block products of matrix_mult222
// (src/commutators.f90:999-1127), commutator_221 (:236-272) and
// construct_cc_ints (src/ME_general.f90:948-949).  A problem sums up to four
// products op(A)op(B) (alpha = +-1, applied while staging A), so several
// reference dgemm calls that add into one sub-block become one pass.
//
// Work-group: 256 threads = 4 waves in a 2x2 grid, each wave owns a 32x32
// sub-tile = 2x2 MFMA 16x16 accumulators.  K is staged through LDS in chunks
// of 16 (4 MFMA k-steps).  MFMA operand layout: A(16x4) lane l holds
// A[l%16][l/16], B(4x16) lane l holds B[l/16][l%16], D(16x16) lane l holds
// D[4*v + l/16][l%16] in component v (FP64: one row per group, four groups).
#include "qd_kernels.h"

namespace qd {

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_gemm(const GemmProb *__restrict__ probs,
                                              const GemmSeg *__restrict__ segs,
                                              const GemmTile *__restrict__ tiles, Slots slots) {
  __shared__ double As[GEMM_TK][GEMM_TM + 1];
  __shared__ double Bs[GEMM_TK][GEMM_TN + 1];
  const GemmTile tl = tiles[blockIdx.x];
  const GemmProb pr = probs[tl.prob];
  const int row0 = tl.tm * GEMM_TM, col0 = tl.tn * GEMM_TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc00 = {0, 0, 0, 0}, acc01 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};

  for (int si = 0; si < pr.nseg; si++) {
    const GemmSeg sg = segs[pr.seg0 + si];
    const double *__restrict__ A = slots.p[sg.a_slot] + sg.a_off;
    const double *__restrict__ B = slots.p[sg.b_slot] + sg.b_off;
    for (int k0 = 0; k0 < sg.K; k0 += GEMM_TK) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int idx = tid + 256 * r;
        int i, k;
        if (!sg.ta) { i = idx & 63; k = idx >> 6; }   // columns of A contiguous in i
        else { k = idx & 15; i = idx >> 4; }          // A stored transposed: contiguous in k
        const int gi = row0 + i, gk = k0 + k;
        double v = 0.0;
        if (gi < pr.M && gk < sg.K) v = sg.ta ? A[gk + (int64_t)gi * sg.lda] : A[gi + (int64_t)gk * sg.lda];
        As[k][i] = sg.alpha * v;
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int idx = tid + 256 * r;
        int j, k;
        if (!sg.tb) { k = idx & 15; j = idx >> 4; }   // columns of B contiguous in k
        else { j = idx & 63; k = idx >> 6; }          // B stored transposed: contiguous in j
        const int gj = col0 + j, gk = k0 + k;
        double v = 0.0;
        if (gj < pr.N && gk < sg.K) v = sg.tb ? B[gj + (int64_t)gk * sg.ldb] : B[gk + (int64_t)gj * sg.ldb];
        Bs[k][j] = v;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < GEMM_TK; kk += 4) {
        const double a0 = As[kk + lk][wm * 32 + lr], a1 = As[kk + lk][wm * 32 + 16 + lr];
        const double b0 = Bs[kk + lk][wn * 32 + lr], b1 = Bs[kk + lk][wn * 32 + 16 + lr];
        acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
        acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
      }
      __syncthreads();
    }
  }
  double *__restrict__ C = slots.p[pr.c_slot] + pr.c_off;
  const d4 *accs[4] = {&acc00, &acc01, &acc10, &acc11};
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int mi = t >> 1, ni = t & 1;
    const d4 a = *accs[t];
    const int c = col0 + wn * 32 + ni * 16 + lr;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = row0 + wm * 32 + mi * 16 + 4 * v + lk;
      if (r < pr.M && c < pr.N) {
        double *p = C + r + (int64_t)c * pr.ldc;
        *p = pr.beta1 ? *p + a[v] : a[v];
      }
    }
  }
}

void launch_gemm(hipStream_t s, const GemmProb *probs, const GemmSeg *segs, const GemmTile *tiles, int ntiles,
                 const Slots &slots) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_gemm, dim3(ntiles), dim3(256), 0, s, probs, segs, tiles, slots);
}

}  // namespace qd
