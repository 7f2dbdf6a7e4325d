// Hartree-Fock on the device: the Fock build of the SCF iteration and the
// basis transform of the two-body blocks (hf_device.hip), driven by the host
// SCF of hf.cpp.
#pragma once
#include <cstdint>
#include <vector>

#include "qd_internal.h"

namespace qd {

// The Coulomb table <ab|V|cd> (hw = 1) in the compressed layout of hf.cpp's
// Spatial: element (a, b, c, d) at tab[start[(a N + b) N + c] + pos_in_m[d]]
// when m_a + m_b = m_c + m_d.  exch_fallback: the ME.bin rule (exchange from
// <ba|V|cd> where <ab|V|dc> is below 1e-10).
struct CoulombView {
  int N;
  const double *tab;
  int64_t tab_size;
  const int64_t *start;
  const int *pos_in_m, *m;
  bool exch_fallback;
};

struct HfDevice;
// Uploads the table and the spin-orbital maps (sp: spatial orbital, spin: +-1,
// per state) on the context's stream.
int hf_device_create(qd_context *ctx, const CoulombView &S, const std::vector<int> &sp, const std::vector<int> &spin,
                     double shw, HfDevice **out);
void hf_device_destroy(HfDevice *d);
// F = diag(Tm) + the direct-minus-exchange mean field of the density rho
// (M x M row-major, block-diagonal by channel); F returned row-major M x M,
// only same-channel entries written (others 0)
int hf_device_fock(HfDevice *d, const std::vector<double> &rho, const std::vector<double> &Tm, std::vector<double> &F);
// The two-body blocks of H(0) in the HF basis written into the operator out:
// per block the pairs (hh | ph | pp of the layout), for each pair x the sparse
// row of the coefficient matrix T (columns z = pair indices of the same block,
// values), H = T V T^T with V[z][y] = <z|V|y>_A.
struct SparseRows {
  std::vector<int> row_start;  // per pair of all blocks (global pair order) + 1
  std::vector<int> col;        // pair index within its block
  std::vector<double> val;
};
int hf_device_transform(HfDevice *d, const SparseRows &T, qd_op *out);
// out[0] = eh - sum over blocks of the Vhhhh diagonal of out (HF_E2)
int hf_device_e0(qd_context *ctx, double eh, qd_op *out);

}  // namespace qd
