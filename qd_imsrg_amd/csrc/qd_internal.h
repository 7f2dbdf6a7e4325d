// Internal declarations shared by the host channel builder, the kernels and
// the C ABI.  Not installed; include/qdimsrg.h is the public surface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qdimsrg.h"

namespace qd {

// sub-block order inside one two-body block = the reference's vectorize order
enum Sub { HHHH = 0, PPPP = 1, PPHH = 2, PHPH = 3, PPPH = 4, PHHH = 5 };
enum PairType { T_HH = 1, T_PH = 2, T_PP = 3 };

// Packed state pair: i in the low 16 bits, j in the high 16 bits (M <= 65535).
__host__ __device__ inline uint32_t pack2(int i, int j) { return (uint32_t)i | ((uint32_t)j << 16); }
__host__ __device__ inline int lo16(uint32_t p) { return (int)(p & 0xffffu); }
__host__ __device__ inline int hi16(uint32_t p) { return (int)(p >> 16); }
// Packed i_array entry: type (2 bits) | block (10 bits) | index within its class (20 bits).
__host__ __device__ inline uint32_t pack_ia(int type, int blk, int idx) {
  return ((uint32_t)type << 30) | ((uint32_t)blk << 20) | (uint32_t)idx;
}
__host__ __device__ inline int ia_type(uint32_t a) { return (int)(a >> 30); }
__host__ __device__ inline int ia_blk(uint32_t a) { return (int)((a >> 20) & 0x3ffu); }
__host__ __device__ inline int ia_idx(uint32_t a) { return (int)(a & 0xfffffu); }

// Host-side channel structure (built once per context).
struct Channels {
  int n = 0, R = 0, M = 0, mp = 0, nblock = 0, ncc = 0;
  int64_t neq = 0, cc_size = 0, p_size = 0;  // p_size: sum r*r (Pandya product)
  int64_t off_fhh = 0, off_fpp = 0, off_fph = 0, off_2b = 0;
  std::vector<int> ml, ms;                  // per state
  std::vector<int> nh, np, nb;              // per block
  std::vector<int64_t> off;                 // per block x 6
  std::vector<uint32_t> pairs;              // all pair lists: per block hh, ph, pp
  std::vector<int> hh_start, ph_start, pp_start;
  std::vector<uint32_t> iarr;               // M(M-1)/2
  // single-particle channels (m_l, m_s): states in ascending order
  int nch = 0;
  std::vector<int> spch, ch_start, ch_states, ch_lookup;  // ch_lookup[(ml+ML0)*2 + (ms>0)]
  int ml0 = 0, mlspan = 0;
  // cross-coupled blocks
  std::vector<int> cc_r, cc_nb;
  std::vector<int64_t> cc_off, cc_poff;
  std::vector<uint32_t> cc_ab, cc_hp;
  std::vector<int> cc_ab_start, cc_hp_start;
  std::vector<int> cc_map;                  // M*M x 2 (block,row), -1 if none
  // Device row order of every CC block: rows sorted by (channel of i, i, j),
  // so the rows (i, j) with i in one single-particle channel (and hence j in
  // one channel) are contiguous.  cc_map, cc_ab and the qd_cc host transfers
  // keep the reference's order; the device tables use this one.
  std::vector<int> cc_dev_row;              // per reference row (cc_ab_start + row): device row
  std::vector<uint32_t> cc_ab_dev;          // pairs in device row order
  // cc_rs[K * nch + c]: first device row of block K whose first state is in
  // channel c (-1 if none).  The device row of (i, j) is then
  // cc_rs[K * nch + ch(i)] + pos(i) * |ch(j)| + pos(j), pos = rank in the channel.
  std::vector<int> cc_rs;
  // per device row (cc_ab_start + row): first row of its run | run length << 16
  // (a run = the rows whose first states share a channel)
  std::vector<uint32_t> cc_run;
  std::vector<int> sp_pos;                  // per state: rank within its channel
  // Class pairs of the two-body pass: per block, the unordered pairs {A, B}
  // (A <= B) of single-particle channels whose state pairs lie in the block.
  // Each has a grid of |A| x |B| ordered (x1, x2), x1 the pos-th state of A
  // and x2 of B, row-major; an entry packs the combined index of the stored
  // pair (min, max) | (x1 > x2) << 16 | (x1 != x2) << 17 | pair type << 18.
  std::vector<int> cp_start;                // per block: first class pair (nblock + 1)
  std::vector<int> cp_a, cp_b, cp_grid_off; // per class pair
  std::vector<int> cp_tmin, cp_tmax;        // smallest / largest pair type (0 hh, 1 ph, 2 pp)
  std::vector<uint32_t> cp_grid;
  // one-body matrices blocked by channel: |c| x |c| row-major at fcls_off[c]
  std::vector<int> fcls_off;
  int fcls_size = 0;
};

// Builds the tables; returns an empty string on success, else the reason.
std::string build_channels(int nbody, int nshell, Channels &c);

// Error text for qd_last_error (thread-local); returns code.
int set_error(int code, const char *fmt, ...);
// Accessors for sources outside capi.hip
hipStream_t context_stream(qd_context *c);
int64_t context_neq(const qd_context *c);
qd_context *op_context(qd_op *o);
double *op_data(qd_op *o);
const Channels &context_channels(const qd_context *c);
struct DevTab;
const DevTab &context_devtab(const qd_context *c);
const double *op_cdata(const qd_op *o);
int op_herm(const qd_op *o);
// cyclic Jacobi eigensolver for a small symmetric matrix (row-major a[n*n]);
// eigenvalues ascending in w, eigenvectors in the columns of v (hf.cpp)
void jacobi(int n, std::vector<double> a, std::vector<double> &w, std::vector<double> &v);

// Device copy of the tables, passed by value to kernels.
struct DevTab {
  int n, M, mp, nblock, ncc, nch, ml0, mlspan;
  int64_t neq, off_fhh, off_fpp, off_fph;
  const int *nh, *np, *nb;
  const int64_t *off;
  const uint32_t *pairs;
  const int *hh_start, *ph_start, *pp_start;
  const uint32_t *iarr;
  const int *spch, *ch_start, *ch_states, *ch_lookup;
  const int *ml, *ms;
  const int *cc_r, *cc_nb;
  const int64_t *cc_off, *cc_poff;
  const uint32_t *cc_ab, *cc_hp;
  const int *cc_ab_start, *cc_hp_start;
  const int2 *cc_map;
  const int *cc_rs;
  const int *cp_a, *cp_b, *cp_grid_off;
  const uint32_t *cp_grid;
  const int *fcls_off;
  int fcls_size;
  // calc_cc as a table (null when a flat index does not fit 30 bits): per
  // CC element in device order, the flat index of the V element it reads |
  // transposed (herm factor) << 30 | pair-sorting sign << 31; CC_ZERO in the
  // index bits for an element that is zero
  const uint32_t *cc_gmap;
  int64_t cc_size;
  int64_t p_size;  // Pandya product entries; element p_size of D is a zero
  const int *cc_rowblk;  // per CC row (M * M, rows of the blocks in order): its block
  const uint32_t *cc_run;  // per CC row: first row of its run | run length << 16
  // per entry of the box list (nblist of them), for the four Pandya terms of
  // the box in its task: the offset in D of the term's sub-block (-1: none)
  // and its leading dimension | stored transposed << 30; [layout][entry][4],
  // layout 0 whole column-major D, 1 run panels
  const int64_t *box_pb;
  const int *box_pl;
  int nblist;
  // per entry of the box list, two int4: (the entry, grid offset in cp_grid,
  // |A1|, |A2|) and (f offset of A1, f offset of A2, A1 == A2, 0)
  const int4 *box_info;
};
constexpr uint32_t CC_ZERO = 0x3fffffffu;

}  // namespace qd
