// Host channel builder: the (M_L, M_S) two-body blocks, their flat offsets,
// the pair -> (type, block, index) map and the cross-coupled blocks.
//
// Semantics follow src/ME_general.f90: arrange_states (:688-744), reindex
// (:956-1116), get_num_elements (:1755-1778), vectorize (:1871-1989) and
// build_cross_coupled (:363-448).  Built with one pass over all pairs into
// (M_S, M_L) buckets instead of the reference's scan per block.
#include <algorithm>
#include <map>

#include "qd_internal.h"

namespace qd {

std::string build_channels(int nbody, int nshell, Channels &c) {
  if (nshell < 1 || nshell > 40) return "nshell out of range";
  c.n = nbody;
  c.R = nshell;
  // states order: shell, then m_l ascending, then m_s = -1, +1
  c.ml.clear();
  c.ms.clear();
  std::vector<int> shell;
  for (int R = 0; R < nshell; R++)
    for (int ML = -R; ML <= R; ML += 2)
      for (int MS = -1; MS <= 1; MS += 2) {
        c.ml.push_back(ML);
        c.ms.push_back(MS);
        shell.push_back(R);
      }
  const int M = (int)c.ml.size();
  c.M = M;
  c.mp = M - nbody;
  if (nbody < 0 || nbody > M) return "nbody out of range";
  // holes = first N states: requires N to fill whole shells (closed shell),
  // which is what makes the reference's Fermi-level split equal to index < N
  if (nbody > 0 && nbody < M && shell[nbody - 1] == shell[nbody])
    return "nbody must fill complete oscillator shells (2, 6, 12, 20, 30, 42, ...)";
  const int n = nbody;

  // ---- two-body blocks
  std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> bucket;
  for (int i = 0; i < M - 1; i++)
    for (int j = i + 1; j < M; j++) bucket[{c.ms[i] + c.ms[j], c.ml[i] + c.ml[j]}].push_back({i, j});
  c.iarr.assign((size_t)M * (M - 1) / 2, 0u);
  c.nh.clear(); c.np.clear(); c.nb.clear();
  c.pairs.clear(); c.hh_start.clear(); c.ph_start.clear(); c.pp_start.clear();
  int q = 0;
  for (int MS = -2; MS <= 2; MS += 2)
    for (int ML = 2 - 2 * nshell; ML <= 2 * nshell - 2; ML++, q++) {
      std::vector<uint32_t> hh, ph, pp;
      auto it = bucket.find({MS, ML});
      if (it != bucket.end())
        for (auto &pr : it->second) {
          int i = pr.first, j = pr.second;
          size_t II = (size_t)i * M - (size_t)i * (i + 1) / 2 + (j - i - 1);
          if (j < n) { c.iarr[II] = pack_ia(T_HH, q, (int)hh.size()); hh.push_back(pack2(i, j)); }
          else if (i < n) { c.iarr[II] = pack_ia(T_PH, q, (int)ph.size()); ph.push_back(pack2(i, j)); }
          else { c.iarr[II] = pack_ia(T_PP, q, (int)pp.size()); pp.push_back(pack2(i, j)); }
        }
      c.nh.push_back((int)hh.size());
      c.nb.push_back((int)ph.size());
      c.np.push_back((int)pp.size());
      c.hh_start.push_back((int)c.pairs.size());
      c.pairs.insert(c.pairs.end(), hh.begin(), hh.end());
      c.ph_start.push_back((int)c.pairs.size());
      c.pairs.insert(c.pairs.end(), ph.begin(), ph.end());
      c.pp_start.push_back((int)c.pairs.size());
      c.pairs.insert(c.pairs.end(), pp.begin(), pp.end());
    }
  c.nblock = q;
  if (c.nblock >= 1024) return "too many blocks for the packed pair map";

  // ---- flat offsets in vectorize order
  c.off_fhh = 1;
  c.off_fpp = 1 + (int64_t)n * n;
  c.off_fph = c.off_fpp + (int64_t)c.mp * c.mp;
  c.off_2b = c.off_fph + (int64_t)c.mp * n;
  c.off.assign((size_t)c.nblock * 6, 0);
  int64_t k = c.off_2b;
  for (int b = 0; b < c.nblock; b++) {
    int64_t h = c.nh[b], p = c.np[b], x = c.nb[b];
    int64_t sz[6] = {h * h, p * p, p * h, x * x, p * x, x * h};
    for (int s = 0; s < 6; s++) { c.off[(size_t)b * 6 + s] = k; k += sz[s]; }
  }
  c.neq = k;

  // ---- single-particle (m_l, m_s) channels
  int mlmax = *std::max_element(c.ml.begin(), c.ml.end());
  c.ml0 = mlmax;
  c.mlspan = 2 * mlmax + 1;
  c.ch_lookup.assign((size_t)c.mlspan * 2, -1);
  c.spch.assign(M, -1);
  c.nch = 0;
  for (int i = 0; i < M; i++) {
    int key = (c.ml[i] + c.ml0) * 2 + (c.ms[i] > 0);
    if (c.ch_lookup[key] < 0) c.ch_lookup[key] = c.nch++;
    c.spch[i] = c.ch_lookup[key];
  }
  c.ch_start.assign(c.nch + 1, 0);
  for (int i = 0; i < M; i++) c.ch_start[c.spch[i] + 1]++;
  for (int h = 0; h < c.nch; h++) c.ch_start[h + 1] += c.ch_start[h];
  c.ch_states.assign(M, 0);
  {
    std::vector<int> fill(c.ch_start.begin(), c.ch_start.end() - 1);
    for (int i = 0; i < M; i++) c.ch_states[fill[c.spch[i]]++] = i;
  }

  // ---- cross-coupled blocks: ordered pairs (i,j), M_L/M_S differences
  int MLmax = 2 * mlmax;
  std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> ccb;
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) ccb[{c.ml[i] - c.ml[j], c.ms[i] - c.ms[j]}].push_back({i, j});
  c.cc_map.assign((size_t)M * M * 2, -1);
  c.cc_r.clear(); c.cc_nb.clear(); c.cc_off.clear(); c.cc_poff.clear();
  c.cc_ab.clear(); c.cc_hp.clear(); c.cc_ab_start.clear(); c.cc_hp_start.clear();
  int64_t coff = 0, poff = 0;
  q = 0;
  for (int ML = -MLmax; ML <= MLmax; ML++)
    for (int MS = -2; MS <= 2; MS++, q++) {
      c.cc_ab_start.push_back((int)c.cc_ab.size());
      c.cc_hp_start.push_back((int)c.cc_hp.size());
      int r = 0, nbb = 0;
      auto it = ccb.find({ML, MS});
      if (it != ccb.end())
        for (auto &pr : it->second) {
          int i = pr.first, j = pr.second;
          c.cc_map[((size_t)M * i + j) * 2] = q;
          c.cc_map[((size_t)M * i + j) * 2 + 1] = r++;
          c.cc_ab.push_back(pack2(i, j));
          if (i < n && j >= n) { c.cc_hp.push_back(pack2(i, j)); nbb++; }
        }
      c.cc_r.push_back(r);
      c.cc_nb.push_back(nbb);
      c.cc_off.push_back(coff);
      c.cc_poff.push_back(poff);
      coff += (int64_t)r * nbb;
      poff += (nbb > 0) ? (int64_t)r * r : 0;
    }
  c.ncc = q;
  c.cc_size = coff;
  c.p_size = poff;
  // device row order: stable sort of each block's rows by the channel of i
  // (the reference order is i-outer, j-inner, both ascending)
  c.cc_dev_row.assign(c.cc_ab.size(), 0);
  c.cc_ab_dev.assign(c.cc_ab.size(), 0u);
  for (int b = 0; b < c.ncc; b++) {
    const int s0 = c.cc_ab_start[b], r = c.cc_r[b];
    std::vector<int> ord(r);
    for (int a = 0; a < r; a++) ord[a] = a;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int u, int v) { return c.spch[lo16(c.cc_ab[s0 + u])] < c.spch[lo16(c.cc_ab[s0 + v])]; });
    for (int d = 0; d < r; d++) {
      c.cc_dev_row[s0 + ord[d]] = d;
      c.cc_ab_dev[s0 + d] = c.cc_ab[s0 + ord[d]];
    }
  }
  c.sp_pos.assign(M, 0);
  for (int h = 0; h < c.nch; h++)
    for (int s = c.ch_start[h]; s < c.ch_start[h + 1]; s++) c.sp_pos[c.ch_states[s]] = s - c.ch_start[h];
  c.cc_rs.assign((size_t)c.ncc * c.nch, -1);
  for (int b = 0; b < c.ncc; b++)
    for (int d = c.cc_r[b] - 1; d >= 0; d--) c.cc_rs[(size_t)b * c.nch + c.spch[lo16(c.cc_ab_dev[c.cc_ab_start[b] + d])]] = d;
  // runs of the device row order: per row, the first row of its run (the rows
  // whose first states share a channel) and the run's length
  c.cc_run.assign(c.cc_ab_dev.size(), 0u);
  for (int b = 0; b < c.ncc; b++) {
    if (c.cc_r[b] > 65535) return "cross-coupled block too large for the run table";
    for (int d = 0; d < c.cc_r[b]; d++) {
      const uint32_t ab = c.cc_ab_dev[c.cc_ab_start[b] + d];
      const int hi = c.spch[lo16(ab)], hj = c.spch[hi16(ab)];
      const int len = (c.ch_start[hi + 1] - c.ch_start[hi]) * (c.ch_start[hj + 1] - c.ch_start[hj]);
      c.cc_run[c.cc_ab_start[b] + d] = (uint32_t)c.cc_rs[(size_t)b * c.nch + hi] | ((uint32_t)len << 16);
    }
  }
  // the closed form of the device row, relied on by the two-body boxes
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      const int b = c.cc_map[((size_t)M * i + j) * 2], row = c.cc_map[((size_t)M * i + j) * 2 + 1];
      const int hj = c.spch[j];
      const int want = c.cc_rs[(size_t)b * c.nch + c.spch[i]] + c.sp_pos[i] * (c.ch_start[hj + 1] - c.ch_start[hj]) + c.sp_pos[j];
      if (c.cc_dev_row[c.cc_ab_start[b] + row] != want) return "internal error: CC device row order";
    }

  // ---- class pairs and their ordered grids
  c.fcls_off.assign(c.nch, 0);
  c.fcls_size = 0;
  for (int h = 0; h < c.nch; h++) {
    const int sz = c.ch_start[h + 1] - c.ch_start[h];
    c.fcls_off[h] = c.fcls_size;
    c.fcls_size += sz * sz;
  }
  c.cp_start.assign(c.nblock + 1, 0);
  c.cp_a.clear(); c.cp_b.clear(); c.cp_grid_off.clear(); c.cp_tmin.clear(); c.cp_tmax.clear(); c.cp_grid.clear();
  for (int b = 0; b < c.nblock; b++) {
    c.cp_start[b] = (int)c.cp_a.size();
    const int base = c.hh_start[b], ntot = c.nh[b] + c.nb[b] + c.np[b];
    if (ntot > 65535) return "two-body block too large for the class-pair grids";
    std::vector<std::pair<int, int>> keys;
    for (int I = 0; I < ntot; I++) {
      const uint32_t pr = c.pairs[base + I];
      const int u = c.spch[lo16(pr)], v = c.spch[hi16(pr)];
      keys.push_back({std::min(u, v), std::max(u, v)});
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (auto &k : keys) {
      const int A = k.first, B = k.second;
      const int sa = c.ch_start[A + 1] - c.ch_start[A], sb = c.ch_start[B + 1] - c.ch_start[B];
      c.cp_a.push_back(A);
      c.cp_b.push_back(B);
      c.cp_grid_off.push_back((int)c.cp_grid.size());
      int tmin = 3, tmax = -1;
      for (int i1 = 0; i1 < sa; i1++)
        for (int i2 = 0; i2 < sb; i2++) {
          const int x1 = c.ch_states[c.ch_start[A] + i1], x2 = c.ch_states[c.ch_start[B] + i2];
          if (x1 == x2) { c.cp_grid.push_back(0u); continue; }
          const int lo = std::min(x1, x2), hi = std::max(x1, x2);
          const uint32_t ia = c.iarr[(size_t)lo * M - (size_t)lo * (lo + 1) / 2 + (hi - lo - 1)];
          if (ia_blk(ia) != b) return "internal error: class pair left its block";
          const int ty = ia_type(ia) - 1;  // 0 hh, 1 ph, 2 pp
          const int comb = ty == 0 ? ia_idx(ia) : ty == 1 ? c.nh[b] + ia_idx(ia) : c.nh[b] + c.nb[b] + ia_idx(ia);
          tmin = std::min(tmin, ty);
          tmax = std::max(tmax, ty);
          c.cp_grid.push_back((uint32_t)comb | ((uint32_t)(x1 > x2) << 16) | (1u << 17) | ((uint32_t)ty << 18));
        }
      c.cp_tmin.push_back(tmin);
      c.cp_tmax.push_back(tmax);
    }
  }
  c.cp_start[c.nblock] = (int)c.cp_a.size();

  return "";
}

}  // namespace qd
