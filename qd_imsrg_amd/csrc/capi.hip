// C ABI of the engine: contexts, device-resident operators, the commutator
// pipeline, BCH_EXPAND and the Magnus flow (s_evolve).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "ode.h"
#include "qd_kernels.h"

using namespace qd;

namespace {
thread_local std::string g_err;
}  // namespace

int qd::set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {
template <class... A>
int fail(int code, const char *fmt, A... a) {
  return qd::set_error(code, fmt, a...);
}

#define HIPC(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return fail(QD_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// BCH_EXPAND coefficients exactly as written at src/magnus_IMSRG.f90:265-267
const double BCH_COF[11] = {1.0, 0.5, 0.166666666666666666, 0.04166666666666666, 0.0083333333333333333,
                            .001388888888888, 1.984126984e-4, 2.48015873e-5, 2.755731922e-6,
                            2.755731922e-7, 2.505210839e-8};
// REAL(4) literals of the reference widened to REAL(8)
const double R4_1EM5 = 9.999999747378752e-06;   // conv, src/magnus_IMSRG.f90:259
const double R4_1EM9 = 9.999999717180685e-10;   // conv_criteria, src/magnus_IMSRG.f90:10
const double R4_0P7 = 0.699999988079071;        // stp = .7, src/magnus_IMSRG.f90:124
const double R4_1EM3 = 0.0010000000474974513;   // + 1e-3, src/magnus_IMSRG.f90:226
const double R4_1EM8 = 9.99999993922529e-09;    // rel = 1e-8, src/gs_IMSRG.f90:92
const double R4_1EM6 = 9.999999974752427e-07;   // abse = 1e-6, src/gs_IMSRG.f90:93
const double R4_0P1 = 0.10000000149011612;      // stp = 0.1, src/gs_IMSRG.f90:97
const double R4_1EM7 = 1.0000000116860974e-07;  // conv_criteria, src/gs_IMSRG.f90:13

constexpr int NCLASS = 7;
enum { TC_PRODUCTS = 0, TC_ACCUM = 1, TC_TWOBODY = 2, TC_ONEBODY = 3, TC_CALCCC = 4, TC_OTHER = 5 };  // 6 reserved

// One grouped-GEMM work list.  Problems [0, npre) are the ladder products;
// [npre, nprob) the Pandya products (empty for the accumulating list).
struct GemmList {
  std::vector<GemmProb> hprobs;
  std::vector<GemmSeg> hsegs;
  GemmProb *probs = nullptr;
  GemmSeg *segs = nullptr;
  int npre = 0;
  GemmTile *tiles_pre = nullptr, *tiles_all = nullptr;  // by shape, then XCD-interleaved (prob -1: empty)
  int ntiles_pre = 0, ntiles_all = 0;
  // per-launch ranges of the two lists: key sh < NSHAPE = that shape alone,
  // key SH_MIX = the mixable shapes together
  int shape_pre[NSHAPE + 2] = {}, shape_all[NSHAPE + 2] = {};
  GemmTile *ptiles = nullptr;                            // grouped per problem
  std::vector<int> ptile_off;
  size_t cap = 0;                                        // capacity of tiles_pre / tiles_all
  double flops = 0, flops_pre = 0;                       // algorithmic 2*M*N*K
  // per segment, for an operand that is a sub-block of r1 or r2: slot << 8 |
  // its type (ZB bit), else -1
  std::vector<int2> hzb;
  // the list without the segments that read sub-block types of r1 absent from
  // a mask (those products are exact zeros), per mask: device problems (same
  // tiles and shapes; a problem left without segments has nseg = -1 and its
  // tiles are skipped, its output region being zeroed once per BCH expansion
  // by the chunks zc) and segments
  struct Variant {
    GemmProb *probs = nullptr;
    GemmSeg *segs = nullptr;
    ZChunk *zc = nullptr;
    int nzc = 0;
  };
  std::map<uint32_t, Variant> var;
};
}  // namespace

struct qd_context {
  Channels ch;
  DevTab dt{};
  hipStream_t stream = nullptr;
  // side stream (default priority) for the one-body and zero-body pieces of a
  // commutator, forked from and joined back to `stream` with these events
  // layout of the Pandya product in P: 0 = whole column-major D (combine
  // pass), else D in run panels with the upper sub-blocks of runs only, a
  // lower element being this sign times the transposed one (fused kernel:
  // -h_H h_eta)
  double d_sign = 0.0;
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_cc = nullptr, ev_acc = nullptr;
  std::vector<void *> allocs;
  double *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *w4 = nullptr, *P = nullptr;
  double *u1 = nullptr, *u2 = nullptr;  // cross-coupled mat-vecs of commutator_121 (M*M each)
  double *partials = nullptr, *scal = nullptr;  // reduction partials, 8 device scalars
  double *host_scal = nullptr;                  // pinned mirror of scal
  BTask *bt1 = nullptr, *bt2 = nullptr, *bt1p = nullptr, *bt2p = nullptr;  // class-pair box passes
  int nbt1 = 0, nbt2 = 0, nbt1p = 0, nbt2p = 0;
  int *blist = nullptr;                                              // class pairs of the box tasks
  int box_cap = BX_VCAP;                                             // grid elements per box task
  int box_rcap = BX_RCAP;                                            // row grid entries per box task
  bool box_wide = false;                                             // tasks for the wide box form
  bool box_pairs = true;                                             // paired box tasks where valid
  // while nz_on (the BCH terms), nz_mask holds the sub-block types of the
  // first operand (Omega, which holds only fph and Vpphh) with a nonzero
  // element (ZB bits, from the scan into nz); the GEMM lists then run without
  // the products of its other sub-block types
  uint32_t *nz = nullptr;
  uint32_t nz_mask = ZB_ALL;
  ZTile *ztiles = nullptr;
  int nztiles = 0;
  bool nz_on = false, zskip = true;  // QD_ZSKIP=0: never skip
  // whole boxes stage their Pandya sub-blocks of D (contiguous in run panels)
  // in LDS instead of reading each term from HBM: chosen with the wide box
  // form (N=42 R=20 two-body class 2.39 -> 2.28 ms); in the narrow form the
  // extra staging rounds cost more than the reads (N=20 R=14 0.202 -> 0.220)
  bool d_stage = false;
  // the box pass forms the square sub-blocks' ladder terms from the products
  // (no k_ladder_sq pass and no round trip through wl); QD_LDIRECT=0: off
  bool ldirect = true;
  double *fcls = nullptr;                                            // channel blocks of f (2 operands)
  TilePair *ptp = nullptr; int nptp = 0;  // 32x32 tile pairs of the Pandya combine
  TilePair *ftp = nullptr; int nftp = 0;  // 64x64 tile pairs of the fused Pandya product + combine
  // BF16 slice planes of the two CC sets for the fused split-precision Pandya
  // pass: per CC block, units of 8 consecutive k of one row (16 bytes per
  // slice), unit (k / 8, x) at cs_off[q] / 8 + (k / 8) r + x
  int64_t *cs_off = nullptr; int64_t cs_len = 0;  // per-block offsets (bf16 elements), plane length
  int2 *cs_units = nullptr; int ncs_units = 0;    // (block, first unit) of each slicing work-group
  unsigned short *cs = nullptr;                   // E planes then H planes, 4 each (on first BF16 use)
  TilePair *lsq = nullptr; int nlsq = 0;  // 32x32 tile pairs of the square sub-blocks' ladder terms
  double *wl = nullptr;                   // their ladder terms (square sub-blocks of the layout)
  Seg2 *pphh_segs = nullptr; Tile2 *pphh_tiles = nullptr; int npphh_segs = 0, npphh_tiles = 0;
  Seg2 *half_segs = nullptr; Tile2 *half_tiles = nullptr; int nhalf_tiles = 0;  // Vpphh, Vppph sub-blocks
  Seg2 *cc_segs = nullptr; Tile2 *cc_tiles = nullptr; int ncc_tiles = 0;
  GemmList g_all, g_acc;  // products (ladder + Pandya) and accumulating products
  GemmList g_1b;          // one-body products of commutator_111 (always FP64)
  GemmTile *ob_tiles = nullptr; int nob_tiles = 0;  // their 16 x 16 tiles
  int *ob221 = nullptr; int nob221 = 0;  // one-body outputs whose states share a channel
  uint32_t *g221 = nullptr;               // their trace elements as codes (velem_code), or null
  bool gather_tables = true;              // build the calc_cc and 221 gather tables
  int launch_mode = 0;    // 0 grouped persistent launch, 1 one launch per channel problem
  int gemm_prec = GEMM_FP64;  // arithmetic of the 2b-2b GEMMs (GemmPrec)
  bool use_graphs = false;  // capture each BCH term (commutator + norms) in a hipGraph
  struct GraphKey {
    const void *g, *prev, *next, *hs, *cca, *ccg;
    int hg, pd;
    double hh, eh;
    int term;       // BCH term k (its coefficient is a kernel argument); 0 otherwise
    uint32_t zm;    // the GEMM lists' mask of r1 (nz_mask)
    bool operator<(const GraphKey &o) const {
      return std::tie(g, prev, next, hs, cca, ccg, hg, pd, hh, eh, term, zm) <
             std::tie(o.g, o.prev, o.next, o.hs, o.cca, o.ccg, o.hg, o.pd, o.hh, o.eh, o.term, o.zm);
    }
  };
  std::map<GraphKey, hipGraphExec_t> graphs;
  int gemm_grid = 1024;   // persistent grid of the grouped launch
  // allocation-free MBPT2: this device's address of the file-scope scratch
  // array and the partial count it may use (NOALLOC_PARTS, lowered by
  // QD_NOALLOC_CAP at creation)
  double *noalloc_scratch = nullptr;
  int noalloc_cap = 0;
  // scratch operators and CC sets of the flow drivers (created on first use)
  qd_op *bch_a = nullptr, *bch_b = nullptr;
  qd_cc *cc_t1 = nullptr, *cc_t2 = nullptr;
  // timing
  bool timing = false;
  struct Ev { hipEvent_t a, b; int cls; };
  std::vector<Ev> events;
  double t_ms[NCLASS] = {};
  int64_t t_n[NCLASS] = {};
};

struct qd_op {
  qd_context *ctx;
  double *d;
  int herm;
};
struct qd_cc {
  qd_context *ctx;
  double *d;
  double herm;
  bool zero;  // known to be all zeros (skips the Pandya work, exactly)
};

// Scratch of qd_mbpt2_noalloc, the entry point that allocates nothing: the
// reduction partials (one per pphh tile) and the result.  One copy per GPU,
// held until the process ends; calls that overlap on one GPU share it and
// race.  Host code never uses the array's name as an address: the device
// address comes from hipGetSymbolAddress once per device (noalloc_setup).
constexpr int NOALLOC_PARTS = 4096;
static __device__ __align__(256) double scratch[NOALLOC_PARTS + 8];

namespace {
std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;  // device -> address of its scratch

// the calling thread's current device is the context's device
int noalloc_setup(double **out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    double *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(QD_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    it = g_scratch_ptr.emplace(dev, ptr).first;
  }
  *out = it->second;
  return QD_OK;
}
}  // namespace

hipStream_t qd::context_stream(qd_context *c) { return c->stream; }
int64_t qd::context_neq(const qd_context *c) { return c->ch.neq; }
qd_context *qd::op_context(qd_op *o) { return o->ctx; }
double *qd::op_data(qd_op *o) { return o->d; }
const Channels &qd::context_channels(const qd_context *c) { return c->ch; }
const DevTab &qd::context_devtab(const qd_context *c) { return c->dt; }
const double *qd::op_cdata(const qd_op *o) { return o->d; }
int qd::op_herm(const qd_op *o) { return o->herm; }

namespace {

template <class T>
int dev_upload(qd_context *c, const std::vector<T> &h, T **out) {
  size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
  void *p = nullptr;
  HIPC(hipMalloc(&p, bytes));
  c->allocs.push_back(p);
  if (!h.empty()) HIPC(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = (T *)p;
  return QD_OK;
}

int dev_alloc(qd_context *c, size_t bytes, void **out) {
  void *p = nullptr;
  HIPC(hipMalloc(&p, std::max<size_t>(bytes, 8)));
  c->allocs.push_back(p);
  HIPC(hipMemset(p, 0, std::max<size_t>(bytes, 8)));
  *out = p;
  return QD_OK;
}

struct Timer {  // brackets one launch with events when timing is on
  qd_context *c;
  int cls;
  hipEvent_t a{}, b{};
  Timer(qd_context *c_, int cls_) : c(c_), cls(cls_) {
    if (c->timing) {
      hipEventCreate(&a);
      hipEventCreate(&b);
      hipEventRecord(a, c->stream);
    }
  }
  ~Timer() {
    if (c->timing) {
      hipEventRecord(b, c->stream);
      c->events.push_back({a, b, cls});
    }
  }
};

uint32_t velem_code(const Channels &ch, int a1, int a2, int a3, int a4);

void add_tiles(std::vector<Tile2> &tiles, int seg, int size) {
  for (int s = 0; s < size; s += 256) tiles.push_back({seg, s});
}

struct GemmBuilder {
  std::vector<GemmProb> probs;
  std::vector<GemmSeg> segs;
  double flops = 0;
  const std::map<int64_t, int> *zb = nullptr;  // operator offset -> sub-block type (ZB bit)
  std::vector<int2> zinfo;  // per segment: the operands' slot << 8 | ZB bit, or -1
  int zbit(int slot, int64_t off) const {
    if (!zb || (slot != S_R1 && slot != S_R2)) return -1;
    auto it = zb->find(off);
    return it == zb->end() ? -1 : (slot << 8 | it->second);
  }
  void begin(int c_slot, int64_t c_off, int ldc, int M, int N, int beta1) {
    GemmProb p{};
    p.c_off = c_off; p.c_slot = c_slot; p.ldc = std::max(ldc, 1); p.M = M; p.N = N; p.beta1 = beta1;
    p.seg0 = (int)segs.size(); p.nseg = 0; p.shape = SH_64x64;
    probs.push_back(p);
  }
  void seg(int a_slot, int64_t a_off, int lda, int ta, int b_slot, int64_t b_off, int ldb, int tb, int K,
           double alpha) {
    GemmSeg s{};
    s.a_off = a_off; s.b_off = b_off; s.a_slot = a_slot; s.b_slot = b_slot;
    s.lda = std::max(lda, 1); s.ldb = std::max(ldb, 1); s.ta = ta; s.tb = tb; s.K = K; s.alpha = alpha;
    zinfo.push_back(make_int2(zbit(a_slot, a_off), zbit(b_slot, b_off)));
    segs.push_back(s);
    probs.back().nseg++;
    flops += 2.0 * probs.back().M * probs.back().N * K;
  }
  void end() {
    const GemmProb &p = probs.back();
    if (p.M <= 0 || p.N <= 0) {
      segs.resize(p.seg0);
      zinfo.resize(p.seg0);
      probs.pop_back();
    }
  }
};

int ktot(const GemmList &g, const GemmProb &p) {
  int k = 0;
  for (int i = 0; i < p.nseg; i++) k += g.hsegs[p.seg0 + i].K;
  return k;
}

// size-driven default before autotuning: skinny shapes for hole-pair sides
// 64x64 everywhere until the autotuner has measured something better: the
// larger and the skinny shapes only pay off for some size buckets
int default_shape(int, int) { return SH_64x64; }

void tiles_of(const GemmList &g, int pi, std::vector<GemmTile> &out) {
  const GemmProb &p = g.hprobs[pi];
  const int TM = SHAPE_TM[p.shape], TN = SHAPE_TN[p.shape];
  for (int tm = 0; tm * TM < p.M; tm++)
    for (int tn = 0; tn * TN < p.N; tn++) out.push_back({pi, tm, tn});
}

// The problems and segments of g without the segments that read a sub-block
// type of r1 that is not in nz1 (zero in every block): their products are
// exact zeros.  Shapes and tile lists are g's.
void variant_of(const GemmList &g, uint32_t nz1, std::vector<GemmProb> &vp, std::vector<GemmSeg> &vs,
                std::vector<ZChunk> *zc = nullptr) {
  auto dead = [&](int info) { return info >= 0 && (info >> 8) == S_R1 && !((nz1 >> (info & 0xff)) & 1u); };
  vp = g.hprobs;
  vs.clear();
  if (zc) zc->clear();
  for (auto &p : vp) {
    const int s0 = p.seg0, n = p.nseg;
    p.seg0 = (int)vs.size();
    p.nseg = 0;
    for (int i = s0; i < s0 + n; i++)
      if (!dead(g.hzb[i].x) && !dead(g.hzb[i].y)) {
        vs.push_back(g.hsegs[i]);
        p.nseg++;
      }
    // no product left: the output is zero; a contiguous output is zeroed
    // once (chunks) and the problem's tiles are skipped, any other keeps its
    // tiles, which write zeros
    if (p.nseg == 0 && (p.beta1 || p.ldc == p.M)) {
      if (!p.beta1 && zc)
        for (int64_t k = 0; k < (int64_t)p.M * p.N; k += 65536)
          zc->push_back({p.c_slot, (int)std::min<int64_t>(65536, (int64_t)p.M * p.N - k), p.c_off + k});
      p.nseg = -1;
    }
  }
  if (vs.empty()) vs.push_back(GemmSeg{});
}

// (re)builds the three tile lists from the problems' current shapes
int upload_tiles(qd_context *c, GemmList &g) {
  const int np = (int)g.hprobs.size();
  auto work = [&](const GemmTile &t) {
    const GemmProb &p = g.hprobs[t.prob];
    return (double)SHAPE_TM[p.shape] * SHAPE_TN[p.shape] * (ktot(g, p) + GEMM_TK);
  };
  std::vector<GemmTile> pre, all, grouped;
  g.ptile_off.assign(np + 1, 0);
  for (int pi = 0; pi < np; pi++) {
    g.ptile_off[pi] = (int)grouped.size();
    tiles_of(g, pi, grouped);
  }
  g.ptile_off[np] = (int)grouped.size();
  all = grouped;
  for (int pi = 0; pi < g.npre; pi++) tiles_of(g, pi, pre);
  auto key = [&](const GemmTile &t) {
    const int sh = g.hprobs[t.prob].shape;
    return gemm_mixable(sh) ? SH_MIX : sh;
  };
  auto bywork = [&](const GemmTile &a, const GemmTile &b) {
    const int sa = key(a), sb = key(b);
    if (sa != sb) return sa < sb;
    return work(a) > work(b);
  };
  std::stable_sort(pre.begin(), pre.end(), bywork);
  std::stable_sort(all.begin(), all.end(), bywork);
  // XCD-aware order within each launch's range: a problem's operands (< 4 MB
  // here) are read by all its tiles, so its tiles should run on one XCD and
  // share that XCD's L2.  The persistent grid is a multiple of 8 work-groups
  // striding the list and work-group i runs on XCD i % 8, so slot t of the
  // list holds a tile of XCD t % 8.  Problems, cut into pieces of whole tile
  // rows of at most 1/16 of the range's work, go to XCDs largest-first
  // (LPT); each XCD keeps its tiles in decreasing work; short XCD queues are
  // padded with empty tiles (prob -1).
  auto xcd_order = [&](std::vector<GemmTile> &t, int *r) {
    std::vector<GemmTile> out;
    for (int sh = 0; sh <= SH_MIX; sh++) {
      r[sh] = (int)out.size();
      size_t b = 0;
      while (b < t.size() && key(t[b]) < sh) b++;
      size_t e = b;
      while (e < t.size() && key(t[e]) == sh) e++;
      if (e == b) continue;
      double total = 0;
      for (size_t k = b; k < e; k++) total += work(t[k]);
      const double cut = total / 16;
      // pieces: runs of one problem's tiles (same tm rows kept together)
      std::vector<std::pair<size_t, size_t>> pieces;
      std::vector<double> pw;
      for (size_t k = b; k < e;) {
        size_t m = k;
        double w = 0;
        while (m < e && t[m].prob == t[k].prob) {
          size_t n = m;
          double rw = 0;
          while (n < e && t[n].prob == t[k].prob && t[n].tm == t[m].tm) rw += work(t[n++]);
          if (m > k && w + rw > cut) break;
          w += rw;
          m = n;
        }
        pieces.push_back({k, m});
        pw.push_back(w);
        k = m;
      }
      std::vector<int> order(pieces.size());
      for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
      std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pw[x] > pw[y]; });
      std::vector<double> load(NXCD, 0.0);
      std::vector<std::vector<GemmTile>> q(NXCD);
      for (int k : order) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += pw[k];
        for (size_t m = pieces[k].first; m < pieces[k].second; m++) q[x].push_back(t[m]);
      }
      size_t len = 0;
      for (auto &v : q) {
        std::stable_sort(v.begin(), v.end(), [&](const GemmTile &x, const GemmTile &y) { return work(x) > work(y); });
        len = std::max(len, v.size());
      }
      for (size_t j = 0; j < len; j++)
        for (int x = 0; x < NXCD; x++) out.push_back(j < q[x].size() ? q[x][j] : GemmTile{-1, 0, 0});
    }
    r[SH_MIX + 1] = (int)out.size();
    t.swap(out);
  };
  xcd_order(pre, g.shape_pre);
  xcd_order(all, g.shape_all);
  // capacity: grown here only (autotuning drops captured graphs first)
  const size_t need = std::max(all.size(), pre.size());
  if (need > g.cap) {
    size_t cap = 0;
    for (auto &p : g.hprobs) cap += (size_t)((p.M + 15) / 16) * ((p.N + 15) / 16);
    cap = std::max(cap, need + need / 4);
    for (GemmTile *old : {g.tiles_all, g.tiles_pre}) {
      if (!old) continue;
      c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void *)old));
      HIPC(hipFree(old));
    }
    g.cap = cap;
    void *q;
    int rc;
    if ((rc = dev_alloc(c, sizeof(GemmTile) * g.cap, &q))) return rc; g.tiles_all = (GemmTile *)q;
    if ((rc = dev_alloc(c, sizeof(GemmTile) * g.cap, &q))) return rc; g.tiles_pre = (GemmTile *)q;
  }
  if (!g.ptiles) {  // every shape covers at least 16 x 64 outputs
    size_t units = 0;
    for (auto &p : g.hprobs) units += (size_t)((p.M + 15) / 16) * ((p.N + 15) / 16);
    void *q;
    int rc;
    if ((rc = dev_alloc(c, sizeof(GemmTile) * std::max<size_t>(units, 1), &q))) return rc;
    g.ptiles = (GemmTile *)q;
  }
  HIPC(hipMemcpy(g.tiles_all, all.data(), sizeof(GemmTile) * all.size(), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(g.tiles_pre, pre.data(), sizeof(GemmTile) * pre.size(), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(g.ptiles, grouped.data(), sizeof(GemmTile) * grouped.size(), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(g.probs, g.hprobs.data(), sizeof(GemmProb) * np, hipMemcpyHostToDevice));
  for (auto &v : g.var) {  // the variants take the new shapes
    std::vector<GemmProb> vp;
    std::vector<GemmSeg> vs;
    variant_of(g, v.first, vp, vs);
    HIPC(hipMemcpy(v.second.probs, vp.data(), sizeof(GemmProb) * np, hipMemcpyHostToDevice));
  }
  g.ntiles_all = (int)all.size();
  g.ntiles_pre = (int)pre.size();
  return QD_OK;
}

int upload_gemm(qd_context *c, GemmBuilder &b, int npre, double flops_pre, GemmList &g) {
  int rc;
  g.hprobs = b.probs;
  g.hsegs = b.segs;
  g.hzb = b.zinfo;
  if (g.hzb.size() != g.hsegs.size()) return fail(QD_E_LAYOUT, "GEMM list: segment tables out of step");
  g.npre = npre;
  g.flops = b.flops;
  g.flops_pre = flops_pre;
  for (auto &p : g.hprobs) p.shape = default_shape(p.M, p.N);
  if ((rc = dev_upload(c, g.hprobs, &g.probs))) return rc;
  if ((rc = dev_upload(c, g.hsegs, &g.segs))) return rc;
  return upload_tiles(c, g);
}

// One-body products of commutator_111 into the (otherwise unused) one-body
// region of w1, w2, w3, with A = r1, B = r2: P0 = A_hh B_hh, Q0 = B_ph^T A_ph,
// P1 = B_pp A_pp, Q1 = A_ph B_ph^T, and R2 = A_ph B_hh - B_ph A_hh + A_pp B_ph
// - B_pp A_ph in three parts (w1: the two hole sums, w2: A_pp B_ph, w3: B_pp A_ph)
void onebody_products(const Channels &ch, GemmBuilder &g) {
  const int n = ch.n, mp = ch.mp;
  const int64_t ohh = ch.off_fhh, opp = ch.off_fpp, oph = ch.off_fph;
  if (n > 0) {
    g.begin(S_W1, ohh, n, n, n, 0); g.seg(S_R1, ohh, n, 0, S_R2, ohh, n, 0, n, 1); g.end();
    if (mp > 0) { g.begin(S_W2, ohh, n, n, n, 0); g.seg(S_R2, oph, mp, 1, S_R1, oph, mp, 0, mp, 1); g.end(); }
  }
  if (mp > 0) {
    g.begin(S_W1, opp, mp, mp, mp, 0); g.seg(S_R2, opp, mp, 0, S_R1, opp, mp, 0, mp, 1); g.end();
    if (n > 0) {
      g.begin(S_W2, opp, mp, mp, mp, 0); g.seg(S_R1, oph, mp, 0, S_R2, oph, mp, 1, n, 1); g.end();
      g.begin(S_W1, oph, mp, mp, n, 0);
      g.seg(S_R1, oph, mp, 0, S_R2, ohh, n, 0, n, 1);
      g.seg(S_R2, oph, mp, 0, S_R1, ohh, n, 0, n, -1);
      g.end();
      g.begin(S_W2, oph, mp, mp, n, 0); g.seg(S_R1, opp, mp, 0, S_R2, oph, mp, 0, mp, 1); g.end();
      g.begin(S_W3, oph, mp, mp, n, 0); g.seg(S_R2, opp, mp, 0, S_R1, oph, mp, 0, mp, 1); g.end();
    }
  }
}

// The Pandya terms of a box (rows: class pair A1 A2, columns: A3 A4), k =
// D(x3,x1|x2,x4), D(x3,x2|x1,x4), D(x4,x2|x1,x3), D(x4,x1|x2,x3): device rows
// of a CC block come in runs by the channel of the first state, so each
// term's box is the dense sub-block of two runs (rows: channel rc0 with
// partner rc1, columns: cc0 with partner cc1).  In run panels (layout 1) the
// sub-block is contiguous, the stored one being the transposed sub-block when
// the row run starts after the column run (flagged 1 << 30 in the leading
// dimension); whole column-major D (layout 0) has leading dimension r.  pb =
// -1: the term has no CC block.
void box_pand_terms(const Channels &ch, int A1, int A2, int A3, int A4, int layout, int64_t *pb, int *pl) {
  const int rc0[4] = {A3, A3, A4, A4}, rc1[4] = {A1, A2, A2, A1}, cc0[4] = {A2, A1, A1, A2},
            cc1[4] = {A4, A4, A3, A3};
  auto csz = [&](int h) { return ch.ch_start[h + 1] - ch.ch_start[h]; };
  for (int k = 0; k < 4; k++) {
    const int x = ch.ch_states[ch.ch_start[rc0[k]]], y = ch.ch_states[ch.ch_start[rc1[k]]];
    const int K = ch.cc_map[2 * ((size_t)x * ch.M + y)];
    pb[k] = -1;
    pl[k] = 0;
    if (K < 0 || ch.cc_nb[K] == 0) continue;
    const int r = ch.cc_r[K];
    const int Lr = ch.cc_rs[(size_t)K * ch.nch + rc0[k]], Lc = ch.cc_rs[(size_t)K * ch.nch + cc0[k]];
    const int lr = csz(rc0[k]) * csz(rc1[k]), lc = csz(cc0[k]) * csz(cc1[k]);
    const int64_t po = ch.cc_poff[K];
    if (layout == 0) {
      pb[k] = po + Lr + (int64_t)Lc * r;
      pl[k] = r;
    } else if (Lr <= Lc) {
      pb[k] = po + (int64_t)Lr * r + (int64_t)Lc * lr;
      pl[k] = lr;
    } else {
      pb[k] = po + (int64_t)Lc * r + (int64_t)Lr * lc;
      pl[k] = lc | (1 << 30);
    }
  }
}

int build_worklists(qd_context *c) {
  const Channels &ch = c->ch;
  int rc;
  // ---- per-block work (stored two-body elements) for the XCD assignment:
  // all tasks of one block run on one XCD so its operands stay in that XCD's
  // L2; blocks are assigned to XCDs largest-first (LPT on the element count)
  std::vector<Seg2> pseg;
  std::vector<Tile2> ptiles;
  std::vector<double> work(ch.nblock, 0.0);
  for (int q = 0; q < ch.nblock; q++) {
    const double nh = ch.nh[q], np = ch.np[q], nb = ch.nb[q];
    work[q] = nh * nh + np * np + np * nh + nb * nb + np * nb + nb * nh;
    if (np * nh) {
      Seg2 sg{ch.off[(size_t)q * 6 + PPHH], (int)np, (int)(np * nh), q, PPHH, ch.pp_start[q], ch.hh_start[q]};
      add_tiles(ptiles, (int)pseg.size(), sg.size);
      pseg.push_back(sg);
    }
  }
  // ---- class-pair box tasks: per block, every column class pair against
  // bundles of row class pairs (a box only if it holds a stored element:
  // row type >= column type somewhere); a box too large for the LDS grids is
  // split into column slabs (bra terms, first launch) and row slabs (ket
  // terms, second launch).  Blocks go to XCDs as for the two-body tasks.
  //
  // Paired form (bt1p): for class pairs a < b of a block whose boxes fit the
  // LDS grid both ways, the box (rows a, columns b) also writes the outputs of
  // the box (rows b, columns a): <J|Z|I> = h_Z <I|Z|J> for the 1b x 2b and
  // Pandya parts, so one staged grid and one contraction serve both.  Its
  // blist entry carries BOX_PAIRED; the box (b, a) is left out.  Where both
  // boxes are too large for one grid, the slabs of (a, b) write both
  // orientations the same way (column slabs in bt1p, row slabs in bt2p, the
  // row slabs' ket terms being the first writers of either orientation).
  std::vector<std::vector<BTask>> bt1(ch.nblock), bt2(ch.nblock), bt1p(ch.nblock), bt2p(ch.nblock);
  std::vector<int> blist;
  {
    const int cap = c->box_cap;
    auto csize = [&](int h) { return ch.ch_start[h + 1] - ch.ch_start[h]; };
    auto whole_box = [&](int gr, int gc) {  // rows gr x columns gc fits one task's grid
      const int nr = csize(ch.cp_a[gr]) * csize(ch.cp_b[gr]), nc = csize(ch.cp_a[gc]) * csize(ch.cp_b[gc]);
      return nr * (nc | 1) <= cap;
    };
    auto pairable = [&](int x, int y) {  // x < y
      return whole_box(x, y) && whole_box(y, x) &&
             (ch.cp_tmax[x] >= ch.cp_tmin[y] || ch.cp_tmax[y] >= ch.cp_tmin[x]);
    };
    auto slab_pairable = [&](int x, int y) {  // x < y, both boxes split into slabs
      return !whole_box(x, y) && !whole_box(y, x) &&
             (ch.cp_tmax[x] >= ch.cp_tmin[y] || ch.cp_tmax[y] >= ch.cp_tmin[x]);
    };
    // column slabs (mode 1, bra terms) and row slabs (mode 2, ket terms) of
    // the box gr x gc into c1 and c2, with blist entry e
    auto slabs = [&](int q, int gr, int gc, int e, std::vector<BTask> *c1a, std::vector<BTask> *c1b,
                     std::vector<BTask> *c2a, std::vector<BTask> *c2b) {
      const int nr = csize(ch.cp_a[gr]) * csize(ch.cp_b[gr]), nc = csize(ch.cp_a[gc]) * csize(ch.cp_b[gc]);
      const int ldc = nc | 1;
      int w = cap / nr;                   // column slab: nr * (w | 1) <= cap
      if ((w | 1) > cap / nr) w--;
      for (int lo = 0; lo < nc; lo += w) {
        const BTask t{q, gc, (int)blist.size(), 1, 1, lo, std::min(w, nc - lo)};
        if (c1a) c1a->push_back(t);
        if (c1b) c1b->push_back(t);
        blist.push_back(e);
      }
      const int h = cap / ldc;            // row slab: h * ldc <= cap
      for (int lo = 0; lo < nr; lo += h) {
        const BTask t{q, gc, (int)blist.size(), 1, 2, lo, std::min(h, nr - lo)};
        if (c2a) c2a->push_back(t);
        if (c2b) c2b->push_back(t);
        blist.push_back(e);
      }
    };
    for (int q = 0; q < ch.nblock; q++)
      for (int gc = ch.cp_start[q]; gc < ch.cp_start[q + 1]; gc++) {
        const int s3 = csize(ch.cp_a[gc]), s4 = csize(ch.cp_b[gc]);
        const int nc = s3 * s4, ldc = nc | 1, fcol = 2 * (s3 * s3 + s4 * s4);
        // whole boxes go to tasks first-fit decreasing (by grid size) under
        // the LDS limits: fuller tasks than bundling in class-pair order
        struct Bin { std::vector<int> g; int goff, foff, roff; };
        std::vector<std::pair<int, int>> whole, wholep;  // (grid size, row class pair [| BOX_PAIRED])
        for (int gr = ch.cp_start[q]; gr < ch.cp_start[q + 1]; gr++) {
          const int s1 = csize(ch.cp_a[gr]), s2 = csize(ch.cp_b[gr]);
          const int nr = s1 * s2, fneed = 2 * (s1 * s1 + s2 * s2);
          if (fcol + fneed > BX_FCAP || 3 * nr > cap || ldc > cap || nc > BX_CMAX || nr > c->box_rcap)
            return fail(QD_E_LAYOUT, "single-particle channel of %d states too large for the two-body boxes",
                        std::max(std::max(s1, s2), std::max(s3, s4)));
          const bool pair_here = gr < gc && pairable(gr, gc), pair_there = gr > gc && pairable(gc, gr);
          const bool spl_here = gr < gc && slab_pairable(gr, gc), spl_there = gr > gc && slab_pairable(gc, gr);
          if (pair_here) wholep.push_back({nr * ldc, gr | BOX_PAIRED});
          if (spl_here) slabs(q, gr, gc, gr | BOX_PAIRED, nullptr, &bt1p[q], nullptr, &bt2p[q]);
          if (ch.cp_tmax[gr] < ch.cp_tmin[gc]) continue;
          if (nr * ldc > cap) {
            const bool pl = spl_here || spl_there;  // the paired lists have these outputs already
            slabs(q, gr, gc, gr, &bt1[q], pl ? nullptr : &bt1p[q], &bt2[q], pl ? nullptr : &bt2p[q]);
            continue;
          }
          whole.push_back({nr * ldc, gr});
          if (!pair_here && !pair_there) wholep.push_back({nr * ldc, gr});
        }
        for (int form = 0; form < 2; form++) {
          auto &wl = form == 0 ? whole : wholep;
          auto &dst = form == 0 ? bt1[q] : bt1p[q];
          std::vector<Bin> bins;
          std::stable_sort(wl.begin(), wl.end(), [](auto &x, auto &y) { return x.first > y.first; });
          for (auto &w : wl) {
            const int gr = w.second & ~BOX_PAIRED;
            const int s1 = csize(ch.cp_a[gr]), s2 = csize(ch.cp_b[gr]);
            const int nr = s1 * s2, fneed = 2 * (s1 * s1 + s2 * s2);
            Bin *fit = nullptr;
            for (auto &bn : bins)
              if (bn.goff + nr * ldc <= cap && bn.foff + fneed <= BX_FCAP && bn.roff + nr <= c->box_rcap &&
                  (int)bn.g.size() < BX_GMAX) {
                fit = &bn;
                break;
              }
            if (!fit) {
              bins.push_back(Bin{{}, 0, fcol, 0});
              fit = &bins.back();
            }
            fit->g.push_back(w.second);
            fit->goff += nr * ldc;
            fit->foff += fneed;
            fit->roff += nr;
          }
          for (auto &bn : bins) {
            dst.push_back(BTask{q, gc, (int)blist.size(), (int)bn.g.size(), 0, 0, 0});
            blist.insert(blist.end(), bn.g.begin(), bn.g.end());
          }
        }
      }
  }
  {  // the Pandya term bases of every box of the first-pass tasks (modes 0, 1)
    const size_t nbl = blist.size();
    std::vector<int64_t> pbv(2 * nbl * 4, -1);
    std::vector<int> plv(2 * nbl * 4, 0);
    for (auto *lists : {&bt1, &bt1p})
      for (auto &l : *lists)
        for (const BTask &t : l) {
          if (t.mode == 2) continue;
          for (int b = 0; b < t.ng; b++) {
            const int g = blist[t.first + b] & ~BOX_PAIRED;
            for (int layout = 0; layout < 2; layout++) {
              const size_t e = ((size_t)layout * nbl + t.first + b) * 4;
              box_pand_terms(ch, ch.cp_a[g], ch.cp_b[g], ch.cp_a[t.gcol], ch.cp_b[t.gcol], layout, &pbv[e], &plv[e]);
            }
          }
        }
    if ((rc = dev_upload(c, pbv, (int64_t **)&c->dt.box_pb)) || (rc = dev_upload(c, plv, (int **)&c->dt.box_pl)))
      return rc;
    c->dt.nblist = (int)nbl;
  }
  {
    auto csz = [&](int h) { return ch.ch_start[h + 1] - ch.ch_start[h]; };
    std::vector<int> order(ch.nblock);
    for (int q = 0; q < ch.nblock; q++) order[q] = q;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    std::vector<double> load(NXCD, 0.0);
    for (int pass = 0; pass < 4; pass++) {
      auto &src = pass == 0 ? bt1 : pass == 1 ? bt2 : pass == 2 ? bt1p : bt2p;
      std::vector<std::vector<BTask>> bx(NXCD);
      std::fill(load.begin(), load.end(), 0.0);
      // within a block, the largest tasks (grid elements) first; the blocks
      // stay contiguous (their operands shared in the XCD's L2)
      auto grid_of = [&](const BTask &t) {
        const int nc = csz(ch.cp_a[t.gcol]) * csz(ch.cp_b[t.gcol]);
        if (t.mode != 0) {
          const int gr = blist[t.first] & ~BOX_PAIRED;
          return t.mode == 1 ? csz(ch.cp_a[gr]) * csz(ch.cp_b[gr]) * t.cnt : t.cnt * nc;
        }
        int e = 0;
        for (int b = 0; b < t.ng; b++) {
          const int gr = blist[t.first + b] & ~BOX_PAIRED;
          e += csz(ch.cp_a[gr]) * csz(ch.cp_b[gr]) * nc;
        }
        return e;
      };
      for (int q : order) {
        if (work[q] == 0) continue;
        int best = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[best] += work[q];
        const size_t at = bx[best].size();
        bx[best].insert(bx[best].end(), src[q].begin(), src[q].end());
        std::stable_sort(bx[best].begin() + at, bx[best].end(),
                         [&](const BTask &u, const BTask &v) { return grid_of(u) > grid_of(v); });
      }
      size_t lng = 0;
      for (auto &l : bx) lng = std::max(lng, l.size());
      std::vector<BTask> bt(lng * NXCD, BTask{-1, 0, 0, 0, 0, 0, 0});
      for (int x = 0; x < NXCD; x++)
        for (size_t k = 0; k < bx[x].size(); k++) bt[k * NXCD + x] = bx[x][k];
      if (bt.empty()) continue;
      if ((rc = dev_upload(c, bt, pass == 0 ? &c->bt1 : pass == 1 ? &c->bt2 : pass == 2 ? &c->bt1p : &c->bt2p)))
        return rc;
      (pass == 0 ? c->nbt1 : pass == 1 ? c->nbt2 : pass == 2 ? c->nbt1p : c->nbt2p) = (int)bt.size();
    }
    if (blist.empty()) blist.push_back(0);
    if ((rc = dev_upload(c, blist, &c->blist))) return rc;
    std::vector<int4> info(2 * blist.size());
    for (size_t e = 0; e < blist.size(); e++) {
      const int g = blist[e] & ~BOX_PAIRED, A1 = ch.cp_a[g], A2 = ch.cp_b[g];
      info[2 * e] = make_int4(blist[e], ch.cp_grid_off[g], csz(A1), csz(A2));
      info[2 * e + 1] = make_int4(ch.fcls_off[A1], ch.fcls_off[A2], A1 == A2, 0);
    }
    if ((rc = dev_upload(c, info, (int4 **)&c->dt.box_info))) return rc;
  }
  if ((rc = dev_upload(c, pseg, &c->pphh_segs))) return rc;
  if ((rc = dev_upload(c, ptiles, &c->pphh_tiles))) return rc;
  {  // the sub-blocks whose summed products come in two halves (w3 + w4 -> wl)
    std::vector<Seg2> hseg;
    std::vector<Tile2> htiles;
    for (int q = 0; q < ch.nblock; q++)
      for (int sub : {PPHH, PPPH}) {
        const int rows = ch.np[q], cols = sub == PPHH ? ch.nh[q] : ch.nb[q];
        if (rows * cols == 0) continue;
        Seg2 sg{ch.off[(size_t)q * 6 + sub], rows, rows * cols, q, sub, 0, 0};
        add_tiles(htiles, (int)hseg.size(), sg.size);
        hseg.push_back(sg);
      }
    if ((rc = dev_upload(c, hseg, &c->half_segs))) return rc;
    if ((rc = dev_upload(c, htiles, &c->half_tiles))) return rc;
    c->nhalf_tiles = (int)htiles.size();
  }
  c->npphh_segs = (int)pseg.size();
  c->npphh_tiles = (int)ptiles.size();
  // ---- CC segments
  std::vector<Seg2> csegs;
  std::vector<Tile2> ctiles;
  for (int q = 0; q < ch.ncc; q++) {
    const int r = ch.cc_r[q], nb = ch.cc_nb[q];
    if (r * nb == 0) continue;
    Seg2 sg{ch.cc_off[q], r, r * nb, q, 0, ch.cc_ab_start[q], ch.cc_hp_start[q]};
    add_tiles(ctiles, (int)csegs.size(), sg.size);
    csegs.push_back(sg);
  }
  if ((rc = dev_upload(c, csegs, &c->cc_segs))) return rc;
  if ((rc = dev_upload(c, ctiles, &c->cc_tiles))) return rc;
  c->ncc_tiles = (int)ctiles.size();

  // ---- operator sub-blocks: the ZB bit of each non-empty sub-block by its
  // offset (for skipping products of all-zero sub-block types) and the tiles
  // of the scan that finds them
  std::map<int64_t, int> zb;
  std::vector<ZTile> zt;
  auto zregion = [&](int64_t off, int64_t size, int bit) {
    if (size <= 0) return;
    zb.emplace(off, bit);
    for (int64_t k = 0; k < size; k += 16384) zt.push_back({off + k, (int)std::min<int64_t>(16384, size - k), bit});
  };
  zregion(ch.off_fhh, (int64_t)ch.n * ch.n, ZB_FHH);
  zregion(ch.off_fpp, (int64_t)ch.mp * ch.mp, ZB_FPP);
  zregion(ch.off_fph, (int64_t)ch.mp * ch.n, ZB_FPH);
  for (int q = 0; q < ch.nblock; q++) {
    const int64_t nh = ch.nh[q], np = ch.np[q], nb = ch.nb[q];
    const int64_t sz[6] = {nh * nh, np * np, np * nh, nb * nb, np * nb, nb * nh};  // enum Sub order
    for (int u = 0; u < 6; u++) zregion(ch.off[(size_t)q * 6 + u], sz[u], u);
  }
  if ((rc = dev_upload(c, zt, &c->ztiles))) return rc;
  c->nztiles = (int)zt.size();

  // ---- GEMM products (matrix_mult222 / commutator_221 intermediates)
  GemmBuilder g1, g2;
  g1.zb = g2.zb = &zb;
  for (int q = 0; q < ch.nblock; q++) {
    const int nh = ch.nh[q], np = ch.np[q], nb = ch.nb[q];
    const int64_t *o = &ch.off[(size_t)q * 6];
    // w1.hhhh = r2.pphh^T r1.pphh
    g1.begin(S_W1, o[HHHH], nh, nh, nh, 0); g1.seg(S_R2, o[PPHH], np, 1, S_R1, o[PPHH], np, 0, np, 1); g1.end();
    // w1.pppp = r1.pphh r2.pphh^T
    g1.begin(S_W1, o[PPPP], np, np, np, 0); g1.seg(S_R1, o[PPHH], np, 0, S_R2, o[PPHH], np, 1, nh, 1); g1.end();
    // w1.phph = r2.ppph^T r1.ppph
    g1.begin(S_W1, o[PHPH], nb, nb, nb, 0); g1.seg(S_R2, o[PPPH], np, 1, S_R1, o[PPPH], np, 0, np, 1); g1.end();
    // w1.phhh = r2.ppph^T r1.pphh
    g1.begin(S_W1, o[PHHH], nb, nb, nh, 0); g1.seg(S_R2, o[PPPH], np, 1, S_R1, o[PPHH], np, 0, np, 1); g1.end();
    // w1.ppph = r1.pphh r2.phhh^T
    g1.begin(S_W1, o[PPPH], np, np, nb, 0); g1.seg(S_R1, o[PPHH], np, 0, S_R2, o[PHHH], nb, 1, nh, 1); g1.end();
    // w2.phph = r1.phhh r2.phhh^T
    g1.begin(S_W2, o[PHPH], nb, nb, nb, 0); g1.seg(S_R1, o[PHHH], nb, 0, S_R2, o[PHHH], nb, 1, nh, 1); g1.end();
    // w2.phhh = r1.ppph^T r2.pphh
    g1.begin(S_W2, o[PHHH], nb, nb, nh, 0); g1.seg(S_R1, o[PPPH], np, 1, S_R2, o[PPHH], np, 0, np, 1); g1.end();
    // w2.ppph = r2.pphh r1.phhh^T
    g1.begin(S_W2, o[PPPH], np, np, nb, 0); g1.seg(S_R2, o[PPHH], np, 0, S_R1, o[PHHH], nb, 1, nh, 1); g1.end();
    // w3.hhhh = r1.hhhh r2.hhhh ; w3.pppp = r2.pppp r1.pppp
    g1.begin(S_W3, o[HHHH], nh, nh, nh, 0); g1.seg(S_R1, o[HHHH], nh, 0, S_R2, o[HHHH], nh, 0, nh, 1); g1.end();
    g1.begin(S_W3, o[PPPP], np, np, np, 0); g1.seg(S_R2, o[PPPP], np, 0, S_R1, o[PPPP], np, 0, np, 1); g1.end();
    // the products that add to Vpphh, Vphhh, Vppph go to the same sub-blocks of
    // w3 (free there); the two-body pass adds them to r3
    // w3.pphh = r1.pppp r2.pphh - r2.pppp r1.pphh - r1.pphh r2.hhhh + r2.pphh r1.hhhh
    // in two halves (w3, w4; their sum goes to wl, which the two-body pass
    // reads), so that the longest tiles walk half the K
    g2.begin(S_W3, o[PPHH], np, np, nh, 0);
    g2.seg(S_R1, o[PPPP], np, 0, S_R2, o[PPHH], np, 0, np, 1);
    g2.seg(S_R1, o[PPHH], np, 0, S_R2, o[HHHH], nh, 0, nh, -1);
    g2.end();
    g2.begin(S_W4, o[PPHH], np, np, nh, 0);
    g2.seg(S_R2, o[PPPP], np, 0, S_R1, o[PPHH], np, 0, np, -1);
    g2.seg(S_R2, o[PPHH], np, 0, S_R1, o[HHHH], nh, 0, nh, 1);
    g2.end();
    // w3.phhh = -r1.phhh r2.hhhh + r2.phhh r1.hhhh
    g2.begin(S_W3, o[PHHH], nb, nb, nh, 0);
    g2.seg(S_R1, o[PHHH], nb, 0, S_R2, o[HHHH], nh, 0, nh, -1);
    g2.seg(S_R2, o[PHHH], nb, 0, S_R1, o[HHHH], nh, 0, nh, 1);
    g2.end();
    // w3.ppph = r1.pppp r2.ppph - r2.pppp r1.ppph
    g2.begin(S_W3, o[PPPH], np, np, nb, 0);
    g2.seg(S_R1, o[PPPP], np, 0, S_R2, o[PPPH], np, 0, np, 1);
    g2.end();
    g2.begin(S_W4, o[PPPH], np, np, nb, 0);
    g2.seg(S_R2, o[PPPP], np, 0, S_R1, o[PPPH], np, 0, np, -1);
    g2.end();
  }
  // the one-body products ride along in the products launch (a few small
  // tiles among the ladder tiles)
  onebody_products(ch, g1);
  const int npre = (int)g1.probs.size();
  const double f1 = g1.flops;
  // Pandya: P(q) = ETACC(q) HCC(q)^T (r x r, K = nb); appended after G1
  for (int q = 0; q < ch.ncc; q++) {
    const int r = ch.cc_r[q], nb = ch.cc_nb[q];
    if (r * nb == 0) continue;
    g1.begin(S_P, ch.cc_poff[q], r, r, r, 0);
    g1.seg(S_CC1, ch.cc_off[q], r, 0, S_CC2, ch.cc_off[q], r, 1, nb, 1);
    g1.end();
  }

  {
    std::vector<TilePair> tp;
    for (int q = 0; q < ch.ncc; q++) {
      const int r = ch.cc_r[q], nb = ch.cc_nb[q];
      if (r * nb == 0) continue;
      const int nt = (r + 31) / 32;
      for (int ti = 0; ti < nt; ti++)
        for (int tj = ti; tj < nt; tj++) tp.push_back({q, ti, tj});
    }
    if ((rc = dev_upload(c, tp, &c->ptp))) return rc;
    c->nptp = (int)tp.size();
    // the fused pass: largest blocks first, so the persistent tail is short
    std::vector<std::pair<int, int>> order;
    for (int q = 0; q < ch.ncc; q++)
      if (ch.cc_r[q] * ch.cc_nb[q]) order.push_back({ch.cc_r[q] * ch.cc_r[q] * ch.cc_nb[q], q});
    std::stable_sort(order.begin(), order.end(), [](auto &a, auto &b) { return a.first > b.first; });
    std::vector<TilePair> ftp;
    for (auto &oq : order) {
      const int q = oq.second, nt = (ch.cc_r[q] + 63) / 64;
      for (int ti = 0; ti < nt; ti++)
        for (int tj = ti; tj < nt; tj++) ftp.push_back({q, ti, tj});
    }
    if ((rc = dev_upload(c, ftp, &c->ftp))) return rc;
    c->nftp = (int)ftp.size();
    std::vector<int64_t> soff(ch.ncc, 0);
    std::vector<int2> units;
    int64_t len = 0;
    for (int q = 0; q < ch.ncc; q++) {
      soff[q] = len;
      const int r = ch.cc_r[q], nb = ch.cc_nb[q];
      if (r * nb == 0) continue;
      const int64_t nu = (int64_t)((nb + 7) / 8) * r;
      for (int64_t u0 = 0; u0 < nu; u0 += CS_UNITS_WG) units.push_back(make_int2(q, (int)u0));
      len += 8 * nu;
    }
    if ((rc = dev_upload(c, soff, &c->cs_off)) || (rc = dev_upload(c, units, &c->cs_units))) return rc;
    c->cs_len = len;
    c->ncs_units = (int)units.size();
  }
  {  // square sub-blocks (HHHH, PHPH, PPPP) of every block, largest first
    std::vector<std::pair<int64_t, int>> order;
    for (int q = 0; q < ch.nblock; q++) {
      const int n3[3] = {ch.nh[q], ch.nb[q], ch.np[q]};
      for (int k = 0; k < 3; k++)
        if (n3[k] > 0) order.push_back({(int64_t)n3[k] * n3[k], 3 * q + k});
    }
    std::stable_sort(order.begin(), order.end(), [](auto &a, auto &b) { return a.first > b.first; });
    std::vector<TilePair> lsq;
    for (auto &o : order) {
      const int q = o.second / 3, k = o.second % 3;
      const int n = k == 0 ? ch.nh[q] : (k == 1 ? ch.nb[q] : ch.np[q]), nt = (n + 31) / 32;
      for (int ti = 0; ti < nt; ti++)
        for (int tj = ti; tj < nt; tj++) lsq.push_back({o.second, ti, tj});
    }
    if ((rc = dev_upload(c, lsq, &c->lsq))) return rc;
    c->nlsq = (int)lsq.size();
  }
  if ((rc = upload_gemm(c, g1, npre, f1, c->g_all))) return rc;
  if ((rc = upload_gemm(c, g2, (int)g2.probs.size(), g2.flops, c->g_acc))) return rc;
  // ---- one-body products of commutator_111 on their own (16 x 16 FMA tiles,
  // FP64 always), for the calls whose products pass did not compute them
  {
    GemmBuilder g3;
    onebody_products(ch, g3);
    const int n = ch.n, mp = ch.mp;
    if ((rc = upload_gemm(c, g3, (int)g3.probs.size(), g3.flops, c->g_1b))) return rc;
    std::vector<GemmTile> t16;
    for (int pi = 0; pi < (int)g3.probs.size(); pi++)
      for (int tm = 0; tm < (g3.probs[pi].M + 15) / 16; tm++)
        for (int tn = 0; tn < (g3.probs[pi].N + 15) / 16; tn++) t16.push_back(GemmTile{pi, tm, tn});
    c->nob_tiles = (int)t16.size();
    if (t16.empty()) t16.push_back(GemmTile{0, 0, 0});
    if ((rc = dev_upload(c, t16, &c->ob_tiles))) return rc;
    std::vector<int> l221;
    const int64_t tot = (int64_t)n * n + (int64_t)mp * mp + (int64_t)mp * n;
    for (int64_t e = 0; e < tot; e++) {
      int x, y;
      if (e < (int64_t)n * n) { x = (int)(e % n); y = (int)(e / n); }
      else if (e < (int64_t)n * n + (int64_t)mp * mp) { const int64_t f = e - (int64_t)n * n; x = n + (int)(f % mp); y = n + (int)(f / mp); }
      else { const int64_t f = e - (int64_t)n * n - (int64_t)mp * mp; x = n + (int)(f % mp); y = (int)(f / mp); }
      if (ch.spch[x] == ch.spch[y]) l221.push_back((int)e);
    }
    c->nob221 = (int)l221.size();
    if (ch.neq < (int64_t)CC_ZERO && c->gather_tables && !l221.empty()) {
      // per listed output (x, y) and state a: the codes of <a x|w|a y> and
      // <a y|w|a x> (w1, w2 share the layout; herm = 1)
      std::vector<uint32_t> g((size_t)l221.size() * ch.M * 2);
      for (size_t k = 0; k < l221.size(); k++) {
        const int64_t e = l221[k];
        int x, y;
        if (e < (int64_t)n * n) { x = (int)(e % n); y = (int)(e / n); }
        else if (e < (int64_t)n * n + (int64_t)mp * mp) { const int64_t f = e - (int64_t)n * n; x = n + (int)(f % mp); y = n + (int)(f / mp); }
        else { const int64_t f = e - (int64_t)n * n - (int64_t)mp * mp; x = n + (int)(f % mp); y = (int)(f / mp); }
        for (int a = 0; a < ch.M; a++) {
          g[(k * ch.M + a) * 2] = velem_code(ch, a, x, a, y);
          g[(k * ch.M + a) * 2 + 1] = velem_code(ch, a, y, a, x);
        }
      }
      if ((rc = dev_upload(c, g, &c->g221))) return rc;
    }
    if (l221.empty()) l221.push_back(0);
    if ((rc = dev_upload(c, l221, &c->ob221))) return rc;
  }
  return QD_OK;
}

int check_op(const qd_context *c, const qd_op *o, const char *what) {
  if (!o) return fail(QD_E_ARG, "%s is null", what);
  if (o->ctx != c) return fail(QD_E_ARG, "%s belongs to another context", what);
  return QD_OK;
}

Slots make_slots(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *out, const qd_cc *ecc,
                 const qd_cc *hcc) {
  Slots s{};
  s.p[S_R1] = r1 ? r1->d : nullptr;
  s.p[S_R2] = r2 ? r2->d : nullptr;
  s.p[S_OUT] = out ? out->d : nullptr;
  s.p[S_W1] = c->w1; s.p[S_W2] = c->w2; s.p[S_W3] = c->w3; s.p[S_W4] = c->w4;
  s.p[S_CC1] = ecc ? ecc->d : nullptr;
  s.p[S_CC2] = hcc ? hcc->d : nullptr;
  s.p[S_P] = c->P;
  return s;
}

// v_elem (src/ME_general.f90:1557-1680) as a code: the flat index of the
// stored element of <a1 a2|V|a3 a4> after sorting both pairs | stored
// transposed (the herm factor applies) << 30 | the sorting sign << 31, or
// CC_ZERO for a repeated state or pairs in different blocks.
uint32_t velem_code(const Channels &ch, int a1, int a2, int a3, int a4) {
  const int M = ch.M;
  if (a1 == a2 || a3 == a4) return CC_ZERO;
  uint32_t sgn = 0;
  if (a1 > a2) { std::swap(a1, a2); sgn ^= 1u; }
  if (a3 > a4) { std::swap(a3, a4); sgn ^= 1u; }
  const uint32_t I = ch.iarr[(size_t)a1 * M - (size_t)a1 * (a1 + 1) / 2 + (a2 - a1 - 1)];
  const uint32_t J = ch.iarr[(size_t)a3 * M - (size_t)a3 * (a3 + 1) / 2 + (a4 - a3 - 1)];
  const int q = ia_blk(I);
  if (q != ia_blk(J)) return CC_ZERO;
  const int64_t w1 = ia_idx(I), w2 = ia_idx(J), nh = ch.nh[q], np = ch.np[q], nb = ch.nb[q];
  const int64_t *o = &ch.off[6 * (size_t)q];
  int64_t k = 0;
  uint32_t tr = 0;
  switch (ia_type(I) * 4 + ia_type(J)) {
    case T_HH * 4 + T_HH: k = o[HHHH] + w1 + w2 * nh; break;
    case T_HH * 4 + T_PH: k = o[PHHH] + w2 + w1 * nb; tr = 1; break;
    case T_HH * 4 + T_PP: k = o[PPHH] + w2 + w1 * np; tr = 1; break;
    case T_PH * 4 + T_HH: k = o[PHHH] + w1 + w2 * nb; break;
    case T_PH * 4 + T_PH: k = o[PHPH] + w1 + w2 * nb; break;
    case T_PH * 4 + T_PP: k = o[PPPH] + w2 + w1 * np; tr = 1; break;
    case T_PP * 4 + T_HH: k = o[PPHH] + w1 + w2 * np; break;
    case T_PP * 4 + T_PH: k = o[PPPH] + w1 + w2 * np; break;
    default: k = o[PPPP] + w1 + w2 * np; break;
  }
  return (uint32_t)k | (tr << 30) | (sgn << 31);
}

// calc_cc (src/ME_general.f90:923-952) as a gather table: for the CC element
// (row (i, j), column (k, l)) of every block in device order, the code of
// the V element <j k|V|i l>.  Requires neq < CC_ZERO.
std::vector<uint32_t> cc_gather_map(const Channels &ch) {
  auto elem = [&](int a1, int a2, int a3, int a4) { return velem_code(ch, a1, a2, a3, a4); };
  std::vector<uint32_t> m((size_t)ch.cc_size);
  for (int q = 0; q < ch.ncc; q++) {
    const int r = ch.cc_r[q], nb = ch.cc_nb[q];
    for (int c = 0; c < nb; c++) {
      const uint32_t hp = ch.cc_hp[ch.cc_hp_start[q] + c];
      for (int a = 0; a < r; a++) {
        const uint32_t ab = ch.cc_ab_dev[ch.cc_ab_start[q] + a];
        m[(size_t)ch.cc_off[q] + a + (size_t)c * r] = elem(hi16(ab), lo16(hp), lo16(ab), hi16(hp));
      }
    }
  }
  return m;
}

bool pandya_on(const qd_cc *hcc, const qd_cc *ecc) {
  return hcc && ecc && !hcc->zero && !ecc->zero;
}

// the variant of g for the context's current mask of r1 (built on first use),
// or null when nothing is skipped
int gemm_variant(qd_context *c, GemmList &g, GemmList::Variant **out) {
  *out = nullptr;
  if (!c->nz_on || c->nz_mask == (uint32_t)ZB_ALL) return QD_OK;
  auto it = g.var.find(c->nz_mask);
  if (it == g.var.end()) {
    std::vector<GemmProb> vp;
    std::vector<GemmSeg> vs;
    std::vector<ZChunk> zc;
    variant_of(g, c->nz_mask, vp, vs, &zc);
    GemmList::Variant v;
    v.nzc = (int)zc.size();
    if (zc.empty()) zc.push_back(ZChunk{});
    int rc;
    if ((rc = dev_upload(c, vp, &v.probs)) || (rc = dev_upload(c, vs, &v.segs)) || (rc = dev_upload(c, zc, &v.zc)))
      return rc;
    it = g.var.emplace(c->nz_mask, v).first;
  }
  *out = &it->second;
  return QD_OK;
}

int gemm_ptrs(qd_context *c, GemmList &g, GemmProb **pp, GemmSeg **ps) {
  GemmList::Variant *v;
  int rc = gemm_variant(c, g, &v);
  if (rc) return rc;
  *pp = v ? v->probs : g.probs;
  *ps = v ? v->segs : g.segs;
  return QD_OK;
}

// One grouped persistent launch over the sorted tile list, or (per-channel
// mode) one launch per problem, the way the reference calls dgemm per block.
int run_gemm_list(qd_context *c, GemmList &g, bool with_pandya, const Slots &s, int prec = -1) {
  if (prec < 0) prec = c->gemm_prec;
  GemmProb *probs;
  GemmSeg *segs;
  int rc = gemm_ptrs(c, g, &probs, &segs);
  if (rc) return rc;
  if (c->launch_mode == 0) {
    const int *r = with_pandya ? g.shape_all : g.shape_pre;
    GemmTile *t = with_pandya ? g.tiles_all : g.tiles_pre;
    for (int sh = 0; sh <= SH_MIX; sh++)
      launch_gemm(c->stream, probs, segs, t + r[sh], r[sh + 1] - r[sh], s, c->gemm_grid, sh, prec);
  } else {
    const int np = with_pandya ? (int)g.hprobs.size() : g.npre;
    for (int pi = 0; pi < np; pi++) {
      const int n = g.ptile_off[pi + 1] - g.ptile_off[pi];
      launch_gemm(c->stream, probs, segs, g.ptiles + g.ptile_off[pi], n, s, 0, g.hprobs[pi].shape, prec);
    }
  }
  HIPC(hipGetLastError());
  return QD_OK;
}

// products phase: G1 (+ Pandya when both CC sets are non-zero)
int run_products(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *out, const qd_cc *hcc, const qd_cc *ecc) {
  const bool pd = pandya_on(hcc, ecc);
  Slots s = make_slots(c, r1, r2, out, ecc, hcc);
  // grouped FP64 and BF16 slices: the Pandya product and its combine run
  // fused; per-channel launches and FP32 keep product GEMM + combine pass
  const bool fused = pd && c->launch_mode == 0 && pandya_fused_prec(c->gemm_prec);
  {
    Timer tm(c, TC_PRODUCTS);
    int rc = run_gemm_list(c, c->g_all, pd && !fused, s);
    if (rc) return rc;
    if (fused) {
      launch_pandya_fused(c->stream, c->dt, ecc->d, hcc->d, c->P, c->ftp, c->nftp, hcc->herm, ecc->herm, c->gemm_prec,
                        CcSlices{c->cs, c->cs_len, c->cs_off, c->cs_units, c->ncs_units});
      HIPC(hipGetLastError());
      c->d_sign = -hcc->herm * ecc->herm;
    }
  }
  if (pd && !fused) {
    Timer tc(c, TC_OTHER);
    c->d_sign = 0.0;
    launch_pandya_combine(c->stream, c->dt, c->P, c->ptp, c->nptp, hcc->herm, ecc->herm);
    HIPC(hipGetLastError());
  }
  return QD_OK;
}

int run_acc(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *out) {
  Slots s = make_slots(c, r1, r2, out, nullptr, nullptr);
  Timer tm(c, TC_ACCUM);
  int rc = run_gemm_list(c, c->g_acc, false, s);
  if (rc) return rc;
  // the Vpphh and Vppph sums come in two halves (shorter K walks); their sum
  // goes to the same sub-blocks of wl, which the two-body pass reads
  launch_sum_halves(c->stream, c->w3, c->w4, c->wl, c->half_segs, c->half_tiles, c->nhalf_tiles);
  HIPC(hipGetLastError());
  return QD_OK;
}

// The two-body pass in two parts: TB_PART_PRE = the channel blocks of f and
// the ket-side slabs of the boxes too large for one task (they read only the
// operands, so they may run early), TB_PART_MAIN = every stored output (1b x
// 2b, ladder and Pandya terms), adding to what the slabs wrote.
enum { TB_PART_PRE = 1, TB_PART_MAIN = 2 };
// host-only flag of run_twobody: the square sub-blocks' ladder terms are
// already in wl (launched by the caller right after the products)
constexpr int TB_LSQ_DONE = 16;
void run_ladder_sq(qd_context *c, const qd_op *r1, const qd_op *r2) {
  if (c->ldirect) return;  // the box pass forms them from the products
  launch_ladder_sq(c->stream, c->dt, c->w1, c->w2, c->w3, c->wl, c->lsq, c->nlsq, r1->herm, r2->herm,
                   c->nz_on ? c->nz_mask : (uint32_t)ZB_ALL);
}
// paired: the main pass runs the paired box tasks (one staged grid writes the
// outputs of both orientations of a class-pair box); valid when the operators
// obey their herm flags and any Pandya input was filled by calc_cc of r1, r2
int run_twobody(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *out, int flags, const qd_cc *hcc,
                const qd_cc *ecc, int part = TB_PART_PRE | TB_PART_MAIN, bool paired = false) {
  Ops o{r1->d, r2->d, r1->herm, r2->herm};
  Timer tm(c, TC_TWOBODY);
  const bool lsq_done = (flags & TB_LSQ_DONE) != 0;
  flags &= ~TB_LSQ_DONE;
  paired = paired && c->box_pairs;  // the same in both parts
  // r1's sub-block types that are nonzero (the BCH terms' Omega: fph, Vpphh)
  flags |= (int)((c->nz_on ? c->nz_mask : (uint32_t)ZB_ALL) << TB_NZ_SHIFT);
  if ((part & TB_PART_PRE) && (flags & TB_122)) {
    launch_fclass(c->stream, c->dt, o, c->fcls);
    launch_box2b(c->stream, c->dt, o, out->d, paired ? c->bt2p : c->bt2, paired ? c->nbt2p : c->nbt2, c->blist, flags,
                 c->w1, c->w2, c->w3, c->wl, c->fcls, nullptr, 0.0, c->box_wide);
  }
  if (part & TB_PART_MAIN) {
    if ((flags & TB_LADDER) && !lsq_done) run_ladder_sq(c, r1, r2);
    if (c->ldirect) flags |= TB_LDIRECT;
    launch_box2b(c->stream, c->dt, o, out->d, paired ? c->bt1p : c->bt1, paired ? c->nbt1p : c->nbt1, c->blist,
                 flags | (c->d_stage ? TB_DSTAGE : 0), c->w1, c->w2, c->w3, c->wl, c->fcls,
                 (flags & TB_PANDYA) ? c->P : nullptr, c->d_sign, c->box_wide);
  }
  HIPC(hipGetLastError());
  return QD_OK;
}

// ecc / hcc: calc_cc of r1 / r2 (the Pandya inputs) or null; products: the
// products pass of the same operands has run (it holds the one-body products,
// in FP64 when the GEMM precision is FP64)
int run_onebody(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *out, int flags,
                const qd_cc *ecc = nullptr, const qd_cc *hcc = nullptr, bool products = false) {
  Ops o{r1->d, r2->d, r1->herm, r2->herm};
  Timer tm(c, TC_ONEBODY);
  if ((flags & OB_111) && !(products && c->gemm_prec == GEMM_FP64))
    launch_gemm_small(c->stream, c->g_1b.probs, c->g_1b.segs, c->ob_tiles, c->nob_tiles,
                      make_slots(c, r1, r2, out, nullptr, nullptr));
  launch_onebody(c->stream, c->dt, o, out->d, flags, c->w1, c->w2, c->w3, c->u1, c->u2, ecc ? ecc->d : nullptr,
                 hcc ? hcc->d : nullptr, c->ob221, c->nob221, c->g221);
  HIPC(hipGetLastError());
  return QD_OK;
}

int read_scalars(qd_context *c, int n, double *out) {
  HIPC(hipMemcpyAsync(c->host_scal, c->scal, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++) out[i] = c->host_scal[i];
  return QD_OK;
}

// fill: FILL_H makes hcc = calc_cc(r2), FILL_E makes ecc = calc_cc(r1) first
// (the dGam sequence, src/gs_IMSRG.f90:223-273).
//
// Schedule.  Main stream: ladder products -> (CC ready) fused Pandya product
// -> (side work ready) two-body box pass.  Side stream (default priority,
// forked and joined with events, which graph capture records as edges): the
// CC fills, the accumulating products (they read only r1, r2 and write the
// free Vpphh/Vphhh/Vppph sub-blocks of w3), the ket-side slabs of the
// two-body pass (operands only), the 121 mat-vecs (u1, u2) and the zero-body
// traces (r3[0]), then, once the products phase is done, the rest of the
// one-body part (r3's one-body region).  The two-body pass writes only r3's
// two-body region.  With the timing events on, everything runs on one stream
// (mat-vecs and zero-body traces after the products) so that each class is
// timed alone.
enum { FILL_H = 1, FILL_E = 2 };
// e_from_r1: the caller filled ecc by calc_cc of r1 for this commutator
int full_commutator(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *r3, const qd_cc *hcc,
                    const qd_cc *ecc, int fill = 0, bool e_from_r1 = false) {
  int rc = QD_OK;
  qd_cc *fh = (fill & FILL_H) ? const_cast<qd_cc *>(hcc) : nullptr;
  qd_cc *fe = (fill & FILL_E) ? const_cast<qd_cc *>(ecc) : nullptr;
  const bool side = !c->timing;
  hipStream_t main = c->stream;
  auto on = [&](hipStream_t st) { c->stream = st; };
  if (fh) fh->zero = false;  // host flags; the fills are launched below
  if (fe) fe->zero = false;
  const bool pd = pandya_on(hcc, ecc);
  const bool fused = pd && c->launch_mode == 0 && pandya_fused_prec(c->gemm_prec);
  const int tbf = TB_122 | TB_LADDER | (pd ? TB_PANDYA : 0);
  // paired boxes: without the Pandya term, or with CC sets that are calc_cc
  // of r2 and r1 under the operators' own herm flags
  const bool paired = !pd || ((fill & FILL_H) && ((fill & FILL_E) || e_from_r1) && hcc->herm == r2->herm &&
                              ecc->herm == r1->herm);
  // Per-channel launches (the plain form): the products go first when they
  // need no CC fills, so that the many small GEMM launches of the main
  // stream start before the side stream's (N=20 R=16 flow 1.28 -> 1.00 s).
  // The grouped form keeps the side stream's work first (issued the other
  // way its flow went 0.121 -> 0.127 s).
  const bool first_products = side && c->launch_mode != 0 && !(fh || fe);  // (never fused)
  if (side) HIPC(hipEventRecord(c->ev_fork, main));
  if (first_products) {
    if ((rc = run_products(c, r1, r2, r3, hcc, ecc))) return rc;
    run_ladder_sq(c, r1, r2);
  }
  if (side) {
    HIPC(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    on(c->aux);
  }
  {
    Timer tm(c, TC_CALCCC);
    if (fh) launch_calc_cc(c->stream, c->dt, r2->d, r2->herm, fh->d, c->cc_segs, c->cc_tiles, c->ncc_tiles);
    if (fe) launch_calc_cc(c->stream, c->dt, r1->d, r1->herm, fe->d, c->cc_segs, c->cc_tiles, c->ncc_tiles);
  }
  if (hipGetLastError() != hipSuccess) rc = fail(QD_E_HIP, "calc_cc launch failed");
  // the 121 mat-vecs and the zero-body traces read only the operands and CC
  // sets: on the side stream right after the work the two-body pass waits
  // for, beside the Pandya product (N=20 R=14 step 0.484 -> 0.479 ms, N=20
  // R=16 without the Pandya term 0.985 -> 0.938 ms; ahead of the accumulating
  // products they delayed the two-body pass: 0.493 ms)
  auto early_1b = [&]() {
    rc = run_onebody(c, r1, r2, r3, OB_121 | OB_MV, pd ? ecc : nullptr, pd ? hcc : nullptr, true);
    if (!rc) {
      Ops o{r1->d, r2->d, r1->herm, r2->herm};
      Timer tm(c, TC_OTHER);
      launch_zerobody(c->stream, c->dt, o, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, c->partials, r3->d, 3);
      if (hipGetLastError() != hipSuccess) rc = fail(QD_E_HIP, "zero-body launch failed");
    }
  };
  if (side) {
    HIPC(hipEventRecord(c->ev_cc, c->aux));
    if (!rc) rc = run_acc(c, r1, r2, r3);
    if (!rc) rc = run_twobody(c, r1, r2, r3, tbf, hcc, ecc, TB_PART_PRE, paired);
    HIPC(hipEventRecord(c->ev_acc, c->aux));
    if (!rc) early_1b();
    on(main);
  }
  if (rc) return rc;
  // products phase (the rest of it)
  if (fused) {
    Timer tm(c, TC_PRODUCTS);
    if (!first_products) {
      if ((rc = run_gemm_list(c, c->g_all, false, make_slots(c, r1, r2, r3, ecc, hcc)))) return rc;
      if (side) run_ladder_sq(c, r1, r2);  // while the CC fills finish on the side stream
    }
    if (side) HIPC(hipStreamWaitEvent(main, c->ev_cc, 0));
    launch_pandya_fused(c->stream, c->dt, ecc->d, hcc->d, c->P, c->ftp, c->nftp, hcc->herm, ecc->herm, c->gemm_prec,
                        CcSlices{c->cs, c->cs_len, c->cs_off, c->cs_units, c->ncs_units});
    HIPC(hipGetLastError());
    c->d_sign = -hcc->herm * ecc->herm;
  } else if (!first_products) {
    // the products read the CC sets only with the Pandya term; wait for the
    // fills only when there are any
    if (side && (fh || fe)) HIPC(hipStreamWaitEvent(main, c->ev_cc, 0));
    if ((rc = run_products(c, r1, r2, r3, hcc, ecc))) return rc;
    if (side) run_ladder_sq(c, r1, r2);
  }
  if (!side && (rc = run_acc(c, r1, r2, r3))) return rc;
  // one-body and zero-body parts
  if (side) {
    HIPC(hipEventRecord(c->ev_fork, main));
    HIPC(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    on(c->aux);
  }
  // with the Pandya inputs on hand, the 121 mat-vecs read them (hcc and ecc
  // are calc_cc of r2 and r1 by the entry point's contract)
  rc = run_onebody(c, r1, r2, r3, OB_111 | OB_121 | OB_221 | (side ? OB_NOMV : 0), pd ? ecc : nullptr,
                   pd ? hcc : nullptr, true);
  if (!rc && !side) {
    Ops o{r1->d, r2->d, r1->herm, r2->herm};
    Timer tm(c, TC_OTHER);
    launch_zerobody(c->stream, c->dt, o, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, c->partials, r3->d, 3);
    if (hipGetLastError() != hipSuccess) rc = fail(QD_E_HIP, "zero-body launch failed");
  }
  if (side) {
    HIPC(hipEventRecord(c->ev_join, c->aux));
    on(main);
    HIPC(hipStreamWaitEvent(main, c->ev_acc, 0));
  }
  if (rc) return rc;
  if ((rc = run_twobody(c, r1, r2, r3, side ? tbf | TB_LSQ_DONE : tbf, hcc, ecc,
                        side ? TB_PART_MAIN : TB_PART_PRE | TB_PART_MAIN, paired)))
    return rc;
  if (side) HIPC(hipStreamWaitEvent(main, c->ev_join, 0));
  return QD_OK;
}

// Pulay (DIIS) weights: minimise |sum_i c_i r_i|^2 subject to sum_i c_i = 1,
// from the Gram matrix B (n x n, row-major) of the residuals.  Gaussian
// elimination with partial pivoting on the bordered (n+1) system, B scaled
// by its largest diagonal.  Returns false if the system is singular.
bool diis_weights(int n, const double *B, double *c) {
  const int m = n + 1;
  std::vector<double> A((size_t)m * (m + 1), 0.0);
  double sc = 0;
  for (int i = 0; i < n; i++) sc = std::max(sc, B[i * n + i]);
  if (!(sc > 0)) return false;
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) A[(size_t)i * (m + 1) + j] = B[i * n + j] / sc;
    A[(size_t)i * (m + 1) + n] = 1.0;
    A[(size_t)n * (m + 1) + i] = 1.0;
  }
  A[(size_t)n * (m + 1) + m] = 1.0;  // right-hand side (0, ..., 0, 1)
  for (int k = 0; k < m; k++) {
    int piv = k;
    for (int i = k + 1; i < m; i++)
      if (std::fabs(A[(size_t)i * (m + 1) + k]) > std::fabs(A[(size_t)piv * (m + 1) + k])) piv = i;
    if (std::fabs(A[(size_t)piv * (m + 1) + k]) < 1e-14) return false;
    if (piv != k)
      for (int j = 0; j <= m; j++) std::swap(A[(size_t)k * (m + 1) + j], A[(size_t)piv * (m + 1) + j]);
    for (int i = k + 1; i < m; i++) {
      const double f = A[(size_t)i * (m + 1) + k] / A[(size_t)k * (m + 1) + k];
      for (int j = k; j <= m; j++) A[(size_t)i * (m + 1) + j] -= f * A[(size_t)k * (m + 1) + j];
    }
  }
  std::vector<double> x(m);
  for (int i = m - 1; i >= 0; i--) {
    double r = A[(size_t)i * (m + 1) + m];
    for (int j = i + 1; j < m; j++) r -= A[(size_t)i * (m + 1) + j] * x[j];
    x[i] = r / A[(size_t)i * (m + 1) + i];
  }
  for (int i = 0; i < n; i++) {
    if (!std::isfinite(x[i])) return false;
    c[i] = x[i];
  }
  return true;
}

int norm_into(qd_context *c, const qd_op *r, int slot) {
  Timer tm(c, TC_OTHER);
  launch_norm2(c->stream, r->d, c->ch.neq, c->partials, c->scal + slot);
  HIPC(hipGetLastError());
  return QD_OK;
}

int ensure_scratch(qd_context *c) {
  int rc;
  if (!c->bch_a) {
    if ((rc = qd_op_create(c, 1, &c->bch_a))) return rc;
    if ((rc = qd_op_create(c, 1, &c->bch_b))) return rc;
    if ((rc = qd_cc_create(c, 1.0, &c->cc_t1))) return rc;
    if ((rc = qd_cc_create(c, -1.0, &c->cc_t2))) return rc;
  }
  return QD_OK;
}

int64_t layout_table(const Channels &ch, int which, int64_t *out, int64_t len) {
  std::vector<int64_t> v;
  switch (which) {
    case 0: for (int q = 0; q < ch.nblock; q++) { v.push_back(ch.nh[q]); v.push_back(ch.np[q]); v.push_back(ch.nb[q]); } break;
    case 1: v.assign(ch.off.begin(), ch.off.end()); break;
    case 2: for (uint32_t a : ch.iarr) { v.push_back(ia_type(a)); v.push_back(ia_blk(a)); v.push_back(ia_idx(a)); } break;
    case 3: v.assign(ch.cc_map.begin(), ch.cc_map.end()); break;
    case 4: for (int q = 0; q < ch.ncc; q++) { v.push_back(ch.cc_r[q]); v.push_back(ch.cc_nb[q]); } break;
    case 5: for (uint32_t pr : ch.pairs) { v.push_back(lo16(pr)); v.push_back(hi16(pr)); } break;
    case 6: v.push_back(ch.neq); v.push_back(ch.cc_size); v.push_back(ch.nblock); v.push_back(ch.ncc); break;
    case 7: for (int i = 0; i < ch.M; i++) { v.push_back(ch.ml[i]); v.push_back(ch.ms[i]); } break;
    case 8: {  // calc_cc gather codes in device order, with each element's flat index in reference order
      if (ch.neq >= (int64_t)CC_ZERO) return fail(QD_E_LAYOUT, "no gather table: neq too large");
      const std::vector<uint32_t> m = cc_gather_map(ch);
      std::vector<int64_t> ref(m.size(), 0);
      for (int q = 0; q < ch.ncc; q++) {
        const int r = ch.cc_r[q], s0 = ch.cc_ab_start[q];
        for (int row = 0; row < r; row++)
          for (int c = 0; c < ch.cc_nb[q]; c++)
            ref[(size_t)ch.cc_off[q] + ch.cc_dev_row[s0 + row] + (size_t)c * r] = ch.cc_off[q] + row + (int64_t)c * r;
      }
      for (size_t e = 0; e < m.size(); e++) { v.push_back(m[e]); v.push_back(ref[e]); }
      break;
    }
    default: return fail(QD_E_ARG, "unknown table %d", which);
  }
  if (out) {
    if (len < (int64_t)v.size()) return fail(QD_E_ARG, "table needs %lld entries", (long long)v.size());
    memcpy(out, v.data(), v.size() * sizeof(int64_t));
  }
  return (int64_t)v.size();
}

}  // namespace

// ===========================================================================
extern "C" {

const char *qd_last_error(void) { return g_err.c_str(); }

int qd_context_create(int nbody, int nshell, qd_context **out) {
  if (!out) return fail(QD_E_ARG, "out is null");
  *out = nullptr;
  qd_context *c = new qd_context();
  std::string why = build_channels(nbody, nshell, c->ch);
  if (!why.empty()) { delete c; return fail(QD_E_LAYOUT, "%s", why.c_str()); }
  int rc = QD_OK;
  auto bail = [&](int r) { qd_context_destroy(c); return r; };
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return fail(QD_E_HIP, "hipStreamCreate failed"); }
  if (hipStreamCreate(&c->aux) != hipSuccess || hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_cc, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_acc, hipEventDisableTiming) != hipSuccess)
    return bail(fail(QD_E_HIP, "side stream or events could not be created"));
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0) {
      // work-groups per CU of the persistent GEMM grid (QD_GEMM_WG_PER_CU, 1..16;
      // default 8: two more than the FP64 mixed launch keeps resident, which
      // then start as the first ones finish and even out the tail.  N=20
      // R=14 products class 0.201 / 0.195 ms and accumulating list 0.076 /
      // 0.079 ms at 6 / 8; N=20 R=16 grouped flow 0.120 / 0.117 s)
      const char *e = getenv("QD_GEMM_WG_PER_CU");
      const int per = e && atoi(e) >= 1 && atoi(e) <= 16 ? atoi(e) : 8;
      c->gemm_grid = per * ncu;
    }
  }
  const Channels &ch = c->ch;
  DevTab &t = c->dt;
  t.n = ch.n; t.M = ch.M; t.mp = ch.mp; t.nblock = ch.nblock; t.ncc = ch.ncc; t.nch = ch.nch;
  t.ml0 = ch.ml0; t.mlspan = ch.mlspan; t.neq = ch.neq;
  t.off_fhh = ch.off_fhh; t.off_fpp = ch.off_fpp; t.off_fph = ch.off_fph;
  std::vector<int2> ccmap((size_t)ch.M * ch.M);
  for (size_t i = 0; i < ccmap.size(); i++) {
    const int blk = ch.cc_map[2 * i], row = ch.cc_map[2 * i + 1];
    ccmap[i] = make_int2(blk, blk < 0 ? -1 : ch.cc_dev_row[ch.cc_ab_start[blk] + row]);
  }
#define UP(vec, field) if ((rc = dev_upload(c, vec, (decltype(vec)::value_type **)&t.field))) return bail(rc)
  UP(ch.nh, nh); UP(ch.np, np); UP(ch.nb, nb); UP(ch.off, off); UP(ch.pairs, pairs);
  UP(ch.hh_start, hh_start); UP(ch.ph_start, ph_start); UP(ch.pp_start, pp_start); UP(ch.iarr, iarr);
  UP(ch.spch, spch); UP(ch.ch_start, ch_start); UP(ch.ch_states, ch_states); UP(ch.ch_lookup, ch_lookup);
  UP(ch.ml, ml); UP(ch.ms, ms); UP(ch.cc_r, cc_r); UP(ch.cc_nb, cc_nb); UP(ch.cc_off, cc_off);
  UP(ch.cc_poff, cc_poff); UP(ch.cc_ab_dev, cc_ab); UP(ch.cc_hp, cc_hp); UP(ch.cc_ab_start, cc_ab_start);
  UP(ch.cc_hp_start, cc_hp_start); UP(ccmap, cc_map);
  UP(ch.cc_rs, cc_rs); UP(ch.cc_run, cc_run);
  UP(ch.cp_a, cp_a); UP(ch.cp_b, cp_b); UP(ch.cp_grid_off, cp_grid_off); UP(ch.cp_grid, cp_grid);
  UP(ch.fcls_off, fcls_off);
  t.fcls_size = ch.fcls_size;
  t.cc_size = ch.cc_size;
  t.p_size = ch.p_size;
  {
    std::vector<int> rowblk((size_t)ch.M * ch.M, 0);
    for (int K = 0; K < ch.ncc; K++)
      for (int d = 0; d < ch.cc_r[K]; d++) rowblk[(size_t)ch.cc_ab_start[K] + d] = K;
    UP(rowblk, cc_rowblk);
  }
  t.cc_gmap = nullptr;
  // QD_GATHER_TABLES=0: no gather tables (the calc_cc and 221 kernels then
  // resolve the pair tables per element, as they must for neq >= CC_ZERO;
  // tests use it to check that path)
  const char *gt = getenv("QD_GATHER_TABLES");
  c->gather_tables = !(gt && gt[0] == '0');
  if (ch.neq < (int64_t)CC_ZERO && c->gather_tables) {
    const std::vector<uint32_t> gmap = cc_gather_map(ch);
    UP(gmap, cc_gmap);
  }
#undef UP
  void *p;
  if ((rc = dev_alloc(c, sizeof(double) * ch.neq, &p))) return bail(rc); c->w1 = (double *)p;
  if ((rc = dev_alloc(c, sizeof(double) * ch.neq, &p))) return bail(rc); c->w2 = (double *)p;
  if ((rc = dev_alloc(c, sizeof(double) * ch.neq, &p))) return bail(rc); c->w3 = (double *)p;
  if ((rc = dev_alloc(c, sizeof(double) * ch.neq, &p))) return bail(rc); c->wl = (double *)p;
  // second halves of the summed products of Vpphh and Vppph
  if ((rc = dev_alloc(c, sizeof(double) * ch.neq, &p))) return bail(rc); c->w4 = (double *)p;
  // one element more than the Pandya products: D[p_size] stays zero (the box
  // pass reads it for a term without a D block)
  if ((rc = dev_alloc(c, sizeof(double) * (ch.p_size + 1), &p))) return bail(rc); c->P = (double *)p;
  if ((rc = dev_alloc(c, sizeof(double) * 2 * std::max(ch.fcls_size, 1), &p))) return bail(rc);
  c->fcls = (double *)p;
  {
    // the wide box form where the narrow one would split more than a tenth
    // of the grid elements into slabs (QD_BOX_WIDE=0/1 decides instead)
    double all = 0, split = 0;
    auto csz = [&](int h) { return ch.ch_start[h + 1] - ch.ch_start[h]; };
    for (int q = 0; q < ch.nblock; q++)
      for (int gr = ch.cp_start[q]; gr < ch.cp_start[q + 1]; gr++)
        for (int gc = ch.cp_start[q]; gc < ch.cp_start[q + 1]; gc++) {
          if (ch.cp_tmax[gr] < ch.cp_tmin[gc]) continue;
          const int nr = csz(ch.cp_a[gr]) * csz(ch.cp_b[gr]), nc = csz(ch.cp_a[gc]) * csz(ch.cp_b[gc]);
          all += (double)nr * nc;
          if (nr * (nc | 1) > BX_VCAP) split += (double)nr * nc;
        }
    c->box_wide = split > 0.1 * all;
    if (const char *env = getenv("QD_BOX_WIDE")) c->box_wide = env[0] == '1';
    c->d_stage = c->box_wide;
    if (const char *env = getenv("QD_DSTAGE")) c->d_stage = env[0] != '0';
    c->box_cap = c->box_wide ? BX_WIDE_VCAP : BX_VCAP;
    c->box_rcap = c->box_wide ? BX_WIDE_RCAP : BX_RCAP;
  }
  if (const char *env = getenv("QD_BOX_CAP")) {
    // grid elements per box task (tests use a small cap to exercise the slabs)
    const int v = atoi(env);
    if (v >= 16 && v <= c->box_cap) c->box_cap = v;
  }
  if (const char *env = getenv("QD_BOX_PAIRS")) c->box_pairs = env[0] != '0';
  if (const char *env = getenv("QD_LDIRECT")) c->ldirect = env[0] != '0';
  if ((rc = dev_alloc(c, sizeof(double) * ch.M * ch.M, &p))) return bail(rc); c->u1 = (double *)p;
  if ((rc = dev_alloc(c, sizeof(double) * ch.M * ch.M, &p))) return bail(rc); c->u2 = (double *)p;
  if ((rc = dev_alloc(c, sizeof(double) * 1024 * 64, &p))) return bail(rc); c->partials = (double *)p;
  if ((rc = dev_alloc(c, sizeof(uint32_t) * 2, &p))) return bail(rc); c->nz = (uint32_t *)p;
  if (const char *env = getenv("QD_ZSKIP")) c->zskip = env[0] != '0';
  if ((rc = dev_alloc(c, sizeof(double) * 8, &p))) return bail(rc); c->scal = (double *)p;
  if (hipHostMalloc((void **)&c->host_scal, sizeof(double) * 8, hipHostMallocDefault) != hipSuccess)
    return bail(fail(QD_E_HIP, "hipHostMalloc failed"));
  if ((rc = build_worklists(c))) return bail(rc);
  if (c->npphh_tiles + 1 > 1024 * 64)
    return bail(fail(QD_E_LAYOUT, "too many pphh tiles for the reduction buffer"));
  if ((rc = noalloc_setup(&c->noalloc_scratch))) return bail(rc);
  c->noalloc_cap = NOALLOC_PARTS;
  if (const char *env = getenv("QD_NOALLOC_CAP")) {
    // a lower cap exercises the size check without a huge channel structure
    const int v = atoi(env);
    if (v >= 1 && v <= NOALLOC_PARTS) c->noalloc_cap = v;
  }
  *out = c;
  return QD_OK;
}

int qd_context_destroy(qd_context *c) {
  if (!c) return QD_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->bch_a) { qd_op_destroy(c->bch_a); qd_op_destroy(c->bch_b); qd_cc_destroy(c->cc_t1); qd_cc_destroy(c->cc_t2); }
  for (auto &e : c->events) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
  for (auto &kv : c->graphs) hipGraphExecDestroy(kv.second);
  for (void *p : c->allocs) hipFree(p);
  if (c->host_scal) hipHostFree(c->host_scal);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->ev_cc) hipEventDestroy(c->ev_cc);
  if (c->ev_acc) hipEventDestroy(c->ev_acc);
  if (c->aux) hipStreamDestroy(c->aux);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return QD_OK;
}

int qd_context_info(const qd_context *c, qd_info *o) {
  if (!c || !o) return fail(QD_E_ARG, "null argument");
  const Channels &ch = c->ch;
  o->nbody = ch.n; o->nshell = ch.R; o->msp = ch.M; o->nblock = ch.nblock; o->ncc = ch.ncc;
  o->neq = ch.neq; o->cc_size = ch.cc_size;
  o->n2b = ch.neq - ch.off_2b;
  return QD_OK;
}

int64_t qd_context_table(const qd_context *c, int which, int64_t *out, int64_t len) {
  if (!c) return fail(QD_E_ARG, "null context");
  return layout_table(c->ch, which, out, len);
}

int64_t qd_layout_table(int nbody, int nshell, int which, int64_t *out, int64_t len) {
  Channels ch;
  std::string why = build_channels(nbody, nshell, ch);
  if (!why.empty()) return fail(QD_E_LAYOUT, "%s", why.c_str());
  return layout_table(ch, which, out, len);
}

int qd_context_sync(qd_context *c) {
  if (!c) return fail(QD_E_ARG, "null context");
  HIPC(hipStreamSynchronize(c->stream));
  return QD_OK;
}

// ---- operators -----------------------------------------------------------
int qd_op_create(qd_context *c, int herm, qd_op **out) {
  if (!c || !out) return fail(QD_E_ARG, "null argument");
  if (herm != 1 && herm != -1) return fail(QD_E_ARG, "herm must be +1 or -1");
  void *p = nullptr;
  HIPC(hipMalloc(&p, sizeof(double) * c->ch.neq));
  HIPC(hipMemsetAsync(p, 0, sizeof(double) * c->ch.neq, c->stream));
  *out = new qd_op{c, (double *)p, herm};
  return QD_OK;
}
// Captured graphs are keyed by the device addresses of their operands: drop
// every one that names memory about to be freed, so the cache does not grow
// with short-lived operators (nor replay into a reused address).
static void evict_graphs(qd_context *c, const void *p) {
  for (auto it = c->graphs.begin(); it != c->graphs.end();) {
    const qd_context::GraphKey &k = it->first;
    if (k.g == p || k.prev == p || k.next == p || k.hs == p || k.cca == p || k.ccg == p) {
      hipGraphExecDestroy(it->second);
      it = c->graphs.erase(it);
    } else {
      ++it;
    }
  }
}

int qd_op_destroy(qd_op *o) {
  if (!o) return QD_OK;
  hipStreamSynchronize(o->ctx->stream);
  evict_graphs(o->ctx, o->d);
  hipFree(o->d);
  delete o;
  return QD_OK;
}
int qd_op_set_herm(qd_op *o, int herm) {
  if (!o || (herm != 1 && herm != -1)) return fail(QD_E_ARG, "bad operator or herm");
  o->herm = herm;
  return QD_OK;
}
int qd_op_get_herm(const qd_op *o, int *herm) {
  if (!o || !herm) return fail(QD_E_ARG, "null argument");
  *herm = o->herm;
  return QD_OK;
}
int qd_op_upload(qd_op *o, const double *vec, int64_t neq) {
  if (!o || !vec) return fail(QD_E_ARG, "null argument");
  if (neq != o->ctx->ch.neq) return fail(QD_E_ARG, "neq %lld != %lld", (long long)neq, (long long)o->ctx->ch.neq);
  HIPC(hipMemcpyAsync(o->d, vec, sizeof(double) * neq, hipMemcpyHostToDevice, o->ctx->stream));
  HIPC(hipStreamSynchronize(o->ctx->stream));
  return QD_OK;
}
int qd_op_download(const qd_op *o, double *vec, int64_t neq) {
  if (!o || !vec) return fail(QD_E_ARG, "null argument");
  if (neq != o->ctx->ch.neq) return fail(QD_E_ARG, "neq %lld != %lld", (long long)neq, (long long)o->ctx->ch.neq);
  HIPC(hipMemcpyAsync(vec, o->d, sizeof(double) * neq, hipMemcpyDeviceToHost, o->ctx->stream));
  HIPC(hipStreamSynchronize(o->ctx->stream));
  return QD_OK;
}
double *qd_op_device_ptr(qd_op *o) { return o ? o->d : nullptr; }

// ---- CC sets ---------------------------------------------------------------
int qd_cc_create(qd_context *c, double herm, qd_cc **out) {
  if (!c || !out) return fail(QD_E_ARG, "null argument");
  void *p = nullptr;
  size_t bytes = sizeof(double) * std::max<int64_t>(c->ch.cc_size, 1);
  HIPC(hipMalloc(&p, bytes));
  HIPC(hipMemsetAsync(p, 0, bytes, c->stream));
  *out = new qd_cc{c, (double *)p, herm, true};
  return QD_OK;
}
int qd_cc_destroy(qd_cc *cc) {
  if (!cc) return QD_OK;
  hipStreamSynchronize(cc->ctx->stream);
  evict_graphs(cc->ctx, cc->d);
  hipFree(cc->d);
  delete cc;
  return QD_OK;
}
int qd_cc_set_herm(qd_cc *cc, double herm) {
  if (!cc) return fail(QD_E_ARG, "null argument");
  cc->herm = herm;
  return QD_OK;
}
// Row permutation between the reference's CC layout (host) and the device
// row order (every block r x nb column-major at cc_off).
static void cc_rows(const Channels &ch, const double *from, double *to, bool to_device) {
  for (int b = 0; b < ch.ncc; b++) {
    const int r = ch.cc_r[b], nb = ch.cc_nb[b], s0 = ch.cc_ab_start[b];
    const int64_t o = ch.cc_off[b];
    for (int col = 0; col < nb; col++)
      for (int a = 0; a < r; a++) {
        const int64_t ref = o + a + (int64_t)col * r, dev = o + ch.cc_dev_row[s0 + a] + (int64_t)col * r;
        if (to_device) to[dev] = from[ref];
        else to[ref] = from[dev];
      }
  }
}

int qd_cc_upload(qd_cc *cc, const double *x, int64_t n) {
  if (!cc || !x) return fail(QD_E_ARG, "null argument");
  if (n != cc->ctx->ch.cc_size) return fail(QD_E_ARG, "cc_size mismatch");
  std::vector<double> dev((size_t)n);
  cc_rows(cc->ctx->ch, x, dev.data(), true);
  HIPC(hipMemcpyAsync(cc->d, dev.data(), sizeof(double) * n, hipMemcpyHostToDevice, cc->ctx->stream));
  HIPC(hipStreamSynchronize(cc->ctx->stream));
  cc->zero = false;
  return QD_OK;
}
int qd_cc_download(const qd_cc *cc, double *x, int64_t n) {
  if (!cc || !x) return fail(QD_E_ARG, "null argument");
  if (n != cc->ctx->ch.cc_size) return fail(QD_E_ARG, "cc_size mismatch");
  std::vector<double> dev((size_t)n);
  HIPC(hipMemcpyAsync(dev.data(), cc->d, sizeof(double) * n, hipMemcpyDeviceToHost, cc->ctx->stream));
  HIPC(hipStreamSynchronize(cc->ctx->stream));
  cc_rows(cc->ctx->ch, dev.data(), x, false);
  return QD_OK;
}
int qd_cc_zero(qd_context *c, qd_cc *cc) {
  if (!c || !cc) return fail(QD_E_ARG, "null argument");
  HIPC(hipMemsetAsync(cc->d, 0, sizeof(double) * std::max<int64_t>(c->ch.cc_size, 1), c->stream));
  cc->zero = true;
  return QD_OK;
}

// ---- commutator pieces -------------------------------------------------------
#define CHK3(a, b, d)                                 \
  do {                                                \
    int rc_;                                          \
    if (!c) return fail(QD_E_ARG, "null context");    \
    if ((rc_ = check_op(c, a, "r1"))) return rc_;     \
    if ((rc_ = check_op(c, b, "r2"))) return rc_;     \
    if (d && (rc_ = check_op(c, d, "r3"))) return rc_; \
  } while (0)

int qd_commutator_110(qd_context *c, const qd_op *r1, const qd_op *r2, double *out) {
  CHK3(r1, r2, (const qd_op *)nullptr);
  Ops o{r1->d, r2->d, r1->herm, r2->herm};
  launch_zerobody(c->stream, c->dt, o, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, c->partials, c->scal, 1);
  HIPC(hipGetLastError());
  return read_scalars(c, 1, out);
}
int qd_commutator_220(qd_context *c, const qd_op *r1, const qd_op *r2, double *out) {
  CHK3(r1, r2, (const qd_op *)nullptr);
  Ops o{r1->d, r2->d, r1->herm, r2->herm};
  launch_zerobody(c->stream, c->dt, o, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, c->partials, c->scal, 2);
  HIPC(hipGetLastError());
  return read_scalars(c, 1, out);
}
int qd_commutator_111(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *r3) {
  CHK3(r1, r2, r3);
  return run_onebody(c, r1, r2, r3, OB_111);
}
int qd_commutator_121(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *r3) {
  CHK3(r1, r2, r3);
  return run_onebody(c, r1, r2, r3, OB_121 | OB_ACCUM);
}
int qd_commutator_221(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *r3) {
  CHK3(r1, r2, r3);
  int rc;
  if ((rc = run_products(c, r1, r2, r3, nullptr, nullptr))) return rc;
  return run_onebody(c, r1, r2, r3, OB_221 | OB_ACCUM);
}
int qd_commutator_122(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *r3) {
  CHK3(r1, r2, r3);
  return run_twobody(c, r1, r2, r3, TB_122, nullptr, nullptr, TB_PART_PRE | TB_PART_MAIN, true);
}
int qd_commutator_222(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *r3, const qd_cc *hcc,
                      const qd_cc *etacc) {
  CHK3(r1, r2, r3);
  int rc;
  const bool pd = pandya_on(hcc, etacc);
  if ((rc = run_products(c, r1, r2, r3, hcc, etacc))) return rc;
  if ((rc = run_acc(c, r1, r2, r3))) return rc;
  return run_twobody(c, r1, r2, r3, TB_LADDER | TB_ACCUM | (pd ? TB_PANDYA : 0), hcc, etacc,
                     TB_PART_PRE | TB_PART_MAIN, !pd);
}
int qd_calc_cc(qd_context *c, const qd_op *h, qd_cc *cc) {
  if (!c || !h || !cc) return fail(QD_E_ARG, "null argument");
  {
    Timer tm(c, TC_CALCCC);
    launch_calc_cc(c->stream, c->dt, h->d, h->herm, cc->d, c->cc_segs, c->cc_tiles, c->ncc_tiles);
    HIPC(hipGetLastError());
  }
  cc->zero = false;
  return QD_OK;
}
int qd_build_white(qd_context *c, qd_op *gen, const qd_op *rec) {
  if (!c || !gen || !rec) return fail(QD_E_ARG, "null argument");
  Timer tm(c, TC_OTHER);
  launch_white(c->stream, c->dt, rec->d, rec->herm, gen->d, c->pphh_segs, c->pphh_tiles, c->npphh_tiles);
  HIPC(hipGetLastError());
  return QD_OK;
}

// ---- operator algebra ---------------------------------------------------------
int qd_add_arrays(qd_context *c, const qd_op *r1, double s1, const qd_op *r2, double s2, qd_op *r3) {
  CHK3(r1, r2, r3);
  Timer tm(c, TC_OTHER);
  launch_axpby(c->stream, c->ch.neq, s1, r1->d, s2, r2->d, r3->d);
  HIPC(hipGetLastError());
  return QD_OK;
}
int qd_copy_arrays(qd_context *c, const qd_op *r1, qd_op *r2) {
  CHK3(r1, r2, (const qd_op *)nullptr);
  if (r1 == r2) return QD_OK;
  HIPC(hipMemcpyAsync(r2->d, r1->d, sizeof(double) * c->ch.neq, hipMemcpyDeviceToDevice, c->stream));
  return QD_OK;
}
int qd_set_to_zero(qd_context *c, qd_op *r) {
  if (!c || !r) return fail(QD_E_ARG, "null argument");
  HIPC(hipMemsetAsync(r->d, 0, sizeof(double) * c->ch.neq, c->stream));
  return QD_OK;
}
int qd_mat_2_norm(qd_context *c, const qd_op *r, double *out) {
  if (!c || !r || !out) return fail(QD_E_ARG, "null argument");
  int rc;
  if ((rc = norm_into(c, r, 0))) return rc;
  double s;
  if ((rc = read_scalars(c, 1, &s))) return rc;
  *out = std::sqrt(s);
  return QD_OK;
}
int qd_mbpt2(qd_context *c, const qd_op *h, double *out) {
  if (!c || !h || !out) return fail(QD_E_ARG, "null argument");
  {
    Timer tm(c, TC_OTHER);
    launch_mbpt2(c->stream, c->dt, h->d, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, c->partials, c->scal);
    HIPC(hipGetLastError());
  }
  return read_scalars(c, 1, out);
}

// Allocation-free MBPT2: the kernels of qd_mbpt2 with the partials and the
// result in the file-scope scratch array instead of the context's buffers; no
// allocation, no timing events.  Every partial is written before it is read,
// so nothing needs clearing between calls.
int qd_mbpt2_noalloc(qd_context *c, const qd_op *h, double *out) {
  if (!c || !h || !out) return fail(QD_E_ARG, "null argument");
  if (h->ctx != c) return fail(QD_E_ARG, "operator belongs to another context");
  if (!c->noalloc_scratch) return fail(QD_E_HIP, "allocation-free scratch not set up for this context");
  if (c->npphh_tiles > c->noalloc_cap)
    return fail(QD_E_ARG, "MBPT2 needs %d reduction partials; the allocation-free scratch holds %d", c->npphh_tiles,
                c->noalloc_cap);
  double *part = c->noalloc_scratch, *dst = c->noalloc_scratch + NOALLOC_PARTS;
  launch_mbpt2(c->stream, c->dt, h->d, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, part, dst);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(out, dst, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return QD_OK;
}

// ---- whole commutator, BCH, Magnus ----------------------------------------------
int qd_commutator(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *r3, const qd_cc *hcc,
                  const qd_cc *etacc, double *e0) {
  CHK3(r1, r2, r3);
  if (r3 == r1 || r3 == r2) return fail(QD_E_ARG, "r3 must not alias r1 or r2");
  int rc;
  if ((rc = full_commutator(c, r1, r2, r3, hcc, etacc))) return rc;
  if (e0) {
    HIPC(hipMemcpyAsync(c->host_scal, r3->d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    *e0 = c->host_scal[0];
  }
  return QD_OK;
}

int qd_commutator_cc(qd_context *c, const qd_op *r1, const qd_op *r2, qd_op *r3, qd_cc *hcc, qd_cc *etacc,
                     double *e0) {
  CHK3(r1, r2, r3);
  if (!hcc || !etacc) return fail(QD_E_ARG, "null CC set");
  if (r3 == r1 || r3 == r2) return fail(QD_E_ARG, "r3 must not alias r1 or r2");
  int rc;
  if (c->use_graphs && !c->timing) {
    // the whole call (both streams) as one graph per operand set; hs = null
    // and the herm code above 100 keep these keys apart from the BCH terms'
    qd_context::GraphKey key{r1->d, r2->d, r3->d, nullptr, hcc->d, etacc->d,
                             100 + 3 * (r1->herm + 1) + (r2->herm + 1), 2, hcc->herm, etacc->herm, 0, ZB_ALL};
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
      hipGraph_t graph;
      HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc = full_commutator(c, r1, r2, r3, hcc, etacc, FILL_H | FILL_E);
      hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (rc) return rc;
      if (ec != hipSuccess) return fail(QD_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
      hipGraphExec_t exec;
      HIPC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      HIPC(hipGraphDestroy(graph));
      it = c->graphs.emplace(key, exec).first;
    }
    HIPC(hipGraphLaunch(it->second, c->stream));
    hcc->zero = etacc->zero = false;
  } else if ((rc = full_commutator(c, r1, r2, r3, hcc, etacc, FILL_H | FILL_E))) {
    return rc;
  }
  if (e0) {
    HIPC(hipMemcpyAsync(c->host_scal, r3->d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    *e0 = c->host_scal[0];
  }
  return QD_OK;
}

int qd_bch_expand(qd_context *c, qd_op *hs, const qd_op *g, const qd_op *h, const qd_cc *hcc, const qd_cc *occ,
                  int *nterms) {
  CHK3(g, h, hs);
  int rc;
  if ((rc = ensure_scratch(c))) return rc;
  // Pandya inputs: the caller's CC sets are used when non-zero; the
  // reference's Magnus driver passes zero sets (no Pandya term).
  const bool pd = pandya_on(hcc, occ);
  qd_cc *ccg = c->cc_t2, *cca = c->cc_t1;
  if (pd) {  // fill CC of G (r1 slot) and of each AD (r2 slot) as calc_cc would
    if ((rc = qd_calc_cc(c, g, ccg))) return rc;
    ccg->herm = occ->herm;
    cca->herm = hcc->herm;
    ccg->zero = cca->zero = false;  // cca is refilled per term by the launch below
  }
  // the sub-block types of G that are zero in every block (a Magnus Omega
  // holds only fph and Vpphh): the products reading them are skipped in every
  // term, their contributions being exact zeros
  struct NzScope {
    qd_context *c;
    ~NzScope() {
      c->nz_on = false;
      c->nz_mask = ZB_ALL;
    }
  } nz_scope{c};
  if (c->zskip) {
    launch_zero_scan(c->stream, g->d, c->ztiles, c->nztiles, c->nz);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(c->host_scal, c->nz, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    memcpy(&c->nz_mask, c->host_scal, sizeof(uint32_t));
    c->nz_on = true;
    // the lists for this mask exist before any term is captured in a graph
    // (no allocation during capture); the outputs of the problems they leave
    // out are zeroed here, once for all terms (nothing else writes them)
    for (GemmList *gl : {&c->g_all, &c->g_acc}) {
      GemmList::Variant *v;
      if ((rc = gemm_variant(c, *gl, &v))) return rc;
      if (v) launch_zero_chunks(c->stream, make_slots(c, nullptr, nullptr, nullptr, nullptr, nullptr), v->zc, v->nzc);
    }
    HIPC(hipGetLastError());
  }
  // HS = H ; AD_0 = H ; then AD_k = [G, AD_{k-1}], HS += cof(k) AD_k
  if ((rc = qd_copy_arrays(c, h, hs))) return rc;
  const qd_op *prev = h;
  qd_op *buf[2] = {c->bch_a, c->bch_b};
  int k, taken = 0;
  for (k = 2; k <= 12; k++) {
    qd_op *next = buf[k & 1];
    next->herm = 1;
    qd_op adh = *prev;  // AD carries herm = +1 (src/magnus_IMSRG.f90:273-275)
    adh.herm = 1;
    // one term: [G, AD] into next, then in one pass ||INT1|| (= HS before
    // the update), ||INT2|| and HS += cof(k) AD
    auto term = [&]() -> int {
      int r;  // with the Pandya term, cca = calc_cc(AD) is filled first
      if ((r = full_commutator(c, g, &adh, next, pd ? cca : nullptr, pd ? ccg : nullptr, pd ? FILL_H : 0, pd)))
        return r;
      Timer tm(c, TC_OTHER);
      launch_bch_update(c->stream, hs->d, next->d, BCH_COF[k - 2], c->ch.neq, c->partials, c->scal);
      HIPC(hipGetLastError());
      return QD_OK;
    };
    if (c->use_graphs && !c->timing) {
      qd_context::GraphKey key{g->d, prev->d, next->d, hs->d, pd ? cca->d : nullptr, pd ? ccg->d : nullptr,
                               g->herm, pd ? 1 : 0, pd ? cca->herm : 0.0, pd ? ccg->herm : 0.0, k,
                               c->nz_on ? c->nz_mask : (uint32_t)ZB_ALL};
      auto it = c->graphs.find(key);
      if (it == c->graphs.end()) {
        hipGraph_t graph;
        HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = term();
        hipError_t ec = hipStreamEndCapture(c->stream, &graph);
        if (rc) return rc;
        if (ec != hipSuccess) return fail(QD_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        hipGraphExec_t exec;
        HIPC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        HIPC(hipGraphDestroy(graph));
        it = c->graphs.emplace(key, exec).first;
      }
      HIPC(hipGraphLaunch(it->second, c->stream));
    } else if ((rc = term())) {
      return rc;
    }
    taken++;
    double nn[2];
    if ((rc = read_scalars(c, 2, nn))) return rc;
    if (std::sqrt(nn[1]) * BCH_COF[k - 2] / std::sqrt(nn[0]) < R4_1EM5) break;
    prev = next;
  }
  if (nterms) *nterms = taken;
  return QD_OK;
}

int qd_magnus_expand(qd_context *c, qd_op *dg, const qd_op *g, const qd_op *eta) {
  (void)g;
  if (!c || !dg || !eta) return fail(QD_E_ARG, "null argument");
  return qd_copy_arrays(c, eta, dg);
}

int qd_euler_step(qd_context *c, qd_op *g, const qd_op *dg, double *s, double stp) {
  if (!c || !g || !dg || !s) return fail(QD_E_ARG, "null argument");
  int rc = qd_add_arrays(c, g, 1.0, dg, stp, g);
  if (rc) return rc;
  *s += stp;
  return QD_OK;
}

int qd_dgam(qd_context *c, const qd_op *h, qd_op *der, qd_cc *hcc, qd_cc *etacc) {
  if (!c || !h || !der || !hcc || !etacc) return fail(QD_E_ARG, "null argument");
  int rc;
  if ((rc = ensure_scratch(c))) return rc;
  qd_op *eta = c->bch_a;
  hcc->herm = 1.0;
  if ((rc = qd_set_to_zero(c, eta))) return rc;
  eta->herm = -1;
  if ((rc = qd_build_white(c, eta, h))) return rc;
  etacc->herm = -1.0;
  // calc_cc(H, HCC) and calc_cc(eta, ETACC), then [eta, H]
  return full_commutator(c, eta, h, der, hcc, etacc, FILL_H | FILL_E);
}

void qd_flow_params_default(qd_flow_params *p) {
  if (!p) return;
  p->stp = R4_0P7;
  p->conv = R4_1EM9;
  p->max_iter = 100000;
  p->pandya = 0;
  p->verbose = 0;
  p->warm_start = 0;
  p->diis = 0;
  p->stop_rule = 0;
  p->conv_deds = 1e-8;
  p->tda_ml = p->tda_ms = p->tda_cut = 0;
  p->spectrum_path = nullptr;
}

int qd_magnus_flow(qd_context *c, const qd_op *h0, qd_op *hs, qd_op *omega, const qd_flow_params *p,
                   qd_flow_result *res, double *trace, int trace_len) {
  if (!c || !h0 || !hs || !omega || !p || !res) return fail(QD_E_ARG, "null argument");
  int rc;
  qd_op *eta = nullptr, *eta0 = nullptr, *g0 = nullptr, *hs0 = nullptr;
  qd_cc *hcc = nullptr, *occ = nullptr;
  if ((rc = qd_op_create(c, -1, &eta)) || (rc = qd_op_create(c, -1, &eta0)) || (rc = qd_op_create(c, -1, &g0)) ||
      (rc = qd_op_create(c, 1, &hs0)) || (rc = qd_cc_create(c, 1.0, &hcc)) || (rc = qd_cc_create(c, -1.0, &occ))) {
    qd_op_destroy(eta); qd_op_destroy(eta0); qd_op_destroy(g0); qd_op_destroy(hs0);
    qd_cc_destroy(hcc); qd_cc_destroy(occ);
    return rc;
  }
  if (p->pandya) { hcc->zero = false; occ->zero = false; }  // request calc_cc-filled inputs
  omega->herm = -1;
  memset(res, 0, sizeof *res);
  double s = 0.0, stp = p->stp, nrm1 = 0, nrm2 = 0, crit = 0, e0 = 0;
  int iteration = 0, rejected = 0, nt = 0, rows = 0;
  int64_t terms = 0;
  auto done = [&](int r) {
    qd_op_destroy(eta); qd_op_destroy(eta0); qd_op_destroy(g0); qd_op_destroy(hs0);
    qd_cc_destroy(hcc); qd_cc_destroy(occ);
    return r;
  };
  // HS = H(0), or BCH(Omega, H(0)) for a warm start
  if (p->warm_start) {
    if ((rc = qd_bch_expand(c, hs, omega, h0, hcc, occ, &nt))) return done(rc);
    terms += nt;
  } else {
    if ((rc = qd_set_to_zero(c, omega))) return done(rc);
    if ((rc = qd_copy_arrays(c, h0, hs))) return done(rc);
  }
  // nrm1 is the step-rejection scale (never updated, src/magnus_IMSRG.f90:
  // 205-231).  A warm start takes it from H(0), the cold start's scale: the
  // warm H(s) is nearly decoupled, and a tiny ||eta|| there would turn every
  // later step into a rejection and a halving of the step.
  if ((rc = qd_build_white(c, eta, p->warm_start ? h0 : hs))) return done(rc);
  if ((rc = qd_mat_2_norm(c, eta, &nrm1))) return done(rc);
  if (p->warm_start && (rc = qd_build_white(c, eta, hs))) return done(rc);
  // stopping quantity: |MBPT2(H(s))| (the reference) or |dE/ds|, the
  // zero-body part of [eta(s), H(s)] (commutator_110 + commutator_220), the
  // energy derivative of the IM-SRG flow equation at s; both are evaluated
  // every accepted step, crit is the one the rule names
  double m2, deds = 0.0;
  const bool rule_deds = p->stop_rule == 1;
  const double conv = rule_deds ? p->conv_deds : p->conv;
  auto eval_crit = [&](const qd_op *h) -> int {
    int r;
    if ((r = qd_mbpt2(c, h, &m2))) return r;
    Ops zo{eta->d, h->d, eta->herm, h->herm};
    launch_zerobody(c->stream, c->dt, zo, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, c->partials, c->scal, 3);
    HIPC(hipGetLastError());
    if ((r = read_scalars(c, 1, &deds))) return r;
    crit = rule_deds ? std::fabs(deds) : std::fabs(m2);
    return QD_OK;
  };
  // the end of a step in one read-back: ||eta|| (rejection test), E0, and
  // both stopping quantities of the new H(s); crit itself is updated by
  // set_crit once the step is accepted
  double m2n = 0.0, dedsn = 0.0;
  auto step_scalars = [&]() -> int {
    Timer tm(c, TC_OTHER);
    launch_norm2(c->stream, eta->d, c->ch.neq, c->partials, c->scal);
    launch_mbpt2(c->stream, c->dt, hs->d, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, c->partials, c->scal + 1);
    Ops zo{eta->d, hs->d, eta->herm, hs->herm};
    launch_zerobody(c->stream, c->dt, zo, c->pphh_segs, c->pphh_tiles, c->npphh_tiles, c->partials, c->scal + 2, 3);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(c->host_scal, c->scal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipMemcpyAsync(c->host_scal + 3, hs->d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    nrm2 = std::sqrt(c->host_scal[0]);
    m2n = c->host_scal[1];
    dedsn = c->host_scal[2];
    e0 = c->host_scal[3];
    if (!std::isfinite(e0) || !std::isfinite(nrm2))
      return fail(QD_E_NUMERIC, "non-finite E0 or ||eta|| at s=%g", s);
    return QD_OK;
  };
  auto set_crit = [&]() {
    m2 = m2n;
    deds = dedsn;
    crit = rule_deds ? std::fabs(deds) : std::fabs(m2);
  };
  if (p->stop_rule != 0 && p->stop_rule != 1) return done(fail(QD_E_ARG, "stop_rule must be 0 or 1"));
  if ((rc = eval_crit(hs))) return done(rc);
  // write_spec (src/magnus_IMSRG.f90:165-189) into spectrum_path
  FILE *spec = nullptr;
  std::vector<double> tda_eig;
  const int nlev = c->ch.n * c->ch.mp + 2;
  if (p->spectrum_path) {
    spec = fopen(p->spectrum_path, "w");
    if (!spec) return done(fail(QD_E_ARG, "cannot create %s", p->spectrum_path));
    tda_eig.resize((size_t)std::max(nlev, 1));
  }
  auto write_spec = [&](double e0now) -> int {
    if (!spec) return QD_OK;
    int d = 0, r;
    if ((r = qd_tda_spectrum(c, hs, p->tda_ml, p->tda_ms, p->tda_cut, tda_eig.data(), (int)tda_eig.size(), &d)))
      return r;
    fprintf(spec, "%12.7f%12.7f", s, e0now);
    for (int k = 2; k < nlev; k++) fprintf(spec, "%12.7f", k - 2 < d ? tda_eig[k - 2] + e0now : 0.0);
    fprintf(spec, "\n");
    return QD_OK;
  };
  auto done_spec = [&](int r) {
    if (spec) fclose(spec);
    spec = nullptr;
    return r;
  };
  {
    HIPC(hipMemcpyAsync(c->host_scal, hs->d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    if ((rc = write_spec(c->host_scal[0]))) return done(done_spec(rc));
  }
  int steps = 0;
  // Pulay (DIIS) acceleration (no counterpart in the reference): once the
  // flow is close to decoupling (crit <= DIIS_START), keep the last m
  // iterates (Omega_i, eta_i) of the map Omega -> Omega + stp eta(Omega) and
  // step from the combination of iterates whose residuals eta_i have the
  // least norm (weights summing to 1).  Further out, and whenever a step of
  // it lets ||eta|| grow past twice the smallest value seen since it took
  // over (the history is then dropped), the reference's step with its
  // rejection rule is taken.  Same stopping rule.
  const double DIIS_START = 1e-4;
  const int m = std::min(std::max(p->diis, 0), 6);
  std::vector<qd_op *> HO(m, nullptr), HR(m, nullptr);
  auto free_hist = [&]() {
    for (auto *o : HO) qd_op_destroy(o);
    for (auto *o : HR) qd_op_destroy(o);
  };
  for (int i = 0; i < m; i++)
    if ((rc = qd_op_create(c, -1, &HO[i])) || (rc = qd_op_create(c, -1, &HR[i]))) {
      free_hist();
      return done(done_spec(rc));
    }
  auto done2 = [&](int r) { free_hist(); return done(done_spec(r)); };
  std::vector<double> B((size_t)std::max(m, 1) * std::max(m, 1), 0.0);
  int count = 0;
  double best = 0.0;
  auto log_step = [&]() {
    if (p->verbose) {
      FILE *vf = p->verbose == 2 ? stderr : stdout;
      fprintf(vf, "%5d%18.9f%18.9f%18.9f%18.9f%18.9f\n", iteration, s, e0, nrm1, nrm2, crit);
      if (p->verbose == 2) fflush(vf);
    }
    if (trace && rows < trace_len) {
      double *r = trace + 6 * (size_t)rows++;
      r[0] = iteration; r[1] = s; r[2] = e0; r[3] = nrm1; r[4] = nrm2; r[5] = crit;
    }
  };
  while (crit > conv && steps < p->max_iter) {
    steps++;
    if (m > 0 && crit <= DIIS_START) {
      const int slot = count % m;
      if ((rc = qd_copy_arrays(c, omega, HO[slot])) || (rc = qd_copy_arrays(c, eta, HR[slot]))) return done2(rc);
      count++;
      const int n = std::min(count, m);
      for (int i = 0; i < n; i++) launch_dot(c->stream, HR[slot]->d, HR[i]->d, c->ch.neq, c->partials, c->scal + i);
      HIPC(hipGetLastError());
      double d[8];
      if ((rc = read_scalars(c, n, d))) return done2(rc);
      for (int i = 0; i < n; i++) B[(size_t)slot * m + i] = B[(size_t)i * m + slot] = d[i];
      std::vector<double> Bn((size_t)n * n), w(n);
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Bn[(size_t)i * n + j] = B[(size_t)i * m + j];
      LinComb lc{};
      if (n >= 2 && diis_weights(n, Bn.data(), w.data())) {
        for (int i = 0; i < n; i++) {
          lc.c[lc.k] = w[i]; lc.x[lc.k++] = HO[i]->d;
          lc.c[lc.k] = w[i] * stp; lc.x[lc.k++] = HR[i]->d;
        }
      } else {
        lc.c[0] = 1.0; lc.x[0] = omega->d; lc.c[1] = stp; lc.x[1] = eta->d; lc.k = 2;
      }
      launch_lincomb(c->stream, c->ch.neq, lc, omega->d);
      HIPC(hipGetLastError());
      s += stp;
      if ((rc = qd_bch_expand(c, hs, omega, h0, hcc, occ, &nt))) return done2(rc);
      terms += nt;
      if ((rc = qd_build_white(c, eta, hs))) return done2(rc);
      if ((rc = step_scalars())) return done2(rc);
      log_step();
      if (count == 1 || nrm2 < best) best = nrm2;
      if (nrm2 > 2.0 * best) count = 0;  // drop the history
      if ((rc = write_spec(e0)) || (rc = write_spec(e0))) return done2(rc);
      set_crit();
      iteration++;
      continue;
    }
    count = 0;
    if ((rc = qd_copy_arrays(c, omega, g0))) return done2(rc);
    // MAGNUS_EXPAND is first order (dOmega = eta, src/magnus_IMSRG.f90:345),
    // so the Euler step takes eta itself
    if ((rc = qd_euler_step(c, omega, eta, &s, stp))) return done2(rc);
    if ((rc = qd_copy_arrays(c, hs, hs0))) return done2(rc);
    if ((rc = qd_bch_expand(c, hs, omega, h0, hcc, occ, &nt))) return done2(rc);
    terms += nt;
    std::swap(eta, eta0);  // eta0 keeps this step's generator for a rejection
    if ((rc = qd_build_white(c, eta, hs))) return done2(rc);
    if ((rc = step_scalars())) return done2(rc);
    log_step();
    if (nrm1 * 1.5 < nrm2) {
      if (p->verbose) fprintf(p->verbose == 2 ? stderr : stdout, " faillll\n");
      s -= stp;
      if ((rc = qd_copy_arrays(c, g0, omega)) || (rc = qd_copy_arrays(c, hs0, hs))) return done2(rc);
      std::swap(eta, eta0);
      stp = stp / 2.0 + R4_1EM3;
      rejected++;
      continue;
    }
    if ((rc = write_spec(e0)) || (rc = write_spec(e0))) return done2(rc);
    set_crit();
    iteration++;
  }
  free_hist();
  done_spec(QD_OK);
  HIPC(hipMemcpyAsync(c->host_scal, hs->d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  res->s = s;
  res->e0 = c->host_scal[0];
  res->crit = crit;
  res->iterations = iteration;
  res->rejected = rejected;
  res->bch_terms = terms;
  res->converged = crit <= conv;
  res->trace_rows = rows;
  res->mbpt2 = m2;
  res->deds = deds;
  if (p->verbose) fflush(stdout);
  return done(QD_OK);
}

// ---- the ODE path: ode(dGam, ...) and the gs_IMSRG decoupling loop ----------------
struct qd_ode {
  qd_context *ctx = nullptr;
  AdamsDevice a;
  qd_cc *hcc = nullptr, *etacc = nullptr;
};

int qd_ode_create(qd_context *c, qd_ode **out) {
  if (!c || !out) return fail(QD_E_ARG, "null argument");
  *out = nullptr;
  qd_ode *o = new qd_ode();
  o->ctx = c;
  int rc;
  if ((rc = o->a.init(c->stream, c->ch.neq)) || (rc = qd_cc_create(c, 1.0, &o->hcc)) ||
      (rc = qd_cc_create(c, -1.0, &o->etacc))) {
    qd_ode_destroy(o);
    return rc;
  }
  *out = o;
  return QD_OK;
}

int qd_ode_destroy(qd_ode *o) {
  if (!o) return QD_OK;
  o->a.release();
  qd_cc_destroy(o->hcc);
  qd_cc_destroy(o->etacc);
  delete o;
  return QD_OK;
}

// Fortran edit descriptor E14.6: sign, "0.", six digits, "E", signed two-digit exponent
static const char *fortran_e(double x, char *buf) {
  if (x == 0.0 || !std::isfinite(x)) {
    snprintf(buf, 32, "%14s", x == 0.0 ? "0.000000E+00" : "NaN");
    return buf;
  }
  int e = (int)std::floor(std::log10(std::fabs(x))) + 1;
  double m = std::fabs(x) / std::pow(10.0, e);
  long long r = std::llround(m * 1e6);
  if (r >= 1000000) { r /= 10; e += 1; }
  if (r < 100000) { r *= 10; e -= 1; }
  char body[32];
  snprintf(body, sizeof body, "%s0.%06lldE%c%02d", x < 0 ? "-" : "", r, e < 0 ? '-' : '+', std::abs(e));
  snprintf(buf, 32, "%14s", body);
  return buf;
}

static OdeRhs dgam_rhs(qd_ode *o) {
  return [o](double, const double *y, double *yp) -> int {
    qd_op H{o->ctx, const_cast<double *>(y), 1}, D{o->ctx, yp, 1};
    return qd_dgam(o->ctx, &H, &D, o->hcc, o->etacc);
  };
}

int qd_ode_integrate(qd_ode *o, qd_op *y, double *t, double tout, double *relerr, double *abserr, int *iflag) {
  if (!o || !y || !t || !relerr || !abserr || !iflag) return fail(QD_E_ARG, "null argument");
  if (y->ctx != o->ctx) return fail(QD_E_ARG, "operator belongs to another context");
  return o->a.ode(dgam_rhs(o), y->d, t, tout, relerr, abserr, iflag);
}

int qd_ode_stats(const qd_ode *o, int *nsteps, int64_t *nfe, int *order, double *h) {
  if (!o) return fail(QD_E_ARG, "null argument");
  if (nsteps) *nsteps = o->a.nsteps;
  if (nfe) *nfe = o->a.nfe;
  if (order) *order = o->a.kold;
  if (h) *h = o->a.h;
  return QD_OK;
}

void qd_gs_params_default(qd_gs_params *p) {
  if (!p) return;
  p->relerr = R4_1EM8;
  p->abserr = R4_1EM6;
  p->stp = R4_0P1;
  p->conv = R4_1EM7;
  p->max_outputs = 100000;
  p->verbose = 0;
  p->maxnum = 500;
}

int qd_gs_flow(qd_context *c, const qd_op *h0, qd_op *hs, const qd_gs_params *p, qd_gs_result *res, double *trace,
               int trace_len) {
  if (!c || !h0 || !hs || !p || !res) return fail(QD_E_ARG, "null argument");
  int rc;
  qd_ode *o;
  if (p->maxnum < 1) return fail(QD_E_ARG, "maxnum must be positive");
  if ((rc = qd_ode_create(c, &o))) return rc;
  o->a.maxnum = p->maxnum;
  memset(res, 0, sizeof *res);
  if ((rc = qd_copy_arrays(c, h0, hs))) { qd_ode_destroy(o); return rc; }
  hs->herm = 1;
  double s = 0.0, rel = p->relerr, ab = p->abserr, crit = 1.0, e_new = 0.0;
  int flag = 1, pr = 0;
  HIPC(hipMemcpyAsync(c->host_scal, hs->d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  e_new = c->host_scal[0];
  while (crit > p->conv && pr < p->max_outputs) {
    const double e_old = e_new;
    if ((rc = o->a.ode(dgam_rhs(o), hs->d, &s, s + p->stp, &rel, &ab, &flag))) { qd_ode_destroy(o); return rc; }
    HIPC(hipMemcpyAsync(c->host_scal, hs->d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    e_new = c->host_scal[0];
    if (!std::isfinite(e_new)) { qd_ode_destroy(o); return fail(QD_E_NUMERIC, "non-finite E0 at s=%g", s); }
    crit = std::fabs((e_old - e_new) / e_old);
    pr++;
    if (p->verbose) {  // the reference's per-output line '(I5,3(e14.6))' (src/gs_IMSRG.f90:117)
      FILE *vf = p->verbose == 2 ? stderr : stdout;
      char b1[32], b2[32], b3[32];
      fprintf(vf, "%5d%s%s%s\n", pr, fortran_e(s, b1), fortran_e(e_new, b2), fortran_e(crit, b3));
      fflush(vf);
    }
    if (trace && pr <= trace_len) {
      double *r = trace + 4 * (size_t)(pr - 1);
      r[0] = pr; r[1] = s; r[2] = e_new; r[3] = crit;
    }
    // As the reference's loop, the next call continues from a flag of 3
    // (tolerances raised), 4 (500 steps) or 5 (stiff) towards the next output
    // point; only 6 (bad input) ends the flow, unconverged.
    if (flag == 6) break;
  }
  res->s = s;
  res->e0 = e_new;
  res->crit = crit;
  res->outputs = pr;
  res->nsteps = o->a.nsteps;
  res->nfe = o->a.nfe;
  res->iflag = flag;
  res->converged = crit <= p->conv;
  qd_ode_destroy(o);
  return QD_OK;
}

// ---- timing ---------------------------------------------------------------------
int qd_timing_enable(qd_context *c, int on) {
  if (!c) return fail(QD_E_ARG, "null context");
  HIPC(hipStreamSynchronize(c->stream));
  for (auto &e : c->events) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
  c->events.clear();
  for (int i = 0; i < NCLASS; i++) { c->t_ms[i] = 0; c->t_n[i] = 0; }
  c->timing = on != 0;
  return QD_OK;
}
int qd_timing_read(qd_context *c, double *ms, int64_t *nl, int nclass) {
  if (!c) return fail(QD_E_ARG, "null context");
  HIPC(hipStreamSynchronize(c->stream));
  for (auto &e : c->events) {
    float t = 0;
    HIPC(hipEventElapsedTime(&t, e.a, e.b));
    c->t_ms[e.cls] += t;
    c->t_n[e.cls] += 1;
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  c->events.clear();
  for (int i = 0; i < nclass && i < NCLASS; i++) {
    if (ms) ms[i] = c->t_ms[i];
    if (nl) nl[i] = c->t_n[i];
  }
  return QD_OK;
}

// captured BCH terms bake in kernels, tile lists and shapes: any change to
// those drops them (they are captured again on next use)
static int drop_graphs(qd_context *c) {
  HIPC(hipStreamSynchronize(c->stream));
  for (auto &kv : c->graphs) hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  return QD_OK;
}

int qd_set_graphs(qd_context *c, int on) {
  if (!c) return fail(QD_E_ARG, "null context");
  int rc = drop_graphs(c);
  c->use_graphs = on != 0;
  return rc;
}

int64_t qd_graph_count(const qd_context *c) {
  if (!c) return fail(QD_E_ARG, "null context");
  return (int64_t)c->graphs.size();
}

int qd_set_launch_mode(qd_context *c, int mode) {
  if (!c || (mode != 0 && mode != 1)) return fail(QD_E_ARG, "launch mode must be 0 (grouped) or 1 (per channel)");
  c->launch_mode = mode;
  drop_graphs(c);
  return QD_OK;
}

int qd_set_precision(qd_context *c, int prec) {
  if (!c || prec < QD_PREC_FP64 || prec > QD_PREC_BF16X4)
    return fail(QD_E_ARG, "precision must be QD_PREC_FP64, QD_PREC_FP32 or QD_PREC_BF16X2..4");
  int rc;
  if ((rc = drop_graphs(c))) return rc;
  if (prec >= GEMM_BF16X2 && !c->cs) {  // slice planes of the fused BF16 Pandya pass
    void *p = nullptr;
    if ((rc = dev_alloc(c, sizeof(unsigned short) * 2 * 4 * std::max<int64_t>(c->cs_len, 8), &p))) return rc;
    c->cs = (unsigned short *)p;
  }
  c->gemm_prec = prec;  // QD_PREC_* equal the GemmPrec values
  return QD_OK;
}

int qd_get_precision(qd_context *c, int *prec) {
  if (!c || !prec) return fail(QD_E_ARG, "null argument");
  *prec = c->gemm_prec;
  return QD_OK;
}

namespace {
int size_class(int x) { return x <= 16 ? 0 : x <= 32 ? 1 : x <= 64 ? 2 : x <= 128 ? 3 : x <= 256 ? 4 : 5; }

// problems [0, np) of a list grouped by size bucket (M, N, K classes),
// largest total work first
std::vector<std::vector<int>> size_buckets(const GemmList &g, int np) {
  std::map<int, std::vector<int>> m;
  for (int pi = 0; pi < np; pi++) {
    const GemmProb &p = g.hprobs[pi];
    m[size_class(p.M) * 100 + size_class(p.N) * 10 + size_class(ktot(g, p))].push_back(pi);
  }
  std::vector<std::pair<double, std::vector<int>>> v;
  for (auto &kv : m) {
    double w = 0;
    for (int pi : kv.second) w += (double)g.hprobs[pi].M * g.hprobs[pi].N * ktot(g, g.hprobs[pi]);
    v.push_back({w, kv.second});
  }
  std::sort(v.begin(), v.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
  std::vector<std::vector<int>> out;
  for (auto &x : v) out.push_back(x.second);
  return out;
}
}  // namespace

// Coordinate descent over size buckets: for each bucket (largest work first)
// every tile shape is tried with all other buckets as they stand, and the one
// giving the shortest device time of the GEMM launches of a commutator with
// the Pandya term (both grouped lists, as launched at the current precision
// and launch form) is kept.  Timing the whole launch rather than a bucket
// alone prices a shape by what it adds to the grouped launch.  Where the
// Pandya products run in the fused kernel, their problems are neither timed
// nor tuned.
int qd_autotune(qd_context *c, int reps, double *ms_before, double *ms_after) {
  if (!c) return fail(QD_E_ARG, "null context");
  if (reps < 1) reps = 3;
  int rc;
  if ((rc = drop_graphs(c))) return rc;
  qd_op *a, *b, *o;
  qd_cc *x, *y;
  if ((rc = qd_op_create(c, -1, &a)) || (rc = qd_op_create(c, 1, &b)) || (rc = qd_op_create(c, 1, &o))) return rc;
  if ((rc = qd_cc_create(c, -1.0, &x)) || (rc = qd_cc_create(c, 1.0, &y))) return rc;
  x->zero = y->zero = false;
  const Slots s1 = make_slots(c, a, b, o, x, y);
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  const bool fused = c->launch_mode == 0 && pandya_fused_prec(c->gemm_prec);
  // the fastest of reps timed runs of both lists (after one untimed run):
  // the lists are deterministic work, so the minimum is the least noisy
  auto total = [&](double *ms) -> int {
    int r2;
    if ((r2 = run_gemm_list(c, c->g_all, !fused, s1)) || (r2 = run_gemm_list(c, c->g_acc, false, s1))) return r2;
    double best = 1e30;
    for (int r = 0; r < reps; r++) {
      HIPC(hipEventRecord(e0, c->stream));
      if ((r2 = run_gemm_list(c, c->g_all, !fused, s1))) return r2;
      if ((r2 = run_gemm_list(c, c->g_acc, false, s1))) return r2;
      HIPC(hipEventRecord(e1, c->stream));
      HIPC(hipEventSynchronize(e1));
      float f = 0;
      HIPC(hipEventElapsedTime(&f, e0, e1));
      best = std::min(best, (double)f);
    }
    *ms = best;
    return QD_OK;
  };
  double before = 0, best_total = 0;
  if ((rc = total(&before))) return rc;
  best_total = before;
  for (GemmList *g : {&c->g_all, &c->g_acc}) {
    const int np = (g == &c->g_all && fused) ? g->npre : (int)g->hprobs.size();
    for (const auto &bk : size_buckets(*g, np)) {
      const int cur = g->hprobs[bk[0]].shape;
      int best = cur;
      for (int sh = 0; sh < NSHAPE; sh++) {
        if (sh == cur) continue;
        for (int pi : bk) g->hprobs[pi].shape = sh;
        if ((rc = upload_tiles(c, *g))) return rc;
        double ms;
        if ((rc = total(&ms))) return rc;
        if (ms < best_total) { best_total = ms; best = sh; }
      }
      for (int pi : bk) g->hprobs[pi].shape = best;
      if ((rc = upload_tiles(c, *g))) return rc;
    }
  }
  double after = 0;
  if ((rc = total(&after))) return rc;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ms_before) *ms_before = before;
  if (ms_after) *ms_after = after;
  qd_op_destroy(a); qd_op_destroy(b); qd_op_destroy(o); qd_cc_destroy(x); qd_cc_destroy(y);
  return QD_OK;
}

int qd_gemm_shapes(const qd_context *c, int *shapes, int64_t len) {
  if (!c) return fail(QD_E_ARG, "null context");
  const int64_t n = (int64_t)(c->g_all.hprobs.size() + c->g_acc.hprobs.size());
  if (shapes) {
    if (len < n) return fail(QD_E_ARG, "needs %lld entries", (long long)n);
    int64_t k = 0;
    for (auto &p : c->g_all.hprobs) shapes[k++] = p.shape;
    for (auto &p : c->g_acc.hprobs) shapes[k++] = p.shape;
  }
  return (int)n;
}

int qd_set_gemm_shapes(qd_context *c, const int *shapes, int64_t len) {
  if (!c || !shapes) return fail(QD_E_ARG, "null argument");
  const int64_t n = (int64_t)(c->g_all.hprobs.size() + c->g_acc.hprobs.size());
  if (len != n) return fail(QD_E_ARG, "%lld shapes given, the lists hold %lld problems", (long long)len, (long long)n);
  for (int64_t k = 0; k < n; k++)
    if (shapes[k] < 0 || shapes[k] >= NSHAPE) return fail(QD_E_ARG, "shape %d of problem %lld", shapes[k], (long long)k);
  int rc;
  if ((rc = drop_graphs(c))) return rc;
  int64_t k = 0;
  for (auto &p : c->g_all.hprobs) p.shape = shapes[k++];
  for (auto &p : c->g_acc.hprobs) p.shape = shapes[k++];
  if ((rc = upload_tiles(c, c->g_all)) || (rc = upload_tiles(c, c->g_acc))) return rc;
  return QD_OK;
}

// Matrix-core flops one commutator's GEMM launches execute at the current
// shapes and precision: every tile covers TM x TN outputs whatever the
// problem's M, N (edge tiles zero-filled); FP64 tiles walk each segment's K
// in MFMA k-steps of 4 (the zero fill past K skipped); FP32 tiles run K
// stages of 32 in full, BF16 tiles S(S+1)/2 slice products per stage of 32.
// The fused Pandya kernels count likewise over 64 x 64 tile pairs (one tile
// on the diagonal).
int qd_gemm_mfma_flops(const qd_context *c, double *products, double *pandya, double *acc) {
  if (!c) return fail(QD_E_ARG, "null context");
  const int prec = c->gemm_prec;
  const double slice = prec >= GEMM_BF16X2 ? prec * (prec + 1) / 2.0 : 1.0;
  auto list = [&](const GemmList &g, int p0, int p1) {
    double f = 0;
    for (int pi = p0; pi < p1; pi++) {
      const GemmProb &p = g.hprobs[pi];
      const int TM = SHAPE_TM[p.shape], TN = SHAPE_TN[p.shape];
      const double tiles = (double)((p.M + TM - 1) / TM) * ((p.N + TN - 1) / TN);
      for (int s = 0; s < p.nseg; s++) {
        const int K = g.hsegs[p.seg0 + s].K;
        const double kk = prec == GEMM_FP64 ? 4.0 * ((K + 3) / 4) : 32.0 * ((K + 31) / 32);
        f += tiles * 2.0 * TM * TN * kk * slice;
      }
    }
    return f;
  };
  const bool fused = c->launch_mode == 0 && pandya_fused_prec(prec);
  const int npre = c->g_all.npre, nall = (int)c->g_all.hprobs.size();
  if (products) *products = list(c->g_all, 0, npre);
  if (acc) *acc = list(c->g_acc, 0, (int)c->g_acc.hprobs.size());
  if (pandya) {
    if (fused) {
      double f = 0;
      for (int q = 0; q < c->ch.ncc; q++) {
        const int r = c->ch.cc_r[q], K = c->ch.cc_nb[q];
        if (r * K == 0) continue;
        const double nt = (r + 63) / 64, pairs = nt * (nt + 1) / 2;
        const double kk = prec == GEMM_FP64 ? 4.0 * ((K + 3) / 4) : 32.0 * ((K + 31) / 32) * slice;
        f += (2 * pairs - nt) * 2.0 * 64 * 64 * kk;  // two tiles per pair, one on the diagonal
      }
      *pandya = f;
    } else {
      *pandya = list(c->g_all, npre, nall);
    }
  }
  return QD_OK;
}

int qd_gemm_flops(const qd_context *c, double *ladder, double *pandya, double *acc) {
  if (!c) return fail(QD_E_ARG, "null context");
  if (ladder) *ladder = c->g_all.flops_pre;
  if (pandya) *pandya = c->g_all.flops - c->g_all.flops_pre;
  if (acc) *acc = c->g_acc.flops;
  return QD_OK;
}

}  // extern "C"
