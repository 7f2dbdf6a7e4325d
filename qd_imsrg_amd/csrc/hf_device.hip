// Hartree-Fock on the device (see hf_device.h): the Fock build of the SCF
// iteration (src/HFQDmod.f90:196-221 with construct_V, :260-290) and the
// transformation of the antisymmetrised two-body elements to the HF basis
// (new_VMAT, src/ME_general.f90:1362-1492), from the Coulomb table uploaded
// once.  The host keeps the SCF control: density, per-channel diagonalisation
// and the convergence test.
#include <algorithm>
#include <cmath>

#include "hf_device.h"
#include "qd_kernels.h"

namespace qd {

struct HfDevView {
  int M, N;
  double shw;
  int fb;
  const double *tab;
  const int64_t *start;
  const int *pos, *m, *sp, *spin;
};

struct HfDevice {
  qd_context *ctx = nullptr;
  hipStream_t st = nullptr;
  HfDevView v{};
  std::vector<void *> mem;
  double *rho = nullptr, *Tm = nullptr, *F = nullptr;
  int2 *fpairs = nullptr;
  int nfp = 0;
  ~HfDevice() {
    for (void *p : mem) (void)hipFree(p);
  }
};

namespace {

#define HIPH(expr)                                                                            \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return set_error(QD_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <class T>
int upload(HfDevice *d, const T *h, size_t n, T **out) {
  void *p = nullptr;
  HIPH(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  d->mem.push_back(p);
  if (n) HIPH(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, d->st));
  *out = (T *)p;
  return QD_OK;
}

__device__ __forceinline__ double cget(const HfDevView &h, int a, int b, int c, int d) {
  if (h.m[a] + h.m[b] - h.m[c] != h.m[d]) return 0.0;
  return h.tab[h.start[((int64_t)a * h.N + b) * h.N + c] + h.pos[d]];
}
// <pq|V|rs>_A of the spin-orbitals (direct minus exchange, spins conserved)
__device__ __forceinline__ double asym(const HfDevView &h, int p, int q, int r, int s) {
  double v = 0.0;
  const int a = h.sp[p], b = h.sp[q], c = h.sp[r], d = h.sp[s];
  if (h.spin[p] == h.spin[r] && h.spin[q] == h.spin[s]) v += cget(h, a, b, c, d);
  if (h.spin[p] == h.spin[s] && h.spin[q] == h.spin[r]) {
    double x = cget(h, a, b, d, c);
    if (h.fb && fabs(x) < 1e-10) x = cget(h, b, a, c, d);
    v -= x;
  }
  return h.shw * v;
}

// F[i][k] = Tm[i] d_ik + sum over channels, j and l of rho[l][j] <ji|V|lk>_A,
// one thread per same-channel (i, k), channels and states in ascending order
__global__ void __launch_bounds__(256) k_hf_fock(HfDevView h, DevTab t, const int2 *__restrict__ fp, int nfp,
                                                 const double *__restrict__ rho, const double *__restrict__ Tm,
                                                 double *__restrict__ F) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nfp) return;
  const int i = fp[e].x, k = fp[e].y, M = h.M;
  double v = 0.0;
  for (int c = 0; c < t.nch; c++)
    for (int x = t.ch_start[c]; x < t.ch_start[c + 1]; x++) {
      const int j = t.ch_states[x];
      for (int y = t.ch_start[c]; y < t.ch_start[c + 1]; y++) {
        const int l = t.ch_states[y];
        const double r = rho[(int64_t)l * M + j];
        if (r != 0.0) v += r * asym(h, j, i, l, k);
      }
    }
  F[(int64_t)i * M + k] = (i == k ? Tm[i] : 0.0) + v;
}

// TV[x][y] = sum_z T[x][z] <z|V|y>_A for one block (pairs in layout order
// hh | ph | pp from hh_start[q]); blocks found by search over tv_off
__global__ void __launch_bounds__(256) k_hf_tv(HfDevView h, DevTab t, const int64_t *__restrict__ tv_off, int nblock,
                                               const int *__restrict__ rs, const int *__restrict__ col,
                                               const double *__restrict__ val, double *__restrict__ TV) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= tv_off[nblock]) return;
  int lo = 0, hi = nblock - 1;
  while (lo < hi) {  // last block with tv_off[q] <= e
    const int mid = (lo + hi + 1) >> 1;
    if (tv_off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  const int q = lo, nt = t.nh[q] + t.nb[q] + t.np[q];
  const int64_t loc = e - tv_off[q];
  const int x = (int)(loc / nt), y = (int)(loc % nt), base = t.hh_start[q];
  const uint32_t py = t.pairs[base + y];
  double s = 0.0;
  for (int k = rs[base + x]; k < rs[base + x + 1]; k++) {
    const uint32_t pz = t.pairs[base + col[k]];
    s += val[k] * asym(h, lo16(pz), hi16(pz), lo16(py), hi16(py));
  }
  TV[e] = s;
}

// H = TV T^T into the stored sub-blocks of the operator
__global__ void __launch_bounds__(256) k_hf_h(DevTab t, const int64_t *__restrict__ tv_off, int nblock,
                                              const int *__restrict__ rs, const int *__restrict__ col,
                                              const double *__restrict__ val, const double *__restrict__ TV,
                                              double *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= tv_off[nblock]) return;
  int lo = 0, hi = nblock - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tv_off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  const int q = lo, nh = t.nh[q], nb = t.nb[q], np = t.np[q], nt = nh + nb + np;
  const int64_t loc = e - tv_off[q];
  const int x = (int)(loc / nt), y = (int)(loc % nt);
  const int cx = x < nh ? 0 : (x < nh + nb ? 1 : 2), cy = y < nh ? 0 : (y < nh + nb ? 1 : 2);
  const int rx = x - (cx == 0 ? 0 : cx == 1 ? nh : nh + nb), ry = y - (cy == 0 ? 0 : cy == 1 ? nh : nh + nb);
  int sub, rows;
  if (cx == 0 && cy == 0) { sub = HHHH; rows = nh; }
  else if (cx == 2 && cy == 2) { sub = PPPP; rows = np; }
  else if (cx == 2 && cy == 0) { sub = PPHH; rows = np; }
  else if (cx == 1 && cy == 1) { sub = PHPH; rows = nb; }
  else if (cx == 2 && cy == 1) { sub = PPPH; rows = np; }
  else if (cx == 1 && cy == 0) { sub = PHHH; rows = nb; }
  else return;  // the transposed sub-blocks are not stored
  const int base = t.hh_start[q];
  const double *tv = TV + tv_off[q] + (int64_t)x * nt;
  double s = 0.0;
  for (int k = rs[base + y]; k < rs[base + y + 1]; k++) s += tv[col[k]] * val[k];
  out[t.off[6 * q + sub] + rx + (int64_t)ry * rows] = s;
}

// E0 = sum of the hole energies - sum over blocks of the Vhhhh diagonal, in
// block order (HF_E2, src/HFQDmod.f90:292-323)
__global__ void k_hf_e0(DevTab t, double eh, double *out) {
  double v = 0.0;
  for (int q = 0; q < t.nblock; q++)
    for (int j = 0; j < t.nh[q]; j++) v += out[t.off[6 * q + HHHH] + j + (int64_t)j * t.nh[q]];
  out[0] = eh - v;
}

}  // namespace

int hf_device_create(qd_context *ctx, const CoulombView &S, const std::vector<int> &sp, const std::vector<int> &spin,
                     double shw, HfDevice **out) {
  *out = nullptr;
  HfDevice *d = new HfDevice();
  d->ctx = ctx;
  d->st = context_stream(ctx);
  const Channels &ch = context_channels(ctx);
  d->v.M = ch.M;
  d->v.N = S.N;
  d->v.shw = shw;
  d->v.fb = S.exch_fallback ? 1 : 0;
  std::vector<int2> fp;
  for (int c = 0; c < ch.nch; c++)
    for (int x = ch.ch_start[c]; x < ch.ch_start[c + 1]; x++)
      for (int y = ch.ch_start[c]; y < ch.ch_start[c + 1]; y++) fp.push_back(make_int2(ch.ch_states[x], ch.ch_states[y]));
  d->nfp = (int)fp.size();
  const size_t N3 = (size_t)S.N * S.N * S.N;
  int rc;
  if ((rc = upload(d, S.tab, (size_t)S.tab_size, (double **)&d->v.tab)) ||
      (rc = upload(d, S.start, N3, (int64_t **)&d->v.start)) || (rc = upload(d, S.pos_in_m, (size_t)S.N, (int **)&d->v.pos)) ||
      (rc = upload(d, S.m, (size_t)S.N, (int **)&d->v.m)) || (rc = upload(d, sp.data(), sp.size(), (int **)&d->v.sp)) ||
      (rc = upload(d, spin.data(), spin.size(), (int **)&d->v.spin)) ||
      (rc = upload(d, fp.data(), fp.size(), &d->fpairs))) {
    delete d;
    return rc;
  }
  void *p;
  for (double **q : {&d->rho, &d->Tm, &d->F}) {
    const size_t n = q == &d->Tm ? (size_t)ch.M : (size_t)ch.M * ch.M;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
      delete d;
      return set_error(QD_E_HIP, "hipMalloc failed (HF buffers)");
    }
    d->mem.push_back(p);
    *q = (double *)p;
  }
  *out = d;
  return QD_OK;
}

void hf_device_destroy(HfDevice *d) {
  if (d) {
    (void)hipStreamSynchronize(d->st);
    delete d;
  }
}

int hf_device_fock(HfDevice *d, const std::vector<double> &rho, const std::vector<double> &Tm, std::vector<double> &F) {
  const int M = d->v.M;
  HIPH(hipMemcpyAsync(d->rho, rho.data(), sizeof(double) * (size_t)M * M, hipMemcpyHostToDevice, d->st));
  HIPH(hipMemcpyAsync(d->Tm, Tm.data(), sizeof(double) * M, hipMemcpyHostToDevice, d->st));
  HIPH(hipMemsetAsync(d->F, 0, sizeof(double) * (size_t)M * M, d->st));
  hipLaunchKernelGGL(k_hf_fock, dim3((unsigned)((d->nfp + 255) / 256)), dim3(256), 0, d->st, d->v,
                     context_devtab(d->ctx), d->fpairs, d->nfp, d->rho, d->Tm, d->F);
  HIPH(hipGetLastError());
  F.resize((size_t)M * M);
  HIPH(hipMemcpyAsync(F.data(), d->F, sizeof(double) * (size_t)M * M, hipMemcpyDeviceToHost, d->st));
  HIPH(hipStreamSynchronize(d->st));
  return QD_OK;
}

int hf_device_transform(HfDevice *d, const SparseRows &T, qd_op *out) {
  const Channels &ch = context_channels(d->ctx);
  const DevTab &t = context_devtab(d->ctx);
  std::vector<int64_t> tv_off(ch.nblock + 1, 0);
  for (int q = 0; q < ch.nblock; q++) {
    const int64_t nt = ch.nh[q] + ch.nb[q] + ch.np[q];
    tv_off[q + 1] = tv_off[q] + nt * nt;
  }
  HfDevice tmp;  // per-call buffers, freed on return
  tmp.st = d->st;
  int64_t *dtv = nullptr;
  int *drs = nullptr, *dcol = nullptr;
  double *dval = nullptr, *TV = nullptr;
  int rc;
  if ((rc = upload(&tmp, tv_off.data(), tv_off.size(), &dtv)) ||
      (rc = upload(&tmp, T.row_start.data(), T.row_start.size(), &drs)) ||
      (rc = upload(&tmp, T.col.data(), T.col.size(), &dcol)) || (rc = upload(&tmp, T.val.data(), T.val.size(), &dval)))
    return rc;
  void *p = nullptr;
  HIPH(hipMalloc(&p, sizeof(double) * (size_t)std::max<int64_t>(tv_off[ch.nblock], 1)));
  tmp.mem.push_back(p);
  TV = (double *)p;
  const int64_t n = tv_off[ch.nblock];
  if (n > 0) {
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_hf_tv, dim3(g), dim3(256), 0, d->st, d->v, t, dtv, ch.nblock, drs, dcol, dval, TV);
    hipLaunchKernelGGL(k_hf_h, dim3(g), dim3(256), 0, d->st, t, dtv, ch.nblock, drs, dcol, dval, TV, op_data(out));
    HIPH(hipGetLastError());
  }
  HIPH(hipStreamSynchronize(d->st));
  return QD_OK;
}

int hf_device_e0(qd_context *ctx, double eh, qd_op *out) {
  hipStream_t st = context_stream(ctx);
  hipLaunchKernelGGL(k_hf_e0, dim3(1), dim3(1), 0, st, context_devtab(ctx), eh, op_data(out));
  HIPH(hipGetLastError());
  HIPH(hipStreamSynchronize(st));
  return QD_OK;
}

}  // namespace qd
