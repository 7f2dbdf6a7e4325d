// On-disk format of an operator: export_hamiltonian / import_hamiltonian
// (src/ME_general.f90:207-361).
//
// The reference writes a full_ham with Fortran unformatted sequential I/O,
// one WRITE per scalar, so every value is one record: a 4-byte length (8),
// the REAL(8) value, the 4-byte length again.  Order: E0; fhh(i,j), then
// fph(i,j), then fpp(i,j), each with the row index i in the outer loop; then
// per block Vhhhh, Vpppp, Vpphh, Vphph, Vphhh, Vppph, each row-major the same
// way.  The device and host vectors of this library are in vectorize order
// (column-major, Vppph before Vphhh), so the record order is a permutation of
// the vector computed here.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qd_internal.h"

namespace qd {
namespace {

// flat (vectorize-order) index of every record of the file, in file order
std::vector<int64_t> record_order(const Channels &ch) {
  std::vector<int64_t> ord;
  ord.reserve((size_t)ch.neq);
  const int n = ch.n, mp = ch.mp;
  ord.push_back(0);
  auto mat = [&](int64_t off, int rows, int cols) {  // element (i, j) at off + i + j*rows
    for (int i = 0; i < rows; i++)
      for (int j = 0; j < cols; j++) ord.push_back(off + i + (int64_t)j * rows);
  };
  mat(ch.off_fhh, n, n);
  mat(ch.off_fph, mp, n);
  mat(ch.off_fpp, mp, mp);
  for (int q = 0; q < ch.nblock; q++) {
    const int nh = ch.nh[q], np = ch.np[q], nb = ch.nb[q];
    const int64_t *o = &ch.off[(size_t)q * 6];
    mat(o[HHHH], nh, nh);
    mat(o[PPPP], np, np);
    mat(o[PPHH], np, nh);
    mat(o[PHPH], nb, nb);
    mat(o[PHHH], nb, nh);
    mat(o[PPPH], np, nb);
  }
  return ord;
}

int write_records(const Channels &ch, const double *vec, const char *path) {
  const std::vector<int64_t> ord = record_order(ch);
  if ((int64_t)ord.size() != ch.neq) return set_error(QD_E_LAYOUT, "record order covers %zu of %lld elements", ord.size(), (long long)ch.neq);
  FILE *f = std::fopen(path, "wb");
  if (!f) return set_error(QD_E_ARG, "cannot create %s", path);
  std::vector<unsigned char> buf((size_t)16 * ord.size());
  const int32_t len = 8;
  for (size_t k = 0; k < ord.size(); k++) {
    unsigned char *r = &buf[16 * k];
    std::memcpy(r, &len, 4);
    std::memcpy(r + 4, &vec[ord[k]], 8);
    std::memcpy(r + 12, &len, 4);
  }
  const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  if (std::fclose(f) != 0 || !ok) return set_error(QD_E_ARG, "write failed: %s", path);
  return QD_OK;
}

int read_records(const Channels &ch, double *vec, const char *path) {
  const std::vector<int64_t> ord = record_order(ch);
  FILE *f = std::fopen(path, "rb");
  if (!f) return set_error(QD_E_ARG, "cannot open %s", path);
  std::vector<unsigned char> buf((size_t)16 * ord.size());
  const size_t got = std::fread(buf.data(), 1, buf.size(), f);
  unsigned char extra;
  const bool longer = std::fread(&extra, 1, 1, f) == 1;
  std::fclose(f);
  if (got != buf.size() || longer)
    return set_error(QD_E_ARG, "%s holds %zu bytes, this channel structure needs %zu (%zu records of 8 bytes)", path,
                     got + (longer ? 1 : 0), buf.size(), ord.size());
  for (size_t k = 0; k < ord.size(); k++) {
    int32_t a, b;
    std::memcpy(&a, &buf[16 * k], 4);
    std::memcpy(&b, &buf[16 * k + 12], 4);
    if (a != 8 || b != 8) return set_error(QD_E_ARG, "%s: record %zu is not one REAL(8)", path, k + 1);
    std::memcpy(&vec[ord[k]], &buf[16 * k + 4], 8);
  }
  return QD_OK;
}

int channels_of(int nbody, int nshell, int64_t neq, Channels &ch) {
  std::string why = build_channels(nbody, nshell, ch);
  if (!why.empty()) return set_error(QD_E_LAYOUT, "%s", why.c_str());
  if (neq != ch.neq) return set_error(QD_E_ARG, "neq %lld != %lld", (long long)neq, (long long)ch.neq);
  return QD_OK;
}

}  // namespace
}  // namespace qd

using namespace qd;

extern "C" {

int qd_export_hamiltonian_host(int nbody, int nshell, const double *vec, int64_t neq, const char *path) {
  if (!vec || !path) return set_error(QD_E_ARG, "null argument");
  Channels ch;
  int rc = channels_of(nbody, nshell, neq, ch);
  return rc ? rc : write_records(ch, vec, path);
}

int qd_import_hamiltonian_host(int nbody, int nshell, double *vec, int64_t neq, const char *path) {
  if (!vec || !path) return set_error(QD_E_ARG, "null argument");
  Channels ch;
  int rc = channels_of(nbody, nshell, neq, ch);
  return rc ? rc : read_records(ch, vec, path);
}

int qd_export_hamiltonian(const qd_op *op, const char *path) {
  if (!op || !path) return set_error(QD_E_ARG, "null argument");
  const Channels &ch = context_channels(op_context(const_cast<qd_op *>(op)));
  std::vector<double> v((size_t)ch.neq);
  int rc = qd_op_download(op, v.data(), ch.neq);
  return rc ? rc : write_records(ch, v.data(), path);
}

int qd_import_hamiltonian(qd_op *op, const char *path) {
  if (!op || !path) return set_error(QD_E_ARG, "null argument");
  const Channels &ch = context_channels(op_context(op));
  std::vector<double> v((size_t)ch.neq);
  int rc = read_records(ch, v.data(), path);
  return rc ? rc : qd_op_upload(op, v.data(), ch.neq);
}

}  // extern "C"
