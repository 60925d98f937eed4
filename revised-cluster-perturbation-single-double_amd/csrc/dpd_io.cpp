// dpd_io.cpp -- the reference's on-disk formats, host side (the pyscf2cc data
// path and the Davidson vector files), so that files written by the reference's
// front end load into the device-resident tensors and results can be written
// back in the layout its programs read.
//
//   * interface node file (pyscf2cc/mod_interface.f90): a stream file
//       [metaPos int64][nNode int64][data ...][nNode x (name char[20], pos int64,
//       size int64)] with metaPos and pos 1-based byte positions; names are
//       lower-cased and blank-padded to 20; a node rewritten with no more bytes
//       keeps its place, a larger one is appended (get_node, :186-216).
//   * moio symmetry-packed lists (libsrint/clmoioso0.f, inipckso.f, updmoiso.f,
//       getallso_d.f, getlstso_d.f): a direct-access data file of records of
//       iprcwd = 4096 doubles (iprcln = 32768 bytes), the irrep blocks of the list
//       (right-pair irrep ascending, numl x numr doubles each) starting on a record
//       boundary, and a control file "<name>.c" of Fortran sequential unformatted
//       records (4-byte length markers around each; INTEGER*8 data): IRREPX,
//       popp(8), popq(8), popr(8), pops(8), ioppq(8), ioprs(8), irrepr(8),
//       irrepi(8), moionuml(10,20), moionumr(10,20), rectmp(10,20) -- list 1.
//   * IO_stream vector files (lib_cc/IO.f90:87-154): raw doubles, element idx
//       at byte 8*(idx-1); the eigenvector file of mod_davidson.f90:388-390 and
//       :588-589 holds root r of state block istate at idx (istate+r-2)*(1+dim)+1
//       as [w, R(1..dim)].
// Device pointers are accepted wherever tensor data goes in or out (staged
// through host memory).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dpd_engine.hpp"

using namespace dpd;

namespace {

constexpr int64_t kRecWords = 4096;  // iprcwd
constexpr int kNameLen = 20;         // mod_interface nameLen
constexpr int kMetaBytes = 36;       // name + pos + size

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool is_device(const void* p) {
  hipPointerAttribute_t a;
  if (!p) return false;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

// host copy of n doubles that may live on the device
int to_host(const double* src, int64_t n, std::vector<double>& out) {
  out.resize((size_t)std::max<int64_t>(n, 0));
  if (n <= 0) return DPD_OK;
  if (is_device(src)) {
    DPD_HIP_TRY(hipMemcpy(out.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost));
  } else {
    std::memcpy(out.data(), src, sizeof(double) * n);
  }
  return DPD_OK;
}

int from_host(double* dst, const double* src, int64_t n) {
  if (n <= 0) return DPD_OK;
  if (is_device(dst)) {
    DPD_HIP_TRY(hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyHostToDevice));
  } else {
    std::memcpy(dst, src, sizeof(double) * n);
  }
  return DPD_OK;
}

// ---- interface node file ------------------------------------------------------
struct Node {
  char name[kNameLen];
  int64_t pos, size;
};

std::string full_name(const char* name) {
  std::string s(name ? name : "");
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  s.resize(kNameLen, ' ');
  return s;
}

int read_meta(FILE* f, int64_t& metaPos, std::vector<Node>& nodes) {
  int64_t hdr[2];
  if (std::fseek(f, 0, SEEK_SET) || std::fread(hdr, 8, 2, f) != 2) return fail(DPD_EINVAL, "node file: bad header");
  metaPos = hdr[0];
  const int64_t n = hdr[1];
  if (n < 0 || n > 1000 || metaPos < 17) return fail(DPD_EINVAL, "node file: bad header (%lld nodes)", (long long)n);
  nodes.resize((size_t)n);
  if (n && std::fseek(f, (long)(metaPos - 1), SEEK_SET)) return fail(DPD_EINVAL, "node file: bad meta position");
  for (auto& nd : nodes) {
    unsigned char b[kMetaBytes];
    if (std::fread(b, 1, kMetaBytes, f) != (size_t)kMetaBytes) return fail(DPD_EINVAL, "node file: truncated meta data");
    std::memcpy(nd.name, b, kNameLen);
    std::memcpy(&nd.pos, b + kNameLen, 8);
    std::memcpy(&nd.size, b + kNameLen + 8, 8);
  }
  return DPD_OK;
}

int find(const std::vector<Node>& nodes, const std::string& nm) {
  for (size_t i = 0; i < nodes.size(); i++)
    if (std::string(nodes[i].name, kNameLen) == nm) return (int)i;
  return -1;
}

int elem_bytes(char type) {
  switch (type) {
    case 'i':
    case 'd':
      return 8;
    case 's':
      return 4;
    case 'c':
      return 1;
    default:
      return 0;
  }
}

// ---- Fortran sequential unformatted records (4-byte markers) -------------------
void put_record(std::vector<unsigned char>& out, const void* data, uint32_t bytes) {
  const unsigned char* m = (const unsigned char*)&bytes;
  out.insert(out.end(), m, m + 4);
  out.insert(out.end(), (const unsigned char*)data, (const unsigned char*)data + bytes);
  out.insert(out.end(), m, m + 4);
}

int get_record(FILE* f, void* data, uint32_t bytes) {
  uint32_t a = 0, b = 0;
  if (std::fread(&a, 4, 1, f) != 1 || a != bytes) return fail(DPD_EINVAL, "moio control file: record of %u bytes, expected %u", a, bytes);
  if (std::fread(data, 1, bytes, f) != bytes || std::fread(&b, 4, 1, f) != 1 || b != a)
    return fail(DPD_EINVAL, "moio control file: truncated record");
  return DPD_OK;
}

struct Control {
  int64_t irrepx = 1;
  int64_t pop[4][8];
  int64_t ioppq[8], ioprs[8], irrepr[8], irrepi[8];
  int64_t numl[20][10], numr[20][10], rec[20][10];  // Fortran (10,20): [list][irrep]
};

int read_control(const std::string& path, Control& c) {
  File cf;
  cf.f = std::fopen((path + ".c").c_str(), "rb");
  if (!cf.f) return fail(DPD_EINVAL, "moio: cannot open %s.c", path.c_str());
  int rc;
  if ((rc = get_record(cf.f, &c.irrepx, 8))) return rc;
  for (int k = 0; k < 4; k++)
    if ((rc = get_record(cf.f, c.pop[k], 64))) return rc;
  if ((rc = get_record(cf.f, c.ioppq, 64)) || (rc = get_record(cf.f, c.ioprs, 64)) ||
      (rc = get_record(cf.f, c.irrepr, 64)) || (rc = get_record(cf.f, c.irrepi, 64)) ||
      (rc = get_record(cf.f, c.numl, sizeof c.numl)) || (rc = get_record(cf.f, c.numr, sizeof c.numr)) ||
      (rc = get_record(cf.f, c.rec, sizeof c.rec)))
    return rc;
  return DPD_OK;
}

}  // namespace

extern "C" {

// ---- node file ------------------------------------------------------------------
int dpd_node_init(const char* path) {
  if (!path) return fail(DPD_EINVAL, "null path");
  File f;
  f.f = std::fopen(path, "wb");
  if (!f.f) return fail(DPD_EINVAL, "cannot create %s", path);
  const int64_t hdr[2] = {17, 0};
  if (std::fwrite(hdr, 8, 2, f.f) != 2) return fail(DPD_EINVAL, "write to %s failed", path);
  return DPD_OK;
}

int64_t dpd_node_size(const char* path, const char* name) {
  if (!path || !name || std::strlen(name) > (size_t)kNameLen) return -1;
  File f;
  f.f = std::fopen(path, "rb");
  if (!f.f) return -1;
  int64_t meta;
  std::vector<Node> nodes;
  if (read_meta(f.f, meta, nodes)) return -1;
  const int i = find(nodes, full_name(name));
  return i < 0 ? -1 : nodes[(size_t)i].size;
}

int dpd_node_write(const char* path, const char* name, char type, const void* data, int64_t count) {
  if (!path || !name || (!data && count > 0) || count < 0) return fail(DPD_EINVAL, "dpd_node_write: bad arguments");
  if (std::strlen(name) > (size_t)kNameLen) return fail(DPD_EINVAL, "node name longer than %d: %s", kNameLen, name);
  const int eb = elem_bytes(type);
  if (!eb) return fail(DPD_EINVAL, "node type must be 'i', 'd', 's' or 'c'");
  {
    File t;
    t.f = std::fopen(path, "rb");
    if (!t.f) {
      int rc = dpd_node_init(path);
      if (rc) return rc;
    }
  }
  File f;
  f.f = std::fopen(path, "r+b");
  if (!f.f) return fail(DPD_EINVAL, "cannot open %s", path);
  int64_t metaPos;
  std::vector<Node> nodes;
  int rc = read_meta(f.f, metaPos, nodes);
  if (rc) return rc;
  const std::string nm = full_name(name);
  const int64_t bytes = eb * count;
  int i = find(nodes, nm);
  if (i < 0) {
    if (nodes.size() >= 1000) return fail(DPD_EINVAL, "node file full (1000 nodes)");
    Node nd;
    std::memcpy(nd.name, nm.data(), kNameLen);
    nd.pos = metaPos;
    nd.size = bytes;
    metaPos += bytes;
    nodes.push_back(nd);
    i = (int)nodes.size() - 1;
  } else if (nodes[(size_t)i].size < bytes) {
    nodes[(size_t)i].pos = metaPos;  // larger: written in new space
    nodes[(size_t)i].size = bytes;
    metaPos += bytes;
  } else {
    nodes[(size_t)i].size = bytes;  // written in the old space
  }
  if (std::fseek(f.f, (long)(nodes[(size_t)i].pos - 1), SEEK_SET) ||
      (bytes && std::fwrite(data, 1, (size_t)bytes, f.f) != (size_t)bytes))
    return fail(DPD_EINVAL, "write to %s failed", path);
  const int64_t hdr[2] = {metaPos, (int64_t)nodes.size()};
  if (std::fseek(f.f, 0, SEEK_SET) || std::fwrite(hdr, 8, 2, f.f) != 2 ||
      std::fseek(f.f, (long)(metaPos - 1), SEEK_SET))
    return fail(DPD_EINVAL, "write to %s failed", path);
  for (const auto& nd : nodes) {
    unsigned char b[kMetaBytes];
    std::memcpy(b, nd.name, kNameLen);
    std::memcpy(b + kNameLen, &nd.pos, 8);
    std::memcpy(b + kNameLen + 8, &nd.size, 8);
    if (std::fwrite(b, 1, kMetaBytes, f.f) != (size_t)kMetaBytes) return fail(DPD_EINVAL, "write to %s failed", path);
  }
  return DPD_OK;
}

int dpd_node_read(const char* path, const char* name, char type, void* data, int64_t count) {
  if (!path || !name || (!data && count > 0) || count < 0) return fail(DPD_EINVAL, "dpd_node_read: bad arguments");
  const int eb = elem_bytes(type);
  if (!eb) return fail(DPD_EINVAL, "node type must be 'i', 'd', 's' or 'c'");
  File f;
  f.f = std::fopen(path, "rb");
  if (!f.f) return fail(DPD_EINVAL, "cannot open %s", path);
  int64_t metaPos;
  std::vector<Node> nodes;
  int rc = read_meta(f.f, metaPos, nodes);
  if (rc) return rc;
  const int i = find(nodes, full_name(name));
  if (i < 0) return fail(DPD_EINVAL, "node '%s' not in %s", name, path);
  const int64_t bytes = eb * count;
  if (bytes > nodes[(size_t)i].size)
    return fail(DPD_EINVAL, "node '%s' holds %lld bytes, %lld requested", name, (long long)nodes[(size_t)i].size,
                (long long)bytes);
  if (std::fseek(f.f, (long)(nodes[(size_t)i].pos - 1), SEEK_SET) ||
      (bytes && std::fread(data, 1, (size_t)bytes, f.f) != (size_t)bytes))
    return fail(DPD_EINVAL, "read of node '%s' failed", name);
  return DPD_OK;
}

// ---- moio lists -----------------------------------------------------------------
int dpd_moio_write(const dpd_ctx* c, const char* path, int64_t irrepx, const int64_t* popp, const int64_t* popq,
                   int64_t ioppq, const int64_t* popr, const int64_t* pops, int64_t ioprs, const double* data) {
  if (!c || !path || !popp || !popq || !popr || !pops) return fail(DPD_EINVAL, "dpd_moio_write: null argument");
  const Sym& S = c->sym;
  if (irrepx < 1 || irrepx > S.nirrep) return fail(DPD_EINVAL, "moio: irrep out of range");
  Control k;
  std::memset(&k, 0, sizeof k);
  k.irrepx = irrepx;
  std::memcpy(k.pop[0], popp, 64);
  std::memcpy(k.pop[1], popq, 64);
  std::memcpy(k.pop[2], popr, 64);
  std::memcpy(k.pop[3], pops, 64);
  k.ioppq[0] = ioppq;
  k.ioprs[0] = ioprs;
  k.irrepr[0] = irrepx;
  // inipckso.f / updmoiso.f: per right-pair irrep, numl x numr words from
  // record rec+1 on, padded to whole records
  int64_t total = 0, nrec = 0;
  for (int r = 1; r <= S.nirrep; r++) {
    const int l = S.dp[r - 1][irrepx - 1] + 1;
    const int64_t numl = S.irpdso(l, popp, popq, (int)ioppq), numr = S.irpdso(r, popr, pops, (int)ioprs);
    k.numl[0][r - 1] = numl;
    k.numr[0][r - 1] = numr;
    k.rec[0][r - 1] = nrec;
    nrec += (numl * numr + kRecWords - 1) / kRecWords;
    total += numl * numr;
  }
  k.rec[1][0] = nrec;  // rectmp(1, list + 1) after the last irrep
  std::vector<double> h;
  int rc = to_host(data, total, h);
  if (rc) return rc;
  std::vector<double> recs((size_t)(nrec * kRecWords), 0.0);
  int64_t off = 0;
  for (int r = 1; r <= S.nirrep; r++) {
    const int64_t n = k.numl[0][r - 1] * k.numr[0][r - 1];
    if (n) std::memcpy(&recs[(size_t)(k.rec[0][r - 1] * kRecWords)], &h[(size_t)off], sizeof(double) * n);
    off += n;
  }
  {
    File f;
    f.f = std::fopen(path, "wb");
    if (!f.f) return fail(DPD_EINVAL, "cannot create %s", path);
    if (!recs.empty() && std::fwrite(recs.data(), sizeof(double), recs.size(), f.f) != recs.size())
      return fail(DPD_EINVAL, "write to %s failed", path);
  }
  std::vector<unsigned char> cb;
  put_record(cb, &k.irrepx, 8);
  for (int q = 0; q < 4; q++) put_record(cb, k.pop[q], 64);
  put_record(cb, k.ioppq, 64);
  put_record(cb, k.ioprs, 64);
  put_record(cb, k.irrepr, 64);
  put_record(cb, k.irrepi, 64);
  put_record(cb, k.numl, sizeof k.numl);
  put_record(cb, k.numr, sizeof k.numr);
  put_record(cb, k.rec, sizeof k.rec);
  File f;
  f.f = std::fopen((std::string(path) + ".c").c_str(), "wb");
  if (!f.f || std::fwrite(cb.data(), 1, cb.size(), f.f) != cb.size())
    return fail(DPD_EINVAL, "cannot write %s.c", path);
  return DPD_OK;
}

// describes the list: irrepx, pops[4][8], iop[2] and its packed length
int dpd_moio_info(const char* path, int64_t* irrepx, int64_t* pops32, int64_t* iop2, int64_t* length) {
  if (!path) return fail(DPD_EINVAL, "null path");
  Control k;
  int rc = read_control(path, k);
  if (rc) return rc;
  if (irrepx) *irrepx = k.irrepr[0];
  if (pops32) std::memcpy(pops32, k.pop, sizeof k.pop);
  if (iop2) {
    iop2[0] = k.ioppq[0];
    iop2[1] = k.ioprs[0];
  }
  if (length) {
    int64_t t = 0;
    for (int r = 0; r < 10; r++) t += k.numl[0][r] * k.numr[0][r];
    *length = t;
  }
  return DPD_OK;
}

// getallso_d.f: every irrep block of list 1 into dst (host or device), in order
int dpd_moio_read(const char* path, double* dst, int64_t length) {
  if (!path || (!dst && length > 0)) return fail(DPD_EINVAL, "dpd_moio_read: bad arguments");
  Control k;
  int rc = read_control(path, k);
  if (rc) return rc;
  int64_t total = 0;
  for (int r = 0; r < 10; r++) total += k.numl[0][r] * k.numr[0][r];
  if (length != total) return fail(DPD_EINVAL, "moio list %s holds %lld doubles, not %lld", path, (long long)total,
                                   (long long)length);
  File f;
  f.f = std::fopen(path, "rb");
  if (!f.f) return fail(DPD_EINVAL, "cannot open %s", path);
  std::vector<double> h((size_t)std::max<int64_t>(total, 1));
  int64_t off = 0;
  for (int r = 0; r < 10; r++) {
    const int64_t n = k.numl[0][r] * k.numr[0][r];
    if (!n) continue;
    if (std::fseek(f.f, (long)(k.rec[0][r] * kRecWords * 8), SEEK_SET) ||
        std::fread(&h[(size_t)off], sizeof(double), (size_t)n, f.f) != (size_t)n)
      return fail(DPD_EINVAL, "moio list %s: truncated block of irrep %d", path, r + 1);
    off += n;
  }
  return from_host(dst, h.data(), total);
}

// ---- IO_stream files ----------------------------------------------------------------
// lib_cc/IO.f90:87-154.  mode: 1 val + vec, 2 val only, 3 vec only; vec may be a
// device pointer; replace: create the file anew (else it must exist)
int dpd_stream_write(const char* path, int64_t idx, double val, const double* vec, int64_t n, int mode, int replace) {
  if (!path || idx < 1 || mode < 1 || mode > 3 || n < 0 || (mode != 2 && n > 0 && !vec))
    return fail(DPD_EINVAL, "dpd_stream_write: bad arguments");
  std::vector<double> h;
  if (mode != 2) {
    int rc = to_host(vec, n, h);
    if (rc) return rc;
  }
  File f;
  f.f = std::fopen(path, replace ? "wb" : "r+b");
  if (!f.f) return fail(DPD_EINVAL, "cannot open %s", path);
  if (std::fseek(f.f, (long)(8 * (idx - 1)), SEEK_SET)) return fail(DPD_EINVAL, "seek in %s failed", path);
  if (mode != 3 && std::fwrite(&val, 8, 1, f.f) != 1) return fail(DPD_EINVAL, "write to %s failed", path);
  if (mode != 2 && n && std::fwrite(h.data(), 8, (size_t)n, f.f) != (size_t)n)
    return fail(DPD_EINVAL, "write to %s failed", path);
  return DPD_OK;
}

int dpd_stream_read(const char* path, int64_t idx, double* val, double* vec, int64_t n, int mode) {
  if (!path || idx < 1 || mode < 1 || mode > 3 || n < 0 || (mode != 3 && !val) || (mode != 2 && n > 0 && !vec))
    return fail(DPD_EINVAL, "dpd_stream_read: bad arguments");
  File f;
  f.f = std::fopen(path, "rb");
  if (!f.f) return fail(DPD_EINVAL, "cannot open %s", path);
  if (std::fseek(f.f, (long)(8 * (idx - 1)), SEEK_SET)) return fail(DPD_EINVAL, "seek in %s failed", path);
  if (mode != 3 && std::fread(val, 8, 1, f.f) != 1) return fail(DPD_EINVAL, "%s: read past the end", path);
  if (mode != 2 && n) {
    std::vector<double> h((size_t)n);
    if (std::fread(h.data(), 8, (size_t)n, f.f) != (size_t)n) return fail(DPD_EINVAL, "%s: read past the end", path);
    return from_host(vec, h.data(), n);
  }
  return DPD_OK;
}

// mod_davidson.f90:388-390 / :588-589: roots of state block istate as [w, R]
int dpd_eigvec_write(const char* path, int64_t istate, int64_t nroot, int64_t dim, const double* w, const double* R,
                     int replace) {
  if (!path || !w || !R || istate < 1 || nroot < 1 || dim < 1) return fail(DPD_EINVAL, "dpd_eigvec_write: bad arguments");
  for (int64_t r = 0; r < nroot; r++) {
    int rc = dpd_stream_write(path, (istate + r - 1) * (1 + dim) + 1, w[r], R + (size_t)r * dim, dim, 1,
                              replace && r == 0);
    if (rc) return rc;
  }
  return DPD_OK;
}

int dpd_eigvec_read(const char* path, int64_t istate, int64_t nroot, int64_t dim, double* w, double* R) {
  if (!path || !w || !R || istate < 1 || nroot < 1 || dim < 1) return fail(DPD_EINVAL, "dpd_eigvec_read: bad arguments");
  for (int64_t r = 0; r < nroot; r++) {
    int rc = dpd_stream_read(path, (istate + r - 1) * (1 + dim) + 1, &w[r], R + (size_t)r * dim, dim, 1);
    if (rc) return rc;
  }
  return DPD_OK;
}

}  // extern "C"
