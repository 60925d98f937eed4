// dpd_capi.cpp -- extern "C" entry points of include/dpd_hip.h (context,
// metadata, engines, contract surface, sort kernels, vector primitives).
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "dpd_engine.hpp"

using namespace dpd;

namespace dpd {
const char* last_error();
}

#define CHECK_CTX(c) \
  if (!(c)) return fail(DPD_EINVAL, "null context")

static int set_device(dpd_ctx* c) {
  if (c->device < 0) return fail(DPD_ENODEV, "host-only context (device -1) cannot run device work");
  DPD_HIP_TRY(hipSetDevice(c->device));
  return DPD_OK;
}

// ---- scratch of the allocation-free GEMM (dpd_xgemm_noalloc) -----------------
// One file-scope device array per GPU, held until the process ends: the task
// and segment tables of the call at its start, the packed operand panels after
// them.  In host code `scratch` names a host-side stand-in, not device memory:
// its device address is looked up once per device (hipGetSymbolAddress, with
// that device current, when the first context on it is created) and only that
// address is handed to kernels.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchTableBytes = 4096;  // GemmTask at 0, GemmSeg at 1024
static_assert(sizeof(GemmTask) <= 1024 && sizeof(GemmSeg) <= 1024, "scratch table slots");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchMaxDevices = 64;
static void* g_scratch_dev[kScratchMaxDevices];
static std::mutex g_scratch_mu;

// device address of this device's scratch (the device is current)
static int scratch_setup(int device) {
  if (device < 0 || device >= kScratchMaxDevices) return fail(DPD_ENODEV, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return DPD_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(DPD_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
  g_scratch_dev[device] = ptr;
  return DPD_OK;
}

// the tables of a one-task GEMM, stored into scratch by a kernel (arguments
// by value: no host buffer has to outlive the call, and a graph capture
// records a kernel node, not a copy from host memory)
struct XgemmTables {
  GemmTask task;
  GemmSeg seg;
};
__global__ void k_put_xgemm_tables(unsigned char* dst, XgemmTables t) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *(GemmTask*)dst = t.task;
    *(GemmSeg*)(dst + 1024) = t.seg;
  }
}

extern "C" {

const char* dpd_last_error(void) { return dpd::last_error(); }

int dpd_ctx_create(dpd_ctx** out, int64_t nirrep, const int64_t* dirprd64, int device) {
  if (!out || !dirprd64 || nirrep < 1 || nirrep > 8) return fail(DPD_EINVAL, "dpd_ctx_create: bad arguments");
  if (device == -1) {
    // host-only context: metadata functions work, device work fails with DPD_ENODEV
    dpd_ctx* c = new dpd_ctx();
    c->sym.set((int)nirrep, dirprd64);
    c->device = -1;
    *out = c;
    return DPD_OK;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
    return fail(DPD_ENODEV, "no HIP device %d available", device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(DPD_ENODEV, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(DPD_ENODEV, "device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) {
      int64_t v = dirprd64[i * 8 + j];
      if ((i < nirrep && j < nirrep) && (v < 1 || v > nirrep))
        return fail(DPD_EINVAL, "dirprd(%d,%d)=%lld out of range", i + 1, j + 1, (long long)v);
    }
  dpd_ctx* c = new dpd_ctx();
  c->sym.set((int)nirrep, dirprd64);
  c->device = device;
  if (set_device(c)) {
    delete c;
    return DPD_EHIP;
  }
  int rc = scratch_setup(device);
  if (!rc) rc = c->red.alloc(sizeof(double) * (kRedBlocks * 72 + 64) + sizeof(int64_t) * (kRedBlocks + 8));
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return DPD_OK;
}

int dpd_ctx_destroy(dpd_ctx* c) {
  if (!c) return DPD_OK;
  if (c->device >= 0) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
  }
  delete c;
  return DPD_OK;
}

int dpd_ctx_set_stream(dpd_ctx* c, void* stream) {
  CHECK_CTX(c);
  c->stream = (hipStream_t)stream;
  return DPD_OK;
}

int dpd_ctx_synchronize(dpd_ctx* c) {
  CHECK_CTX(c);
  if (set_device(c)) return DPD_ENODEV;
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  return DPD_OK;
}

size_t dpd_ctx_workspace_bytes(const dpd_ctx* c) { return c ? c->ws.bytes : 0; }

// ---- metadata ---------------------------------------------------------------
int64_t dpd_irpdso(const dpd_ctx* c, int64_t sym, const int64_t* popp, const int64_t* popq, int64_t iop) {
  if (!c || sym < 1 || sym > c->sym.nirrep) return 0;
  return c->sym.irpdso((int)sym, popp, popq, (int)iop);
}

int64_t dpd_idsymsoc(const dpd_ctx* c, int64_t irrepx, const int64_t* popp, const int64_t* popq, int64_t ioppq,
                     const int64_t* popr, const int64_t* pops, int64_t ioprs) {
  if (!c || irrepx < 1 || irrepx > c->sym.nirrep) return 0;
  return c->sym.idsymsoc((int)irrepx, popp, popq, (int)ioppq, popr, pops, (int)ioprs);
}

int64_t dpd_isymoffso(const dpd_ctx* c, int64_t irrepq1, int64_t irrep, const int64_t* popp, const int64_t* popq,
                      int64_t iop) {
  if (!c) return 0;
  return c->sym.isymoffso((int)irrepq1, (int)irrep, popp, popq, (int)iop);
}

int64_t dpd_tensor_len(const dpd_ctx* c, const dpd_tensor* t) {
  if (!c || !t) return 0;
  return c->sym.idsymsoc((int)t->sym, t->pop[0], t->pop[1], (int)t->iop[0], t->pop[2], t->pop[3], (int)t->iop[1]);
}

int dpd_getidx(const dpd_ctx* c, int64_t idx, int64_t* info8, const dpd_tensor* t) {
  if (!c || !t || !info8) return fail(DPD_EINVAL, "dpd_getidx: null argument");
  DevLayout L = make_layout(c->sym, Desc::from(t, 4));
  if (idx < 1 || idx > L.len) return fail(DPD_EINVAL, "getidx: index %lld out of range", (long long)idx);
  int h[4];
  int64_t x[4];
  decode_elem(L, idx - 1, h, x);
  for (int i = 0; i < 4; i++) {
    info8[i] = h[i] + 1;
    info8[4 + i] = x[i] + 1;
  }
  return DPD_OK;
}

int64_t dpd_putidx(const dpd_ctx* c, const int64_t* info8, const dpd_tensor* t) {
  if (!c || !t || !info8) return 0;
  DevLayout L = make_layout(c->sym, Desc::from(t, 4));
  double sg;
  for (int i = 0; i < 4; i++)
    if (info8[i] < 1 || info8[i] > c->sym.nirrep) return 0;
  int64_t ad = fetch_addr(L, (int)info8[0] - 1, (int)info8[1] - 1, (int)info8[2] - 1, (int)info8[3] - 1,
                          info8[4] - 1, info8[5] - 1, info8[6] - 1, info8[7] - 1, &sg);
  // putidx addresses stored elements only (no mirroring)
  if (ad < 0) return 0;
  int h[4];
  int64_t x[4];
  decode_elem(L, ad, h, x);
  for (int i = 0; i < 4; i++)
    if (h[i] != info8[i] - 1 || x[i] != info8[4 + i] - 1) return 0;
  return ad + 1;
}

// ---- engines ------------------------------------------------------------------
static int check_iop(const dpd_tensor* t, int rank) {
  for (int s = 0; s < (rank == 4 ? 2 : 1); s++)
    if (t->iop[s] < 0 || t->iop[s] > 2) return fail(DPD_EINVAL, "iop must be 0, 1 or 2");
  return DPD_OK;
}

static int make_operand(dpd_ctx* c, const dpd_tensor* t, const char* lab, int rank, Operand& o) {
  if (!t) return fail(DPD_EINVAL, "null tensor");
  int rc = check_iop(t, rank);
  if (rc) return rc;
  if (t->sym < 1 || t->sym > c->sym.nirrep) return fail(DPD_EINVAL, "tensor symmetry out of range");
  if (!lab) return fail(DPD_EINVAL, "null label string");
  rc = check_labels(lab, rank);
  if (rc) return rc;
  o.d = Desc::from(t, rank);
  o.rank = rank;
  std::memset(o.lab, 0, 4);
  std::memcpy(o.lab, lab, rank);
  o.data = t->data;
  return DPD_OK;
}

static int engine_entry(dpd_ctx* c, int kind, const dpd_tensor* v1, const char* t1, const dpd_tensor* v2,
                        const char* t2, dpd_tensor* v3, const char* t3, double f1, double f2) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  Operand o1, o2, o3;
  int r2 = (kind == 424) ? 2 : 4, r3 = (kind == 442) ? 2 : 4;
  int rc = make_operand(c, v1, t1, 4, o1);
  if (!rc) rc = make_operand(c, v2, t2, r2, o2);
  if (!rc) rc = make_operand(c, v3, t3, r3, o3);
  if (rc) return rc;
  return engine_call(c, kind, o1, o2, o3, f1, f2, 1, c->data_prec == 4);
}

int dpd_vv2v_dpd_444(dpd_ctx* c, const dpd_tensor* v1, const char* t1, const dpd_tensor* v2, const char* t2,
                     dpd_tensor* v3, const char* t3, double f1, double f2) {
  return engine_entry(c, 444, v1, t1, v2, t2, v3, t3, f1, f2);
}
int dpd_vv2v_dpd_424(dpd_ctx* c, const dpd_tensor* v1, const char* t1, const dpd_tensor* v2, const char* t2,
                     dpd_tensor* v3, const char* t3, double f1, double f2) {
  return engine_entry(c, 424, v1, t1, v2, t2, v3, t3, f1, f2);
}
int dpd_vv2v_dpd_442(dpd_ctx* c, const dpd_tensor* v1, const char* t1, const dpd_tensor* v2, const char* t2,
                     dpd_tensor* v3, const char* t3, double f1, double f2) {
  return engine_entry(c, 442, v1, t1, v2, t2, v3, t3, f1, f2);
}

// libsrint/vv2v_dpd_323.f90:22-145: 3-index x 2-index -> 3-index through the 444
// engine.  The third index of v1 is the contracted one; each rank-3 tensor's
// single index is paired with a dummy index of population 1 in irrep 1, and v2
// with two of them, so that v1(pq,m c1) x v2(m c1,r c2) -> v3(pq,r c2) is the
// engine's standard contraction.  An intact pair (v1 and v3 open with the same
// two labels) may be packed; a pair that is reordered or split needs iop 1.
int dpd_vv2v_dpd_323(dpd_ctx* c, const dpd_tensor* v1, const char* type1, const dpd_tensor* v2, const char* type2,
                     dpd_tensor* v3, const char* type3, double f1, double f2) {
  CHECK_CTX(c);
  if (!v1 || !v2 || !v3) return fail(DPD_EINVAL, "null tensor");
  if (!type1 || !type2 || !type3) return fail(DPD_EINVAL, "null label string");
  const char cc = type1[2];
  for (int i = 0; i < 3; i++)
    if (type3[i] == cc)
      return fail(DPD_EINVAL, "vv2v_dpd_323: the third index of v1 (%c) must be the contracted one", cc);
  const bool intact = type1[0] == type3[0] && type1[1] == type3[1];
  if (v2->iop[0] != 1) return fail(DPD_EINVAL, "vv2v_dpd_323: the 2-index tensor v2 must be uncompressed (iop 1)");
  if (!intact && (v1->iop[0] != 1 || v3->iop[0] != 1))
    return fail(DPD_EINVAL, "vv2v_dpd_323: pair %.2s -> %.2s is reordered or split: v1 and v3 need iop 1", type1,
                type3);
  // two letters used by none of the label strings name the dummy indices
  char d1 = 0, d2 = 0;
  for (char l = 'a'; l <= 'z' && !d2; l++) {
    bool used = false;
    for (int i = 0; i < 3; i++) used = used || type1[i] == l || type3[i] == l;
    used = used || type2[0] == l || type2[1] == l;
    if (used) continue;
    if (!d1) d1 = l;
    else d2 = l;
  }
  if (!d2) return fail(DPD_EINVAL, "vv2v_dpd_323: no free letters for the dummy indices");
  char l1[4] = {type1[0], type1[1], type1[2], d1};
  char l3[4] = {type3[0], type3[1], type3[2], d2};
  char l2[4];
  if (type2[0] == cc) {
    l2[0] = type2[0], l2[1] = d1, l2[2] = type2[1], l2[3] = d2;
  } else if (type2[1] == cc) {
    l2[0] = type2[0], l2[1] = d2, l2[2] = type2[1], l2[3] = d1;
  } else {
    return fail(DPD_EINVAL, "vv2v_dpd_323: contracted index '%c' not found in v2 (%.2s)", cc, type2);
  }
  int64_t pdum[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  dpd_tensor a = *v1, b, d = *v3;
  std::memcpy(a.pop[3], pdum, sizeof pdum);
  a.iop[1] = 1;
  std::memcpy(d.pop[3], pdum, sizeof pdum);
  d.iop[1] = 1;
  std::memset(&b, 0, sizeof b);
  std::memcpy(b.pop[0], v2->pop[0], sizeof pdum);
  std::memcpy(b.pop[1], pdum, sizeof pdum);
  std::memcpy(b.pop[2], v2->pop[1], sizeof pdum);
  std::memcpy(b.pop[3], pdum, sizeof pdum);
  b.iop[0] = b.iop[1] = 1;
  b.sym = v2->sym;
  b.data = v2->data;
  return engine_entry(c, 444, &a, l1, &b, l2, &d, l3, f1, f2);
}

// ---- contract surface ---------------------------------------------------------
static int contract_entry(dpd_ctx* c, int kind, const dpd_tensor* v1, const char* t1, int r1,
                          const dpd_tensor* v2, const char* t2, int r2, dpd_tensor* v3, const char* t3,
                          int r3, double fact, int64_t anti) {
  CHECK_CTX(c);
  if (c->data_prec != 8)
    return fail(DPD_EINVAL, "the contract surface takes FP64 tensors only (lib_cc sets prec = 8, "
                            "readmodule.f90:64); data precision is %d", c->data_prec);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  int maxanti = (kind == 444) ? 3 : (kind == 424 ? 2 : 0);
  if (anti < 0 || anti > maxanti)
    return fail(DPD_EINVAL, "ERROR: vv2v_%d error in antisymmetrize. %lld", kind, (long long)anti);
  Operand o1, o2, o3;
  int rc = make_operand(c, v1, t1, r1, o1);
  if (!rc) rc = make_operand(c, v2, t2, r2, o2);
  if (!rc) rc = make_operand(c, v3, t3, r3, o3);
  if (rc) return rc;
  return contract_call(c, kind, o1, o2, o3, fact, (int)anti, 1, PREC_F64);
}

int dpd_vv2v_444(dpd_ctx* c, const dpd_tensor* v1, const dpd_tensor* v2, dpd_tensor* v3, double fact,
                 const char* t1, const char* t2, const char* t3, int64_t anti) {
  return contract_entry(c, 444, v1, t1, 4, v2, t2, 4, v3, t3, 4, fact, anti);
}

// mod_contract.f90:166-219: the rank-2 wrappers pad with '=','+','-'
int dpd_vv2v_222(dpd_ctx* c, const dpd_tensor* v1, const dpd_tensor* v2, dpd_tensor* v3, double fact,
                 const char* t1, const char* t2, const char* t3) {
  if (!t1 || !t2 || !t3) return fail(DPD_EINVAL, "null label string");
  char a[4] = {t1[0], t1[1], '=', '+'}, b[4] = {t2[0], t2[1], '=', '-'}, d[4] = {t3[0], t3[1], '+', '-'};
  return contract_entry(c, 444, v1, a, 4, v2, b, 4, v3, d, 4, fact, 0);
}
int dpd_vv2v_422(dpd_ctx* c, const dpd_tensor* v1, const dpd_tensor* v2, dpd_tensor* v3, double fact,
                 const char* t1, const char* t2, const char* t3) {
  if (!t1 || !t2 || !t3) return fail(DPD_EINVAL, "null label string");
  char b[4] = {t2[0], t2[1], '+', '-'}, d[4] = {t3[0], t3[1], '+', '-'};
  return contract_entry(c, 444, v1, t1, 4, v2, b, 4, v3, d, 4, fact, 0);
}
int dpd_vv2v_224(dpd_ctx* c, const dpd_tensor* v1, const dpd_tensor* v2, dpd_tensor* v3, double fact,
                 const char* t1, const char* t2, const char* t3, int64_t anti) {
  if (!t1 || !t2 || !t3) return fail(DPD_EINVAL, "null label string");
  char a[4] = {t1[0], t1[1], '+', '-'}, b[4] = {t2[0], t2[1], '+', '-'};
  return contract_entry(c, 444, v1, a, 4, v2, b, 4, v3, t3, 4, fact, anti);
}
int dpd_vv2v_424(dpd_ctx* c, const dpd_tensor* v1, const dpd_tensor* v2, dpd_tensor* v3, double fact,
                 const char* t1, const char* t2, const char* t3, int64_t anti) {
  return contract_entry(c, 424, v1, t1, 4, v2, t2, 2, v3, t3, 4, fact, anti);
}
int dpd_vv2v_442(dpd_ctx* c, const dpd_tensor* v1, const dpd_tensor* v2, dpd_tensor* v3, double fact,
                 const char* t1, const char* t2, const char* t3) {
  return contract_entry(c, 442, v1, t1, 4, v2, t2, 4, v3, t3, 2, fact, 0);
}

// ---- sort kernels -------------------------------------------------------------
static Desc mk(const int64_t* p, const int64_t* q, const int64_t* r, const int64_t* s, int i1, int i2, int sym) {
  Desc d;
  std::memcpy(d.pop[0], p, 64);
  std::memcpy(d.pop[1], q, 64);
  std::memcpy(d.pop[2], r, 64);
  std::memcpy(d.pop[3], s, 64);
  d.iop[0] = i1;
  d.iop[1] = i2;
  d.sym = sym;
  return d;
}

static bool same8(const int64_t* a, const int64_t* b) { return std::memcmp(a, b, 64) == 0; }

static int run_sort(dpd_ctx* c, const Desc& dout, const Desc& din, const int8_t* perm, const double* in,
                    double* out) {
  { int rc_ = set_device(c); if (rc_) return rc_; }
  if (din.sym < 1 || din.sym > c->sym.nirrep) return fail(DPD_EINVAL, "symmetry out of range");
  MapArgs a;
  std::memset(&a, 0, sizeof a);
  a.out = make_layout(c->sym, dout);
  a.in = make_layout(c->sym, din);
  a.nterm = 1;
  for (int i = 0; i < 4; i++) a.term[0].perm[i] = perm[i];
  a.term[0].coef = 1.0;
  a.alpha = 1.0;
  a.beta = 0.0;
  a.nbatch = 1;
  if (c->data_prec == 4) return gather_f32(c, a, (const float*)in, (float*)out);
  return gather(c, a, in, out);
}

int dpd_iexpso(dpd_ctx* c, const double* in, double* out, const int64_t* popp, const int64_t* popq, int64_t ioppq,
               int64_t iexppq, const int64_t* popr, const int64_t* pops, int64_t ioprs, int64_t iexprs, int64_t sym) {
  CHECK_CTX(c);
  if (ioppq < 0 || ioppq > 2) return fail(DPD_EINVAL, "fatal error: ioppq");
  if (ioprs < 0 || ioprs > 2) return fail(DPD_EINVAL, "fatal error: ioprs");
  if (iexppq == 1 && ioppq != 1 && !same8(popp, popq)) return fail(DPD_EINVAL, "fatal error: popp /= popq");
  if (iexprs == 1 && ioprs != 1 && !same8(popr, pops)) return fail(DPD_EINVAL, "fatal error: popr /= pops");
  bool ok = (iexppq == 1 && ioppq == 0 && iexprs == 0) || (ioprs == 0 && iexprs == 1 && iexppq == 0) ||
            (iexppq == 1 && ioppq == 2 && iexprs == 0) || (ioprs == 2 && iexprs == 1 && iexppq == 0) ||
            (iexppq == 1 && ioppq != 1 && iexprs == 1 && ioprs != 1);
  if (!ok) return fail(DPD_EINVAL, "fatal error: iexpso must expand exactly one side, or both packed sides");
  Desc din = mk(popp, popq, popr, pops, (int)ioppq, (int)ioprs, (int)sym);
  Desc dout = mk(popp, popq, popr, pops, iexppq ? 1 : (int)ioppq, iexprs ? 1 : (int)ioprs, (int)sym);
  const int8_t id[4] = {0, 1, 2, 3};
  return run_sort(c, dout, din, id, in, out);
}

int dpd_icompso(dpd_ctx* c, const double* in, double* out, const int64_t* popp, const int64_t* popq,
                int64_t ioppq, int64_t icpq, const int64_t* popr, const int64_t* pops, int64_t ioprs, int64_t icrs,
                int64_t sym) {
  CHECK_CTX(c);
  if (ioppq < 0 || ioppq > 2 || ioprs < 0 || ioprs > 2) return fail(DPD_EINVAL, "error of iop in icompso");
  bool ok = (icpq == 0 && icrs == 1) || (icpq == 1 && icrs == 0) || (icpq == 2 && icrs == 1) ||
            (icpq == 1 && icrs == 2);
  if (!ok) return fail(DPD_EINVAL, "attention!!!, error in icompso");
  if ((icpq != 1 && ioppq != 1) || (icrs != 1 && ioprs != 1))
    return fail(DPD_EINVAL, "icompso: the side being compressed must be full (iop=1)");
  Desc din = mk(popp, popq, popr, pops, (int)ioppq, (int)ioprs, (int)sym);
  Desc dout = mk(popp, popq, popr, pops, icpq == 1 ? (int)ioppq : (int)icpq, icrs == 1 ? (int)ioprs : (int)icrs,
                 (int)sym);
  const int8_t id[4] = {0, 1, 2, 3};
  return run_sort(c, dout, din, id, in, out);
}

int dpd_isymtrso(dpd_ctx* c, const double* in, double* out, const int64_t* popp, const int64_t* popq,
                 int64_t ioppq, int64_t icpq, const int64_t* popr, const int64_t* pops, int64_t ioprs, int64_t icrs,
                 int64_t sym) {
  CHECK_CTX(c);
  if (ioppq < 0 || ioppq > 2 || ioprs < 0 || ioprs > 2) return fail(DPD_EINVAL, "error of iop in isymtrso");
  if ((icpq == 1 && ioppq != 1) || (icrs == 1 && ioprs != 1))
    return fail(DPD_EINVAL, "isymtrso: a swapped pair must be full (iop=1)");
  Desc din = mk(popp, popq, popr, pops, (int)ioppq, (int)ioprs, (int)sym);
  Desc dout = mk(icpq == 1 ? popq : popp, icpq == 1 ? popp : popq, icrs == 1 ? pops : popr,
                 icrs == 1 ? popr : pops, (int)ioppq, (int)ioprs, (int)sym);
  const int8_t perm[4] = {(int8_t)(icpq == 1 ? 1 : 0), (int8_t)(icpq == 1 ? 0 : 1), (int8_t)(icrs == 1 ? 3 : 2),
                          (int8_t)(icrs == 1 ? 2 : 3)};
  return run_sort(c, dout, din, perm, in, out);
}

int dpd_itranspso(dpd_ctx* c, const double* in, double* out, const int64_t* popp, const int64_t* popq,
                  int64_t ioppq, const int64_t* popr, const int64_t* pops, int64_t ioprs, int64_t sym) {
  CHECK_CTX(c);
  if (ioppq < 0 || ioppq > 2 || ioprs < 0 || ioprs > 2) return fail(DPD_EINVAL, "error of iop in itranspso");
  Desc din = mk(popp, popq, popr, pops, (int)ioppq, (int)ioprs, (int)sym);
  Desc dout = mk(popr, pops, popp, popq, (int)ioprs, (int)ioppq, (int)sym);
  const int8_t perm[4] = {2, 3, 0, 1};
  return run_sort(c, dout, din, perm, in, out);
}

// libsrint/iassymso.f:1-142 (assym2so.f, assym3so.f, dasymrhfso.f, dsymrhfso.f):
//   iop 0: out(p<q,r<s) = in(p<q,r,s) - in(p<q,s,r)
//   iop 1: out(p,q,r<s) = in(p,q,r,s) - in(p,q,s,r)
//   iop 2: out(p,q,r,s) = in(p,q,r,s) - in(q,p,s,r)
//   iop 3: out(p,q,r,s) = in(p,q,r,s) + in(q,p,s,r)
//   iop 4: out(p<q,r<s) = in(p,q,r<s) - in(q,p,r<s)
//   iop 5: out(p<q,r,s) = in(p,q,r,s) - in(q,p,r,s)
// one two-term gather; a pair that is mirrored must have equal populations
int dpd_iassymso(dpd_ctx* c, const double* in, double* out, int64_t sym, const int64_t* popp,
                 const int64_t* popq, const int64_t* popr, const int64_t* pops, int64_t iop) {
  CHECK_CTX(c);
  if (iop < 0 || iop > 5) return fail(DPD_EINVAL, "error of iop in iassymso");
  const bool swap_pq = iop >= 2 && iop != 1, swap_rs = iop <= 3;
  if (swap_pq && !same8(popp, popq)) return fail(DPD_EINVAL, "iassymso: popp /= popq");
  if (swap_rs && !same8(popr, pops)) return fail(DPD_EINVAL, "iassymso: popr /= pops");
  static const int io[6][4] = {// in iop pq, rs; out iop pq, rs
                               {0, 1, 0, 0}, {1, 1, 1, 0}, {1, 1, 1, 1},
                               {1, 1, 1, 1}, {1, 0, 0, 0}, {1, 1, 0, 1}};
  { int rc_ = set_device(c); if (rc_) return rc_; }
  if (sym < 1 || sym > c->sym.nirrep) return fail(DPD_EINVAL, "symmetry out of range");
  MapArgs a;
  std::memset(&a, 0, sizeof a);
  a.in = make_layout(c->sym, mk(popp, popq, popr, pops, io[iop][0], io[iop][1], (int)sym));
  a.out = make_layout(c->sym, mk(popp, popq, popr, pops, io[iop][2], io[iop][3], (int)sym));
  a.nterm = 2;
  const int8_t id[4] = {0, 1, 2, 3};
  const int8_t sw[4] = {(int8_t)(swap_pq ? 1 : 0), (int8_t)(swap_pq ? 0 : 1), (int8_t)(swap_rs ? 3 : 2),
                        (int8_t)(swap_rs ? 2 : 3)};
  for (int i = 0; i < 4; i++) {
    a.term[0].perm[i] = id[i];
    a.term[1].perm[i] = sw[i];
  }
  a.term[0].coef = 1.0;
  a.term[1].coef = iop == 3 ? 1.0 : -1.0;
  a.alpha = 1.0;
  a.beta = 0.0;
  a.nbatch = 1;
  // FP32 data: the one sum or difference is formed in FP64 and rounded once,
  // which is the FP32 result itself
  if (c->data_prec == 4) return gather_f32(c, a, (const float*)in, (float*)out);
  return gather(c, a, in, out);
}

int dpd_sstgenso(dpd_ctx* c, const double* win, double* wout, const int64_t* pop1, const int64_t* pop2,
                 const int64_t* pop3, const int64_t* pop4, int64_t irrepx, const char* type4) {
  CHECK_CTX(c);
  if (!type4) return fail(DPD_EINVAL, "null type");
  int8_t perm[4];
  const int64_t* pin[4] = {pop1, pop2, pop3, pop4};
  int seen = 0;
  for (int i = 0; i < 4; i++) {
    int p = type4[i] - '1';
    if (p < 0 || p > 3 || (seen & (1 << p)))
      return fail(DPD_EINVAL, "fatal error in sstgenso: invalid type received: %.4s", type4);
    seen |= 1 << p;
    perm[i] = (int8_t)p;
  }
  Desc din = mk(pop1, pop2, pop3, pop4, 1, 1, (int)irrepx);
  Desc dout = mk(pin[perm[0]], pin[perm[1]], pin[perm[2]], pin[perm[3]], 1, 1, (int)irrepx);
  return run_sort(c, dout, din, perm, win, wout);
}

// ---- dense GEMM (libsrint/xgemmso.f:1-96 router: XGEMMD/XGEMMDG/XGEMMSG) ----------
}  // extern "C"

// checks the arguments of a column-major GEMM call and describes it as one
// task with one K segment; *empty is set when C is empty (nothing to do)
static int xgemm_describe(dpd_ctx* c, const char* transa, const char* transb, int64_t m, int64_t n, int64_t k,
                          int64_t lda, int64_t ldb, int64_t ldc, int prec, GemmPlan& p, bool* empty) {
  *empty = false;
  if (!transa || !transb) return fail(DPD_EINVAL, "null trans flag");
  char ta = transa[0], tb = transb[0];
  if ((ta != 'N' && ta != 'T') || (tb != 'N' && tb != 'T')) return fail(DPD_EINVAL, "trans must be 'N' or 'T'");
  if (m < 0 || n < 0 || k < 0 || m > INT32_MAX || n > INT32_MAX || k > INT32_MAX)
    return fail(DPD_EINVAL, "bad GEMM sizes");
  if (m == 0 || n == 0) {  // empty C (BLAS quick return, before the leading-dimension rules)
    *empty = true;
    return DPD_OK;
  }
  if (ldc < std::max<int64_t>(1, m) || lda < std::max<int64_t>(1, ta == 'N' ? m : k) ||
      ldb < std::max<int64_t>(1, tb == 'N' ? k : n))
    return fail(DPD_EINVAL, "bad leading dimension");
  if (prec != PREC_F64 && prec != PREC_F32 && prec != PREC_BF16X3)
    return fail(DPD_EINVAL, "precision %d is not available", prec);
  GemmTask tk;
  std::memset(&tk, 0, sizeof tk);
  tk.M = (int32_t)m;
  tk.N = (int32_t)n;
  tk.md[0] = (int32_t)m;
  tk.md[1] = 1;
  tk.nd[0] = (int32_t)n;
  tk.nd[1] = 1;
  tk.sc_m[0] = 1;
  tk.sc_n[0] = ldc;
  GemmSeg sg;
  std::memset(&sg, 0, sizeof sg);
  sg.K = (int32_t)k;
  sg.kd[0] = (int32_t)std::max<int64_t>(k, 1);
  sg.kd[1] = 1;
  sg.sa_m[0] = ta == 'N' ? 1 : lda;
  sg.sa_k[0] = ta == 'N' ? lda : 1;
  sg.sb_k[0] = tb == 'N' ? 1 : ldb;
  sg.sb_n[0] = tb == 'N' ? ldb : 1;
  sg.ktiles = (int32_t)((k + 15) / 16);
  tk.nseg = 1;
  tk.ktiles_total = sg.ktiles;
  p.cfg = choose_gemm_cfg({tk}, c->gemm_cfg);
  tk.tiles_m = (int32_t)((m + gemm_bm(p.cfg) - 1) / gemm_bm(p.cfg));
  tk.tiles_n = (int32_t)((n + gemm_bn(p.cfg) - 1) / gemm_bn(p.cfg));
  tk.ksplit = 1;
  p.tasks.push_back(tk);
  p.segs.push_back(sg);
  p.nwg = tk.tiles_m * tk.tiles_n;
  return DPD_OK;
}

extern "C" {

int dpd_xgemm(dpd_ctx* c, const char* transa, const char* transb, int64_t m, int64_t n, int64_t k, double alpha,
              const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
              int prec) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  GemmPlan p;
  bool empty = false;
  // FP32 arrays (XGEMMSG, xgemmso.f:21-22): FP32 MFMA, FP64 alpha and beta
  const bool f32data = c->data_prec == 4;
  if (f32data) prec = PREC_F32;
  int rc = xgemm_describe(c, transa, transb, m, n, k, lda, ldb, ldc, prec, p, &empty);
  if (rc || empty) return rc;
  if ((rc = p.upload())) return rc;
  if (f32data) {
    ws_reset(c);
    if ((rc = ws_reserve(c, sizeof(double) * p.partial_elems + 256))) return rc;
    rc = run_gemm(c, p, A, 1, B, 1, C, alpha, beta, PREC_F32, true);
  } else if (prec == PREC_BF16X3 || c->packing == 2) {
    if ((rc = build_variant(p, prec))) return rc;
    const GemmVariant& V = p.pk[prec];
    ws_reset(c);
    rc = ws_reserve(c, pk_esize(prec) * (V.pa_elems + V.pb_elems) + sizeof(double) * V.partial_elems + 1024);
    if (rc) return rc;
    PackSrc a, b;
    a.data = A;
    b.data = B;
    rc = run_gemm_packed(c, p, prec, a, b, C, alpha, beta);
  } else {
    rc = run_gemm(c, p, A, 0, B, 0, C, alpha, beta, prec);
  }
  if (rc) return rc;
  // the plan's device arrays are freed on return: wait for the launch
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  return DPD_OK;
}

// dpd_xgemm without device allocation or synchronisation: the tables and the
// packed panels live in the device's static scratch (kScratchBytes); long-K
// tiles are not split over workgroups (no partial sums to hold)
int dpd_xgemm_noalloc(dpd_ctx* c, const char* transa, const char* transb, int64_t m, int64_t n, int64_t k,
                      double alpha, const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                      double* C, int64_t ldc, int prec) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  void* sp = g_scratch_dev[c->device];
  if (!sp) return fail(DPD_EINVAL, "dpd_xgemm_noalloc: no scratch address for device %d (internal)", c->device);
  if (c->data_prec != 8) return fail(DPD_EINVAL, "dpd_xgemm_noalloc: FP64 arrays only (data precision 8)");
  GemmPlan p;
  bool empty = false;
  int rc = xgemm_describe(c, transa, transb, m, n, k, lda, ldb, ldc, prec, p, &empty);
  if (rc || empty) return rc;
  unsigned char* base = (unsigned char*)sp;
  const bool packed = prec == PREC_BF16X3 || c->packing == 2;
  GemmVariant V;
  size_t need = kScratchTableBytes;
  if (packed) {
    // the tile dpd_xgemm's packed path takes (build_variant's choice)
    int cfg = choose_gemm_cfg(p.tasks, -1);
    if (cfg == GEMM_V1) cfg = GEMM_64x64;
    if ((rc = layout_variant(p, V, prec, cfg, false))) return rc;
    need += pk_esize(prec) * (size_t)(V.pa_elems + V.pb_elems) + 512;
  }
  if (need > kScratchBytes)
    return fail(DPD_EINVAL,
                "dpd_xgemm_noalloc: m=%lld n=%lld k=%lld needs %zu bytes of scratch, more than its %zu; "
                "use dpd_xgemm",
                (long long)m, (long long)n, (long long)k, need, kScratchBytes);
  XgemmTables t;
  std::memset(&t, 0, sizeof t);
  t.task = packed ? V.tasks[0] : p.tasks[0];
  t.seg = p.segs[0];
  hipLaunchKernelGGL(k_put_xgemm_tables, dim3(1), dim3(64), 0, c->stream, base, t);
  DPD_HIP_TRY(hipGetLastError());
  p.d_tasks = (GemmTask*)base;
  p.d_segs = (GemmSeg*)(base + 1024);
  if (!packed) return run_gemm(c, p, A, 0, B, 0, C, alpha, beta, prec);
  V.d_tasks = (GemmTask*)base;
  V.d_split = nullptr;
  unsigned char* pa = base + kScratchTableBytes;
  unsigned char* pb = pa + ((pk_esize(prec) * (size_t)V.pa_elems + 255) & ~size_t(255));
  PackSrc a, b;
  a.data = A;
  b.data = B;
  if (V.pa_elems && (rc = pack_side(c, p, V, 0, a, pa))) return rc;
  if (V.pb_elems && (rc = pack_side(c, p, V, 1, b, pb))) return rc;
  if (V.nwg == 0) return DPD_OK;
  GemmLaunch g;
  std::memset(&g, 0, sizeof g);
  g.C = C;
  g.tasks = V.d_tasks;
  g.segs = p.d_segs;
  g.ntask = 1;
  g.nwg = V.nwg;
  g.alpha = alpha;
  g.beta = beta;
  g.prec = prec;
  g.cfg = V.cfg;
  DPD_HIP_TRY(launch_gemm_pk_only(g, pa, pb, c->stream));
  return DPD_OK;
}

int dpd_ctx_set_gemm_config(dpd_ctx* c, int cfg, int packing) {
  CHECK_CTX(c);
  if (cfg < -2 || cfg > kGemmCfgMax) return fail(DPD_EINVAL, "GEMM configuration must be -2..%d", kGemmCfgMax);
  if (packing < 0 || packing > 2) return fail(DPD_EINVAL, "packing must be 0, 1 or 2");
  if (c->device >= 0) DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  c->gemm_cfg = cfg;
  c->packing = packing;
  c->group_cache.clear();  // holds pointers into contract_cache
  dpd::bump_mem_gen();
  c->contract_cache.clear();
  c->engine_cache.clear();
  c->map_cache.clear();
  c->map_bytes = 0;
  return DPD_OK;
}

int dpd_ctx_set_split_k(dpd_ctx* c, int mode) {
  CHECK_CTX(c);
  if (mode < -1 || mode > 1) return fail(DPD_EINVAL, "split-K mode must be -1, 0 or 1");
  if (c->device >= 0) DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  c->split_k = mode;
  c->group_cache.clear();
  dpd::bump_mem_gen();
  c->contract_cache.clear();
  c->engine_cache.clear();
  c->map_cache.clear();
  c->map_bytes = 0;
  return DPD_OK;
}

// libsrint/readinp.f90:15,120-127 (the mod_iop `prec` switch): 8 FP64 tensors,
// 4 FP32 tensors for the engines, sorts and xgemm (vv2v_dpd_444.f90:206-229
// routes to the _s/_sg engines, sstgenso.f90:35-42, xgemmso.f:14-22)
int dpd_ctx_set_data_precision(dpd_ctx* c, int prec) {
  CHECK_CTX(c);
  if (prec != 8 && prec != 4) return fail(DPD_EINVAL, "prec must be 8 or 4 (got %d)", prec);
  c->data_prec = prec;
  return DPD_OK;
}

int dpd_ctx_data_precision(const dpd_ctx* c) { return c ? c->data_prec : 0; }

// ---- profiler ------------------------------------------------------------------
int dpd_profile_enable(dpd_ctx* c, int on, int max_records) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  for (auto e : c->prof.pool) (void)hipEventDestroy(e);
  c->prof.pool.clear();
  c->prof.recs.clear();
  c->prof.used = 0;
  c->prof.on = on != 0;
  if (c->prof.on) {
    int n = 2 * std::max(1, max_records);
    c->prof.pool.resize(n);
    for (int i = 0; i < n; i++) DPD_HIP_TRY(hipEventCreate(&c->prof.pool[i]));
    c->prof.recs.reserve(max_records);
  }
  return DPD_OK;
}

// aggregate per tag: count, total milliseconds, flops per launch (of the last
// record), precision; JSON written into buf
int dpd_profile_report(dpd_ctx* c, char* buf, int64_t len) {
  CHECK_CTX(c);
  if (!buf || len < 3) return fail(DPD_EINVAL, "bad report buffer");
  { int rc_ = set_device(c); if (rc_) return rc_; }
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<std::string> tags;
  std::vector<double> ms;
  std::vector<int64_t> cnt, fl, fltot, by;
  std::vector<int> pr, cf, pk, sp, kd;
  for (auto& r : c->prof.recs) {
    float t = 0.f;
    DPD_HIP_TRY(hipEventElapsedTime(&t, r.a, r.b));
    // one entry per tag, precision and tile shape (a tag's plans for different
    // vector counts may be tuned to different tiles)
    std::string key = r.kind == 1 ? r.tag
                                  : r.tag + (r.prec == PREC_F64 ? ":fp64" : r.prec == PREC_F32 ? ":fp32" : ":bf16x3") +
                                        ":t" + std::to_string(r.cfg) + (r.split ? "s" : "");
    size_t k = std::find(tags.begin(), tags.end(), key) - tags.begin();
    if (k == tags.size()) {
      tags.push_back(key);
      ms.push_back(0);
      cnt.push_back(0);
      fl.push_back(0);
      fltot.push_back(0);
      by.push_back(0);
      pr.push_back(r.prec);
      cf.push_back(r.cfg);
      pk.push_back(r.packed);
      sp.push_back(r.split);
      kd.push_back(r.kind);
    }
    ms[k] += t;
    cnt[k] += 1;
    fl[k] = r.flops;
    fltot[k] += r.flops;
    by[k] += r.bytes;
    pr[k] = r.prec;
    cf[k] = r.cfg;
    pk[k] = r.packed;
  }
  std::string js = "{";
  for (size_t k = 0; k < tags.size(); k++) {
    char item[512];
    snprintf(item, sizeof item,
             "%s\"%s\": {\"count\": %lld, \"ms\": %.6f, \"flops_total\": %lld, \"bytes_total\": %lld, "
             "\"prec\": %d, \"cfg\": %d, \"packed\": %d, \"split\": %d, \"kind\": \"%s\"}",
             k ? ", " : "", tags[k].c_str(), (long long)cnt[k], ms[k], (long long)fltot[k], (long long)by[k], pr[k],
             cf[k], pk[k], sp[k], kd[k] ? "gather" : "gemm");
    js += item;
  }
  js += "}";
  if ((int64_t)js.size() + 1 > len) return fail(DPD_EINVAL, "report buffer too small (%zu)", js.size() + 1);
  std::memcpy(buf, js.c_str(), js.size() + 1);
  c->prof.recs.clear();
  c->prof.used = 0;
  return DPD_OK;
}

// ---- vector primitives ----------------------------------------------------------
static double* red_d(dpd_ctx* c) { return (double*)c->red.p; }

int dpd_dot(dpd_ctx* c, int64_t n, const double* x, const double* y, double* result) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  double* part = red_d(c);
  double* out = part + kRedBlocks * 64;
  DPD_HIP_TRY(launch_multidot(n, x, y, 0, 1, part, out, c->stream));
  DPD_HIP_TRY(hipMemcpyAsync(result, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  return DPD_OK;
}

int dpd_nrm2(dpd_ctx* c, int64_t n, const double* x, double* result) {
  double s;
  int rc = dpd_dot(c, n, x, x, &s);
  if (rc) return rc;
  *result = std::sqrt(s);
  return DPD_OK;
}

int dpd_amax(dpd_ctx* c, int64_t n, const double* x, int64_t* index, double* value) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  if (n <= 0) {
    *index = 0;
    *value = 0.0;
    return DPD_OK;
  }
  double* pv = red_d(c);
  int64_t* pi = (int64_t*)(pv + kRedBlocks * 72 + 64);
  double* ov = pv + kRedBlocks * 64 + 8;
  int64_t* oi = pi + kRedBlocks;
  DPD_HIP_TRY(launch_amax(n, x, pv, pi, ov, oi, c->stream));
  DPD_HIP_TRY(hipMemcpyAsync(value, ov, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DPD_HIP_TRY(hipMemcpyAsync(index, oi, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  *index += 1;
  return DPD_OK;
}

int dpd_axpy(dpd_ctx* c, int64_t n, double a, const double* x, double* y) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  DPD_HIP_TRY(launch_axpy(n, a, x, y, c->stream));
  return DPD_OK;
}

int dpd_scal(dpd_ctx* c, int64_t n, double a, double* x) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  DPD_HIP_TRY(launch_scal(n, a, x, c->stream));
  return DPD_OK;
}

int dpd_precondition(dpd_ctx* c, int64_t n, double w, const double* diag, const double* e, double* p) {
  CHECK_CTX(c);
  { int rc_ = set_device(c); if (rc_) return rc_; }
  DPD_HIP_TRY(launch_precondition(n, w, diag, e, p, c->stream));
  return DPD_OK;
}

}  // extern "C"
