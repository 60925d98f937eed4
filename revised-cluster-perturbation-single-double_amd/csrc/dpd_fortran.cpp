// Fortran-callable entry points with the reference's own names and argument
// lists, so that lib_cc and the method programs link against this library
// unchanged (in place of the libsrint objects that define the same names).
//
// Calling convention of the reference build (src/config.h:7-11, -i8): external
// names in lower case with one trailing underscore, every argument by
// reference, INTEGER = int64, and one hidden length per CHARACTER argument
// appended after the explicit arguments.  The labels are CHARACTER(len=4) (or
// len=1 for the GEMM transposes); only their first characters are read, so the
// type of the hidden lengths (size_t for gfortran, another integer for other
// compilers) does not matter.
//
// Arrays may live in device memory (the reference's igpu = 1 contract,
// libsrint/vv2v_dpd_444.f90:45-46) or in host memory (igpu = 0): host operands
// are staged through grow-only device buffers and the output is copied back.
// Every call returns after its result is complete, as the reference's do.  An
// error prints the message and stops the program, as the reference's
// `write(6,*) ...; stop` does, but with exit status 1.  Data are FP64 (the
// reference's prec = 8) or, after dpd_fortran_set_prec_(4), FP32 arrays passed
// through the same dummy arguments (the reference's prec = 4).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "dpd_engine.hpp"
#include "dpd_fortran.h"

using namespace dpd;

namespace {

dpd_ctx* g_ctx = nullptr;
DevBuf g_stage[3];

[[noreturn]] void stop_with(const char* where) {
  std::printf(" %s: %s\n", where, dpd_last_error());
  std::fflush(stdout);
  std::exit(1);
}

void need_ctx(const char* where) {
  if (!g_ctx) {
    std::printf(" %s: call dpd_fortran_init(nirrep, dirprd, device) first\n", where);
    std::fflush(stdout);
    std::exit(1);
  }
}

bool on_device(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // plain host memory is unknown to the runtime
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged;
}

// bytes of one tensor element under the current prec (8 or 4)
size_t elem_bytes() { return g_ctx && dpd_ctx_data_precision(g_ctx) == 4 ? sizeof(float) : sizeof(double); }

// device view of an operand of n elements in slot k; host data is copied in
// when `load` is set
double* view(int k, const double* p, int64_t n, bool load, const char* where) {
  if (on_device(p) || n <= 0) return const_cast<double*>(p);
  DevBuf& b = g_stage[k];
  const size_t bytes = elem_bytes() * (size_t)n;
  if (b.bytes < bytes) {
    b.reset();
    if (b.alloc(bytes)) stop_with(where);
  }
  if (load && hipMemcpyAsync(b.p, p, bytes, hipMemcpyHostToDevice, g_ctx->stream) != hipSuccess)
    stop_with(where);
  return (double*)b.p;
}

void finish(double* host, const double* dev, int64_t n, const char* where) {
  if (host != dev && n > 0 &&
      hipMemcpyAsync(host, dev, elem_bytes() * (size_t)n, hipMemcpyDeviceToHost, g_ctx->stream) != hipSuccess)
    stop_with(where);
  if (hipStreamSynchronize(g_ctx->stream) != hipSuccess) stop_with(where);
}

dpd_tensor tensor4(const int64_t* p1, const int64_t* p2, const int64_t* iop12, const int64_t* p3, const int64_t* p4,
                   const int64_t* iop34, const int64_t* sym) {
  dpd_tensor t;
  std::memset(&t, 0, sizeof t);
  std::memcpy(t.pop[0], p1, 64);
  std::memcpy(t.pop[1], p2, 64);
  std::memcpy(t.pop[2], p3, 64);
  std::memcpy(t.pop[3], p4, 64);
  t.iop[0] = *iop12;
  t.iop[1] = *iop34;
  t.sym = *sym;
  return t;
}

dpd_tensor tensor2(const int64_t* p1, const int64_t* p2, const int64_t* iop12, const int64_t* sym) {
  dpd_tensor t;
  std::memset(&t, 0, sizeof t);
  std::memcpy(t.pop[0], p1, 64);
  std::memcpy(t.pop[1], p2, 64);
  t.pop[2][0] = 1;  // the 1-sized dummy pair of the rank-2 operands
  t.pop[3][0] = 1;
  t.iop[0] = *iop12;
  t.iop[1] = 1;
  t.sym = *sym;
  return t;
}

// a rank-3 operand (vv2v_dpd_323): pair (p,q) and a single index r, stored as
// the 4-index tensor with a 1-sized dummy fourth index
dpd_tensor tensor3(const int64_t* p1, const int64_t* p2, const int64_t* iop12, const int64_t* p3, const int64_t* sym) {
  dpd_tensor t;
  std::memset(&t, 0, sizeof t);
  std::memcpy(t.pop[0], p1, 64);
  std::memcpy(t.pop[1], p2, 64);
  std::memcpy(t.pop[2], p3, 64);
  t.pop[3][0] = 1;
  t.iop[0] = *iop12;
  t.iop[1] = 1;
  t.sym = *sym;
  return t;
}

int64_t len_of(const dpd_tensor& t) { return dpd_tensor_len(g_ctx, &t); }

// one engine call: stage, run, copy back
template <class F>
void engine(const char* where, dpd_tensor& a, const double* v1, dpd_tensor& b, const double* v2, dpd_tensor& c,
            double* v3, F&& call) {
  need_ctx(where);
  const int64_t n1 = len_of(a), n2 = len_of(b), n3 = len_of(c);
  if (n1 < 0 || n2 < 0 || n3 < 0) stop_with(where);
  a.data = view(0, v1, n1, true, where);
  b.data = view(1, v2, n2, true, where);
  c.data = view(2, v3, n3, true, where);
  if (call()) stop_with(where);
  finish(v3, c.data, n3, where);
}

int64_t pair_len(int64_t sym, const int64_t* pp, const int64_t* pq, int64_t iop, const int64_t* pr,
                 const int64_t* ps, int64_t iop2) {
  return dpd_idsymsoc(g_ctx, sym, pp, pq, iop, pr, ps, iop2);
}

template <class F>
void sort(const char* where, const double* in, int64_t nin, double* out, int64_t nout, F&& call) {
  need_ctx(where);
  if (nin < 0 || nout < 0) stop_with(where);
  const double* din = view(0, in, nin, true, where);
  double* dout = view(1, out, nout, false, where);
  if (call(din, dout)) stop_with(where);
  finish(out, dout, nout, where);
}

}  // namespace

extern "C" {

// set-up in place of the mod_symm globals (lib_mod/mod_symm.f): nirrep and the
// 8x8 direct-product table as Fortran stores it (column major)
void dpd_fortran_init_(const int64_t* nirrep, const int64_t* dirprd, const int64_t* device) {
  int64_t rm[64];
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) rm[i * 8 + j] = dirprd[j * 8 + i];
  if (g_ctx) dpd_ctx_destroy(g_ctx);
  g_ctx = nullptr;
  if (dpd_ctx_create(&g_ctx, *nirrep, rm, (int)*device)) stop_with("dpd_fortran_init");
}

void dpd_fortran_set_prec_(const int64_t* prec) {
  need_ctx("dpd_fortran_set_prec");
  if (dpd_ctx_set_data_precision(g_ctx, (int)*prec)) stop_with("dpd_fortran_set_prec");
}

void dpd_fortran_finalize_() {
  for (auto& b : g_stage) b.reset();
  if (g_ctx) dpd_ctx_destroy(g_ctx);
  g_ctx = nullptr;
}

// libsrint/vv2v_dpd_444.f90:55-57
void vv2v_dpd_444_(const double* v1, const int64_t* p1_1, const int64_t* p1_2, const int64_t* iop1_12,
                   const int64_t* p1_3, const int64_t* p1_4, const int64_t* iop1_34, const int64_t* sym1,
                   const char* type1, const double* v2, const int64_t* p2_1, const int64_t* p2_2,
                   const int64_t* iop2_12, const int64_t* p2_3, const int64_t* p2_4, const int64_t* iop2_34,
                   const int64_t* sym2, const char* type2, double* v3, const int64_t* p3_1, const int64_t* p3_2,
                   const int64_t* iop3_12, const int64_t* p3_3, const int64_t* p3_4, const int64_t* iop3_34,
                   const int64_t* sym3, const char* type3, const double* fact1, const double* fact2, size_t,
                   size_t, size_t) {
  dpd_tensor a = tensor4(p1_1, p1_2, iop1_12, p1_3, p1_4, iop1_34, sym1);
  dpd_tensor b = tensor4(p2_1, p2_2, iop2_12, p2_3, p2_4, iop2_34, sym2);
  dpd_tensor c = tensor4(p3_1, p3_2, iop3_12, p3_3, p3_4, iop3_34, sym3);
  engine("vv2v_dpd_444", a, v1, b, v2, c, v3,
         [&] { return dpd_vv2v_dpd_444(g_ctx, &a, type1, &b, type2, &c, type3, *fact1, *fact2); });
}

// libsrint/vv2v_dpd_424.f90:21-23
void vv2v_dpd_424_(const double* v1, const int64_t* p1_1, const int64_t* p1_2, const int64_t* iop1_12,
                   const int64_t* p1_3, const int64_t* p1_4, const int64_t* iop1_34, const int64_t* sym1,
                   const char* type1, const double* v2, const int64_t* p2_1, const int64_t* p2_2,
                   const int64_t* iop2_12, const int64_t* sym2, const char* type2, double* v3, const int64_t* p3_1,
                   const int64_t* p3_2, const int64_t* iop3_12, const int64_t* p3_3, const int64_t* p3_4,
                   const int64_t* iop3_34, const int64_t* sym3, const char* type3, const double* fact1,
                   const double* fact2, size_t, size_t, size_t) {
  dpd_tensor a = tensor4(p1_1, p1_2, iop1_12, p1_3, p1_4, iop1_34, sym1);
  dpd_tensor b = tensor2(p2_1, p2_2, iop2_12, sym2);
  dpd_tensor c = tensor4(p3_1, p3_2, iop3_12, p3_3, p3_4, iop3_34, sym3);
  engine("vv2v_dpd_424", a, v1, b, v2, c, v3,
         [&] { return dpd_vv2v_dpd_424(g_ctx, &a, type1, &b, type2, &c, type3, *fact1, *fact2); });
}

// libsrint/vv2v_dpd_442.f90:22-24
void vv2v_dpd_442_(const double* v1, const int64_t* p1_1, const int64_t* p1_2, const int64_t* iop1_12,
                   const int64_t* p1_3, const int64_t* p1_4, const int64_t* iop1_34, const int64_t* sym1,
                   const char* type1, const double* v2, const int64_t* p2_1, const int64_t* p2_2,
                   const int64_t* iop2_12, const int64_t* p2_3, const int64_t* p2_4, const int64_t* iop2_34,
                   const int64_t* sym2, const char* type2, double* v3, const int64_t* p3_1, const int64_t* p3_2,
                   const int64_t* iop3_12, const int64_t* sym3, const char* type3, const double* fact1,
                   const double* fact2, size_t, size_t, size_t) {
  dpd_tensor a = tensor4(p1_1, p1_2, iop1_12, p1_3, p1_4, iop1_34, sym1);
  dpd_tensor b = tensor4(p2_1, p2_2, iop2_12, p2_3, p2_4, iop2_34, sym2);
  dpd_tensor c = tensor2(p3_1, p3_2, iop3_12, sym3);
  engine("vv2v_dpd_442", a, v1, b, v2, c, v3,
         [&] { return dpd_vv2v_dpd_442(g_ctx, &a, type1, &b, type2, &c, type3, *fact1, *fact2); });
}

// libsrint/vv2v_dpd_323.f90:22-25 (CHARACTER*3 / *2 / *3 label strings)
void vv2v_dpd_323_(const double* v1, const int64_t* p1_1, const int64_t* p1_2, const int64_t* iop1_12,
                   const int64_t* p1_3, const int64_t* sym1, const char* type1, const double* v2,
                   const int64_t* p2_1, const int64_t* p2_2, const int64_t* iop2_12, const int64_t* sym2,
                   const char* type2, double* v3, const int64_t* p3_1, const int64_t* p3_2, const int64_t* iop3_12,
                   const int64_t* p3_3, const int64_t* sym3, const char* type3, const double* fact1,
                   const double* fact2, size_t, size_t, size_t) {
  dpd_tensor a = tensor3(p1_1, p1_2, iop1_12, p1_3, sym1);
  dpd_tensor b = tensor2(p2_1, p2_2, iop2_12, sym2);
  dpd_tensor c = tensor3(p3_1, p3_2, iop3_12, p3_3, sym3);
  engine("vv2v_dpd_323", a, v1, b, v2, c, v3,
         [&] { return dpd_vv2v_dpd_323(g_ctx, &a, type1, &b, type2, &c, type3, *fact1, *fact2); });
}

// libsrint/xgemmso.f:1-25 (XGEMMSO, the prec/igpu router), column major
void xgemmso_(const char* transa, const char* transb, const int64_t* m, const int64_t* n, const int64_t* k,
              const double* alpha, const double* A, const int64_t* lda, const double* B, const int64_t* ldb,
              const double* beta, double* C, const int64_t* ldc, size_t, size_t) {
  need_ctx("xgemmso");
  const bool ta = *transa == 'T' || *transa == 't' || *transa == 'C' || *transa == 'c';
  const bool tb = *transb == 'T' || *transb == 't' || *transb == 'C' || *transb == 'c';
  // the reference wraps every leading dimension as MAX(1,LD) (xgemmso.f:11-22):
  // its callers pass block sizes that are 0 for empty irrep blocks
  const int64_t la = std::max<int64_t>(1, *lda), lb = std::max<int64_t>(1, *ldb), lc = std::max<int64_t>(1, *ldc);
  if (*m <= 0 || *n <= 0) return;  // empty C: nothing is read or written (DGEMM's quick return)
  const int64_t na = la * (ta ? *m : *k), nb = lb * (tb ? *k : *n), nc = lc * *n;
  const double* dA = view(0, A, na, true, "xgemmso");
  const double* dB = view(1, B, nb, true, "xgemmso");
  double* dC = view(2, C, nc, true, "xgemmso");
  const char sa[2] = {ta ? 'T' : 'N', 0}, sb[2] = {tb ? 'T' : 'N', 0};
  if (dpd_xgemm(g_ctx, sa, sb, *m, *n, *k, *alpha, dA, la, dB, lb, *beta, dC, lc, 0)) stop_with("xgemmso");
  finish(C, dC, nc, "xgemmso");
}

// libsrint/sstgenso.f90:22 (sstgenso1) and :1 (sstgenso, whose nsize/ipos are unused)
void sstgenso1_(const double* win, double* wout, const int64_t* pop1, const int64_t* pop2, const int64_t* pop3,
                const int64_t* pop4, const int64_t* irrepx, const char* type, size_t) {
  need_ctx("sstgenso1");
  const int64_t n = pair_len(*irrepx, pop1, pop2, 1, pop3, pop4, 1);
  sort("sstgenso1", win, n, wout, n, [&](const double* i, double* o) {
    char t[5] = {type[0], type[1], type[2], type[3], 0};
    return dpd_sstgenso(g_ctx, i, o, pop1, pop2, pop3, pop4, *irrepx, t);
  });
}

void sstgenso_(const double* win, double* wout, const int64_t*, const int64_t* pop1, const int64_t* pop2,
               const int64_t* pop3, const int64_t* pop4, const double*, const int64_t* irrepx, const char* type,
               size_t l) {
  sstgenso1_(win, wout, pop1, pop2, pop3, pop4, irrepx, type, l);
}

// libsrint/iexpso.f90:24
void iexpso_(const double* iin, double* iout, const int64_t* popp, const int64_t* popq, const int64_t* ioppq,
             const int64_t* iexppq, const int64_t* popr, const int64_t* pops, const int64_t* ioprs,
             const int64_t* iexprs, const int64_t* sym) {
  need_ctx("iexpso");
  const int64_t nin = pair_len(*sym, popp, popq, *ioppq, popr, pops, *ioprs);
  const int64_t nout = pair_len(*sym, popp, popq, *iexppq ? 1 : *ioppq, popr, pops, *iexprs ? 1 : *ioprs);
  sort("iexpso", iin, nin, iout, nout, [&](const double* i, double* o) {
    return dpd_iexpso(g_ctx, i, o, popp, popq, *ioppq, *iexppq, popr, pops, *ioprs, *iexprs, *sym);
  });
}

// libsrint/icompso.f:1
void icompso_(const double* iin, double* iout, const int64_t* popp, const int64_t* popq, const int64_t* ioppq,
              const int64_t* icpq, const int64_t* popr, const int64_t* pops, const int64_t* ioprs,
              const int64_t* icrs, const int64_t* sym) {
  need_ctx("icompso");
  const int64_t nin = pair_len(*sym, popp, popq, *ioppq, popr, pops, *ioprs);
  const int64_t nout =
      pair_len(*sym, popp, popq, *icpq == 1 ? *ioppq : *icpq, popr, pops, *icrs == 1 ? *ioprs : *icrs);
  sort("icompso", iin, nin, iout, nout, [&](const double* i, double* o) {
    return dpd_icompso(g_ctx, i, o, popp, popq, *ioppq, *icpq, popr, pops, *ioprs, *icrs, *sym);
  });
}

// libsrint/isymtrso.f:1
void isymtrso_(const double* iin, double* iout, const int64_t* popp, const int64_t* popq, const int64_t* ioppq,
               const int64_t* icpq, const int64_t* popr, const int64_t* pops, const int64_t* ioprs,
               const int64_t* icrs, const int64_t* sym) {
  need_ctx("isymtrso");
  const int64_t n = pair_len(*sym, popp, popq, *ioppq, popr, pops, *ioprs);
  sort("isymtrso", iin, n, iout, n, [&](const double* i, double* o) {
    return dpd_isymtrso(g_ctx, i, o, popp, popq, *ioppq, *icpq, popr, pops, *ioprs, *icrs, *sym);
  });
}

// libsrint/itranspso.f:1
void itranspso_(const double* iin, double* iout, const int64_t* popp, const int64_t* popq, const int64_t* ioppq,
                const int64_t* popr, const int64_t* pops, const int64_t* ioprs, const int64_t* sym) {
  need_ctx("itranspso");
  const int64_t n = pair_len(*sym, popp, popq, *ioppq, popr, pops, *ioprs);
  sort("itranspso", iin, n, iout, n, [&](const double* i, double* o) {
    return dpd_itranspso(g_ctx, i, o, popp, popq, *ioppq, popr, pops, *ioprs, *sym);
  });
}

// libsrint/iassymso.f:1
void iassymso_(const double* vin, double* vout, const int64_t* symv, const int64_t* popp, const int64_t* popq,
               const int64_t* popr, const int64_t* pops, const int64_t* iop) {
  need_ctx("iassymso");
  static const int io[6][4] = {{0, 1, 0, 0}, {1, 1, 1, 0}, {1, 1, 1, 1}, {1, 1, 1, 1}, {1, 0, 0, 0}, {1, 1, 0, 1}};
  if (*iop < 0 || *iop > 5) {
    std::printf(" error of iop in iassymso\n");
    std::fflush(stdout);
    std::exit(1);
  }
  const int* t = io[*iop];
  const int64_t nin = pair_len(*symv, popp, popq, t[0], popr, pops, t[1]);
  const int64_t nout = pair_len(*symv, popp, popq, t[2], popr, pops, t[3]);
  sort("iassymso", vin, nin, vout, nout, [&](const double* i, double* o) {
    return dpd_iassymso(g_ctx, i, o, *symv, popp, popq, popr, pops, *iop);
  });
}

// libsrint/irpdso.f:2, idsymsoc.f:2, isymoffso.f:2 (INTEGER*8 functions)
int64_t irpdso_(const int64_t* sym, const int64_t* popp, const int64_t* popq, const int64_t* ioppq) {
  need_ctx("irpdso");
  return dpd_irpdso(g_ctx, *sym, popp, popq, *ioppq);
}

int64_t idsymsoc_(const int64_t* irrepx, const int64_t* popp, const int64_t* popq, const int64_t* ioppq,
                  const int64_t* popr, const int64_t* pops, const int64_t* ioprs) {
  need_ctx("idsymsoc");
  return dpd_idsymsoc(g_ctx, *irrepx, popp, popq, *ioppq, popr, pops, *ioprs);
}

int64_t isymoffso_(const int64_t* irrepq1, const int64_t* irrep, const int64_t* popp, const int64_t* popq,
                   const int64_t* iop) {
  need_ctx("isymoffso");
  return dpd_isymoffso(g_ctx, *irrepq1, *irrep, popp, popq, *iop);
}

}  // extern "C"
