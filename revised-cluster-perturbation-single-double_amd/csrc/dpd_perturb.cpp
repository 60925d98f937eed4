// dpd_perturb.cpp -- perturbation drivers on the device-resident problem.
//
//   * ground-state CPS(D) (CPSd.SF/run.f90, run1.f90, run2.f90, run3.f90):
//       T1^(1) = 0, T2^(1) = <ab||ij>/(e_i+e_j-e_a-e_b), E^(2) = <T2^(1)|abij>;
//       at n = 2, 3:  C1^(n) = <a,i|[V,T2^(n-1)]|0>          (calC1.f90, calC1_VR2)
//                     C2^(n) = <ab,ij|[V,T^(n-1)]|0> (+ [[V,T2^(1)],T2^(1)] at n = 3,
//                              calC2.f90, calC2_VR.f90, calC2_VTT.f90)
//                     T2^(n) = C2^(n)/(e_i+e_j-e_a-e_b)      (denomT2)
//                     -e(a,i) T1^(n) = <a,i|[V,T1^(n)]|0> + C1^(n)   by Jacobi
//                     steps on the orbital-energy denominators with DIIS
//                     E^(n+1) = <T2^(n)|abij>                (calEcorr.f90)
//   * EE-CPS(D) excitation energies of CIS roots order by order
//     (EE-CPSd.SF/run.f90, run_order.f90, calJ.f90, calE_C.f90, calV_C.f90,
//     calC_JR.f90, calC_VTR.f90, solver.f90, orthogonalR2L.f90):
//       [J^(0) - w0] R^(n) = sum_{k=2}^{n} w^(k) R^(n-k) - C^(n),  R^(n) orthogonal
//       to the CIS vector L, C^(n) = sum_k J^(k) R^(n-k), w^(n) = <L|C1^(n)>,
//       orders 1-3 with the terms CPS(D-3) keeps.
//   * grCPS(D) ground state (grCPSd.SF/run.f90, solver.f90): the CIS(Doo)
//     Jacobian as zeroth order, T^(1..3) by DIIS on (T1 | T2) at every order.
// Every contraction is one call of the contract surface on the problem's
// integral blocks; the DIIS history (libsrint/dosodiis2_d.f, which keeps it on
// disk) stays in HBM, only its small error matrix is solved on the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "dpd_sigma.hpp"

namespace dpd {
namespace {

Desc mk(const int64_t* a, const int64_t* b, const int64_t* c, const int64_t* d, int i1, int i2, int sym) {
  Desc x;
  std::memcpy(x.pop[0], a, 64);
  std::memcpy(x.pop[1], b, 64);
  std::memcpy(x.pop[2], c, 64);
  std::memcpy(x.pop[3], d, 64);
  x.iop[0] = i1;
  x.iop[1] = i2;
  x.sym = sym;
  return x;
}

Operand opnd(const Desc& d, const char* lab, int rank, const double* data) {
  Operand o;
  o.d = d;
  o.rank = rank;
  std::memset(o.lab, 0, 4);
  std::memcpy(o.lab, lab, (size_t)rank);
  o.data = data;
  return o;
}

struct Work {
  dpd_problem* p;
  dpd_ctx* c;
  hipStream_t st;
  const int64_t *v, *o, *z;
  DevBuf red;  // reduction partials and results
  Desc dR1(int s) const { return mk(v, o, z, z, 1, 1, s); }
  Desc dR2(int s) const { return mk(v, v, o, o, 0, 0, s); }
  Desc dFoo(int s) const { return mk(o, o, z, z, 1, 1, s); }
  Desc dFvv(int s) const { return mk(v, v, z, z, 1, 1, s); }
  Desc dWovvo(int s) const { return mk(v, o, v, o, 1, 1, s); }
  Desc dWovoo(int s) const { return mk(o, o, o, v, 0, 1, s); }
  Desc dWoooo(int s) const { return mk(o, o, o, o, 0, 0, s); }
  int64_t len(const Desc& d) const { return make_layout(c->sym, d).len; }
  Operand I(const char* which, const char* lab) const {
    const dpd_problem* q = p;
    if (!std::strcmp(which, "iabj")) return opnd(mk(v, o, v, o, 1, 1, 1), lab, 4, (const double*)q->iabj.p);
    if (!std::strcmp(which, "ijka")) return opnd(mk(o, o, o, v, 0, 1, 1), lab, 4, (const double*)q->ijka.p);
    if (!std::strcmp(which, "ijkl")) return opnd(mk(o, o, o, o, 0, 0, 1), lab, 4, (const double*)q->ijkl.p);
    if (!std::strcmp(which, "abij")) return opnd(mk(v, v, o, o, 0, 0, 1), lab, 4, (const double*)q->abij.p);
    if (!std::strcmp(which, "abci")) return opnd(mk(v, v, v, o, 0, 1, 1), lab, 4, (const double*)q->abci.p);
    return opnd(mk(v, v, v, v, 0, 0, 1), lab, 4, (const double*)q->abcd.p);
  }
  int k(int kind, const Operand& a, const Operand& b, const Operand& cc, double f, int anti, const char* tag) {
    return contract_call(c, kind, a, b, cc, f, anti, 1, PREC_F64, nullptr, tag);
  }
  // ---- named contractions (lib_cc/*.f90; packed (iop 0) pairs) -------------
  // c1_r2abci.f90 / c1_r2ijka.f90: c(a,i) += 1/2 <fe||am> r(fe,im) - 1/2 r(ae,mn) <mn||ie>
  int c1_vr2(double* C1, const double* R2, int s, double f) {
    int rc = k(442, I("abci", "feam"), opnd(dR2(s), "feim", 4, R2), opnd(dR1(s), "ai", 2, C1), f, 0, "c1_r2abci");
    if (!rc) rc = k(442, opnd(dR2(s), "aemn", 4, R2), I("ijka", "mnie"), opnd(dR1(s), "ai", 2, C1), -f, 0, "c1_r2ijka");
    return rc;
  }
  // calC2_VR1.f90: c2_r1ijka(ijka) + c2_r1abci
  int c2_vr1(double* C2, const double* R1, int s, double f) {
    int rc = c2_r1ijka(C2, R1, s, I("ijka", "ijmb"), f);
    if (!rc) rc = k(424, I("abci", "abej"), opnd(dR1(s), "ei", 2, R1), opnd(dR2(s), "abij", 4, C2), f, 2, "c2_r1abci");
    return rc;
  }
  // calC2_VR2.f90: ring (c2_r2Wmbej), hh-ladder (c2_r2Wmnij), pp-ladder (c2_r2abcd)
  int c2_vr2(double* C2, const double* R2, int s, double f) {
    int rc = c2_r2Wmbej(C2, s, R2, s, I("iabj", "embj"), f);
    if (!rc) rc = c2_r2Wmnij(C2, s, R2, s, I("ijkl", "mnij"), f);
    if (!rc && p->abcd_host) rc = problem_pp_ladder_streamed(p, R2, C2, 1, 0, s, f);  // abcd out of core
    else if (!rc)
      rc = k(444, I("abcd", "abef"), opnd(dR2(s), "efij", 4, R2), opnd(dR2(s), "abij", 4, C2), f, 0, "c2_r2abcd");
    return rc;
  }
  // the doubles-from-doubles terms take the symmetries of C2 (sc) and R2 (sr):
  // in the <u|[[V,T],R]|0> terms R2 is the totally symmetric T2
  int c2_r2Wmbej(double* C2, int sc, const double* R2, int sr, const Operand& W, double f) {
    return k(444, opnd(dR2(sr), "aeim", 4, R2), W, opnd(dR2(sc), "abij", 4, C2), f, 3, "c2_r2Wmbej");
  }
  int c2_r2Wmnij(double* C2, int sc, const double* R2, int sr, const Operand& W, double f) {
    return k(444, opnd(dR2(sr), "abmn", 4, R2), W, opnd(dR2(sc), "abij", 4, C2), f, 0, "c2_r2Wmnij");
  }
  int c2_r1ijka(double* C2, const double* R1, int s, const Operand& W, double f) {
    return k(424, W, opnd(dR1(s), "am", 2, R1), opnd(dR2(s), "abij", 4, C2), -f, 1, "c2_r1ijka");
  }
  // c1_r1iabj.f90 (vv2v_422 padded as the contract surface pads it)
  int c1_r1iabj(double* C1, const double* R1, int s, const Operand& W, double f) {
    return k(444, W, opnd(dR1(s), "fn+-", 4, R1), opnd(dR1(s), "ai+-", 4, C1), f, 0, "c1_r1iabj");
  }
  int c1_r1Fmi(double* C1, const double* R1, int s, const double* F, int sf, double f) {
    return k(444, opnd(dR1(s), "am=+", 4, R1), opnd(dFoo(sf), "mi=-", 4, F), opnd(dR1(s), "ai+-", 4, C1), -f, 0,
             "c1_r1Fmi");
  }
  int c1_r1Fae(double* C1, const double* R1, int s, const double* F, int sf, double f) {
    return k(444, opnd(dFvv(sf), "ae=+", 4, F), opnd(dR1(s), "ei=-", 4, R1), opnd(dR1(s), "ai+-", 4, C1), f, 0,
             "c1_r1Fae");
  }
  int c2_r2Fmi(double* C2, int sc, const double* R2, int sr, const double* F, int sf, double f) {
    return k(424, opnd(dR2(sr), "abim", 4, R2), opnd(dFoo(sf), "mj", 2, F), opnd(dR2(sc), "abij", 4, C2), -f, 2,
             "c2_r2Fmi");
  }
  int c2_r2Fae(double* C2, int sc, const double* R2, int sr, const double* F, int sf, double f) {
    return k(424, opnd(dR2(sr), "aeij", 4, R2), opnd(dFvv(sf), "be", 2, F), opnd(dR2(sc), "abij", 4, C2), f, 1,
             "c2_r2Fae");
  }
  // intermediates
  int Fmi_r2abij(double* F, const double* R2, int s, double f) {
    return k(442, I("abij", "efmn"), opnd(dR2(s), "efin", 4, R2), opnd(dFoo(s), "mi", 2, F), f, 0, "Fmi_r2abij");
  }
  int Fae_r2abij(double* F, const double* R2, int s, double f) {
    return k(442, opnd(dR2(s), "afmn", 4, R2), I("abij", "efmn"), opnd(dFvv(s), "ae", 2, F), -f, 0, "Fae_r2abij");
  }
  int Wmbej_r2abij(double* W, const double* R2, int s, double f) {
    return k(444, I("abij", "fenm"), opnd(dR2(s), "bfjn", 4, R2), opnd(dWovvo(s), "embj", 4, W), 0.5 * f, 0,
             "Wmbej_r2abij");
  }
  int Wmnij_r2abij(double* W, const double* R2, int s, double f) {
    return k(444, I("abij", "efmn"), opnd(dR2(s), "efij", 4, R2), opnd(dWoooo(s), "mnij", 4, W), f, 0,
             "Wmnij_r2abij");
  }
  int Fmi_r1ijka(double* F, const double* R1, int s, double f) {
    return k(444, I("ijka", "mnie"), opnd(dR1(s), "en+-", 4, R1), opnd(dFoo(s), "mi+-", 4, F), f, 0, "Fmi_r1ijka");
  }
  int Fae_r1abci(double* F, const double* R1, int s, double f) {
    return k(444, I("abci", "efam"), opnd(dR1(s), "fm+-", 4, R1), opnd(dFvv(s), "ae+-", 4, F), f, 0, "Fae_r1abci");
  }
  // Wkbij_r2ijka.f90 (P(ij) on the output pair)
  int Wkbij_r2ijka(double* W, const double* R2, int s, double f) {
    return k(444, I("ijka", "klic"), opnd(dR2(s), "bcjl", 4, R2), opnd(dWovoo(s), "ijkb", 4, W), f, 1,
             "Wkbij_r2ijka");
  }
  // Wkbij_r2abci.f90: X(ij,bk) = -1/2 r(cd,ij) <cd||bk> into a scratch tensor of
  // layout (o,o,v,o), then W(ij,kb) += X(ij,bk) (isymtrso swap of the second pair
  // and daxpy): the engine's packed-pair rule forbids writing (ij,kb) directly
  int Wkbij_r2abci(double* W, const double* R2, int s, double f) {
    const Desc dx = mk(o, o, v, o, 0, 1, s);
    DevBuf xb;
    const int64_t lx = len(dx);
    int rc = xb.alloc(sizeof(double) * (size_t)std::max<int64_t>(lx, 1));
    if (rc) return rc;
    double* X = (double*)xb.p;
    DPD_HIP_TRY(launch_fill(lx, 0.0, X, st));
    if ((rc = k(444, opnd(dR2(s), "cdij", 4, R2), I("abci", "cdbk"), opnd(dx, "ijbk", 4, X), -f, 0,
                "Wkbij_r2abci")))
      return rc;
    MapArgs a;
    std::memset(&a, 0, sizeof a);
    a.out = make_layout(c->sym, dWovoo(s));
    a.in = make_layout(c->sym, dx);
    a.nterm = 1;
    const int8_t perm[4] = {0, 1, 3, 2};  // W(i,j,k,b) <- X(i,j,b,k)
    for (int d = 0; d < 4; d++) a.term[0].perm[d] = perm[d];
    a.term[0].coef = 1.0;
    a.alpha = 1.0;
    a.beta = 1.0;
    a.nbatch = 1;
    if ((rc = gather(c, a, X, W))) return rc;
    DPD_HIP_TRY(hipStreamSynchronize(st));  // X is freed on return
    return DPD_OK;
  }
  int Wmbej_r1ijka(double* W, const double* R1, int s, double f) {
    return k(424, I("ijka", "nmje"), opnd(dR1(s), "bn", 2, R1), opnd(dWovvo(s), "embj", 4, W), -f, 0,
             "Wmbej_r1ijka");
  }
  int Wmbej_r1abci(double* W, const double* R1, int s, double f) {
    return k(424, I("abci", "febm"), opnd(dR1(s), "fj", 2, R1), opnd(dWovvo(s), "embj", 4, W), f, 0,
             "Wmbej_r1abci");
  }
  int Wmnij_r1ijka(double* W, const double* R1, int s, double f) {
    return k(424, I("ijka", "mnie"), opnd(dR1(s), "ej", 2, R1), opnd(dWoooo(s), "mnij", 4, W), f, 2,
             "Wmnij_r1ijka");
  }
  // ---- grCPSd.SF terms (lib_cc/*.f90) ------------------------------------------
  Desc dWvvvo(int s) const { return mk(v, v, v, o, 0, 1, s); }
  int Wkbij_r1iabj(double* W, const double* R1, int s, double f) {
    return k(424, I("iabj", "ckbj"), opnd(dR1(s), "ci", 2, R1), opnd(dWovoo(s), "ijkb", 4, W), -f, 1,
             "Wkbij_r1iabj");
  }
  int Wabdj_r1abcd(double* W, const double* R1, int s, double f) {  // both in core
    return k(424, I("abcd", "abdc"), opnd(dR1(s), "cj", 2, R1), opnd(dWvvvo(s), "abdj", 4, W), f, 0,
             "Wabdj_r1abcd");
  }
  int Wkbij_r1Uklij(double* W, const double* R1, int s, double f) {
    return k(424, I("ijkl", "klij"), opnd(dR1(s), "bl", 2, R1), opnd(dWovoo(s), "ijkb", 4, W), -f, 0,
             "Wkbij_r1Uklij");
  }
  int Fme_r1abij(double* F, const double* R1, int s, double f) {
    return k(444, I("abij", "efmn"), opnd(dR1(s), "fn+-", 4, R1), opnd(dR1(s), "em+-", 4, F), f, 0, "Fme_r1abij");
  }
  int c1_r2Fme(double* C1, const double* R2, int s, const double* F, int sf, double f) {
    return k(444, opnd(dR2(s), "aeim", 4, R2), opnd(dR1(sf), "em+-", 4, F), opnd(dR1(s), "ai+-", 4, C1), f, 0,
             "c1_r2Fme");
  }
  // effectT2.f90 (p_effectT2 types 1 and 2): T2 += f [t1a(a,i) t1b(b,j)]
  // antisymmetrised in ij (type 1, t1a = t1b) or in ab and ij (type 2)
  int effectT2(double* T2, const double* T1a, const double* T1b, double f, int type) {
    return k(444, opnd(dR1(1), "ai+-", 4, T1a), opnd(dR1(1), "bj+-", 4, T1b), opnd(dR2(1), "abij", 4, T2), f,
             type == 1 ? 2 : 3, "effectT2");
  }
  // ---- vector helpers --------------------------------------------------------
  EorbOff eoff() const {
    EorbOff e;
    const int n = c->sym.nirrep;
    int64_t q = 0;
    for (int h = 0; h < 8; h++) {
      e.offo[h] = q;
      if (h < n) q += o[h];
    }
    for (int h = 0; h < 8; h++) {
      e.offv[h] = q;
      if (h < n) q += v[h];
    }
    return e;
  }
  int denominators(const Desc& d, int mode, double* out) {
    DPD_HIP_TRY(launch_orbital_energy(make_layout(c->sym, d), (const double*)p->eorb.p, eoff(), mode, out, st));
    return DPD_OK;
  }
  // dot products of x with y_0..y_{nv-1} (stride ys), to the host
  int dots(int64_t n, const double* x, const double* y, int64_t ys, int nv, double* host) {
    double* part = (double*)red.p;
    double* out = part + (size_t)kRedBlocks * 64;
    if (nv > 64) return fail(DPD_EINVAL, "too many dot products (internal)");
    DPD_HIP_TRY(launch_multidot(n, x, y, ys, nv, part, out, st));
    DPD_HIP_TRY(hipMemcpyAsync(host, out, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
    DPD_HIP_TRY(hipStreamSynchronize(st));
    return DPD_OK;
  }
  // signed element of largest magnitude (first one: idamax)
  int amax(int64_t n, const double* x, double* val) {
    double* pv = (double*)red.p;
    int64_t* pi = (int64_t*)(pv + (size_t)kRedBlocks * 72 + 64);
    double* ov = pv + (size_t)kRedBlocks * 64 + 64;
    int64_t* oi = pi + kRedBlocks;
    if (n <= 0) {
      *val = 0.0;
      return DPD_OK;
    }
    DPD_HIP_TRY(launch_amax(n, x, pv, pi, ov, oi, st));
    DPD_HIP_TRY(hipMemcpyAsync(val, ov, sizeof(double), hipMemcpyDeviceToHost, st));
    DPD_HIP_TRY(hipStreamSynchronize(st));
    return DPD_OK;
  }
};

// dodiis.f:1-25: coefficients c (summing to 1) from the error matrix B (n x n,
// row-major); the Lagrange row and column carry -1e-4 as in the reference.
// Gaussian elimination with partial pivoting on the (n+1) x (n+1) system.
int diis_coefficients(const std::vector<double>& B, int n, std::vector<double>& c) {
  const int m = n + 1;
  std::vector<double> A((size_t)m * m, 0.0), b(m, 0.0);
  for (int i = 1; i < m; i++) {
    A[i] = -1.0e-4;
    A[(size_t)i * m] = -1.0e-4;
    for (int j = 1; j < m; j++) A[(size_t)i * m + j] = B[(size_t)(i - 1) * n + (j - 1)];
  }
  b[0] = -1.0e-4;
  for (int k = 0; k < m; k++) {
    int piv = k;
    for (int i = k + 1; i < m; i++)
      if (std::fabs(A[(size_t)i * m + k]) > std::fabs(A[(size_t)piv * m + k])) piv = i;
    if (A[(size_t)piv * m + k] == 0.0) return fail(DPD_EINVAL, "DIIS: singular error matrix");
    if (piv != k) {
      for (int j = 0; j < m; j++) std::swap(A[(size_t)k * m + j], A[(size_t)piv * m + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < m; i++) {
      const double f = A[(size_t)i * m + k] / A[(size_t)k * m + k];
      for (int j = k; j < m; j++) A[(size_t)i * m + j] -= f * A[(size_t)k * m + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = m - 1; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < m; j++) s -= A[(size_t)i * m + j] * b[j];
    b[i] = s / A[(size_t)i * m + i];
  }
  c.assign(b.begin() + 1, b.end());
  return DPD_OK;
}

// libsrint/dosodiis2_d.f on device vectors: records of error vectors and of the
// previous vectors (ring of maxdim), the previous extrapolation, the error
// matrix on the host (only the row of the record written this step changes)
struct Diis {
  int maxdim = 10;
  int64_t n = 0;
  int ndimr = 1, idimr = 1;
  DevBuf err, vec, prev, coef, tmp;
  std::vector<double> B;
  int init(int64_t n_, int maxdim_) {
    n = n_;
    maxdim = std::max(1, maxdim_);
    B.assign((size_t)maxdim * maxdim, 0.0);
    const size_t vb = sizeof(double) * (size_t)std::max<int64_t>(n, 1);
    int rc;
    if ((rc = err.alloc(vb * maxdim)) || (rc = vec.alloc(vb * maxdim)) || (rc = prev.alloc(vb)) ||
        (rc = tmp.alloc(vb)) || (rc = coef.alloc(sizeof(double) * maxdim)))
      return rc;
    return DPD_OK;
  }
  double* e(int k) { return (double*)err.p + (size_t)k * n; }
  double* x(int k) { return (double*)vec.p + (size_t)k * n; }
  // x (device, n) in: this iteration's vector, out: the extrapolation;
  // resm: signed largest element of the extrapolated residual (1 at it == 1)
  int step(Work& w, int it, double* xv, double* resm) {
    hipStream_t st = w.st;
    if (it == 1) {
      ndimr = idimr = 1;
      DPD_HIP_TRY(launch_copy(n, xv, (double*)prev.p, st));
      *resm = 1.0;
      return DPD_OK;
    }
    const int kk = idimr - 1;
    DPD_HIP_TRY(launch_copy(n, (const double*)prev.p, x(kk), st));
    DPD_HIP_TRY(launch_copy(n, xv, e(kk), st));
    DPD_HIP_TRY(launch_axpy(n, -1.0, (const double*)prev.p, e(kk), st));
    const int nd = ndimr;
    std::vector<double> d(nd);
    int rc = w.dots(n, e(kk), e(0), n, nd, d.data());
    if (rc) return rc;
    for (int j = 0; j < nd; j++) B[(size_t)kk * maxdim + j] = B[(size_t)j * maxdim + kk] = d[j];
    std::vector<double> Bn((size_t)nd * nd), c;
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++) Bn[(size_t)i * nd + j] = B[(size_t)i * maxdim + j];
    if ((rc = diis_coefficients(Bn, nd, c))) return rc;
    DPD_HIP_TRY(hipMemcpyAsync(coef.p, c.data(), sizeof(double) * nd, hipMemcpyHostToDevice, st));
    DPD_HIP_TRY(launch_lincomb(n, e(0), n, nd, (const double*)coef.p, nd, 1, xv, st));
    if ((rc = w.amax(n, xv, resm))) return rc;  // (synchronises: coef is consumed)
    DPD_HIP_TRY(launch_lincomb(n, x(0), n, nd, (const double*)coef.p, nd, 1, (double*)tmp.p, st));
    DPD_HIP_TRY(launch_axpy(n, 1.0, (const double*)tmp.p, xv, st));
    DPD_HIP_TRY(launch_copy(n, xv, (double*)prev.p, st));
    ndimr = std::min(maxdim, nd + 1);
    if (++idimr > maxdim) idimr = 1;
    return DPD_OK;
  }
};

void defaults(dpd_cpsd_opts* o) {
  if (o->diis_dim <= 0) o->diis_dim = 10;
  if (o->diis_cyc <= 0) o->diis_cyc = 50;
  if (!(o->diis_max_err > 0.0)) o->diis_max_err = 1.0e-7;
}

int setup(dpd_problem* P, Work& w) {
  if (!P) return fail(DPD_EINVAL, "null problem");
  if (P->method != 2) return fail(DPD_EINVAL, "perturbation drivers need a problem of method 2 (every integral block)");
  if (P->stale) return fail(DPD_EINVAL, "problem inputs changed: call dpd_problem_refresh first");
  w.p = P;
  w.c = P->ctx;
  w.st = P->ctx->stream;
  w.v = P->popv;
  w.o = P->popo;
  w.z = P->zorb;
  DPD_HIP_TRY(hipSetDevice(w.c->device));
  return w.red.alloc(sizeof(double) * ((size_t)kRedBlocks * 72 + 128) + sizeof(int64_t) * ((size_t)kRedBlocks + 8));
}

// CPSd.SF/run2.f90:43-71: -e(a,i) T1 = <a,i|[V,T1]|0> + C1 by Jacobi steps with DIIS
int t1_diis(Work& w, const dpd_cpsd_opts* o, const double* C1, const double* den1, double* T1, double* R1,
            int64_t l1, int64_t* iters) {
  Diis d;
  int rc = d.init(l1, (int)o->diis_dim);
  if (rc) return rc;
  const Operand W = w.I("iabj", "fnai");
  DPD_HIP_TRY(launch_copy(l1, C1, R1, w.st));
  int it = 0;
  while (true) {
    it++;
    DPD_HIP_TRY(launch_div(l1, R1, den1, T1, w.st));  // denomT1: r / (e_i - e_a)
    double err = 0.0;
    if ((rc = d.step(w, it, T1, &err))) return rc;
    if (it > 1 && std::fabs(err) < o->diis_max_err) break;
    if (it > o->diis_cyc) break;
    DPD_HIP_TRY(launch_copy(l1, C1, R1, w.st));
    if ((rc = w.c1_r1iabj(R1, T1, 1, W, 1.0))) return rc;
  }
  *iters = it;
  return DPD_OK;
}

// ground state; T (device) receives T^(1..3) as (T1 | T2) per order
int cpsd_run(Work& w, dpd_cpsd_opts* o, double* T) {
  const int64_t l1 = w.len(w.dR1(1)), l2 = w.len(w.dR2(1)), lt = l1 + l2;
  DevBuf den1, den2, C1, C2, R1, tmp, ws;
  int rc;
  if ((rc = den1.alloc(sizeof(double) * std::max<int64_t>(l1, 1))) ||
      (rc = den2.alloc(sizeof(double) * std::max<int64_t>(l2, 1))) ||
      (rc = C1.alloc(sizeof(double) * std::max<int64_t>(l1, 1))) ||
      (rc = C2.alloc(sizeof(double) * std::max<int64_t>(l2, 1))) ||
      (rc = R1.alloc(sizeof(double) * std::max<int64_t>(l1, 1))))
    return rc;
  double* d1 = (double*)den1.p;
  double* d2 = (double*)den2.p;
  if ((rc = w.denominators(w.dR1(1), 4, d1)) || (rc = w.denominators(w.dR2(1), 3, d2))) return rc;
  auto t1 = [&](int n) { return T + (size_t)(n - 1) * lt; };
  auto t2 = [&](int n) { return T + (size_t)(n - 1) * lt + l1; };
  const double* abij = (const double*)w.p->abij.p;
  // run1.f90: T1^(1) = 0, T2^(1) = <ab||ij> / (e_i + e_j - e_a - e_b)
  DPD_HIP_TRY(launch_fill(l1, 0.0, t1(1), w.st));
  DPD_HIP_TRY(launch_div(l2, abij, d2, t2(1), w.st));
  if ((rc = w.dots(l2, t2(1), abij, 0, 1, &o->energy[0]))) return rc;  // calEcorr.f90
  // calC2_VTT.f90 intermediates (symmetry 1)
  const int64_t lfoo = w.len(w.dFoo(1)), lfvv = w.len(w.dFvv(1)), lw = w.len(w.dWovvo(1)), lkl = w.len(w.dWoooo(1));
  if ((rc = tmp.alloc(sizeof(double) * (size_t)(lfoo + lfvv + lw + lkl + 4)))) return rc;
  double* Fli = (double*)tmp.p;
  double* Fad = Fli + lfoo + 1;
  double* Wk = Fad + lfvv + 1;
  double* Wkl = Wk + lw + 1;
  for (int n = 2; n <= 3; n++) {
    // calC1.f90: C1^(n) = <a,i|[V,T2^(n-1)]|0>
    DPD_HIP_TRY(launch_fill(l1, 0.0, (double*)C1.p, w.st));
    if ((rc = w.c1_vr2((double*)C1.p, t2(n - 1), 1, 1.0))) return rc;
    // calC2.f90 / calC2_VR.f90: C2^(n) = <ab,ij|[V,T^(n-1)]|0>
    DPD_HIP_TRY(launch_fill(l2, 0.0, (double*)C2.p, w.st));
    if (n > 2 && (rc = w.c2_vr1((double*)C2.p, t1(n - 1), 1, 1.0))) return rc;
    if ((rc = w.c2_vr2((double*)C2.p, t2(n - 1), 1, 1.0))) return rc;
    if (n == 3) {
      // calC2_VTT.f90 with T^(1) (T1^(1) = 0)
      DPD_HIP_TRY(launch_fill(lfoo + lfvv + lw + lkl + 4, 0.0, Fli, w.st));
      if ((rc = w.Fmi_r2abij(Fli, t2(1), 1, 1.0)) || (rc = w.c2_r2Fmi((double*)C2.p, 1, t2(1), 1, Fli, 1, 1.0)) ||
          (rc = w.Fae_r2abij(Fad, t2(1), 1, 1.0)) || (rc = w.c2_r2Fae((double*)C2.p, 1, t2(1), 1, Fad, 1, 1.0)) ||
          (rc = w.Wmbej_r2abij(Wk, t2(1), 1, 1.0)) ||
          (rc = w.c2_r2Wmbej((double*)C2.p, 1, t2(1), 1, opnd(w.dWovvo(1), "embj", 4, Wk), 1.0)) ||
          (rc = w.Wmnij_r2abij(Wkl, t2(1), 1, 0.5)) ||
          (rc = w.c2_r2Wmnij((double*)C2.p, 1, t2(1), 1, opnd(w.dWoooo(1), "mnij", 4, Wkl), 2.0)))
        return rc;
    }
    // denomT2: T2^(n) = C2^(n) / (e_i + e_j - e_a - e_b)
    DPD_HIP_TRY(launch_div(l2, (const double*)C2.p, d2, t2(n), w.st));
    if ((rc = t1_diis(w, o, (const double*)C1.p, d1, t1(n), (double*)R1.p, l1, &o->iterations[n - 2]))) return rc;
    if ((rc = w.dots(l2, t2(n), abij, 0, 1, &o->energy[n - 1]))) return rc;
  }
  return DPD_OK;
}

// grCPSd.SF/run.f90 (run1-3.f90, solver.f90, calA0R1.f90, calA0R2.f90,
// calA1R1.f90 (calA2R1), calA1R2.f90, calC1.f90, calC2.f90, calL1_VTT.f90,
// calL2_VTT.f90, calEcorr.f90): the ground state with the CIS(Doo) Jacobian as
// zeroth order.  For n = 1, 2, 3:  T = (C + A0' T) / D  (A0' = A0 without its
// orbital-energy diagonal D) by Jacobi steps with DIIS on (T1 | T2);
// C^(1) = (0, <ab||ij>), C^(2) = (0, [V,T2^(1)]), C^(3) = (-D1 T1^(1) + L1, [V,T2^(2)] + L2)
// with D1 the MP2 dressing of A0's singles block and L1/L2 the [[V,T^(1)],T^(1)]
// terms; E^(n+1) = <abij| T2^(n) + [T1 T1]^(n)>.
struct GrInt {
  DevBuf buf;
  double *mp2, *fli, *fad, *wk, *wdiff, *gli, *gad, *gld, *tmp, *t2e;
};

int gr_setup(Work& w, GrInt& X) {
  const int64_t l2 = w.len(w.dR2(1)), lfoo = w.len(w.dFoo(1)), lfvv = w.len(w.dFvv(1)), lw = w.len(w.dWovvo(1));
  const int64_t l1 = w.len(w.dR1(1));
  const int64_t ltmp = std::max({w.len(w.dWovoo(1)), w.len(w.dWvvvo(1)), lfoo, lfvv, lw, w.len(w.dWoooo(1))});
  const int64_t tot = l2 + lfoo + lfvv + 2 * lw + lfoo + lfvv + l1 + ltmp + l2 + 16;
  int rc = X.buf.alloc(sizeof(double) * (size_t)tot);
  if (rc) return rc;
  double* q = (double*)X.buf.p;
  auto take = [&](int64_t n) {
    double* r = q;
    q += n + 1;
    return r;
  };
  X.mp2 = take(l2);
  X.fli = take(lfoo);
  X.fad = take(lfvv);
  X.wk = take(lw);
  X.wdiff = take(lw);
  X.gli = take(lfoo);
  X.gad = take(lfvv);
  X.gld = take(l1);
  X.tmp = take(ltmp);
  X.t2e = take(l2);
  return DPD_OK;
}

// C1 += -D1 T1 + L1 (calA2R1 in calA1R1.f90, calL1_VTT.f90), T = T^(1)
int gr_c1_order3(Work& w, GrInt& X, double* C1, const double* T1, const double* T2) {
  const int64_t lw = w.len(w.dWovvo(1)), lfoo = w.len(w.dFoo(1)), lfvv = w.len(w.dFvv(1)), l1 = w.len(w.dR1(1));
  const double* iabj = (const double*)w.p->iabj.p;
  int rc;
  // calA2R1: Wkaci - iabj (the MP2 part), c1_r1iabj x -1, c1_r1Fmi / c1_r1Fae x +1
  DPD_HIP_TRY(launch_copy(lw, X.wk, X.wdiff, w.st));
  DPD_HIP_TRY(launch_axpy(lw, -1.0, iabj, X.wdiff, w.st));
  if ((rc = w.c1_r1iabj(C1, T1, 1, opnd(w.dWovvo(1), "fnai", 4, X.wdiff), -1.0)) ||
      (rc = w.c1_r1Fmi(C1, T1, 1, X.fli, 1, 1.0)) || (rc = w.c1_r1Fae(C1, T1, 1, X.fad, 1, 1.0)))
    return rc;
  // calL1_VTT: Gli, Gad, Gld from T^(1), then their singles terms
  DPD_HIP_TRY(launch_fill(lfoo, 0.0, X.gli, w.st));
  DPD_HIP_TRY(launch_fill(lfvv, 0.0, X.gad, w.st));
  DPD_HIP_TRY(launch_fill(l1, 0.0, X.gld, w.st));
  if ((rc = w.Fmi_r1ijka(X.gli, T1, 1, -1.0)) || (rc = w.Fmi_r2abij(X.gli, T2, 1, -1.0)) ||
      (rc = w.Fae_r1abci(X.gad, T1, 1, -1.0)) || (rc = w.Fae_r2abij(X.gad, T2, 1, -1.0)) ||
      (rc = w.Fme_r1abij(X.gld, T1, 1, 1.0)) || (rc = w.c1_r1Fmi(C1, T1, 1, X.gli, 1, -1.0)) ||
      (rc = w.c1_r1Fae(C1, T1, 1, X.gad, 1, -1.0)) || (rc = w.c1_r2Fme(C1, T2, 1, X.gld, 1, 1.0)))
    return rc;
  return DPD_OK;
}

// C2 += L2 (calL2_VTT.f90), T = T^(1)
int gr_l2_vtt(Work& w, GrInt& X, double* C2, const double* T1, const double* T2) {
  int rc;
  double* t = X.tmp;
  auto clear = [&](const Desc& d) { return launch_fill(w.len(d), 0.0, t, w.st); };
  DPD_HIP_TRY(clear(w.dWovoo(1)));
  if ((rc = w.Wkbij_r1iabj(t, T1, 1, 1.0)) ||
      (rc = w.c2_r1ijka(C2, T1, 1, opnd(w.dWovoo(1), "ijmb", 4, t), -1.0)))
    return rc;
  DPD_HIP_TRY(clear(w.dWvvvo(1)));
  if ((rc = w.Wabdj_r1abcd(t, T1, 1, 1.0)) ||
      (rc = w.k(424, opnd(w.dWvvvo(1), "abej", 4, t), opnd(w.dR1(1), "ei", 2, T1), opnd(w.dR2(1), "abij", 4, C2),
                0.5, 2, "c2_r1abci")))
    return rc;
  DPD_HIP_TRY(clear(w.dWovoo(1)));
  if ((rc = w.Wkbij_r1Uklij(t, T1, 1, 1.0)) ||
      (rc = w.c2_r1ijka(C2, T1, 1, opnd(w.dWovoo(1), "ijmb", 4, t), 0.5)))
    return rc;
  DPD_HIP_TRY(clear(w.dFoo(1)));
  if ((rc = w.Fmi_r1ijka(t, T1, 1, 1.0)) || (rc = w.Fmi_r2abij(t, T2, 1, 1.0)) ||
      (rc = w.c2_r2Fmi(C2, 1, T2, 1, t, 1, 1.0)))
    return rc;
  DPD_HIP_TRY(clear(w.dFvv(1)));
  if ((rc = w.Fae_r1abci(t, T1, 1, 1.0)) || (rc = w.Fae_r2abij(t, T2, 1, 1.0)) ||
      (rc = w.c2_r2Fae(C2, 1, T2, 1, t, 1, 1.0)))
    return rc;
  DPD_HIP_TRY(clear(w.dWovvo(1)));
  if ((rc = w.Wmbej_r1ijka(t, T1, 1, 1.0)) || (rc = w.Wmbej_r1abci(t, T1, 1, 1.0)) ||
      (rc = w.Wmbej_r2abij(t, T2, 1, 1.0)) ||
      (rc = w.c2_r2Wmbej(C2, 1, T2, 1, opnd(w.dWovvo(1), "embj", 4, t), 1.0)))
    return rc;
  DPD_HIP_TRY(clear(w.dWovoo(1)));
  if ((rc = w.Wkbij_r2abci(t, T2, 1, 1.0)) ||
      (rc = w.c2_r1ijka(C2, T1, 1, opnd(w.dWovoo(1), "ijmb", 4, t), 1.0)))
    return rc;
  DPD_HIP_TRY(clear(w.dWoooo(1)));
  if ((rc = w.Wmnij_r1ijka(t, T1, 1, 0.5)) || (rc = w.Wmnij_r2abij(t, T2, 1, 0.5)) ||
      (rc = w.c2_r2Wmnij(C2, 1, T2, 1, opnd(w.dWoooo(1), "mnij", 4, t), 2.0)))
    return rc;
  return DPD_OK;
}

// T (device): T^(1..3) as (T1 | T2) per order; it3 (host, 3): DIIS iterations
int grcpsd_run(Work& w, dpd_cpsd_opts* o, double* T, int64_t* it3) {
  if (w.p->abcd_host) return fail(DPD_EINVAL, "grCPS(D) needs <ab||ef> on the device (Wabdj_r1abcd)");
  const int64_t l1 = w.len(w.dR1(1)), l2 = w.len(w.dR2(1)), lt = l1 + l2, lw = w.len(w.dWovvo(1));
  DevBuf den1, den2, Cb, Rb;
  GrInt X;
  int rc;
  if ((rc = den1.alloc(sizeof(double) * std::max<int64_t>(l1, 1))) ||
      (rc = den2.alloc(sizeof(double) * std::max<int64_t>(l2, 1))) ||
      (rc = Cb.alloc(sizeof(double) * (size_t)(lt + 1))) || (rc = Rb.alloc(sizeof(double) * (size_t)(lt + 1))) ||
      (rc = gr_setup(w, X)))
    return rc;
  double* d1 = (double*)den1.p;
  double* d2 = (double*)den2.p;
  if ((rc = w.denominators(w.dR1(1), 4, d1)) || (rc = w.denominators(w.dR2(1), 3, d2))) return rc;
  double *C1 = (double*)Cb.p, *C2 = C1 + l1, *R1 = (double*)Rb.p, *R2 = R1 + l1;
  auto t1 = [&](int n) { return T + (size_t)(n - 1) * lt; };
  auto t2 = [&](int n) { return T + (size_t)(n - 1) * lt + l1; };
  const double* abij = (const double*)w.p->abij.p;
  // run.f90:29-47: MP2t, Fli = -Fmi(MP2t), Fad = -Fae(MP2t), Wkaci = iabj + 2 Wmbej(MP2t)
  DPD_HIP_TRY(launch_div(l2, abij, d2, X.mp2, w.st));
  DPD_HIP_TRY(launch_fill(w.len(w.dFoo(1)), 0.0, X.fli, w.st));
  DPD_HIP_TRY(launch_fill(w.len(w.dFvv(1)), 0.0, X.fad, w.st));
  DPD_HIP_TRY(launch_copy(lw, (const double*)w.p->iabj.p, X.wk, w.st));
  if ((rc = w.Fmi_r2abij(X.fli, X.mp2, 1, -1.0)) || (rc = w.Fae_r2abij(X.fad, X.mp2, 1, -1.0)) ||
      (rc = w.Wmbej_r2abij(X.wk, X.mp2, 1, 2.0)))
    return rc;
  const Operand Wk = opnd(w.dWovvo(1), "fnai", 4, X.wk);
  for (int n = 1; n <= 3; n++) {
    DPD_HIP_TRY(launch_fill(lt, 0.0, C1, w.st));
    if (n == 1) {
      DPD_HIP_TRY(launch_copy(l2, abij, C2, w.st));  // run1.f90:18-19
    } else {
      if (n == 3 && (rc = gr_c1_order3(w, X, C1, t1(1), t2(1)))) return rc;
      if ((rc = w.c2_vr2(C2, t2(n - 1), 1, 1.0))) return rc;  // calA1R2: ring + hh + pp ladder
      if (n == 3 && (rc = gr_l2_vtt(w, X, C2, t1(1), t2(1)))) return rc;
    }
    // solver.f90
    Diis d;
    if ((rc = d.init(lt, (int)o->diis_dim))) return rc;
    DPD_HIP_TRY(launch_copy(lt, C1, R1, w.st));
    int it = 0;
    double E = 0.0;
    while (true) {
      it++;
      DPD_HIP_TRY(launch_div(l1, R1, d1, t1(n), w.st));  // denomT1
      DPD_HIP_TRY(launch_div(l2, R2, d2, t2(n), w.st));  // denomT2
      double err = 0.0;
      if ((rc = d.step(w, it, t1(n), &err))) return rc;
      // calEcorr.f90: E = <abij| T2 + [T1 T1]>
      DPD_HIP_TRY(launch_copy(l2, t2(n), X.t2e, w.st));
      if (n == 2 && (rc = w.effectT2(X.t2e, t1(1), t1(1), 1.0, 1))) return rc;
      if (n == 3 && (rc = w.effectT2(X.t2e, t1(2), t1(1), 1.0, 2))) return rc;
      if ((rc = w.dots(l2, X.t2e, abij, 0, 1, &E))) return rc;
      if (it > 1 && std::fabs(err) < o->diis_max_err) break;
      if (it > o->diis_cyc) break;
      // R = C + A0' T (calA0R1.f90, calA0R2.f90)
      DPD_HIP_TRY(launch_copy(lt, C1, R1, w.st));
      if ((rc = w.c1_r1iabj(R1, t1(n), 1, Wk, 1.0)) || (rc = w.c1_r1Fmi(R1, t1(n), 1, X.fli, 1, -1.0)) ||
          (rc = w.c1_r1Fae(R1, t1(n), 1, X.fad, 1, -1.0)) || (rc = w.c1_vr2(R1, t2(n), 1, 1.0)) ||
          (rc = w.c2_vr1(R2, t1(n), 1, 1.0)))
        return rc;
    }
    o->energy[n - 1] = E;
    if (n >= 2) o->iterations[n - 2] = it;
    if (it3) it3[n - 1] = it;
  }
  return DPD_OK;
}

// EE-CPSd.SF intermediates (calJ.f90), symmetry 1
struct Jint {
  DevBuf buf;
  double *j2oo, *j2vv, *j2ovvo, *j2ovoo, *j3oo, *j3vv, *j3ovvo;
};

int build_j(Work& w, const double* T, Jint& J) {
  const int64_t l1 = w.len(w.dR1(1)), lt = l1 + w.len(w.dR2(1));
  const int64_t loo = w.len(w.dFoo(1)), lvv = w.len(w.dFvv(1)), lw = w.len(w.dWovvo(1)), lwo = w.len(w.dWovoo(1));
  const int64_t tot = 2 * loo + 2 * lvv + 2 * lw + lwo + 8;
  int rc = J.buf.alloc(sizeof(double) * (size_t)tot);
  if (rc) return rc;
  double* q = (double*)J.buf.p;
  J.j2oo = q;
  J.j2vv = J.j2oo + loo + 1;
  J.j2ovvo = J.j2vv + lvv + 1;
  J.j2ovoo = J.j2ovvo + lw + 1;
  J.j3oo = J.j2ovoo + lwo + 1;
  J.j3vv = J.j3oo + loo + 1;
  J.j3ovvo = J.j3vv + lvv + 1;
  DPD_HIP_TRY(launch_fill(tot, 0.0, q, w.st));
  const double* T2a = T + l1;            // T2^(1)
  const double* T1b = T + lt;            // T1^(2)
  const double* T2b = T + lt + l1;       // T2^(2)
  // calJ2 (T1^(1) = 0)
  if ((rc = w.Fmi_r2abij(J.j2oo, T2a, 1, 1.0)) || (rc = w.Fae_r2abij(J.j2vv, T2a, 1, 1.0)) ||
      (rc = w.Wkbij_r2ijka(J.j2ovoo, T2a, 1, 1.0)) || (rc = w.Wkbij_r2abci(J.j2ovoo, T2a, 1, 1.0)) ||
      (rc = w.Wmbej_r2abij(J.j2ovvo, T2a, 1, 2.0)))
    return rc;
  // calJ3
  if ((rc = w.Fmi_r1ijka(J.j3oo, T1b, 1, 1.0)) || (rc = w.Fmi_r2abij(J.j3oo, T2b, 1, 1.0)) ||
      (rc = w.Fae_r1abci(J.j3vv, T1b, 1, 1.0)) || (rc = w.Fae_r2abij(J.j3vv, T2b, 1, 1.0)) ||
      (rc = w.Wmbej_r1ijka(J.j3ovvo, T1b, 1, 1.0)) || (rc = w.Wmbej_r1abci(J.j3ovvo, T1b, 1, 1.0)) ||
      (rc = w.Wmbej_r2abij(J.j3ovvo, T2b, 1, 2.0)))
    return rc;
  return DPD_OK;
}

// one CIS root: w[0..2] = w^(1), w^(2), w^(3)
int ee_root(Work& w, const dpd_cpsd_opts* o, const double* T, const Jint& J, double w0, const double* L, double* wn,
            int64_t* iters, std::deque<DevBuf>& bufs) {
  const int s = w.p->symR;
  const int64_t l1 = w.len(w.dR1(s)), l2 = w.len(w.dR2(s));
  const int64_t l1g = w.len(w.dR1(1));
  const double* T2a = T + l1g;  // T2^(1)
  // R^(0..2), C, Ctilde, OED2, scratch of the VTR intermediates
  double* R1[3];
  double* R2[3];
  size_t k = 0;
  auto buf = [&](int64_t n) -> double* {
    if (bufs.size() <= k) bufs.emplace_back();
    DevBuf& b = bufs[k++];
    if (b.bytes < sizeof(double) * (size_t)std::max<int64_t>(n, 1) && b.alloc(sizeof(double) * std::max<int64_t>(n, 1)))
      return nullptr;
    return (double*)b.p;
  };
  for (int i = 0; i < 3; i++) {
    R1[i] = buf(l1);
    R2[i] = buf(l2);
  }
  double *C1 = buf(l1), *C2 = buf(l2), *oed2 = buf(l2), *den1 = buf(l1), *pR1 = buf(l1);
  const int64_t lkl = w.len(w.dWoooo(s)), lw = w.len(w.dWovvo(s)), loo = w.len(w.dFoo(s)), lvv = w.len(w.dFvv(s));
  double* sc = buf(lkl + lw + loo + lvv + 4);
  for (int i = 0; i < 3; i++)
    if (!R1[i] || !R2[i]) return fail(DPD_ENOMEM, "EE-CPS(D): device allocation failed");
  if (!C1 || !C2 || !oed2 || !den1 || !pR1 || !sc) return fail(DPD_ENOMEM, "EE-CPS(D): device allocation failed");
  int rc;
  if ((rc = w.denominators(w.dR2(s), 2, oed2)) || (rc = w.denominators(w.dR1(s), 4, den1))) return rc;
  // run.f90:57-60: R^(0) = the CIS vector, renormalised when |R| != 1 (the
  // reference scales all but its last element); L stays as given
  double nn = 0.0;
  if ((rc = w.dots(l1, L, L, 0, 1, &nn))) return rc;
  const double nrm = std::sqrt(nn);
  DPD_HIP_TRY(launch_copy(l1, L, R1[0], w.st));
  if (std::fabs(nrm - 1.0) > 1.0e-10) DPD_HIP_TRY(launch_scal(l1 - 1, 1.0 / nrm, R1[0], w.st));
  DPD_HIP_TRY(launch_fill(l2, 0.0, R2[0], w.st));
  const Operand iabjW = w.I("iabj", "fnai");
  // calE_C.f90: C^(n) = sum_k J^(k) R^(n-k)
  auto calE_C = [&](int order) -> int {
    DPD_HIP_TRY(launch_fill(l1, 0.0, C1, w.st));
    DPD_HIP_TRY(launch_fill(l2, 0.0, C2, w.st));
    const double* r1 = R1[order - 1];
    const double* r2 = R2[order - 1];
    int r;
    if (order != 1 && (r = w.c1_vr2(C1, r2, s, 1.0))) return r;  // calC1_VR2 (R2^(0) = 0)
    if (order != 3) {
      if (order != 2 && (r = w.c2_vr1(C2, r1, s, 1.0))) return r;  // calC2_VR1 (R1^(1) = 0)
      if (order != 1 && (r = w.c2_vr2(C2, r2, s, 1.0))) return r;  // calC2_VR2
    }
    if (order == 1) return DPD_OK;
    if (order != 3) {
      const double* q1 = R1[order - 2];
      // calC_J2R
      if ((r = w.c2_r1ijka(C2, q1, s, opnd(w.dWovoo(1), "ijmb", 4, J.j2ovoo), 1.0)) ||
          (r = w.c1_r1iabj(C1, q1, s, opnd(w.dWovvo(1), "fnai", 4, J.j2ovvo), 1.0)) ||
          (r = w.c1_r1Fmi(C1, q1, s, J.j2oo, 1, 1.0)) || (r = w.c1_r1Fae(C1, q1, s, J.j2vv, 1, 1.0)))
        return r;
      // calC_VTR with T^(1) (T1^(1) = 0, R2^(0) = 0)
      double* Wkl = sc;
      double* Wb = Wkl + lkl + 1;
      double* G = Wb + lw + 1;
      double* Gad = G + loo + 1;
      DPD_HIP_TRY(launch_fill(lkl + lw + loo + lvv + 4, 0.0, sc, w.st));
      if ((r = w.Wmnij_r1ijka(Wkl, q1, s, 1.0)) ||
          (r = w.c2_r2Wmnij(C2, s, T2a, 1, opnd(w.dWoooo(s), "mnij", 4, Wkl), 1.0)) ||
          (r = w.Wmbej_r1abci(Wb, q1, s, 1.0)) ||
          (r = w.c2_r2Wmbej(C2, s, T2a, 1, opnd(w.dWovvo(s), "embj", 4, Wb), 1.0)) ||
          (r = w.Fmi_r1ijka(G, q1, s, 1.0)) || (r = w.c2_r2Fmi(C2, s, T2a, 1, G, s, 1.0)) ||
          (r = w.Fae_r1abci(Gad, q1, s, 1.0)) || (r = w.c2_r2Fae(C2, s, T2a, 1, Gad, s, 1.0)))
        return r;
    }
    if (order == 2) return DPD_OK;
    const double* q1 = R1[order - 3];
    // calC_J3R
    if ((r = w.c1_r1iabj(C1, q1, s, opnd(w.dWovvo(1), "fnai", 4, J.j3ovvo), 1.0)) ||
        (r = w.c1_r1Fmi(C1, q1, s, J.j3oo, 1, 1.0)) || (r = w.c1_r1Fae(C1, q1, s, J.j3vv, 1, 1.0)))
      return r;
    return DPD_OK;
  };
  // orthogonalR2L.f90: R -= (<L|R> / |L|) L
  auto orthogonal = [&](double* R) -> int {
    double d[2];
    int r = w.dots(l1, L, L, 0, 1, &d[0]);
    if (!r) r = w.dots(l1, L, R, 0, 1, &d[1]);
    if (r) return r;
    DPD_HIP_TRY(launch_axpy(l1, -d[1] / std::sqrt(d[0]), L, R, w.st));
    return DPD_OK;
  };
  // solver.f90: (J_CIS - w0) R1 = Ct1 (Ct1 in C1, overwritten with -Ct1)
  auto solver = [&](double* R1out, int64_t* its) -> int {
    Diis d;
    int r = d.init(l1, (int)o->diis_dim);
    if (r) return r;
    DPD_HIP_TRY(launch_scal(l1, -1.0, C1, w.st));
    DPD_HIP_TRY(launch_copy(l1, C1, pR1, w.st));
    int it = 0;
    while (true) {
      it++;
      DPD_HIP_TRY(launch_div(l1, pR1, den1, R1out, w.st));  // denomT1
      if ((r = orthogonal(R1out))) return r;
      double err = 0.0;
      if ((r = d.step(w, it, R1out, &err))) return r;
      if ((r = orthogonal(R1out))) return r;
      if (it > 1 && std::fabs(err) < o->diis_max_err) break;
      if (it > o->diis_cyc) break;
      DPD_HIP_TRY(launch_copy(l1, C1, pR1, w.st));
      DPD_HIP_TRY(launch_axpy(l1, -w0, R1out, pR1, w.st));
      if ((r = w.c1_r1iabj(pR1, R1out, s, iabjW, 1.0))) return r;
    }
    *its = it;
    return DPD_OK;
  };
  *iters = 0;
  for (int order = 1; order <= 3; order++) {
    if (order != 1) {
      // calV_C: Ctilde^(n-1) = sum_{k=2}^{n-1} w^(k) R^(n-1-k) - C^(n-1) (in place in C)
      const int n1 = order - 1;
      DPD_HIP_TRY(launch_scal(l1, -1.0, C1, w.st));
      DPD_HIP_TRY(launch_scal(l2, -1.0, C2, w.st));
      for (int kk = 2; kk <= n1; kk++) {
        DPD_HIP_TRY(launch_axpy(l1, wn[kk - 1], R1[n1 - kk], C1, w.st));
        DPD_HIP_TRY(launch_axpy(l2, wn[kk - 1], R2[n1 - kk], C2, w.st));
      }
      // R2^(n-1) = Ctilde2 / (e(ab,ij) - w0)
      DPD_HIP_TRY(launch_div_shift(l2, C2, oed2, w0, R2[order - 1], w.st));
      if (order != 2) {
        if ((rc = solver(R1[order - 1], iters))) return rc;
      } else {
        DPD_HIP_TRY(launch_fill(l1, 0.0, R1[order - 1], w.st));  // R1^(1) = 0
      }
    }
    if ((rc = calE_C(order))) return rc;
    if ((rc = w.dots(l1, L, C1, 0, 1, &wn[order - 1]))) return rc;
  }
  return DPD_OK;
}

}  // namespace
}  // namespace dpd

using namespace dpd;

extern "C" int dpd_problem_cpsd(dpd_problem* P, dpd_cpsd_opts* o, double* T) {
  if (!o) return fail(DPD_EINVAL, "null options");
  Work w;
  int rc = setup(P, w);
  if (rc) return rc;
  defaults(o);
  const int64_t lt = w.len(w.dR1(1)) + w.len(w.dR2(1));
  DevBuf own;
  if (!T) {
    if ((rc = own.alloc(sizeof(double) * (size_t)(3 * lt)))) return rc;
    T = (double*)own.p;
  }
  rc = cpsd_run(w, o, T);
  if (rc) return rc;
  DPD_HIP_TRY(hipStreamSynchronize(w.st));
  return DPD_OK;
}

extern "C" int dpd_problem_grcpsd(dpd_problem* P, dpd_cpsd_opts* o, double* T, int64_t* iterations) {
  if (!o) return fail(DPD_EINVAL, "null options");
  Work w;
  int rc = setup(P, w);
  if (rc) return rc;
  defaults(o);
  const int64_t lt = w.len(w.dR1(1)) + w.len(w.dR2(1));
  DevBuf own;
  if (!T) {
    if ((rc = own.alloc(sizeof(double) * (size_t)(3 * lt)))) return rc;
    T = (double*)own.p;
  }
  rc = grcpsd_run(w, o, T, iterations);
  if (rc) return rc;
  DPD_HIP_TRY(hipStreamSynchronize(w.st));
  return DPD_OK;
}

extern "C" int64_t dpd_problem_cpsd_len(const dpd_problem* P) {
  if (!P) return 0;
  const Sym& S = P->ctx->sym;
  return S.irpdso(1, P->popv, P->popo, 1) + S.idsymsoc(1, P->popv, P->popv, 0, P->popo, P->popo, 0);
}

extern "C" int dpd_problem_eecpsd(dpd_problem* P, dpd_cpsd_opts* o, const double* T, int64_t nroot,
                                  const double* w0, const double* L, double* w, int64_t* iterations) {
  if (!o || !T || !w0 || !L || !w || nroot < 1) return fail(DPD_EINVAL, "dpd_problem_eecpsd: bad arguments");
  Work w_;
  int rc = setup(P, w_);
  if (rc) return rc;
  defaults(o);
  Jint J;
  if ((rc = build_j(w_, T, J))) return rc;
  std::deque<DevBuf> bufs;
  const int64_t l1 = w_.len(w_.dR1(P->symR));
  for (int64_t r = 0; r < nroot; r++) {
    int64_t its = 0;
    if ((rc = ee_root(w_, o, T, J, w0[r], L + (size_t)r * l1, w + 3 * r, &its, bufs))) return rc;
    if (iterations) iterations[r] = its;
  }
  DPD_HIP_TRY(hipStreamSynchronize(w_.st));
  return DPD_OK;
}
