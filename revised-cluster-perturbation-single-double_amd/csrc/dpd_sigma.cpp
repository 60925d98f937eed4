// dpd_sigma.cpp -- device-resident problem setup and the batched sigma builders.
//
// Operators (J R for a batch of R = (R1(a,i), R2(a<b,i<j)) vectors):
//   method 0, SF-CIS  (EE-CIS.SF/calAR.f90:6-34):
//       c(a,i) = e(a,i) r(a,i) + <fn||ai> r(f,n)                 [c1_r1iabj]
//   method 1, CIS(Doo) (CISdoo.SD.SF/calAR.f90:9-79):
//       c(a,i)  = Fac r(c,i) - Fki r(a,k) + Wkaci r(c,k)
//               + 1/2 <fe||am> r(fe,im) - 1/2 r(ae,mn) <mn||ie>
//       c(ab,ij)= e(ab,ij) r(ab,ij) + P(ij) <ab||ej> r(e,i) - P(ab) <ij||mb> r(a,m)
//     with the MP2 intermediates of CISdoo.SD.SF/run.f90:31-49
//       t(ab,ij) = <ab||ij> / (e_i+e_j-e_a-e_b),   Fac = f_ac - 1/2 t(af,mn)<ef||mn>,
//       Fki = f_ki + 1/2 <ef||mn> t(ef,in),        Wkaci = <ka||ci> + <fe||nm> t(bf,jn)
//   method 2, SF-CPSd = method 1 + the doubles-doubles first-order coupling of
//     EE-CPSd.SF/calC2_VR2.f90:4-47 (ring P(ab)P(ij) r(ae,im) <mb||ej>,
//     hh-ladder 1/2 r(ab,mn) <mn||ij>, pp-ladder 1/2 <ab||ef> r(ef,ij)).
// Every term is one mod_contract-style call on the engine with the vector batch
// folded into the GEMM, in the order of the reference's calAR.
#include "dpd_sigma.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <numeric>

namespace dpd {

static Desc mkdesc(const int64_t* a, const int64_t* b, const int64_t* c, const int64_t* d, int i1, int i2,
                   int sym) {
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

static Operand op(const Desc& d, const char* lab, int rank, const double* data, const float* f32 = nullptr) {
  Operand o;
  o.d = d;
  o.rank = rank;
  std::memset(o.lab, 0, 4);
  std::memcpy(o.lab, lab, std::min<size_t>((size_t)rank, std::strlen(lab)));
  o.data = data;
  o.data_f32 = f32;
  return o;
}

// static-operand ids are uid * 64 + tensor + 1: does id belong to problem uid?
static bool id_of_problem(uint64_t id, uint64_t uid) { return (id - 1) / 64 == uid; }

// a constant tensor of the problem: packed GEMM panels of it are built once
static Operand konst(Operand o, const dpd_problem* p, int which) {
  o.static_id = p->uid * 64 + (uint64_t)which + 1;
  return o;
}

static Operand batched(Operand o, int64_t stride) {
  o.batched = true;
  o.bstride = stride;
  return o;
}

static EorbOff eorb_off(const dpd_problem* p) {
  EorbOff e;
  int n = p->ctx->sym.nirrep;
  int64_t k = 0;
  for (int h = 0; h < 8; h++) {
    e.offo[h] = k;
    if (h < n) k += p->popo[h];
  }
  for (int h = 0; h < 8; h++) {
    e.offv[h] = k;
    if (h < n) k += p->popv[h];
  }
  return e;
}

// one step of a sigma build: a contraction (kind 444/424/442), the clearing of
// the output (kind 0) or a diagonal term y += d * x over nvec vectors (kind 1)
struct SigmaTerm {
  int kind = 0;
  Operand a, b, c;
  double fact = 1.0;
  int anti = 0;
  int nb = 1;
  const char* tag = nullptr;
  int64_t n = 0, stride = 0;
  const double* d = nullptr;
  const double* x = nullptr;
  double* y = nullptr;
  SinglesArgs sg;  // kind 3
};

// kind 3: the singles-from-singles terms W r (+ Fac r - r Fki) in one launch;
// W is the (v,o,v,o) operator of symmetry 1 whose pairs have R1's symmetry S
static SinglesArgs singles_args(const dpd_problem* p, const double* W, bool fock) {
  const Sym& Sy = p->ctx->sym;
  const int X = p->symR - 1;
  const int64_t *v = p->popv, *o = p->popo, *z = p->zorb;
  const DevLayout LR = make_layout(Sy, mkdesc(v, o, z, z, 1, 1, p->symR));
  const DevLayout LW = make_layout(Sy, mkdesc(v, o, v, o, 1, 1, 1));
  const DevLayout LF = make_layout(Sy, mkdesc(v, v, z, z, 1, 1, 1));
  const DevLayout LK = make_layout(Sy, mkdesc(o, o, z, z, 1, 1, 1));
  SinglesArgs g;
  std::memset(&g, 0, sizeof g);
  g.l1 = LR.len;
  g.stride = p->dim;
  for (int hi = 0; hi < Sy.nirrep; hi++) {
    const int ha = Sy.dp[hi][X];
    if (v[ha] == 0 || o[hi] == 0) continue;
    g.start[g.nblk] = LR.blk[0] + LR.offL[X][hi];
    g.na[g.nblk] = (int32_t)v[ha];
    g.ni[g.nblk] = (int32_t)o[hi];
    g.fac_off[g.nblk] = LF.offL[0][ha];
    g.fki_off[g.nblk] = LK.offL[0][hi];
    g.nblk++;
  }
  g.start[g.nblk] = g.l1;
  g.W = W + LW.blk[X];
  g.Fac = fock ? (const double*)p->fac.p : nullptr;
  g.Fki = fock ? (const double*)p->fki.p : nullptr;
  return g;
}

// the terms of sigma = J R for one problem, in order
static int sigma_terms(dpd_problem* p, const double* R, double* sigma, int64_t nvec, std::vector<SigmaTerm>& T) {
  dpd_ctx* c = p->ctx;
  const int64_t *v = p->popv, *o = p->popo, *z = p->zorb;
  const int S = p->symR;
  const int nb = (int)nvec;
  // FP32 copies feed the gather kernels; the packed-panel kernels convert on packing
  const bool f32 = p->prec == PREC_F32 && c->packing == 0;
  if (f32 && !p->iabj32.p) {
    int rc0 = problem_make_f32(p);
    if (rc0) return rc0;
  }
  auto F = [&](const DevBuf& b) -> const float* { return f32 ? (const float*)b.p : nullptr; };
  auto lab = [](Operand x, const char* l, int rank) {
    x.rank = rank;
    std::memset(x.lab, 0, 4);
    std::memcpy(x.lab, l, rank);
    return x;
  };
  auto con = [&](int kind, const Operand& a, const Operand& b, const Operand& cc, double fact, int anti,
                 const char* tag) {
    SigmaTerm t;
    t.kind = kind;
    t.a = a;
    t.b = b;
    t.c = cc;
    t.fact = fact;
    t.anti = anti;
    t.nb = nb;
    t.tag = tag;
    T.push_back(t);
  };
  auto singles = [&](const double* W, bool fock) {
    SigmaTerm t;
    t.kind = 3;
    t.sg = singles_args(p, W, fock);
    t.sg.nvec = nb;
    t.x = R;
    t.y = sigma;
    T.push_back(t);
  };
  auto diag = [&](int64_t n, const double* d, const double* x, double* y) {
    SigmaTerm t;
    t.kind = 1;
    t.n = n;
    t.d = d;
    t.x = x;
    t.y = y;
    t.nb = nb;
    t.stride = p->dim;
    T.push_back(t);
  };
  {
    SigmaTerm t;
    t.kind = 0;
    t.n = p->dim * nvec;
    t.y = sigma;
    T.push_back(t);
  }
  Desc dR1 = mkdesc(v, o, z, z, 1, 1, S);
  Operand R1 = batched(op(dR1, "", 4, R), p->dim);
  Operand C1 = batched(op(dR1, "", 4, sigma), p->dim);
  if (p->method == 0) {
    // e(a,i) r(a,i)
    diag(p->l1, (const double*)p->diag.p, R, sigma);
    // c1_r1iabj: vv2v_422(iabj, r1, c1, 'fnai','fn','ai') as one FP64 launch
    singles((const double*)p->iabj.p, false);
    return DPD_OK;
  }
  Desc dR2 = mkdesc(v, v, o, o, 0, 0, S);
  Operand R2 = batched(op(dR2, "", 4, R + p->l1), p->dim);
  Operand C2 = batched(op(dR2, "", 4, sigma + p->l1), p->dim);
  Operand abci = konst(op(mkdesc(v, v, v, o, 0, 1, 1), "feam", 4, (const double*)p->abci.p, F(p->abci32)), p, 5);
  Operand ijka = konst(op(mkdesc(o, o, o, v, 0, 1, 1), "mnie", 4, (const double*)p->ijka.p, F(p->ijka32)), p, 6);
  // ---- singles (CISdoo.SD.SF/calAR.f90:41-56)
  // c1_r1Fae: vv2v_222(Fac, r1, c1, 'ae','ei','ai'), c1_r1Fmi: vv2v_222(r1, Fki, c1,
  // -1, 'am','mi','ai') and c1_r1iabj: vv2v_422(Wkaci, r1, c1, 'fnai','fn','ai'):
  // three small products on R1 (a few hundred elements per vector), one FP64 launch
  singles((const double*)p->wkaci.p, true);
  // c1_r2abci: vv2v_442(abci, r2, c1, 'feam','feim','ai')  (r2 iop 0: fact)
  con(442, abci, lab(R2, "feim", 4), lab(C1, "ai", 2), 1.0, 0, "c1_r2abci");
  // c1_r2ijka: vv2v_442(r2, ijka, c1, -1, 'aemn','mnie','ai')
  con(442, lab(R2, "aemn", 4), ijka, lab(C1, "ai", 2), -1.0, 0, "c1_r2ijka");
  // ---- doubles (CISdoo.SD.SF/calAR.f90:63-77)
  diag(p->l2, (const double*)p->diag.p + p->l1, R + p->l1, sigma + p->l1);
  // c2_r1abci: vv2v_424(abci, r1, c2, 'abej','ei','abij', anti=2)
  con(424, lab(abci, "abej", 4), lab(R1, "ei", 2), lab(C2, "abij", 4), 1.0, 2, "c2_r1abci");
  // c2_r1ijka: vv2v_424(ijka, r1, c2, -1, 'ijmb','am','abij', anti=1)
  con(424, lab(ijka, "ijmb", 4), lab(R1, "am", 2), lab(C2, "abij", 4), -1.0, 1, "c2_r1ijka");
  if (p->method == 2) {
    Operand W = konst(op(mkdesc(v, o, v, o, 1, 1, 1), "embj", 4, (const double*)p->iabj.p, F(p->iabj32)), p, 7);
    Operand ijkl = konst(op(mkdesc(o, o, o, o, 0, 0, 1), "mnij", 4, (const double*)p->ijkl.p, F(p->ijkl32)), p, 8);
    // c2_r2Wmbej: vv2v_444(r2, W, c2, 'aeim','embj','abij', anti=3)
    con(444, lab(R2, "aeim", 4), W, lab(C2, "abij", 4), 1.0, 3, "c2_r2Wmbej");
    // c2_r2Wmnij: vv2v_444(r2, ijkl, c2, 'abmn','mnij','abij')
    con(444, lab(R2, "abmn", 4), ijkl, lab(C2, "abij", 4), 1.0, 0, "c2_r2Wmnij");
    if (p->abcd_host) {
      // c2_r2abcd with <ab||ef> streamed from host memory (c2_r2abcd.f90:29-104)
      SigmaTerm t;
      t.kind = 2;
      t.x = R + p->l1;
      t.y = sigma + p->l1;
      t.nb = nb;
      t.stride = p->dim;
      t.fact = 1.0;
      T.push_back(t);
    } else {
      Operand abcd =
          konst(op(mkdesc(v, v, v, v, 0, 0, 1), "abef", 4, (const double*)p->abcd.p, F(p->abcd32)), p, 9);
      // c2_r2abcd: vv2v_444(abcd, r2, c2, 'abef','efij','abij')
      con(444, abcd, lab(R2, "efij", 4), lab(C2, "abij", 4), 1.0, 0, "c2_r2abcd");
    }
  }
  return DPD_OK;
}

static int run_simple_term(dpd_ctx* c, const SigmaTerm& t) {
  if (t.kind == 0) {
    DPD_HIP_TRY(launch_fill(t.n, 0.0, t.y, c->stream));
  } else if (t.kind == 3) {
    DPD_HIP_TRY(launch_singles(t.sg, t.x, t.y, c->stream));
  } else {
    DPD_HIP_TRY(launch_hadamard_acc(t.n, t.d, t.x, t.y, t.nb, t.stride, t.stride, c->stream));
  }
  return DPD_OK;
}

int problem_sigma(dpd_problem* p, const double* R, double* sigma, int64_t nvec) {
  std::vector<SigmaTerm> T;
  int rc = sigma_terms(p, R, sigma, nvec, T);
  if (rc) return rc;
  dpd_ctx* c = p->ctx;
  // the antisymmetrising epilogues of consecutive contractions into sigma2
  // (c2_r1abci, c2_r1ijka, c2_r2Wmbej) are held back and finished by one
  // launch, in the same order, before the next term that touches sigma
  EpiDefer held[kEpiMax];
  int nheld = 0;
  for (const SigmaTerm& t : T) {
    const bool deferrable = (t.kind == 424 || t.kind == 444) && t.anti != 0 && nheld < kEpiMax &&
                            (nheld == 0 || held[0].out == t.c.data);
    if (!deferrable && nheld) {
      if ((rc = epi_flush(c, held, nheld, "c2:epi"))) return rc;
      nheld = 0;
    }
    if (t.kind == 0 || t.kind == 1 || t.kind == 3) {
      rc = run_simple_term(c, t);
    } else if (t.kind == 2) {
      rc = problem_pp_ladder_streamed(p, t.x, t.y, t.nb, t.stride, p->symR, t.fact);
    } else if (deferrable) {
      held[nheld] = EpiDefer();
      held[nheld].slot = nheld;
      rc = contract_call(c, t.kind, t.a, t.b, t.c, t.fact, t.anti, t.nb, p->prec, nullptr, t.tag, &held[nheld]);
      if (!rc && held[nheld].deferred) nheld++;
    } else {
      rc = contract_call(c, t.kind, t.a, t.b, t.c, t.fact, t.anti, t.nb, p->prec, nullptr, t.tag);
    }
    if (rc) return rc;
  }
  if (nheld) rc = epi_flush(c, held, nheld, "c2:epi");
  return rc;
}

// sigma of n problems of one method and context, each contraction of all of
// them in one grouped GEMM launch
int problem_sigma_group(dpd_problem* const* ps, const double* const* R, double* const* S, const int64_t* nvec,
                        int n) {
  if (n <= 0) return DPD_OK;
  dpd_ctx* c = ps[0]->ctx;
  std::vector<std::vector<SigmaTerm>> T(n);
  for (int i = 0; i < n; i++) {
    if (ps[i]->ctx != c || ps[i]->method != ps[0]->method || ps[i]->prec != ps[0]->prec)
      return fail(DPD_EINVAL, "grouped sigma: problems must share context, method and precision");
    if (ps[i]->abcd_host) return fail(DPD_EINVAL, "grouped sigma: abcd out of core is not supported");
    int rc = sigma_terms(ps[i], R[i], S[i], nvec[i], T[i]);
    if (rc) return rc;
    if (T[i].size() != T[0].size()) return fail(DPD_EINVAL, "grouped sigma: term lists differ (internal)");
  }
  std::vector<Operand> a(n), b(n), cc(n);
  std::vector<int> nb(n);
  for (size_t k = 0; k < T[0].size(); k++) {
    const SigmaTerm& t0 = T[0][k];
    if (t0.kind == 0 || t0.kind == 1 || t0.kind == 3) {
      for (int i = 0; i < n; i++) {
        int rc = run_simple_term(c, T[i][k]);
        if (rc) return rc;
      }
      continue;
    }
    for (int i = 0; i < n; i++) {
      a[i] = T[i][k].a;
      b[i] = T[i][k].b;
      cc[i] = T[i][k].c;
      nb[i] = T[i][k].nb;
    }
    int rc = contract_group(c, t0.kind, a.data(), b.data(), cc.data(), n, t0.fact, t0.anti, nb.data(),
                            ps[0]->prec, t0.tag);
    if (rc) return rc;
  }
  return DPD_OK;
}
static int upload(DevBuf& b, const std::vector<double>& h, hipStream_t s) {
  int rc = b.alloc(std::max<size_t>(8, h.size() * sizeof(double)));
  if (rc) return rc;
  if (!h.empty()) DPD_HIP_TRY(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  DPD_HIP_TRY(hipStreamSynchronize(s));
  return DPD_OK;
}

static int make_f32(DevBuf& dst, const DevBuf& src, int64_t n, hipStream_t s) {
  int rc = dst.alloc(std::max<size_t>(8, n * sizeof(float)));
  if (rc) return rc;
  DPD_HIP_TRY(launch_convert_f32(n, (const double*)src.p, (float*)dst.p, s));
  return DPD_OK;
}

int problem_make_f32(dpd_problem* p) {
  hipStream_t s = p->ctx->stream;
  int rc;
  if ((rc = make_f32(p->iabj32, p->iabj, p->len_iabj, s))) return rc;
  if (p->method >= 1) {
    if ((rc = make_f32(p->ijka32, p->ijka, p->len_ijka, s))) return rc;
    if ((rc = make_f32(p->abci32, p->abci, p->len_abci, s))) return rc;
    if ((rc = make_f32(p->fac32, p->fac, p->len_fac, s))) return rc;
    if ((rc = make_f32(p->fki32, p->fki, p->len_fki, s))) return rc;
    if ((rc = make_f32(p->wkaci32, p->wkaci, p->len_iabj, s))) return rc;
  }
  if (p->method == 2) {
    if ((rc = make_f32(p->ijkl32, p->ijkl, p->len_ijkl, s))) return rc;
    if (!p->abcd_host && (rc = make_f32(p->abcd32, p->abcd, p->len_abcd, s))) return rc;
  }
  DPD_HIP_TRY(hipStreamSynchronize(s));
  return DPD_OK;
}
}  // namespace dpd

using namespace dpd;

// diagonal, MP2 amplitudes and the CIS(Doo) intermediates from the current
// orbital energies and integrals (also after a caller changed them)
static int problem_finalize(dpd_problem* p) {
  dpd_ctx* c = p->ctx;
  const Sym& S = c->sym;
  const int64_t *v = p->popv, *o = p->popo, *z = p->zorb;
  hipStream_t st = c->stream;
  int rc;
  // diagonal (OED1, OED2): lib_cc/Eorb_diff.f90
  if ((rc = p->diag.alloc(sizeof(double) * std::max<int64_t>(1, p->dim)))) return rc;
  EorbOff eo = eorb_off(p);
  DPD_HIP_TRY(launch_orbital_energy(make_layout(S, mkdesc(v, o, z, z, 1, 1, p->symR)), (const double*)p->eorb.p,
                                    eo, 1, (double*)p->diag.p, st));
  if (p->method >= 1) {
    DPD_HIP_TRY(launch_orbital_energy(make_layout(S, mkdesc(v, v, o, o, 0, 0, p->symR)),
                                      (const double*)p->eorb.p, eo, 2, (double*)p->diag.p + p->l1, st));
    // MP2 amplitudes t = <ab||ij> / (e_i+e_j-e_a-e_b)   (denomT.f90:94-249)
    Desc dT = mkdesc(v, v, o, o, 0, 0, 1);
    if ((rc = p->mp2t.alloc(sizeof(double) * std::max<int64_t>(1, p->len_abij)))) return rc;
    DPD_HIP_TRY(launch_orbital_energy(make_layout(S, dT), (const double*)p->eorb.p, eo, 3, (double*)p->mp2t.p, st));
    DPD_HIP_TRY(launch_div(p->len_abij, (const double*)p->abij.p, (const double*)p->mp2t.p, (double*)p->mp2t.p, st));
    // Fac = f_vv (putDiagFock 'V') + Fae_r2abij(MP2t, abij)
    Desc dFac = mkdesc(v, v, z, z, 1, 1, 1), dFki = mkdesc(o, o, z, z, 1, 1, 1);
    p->len_fac = make_layout(S, dFac).len;
    p->len_fki = make_layout(S, dFki).len;
    std::vector<double> hf(p->len_fac, 0.0), hk(p->len_fki, 0.0);
    {
      int64_t offv = 0, offo = 0, ev = 0;
      for (int h = 0; h < S.nirrep; h++) ev += p->popo[h];
      for (int h = 0; h < S.nirrep; h++) {
        int64_t nv = p->popv[h], no = p->popo[h];
        int64_t bv = S.isymoffso(h + 1, 1, v, v, 1) - 1, bo = S.isymoffso(h + 1, 1, o, o, 1) - 1;
        for (int64_t a = 0; a < nv; a++) hf[bv + a * nv + a] = p->eorb_host[ev + offv + a];
        for (int64_t i = 0; i < no; i++) hk[bo + i * no + i] = p->eorb_host[offo + i];
        offv += nv;
        offo += no;
      }
    }
    if ((rc = upload(p->fac, hf, st))) return rc;
    if ((rc = upload(p->fki, hk, st))) return rc;
    Operand T = op(dT, "afmn", 4, (const double*)p->mp2t.p);
    Operand Vabij = op(dT, "efmn", 4, (const double*)p->abij.p);
    // Fae_r2abij.f90:16: vv2v_442(r2, abij, Fae, -fact, 'afmn','efmn','ae')
    rc = contract_call(c, 442, T, Vabij, op(dFac, "ae", 2, (const double*)p->fac.p), -1.0, 0, 1, PREC_F64);
    if (rc) return rc;
    // Fmi_r2abij.f90:18: vv2v_442(abij, r2, Fmi, fact, 'efmn','efin','mi')
    T = op(dT, "efin", 4, (const double*)p->mp2t.p);
    rc = contract_call(c, 442, Vabij, T, op(dFki, "mi", 2, (const double*)p->fki.p), 1.0, 0, 1, PREC_F64);
    if (rc) return rc;
    // Wkaci = iabj + Wmbej_r2abij(MP2t, abij, 2.0): vv2v_444(abij, r2, W, 0.5*2, 'fenm','bfjn','embj')
    if ((rc = p->wkaci.alloc(sizeof(double) * std::max<int64_t>(1, p->len_iabj)))) return rc;
    DPD_HIP_TRY(launch_copy(p->len_iabj, (const double*)p->iabj.p, (double*)p->wkaci.p, st));
    rc = contract_call(c, 444, op(dT, "fenm", 4, (const double*)p->abij.p), op(dT, "bfjn", 4, (const double*)p->mp2t.p),
                       op(mkdesc(v, o, v, o, 1, 1, 1), "embj", 4, (const double*)p->wkaci.p), 1.0, 0, 1, PREC_F64);
    if (rc) return rc;
  }
  DPD_HIP_TRY(hipStreamSynchronize(st));
  return DPD_OK;
}


// ---------------------------------------------------------------------------
// Transpose of the sigma operator and the adjoint (transpose-contraction)
// gradient of L.sigma(R).  Every term c_out += f * (V x R) of problem_sigma has
//   transpose:  t_R  += f * (V x l_out)      (the expansions and
//                antisymmetrisers of the forward call become, in the transposed
//                call, the expansion of l_out and the antisymmetrise+compress
//                of t_R: the surface applies them from the labels)
//   gradient:   Vbar += f * (l_out x R)     in V's labels and layout
// The MP2 amplitudes and the CIS(Doo) intermediates are recomputed here from the
// integrals and orbital energies, not stored from the forward pass.
// ---------------------------------------------------------------------------
static Operand lab_of(Operand x, const char* l, int rank) {
  x.rank = rank;
  std::memset(x.lab, 0, 4);
  std::memcpy(x.lab, l, rank);
  return x;
}

int dpd::problem_sigma_t(dpd_problem* p, const double* L, double* t, int64_t nvec) {
  if (p->abcd_host) return fail(DPD_EINVAL, "transposed operator: abcd out of core is not supported");
  dpd_ctx* c = p->ctx;
  const int64_t *v = p->popv, *o = p->popo, *z = p->zorb;
  const int S = p->symR;
  const int nb = (int)nvec;
  const int pr = p->prec;
  DPD_HIP_TRY(launch_fill(p->dim * nvec, 0.0, t, c->stream));
  Desc dR1 = mkdesc(v, o, z, z, 1, 1, S);
  Operand L1 = batched(op(dR1, "", 4, L), p->dim);
  Operand T1 = batched(op(dR1, "", 4, t), p->dim);
  int rc;
  if (p->method == 0) {
    DPD_HIP_TRY(launch_hadamard_acc(p->l1, (const double*)p->diag.p, L, t, nb, p->dim, p->dim, c->stream));
    Operand W = konst(op(mkdesc(v, o, v, o, 1, 1, 1), "fnai", 4, (const double*)p->iabj.p), p, 1);
    return contract_call(c, 444, W, lab_of(L1, "ai+-", 4), lab_of(T1, "fn+-", 4), 1.0, 0, nb, pr, nullptr,
                         "t_c1_r1iabj");
  }
  Desc dR2 = mkdesc(v, v, o, o, 0, 0, S);
  Operand L2 = batched(op(dR2, "", 4, L + p->l1), p->dim);
  Operand T2 = batched(op(dR2, "", 4, t + p->l1), p->dim);
  Operand Fac = konst(op(mkdesc(v, v, z, z, 1, 1, 1), "ae=+", 4, (const double*)p->fac.p), p, 2);
  Operand Fki = konst(op(mkdesc(o, o, z, z, 1, 1, 1), "mi=-", 4, (const double*)p->fki.p), p, 3);
  Operand Wk = konst(op(mkdesc(v, o, v, o, 1, 1, 1), "fnai", 4, (const double*)p->wkaci.p), p, 4);
  Operand abci = konst(op(mkdesc(v, v, v, o, 0, 1, 1), "feam", 4, (const double*)p->abci.p), p, 5);
  Operand ijka = konst(op(mkdesc(o, o, o, v, 0, 1, 1), "mnie", 4, (const double*)p->ijka.p), p, 6);
  if ((rc = contract_call(c, 444, Fac, lab_of(L1, "ai+-", 4), lab_of(T1, "ei=-", 4), 1.0, 0, nb, pr, nullptr,
                          "t_c1_r1Fae")))
    return rc;
  if ((rc = contract_call(c, 444, lab_of(L1, "ai+-", 4), Fki, lab_of(T1, "am=+", 4), -1.0, 0, nb, pr, nullptr,
                          "t_c1_r1Fmi")))
    return rc;
  if ((rc = contract_call(c, 444, Wk, lab_of(L1, "ai+-", 4), lab_of(T1, "fn+-", 4), 1.0, 0, nb, pr, nullptr,
                          "t_c1_r1iabj")))
    return rc;
  if ((rc = contract_call(c, 424, abci, lab_of(L1, "ai", 2), lab_of(T2, "feim", 4), 1.0, 2, nb, pr, nullptr,
                          "t_c1_r2abci")))
    return rc;
  if ((rc = contract_call(c, 424, ijka, lab_of(L1, "ai", 2), lab_of(T2, "aemn", 4), -1.0, 1, nb, pr, nullptr,
                          "t_c1_r2ijka")))
    return rc;
  DPD_HIP_TRY(launch_hadamard_acc(p->l2, (const double*)p->diag.p + p->l1, L + p->l1, t + p->l1, nb, p->dim,
                                  p->dim, c->stream));
  if ((rc = contract_call(c, 442, lab_of(abci, "abej", 4), lab_of(L2, "abij", 4), lab_of(T1, "ei", 2), 1.0, 0, nb,
                          pr, nullptr, "t_c2_r1abci")))
    return rc;
  if ((rc = contract_call(c, 442, lab_of(L2, "abij", 4), lab_of(ijka, "ijmb", 4), lab_of(T1, "am", 2), -1.0, 0, nb,
                          pr, nullptr, "t_c2_r1ijka")))
    return rc;
  if (p->method == 2) {
    Operand W = konst(op(mkdesc(v, o, v, o, 1, 1, 1), "embj", 4, (const double*)p->iabj.p), p, 7);
    Operand ijkl = konst(op(mkdesc(o, o, o, o, 0, 0, 1), "mnij", 4, (const double*)p->ijkl.p), p, 8);
    Operand abcd = konst(op(mkdesc(v, v, v, v, 0, 0, 1), "abef", 4, (const double*)p->abcd.p), p, 9);
    if ((rc = contract_call(c, 444, lab_of(L2, "abij", 4), W, lab_of(T2, "aeim", 4), 1.0, 3, nb, pr, nullptr,
                            "t_c2_r2Wmbej")))
      return rc;
    if ((rc = contract_call(c, 444, lab_of(L2, "abij", 4), ijkl, lab_of(T2, "abmn", 4), 1.0, 0, nb, pr, nullptr,
                            "t_c2_r2Wmnij")))
      return rc;
    if ((rc = contract_call(c, 444, abcd, lab_of(L2, "abij", 4), lab_of(T2, "efij", 4), 1.0, 0, nb, pr, nullptr,
                            "t_c2_r2abcd")))
      return rc;
  }
  return DPD_OK;
}

// gradient buffers: eorb and the integral blocks, in the problem's layouts
struct GradOut {
  double* eorb = nullptr;
  double *iabj = nullptr, *ijka = nullptr, *ijkl = nullptr, *abij = nullptr, *abci = nullptr, *abcd = nullptr;
};

// scale * d(L . sigma(R)) / d(theta), accumulated into g (single vectors L, R)
static int problem_grad(dpd_problem* p, const double* L, const double* R, double scale, const GradOut& g) {
  dpd_ctx* c = p->ctx;
  const Sym& Sy = c->sym;
  const int64_t *v = p->popv, *o = p->popo, *z = p->zorb;
  const int S = p->symR;
  const int pr = p->prec;
  hipStream_t st = c->stream;
  EorbOff eo = eorb_off(p);
  Desc dR1 = mkdesc(v, o, z, z, 1, 1, S);
  Operand L1 = op(dR1, "", 4, L), R1 = op(dR1, "", 4, R);
  Desc dW = mkdesc(v, o, v, o, 1, 1, 1);
  int rc;
  auto need = [&](const double* ptr, const char* what) -> int {
    return ptr ? DPD_OK : fail(DPD_EINVAL, "gradient buffer for %s is required", what);
  };
  if ((rc = need(g.eorb, "eorb")) || (rc = need(g.iabj, "iabj"))) return rc;
  if (p->method == 0) {
    // c1_r1iabj: iabj_bar(fn,ai) += R1(fn) L1(ai)
    rc = contract_call(c, 444, lab_of(R1, "fn+-", 4), lab_of(L1, "ai+-", 4), op(dW, "fnai", 4, g.iabj), scale, 0, 1,
                       pr);
    if (rc) return rc;
    // OED1 (e_a - e_i) r(a,i)
    EorbAdj adj = {{1, 2, 0, 0}, {1.0, -1.0, 0.0, 0.0}, scale};
    DPD_HIP_TRY(launch_eorb_adjoint(make_layout(Sy, dR1), L, R, eo, adj, g.eorb, st));
    return DPD_OK;
  }
  if ((rc = need(g.ijka, "ijka")) || (rc = need(g.abij, "abij")) || (rc = need(g.abci, "abci"))) return rc;
  if (p->method == 2 && ((rc = need(g.ijkl, "ijkl")) || (rc = need(g.abcd, "abcd")))) return rc;
  Desc dR2 = mkdesc(v, v, o, o, 0, 0, S), dT = mkdesc(v, v, o, o, 0, 0, 1);
  Operand L2 = op(dR2, "", 4, L + p->l1), R2 = op(dR2, "", 4, R + p->l1);
  Desc dFac = mkdesc(v, v, z, z, 1, 1, 1), dFki = mkdesc(o, o, z, z, 1, 1, 1);
  Desc dabci = mkdesc(v, v, v, o, 0, 1, 1), dijka = mkdesc(o, o, o, v, 0, 1, 1);
  DevBuf fac_b, fki_b, wk_b, t2_b, t2bar_b, den_b, tmp_b;
  if ((rc = fac_b.alloc(sizeof(double) * p->len_fac)) || (rc = fki_b.alloc(sizeof(double) * p->len_fki)) ||
      (rc = wk_b.alloc(sizeof(double) * p->len_iabj)) || (rc = t2_b.alloc(sizeof(double) * p->len_abij)) ||
      (rc = t2bar_b.alloc(sizeof(double) * p->len_abij)) || (rc = den_b.alloc(sizeof(double) * p->len_abij)) ||
      (rc = tmp_b.alloc(sizeof(double) * p->len_abij)))
    return rc;
  double *facb = (double*)fac_b.p, *fkib = (double*)fki_b.p, *wkb = (double*)wk_b.p;
  double *t2 = (double*)t2_b.p, *t2b = (double*)t2bar_b.p, *den = (double*)den_b.p, *tmp = (double*)tmp_b.p;
  DPD_HIP_TRY(launch_fill(p->len_fac, 0.0, facb, st));
  DPD_HIP_TRY(launch_fill(p->len_fki, 0.0, fkib, st));
  DPD_HIP_TRY(launch_fill(p->len_iabj, 0.0, wkb, st));
  DPD_HIP_TRY(launch_fill(p->len_abij, 0.0, t2b, st));
  auto call = [&](int kind, const Operand& a, const Operand& b, const Operand& out, double f, int anti) {
    return contract_call(c, kind, a, b, out, f, anti, 1, pr);
  };
  // ---- singles terms
  if ((rc = call(444, lab_of(L1, "ai+-", 4), lab_of(R1, "ei=-", 4), op(dFac, "ae=+", 4, facb), scale, 0))) return rc;
  if ((rc = call(444, lab_of(R1, "am=+", 4), lab_of(L1, "ai+-", 4), op(dFki, "mi=-", 4, fkib), -scale, 0))) return rc;
  if ((rc = call(444, lab_of(R1, "fn+-", 4), lab_of(L1, "ai+-", 4), op(dW, "fnai", 4, wkb), scale, 0))) return rc;
  if ((rc = call(424, lab_of(R2, "feim", 4), lab_of(L1, "ai", 2), op(dabci, "feam", 4, g.abci), scale, 0))) return rc;
  if ((rc = call(424, lab_of(R2, "aemn", 4), lab_of(L1, "ai", 2), op(dijka, "mnie", 4, g.ijka), -scale, 0))) return rc;
  // ---- doubles terms
  {
    EorbAdj adj = {{1, 1, 2, 2}, {1.0, 1.0, -1.0, -1.0}, scale};  // OED2: e_a + e_b - e_i - e_j
    DPD_HIP_TRY(launch_eorb_adjoint(make_layout(Sy, dR2), L + p->l1, R + p->l1, eo, adj, g.eorb, st));
  }
  if ((rc = call(424, lab_of(L2, "abij", 4), lab_of(R1, "ei", 2), op(dabci, "abej", 4, g.abci), scale, 0))) return rc;
  if ((rc = call(424, lab_of(L2, "abij", 4), lab_of(R1, "am", 2), op(dijka, "ijmb", 4, g.ijka), -scale, 0))) return rc;
  if (p->method == 2) {
    if ((rc = call(444, lab_of(R2, "aeim", 4), lab_of(L2, "abij", 4), op(dW, "embj", 4, g.iabj), scale, 0)))
      return rc;
    if ((rc = call(444, lab_of(R2, "abmn", 4), lab_of(L2, "abij", 4),
                   op(mkdesc(o, o, o, o, 0, 0, 1), "mnij", 4, g.ijkl), scale, 0)))
      return rc;
    if ((rc = call(444, lab_of(L2, "abij", 4), lab_of(R2, "efij", 4),
                   op(mkdesc(v, v, v, v, 0, 0, 1), "abef", 4, g.abcd), scale, 0)))
      return rc;
  }
  // ---- intermediates, recomputed: D = e_i + e_j - e_a - e_b, t = abij / D
  const DevLayout LT = make_layout(Sy, dT);
  DPD_HIP_TRY(launch_orbital_energy(LT, (const double*)p->eorb.p, eo, 3, den, st));
  DPD_HIP_TRY(launch_div(p->len_abij, (const double*)p->abij.p, den, t2, st));
  const double* V = (const double*)p->abij.p;
  // Fac = f_vv - sum t(af,mn) v(ef,mn): diagonal -> eorb (virtual), then t and v
  if ((rc = call(424, op(dT, "efmn", 4, V), op(dFac, "ae", 2, facb), op(dT, "afmn", 4, t2b), -1.0, 1))) return rc;
  if ((rc = call(424, op(dT, "afmn", 4, t2), op(dFac, "ae", 2, facb), op(dT, "efmn", 4, g.abij), -1.0, 1))) return rc;
  // Fki = f_oo + sum v(ef,mn) t(ef,in)
  if ((rc = call(424, op(dT, "efin", 4, t2), op(dFki, "mi", 2, fkib), op(dT, "efmn", 4, g.abij), 1.0, 2))) return rc;
  if ((rc = call(424, op(dT, "efmn", 4, V), op(dFki, "mi", 2, fkib), op(dT, "efin", 4, t2b), 1.0, 2))) return rc;
  // Wk = iabj + sum v(fe,nm) t(bf,jn)
  DPD_HIP_TRY(launch_axpy(p->len_iabj, 1.0, wkb, g.iabj, st));
  if ((rc = call(444, op(dW, "embj", 4, wkb), op(dT, "bfjn", 4, t2), op(dT, "fenm", 4, g.abij), 1.0, 3))) return rc;
  if ((rc = call(444, op(dT, "fenm", 4, V), op(dW, "embj", 4, wkb), op(dT, "bfjn", 4, t2b), 1.0, 3))) return rc;
  // t = v / D: v_bar += t_bar / D; e_bar from dD (t_bar t / D, + for a,b and - for i,j)
  DPD_HIP_TRY(launch_div(p->len_abij, t2b, den, tmp, st));
  DPD_HIP_TRY(launch_axpy(p->len_abij, 1.0, tmp, g.abij, st));
  {
    EorbAdj adj = {{1, 1, 2, 2}, {1.0, 1.0, -1.0, -1.0}, 1.0};
    DPD_HIP_TRY(launch_eorb_adjoint(LT, tmp, t2, eo, adj, g.eorb, st));
  }
  // diagonals of Fac / Fki -> eorb (host: tiny)
  std::vector<double> hf(p->len_fac), hk(p->len_fki), de(p->eorb_host.size(), 0.0);
  DPD_HIP_TRY(hipMemcpyAsync(hf.data(), facb, sizeof(double) * hf.size(), hipMemcpyDeviceToHost, st));
  DPD_HIP_TRY(hipMemcpyAsync(hk.data(), fkib, sizeof(double) * hk.size(), hipMemcpyDeviceToHost, st));
  DPD_HIP_TRY(hipStreamSynchronize(st));
  {
    int64_t offv = 0, offo = 0, ev = 0;
    for (int h = 0; h < Sy.nirrep; h++) ev += p->popo[h];
    for (int h = 0; h < Sy.nirrep; h++) {
      const int64_t nv = p->popv[h], no = p->popo[h];
      const int64_t bv = Sy.isymoffso(h + 1, 1, v, v, 1) - 1, bo = Sy.isymoffso(h + 1, 1, o, o, 1) - 1;
      for (int64_t a = 0; a < nv; a++) de[ev + offv + a] += hf[bv + a * nv + a];
      for (int64_t i = 0; i < no; i++) de[offo + i] += hk[bo + i * no + i];
      offv += nv;
      offo += no;
    }
  }
  DevBuf de_b;
  if ((rc = upload(de_b, de, st))) return rc;
  DPD_HIP_TRY(launch_axpy((int64_t)de.size(), 1.0, (const double*)de_b.p, g.eorb, st));
  DPD_HIP_TRY(hipStreamSynchronize(st));
  return DPD_OK;
}

extern "C" {

int dpd_synth_fill(const dpd_synth_spec* spec, const char* which, int64_t iop, double* out, int64_t len);

int dpd_problem_create(dpd_ctx* c, const dpd_synth_spec* spec_in, int method, int64_t symR, dpd_problem** out) {
  if (!c || !spec_in || !out) return fail(DPD_EINVAL, "null argument");
  if (method < 0 || method > 2) return fail(DPD_EINVAL, "method must be 0 (SF-CIS), 1 (CIS(Doo)) or 2 (SF-CPSd)");
  DPD_HIP_TRY(hipSetDevice(c->device));
  dpd_synth_spec spec = *spec_in;
  int rc = dpd_synth_layout(&spec);
  if (rc) return rc;
  if (spec.nirrep != c->sym.nirrep) return fail(DPD_EINVAL, "spec nirrep differs from the context");
  if (symR < 1 || symR > spec.nirrep) return fail(DPD_EINVAL, "symR out of range");
  auto p = std::make_unique<dpd_problem>();
  static std::atomic<uint64_t> next_uid{1};
  p->uid = next_uid++;
  p->ctx = c;
  p->spec = spec;
  p->method = method;
  p->symR = (int)symR;
  const Sym& S = c->sym;
  for (int h = 0; h < 8; h++) {
    p->popo[h] = h < S.nirrep ? spec.popo[h] : 0;
    p->popv[h] = h < S.nirrep ? spec.popv[h] : 0;
    p->zorb[h] = h == 0;
  }
  const int64_t *v = p->popv, *o = p->popo;
  p->l1 = S.irpdso((int)symR, v, o, 1);
  p->l2 = method == 0 ? 0 : S.idsymsoc((int)symR, v, v, 0, o, o, 0);
  p->dim = p->l1 + p->l2;
  hipStream_t st = c->stream;
  // orbital energies
  int64_t ne = 0;
  for (int h = 0; h < S.nirrep; h++) ne += p->popo[h] + p->popv[h];
  p->eorb_host.assign(ne, 0.0);
  rc = dpd_synth_fill(&spec, "eorb", 0, p->eorb_host.data(), ne);
  if (rc) return rc;
  if ((rc = upload(p->eorb, p->eorb_host, st))) return rc;
  auto gen = [&](const char* which, int iop, const Desc& d, DevBuf& dst, int64_t& len) -> int {
    len = make_layout(S, d).len;
    std::vector<double> h(len);
    int r = dpd_synth_fill(&spec, which, iop, h.data(), len);
    if (r) return r;
    return upload(dst, h, st);
  };
  if ((rc = gen("iabj", 1, mkdesc(v, o, v, o, 1, 1, 1), p->iabj, p->len_iabj))) return rc;
  if (method >= 1) {
    if ((rc = gen("ijka", 0, mkdesc(o, o, o, v, 0, 1, 1), p->ijka, p->len_ijka))) return rc;
    if ((rc = gen("abij", 0, mkdesc(v, v, o, o, 0, 0, 1), p->abij, p->len_abij))) return rc;
    if ((rc = gen("abci", 0, mkdesc(v, v, v, o, 0, 1, 1), p->abci, p->len_abci))) return rc;
    if (method == 2) {
      if ((rc = gen("ijkl", 0, mkdesc(o, o, o, o, 0, 0, 1), p->ijkl, p->len_ijkl))) return rc;
      if ((rc = gen("abcd", 0, mkdesc(v, v, v, v, 0, 0, 1), p->abcd, p->len_abcd))) return rc;
    }
  }
  if ((rc = problem_finalize(p.get()))) return rc;
  DPD_HIP_TRY(hipStreamSynchronize(st));
  *out = p.release();
  return DPD_OK;
}

int dpd_problem_destroy(dpd_problem* p) {
  if (!p) return DPD_OK;
  (void)hipStreamSynchronize(p->ctx->stream);
  // its integral panels go with it (plans are shared by problems of one shape)
  dpd::drop_pack_entries(p->ctx, id_of_problem, p->uid);
  delete p;
  return DPD_OK;
}

int64_t dpd_problem_dim(const dpd_problem* p) { return p ? p->dim : 0; }


int dpd_problem_set_precision(dpd_problem* p, int prec) {
  if (!p) return fail(DPD_EINVAL, "null problem");
  if (prec != PREC_F64 && prec != PREC_F32 && prec != PREC_BF16X3)
    return fail(DPD_EINVAL, "precision %d is not available", prec);
  p->prec = prec;
  return DPD_OK;
}

int dpd_problem_sigma(dpd_problem* p, const double* R, double* sigma, int64_t nvec) {
  if (!p || !R || !sigma || nvec < 1) return fail(DPD_EINVAL, "dpd_problem_sigma: bad arguments");
  if (p->stale) return fail(DPD_EINVAL, "problem inputs changed: call dpd_problem_refresh first");
  DPD_HIP_TRY(hipSetDevice(p->ctx->device));
  return problem_sigma(p, R, sigma, nvec);
}

int dpd_problem_diag(dpd_problem* p, double* diag) {
  if (!p || !diag) return fail(DPD_EINVAL, "null argument");
  DPD_HIP_TRY(hipMemcpyAsync(diag, p->diag.p, sizeof(double) * p->dim, hipMemcpyDeviceToDevice, p->ctx->stream));
  return DPD_OK;
}

// lib_cc/preDavidson.f90:120-223 (spin_flip=ab): unit vectors at the nroot lowest
// e(a_beta) - e(i_alpha) of the excited-state symmetry, a placed after the alpha
// virtuals of its irrep.
int dpd_problem_guess(dpd_problem* p, int64_t nroot, double* R, double* energies) {
  if (!p || !R || nroot < 1) return fail(DPD_EINVAL, "bad arguments");
  const Sym& S = p->ctx->sym;
  const dpd_synth_spec& sp = p->spec;
  const int sym = p->symR;
  // energies of alpha occupied / beta virtual spin orbitals per irrep
  int64_t offo[8], offv[8], k = 0;
  for (int h = 0; h < S.nirrep; h++) {
    offo[h] = k;
    k += p->popo[h];
  }
  for (int h = 0; h < S.nirrep; h++) {
    offv[h] = k;
    k += p->popv[h];
  }
  struct Cand {
    double e;
    int64_t lin;
    int ha, hi;
    int64_t a, i;
  };
  std::vector<Cand> cand;
  // (vrtb, popa) pair block of irrep sym: irrep of i ascending, a fastest
  for (int hi = 0; hi < S.nirrep; hi++) {
    int ha = S.dp[hi][sym - 1];
    for (int64_t i = 0; i < sp.popa[hi]; i++)
      for (int64_t a = 0; a < sp.vrtb[ha]; a++) {
        double e = p->eorb_host[offv[ha] + sp.vrta[ha] + a] - p->eorb_host[offo[hi] + i];
        cand.push_back({e, (int64_t)cand.size(), ha, hi, a, i});
      }
  }
  if ((int64_t)cand.size() < nroot) return fail(DPD_EINVAL, "ERROR: nSize < nlowest in getLowest");
  std::stable_sort(cand.begin(), cand.end(), [](const Cand& x, const Cand& y) { return x.e < y.e; });
  DevLayout L = make_layout(S, mkdesc(p->popv, p->popo, p->zorb, p->zorb, 1, 1, sym));
  std::vector<double> h((size_t)(nroot * p->dim), 0.0);
  for (int64_t r = 0; r < nroot; r++) {
    const Cand& c = cand[r];
    double sg;
    int64_t ad = fetch_addr(L, c.ha, c.hi, 0, 0, sp.vrta[c.ha] + c.a, c.i, 0, 0, &sg);
    if (ad < 0) return fail(DPD_EINVAL, "guess index not stored (internal)");
    h[r * p->dim + ad] = 1.0;
    if (energies) energies[r] = c.e;
  }
  DPD_HIP_TRY(hipMemcpyAsync(R, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, p->ctx->stream));
  DPD_HIP_TRY(hipStreamSynchronize(p->ctx->stream));
  return DPD_OK;
}


int dpd_problem_sigma_t(dpd_problem* p, const double* L, double* t, int64_t nvec) {
  if (!p || !L || !t || nvec < 1) return fail(DPD_EINVAL, "dpd_problem_sigma_t: bad arguments");
  if (p->stale) return fail(DPD_EINVAL, "problem inputs changed: call dpd_problem_refresh first");
  DPD_HIP_TRY(hipSetDevice(p->ctx->device));
  return dpd::problem_sigma_t(p, L, t, nvec);
}

// named device tensors of a problem: the inputs (orbital energies and
// integrals) and the intermediates finalize derives from them
static int problem_tensor(dpd_problem* p, const char* name, DevBuf** buf, int64_t* len, bool* input) {
  if (p->abcd_host && std::strcmp(name, "abcd") == 0)
    return fail(DPD_EINVAL, "abcd is out of core (dpd_problem_set_abcd_out_of_core)");
  struct {
    const char* n;
    DevBuf* b;
    int64_t l;
    bool in;
  } tab[] = {{"eorb", &p->eorb, (int64_t)p->eorb_host.size(), true}, {"iabj", &p->iabj, p->len_iabj, true},
             {"ijka", &p->ijka, p->len_ijka, true},   {"ijkl", &p->ijkl, p->len_ijkl, true},
             {"abij", &p->abij, p->len_abij, true},   {"abci", &p->abci, p->len_abci, true},
             {"abcd", &p->abcd, p->len_abcd, true},   {"fac", &p->fac, p->len_fac, false},
             {"fki", &p->fki, p->len_fki, false},     {"wkaci", &p->wkaci, p->len_iabj, false},
             {"mp2t", &p->mp2t, p->len_abij, false},  {"diag", &p->diag, p->dim, false}};
  for (auto& e : tab)
    if (std::strcmp(e.n, name) == 0) {
      *buf = e.b;
      *len = e.b->p ? e.l : 0;
      *input = e.in;
      return DPD_OK;
    }
  return fail(DPD_EINVAL, "unknown problem tensor '%s'", name);
}

int64_t dpd_problem_tensor_len(dpd_problem* p, const char* name) {
  DevBuf* b;
  int64_t len;
  bool in;
  if (!p || !name || problem_tensor(p, name, &b, &len, &in)) return -1;
  return len;
}

int dpd_problem_get(dpd_problem* p, const char* name, double* dst) {
  if (!p || !name || !dst) return fail(DPD_EINVAL, "dpd_problem_get: null argument");
  DevBuf* b;
  int64_t len;
  bool in;
  int rc = problem_tensor(p, name, &b, &len, &in);
  if (rc) return rc;
  DPD_HIP_TRY(hipSetDevice(p->ctx->device));
  if (len > 0)
    DPD_HIP_TRY(hipMemcpyAsync(dst, b->p, sizeof(double) * len, hipMemcpyDeviceToDevice, p->ctx->stream));
  return DPD_OK;
}

int dpd_problem_set(dpd_problem* p, const char* name, const double* src) {
  if (!p || !name || !src) return fail(DPD_EINVAL, "dpd_problem_set: null argument");
  DevBuf* b;
  int64_t len;
  bool in;
  int rc = problem_tensor(p, name, &b, &len, &in);
  if (rc) return rc;
  if (!in) return fail(DPD_EINVAL, "'%s' is derived from the inputs; set the inputs and refresh", name);
  if (len == 0) return fail(DPD_EINVAL, "'%s' is not used by this method", name);
  DPD_HIP_TRY(hipSetDevice(p->ctx->device));
  DPD_HIP_TRY(hipMemcpyAsync(b->p, src, sizeof(double) * len, hipMemcpyDeviceToDevice, p->ctx->stream));
  p->stale = true;
  return DPD_OK;
}

int dpd_problem_refresh(dpd_problem* p) {
  if (!p) return fail(DPD_EINVAL, "null problem");
  DPD_HIP_TRY(hipSetDevice(p->ctx->device));
  hipStream_t st = p->ctx->stream;
  DPD_HIP_TRY(hipMemcpyAsync(p->eorb_host.data(), p->eorb.p, sizeof(double) * p->eorb_host.size(),
                             hipMemcpyDeviceToHost, st));
  DPD_HIP_TRY(hipStreamSynchronize(st));
  static std::atomic<uint64_t> next_gen{1ull << 40};
  dpd::drop_pack_entries(p->ctx, id_of_problem, p->uid);
  p->uid = next_gen++;  // constant tensors changed: packed panels are rebuilt under new ids
  p->iabj32.reset();   // FP32 copies are remade on the next FP32 sigma
  p->ijka32.reset();
  p->ijkl32.reset();
  p->abci32.reset();
  p->abcd32.reset();
  p->fac32.reset();
  p->fki32.reset();
  p->wkaci32.reset();
  int rc = problem_finalize(p);
  if (rc) return rc;
  p->stale = false;
  return DPD_OK;
}

int dpd_problem_grad(dpd_problem* p, const double* L, const double* R, double scale, double* eorb_bar,
                     double* iabj_bar, double* ijka_bar, double* ijkl_bar, double* abij_bar, double* abci_bar,
                     double* abcd_bar) {
  if (!p || !L || !R) return fail(DPD_EINVAL, "dpd_problem_grad: null argument");
  if (p->stale) return fail(DPD_EINVAL, "problem inputs changed: call dpd_problem_refresh first");
  if (p->abcd_host) return fail(DPD_EINVAL, "gradient: abcd out of core is not supported");
  DPD_HIP_TRY(hipSetDevice(p->ctx->device));
  GradOut g;
  g.eorb = eorb_bar;
  g.iabj = iabj_bar;
  g.ijka = ijka_bar;
  g.ijkl = ijkl_bar;
  g.abij = abij_bar;
  g.abci = abci_bar;
  g.abcd = abcd_bar;
  return problem_grad(p, L, R, scale, g);
}
}  // extern "C"
