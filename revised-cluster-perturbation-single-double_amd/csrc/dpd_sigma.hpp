// dpd_sigma.hpp -- device-resident problem (integrals, intermediates, diagonal)
// and the batched sigma builders.
#pragma once
#include <memory>
#include <vector>

#include "dpd_engine.hpp"

namespace dpd {
// Davidson workspace of a problem, kept between solves (grow-only): the
// subspace (b, Ab and their collapsed copies), the per-root vectors, the
// reduction scratch and the pinned host staging of the small transfers
struct DavWork {
  DevBuf B, AB, NB, NAB, AQ, E, P, coef, part, partA, partB, partC, partD, dots, amaxp, amaxpi, amaxv, amaxi, wdev, tri,
      idx, cgs, gram, qpart, ppart, rstat;
  void* pin_out = nullptr;
  void* pin_in = nullptr;
  size_t pin_bytes = 0;
  ~DavWork() {
    if (pin_out) (void)hipHostFree(pin_out);
    if (pin_in) (void)hipHostFree(pin_in);
  }
};
}  // namespace dpd

struct dpd_problem {
  dpd_ctx* ctx = nullptr;
  dpd_synth_spec spec;
  int method = 2;
  int symR = 1;
  int prec = dpd::PREC_F64;
  int64_t popo[8], popv[8], zorb[8];
  int64_t l1 = 0, l2 = 0, dim = 0;
  // device tensors (FP64)
  dpd::DevBuf eorb, iabj, ijka, ijkl, abij, abci, abcd, fac, fki, wkaci, mp2t, diag;
  // FP32 copies used by the mixed-precision contractions
  dpd::DevBuf iabj32, ijka32, ijkl32, abci32, abcd32, fac32, fki32, wkaci32;
  std::vector<double> eorb_host;
  int64_t len_iabj = 0, len_ijka = 0, len_ijkl = 0, len_abij = 0, len_abci = 0, len_abcd = 0;
  int64_t len_fac = 0, len_fki = 0;
  int64_t sigma_flops = 0;  // algorithmic GEMM flops of one sigma (one vector)
  uint64_t uid = 0;         // process-unique id: constant tensors are packed once under it
  bool stale = false;       // an input was set after the last finalize
  // <ab||ef> out of core: a pinned host copy streamed irrep block by irrep block
  // through two device staging buffers (abcd itself is then freed)
  double* abcd_host = nullptr;
  dpd::DevBuf abcd_stage[2];
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_copied[2] = {nullptr, nullptr}, ev_used[2] = {nullptr, nullptr};
  std::vector<std::unique_ptr<dpd::GemmPlan>> ooc_plans;  // per (nvec, stride, symmetry, block)
  std::vector<int64_t> ooc_keys;
  std::unique_ptr<dpd::DavWork> dav;  // Davidson workspace (grow-only, freed with the problem)
  ~dpd_problem();
};

namespace dpd {
int problem_sigma(dpd_problem* p, const double* R, double* sigma, int64_t nvec);
int problem_make_f32(dpd_problem* p);
int problem_sigma_t(dpd_problem* p, const double* L, double* t, int64_t nvec);
int problem_sigma_group(dpd_problem* const* ps, const double* const* R, double* const* S, const int64_t* nvec,
                        int n);
// c2_r2abcd (pp-ladder) with <ab||ef> streamed from the pinned host copy:
// C2(ab,ij) += fact * sum_{e<f} <ab||ef> R2(ef,ij) for nvec vectors of symmetry
// sym (R2 / C2 of vector v at R2 + v*stride / C2 + v*stride)
int problem_pp_ladder_streamed(dpd_problem* p, const double* R2, double* C2, int64_t nvec, int64_t stride, int sym,
                               double fact);
}
