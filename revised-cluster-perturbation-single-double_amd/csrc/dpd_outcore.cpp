// dpd_outcore.cpp -- the particle-particle ladder with <ab||ef> out of core.
//
// The reference keeps abcd on disk when it does not fit in memory and reads it
// irrep block by irrep block inside the ladder (c2_r2abcd.f90:29-104,
// p_c2_r2abcd_h: one DGEMM per block).  Here the blocks live in pinned host
// memory and are copied over PCIe on a copy stream into two device staging
// buffers, while the GEMM of the previous block runs on the compute stream:
// block k is copied once the GEMM of block k-2 released its buffer, and its
// GEMM waits for its copy.  Per block (right-pair irrep r of abcd, ab and ef
// both of irrep r) and for the R2 block q = r (x) sym:
//     C2(ab,ij)_q += fact * abcd(ab,ef)_r R2(ef,ij)_q
// one GEMM over all vectors (N = ij pairs x vectors), accumulated (beta = 1).
// The GEMM is the gathering kernel in the problem's precision (BF16x3 runs it
// in FP32: the packed panels of a streamed operand would be rebuilt per call).
#include <algorithm>
#include <cstring>

#include "dpd_sigma.hpp"

dpd_problem::~dpd_problem() {
  if (copy_stream) (void)hipStreamSynchronize(copy_stream);
  for (auto& e : ev_copied)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ev_used)
    if (e) (void)hipEventDestroy(e);
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  if (abcd_host) (void)hipHostFree(abcd_host);
}

namespace dpd {

namespace {

// the GEMM of one abcd block against R2 block q for nvec vectors
int ooc_plan(dpd_problem* p, int64_t nvec, int64_t stride, int sym, int r, GemmPlan** out) {
  const int64_t key = (((nvec * 1000003 + stride) * 9 + sym) * 9 + r);
  for (size_t i = 0; i < p->ooc_keys.size(); i++)
    if (p->ooc_keys[i] == key) {
      *out = p->ooc_plans[i].get();
      return DPD_OK;
    }
  const Sym& S = p->ctx->sym;
  const int64_t *v = p->popv, *o = p->popo;
  const int q = S.dp[r - 1][sym - 1] + 1;  // ij irrep of the R2 block whose ef irrep is r
  const int64_t nr = S.irpdso(r, v, v, 0), nij = S.irpdso(q, o, o, 0);
  int64_t off_q = 0;
  for (int h = 1; h < q; h++) off_q += S.irpdso(S.dp[h - 1][sym - 1] + 1, v, v, 0) * S.irpdso(h, o, o, 0);
  auto plan = std::make_unique<GemmPlan>();
  GemmTask tk;
  std::memset(&tk, 0, sizeof tk);
  tk.M = (int32_t)nr;
  tk.N = (int32_t)(nij * nvec);
  tk.md[0] = (int32_t)nr;
  tk.md[1] = 1;
  tk.nd[0] = (int32_t)std::max<int64_t>(nij, 1);
  tk.nd[1] = (int32_t)nvec;
  tk.sc_m[0] = 1;
  tk.sc_n[0] = nr;
  tk.sc_n[1] = stride;
  tk.c_off = off_q;
  GemmSeg sg;
  std::memset(&sg, 0, sizeof sg);
  sg.K = (int32_t)nr;
  sg.kd[0] = (int32_t)std::max<int64_t>(nr, 1);
  sg.kd[1] = 1;
  sg.sa_m[0] = 1;
  sg.sa_k[0] = nr;
  sg.sb_k[0] = 1;
  sg.sb_n[0] = nr;
  sg.sb_n[1] = stride;
  sg.b_off = off_q;
  sg.ktiles = (int32_t)((nr + 15) / 16);
  tk.nseg = 1;
  tk.ktiles_total = sg.ktiles;
  tk.ksplit = 1;
  plan->cfg = choose_gemm_cfg({tk}, p->ctx->gemm_cfg);
  if (plan->cfg == GEMM_64x192) plan->cfg = GEMM_64x128;
  if (plan->cfg == GEMM_128x64D) plan->cfg = GEMM_128x64;
  tk.tiles_m = (int32_t)((nr + gemm_bm(plan->cfg) - 1) / gemm_bm(plan->cfg));
  tk.tiles_n = (int32_t)((tk.N + gemm_bn(plan->cfg) - 1) / gemm_bn(plan->cfg));
  plan->tasks.push_back(tk);
  plan->segs.push_back(sg);
  plan->nwg = (nr > 0 && tk.N > 0) ? tk.tiles_m * tk.tiles_n : 0;
  plan->flops = 2 * nr * nr * nij * nvec;
  int rc = plan->upload();
  if (rc) return rc;
  *out = plan.get();
  p->ooc_plans.push_back(std::move(plan));
  p->ooc_keys.push_back(key);
  return DPD_OK;
}

}  // namespace

int problem_pp_ladder_streamed(dpd_problem* p, const double* R2, double* C2, int64_t nvec, int64_t stride, int sym,
                               double fact) {
  if (!p->abcd_host) return fail(DPD_EINVAL, "abcd is not out of core (internal)");
  dpd_ctx* c = p->ctx;
  const Sym& S = c->sym;
  hipStream_t st = c->stream;
  const int prec = p->prec == PREC_F64 ? PREC_F64 : PREC_F32;
  int64_t off_r = 0;
  int k = 0;
  for (int r = 1; r <= S.nirrep; r++) {
    const int64_t nr = S.irpdso(r, p->popv, p->popv, 0);
    const int64_t bytes = sizeof(double) * nr * nr;
    if (nr == 0) continue;
    GemmPlan* plan;
    int rc = ooc_plan(p, nvec, stride, sym, r, &plan);
    if (rc) return rc;
    const int b = k & 1;
    // the buffer is free once the GEMM of block k-2 is done
    DPD_HIP_TRY(hipStreamWaitEvent(p->copy_stream, p->ev_used[b], 0));
    DPD_HIP_TRY(hipMemcpyAsync(p->abcd_stage[b].p, p->abcd_host + off_r, bytes, hipMemcpyHostToDevice,
                               p->copy_stream));
    DPD_HIP_TRY(hipEventRecord(p->ev_copied[b], p->copy_stream));
    DPD_HIP_TRY(hipStreamWaitEvent(st, p->ev_copied[b], 0));
    if ((rc = run_gemm(c, *plan, p->abcd_stage[b].p, 0, R2, 0, C2, fact, 1.0, prec))) return rc;
    DPD_HIP_TRY(hipEventRecord(p->ev_used[b], st));
    off_r += nr * nr;
    k++;
  }
  return DPD_OK;
}

}  // namespace dpd

using namespace dpd;

extern "C" int dpd_problem_set_abcd_out_of_core(dpd_problem* p, int on) {
  if (!p) return fail(DPD_EINVAL, "null problem");
  if (p->method != 2) return fail(DPD_EINVAL, "abcd exists only for method 2");
  dpd_ctx* c = p->ctx;
  DPD_HIP_TRY(hipSetDevice(c->device));
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  const size_t bytes = sizeof(double) * (size_t)p->len_abcd;
  if (on && !p->abcd_host) {
    if (g_capturing) return fail(DPD_EINVAL, "abcd residency changed during graph capture");
    void* h = nullptr;
    DPD_HIP_TRY(hipHostMalloc(&h, std::max<size_t>(bytes, 8), hipHostMallocDefault));
    if (bytes) {
      hipError_t e = hipMemcpy(h, p->abcd.p, bytes, hipMemcpyDeviceToHost);
      if (e != hipSuccess) {
        (void)hipHostFree(h);
        return fail(DPD_EHIP, "abcd to host: %s", hipGetErrorString(e));
      }
    }
    // the largest irrep block sizes the staging buffers
    const Sym& S = c->sym;
    int64_t mx = 1;
    for (int r = 1; r <= S.nirrep; r++) mx = std::max(mx, S.irpdso(r, p->popv, p->popv, 0));
    int rc;
    if ((rc = p->abcd_stage[0].alloc(sizeof(double) * mx * mx)) || (rc = p->abcd_stage[1].alloc(sizeof(double) * mx * mx))) {
      (void)hipHostFree(h);
      return rc;
    }
    if (!p->copy_stream) {
      DPD_HIP_TRY(hipStreamCreateWithFlags(&p->copy_stream, hipStreamNonBlocking));
      for (int b = 0; b < 2; b++) {
        DPD_HIP_TRY(hipEventCreateWithFlags(&p->ev_copied[b], hipEventDisableTiming));
        DPD_HIP_TRY(hipEventCreateWithFlags(&p->ev_used[b], hipEventDisableTiming));
      }
    }
    for (int b = 0; b < 2; b++) DPD_HIP_TRY(hipEventRecord(p->ev_used[b], c->stream));
    p->abcd_host = (double*)h;
    // the device copy and its packed panels go
    drop_pack_entries(c, [](uint64_t id, uint64_t key) { return id == key; }, p->uid * 64 + 9 + 1);
    p->abcd.reset();
    p->abcd32.reset();
    bump_mem_gen();
  } else if (!on && p->abcd_host) {
    int rc = p->abcd.alloc(std::max<size_t>(bytes, 8));
    if (rc) return rc;
    if (bytes) DPD_HIP_TRY(hipMemcpy(p->abcd.p, p->abcd_host, bytes, hipMemcpyHostToDevice));
    DPD_HIP_TRY(hipStreamSynchronize(p->copy_stream));
    (void)hipHostFree(p->abcd_host);
    p->abcd_host = nullptr;
    p->abcd_stage[0].reset();
    p->abcd_stage[1].reset();
    if (p->prec != PREC_F64 || p->abcd32.p == nullptr) p->abcd32.reset();
    bump_mem_gen();
  }
  return DPD_OK;
}
