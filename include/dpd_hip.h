/*
 * dpd_hip.h -- C-ABI of the MI355X-native DPD tensor-contraction engine.
 *
 * This is the drop-in boundary for the reference's lib_cc contract/sort call
 * surface and the libsrint engines it runs on.  Every entry point takes plain
 * pointers and sizes; tensor data pointers are DEVICE pointers (HBM, FP64) on
 * the context's device, exactly as the reference's engines take device
 * pointers when igpu=1 (libsrint/vv2v_dpd_444.f90:45-46).  Host-side inputs
 * are marked as such.
 *
 * Conventions (same as the reference):
 *   irreps are 1-based; population arrays have 8 entries; integers are int64
 *   (the reference builds with -i8); tensors are column-major DPD, blocks
 *   ordered by right-pair irrep, left pair fastest; pair packing iop:
 *   0 = p<q (antisymmetric), 1 = full, 2 = p<=q (symmetric).
 *   Label strings are NOT NUL-terminated: 4 characters for a 4-index tensor,
 *   2 for a 2-index one.
 * Errors: every function returns 0 on success or a negative DPD_E* code; the
 * reference's `write(6,*) ...; stop` becomes an error return with the message
 * available from dpd_last_error() (per host thread).
 * Threading: a context is not re-entrant; use one context per host thread.
 * All device work is enqueued on the context's stream (dpd_ctx_set_stream);
 * calls return when the work is enqueued unless stated otherwise.
 */
#ifndef DPD_HIP_H
#define DPD_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DPD_OK = 0,
  DPD_EINVAL = -1,     /* illegal argument (iop, anti, label string, sizes) */
  DPD_ESYM = -2,       /* symmetry mismatch (e.g. Sym1*Sym2 != Sym3) */
  DPD_EPOP = -3,       /* contracted-index populations differ */
  DPD_EIOP = -4,       /* compressed pair used where it must be full (IOP barrier) */
  DPD_ENOMEM = -5,     /* device allocation failed */
  DPD_EHIP = -6,       /* HIP runtime error */
  DPD_ENODEV = -7,     /* no usable gfx950 device */
  DPD_ENOTCONV = -8    /* Davidson: not converged within the cycle limit (results still written) */
};

typedef struct dpd_ctx dpd_ctx;

/* Descriptor of one DPD tensor: restates `type opt` (lib_cc/mod_operator.f90:8-25)
 * with the storage pointer (`arr`) as a device pointer.  For 2-index operands
 * (vv2v_424's v2, vv2v_442's v3) only pop[0..1], iop[0] and sym are read. */
typedef struct {
  int64_t pop[4][8]; /* popp, popq, popr, pops */
  int64_t iop[2];    /* ioppq, ioprs */
  int64_t sym;       /* 1-based irrep of the tensor */
  double *data;      /* device pointer, FP64 (FP32 under dpd_ctx_set_data_precision(ctx, 4)),
                        length dpd_tensor_len() */
} dpd_tensor;

/* ---- context ------------------------------------------------------------ */
/* Replaces the global mod_symm state (lib_mod/mod_symm.f): nirrep and the
 * 8x8 direct-product table (host, row-major dirprd[i*8+j] = dirprd(i+1,j+1)).
 * device = -1 creates a host-only context: metadata calls work, every call that
 * needs the GPU returns DPD_ENODEV. */
int dpd_ctx_create(dpd_ctx **ctx, int64_t nirrep, const int64_t *dirprd64, int device);
int dpd_ctx_destroy(dpd_ctx *ctx);
/* hipStream_t as void*; NULL = the default stream */
int dpd_ctx_set_stream(dpd_ctx *ctx, void *stream);
int dpd_ctx_synchronize(dpd_ctx *ctx);
/* GEMM kernels of every later plan.  cfg: -1 automatic (default: tile shape of
 * the packed-panel kernel picked by timing every shape on a plan's first call),
 * -2 automatic by a padding cost model without timing, 0 the first-generation
 * 64x64 kernel, 1..6 64x64, 128x64, 64x128, 128x128, 32x32, 64x192 tiles (64x192
 * packed-panel kernel only), 7 the BF16x3 128x64 tile with A read straight into
 * registers (other precisions: 128x64).
 * packing: 0 operands gathered inside the GEMM, 1 (default) packed panels when
 * an operand is a problem's constant tensor (packed once and kept), 2 packed
 * panels always (BF16x3 always uses panels).  Clears the plan caches.  No
 * counterpart in the reference (its GEMMs are cuBLAS/BLAS calls). */
int dpd_ctx_set_gemm_config(dpd_ctx *ctx, int cfg, int packing);
/* No reference counterpart: split-K of the packed-panel GEMMs.  mode -1: timed
 * together with the tile shape when the tile is autotuned, else on; 0: off (one
 * workgroup per output tile); 1: long-K tasks split while a launch has fewer
 * than 1024 tiles, partial sums reduced by a second kernel.  Clears the plans. */
int dpd_ctx_set_split_k(dpd_ctx *ctx, int mode);
/* libsrint/readinp.f90:15,120-127 (the mod_iop `prec` switch that routes
 * vv2v_dpd_444.f90:206-229, sstgenso.f90:35-42 and xgemmso.f:11-22 to the _s/_sg
 * code).  prec 8 (default): tensor data are FP64.  prec 4: the `data` pointers
 * given to the engines (dpd_vv2v_dpd_444/424/442/323), the sort kernels
 * (dpd_iexpso ... dpd_sstgenso) and dpd_xgemm point at FP32 arrays (the
 * reference's real*4 storage, vv2v_dpd_444_sg.f90:20-21); their GEMMs read FP32
 * and run FP32 MFMA, and the result update v3 = fact2*v3 + fact1*C is formed
 * in FP64 and rounded to FP32 once (vv2v_dpd_444_sg.f90:431,445).  Sorts move
 * FP32 values exactly (iassymso's one sum or difference is rounded once).  The
 * contract surface (dpd_vv2v_444 ...) stays FP64, as lib_cc always runs with
 * prec = 8 (lib_cc/readmodule.f90:64), and returns DPD_EINVAL under prec 4;
 * dpd_problem_* own their storage and are not affected. */
int dpd_ctx_set_data_precision(dpd_ctx *ctx, int prec);
int dpd_ctx_data_precision(const dpd_ctx *ctx);
const char *dpd_last_error(void);
/* bytes of device scratch the context currently holds (grows on demand) */
size_t dpd_ctx_workspace_bytes(const dpd_ctx *ctx);
/* Kernel timing (the reference's cpu_time prints, on the device): when enabled,
 * every tagged GEMM stage of the sigma builders is bracketed by HIP events on
 * the context stream (up to max_records launches; packed-panel GEMMs: the GEMM
 * kernel alone, without the packing).  The report aggregates per
 * "tag:precision:t<tile shape>" {count, ms, flops_total, bytes_total, prec, cfg,
 * packed} as JSON and clears the records. */
int dpd_profile_enable(dpd_ctx *ctx, int on, int max_records);
int dpd_profile_report(dpd_ctx *ctx, char *json, int64_t len);

/* ---- metadata (host, no device work) ---------------------------------- */
/* libsrint/irpdso.f:1-51 */
int64_t dpd_irpdso(const dpd_ctx *ctx, int64_t sym, const int64_t *popp, const int64_t *popq,
                   int64_t iop);
/* libsrint/idsymsoc.f:1-29 */
int64_t dpd_idsymsoc(const dpd_ctx *ctx, int64_t irrepx, const int64_t *popp,
                     const int64_t *popq, int64_t ioppq, const int64_t *popr,
                     const int64_t *pops, int64_t ioprs);
/* libsrint/isymoffso.f:1-53 (1-based offset) */
int64_t dpd_isymoffso(const dpd_ctx *ctx, int64_t irrepq1, int64_t irrep, const int64_t *popp,
                      const int64_t *popq, int64_t iop);
/* length of a descriptor (lib_cc/mod_operator.f90:93, init_opt) */
int64_t dpd_tensor_len(const dpd_ctx *ctx, const dpd_tensor *t);
/* lib_cc/getidx.f90: 1-based linear index -> info8 = (irrp,irrq,irrr,irrs,p,q,r,s) */
int dpd_getidx(const dpd_ctx *ctx, int64_t idx, int64_t *info8, const dpd_tensor *t);
/* lib_cc/putidx.f90: info8 -> 1-based linear index (0 if not stored) */
int64_t dpd_putidx(const dpd_ctx *ctx, const int64_t *info8, const dpd_tensor *t);

/* ---- engines: v3 = fact1 * (v1 (x) v2) + fact2 * v3 -------------------- */
/* libsrint/vv2v_dpd_444.f90:55-232 (routes to the fast path or the fallback) */
int dpd_vv2v_dpd_444(dpd_ctx *ctx, const dpd_tensor *v1, const char *type1,
                     const dpd_tensor *v2, const char *type2, dpd_tensor *v3,
                     const char *type3, double fact1, double fact2);
/* libsrint/vv2v_dpd_424.f90:21-  (v2 is 2-index) */
int dpd_vv2v_dpd_424(dpd_ctx *ctx, const dpd_tensor *v1, const char *type1,
                     const dpd_tensor *v2, const char *type2, dpd_tensor *v3,
                     const char *type3, double fact1, double fact2);
/* libsrint/vv2v_dpd_442.f90:22-  (v3 is 2-index) */
int dpd_vv2v_dpd_442(dpd_ctx *ctx, const dpd_tensor *v1, const char *type1,
                     const dpd_tensor *v2, const char *type2, dpd_tensor *v3,
                     const char *type3, double fact1, double fact2);

/* libsrint/vv2v_dpd_323.f90:22-145: 3-index x 2-index -> 3-index
 * (e.g. 'pqm' x 'mr' -> 'pqr').  Rank-3 tensors use pop[0..2], iop[0] (the pair
 * of their first two indices) and sym; v2 uses pop[0..1] and must have iop 1.
 * The third index of v1 is the contracted one.  A pair kept intact may be packed
 * (iop 0/2); when it is reordered or split, v1 and v3 need iop 1. */
int dpd_vv2v_dpd_323(dpd_ctx *ctx, const dpd_tensor *v1, const char *type1,
                     const dpd_tensor *v2, const char *type2, dpd_tensor *v3,
                     const char *type3, double fact1, double fact2);

/* ---- contract surface: v3 += fact * P{v1 (x) v2} (lib_cc/mod_contract.f90) */
/* Scratch (the reference's iCore/maxCore) comes from the context's workspace. */
/* mod_contract.f90:10-162, anti = 0..3 antisymmetrises none/pq/rs/both of v3 */
int dpd_vv2v_444(dpd_ctx *ctx, const dpd_tensor *v1, const dpd_tensor *v2, dpd_tensor *v3,
                 double fact, const char *type1, const char *type2, const char *type3,
                 int64_t anti);
/* mod_contract.f90:166-180 (labels of 2 characters each) */
int dpd_vv2v_222(dpd_ctx *ctx, const dpd_tensor *v1, const dpd_tensor *v2, dpd_tensor *v3,
                 double fact, const char *type1, const char *type2, const char *type3);
/* mod_contract.f90:184-199 (type1 4 chars, type2/type3 2 chars) */
int dpd_vv2v_422(dpd_ctx *ctx, const dpd_tensor *v1, const dpd_tensor *v2, dpd_tensor *v3,
                 double fact, const char *type1, const char *type2, const char *type3);
/* mod_contract.f90:203-219 (type1/type2 2 chars, type3 4 chars) */
int dpd_vv2v_224(dpd_ctx *ctx, const dpd_tensor *v1, const dpd_tensor *v2, dpd_tensor *v3,
                 double fact, const char *type1, const char *type2, const char *type3,
                 int64_t anti);
/* mod_contract.f90:223-368, anti = 0..2 */
int dpd_vv2v_424(dpd_ctx *ctx, const dpd_tensor *v1, const dpd_tensor *v2, dpd_tensor *v3,
                 double fact, const char *type1, const char *type2, const char *type3,
                 int64_t anti);
/* mod_contract.f90:372-455 */
int dpd_vv2v_442(dpd_ctx *ctx, const dpd_tensor *v1, const dpd_tensor *v2, dpd_tensor *v3,
                 double fact, const char *type1, const char *type2, const char *type3);

/* ---- sort kernels (device in/out, FP64) -------------------------------- */
/* libsrint/iexpso.f90:24-120 */
int dpd_iexpso(dpd_ctx *ctx, const double *in, double *out, const int64_t *popp,
               const int64_t *popq, int64_t ioppq, int64_t iexppq, const int64_t *popr,
               const int64_t *pops, int64_t ioprs, int64_t iexprs, int64_t sym);
/* libsrint/icompso.f:1-93 */
int dpd_icompso(dpd_ctx *ctx, const double *in, double *out, const int64_t *popp,
                const int64_t *popq, int64_t ioppq, int64_t icpq, const int64_t *popr,
                const int64_t *pops, int64_t ioprs, int64_t icrs, int64_t sym);
/* libsrint/isymtrso.f:1-120 */
int dpd_isymtrso(dpd_ctx *ctx, const double *in, double *out, const int64_t *popp,
                 const int64_t *popq, int64_t ioppq, int64_t icpq, const int64_t *popr,
                 const int64_t *pops, int64_t ioprs, int64_t icrs, int64_t sym);
/* libsrint/itranspso.f:1-53 */
int dpd_itranspso(dpd_ctx *ctx, const double *in, double *out, const int64_t *popp,
                  const int64_t *popq, int64_t ioppq, const int64_t *popr, const int64_t *pops,
                  int64_t ioprs, int64_t sym);
/* libsrint/iassymso.f:1-142: antisymmetrise (iop 0,1,2,4,5) or symmetrise (iop 3)
 * a pair of a 4-index tensor and compress where the output pair is packed. */
int dpd_iassymso(dpd_ctx *ctx, const double *in, double *out, int64_t sym, const int64_t *popp,
                 const int64_t *popq, const int64_t *popr, const int64_t *pops, int64_t iop);
/* libsrint/sstgenso.f90:22-45 (sstgenso1), type4 e.g. "3142" */
int dpd_sstgenso(dpd_ctx *ctx, const double *win, double *wout, const int64_t *pop1,
                 const int64_t *pop2, const int64_t *pop3, const int64_t *pop4, int64_t irrepx,
                 const char *type4);

/* ---- dense GEMM: libsrint/xgemmso.f:1-96 (XGEMMD / XGEMMDG / XGEMMSG) ---- */
/* column-major C = alpha*op(A)*op(B) + beta*C on device FP64 arrays; prec 0 runs
 * FP64 MFMA, prec 1 FP32 MFMA (operands rounded to FP32, FP32 accumulation),
 * prec 2 BF16x3 (operands split into bf16 hi + lo, three BF16 MFMAs, FP32
 * accumulation); synchronises the stream before returning. */
int dpd_xgemm(dpd_ctx *ctx, const char *transa, const char *transb, int64_t m, int64_t n,
              int64_t k, double alpha, const double *A, int64_t lda, const double *B,
              int64_t ldb, double beta, double *C, int64_t ldc, int prec);

/* No counterpart in the reference.  dpd_xgemm for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): the same arguments and result, but no hipMalloc,
 * hipFree or stream synchronisation.  The task tables and the packed operand
 * panels (BF16x3, or packing 2) live in one static 64 MiB device array per GPU,
 * held until the process ends; a call whose panels do not fit returns
 * DPD_EINVAL (FP64/FP32 on the gather kernel need only the tables and take any
 * size).  Long-K tiles are not split over workgroups, so results may differ from
 * dpd_xgemm's in the last bits.  Every call on one GPU uses that GPU's array:
 * calls on one stream are ordered, but calls that overlap on one GPU (other
 * streams, other threads or contexts) race. */
int dpd_xgemm_noalloc(dpd_ctx *ctx, const char *transa, const char *transb, int64_t m,
                      int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                      const double *B, int64_t ldb, double beta, double *C, int64_t ldc,
                      int prec);

/* ---- Davidson vector primitives (device vectors, FP64) ----------------- */
/* result written to a host double (synchronises the stream) */
int dpd_dot(dpd_ctx *ctx, int64_t n, const double *x, const double *y, double *result);
int dpd_nrm2(dpd_ctx *ctx, int64_t n, const double *x, double *result);
/* idamax: 1-based index of the first max |x_i| and the signed value there */
int dpd_amax(dpd_ctx *ctx, int64_t n, const double *x, int64_t *index, double *value);
int dpd_axpy(dpd_ctx *ctx, int64_t n, double a, const double *x, double *y);
int dpd_scal(dpd_ctx *ctx, int64_t n, double a, double *x);
/* Davidson preconditioner (lib_cc/mod_davidson.f90:417-428):
 * p_i = e_i / (w - diag_i) with |w - diag_i| clamped to 1e-4 (sign kept) */
int dpd_precondition(dpd_ctx *ctx, int64_t n, double w, const double *diag, const double *e,
                     double *p);

/* ---- synthetic Hamiltonians (host; the pyscf2cc/getHF data path) -------- */
/* Seeded synthetic spin-free DPD Hamiltonian: see dpd_synth_* in the source. */
typedef struct {
  int64_t nirrep;
  int64_t nso[8];    /* spatial orbitals per irrep */
  int64_t nalpha, nbeta;
  uint64_t seed;
  /* outputs (filled by dpd_synth_layout) */
  int64_t popa[8], popb[8], vrta[8], vrtb[8];
  int64_t popo[8], popv[8]; /* spin-free occupied / virtual per irrep (alpha then beta) */
} dpd_synth_spec;
int dpd_synth_layout(dpd_synth_spec *spec);
/* Fill one integral tensor in host memory.  which: "fij","fab","fia","ijkl","ijka",
 * "abij","iabj","abci","abcd" (layouts of CISdoo.SD.SF/allocate_mem.f90:17-54 with
 * iop as given); eorb (host) receives the spin-free orbital energies when
 * which == "eorb" (length sum(popo)+sum(popv), occupied irreps then virtual). */
int dpd_synth_fill(const dpd_synth_spec *spec, const char *which, int64_t iop, double *host_out,
                   int64_t len);

/* ---- sigma builders and Davidson (device-resident) --------------------- */
typedef struct dpd_problem dpd_problem;
/* method: 0 = SF-CIS (EE-CIS.SF/calAR.f90), 1 = CIS(Doo) (CISdoo.SD.SF/calAR.f90),
 *         2 = SF-CPSd = CIS(Doo) + doubles-doubles first-order coupling
 *             (EE-CPSd.SF/calC2_VR2.f90)
 * precision: 0 = FP64 MFMA, 1 = FP32 MFMA (FP32 operands, FP32 accumulate),
 *            2 = BF16x3 split MFMA (FP32 accumulate) */
int dpd_problem_create(dpd_ctx *ctx, const dpd_synth_spec *spec, int method, int64_t symR,
                       dpd_problem **prob);
int dpd_problem_destroy(dpd_problem *prob);
int64_t dpd_problem_dim(const dpd_problem *prob);
int dpd_problem_set_precision(dpd_problem *prob, int precision);
/* sigma = J * R for nvec vectors stored contiguously (device, stride dim) */
int dpd_problem_sigma(dpd_problem *prob, const double *R, double *sigma, int64_t nvec);
/* diagonal preconditioner (device, length dim): (OED1, OED2) */
int dpd_problem_diag(dpd_problem *prob, double *diag);
/* spin-flip alpha->beta initial guesses (lib_cc/preDavidson.f90:120-223);
 * writes nroot unit vectors (device, contiguous) and their OED energies (host) */
int dpd_problem_guess(dpd_problem *prob, int64_t nroot, double *R, double *energies);

/* ---- adjoint: transposed operator and gradients -----------------------
 * No single reference routine: the reference has no gradient code for these
 * methods.  The transposed operator is built from the same contraction terms as
 * dpd_problem_sigma (CISdoo.SD.SF/calAR.f90:41-77, EE-CPSd.SF/calC2_VR2.f90)
 * with operand roles exchanged; the gradient contracts the two vectors into the
 * shape of each integral block (the transpose-contraction rule). */
/* t = J^T L for nvec vectors (device, stride dim) */
int dpd_problem_sigma_t(dpd_problem *prob, const double *L, double *t, int64_t nvec);
/* named device tensors: inputs "eorb" (spin-free orbital energies, occupied
 * irreps then virtual), "iabj", "ijka", "ijkl", "abij", "abci", "abcd" (layouts
 * of dpd_synth_fill) and, read only, the intermediates "fac", "fki", "wkaci",
 * "mp2t", "diag".  Length -1: unknown name, 0: not used by the method. */
int64_t dpd_problem_tensor_len(dpd_problem *prob, const char *name);
int dpd_problem_get(dpd_problem *prob, const char *name, double *dst_device);
/* overwrite an input; the problem refuses sigma/Davidson until refreshed */
int dpd_problem_set(dpd_problem *prob, const char *name, const double *src_device);
/* recompute the diagonal, MP2 amplitudes and the CIS(Doo) intermediates from
 * the current inputs (the steps of dpd_problem_create after the integrals) */
int dpd_problem_refresh(dpd_problem *prob);
/* scale * d(L . J(theta) R)/d(theta) accumulated (+=) into the gradient buffers
 * (device, the layouts and lengths of the named inputs; a buffer the method does
 * not use may be NULL).  theta are the stored elements of the inputs, so for a
 * packed pair (p<q) the derivative is with respect to the stored element.
 * With L, R the left and right eigenvectors of one root this is L^T R times the
 * gradient of its eigenvalue (first-order perturbation theory of a
 * non-symmetric matrix).  The MP2 amplitudes and intermediates are recomputed
 * from the inputs, not stored. */
int dpd_problem_grad(dpd_problem *prob, const double *L, const double *R, double scale, double *eorb_bar,
                     double *iabj_bar, double *ijka_bar, double *ijkl_bar, double *abij_bar,
                     double *abci_bar, double *abcd_bar);

typedef struct {
  int64_t nroot;
  int64_t max_dim;   /* davidson_dim (default 20) */
  int64_t max_cyc;   /* davidson_cyc (default 50) */
  double conv;       /* max-abs residual tolerance, 10^-davidson_conv */
  int lowest;        /* 0: root tracking by overlap (default), 1: lowest */
  int final_check;   /* 1: recompute |A R - w R|_inf for every root at the end */
  int transpose;     /* 1: solve with the transposed operator J^T (left eigenvectors) */
  /* outputs */
  int64_t iterations;
  int64_t sigma_calls;
  int64_t skipped;   /* new directions dropped by the norm < 1e-6 test (:437) */
  double residual[64];
  double final_residual[64];
  /* 1: the sigma step of every iteration is replayed from a hipGraph captured
   * per vector count (on a private stream ordered with the context's; staged
   * through fixed device vectors; eager while the context profiler is on) */
  int graph;
  int64_t graph_replays; /* out: sigma steps replayed from a graph */
} dpd_davidson_opts;
/* Subspace eigensolve (the dgeev call of lib_cc/mod_davidson.f90:186), host FP64:
 * eigenvalues (wr, wi) and unit-norm right eigenvectors (columns of V, column
 * major) of a real nonsymmetric n x n matrix given row-major. */
int dpd_small_eig(int64_t n, const double *A_rowmajor, double *wr, double *wi, double *V_colmajor);
/* Non-Hermitian Davidson (lib_cc/mod_davidson.f90:40-561) on the problem's
 * operator; R (device, nroot*dim) holds the guesses on entry and the
 * eigenvectors on exit; w (host, nroot) receives the eigenvalues. */
int dpd_davidson(dpd_problem *prob, dpd_davidson_opts *opts, double *R, double *w);

/* ---- batches of independent problems ------------------------------------
 * No single reference routine: the reference solves one problem per run.
 * sigma of n problems (same context, method and precision; sizes differ):
 * every contraction of all of them runs in one grouped GEMM launch (tile shape
 * timed for the group on first use).  R[i], sigma[i]: device, nvec[i] vectors. */
int dpd_problem_sigma_group(dpd_problem *const *probs, const double *const *R, double *const *sigma,
                            const int64_t *nvec, int64_t n);
/* n Davidson solves (lib_cc/mod_davidson.f90:40-561 each) advanced in lockstep:
 * the sigma requests of one step are applied together.  flags bit 0: grouped
 * GEMM launches (else one problem after another); bit 1: the grouped sigma step
 * is captured once per shape in a hipGraph and replayed.  opts[i], R[i], w[i]
 * as for dpd_davidson.  stats (host, 4, may be NULL): batched sigma steps, graph
 * replays, graphs captured, captures abandoned (run eagerly instead).  Returns
 * DPD_ENOTCONV if any member did not converge. */
int dpd_davidson_batch(dpd_problem *const *probs, int64_t n, dpd_davidson_opts *opts, double *const *R,
                       double *const *w, int64_t flags, int64_t *stats);

/* Out-of-core <ab||ef> (the reference's c2_r2abcd.f90:29-104 / p_c2_r2abcd_h
 * path for integrals beyond memory): on = 1 moves abcd to pinned host memory
 * and the pp-ladder streams it irrep block by irrep block through two device
 * staging buffers, the copy of one block overlapping the GEMM of the previous;
 * on = 0 brings it back.  While out of core, dpd_problem_sigma_t,
 * dpd_problem_grad, grouped sigma and get/set of "abcd" return DPD_EINVAL. */
int dpd_problem_set_abcd_out_of_core(dpd_problem *prob, int on);

/* ---- perturbation drivers (CPSd.SF, EE-CPSd.SF) ---------------------------
 * Both need a problem of method 2 (every integral block).  DIIS settings as in
 * lib_cc/mod_common.f90:15-18 (diis_dim 10, diis_cyc 50, diis_max_err 1e-7 when
 * left 0); the DIIS history (libsrint/dosodiis2_d.f) stays in device memory. */
typedef struct {
  int64_t diis_dim;      /* DIIS subspace size */
  int64_t diis_cyc;      /* iteration cap of each DIIS solve */
  double diis_max_err;   /* |largest element of the extrapolated residual| to stop */
  double energy[3];      /* out (ground state): E^(2), E^(3), E^(4) */
  int64_t iterations[2]; /* out (ground state): DIIS iterations of the T1 equations, orders 2, 3 */
} dpd_cpsd_opts;
/* Length of one order of the ground-state amplitudes: |T1| + |T2(a<b,i<j)| of
 * the totally symmetric irrep. */
int64_t dpd_problem_cpsd_len(const dpd_problem *prob);
/* CPSd.SF/run.f90 (run1/run2/run3.f90): ground-state CPS(D) amplitudes
 * T^(1..3) and correlation energies E^(2..4).  T (device, may be NULL): receives
 * 3 x dpd_problem_cpsd_len() doubles, (T1 | T2) of orders 1, 2, 3 -- the f_T
 * file's records in order. */
int dpd_problem_cpsd(dpd_problem *prob, dpd_cpsd_opts *opts, double *T);
/* grCPSd.SF/run.f90 (run1-3.f90, solver.f90): ground-state grCPS(D), the
 * CIS(Doo) Jacobian as zeroth order -- T^(n) from A0 T^(n) + C^(n) = 0 by DIIS
 * on (T1 | T2) at n = 1, 2, 3, with the [[V,T^(1)],T^(1)] terms of
 * calL1_VTT.f90/calL2_VTT.f90 at n = 3; E^(n+1) = <abij| T2^(n) + [T1 T1]^(n)>
 * (calEcorr.f90, effectT2.f90).  T as for dpd_problem_cpsd (the iCPSd_T_
 * records); opts->energy: E^(2..4), opts->iterations: DIIS iterations of orders
 * 2, 3; iterations (host, 3, may be NULL): those of orders 1, 2, 3.  Needs
 * <ab||ef> on the device. */
int dpd_problem_grcpsd(dpd_problem *prob, dpd_cpsd_opts *opts, double *T, int64_t *iterations);
/* EE-CPSd.SF/run.f90 + run_order.f90: EE-CPS(D) excitation energies of nroot
 * CIS roots of the problem's symmetry, order by order.  T: the ground-state
 * amplitudes of dpd_problem_cpsd (device); w0 (host, nroot): CIS energies;
 * L (device, nroot x |R1|): CIS vectors (f_CIS).  w (host, 3*nroot): w^(1),
 * w^(2), w^(3) per root (the total is w0 + w^(2) + w^(3)); iterations (host,
 * nroot, may be NULL): DIIS iterations of the order-2 R1 solve. */
int dpd_problem_eecpsd(dpd_problem *prob, dpd_cpsd_opts *opts, const double *T, int64_t nroot,
                       const double *w0, const double *L, double *w, int64_t *iterations);

/* ---- on-disk formats of the reference (host side; no GPU needed) ------------
 * Tensor data may be host or device pointers. */
/* pyscf2cc/mod_interface.f90: the "interface" node file (stream file: metaPos,
 * nNode, data, 36-byte node records; names lower-cased, at most 20 characters).
 * type: 'i' int64, 'd' double, 's' float, 'c' characters; count elements. */
int dpd_node_init(const char *path);
int dpd_node_write(const char *path, const char *name, char type, const void *data, int64_t count);
int dpd_node_read(const char *path, const char *name, char type, void *data, int64_t count);
/* bytes held by the node, -1 when absent */
int64_t dpd_node_size(const char *path, const char *name);
/* libsrint/clmoioso0.f + putallso / getallso_d.f: a symmetry-packed list as the
 * data file <path> (records of 4096 doubles, irrep blocks record-aligned) and
 * its control file <path>.c (Fortran sequential records, INTEGER*8). */
int dpd_moio_write(const dpd_ctx *ctx, const char *path, int64_t irrepx, const int64_t *popp,
                   const int64_t *popq, int64_t ioppq, const int64_t *popr, const int64_t *pops,
                   int64_t ioprs, const double *data);
/* irrepx, pops (4 x 8), iop (2) and the packed length of a list (each may be NULL) */
int dpd_moio_info(const char *path, int64_t *irrepx, int64_t *pops, int64_t *iop, int64_t *length);
int dpd_moio_read(const char *path, double *dst, int64_t length);
/* lib_cc/IO.f90:87-154 (IO_stream): element idx (1-based, 8-byte units); mode 1
 * val and vec, 2 val only, 3 vec only; replace creates the file anew. */
int dpd_stream_write(const char *path, int64_t idx, double val, const double *vec, int64_t n, int mode,
                     int replace);
int dpd_stream_read(const char *path, int64_t idx, double *val, double *vec, int64_t n, int mode);
/* the Davidson eigenvector file (lib_cc/mod_davidson.f90:388-390 writes, :588-589
 * reads the guesses): root r (0-based) of state block istate as [w, R(dim)] at
 * element (istate + r - 1) * (1 + dim) + 1. */
int dpd_eigvec_write(const char *path, int64_t istate, int64_t nroot, int64_t dim, const double *w,
                     const double *R, int replace);
int dpd_eigvec_read(const char *path, int64_t istate, int64_t nroot, int64_t dim, double *w, double *R);

#ifdef __cplusplus
}
#endif
#endif
