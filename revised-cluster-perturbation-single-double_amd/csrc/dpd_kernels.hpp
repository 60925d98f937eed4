// dpd_kernels.hpp -- host-side launch interface of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dpd_common.hpp"

namespace dpd {

// ---- generic DPD gather-map (all sort kernels) -----------------------------
// out[o] = beta*out[o] + alpha * sum_t coef_t * in(perm_t(decode_out(o)))
// for every element o of the output layout; `perm[i]` names the input dim
// that output dim i comes from.  Packed input pairs are mirrored with their
// sign; structural zeros read as 0.  `nbatch` tensors with element strides.
struct MapTerm {
  int8_t perm[4];
  double coef;
};
constexpr int kMaxTerms = 4;
struct MapArgs {
  DevLayout out, in;
  MapTerm term[kMaxTerms];
  int nterm;
  double alpha, beta;
  int64_t in_bstride, out_bstride;
  int nbatch;
};
hipError_t launch_gather_map(const MapArgs& a, const double* in, double* out, hipStream_t s);
// the same from an FP32 source (scratch results of the FP32-accumulating GEMMs)
hipError_t launch_gather_map_f32in(const MapArgs& a, const float* in, double* out, hipStream_t s);
// the same sort writing FP32 (an expanded operand that only feeds an FP32 or
// BF16x3 panel pack)
hipError_t launch_gather_map_f32out(const MapArgs& a, const double* in, float* out, hipStream_t s);
// FP32 storage on both sides (tensors of a context whose data precision is 4)
hipError_t launch_gather_map_f32(const MapArgs& a, const float* in, float* out, hipStream_t s);
// the same gathers through a precomputed code per output element and term
// (nterm int32 each): launch_map_build writes it, launch_gather_code* apply it
hipError_t launch_map_build(const MapArgs& a, int32_t* code, hipStream_t s);
hipError_t launch_gather_code(const MapArgs& a, const int32_t* code, const double* in, double* out, hipStream_t s);
hipError_t launch_gather_code_f32in(const MapArgs& a, const int32_t* code, const float* in, double* out,
                                    hipStream_t s);
hipError_t launch_gather_code_f32out(const MapArgs& a, const int32_t* code, const double* in, float* out,
                                     hipStream_t s);

// ---- fused accumulating epilogues ------------------------------------------
// out[o] += alpha_s * sum_t coef_{s,t} * in_s(code_s[o*nt_s + t]) for the sources
// s in order (the same products and order as one k_gather_code launch per
// source with beta = 1), every tensor of the batch, out read and written once
struct EpiSrc {
  const int32_t* code;
  const void* in;  // float or double, as the launch says
  int64_t in_bstride;
  int nt;
  double coef[kMaxTerms];
  double alpha;
};
constexpr int kEpiMax = 4;
struct EpiFuse {
  EpiSrc s[kEpiMax];
  int ns;
  int64_t n, out_bstride;
  int nbatch;
};
hipError_t launch_epi_fused(const EpiFuse& g, bool in_f32, double* out, hipStream_t s);

// ---- tiled DPD permutation (sorts between full-pair layouts) ---------------
// Within one irrep quadruple both tensors are dense 4-index arrays with their
// own strides.  The output digits split into A (output-contiguous: output
// dim 0, with dim 1 when A is short), B (input-contiguous: the output dims of
// input dims 0 and 1) and C (the rest, up to three digits).  A work item is a
// TC1 x TC0 x TA x TB tile (C digits 1 and 0, A, B), read with B fastest into
// LDS and written with A fastest; without B (the first input and output dims
// coincide) every tile is copied straight through (k_sort_copy).
struct SortTask {
  int64_t first;              // first work item of the task
  int64_t in_base, out_base;  // element (0,0,0,0) of the irrep block in each tensor
  int64_t inA[2];             // input strides of the A digits (output: contiguous)
  int64_t outB[2];            // output strides of the B digits (input: contiguous)
  int64_t inC[3], outC[3];    // strides of the C digits
  int32_t nA0, nA, nB0, nB;   // first digit and flattened extent of A and B (nB = 0: no B)
  int32_t nC[3];
  int32_t TA, TB, TC0, TC1;   // tile extents along A, B and C digits 0 and 1
  int32_t tA, tB, tC0, tC1;   // tiles along A, B and C digits 0 and 1
  int32_t pad_;
};
constexpr int kSortLds = 2304;     // LDS elements of one tile: TC1 * TC0 * TA * (TB + 1) at most
constexpr int kSortTileB = 2048;   // TA * TB at most
constexpr int kSortCopy = 4096;    // elements of a copied tile at most
constexpr int kSortRows = 1024;    // C digit combinations of a copied tile at most
// item_task[i]: the task of work item i; direct: every task without B
// (k_sort_copy), else k_sort_tiles
hipError_t launch_sort_tiles(const SortTask* tasks, const int32_t* item_task, int64_t nitems, bool direct,
                             const double* in, double* out, hipStream_t s);
hipError_t launch_sort_tiles_f32(const SortTask* tasks, const int32_t* item_task, int64_t nitems, bool direct,
                                 const float* in, float* out, hipStream_t s);

// ---- singles-from-singles sigma terms (W r, Fac r, r Fki) -----------------
struct SinglesArgs {
  int64_t l1, stride;  // R1 elements per vector; distance between vectors
  int nvec, nblk;      // vectors; irrep sub-blocks of R1 (by the irrep of i)
  int64_t start[9];    // first element of each sub-block (start[nblk] = l1)
  int32_t na[8], ni[8];
  int64_t fac_off[8], fki_off[8];  // Fac block of a's irrep, Fki block of i's irrep
  const double* W;     // dense l1 x l1 block (column (a,i), row (f,n)), or null
  const double* Fac;   // null: no Fock terms
  const double* Fki;
};
hipError_t launch_singles(const SinglesArgs& g, const double* R, double* S, hipStream_t s);

// ---- grouped DPD GEMM -------------------------------------------------------
// Every task writes C(m,n) (m,n factorised into up to 3 mixed-radix digits with
// their own strides) = alpha * sum_seg A_seg(m,k) B_seg(k,n) + beta * C(m,n).
struct GemmSeg {
  int64_t a_off, b_off;
  int64_t sa_m[3], sa_k[3];
  int64_t sb_k[3], sb_n[3];
  int32_t kd[2];  // mixed-radix sizes of the first two K digits
  int32_t K;
  int32_t ktile_begin;  // first K tile of this segment inside the task
  int32_t ktiles;
  int32_t pad_;
};
struct GemmTask {
  int64_t c_off;
  int64_t sc_m[3], sc_n[3];
  int32_t md[2], nd[2];
  int32_t M, N;
  int32_t seg_begin, nseg;
  int32_t ktiles_total;
  int32_t ksplit;
  int32_t tiles_m, tiles_n;
  int32_t wg_begin;
  int32_t group;  // grouped launches: index into the launch's GroupTab
  int64_t part_off;  // split-K partial buffer offset (elements) if ksplit > 1
  // packed-panel kernel: K steps, first panel ids and panel offsets (elements)
  int32_t steps;
  int32_t pan_a, pan_b;
  int32_t pad2_;
  int64_t pa_off, pb_off;
};

enum Prec { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16X3 = 2 };
// tile configurations: GEMM_V1 is the first-generation 64x64 kernel, the others
// the pipelined kernel of dpd_gemm.hip with BM x BN tiles
// (GEMM_64x192, the wide-N tile that covers a whole narrow N in one tile, exists
// for the packed-panel kernel only)
enum GemmCfg {
  GEMM_V1 = 0,
  GEMM_64x64 = 1,
  GEMM_128x64 = 2,
  GEMM_64x128 = 3,
  GEMM_128x128 = 4,
  GEMM_32x32 = 5,
  GEMM_64x192 = 6,
  // BF16x3 packed-panel kernel with A fragments loaded straight into registers
  // (other precisions take GEMM_128x64 in its place)
  GEMM_128x64D = 7
};
constexpr int kGemmCfgMax = GEMM_128x64D;
// width of the direct-A tile (build-time choice: 64, or 192 for a tile that
// covers a whole narrow N; 64 is faster on the D2h/120 ring and equal elsewhere)
#ifndef DPD_PKD_BN
#define DPD_PKD_BN 64
#endif
constexpr int kPkdBN = DPD_PKD_BN;
inline int gemm_bm(int cfg) {
  return (cfg == GEMM_128x64 || cfg == GEMM_128x128 || cfg == GEMM_128x64D) ? 128 : cfg == GEMM_32x32 ? 32 : 64;
}
inline int gemm_bn(int cfg) {
  return (cfg == GEMM_64x128 || cfg == GEMM_128x128) ? 128
         : cfg == GEMM_32x32                         ? 32
         : cfg == GEMM_64x192                        ? 192
         : cfg == GEMM_128x64D                      ? kPkdBN
                                                     : 64;
}

struct GemmLaunch {
  const void* A;  // storage of operand 1 (double, or float when a_f32)
  const void* B;
  double* C;
  double* partial;  // split-K scratch
  const GemmTask* tasks;  // device
  const GemmSeg* segs;    // device
  int ntask;
  int nwg;
  double alpha, beta;
  int prec;
  int a_f32, b_f32;  // operand storage is FP32 (mixed precision)
  // split-K reduce
  const int32_t* split_tasks;  // device list of task ids with ksplit > 1
  int nsplit;
  int max_split_mn;
  int cfg;  // GemmCfg
  int c_f32;  // C is FP32: C = fl32(beta * C + alpha * AB), formed in FP64
};
hipError_t launch_gemm(const GemmLaunch& g, hipStream_t s);
hipError_t launch_gemm2(const GemmLaunch& g, hipStream_t s);
hipError_t launch_splitk_reduce(const GemmLaunch& g, hipStream_t s);

// ---- packed-panel GEMM (dpd_gemm_pk.hip) ----------------------------------
// copies one operand side of every task into GEMM panels of precision `mode`
struct PackArgs {
  const double* src;
  void* dst;
  const GemmTask* tasks;  // device
  const GemmSeg* segs;    // device
  int ntask;
  int npanel;
  int side;  // 0: A (rows = M), 1: B (rows = N)
  int mode;  // Prec
  int tile;  // BM or BN
  // optional fused sort: the GEMM operand is map.out (an expanded/permuted copy
  // of map.in, xlen elements per batch member); the pack reads map.in directly
  int use_map;
  int rowmask;  // bit d: dim d of map.out is a row (M or N) index, else a K index
  int rfast;    // read the source with rows (1) or k's (0) varying fastest
  int64_t xlen, src_bstride;
  MapArgs map;
  int src_f32;  // src holds floats (no fused sort)
  // code mode: write, per panel element (panel pid, element e at pid*pan + e),
  // +-(source address + 1) or 0 instead of packing values; with gcode the
  // operand is the gather of src through gcode (one term): X element o of batch
  // member b is gcode[o] applied to member b of src (stride gbstride)
  int32_t* code_out;
  const int32_t* gcode;
  int64_t gxlen, gbstride;
};
hipError_t launch_pack(const PackArgs& p, hipStream_t s);
// panels from a code written by launch_pack's code mode: npanel panels of
// pan_elems elements from src (FP64, or FP32 when src_f32) into dst (mode)
hipError_t launch_pack_code(int mode, int src_f32, int64_t npanel, int64_t pan_elems, const int32_t* code,
                            const void* src, void* dst, hipStream_t s);
// K step of a packed panel and its element count factor (BF16x3 keeps two planes)
inline int pk_kstep(int mode) { return mode == PREC_F64 ? 16 : 32; }
inline int pk_planes(int mode) { return mode == PREC_BF16X3 ? 2 : 1; }
inline int pk_esize(int mode) { return mode == PREC_F64 ? 8 : mode == PREC_F32 ? 4 : 2; }
hipError_t launch_gemm_pk(const GemmLaunch& g, const void* PA, const void* PB, hipStream_t s);
// the GEMM kernel alone (split-K tasks leave partials for launch_splitk_reduce)
hipError_t launch_gemm_pk_only(const GemmLaunch& g, const void* PA, const void* PB, hipStream_t s);
// grouped launch: the tasks of several independent contractions (one per
// problem of a batch) in one kernel; task.group selects the panels, result and
// beta of its contraction; split-K partials of all tasks share g.partial.
constexpr int kMaxGroup = 32;
struct GroupEnt {
  const void* pa;
  const void* pb;
  double* c;
  double beta;
  int c_f32;  // c is FP32 (update formed in FP64, rounded once)
  int pad_;
};
struct GroupTab {
  GroupEnt e[kMaxGroup];
};
hipError_t launch_gemm_pk_grouped(const GemmLaunch& g, const GroupTab& gt, hipStream_t s);
hipError_t launch_splitk_reduce_grouped(const GemmLaunch& g, const GroupTab& gt, hipStream_t s);

// ---- vector kernels (FP64) ------------------------------------------------
constexpr int kRedBlocks = 512;
// partial[b*nv + v] over blocks b: dot(x, y_v) for v < nv, y_v = y + v*ystride
hipError_t launch_multidot(int64_t n, const double* x, const double* y, int64_t ystride, int nv,
                           double* partial, double* out, hipStream_t s);
hipError_t launch_amax(int64_t n, const double* x, double* partial_val, int64_t* partial_idx,
                       double* out_val, int64_t* out_idx, hipStream_t s);
hipError_t launch_axpy(int64_t n, double a, const double* x, double* y, hipStream_t s);
// y = a_dev[0]*x + y with the coefficient read from device memory (scaled by mul)
hipError_t launch_axpy_dev(int64_t n, const double* a_dev, double mul, const double* x, double* y,
                           hipStream_t s);
hipError_t launch_scal(int64_t n, double a, double* x, hipStream_t s);
hipError_t launch_scal_dev(int64_t n, const double* a_dev, int invert_sqrt, double* x,
                           hipStream_t s);
hipError_t launch_fill(int64_t n, double v, double* x, hipStream_t s);
hipError_t launch_copy(int64_t n, const double* x, double* y, hipStream_t s);
// out[j*n + i] = sum_k coef[k + j*ldc] * X[k*xstride + i]  for j < nout
hipError_t launch_lincomb(int64_t n, const double* X, int64_t xstride, int nk, const double* coef,
                          int ldc, int nout, double* out, hipStream_t s);
hipError_t launch_precondition(int64_t n, double w, const double* diag, const double* e, double* p,
                               hipStream_t s);
hipError_t launch_hadamard_acc(int64_t n, const double* a, const double* b, double* c,
                               int nbatch, int64_t bstride_b, int64_t bstride_c, hipStream_t s);
hipError_t launch_convert_f32(int64_t n, const double* x, float* y, hipStream_t s);

}  // namespace dpd

namespace dpd {
// orbital-energy differences over a DPD layout (Eorb_diff.f90, denomT.f90):
// mode 1: e(a,i) = e_a - e_i        (layout (v,o,-,-))
// mode 2: e(ab,ij) = e_a+e_b-e_i-e_j (layout (v,v,o,o))
// mode 3: D(ab,ij) = e_i+e_j-e_a-e_b
// mode 4: D(a,i) = e_i - e_a         (layout (v,o,-,-))
struct EorbOff {
  int64_t offo[8], offv[8];
};
// adjoint of orbital-energy sums over a DPD layout (see k_eorb_adjoint)
struct EorbAdj {
  int space[4];    // 0 none, 1 virtual, 2 occupied
  double sign[4];
  double coef;
};
hipError_t launch_eorb_adjoint(const DevLayout& L, const double* x, const double* y, const EorbOff& off,
                               const EorbAdj& adj, double* eorb_bar, hipStream_t s);
hipError_t launch_orbital_energy(const DevLayout& L, const double* eorb, const EorbOff& off, int mode,
                                 double* out, hipStream_t s);
hipError_t launch_div(int64_t n, const double* num, const double* den, double* out, hipStream_t s);
// out = num / (den - shift)
hipError_t launch_div_shift(int64_t n, const double* num, const double* den, double shift, double* out,
                            hipStream_t s);
}  // namespace dpd

namespace dpd {
// ---- batched Davidson kernels --------------------------------------------
// out[xi + yj*nx] = X_xi . Y_yj  (partial: kRedBlocks*nx*ny)
hipError_t launch_block_dots(int64_t n, const double* X, int64_t xs, int nx, const double* Y, int64_t ys, int ny,
                             double* partial, double* out, hipStream_t s);
// x_v, y_v scaled by 1/sqrt(nrm2[v]) unless that norm is 1 within 1e-10
hipError_t launch_renorm_pair(int64_t n, const double* nrm2, double* x, double* y, int nv, hipStream_t s);
// small per-iteration coefficient sets passed to the kernels by value (kernel
// arguments: no host-to-device copy ahead of the launch)
constexpr int kRootVals = 64;
struct RootVals {
  double w[kRootVals];
};
constexpr int kLinCoef = 256;
struct LinCoef {
  double v[kLinCoef];  // v[k + j*nk]
};
struct RowList {
  int32_t src[kRootVals];
  int n;
};
// preconditioned directions P of the residuals e_v = Aq_v - w_v q_v (e not
// stored) of every root in one pass (nv <= kRootVals), and in one more launch stat[v] = max|e_v|
// (first index oi[v]), stat[nv + v] = |q_v|^2 and stat[2 nv + v] = |p_v|^2, each
// summed in a fixed order (qpart, ppart: kRedBlocks * nv partials).  q and Aq
// are used as they are: e, p and max|e| scale with 1/|q| when the caller
// renormalises q.
hipError_t launch_ritz_resid_part(int64_t n, const RootVals& w, const double* d, const double* q, const double* aq,
                                  double* p, int nv, double* pv, int64_t* pi, double* qpart, double* ppart,
                                  double* stat, int64_t* oi, hipStream_t s);
// launch_lincomb of two sets (X -> out and X2 -> out2) with the same coefficients
hipError_t launch_lincomb2(int64_t n, const double* X, const double* X2, int64_t xs, int nk, const double* coef,
                           int ldc, int nout, double* out, double* out2, hipStream_t s);
// the same with the coefficients by value (nk * nout <= kLinCoef)
hipError_t launch_lincomb2_v(int64_t n, const double* X, const double* X2, int64_t xs, int nk, const LinCoef& cf,
                             int nout, double* out, double* out2, hipStream_t s);
// rows P_k = P_src[k] (k ascending; src[k] >= k, increasing): in place
hipError_t launch_compact_rows(int64_t n, double* P, const RowList& rows, hipStream_t s);
// partial[0..kRedBlocks) of dot(x, y); out[r] = sum of row r of kRedBlocks partials
hipError_t launch_pair_dots_partial(int64_t n, const double* x, const double* y, double* partial, hipStream_t s);
hipError_t launch_reduce_partials(const double* partial, int nrows, double* out, hipStream_t s);
// e_v = aq_v - w_v q_v and max|e_v| (first index) for v < nv; w on device
hipError_t launch_resid_amax(int64_t n, const double* aq, const double* q, const double* w, double* e, int nv,
                             double* pv, int64_t* pi, double* ov, int64_t* oi, hipStream_t s);
// one fused modified Gram-Schmidt step on p:
//   c = sum(part_in[0..nb)), mode 0: p *= 1/sqrt(c); mode 1: p -= c*b_prev;
//   then part_out[block] = dot(p, b_next) (b_next == nullptr: dot(p, p))
hipError_t launch_mgs_step(int64_t n, double* p, const double* b_prev, const double* b_next, const double* part_in,
                           int mode, double* part_out, hipStream_t s);
// directions handled together by the block kernels below
constexpr int kMgsMulti = 8;
// out[j + k*nv] = P_j . B_k (partial: kRedBlocks * nv * nb), nv <= kMgsMulti
hipError_t launch_dots_multi(int64_t n, const double* P, int nv, const double* B, int nb, double* partial,
                             double* out, hipStream_t s,
                             const int* need = nullptr);
// two such sets in one launch and one reduction: out[j + k*nv] = P1_j . B1_k
// (k < nb1), then out[nv*nb1 + j + k*nv] = P2_j . B2_k (k < nb2); partial:
// kRedBlocks * nv * (nb1 + nb2)
hipError_t launch_dots_pair(int64_t n, int nv, const double* P1, const double* B1, int nb1, const double* P2,
                            const double* B2, int nb2, double* partial, double* out, hipStream_t s);
// the root of each of up to kMgsMulti directions (an index into per-root arrays)
struct MgsSrc {
  int32_t src[kMgsMulti];
};
// one classical Gram-Schmidt projection: P_j <- f_j (P_j - sum_k C[j + k*nv] B_k)
// with f_j = 1/sqrt(pn2[src[j]]) (pn2 given; C the projections of the
// unnormalised P_j) or 1, and the Gram partials of the result (rows j + k*nv,
// j <= k, kRedBlocks each, as launch_tri_apply writes them); need given and
// *need == 0: nothing done.  nb may be 0 (normalisation only).
hipError_t launch_cgs_update(int64_t n, const double* B, int nb, const double* C, int nv, double* P,
                             const int* need, const double* pn2, const MgsSrc& src, double* gram, hipStream_t s);
// need = 1 when any of nv unit vectors lost more than half its squared norm to
// its projections C on an orthonormal basis (a second projection is needed);
// pn2 given: the vectors P_j are unnormalised, |P_j|^2 = pn2[src[j]]
hipError_t launch_cgs_need(const double* C, int nv, int nb, int* need, hipStream_t s,
                           const double* pn2 = nullptr, const MgsSrc& src = MgsSrc{});
// Cholesky QR step taken inside launch_tri_apply: from the Gram partials gpart
// (rows j + k*nv, j <= k) of nv vectors, G = L L^T and T = L^-T, stopping at
// the first remaining norm (times piv_in[j] when given) below 1e-6 (columns
// from there on zero); piv_out / nrm2_out: the norms and their squares (0 from
// the stopping column on); T_out: T; T_pre given: the vectors are multiplied
// by T_pre T instead of T
struct TriChol {
  const double* gpart = nullptr;
  const double* piv_in = nullptr;
  const double* T_pre = nullptr;
  double* piv_out = nullptr;
  double* nrm2_out = nullptr;
  double* T_out = nullptr;
};
// Y_j = sum_{i<=j} T[i + j*nv] X_i (T null: identity; Y null: not stored; Y may
// be X) and, with partial given, the Gram partials of Y (rows j + k*nv, j <= k);
// with ch.gpart given T is the factor of that Cholesky QR step instead
hipError_t launch_tri_apply(int64_t n, const double* X, int nv, const double* T, double* Y, double* partial,
                            hipStream_t s, const TriChol& ch = TriChol{});
// dst = p * (1/sqrt(sum(part))) ; returns nothing (norm read by the host separately)
hipError_t launch_scale_by_norm(int64_t n, const double* p, const double* part, double* dst, hipStream_t s);
// p = e / clamp(w - d) for the roots listed in idx (w on device)
hipError_t launch_precond_batched(int64_t n, const double* w, const int32_t* idx, int nv, const double* d,
                                  const double* e, double* p, hipStream_t s);
// the same from the Ritz pair: e = aq_v - w_v q_v formed on the fly
hipError_t launch_precond_ritz(int64_t n, const double* w, const int32_t* idx, int nv, const double* d,
                               const double* aq, const double* q, double* p, hipStream_t s);
}  // namespace dpd
