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
  int32_t pad_;
  int64_t part_off;  // split-K partial buffer offset (elements) if ksplit > 1
};

enum Prec { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16X3 = 2 };

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
};
hipError_t launch_gemm(const GemmLaunch& g, hipStream_t s);

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
struct EorbOff {
  int64_t offo[8], offv[8];
};
hipError_t launch_orbital_energy(const DevLayout& L, const double* eorb, const EorbOff& off, int mode,
                                 double* out, hipStream_t s);
hipError_t launch_div(int64_t n, const double* num, const double* den, double* out, hipStream_t s);
}  // namespace dpd
