// dpd_kernels.hip -- hand-written gfx950 (CDNA4) kernels of the DPD engine.
//
//  * k_gather_map: one generic, element-parallel DPD sort kernel.  Each thread
//    owns one OUTPUT element (coalesced writes), decodes its irreps/indices from
//    the output layout, permutes them and gathers up to four signed terms from
//    the input layout.  It implements iexpso/icompso/isymtrso/itranspso/sstgenso
//    and the fused antisymmetrise+compress+accumulate epilogue of mod_contract.
//  * k_gemm: grouped GEMM over DPD irrep tasks on MFMA (FP64
//    v_mfma_f64_16x16x4f64 or FP32 v_mfma_f32_16x16x4f32), 64x64x16 LDS tiles,
//    4 waves of 32x32, K segments per task (several irrep tuples accumulate into
//    one output block) and deterministic split-K.
//  * reductions for the Davidson vector algebra (FP64): multi-dot, max-abs,
//    axpy/scal/lincomb, the clamped preconditioner.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "dpd_kernels.hpp"

namespace dpd {

// ===========================================================================
// gather-map
// ===========================================================================
__device__ inline void stage_map(const MapArgs& a, MapArgs& sa);

template <typename TIN, typename TOUT>
__global__ __launch_bounds__(256) void k_gather_map(MapArgs a_arg, const TIN* __restrict__ in,
                                                    TOUT* __restrict__ out) {
  __shared__ MapArgs a;
  stage_map(a_arg, a);
  const int64_t n = a.out.len;
  const int b = blockIdx.y;
  const TIN* inb = in + (int64_t)b * a.in_bstride;
  TOUT* outb = out + (int64_t)b * a.out_bstride;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n;
       o += (int64_t)gridDim.x * blockDim.x) {
    int h[4];
    int64_t x[4];
    decode_elem(a.out, o, h, x);
    double acc = 0.0;
    for (int t = 0; t < a.nterm; t++) {
      int hi[4];
      int64_t xi[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        hi[a.term[t].perm[i]] = h[i];
        xi[a.term[t].perm[i]] = x[i];
      }
      double sg;
      int64_t ad = fetch_addr(a.in, hi[0], hi[1], hi[2], hi[3], xi[0], xi[1], xi[2], xi[3], &sg);
      if (ad >= 0) acc += a.term[t].coef * sg * (double)inb[ad];
    }
    double v = a.alpha * acc;
    if (a.beta != 0.0) v = a.beta * (double)outb[o] + v;
    outb[o] = (TOUT)v;
  }
}

// Column-tiled variant: every block owns a chunk of kGatherChunk consecutive
// elements of one column (one right pair of one irrep block) of the output, so
// the right pair is decoded once per block and every element only decodes its
// left pair (no search over blocks, no 64-bit division).
constexpr int kGatherChunk = 1024;

// small outputs use chunks of one element per thread (more blocks in flight)
constexpr int kGatherChunkSmall = 256;
__host__ __device__ inline int64_t gather_chunks(const DevLayout& L, int xr, int chunk) {
  const int xl = L.dp[xr][L.sym];
  return L.nR[xr] * ((L.nL[xl] + chunk - 1) / chunk);
}

// the layouts are indexed by per-lane irreps: staged in LDS so that those
// lookups are LDS reads rather than dependent loads from the argument segment
__device__ inline void stage_map(const MapArgs& a, MapArgs& sa) {
  static_assert(sizeof(MapArgs) % 4 == 0, "MapArgs size");
  for (int i = threadIdx.x; i < (int)(sizeof(MapArgs) / 4); i += blockDim.x)
    ((int32_t*)&sa)[i] = ((const int32_t*)&a)[i];
  __syncthreads();
}

template <typename TIN, typename TOUT>
__global__ __launch_bounds__(256) void k_gather_cols(MapArgs a_arg, const TIN* __restrict__ in,
                                                     TOUT* __restrict__ out, int chunk) {
  __shared__ MapArgs a;
  stage_map(a_arg, a);
  const DevLayout& L = a.out;
  int64_t b = blockIdx.x;
  int xr = 0;
  for (; xr < L.nirrep; xr++) {
    const int64_t c = gather_chunks(L, xr, chunk);
    if (b < c) break;
    b -= c;
  }
  if (xr >= L.nirrep) return;
  const int xl = L.dp[xr][L.sym];
  const int64_t nl = L.nL[xl];
  const int64_t per = (nl + chunk - 1) / chunk;
  const int64_t r = b / per, l0 = (b % per) * chunk;
  int h[4];
  int64_t x[4];
  decode_pair(L.pop[2], L.pop[3], L.iop[1], L.dp, L.offR[xr], L.nirrep, xr, r, &h[2], &x[2], &h[3], &x[3]);
  const int64_t col = L.blk[xr] + r * nl;
  const int64_t l1 = l0 + chunk < nl ? l0 + chunk : nl;
  // the element's source addresses are decoded once and serve every tensor of
  // the batch (the decode, not the memory traffic, bounds this kernel)
  for (int64_t l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
    decode_pair(L.pop[0], L.pop[1], L.iop[0], L.dp, L.offL[xl], L.nirrep, xl, l, &h[0], &x[0], &h[1], &x[1]);
    int64_t ad[kMaxTerms];
    double cf[kMaxTerms];
#pragma unroll
    for (int t = 0; t < kMaxTerms; t++) {
      ad[t] = -1;
      cf[t] = 0.0;
      if (t < a.nterm) {
        int hi[4];
        int64_t xi[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          hi[a.term[t].perm[i]] = h[i];
          xi[a.term[t].perm[i]] = x[i];
        }
        double sg;
        ad[t] = fetch_addr(a.in, hi[0], hi[1], hi[2], hi[3], xi[0], xi[1], xi[2], xi[3], &sg);
        cf[t] = a.term[t].coef * sg;
      }
    }
    const int64_t o = col + l;
    for (int bt = 0; bt < a.nbatch; bt++) {
      const TIN* inb = in + (int64_t)bt * a.in_bstride;
      TOUT* outb = out + (int64_t)bt * a.out_bstride;
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < kMaxTerms; t++)
        if (ad[t] >= 0) acc += cf[t] * (double)inb[ad[t]];
      double v = a.alpha * acc;
      if (a.beta != 0.0) v = a.beta * (double)outb[o] + v;
      outb[o] = (TOUT)v;
    }
  }
}

template <typename TIN, typename TOUT>
static hipError_t gather_map_t(const MapArgs& a, const TIN* in, TOUT* out, hipStream_t s) {
  if (a.out.len == 0 || a.nbatch == 0) return hipSuccess;
  // one block per (column, chunk): the same count the kernel derives; a
  // launch of fewer than 1024 blocks takes one element per thread instead
  int64_t blocks = 0;
  int chunk = kGatherChunk;
  for (int xr = 0; xr < a.out.nirrep; xr++) blocks += gather_chunks(a.out, xr, chunk);
  if (blocks < 1024) {
    chunk = kGatherChunkSmall;
    blocks = 0;
    for (int xr = 0; xr < a.out.nirrep; xr++) blocks += gather_chunks(a.out, xr, chunk);
  }
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffff) {
    int64_t b2 = (a.out.len + 255) / 256;
    if (b2 > 8192) b2 = 8192;
    hipLaunchKernelGGL((k_gather_map<TIN, TOUT>), dim3((unsigned)b2, (unsigned)a.nbatch), dim3(256), 0, s, a, in,
                       out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_gather_cols<TIN, TOUT>), dim3((unsigned)blocks, 1), dim3(256), 0, s, a, in, out, chunk);
  return hipGetLastError();
}

hipError_t launch_gather_map(const MapArgs& a, const double* in, double* out, hipStream_t s) {
  return gather_map_t<double, double>(a, in, out, s);
}

hipError_t launch_gather_map_f32in(const MapArgs& a, const float* in, double* out, hipStream_t s) {
  return gather_map_t<float, double>(a, in, out, s);
}

hipError_t launch_gather_map_f32out(const MapArgs& a, const double* in, float* out, hipStream_t s) {
  return gather_map_t<double, float>(a, in, out, s);
}

hipError_t launch_gather_map_f32(const MapArgs& a, const float* in, float* out, hipStream_t s) {
  return gather_map_t<float, float>(a, in, out, s);
}

// ---------------------------------------------------------------------------
// precomputed gathers: the source of every output element and term decoded
// once per layout (k_map_build) into a code, +(address + 1) or -(address + 1)
// for the sign of the mirrored element, 0 where the term contributes nothing;
// k_gather_code then needs no index arithmetic and reads each code once for
// all tensors of the batch.  The terms are summed in the same order and with
// the same products as k_gather_cols, so the results are the same bits.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_map_build(MapArgs a_arg, int32_t* __restrict__ code) {
  __shared__ MapArgs a;
  stage_map(a_arg, a);
  const int64_t n = a.out.len;
  const int nt = a.nterm;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (int64_t)gridDim.x * blockDim.x) {
    int h[4];
    int64_t x[4];
    decode_elem(a.out, o, h, x);
    for (int t = 0; t < nt; t++) {
      int hi[4];
      int64_t xi[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        hi[a.term[t].perm[i]] = h[i];
        xi[a.term[t].perm[i]] = x[i];
      }
      double sg;
      const int64_t ad = fetch_addr(a.in, hi[0], hi[1], hi[2], hi[3], xi[0], xi[1], xi[2], xi[3], &sg);
      int32_t cd = 0;
      if (ad >= 0 && sg != 0.0) cd = sg > 0.0 ? (int32_t)(ad + 1) : -(int32_t)(ad + 1);
      code[o * nt + t] = cd;
    }
  }
}

hipError_t launch_map_build(const MapArgs& a, int32_t* code, hipStream_t s) {
  if (a.out.len == 0 || a.nterm < 1) return hipSuccess;
  int64_t b = (a.out.len + 255) / 256;
  if (b > 16384) b = 16384;
  hipLaunchKernelGGL(k_map_build, dim3((unsigned)b), dim3(256), 0, s, a, code);
  return hipGetLastError();
}

struct CodeArgs {
  double coef[kMaxTerms];
  double alpha, beta;
  int64_t n, in_bstride, out_bstride;
  int nbatch;
};

template <typename TIN, typename TOUT, int NT>
__global__ __launch_bounds__(256) void k_gather_code(CodeArgs g, const int32_t* __restrict__ code,
                                                     const TIN* __restrict__ in, TOUT* __restrict__ out) {
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < g.n; o += (int64_t)gridDim.x * blockDim.x) {
    int32_t cd[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) cd[t] = code[o * NT + t];
    int64_t ad[NT];
    double cf[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
      ad[t] = cd[t] > 0 ? (int64_t)cd[t] - 1 : (int64_t)(-cd[t]) - 1;
      cf[t] = cd[t] > 0 ? g.coef[t] : -g.coef[t];
    }
#pragma unroll 2
    for (int b = 0; b < g.nbatch; b++) {
      const TIN* inb = in + (int64_t)b * g.in_bstride;
      TOUT* outb = out + (int64_t)b * g.out_bstride;
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < NT; t++)
        if (cd[t] != 0) acc += cf[t] * (double)inb[ad[t]];
      double v = g.alpha * acc;
      if (g.beta != 0.0) v = g.beta * (double)outb[o] + v;
      outb[o] = (TOUT)v;
    }
  }
}

template <typename TIN, typename TOUT>
static hipError_t gather_code_t(const MapArgs& a, const int32_t* code, const TIN* in, TOUT* out, hipStream_t s) {
  if (a.out.len == 0 || a.nbatch == 0) return hipSuccess;
  CodeArgs g;
  for (int t = 0; t < kMaxTerms; t++) g.coef[t] = t < a.nterm ? a.term[t].coef : 0.0;
  g.alpha = a.alpha;
  g.beta = a.beta;
  g.n = a.out.len;
  g.in_bstride = a.in_bstride;
  g.out_bstride = a.out_bstride;
  g.nbatch = a.nbatch;
  int64_t b = (a.out.len + 255) / 256;
  if (b > 16384) b = 16384;
  const dim3 grid((unsigned)b), blk(256);
  switch (a.nterm) {
    case 1: hipLaunchKernelGGL((k_gather_code<TIN, TOUT, 1>), grid, blk, 0, s, g, code, in, out); break;
    case 2: hipLaunchKernelGGL((k_gather_code<TIN, TOUT, 2>), grid, blk, 0, s, g, code, in, out); break;
    case 3: hipLaunchKernelGGL((k_gather_code<TIN, TOUT, 3>), grid, blk, 0, s, g, code, in, out); break;
    case 4: hipLaunchKernelGGL((k_gather_code<TIN, TOUT, 4>), grid, blk, 0, s, g, code, in, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_gather_code(const MapArgs& a, const int32_t* code, const double* in, double* out, hipStream_t s) {
  return gather_code_t<double, double>(a, code, in, out, s);
}
hipError_t launch_gather_code_f32in(const MapArgs& a, const int32_t* code, const float* in, double* out,
                                    hipStream_t s) {
  return gather_code_t<float, double>(a, code, in, out, s);
}
hipError_t launch_gather_code_f32out(const MapArgs& a, const int32_t* code, const double* in, float* out,
                                     hipStream_t s) {
  return gather_code_t<double, float>(a, code, in, out, s);
}

// ===========================================================================
// fused accumulating epilogues: the antisymmetrising epilogues of several
// contractions into one output, which is then read and written once
// ===========================================================================
template <typename TIN>
__global__ __launch_bounds__(256) void k_epi_fused(EpiFuse g, double* __restrict__ out) {
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < g.n; o += (int64_t)gridDim.x * blockDim.x) {
    int32_t cd[kEpiMax][kMaxTerms];
#pragma unroll
    for (int s = 0; s < kEpiMax; s++)
#pragma unroll
      for (int t = 0; t < kMaxTerms; t++)
        cd[s][t] = (s < g.ns && t < g.s[s].nt) ? g.s[s].code[o * g.s[s].nt + t] : 0;
#pragma unroll 2
    for (int b = 0; b < g.nbatch; b++) {
      double* outb = out + (int64_t)b * g.out_bstride;
      double v = outb[o];
#pragma unroll
      for (int s = 0; s < kEpiMax; s++) {
        if (s >= g.ns) break;
        const TIN* inb = (const TIN*)g.s[s].in + (int64_t)b * g.s[s].in_bstride;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < kMaxTerms; t++) {
          const int32_t c = cd[s][t];
          if (c != 0) {
            const int64_t ad = c > 0 ? (int64_t)c - 1 : (int64_t)(-c) - 1;
            const double cf = c > 0 ? g.s[s].coef[t] : -g.s[s].coef[t];
            acc += cf * (double)inb[ad];
          }
        }
        v = 1.0 * v + g.s[s].alpha * acc;
      }
      outb[o] = v;
    }
  }
}

hipError_t launch_epi_fused(const EpiFuse& g, bool in_f32, double* out, hipStream_t s) {
  if (g.n <= 0 || g.nbatch <= 0 || g.ns <= 0) return hipSuccess;
  if (g.ns > kEpiMax) return hipErrorInvalidValue;
  int64_t b = (g.n + 255) / 256;
  if (b > 16384) b = 16384;
  if (in_f32)
    hipLaunchKernelGGL((k_epi_fused<float>), dim3((unsigned)b), dim3(256), 0, s, g, out);
  else
    hipLaunchKernelGGL((k_epi_fused<double>), dim3((unsigned)b), dim3(256), 0, s, g, out);
  return hipGetLastError();
}

// ===========================================================================
// tiled DPD permutation: one workgroup per work item of a SortTask (the task
// found by bisection over the tasks' first items).  A transposing tile is read
// with the input-contiguous digits fastest (coalesced reads), staged in LDS
// with rows padded to an odd length, and written with the output-contiguous
// digits fastest (coalesced writes); tasks without B are copied straight.
// Element indices are split into digits by multiplication with a reciprocal
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", 1994: q = (umulhi(n, m) + n) >> l for n < 2^31), and every
// thread keeps four loads in flight.
// ===========================================================================
struct FastDiv {
  uint32_t d, m, l;
  __device__ void init(uint32_t dd) {
    d = dd;
    l = 0;
    while ((1u << l) < d) l++;
    m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
  }
  __device__ uint32_t div(uint32_t n) const { return (__umulhi(n, m) + n) >> l; }
};

struct SortItem {
  int64_t ib0, ob0;  // bases of the item (C digits' tile origin and digit 2)
  int a_lo, a_n, b_lo, b_n, c0_n, c1_n;
};

// the task of work item `item` (from the plan's item -> task table) staged in
// LDS, and the item's tile
__device__ inline void sort_item(const SortTask* __restrict__ tasks, const int32_t* __restrict__ item_task,
                                 SortTask& st, SortItem& it) {
  const int64_t item = blockIdx.x;
  const int k = item_task[item];
  static_assert(sizeof(SortTask) % 4 == 0, "SortTask size");
  for (int i = threadIdx.x; i < (int)(sizeof(SortTask) / 4); i += blockDim.x)
    ((int32_t*)&st)[i] = ((const int32_t*)&tasks[k])[i];
  __syncthreads();
  int local = (int)(item - st.first);
  const int tb = local % st.tB;
  local /= st.tB;
  const int ta = local % st.tA;
  local /= st.tA;
  const int tc0 = local % st.tC0;
  local /= st.tC0;
  const int tc1 = local % st.tC1;
  const int c2 = local / st.tC1;
  it.a_lo = ta * st.TA;
  it.a_n = min(st.TA, st.nA - it.a_lo);
  it.b_lo = tb * st.TB;
  it.b_n = st.nB ? min(st.TB, st.nB - it.b_lo) : 1;
  const int c0_lo = tc0 * st.TC0, c1_lo = tc1 * st.TC1;
  it.c0_n = min(st.TC0, st.nC[0] - c0_lo);
  it.c1_n = min(st.TC1, st.nC[1] - c1_lo);
  it.ib0 = st.in_base + (int64_t)c0_lo * st.inC[0] + (int64_t)c1_lo * st.inC[1] + (int64_t)c2 * st.inC[2];
  it.ob0 = st.out_base + (int64_t)c0_lo * st.outC[0] + (int64_t)c1_lo * st.outC[1] + (int64_t)c2 * st.outC[2];
}

constexpr int kSortU = 4;  // loads in flight per thread

// Offsets of a tile's rows: every digit combination outside the contiguous
// group is turned into its input and output offset once per work item (LDS
// tables of int32, relative to the item's bases; tensors stay below 2^31
// elements), so that an element costs one reciprocal split and table reads.
template <typename T>
__global__ __launch_bounds__(256) void k_sort_copy(const SortTask* __restrict__ tasks,
                                                   const int32_t* __restrict__ item_task, const T* __restrict__ in,
                                                   T* __restrict__ out) {
  __shared__ SortTask st;
  __shared__ int32_t rin[kSortRows], rout[kSortRows];
  SortItem it;
  sort_item(tasks, item_task, st, it);
  // rows: C digits 0 and 1 of the tile
  const int nr = it.c0_n * it.c1_n;
  for (int r = threadIdx.x; r < nr; r += blockDim.x) {
    const int i0 = r % it.c0_n, i1 = r / it.c0_n;
    rin[r] = (int32_t)((int64_t)i0 * st.inC[0] + (int64_t)i1 * st.inC[1]);
    rout[r] = (int32_t)((int64_t)i0 * st.outC[0] + (int64_t)i1 * st.outC[1]);
  }
  __syncthreads();
  FastDiv fa;
  fa.init(it.a_n);
  const T* inb = in + it.ib0 + it.a_lo;
  T* outb = out + it.ob0 + it.a_lo;
  const int n = it.a_n * nr;
  for (int e0 = threadIdx.x; e0 < n; e0 += kSortU * blockDim.x) {
    T v[kSortU];
    int oa[kSortU];
#pragma unroll
    for (int u = 0; u < kSortU; u++) {
      const int e = e0 + u * blockDim.x;
      oa[u] = -1;
      if (e < n) {
        const int r = (int)fa.div(e), ia = e - r * it.a_n;
        v[u] = inb[rin[r] + ia];
        oa[u] = rout[r] + ia;
      }
    }
#pragma unroll
    for (int u = 0; u < kSortU; u++)
      if (oa[u] >= 0) outb[oa[u]] = v[u];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sort_tiles(const SortTask* __restrict__ tasks,
                                                    const int32_t* __restrict__ item_task, const T* __restrict__ in,
                                                    T* __restrict__ out) {
  __shared__ T lds[kSortLds];
  __shared__ int32_t tab[kSortLds];  // input rows, then output columns
  __shared__ SortTask st;
  SortItem it;
  sort_item(tasks, item_task, st, it);
  // an odd row pitch (in elements): the A-fastest reads of the write phase then
  // fall on distinct LDS banks
  const int pitch = it.b_n | 1;
  const int nc = it.c0_n * it.c1_n;
  // input rows (A digit, C digits: B contiguous within) and output columns
  // (B digit, C digits: A contiguous within)
  const int nrow = it.a_n * nc, ncol = it.b_n * nc;
  int32_t* rin = tab;
  int32_t* cout = tab + nrow;
  for (int r = threadIdx.x; r < nrow; r += blockDim.x) {
    const int ia = r % it.a_n, q = r / it.a_n;
    const int i0 = q % it.c0_n, i1 = q / it.c0_n;
    const int A = it.a_lo + ia;
    const int a0 = A % st.nA0, a1 = A / st.nA0;
    rin[r] = (int32_t)((int64_t)i0 * st.inC[0] + (int64_t)i1 * st.inC[1] + (int64_t)a0 * st.inA[0] +
                       (int64_t)a1 * st.inA[1]);
  }
  for (int q = threadIdx.x; q < ncol; q += blockDim.x) {
    const int ib = q % it.b_n, r2 = q / it.b_n;
    const int i0 = r2 % it.c0_n, i1 = r2 / it.c0_n;
    const int B = it.b_lo + ib;
    const int b0 = B % st.nB0, b1 = B / st.nB0;
    cout[q] = (int32_t)((int64_t)i0 * st.outC[0] + (int64_t)i1 * st.outC[1] + (int64_t)b0 * st.outB[0] +
                        (int64_t)b1 * st.outB[1]);
  }
  __syncthreads();
  FastDiv fa, fb;
  fa.init(it.a_n);
  fb.init(it.b_n);
  const T* inb = in + it.ib0 + it.b_lo;
  T* outb = out + it.ob0 + it.a_lo;
  const int n = it.a_n * it.b_n * nc;
  // read: B fastest; row r = (C1 * c0_n + C0) * a_n + A
  for (int e0 = threadIdx.x; e0 < n; e0 += kSortU * blockDim.x) {
    T v[kSortU];
    int li[kSortU];
#pragma unroll
    for (int u = 0; u < kSortU; u++) {
      const int e = e0 + u * blockDim.x;
      li[u] = -1;
      if (e < n) {
        const int r = (int)fb.div(e), ib = e - r * it.b_n;
        v[u] = inb[rin[r] + ib];
        li[u] = r * pitch + ib;
      }
    }
#pragma unroll
    for (int u = 0; u < kSortU; u++)
      if (li[u] >= 0) lds[li[u]] = v[u];
  }
  __syncthreads();
  // write: A fastest; column q = (C1 * c0_n + C0) * b_n + B
  for (int e0 = threadIdx.x; e0 < n; e0 += kSortU * blockDim.x) {
    T v[kSortU];
    int oa[kSortU];
#pragma unroll
    for (int u = 0; u < kSortU; u++) {
      const int e = e0 + u * blockDim.x;
      oa[u] = -1;
      if (e < n) {
        const int q = (int)fa.div(e), ia = e - q * it.a_n;
        const int r2 = (int)fb.div(q), ib = q - r2 * it.b_n;
        v[u] = lds[(r2 * it.a_n + ia) * pitch + ib];
        oa[u] = cout[q] + ia;
      }
    }
#pragma unroll
    for (int u = 0; u < kSortU; u++)
      if (oa[u] >= 0) outb[oa[u]] = v[u];
  }
}

template <typename T>
static hipError_t sort_tiles_t(const SortTask* tasks, const int32_t* item_task, int64_t nitems, bool direct,
                               const T* in, T* out, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (nitems > 0x7fffffff) return hipErrorInvalidValue;
  if (direct)
    hipLaunchKernelGGL((k_sort_copy<T>), dim3((unsigned)nitems), dim3(256), 0, s, tasks, item_task, in, out);
  else
    hipLaunchKernelGGL((k_sort_tiles<T>), dim3((unsigned)nitems), dim3(256), 0, s, tasks, item_task, in, out);
  return hipGetLastError();
}
hipError_t launch_sort_tiles(const SortTask* tasks, const int32_t* item_task, int64_t nitems, bool direct,
                             const double* in, double* out, hipStream_t s) {
  return sort_tiles_t<double>(tasks, item_task, nitems, direct, in, out, s);
}
hipError_t launch_sort_tiles_f32(const SortTask* tasks, const int32_t* item_task, int64_t nitems, bool direct,
                                 const float* in, float* out, hipStream_t s) {
  return sort_tiles_t<float>(tasks, item_task, nitems, direct, in, out, s);
}

// ===========================================================================
// singles-from-singles terms of the sigma builders in one launch:
//   c(a,i) += sum_fn W(fn,ai) r(f,n) + sum_e Fac(a,e) r(e,i) - sum_m r(a,m) Fki(m,i)
// One block per output element (a,i) of every vector; W is the dense l1 x l1
// block of the (v,o,v,o) operator whose pairs have R1's symmetry (rows and
// columns both in R1's element order), Fac/Fki the (v,v)/(o,o) Fock-like
// blocks (absent for SF-CIS).  FP64; fixed-order block reductions.
// ===========================================================================
template <int NV>
__global__ __launch_bounds__(256) void k_singles(SinglesArgs g, const double* __restrict__ R, double* __restrict__ S) {
  const int64_t col = blockIdx.x;
  const int v0 = blockIdx.y * NV;
  const int nv = min(NV, g.nvec - v0);
  int h = 0;
  while (h + 1 < g.nblk && col >= g.start[h + 1]) h++;
  const int64_t loc = col - g.start[h];
  const int na = g.na[h];
  const int a = (int)(loc % na), i = (int)(loc / na);
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) acc[v] = 0.0;
  const double* Rv = R + (int64_t)v0 * g.stride;
  if (g.W) {
    const double* wc = g.W + col * g.l1;
    for (int64_t k = threadIdx.x; k < g.l1; k += blockDim.x) {
      const double w = wc[k];
#pragma unroll
      for (int v = 0; v < NV; v++)
        if (v < nv) acc[v] += w * Rv[(int64_t)v * g.stride + k];
    }
  }
  if (g.Fac) {
    const double* f = g.Fac + g.fac_off[h] + a;  // Fac(a,e) at + na*e
    const double* ri = Rv + g.start[h] + (int64_t)na * i;  // r(e,i) at + e
    for (int e = threadIdx.x; e < na; e += blockDim.x) {
      const double fe = f[(int64_t)na * e];
#pragma unroll
      for (int v = 0; v < NV; v++)
        if (v < nv) acc[v] += fe * ri[(int64_t)v * g.stride + e];
    }
    const int ni = g.ni[h];
    const double* k2 = g.Fki + g.fki_off[h] + (int64_t)ni * i;  // Fki(m,i) at + m
    const double* ra = Rv + g.start[h] + a;                     // r(a,m) at + na*m
    for (int m = threadIdx.x; m < ni; m += blockDim.x) {
      const double km = k2[m];
#pragma unroll
      for (int v = 0; v < NV; v++)
        if (v < nv) acc[v] -= ra[(int64_t)v * g.stride + (int64_t)na * m] * km;
    }
  }
  __shared__ double red[NV][256];
#pragma unroll
  for (int v = 0; v < NV; v++) red[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int v = 0; v < NV; v++) red[v][threadIdx.x] += red[v][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < nv) S[(int64_t)(v0 + threadIdx.x) * g.stride + col] += red[threadIdx.x][0];
}

hipError_t launch_singles(const SinglesArgs& g, const double* R, double* S, hipStream_t s) {
  if (g.l1 <= 0 || g.nvec <= 0) return hipSuccess;
  if (g.l1 > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_singles<8>), dim3((unsigned)g.l1, (g.nvec + 7) / 8), dim3(256), 0, s, g, R, S);
  return hipGetLastError();
}

// ===========================================================================
// grouped GEMM
// ===========================================================================
typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int PAD = 2;

template <typename CT>
struct Acc;
// Result layout of the 16x16 output tile: lane l, register r holds row
// row(l,r), column l%16.  A and B operands: lane l supplies A[l%16][l/16] and
// B[l/16][l%16] for both instructions.
template <>
struct Acc<double> {
  typedef double4_t T;
  static __device__ inline T mfma(double a, double b, T c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Acc<float> {
  typedef float4_t T;
  static __device__ inline T mfma(float a, float b, T c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

__device__ inline int64_t digits3(int64_t v, const int32_t* d, const int64_t* s) {
  int64_t d0 = v % d[0];
  int64_t t = v / d[0];
  int64_t d1 = t % d[1];
  int64_t d2 = t / d[1];
  return d0 * s[0] + d1 * s[1] + d2 * s[2];
}

template <typename T>
__device__ inline double ldg_as_double(const T* p, int64_t i);
template <>
__device__ inline double ldg_as_double<double>(const double* p, int64_t i) {
  return p[i];
}
template <>
__device__ inline double ldg_as_double<float>(const float* p, int64_t i) {
  return (double)p[i];
}

template <typename CT, typename TA, typename TB>
__global__ __launch_bounds__(256) void k_gemm(GemmLaunch g) {
  __shared__ CT As[BK][BM + PAD];
  __shared__ CT Bs[BK][BN + PAD];
  __shared__ int64_t moffA[BM], noffB[BN], koffA[BK], koffB[BK];
  __shared__ int s_mfast, s_nfast;

  const int tid = threadIdx.x;
  const int wg = blockIdx.x;
  // locate the task (binary search on wg_begin)
  int lo = 0, hi = g.ntask - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (g.tasks[mid].wg_begin <= wg) lo = mid;
    else hi = mid - 1;
  }
  const GemmTask& T = g.tasks[lo];
  int local = wg - T.wg_begin;
  const int kp = local % T.ksplit;
  local /= T.ksplit;
  const int tm = local % T.tiles_m;
  const int tn = local / T.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;

  const TA* A = (const TA*)g.A;
  const TB* B = (const TB*)g.B;

  typedef typename Acc<CT>::T AT;
  AT acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = AT{0, 0, 0, 0};

  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

  const int64_t kt0 = (int64_t)kp * T.ktiles_total / T.ksplit;
  const int64_t kt1 = (int64_t)(kp + 1) * T.ktiles_total / T.ksplit;
  int s = T.seg_begin;
  int cur_seg = -1;
  for (int64_t kt = kt0; kt < kt1; kt++) {
    while (kt >= g.segs[s].ktile_begin + g.segs[s].ktiles) s++;
    const GemmSeg& S = g.segs[s];
    if (s != cur_seg) {
      __syncthreads();
      if (tid < BM) {
        int m = m0 + tid;
        moffA[tid] = (m < T.M) ? digits3(m, T.md, S.sa_m) : -1;
      } else if (tid < BM + BN) {
        int n = n0 + tid - BM;
        noffB[tid - BM] = (n < T.N) ? digits3(n, T.nd, S.sb_n) : -1;
      }
      if (tid == 0) {
        s_mfast = (S.sa_m[0] == 1) || (S.sa_k[0] != 1);
        s_nfast = (S.sb_n[0] == 1) || (S.sb_k[0] != 1);
      }
      cur_seg = s;
    }
    const int k0 = (int)(kt - S.ktile_begin) * BK;
    __syncthreads();
    if (tid < BK) {
      int k = k0 + tid;
      koffA[tid] = (k < S.K) ? digits3(k, S.kd, S.sa_k) : -1;
    } else if (tid < 2 * BK) {
      int k = k0 + tid - BK;
      koffB[tid - BK] = (k < S.K) ? digits3(k, S.kd, S.sb_k) : -1;
    }
    __syncthreads();
    // A tile
    if (s_mfast) {
      const int mm = tid & 63, kb = tid >> 6;
      const int64_t mo = moffA[mm];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int kk = kb + 4 * j;
        const int64_t ko = koffA[kk];
        double v = (mo >= 0 && ko >= 0) ? ldg_as_double(A, S.a_off + mo + ko) : 0.0;
        As[kk][mm] = (CT)v;
      }
    } else {
      const int kk = tid & 15, mb = tid >> 4;
      const int64_t ko = koffA[kk];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int mm = mb + 16 * j;
        const int64_t mo = moffA[mm];
        double v = (mo >= 0 && ko >= 0) ? ldg_as_double(A, S.a_off + mo + ko) : 0.0;
        As[kk][mm] = (CT)v;
      }
    }
    // B tile
    if (s_nfast) {
      const int nn = tid & 63, kb = tid >> 6;
      const int64_t no = noffB[nn];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int kk = kb + 4 * j;
        const int64_t ko = koffB[kk];
        double v = (no >= 0 && ko >= 0) ? ldg_as_double(B, S.b_off + no + ko) : 0.0;
        Bs[kk][nn] = (CT)v;
      }
    } else {
      const int kk = tid & 15, nb = tid >> 4;
      const int64_t ko = koffB[kk];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int nn = nb + 16 * j;
        const int64_t no = noffB[nn];
        double v = (no >= 0 && ko >= 0) ? ldg_as_double(B, S.b_off + no + ko) : 0.0;
        Bs[kk][nn] = (CT)v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK / 4; ks++) {
      const int kr = ks * 4 + (lane >> 4);
      CT a0 = As[kr][wm + (lane & 15)];
      CT a1 = As[kr][wm + 16 + (lane & 15)];
      CT b0 = Bs[kr][wn + (lane & 15)];
      CT b1 = Bs[kr][wn + 16 + (lane & 15)];
      acc[0][0] = Acc<CT>::mfma(a0, b0, acc[0][0]);
      acc[0][1] = Acc<CT>::mfma(a0, b1, acc[0][1]);
      acc[1][0] = Acc<CT>::mfma(a1, b0, acc[1][0]);
      acc[1][1] = Acc<CT>::mfma(a1, b1, acc[1][1]);
    }
  }

  // epilogue: lane holds D[row(lane,r)][lane%16] of each 16x16 sub-tile
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + 16 * i + Acc<CT>::row(lane, r);
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m >= T.M || n >= T.N) continue;
        const double v = g.alpha * (double)acc[i][j][r];
        if (T.ksplit > 1) {
          g.partial[T.part_off + (int64_t)kp * T.M * T.N + m + (int64_t)n * T.M] = v;
        } else {
          const int64_t ad = T.c_off + digits3(m, T.md, T.sc_m) + digits3(n, T.nd, T.sc_n);
          if (g.c_f32) {
            float* Cf = (float*)g.C;
            Cf[ad] = (float)((g.beta == 0.0) ? v : g.beta * (double)Cf[ad] + v);
          } else {
            g.C[ad] = (g.beta == 0.0) ? v : g.beta * g.C[ad] + v;
          }
        }
      }
}

__global__ __launch_bounds__(256) void k_splitk_reduce(GemmLaunch g) {
  const GemmTask& T = g.tasks[g.split_tasks[blockIdx.y]];
  const int64_t mn = (int64_t)T.M * T.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    double sum = 0.0;
    for (int kp = 0; kp < T.ksplit; kp++) sum += g.partial[T.part_off + kp * mn + e];
    const int m = (int)(e % T.M), n = (int)(e / T.M);
    const int64_t ad = T.c_off + digits3(m, T.md, T.sc_m) + digits3(n, T.nd, T.sc_n);
    if (g.c_f32) {
      float* Cf = (float*)g.C;
      Cf[ad] = (float)((g.beta == 0.0) ? sum : g.beta * (double)Cf[ad] + sum);
    } else {
      g.C[ad] = (g.beta == 0.0) ? sum : g.beta * g.C[ad] + sum;
    }
  }
}

// grouped launches: the result and beta of a task come from its group entry
__global__ __launch_bounds__(256) void k_splitk_reduce_grp(GemmLaunch g, GroupTab gt) {
  const GemmTask& T = g.tasks[g.split_tasks[blockIdx.y]];
  const GroupEnt& ge = gt.e[T.group];
  double* C = ge.c;
  const double beta = ge.beta;
  const int64_t mn = (int64_t)T.M * T.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    double sum = 0.0;
    for (int kp = 0; kp < T.ksplit; kp++) sum += g.partial[T.part_off + kp * mn + e];
    const int m = (int)(e % T.M), n = (int)(e / T.M);
    const int64_t ad = T.c_off + digits3(m, T.md, T.sc_m) + digits3(n, T.nd, T.sc_n);
    if (ge.c_f32) {
      float* Cf = (float*)C;
      Cf[ad] = (float)((beta == 0.0) ? sum : beta * (double)Cf[ad] + sum);
    } else {
      C[ad] = (beta == 0.0) ? sum : beta * C[ad] + sum;
    }
  }
}

hipError_t launch_splitk_reduce_grouped(const GemmLaunch& g, const GroupTab& gt, hipStream_t s) {
  if (g.nsplit == 0) return hipSuccess;
  int bx = (g.max_split_mn + 255) / 256;
  if (bx > 1024) bx = 1024;
  if (bx < 1) bx = 1;
  hipLaunchKernelGGL(k_splitk_reduce_grp, dim3(bx, g.nsplit), dim3(256), 0, s, g, gt);
  return hipGetLastError();
}

hipError_t launch_gemm(const GemmLaunch& g, hipStream_t s) {
  if (g.nwg == 0) return hipSuccess;
  if (g.cfg != GEMM_V1) {
    hipError_t e2 = launch_gemm2(g, s);
    if (e2 != hipSuccess) return e2;
  } else if (g.prec == PREC_F64) {
    if (g.a_f32 || g.b_f32) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_gemm<double, double, double>), dim3(g.nwg), dim3(256), 0, s, g);
  } else {
    if (g.a_f32 && g.b_f32)
      hipLaunchKernelGGL((k_gemm<float, float, float>), dim3(g.nwg), dim3(256), 0, s, g);
    else if (g.a_f32)
      hipLaunchKernelGGL((k_gemm<float, float, double>), dim3(g.nwg), dim3(256), 0, s, g);
    else if (g.b_f32)
      hipLaunchKernelGGL((k_gemm<float, double, float>), dim3(g.nwg), dim3(256), 0, s, g);
    else
      hipLaunchKernelGGL((k_gemm<float, double, double>), dim3(g.nwg), dim3(256), 0, s, g);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_splitk_reduce(g, s);
}

hipError_t launch_splitk_reduce(const GemmLaunch& g, hipStream_t s) {
  if (g.nsplit == 0) return hipSuccess;
  int bx = (g.max_split_mn + 255) / 256;
  if (bx > 1024) bx = 1024;
  if (bx < 1) bx = 1;
  hipLaunchKernelGGL(k_splitk_reduce, dim3(bx, g.nsplit), dim3(256), 0, s, g);
  return hipGetLastError();
}

// ===========================================================================
// vector kernels
// ===========================================================================
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// dot products of x with nv vectors; grid (kRedBlocks, nv); fixed reduction order
__global__ __launch_bounds__(256) void k_multidot_partial(int64_t n, const double* __restrict__ x,
                                                          const double* __restrict__ y,
                                                          int64_t ystride, double* partial) {
  __shared__ double red[4];
  const int v = blockIdx.y;
  const double* yv = y + (int64_t)v * ystride;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * yv[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[(int64_t)v * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_reduce_rows(const double* partial, int nb, double* out) {
  __shared__ double red[4];
  const int v = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[(int64_t)v * nb + i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[v] = (red[0] + red[1]) + (red[2] + red[3]);
}

hipError_t launch_multidot(int64_t n, const double* x, const double* y, int64_t ystride, int nv,
                           double* partial, double* out, hipStream_t s) {
  if (nv == 0) return hipSuccess;
  hipLaunchKernelGGL(k_multidot_partial, dim3(kRedBlocks, nv), dim3(256), 0, s, n, x, y, ystride,
                     partial);
  hipLaunchKernelGGL(k_reduce_rows, dim3(nv), dim3(256), 0, s, partial, kRedBlocks, out);
  return hipGetLastError();
}

// max |x| with the FIRST index attaining it (BLAS idamax semantics)
__device__ inline void amax_merge(double& v, int64_t& i, double v2, int64_t i2) {
  double a = fabs(v), b = fabs(v2);
  if (b > a || (b == a && i2 < i)) {
    v = v2;
    i = i2;
  }
}

__global__ __launch_bounds__(256) void k_amax_partial(int64_t n, const double* __restrict__ x,
                                                      double* pv, int64_t* pi) {
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  double v = 0.0;
  int64_t idx = INT64_MAX;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    amax_merge(v, idx, x[i], i);
  sv[threadIdx.x] = v;
  si[threadIdx.x] = idx;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      double v2 = sv[threadIdx.x + off];
      int64_t i2 = si[threadIdx.x + off];
      double v1 = sv[threadIdx.x];
      int64_t i1 = si[threadIdx.x];
      amax_merge(v1, i1, v2, i2);
      sv[threadIdx.x] = v1;
      si[threadIdx.x] = i1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

// one block: the largest |pv[i]| (first index on ties) of nb partials, in a fixed order
__device__ inline void k_amax_final_body(int nb, const double* pv, const int64_t* pi, double* ov, int64_t* oi) {
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  double v = 0.0;
  int64_t idx = INT64_MAX;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) amax_merge(v, idx, pv[i], pi[i]);
  sv[threadIdx.x] = v;
  si[threadIdx.x] = idx;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      double v1 = sv[threadIdx.x];
      int64_t i1 = si[threadIdx.x];
      amax_merge(v1, i1, sv[threadIdx.x + off], si[threadIdx.x + off]);
      sv[threadIdx.x] = v1;
      si[threadIdx.x] = i1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *ov = sv[0];
    *oi = si[0];
  }
}

// block b reduces the nb partials of vector b
__global__ __launch_bounds__(256) void k_amax_final(int nb, const double* pv, const int64_t* pi,
                                                    double* ov, int64_t* oi) {
  k_amax_final_body(nb, pv + (int64_t)blockIdx.x * nb, pi + (int64_t)blockIdx.x * nb, ov + blockIdx.x,
                    oi + blockIdx.x);
}

hipError_t launch_amax(int64_t n, const double* x, double* pv, int64_t* pi, double* ov,
                       int64_t* oi, hipStream_t s) {
  hipLaunchKernelGGL(k_amax_partial, dim3(kRedBlocks), dim3(256), 0, s, n, x, pv, pi);
  hipLaunchKernelGGL(k_amax_final, dim3(1), dim3(256), 0, s, kRedBlocks, pv, pi, ov, oi);
  return hipGetLastError();
}

__global__ void k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
__global__ void k_axpy_dev(int64_t n, const double* a, double mul, const double* __restrict__ x,
                           double* __restrict__ y) {
  const double c = mul * a[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += c * x[i];
}
__global__ void k_scal(int64_t n, double a, double* x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}
__global__ void k_scal_dev(int64_t n, const double* a, int invert_sqrt, double* x) {
  const double c = invert_sqrt ? 1.0 / sqrt(a[0]) : a[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= c;
}
__global__ void k_fill(int64_t n, double v, double* x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}
// out_j = sum_k coef[k + j*ldc] X_k  (in the reference's order: k ascending); every thread
// forms up to kLinOut outputs so each X element is read once per grid row
constexpr int kLinOut = 8;
// blockIdx.z = 1 forms the same combinations of a second set (X2 -> out2)
__global__ void k_lincomb(int64_t n, const double* __restrict__ X, int64_t xs, int nk,
                          const double* __restrict__ coef, int ldc, int nout, double* __restrict__ out,
                          const double* __restrict__ X2, double* __restrict__ out2) {
  if (blockIdx.z == 1) {
    X = X2;
    out = out2;
  }
  const int j0 = blockIdx.y * kLinOut;
  const int nj = min(kLinOut, nout - j0);
  // columns past nj repeat the last one (computed, never stored): no branches,
  // so the loads of four basis vectors are in flight together
  const double* cj[kLinOut];
#pragma unroll
  for (int j = 0; j < kLinOut; j++) cj[j] = coef + (int64_t)(j0 + min(j, nj - 1)) * ldc;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s[kLinOut];
#pragma unroll
    for (int j = 0; j < kLinOut; j++) s[j] = 0.0;
    int k = 0;
    for (; k + 4 <= nk; k += 4) {
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; u++) x[u] = X[(int64_t)(k + u) * xs + i];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int j = 0; j < kLinOut; j++) s[j] += cj[j][k + u] * x[u];
    }
    for (; k < nk; k++) {
      const double x = X[(int64_t)k * xs + i];
#pragma unroll
      for (int j = 0; j < kLinOut; j++) s[j] += cj[j][k] * x;
    }
#pragma unroll
    for (int j = 0; j < kLinOut; j++)
      if (j < nj) out[(int64_t)(j0 + j) * n + i] = s[j];
  }
}
// k_lincomb with the coefficients in the kernel arguments (v[k + j*nk])
__global__ void k_lincomb_v(int64_t n, const double* __restrict__ X, int64_t xs, int nk, LinCoef cf, int nout,
                            double* __restrict__ out, const double* __restrict__ X2, double* __restrict__ out2) {
  if (blockIdx.z == 1) {
    X = X2;
    out = out2;
  }
  const int j0 = blockIdx.y * kLinOut;
  const int nj = min(kLinOut, nout - j0);
  int jc[kLinOut];
#pragma unroll
  for (int j = 0; j < kLinOut; j++) jc[j] = (j0 + min(j, nj - 1)) * nk;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s[kLinOut];
#pragma unroll
    for (int j = 0; j < kLinOut; j++) s[j] = 0.0;
    int k = 0;
    for (; k + 4 <= nk; k += 4) {
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; u++) x[u] = X[(int64_t)(k + u) * xs + i];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int j = 0; j < kLinOut; j++) s[j] += cf.v[jc[j] + k + u] * x[u];
    }
    for (; k < nk; k++) {
      const double x = X[(int64_t)k * xs + i];
#pragma unroll
      for (int j = 0; j < kLinOut; j++) s[j] += cf.v[jc[j] + k] * x;
    }
#pragma unroll
    for (int j = 0; j < kLinOut; j++)
      if (j < nj) out[(int64_t)(j0 + j) * n + i] = s[j];
  }
}

__global__ void k_compact_rows(int64_t n, double* __restrict__ P, RowList rows) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    for (int k = 0; k < rows.n; k++)
      if (rows.src[k] != k) P[(int64_t)k * n + i] = P[(int64_t)rows.src[k] * n + i];
}

// lib_cc/mod_davidson.f90:417-428
__global__ void k_precond(int64_t n, double w, const double* __restrict__ d,
                          const double* __restrict__ e, double* __restrict__ p) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double t = w - d[i];
    if (fabs(t) < 1.0e-4) t = (t < 0.0) ? -1.0e-4 : 1.0e-4;
    p[i] = e[i] / t;
  }
}
__global__ void k_hadamard_acc(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                               double* __restrict__ c, int64_t bsb, int64_t bsc) {
  const int v = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    c[(int64_t)v * bsc + i] += a[i] * b[(int64_t)v * bsb + i];
}
__global__ void k_convert_f32(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

static inline unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

hipError_t launch_axpy(int64_t n, double a, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, s, n, a, x, y);
  return hipGetLastError();
}
hipError_t launch_axpy_dev(int64_t n, const double* a, double mul, const double* x, double* y,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_dev, dim3(grid_for(n)), dim3(256), 0, s, n, a, mul, x, y);
  return hipGetLastError();
}
hipError_t launch_scal(int64_t n, double a, double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_scal, dim3(grid_for(n)), dim3(256), 0, s, n, a, x);
  return hipGetLastError();
}
hipError_t launch_scal_dev(int64_t n, const double* a, int inv, double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_scal_dev, dim3(grid_for(n)), dim3(256), 0, s, n, a, inv, x);
  return hipGetLastError();
}
hipError_t launch_fill(int64_t n, double v, double* x, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, n, v, x);
  return hipGetLastError();
}
hipError_t launch_copy(int64_t n, const double* x, double* y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, s, n, x, y);
  return hipGetLastError();
}
hipError_t launch_lincomb(int64_t n, const double* X, int64_t xs, int nk, const double* coef,
                          int ldc, int nout, double* out, hipStream_t s) {
  if (nout == 0) return hipSuccess;
  hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n), (nout + kLinOut - 1) / kLinOut), dim3(256), 0, s, n, X, xs,
                     nk, coef, ldc, nout, out, (const double*)nullptr, (double*)nullptr);
  return hipGetLastError();
}
hipError_t launch_lincomb2(int64_t n, const double* X, const double* X2, int64_t xs, int nk, const double* coef,
                           int ldc, int nout, double* out, double* out2, hipStream_t s) {
  if (nout == 0) return hipSuccess;
  hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n), (nout + kLinOut - 1) / kLinOut, 2), dim3(256), 0, s, n, X, xs,
                     nk, coef, ldc, nout, out, X2, out2);
  return hipGetLastError();
}
hipError_t launch_lincomb2_v(int64_t n, const double* X, const double* X2, int64_t xs, int nk, const LinCoef& cf,
                             int nout, double* out, double* out2, hipStream_t s) {
  if (nout == 0) return hipSuccess;
  if ((int64_t)nk * nout > kLinCoef) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_lincomb_v, dim3(grid_for(n), (nout + kLinOut - 1) / kLinOut, 2), dim3(256), 0, s, n, X, xs,
                     nk, cf, nout, out, X2, out2);
  return hipGetLastError();
}
hipError_t launch_compact_rows(int64_t n, double* P, const RowList& rows, hipStream_t s) {
  if (rows.n <= 0) return hipSuccess;
  if (rows.n > kRootVals) return hipErrorInvalidValue;
  for (int k = 0; k < rows.n; k++)
    if (rows.src[k] < k || (k > 0 && rows.src[k] <= rows.src[k - 1])) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_compact_rows, dim3(grid_for(n)), dim3(256), 0, s, n, P, rows);
  return hipGetLastError();
}
hipError_t launch_precondition(int64_t n, double w, const double* d, const double* e, double* p,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_precond, dim3(grid_for(n)), dim3(256), 0, s, n, w, d, e, p);
  return hipGetLastError();
}
hipError_t launch_hadamard_acc(int64_t n, const double* a, const double* b, double* c, int nbatch,
                               int64_t bsb, int64_t bsc, hipStream_t s) {
  if (n <= 0 || nbatch <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_hadamard_acc, dim3(grid_for(n), nbatch), dim3(256), 0, s, n, a, b, c, bsb,
                     bsc);
  return hipGetLastError();
}
hipError_t launch_convert_f32(int64_t n, const double* x, float* y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_convert_f32, dim3(grid_for(n)), dim3(256), 0, s, n, x, y);
  return hipGetLastError();
}

}  // namespace dpd

namespace dpd {

__global__ __launch_bounds__(256) void k_orbital_energy(DevLayout L, const double* __restrict__ e,
                                                        EorbOff off, int mode, double* __restrict__ out) {
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < L.len;
       o += (int64_t)gridDim.x * blockDim.x) {
    int h[4];
    int64_t x[4];
    decode_elem(L, o, h, x);
    double v;
    if (mode == 1 || mode == 4) {
      v = e[off.offv[h[0]] + x[0]] - e[off.offo[h[1]] + x[1]];
      if (mode == 4) v = -v;
    } else {
      const double ea = e[off.offv[h[0]] + x[0]], eb = e[off.offv[h[1]] + x[1]];
      const double ei = e[off.offo[h[2]] + x[2]], ej = e[off.offo[h[3]] + x[3]];
      v = (mode == 2) ? ea + eb - ei - ej : ei + ej - ea - eb;
    }
    out[o] = v;
  }
}

__global__ void k_div(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ c) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    c[i] = a[i] / b[i];
}

// adjoint of the orbital-energy sums: for every element e of layout L,
// w = coef * x[e] * y[e], and eorb_bar[p_d] += sign[d] * w for every dim d with
// space[d] != 0 (1 virtual, 2 occupied) at its orbital p_d.  Global atomic adds
// (vector memory), so the summation order is not fixed.
__global__ __launch_bounds__(256) void k_eorb_adjoint(DevLayout L, const double* __restrict__ x,
                                                      const double* __restrict__ y, EorbOff off, EorbAdj adj,
                                                      double* eorb_bar) {
  __shared__ DevLayout sl;
  for (int i = threadIdx.x; i < (int)(sizeof(DevLayout) / 4); i += blockDim.x)
    ((int32_t*)&sl)[i] = ((const int32_t*)&L)[i];
  __syncthreads();
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < sl.len; o += (int64_t)gridDim.x * blockDim.x) {
    const double w = adj.coef * x[o] * y[o];
    if (w == 0.0) continue;
    int h[4];
    int64_t xi[4];
    decode_elem(sl, o, h, xi);
#pragma unroll
    for (int d = 0; d < 4; d++) {
      if (adj.space[d] == 0) continue;
      const int64_t pos = (adj.space[d] == 1 ? off.offv[h[d]] : off.offo[h[d]]) + xi[d];
      atomicAdd(&eorb_bar[pos], adj.sign[d] * w);
    }
  }
}

hipError_t launch_eorb_adjoint(const DevLayout& L, const double* x, const double* y, const EorbOff& off,
                               const EorbAdj& adj, double* eorb_bar, hipStream_t s) {
  if (L.len <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_eorb_adjoint, dim3(grid_for(L.len)), dim3(256), 0, s, L, x, y, off, adj, eorb_bar);
  return hipGetLastError();
}

hipError_t launch_orbital_energy(const DevLayout& L, const double* eorb, const EorbOff& off, int mode,
                                 double* out, hipStream_t s) {
  if (L.len == 0) return hipSuccess;
  hipLaunchKernelGGL(k_orbital_energy, dim3(grid_for(L.len)), dim3(256), 0, s, L, eorb, off, mode, out);
  return hipGetLastError();
}

__global__ void k_div_shift(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double shift,
                            double* __restrict__ c) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    c[i] = a[i] / (b[i] - shift);
}

hipError_t launch_div_shift(int64_t n, const double* a, const double* b, double shift, double* c, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_div_shift, dim3(grid_for(n)), dim3(256), 0, s, n, a, b, shift, c);
  return hipGetLastError();
}

hipError_t launch_div(int64_t n, const double* a, const double* b, double* c, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_div, dim3(grid_for(n)), dim3(256), 0, s, n, a, b, c);
  return hipGetLastError();
}

}  // namespace dpd

namespace dpd {

__device__ inline double block_sum_256(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return s;
}

// every block reduces the same kRedBlocks partials in the same order
__device__ inline double reduce_parts(const double* part, double* red) {
  double s = 0.0;
  for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) s += part[i];
  return block_sum_256(s, red);
}

__global__ __launch_bounds__(256) void k_pair_dots(int64_t n, const double* __restrict__ x, int64_t xs,
                                                   const double* __restrict__ y, int64_t ys, double* partial) {
  __shared__ double red[4];
  const int v = blockIdx.y;
  const double* xv = x + (int64_t)v * xs;
  const double* yv = y + (int64_t)v * ys;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += xv[i] * yv[i];
  s = block_sum_256(s, red);
  if (threadIdx.x == 0) partial[(int64_t)v * gridDim.x + blockIdx.x] = s;
}

// out[xi + yj*nx] = X_xi . Y_yj; grid (kRedBlocks, ceil(nx/4), ceil(ny/4)), 4x4 dots per thread
__global__ __launch_bounds__(256) void k_block_dots(int64_t n, const double* __restrict__ X, int64_t xs, int nx,
                                                    const double* __restrict__ Y, int64_t ys, int ny,
                                                    double* partial) {
  __shared__ double red[4][16];
  const int x0 = blockIdx.y * 4, y0 = blockIdx.z * 4;
  const int nxl = min(4, nx - x0), nyl = min(4, ny - y0);
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double xv[4], yv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) xv[a] = a < nxl ? X[(int64_t)(x0 + a) * xs + i] : 0.0;
#pragma unroll
    for (int b = 0; b < 4; b++) yv[b] = b < nyl ? Y[(int64_t)(y0 + b) * ys + i] : 0.0;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) acc[a][b] += xv[a] * yv[b];
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      double v = wave_sum(acc[a][b]);
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][a * 4 + b] = v;
    }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int a = threadIdx.x >> 2, b = threadIdx.x & 3;
    if (a < nxl && b < nyl) {
      const double v = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
      partial[((int64_t)(x0 + a) + (int64_t)(y0 + b) * nx) * gridDim.x + blockIdx.x] = v;
    }
  }
}

hipError_t launch_block_dots(int64_t n, const double* X, int64_t xs, int nx, const double* Y, int64_t ys, int ny,
                             double* partial, double* out, hipStream_t s) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_block_dots, dim3(kRedBlocks, (nx + 3) / 4, (ny + 3) / 4), dim3(256), 0, s, n, X, xs, nx, Y,
                     ys, ny, partial);
  hipLaunchKernelGGL(k_reduce_rows, dim3(nx * ny), dim3(256), 0, s, partial, kRedBlocks, out);
  return hipGetLastError();
}

// x_v *= 1/t and y_v *= 1/t with t = sqrt(nrm2[v]) when |t - 1| > 1e-10
// (lib_cc/mod_davidson.f90:378-386: the Ritz vector is renormalised only then)
__global__ void k_renorm_pair(int64_t n, const double* nrm2, double* __restrict__ x, double* __restrict__ y) {
  const int v = blockIdx.y;
  const double t = sqrt(nrm2[v]);
  if (!(fabs(t - 1.0) > 1e-10)) return;
  const double f = 1.0 / t;
  double* xv = x + (int64_t)v * n;
  double* yv = y + (int64_t)v * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    xv[i] *= f;
    yv[i] *= f;
  }
}

hipError_t launch_renorm_pair(int64_t n, const double* nrm2, double* x, double* y, int nv, hipStream_t s) {
  if (nv == 0) return hipSuccess;
  hipLaunchKernelGGL(k_renorm_pair, dim3(grid_for(n), nv), dim3(256), 0, s, n, nrm2, x, y);
  return hipGetLastError();
}

hipError_t launch_pair_dots_partial(int64_t n, const double* x, const double* y, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_dots, dim3(kRedBlocks, 1), dim3(256), 0, s, n, x, (int64_t)0, y, (int64_t)0, partial);
  return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* partial, int nrows, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_rows, dim3(nrows), dim3(256), 0, s, partial, kRedBlocks, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_resid_amax_partial(int64_t n, const double* __restrict__ aq,
                                                            const double* __restrict__ q, const double* w,
                                                            double* __restrict__ e, double* pv, int64_t* pi) {
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  const int v = blockIdx.y;
  const double wv = w[v];
  const double* aqv = aq + (int64_t)v * n;
  const double* qv = q + (int64_t)v * n;
  double* ev = e + (int64_t)v * n;
  double best = 0.0;
  int64_t bi = INT64_MAX;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double r = aqv[i] + (-wv) * qv[i];
    ev[i] = r;
    amax_merge(best, bi, r, i);
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      double v1 = sv[threadIdx.x];
      int64_t i1 = si[threadIdx.x];
      amax_merge(v1, i1, sv[threadIdx.x + off], si[threadIdx.x + off]);
      sv[threadIdx.x] = v1;
      si[threadIdx.x] = i1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[(int64_t)v * gridDim.x + blockIdx.x] = sv[0];
    pi[(int64_t)v * gridDim.x + blockIdx.x] = si[0];
  }
}

// Residual and preconditioned direction of every root in one pass
// (lib_cc/mod_davidson.f90:396-406, :417-428): e = Aq - w q (not stored),
// p = e / (w - d) (|w - d| clamped to 1e-4) into P, the largest |e| per root,
// and the partial sums of |q|^2 and |p|^2 (kRedBlocks per root).  q and Aq
// are taken as they are: their renormalisation (:378-386) is applied by the
// caller afterwards (e, p and max|e| are linear in it).
__global__ __launch_bounds__(256) void k_ritz_resid(int64_t n, RootVals w, const double* __restrict__ d,
                                                    const double* __restrict__ q, const double* __restrict__ aq,
                                                    double* __restrict__ p, double* pv, int64_t* pi, double* qpart,
                                                    double* ppart) {
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  __shared__ double red[4];
  const int v = blockIdx.y;
  const double wv = w.w[v];
  const double* qv = q + (int64_t)v * n;
  const double* aqv = aq + (int64_t)v * n;
  double* pvv = p + (int64_t)v * n;
  double best = 0.0, sq = 0.0, sp = 0.0;
  int64_t bi = INT64_MAX;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double x = qv[i], y = aqv[i];
    const double r = y + (-wv) * x;
    double tt = wv - d[i];
    if (fabs(tt) < 1.0e-4) tt = (tt < 0.0) ? -1.0e-4 : 1.0e-4;
    const double pp = r / tt;
    pvv[i] = pp;
    sq += x * x;
    sp += pp * pp;
    amax_merge(best, bi, r, i);
  }
  sq = block_sum_256(sq, red);
  sp = block_sum_256(sp, red);
  if (threadIdx.x == 0) {
    qpart[(int64_t)v * gridDim.x + blockIdx.x] = sq;
    ppart[(int64_t)v * gridDim.x + blockIdx.x] = sp;
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      double v1 = sv[threadIdx.x];
      int64_t i1 = si[threadIdx.x];
      amax_merge(v1, i1, sv[threadIdx.x + off], si[threadIdx.x + off]);
      sv[threadIdx.x] = v1;
      si[threadIdx.x] = i1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[(int64_t)v * gridDim.x + blockIdx.x] = sv[0];
    pi[(int64_t)v * gridDim.x + blockIdx.x] = si[0];
  }
}

// per root v (one block each): the largest |e| with its first index, and the
// sums of the |q|^2 and |p|^2 partials, all in a fixed order, into
// stat[v], stat[nv + v], stat[2 nv + v] (and its index into oi[v])
__global__ __launch_bounds__(256) void k_ritz_final(const double* pv, const int64_t* pi, const double* qpart,
                                                    const double* ppart, double* stat, int64_t* oi) {
  __shared__ double red[4];
  const int v = blockIdx.x, nv = gridDim.x;
  const double sq = block_sum_256(qpart[(int64_t)v * kRedBlocks + threadIdx.x] +
                                      qpart[(int64_t)v * kRedBlocks + threadIdx.x + 256], red);
  const double sp = block_sum_256(ppart[(int64_t)v * kRedBlocks + threadIdx.x] +
                                      ppart[(int64_t)v * kRedBlocks + threadIdx.x + 256], red);
  if (threadIdx.x == 0) {
    stat[nv + v] = sq;
    stat[2 * nv + v] = sp;
  }
  k_amax_final_body(kRedBlocks, pv + (int64_t)v * kRedBlocks, pi + (int64_t)v * kRedBlocks, stat + v, oi + v);
}

hipError_t launch_ritz_resid_part(int64_t n, const RootVals& w, const double* d, const double* q, const double* aq,
                                  double* p, int nv, double* pv, int64_t* pi, double* qpart, double* ppart,
                                  double* stat, int64_t* oi, hipStream_t s) {
  if (nv == 0) return hipSuccess;
  if (nv > kRootVals) return hipErrorInvalidValue;
  static_assert(kRedBlocks == 512, "k_ritz_final sums two partials per thread");
  hipLaunchKernelGGL(k_ritz_resid, dim3(kRedBlocks, nv), dim3(256), 0, s, n, w, d, q, aq, p, pv, pi, qpart, ppart);
  hipLaunchKernelGGL(k_ritz_final, dim3(nv), dim3(256), 0, s, pv, pi, qpart, ppart, stat, oi);
  return hipGetLastError();
}

hipError_t launch_resid_amax(int64_t n, const double* aq, const double* q, const double* w, double* e, int nv,
                             double* pv, int64_t* pi, double* ov, int64_t* oi, hipStream_t s) {
  hipLaunchKernelGGL(k_resid_amax_partial, dim3(kRedBlocks, nv), dim3(256), 0, s, n, aq, q, w, e, pv, pi);
  hipLaunchKernelGGL(k_amax_final, dim3(nv), dim3(256), 0, s, kRedBlocks, pv, pi, ov, oi);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_mgs_step(int64_t n, double* __restrict__ p,
                                                  const double* __restrict__ bprev,
                                                  const double* __restrict__ bnext, const double* part_in,
                                                  int mode, double* part_out) {
  __shared__ double red[4];
  const double c = reduce_parts(part_in, red);
  const double f = (mode == 0) ? 1.0 / sqrt(c) : c;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double x = p[i];
    if (mode == 0) x = x * f;
    else x = x + (-f) * bprev[i];
    p[i] = x;
    s += x * (bnext ? bnext[i] : x);
  }
  s = block_sum_256(s, red);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
}

hipError_t launch_mgs_step(int64_t n, double* p, const double* bprev, const double* bnext, const double* part_in,
                           int mode, double* part_out, hipStream_t s) {
  hipLaunchKernelGGL(k_mgs_step, dim3(kRedBlocks), dim3(256), 0, s, n, p, bprev, bnext, part_in, mode, part_out);
  return hipGetLastError();
}

typedef double dots_d4_t __attribute__((ext_vector_type(4)));

// Block dot products C[j + k*nv] = P_j . B_k of NV (<= 8) vectors P (stride n)
// against nb basis vectors B (stride n) on the FP64 matrix cores.  A block
// column covers kDotsMB = 32 basis vectors (two 16-row MFMA tiles): each basis
// vector is read once and the P vectors once per 32 basis vectors (a VALU
// version holding NV x 8 sums a lane read them once per 8, 0.6x the speed at
// 5 vectors and 35 basis vectors).  Per step a block stages a 64-element
// slice of its 32 basis vectors and of the NV P vectors in LDS (loaded a row
// per instruction, 512 contiguous bytes), and each wave multiplies 16 of the 64
// elements with v_mfma_f64_16x16x4f64: A = basis (row m = lane & 15, k =
// lane >> 4), B = P (k = lane >> 4, column n = lane & 15, zero for n >= NV), the
// 16 x 16 result in 4 registers a lane (row (lane >> 4) + 4 r, column lane & 15).
// LDS rows are 66 doubles apart, which makes the fragment reads and the row
// writes free of bank conflicts; the staging is double buffered with one barrier
// per step.  The four waves' results are added in a fixed order.  Partials:
// row j + k*NV, kRedBlocks a row (block x of the grid in slot x).  Block
// columns from cols1 on compute a second, independent set of the same kind
// (P2, B2, nb2 into partial2): two sets in one launch.  Grid (kRedBlocks,
// columns of set 1 + set 2).
constexpr int kDotsMB = 32;
constexpr int kDotsTE = 64;
constexpr int kDotsLd = kDotsTE + 2;
template <int NV>
__global__ __launch_bounds__(256) void k_dots_mfma(int64_t n, const double* __restrict__ P,
                                                   const double* __restrict__ B, int nb, double* partial,
                                                   const int* __restrict__ need, int cols1,
                                                   const double* __restrict__ P2, const double* __restrict__ B2,
                                                   int nb2, double* partial2) {
  static_assert(NV >= 1 && NV <= 8, "up to 8 vectors (rows 0-7 of the B fragment)");
  __shared__ double bs[2][kDotsMB * kDotsLd];
  __shared__ double ps[2][NV * kDotsLd];
  if (need && *need == 0) return;
  int by = blockIdx.y;
  if (by >= cols1) {
    P = P2;
    B = B2;
    nb = nb2;
    partial = partial2;
    by -= cols1;
  }
  const int m0 = by * kDotsMB;
  const int nm = min(kDotsMB, nb - m0);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
  // staging: wave w loads basis rows w + 4 i (i < 8) and P rows w, w + 4
  double rb[kDotsMB / 4], rp[2];
  const int64_t ntiles = (n + kDotsTE - 1) / kDotsTE;
  auto load = [&](int64_t tile) {
    const int64_t e = tile * kDotsTE + lane;
    const bool in = e < n;
#pragma unroll
    for (int i = 0; i < kDotsMB / 4; i++) {
      const int m = wave + 4 * i;
      rb[i] = (in && m < nm) ? B[(int64_t)(m0 + m) * n + e] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int j = wave + 4 * i;
      rp[i] = (in && j < NV) ? P[(int64_t)j * n + e] : 0.0;
    }
  };
  dots_d4_t acc[2];
#pragma unroll
  for (int t = 0; t < 2; t++) acc[t] = dots_d4_t{0.0, 0.0, 0.0, 0.0};
  int64_t tile = blockIdx.x;
  if (tile < ntiles) load(tile);
  for (int buf = 0; tile < ntiles; tile += gridDim.x, buf ^= 1) {
#pragma unroll
    for (int i = 0; i < kDotsMB / 4; i++) bs[buf][(wave + 4 * i) * kDotsLd + lane] = rb[i];
#pragma unroll
    for (int i = 0; i < 2; i++)
      if (wave + 4 * i < NV) ps[buf][(wave + 4 * i) * kDotsLd + lane] = rp[i];
    __syncthreads();
    if (tile + gridDim.x < ntiles) load(tile + gridDim.x);
    // this wave's 16 elements: 4 k-steps of 4
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
      const int k = 16 * wave + 4 * s4 + g;
      const double b = c < NV ? ps[buf][c * kDotsLd + k] : 0.0;
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const double a = bs[buf][(16 * t + c) * kDotsLd + k];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  // the four waves' 32 x NV results, added in a fixed order (staging buffers reused)
  __syncthreads();
  double* red = &bs[0][0];  // [wave][t][r][lane]: 4 x 2 x 4 x 64 doubles
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) red[((wave * 2 + t) * 4 + r) * 64 + lane] = acc[t][r];
  __syncthreads();
  for (int q = threadIdx.x; q < 2 * 4 * 64; q += blockDim.x) {
    const int t = q / 256, r = (q / 64) % 4, l = q % 64;
    const int m = 16 * t + (l >> 4) + 4 * r, j = l & 15;
    if (j < NV && m < nm) {
      const double v = (red[((0 * 2 + t) * 4 + r) * 64 + l] + red[((1 * 2 + t) * 4 + r) * 64 + l]) +
                       (red[((2 * 2 + t) * 4 + r) * 64 + l] + red[((3 * 2 + t) * 4 + r) * 64 + l]);
      partial[((int64_t)j + (int64_t)(m0 + m) * NV) * kRedBlocks + blockIdx.x] = v;
    }
  }
}

template <int NV>
static void launch_dots_nv(int64_t n, const double* P, const double* B, int nb, double* partial, hipStream_t s,
                           const int* need, const double* P2, const double* B2, int nb2) {
  const int cols1 = (nb + kDotsMB - 1) / kDotsMB, cols2 = (nb2 + kDotsMB - 1) / kDotsMB;
  hipLaunchKernelGGL((k_dots_mfma<NV>), dim3(kRedBlocks, cols1 + cols2), dim3(256), 0, s, n, P, B, nb, partial,
                     need, cols1, P2, B2, nb2, partial + (int64_t)NV * nb * kRedBlocks);
}

static void launch_dots_any(int64_t n, const double* P, int nv, const double* B, int nb, double* partial,
                            hipStream_t s, const int* need, const double* P2, const double* B2, int nb2) {
  static_assert(kMgsMulti == 8, "one instantiation per vector count");
  switch (nv) {
    case 1: launch_dots_nv<1>(n, P, B, nb, partial, s, need, P2, B2, nb2); break;
    case 2: launch_dots_nv<2>(n, P, B, nb, partial, s, need, P2, B2, nb2); break;
    case 3: launch_dots_nv<3>(n, P, B, nb, partial, s, need, P2, B2, nb2); break;
    case 4: launch_dots_nv<4>(n, P, B, nb, partial, s, need, P2, B2, nb2); break;
    case 5: launch_dots_nv<5>(n, P, B, nb, partial, s, need, P2, B2, nb2); break;
    case 6: launch_dots_nv<6>(n, P, B, nb, partial, s, need, P2, B2, nb2); break;
    case 7: launch_dots_nv<7>(n, P, B, nb, partial, s, need, P2, B2, nb2); break;
    default: launch_dots_nv<8>(n, P, B, nb, partial, s, need, P2, B2, nb2); break;
  }
}

hipError_t launch_dots_multi(int64_t n, const double* P, int nv, const double* B, int nb, double* partial,
                             double* out, hipStream_t s, const int* need) {
  if (nv <= 0 || nb <= 0) return hipSuccess;
  if (nv > kMgsMulti) return hipErrorInvalidValue;
  launch_dots_any(n, P, nv, B, nb, partial, s, need, nullptr, nullptr, 0);
  hipLaunchKernelGGL(k_reduce_rows, dim3(nv * nb), dim3(256), 0, s, partial, kRedBlocks, out);
  return hipGetLastError();
}

hipError_t launch_dots_pair(int64_t n, int nv, const double* P1, const double* B1, int nb1, const double* P2,
                            const double* B2, int nb2, double* partial, double* out, hipStream_t s) {
  if (nv <= 0 || nv > kMgsMulti || nb1 < 0 || nb2 < 0) return hipErrorInvalidValue;
  if (nb1 + nb2 == 0) return hipSuccess;
  launch_dots_any(n, P1, nv, B1, nb1, partial, s, nullptr, P2, B2, nb2);
  hipLaunchKernelGGL(k_reduce_rows, dim3(nv * (nb1 + nb2)), dim3(256), 0, s, partial, kRedBlocks, out);
  return hipGetLastError();
}

// one projection of classical Gram-Schmidt for nv <= kMgsMulti directions:
// P_j <- f_j (P_j - sum_k C[j + k*nv] B_k) (k ascending) with f_j =
// 1/sqrt(pn2[src[j]]) when pn2 is given (C then holds the projections of the
// unnormalised P_j: the result is the projection of P_j / |P_j|), else 1; and
// the partials of the Gram matrix of the new P_j (rows j + k*nv, j <= k, as
// k_tri_apply writes them).  need given and *need == 0: nothing is done.
// Grid: kRedBlocks blocks of 256.
template <int NV>
__global__ __launch_bounds__(256) void k_cgs_update(int64_t n, const double* __restrict__ B, int nb,
                                                    const double* __restrict__ C, double* __restrict__ P,
                                                    const int* __restrict__ need, const double* pn2, MgsSrc src,
                                                    double* __restrict__ gram) {
  constexpr int KU = 8;  // basis vectors loaded together (in flight) per element
  __shared__ double red[4][NV * NV];
  if (need && *need == 0) return;
  double f[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) f[j] = pn2 ? 1.0 / sqrt(pn2[src.src[j]]) : 1.0;
  double g[NV][NV];
#pragma unroll
  for (int j = 0; j < NV; j++)
#pragma unroll
    for (int k = 0; k < NV; k++) g[j][k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) s[j] = 0.0;
    int k = 0;
    for (; k + KU <= nb; k += KU) {
      double b[KU];
#pragma unroll
      for (int u = 0; u < KU; u++) b[u] = B[(int64_t)(k + u) * n + i];
#pragma unroll
      for (int u = 0; u < KU; u++)
#pragma unroll
        for (int j = 0; j < NV; j++) s[j] += C[j + (int64_t)(k + u) * NV] * b[u];
    }
    for (; k < nb; k++) {
      const double b = B[(int64_t)k * n + i];
#pragma unroll
      for (int j = 0; j < NV; j++) s[j] += C[j + (int64_t)k * NV] * b;
    }
    double x[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) {
      x[j] = (P[(int64_t)j * n + i] - s[j]) * f[j];
      P[(int64_t)j * n + i] = x[j];
    }
#pragma unroll
    for (int k2 = 0; k2 < NV; k2++)
#pragma unroll
      for (int j = 0; j <= k2; j++) g[j][k2] += x[j] * x[k2];
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NV; k++)
#pragma unroll
    for (int j = 0; j <= k; j++) {
      const double v = wave_sum(g[j][k]);
      if (lane == 0) red[wave][j + k * NV] = v;
    }
  __syncthreads();
  for (int t = threadIdx.x; t < NV * NV; t += blockDim.x) {
    const int j = t % NV, k = t / NV;
    if (j <= k)
      gram[((int64_t)j + (int64_t)k * NV) * kRedBlocks + blockIdx.x] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
  }
}

hipError_t launch_cgs_update(int64_t n, const double* B, int nb, const double* C, int nv, double* P,
                             const int* need, const double* pn2, const MgsSrc& src, double* gram, hipStream_t s) {
  if (nv <= 0 || nv > kMgsMulti || nb < 0 || (nb > 0 && !C) || !gram) return hipErrorInvalidValue;
  for (int j = 0; pn2 && j < nv; j++)
    if (src.src[j] < 0) return hipErrorInvalidValue;
  static_assert(kMgsMulti == 8, "one instantiation per vector count");
#define DPD_CGS(NV_)                                                                                               \
  hipLaunchKernelGGL(k_cgs_update<NV_>, dim3(kRedBlocks), dim3(256), 0, s, n, B, nb, C, P, need, pn2, src, gram); \
  break;
  switch (nv) {
    case 1: DPD_CGS(1)
    case 2: DPD_CGS(2)
    case 3: DPD_CGS(3)
    case 4: DPD_CGS(4)
    case 5: DPD_CGS(5)
    case 6: DPD_CGS(6)
    case 7: DPD_CGS(7)
    default: DPD_CGS(8)
  }
#undef DPD_CGS
  return hipGetLastError();
}

// re-orthogonalisation test of classical Gram-Schmidt ("twice is enough",
// Kahan-Parlett): vectors of unit norm whose projections C (nv x nb, [j + k*nv])
// on an orthonormal basis all have sum_k c_jk^2 <= 1/2 keep at least 1/sqrt(2)
// of their norm, and one projection is as accurate as two; need = 0 then.
// pn2 given: C are the projections of unnormalised vectors P_j with
// |P_j|^2 = pn2[src[j]], and the test is sum_k c_jk^2 <= |P_j|^2 / 2.
__global__ void k_cgs_need(const double* __restrict__ C, int nv, int nb, int* need, const double* pn2,
                           MgsSrc src) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  if (threadIdx.x < nv) {
    const int j = threadIdx.x;
    int r = 0;
#pragma unroll
    for (int q = 0; q < kMgsMulti; q++)
      if (q == j) r = src.src[q];
    const double lim = pn2 ? 0.5 * pn2[r] : 0.5;
    double s = 0.0;
    for (int k = 0; k < nb; k++) s += C[j + (int64_t)k * nv] * C[j + (int64_t)k * nv];
    if (!(s <= lim)) atomicOr(&any, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *need = any;
}

hipError_t launch_cgs_need(const double* C, int nv, int nb, int* need, hipStream_t s, const double* pn2,
                           const MgsSrc& src) {
  if (nv <= 0 || nv > kMgsMulti) return hipErrorInvalidValue;
  for (int j = 0; pn2 && j < nv; j++)
    if (src.src[j] < 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_cgs_need, dim3(1), dim3(64), 0, s, C, nv, nb, need, pn2, src);
  return hipGetLastError();
}

// Cholesky QR step for nv <= kMgsMulti vectors from the partials of their Gram
// matrix G (rows j + k*nv, j <= k): G = L L^T and T = L^-T (upper triangular,
// [i + j*nv]), so that Q = P T has orthonormal columns and L_jj is the norm of
// p_j left after its components along q_0 .. q_j-1 are removed -- the norm the
// sequential Gram-Schmidt compares with 1e-6.  The factorisation stops at the
// first j whose remaining norm, times piv_in[j] when given (the pivot of an
// earlier pass), is below 1e-6 or not a number; columns j.. of T are zero.
// piv_out[j] (if given) = that combined norm, nrm2_out[j] (if given) = its
// square, both 0 from the stopping column on.  Run by all 256 threads of a
// block; T ([i + j*nv]) lands in Ts (LDS, visible to all threads on return).
__device__ inline void chol_tri_block(const double* partial, int nv, const double* piv_in, double* Ts,
                                      double* piv_out, double* nrm2_out) {
  // the nv(nv+1)/2 Gram rows are dealt round-robin to the 4 waves; every load
  // of a wave is issued before the first sum (one memory latency, not one per row)
  constexpr int kRowsPerWave = (kMgsMulti * (kMgsMulti + 1) / 2 + 3) / 4;
  constexpr int kPer = kRedBlocks / 64;
  __shared__ double Gs[kMgsMulti * kMgsMulti];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nr = nv * (nv + 1) / 2;
  double v[kRowsPerWave][kPer];
#pragma unroll
  for (int q = 0; q < kRowsPerWave; q++) {
    const int r = wave + 4 * q;
    int k = 0;
    while ((k + 1) * (k + 2) / 2 <= r) k++;
    const int j = r - k * (k + 1) / 2;
    // rows past nr read row 0 (always present) and are discarded
    const double* p = partial + (r < nr ? (int64_t)j + (int64_t)k * nv : 0) * kRedBlocks;
#pragma unroll
    for (int i = 0; i < kPer; i++) v[q][i] = p[lane + 64 * i];
  }
#pragma unroll
  for (int q = 0; q < kRowsPerWave; q++) {
    const int r = wave + 4 * q;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kPer; i++) s += v[q][i];
    s = wave_sum(s);
    if (lane == 0 && r < nr) {
      int k = 0;
      while ((k + 1) * (k + 2) / 2 <= r) k++;
      const int j = r - k * (k + 1) / 2;
      Gs[j + k * nv] = s;
      Gs[k + j * nv] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
  double G[kMgsMulti][kMgsMulti];
#pragma unroll
  for (int i = 0; i < kMgsMulti; i++)
#pragma unroll
    for (int j = 0; j < kMgsMulti; j++) G[i][j] = (i < nv && j < nv) ? Gs[i + j * nv] : 0.0;
  double L[kMgsMulti][kMgsMulti], Ld[kMgsMulti], f[kMgsMulti];
#pragma unroll
  for (int i = 0; i < kMgsMulti; i++) {
    Ld[i] = 0.0;
    f[i] = (piv_in && i < nv) ? piv_in[i] : 1.0;
#pragma unroll
    for (int j = 0; j < kMgsMulti; j++) L[i][j] = 0.0;
  }
  int nvalid = nv;
#pragma unroll
  for (int j = 0; j < kMgsMulti; j++) {
    if (j < nvalid) {
      double d = G[j][j];
#pragma unroll
      for (int m = 0; m < j; m++) d -= L[j][m] * L[j][m];
      if (!(d * f[j] * f[j] >= 1e-12)) {
        nvalid = j;
      } else {
        const double l = sqrt(d);
        L[j][j] = l;
        Ld[j] = 1.0 / l;
#pragma unroll
        for (int i = j + 1; i < kMgsMulti; i++) {
          if (i < nv) {
            double s = G[i][j];
#pragma unroll
            for (int m = 0; m < j; m++) s -= L[i][m] * L[j][m];
            L[i][j] = s * Ld[j];
          }
        }
      }
    }
  }
  // L^-1 (lower triangular) of the leading nvalid columns
  double Li[kMgsMulti][kMgsMulti];
#pragma unroll
  for (int i = 0; i < kMgsMulti; i++)
#pragma unroll
    for (int j = 0; j < kMgsMulti; j++) Li[i][j] = 0.0;
#pragma unroll
  for (int j = 0; j < kMgsMulti; j++) {
    if (j < nvalid) {
      Li[j][j] = Ld[j];
#pragma unroll
      for (int i = j + 1; i < kMgsMulti; i++) {
        if (i < nvalid) {
          double s = 0.0;
#pragma unroll
          for (int m = j; m < i; m++) s += L[i][m] * Li[m][j];
          Li[i][j] = -s * Ld[i];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kMgsMulti; j++) {
    if (j < nv) {
#pragma unroll
      for (int i = 0; i < kMgsMulti; i++)
        if (i < nv) Ts[i + j * nv] = (j < nvalid && i <= j) ? Li[j][i] : 0.0;
      const double pv = j < nvalid ? L[j][j] * f[j] : 0.0;
      if (piv_out) piv_out[j] = pv;
      if (nrm2_out) nrm2_out[j] = pv * pv;
    }
  }
  }
  __syncthreads();
}

// Y_j = sum_{i<=j} T[i + j*nv] X_i for j < nv <= kMgsMulti (T == nullptr: Y = X;
// Y == nullptr: nothing stored; Y may be X: each thread reads its elements of
// every X_i before it writes); with partial given, also the partials of the
// Gram matrix Y_j . Y_k, j <= k, in rows j + k*nv of kRedBlocks (block x in
// slot x).  With ch.gpart given, T is not read: every block first takes the
// Cholesky QR step of chol_tri_block on those Gram partials (with ch.piv_in),
// and T is that factor, times ch.T_pre on the left when given (a product of
// upper triangular factors); block 0 writes the factor alone to ch.T_out and
// the norms to ch.piv_out / ch.nrm2_out.  Grid: kRedBlocks blocks of 256.
__global__ __launch_bounds__(256) void k_tri_apply(int64_t n, const double* X, int nv, const double* T, double* Y,
                                                   double* partial, TriChol ch) {
  __shared__ double red[4][kMgsMulti * kMgsMulti];
  __shared__ double Ts[kMgsMulti * kMgsMulti], Te[kMgsMulti * kMgsMulti];
  if (ch.gpart) {
    const bool b0 = blockIdx.x == 0;
    chol_tri_block(ch.gpart, nv, ch.piv_in, Ts, b0 ? ch.piv_out : nullptr, b0 ? ch.nrm2_out : nullptr);
    const int q = threadIdx.x;
    if (q < nv * nv) {
      const int i = q % nv, j = q / nv;
      double v = 0.0;
      if (!ch.T_pre) {
        v = Ts[q];
      } else if (i <= j) {
        for (int k = i; k <= j; k++) v += ch.T_pre[i + k * nv] * Ts[k + j * nv];
      }
      Te[q] = v;
      if (b0 && ch.T_out) ch.T_out[q] = Ts[q];
    }
    __syncthreads();
    T = Te;
  }
  double t[kMgsMulti][kMgsMulti];
#pragma unroll
  for (int j = 0; j < kMgsMulti; j++)
#pragma unroll
    for (int i = 0; i < kMgsMulti; i++) t[i][j] = (i <= j && j < nv) ? (T ? T[i + j * nv] : (i == j ? 1.0 : 0.0)) : 0.0;
  double g[kMgsMulti][kMgsMulti];
#pragma unroll
  for (int j = 0; j < kMgsMulti; j++)
#pragma unroll
    for (int k = 0; k < kMgsMulti; k++) g[j][k] = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double x[kMgsMulti], y[kMgsMulti];
#pragma unroll
    for (int i = 0; i < kMgsMulti; i++) x[i] = i < nv ? X[(int64_t)i * n + e] : 0.0;
#pragma unroll
    for (int j = 0; j < kMgsMulti; j++) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i <= j; i++) s += t[i][j] * x[i];
      y[j] = s;
    }
    if (Y) {
#pragma unroll
      for (int j = 0; j < kMgsMulti; j++)
        if (j < nv) Y[(int64_t)j * n + e] = y[j];
    }
    if (partial) {
#pragma unroll
      for (int k = 0; k < kMgsMulti; k++)
#pragma unroll
        for (int j = 0; j <= k; j++) g[j][k] += y[j] * y[k];
    }
  }
  if (!partial) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < kMgsMulti; k++)
#pragma unroll
    for (int j = 0; j <= k; j++)
      if (k < nv) {
        const double v = wave_sum(g[j][k]);
        if (lane == 0) red[wave][j + k * kMgsMulti] = v;
      }
  __syncthreads();
  for (int t = threadIdx.x; t < kMgsMulti * kMgsMulti; t += blockDim.x) {
    const int j = t % kMgsMulti, k = t / kMgsMulti;
    if (j <= k && k < nv)
      partial[((int64_t)j + (int64_t)k * nv) * kRedBlocks + blockIdx.x] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
  }
}

hipError_t launch_tri_apply(int64_t n, const double* X, int nv, const double* T, double* Y, double* partial,
                            hipStream_t s, const TriChol& ch) {
  if (nv <= 0 || nv > kMgsMulti) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tri_apply, dim3(kRedBlocks), dim3(256), 0, s, n, X, nv, T, Y, partial, ch);
  return hipGetLastError();
}


__global__ __launch_bounds__(256) void k_scale_by_norm(int64_t n, const double* __restrict__ p, const double* part,
                                                       double* __restrict__ dst) {
  __shared__ double red[4];
  const double c = reduce_parts(part, red);
  const double f = 1.0 / sqrt(c);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = p[i] * f;
}

hipError_t launch_scale_by_norm(int64_t n, const double* p, const double* part, double* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_by_norm, dim3(kRedBlocks), dim3(256), 0, s, n, p, part, dst);
  return hipGetLastError();
}

__global__ void k_precond_batched(int64_t n, const double* w, const int32_t* idx, const double* __restrict__ d,
                                  const double* __restrict__ e, double* __restrict__ p) {
  const int v = idx[blockIdx.y];
  const double wv = w[v];
  const double* ev = e + (int64_t)v * n;
  double* pv = p + (int64_t)blockIdx.y * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double t = wv - d[i];
    if (fabs(t) < 1.0e-4) t = (t < 0.0) ? -1.0e-4 : 1.0e-4;
    pv[i] = ev[i] / t;
  }
}

// p = (aq_v - w_v q_v) / clamp(w_v - d) for the root v = idx[k] of each output k
__global__ void k_precond_ritz(int64_t n, const double* w, const int32_t* idx, const double* __restrict__ d,
                               const double* __restrict__ aq, const double* __restrict__ q, double* __restrict__ p) {
  const int v = idx[blockIdx.y];
  const double wv = w[v];
  const double* aqv = aq + (int64_t)v * n;
  const double* qv = q + (int64_t)v * n;
  double* pv = p + (int64_t)blockIdx.y * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double r = aqv[i] + (-wv) * qv[i];
    double t = wv - d[i];
    if (fabs(t) < 1.0e-4) t = (t < 0.0) ? -1.0e-4 : 1.0e-4;
    pv[i] = r / t;
  }
}

hipError_t launch_precond_ritz(int64_t n, const double* w, const int32_t* idx, int nv, const double* d,
                               const double* aq, const double* q, double* p, hipStream_t s) {
  if (nv == 0) return hipSuccess;
  hipLaunchKernelGGL(k_precond_ritz, dim3(grid_for(n), nv), dim3(256), 0, s, n, w, idx, d, aq, q, p);
  return hipGetLastError();
}

hipError_t launch_precond_batched(int64_t n, const double* w, const int32_t* idx, int nv, const double* d,
                                  const double* e, double* p, hipStream_t s) {
  if (nv == 0) return hipSuccess;
  hipLaunchKernelGGL(k_precond_batched, dim3(grid_for(n), nv), dim3(256), 0, s, n, w, idx, d, e, p);
  return hipGetLastError();
}

}  // namespace dpd
