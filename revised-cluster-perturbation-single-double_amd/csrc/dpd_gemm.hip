// dpd_gemm.hip -- pipelined grouped DPD GEMM for gfx950 (CDNA4) MFMA.
//
// Same task/segment model as the first-generation kernel in dpd_kernels.hip
// (every task is one output irrep block C(m,n); its K segments are the irrep
// tuples that contract into it; M, N and K are mixed-radix with up to three
// digits and their own strides), restructured for throughput:
//
//  * BM x BN x 16 tiles with BM, BN in {64, 128}; four waves in a 2 x 2 grid,
//    each wave (BM/2) x (BN/2) built from 16x16 MFMA tiles (FP64
//    v_mfma_f64_16x16x4f64 or FP32 v_mfma_f32_16x16x4f32);
//  * two LDS buffers and a register prefetch: the global loads of K tile t+1
//    are issued before the MFMAs of tile t and stored to the other buffer
//    after them, one barrier per K tile;
//  * the K-digit offsets of tile t+2 are decoded into a three-slot LDS ring
//    during tile t by 32 threads, so no thread divides on the load path; the
//    M/N offsets a thread loads from live in its registers and are recomputed
//    only when the K loop enters a new segment;
//  * XCD-aware workgroup order: the hardware deals workgroups round-robin to
//    the eight XCDs; the kernel renumbers them so that consecutive tiles of a
//    task (which share the A row block) run on one XCD and share its L2.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dpd_kernels.hpp"

namespace dpd {

namespace {
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename CT>
struct Mfma;
// lane l supplies A[l%16][l/16] and B[l/16][l%16]; result register r of lane l
// is D[row(l, r)][l%16]
template <>
struct Mfma<double> {
  typedef d4_t T;
  static __device__ inline T op(double a, double b, T c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef f4_t T;
  static __device__ inline T op(float a, float b, T c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename CT, typename T>
__device__ inline CT load_as(const T* p, int64_t i) {
  return (CT)p[i];
}

// mixed-radix offset with 32-bit digit arithmetic (all M/N/K sizes < 2^31)
__device__ inline int64_t digit_off(int v, const int32_t* d, const int64_t* s) {
  const int d0 = v % d[0];
  const int t = v / d[0];
  const int d1 = t % d[1];
  const int d2 = t / d[1];
  return (int64_t)d0 * s[0] + (int64_t)d1 * s[1] + (int64_t)d2 * s[2];
}

constexpr int kBK = 16;
constexpr int kThreads = 256;
}  // namespace

template <typename CT, typename TA, typename TB, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm2(GemmLaunch g) {
  constexpr int PADA = 16, PADB = 16;
  constexpr int EA = BM * kBK / kThreads;  // A elements per thread per tile
  constexpr int EB = BN * kBK / kThreads;
  constexpr int TM = BM / 32, TN = BN / 32;  // 16x16 MFMA tiles per wave
  __shared__ CT As[2][kBK][BM + PADA];
  __shared__ CT Bs[2][kBK][BN + PADB];
  __shared__ int64_t kA[3][kBK], kB[3][kBK];
  __shared__ int64_t cm[BM], cn[BN];

  const int tid = threadIdx.x;
  // XCD-aware renumbering of the workgroup id
  int wg;
  {
    const int x = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int q = g.nwg >> 3, r = g.nwg & 7;
    wg = x * q + (x < r ? x : r) + slot;
  }
  int lo = 0, hi = g.ntask - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (g.tasks[mid].wg_begin <= wg) lo = mid;
    else hi = mid - 1;
  }
  const GemmTask& T = g.tasks[lo];
  int local = wg - T.wg_begin;
  const int tn = local % T.tiles_n;
  local /= T.tiles_n;
  const int tm = local % T.tiles_m;
  const int kp = local / T.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const TA* A = (const TA*)g.A;
  const TB* B = (const TB*)g.B;

  typedef typename Mfma<CT>::T AT;
  AT acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = AT{0, 0, 0, 0};
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);

  const int kt0 = (int)((int64_t)kp * T.ktiles_total / T.ksplit);
  const int kt1 = (int)((int64_t)(kp + 1) * T.ktiles_total / T.ksplit);
  const GemmSeg* segs = g.segs;

  // K offsets of tile t into ring slot `slot` (threads 0..2*kBK-1)
  int sk = T.seg_begin;  // segment tracker of the ring writer
  auto fill_k = [&](int t, int slot) {
    while (t >= segs[sk].ktile_begin + segs[sk].ktiles) sk++;
    if (tid < 2 * kBK) {
      const GemmSeg& S = segs[sk];
      const int k = (t - S.ktile_begin) * kBK + (tid & (kBK - 1));
      const bool isA = tid < kBK;
      int64_t o = -1;
      if (k < S.K) o = digit_off(k, S.kd, isA ? S.sa_k : S.sb_k);
      if (isA) kA[slot][tid] = o;
      else kB[slot][tid - kBK] = o;
    }
  };

  // register-resident M/N offsets of the loader, per segment
  int sl = -1;  // segment the registers describe
  bool amf = true, bnf = true;
  int64_t aoff = 0, boff = 0;
  int64_t mo[EA], no[EB];
  auto enter_seg = [&](int s) {
    const GemmSeg& S = segs[s];
    amf = (S.sa_m[0] == 1) || (S.sa_k[0] != 1);
    bnf = (S.sb_n[0] == 1) || (S.sb_k[0] != 1);
    aoff = S.a_off;
    boff = S.b_off;
    if (amf) {
      const int m = m0 + tid % BM;
      mo[0] = m < T.M ? digit_off(m, T.md, S.sa_m) : -1;
    } else {
#pragma unroll
      for (int j = 0; j < EA; j++) {
        const int m = m0 + tid / kBK + j * (kThreads / kBK);
        mo[j] = m < T.M ? digit_off(m, T.md, S.sa_m) : -1;
      }
    }
    if (bnf) {
      const int n = n0 + tid % BN;
      no[0] = n < T.N ? digit_off(n, T.nd, S.sb_n) : -1;
    } else {
#pragma unroll
      for (int j = 0; j < EB; j++) {
        const int n = n0 + tid / kBK + j * (kThreads / kBK);
        no[j] = n < T.N ? digit_off(n, T.nd, S.sb_n) : -1;
      }
    }
    sl = s;
  };
  int sp = T.seg_begin;  // segment tracker of the loader
  CT ra[EA], rb[EB];
  auto load_tile = [&](int t, int slot) {
    while (t >= segs[sp].ktile_begin + segs[sp].ktiles) sp++;
    if (sp != sl) enter_seg(sp);
    if (amf) {
#pragma unroll
      for (int j = 0; j < EA; j++) {
        const int64_t ko = kA[slot][tid / BM + j * (kThreads / BM)];
        ra[j] = (mo[0] >= 0 && ko >= 0) ? load_as<CT>(A, aoff + mo[0] + ko) : (CT)0;
      }
    } else {
      const int64_t ko = kA[slot][tid % kBK];
#pragma unroll
      for (int j = 0; j < EA; j++) ra[j] = (mo[j] >= 0 && ko >= 0) ? load_as<CT>(A, aoff + mo[j] + ko) : (CT)0;
    }
    if (bnf) {
#pragma unroll
      for (int j = 0; j < EB; j++) {
        const int64_t ko = kB[slot][tid / BN + j * (kThreads / BN)];
        rb[j] = (no[0] >= 0 && ko >= 0) ? load_as<CT>(B, boff + no[0] + ko) : (CT)0;
      }
    } else {
      const int64_t ko = kB[slot][tid % kBK];
#pragma unroll
      for (int j = 0; j < EB; j++) rb[j] = (no[j] >= 0 && ko >= 0) ? load_as<CT>(B, boff + no[j] + ko) : (CT)0;
    }
  };
  auto store_tile = [&](int buf) {
    if (amf) {
#pragma unroll
      for (int j = 0; j < EA; j++) As[buf][tid / BM + j * (kThreads / BM)][tid % BM] = ra[j];
    } else {
#pragma unroll
      for (int j = 0; j < EA; j++) As[buf][tid % kBK][tid / kBK + j * (kThreads / kBK)] = ra[j];
    }
    if (bnf) {
#pragma unroll
      for (int j = 0; j < EB; j++) Bs[buf][tid / BN + j * (kThreads / BN)][tid % BN] = rb[j];
    } else {
#pragma unroll
      for (int j = 0; j < EB; j++) Bs[buf][tid % kBK][tid / kBK + j * (kThreads / kBK)] = rb[j];
    }
  };

  if (kt0 < kt1) {
    fill_k(kt0, 0);
    if (kt0 + 1 < kt1) fill_k(kt0 + 1, 1);
    __syncthreads();
    load_tile(kt0, 0);
    store_tile(0);
    __syncthreads();
    for (int t = kt0; t < kt1; t++) {
      const int i = t - kt0;
      const int buf = i & 1;
      const bool more = t + 1 < kt1;
      if (t + 2 < kt1) fill_k(t + 2, (i + 2) % 3);
      if (more) load_tile(t + 1, (i + 1) % 3);
#pragma unroll
      for (int ks = 0; ks < kBK / 4; ks++) {
        const int kr = ks * 4 + (lane >> 4);
        CT a[TM], b[TN];
#pragma unroll
        for (int x = 0; x < TM; x++) a[x] = As[buf][kr][wm + 16 * x + (lane & 15)];
#pragma unroll
        for (int y = 0; y < TN; y++) b[y] = Bs[buf][kr][wn + 16 * y + (lane & 15)];
#pragma unroll
        for (int x = 0; x < TM; x++)
#pragma unroll
          for (int y = 0; y < TN; y++) acc[x][y] = Mfma<CT>::op(a[x], b[y], acc[x][y]);
      }
      if (more) store_tile(buf ^ 1);
      __syncthreads();
    }
  }

  // epilogue: output offsets of this tile's rows and columns
  const bool split = T.ksplit > 1;
  if (!split) {
    for (int i = tid; i < BM + BN; i += kThreads) {
      if (i < BM) {
        const int m = m0 + i;
        cm[i] = m < T.M ? digit_off(m, T.md, T.sc_m) : -1;
      } else {
        const int n = n0 + i - BM;
        cn[i - BM] = n < T.N ? digit_off(n, T.nd, T.sc_n) : -1;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < TM; x++)
#pragma unroll
    for (int y = 0; y < TN; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ml = wm + 16 * x + Mfma<CT>::row(lane, r);
        const int nl = wn + 16 * y + (lane & 15);
        const int m = m0 + ml, n = n0 + nl;
        if (m >= T.M || n >= T.N) continue;
        const double v = g.alpha * (double)acc[x][y][r];
        if (split) {
          g.partial[T.part_off + (int64_t)kp * T.M * T.N + m + (int64_t)n * T.M] = v;
        } else {
          const int64_t ad = T.c_off + cm[ml] + cn[nl];
          if (g.c_f32) {
            float* Cf = (float*)g.C;
            Cf[ad] = (float)((g.beta == 0.0) ? v : g.beta * (double)Cf[ad] + v);
          } else {
            g.C[ad] = (g.beta == 0.0) ? v : g.beta * g.C[ad] + v;
          }
        }
      }
}

template <typename CT, typename TA, typename TB>
static hipError_t launch_cfg(const GemmLaunch& g, hipStream_t s) {
  switch (g.cfg) {
    case GEMM_64x64:
      hipLaunchKernelGGL((k_gemm2<CT, TA, TB, 64, 64>), dim3(g.nwg), dim3(256), 0, s, g);
      break;
    case GEMM_128x64:
      hipLaunchKernelGGL((k_gemm2<CT, TA, TB, 128, 64>), dim3(g.nwg), dim3(256), 0, s, g);
      break;
    case GEMM_64x128:
      hipLaunchKernelGGL((k_gemm2<CT, TA, TB, 64, 128>), dim3(g.nwg), dim3(256), 0, s, g);
      break;
    case GEMM_128x128:
      hipLaunchKernelGGL((k_gemm2<CT, TA, TB, 128, 128>), dim3(g.nwg), dim3(256), 0, s, g);
      break;
    case GEMM_32x32:
      hipLaunchKernelGGL((k_gemm2<CT, TA, TB, 32, 32>), dim3(g.nwg), dim3(256), 0, s, g);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_gemm2(const GemmLaunch& g, hipStream_t s) {
  if (g.prec == PREC_F64) {
    if (g.a_f32 || g.b_f32) return hipErrorInvalidValue;
    return launch_cfg<double, double, double>(g, s);
  }
  if (g.a_f32 && g.b_f32) return launch_cfg<float, float, float>(g, s);
  if (g.a_f32) return launch_cfg<float, float, double>(g, s);
  if (g.b_f32) return launch_cfg<float, double, float>(g, s);
  return launch_cfg<float, double, double>(g, s);
}

}  // namespace dpd
