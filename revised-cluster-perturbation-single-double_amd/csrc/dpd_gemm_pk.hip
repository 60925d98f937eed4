// dpd_gemm_pk.hip -- grouped DPD GEMM on pre-packed operand panels (gfx950).
//
// The DPD operands of a contraction are irrep blocks addressed through
// mixed-radix digits; gathering them inside the GEMM costs address arithmetic
// and narrow loads on every K tile.  Here a pack kernel first copies each
// operand into panels laid out exactly as the GEMM stages them in LDS:
//
//   A panel (task, m tile, K step):  BM x KS elements,  B panel: BN x KS
//   FP64 (KS = 16) and FP32 (KS = 32): row-major [k][m]  (m contiguous)
//   BF16x3 (KS = 32): hi plane [m][k] then lo plane [m][k] (k contiguous),
//                     x = hi + lo with hi = bf16(x), lo = bf16(x - hi)
//
// Panels pad M/N to the tile and every K segment to 16 with zeros, so the GEMM
// main loop is 16-byte loads of contiguous memory, LDS double buffering with a
// register prefetch and one barrier per K step.  Operands that do not change
// between calls (integrals, intermediates) are packed once and kept.
//
// MFMA: FP64 v_mfma_f64_16x16x4f64, FP32 v_mfma_f32_16x16x4f32, BF16x3
// v_mfma_f32_16x16x32_bf16 three times (hi*hi + hi*lo + lo*hi, FP32
// accumulation).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dpd_kernels.hpp"

namespace dpd {

namespace {
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u4_t __attribute__((ext_vector_type(4)));  // one 16-byte chunk

__device__ inline int64_t digit_off(int v, const int32_t* d, const int64_t* s) {
  const int d0 = v % d[0];
  const int t = v / d[0];
  const int d1 = t % d[1];
  const int d2 = t / d[1];
  return (int64_t)d0 * s[0] + (int64_t)d1 * s[1] + (int64_t)d2 * s[2];
}

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__device__ inline uint16_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kThreads = 256;

__device__ inline int find_task(const GemmTask* tasks, int ntask, int key, int which) {
  int lo = 0, hi = ntask - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    int b = which == 0 ? tasks[mid].wg_begin : which == 1 ? tasks[mid].pan_a : tasks[mid].pan_b;
    if (b <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
}  // namespace

// ---------------------------------------------------------------------------
// pack: one block per panel.  The block first tabulates, per K granule of the
// panel (16 k's share a segment), the offset of each of its rows and of each of
// its k's (one mixed-radix decode per row and per k instead of one per element);
// every element is then one add, or, for a fused sort, one fetch from the
// unsorted tensor with the row's and the k's indices.
// ---------------------------------------------------------------------------
template <int MODE, int MAXT>
__global__ __launch_bounds__(256) void k_pack(PackArgs p) {
  // the sort's layouts are indexed by per-lane irreps: keep them in LDS
  __shared__ MapArgs smap;
  if (p.use_map) {
    for (int i = threadIdx.x; i < (int)(sizeof(MapArgs) / 4); i += blockDim.x)
      ((int32_t*)&smap)[i] = ((const int32_t*)&p.map)[i];
  }
  constexpr int KS = MODE == PREC_F64 ? 16 : 32;
  constexpr int NG = KS / 16;    // granules per panel
  __shared__ int64_t rowoff[NG][MAXT];  // row part of the address (-1: padding row)
  __shared__ int64_t koff[KS];          // k part incl. the segment base (-1: padding k)
  __shared__ int8_t rh[NG][MAXT][4], kh[KS][4];
  __shared__ int32_t rx[NG][MAXT][4], kx[KS][4], rb[NG][MAXT];
  __shared__ int sgi[NG];
  __shared__ double tv[KS][MAXT + 1];  // the panel, [k][row]
  const int pid = blockIdx.x;
  const int ti = find_task(p.tasks, p.ntask, pid, 1 + p.side);
  const GemmTask& T = p.tasks[ti];
  const int local = pid - (p.side == 0 ? T.pan_a : T.pan_b);
  const int tile = local / T.steps, step = local % T.steps;
  const int R = p.side == 0 ? T.M : T.N;
  const int r0 = tile * p.tile;
  const int64_t pan_elems = (int64_t)p.tile * KS;
  const int64_t base = (p.side == 0 ? T.pa_off : T.pb_off) + (int64_t)local * pan_elems * (MODE == PREC_BF16X3 ? 2 : 1);
  const int32_t* rd = p.side == 0 ? T.md : T.nd;
  if (threadIdx.x < NG) {
    const int gran = step * NG + threadIdx.x;
    int sg = -1;
    if (gran < T.ktiles_total) {
      sg = T.seg_begin;
      while (gran >= p.segs[sg].ktile_begin + p.segs[sg].ktiles) sg++;
    }
    sgi[threadIdx.x] = sg;
  }
  __syncthreads();
  // rows (one thread per granule and row)
  for (int t = threadIdx.x; t < NG * p.tile; t += kThreads) {
    const int g = t / p.tile, r = t % p.tile;
    const int sg = sgi[g];
    int64_t ro = -1;
    if (sg >= 0 && r0 + r < R) {
      const GemmSeg& S = p.segs[sg];
      ro = digit_off(r0 + r, rd, p.side == 0 ? S.sa_m : S.sb_n);
      if (p.use_map) {
        // indices of the row: decode the element (row, first k of the granule)
        const int gran = step * NG + g;
        const int k0 = (gran - S.ktile_begin) * 16;
        const int64_t ad = (p.side == 0 ? S.a_off : S.b_off) + ro +
                           digit_off(k0, S.kd, p.side == 0 ? S.sa_k : S.sb_k);
        const int64_t b = ad / p.xlen;
        int h[4];
        int64_t x[4];
        decode_elem(smap.out, ad - b * p.xlen, h, x);
        for (int d = 0; d < 4; d++) {
          rh[g][r][d] = (int8_t)h[d];
          rx[g][r][d] = (int32_t)x[d];
        }
        rb[g][r] = (int32_t)b;
      }
    }
    rowoff[g][r] = ro;
  }
  // k's (one thread per k of the panel)
  if (threadIdx.x < KS) {
    const int k = threadIdx.x;
    const int g = k >> 4;
    const int sg = sgi[g];
    int64_t ko = -1;
    if (sg >= 0) {
      const GemmSeg& S = p.segs[sg];
      const int gran = step * NG + g;
      const int ks = (gran - S.ktile_begin) * 16 + (k & 15);
      if (ks < S.K) {
        ko = (p.side == 0 ? S.a_off : S.b_off) + digit_off(ks, S.kd, p.side == 0 ? S.sa_k : S.sb_k);
        if (p.use_map) {
          // indices of the k: decode the element (first row of the task, k)
          const int64_t ad = ko + digit_off(0, rd, p.side == 0 ? S.sa_m : S.sb_n);
          const int64_t b = ad / p.xlen;
          int h[4];
          int64_t x[4];
          decode_elem(smap.out, ad - b * p.xlen, h, x);
          for (int d = 0; d < 4; d++) {
            kh[k][d] = (int8_t)h[d];
            kx[k][d] = (int32_t)x[d];
          }
        }
      }
    }
    koff[k] = ko;
  }
  __syncthreads();
  // read the panel in the order in which the source is contiguous (rows or
  // k's fastest) into LDS, then write it in panel order
  for (int idx = threadIdx.x; idx < pan_elems; idx += kThreads) {
    int r, k;
    if (p.rfast) {
      r = idx % p.tile;
      k = idx / p.tile;
    } else {
      k = idx % KS;
      r = idx / KS;
    }
    const int g = k >> 4;
    double v = 0.0;
    const int64_t ro = rowoff[g][r], ko = koff[k];
    if (p.code_out) {
      // code mode (no fused sort): the source of this element
      int32_t cd = 0;
      if (ro >= 0 && ko >= 0) {
        const int64_t ad = ro + ko;
        if (p.gcode) {
          const int64_t b = ad / p.gxlen, o = ad - b * p.gxlen;
          const int32_t gc = p.gcode[o];
          if (gc != 0) {
            const int64_t a1 = b * p.gbstride + (gc > 0 ? (int64_t)gc : -(int64_t)gc);
            cd = gc > 0 ? (int32_t)a1 : -(int32_t)a1;
          }
        } else {
          cd = (int32_t)(ad + 1);
        }
      }
      const int e = MODE == PREC_BF16X3 ? r * KS + k : k * p.tile + r;
      p.code_out[(int64_t)pid * pan_elems + e] = cd;
      continue;
    }
    if (ro >= 0 && ko >= 0) {
      if (!p.use_map) {
        v = p.src_f32 ? (double)((const float*)p.src)[ro + ko] : p.src[ro + ko];
      } else {
        int hi[4];
        int64_t xi[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
          const bool row = (p.rowmask >> d) & 1;
          hi[smap.term[0].perm[d]] = row ? rh[g][r][d] : kh[k][d];
          xi[smap.term[0].perm[d]] = row ? rx[g][r][d] : kx[k][d];
        }
        double sg;
        const int64_t ia = fetch_addr(smap.in, hi[0], hi[1], hi[2], hi[3], xi[0], xi[1], xi[2], xi[3], &sg);
        if (ia >= 0) v = smap.alpha * (smap.term[0].coef * sg * p.src[(int64_t)rb[g][r] * p.src_bstride + ia]);
      }
    }
    tv[k][r] = v;
  }
  if (p.code_out) return;  // uniform: every thread of the block took the code branch
  __syncthreads();
  for (int e = threadIdx.x; e < pan_elems; e += kThreads) {
    int r, k;
    if (MODE == PREC_BF16X3) {
      r = e / KS;
      k = e % KS;
    } else {
      k = e / p.tile;
      r = e % p.tile;
    }
    const double v = tv[k][r];
    if (MODE == PREC_F64) {
      ((double*)p.dst)[base + e] = v;
    } else if (MODE == PREC_F32) {
      ((float*)p.dst)[base + e] = (float)v;
    } else {
      const float f = (float)v;
      const uint16_t h = bf16_bits(f);
      const uint16_t l = bf16_bits((float)(v - (double)bf16_float(h)));
      ((uint16_t*)p.dst)[base + e] = h;
      ((uint16_t*)p.dst)[base + pan_elems + e] = l;
    }
  }
}

// panels from precomputed source codes: one thread per element of the hi (or
// only) plane, panels back to back (panel pid at pid * pan_elems * planes).
// With dm != 0 (total < 2^31) the panel of element f is found by multiplying
// with a reciprocal, q = (umulhi(f, dm) + f) >> dl (Granlund & Montgomery,
// "Division by invariant integers using multiplication", 1994), instead of a
// 64-bit division.
template <int MODE, typename TIN>
__global__ __launch_bounds__(256) void k_pack_code(int64_t total, int64_t pan_elems, const int32_t* __restrict__ code,
                                                   const TIN* __restrict__ src, void* __restrict__ dst, uint32_t dm,
                                                   uint32_t dl) {
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int32_t cd = code[f];
    double v = 0.0;
    if (cd > 0) v = (double)src[cd - 1];
    else if (cd < 0) v = -(double)src[-(int64_t)cd - 1];
    if (MODE == PREC_F64) {
      ((double*)dst)[f] = v;
    } else if (MODE == PREC_F32) {
      ((float*)dst)[f] = (float)v;
    } else {
      int64_t pid;
      if (dm) {
        const uint32_t f32 = (uint32_t)f;
        pid = (__umulhi(f32, dm) + f32) >> dl;
      } else {
        pid = f / pan_elems;
      }
      const int64_t e = f - pid * pan_elems;
      const int64_t base = pid * 2 * pan_elems + e;
      const float fv = (float)v;
      const uint16_t h = bf16_bits(fv);
      const uint16_t l = bf16_bits((float)(v - (double)bf16_float(h)));
      ((uint16_t*)dst)[base] = h;
      ((uint16_t*)dst)[base + pan_elems] = l;
    }
  }
}

hipError_t launch_pack_code(int mode, int src_f32, int64_t npanel, int64_t pan_elems, const int32_t* code,
                            const void* src, void* dst, hipStream_t s) {
  const int64_t total = npanel * pan_elems;
  if (total <= 0) return hipSuccess;
  int64_t b = (total + 255) / 256;
  if (b > 16384) b = 16384;
  const dim3 grid((unsigned)b), blk(256);
  // the reciprocal of pan_elems (valid for dividends below 2^31)
  uint32_t dm = 0, dl = 0;
  if (total < ((int64_t)1 << 31) && pan_elems >= 1 && pan_elems < ((int64_t)1 << 31)) {
    const uint64_t d = (uint64_t)pan_elems;
    while (((uint64_t)1 << dl) < d) dl++;
    dm = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << dl) - d)) / d + 1);
  }
#define PKC(M)                                                                                                   \
  if (src_f32)                                                                                                   \
    hipLaunchKernelGGL((k_pack_code<M, float>), grid, blk, 0, s, total, pan_elems, code, (const float*)src, dst,  \
                       dm, dl);                                                                                  \
  else                                                                                                           \
    hipLaunchKernelGGL((k_pack_code<M, double>), grid, blk, 0, s, total, pan_elems, code, (const double*)src, dst, \
                       dm, dl);
  switch (mode) {
    case PREC_F64: PKC(PREC_F64) break;
    case PREC_F32: PKC(PREC_F32) break;
    case PREC_BF16X3: PKC(PREC_BF16X3) break;
    default: return hipErrorInvalidValue;
  }
#undef PKC
  return hipGetLastError();
}

template <int MODE>
static hipError_t launch_pack_mode(const PackArgs& p, hipStream_t s) {
  if (p.tile <= 32) hipLaunchKernelGGL((k_pack<MODE, 32>), dim3(p.npanel), dim3(256), 0, s, p);
  else if (p.tile <= 64) hipLaunchKernelGGL((k_pack<MODE, 64>), dim3(p.npanel), dim3(256), 0, s, p);
  else if (p.tile <= 128) hipLaunchKernelGGL((k_pack<MODE, 128>), dim3(p.npanel), dim3(256), 0, s, p);
  else if (p.tile <= 192) hipLaunchKernelGGL((k_pack<MODE, 192>), dim3(p.npanel), dim3(256), 0, s, p);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_pack(const PackArgs& p, hipStream_t s) {
  if (p.npanel == 0) return hipSuccess;
  switch (p.mode) {
    case PREC_F64:
      return launch_pack_mode<PREC_F64>(p, s);
    case PREC_F32:
      return launch_pack_mode<PREC_F32>(p, s);
    case PREC_BF16X3:
      return launch_pack_mode<PREC_BF16X3>(p, s);
    default:
      return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// packed GEMM
// ---------------------------------------------------------------------------
template <int MODE>
struct PkTraits;
template <>
struct PkTraits<PREC_F64> {
  typedef double E;  // element type in LDS
  typedef double CT;
  typedef d4_t AT;
  static constexpr int KS = 16;
  static constexpr int PAD = 16;  // elements; rows land 32 banks apart
  static constexpr int PLANES = 1;
};
template <>
struct PkTraits<PREC_F32> {
  typedef float E;
  typedef float CT;
  typedef f4_t AT;
  static constexpr int KS = 32;
  static constexpr int PAD = 16;
  static constexpr int PLANES = 1;
};
template <>
struct PkTraits<PREC_BF16X3> {
  typedef uint16_t E;
  typedef float CT;
  typedef f4_t AT;
  static constexpr int KS = 32;
  static constexpr int PAD = 8;  // row = 40 bf16 = 80 bytes: 16 rows of a b128 read hit distinct banks
  static constexpr int PLANES = 2;
};

// LDS buffers of the GEMM tile loop: one (measured on the D2h/120 benchmark:
// the 128x64 tiles run 4 instead of 2 waves per SIMD; the ring GEMM gains 16%
// in BF16x3 and the FP64 solve 3%, the pp-ladder is unchanged)
constexpr int pk_nbuf(int mode, int bm, int bn) { return (void)mode, (void)bm, (void)bn, 1; }
#ifndef DPD_PK_STAGES
#define DPD_PK_STAGES 3
#endif
// K steps held in registers ahead of the one being multiplied
constexpr int pk_stages(int mode, int bm, int bn) { return (void)mode, (void)bm, (void)bn, DPD_PK_STAGES; }

// one workgroup's tile of task T (local = workgroup index inside the task)
template <int MODE, int BM, int BN>
__device__ __forceinline__ void gemm_pk_tile(const GemmLaunch& g, const GemmTask& T, int local,
                                             const void* __restrict__ PA, const void* __restrict__ PB,
                                             double* __restrict__ C, const double beta, const bool c32) {
  typedef PkTraits<MODE> Tr;
  typedef typename Tr::E E;
  typedef typename Tr::AT AT;
  constexpr int KS = Tr::KS;
  constexpr bool BF = MODE == PREC_BF16X3;
  // LDS tiles: FP64/FP32 [k][m] rows of BM(+PAD); BF16 [plane][m][k] rows of KS(+PAD)
  constexpr int A_ELEMS = BF ? 2 * BM * (KS + Tr::PAD) : KS * (BM + Tr::PAD);
  constexpr int B_ELEMS = BF ? 2 * BN * (KS + Tr::PAD) : KS * (BN + Tr::PAD);
  constexpr int A_PANEL = BM * KS * Tr::PLANES;  // elements per packed panel
  constexpr int B_PANEL = BN * KS * Tr::PLANES;
  constexpr int EPC = 16 / sizeof(E);  // elements per 16-byte chunk
  constexpr int CA = A_PANEL / EPC / kThreads, CB = B_PANEL / EPC / kThreads;
  static_assert(CA * EPC * kThreads == A_PANEL && CB * EPC * kThreads == B_PANEL, "panel/threads mismatch");
  constexpr int TM = BM / 32, TN = BN / 32;
  // NBUF 2: the next K step is stored into the other LDS buffer while this one
  // is read; NBUF 1: one buffer (half the LDS, twice the workgroups per CU),
  // stored after a barrier (the global loads of the next step still overlap)
  constexpr int NBUF = pk_nbuf(MODE, BM, BN);
  __shared__ __attribute__((aligned(16))) E As[NBUF][A_ELEMS];
  __shared__ __attribute__((aligned(16))) E Bs[NBUF][B_ELEMS];
  __shared__ int64_t cm[BM], cn[BN];

  const int tid = threadIdx.x;
  const int tn = local % T.tiles_n;
  local /= T.tiles_n;
  const int tm = local % T.tiles_m;
  const int kp = local / T.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int s0 = (int)((int64_t)kp * T.steps / T.ksplit);
  const int s1 = (int)((int64_t)(kp + 1) * T.steps / T.ksplit);
  const u4_t* pa = (const u4_t*)((const E*)PA + T.pa_off + ((int64_t)tm * T.steps + s0) * A_PANEL);
  const u4_t* pb = (const u4_t*)((const E*)PB + T.pb_off + ((int64_t)tn * T.steps + s0) * B_PANEL);

  AT acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = AT{0, 0, 0, 0};
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);

  // LDS position of element e of a panel with `rows` rows (FP64/FP32 [k][r],
  // BF16x3 [plane][r][k])
#define PK_LDS_POS(e, rows)                                                                                  \
  (BF ? ((e) / ((rows) * KS)) * (rows) * (KS + Tr::PAD) + (((e) % ((rows) * KS)) / KS) * (KS + Tr::PAD) +    \
            ((e) % KS)                                                                                       \
      : ((e) / (rows)) * ((rows) + Tr::PAD) + ((e) % (rows)))
  // K steps are prefetched STAGES ahead into registers (ring of register
  // slots, indexed at compile time by unrolling the loop STAGES times); the
  // slot of step i is stored into the single LDS buffer before step i runs
  constexpr int STAGES = pk_stages(MODE, BM, BN);
  u4_t ra[STAGES][CA], rb[STAGES][CB];
  const int nst = s1 - s0;
  auto load_step = [&](int slot_unused, int step, u4_t* xa, u4_t* xb) {
    (void)slot_unused;
    const u4_t* qa = pa + (int64_t)step * (A_PANEL / EPC);
    const u4_t* qb = pb + (int64_t)step * (B_PANEL / EPC);
#pragma unroll
    for (int j = 0; j < CA; j++) xa[j] = qa[tid + j * kThreads];
#pragma unroll
    for (int j = 0; j < CB; j++) xb[j] = qb[tid + j * kThreads];
  };
  auto store_step = [&](const u4_t* xa, const u4_t* xb) {
#pragma unroll
    for (int j = 0; j < CA; j++) *(u4_t*)&As[0][PK_LDS_POS((tid + j * kThreads) * EPC, BM)] = xa[j];
#pragma unroll
    for (int j = 0; j < CB; j++) *(u4_t*)&Bs[0][PK_LDS_POS((tid + j * kThreads) * EPC, BN)] = xb[j];
  };
  if (nst > 0) {
#pragma unroll
    for (int u = 0; u < STAGES; u++)
      if (u < nst) load_step(u, u, ra[u], rb[u]);
    store_step(ra[0], rb[0]);
    __syncthreads();
    for (int i0 = 0; i0 < nst; i0 += STAGES) {
#pragma unroll
      for (int u = 0; u < STAGES; u++) {
        const int i = i0 + u;
        if (i < nst) {
          // slot u (step i) is in LDS already: refill it with step i + STAGES
          if (i + STAGES < nst) load_step(u, i + STAGES, ra[u], rb[u]);
      if constexpr (!BF) {
        typedef typename Tr::CT CT;
        const CT* as = (const CT*)As[0];
        const CT* bs = (const CT*)Bs[0];
#pragma unroll
        for (int ks = 0; ks < KS / 4; ks++) {
          const int kr = ks * 4 + (lane >> 4);
          CT a[TM], b[TN];
#pragma unroll
          for (int x = 0; x < TM; x++) a[x] = as[kr * (BM + Tr::PAD) + wm + 16 * x + (lane & 15)];
#pragma unroll
          for (int y = 0; y < TN; y++) b[y] = bs[kr * (BN + Tr::PAD) + wn + 16 * y + (lane & 15)];
#pragma unroll
          for (int x = 0; x < TM; x++)
#pragma unroll
            for (int y = 0; y < TN; y++) {
              if constexpr (MODE == PREC_F64)
                acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
              else
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[x], b[y], acc[x][y], 0, 0, 0);
            }
        }
      } else {
        const int kq = 8 * (lane >> 4);
        bf8_t ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
        for (int x = 0; x < TM; x++) {
          const int row = wm + 16 * x + (lane & 15);
          ah[x] = *(const bf8_t*)&As[0][row * (KS + Tr::PAD) + kq];
          al[x] = *(const bf8_t*)&As[0][BM * (KS + Tr::PAD) + row * (KS + Tr::PAD) + kq];
        }
#pragma unroll
        for (int y = 0; y < TN; y++) {
          const int row = wn + 16 * y + (lane & 15);
          bh[y] = *(const bf8_t*)&Bs[0][row * (KS + Tr::PAD) + kq];
          bl[y] = *(const bf8_t*)&Bs[0][BN * (KS + Tr::PAD) + row * (KS + Tr::PAD) + kq];
        }
#pragma unroll
        for (int x = 0; x < TM; x++)
#pragma unroll
          for (int y = 0; y < TN; y++) {
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
          }
      }
          __syncthreads();  // every wave is done reading the buffer
          if (i + 1 < nst) {
            store_step(ra[(u + 1) % STAGES], rb[(u + 1) % STAGES]);
            __syncthreads();
          }
        }
      }
    }
  }
#undef PK_LDS_POS

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
        // result register r of lane l: row (l>>4)+4r for FP64, 4(l>>4)+r for FP32 accumulators
        const int row = MODE == PREC_F64 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
        const int ml = wm + 16 * x + row;
        const int nl = wn + 16 * y + (lane & 15);
        const int m = m0 + ml, n = n0 + nl;
        if (m >= T.M || n >= T.N) continue;
        const double v = g.alpha * (double)acc[x][y][r];
        if (split) {
          g.partial[T.part_off + (int64_t)kp * T.M * T.N + m + (int64_t)n * T.M] = v;
        } else {
          const int64_t ad = T.c_off + cm[ml] + cn[nl];
          if (c32) {  // FP32 result: the update is formed in FP64 and rounded once
            float* Cf = (float*)C;
            Cf[ad] = (float)((beta == 0.0) ? v : beta * (double)Cf[ad] + v);
          } else {
            C[ad] = (beta == 0.0) ? v : beta * C[ad] + v;
          }
        }
      }
}

// ---------------------------------------------------------------------------
// BF16x3 tile with the A operand read straight from HBM into registers
// (GEMM_128x64D; BN = DPD_PKD_BN).  Each of the 4 waves owns 32 rows of the
// 128-row tile and loads its A fragments from the packed panel (16 rows x 32 k
// per MFMA tile, hi and lo planes: one contiguous KiB each, the layout the pack
// kernel writes), so A never passes through LDS and is never read by two
// waves.  Only B (BN x 32 per K step, hi and lo) is staged, double buffered
// with one barrier per K step; the A fragments of DPD_PKD_PF steps are held in
// registers ahead of the one multiplied.  Measured at D2h/120: the ring GEMM
// 103 instead of 115 us per launch, the pp-ladder 148 instead of 138 (the
// autotuner keeps the staged 128x64 tile there).
// Products are accumulated in the order of gemm_pk_tile (lo*hi, hi*lo, hi*hi).
// ---------------------------------------------------------------------------
#ifndef DPD_PKD_PF
#define DPD_PKD_PF 2
#endif
template <int BN>
__device__ __forceinline__ void gemm_pkd_tile(const GemmLaunch& g, const GemmTask& T, int local,
                                              const void* __restrict__ PA, const void* __restrict__ PB,
                                              double* __restrict__ C, const double beta, const bool c32) {
  constexpr int BM = 128, KS = 32, LDK = KS + 8;  // LDS row = 40 bf16 = 80 bytes
  constexpr int NT = BN / 16;                     // 16-wide column tiles per wave
  constexpr int A_PANEL = BM * KS * 2;            // bf16 elements per packed A panel (hi + lo)
  constexpr int B_PANEL = BN * KS * 2;
  constexpr int CB = B_PANEL / 8 / kThreads;      // 16-byte chunks of B per thread per K step
  static_assert(CB * 8 * kThreads == B_PANEL, "B panel / threads");
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][2 * BN * LDK];
  __shared__ int64_t cm[BM], cn[BN];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tn = local % T.tiles_n;
  local /= T.tiles_n;
  const int tm = local % T.tiles_m;
  const int kp = local / T.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const int s0 = (int)((int64_t)kp * T.steps / T.ksplit);
  const int s1 = (int)((int64_t)(kp + 1) * T.steps / T.ksplit);
  const int nst = s1 - s0;
  const uint16_t* pa = (const uint16_t*)PA + T.pa_off + ((int64_t)tm * T.steps + s0) * A_PANEL;
  const u4_t* pb = (const u4_t*)((const uint16_t*)PB + T.pb_off + ((int64_t)tn * T.steps + s0) * B_PANEL);
  // this lane's A fragment offset inside a panel: row (lane & 15) of the
  // wave's 16-row tile t, k chunk 8 * (lane >> 4)
  const int aoff = (32 * wave + (lane & 15)) * KS + 8 * (lane >> 4);
  const int kq = 8 * (lane >> 4);

  f4_t acc[2][NT];
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int y = 0; y < NT; y++) acc[t][y] = f4_t{0, 0, 0, 0};

  // A fragments of PF K steps in registers (a ring of slots indexed at compile
  // time: the loop below is unrolled PF times)
  constexpr int PF = DPD_PKD_PF;
  bf8_t rh[PF][2], rl[PF][2];
  u4_t rb[CB];
  auto load_a = [&](int step, bf8_t* h, bf8_t* l) {
    const uint16_t* q = pa + (int64_t)step * A_PANEL + aoff;
#pragma unroll
    for (int t = 0; t < 2; t++) {
      h[t] = *(const bf8_t*)(q + 16 * t * KS);
      l[t] = *(const bf8_t*)(q + 16 * t * KS + BM * KS);
    }
  };
  auto load_b = [&](int step) {
    const u4_t* q = pb + (int64_t)step * (B_PANEL / 8);
#pragma unroll
    for (int j = 0; j < CB; j++) rb[j] = q[tid + j * kThreads];
  };
  auto store_b = [&](int buf) {
#pragma unroll
    for (int j = 0; j < CB; j++) {
      const int e = (tid + j * kThreads) * 8;  // element of the panel: [plane][n][k]
      const int plane = e / (BN * KS), rem = e % (BN * KS);
      *(u4_t*)&Bs[buf][plane * BN * LDK + (rem / KS) * LDK + (rem % KS)] = rb[j];
    }
  };
  if (nst > 0) {
    load_b(0);
#pragma unroll
    for (int u = 0; u < PF; u++)
      if (u < nst) load_a(u, rh[u], rl[u]);
    store_b(0);
    __syncthreads();
    for (int i0 = 0; i0 < nst; i0 += PF) {
#pragma unroll
      for (int u = 0; u < PF; u++) {
        const int i = i0 + u;
        if (i < nst) {
          const int buf = i & 1;
          const bool more = i + 1 < nst;
          if (more) load_b(i + 1);
          const uint16_t* bs = Bs[buf];
#pragma unroll
          for (int y = 0; y < NT; y++) {
            const bf8_t bh = *(const bf8_t*)&bs[(16 * y + (lane & 15)) * LDK + kq];
            const bf8_t bl = *(const bf8_t*)&bs[BN * LDK + (16 * y + (lane & 15)) * LDK + kq];
#pragma unroll
            for (int t = 0; t < 2; t++) {
              acc[t][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rl[u][t], bh, acc[t][y], 0, 0, 0);
              acc[t][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh[u][t], bl, acc[t][y], 0, 0, 0);
              acc[t][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh[u][t], bh, acc[t][y], 0, 0, 0);
            }
          }
          // slot u is free again: refill it with step i + PF
          if (i + PF < nst) load_a(i + PF, rh[u], rl[u]);
          if (more) store_b(buf ^ 1);  // the other buffer: every wave left it at the last barrier
          __syncthreads();
        }
      }
    }
  }

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
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int y = 0; y < NT; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ml = 32 * wave + 16 * t + 4 * (lane >> 4) + r;  // FP32 accumulator row
        const int nl = 16 * y + (lane & 15);
        const int m = m0 + ml, n = n0 + nl;
        if (m >= T.M || n >= T.N) continue;
        const double v = g.alpha * (double)acc[t][y][r];
        if (split) {
          g.partial[T.part_off + (int64_t)kp * T.M * T.N + m + (int64_t)n * T.M] = v;
        } else {
          const int64_t ad = T.c_off + cm[ml] + cn[nl];
          if (c32) {  // FP32 result: the update is formed in FP64 and rounded once
            float* Cf = (float*)C;
            Cf[ad] = (float)((beta == 0.0) ? v : beta * (double)Cf[ad] + v);
          } else {
            C[ad] = (beta == 0.0) ? v : beta * C[ad] + v;
          }
        }
      }
}

// XCD-aware order: consecutive workgroup ids (neighbouring tiles that share
// panels) go to one XCD's L2 (dispatch deals workgroups round-robin over 8 XCDs)
__device__ __forceinline__ int xcd_wg(int nwg) {
  const int x = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int q = nwg >> 3, r = nwg & 7;
  return x * q + (x < r ? x : r) + slot;
}

template <int MODE, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_pk(GemmLaunch g, const void* __restrict__ PA,
                                                 const void* __restrict__ PB) {
  const int wg = xcd_wg(g.nwg);
  const GemmTask& T = g.tasks[find_task(g.tasks, g.ntask, wg, 0)];
  gemm_pk_tile<MODE, BM, BN>(g, T, wg - T.wg_begin, PA, PB, g.C, g.beta, g.c_f32 != 0);
}

template <int MODE, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_pk_grp(GemmLaunch g, GroupTab gt) {
  const int wg = xcd_wg(g.nwg);
  const GemmTask& T = g.tasks[find_task(g.tasks, g.ntask, wg, 0)];
  const GroupEnt& e = gt.e[T.group];
  gemm_pk_tile<MODE, BM, BN>(g, T, wg - T.wg_begin, e.pa, e.pb, e.c, e.beta, e.c_f32 != 0);
}

template <int BN>
#ifndef DPD_PKD_OCC
#define DPD_PKD_OCC 2
#endif
__global__ __launch_bounds__(256, DPD_PKD_OCC) void k_gemm_pkd(GemmLaunch g, const void* __restrict__ PA,
                                                     const void* __restrict__ PB) {
  const int wg = xcd_wg(g.nwg);
  const GemmTask& T = g.tasks[find_task(g.tasks, g.ntask, wg, 0)];
  gemm_pkd_tile<BN>(g, T, wg - T.wg_begin, PA, PB, g.C, g.beta, g.c_f32 != 0);
}

template <int BN>
__global__ __launch_bounds__(256, DPD_PKD_OCC) void k_gemm_pkd_grp(GemmLaunch g, GroupTab gt) {
  const int wg = xcd_wg(g.nwg);
  const GemmTask& T = g.tasks[find_task(g.tasks, g.ntask, wg, 0)];
  const GroupEnt& e = gt.e[T.group];
  gemm_pkd_tile<BN>(g, T, wg - T.wg_begin, e.pa, e.pb, e.c, e.beta, e.c_f32 != 0);
}

template <int MODE>
static hipError_t launch_pk_grp_mode(const GemmLaunch& g, const GroupTab& gt, hipStream_t s) {
  if (g.cfg == GEMM_128x64D) {
    if (MODE != PREC_BF16X3) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_gemm_pkd_grp<kPkdBN>), dim3(g.nwg), dim3(256), 0, s, g, gt);
    return hipGetLastError();
  }
  switch (g.cfg) {
    case GEMM_64x64:
      hipLaunchKernelGGL((k_gemm_pk_grp<MODE, 64, 64>), dim3(g.nwg), dim3(256), 0, s, g, gt);
      break;
    case GEMM_128x64:
      hipLaunchKernelGGL((k_gemm_pk_grp<MODE, 128, 64>), dim3(g.nwg), dim3(256), 0, s, g, gt);
      break;
    case GEMM_64x128:
      hipLaunchKernelGGL((k_gemm_pk_grp<MODE, 64, 128>), dim3(g.nwg), dim3(256), 0, s, g, gt);
      break;
    case GEMM_128x128:
      hipLaunchKernelGGL((k_gemm_pk_grp<MODE, 128, 128>), dim3(g.nwg), dim3(256), 0, s, g, gt);
      break;
    case GEMM_32x32:
      hipLaunchKernelGGL((k_gemm_pk_grp<MODE, 32, 32>), dim3(g.nwg), dim3(256), 0, s, g, gt);
      break;
    case GEMM_64x192:
      hipLaunchKernelGGL((k_gemm_pk_grp<MODE, 64, 192>), dim3(g.nwg), dim3(256), 0, s, g, gt);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_gemm_pk_grouped(const GemmLaunch& g, const GroupTab& gt, hipStream_t s) {
  if (g.nwg == 0) return hipSuccess;
  hipError_t e;
  switch (g.prec) {
    case PREC_F64:
      e = launch_pk_grp_mode<PREC_F64>(g, gt, s);
      break;
    case PREC_F32:
      e = launch_pk_grp_mode<PREC_F32>(g, gt, s);
      break;
    case PREC_BF16X3:
      e = launch_pk_grp_mode<PREC_BF16X3>(g, gt, s);
      break;
    default:
      return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  return launch_splitk_reduce_grouped(g, gt, s);
}

template <int MODE>
static hipError_t launch_pk_mode(const GemmLaunch& g, const void* PA, const void* PB, hipStream_t s) {
  if (g.cfg == GEMM_128x64D) {
    if (MODE != PREC_BF16X3) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_gemm_pkd<kPkdBN>), dim3(g.nwg), dim3(256), 0, s, g, PA, PB);
    return hipGetLastError();
  }
  switch (g.cfg) {
    case GEMM_64x64:
      hipLaunchKernelGGL((k_gemm_pk<MODE, 64, 64>), dim3(g.nwg), dim3(256), 0, s, g, PA, PB);
      break;
    case GEMM_128x64:
      hipLaunchKernelGGL((k_gemm_pk<MODE, 128, 64>), dim3(g.nwg), dim3(256), 0, s, g, PA, PB);
      break;
    case GEMM_64x128:
      hipLaunchKernelGGL((k_gemm_pk<MODE, 64, 128>), dim3(g.nwg), dim3(256), 0, s, g, PA, PB);
      break;
    case GEMM_128x128:
      hipLaunchKernelGGL((k_gemm_pk<MODE, 128, 128>), dim3(g.nwg), dim3(256), 0, s, g, PA, PB);
      break;
    case GEMM_32x32:
      hipLaunchKernelGGL((k_gemm_pk<MODE, 32, 32>), dim3(g.nwg), dim3(256), 0, s, g, PA, PB);
      break;
    case GEMM_64x192:
      hipLaunchKernelGGL((k_gemm_pk<MODE, 64, 192>), dim3(g.nwg), dim3(256), 0, s, g, PA, PB);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_gemm_pk(const GemmLaunch& g, const void* PA, const void* PB, hipStream_t s) {
  hipError_t e = launch_gemm_pk_only(g, PA, PB, s);
  if (e != hipSuccess || g.nwg == 0) return e;
  return launch_splitk_reduce(g, s);
}

hipError_t launch_gemm_pk_only(const GemmLaunch& g, const void* PA, const void* PB, hipStream_t s) {
  if (g.nwg == 0) return hipSuccess;
  hipError_t e;
  switch (g.prec) {
    case PREC_F64:
      e = launch_pk_mode<PREC_F64>(g, PA, PB, s);
      break;
    case PREC_F32:
      e = launch_pk_mode<PREC_F32>(g, PA, PB, s);
      break;
    case PREC_BF16X3:
      e = launch_pk_mode<PREC_BF16X3>(g, PA, PB, s);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return e;
}

}  // namespace dpd
