// dpd_engine.hpp -- context, device memory and the contraction planner.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "dpd_common.hpp"
#include "dpd_kernels.hpp"

struct dpd_ctx;

namespace dpd {

#define DPD_HIP_TRY(expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return fail(DPD_EHIP, "%s: %s (hip %d)", #expr, hipGetErrorString(e_), (int)e_); \
  } while (0)

// device buffer owned by RAII
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t b);
};

// One GEMM stage of a contraction, independent of the data pointers.
// packed copy of one operand side, kept for operands that do not change
// between calls (identified by data pointer and a caller-given id)
struct PackCache {
  const void* src = nullptr;
  uint64_t id = 0;
  DevBuf buf;
};
// the tiling of a plan for the packed-panel kernel in one precision
struct GemmVariant {
  bool built = false;
  bool tuned = false;  // tile shape chosen by timing (autotuner)
  bool split_k = true;  // long-K tasks split over workgroups (else one per tile)
  int mode = 0, cfg = 0;
  std::vector<GemmTask> tasks;
  std::vector<int32_t> split;
  int nwg = 0, max_split_mn = 0, npan_a = 0, npan_b = 0;
  int64_t partial_elems = 0, pa_elems = 0, pb_elems = 0;  // panel elements (BF16x3: both planes)
  DevBuf dev;
  GemmTask* d_tasks = nullptr;
  int32_t* d_split = nullptr;
  // per side, one packed copy per static operand (problems of the same shape
  // share the plan, each keeps its own panels)
  std::vector<std::unique_ptr<PackCache>> cache[2];
  // per side, the source code of every panel element of the per-call operand
  // (built on its first pack; composed with the operand's gather when comp)
  DevBuf pcode[2];
  int pcode_state[2] = {0, 0};  // 0 not built, 1 built, -1 not available
  bool pcode_comp[2] = {false, false};
};
// drops the packed copies of V (buffers freed)
void clear_pack_cache(GemmVariant& V);

// Generation of the device memory that captured graphs may reference: bumped
// whenever the workspace, a plan's task tables or a packed-panel buffer is
// freed, moved or repacked.  A graph captured at another generation is stale.
extern std::atomic<uint64_t> g_mem_gen;
inline void bump_mem_gen() { g_mem_gen.fetch_add(1, std::memory_order_relaxed); }
// set by a thread while it captures a graph: every path that would allocate,
// free or repack fails instead (before any HIP call)
extern thread_local bool g_capturing;

struct GemmPlan {
  std::vector<GemmTask> tasks;
  std::vector<GemmSeg> segs;
  std::vector<int32_t> split;
  int nwg = 0;
  int cfg = 0;       // GemmCfg tile configuration chosen by the planner
  int cfg_req = -1;  // requested configuration (-1: planner's choice)
  int split_req = -1;  // split-K: -1 automatic (timed with the tile), 0 off, 1 on
  int max_split_mn = 0;
  int64_t partial_elems = 0;
  int64_t flops = 0;    // 2*M*N*K summed (algorithmic, per launch)
  int64_t c_elems = 0;  // M*N summed: result elements written per launch
  DevBuf dev;         // tasks | segs | split
  GemmTask* d_tasks = nullptr;
  GemmSeg* d_segs = nullptr;
  int32_t* d_split = nullptr;
  GemmVariant pk[3];  // packed-panel tilings per precision (built on first use)
  int upload();
};

// Operand of a contraction: descriptor (rank 2 or 4), labels, data, and an
// optional batch of `nbatch` tensors `bstride` elements apart.
struct Operand {
  Desc d;
  int rank = 4;
  char lab[4];
  const double* data = nullptr;
  const float* data_f32 = nullptr;  // optional FP32 copy (mixed precision)
  uint64_t static_id = 0;  // nonzero: contents never change under this id (packed once)
  int64_t bstride = 0;
  bool batched = false;
};

// Build the grouped-GEMM plan of v3 = f1*(v1 x v2) + f2*v3 (labels already checked).
int build_gemm_plan(const Sym& S, const Operand& o1, const Operand& o2, const Operand& o3,
                    int nbatch, GemmPlan& plan);

// Full description of a mod_contract call (vv2v_444 / 424 / 442), reusable.
struct ContractPlan {
  int kind = 444;  // 444, 424, 442
  int anti = 0;
  // expansions of operands (iexpso semantics)
  bool exp1 = false, exp2 = false;
  Desc d1x, d2x;    // descriptors after expansion
  Desc d3x;         // descriptor of the scratch result (anti != 0)
  MapArgs map1, map2;  // expansion gathers
  MapArgs epi;         // antisymmetrise + compress + accumulate into v3
  // precomputed source codes of map1, map2, epi (shared by plans with the same
  // layouts through the context's map cache); null until first used, or when
  // a gather keeps decoding its indices (input too long for the code)
  const int32_t* code[3] = {nullptr, nullptr, nullptr};
  bool code_tried[3] = {false, false, false};
  bool epi_on = false;  // the GEMM writes a scratch result that epi maps into v3
  // labels of the GEMM operands (differ from the caller's when regrouped)
  char lab1x[4], lab2x[4], lab3x[4];
  GemmPlan gemm;
  int64_t len1x = 0, len2x = 0, len3x = 0;
  int nbatch = 1;
  bool batch1 = false, batch2 = false;
  int64_t bstride1 = 0, bstride2 = 0, bstride3 = 0;
};
// one grouped contraction (the same contraction of several problems): the
// members' plans, the tile shape they share and the merged task list
struct GroupTasks {
  DevBuf dev;  // merged tasks, then the ids of the split-K tasks
  GemmTask* d_tasks = nullptr;
  int32_t* d_split = nullptr;
  int nwg = 0, ntask = 0, nsplit = 0, max_split_mn = 0;
  int64_t partial_elems = 0;
};
// task table of a tiled DPD permutation (plan_sort), kept per layouts and permutation
struct SortPlan {
  std::vector<SortTask> tasks;
  std::vector<char> host;  // the tables as uploaded: tasks, then the task of every work item
  DevBuf dev;
  int64_t nitems = 0;
  bool direct = false;  // no transposing digits: tiles are copied straight
  bool ok = false;      // false: the sort keeps the generic gather kernels
};

struct GroupPlan {
  int cfg = -1;
  bool tuned = false;
  int built_cfg = -1;
  GroupTasks gt;
  std::vector<ContractPlan*> members;
};

}  // namespace dpd

namespace dpd {
// per-launch timing of tagged GEMM stages (HIP events on the context stream)
struct ProfRec {
  std::string tag;
  hipEvent_t a, b;
  int64_t flops;
  int prec;
  int cfg;        // GemmCfg tile shape of the launch
  int packed;     // packed-panel kernel (1) or gathering kernel (0)
  int64_t bytes;  // algorithmic bytes: operands as the kernel reads them once + result
  int split;      // some task of the launch is split over K (partials + reduce)
  int kind;       // 0 GEMM launch, 1 gather (expansion, sort or epilogue)
};
struct Profiler {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<ProfRec> recs;
};
}  // namespace dpd

// the opaque context of the C-ABI
struct dpd_ctx {
  dpd::Profiler prof;
  dpd::Sym sym;
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t max_threads_block = 256;
  int gemm_cfg = -1;  // GEMM tile configuration override (-1: automatic)
  int packing = 1;    // 0 gather kernels only, 1 packed panels when an operand is static, 2 always
  int split_k = -1;   // packed-panel split-K: -1 automatic, 0 off, 1 on
  int data_prec = 8;  // element size of the tensors given to the engines, sorts and xgemm (8 or 4)
  // scratch: grow-only device workspace (bump allocated per top-level call)
  dpd::DevBuf ws;
  size_t ws_top = 0;
  // small device scratch for reductions
  dpd::DevBuf red;
  // plan caches
  std::unordered_map<std::string, std::unique_ptr<dpd::ContractPlan>> contract_cache;
  std::unordered_map<std::string, std::unique_ptr<dpd::GemmPlan>> engine_cache;
  std::unordered_map<std::string, std::unique_ptr<dpd::GroupPlan>> group_cache;
  // source codes of the plans' gathers, keyed by layouts and terms
  std::unordered_map<std::string, std::unique_ptr<dpd::DevBuf>> map_cache;
  size_t map_bytes = 0;
  // scratch results of held-back epilogues (grow-only, one per slot)
  dpd::DevBuf epi_slot[dpd::kEpiMax];
  // tiled permutation plans of the sorts, keyed by layouts and permutation
  std::unordered_map<std::string, std::unique_ptr<dpd::SortPlan>> sort_cache;
  std::vector<char> pinned_scratch;
};

namespace dpd {
// workspace helpers: reserve before use (may synchronise + reallocate)
int ws_reserve(dpd_ctx* c, size_t bytes);
void* ws_take(dpd_ctx* c, size_t bytes);
inline void ws_reset(dpd_ctx* c) { c->ws_top = 0; }
// frees the packed panels of every cached plan whose static-operand id the
// predicate accepts (a destroyed or refreshed problem's integrals)
void drop_pack_entries(dpd_ctx* c, bool (*match)(uint64_t id, uint64_t key), uint64_t key);

// engine level (libsrint vv2v_dpd_* semantics)
// f32data: the three tensors are FP32 arrays (the reference's prec = 4 storage);
// the GEMM then reads FP32, accumulates in FP32 MFMA and writes
// v3 = fl32(f2 * v3 + f1 * C) formed in FP64
int engine_call(dpd_ctx* c, int kind, const Operand& o1, const Operand& o2, const Operand& o3,
                double f1, double f2, int nbatch, bool f32data = false);
// an accumulating epilogue held back so that several contractions into one
// output are finished by one launch (launch_epi_fused): the source code of the
// gather, the scratch result (kept in a context slot until the launch) and the
// output; deferred is false when the call finished its own epilogue
struct EpiDefer {
  int slot = 0;  // context scratch slot for the result (0 .. kEpiMax-1)
  bool deferred = false;
  bool f32 = false;
  EpiSrc src{};
  int64_t n = 0, out_bstride = 0;
  int nbatch = 0;
  double* out = nullptr;
};
// contract surface (mod_contract semantics) with optional batching of one operand;
// defer: hold back an accumulating epilogue that has a source code (see EpiDefer)
int contract_call(dpd_ctx* c, int kind, const Operand& o1, const Operand& o2, Operand o3,
                  double fact, int anti, int nbatch, int prec, int64_t* flops = nullptr,
                  const char* tag = nullptr, EpiDefer* defer = nullptr);
// finishes held-back epilogues of one output in order (one launch)
int epi_flush(dpd_ctx* c, EpiDefer* d, int n, const char* tag);
// the same contraction of n problems (kind, labels and anti equal; sizes may
// differ): one grouped GEMM launch for all of them where every member is packed
int contract_group(dpd_ctx* c, int kind, const Operand* o1, const Operand* o2, const Operand* o3, int n,
                   double fact, int anti, const int* nbatch, int prec, const char* tag = nullptr);
int run_gemm(dpd_ctx* c, GemmPlan& p, const void* A, int a_f32, const void* B, int b_f32, double* C,
             double f1, double f2, int prec, bool c_f32 = false);
int choose_gemm_cfg(const std::vector<GemmTask>& tasks, int requested);
int build_variant(GemmPlan& p, int mode);
struct PackSrc;
int tune_variant(dpd_ctx* c, GemmPlan& p, int mode, const PackSrc& A, const PackSrc& B, double f1, double f2);
// one operand of the packed-panel GEMM: its data, a nonzero id when the contents
// never change (packed once and cached), and optionally the sort that turns it
// into the GEMM operand (fused into the pack: `data` is then the unsorted tensor,
// xlen the sorted length per batch member, bstride the unsorted batch stride)
struct PackSrc {
  const double* data = nullptr;  // floats when f32
  bool f32 = false;
  uint64_t id = 0;
  const MapArgs* map = nullptr;
  int rowmask = 0;  // dims of the sorted operand that are GEMM rows (M or N)
  int64_t xlen = 0, bstride = 0;
  // per-call operands of cached plans: pack through a source code kept with the
  // variant (coded); gcode: the operand is the gather of `data` through the
  // plan's gather code (one term), gxlen elements per batch member, members of
  // `data` gbstride apart; src_elems bounds every source address
  bool coded = false;
  const int32_t* gcode = nullptr;
  int64_t gxlen = 0, gbstride = 0, src_elems = 0;
};
// packed-panel GEMM; uncached operands are packed into workspace the caller reserved
// ev0/ev1 bracket the GEMM kernel; ev_pack is recorded when the packing is done
int run_gemm_packed(dpd_ctx* c, GemmPlan& p, int mode, const PackSrc& A, const PackSrc& B, double* C, double f1,
                    double f2, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                    bool c_f32 = false, hipEvent_t ev_pack = nullptr);
// host part of a packed-panel tiling (task tables, panel offsets, split-K);
// touches no device memory: V.d_tasks / V.d_split stay for the caller to set
int layout_variant(const GemmPlan& p, GemmVariant& V, int mode, int cfg, bool split_k);
// packs one operand side (0: A, 1: B) of variant V into dst (V.d_tasks and
// p.d_segs must point at device copies of the tables)
int pack_side(dpd_ctx* c, GemmPlan& p, GemmVariant& V, int side, const PackSrc& src, void* dst);
int gather(dpd_ctx* c, MapArgs a, const double* in, double* out);
// the same on FP32 arrays (sums formed in FP64, rounded once)
int gather_f32(dpd_ctx* c, MapArgs a, const float* in, float* out);
// labels / checks shared by the C-ABI and the sigma builders
int check_labels(const char* t, int n);
int map444(const char* type1, const char* type2, const char* type3, char* t1, char* t2, char* t3);
}  // namespace dpd
