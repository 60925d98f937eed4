// dpd_engine.cpp -- contraction planner and mod_contract surface on the GPU.
//
// A contraction v3 = f1 * v1 (x) v2 + f2 * v3 over DPD tensors is turned into
// one grouped-GEMM launch:
//   * every tensor dimension becomes a "mode": a single orbital index, or a
//     packed pair (iop 0/2) that travels intact between two tensors (the
//     reference's fast paths and macro pairs, vv2v_dpd_444_d.f90:120-235,
//     vv2v_dpd_442.f90:150-200); a batch of Davidson vectors is one more mode;
//   * modes shared by v1 and v3 form the GEMM rows (M), v2 and v3 the columns
//     (N), v1 and v2 the reduction (K);
//   * for every irrep assignment of the M and N modes (one output block) the
//     planner lists every compatible irrep assignment of the K modes as a K
//     segment; inside a segment each tensor is affine in every index (the
//     reference fallback's "task" decomposition, vv2v_dpd_444_d.f90:256-517),
//     so the kernel addresses A, B, C with mixed-radix strides and never packs.
// Plans hold no data pointers and are cached per signature.
#include "dpd_engine.hpp"
#include "dpd_sortplan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>

namespace dpd {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
const char* last_error() { return g_err.c_str(); }

int DevBuf::alloc(size_t b) {
  reset();
  if (b == 0) return DPD_OK;
  if (hipMalloc(&p, b) != hipSuccess) {
    p = nullptr;
    return fail(DPD_ENOMEM, "hipMalloc of %zu bytes failed", b);
  }
  bytes = b;
  return DPD_OK;
}

std::atomic<uint64_t> g_mem_gen{0};
thread_local bool g_capturing = false;

int ws_reserve(dpd_ctx* c, size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (c->ws.bytes >= c->ws_top + bytes) return DPD_OK;
  if (c->ws_top != 0)
    return fail(DPD_ENOMEM, "workspace must be reserved before it is used (internal)");
  if (g_capturing) return fail(DPD_EINVAL, "workspace growth during graph capture (internal)");
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  size_t nb = std::max(bytes, c->ws.bytes * 2);
  bump_mem_gen();
  return c->ws.alloc(nb);
}

void drop_pack_entries(dpd_ctx* c, bool (*match)(uint64_t id, uint64_t key), uint64_t key) {
  auto drop = [&](GemmPlan& g) {
    for (auto& V : g.pk)
      for (auto& side : V.cache) {
        const size_t before = side.size();
        side.erase(std::remove_if(side.begin(), side.end(),
                                  [&](const std::unique_ptr<PackCache>& e) { return e->id && match(e->id, key); }),
                   side.end());
        if (side.size() != before) bump_mem_gen();
      }
  };
  for (auto& kv : c->contract_cache) drop(kv.second->gemm);
  for (auto& kv : c->engine_cache) drop(*kv.second);
}

void clear_pack_cache(GemmVariant& V) {
  bool any = false;
  for (auto& side : V.cache) {
    any = any || !side.empty();
    side.clear();
  }
  if (any) bump_mem_gen();
}

void* ws_take(dpd_ctx* c, size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (c->ws_top + bytes > c->ws.bytes) {
    fail(DPD_ENOMEM, "workspace overrun (internal): %zu + %zu > %zu", c->ws_top, bytes, c->ws.bytes);
    return nullptr;
  }
  void* p = (char*)c->ws.p + c->ws_top;
  c->ws_top += bytes;
  return p;
}

int check_labels(const char* t, int n) {
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (t[i] == t[j]) return fail(DPD_EINVAL, "label string '%.*s' repeats an index", n, t);
  return DPD_OK;
}

// label mapper of libsrint/vv2v_dpd_444.f90:88-133: outputs -> 1,2,5,6,
// contracted -> 3,4 in order of appearance in type1
int map444(const char* type1, const char* type2, const char* type3, char* t1, char* t2, char* t3) {
  char from[6], to[6];
  int nmap = 4;
  const char outs[4] = {'1', '2', '5', '6'};
  for (int i = 0; i < 4; i++) {
    from[i] = type3[i];
    to[i] = outs[i];
  }
  for (int i = 0; i < 4; i++) {
    bool contracted = true;
    for (int j = 0; j < 4; j++)
      if (type1[i] == type3[j]) contracted = false;
    if (contracted) {
      if (nmap >= 6) return fail(DPD_EINVAL, "vv2v_dpd_444: more than two contracted indices");
      from[nmap] = type1[i];
      to[nmap] = nmap == 4 ? '3' : '4';
      nmap++;
    }
  }
  if (nmap != 6) return fail(DPD_EINVAL, "vv2v_dpd_444: exactly two contracted indices are required");
  std::memcpy(t1, type1, 4);
  std::memcpy(t2, type2, 4);
  std::memcpy(t3, type3, 4);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < nmap; j++) {
      if (type1[i] == from[j]) t1[i] = to[j];
      if (type2[i] == from[j]) t2[i] = to[j];
      if (type3[i] == from[j]) t3[i] = to[j];
    }
  return DPD_OK;
}

// ---------------------------------------------------------------------------
// planner
// ---------------------------------------------------------------------------
namespace {

constexpr int kMaxMode = 14;

struct Mode {
  char lab[2];
  bool pair = false;
  bool batch = false;
  int64_t size[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int cls = -1;  // 0 M, 1 N, 2 K, 3 dummy (1-sized)
  int64_t total() const {
    int64_t t = 0;
    for (int h = 0; h < 8; h++) t += size[h];
    return t;
  }
};

struct TV {
  Desc d;
  char lab[4];
  int mode_of_dim[4];
  int pair_mode[2];
  int batch_mode = -1;
  int64_t bstride = 0;
  DevLayout L;
};

int find_mode(const std::vector<Mode>& m, const char* lab, bool pair) {
  for (size_t i = 0; i < m.size(); i++) {
    if (m[i].pair != pair || m[i].batch) continue;
    if (m[i].lab[0] == lab[0] && (!pair || m[i].lab[1] == lab[1])) return (int)i;
  }
  return -1;
}

void pad_labels(TV& t, const Operand& o, char d0, char d1) {
  t.d = o.d;
  if (o.rank == 2) {
    for (int k = 2; k < 4; k++)
      for (int h = 0; h < 8; h++) t.d.pop[k][h] = (h == 0);
    t.d.iop[1] = 1;
    t.lab[0] = o.lab[0];
    t.lab[1] = o.lab[1];
    t.lab[2] = d0;
    t.lab[3] = d1;
  } else {
    std::memcpy(t.lab, o.lab, 4);
  }
}

struct Affine {
  int64_t base = 0;
  int64_t stride[kMaxMode];
  bool valid = false;
};

bool tensor_sym_ok(const Sym& S, const TV& t, const int* mirr) {
  int x = 0;
  for (int s = 0; s < 2; s++) {
    if (t.pair_mode[s] >= 0) x = S.dp[x][mirr[t.pair_mode[s]]];
    else x = S.dp[S.dp[x][mirr[t.mode_of_dim[2 * s]]]][mirr[t.mode_of_dim[2 * s + 1]]];
  }
  return x == t.d.sym - 1;
}

Affine tensor_affine(const Sym& S, const TV& t, const int* mirr, int nmode) {
  Affine a;
  for (int i = 0; i < nmode; i++) a.stride[i] = 0;
  int h[4] = {0, 0, 0, 0};
  int xl, xr;
  if (t.pair_mode[0] >= 0) {
    xl = mirr[t.pair_mode[0]];
  } else {
    h[0] = mirr[t.mode_of_dim[0]];
    h[1] = mirr[t.mode_of_dim[1]];
    xl = S.dp[h[0]][h[1]];
  }
  if (t.pair_mode[1] >= 0) {
    xr = mirr[t.pair_mode[1]];
  } else {
    h[2] = mirr[t.mode_of_dim[2]];
    h[3] = mirr[t.mode_of_dim[3]];
    xr = S.dp[h[2]][h[3]];
  }
  if (S.dp[xl][xr] != t.d.sym - 1) return a;
  const DevLayout& L = t.L;
  const int64_t nl = L.nL[xl];
  a.base = L.blk[xr];
  if (t.pair_mode[0] >= 0) {
    a.stride[t.pair_mode[0]] += 1;
  } else {
    a.base += L.offL[xl][h[1]];
    a.stride[t.mode_of_dim[0]] += 1;
    a.stride[t.mode_of_dim[1]] += L.pop[0][h[0]];
  }
  if (t.pair_mode[1] >= 0) {
    a.stride[t.pair_mode[1]] += nl;
  } else {
    a.base += nl * (int64_t)L.offR[xr][h[3]];
    a.stride[t.mode_of_dim[2]] += nl;
    a.stride[t.mode_of_dim[3]] += nl * (int64_t)L.pop[2][h[2]];
  }
  if (t.batch_mode >= 0) a.stride[t.batch_mode] += t.bstride;
  a.valid = true;
  return a;
}

}  // namespace

int GemmPlan::upload() {
  size_t bt = tasks.size() * sizeof(GemmTask);
  size_t bs = segs.size() * sizeof(GemmSeg);
  size_t bp = split.size() * sizeof(int32_t);
  size_t tot = ((bt + 255) & ~size_t(255)) + ((bs + 255) & ~size_t(255)) + bp + 256;
  int rc = dev.alloc(tot);
  if (rc) return rc;
  char* base = (char*)dev.p;
  d_tasks = (GemmTask*)base;
  d_segs = (GemmSeg*)(base + ((bt + 255) & ~size_t(255)));
  d_split = (int32_t*)((char*)d_segs + ((bs + 255) & ~size_t(255)));
  if (bt) DPD_HIP_TRY(hipMemcpy(d_tasks, tasks.data(), bt, hipMemcpyHostToDevice));
  if (bs) DPD_HIP_TRY(hipMemcpy(d_segs, segs.data(), bs, hipMemcpyHostToDevice));
  if (bp) DPD_HIP_TRY(hipMemcpy(d_split, split.data(), bp, hipMemcpyHostToDevice));
  return DPD_OK;
}

int build_gemm_plan(const Sym& S, const Operand& o1, const Operand& o2, const Operand& o3,
                    int nbatch, GemmPlan& plan) {
  TV T[3];
  pad_labels(T[0], o1, '\x01', '\x02');
  pad_labels(T[1], o2, '\x03', '\x04');
  pad_labels(T[2], o3, '\x05', '\x06');
  std::vector<Mode> modes;
  // a full (iop=1) pair whose two labels appear only as that same ordered pair
  // in every tensor is one GEMM digit (the pair index of its irrep block): the
  // GEMM then runs per pair irrep instead of per irrep of each index
  auto intact_full = [&](int t, int s) -> bool {
    const char p0 = T[t].lab[2 * s], q0 = T[t].lab[2 * s + 1];
    for (int t2 = 0; t2 < 3; t2++)
      for (int s2 = 0; s2 < 2; s2++) {
        const char a = T[t2].lab[2 * s2], b = T[t2].lab[2 * s2 + 1];
        const bool has = a == p0 || a == q0 || b == p0 || b == q0;
        if (has && !(a == p0 && b == q0 && T[t2].d.iop[s2] == 1)) return false;
        // populations must agree index by index (else the per-index modes report it)
        if (has && (std::memcmp(T[t2].d.pop[2 * s2], T[t].d.pop[2 * s], sizeof(int64_t) * 8) != 0 ||
                    std::memcmp(T[t2].d.pop[2 * s2 + 1], T[t].d.pop[2 * s + 1], sizeof(int64_t) * 8) != 0))
          return false;
      }
    return true;
  };
  for (int t = 0; t < 3; t++) {
    TV& tv = T[t];
    for (int s = 0; s < 2; s++) {
      tv.pair_mode[s] = -1;
      if (tv.d.iop[s] != 1 || intact_full(t, s)) {
        char lab[2] = {tv.lab[2 * s], tv.lab[2 * s + 1]};
        int id = find_mode(modes, lab, true);
        int64_t sz[8];
        for (int x = 1; x <= 8; x++)
          sz[x - 1] = (x <= S.nirrep) ? S.irpdso(x, tv.d.pop[2 * s], tv.d.pop[2 * s + 1], tv.d.iop[s]) : 0;
        if (id < 0) {
          Mode m;
          m.lab[0] = lab[0];
          m.lab[1] = lab[1];
          m.pair = true;
          std::memcpy(m.size, sz, sizeof sz);
          modes.push_back(m);
          id = (int)modes.size() - 1;
        } else {
          for (int h = 0; h < 8; h++)
            if (modes[id].size[h] != sz[h])
              return fail(DPD_EPOP, "packed pair '%c%c' has mismatched populations or iop", lab[0], lab[1]);
        }
        tv.pair_mode[s] = id;
        tv.mode_of_dim[2 * s] = tv.mode_of_dim[2 * s + 1] = id;
      } else {
        for (int k = 0; k < 2; k++) {
          char lab[2] = {tv.lab[2 * s + k], 0};
          int id = find_mode(modes, lab, false);
          if (id < 0) {
            Mode m;
            m.lab[0] = lab[0];
            m.lab[1] = 0;
            for (int h = 0; h < 8; h++) m.size[h] = (h < S.nirrep) ? tv.d.pop[2 * s + k][h] : 0;
            modes.push_back(m);
            id = (int)modes.size() - 1;
          } else {
            for (int h = 0; h < S.nirrep; h++)
              if (modes[id].size[h] != tv.d.pop[2 * s + k][h])
                return fail(DPD_EPOP, "index '%c' has mismatched populations in two tensors", lab[0]);
          }
          tv.mode_of_dim[2 * s + k] = id;
        }
      }
    }
  }
  for (size_t i = 0; i < modes.size(); i++)
    if (modes[i].pair)
      for (size_t j = 0; j < modes.size(); j++)
        if (!modes[j].pair && (modes[j].lab[0] == modes[i].lab[0] || modes[j].lab[0] == modes[i].lab[1]))
          return fail(DPD_EIOP,
                      "IOP barrier: packed pair '%c%c' is split or reordered; expand it to iop=1 first",
                      modes[i].lab[0], modes[i].lab[1]);
  // batch mode
  int bm = -1;
  if (nbatch > 1) {
    Mode m;
    m.lab[0] = '#';
    m.lab[1] = 0;
    m.batch = true;
    m.size[0] = nbatch;
    modes.push_back(m);
    bm = (int)modes.size() - 1;
    int nb = 0;
    for (int t = 0; t < 2; t++) {
      const Operand& o = t == 0 ? o1 : o2;
      if (o.batched) {
        T[t].batch_mode = bm;
        T[t].bstride = o.bstride;
        nb++;
      }
    }
    if (nb != 1) return fail(DPD_EINVAL, "exactly one input operand must be batched");
    T[2].batch_mode = bm;
    T[2].bstride = o3.bstride;
  }
  const int nm = (int)modes.size();
  if (nm > kMaxMode) return fail(DPD_EINVAL, "too many indices in contraction");
  for (int i = 0; i < nm; i++) {
    bool in[3] = {false, false, false};
    for (int t = 0; t < 3; t++) {
      for (int d = 0; d < 4; d++)
        if (T[t].mode_of_dim[d] == i) in[t] = true;
      if (T[t].batch_mode == i) in[t] = true;
    }
    Mode& m = modes[i];
    if (!m.batch && m.total() == 1 && m.size[0] == 1) m.cls = 3;
    else if (in[0] && in[2] && !in[1]) m.cls = 0;
    else if (in[1] && in[2] && !in[0]) m.cls = 1;
    else if (in[0] && in[1] && !in[2]) m.cls = 2;
    else if (in[0] && in[1] && in[2])
      return fail(DPD_EINVAL, "index '%c' appears in all three tensors", m.lab[0]);
    else
      return fail(DPD_EINVAL, "index '%c' appears in one tensor only", m.lab[0]);
  }
  for (int t = 0; t < 3; t++) T[t].L = make_layout(S, T[t].d);

  // digit order: M and N in order of appearance in v3 (coalesced C), batch last
  std::vector<int> idM, idN, idK, idDum;
  for (int d = 0; d < 4; d++) {
    int id = T[2].mode_of_dim[d];
    if (modes[id].cls == 0 && std::find(idM.begin(), idM.end(), id) == idM.end()) idM.push_back(id);
    if (modes[id].cls == 1 && std::find(idN.begin(), idN.end(), id) == idN.end()) idN.push_back(id);
  }
  for (int d = 0; d < 4; d++) {
    int id = T[0].mode_of_dim[d];
    if (modes[id].cls == 2 && std::find(idK.begin(), idK.end(), id) == idK.end()) idK.push_back(id);
  }
  for (int i = 0; i < nm; i++) {
    if (modes[i].cls == 3) idDum.push_back(i);
    if (modes[i].batch) (modes[i].cls == 0 ? idM : idN).push_back(i);
  }
  if (idM.size() > 3 || idN.size() > 3 || idK.size() > 3)
    return fail(DPD_EINVAL, "contraction needs more than three digits per GEMM dimension");

  // enumerate
  std::vector<int> ext;
  for (int id : idM)
    if (!modes[id].batch) ext.push_back(id);
  for (int id : idN)
    if (!modes[id].batch) ext.push_back(id);
  const int nir = S.nirrep;
  int64_t ecount = 1, kcount = 1;
  for (size_t i = 0; i < ext.size(); i++) ecount *= nir;
  for (size_t i = 0; i < idK.size(); i++) kcount *= nir;
  int mirr[kMaxMode];
  for (int i = 0; i < nm; i++) mirr[i] = 0;

  auto size_of = [&](int id) -> int64_t { return modes[id].batch ? nbatch : modes[id].size[mirr[id]]; };

  plan.tasks.clear();
  plan.segs.clear();
  plan.split.clear();
  plan.flops = 0;
  plan.c_elems = 0;
  for (int64_t e = 0; e < ecount; e++) {
    int64_t r = e;
    bool skip = false;
    for (int id : ext) {
      mirr[id] = (int)(r % nir);
      r /= nir;
      if (modes[id].size[mirr[id]] == 0) skip = true;
    }
    if (skip) continue;
    for (int id : idK) mirr[id] = 0;
    if (!tensor_sym_ok(S, T[2], mirr)) continue;
    Affine a3 = tensor_affine(S, T[2], mirr, nm);
    if (!a3.valid) continue;
    GemmTask tk;
    std::memset(&tk, 0, sizeof tk);
    int64_t M = 1, N = 1;
    for (size_t i = 0; i < 3; i++) {
      int64_t szm = i < idM.size() ? size_of(idM[i]) : 1;
      int64_t szn = i < idN.size() ? size_of(idN[i]) : 1;
      if (i < 2) {
        tk.md[i] = (int32_t)szm;
        tk.nd[i] = (int32_t)szn;
      }
      tk.sc_m[i] = i < idM.size() ? a3.stride[idM[i]] : 0;
      tk.sc_n[i] = i < idN.size() ? a3.stride[idN[i]] : 0;
      M *= szm;
      N *= szn;
    }
    if (M == 0 || N == 0) continue;
    if (M > INT32_MAX || N > INT32_MAX) return fail(DPD_EINVAL, "GEMM block too large");
    plan.c_elems += M * N;
    tk.M = (int32_t)M;
    tk.N = (int32_t)N;
    tk.c_off = a3.base;
    tk.seg_begin = (int32_t)plan.segs.size();
    int ktile = 0;
    for (int64_t kk = 0; kk < kcount; kk++) {
      int64_t rr = kk;
      bool sk = false;
      for (int id : idK) {
        mirr[id] = (int)(rr % nir);
        rr /= nir;
        if (modes[id].size[mirr[id]] == 0) sk = true;
      }
      if (sk) continue;
      if (!tensor_sym_ok(S, T[0], mirr) || !tensor_sym_ok(S, T[1], mirr)) continue;
      Affine a1 = tensor_affine(S, T[0], mirr, nm), a2 = tensor_affine(S, T[1], mirr, nm);
      if (!a1.valid || !a2.valid) continue;
      GemmSeg sg;
      std::memset(&sg, 0, sizeof sg);
      int64_t K = 1;
      for (size_t i = 0; i < 3; i++) {
        int64_t szk = i < idK.size() ? modes[idK[i]].size[mirr[idK[i]]] : 1;
        if (i < 2) sg.kd[i] = (int32_t)szk;
        sg.sa_k[i] = i < idK.size() ? a1.stride[idK[i]] : 0;
        sg.sb_k[i] = i < idK.size() ? a2.stride[idK[i]] : 0;
        sg.sa_m[i] = i < idM.size() ? a1.stride[idM[i]] : 0;
        sg.sb_n[i] = i < idN.size() ? a2.stride[idN[i]] : 0;
        K *= szk;
      }
      if (K == 0) continue;
      if (K > INT32_MAX) return fail(DPD_EINVAL, "GEMM reduction too large");
      sg.K = (int32_t)K;
      sg.a_off = a1.base;
      sg.b_off = a2.base;
      sg.ktiles = (int32_t)((K + 15) / 16);
      sg.ktile_begin = ktile;
      ktile += sg.ktiles;
      plan.segs.push_back(sg);
      plan.flops += 2 * M * N * K;
    }
    tk.nseg = (int32_t)plan.segs.size() - tk.seg_begin;
    tk.ktiles_total = ktile;
    tk.ksplit = 1;
    plan.tasks.push_back(tk);
  }
  // tasks without any K segment still apply beta (C = f2*C): give them one empty
  // segment so the kernel writes alpha*0 + beta*C.
  for (auto& tk : plan.tasks) {
    if (tk.nseg == 0) {
      GemmSeg sg;
      std::memset(&sg, 0, sizeof sg);
      sg.kd[0] = sg.kd[1] = 1;
      sg.K = 0;
      sg.ktiles = 0;
      tk.seg_begin = (int32_t)plan.segs.size();
      tk.nseg = 1;
      plan.segs.push_back(sg);
    }
  }
  // tile shape: the largest tiles whose padding waste stays small; the
  // pipelined kernel handles every task shape, the first-generation kernel is
  // kept selectable (dpd_ctx_set_gemm_config) for comparison
  plan.cfg = choose_gemm_cfg(plan.tasks, plan.cfg_req);
  // the gathering kernel has no 192-wide tile
  if (plan.cfg == GEMM_64x192) plan.cfg = GEMM_64x128;
  if (plan.cfg == GEMM_128x64D) plan.cfg = GEMM_128x64;
  const int BM = gemm_bm(plan.cfg), BN = gemm_bn(plan.cfg);
  for (auto& tk : plan.tasks) {
    tk.tiles_m = (tk.M + BM - 1) / BM;
    tk.tiles_n = (tk.N + BN - 1) / BN;
  }
  // split-K: balance K tiles so the launch fills the chip
  int64_t work = 0, tiles = 0;
  for (auto& tk : plan.tasks) {
    work += (int64_t)tk.tiles_m * tk.tiles_n * std::max(tk.ktiles_total, 1);
    tiles += (int64_t)tk.tiles_m * tk.tiles_n;
  }
  const int64_t target = 1024;
  int64_t per_wg = std::max<int64_t>(8, (work + target - 1) / target);
  int64_t partial = 0;
  int max_mn = 0;
  int wg = 0;
  for (size_t i = 0; i < plan.tasks.size(); i++) {
    GemmTask& tk = plan.tasks[i];
    int ks = 1;
    if (tiles < target && tk.ktiles_total > per_wg) {
      ks = (int)std::min<int64_t>(64, (tk.ktiles_total + per_wg - 1) / per_wg);
      if (ks < 1) ks = 1;
    }
    tk.ksplit = ks;
    if (ks > 1) {
      tk.part_off = partial;
      partial += (int64_t)ks * tk.M * tk.N;
      plan.split.push_back((int32_t)i);
      max_mn = std::max<int64_t>(max_mn, (int64_t)tk.M * tk.N);
    }
    tk.wg_begin = wg;
    wg += tk.tiles_m * tk.tiles_n * ks;
  }
  plan.nwg = wg;
  plan.partial_elems = partial;
  plan.max_split_mn = max_mn;
  (void)idDum;
  // host-side bounds proof: every address a task or segment can form lies
  // inside its tensor (all strides are non-negative)
  int64_t lim[3];
  for (int t = 0; t < 3; t++) lim[t] = T[t].L.len + (T[t].batch_mode >= 0 ? (int64_t)(nbatch - 1) * T[t].bstride : 0);
  auto span = [](int64_t n, const int32_t* d, const int64_t* s) {
    int64_t d2 = n / ((int64_t)d[0] * d[1]);
    return (d[0] - 1) * s[0] + (d[1] - 1) * s[1] + (d2 - 1) * s[2];
  };
  for (const auto& tk : plan.tasks) {
    if ((int64_t)tk.md[0] * tk.md[1] * (tk.M / ((int64_t)tk.md[0] * tk.md[1])) != tk.M ||
        (int64_t)tk.nd[0] * tk.nd[1] * (tk.N / ((int64_t)tk.nd[0] * tk.nd[1])) != tk.N)
      return fail(DPD_EINVAL, "internal: GEMM digit sizes do not factor M/N");
    if (tk.c_off < 0 || tk.c_off + span(tk.M, tk.md, tk.sc_m) + span(tk.N, tk.nd, tk.sc_n) >= lim[2])
      return fail(DPD_EINVAL, "internal: output block out of bounds");
    for (int si = tk.seg_begin; si < tk.seg_begin + tk.nseg; si++) {
      const GemmSeg& sg = plan.segs[si];
      if (sg.K == 0) continue;
      if ((int64_t)sg.kd[0] * sg.kd[1] * (sg.K / ((int64_t)sg.kd[0] * sg.kd[1])) != sg.K)
        return fail(DPD_EINVAL, "internal: GEMM digit sizes do not factor K");
      if (sg.a_off < 0 || sg.a_off + span(tk.M, tk.md, sg.sa_m) + span(sg.K, sg.kd, sg.sa_k) >= lim[0])
        return fail(DPD_EINVAL, "internal: A segment out of bounds");
      if (sg.b_off < 0 || sg.b_off + span(sg.K, sg.kd, sg.sb_k) + span(tk.N, tk.nd, sg.sb_n) >= lim[1])
        return fail(DPD_EINVAL, "internal: B segment out of bounds");
    }
  }
  return plan.upload();
}

int choose_gemm_cfg(const std::vector<GemmTask>& tasks, int requested) {
  if (requested >= GEMM_V1 && requested <= kGemmCfgMax) return requested;
  // cost ~ padded MFMA work / relative tile efficiency, plus a penalty when the
  // tile count cannot fill 256 CUs (split-K recovers part of that)
  static const int cand[5] = {GEMM_128x128, GEMM_128x64, GEMM_64x128, GEMM_64x64, GEMM_32x32};
  static const double eff[5] = {1.0, 0.85, 0.85, 0.7, 0.4};
  int best = GEMM_64x64;
  double best_cost = 0.0;
  for (int ci = 0; ci < 5; ci++) {
    const int bm = gemm_bm(cand[ci]), bn = gemm_bn(cand[ci]);
    double padded = 0.0, ntile = 0.0;
    for (const auto& tk : tasks) {
      const double tmn = (double)((tk.M + bm - 1) / bm) * ((tk.N + bn - 1) / bn);
      padded += tmn * bm * bn * std::max(tk.ktiles_total, 1);
      ntile += tmn;
    }
    double cost = padded / eff[ci];
    if (ntile < 512) cost *= 1.0 + 0.25 * (512 - ntile) / 512;
    if (ci == 0 || cost < best_cost) {
      best = cand[ci];
      best_cost = cost;
    }
  }
  return best;
}

// split_k: split long-K tasks over several workgroups while the launch has
// fewer than 1024 tiles (partials summed by a reduce kernel); off: one
// workgroup per tile
int layout_variant(const GemmPlan& p, GemmVariant& V, int mode, int cfg, bool split_k) {
  if (cfg == GEMM_128x64D && mode != PREC_BF16X3) cfg = GEMM_128x64;  // direct-A tile: BF16x3 only
  V.mode = mode;
  V.tasks = p.tasks;
  V.split.clear();
  V.cfg = cfg;
  V.split_k = split_k;
  const int BM = gemm_bm(V.cfg), BN = gemm_bn(V.cfg), KS = pk_kstep(mode), PL = pk_planes(mode);
  for (auto& tk : V.tasks) {
    tk.steps = (int32_t)(((int64_t)tk.ktiles_total * 16 + KS - 1) / KS);
    tk.tiles_m = (tk.M + BM - 1) / BM;
    tk.tiles_n = (tk.N + BN - 1) / BN;
  }
  // split-K over K steps so that the launch fills the chip
  int64_t work = 0, tiles = 0;
  for (auto& tk : V.tasks) {
    work += (int64_t)tk.tiles_m * tk.tiles_n * std::max(tk.steps, 1);
    tiles += (int64_t)tk.tiles_m * tk.tiles_n;
  }
  const int64_t target = 1024;
  const int64_t min_steps = mode == PREC_F64 ? 8 : 4;
  int64_t per_wg = std::max<int64_t>(min_steps, (work + target - 1) / target);
  int64_t partial = 0, pa = 0, pb = 0;
  int max_mn = 0, wg = 0, pan_a = 0, pan_b = 0;
  for (size_t i = 0; i < V.tasks.size(); i++) {
    GemmTask& tk = V.tasks[i];
    int ks = 1;
    if (split_k && tiles < target && tk.steps > per_wg)
      ks = (int)std::min<int64_t>(64, (tk.steps + per_wg - 1) / per_wg);
    tk.ksplit = ks;
    tk.part_off = 0;
    if (ks > 1) {
      tk.part_off = partial;
      partial += (int64_t)ks * tk.M * tk.N;
      V.split.push_back((int32_t)i);
      max_mn = std::max<int64_t>(max_mn, (int64_t)tk.M * tk.N);
    }
    tk.wg_begin = wg;
    wg += tk.tiles_m * tk.tiles_n * ks;
    tk.pan_a = pan_a;
    tk.pan_b = pan_b;
    tk.pa_off = pa;
    tk.pb_off = pb;
    pan_a += tk.tiles_m * tk.steps;
    pan_b += tk.tiles_n * tk.steps;
    pa += (int64_t)tk.tiles_m * tk.steps * BM * KS * PL;
    pb += (int64_t)tk.tiles_n * tk.steps * BN * KS * PL;
    if (wg < 0 || pan_a < 0 || pan_b < 0) return fail(DPD_EINVAL, "GEMM plan too large");
  }
  V.nwg = wg;
  V.npan_a = pan_a;
  V.npan_b = pan_b;
  V.pa_elems = pa;
  V.pb_elems = pb;
  V.partial_elems = partial;
  V.max_split_mn = max_mn;
  return DPD_OK;
}

static int build_variant_cfg(GemmPlan& p, GemmVariant& V, int mode, int cfg, bool split_k = true) {
  if (g_capturing) return fail(DPD_EINVAL, "tile variant built during graph capture (internal)");
  bump_mem_gen();  // the variant's task tables are rewritten and reallocated
  for (int sd = 0; sd < 2; sd++) {  // panel codes follow the tiling
    V.pcode[sd].reset();
    V.pcode_state[sd] = 0;
  }
  int rc = layout_variant(p, V, mode, cfg, split_k);
  if (rc) return rc;
  size_t bt = V.tasks.size() * sizeof(GemmTask), bp = V.split.size() * sizeof(int32_t);
  rc = V.dev.alloc(((bt + 255) & ~size_t(255)) + bp + 256);
  if (rc) return rc;
  V.d_tasks = (GemmTask*)V.dev.p;
  V.d_split = (int32_t*)((char*)V.dev.p + ((bt + 255) & ~size_t(255)));
  if (bt) DPD_HIP_TRY(hipMemcpy(V.d_tasks, V.tasks.data(), bt, hipMemcpyHostToDevice));
  if (bp) DPD_HIP_TRY(hipMemcpy(V.d_split, V.split.data(), bp, hipMemcpyHostToDevice));
  V.built = true;
  return DPD_OK;
}

int build_variant(GemmPlan& p, int mode) {
  GemmVariant& V = p.pk[mode];
  if (V.built) return DPD_OK;
  int cfg = choose_gemm_cfg(p.tasks, p.cfg_req >= GEMM_64x64 ? p.cfg_req : -1);
  if (cfg == GEMM_V1) cfg = GEMM_64x64;
  return build_variant_cfg(p, V, mode, cfg, p.split_req != 0);
}

// one past the largest element of C a plan can write
static int64_t c_span(const GemmPlan& p) {
  auto span = [](int64_t n, const int32_t* d, const int64_t* s) {
    int64_t d2 = n / ((int64_t)d[0] * d[1]);
    return (d[0] - 1) * s[0] + (d[1] - 1) * s[1] + (d2 - 1) * s[2];
  };
  int64_t m = 0;
  for (const auto& tk : p.tasks) m = std::max(m, tk.c_off + span(tk.M, tk.md, tk.sc_m) + span(tk.N, tk.nd, tk.sc_n) + 1);
  return m;
}

// pack one operand side of a variant into dst
int pack_side(dpd_ctx* c, GemmPlan& p, GemmVariant& V, int side, const PackSrc& src, void* dst) {
  const int tile = side == 0 ? gemm_bm(V.cfg) : gemm_bn(V.cfg);
  const int64_t npanel = side == 0 ? V.npan_a : V.npan_b;
  const int64_t pan = (int64_t)tile * pk_kstep(V.mode);
  // per-call operand of a cached plan: build the source code of every panel
  // element once (the pack kernel's own address logic in code mode), then pack
  // with one flat pass; not while capturing, and only for addresses below 2^31
  if (src.coded && V.pcode_state[side] == 1 && V.pcode_comp[side] != (src.gcode != nullptr) && !g_capturing) {
    V.pcode[side].reset();  // built for the other form of the operand
    V.pcode_state[side] = 0;
    bump_mem_gen();
  }
  if (src.coded && !src.map && !src.id && !g_capturing && V.pcode_state[side] == 0 &&
      src.src_elems < (int64_t)INT32_MAX - 1 && (!src.gcode || src.gxlen > 0)) {
    V.pcode_state[side] = -1;
    if (!V.pcode[side].alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(npanel * pan, 1))) {
      PackArgs a;
      std::memset(&a, 0, sizeof a);
      a.src = src.data;
      if (!V.tasks.empty()) {
        const GemmSeg& s0 = p.segs[V.tasks[0].seg_begin];
        a.rfast = side == 0 ? (s0.sa_m[0] == 1) : (s0.sb_n[0] == 1);
      }
      a.dst = nullptr;
      a.tasks = V.d_tasks;
      a.segs = p.d_segs;
      a.ntask = (int)V.tasks.size();
      a.npanel = (int)npanel;
      a.side = side;
      a.mode = V.mode;
      a.tile = tile;
      a.code_out = (int32_t*)V.pcode[side].p;
      a.gcode = src.gcode;
      a.gxlen = src.gxlen;
      a.gbstride = src.gbstride;
      if (launch_pack(a, c->stream) == hipSuccess) {
        V.pcode_state[side] = 1;
        V.pcode_comp[side] = src.gcode != nullptr;
      }
    } else {
      (void)hipGetLastError();
    }
  }
  if (src.coded && !src.map && !src.id && V.pcode_state[side] == 1 && V.pcode_comp[side] == (src.gcode != nullptr)) {
    DPD_HIP_TRY(launch_pack_code(V.mode, src.f32 ? 1 : 0, npanel, pan, (const int32_t*)V.pcode[side].p, src.data,
                                 dst, c->stream));
    return DPD_OK;
  }
  if (src.gcode) return fail(DPD_EINVAL, "composed pack without its code (internal)");
  PackArgs a;
  std::memset(&a, 0, sizeof a);
  a.src = src.data;
  // read order: the source's unit stride decides (rows or k's fastest)
  if (!V.tasks.empty()) {
    const GemmSeg& s0 = p.segs[V.tasks[0].seg_begin];
    a.rfast = side == 0 ? (s0.sa_m[0] == 1) : (s0.sb_n[0] == 1);
  }
  if (src.map) {
    a.use_map = 1;
    a.rowmask = src.rowmask;
    a.map = *src.map;
    // the unsorted tensor's dim 0 is its fastest: is it a row or a k index?
    for (int d = 0; d < 4; d++)
      if (src.map->term[0].perm[d] == 0) a.rfast = (src.rowmask >> d) & 1;
    a.xlen = src.xlen;
    a.src_bstride = src.bstride;
  }
  a.src_f32 = src.f32 && !src.map ? 1 : 0;
  a.dst = dst;
  a.tasks = V.d_tasks;
  a.segs = p.d_segs;
  a.ntask = (int)V.tasks.size();
  a.npanel = side == 0 ? V.npan_a : V.npan_b;
  a.side = side;
  a.mode = V.mode;
  a.tile = tile;
  DPD_HIP_TRY(launch_pack(a, c->stream));
  return DPD_OK;
}

// panels of both operand sides of variant V: static operands from V's cache
// (packed on first use; one entry per operand, found by data pointer and id),
// per-call operands packed into the workspace reserved by the caller
static int pack_sides(dpd_ctx* c, GemmPlan& p, GemmVariant& V, int mode, const PackSrc& A, const PackSrc& B,
                      const void* side_ptr[2]) {
  const size_t es = pk_esize(mode);
  const PackSrc* src[2] = {&A, &B};
  const int64_t elems[2] = {V.pa_elems, V.pb_elems};
  for (int sd = 0; sd < 2; sd++) {
    if (elems[sd] == 0) {
      side_ptr[sd] = nullptr;
      continue;
    }
    if (src[sd]->id) {
      PackCache* hit = nullptr;
      PackCache* same_src = nullptr;
      for (auto& e : V.cache[sd]) {
        if (e->src == src[sd]->data && e->id == src[sd]->id) hit = e.get();
        else if (e->src == src[sd]->data) same_src = e.get();
      }
      if (!hit) {
        if (g_capturing) return fail(DPD_EINVAL, "operand panels packed during graph capture (internal)");
        if (same_src) {
          // the operand's contents changed (new id at the same address): repack in place
          hit = same_src;
          bump_mem_gen();
        } else {
          V.cache[sd].push_back(std::make_unique<PackCache>());
          hit = V.cache[sd].back().get();
          int rc = hit->buf.alloc(es * elems[sd]);
          if (rc) {
            V.cache[sd].pop_back();
            return rc;
          }
        }
        hit->src = nullptr;
        hit->id = 0;
        int rc = pack_side(c, p, V, sd, *src[sd], hit->buf.p);
        if (rc) return rc;
        hit->src = src[sd]->data;
        hit->id = src[sd]->id;
      }
      side_ptr[sd] = hit->buf.p;
    } else {
      void* w = ws_take(c, es * elems[sd]);
      if (!w) return DPD_ENOMEM;
      int rc = pack_side(c, p, V, sd, *src[sd], w);
      if (rc) return rc;
      side_ptr[sd] = w;
    }
  }
  return DPD_OK;
}

int run_gemm_packed(dpd_ctx* c, GemmPlan& p, int mode, const PackSrc& A, const PackSrc& B, double* C, double f1,
                    double f2, hipEvent_t ev0, hipEvent_t ev1, bool c_f32, hipEvent_t ev_pack) {
  GemmVariant& V = p.pk[mode];
  const void* side_ptr[2];
  int rc0 = pack_sides(c, p, V, mode, A, B, side_ptr);
  if (rc0) return rc0;
  if (ev_pack) DPD_HIP_TRY(hipEventRecord(ev_pack, c->stream));
  if (V.nwg == 0) {
    if (ev0) DPD_HIP_TRY(hipEventRecord(ev0, c->stream));
    if (ev1) DPD_HIP_TRY(hipEventRecord(ev1, c->stream));
    return DPD_OK;
  }
  GemmLaunch g;
  std::memset(&g, 0, sizeof g);
  g.C = C;
  g.tasks = V.d_tasks;
  g.segs = p.d_segs;
  g.ntask = (int)V.tasks.size();
  g.nwg = V.nwg;
  g.alpha = f1;
  g.beta = f2;
  g.prec = mode;
  g.split_tasks = V.d_split;
  g.nsplit = (int)V.split.size();
  g.max_split_mn = V.max_split_mn;
  g.cfg = V.cfg;
  g.c_f32 = c_f32 ? 1 : 0;
  if (V.partial_elems > 0 && !(g.partial = (double*)ws_take(c, sizeof(double) * V.partial_elems))) return DPD_ENOMEM;
  // the events bracket the GEMM kernel alone (the split-K reduce follows)
  if (ev0) DPD_HIP_TRY(hipEventRecord(ev0, c->stream));
  DPD_HIP_TRY(launch_gemm_pk_only(g, side_ptr[0], side_ptr[1], c->stream));
  if (ev1) DPD_HIP_TRY(hipEventRecord(ev1, c->stream));
  DPD_HIP_TRY(launch_splitk_reduce(g, c->stream));
  return DPD_OK;
}

// Autotuner: times every tile shape of the packed-panel kernel on the call's
// own operands (packed into temporary buffers, results into a scratch C) and
// keeps the fastest for this plan and precision.  Runs once per plan.
int tune_variant(dpd_ctx* c, GemmPlan& p, int mode, const PackSrc& A, const PackSrc& B, double f1, double f2) {
  // candidates in order of preference: among shapes within 5% of the fastest
  // the first listed wins, which keeps the choice stable from run to run; each
  // tile shape with and without split-K (the panels are the same for both)
  static const int cand[] = {GEMM_128x64, GEMM_64x192, GEMM_128x64D, GEMM_128x128, GEMM_64x128, GEMM_64x64,
                             GEMM_32x32};
  constexpr int NC = (int)(sizeof cand / sizeof cand[0]);
  float tms[NC][2];
  const size_t es = pk_esize(mode);
  DevBuf cs;
  int rc = cs.alloc(sizeof(double) * std::max<int64_t>(c_span(p), 1));
  if (rc) return rc;
  hipEvent_t e0, e1;
  DPD_HIP_TRY(hipEventCreate(&e0));
  DPD_HIP_TRY(hipEventCreate(&e1));
  float best_ms = -1.0f;
  for (int ci = 0; ci < NC && !rc; ci++) {
    tms[ci][0] = tms[ci][1] = 1e30f;
    const int cfg = cand[ci];
    if (cfg == GEMM_128x64D && mode != PREC_BF16X3) continue;  // the same as GEMM_128x64 there
    DevBuf ta, tb;
    bool had_split = false;
    for (int sk = 0; sk < 2 && !rc; sk++) {
      if ((sk == 0 && p.split_req == 0) || (sk == 1 && p.split_req == 1)) continue;  // requested split mode
      GemmVariant T;
      if ((rc = build_variant_cfg(p, T, mode, cfg, sk == 0))) break;
      if (sk == 0) had_split = !T.split.empty();
      if (sk == 1 && !had_split && p.split_req != 0) {
        tms[ci][1] = tms[ci][0];  // nothing was split: the same variant
        continue;
      }
      DevBuf tp;
      // the panels do not depend on the split mode: packed by whichever pass runs first
      if ((T.pa_elems && !ta.p && ((rc = ta.alloc(es * T.pa_elems)) || (rc = pack_side(c, p, T, 0, A, ta.p)))) ||
          (T.pb_elems && !tb.p && ((rc = tb.alloc(es * T.pb_elems)) || (rc = pack_side(c, p, T, 1, B, tb.p)))))
        break;
      if ((T.pa_elems && !ta.p) || (T.pb_elems && !tb.p)) {
        rc = fail(DPD_EINVAL, "autotuner: operand panel missing (internal)");
        break;
      }
      if (T.partial_elems && (rc = tp.alloc(sizeof(double) * T.partial_elems))) break;
      GemmLaunch g;
      std::memset(&g, 0, sizeof g);
      g.C = (double*)cs.p;
      g.tasks = T.d_tasks;
      g.segs = p.d_segs;
      g.ntask = (int)T.tasks.size();
      g.nwg = T.nwg;
      g.alpha = f1;
      g.beta = f2;
      g.prec = mode;
      g.split_tasks = T.d_split;
      g.nsplit = (int)T.split.size();
      g.max_split_mn = T.max_split_mn;
      g.cfg = cfg;
      g.partial = (double*)tp.p;
      DPD_HIP_TRY(launch_gemm_pk(g, ta.p, tb.p, c->stream));
      DPD_HIP_TRY(hipEventRecord(e0, c->stream));
      for (int r = 0; r < 3; r++) DPD_HIP_TRY(launch_gemm_pk(g, ta.p, tb.p, c->stream));
      DPD_HIP_TRY(hipEventRecord(e1, c->stream));
      DPD_HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.0f;
      DPD_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      tms[ci][sk] = ms;
      if (best_ms < 0.0f || ms < best_ms) best_ms = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  int best = GEMM_64x64;
  bool best_split = true;
  if (best_ms >= 0.0f) {
    bool found = false;
    for (int ci = 0; ci < NC && !found; ci++)
      for (int sk = 0; sk < 2 && !found; sk++)
        if (tms[ci][sk] <= 1.05f * best_ms) {
          best = cand[ci];
          best_split = sk == 0;
          found = true;
        }
  }
  GemmVariant& V = p.pk[mode];
  if (V.built && V.cfg == best && V.split_k == best_split) return DPD_OK;
  // panels cached for the old tile shape are rebuilt on the next call
  if (!(V.built && V.cfg == best)) clear_pack_cache(V);
  return build_variant_cfg(p, V, mode, best, best_split);
}

int run_gemm(dpd_ctx* c, GemmPlan& p, const void* A, int a_f32, const void* B, int b_f32, double* C,
             double f1, double f2, int prec, bool c_f32) {
  if (p.nwg == 0) return DPD_OK;
  GemmLaunch g;
  std::memset(&g, 0, sizeof g);
  g.A = A;
  g.B = B;
  g.C = C;
  g.tasks = p.d_tasks;
  g.segs = p.d_segs;
  g.ntask = (int)p.tasks.size();
  g.nwg = p.nwg;
  g.alpha = f1;
  g.beta = f2;
  g.prec = prec;
  g.a_f32 = a_f32;
  g.b_f32 = b_f32;
  g.c_f32 = c_f32 ? 1 : 0;
  g.split_tasks = p.d_split;
  g.nsplit = (int)p.split.size();
  g.max_split_mn = p.max_split_mn;
  g.cfg = p.cfg;
  if (p.partial_elems > 0 && !(g.partial = (double*)ws_take(c, sizeof(double) * p.partial_elems))) return DPD_ENOMEM;
  DPD_HIP_TRY(launch_gemm(g, c->stream));
  return DPD_OK;
}

// the tiled permutation plan of a plain copy between full-pair layouts, built
// and uploaded on first use (outside graph capture); null when the generic
// gather kernels apply
constexpr size_t kSortCacheEntries = 512;
static const SortPlan* sort_plan(dpd_ctx* c, const MapArgs& a) {
  if (!sort_tiles_apply(a)) return nullptr;
  std::string key((const char*)&a.out, sizeof(DevLayout));
  key.append((const char*)&a.in, sizeof(DevLayout));
  key.append((const char*)a.term[0].perm, 4);
  auto it = c->sort_cache.find(key);
  if (it != c->sort_cache.end()) return it->second->ok ? it->second.get() : nullptr;
  if (g_capturing) return nullptr;
  if (c->sort_cache.size() >= kSortCacheEntries) {
    if (hipStreamSynchronize(c->stream) != hipSuccess) return nullptr;
    bump_mem_gen();  // captured graphs may hold the task tables
    c->sort_cache.clear();
  }
  auto P = std::make_unique<SortPlan>();
  if (plan_sort(a, P->tasks, &P->nitems, &P->direct) && !P->tasks.empty() && P->nitems > 0) {
    // the task table, then the task of every work item (one int32 each)
    const size_t nt = P->tasks.size(), tb = sizeof(SortTask) * nt;
    P->host.resize(tb + sizeof(int32_t) * (size_t)P->nitems);
    std::memcpy(P->host.data(), P->tasks.data(), tb);
    sort_item_map(P->tasks, P->nitems, (int32_t*)(P->host.data() + tb));
    if (P->dev.alloc(P->host.size()) == DPD_OK &&
        hipMemcpyAsync(P->dev.p, P->host.data(), P->host.size(), hipMemcpyHostToDevice, c->stream) == hipSuccess)
      P->ok = true;
  }
  (void)hipGetLastError();
  const SortPlan* r = P->ok ? P.get() : nullptr;
  c->sort_cache.emplace(std::move(key), std::move(P));
  return r;
}

static const int32_t* sort_item_task(const SortPlan* P) {
  return (const int32_t*)((const char*)P->dev.p + sizeof(SortTask) * P->tasks.size());
}

int gather(dpd_ctx* c, MapArgs a, const double* in, double* out) {
  if (const SortPlan* P = sort_plan(c, a)) {
    DPD_HIP_TRY(launch_sort_tiles((const SortTask*)P->dev.p, sort_item_task(P), P->nitems, P->direct, in, out,
                                  c->stream));
    return DPD_OK;
  }
  DPD_HIP_TRY(launch_gather_map(a, in, out, c->stream));
  return DPD_OK;
}

int gather_f32(dpd_ctx* c, MapArgs a, const float* in, float* out) {
  if (const SortPlan* P = sort_plan(c, a)) {
    DPD_HIP_TRY(launch_sort_tiles_f32((const SortTask*)P->dev.p, sort_item_task(P), P->nitems, P->direct, in,
                                      out, c->stream));
    return DPD_OK;
  }
  DPD_HIP_TRY(launch_gather_map_f32(a, in, out, c->stream));
  return DPD_OK;
}

// ---------------------------------------------------------------------------
// engine level: libsrint/vv2v_dpd_{444,424,442}.f90 dispatcher checks
// ---------------------------------------------------------------------------
static bool pops_eq(const int64_t* a, const int64_t* b) {
  for (int i = 0; i < 8; i++)
    if (a[i] != b[i]) return false;
  return true;
}

static int check_engine(const Sym& S, int kind, const Operand& o1, const Operand& o2,
                        const Operand& o3, Operand& m1, Operand& m2, Operand& m3) {
  m1 = o1;
  m2 = o2;
  m3 = o3;
  if (kind == 444) {
    int rc = map444(o1.lab, o2.lab, o3.lab, m1.lab, m2.lab, m3.lab);
    if (rc) return rc;
    // contracted populations (vv2v_dpd_444.f90:152-174)
    for (char cc : {'3', '4'}) {
      int p1 = -1, p2 = -1;
      for (int i = 0; i < 4; i++) {
        if (m1.lab[i] == cc) p1 = i;
        if (m2.lab[i] == cc) p2 = i;
      }
      if (p1 >= 0 && p2 >= 0 && !pops_eq(o1.d.pop[p1], o2.d.pop[p2]))
        return fail(DPD_EPOP, "FATAL ERROR in vv2v_dpd_444: contracted index populations differ (%.4s * %.4s -> %.4s)",
                    o1.lab, o2.lab, o3.lab);
    }
    // fast-path detection (vv2v_dpd_444.f90:177-187) and the IOP wall (:189-201)
    const char* t1 = m1.lab;
    const char* t2 = m2.lab;
    const char* t3 = m3.lab;
    const int* i1 = o1.d.iop;
    const int* i2 = o2.d.iop;
    const int* i3 = o3.d.iop;
    bool fast = false;
    if (!std::memcmp(t3, "1256", 4)) {
      if (!std::memcmp(t1, "3412", 4) && !std::memcmp(t2, "3456", 4))
        fast = i1[0] == i2[0] && i1[1] == i3[0] && i2[1] == i3[1];
      else if (!std::memcmp(t1, "1234", 4) && !std::memcmp(t2, "3456", 4))
        fast = i1[1] == i2[0] && i1[0] == i3[0] && i2[1] == i3[1];
      else if (!std::memcmp(t1, "1234", 4) && !std::memcmp(t2, "5634", 4))
        fast = i1[1] == i2[1] && i1[0] == i3[0] && i2[0] == i3[1];
      else if (!std::memcmp(t1, "3412", 4) && !std::memcmp(t2, "5634", 4))
        fast = i1[0] == i2[1] && i1[1] == i3[0] && i2[0] == i3[1];
    }
    if (!fast && (i1[0] != 1 || i1[1] != 1 || i2[0] != 1 || i2[1] != 1 || i3[0] != 1 || i3[1] != 1))
      return fail(DPD_EIOP, "FATAL ERROR in vv2v_dpd_444: the fallback path forbids compressed IOP (%.4s * %.4s -> %.4s)",
                  o1.lab, o2.lab, o3.lab);
    return DPD_OK;
  }
  if (S.D(o1.d.sym, o2.d.sym) != o3.d.sym)
    return fail(DPD_ESYM, "FATAL ERROR in vv2v_dpd_%d: Sym1*Sym2 != Sym3", kind);
  if (kind == 424) {
    int ncon = 0;
    char cc = 0;
    for (int i = 0; i < 4; i++) {
      bool in3 = false;
      for (int j = 0; j < 4; j++)
        if (o1.lab[i] == o3.lab[j]) in3 = true;
      if (!in3) {
        ncon++;
        cc = o1.lab[i];
      }
    }
    if (ncon != 1) return fail(DPD_EINVAL, "FATAL ERROR: vv2v_dpd_424 requires EXACTLY 1 contracted index.");
    if (o2.lab[0] != cc && o2.lab[1] != cc)
      return fail(DPD_EINVAL, "FATAL ERROR: contracted index '%c' missing in V2.", cc);
    // IOP barrier (vv2v_dpd_424.f90:164-196)
    for (int s = 0; s < 2; s++) {
      if (o1.d.iop[s] != 1) {
        bool ok = false;
        for (int u = 0; u < 2; u++)
          if (!std::memcmp(o1.lab + 2 * s, o3.lab + 2 * u, 2) && o1.d.iop[s] == o3.d.iop[u]) ok = true;
        if (!ok) return fail(DPD_EIOP, "FATAL ERROR in vv2v_dpd_424: IOP barrier");
      }
      if (o3.d.iop[s] != 1) {
        bool ok = false;
        for (int u = 0; u < 2; u++)
          if (!std::memcmp(o3.lab + 2 * s, o1.lab + 2 * u, 2) && o3.d.iop[s] == o1.d.iop[u]) ok = true;
        if (!std::memcmp(o3.lab + 2 * s, o2.lab, 2) && o3.d.iop[s] == o2.d.iop[0]) ok = true;
        if (!ok) return fail(DPD_EIOP, "FATAL ERROR in vv2v_dpd_424: IOP barrier");
      }
    }
    if (o2.d.iop[0] != 1) {
      bool ok = false;
      for (int u = 0; u < 2; u++)
        if (!std::memcmp(o2.lab, o3.lab + 2 * u, 2) && o2.d.iop[0] == o3.d.iop[u]) ok = true;
      if (!ok) return fail(DPD_EIOP, "FATAL ERROR in vv2v_dpd_424: IOP barrier");
    }
    return DPD_OK;
  }
  // 442
  int ncon = 0;
  for (int i = 0; i < 4; i++) {
    if (o1.lab[i] == o3.lab[0] || o1.lab[i] == o3.lab[1]) continue;
    ncon++;
    bool in2 = false;
    for (int j = 0; j < 4; j++)
      if (o2.lab[j] == o1.lab[i]) in2 = true;
    if (!in2) return fail(DPD_EINVAL, "FATAL ERROR: Contracted index missing in V2.");
  }
  if (ncon != 3) return fail(DPD_EINVAL, "FATAL ERROR: vv2v_dpd_442 requires EXACTLY 3 contracted indices.");
  return DPD_OK;
}

static std::string sig_of(int kind, const Operand& o1, const Operand& o2, const Operand& o3,
                          int nbatch, int anti, int extra) {
  std::string s;
  auto put = [&](const void* p, size_t n) { s.append((const char*)p, n); };
  put(&kind, sizeof kind);
  for (const Operand* o : {&o1, &o2, &o3}) {
    put(o->d.pop, sizeof o->d.pop);
    put(o->d.iop, sizeof o->d.iop);
    put(&o->d.sym, sizeof o->d.sym);
    put(o->lab, 4);
    put(&o->rank, sizeof o->rank);
    put(&o->bstride, sizeof o->bstride);
    bool b = o->batched;
    put(&b, 1);
  }
  put(&nbatch, sizeof nbatch);
  put(&anti, sizeof anti);
  put(&extra, sizeof extra);
  return s;
}

int engine_call(dpd_ctx* c, int kind, const Operand& o1, const Operand& o2, const Operand& o3,
                double f1, double f2, int nbatch, bool f32data) {
  Operand m1, m2, m3;
  int rc = check_engine(c->sym, kind, o1, o2, o3, m1, m2, m3);
  if (rc) return rc;
  std::string key = sig_of(kind, m1, m2, m3, nbatch, -1, 0);
  auto it = c->engine_cache.find(key);
  GemmPlan* p;
  if (it == c->engine_cache.end()) {
    auto np = std::make_unique<GemmPlan>();
    np->cfg_req = c->gemm_cfg;
    rc = build_gemm_plan(c->sym, m1, m2, m3, nbatch, *np);
    if (rc) return rc;
    p = np.get();
    c->engine_cache.emplace(key, std::move(np));
  } else {
    p = it->second.get();
  }
  ws_reset(c);
  if (f32data) {
    // FP32 storage (libsrint/vv2v_dpd_444_sg.f90): the gathering kernel reads the
    // FP32 tensors where they lie and runs FP32 MFMA
    rc = ws_reserve(c, sizeof(double) * (p->partial_elems + 32));
    if (rc) return rc;
    return run_gemm(c, *p, o1.data, 1, o2.data, 1, const_cast<double*>(o3.data), f1, f2, PREC_F32, true);
  }
  if (c->packing == 2) {
    if ((rc = build_variant(*p, PREC_F64))) return rc;
    const GemmVariant& V = p->pk[PREC_F64];
    rc = ws_reserve(c, pk_esize(PREC_F64) * (V.pa_elems + V.pb_elems) + sizeof(double) * V.partial_elems + 1024);
    if (rc) return rc;
    PackSrc a, b;
    a.data = o1.data;
    b.data = o2.data;
    return run_gemm_packed(c, *p, PREC_F64, a, b, const_cast<double*>(o3.data), f1, f2);
  }
  rc = ws_reserve(c, sizeof(double) * (p->partial_elems + 32));
  if (rc) return rc;
  return run_gemm(c, *p, o1.data, 0, o2.data, 0, const_cast<double*>(o3.data), f1, f2, PREC_F64);
}

// ---------------------------------------------------------------------------
// contract surface: lib_cc/mod_contract.f90
// ---------------------------------------------------------------------------
static bool flat_pair(const char (*pairs)[2], int n, int self) {
  for (int i = 0; i < n; i++) {
    if (i == self) continue;
    if (pairs[i][0] == pairs[self][0] && pairs[i][1] == pairs[self][1]) return false;
  }
  return true;
}

static MapArgs identity_map(const Sym& S, const Desc& dout, const Desc& din) {
  MapArgs a;
  std::memset(&a, 0, sizeof a);
  a.out = make_layout(S, dout);
  a.in = make_layout(S, din);
  a.nterm = 1;
  a.term[0].perm[0] = 0;
  a.term[0].perm[1] = 1;
  a.term[0].perm[2] = 2;
  a.term[0].perm[3] = 3;
  a.term[0].coef = 1.0;
  a.alpha = 1.0;
  a.beta = 0.0;
  a.nbatch = 1;
  return a;
}

// mod_contract.f90:111-160 fused: v3 += X - P(pq)X - P(rs)X + P(pq)P(rs)X,
// gathered straight into v3's (possibly packed) layout
// gather of din into dout where output dim i comes from input dim perm[i]
static MapArgs perm_map(const Sym& S, const Desc& dout, const Desc& din, const int* perm) {
  MapArgs a = identity_map(S, dout, din);
  for (int i = 0; i < 4; i++) a.term[0].perm[i] = (int8_t)perm[i];
  return a;
}

// v3 += X - P(pq)X - P(rs)X + P(pq)P(rs)X (the swaps as requested) where the
// scratch X holds v3's indices in another order: v3 dim i is X dim pos[i]
static MapArgs anti_map_perm(const Sym& S, const Desc& d3, const Desc& dx, const int* pos, bool apq, bool ars) {
  MapArgs a = identity_map(S, d3, dx);
  a.nterm = 0;
  for (int i = 0; i < (apq ? 2 : 1); i++)
    for (int j = 0; j < (ars ? 2 : 1); j++) {
      MapTerm& t = a.term[a.nterm++];
      const int src[4] = {i ? 1 : 0, i ? 0 : 1, j ? 3 : 2, j ? 2 : 3};
      for (int d = 0; d < 4; d++) t.perm[d] = (int8_t)pos[src[d]];
      t.coef = ((i + j) & 1) ? -1.0 : 1.0;
    }
  a.beta = 1.0;
  return a;
}

static MapArgs anti_map(const Sym& S, const Desc& d3, const Desc& dx, bool apq, bool ars) {
  MapArgs a = identity_map(S, d3, dx);
  a.nterm = 0;
  for (int i = 0; i < (apq ? 2 : 1); i++)
    for (int j = 0; j < (ars ? 2 : 1); j++) {
      MapTerm& t = a.term[a.nterm++];
      t.perm[0] = i ? 1 : 0;
      t.perm[1] = i ? 0 : 1;
      t.perm[2] = j ? 3 : 2;
      t.perm[3] = j ? 2 : 3;
      t.coef = ((i + j) & 1) ? -1.0 : 1.0;
    }
  a.beta = 1.0;
  return a;
}

// Regrouping of a 444 contraction whose operands split pairs between GEMM
// dimensions (e.g. the ring term R(ae,im) W(em,bj) -> X(ab,ij)): the operands
// are gathered into full-pair layouts (M M | K K), (K K | N N) and the GEMM
// writes X(M M | N N), so that every pair is one GEMM digit and the GEMM runs
// per pair irrep (few large blocks) instead of per irrep of every index.  X is
// then mapped into v3 with the requested antisymmetrisation.  The result is the
// same contraction; only the GEMM blocking differs.
static int regroup_444(const Sym& S, const Operand& o1, const Operand& o2, const Operand& o3, bool apq, bool ars,
                       ContractPlan& P) {
  const char* L[3] = {o1.lab, o2.lab, o3.lab};
  const Desc* D[3] = {&o1.d, &o2.d, &o3.d};
  auto find = [&](int t, char c) -> int {
    for (int i = 0; i < 4; i++)
      if (L[t][i] == c) return i;
    return -1;
  };
  // classes: 0 = M (v1,v3), 1 = N (v2,v3), 2 = K (v1,v2)
  auto cls = [&](char c) -> int {
    const bool a = find(0, c) >= 0, b = find(1, c) >= 0, d = find(2, c) >= 0;
    if (a && d && !b) return 0;
    if (b && d && !a) return 1;
    if (a && b && !d) return 2;
    return -1;
  };
  int cnt[3][3] = {{0}};
  for (int t = 0; t < 3; t++)
    for (int i = 0; i < 4; i++) {
      const int k = cls(L[t][i]);
      if (k < 0) return DPD_OK;
      cnt[t][k]++;
    }
  if (cnt[0][0] != 2 || cnt[0][2] != 2 || cnt[1][2] != 2 || cnt[1][1] != 2 || cnt[2][0] != 2 || cnt[2][1] != 2)
    return DPD_OK;
  bool mixed = false;
  for (int t = 0; t < 3; t++)
    for (int s2 = 0; s2 < 2; s2++)
      if (cls(L[t][2 * s2]) != cls(L[t][2 * s2 + 1])) mixed = true;
  if (!mixed) return DPD_OK;
  // the ordered pair of class k: taken from a tensor that already holds it as a pair
  auto pair_of = [&](int k, const int* pref, int npref, char* out) {
    for (int pi = 0; pi < npref; pi++) {
      const int t = pref[pi];
      for (int s2 = 0; s2 < 2; s2++)
        if (cls(L[t][2 * s2]) == k && cls(L[t][2 * s2 + 1]) == k) {
          out[0] = L[t][2 * s2];
          out[1] = L[t][2 * s2 + 1];
          return;
        }
    }
    int n = 0;
    for (int i = 0; i < 4 && n < 2; i++)
      if (cls(L[pref[0]][i]) == k) out[n++] = L[pref[0]][i];
  };
  char pm[2], pn[2], pk[2];
  const int prefM[2] = {0, 2}, prefN[2] = {1, 2}, prefK[2] = {1, 0};
  pair_of(0, prefM, 2, pm);
  pair_of(1, prefN, 2, pn);
  pair_of(2, prefK, 2, pk);
  // target layouts; an operand already in one of its target forms stays as is
  auto build = [&](int t, const char* x, const char* y, char* lab, Desc& dx, bool& moved, MapArgs& m,
                   int64_t& len) {
    const char* cur = L[t];
    const bool xy = cur[0] == x[0] && cur[1] == x[1] && cur[2] == y[0] && cur[3] == y[1];
    const bool yx = cur[0] == y[0] && cur[1] == y[1] && cur[2] == x[0] && cur[3] == x[1];
    if ((xy || yx) && D[t]->iop[0] == 1 && D[t]->iop[1] == 1) {
      std::memcpy(lab, cur, 4);
      dx = *D[t];
      moved = false;
      return;
    }
    const char want[4] = {x[0], x[1], y[0], y[1]};
    int perm[4];
    dx = *D[t];
    for (int i = 0; i < 4; i++) {
      perm[i] = find(t, want[i]);
      std::memcpy(dx.pop[i], D[t]->pop[perm[i]], sizeof dx.pop[i]);
      lab[i] = want[i];
    }
    dx.iop[0] = dx.iop[1] = 1;
    m = perm_map(S, dx, *D[t], perm);
    len = m.out.len;
    moved = true;
  };
  build(0, pm, pk, P.lab1x, P.d1x, P.exp1, P.map1, P.len1x);
  build(1, pk, pn, P.lab2x, P.d2x, P.exp2, P.map2, P.len2x);
  // result: v3 itself when it already is (M M | N N) full and needs no antisymmetrisation
  const char want3[4] = {pm[0], pm[1], pn[0], pn[1]};
  const bool same3 = std::memcmp(o3.lab, want3, 4) == 0 && o3.d.iop[0] == 1 && o3.d.iop[1] == 1;
  if (same3 && !apq && !ars) {
    P.epi_on = false;
    std::memcpy(P.lab3x, o3.lab, 4);
    P.d3x = o3.d;
    return DPD_OK;
  }
  Desc dx = o3.d;
  int pos[4];
  for (int i = 0; i < 4; i++) {
    const int src = find(2, want3[i]);
    std::memcpy(dx.pop[i], o3.d.pop[src], sizeof dx.pop[i]);
    P.lab3x[i] = want3[i];
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      if (want3[j] == o3.lab[i]) pos[i] = j;
  dx.iop[0] = dx.iop[1] = 1;
  P.d3x = dx;
  P.epi = anti_map_perm(S, o3.d, dx, pos, apq, ars);
  P.len3x = P.epi.in.len;
  P.epi_on = true;
  return DPD_OK;
}

static int build_contract_plan(const Sym& S, int kind, const Operand& o1, const Operand& o2,
                               const Operand& o3, int anti, int nbatch, ContractPlan& P) {
  P.kind = kind;
  P.anti = anti;
  P.nbatch = nbatch;
  // pairs whose labels must be flat (check_flat, mod_contract.f90:493-510)
  int iexp1[2] = {0, 0}, iexp2[2] = {0, 0}, iexp3[2] = {0, 0};
  if (kind == 444) {
    char pairs[6][2] = {{o1.lab[0], o1.lab[1]}, {o1.lab[2], o1.lab[3]}, {o2.lab[0], o2.lab[1]},
                        {o2.lab[2], o2.lab[3]}, {o3.lab[0], o3.lab[1]}, {o3.lab[2], o3.lab[3]}};
    int iop[6] = {o1.d.iop[0], o1.d.iop[1], o2.d.iop[0], o2.d.iop[1], o3.d.iop[0], o3.d.iop[1]};
    int* ex[6] = {&iexp1[0], &iexp1[1], &iexp2[0], &iexp2[1], &iexp3[0], &iexp3[1]};
    for (int i = 0; i < 6; i++)
      if (flat_pair(pairs, 6, i) && iop[i] != 1) *ex[i] = 1;
    if (anti == 0 && (iexp3[0] || iexp3[1]))
      return fail(DPD_EINVAL, "vv2v_444: an output pair must be expanded but anti=0");
  } else if (kind == 424) {
    char pairs[4][2] = {{o1.lab[0], o1.lab[1]}, {o1.lab[2], o1.lab[3]}, {o3.lab[0], o3.lab[1]},
                        {o3.lab[2], o3.lab[3]}};
    int iop[4] = {o1.d.iop[0], o1.d.iop[1], o3.d.iop[0], o3.d.iop[1]};
    int* ex[4] = {&iexp1[0], &iexp1[1], &iexp3[0], &iexp3[1]};
    for (int i = 0; i < 4; i++)
      if (flat_pair(pairs, 4, i) && iop[i] != 1) *ex[i] = 1;
    if (iexp3[0] && anti != 1)
      return fail(DPD_EINVAL, "ERROR: vv2v_424 error for the 1st pair not antisymmertize but be expand.");
    if (iexp3[1] && anti != 2)
      return fail(DPD_EINVAL, "ERROR: vv2v_424 error for the 2nd pair not antisymmertize but be expand.");
  } else {
    char pairs[4][2] = {{o1.lab[0], o1.lab[1]}, {o1.lab[2], o1.lab[3]}, {o2.lab[0], o2.lab[1]},
                        {o2.lab[2], o2.lab[3]}};
    int iop[4] = {o1.d.iop[0], o1.d.iop[1], o2.d.iop[0], o2.d.iop[1]};
    int* ex[4] = {&iexp1[0], &iexp1[1], &iexp2[0], &iexp2[1]};
    for (int i = 0; i < 4; i++)
      if (flat_pair(pairs, 4, i) && iop[i] != 1) *ex[i] = 1;
  }
  // expansions (iexpso semantics, popp==popq required)
  auto expand = [&](const Operand& o, int* ie, bool& flag, Desc& dx, MapArgs& m, int64_t& len) -> int {
    dx = o.d;
    flag = ie[0] || ie[1];
    if (!flag) return DPD_OK;
    if (ie[0] && !pops_eq(o.d.pop[0], o.d.pop[1])) return fail(DPD_EINVAL, "fatal error: popp /= popq");
    if (ie[1] && !pops_eq(o.d.pop[2], o.d.pop[3])) return fail(DPD_EINVAL, "fatal error: popr /= pops");
    if (ie[0]) dx.iop[0] = 1;
    if (ie[1]) dx.iop[1] = 1;
    m = identity_map(S, dx, o.d);
    len = m.out.len;
    return DPD_OK;
  };
  int rc = expand(o1, iexp1, P.exp1, P.d1x, P.map1, P.len1x);
  if (rc) return rc;
  if (kind != 424) {
    rc = expand(o2, iexp2, P.exp2, P.d2x, P.map2, P.len2x);
    if (rc) return rc;
  } else {
    P.d2x = o2.d;
  }
  P.d3x = o3.d;
  if (iexp3[0]) P.d3x.iop[0] = 1;
  if (iexp3[1]) P.d3x.iop[1] = 1;
  bool apq = (kind == 444) ? (anti == 1 || anti == 3) : (kind == 424 && anti == 1);
  bool ars = (kind == 444) ? (anti == 2 || anti == 3) : (kind == 424 && anti == 2);
  if (apq && (P.d3x.iop[0] != 1 || !pops_eq(o3.d.pop[0], o3.d.pop[1])))
    return fail(DPD_EINVAL, "ERROR: vv2v_%d anti 1st pair error.", kind);
  if (ars && (P.d3x.iop[1] != 1 || !pops_eq(o3.d.pop[2], o3.d.pop[3])))
    return fail(DPD_EINVAL, "ERROR: vv2v_%d anti 2nd pair error.", kind);
  std::memcpy(P.lab1x, o1.lab, 4);
  std::memcpy(P.lab2x, o2.lab, 4);
  std::memcpy(P.lab3x, o3.lab, 4);
  P.epi_on = anti != 0;
  if (anti != 0) {
    P.epi = anti_map(S, o3.d, P.d3x, apq, ars);
    P.len3x = P.epi.in.len;
  }
  if (kind == 444) {
    rc = regroup_444(S, o1, o2, o3, apq, ars, P);
    if (rc) return rc;
  }
  // batch strides of the scratch copies
  P.batch1 = o1.batched;
  P.batch2 = o2.batched;
  P.bstride1 = P.exp1 ? P.len1x : o1.bstride;
  P.bstride2 = P.exp2 ? P.len2x : o2.bstride;
  P.bstride3 = P.epi_on ? P.len3x : o3.bstride;
  Operand e1 = o1, e2 = o2, e3 = o3;
  e1.d = P.d1x;
  e2.d = P.d2x;
  e3.d = P.d3x;
  std::memcpy(e1.lab, P.lab1x, 4);
  std::memcpy(e2.lab, P.lab2x, 4);
  std::memcpy(e3.lab, P.lab3x, 4);
  e1.bstride = P.bstride1;
  e2.bstride = P.bstride2;
  e3.bstride = P.bstride3;
  Operand m1, m2, m3;
  rc = check_engine(S, kind, e1, e2, e3, m1, m2, m3);
  if (rc) return rc;
  return build_gemm_plan(S, m1, m2, m3, nbatch, P.gemm);
}

// one contraction between plan lookup and the GEMM: which operands are packed,
// gathered or fused, where the (scratch) result goes
struct CallPrep {
  ContractPlan* P = nullptr;
  bool packed = false, fuse1 = false, fuse2 = false, gx1 = false, gx2 = false;
  int nb1 = 1, nb2 = 1;
  const double* p1 = nullptr;
  const double* p2 = nullptr;
  const float* p1f = nullptr;
  const float* p2f = nullptr;
  double* p3 = nullptr;
  double beta = 1.0;
  bool p3_f32 = false;  // the scratch result is FP32 (FP32-accumulating packed GEMM)
  bool x1_f32 = false, x2_f32 = false;  // expanded operands held in FP32 (they only feed the pack)
  void* p3_ext = nullptr;  // scratch result outside the workspace (a held-back epilogue's slot)
  PackSrc a, b;
};

static int plan_of(dpd_ctx* c, int kind, const Operand& o1, const Operand& o2, const Operand& o3, int nbatch,
                   int anti, ContractPlan** out, std::string* key_out = nullptr) {
  std::string key = sig_of(kind, o1, o2, o3, nbatch, anti, 0);
  auto it = c->contract_cache.find(key);
  if (it == c->contract_cache.end()) {
    auto np = std::make_unique<ContractPlan>();
    np->gemm.cfg_req = c->gemm_cfg;
    np->gemm.split_req = c->split_k;
    int rc = build_contract_plan(c->sym, kind, o1, o2, o3, anti, nbatch, *np);
    if (rc) return rc;
    *out = np.get();
    c->contract_cache.emplace(key, std::move(np));
  } else {
    *out = it->second.get();
  }
  if (key_out) *key_out = std::move(key);
  return DPD_OK;
}

// decisions of a call and the workspace it needs
static int call_setup(dpd_ctx* c, ContractPlan* P, const Operand& o1, const Operand& o2, int nbatch, int prec,
                      CallPrep& r, size_t* need) {
  r.P = P;
  r.nb1 = o1.batched ? nbatch : 1;
  r.nb2 = o2.batched ? nbatch : 1;
  // packed panels: always for BF16x3, otherwise when an operand is static
  r.packed = prec == PREC_BF16X3 || c->packing == 2 || (c->packing == 1 && (o1.static_id || o2.static_id));
  if (prec == PREC_BF16X3 && c->packing == 0) return fail(DPD_EINVAL, "BF16x3 needs packed panels (packing 0)");
  // packed panels read sorted constant operands straight from the unsorted
  // tensors (packed once); sorted per-call operands are expanded by the gather
  // kernel first (measured as fast as the fused sort for them, and simpler)
  r.fuse1 = r.packed && P->exp1 && o1.static_id;
  r.fuse2 = r.packed && P->exp2 && o2.static_id;
  r.gx1 = P->exp1 && !r.fuse1;
  r.gx2 = P->exp2 && !r.fuse2;
  size_t n = 0;
  if (r.packed) {
    int rc0 = build_variant(P->gemm, prec);
    if (rc0) return rc0;
    const GemmVariant& V = P->gemm.pk[prec];
    if (!o1.static_id) n += pk_esize(prec) * V.pa_elems + 256;
    if (!o2.static_id) n += pk_esize(prec) * V.pb_elems + 256;
    n += sizeof(double) * V.partial_elems + 256;
  }
  // an expanded operand of an FP32 or BF16x3 packed GEMM is only read by the
  // pack, which rounds it to FP32 or BF16 hi + lo anyway: it is gathered in FP32
  r.x1_f32 = r.gx1 && r.packed && prec != PREC_F64;
  r.x2_f32 = r.gx2 && r.packed && prec != PREC_F64;
  if (r.gx1) n += (r.x1_f32 ? sizeof(float) : sizeof(double)) * P->len1x * r.nb1 + 256;
  if (r.gx2) n += (r.x2_f32 ? sizeof(float) : sizeof(double)) * P->len2x * r.nb2 + 256;
  // the scratch result of an FP32-accumulating packed GEMM is kept in FP32:
  // the accumulators are FP32, so nothing is lost and half the bytes move
  r.p3_f32 = P->epi_on && r.packed && prec != PREC_F64;
  if (P->epi_on) n += (r.p3_f32 ? sizeof(float) : sizeof(double)) * P->len3x * nbatch + 256;
  if (!r.packed) n += sizeof(double) * P->gemm.partial_elems + 256;
  *need += n;
  return DPD_OK;
}

static ProfRec* prof_open(dpd_ctx* c, const char* tag, int64_t flops, int prec, int cfg, int packed,
                          int64_t bytes, int split, int kind = 0) {
  if (!c->prof.on || !tag || c->prof.used + 2 > c->prof.pool.size()) return nullptr;
  c->prof.recs.push_back({tag, c->prof.pool[c->prof.used], c->prof.pool[c->prof.used + 1], flops, prec, cfg, packed,
                          bytes, split, kind});
  c->prof.used += 2;
  return &c->prof.recs.back();
}

// a gather of a tagged contraction, bracketed by events when profiling: its
// algorithmic bytes are the input once, the output once (twice when it is
// accumulated into)
static ProfRec* prof_gather(dpd_ctx* c, const char* tag, const char* what, const MapArgs& m, int esz_in,
                            int esz_out) {
  if (!c->prof.on || !tag) return nullptr;
  thread_local std::string name;
  name = std::string(tag) + ":" + what;
  const int64_t nb = m.nbatch;
  const int64_t bytes = nb * (m.in.len * esz_in + m.out.len * esz_out * (m.beta != 0.0 ? 2 : 1));
  ProfRec* pr = prof_open(c, name.c_str(), 0, PREC_F64, 0, 0, bytes, 0, 1);
  if (pr) (void)hipEventRecord(pr->a, c->stream);
  return pr;
}
static void prof_close(dpd_ctx* c, ProfRec* pr) {
  if (pr) (void)hipEventRecord(pr->b, c->stream);
}

// Source codes of a plan's gather (slot 0/1: expansions of operand 1/2, 2: the
// epilogue), built on the first call outside graph capture and shared through
// the context's cache by every plan with the same layouts and terms.  Inputs
// of 2^31 elements or more, and codes beyond kMapCacheBytes in total, keep the
// decoding kernel.
constexpr size_t kMapCacheBytes = size_t(2) << 30;
static const int32_t* plan_code(dpd_ctx* c, ContractPlan* P, int slot, const MapArgs& m) {
  if (P->code[slot] || P->code_tried[slot] || g_capturing) return P->code[slot];
  P->code_tried[slot] = true;
  if (m.nterm < 1 || m.nterm > kMaxTerms || m.in.len >= INT32_MAX || m.out.len <= 0) return nullptr;
  std::string key((const char*)&m.out, sizeof(DevLayout));
  key.append((const char*)&m.in, sizeof(DevLayout));
  for (int t = 0; t < m.nterm; t++) key.append((const char*)m.term[t].perm, 4);
  key.push_back((char)m.nterm);
  auto it = c->map_cache.find(key);
  if (it != c->map_cache.end()) return P->code[slot] = (const int32_t*)it->second->p;
  const size_t bytes = sizeof(int32_t) * (size_t)m.nterm * (size_t)m.out.len;
  if (c->map_bytes + bytes > kMapCacheBytes) return nullptr;
  auto buf = std::make_unique<DevBuf>();
  if (buf->alloc(bytes)) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (launch_map_build(m, (int32_t*)buf->p, c->stream) != hipSuccess) return nullptr;
  c->map_bytes += bytes;
  P->code[slot] = (const int32_t*)buf->p;
  c->map_cache.emplace(key, std::move(buf));
  return P->code[slot];
}

// expansions into the workspace, the scratch result, the packing sources
static int call_pre(dpd_ctx* c, CallPrep& r, const Operand& o1, const Operand& o2, const Operand& o3, int nbatch,
                    const char* tag) {
  ContractPlan* P = r.P;
  r.p1 = o1.data;
  r.p2 = o2.data;
  r.p1f = o1.data_f32;
  r.p2f = o2.data_f32;
  // expansion of operand 1 or 2 into the workspace (FP32 when it only feeds a pack)
  auto expand = [&](const MapArgs& map, const Operand& o, int nb, int64_t lenx, bool f32, const double** dst,
                    const char* what, int slot) -> int {
    void* x = ws_take(c, (f32 ? sizeof(float) : sizeof(double)) * lenx * nb);
    if (!x) return DPD_ENOMEM;
    MapArgs m = map;
    m.nbatch = nb;
    m.in_bstride = o.bstride;
    m.out_bstride = lenx;
    const int32_t* code = plan_code(c, P, slot, m);
    ProfRec* pr = prof_gather(c, tag, what, m, 8, f32 ? 4 : 8);
    if (code && f32) {
      DPD_HIP_TRY(launch_gather_code_f32out(m, code, o.data, (float*)x, c->stream));
    } else if (code) {
      DPD_HIP_TRY(launch_gather_code(m, code, o.data, (double*)x, c->stream));
    } else if (f32) {
      if (m.out.len > 0 && m.nbatch > 0) DPD_HIP_TRY(launch_gather_map_f32out(m, o.data, (float*)x, c->stream));
    } else {
      int rc = gather(c, m, o.data, (double*)x);
      if (rc) return rc;
    }
    prof_close(c, pr);
    *dst = (const double*)x;
    return DPD_OK;
  };
  // an expansion that only feeds a packed side is composed into the pack: the
  // panels are written straight from the operand through the gather's code
  auto composable = [&](int slot, const MapArgs& map) -> const int32_t* {
    if (!r.packed || map.nterm != 1 || map.term[0].coef != 1.0 || map.alpha != 1.0 || map.beta != 0.0)
      return nullptr;
    return plan_code(c, P, slot, map);
  };
  const int32_t* comp[2] = {r.gx1 ? composable(0, P->map1) : nullptr, r.gx2 ? composable(1, P->map2) : nullptr};
  if (r.gx1 && !comp[0]) {
    int rc = expand(P->map1, o1, r.nb1, P->len1x, r.x1_f32, &r.p1, "x1", 0);
    if (rc) return rc;
    r.p1f = nullptr;
  }
  if (r.gx2 && !comp[1]) {
    int rc = expand(P->map2, o2, r.nb2, P->len2x, r.x2_f32, &r.p2, "x2", 1);
    if (rc) return rc;
    r.p2f = nullptr;
  }
  // scratch result: every element the epilogue reads lies in a block some GEMM
  // task writes (tasks without K segments write zeros), so it is written with
  // beta = 0 and needs no clearing
  r.p3 = const_cast<double*>(o3.data);
  r.beta = 1.0;
  if (P->epi_on) {
    r.p3 = r.p3_ext ? (double*)r.p3_ext
                    : (double*)ws_take(c, (r.p3_f32 ? sizeof(float) : sizeof(double)) * P->len3x * nbatch);
    if (!r.p3) return DPD_ENOMEM;
    r.beta = 0.0;
  }
  if (r.packed) {
    r.a = PackSrc();
    r.b = PackSrc();
    r.a.data = r.fuse1 ? o1.data : r.p1;
    r.a.f32 = !r.fuse1 && r.x1_f32;
    r.a.id = o1.static_id;
    // dims of a sorted operand that are GEMM rows: the ones it shares with the result
    auto rows_of = [&](const char* lab) {
      int m = 0;
      for (int d = 0; d < 4; d++)
        if (std::memchr(P->lab3x, lab[d], 4)) m |= 1 << d;
      return m;
    };
    if (r.fuse1) {
      r.a.map = &P->map1;
      r.a.rowmask = rows_of(P->lab1x);
      r.a.xlen = P->len1x;
      r.a.bstride = o1.bstride;
    }
    r.b.data = r.fuse2 ? o2.data : r.p2;
    r.b.f32 = !r.fuse2 && r.x2_f32;
    r.b.id = o2.static_id;
    if (r.fuse2) {
      r.b.map = &P->map2;
      r.b.rowmask = rows_of(P->lab2x);
      r.b.xlen = P->len2x;
      r.b.bstride = o2.bstride;
    }
    // per-call operands pack through source codes kept with the plan
    auto coded = [&](PackSrc& ps, const Operand& o, int nb, const int32_t* gc, int64_t lenx) {
      if (o.static_id) return;
      ps.coded = true;
      const int64_t len = make_layout(c->sym, o.d).len;
      ps.src_elems = nb > 1 ? (int64_t)(nb - 1) * o.bstride + len : len;
      if (gc) {
        ps.data = o.data;
        ps.f32 = false;
        ps.gcode = gc;
        ps.gxlen = lenx;
        ps.gbstride = o.bstride;
      } else if (ps.data != o.data) {
        ps.src_elems = lenx * nb;  // an expanded copy in the workspace
      }
    };
    coded(r.a, o1, r.nb1, comp[0], P->len1x);
    coded(r.b, o2, r.nb2, comp[1], P->len2x);
  }
  return DPD_OK;
}

// antisymmetrise + compress + accumulate the scratch result into v3
static int call_post(dpd_ctx* c, const CallPrep& r, const Operand& o3, int nbatch, const char* tag) {
  const ContractPlan* P = r.P;
  if (!P->epi_on) return DPD_OK;
  MapArgs m = P->epi;
  m.nbatch = nbatch;
  m.in_bstride = P->len3x;
  m.out_bstride = o3.bstride;
  const int32_t* code = plan_code(c, const_cast<ContractPlan*>(P), 2, m);
  ProfRec* pr = prof_gather(c, tag, "epi", m, r.p3_f32 ? 4 : 8, 8);
  if (code && r.p3_f32) {
    DPD_HIP_TRY(launch_gather_code_f32in(m, code, (const float*)r.p3, const_cast<double*>(o3.data), c->stream));
  } else if (code) {
    DPD_HIP_TRY(launch_gather_code(m, code, r.p3, const_cast<double*>(o3.data), c->stream));
  } else if (r.p3_f32) {
    DPD_HIP_TRY(launch_gather_map_f32in(m, (const float*)r.p3, const_cast<double*>(o3.data), c->stream));
  } else {
    int rc = gather(c, m, r.p3, const_cast<double*>(o3.data));
    if (rc) return rc;
  }
  prof_close(c, pr);
  return DPD_OK;
}

// algorithmic bytes of one call: both operands once as the kernel reads them
// (panels, or the tensors in their storage precision) and the result (read too
// when accumulated)
static int64_t call_bytes(dpd_ctx* c, const CallPrep& r, const Operand& o1, const Operand& o2, int prec) {
  const ContractPlan* P = r.P;
  const bool use_f32 = prec != PREC_F64;
  int64_t bytes = P->gemm.c_elems * (r.p3_f32 ? 4 : 8) * (r.beta != 0.0 ? 2 : 1);
  if (r.packed) {
    const GemmVariant& V = P->gemm.pk[prec];
    bytes += (V.pa_elems + V.pb_elems) * pk_esize(prec);
  } else {
    const int64_t l1 = (r.gx1 ? P->len1x : make_layout(c->sym, o1.d).len) * r.nb1;
    const int64_t l2 = (r.gx2 ? P->len2x : make_layout(c->sym, o2.d).len) * r.nb2;
    bytes += l1 * ((use_f32 && r.p1f) ? 4 : 8) + l2 * ((use_f32 && r.p2f) ? 4 : 8);
  }
  return bytes;
}

// algorithmic bytes of the packing of a call's per-call operands: each source
// element read once, each panel element written once (cached sides cost nothing)
static int64_t pack_bytes(dpd_ctx* c, const CallPrep& r, const Operand& o1, const Operand& o2, int prec) {
  const GemmVariant& V = r.P->gemm.pk[prec];
  int64_t bytes = 0;
  auto side = [&](const PackSrc& s, const Operand& o, bool gx, int nb, int64_t lenx, int64_t panel) {
    if (s.id) return;
    const bool from_operand = !gx || s.map || s.gcode;
    const int64_t src = (from_operand ? make_layout(c->sym, o.d).len : lenx) * nb;
    bytes += src * (s.f32 ? 4 : 8) + panel * (int64_t)pk_esize(prec);
  };
  side(r.a, o1, r.gx1, r.nb1, r.P->len1x, V.pa_elems);
  side(r.b, o2, r.gx2, r.nb2, r.P->len2x, V.pb_elems);
  return bytes;
}

// the held-back epilogue of a call (its scratch result in the context slot
// d->slot): only an accumulating epilogue that runs through a source code;
// false when the call is to finish its own epilogue
static bool epi_defer_setup(dpd_ctx* c, ContractPlan* P, CallPrep& r, const Operand& o3, int nbatch,
                            EpiDefer* d) {
  d->deferred = false;
  if (!P->epi_on || d->slot < 0 || d->slot >= kEpiMax) return false;
  MapArgs m = P->epi;
  m.nbatch = nbatch;
  m.in_bstride = P->len3x;
  m.out_bstride = o3.bstride;
  if (m.beta != 1.0 || m.nterm < 1 || m.nterm > kMaxTerms || m.out.len <= 0) return false;
  const int32_t* code = plan_code(c, P, 2, m);
  if (!code) return false;
  const size_t bytes = (r.p3_f32 ? sizeof(float) : sizeof(double)) * (size_t)P->len3x * nbatch;
  DevBuf& b = c->epi_slot[d->slot];
  if (!b.p || b.bytes < bytes) {
    if (g_capturing) return false;  // grown on an eager call only
    if (hipStreamSynchronize(c->stream) != hipSuccess) return false;
    bump_mem_gen();  // captured graphs may hold the old slot
    b.reset();
    if (b.alloc(bytes) != DPD_OK) {
      (void)hipGetLastError();
      return false;
    }
  }
  r.p3_ext = b.p;
  d->f32 = r.p3_f32;
  d->src.code = code;
  d->src.in = b.p;
  d->src.in_bstride = m.in_bstride;
  d->src.nt = m.nterm;
  for (int t = 0; t < kMaxTerms; t++) d->src.coef[t] = t < m.nterm ? m.term[t].coef : 0.0;
  d->src.alpha = m.alpha;
  d->n = m.out.len;
  d->out_bstride = m.out_bstride;
  d->nbatch = nbatch;
  d->out = const_cast<double*>(o3.data);
  return true;
}

int epi_flush(dpd_ctx* c, EpiDefer* d, int n, const char* tag) {
  int i = 0;
  while (i < n) {
    if (!d[i].deferred) {
      i++;
      continue;
    }
    // a run of held-back epilogues into the same output, tensors and precision
    EpiFuse g;
    std::memset(&g, 0, sizeof g);
    g.n = d[i].n;
    g.out_bstride = d[i].out_bstride;
    g.nbatch = d[i].nbatch;
    int j = i;
    while (j < n && d[j].deferred && g.ns < kEpiMax && d[j].out == d[i].out && d[j].n == d[i].n &&
           d[j].out_bstride == d[i].out_bstride && d[j].nbatch == d[i].nbatch && d[j].f32 == d[i].f32)
      g.s[g.ns++] = d[j++].src;
    ProfRec* pr = nullptr;
    if (c->prof.on && tag) {
      // algorithmic bytes: the output read and written once, every scratch result read once
      int64_t bytes = 2 * 8 * g.n * g.nbatch;
      for (int s = 0; s < g.ns; s++) bytes += (d[i].f32 ? 4 : 8) * g.s[s].in_bstride * g.nbatch;
      pr = prof_open(c, tag, 0, PREC_F64, 0, 0, bytes, 0, 1);
      if (pr) DPD_HIP_TRY(hipEventRecord(pr->a, c->stream));
    }
    DPD_HIP_TRY(launch_epi_fused(g, d[i].f32, d[i].out, c->stream));
    prof_close(c, pr);
    for (int k = i; k < j; k++) d[k].deferred = false;
    i = j;
  }
  return DPD_OK;
}

int contract_call(dpd_ctx* c, int kind, const Operand& o1, const Operand& o2, Operand o3,
                  double fact, int anti, int nbatch, int prec, int64_t* flops, const char* tag, EpiDefer* defer) {
  ContractPlan* P;
  int rc = plan_of(c, kind, o1, o2, o3, nbatch, anti, &P);
  if (rc) return rc;
  if (flops) *flops += P->gemm.flops;
  CallPrep r;
  size_t need = 0;
  if ((rc = call_setup(c, P, o1, o2, nbatch, prec, r, &need))) return rc;
  const bool held = defer && epi_defer_setup(c, P, r, o3, nbatch, defer);
  ws_reset(c);
  if ((rc = ws_reserve(c, need))) return rc;
  if ((rc = call_pre(c, r, o1, o2, o3, nbatch, tag))) return rc;
  const bool use_f32 = (prec != PREC_F64);
  const void* A = (use_f32 && r.p1f) ? (const void*)r.p1f : (const void*)r.p1;
  const void* B = (use_f32 && r.p2f) ? (const void*)r.p2f : (const void*)r.p2;
  ProfRec* pr = prof_open(c, tag, P->gemm.flops, prec, r.packed ? P->gemm.pk[prec].cfg : P->gemm.cfg,
                          r.packed ? 1 : 0, c->prof.on && tag ? call_bytes(c, r, o1, o2, prec) : 0,
                          r.packed ? !P->gemm.pk[prec].split.empty() : !P->gemm.split.empty());
  // the packed path records the events around the GEMM kernel itself
  if (pr && !r.packed) DPD_HIP_TRY(hipEventRecord(pr->a, c->stream));
  if (r.packed) {
    // the packing of the per-call operands, timed as a gather entry "<tag>:pack"
    ProfRec* pp = nullptr;
    if (pr) {
      thread_local std::string pname;
      pname = std::string(tag) + ":pack";
      pp = prof_open(c, pname.c_str(), 0, prec, 0, 1, pack_bytes(c, r, o1, o2, prec), 0, 1);
      if (pp) DPD_HIP_TRY(hipEventRecord(pp->a, c->stream));
    }
    rc = run_gemm_packed(c, P->gemm, prec, r.a, r.b, r.p3, fact, r.beta, pr ? pr->a : nullptr,
                         pr ? pr->b : nullptr, r.p3_f32, pp ? pp->b : nullptr);
    // first call of this plan in this precision: time every tile shape on these
    // operands (results go to a scratch C) and keep the fastest for later calls
    if (!rc && c->gemm_cfg == -1 && !P->gemm.pk[prec].tuned) {
      rc = tune_variant(c, P->gemm, prec, r.a, r.b, fact, r.beta);
      P->gemm.pk[prec].tuned = true;
    }
  } else {
    rc = run_gemm(c, P->gemm, A, (use_f32 && r.p1f) ? 1 : 0, B, (use_f32 && r.p2f) ? 1 : 0, r.p3, fact, r.beta,
                  use_f32 ? PREC_F32 : PREC_F64);
  }
  if (rc) return rc;
  if (pr && !r.packed) DPD_HIP_TRY(hipEventRecord(pr->b, c->stream));
  if (held) {
    defer->deferred = true;
    return DPD_OK;
  }
  return call_post(c, r, o3, nbatch, tag);
}

// ---------------------------------------------------------------------------
// Grouped contraction: the same contraction (kind, labels, anti) of n
// independent problems whose tensors differ in size.  The pre-stages (gathers,
// packing) run per problem; every GEMM task of all of them runs in ONE
// packed-panel launch with one tile shape (task.group selects the problem's
// panels and result); the epilogues run per problem.  The tile shape is timed
// over the whole group on its first call (autotuner) unless one is requested.
// ---------------------------------------------------------------------------
static int group_variants(GroupPlan& G, int prec, int cfg) {
  for (ContractPlan* P : G.members) {
    GemmVariant& V = P->gemm.pk[prec];
    if (V.built && V.cfg == cfg) continue;
    if (g_capturing) return fail(DPD_EINVAL, "tile variant rebuilt during graph capture (internal)");
    clear_pack_cache(V);
    int rc = build_variant_cfg(P->gemm, V, prec, cfg);
    if (rc) return rc;
    V.tuned = true;  // a member's tile is the group's
  }
  return DPD_OK;
}

// merged task list of variants (one per member, all built with one tile shape);
// split-K is decided again over the whole group (the same rule as for one
// plan: split long-K tasks while the launch has fewer than 1024 tiles)
static int group_tasks(const std::vector<const GemmVariant*>& Vs, int mode, bool allow_split, GroupTasks& G) {
  if (g_capturing) return fail(DPD_EINVAL, "grouped task list built during graph capture (internal)");
  bump_mem_gen();  // the group's task tables are reallocated
  std::vector<GemmTask> all;
  for (size_t i = 0; i < Vs.size(); i++)
    for (GemmTask tk : Vs[i]->tasks) {
      tk.group = (int32_t)i;
      all.push_back(tk);
    }
  int64_t work = 0, tiles = 0;
  for (const auto& tk : all) {
    work += (int64_t)tk.tiles_m * tk.tiles_n * std::max(tk.steps, 1);
    tiles += (int64_t)tk.tiles_m * tk.tiles_n;
  }
  const int64_t target = 1024;
  const int64_t min_steps = mode == PREC_F64 ? 8 : 4;
  const int64_t per_wg = std::max<int64_t>(min_steps, (work + target - 1) / target);
  std::vector<int32_t> split;
  int64_t partial = 0;
  int wg = 0, max_mn = 0;
  for (size_t i = 0; i < all.size(); i++) {
    GemmTask& tk = all[i];
    int ks = 1;
    if (allow_split && tiles < target && tk.steps > per_wg)
      ks = (int)std::min<int64_t>(64, (tk.steps + per_wg - 1) / per_wg);
    tk.ksplit = ks;
    tk.part_off = 0;
    if (ks > 1) {
      tk.part_off = partial;
      partial += (int64_t)ks * tk.M * tk.N;
      split.push_back((int32_t)i);
      max_mn = std::max<int64_t>(max_mn, (int64_t)tk.M * tk.N);
    }
    tk.wg_begin = wg;
    wg += tk.tiles_m * tk.tiles_n * ks;
    if (wg < 0) return fail(DPD_EINVAL, "grouped GEMM too large");
  }
  G.nwg = wg;
  G.ntask = (int)all.size();
  G.nsplit = (int)split.size();
  G.max_split_mn = max_mn;
  G.partial_elems = partial;
  if (all.empty()) return DPD_OK;
  const size_t bt = ((sizeof(GemmTask) * all.size() + 255) & ~size_t(255)), bs = sizeof(int32_t) * split.size();
  int rc = G.dev.alloc(bt + bs + 256);
  if (rc) return rc;
  G.d_tasks = (GemmTask*)G.dev.p;
  G.d_split = (int32_t*)((char*)G.dev.p + bt);
  DPD_HIP_TRY(hipMemcpy(G.d_tasks, all.data(), sizeof(GemmTask) * all.size(), hipMemcpyHostToDevice));
  if (bs) DPD_HIP_TRY(hipMemcpy(G.d_split, split.data(), bs, hipMemcpyHostToDevice));
  return DPD_OK;
}

static void group_launch_args(const GroupTasks& T, int prec, int cfg, double fact, double* partial, GemmLaunch& g) {
  std::memset(&g, 0, sizeof g);
  g.tasks = T.d_tasks;
  g.ntask = T.ntask;
  g.nwg = T.nwg;
  g.alpha = fact;
  g.prec = prec;
  g.cfg = cfg;
  g.partial = partial;
  g.split_tasks = T.d_split;
  g.nsplit = T.nsplit;
  g.max_split_mn = T.max_split_mn;
}

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st == hipStreamCaptureStatusActive;
}

static int tune_group(dpd_ctx* c, GroupPlan& G, int prec, const std::vector<CallPrep>& R, double fact) {
  static const int cand[] = {GEMM_128x64, GEMM_64x192, GEMM_128x64D, GEMM_128x128, GEMM_64x128, GEMM_64x64,
                             GEMM_32x32};
  const int ncand = (int)(sizeof cand / sizeof cand[0]);
  float tms[ncand];
  const size_t es = pk_esize(prec);
  const size_t n = G.members.size();
  // scratch results (beta = 0) so the tuning leaves the outputs alone
  std::vector<DevBuf> cs(n);
  for (size_t i = 0; i < n; i++) {
    int rc = cs[i].alloc(sizeof(double) * std::max<int64_t>(c_span(G.members[i]->gemm), 1));
    if (rc) return rc;
  }
  hipEvent_t e0, e1;
  DPD_HIP_TRY(hipEventCreate(&e0));
  DPD_HIP_TRY(hipEventCreate(&e1));
  int rc = DPD_OK, best = -1;
  float best_ms = 0.0f;
  for (int ci = 0; ci < ncand && !rc; ci++) {
    tms[ci] = 1e30f;
    const int cfg = cand[ci];
    if (cfg == GEMM_128x64D && prec != PREC_BF16X3) continue;  // the same as GEMM_128x64 there
    std::vector<GemmVariant> T(n);
    std::vector<DevBuf> pa(n), pb(n);
    std::vector<const GemmVariant*> Vs(n);
    GroupTab gt;
    std::memset(&gt, 0, sizeof gt);
    for (size_t i = 0; i < n && !rc; i++) {
      GemmPlan& p = G.members[i]->gemm;
      if ((rc = build_variant_cfg(p, T[i], prec, cfg))) break;
      if ((T[i].pa_elems && (rc = pa[i].alloc(es * T[i].pa_elems))) ||
          (T[i].pb_elems && (rc = pb[i].alloc(es * T[i].pb_elems))))
        break;
      if ((T[i].pa_elems && (rc = pack_side(c, p, T[i], 0, R[i].a, pa[i].p))) ||
          (T[i].pb_elems && (rc = pack_side(c, p, T[i], 1, R[i].b, pb[i].p))))
        break;
      Vs[i] = &T[i];
      gt.e[i] = {pa[i].p, pb[i].p, (double*)cs[i].p, 0.0, 0, 0};
    }
    if (rc) break;
    GroupTasks TT;
    if ((rc = group_tasks(Vs, prec, c->split_k != 0, TT))) break;
    if (TT.nwg == 0) continue;
    DevBuf part;
    if (TT.partial_elems && (rc = part.alloc(sizeof(double) * TT.partial_elems))) break;
    GemmLaunch g;
    group_launch_args(TT, prec, cfg, fact, (double*)part.p, g);
    DPD_HIP_TRY(launch_gemm_pk_grouped(g, gt, c->stream));
    DPD_HIP_TRY(hipEventRecord(e0, c->stream));
    for (int r = 0; r < 3; r++) DPD_HIP_TRY(launch_gemm_pk_grouped(g, gt, c->stream));
    DPD_HIP_TRY(hipEventRecord(e1, c->stream));
    DPD_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    DPD_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    tms[ci] = ms;
    if (best < 0 || ms < best_ms) {
      best = cfg;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (best >= 0) {
    // among shapes within 5% of the fastest the first listed wins (stable choice)
    for (int ci = 0; ci < ncand; ci++)
      if (tms[ci] <= 1.05f * best_ms) {
        best = cand[ci];
        break;
      }
    G.cfg = best;
  }
  return DPD_OK;
}

int contract_group(dpd_ctx* c, int kind, const Operand* o1, const Operand* o2, const Operand* o3, int n,
                   double fact, int anti, const int* nbatch, int prec, const char* tag) {
  if (n <= 0) return DPD_OK;
  if (n > kMaxGroup) {
    for (int i0 = 0; i0 < n; i0 += kMaxGroup) {
      const int m = std::min(kMaxGroup, n - i0);
      int rc = contract_group(c, kind, o1 + i0, o2 + i0, o3 + i0, m, fact, anti, nbatch + i0, prec, tag);
      if (rc) return rc;
    }
    return DPD_OK;
  }
  std::vector<ContractPlan*> plans(n);
  std::string gkey = std::to_string(prec) + "#";
  bool all_packed = true;
  for (int i = 0; i < n; i++) {
    std::string k;
    int rc = plan_of(c, kind, o1[i], o2[i], o3[i], nbatch[i], anti, &plans[i], &k);
    if (rc) return rc;
    gkey += k;
    gkey += '|';
    all_packed = all_packed && (prec == PREC_BF16X3 || c->packing == 2 ||
                                (c->packing == 1 && (o1[i].static_id || o2[i].static_id)));
  }
  if (!all_packed) {
    // the gathering kernels have no grouped form: one launch per problem
    for (int i = 0; i < n; i++) {
      int rc = contract_call(c, kind, o1[i], o2[i], o3[i], fact, anti, nbatch[i], prec, nullptr, tag);
      if (rc) return rc;
    }
    return DPD_OK;
  }
  const bool cap = capturing(c->stream);
  auto it = c->group_cache.find(gkey);
  GroupPlan* G;
  if (it == c->group_cache.end()) {
    if (cap) return fail(DPD_EINVAL, "grouped contraction first seen during graph capture (internal)");
    auto g = std::make_unique<GroupPlan>();
    g->members = plans;
    std::vector<GemmTask> all;
    for (ContractPlan* P : plans) all.insert(all.end(), P->gemm.tasks.begin(), P->gemm.tasks.end());
    g->cfg = c->gemm_cfg >= GEMM_64x64 ? c->gemm_cfg : choose_gemm_cfg(all, -1);
    if (g->cfg == GEMM_V1) g->cfg = GEMM_64x64;
    g->tuned = c->gemm_cfg != -1;
    // members already built with one common tile (e.g. a batch whose member
    // set shrinks) keep it: no re-tuning, no rebuilt panels
    int common = -1;
    bool same = c->gemm_cfg == -1;
    for (ContractPlan* P : plans) {
      const GemmVariant& V = P->gemm.pk[prec];
      if (!V.built || (common >= 0 && V.cfg != common)) same = false;
      common = V.cfg;
    }
    if (same && common >= GEMM_64x64) {
      g->cfg = common;
      g->tuned = true;
    }
    G = g.get();
    c->group_cache.emplace(gkey, std::move(g));
  } else {
    G = it->second.get();
  }
  int rc = group_variants(*G, prec, G->cfg);
  if (rc) return rc;
  if (G->built_cfg != G->cfg) {
    if (cap) return fail(DPD_EINVAL, "grouped task list built during graph capture (internal)");
    std::vector<const GemmVariant*> Vs;
    for (ContractPlan* P : plans) Vs.push_back(&P->gemm.pk[prec]);
    if ((rc = group_tasks(Vs, prec, c->split_k != 0, G->gt))) return rc;
    G->built_cfg = G->cfg;
  }
  std::vector<CallPrep> R(n);
  size_t need = sizeof(double) * G->gt.partial_elems + 256;
  for (int i = 0; i < n; i++)
    if ((rc = call_setup(c, plans[i], o1[i], o2[i], nbatch[i], prec, R[i], &need))) return rc;
  ws_reset(c);
  if (cap && c->ws.bytes < ((need + 255) & ~size_t(255)))
    return fail(DPD_EINVAL, "workspace growth during graph capture (internal)");
  if ((rc = ws_reserve(c, need))) return rc;
  GroupTab gt;
  std::memset(&gt, 0, sizeof gt);
  int64_t flops = 0, bytes = 0;
  for (int i = 0; i < n; i++) {
    if ((rc = call_pre(c, R[i], o1[i], o2[i], o3[i], nbatch[i], cap ? nullptr : tag))) return rc;
    const void* side[2];
    if ((rc = pack_sides(c, plans[i]->gemm, plans[i]->gemm.pk[prec], prec, R[i].a, R[i].b, side))) return rc;
    gt.e[i] = {side[0], side[1], R[i].p3, R[i].beta, R[i].p3_f32 ? 1 : 0, 0};
    flops += plans[i]->gemm.flops;
    if (c->prof.on && tag) bytes += call_bytes(c, R[i], o1[i], o2[i], prec);
  }
  double* partial = nullptr;
  if (G->gt.partial_elems && !(partial = (double*)ws_take(c, sizeof(double) * G->gt.partial_elems)))
    return DPD_ENOMEM;
  GemmLaunch g;
  group_launch_args(G->gt, prec, G->cfg, fact, partial, g);
  ProfRec* pr = cap ? nullptr : prof_open(c, tag, flops, prec, G->cfg, 1, bytes, G->gt.nsplit > 0);
  if (pr) DPD_HIP_TRY(hipEventRecord(pr->a, c->stream));
  DPD_HIP_TRY(launch_gemm_pk_grouped(g, gt, c->stream));
  if (pr) DPD_HIP_TRY(hipEventRecord(pr->b, c->stream));
  for (int i = 0; i < n; i++)
    if ((rc = call_post(c, R[i], o3[i], nbatch[i], cap ? nullptr : tag))) return rc;
  if (!G->tuned && !cap) {
    G->tuned = true;
    const int before = G->cfg;
    if ((rc = tune_group(c, *G, prec, R, fact))) return rc;
    if (G->cfg != before) {
      if ((rc = group_variants(*G, prec, G->cfg))) return rc;
      std::vector<const GemmVariant*> Vs;
      for (ContractPlan* P : plans) Vs.push_back(&P->gemm.pk[prec]);
      if ((rc = group_tasks(Vs, prec, c->split_k != 0, G->gt))) return rc;
      G->built_cfg = G->cfg;
    }
  }
  return DPD_OK;
}

}  // namespace dpd
