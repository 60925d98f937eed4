// dpd_synth.cpp -- seeded synthetic spin-free DPD Hamiltonians (host).
//
// Stand-in for the reference's data path pyscf2cc -> getHF -> moio files
// (pyscf2cc/py2cc.py:388-530, getHF/main.f90:94-145): it produces the same
// spin-free DPD tensors the .SF programs load (CISdoo.SD.SF/allocate_mem.f90:
// 17-54) from a seeded model instead of PySCF:
//   * spatial orbitals per irrep; every spatial orbital gives an alpha and a
//     beta spin orbital; inside an irrep the occupied space lists alpha then
//     beta orbitals, likewise the virtual space (lib_cc/readmodule.f90:40-43);
//   * alpha occupations by the largest-remainder split of n_alpha over the
//     irreps (ties to the lower irrep), beta = alpha minus one electron in each
//     of the (n_alpha - n_beta) irreps with the most alpha electrons (ties to
//     the lower irrep): a high-spin open-shell reference;
//   * orbital energies per spin orbital: occupied U[-1.5,-0.3], virtual
//     U[0.2,3.0] Eh, ascending inside each (irrep, spin, occ/vir) group;
//     canonical Fock (f_ij, f_ab diagonal, f_ia = 0);
//   * spatial ERIs (PQ|RS) = 0.02 * N(0,1), 8-fold symmetric, zero unless
//     irrep(P)^irrep(Q)^irrep(R)^irrep(S) is totally symmetric, plus a Coulomb
//     diagonal (PP|QQ) += U[0.3,0.8];
//   * antisymmetrised spin-orbital integrals <pq||rs> = <pq|rs> - <pq|sr> with
//     <pq|rs> = (PR|QS) d(sp,sr) d(sq,ss); `iabj` is stored as (b,i,a,j) =
//     <ia||bj> (pyscf2cc/py2cc.py:503-516).
// Random numbers come from a counter-based hash of (seed, stream, key), so any
// element can be generated independently and in parallel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "dpd_common.hpp"

namespace dpd {

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t key) {
  return mix64(mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ull)) ^ key);
}
// uniform in [0,1) with 53 random bits
inline double uni(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
inline double normal(uint64_t seed, uint64_t stream, uint64_t key) {
  double u1 = 1.0 - uni(hash3(seed, stream, 2 * key));  // (0,1]
  double u2 = uni(hash3(seed, stream, 2 * key + 1));
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

enum : uint64_t { S_ERI = 1, S_COUL = 2, S_EORB = 3 };

struct Model {
  Sym S;
  dpd_synth_spec spec;
  int64_t soff[9];  // spatial offset per irrep
  // energies per (irrep, spin, occ=0/vir=1) group in ascending order
  std::vector<double> eps[8][2][2];

  double eri(int64_t P, int64_t Q, int64_t R, int64_t Sx, int hP, int hQ, int hR, int hS) const {
    // (PQ|RS) with 8-fold symmetry
    double v = 0.0;
    if (S.dp[S.dp[hP][hQ]][S.dp[hR][hS]] == 0) {
      uint64_t pq = P >= Q ? (uint64_t)(P * (P + 1) / 2 + Q) : (uint64_t)(Q * (Q + 1) / 2 + P);
      uint64_t rs = R >= Sx ? (uint64_t)(R * (R + 1) / 2 + Sx) : (uint64_t)(Sx * (Sx + 1) / 2 + R);
      uint64_t a = std::max(pq, rs), b = std::min(pq, rs);
      uint64_t key = a * (a + 1) / 2 + b;
      v = 0.02 * normal(spec.seed, S_ERI, key);
    }
    if (P == Q && R == Sx) {
      uint64_t a = (uint64_t)std::max(P, R), b = (uint64_t)std::min(P, R);
      v += 0.3 + 0.5 * uni(hash3(spec.seed, S_COUL, a * (a + 1) / 2 + b));
    }
    return v;
  }
};

// a spin orbital of the occupied (space=0) or virtual (space=1) DPD space
struct SO {
  int h;        // irrep (0-based)
  int spin;     // 0 alpha, 1 beta
  int64_t P;    // global spatial index
};

inline SO so_of(const Model& m, int space, int h, int64_t l) {
  const dpd_synth_spec& s = m.spec;
  SO o;
  o.h = h;
  if (space == 0) {
    if (l < s.popa[h]) {
      o.spin = 0;
      o.P = m.soff[h] + l;
    } else {
      o.spin = 1;
      o.P = m.soff[h] + (l - s.popa[h]);
    }
  } else {
    if (l < s.vrta[h]) {
      o.spin = 0;
      o.P = m.soff[h] + s.popa[h] + l;
    } else {
      o.spin = 1;
      o.P = m.soff[h] + s.popb[h] + (l - s.vrta[h]);
    }
  }
  return o;
}

// <pq|rs> = (PR|QS) d(sp,sr) d(sq,ss)
inline double phys(const Model& m, const SO& p, const SO& q, const SO& r, const SO& s) {
  if (p.spin != r.spin || q.spin != s.spin) return 0.0;
  return m.eri(p.P, r.P, q.P, s.P, p.h, r.h, q.h, s.h);
}
inline double anti(const Model& m, const SO& p, const SO& q, const SO& r, const SO& s) {
  return phys(m, p, q, r, s) - phys(m, p, q, s, r);
}

int build_model(const dpd_synth_spec* in, Model& m) {
  m.spec = *in;
  dpd_synth_spec& s = m.spec;
  int n = (int)s.nirrep;
  if (n != 1 && n != 2 && n != 4 && n != 8) return fail(DPD_EINVAL, "synth: nirrep must be 1, 2, 4 or 8");
  int64_t dp[64];
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) dp[i * 8 + j] = (i < n && j < n) ? ((i ^ j) + 1) : 1;
  m.S.set(n, dp);
  int64_t N = 0;
  for (int h = 0; h < n; h++) {
    if (s.nso[h] < 0) return fail(DPD_EINVAL, "synth: negative orbital count");
    N += s.nso[h];
  }
  for (int h = n; h < 8; h++) s.nso[h] = 0;
  if (s.nalpha < s.nbeta || s.nbeta < 0 || s.nalpha > N)
    return fail(DPD_EINVAL, "synth: need 0 <= n_beta <= n_alpha <= n_orbitals");
  // largest remainder for alpha (exact integer arithmetic)
  int64_t rem[8];
  int64_t tot = 0;
  for (int h = 0; h < 8; h++) {
    s.popa[h] = (h < n && N > 0) ? s.nalpha * s.nso[h] / N : 0;
    rem[h] = (h < n && N > 0) ? s.nalpha * s.nso[h] % N : -1;
    tot += s.popa[h];
  }
  while (tot < s.nalpha) {
    int best = -1;
    for (int h = 0; h < n; h++)
      if (s.popa[h] < s.nso[h] && (best < 0 || rem[h] > rem[best])) best = h;
    if (best < 0) return fail(DPD_EINVAL, "synth: cannot place alpha electrons");
    s.popa[best]++;
    rem[best] = -1;
    tot++;
  }
  for (int h = 0; h < 8; h++) s.popb[h] = s.popa[h];
  bool used[8] = {false, false, false, false, false, false, false, false};
  for (int64_t k = 0; k < s.nalpha - s.nbeta; k++) {
    int best = -1;
    for (int h = 0; h < n; h++)
      if (!used[h] && s.popa[h] > 0 && (best < 0 || s.popa[h] > s.popa[best])) best = h;
    if (best < 0) return fail(DPD_EINVAL, "synth: n_alpha - n_beta exceeds the occupied irreps");
    used[best] = true;
    s.popb[best]--;
  }
  int64_t off = 0;
  for (int h = 0; h < 8; h++) {
    s.vrta[h] = s.nso[h] - s.popa[h];
    s.vrtb[h] = s.nso[h] - s.popb[h];
    s.popo[h] = s.popa[h] + s.popb[h];
    s.popv[h] = s.vrta[h] + s.vrtb[h];
    m.soff[h] = off;
    off += s.nso[h];
  }
  m.soff[8] = off;
  // orbital energies
  for (int h = 0; h < n; h++)
    for (int sp = 0; sp < 2; sp++)
      for (int ov = 0; ov < 2; ov++) {
        int64_t cnt = ov == 0 ? (sp == 0 ? s.popa[h] : s.popb[h]) : (sp == 0 ? s.vrta[h] : s.vrtb[h]);
        std::vector<double>& e = m.eps[h][sp][ov];
        e.resize(cnt);
        for (int64_t k = 0; k < cnt; k++) {
          uint64_t key = (((uint64_t)h * 2 + sp) * 2 + ov) * 1000003ull + (uint64_t)k;
          double u = uni(hash3(s.seed, S_EORB, key));
          e[k] = ov == 0 ? -1.5 + 1.2 * u : 0.2 + 2.8 * u;
        }
        std::sort(e.begin(), e.end());
      }
  return DPD_OK;
}

inline double eps_of(const Model& m, int space, int h, int64_t l) {
  const dpd_synth_spec& s = m.spec;
  if (space == 0) return l < s.popa[h] ? m.eps[h][0][0][l] : m.eps[h][1][0][l - s.popa[h]];
  return l < s.vrta[h] ? m.eps[h][0][1][l] : m.eps[h][1][1][l - s.vrta[h]];
}

template <class F>
void parallel_for(int64_t n, F f) {
  unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
  if (n < 100000) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      int64_t b = n * t / nt, e = n * (t + 1) / nt;
      for (int64_t i = b; i < e; i++) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace

}  // namespace dpd

using namespace dpd;

extern "C" int dpd_synth_layout(dpd_synth_spec* spec) {
  if (!spec) return fail(DPD_EINVAL, "null spec");
  Model m;
  int rc = build_model(spec, m);
  if (rc) return rc;
  *spec = m.spec;
  return DPD_OK;
}

extern "C" int dpd_synth_fill(const dpd_synth_spec* spec, const char* which, int64_t iop, double* out,
                              int64_t len) {
  if (!spec || !which || !out) return fail(DPD_EINVAL, "null argument");
  Model m;
  int rc = build_model(spec, m);
  if (rc) return rc;
  const dpd_synth_spec& s = m.spec;
  const int n = (int)s.nirrep;
  std::string w(which);
  if (w == "eorb") {
    int64_t k = 0, need = 0;
    for (int h = 0; h < n; h++) need += s.popo[h] + s.popv[h];
    if (len < need) return fail(DPD_EINVAL, "eorb buffer too small (%lld < %lld)", (long long)len, (long long)need);
    for (int h = 0; h < n; h++)
      for (int64_t l = 0; l < s.popo[h]; l++) out[k++] = eps_of(m, 0, h, l);
    for (int h = 0; h < n; h++)
      for (int64_t l = 0; l < s.popv[h]; l++) out[k++] = eps_of(m, 1, h, l);
    return DPD_OK;
  }
  // spaces per dim: 0 occupied, 1 virtual, 2 dummy (1 element in irrep 1)
  int sp[4];
  int iop12 = 1, iop34 = 1;
  if (w == "fij") { sp[0] = 0; sp[1] = 0; sp[2] = 2; sp[3] = 2; }
  else if (w == "fab") { sp[0] = 1; sp[1] = 1; sp[2] = 2; sp[3] = 2; }
  else if (w == "fia") { sp[0] = 1; sp[1] = 0; sp[2] = 2; sp[3] = 2; }
  else if (w == "ijkl") { sp[0] = 0; sp[1] = 0; sp[2] = 0; sp[3] = 0; iop12 = iop34 = (int)iop; }
  else if (w == "ijka") { sp[0] = 0; sp[1] = 0; sp[2] = 0; sp[3] = 1; iop12 = (int)iop; }
  else if (w == "abij") { sp[0] = 1; sp[1] = 1; sp[2] = 0; sp[3] = 0; iop12 = iop34 = (int)iop; }
  else if (w == "iabj") { sp[0] = 1; sp[1] = 0; sp[2] = 1; sp[3] = 0; }
  else if (w == "abci") { sp[0] = 1; sp[1] = 1; sp[2] = 1; sp[3] = 0; iop12 = (int)iop; }
  else if (w == "abcd") { sp[0] = 1; sp[1] = 1; sp[2] = 1; sp[3] = 1; iop12 = iop34 = (int)iop; }
  else return fail(DPD_EINVAL, "synth: unknown tensor '%s'", which);
  if (iop < 0 || iop > 2) return fail(DPD_EINVAL, "synth: iop must be 0, 1 or 2");
  Desc d;
  for (int k = 0; k < 4; k++)
    for (int h = 0; h < 8; h++) d.pop[k][h] = sp[k] == 0 ? s.popo[h] : sp[k] == 1 ? s.popv[h] : (h == 0);
  d.iop[0] = iop12;
  d.iop[1] = iop34;
  d.sym = 1;
  DevLayout L = make_layout(m.S, d);
  if (len != L.len) return fail(DPD_EINVAL, "synth: '%s' has %lld elements, buffer holds %lld", which,
                                (long long)L.len, (long long)len);
  parallel_for(L.len, [&](int64_t o) {
    int h[4];
    int64_t x[4];
    decode_elem(L, o, h, x);
    double v = 0.0;
    if (w == "fij" || w == "fab") {
      v = (x[0] == x[1] && h[0] == h[1]) ? eps_of(m, sp[0], h[0], x[0]) : 0.0;
    } else if (w == "fia") {
      v = 0.0;
    } else {
      SO a = so_of(m, sp[0], h[0], x[0]), b = so_of(m, sp[1], h[1], x[1]);
      SO c = so_of(m, sp[2], h[2], x[2]), e = so_of(m, sp[3], h[3], x[3]);
      if (w == "iabj") v = anti(m, b, c, a, e);  // element (b,i,a,j) = <ia||bj>
      else v = anti(m, a, b, c, e);
    }
    out[o] = v;
  });
  return DPD_OK;
}

// dense spatial ERI access for tests: (PQ|RS) over global spatial indices
extern "C" int dpd_synth_eri(const dpd_synth_spec* spec, int64_t n, const int64_t* idx4, double* out) {
  if (!spec || !idx4 || !out) return fail(DPD_EINVAL, "null argument");
  Model m;
  int rc = build_model(spec, m);
  if (rc) return rc;
  auto irr = [&](int64_t P) {
    int h = 0;
    while (h < 7 && m.soff[h + 1] <= P) h++;
    return h;
  };
  for (int64_t k = 0; k < n; k++) {
    const int64_t* q = idx4 + 4 * k;
    out[k] = m.eri(q[0], q[1], q[2], q[3], irr(q[0]), irr(q[1]), irr(q[2]), irr(q[3]));
  }
  return DPD_OK;
}
