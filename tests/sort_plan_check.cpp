// CPU check of the tiled DPD permutation's host plan (csrc/dpd_sortplan.hpp):
// the index arithmetic of k_sort_copy / k_sort_tiles (csrc/dpd_kernels.hip),
// digits split by the same reciprocal multiplication, is replayed over every
// work item of plan_sort's tasks; every address must lie in the tensor, every
// LDS slot within kSortLds, every output element be written exactly once, and
// each from the input element that fetch_addr (the generic gather's address)
// names.  Random populations for 1, 2, 4 and 8 irreps, every symmetry and all
// 24 permutations; "full" adds blocks of up to 125 per irrep (several tiles).
// Prints "<cases> cases, <failures> failures".
#include <cstdio>
#include <random>
#include <vector>
#include <algorithm>
#include <climits>
#include "dpd_sortplan.hpp"
using namespace dpd;
namespace dpd { void set_error(const std::string&) {} int fail(int c, const char*, ...) { return c; } }

static int check(const Sym& S, const Desc& din, const int8_t perm[4], long* nitems_out) {
  Desc dout = din;
  for (int i = 0; i < 4; i++) for (int h = 0; h < 8; h++) dout.pop[i][h] = din.pop[perm[i]][h];
  MapArgs a; memset(&a, 0, sizeof a);
  a.out = make_layout(S, dout); a.in = make_layout(S, din);
  a.nterm = 1; for (int i = 0; i < 4; i++) a.term[0].perm[i] = perm[i];
  a.term[0].coef = 1.0; a.alpha = 1.0; a.beta = 0.0; a.nbatch = 1;
  if (!sort_tiles_apply(a)) { if (a.out.len == 0) return 0; printf("not applicable\n"); return 1; }
  std::vector<SortTask> T; int64_t nitems; bool direct;
  if (!plan_sort(a, T, &nitems, &direct)) { printf("plan failed\n"); return 1; }
  *nitems_out = nitems;
  const int64_t len = a.out.len;
  std::vector<int64_t> src(len, -1);
  struct FD { uint32_t d, m, l; void init(uint32_t dd){ d=dd; l=0; while ((1u<<l) < d) l++; m=(uint32_t)((((uint64_t)1<<32)*(((uint64_t)1<<l)-d))/d+1);} uint32_t div(uint32_t n) const { return (uint32_t)((((uint64_t)n*m)>>32) + n) >> l; } };
  for (const SortTask& t : T) if ((t.nB == 0) != direct) { printf("mixed modes\n"); return 1; }
  std::vector<int64_t> lds(kSortLds);
  std::vector<int32_t> item_task(nitems);
  sort_item_map(T, nitems, item_task.data());
  for (int64_t item = 0; item < nitems; item++) {
    const int k = item_task[item];
    if (k < 0 || k >= (int)T.size() || T[k].first > item) { printf("item map\n"); return 1; }
    const SortTask& st = T[k];
    int local = (int)(item - st.first);
    const int tb = local % st.tB; local /= st.tB;
    const int ta = local % st.tA; local /= st.tA;
    const int tc0 = local % st.tC0; local /= st.tC0;
    const int tc1 = local % st.tC1; const int c2 = local / st.tC1;
    if (c2 >= st.nC[2]) { printf("c2 out of range\n"); return 1; }
    const int a_lo = ta * st.TA, a_n = std::min(st.TA, st.nA - a_lo);
    const int b_lo = tb * st.TB, b_n = st.nB ? std::min(st.TB, st.nB - b_lo) : 1;
    const int c0_lo = tc0 * st.TC0, c1_lo = tc1 * st.TC1;
    const int c0_n = std::min(st.TC0, st.nC[0] - c0_lo), c1_n = std::min(st.TC1, st.nC[1] - c1_lo);
    if (a_n <= 0 || b_n <= 0 || c0_n <= 0 || c1_n <= 0) { printf("empty tile\n"); return 1; }
    const int64_t ib0 = st.in_base + (int64_t)c0_lo * st.inC[0] + (int64_t)c1_lo * st.inC[1] + (int64_t)c2 * st.inC[2];
    const int64_t ob0 = st.out_base + (int64_t)c0_lo * st.outC[0] + (int64_t)c1_lo * st.outC[1] + (int64_t)c2 * st.outC[2];
    auto put = [&](int64_t o, int64_t i) -> int {
      if (o < 0 || o >= len || i < 0 || i >= len) { printf("out of bounds o=%ld i=%ld len=%ld\n", (long)o, (long)i, (long)len); return 1; }
      if (src[o] != -1) { printf("output %ld written twice\n", (long)o); return 1; }
      src[o] = i; return 0;
    };
    if (direct) {
      const int nr = c0_n * c1_n, n = a_n * nr;
      if (n > kSortCopy || nr > kSortRows) { printf("copy tile too big\n"); return 1; }
      std::vector<int64_t> rin(nr), rout(nr);
      for (int r = 0; r < nr; r++) {
        const int i0 = r % c0_n, i1 = r / c0_n;
        rin[r] = (int64_t)i0 * st.inC[0] + (int64_t)i1 * st.inC[1];
        rout[r] = (int64_t)i0 * st.outC[0] + (int64_t)i1 * st.outC[1];
        if (rin[r] > INT32_MAX || rout[r] > INT32_MAX) { printf("offset exceeds int32\n"); return 1; }
      }
      FD fa; fa.init(a_n);
      for (int e = 0; e < n; e++) {
        const int r = (int)fa.div(e), ia = e - r * a_n;
        if (ia != e % a_n || r >= nr) { printf("digit error\n"); return 1; }
        if (put(ob0 + a_lo + rout[r] + ia, ib0 + a_lo + rin[r] + ia)) return 1;
      }
      continue;
    }
    const int pitch = b_n | 1, nc = c0_n * c1_n, n = a_n * b_n * nc;
    const int nrow = a_n * nc, ncol = b_n * nc;
    if (nc * a_n * pitch > kSortLds || nrow + ncol > kSortLds) { printf("LDS overflow\n"); return 1; }
    std::vector<int64_t> rin(nrow), cout(ncol);
    for (int r = 0; r < nrow; r++) {
      const int ia = r % a_n, q = r / a_n, i0 = q % c0_n, i1 = q / c0_n, A = a_lo + ia;
      const int a0 = A % st.nA0, a1 = A / st.nA0;
      rin[r] = (int64_t)i0 * st.inC[0] + (int64_t)i1 * st.inC[1] + (int64_t)a0 * st.inA[0] + (int64_t)a1 * st.inA[1];
      if (rin[r] > INT32_MAX) { printf("offset exceeds int32\n"); return 1; }
    }
    for (int q = 0; q < ncol; q++) {
      const int ib = q % b_n, r2 = q / b_n, i0 = r2 % c0_n, i1 = r2 / c0_n, B = b_lo + ib;
      const int b0 = B % st.nB0, b1 = B / st.nB0;
      cout[q] = (int64_t)i0 * st.outC[0] + (int64_t)i1 * st.outC[1] + (int64_t)b0 * st.outB[0] + (int64_t)b1 * st.outB[1];
      if (cout[q] > INT32_MAX) { printf("offset exceeds int32\n"); return 1; }
    }
    FD fa, fb; fa.init(a_n); fb.init(b_n);
    std::fill(lds.begin(), lds.end(), -7);
    for (int e = 0; e < n; e++) {
      const int r = (int)fb.div(e), ib = e - r * b_n;
      if (r >= nrow) { printf("row range\n"); return 1; }
      const int li = r * pitch + ib;
      if (li < 0 || li >= kSortLds) { printf("LDS index %d\n", li); return 1; }
      lds[li] = ib0 + b_lo + rin[r] + ib;
    }
    for (int e = 0; e < n; e++) {
      const int q = (int)fa.div(e), ia = e - q * a_n;
      const int r2 = (int)fb.div(q), ib = q - r2 * b_n;
      if (q >= ncol) { printf("column range\n"); return 1; }
      const int li = (r2 * a_n + ia) * pitch + ib;
      if (li < 0 || li >= kSortLds || lds[li] == -7) { printf("LDS read %d\n", li); return 1; }
      if (put(ob0 + a_lo + cout[q] + ia, lds[li])) return 1;
    }
  }
  // compare with the generic map
  for (int64_t o = 0; o < len; o++) {
    int h[4]; int64_t x[4]; decode_elem(a.out, o, h, x);
    int hi[4]; int64_t xi[4]; for (int i = 0; i < 4; i++) { hi[perm[i]] = h[i]; xi[perm[i]] = x[i]; }
    double sg; int64_t ad = fetch_addr(a.in, hi[0], hi[1], hi[2], hi[3], xi[0], xi[1], xi[2], xi[3], &sg);
    if (ad != src[o]) { printf("mismatch at %ld: %ld vs %ld\n", (long)o, (long)src[o], (long)ad); return 1; }
  }
  return 0;
}

int main(int argc, char** argv) {
  const bool full = argc > 1 && std::string(argv[1]) == "full";
  std::mt19937 rng(5);
  int bad = 0; long cases = 0;
  for (int nirrep : {1, 2, 4, 8}) {
    Sym S; int64_t dp64[64];
    for (int i = 0; i < 8; i++) for (int j = 0; j < 8; j++) dp64[i * 8 + j] = (i ^ j) + 1;
    S.set(nirrep, dp64);
    for (int trial = 0; trial < 6; trial++) {
      int64_t pops[4][8];
      int big = trial % 3;  // sizes up to 70 to exercise the tiling
      for (int k = 0; k < 4; k++) for (int h = 0; h < 8; h++) {
        int maxv = (big == 2 ? 72 : big == 1 ? 24 : 8) / nirrep;
        pops[k][h] = h < nirrep ? (int64_t)(rng() % (maxv + 1)) : 0;
      }
      // spaces: keep the 4 pops independent
      for (int sym = 1; sym <= nirrep; sym++) {
        Desc d; memset(&d, 0, sizeof d);
        for (int k = 0; k < 4; k++) for (int h = 0; h < 8; h++) d.pop[k][h] = pops[k][h];
        d.iop[0] = d.iop[1] = 1; d.sym = sym;
        int8_t perm[4] = {0, 1, 2, 3};
        do {
          long ni = 0; cases++;
          if (check(S, d, perm, &ni)) { printf("FAIL nirrep %d trial %d sym %d perm %d%d%d%d\n", nirrep, trial, sym, perm[0]+1, perm[1]+1, perm[2]+1, perm[3]+1); bad++; if (bad > 5) return 1; }
        } while (std::next_permutation(perm, perm + 4));
      }
    }
  }
  // large irrep blocks (several tiles along A and B)
  if (full) {
    Sym S; int64_t dp64[64];
    for (int i = 0; i < 8; i++) for (int j = 0; j < 8; j++) dp64[i * 8 + j] = (i ^ j) + 1;
    S.set(4, dp64);
    int64_t V[8] = {125, 30, 60, 100, 0, 0, 0, 0}, Oo[8] = {8, 3, 4, 6, 0, 0, 0, 0};
    const int64_t* sp[3][4] = {{V, Oo, V, Oo}, {V, V, Oo, Oo}, {Oo, V, V, Oo}};
    for (auto& q : sp) for (int sym = 1; sym <= 4; sym += 3) {
      Desc d; memset(&d, 0, sizeof d);
      for (int k = 0; k < 4; k++) for (int h = 0; h < 8; h++) d.pop[k][h] = q[k][h];
      d.iop[0] = d.iop[1] = 1; d.sym = sym;
      int8_t perm[4] = {0, 1, 2, 3};
      do { long ni = 0; cases++;
        if (check(S, d, perm, &ni)) { printf("FAIL large perm %d%d%d%d\n", perm[0]+1, perm[1]+1, perm[2]+1, perm[3]+1); bad++; }
      } while (std::next_permutation(perm, perm + 4));
    }
  }
  printf("%ld cases, %d failures\n", cases, bad);
  return bad != 0;
}
