// dpd_sortplan.hpp -- host planning of the tiled DPD permutation (k_sort_tiles).
//
// A sort between two layouts whose pairs are both stored in full (iop = 1 on
// both sides; libsrint/sstgenso.f90, itranspso.f, isymtrso.f on full pairs) is
// a set of dense 4-index permutations, one per irrep quadruple: element
// (x0,x1,x2,x3) of the quadruple (h0,h1,h2,h3) lies at
//   blk[xr] + offL[xl][h1] + nL[xl] * offR[xr][h3] + x0 + n0 x1 + nL[xl] (x2 + n2 x3)
// with xl = h0 x h1 and xr = h2 x h3 (the address fetch_addr computes for full
// pairs).  plan_sort turns each quadruple into one SortTask.
#pragma once
#include <algorithm>
#include <climits>
#include <vector>

#include "dpd_kernels.hpp"

namespace dpd {

// a plain single-term copy between full-pair layouts (the case k_sort_tiles
// covers); other gathers keep the generic kernels
inline bool sort_tiles_apply(const MapArgs& a) {
  if (a.nterm != 1 || a.term[0].coef != 1.0 || a.alpha != 1.0 || a.beta != 0.0 || a.nbatch != 1) return false;
  if (a.out.iop[0] != 1 || a.out.iop[1] != 1 || a.in.iop[0] != 1 || a.in.iop[1] != 1) return false;
  if (a.out.sym != a.in.sym || a.out.nirrep != a.in.nirrep || a.out.len != a.in.len) return false;
  if (a.out.len >= INT32_MAX) return false;  // the kernels keep offsets in int32
  int seen = 0;
  for (int i = 0; i < 4; i++) {
    const int p = a.term[0].perm[i];
    if (p < 0 || p > 3 || (seen & (1 << p))) return false;
    seen |= 1 << p;
  }
  return a.out.len > 0;
}

// tasks of a sort for which sort_tiles_apply holds; returns false if the
// layouts disagree (an input extent differs from its output dim) or the work
// does not fit the launch
inline bool plan_sort(const MapArgs& a, std::vector<SortTask>& tasks, int64_t* nitems, bool* direct) {
  const DevLayout& O = a.out;
  const DevLayout& I = a.in;
  const int8_t* perm = a.term[0].perm;
  int inv[4];  // output dim of each input dim
  for (int i = 0; i < 4; i++) inv[perm[i]] = i;
  const int j = inv[0], j2 = inv[1];
  tasks.clear();
  int64_t total = 0;
  for (int xr = 0; xr < O.nirrep; xr++) {
    const int xl = O.dp[xr][O.sym];
    if (O.nL[xl] == 0 || O.nR[xr] == 0) continue;
    for (int h1 = 0; h1 < O.nirrep; h1++) {
      const int h0 = O.dp[h1][xl];
      for (int h3 = 0; h3 < O.nirrep; h3++) {
        const int h2 = O.dp[h3][xr];
        const int g[4] = {h0, h1, h2, h3};
        int64_t n[4];
        bool empty = false;
        for (int i = 0; i < 4; i++) {
          n[i] = O.pop[i][g[i]];
          empty |= n[i] == 0;
        }
        if (empty) continue;
        const int64_t os[4] = {1, n[0], O.nL[xl], O.nL[xl] * n[2]};
        const int64_t obase = O.blk[xr] + O.offL[xl][h1] + O.nL[xl] * O.offR[xr][h3];
        int hin[4];
        for (int i = 0; i < 4; i++) hin[perm[i]] = g[i];
        const int ixl = I.dp[hin[0]][hin[1]], ixr = I.dp[hin[2]][hin[3]];
        if (I.dp[ixl][ixr] != I.sym) return false;
        for (int k = 0; k < 4; k++)
          if (I.pop[k][hin[k]] != n[inv[k]]) return false;
        const int64_t isd[4] = {1, I.pop[0][hin[0]], I.nL[ixl], I.nL[ixl] * I.pop[2][hin[2]]};
        const int64_t ibase = I.blk[ixr] + I.offL[ixl][hin[1]] + I.nL[ixl] * I.offR[ixr][hin[3]];
        int64_t is[4];  // input stride of each output dim
        for (int i = 0; i < 4; i++) is[i] = isd[perm[i]];
        // digit groups: A output-contiguous, B input-contiguous, C the rest
        int A[2], B[2], C[3], na = 0, nb = 0, nc = 0;
        bool used[4] = {false, false, false, false};
        A[na++] = 0;
        used[0] = true;
        if (j == 0) {
          if (perm[1] == 1) {  // output dims 0 and 1 are input dims 0 and 1
            A[na++] = 1;
            used[1] = true;
          }
        } else {
          B[nb++] = j;
          used[j] = true;
          // a short side takes the next dim of its own order when that stays contiguous
          const bool a1 = j != 1 && n[0] < 32;
          const bool b2 = j2 != 0 && !used[j2] && n[j] < 32;
          if (a1 && b2 && j2 == 1) {
            if (n[0] <= n[j]) A[na++] = 1;
            else B[nb++] = 1;
            used[1] = true;
          } else {
            if (a1) {
              A[na++] = 1;
              used[1] = true;
            }
            if (b2 && !used[j2]) {
              B[nb++] = j2;
              used[j2] = true;
            }
          }
        }
        for (int i = 0; i < 4; i++)
          if (!used[i]) C[nc++] = i;
        SortTask t{};
        t.in_base = ibase;
        t.out_base = obase;
        t.nA0 = (int32_t)n[A[0]];
        int64_t ea = n[A[0]];
        t.inA[0] = is[A[0]];
        if (na == 2) {
          ea *= n[A[1]];
          t.inA[1] = is[A[1]];
        }
        int64_t eb = 0;
        if (nb > 0) {
          eb = n[B[0]];
          t.nB0 = (int32_t)n[B[0]];
          t.outB[0] = os[B[0]];
          if (nb == 2) {
            eb *= n[B[1]];
            t.outB[1] = os[B[1]];
          }
        }
        for (int k = 0; k < 3; k++) {
          t.nC[k] = 1;
          if (k < nc) {
            t.nC[k] = (int32_t)n[C[k]];
            t.inC[k] = is[C[k]];
            t.outC[k] = os[C[k]];
          }
        }
        if (ea > 0x7fffffff || eb > 0x7fffffff) return false;
        t.nA = (int32_t)ea;
        t.nB = (int32_t)eb;
        // tiles of balanced size: A (and B) first, C digits 0 and 1 fill the rest
        auto split = [](int64_t ext, int64_t cap, int32_t* T, int32_t* cnt) {
          cap = std::max<int64_t>(1, cap);
          const int64_t c = (ext + cap - 1) / cap;
          *cnt = (int32_t)c;
          *T = (int32_t)((ext + c - 1) / c);
        };
        // (a transposing tile's LDS rows are padded by one at most; its row and
        // column offset tables, TC1 TC0 (TA + TB) entries, fit the same bound)
        int64_t row;  // tile elements per C index
        int64_t budget, rows = kSortLds;
        if (nb == 0) {
          split(ea, kSortCopy, &t.TA, &t.tA);
          t.TB = 1;
          t.tB = 1;
          row = t.TA;
          budget = kSortCopy;
          rows = kSortRows;
        } else {
          split(ea, 64, &t.TA, &t.tA);
          split(eb, kSortTileB / t.TA, &t.TB, &t.tB);
          row = (int64_t)t.TA * (t.TB + 1);
          budget = kSortLds;
        }
        split(t.nC[0], std::min(budget / row, rows), &t.TC0, &t.tC0);
        split(t.nC[1], std::min(budget / (row * t.TC0), rows / t.TC0), &t.TC1, &t.tC1);
        const int64_t items = (int64_t)t.tA * t.tB * t.tC0 * t.tC1 * t.nC[2];
        t.first = total;
        total += items;
        tasks.push_back(t);
      }
    }
  }
  if (total > 0x7fffffff) return false;
  *nitems = total;
  *direct = j == 0;
  return true;
}

// the task of every work item (k_sort_copy / k_sort_tiles read it)
inline void sort_item_map(const std::vector<SortTask>& tasks, int64_t nitems, int32_t* map) {
  const size_t nt = tasks.size();
  for (size_t k = 0; k < nt; k++) {
    const int64_t end = k + 1 < nt ? tasks[k + 1].first : nitems;
    for (int64_t i = tasks[k].first; i < end; i++) map[i] = (int32_t)k;
  }
}

}  // namespace dpd
