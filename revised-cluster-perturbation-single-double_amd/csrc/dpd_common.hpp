// dpd_common.hpp -- DPD symmetry metadata and storage layout shared by the host
// planner and the HIP kernels of the MI355X engine.
//
// Storage follows the reference's DPD convention (libsrint/build_dpd_tables.f90:
// 5-45): a tensor I(pq,rs) of irrep `sym` is stored as one column-major block per
// right-pair irrep XR (ascending), of shape nL(XL) x nR(XR) with XL = XR (x) sym;
// inside a pair block the sub-blocks follow the irrep of the SECOND index
// ascending, first index fastest.  Pair packing iop: 0 = p<q (a diagonal-irrep
// sub-block is the strict upper triangle, column by column: offset q(q-1)/2+p),
// 1 = full, 2 = p<=q (offset q(q+1)/2+p).  Counts follow libsrint/irpdso.f:1-51,
// offsets libsrint/isymoffso.f:1-53.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>

#include "dpd_hip.h"

namespace dpd {

using i64 = int64_t;

// Symmetry context: replaces the reference's global mod_symm (nirrep, dirprd).
struct Sym {
  int nirrep = 1;
  int dp[8][8];  // 0-based: dp[a][b] = irrep (0-based) of a (x) b

  void set(int n, const int64_t* dirprd64) {
    nirrep = n;
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) dp[i][j] = (int)dirprd64[i * 8 + j] - 1;
  }
  // 1-based wrapper used by the metadata functions
  int D(int a, int b) const { return dp[a - 1][b - 1] + 1; }

  // pair count of irrep `sym` (1-based), libsrint/irpdso.f:1-51
  i64 irpdso(int sym, const i64* popp, const i64* popq, int iop) const {
    i64 n = 0;
    if (iop == -1) return popp[sym - 1];
    for (int r = 1; r <= nirrep; r++) {
      int l = D(r, sym);
      if (iop == 1) {
        n += popp[l - 1] * popq[r - 1];
      } else if (iop == 0 || iop == 2) {
        if (l > r) continue;
        if (l < r) n += popp[l - 1] * popq[r - 1];
        else n += popp[l - 1] * (popq[r - 1] + (iop == 0 ? -1 : 1)) / 2;
      }
    }
    return n;
  }
  // libsrint/idsymsoc.f:1-29
  i64 idsymsoc(int x, const i64* pp, const i64* pq, int iop12, const i64* pr, const i64* ps,
               int iop34) const {
    i64 n = 0;
    for (int r = 1; r <= nirrep; r++) n += irpdso(D(r, x), pp, pq, iop12) * irpdso(r, pr, ps, iop34);
    return n;
  }
  // libsrint/isymoffso.f:1-53 (1-based)
  i64 isymoffso(int hq1, int x, const i64* pp, const i64* pq, int iop) const {
    i64 off = 1;
    for (int q = 1; q < hq1; q++) {
      int p = D(q, x);
      if (iop == 1) off += pp[p - 1] * pq[q - 1];
      else if (p < q) off += pp[p - 1] * pq[q - 1];
      else if (p == q) off += pp[p - 1] * (pp[q - 1] + (iop == 0 ? -1 : 1)) / 2;
    }
    return off;
  }
};

// POD layout descriptor passed by value to kernels.
struct DevLayout {
  int32_t nirrep, sym;  // sym 0-based
  int32_t iop[2];
  int32_t pop[4][8];
  int8_t dp[8][8];
  int64_t blk[8];   // start of the block with right irrep x (0-based)
  int64_t nL[8], nR[8];
  int32_t offL[8][8];  // [X][hq] sub-block start inside the left pair block
  int32_t offR[8][8];
  int64_t len;
};

// 4-index descriptor on the host (2-index tensors are padded with 1-sized
// dummy dims in irrep 1, exactly as mod_contract.f90:177,196,216 pads them).
struct Desc {
  i64 pop[4][8];
  int iop[2];
  int sym;  // 1-based

  static Desc from(const dpd_tensor* t, int rank) {
    Desc d;
    std::memcpy(d.pop, t->pop, sizeof d.pop);
    d.iop[0] = (int)t->iop[0];
    d.iop[1] = (int)t->iop[1];
    d.sym = (int)t->sym;
    if (rank == 2) {
      for (int k = 2; k < 4; k++)
        for (int h = 0; h < 8; h++) d.pop[k][h] = (h == 0);
      d.iop[1] = 1;
    }
    return d;
  }
};

inline DevLayout make_layout(const Sym& S, const Desc& d) {
  DevLayout L;
  std::memset(&L, 0, sizeof L);
  L.nirrep = S.nirrep;
  L.sym = d.sym - 1;
  L.iop[0] = d.iop[0];
  L.iop[1] = d.iop[1];
  for (int k = 0; k < 4; k++)
    for (int h = 0; h < 8; h++) L.pop[k][h] = (int32_t)d.pop[k][h];
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 8; b++) L.dp[a][b] = (int8_t)S.dp[a][b];
  for (int x = 1; x <= S.nirrep; x++) {
    L.nL[x - 1] = S.irpdso(x, d.pop[0], d.pop[1], d.iop[0]);
    L.nR[x - 1] = S.irpdso(x, d.pop[2], d.pop[3], d.iop[1]);
    for (int h = 1; h <= S.nirrep; h++) {
      L.offL[x - 1][h - 1] = (int32_t)(S.isymoffso(h, x, d.pop[0], d.pop[1], d.iop[0]) - 1);
      L.offR[x - 1][h - 1] = (int32_t)(S.isymoffso(h, x, d.pop[2], d.pop[3], d.iop[1]) - 1);
    }
  }
  i64 off = 0;
  for (int xr = 0; xr < S.nirrep; xr++) {
    L.blk[xr] = off;
    off += L.nL[S.dp[xr][L.sym]] * L.nR[xr];
  }
  for (int xr = S.nirrep; xr < 8; xr++) L.blk[xr] = off;
  L.len = off;
  return L;
}

// position of (hp,p,hq,q) (0-based irreps/indices) in its pair block, -1 if the
// element is not stored under the packing (host+device)
__host__ __device__ inline int64_t pair_pos(const int32_t* popp, int iop, int32_t off_sub, int hp,
                                            int64_t p, int hq, int64_t q) {
  if (iop == 1) return off_sub + p + q * popp[hp];
  if (hp > hq) return -1;
  if (hp < hq) return off_sub + p + q * popp[hp];
  if (iop == 0) return (p < q) ? off_sub + q * (q - 1) / 2 + p : -1;
  return (p <= q) ? off_sub + q * (q + 1) / 2 + p : -1;
}

// address of element (h[4] 0-based irreps, x[4] 0-based indices) of a tensor that
// may hold packed pairs; returns -1 for a structural zero and sets the mirror
// sign (iop=0 pairs are antisymmetric, iop=2 symmetric).
__host__ __device__ inline int64_t fetch_addr(const DevLayout& L, int h0, int h1, int h2, int h3,
                                              int64_t x0, int64_t x1, int64_t x2, int64_t x3,
                                              double* sign) {
  double sg = 1.0;
  if (L.iop[0] != 1) {
    if (L.iop[0] == 0 && h0 == h1 && x0 == x1) return -1;
    if (h0 > h1 || (h0 == h1 && x0 > x1)) {
      int th = h0; h0 = h1; h1 = th;
      int64_t tx = x0; x0 = x1; x1 = tx;
      if (L.iop[0] == 0) sg = -sg;
    }
  }
  if (L.iop[1] != 1) {
    if (L.iop[1] == 0 && h2 == h3 && x2 == x3) return -1;
    if (h2 > h3 || (h2 == h3 && x2 > x3)) {
      int th = h2; h2 = h3; h3 = th;
      int64_t tx = x2; x2 = x3; x3 = tx;
      if (L.iop[1] == 0) sg = -sg;
    }
  }
  if (x0 < 0 || x1 < 0 || x2 < 0 || x3 < 0 || x0 >= L.pop[0][h0] || x1 >= L.pop[1][h1] ||
      x2 >= L.pop[2][h2] || x3 >= L.pop[3][h3])
    return -1;
  int xl = L.dp[h0][h1], xr = L.dp[h2][h3];
  if (L.dp[xl][xr] != L.sym) return -1;
  int64_t pl = pair_pos(L.pop[0], L.iop[0], L.offL[xl][h1], h0, x0, h1, x1);
  int64_t pr = pair_pos(L.pop[2], L.iop[1], L.offR[xr][h3], h2, x2, h3, x3);
  if (pl < 0 || pr < 0) return -1;
  *sign = sg;
  return L.blk[xr] + pl + L.nL[xl] * pr;
}

// invert a position inside a pair block of irrep X (0-based) into (hp,p,hq,q)
__host__ __device__ inline void decode_pair(const int32_t* popp, const int32_t* popq, int iop,
                                            const int8_t (*dp)[8], const int32_t* off, int nirrep,
                                            int X, int64_t l, int* hp, int64_t* p, int* hq,
                                            int64_t* q) {
  int h = nirrep - 1;
  while (h > 0 && off[h] > l) h--;
  int a = dp[h][X];
  int64_t l2 = l - off[h];
  *hq = h;
  *hp = a;
  if (iop == 1 || a < h) {
    const int64_t np = popp[a];
    if (l2 < 0x7fffffff) {  // 32-bit division where it fits
      const int32_t l32 = (int32_t)l2, n32 = (int32_t)np;
      *p = l32 % n32;
      *q = l32 / n32;
    } else {
      *p = l2 % np;
      *q = l2 / np;
    }
  } else if (iop == 0) {
    int64_t qq = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)l2)) * 0.5);
    while (qq * (qq - 1) / 2 > l2) qq--;
    while ((qq + 1) * qq / 2 <= l2) qq++;
    *q = qq;
    *p = l2 - qq * (qq - 1) / 2;
  } else {
    int64_t qq = (int64_t)((-1.0 + sqrt(1.0 + 8.0 * (double)l2)) * 0.5);
    if (qq < 0) qq = 0;
    while (qq * (qq + 1) / 2 > l2) qq--;
    while ((qq + 1) * (qq + 2) / 2 <= l2) qq++;
    *q = qq;
    *p = l2 - qq * (qq + 1) / 2;
  }
  (void)popq;
}

// linear storage index -> irreps and indices (0-based)
__host__ __device__ inline void decode_elem(const DevLayout& L, int64_t o, int* h, int64_t* x) {
  int xr = L.nirrep - 1;
  while (xr > 0 && L.blk[xr] > o) xr--;
  int xl = L.dp[xr][L.sym];
  int64_t nl = L.nL[xl];
  int64_t loc = o - L.blk[xr];
  int64_t l = loc % nl, r = loc / nl;
  decode_pair(L.pop[0], L.pop[1], L.iop[0], L.dp, L.offL[xl], L.nirrep, xl, l, &h[0], &x[0], &h[1],
              &x[1]);
  decode_pair(L.pop[2], L.pop[3], L.iop[1], L.dp, L.offR[xr], L.nirrep, xr, r, &h[2], &x[2], &h[3],
              &x[3]);
}

// error reporting (thread-local message)
void set_error(const std::string& msg);
int fail(int code, const char* fmt, ...);

}  // namespace dpd
