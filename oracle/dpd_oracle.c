/*
 * dpd_oracle.c -- CPU oracle for the DPD contraction engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the reference's
 * CPU path for the DPD tensor-contraction engine so that the HIP product can be
 * checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load the shared object built from it.
 *
 * Restated reference files (paths relative to the reference's src/):
 *   libsrint/irpdso.f:1-51            -> orc_irpdso
 *   libsrint/idsymsoc.f:1-29          -> orc_idsymsoc
 *   libsrint/isymoffso.f:1-53         -> orc_isymoffso
 *   lib_cc/getidx.f90, putidx.f90     -> orc_getidx / orc_putidx
 *   libsrint/iexpso.f90:24-120 (+exp_pq_d :221-282, exp_pq2_d) -> orc_iexpso
 *   libsrint/icompso.f:1-93 (+sqsymso.f ...)                    -> orc_icompso
 *   libsrint/isymtrso.f:1-120 (+symtr1so.f, symtr3so.f)         -> orc_isymtrso
 *   libsrint/itranspso.f:1-53                                   -> orc_itranspso
 *   libsrint/iassymso.f:1-142                                   -> orc_iassymso
 *   libsrint/sstgenso.f90:22-206                                -> orc_sstgenso
 *   libsrint/vv2v_dpd_444.f90:55-232 (dispatcher)               -> orc_vv2v_dpd_444
 *   libsrint/vv2v_dpd_444_d.f90:120-235 (fast path)             -> orc_vv2v_dpd_444_fast
 *   libsrint/vv2v_dpd_444_d.f90:256-517 (fallback),
 *   libsrint/vv2v_dpd_424_d.f90, vv2v_dpd_442_d.f90             -> orc_contract_general
 *   libsrint/vv2v_dpd_424.f90:21-200, vv2v_dpd_442.f90:22-200   -> orc_vv2v_dpd_424/442
 *   lib_cc/mod_contract.f90:10-455                              -> orc_vv2v_444/424/442
 *   lib_cc/Eorb_diff.f90, denomT.f90, putDiagFock.f90           -> elementwise helpers
 *
 * The fallback/424/442 engines are restated by one routine, orc_contract_general:
 * the reference builds a task list over irrep tuples, gathers A(M,K) and B(K,N)
 * per task, calls DGEMM and scatters v3 = fact2*v3 + fact1*C
 * (vv2v_dpd_444_d.f90:256-517); the routine below does the same over every irrep
 * assignment of the contraction's modes, which covers all of the 424/442 routes.
 */
#include "dpd_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* global state: lib_mod/mod_symm.f (nirrep, dirprd)                          */
/* ------------------------------------------------------------------------- */
static i64 g_nirrep = 1;
static i64 g_dirprd[8][8];
static char g_err[512];

typedef void (*dgemm_fn)(const char *, const char *, const i64 *, const i64 *, const i64 *,
                         const double *, const double *, const i64 *, const double *,
                         const i64 *, const double *, double *, const i64 *, size_t, size_t);
static dgemm_fn g_dgemm = NULL;
typedef void (*blas_threads_fn)(int);
static blas_threads_fn g_blas_threads = NULL;
static int g_threads = 0; /* 0: OpenMP default */

static int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *orc_last_error(void) { return g_err; }

void orc_set_symmetry(i64 nirrep, const i64 *dirprd64) {
  g_nirrep = nirrep;
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) g_dirprd[i][j] = dirprd64[i * 8 + j];
}

void orc_set_dgemm(void *fn) { g_dgemm = (dgemm_fn)fn; }
void orc_set_blas_threads_fn(void *fn) { g_blas_threads = (blas_threads_fn)fn; }
static int max_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}
int orc_has_blas(void) { return g_dgemm != NULL; }
void orc_set_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}

static inline i64 DP(i64 a, i64 b) { return g_dirprd[a - 1][b - 1]; }

/* ------------------------------------------------------------------------- */
/* metadata                                                                   */
/* ------------------------------------------------------------------------- */
/* irpdso.f:1-51 */
i64 orc_irpdso(i64 sym, const i64 *popp, const i64 *popq, i64 iop) {
  i64 n = 0;
  if (iop == 1) {
    for (i64 r = 1; r <= g_nirrep; r++) n += popp[DP(r, sym) - 1] * popq[r - 1];
  } else if (iop == 0) {
    for (i64 r = 1; r <= g_nirrep; r++) {
      i64 l = DP(r, sym);
      if (l > r) continue;
      if (l < r) n += popp[l - 1] * popq[r - 1];
      else n += popp[l - 1] * (popq[r - 1] - 1) / 2;
    }
  } else if (iop == -1) {
    n = popp[sym - 1];
  } else if (iop == 2) {
    for (i64 r = 1; r <= g_nirrep; r++) {
      i64 l = DP(r, sym);
      if (l > r) continue;
      if (l < r) n += popp[l - 1] * popq[r - 1];
      else n += popp[l - 1] * (popq[r - 1] + 1) / 2;
    }
  }
  return n;
}

/* idsymsoc.f:1-29 */
i64 orc_idsymsoc(i64 irrepx, const i64 *popp, const i64 *popq, i64 ioppq, const i64 *popr,
                 const i64 *pops, i64 ioprs) {
  i64 n = 0;
  for (i64 r = 1; r <= g_nirrep; r++)
    n += orc_irpdso(DP(r, irrepx), popp, popq, ioppq) * orc_irpdso(r, popr, pops, ioprs);
  return n;
}

/* isymoffso.f:1-53 -- 1-based start of sub-block irrep_q inside pair irrep `irrep` */
i64 orc_isymoffso(i64 irrepq1, i64 irrep, const i64 *popp, const i64 *popq, i64 iop) {
  i64 off = 1;
  for (i64 q = 1; q < irrepq1; q++) {
    i64 p = DP(q, irrep);
    if (iop == 1) {
      off += popp[p - 1] * popq[q - 1];
    } else if (iop == 0) {
      if (p < q) off += popp[p - 1] * popq[q - 1];
      else if (p == q) off += popp[p - 1] * (popp[q - 1] - 1) / 2;
    } else if (iop == 2) {
      if (p < q) off += popp[p - 1] * popq[q - 1];
      else if (p == q) off += popp[p - 1] * (popp[q - 1] + 1) / 2;
    }
  }
  return off;
}

i64 orc_desc_len(const orc_desc *d) {
  return orc_idsymsoc(d->sym, d->pop[0], d->pop[1], d->iop[0], d->pop[2], d->pop[3], d->iop[1]);
}

/* ------------------------------------------------------------------------- */
/* layout tables + element addressing                                         */
/* ------------------------------------------------------------------------- */
typedef struct {
  const orc_desc *d;
  i64 nL[9], nR[9];     /* pair counts per pair irrep */
  i64 blk[9];           /* block offset per right irrep (0-based) */
  i64 offL[9][9];       /* [pair irrep][irrep of 2nd index] sub-block offset (0-based) */
  i64 offR[9][9];
} orc_lay;

static void lay_init(orc_lay *L, const orc_desc *d) {
  L->d = d;
  for (i64 x = 1; x <= g_nirrep; x++) {
    L->nL[x] = orc_irpdso(x, d->pop[0], d->pop[1], d->iop[0]);
    L->nR[x] = orc_irpdso(x, d->pop[2], d->pop[3], d->iop[1]);
    for (i64 h = 1; h <= g_nirrep; h++) {
      L->offL[x][h] = orc_isymoffso(h, x, d->pop[0], d->pop[1], d->iop[0]) - 1;
      L->offR[x][h] = orc_isymoffso(h, x, d->pop[2], d->pop[3], d->iop[1]) - 1;
    }
  }
  i64 off = 0;
  for (i64 xr = 1; xr <= g_nirrep; xr++) {
    L->blk[xr] = off;
    off += L->nL[DP(xr, d->sym)] * L->nR[xr];
  }
}

/* position of (hp,p,hq,q) (0-based p,q) inside its pair block; -1 if not stored */
static i64 pair_pos(const i64 *popp, i64 iop, i64 off_sub, i64 hp, i64 p, i64 hq, i64 q) {
  if (iop == 1) return off_sub + p + q * popp[hp - 1];
  if (hp > hq) return -1;
  if (hp < hq) return off_sub + p + q * popp[hp - 1];
  if (iop == 0) {
    if (p >= q) return -1;
    return off_sub + q * (q - 1) / 2 + p;
  }
  if (p > q) return -1; /* iop == 2 */
  return off_sub + q * (q + 1) / 2 + p;
}

/* address of element (h[4] 1-based irreps, x[4] 0-based indices); -1 if not stored */
static i64 lay_addr(const orc_lay *L, const i64 *h, const i64 *x) {
  const orc_desc *d = L->d;
  i64 xl = DP(h[0], h[1]), xr = DP(h[2], h[3]);
  if (DP(xl, xr) != d->sym) return -1;
  i64 pl = pair_pos(d->pop[0], d->iop[0], L->offL[xl][h[1]], h[0], x[0], h[1], x[1]);
  i64 pr = pair_pos(d->pop[2], d->iop[1], L->offR[xr][h[3]], h[2], x[2], h[3], x[3]);
  if (pl < 0 || pr < 0) return -1;
  return L->blk[xr] + pl + L->nL[xl] * pr;
}

/* iterate every stored element of a tensor in storage order */
typedef void (*elem_cb)(void *u, i64 addr, const i64 *h, const i64 *x);

static void pair_iter_count(const i64 *popp, const i64 *popq, i64 iop, i64 X, i64 hq, i64 *hp,
                            i64 *n) {
  *hp = DP(hq, X);
  if (iop != 1 && *hp > hq) { *n = 0; return; }
  *n = popp[*hp - 1] * popq[hq - 1];
}

static void for_each_elem(const orc_lay *L, elem_cb cb, void *u) {
  const orc_desc *d = L->d;
  i64 h[4], x[4];
  for (i64 xr = 1; xr <= g_nirrep; xr++) {
    i64 xl = DP(xr, d->sym);
    if (L->nL[xl] == 0 || L->nR[xr] == 0) continue;
    for (i64 hs = 1; hs <= g_nirrep; hs++) {
      i64 hr, nrs;
      pair_iter_count(d->pop[2], d->pop[3], d->iop[1], xr, hs, &hr, &nrs);
      if (nrs == 0) continue;
      for (i64 s = 0; s < d->pop[3][hs - 1]; s++)
        for (i64 r = 0; r < d->pop[2][hr - 1]; r++) {
          if (d->iop[1] != 1 && hr == hs && (d->iop[1] == 0 ? r >= s : r > s)) continue;
          for (i64 hq = 1; hq <= g_nirrep; hq++) {
            i64 hp, npq;
            pair_iter_count(d->pop[0], d->pop[1], d->iop[0], xl, hq, &hp, &npq);
            if (npq == 0) continue;
            for (i64 q = 0; q < d->pop[1][hq - 1]; q++)
              for (i64 p = 0; p < d->pop[0][hp - 1]; p++) {
                if (d->iop[0] != 1 && hp == hq && (d->iop[0] == 0 ? p >= q : p > q)) continue;
                h[0] = hp; h[1] = hq; h[2] = hr; h[3] = hs;
                x[0] = p; x[1] = q; x[2] = r; x[3] = s;
                cb(u, lay_addr(L, h, x), h, x);
              }
          }
        }
    }
  }
}

/* getidx.f90:6-209 (1-based linear index -> irreps and 1-based indices) */
typedef struct { i64 target; i64 *info; int found; } getidx_u;
static void getidx_cb(void *uu, i64 addr, const i64 *h, const i64 *x) {
  getidx_u *u = (getidx_u *)uu;
  if (u->found || addr + 1 != u->target) return;
  for (int i = 0; i < 4; i++) { u->info[i] = h[i]; u->info[4 + i] = x[i] + 1; }
  u->found = 1;
}
int orc_getidx(i64 idx, i64 *info8, const orc_desc *d) {
  orc_lay L;
  lay_init(&L, d);
  getidx_u u = {idx, info8, 0};
  for_each_elem(&L, getidx_cb, &u);
  if (!u.found) return set_err(-1, "getidx: index %lld out of range", (long long)idx);
  return 0;
}
i64 orc_putidx(const i64 *info8, const orc_desc *d) {
  orc_lay L;
  lay_init(&L, d);
  i64 h[4], x[4];
  for (int i = 0; i < 4; i++) { h[i] = info8[i]; x[i] = info8[4 + i] - 1; }
  i64 a = lay_addr(&L, h, x);
  return a < 0 ? 0 : a + 1;
}

/* ------------------------------------------------------------------------- */
/* sort kernels                                                               */
/* ------------------------------------------------------------------------- */
/* fetch element (h,x) of a tensor that may be packed; applies the mirror sign of
 * iop=0 pairs (antisymmetric), iop=2 pairs (symmetric) and the zero diagonal. */
static double fetch_packed(const orc_lay *L, const double *v, const i64 *h0, const i64 *x0) {
  const orc_desc *d = L->d;
  i64 h[4] = {h0[0], h0[1], h0[2], h0[3]}, x[4] = {x0[0], x0[1], x0[2], x0[3]};
  double sign = 1.0;
  for (int pr = 0; pr < 2; pr++) {
    i64 iop = d->iop[pr];
    int a = 2 * pr, b = 2 * pr + 1;
    if (iop == 1) continue;
    int swap = (h[a] > h[b]) || (h[a] == h[b] && x[a] > x[b]);
    if (iop == 0 && h[a] == h[b] && x[a] == x[b]) return 0.0;
    if (swap) {
      i64 t = h[a]; h[a] = h[b]; h[b] = t;
      t = x[a]; x[a] = x[b]; x[b] = t;
      if (iop == 0) sign = -sign;
    }
  }
  i64 ad = lay_addr(L, h, x);
  if (ad < 0) return 0.0;
  return sign * v[ad];
}

typedef struct {
  const orc_lay *Lin;
  const double *in;
  double *out;
  int perm[4]; /* out dim i <- in dim perm[i] */
  int mode;    /* 0: plain gather with packing signs */
} sort_u;

static void sort_cb(void *uu, i64 addr, const i64 *h, const i64 *x) {
  sort_u *u = (sort_u *)uu;
  i64 hi[4], xi[4];
  for (int i = 0; i < 4; i++) { hi[u->perm[i]] = h[i]; xi[u->perm[i]] = x[i]; }
  u->out[addr] = fetch_packed(u->Lin, u->in, hi, xi);
}

static int pops_equal(const i64 *a, const i64 *b) {
  for (int i = 0; i < 8; i++) if (a[i] != b[i]) return 0;
  return 1;
}

static void make_desc(orc_desc *d, const i64 *p1, const i64 *p2, const i64 *p3, const i64 *p4,
                      i64 iop12, i64 iop34, i64 sym) {
  memcpy(d->pop[0], p1, 8 * sizeof(i64));
  memcpy(d->pop[1], p2, 8 * sizeof(i64));
  memcpy(d->pop[2], p3, 8 * sizeof(i64));
  memcpy(d->pop[3], p4, 8 * sizeof(i64));
  d->iop[0] = iop12; d->iop[1] = iop34; d->sym = sym;
}

/* decode a linear storage index into irreps/indices (inverse of lay_addr) */
static void pair_decode(const i64 *popp, i64 iop, const i64 (*off)[9], i64 X, i64 l, i64 *hp, i64 *p, i64 *hq,
                        i64 *q) {
  i64 h = g_nirrep;
  while (h > 1 && off[X][h] > l) h--;
  i64 a = DP(h, X);
  i64 l2 = l - off[X][h];
  *hq = h;
  *hp = a;
  if (iop == 1 || a < h) {
    *p = l2 % popp[a - 1];
    *q = l2 / popp[a - 1];
  } else if (iop == 0) {
    i64 qq = 1;
    while ((qq + 1) * qq / 2 <= l2) qq++;
    *q = qq;
    *p = l2 - qq * (qq - 1) / 2;
  } else {
    i64 qq = 0;
    while ((qq + 1) * (qq + 2) / 2 <= l2) qq++;
    *q = qq;
    *p = l2 - qq * (qq + 1) / 2;
  }
}

static void decode_elem(const orc_lay *L, i64 o, i64 *h, i64 *x) {
  const orc_desc *d = L->d;
  i64 xr = g_nirrep;
  while (xr > 1 && L->blk[xr] > o) xr--;
  i64 xl = DP(xr, d->sym);
  i64 loc = o - L->blk[xr];
  i64 l = loc % L->nL[xl], r = loc / L->nL[xl];
  pair_decode(d->pop[0], d->iop[0], (const i64(*)[9])L->offL, xl, l, &h[0], &x[0], &h[1], &x[1]);
  pair_decode(d->pop[2], d->iop[1], (const i64(*)[9])L->offR, xr, r, &h[2], &x[2], &h[3], &x[3]);
}

static void run_sort(const orc_desc *din, const double *in, const orc_desc *dout, double *out,
                     const int *perm) {
  orc_lay Li, Lo;
  lay_init(&Li, din);
  lay_init(&Lo, dout);
  i64 n = orc_desc_len(dout);
#pragma omp parallel for schedule(static) num_threads(max_threads())
  for (i64 o = 0; o < n; o++) {
    i64 h[4], x[4], hi[4], xi[4];
    decode_elem(&Lo, o, h, x);
    for (int i = 0; i < 4; i++) { hi[perm[i]] = h[i]; xi[perm[i]] = x[i]; }
    out[o] = fetch_packed(&Li, in, hi, xi);
  }
}

/* iexpso.f90:24-120: expand packed pairs (iop 0/2) of one or both sides to iop=1 */
int orc_iexpso(const double *in, double *out, const i64 *popp, const i64 *popq, i64 ioppq,
               i64 iexppq, const i64 *popr, const i64 *pops, i64 ioprs, i64 iexprs, i64 sym) {
  if (ioppq < 0 || ioppq > 2) return set_err(-1, "fatal error: ioppq");
  if (ioprs < 0 || ioprs > 2) return set_err(-1, "fatal error: ioprs");
  if (iexppq == 1 && ioppq != 1 && !pops_equal(popp, popq))
    return set_err(-1, "fatal error: popp /= popq");
  if (iexprs == 1 && ioprs != 1 && !pops_equal(popr, pops))
    return set_err(-1, "fatal error: popr /= pops");
  int ok = (iexppq == 1 && ioppq == 0 && iexprs == 0) || (ioprs == 0 && iexprs == 1 && iexppq == 0) ||
           (iexppq == 1 && ioppq == 2 && iexprs == 0) || (ioprs == 2 && iexprs == 1 && iexppq == 0) ||
           (iexppq == 1 && ioppq != 1 && iexprs == 1 && ioprs != 1);
  if (!ok) return set_err(-1, "fatal error: iexpso must expand exactly one side, or both packed sides");
  orc_desc din, dout;
  make_desc(&din, popp, popq, popr, pops, ioppq, ioprs, sym);
  make_desc(&dout, popp, popq, popr, pops, iexppq ? 1 : ioppq, iexprs ? 1 : ioprs, sym);
  int perm[4] = {0, 1, 2, 3};
  run_sort(&din, in, &dout, out, perm);
  return 0;
}

/* icompso.f:1-93: compress one full side to p<q (0) or p<=q (2) */
int orc_icompso(const double *in, double *out, const i64 *popp, const i64 *popq, i64 ioppq,
                i64 icpq, const i64 *popr, const i64 *pops, i64 ioprs, i64 icrs, i64 sym) {
  if (ioppq < 0 || ioppq > 2) return set_err(-1, "error of ioppq in icompso");
  if (ioprs < 0 || ioprs > 2) return set_err(-1, "error of ioprs in icompso");
  int ok = (icpq == 0 && icrs == 1) || (icpq == 1 && icrs == 0) || (icpq == 2 && icrs == 1) ||
           (icpq == 1 && icrs == 2);
  if (!ok) return set_err(-1, "attention!!!, error in icompso");
  if ((icpq != 1 && ioppq != 1) || (icrs != 1 && ioprs != 1))
    return set_err(-1, "icompso: the side being compressed must be full (iop=1)");
  orc_desc din, dout;
  make_desc(&din, popp, popq, popr, pops, ioppq, ioprs, sym);
  make_desc(&dout, popp, popq, popr, pops, icpq == 1 ? ioppq : icpq, icrs == 1 ? ioprs : icrs, sym);
  int perm[4] = {0, 1, 2, 3};
  run_sort(&din, in, &dout, out, perm);
  return 0;
}

/* isymtrso.f:1-120: swap inside pq (icpq=1) and/or rs (icrs=1); no sign */
int orc_isymtrso(const double *in, double *out, const i64 *popp, const i64 *popq, i64 ioppq,
                 i64 icpq, const i64 *popr, const i64 *pops, i64 ioprs, i64 icrs, i64 sym) {
  if (ioppq < 0 || ioppq > 2) return set_err(-1, "error of ioppq in isymtrso");
  if (ioprs < 0 || ioprs > 2) return set_err(-1, "error of ioprs in isymtrso");
  if ((icpq == 1 && ioppq != 1) || (icrs == 1 && ioprs != 1))
    return set_err(-1, "isymtrso: a swapped pair must be full (iop=1)");
  orc_desc din, dout;
  make_desc(&din, popp, popq, popr, pops, ioppq, ioprs, sym);
  make_desc(&dout, icpq == 1 ? popq : popp, icpq == 1 ? popp : popq, icrs == 1 ? pops : popr,
            icrs == 1 ? popr : pops, ioppq, ioprs, sym);
  int perm[4] = {icpq == 1 ? 1 : 0, icpq == 1 ? 0 : 1, icrs == 1 ? 3 : 2, icrs == 1 ? 2 : 3};
  run_sort(&din, in, &dout, out, perm);
  return 0;
}

/* itranspso.f:1-53: I(pq,rs) -> I(rs,pq) */
int orc_itranspso(const double *in, double *out, const i64 *popp, const i64 *popq, i64 ioppq,
                  const i64 *popr, const i64 *pops, i64 ioprs, i64 sym) {
  if (ioppq < 0 || ioppq > 2 || ioprs < 0 || ioprs > 2) return set_err(-1, "error of iop in itranspso");
  orc_desc din, dout;
  make_desc(&din, popp, popq, popr, pops, ioppq, ioprs, sym);
  make_desc(&dout, popr, pops, popp, popq, ioprs, ioppq, sym);
  int perm[4] = {2, 3, 0, 1};
  run_sort(&din, in, &dout, out, perm);
  return 0;
}

/* iassymso.f:1-142 (assym2so.f, assym3so.f, dasymrhfso.f, dsymrhfso.f):
 * out = in -/+ in with one or both pairs swapped, compressed where the output
 * pair is packed:
 *   iop 0: (p<q,r,s) -> (p<q,r<s), swap rs     iop 1: (p,q,r,s) -> (p,q,r<s), swap rs
 *   iop 2: (p,q,r,s) -> (p,q,r,s), swap both   iop 3: as 2 with +
 *   iop 4: (p,q,r<s) -> (p<q,r<s), swap pq     iop 5: (p,q,r,s) -> (p<q,r,s), swap pq */
int orc_iassymso(const double *in, double *out, i64 sym, const i64 *popp, const i64 *popq,
                 const i64 *popr, const i64 *pops, i64 iop) {
  static const int io[6][4] = {{0, 1, 0, 0}, {1, 1, 1, 0}, {1, 1, 1, 1},
                               {1, 1, 1, 1}, {1, 0, 0, 0}, {1, 1, 0, 1}};
  if (iop < 0 || iop > 5) return set_err(-1, "error of iop in iassymso");
  int spq = (iop >= 2), srs = (iop <= 3);
  if (spq && !pops_equal(popp, popq)) return set_err(-1, "iassymso: popp /= popq");
  if (srs && !pops_equal(popr, pops)) return set_err(-1, "iassymso: popr /= pops");
  orc_desc din, dout;
  make_desc(&din, popp, popq, popr, pops, io[iop][0], io[iop][1], sym);
  make_desc(&dout, popp, popq, popr, pops, io[iop][2], io[iop][3], sym);
  orc_lay Li, Lo;
  lay_init(&Li, &din);
  lay_init(&Lo, &dout);
  const double sgn = iop == 3 ? 1.0 : -1.0;
  i64 n = orc_desc_len(&dout);
#pragma omp parallel for schedule(static) num_threads(max_threads())
  for (i64 o = 0; o < n; o++) {
    i64 h[4], x[4], h2[4], x2[4];
    decode_elem(&Lo, o, h, x);
    for (int i = 0; i < 4; i++) { h2[i] = h[i]; x2[i] = x[i]; }
    if (spq) { h2[0] = h[1]; x2[0] = x[1]; h2[1] = h[0]; x2[1] = x[0]; }
    if (srs) { h2[2] = h[3]; x2[2] = x[3]; h2[3] = h[2]; x2[3] = x[2]; }
    out[o] = fetch_packed(&Li, in, h, x) + sgn * fetch_packed(&Li, in, h2, x2);
  }
  return 0;
}

/* sstgenso.f90:22-206: new index i takes old index type(i) (e.g. '3142') */
int orc_sstgenso(const double *win, double *wout, const i64 *pop1, const i64 *pop2,
                 const i64 *pop3, const i64 *pop4, i64 irrepx, const char *type4) {
  int perm[4];
  const i64 *pin[4] = {pop1, pop2, pop3, pop4};
  for (int i = 0; i < 4; i++) {
    perm[i] = type4[i] - '1';
    if (perm[i] < 0 || perm[i] > 3)
      return set_err(-1, "fatal error in sstgend_opt: invalid type received: %.4s", type4);
  }
  orc_desc din, dout;
  make_desc(&din, pop1, pop2, pop3, pop4, 1, 1, irrepx);
  make_desc(&dout, pin[perm[0]], pin[perm[1]], pin[perm[2]], pin[perm[3]], 1, 1, irrepx);
  run_sort(&din, win, &dout, wout, perm);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* GEMM: column-major C(MxN) = alpha*op(A)*op(B) + beta*C                     */
/* ------------------------------------------------------------------------- */
static void gemm(char ta, char tb, i64 m, i64 n, i64 k, double alpha, const double *a, i64 lda,
                 const double *b, i64 ldb, double beta, double *c, i64 ldc) {
  if (m <= 0 || n <= 0) return;
  if (g_dgemm) {
    g_dgemm(&ta, &tb, &m, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc, 1, 1);
    return;
  }
#pragma omp parallel for schedule(static) if (!omp_in_parallel())
  for (i64 j = 0; j < n; j++) {
    for (i64 i = 0; i < m; i++) {
      double s = 0.0;
      for (i64 l = 0; l < k; l++) {
        double av = (ta == 'N') ? a[i + l * lda] : a[l + i * lda];
        double bv = (tb == 'N') ? b[l + j * ldb] : b[j + l * ldb];
        s += av * bv;
      }
      c[i + j * ldc] = (beta == 0.0) ? alpha * s : alpha * s + beta * c[i + j * ldc];
    }
  }
}

/* ------------------------------------------------------------------------- */
/* general contraction engine                                                 */
/* ------------------------------------------------------------------------- */
#define MAXMODE 12
static const i64 ZORB[8] = {1, 0, 0, 0, 0, 0, 0, 0};

typedef struct {
  char lab[2];  /* label(s); lab[1]==0 for a single index */
  int is_pair;  /* intact packed pair */
  i64 size[9];  /* per irrep */
  int cls;      /* 0:M(v1&v3) 1:N(v2&v3) 2:K(v1&v2) 3:only v1 4:only v2 5:only v3 */
} mode_t_;

typedef struct {
  orc_desc d;        /* 4-index (padded) descriptor */
  char lab[4];
  int mode_of_dim[4]; /* mode id per dim (pair mode: both dims point at it) */
  int pair_mode[2];   /* mode id if the pair slot is an intact packed pair else -1 */
  orc_lay L;
} tens_t;

static int find_mode(mode_t_ *m, int nm, const char *lab, int is_pair) {
  for (int i = 0; i < nm; i++) {
    if (m[i].is_pair != is_pair) continue;
    if (m[i].lab[0] == lab[0] && (!is_pair || m[i].lab[1] == lab[1])) return i;
  }
  return -1;
}

/* pad a rank-2 tensor to rank 4 with 1-sized dummy dims in irrep 1 (the
 * mod_contract.f90:177,196,216 '=','+','-' padding). */
static void pad_tensor(tens_t *t, const orc_desc *d, const char *type, int rank, char dum0,
                       char dum1) {
  t->d = *d;
  if (rank == 2) {
    memcpy(t->d.pop[2], ZORB, sizeof ZORB);
    memcpy(t->d.pop[3], ZORB, sizeof ZORB);
    t->d.iop[1] = 1;
    t->lab[0] = type[0]; t->lab[1] = type[1]; t->lab[2] = dum0; t->lab[3] = dum1;
  } else {
    memcpy(t->lab, type, 4);
  }
}

/* per-tensor affine addressing inside one irrep assignment */
typedef struct {
  i64 base;
  i64 stride[MAXMODE];
  int valid;
} aff_t;

static void tensor_affine(const tens_t *t, const mode_t_ *modes, const i64 *mirr, aff_t *af) {
  const orc_desc *d = &t->d;
  memset(af, 0, sizeof *af);
  i64 h[4];
  i64 xl, xr;
  /* irreps of the four dims (pair modes: a pair irrep, dims unknown) */
  for (int s = 0; s < 2; s++) {
    int pm = t->pair_mode[s];
    if (pm >= 0) {
      if (s == 0) xl = mirr[pm]; else xr = mirr[pm];
    } else {
      h[2 * s] = mirr[t->mode_of_dim[2 * s]];
      h[2 * s + 1] = mirr[t->mode_of_dim[2 * s + 1]];
      if (s == 0) xl = DP(h[0], h[1]); else xr = DP(h[2], h[3]);
    }
  }
  if (DP(xl, xr) != d->sym) return;
  const orc_lay *L = &t->L;
  i64 nl = L->nL[xl];
  af->base = L->blk[xr];
  /* left pair */
  if (t->pair_mode[0] >= 0) {
    af->stride[t->pair_mode[0]] += 1;
  } else {
    af->base += L->offL[xl][h[1]];
    af->stride[t->mode_of_dim[0]] += 1;
    af->stride[t->mode_of_dim[1]] += d->pop[0][h[0] - 1];
  }
  if (t->pair_mode[1] >= 0) {
    af->stride[t->pair_mode[1]] += nl;
  } else {
    af->base += nl * L->offR[xr][h[3]];
    af->stride[t->mode_of_dim[2]] += nl;
    af->stride[t->mode_of_dim[3]] += nl * d->pop[2][h[2] - 1];
  }
  (void)modes;
  af->valid = 1;
}

static int sym_ok(const tens_t *t, const i64 *mirr) {
  i64 x = 1;
  for (int s = 0; s < 2; s++) {
    if (t->pair_mode[s] >= 0) x = DP(x, mirr[t->pair_mode[s]]);
    else x = DP(DP(x, mirr[t->mode_of_dim[2 * s]]), mirr[t->mode_of_dim[2 * s + 1]]);
  }
  return x == t->d.sym;
}

/* mixed-radix enumeration helper */
static i64 prod_sizes(const mode_t_ *modes, const int *ids, int n, const i64 *mirr) {
  i64 p = 1;
  for (int i = 0; i < n; i++) p *= modes[ids[i]].size[mirr[ids[i]]];
  return p;
}

static void offsets_for(const mode_t_ *modes, const int *ids, int n, const i64 *mirr,
                        const i64 *stride, i64 count, i64 *out) {
  /* out[c] = sum_i stride[id_i] * digit_i(c), first id fastest */
  i64 dig[MAXMODE] = {0};
  for (i64 c = 0; c < count; c++) {
    i64 o = 0;
    for (int i = 0; i < n; i++) o += stride[ids[i]] * dig[i];
    out[c] = o;
    for (int i = 0; i < n; i++) {
      if (++dig[i] < modes[ids[i]].size[mirr[ids[i]]]) break;
      dig[i] = 0;
    }
  }
}

int orc_contract_general(const double *v1, const orc_desc *d1, const char *type1, int rank1,
                         const double *v2, const orc_desc *d2, const char *type2, int rank2,
                         double *v3, const orc_desc *d3, const char *type3, int rank3,
                         double fact1, double fact2) {
  tens_t T[3];
  pad_tensor(&T[0], d1, type1, rank1, '\x01', '\x02');
  pad_tensor(&T[1], d2, type2, rank2, '\x03', '\x04');
  pad_tensor(&T[2], d3, type3, rank3, '\x05', '\x06');
  mode_t_ modes[MAXMODE];
  int nm = 0;
  /* build modes */
  for (int t = 0; t < 3; t++) {
    for (int s = 0; s < 2; s++) {
      T[t].pair_mode[s] = -1;
      if (T[t].d.iop[s] != 1) {
        char lab[2] = {T[t].lab[2 * s], T[t].lab[2 * s + 1]};
        int id = find_mode(modes, nm, lab, 1);
        if (id < 0) {
          if (nm >= MAXMODE) return set_err(-2, "too many modes");
          id = nm++;
          memset(&modes[id], 0, sizeof modes[id]);
          modes[id].lab[0] = lab[0]; modes[id].lab[1] = lab[1]; modes[id].is_pair = 1;
          for (i64 x = 1; x <= g_nirrep; x++)
            modes[id].size[x] = orc_irpdso(x, T[t].d.pop[2 * s], T[t].d.pop[2 * s + 1], T[t].d.iop[s]);
        } else {
          for (i64 x = 1; x <= g_nirrep; x++)
            if (modes[id].size[x] != orc_irpdso(x, T[t].d.pop[2 * s], T[t].d.pop[2 * s + 1], T[t].d.iop[s]))
              return set_err(-3, "packed pair '%c%c' has mismatched populations or iop", lab[0], lab[1]);
        }
        T[t].pair_mode[s] = id;
        T[t].mode_of_dim[2 * s] = T[t].mode_of_dim[2 * s + 1] = id;
      } else {
        for (int k = 0; k < 2; k++) {
          char lab[2] = {T[t].lab[2 * s + k], 0};
          int id = find_mode(modes, nm, lab, 0);
          if (id < 0) {
            if (nm >= MAXMODE) return set_err(-2, "too many modes");
            id = nm++;
            memset(&modes[id], 0, sizeof modes[id]);
            modes[id].lab[0] = lab[0];
            for (i64 x = 1; x <= g_nirrep; x++) modes[id].size[x] = T[t].d.pop[2 * s + k][x - 1];
          } else {
            for (i64 x = 1; x <= g_nirrep; x++)
              if (modes[id].size[x] != T[t].d.pop[2 * s + k][x - 1])
                return set_err(-3, "contracted/external index '%c' has mismatched populations", lab[0]);
          }
          T[t].mode_of_dim[2 * s + k] = id;
        }
      }
    }
  }
  /* a label must not appear both inside a packed pair and as a single index */
  for (int i = 0; i < nm; i++)
    if (modes[i].is_pair)
      for (int j = 0; j < nm; j++)
        if (!modes[j].is_pair && (modes[j].lab[0] == modes[i].lab[0] || modes[j].lab[0] == modes[i].lab[1]))
          return set_err(-4, "IOP barrier: packed pair '%c%c' is split or reordered; expand it to iop=1 first",
                         modes[i].lab[0], modes[i].lab[1]);
  /* classify */
  for (int i = 0; i < nm; i++) {
    int in[3] = {0, 0, 0};
    for (int t = 0; t < 3; t++)
      for (int dd = 0; dd < 4; dd++) if (T[t].mode_of_dim[dd] == i) in[t] = 1;
    if (in[0] && in[2] && !in[1]) modes[i].cls = 0;
    else if (in[1] && in[2] && !in[0]) modes[i].cls = 1;
    else if (in[0] && in[1] && !in[2]) modes[i].cls = 2;
    else if (in[0] && !in[1] && !in[2]) modes[i].cls = 3;
    else if (in[1] && !in[0] && !in[2]) modes[i].cls = 4;
    else if (in[2] && !in[0] && !in[1]) modes[i].cls = 5;
    else return set_err(-5, "index '%c' appears in all three tensors", modes[i].lab[0]);
    if (modes[i].cls >= 3) {
      i64 tot = 0;
      for (i64 x = 1; x <= g_nirrep; x++) tot += modes[i].size[x];
      if (tot != 1) return set_err(-6, "index '%c' appears in one tensor only and is not a 1-sized dummy",
                                   modes[i].lab[0]);
    }
  }
  for (int t = 0; t < 3; t++) lay_init(&T[t].L, &T[t].d);

  int idsM[MAXMODE], idsN[MAXMODE], idsK[MAXMODE], ids3[MAXMODE];
  int nM = 0, nN = 0, nK = 0, n3 = 0;
  for (int i = 0; i < nm; i++) {
    if (modes[i].cls == 0 || modes[i].cls == 5) idsM[nM++] = i;
    else if (modes[i].cls == 1) idsN[nN++] = i;
    else idsK[nK++] = i; /* K plus v1-only / v2-only dummies */
    if (modes[i].cls == 0 || modes[i].cls == 1 || modes[i].cls == 5) ids3[n3++] = i;
  }
  (void)ids3;

  /* enumerate output tasks: irrep assignment of M and N modes
   * (the task list of vv2v_dpd_444_d.f90:299-318) */
  int ext[MAXMODE], next = 0;
  for (int i = 0; i < nM; i++) ext[next++] = idsM[i];
  for (int i = 0; i < nN; i++) ext[next++] = idsN[i];
  i64 ecount = 1;
  for (int i = 0; i < next; i++) ecount *= g_nirrep;
  i64 kcount = 1;
  for (int i = 0; i < nK; i++) kcount *= g_nirrep;
  i64 *tasks = (i64 *)malloc(sizeof(i64) * (size_t)(ecount > 0 ? ecount : 1));
  i64 ntask = 0;
  for (i64 e = 0; e < ecount; e++) {
    i64 mirr[MAXMODE] = {0}, r = e;
    int skip = 0;
    for (int i = 0; i < next; i++) {
      mirr[ext[i]] = 1 + r % g_nirrep;
      r /= g_nirrep;
      if (modes[ext[i]].size[mirr[ext[i]]] == 0) skip = 1;
    }
    for (int i = 0; i < nm; i++) if (mirr[i] == 0) mirr[i] = 1;
    if (skip || !sym_ok(&T[2], mirr)) continue;
    tasks[ntask++] = e;
  }
  /* many tasks: one task per thread with single-threaded BLAS
   * (vv2v_dpd_444_d.f90:322-363); few tasks: threaded BLAS inside each */
  int nthr = max_threads();
  int par = ntask >= 2 * nthr;
  if (par && g_blas_threads) g_blas_threads(1);
#pragma omp parallel for schedule(dynamic, 1) num_threads(par ? nthr : 1)
  for (i64 t = 0; t < ntask; t++) {
    i64 mirr[MAXMODE] = {0}, r = tasks[t];
    for (int i = 0; i < next; i++) {
      mirr[ext[i]] = 1 + r % g_nirrep;
      r /= g_nirrep;
    }
    for (int i = 0; i < nm; i++) if (mirr[i] == 0) mirr[i] = 1;
    i64 M = prod_sizes(modes, idsM, nM, mirr), N = prod_sizes(modes, idsN, nN, mirr);
    aff_t a3;
    tensor_affine(&T[2], modes, mirr, &a3);
    if (!a3.valid || M == 0 || N == 0) continue;
    i64 *o3m = (i64 *)malloc(sizeof(i64) * M), *o3n = (i64 *)malloc(sizeof(i64) * N);
    offsets_for(modes, idsM, nM, mirr, a3.stride, M, o3m);
    offsets_for(modes, idsN, nN, mirr, a3.stride, N, o3n);
    double *C = (double *)calloc((size_t)(M * N), sizeof(double));
    for (i64 kk = 0; kk < kcount; kk++) {
      i64 rr = kk;
      int skk = 0;
      for (int i = 0; i < nK; i++) {
        mirr[idsK[i]] = 1 + rr % g_nirrep;
        rr /= g_nirrep;
        if (modes[idsK[i]].size[mirr[idsK[i]]] == 0) skk = 1;
      }
      if (skk) continue;
      if (!sym_ok(&T[0], mirr) || !sym_ok(&T[1], mirr)) continue;
      i64 K = prod_sizes(modes, idsK, nK, mirr);
      if (K == 0) continue;
      aff_t a1, a2;
      tensor_affine(&T[0], modes, mirr, &a1);
      tensor_affine(&T[1], modes, mirr, &a2);
      if (!a1.valid || !a2.valid) continue;
      i64 *o1m = (i64 *)malloc(sizeof(i64) * M), *o1k = (i64 *)malloc(sizeof(i64) * K);
      i64 *o2k = (i64 *)malloc(sizeof(i64) * K), *o2n = (i64 *)malloc(sizeof(i64) * N);
      offsets_for(modes, idsM, nM, mirr, a1.stride, M, o1m);
      offsets_for(modes, idsK, nK, mirr, a1.stride, K, o1k);
      offsets_for(modes, idsK, nK, mirr, a2.stride, K, o2k);
      offsets_for(modes, idsN, nN, mirr, a2.stride, N, o2n);
      double *A = (double *)malloc(sizeof(double) * (size_t)(M * K));
      double *B = (double *)malloc(sizeof(double) * (size_t)(K * N));
#pragma omp parallel for schedule(static) if (!par) num_threads(nthr)
      for (i64 k = 0; k < K; k++)
        for (i64 m = 0; m < M; m++) A[m + M * k] = v1[a1.base + o1m[m] + o1k[k]];
#pragma omp parallel for schedule(static) if (!par) num_threads(nthr)
      for (i64 n = 0; n < N; n++)
        for (i64 k = 0; k < K; k++) B[k + K * n] = v2[a2.base + o2k[k] + o2n[n]];
      gemm('N', 'N', M, N, K, 1.0, A, M, B, K, 1.0, C, M);
      free(A); free(B); free(o1m); free(o1k); free(o2k); free(o2n);
    }
    /* scatter: v3 = fact2*v3 + fact1*C  (vv2v_dpd_444_d.f90:486-512) */
    for (i64 n = 0; n < N; n++)
      for (i64 m = 0; m < M; m++) {
        i64 ad = a3.base + o3m[m] + o3n[n];
        if (fact2 == 0.0) v3[ad] = fact1 * C[m + M * n];
        else v3[ad] = fact2 * v3[ad] + fact1 * C[m + M * n];
      }
    free(C); free(o3m); free(o3n);
  }
  if (par && g_blas_threads) g_blas_threads(nthr);
  free(tasks);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* 444 dispatcher + fast path                                                 */
/* ------------------------------------------------------------------------- */
/* label mapper of vv2v_dpd_444.f90:88-133: outputs -> 1,2,5,6; contracted -> 3,4 */
static void map444(const char *type1, const char *type2, const char *type3, char *t1, char *t2,
                   char *t3) {
  char from[6], to[6];
  int nmap = 4;
  from[0] = type3[0]; to[0] = '1';
  from[1] = type3[1]; to[1] = '2';
  from[2] = type3[2]; to[2] = '5';
  from[3] = type3[3]; to[3] = '6';
  for (int i = 0; i < 4; i++) {
    int contracted = 1;
    for (int j = 0; j < 4; j++) if (type1[i] == type3[j]) { contracted = 0; break; }
    if (contracted && nmap < 6) {
      from[nmap] = type1[i];
      to[nmap] = nmap == 4 ? '3' : '4';
      nmap++;
    }
  }
  memcpy(t1, type1, 4); memcpy(t2, type2, 4); memcpy(t3, type3, 4);
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < nmap; j++) if (type1[i] == from[j]) { t1[i] = to[j]; break; }
    for (int j = 0; j < nmap; j++) if (type2[i] == from[j]) { t2[i] = to[j]; break; }
    for (int j = 0; j < nmap; j++) if (type3[i] == from[j]) { t3[i] = to[j]; break; }
  }
}

static int fast_kind(const char *t1, const char *t2, const char *t3, const orc_desc *d1,
                     const orc_desc *d2, const orc_desc *d3, char *tr1, char *tr2, i64 *iop_mn) {
  if (memcmp(t3, "1256", 4) != 0) return 0;
  if (!memcmp(t1, "3412", 4) && !memcmp(t2, "3456", 4)) {
    if (d1->iop[0] == d2->iop[0] && d1->iop[1] == d3->iop[0] && d2->iop[1] == d3->iop[1]) {
      *tr1 = 'T'; *tr2 = 'N'; *iop_mn = d1->iop[0]; return 1;
    }
  } else if (!memcmp(t1, "1234", 4) && !memcmp(t2, "3456", 4)) {
    if (d1->iop[1] == d2->iop[0] && d1->iop[0] == d3->iop[0] && d2->iop[1] == d3->iop[1]) {
      *tr1 = 'N'; *tr2 = 'N'; *iop_mn = d1->iop[1]; return 1;
    }
  } else if (!memcmp(t1, "1234", 4) && !memcmp(t2, "5634", 4)) {
    if (d1->iop[1] == d2->iop[1] && d1->iop[0] == d3->iop[0] && d2->iop[0] == d3->iop[1]) {
      *tr1 = 'N'; *tr2 = 'T'; *iop_mn = d1->iop[1]; return 1;
    }
  } else if (!memcmp(t1, "3412", 4) && !memcmp(t2, "5634", 4)) {
    if (d1->iop[0] == d2->iop[1] && d1->iop[1] == d3->iop[0] && d2->iop[0] == d3->iop[1]) {
      *tr1 = 'T'; *tr2 = 'T'; *iop_mn = d1->iop[0]; return 1;
    }
  }
  return 0;
}

/* vv2v_dpd_444_d.f90:120-235 (t1,t2,t3 already canonical) */
int orc_vv2v_dpd_444_fast(const double *v1, const orc_desc *d1, const char *t1,
                          const double *v2, const orc_desc *d2, const char *t2, double *v3,
                          const orc_desc *d3, const char *t3, double fact1, double fact2) {
  char tr1, tr2;
  i64 iop_mn;
  if (!fast_kind(t1, t2, t3, d1, d2, d3, &tr1, &tr2, &iop_mn)) return 1;
  i64 inumpq[9], inumrs[9], inummn[9], off1[10], off2[10], off3[10];
  for (i64 i = 1; i <= g_nirrep; i++) {
    inumpq[i] = orc_irpdso(i, d3->pop[0], d3->pop[1], d3->iop[0]);
    inumrs[i] = orc_irpdso(i, d3->pop[2], d3->pop[3], d3->iop[1]);
    inummn[i] = (tr1 == 'T') ? orc_irpdso(i, d1->pop[0], d1->pop[1], iop_mn)
                             : orc_irpdso(i, d1->pop[2], d1->pop[3], iop_mn);
  }
  off1[1] = off2[1] = off3[1] = 0;
  for (i64 i = 1; i < g_nirrep; i++) {
    i64 l1 = DP(i, d1->sym), l2 = DP(i, d2->sym), l3 = DP(i, d3->sym);
    off1[i + 1] = off1[i] + (tr1 == 'T' ? inummn[l1] * inumpq[i] : inumpq[l1] * inummn[i]);
    off2[i + 1] = off2[i] + (tr2 == 'N' ? inummn[l2] * inumrs[i] : inumrs[l2] * inummn[i]);
    off3[i + 1] = off3[i] + inumpq[l3] * inumrs[i];
  }
  for (i64 rs = 1; rs <= g_nirrep; rs++) {
    i64 pq = DP(rs, d3->sym), mn = DP(pq, d1->sym);
    i64 o1 = (tr1 == 'T') ? off1[pq] : off1[mn];
    i64 o2 = (tr2 == 'N') ? off2[rs] : off2[mn];
    i64 o3 = off3[rs];
    i64 m = inumpq[pq], n = inumrs[rs], k = inummn[mn];
    if (m == 0 || n == 0) continue;
    if (k == 0) {
      for (i64 i = 0; i < m * n; i++) v3[o3 + i] = (fact2 == 0.0) ? 0.0 : fact2 * v3[o3 + i];
      continue;
    }
    i64 lda = (tr1 == 'N') ? m : k;
    i64 ldb = (tr2 == 'T') ? n : k;
    gemm(tr1, tr2, m, n, k, fact1, v1 + o1, lda, v2 + o2, ldb, fact2, v3 + o3, m);
  }
  return 0;
}

/* vv2v_dpd_444.f90:55-232 */
int orc_vv2v_dpd_444(const double *v1, const orc_desc *d1, const char *type1,
                     const double *v2, const orc_desc *d2, const char *type2, double *v3,
                     const orc_desc *d3, const char *type3, double fact1, double fact2) {
  char t1[4], t2[4], t3[4];
  map444(type1, type2, type3, t1, t2, t3);
  int p13 = -1, p14 = -1, p23 = -1, p24 = -1;
  for (int i = 0; i < 4; i++) {
    if (t1[i] == '3') p13 = i;
    if (t1[i] == '4') p14 = i;
    if (t2[i] == '3') p23 = i;
    if (t2[i] == '4') p24 = i;
  }
  if (p13 >= 0 && p23 >= 0 && !pops_equal(d1->pop[p13], d2->pop[p23]))
    return set_err(-3, "FATAL ERROR in vv2v_dpd_444: contracted index 1 populations differ (%.4s * %.4s -> %.4s)",
                   type1, type2, type3);
  if (p14 >= 0 && p24 >= 0 && !pops_equal(d1->pop[p14], d2->pop[p24]))
    return set_err(-3, "FATAL ERROR in vv2v_dpd_444: contracted index 2 populations differ (%.4s * %.4s -> %.4s)",
                   type1, type2, type3);
  char tr1, tr2;
  i64 iop_mn;
  if (fast_kind(t1, t2, t3, d1, d2, d3, &tr1, &tr2, &iop_mn))
    return orc_vv2v_dpd_444_fast(v1, d1, t1, v2, d2, t2, v3, d3, t3, fact1, fact2);
  if (d1->iop[0] != 1 || d1->iop[1] != 1 || d2->iop[0] != 1 || d2->iop[1] != 1 || d3->iop[0] != 1 ||
      d3->iop[1] != 1)
    return set_err(-4, "FATAL ERROR in vv2v_dpd_444: the fallback path forbids compressed IOP (%.4s * %.4s -> %.4s)",
                   type1, type2, type3);
  return orc_contract_general(v1, d1, t1, 4, v2, d2, t2, 4, v3, d3, t3, 4, fact1, fact2);
}

/* vv2v_dpd_424.f90:21-200: sym check, exactly one contracted index, IOP barrier */
int orc_vv2v_dpd_424(const double *v1, const orc_desc *d1, const char *type1,
                     const double *v2, const orc_desc *d2, const char *type2, double *v3,
                     const orc_desc *d3, const char *type3, double fact1, double fact2) {
  if (DP(d1->sym, d2->sym) != d3->sym)
    return set_err(-7, "FATAL ERROR in vv2v_dpd_424: Sym1*Sym2 != Sym3");
  int ncon = 0;
  char c = 0;
  for (int i = 0; i < 4; i++) {
    int in3 = 0;
    for (int j = 0; j < 4; j++) if (type1[i] == type3[j]) in3 = 1;
    if (!in3) { ncon++; c = type1[i]; }
  }
  if (ncon != 1) return set_err(-8, "FATAL ERROR: vv2v_dpd_424 requires EXACTLY 1 contracted index.");
  if (type2[0] != c && type2[1] != c) return set_err(-8, "FATAL ERROR: contracted index missing in V2.");
  /* IOP barrier: a packed pair must flow intact (same order, same iop) */
  const char *t1p[2] = {type1, type1 + 2}, *t3p[2] = {type3, type3 + 2};
  for (int s = 0; s < 2; s++) {
    if (d1->iop[s] != 1) {
      int ok = 0;
      for (int u = 0; u < 2; u++)
        if (!memcmp(t1p[s], t3p[u], 2) && d1->iop[s] == d3->iop[u]) ok = 1;
      if (!ok) return set_err(-4, "FATAL ERROR in vv2v_dpd_424: IOP barrier (%.4s * %.2s -> %.4s)", type1, type2, type3);
    }
    if (d3->iop[s] != 1) {
      int ok = 0;
      for (int u = 0; u < 2; u++)
        if (!memcmp(t3p[s], t1p[u], 2) && d3->iop[s] == d1->iop[u]) ok = 1;
      if (!memcmp(t3p[s], type2, 2) && d3->iop[s] == d2->iop[0]) ok = 1;
      if (!ok) return set_err(-4, "FATAL ERROR in vv2v_dpd_424: IOP barrier (%.4s * %.2s -> %.4s)", type1, type2, type3);
    }
  }
  if (d2->iop[0] != 1) {
    int ok = 0;
    for (int u = 0; u < 2; u++) if (!memcmp(type2, t3p[u], 2) && d2->iop[0] == d3->iop[u]) ok = 1;
    if (!ok) return set_err(-4, "FATAL ERROR in vv2v_dpd_424: IOP barrier (%.4s * %.2s -> %.4s)", type1, type2, type3);
  }
  return orc_contract_general(v1, d1, type1, 4, v2, d2, type2, 2, v3, d3, type3, 4, fact1, fact2);
}

/* vv2v_dpd_442.f90:22-200: sym check, exactly three contracted indices */
int orc_vv2v_dpd_442(const double *v1, const orc_desc *d1, const char *type1,
                     const double *v2, const orc_desc *d2, const char *type2, double *v3,
                     const orc_desc *d3, const char *type3, double fact1, double fact2) {
  if (DP(d1->sym, d2->sym) != d3->sym)
    return set_err(-7, "FATAL ERROR in vv2v_dpd_442: Sym1*Sym2 != Sym3");
  int ncon = 0;
  for (int i = 0; i < 4; i++) {
    int in3 = (type1[i] == type3[0] || type1[i] == type3[1]);
    if (!in3) {
      ncon++;
      int in2 = 0;
      for (int j = 0; j < 4; j++) if (type2[j] == type1[i]) in2 = 1;
      if (!in2) return set_err(-8, "FATAL ERROR: Contracted index missing in V2.");
    }
  }
  if (ncon != 3) return set_err(-8, "FATAL ERROR: vv2v_dpd_442 requires EXACTLY 3 contracted indices.");
  return orc_contract_general(v1, d1, type1, 4, v2, d2, type2, 4, v3, d3, type3, 2, fact1, fact2);
}

/* ------------------------------------------------------------------------- */
/* contract surface: lib_cc/mod_contract.f90                                  */
/* ------------------------------------------------------------------------- */
/* check_flat, mod_contract.f90:493-510 */
static int check_flat(const char *cur, const char (*pairs)[2], int n, int self) {
  for (int i = 0; i < n; i++) {
    if (i == self) continue;
    if (pairs[i][0] == cur[0] && pairs[i][1] == cur[1]) return 0;
  }
  return 1;
}

static void daxpy_(i64 n, double a, const double *x, double *y) {
  for (i64 i = 0; i < n; i++) y[i] += a * x[i];
}

/* after the engine: antisymmetrise/compress v3 pairs (mod_contract.f90:111-160) and add */
static int finish_anti(double *ptr3, const orc_desc *d3, const i64 *iop3, const i64 *iexp3,
                       i64 anti_pq, i64 anti_rs, double *v3out) {
  i64 n;
  double *cur = ptr3;
  double *tmp1 = NULL, *tmp2 = NULL;
  i64 iop5 = iop3[0], iop6 = iop3[1];
  if (anti_pq) {
    if (iop5 != 1 || !pops_equal(d3->pop[0], d3->pop[1]))
      return set_err(-9, "ERROR: vv2v anti 1st pair error.");
    n = orc_idsymsoc(d3->sym, d3->pop[0], d3->pop[1], iop5, d3->pop[2], d3->pop[3], iop6);
    double *sw = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    orc_isymtrso(cur, sw, d3->pop[0], d3->pop[1], iop5, 1, d3->pop[2], d3->pop[3], iop6, 0, d3->sym);
    daxpy_(n, -1.0, sw, cur);
    free(sw);
  }
  if (iexp3[0]) {
    n = orc_idsymsoc(d3->sym, d3->pop[0], d3->pop[1], d3->iop[0], d3->pop[2], d3->pop[3], iop6);
    tmp1 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    orc_icompso(cur, tmp1, d3->pop[0], d3->pop[1], 1, d3->iop[0], d3->pop[2], d3->pop[3], iop6, 1, d3->sym);
    cur = tmp1;
  }
  i64 iopL = d3->iop[0];
  if (anti_rs) {
    if (iop6 != 1 || !pops_equal(d3->pop[2], d3->pop[3])) {
      free(tmp1);
      return set_err(-9, "ERROR: vv2v anti 2nd pair error.");
    }
    n = orc_idsymsoc(d3->sym, d3->pop[0], d3->pop[1], iopL, d3->pop[2], d3->pop[3], iop6);
    double *sw = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    orc_isymtrso(cur, sw, d3->pop[0], d3->pop[1], iopL, 0, d3->pop[2], d3->pop[3], iop6, 1, d3->sym);
    daxpy_(n, -1.0, sw, cur);
    free(sw);
  }
  if (iexp3[1]) {
    n = orc_idsymsoc(d3->sym, d3->pop[0], d3->pop[1], iopL, d3->pop[2], d3->pop[3], d3->iop[1]);
    tmp2 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    orc_icompso(cur, tmp2, d3->pop[0], d3->pop[1], iopL, 1, d3->pop[2], d3->pop[3], 1, d3->iop[1], d3->sym);
    cur = tmp2;
  }
  n = orc_desc_len(d3);
  daxpy_(n, 1.0, cur, v3out);
  free(tmp1);
  free(tmp2);
  return 0;
}

static double *expand_copy(const double *v, const orc_desc *d, const i64 *iexp, i64 *iop_new) {
  iop_new[0] = iexp[0] ? 1 : d->iop[0];
  iop_new[1] = iexp[1] ? 1 : d->iop[1];
  if (!iexp[0] && !iexp[1]) return NULL;
  i64 n = orc_idsymsoc(d->sym, d->pop[0], d->pop[1], iop_new[0], d->pop[2], d->pop[3], iop_new[1]);
  double *o = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  if (orc_iexpso(v, o, d->pop[0], d->pop[1], d->iop[0], iexp[0], d->pop[2], d->pop[3], d->iop[1],
                 iexp[1], d->sym) != 0) {
    free(o);
    return NULL;
  }
  return o;
}

/* mod_contract.f90:10-162 */
int orc_vv2v_444(const double *v1, const orc_desc *d1, const double *v2, const orc_desc *d2,
                 double *v3, const orc_desc *d3, double fact, const char *type1,
                 const char *type2, const char *type3, i64 anti) {
  if (anti < 0 || anti > 3) return set_err(-10, "ERROR: vv2v_444 error in antisymmetrize. %lld", (long long)anti);
  char pairs[6][2] = {{type1[0], type1[1]}, {type1[2], type1[3]}, {type2[0], type2[1]},
                      {type2[2], type2[3]}, {type3[0], type3[1]}, {type3[2], type3[3]}};
  i64 iop[6] = {d1->iop[0], d1->iop[1], d2->iop[0], d2->iop[1], d3->iop[0], d3->iop[1]};
  i64 iexp[6] = {0};
  for (int i = 0; i < 6; i++)
    if (check_flat(pairs[i], (const char(*)[2])pairs, 6, i) && iop[i] != 1) { iexp[i] = 1; iop[i] = 1; }
  if (anti == 0 && (iexp[4] || iexp[5]))
    return set_err(-11, "ERROR: vv2v_444 output pair must be expanded but anti=0 (result would be lost)");
  i64 io1[2], io2[2];
  double *e1 = expand_copy(v1, d1, iexp, io1);
  double *e2 = expand_copy(v2, d2, iexp + 2, io2);
  orc_desc dd1 = *d1, dd2 = *d2, dd3 = *d3;
  dd1.iop[0] = io1[0]; dd1.iop[1] = io1[1];
  dd2.iop[0] = io2[0]; dd2.iop[1] = io2[1];
  dd3.iop[0] = iop[4]; dd3.iop[1] = iop[5];
  double *p3 = v3;
  double *scratch = NULL;
  if (anti != 0) {
    i64 n = orc_desc_len(&dd3);
    scratch = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    p3 = scratch;
  }
  int rc = orc_vv2v_dpd_444(e1 ? e1 : v1, &dd1, type1, e2 ? e2 : v2, &dd2, type2, p3, &dd3, type3, fact, 1.0);
  if (rc == 0 && anti != 0)
    rc = finish_anti(p3, d3, iop + 4, iexp + 4, anti == 1 || anti == 3, anti == 2 || anti == 3, v3);
  free(e1); free(e2); free(scratch);
  return rc;
}

/* mod_contract.f90:223-368 */
int orc_vv2v_424(const double *v1, const orc_desc *d1, const double *v2, const orc_desc *d2,
                 double *v3, const orc_desc *d3, double fact, const char *type1,
                 const char *type2, const char *type3, i64 anti) {
  if (anti < 0 || anti > 2) return set_err(-10, "ERROR: vv2v_424 error in antisymmetrize. %lld", (long long)anti);
  char pairs[4][2] = {{type1[0], type1[1]}, {type1[2], type1[3]}, {type3[0], type3[1]}, {type3[2], type3[3]}};
  i64 iop[4] = {d1->iop[0], d1->iop[1], d3->iop[0], d3->iop[1]};
  i64 iexp[4] = {0};
  for (int i = 0; i < 4; i++)
    if (check_flat(pairs[i], (const char(*)[2])pairs, 4, i) && iop[i] != 1) { iexp[i] = 1; iop[i] = 1; }
  if (iexp[2] && anti != 1)
    return set_err(-12, "ERROR: vv2v_424 error for the 1st pair not antisymmertize but be expand.");
  if (iexp[3] && anti != 2)
    return set_err(-12, "ERROR: vv2v_424 error for the 2nd pair not antisymmertize but be expand.");
  i64 io1[2];
  double *e1 = expand_copy(v1, d1, iexp, io1);
  orc_desc dd1 = *d1, dd3 = *d3;
  dd1.iop[0] = io1[0]; dd1.iop[1] = io1[1];
  dd3.iop[0] = iop[2]; dd3.iop[1] = iop[3];
  double *p3 = v3, *scratch = NULL;
  if (anti != 0) {
    i64 n = orc_desc_len(&dd3);
    scratch = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    p3 = scratch;
  }
  int rc = orc_vv2v_dpd_424(e1 ? e1 : v1, &dd1, type1, v2, d2, type2, p3, &dd3, type3, fact, 1.0);
  if (rc == 0 && anti != 0)
    rc = finish_anti(p3, d3, iop + 2, iexp + 2, anti == 1, anti == 2, v3);
  free(e1); free(scratch);
  return rc;
}

/* mod_contract.f90:372-455 */
int orc_vv2v_442(const double *v1, const orc_desc *d1, const double *v2, const orc_desc *d2,
                 double *v3, const orc_desc *d3, double fact, const char *type1,
                 const char *type2, const char *type3) {
  char pairs[4][2] = {{type1[0], type1[1]}, {type1[2], type1[3]}, {type2[0], type2[1]}, {type2[2], type2[3]}};
  i64 iop[4] = {d1->iop[0], d1->iop[1], d2->iop[0], d2->iop[1]};
  i64 iexp[4] = {0};
  for (int i = 0; i < 4; i++)
    if (check_flat(pairs[i], (const char(*)[2])pairs, 4, i) && iop[i] != 1) { iexp[i] = 1; iop[i] = 1; }
  i64 io1[2], io2[2];
  double *e1 = expand_copy(v1, d1, iexp, io1);
  double *e2 = expand_copy(v2, d2, iexp + 2, io2);
  orc_desc dd1 = *d1, dd2 = *d2;
  dd1.iop[0] = io1[0]; dd1.iop[1] = io1[1];
  dd2.iop[0] = io2[0]; dd2.iop[1] = io2[1];
  int rc = orc_vv2v_dpd_442(e1 ? e1 : v1, &dd1, type1, e2 ? e2 : v2, &dd2, type2, v3, d3, type3, fact, 1.0);
  free(e1); free(e2);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* elementwise helpers                                                        */
/* ------------------------------------------------------------------------- */
/* orbital-energy offsets: occupied irreps 1..n, then virtual irreps 1..n
 * (lib_cc/Eorb_diff.f90:26-35) */
static void eorb_offsets(const i64 *pop_a, const i64 *pop_i, i64 *offv, i64 *offo) {
  offo[1] = 0;
  for (i64 h = 1; h < g_nirrep; h++) offo[h + 1] = offo[h] + pop_i[h - 1];
  offv[1] = offo[g_nirrep] + pop_i[g_nirrep - 1];
  for (i64 h = 1; h < g_nirrep; h++) offv[h + 1] = offv[h] + pop_a[h - 1];
}

typedef struct { const double *e; i64 offv[9], offo[9]; double *out; const double *in; int kind; } ew_u;

static void ew_cb(void *uu, i64 addr, const i64 *h, const i64 *x) {
  ew_u *u = (ew_u *)uu;
  if (u->kind == 1) { /* OED1 e(a,i) = e_a - e_i, Eorb_diff.f90:6-63 */
    u->out[addr] = u->e[u->offv[h[0]] + x[0]] - u->e[u->offo[h[1]] + x[1]];
  } else { /* OED2 e(ab,ij) = e_a + e_b - e_i - e_j  or  t = r / D */
    double ea = u->e[u->offv[h[0]] + x[0]], eb = u->e[u->offv[h[1]] + x[1]];
    double ei = u->e[u->offo[h[2]] + x[2]], ej = u->e[u->offo[h[3]] + x[3]];
    if (u->kind == 2) u->out[addr] = ea + eb - ei - ej;
    else u->out[addr] = u->in[addr] / (ei + ej - ea - eb); /* denomT.f90:130-136 */
  }
}

int orc_eorb_diff1(double *oed1, const double *eorb, const i64 *pop_a, const i64 *pop_i, i64 sym) {
  orc_desc d;
  make_desc(&d, pop_a, pop_i, ZORB, ZORB, 1, 1, sym);
  orc_lay L;
  lay_init(&L, &d);
  ew_u u;
  u.e = eorb; u.out = oed1; u.in = NULL; u.kind = 1;
  eorb_offsets(pop_a, pop_i, u.offv, u.offo);
  for_each_elem(&L, ew_cb, &u);
  return 0;
}

int orc_eorb_diff2(double *oed2, const double *eorb, const i64 *pop_a, const i64 *pop_i, i64 sym, i64 iop) {
  orc_desc d;
  make_desc(&d, pop_a, pop_a, pop_i, pop_i, iop, iop, sym);
  orc_lay L;
  lay_init(&L, &d);
  ew_u u;
  u.e = eorb; u.out = oed2; u.in = NULL; u.kind = 2;
  eorb_offsets(pop_a, pop_i, u.offv, u.offo);
  for_each_elem(&L, ew_cb, &u);
  return 0;
}

int orc_denomT2(double *t2, const double *r2, const double *eorb, const i64 *pop_a, const i64 *pop_i,
                i64 sym, i64 iop) {
  orc_desc d;
  make_desc(&d, pop_a, pop_a, pop_i, pop_i, iop, iop, sym);
  orc_lay L;
  lay_init(&L, &d);
  ew_u u;
  u.e = eorb; u.out = t2; u.in = r2; u.kind = 3;
  eorb_offsets(pop_a, pop_i, u.offv, u.offo);
  for_each_elem(&L, ew_cb, &u);
  return 0;
}

/* putDiagFock.f90:6-55: f(p,p) = fact*e(p) in a (p,q) iop=1 sym=1 matrix */
int orc_putDiagFock(double *fock, const double *eorb, const i64 *pop_a, const i64 *pop_i, char type,
                    double fact) {
  i64 offv[9], offo[9];
  eorb_offsets(pop_a, pop_i, offv, offo);
  const i64 *pop_p = (type == 'O') ? pop_i : pop_a;
  const i64 *offp = (type == 'O') ? offo : offv;
  for (i64 h = 1; h <= g_nirrep; h++) {
    i64 n = pop_p[h - 1];
    i64 base = orc_isymoffso(h, 1, pop_p, pop_p, 1) - 1;
    for (i64 p = 0; p < n; p++) fock[base + n * p + p] = fact * eorb[offp[h] + p];
  }
  return 0;
}
