/*
 * dpd_oracle.h -- CPU oracle for the DPD contraction engine (TEST INFRASTRUCTURE ONLY).
 *
 * This header declares a plain-C restatement of the reference's CPU path
 * (libsrint DPD metadata, sort kernels, vv2v_dpd_* engines and the lib_cc
 * mod_contract surface).  It is the checker that the HIP product is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product never links, imports or calls it.
 *
 * Conventions (mirroring the reference):
 *   - irreps are 1-based (1..nirrep), dirprd is the 8x8 direct-product table;
 *   - every population array has 8 entries; index arrays are int64;
 *   - tensors are FP64, column-major DPD: blocks ordered by right-pair irrep,
 *     left pair fastest inside a block;
 *   - pair packing iop: 0 = p<q, 1 = full (p,q), 2 = p<=q.
 * Functions return 0 on success and a negative code on an error that the
 * reference reports with `write(6,*)...; stop` (the message is kept in
 * orc_last_error()).
 */
#ifndef DPD_ORACLE_H
#define DPD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t i64;

/* 4-index DPD tensor descriptor: restates `type opt` (lib_cc/mod_operator.f90:8-25). */
typedef struct {
  i64 pop[4][8]; /* popp, popq, popr, pops */
  i64 iop[2];    /* ioppq, ioprs */
  i64 sym;       /* 1-based irrep of the tensor */
} orc_desc;

const char *orc_last_error(void);

/* global symmetry state (lib_mod/mod_symm.f) */
void orc_set_symmetry(i64 nirrep, const i64 *dirprd64);
/* optional BLAS: Fortran ILP64 dgemm_ pointer (e.g. OpenBLAS scipy_dgemm_64_) */
void orc_set_dgemm(void *fn);
int orc_has_blas(void);
/* optional: openblas_set_num_threads-style setter, used around task-parallel loops */
void orc_set_blas_threads_fn(void *fn);
void orc_set_threads(int n);

/* metadata: irpdso.f, idsymsoc.f, isymoffso.f, getidx.f90, putidx.f90 */
i64 orc_irpdso(i64 sym, const i64 *popp, const i64 *popq, i64 iop);
i64 orc_idsymsoc(i64 irrepx, const i64 *popp, const i64 *popq, i64 ioppq,
                 const i64 *popr, const i64 *pops, i64 ioprs);
i64 orc_isymoffso(i64 irrepq1, i64 irrep, const i64 *popp, const i64 *popq, i64 iop);
i64 orc_desc_len(const orc_desc *d);
/* linear (1-based) index -> (irrp,irrq,irrr,irrs,p,q,r,s) (all 1-based), getidx.f90 */
int orc_getidx(i64 idx, i64 *info8, const orc_desc *d);
/* (irrp,irrq,irrr,irrs,p,q,r,s) -> 1-based linear index, putidx.f90 (0 if not stored) */
i64 orc_putidx(const i64 *info8, const orc_desc *d);

/* sort kernels: iexpso.f90, icompso.f, isymtrso.f, itranspso.f, sstgenso.f90 */
int orc_iexpso(const double *in, double *out, const i64 *popp, const i64 *popq, i64 ioppq,
               i64 iexppq, const i64 *popr, const i64 *pops, i64 ioprs, i64 iexprs, i64 sym);
int orc_icompso(const double *in, double *out, const i64 *popp, const i64 *popq, i64 ioppq,
                i64 icpq, const i64 *popr, const i64 *pops, i64 ioprs, i64 icrs, i64 sym);
int orc_isymtrso(const double *in, double *out, const i64 *popp, const i64 *popq, i64 ioppq,
                 i64 icpq, const i64 *popr, const i64 *pops, i64 ioprs, i64 icrs, i64 sym);
int orc_itranspso(const double *in, double *out, const i64 *popp, const i64 *popq, i64 ioppq,
                  const i64 *popr, const i64 *pops, i64 ioprs, i64 sym);
int orc_iassymso(const double *in, double *out, i64 sym, const i64 *popp, const i64 *popq,
                 const i64 *popr, const i64 *pops, i64 iop);
int orc_sstgenso(const double *win, double *wout, const i64 *pop1, const i64 *pop2,
                 const i64 *pop3, const i64 *pop4, i64 irrepx, const char *type4);

/* engines: vv2v_dpd_444.f90 (+_d), vv2v_dpd_424.f90 (+_d), vv2v_dpd_442.f90 (+_d).
 * v3 = fact1 * contraction(v1, v2) + fact2 * v3.  2-index operands use pop[0..1],
 * iop[0] and sym; their type string has 2 characters. */
int orc_vv2v_dpd_444(const double *v1, const orc_desc *d1, const char *type1,
                     const double *v2, const orc_desc *d2, const char *type2,
                     double *v3, const orc_desc *d3, const char *type3,
                     double fact1, double fact2);
int orc_vv2v_dpd_424(const double *v1, const orc_desc *d1, const char *type1,
                     const double *v2, const orc_desc *d2, const char *type2,
                     double *v3, const orc_desc *d3, const char *type3,
                     double fact1, double fact2);
int orc_vv2v_dpd_442(const double *v1, const orc_desc *d1, const char *type1,
                     const double *v2, const orc_desc *d2, const char *type2,
                     double *v3, const orc_desc *d3, const char *type3,
                     double fact1, double fact2);
/* the general einsum-over-irrep-tasks engine all three share (exposed for tests) */
int orc_contract_general(const double *v1, const orc_desc *d1, const char *type1, int rank1,
                         const double *v2, const orc_desc *d2, const char *type2, int rank2,
                         double *v3, const orc_desc *d3, const char *type3, int rank3,
                         double fact1, double fact2);
/* the restated 444 fast path only (returns 1 if the pattern is not a fast path) */
int orc_vv2v_dpd_444_fast(const double *v1, const orc_desc *d1, const char *t1,
                          const double *v2, const orc_desc *d2, const char *t2,
                          double *v3, const orc_desc *d3, const char *t3,
                          double fact1, double fact2);

/* contract surface: lib_cc/mod_contract.f90 (v3 += fact * P{v1*v2}) */
int orc_vv2v_444(const double *v1, const orc_desc *d1, const double *v2, const orc_desc *d2,
                 double *v3, const orc_desc *d3, double fact, const char *type1,
                 const char *type2, const char *type3, i64 anti);
int orc_vv2v_424(const double *v1, const orc_desc *d1, const double *v2, const orc_desc *d2,
                 double *v3, const orc_desc *d3, double fact, const char *type1,
                 const char *type2, const char *type3, i64 anti);
int orc_vv2v_442(const double *v1, const orc_desc *d1, const double *v2, const orc_desc *d2,
                 double *v3, const orc_desc *d3, double fact, const char *type1,
                 const char *type2, const char *type3);

/* elementwise helpers used by the sigma builders */
int orc_eorb_diff1(double *oed1, const double *eorb, const i64 *pop_a, const i64 *pop_i, i64 sym);
int orc_eorb_diff2(double *oed2, const double *eorb, const i64 *pop_a, const i64 *pop_i,
                   i64 sym, i64 iop);
int orc_denomT2(double *t2, const double *r2, const double *eorb, const i64 *pop_a,
                const i64 *pop_i, i64 sym, i64 iop);
int orc_putDiagFock(double *fock, const double *eorb, const i64 *pop_a, const i64 *pop_i,
                    char type, double fact);

#ifdef __cplusplus
}
#endif
#endif
