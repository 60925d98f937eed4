/* Fortran-callable entry points: the reference's own external names and
 * argument lists (src/config.h:7-11 convention with -i8: lower case + trailing
 * underscore, every argument by reference, INTEGER = int64, one hidden length
 * per CHARACTER argument after the explicit ones), so that lib_cc and the
 * method programs link against libdpdhip.so unchanged in place of the libsrint
 * objects that define these names.  Arrays may be device pointers (the
 * reference's igpu = 1 contract) or host memory (staged through the device);
 * each call returns with its result complete; an error prints the message and
 * stops the program (the reference's `write(6,*); stop`, exit status 1).  Data
 * are FP64 (prec = 8) unless dpd_fortran_set_prec_ selects FP32 (prec = 4): the
 * engines, sorts and xgemmso then take REAL*4 arrays through the same
 * DOUBLE PRECISION dummy arguments, as the reference's routers do. */
#ifndef DPD_FORTRAN_H
#define DPD_FORTRAN_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* No reference counterpart: sets the state the reference keeps in the mod_symm
 * globals (lib_mod/mod_symm.f: nirrep, dirprd(8,8) column major) and the GPU to
 * use (device -1: host-only, metadata functions only).  Called once by the
 * program before the first engine call. */
void dpd_fortran_init_(const int64_t *nirrep, const int64_t *dirprd, const int64_t *device);
void dpd_fortran_finalize_(void);
/* No reference counterpart as a call: sets what the reference keeps in the
 * mod_iop variable `prec` (libsrint/readinp.f90:120-127), 8 or 4. */
void dpd_fortran_set_prec_(const int64_t *prec);

/* libsrint/vv2v_dpd_444.f90:55-57 */
void vv2v_dpd_444_(const double *v1, const int64_t *p1_1, const int64_t *p1_2, const int64_t *iop1_12,
                   const int64_t *p1_3, const int64_t *p1_4, const int64_t *iop1_34, const int64_t *sym1,
                   const char *type1, const double *v2, const int64_t *p2_1, const int64_t *p2_2,
                   const int64_t *iop2_12, const int64_t *p2_3, const int64_t *p2_4, const int64_t *iop2_34,
                   const int64_t *sym2, const char *type2, double *v3, const int64_t *p3_1, const int64_t *p3_2,
                   const int64_t *iop3_12, const int64_t *p3_3, const int64_t *p3_4, const int64_t *iop3_34,
                   const int64_t *sym3, const char *type3, const double *fact1, const double *fact2, size_t len1,
                   size_t len2, size_t len3);
/* libsrint/vv2v_dpd_424.f90:21-23 */
void vv2v_dpd_424_(const double *v1, const int64_t *p1_1, const int64_t *p1_2, const int64_t *iop1_12,
                   const int64_t *p1_3, const int64_t *p1_4, const int64_t *iop1_34, const int64_t *sym1,
                   const char *type1, const double *v2, const int64_t *p2_1, const int64_t *p2_2,
                   const int64_t *iop2_12, const int64_t *sym2, const char *type2, double *v3, const int64_t *p3_1,
                   const int64_t *p3_2, const int64_t *iop3_12, const int64_t *p3_3, const int64_t *p3_4,
                   const int64_t *iop3_34, const int64_t *sym3, const char *type3, const double *fact1,
                   const double *fact2, size_t len1, size_t len2, size_t len3);
/* libsrint/vv2v_dpd_442.f90:22-24 */
void vv2v_dpd_442_(const double *v1, const int64_t *p1_1, const int64_t *p1_2, const int64_t *iop1_12,
                   const int64_t *p1_3, const int64_t *p1_4, const int64_t *iop1_34, const int64_t *sym1,
                   const char *type1, const double *v2, const int64_t *p2_1, const int64_t *p2_2,
                   const int64_t *iop2_12, const int64_t *p2_3, const int64_t *p2_4, const int64_t *iop2_34,
                   const int64_t *sym2, const char *type2, double *v3, const int64_t *p3_1, const int64_t *p3_2,
                   const int64_t *iop3_12, const int64_t *sym3, const char *type3, const double *fact1,
                   const double *fact2, size_t len1, size_t len2, size_t len3);
/* libsrint/vv2v_dpd_323.f90:22-25 (type1*3, type2*2, type3*3) */
void vv2v_dpd_323_(const double *v1, const int64_t *p1_1, const int64_t *p1_2, const int64_t *iop1_12,
                   const int64_t *p1_3, const int64_t *sym1, const char *type1, const double *v2,
                   const int64_t *p2_1, const int64_t *p2_2, const int64_t *iop2_12, const int64_t *sym2,
                   const char *type2, double *v3, const int64_t *p3_1, const int64_t *p3_2,
                   const int64_t *iop3_12, const int64_t *p3_3, const int64_t *sym3, const char *type3,
                   const double *fact1, const double *fact2, size_t len1, size_t len2, size_t len3);
/* libsrint/xgemmso.f:1-25 (XGEMMSO) */
void xgemmso_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
              const double *alpha, const double *a, const int64_t *lda, const double *b, const int64_t *ldb,
              const double *beta, double *c, const int64_t *ldc, size_t lena, size_t lenb);
/* libsrint/sstgenso.f90:22 (sstgenso1) and :1 (sstgenso; nsize and ipos unused) */
void sstgenso1_(const double *win, double *wout, const int64_t *pop1, const int64_t *pop2, const int64_t *pop3,
                const int64_t *pop4, const int64_t *irrepx, const char *type, size_t lentype);
void sstgenso_(const double *win, double *wout, const int64_t *nsize, const int64_t *pop1, const int64_t *pop2,
               const int64_t *pop3, const int64_t *pop4, const double *ipos, const int64_t *irrepx,
               const char *type, size_t lentype);
/* libsrint/iexpso.f90:24 */
void iexpso_(const double *iin, double *iout, const int64_t *popp, const int64_t *popq, const int64_t *ioppq,
             const int64_t *iexppq, const int64_t *popr, const int64_t *pops, const int64_t *ioprs,
             const int64_t *iexprs, const int64_t *sym);
/* libsrint/icompso.f:1 */
void icompso_(const double *iin, double *iout, const int64_t *popp, const int64_t *popq, const int64_t *ioppq,
              const int64_t *icpq, const int64_t *popr, const int64_t *pops, const int64_t *ioprs,
              const int64_t *icrs, const int64_t *sym);
/* libsrint/isymtrso.f:1 */
void isymtrso_(const double *iin, double *iout, const int64_t *popp, const int64_t *popq, const int64_t *ioppq,
               const int64_t *icpq, const int64_t *popr, const int64_t *pops, const int64_t *ioprs,
               const int64_t *icrs, const int64_t *sym);
/* libsrint/itranspso.f:1 */
void itranspso_(const double *iin, double *iout, const int64_t *popp, const int64_t *popq, const int64_t *ioppq,
                const int64_t *popr, const int64_t *pops, const int64_t *ioprs, const int64_t *sym);
/* libsrint/iassymso.f:1 */
void iassymso_(const double *vin, double *vout, const int64_t *symv, const int64_t *popp, const int64_t *popq,
               const int64_t *popr, const int64_t *pops, const int64_t *iop);
/* libsrint/irpdso.f:2, idsymsoc.f:2, isymoffso.f:2 */
int64_t irpdso_(const int64_t *sym, const int64_t *popp, const int64_t *popq, const int64_t *ioppq);
int64_t idsymsoc_(const int64_t *irrepx, const int64_t *popp, const int64_t *popq, const int64_t *ioppq,
                  const int64_t *popr, const int64_t *pops, const int64_t *ioprs);
int64_t isymoffso_(const int64_t *irrepq1, const int64_t *irrep, const int64_t *popp, const int64_t *popq,
                   const int64_t *iop);

#ifdef __cplusplus
}
#endif
#endif
