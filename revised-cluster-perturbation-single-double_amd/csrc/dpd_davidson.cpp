// dpd_davidson.cpp -- non-Hermitian Davidson solver on the device-resident
// problem (lib_cc/mod_davidson.f90:40-726), FP64 throughout.
//
// The b and Ab vectors live in HBM (the reference streams them from disk through
// IO_stream); the subspace matrix Abar and its eigenproblem are small and are
// solved on the host in FP64.  Semantics follow the reference step by step:
//   * initial subspace = the guesses, Abar(j,i) = b_j . A b_i   (initDavidson)
//   * every iteration: eigenpairs of Abar (dgeev stand-in below), eigenvalues
//     sorted ascending by exchange sort, roots picked by the largest |component|
//     on basis vector `iroot` unless `lowest`                   (:186-259)
//   * collapse: when the subspace was full at the end of the previous iteration
//     the nroot+5 selected Ritz vectors are Gram-Schmidt orthonormalised in the
//     subspace and become the new basis                        (:268-341)
//   * per root: q = sum r_i b_i, Aq = sum r_i Ab_i, renormalise if |q| != 1,
//     e = Aq - w q, converged if max|e_i| < conv                (:352-406)
//   * new vector p_i = e_i / (w - A_ii) with |w - A_ii| clamped to 1e-4,
//     modified Gram-Schmidt against every basis vector, skipped if the
//     remaining norm is below 1e-6                              (:417-439, :687-726)
//   * all roots must converge in the same iteration             (:451-458)
//   * optional final check |A R - w R|_inf per root              (:527-542)
// The sigma of every new vector of one iteration is evaluated in one batched
// launch after the root loop; the reference evaluates the same vectors one by
// one inside the loop, and nothing in the loop reads them.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dpd_sigma.hpp"

namespace dpd {

// ---------------------------------------------------------------------------
// small dense nonsymmetric eigenproblem (FP64, row-major n x n)
// Householder reduction to Hessenberg form, Francis double-shift QR for the
// eigenvalues, inverse iteration on the original matrix for the eigenvectors
// of real eigenvalues; vectors normalised to unit 2-norm (as dgeev does).
// ---------------------------------------------------------------------------
static void hessenberg(std::vector<double>& a, int n) {
  for (int k = 0; k < n - 2; k++) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; i++) alpha += a[i * n + k] * a[i * n + k];
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (a[(k + 1) * n + k] > 0) alpha = -alpha;
    std::vector<double> v(n, 0.0);
    v[k + 1] = a[(k + 1) * n + k] - alpha;
    for (int i = k + 2; i < n; i++) v[i] = a[i * n + k];
    double vv = 0.0;
    for (int i = k + 1; i < n; i++) vv += v[i] * v[i];
    if (vv == 0.0) continue;
    // A = H A H with H = I - 2 v v^T / (v^T v)
    for (int j = 0; j < n; j++) {
      double s = 0.0;
      for (int i = k + 1; i < n; i++) s += v[i] * a[i * n + j];
      s = 2.0 * s / vv;
      for (int i = k + 1; i < n; i++) a[i * n + j] -= s * v[i];
    }
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = k + 1; j < n; j++) s += a[i * n + j] * v[j];
      s = 2.0 * s / vv;
      for (int j = k + 1; j < n; j++) a[i * n + j] -= s * v[j];
    }
    for (int i = k + 2; i < n; i++) a[i * n + k] = 0.0;
  }
}

static inline double sgn(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// Francis double-shift QR on an upper Hessenberg matrix (eigenvalues only)
static int hqr(std::vector<double>& hm, int n, double* wr, double* wi) {
  auto h = [&](int i, int j) -> double& { return hm[i * n + j]; };
  double anorm = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = std::max(i - 1, 0); j < n; j++) anorm += std::fabs(h(i, j));
  int nn = n - 1;
  double t = 0.0, p = 0, q = 0, r = 0, s, w, x, y, z;
  while (nn >= 0) {
    int its = 0, l;
    do {
      for (l = nn; l >= 1; l--) {
        s = std::fabs(h(l - 1, l - 1)) + std::fabs(h(l, l));
        if (s == 0.0) s = anorm;
        if (std::fabs(h(l, l - 1)) + s == s) {
          h(l, l - 1) = 0.0;
          break;
        }
      }
      x = h(nn, nn);
      if (l == nn) {
        wr[nn] = x + t;
        wi[nn--] = 0.0;
      } else {
        y = h(nn - 1, nn - 1);
        w = h(nn, nn - 1) * h(nn - 1, nn);
        if (l == nn - 1) {
          p = 0.5 * (y - x);
          q = p * p + w;
          z = std::sqrt(std::fabs(q));
          x += t;
          if (q >= 0.0) {
            z = p + sgn(z, p);
            wr[nn - 1] = wr[nn] = x + z;
            if (z != 0.0) wr[nn] = x - w / z;
            wi[nn - 1] = wi[nn] = 0.0;
          } else {
            wr[nn - 1] = wr[nn] = x + p;
            wi[nn - 1] = -(wi[nn] = z);
          }
          nn -= 2;
        } else {
          if (its == 60) return -1;
          if (its == 10 || its == 20 || its == 40) {
            t += x;
            for (int i = 0; i <= nn; i++) h(i, i) -= x;
            s = std::fabs(h(nn, nn - 1)) + std::fabs(h(nn - 1, nn - 2));
            y = x = 0.75 * s;
            w = -0.4375 * s * s;
          }
          ++its;
          int m;
          for (m = nn - 2; m >= l; m--) {
            z = h(m, m);
            r = x - z;
            s = y - z;
            p = (r * s - w) / h(m + 1, m) + h(m, m + 1);
            q = h(m + 1, m + 1) - z - r - s;
            r = h(m + 2, m + 1);
            s = std::fabs(p) + std::fabs(q) + std::fabs(r);
            p /= s;
            q /= s;
            r /= s;
            if (m == l) break;
            double u = std::fabs(h(m, m - 1)) * (std::fabs(q) + std::fabs(r));
            double v = std::fabs(p) * (std::fabs(h(m - 1, m - 1)) + std::fabs(z) + std::fabs(h(m + 1, m + 1)));
            if (u + v == v) break;
          }
          for (int i = m + 2; i <= nn; i++) {
            h(i, i - 2) = 0.0;
            if (i != m + 2) h(i, i - 3) = 0.0;
          }
          for (int k = m; k <= nn - 1; k++) {
            if (k != m) {
              p = h(k, k - 1);
              q = h(k + 1, k - 1);
              r = 0.0;
              if (k != nn - 1) r = h(k + 2, k - 1);
              if ((x = std::fabs(p) + std::fabs(q) + std::fabs(r)) != 0.0) {
                p /= x;
                q /= x;
                r /= x;
              }
            }
            if ((s = sgn(std::sqrt(p * p + q * q + r * r), p)) != 0.0) {
              if (k == m) {
                if (l != m) h(k, k - 1) = -h(k, k - 1);
              } else {
                h(k, k - 1) = -s * x;
              }
              p += s;
              x = p / s;
              y = q / s;
              z = r / s;
              q /= p;
              r /= p;
              for (int j = k; j <= nn; j++) {
                p = h(k, j) + q * h(k + 1, j);
                if (k != nn - 1) {
                  p += r * h(k + 2, j);
                  h(k + 2, j) -= p * z;
                }
                h(k + 1, j) -= p * y;
                h(k, j) -= p * x;
              }
              int mmin = nn < k + 3 ? nn : k + 3;
              for (int i = l; i <= mmin; i++) {
                p = x * h(i, k) + y * h(i, k + 1);
                if (k != nn - 1) {
                  p += z * h(i, k + 2);
                  h(i, k + 2) -= p * r;
                }
                h(i, k + 1) -= p * q;
                h(i, k) -= p;
              }
            }
          }
        }
      }
    } while (l < nn - 1);
  }
  return 0;
}

// solve (A - mu I) x = b by LU with partial pivoting (in place on a copy)
static void shifted_solve(const std::vector<double>& A, int n, double mu, std::vector<double>& b, double tiny) {
  std::vector<double> M(A);
  std::vector<int> piv(n);
  for (int i = 0; i < n; i++) M[i * n + i] -= mu;
  for (int k = 0; k < n; k++) {
    int pr = k;
    for (int i = k + 1; i < n; i++)
      if (std::fabs(M[i * n + k]) > std::fabs(M[pr * n + k])) pr = i;
    piv[k] = pr;
    if (pr != k) {
      for (int j = 0; j < n; j++) std::swap(M[k * n + j], M[pr * n + j]);
      std::swap(b[k], b[pr]);
    }
    if (std::fabs(M[k * n + k]) < tiny) M[k * n + k] = M[k * n + k] < 0 ? -tiny : tiny;
    for (int i = k + 1; i < n; i++) {
      double f = M[i * n + k] / M[k * n + k];
      M[i * n + k] = 0.0;
      for (int j = k + 1; j < n; j++) M[i * n + j] -= f * M[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < n; j++) s -= M[i * n + j] * b[j];
    b[i] = s / M[i * n + i];
  }
}

// eigenvalues (wr, wi) and right eigenvectors (columns of V, column-major n x n)
int small_eig(int n, const double* Arow, double* wr, double* wi, double* V) {
  std::vector<double> A(Arow, Arow + (size_t)n * n), H(A);
  if (n == 0) return 0;
  hessenberg(H, n);
  if (hqr(H, n, wr, wi) != 0) return fail(DPD_EINVAL, "subspace eigensolver did not converge");
  double anorm = 0.0;
  for (int i = 0; i < n * n; i++) anorm = std::max(anorm, std::fabs(A[i]));
  if (anorm == 0.0) anorm = 1.0;
  const double eps = 2.220446049250313e-16;
  for (int k = 0; k < n; k++) {
    std::vector<double> x(n);
    for (int i = 0; i < n; i++) x[i] = 1.0 + 0.01 * ((i * 7 + 3) % 11);
    double mu = wr[k] + eps * anorm * 4.0;
    for (int it = 0; it < 3; it++) {
      shifted_solve(A, n, mu, x, eps * anorm);
      double nrm = 0.0;
      for (int i = 0; i < n; i++) nrm += x[i] * x[i];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < n; i++) x[i] /= nrm;
    }
    for (int i = 0; i < n; i++) V[i + (size_t)k * n] = x[i];
  }
  return DPD_OK;
}

// ---------------------------------------------------------------------------
// Davidson
// ---------------------------------------------------------------------------
struct DavWork {
  DevBuf B, AB, NB, NAB, q, aq, e, p, coef, part, dots, amaxv, amaxi, scal;
};

static int dsync(dpd_ctx* c) {
  DPD_HIP_TRY(hipStreamSynchronize(c->stream));
  return DPD_OK;
}

}  // namespace dpd

using namespace dpd;

extern "C" int dpd_small_eig(int64_t n, const double* A_rowmajor, double* wr, double* wi, double* V_colmajor) {
  if (n < 0 || n > 4096 || (n && (!A_rowmajor || !wr || !wi || !V_colmajor))) return fail(DPD_EINVAL, "bad args");
  return small_eig((int)n, A_rowmajor, wr, wi, V_colmajor);
}

extern "C" int dpd_davidson(dpd_problem* P, dpd_davidson_opts* o, double* R, double* w_out) {
  if (!P || !o || !R || !w_out) return fail(DPD_EINVAL, "dpd_davidson: null argument");
  dpd_ctx* c = P->ctx;
  DPD_HIP_TRY(hipSetDevice(c->device));
  const int64_t n = P->dim;
  const int nroot = (int)o->nroot;
  const int dimAbar = (int)(o->max_dim > 0 ? o->max_dim : 20);
  const int maxcyc = (int)(o->max_cyc > 0 ? o->max_cyc : 50);
  const double maxError = o->conv > 0 ? o->conv : 1e-8;
  if (nroot < 1 || nroot > 64) return fail(DPD_EINVAL, "nroot must be 1..64");
  if (nroot >= dimAbar) return fail(DPD_EINVAL, "ERROR: too much roots required %d %d", nroot, dimAbar);
  const int dimDel = dimAbar - nroot - 5;
  hipStream_t st = c->stream;
  DavWork W;
  int rc;
  const size_t vb = sizeof(double) * (size_t)n;
  if ((rc = W.B.alloc(vb * dimAbar)) || (rc = W.AB.alloc(vb * dimAbar)) || (rc = W.NB.alloc(vb * dimAbar)) ||
      (rc = W.NAB.alloc(vb * dimAbar)) || (rc = W.q.alloc(vb)) || (rc = W.aq.alloc(vb)) || (rc = W.e.alloc(vb)) ||
      (rc = W.p.alloc(vb)) || (rc = W.coef.alloc(sizeof(double) * dimAbar * dimAbar)) ||
      (rc = W.part.alloc(sizeof(double) * kRedBlocks * dimAbar)) || (rc = W.dots.alloc(sizeof(double) * 2 * dimAbar)) ||
      (rc = W.amaxv.alloc(sizeof(double) * (kRedBlocks + 8))) || (rc = W.amaxi.alloc(sizeof(int64_t) * (kRedBlocks + 8))))
    return rc;
  double* B = (double*)W.B.p;
  double* AB = (double*)W.AB.p;
  double* q = (double*)W.q.p;
  double* aq = (double*)W.aq.p;
  double* e = (double*)W.e.p;
  double* p = (double*)W.p.p;
  double* dots = (double*)W.dots.p;
  double* part = (double*)W.part.p;
  const double* diag = (const double*)P->diag.p;
  auto bv = [&](int i) { return B + (size_t)i * n; };
  auto abv = [&](int i) { return AB + (size_t)i * n; };
  // host helpers
  auto dot_many = [&](const double* x, const double* y0, int nv, double* host) -> int {
    DPD_HIP_TRY(launch_multidot(n, x, y0, n, nv, part, dots, st));
    DPD_HIP_TRY(hipMemcpyAsync(host, dots, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
    return dsync(c);
  };
  auto nrm2 = [&](const double* x, double* out) -> int {
    double s;
    int r = dot_many(x, x, 1, &s);
    *out = std::sqrt(s);
    return r;
  };
  int64_t nsig = 0;
  auto sigma = [&](const double* x, double* y, int nv) -> int {
    nsig += nv;
    return problem_sigma(P, x, y, nv);
  };

  std::vector<double> Abar((size_t)dimAbar * dimAbar, 0.0);  // column-major Abar(i,j) = [i + j*dimAbar]
  auto A_ = [&](int i, int j) -> double& { return Abar[i + (size_t)j * dimAbar]; };
  std::vector<double> vecNeed((size_t)dimAbar * nroot, 0.0), eigenVal(nroot, 0.0), res(nroot, 0.0);
  // initDavidson (:563-610): b_i = guess_i, Ab_i = sigma(b_i), Abar(j,i) = g_j . Ab_i
  DPD_HIP_TRY(hipMemcpyAsync(B, R, vb * nroot, hipMemcpyDeviceToDevice, st));
  if ((rc = sigma(B, AB, nroot))) return rc;
  {
    std::vector<double> h(nroot);
    for (int i = 0; i < nroot; i++) {
      if ((rc = dot_many(abv(i), R, nroot, h.data()))) return rc;
      for (int j = 0; j < nroot; j++) A_(j, i) = h[j];
      vecNeed[i + (size_t)i * dimAbar] = 1.0;
    }
  }
  int iterations = 1;
  int dimIter = nroot, dimFree = dimAbar - dimIter;
  bool dim_exceed = false, converged = false;
  int ndimIter = 0, ndimFree = 0;
  std::vector<double> wr(dimAbar), wi(dimAbar), V((size_t)dimAbar * dimAbar), sub((size_t)dimAbar * dimAbar);
  std::vector<int> lowestE(dimAbar);
  while (true) {
    if (iterations > maxcyc) break;
    // eigenproblem of Abar(1:dimIter,1:dimIter) (:185-219)
    for (int i = 0; i < dimIter; i++)
      for (int j = 0; j < dimIter; j++) sub[(size_t)i * dimIter + j] = A_(i, j);
    if ((rc = small_eig(dimIter, sub.data(), wr.data(), wi.data(), V.data()))) return rc;
    auto vec = [&](int row, int col) { return V[row + (size_t)col * dimIter]; };
    // sort (:614-635) and root selection (:225-251)
    for (int i = 0; i < dimIter; i++) lowestE[i] = i;
    for (int i = 0; i < dimIter - 1; i++)
      for (int j = i + 1; j < dimIter; j++)
        if (wr[lowestE[j]] < wr[lowestE[i]]) std::swap(lowestE[i], lowestE[j]);
    if (!o->lowest) {
      for (int iroot = 0; iroot < nroot; iroot++) {
        double t1 = 0.0;
        int ii = 0;
        for (int j = 0; j < dimIter; j++) {
          double t2 = std::fabs(vec(iroot, lowestE[j]));
          if (t2 > t1) {
            t1 = t2;
            ii = j;
          }
        }
        int sel = lowestE[ii];
        for (int j = 1; j <= ii - iroot; j++) lowestE[ii - j + 1] = lowestE[ii - j];
        lowestE[iroot] = sel;
      }
    }
    for (int iroot = 0; iroot < nroot; iroot++) {
      for (int i = 0; i < dimIter; i++) vecNeed[i + (size_t)iroot * dimAbar] = vec(i, lowestE[iroot]);
      eigenVal[iroot] = wr[lowestE[iroot]];
    }
    // collapse (:268-341)
    if (dim_exceed) {
      std::vector<double> nw((size_t)dimIter * ndimIter);
      for (int k = 0; k < ndimIter; k++)
        for (int i = 0; i < dimIter; i++) nw[i + (size_t)k * dimIter] = vec(i, lowestE[k]);
      // Schmidt_2D (:640-679)
      for (int a = 1; a < ndimIter; a++) {
        for (int b = 0; b < a; b++) {
          double ov = 0.0;
          for (int i = 0; i < dimIter; i++) ov += nw[i + (size_t)b * dimIter] * nw[i + (size_t)a * dimIter];
          for (int i = 0; i < dimIter; i++) nw[i + (size_t)a * dimIter] -= ov * nw[i + (size_t)b * dimIter];
        }
        double nr = 0.0;
        for (int i = 0; i < dimIter; i++) nr += nw[i + (size_t)a * dimIter] * nw[i + (size_t)a * dimIter];
        nr = std::sqrt(nr);
        for (int i = 0; i < dimIter; i++) nw[i + (size_t)a * dimIter] /= nr;
      }
      DPD_HIP_TRY(hipMemcpyAsync(W.coef.p, nw.data(), sizeof(double) * nw.size(), hipMemcpyHostToDevice, st));
      DPD_HIP_TRY(launch_lincomb(n, AB, n, dimIter, (const double*)W.coef.p, dimIter, ndimIter, (double*)W.NAB.p, st));
      DPD_HIP_TRY(launch_lincomb(n, B, n, dimIter, (const double*)W.coef.p, dimIter, ndimIter, (double*)W.NB.p, st));
      std::swap(W.B.p, W.NB.p);
      std::swap(W.AB.p, W.NAB.p);
      B = (double*)W.B.p;
      AB = (double*)W.AB.p;
      if ((rc = dsync(c))) return rc;
      dimFree = ndimFree;
      dimIter = ndimIter;
      std::vector<double> h(dimIter);
      for (int i = 0; i < dimIter; i++) {
        if ((rc = dot_many(abv(i), B, dimIter, h.data()))) return rc;
        for (int j = 0; j < dimIter; j++) A_(j, i) = h[j];
      }
    }
    int nNewBase = 0;
    int nconv = 0;
    for (int iroot = 0; iroot < nroot; iroot++) {
      // q and Aq (:352-377)
      if (dim_exceed) {
        DPD_HIP_TRY(launch_copy(n, abv(iroot), aq, st));
        DPD_HIP_TRY(launch_copy(n, bv(iroot), q, st));
      } else {
        DPD_HIP_TRY(hipMemcpyAsync(W.coef.p, &vecNeed[(size_t)iroot * dimAbar], sizeof(double) * dimIter,
                                   hipMemcpyHostToDevice, st));
        DPD_HIP_TRY(launch_lincomb(n, AB, n, dimIter, (const double*)W.coef.p, dimIter, 1, aq, st));
        DPD_HIP_TRY(launch_lincomb(n, B, n, dimIter, (const double*)W.coef.p, dimIter, 1, q, st));
      }
      double t;
      if ((rc = nrm2(q, &t))) return rc;
      if (std::fabs(t - 1.0) > 1e-10) {
        DPD_HIP_TRY(launch_scal(n, 1.0 / t, q, st));
        DPD_HIP_TRY(launch_scal(n, 1.0 / t, aq, st));
      }
      // eigenvector written out every iteration (:388-390)
      DPD_HIP_TRY(launch_copy(n, q, R + (size_t)iroot * n, st));
      // residual (:396-406)
      DPD_HIP_TRY(launch_copy(n, aq, e, st));
      DPD_HIP_TRY(launch_axpy(n, -eigenVal[iroot], q, e, st));
      double mv;
      int64_t mi;
      DPD_HIP_TRY(launch_amax(n, e, (double*)W.amaxv.p, (int64_t*)W.amaxi.p, (double*)W.amaxv.p + kRedBlocks,
                              (int64_t*)W.amaxi.p + kRedBlocks, st));
      DPD_HIP_TRY(hipMemcpyAsync(&mv, (double*)W.amaxv.p + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
      DPD_HIP_TRY(hipMemcpyAsync(&mi, (int64_t*)W.amaxi.p + kRedBlocks, sizeof(int64_t), hipMemcpyDeviceToHost, st));
      if ((rc = dsync(c))) return rc;
      res[iroot] = std::fabs(mv);
      if (std::fabs(mv) < maxError) {
        nconv++;
        continue;
      }
      if (dimFree == 0) continue;
      // preconditioner (:417-428)
      DPD_HIP_TRY(launch_precondition(n, eigenVal[iroot], diag, e, p, st));
      // SchmidtS (:687-726): normalise, modified Gram-Schmidt, normalise
      double nrm;
      if ((rc = nrm2(p, &nrm))) return rc;
      DPD_HIP_TRY(launch_scal(n, 1.0 / nrm, p, st));
      for (int i = 0; i < dimIter + nNewBase; i++) {
        DPD_HIP_TRY(launch_multidot(n, p, bv(i), n, 1, part, dots, st));
        DPD_HIP_TRY(launch_axpy_dev(n, dots, -1.0, bv(i), p, st));
      }
      if ((rc = nrm2(p, &nrm))) return rc;
      if (nrm < 1e-6) continue;
      DPD_HIP_TRY(launch_scal(n, 1.0 / nrm, p, st));
      DPD_HIP_TRY(launch_copy(n, p, bv(dimIter + nNewBase), st));
      dimFree--;
      nNewBase++;
    }
    // sigma of this iteration's new vectors, one batched evaluation (:442-444)
    if (nNewBase > 0)
      if ((rc = sigma(bv(dimIter), abv(dimIter), nNewBase))) return rc;
    if (nconv == nroot) {
      converged = true;
      break;
    }
    if (dim_exceed) dim_exceed = false;
    if (dimFree == 0) {
      dim_exceed = true;
      ndimIter = dimAbar - dimDel;
      if (ndimIter < nroot) {
        ndimIter = nroot;
        ndimFree = dimAbar - ndimIter;
      } else {
        ndimFree = dimDel;
      }
    }
    // Abar update (:484-509)
    if (nNewBase > 0) {
      std::vector<double> h(dimIter + nNewBase);
      for (int j = 0; j < dimIter; j++) {
        if ((rc = dot_many(abv(j), bv(dimIter), nNewBase, h.data()))) return rc;
        for (int i = 0; i < nNewBase; i++) A_(dimIter + i, j) = h[i];
      }
      for (int j = dimIter; j < dimIter + nNewBase; j++) {
        if ((rc = dot_many(abv(j), B, dimIter + nNewBase, h.data()))) return rc;
        for (int i = 0; i < dimIter + nNewBase; i++) A_(i, j) = h[i];
      }
    }
    dimIter += nNewBase;
    iterations++;
  }
  for (int i = 0; i < nroot; i++) w_out[i] = eigenVal[i];
  o->iterations = iterations;
  for (int i = 0; i < nroot && i < 64; i++) o->residual[i] = res[i];
  // final check (:527-542): e = A R - w R
  if (o->final_check) {
    if ((rc = sigma(R, (double*)W.NAB.p, nroot))) return rc;
    for (int i = 0; i < nroot; i++) {
      double* ev = (double*)W.NAB.p + (size_t)i * n;
      DPD_HIP_TRY(launch_axpy(n, -eigenVal[i], R + (size_t)i * n, ev, st));
      double mv;
      DPD_HIP_TRY(launch_amax(n, ev, (double*)W.amaxv.p, (int64_t*)W.amaxi.p, (double*)W.amaxv.p + kRedBlocks,
                              (int64_t*)W.amaxi.p + kRedBlocks, st));
      DPD_HIP_TRY(hipMemcpyAsync(&mv, (double*)W.amaxv.p + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
      if ((rc = dsync(c))) return rc;
      if (i < 64) o->final_residual[i] = std::fabs(mv);
    }
  }
  o->sigma_calls = nsig;
  if ((rc = dsync(c))) return rc;
  return converged ? DPD_OK : DPD_ENOTCONV;
}
