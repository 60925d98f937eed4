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
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "dpd_sigma.hpp"

namespace dpd {

// ---------------------------------------------------------------------------
// small dense nonsymmetric eigenproblem (FP64, row-major n x n)
// Householder reduction A = Q H Q^T to upper Hessenberg form (Golub & Van Loan,
// Algorithm 7.4.2; the reflectors are kept, Q is never formed), the Francis
// double-shift QR algorithm for the eigenvalues, and for the requested columns
// inverse iteration on H (one LU of the shifted H per vector, O(n^2)) mapped
// back by the reflectors; vectors have unit 2-norm (as dgeev's).
// ---------------------------------------------------------------------------
// reflector k (k < n - 2): I - tau[k] v v^T with v[k+1..n) in column k of `vr`
// (row-major n x n, entry (i, k)), zero elsewhere; tau[k] = 0: no reflection
static void hessenberg(std::vector<double>& a, int n, std::vector<double>& vr, std::vector<double>& tau) {
  vr.assign((size_t)n * n, 0.0);
  tau.assign((size_t)std::max(n, 1), 0.0);
  std::vector<double> v(n), w(n);
  for (int k = 0; k < n - 2; k++) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; i++) alpha += a[i * n + k] * a[i * n + k];
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (a[(k + 1) * n + k] > 0) alpha = -alpha;
    v[k + 1] = a[(k + 1) * n + k] - alpha;
    for (int i = k + 2; i < n; i++) v[i] = a[i * n + k];
    double vv = 0.0;
    for (int i = k + 1; i < n; i++) vv += v[i] * v[i];
    if (vv == 0.0) continue;
    const double tk = 2.0 / vv;
    tau[k] = tk;
    for (int i = k + 1; i < n; i++) vr[(size_t)i * n + k] = v[i];
    // A = H A H with H = I - tau v v^T.  From the left on columns k..n-1 (the
    // columns before k are zero in rows k+1..n-1): w^T = v^T A accumulated row
    // by row (contiguous), then A -= tau v w^T
    for (int j = k; j < n; j++) w[j] = 0.0;
    for (int i = k + 1; i < n; i++) {
      const double vi = v[i];
      const double* ai = &a[(size_t)i * n];
      for (int j = k; j < n; j++) w[j] += vi * ai[j];
    }
    for (int i = k + 1; i < n; i++) {
      const double f = tk * v[i];
      double* ai = &a[(size_t)i * n];
      for (int j = k; j < n; j++) ai[j] -= f * w[j];
    }
    // from the right, row by row: A(i,:) -= tau (A(i,:) v) v^T
    for (int i = 0; i < n; i++) {
      double* ai = &a[(size_t)i * n];
      double s = 0.0;
      for (int j = k + 1; j < n; j++) s += ai[j] * v[j];
      s *= tk;
      for (int j = k + 1; j < n; j++) ai[j] -= s * v[j];
    }
    for (int i = k + 2; i < n; i++) a[i * n + k] = 0.0;
  }
}

// Householder vector of a 2- or 3-vector x (Golub & Van Loan, Matrix
// Computations, Algorithm 5.1.1): v[0] = 1 and beta such that
// (I - beta v v^T) x is a multiple of e_1; beta = 0 when x already is one.
static void house_small(const double* x, int len, double* v, double& beta) {
  double sigma = 0.0;
  for (int i = 1; i < len; i++) sigma += x[i] * x[i];
  v[0] = 1.0;
  for (int i = 1; i < len; i++) v[i] = x[i];
  if (sigma == 0.0) {
    beta = 0.0;
    return;
  }
  const double mu = std::sqrt(x[0] * x[0] + sigma);
  const double v0 = x[0] <= 0.0 ? x[0] - mu : -sigma / (x[0] + mu);
  beta = 2.0 * v0 * v0 / (sigma + v0 * v0);
  for (int i = 1; i < len; i++) v[i] /= v0;
}

// rows r0..r0+len-1 of H, columns c0..c1:  H <- (I - beta v v^T) H
static void reflect_rows(std::vector<double>& H, int n, int r0, int len, int c0, int c1, const double* v,
                         double beta) {
  if (beta == 0.0) return;
  for (int j = c0; j <= c1; j++) {
    double s = 0.0;
    for (int i = 0; i < len; i++) s += v[i] * H[(size_t)(r0 + i) * n + j];
    s *= beta;
    for (int i = 0; i < len; i++) H[(size_t)(r0 + i) * n + j] -= s * v[i];
  }
}

// columns c0..c0+len-1 of H, rows r0..r1:  H <- H (I - beta v v^T)
static void reflect_cols(std::vector<double>& H, int n, int c0, int len, int r0, int r1, const double* v,
                         double beta) {
  if (beta == 0.0) return;
  for (int i = r0; i <= r1; i++) {
    double* hi = &H[(size_t)i * n + c0];
    double s = 0.0;
    for (int j = 0; j < len; j++) s += hi[j] * v[j];
    s *= beta;
    for (int j = 0; j < len; j++) hi[j] -= s * v[j];
  }
}

// One implicit double-shift (Francis) QR step on the unreduced window
// lo..hi (hi - lo >= 2) of the upper Hessenberg matrix H, with the shift pair
// given as the trace `tr` and determinant `det` of its 2x2 generator (Golub &
// Van Loan, Algorithm 7.5.1).  Only the window is transformed: the eigenvalues
// are all that is wanted, and the window's are independent of the rest of H
// once the subdiagonal entries around it are zero.
static void francis_step(std::vector<double>& H, int n, int lo, int hi, double tr, double det) {
  auto h = [&](int i, int j) -> double& { return H[(size_t)i * n + j]; };
  // first column of (H - s1 I)(H - s2 I), which has three nonzeros
  double x[3];
  x[0] = h(lo, lo) * h(lo, lo) + h(lo, lo + 1) * h(lo + 1, lo) - tr * h(lo, lo) + det;
  x[1] = h(lo + 1, lo) * (h(lo, lo) + h(lo + 1, lo + 1) - tr);
  x[2] = h(lo + 1, lo) * h(lo + 2, lo + 1);
  double v[3], beta;
  for (int k = lo; k <= hi - 2; k++) {
    house_small(x, 3, v, beta);
    reflect_rows(H, n, k, 3, std::max(lo, k - 1), hi, v, beta);
    reflect_cols(H, n, k, 3, lo, std::min(k + 3, hi), v, beta);
    if (k > lo) h(k + 1, k - 1) = h(k + 2, k - 1) = 0.0;  // the bulge has moved on
    x[0] = h(k + 1, k);
    x[1] = h(k + 2, k);
    x[2] = k < hi - 2 ? h(k + 3, k) : 0.0;
  }
  house_small(x, 2, v, beta);
  reflect_rows(H, n, hi - 1, 2, hi - 2, hi, v, beta);
  reflect_cols(H, n, hi - 1, 2, lo, hi, v, beta);
  h(hi, hi - 2) = 0.0;
}

// eigenvalues of the 2x2 block [[a, b], [c, d]]: with p = (a - d)/2 they are
// d + mu for the roots mu of mu^2 - 2 p mu - b c = 0; the root of larger size
// is formed without cancellation and the other from the product -b c.
static void eig2x2(double a, double b, double c, double d, double* re, double* im) {
  const double p = 0.5 * (a - d);
  const double disc = p * p + b * c;
  if (disc >= 0.0) {
    const double t = p + std::copysign(std::sqrt(disc), p);
    re[0] = d + t;
    re[1] = t != 0.0 ? d - b * c / t : d;
    im[0] = im[1] = 0.0;
  } else {
    re[0] = re[1] = d + p;
    im[0] = std::sqrt(-disc);
    im[1] = -im[0];
  }
}

// Eigenvalues (wr, wi) of an upper Hessenberg matrix by the Francis QR
// algorithm (Golub & Van Loan, Algorithm 7.5.2): the active window is the
// trailing unreduced block; a subdiagonal entry below eps times its two
// diagonal neighbours is set to zero, a 1x1 or 2x2 trailing block is read off
// and removed.  The standard shifts are the eigenvalues of the window's
// trailing 2x2 block; after every 10 steps without a deflation the ad hoc
// shift pair of EISPACK's hqr (public domain) is used once: mu +- i nu with
// mu = h(hi,hi) + 0.75 s and nu^2 = 0.4375 s^2, s = |h(hi,hi-1)| + |h(hi-1,hi-2)|.
static int francis_eigenvalues(std::vector<double>& H, int n, double* wr, double* wi) {
  auto h = [&](int i, int j) -> double& { return H[(size_t)i * n + j]; };
  const double eps = 2.220446049250313e-16;
  double hnorm = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = std::max(i - 1, 0); j < n; j++) hnorm = std::max(hnorm, std::fabs(h(i, j)));
  if (hnorm == 0.0) hnorm = 1.0;
  int hi = n - 1, steps = 0;
  const int max_steps = 30 * std::max(n, 1);
  int since_deflation = 0;
  while (hi >= 0) {
    int lo = hi;
    for (; lo > 0; lo--) {
      double s = std::fabs(h(lo - 1, lo - 1)) + std::fabs(h(lo, lo));
      if (s == 0.0) s = hnorm;
      if (std::fabs(h(lo, lo - 1)) <= eps * s) {
        h(lo, lo - 1) = 0.0;
        break;
      }
    }
    if (lo == hi) {
      wr[hi] = h(hi, hi);
      wi[hi] = 0.0;
      hi -= 1;
      since_deflation = 0;
      continue;
    }
    if (lo == hi - 1) {
      double re[2], im[2];
      eig2x2(h(hi - 1, hi - 1), h(hi - 1, hi), h(hi, hi - 1), h(hi, hi), re, im);
      wr[hi - 1] = re[0];
      wi[hi - 1] = im[0];
      wr[hi] = re[1];
      wi[hi] = im[1];
      hi -= 2;
      since_deflation = 0;
      continue;
    }
    if (++steps > max_steps) return -1;
    double tr, det;
    if (++since_deflation % 10 == 0) {
      const double s = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2));
      const double mu = h(hi, hi) + 0.75 * s;
      tr = 2.0 * mu;
      det = mu * mu + 0.4375 * s * s;
    } else {
      tr = h(hi - 1, hi - 1) + h(hi, hi);
      det = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
    }
    francis_step(H, n, lo, hi, tr, det);
  }
  return 0;
}

// LU factors of (H - mu I) for upper Hessenberg H with partial pivoting between
// neighbouring rows (O(n^2)): U in M (row-major), the multiplier of step k in
// f[k] and whether rows k, k+1 were swapped in sw[k]; tiny pivots are replaced
// by +-tiny
static void hess_factor(const std::vector<double>& H, int n, double mu, double tiny, std::vector<double>& M,
                        std::vector<double>& f, std::vector<char>& sw) {
  M = H;
  f.assign((size_t)std::max(n, 1), 0.0);
  sw.assign((size_t)std::max(n, 1), 0);
  for (int i = 0; i < n; i++) M[(size_t)i * n + i] -= mu;
  for (int k = 0; k < n - 1; k++) {
    double* rk = &M[(size_t)k * n];
    double* r1 = &M[(size_t)(k + 1) * n];
    if (std::fabs(r1[k]) > std::fabs(rk[k])) {
      for (int j = k; j < n; j++) std::swap(rk[j], r1[j]);
      sw[k] = 1;
    }
    if (std::fabs(rk[k]) < tiny) rk[k] = rk[k] < 0 ? -tiny : tiny;
    const double fk = r1[k] / rk[k];
    f[k] = fk;
    r1[k] = 0.0;
    for (int j = k + 1; j < n; j++) r1[j] -= fk * rk[j];
  }
  double* rl = &M[(size_t)(n - 1) * n];
  if (n > 0 && std::fabs(rl[n - 1]) < tiny) rl[n - 1] = rl[n - 1] < 0 ? -tiny : tiny;
}

// b <- (H - mu I)^-1 b from hess_factor's factors
static void hess_apply(const std::vector<double>& M, int n, const std::vector<double>& f,
                       const std::vector<char>& sw, double* b) {
  for (int k = 0; k < n - 1; k++) {
    if (sw[k]) std::swap(b[k], b[k + 1]);
    b[k + 1] -= f[k] * b[k];
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    const double* ri = &M[(size_t)i * n];
    for (int j = i + 1; j < n; j++) s -= ri[j] * b[j];
    b[i] = s / ri[i];
  }
}

// two-phase solver: eigenvalues first, then eigenvectors of chosen real eigenvalues
struct SmallEig {
  int n = 0;
  std::vector<double> H, VR, tau, T, y, M, f;
  std::vector<char> sw;
  double anorm = 1.0;
  const double* wr = nullptr;
  int values(int n_, const double* Arow, double* wr_, double* wi_) {
    n = n_;
    wr = wr_;
    if (n == 0) return DPD_OK;
    H.assign(Arow, Arow + (size_t)n * n);
    hessenberg(H, n, VR, tau);
    T = H;
    if (francis_eigenvalues(T, n, wr_, wi_) != 0) return fail(DPD_EINVAL, "subspace eigensolver did not converge");
    anorm = 0.0;
    for (int i = 0; i < n * n; i++) anorm = std::max(anorm, std::fabs(Arow[i]));
    if (anorm == 0.0) anorm = 1.0;
    return DPD_OK;
  }
  // unit right eigenvector of eigenvalue k into x[0..n)
  void vector(int k, double* x) {
    const double eps = 2.220446049250313e-16;
    y.resize(n);
    for (int i = 0; i < n; i++) y[i] = 1.0 + 0.01 * ((i * 7 + 3) % 11);
    const double mu = wr[k] + eps * anorm * 4.0;
    hess_factor(H, n, mu, eps * anorm, M, f, sw);
    for (int it = 0; it < 3; it++) {
      hess_apply(M, n, f, sw, y.data());
      double nrm = 0.0;
      for (int i = 0; i < n; i++) nrm += y[i] * y[i];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < n; i++) y[i] /= nrm;
    }
    // x = Q y = H_0 H_1 ... H_{n-3} y (orthogonal: unit norm kept)
    for (int kk = n - 3; kk >= 0; kk--) {
      const double tk = tau[kk];
      if (tk == 0.0) continue;
      double s = 0.0;
      for (int i = kk + 1; i < n; i++) s += VR[(size_t)i * n + kk] * y[i];
      s *= tk;
      for (int i = kk + 1; i < n; i++) y[i] -= s * VR[(size_t)i * n + kk];
    }
    for (int i = 0; i < n; i++) x[i] = y[i];
  }
};

// eigenvalues (wr, wi) of the row-major matrix and right eigenvectors of every
// eigenvalue (columns of V, column-major n x n)
int small_eig(int n, const double* Arow, double* wr, double* wi, double* V) {
  SmallEig E;
  int rc = E.values(n, Arow, wr, wi);
  if (rc) return rc;
  for (int k = 0; k < n; k++) E.vector(k, V + (size_t)k * n);
  return DPD_OK;
}

// ---------------------------------------------------------------------------
// Davidson
// ---------------------------------------------------------------------------
// buf holds at least `bytes` (reallocated only to grow)
static int grow(DevBuf& buf, size_t bytes) {
  if (buf.p && buf.bytes >= bytes) return DPD_OK;
  buf.reset();
  return buf.alloc(bytes);
}

static int dsync(hipStream_t s) {
  DPD_HIP_TRY(hipStreamSynchronize(s));
  return DPD_OK;
}

}  // namespace dpd

using namespace dpd;

extern "C" int dpd_small_eig(int64_t n, const double* A_rowmajor, double* wr, double* wi, double* V_colmajor) {
  if (n < 0 || n > 4096 || (n && (!A_rowmajor || !wr || !wi || !V_colmajor))) return fail(DPD_EINVAL, "bad args");
  return small_eig((int)n, A_rowmajor, wr, wi, V_colmajor);
}

// slots at the start of the transfer buffer for the norms of the new directions
constexpr size_t kNormSlots = 64;

// sigma of nv vectors for one Davidson; null: the problem's own operator
typedef std::function<int(const double* x, double* y, int nv)> SigmaHook;

// stream: where the solver's own work goes (null: the context's stream, which
// the problem's operator uses too)
static int davidson_run(dpd_problem* P, dpd_davidson_opts* o, double* R, double* w_out, const SigmaHook* hook,
                        hipStream_t stream);

static int davidson_graph(dpd_problem* P, dpd_davidson_opts* o, double* R, double* w_out);

extern "C" int dpd_davidson(dpd_problem* P, dpd_davidson_opts* o, double* R, double* w_out) {
  if (o) o->graph_replays = 0;
  if (o && o->graph && P) return davidson_graph(P, o, R, w_out);
  return davidson_run(P, o, R, w_out, nullptr, nullptr);
}

static int davidson_run(dpd_problem* P, dpd_davidson_opts* o, double* R, double* w_out, const SigmaHook* hook,
                        hipStream_t stream) {
  if (!P || !o || !R || !w_out) return fail(DPD_EINVAL, "dpd_davidson: null argument");
  if (P->stale) return fail(DPD_EINVAL, "problem inputs changed: call dpd_problem_refresh first");
  dpd_ctx* c = P->ctx;
  DPD_HIP_TRY(hipSetDevice(c->device));
  const int64_t n = P->dim;
  const int nroot = (int)o->nroot;
  const int dimAbar = (int)(o->max_dim > 0 ? o->max_dim : 20);
  const int maxcyc = (int)(o->max_cyc > 0 ? o->max_cyc : 50);
  const double maxError = o->conv > 0 ? o->conv : 1e-8;
  if (nroot < 1 || nroot > 64) return fail(DPD_EINVAL, "nroot must be 1..64");
  if (nroot >= dimAbar) return fail(DPD_EINVAL, "ERROR: too much roots required %d %d", nroot, dimAbar);
  if (dimAbar > 1024) return fail(DPD_EINVAL, "max_dim must be <= 1024");
  const int dimDel = dimAbar - nroot - 5;
  hipStream_t st = stream ? stream : c->stream;
  // the workspace of a single solve stays with the problem between solves
  // (allocation and the pinned staging cost about 3 ms per solve otherwise);
  // a member of a batched solve (own stream, run from a worker thread) takes
  // a workspace of its own for the call
  std::unique_ptr<DavWork> own;
  if (stream) {
    own.reset(new DavWork);
  } else if (!P->dav) {
    P->dav.reset(new DavWork);
  }
  DavWork& W = stream ? *own : *P->dav;
  // a call's own workspace is freed only after its last transfer has completed
  struct SyncOnExit {
    hipStream_t s;
    bool on;
    ~SyncOnExit() {
      if (on) (void)hipStreamSynchronize(s);
    }
  } sync_on_exit{st, (bool)own};
  int rc;
  const size_t vb = sizeof(double) * (size_t)n;
  const size_t nd2 = (size_t)dimAbar * dimAbar;
  // B/NB and AB/NAB swap pointers at a collapse: each pair is grown to one size
  const size_t subb = std::max({vb * dimAbar, W.B.bytes, W.NB.bytes, W.AB.bytes, W.NAB.bytes});
  if ((rc = grow(W.B, subb)) || (rc = grow(W.AB, subb)) || (rc = grow(W.NB, subb)) || (rc = grow(W.NAB, subb)) ||
      (rc = grow(W.AQ, vb * nroot)) || (rc = grow(W.E, vb * nroot)) || (rc = grow(W.P, vb * nroot)) ||
      (rc = grow(W.coef, sizeof(double) * std::max<size_t>(nd2, (size_t)kMgsMulti * dimAbar))) ||
      (rc = grow(W.part, sizeof(double) * kRedBlocks * std::max<size_t>(nd2, (size_t)kMgsMulti * dimAbar))) ||
      (rc = grow(W.partA, sizeof(double) * kRedBlocks * kMgsMulti)) ||
      (rc = grow(W.partB, sizeof(double) * kRedBlocks * kMgsMulti)) ||
      (rc = grow(W.partC, sizeof(double) * kRedBlocks)) || (rc = grow(W.partD, sizeof(double) * kRedBlocks)) ||
      (rc = grow(W.dots, sizeof(double) * (2 * nd2 + kNormSlots))) ||
      (rc = grow(W.amaxp, sizeof(double) * kRedBlocks * nroot)) ||
      (rc = grow(W.amaxpi, sizeof(int64_t) * kRedBlocks * nroot)) || (rc = grow(W.amaxv, sizeof(double) * nroot)) ||
      (rc = grow(W.amaxi, sizeof(int64_t) * nroot)) || (rc = grow(W.wdev, sizeof(double) * nroot)) ||
      (rc = grow(W.idx, sizeof(int32_t) * nroot)) || (rc = grow(W.cgs, 256)) ||
      (rc = grow(W.tri, sizeof(double) * (2 * kMgsMulti * kMgsMulti + kMgsMulti))) ||
      (rc = grow(W.gram, sizeof(double) * 2 * kRedBlocks * kMgsMulti * kMgsMulti)) ||
      (rc = grow(W.qpart, sizeof(double) * kRedBlocks * nroot)) ||
      (rc = grow(W.ppart, sizeof(double) * kRedBlocks * nroot)) || (rc = grow(W.rstat, sizeof(double) * 3 * nroot)))
    return rc;
  double* B = (double*)W.B.p;
  double* AB = (double*)W.AB.p;
  double* AQ = (double*)W.AQ.p;
  double* E = (double*)W.E.p;
  double* PV = (double*)W.P.p;
  double* dots = (double*)W.dots.p;
  double* part = (double*)W.part.p;
  double* wdev = (double*)W.wdev.p;
  double* rstat = (double*)W.rstat.p;  // per root: max|e|, |q|^2, |p|^2 of the last Ritz pass
  double* gram = (double*)W.gram.p;
  const double* diag = (const double*)P->diag.p;
  auto bv = [&](int i) { return B + (size_t)i * n; };
  auto abv = [&](int i) { return AB + (size_t)i * n; };
  int64_t nsig = 0, nskip = 0;
  // small host<->device transfers go through pinned buffers: device-to-host
  // copies land in pin_out and are read after the synchronisation; uploads are
  // staged in pin_in slots that stay untouched until the next synchronisation
  const size_t pin_bytes = sizeof(double) * (4 * nd2 + 4 * (size_t)dimAbar * nroot + 1024);
  if (W.pin_bytes < pin_bytes) {
    // no transfer of an earlier solve is in flight: every solve ends synchronised
    if (W.pin_out) (void)hipHostFree(W.pin_out);
    if (W.pin_in) (void)hipHostFree(W.pin_in);
    W.pin_out = W.pin_in = nullptr;
    W.pin_bytes = 0;
    DPD_HIP_TRY(hipHostMalloc(&W.pin_out, pin_bytes, hipHostMallocDefault));
    DPD_HIP_TRY(hipHostMalloc(&W.pin_in, pin_bytes, hipHostMallocDefault));
    W.pin_bytes = pin_bytes;
  }
  struct {
    void* p;
    size_t bytes;
  } pin_out{W.pin_out, W.pin_bytes}, pin_in{W.pin_in, W.pin_bytes};
  size_t in_top = 0;
  auto fetch = [&](void* host, const void* dev, size_t bytes) -> int {
    if (bytes > pin_out.bytes) return fail(DPD_EINVAL, "transfer too large (internal)");
    if (bytes) DPD_HIP_TRY(hipMemcpyAsync(pin_out.p, dev, bytes, hipMemcpyDeviceToHost, st));
    int r = dsync(st);
    if (r) return r;
    if (bytes) std::memcpy(host, pin_out.p, bytes);
    in_top = 0;  // every staged upload has completed
    return DPD_OK;
  };
  auto upload = [&](void* dev, const void* host, size_t bytes) -> int {
    if (bytes > pin_in.bytes) return fail(DPD_EINVAL, "transfer too large (internal)");
    if (in_top + bytes > pin_in.bytes) {
      int r = dsync(st);
      if (r) return r;
      in_top = 0;
    }
    void* slot = (char*)pin_in.p + in_top;
    std::memcpy(slot, host, bytes);
    DPD_HIP_TRY(hipMemcpyAsync(dev, slot, bytes, hipMemcpyHostToDevice, st));
    in_top += (bytes + 255) & ~size_t(255);
    return DPD_OK;
  };
  auto sigma = [&](const double* x, double* y, int nv) -> int {
    nsig += nv;
    if (hook) return (*hook)(x, y, nv);
    return o->transpose ? problem_sigma_t(P, x, y, nv) : problem_sigma(P, x, y, nv);
  };
  // all dots X_i . Y_j (i < nx, j < ny) -> host[i + j*nx], one synchronisation
  auto block_dots = [&](const double* X, int nx, const double* Y, int ny, double* host) -> int {
    DPD_HIP_TRY(launch_block_dots(n, X, n, nx, Y, n, ny, part, dots, st));
    return fetch(host, dots, sizeof(double) * nx * ny);
  };
  // e_v = A r_v - w_v r_v and max|e_v| for v < nroot (w already on the device)
  auto residuals = [&](const double* AR, double* res_host) -> int {
    DPD_HIP_TRY(launch_resid_amax(n, AR, R, wdev, E, nroot, (double*)W.amaxp.p, (int64_t*)W.amaxpi.p,
                                  (double*)W.amaxv.p, (int64_t*)W.amaxi.p, st));
    return fetch(res_host, W.amaxv.p, sizeof(double) * nroot);
  };

  std::vector<double> Abar(nd2, 0.0);  // column-major Abar(i,j) = [i + j*dimAbar]
  auto A_ = [&](int i, int j) -> double& { return Abar[i + (size_t)j * dimAbar]; };
  std::vector<double> vecNeed((size_t)dimAbar * nroot, 0.0), eigenVal(nroot, 0.0), res(nroot, 0.0);
  std::vector<double> hd(2 * nd2 + kNormSlots);
  // initDavidson (:563-610): b_i = guess_i, Ab_i = sigma(b_i), Abar(j,i) = g_j . Ab_i
  DPD_HIP_TRY(hipMemcpyAsync(B, R, vb * nroot, hipMemcpyDeviceToDevice, st));
  if ((rc = sigma(B, AB, nroot))) return rc;
  if ((rc = block_dots(AB, nroot, B, nroot, hd.data()))) return rc;
  for (int i = 0; i < nroot; i++) {
    for (int j = 0; j < nroot; j++) A_(j, i) = hd[i + (size_t)j * nroot];
    vecNeed[i + (size_t)i * dimAbar] = 1.0;
  }
  int iterations = 1;
  int dimIter = nroot, dimFree = dimAbar - dimIter;
  bool dim_exceed = false, converged = false;
  int ndimIter = 0, ndimFree = 0;
  std::vector<double> wr(dimAbar), wi(dimAbar), V(nd2), sub(nd2), mv(2 * nroot);
  bool r_unscaled = false;  // R of the last Ritz pass still awaits its renormalisation
  std::vector<int> lowestE(dimAbar);
  std::vector<int32_t> cand;
  SmallEig eig;
  while (true) {
    if (iterations > maxcyc) break;
    // eigenproblem of Abar(1:dimIter,1:dimIter) (:185-219)
    for (int i = 0; i < dimIter; i++)
      for (int j = 0; j < dimIter; j++) sub[(size_t)i * dimIter + j] = A_(i, j);
    if ((rc = eig.values(dimIter, sub.data(), wr.data(), wi.data()))) return rc;
    auto vec = [&](int row, int col) { return V[row + (size_t)col * dimIter]; };
    // sort (:614-635) and root selection (:225-251)
    for (int i = 0; i < dimIter; i++) lowestE[i] = i;
    for (int i = 0; i < dimIter - 1; i++)
      for (int j = i + 1; j < dimIter; j++)
        if (wr[lowestE[j]] < wr[lowestE[i]]) std::swap(lowestE[i], lowestE[j]);
    // eigenvectors: every one when roots are tracked by overlap, else only the
    // lowest nroot (nroot + 5 when the subspace collapses)
    {
      const int need = o->lowest ? std::min(dimIter, dim_exceed ? std::max(nroot, ndimIter) : nroot) : dimIter;
      for (int j = 0; j < need; j++) eig.vector(lowestE[j], V.data() + (size_t)lowestE[j] * dimIter);
    }
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
    // collapse (:268-341); ndimIter = nroot + 5 and the reference reads past
    // the subspace when davidson_dim < nroot + 5: an input error here
    if (dim_exceed && ndimIter > dimIter)
      return fail(DPD_EINVAL, "subspace collapse needs davidson_dim >= nroot + 5 (%d < %d)", dimAbar, ndimIter);
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
      if ((rc = upload(W.coef.p, nw.data(), sizeof(double) * nw.size()))) return rc;
      DPD_HIP_TRY(launch_lincomb(n, AB, n, dimIter, (const double*)W.coef.p, dimIter, ndimIter, (double*)W.NAB.p, st));
      DPD_HIP_TRY(launch_lincomb(n, B, n, dimIter, (const double*)W.coef.p, dimIter, ndimIter, (double*)W.NB.p, st));
      std::swap(W.B.p, W.NB.p);
      std::swap(W.AB.p, W.NAB.p);
      B = (double*)W.B.p;
      AB = (double*)W.AB.p;
      dimFree = ndimFree;
      dimIter = ndimIter;
      if ((rc = block_dots(AB, dimIter, B, dimIter, hd.data()))) return rc;
      for (int i = 0; i < dimIter; i++)
        for (int j = 0; j < dimIter; j++) A_(j, i) = hd[i + (size_t)j * dimIter];
    }
    // q and Aq of every root (:352-377); q is written straight into R, the
    // eigenvector output the reference rewrites every iteration (:388-390)
    if (dim_exceed) {
      DPD_HIP_TRY(hipMemcpyAsync(R, B, vb * nroot, hipMemcpyDeviceToDevice, st));
      DPD_HIP_TRY(hipMemcpyAsync(AQ, AB, vb * nroot, hipMemcpyDeviceToDevice, st));
    } else if ((int64_t)dimIter * nroot <= kLinCoef) {
      // the coefficients travel in the kernel arguments
      LinCoef cf;
      for (int j = 0; j < nroot; j++)
        for (int k = 0; k < dimIter; k++) cf.v[k + j * dimIter] = vecNeed[k + (size_t)j * dimAbar];
      DPD_HIP_TRY(launch_lincomb2_v(n, AB, B, n, dimIter, cf, nroot, AQ, R, st));
    } else {
      if ((rc = upload(W.coef.p, vecNeed.data(), sizeof(double) * dimAbar * nroot))) return rc;
      DPD_HIP_TRY(launch_lincomb2(n, AB, B, n, dimIter, (const double*)W.coef.p, dimAbar, nroot, AQ, R, st));
    }
    // in one pass: the residuals e = Aq - w q and their largest elements
    // (:396-406), the preconditioned directions of every root (:417-428), kept
    // for the unconverged ones, and |q|^2 and |p|^2.  The renormalisation of q
    // and Aq (:378-386, by 1/|q| when ||q| - 1| > 1e-10) is applied afterwards:
    // e and max|e| scale with it (the host scales max|e|), the directions are
    // normalised anyway, and q (the eigenvector output R) is rescaled on exit.
    // e itself is not stored: the rare redo below forms it again from q and Aq
    RootVals wv;
    for (int v = 0; v < nroot; v++) wv.w[v] = eigenVal[v];
    DPD_HIP_TRY(launch_ritz_resid_part(n, wv, diag, R, AQ, PV, nroot, (double*)W.amaxp.p, (int64_t*)W.amaxpi.p,
                                       (double*)W.qpart.p, (double*)W.ppart.p, rstat, (int64_t*)W.amaxi.p, st));
    bool roots_on_device = false;  // eigenvalues and candidates uploaded (the rare redo path needs them)
    if ((rc = fetch(mv.data(), rstat, sizeof(double) * 2 * nroot))) return rc;
    int nconv = 0;
    cand.clear();
    r_unscaled = false;
    for (int iroot = 0; iroot < nroot; iroot++) {
      const double t = std::sqrt(mv[nroot + iroot]);
      double m = std::fabs(mv[iroot]);
      if (std::fabs(t - 1.0) > 1e-10) {
        m /= t;
        r_unscaled = true;
      }
      res[iroot] = m;
      if (m < maxError) nconv++;
      else cand.push_back(iroot);
    }
    // preconditioned directions of every unconverged root (:417-428)
    int nNewBase = 0;
    size_t spec = 0;  // directions orthogonalised this iteration, assumed kept
    // new Abar entries of nnew new vectors: rows against the old columns, then
    // the new columns against every basis vector (each basis vector is read once
    // per group of up to kMgsMulti new vectors); written to dots from `off` on
    auto abar_dots = [&](int nnew, size_t& off, std::vector<size_t>& off1, std::vector<size_t>& off2) -> int {
      const int nt = dimIter + nnew;
      if (nnew <= kMgsMulti) {
        // one launch for both: the second block lands right after the first
        DPD_HIP_TRY(launch_dots_pair(n, nnew, bv(dimIter), AB, dimIter, abv(dimIter), B, nt, part, dots + off, st));
        off1.push_back(off);
        off += (size_t)nnew * dimIter;
        off2.push_back(off);
        off += (size_t)nnew * nt;
        return DPD_OK;
      }
      for (int c0 = 0; c0 < nnew; c0 += kMgsMulti) {
        const int nv = std::min(kMgsMulti, nnew - c0);
        DPD_HIP_TRY(launch_dots_multi(n, bv(dimIter + c0), nv, AB, dimIter, part, dots + off, st));
        off1.push_back(off);
        off += (size_t)nv * dimIter;
      }
      for (int c0 = 0; c0 < nnew; c0 += kMgsMulti) {
        const int nv = std::min(kMgsMulti, nnew - c0);
        DPD_HIP_TRY(launch_dots_multi(n, abv(dimIter + c0), nv, B, nt, part, dots + off, st));
        off2.push_back(off);
        off += (size_t)nv * nt;
      }
      return DPD_OK;
    };
    auto abar_fill = [&](int nnew, const std::vector<size_t>& off1, const std::vector<size_t>& off2) {
      const int nt = dimIter + nnew;
      for (int c0 = 0, ci = 0; c0 < nnew; c0 += kMgsMulti, ci++) {
        const int nv = std::min(kMgsMulti, nnew - c0);
        for (int j = 0; j < dimIter; j++)
          for (int i = 0; i < nv; i++) A_(dimIter + c0 + i, j) = hd[off1[ci] + i + (size_t)j * nv];
        for (int j = 0; j < nv; j++)
          for (int i = 0; i < nt; i++) A_(i, dimIter + c0 + j) = hd[off2[ci] + j + (size_t)i * nv];
      }
    };
    if (dimFree > 0 && !cand.empty()) {
      // the candidates' directions moved to the front of PV (root cand[k] to
      // slot k; cand is increasing, so no source is overwritten before it is read)
      RowList rows;
      rows.n = (int)cand.size();
      bool moved = false;
      for (size_t k = 0; k < cand.size(); k++) {
        rows.src[k] = cand[k];
        moved = moved || cand[k] != (int32_t)k;
      }
      if (moved) DPD_HIP_TRY(launch_compact_rows(n, PV, rows, st));
    }
    // SchmidtS (:687-726) root by root: normalise, remove the components along
    // every basis vector (the new ones of this iteration included), normalise;
    // skipped when the remaining norm is below 1e-6.  Against the old basis the
    // directions are projected together with block operations (classical
    // Gram-Schmidt applied twice); among this iteration's new vectors by
    // Cholesky QR (twice), whose pivots are the sequential remaining norms.
    // The roots are orthogonalised assuming none is skipped; the norms are
    // checked once at the end and, in the rare case that one fell below 1e-6,
    // the roots from there on are redone one at a time (sequential MGS: one
    // launch per step that finishes the previous reduction, updates p and
    // starts the next).
    auto schmidt = [&](size_t k, int slot, double* nrm_dev) -> int {
      double* p = PV + k * n;
      double* pin = (double*)W.partA.p;
      double* pout = (double*)W.partB.p;
      DPD_HIP_TRY(launch_pair_dots_partial(n, p, p, pin, st));
      const int nb = dimIter + slot;
      for (int i = 0; i <= nb; i++) {
        const double* bprev = i > 0 ? bv(i - 1) : nullptr;
        const double* bnext = i < nb ? bv(i) : nullptr;
        DPD_HIP_TRY(launch_mgs_step(n, p, bprev, bnext, pin, i == 0 ? 0 : 1, pout, st));
        std::swap(pin, pout);
      }
      DPD_HIP_TRY(launch_reduce_partials(pin, 1, nrm_dev, st));
      DPD_HIP_TRY(launch_scale_by_norm(n, p, pin, bv(dimIter + slot), st));
      return DPD_OK;
    };
    {
      const size_t nopt = std::min<size_t>(cand.size(), (size_t)std::max(dimFree, 0));
      // up to kMgsMulti directions at a time: normalised and projected together
      // on the basis so far (each basis vector read once), then orthonormalised
      // among themselves by Cholesky QR applied twice -- in exact arithmetic the
      // vectors and remaining norms of the sequential Gram-Schmidt
      double* T1 = (double*)W.tri.p;
      double* piv = T1 + 2 * kMgsMulti * kMgsMulti;
      for (size_t c0 = 0; c0 < nopt; c0 += kMgsMulti) {
        const int nv = (int)std::min<size_t>(kMgsMulti, nopt - c0);
        double* P0 = PV + c0 * n;
        const int nb = dimIter + (int)c0;  // old basis and the new vectors of earlier groups
        // normalised and projected on the (orthonormal) basis twice with block
        // operations (classical Gram-Schmidt with one re-orthogonalisation).
        // The first projections are taken of the unnormalised directions (whose
        // squared norms the Ritz pass left in rstat) and the normalisation is
        // applied with the first update; the second projection runs only when
        // some direction lost more than half its squared norm to the first
        // (decided on the device).  Each update leaves the Gram partials of
        // its result for the Cholesky QR below.
        int* need = (int*)W.cgs.p;
        const double* pn2 = rstat + 2 * nroot;
        MgsSrc src;
        for (int j = 0; j < kMgsMulti; j++) src.src[j] = j < nv ? cand[c0 + j] : 0;
        if (nb > 0) {
          double* C = (double*)W.coef.p;
          DPD_HIP_TRY(launch_dots_multi(n, P0, nv, B, nb, part, C, st));
          DPD_HIP_TRY(launch_cgs_need(C, nv, nb, need, st, pn2, src));
          DPD_HIP_TRY(launch_cgs_update(n, B, nb, C, nv, P0, nullptr, pn2, src, gram, st));
          DPD_HIP_TRY(launch_dots_multi(n, P0, nv, B, nb, part, C, st, need));
          DPD_HIP_TRY(launch_cgs_update(n, B, nb, C, nv, P0, need, nullptr, src, gram, st));
        } else {
          DPD_HIP_TRY(launch_cgs_update(n, B, 0, nullptr, nv, P0, nullptr, pn2, src, gram, st));
        }
        // Q = P L^-T with G = P^T P = L L^T; the pivots L_jj are the remaining
        // norms (a direction below 1e-6 zeroes itself and the ones after it).
        // A second pass on Q restores the orthogonality the first loses to
        // the conditioning of G; its squared pivots times the first ones' are
        // the norms the host checks (dots[c0 + j]).  Each factor is taken by
        // the launch that applies it (every block factors the small Gram
        // matrix itself); Q1 = P T1 is formed only for its Gram matrix, and
        // the new basis vectors are P (T1 T2)
        TriChol c1;
        c1.gpart = gram;
        if (nv == 1) {
          c1.nrm2_out = dots + c0;
          DPD_HIP_TRY(launch_tri_apply(n, P0, nv, nullptr, bv(dimIter + (int)c0), nullptr, st, c1));
        } else {
          double* gram2 = gram + (size_t)kRedBlocks * kMgsMulti * kMgsMulti;
          c1.piv_out = piv;
          c1.T_out = T1;
          DPD_HIP_TRY(launch_tri_apply(n, P0, nv, nullptr, nullptr, gram2, st, c1));
          TriChol c2;
          c2.gpart = gram2;
          c2.piv_in = piv;
          c2.T_pre = T1;
          c2.nrm2_out = dots + c0;
          DPD_HIP_TRY(launch_tri_apply(n, P0, nv, nullptr, bv(dimIter + (int)c0), nullptr, st, c2));
        }
      }
      spec = nopt;
    }
    // sigma of this iteration's new vectors, one batched evaluation (:442-444).
    // Optimistically every direction is kept: sigma and the new rows and
    // columns of Abar are launched before the norms are read, and norms and
    // Abar entries come back in one transfer.  If a direction fell below 1e-6
    // (rare) the ones from it on are dropped and redone one at a time.
    bool abar_done = false;
    if (spec > 0) {
      if ((rc = sigma(bv(dimIter), abv(dimIter), (int)spec))) return rc;
      std::vector<size_t> off1, off2;
      size_t off = kNormSlots;
      if ((rc = abar_dots((int)spec, off, off1, off2))) return rc;
      if ((rc = fetch(hd.data(), dots, sizeof(double) * off))) return rc;
      size_t k = 0;
      while (k < spec && std::sqrt(hd[k]) >= 1e-6) k++;
      dimFree -= (int)k;
      nNewBase = (int)k;
      if (k == spec) {
        abar_fill((int)spec, off1, off2);
        abar_done = true;
      } else {
        nskip++;
        // direction k is dropped: the ones after it were projected against it
        for (k = k + 1; k < cand.size(); k++) {
          if (dimFree == 0) continue;
          if (!roots_on_device) {
            if ((rc = upload(wdev, eigenVal.data(), sizeof(double) * nroot)) ||
                (rc = upload(W.idx.p, cand.data(), sizeof(int32_t) * cand.size())))
              return rc;
            roots_on_device = true;
          }
          DPD_HIP_TRY(launch_precond_ritz(n, wdev, (const int32_t*)W.idx.p + k, 1, diag, AQ, R, PV + k * n, st));
          if ((rc = schmidt(k, nNewBase, dots))) return rc;
          double t2;
          if ((rc = fetch(&t2, dots, sizeof(double)))) return rc;
          if (std::sqrt(t2) < 1e-6) {
            nskip++;
            continue;
          }
          // its sigma now (the optimistic one was computed before the rollback)
          if ((rc = sigma(bv(dimIter + nNewBase), abv(dimIter + nNewBase), 1))) return rc;
          dimFree--;
          nNewBase++;
        }
      }
    }
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
    // Abar update (:484-509): new rows against the old columns, then the new
    // columns against every basis vector
    if (nNewBase > 0 && !abar_done) {
      std::vector<size_t> off1, off2;
      size_t off = 0;
      if ((rc = abar_dots(nNewBase, off, off1, off2))) return rc;
      if ((rc = fetch(hd.data(), dots, sizeof(double) * off))) return rc;
      abar_fill(nNewBase, off1, off2);
    }
    dimIter += nNewBase;
    iterations++;
  }
  // the last Ritz pass's q is the eigenvector output: renormalised now where its
  // norm was off 1 by more than 1e-10 (the same test on the device)
  if (r_unscaled) DPD_HIP_TRY(launch_renorm_pair(n, rstat + nroot, R, AQ, nroot, st));
  for (int i = 0; i < nroot; i++) w_out[i] = eigenVal[i];
  o->iterations = iterations;
  for (int i = 0; i < nroot && i < 64; i++) o->residual[i] = res[i];
  // final check (:527-542): e = A R - w R
  if (o->final_check) {
    if ((rc = sigma(R, (double*)W.NAB.p, nroot))) return rc;
    if ((rc = upload(wdev, eigenVal.data(), sizeof(double) * nroot))) return rc;
    if ((rc = residuals((const double*)W.NAB.p, mv.data()))) return rc;
    for (int i = 0; i < nroot && i < 64; i++) o->final_residual[i] = std::fabs(mv[i]);
  }
  o->sigma_calls = nsig;
  o->skipped = nskip;
  if ((rc = dsync(st))) return rc;
  return converged ? DPD_OK : DPD_ENOTCONV;
}

// ---------------------------------------------------------------------------
// Batched Davidson: n independent problems solved in lockstep.  Each problem's
// Davidson runs in its own host thread; its sigma requests wait at a barrier
// until every unfinished problem has one, and the last to arrive applies the
// operator of all of them at once (problem_sigma_group: one grouped GEMM
// launch per contraction), optionally replaying a hipGraph of that step.
// The reference solves one problem per program run (lib_cc/mod_davidson.f90).
// ---------------------------------------------------------------------------
namespace {

struct SigmaReq {
  int id = 0;
  const double* x = nullptr;
  double* y = nullptr;
  int nv = 0;
  bool done = false;
  int rc = 0;
};

struct Batcher {
  std::mutex m;
  std::condition_variable cv;
  int active = 0;
  std::vector<SigmaReq*> pending;
  std::function<int(std::vector<SigmaReq*>&)> run;
  std::string err;  // message of a failed batch (set in the thread that ran it)

  int submit(SigmaReq* r) {
    std::unique_lock<std::mutex> lk(m);
    r->done = false;
    pending.push_back(r);
    if ((int)pending.size() == active)
      flush();
    else
      cv.wait(lk, [&] { return r->done; });
    return r->rc;
  }
  // with the lock held: every unfinished problem waits in submit
  void flush() {
    std::vector<SigmaReq*> batch;
    batch.swap(pending);
    std::sort(batch.begin(), batch.end(), [](const SigmaReq* a, const SigmaReq* b) { return a->id < b->id; });
    const int rc = run(batch);
    if (rc) err = dpd_last_error();
    for (SigmaReq* q : batch) {
      q->rc = rc;
      q->done = true;
    }
    cv.notify_all();
  }
  void leave() {
    std::unique_lock<std::mutex> lk(m);
    active--;
    if (active > 0 && (int)pending.size() == active) flush();
  }
};

struct GraphCache {
  std::unordered_map<std::string, int> seen;
  std::unordered_map<std::string, hipGraphExec_t> exec;
  std::vector<hipGraph_t> graphs;
  uint64_t gen = 0;  // device-memory generation the graphs were captured at
  void drop() {
    for (auto& e : exec) (void)hipGraphExecDestroy(e.second);
    for (hipGraph_t g : graphs) (void)hipGraphDestroy(g);
    exec.clear();
    graphs.clear();
    for (auto& kv : seen) kv.second = 0;
  }
  ~GraphCache() { drop(); }
};

}  // namespace

// dpd_davidson with the sigma step replayed from hipGraphs: the solver runs on
// a private stream (capture needs one that is not the legacy default stream),
// ordered after the context's stream on entry and before it on exit; the new
// vectors of a step are copied into fixed device vectors, the sigma of that
// vector count runs eagerly twice (plans, tuned tiles, panels and workspace
// settle) and is then captured once and replayed; graphs are dropped whenever
// device memory they reference moved (g_mem_gen).  With the context profiler on
// the steps stay eager (its events are recorded per launch).
static int davidson_graph(dpd_problem* P, dpd_davidson_opts* o, double* R, double* w_out) {
  if (!R || !w_out) return fail(DPD_EINVAL, "dpd_davidson: null argument");
  dpd_ctx* c = P->ctx;
  DPD_HIP_TRY(hipSetDevice(c->device));
  const int64_t maxv = std::max<int64_t>(1, o->nroot);
  const size_t bytes = sizeof(double) * (size_t)P->dim * (size_t)maxv;
  DevBuf X, Y;
  int rc;
  if ((rc = X.alloc(bytes)) || (rc = Y.alloc(bytes))) return rc;
  hipStream_t caller = c->stream, st = nullptr;
  hipEvent_t ev = nullptr;
  DPD_HIP_TRY(hipStreamCreate(&st));
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamDestroy(st);
    return fail(DPD_EHIP, "event creation failed");
  }
  (void)hipEventRecord(ev, caller);
  (void)hipStreamWaitEvent(st, ev, 0);
  c->stream = st;
  int64_t replays = 0;
  {
    GraphCache gc;
    SigmaHook hook = [&](const double* x, double* y, int nv) -> int {
      if (nv > maxv || c->prof.on) return o->transpose ? problem_sigma_t(P, x, y, nv) : problem_sigma(P, x, y, nv);
      auto sig = [&]() { return o->transpose ? problem_sigma_t(P, (const double*)X.p, (double*)Y.p, nv)
                                             : problem_sigma(P, (const double*)X.p, (double*)Y.p, nv); };
      const uint64_t gen = g_mem_gen.load(std::memory_order_relaxed);
      if (gen != gc.gen) {
        if (!gc.exec.empty()) DPD_HIP_TRY(hipStreamSynchronize(st));
        gc.drop();
        gc.gen = gen;
      }
      DPD_HIP_TRY(hipMemcpyAsync(X.p, x, sizeof(double) * (size_t)P->dim * nv, hipMemcpyDeviceToDevice, st));
      const std::string key = std::to_string(nv);
      auto it = gc.exec.find(key);
      int& seen = gc.seen[key];
      if (it != gc.exec.end()) {
        DPD_HIP_TRY(hipGraphLaunch(it->second, st));
        replays++;
      } else if (seen < 2) {
        seen++;
        int r = sig();
        if (r) return r;
      } else {
        bool ok = seen == 2;
        if (ok) {
          hipGraph_t g = nullptr;
          DPD_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          g_capturing = true;
          int r = sig();
          g_capturing = false;
          hipError_t e = hipStreamEndCapture(st, &g);
          ok = r == DPD_OK && e == hipSuccess && g && g_mem_gen.load(std::memory_order_relaxed) == gc.gen;
          if (ok) {
            gc.graphs.push_back(g);
            hipGraphExec_t ge = nullptr;
            DPD_HIP_TRY(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            gc.exec.emplace(key, ge);
            DPD_HIP_TRY(hipGraphLaunch(ge, st));
            replays++;
          } else {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            seen = 3;  // stays eager for this vector count
          }
        }
        if (!ok) {
          int r = sig();
          if (r) return r;
        }
      }
      DPD_HIP_TRY(hipMemcpyAsync(y, Y.p, sizeof(double) * (size_t)P->dim * nv, hipMemcpyDeviceToDevice, st));
      return DPD_OK;
    };
    rc = davidson_run(P, o, R, w_out, &hook, st);
    (void)hipStreamSynchronize(st);
  }
  c->stream = caller;
  (void)hipEventRecord(ev, st);
  (void)hipStreamWaitEvent(caller, ev, 0);
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(st);
  o->graph_replays = replays;
  return rc;
}

extern "C" int dpd_problem_sigma_group(dpd_problem* const* probs, const double* const* R, double* const* sigma,
                                       const int64_t* nvec, int64_t n) {
  if (!probs || !R || !sigma || !nvec || n < 1) return fail(DPD_EINVAL, "dpd_problem_sigma_group: bad arguments");
  for (int64_t i = 0; i < n; i++) {
    if (!probs[i] || !R[i] || !sigma[i] || nvec[i] < 1) return fail(DPD_EINVAL, "dpd_problem_sigma_group: member %lld",
                                                                    (long long)i);
    if (probs[i]->stale) return fail(DPD_EINVAL, "problem inputs changed: call dpd_problem_refresh first");
  }
  DPD_HIP_TRY(hipSetDevice(probs[0]->ctx->device));
  return problem_sigma_group(probs, R, sigma, nvec, (int)n);
}

extern "C" int dpd_davidson_batch(dpd_problem* const* probs, int64_t n, dpd_davidson_opts* opts, double* const* R,
                                  double* const* w, int64_t flags, int64_t* stats) {
  int64_t stat[4] = {0, 0, 0, 0};  // batched steps, graph replays, captures, failed captures
  if (!probs || !opts || !R || !w || n < 1 || n > 1024) return fail(DPD_EINVAL, "dpd_davidson_batch: bad arguments");
  dpd_ctx* c = probs[0]->ctx;
  for (int64_t i = 0; i < n; i++) {
    if (!probs[i] || !R[i] || !w[i]) return fail(DPD_EINVAL, "dpd_davidson_batch: null member %lld", (long long)i);
    if (probs[i]->ctx != c || probs[i]->method != probs[0]->method || probs[i]->prec != probs[0]->prec)
      return fail(DPD_EINVAL, "dpd_davidson_batch: problems must share context, method and precision");
    if (opts[i].transpose) return fail(DPD_EINVAL, "dpd_davidson_batch: transpose is not supported");
  }
  const bool grouped = flags & 1, graph = (flags & 2) && grouped;
  DPD_HIP_TRY(hipSetDevice(c->device));
  // a private stream (a graph cannot be captured on the legacy default stream),
  // ordered after the caller's stream on entry and before it on exit
  hipStream_t caller = c->stream, st = nullptr;
  hipEvent_t ev = nullptr;
  DPD_HIP_TRY(hipStreamCreate(&st));
  DPD_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  DPD_HIP_TRY(hipEventRecord(ev, caller));
  DPD_HIP_TRY(hipStreamWaitEvent(st, ev, 0));
  c->stream = st;
  // fixed staging vectors for graph replays (the solver's vectors move)
  std::vector<DevBuf> X(graph ? n : 0), Y(graph ? n : 0);
  int rc = DPD_OK;
  for (int64_t i = 0; graph && i < n && !rc; i++) {
    const size_t bytes = sizeof(double) * (size_t)probs[i]->dim * (size_t)std::max<int64_t>(1, opts[i].nroot);
    if ((rc = X[i].alloc(bytes)) || (rc = Y[i].alloc(bytes))) break;
    // a member of a graph's shape that has no request in a step is computed
    // on whatever its staging vectors hold: start them finite
    if (hipMemsetAsync(X[i].p, 0, bytes, st) != hipSuccess) rc = fail(DPD_EHIP, "staging clear failed");
  }
  // each solver's own work on a stream of its own, so that the vector algebra
  // and host synchronisations of different problems overlap; the sigma step
  // on the context stream waits for every member and they wait for it
  std::vector<hipStream_t> ms(n, nullptr);
  std::vector<hipEvent_t> mev(n, nullptr);
  hipEvent_t done = nullptr;
  for (int64_t i = 0; i < n && !rc; i++) {
    if (hipStreamCreate(&ms[i]) != hipSuccess || hipEventCreateWithFlags(&mev[i], hipEventDisableTiming) != hipSuccess)
      rc = fail(DPD_EHIP, "stream creation failed");
    else if (hipStreamWaitEvent(ms[i], ev, 0) != hipSuccess)
      rc = fail(DPD_EHIP, "stream wait failed");
  }
  if (!rc && hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess)
    rc = fail(DPD_EHIP, "event creation failed");
  GraphCache gc;
  Batcher B;
  B.active = (int)n;
  // graph shape: the members a captured step covers, each with all its roots
  // (a member's request never exceeds them), so that members converging or
  // asking for fewer vectors replay the same graph; a new shape (the members
  // still asking) is taken when a request falls outside it or the requested
  // work drops below a quarter of the shape's (computing converged members
  // costs more than the two eager steps a new shape takes beyond that)
#ifndef DPD_BATCH_SHAPE_FRAC
#define DPD_BATCH_SHAPE_FRAC 0.25
#endif
  std::vector<char> in_shape(n, 1);
  double shape_work = 0.0;
  for (int64_t i = 0; i < n; i++) shape_work += (double)probs[i]->dim * std::max<int64_t>(1, opts[i].nroot);
  int shape_id = 0;
  // one batched sigma step on the context stream
  auto run_step = [&](std::vector<SigmaReq*>& reqs) -> int {
    bool staged = graph;  // every request fits its member's staging vectors
    for (SigmaReq* q : reqs) staged = staged && q->nv <= std::max<int64_t>(1, opts[q->id].nroot);
    if (!staged) {
      const int m = (int)reqs.size();
      std::vector<dpd_problem*> ps(m);
      std::vector<const double*> xs(m);
      std::vector<double*> ys(m);
      std::vector<int64_t> nv(m);
      for (int i = 0; i < m; i++) {
        ps[i] = probs[reqs[i]->id];
        nv[i] = reqs[i]->nv;
        xs[i] = reqs[i]->x;
        ys[i] = reqs[i]->y;
      }
      if (!grouped) {
        for (int i = 0; i < m; i++) {
          int r = problem_sigma(ps[i], xs[i], ys[i], nv[i]);
          if (r) return r;
        }
        return DPD_OK;
      }
      return problem_sigma_group(ps.data(), xs.data(), ys.data(), nv.data(), m);
    }
    double work = 0.0;
    bool fits = true;
    for (SigmaReq* q : reqs) {
      work += (double)probs[q->id]->dim * q->nv;
      fits = fits && in_shape[q->id];
    }
    if (!fits || work < DPD_BATCH_SHAPE_FRAC * shape_work) {
      std::fill(in_shape.begin(), in_shape.end(), 0);
      shape_work = 0.0;
      for (SigmaReq* q : reqs) {
        in_shape[q->id] = 1;
        shape_work += (double)probs[q->id]->dim * std::max<int64_t>(1, opts[q->id].nroot);
      }
      shape_id++;
    }
    std::vector<dpd_problem*> ps;
    std::vector<const double*> xs;
    std::vector<double*> ys;
    std::vector<int64_t> nv;
    for (int64_t i = 0; i < n; i++)
      if (in_shape[i]) {
        ps.push_back(probs[i]);
        nv.push_back(std::max<int64_t>(1, opts[i].nroot));
        xs.push_back((const double*)X[i].p);
        ys.push_back((double*)Y[i].p);
      }
    const int m = (int)ps.size();
    const std::string key = std::to_string(shape_id);
    // graphs hold addresses of the workspace, task tables and packed panels:
    // once any of those was freed, moved or repacked they are dropped and the
    // shapes learnt again (two eager runs, then a capture)
    const uint64_t gen = g_mem_gen.load(std::memory_order_relaxed);
    if (gen != gc.gen) {
      if (!gc.exec.empty()) DPD_HIP_TRY(hipStreamSynchronize(st));
      gc.drop();
      gc.gen = gen;
    }
    for (SigmaReq* q : reqs)
      DPD_HIP_TRY(hipMemcpyAsync(X[q->id].p, q->x, sizeof(double) * probs[q->id]->dim * q->nv,
                                 hipMemcpyDeviceToDevice, st));
    auto it = gc.exec.find(key);
    int& seen = gc.seen[key];
    if (it != gc.exec.end()) {
      DPD_HIP_TRY(hipGraphLaunch(it->second, st));
      stat[1]++;
    } else if (seen < 2) {
      // the first two times eagerly: plans, tiles (timed on the first call) and
      // the panels and workspace of the chosen tiles are made here
      seen++;
      int r = problem_sigma_group(ps.data(), xs.data(), ys.data(), nv.data(), m);
      if (r) return r;
    } else {
      bool ok = seen == 2;
      if (ok) {
        // capture; a step that would still allocate, free, repack or
        // synchronise stops before doing so (g_capturing), fails the capture
        // and stays eager
        hipGraph_t g = nullptr;
        DPD_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        g_capturing = true;
        int r = problem_sigma_group(ps.data(), xs.data(), ys.data(), nv.data(), m);
        g_capturing = false;
        hipError_t e = hipStreamEndCapture(st, &g);
        ok = r == DPD_OK && e == hipSuccess && g && g_mem_gen.load(std::memory_order_relaxed) == gc.gen;
        if (ok) {
          gc.graphs.push_back(g);
          hipGraphExec_t ge = nullptr;
          DPD_HIP_TRY(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
          gc.exec.emplace(key, ge);
          DPD_HIP_TRY(hipGraphLaunch(ge, st));
          stat[1]++;
          stat[2]++;
        } else {
          stat[3]++;
          if (g) (void)hipGraphDestroy(g);
          (void)hipGetLastError();
          seen = 3;  // never again for this shape
        }
      }
      if (!ok) {
        int r = problem_sigma_group(ps.data(), xs.data(), ys.data(), nv.data(), m);
        if (r) return r;
      }
    }
    for (SigmaReq* q : reqs)
      DPD_HIP_TRY(hipMemcpyAsync(q->y, Y[q->id].p, sizeof(double) * probs[q->id]->dim * q->nv,
                                 hipMemcpyDeviceToDevice, st));
    return DPD_OK;
  };
  B.run = [&](std::vector<SigmaReq*>& reqs) -> int {
    stat[0]++;
    for (SigmaReq* q : reqs) {
      DPD_HIP_TRY(hipEventRecord(mev[q->id], ms[q->id]));
      DPD_HIP_TRY(hipStreamWaitEvent(st, mev[q->id], 0));
    }
    int brc = run_step(reqs);
    if (brc) return brc;
    DPD_HIP_TRY(hipEventRecord(done, st));
    for (SigmaReq* q : reqs) DPD_HIP_TRY(hipStreamWaitEvent(ms[q->id], done, 0));
    return DPD_OK;
  };
  std::vector<int> rcs(n, DPD_OK);
  std::vector<std::string> errs(n);
  if (!rc) {
    std::vector<std::thread> th;
    th.reserve(n);
    for (int64_t i = 0; i < n; i++)
      th.emplace_back([&, i] {
        SigmaReq req;
        req.id = (int)i;
        SigmaHook hook = [&](const double* x, double* y, int nv) {
          req.x = x;
          req.y = y;
          req.nv = nv;
          const int r = B.submit(&req);
          if (r) fail(r, "batched sigma: %s", B.err.c_str());
          return r;
        };
        if (hipSetDevice(c->device) != hipSuccess)
          rcs[i] = DPD_EHIP;
        else
          rcs[i] = davidson_run(probs[i], &opts[i], R[i], w[i], &hook, ms[i]);
        if (rcs[i] && rcs[i] != DPD_ENOTCONV) errs[i] = dpd_last_error();
        B.leave();
      });
    for (auto& t : th) t.join();
  }
  for (int64_t i = 0; i < n; i++) {
    if (ms[i]) {
      (void)hipStreamSynchronize(ms[i]);
      (void)hipStreamDestroy(ms[i]);
    }
    if (mev[i]) (void)hipEventDestroy(mev[i]);
  }
  if (done) (void)hipEventDestroy(done);
  (void)hipStreamSynchronize(st);
  c->stream = caller;
  if (stats)
    for (int k = 0; k < 4; k++) stats[k] = stat[k];
  X.clear();
  Y.clear();
  (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(st);
  if (rc) return rc;
  // first hard error wins; otherwise not-converged if any member did not converge
  int worst = DPD_OK;
  for (int64_t i = 0; i < n; i++)
    if (rcs[i] && rcs[i] != DPD_ENOTCONV) return fail(rcs[i], "member %lld: %s", (long long)i, errs[i].c_str());
  for (int64_t i = 0; i < n; i++)
    if (rcs[i] == DPD_ENOTCONV) worst = DPD_ENOTCONV;
  return worst;
}
