// Small dense eigenproblems of the fast-diagonalization preconditioner and
// the Radau IIA stage matrix (n <= 8): eigenvalues of a real nonsymmetric
// matrix by the shifted QR algorithm on its complex Hessenberg form,
// eigenvectors by inverse iteration, and the real-pair storage
// (column pairs Re v, Im v for every conjugate pair lambda = a + ib, b > 0,
// then the real eigenvectors) that extract_real_eigenvalues produces in
// advection_solver_irk_real.cc:293-340. All work runs in long double; the
// eigenvector matrices of the upwind DG derivative have condition numbers
// up to ~1e4 at degree 7, which the extra precision absorbs.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <vector>

namespace dgops
{
  using ereal = long double;
  using cplx  = std::complex<ereal>;

  inline std::vector<cplx>
  eigenvalues(const std::vector<double> &A, int n)
  {
    std::vector<cplx> H(n * n);
    for (int i = 0; i < n * n; ++i)
      H[i] = (ereal)A[i];
    auto h = [&](int i, int j) -> cplx & { return H[i * n + j]; };
    // Householder reduction to upper Hessenberg form
    for (int k = 0; k + 2 < n; ++k)
      {
        ereal alpha = 0;
        for (int i = k + 1; i < n; ++i)
          alpha += std::norm(h(i, k));
        alpha = std::sqrt(alpha);
        if (alpha == 0)
          continue;
        std::vector<cplx> v(n, 0.);
        const cplx        x0 = h(k + 1, k);
        const cplx        ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cplx(1);
        v[k + 1]             = x0 + ph * alpha;
        for (int i = k + 2; i < n; ++i)
          v[i] = h(i, k);
        ereal vn = 0;
        for (int i = k + 1; i < n; ++i)
          vn += std::norm(v[i]);
        if (vn == 0)
          continue;
        for (int j = 0; j < n; ++j)
          {
            cplx s = 0;
            for (int i = k + 1; i < n; ++i)
              s += std::conj(v[i]) * h(i, j);
            s *= ereal(2) / vn;
            for (int i = k + 1; i < n; ++i)
              h(i, j) -= v[i] * s;
          }
        for (int i = 0; i < n; ++i)
          {
            cplx s = 0;
            for (int j = k + 1; j < n; ++j)
              s += h(i, j) * v[j];
            s *= ereal(2) / vn;
            for (int j = k + 1; j < n; ++j)
              h(i, j) -= s * std::conj(v[j]);
          }
      }
    // single-shift QR with Wilkinson shifts and deflation at the bottom
    std::vector<cplx> ev;
    ereal             scale = 0;
    for (int i = 0; i < n * n; ++i)
      scale = std::max(scale, std::abs(H[i]));
    int m = n, iter = 0;
    while (m > 0)
      {
        if (m == 1)
          {
            ev.push_back(h(0, 0));
            break;
          }
        const ereal tol = ereal(1e-19) *
                          (std::abs(h(m - 1, m - 1)) +
                           std::abs(h(m - 2, m - 2)) + ereal(1e-300));
        if (std::abs(h(m - 1, m - 2)) <= std::max(tol, ereal(1e-300) * scale))
          {
            ev.push_back(h(m - 1, m - 1));
            --m;
            iter = 0;
            continue;
          }
        if (++iter > 1000)
          throw std::runtime_error("eigenvalue iteration did not converge");
        const cplx a = h(m - 2, m - 2), b = h(m - 2, m - 1),
                   c = h(m - 1, m - 2), d = h(m - 1, m - 1);
        const cplx tr = a + d, det = a * d - b * c;
        const cplx disc = std::sqrt(tr * tr - ereal(4) * det);
        const cplx l1 = (tr + disc) / ereal(2), l2 = (tr - disc) / ereal(2);
        cplx       mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
        if (iter % 11 == 10)
          mu += std::abs(h(m - 1, m - 2)) * cplx(0.75, 0.3); // exceptional
        for (int i = 0; i < m; ++i)
          h(i, i) -= mu;
        std::vector<cplx> cs(m), sn(m);
        for (int k = 0; k + 1 < m; ++k)
          {
            const cplx  x = h(k, k), y = h(k + 1, k);
            const ereal r = std::sqrt(std::norm(x) + std::norm(y));
            const cplx  cc = r == 0 ? cplx(1) : x / r;
            const cplx  ss = r == 0 ? cplx(0) : y / r;
            cs[k]          = cc;
            sn[k]          = ss;
            for (int j = k; j < m; ++j)
              {
                const cplx t1 = h(k, j), t2 = h(k + 1, j);
                h(k, j)       = std::conj(cc) * t1 + std::conj(ss) * t2;
                h(k + 1, j)   = -ss * t1 + cc * t2;
              }
          }
        for (int k = 0; k + 1 < m; ++k)
          {
            const cplx cc = cs[k], ss = sn[k];
            for (int i = 0; i <= std::min(k + 1, m - 1); ++i)
              {
                const cplx t1 = h(i, k), t2 = h(i, k + 1);
                h(i, k)       = t1 * cc + t2 * ss;
                h(i, k + 1)   = -t1 * std::conj(ss) + t2 * std::conj(cc);
              }
          }
        for (int i = 0; i < m; ++i)
          h(i, i) += mu;
      }
    return ev;
  }

  // solve (A - lambda I) x = b with partial pivoting; a zero pivot is
  // replaced by a tiny one (inverse iteration)
  inline std::vector<cplx>
  shifted_solve(const std::vector<double> &A, int n, cplx lambda,
                std::vector<cplx> b)
  {
    std::vector<cplx> M(n * n);
    ereal             nrm = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        {
          M[i * n + j] = (ereal)A[i * n + j] - (i == j ? lambda : cplx(0));
          nrm          = std::max(nrm, (ereal)std::fabs(A[i * n + j]));
        }
    for (int c = 0; c < n; ++c)
      {
        int p = c;
        for (int r = c + 1; r < n; ++r)
          if (std::abs(M[r * n + c]) > std::abs(M[p * n + c]))
            p = r;
        for (int k = 0; k < n; ++k)
          std::swap(M[c * n + k], M[p * n + k]);
        std::swap(b[c], b[p]);
        if (std::abs(M[c * n + c]) < ereal(1e-18) * nrm)
          M[c * n + c] = ereal(1e-18) * nrm;
        for (int r = c + 1; r < n; ++r)
          {
            const cplx f = M[r * n + c] / M[c * n + c];
            for (int k = c; k < n; ++k)
              M[r * n + k] -= f * M[c * n + k];
            b[r] -= f * b[c];
          }
      }
    for (int r = n - 1; r >= 0; --r)
      {
        cplx s = b[r];
        for (int k = r + 1; k < n; ++k)
          s -= M[r * n + k] * b[k];
        b[r] = s / M[r * n + r];
      }
    return b;
  }

  inline std::vector<cplx>
  eigenvector(const std::vector<double> &A, int n, cplx lambda)
  {
    std::vector<cplx> x(n);
    for (int i = 0; i < n; ++i)
      x[i] = cplx(ereal(1) + ereal(0.1) * i, ereal(0.05) * (i % 3));
    for (int it = 0; it < 4; ++it)
      {
        x          = shifted_solve(A, n, lambda, x);
        ereal nrm2 = 0;
        int   imax = 0;
        for (int i = 0; i < n; ++i)
          {
            nrm2 += std::norm(x[i]);
            if (std::abs(x[i]) > std::abs(x[imax]))
              imax = i;
          }
        // unit norm, largest entry real and positive
        const cplx ph = std::abs(x[imax]) / x[imax] / std::sqrt(nrm2);
        for (auto &v : x)
          v *= ph;
      }
    return x;
  }

  template <typename R>
  inline std::vector<R>
  invert_dense_t(std::vector<R> a, int n)
  {
    std::vector<R> inv(n * n, R(0));
    for (int i = 0; i < n; ++i)
      inv[i * n + i] = 1;
    for (int c = 0; c < n; ++c)
      {
        int p = c;
        for (int r = c + 1; r < n; ++r)
          if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c]))
            p = r;
        for (int k = 0; k < n; ++k)
          {
            std::swap(a[c * n + k], a[p * n + k]);
            std::swap(inv[c * n + k], inv[p * n + k]);
          }
        const R d = a[c * n + c];
        if (d == 0)
          throw std::runtime_error("singular matrix");
        for (int k = 0; k < n; ++k)
          {
            a[c * n + k] /= d;
            inv[c * n + k] /= d;
          }
        for (int r = 0; r < n; ++r)
          if (r != c && a[r * n + c] != R(0))
            {
              const R f = a[r * n + c];
              for (int k = 0; k < n; ++k)
                {
                  a[r * n + k] -= f * a[c * n + k];
                  inv[r * n + k] -= f * inv[c * n + k];
                }
            }
      }
    return inv;
  }

  inline std::vector<double>
  invert_dense(const std::vector<double> &a, int n)
  {
    std::vector<ereal> al(a.begin(), a.end());
    const auto         il = invert_dense_t<ereal>(al, n);
    return std::vector<double>(il.begin(), il.end());
  }

  // Real-pair eigendecomposition A = V B V^-1 with B block diagonal:
  // columns 2p, 2p+1 = (Re v, Im v) of lambda_p = a + ib (b > 0), then the
  // real eigenvectors. kind[j]: 0 real, 1 first of a pair, 2 second.
  struct RealPairEig
  {
    int                 n;
    std::vector<double> V, Vinv; // row-major n x n
    std::vector<double> lam_re, lam_im;
    std::vector<int>    kind;
  };

  inline RealPairEig
  real_pair_eigen(const std::vector<double> &A, int n)
  {
    const auto ev    = eigenvalues(A, n);
    ereal      scale = 0;
    for (auto e : ev)
      scale = std::max(scale, std::abs(e));
    const ereal       tol = ereal(1e-10) * std::max(scale, ereal(1e-300));
    std::vector<cplx> pairs, reals;
    std::vector<bool> used(ev.size(), false);
    for (size_t i = 0; i < ev.size(); ++i)
      {
        if (used[i])
          continue;
        if (std::fabs(ev[i].imag()) <= tol)
          {
            reals.push_back(ev[i].real());
            used[i] = true;
            continue;
          }
        size_t best = ev.size();
        for (size_t j = 0; j < ev.size(); ++j)
          if (!used[j] && j != i &&
              (best == ev.size() ||
               std::abs(ev[j] - std::conj(ev[i])) <
                 std::abs(ev[best] - std::conj(ev[i]))))
            best = j;
        if (best == ev.size() ||
            std::abs(ev[best] - std::conj(ev[i])) > ereal(1e-8) * scale)
          throw std::runtime_error(
            "eigenvalues do not come in complex-conjugate pairs");
        used[i] = used[best] = true;
        const cplx l = ev[i].imag() > 0 ? ev[i] : ev[best];
        const cplx o = ev[i].imag() > 0 ? ev[best] : ev[i];
        pairs.push_back(ereal(0.5) * (l + std::conj(o)));
      }
    std::sort(pairs.begin(), pairs.end(), [](cplx a, cplx b) {
      return a.real() != b.real() ? a.real() < b.real() : a.imag() < b.imag();
    });
    std::sort(reals.begin(), reals.end(),
              [](cplx a, cplx b) { return a.real() < b.real(); });
    RealPairEig R;
    R.n = n;
    std::vector<ereal> Vl(n * n, 0.);
    R.lam_re.assign(n, 0.);
    R.lam_im.assign(n, 0.);
    R.kind.assign(n, 0);
    int col = 0;
    for (const cplx l : pairs)
      {
        const auto v = eigenvector(A, n, l);
        for (int i = 0; i < n; ++i)
          {
            Vl[i * n + col]     = v[i].real();
            Vl[i * n + col + 1] = v[i].imag();
          }
        R.lam_re[col] = R.lam_re[col + 1] = (double)l.real();
        R.lam_im[col] = R.lam_im[col + 1] = (double)l.imag();
        R.kind[col]                       = 1;
        R.kind[col + 1]                   = 2;
        col += 2;
      }
    for (const cplx l : reals)
      {
        const auto v = eigenvector(A, n, l.real());
        for (int i = 0; i < n; ++i)
          Vl[i * n + col] = v[i].real();
        R.lam_re[col] = (double)l.real();
        R.kind[col]   = 0;
        ++col;
      }
    const auto Vi = invert_dense_t<ereal>(Vl, n);
    R.V.assign(Vl.begin(), Vl.end());
    R.Vinv.assign(Vi.begin(), Vi.end());
    return R;
  }
} // namespace dgops
