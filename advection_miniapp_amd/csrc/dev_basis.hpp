// Device copy of the 1D tables of one degree, in the arithmetic type of a
// kernel. Kernels receive a pointer to it and read it with uniform
// addresses, so the values come through the scalar cache.
#pragma once

#include "basis.hpp"

namespace dgops
{
  constexpr int max_h = max_n / 2;

  // Even-odd form of an n x n matrix M that is centrosymmetric
  // (M[n-1-q][n-1-i] = M[q][i]: S, S^-1) or skew-centrosymmetric (= -M[q][i]:
  // the collocation derivative D), for h = n/2 and, for odd n, the middle
  // index m = h:
  //   e[q][i] = (M[q][i] + M[q][n-1-i]) / 2,  o[q][i] = (M[q][i] - M[q][n-1-i]) / 2
  //   (q, i < h), mc[q] = M[q][m], mr[i] = M[m][i], mm = M[m][m].
  // A product then needs h*h + h*h (+ 2h + 1) multiplies instead of n*n
  // (the even-odd evaluation of deal.II's EvaluatorTensorProduct).
  template <typename T>
  struct EvenOdd
  {
    T e[max_h][max_h], o[max_h][max_h];
    T mc[max_h], mr[max_h], mm;
  };

  template <typename T>
  struct DevBasis
  {
    T S[max_n][max_n];    // nodal basis at Gauss points [q][i]
    T Sinv[max_n][max_n]; // inverse of S [i][q]
    T D[max_n][max_n];    // collocation derivative [q][p]
    T G[max_n][max_n];    // nodal basis derivative at Gauss points [q][i]
    T l0[max_n], l1[max_n];     // Gauss-Lagrange values at 0 and 1
    T dl0[max_n], dl1[max_n];   // Gauss-Lagrange derivatives at 0 and 1
    T dphi0[max_n], dphi1[max_n];
    T w[max_n];    // Gauss weights on [0,1]
    T winv[max_n]; // their reciprocals (0 past n)
    T xq[max_n];   // Gauss points on [0,1]
    T xn[max_n]; // Gauss-Lobatto nodes on [0,1]
    EvenOdd<T> eo_S, eo_Sinv, eo_D; // even-odd forms of S, S^-1, D
  };

  template <typename T>
  inline EvenOdd<T>
  make_even_odd(const double (&M)[max_n][max_n], int n)
  {
    EvenOdd<T> r = {};
    const int  h = n / 2, m = h;
    for (int q = 0; q < h; ++q)
      for (int i = 0; i < h; ++i)
        {
          r.e[q][i] = (T)(0.5 * (M[q][i] + M[q][n - 1 - i]));
          r.o[q][i] = (T)(0.5 * (M[q][i] - M[q][n - 1 - i]));
        }
    if (n % 2)
      {
        for (int k = 0; k < h; ++k)
          {
            r.mc[k] = (T)M[k][m];
            r.mr[k] = (T)M[m][k];
          }
        r.mm = (T)M[m][m];
      }
    return r;
  }

  template <typename T>
  inline DevBasis<T>
  make_dev_basis(const Basis1D &b)
  {
    DevBasis<T> d = {};
    for (int i = 0; i < max_n; ++i)
      {
        for (int j = 0; j < max_n; ++j)
          {
            d.S[i][j]    = (T)b.S[i][j];
            d.Sinv[i][j] = (T)b.Sinv[i][j];
            d.D[i][j]    = (T)b.D[i][j];
            d.G[i][j]    = (T)b.G[i][j];
          }
        d.l0[i]    = (T)b.l0[i];
        d.l1[i]    = (T)b.l1[i];
        d.dl0[i]   = (T)b.dl0[i];
        d.dl1[i]   = (T)b.dl1[i];
        d.dphi0[i] = (T)b.dphi0[i];
        d.dphi1[i] = (T)b.dphi1[i];
        d.w[i]     = (T)b.gauss_wts[i];
        d.winv[i]  = i < b.n ? (T)(1. / b.gauss_wts[i]) : (T)0;
        d.xq[i]    = (T)b.gauss_pts[i];
        d.xn[i]    = (T)b.gl_pts[i];
      }
    d.eo_S    = make_even_odd<T>(b.S, b.n);
    d.eo_Sinv = make_even_odd<T>(b.Sinv, b.n);
    d.eo_D    = make_even_odd<T>(b.D, b.n);
    return d;
  }
} // namespace dgops
