// Tensor-product Lagrange basis on cosine (Chebyshev-Gauss-Lobatto) knots, keeping the element API
// of the reference (element/LagrangeShapeFuncs.h:7-252):
//   getCosinePoint(knotInd, order)                       knot x_j = -cos(pi*j/order)
//   lagrangePoly1d<T, order>(x, nodeInd)                 N_j(x)
//   lagrangePoly1dDeriv<T, order>(x, nodeInd, N, dNdx)   N_j(x), dN_j/dx
//   lagrangePoly2d<T, order>(x[2], nodeXInd, nodeYInd)
//   lagrangePoly2dDeriv<T, order>(x[2], nodeXInd, nodeYInd, N, deriv[2])
//     deriv[0] = N_y * dN_x/dx, deriv[1] = N_x * dN_y/dy  (reference :244-252)
//
// For T = double and orders 1..4 the polynomials are evaluated from monomial coefficient tables
// with the same fma nesting as the reference (:37-108, :129-216), so host and device produce the
// reference's basis values bit for bit. The coefficient values are the reference's, including the
// truncated sqrt(2) = 1.41421356237310 of the order-4 table (:93-95); they are the expansion of the
// product formula below on the cosine knots. Other types/orders use the product formula
// N_j(x) = prod_{i != j} (x - x_i) / (x_j - x_i) with a forward-mode product rule for dN/dx
// (the reference differentiates the same product with a forward AD scalar, :120-127).
#pragma once

#include <math.h>

#include "fem_hd.h"

FEM_HD inline double getCosinePoint(const int knotInd, const int order) {
  const double xLinear = -1.0 + 2.0 * double(knotInd) / order;
  return -cos(M_PI / 2 * (xLinear + 1.0));
}

namespace fem_basis_detail {

// Monomial coefficients, highest degree first: N_j(x) = sum_k c[j][k] x^(order-k).
template <int order>
struct MonomialTable;

template <>
struct MonomialTable<1> {
  static constexpr double c[2][2] = {{-0.5, 0.5}, {0.5, 0.5}};
};
template <>
struct MonomialTable<2> {
  static constexpr double c[3][3] = {{0.5, -0.5, 0.0}, {-1.0, 0.0, 1.0}, {0.5, 0.5, 0.0}};
};
template <>
struct MonomialTable<3> {
  static constexpr double c[4][4] = {{-2.0 / 3.0, 2.0 / 3.0, 1.0 / 6.0, -1.0 / 6.0},
                                     {4.0 / 3.0, -2.0 / 3.0, -4.0 / 3.0, 2.0 / 3.0},
                                     {-4.0 / 3.0, -2.0 / 3.0, 4.0 / 3.0, 2.0 / 3.0},
                                     {2.0 / 3.0, 2.0 / 3.0, -1.0 / 6.0, -1.0 / 6.0}};
};
template <>
struct MonomialTable<4> {
  static constexpr double c[5][5] = {{1.0, -1.0, -0.5, 0.5, 0.0},
                                     {-2.0, 1.41421356237310, 2.0, -1.41421356237310, 0.0},
                                     {2.0, 0.0, -3.0, 0.0, 1.0},
                                     {-2.0, -1.41421356237310, 2.0, 1.41421356237310, 0.0},
                                     {1.0, 1.0, -0.5, -0.5, 0.0}};
};

// Powers x^1..x^4 formed as x2 = x*x, x3 = x2*x, x4 = x2*x2.
FEM_HD inline void powers(const double x, double pw[5]) {
  pw[0] = 1.0;
  pw[1] = x;
  pw[2] = x * x;
  pw[3] = pw[2] * x;
  pw[4] = pw[2] * pw[2];
}

// N = fma(c0, x^p, fma(c1, x^(p-1), ... fma(c_{p-1}, x, c_p)))
template <int order>
FEM_HD inline double evalMonomial(const double *c, const double pw[5]) {
  double acc = c[order];
  for (int k = order - 1; k >= 0; k--) {
    acc = fma(c[k], pw[order - k], acc);
  }
  return acc;
}

// dN = fma(p*c0, x^(p-1), fma((p-1)*c1, x^(p-2), ... fma(2*c_{p-2}, x, c_{p-1})))
template <int order>
FEM_HD inline double evalMonomialDeriv(const double *c, const double pw[5]) {
  double acc = c[order - 1];
  for (int k = order - 2; k >= 0; k--) {
    acc = fma((order - k) * c[k], pw[order - 1 - k], acc);
  }
  return acc;
}

}  // namespace fem_basis_detail

// Generic product-formula evaluation.
template <typename numType, int order>
FEM_HD inline numType lagrangePoly1d(const numType x, const int nodeInd) {
  numType result = 1.0;
  const double xj = getCosinePoint(nodeInd, order);
  for (int ii = 0; ii <= order; ii++) {
    if (ii != nodeInd) {
      const double xi = getCosinePoint(ii, order);
      result *= (x - xi) / (xj - xi);
    }
  }
  return result;
}

template <typename numType, int order>
FEM_HD inline void lagrangePoly1dDeriv(const numType x, const int nodeInd, numType &N, numType &dNdx) {
  // Forward-mode product rule over the factors (x - x_i)/(x_j - x_i).
  numType val = 1.0;
  numType der = 0.0;
  const double xj = getCosinePoint(nodeInd, order);
  for (int ii = 0; ii <= order; ii++) {
    if (ii != nodeInd) {
      const double xi = getCosinePoint(ii, order);
      const double inv = 1.0 / (xj - xi);
      der = der * ((x - xi) * inv) + val * inv;
      val = val * ((x - xi) * inv);
    }
  }
  N = val;
  dNdx = der;
}

#define FEM_LAGRANGE_MONOMIAL_SPECIALISATION(ORD)                                                          \
  template <>                                                                                              \
  FEM_HD inline double lagrangePoly1d<double, ORD>(const double x, const int nodeInd) {                   \
    double pw[5];                                                                                          \
    fem_basis_detail::powers(x, pw);                                                                       \
    return fem_basis_detail::evalMonomial<ORD>(fem_basis_detail::MonomialTable<ORD>::c[nodeInd], pw);     \
  }                                                                                                        \
  template <>                                                                                              \
  FEM_HD inline void lagrangePoly1dDeriv<double, ORD>(const double x, const int nodeInd, double &N,       \
                                                      double &dNdx) {                                      \
    double pw[5];                                                                                          \
    fem_basis_detail::powers(x, pw);                                                                       \
    const double *c = fem_basis_detail::MonomialTable<ORD>::c[nodeInd];                                    \
    N = fem_basis_detail::evalMonomial<ORD>(c, pw);                                                        \
    dNdx = fem_basis_detail::evalMonomialDeriv<ORD>(c, pw);                                                \
  }

FEM_LAGRANGE_MONOMIAL_SPECIALISATION(1)
FEM_LAGRANGE_MONOMIAL_SPECIALISATION(2)
FEM_LAGRANGE_MONOMIAL_SPECIALISATION(3)
FEM_LAGRANGE_MONOMIAL_SPECIALISATION(4)
#undef FEM_LAGRANGE_MONOMIAL_SPECIALISATION

template <typename numType, int order>
FEM_HD inline numType lagrangePoly2d(const numType x[2], const int nodeXInd, const int nodeYInd) {
  return lagrangePoly1d<numType, order>(x[0], nodeXInd) * lagrangePoly1d<numType, order>(x[1], nodeYInd);
}

template <typename numType, int order>
FEM_HD inline void lagrangePoly2dDeriv(const numType x[2], const int nodeXInd, const int nodeYInd, numType &N,
                                       numType deriv[2]) {
  numType Nx, Ny, dNdx, dNdy;
  lagrangePoly1dDeriv<numType, order>(x[0], nodeXInd, Nx, dNdx);
  lagrangePoly1dDeriv<numType, order>(x[1], nodeYInd, Ny, dNdy);
  N = Nx * Ny;
  deriv[0] = Ny * dNdx;
  deriv[1] = Nx * dNdy;
}
