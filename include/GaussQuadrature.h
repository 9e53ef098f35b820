// 1-D Gauss-Legendre rules on [-1, 1], keeping the element API of the reference:
//   getGaussQuadWeight<order>(ptInd), getGaussQuadCoord<order>(ptInd)
// where `order` selects the rule with order+1 points (exact for degree 2*order+1).
// Replaces element/GaussQuadrature.h:23-174 of the reference.
//
// The point/weight values for orders 1..4 are the reference's 15-significant-digit
// constants (element/GaussQuadrature.h:46-99, themselves TACS's table), kept digit for digit
// so that host and device evaluate exactly the quadrature the reference evaluates. Orders
// 5..7 carry the longer constants the reference lists for those rules. An order outside
// 0..7 returns 0.0, as the reference's generic template does (element/GaussQuadrature.h:23-31).
#pragma once

#include "fem_hd.h"

namespace fem_quadrature_detail {

// Rule with `n` points, stored symmetric-half-free (all points, ascending).
template <int n>
struct GaussRule;

template <>
struct GaussRule<1> {
  static constexpr double x[1] = {0.0};
  static constexpr double w[1] = {2.0};
};
template <>
struct GaussRule<2> {
  static constexpr double x[2] = {-0.577350269189626, 0.577350269189626};
  static constexpr double w[2] = {1.0, 1.0};
};
template <>
struct GaussRule<3> {
  static constexpr double x[3] = {-0.774596669241483, 0.0, 0.774596669241483};
  static constexpr double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
};
template <>
struct GaussRule<4> {
  static constexpr double x[4] = {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053};
  static constexpr double w[4] = {0.347854845137454, 0.652145154862546, 0.652145154862546, 0.347854845137454};
};
template <>
struct GaussRule<5> {
  static constexpr double x[5] = {-0.906179845938664, -0.538469310105683, 0.0, 0.538469310105683,
                                  0.906179845938664};
  static constexpr double w[5] = {0.236926885056189, 0.478628670499366, 0.568888888888889, 0.478628670499366,
                                  0.236926885056189};
};
template <>
struct GaussRule<6> {
  static constexpr double x[6] = {-0.9324695142031520278123016, -0.6612093864662645136613996,
                                  -0.2386191860831969086305017, 0.2386191860831969086305017,
                                  0.6612093864662645136613996,  0.9324695142031520278123016};
  static constexpr double w[6] = {0.1713244923791703450402961, 0.3607615730481386075698335,
                                  0.4679139345726910473898703, 0.4679139345726910473898703,
                                  0.3607615730481386075698335, 0.1713244923791703450402961};
};
template <>
struct GaussRule<7> {
  static constexpr double x[7] = {-0.9491079123427585245261897, -0.7415311855993944398638648,
                                  -0.4058451513773971669066064, 0.0,
                                  0.4058451513773971669066064,  0.7415311855993944398638648,
                                  0.9491079123427585245261897};
  static constexpr double w[7] = {0.1294849661688696932706114, 0.2797053914892766679014678,
                                  0.3818300505051189449503698, 0.4179591836734693877551020,
                                  0.3818300505051189449503698, 0.2797053914892766679014678,
                                  0.1294849661688696932706114};
};
template <>
struct GaussRule<8> {
  static constexpr double x[8] = {-0.9602898564975362316835609, -0.7966664774136267395915539,
                                  -0.5255324099163289858177390, -0.1834346424956498049394761,
                                  0.1834346424956498049394761,  0.5255324099163289858177390,
                                  0.7966664774136267395915539,  0.9602898564975362316835609};
  static constexpr double w[8] = {0.1012285362903762591525314, 0.2223810344533744705443560,
                                  0.3137066458778872873379622, 0.3626837833783619829651504,
                                  0.3626837833783619829651504, 0.3137066458778872873379622,
                                  0.2223810344533744705443560, 0.1012285362903762591525314};
};

}  // namespace fem_quadrature_detail

template <int order>
FEM_HD constexpr double getGaussQuadWeight(const int ptInd) {
  if constexpr (order >= 0 && order <= 7) {
    return fem_quadrature_detail::GaussRule<order + 1>::w[ptInd];
  } else {
    return 0.0;
  }
}

template <int order>
FEM_HD constexpr double getGaussQuadCoord(const int ptInd) {
  if constexpr (order >= 0 && order <= 7) {
    return fem_quadrature_detail::GaussRule<order + 1>::x[ptInd];
  } else {
    return 0.0;
  }
}
