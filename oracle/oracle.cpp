// ORACLE — test infrastructure only. Nothing in the product (libfem_amd) links, loads or calls this
// file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the checker.
//
// A CPU restatement of the reference's CPU assembly path (A-CGray/GPU-FEM-Project):
//   element/GaussQuadrature.h:46-99        Gauss-Legendre constants (15 digits, TACS table)
//   element/LagrangeShapeFuncs.h:7-252     cosine-knot Lagrange basis, monomial tables, 2-D derivative
//   element/FEKernels.h:25-61              gatherElementData, scatterResidual
//   element/FEKernels.h:63-104             scatterMat (made atomic under OpenMP; the reference races, :97-98)
//   element/FEKernels.h:167-196            addTransformStateGradSens (the branch both cost estimates select)
//   element/FEKernels.h:229-284            interpParamGradient, interpRealGradient
//   element/FEKernels.h:286-315            planeStressWeakRes, LINEAR strain, in closed form
//   element/FEKernels.h:317-377            assemblePlaneStressResidual
//   element/FEKernels.h:379-508            assemblePlaneStressJacobian (per-column forward tangent); the
//                                          forward seed of state s is row s = gradN_a, as the reference's
//                                          GPU kernel writes it (:936). The CPU loop at :469 indexes the
//                                          1x2 seed with the state row, an out-of-bounds read for s = 1,
//                                          which is not reproduced.
//   verification/TACSHelpers.h:138-182     generateElementBCSRMap (linear column search)
// The A2D calls of the reference (un-vendored, no pinned version) are replaced by closed forms:
//   MatInv/MatDet of a 2x2; MatMatMult; weakRes = scale*t*sigma(sym(grad u)), whose forward tangent
//   is scale*t*sigma(sym(grad u_dot)) (linear strain).
// Also: an independent "Gram" assembler (K from weighted gradient Gram sums) used to cross-check the
// faithful form, a brute-force BSR pattern builder, and a timing entry point for the CPU baseline.
//
// Parity pinning (tests/test_oracle_*.py): Gauss rules vs closed forms and exact monomial integration;
// basis vs Kronecker/partition-of-unity/linear reproduction and independently derived coefficients;
// element stiffness vs the textbook unit-square 4-node plane-stress matrix; global K symmetric, rigid
// modes in its null space, R = K u, patch test, uniform-strain energy; assembly vs an independent
// numpy dense assembler on tiny meshes. Bit-exact rounding of A2D's AD stack is parity unpinned
// (no A2D available); closed forms agree with it to rounding.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---- element/GaussQuadrature.h:46-99 ----
const double kGaussX[5][5] = {
    {0, 0, 0, 0, 0},
    {-0.577350269189626, 0.577350269189626, 0, 0, 0},
    {-0.774596669241483, 0.0, 0.774596669241483, 0, 0},
    {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053, 0},
    {-0.906179845938664, -0.538469310105683, 0.0, 0.538469310105683, 0.906179845938664}};
const double kGaussW[5][5] = {
    {2.0, 0, 0, 0, 0},
    {1.0, 1.0, 0, 0, 0},
    {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0, 0, 0},
    {0.347854845137454, 0.652145154862546, 0.652145154862546, 0.347854845137454, 0},
    {0.236926885056189, 0.478628670499366, 0.568888888888889, 0.478628670499366, 0.236926885056189}};

// ---- element/LagrangeShapeFuncs.h:37-108, 129-216: coefficients highest degree first ----
const double kC1[2][2] = {{-0.5, 0.5}, {0.5, 0.5}};
const double kC2[3][3] = {{0.5, -0.5, 0.0}, {-1.0, 0.0, 1.0}, {0.5, 0.5, 0.0}};
const double kC3[4][4] = {{-2.0 / 3.0, 2.0 / 3.0, 1.0 / 6.0, -1.0 / 6.0},
                          {4.0 / 3.0, -2.0 / 3.0, -4.0 / 3.0, 2.0 / 3.0},
                          {-4.0 / 3.0, -2.0 / 3.0, 4.0 / 3.0, 2.0 / 3.0},
                          {2.0 / 3.0, 2.0 / 3.0, -1.0 / 6.0, -1.0 / 6.0}};
const double kC4[5][5] = {{1.0, -1.0, -0.5, 0.5, 0.0},
                          {-2.0, 1.41421356237310, 2.0, -1.41421356237310, 0.0},
                          {2.0, 0.0, -3., 0.0, 1.0},
                          {-2.0, -1.41421356237310, 2.0, 1.41421356237310, 0.0},
                          {1.0, 1.0, -0.5, -0.5, 0.0}};

// lagrangePoly1dDeriv<double, order> (LagrangeShapeFuncs.h:129-216), same fma nesting.
template <int order>
inline void lagrange1dDeriv(const double x, const int k, double &N, double &dN) {
  const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
  if constexpr (order == 1) {
    const double *c = kC1[k];
    N = std::fma(c[0], x, c[1]);
    dN = c[0];
  } else if constexpr (order == 2) {
    const double *c = kC2[k];
    N = std::fma(c[0], x * x, std::fma(c[1], x, c[2]));
    dN = std::fma(2 * c[0], x, c[1]);
  } else if constexpr (order == 3) {
    const double *c = kC3[k];
    N = std::fma(c[0], x3, std::fma(c[1], x2, std::fma(c[2], x, c[3])));
    dN = std::fma(3 * c[0], x2, std::fma(2 * c[1], x, c[2]));
  } else {
    const double *c = kC4[k];
    N = std::fma(c[0], x4, std::fma(c[1], x3, std::fma(c[2], x2, std::fma(c[3], x, c[4]))));
    dN = std::fma(4 * c[0], x3, std::fma(3 * c[1], x2, std::fma(2 * c[2], x, c[3])));
  }
}

// lagrangePoly2dDeriv (LagrangeShapeFuncs.h:244-252).
template <int order>
inline void lagrange2dDeriv(const double xi[2], int nx, int ny, double deriv[2]) {
  double Nx, dNx, Ny, dNy;
  lagrange1dDeriv<order>(xi[0], nx, Nx, dNx);
  lagrange1dDeriv<order>(xi[1], ny, Ny, dNy);
  deriv[0] = Ny * dNx;
  deriv[1] = Nx * dNy;
}

// interpParamGradient (FEKernels.h:229-257): out(i,j) = sum_a vals[a*2+i] * dN_a/dxi_j.
template <int order>
void interpParamGradient(const double xi[2], const double *vals, double out[4]) {
  for (int k = 0; k < 4; k++) out[k] = 0.0;
  for (int ny = 0; ny <= order; ny++) {
    for (int nx = 0; nx <= order; nx++) {
      const int a = ny * (order + 1) + nx;
      double d[2];
      lagrange2dDeriv<order>(xi, nx, ny, d);
      for (int i = 0; i < 2; i++) {
        for (int j = 0; j < 2; j++) out[i * 2 + j] += vals[a * 2 + i] * d[j];
      }
    }
  }
}

// 2x2 MatDet / MatInv (closed forms of the A2D calls at FEKernels.h:357-359).
inline double det2(const double J[4]) { return J[0] * J[3] - J[1] * J[2]; }
inline void inv2(const double J[4], double Ji[4]) {
  const double inv = 1.0 / det2(J);
  Ji[0] = J[3] * inv;
  Ji[1] = -J[1] * inv;
  Ji[2] = -J[2] * inv;
  Ji[3] = J[0] * inv;
}
// C = A B, 2x2 row-major.
inline void matmul2(const double A[4], const double B[4], double C[4]) {
  for (int i = 0; i < 2; i++) {
    for (int j = 0; j < 2; j++) C[i * 2 + j] = A[i * 2] * B[j] + A[i * 2 + 1] * B[2 + j];
  }
}

// planeStressWeakRes, LINEAR strain (FEKernels.h:286-315): scale*t*sigma(sym(Ux)).
inline void weakRes(const double Ux[4], double E, double nu, double t, double scale, double out[4]) {
  const double mu = 0.5 * E / (1.0 + nu);
  const double lambda = 2 * mu * nu / (1.0 - nu);
  const double exx = Ux[0], eyy = Ux[3], exy = 0.5 * (Ux[1] + Ux[2]);
  const double tr = exx + eyy;
  const double f = scale * t;
  out[0] = f * (2 * mu * exx + lambda * tr);
  out[3] = f * (2 * mu * eyy + lambda * tr);
  out[1] = f * (2 * mu * exy);
  out[2] = out[1];
}

// planeStressWeakRes, NONLINEAR strain (FEKernels.h:286-315 with GreenStrainType::NONLINEAR): the Green strain
// E = (Ux + Ux^T + Ux^T Ux) / 2, S = 2 mu E + lambda tr(E) I, energy tr(E S) seeded with scale*t/2, so the
// reverse sweep gives dW/dUx = scale*t * F S with F = I + Ux (dE = sym(F^T dUx), tr(dE S) = (F S) : dUx).
inline void greenStress(const double Ux[4], double mu, double lambda, double S[4]) {
  const double exx = Ux[0] + 0.5 * (Ux[0] * Ux[0] + Ux[2] * Ux[2]);
  const double eyy = Ux[3] + 0.5 * (Ux[1] * Ux[1] + Ux[3] * Ux[3]);
  const double exy = 0.5 * (Ux[1] + Ux[2] + Ux[0] * Ux[1] + Ux[2] * Ux[3]);
  const double tr = exx + eyy;
  S[0] = 2 * mu * exx + lambda * tr;
  S[3] = 2 * mu * eyy + lambda * tr;
  S[1] = S[2] = 2 * mu * exy;
}

inline void weakResNonlinear(const double Ux[4], double E, double nu, double t, double scale, double out[4]) {
  const double mu = 0.5 * E / (1.0 + nu);
  const double lambda = 2 * mu * nu / (1.0 - nu);
  double S[4];
  greenStress(Ux, mu, lambda, S);
  const double F[4] = {1.0 + Ux[0], Ux[1], Ux[2], 1.0 + Ux[3]};
  double FS[4];
  matmul2(F, S, FS);
  const double f = scale * t;
  for (int i = 0; i < 4; i++) out[i] = f * FS[i];
}

// Forward (tangent) derivative of weakResNonlinear at Ux in direction Udot, as the reference's ADScalar
// sweep computes it (FEKernels.h:440-497): scale*t * (Udot S + F Sdot), Sdot = C : sym(F^T Udot).
inline void weakResNonlinearFwd(const double Ux[4], const double Udot[4], double E, double nu, double t,
                                double scale, double out[4]) {
  const double mu = 0.5 * E / (1.0 + nu);
  const double lambda = 2 * mu * nu / (1.0 - nu);
  double S[4];
  greenStress(Ux, mu, lambda, S);
  const double F[4] = {1.0 + Ux[0], Ux[1], Ux[2], 1.0 + Ux[3]};
  // Edot = (F^T Udot + Udot^T F) / 2
  const double exx = F[0] * Udot[0] + F[2] * Udot[2];
  const double eyy = F[1] * Udot[1] + F[3] * Udot[3];
  const double exy = 0.5 * (F[0] * Udot[1] + F[2] * Udot[3] + Udot[0] * F[1] + Udot[2] * F[3]);
  const double tr = exx + eyy;
  const double Sd[4] = {2 * mu * exx + lambda * tr, 2 * mu * exy, 2 * mu * exy, 2 * mu * eyy + lambda * tr};
  double a[4], b[4];
  matmul2(Udot, S, a);
  matmul2(F, Sd, b);
  const double f = scale * t;
  for (int i = 0; i < 4; i++) out[i] = f * (a[i] + b[i]);
}

// addTransformStateGradSens, the branch selected for every order (FEKernels.h:167-196):
// sens[a*2+i] += sum_k (dN_a . Jinv)_k * wr(i, k).
template <int order>
void addTransformStateGradSens(const double xi[2], const double wr[4], const double Ji[4], double *sens) {
  constexpr int N = (order + 1) * (order + 1);
  double temp[N][2];
  for (int a = 0; a < N; a++) temp[a][0] = temp[a][1] = 0.0;
  for (int ny = 0; ny <= order; ny++) {
    for (int nx = 0; nx <= order; nx++) {
      const int a = ny * (order + 1) + nx;
      double d[2];
      lagrange2dDeriv<order>(xi, nx, ny, d);
      for (int j = 0; j < 2; j++) {
        for (int k = 0; k < 2; k++) temp[a][j] += d[k] * Ji[k * 2 + j];
      }
    }
  }
  for (int a = 0; a < N; a++) {
    for (int i = 0; i < 2; i++) {
      double sum = 0.0;
      for (int k = 0; k < 2; k++) sum += temp[a][k] * wr[i * 2 + k];
      sens[a * 2 + i] += sum;
    }
  }
}

inline void atomic_add(double *p, double v, bool atomic) {
  if (atomic) {
#pragma omp atomic
    *p += v;
  } else {
    *p += v;
  }
}

// gatherElementData (FEKernels.h:25-42).
template <int order>
void gather(const int *connPtr, const int *conn, int e, const double *states, const double *coords, double *lu,
            double *lx) {
  constexpr int N = (order + 1) * (order + 1);
  for (int a = 0; a < N; a++) {
    const int n = conn[connPtr[e] + a];
    lx[a * 2] = coords[2 * n];
    lx[a * 2 + 1] = coords[2 * n + 1];
    lu[a * 2] = states[2 * n];
    lu[a * 2 + 1] = states[2 * n + 1];
  }
}

// Residual of one element (FEKernels.h:317-377 loop body); NL selects the NONLINEAR strain.
template <int order, bool NL = false>
void elementResidual(const double *lu, const double *lx, double E, double nu, double t, double *lres) {
  constexpr int N = (order + 1) * (order + 1);
  for (int k = 0; k < 2 * N; k++) lres[k] = 0.0;
  for (int qx = 0; qx <= order; qx++) {
    for (int qy = 0; qy <= order; qy++) {
      const double w = kGaussW[order][qx] * kGaussW[order][qy];
      const double xi[2] = {kGaussX[order][qx], kGaussX[order][qy]};
      double J[4], Ji[4], dudxi[4], dudx[4], wr[4];
      interpParamGradient<order>(xi, lx, J);
      inv2(J, Ji);
      const double detJ = det2(J);
      interpParamGradient<order>(xi, lu, dudxi);
      matmul2(dudxi, Ji, dudx);
      if (NL) weakResNonlinear(dudx, E, nu, t, w * detJ, wr);
      else weakRes(dudx, E, nu, t, w * detJ, wr);
      addTransformStateGradSens<order>(xi, wr, Ji, lres);
    }
  }
}

// Residual and Jacobian of one element (FEKernels.h:379-508 loop body). localMat is stored
// transposed, localMat[col*2N + row], as the reference does (:492). NL selects the NONLINEAR strain, whose
// tangent depends on the state.
template <int order, bool NL = false>
void elementJacobian(const double *lu, const double *lx, double E, double nu, double t, double *lres,
                     double *localMat) {
  constexpr int N = (order + 1) * (order + 1);
  constexpr int D = 2 * N;
  for (int k = 0; k < D; k++) lres[k] = 0.0;
  for (int k = 0; k < D * D; k++) localMat[k] = 0.0;
  for (int qx = 0; qx <= order; qx++) {
    for (int qy = 0; qy <= order; qy++) {
      const double w = kGaussW[order][qx] * kGaussW[order][qy];
      const double xi[2] = {kGaussX[order][qx], kGaussX[order][qy]};
      double J[4], Ji[4], dudxi[4], dudx[4], wr[4];
      interpParamGradient<order>(xi, lx, J);
      inv2(J, Ji);
      const double detJ = det2(J);
      interpParamGradient<order>(xi, lu, dudxi);
      matmul2(dudxi, Ji, dudx);
      if (NL) weakResNonlinear(dudx, E, nu, t, w * detJ, wr);
      else weakRes(dudx, E, nu, t, w * detJ, wr);
      addTransformStateGradSens<order>(xi, wr, Ji, lres);
      for (int ny = 0; ny <= order; ny++) {
        for (int nx = 0; nx <= order; nx++) {
          const int a = ny * (order + 1) + nx;
          double dxiDot[2];
          lagrange2dDeriv<order>(xi, nx, ny, dxiDot);
          double dxDot[2];
          for (int j = 0; j < 2; j++) dxDot[j] = dxiDot[0] * Ji[j] + dxiDot[1] * Ji[2 + j];
          for (int s = 0; s < 2; s++) {
            double Udot[4] = {0, 0, 0, 0};
            Udot[s * 2] = dxDot[0];
            Udot[s * 2 + 1] = dxDot[1];
            double wrDot[4];
            if (NL) weakResNonlinearFwd(dudx, Udot, E, nu, t, w * detJ, wrDot);
            else weakRes(Udot, E, nu, t, w * detJ, wrDot);
            double col[D];
            for (int k = 0; k < D; k++) col[k] = 0.0;
            addTransformStateGradSens<order>(xi, wrDot, Ji, col);
            const int c = a * 2 + s;
            for (int r = 0; r < D; r++) localMat[c * D + r] += col[r];
          }
        }
      }
    }
  }
}

// Independent Gram-form element: gradN at every quadrature point, then
// K_(a,i),(b,k) = sum_q s [lambda gN_ai gN_bk + mu (delta_ik gN_a.gN_b + gN_ak gN_bi)], stored row-major.
template <int order>
void elementGram(const double *lu, const double *lx, double E, double nu, double t, double *lres, double *K) {
  constexpr int NP = order + 1, N = NP * NP, D = 2 * N;
  const double mu = 0.5 * E / (1.0 + nu);
  const double lambda = 2 * mu * nu / (1.0 - nu);
  for (int k = 0; k < D; k++) lres[k] = 0.0;
  for (int k = 0; k < D * D; k++) K[k] = 0.0;
  for (int qy = 0; qy < NP; qy++) {
    for (int qx = 0; qx < NP; qx++) {
      const double xi[2] = {kGaussX[order][qx], kGaussX[order][qy]};
      double g[N][2];
      for (int a = 0; a < N; a++) lagrange2dDeriv<order>(xi, a % NP, a / NP, g[a]);
      double J[4] = {0, 0, 0, 0};
      for (int a = 0; a < N; a++) {
        J[0] += lx[2 * a] * g[a][0];
        J[1] += lx[2 * a] * g[a][1];
        J[2] += lx[2 * a + 1] * g[a][0];
        J[3] += lx[2 * a + 1] * g[a][1];
      }
      const double detJ = J[0] * J[3] - J[1] * J[2];
      const double Ji[4] = {J[3] / detJ, -J[1] / detJ, -J[2] / detJ, J[0] / detJ};
      double gN[N][2];
      double Gu[4] = {0, 0, 0, 0};
      for (int a = 0; a < N; a++) {
        gN[a][0] = g[a][0] * Ji[0] + g[a][1] * Ji[2];
        gN[a][1] = g[a][0] * Ji[1] + g[a][1] * Ji[3];
        for (int i = 0; i < 2; i++) {
          for (int j = 0; j < 2; j++) Gu[i * 2 + j] += lu[2 * a + i] * gN[a][j];
        }
      }
      const double s = kGaussW[order][qx] * kGaussW[order][qy] * detJ * t;
      const double tr = Gu[0] + Gu[3];
      const double sig[2][2] = {{2 * mu * Gu[0] + lambda * tr, mu * (Gu[1] + Gu[2])},
                                {mu * (Gu[1] + Gu[2]), 2 * mu * Gu[3] + lambda * tr}};
      for (int a = 0; a < N; a++) {
        for (int i = 0; i < 2; i++) lres[2 * a + i] += s * (sig[i][0] * gN[a][0] + sig[i][1] * gN[a][1]);
      }
      for (int a = 0; a < N; a++) {
        for (int b = 0; b < N; b++) {
          const double dot = gN[a][0] * gN[b][0] + gN[a][1] * gN[b][1];
          for (int i = 0; i < 2; i++) {
            for (int k = 0; k < 2; k++) {
              const double v = lambda * gN[a][i] * gN[b][k] + mu * ((i == k ? dot : 0.0) + gN[a][k] * gN[b][i]);
              K[(2 * a + i) * D + 2 * b + k] += s * v;
            }
          }
        }
      }
    }
  }
}

int set_threads(int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
  return threads;
#else
  (void)threads;
  return 1;
#endif
}

template <int order, bool NL = false>
int residualT(const int *connPtr, const int *conn, int first, int count, const double *states, const double *coords,
              double E, double nu, double t, double *residual, int threads) {
  constexpr int N = (order + 1) * (order + 1);
  const int nt = set_threads(threads);
  const bool atomic = nt > 1;
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
  for (int e = first; e < first + count; e++) {
    double lu[2 * N], lx[2 * N], lres[2 * N];
    gather<order>(connPtr, conn, e, states, coords, lu, lx);
    elementResidual<order, NL>(lu, lx, E, nu, t, lres);
    for (int a = 0; a < N; a++) {
      const int n = conn[connPtr[e] + a];
      for (int s = 0; s < 2; s++) atomic_add(&residual[2 * n + s], lres[2 * a + s], atomic);
    }
  }
  return 0;
}

template <int order, bool gram, bool NL = false>
int jacobianT(const int *connPtr, const int *conn, int first, int count, const double *states, const double *coords,
              double E, double nu, double t, const int *map, double *residual, double *values, int threads) {
  constexpr int N = (order + 1) * (order + 1);
  constexpr int D = 2 * N;
  const int nt = set_threads(threads);
  const bool atomic = nt > 1;
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
  for (int e = first; e < first + count; e++) {
    double lu[D], lx[D], lres[D];
    std::vector<double> M(size_t(D) * D);
    gather<order>(connPtr, conn, e, states, coords, lu, lx);
    if (gram) {
      elementGram<order>(lu, lx, E, nu, t, lres, M.data());
    } else {
      elementJacobian<order, NL>(lu, lx, E, nu, t, lres, M.data());
    }
    // scatterMat (FEKernels.h:63-104): block (i, j) entry (r, c) += localMat[(2i+r)*2N + 2j+c].
    const int *m = map + int64_t(e) * N * N;
    for (int i = 0; i < N; i++) {
      for (int j = 0; j < N; j++) {
        double *blk = values + m[i * N + j];
        for (int r = 0; r < 2; r++) {
          for (int c = 0; c < 2; c++) atomic_add(&blk[r * 2 + c], M[(2 * i + r) * D + 2 * j + c], atomic);
        }
      }
    }
    if (residual) {
      for (int a = 0; a < N; a++) {
        const int n = conn[connPtr[e] + a];
        for (int s = 0; s < 2; s++) atomic_add(&residual[2 * n + s], lres[2 * a + s], atomic);
      }
    }
  }
  return 0;
}

template <int order, bool NL = false>
int elementT(const double *lx, const double *lu, double E, double nu, double t, int gram, double *lres, double *K) {
  constexpr int D = 2 * (order + 1) * (order + 1);
  if (gram && !NL) {
    elementGram<order>(lu, lx, E, nu, t, lres, K);
  } else {
    std::vector<double> M(size_t(D) * D);
    elementJacobian<order, NL>(lu, lx, E, nu, t, lres, M.data());
    // Return row-major K: K[r*D + c] = localMat[c*D + r].
    for (int r = 0; r < D; r++) {
      for (int c = 0; c < D; c++) K[r * D + c] = M[c * D + r];
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int oracle_gauss(int order, int i, double *x, double *w) {
  if (order < 1 || order > 4 || i < 0 || i > order) return -1;
  *x = kGaussX[order][i];
  *w = kGaussW[order][i];
  return 0;
}

int oracle_lagrange1d(int order, double x, int k, double *N, double *dN) {
  if (order < 1 || order > 4 || k < 0 || k > order) return -1;
  switch (order) {
    case 1: lagrange1dDeriv<1>(x, k, *N, *dN); break;
    case 2: lagrange1dDeriv<2>(x, k, *N, *dN); break;
    case 3: lagrange1dDeriv<3>(x, k, *N, *dN); break;
    default: lagrange1dDeriv<4>(x, k, *N, *dN); break;
  }
  return 0;
}

// Residual and stiffness of one element from its nodal coords/states (2N each); K row-major 2N x 2N.
// gram = 0: the faithful per-column loop nest; gram = 1: the independent Gram form.
int oracle_element(int order, const double *lx, const double *lu, double E, double nu, double t, int gram,
                   double *lres, double *K) {
  switch (order) {
    case 1: return elementT<1>(lx, lu, E, nu, t, gram, lres, K);
    case 2: return elementT<2>(lx, lu, E, nu, t, gram, lres, K);
    case 3: return elementT<3>(lx, lu, E, nu, t, gram, lres, K);
    case 4: return elementT<4>(lx, lu, E, nu, t, gram, lres, K);
    default: return -1;
  }
}

// assemblePlaneStressResidual over elements [first, first+count): residual += R.
int oracle_residual(int order, const int *connPtr, const int *conn, int first, int count, const double *states,
                    const double *coords, double E, double nu, double t, double *residual, int threads) {
  switch (order) {
    case 1: return residualT<1>(connPtr, conn, first, count, states, coords, E, nu, t, residual, threads);
    case 2: return residualT<2>(connPtr, conn, first, count, states, coords, E, nu, t, residual, threads);
    case 3: return residualT<3>(connPtr, conn, first, count, states, coords, E, nu, t, residual, threads);
    case 4: return residualT<4>(connPtr, conn, first, count, states, coords, E, nu, t, residual, threads);
    default: return -1;
  }
}

// assemblePlaneStressJacobian over elements [first, first+count): values += K through map; residual
// (if non-NULL) += R. gram selects the cross-check formulation.
int oracle_jacobian(int order, const int *connPtr, const int *conn, int first, int count, const double *states,
                    const double *coords, double E, double nu, double t, const int *map, double *residual,
                    double *values, int gram, int threads) {
#define ORACLE_JAC(P)                                                                                        \
  return gram ? jacobianT<P, true>(connPtr, conn, first, count, states, coords, E, nu, t, map, residual,   \
                                   values, threads)                                                          \
              : jacobianT<P, false>(connPtr, conn, first, count, states, coords, E, nu, t, map, residual,  \
                                    values, threads)
  switch (order) {
    case 1: ORACLE_JAC(1);
    case 2: ORACLE_JAC(2);
    case 3: ORACLE_JAC(3);
    case 4: ORACLE_JAC(4);
    default: return -1;
  }
#undef ORACLE_JAC
}

// NONLINEAR-strain versions of oracle_element / oracle_residual / oracle_jacobian (the reference's kernels
// with GreenStrainType::NONLINEAR; its dispatchers instantiate LINEAR only).
int oracle_element_nl(int order, const double *lx, const double *lu, double E, double nu, double t, double *lres,
                      double *K) {
  switch (order) {
    case 1: return elementT<1, true>(lx, lu, E, nu, t, 0, lres, K);
    case 2: return elementT<2, true>(lx, lu, E, nu, t, 0, lres, K);
    case 3: return elementT<3, true>(lx, lu, E, nu, t, 0, lres, K);
    case 4: return elementT<4, true>(lx, lu, E, nu, t, 0, lres, K);
    default: return -1;
  }
}

int oracle_residual_nl(int order, const int *connPtr, const int *conn, int first, int count, const double *states,
                       const double *coords, double E, double nu, double t, double *residual, int threads) {
  switch (order) {
    case 1: return residualT<1, true>(connPtr, conn, first, count, states, coords, E, nu, t, residual, threads);
    case 2: return residualT<2, true>(connPtr, conn, first, count, states, coords, E, nu, t, residual, threads);
    case 3: return residualT<3, true>(connPtr, conn, first, count, states, coords, E, nu, t, residual, threads);
    case 4: return residualT<4, true>(connPtr, conn, first, count, states, coords, E, nu, t, residual, threads);
    default: return -1;
  }
}

int oracle_jacobian_nl(int order, const int *connPtr, const int *conn, int first, int count, const double *states,
                       const double *coords, double E, double nu, double t, const int *map, double *residual,
                       double *values, int threads) {
#define ORACLE_JAC_NL(P)                                                                                     \
  return jacobianT<P, false, true>(connPtr, conn, first, count, states, coords, E, nu, t, map, residual,    \
                                   values, threads)
  switch (order) {
    case 1: ORACLE_JAC_NL(1);
    case 2: ORACLE_JAC_NL(2);
    case 3: ORACLE_JAC_NL(3);
    case 4: ORACLE_JAC_NL(4);
    default: return -1;
  }
#undef ORACLE_JAC_NL
}

// Brute-force block pattern: the set of (conn[e,i], conn[e,j]) pairs, rows sorted. Two calls: with
// cols == NULL it fills rowp only.
int oracle_bsr_pattern(int numNodes, int numElements, int npe, const int *connPtr, const int *conn, int *rowp,
                       int *cols) {
  std::vector<std::set<int>> rows(static_cast<size_t>(numNodes));
  for (int e = 0; e < numElements; e++) {
    for (int i = 0; i < npe; i++) {
      for (int j = 0; j < npe; j++) rows[size_t(conn[connPtr[e] + i])].insert(conn[connPtr[e] + j]);
    }
  }
  rowp[0] = 0;
  for (int n = 0; n < numNodes; n++) {
    if (rows[size_t(n)].empty()) rows[size_t(n)].insert(n);
    rowp[n + 1] = rowp[n] + int(rows[size_t(n)].size());
  }
  if (cols) {
    for (int n = 0; n < numNodes; n++) std::copy(rows[size_t(n)].begin(), rows[size_t(n)].end(), cols + rowp[n]);
  }
  return 0;
}

// generateElementBCSRMap (TACSHelpers.h:138-182): linear search along the row. Returns the number of
// (element, i, j) pairs with no block (the reference prints and leaves them unset; here they get -1).
int oracle_element_bcsr_map(const int *connPtr, const int *conn, int numElements, int npe, const int *rowp,
                            const int *cols, int bsize, int *map) {
  int misses = 0;
  const int blockLength = bsize * bsize;
  for (int e = 0; e < numElements; e++) {
    const int *en = &conn[connPtr[e]];
    for (int i = 0; i < npe; i++) {
      const int r = en[i];
      for (int j = 0; j < npe; j++) {
        int found = -1;
        for (int jj = rowp[r]; jj < rowp[r + 1]; jj++) {
          if (cols[jj] == en[j]) {
            found = jj * blockLength;
            break;
          }
        }
        if (found < 0) misses++;
        map[int64_t(e) * npe * npe + i * npe + j] = found;
      }
    }
  }
  return misses;
}

// CPU baseline: wall seconds of one residual + Jacobian assembly (faithful form) over elements
// [first, first+count) with `threads` OpenMP threads (<= 0: all). Outputs are accumulated.
double oracle_time_res_jac(int order, const int *connPtr, const int *conn, int first, int count,
                           const double *states, const double *coords, double E, double nu, double t, const int *map,
                           double *residual, double *values, int threads) {
  const auto t0 = std::chrono::steady_clock::now();
  if (oracle_residual(order, connPtr, conn, first, count, states, coords, E, nu, t, residual, threads)) return -1.0;
  if (oracle_jacobian(order, connPtr, conn, first, count, states, coords, E, nu, t, map, nullptr, values, 0,
                      threads)) {
    return -1.0;
  }
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
