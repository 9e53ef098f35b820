"""Independent textbook plane-stress assembler (numpy), used to pin the oracle.

It shares no code or formulation with oracle/oracle.cpp or the library:
  * Gauss points/weights from numpy.polynomial.legendre.leggauss (full double precision);
  * 1-D Lagrange basis from the product formula on the cosine knots -cos(pi j / p), derivative by
    numpy polynomial differentiation;
  * stiffness as K_e = sum_q w_q detJ_q t B_q^T D B_q with the Voigt strain-displacement matrix B and the
    plane-stress constitutive matrix D = E/(1-nu^2) [[1, nu, 0], [nu, 1, 0], [0, 0, (1-nu)/2]];
  * dense global assembly; the residual is K u (linear strain).
Meant for meshes of at most a few hundred nodes.
"""

from __future__ import annotations

import numpy as np
from numpy.polynomial import legendre
from numpy.polynomial import polynomial as P


def cosine_knots(p: int) -> np.ndarray:
    return -np.cos(np.pi * np.arange(p + 1) / p)


def lagrange_polys(p: int):
    """Monomial coefficients (lowest degree first) of the p+1 Lagrange polynomials."""
    x = cosine_knots(p)
    polys = []
    for j in range(p + 1):
        c = np.array([1.0])
        for i in range(p + 1):
            if i != j:
                c = P.polymul(c, np.array([-x[i], 1.0]) / (x[j] - x[i]))
        polys.append(c)
    return polys


def basis_1d(p: int, xi: np.ndarray):
    """N[k, i] = N_k(xi_i), dN[k, i] = N_k'(xi_i)."""
    polys = lagrange_polys(p)
    N = np.array([P.polyval(xi, c) for c in polys])
    dN = np.array([P.polyval(xi, P.polyder(c)) for c in polys])
    return N, dN


def plane_stress_D(E: float, nu: float) -> np.ndarray:
    return E / (1 - nu**2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])


def element_stiffness(p: int, xy: np.ndarray, E: float, nu: float, t: float) -> np.ndarray:
    """K_e (2N x 2N, DOF order (node a, component i) -> 2a+i) for node coords xy[N, 2]."""
    NP = p + 1
    N = NP * NP
    xg, wg = legendre.leggauss(NP)
    Nb, dNb = basis_1d(p, xg)
    D = plane_stress_D(E, nu)
    K = np.zeros((2 * N, 2 * N))
    for qy in range(NP):
        for qx in range(NP):
            # dN_a/dxi, dN_a/deta for a = ay*NP + ax
            g = np.zeros((N, 2))
            for ay in range(NP):
                for ax in range(NP):
                    a = ay * NP + ax
                    g[a, 0] = dNb[ax, qx] * Nb[ay, qy]
                    g[a, 1] = Nb[ax, qx] * dNb[ay, qy]
            J = xy.T @ g  # J[i, j] = d x_i / d xi_j
            detJ = np.linalg.det(J)
            gN = g @ np.linalg.inv(J)  # rows: physical gradient of N_a
            B = np.zeros((3, 2 * N))
            B[0, 0::2] = gN[:, 0]
            B[1, 1::2] = gN[:, 1]
            B[2, 0::2] = gN[:, 1]
            B[2, 1::2] = gN[:, 0]
            K += wg[qx] * wg[qy] * detJ * t * (B.T @ D @ B)
    return K


def assemble_dense(p: int, coords: np.ndarray, conn: np.ndarray, E: float, nu: float, t: float) -> np.ndarray:
    n = coords.shape[0]
    K = np.zeros((2 * n, 2 * n))
    for e in range(conn.shape[0]):
        nodes = conn[e]
        Ke = element_stiffness(p, coords[nodes], E, nu, t)
        dof = np.empty(2 * len(nodes), dtype=np.int64)
        dof[0::2] = 2 * nodes
        dof[1::2] = 2 * nodes + 1
        K[np.ix_(dof, dof)] += Ke
    return K


def assemble_bsr(p: int, coords: np.ndarray, conn: np.ndarray, rowp: np.ndarray, cols: np.ndarray, E: float,
                 nu: float, t: float) -> np.ndarray:
    """BSR values (2x2 row-major blocks) of the global stiffness, each element's K_e added block by block at the
    slot found by a search of the sorted columns of its row (for meshes too large for the dense form)."""
    vals = np.zeros(len(cols) * 4)
    for e in range(conn.shape[0]):
        nodes = conn[e]
        Ke = element_stiffness(p, coords[nodes], E, nu, t)
        for i, r in enumerate(nodes):
            row = cols[rowp[r]:rowp[r + 1]]
            slots = rowp[r] + np.searchsorted(row, nodes)
            assert np.array_equal(cols[slots], nodes)
            for j, k in enumerate(slots):
                vals[4 * k: 4 * k + 4] += Ke[2 * i: 2 * i + 2, 2 * j: 2 * j + 2].reshape(-1)
    return vals


def dense_to_bsr_values(K: np.ndarray, rowp: np.ndarray, cols: np.ndarray) -> np.ndarray:
    """Extract 2x2 blocks of a dense matrix in BSR (row-major block) order."""
    vals = np.zeros(len(cols) * 4)
    for r in range(len(rowp) - 1):
        for k in range(rowp[r], rowp[r + 1]):
            c = cols[k]
            vals[4 * k: 4 * k + 4] = K[2 * r: 2 * r + 2, 2 * c: 2 * c + 2].reshape(-1)
    return vals


def bsr_to_dense(n: int, rowp: np.ndarray, cols: np.ndarray, vals: np.ndarray) -> np.ndarray:
    K = np.zeros((2 * n, 2 * n))
    for r in range(n):
        for k in range(rowp[r], rowp[r + 1]):
            c = cols[k]
            K[2 * r: 2 * r + 2, 2 * c: 2 * c + 2] += vals[4 * k: 4 * k + 4].reshape(2, 2)
    return K


def bsr_spmv(rowp, cols, vals, x):
    n = len(rowp) - 1
    A = vals.reshape(-1, 2, 2)
    xb = x.reshape(-1, 2)
    y = np.zeros((n, 2))
    rows = np.repeat(np.arange(n), np.diff(rowp))
    np.add.at(y, rows, np.einsum("kij,kj->ki", A, xb[cols]))
    return y.reshape(-1)


def normwise_rel(x: np.ndarray, ref: np.ndarray) -> float:
    d = np.max(np.abs(np.asarray(x) - np.asarray(ref)))
    s = np.max(np.abs(ref))
    return float(d / s) if s > 0 else float(d)
