"""Oracle pins: quadrature, basis, geometry, element matrices.

Each check is independent of the oracle's own formulas: golden logs of the
reference (tests/golden/reference_logs.json), exact rational integration,
closed forms and invariants.
"""
import itertools
import json
import math
import os
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_logs.json")))


def exact_monomial(a, b, c):
    # integral of x^a y^b z^c over the unit tetrahedron
    return Fr(math.factorial(a) * math.factorial(b) * math.factorial(c),
              math.factorial(a + b + c + 3))


@pytest.mark.parametrize("npts,degree", [(1, 1), (4, 2), (11, 4), (15, 5)])
def test_quadrature_exactness(npts, degree):
    pts, w = O.quad_rule(npts)
    assert abs(w.sum() - 1.0 / 6.0) < 1e-15
    for a, b, c in itertools.product(range(degree + 1), repeat=3):
        if a + b + c > degree:
            continue
        q = float(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c))
        assert abs(q - float(exact_monomial(a, b, c))) < 1e-15, (a, b, c)


@pytest.mark.parametrize("npts,degree", [(1, 1), (4, 2), (11, 4), (15, 5)])
def test_quadrature_not_exact_beyond(npts, degree):
    # a plausible table mistake (wrong point, wrong weight) that keeps lower
    # degrees exact is caught by the degree above failing in a known way
    pts, w = O.quad_rule(npts)
    worst = 0.0
    for a, b, c in itertools.product(range(degree + 2), repeat=3):
        if a + b + c != degree + 1:
            continue
        q = float(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c))
        worst = max(worst, abs(q - float(exact_monomial(a, b, c))))
    assert worst > 1e-8


def test_keast_negative_centroid_weight():
    pts, w = O.quad_rule(11)
    assert np.allclose(pts[0], 0.25) and w[0] < 0 and abs(w[0] + 74.0 / 5625.0) < 1e-18


@pytest.mark.parametrize("name,order", [("fePr3dTest2", 1), ("fePr3dTest3", 2)])
def test_basis_golden_logs(name, order):
    g = GOLD[name]
    v, grad = O.basis(order, g["Y"])
    np.testing.assert_allclose(v, g["LocEval"], atol=1e-12)
    np.testing.assert_allclose(grad[0], g["LocGradX"], atol=1e-12)
    np.testing.assert_allclose(grad[1], g["LocGradY"], atol=1e-12)
    np.testing.assert_allclose(grad[2], g["LocGradZ"], atol=1e-12)
    gg = O.global_grad(order, np.array(g["nodes"], float), g["Y"])
    np.testing.assert_allclose(gg[0], g["GlobalGradX"], atol=1e-12)
    np.testing.assert_allclose(gg[1], g["GlobalGradY"], atol=1e-12)
    np.testing.assert_allclose(gg[2], g["GlobalGradZ"], atol=1e-12)


def test_geometry_golden_log():
    g = GOLD["geoMapInterfaceTest5"]
    P = np.array(g["nodes"], float)
    J = O.jacobian(P)
    np.testing.assert_allclose(J, g["gradient"], atol=1e-15)
    # position x(Y) = P1 + J Y
    np.testing.assert_allclose(P[0] + J @ np.array(g["Y"]), g["eval"], atol=1e-15)


def test_jacobian_columns_not_rows():
    # a transposed J would pass the identity golden log; this tet is not symmetric
    P = np.array([[0, 0, 0], [1, 2, 0], [0, 1, 0], [3, 0, 1]], float)
    J = O.jacobian(P)
    np.testing.assert_allclose(J[:, 0], P[1] - P[0])
    np.testing.assert_allclose(J[:, 2], P[3] - P[0])


def test_basis_partition_and_kronecker():
    rng = np.random.default_rng(0)
    for _ in range(5):
        y = rng.random(3) / 3
        for order in (1, 2):
            v, g = O.basis(order, y)
            assert abs(v.sum() - 1.0) < 1e-14
            np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-13)
    nodes = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (.5, 0, 0), (.5, .5, 0), (0, .5, 0),
             (0, 0, .5), (.5, 0, .5), (0, .5, .5)]
    for i, y in enumerate(nodes):
        v, _ = O.basis(2, y)
        e = np.zeros(10)
        e[i] = 1
        np.testing.assert_allclose(v, e, atol=1e-15)


# ---- exact element matrices with rational polynomial algebra ---------------------
def poly_mul(p, q):
    r = {}
    for (a, ca) in p.items():
        for (b, cb) in q.items():
            k = (a[0] + b[0], a[1] + b[1], a[2] + b[2])
            r[k] = r.get(k, 0) + ca * cb
    return r


def poly_int(p):
    return sum(c * exact_monomial(*k) for k, c in p.items())


def poly_d(p, axis):
    r = {}
    for k, c in p.items():
        if k[axis] > 0:
            kk = list(k)
            kk[axis] -= 1
            r[tuple(kk)] = r.get(tuple(kk), 0) + c * k[axis]
    return r


def lam(i):
    return {0: {(0, 0, 0): Fr(1), (1, 0, 0): Fr(-1), (0, 1, 0): Fr(-1), (0, 0, 1): Fr(-1)},
            1: {(1, 0, 0): Fr(1)}, 2: {(0, 1, 0): Fr(1)}, 3: {(0, 0, 1): Fr(1)}}[i]


def exact_basis(order):
    if order == 1:
        return [lam(i) for i in range(4)]
    B = []
    for i in range(4):
        li = lam(i)
        t = {k: 2 * c for k, c in li.items()}
        t[(0, 0, 0)] = t.get((0, 0, 0), 0) - 1
        B.append(poly_mul(li, t))
    for a, b in [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]:
        B.append({k: 4 * c for k, c in poly_mul(lam(a), lam(b)).items()})
    return B


def exact_tensors(order):
    B = exact_basis(order)
    n = len(B)
    dB = [[poly_d(b, ax) for ax in range(3)] for b in B]
    S = np.zeros((3, 3, n, n))
    M = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            M[i, j] = float(poly_int(poly_mul(B[i], B[j])))
            for a in range(3):
                for b in range(3):
                    S[a, b, i, j] = float(poly_int(poly_mul(dB[i][a], dB[j][b])))
    return S, M


@pytest.fixture(scope="module")
def tensors():
    return {1: exact_tensors(1), 2: exact_tensors(2)}


def random_tet(rng):
    while True:
        P = rng.random((4, 3))
        if np.linalg.det((P[1:] - P[0]).T) > 0.02:
            return P


@pytest.mark.parametrize("order", [1, 2])
def test_element_matrices_exact(order, tensors):
    S, M = tensors[order]
    rng = np.random.default_rng(1)
    for _ in range(4):
        P = random_tet(rng)
        J = (P[1:] - P[0]).T
        det = np.linalg.det(J)
        Ji = np.linalg.inv(J)
        A = Ji @ Ji.T
        Kx = det * np.einsum("ab,abij->ij", A, S)
        Mx = det * M
        c = 1.7
        for rule in ([1, 4, 11, 15] if order == 1 else [4, 11, 15]):
            K = O.element_matrix(O.OP_STIFF, order, P, c_stiff=c, rule_stiff=rule)
            np.testing.assert_allclose(K, c * Kx, rtol=0, atol=1e-12 * np.abs(Kx).max())
        for rule in ([4, 11, 15] if order == 1 else [11, 15]):
            Mm = O.element_matrix(O.OP_MASS, order, P, c_mass=c, rule_mass=rule)
            np.testing.assert_allclose(Mm, c * Mx, rtol=0, atol=1e-12 * np.abs(Mx).max())


def test_p1_closed_forms():
    rng = np.random.default_rng(2)
    P = random_tet(rng)
    J = (P[1:] - P[0]).T
    vol = np.linalg.det(J) / 6
    G = np.vstack([-np.linalg.inv(J).sum(axis=0), np.linalg.inv(J)])  # rows grad lambda
    K = O.element_matrix(O.OP_STIFF, 1, P, rule_stiff=1)
    np.testing.assert_allclose(K, vol * G @ G.T, atol=1e-13)
    M = O.element_matrix(O.OP_MASS, 1, P, rule_mass=4)
    np.testing.assert_allclose(M, vol / 20 * (np.ones((4, 4)) + np.eye(4)), atol=1e-15)


@pytest.mark.parametrize("order", [1, 2])
def test_element_invariants(order):
    rng = np.random.default_rng(3)
    P = random_tet(rng)
    n = 4 if order == 1 else 10
    K = O.element_matrix(O.OP_STIFF, order, P, rule_stiff=15)
    M = O.element_matrix(O.OP_MASS, order, P, rule_mass=15)
    vol = np.linalg.det((P[1:] - P[0]).T) / 6
    np.testing.assert_allclose(K, K.T, atol=1e-13)
    np.testing.assert_allclose(K @ np.ones(n), 0, atol=1e-12)
    assert abs(M.sum() - vol) < 1e-14
    assert np.linalg.eigvalsh(K).min() > -1e-12


def test_reference_p2_stiffness_known_entries(tensors):
    # reference tet, k = 1: diagonal of vertex 2 (x(2x-1)) is int (4x-1)^2 = 1/10
    S, _ = tensors[2]
    K = S[0, 0] + S[1, 1] + S[2, 2]
    assert abs(K[1, 1] - 0.1) < 1e-15
    Ko = O.element_matrix(O.OP_STIFF, 2, np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.]]),
                          rule_stiff=4)
    np.testing.assert_allclose(Ko, K, atol=1e-14)
