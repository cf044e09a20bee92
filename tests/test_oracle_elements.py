"""Pins of the oracle's element-level functions by mathematics that does not
reuse the oracle's formulas: exact Gauss quadrature of the trilinear map,
finite differences, closed forms on cubes and parallelepipeds, and
invariants (translation, linear completeness, Euler's theorem).

Each check is chosen so that a dropped term, a wrong sign, index or operand
order in the oracle makes it fail.
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import dptr, lib

# Reference local node order (lulesh-init.cc:237-244): (xi, eta, zeta) signs
SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
UNIT = (SIGNS + 1) / 2  # unit cube corners

# Faces in the reference's order (lulesh.cc:421-450)
FACES = [(0, 1, 2, 3), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0), (4, 7, 6, 5)]


def xyz(P):
    P = np.ascontiguousarray(P, dtype=np.float64)
    return [np.ascontiguousarray(P[:, k]) for k in range(3)]


def random_hex(rng, distortion=0.15, scale=1.0):
    A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    P = UNIT @ A.T * scale + rng.standard_normal(3)
    return P + distortion * scale * rng.standard_normal((8, 3))


def jac(P, xi):
    """Jacobian of the trilinear map at reference point xi in [-1,1]^3."""
    J = np.zeros((3, 3))
    for a in range(8):
        s = SIGNS[a]
        dN = np.array([s[0] * (1 + s[1] * xi[1]) * (1 + s[2] * xi[2]),
                       s[1] * (1 + s[0] * xi[0]) * (1 + s[2] * xi[2]),
                       s[2] * (1 + s[0] * xi[0]) * (1 + s[1] * xi[1])]) / 8.0
        J += np.outer(P[a], dN)
    return J


def gauss_volume(P):
    """Exact volume of the trilinear hexahedron: det J is a polynomial of
    degree <= 2 in each reference coordinate, so 2-point Gauss is exact."""
    g = 1 / np.sqrt(3)
    return sum(np.linalg.det(jac(P, (a, b, c))) for a in (-g, g) for b in (-g, g) for c in (-g, g))


def ora_volume(P):
    x, y, z = xyz(P)
    return lib().ora_calc_elem_volume(dptr(x), dptr(y), dptr(z))


def test_volume_unit_cube():
    assert ora_volume(UNIT) == 1.0


def test_volume_parallelepiped_is_determinant():
    rng = np.random.default_rng(1)
    for _ in range(20):
        A = rng.standard_normal((3, 3))
        if np.linalg.det(A) < 0.1:
            A[:, 0] *= -1
        P = UNIT @ A.T + rng.standard_normal(3)
        assert ora_volume(P) == pytest.approx(np.linalg.det(A), rel=1e-13)


def test_volume_general_hex_matches_gauss():
    rng = np.random.default_rng(2)
    for _ in range(50):
        P = random_hex(rng)
        assert ora_volume(P) == pytest.approx(gauss_volume(P), rel=1e-12)


def test_volume_derivative_is_exact_gradient():
    """dV/dx_a: the volume is linear in any single coordinate, so a central
    difference of the oracle's own volume is exact up to rounding."""
    rng = np.random.default_rng(3)
    L = lib()
    for _ in range(20):
        P = random_hex(rng)
        x, y, z = xyz(P)
        dv = np.zeros(24)
        L.ora_calc_elem_volume_derivative(dptr(dv), dptr(x), dptr(y), dptr(z))
        h = 1e-3
        for comp in range(3):
            for a in range(8):
                Pp, Pm = P.copy(), P.copy()
                Pp[a, comp] += h
                Pm[a, comp] -= h
                fd = (ora_volume(Pp) - ora_volume(Pm)) / (2 * h)
                assert dv[8 * comp + a] == pytest.approx(fd, abs=1e-10)
        # Euler: V is homogeneous of degree 1 in the x components
        assert np.dot(dv[0:8], x) == pytest.approx(ora_volume(P), rel=1e-12)
        assert abs(np.dot(dv[0:8], y)) < 1e-12
        assert abs(dv[0:8].sum()) < 1e-13


def shape(P):
    x, y, z = xyz(P)
    b = np.zeros(24)
    vol = ctypes.c_double()
    lib().ora_calc_elem_shape_function_derivatives(dptr(x), dptr(y), dptr(z), dptr(b),
                                                   ctypes.byref(vol))
    return b.reshape(3, 8), vol.value


def test_shape_function_determinant_is_centre_jacobian():
    rng = np.random.default_rng(4)
    for _ in range(30):
        P = random_hex(rng)
        b, det = shape(P)
        assert det == pytest.approx(8 * np.linalg.det(jac(P, (0, 0, 0))), rel=1e-12)


def test_shape_function_derivatives_linear_completeness():
    """Centre-point derivatives reproduce linear fields: sum_a b[i][a] x_a[j]
    = det * delta_ij, and sum_a b[i][a] = 0, for any hexahedron."""
    rng = np.random.default_rng(5)
    for _ in range(30):
        P = random_hex(rng)
        b, det = shape(P)
        M = b @ P  # M[i][j] = sum_a b[i][a] P[a][j]
        np.testing.assert_allclose(M, det * np.eye(3), atol=1e-12 * abs(det))
        np.testing.assert_allclose(b.sum(axis=1), 0, atol=1e-13)


def normals(P):
    x, y, z = xyz(P)
    pf = np.zeros(24)
    lib().ora_calc_elem_node_normals(dptr(pf), dptr(x), dptr(y), dptr(z))
    return pf.reshape(3, 8)


def test_node_normals_from_face_diagonals():
    """Each face contributes a quarter of its area vector to its four nodes;
    the quad area vector is (1/2) d1 x d2 with d1 = p2-p0, d2 = p3-p1, and in
    the reference's face orientation the quarter enters with sign -1/4 (the
    normal points out of the element)."""
    rng = np.random.default_rng(6)
    for _ in range(30):
        P = random_hex(rng)
        expect = np.zeros((3, 8))
        for f in FACES:
            area = -0.125 * np.cross(P[f[2]] - P[f[0]], P[f[3]] - P[f[1]])
            for a in f:
                expect[:, a] += area
        np.testing.assert_allclose(normals(P), expect, atol=1e-14)


def test_node_normals_unit_cube_outward():
    pf = normals(UNIT)
    # node 6 = (1,1,1): three outward faces -> +1/4 each; node 0 -> -1/4
    np.testing.assert_array_equal(pf[:, 6], [0.25, 0.25, 0.25])
    np.testing.assert_array_equal(pf[:, 0], [-0.25, -0.25, -0.25])
    np.testing.assert_array_equal(pf[:, 1], [0.25, -0.25, -0.25])


def test_node_normals_parallelepiped_divergence_theorem():
    """Closed surface: sum of normals is 0; integral of x n_x over planar
    faces is the volume, and the quarter rule integrates it exactly."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        P = UNIT @ A.T + rng.standard_normal(3)
        pf = normals(P)
        V = np.linalg.det(A)
        np.testing.assert_allclose(pf.sum(axis=1), 0, atol=1e-14)
        np.testing.assert_allclose(pf @ P, V * np.eye(3), atol=1e-13)


def char_length(P):
    x, y, z = xyz(P)
    return lib().ora_calc_elem_characteristic_length(dptr(x), dptr(y), dptr(z), ora_volume(P))


def test_characteristic_length():
    assert char_length(UNIT * 0.3) == pytest.approx(0.3, rel=1e-15)
    rng = np.random.default_rng(8)
    for _ in range(30):
        P = random_hex(rng)
        amax = max(0.5 * np.linalg.norm(np.cross(P[f[2]] - P[f[0]], P[f[3]] - P[f[1]]))
                   for f in [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5),
                             (2, 3, 7, 6), (3, 0, 4, 7)])
        assert char_length(P) == pytest.approx(ora_volume(P) / amax, rel=1e-12)


def test_velocity_gradient_affine_field():
    """For v = A x + c on any hexahedron, the centre-point gradient is
    exact: D = sym(A)."""
    rng = np.random.default_rng(9)
    L = lib()
    for _ in range(30):
        P = random_hex(rng)
        A = rng.standard_normal((3, 3))
        V = P @ A.T + rng.standard_normal(3)
        b, det = shape(P)
        d = np.zeros(6)
        xv, yv, zv = xyz(V)
        L.ora_calc_elem_velocity_gradient(dptr(xv), dptr(yv), dptr(zv),
                                          dptr(np.ascontiguousarray(b.ravel())), det, dptr(d))
        S = 0.5 * (A + A.T)
        np.testing.assert_allclose(d, [A[0, 0], A[1, 1], A[2, 2], S[1, 2], S[0, 2], S[0, 1]],
                                   atol=1e-11)


def hourglass(P, V, determ=None, ss=1.0, mass=1.0, hgcoef=3.0):
    x, y, z = xyz(P)
    xd, yd, zd = xyz(V)
    if determ is None:
        determ = ora_volume(P)
    out = np.zeros(24)
    lib().ora_calc_elem_fb_hourglass(dptr(x), dptr(y), dptr(z), dptr(xd), dptr(yd), dptr(zd),
                                     determ, ss, mass, hgcoef, dptr(out))
    return out.reshape(3, 8)


GAMMA = np.array([[1, 1, -1, -1, -1, -1, 1, 1], [1, -1, -1, 1, -1, 1, 1, -1],
                  [1, -1, 1, -1, 1, -1, 1, -1], [-1, 1, -1, 1, 1, -1, 1, -1]], dtype=float)


def test_hourglass_base_vectors_orthogonal_to_linear_modes():
    """The four hourglass vectors are orthogonal to 1, x, y, z on the cube
    and to each other (independent check of the gamma table)."""
    lin = np.column_stack([np.ones(8), UNIT])
    np.testing.assert_array_equal(GAMMA @ lin, 0)
    np.testing.assert_array_equal(GAMMA @ GAMMA.T, 8 * np.eye(4))


def test_hourglass_zero_for_zero_and_affine_velocity():
    rng = np.random.default_rng(10)
    for _ in range(30):
        P = random_hex(rng)
        assert np.all(hourglass(P, np.zeros((8, 3))) == 0)
        A = rng.standard_normal((3, 3))
        V = P @ A.T + rng.standard_normal(3)
        F = hourglass(P, V)
        assert np.max(np.abs(F)) < 1e-11 * np.max(np.abs(V))


def test_hourglass_cube_closed_form():
    """On a cube the volume correction vanishes and F = coef * G^T (G v) with
    coef = -hgcoef * 0.01 * ss * mass / V^(1/3) (Flanagan-Belytschko)."""
    rng = np.random.default_rng(11)
    h = 0.37
    P = UNIT * h
    for _ in range(10):
        V = rng.standard_normal((8, 3))
        ss, mass = 2.5, 0.7
        F = hourglass(P, V, ss=ss, mass=mass)
        coef = -3.0 * 0.01 * ss * mass / np.cbrt(ora_volume(P))
        expect = coef * (GAMMA.T @ (GAMMA @ V))  # (8,3)
        np.testing.assert_allclose(F, expect.T, rtol=1e-13, atol=1e-15)


def test_hourglass_general_hex_against_fd_projection():
    """General hexahedron: hourgam = gamma - (1/V) grad V (gamma . x), with
    grad V from finite differences of the Gauss-quadrature volume."""
    rng = np.random.default_rng(12)
    for _ in range(10):
        P = random_hex(rng)
        V = rng.standard_normal((8, 3))
        vol = gauss_volume(P)
        grad = np.zeros((8, 3))
        h = 1e-4
        for a in range(8):
            for c in range(3):
                Pp, Pm = P.copy(), P.copy()
                Pp[a, c] += h
                Pm[a, c] -= h
                grad[a, c] = (gauss_volume(Pp) - gauss_volume(Pm)) / (2 * h)
        hg = GAMMA.T - (grad @ (GAMMA @ P).T) / vol  # (8,4)
        coef = -3.0 * 0.01 * 1.0 * 1.0 / np.cbrt(vol)
        expect = coef * hg @ (hg.T @ V)
        F = hourglass(P, V)
        np.testing.assert_allclose(F, expect.T, rtol=1e-8, atol=1e-9 * np.max(np.abs(expect)))


def test_hourglass_forces_sum_to_zero():
    rng = np.random.default_rng(13)
    for _ in range(30):
        P = random_hex(rng)
        F = hourglass(P, rng.standard_normal((8, 3)))
        np.testing.assert_allclose(F.sum(axis=1), 0, atol=1e-13 * np.max(np.abs(F)))


def eos(e, delv, p, q, qq, ql, vnew):
    out = np.zeros(4)
    lib().ora_eval_eos_elem(dptr(np.array([e, delv, p, q, qq, ql, vnew], dtype=np.float64)),
                            dptr(out))
    return out


def test_eos_static_state():
    """delv = 0, q = 0: energy unchanged, p = (2/3) e / v, ss^2 = (10/9) e
    (ideal gas, gamma = 5/3)."""
    for e, v in [(1.0, 1.0), (3.5e6, 0.8), (42.0, 1.3)]:
        en, pn, qn, ss = eos(e, 0.0, 2 / 3 * e / v, 0.0, 0.0, 0.0, v)
        assert en == e
        assert pn == pytest.approx(2 / 3 * e / v, rel=1e-15)
        assert qn == 0
        assert ss == pytest.approx(np.sqrt(10 / 9 * e), rel=1e-15)


@pytest.mark.parametrize("dv", [-1e-3, -1e-2, 1e-3, 1e-2])
def test_eos_follows_adiabat(dv):
    """Without viscosity the predictor-corrector integrates de = -p dv along
    the ideal-gas adiabat e ~ v^(-2/3) with Simpson weights: the error is
    O(dv^3), a wrong weight would leave O(dv)."""
    e0, v0 = 1.0, 1.0
    v1 = v0 + dv
    en, pn, _, _ = eos(e0, dv, 2 / 3 * e0 / v0, 0.0, 0.0, 0.0, v1)
    exact = e0 * (v0 / v1) ** (2 / 3)
    assert abs(en - exact) < 2 * abs(dv) ** 3
    assert pn == pytest.approx(2 / 3 * en / v1, rel=1e-14)


def test_eos_viscous_heating():
    """Compression with a linear q term heats more than the adiabat and
    the returned q is ssc*ql + qq (ssc the post-step sound speed)."""
    e0, dv = 1.0, -1e-2
    v1 = 1 + dv
    en0, _, _, _ = eos(e0, dv, 2 / 3, 0.0, 0.0, 0.0, v1)
    en, pn, qn, ss = eos(e0, dv, 2 / 3, 0.0, 0.1, 0.2, v1)
    assert en > en0
    assert qn == pytest.approx(ss * 0.2 + 0.1, rel=1e-15)


def test_eos_cutoffs():
    en, pn, qn, ss = eos(5e-8, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    assert en == 0.0 and pn == 0.0  # |e| < e_cut -> 0, then p = 0
    assert ss == pytest.approx(0.3333333e-18)  # floor of lulesh.cc:1779-1783
    en, pn, _, _ = eos(1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0e9)
    assert pn == 0.0  # v >= eosvmax -> p = 0
