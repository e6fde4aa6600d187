"""Newton<T>::prepareForSolve's right-hand side in the oracle (oracle.newton_rhs, restating
src/Solver/Nonlocal/Newton.H:30-140), pinned by closed forms:

* u = 0 and F0 = 0: rhs = b exactly (a linear model from u0 = 0 is one Linear solve, Newton.H:142-209);
* uniform stiffness and uniform eigenstrain F0 at u = 0: dw = -C:F0 is uniform, so inside rhs = b exactly,
  and a traction row reads b + (C:F0) n with the reference's truncated normals;
* quadratic u with uniform stiffness: dw is affine, so its central differences are exact and the rhs is
  b - C_ijkl u_k,lj (the strong form) at nodes two or more from every wall (next to a wall the central
  difference reads dw on the wall, where the gradient is one-sided and not exact for a quadratic);
* F0 linear in x with uniform stiffness, u = 0: interior rhs = b + C_ijkl dF0_kl/dx_j exactly;
* smooth non-uniform stiffness and smooth u: the interior rhs converges to b - div(C : grad u) at
  second order.
"""
import numpy as np
import pytest

from tests import helpers as H

ISO, MM, DIAG = 0, 1, 2
S2, S3 = 0.70710678118, 0.57735026919


def iso_tensor(lam, mu):
    d = np.eye(3)
    return lam * np.einsum("ij,kl->ijkl", d, d) + mu * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d))


def normal_of(face):
    n = np.zeros(3)
    parts = face.split("_")
    s = {1: 1.0, 2: S2, 3: S3}[len(parts)]
    for p in parts:
        n["XYZ".index(p[0])] = s if p[1:] == "HI" else -s
    return n


def test_zero_state_returns_b(oracle):
    n = (6, 5, 7)
    g = oracle.Grid(n)
    rng = np.random.default_rng(0)
    C = np.zeros(g.shape(2))
    C[0], C[1] = 1.0 + rng.random(C[0].shape), 0.5 + rng.random(C[1].shape)
    t, _ = oracle.bc_arrays(3, {"ZHI": {0: (1, 0), 1: (3, 0), 2: (1, 0)}}, default=(1, 0.0))
    b = rng.standard_normal(g.shape(3))
    out = oracle.newton_rhs(g, ISO, C, t, g.zeros(), b)
    assert np.array_equal(out, b)


def test_uniform_eigenstrain_traction_rows(oracle):
    n = (5, 6, 4)
    g = oracle.Grid(n)
    lam, mu = 6.0, 2.6
    C = np.zeros(g.shape(2))
    C[0], C[1] = lam, mu
    F0v = np.array([[0.01, 0.002, -0.003], [0.004, -0.02, 0.001], [0.0, 0.005, 0.015]])
    F0 = np.zeros(g.shape(9))
    for a in range(3):
        for c in range(3):
            F0[3 * a + c] = F0v[a, c]
    # every class Traction except ZLO-containing classes (Displacement)
    names = oracle.FACES3
    spec = {f: ({0: (0, 0.0), 1: (0, 0.0), 2: (0, 0.0)} if "ZLO" in f.split("_") else
                {0: (1, 0.0), 1: (1, 0.0), 2: (1, 0.0)}) for f in names}
    t, _ = oracle.bc_arrays(3, spec)
    b = np.random.default_rng(1).standard_normal(g.shape(3))
    out = oracle.newton_rhs(g, ISO, C, t, g.zeros(), b, F0)
    sig0 = np.einsum("ijkl,kl->ij", iso_tensor(lam, mu), F0v)
    assert np.array_equal(out[:, 1:-1, 1:-1, 1:-1], b[:, 1:-1, 1:-1, 1:-1])
    for k in range(g.nodes[2]):
        for j in range(g.nodes[1]):
            for i in range(g.nodes[0]):
                nrm = np.zeros(3)
                face = oracle.lib().or_classify(oracle.ctypes.byref(g.c()), i, j, k, nrm.ctypes.data_as(oracle.ctypes.c_void_p))
                if face < 0:
                    continue
                nm = names[face]
                if "ZLO" in nm.split("_"):
                    assert np.array_equal(out[:, k, j, i], b[:, k, j, i])
                else:
                    expect = b[:, k, j, i] + sig0 @ normal_of(nm)
                    assert np.allclose(out[:, k, j, i], expect, rtol=0, atol=1e-14)


def test_quadratic_displacement_strong_form(oracle):
    n = (6, 6, 6)
    g = oracle.Grid(n)
    rng = np.random.default_rng(2)
    R = H.random_rotation(rng)
    c21 = H.cubic_table_numpy(1.68, 1.21, 0.75, R)
    C = np.zeros(g.shape(21))
    C[:] = c21[:, None, None, None]
    Ct = H.full_tensor(3, MM, c21)
    X = g.coords()
    Q = rng.standard_normal((3, 3, 3))
    Q = 0.5 * (Q + Q.transpose(0, 2, 1))  # u_p = 0.5 x^T Q_p x
    u = np.stack([0.5 * sum(Q[p, a, c] * X[a] * X[c] for a in range(3) for c in range(3)) for p in range(3)])
    b = rng.standard_normal(g.shape(3))
    out = oracle.newton_rhs(g, MM, C, oracle.bc_arrays(3)[0], u, b)
    f = np.einsum("ijkl,klj->i", Ct, Q)  # C_ijkl u_k,lj with u_k,lj = Q_k,lj
    for p in range(3):
        assert np.abs(out[p, 2:-2, 2:-2, 2:-2] - (b[p, 2:-2, 2:-2, 2:-2] - f[p])).max() < 1e-11
    # next to a wall the one-sided wall gradient enters: the rows differ there (the reference's scheme)
    assert np.abs(out[:, 1, 2:-2, 2:-2] - (b[:, 1, 2:-2, 2:-2] - f[:, None, None])).max() > 1e-3


def test_linear_eigenstrain_divergence(oracle):
    n = (6, 5, 6)
    g = oracle.Grid(n)
    lam, mu = 1.3, 0.7
    C = np.zeros(g.shape(2))
    C[0], C[1] = lam, mu
    X = g.coords()
    G = np.random.default_rng(3).standard_normal((3, 3, 3)) * 0.01  # F0_kl = G_klj x_j
    F0 = np.stack([sum(G[a, c, j] * X[j] for j in range(3)) for a in range(3) for c in range(3)])
    b = np.zeros(g.shape(3))
    out = oracle.newton_rhs(g, ISO, C, oracle.bc_arrays(3)[0], g.zeros(), b, F0)
    div = np.einsum("ijkl,klj->i", iso_tensor(lam, mu), G)  # d/dx_j (C_ijkl F0_kl)
    for p in range(3):
        assert np.abs(out[p, 1:-1, 1:-1, 1:-1] - div[p]).max() < 1e-13


def test_smooth_fields_second_order(oracle):
    errs = []
    for m in (8, 16):
        g = oracle.Grid((m, m, m))
        X = g.coords()
        C = np.zeros(g.shape(2))
        C[0] = 1.0 + 0.3 * np.sin(X[0] + 2 * X[1])
        C[1] = 0.8 + 0.2 * np.cos(X[2])
        u = np.stack([np.sin(X[1]) * X[2], np.cos(X[0]) * 0.5, X[0] * X[1] * X[2]])
        b = np.zeros(g.shape(3))
        out = oracle.newton_rhs(g, ISO, C, oracle.bc_arrays(3)[0], u, b)
        # analytic div(C : grad u) by a fine finite difference of the analytic stress
        e = 1e-5

        def stress(x, y, z):
            lam = 1.0 + 0.3 * np.sin(x + 2 * y)
            mu = 0.8 + 0.2 * np.cos(z)
            gu = np.array([[0, np.cos(y) * z, np.sin(y)], [-0.5 * np.sin(x), 0, 0], [y * z, x * z, x * y]])
            return lam * np.trace(gu) * np.eye(3) + mu * (gu + gu.T)

        worst = 0.0
        idx = [(i, j, k) for i in range(2, m - 1) for j in range(2, m - 1) for k in range(2, m - 1)][::7]
        for (i, j, k) in idx:
            x, y, z = i / m, j / m, k / m
            d = np.zeros(3)
            for a, dx in enumerate(np.eye(3) * e):
                d += (stress(x + dx[0], y + dx[1], z + dx[2])[:, a] - stress(x - dx[0], y - dx[1], z - dx[2])[:, a]) / (2 * e)
            worst = max(worst, np.abs(out[:, k, j, i] + d).max())
        errs.append(worst)
    assert errs[0] / errs[1] > 3.0, errs
