"""Pins of the oracle's model-field assembly and of Newton's W / DW -- no GPU.

The functions under test restate DegradeMaterial's node average
(PolymerDegradation.cpp:817-828), PhaseFieldMicrostructure's x-y average with
Affine::Cubic::Combine (PFM.cpp:529-543, Affine/Cubic.H:70-94) and
Newton<T>::W / DW (Newton.H:233-333).  Each check is a closed form or an
invariant that a dropped cell, a shifted ghost index, a missing 1/2 or a
transposed gradient would fail.
"""
import numpy as np
import pytest

from tests import helpers as H

ISO, MM = H.ISO, H.MM


def cell_centres(g):
    """Cell-centre coordinates of cells -1..n (one ghost layer), each of shape reversed(n + 2)."""
    axes = [(np.arange(g.n[d] + 2) - 0.5) * g.dx[d] for d in range(g.dim)]
    mesh = np.meshgrid(*reversed(axes), indexing="ij")
    return list(reversed(mesh))


@pytest.mark.parametrize("n", [(6, 5), (4, 5, 3)])
def test_degrade_material_average_reproduces_affine_fields(oracle, n):
    """The mean of the 2^dim cells around a node is exact for an affine field: the node value equals
    the field at the node.  A missing cell, a wrong weight (1/6 for 1/8) or a ghost index shifted by one
    breaks it."""
    g = oracle.Grid(n)
    xc = cell_centres(g)
    a = np.array([0.3, -0.7, 0.45][: g.dim])
    eta = (0.2 + sum(a[d] * xc[d] for d in range(g.dim)))[None]
    got = oracle.degrade_material_nodes(g, eta)[0]
    xn = g.coords()
    want = 0.2 + sum(a[d] * xn[d] for d in range(g.dim))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


def test_degrade_material_sees_every_cell_once(oracle):
    """A single nonzero cell contributes 1/2^dim to exactly its 2^dim corner nodes."""
    g = oracle.Grid((4, 3, 5))
    eta = np.zeros((1, 7, 5, 6))
    eta[0, 3, 2, 4] = 1.0  # cell (i, j, k) = (3, 1, 2)
    got = oracle.degrade_material_nodes(g, eta)[0]
    assert np.count_nonzero(got) == 8
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                assert got[2 + dk, 1 + dj, 3 + di] == 0.125


def test_pfm_average_is_the_xy_plane_of_layer_k(oracle):
    """PFM.cpp:537 averages the 4 cells of the x-y plane at cell layer k, also in 3D: a field that varies
    only in z takes the value of cell layer k at node k (not the mean of layers k-1 and k), while the
    x-y average of an affine x-y field is exact."""
    g = oracle.Grid((4, 4, 3))
    xc = cell_centres(g)
    eta = (1.0 + 10.0 * xc[2])[None]
    got = oracle.pfm_eta_nodes(g, eta)[0]
    for k in range(g.n[2] + 1):
        np.testing.assert_allclose(got[k], 1.0 + 10.0 * (k + 0.5) * g.dx[2], rtol=0, atol=1e-14)
    eta2 = (0.5 + 0.3 * xc[0] - 0.2 * xc[1])[None]
    xn = g.coords()
    np.testing.assert_allclose(oracle.pfm_eta_nodes(g, eta2)[0], 0.5 + 0.3 * xn[0] - 0.2 * xn[1], atol=1e-15)


def test_combine_weights(oracle):
    """One phase present -> its own record and F0 exactly; equal weights -> the plain mean; the weights
    are normalized (eta scaled by any factor gives the same record)."""
    rng = np.random.default_rng(1)
    ddw = rng.uniform(0.5, 2.0, size=(3, 21))
    F0 = rng.normal(size=(3, 3, 3))
    etas = np.zeros((3, 2, 2, 2))
    etas[1] = 0.7
    C, F = oracle.combine(ddw, F0, etas)
    np.testing.assert_allclose(C[:, 0, 0, 0], ddw[1], rtol=1e-15)
    np.testing.assert_allclose(F[:, 1, 1, 1], F0[1].reshape(-1), rtol=1e-15)
    etas[:] = 0.25
    C, F = oracle.combine(ddw, F0, etas)
    np.testing.assert_allclose(C[:, 0, 0, 0], ddw.mean(axis=0), rtol=1e-14)
    C2, _ = oracle.combine(ddw, F0, 3.0 * etas)
    np.testing.assert_allclose(C2, C, rtol=1e-14)
    C0, _ = oracle.combine(ddw, None, np.zeros((3, 1, 1, 1)))
    assert np.isnan(C0).all()


@pytest.mark.parametrize("dim,kind", [(2, ISO), (3, ISO), (3, MM)])
def test_w_dw_of_affine_displacement(oracle, dim, kind):
    """u = G x with uniform C: every node's gradient is G (one-sided and central differences are exact),
    so DW = C : (G - F0) and W = wfac (G - F0) : C : (G - F0) at every node, walls included.  The
    stress is the textbook contraction of the full 81-entry tensor."""
    rng = np.random.default_rng(dim + 10 * kind)
    n = (5, 4, 3)[:dim]
    g = oracle.Grid(n)
    nc = H.ncomp(dim, kind)
    c = H.random_stiffness(rng, (1,) * dim, kind, dim).reshape(nc, -1)[:, 0]
    C = np.broadcast_to(c.reshape((nc,) + (1,) * dim), g.shape(nc)).copy()
    G = rng.normal(size=(dim, dim))
    F0 = rng.normal(size=(dim, dim))
    x = g.coords()
    u = np.stack([sum(G[p, d] * x[d] for d in range(dim)) for p in range(dim)])
    Ct = H.full_tensor(dim, kind, c)
    for f0, wfac in ((None, 0.5), (F0, 1.0)):
        E = G - (0 if f0 is None else f0)
        sig = np.einsum("ijkl,kl->ij", Ct, E)
        F0p = None if f0 is None else np.broadcast_to(f0.reshape((dim * dim,) + (1,) * dim), g.shape(dim * dim)).copy()
        w, dw = oracle.newton_w_dw(g, kind, C, u, F0p, wfac)
        for a in range(dim):
            for b in range(dim):
                np.testing.assert_allclose(dw[a * dim + b], sig[a, b], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(w, wfac * np.sum(E * sig), rtol=1e-12)


def test_w_is_half_the_reference_energy_for_linear_models(oracle):
    """Elastic::Energy (Elastic.cpp:502-549) is W(grad u) + sigma:eps/2; for the symmetric stress of these
    models sigma:eps = grad u : sigma = 2 W, so Energy = 2 W (wfac 1/2) node by node, on a random field
    with a random per-node stiffness (the oracle's Energy is a separate loop)."""
    rng = np.random.default_rng(5)
    n = (6, 5, 4)
    g = oracle.Grid(n)
    C = H.random_stiffness(rng, n, MM, 3)
    u = H.random_field(rng, n, 3)
    w, dw = oracle.newton_w_dw(g, MM, C, u)
    e = oracle.post(g, MM, C, u, 2)
    np.testing.assert_allclose(e, 2.0 * w, rtol=1e-12, atol=1e-12 * np.abs(e).max())
    np.testing.assert_allclose(dw, oracle.post(g, MM, C, u, 1), rtol=1e-13, atol=1e-13 * np.abs(dw).max())
