"""Seeded synthetic inputs for the nodal elasticity workloads.

Every generator here is this package's own and is deterministic given its
seed.  The reference builds its microstructures with ``rand()``
(src/IC/Voronoi.H:39-107, src/Set/Set.cpp:25-28), which is not reproducible
across C libraries, so these replace it.

Voronoi seeds are quantized to 1/1024 of a cell so that the nearest-seed
search is exact integer arithmetic (the GPU kernel and the checker agree bit
for bit on every node's grain, ties going to the lowest seed index).
"""
import numpy as np

SEED = 20261203
VORONOI_Q = 1024  # seed positions are integers in units of (cell size / VORONOI_Q)

DISP, TRAC, PERIODIC, NEUMANN = 0, 1, 2, 3

FACES3 = ["XLO", "YLO", "ZLO", "XHI", "YHI", "ZHI",
          "YLO_ZLO", "YLO_ZHI", "YHI_ZLO", "YHI_ZHI",
          "ZLO_XLO", "ZLO_XHI", "ZHI_XLO", "ZHI_XHI",
          "XLO_YLO", "XLO_YHI", "XHI_YLO", "XHI_YHI",
          "XLO_YLO_ZLO", "XLO_YLO_ZHI", "XLO_YHI_ZLO", "XLO_YHI_ZHI",
          "XHI_YLO_ZLO", "XHI_YLO_ZHI", "XHI_YHI_ZLO", "XHI_YHI_ZHI"]
FACES2 = ["XLO", "YLO", "XHI", "YHI", "XLO_YLO", "XLO_YHI", "XHI_YLO", "XHI_YHI"]

# Cubic constants used by the cubic workloads (the reference's defaults, Cubic.H:97).
C11, C12, C44 = 1.68, 1.21, 0.75


def faces(dim):
    return FACES3 if dim == 3 else FACES2


def bc_tables(dim, spec):
    """(type[nfaces,3] int32, value[nfaces,3] f64) from {face name: {comp: (type, value)}}.

    Every face class not named defaults to Displacement 0, as in the reference
    (BC/Operator/Elastic.H:46-55).
    """
    names = faces(dim)
    t = np.zeros((len(names), 3), dtype=np.int32)
    v = np.zeros((len(names), 3))
    for face, comps in spec.items():
        fi = names.index(face)
        for c, (ty, val) in comps.items():
            t[fi, c] = ty
            v[fi, c] = val
    return t, v


def lame(E, nu):
    """(lambda, mu) from (E, nu) in the reference's arithmetic order (Affine/Isotropic.H:67-74)."""
    lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)
    mu = E / 2.0 / (1.0 + nu)
    return lam, mu


def node_coords(n):
    """Nodal coordinates on [0,1]^dim for n cells per side; arrays of shape reversed(nodes)."""
    dim = len(n)
    axes = [np.arange(n[d] + 1) / n[d] for d in range(dim)]
    mesh = np.meshgrid(*reversed(axes), indexing="ij")
    return list(reversed(mesh))


# ------------------------------------------------------------------ 2D inclusion
def inclusion_2d(n=64):
    """2D, n x n cells on [0,1]^2, isotropic E=210 nu=0.3 with a centred r=0.2 inclusion of 10x stiffness.

    BCs: YLO displacement (0,0) (its corners too, the reference default);
    YHI displacement (0, 0.01) including the YHI corners; XLO/XHI traction 0.
    Returns (n, C[2, ny+1, nx+1] as (lambda, mu) planes, bc_type, bc_value).
    """
    lam, mu = lame(210.0, 0.3)
    x, y = node_coords((n, n))
    inc = (x - 0.5) ** 2 + (y - 0.5) ** 2 < 0.2 ** 2
    scale = np.where(inc, 10.0, 1.0)
    C = np.stack([lam * scale, mu * scale])
    spec = {
        "YHI": {0: (DISP, 0.0), 1: (DISP, 0.01)},
        "XLO_YHI": {0: (DISP, 0.0), 1: (DISP, 0.01)},
        "XHI_YHI": {0: (DISP, 0.0), 1: (DISP, 0.01)},
        "XLO": {0: (TRAC, 0.0), 1: (TRAC, 0.0)},
        "XHI": {0: (TRAC, 0.0), 1: (TRAC, 0.0)},
    }
    t, v = bc_tables(2, spec)
    return (n, n), C, t, v


# ------------------------------------------------------------------ grains
def haar_rotations(rng, count):
    """Haar-uniform rotation matrices from normalized Gaussian quaternions; shape (count, 3, 3)."""
    q = rng.standard_normal((count, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((count, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def voronoi_seeds(rng, count, n):
    """Uniform seeds in [0,1]^3 quantized to 1/VORONOI_Q of a cell: int64 array (count, 3)."""
    u = rng.random((count, 3))
    scale = np.array([n[0], n[1], n[2]], dtype=np.float64) * VORONOI_Q
    return np.rint(u * scale).astype(np.int64)


def voronoi_grain_ids(seeds_q, n, chunk=1 << 16):
    """Nearest-seed grain id per node by exact integer distance (ties -> lowest index).

    Returns int32 array of shape (nz+1, ny+1, nx+1).  Used by tests as the
    checker of the GPU grain assignment.
    """
    nx, ny, nz = (v + 1 for v in n)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.int64) * VORONOI_Q
    out = np.empty(P.shape[0], dtype=np.int32)
    S = seeds_q.astype(np.int64)
    for s in range(0, P.shape[0], chunk):
        d = ((P[s:s + chunk, None, :] - S[None, :, :]) ** 2).sum(axis=2)
        out[s:s + chunk] = np.argmin(d, axis=1)
    return out.reshape(nz, ny, nx)


def uniaxial_z_bcs(strain=0.01):
    """ZLO fixed (all comps), ZHI u_z = strain with x/y traction free, lateral faces traction free.

    Edges and corners follow the face whose z-condition applies.
    """
    zlo = {0: (DISP, 0.0), 1: (DISP, 0.0), 2: (DISP, 0.0)}
    zhi = {0: (TRAC, 0.0), 1: (TRAC, 0.0), 2: (DISP, strain)}
    free = {0: (TRAC, 0.0), 1: (TRAC, 0.0), 2: (TRAC, 0.0)}
    spec = {}
    for f in FACES3:
        parts = f.split("_")
        if "ZLO" in parts:
            spec[f] = zlo
        elif "ZHI" in parts:
            spec[f] = zhi
        else:
            spec[f] = free
    return bc_tables(3, spec)


def polycrystal(n=256, grains=64, seed=SEED):
    """3D, n^3 cells, cubic (1.68,1.21,0.75) Voronoi polycrystal with Haar orientations, uniaxial z strain 0.01.

    Returns a dict with the grid size, quantized seeds, rotations and BC tables.
    """
    rng = np.random.default_rng(seed)
    size = (n, n, n)
    seeds = voronoi_seeds(rng, grains, size)
    R = haar_rotations(rng, grains)
    t, v = uniaxial_z_bcs(0.01)
    return {"n": size, "seeds": seeds, "R": R, "C11": C11, "C12": C12, "C44": C44, "bc_type": t, "bc_value": v}


# ------------------------------------------------------------------ AMR crack (composite solve)
def chop_box(lo, hi, max_grid_size):
    """A box of nodes [lo, hi] cut into boxes of at most max_grid_size cells per direction, the pieces of a
    direction as equal as even corners allow (the role of AMReX's BoxArray::maxSize): e.g. 130 cells at
    32 give 5 pieces of 26."""
    cuts = []
    for a, b in zip(lo, hi):
        cells = b - a
        npiece = -(-cells // max_grid_size)
        pairs = cells // 2
        edges = [a + 2 * ((q * pairs) // npiece) for q in range(npiece)] + [b]
        cuts.append(edges)
    out = []
    for idx in np.ndindex(*[len(c) - 1 for c in cuts]):
        out.append((tuple(cuts[d][idx[d]] for d in range(len(lo))), tuple(cuts[d][idx[d] + 1] for d in range(len(lo)))))
    return out


def crack_nest_boxes(n=128, max_grid_size=None, front="physical"):
    """The refined regions of the mode-I crack on an n^3 base (ratio 2, node indices of each level).

    Level 1 covers |z - 0.5| <= 0.125, full x and y.  Level 2 covers |z - 0.5| <= 0.0625 and the crack
    front |x - 0.4| <= 0.125 (x rounded out to even nodes), full y.  With max_grid_size each region is a
    list of boxes of at most that many cells per direction (uneven where a region is not a multiple of
    it); without, one box per level.
    """
    n2 = 4 * n
    x0 = int(np.floor(0.275 * n2 / 2)) * 2
    x1 = int(np.ceil(0.525 * n2 / 2)) * 2
    if front == "blocks":  # rounded out to whole 32-cell blocks at n = 128: x in [0.25, 0.5625]
        x0, x1 = n, 9 * n // 4
    l1 = ((0, 0, 3 * n // 4), (2 * n, 2 * n, 5 * n // 4))
    l2 = ((x0, 0, 7 * n // 4), (x1, 4 * n, 9 * n // 4))
    if max_grid_size is None:
        return [l1, l2]
    return [chop_box(l1[0], l1[1], max_grid_size), chop_box(l2[0], l2[1], max_grid_size)]


def crack_eta(X, Y, Z, width=0.01, z0=0.5, x_tip=0.4):
    """eta = 0.5 (1 - tanh(d / width)), d the distance to the half-plane {z = z0, x < x_tip}."""
    dx = np.maximum(X - x_tip, 0.0)
    d = np.sqrt(dx * dx + (Z - z0) ** 2)
    return 0.5 * (1.0 - np.tanh(d / width))


def mode1_bcs(load=0.005):
    """ZHI u_z = +load, ZLO u_z = -load (ZLO also pins u_x = u_y = 0 against rigid motion), the
    lateral faces traction free; edges and corners follow the z face."""
    zlo = {0: (DISP, 0.0), 1: (DISP, 0.0), 2: (DISP, -load)}
    zhi = {0: (TRAC, 0.0), 1: (TRAC, 0.0), 2: (DISP, load)}
    free = {0: (TRAC, 0.0), 1: (TRAC, 0.0), 2: (TRAC, 0.0)}
    spec = {}
    for f in FACES3:
        parts = f.split("_")
        spec[f] = zlo if "ZLO" in parts else (zhi if "ZHI" in parts else free)
    return bc_tables(3, spec)


G_LITERAL = (1.0, 0.0, 1e-4)          # eta^2 + 1e-4: soft bulk, stiff crack plane (contrast ~2.5e3)
G_DAMAGE = (1.0, -2.0, 1.0 + 1e-4)    # (1 - eta)^2 + 1e-4: eta as damage (Degradable/Isotropic.H:78-79)


def crack_nest(n=128, g=G_DAMAGE, E=1.0, nu=0.3, width=0.01, load=0.005, max_grid_size=32, front="physical"):
    """3D AMR mode-I crack: base n^3 cells on [0,1]^3, 3 levels (ratio 2).

    Stiffness g(eta) C_iso(E, nu) per node with g(eta) = a eta^2 + b eta + c.  The default is the
    reference's convention, eta as damage: g = (1 - eta)^2 + 1e-4, the crack plane soft (g = 0.25 on
    it) in a stiff bulk, a contrast of 4.  The formula eta^2 + 1e-4 read literally (G_LITERAL) makes
    the bulk soft and the crack plane stiff, a contrast of about 2.5e3 across a layer thinner than a
    cell; with the strong-form operator's grad(C) terms the Jacobi-smoothed multigrid then diverges
    (already at a contrast of 26), in the oracle as on the device, so it is not a solvable workload.
    """
    lam, mu = lame(E, nu)
    t, v = mode1_bcs(load)
    return {"n": (n, n, n), "boxes": crack_nest_boxes(n, max_grid_size, front), "regions": crack_nest_boxes(n, None, front),
            "max_grid_size": max_grid_size, "lame": (lam, mu), "g": tuple(g), "width": width,
            "bc_type": t, "bc_value": v}


def crack_nest_model(cfg, X, Y, Z):
    """(lambda, mu) planes of the degraded isotropic stiffness at the nodes (X, Y, Z)."""
    a, b, c = cfg["g"]
    eta = crack_eta(X, Y, Z, cfg["width"])
    gg = a * eta * eta + b * eta + c
    lam, mu = cfg["lame"]
    return np.stack([lam * gg, mu * gg])


# ------------------------------------------------------------------ batched RVEs
RVE_SIDES = (32, 48, 64, 96)


def fixed_corner_gradient(strain=0.01):
    """Displacement gradient of the RVE loading u = strain * z * e_z (applied on every face)."""
    G = np.zeros((3, 3))
    G[2, 2] = strain
    return G


def rve_batch(count=512, seed=SEED, sides=RVE_SIDES, grains=(8, 32)):
    """Batched RVEs: side n drawn uniformly from `sides`, 8-32 Voronoi grains with
    Haar orientations, cubic (1.68, 1.21, 0.75), Displacement on every face with
    u = 0.01 z e_z there (all corners included).

    Returns a list of dicts with n, quantized seeds, rotations and G.
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sides))
        ng = int(rng.integers(grains[0], grains[1] + 1))
        seeds = voronoi_seeds(rng, ng, (n, n, n))
        R = haar_rotations(rng, ng)
        out.append({"n": n, "seeds": seeds, "R": R, "G": fixed_corner_gradient(0.01)})
    return out


# ------------------------------------------------------------------ serving stream
def dense_coarsest_unknowns(n):
    """Unknowns of the coarsest level of a member of side n (n halved while even and n / 2 >= 2):
    3 (c+1)^3; above 1536 the batch solves that level by BiCGStab instead of a dense inverse."""
    while n % 2 == 0 and n // 2 >= 2:
        n //= 2
    return 3 * (n + 1) ** 3


def random_crack(rng, n):
    """A random half-plane crack: point p in [0.25, 0.75]^3, unit normal, in-plane direction tv
    (the crack covers tv.(x - p) <= 0), width max(0.01, 2 h), degradation g = (1 - eta)^2 + 1e-4."""
    p = rng.random(3) * 0.5 + 0.25
    nrm = rng.standard_normal(3)
    nrm /= np.linalg.norm(nrm)
    tv = np.cross(nrm, rng.standard_normal(3))
    tv /= np.linalg.norm(tv)
    return (p, nrm, tv, max(0.01, 2.0 / n), (1.0, -2.0, 1.0 + 1e-4))


def request_stream(rate=200.0, duration=60.0, seed=SEED, nmin=32, nmax=192, grains=(8, 64), p_crack=0.5, strain=0.01,
                   blocking=4):
    """Serving stream: Poisson arrivals (rate per s) over `duration` s; side log-uniform in
    [nmin, nmax] (kept as n_requested), solved at the nearest multiple of `blocking` cells -- AMReX
    requires the domain to be a multiple of amr.blocking_factor, which the reference's elastic inputs
    set to 4 (tests/UniaxialTension/input:13, tests/Eshelby/input:13); blocking=1 solves every request at
    its integer side; 8-64 Voronoi grains (Haar orientations, cubic constants above); a crack with
    probability p_crack; load u = strain (a.x) a on every face for a random unit vector a.  Sorted by
    arrival."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        n_req = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        n = max(blocking, blocking * int(np.rint(n_req / blocking)))
        ng = int(rng.integers(grains[0], grains[1] + 1))
        seeds = voronoi_seeds(rng, ng, (n, n, n))
        R = haar_rotations(rng, ng)
        a = rng.standard_normal(3)
        a /= np.linalg.norm(a)
        crack = random_crack(rng, n) if rng.random() < p_crack else None
        out.append({"arrival": t, "n": n, "n_requested": n_req, "seeds": seeds, "R": R,
                    "G": strain * np.outer(a, a), "crack": crack})
    return out
