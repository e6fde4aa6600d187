"""Composite (AMR) restatement of the nodal elasticity problem over a nest of boxes.

Test infrastructure only (the checker of the composite solve): imported by tests/ and by
bench.py's checks, never by the product package.

The reference builds a composite solve from Operator<Node>'s AMR pieces and AMReX's MLMG:
  * reflux (src/Operator/Operator.cpp:728-839): every coarse node under the coarsened fine
    box takes the full-weighting restriction of the fine residual, with the corner / edge /
    face / interior branches chosen by the coarse *domain*; the fine residual is read on the
    fine box grown by one ghost layer (fine_res carries 2 ghost nodes);
  * averageDownSolutionRHS (:626-641): coarse nodes under the fine box take the coincident
    fine solution value (nodal amrex::average_down);
  * the fine level's ghost nodes hold the coarse solution interpolated trilinearly
    (interpolation, :593-614: coincident = injection, edge = mean of 2, face = mean of 4,
    centre = mean of 8), the "ghost nodes" coarse/fine strategy that
    Solver::Nonlocal::Linear selects (src/Solver/Nonlocal/Linear.H:23-29).

Composite problem stated here (and solved by the device's ela_amr_solve):
  level l >= 1 is one box B_l = [lo_l, hi_l] of level-l nodes (even corners), stored on the
  array box G_l = B_l grown by 2 nodes, clipped to the domain.  A node of B_l lying on a face
  of B_l that is not on the domain boundary is a *slave*; slaves and the nodes of G_l outside
  B_l hold the interpolation of level l-1.  Level l-1 nodes under B_l / 2 are *covered*;
  those on the boundary of B_l / 2 away from the domain boundary are *interface* nodes.
  Unknowns: every level-l node of B_l that is not a slave and not strictly covered.
  Equations: b - A_l u = 0 at the unknowns (A_l the single-box operator on G_l, with the
  domain's boundary rows on domain faces); at an interface node, the reflux of level l+1's
  residual = 0.  Strictly covered nodes take the coincident fine value.

Parity anchors: the reference's own multi-level known-answer tests (UniaxialTest with 2 and 3
levels and "2 non-centered levels", Elastic_UniaxialTest.cpp:13-206, src/test.cc:140-150;
TrigTest with 2 and 3 levels, Elastic_TrigTest.cpp:17-140, src/test.cc:128-134) and RefluxTest
(Elastic_RefluxTest.cpp:14-108).  The exact composite iterate of AMReX's MLMG is not pinned
(AMReX is not vendored).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O


def _face_sets(dim):
    names = O.FACES3 if dim == 3 else O.FACES2
    out = []
    for nm in names:
        s = set()
        for part in nm.split("_"):
            s.add(("XYZ".index(part[0]), part[1:] == "HI"))
        out.append(s)
    return out


class Nest:
    """Levels 0..L.  n0: cells of level 0 per direction; boxes: for levels 1..L, either one box (lo, hi)
    or a list of boxes [(lo, hi), ...] whose union is the level's refined region (node indices in the
    level's own index space, even corners; the boxes of a level may share faces); extent: domain
    lengths.  A level's arrays cover G_l, the bounding box of its union grown by 2 nodes, clipped to
    the domain."""

    def __init__(self, n0, boxes, extent=None):
        self.dim = len(n0)
        dim = self.dim
        self.nlev = len(boxes) + 1
        ext = tuple(extent) if extent is not None else (1.0,) * dim
        self.N = [tuple(int(c) << l for c in n0) for l in range(self.nlev)]
        self.dx = [tuple(ext[d] / self.N[l][d] for d in range(dim)) for l in range(self.nlev)]
        self.boxes = [[((0,) * dim, self.N[0])]]
        for b in boxes:
            lst = [b] if np.ndim(b[0]) == 1 else list(b)
            self.boxes.append([(tuple(int(v) for v in lo), tuple(int(v) for v in hi)) for lo, hi in lst])
        self.lo = [tuple(min(bx[0][d] for bx in self.boxes[l]) for d in range(dim)) for l in range(self.nlev)]
        self.hi = [tuple(max(bx[1][d] for bx in self.boxes[l]) for d in range(dim)) for l in range(self.nlev)]
        self.glo, self.ghi = [], []
        for l in range(self.nlev):
            if l == 0:
                self.glo.append((0,) * dim)
                self.ghi.append(self.N[0])
            else:
                self.glo.append(tuple(max(self.lo[l][d] - 2, 0) for d in range(dim)))
                self.ghi.append(tuple(min(self.hi[l][d] + 2, self.N[l][d]) for d in range(dim)))
        self.cells = [self._union_cells(l) for l in range(self.nlev)]
        for l in range(1, self.nlev):
            for lo, hi in self.boxes[l]:
                for d in range(dim):
                    assert lo[d] % 2 == 0 and hi[d] % 2 == 0 and 0 <= lo[d] < hi[d] <= self.N[l][d], \
                        "box corners must be even"
                # proper nesting: the coarsened box grown by 2 coarse cells (clipped to the domain) lies in
                # the coarser level's union
                rng = [(max(lo[d] // 2 - 2, 0), min(hi[d] // 2 + 2, self.N[l - 1][d])) for d in range(dim)]
                sl = tuple(reversed([slice(a - self.glo[l - 1][d], b - self.glo[l - 1][d])
                                     for d, (a, b) in enumerate(rng)]))
                assert self.cells[l - 1][sl].all(), "box must sit 2 coarse cells inside its parent"

    def _union_cells(self, l):
        """Cells of G_l covered by the level's boxes (cell c spans nodes c..c+1)."""
        shp = tuple(reversed([self.ghi[l][d] - self.glo[l][d] for d in range(self.dim)]))
        m = np.zeros(shp, bool)
        for lo, hi in self.boxes[l]:
            sl = tuple(reversed([slice(max(lo[d], self.glo[l][d]) - self.glo[l][d],
                                       min(hi[d], self.ghi[l][d]) - self.glo[l][d]) for d in range(self.dim)]))
            m[sl] = True
        return m

    def _touch(self, l, cells):
        """Per node of G_l: (some touching cell in `cells`, every touching in-domain cell in `cells`);
        cells outside G_l count as not in `cells`, cells outside the domain do not count."""
        dim = self.dim
        pad = np.pad(cells, 1, constant_values=False)
        dom = np.zeros(pad.shape, bool)
        sl = []
        for d in reversed(range(dim)):
            # padded index j is global cell glo + j - 1, in the domain when 0 <= glo + j - 1 < N
            a = max(0, 1 - self.glo[l][d])
            b = min(pad.shape[dim - 1 - d], self.N[l][d] - self.glo[l][d] + 1)
            sl.append(slice(a, b))
        dom[tuple(sl)] = True
        nshape = tuple(v + 1 for v in cells.shape)
        anyc = np.zeros(nshape, bool)
        allc = np.ones(nshape, bool)
        for off in np.ndindex(*(2,) * dim):
            w = tuple(slice(o, o + n) for o, n in zip(off, nshape))
            anyc |= pad[w]
            allc &= pad[w] | ~dom[w]
        return anyc, allc

    # -------------------------------------------------------------- geometry
    def grid(self, l):
        cells = tuple(self.ghi[l][d] - self.glo[l][d] for d in range(self.dim))
        return O.Grid(cells, tuple(cells[d] * self.dx[l][d] for d in range(self.dim)))

    def shape(self, l, nc):
        return (nc,) + tuple(reversed([self.ghi[l][d] - self.glo[l][d] + 1 for d in range(self.dim)]))

    def gidx(self, l):
        """Global node indices of G_l, one array per direction (shape of a plane of G_l)."""
        axes = [np.arange(self.glo[l][d], self.ghi[l][d] + 1) for d in range(self.dim)]
        mesh = np.meshgrid(*reversed(axes), indexing="ij")
        return list(reversed(mesh))

    def coords(self, l):
        return [g * self.dx[l][d] for d, g in enumerate(self.gidx(l))]

    def _under_cells(self, l):
        """Cells of G_l under level l+1's union (coarse cell c is under it when fine cell 2c is)."""
        fc = self.cells[l + 1]
        shp = self.cells[l].shape
        P = [np.arange(shp[self.dim - 1 - d]) + self.glo[l][d] for d in range(self.dim)]  # global coarse cells
        idx = []
        for d in range(self.dim):
            f = 2 * P[d] - self.glo[l + 1][d]
            idx.append(f)
        grids = np.meshgrid(*reversed(idx), indexing="ij")
        valid = np.ones(shp, bool)
        for gdim, gg in enumerate(grids):
            valid &= (gg >= 0) & (gg < fc.shape[gdim])
        clipped = [np.clip(gg, 0, fc.shape[gdim] - 1) for gdim, gg in enumerate(grids)]
        out = valid & fc[tuple(clipped)]
        return out

    def under(self, l):
        """Nodes of G_l under level l+1's union (touching at least one cell under it)."""
        return self._touch(l, self._under_cells(l))[0]

    def classify(self, l):
        """Boolean masks over G_l: inB (in the union), slave (on the union's boundary away from the
        domain boundary), covered (every touching in-domain cell under level l+1), interface (under
        level l+1 but not covered), unknown."""
        inB, allin = self._touch(l, self.cells[l])
        slave = inB & ~allin if l > 0 else np.zeros(inB.shape, bool)
        covered = np.zeros(inB.shape, bool)
        iface = np.zeros(inB.shape, bool)
        if l + 1 < self.nlev:
            anyu, allu = self._touch(l, self._under_cells(l))
            covered = allu & inB
            iface = anyu & ~allu
        unknown = inB & ~slave & ~covered
        return dict(inB=inB, slave=slave, covered=covered, iface=iface, unknown=unknown)

    def bct_array(self, l, bct):
        """BC type table of G_l: the domain's entry for classes whose faces are all domain
        faces of G_l, Displacement for classes touching an artificial face."""
        t = np.array(bct, dtype=np.int32).copy()
        for c, fs in enumerate(_face_sets(self.dim)):
            for (d, hi) in fs:
                art = (self.ghi[l][d] < self.N[l][d]) if hi else (self.glo[l][d] > 0)
                if art:
                    t[c, :] = O.DISP
                    break
        return t

    def bc_init(self, l, bct, bcv, rhs):
        """BC::Operator::Elastic::Init on G_l: every node on the domain boundary takes the
        prescribed value of its class (src/BC/Operator/Elastic.H:106-181)."""
        out = np.array(rhs, dtype=np.float64).copy()
        P = self.gidx(l)
        code = np.zeros(P[0].shape, dtype=np.int64)
        for d in range(self.dim):
            code += ((P[d] == 0) * 1 + (P[d] == self.N[l][d]) * 2) * 4 ** d
        table = -np.ones(4 ** self.dim, dtype=np.int64)
        for c, fs in enumerate(_face_sets(self.dim)):
            table[sum((2 if hi else 1) * 4 ** d for (d, hi) in fs)] = c
        cls = table[code]
        on = code > 0
        vals = np.asarray(bcv, dtype=np.float64)
        for p in range(self.dim):
            out[p][on] = vals[cls[on], p]
        return out

    # -------------------------------------------------------------- transfers
    def prolong(self, l, crse):
        """Trilinear interpolation of level l-1's array onto every node of G_l
        (Operator.cpp:593-614: mean of the 1, 2, 4 or 8 coincident/neighbouring coarse nodes)."""
        P = self.gidx(l)
        dim = self.dim
        q = [P[d] // 2 - self.glo[l - 1][d] for d in range(dim)]
        odd = [(P[d] % 2).astype(bool) for d in range(dim)]
        nc = crse.shape[0]
        out = np.zeros((nc,) + P[0].shape)
        cnt = np.zeros(P[0].shape)
        for delta in np.ndindex(*(2,) * dim):
            use = np.ones(P[0].shape, bool)
            idx = []
            for d in range(dim):
                if delta[d]:
                    use &= odd[d]
                idx.append(np.where(use, q[d] + delta[d], q[d]) if delta[d] else q[d])
            sl = tuple(reversed(idx))
            for c in range(nc):
                out[c] += np.where(use, crse[c][sl], 0.0)
            cnt += use
        return out / cnt

    def _under_nodes(self, l):
        """Global indices (one array per direction) of the level-(l-1) nodes under level l's union."""
        m = self.under(l - 1)
        nz = np.nonzero(m)
        return [nz[self.dim - 1 - d] + self.glo[l - 1][d] for d in range(self.dim)]

    def reflux(self, l, crse_res, fine_res):
        """Coarse nodes under level l's union take the restriction of the fine residual (read on G_l)."""
        dim = self.dim
        out = np.array(crse_res, dtype=np.float64).copy()
        C = self._under_nodes(l)
        acc = np.zeros((fine_res.shape[0],) + C[0].shape)
        for off in np.ndindex(*(3,) * dim):
            w = np.ones(C[0].shape)
            idx = []
            for d in range(dim):
                o = off[d] - 1
                bnd = (C[d] == 0) | (C[d] == self.N[l - 1][d])
                wd = np.where(bnd, 1.0 if o == 0 else 0.0, 0.5 if o == 0 else 0.25)
                w = w * wd
                f = 2 * C[d] + o - self.glo[l][d]
                f = np.clip(f, 0, self.ghi[l][d] - self.glo[l][d])
                idx.append(f)
            sl = tuple(reversed(idx))
            for c in range(fine_res.shape[0]):
                acc[c] += w * fine_res[c][sl]
        dst = tuple(reversed([C[d] - self.glo[l - 1][d] for d in range(dim)]))
        for c in range(out.shape[0]):
            out[c][dst] = acc[c]
        return out

    def inject(self, l, crse, fine):
        """averageDownSolutionRHS: coarse nodes under level l's union take the coincident fine value."""
        dim = self.dim
        out = np.array(crse, dtype=np.float64).copy()
        C = self._under_nodes(l)
        dst = tuple(reversed([C[d] - self.glo[l - 1][d] for d in range(dim)]))
        src = tuple(reversed([2 * C[d] - self.glo[l][d] for d in range(dim)]))
        for c in range(out.shape[0]):
            out[c][dst] = fine[c][src]
        return out

    # -------------------------------------------------------------- composite operator
    def fill(self, u):
        """Slaves and ghost nodes from the coarser level; covered nodes from the finer one."""
        u = [np.array(a, dtype=np.float64).copy() for a in u]
        for l in range(1, self.nlev):
            m = self.classify(l)
            keep = m["inB"] & ~m["slave"]
            p = self.prolong(l, u[l - 1])
            for c in range(u[l].shape[0]):
                u[l][c] = np.where(keep, u[l][c], p[c])
        for l in range(self.nlev - 1, 0, -1):
            u[l - 1] = self.inject(l, u[l - 1], u[l])
        return u

    def residual(self, kind, C, bct, b, u, uniform=False, masked=True):
        """Composite residual: per level b - A u on G_l, reflux of each finer level into the
        coarser one (finest first), zero off the unknowns (masked)."""
        u = self.fill(u)
        r = []
        for l in range(self.nlev):
            g = self.grid(l)
            r.append(b[l] - O.fapply(g, kind, C[l], self.bct_array(l, bct), u[l], uniform))
        for l in range(self.nlev - 1, 0, -1):
            r[l - 1] = self.reflux(l, r[l - 1], r[l])
        if masked:
            for l in range(self.nlev):
                m = self.classify(l)["unknown"]
                r[l] = r[l] * m
        return r, u

    # -------------------------------------------------------------- sparse form and direct solve
    def _level_matrix(self, l, kind, C, bct, uniform):
        """A_l on G_l as a sparse matrix, probed with a 3^dim colouring: every row of a column
        lies within one node of it, so each row of a probe belongs to the one probed node of
        that colour within one node of it."""
        g = self.grid(l)
        dim = self.dim
        shp = self.shape(l, dim)
        dims = list(reversed(shp[1:]))  # node counts per direction x, y, z
        nn = int(np.prod(dims))
        mesh = list(reversed(np.meshgrid(*[np.arange(n) for n in reversed(dims)], indexing="ij")))
        t = self.bct_array(l, bct)
        lin = np.arange(nn).reshape(shp[1:])
        rows, cols, vals = [], [], []
        for color in np.ndindex(*(3,) * dim):
            sel = np.ones(shp[1:], bool)
            own = []
            valid = np.ones(shp[1:], bool)
            for d in range(dim):
                m = mesh[d]
                sel &= (m % 3) == color[d]
                delta = (color[d] - m) % 3
                o = np.where(delta == 0, m, np.where(delta == 1, m + 1, m - 1))
                valid &= (o >= 0) & (o < dims[d])
                own.append(np.clip(o, 0, dims[d] - 1))
            ocl = np.zeros(shp[1:], dtype=np.int64)
            stride = 1
            for d in range(dim):
                ocl += own[d] * stride
                stride *= dims[d]
            for p in range(dim):
                e = np.zeros(shp)
                e[p][sel] = 1.0
                f = O.fapply(g, kind, C, t, e, uniform)
                for q in range(dim):
                    nz = (f[q] != 0) & valid
                    rows.append(q * nn + lin[nz])
                    cols.append(p * nn + ocl[nz])
                    vals.append(f[q][nz])
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(dim * nn, dim * nn))

    def solve_direct(self, kind, C, bct, b, uniform=False):
        """The composite solution by a sparse direct solve (scipy).  Returns per-level arrays
        on G_l with slaves, ghosts and covered nodes filled."""
        dim, L = self.dim, self.nlev
        masks = [self.classify(l) for l in range(L)]
        nn = [int(np.prod(self.shape(l, 1)[1:])) for l in range(L)]
        # unknown numbering
        base, ids = 0, []
        for l in range(L):
            m = masks[l]["unknown"].ravel()
            idx = -np.ones(dim * nn[l], dtype=np.int64)
            cnt = int(m.sum())
            for p in range(dim):
                idx[p * nn[l]:(p + 1) * nn[l]][m] = base + p * cnt + np.arange(cnt)
            ids.append(idx)
            base += dim * cnt
        nx = base
        # E_l: x -> full arrays on G_l
        E = [None] * L
        Pm = [None] * L
        for l in range(L):
            rows, cols, vals = [], [], []
            sel = ids[l] >= 0
            rows.append(np.nonzero(sel)[0])
            cols.append(ids[l][sel])
            vals.append(np.ones(int(sel.sum())))
            Ei = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(dim * nn[l], nx))
            if l > 0:
                m = masks[l]
                fillm = ~(m["inB"] & ~m["slave"])
                Pl = self._prolong_matrix(l)
                Pm[l] = Pl
                Pd = sp.kron(sp.identity(dim), Pl, format="csr")
                rsel = np.concatenate([fillm.ravel()] * dim).astype(float)
                Ei = Ei + sp.diags(rsel) @ Pd @ E[l - 1]
            E[l] = Ei.tocsr()
        for l in range(L - 2, -1, -1):
            Inj = self._inject_matrix(l + 1)
            Id = sp.kron(sp.identity(dim), Inj, format="csr")
            cov = np.concatenate([masks[l]["covered"].ravel()] * dim).astype(float)
            E[l] = (E[l] + sp.diags(cov) @ Id @ E[l + 1]).tocsr()
        A = [self._level_matrix(l, kind, C[l], bct, uniform) for l in range(L)]
        blocks, rhs = [], []
        for l in range(L):
            m = masks[l]
            rown = m["unknown"] & ~m["iface"]
            Sel = np.concatenate([rown.ravel()] * dim)
            blocks.append((A[l] @ E[l])[Sel])
            rhs.append(np.asarray(b[l]).ravel()[Sel])
            if l + 1 < L:
                Rf = sp.kron(sp.identity(dim), self._reflux_matrix(l + 1), format="csr")
                Sif = np.concatenate([m["iface"].ravel()] * dim)
                blocks.append((Rf @ A[l + 1] @ E[l + 1])[Sif])
                rhs.append((Rf @ np.asarray(b[l + 1]).ravel())[Sif])
        M = sp.vstack(blocks).tocsc()
        c = np.concatenate(rhs)
        # rows are in (level, [non-interface | interface]) order; columns in id order -- square
        x = spla.spsolve(M, c)
        u = []
        for l in range(L):
            u.append((E[l] @ x).reshape(self.shape(l, dim)))
        return u

    def _prolong_matrix(self, l):
        Pg = self.gidx(l)
        dim = self.dim
        nf = Pg[0].size
        cshape = self.shape(l - 1, 1)[1:]
        rows, cols, vals = [], [], []
        q = [(Pg[d] // 2 - self.glo[l - 1][d]).ravel() for d in range(dim)]
        odd = [(Pg[d] % 2).astype(bool).ravel() for d in range(dim)]
        nodd = sum(o.astype(int) for o in odd)
        w = 1.0 / (2.0 ** nodd)
        fr = np.arange(nf)
        for delta in np.ndindex(*(2,) * dim):
            use = np.ones(nf, bool)
            for d in range(dim):
                if delta[d]:
                    use &= odd[d]
            lin = np.zeros(nf, dtype=np.int64)
            stride = 1
            for d in range(dim):
                lin += (q[d] + delta[d] * use) * stride
                stride *= cshape[dim - 1 - d]
            rows.append(fr[use])
            cols.append(lin[use])
            vals.append(w[use])
        nc = int(np.prod(cshape))
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nf, nc))

    def _inject_matrix(self, l):
        """(coarse G_{l-1} nodes) x (fine G_l nodes): coincident values under level l's union."""
        dim = self.dim
        cshape = self.shape(l - 1, 1)[1:]
        fshape = self.shape(l, 1)[1:]
        C = self._under_nodes(l)
        rl = np.zeros(C[0].size, dtype=np.int64)
        fl = np.zeros(C[0].size, dtype=np.int64)
        rs = fs = 1
        for d in range(dim):
            rl += (C[d] - self.glo[l - 1][d]) * rs
            fl += (2 * C[d] - self.glo[l][d]) * fs
            rs *= cshape[dim - 1 - d]
            fs *= fshape[dim - 1 - d]
        return sp.csr_matrix((np.ones(rl.size), (rl, fl)), shape=(int(np.prod(cshape)), int(np.prod(fshape))))

    def _reflux_matrix(self, l):
        dim = self.dim
        cshape = self.shape(l - 1, 1)[1:]
        fshape = self.shape(l, 1)[1:]
        C = self._under_nodes(l)
        rl = np.zeros(C[0].size, dtype=np.int64)
        rs = 1
        for d in range(dim):
            rl += (C[d] - self.glo[l - 1][d]) * rs
            rs *= cshape[dim - 1 - d]
        rows, cols, vals = [], [], []
        for off in np.ndindex(*(3,) * dim):
            w = np.ones(C[0].size)
            fl = np.zeros(C[0].size, dtype=np.int64)
            fs = 1
            for d in range(dim):
                o = off[d] - 1
                bnd = (C[d] == 0) | (C[d] == self.N[l - 1][d])
                w = w * np.where(bnd, 1.0 if o == 0 else 0.0, 0.5 if o == 0 else 0.25)
                f = np.clip(2 * C[d] + o - self.glo[l][d], 0, fshape[dim - 1 - d] - 1)
                fl += f * fs
                fs *= fshape[dim - 1 - d]
            nz = w != 0
            rows.append(rl[nz])
            cols.append(fl[nz])
            vals.append(w[nz])
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(int(np.prod(cshape)), int(np.prod(fshape))))
