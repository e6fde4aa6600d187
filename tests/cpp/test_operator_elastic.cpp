// C++ restatement of the reference's operator tests, driven through the
// C++ interface of include/alamo_elastic.hpp (which calls libela.so on the GPU):
//   UniaxialTest (src/Test/Operator/Elastic_UniaxialTest.cpp:13-206), component 0, 32^3, 1 level;
//   TrigTest     (src/Test/Operator/Elastic_TrigTest.cpp:17-140), Laplacian model, 32^3, 1 level;
//   and, as src/test.cc:128-150 runs them, UniaxialTest with 2 levels, 3 levels and "2 non-centered
//   levels" and TrigTest with 2 and 3 levels (the levels of Test/Operator/Elastic.cpp:56-89: each the
//   refinement of the middle half of the one below; Grid::YZ keeps the full x range);
//   and the affine Newton path (Newton.H:30-140 with Affine::Isotropic, Affine/Isotropic.H): a free
//   thermal expansion, uniform F0 = e0 I with traction-free faces and ZLO held at e0 (x, y, 0),
//   whose exact discrete solution is u = e0 x (the stress C:(grad u - F0) vanishes everywhere).
// Prints one line per test and returns the number of failures, like src/test.cc.
#include <cmath>
#include <cstdio>
#include <vector>

#include "alamo_elastic.hpp"

template <class M>
static void uniaxial_bcs(BC::Operator::Elastic<M>& bc) {
    using B = BC::Operator::Elastic<M>;
    bc.Set(B::XHI, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI, B::Y, B::Neumann, 0.0);
    bc.Set(B::YLO, B::X, B::Neumann, 0.0);
    bc.Set(B::YHI, B::X, B::Neumann, 0.0);
    bc.Set(B::XHI_YHI, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YLO, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI, B::Z, B::Neumann, 0.0);
    bc.Set(B::ZLO, B::X, B::Neumann, 0.0);
    bc.Set(B::ZHI, B::X, B::Neumann, 0.0);
    bc.Set(B::ZHI_XHI, B::X, B::Displacement, 0.1);
    bc.Set(B::ZLO_XHI, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YLO_ZLO, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YLO_ZHI, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YHI_ZLO, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YHI_ZHI, B::X, B::Displacement, 0.1);
    for (auto f : {B::YLO_ZLO, B::YLO_ZHI, B::YHI_ZLO, B::YHI_ZHI}) {
        bc.Set(f, B::X, B::Neumann, 0.0);
        bc.Set(f, B::Y, B::Neumann, 0.0);
    }
}

// The levels of the reference's test geometry: nlevels - 1 boxes, each the refinement of the middle
// half (or, for Grid::YZ, the middle half in y and z with the full x range) of the level below.
static std::vector<Box> test_boxes(int n, int nlevels, bool yz) {
    std::vector<Box> out;
    int lo[3] = {0, 0, 0}, hi[3] = {n, n, n};
    for (int l = 1; l < nlevels; ++l) {
        Box b;
        for (int d = 0; d < 3; ++d) {
            const int w = (hi[d] - lo[d]);
            const bool grow = !(yz && d == 0);
            const int clo = grow ? lo[d] + w / 4 : lo[d], chi = grow ? hi[d] - w / 4 : hi[d];
            b.lo[d] = 2 * clo;
            b.hi[d] = 2 * chi;
            lo[d] = b.lo[d];
            hi[d] = b.hi[d];
        }
        out.push_back(b);
    }
    return out;
}

// Global coordinates of node `id` of AMR level l's arrays.
template <class Op>
static void coords(Op& op, int l, int id, int n0, double x[3]) {
    const auto nn = op.levelNodes(l);
    const auto o = op.levelOrigin(l);
    const int i = id % nn[0], j = (id / nn[0]) % nn[1], k = id / (nn[0] * nn[1]);
    const double h = 1.0 / (n0 << l);
    x[0] = (o[0] + i) * h;
    x[1] = (o[1] + j) * h;
    x[2] = (o[2] + k) * h;
}

static int uniaxial_amr(int n, int nlevels, bool yz) {
    using model_type = Model::Solid::Linear::Isotropic;
    model_type model(2.6, 6.0);
    Geometry g;
    g.n = {n, n, n};
    Operator::Elastic<model_type> elastic(g, test_boxes(n, nlevels, yz));
    elastic.SetUniform(false);
    BC::Operator::Elastic<model_type> bc;
    uniaxial_bcs(bc);
    elastic.SetBC(&bc);
    std::vector<std::vector<double>> rhs(nlevels), u(nlevels);
    std::vector<std::vector<model_type>> modelfab(nlevels);
    for (int l = 0; l < nlevels; ++l) {
        const int nn = elastic.levelSize(l);
        modelfab[l].assign(nn, model);
        rhs[l].assign(3 * (size_t)nn, 0.0);
        u[l].assign(3 * (size_t)nn, 0.0);
        elastic.InitRHS(l, rhs[l]);
    }
    Solver::Nonlocal::Newton<model_type> mlmg(elastic);
    mlmg.solve(u, rhs, modelfab, 1e-10, 0.0);
    double err = 0;
    for (int l = 0; l < nlevels; ++l) {
        const int nn = elastic.levelSize(l);
        for (int id = 0; id < nn; ++id) {
            double x[3];
            coords(elastic, l, id, n, x);
            err = std::fmax(err, std::fabs(u[l][id] - 0.1 * x[0]));
            err = std::fmax(err, std::fabs(u[l][nn + id]));
            err = std::fmax(err, std::fabs(u[l][2 * (size_t)nn + id]));
        }
    }
    const double rel = err / 0.1;
    const int failed = rel > 1e-7;
    std::printf("UniaxialTest %d levels%s %d^3: relative error = %.3e %% %s\n", nlevels, yz ? " (non-centered)" : "",
                n, 100 * rel, failed ? "FAIL" : "PASS");
    return failed;
}

static int trig_amr(int n, int nlevels) {
    using model_type = Model::Solid::Linear::Laplacian;
    Geometry g;
    g.n = {n, n, n};
    Operator::Elastic<model_type> elastic(g, test_boxes(n, nlevels, false));
    elastic.SetUniform(false);
    BC::Operator::Elastic<model_type> bc;
    elastic.SetBC(&bc);
    const double pi = 3.14159265358979323846;
    std::vector<std::vector<double>> rhs(nlevels), u(nlevels);
    std::vector<std::vector<model_type>> modelfab(nlevels);
    for (int l = 0; l < nlevels; ++l) {
        const int nn = elastic.levelSize(l);
        modelfab[l].assign(nn, model_type(3));
        rhs[l].assign(3 * (size_t)nn, 0.0);
        u[l].assign(3 * (size_t)nn, 0.0);
        for (int id = 0; id < nn; ++id) {
            double x[3];
            coords(elastic, l, id, n, x);
            rhs[l][id] = std::sin(pi * x[0]) * std::sin(pi * x[1]) * std::sin(pi * x[2]);
        }
        elastic.InitRHS(l, rhs[l]);
    }
    Solver::Nonlocal::Newton<model_type> mlmg(elastic);
    mlmg.solve(u, rhs, modelfab, 1e-8, 0.0);
    double e = 0, m = 0;  // level 0, as the reference's TrigTest measures it
    const int nn = elastic.levelSize(0);
    for (int id = 0; id < nn; ++id) {
        double x[3];
        coords(elastic, 0, id, n, x);
        const double ex = -std::sin(pi * x[0]) * std::sin(pi * x[1]) * std::sin(pi * x[2]) / (3 * pi * pi);
        e = std::fmax(e, std::fabs(u[0][id] - ex));
        m = std::fmax(m, std::fabs(ex));
    }
    const int failed = e / m > 0.02;
    std::printf("TrigTest %d levels %d^3: relative error = %.3e %% %s\n", nlevels, n, 100 * e / m, failed ? "FAIL" : "PASS");
    return failed;
}

static int affine_free_expansion(int n) {
    using model_type = Model::Solid::Affine::Isotropic;
    const double e0 = 0.01;
    model_type model(2.6, 6.0, {e0, 0, 0, 0, e0, 0, 0, 0, e0});
    Geometry g;
    g.n = {n, n, n};
    Operator::Elastic<model_type> elastic(g);
    elastic.SetUniform(false);
    BC::Operator::Elastic<model_type> bc;
    using B = BC::Operator::Elastic<model_type>;
    // every class Traction 0, except the classes on ZLO: Displacement in all components
    for (int f = 0; f < 26; ++f) {
        const bool zlo = f == B::ZLO || f == B::YLO_ZLO || f == B::YHI_ZLO || f == B::ZLO_XLO || f == B::ZLO_XHI ||
                         f == B::XLO_YLO_ZLO || f == B::XLO_YHI_ZLO || f == B::XHI_YLO_ZLO || f == B::XHI_YHI_ZLO;
        for (auto d : {B::X, B::Y, B::Z}) bc.Set((typename B::Face)f, d, zlo ? B::Displacement : B::Traction, 0.0);
    }
    elastic.SetBC(&bc);
    const int nn = (n + 1) * (n + 1) * (n + 1);
    std::vector<model_type> modelfab(nn, model);
    std::vector<double> rhs(3 * (size_t)nn, 0.0), u(3 * (size_t)nn, 0.0);
    elastic.InitRHS(rhs);
    for (int j = 0; j <= n; ++j)  // the prescribed displacement on ZLO: e0 (x, y, 0)
        for (int i = 0; i <= n; ++i) {
            const size_t id = i + (size_t)(n + 1) * j;
            rhs[id] = e0 * i / n;
            rhs[(size_t)nn + id] = e0 * j / n;
        }
    Solver::Nonlocal::Newton<model_type> mlmg(elastic);
    mlmg.solve(u, rhs, modelfab, 1e-11, 0.0);
    double err = 0;
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                const size_t id = i + (size_t)(n + 1) * (j + (size_t)(n + 1) * k);
                const double ex[3] = {e0 * i / n, e0 * j / n, e0 * k / n};
                for (int p = 0; p < 3; ++p) err = std::fmax(err, std::fabs(u[p * (size_t)nn + id] - ex[p]));
            }
    const double rel = err / e0;
    const int failed = rel > 1e-7;
    std::printf("Affine free expansion %d^3: relative error = %.3e %% %s\n", n, 100 * rel, failed ? "FAIL" : "PASS");
    return failed;
}

// Affine::Cubic::Combine (Affine/Cubic.H:70-94): eta-weighted mean of the records and of F0.
static int combine_check() {
    using M = Model::Solid::Affine::Cubic;
    M a, b;
    a.Define(1.68, 1.21, 0.75, {1, 0, 0, 0, 1, 0, 0, 0, 1}, {0.01, 0, 0, 0, 0, 0, 0, 0, 0});
    const double c = std::cos(0.3), s = std::sin(0.3);
    b.Define(2.0, 1.0, 0.5, {c, -s, 0, s, c, 0, 0, 0, 1}, {0, 0, 0, 0, 0.02, 0, 0, 0, -0.01});
    const M m = M::Combine({a, b}, {1.0, 3.0});
    double err = 0;
    for (size_t q = 0; q < m.data.size(); ++q) err = std::fmax(err, std::fabs(m.data[q] - (0.25 * a.data[q] + 0.75 * b.data[q])));
    for (int q = 0; q < 9; ++q) err = std::fmax(err, std::fabs(m.F0[q] - (0.25 * a.F0[q] + 0.75 * b.F0[q])));
    const int failed = err > 1e-15;
    std::printf("Affine::Cubic::Combine: max deviation = %.3e %s\n", err, failed ? "FAIL" : "PASS");
    return failed;
}

// The rest of the operator interface (Elastic.H:48-136, Operator.H:79-100, Newton.H:211-333) on known
// answers.  Uniform isotropic (lambda, mu), default BCs (Displacement 0 on every class), u = G x.
static int interface_known_answers() {
    using M = Model::Solid::Linear::Isotropic;
    const int n = 8, nn = (n + 1) * (n + 1) * (n + 1);
    const double mu = 1.3, lambda = 2.1, h = 1.0 / n;
    const double G[3][3] = {{0.01, 0.002, -0.003}, {0.004, -0.02, 0.001}, {0.0, 0.005, 0.015}};
    Geometry g;
    g.n = {n, n, n};
    Operator::Elastic<M> op(g);
    BC::Operator::Elastic<M> bc;
    op.SetModel(M(mu, lambda));
    op.SetBC(&bc);
    op.SetUniform(false);
    op.prepareForSolve();
    int failed = 0;
    auto report = [&](const char* what, double err, double tol) {
        const int f = !(err <= tol);
        failed += f;
        std::printf("  %-44s max deviation = %.3e %s\n", what, err, f ? "FAIL" : "PASS");
    };
    auto id = [&](int i, int j, int k, int m) { return i + (size_t)(m + 1) * (j + (size_t)(m + 1) * k); };
    std::vector<double> u(3 * (size_t)nn);
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                const double x[3] = {i * h, j * h, k * h};
                for (int p = 0; p < 3; ++p) u[p * (size_t)nn + id(i, j, k, n)] = G[p][0] * x[0] + G[p][1] * x[1] + G[p][2] * x[2];
            }
    auto wall = [&](int i, int j, int k, int m) { return i == 0 || j == 0 || k == 0 || i == m || j == m || k == m; };
    std::printf("Operator interface known answers (%d^3, uniform isotropic, u = G x):\n", n);
    report("getNComp / isCrossStencil / getNGrow", std::fabs(op.getNComp() - 3.0) + op.isCrossStencil() +
                                                     std::fabs(op.getNGrow() - 2.0), 0.0);
    op.SetHomogeneous(true);
    report("SetHomogeneous stores the flag", op.GetHomogeneous() ? 0.0 : 1.0, 0.0);
    // Fapply: interior rows of an affine field vanish, Displacement rows return u
    std::vector<double> f;
    op.Fapply(0, 0, f, u);
    double e = 0;
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i)
                for (int p = 0; p < 3; ++p) {
                    const size_t q = p * (size_t)nn + id(i, j, k, n);
                    e = std::fmax(e, std::fabs(f[q] - (wall(i, j, k, n) ? u[q] : 0.0)));
                }
    report("Fapply(amrlev, mglev, out, in)", e, 1e-12);
    // residuals of the exact solution
    std::vector<double> r;
    const double rn = op.solutionResidual(0, r, u, f);
    e = rn;
    for (double v : r) e = std::fmax(e, std::fabs(v));
    report("solutionResidual of the exact solution", e, 1e-12);
    // Diagonal: 1 on Displacement rows, -2/h^2 (C_pppp + sum_{d != p} C_pdpd) inside
    std::vector<double> dg;
    op.Diagonal(0, 0, dg);
    e = 0;
    for (int p = 0; p < 3; ++p) {
        e = std::fmax(e, std::fabs(dg[p * (size_t)nn + id(0, 3, 5, n)] - 1.0));
        const double want = -2.0 / (h * h) * ((lambda + 2 * mu) + 2 * mu);
        e = std::fmax(e, std::fabs(dg[p * (size_t)nn + id(4, 4, 4, n)] - want) / std::fabs(want));
    }
    report("Diagonal(amrlev, mglev, diag)", e, 1e-14);
    // Fsmooth: the exact solution is a fixed point of the Jacobi sweeps
    std::vector<double> x = u;
    op.Fsmooth(0, 0, x, f);
    e = 0;
    for (size_t q = 0; q < x.size(); ++q) e = std::fmax(e, std::fabs(x[q] - u[q]));
    report("Fsmooth keeps the exact solution", e, 1e-15);
    // normalize: component 0 divided by diagonal component 0 once per component (3 times)
    std::vector<double> y = u;
    op.normalize(0, 0, y);
    e = 0;
    for (int q = 0; q < nn; ++q) {
        const double d0 = dg[q];
        e = std::fmax(e, std::fabs(y[q] - u[q] / d0 / d0 / d0) / (std::fabs(u[q] / d0 / d0 / d0) + 1e-300));
        e = std::fmax(e, std::fabs(y[nn + q] - u[nn + q]) + std::fabs(y[2 * nn + q] - u[2 * nn + q]));
    }
    report("normalize (the reference's component indexing)", e, 1e-15);
    // restriction / interpolation reproduce affine fields exactly
    const int nc = n / 2, nnc = (nc + 1) * (nc + 1) * (nc + 1);
    std::vector<double> crse;
    std::vector<double> fine = u;
    op.restriction(0, 1, crse, fine);
    e = 0;
    for (int k = 0; k <= nc; ++k)
        for (int j = 0; j <= nc; ++j)
            for (int i = 0; i <= nc; ++i)
                for (int p = 0; p < 3; ++p)
                    e = std::fmax(e, std::fabs(crse[p * (size_t)nnc + id(i, j, k, nc)] - u[p * (size_t)nn + id(2 * i, 2 * j, 2 * k, n)]));
    report("restriction of an affine field", e, 1e-15);
    std::vector<double> back(u.size(), 0.0);
    op.interpolation(0, 0, back, crse);
    e = 0;
    for (size_t q = 0; q < u.size(); ++q) e = std::fmax(e, std::fabs(back[q] - u[q]));
    report("interpolation of an affine field", e, 1e-15);
    // Strain / Stress / Energy
    double sig[3][3], eps[3][3], W = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            eps[a][b] = 0.5 * (G[a][b] + G[b][a]);
            sig[a][b] = mu * (G[a][b] + G[b][a]) + (a == b ? lambda * (G[0][0] + G[1][1] + G[2][2]) : 0.0);
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) W += 0.5 * G[a][b] * sig[a][b];
    std::vector<double> st, ss, en;
    op.Strain(0, st, u);
    op.Stress(0, ss, u);
    op.Energy(0, en, u);
    e = 0;
    for (int q = 0; q < nn; q += 37)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                e = std::fmax(e, std::fabs(st[(a * 3 + b) * (size_t)nn + q] - eps[a][b]));
                e = std::fmax(e, std::fabs(ss[(a * 3 + b) * (size_t)nn + q] - sig[a][b]) / 0.1);
            }
    for (int q = 0; q < nn; q += 37) e = std::fmax(e, std::fabs(en[q] - 2 * W) / (2 * W));
    report("Strain / Stress / Energy", e, 1e-12);
    // Newton W / DW / compResidual
    std::vector<M> model(nn, M(mu, lambda));
    Solver::Nonlocal::Newton<M> newton(op);
    std::vector<double> w, dw, res, zero(u.size(), 0.0);
    newton.W(w, u, model);
    newton.DW(dw, u, model);
    newton.compResidual(res, u, zero, model);
    e = 0;
    for (int q = 0; q < nn; ++q) {
        e = std::fmax(e, std::fabs(w[q] - W) / W);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) e = std::fmax(e, std::fabs(dw[(a * 3 + b) * (size_t)nn + q] - sig[a][b]) / 0.1);
    }
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i)
                for (int p = 0; p < 3; ++p) {
                    const size_t q = p * (size_t)nn + id(i, j, k, n);
                    e = std::fmax(e, std::fabs(res[q] - (wall(i, j, k, n) ? -u[q] : 0.0)));
                }
    report("Newton W / DW / compResidual", e, 1e-12);
    // reflux over two AMR levels: a residual of ones on the fine box restricts to 1 inside and to the
    // zero-ghost weights on the coarsened box's faces, edges and corners
    {
        const int n0 = 16;
        Geometry g2;
        g2.n = {n0, n0, n0};
        const std::vector<Box> fb = test_boxes(n0, 2, false);
        Operator::Elastic<M> amr(g2, fb);
        std::vector<double> cres(3 * (size_t)amr.levelSize(0), -7.0), fres(3 * (size_t)amr.levelSize(1), 1.0), none;
        amr.Reflux(0, cres, none, none, fres, none, none);
        const auto cn = amr.levelNodes(0);
        auto at = [&](int i, int j, int k) { return cres[i + (size_t)cn[0] * (j + (size_t)cn[1] * k)]; };
        const int lo = fb[0].lo[0] / 2, hi = fb[0].hi[0] / 2, mid = (lo + hi) / 2;
        e = std::fabs(at(mid, mid, mid) - 1.0) + std::fabs(at(lo, mid, mid) - 0.75) +
            std::fabs(at(lo, lo, mid) - 0.5625) + std::fabs(at(lo, lo, lo) - 27.0 / 64.0) +
            std::fabs(at(hi, hi, hi) - 27.0 / 64.0) + std::fabs(at(lo - 1, mid, mid) + 7.0);
        report("Reflux (partition of unity, zero ghosts)", e, 1e-15);
    }
    return failed;
}

static int uniaxial(int n, bool mixed) {
    using model_type = Model::Solid::Linear::Isotropic;
    model_type model(2.6, 6.0);  // (mu, lambda) as the reference passes (lame, shear)
    Geometry g;
    g.n = {n, n, n};
    Operator::Elastic<model_type> elastic(g);
    elastic.SetUniform(false);
    BC::Operator::Elastic<model_type> bc;
    using B = BC::Operator::Elastic<model_type>;
    bc.Set(B::XHI, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI, B::Y, B::Neumann, 0.0);
    bc.Set(B::YLO, B::X, B::Neumann, 0.0);
    bc.Set(B::YHI, B::X, B::Neumann, 0.0);
    bc.Set(B::XHI_YHI, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YLO, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI, B::Z, B::Neumann, 0.0);
    bc.Set(B::ZLO, B::X, B::Neumann, 0.0);
    bc.Set(B::ZHI, B::X, B::Neumann, 0.0);
    bc.Set(B::ZHI_XHI, B::X, B::Displacement, 0.1);
    bc.Set(B::ZLO_XHI, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YLO_ZLO, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YLO_ZHI, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YHI_ZLO, B::X, B::Displacement, 0.1);
    bc.Set(B::XHI_YHI_ZHI, B::X, B::Displacement, 0.1);
    for (auto f : {B::YLO_ZLO, B::YLO_ZHI, B::YHI_ZLO, B::YHI_ZHI}) {
        bc.Set(f, B::X, B::Neumann, 0.0);
        bc.Set(f, B::Y, B::Neumann, 0.0);
    }
    elastic.SetBC(&bc);
    const int nn = (n + 1) * (n + 1) * (n + 1);
    std::vector<model_type> modelfab(nn, model);
    std::vector<double> rhs(3 * (size_t)nn, 0.0), u(3 * (size_t)nn, 0.0);
    elastic.InitRHS(rhs);
    Solver::Nonlocal::Newton<model_type> mlmg(elastic);
    mlmg.setMixedPrecision(mixed);
    mlmg.solve(u, rhs, modelfab, 1e-10, 0.0);
    double norm = 0, total = 0;  // relative error as the reference computes it (norm0 sums)
    double err[3] = {0, 0, 0}, ex[3] = {0, 0, 0};
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                const size_t id = i + (size_t)(n + 1) * (j + (size_t)(n + 1) * k);
                const double exact[3] = {0.1 * i / n, 0.0, 0.0};
                for (int p = 0; p < 3; ++p) {
                    err[p] = std::fmax(err[p], std::fabs(u[p * (size_t)nn + id] - exact[p]));
                    ex[p] = std::fmax(ex[p], std::fabs(exact[p]));
                }
            }
    for (int p = 0; p < 3; ++p) { norm += err[p]; total += ex[p]; }
    const double rel = norm / total;
    const int failed = rel > 0.01 || rel > 1e-7;
    std::printf("UniaxialTest 1 level %d^3 %s: relative error = %.3e %% %s\n", n, mixed ? "mixed" : "fp64", 100 * rel,
                failed ? "FAIL" : "PASS");
    return failed;
}

static int trig(int n) {
    using model_type = Model::Solid::Linear::Laplacian;
    Geometry g;
    g.n = {n, n, n};
    Operator::Elastic<model_type> elastic(g);
    elastic.SetUniform(false);
    BC::Operator::Elastic<model_type> bc;  // dim == 3: displacement everywhere
    elastic.SetBC(&bc);
    const int nn = (n + 1) * (n + 1) * (n + 1);
    std::vector<model_type> modelfab(nn, model_type(3));
    std::vector<double> rhs(3 * (size_t)nn, 0.0), u(3 * (size_t)nn, 0.0), exact(nn);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                const size_t id = i + (size_t)(n + 1) * (j + (size_t)(n + 1) * k);
                const double t = std::sin(pi * i / n) * std::sin(pi * j / n) * std::sin(pi * k / n);
                rhs[id] = t;
                exact[id] = -t / (3 * pi * pi);
            }
    elastic.InitRHS(rhs);
    Solver::Nonlocal::Newton<model_type> mlmg(elastic);
    mlmg.solve(u, rhs, modelfab, 1e-8, 0.0);
    double e = 0, m = 0;
    for (int id = 0; id < nn; ++id) {
        e = std::fmax(e, std::fabs(u[id] - exact[id]));
        m = std::fmax(m, std::fabs(exact[id]));
    }
    const int failed = e / m > 0.02;
    std::printf("TrigTest 1 level %d^3: relative error = %.3e %% %s\n", n, 100 * e / m, failed ? "FAIL" : "PASS");
    return failed;
}

int main() {
    int failed = 0;
    failed += uniaxial(32, false);
    failed += uniaxial(32, true);
    failed += trig(32);
    failed += uniaxial_amr(32, 2, false);
    failed += uniaxial_amr(32, 3, false);
    failed += uniaxial_amr(32, 2, true);
    failed += trig_amr(32, 2);
    failed += trig_amr(32, 3);
    failed += affine_free_expansion(16);
    failed += combine_check();
    failed += interface_known_answers();
    std::printf("%d tests failed\n", failed);
    return failed;
}
