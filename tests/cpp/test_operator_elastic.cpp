// C++ restatement of the reference's operator tests, driven through the
// C++ interface of include/alamo_elastic.hpp (which calls libela.so on the GPU):
//   UniaxialTest (src/Test/Operator/Elastic_UniaxialTest.cpp:13-206), component 0, 32^3, 1 level;
//   TrigTest     (src/Test/Operator/Elastic_TrigTest.cpp:17-140), Laplacian model, 32^3, 1 level.
// Prints one line per test and returns the number of failures, like src/test.cc.
#include <cmath>
#include <cstdio>
#include <vector>

#include "alamo_elastic.hpp"

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
    std::printf("%d tests failed\n", failed);
    return failed;
}
