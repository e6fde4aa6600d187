// alamo_elastic.hpp -- C++ host interface of the nodal elasticity solve,
// shaped like the reference's classes so a caller of
//   Operator::Elastic<T>            (src/Operator/Elastic.H:20-171)
//   BC::Operator::Elastic<T>        (src/BC/Operator/Elastic.H:12-407)
//   Model::Solid::Linear::{Isotropic,Cubic,Laplacian}
//   Model::Solid::Affine::{Isotropic,Cubic}
//   Model::Solid::LinearElastic::Degradable::Isotropic
//   Solver::Nonlocal::{Linear, Newton<T>} (src/Solver/Nonlocal/*.H)
// can switch to the gfx950 implementation.  Header-only; every computation
// goes through the C ABI of ela.h (libela.so).  AMReX containers are replaced
// by project types: a Geometry (cells and cell size of one box) and dense host
// fields (std::vector<double>, component-major, x fastest).  Where the
// reference aborts (Util::Abort), these classes throw std::runtime_error.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ela.h"

namespace Set {
using Scalar = double;
}

namespace ela_detail {
inline void check(int rc, const char* what) {
    if (rc != ELA_OK) throw std::runtime_error(std::string(what) + ": " + ela_last_error());
}
inline ela_ctx* context() {
    static ela_ctx* ctx = nullptr;
    if (!ctx) check(ela_ctx_create(0, 0, &ctx), "ela_ctx_create");
    return ctx;
}
}  // namespace ela_detail

// ---------------------------------------------------------------- models
namespace Model {
namespace Solid {
namespace Linear {

// Model::Solid::Linear::Isotropic (src/Model/Solid/Linear/Isotropic.H:13-58).
// The constructor takes (mu, lambda) as the reference's does; storage is (lambda, mu).
class Isotropic {
public:
    static constexpr int kind = ELA_MODEL_ISO;
    Isotropic() {}
    Isotropic(Set::Scalar mu, Set::Scalar lambda) { Define(mu, lambda); }
    void Define(Set::Scalar mu, Set::Scalar lambda) { data = {lambda, mu}; }
    std::vector<double> data{NAN, NAN};
};

// Model::Solid::Linear::Cubic (src/Model/Solid/Linear/Cubic.H:13-114): 21 values,
// rotated from (C11, C12, C44) exactly as the reference's Define does (its
// unrotated tensor sets C44 only where i==k && j==l, and each storage slot
// keeps the last (p,q,s,t) of its symmetry class).
class Cubic {
public:
    static constexpr int kind = ELA_MODEL_MM;
    Cubic() : data(21, NAN) {}
    void Define(Set::Scalar C11, Set::Scalar C12, Set::Scalar C44, Set::Scalar phi1, Set::Scalar Phi,
                Set::Scalar phi2) {
        auto rx = [](double a) { return std::array<double, 9>{1, 0, 0, 0, std::cos(a), -std::sin(a), 0, std::sin(a), std::cos(a)}; };
        auto rz = [](double a) { return std::array<double, 9>{std::cos(a), -std::sin(a), 0, std::sin(a), std::cos(a), 0, 0, 0, 1}; };
        auto mul = [](const std::array<double, 9>& a, const std::array<double, 9>& b) {
            std::array<double, 9> o{};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) o[3 * i + j] += a[3 * i + k] * b[3 * k + j];
            return o;
        };
        // AngleAxis(phi2, X) * AngleAxis(Phi, Z) * AngleAxis(phi1, X)  (Cubic.H:24-27)
        Define(C11, C12, C44, mul(mul(rx(phi2), rz(Phi)), rx(phi1)));
    }
    void Define(Set::Scalar C11, Set::Scalar C12, Set::Scalar C44, const std::array<double, 9>& R) {
        double Ct[3][3][3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l)
                        Ct[i][j][k][l] = (i == j && j == k && k == l) ? C11
                                         : (i == k && j == l)       ? C44
                                         : (i == j && k == l)       ? C12
                                                                    : 0.0;
        auto voigt = [](int i, int j) {
            static const int t[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
            return t[i][j];
        };
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q)
                for (int s = 0; s < 3; ++s)
                    for (int t = 0; t < 3; ++t) {
                        double acc = 0;
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j)
                                for (int k = 0; k < 3; ++k)
                                    for (int l = 0; l < 3; ++l)
                                        acc += R[3 * p + i] * R[3 * s + k] * Ct[i][j][k][l] * R[3 * q + j] * R[3 * t + l];
                        int I = voigt(p, q), J = voigt(s, t);
                        if (I > J) std::swap(I, J);
                        data[I * 6 - I * (I - 1) / 2 + (J - I)] = acc;
                    }
    }
    std::vector<double> data;
};

// Model::Solid::Linear::Laplacian (src/Model/Solid/Linear/Laplacian.H:25-57): Diagonal, A = ones.
class Laplacian {
public:
    static constexpr int kind = ELA_MODEL_DIAG;
    explicit Laplacian(int dim = 3) : data(dim * dim, 1.0) {}
    std::vector<double> data;
};

}  // namespace Linear

namespace Affine {
// Model::Solid::Affine::Isotropic (src/Model/Solid/Affine/Isotropic.H:12-84): the linear stiffness
// plus an eigenstrain F0 (row-major 3x3), DW(gradu) = DDW : (gradu - F0).
class Isotropic : public Linear::Isotropic {
public:
    static constexpr bool affine = true;
    static constexpr double w_factor = 1.0;  // W = (F-F0):C:(F-F0), no 1/2 (Affine/Isotropic.H:31-34)
    Isotropic() {}
    Isotropic(Set::Scalar mu, Set::Scalar lambda, const std::array<double, 9>& a_F0 = {})
        : Linear::Isotropic(mu, lambda), F0(a_F0) {}
    std::array<double, 9> F0{};
};
// Model::Solid::Affine::Cubic (src/Model/Solid/Affine/Cubic.H:13-97).
class Cubic : public Linear::Cubic {
public:
    static constexpr bool affine = true;
    Cubic() {}
    void Define(Set::Scalar C11, Set::Scalar C12, Set::Scalar C44, const std::array<double, 9>& R,
                const std::array<double, 9>& a_F0 = {}) {
        Linear::Cubic::Define(C11, C12, C44, R);
        F0 = a_F0;
    }
    // Combine (Affine/Cubic.H:70-94): the eta-weighted mean of the stiffness records and of F0,
    // weights eta[n] / sum(eta), accumulated in model order.
    static Cubic Combine(const std::vector<Cubic>& models, const std::vector<Set::Scalar>& eta) {
        Cubic ret;
        std::fill(ret.data.begin(), ret.data.end(), 0.0);
        ret.F0.fill(0.0);
        Set::Scalar etasum = 0.;
        for (size_t n = 0; n < models.size(); ++n) etasum += eta[n];
        for (size_t n = 0; n < models.size(); ++n) {
            const Set::Scalar w = eta[n] / etasum;
            for (size_t q = 0; q < ret.data.size(); ++q) ret.data[q] += models[n].data[q] * w;
            for (int q = 0; q < 9; ++q) ret.F0[q] += models[n].F0[q] * w;
        }
        return ret;
    }
    std::array<double, 9> F0{};
};
}  // namespace Affine

namespace LinearElastic {
namespace Degradable {
// Degradable::Isotropic (src/Model/Solid/LinearElastic/Degradable/Isotropic.H:21-117).
class Isotropic {
public:
    static constexpr int kind = ELA_MODEL_ISO;
    Isotropic() {}
    Isotropic(Set::Scalar mu, Set::Scalar lambda) : mu0(mu), lambda0(lambda), data{lambda, mu} {}
    void DegradeModulus(Set::Scalar eta) {  // Isotropic.H:74-81
        if (eta < 0.0 || eta > 1.0) throw std::runtime_error("Invalid value of eta");
        data = {lambda0 * (1.0 - eta), mu0 * (1.0 - eta)};
    }
    // Isotropic.H:83-104: one eta degrades both moduli, two degrade lambda and mu separately;
    // any other count leaves the moduli unchanged.  The checks are the reference's, in its order.
    void DegradeModulus(const std::vector<Set::Scalar>& eta) {
        if (eta.size() == 1) {
            if (std::isnan(eta[0])) throw std::runtime_error("Nans in eta");
            if (std::isinf(eta[0])) throw std::runtime_error("Infs in eta");
            if (eta[0] < 0.0 || eta[0] > 1.0) throw std::runtime_error("Invalid value of eta");
            data = {lambda0 * (1.0 - eta[0]), mu0 * (1.0 - eta[0])};
        } else if (eta.size() == 2) {
            if (std::isnan(eta[0]) || std::isnan(eta[1])) throw std::runtime_error("Nans in eta");
            if (std::isinf(eta[0])) throw std::runtime_error("Infs in eta");
            if (eta[0] < 0.0 || eta[0] > 1.0 || eta[1] < 0.0 || eta[1] > 1.0)
                throw std::runtime_error("Invalid value of eta");
            data = {lambda0 * (1.0 - eta[0]), mu0 * (1.0 - eta[1])};
        }
    }
    Set::Scalar mu0 = NAN, lambda0 = NAN;
    std::vector<double> data{NAN, NAN};
};
}  // namespace Degradable
}  // namespace LinearElastic
}  // namespace Solid
}  // namespace Model

// ---------------------------------------------------------------- boundary conditions
namespace BC {
namespace Operator {
// BC::Operator::Elastic<T> (src/BC/Operator/Elastic.H): face classes and
// component types; every class defaults to Displacement 0 (:46-55).
template <class T>
class Elastic {
public:
    enum Type { Displacement = ELA_BC_DISP, Traction = ELA_BC_TRAC, Periodic = ELA_BC_PERIODIC, Neumann = ELA_BC_NEUMANN };
    enum Face {
        XLO, YLO, ZLO, XHI, YHI, ZHI,
        YLO_ZLO, YLO_ZHI, YHI_ZLO, YHI_ZHI,
        ZLO_XLO, ZLO_XHI, ZHI_XLO, ZHI_XHI,
        XLO_YLO, XLO_YHI, XHI_YLO, XHI_YHI,
        XLO_YLO_ZLO, XLO_YLO_ZHI, XLO_YHI_ZLO, XLO_YHI_ZHI,
        XHI_YLO_ZLO, XHI_YLO_ZHI, XHI_YHI_ZLO, XHI_YHI_ZHI,
        INT
    };
    enum Direction { X = 0, Y = 1, Z = 2 };
    Elastic() : types(26 * 3, Displacement), values(26 * 3, 0.0) {}
    void Set(Face face, Direction dir, Type type, Set::Scalar value) {
        types[3 * face + dir] = type;
        values[3 * face + dir] = value;
    }
    std::vector<int> types;
    std::vector<double> values;
};
}  // namespace Operator
}  // namespace BC

// ---------------------------------------------------------------- operator
struct Geometry {
    int dim = 3;
    std::array<int, 3> n{0, 0, 0};       // cells
    std::array<double, 3> L{1, 1, 1};    // domain lengths
};

// The refined region of an AMR level (ratio 2): node indices of the level,
// even corners (the reference's BoxArray of that level, as one box).
struct Box {
    std::array<int, 3> lo{0, 0, 0}, hi{0, 0, 0};
};

namespace ela_detail {
// A device copy of a dense host array, freed with the object.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    explicit DevBuf(size_t n) : bytes(n) { check(ela_dev_alloc(context(), n, &p), "ela_dev_alloc"); }
    DevBuf(const std::vector<double>& host) : DevBuf(host.size() * sizeof(double)) {
        check(ela_memcpy(p, host.data(), bytes), "ela_memcpy");
    }
    void to(std::vector<double>& host) const { check(ela_memcpy(host.data(), p, bytes), "ela_memcpy"); }
    ~DevBuf() {
        if (p) ela_dev_free(context(), p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};
inline ela_level_layout dense_layout(int dim, const std::array<int, 3>& nodes) {
    ela_level_layout l{};
    l.dim = dim;
    for (int d = 0; d < 3; ++d) {
        l.nodes[d] = d < dim ? nodes[d] : 1;
        l.n[d] = l.nodes[d] - 1;
    }
    l.pitch = l.nodes[0];
    l.plane = (int64_t)l.nodes[0] * l.nodes[1];
    l.cstride = l.plane * l.nodes[2];
    return l;
}
}  // namespace ela_detail

namespace Operator {
// Operator::Elastic<T> (src/Operator/Elastic.H:20-171) with the virtuals of
// Operator<Grid::Node> (src/Operator/Operator.H:79-100).  define() builds the
// MG hierarchy; SetModel/SetBC/SetUniform/SetHomogeneous as in the reference.
// Fields are dense host vectors of the MG level they belong to (dim
// components, x fastest); the per-level operations (Fapply, Diagonal,
// Fsmooth, normalize, restriction, interpolation, the residuals, Strain /
// Stress / Energy) run on one AMR level's hierarchy (amrlev 0 of a single-box
// operator); reflux runs between two levels of a several-level operator.
template <class T>
class Elastic {
public:
    Elastic() {}
    Elastic(const Geometry& g, int mode = ELA_ALLOC_BOTH) { define(g, mode); }
    // several AMR levels: level 0 on g, level l >= 1 on fine[l - 1]
    Elastic(const Geometry& g, const std::vector<Box>& fine, int mode = ELA_SOLVE_MIXED) { define(g, fine, mode); }
    ~Elastic() {
        if (h) ela_hier_destroy(h);
        if (a) ela_amr_destroy(a);
    }
    Elastic(const Elastic&) = delete;
    Elastic& operator=(const Elastic&) = delete;

    void define(const Geometry& g, int mode = ELA_ALLOC_BOTH) {
        geom = g;
        ela_domain d{};
        d.dim = g.dim;
        for (int q = 0; q < 3; ++q) {
            d.n[q] = q < g.dim ? g.n[q] : 0;
            d.dx[q] = q < g.dim ? g.L[q] / g.n[q] : 1.0;
        }
        ela_detail::check(ela_hier_create(ela_detail::context(), &d, T::kind, 0, mode, &h), "define");
    }
    // Elastic<T>::define(geom[], grids[], dmap[], info) with several AMR levels (Elastic.cpp:30-55)
    void define(const Geometry& g, const std::vector<Box>& fine, int mode = ELA_SOLVE_MIXED) {
        geom = g;
        ela_domain d{};
        d.dim = g.dim;
        for (int q = 0; q < 3; ++q) {
            d.n[q] = q < g.dim ? g.n[q] : 0;
            d.dx[q] = q < g.dim ? g.L[q] / g.n[q] : 1.0;
        }
        std::vector<int> boxes;
        for (const Box& b : fine) {
            boxes.insert(boxes.end(), b.lo.begin(), b.lo.end());
            boxes.insert(boxes.end(), b.hi.begin(), b.hi.end());
        }
        ela_detail::check(ela_amr_create(ela_detail::context(), &d, (int)fine.size() + 1, boxes.data(), T::kind, 4,
                                         mode, &a),
                          "define");
        nlev = (int)fine.size() + 1;
        fine_boxes = fine;
    }
    int NumAMRLevels() const { return nlev; }
    // node counts and first global node index of an AMR level's arrays (the level's box grown by 2 nodes)
    std::array<int, 3> levelNodes(int amrlev) const {
        ela_level_layout l;
        ela_detail::check(ela_amr_level_info(a, amrlev, &l, nullptr), "ela_amr_level_info");
        return {l.nodes[0], l.nodes[1], l.nodes[2]};
    }
    std::array<int, 3> levelOrigin(int amrlev) const {
        ela_level_layout l;
        int o[3];
        ela_detail::check(ela_amr_level_info(a, amrlev, &l, o), "ela_amr_level_info");
        return {o[0], o[1], o[2]};
    }
    int levelSize(int amrlev) const {
        const auto n = levelNodes(amrlev);
        return n[0] * n[1] * n[2];
    }
    // SetModel(amrlev, FabArray) (Elastic.H:49-54): one model per node of the level's arrays
    void SetModel(int amrlev, const std::vector<T>& a_model) {
        if (!a) {
            if (amrlev != 0) throw std::runtime_error("SetModel: one AMR level");
            SetModel(a_model);
            return;
        }
        const int nn = levelSize(amrlev);
        if ((int)a_model.size() != nn) throw std::runtime_error("Inconsistent model field size");
        const int nc = (int)a_model[0].data.size();
        std::vector<double> C((size_t)nc * nn);
        for (int i = 0; i < nn; ++i)
            for (int q = 0; q < nc; ++q) C[(size_t)q * nn + i] = a_model[i].data[q];
        ela_detail::check(ela_amr_set_model_field(a, amrlev, C.data()), "SetModel");
    }
    void InitRHS(int amrlev, std::vector<double>& rhs, bool homogeneous = false) {
        if (!a) {
            InitRHS(rhs, homogeneous);
            return;
        }
        ela_detail::check(ela_amr_bc_init(a, amrlev, rhs.data(), homogeneous), "Init");
    }
    int nodes(int mglev = 0) const {
        ela_level_layout l;
        ela_detail::check(ela_level_info(h, mglev, &l), "ela_level_info");
        return l.nodes[0] * l.nodes[1] * l.nodes[2];
    }
    void SetModel(const T& a_model) {  // Elastic.cpp:57-82: the same model on every node
        const int nc = (int)a_model.data.size(), nn = nodes();
        std::vector<double> C((size_t)nc * nn);
        for (int q = 0; q < nc; ++q)
            for (int i = 0; i < nn; ++i) C[(size_t)q * nn + i] = a_model.data[q];
        ela_detail::check(ela_set_model_field(h, C.data()), "SetModel");
    }
    void SetModel(const std::vector<T>& a_model) {  // Elastic.cpp:84-118: one model per node
        const int nn = nodes();
        if ((int)a_model.size() != nn) throw std::runtime_error("Inconsistent model field size");
        const int nc = (int)a_model[0].data.size();
        std::vector<double> C((size_t)nc * nn);
        for (int i = 0; i < nn; ++i)
            for (int q = 0; q < nc; ++q) C[(size_t)q * nn + i] = a_model[i].data[q];
        ela_detail::check(ela_set_model_field(h, C.data()), "SetModel");
    }
    void SetBC(::BC::Operator::Elastic<T>* a_bc) {
        m_bc = a_bc;
        if (a) ela_detail::check(ela_amr_set_bc(a, a_bc->types.data(), a_bc->values.data()), "SetBC");
        else ela_detail::check(ela_set_bc(h, a_bc->types.data(), a_bc->values.data()), "SetBC");
    }
    ::BC::Operator::Elastic<T>& GetBC() { return *m_bc; }
    void SetUniform(bool a_uniform) {
        if (a) ela_detail::check(ela_amr_set_uniform(a, a_uniform), "SetUniform");
        else ela_detail::check(ela_set_uniform(h, a_uniform), "SetUniform");
    }
    void prepareForSolve() {
        if (a) ela_detail::check(ela_amr_prepare(a), "prepareForSolve");
        else ela_detail::check(ela_prepare(h), "prepareForSolve");
    }
    // BC::Operator::Elastic::Init on level 0
    void InitRHS(std::vector<double>& rhs, bool homogeneous = false) {
        ela_detail::check(ela_bc_init(h, rhs.data(), homogeneous), "Init");
    }
    // ------------------------------------------------ per-level operations (Elastic.H:116-136, Operator.H:79-100)
    int getNComp() const { return geom.dim; }         // Elastic.H:127
    bool isCrossStencil() const { return false; }     // Elastic.H:128
    int getNGrow() const { return 2; }                // Operator.H:95
    int NMGLevels(int amrlev = 0) const {
        single(amrlev);
        return ela_hier_levels(h);
    }
    // Node counts of MG level mglev (a field there has getNComp() * nodes(mglev) values).
    std::array<int, 3> levelShape(int mglev) const {
        ela_level_layout l;
        ela_detail::check(ela_level_info(h, mglev, &l), "ela_level_info");
        return {l.nodes[0], l.nodes[1], l.nodes[2]};
    }
    // SetHomogeneous (Elastic.H:48): stored as in the reference.  The models here ignore it in
    // their operator() (Solid.H:24-34), so it changes no operator here.
    void SetHomogeneous(bool a_homogeneous) { m_homogeneous = a_homogeneous; }
    bool GetHomogeneous() const { return m_homogeneous; }
    // Fapply(amrlev, mglev, out, in) (Elastic.cpp:151-260)
    void Fapply(int amrlev, int mglev, std::vector<double>& out, const std::vector<double>& in, int prec = ELA_F64) const {
        single(amrlev);
        Field du(h, mglev, prec, geom.dim), df(h, mglev, prec, geom.dim);
        du.up(in);
        ela_detail::check(ela_fapply(h, mglev, prec, du.p, df.p, nullptr), "Fapply");
        out.resize(in.size());
        df.down(out);
    }
    // Diagonal(amrlev, mglev, diag) (Elastic.cpp:264-352)
    void Diagonal(int amrlev, int mglev, std::vector<double>& diag) const {
        single(amrlev);
        Field dd(h, mglev, ELA_F64, geom.dim);
        ela_detail::check(ela_diagonal(h, mglev, ELA_F64, dd.p, nullptr), "Diagonal");
        diag.resize((size_t)geom.dim * nodes(mglev));
        dd.down(diag);
    }
    // Fsmooth(amrlev, mglev, x, b) (Operator.cpp:90-162): 2 weighted-Jacobi sweeps, omega = 2/3
    void Fsmooth(int amrlev, int mglev, std::vector<double>& x, const std::vector<double>& b, int prec = ELA_F64) const {
        single(amrlev);
        Field dx(h, mglev, prec, geom.dim), db(h, mglev, prec, geom.dim);
        dx.up(x);
        db.up(b);
        ela_detail::check(ela_fsmooth(h, mglev, prec, dx.p, db.p, 2, 2.0 / 3.0, nullptr), "Fsmooth");
        dx.down(x);
    }
    // normalize(amrlev, mglev, mf) (Operator.cpp:164-195), with the reference's component indexing
    void normalize(int amrlev, int mglev, std::vector<double>& mf) const {
        single(amrlev);
        Field d(h, mglev, ELA_F64, geom.dim);
        d.up(mf);
        ela_detail::check(ela_normalize(h, mglev, ELA_F64, d.p, nullptr), "normalize");
        d.down(mf);
    }
    // restriction(amrlev, cmglev, crse, fine) (Operator.cpp:454-554)
    void restriction(int amrlev, int cmglev, std::vector<double>& crse, std::vector<double>& fine) const {
        single(amrlev);
        Field df(h, cmglev - 1, ELA_F64, geom.dim), dc(h, cmglev, ELA_F64, geom.dim);
        df.up(fine);
        ela_detail::check(ela_restriction(h, cmglev, ELA_F64, df.p, dc.p, nullptr), "restriction");
        crse.resize((size_t)geom.dim * nodes(cmglev));
        dc.down(crse);
    }
    // interpolation(amrlev, fmglev, fine, crse) (Operator.cpp:556-624): fine += P crse
    void interpolation(int amrlev, int fmglev, std::vector<double>& fine, const std::vector<double>& crse) const {
        single(amrlev);
        Field df(h, fmglev, ELA_F64, geom.dim), dc(h, fmglev + 1, ELA_F64, geom.dim);
        df.up(fine);
        dc.up(crse);
        ela_detail::check(ela_interpolation(h, fmglev, ELA_F64, df.p, dc.p, nullptr), "interpolation");
        df.down(fine);
    }
    // solutionResidual(amrlev, resid, x, b) (Operator.cpp:841-850): resid = b - A x; returns |resid|inf
    double solutionResidual(int amrlev, std::vector<double>& resid, const std::vector<double>& x,
                            const std::vector<double>& b) const {
        return correctionResidual(amrlev, 0, resid, x, b);
    }
    // correctionResidual(amrlev, mglev, resid, x, b, ...) (Operator.cpp:852-863)
    double correctionResidual(int amrlev, int mglev, std::vector<double>& resid, const std::vector<double>& x,
                              const std::vector<double>& b) const {
        single(amrlev);
        Field dx(h, mglev, ELA_F64, geom.dim), db(h, mglev, ELA_F64, geom.dim), dr(h, mglev, ELA_F64, geom.dim);
        dx.up(x);
        db.up(b);
        double nrm = 0;
        ela_detail::check(ela_residual(h, mglev, ELA_F64, dx.p, db.p, dr.p, &nrm, nullptr), "residual");
        resid.resize(x.size());
        dr.down(resid);
        return nrm;
    }
    // Strain / Stress / Energy(amrlev, out, u, voigt) (Elastic.H:76-92, Elastic.cpp:399-549) on level 0
    void Strain(int amrlev, std::vector<double>& out, const std::vector<double>& u, bool voigt = false) const {
        post(amrlev, 0, out, u, voigt);
    }
    void Stress(int amrlev, std::vector<double>& out, const std::vector<double>& u, bool voigt = false) const {
        post(amrlev, 1, out, u, voigt);
    }
    void Energy(int amrlev, std::vector<double>& out, const std::vector<double>& u) const { post(amrlev, 2, out, u, false); }
    // Reflux(crse_amrlev, res, crse_sol, crse_rhs, fine_res, fine_sol, fine_rhs) (Elastic.H:104-107,
    // Operator.cpp:728-839) of a several-level operator: every coarse node under the fine box takes the
    // full-weighting restriction of the fine residual (fine values outside the box: zero ghosts).  Fields
    // are the levels' arrays over their grown boxes (levelNodes); only res and fine_res are read, as in
    // the reference, whose solution and rhs arguments are unused.
    void reflux(int crse_amrlev, std::vector<double>& res, const std::vector<double>& /*crse_sol*/,
                const std::vector<double>& /*crse_rhs*/, std::vector<double>& fine_res,
                const std::vector<double>& /*fine_sol*/, const std::vector<double>& /*fine_rhs*/) const {
        if (!a) throw std::runtime_error("reflux needs a several-level operator");
        if (crse_amrlev < 0 || crse_amrlev + 1 >= nlev) throw std::runtime_error("reflux: no finer level");
        const int dim = geom.dim, f = crse_amrlev + 1;
        const auto cn = levelNodes(crse_amrlev), co = levelOrigin(crse_amrlev);
        const auto fn = levelNodes(f), fo = levelOrigin(f);
        const Box& B = fine_boxes[f - 1];
        std::array<int, 3> dom{1, 1, 1}, bn{1, 1, 1};
        for (int d = 0; d < dim; ++d) {
            dom[d] = (geom.n[d] << crse_amrlev) + 1;
            bn[d] = B.hi[d] - B.lo[d] + 1;
        }
        const ela_level_layout Lc = ela_detail::dense_layout(dim, dom), Lw = ela_detail::dense_layout(dim, cn);
        ela_level_layout Lf = ela_detail::dense_layout(dim, bn);
        const ela_level_layout Lg = ela_detail::dense_layout(dim, fn);
        Lf.pitch = Lg.pitch;  // the fine box inside the fine level's grown-box array
        Lf.plane = Lg.plane;
        Lf.cstride = Lg.cstride;
        ela_detail::DevBuf dc(res), dfb(fine_res);
        long long off = 0;
        for (int d = 0; d < dim; ++d) off += (long long)(B.lo[d] - fo[d]) * (d == 0 ? 1 : d == 1 ? Lg.pitch : Lg.plane);
        int forig[3] = {B.lo[0], B.lo[1], dim == 3 ? B.lo[2] : 0}, corig[3] = {co[0], co[1], dim == 3 ? co[2] : 0};
        ela_detail::check(ela_reflux_window(&Lc, &Lw, corig, dc.p, &Lf, (const double*)dfb.p + off, forig, dim, ELA_F64,
                                            nullptr),
                          "reflux");
        dc.to(res);
    }
    void Reflux(int crse_amrlev, std::vector<double>& res, const std::vector<double>& crse_sol,
                const std::vector<double>& crse_rhs, std::vector<double>& fine_res, const std::vector<double>& fine_sol,
                const std::vector<double>& fine_rhs) const {
        reflux(crse_amrlev, res, crse_sol, crse_rhs, fine_res, fine_sol, fine_rhs);
    }

    std::vector<double> Fapply(int mglev, const std::vector<double>& u, int prec = ELA_F64) {
        void *du = nullptr, *df = nullptr;
        ela_detail::check(ela_field_alloc(h, mglev, prec, geom.dim, &du), "alloc");
        ela_detail::check(ela_field_alloc(h, mglev, prec, geom.dim, &df), "alloc");
        ela_detail::check(ela_upload(h, mglev, prec, geom.dim, u.data(), du), "upload");
        ela_detail::check(ela_fapply(h, mglev, prec, du, df, nullptr), "Fapply");
        std::vector<double> f(u.size());
        ela_detail::check(ela_download(h, mglev, prec, geom.dim, df, f.data()), "download");
        ela_field_free(h, du);
        ela_field_free(h, df);
        return f;
    }
    ela_hier* handle() { return h; }
    ela_amr* amr_handle() { return a; }
    Geometry geom;

private:
    // A device field of one MG level, freed with the object.
    struct Field {
        ela_hier* h;
        int mglev, prec, nc;
        void* p = nullptr;
        Field(ela_hier* a_h, int a_mglev, int a_prec, int a_nc) : h(a_h), mglev(a_mglev), prec(a_prec), nc(a_nc) {
            ela_detail::check(ela_field_alloc(h, mglev, prec, nc, &p), "ela_field_alloc");
        }
        ~Field() {
            if (p) ela_field_free(h, p);
        }
        void up(const std::vector<double>& host) { ela_detail::check(ela_upload(h, mglev, prec, nc, host.data(), p), "upload"); }
        void down(std::vector<double>& host) { ela_detail::check(ela_download(h, mglev, prec, nc, p, host.data()), "download"); }
        Field(const Field&) = delete;
        Field& operator=(const Field&) = delete;
    };
    void single(int amrlev) const {
        if (!h || amrlev != 0) throw std::runtime_error("per-level operations run on a single-box operator's amrlev 0");
    }
    void post(int amrlev, int what, std::vector<double>& out, const std::vector<double>& u, bool voigt) const {
        single(amrlev);
        const int dim = geom.dim;
        const int nout = what == 2 ? 1 : (voigt ? (dim == 3 ? 6 : 3) : dim * dim);
        Field du(h, 0, ELA_F64, dim), dout(h, 0, ELA_F64, nout);
        du.up(u);
        ela_detail::check(ela_postprocess(h, what, voigt, du.p, dout.p, nullptr), "postprocess");
        out.resize((size_t)nout * nodes(0));
        dout.down(out);
    }
    bool m_homogeneous = false;
    std::vector<Box> fine_boxes;
    ela_hier* h = nullptr;
    ela_amr* a = nullptr;  // several AMR levels
    int nlev = 1;
    ::BC::Operator::Elastic<T>* m_bc = nullptr;  // non-owning, as in the reference (Elastic.H:161)
};
}  // namespace Operator

// ---------------------------------------------------------------- solvers
namespace Solver {
namespace Nonlocal {
// The factor of W's trace: 1/2 for the Linear models and Affine::Cubic, the model's own otherwise.
template <class M, class = void>
struct w_factor {
    static constexpr double value = 0.5;
};
template <class M>
struct w_factor<M, decltype((void)M::w_factor)> {
    static constexpr double value = M::w_factor;
};
template <class M, class = void>
struct has_F0 {
    static constexpr bool value = false;
};
template <class M>
struct has_F0<M, decltype((void)M::affine)> {
    static constexpr bool value = M::affine;
};
// Solver::Nonlocal::Linear (src/Solver/Nonlocal/Linear.H:19-143): holds a reference to the operator.
template <class T>
class Linear {
public:
    explicit Linear(::Operator::Elastic<T>& a_op) : linop(a_op) {}
    void setMaxIter(int n) { max_iter = n; }
    void setVerbose(int v) { verbose = v; }
    void setMixedPrecision(bool on) { mode = on ? ELA_SOLVE_MIXED : ELA_SOLVE_FP64; }
    void setKrylov(int restart) { ela_detail::check(ela_set_krylov(linop.handle(), restart), "setKrylov"); }
    // solve(sol, rhs, tol_rel, tol_abs) -> final |r|inf (Linear.H:102-112); sol holds the initial guess.
    Set::Scalar solve(std::vector<double>& sol, const std::vector<double>& rhs, Set::Scalar tol_rel,
                      Set::Scalar tol_abs) {
        const int rc = ela_solve(linop.handle(), rhs.data(), sol.data(), tol_rel, tol_abs, mode, max_iter, &stats);
        if (rc != ELA_OK && rc != ELA_ERR_NONCONVERGED) ela_detail::check(rc, "solve");
        if (rc == ELA_ERR_NONCONVERGED) throw std::runtime_error(std::string("MLMG failed to converge: ") + ela_last_error());
        return stats.res;
    }
    // solve(Vector<MultiFab*> sol, rhs, tol_rel, tol_abs) over several AMR levels (Linear.H:102-112):
    // one host vector per level (the level's grown box); sol holds the initial guess.
    Set::Scalar solve(std::vector<std::vector<double>>& sol, const std::vector<std::vector<double>>& rhs,
                      Set::Scalar tol_rel, Set::Scalar tol_abs) {
        ela_amr* a = linop.amr_handle();
        if (!a) throw std::runtime_error("solve: the operator has one AMR level");
        for (int l = 0; l < (int)sol.size(); ++l) {
            ela_detail::check(ela_amr_upload(a, l, ELA_AMR_U, sol[l].data()), "upload");
            ela_detail::check(ela_amr_upload(a, l, ELA_AMR_B, rhs[l].data()), "upload");
        }
        const int rc = ela_amr_solve(a, tol_rel, tol_abs, max_iter, &stats);
        if (rc != ELA_OK && rc != ELA_ERR_NONCONVERGED) ela_detail::check(rc, "solve");
        if (rc == ELA_ERR_NONCONVERGED) throw std::runtime_error(std::string("MLMG failed to converge: ") + ela_last_error());
        for (int l = 0; l < (int)sol.size(); ++l)
            ela_detail::check(ela_amr_download(a, l, ELA_AMR_U, sol[l].data()), "download");
        return stats.res;
    }
    ela_stats stats{};

protected:
    ::Operator::Elastic<T>& linop;
    int max_iter = 200, verbose = 0, mode = ELA_SOLVE_FP64;
    Set::Scalar m_tol_rel = 1E-8, m_tol_abs = 1E-8;  // Linear.H:118
};

// Solver::Nonlocal::Newton<T> (src/Solver/Nonlocal/Newton.H:15-349) for the
// models whose kinematic variable is grad u: each step builds
// rhs = b - div(DDW : (grad u - F0)) (prepareForSolve, on the GPU: ela_newton_rhs),
// solves A du = rhs from du = 0 and adds du.  For the linear models from
// u = 0 the rhs is b, so this is exactly one Linear::solve, as in the
// reference; Affine models bring their eigenstrain F0.
template <class T>
class Newton : public Linear<T> {
public:
    explicit Newton(::Operator::Elastic<T>& a_op) : Linear<T>(a_op) {}
    void setNRIters(int n) { nriters = n; }
    Set::Scalar solve(std::vector<double>& u, const std::vector<double>& b, const std::vector<T>& model,
                      Set::Scalar tol_rel, Set::Scalar tol_abs) {
        this->linop.SetModel(model);
        this->linop.prepareForSolve();
        ela_hier* h = this->linop.handle();
        const int dim = this->linop.geom.dim, nn = this->linop.nodes();
        std::vector<double> F0;
        if constexpr (has_F0<T>::value) {
            F0.assign((size_t)dim * dim * nn, 0.0);
            for (int i = 0; i < nn; ++i)
                for (int a = 0; a < dim; ++a)
                    for (int c = 0; c < dim; ++c) F0[(size_t)(a * dim + c) * nn + i] = model[i].F0[3 * a + c];
        }
        void *du = nullptr, *db = nullptr, *dr = nullptr, *df = nullptr;
        ela_detail::check(ela_field_alloc(h, 0, ELA_F64, dim, &du), "alloc");
        ela_detail::check(ela_field_alloc(h, 0, ELA_F64, dim, &db), "alloc");
        ela_detail::check(ela_field_alloc(h, 0, ELA_F64, dim, &dr), "alloc");
        if (!F0.empty()) {
            ela_detail::check(ela_field_alloc(h, 0, ELA_F64, dim * dim, &df), "alloc");
            ela_detail::check(ela_upload(h, 0, ELA_F64, dim * dim, F0.data(), df), "upload");
        }
        ela_detail::check(ela_upload(h, 0, ELA_F64, dim, b.data(), db), "upload");
        Set::Scalar res = 0;
        std::vector<double> rhs(b.size());
        for (int it = 0; it < nriters; ++it) {
            ela_detail::check(ela_upload(h, 0, ELA_F64, dim, u.data(), du), "upload");
            ela_detail::check(ela_newton_rhs(h, du, db, df, dr, nullptr), "prepareForSolve");
            ela_detail::check(ela_download(h, 0, ELA_F64, dim, dr, rhs.data()), "download");
            std::vector<double> d(b.size(), 0.0);
            res = Linear<T>::solve(d, rhs, tol_rel, tol_abs);
            for (size_t q = 0; q < u.size(); ++q) u[q] += d[q];
        }
        for (void* p : {du, db, dr, df})
            if (p) ela_field_free(h, p);
        return res;
    }
    // several AMR levels, linear models: the Newton step from u solves the composite system from u
    Set::Scalar solve(std::vector<std::vector<double>>& u, const std::vector<std::vector<double>>& b,
                      const std::vector<std::vector<T>>& model, Set::Scalar tol_rel, Set::Scalar tol_abs) {
        for (int l = 0; l < (int)model.size(); ++l) this->linop.SetModel(l, model[l]);
        this->linop.prepareForSolve();
        return Linear<T>::solve(u, b, tol_rel, tol_abs);
    }

    // compResidual(res, u, b, model) (Newton.H:211-229): res = b - div DW(grad u) inside and
    // b - BC row(u, grad u, dw) on the boundary, the right-hand side of a Newton step from u.
    void compResidual(std::vector<double>& res, const std::vector<double>& u, const std::vector<double>& b,
                      const std::vector<T>& model) {
        this->linop.SetModel(model);
        Fields f(this->linop, model);
        f.u.up(u);
        f.b.up(b);
        ela_detail::check(ela_newton_rhs(f.h, f.u.p, f.b.p, f.F0p(), f.r.p, nullptr), "compResidual");
        res.resize(u.size());
        f.r.down(res);
    }
    // W(w, u, model) (Newton.H:231-275): the strain energy density W(grad u) per node.
    void W(std::vector<double>& w, const std::vector<double>& u, const std::vector<T>& model) {
        this->linop.SetModel(model);
        Fields f(this->linop, model);
        f.u.up(u);
        Buf wb(f.h, 1);
        ela_detail::check(ela_newton_w_dw(f.h, f.u.p, f.F0p(), w_factor<T>::value, wb.p, nullptr, nullptr), "W");
        w.resize((size_t)this->linop.nodes());
        wb.down(w);
    }
    // DW(dw, u, model) (Newton.H:277-333): dw = DW(grad u) per node, dim*dim components (row-major).
    void DW(std::vector<double>& dw, const std::vector<double>& u, const std::vector<T>& model) {
        this->linop.SetModel(model);
        Fields f(this->linop, model);
        f.u.up(u);
        const int dim = this->linop.geom.dim;
        Buf db(f.h, dim * dim);
        ela_detail::check(ela_newton_w_dw(f.h, f.u.p, f.F0p(), 0.0, nullptr, db.p, nullptr), "DW");
        dw.resize((size_t)dim * dim * this->linop.nodes());
        db.down(dw);
    }

private:
    struct Buf {
        ela_hier* h;
        int nc;
        void* p = nullptr;
        Buf(ela_hier* a_h, int a_nc) : h(a_h), nc(a_nc) { ela_detail::check(ela_field_alloc(h, 0, ELA_F64, nc, &p), "alloc"); }
        ~Buf() {
            if (p) ela_field_free(h, p);
        }
        void up(const std::vector<double>& v) { ela_detail::check(ela_upload(h, 0, ELA_F64, nc, v.data(), p), "upload"); }
        void down(std::vector<double>& v) { ela_detail::check(ela_download(h, 0, ELA_F64, nc, p, v.data()), "download"); }
        Buf(const Buf&) = delete;
        Buf& operator=(const Buf&) = delete;
    };
    // u, b, rhs and (affine models) the per-node F0 planes on level 0
    struct Fields {
        ela_hier* h;
        Buf u, b, r;
        std::unique_ptr<Buf> F0;
        Fields(::Operator::Elastic<T>& op, const std::vector<T>& model)
            : h(op.handle()), u(h, op.geom.dim), b(h, op.geom.dim), r(h, op.geom.dim) {
            if constexpr (has_F0<T>::value) {
                const int dim = op.geom.dim, nn = op.nodes();
                std::vector<double> f((size_t)dim * dim * nn);
                for (int i = 0; i < nn; ++i)
                    for (int a = 0; a < dim; ++a)
                        for (int c = 0; c < dim; ++c) f[(size_t)(a * dim + c) * nn + i] = model[i].F0[3 * a + c];
                F0.reset(new Buf(h, dim * dim));
                F0->up(f);
            }
        }
        void* F0p() const { return F0 ? F0->p : nullptr; }
    };
    int nriters = 1;  // Newton.H:337
};
}  // namespace Nonlocal
}  // namespace Solver
