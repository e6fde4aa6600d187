// ela_plotfile.cpp -- nodal plot files in AMReX's plotfile layout, the format
// the reference's integrators write (Integrator::WritePlotFile,
// src/Integrator/Integrator.cpp:582-687, through amrex::WriteMultiLevelPlotfile)
// so that a run of this library can be compared with a reference run by
// AMReX's fcompare.  Host code only.
//
// Layout (AMReX plotfile "HyperCLaw-V1.1", as AMReX documents it):
//   <dir>/Header                 the global header (variables, levels, domains,
//                                cell sizes, per-level box lists in physical
//                                coordinates, the path of each level's data);
//   <dir>/Level_<l>/Cell_H       the level's MultiFab header (VisMF version 1:
//                                box array, the file and byte offset of every
//                                box's data, per-box component minima and maxima);
//   <dir>/Level_<l>/Cell_D_00000 every box's data: a text FAB header (real
//                                descriptor, box, component count) followed by
//                                the values in FP64, x fastest, component-major.
// Boxes are nodal (index type (1,1,1)): a box of nodes lo..hi per direction.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ela.h"
#include "ela_internal.hpp"

namespace {

int pfail(int code, const std::string& m) { return ela::set_error(code, m); }

bool make_dir(const std::string& d) {
    if (mkdir(d.c_str(), 0755) == 0) return true;
    if (errno != EEXIST) return false;
    struct stat st;
    return stat(d.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::string ivec(const int* v, int dim) {
    std::ostringstream o;
    o << '(';
    for (int d = 0; d < dim; ++d) o << (d ? "," : "") << v[d];
    o << ')';
    return o.str();
}

// ((lo) (hi) (type)) of a box, AMReX's text form.
std::string box_text(const int* lo, const int* hi, int dim, int nodal) {
    const int t[3] = {nodal, nodal, nodal};
    return "(" + ivec(lo, dim) + " " + ivec(hi, dim) + " " + ivec(t, dim) + ")";
}

// The real descriptor of little-endian IEEE doubles, AMReX's FPC::NativeRealDescriptor on x86-64.
const char* kRealDesc = "((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))";

}  // namespace

extern "C" {

int ela_write_plotfile(const char* dir, int dim, int nlev, const ela_plot_level* levels, int ncomp,
                       const char* const* names, double time, int step, const double* prob_lo,
                       const double* prob_hi) {
    if (!dir || !levels || ncomp < 1 || !names || nlev < 1 || (dim != 2 && dim != 3))
        return pfail(ELA_ERR_BADARG, "bad arguments");
    for (int l = 0; l < nlev; ++l) {
        const ela_plot_level& L = levels[l];
        if (L.nboxes < 1 || !L.boxes || !L.data) return pfail(ELA_ERR_BADARG, "level without boxes or data");
        for (int b = 0; b < L.nboxes; ++b)
            for (int d = 0; d < dim; ++d) {
                const int lo = L.boxes[6 * b + d], hi = L.boxes[6 * b + 3 + d];
                if (lo > hi || lo < L.origin[d] || hi >= L.origin[d] + L.nodes[d])
                    return pfail(ELA_ERR_BADARG, "level " + std::to_string(l) + ": box outside its data array");
            }
    }
    const std::string root(dir);
    if (!make_dir(root)) return pfail(ELA_ERR_BADARG, "cannot create " + root);
    // ------------------------------------------------ the global header
    std::ofstream H(root + "/Header");
    if (!H) return pfail(ELA_ERR_BADARG, "cannot write " + root + "/Header");
    H.precision(17);
    H << "HyperCLaw-V1.1\n" << ncomp << '\n';
    for (int c = 0; c < ncomp; ++c) H << names[c] << '\n';
    H << dim << '\n' << time << '\n' << nlev - 1 << '\n';
    for (int d = 0; d < dim; ++d) H << prob_lo[d] << ' ';
    H << '\n';
    for (int d = 0; d < dim; ++d) H << prob_hi[d] << ' ';
    H << '\n';
    for (int l = 0; l + 1 < nlev; ++l) H << 2 << ' ';  // refinement ratio 2 between levels
    H << '\n';
    for (int l = 0; l < nlev; ++l) {  // the level's cell domain
        int lo[3] = {0, 0, 0}, hi[3];
        for (int d = 0; d < dim; ++d) hi[d] = levels[l].domain[d] - 1;
        H << box_text(lo, hi, dim, 0) << ' ';
    }
    H << '\n';
    for (int l = 0; l < nlev; ++l) H << step << ' ';
    H << '\n';
    for (int l = 0; l < nlev; ++l) {
        for (int d = 0; d < dim; ++d) H << (prob_hi[d] - prob_lo[d]) / levels[l].domain[d] << ' ';
        H << '\n';
    }
    H << 0 << '\n' << 0 << '\n';  // Cartesian coordinates; boundary width 0
    for (int l = 0; l < nlev; ++l) {
        const ela_plot_level& L = levels[l];
        H << l << ' ' << L.nboxes << ' ' << time << '\n' << step << '\n';
        for (int b = 0; b < L.nboxes; ++b)
            for (int d = 0; d < dim; ++d) {
                const double dx = (prob_hi[d] - prob_lo[d]) / L.domain[d];
                H << prob_lo[d] + dx * L.boxes[6 * b + d] << ' ' << prob_lo[d] + dx * L.boxes[6 * b + 3 + d] << '\n';
            }
        H << "Level_" << l << "/Cell\n";
    }
    H.close();
    if (!H) return pfail(ELA_ERR_BADARG, "writing " + root + "/Header failed");
    // ------------------------------------------------ per level: data and MultiFab header
    for (int l = 0; l < nlev; ++l) {
        const ela_plot_level& L = levels[l];
        const std::string ld = root + "/Level_" + std::to_string(l);
        if (!make_dir(ld)) return pfail(ELA_ERR_BADARG, "cannot create " + ld);
        std::ofstream D(ld + "/Cell_D_00000", std::ios::binary);
        if (!D) return pfail(ELA_ERR_BADARG, "cannot write " + ld + "/Cell_D_00000");
        const long long nx = L.nodes[0], ny = L.nodes[1], nz = dim == 3 ? L.nodes[2] : 1;
        const long long cs = nx * ny * nz;
        std::vector<long long> offs(L.nboxes);
        std::vector<double> mins((size_t)L.nboxes * ncomp), maxs((size_t)L.nboxes * ncomp);
        std::vector<double> buf;
        for (int b = 0; b < L.nboxes; ++b) {
            const int* lo = L.boxes + 6 * b;
            const int* hi = L.boxes + 6 * b + 3;
            offs[b] = (long long)D.tellp();
            D << "FAB " << kRealDesc << box_text(lo, hi, dim, 1) << ' ' << ncomp << '\n';
            const long long bx = hi[0] - lo[0] + 1, by = hi[1] - lo[1] + 1, bz = dim == 3 ? hi[2] - lo[2] + 1 : 1;
            buf.resize((size_t)(bx * by * bz));
            for (int c = 0; c < ncomp; ++c) {
                double mn = std::numeric_limits<double>::infinity(), mx = -mn;
                size_t q = 0;
                for (long long k = 0; k < bz; ++k)
                    for (long long j = 0; j < by; ++j)
                        for (long long i = 0; i < bx; ++i) {
                            const long long gi = lo[0] + i - L.origin[0], gj = lo[1] + j - L.origin[1],
                                            gk = dim == 3 ? lo[2] + k - L.origin[2] : 0;
                            const double v = L.data[c * cs + gi + nx * (gj + ny * gk)];
                            buf[q++] = v;
                            mn = std::fmin(mn, v);
                            mx = std::fmax(mx, v);
                        }
                D.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
                mins[(size_t)b * ncomp + c] = mn;
                maxs[(size_t)b * ncomp + c] = mx;
            }
        }
        D.close();
        if (!D) return pfail(ELA_ERR_BADARG, "writing " + ld + "/Cell_D_00000 failed");
        std::ofstream M(ld + "/Cell_H");
        if (!M) return pfail(ELA_ERR_BADARG, "cannot write " + ld + "/Cell_H");
        M.precision(17);
        M << 1 << '\n' << 0 << '\n' << ncomp << '\n' << 0 << '\n';  // VisMF version 1, one file, no ghosts
        M << '(' << L.nboxes << " 0\n";
        for (int b = 0; b < L.nboxes; ++b) M << box_text(L.boxes + 6 * b, L.boxes + 6 * b + 3, dim, 1) << '\n';
        M << ")\n" << L.nboxes << '\n';
        for (int b = 0; b < L.nboxes; ++b) M << "FabOnDisk: Cell_D_00000 " << offs[b] << '\n';
        M << '\n' << L.nboxes << ',' << ncomp << '\n';
        for (int b = 0; b < L.nboxes; ++b) {
            for (int c = 0; c < ncomp; ++c) M << mins[(size_t)b * ncomp + c] << ',';
            M << '\n';
        }
        M << '\n' << L.nboxes << ',' << ncomp << '\n';
        for (int b = 0; b < L.nboxes; ++b) {
            for (int c = 0; c < ncomp; ++c) M << maxs[(size_t)b * ncomp + c] << ',';
            M << '\n';
        }
        M.close();
        if (!M) return pfail(ELA_ERR_BADARG, "writing " + ld + "/Cell_H failed");
    }
    return ELA_OK;
}

}  // extern "C"
