// fdtd_run: command-line driver with the reference program's inputs and
// outputs, over the C ABI (include/fdtd.h).  The reference's main()
// (main.cu:2341-3192) reads Input.txt (READ_DATA_FILE, setup.cu:10-58),
// writes SourceFile.txt (main.cu:2629-2644), runs the loop t = 0..Tend
// (main.cu:2724) and after step t, whenever t % (Tend/10) == 0, writes
// ScatAbs.<t>.txt with the 19 columns of CalculateAbsScatt
// (Spectral.cu:1153-1171).  This driver does the same with fdtd_step and
// fdtd_flux.  Every t_skip steps it writes the snapshots of the reference's
// loop (main.cu:2948-2990) in the formats of snapshot.cu: Snap_in 1 the
// column ey(2, 2, k) and the 1D line (aux_sim.<n>.txt, aux_sim.<n>_2.txt);
// 2 |P_d| of the Drude pole and the hydrodynamic density on the plane
// i = NX/2 (videofiles/YZ_sim.<n>.txt, YZ_sim.<n+1>_N.txt); 3 the 1D files
// and |E| on that plane; other values none.
//
//   fdtd_run [Input.txt] [--dir DIR] [--parse-only] [--xz]
//
// --parse-only prints the parsed parameters ("name value" lines, then the
// tokens the solver does not use) and exits without touching the device.
// --xz also writes the XZ planes j = NY/2 (XZ_sim.<n>.txt, XZ_sim.<n>_N.txt,
// every x_skip-th row), the branch the reference compiles out (snapshot.cu:32).
// Input.txt: one header token, then 41 "value name" pairs in the order
// below; the names are not checked, as in the reference.
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/fdtd.h"

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kC0 = 2.99792458e8;

// One token of READ_DATA_FILE (setup.cu:17-58): its name, whether it is read
// with %lf, and where it lands in fdtd_params (null: read and not used).
struct Token {
    const char *name;
    bool real;
    int *i;
    double *d;
};

std::vector<Token> tokens(fdtd_params &p) {
    return {
        {"Tend", false, &p.Tend, nullptr},
        {"dx", true, nullptr, &p.dx},
        {"dy", true, nullptr, &p.dy},
        {"dz", true, nullptr, &p.dz},
        {"num_trials", false, nullptr, nullptr},
        {"min_trials", false, nullptr, nullptr},
        {"max_trials", false, nullptr, nullptr},
        {"TE_TM", false, nullptr, nullptr},
        {"NCELLX", false, &p.NCELLX, nullptr},
        {"NCELLY", false, &p.NCELLY, nullptr},
        {"NCELLZ", false, &p.NCELLZ, nullptr},
        {"NcpmlX", false, &p.NcpmlX, nullptr},
        {"NcpmlY", false, &p.NcpmlY, nullptr},
        {"NcpmlZ", false, &p.NcpmlZ, nullptr},
        {"inf_disp_slab", false, nullptr, nullptr},
        {"nano_sphere", false, &p.nano_sphere, nullptr},
        {"nano_sphere_radius", true, nullptr, &p.nano_sphere_radius},
        {"material", false, &p.material, nullptr},
        {"first_medium", false, &p.first_medium, nullptr},
        {"inc_plane", false, nullptr, nullptr},
        {"dispersive_slab", false, nullptr, nullptr},
        {"NUM_freq", false, &p.NUM_freq, nullptr},
        {"WL_or_freq", false, &p.WL_or_freq, nullptr},
        {"NUM_freq_inc", false, nullptr, nullptr},
        {"Tend_inc", false, nullptr, nullptr},
        {"Snap_in", false, nullptr, nullptr},
        {"Freq_start", false, nullptr, nullptr},
        {"Test_offset", false, nullptr, nullptr},
        {"Spect_loc", false, nullptr, nullptr},
        {"BandWidth", true, nullptr, &p.BandWidth},
        {"NONLOCAL", false, &p.NONLOCAL, nullptr},
        {"Diverge_Gradient", false, nullptr, nullptr},
        {"t_skip", false, nullptr, nullptr},
        {"x_skip", false, nullptr, nullptr},
        {"y_skip", false, nullptr, nullptr},
        {"z_skip", false, nullptr, nullptr},
        {"Nonlocalend", true, nullptr, nullptr},
        {"MAX_AMP", true, nullptr, &p.MAX_AMP},
        {"TimeFactor", true, nullptr, &p.TimeFactor},
        {"lam_0", true, nullptr, &p.lam_0},
        {"StaticField", false, &p.StaticField, nullptr},
    };
}

// Parses Input.txt into p (defaults first); unused tokens go to `unused`.
bool read_input(const char *path, fdtd_params &p, std::vector<std::pair<std::string, std::string>> &unused,
                std::string &err) {
    unused.clear();
    std::ifstream in(path);
    if (!in) { err = std::string("cannot open ") + path; return false; }
    fdtd_default_params(&p);
    std::string header, value, name;
    if (!(in >> header)) { err = "empty input file"; return false; }
    for (const Token &t : tokens(p)) {
        if (!(in >> value >> name)) { err = std::string("input ends before ") + t.name; return false; }
        char *end = nullptr;
        if (t.real) {
            const double v = std::strtod(value.c_str(), &end);
            if (end == value.c_str()) { err = std::string("bad number for ") + t.name + ": " + value; return false; }
            if (t.d) *t.d = v;
            else unused.push_back({t.name, value});
        } else {
            const long v = std::strtol(value.c_str(), &end, 10);
            if (end == value.c_str()) { err = std::string("bad integer for ") + t.name + ": " + value; return false; }
            if (t.i) *t.i = (int)v;
            else unused.push_back({t.name, value});
        }
    }
    return true;
}

// The Gaussian-modulated sine written to SourceFile.txt (main.cu:2636-2639),
// with the float-converted square of its envelope argument.
double pulse(double amp, double delay, double width, double f0, double dt, int t) {
    const float b = (float)(((double)t - delay) / width);
    return amp * std::exp(-(double)(b * b) / 2.0) * std::sin(2 * kPi * f0 * t * dt);
}

// ScatAbs.<t>.txt (Spectral.cu:1153-1171): per frequency the wavelength,
// total scattering, absorption and extinction, the 12 face fluxes (near/far x,
// y, z of the scattering box, then of the absorption box) and |E_inc|,
// Re E_inc, Im E_inc.  With nano_sphere == 1 the totals are normalised by
// pi r^2 and the absorption sign flipped; otherwise the reference leaves the
// extinction unset and this writes 0 there.
bool write_scatabs(const std::string &dir, int t, const fdtd_params &p, const std::vector<double> &freq,
                   const std::vector<double> &einc, const std::vector<double> &raw) {
    const int F = (int)freq.size();
    const std::string path = dir + "/ScatAbs." + std::to_string(t) + ".txt";
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (int m = 0; m < F; m++) {
        double q[12];
        for (int face = 0; face < 12; face++) q[face] = raw[(size_t)face * F + m];
        double sca = -q[0] + q[1] - q[2] + q[3] - q[4] + q[5];
        double abs_ = -q[6] + q[7] - q[8] + q[9] - q[10] + q[11];
        double ext = 0.0;
        if (p.nano_sphere == 1) {
            const double a = kPi * p.nano_sphere_radius * p.nano_sphere_radius;
            sca /= a;
            abs_ = -abs_ / a;
            ext = abs_ + sca;
        }
        const std::complex<double> e(einc[2 * m], einc[2 * m + 1]);
        std::fprintf(f, "%e\t%e\t%e\t%e", kC0 / freq[m], sca, abs_, ext);
        for (int face = 0; face < 12; face++) std::fprintf(f, "\t%e", q[face]);
        std::fprintf(f, "\t%e\t%e\t%e\n", std::abs(e), e.real(), e.imag());
    }
    return std::fclose(f) == 0;
}

// Snapshot files of the reference (snapshot.cu), from planes fdtd_get_slice
// copies off the device.  snapshot_count numbers the files and advances once
// per SNAPSHOT_* call, as in the reference.
struct Snapshots {
    fdtd_sim *sim;
    fdtd_grid_info gi;
    std::string dir;
    int xskip, yskip, zskip;
    bool xz;       // also the XZ planes (the reference's XZ switch, snapshot.cu:32, 89)
    int count = 0;

    // comps' plane at i = NX/2 (axis 0, [j][k]) or j = NY/2 (axis 1, [i][k])
    bool plane(int comp, int axis, std::vector<float> &out) const {
        const int idx = axis == 0 ? gi.NX / 2 : gi.NY / 2;
        out.assign((size_t)(axis == 0 ? gi.NY : gi.NX) * gi.NZ, 0.f);
        return fdtd_get_slice(sim, comp, axis, idx, out.data()) == FDTD_OK;
    }
    // rows every rskip-th of n, columns every zskip-th: sqrt of the sum of squares
    // of the components (mag) or their value (one component), "%e " separated
    bool write(const std::string &path, const std::vector<float> *c, int ncomp, int n, int rskip) const {
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return false;
        for (int r = 0; r < n; r += rskip) {
            for (int k = 0; k < gi.NZ; k += zskip) {
                const size_t q = (size_t)r * gi.NZ + k;
                double v;
                if (ncomp == 3) v = std::sqrt((double)c[0][q] * c[0][q] + (double)c[1][q] * c[1][q] + (double)c[2][q] * c[2][q]);
                else v = c[0][q];
                std::fprintf(f, "%e ", v);
            }
            std::fprintf(f, "\n");
        }
        return std::fclose(f) == 0;
    }
    // SNAPSHOT_2D (snapshot.cu:25-76): |E| (first = 0) or |P_d| (first = 7, the
    // Snap_in == 2 copies of Px_d, Py_d, Pz_d into ex, ey, ez, main.cu:2962-2964)
    // on the YZ plane i = NX/2 (rows j every y_skip-th) and, with xz, the XZ
    // plane j = NY/2 (rows i every x_skip-th); columns k every z_skip-th
    bool shot_2d(int first) {
        const std::string n = std::to_string(count);
        std::vector<float> c[3];
        if (xz) {
            for (int q = 0; q < 3; q++)
                if (!plane(first + q, 1, c[q])) return false;
            if (!write(dir + "/videofiles/XZ_sim." + n + ".txt", c, 3, gi.NX, xskip)) return false;
        }
        for (int q = 0; q < 3; q++)
            if (!plane(first + q, 0, c[q])) return false;
        if (!write(dir + "/videofiles/YZ_sim." + n + ".txt", c, 3, gi.NY, yskip)) return false;
        count++;
        return true;
    }
    // SNAPSHOT_2D_N (snapshot.cu:80-133): (NDx + NDy + NDz) / 3; zero without
    // the hydrodynamic model
    bool shot_2d_n() {
        const std::string n = std::to_string(count);
        std::vector<float> c[1];
        for (int axis = xz ? 1 : 0; axis >= 0; axis--) {
            const int rows = axis == 0 ? gi.NY : gi.NX;
            if (gi.hydro) {
                if (!plane(6, axis, c[0])) return false;
            } else {
                c[0].assign((size_t)rows * gi.NZ, 0.f);
            }
            const std::string name = std::string(axis == 0 ? "YZ" : "XZ") + "_sim." + n + "_N.txt";
            if (!write(dir + "/videofiles/" + name, c, 1, rows, axis == 0 ? yskip : xskip)) return false;
        }
        count++;
        return true;
    }
    // SNAPSHOT_1D (snapshot.cu:137-175): ey on the column i = j = 2 (every
    // z_skip-th k) to aux_sim.<n>.txt and the 1D line e_inc (every z_skip-th
    // cell) to aux_sim.<n>_2.txt, one value per line, in the run directory
    bool shot_1d() {
        const std::string n = std::to_string(count);
        std::vector<float> ey((size_t)gi.NY * gi.NZ, 0.f);   // plane i = 2, row j = 2 read below
        if (fdtd_get_slice(sim, 1, 0, 2, ey.data()) != FDTD_OK) return false;
        FILE *f = std::fopen((dir + "/aux_sim." + n + ".txt").c_str(), "w");
        if (!f) return false;
        for (int k = 0; k < gi.NZ; k += zskip) std::fprintf(f, "%e\n", (double)ey[(size_t)2 * gi.NZ + k]);
        if (std::fclose(f) != 0) return false;
        std::vector<double> e(gi.inc_length), h(gi.inc_length);
        if (fdtd_get_line(sim, e.data(), h.data()) != FDTD_OK) return false;
        f = std::fopen((dir + "/aux_sim." + n + "_2.txt").c_str(), "w");
        if (!f) return false;
        for (int m = 0; m < gi.inc_length; m += zskip) std::fprintf(f, "%e\n", e[m]);
        if (std::fclose(f) != 0) return false;
        count++;
        return true;
    }
    // the snapshot block of the reference loop for one Snap_in value (main.cu:2948-2990)
    bool take(int snap_in) {
        if (snap_in == 1) return shot_1d();
        if (snap_in == 2) return shot_2d(7) && shot_2d_n();
        if (snap_in == 3) return shot_1d() && shot_2d(0);
        return true;
    }
};

int unused_int(const std::vector<std::pair<std::string, std::string>> &u, const char *name, int dflt) {
    for (const auto &x : u)
        if (x.first == name) return std::atoi(x.second.c_str());
    return dflt;
}

int fail(const std::string &m) {
    std::fprintf(stderr, "fdtd_run: %s\n", m.c_str());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    const char *input = "Input.txt";
    std::string dir = ".";
    bool parse_only = false, xz = false;
    for (int a = 1; a < argc; a++) {
        if (!std::strcmp(argv[a], "--parse-only")) parse_only = true;
        else if (!std::strcmp(argv[a], "--xz")) xz = true;
        else if (!std::strcmp(argv[a], "--dir") && a + 1 < argc) dir = argv[++a];
        else if (argv[a][0] == '-') return fail(std::string("unknown option ") + argv[a]);
        else input = argv[a];
    }
    fdtd_params p;
    std::vector<std::pair<std::string, std::string>> unused;
    std::string err;
    if (!read_input(input, p, unused, err)) return fail(err);
    if (parse_only) {
        for (const Token &t : tokens(p)) {
            if (t.i) std::printf("%s %d\n", t.name, *t.i);
            if (t.d) std::printf("%s %.17g\n", t.name, *t.d);
        }
        for (const auto &u : unused) std::printf("unused %s %s\n", u.first.c_str(), u.second.c_str());
        return 0;
    }
    // the dispersive half-space and slab of inf_disp_slab 1 and 2 (setup.cu:156-200)
    // are not built by this driver: refuse rather than run another geometry
    if (unused_int(unused, "inf_disp_slab", 0) != 0)
        return fail("inf_disp_slab != 0 (dispersive half-space or slab) is not supported; "
                    "periodic runs take a metal layer through fdtd_add_layer");
    fdtd_sim *sim = nullptr;
    if (fdtd_create(&p, &sim) != FDTD_OK) return fail(fdtd_last_error(nullptr));
    auto bail = [&](const char *what) {
        const std::string m = std::string(what) + ": " + fdtd_last_error(sim);
        fdtd_destroy(sim);
        return fail(m);
    };
    fdtd_grid_info gi;
    if (fdtd_get_info(sim, &gi) != FDTD_OK) return bail("fdtd_get_info");
    const int F = p.NUM_freq;
    std::vector<double> freq(F), einc(2 * (size_t)F), raw(12 * (size_t)F);
    if (fdtd_get_spectrum_setup(sim, freq.data(), einc.data()) != FDTD_OK) return bail("fdtd_get_spectrum_setup");
    {
        FILE *f = std::fopen((dir + "/SourceFile.txt").c_str(), "w");
        if (!f) return bail("SourceFile.txt");
        for (int t = 0; t < p.Tend; t++)
            std::fprintf(f, "%e\t%e\n", t * gi.dt, pulse(p.MAX_AMP, gi.delay, gi.width, gi.f_0, gi.dt, t));
        std::fclose(f);
    }
    // the loop t = 0..Tend; after step t the spectra when t % (Tend/10) == 0
    // and, with Snap_in != 0, a snapshot when t % t_skip == 0
    const int every = p.Tend / 10 > 0 ? p.Tend / 10 : 1;
    const int snap_in = unused_int(unused, "Snap_in", 0), t_skip = unused_int(unused, "t_skip", 0);
    Snapshots shots{sim, gi, dir, std::max(1, unused_int(unused, "x_skip", 1)), std::max(1, unused_int(unused, "y_skip", 1)),
                    std::max(1, unused_int(unused, "z_skip", 1)), xz};
    const bool snap = snap_in >= 1 && snap_in <= 3 && t_skip > 0;
    if (snap && snap_in != 1) {
        const std::string vd = dir + "/videofiles";
        if (::mkdir(vd.c_str(), 0755) != 0 && errno != EEXIST) return bail("videofiles");
    }
    int done = 0;
    for (int t = 0; t <= p.Tend; t++) {
        const bool spectra = t % every == 0 && p.StaticField == 0, shot = snap && t % t_skip == 0;
        if (!spectra && !shot) continue;
        if (fdtd_step(sim, t + 1 - done) != FDTD_OK) return bail("fdtd_step");
        done = t + 1;
        if (spectra) {
            if (fdtd_flux(sim, nullptr, nullptr, nullptr, raw.data()) != FDTD_OK) return bail("fdtd_flux");
            if (!write_scatabs(dir, t, p, freq, einc, raw)) return bail("ScatAbs");
            std::printf("t = %d: ScatAbs.%d.txt\n", t, t);
        }
        if (shot && !shots.take(snap_in)) return bail("snapshot");
    }
    if (done < p.Tend + 1 && fdtd_step(sim, p.Tend + 1 - done) != FDTD_OK) return bail("fdtd_step");
    if (fdtd_sync(sim) != FDTD_OK) return bail("fdtd_sync");
    fdtd_destroy(sim);
    return 0;
}
