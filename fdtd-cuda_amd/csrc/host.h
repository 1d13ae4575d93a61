// Host-side FP64 setup of the FDTD solver (no HIP types).
#ifndef FDTD_HOST_H_
#define FDTD_HOST_H_

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fdtd.h"

namespace fdtd {

// Physical constants as the reference defines them (define.h:22-29).
constexpr double kPI = 3.1415926535897932384626433832795;
constexpr double kEP0 = 8.85418782E-12;
constexpr double kMU0 = 1.25663706E-06;
constexpr double kC0 = 2.99792458E+08;
constexpr double kZ0 = 376.730313462;
constexpr double kE0 = 1.60217662E-19;
constexpr double kME = 9.10938356E-31;
constexpr double kPLANKS = 4.13566553853599E-15;   // eV s

// One CPML axis: recursion coefficients (b, c) at E positions (Nc+1 per
// slab) and H positions (Nc per slab), and the kappa-scaled spacings.
struct CpmlAxis {
    int Nc = 0;
    int F = 0;                      // first cell of the far slab (cpml_F_*)
    std::vector<double> beN, ceN, beF, ceF;   // Nc+1
    std::vector<double> bhN, chN, bhF, chF;   // Nc
    std::vector<double> ked, khd;             // N
};

// Drude + 2 critical-point ADE coefficients (setup.cu:1282-1373).
struct Ade {
    double eps_inf = 0, N_EQ = 0;
    double C_E = 0, C_E1 = 0, C_E2 = 0;
    double d[5] = {0, 0, 0, 0, 0};  // Drude pole d1..d5
    double dNL = 0;
    double c[2][5] = {{0}};         // CP poles C1..C5
};

struct Setup {
    fdtd_params P;
    int N[3] = {0, 0, 0};           // augmented grid
    double dt = 0;
    CpmlAxis cp[3];
    int tf0[3] = {0, 0, 0}, tf1[3] = {0, 0, 0};
    int abs_s[3], abs_e[3], sca_s[3], sca_e[3];
    int L = 0, m0 = 0, src = 0;
    double width = 0, delay = 0, f0 = 0;
    double theta = 0, phi = 0, psi = 0;
    double Chh = 1, Che = 0, Cee = 1, Ceh = 0;
    bool hydro = false;
    bool nonlocal1 = false;         // NONLOCAL == 1: grad-div term on the Drude pole
    Ade ade;
    std::vector<double> freq;
    std::vector<std::complex<double>> Einc;
    std::vector<uint8_t> mat[4];    // mat_matrix, X, Y, Z
};

// Dispersive E-component cells: the reference's 1 < id < 6 (main.cu:1175),
// plus id 7 so the Lorentz-Drude material (setup.cu:1256-1275), which the
// reference's kernels never treat as dispersive, takes the ADE branch.
inline bool dispersive_id(int v) { return (v > 1 && v < 6) || v == 7; }

void default_params(fdtd_params *p);
// Builds everything fdtd_create needs; returns FDTD_OK, or FDTD_ERR_INVALID /
// FDTD_ERR_UNSUPPORTED with a message.
int build_setup(const fdtd_params &p, Setup &s, std::string &err);
// Staircased sphere masks (setup.cu:219-248).
void sphere_masks(const Setup &s, double radius, std::vector<uint8_t> m[4]);
// Same rule around an arbitrary centre (cells).
void sphere_masks_at(const Setup &s, const double centre[3], double radius, std::vector<uint8_t> m[4]);
void rod_masks_at(const Setup &s, const double centre[3], int axis, double length, double radius,
                  std::vector<uint8_t> m[4]);
double source_pulse(const Setup &s, int t);
// DFT of the incident line at the source plane (e_inc[m0]) over the first
// `steps` steps, sampled and phased as the DFT monitors; one value per frequency.
std::vector<std::complex<double>> line_dft(const Setup &s, int steps);

}  // namespace fdtd

#endif
