// FP64 host precompute: grid, CPML profiles, dispersive coefficients, TFSF
// source, frequency list and source spectrum.  Follows SETUP_CONST
// (setup.cu:347-510), SETUP_CPML_X/Y/Z (setup.cu:513-977), SETUP_Drude_CP
// (setup.cu:979-1378), SETUP_TFSF (tfsf2.cu:6-56), the frequency list
// (main.cu:2446-2463) and the E_incident sum (main.cu:2629-2644).
#include "host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fdtd {

void default_params(fdtd_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->Tend = 1000;
    p->dx = p->dy = p->dz = 2e-9;
    p->NCELLX = p->NCELLY = p->NCELLZ = 64;
    p->NcpmlX = p->NcpmlY = p->NcpmlZ = 10;
    p->nano_sphere = 1;
    p->nano_sphere_radius = 20e-9;
    p->material = 3;
    p->first_medium = 0;
    p->NUM_freq = 16;
    p->WL_or_freq = 1;
    p->BandWidth = 6.79e14;
    p->NONLOCAL = 0;
    p->MAX_AMP = 1.0;
    p->TimeFactor = 1.7320508075688772;
    p->lam_0 = 600e-9;
    p->StaticField = 0;
    p->tfsf_inset = 10;
    p->abs_offset = 5;
    p->sca_offset = 5;
    p->cpml_exp = 3;
    p->kappa_max = 5;
    p->alpha_max = 0.05;
    p->sigma_factor = 0.8;
    p->lam_min = 200e-9;
    p->lam_max = 700e-9;
    p->dft_stride = 4;
    p->dft_block = 64;
    p->source_offset = 50;
    p->hydro_beta = 1.39e6 * std::sqrt(0.75);
    p->xs_area = 0;
    p->dft_exact = 0;
}

namespace {

// Graded CPML of one axis: sigma = smax*(rho/Nc)^m, kappa = 1+(kmax-1)*(rho/Nc)^m,
// alpha = amax*(1-rho/Nc); b = exp(-(sigma/kappa+alpha)*dt/eps0),
// c = sigma*(b-1)/(kappa*(sigma+kappa*alpha)).  E samples at integer depth,
// H at half-integer depth (setup.cu:540-681).
void cpml_profile(const fdtd_params &p, int Nc, int N, double d, double dt, CpmlAxis &a) {
    const double m = p.cpml_exp, kmax = p.kappa_max, amax = p.alpha_max;
    const double smax = p.sigma_factor * (m + 1) / (kZ0 * d);
    a.Nc = Nc;
    a.F = N - 1 - Nc;
    a.ked.assign(N, d);
    a.khd.assign(N, d);
    // xs: graded depth for sigma and kappa, xa: complementary depth for alpha,
    // both written as the reference evaluates them so the tables match bit for bit
    auto coef = [&](double xs, double xa, double &b, double &c) -> double {
        const double S = std::pow(xs, m) * smax;
        const double A = amax * std::pow(xa, 1.0);
        const double K = 1 + (kmax - 1) * std::pow(xs, m);
        b = std::exp(-(S / K + A) * (dt / kEP0));
        c = (K * A + S == 0.0) ? 0.0 : S * (b - 1) / (K * (S + K * A));
        return K;
    };
    const double n = (double)Nc;
    a.beN.assign(Nc + 1, 0); a.ceN.assign(Nc + 1, 0); a.beF.assign(Nc + 1, 0); a.ceF.assign(Nc + 1, 0);
    a.bhN.assign(Nc, 0); a.chN.assign(Nc, 0); a.bhF.assign(Nc, 0); a.chF.assign(Nc, 0);
    if (Nc == 0) return;   // periodic axis: no CPML, 1/(kappa d) stays 1/d
    for (int i = 0; i <= Nc; i++) {
        a.ked[i] *= coef((double)(Nc - i) / n, (double)i / n, a.beN[i], a.ceN[i]);
        a.ked[i + a.F] *= coef((double)i / n, (double)(Nc - i) / n, a.beF[i], a.ceF[i]);
    }
    for (int i = 0; i < Nc; i++) {
        a.khd[i] *= coef((n - (double)i - 0.5) / n, ((double)i + 0.5) / n, a.bhN[i], a.chN[i]);
        a.khd[i + a.F] *= coef(((double)i + 0.5) / n, (n - (double)i - 0.5) / n, a.bhF[i], a.chF[i]);
    }
}

}  // namespace

double source_pulse(const Setup &s, int t) {
    if (s.P.StaticField == 1) return (t < 1000) ? s.P.MAX_AMP * t / 1000.0 : s.P.MAX_AMP;
    // powf((t-delay)/width, 2) on a float-converted argument (main.cu:2636)
    const float b = (float)(((double)t - s.delay) / s.width);
    const float b2 = b * b;
    return s.P.MAX_AMP * std::exp(-(double)b2 / 2.0) * std::sin(2 * kPI * s.f0 * t * s.dt);
}

std::vector<std::complex<double>> line_dft(const Setup &s, int steps) {
    // the 1D line of k_line (UPDATE_h_inc, UPDATE_e_inc, Mur ends, additive
    // source: main.cu:2280-2337, 2831-2859) in FP64, and the DFT of e_inc[m0]
    // after every sampling step t (t % dft_stride == 0) with phase (t+1) dt
    const int L = s.L, F = (int)s.freq.size();
    std::vector<double> e(L, 0.0), en(L, 0.0), h(L, 0.0);
    double mur[2] = {0.0, 0.0};
    const double d1 = s.P.dx;
    const double ch = kZ0 * s.dt / (kMU0 * d1), ce = s.dt / (kZ0 * kEP0 * d1);
    const double fac = kC0 * s.dt, mc = (fac - d1) / (fac + d1);
    std::vector<std::complex<double>> out(F, 0.0);
    for (int t = 0; t < steps; t++) {
        for (int i = 0; i < L - 1; i++) h[i] = h[i] - ch * (e[i + 1] - e[i]);
        for (int i = 1; i < L - 1; i++) en[i] = e[i] - ce * (h[i] - h[i - 1]);
        en[0] = mur[0] + mc * (en[1] - e[0]);
        en[L - 1] = mur[1] + mc * (en[L - 2] - e[L - 1]);
        mur[0] = en[1];
        mur[1] = en[L - 2];
        en[s.src] += source_pulse(s, t);
        e.swap(en);
        if (s.P.StaticField == 0 && t % s.P.dft_stride == 0)
            for (int m = 0; m < F; m++) out[m] += e[s.m0] * std::exp(std::complex<double>(0, 2 * kPI * s.freq[m] * (t + 1.0) * s.dt));
    }
    return out;
}

void sphere_masks_at(const Setup &s, const double centre[3], double radius, std::vector<uint8_t> m[4]) {
    const int NX = s.N[0], NY = s.N[1], NZ = s.N[2];
    const int r = (int)(radius / s.P.dx);
    const double cx = centre[0], cy = centre[1], cz = centre[2];
    const uint8_t mat = (uint8_t)s.P.material;
    auto in = [&](double x, double y, double z) {
        return std::pow((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz), 0.5) < r;
    };
    const int lo[3] = {std::max(0, (int)(cx - r - 2)), std::max(0, (int)(cy - r - 2)), std::max(0, (int)(cz - r - 2))};
    const int hi[3] = {std::min(NX, (int)(cx + r + 3)), std::min(NY, (int)(cy + r + 3)), std::min(NZ, (int)(cz + r + 3))};
    for (int i = lo[0]; i < hi[0]; i++)
        for (int j = lo[1]; j < hi[1]; j++)
            for (int k = lo[2]; k < hi[2]; k++) {
                const size_t id = ((size_t)i * NY + j) * NZ + k;
                if (in(i, j, k)) m[0][id] = mat;
                if (in(i + 0.5, j, k)) m[1][id] = mat;
                if (in(i, j + 0.5, k)) m[2][id] = mat;
                if (in(i, j, k + 0.5)) m[3][id] = mat;
            }
}

// Nanorod: cylinder with hemispherical caps, total length `length`, along
// `axis`.  A Yee position is inside when its distance to the axis segment
// between the two cap centres is below radius/dx (cells; not truncated, so
// continuous diameters stay distinct).  No reference counterpart: the
// reference paints only spheres (setup.cu:219-248); the test per component
// position is the same.
void rod_masks_at(const Setup &s, const double centre[3], int axis, double length, double radius,
                  std::vector<uint8_t> m[4]) {
    const int N[3] = {s.N[0], s.N[1], s.N[2]};
    const double r = radius / s.P.dx;
    const double half = std::max(0.0, 0.5 * length / s.P.dx - r);   // half-length of the cylindrical part
    const uint8_t mat = (uint8_t)s.P.material;
    auto in = [&](double x, double y, double z) {
        double d[3] = {x - centre[0], y - centre[1], z - centre[2]};
        const double t = std::min(half, std::max(-half, d[axis]));
        d[axis] -= t;
        return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < r;
    };
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        const double ext = r + (a == axis ? half : 0.0) + 2.0;
        lo[a] = std::max(0, (int)std::floor(centre[a] - ext));
        hi[a] = std::min(N[a], (int)std::ceil(centre[a] + ext) + 1);
    }
    for (int i = lo[0]; i < hi[0]; i++)
        for (int j = lo[1]; j < hi[1]; j++)
            for (int k = lo[2]; k < hi[2]; k++) {
                const size_t id = ((size_t)i * N[1] + j) * N[2] + k;
                if (in(i, j, k)) m[0][id] = mat;
                if (in(i + 0.5, j, k)) m[1][id] = mat;
                if (in(i, j + 0.5, k)) m[2][id] = mat;
                if (in(i, j, k + 0.5)) m[3][id] = mat;
            }
}

// The reference centres its sphere on the cell NCELL/2 of the augmented grid.
void sphere_masks(const Setup &s, double radius, std::vector<uint8_t> m[4]) {
    const double c[3] = {(double)(s.N[0] / 2), (double)(s.N[1] / 2), (double)(s.N[2] / 2)};
    sphere_masks_at(s, c, radius, m);
}

int build_setup(const fdtd_params &p, Setup &s, std::string &err) {
    const int bad = FDTD_ERR_INVALID;
    s.P = p;
    if (!(p.dx > 0) || p.dx != p.dy || p.dx != p.dz) { err = "cubic cells required (dx == dy == dz > 0)"; return bad; }
    if (p.NCELLX < 4 || p.NCELLY < 4 || p.NCELLZ < 18) { err = "grid too small (NCELLZ must exceed 17 for the source cell)"; return bad; }
    if (p.periodic_xy) {
        if (p.periodic_xy != 1) { err = "periodic_xy must be 0 or 1"; return bad; }
        if (p.NcpmlX != 0 || p.NcpmlY != 0) { err = "periodic_xy needs NcpmlX = NcpmlY = 0 (CPML along z only)"; return bad; }
        if (p.inc_theta != 0.0) { err = "periodic_xy runs normal incidence only (inc_theta = 0)"; return FDTD_ERR_UNSUPPORTED; }
        if (p.NONLOCAL != 0) { err = "periodic_xy runs the local models only (NONLOCAL 0)"; return FDTD_ERR_UNSUPPORTED; }
        if (p.NcpmlZ < 1) { err = "Ncpml must be >= 1"; return bad; }
    } else if (p.NcpmlX < 1 || p.NcpmlY < 1 || p.NcpmlZ < 1) {
        err = "Ncpml must be >= 1";
        return bad;
    }
    if (p.NUM_freq < 1 || p.NUM_freq > 512) { err = "NUM_freq must be in [1, 512]"; return bad; }
    if (p.first_medium > 1) { err = "first_medium must be <= 1"; return bad; }
    if (p.dft_stride < 1) { err = "dft_stride must be >= 1"; return bad; }
    if (p.dft_block < 1 || p.dft_block > 256) { err = "dft_block must be in [1, 256]"; return bad; }
    if (p.dft_exact < 0 || p.dft_exact > 2) { err = "dft_exact must be 0, 1 or 2"; return bad; }
    if (!(p.TimeFactor > 0)) { err = "TimeFactor must be > 0"; return bad; }
    // dt = 0.5 dy / c / TimeFactor (setup.cu:359) must not exceed the 3D Courant
    // limit dx / (c sqrt(3)) for cubic cells, i.e. TimeFactor >= sqrt(3)/2: a
    // larger step diverges (the reference runs it and stops on a NaN probe,
    // main.cu:2728-2731)
    if (p.TimeFactor < 0.5 * std::sqrt(3.0)) {
        err = "TimeFactor below sqrt(3)/2: dt exceeds the 3D Courant limit dx/(c*sqrt(3))";
        return bad;
    }
    if (!(p.BandWidth > 0)) { err = "BandWidth must be > 0"; return bad; }
    if (p.NONLOCAL != 0 && p.NONLOCAL != 1 && p.NONLOCAL != 3) {
        err = "NONLOCAL must be 0 (local), 1 (grad-div nonlocal Drude pole) or 3 (hydrodynamic)";
        return FDTD_ERR_UNSUPPORTED;
    }
    if (p.material == 7 && p.NONLOCAL == 3) {
        err = "material 7 (Lorentz-Drude) has no hydrodynamic model (the reference defines N_EQ for 2 and 3 only)";
        return FDTD_ERR_UNSUPPORTED;
    }
    const int Nint[3] = {p.NCELLX, p.NCELLY, p.NCELLZ};
    const int Nc[3] = {p.NcpmlX, p.NcpmlY, p.NcpmlZ};
    // SETUP_CONST (setup.cu:347-400): dt, grid augmentation by 2*Ncpml + 1
    s.dt = 0.5 * p.dy / kC0 / p.TimeFactor;
    s.L = p.NCELLZ * 6;
    for (int a = 0; a < 3; a++) {
        s.N[a] = Nint[a] + 2 * Nc[a] + 1;
        s.tf0[a] = p.tfsf_inset + Nc[a];
        s.tf1[a] = s.N[a] - s.tf0[a];
        s.abs_s[a] = s.tf0[a] + p.abs_offset;
        s.abs_e[a] = s.N[a] - s.tf0[a] - p.abs_offset;
        s.sca_s[a] = s.tf0[a] - p.sca_offset;
        s.sca_e[a] = s.N[a] - s.tf0[a] + p.sca_offset;
        if (p.periodic_xy && a < 2) {
            // setup.cu:401-405, 429: x and y keep NCELL cells (no +1), the TFSF box and
            // the monitors span them
            s.N[a] = Nint[a];
            s.tf0[a] = 0;
            s.tf1[a] = s.N[a];
            s.abs_s[a] = s.sca_s[a] = 0;
            s.abs_e[a] = s.sca_e[a] = s.N[a];
            continue;
        }
        if (p.periodic_xy) {
            // reflection plane NtfsfZ - sca_offset, transmission plane NZ - NtfsfZ - sca_offset
            // (Spectral.cu:11-15), the z faces of the scattering box
            s.sca_e[a] = s.N[a] - s.tf0[a] - p.sca_offset + 1;
            if (s.sca_s[a] <= Nc[a] || s.sca_e[a] - 1 >= s.N[a] - 1 - Nc[a] || s.sca_e[a] - 1 <= s.tf0[a] + 1) {
                err = "periodic reflection/transmission planes do not fit between the z CPML slabs";
                return bad;
            }
            continue;
        }
        if (s.sca_s[a] < 1 || s.sca_e[a] > s.N[a] - 1 || s.abs_e[a] - s.abs_s[a] < 2 || s.tf0[a] <= Nc[a]) {
            err = "TFSF or monitor boxes do not fit the grid";
            return bad;
        }
    }
    const double d = p.dx;
    for (int a = 0; a < 3; a++) cpml_profile(p, Nc[a], s.N[a], d, s.dt, s.cp[a]);
    // frequency list (main.cu:2446-2463)
    const int F = p.NUM_freq;
    s.freq.assign(F, 0.0);
    for (int i = 0; i < F; i++) {
        const double den = F > 1 ? (F - 1.0) : 1.0;
        if (p.WL_or_freq == 1) s.freq[i] = kC0 / (p.lam_min + i * (p.lam_max - p.lam_min) / den);
        else s.freq[i] = kC0 / p.lam_max + i * (kC0 / p.lam_min - kC0 / p.lam_max) / den;
    }
    // SETUP_Drude_CP (setup.cu:1164-1373)
    s.hydro = (p.NONLOCAL == 3);
    double wD = 0, gD = 0, A[2] = {0, 0}, ph[2] = {0, 0}, Om[2] = {0, 0}, Ga[2] = {0, 0};
    if (p.material == 2) {
        s.ade.eps_inf = 1.4447; wD = 1.3280e16; gD = 9.1269e13;
        A[0] = -1.5951; ph[0] = 3.1288; Om[0] = 8.2749e15; Ga[0] = 5.1770e15;
        A[1] = 0.25261; ph[1] = -1.5066; Om[1] = 6.1998e15; Ga[1] = 5.4126e14;
        s.ade.N_EQ = 5.541318196761572e28;
    } else if (p.material == 3) {
        s.ade.eps_inf = 1.1431; wD = 1.3202e16; gD = 1.0805e14;
        A[0] = 0.26698; ph[0] = -1.2371; Om[0] = 3.8711e15; Ga[0] = 4.4642e14;
        A[1] = 3.0834; ph[1] = -1.0968; Om[1] = 4.1684e15; Ga[1] = 2.3555e15;
        s.ade.N_EQ = 5.476415562682574e28;
    } else if (p.material == 7) {
        // Lorentz-Drude (setup.cu:1256-1275), coefficients below
    } else {
        err = "material must be 2 (silver), 3 (gold) or 7 (Lorentz-Drude)";
        return bad;
    }
    const double dt = s.dt;
    s.nonlocal1 = p.NONLOCAL == 1;
    double s3 = 0, s4 = 0, s5 = 0;
    if (p.material == 7) {
        // setup.cu:1384-1402: Lorentz poles P^{n+1} = alpha P^n + psi P^{n-1} + eta <E> and
        // the Drude pole with alpha_HD1, psi_HD, eta_HD; the E update's C1_NL = eps_inf eps0 / dt
        // and C2_NL = sum eta / (4 dt) average <E> = (E^{n+1} + 2 E^n + E^{n-1}) / 4, so every
        // pole takes the five-coefficient form k1 = alpha, k2 = psi, k3 = k5 = eta/4, k4 = eta/2
        const double w = 2.0 * kPI / kPLANKS;
        s.ade.eps_inf = 3.559;
        const double deps[2] = {2.912, 1.272}, delta[2] = {w * 1.541, w * 0.525}, omg[2] = {w * 4.693, w * 3.112};
        const double wd = w * 8.812, gd = w * 0.0752;
        for (int n = 0; n < 2; n++) {
            const double den = 1.0 + delta[n] * dt;
            const double eta = kEP0 * deps[n] * omg[n] * omg[n] * dt * dt / den;
            double *c = s.ade.c[n];
            c[0] = (2.0 - omg[n] * omg[n] * dt * dt) / den;
            c[1] = -1.0 * (1.0 - delta[n] * dt) / den;
            c[2] = eta / 4.0; c[3] = eta / 2.0; c[4] = eta / 4.0;
        }
        const double eta_d = 2.0 * kEP0 * wd * wd * dt * dt / (2.0 + gd * dt);
        s.ade.d[0] = 4.0 / (2.0 + gd * dt);
        s.ade.d[1] = -1.0 * (2.0 - gd * dt) / (2.0 + gd * dt);
        s.ade.d[2] = eta_d / 4.0; s.ade.d[3] = eta_d / 2.0; s.ade.d[4] = eta_d / 4.0;
        // alpha_HD2 = 2 dt^2 beta^2 / (2 + gamma dt) (setup.cu:1396)
        s.ade.dNL = s.nonlocal1 ? 2.0 * dt * dt * p.hydro_beta * p.hydro_beta / (2.0 + gd * dt) : 0.0;
        s.ade.N_EQ = 0.0;
        s3 = s.ade.d[2] + s.ade.c[0][2] + s.ade.c[1][2];
        s4 = s.ade.d[3] + s.ade.c[0][3] + s.ade.c[1][3];
        s5 = s.ade.d[4] + s.ade.c[0][4] + s.ade.c[1][4];
    }
    for (int n = 0; n < 2 && p.material != 7; n++) {
        // second-order CP pole, central differences at n with E averaged over n+1, n, n-1
        const double a0 = 2.0 * kEP0 * A[n] * Om[n] * (Om[n] * std::cos(ph[n]) - Ga[n] * std::sin(ph[n]));
        const double a1 = -2.0 * kEP0 * A[n] * Om[n] * std::sin(ph[n]);
        const double b0 = Ga[n] * Ga[n] + Om[n] * Om[n], b1 = 2.0 * Ga[n], b2 = 1.0;
        const double C = b2 / (dt * dt) + b1 / (2.0 * dt) + b0 / 4.0;
        double *c = s.ade.c[n];
        c[0] = (2 * b2 / (dt * dt) - b0 / 2.0) / C;
        c[1] = (b1 / (2.0 * dt) - b2 / (dt * dt) - b0 / 4.0) / C;
        c[2] = (a0 / 4.0 + a1 / (2.0 * dt)) / C;
        c[3] = a0 / (2.0 * C);
        c[4] = (a0 / 4.0 - a1 / (2.0 * dt)) / C;
    }
    const double dd = 1.0 / (dt * dt) + gD / (2.0 * dt);
    if (p.material == 7) {
        // coefficients set above
    } else if (!s.hydro) {
        s.ade.d[0] = 2.0 / (dd * dt * dt);
        s.ade.d[1] = (gD / (2.0 * dt) - 1.0 / (dt * dt)) / dd;
        s.ade.d[2] = s.ade.d[4] = kEP0 * wD * wD / (4.0 * dd);
        s.ade.d[3] = kEP0 * wD * wD / (2.0 * dd);
        // NONLOCAL == 1: d_NL = NONLOC^2 / d_d, NONLOC = v_F sqrt(3/4) (setup.cu:1179, 1305)
        s.ade.dNL = s.nonlocal1 ? p.hydro_beta * p.hydro_beta / dd : 0.0;
        // summed Drude first, then the CP poles, as the reference does (setup.cu:1306-1332)
        s3 = s.ade.d[2] + s.ade.c[0][2] + s.ade.c[1][2];
        s4 = s.ade.d[3] + s.ade.c[0][3] + s.ade.c[1][3];
        s5 = s.ade.d[4] + s.ade.c[0][4] + s.ade.c[1][4];
    } else {
        s.ade.d[0] = 2.0 / (dd * dt * dt);
        s.ade.d[1] = (gD / (2.0 * dt) - 1.0 / (dt * dt)) / dd;
        s.ade.d[2] = -1.0 * kE0 / kME / dd / dt;
        s.ade.d[3] = 1.0 * kE0 / kME / dd / dt;
        s.ade.d[4] = -1.0 / dd;
        s.ade.dNL = std::pow(s.ade.N_EQ, 1.0 / 3.0) * p.hydro_beta * p.hydro_beta / dd;
        s3 = s.ade.c[0][2] + s.ade.c[1][2];
        s4 = s.ade.c[0][3] + s.ade.c[1][3];
        s5 = s.ade.c[0][4] + s.ade.c[1][4];
    }
    s.ade.C_E = kEP0 * s.ade.eps_inf + s3;
    s.ade.C_E1 = -s4 + kEP0 * s.ade.eps_inf;
    s.ade.C_E2 = s5;
    // uniform media (setup.cu:1943-2062): eps = EP0, mu = MU0, sigma = 0
    s.Chh = 1.0; s.Che = kZ0 * dt / kMU0; s.Cee = 1.0; s.Ceh = dt / (kZ0 * kEP0);
    // SETUP_TFSF (tfsf2.cu:6-56)
    s.theta = p.inc_theta * kPI / 180.0;
    s.phi = p.inc_phi * kPI / 180.0;
    s.psi = p.polar_psi * kPI / 180.0;
    s.m0 = (int)std::floor(s.L / 2);
    s.src = s.m0 - p.source_offset;
    if (s.src < 1) { err = "source cell outside the incident line"; return bad; }
    s.width = std::ceil((6 / (p.BandWidth * 2 * kPI)) / dt);
    s.delay = 6 * s.width;
    s.f0 = kC0 / p.lam_0;
    // the TFSF interpolation reaches m0 + (box diagonal projection); keep it inside the line
    {
        const double st = std::sin(s.theta), ct = std::cos(s.theta);
        const double ux = std::fabs(st * std::cos(s.phi)), uy = std::fabs(st * std::sin(s.phi)), uz = std::fabs(ct);
        const double reach = ux * (s.tf1[0] - s.tf0[0] + 1) + uy * (s.tf1[1] - s.tf0[1] + 1) + uz * (s.tf1[2] - s.tf0[2] + 1);
        if (s.m0 + reach + 2 >= s.L || s.m0 - reach - 2 < 0) { err = "incident line too short for the TFSF box"; return bad; }
    }
    // source spectrum over t < Tend (main.cu:2629-2644)
    s.Einc.assign(F, 0.0);
    for (int t = 0; t < p.Tend; t++) {
        const double pl = source_pulse(s, t);
        for (int m = 0; m < F; m++) s.Einc[m] += pl * std::exp(std::complex<double>(0, 2 * kPI * t * dt * s.freq[m]));
    }
    // MATERIAL_MATRIX (setup.cu:135-248) on the augmented grid
    const size_t Ncell = (size_t)s.N[0] * s.N[1] * s.N[2];
    for (int c = 0; c < 4; c++) s.mat[c].assign(Ncell, (uint8_t)p.first_medium);
    if (p.nano_sphere == 1) sphere_masks(s, p.nano_sphere_radius, s.mat);
    return FDTD_OK;
}

}  // namespace fdtd
