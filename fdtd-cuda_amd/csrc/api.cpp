// C ABI implementation (include/fdtd.h): owns device memory, the solver
// stream and the step sequencing.  Replaces the reference's main() time loop
// (main.cu:2724-3175) and its globals with one handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fdtd.h"
#include "host.h"
#include "internal.h"

#include "sim.h"

using fdtd::Setup;
using namespace fdtd_host;

namespace {
thread_local std::string g_create_err = "no error";
}

namespace fdtd_host {

int set_err(fdtd_sim *s, int code, const std::string &m) {
    if (s) s->err = m;
    else g_create_err = m;
    return code;
}


std::vector<float> to_f(const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); }

StepDev make_step(fdtd_sim *s, bool fused) {
    const Setup &S = s->S;
    StepDev d{};
    d.g = s->g;
    d.c.Chh = (float)S.Chh; d.c.Che = (float)S.Che; d.c.Cee = (float)S.Cee; d.c.Ceh = (float)S.Ceh;
    for (int a = 0; a < 3; a++) {
        d.cp[a].ik_e = s->ik_e[a]; d.cp[a].ik_h = s->ik_h[a];
        d.cp[a].be = s->be[a]; d.cp[a].ce = s->ce[a]; d.cp[a].bh = s->bh[a]; d.cp[a].ch = s->ch[a];
        d.cp[a].Nc = S.cp[a].Nc; d.cp[a].F = S.cp[a].F;
    }
    TfsfDev &t = d.tf;
    for (int a = 0; a < 3; a++) { t.t0[a] = S.tf0[a]; t.t1[a] = S.tf1[a]; t.inv_d[a] = (float)(1.0 / S.P.dx); }
    if (S.P.periodic_xy) t.t1[2] = 1 << 30;   // the far z plane is not corrected (main.cu:698)
    const double st = std::sin(S.theta), ct = std::cos(S.theta), sp = std::sin(S.phi), cph = std::cos(S.phi);
    const double sps = std::sin(S.psi), cps = std::cos(S.psi);
    t.s[0] = (float)(st * cph); t.s[1] = (float)(st * sp); t.s[2] = (float)ct;
    // polarisation factors of the incident E and H components (main.cu:395-470)
    t.ef[0] = (float)(cps * sp - sps * ct * cph);
    t.ef[1] = (float)(-cps * cph - sps * ct * sp);
    t.ef[2] = (float)(sps * st);
    t.hf[0] = (float)(sps * sp + cps * ct * cph);
    t.hf[1] = (float)(-sps * cph + cps * ct * sp);
    t.hf[2] = (float)(-cps * st);
    t.cH = (float)S.Che; t.cE = (float)S.Ceh;
    t.dscale = 1.0f;  // d_1D = dx (tfsf2.cu:19)
    t.m0 = S.m0;
    t.e_inc = s->line_e[s->le];
    t.h_inc = s->line_h;
    d.tfsf_on = fused ? 1 : 0;
    d.fd = s->fd;
    d.sample_row = nullptr;
    return d;
}

LineDev make_line(fdtd_sim *s) {
    const Setup &S = s->S;
    LineDev l{};
    l.e_old = s->line_e[s->le];
    l.e_new = s->line_e[1 - s->le];
    l.h = s->line_h;
    l.mur = s->mur;
    l.L = S.L;
    l.src = S.src;
    const double d1 = S.P.dx;
    l.ch = fdtd::kZ0 * S.dt / (fdtd::kMU0 * d1);
    l.ce = S.dt / (fdtd::kZ0 * fdtd::kEP0 * d1);
    const double factor = fdtd::kC0 * S.dt;
    l.mur_coef = (factor - d1) / (factor + d1);
    l.amp = S.P.MAX_AMP;
    l.delay = S.delay;
    l.width = S.width;
    l.w0dt = 2 * fdtd::kPI * S.f0 * S.dt;
    l.static_field = S.P.StaticField == 1;
    l.t = s->t;
    l.probe = s->E[s->p][0] + s->probe_cell;
    l.status = s->status_d;
    return l;
}

int check_finite(fdtd_sim *s) {
    if (s->status_h && *(volatile unsigned *)s->status_h)
        return set_err(s, FDTD_ERR_NONFINITE,
                       "non-finite field value (NaN or Inf) detected at the probe cell or on a monitor face by step " +
                           std::to_string(s->t));
    return FDTD_OK;
}

DispDev make_disp(fdtd_sim *s, int parity) {
    const Setup &S = s->S;
    DispDev d{};
    d.g = s->g;
    for (int a = 0; a < 3; a++) { d.lo[a] = s->box_lo[a]; d.n[a] = s->box_n[a]; d.ik_e[a] = s->ik_e[a]; }
    for (int c = 0; c < 3; c++) {
        d.id[c] = s->disp_id[c];
        for (int q = 0; q < 8; q++) d.st[c][q] = s->disp_st[c][q];
        d.H[c] = s->H[parity][c];
        d.E[c] = s->E[parity][c];
    }
    // normalised ADE: P~ = P/eps0, 1/C_E folded into every coefficient
    const fdtd::Ade &A = S.ade;
    const double CE = A.C_E, e0 = fdtd::kEP0;
    d.a0 = (float)(S.dt / (fdtd::kZ0 * CE));
    d.a1 = (float)(A.C_E1 / CE);
    d.a2 = (float)(A.C_E2 / CE);
    const double *k[3] = {A.d, A.c[0], A.c[1]};
    for (int q = 0; q < 3; q++) {
        // (k1 + k2 - 1) vanishes identically for the Drude pole (setup.cu:1300-1301)
        const double sum = q == 0 ? 0.0 : (k[q][0] - 1.0) + k[q][1];
        d.e0[q] = (float)(sum * e0 / CE);
        d.e1[q] = (float)(-k[q][1] * e0 / CE);
        d.g0[q] = (float)sum;
        d.g1[q] = (float)(-k[q][1]);
        d.p3[q] = (float)(k[q][2] / e0);
        d.p4[q] = (float)(k[q][3] / e0);
        d.p5[q] = (float)(k[q][4] / e0);
    }
    d.rho = s->rho_box;
    {
        const double S3 = A.d[2] + A.c[0][2] + A.c[1][2], S4 = A.d[3] + A.c[0][3] + A.c[1][3], S5 = A.d[4] + A.c[0][4] + A.c[1][4];
        d.u1 = (float)((A.eps_inf - 1.0) + S3 / e0);
        d.u2 = (float)((A.eps_inf - 1.0) - S4 / e0);
        d.u3 = (float)(S5 / e0);
        d.A0 = (float)(S.dt / (fdtd::kZ0 * e0));
        for (int q = 0; q < 3; q++) {
            d.G0[q] = q == 0 ? 0.f : (float)((k[q][0] - 1.0) + k[q][1]);
            d.K2[q] = (float)k[q][1];
        }
    }
    for (int c = 0; c < 3; c++) {
        d.traj[c] = s->traj_step[c];
        d.traj_prev[0][c] = s->traj_first[0][c];
        d.traj_prev[1][c] = s->traj_first[1][c];
    }
    if (S.nonlocal1) {
        for (int c = 0; c < 3; c++) d.gd[c] = s->nl_gd[c];
        d.nl_e = (float)(A.dNL * e0 / CE);
        d.nl_p = (float)A.dNL;
        d.nl_i2 = (float)(1.0 / (S.P.dx * S.P.dx));
    }
    d.hydro = S.hydro ? 1 : 0;
    if (S.hydro) {
        const double dt = S.dt, dx = S.P.dx;
        for (int c = 0; c < 3; c++) { d.vnew[c] = s->h_vnew[c]; d.nd[c] = s->h_nd[c]; d.nd_prev[c] = s->h_ndp[c]; }
        d.h_d1m1 = A.d[0] - 1.0;
        d.h_d3 = A.d[2];
        d.h_d4 = A.d[3];
        // dNL / N_EQ^{1/3} = beta^2 / dd with n~ = ND / N_EQ (setup.cu:1362)
        d.h_nl = A.dNL / std::cbrt(A.N_EQ);
        d.h_inv_d = 1.0 / dx;
        d.h_cont = 2.0 * dt / dx;
        d.h_cur = dt * fdtd::kE0 * A.N_EQ / CE;
        d.h_inv_neq = 1.0 / A.N_EQ;
        // the E update and CP poles of the box in FP64 (the float set above, unrounded)
        d.ha0 = S.dt / (fdtd::kZ0 * CE);
        d.ha1 = A.C_E1 / CE;
        d.ha2 = A.C_E2 / CE;
        for (int q = 0; q < 3; q++) {
            const double sum = q == 0 ? 0.0 : (k[q][0] - 1.0) + k[q][1];
            d.he0[q] = sum * e0 / CE;
            d.he1[q] = -k[q][1] * e0 / CE;
            d.hg0[q] = sum;
            d.hg1[q] = -k[q][1];
            d.hp3[q] = k[q][2] / e0;
            d.hp4[q] = k[q][3] / e0;
            d.hp5[q] = k[q][4] / e0;
        }
    }
    return d;
}

// Frees the adjoint's buffers (sized by the grid and the dispersive box),
// so a new design or box allocates them afresh instead of beside them.
void adj_release(fdtd_sim *s) {
    (void)hipDeviceSynchronize();
    for (int q = 0; q < 2; q++)
        for (int c = 0; c < 3; c++) {
            if (s->adj.E[q][c]) { float *b = s->adj.E[q][c] - s->fbase; dfree(s, b); s->adj.E[q][c] = nullptr; }
            if (s->adj.H[q][c]) { float *b = s->adj.H[q][c] - s->fbase; dfree(s, b); s->adj.H[q][c] = nullptr; }
        }
    for (int q = 0; q < 2; q++)
        for (int a = 0; a < 6; a++) { dfree(s, s->adj.psiE[q][a]); dfree(s, s->adj.psiH[q][a]); }
    for (int c = 0; c < 3; c++) {
        for (int q = 0; q < 7; q++) dfree(s, s->adj.st[c][q]);
        dfree(s, s->adj.Qb[c]); dfree(s, s->adj.Eb[c]); dfree(s, s->adj.dq[c]); dfree(s, s->adj.db[c]);
    }
    dfree(s, s->adj.grad);
    dfree(s, s->adj.corr);
    dfree(s, s->adj.aux_d);
    s->adj.ready = false;
    s->adj.rows = s->adj.rows_slabs = false;
}

// Builds the dispersive box from the masks and uploads ids; state zeroed.
// The buffers of an earlier box (and the adjoint's, sized by it) are freed first.
int build_dispersive(fdtd_sim *s) {
    const Setup &S = s->S;
    const int NX = S.N[0], NY = S.N[1], NZ = S.N[2];
    adj_release(s);
    for (int c = 0; c < 3; c++) {
        dfree(s, s->disp_id[c]);
        for (int q = 0; q < 8; q++) dfree(s, s->disp_st[c][q]);
        dfree(s, s->h_vnew[c]); dfree(s, s->h_nd[c]); dfree(s, s->h_ndp[c]);
        dfree(s, s->nl_gd[c]);
    }
    dfree(s, s->rho_box);
    int lo[3] = {NX, NY, NZ}, hi[3] = {-1, -1, -1};
    for (int c = 0; c < 3; c++) s->n_disp[c] = 0;
    for (int c = 0; c < 3; c++) {
        const std::vector<uint8_t> &m = S.mat[c + 1];
        for (int i = 0; i < NX; i++)
            for (int j = 0; j < NY; j++)
                for (int k = 0; k < NZ; k++) {
                    const uint8_t v = m[((size_t)i * NY + j) * NZ + k];
                    if (fdtd::dispersive_id(v)) {
                        s->n_disp[c]++;
                        lo[0] = std::min(lo[0], i); lo[1] = std::min(lo[1], j); lo[2] = std::min(lo[2], k);
                        hi[0] = std::max(hi[0], i); hi[1] = std::max(hi[1], j); hi[2] = std::max(hi[2], k);
                    }
                }
    }
    const long total = (long)s->n_disp[0] + s->n_disp[1] + s->n_disp[2];
    if (total == 0) {
        for (int a = 0; a < 3; a++) { s->box_lo[a] = 0; s->box_n[a] = 0; }
        s->nbox = 0;
        return FDTD_OK;
    }
    if (S.P.periodic_xy) {
        // between the source plane and the transmission plane (off the z CPML, the TFSF
        // plane and both monitor planes); x and y unrestricted, no halo across the wrap
        if (lo[2] <= S.tf0[2] || hi[2] >= S.sca_e[2] - 1)
            return set_err(s, FDTD_ERR_INVALID, "periodic: dispersive cells must lie between the source plane and the transmission plane");
        for (int a = 0; a < 2; a++) { s->box_lo[a] = lo[a]; s->box_n[a] = hi[a] - lo[a] + 1; }
        s->box_lo[2] = lo[2] - 1;
        s->box_n[2] = hi[2] - lo[2] + 3;
    } else {
        for (int a = 0; a < 3; a++) {
            // dispersive cells must sit strictly inside the absorption monitor box, which keeps
            // them off the CPML, the TFSF faces and every monitor face
            if (lo[a] <= S.abs_s[a] || hi[a] >= S.abs_e[a] - 1)
                return set_err(s, FDTD_ERR_INVALID, "dispersive cells must lie strictly inside the absorption box");
        }
        // one halo cell on every side: hydro stencils read neighbours of metal cells
        for (int a = 0; a < 3; a++) { s->box_lo[a] = lo[a] - 1; s->box_n[a] = hi[a] - lo[a] + 3; }
    }
    s->nbox = (long)s->box_n[0] * s->box_n[1] * s->box_n[2];
    for (int c = 0; c < 3; c++) {
        std::vector<uint8_t> ids(s->nbox, 0);
        const std::vector<uint8_t> &m = S.mat[c + 1];
        for (int bi = 0; bi < s->box_n[0]; bi++)
            for (int bj = 0; bj < s->box_n[1]; bj++)
                for (int bk = 0; bk < s->box_n[2]; bk++) {
                    const uint8_t v = m[((size_t)(s->box_lo[0] + bi) * NY + s->box_lo[1] + bj) * NZ + s->box_lo[2] + bk];
                    ids[((size_t)bi * s->box_n[1] + bj) * s->box_n[2] + bk] = fdtd::dispersive_id(v) ? 1 : 0;
                }
        int rc = upload(s, &s->disp_id[c], ids);
        if (rc) return rc;
        for (int q = 0; q < 8; q++) {
            rc = dalloc(s, &s->disp_st[c][q], s->nbox);
            if (rc) return rc;
        }
        if (S.hydro) {
            if ((rc = dalloc(s, &s->h_vnew[c], s->nbox)) || (rc = dalloc(s, &s->h_nd[c], s->nbox)) ||
                (rc = dalloc(s, &s->h_ndp[c], s->nbox)))
                return rc;
        }
        if (S.nonlocal1 && (rc = dalloc(s, &s->nl_gd[c], s->nbox))) return rc;
    }
    s->rho_box = nullptr;
    if (!s->des_rho.empty()) {
        std::vector<float> rb(s->nbox, 1.0f);
        for (int a = 0; a < s->des_n[0]; a++)
            for (int b = 0; b < s->des_n[1]; b++)
                for (int c = 0; c < s->des_n[2]; c++) {
                    const int bi = s->des_lo[0] + a - s->box_lo[0], bj = s->des_lo[1] + b - s->box_lo[1], bk = s->des_lo[2] + c - s->box_lo[2];
                    rb[((size_t)bi * s->box_n[1] + bj) * s->box_n[2] + bk] = s->des_rho[((size_t)a * s->des_n[1] + b) * s->des_n[2] + c];
                }
        int rc2 = upload(s, &s->rho_box, rb);
        if (rc2) return rc2;
    }
    s->adj.ready = false;
    return FDTD_OK;
}

void build_faces(fdtd_sim *s) {
    const Setup &S = s->S;
    FaceDev &f = s->fd;
    long off = 0;
    for (int q = 0; q < 12; q++) {
        const int box = q / 6, axis = (q / 2) % 3, side = q % 2;
        const int *st = box ? S.abs_s : S.sca_s, *en = box ? S.abs_e : S.sca_e;
        const int outer = axis == 0 ? 1 : 0, inner = axis == 2 ? 1 : 2;
        f.plane[q] = side ? en[axis] - 1 : st[axis];
        f.lo_a[q] = st[outer];
        f.na[q] = en[outer] - st[outer];
        f.lo_b[q] = st[inner];
        f.nb[q] = en[inner] - st[inner];
        // periodic x, y: the reflection and transmission planes are the z faces of box 0
        if (S.P.periodic_xy && (box == 1 || axis != 2)) f.na[q] = f.nb[q] = 0;
        f.off[q] = off;
        off += 2L * f.na[q] * f.nb[q];
    }
    f.mE = off;
    f.mE_pad = (off + 255) / 256 * 256;
    s->mtot = f.mE_pad * 2;
}

hipEvent_t next_event(fdtd_sim *s, int which);

int contract(fdtd_sim *s) {
    if (s->slot == 0) return FDTD_OK;
    const int F = s->S.P.NUM_freq;
    const bool timed = s->timing && s->timing_what == 3;
    if (timed) {
        hipEvent_t a = next_event(s, 0);
        if (!a) return set_err(s, FDTD_ERR_DEVICE, "hipEventCreate failed");
        HIPCHK(s, hipEventRecord(a, s->stream));
    }
    if (s->S.P.dft_exact == 1) {
        HIPCHK(s, launch_twiddle(s->W, s->slot_t.data(), s->K, s->slot, F, s->freq_d, s->S.dt, s->stream));
        HIPCHK(s, launch_dft_contract(s->X, s->W, s->acc, s->mtot, s->fd.mE_pad, s->K, F, s->stream));
    } else {
        const int parts = s->S.P.dft_exact == 2 ? 3 : 2;
        HIPCHK(s, launch_twiddle_frag(s->Wf, s->slot_t.data(), s->K, s->slot, F, s->freq_d, s->S.dt, parts, s->stream));
        HIPCHK(s, launch_dft_mfma(s->X, s->Wf, s->acc, s->mtot, s->fd.mE_pad, s->K, F, parts, s->stream));
    }
    if (timed) {
        HIPCHK(s, hipEventRecord(next_event(s, 1), s->stream));
        s->ev_used++;
    }
    s->slot = 0;
    return FDTD_OK;
}

bool sampling_step(const fdtd_sim *s) { return s->S.P.StaticField == 0 && s->t % s->S.P.dft_stride == 0; }

int after_sample(fdtd_sim *s) {
    s->slot_t[s->slot] = s->t;
    s->slot++;
    s->samples++;
    if (s->slot == s->K) return contract(s);
    return FDTD_OK;
}

hipEvent_t next_event(fdtd_sim *s, int which) {
    if (s->ev_used >= s->events.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return nullptr;
        s->events.push_back({a, b});
    }
    return which == 0 ? s->events[s->ev_used].first : s->events[s->ev_used].second;
}


// The row-blocked kernel (k_step_rows) runs the whole grid with TFSF left
// out; k_tfsf_fix then adds it, which needs the TFSF cells and their
// neighbours free of CPML and with 1/(kappa d) at its interior value.  The
// kernel's table image (layout in kernels.hip) needs 1/(kappa d) constant
// outside 64 cells from either end of an axis and at most 31 CPML cells.
// When any of this fails, or the fields exceed 32-bit byte offsets, the
// general kernel (k_fused_step, TFSF inside) runs alone.
int plan_rows(fdtd_sim *s) {
    const Setup &S = s->S;
    s->use_rows = false;
    if (S.P.periodic_xy) return FDTD_OK;   // periodic steps run their own kernels
    if (const char *v = std::getenv("FDTD_PATH"))
        if (std::string(v) == "general") return FDTD_OK;
    if (const char *v = std::getenv("FDTD_ROWS")) s->rows = std::min(3, std::max(2, std::atoi(v)));
    if (const char *v = std::getenv("FDTD_CCHUNK")) s->cchunk = std::max(1, std::atoi(v));
    if (const char *v = std::getenv("FDTD_SCHUNK")) s->schunk = std::max(1, std::atoi(v));
    if (const char *v = std::getenv("FDTD_XCD")) s->xcd = std::atoi(v) != 0;
    if ((double)s->fsz * 4.0 >= 4294967296.0) return FDTD_OK;
    std::vector<uint32_t> img(kRowsAuxWords, 0u);
    auto put = [&](int q, float v) { std::memcpy(&img[q], &v, 4); };
    for (int a = 0; a < 3; a++) {
        const fdtd::CpmlAxis &c = S.cp[a];
        const int N = S.N[a], mid = N / 2;
        const std::vector<float> &ike = s->t_ike[a], &ikh = s->t_ikh[a];
        const int ew = std::max(c.Nc + 2, N - c.F + 1);
        if (ew > kRowsAuxMaxEdge || 2 * (c.Nc + 1) > kRowsAuxMaxLayers) return FDTD_OK;
        for (int x = ew; x < N - ew; x++)
            if (ike[x] != ike[mid] || ikh[x] != ikh[mid]) return FDTD_OK;
        auto plain = [&](int x) {
            const bool hl = x < c.Nc || x >= c.F, el = x <= c.Nc || x >= c.F;
            return x > 0 && x < N - 1 && !hl && !el && ikh[x] == ikh[mid] && ike[x] == ike[mid];
        };
        for (int x = S.tf0[a] - 2; x <= S.tf1[a] + 1; x++)
            if (!plain(x)) return FDTD_OK;
        const int base = a * kRowsAuxAxis;
        for (int x = 0; x < ew; x++) {
            const int xn = std::min(x, N - 1), xf = std::max(0, N - ew + x);
            put(base + x, ike[xn]); put(base + 64 + x, ike[xf]);
            put(base + 128 + x, ikh[xn]); put(base + 192 + x, ikh[xf]);
        }
        for (size_t q = 0; q < s->t_be[a].size(); q++) { put(base + 256 + (int)q, s->t_be[a][q]); put(base + 320 + (int)q, s->t_ce[a][q]); }
        for (size_t q = 0; q < s->t_bh[a].size(); q++) { put(base + 384 + (int)q, s->t_bh[a][q]); put(base + 448 + (int)q, s->t_ch[a][q]); }
        put(1536 + a, ike[mid]);
        put(1539 + a, ikh[mid]);
        img[1542 + a] = (uint32_t)ew;
    }
    // clean box: planes and rows with no CPML, interior 1/(kappa d) and every
    // domain test true, whose lower neighbour (the halo) is H-clean
    int lo[2], hi[2];
    bool clean = true;
    for (int a = 0; a < 2 && clean; a++) {
        const fdtd::CpmlAxis &c = S.cp[a];
        const int N = S.N[a], mid = N / 2;
        const std::vector<float> &ike = s->t_ike[a], &ikh = s->t_ikh[a];
        auto hl = [&](int x) { return x < c.Nc || x >= c.F; };
        auto el = [&](int x) { return x <= c.Nc || x >= c.F; };
        auto hclean = [&](int x) { return x >= 0 && x < N - 1 && !hl(x) && ikh[x] == ikh[mid]; };
        auto eclean = [&](int x) { return x > 0 && x < N - 1 && !el(x) && ike[x] == ike[mid]; };
        if (!hclean(mid) || !eclean(mid)) { clean = false; break; }
        int l = mid, h = mid + 1;
        while (l - 1 >= 0 && hclean(l - 1) && eclean(l - 1)) l--;
        while (h < N && hclean(h) && eclean(h)) h++;
        if (!hclean(l - 1)) l++;
        if (l >= h) clean = false;
        lo[a] = l; hi[a] = h;
        s->ik0[0][a] = ike[mid];
        s->ik0[1][a] = ikh[mid];
    }
    const int NX = S.N[0], NY = S.N[1];
    s->nxbox = s->nybox = s->ncorner = 0;
    s->has_clean = false;
    if (!clean) return FDTD_OK;   // the general kernel runs alone
    for (int a = 0; a < 3; a++)
        if (S.cp[a].Nc < 1) return FDTD_OK;
    s->has_clean = true;
    s->cbox[0] = lo[0]; s->cbox[1] = hi[0]; s->cbox[2] = lo[1]; s->cbox[3] = hi[1];
    auto add = [&](int (*b)[4], int &n, int i0, int i1, int j0, int j1) {
        if (i0 < i1 && j0 < j1) { b[n][0] = i0; b[n][1] = i1; b[n][2] = j0; b[n][3] = j1; n++; }
    };
    add(s->xbox, s->nxbox, 0, lo[0], lo[1], hi[1]);
    add(s->xbox, s->nxbox, hi[0], NX, lo[1], hi[1]);
    add(s->ybox, s->nybox, lo[0], hi[0], 0, lo[1]);
    add(s->ybox, s->nybox, lo[0], hi[0], hi[1], NY);
    add(s->cbox4, s->ncorner, 0, lo[0], 0, lo[1]);
    add(s->cbox4, s->ncorner, 0, lo[0], hi[1], NY);
    add(s->cbox4, s->ncorner, hi[0], NX, 0, lo[1]);
    add(s->cbox4, s->ncorner, hi[0], NX, hi[1], NY);
    float *ptrs[18];
    for (int q = 0; q < 2; q++)
        for (int x = 0; x < 6; x++) ptrs[q * 6 + x] = s->psiH[q][x];
    for (int x = 0; x < 6; x++) ptrs[12 + x] = s->psiE[x];
    std::memcpy(&img[1548], ptrs, sizeof(ptrs));
    s->aux_img = img;
    int rc = upload(s, &s->aux_d, img);
    if (rc) return rc;
    if (const char *v = std::getenv("FDTD_CONCURRENT")) s->concurrent = std::atoi(v) != 0;
    if (const char *v = std::getenv("FDTD_EARLY")) s->early = std::atoi(v) != 0;
    if (const char *v = std::getenv("FDTD_TILES")) s->tiles = std::min(3, std::max(1, std::atoi(v)));
    if (const char *v = std::getenv("FDTD_STILES")) s->stiles = std::atoi(v) == 2 ? 2 : 1;
    if (!s->tiles) s->tiles = rows_tiles(s->g.NZ);
    // x chunk of the clean box: 16, or 32 for two-tile blocks when that still
    // leaves at least 512 blocks (two per CU; fewer prologue planes, measured
    // best on the 281^3 grid; on small grids 32 leaves the device half idle)
    if (!s->cchunk) {
        s->cchunk = 16;
        const int W = rows_waves(s->g.NZ), T = s->tiles * W <= 16 ? s->tiles : 1;
        if (T >= 2) {
            const long njb = ((s->cbox[3] - s->cbox[2] + s->rows - 1) / s->rows + T - 1) / T;
            const long nseg = (s->g.NZ + 1 + 64 * W - 3) / (64 * W - 2);
            if (njb * nseg * ((s->cbox[1] - s->cbox[0] + 31) / 32) >= 512) s->cchunk = 32;
        }
    }
    if (!s->stiles) s->stiles = std::min(2, rows_tiles(s->g.NZ));
    if (s->concurrent && s->has_clean) {
        if (hipStreamCreateWithFlags(&s->stream2, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_line, hipEventDisableTiming) != hipSuccess)
            return set_err(s, FDTD_ERR_DEVICE, "second stream or events could not be created");
        if (const char *v = std::getenv("FDTD_SIDE_STREAMS")) s->side_streams = std::atoi(v) == 1 ? 1 : 3;
        if (const char *v = std::getenv("FDTD_CHAIN_JOIN")) s->chain_join = std::atoi(v) != 0;
        if (const char *v = std::getenv("FDTD_POST_AFTER_JOIN")) s->post_after_join = std::atoi(v) != 0;
        if (const char *v = std::getenv("FDTD_EARLY_SIDE")) s->early_side = std::min(2, std::max(0, std::atoi(v)));
        if (s->side_streams == 3 &&
            (hipStreamCreateWithFlags(&s->stream3, hipStreamNonBlocking) != hipSuccess ||
             hipStreamCreateWithFlags(&s->stream4, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreateWithFlags(&s->ev_join3, hipEventDisableTiming) != hipSuccess ||
             hipEventCreateWithFlags(&s->ev_join4, hipEventDisableTiming) != hipSuccess ||
             hipEventCreateWithFlags(&s->ev_early, hipEventDisableTiming) != hipSuccess))
            return set_err(s, FDTD_ERR_DEVICE, "side streams or events could not be created");
    }
    s->use_rows = s->has_clean;
    return FDTD_OK;
}

RowsArgs make_rows(fdtd_sim *s) {
    const Setup &S = s->S;
    RowsArgs r{};
    const int q = s->p;
    const long fb = s->fbase;
    for (int c = 0; c < 3; c++) {
        r.Ein[c] = (const char *)(s->E[q][c] - fb); r.Hin[c] = (const char *)(s->H[q][c] - fb);
        r.Eout[c] = (char *)(s->E[1 - q][c] - fb); r.Hout[c] = (char *)(s->H[1 - q][c] - fb);
    }
    r.aux = s->aux_d;
    r.Chh = (float)S.Chh; r.Che = (float)S.Che; r.Cee = (float)S.Cee; r.Ceh = (float)S.Ceh;
    r.NX = s->g.NX; r.NY = s->g.NY; r.NZ = s->g.NZ; r.sz = s->g.sz;
    r.sxy4 = (unsigned)(s->g.sxy * 4);
    for (int a = 0; a < 3; a++) { r.Nc[a] = S.cp[a].Nc; r.F[a] = S.cp[a].F; }
    r.pp = s->pp;
    r.ichunk = s->cchunk;
    r.xcd = s->xcd;
    for (int e = 0; e < 2; e++)
        for (int a = 0; a < 2; a++) r.ik0[e][a] = s->ik0[e][a];
    return r;
}

// Default row tiles per block of the row-blocked kernel: two tiles (whose
// halo rows the block then reads once) where a z column takes at most 5
// waves, so a block holds at most 10; with 8-wave columns the 16-wave block
// of two tiles measured slower than two 8-wave blocks per CU.
int rows_tiles(int NZ) { return rows_waves(NZ) <= 5 ? 2 : 1; }

// Whether the TFSF fix may run right after the clean-box kernel, before the
// side pieces finish: its cells (planes t0-1, t0, t1, t1+1 of each axis over
// [t0-1, t1+1] of the others) lie in the clean box in x and y, so no side
// piece reads or writes them, and the fix reads no field values.
bool tfsf_in_clean(const fdtd_sim *s) {
    const Setup &S = s->S;
    for (int a = 0; a < 2; a++)
        if (S.tf0[a] - 1 < s->cbox[2 * a] || S.tf1[a] + 1 >= s->cbox[2 * a + 1]) return false;
    return true;
}

// The step's update of the dispersive box: the local model, or the
// hydrodynamic launches and then their level shift, a swap of the roles of the
// velocity-state and density buffers (k_hydro_ce leaves v^{n+1} in h_vnew and
// n~^{n+1} in h_nd, and reads neither the old velocity state nor h_ndp of the
// component it updates at a neighbour after writing it).
hipError_t launch_box(fdtd_sim *s, const DispDev &dd) {
    if (!s->S.hydro) return launch_dispersive(dd, s->stream);
    const hipError_t e = launch_hydro(dd, s->stream);
    if (e == hipSuccess)
        for (int c = 0; c < 3; c++) {
            std::swap(s->disp_st[c][2], s->h_vnew[c]);
            std::swap(s->h_nd[c], s->h_ndp[c]);
        }
    return e;
}

// Whether the cells the TFSF fix-up writes ([t0-1, t1+1] on x and y) keep off
// the clean box's outer planes and rows, which the side pieces of the next
// step read.
bool tfsf_off_clean_edges(const fdtd_sim *s) {
    const Setup &S = s->S;
    for (int a = 0; a < 2; a++)
        if (S.tf0[a] - 1 <= s->cbox[2 * a] || S.tf1[a] + 1 >= s->cbox[2 * a + 1] - 1) return false;
    return true;
}

// Whether the dispersive (or hydrodynamic) box may run right after the
// clean-box kernel and ahead of the TFSF fix: the box with a 2-cell margin
// (its kernels read fields one cell beyond their own) lies in the clean box in
// x and y, and either strictly inside the TFSF planes on every axis or clear
// of [t0-1, t1+1] on one, so it shares no cell with the fix in either order.
bool disp_in_clean(const fdtd_sim *s) {
    if (s->nbox <= 0) return false;
    const Setup &S = s->S;
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) { lo[a] = s->box_lo[a] - 2; hi[a] = s->box_lo[a] + s->box_n[a] + 2; }   // [lo, hi)
    for (int a = 0; a < 2; a++)
        if (lo[a] < s->cbox[2 * a] || hi[a] > s->cbox[2 * a + 1]) return false;
    bool inside = true, clear = false;
    for (int a = 0; a < 3; a++) {
        if (lo[a] < S.tf0[a] + 1 || hi[a] - 1 > S.tf1[a] - 1) inside = false;
        if (hi[a] - 1 < S.tf0[a] - 1 || lo[a] > S.tf1[a] + 1) clear = true;
    }
    return inside || clear;
}

// Whether the box [lo, hi) meets a cell of the TFSF fix-up: planes t0-1, t0,
// t1, t1+1 of each axis over [t0-1, t1+1] of the other two.
static bool meets_tfsf(const Setup &S, const int lo[3], const int hi[3]) {
    for (int a = 0; a < 3; a++) {
        const int xs[4] = {S.tf0[a] - 1, S.tf0[a], S.tf1[a], S.tf1[a] + 1};
        for (int x : xs) {
            if (x < lo[a] || x >= hi[a]) continue;
            bool all = true;
            for (int b = 0; b < 3; b++)
                if (b != a && (hi[b] <= S.tf0[b] - 1 || lo[b] >= S.tf1[b] + 2)) all = false;
            if (all) return true;
        }
    }
    return false;
}

// Whether the DFT sample may run with the post-update work, before the side
// pieces finish: every monitor face lies in the clean box in x and y (no
// side piece writes it) and shares no cell with the TFSF fix-up or with the
// dispersive box and its 2-cell margin, so it reads the final fields.
bool sample_in_clean(const fdtd_sim *s) {
    const FaceDev &f = s->fd;
    for (int q = 0; q < 12; q++) {
        const int axis = (q >> 1) % 3;
        int lo[3], hi[3];
        const int ra[2] = {f.lo_a[q], f.lo_a[q] + f.na[q]}, rb[2] = {f.lo_b[q], f.lo_b[q] + f.nb[q]};
        // in-plane axes: x (y, z), y (x, z), z (x, y), as the sample kernel reads them
        const int oa = axis == 0 ? 1 : 0, ob = axis == 2 ? 1 : 2;
        lo[axis] = f.plane[q]; hi[axis] = f.plane[q] + 1;
        lo[oa] = ra[0]; hi[oa] = ra[1];
        lo[ob] = rb[0]; hi[ob] = rb[1];
        if (f.na[q] <= 0 || f.nb[q] <= 0) continue;
        for (int a = 0; a < 2; a++)
            if (lo[a] < s->cbox[2 * a] || hi[a] > s->cbox[2 * a + 1]) return false;
        if (meets_tfsf(s->S, lo, hi)) return false;
        if (s->nbox > 0) {
            bool apart = false;
            for (int a = 0; a < 3; a++)
                if (hi[a] <= s->box_lo[a] - 2 || lo[a] >= s->box_lo[a] + s->box_n[a] + 2) apart = true;
            if (!apart) return false;
        }
    }
    return true;
}

// The row-blocked field update over independent cell sets: the clean
// variant on the clean box, the x- and y-general variants on the slabs beside
// it, and the general kernel (TFSF off) on the four corner columns, then the
// TFSF fix.  With a second stream (FDTD_CONCURRENT=1) the 1D line and the
// side pieces run there beside the clean box, and the TFSF fix and the
// dispersive box (dd, when given) follow the clean box on the main stream
// without waiting for the side pieces wherever tfsf_in_clean and
// disp_in_clean allow, in one launch together with the DFT sample of the
// step (samp; d.sample_row set) where sample_in_clean allows.  Returns with
// *disp_done and *samp_done set when the dispersive box and the sample were
// launched here.
int launch_rows_update(fdtd_sim *s, const StepDev &d, const LineDev &l, const DispDev *dd, bool samp, bool *disp_done,
                       bool *samp_done, hipEvent_t ev_clean0, hipEvent_t ev_clean1) {
    RowsArgs r = make_rows(s);
    const bool fork = s->stream2 != nullptr;
    hipStream_t side = fork ? s->stream2 : s->stream;
    const bool tf_early = fork && s->early && tfsf_in_clean(s);
    const bool disp_early = fork && s->early && dd && disp_in_clean(s);
    *disp_done = false;
    *samp_done = false;
    // three side streams: x slabs, y slabs, and the corner columns (with the line first under early_side 2)
    const bool three = fork && s->stream3 && s->stream4;
    hipStream_t side_y = three ? s->stream3 : side, side_c = three ? s->stream4 : side;
    // early_side 2 (three side streams): the line goes first on the corner
    // stream, which keeps the plain fork, and the x slabs may start early too
    const bool line_c = three && s->early_side == 2;
    hipStream_t side_l = line_c ? s->stream4 : side;
    if (fork) {
        HIPCHK(s, hipEventRecord(s->ev_fork, s->stream));
        // the side pieces may start at the last step's join (ev_early); the stream with the line may not
        hipEvent_t e = s->early_ready ? s->ev_early : s->ev_fork;
        HIPCHK(s, hipStreamWaitEvent(s->stream2, line_c ? e : s->ev_fork, 0));
        if (three) {
            HIPCHK(s, hipStreamWaitEvent(s->stream3, e, 0));
            HIPCHK(s, hipStreamWaitEvent(s->stream4, line_c ? s->ev_fork : e, 0));
        }
    }
    s->early_ready = false;
    HIPCHK(s, launch_line(l, side_l));
    if (tf_early) HIPCHK(s, hipEventRecord(s->ev_line, side_l));
    if (s->nxbox) {
        RowsArgs rx = r;
        rx.ichunk = s->schunk;
        rx.tiles = s->stiles * rows_waves(s->g.NZ) <= 16 ? s->stiles : 1;
        rx.npiece = s->nxbox;
        std::memcpy(rx.box, s->xbox, sizeof(s->xbox));
        HIPCHK(s, launch_step_rows(rx, 2, 1, side));
    }
    if (s->nybox) {
        RowsArgs ry = r;
        ry.ichunk = s->schunk;
        ry.tiles = s->stiles * rows_waves(s->g.NZ) <= 16 ? s->stiles : 1;
        ry.npiece = s->nybox;
        std::memcpy(ry.box, s->ybox, sizeof(s->ybox));
        HIPCHK(s, launch_step_rows(ry, 2, 2, side_y));
    }
    // thin columns: short x chunks, 4 rows per block (tile rows 4, 8, 16 measured 0.381, 0.386, 0.386 ms per step)
    if (s->ncorner) HIPCHK(s, launch_fused_step(d, s->ctj, 2, s->cbox4, s->ncorner, side_c));
    if (three) {
        HIPCHK(s, hipEventRecord(s->ev_join3, s->stream3));
        HIPCHK(s, hipEventRecord(s->ev_join4, s->stream4));
        if (s->chain_join) {
            // the y slabs and corners join the x-slab stream, so the main stream waits once
            HIPCHK(s, hipStreamWaitEvent(s->stream2, s->ev_join3, 0));
            HIPCHK(s, hipStreamWaitEvent(s->stream2, s->ev_join4, 0));
        }
    }
    if (fork) HIPCHK(s, hipEventRecord(s->ev_join, s->stream2));
    r.npiece = 1;
    std::memcpy(r.box[0], s->cbox, sizeof(s->cbox));
    r.tiles = s->tiles * rows_waves(s->g.NZ) <= 16 ? s->tiles : 1;
    if (ev_clean0) HIPCHK(s, hipEventRecord(ev_clean0, s->stream));
    HIPCHK(s, launch_step_rows(r, s->rows, 0, s->stream));
    if (ev_clean1) HIPCHK(s, hipEventRecord(ev_clean1, s->stream));
    // with the side streams chained, the post-update launch follows the join (one
    // barrier on the main stream; the side pieces end before the clean box anyway)
    const bool after_join = three && s->chain_join && s->post_after_join;
    if (after_join) HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_join, 0));
    else if (tf_early) HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_line, 0));
    // The next step's side pieces may start here, beside this step's
    // post-update work, when that work writes no cell they read: they read
    // their own cells and the clean box's outer planes and rows (one cell deep)
    // of this step's outputs, and write the other ping-pong buffers; the TFSF
    // fix-up writes [t0-1, t1+1] on x and y (kept off those planes and rows
    // here), the dispersive box and the monitor sample stay inside the clean
    // box with a 2-cell margin.  The stream that runs the next step's line
    // keeps the fork after the post-update launch (the line would overwrite
    // the incident field the fix-up reads): the x-slab stream with
    // early_side 1, the corner stream with 2.
    if (after_join && s->early_side && tf_early && (s->nbox <= 0 || disp_early) && tfsf_off_clean_edges(s)) {
        HIPCHK(s, hipEventRecord(s->ev_early, s->stream));
        s->early_ready = true;
    }
    const bool merge_disp = disp_early && tf_early && !s->S.hydro;
    if (disp_early && !merge_disp) {
        HIPCHK(s, launch_box(s, *dd));
        *disp_done = true;
    }
    const bool merge_samp = tf_early && samp && d.sample_row && (s->nbox <= 0 || disp_early) && sample_in_clean(s);
    if (tf_early) {
        if (merge_disp || merge_samp) HIPCHK(s, launch_post(d, merge_disp ? dd : nullptr, merge_samp, s->stream));
        else HIPCHK(s, launch_tfsf_fix(d, s->stream));
        *disp_done = *disp_done || merge_disp;
        *samp_done = merge_samp;
    }
    if (fork && !after_join) HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_join, 0));
    if (three && !s->chain_join) {
        HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_join3, 0));
        HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_join4, 0));
    }
    if (!tf_early) HIPCHK(s, launch_tfsf_fix(d, s->stream));
    return FDTD_OK;
}

}  // namespace fdtd_host


// The entry points take C linkage from their declarations in fdtd.h.

void fdtd_default_params(fdtd_params *p) { fdtd::default_params(p); }

const char *fdtd_last_error(const fdtd_sim *sim) { return sim ? sim->err.c_str() : g_create_err.c_str(); }


int fdtd_create(const fdtd_params *p, fdtd_sim **out) {
    if (!p || !out) return set_err(nullptr, FDTD_ERR_INVALID, "null argument");
    *out = nullptr;
    fdtd_sim *s = new (std::nothrow) fdtd_sim();
    if (!s) return set_err(nullptr, FDTD_ERR_ALLOC, "host allocation failed");
    // parameters are checked before the device is touched, so invalid input
    // reports the same error with or without a GPU
    std::string err;
    if (const int rc = fdtd::build_setup(*p, s->S, err)) {
        delete s;
        return set_err(nullptr, rc, err);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete s;
        return set_err(nullptr, FDTD_ERR_DEVICE, "no HIP device");
    }
    const Setup &S = s->S;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return set_err(nullptr, FDTD_ERR_DEVICE, "hipGetDevice failed");
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return set_err(nullptr, FDTD_ERR_DEVICE, "hipStreamCreate failed");
    }
    if (const char *v = std::getenv("FDTD_TJ")) s->tj = std::atoi(v);
    if (const char *v = std::getenv("FDTD_CTJ")) s->ctj = std::atoi(v);
    if (s->ctj != 4 && s->ctj != 8 && s->ctj != 16) s->ctj = 4;
    if (const char *v = std::getenv("FDTD_ICHUNK")) s->ichunk = std::max(1, std::atoi(v));
    auto fail = [&](int rc) {
        std::string m = s->err;
        fdtd_destroy(s);
        return set_err(nullptr, rc, m);
    };
    int rc;
    Geom &g = s->g;
    g.NX = S.N[0]; g.NY = S.N[1]; g.NZ = S.N[2];
    g.sz = (g.NZ + 2 + 31) / 32 * 32;
    g.sxy = (long)g.sz * (g.NY + 2);
    const size_t fsz = (size_t)g.sxy * (g.NX + 2);
    const long base = g.sxy + g.sz + 1;
    s->fsz = fsz;
    s->fbase = base;
    for (int q = 0; q < 2; q++)
        for (int c = 0; c < 3; c++) {
            float *a, *b;
            if ((rc = dalloc(s, &a, fsz)) || (rc = dalloc(s, &b, fsz))) return fail(rc);
            s->E[q][c] = a + base;
            s->H[q][c] = b + base;
        }
    // psi slabs (ordering in internal.h)
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    const int ncx = S.cp[0].Nc, ncy = S.cp[1].Nc, ncz = S.cp[2].Nc;
    const size_t hz_ = (size_t)NX * NY * 2 * ncz, hy_ = (size_t)NX * 2 * ncy * NZ, hx_ = (size_t)2 * ncx * NY * NZ;
    const size_t ez_ = (size_t)NX * NY * 2 * (ncz + 1), ey_ = (size_t)NX * 2 * (ncy + 1) * NZ, ex_ = (size_t)2 * (ncx + 1) * NY * NZ;
    const size_t hsz[6] = {hz_, hy_, hz_, hx_, hx_, hy_};
    const size_t esz[6] = {ez_, ey_, ez_, ex_, ey_, ex_};
    for (int a = 0; a < 6; a++) { s->hsz[a] = hsz[a]; s->esz[a] = esz[a]; }
    for (int q = 0; q < 2; q++)
        for (int a = 0; a < 6; a++)
            if ((rc = dalloc(s, &s->psiH[q][a], hsz[a]))) return fail(rc);
    for (int a = 0; a < 6; a++)
        if ((rc = dalloc(s, &s->psiE[a], esz[a]))) return fail(rc);
    // CPML tables: 1/(kappa d), recursion b and c/d, near slab then far slab
    for (int a = 0; a < 3; a++) {
        const fdtd::CpmlAxis &c = S.cp[a];
        const double d = S.P.dx;
        std::vector<float> ike(c.ked.size()), ikh(c.khd.size());
        for (size_t i = 0; i < c.ked.size(); i++) { ike[i] = (float)(1.0 / c.ked[i]); ikh[i] = (float)(1.0 / c.khd[i]); }
        std::vector<float> be, ce, bh, ch;
        for (int i = 0; i <= c.Nc; i++) { be.push_back((float)c.beN[i]); ce.push_back((float)(c.ceN[i] / d)); }
        for (int i = 0; i <= c.Nc; i++) { be.push_back((float)c.beF[i]); ce.push_back((float)(c.ceF[i] / d)); }
        for (int i = 0; i < c.Nc; i++) { bh.push_back((float)c.bhN[i]); ch.push_back((float)(c.chN[i] / d)); }
        for (int i = 0; i < c.Nc; i++) { bh.push_back((float)c.bhF[i]); ch.push_back((float)(c.chF[i] / d)); }
        if ((rc = upload(s, &s->ik_e[a], ike)) || (rc = upload(s, &s->ik_h[a], ikh)) || (rc = upload(s, &s->be[a], be)) ||
            (rc = upload(s, &s->ce[a], ce)) || (rc = upload(s, &s->bh[a], bh)) || (rc = upload(s, &s->ch[a], ch)))
            return fail(rc);
        s->t_ike[a] = ike; s->t_ikh[a] = ikh; s->t_be[a] = be; s->t_ce[a] = ce; s->t_bh[a] = bh; s->t_ch[a] = ch;
    }
    if ((rc = plan_rows(s))) return fail(rc);
    // 1D line
    if ((rc = dalloc(s, &s->line_e[0], S.L)) || (rc = dalloc(s, &s->line_e[1], S.L)) || (rc = dalloc(s, &s->line_h, S.L)) ||
        (rc = dalloc(s, &s->mur, 2)))
        return fail(rc);
    // dispersive box
    if ((rc = build_dispersive(s))) return fail(rc);
    // divergence watch word in pinned host memory, written by the kernels
    if (hipHostMalloc((void **)&s->status_h, 256, hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void **)&s->status_d, s->status_h, 0) != hipSuccess) {
        s->err = "pinned status word could not be allocated";
        return fail(FDTD_ERR_ALLOC);
    }
    *(volatile unsigned *)s->status_h = 0u;
    s->probe_cell = gidx(s->g, g.NX / 2, g.NY / 2, g.NZ / 2 - g.NZ / 4);
    // DFT buffers
    build_faces(s);
    s->fd.status = s->status_d;
    s->K = S.P.dft_block;
    s->slot_t.assign(s->K, 0);
    const int F = S.P.NUM_freq;
    // the sample ring holds K rows padded to the MFMA k-step (16); padding rows stay zero
    const int kpad = (s->K + 15) / 16 * 16;
    const size_t wf_bytes = (size_t)2 * (kpad / 16) * ((2 * F + 31) / 32) * 3 * 64 * 16;   // up to three parts
    uint8_t *wf = nullptr;
    if ((rc = dalloc(s, &wf, wf_bytes))) return fail(rc);
    s->Wf = wf;
    if ((rc = dalloc(s, &s->X, (size_t)kpad * s->mtot)) || (rc = dalloc(s, &s->acc, (size_t)2 * F * s->mtot)) ||
        (rc = dalloc(s, &s->W, (size_t)2 * s->K * 2 * F)) || (rc = upload(s, &s->freq_d, S.freq)) ||
        (rc = dalloc(s, &s->flux_d, (size_t)12 * F)))
        return fail(rc);
    if ((rc = dalloc(s, &s->dense, (size_t)NX * NY * NZ))) return fail(rc);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(FDTD_ERR_DEVICE);
    *out = s;
    return FDTD_OK;
}

int fdtd_destroy(fdtd_sim *s) {
    if (!s) return FDTD_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto &e : s->events) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    if (s->stream2) (void)hipStreamSynchronize(s->stream2);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    if (s->ev_line) (void)hipEventDestroy(s->ev_line);
    for (hipStream_t q : {s->stream3, s->stream4})
        if (q) (void)hipStreamSynchronize(q);
    if (s->ev_join3) (void)hipEventDestroy(s->ev_join3);
    if (s->ev_join4) (void)hipEventDestroy(s->ev_join4);
    if (s->ev_early) (void)hipEventDestroy(s->ev_early);
    for (void *p : s->allocs) (void)hipFree(p);
    if (s->status_h) (void)hipHostFree(s->status_h);
    if (s->stream2) (void)hipStreamDestroy(s->stream2);
    if (s->stream3) (void)hipStreamDestroy(s->stream3);
    if (s->stream4) (void)hipStreamDestroy(s->stream4);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return FDTD_OK;
}

int fdtd_get_info(const fdtd_sim *s, fdtd_grid_info *g) {
    if (!s || !g) return FDTD_ERR_INVALID;
    const Setup &S = s->S;
    std::memset(g, 0, sizeof(*g));
    g->NX = S.N[0]; g->NY = S.N[1]; g->NZ = S.N[2];
    for (int a = 0; a < 3; a++) {
        g->Ncpml[a] = S.cp[a].Nc;
        g->tfsf[a] = S.tf0[a];
        g->abs_start[a] = S.abs_s[a]; g->abs_end[a] = S.abs_e[a];
        g->sca_start[a] = S.sca_s[a]; g->sca_end[a] = S.sca_e[a];
        g->n_dispersive[a] = s->n_disp[a];
        g->box_lo[a] = s->box_lo[a];
        g->box_hi[a] = s->box_lo[a] + s->box_n[a];
    }
    g->dt = S.dt;
    g->inc_length = S.L; g->m0 = S.m0;
    g->width = S.width; g->delay = S.delay; g->f_0 = S.f0;
    g->num_freq = S.P.NUM_freq;
    g->t = s->t;
    g->hydro = S.hydro;
    g->sample_count = s->samples;
    g->fused_path = s->use_rows ? 1 : 0;
    for (int a = 0; a < 4; a++) g->clean[a] = s->use_rows && s->has_clean ? s->cbox[a] : -1;
    const bool two = s->use_rows && s->stream2 && s->early;
    const bool tf = two && tfsf_in_clean(s), dp = two && disp_in_clean(s);
    g->schedule = (tf ? 1 : 0) | (dp ? 2 : 0) | (tf && (s->nbox <= 0 || dp) && sample_in_clean(s) ? 4 : 0);
    // bit 3: the next step's y slabs and corner columns start at the join (launch_rows_update)
    const bool three = two && s->stream3 && s->stream4 && s->chain_join && s->post_after_join;
    if (three && s->early_side && tf && (s->nbox <= 0 || dp) && tfsf_off_clean_edges(s)) g->schedule |= 8;
    if (s->use_rows) {
        const int W = rows_waves(s->g.NZ);
        g->tile[0] = s->rows;
        g->tile[1] = s->cchunk;
        g->tile[2] = s->schunk;
        g->tile[3] = s->tiles * W <= 16 ? s->tiles : 1;
        g->tile[4] = s->stiles * W <= 16 ? s->stiles : 1;
        g->tile[5] = s->ctj;
        g->tile[6] = s->stream2 ? (s->stream3 ? 3 : 1) : 0;
    } else {
        g->tile[0] = s->tj;
        g->tile[1] = s->ichunk;
    }
    return FDTD_OK;
}

int fdtd_set_materials(fdtd_sim *s, const uint8_t *mx, const uint8_t *my, const uint8_t *mz) {
    if (!s || !mx || !my || !mz) return FDTD_ERR_INVALID;
    const size_t N = (size_t)s->g.NX * s->g.NY * s->g.NZ;
    Setup &S = s->S;
    std::vector<uint8_t> old[3] = {S.mat[1], S.mat[2], S.mat[3]};
    S.mat[1].assign(mx, mx + N);
    S.mat[2].assign(my, my + N);
    S.mat[3].assign(mz, mz + N);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    int rc = build_dispersive(s);
    if (rc) {
        for (int c = 0; c < 3; c++) S.mat[c + 1] = old[c];
        std::string m = s->err;
        build_dispersive(s);
        s->err = m;
        return rc;
    }
    // the new dispersive cells start from the current field as E^n
    if (s->nbox) {
        DispDev d = make_disp(s, s->p);
        HIPCHK(s, launch_disp_sync_state(d, 1, s->stream));
    }
    return FDTD_OK;
}

int fdtd_add_sphere(fdtd_sim *s, const double centre[3], double radius) {
    if (!s || !centre || !(radius > 0)) return FDTD_ERR_INVALID;
    Setup &S = s->S;
    std::vector<uint8_t> m[4] = {S.mat[0], S.mat[1], S.mat[2], S.mat[3]};
    fdtd::sphere_masks_at(S, centre, radius, m);
    const int rc = fdtd_set_materials(s, m[1].data(), m[2].data(), m[3].data());
    if (rc) return rc;
    S.mat[0] = m[0];
    return FDTD_OK;
}

int fdtd_add_rod(fdtd_sim *s, const double centre[3], int axis, double length, double radius) {
    if (!s || !centre || axis < 0 || axis > 2 || !(radius > 0) || !(length > 0)) return FDTD_ERR_INVALID;
    Setup &S = s->S;
    std::vector<uint8_t> m[4] = {S.mat[0], S.mat[1], S.mat[2], S.mat[3]};
    fdtd::rod_masks_at(S, centre, axis, length, radius, m);
    const int rc = fdtd_set_materials(s, m[1].data(), m[2].data(), m[3].data());
    if (rc) return rc;
    S.mat[0] = m[0];
    return FDTD_OK;
}

int fdtd_set_design(fdtd_sim *s, const int lo[3], const int n[3], const float *rho) {
    if (!s || !lo || !n || !rho) return FDTD_ERR_INVALID;
    Setup &S = s->S;
    for (int a = 0; a < 3; a++) {
        if (n[a] < 1 || lo[a] <= S.abs_s[a] || lo[a] + n[a] - 1 >= S.abs_e[a] - 1)
            return set_err(s, FDTD_ERR_INVALID, "design block must lie strictly inside the absorption box");
    }
    if (S.hydro || S.nonlocal1) return set_err(s, FDTD_ERR_UNSUPPORTED, "design densities need the local model (NONLOCAL 0)");
    const size_t nv = (size_t)n[0] * n[1] * n[2];
    for (size_t v = 0; v < nv; v++)
        if (!(rho[v] >= 0.f && rho[v] <= 1.f)) return set_err(s, FDTD_ERR_INVALID, "densities must lie in [0, 1]");
    const int NY = S.N[1], NZ = S.N[2];
    std::vector<uint8_t> m[3] = {S.mat[1], S.mat[2], S.mat[3]};
    for (int a = 0; a < n[0]; a++)
        for (int b = 0; b < n[1]; b++)
            for (int c = 0; c < n[2]; c++) {
                const size_t id = ((size_t)(lo[0] + a) * NY + lo[1] + b) * NZ + lo[2] + c;
                for (int q = 0; q < 3; q++) m[q][id] = (uint8_t)S.P.material;
            }
    for (int a = 0; a < 3; a++) { s->des_lo[a] = lo[a]; s->des_n[a] = n[a]; }
    s->des_rho.assign(rho, rho + nv);
    return fdtd_set_materials(s, m[0].data(), m[1].data(), m[2].data());
}

namespace {

// Full forward state at a step boundary (fields at the current parity, psi,
// dispersive box, 1D line) in device memory.
struct Checkpoint {
    std::vector<void *> bufs;
    int t = 0, p = 0, pp = 0, le = 0;
};

int ckpt_copy(fdtd_sim *s, Checkpoint &ck, bool save) {
    std::vector<std::pair<void *, size_t>> items;   // (live pointer, bytes)
    for (int c = 0; c < 3; c++) {
        items.push_back({s->E[s->p][c] - s->fbase, s->fsz * sizeof(float)});
        items.push_back({s->H[s->p][c] - s->fbase, s->fsz * sizeof(float)});
    }
    for (int a = 0; a < 6; a++) items.push_back({s->psiH[s->pp][a], s->hsz[a] * sizeof(float)});
    for (int a = 0; a < 6; a++) items.push_back({s->psiE[a], s->esz[a] * sizeof(float)});
    if (s->nbox)
        for (int c = 0; c < 3; c++)
            for (int q = 0; q < 8; q++) items.push_back({s->disp_st[c][q], s->nbox * sizeof(float)});
    items.push_back({s->line_e[s->le], s->S.L * sizeof(double)});
    items.push_back({s->line_h, s->S.L * sizeof(double)});
    items.push_back({s->mur, 2 * sizeof(double)});
    // one allocation per checkpoint, items at 256-byte aligned offsets
    std::vector<size_t> off(items.size());
    size_t total = 0;
    for (size_t n = 0; n < items.size(); n++) {
        off[n] = total;
        total += (items[n].second + 255) / 256 * 256;
    }
    if (save) {
        void *b = nullptr;
        if (hipMalloc(&b, total) != hipSuccess) return set_err(s, FDTD_ERR_ALLOC, "checkpoint allocation failed");
        ck.bufs.push_back(b);
        for (size_t n = 0; n < items.size(); n++)
            HIPCHK(s, hipMemcpyAsync((char *)b + off[n], items[n].first, items[n].second, hipMemcpyDeviceToDevice, s->stream));
        ck.t = s->t; ck.p = s->p; ck.pp = s->pp; ck.le = s->le;
    } else {
        char *b = (char *)ck.bufs.at(0);
        for (size_t n = 0; n < items.size(); n++)
            HIPCHK(s, hipMemcpyAsync(items[n].first, b + off[n], items[n].second, hipMemcpyDeviceToDevice, s->stream));
    }
    return FDTD_OK;
}

void ckpt_free(Checkpoint &ck) {
    for (void *b : ck.bufs) (void)hipFree(b);
    ck.bufs.clear();
}

int adj_alloc(fdtd_sim *s) {
    if (s->adj.ready) return FDTD_OK;
    adj_release(s);
    int rc;
    for (int q = 0; q < 2; q++)
        for (int c = 0; c < 3; c++) {
            float *a, *b;
            if ((rc = dalloc(s, &a, s->fsz)) || (rc = dalloc(s, &b, s->fsz))) return rc;
            s->adj.E[q][c] = a + s->fbase;
            s->adj.H[q][c] = b + s->fbase;
        }
    for (int q = 0; q < 2; q++)
        for (int a = 0; a < 6; a++)
            if ((rc = dalloc(s, &s->adj.psiE[q][a], s->esz[a])) || (rc = dalloc(s, &s->adj.psiH[q][a], s->hsz[a]))) return rc;
    for (int c = 0; c < 3; c++) {
        for (int q = 0; q < 7; q++)
            if ((rc = dalloc(s, &s->adj.st[c][q], s->nbox))) return rc;
        if ((rc = dalloc(s, &s->adj.Qb[c], s->nbox)) || (rc = dalloc(s, &s->adj.Eb[c], s->nbox)) ||
            (rc = dalloc(s, &s->adj.dq[c], s->nbox)) || (rc = dalloc(s, &s->adj.db[c], s->nbox)))
            return rc;
    }
    if ((rc = dalloc(s, &s->adj.grad, s->nbox))) return rc;
    // transposed step on the clean box with the row-blocked kernel (variant 3):
    // needs the row path, wave-edge lanes outside the z slabs of E (their
    // published z weights carry no psi), and the dispersive box grown by two
    // cells inside the clean box and free of z CPML (launch_adj_box_fix)
    s->adj.rows = false;
    if (const char *v = std::getenv("FDTD_ADJ_SCHUNK")) s->adj_schunk = std::max(1, std::atoi(v));
    if (const char *v = std::getenv("FDTD_ADJ_RCHUNK")) s->adj_rchunk = std::max(1, std::atoi(v));
    if (const char *v = std::getenv("FDTD_ADJ_SCHED")) s->adj_sched = std::min(3, std::max(0, std::atoi(v)));
    if (const char *v = std::getenv("FDTD_ADJ_TILES")) s->adj_tiles = std::atoi(v) == 2 ? 2 : 1;
    if (const char *v = std::getenv("FDTD_ADJ_ROWS")) s->adj_rows = std::atoi(v) == 3 ? 3 : 2;
    if (const char *v = std::getenv("FDTD_ADJ_CTJ")) {
        const int tj = std::atoi(v);
        if (tj == 4 || tj == 8 || tj == 16) s->adj_ctj = tj;
    }
    if (const char *v = std::getenv("FDTD_ADJ_CICH")) s->adj_cich = std::max(1, std::atoi(v));
    if (const char *v = std::getenv("FDTD_ADJ_CORNER_ROWS")) s->adj_corner_rows = std::atoi(v) != 0;
    if (const char *v = std::getenv("FDTD_ADJ_CRCHUNK")) s->adj_crchunk = std::max(1, std::atoi(v));
    if (const char *v = std::getenv("FDTD_ADJ_CCHUNK")) s->adj_cchunk = std::max(1, std::atoi(v));
    if (const char *v = std::getenv("FDTD_ADJ_STILES")) s->adj_stiles = std::atoi(v) == 2 ? 2 : 1;
    if (const char *v = std::getenv("FDTD_ADJ_FUSE_INJECT")) s->adj_fuse_inject = std::atoi(v) != 0;
    if (const char *v = std::getenv("FDTD_ADJ_EARLY")) s->adj_early = std::atoi(v) != 0;
    if (const char *v = std::getenv("FDTD_ADJ_STJ")) {
        const int tj = std::atoi(v);
        if (tj == 4 || tj == 8 || tj == 16) s->adj_stj = tj;
    }
    const char *ap = std::getenv("FDTD_ADJ_PATH");
    if (s->use_rows && !(ap && std::string(ap) == "general") && s->S.Chh == 1.0) {
        const Setup &S = s->S;
        const int NZ = S.N[2], Ncz = S.cp[2].Nc, Fz = S.cp[2].F;
        const int W = rows_waves(NZ), nseg = (NZ + 1 + 64 * W - 3) / (64 * W - 2);
        bool ok = true;
        for (int seg = 0; seg < nseg && ok; seg++)
            for (int w = 1; w < W && ok; w++) {
                const int k = seg * (64 * W - 2) - 1 + 64 * w;
                if (k >= 0 && k < NZ && (k <= Ncz + 1 || k >= Fz - 1)) ok = false;
            }
        if (ok && s->nbox > 0) {
            for (int a = 0; a < 2; a++)
                if (s->box_lo[a] - 2 < s->cbox[2 * a] || s->box_lo[a] + s->box_n[a] + 2 > s->cbox[2 * a + 1]) ok = false;
            if (s->box_lo[2] - 2 <= Ncz + 1 || s->box_lo[2] + s->box_n[2] + 2 >= Fz - 1) ok = false;
        }
        if (ok) {
            std::vector<uint32_t> img = s->aux_img;
            float *ptrs[24];
            for (int q = 0; q < 2; q++)
                for (int x = 0; x < 6; x++) { ptrs[q * 6 + x] = s->adj.psiH[q][x]; ptrs[12 + q * 6 + x] = s->adj.psiE[q][x]; }
            std::memcpy(&img[1548], ptrs, sizeof(ptrs));
            if ((rc = upload(s, &s->adj.aux_d, img))) return rc;
            s->adj.rows = true;
            // x and y slabs on the row-blocked transposed variants (FDTD_ADJ_SLABS=general:
            // the general transposed kernel)
            const char *sl = std::getenv("FDTD_ADJ_SLABS");
            s->adj.rows_slabs = !(sl && std::string(sl) == "general");
            // amounts of the dispersive-cell correction (schedule 3)
            if (s->nbox > 0) {
                const long nc = (long)(s->box_n[0] + 2) * (s->box_n[1] + 2) * (s->box_n[2] + 2);
                if ((rc = dalloc(s, &s->adj.corr, 6 * nc))) return rc;
            }
        }
    }
    s->adj.ready = true;
    return FDTD_OK;
}

int adj_zero(fdtd_sim *s) {
    for (int q = 0; q < 2; q++)
        for (int c = 0; c < 3; c++) {
            HIPCHK(s, hipMemsetAsync(s->adj.E[q][c] - s->fbase, 0, s->fsz * sizeof(float), s->stream));
            HIPCHK(s, hipMemsetAsync(s->adj.H[q][c] - s->fbase, 0, s->fsz * sizeof(float), s->stream));
        }
    for (int q = 0; q < 2; q++)
        for (int a = 0; a < 6; a++) {
            HIPCHK(s, hipMemsetAsync(s->adj.psiE[q][a], 0, s->esz[a] * sizeof(float), s->stream));
            HIPCHK(s, hipMemsetAsync(s->adj.psiH[q][a], 0, s->hsz[a] * sizeof(float), s->stream));
        }
    for (int c = 0; c < 3; c++)
        for (int q = 0; q < 7; q++) HIPCHK(s, hipMemsetAsync(s->adj.st[c][q], 0, s->nbox * sizeof(float), s->stream));
    HIPCHK(s, hipMemsetAsync(s->adj.grad, 0, s->nbox * sizeof(double), s->stream));
    return FDTD_OK;
}

}  // namespace

namespace {

struct AdjScratch {
    float *probe = nullptr;             // E(r0) per forward step
    float *traj = nullptr;              // E on the box for one segment: slots E^{t-1}, E^t, ..., E^{t+len}
    std::vector<Checkpoint> cks;
};

int adjoint_run(fdtd_sim *s, AdjScratch &W, int nsteps, int ckpt, const int r0[3], int nf, const double *freq,
                double *J, double *grad);

}  // namespace

int fdtd_adjoint(fdtd_sim *s, int nsteps, int ckpt, const int r0[3], int nf, const double *freq, double *J,
                 double *grad) {
    if (!s || !r0 || !freq || !J || !grad || nsteps < 1 || nf < 1) return FDTD_ERR_INVALID;
    const Setup &S = s->S;
    if (s->des_rho.empty() || !s->rho_box) return set_err(s, FDTD_ERR_INVALID, "adjoint: no design region set");
    if (S.hydro || S.nonlocal1) return set_err(s, FDTD_ERR_UNSUPPORTED, "adjoint: nonlocal models not supported");
    if (S.P.periodic_xy) return set_err(s, FDTD_ERR_UNSUPPORTED, "adjoint: periodic simulations not supported");
    for (int a = 0; a < 3; a++)
        if (r0[a] < 0 || r0[a] >= S.N[a]) return set_err(s, FDTD_ERR_INVALID, "adjoint: objective cell outside the grid");
    for (int f = 0; f < nf; f++)
        if (!std::isfinite(freq[f])) return set_err(s, FDTD_ERR_INVALID, "adjoint: non-finite frequency");
    // the transposed dispersive update assumes no dispersive cell carries CPML terms
    for (int a = 0; a < 3; a++)
        if (s->box_lo[a] < S.cp[a].Nc || s->box_lo[a] + s->box_n[a] > S.cp[a].F)
            return set_err(s, FDTD_ERR_UNSUPPORTED, "adjoint: dispersive cells inside the CPML");
    AdjScratch W;
    const int rc = adjoint_run(s, W, nsteps, ckpt, r0, nf, freq, J, grad);
    (void)hipStreamSynchronize(s->stream);
    for (auto &c : W.cks) ckpt_free(c);
    if (W.probe) (void)hipFree(W.probe);
    if (W.traj) (void)hipFree(W.traj);
    s->probe = nullptr;
    s->sampling = true;
    for (int c = 0; c < 3; c++) s->traj_step[c] = s->traj_first[0][c] = s->traj_first[1][c] = nullptr;
    return rc;
}

namespace {

// One transposed step: adjoint level pa -> 1 - pa, psi_E pe, psi_H ph.  On
// the row path the clean box runs on the row-blocked kernel (variant 3) with
// the dispersive-cell correction after it, the x and y slabs and the corner
// columns on its slab and corner variants (4, 5, 6), beside it on the side
// streams.
//
// ad: the transposed dispersive update of this step (before the correction of
// the dispersive cells, which reads the box state it writes).  next: the
// injection of the following step, folded into this step's correction when
// it is on (the caller turns it on only where the objective cell lies in the
// clean box, which no side launch writes); *injected reports whether it was.
int adjoint_step(fdtd_sim *s, int pa, int pe, int ph, const AdjDispDev &ad, const AdjInject &next, bool *injected) {
    const Setup &S = s->S;
    AdjDev av{};
    const StepDev sd = make_step(s, false);
    av.g = sd.g; av.c = sd.c;
    for (int x = 0; x < 3; x++) av.cp[x] = sd.cp[x];
    for (int c = 0; c < 3; c++) {
        av.Ein[c] = s->adj.E[pa][c]; av.Hin[c] = s->adj.H[pa][c];
        av.Eout[c] = s->adj.E[1 - pa][c]; av.Hout[c] = s->adj.H[1 - pa][c];
        av.bid[c] = s->disp_id[c]; av.Qb[c] = s->adj.Qb[c]; av.Eb[c] = s->adj.Eb[c];
        av.dq[c] = s->adj.dq[c]; av.db[c] = s->adj.db[c];
    }
    for (int x = 0; x < 6; x++) {
        av.psiE_in.p[x] = s->adj.psiE[pe][x]; av.psiE_out.p[x] = s->adj.psiE[1 - pe][x];
        av.psiH_in.p[x] = s->adj.psiH[ph][x]; av.psiH_out.p[x] = s->adj.psiH[1 - ph][x];
    }
    for (int x = 0; x < 3; x++) { av.blo[x] = s->box_lo[x]; av.bn[x] = s->box_n[x]; }
    *injected = false;
    if (!s->adj.rows) {
        const int full[1][4] = {{0, s->g.NX, 0, s->g.NY}};
        HIPCHK(s, launch_adj_disp(ad, s->stream));
        HIPCHK(s, launch_adj_step(av, s->tj, s->ichunk, full, 1, s->stream));
        return FDTD_OK;
    }
    const bool fork = s->stream2 != nullptr;
    // with three side streams the x slabs, the y slabs and the corner columns each get one
    const bool three = fork && s->stream3 && s->stream4;
    hipStream_t side = fork ? s->stream2 : s->stream;
    // schedules 1-3: the transposed dispersive update beside the clean box,
    // first on the corner stream (it reads the input level only, and the row
    // kernels do not read the box state); the correction of the dispersive
    // cells after the clean box on that stream (1) or on the main stream (2),
    // or its amounts computed on the corner stream after the update and only
    // added to the outputs on the main stream after the clean box (3)
    const int sched = three && s->nbox > 0 && (s->adj_sched != 3 || s->adj.corr) ? s->adj_sched : 0;
    if (sched == 0) HIPCHK(s, launch_adj_disp(ad, s->stream));
    // early start (schedule 3): the x and y slabs of this step may start at the
    // last step's join (ev_early, recorded before its correction's addition);
    // the corner stream, whose dispersive update and correction amounts read
    // what the addition writes, keeps the fork
    const bool early_in = sched == 3 && s->adj_early_ready;
    s->adj_early_ready = false;
    if (fork) {
        HIPCHK(s, hipEventRecord(s->ev_fork, s->stream));
        HIPCHK(s, hipStreamWaitEvent(s->stream2, early_in ? s->ev_early : s->ev_fork, 0));
        if (three) {
            HIPCHK(s, hipStreamWaitEvent(s->stream3, early_in ? s->ev_early : s->ev_fork, 0));
            HIPCHK(s, hipStreamWaitEvent(s->stream4, s->ev_fork, 0));
        }
    }
    if (sched) {
        HIPCHK(s, launch_adj_disp(ad, s->stream4));
        if (sched == 3) HIPCHK(s, launch_adj_box_corr(av, s->adj.corr, s->stream4));
        if (sched >= 2) HIPCHK(s, hipEventRecord(s->ev_line, s->stream4));
    }
    RowsArgs r = make_rows(s);
    const long fb = s->fbase;
    for (int c = 0; c < 3; c++) {
        r.Ein[c] = (const char *)(s->adj.E[pa][c] - fb); r.Hin[c] = (const char *)(s->adj.H[pa][c] - fb);
        r.Eout[c] = (char *)(s->adj.E[1 - pa][c] - fb); r.Hout[c] = (char *)(s->adj.H[1 - pa][c] - fb);
    }
    r.aux = s->adj.aux_d;
    r.Chh = 1.0f;
    r.Che = (float)-S.Ceh;
    r.Cee = (float)S.Cee;
    r.Ceh = (float)-S.Che;
    r.pp = ph;
    r.pe = pe;
    // x and y slabs on the row-blocked transposed variants (two rows, one row
    // tile, short x chunks so the thin slabs fill the device), or all four
    // slabs on the general transposed kernel
    auto slab_rows = [&](const int (*b)[4], int n, int variant, hipStream_t st) {
        RowsArgs q = r;
        q.ichunk = s->adj_rchunk;
        q.tiles = s->adj_stiles * rows_waves(s->g.NZ) <= 16 ? s->adj_stiles : 1;
        q.npiece = n;
        std::memcpy(q.box, b, sizeof(q.box[0]) * n);
        return launch_step_rows(q, 2, variant, st);
    };
    const bool rs = s->adj.rows_slabs;
    hipStream_t sx = three ? s->stream2 : side, sy = three ? s->stream3 : side, sc = three ? s->stream4 : side;
    if (rs) {
        if (s->nxbox) HIPCHK(s, slab_rows(s->xbox, s->nxbox, 4, sx));
        if (s->nybox) HIPCHK(s, slab_rows(s->ybox, s->nybox, 5, sy));
    } else if (three) {
        if (s->nxbox) HIPCHK(s, launch_adj_step(av, s->adj_stj, s->adj_schunk, s->xbox, s->nxbox, sx));
        if (s->nybox) HIPCHK(s, launch_adj_step(av, s->adj_stj, s->adj_schunk, s->ybox, s->nybox, sy));
    } else {
        int slabs[4][4], ns = 0;
        for (int q = 0; q < s->nxbox; q++) std::memcpy(slabs[ns++], s->xbox[q], sizeof(slabs[0]));
        for (int q = 0; q < s->nybox; q++) std::memcpy(slabs[ns++], s->ybox[q], sizeof(slabs[0]));
        if (ns) HIPCHK(s, launch_adj_step(av, s->adj_stj, s->adj_schunk, slabs, ns, side));
    }
    if (s->ncorner) {
        if (rs && s->adj_corner_rows) {
            RowsArgs q = r;
            q.ichunk = s->adj_crchunk;
            q.tiles = 1;
            q.npiece = s->ncorner;
            std::memcpy(q.box, s->cbox4, sizeof(q.box[0]) * s->ncorner);
            HIPCHK(s, launch_step_rows(q, 2, 6, sc));
        } else {
            HIPCHK(s, launch_adj_step(av, s->adj_ctj, s->adj_cich, s->cbox4, s->ncorner, sc));
        }
    }
    r.npiece = 1;
    std::memcpy(r.box[0], s->cbox, sizeof(s->cbox));
    const int arows = s->adj_rows ? s->adj_rows : s->rows;
    r.tiles = arows == 2 && s->adj_tiles * rows_waves(s->g.NZ) <= 16 ? s->adj_tiles : 1;
    if (s->adj_cchunk) r.ichunk = s->adj_cchunk;
    HIPCHK(s, launch_step_rows(r, arows, 3, s->stream));
    AdjInject none{};
    const AdjInject &inj = s->nbox > 0 && next.on ? next : none;
    // early start of the next step's slabs: the join first, then the addition
    // (its cells, the box grown by one and the injection cell, lie off the
    // clean box's outer planes and rows, which the slabs read; the caller
    // allows it only then and only when the next injection rides on it)
    const bool early_out = sched == 3 && s->adj_early_allowed && s->nbox > 0;
    if (early_out) {
        HIPCHK(s, hipEventRecord(s->ev_join3, s->stream3));
        HIPCHK(s, hipEventRecord(s->ev_join4, s->stream4));
        HIPCHK(s, hipStreamWaitEvent(s->stream2, s->ev_join3, 0));
        HIPCHK(s, hipStreamWaitEvent(s->stream2, s->ev_join4, 0));
        HIPCHK(s, hipEventRecord(s->ev_join, s->stream2));
        HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_join, 0));
        HIPCHK(s, hipEventRecord(s->ev_early, s->stream));
        s->adj_early_ready = true;
        HIPCHK(s, launch_adj_box_add(av, inj, s->adj.corr, s->stream));
        *injected = inj.on != 0;
        return FDTD_OK;
    }
    if (s->nbox) {
        if (sched == 1) {
            HIPCHK(s, hipEventRecord(s->ev_line, s->stream));
            HIPCHK(s, hipStreamWaitEvent(s->stream4, s->ev_line, 0));
            HIPCHK(s, launch_adj_box_fix(av, inj, s->stream4));
        } else if (sched == 3) {
            HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_line, 0));
            HIPCHK(s, launch_adj_box_add(av, inj, s->adj.corr, s->stream));
        } else {
            if (sched == 2) HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_line, 0));
            HIPCHK(s, launch_adj_box_fix(av, inj, s->stream));
        }
        *injected = inj.on != 0;
    }
    // the y-slab and corner streams join through the x-slab stream: one wait on the main stream
    if (three) {
        HIPCHK(s, hipEventRecord(s->ev_join3, s->stream3));
        HIPCHK(s, hipEventRecord(s->ev_join4, s->stream4));
        HIPCHK(s, hipStreamWaitEvent(s->stream2, s->ev_join3, 0));
        HIPCHK(s, hipStreamWaitEvent(s->stream2, s->ev_join4, 0));
    }
    if (fork) {
        HIPCHK(s, hipEventRecord(s->ev_join, s->stream2));
        HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_join, 0));
    }
    return FDTD_OK;
}

int adjoint_run(fdtd_sim *s, AdjScratch &W, int nsteps, int ckpt, const int r0[3], int nf, const double *freq,
                double *J, double *grad) {
    const Setup &S = s->S;
    int rc;
    if ((rc = adj_alloc(s))) return rc;
    const int t0 = s->t;
    // ckpt <= 0: the box trajectory of every step is kept (one forward pass);
    // else checkpoints every ckpt steps and the trajectory of one segment at a time
    const bool store = ckpt <= 0;
    const int seglen = store ? nsteps : ckpt;
    const int nseg = store ? 1 : (nsteps + ckpt - 1) / ckpt;
    const long nbox = s->nbox;
    if (hipMalloc((void **)&W.probe, (size_t)3 * nsteps * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&W.traj, (size_t)(seglen + 2) * 3 * std::max(nbox, 1L) * sizeof(float)) != hipSuccess)
        return set_err(s, FDTD_ERR_ALLOC, "adjoint: scratch allocation failed");
    float *const probe = W.probe, *const traj = W.traj;
    std::vector<Checkpoint> &cks = W.cks;
    cks.resize(store ? 0 : nseg + 1);
    auto record = [&](int st) {
        for (int c = 0; c < 3; c++) {
            s->traj_step[c] = traj + ((size_t)(st + 2) * 3 + c) * nbox;
            s->traj_first[0][c] = st == 0 ? traj + ((size_t)0 * 3 + c) * nbox : nullptr;
            s->traj_first[1][c] = st == 0 ? traj + ((size_t)1 * 3 + c) * nbox : nullptr;
        }
    };
    auto unrecord = [&]() {
        for (int c = 0; c < 3; c++) s->traj_step[c] = s->traj_first[0][c] = s->traj_first[1][c] = nullptr;
    };
    // forward with checkpoints (or the full trajectory) and the objective probe
    s->probe = probe;
    s->probe_p = gidx(s->g, r0[0], r0[1], r0[2]);
    s->probe_t0 = t0;
    // the recorded loops step one at a time (record() only moves host pointers),
    // keeping the early start of the side pieces across the steps of a loop
    if (store) {
        s->early_ready = false;
        for (int st = 0; st < nsteps; st++) {
            record(st);
            if ((rc = step_steps(s, 1))) return rc;
        }
        unrecord();
    } else {
        for (int seg = 0; seg < nseg; seg++) {
            const int len = std::min(ckpt, nsteps - seg * ckpt);
            if (seg + 1 < nseg) {
                if ((rc = ckpt_copy(s, cks[seg], true))) return rc;
                if ((rc = fdtd_step(s, len))) return rc;
            } else {
                // the last segment records its box trajectory on the way, so the
                // reverse sweep starts on it without a checkpoint or a recompute
                s->early_ready = false;
                for (int st = 0; st < len; st++) {
                    record(st);
                    if ((rc = step_steps(s, 1))) return rc;
                }
                unrecord();
            }
        }
    }
    s->probe = nullptr;
    if (!store && (rc = ckpt_copy(s, cks[nseg], true))) return rc;
    std::vector<float> pr((size_t)3 * nsteps);
    if (hipStreamSynchronize(s->stream) != hipSuccess ||
        hipMemcpy(pr.data(), probe, pr.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return set_err(s, FDTD_ERR_DEVICE, "adjoint: probe copy failed");
    // objective in FP64
    std::vector<std::complex<double>> gsum((size_t)nf * 3, 0.0);
    for (int t = t0; t < t0 + nsteps; t++)
        for (int f = 0; f < nf; f++) {
            const std::complex<double> w = std::exp(std::complex<double>(0, 2 * fdtd::kPI * freq[f] * (t + 1.0) * S.dt));
            for (int c = 0; c < 3; c++) gsum[(size_t)f * 3 + c] += (double)pr[(size_t)3 * (t - t0) + c] * w;
        }
    double Jv = 0;
    for (auto &v : gsum) Jv += std::norm(v);
    *J = Jv;
    // reverse sweep
    if ((rc = adj_zero(s))) return rc;
    s->sampling = false;
    int pa = 0, pe = 0, ph = 0;
    const long p0 = gidx(s->g, r0[0], r0[1], r0[2]);
    // dJ/dE^{t+1}(r0) as the reverse sweep adds it before the transposed step of t
    auto inject_at = [&](int t) {
        AdjInject v{};
        double a[3] = {0, 0, 0};
        for (int f = 0; f < nf; f++) {
            const std::complex<double> w = std::exp(std::complex<double>(0, 2 * fdtd::kPI * freq[f] * (t + 1.0) * S.dt));
            for (int c = 0; c < 3; c++) a[c] += 2.0 * std::real(std::conj(gsum[(size_t)f * 3 + c]) * w);
        }
        v.p0 = p0;
        for (int c = 0; c < 3; c++) v.a[c] = (float)a[c];
        v.on = 1;
        v.inside = 1;
        for (int x = 0; x < 3; x++)
            if (r0[x] < s->box_lo[x] - 1 || r0[x] >= s->box_lo[x] + s->box_n[x] + 1) v.inside = 0;
        return v;
    };
    // the injection of step t-1 rides on the dispersive-cell correction of step t
    // where the objective cell lies in the clean box (written there by the clean box alone)
    const bool fuse = s->adj.rows && s->nbox > 0 && s->adj_fuse_inject && r0[0] >= s->cbox[0] &&
                      r0[0] < s->cbox[1] && r0[1] >= s->cbox[2] && r0[1] < s->cbox[3];
    // the next step's slabs may start before this step's correction is added
    // when every injection rides on it and its cell is off the clean box's
    // outer planes and rows (which the slabs read)
    s->adj_early_allowed = fuse && s->adj_early && r0[0] > s->cbox[0] && r0[0] < s->cbox[1] - 1 &&
                           r0[1] > s->cbox[2] && r0[1] < s->cbox[3] - 1;
    s->adj_early_ready = false;
    bool injected = false;
    for (int seg = nseg - 1; seg >= 0; seg--) {
        int seg_t0 = t0, len = nsteps;
        if (!store) {
            seg_t0 = t0 + seg * ckpt;
            len = std::min(ckpt, nsteps - seg * ckpt);
            if (seg + 1 < nseg) {
                // restore the segment's checkpoint and recompute its box trajectory
                Checkpoint &ck = cks[seg];
                s->t = ck.t; s->p = ck.p; s->pp = ck.pp; s->le = ck.le;
                if ((rc = ckpt_copy(s, ck, false))) return rc;
                s->early_ready = false;
                for (int st = 0; st < len; st++) {
                    record(st);
                    if ((rc = step_steps(s, 1))) return rc;
                }
                unrecord();
                s->early_ready = false;
                s->adj_early_ready = false;
            }
        }
        for (int st = len - 1; st >= 0; st--) {
            const int t = seg_t0 + st;
            if (!injected) {
                const AdjInject v = inject_at(t);
                HIPCHK(s, launch_adj_inject(s->adj.E[pa], p0, v.a[0], v.a[1], v.a[2], s->stream));
            }
            AdjDispDev ad{};
            ad.d = make_disp(s, s->p);
            for (int c = 0; c < 3; c++) {
                ad.Ehat[c] = s->adj.E[pa][c];
                for (int q = 0; q < 7; q++) ad.a[c][q] = s->adj.st[c][q];
                ad.Qb[c] = s->adj.Qb[c];
                ad.Eb[c] = s->adj.Eb[c];
                ad.dq[c] = s->adj.dq[c];
                ad.db[c] = s->adj.db[c];
                ad.tm1[c] = traj + ((size_t)st * 3 + c) * nbox;
                ad.t0[c] = traj + ((size_t)(st + 1) * 3 + c) * nbox;
                ad.tp1[c] = traj + ((size_t)(st + 2) * 3 + c) * nbox;
            }
            ad.grad = s->adj.grad;
            ad.Ceh = (float)S.Ceh;
            ad.Cee = (float)S.Cee;
            const AdjInject nx = fuse && t > t0 ? inject_at(t - 1) : AdjInject{};
            if ((rc = adjoint_step(s, pa, pe, ph, ad, nx, &injected))) return rc;
            pa ^= 1; pe ^= 1; ph ^= 1;
        }
    }
    // gradient on the design block
    std::vector<double> gb(nbox);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(gb.data(), s->adj.grad, nbox * sizeof(double), hipMemcpyDeviceToHost));
    size_t v = 0;
    for (int a = 0; a < s->des_n[0]; a++)
        for (int b = 0; b < s->des_n[1]; b++)
            for (int c = 0; c < s->des_n[2]; c++) {
                const int bi = s->des_lo[0] + a - s->box_lo[0], bj = s->des_lo[1] + b - s->box_lo[1], bk = s->des_lo[2] + c - s->box_lo[2];
                grad[v++] = gb[((size_t)bi * s->box_n[1] + bj) * s->box_n[2] + bk];
            }
    if (store) return FDTD_OK;   // the forward pass left the final state in place
    // leave the forward state after nsteps
    Checkpoint &fin = cks[nseg];
    s->t = fin.t; s->p = fin.p; s->pp = fin.pp; s->le = fin.le;
    return ckpt_copy(s, fin, false);
}

}  // namespace

int fdtd_get_materials(const fdtd_sim *s, int comp, uint8_t *out) {
    if (!s || !out || comp < 0 || comp > 3) return FDTD_ERR_INVALID;
    std::memcpy(out, s->S.mat[comp].data(), s->S.mat[comp].size());
    return FDTD_OK;
}

// Split-path arguments updating the current level in place.
static StepDev inplace_step(fdtd_sim *s, bool tfsf) {
    StepDev d = make_step(s, tfsf);
    for (int c = 0; c < 3; c++) {
        d.Ein[c] = s->E[s->p][c]; d.Hin[c] = s->H[s->p][c];
        d.Eout[c] = s->E[s->p][c]; d.Hout[c] = s->H[s->p][c];
    }
    for (int a = 0; a < 6; a++) { d.psiH_in.p[a] = d.psiH_out.p[a] = s->psiH[s->pp][a]; d.psiE.p[a] = s->psiE[a]; }
    return d;
}

int fdtd_update_h(fdtd_sim *s) {
    if (!s) return FDTD_ERR_INVALID;
    const StepDev d = inplace_step(s, false);
    HIPCHK(s, s->S.P.periodic_xy ? launch_periodic_h(d, s->stream) : launch_split_h(d, s->stream));
    return FDTD_OK;
}

int fdtd_tfsf(fdtd_sim *s) {
    if (!s) return FDTD_ERR_INVALID;
    LineDev l = make_line(s);
    HIPCHK(s, launch_line(l, s->stream));
    StepDev d = make_step(s, false);
    for (int c = 0; c < 3; c++) { d.Eout[c] = s->E[s->p][c]; d.Hout[c] = s->H[s->p][c]; }
    HIPCHK(s, s->S.P.periodic_xy ? launch_periodic_tfsf(d, s->stream) : launch_tfsf_split(d, s->stream));
    s->le ^= 1;
    return FDTD_OK;
}

int fdtd_update_e(fdtd_sim *s) {
    if (!s) return FDTD_ERR_INVALID;
    const StepDev d = inplace_step(s, false);
    if (s->nbox && s->S.nonlocal1) HIPCHK(s, launch_nonlocal_gd(make_disp(s, s->p), s->stream));
    HIPCHK(s, s->S.P.periodic_xy ? launch_periodic_e(d, s->stream) : launch_split_e(d, s->stream));
    if (s->nbox) {
        DispDev dd = make_disp(s, s->p);
        HIPCHK(s, launch_box(s, dd));
    }
    return FDTD_OK;
}

int fdtd_dft_accumulate(fdtd_sim *s) {
    if (!s) return FDTD_ERR_INVALID;
    int rc = FDTD_OK;
    if (s->mon && s->t - s->mon_t0 < s->mon_cap) {
        const float *E1[3] = {s->E[s->p][0], s->E[s->p][1], s->E[s->p][2]};
        HIPCHK(s, launch_probe(E1, s->mon_p, s->mon + 3 * (size_t)(s->t - s->mon_t0), s->stream));
    }
    if (sampling_step(s)) {
        StepDev d = make_step(s, false);
        for (int c = 0; c < 3; c++) { d.Ein[c] = s->E[s->p][c]; d.Hin[c] = s->H[s->p][c]; }
        d.sample_row = s->X + (size_t)s->slot * s->mtot;
        HIPCHK(s, launch_dft_sample_split(d, s->stream));
        rc = after_sample(s);
    }
    s->t++;
    return rc;
}

// One periodic step (Periodic_XY): line, E ghosts + H with the source
// plane, H ghosts + E with the source plane, dispersive box, monitors; the
// reference's order (main.cu:2724-2934) with its corrections folded into the
// updates.  Fields update in place.
static int periodic_step(fdtd_sim *s) {
    const LineDev l = make_line(s);
    HIPCHK(s, launch_line(l, s->stream));
    const StepDev d = inplace_step(s, true);
    HIPCHK(s, launch_periodic_h(d, s->stream));
    HIPCHK(s, launch_periodic_e(d, s->stream));
    s->le ^= 1;
    if (s->nbox) HIPCHK(s, launch_dispersive(make_disp(s, s->p), s->stream));
    return fdtd_dft_accumulate(s);
}

int fdtd_step(fdtd_sim *s, int nsteps) {
    if (!s || nsteps < 0) return FDTD_ERR_INVALID;
    // an early start reaches only the steps of this call (anything the caller
    // did on the main stream since the last one is ordered by the plain fork)
    s->early_ready = false;
    return fdtd_host::step_steps(s, nsteps);
}

// fdtd_step without clearing the early start: for loops of single steps that
// launch nothing else between them (the adjoint's recorded steps).
int fdtd_host::step_steps(fdtd_sim *s, int nsteps) {
    if (int rc = check_finite(s)) return rc;
    if (s->S.P.periodic_xy) {
        for (int n = 0; n < nsteps; n++)
            if (int rc = periodic_step(s)) return rc;
        return FDTD_OK;
    }
    for (int n = 0; n < nsteps; n++) {
        LineDev l = make_line(s);
        StepDev d = make_step(s, true);
        const int q = s->p, pp = s->pp;
        for (int c = 0; c < 3; c++) {
            d.Ein[c] = s->E[q][c]; d.Hin[c] = s->H[q][c];
            d.Eout[c] = s->E[1 - q][c]; d.Hout[c] = s->H[1 - q][c];
        }
        for (int a = 0; a < 6; a++) { d.psiH_in.p[a] = s->psiH[pp][a]; d.psiH_out.p[a] = s->psiH[1 - pp][a]; d.psiE.p[a] = s->psiE[a]; }
        const bool samp = s->sampling && sampling_step(s);
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        const bool step_timing = s->timing && s->timing_what != 3;
        if (step_timing) {
            ev0 = next_event(s, 0);
            ev1 = next_event(s, 1);
            if (!ev0 || !ev1) return set_err(s, FDTD_ERR_DEVICE, "hipEventCreate failed");
        }
        const bool only_clean = step_timing && s->timing_what == 2 && s->use_rows;
        DispDev dd{};
        if (s->nbox) dd = make_disp(s, 1 - q);
        // NONLOCAL == 1: grad(div P_d^n) from the box state of the last step, ahead of every
        // launch that can update the box
        if (s->nbox && s->S.nonlocal1) HIPCHK(s, launch_nonlocal_gd(dd, s->stream));
        bool disp_done = false, samp_done = false;
        if (s->use_rows) {
            d.tfsf_on = 0;
            if (samp) d.sample_row = s->X + (size_t)s->slot * s->mtot;
            if (step_timing && !only_clean) HIPCHK(s, hipEventRecord(ev0, s->stream));
            int rc2 = launch_rows_update(s, d, l, s->nbox ? &dd : nullptr, samp, &disp_done, &samp_done,
                                         only_clean ? ev0 : nullptr, only_clean ? ev1 : nullptr);
            if (rc2) return rc2;
            if (step_timing && !only_clean) HIPCHK(s, hipEventRecord(ev1, s->stream));
        } else {
            HIPCHK(s, launch_line(l, s->stream));
            const int full[1][4] = {{0, s->g.NX, 0, s->g.NY}};
            if (step_timing) HIPCHK(s, hipEventRecord(ev0, s->stream));
            HIPCHK(s, launch_fused_step(d, s->tj, s->ichunk, full, 1, s->stream));
            if (step_timing) HIPCHK(s, hipEventRecord(ev1, s->stream));
        }
        if (step_timing) s->ev_used++;
        if (s->nbox && !disp_done) HIPCHK(s, launch_box(s, dd));
        if (s->probe) {
            const float *E1[3] = {s->E[1 - q][0], s->E[1 - q][1], s->E[1 - q][2]};
            HIPCHK(s, launch_probe(E1, s->probe_p, s->probe + 3 * (size_t)(s->t - s->probe_t0), s->stream));
        }
        if (s->mon && s->t - s->mon_t0 < s->mon_cap) {
            const float *E1[3] = {s->E[1 - q][0], s->E[1 - q][1], s->E[1 - q][2]};
            HIPCHK(s, launch_probe(E1, s->mon_p, s->mon + 3 * (size_t)(s->t - s->mon_t0), s->stream));
        }
        s->p ^= 1;
        s->pp ^= 1;
        s->le ^= 1;
        if (samp) {
            // tangential E^{n+1} and H^{n+1/2} on the monitor faces (ScattAbs, main.cu:2920), unless the
            // post-update launch took them
            if (!samp_done) {
                StepDev ds = make_step(s, false);
                for (int c = 0; c < 3; c++) { ds.Ein[c] = s->E[s->p][c]; ds.Hin[c] = s->H[s->p][c]; }
                ds.sample_row = s->X + (size_t)s->slot * s->mtot;
                HIPCHK(s, launch_dft_sample_split(ds, s->stream));
            }
            int rc = after_sample(s);
            if (rc) return rc;
        }
        s->t++;
    }
    return FDTD_OK;
}

int fdtd_set_point_monitor(fdtd_sim *s, const int r[3], int capacity) {
    if (!s || !r || capacity < 1) return FDTD_ERR_INVALID;
    for (int a = 0; a < 3; a++)
        if (r[a] < 0 || r[a] >= s->S.N[a]) return set_err(s, FDTD_ERR_INVALID, "point monitor outside the grid");
    if (capacity > s->mon_cap_alloc) {
        // the old buffer goes once the launches queued into it have run
        HIPCHK(s, hipStreamSynchronize(s->stream));
        dfree(s, s->mon);
        s->mon_cap_alloc = 0;
        float *m = nullptr;
        if (int rc = dalloc(s, &m, (size_t)3 * capacity)) return rc;
        s->mon = m;
        s->mon_cap_alloc = capacity;
    }
    s->mon_cap = capacity;
    s->mon_p = gidx(s->g, r[0], r[1], r[2]);
    s->mon_t0 = s->t;
    return FDTD_OK;
}

int fdtd_point_spectrum(fdtd_sim *s, int nf, const double *freq, double *e2, double *amp) {
    if (!s || nf < 1 || !freq || (!e2 && !amp)) return FDTD_ERR_INVALID;
    if (!s->mon) return set_err(s, FDTD_ERR_INVALID, "no point monitor set");
    const int n = std::min(s->t - s->mon_t0, s->mon_cap);
    std::vector<float> e((size_t)3 * std::max(n, 1));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (int rc = check_finite(s)) return rc;
    if (n > 0) HIPCHK(s, hipMemcpy(e.data(), s->mon, (size_t)3 * n * sizeof(float), hipMemcpyDeviceToHost));
    const double dt = s->S.dt;
    for (int f = 0; f < nf; f++) {
        std::complex<double> g[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < n; q++) {
            const std::complex<double> w = std::exp(std::complex<double>(0, 2 * fdtd::kPI * freq[f] * (s->mon_t0 + q + 1.0) * dt));
            for (int c = 0; c < 3; c++) g[c] += (double)e[(size_t)3 * q + c] * w;
        }
        if (e2) e2[f] = std::norm(g[0]) + std::norm(g[1]) + std::norm(g[2]);
        if (amp)
            for (int c = 0; c < 3; c++) { amp[(f * 3 + c) * 2] = g[c].real(); amp[(f * 3 + c) * 2 + 1] = g[c].imag(); }
    }
    return FDTD_OK;
}

int fdtd_sync(fdtd_sim *s) {
    if (!s) return FDTD_ERR_INVALID;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return check_finite(s);
}

int fdtd_get_field(fdtd_sim *s, int comp, float *out) {
    if (!s || !out || comp < 0 || comp > 5) return FDTD_ERR_INVALID;
    const float *f = comp < 3 ? s->E[s->p][comp] : s->H[s->p][comp - 3];
    HIPCHK(s, launch_pack_field(f, s->g, s->dense, 1, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(out, s->dense, (size_t)s->g.NX * s->g.NY * s->g.NZ * sizeof(float), hipMemcpyDeviceToHost));
    return FDTD_OK;
}

int fdtd_get_slice(fdtd_sim *s, int comp, int axis, int index, float *out) {
    if (!s || !out || comp < 0 || comp > 9 || axis < 0 || axis > 2) return FDTD_ERR_INVALID;
    const Geom &g = s->g;
    const int n[3] = {g.NX, g.NY, g.NZ};
    if (index < 0 || index >= n[axis]) return set_err(s, FDTD_ERR_INVALID, "fdtd_get_slice: index outside the grid");
    if (comp >= 6) {
        // state of the dispersive box, zero outside it.  6: hydrodynamic
        // density perturbation N_EQ (n~x + n~y + n~z) / 3 at time level n+1 (the
        // level the last step left in the n~^n buffer), as the ND snapshots
        // average the three component densities (snapshot.cu:80-133).  7-9: the
        // Drude pole of ex, ey, ez at level n+1 in the reference's units (P_d,
        // the Px_d/Py_d/Pz_d arrays the Snap_in == 2 snapshots copy,
        // main.cu:2962-2964): eps0 P~ of the local model, the electron velocity
        // of the hydrodynamic one
        if (comp == 6 && !s->S.hydro) return set_err(s, FDTD_ERR_INVALID, "fdtd_get_slice: no hydrodynamic model");
        const size_t nb = (size_t)std::max(s->nbox, 0L);
        std::vector<float> st;
        if (nb) {
            st.resize(3 * nb);
            HIPCHK(s, hipStreamSynchronize(s->stream));
            if (comp == 6)
                for (int c = 0; c < 3; c++) HIPCHK(s, hipMemcpy(st.data() + c * nb, s->h_ndp[c], nb * sizeof(float), hipMemcpyDeviceToHost));
            else
                HIPCHK(s, hipMemcpy(st.data(), s->disp_st[comp - 7][2], nb * sizeof(float), hipMemcpyDeviceToHost));
        }
        const double scale = comp == 6 ? s->S.ade.N_EQ / 3.0 : (s->S.hydro ? 1.0 : fdtd::kEP0);
        const int ua = axis == 0 ? 1 : 0, ub = axis == 2 ? 1 : 2;   // plane axes: out[u * n[ub] + v]
        for (int u = 0; u < n[ua]; u++)
            for (int v = 0; v < n[ub]; v++) {
                int x[3];
                x[axis] = index; x[ua] = u; x[ub] = v;
                float val = 0.f;
                bool in = nb > 0;
                for (int a = 0; a < 3; a++) in = in && x[a] >= s->box_lo[a] && x[a] < s->box_lo[a] + s->box_n[a];
                if (in) {
                    const size_t q = ((size_t)(x[0] - s->box_lo[0]) * s->box_n[1] + (x[1] - s->box_lo[1])) * s->box_n[2] +
                                     (x[2] - s->box_lo[2]);
                    const double sum = comp == 6 ? (double)st[q] + st[nb + q] + st[2 * nb + q] : (double)st[q];
                    val = (float)(scale * sum);
                }
                out[(size_t)u * n[ub] + v] = val;
            }
        return FDTD_OK;
    }
    const float *f = comp < 3 ? s->E[s->p][comp] : s->H[s->p][comp - 3];
    const size_t fl = sizeof(float);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (axis == 0)   // rows j of NZ values, pitch sz
        HIPCHK(s, hipMemcpy2D(out, g.NZ * fl, f + (size_t)index * g.sxy, g.sz * fl, g.NZ * fl, g.NY, hipMemcpyDeviceToHost));
    else if (axis == 1)   // rows i of NZ values, pitch sxy
        HIPCHK(s, hipMemcpy2D(out, g.NZ * fl, f + (size_t)index * g.sz, g.sxy * fl, g.NZ * fl, g.NX, hipMemcpyDeviceToHost));
    else   // per i, NY single values with pitch sz
        for (int i = 0; i < g.NX; i++)
            HIPCHK(s, hipMemcpy2D(out + (size_t)i * g.NY, fl, f + (size_t)i * g.sxy + index, g.sz * fl, fl, g.NY,
                                  hipMemcpyDeviceToHost));
    return FDTD_OK;
}

int fdtd_set_field(fdtd_sim *s, int comp, const float *in) {
    if (!s || !in || comp < 0 || comp > 5) return FDTD_ERR_INVALID;
    float *f = comp < 3 ? s->E[s->p][comp] : s->H[s->p][comp - 3];
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpyAsync(s->dense, in, (size_t)s->g.NX * s->g.NY * s->g.NZ * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, launch_pack_field(f, s->g, s->dense, 0, s->stream));
    if (comp < 3 && s->nbox) {
        DispDev d = make_disp(s, s->p);
        HIPCHK(s, launch_disp_sync_state(d, 1, s->stream));
    }
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

int fdtd_get_line(fdtd_sim *s, double *e, double *h) {
    if (!s) return FDTD_ERR_INVALID;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (e) HIPCHK(s, hipMemcpy(e, s->line_e[s->le], s->S.L * sizeof(double), hipMemcpyDeviceToHost));
    if (h) HIPCHK(s, hipMemcpy(h, s->line_h, s->S.L * sizeof(double), hipMemcpyDeviceToHost));
    return FDTD_OK;
}

int fdtd_get_spectrum_setup(const fdtd_sim *s, double *freq, double *einc) {
    if (!s) return FDTD_ERR_INVALID;
    for (int m = 0; m < s->S.P.NUM_freq; m++) {
        if (freq) freq[m] = s->S.freq[m];
        if (einc) { einc[2 * m] = s->S.Einc[m].real(); einc[2 * m + 1] = s->S.Einc[m].imag(); }
    }
    return FDTD_OK;
}

int fdtd_get_dft(fdtd_sim *s, int face, int comp, double *out) {
    if (!s || !out || face < 0 || face > 11 || comp < 0 || comp > 3) return FDTD_ERR_INVALID;
    int rc = contract(s);
    if (rc) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if ((rc = check_finite(s))) return rc;
    const FaceDev &f = s->fd;
    const long blk = (long)f.na[face] * f.nb[face];
    const long col = (comp < 2 ? 0 : f.mE_pad) + f.off[face] + (comp % 2) * blk;
    const int F = s->S.P.NUM_freq;
    std::vector<double> re(blk), im(blk);
    for (int m = 0; m < F; m++) {
        HIPCHK(s, hipMemcpy(re.data(), s->acc + (size_t)m * s->mtot + col, blk * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(s, hipMemcpy(im.data(), s->acc + (size_t)(F + m) * s->mtot + col, blk * sizeof(double), hipMemcpyDeviceToHost));
        for (long c = 0; c < blk; c++) {
            out[((size_t)m * blk + c) * 2] = re[c];
            out[((size_t)m * blk + c) * 2 + 1] = im[c];
        }
    }
    return FDTD_OK;
}

int fdtd_flux(fdtd_sim *s, double *sca, double *abs_, double *ext, double *raw) {
    if (!s) return FDTD_ERR_INVALID;
    int rc = contract(s);
    if (rc) return rc;
    const Setup &S = s->S;
    const int F = S.P.NUM_freq;
    HIPCHK(s, launch_flux_reduce(s->acc, s->fd, s->mtot, s->fd.mE_pad, F, S.P.dx * S.P.dy, s->flux_d, s->stream));
    std::vector<double> Q(12 * F);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(Q.data(), s->flux_d, Q.size() * sizeof(double), hipMemcpyDeviceToHost));
    if ((rc = check_finite(s))) return rc;
    for (double v : Q)
        if (!std::isfinite(v)) return set_err(s, FDTD_ERR_NONFINITE, "non-finite face flux (the monitor spectra diverged)");
    flux_combine(s, Q.data(), sca, abs_, ext);
    if (raw) std::memcpy(raw, Q.data(), Q.size() * sizeof(double));
    return FDTD_OK;
}

void fdtd_host::flux_combine(const fdtd_sim *s, double *Q, double *sca, double *abs_, double *ext) {
    const Setup &S = s->S;
    const int F = S.P.NUM_freq;
    // normalise by |E_incident|^2 and combine the faces (Spectral.cu:1162-1168)
    for (int m = 0; m < F; m++) {
        const double a = std::abs(S.Einc[m]);
        for (int q = 0; q < 12; q++) Q[q * F + m] /= a * a;
    }
    const double area = S.P.xs_area > 0 ? S.P.xs_area : fdtd::kPI * S.P.nano_sphere_radius * S.P.nano_sphere_radius;
    for (int m = 0; m < F; m++) {
        double sc = 0, ab = 0;
        for (int ax = 0; ax < 3; ax++) {
            sc += -Q[(ax * 2 + 0) * F + m] + Q[(ax * 2 + 1) * F + m];
            ab += -Q[(6 + ax * 2 + 0) * F + m] + Q[(6 + ax * 2 + 1) * F + m];
        }
        sc /= area;
        ab = -ab / area;
        if (sca) sca[m] = sc;
        if (abs_) abs_[m] = ab;
        if (ext) ext[m] = sc + ab;
    }
}

int fdtd_kernel_timing(fdtd_sim *s, int enable, double *total_ms, long *launches) {
    if (!s) return FDTD_ERR_INVALID;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    double tot = 0;
    for (size_t i = 0; i < s->ev_used; i++) {
        float ms = 0;
        HIPCHK(s, hipEventElapsedTime(&ms, s->events[i].first, s->events[i].second));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = (long)s->ev_used;
    s->ev_used = 0;
    if (enable < 0 || enable > 3) return set_err(s, FDTD_ERR_INVALID, "kernel_timing: enable must be 0, 1, 2 or 3");
    s->timing = enable != 0;
    if (enable) s->timing_what = enable;
    return FDTD_OK;
}


int fdtd_reflectance(fdtd_sim *s, double *R, double *T, double *ref_r, double *ref_t) {
    if (!s) return FDTD_ERR_INVALID;
    const Setup &S = s->S;
    if (!S.P.periodic_xy) return set_err(s, FDTD_ERR_INVALID, "fdtd_reflectance needs periodic_xy");
    const int F = S.P.NUM_freq;
    std::vector<double> raw(12 * (size_t)F);
    if (int rc = fdtd_flux(s, nullptr, nullptr, nullptr, raw.data())) return rc;
    // the incident wave at the source plane, sampled like the monitors: the same
    // FP64 line replayed on the host for the steps run so far
    const std::vector<std::complex<double>> inc = fdtd::line_dft(S, s->t);
    const double area = S.N[0] * S.P.dx * S.N[1] * S.P.dy;
    for (int m = 0; m < F; m++) {
        const double a2 = std::norm(S.Einc[m]);
        // faces 4, 5: S_z = sum Re(Ex Hy* - Ey Hx*) dA, stored over |E_incident|^2
        const double sr = raw[4 * (size_t)F + m] * a2, st = raw[5 * (size_t)F + m] * a2;
        const double pin = std::norm(inc[m]) * area;
        if (R) R[m] = pin > 0 ? -sr / pin : 0.0;
        if (T) T[m] = pin > 0 ? st / pin : 0.0;
        if (ref_r) ref_r[m] = -sr / area / a2;
        if (ref_t) ref_t[m] = -st / area / a2;
    }
    return FDTD_OK;
}

int fdtd_add_layer(fdtd_sim *s, int k0, int k1) {
    if (!s || k0 >= k1) return FDTD_ERR_INVALID;
    Setup &S = s->S;
    const int NX = S.N[0], NY = S.N[1], NZ = S.N[2];
    if (k0 < 0 || k1 > NZ) return set_err(s, FDTD_ERR_INVALID, "fdtd_add_layer: layer outside the grid");
    std::vector<uint8_t> m[4] = {S.mat[0], S.mat[1], S.mat[2], S.mat[3]};
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
            for (int k = k0; k < k1; k++)
                for (int c = 0; c < 4; c++) m[c][((size_t)i * NY + j) * NZ + k] = (uint8_t)S.P.material;
    const int rc = fdtd_set_materials(s, m[1].data(), m[2].data(), m[3].data());
    if (rc) return rc;
    S.mat[0] = m[0];
    return FDTD_OK;
}
