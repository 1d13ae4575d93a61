// HIP kernels of the FDTD hot path for gfx950 (MI355X).
//
// Fused fast path, one launch per time step:
//   k_fused_step   H update + CPML psi_H + TFSF + E update + CPML psi_E +
//                  DFT face sampling, marching along x through a (j,k) tile.
//                  E and H are ping-ponged so every field is read once from
//                  buffer A and written once to buffer B (48 B per vacuum cell).
//   k_dispersive   Drude + 2 CP ADE on the dense box holding the metal; runs
//                  after k_fused_step and overwrites E at dispersive cells.
//   k_line         1D incident line (h_inc, e_inc, Mur, source) on device.
//   k_dft_*        twiddle build, buffered DFT contraction, flux reduction.
// Split path (reference call structure): k_split_h, k_tfsf_split, k_split_e,
// k_dft_sample_split, sharing the per-cell device functions below.
//
// Update equations follow the reference kernels (main.cu:762-1044 H,
// main.cu:1048-1912 E, main.cu:369-758 TFSF, main.cu:2280-2337 line,
// main.cu:35-366 DFT); see each function.

#include <hip/hip_runtime.h>
#include "internal.h"

namespace {

constexpr double kPI = 3.1415926535897932384626433832795;

// ----------------------------------------------------------------- CPML index helpers

// H slabs: 2*Nc layers (near [0,Nc), far Nc+k2); E slabs: 2*(Nc+1) layers (near [0,Nc], far Nc+1+k2)
__device__ inline long zslab(const Geom &g, int layers, int i, int j, int kk) {
    return ((long)i * g.NY + j) * layers + kk;
}
__device__ inline long yslab(const Geom &g, int layers, int i, int jj, int k) {
    return ((long)i * layers + jj) * g.NZ + k;
}
__device__ inline long xslab(const Geom &g, int ii, int j, int k) {
    return ((long)ii * g.NY + j) * g.NZ + k;
}

__device__ inline float psi_step(const float *in, float *out, long idx, float b, float c, float dE, bool wr) {
    float p = b * in[idx] + c * dE;
    if (wr) out[idx] = p;
    return p;
}

// ----------------------------------------------------------------- TFSF (main.cu:369-758)

// Linear interpolation of the 1D line at projected distance d from the TFSF
// corner: e_inc at integer offsets (main.cu:387-391), h_inc half-cell shifted
// (main.cu:420-424).
__device__ inline float inc_interp(const TfsfDev &t, bool hline, float di, float dj, float dk) {
    float d = (di * t.s[0] + dj * t.s[1] + dk * t.s[2]) * t.dscale;
    if (hline) d += 0.5f;
    const int id = (int)d;
    const float w = d - (float)id;
    const double *line = hline ? t.h_inc + (t.m0 - 1) : t.e_inc + t.m0;
    return (1.0f - w) * (float)line[id] + w * (float)line[id + 1];
}

// Sum of the TFSF corrections to field component c (E if efield, else H) at
// (i,j,k), accumulated in the reference's CORRECT_X, _Y, _Z order.
//   H component c on face normal a uses e_inc component b = 3-a-c; it sits at
//   plane t0[a]-1 (near) or t1[a] (far), spans [t0,t1] along c and [t0,t1)
//   along b; sign +/- (near/far) when a follows c cyclically, else -/+.
//   E component c on face normal a uses h_inc component b; it sits on planes
//   t0[a] and t1[a], spans [t0,t1) along c and [t0,t1] along b; sign -/+ when a
//   follows c cyclically, else +/-.
// Applied to H after the H update and to E^n before the E update, as the
// reference's CORRECT kernels run between UPDATE_h* and UPDATE_e*.
__device__ inline int pick3(int a, int x, int y, int z) { return a == 0 ? x : (a == 1 ? y : z); }
__device__ inline float pick3f(int a, float x, float y, float z) { return a == 0 ? x : (a == 1 ? y : z); }

// One face term: H component c (or E if efield) on the face normal to axis a.
// Indices are picked with selects, not local arrays, so a runtime (c, a)
// never puts the cell on the stack.
__device__ inline float tfsf_term(const TfsfDev &t, bool efield, int c, int a, int i, int j, int k) {
    const int b = 3 - a - c;
    const int xa = pick3(a, i, j, k), xb = pick3(b, i, j, k), xc = pick3(c, i, j, k);
    const int t0a = pick3(a, t.t0[0], t.t0[1], t.t0[2]), t1a = pick3(a, t.t1[0], t.t1[1], t.t1[2]);
    const int t0b = pick3(b, t.t0[0], t.t0[1], t.t0[2]), t1b = pick3(b, t.t1[0], t.t1[1], t.t1[2]);
    const int t0c = pick3(c, t.t0[0], t.t0[1], t.t0[2]), t1c = pick3(c, t.t1[0], t.t1[1], t.t1[2]);
    const int near_plane = efield ? t0a : t0a - 1;
    const bool nr = xa == near_plane, fr = xa == t1a;
    if (!nr && !fr) return 0.f;
    const bool c_ok = efield ? (xc >= t0c && xc < t1c) : (xc >= t0c && xc <= t1c);
    const bool b_ok = efield ? (xb >= t0b && xb <= t1b) : (xb >= t0b && xb < t1b);
    if (!c_ok || !b_ok) return 0.f;
    const float ddc = (float)(xc - t0c) + (efield ? 0.5f : 0.f);
    const float ddb = (float)(xb - t0b) + (efield ? 0.f : 0.5f);
    const float span = (float)(t1a - t0a);
    const float dda = efield ? (nr ? -0.5f : span + 0.5f) : (nr ? 0.f : span);
    const float d0 = a == 0 ? dda : (b == 0 ? ddb : ddc);
    const float d1 = a == 1 ? dda : (b == 1 ? ddb : ddc);
    const float d2 = a == 2 ? dda : (b == 2 ? ddb : ddc);
    const bool cyc = a == (c + 1) % 3;
    float sg = (cyc != efield) ? 1.f : -1.f;
    if (fr) sg = -sg;
    const float pol = efield ? pick3f(b, t.hf[0], t.hf[1], t.hf[2]) : pick3f(b, t.ef[0], t.ef[1], t.ef[2]);
    const float v = inc_interp(t, efield, d0, d1, d2) * pol;
    const float term = sg * (efield ? t.cE : t.cH) * v * pick3f(a, t.inv_d[0], t.inv_d[1], t.inv_d[2]);
    return term;
}

__device__ inline float tfsf_corr(const TfsfDev &t, bool efield, int c, int i, int j, int k) {
    float r = 0.f;
#pragma unroll 1
    for (int a = 0; a < 3; a++) {
        if (a == c) continue;
        r += tfsf_term(t, efield, c, a, i, j, k);
    }
    return r;
}

// Corrections of all three components of (i,j,k) from the faces flagged per
// normal axis (wave-uniform flags); per component the faces add in ascending
// axis order, as in tfsf_corr.  One call site of tfsf_term keeps it cheap.
__device__ inline void tfsf_faces(const TfsfDev &t, bool efield, bool fx, bool fy, bool fz, int i, int j, int k,
                                  float &rx, float &ry, float &rz) {
    rx = ry = rz = 0.f;
#pragma unroll 1
    for (int n = 0; n < 6; n++) {
        const int a = n >> 1;
        if (!(a == 0 ? fx : (a == 1 ? fy : fz))) continue;
        const int c = (n & 1) ? (a + 2) % 3 : (a + 1) % 3;
        const float v = tfsf_term(t, efield, c, a, i, j, k);
        rx += c == 0 ? v : 0.f;
        ry += c == 1 ? v : 0.f;
        rz += c == 2 ? v : 0.f;
    }
}

__device__ inline float tfsf_h(const TfsfDev &t, int c, int i, int j, int k) {
    if (i < t.t0[0] - 1 || i > t.t1[0] || j < t.t0[1] - 1 || j > t.t1[1] || k < t.t0[2] - 1 || k > t.t1[2]) return 0.f;
    return tfsf_corr(t, false, c, i, j, k);
}
__device__ inline float tfsf_e(const TfsfDev &t, int c, int i, int j, int k) {
    if (i < t.t0[0] || i > t.t1[0] || j < t.t0[1] || j > t.t1[1] || k < t.t0[2] || k > t.t1[2]) return 0.f;
    return tfsf_corr(t, true, c, i, j, k);
}

// tfsf_term, tfsf_h and tfsf_e with the component and face axis fixed at
// compile time (the same expressions and summation order, so the same
// values): the index picks fold away, which the fix-up kernel, evaluating
// twelve of these per cell, needs; the fused kernel keeps the runtime form
// for code size.
template <bool EF, int c, int a>
__device__ __forceinline__ float tfsf_term_ct(const TfsfDev &t, int i, int j, int k) {
    constexpr int b = 3 - a - c;
    const int x[3] = {i, j, k};
    const int xa = x[a], xb = x[b], xc = x[c];
    const int t0a = t.t0[a], t1a = t.t1[a], t0b = t.t0[b], t1b = t.t1[b], t0c = t.t0[c], t1c = t.t1[c];
    const int near_plane = EF ? t0a : t0a - 1;
    const bool nr = xa == near_plane, fr = xa == t1a;
    if (!nr && !fr) return 0.f;
    const bool c_ok = EF ? (xc >= t0c && xc < t1c) : (xc >= t0c && xc <= t1c);
    const bool b_ok = EF ? (xb >= t0b && xb <= t1b) : (xb >= t0b && xb < t1b);
    if (!c_ok || !b_ok) return 0.f;
    const float ddc = (float)(xc - t0c) + (EF ? 0.5f : 0.f);
    const float ddb = (float)(xb - t0b) + (EF ? 0.f : 0.5f);
    const float span = (float)(t1a - t0a);
    const float dda = EF ? (nr ? -0.5f : span + 0.5f) : (nr ? 0.f : span);
    const float d0 = a == 0 ? dda : (b == 0 ? ddb : ddc);
    const float d1 = a == 1 ? dda : (b == 1 ? ddb : ddc);
    const float d2 = a == 2 ? dda : (b == 2 ? ddb : ddc);
    constexpr bool cyc = a == (c + 1) % 3;
    float sg = (cyc != EF) ? 1.f : -1.f;
    if (fr) sg = -sg;
    const float pol = EF ? t.hf[b] : t.ef[b];
    const float v = inc_interp(t, EF, d0, d1, d2) * pol;
    return sg * (EF ? t.cE : t.cH) * v * t.inv_d[a];
}
template <bool EF, int c>
__device__ __forceinline__ float tfsf_corr_ct(const TfsfDev &t, int i, int j, int k) {
    float r = 0.f;
    if constexpr (c != 0) r += tfsf_term_ct<EF, c, 0>(t, i, j, k);
    if constexpr (c != 1) r += tfsf_term_ct<EF, c, 1>(t, i, j, k);
    if constexpr (c != 2) r += tfsf_term_ct<EF, c, 2>(t, i, j, k);
    return r;
}
template <int c>
__device__ __forceinline__ float tfsf_h_ct(const TfsfDev &t, int i, int j, int k) {
    if (i < t.t0[0] - 1 || i > t.t1[0] || j < t.t0[1] - 1 || j > t.t1[1] || k < t.t0[2] - 1 || k > t.t1[2]) return 0.f;
    return tfsf_corr_ct<false, c>(t, i, j, k);
}
template <int c>
__device__ __forceinline__ float tfsf_e_ct(const TfsfDev &t, int i, int j, int k) {
    if (i < t.t0[0] || i > t.t1[0] || j < t.t0[1] || j > t.t1[1] || k < t.t0[2] || k > t.t1[2]) return 0.f;
    return tfsf_corr_ct<true, c>(t, i, j, k);
}

// ----------------------------------------------------------------- per-cell H update (main.cu:762-1044)

// Neighbour values of E^n needed by the H update at (i,j,k).
struct ENbr {
    float ex, ey, ez;        // (i,j,k)
    float ex_j1, ez_j1;      // (i,j+1,k)
    float ex_k1, ey_k1;      // (i,j,k+1)
    float ey_i1, ez_i1;      // (i+1,j,k)
};

// CPML slab layer of H position x along one axis (near [0,Nc), far Nc+x-F), or -1.
__device__ inline int h_layer(const CpmlDev &c, int x) {
    return x < c.Nc ? x : (x >= c.F ? c.Nc + x - c.F : -1);
}
// CPML slab layer of E position x (near [0,Nc], far Nc+1+x-F), or -1.  The
// reference's ez near-y test is j < Ncpml (main.cu:1877); at j == Ncpml the
// recursion has ce == 0, so psi stays zero and both tests give the same field.
__device__ inline int e_layer(const CpmlDev &c, int x) {
    return x <= c.Nc ? x : (x >= c.F ? c.Nc + 1 + x - c.F : -1);
}

// New H at (i,j,k).  wr: this thread owns the cell and writes psi_H.
__device__ inline void h_cell(const StepDev &s, int i, int j, int k, const ENbr &e, float hxo, float hyo,
                              float hzo, bool wr, float &hx, float &hy, float &hz) {
    const Geom &g = s.g;
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    const float Chh = s.c.Chh, Che = s.c.Che;
    const CpmlDev &cx = s.cp[0], &cy = s.cp[1], &cz = s.cp[2];
    const bool in = i >= 0 && j >= 0 && k >= 0;
    // psi_H is updated only on the interior i < NX-1, j < NY-1, k < NZ-1 (main.cu:834-1037)
    const bool pin = in && i < NX - 1 && j < NY - 1 && k < NZ - 1;
    const int lx = pin ? h_layer(cx, i) : -1, ly = pin ? h_layer(cy, j) : -1, lz = pin ? h_layer(cz, k) : -1;
    const int Lz = 2 * cz.Nc, Ly = 2 * cy.Nc;
    // hx: i < NX, j < NY-1, k < NZ-1 (main.cu:829)
    hx = hxo;
    if (in && i < NX && j < NY - 1 && k < NZ - 1) {
        const float dEy = e.ey_k1 - e.ey, dEz = e.ez_j1 - e.ez;
        float h = Chh * hxo + Che * (dEy * cz.ik_h[k] - dEz * cy.ik_h[j]);
        if (lz >= 0) h += Che * psi_step(s.psiH_in.p[0], s.psiH_out.p[0], zslab(g, Lz, i, j, lz), cz.bh[lz], cz.ch[lz], dEy, wr);
        if (ly >= 0) h -= Che * psi_step(s.psiH_in.p[1], s.psiH_out.p[1], yslab(g, Ly, i, ly, k), cy.bh[ly], cy.ch[ly], dEz, wr);
        if (s.tfsf_on) h += tfsf_h(s.tf, 0, i, j, k);
        hx = h;
    }
    // hy: i < NX-1, j < NY, k < NZ-1 (main.cu:918)
    hy = hyo;
    if (in && i < NX - 1 && j < NY && k < NZ - 1) {
        const float dEz = e.ez_i1 - e.ez, dEx = e.ex_k1 - e.ex;
        float h = Chh * hyo + Che * (dEz * cx.ik_h[i] - dEx * cz.ik_h[k]);
        if (lz >= 0) h -= Che * psi_step(s.psiH_in.p[2], s.psiH_out.p[2], zslab(g, Lz, i, j, lz), cz.bh[lz], cz.ch[lz], dEx, wr);
        if (lx >= 0) h += Che * psi_step(s.psiH_in.p[3], s.psiH_out.p[3], xslab(g, lx, j, k), cx.bh[lx], cx.ch[lx], dEz, wr);
        if (s.tfsf_on) h += tfsf_h(s.tf, 1, i, j, k);
        hy = h;
    }
    // hz: i < NX-1, j < NY-1, k < NZ (main.cu:1014)
    hz = hzo;
    if (in && i < NX - 1 && j < NY - 1 && k < NZ) {
        const float dEx = e.ex_j1 - e.ex, dEy = e.ey_i1 - e.ey;
        float h = Chh * hzo + Che * (dEx * cy.ik_h[j] - dEy * cx.ik_h[i]);
        if (lx >= 0) h -= Che * psi_step(s.psiH_in.p[4], s.psiH_out.p[4], xslab(g, lx, j, k), cx.bh[lx], cx.ch[lx], dEy, wr);
        if (ly >= 0) h += Che * psi_step(s.psiH_in.p[5], s.psiH_out.p[5], yslab(g, Ly, i, ly, k), cy.bh[ly], cy.ch[ly], dEx, wr);
        if (s.tfsf_on) h += tfsf_h(s.tf, 2, i, j, k);
        hz = h;
    }
}

// ----------------------------------------------------------------- per-cell E update (main.cu:1048-1912)

// H^{n+1/2} values around (i,j,k) needed by the E update.
struct HNbr {
    float hx, hy, hz;        // (i,j,k)
    float hz_jm, hx_jm;      // (i,j-1,k)
    float hy_km, hx_km;      // (i,j,k-1)
    float hz_im, hy_im;      // (i-1,j,k)
};

__device__ inline void e_cell(const StepDev &s, int i, int j, int k, const HNbr &h, float exo, float eyo,
                              float ezo, float &ex, float &ey, float &ez) {
    const Geom &g = s.g;
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    const float Cee = s.c.Cee, Ceh = s.c.Ceh;
    const CpmlDev &cx = s.cp[0], &cy = s.cp[1], &cz = s.cp[2];
    const bool in = i >= 0 && j >= 0 && k >= 0;
    // psi_E runs on i < NX-1, j < NY-1, k < NZ-1 intersected with each component's domain
    const bool pin = in && i < NX - 1 && j < NY - 1 && k < NZ - 1;
    const int lx = pin ? e_layer(cx, i) : -1, ly = pin ? e_layer(cy, j) : -1, lz = pin ? e_layer(cz, k) : -1;
    const int Lz = 2 * (cz.Nc + 1), Ly = 2 * (cy.Nc + 1);
    // ex: i < NX-1, 0 < j < NY-1, 0 < k < NZ-1 (main.cu:1172)
    ex = exo;
    if (in && i < NX - 1 && j > 0 && j < NY - 1 && k > 0 && k < NZ - 1) {
        float eo = exo;
        if (s.tfsf_on) eo += tfsf_e(s.tf, 0, i, j, k);
        const float dHz = h.hz - h.hz_jm, dHy = h.hy - h.hy_km;
        float e = Cee * eo + Ceh * (dHz * cy.ik_e[j] - dHy * cz.ik_e[k]);
        if (lz >= 0) e -= Ceh * psi_step(s.psiE.p[0], s.psiE.p[0], zslab(g, Lz, i, j, lz), cz.be[lz], cz.ce[lz], dHy, true);
        if (ly >= 0) e += Ceh * psi_step(s.psiE.p[1], s.psiE.p[1], yslab(g, Ly, i, ly, k), cy.be[ly], cy.ce[ly], dHz, true);
        ex = e;
    }
    // ey: 0 < i < NX-1, j < NY-1, 0 < k < NZ-1 (main.cu:1476)
    ey = eyo;
    if (in && i > 0 && i < NX - 1 && j < NY - 1 && k > 0 && k < NZ - 1) {
        float eo = eyo;
        if (s.tfsf_on) eo += tfsf_e(s.tf, 1, i, j, k);
        const float dHx = h.hx - h.hx_km, dHz = h.hz - h.hz_im;
        float e = Cee * eo + Ceh * (dHx * cz.ik_e[k] - dHz * cx.ik_e[i]);
        if (lz >= 0) e += Ceh * psi_step(s.psiE.p[2], s.psiE.p[2], zslab(g, Lz, i, j, lz), cz.be[lz], cz.ce[lz], dHx, true);
        if (lx >= 0) e -= Ceh * psi_step(s.psiE.p[3], s.psiE.p[3], xslab(g, lx, j, k), cx.be[lx], cx.ce[lx], dHz, true);
        ey = e;
    }
    // ez: 0 < i < NX-1, 0 < j < NY-1, k < NZ-1 (main.cu:1767)
    ez = ezo;
    if (in && i > 0 && i < NX - 1 && j > 0 && j < NY - 1 && k < NZ - 1) {
        float eo = ezo;
        if (s.tfsf_on) eo += tfsf_e(s.tf, 2, i, j, k);
        const float dHy = h.hy - h.hy_im, dHx = h.hx - h.hx_jm;
        float e = Cee * eo + Ceh * (dHy * cx.ik_e[i] - dHx * cy.ik_e[j]);
        if (ly >= 0) e -= Ceh * psi_step(s.psiE.p[4], s.psiE.p[4], yslab(g, Ly, i, ly, k), cy.be[ly], cy.ce[ly], dHx, true);
        if (lx >= 0) e += Ceh * psi_step(s.psiE.p[5], s.psiE.p[5], xslab(g, lx, j, k), cx.be[lx], cx.ce[lx], dHy, true);
        ez = e;
    }
}

// ----------------------------------------------------------------- DFT face sampling (main.cu:35-366)

// Writes the tangential E and H of (i,j,k) into the sample row for every
// monitor face containing the cell.  Column layout: E half [0, mE), H half at
// mE_pad + the same offset; per face: component a block then component b.
__device__ inline void dft_sample_cell(const FaceDev &f, float *row, int i, int j, int k, float ex, float ey,
                                       float ez, float hx, float hy, float hz) {
#pragma unroll
    for (int q = 0; q < 12; q++) {
        const int axis = (q >> 1) % 3;
        int n, a, b;
        if (axis == 0) { n = i; a = j; b = k; }
        else if (axis == 1) { n = j; a = i; b = k; }
        else { n = k; a = i; b = j; }
        if (n != f.plane[q]) continue;
        a -= f.lo_a[q];
        b -= f.lo_b[q];
        if (a < 0 || a >= f.na[q] || b < 0 || b >= f.nb[q]) continue;
        const long col = f.off[q] + (long)a * f.nb[q] + b;
        const long blk = (long)f.na[q] * f.nb[q];
        float ea, eb, ha, hb;
        if (axis == 0) { ea = ey; eb = ez; ha = hy; hb = hz; }
        else if (axis == 1) { ea = ez; eb = ex; ha = hz; hb = hx; }
        else { ea = ex; eb = ey; ha = hx; hb = hy; }
        row[col] = ea;
        row[col + blk] = eb;
        row[f.mE_pad + col] = ha;
        row[f.mE_pad + col + blk] = hb;
    }
}

// ----------------------------------------------------------------- fused step kernel

// Blocks are dealt round-robin over the 8 XCDs; this renumbering gives each
// XCD a contiguous run of logical blocks, so neighbouring tiles share an L2.
__device__ inline int xcd_order(int b, int nblk) {
    const int q = nblk >> 3, r = nblk & 7, x = b & 7, slot = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
}

// Lane l of a wave receives v from lane l-1 (DPP wave_shr:1); lane 0 gets 0.
__device__ inline float from_lower_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

// Tile: 64 lanes along z (k = kt*63 - 1 + lane) by TJ wave rows along y
// (j = jt*(TJ-1) - 1 + row); a block marches x over [i0, i1).  Lane 0 and row
// 0 are halo: they recompute H there (from the read-only input buffers) so E
// at the tile's low faces has its k-1 and j-1 neighbours, and own nothing.
// H at k-1 comes from the next-lower lane through DPP, H at j-1 from the row
// below through LDS (double buffered, one barrier per x step), H at i-1 stays
// in registers.
//
// Everything that depends only on (j,k) is hoisted out of the x march: the
// component domains, 1/(kappa d) along y and z, the CPML layers and
// recursion coefficients of the y and z slabs, psi offsets.  CPML and TFSF
// work sits behind wave-uniform tests (a z-slab lane in the wave, a y-slab
// row, an x-slab plane, a TFSF face), so a wave in the vacuum interior runs
// the bare Yee update.  Field loads of plane i+1 are issued one x step ahead
// of their use; psi_H of plane i+1 is fetched right after H of plane i is
// done and psi_E of plane i at the top of its step.
//
// Blocks are renumbered so each XCD (blocks are dealt round-robin over 8)
// receives a contiguous run of tiles: tiles adjacent in j, whose halo rows
// overlap, then share that XCD's L2.
//
// The body takes the block's index within its own cell set (L), so batched
// launches run the same code over many simulations.
template <int TJ>
__device__ __forceinline__ void fused_body(const StepDev &s, int ichunk, int nkt, const Pieces &pc, int L) {
    __shared__ float sH[2][2][TJ][64];   // hz, hx of the plane for the j-1 neighbour
    const Geom &g = s.g;
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    int jt, kt, ch, ia, ib, ja, jb;
    {
        int q = 0;
        while (q + 1 < pc.n && L >= pc.start[q + 1]) q++;
        L -= pc.start[q];
        jt = L % pc.njt[q];
        kt = (L / pc.njt[q]) % nkt;
        ch = L / (pc.njt[q] * nkt);
        ia = pc.ia[q]; ib = pc.ib[q]; ja = pc.ja[q]; jb = pc.jb[q];
    }
    const int lane = threadIdx.x, row = threadIdx.y;
    const int k = kt * 63 - 1 + lane;
    const int j = __builtin_amdgcn_readfirstlane(ja + jt * (TJ - 1) - 1 + row);
    const int i0 = ia + ch * ichunk, i1 = min(i0 + ichunk, ib);
    const bool own = lane > 0 && row > 0 && k < NZ && j < jb;
    const int kc = min(k, NZ - 1), jc = min(j, NY - 1);
    const int kk = k < NZ ? k : NZ, jj = j < NY ? j : NY;   // beyond the grid: outside every domain
    const int sz = g.sz;
    const long sxy = g.sxy;
    const int cidx = jc * sz + kc;

    // (j,k) parts of the component domains (reference loop bounds, main.cu:829-1014, 1172-1767)
    const bool jk_in = jj >= 0 && kk >= 0;
    const bool dhx = jk_in && jj < NY - 1 && kk < NZ - 1;   // also the psi (j,k) range
    const bool dhy = jk_in && jj < NY && kk < NZ - 1;
    const bool dhz = jk_in && jj < NY - 1 && kk < NZ;
    const bool dex = dhx && jj > 0 && kk > 0;
    const bool dey = dhx && kk > 0;
    const bool dez = dhx && jj > 0;

    const CpmlDev &cx = s.cp[0], &cy = s.cp[1], &cz = s.cp[2];
    const int kq = max(kc, 0), jq = max(jc, 0);
    const float ikz_h = cz.ik_h[kq], ikz_e = cz.ik_e[kq];
    const float iky_h = cy.ik_h[jq], iky_e = cy.ik_e[jq];
    // z slabs: per lane
    const int lzh = dhx ? h_layer(cz, kk) : -1, lze = dhx ? e_layer(cz, kk) : -1;
    const bool zh = __ballot(lzh >= 0) != 0, ze = __ballot(lze >= 0) != 0;
    const int Lzh = 2 * cz.Nc, Lze = 2 * (cz.Nc + 1);
    const float bzh = lzh >= 0 ? cz.bh[lzh] : 0.f, czh = lzh >= 0 ? cz.ch[lzh] : 0.f;
    const float bze = lze >= 0 ? cz.be[lze] : 0.f, cze = lze >= 0 ? cz.ce[lze] : 0.f;
    const int ozh = jq * Lzh + max(lzh, 0), oze = jq * Lze + max(lze, 0);
    const long szh = (long)NY * Lzh, sze = (long)NY * Lze;
    // y slabs: per row (wave-uniform)
    const int lyh = (j >= 0 && j < NY - 1) ? h_layer(cy, j) : -1, lye = (j >= 0 && j < NY - 1) ? e_layer(cy, j) : -1;
    const int Lyh = 2 * cy.Nc, Lye = 2 * (cy.Nc + 1);
    const float byh = lyh >= 0 ? cy.bh[lyh] : 0.f, cyh = lyh >= 0 ? cy.ch[lyh] : 0.f;
    const float bye = lye >= 0 ? cy.be[lye] : 0.f, cye = lye >= 0 ? cy.ce[lye] : 0.f;
    const int oyh = max(lyh, 0) * NZ + kq, oye = max(lye, 0) * NZ + kq;
    const long syh = (long)Lyh * NZ, sye = (long)Lye * NZ;
    // x slabs: offset within a layer
    const int ox = jq * NZ + kq;
    const long sxl = (long)NY * NZ;
    // TFSF faces touching this wave (H faces sit at t0-1 and t1, E faces at t0 and t1)
    const TfsfDev &tf = s.tf;
    const bool tf_on = s.tfsf_on != 0;
    const bool tyh = tf_on && (j == tf.t0[1] - 1 || j == tf.t1[1]), tye = tf_on && (j == tf.t0[1] || j == tf.t1[1]);
    const bool tzh = tf_on && __ballot(k == tf.t0[2] - 1 || k == tf.t1[2]) != 0;
    const bool tze = tf_on && __ballot(k == tf.t0[2] || k == tf.t1[2]) != 0;

    const float Chh = s.c.Chh, Che = s.c.Che, Cee = s.c.Cee, Ceh = s.c.Ceh;
    const float *Ex = s.Ein[0], *Ey = s.Ein[1], *Ez = s.Ein[2];
    const float *Hx = s.Hin[0], *Hy = s.Hin[1], *Hz = s.Hin[2];
    const PsiSet &pHi = s.psiH_in, &pHo = s.psiH_out, &pE = s.psiE;

    // H^{n+1/2} at plane i0-1 (values only)
    float hyP = 0.f, hzP = 0.f;
    if (i0 > 0) {
        const long p = (long)(i0 - 1) * sxy + cidx;
        ENbr e;
        e.ex = Ex[p]; e.ey = Ey[p]; e.ez = Ez[p];
        e.ex_j1 = Ex[p + sz]; e.ez_j1 = Ez[p + sz];
        e.ex_k1 = Ex[p + 1]; e.ey_k1 = Ey[p + 1];
        e.ey_i1 = Ey[p + sxy]; e.ez_i1 = Ez[p + sxy];
        float hx, hy, hz;
        h_cell(s, i0 - 1, jj, kk, e, Hx[p], Hy[p], Hz[p], false, hx, hy, hz);
        hyP = hy; hzP = hz;
    }
    // pipeline fill: E^n at i0 and i0+1, E^n neighbours and H at i0, psi_H at i0
    float exo, eyo, ezo, e1x, e1y, e1z, nxj, nzj, nxk, nyk, hxo, hyo, hzo;
    {
        const long p = (long)i0 * sxy + cidx;
        exo = Ex[p]; eyo = Ey[p]; ezo = Ez[p];
        e1x = Ex[p + sxy]; e1y = Ey[p + sxy]; e1z = Ez[p + sxy];
        nxj = Ex[p + sz]; nzj = Ez[p + sz];
        nxk = Ex[p + 1]; nyk = Ey[p + 1];
        hxo = Hx[p]; hyo = Hy[p]; hzo = Hz[p];
    }
    float qhxz = 0.f, qhyz = 0.f, qhxy = 0.f, qhzy = 0.f, qhyx = 0.f, qhzx = 0.f;   // psi_H of the current plane
    auto fetch_psiH = [&](int ip) {
        if (ip >= NX - 1) return;
        if (zh && lzh >= 0) {
            const long q = ip * szh + ozh;
            qhxz = pHi.p[0][q]; qhyz = pHi.p[2][q];
        }
        if (lyh >= 0 && dhx) {
            const long q = ip * syh + oyh;
            qhxy = pHi.p[1][q]; qhzy = pHi.p[5][q];
        }
        const int lx = h_layer(cx, ip);
        if (lx >= 0 && dhx) {
            const long q = lx * sxl + ox;
            qhyx = pHi.p[3][q]; qhzx = pHi.p[4][q];
        }
    };
    fetch_psiH(i0);

    int buf = 0;
    for (int i = i0; i < i1; i++) {
        const long p = (long)i * sxy + cidx;
        const bool more = i + 1 < i1;
        // fields of the next step
        float f_e1x = 0.f, f_e1y = 0.f, f_e1z = 0.f, f_nxj = 0.f, f_nzj = 0.f, f_nxk = 0.f, f_nyk = 0.f;
        float f_hx = 0.f, f_hy = 0.f, f_hz = 0.f;
        if (more) {
            const long p1 = p + sxy, p2 = min(i + 2, NX) * sxy + cidx;
            f_e1x = Ex[p2]; f_e1y = Ey[p2]; f_e1z = Ez[p2];
            f_nxj = Ex[p1 + sz]; f_nzj = Ez[p1 + sz];
            f_nxk = Ex[p1 + 1]; f_nyk = Ey[p1 + 1];
            f_hx = Hx[p1]; f_hy = Hy[p1]; f_hz = Hz[p1];
        }
        const bool pin = i < NX - 1;
        // psi_E of this plane
        const bool xin = i > 0 && pin;   // ey, ez need 0 < i < NX-1
        const int lxe = pin ? e_layer(cx, i) : -1;
        float qexz = 0.f, qeyz = 0.f, qexy = 0.f, qezy = 0.f, qeyx = 0.f, qezx = 0.f;
        if (pin && ze && lze >= 0) {
            const long q = i * sze + oze;
            if (dex) qexz = pE.p[0][q];
            if (dey && xin) qeyz = pE.p[2][q];
        }
        if (pin && lye >= 0) {
            const long q = i * sye + oye;
            if (dex) qexy = pE.p[1][q];
            if (dez && xin) qezy = pE.p[4][q];
        }
        if (lxe >= 0 && xin) {
            const long q = lxe * sxl + ox;
            if (dey) qeyx = pE.p[3][q];
            if (dez) qezx = pE.p[5][q];
        }

        // ---- H^{n+1/2} at plane i
        const float ikx_h = cx.ik_h[i];
        const int lxh = pin ? h_layer(cx, i) : -1;
        float hx, hy, hz;
        {
            const float dEy_z = nyk - eyo, dEz_y = nzj - ezo;   // hx
            const float dEz_x = e1z - ezo, dEx_z = nxk - exo;   // hy
            const float dEx_y = nxj - exo, dEy_x = e1y - eyo;   // hz
            hx = Chh * hxo + Che * (dEy_z * ikz_h - dEz_y * iky_h);
            hy = Chh * hyo + Che * (dEz_x * ikx_h - dEx_z * ikz_h);
            hz = Chh * hzo + Che * (dEx_y * iky_h - dEy_x * ikx_h);
            if (pin && zh && lzh >= 0) {
                const long q = i * szh + ozh;
                const float a = bzh * qhxz + czh * dEy_z;
                const float b = bzh * qhyz + czh * dEx_z;
                hx += Che * a;
                hy -= Che * b;
                if (own) { pHo.p[0][q] = a; pHo.p[2][q] = b; }
            }
            if (pin && lyh >= 0 && dhx) {
                const long q = i * syh + oyh;
                const float a = byh * qhxy + cyh * dEz_y;
                hx -= Che * a;
                if (own) pHo.p[1][q] = a;
            }
            if (lxh >= 0 && dhx) {
                const long q = lxh * sxl + ox;
                const float bx = cx.bh[lxh], cxv = cx.ch[lxh];
                const float a = bx * qhyx + cxv * dEz_x;
                const float b = bx * qhzx + cxv * dEy_x;
                hy += Che * a;
                hz -= Che * b;
                if (own) { pHo.p[3][q] = a; pHo.p[4][q] = b; }
            }
            if (pin && lyh >= 0 && dhx) {
                const long q = i * syh + oyh;
                const float a = byh * qhzy + cyh * dEx_y;
                hz += Che * a;
                if (own) pHo.p[5][q] = a;
            }
            const bool txh = tf_on && (i == tf.t0[0] - 1 || i == tf.t1[0]);
            if (txh || tyh || tzh) {
                float rx, ry, rz;
                tfsf_faces(tf, false, txh, tyh, tzh, i, jj, kk, rx, ry, rz);
                hx += rx; hy += ry; hz += rz;
            }
            hx = dhx ? hx : hxo;
            hy = (dhy && pin) ? hy : hyo;
            hz = (dhz && pin) ? hz : hzo;
        }
        if (more) fetch_psiH(i + 1);
        if (own) { s.Hout[0][p] = hx; s.Hout[1][p] = hy; s.Hout[2][p] = hz; }
        sH[buf][0][row][lane] = hz;
        sH[buf][1][row][lane] = hx;
        const float hy_km = from_lower_lane(hy), hx_km = from_lower_lane(hx);
        __syncthreads();
        const int rm = row > 0 ? row - 1 : 0;
        const float hz_jm = sH[buf][0][rm][lane], hx_jm = sH[buf][1][rm][lane];

        // ---- E^{n+1} at plane i
        if (own) {
            const float ikx_e = cx.ik_e[i];
            float eox = exo, eoy = eyo, eoz = ezo;
            const bool txe = tf_on && (i == tf.t0[0] || i == tf.t1[0]);
            if (txe || tye || tze) {
                float rx, ry, rz;
                tfsf_faces(tf, true, txe, tye, tze, i, j, k, rx, ry, rz);
                eox += rx; eoy += ry; eoz += rz;
            }
            const float dHz_y = hz - hz_jm, dHy_z = hy - hy_km;   // ex
            const float dHx_z = hx - hx_km, dHz_x = hz - hzP;     // ey
            const float dHy_x = hy - hyP, dHx_y = hx - hx_jm;     // ez
            float ex = Cee * eox + Ceh * (dHz_y * iky_e - dHy_z * ikz_e);
            float ey = Cee * eoy + Ceh * (dHx_z * ikz_e - dHz_x * ikx_e);
            float ez = Cee * eoz + Ceh * (dHy_x * ikx_e - dHx_y * iky_e);
            const bool ux = dex && pin, uy = dey && xin, uz = dez && xin;
            if (pin && ze && lze >= 0) {
                const long q = i * sze + oze;
                if (ux) { const float a = bze * qexz + cze * dHy_z; ex -= Ceh * a; pE.p[0][q] = a; }
                if (uy) { const float a = bze * qeyz + cze * dHx_z; ey += Ceh * a; pE.p[2][q] = a; }
            }
            if (pin && lye >= 0) {
                const long q = i * sye + oye;
                if (ux) { const float a = bye * qexy + cye * dHz_y; ex += Ceh * a; pE.p[1][q] = a; }
            }
            if (lxe >= 0 && xin) {
                const long q = lxe * sxl + ox;
                const float bx = cx.be[lxe], cxv = cx.ce[lxe];
                if (uy) { const float a = bx * qeyx + cxv * dHz_x; ey -= Ceh * a; pE.p[3][q] = a; }
            }
            if (pin && lye >= 0) {
                const long q = i * sye + oye;
                if (uz) { const float a = bye * qezy + cye * dHx_y; ez -= Ceh * a; pE.p[4][q] = a; }
            }
            if (lxe >= 0 && xin) {
                const long q = lxe * sxl + ox;
                const float bx = cx.be[lxe], cxv = cx.ce[lxe];
                if (uz) { const float a = bx * qezx + cxv * dHy_x; ez += Ceh * a; pE.p[5][q] = a; }
            }
            s.Eout[0][p] = ux ? ex : exo;
            s.Eout[1][p] = uy ? ey : eyo;
            s.Eout[2][p] = uz ? ez : ezo;
        }
        hyP = hy; hzP = hz;
        exo = e1x; eyo = e1y; ezo = e1z;
        e1x = f_e1x; e1y = f_e1y; e1z = f_e1z;
        nxj = f_nxj; nzj = f_nzj; nxk = f_nxk; nyk = f_nyk;
        hxo = f_hx; hyo = f_hy; hzo = f_hz;
        buf ^= 1;
    }
}

template <int TJ>
__global__ __launch_bounds__(64 * TJ) void k_fused_step(StepDev s, int ichunk, int nkt, Pieces pc) {
    fused_body<TJ>(s, ichunk, nkt, pc, xcd_order(blockIdx.x, pc.start[pc.n]));
}

// ----------------------------------------------------------------- row-blocked step kernel

__device__ inline float ldb(const char *b, unsigned o) { return *(const float *)(b + o); }
__device__ inline void stb(char *b, unsigned o, float v) { *(float *)(b + o) = v; }
// Lane l of a wave receives v from lane l+1 (DPP wave_shl:1); lane 63 gets 0.
__device__ inline float from_upper_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// Layout of the row-blocked kernel's table image (32-bit words), copied to
// LDS by every block: per axis a at a*512: 1/(kappa d) for E at x < ew
// (near, @0) and x >= N-ew (far, @64), the same for H (@128, @192), CPML
// recursion be, ce/d (@256, @320: near layers then far), bh, ch/d (@384,
// @448); then the interior 1/(kappa d) for E (@1536+a) and H (@1539+a), ew
// (@1542+a), and the psi pointers at word 1548: psi_H sets 0 and 1 (6 each),
// psi_E (6), and for the adjoint a second psi_E set (6), 64-bit each.
constexpr int kAuxAxis = 512, kAuxIkE0 = 1536, kAuxIkH0 = 1539, kAuxEw = 1542, kAuxPsi = 1548;
typedef __attribute__((address_space(1))) float gfloat;   // global memory (psi pointers read from LDS)

// Workgroup barrier ordering LDS only: global loads in flight (prefetches)
// are not waited for, as a full __syncthreads() fence would do.
__device__ inline void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <bool NT>
__device__ inline void stn(char *b, unsigned o, float v) {
    float *p = (float *)(b + o);
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// One full leapfrog step of the whole grid (TFSF excluded: k_tfsf_fix adds
// it).  A block spans a z segment with W waves, k = seg*(64W-2) - 1 + t for
// t = 64w + lane; its first and last lanes are halo.  Each thread holds R
// rows of y in registers (row 0 is halo, recomputed) and the block marches x
// over one chunk.  Neighbours: E at k+1 and H at k-1 from the adjacent lane
// (DPP) or, across a wave edge, from the adjacent wave through LDS; E at j+1
// and H at j-1 from the adjacent row's registers; H at i-1 from the previous
// x step.  One barrier per x step.  Loads of plane i+1 are issued while
// plane i is computed; fields are stored non-temporally (each value is
// written once and not read again in this step).  Field offsets are 32-bit
// byte offsets from the allocation base, so every field access is a scalar
// base plus a vector offset.  CPML and the domain tests (main.cu:829-1014,
// 1172-1767) follow the uniformity of their axis: x per plane, y per row, z
// per lane (only the edge waves of a segment hold z-slab lanes); per
// component the psi terms add in the reference's order.  CPML tables and psi
// pointers live in LDS (the table image above), which keeps scalar registers
// for the field pointers.
//
// ADJ: the transposed step of the discrete adjoint on a clean box (no x or y
// CPML).  Written with the weights of k_adj_step, the adjoint of a step is
//   Hhat_tot = Hhat + forward differences of V,  V = Q ik (+ z CPML), Q = Ceh Ehat
//   Ehat'    = Cee Ehat + backward differences of W,  W = G ik (+ z CPML), G = Che Hhat_tot
// which is the forward step with Che -> -Ceh and Ceh -> -Che (the host passes
// them swapped) wherever 1/(kappa d) is constant; along z each lane first
// forms its own z weights (Vyz/Ceh, Vxz/Ceh before the H half, Wyz/Che,
// Wxz/Che before the E half, with the adjoint psi recursions of its z-slab
// cells) and the z differences are taken of those.  Both adjoint psi sets are
// double buffered (pp: psi_H set read, pe: psi_E set read).  The lanes at
// wave edges must lie outside the z slabs (checked by the host).
//
// ADJ with XG (x slabs) or YG (y slabs), or both (corner columns): the weights
// along each general axis are formed per cell with their adjoint psi recursions (V of rows
// 0..R+1 of the plane along y; V of plane i+1, carried to the next step, and
// W of plane i-1 along x) and the differences are taken of them; the other
// in-plane axis is clean and keeps the swapped-coefficient form.  The slabs
// hold no dispersive cell (the host keeps the dispersive box and its margin
// inside the clean box).
//
// The body takes the block's index within its simulation's cell set (L), so
// batched launches run it over many simulations.
//
// T row tiles per block: waves T*W; tile wg / W takes row tile jt = T*jb + tile
// and the tiles march the same x chunk in step (one barrier per x step for
// the whole block).  A workgroup is placed only where every SIMD has
// ceil(waves / 4) free wave slots, so at four waves per SIMD a 5-wave block
// holds two per CU (10 waves) while one 15-wave block (T = 3) holds 15.
// Tiles past the last row tile of the box repeat its loads and store nothing.
template <int W, int R, bool XG, bool YG, bool ADJ, int T = 1>
__device__ __forceinline__ void rows_body(const RowsArgs &ra, int L) {
    static_assert(ADJ || !XG || !YG, "corner columns of the forward step run on the general kernel");
    constexpr bool AS = ADJ && (XG || YG);   // adjoint on a side slab or corner column
    static_assert(T >= 1 && W * T <= 16, "at most 16 waves per block");
    __shared__ float sK[2][2][W * T][R + 1];   // last lane's hy, hx per wave and row (H at k-1 of the next wave)
    __shared__ float sE[2][2][W * T][R + 1];   // first lane's ex, ey per wave and row (E at k+1 of the previous wave)
    constexpr int kTab = kAuxPsi + (ADJ ? 48 : 36);
    __shared__ __align__(16) unsigned sT[kTab];
    const int NX = ra.NX, NY = ra.NY, NZ = ra.NZ;
    for (int q = threadIdx.y * 64 + threadIdx.x; q < kTab; q += 64 * W * T) sT[q] = ra.aux[q];
    const float *sTf = (const float *)sT;
    int pc = 0;
    while (pc + 1 < ra.npiece && L >= ra.start[pc + 1]) pc++;
    L -= ra.start[pc];
    const int ia = ra.box[pc][0], ib = ra.box[pc][1], ja = ra.box[pc][2], jbox = ra.box[pc][3];
    const int njt = (jbox - ja + R - 1) / R;
    const int njb = (njt + T - 1) / T;
    const int jtb = L % njb;
    L /= njb;
    const int seg = L % ra.nseg;
    const int ch = L / ra.nseg;
    // the wave index is uniform across a wave (blocks are 64 wide): keep it,
    // and the row tile and row offsets derived from it, in scalar registers
    const int lane = threadIdx.x, wg = __builtin_amdgcn_readfirstlane((int)threadIdx.y), w = T == 1 ? wg : wg % W;
    const int t = w * 64 + lane;
    const int jt0 = T == 1 ? jtb : jtb * T + wg / W;
    const bool live = jt0 < njt;
    const int jt = live ? jt0 : njt - 1;
    const int k = seg * (64 * W - 2) - 1 + t;
    const int kc = min(k, NZ - 1);
    const bool kown = live && t > 0 && t < 64 * W - 1 && k < NZ;
    const int i0 = ia + ch * ra.ichunk, i1 = min(i0 + ra.ichunk, ib);
    const int pp = ra.pp;
    auto hlay = [&](int a, int x) { return x < ra.Nc[a] ? x : (x >= ra.F[a] ? ra.Nc[a] + x - ra.F[a] : -1); };
    auto elay = [&](int a, int x) { return x <= ra.Nc[a] ? x : (x >= ra.F[a] ? ra.Nc[a] + 1 + x - ra.F[a] : -1); };
    // 1/(kappa d) of axis a (length n) at x, E positions if e
    auto ik = [&](int a, bool e, int x, int n) {
        const int ew = (int)sT[kAuxEw + a];
        const int base = a * kAuxAxis + (e ? 0 : 128);
        const int q = x < ew ? base + x : (x >= n - ew ? base + 64 + x - (n - ew) : (e ? kAuxIkE0 : kAuxIkH0) + a);
        return sTf[q];
    };
    auto coef = [&](int a, int which, int layer) { return sTf[a * kAuxAxis + 256 + which * 64 + layer]; };
    auto psi = [&](int idx) { return *(gfloat *const *)&sT[kAuxPsi + 2 * idx]; };   // 0-5 set 0, 6-11 set 1, 12-17 psi_E
    // adjoint: psi_E read from set pe (12 + 6 pe), written to the other
    const int ein = ADJ ? 12 + 6 * ra.pe : 12, eout = ADJ ? 18 - 6 * ra.pe : 12;
    const int hout = (1 - pp) * 6;
    lds_sync();

    // z: per lane
    const bool kA = k >= 0 && k < NZ - 1, kB = k >= 0 && k < NZ, kC = k > 0 && k < NZ - 1;
    const int kq = max(kc, 0);
    const int lzh = kA ? hlay(2, k) : -1, lze = kC ? elay(2, k) : -1;
    const bool zh = __ballot(lzh >= 0) != 0, ze = __ballot(lze >= 0) != 0;
    const float ikz_h = ik(2, false, kq, NZ), ikz_e = ik(2, true, kq, NZ);
    const float bzh = lzh >= 0 ? coef(2, 2, lzh) : 0.f, czh = lzh >= 0 ? coef(2, 3, lzh) : 0.f;
    const float bze = lze >= 0 ? coef(2, 0, lze) : 0.f, cze = lze >= 0 ? coef(2, 1, lze) : 0.f;
    const int Lzh = 2 * ra.Nc[2], Lze = 2 * (ra.Nc[2] + 1), Lyh = 2 * ra.Nc[1], Lye = 2 * (ra.Nc[1] + 1);

    // y: per row, j_r = ja + jt*R - 1 + r (row R+1 only feeds the j+1 neighbour); own rows < jbox
    const int jb = ja + jt * R - 1;
    const unsigned sz4 = (unsigned)ra.sz * 4u, sxy4 = ra.sxy4;
    unsigned ro[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; r++) ro[r] = (unsigned)(min(jb + r, NY) + 1) * sz4 + (unsigned)(kc + 1) * 4u;
    auto jA = [&](int r) { return !YG || (jb + r >= 0 && jb + r < NY - 1); };
    auto jB = [&](int r) { return !YG || (jb + r >= 0 && jb + r < NY); };
    auto jC = [&](int r) { return !YG || (jb + r > 0 && jb + r < NY - 1); };
    auto jq = [&](int r) { return min(max(jb + r, 0), NY - 1); };
    auto lyh = [&](int r) { return YG && jA(r) ? hlay(1, jb + r) : -1; };
    auto lye = [&](int r) { return YG && jC(r) ? elay(1, jb + r) : -1; };
    auto lxh = [&](int ip) { return XG && ip < NX - 1 ? hlay(0, ip) : -1; };
    auto lxe = [&](int ip) { return XG && ip > 0 && ip < NX - 1 ? elay(0, ip) : -1; };

    const float Chh = ra.Chh, Che = ra.Che, Cee = ra.Cee, Ceh = ra.Ceh;
    // adjoint: the forward Ceh and Che (Q = cE Ehat, G = cH Hhat_tot) and their inverses
    const float cE = -Che, cH = -Ceh, icE = ADJ ? 1.f / cE : 0.f, icH = ADJ ? 1.f / cH : 0.f;
    const char *bEx = ra.Ein[0], *bEy = ra.Ein[1], *bEz = ra.Ein[2];
    const char *bHx = ra.Hin[0], *bHy = ra.Hin[1], *bHz = ra.Hin[2];
    char *oEx = ra.Eout[0], *oEy = ra.Eout[1], *oEz = ra.Eout[2];
    char *oHx = ra.Hout[0], *oHy = ra.Hout[1], *oHz = ra.Hout[2];
    const bool last_lane = lane == 63 && w < W - 1, first_lane = lane == 0 && w > 0;

    auto e_up = [&](const float *x, const float *y, int b, float *xk, float *yk) {
#pragma unroll
        for (int r = 0; r <= R; r++) { xk[r] = from_upper_lane(x[r]); yk[r] = from_upper_lane(y[r]); }
        if (last_lane) {
#pragma unroll
            for (int r = 0; r <= R; r++) { xk[r] = sE[b][0][wg + 1][r]; yk[r] = sE[b][1][wg + 1][r]; }
        }
    };
    // adjoint: the first lane (outside the z slabs) publishes its z weights
    auto e_publish = [&](const float *x, const float *y, int b) {
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r <= R; r++) {
                sE[b][0][wg][r] = ADJ ? (kC ? x[r] * ikz_e : 0.f) : x[r];
                sE[b][1][wg][r] = ADJ ? (kC ? y[r] * ikz_e : 0.f) : y[r];
            }
        }
    };

    // psi of one plane: z (psi_H hx_z, hy_z rows 0..R; psi_E ex_z, ey_z rows
    // 1..R), y (hx_y, hz_y; ex_y, ez_y) when YG, x (hy_x, hz_x; ey_x, ez_x)
    // when XG.  The y and x loads are issued unconditionally (indices clamped
    // into the slab arrays, values unused where no slab applies) so the wait
    // counts after them stay exact.
    // Adjoint slabs: exy, ezy hold rows 0..R+1 (V along y) and eyx, ezx rows
    // 0..R of plane i+1 (V along x of the next plane).
    struct PSet {
        float hxz[R + 1], hyz[R + 1], exz[R + 1], eyz[R + 1];
        float hxy[R + 1], hzy[R + 1], exy[ADJ ? R + 2 : R], ezy[ADJ ? R + 2 : R];
        float hyx[R + 1], hzx[R + 1], eyx[ADJ ? R + 1 : R], ezx[ADJ ? R + 1 : R];
    };
    auto load_psi = [&](int ip, PSet &P) {
#pragma unroll
        for (int r = 0; r <= R; r++) { P.hxz[r] = 0.f; P.hyz[r] = 0.f; P.exz[r] = 0.f; P.eyz[r] = 0.f; }
        const bool iA = !XG || ip < NX - 1, iC = !XG || (ip > 0 && ip < NX - 1);
        if (ADJ) {
            // adjoint: psi_E of rows 0..R (z weights of the H half), psi_H of rows 1..R (E half).
            // Rows clamped into the grid: on a y slab or corner column a tile's halo row
            // can be j = -1 or j = NY (at planes 0 and NX-1 outside the arrays); their
            // values are masked out by the domain tests
            if (ze && lze >= 0) {
                const gfloat *pa = psi(ein + 0), *pb = psi(ein + 2);
#pragma unroll
                for (int r = 0; r <= R; r++) {
                    const long q = ((long)ip * NY + jq(r)) * Lze + lze;
                    P.exz[r] = pa[q];
                    P.eyz[r] = pb[q];
                }
            }
            if (zh && lzh >= 0) {
                const gfloat *pa = psi(pp * 6 + 0), *pb = psi(pp * 6 + 2);
#pragma unroll
                for (int r = 1; r <= R; r++) {
                    const long q = ((long)ip * NY + jq(r)) * Lzh + lzh;
                    P.hxz[r] = pa[q];
                    P.hyz[r] = pb[q];
                }
            }
            if (YG) {
                // psi_E along y of rows 0..R+1, psi_H along y of rows 0..R (clamped into the slab arrays)
                const gfloat *e1 = psi(ein + 1), *e4 = psi(ein + 4), *h1 = psi(pp * 6 + 1), *h5 = psi(pp * 6 + 5);
#pragma unroll
                for (int r = 0; r <= R + 1; r++) {
                    const long q = ((long)ip * Lye + max(lye(r), 0)) * NZ + kq;
                    P.exy[r] = e1[q];
                    P.ezy[r] = e4[q];
                }
#pragma unroll
                for (int r = 0; r <= R; r++) {
                    const long q = ((long)ip * Lyh + max(lyh(r), 0)) * NZ + kq;
                    P.hxy[r] = h1[q];
                    P.hzy[r] = h5[q];
                }
            }
            if (XG) {
                // psi_E along x of plane ip+1, psi_H along x of plane ip
                const gfloat *e3 = psi(ein + 3), *e5 = psi(ein + 5), *h3 = psi(pp * 6 + 3), *h4 = psi(pp * 6 + 4);
                const int le = max(lxe(ip + 1), 0), lh = max(lxh(ip), 0);
#pragma unroll
                for (int r = 0; r <= R; r++) {
                    const long qe = ((long)le * NY + jq(r)) * NZ + kq, qh = ((long)lh * NY + jq(r)) * NZ + kq;
                    P.eyx[r] = e3[qe];
                    P.ezx[r] = e5[qe];
                    P.hyx[r] = h3[qh];
                    P.hzx[r] = h4[qh];
                }
            }
            return;
        }
        if (zh && iA && lzh >= 0) {
            const gfloat *pa = psi(pp * 6 + 0), *pb = psi(pp * 6 + 2);
#pragma unroll
            for (int r = 0; r <= R; r++)
                if (jA(r)) {
                    const long q = ((long)ip * NY + jb + r) * Lzh + lzh;
                    P.hxz[r] = pa[q];
                    P.hyz[r] = pb[q];
                }
        }
        if (ze && iA && lze >= 0) {
            const gfloat *pa = psi(12 + 0), *pb = psi(12 + 2);
#pragma unroll
            for (int r = 1; r <= R; r++) {
                const long q = ((long)ip * NY + jb + r) * Lze + lze;
                if (jC(r)) P.exz[r - 1] = pa[q];
                if (iC && jA(r)) P.eyz[r - 1] = pb[q];
            }
        }
        if (YG) {
            const gfloat *h1 = psi(pp * 6 + 1), *h5 = psi(pp * 6 + 5), *e1 = psi(12 + 1), *e4 = psi(12 + 4);
#pragma unroll
            for (int r = 0; r <= R; r++) {
                const long q = ((long)ip * Lyh + max(lyh(r), 0)) * NZ + kq;
                P.hxy[r] = h1[q];
                P.hzy[r] = h5[q];
            }
#pragma unroll
            for (int r = 1; r <= R; r++) {
                const long q = ((long)ip * Lye + max(lye(r), 0)) * NZ + kq;
                P.exy[r - 1] = e1[q];
                P.ezy[r - 1] = e4[q];
            }
        }
        if (XG) {
            const gfloat *h3 = psi(pp * 6 + 3), *h4 = psi(pp * 6 + 4), *e3 = psi(12 + 3), *e5 = psi(12 + 5);
            const int lh = max(lxh(ip), 0), le = max(lxe(ip), 0);
#pragma unroll
            for (int r = 0; r <= R; r++) {
                const long q = ((long)lh * NY + jq(r)) * NZ + kq;
                P.hyx[r] = h3[q];
                P.hzx[r] = h4[q];
            }
#pragma unroll
            for (int r = 1; r <= R; r++) {
                const long q = ((long)le * NY + jq(r)) * NZ + kq;
                P.eyx[r - 1] = e3[q];
                P.ezx[r - 1] = e5[q];
            }
        }
    };

    // Adjoint z weights.  E side (rows 0..R): Vxz/Ceh = ex ik - c (psi/Ceh - ex),
    // Vyz/Ceh = ey ik + c (psi/Ceh + ey), and the own cells' psi_E <- b (psi -+ Ceh e).
    // H side (rows 1..R, from Hhat_tot): Wxz/Che = hx ik + c (psi/Che + hx),
    // Wyz/Che = hy ik - c (psi/Che - hy), psi_H <- b (psi +- Che h).  Zero
    // outside the component domains (k_adj_step's adj_v, adj_w).
    auto zweights = [&](int ip, const float *x, const float *y, const PSet &P, bool wr, float *zx, float *zy) {
#pragma unroll
        for (int r = 0; r <= R; r++) {
            // x and y parts of the ex and ey domains (always true on a clean box)
            const bool dx = (!XG || ip < NX - 1) && jC(r), dy = (!XG || (ip > 0 && ip < NX - 1)) && jA(r);
            float ax = kC ? x[r] * ikz_e : 0.f, ay = kC ? y[r] * ikz_e : 0.f;
            if (ze && lze >= 0) {
                ax -= cze * (P.exz[r] * icE - x[r]);
                ay += cze * (P.eyz[r] * icE + y[r]);
                if (wr && r > 0 && kown && jb + r < jbox) {
                    const long q = ((long)ip * NY + jb + r) * Lze + lze;
                    if (dx) psi(eout + 0)[q] = bze * (P.exz[r] - cE * x[r]);
                    if (dy) psi(eout + 2)[q] = bze * (P.eyz[r] + cE * y[r]);
                }
            }
            zx[r] = dx ? ax : 0.f;
            zy[r] = dy ? ay : 0.f;
        }
    };
    auto hweights = [&](int ip, const float *hx, const float *hy, const PSet &P, float *wx, float *wy) {
#pragma unroll
        for (int r = 1; r <= R; r++) {
            const bool dx = jA(r), dy = (!XG || ip < NX - 1) && jB(r);   // x and y parts of the hx, hy domains
            float ax = kA ? hx[r] * ikz_h : 0.f, ay = kA ? hy[r] * ikz_h : 0.f;
            if (zh && lzh >= 0) {
                ax += czh * (P.hxz[r] * icH + hx[r]);
                ay -= czh * (P.hyz[r] * icH - hy[r]);
                if (kown && jb + r < jbox) {
                    const long q = ((long)ip * NY + jb + r) * Lzh + lzh;
                    if (dx) psi(hout + 0)[q] = bzh * (P.hxz[r] + cH * hx[r]);
                    if (dy) psi(hout + 2)[q] = bzh * (P.hyz[r] - cH * hy[r]);
                }
            }
            wx[r] = dx ? ax : 0.f;
            wy[r] = dy ? ay : 0.f;
        }
    };
    // Adjoint slabs, weights along the slab axis (k_adj_step's adj_v, adj_w
    // with Q = cE E, G = cH Hhat_tot; zero outside the component domains; the
    // own cells write the psi adjoints).
    // y, V of rows 0..R+1 of plane ip: Vxy = Q ik + c (psi + Q), Vzy = Q ik - c (psi - Q)
    auto vyweights = [&](int ip, const float *ex, const float *ez, float exj, float ezj, const PSet &P, bool wr,
                         float *vxy, float *vzy) {
#pragma unroll
        for (int r = 0; r <= R + 1; r++) {
            const float Qx = cE * (r <= R ? ex[r] : exj), Qz = cE * (r <= R ? ez[r] : ezj);
            const float iky = ik(1, true, jq(r), NY);
            float ax = Qx * iky, az = Qz * iky;
            const int ly = lye(r);
            if (ly >= 0) {
                const float by = coef(1, 0, ly), cy = coef(1, 1, ly);
                const float tx = P.exy[r] + Qx, tz = P.ezy[r] - Qz;
                ax += cy * tx;
                az -= cy * tz;
                if (wr && r > 0 && r <= R && kown && jb + r < jbox) {
                    const long q = ((long)ip * Lye + ly) * NZ + k;
                    if (kC) psi(eout + 1)[q] = by * tx;
                    if (kA) psi(eout + 4)[q] = by * tz;
                }
            }
            vxy[r] = jC(r) && kC ? ax : 0.f;
            vzy[r] = jC(r) && kA ? az : 0.f;
        }
    };
    // x, V of rows 0..R of plane ip: Vyx = Q ik - c (psi - Q), Vzx = Q ik + c (psi + Q)
    auto vxweights = [&](int ip, const float *ey, const float *ez, const float *pyx, const float *pzx, bool wr,
                         float *vyx, float *vzx) {
        const bool iC = ip > 0 && ip < NX - 1;
        const float ikx = ik(0, true, min(max(ip, 0), NX - 1), NX);
        const int lx = lxe(ip);
        const float bx = lx >= 0 ? coef(0, 0, lx) : 0.f, cx = lx >= 0 ? coef(0, 1, lx) : 0.f;
#pragma unroll
        for (int r = 0; r <= R; r++) {
            const float Qy = cE * ey[r], Qz = cE * ez[r];
            float ay = Qy * ikx, az = Qz * ikx;
            if (lx >= 0) {
                const float ty = pyx[r] - Qy, tz = pzx[r] + Qz;
                ay -= cx * ty;
                az += cx * tz;
                if (wr && r > 0 && kown && jb + r < jbox) {
                    const long q = ((long)lx * NY + jb + r) * NZ + k;
                    if (kC) psi(eout + 3)[q] = bx * ty;
                    if (kA) psi(eout + 5)[q] = bx * tz;
                }
            }
            vyx[r] = iC && jA(r) && kC ? ay : 0.f;
            vzx[r] = iC && jC(r) && kA ? az : 0.f;
        }
    };
    // y, W of rows 0..R of plane ip: Wxy = G ik - c (psi - G), Wzy = G ik + c (psi + G)
    auto wyweights = [&](int ip, const float *hx, const float *hz, const PSet &P, float *wxy, float *wzy) {
#pragma unroll
        for (int r = 0; r <= R; r++) {
            const float Gx = cH * hx[r], Gz = cH * hz[r], iky = ik(1, false, jq(r), NY);
            float ax = Gx * iky, az = Gz * iky;
            const int ly = lyh(r);
            if (ly >= 0 && kA) {
                const float by = coef(1, 2, ly), cy = coef(1, 3, ly);
                const float tx = P.hxy[r] - Gx, tz = P.hzy[r] + Gz;
                ax -= cy * tx;
                az += cy * tz;
                if (r > 0 && kown && jb + r < jbox) {
                    const long q = ((long)ip * Lyh + ly) * NZ + k;
                    psi(hout + 1)[q] = by * tx;
                    psi(hout + 5)[q] = by * tz;
                }
            }
            wxy[r] = jA(r) && kA ? ax : 0.f;
            wzy[r] = jA(r) && kB ? az : 0.f;
        }
    };
    // x, W of rows 1..R of plane ip: Wyx = G ik + c (psi + G), Wzx = G ik - c (psi - G)
    auto wxweights = [&](int ip, const float *hy, const float *hz, const PSet &P, bool wr, float *wyx, float *wzx) {
        const bool iA = ip < NX - 1;
        const float ikx = ik(0, false, min(max(ip, 0), NX - 1), NX);
        const int lx = lxh(ip);
        const float bx = lx >= 0 ? coef(0, 2, lx) : 0.f, cx = lx >= 0 ? coef(0, 3, lx) : 0.f;
        wyx[0] = 0.f;
        wzx[0] = 0.f;
#pragma unroll
        for (int r = 1; r <= R; r++) {
            const float Gy = cH * hy[r], Gz = cH * hz[r];
            float ay = Gy * ikx, az = Gz * ikx;
            if (lx >= 0 && kA) {
                const float ty = P.hyx[r] + Gy, tz = P.hzx[r] - Gz;
                ay += cx * ty;
                az -= cx * tz;
                if (wr && kown && jb + r < jbox) {
                    const long q = ((long)lx * NY + jb + r) * NZ + k;
                    psi(hout + 3)[q] = bx * ty;
                    psi(hout + 4)[q] = bx * tz;
                }
            }
            wyx[r] = iA && jB(r) && kA ? ay : 0.f;
            wzx[r] = iA && jA(r) && kB ? az : 0.f;
        }
    };

    // H of plane ip for rows 0..R from E^n of plane ip (rows 0..R, ex and ez of
    // row R+1, ex and ey at k+1), ey and ez of plane ip+1, H^{n-1/2} and psi_H
    // of plane ip.  wr: write psi_H of the own cells.
    // Adjoint: exk, eyk are the z weights of the k+1 neighbours and zx, zy the
    // own ones; on the slabs vxy, vzy are the V weights along y of rows 0..R+1
    // and vyxc, vzxc (plane ip), vyxn, vzxn (plane ip+1) those along x.
    auto hplane = [&](int ip, const float *ex, const float *ey, const float *ez, float exj, float ezj,
                      const float *exk, const float *eyk, const float *ny, const float *nz, const float (*ho)[R + 1],
                      const PSet &P, bool wr, const float *zx, const float *zy, const float *vxy, const float *vzy,
                      const float *vyxc, const float *vzxc, const float *vyxn, const float *vzxn, float *hx, float *hy,
                      float *hz) {
        const bool iA = !XG || ip < NX - 1;
        const float ikx_h = XG ? ik(0, false, ip, NX) : ra.ik0[1][0];
        const int lx = lxh(ip);
#pragma unroll
        for (int r = 0; r <= R; r++) {
            const int j = jb + r;
            const bool own = wr && r > 0 && kown && j < jbox;
            const float iky_h = YG ? ik(1, false, jq(r), NY) : ra.ik0[1][1];
            const float xj = r < R ? ex[r + 1] : exj, zj = r < R ? ez[r + 1] : ezj;
            const float dEy_z = eyk[r] - ey[r], dEz_y = zj - ez[r];
            const float dEz_x = nz[r] - ez[r], dEx_z = exk[r] - ex[r];
            const float dEx_y = xj - ex[r], dEy_x = ny[r] - ey[r];
            float vx, vy, vz;
            if constexpr (AS) {
                // Hhat_tot = Hhat + forward differences of V (V along z from the lane weights)
                vx = Chh * ho[0][r] + Che * (eyk[r] - zy[r]);
                vy = Chh * ho[1][r] - Che * (exk[r] - zx[r]);
                vz = Chh * ho[2][r];
                if constexpr (YG) { vx += vzy[r + 1] - vzy[r]; vz += vxy[r] - vxy[r + 1]; }
                else { vx -= Che * dEz_y * iky_h; vz += Che * dEx_y * iky_h; }
                if constexpr (XG) { vy += vzxc[r] - vzxn[r]; vz += vyxn[r] - vyxc[r]; }
                else { vy += Che * dEz_x * ikx_h; vz -= Che * dEy_x * ikx_h; }
            } else if constexpr (ADJ) {
                vx = Chh * ho[0][r] + Che * ((eyk[r] - zy[r]) - dEz_y * iky_h);
                vy = Chh * ho[1][r] + Che * (dEz_x * ikx_h - (exk[r] - zx[r]));
                vz = Chh * ho[2][r] + Che * (dEx_y * iky_h - dEy_x * ikx_h);
            } else {
                vx = Chh * ho[0][r] + Che * (dEy_z * ikz_h - dEz_y * iky_h);
                vy = Chh * ho[1][r] + Che * (dEz_x * ikx_h - dEx_z * ikz_h);
                vz = Chh * ho[2][r] + Che * (dEx_y * iky_h - dEy_x * ikx_h);
            }
            // per component the psi terms add in the reference's order:
            // hx +z -y, hy -z +x, hz -x +y (main.cu:834-1037)
            if (!ADJ && zh && iA && jA(r) && lzh >= 0) {
                const long q = ((long)ip * NY + j) * Lzh + lzh;
                const float a = bzh * P.hxz[r] + czh * dEy_z;
                const float b = bzh * P.hyz[r] + czh * dEx_z;
                vx += Che * a;
                vy -= Che * b;
                if (own) { psi(hout + 0)[q] = a; psi(hout + 2)[q] = b; }
            }
            const int ly = lyh(r);
            float pyz = 0.f;
            if (!ADJ && YG && iA && ly >= 0 && kA) {
                const long q = ((long)ip * Lyh + ly) * NZ + k;
                const float by = coef(1, 2, ly), cyv = coef(1, 3, ly);
                const float a = by * P.hxy[r] + cyv * dEz_y;
                pyz = by * P.hzy[r] + cyv * dEx_y;
                vx -= Che * a;
                if (own) { psi(hout + 1)[q] = a; psi(hout + 5)[q] = pyz; }
            }
            if (!ADJ && XG && lx >= 0 && jA(r) && kA) {
                const long q = ((long)lx * NY + j) * NZ + k;
                const float bx = coef(0, 2, lx), cxv = coef(0, 3, lx);
                const float a = bx * P.hyx[r] + cxv * dEz_x;
                const float b = bx * P.hzx[r] + cxv * dEy_x;
                vy += Che * a;
                vz -= Che * b;
                if (own) { psi(hout + 3)[q] = a; psi(hout + 4)[q] = b; }
            }
            if (!ADJ && YG && iA && ly >= 0 && kA) vz += Che * pyz;
            hx[r] = jA(r) && kA ? vx : ho[0][r];
            hy[r] = iA && jB(r) && kA ? vy : ho[1][r];
            hz[r] = iA && jA(r) && kB ? vz : ho[2][r];
        }
    };

    // Register sets: E^n of one plane (rows 0..R, and ex, ez of row R+1) and
    // H^{n-1/2} of one plane.  Step i uses E of planes i and i+1 and H of
    // plane i, and loads E of plane i+2 and H of plane i+1 for later steps;
    // the x march is unrolled six times over three E sets and two H sets so no
    // set is copied while its loads are in flight.  psi of plane i is loaded at
    // the top of step i (one set).
    struct ESet { float x[R + 1], y[R + 1], z[R + 1], xj, zj; };
    struct HSet { float h[3][R + 1]; };
    float hyP[R + 1], hzP[R + 1];                         // H^{n+1/2} of plane i-1 (adjoint x slabs: Wyx, Wzx)
    float vyxC[R + 1], vzxC[R + 1];                       // adjoint x slabs: Vyx, Vzx of plane i
    auto load_e = [&](int ip, ESet &e) {
        const unsigned pl = (unsigned)(ip + 1) * sxy4;
#pragma unroll
        for (int r = 0; r <= R; r++) { e.x[r] = ldb(bEx, pl + ro[r]); e.y[r] = ldb(bEy, pl + ro[r]); e.z[r] = ldb(bEz, pl + ro[r]); }
        e.xj = ldb(bEx, pl + ro[R + 1]);
        e.zj = ldb(bEz, pl + ro[R + 1]);
    };
    auto load_h = [&](int ip, HSet &h) {
        const unsigned pl = (unsigned)(ip + 1) * sxy4;
#pragma unroll
        for (int r = 0; r <= R; r++) { h.h[0][r] = ldb(bHx, pl + ro[r]); h.h[1][r] = ldb(bHy, pl + ro[r]); h.h[2][r] = ldb(bHz, pl + ro[r]); }
    };
    ESet EA, EB, EC;
    HSet H0, H1;
    PSet P0, P1;   // P1 only for the prologue plane
    // prologue: H^{n+1/2} of plane i0-1 (not needed at i0 = 0: ey, ez are not
    // updated there), E^n of planes i0 and i0+1, H and psi of plane i0
#pragma unroll
    for (int r = 0; r <= R; r++) { hyP[r] = 0.f; hzP[r] = 0.f; vyxC[r] = 0.f; vzxC[r] = 0.f; }
    load_e(i0, EA);
    if (i0 > 0) {
        ESet P;
        float pxk[R + 1], pyk[R + 1], hx[R + 1], hy[R + 1], hz[R + 1];
        float vxy[R + 2], vzy[R + 2], vyx[R + 1], vzx[R + 1], pyx[R + 1], pzx[R + 1];
        load_e(i0 - 1, P);
        load_h(i0 - 1, H1);
        load_psi(i0 - 1, P1);
        if constexpr (AS && XG) {
            // psi_E along x of the prologue plane itself (load_psi brings plane i0's)
            const gfloat *e3 = psi(ein + 3), *e5 = psi(ein + 5);
            const int le = max(lxe(i0 - 1), 0);
#pragma unroll
            for (int r = 0; r <= R; r++) {
                const long q = ((long)le * NY + jq(r)) * NZ + kq;
                pyx[r] = e3[q];
                pzx[r] = e5[q];
            }
        }
        e_publish(P.x, P.y, 1);
        lds_sync();
        float zx[R + 1], zy[R + 1];
        if constexpr (ADJ) {
            zweights(i0 - 1, P.x, P.y, P1, false, zx, zy);
            e_up(zx, zy, 1, pxk, pyk);
        } else {
            e_up(P.x, P.y, 1, pxk, pyk);
        }
        if constexpr (AS && YG) vyweights(i0 - 1, P.x, P.z, P.xj, P.zj, P1, false, vxy, vzy);
        if constexpr (AS && XG) {
            vxweights(i0 - 1, P.y, P.z, pyx, pzx, false, vyx, vzx);
            vxweights(i0, EA.y, EA.z, P1.eyx, P1.ezx, true, vyxC, vzxC);   // plane i0 is this chunk's: psi written
        }
        hplane(i0 - 1, P.x, P.y, P.z, P.xj, P.zj, pxk, pyk, EA.y, EA.z, H1.h, P1, false, zx, zy, vxy, vzy, vyx, vzx,
               vyxC, vzxC, hx, hy, hz);
        if constexpr (AS && XG) {
            wxweights(i0 - 1, hy, hz, P1, false, hyP, hzP);
        } else {
#pragma unroll
            for (int r = 0; r <= R; r++) { hyP[r] = hy[r]; hzP[r] = hz[r]; }
        }
    }
    e_publish(EA.x, EA.y, 0);
    load_e(i0 + 1, EB);
    load_h(i0, H0);
    lds_sync();

    // one x step: C = E of plane i, N = E of plane i+1, Fr <- E of plane i+2,
    // Hc = H of plane i, Hn <- H of plane i+1, Pc <- psi of plane i
    auto step = [&](int i, ESet &C, ESet &N, ESet &Fr, HSet &Hc, HSet &Hn, PSet &Pc) {
        const bool more = i + 1 < i1;
        const int buf = (i - i0) & 1;
        const unsigned pl = (unsigned)(i + 1) * sxy4;
        // psi of this plane first (waited for after the field loads issued
        // behind it, not with them), then the fields of the next step; all
        // loads are issued on every step (the last one of the chunk reloads
        // plane i, which is cached) so the wait counts after them are exact
        load_psi(i, Pc);
        load_e(more ? i + 2 : i, Fr);
        load_h(more ? i + 1 : i, Hn);
        float hx[R + 1], hy[R + 1], hz[R + 1], exk[R + 1], eyk[R + 1], zx[R + 1], zy[R + 1];
        if constexpr (ADJ) {
            zweights(i, C.x, C.y, Pc, true, zx, zy);
            e_up(zx, zy, buf, exk, eyk);
        } else {
            e_up(C.x, C.y, buf, exk, eyk);
        }
        float vxy[R + 2], vzy[R + 2], vyxN[R + 1], vzxN[R + 1];
        if constexpr (AS && YG) vyweights(i, C.x, C.z, C.xj, C.zj, Pc, true, vxy, vzy);
        if constexpr (AS && XG) vxweights(i + 1, N.y, N.z, Pc.eyx, Pc.ezx, more, vyxN, vzxN);
        hplane(i, C.x, C.y, C.z, C.xj, C.zj, exk, eyk, N.y, N.z, Hc.h, Pc, true, zx, zy, vxy, vzy, vyxC, vzxC, vyxN,
               vzxN, hx, hy, hz);
        // adjoint: the z differences of the E half are taken of the H-side weights
        float wx[R + 1], wy[R + 1], wxy[R + 1], wzy[R + 1], wyx[R + 1], wzx[R + 1];
        if constexpr (ADJ) hweights(i, hx, hy, Pc, wx, wy);
        if constexpr (AS && YG) wyweights(i, hx, hz, Pc, wxy, wzy);
        if constexpr (AS && XG) wxweights(i, hy, hz, Pc, true, wyx, wzx);
        const float *sy = ADJ ? wy : hy, *sx = ADJ ? wx : hx;
#pragma unroll
        for (int r = 1; r <= R; r++) {
            if (kown && jb + r < jbox) { stn<true>(oHx, pl + ro[r], hx[r]); stn<true>(oHy, pl + ro[r], hy[r]); stn<true>(oHz, pl + ro[r], hz[r]); }
            if (lane == 63) { sK[buf][0][wg][r] = sy[r]; sK[buf][1][wg][r] = sx[r]; }
        }
        e_publish(N.x, N.y, buf ^ 1);
        float hyk[R + 1], hxk[R + 1];
#pragma unroll
        for (int r = 1; r <= R; r++) { hyk[r] = from_lower_lane(sy[r]); hxk[r] = from_lower_lane(sx[r]); }
        lds_sync();
        if (first_lane) {
#pragma unroll
            for (int r = 1; r <= R; r++) { hyk[r] = sK[buf][0][wg - 1][r]; hxk[r] = sK[buf][1][wg - 1][r]; }
        }
        const bool iA = !XG || i < NX - 1, iC = !XG || (i > 0 && i < NX - 1);
        const float ikx_e = XG ? ik(0, true, i, NX) : ra.ik0[0][0];
        const int lx = lxe(i);
#pragma unroll
        for (int r = 1; r <= R; r++) {
            const int j = jb + r;
            const bool own = kown && j < jbox;
            const float iky_e = YG ? ik(1, true, jq(r), NY) : ra.ik0[0][1];
            const float dHz_y = hz[r] - hz[r - 1], dHy_z = hy[r] - hyk[r];
            const float dHx_z = hx[r] - hxk[r], dHz_x = hz[r] - hzP[r];
            const float dHy_x = hy[r] - hyP[r], dHx_y = hx[r] - hx[r - 1];
            float vx, vy, vz;
            if constexpr (AS) {
                // Ehat = Cee Ehat + backward differences of W (W along z from the lane weights)
                vx = Cee * C.x[r] - Ceh * (wy[r] - hyk[r]);
                vy = Cee * C.y[r] + Ceh * (wx[r] - hxk[r]);
                vz = Cee * C.z[r];
                if constexpr (YG) { vx += wzy[r - 1] - wzy[r]; vz += wxy[r] - wxy[r - 1]; }
                else { vx += Ceh * dHz_y * iky_e; vz -= Ceh * dHx_y * iky_e; }
                if constexpr (XG) { vy += wzx[r] - hzP[r]; vz += hyP[r] - wyx[r]; }
                else { vy -= Ceh * dHz_x * ikx_e; vz += Ceh * dHy_x * ikx_e; }
            } else if constexpr (ADJ) {
                vx = Cee * C.x[r] + Ceh * (dHz_y * iky_e - (wy[r] - hyk[r]));
                vy = Cee * C.y[r] + Ceh * ((wx[r] - hxk[r]) - dHz_x * ikx_e);
                vz = Cee * C.z[r] + Ceh * (dHy_x * ikx_e - dHx_y * iky_e);
            } else {
                vx = Cee * C.x[r] + Ceh * (dHz_y * iky_e - dHy_z * ikz_e);
                vy = Cee * C.y[r] + Ceh * (dHx_z * ikz_e - dHz_x * ikx_e);
                vz = Cee * C.z[r] + Ceh * (dHy_x * ikx_e - dHx_y * iky_e);
            }
            const bool ux = iA && jC(r) && kC, uy = iC && jA(r) && kC, uz = iC && jC(r) && kA;
            // ex -z +y, ey +z -x, ez -y +x (main.cu:1172-1906)
            if (!ADJ && ze && iA && lze >= 0) {
                const long q = ((long)i * NY + j) * Lze + lze;
                if (jC(r)) { const float a = bze * Pc.exz[r - 1] + cze * dHy_z; vx -= Ceh * a; if (own) psi(12 + 0)[q] = a; }
                if (iC && jA(r)) { const float a = bze * Pc.eyz[r - 1] + cze * dHx_z; vy += Ceh * a; if (own) psi(12 + 2)[q] = a; }
            }
            const int ly = lye(r);
            if (!ADJ && YG && iA && ly >= 0) {
                const long q = ((long)i * Lye + ly) * NZ + k;
                const float by = coef(1, 0, ly), cyv = coef(1, 1, ly);
                if (kC) { const float a = by * Pc.exy[r - 1] + cyv * dHz_y; vx += Ceh * a; if (own) psi(12 + 1)[q] = a; }
                if (iC && kA) { const float a = by * Pc.ezy[r - 1] + cyv * dHx_y; vz -= Ceh * a; if (own) psi(12 + 4)[q] = a; }
            }
            if (!ADJ && XG && lx >= 0) {
                const long q = ((long)lx * NY + j) * NZ + k;
                const float bx = coef(0, 0, lx), cxv = coef(0, 1, lx);
                if (jA(r) && kC) { const float a = bx * Pc.eyx[r - 1] + cxv * dHz_x; vy -= Ceh * a; if (own) psi(12 + 3)[q] = a; }
                if (jC(r) && kA) { const float a = bx * Pc.ezx[r - 1] + cxv * dHy_x; vz += Ceh * a; if (own) psi(12 + 5)[q] = a; }
            }
            if (own) {
                stn<true>(oEx, pl + ro[r], ux ? vx : C.x[r]);
                stn<true>(oEy, pl + ro[r], uy ? vy : C.y[r]);
                stn<true>(oEz, pl + ro[r], uz ? vz : C.z[r]);
            }
        }
        if constexpr (AS && XG) {
#pragma unroll
            for (int r = 0; r <= R; r++) { hyP[r] = wyx[r]; hzP[r] = wzx[r]; vyxC[r] = vyxN[r]; vzxC[r] = vzxN[r]; }
        } else {
#pragma unroll
            for (int r = 0; r <= R; r++) { hyP[r] = hy[r]; hzP[r] = hz[r]; }
        }
    };
    for (int i = i0;;) {
        step(i, EA, EB, EC, H0, H1, P0); if (++i >= i1) break;
        step(i, EB, EC, EA, H1, H0, P0); if (++i >= i1) break;
        step(i, EC, EA, EB, H0, H1, P0); if (++i >= i1) break;
        step(i, EA, EB, EC, H1, H0, P0); if (++i >= i1) break;
        step(i, EB, EC, EA, H0, H1, P0); if (++i >= i1) break;
        step(i, EC, EA, EB, H1, H0, P0); if (++i >= i1) break;
    }
}

// Two rows per thread fit in 128 VGPRs (four waves per SIMD) on the clean
// box.  The side-slab variants are not capped: their launches hold one block
// per CU at most (a few hundred blocks), so a fourth wave per SIMD gains
// nothing there, while the cap spilled the x-general variant to scratch.
template <int W, int R, bool XG, bool YG, bool ADJ = false, int T = 1>
__global__ __launch_bounds__(64 * W * T) __attribute__((amdgpu_waves_per_eu(R == 2 && !XG && !YG ? 4 : 1))) void k_step_rows(RowsArgs ra) {
    rows_body<W, R, XG, YG, ADJ, T>(ra, ra.xcd ? xcd_order(blockIdx.x, ra.start[ra.npiece]) : (int)blockIdx.x);
}

// ----------------------------------------------------------------- split path kernels

__global__ void k_split_h(StepDev s) {
    const Geom &g = s.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int i = blockIdx.z;
    if (k >= g.NZ) return;
    const long p = gidx(g, i, j, k);
    ENbr e;
    e.ex = s.Ein[0][p]; e.ey = s.Ein[1][p]; e.ez = s.Ein[2][p];
    e.ex_j1 = s.Ein[0][p + g.sz]; e.ez_j1 = s.Ein[2][p + g.sz];
    e.ex_k1 = s.Ein[0][p + 1]; e.ey_k1 = s.Ein[1][p + 1];
    e.ey_i1 = s.Ein[1][p + g.sxy]; e.ez_i1 = s.Ein[2][p + g.sxy];
    float hx, hy, hz;
    h_cell(s, i, j, k, e, s.Hin[0][p], s.Hin[1][p], s.Hin[2][p], true, hx, hy, hz);
    s.Hout[0][p] = hx; s.Hout[1][p] = hy; s.Hout[2][p] = hz;
}

__global__ void k_split_e(StepDev s) {
    const Geom &g = s.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int i = blockIdx.z;
    if (k >= g.NZ) return;
    const long p = gidx(g, i, j, k);
    HNbr h;
    h.hx = s.Hin[0][p]; h.hy = s.Hin[1][p]; h.hz = s.Hin[2][p];
    h.hz_jm = s.Hin[2][p - g.sz]; h.hx_jm = s.Hin[0][p - g.sz];
    h.hy_km = s.Hin[1][p - 1]; h.hx_km = s.Hin[0][p - 1];
    h.hz_im = s.Hin[2][p - g.sxy]; h.hy_im = s.Hin[1][p - g.sxy];
    float ex, ey, ez;
    e_cell(s, i, j, k, h, s.Ein[0][p], s.Ein[1][p], s.Ein[2][p], ex, ey, ez);
    s.Eout[0][p] = ex; s.Eout[1][p] = ey; s.Eout[2][p] = ez;
}

// ----------------------------------------------------------------- periodic x, y (Periodic_XY)
//
// x and y wrap around (the i == NCELLX-1 / j == NCELLY-1 and i == 0 / j == 0
// branches of UPDATE_h* and UPDATE_e*, main.cu:781-1044, 1067-1912, real
// build): every column is updated, the neighbours across the x and y edges
// are read from the ghost layers, which k_wrap_e / k_wrap_h fill with the
// opposite face before each half step, so the stencils need no wrap logic;
// CPML along z only.  With tfsf_on the near z plane of the TFSF box adds the
// incident wave (CORRECT_Z near face, main.cu:638-697; Periodic_XY skips the
// far face, main.cu:698): H at k0-1 after its update, E^n at k0 before.
// Updates run in place (E and H of the current level), as the split path.

// E ghosts for the H update: plane x = NX <- x = 0 (rows 0..NY-1), then row
// y = NY <- y = 0 for x = 0..NX (the corner follows from the ghost plane).
__global__ void k_wrap_e(float *ex, float *ey, float *ez, Geom g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int u = blockIdx.y;
    if (k >= g.NZ) return;
    float *f[3] = {ex, ey, ez};
    if (blockIdx.z == 0) {
        if (u >= g.NY) return;
        const long d = gidx(g, g.NX, u, k), src = gidx(g, 0, u, k);
        for (int c = 0; c < 3; c++) f[c][d] = f[c][src];
    } else {
        if (u > g.NX) return;
        const long d = gidx(g, u, g.NY, k), src = gidx(g, u == g.NX ? 0 : u, 0, k);
        for (int c = 0; c < 3; c++) f[c][d] = f[c][src];
    }
}

// H ghosts for the E update: plane x = -1 <- x = NX-1, row y = -1 <- y = NY-1
// for x = -1..NX-1.
__global__ void k_wrap_h(float *hx, float *hy, float *hz, Geom g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int u = blockIdx.y;
    if (k >= g.NZ) return;
    float *f[3] = {hx, hy, hz};
    if (blockIdx.z == 0) {
        if (u >= g.NY) return;
        const long d = gidx(g, -1, u, k), src = gidx(g, g.NX - 1, u, k);
        for (int c = 0; c < 3; c++) f[c][d] = f[c][src];
    } else {
        if (u > g.NX) return;
        const int i = u - 1;
        const long d = gidx(g, i, -1, k), src = gidx(g, i < 0 ? g.NX - 1 : i, g.NY - 1, k);
        for (int c = 0; c < 3; c++) f[c][d] = f[c][src];
    }
}

__global__ void k_per_h(StepDev s) {
    const Geom &g = s.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= g.NZ) return;
    const long p = gidx(g, i, j, k);
    const CpmlDev &cz = s.cp[2];
    const float Chh = s.c.Chh, Che = s.c.Che;
    const float ikx = s.cp[0].ik_h[i], iky = s.cp[1].ik_h[j];
    const float *ex = s.Ein[0], *ey = s.Ein[1], *ez = s.Ein[2];
    if (k < g.NZ - 1) {
        const int lz = h_layer(cz, k), Lz = 2 * cz.Nc;
        const float ikz = cz.ik_h[k];
        // hx (main.cu:781-865)
        const float dEy = ey[p + 1] - ey[p], dEzj = ez[p + g.sz] - ez[p];
        float h = Chh * s.Hin[0][p] + Che * (dEy * ikz - dEzj * iky);
        if (lz >= 0) h += Che * psi_step(s.psiH_in.p[0], s.psiH_out.p[0], zslab(g, Lz, i, j, lz), cz.bh[lz], cz.ch[lz], dEy, true);
        if (s.tfsf_on) h += tfsf_term_ct<false, 0, 2>(s.tf, i, j, k);
        s.Hout[0][p] = h;
        // hy (main.cu:877-948)
        const float dEzi = ez[p + g.sxy] - ez[p], dEx = ex[p + 1] - ex[p];
        h = Chh * s.Hin[1][p] + Che * (dEzi * ikx - dEx * ikz);
        if (lz >= 0) h -= Che * psi_step(s.psiH_in.p[2], s.psiH_out.p[2], zslab(g, Lz, i, j, lz), cz.bh[lz], cz.ch[lz], dEx, true);
        if (s.tfsf_on) h += tfsf_term_ct<false, 1, 2>(s.tf, i, j, k);
        s.Hout[1][p] = h;
    }
    // hz (main.cu:962-1044), every k
    const float dExj = ex[p + g.sz] - ex[p], dEyi = ey[p + g.sxy] - ey[p];
    s.Hout[2][p] = Chh * s.Hin[2][p] + Che * (dExj * iky - dEyi * ikx);
}

__global__ void k_per_e(StepDev s) {
    const Geom &g = s.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= g.NZ - 1) return;
    const long p = gidx(g, i, j, k);
    const CpmlDev &cz = s.cp[2];
    const float Cee = s.c.Cee, Ceh = s.c.Ceh;
    const float ikx = s.cp[0].ik_e[i], iky = s.cp[1].ik_e[j];
    const float *hx = s.Hin[0], *hy = s.Hin[1], *hz = s.Hin[2];
    if (k > 0) {
        const int lz = e_layer(cz, k), Lz = 2 * (cz.Nc + 1);
        const float ikz = cz.ik_e[k];
        // ex (main.cu:1067-1170)
        float eo = s.Ein[0][p];
        if (s.tfsf_on) eo += tfsf_term_ct<true, 0, 2>(s.tf, i, j, k);
        const float dHzj = hz[p] - hz[p - g.sz], dHy = hy[p] - hy[p - 1];
        float e = Cee * eo + Ceh * (dHzj * iky - dHy * ikz);
        if (lz >= 0) e -= Ceh * psi_step(s.psiE.p[0], s.psiE.p[0], zslab(g, Lz, i, j, lz), cz.be[lz], cz.ce[lz], dHy, true);
        s.Eout[0][p] = e;
        // ey (main.cu:1359-1475)
        eo = s.Ein[1][p];
        if (s.tfsf_on) eo += tfsf_term_ct<true, 1, 2>(s.tf, i, j, k);
        const float dHx = hx[p] - hx[p - 1], dHzi = hz[p] - hz[p - g.sxy];
        e = Cee * eo + Ceh * (dHx * ikz - dHzi * ikx);
        if (lz >= 0) e += Ceh * psi_step(s.psiE.p[2], s.psiE.p[2], zslab(g, Lz, i, j, lz), cz.be[lz], cz.ce[lz], dHx, true);
        s.Eout[1][p] = e;
    }
    // ez (main.cu:1663-1766), k < NZ-1
    const float dHyi = hy[p] - hy[p - g.sxy], dHxj = hx[p] - hx[p - g.sz];
    s.Eout[2][p] = Cee * s.Ein[2][p] + Ceh * (dHyi * ikx - dHxj * iky);
}

// Split path of the periodic mode: the near z plane corrections in place
// (H at k0-1 with e_inc^n, E at k0 with h_inc^{n+1/2}), after the H update.
__global__ void k_per_tfsf(StepDev s) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= s.g.NY) return;
    const int k0 = s.tf.t0[2];
    const long ph = gidx(s.g, i, j, k0 - 1), pe = gidx(s.g, i, j, k0);
    s.Hout[0][ph] += tfsf_term_ct<false, 0, 2>(s.tf, i, j, k0 - 1);
    s.Hout[1][ph] += tfsf_term_ct<false, 1, 2>(s.tf, i, j, k0 - 1);
    s.Eout[0][pe] += tfsf_term_ct<true, 0, 2>(s.tf, i, j, k0);
    s.Eout[1][pe] += tfsf_term_ct<true, 1, 2>(s.tf, i, j, k0);
}

// CORRECT_X/Y/Z in place over the bounding box of the TFSF planes.
__global__ void k_tfsf_split(StepDev s) {
    const TfsfDev &t = s.tf;
    const int k = t.t0[2] - 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t.t0[1] - 1 + blockIdx.y;
    const int i = t.t0[0] - 1 + blockIdx.z;
    if (k > t.t1[2]) return;
    const bool onx = i == t.t0[0] - 1 || i == t.t0[0] || i == t.t1[0];
    const bool ony = j == t.t0[1] - 1 || j == t.t0[1] || j == t.t1[1];
    const bool onz = k == t.t0[2] - 1 || k == t.t0[2] || k == t.t1[2];
    if (!(onx || ony || onz)) return;
    const long p = gidx(s.g, i, j, k);
    for (int c = 0; c < 3; c++) {
        float dh = tfsf_h(t, c, i, j, k);
        if (dh != 0.f) s.Hout[c][p] += dh;
        float de = tfsf_e(t, c, i, j, k);
        if (de != 0.f) s.Eout[c][p] += de;
    }
}

// TFSF corrections after a fused step that left them out (tfsf_on = 0):
// H += dH on the H face layers and E += Cee dE + Ceh curl(dH) on the E cells
// whose update saw those H values.  The reference adds dH to H before the E
// update and dE to E^n (main.cu:2791-2901); the update is linear, so adding
// the propagated terms afterwards gives the same fields up to rounding.
// Cells: per normal axis a the planes t0-1, t0, t1, t1+1 spanning [t0-1, t1+1]
// along the other axes; a cell lying on planes of several axes is taken by
// the lowest one.  Requires the region free of CPML and inside every domain.
__device__ __forceinline__ void tfsf_fix_body(const StepDev &s, int bx, int by, int bz, int lane) {
    const TfsfDev &t = s.tf;
    const int a = bz >> 2, q = bz & 3;
    const int b0 = a == 0 ? 1 : 0, b1 = a == 2 ? 1 : 2;   // other axes, b1 fastest in memory
    const int u = pick3(b0, t.t0[0], t.t0[1], t.t0[2]) - 1 + by;
    const int v = pick3(b1, t.t0[0], t.t0[1], t.t0[2]) - 1 + bx * 64 + lane;
    if (u > pick3(b0, t.t1[0], t.t1[1], t.t1[2]) + 1 || v > pick3(b1, t.t1[0], t.t1[1], t.t1[2]) + 1) return;
    const int t0a = pick3(a, t.t0[0], t.t0[1], t.t0[2]), t1a = pick3(a, t.t1[0], t.t1[1], t.t1[2]);
    const int xa = q == 0 ? t0a - 1 : (q == 1 ? t0a : (q == 2 ? t1a : t1a + 1));
    const int i = a == 0 ? xa : u;
    const int j = a == 1 ? xa : (a == 0 ? u : v);
    const int k = a == 2 ? xa : v;
    auto on = [&](int ax, int x) {
        const int lo = pick3(ax, t.t0[0], t.t0[1], t.t0[2]), hi = pick3(ax, t.t1[0], t.t1[1], t.t1[2]);
        return x == lo - 1 || x == lo || x == hi || x == hi + 1;
    };
    if (a >= 1 && on(0, i)) return;
    if (a >= 2 && on(1, j)) return;
    const long p = gidx(s.g, i, j, k);
    const float dhx = tfsf_h_ct<0>(t, i, j, k), dhy = tfsf_h_ct<1>(t, i, j, k), dhz = tfsf_h_ct<2>(t, i, j, k);
    if (dhx != 0.f) s.Hout[0][p] += dhx;
    if (dhy != 0.f) s.Hout[1][p] += dhy;
    if (dhz != 0.f) s.Hout[2][p] += dhz;
    const float Cee = s.c.Cee, Ceh = s.c.Ceh;
    const float ikx = s.cp[0].ik_e[i], iky = s.cp[1].ik_e[j], ikz = s.cp[2].ik_e[k];
    const float dz_jm = tfsf_h_ct<2>(t, i, j - 1, k), dx_jm = tfsf_h_ct<0>(t, i, j - 1, k);
    const float dy_km = tfsf_h_ct<1>(t, i, j, k - 1), dx_km = tfsf_h_ct<0>(t, i, j, k - 1);
    const float dz_im = tfsf_h_ct<2>(t, i - 1, j, k), dy_im = tfsf_h_ct<1>(t, i - 1, j, k);
    const float fx = Cee * tfsf_e_ct<0>(t, i, j, k) + Ceh * ((dhz - dz_jm) * iky - (dhy - dy_km) * ikz);
    const float fy = Cee * tfsf_e_ct<1>(t, i, j, k) + Ceh * ((dhx - dx_km) * ikz - (dhz - dz_im) * ikx);
    const float fz = Cee * tfsf_e_ct<2>(t, i, j, k) + Ceh * ((dhy - dy_im) * ikx - (dhx - dx_jm) * iky);
    if (fx != 0.f) s.Eout[0][p] += fx;
    if (fy != 0.f) s.Eout[1][p] += fy;
    if (fz != 0.f) s.Eout[2][p] += fz;
}

__global__ __launch_bounds__(64) void k_tfsf_fix(StepDev s) { tfsf_fix_body(s, blockIdx.x, blockIdx.y, blockIdx.z, threadIdx.x); }

// One thread per sample column of the E half (H taken at the same cell).
__device__ __forceinline__ void sample_body(const Geom &g, const FaceDev &f, const float *const *E, const float *const *H,
                                            float *row, long col) {
    if (col >= f.mE) return;
    int q = 0;
    while (q < 11 && col >= f.off[q + 1]) q++;
    const long blk = (long)f.na[q] * f.nb[q];
    long r = col - f.off[q];
    const int comp = r >= blk ? 1 : 0;
    r -= comp * blk;
    const int a = (int)(r / f.nb[q]) + f.lo_a[q], b = (int)(r % f.nb[q]) + f.lo_b[q];
    const int axis = (q >> 1) % 3;
    int i, j, k;
    if (axis == 0) { i = f.plane[q]; j = a; k = b; }
    else if (axis == 1) { j = f.plane[q]; i = a; k = b; }
    else { k = f.plane[q]; i = a; j = b; }
    const long p = gidx(g, i, j, k);
    // E_a/E_b and H_a/H_b components per normal axis: x (y,z), y (z,x), z (x,y)
    const int ca = (axis + 1) % 3, cb = (axis + 2) % 3;
    const int c = comp ? cb : ca;
    const float ev = E[c][p], hv = H[c][p];
    row[col] = ev;
    row[f.mE_pad + col] = hv;
    // divergence watch: every writer stores the same word, so no atomic is needed
    if (f.status && !(__builtin_isfinite(ev) && __builtin_isfinite(hv))) *f.status = 1u;
}

__global__ void k_dft_sample_split(StepDev s) {
    sample_body(s.g, s.fd, s.Ein, s.Hin, s.sample_row, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ----------------------------------------------------------------- dispersive box (main.cu:1178-1219)

// The reference advances each pole as P^{n+1} = k1 P^n + k2 P^{n-1} + drive and
// subtracts (k1-1) P^n + k2 P^{n-1} in the E update.  The same recurrence is
// carried here in (P, J = P^n - P^{n-1}) form: (k1-1) P^n + k2 P^{n-1} =
// (k1+k2-1) P^n - k2 J^n, and J^{n+1} = -k2 J^n + (k1+k2-1) P^n + drive.  For
// the Drude pole k1+k2-1 = 0 exactly, so the secularly growing P drops out and
// the FP32 update avoids the near-total cancellation of the two-level form.

__device__ __forceinline__ void disp_body(const DispDev &d, long idx) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    if (idx >= nb) return;
    const int bk = (int)(idx % d.n[2]);
    const int bj = (int)((idx / d.n[2]) % d.n[1]);
    const int bi = (int)(idx / ((long)d.n[2] * d.n[1]));
    const int i = d.lo[0] + bi, j = d.lo[1] + bj, k = d.lo[2] + bk;
    const Geom &g = d.g;
    const long p = gidx(g, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (!d.id[c][idx]) continue;
        float curl;
        if (c == 0) curl = (d.H[2][p] - d.H[2][p - g.sz]) * d.ik_e[1][j] - (d.H[1][p] - d.H[1][p - 1]) * d.ik_e[2][k];
        else if (c == 1) curl = (d.H[0][p] - d.H[0][p - 1]) * d.ik_e[2][k] - (d.H[2][p] - d.H[2][p - g.sxy]) * d.ik_e[0][i];
        else curl = (d.H[1][p] - d.H[1][p - g.sxy]) * d.ik_e[0][i] - (d.H[0][p] - d.H[0][p - g.sz]) * d.ik_e[1][j];
        float *const *st = d.st[c];
        const float en = st[0][idx], en1 = st[1][idx];
        float P[3], J[3];
#pragma unroll
        for (int q = 0; q < 3; q++) { P[q] = st[2 + 2 * q][idx]; J[q] = st[3 + 2 * q][idx]; }
        float a0 = d.a0, a1 = d.a1, a2 = d.a2, r = 1.f;
        float e0[3] = {d.e0[0], d.e0[1], d.e0[2]}, e1[3] = {d.e1[0], d.e1[1], d.e1[2]};
        if (d.rho) {
            // vacuum-to-metal blend of the ADE coefficients at density r
            r = d.rho[idx];
            const float inv = 1.f / (1.f + r * d.u1);
            a0 = d.A0 * inv;
            a1 = (1.f + r * d.u2) * inv;
            a2 = r * d.u3 * inv;
#pragma unroll
            for (int q = 0; q < 3; q++) { e0[q] = d.G0[q] * inv; e1[q] = -d.K2[q] * inv; }
        }
        float e = a0 * curl + a1 * en - a2 * en1;
#pragma unroll
        for (int q = 0; q < 3; q++) e -= e0[q] * P[q] + e1[q] * J[q];
        // NONLOCAL == 1: d_NL grad(div P_d^n) in the E update and the Drude pole (updatefields.cu:876-891)
        const float gdv = d.gd[0] ? d.gd[c][idx] : 0.f;
        e -= d.nl_e * gdv;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            float jn = d.g1[q] * J[q] + d.g0[q] * P[q] + r * (d.p3[q] * e + d.p4[q] * en + d.p5[q] * en1);
            if (q == 0) jn += d.nl_p * gdv;
            st[2 + 2 * q][idx] = P[q] + jn;
            st[3 + 2 * q][idx] = jn;
        }
        if (d.traj[c]) d.traj[c][idx] = e;
        if (d.traj_prev[0][c]) { d.traj_prev[0][c][idx] = en1; d.traj_prev[1][c][idx] = en; }
        st[1][idx] = en;
        st[0][idx] = e;
        d.E[c][p] = e;
    }
}

__global__ void k_dispersive(DispDev d) { disp_body(d, (long)blockIdx.x * blockDim.x + threadIdx.x); }

// NONLOCAL == 1: grad(div P~_d) of the Drude pole at every dispersive cell
// of each component, all neighbours at time level n (Calc_DIV_GRADx/y/z,
// updatefields.cu:1436-1604: the x component is (Px[i+1] - 2 Px[i] +
// Px[i-1])/dx^2 + (Py[i+1,j] - Py[i,j] - Py[i+1,j-1] + Py[i,j-1])/(dx dy) +
// (Pz[i+1,k] - Pz[i,k] - Pz[i+1,k-1] + Pz[i,k-1])/(dx dz), y and z by
// symmetry).  P~_d is zero on non-metal cells (hard wall).  One thread per
// box cell; the box keeps a one-cell halo, so every stencil stays inside.
__global__ void k_nonlocal_gd(DispDev d) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb) return;
    const int bk = (int)(idx % d.n[2]);
    const int bj = (int)((idx / d.n[2]) % d.n[1]);
    const int bi = (int)(idx / ((long)d.n[2] * d.n[1]));
    if (!(bi > 0 && bj > 0 && bk > 0 && bi < d.n[0] - 1 && bj < d.n[1] - 1 && bk < d.n[2] - 1)) return;
    const long oi = (long)d.n[1] * d.n[2], oj = d.n[2], ok = 1;
    const float *P0 = d.st[0][2], *P1 = d.st[1][2], *P2 = d.st[2][2];
    if (d.id[0][idx])
        d.gd[0][idx] = d.nl_i2 * ((P0[idx + oi] - 2.f * P0[idx] + P0[idx - oi]) +
                                  (P1[idx + oi] - P1[idx] - P1[idx + oi - oj] + P1[idx - oj]) +
                                  (P2[idx + oi] - P2[idx] - P2[idx + oi - ok] + P2[idx - ok]));
    if (d.id[1][idx])
        d.gd[1][idx] = d.nl_i2 * ((P1[idx + oj] - 2.f * P1[idx] + P1[idx - oj]) +
                                  (P0[idx + oj] - P0[idx] - P0[idx - oi + oj] + P0[idx - oi]) +
                                  (P2[idx + oj] - P2[idx] - P2[idx + oj - ok] + P2[idx - ok]));
    if (d.id[2][idx])
        d.gd[2][idx] = d.nl_i2 * ((P2[idx + ok] - 2.f * P2[idx] + P2[idx - ok]) +
                                  (P1[idx + ok] - P1[idx] - P1[idx - oj + ok] + P1[idx - oj]) +
                                  (P0[idx + ok] - P0[idx] - P0[idx - oi + ok] + P0[idx - oi]));
}

// The work that follows the field update of a step in one launch, for cell
// sets that are pairwise disjoint (the TFSF fix-up adds to its own cells only
// and reads no field; the dispersive box reads H beside its cells and
// rewrites E in them; the DFT sample reads E and H on the monitor faces):
// blocks [0, nd) run the dispersive box, 256 cells each, the next nt the
// fix-up, one 64-lane row of its planes per wave, the last ns the sample of
// the new fields (s.Eout, s.Hout) into s.sample_row, 256 columns each.
__global__ __launch_bounds__(256) void k_post(StepDev s, DispDev d, int nd, int nbx, int span, int nt, int ns) {
    const int b = (int)blockIdx.x;
    if (b < nd) {
        disp_body(d, (long)b * 256 + threadIdx.x);
        return;
    }
    if (b < nd + nt) {
        const int L = (b - nd) * 4 + (int)(threadIdx.x >> 6);
        if (L >= nbx * span * 12) return;
        tfsf_fix_body(s, L % nbx, (L / nbx) % span, L / (nbx * span), threadIdx.x & 63);
        return;
    }
    if (b < nd + nt + ns)
        sample_body(s.g, s.fd, s.Eout, s.Hout, s.sample_row, (long)(b - nd - nt) * 256 + threadIdx.x);
}

// ----------------------------------------------------------------- hydrodynamic model (main.cu:1916-2276)

__device__ inline bool box_interior(const DispDev &d, long idx, int &bi, int &bj, int &bk) {
    bk = (int)(idx % d.n[2]);
    bj = (int)((idx / d.n[2]) % d.n[1]);
    bi = (int)(idx / ((long)d.n[2] * d.n[1]));
    return bi > 0 && bj > 0 && bk > 0 && bi < d.n[0] - 1 && bj < d.n[1] - 1 && bk < d.n[2] - 1;
}

// The hydrodynamic box stores its state in FP32 (E^n, E^{n-1}, v,
// v^n - v^{n-1}, the CP poles, n~) and does its arithmetic in FP64 with FP64
// coefficients: the model differences nearly equal terms every step (the
// drive d3 E^n + d4 E^{n-1} with d4 = -d3, the curl against the current
// inside the metal), and FP32 arithmetic there builds up in the absorption
// spectrum over tens of thousands of steps (5.4e-4 of FP64 after 40,000 steps
// on the 96^3 copy of the hydrodynamic workload, against 3e-5 this way).

// Momentum update of the electron velocity (UpdateHydroPx/Py/Pz, general
// branch): v^{n+1} = d1 v^n + d2 v^{n-1} + d3 E^n + d4 E^{n-1}
//                  + beta^2/dd (GradDiv v + GradDiv (v n~)) / cbrt(1 + n~),
// every neighbour at time level n.  With d1 + d2 = 1 the first two terms are
// v^n + (d1 - 1)(v^n - v^{n-1}).  Convection and magnetic terms are off, as
// in every configuration of the reference (setup.cu:992-999).
__device__ __forceinline__ void hydro_v_cell(const DispDev &d, long idx) {
    const long oi = (long)d.n[1] * d.n[2], oj = d.n[2], ok = 1;
    const double I2 = d.h_inv_d * d.h_inv_d;
    const float *V0 = d.st[0][2], *V1 = d.st[1][2], *V2 = d.st[2][2];
    const float *N0 = d.nd_prev[0], *N1 = d.nd_prev[1], *N2 = d.nd_prev[2];
#define VN(V, N, o) ((double)(V)[idx + (o)] * (N)[idx + (o)])
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (!d.id[c][idx]) continue;
        double gd, gd2;
        if (c == 0) {
            gd = ((double)V0[idx + oi] - 2.0 * V0[idx] + V0[idx - oi]) + ((double)V1[idx + oi] - V1[idx] - V1[idx + oi - oj] + V1[idx - oj])
               + ((double)V2[idx + oi] - V2[idx] - V2[idx + oi - ok] + V2[idx - ok]);
            gd2 = (VN(V0, N0, oi) - 2.0 * VN(V0, N0, 0) + VN(V0, N0, -oi)) + (VN(V1, N1, oi) - VN(V1, N1, 0) - VN(V1, N1, oi - oj) + VN(V1, N1, -oj))
                + (VN(V2, N2, oi) - VN(V2, N2, 0) - VN(V2, N2, oi - ok) + VN(V2, N2, -ok));
        } else if (c == 1) {
            gd = ((double)V1[idx + oj] - 2.0 * V1[idx] + V1[idx - oj]) + ((double)V0[idx + oj] - V0[idx] - V0[idx - oi + oj] + V0[idx - oi])
               + ((double)V2[idx + oj] - V2[idx] - V2[idx + oj - ok] + V2[idx - ok]);
            gd2 = (VN(V1, N1, oj) - 2.0 * VN(V1, N1, 0) + VN(V1, N1, -oj)) + (VN(V0, N0, oj) - VN(V0, N0, 0) - VN(V0, N0, oj - oi) + VN(V0, N0, -oi))
                + (VN(V2, N2, oj) - VN(V2, N2, 0) - VN(V2, N2, oj - ok) + VN(V2, N2, -ok));
        } else {
            gd = ((double)V2[idx + ok] - 2.0 * V2[idx] + V2[idx - ok]) + ((double)V1[idx + ok] - V1[idx] - V1[idx - oj + ok] + V1[idx - oj])
               + ((double)V0[idx + ok] - V0[idx] - V0[idx - oi + ok] + V0[idx - oi]);
            gd2 = (VN(V2, N2, ok) - 2.0 * VN(V2, N2, 0) + VN(V2, N2, -ok)) + (VN(V1, N1, ok) - VN(V1, N1, 0) - VN(V1, N1, ok - oj) + VN(V1, N1, -oj))
                + (VN(V0, N0, ok) - VN(V0, N0, 0) - VN(V0, N0, ok - oi) + VN(V0, N0, -oi));
        }
        const float *const *st = d.st[c];
        // n~ is a small perturbation: the pressure term's 1/cbrt(1 + n~) is taken in FP32
        const double np = d.nd_prev[c][idx];
        d.vnew[c][idx] = (float)((double)st[2][idx] + d.h_d1m1 * st[3][idx] + (d.h_d3 * st[0][idx] + d.h_d4 * st[1][idx])
                                 + d.h_nl * I2 * (gd + gd2) * (double)(1.f / cbrtf(1.f + (float)np)));
    }
#undef VN
}

// Continuity sum of component c at box cell idx (main.cu:1224-1239,
// 1537-1554, 1818-1837): n~^{n+1} = n~^{n-1} - 2 dt/dx * (div(n~ v) + div v),
// the normal-axis terms halved.  P and N are the velocity and density of each
// component at the level the reference's x -> y -> z order reads: the
// components before c at n+1, c and the later ones at n.  The z update keeps
// the reference's 0.5 (Vz1 - Vz2) in place of the density-weighted normal
// term (main.cu:1837); that term is in absolute density units there, so here
// it is scaled by 1/N_EQ.
__device__ __forceinline__ double hydro_cont_sum(const DispDev &d, int c, long idx, const float *Px, const float *Py,
                                                 const float *Pz, const float *Nx, const float *Ny, const float *Nz) {
    const long oi = (long)d.n[1] * d.n[2], oj = d.n[2], ok = 1;
    double r;
    if (c == 0) {
        const double Vx1 = Px[idx + oi], Vx2 = Px[idx - oi], Nx1 = Nx[idx + oi], Nx2 = Nx[idx - oi];
        const double Vy1 = 0.5 * ((double)Py[idx + oi] + Py[idx]), Vy2 = 0.5 * ((double)Py[idx + oi - oj] + Py[idx - oj]);
        const double Ny1 = 0.5 * ((double)Ny[idx + oi] + Ny[idx]), Ny2 = 0.5 * ((double)Ny[idx + oi - oj] + Ny[idx - oj]);
        const double Vz1 = 0.5 * ((double)Pz[idx + oi] + Pz[idx]), Vz2 = 0.5 * ((double)Pz[idx + oi - ok] + Pz[idx - ok]);
        const double Nz1 = 0.5 * ((double)Nz[idx + oi] + Nz[idx]), Nz2 = 0.5 * ((double)Nz[idx + oi - ok] + Nz[idx - ok]);
        r = 0.5 * (Nx1 * Vx1 - Nx2 * Vx2) + (Ny1 * Vy1 - Ny2 * Vy2) + (Nz1 * Vz1 - Nz2 * Vz2)
          + (0.5 * (Vx1 - Vx2) + (Vy1 - Vy2) + (Vz1 - Vz2));
    } else if (c == 1) {
        const double Vy1 = Py[idx + oj], Vy2 = Py[idx - oj], Ny1 = Ny[idx + oj], Ny2 = Ny[idx - oj];
        const double Vx1 = 0.5 * ((double)Px[idx + oj] + Px[idx]), Vx2 = 0.5 * ((double)Px[idx - oi + oj] + Px[idx - oi]);
        const double Nx1 = 0.5 * ((double)Nx[idx + oj] + Nx[idx]), Nx2 = 0.5 * ((double)Nx[idx - oi + oj] + Nx[idx - oi]);
        const double Vz1 = 0.5 * ((double)Pz[idx + oj] + Pz[idx]), Vz2 = 0.5 * ((double)Pz[idx + oj - ok] + Pz[idx - ok]);
        const double Nz1 = 0.5 * ((double)Nz[idx + oj] + Nz[idx]), Nz2 = 0.5 * ((double)Nz[idx + oj - ok] + Nz[idx - ok]);
        r = (Nx1 * Vx1 - Nx2 * Vx2) + 0.5 * (Ny1 * Vy1 - Ny2 * Vy2) + (Nz1 * Vz1 - Nz2 * Vz2)
          + ((Vx1 - Vx2) + 0.5 * (Vy1 - Vy2) + (Vz1 - Vz2));
    } else {
        const double Vz1 = Pz[idx + ok], Vz2 = Pz[idx - ok];
        const double Vx1 = 0.5 * ((double)Px[idx + ok] + Px[idx]), Vx2 = 0.5 * ((double)Px[idx - oi + ok] + Px[idx - oi]);
        const double Nx1 = 0.5 * ((double)Nx[idx + ok] + Nx[idx]), Nx2 = 0.5 * ((double)Nx[idx - oi + ok] + Nx[idx - oi]);
        const double Vy1 = 0.5 * ((double)Py[idx + ok] + Py[idx]), Vy2 = 0.5 * ((double)Py[idx - oj + ok] + Py[idx - oj]);
        const double Ny1 = 0.5 * ((double)Ny[idx + ok] + Ny[idx]), Ny2 = 0.5 * ((double)Ny[idx - oj + ok] + Ny[idx - oj]);
        r = (Nx1 * Vx1 - Nx2 * Vx2) + (Ny1 * Vy1 - Ny2 * Vy2) + 0.5 * (Vz1 - Vz2) * d.h_inv_neq
          + ((Vx1 - Vx2) + (Vy1 - Vy2) + 0.5 * (Vz1 - Vz2));
    }
    return r;
}

// Continuity and E update of component c at one cell (main.cu:1224-1271
// and the y, z equivalents): n~^{n+1} of the cell, then
// E^{n+1} = a0 curl + h_cur v^{n+1} (1 + n~^{n+1}) + a1 E^n - a2 E^{n-1} - CP terms,
// then the CP poles advance.  Nothing a launch of it reads at a neighbour is
// written by that launch: the components before c are read at n+1 from vnew
// and nd (written by the earlier launches of the step), c and the later ones
// at n from the velocity state and nd_prev, which no launch of the step
// writes; the cell's own n~ goes from n-1 to n+1 in nd.  The level shift
// (velocity state <- vnew, nd_prev <-> nd) is a swap of the buffers' roles on
// the host after the step's launches.
__device__ __forceinline__ void hydro_ce_cell(const DispDev &d, int c, long idx, int bi, int bj, int bk) {
    const float *V[3], *N[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        V[q] = q < c ? d.vnew[q] : d.st[q][2];
        N[q] = q < c ? d.nd[q] : d.nd_prev[q];
    }
    const double r = hydro_cont_sum(d, c, idx, V[0], V[1], V[2], N[0], N[1], N[2]);
    const float ndf = (float)((double)d.nd[c][idx] - d.h_cont * r);
    d.nd[c][idx] = ndf;
    const int i = d.lo[0] + bi, j = d.lo[1] + bj, k = d.lo[2] + bk;
    const Geom &g = d.g;
    const long p = gidx(g, i, j, k);
    const float *const *H = d.H;
    double curl;
    if (c == 0) curl = ((double)H[2][p] - H[2][p - g.sz]) * d.ik_e[1][j] - ((double)H[1][p] - H[1][p - 1]) * d.ik_e[2][k];
    else if (c == 1) curl = ((double)H[0][p] - H[0][p - 1]) * d.ik_e[2][k] - ((double)H[2][p] - H[2][p - g.sxy]) * d.ik_e[0][i];
    else curl = ((double)H[1][p] - H[1][p - g.sxy]) * d.ik_e[0][i] - ((double)H[0][p] - H[0][p - g.sz]) * d.ik_e[1][j];
    float *const *st = d.st[c];
    const double en = st[0][idx], en1 = st[1][idx];
    const double vn = d.vnew[c][idx], ndn = ndf;
    double P[2], J[2];
#pragma unroll
    for (int q = 1; q < 3; q++) { P[q - 1] = st[2 + 2 * q][idx]; J[q - 1] = st[3 + 2 * q][idx]; }
    double e = d.ha0 * curl + d.h_cur * vn * (1.0 + ndn) + d.ha1 * en - d.ha2 * en1;
#pragma unroll
    for (int q = 1; q < 3; q++) e -= d.he0[q] * P[q - 1] + d.he1[q] * J[q - 1];
#pragma unroll
    for (int q = 1; q < 3; q++) {
        const double jn = d.hg1[q] * J[q - 1] + d.hg0[q] * P[q - 1] + d.hp3[q] * e + d.hp4[q] * en + d.hp5[q] * en1;
        st[2 + 2 * q][idx] = (float)(P[q - 1] + jn);
        st[3 + 2 * q][idx] = (float)jn;
    }
    st[3][idx] = (float)(vn - (double)st[2][idx]);
    st[1][idx] = (float)en;
    st[0][idx] = (float)e;
    d.E[c][p] = (float)e;
}

// The momentum update of every component and the x continuity + E update in
// one launch: the x update reads the new velocity only at its own cell (the
// same thread computed it just before), and at neighbours only level-n values,
// which no launch of the step writes.
__global__ void k_hydro_v_ce0(DispDev d) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb) return;
    int bi, bj, bk;
    if (!box_interior(d, idx, bi, bj, bk)) return;
    hydro_v_cell(d, idx);
    if (d.id[0][idx]) hydro_ce_cell(d, 0, idx, bi, bj, bk);
}

// Continuity + E update of component c = 1 (y) or 2 (z).
__global__ void k_hydro_ce(DispDev d, int c) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb) return;
    int bi, bj, bk;
    if (!box_interior(d, idx, bi, bj, bk) || !d.id[c][idx]) return;
    hydro_ce_cell(d, c, idx, bi, bj, bk);
}

// Copies E^n from the field into the box state (after a field overwrite) or
// back (from_field = 0).
__global__ void k_disp_sync(DispDev d, int from_field) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb) return;
    const int bk = (int)(idx % d.n[2]);
    const int bj = (int)((idx / d.n[2]) % d.n[1]);
    const int bi = (int)(idx / ((long)d.n[2] * d.n[1]));
    const long p = gidx(d.g, d.lo[0] + bi, d.lo[1] + bj, d.lo[2] + bk);
    for (int c = 0; c < 3; c++) {
        if (!d.id[c][idx]) continue;
        if (from_field) d.st[c][0][idx] = d.E[c][p];
        else d.E[c][p] = d.st[c][0][idx];
    }
}

// ----------------------------------------------------------------- 1D incident line

// One block: h_inc (main.cu:2328-2336), e_inc (main.cu:2280-2326) into the
// other buffer, first-order Mur ends and the additive source (main.cu:2831-2859).
__device__ __forceinline__ void line_body(const LineDev &l, int t) {
    const int L = l.L;
    for (int i = threadIdx.x; i < L - 1; i += blockDim.x) l.h[i] = l.h[i] - l.ch * (l.e_old[i + 1] - l.e_old[i]);
    __syncthreads();
    for (int i = threadIdx.x + 1; i < L - 1; i += blockDim.x) l.e_new[i] = l.e_old[i] - l.ce * (l.h[i] - l.h[i - 1]);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double e1 = l.mur[0], e2 = l.mur[1];
        l.e_new[0] = e1 + l.mur_coef * (l.e_new[1] - l.e_old[0]);
        l.e_new[L - 1] = e2 + l.mur_coef * (l.e_new[L - 2] - l.e_old[L - 1]);
        l.mur[0] = l.e_new[1];
        l.mur[1] = l.e_new[L - 2];
        double pulse;
        if (l.static_field) {
            pulse = (t < 1000) ? l.amp * t / 1000.0 : l.amp;  // main.cu:2852-2857
        } else {
            // powf on a float-converted argument, as the reference (main.cu:2848)
            const float b = (float)(((double)t - l.delay) / l.width);
            const float b2 = b * b;
            pulse = l.amp * exp(-(double)b2 / 2.0) * sin(l.w0dt * (double)t);
        }
        l.e_new[l.src] += pulse;
        if (l.probe && l.status && !__builtin_isfinite(*l.probe)) *l.status = 1u;
    }
}

__global__ __launch_bounds__(1024) void k_line(LineDev l) { line_body(l, l.t); }

// ----------------------------------------------------------------- DFT contraction

struct SlotTimes {
    int t[256];
};

// W[half][s][r]: half 0 = E (phase (t+1)dt), 1 = H (phase (t+1/2)dt); r < F cos, r >= F sin.
__global__ void k_twiddle(double *W, SlotTimes ts, int K, int nvalid, int F, const double *freq, double dt) {
    const int n = 2 * K * 2 * F;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x) {
        const int r = x % (2 * F);
        const int sidx = (x / (2 * F)) % K;
        const int half = x / (2 * F * K);
        double v = 0.0;
        if (sidx < nvalid) {
            const int f = r % F;
            const double tt = ts.t[sidx] + (half == 0 ? 1.0 : 0.5);
            const double ph = 2.0 * kPI * tt * dt * freq[f];
            v = r < F ? cos(ph) : sin(ph);
        }
        W[x] = v;
    }
}

// acc[r][col] += sum_s X[s][col] * W[half(col)][s][r]; FP64 accumulate.
// Block: 256 columns x one chunk of 16 rows r.
constexpr int kRC = 16;
__global__ __launch_bounds__(256) void k_dft_contract(const float *X, const double *W, double *acc, long mtot,
                                                      long mE_pad, int K, int F) {
    __shared__ double sW[256 * kRC];
    const long col0 = (long)blockIdx.x * 256;
    const int r0 = blockIdx.y * kRC;
    const int half = col0 >= mE_pad ? 1 : 0;
    const double *Wh = W + (long)half * K * 2 * F;
    for (int x = threadIdx.x; x < K * kRC; x += 256) {
        const int sidx = x / kRC, rr = x % kRC;
        sW[x] = (r0 + rr < 2 * F) ? Wh[(long)sidx * 2 * F + r0 + rr] : 0.0;
    }
    __syncthreads();
    const long col = col0 + threadIdx.x;
    if (col >= mtot) return;
    double a[kRC];
#pragma unroll
    for (int rr = 0; rr < kRC; rr++) a[rr] = 0.0;
    for (int sidx = 0; sidx < K; sidx++) {
        const double x = (double)X[(long)sidx * mtot + col];
#pragma unroll
        for (int rr = 0; rr < kRC; rr++) a[rr] += x * sW[sidx * kRC + rr];
    }
#pragma unroll
    for (int rr = 0; rr < kRC; rr++)
        if (r0 + rr < 2 * F) acc[(long)(r0 + rr) * mtot + col] += a[rr];
}

// ----------------------------------------------------------------- MFMA split-BF16 contraction

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Round-to-nearest-even FP32 -> BF16 (finite inputs).
__device__ inline __bf16 to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (unsigned short)(u >> 16));
}
__device__ inline float from_bf16(__bf16 b) {
    return __uint_as_float((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
}

// Split of a value into P BF16 parts v ~ p0 + p1 + ... (each the rounded
// remainder of the previous ones): two parts carry ~16 significant bits,
// three ~24 (FP32).
template <int P>
__device__ inline void split_bf16(double v, __bf16 *out) {
    double r = v;
#pragma unroll
    for (int q = 0; q < P; q++) {
        out[q] = to_bf16((float)r);
        r -= (double)from_bf16(out[q]);
    }
}

// Twiddles in the A-operand fragment order of v_mfma_f32_32x32x16_bf16:
// lane l of m-tile mt and k-step ks holds W[s][r] for r = 32 mt + (l % 32) and
// s = 16 ks + 8 (l / 32) + j, j = 0..7, split into P BF16 parts.
// Wf index: (((half * nks + ks) * nmt + mt) * P + part) * 64 + lane.
template <int P>
__global__ void k_twiddle_frag(bf16x8 *Wf, SlotTimes ts, int nks, int nmt, int nvalid, int F, const double *freq,
                               double dt) {
    const int n = 2 * nks * nmt * 64;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int l = x % 64, mt = (x / 64) % nmt, ks = (x / (64 * nmt)) % nks, half = x / (64 * nmt * nks);
    const int r = 32 * mt + (l % 32);
    bf16x8 part[P];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int sidx = 16 * ks + 8 * (l / 32) + j;
        double v = 0.0;
        if (sidx < nvalid && r < 2 * F) {
            const int f = r % F;
            const double tt = ts.t[sidx] + (half == 0 ? 1.0 : 0.5);
            const double ph = 2.0 * kPI * tt * dt * freq[f];
            v = r < F ? cos(ph) : sin(ph);
        }
        __bf16 b[P];
        split_bf16<P>(v, b);
#pragma unroll
        for (int q = 0; q < P; q++) part[q][j] = b[q];
    }
    const long base = ((((long)half * nks + ks) * nmt + mt) * P) * 64 + l;
#pragma unroll
    for (int q = 0; q < P; q++) Wf[base + 64 * q] = part[q];
}

// acc[r][col] += sum_s W[half][s][r] X[s][col] with split-BF16 products
// accumulated in FP32 per block, then added to the FP64 accumulator.  P = 2:
// hi*hi + hi*lo + lo*hi (relative error ~2^-16); P = 3: the six products of
// total order <= 2 of the three-part splits (~2^-24, FP32 level).  One wave
// = 32 columns x up to 4 m-tiles (128 rows); grid.y walks groups of 4
// m-tiles.  B operand: lane l holds column l % 32, slots 16 ks + 8 (l / 32)
// + j.  C/D: lane l holds column l % 32, rows 8 i + 4 (l / 32) + j in
// element 4 i + j.
template <int P>
__global__ __launch_bounds__(256) void k_dft_mfma(const float *X, const bf16x8 *Wf, double *acc, long mtot,
                                                 long mE_pad, int nks, int nmt, int rows) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long col0 = ((long)blockIdx.x * 4 + w) * 32;
    if (col0 >= mtot) return;
    const int half = col0 >= mE_pad ? 1 : 0;
    const int mg = blockIdx.y;
    const int nm = min(4, nmt - 4 * mg);
    const long col = col0 + (l & 31);
    const bool cok = col < mtot;
    f32x16 c[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int e = 0; e < 16; e++) c[q][e] = 0.f;
    for (int ks = 0; ks < nks; ks++) {
        bf16x8 b[P];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int sidx = 16 * ks + 8 * (l >> 5) + j;
            const float x = cok ? X[(long)sidx * mtot + col] : 0.f;
            __bf16 bs[P];
            split_bf16<P>((double)x, bs);
#pragma unroll
            for (int q = 0; q < P; q++) b[q][j] = bs[q];
        }
        const long wb = (((long)half * nks + ks) * nmt + 4 * mg) * P * 64 + l;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (q < nm) {
                bf16x8 a[P];
#pragma unroll
                for (int u = 0; u < P; u++) a[u] = Wf[wb + (long)q * P * 64 + 64 * u];
                c[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[0], c[q], 0, 0, 0);
                c[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[1], c[q], 0, 0, 0);
                c[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[0], c[q], 0, 0, 0);
                if constexpr (P == 3) {
                    c[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[2], c[q], 0, 0, 0);
                    c[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[2], b[0], c[q], 0, 0, 0);
                    c[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[1], c[q], 0, 0, 0);
                }
            }
        }
    }
    if (!cok) return;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q >= nm) continue;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = 32 * (4 * mg + q) + 8 * i + 4 * (l >> 5) + j;
                if (r < rows) acc[(long)r * mtot + col] += (double)c[q][4 * i + j];
            }
    }
}

// Flux through one face at one frequency: sum Re(Ea conj(Hb) - Eb conj(Ha)) * dA (Spectral.cu:737-1132).
__global__ __launch_bounds__(256) void k_flux(const double *acc, FaceDev fd, long mtot, long mE_pad, int F,
                                              double dA, double *out) {
    __shared__ double red[256];
    const int q = blockIdx.x, f = blockIdx.y;
    const long blk = (long)fd.na[q] * fd.nb[q];
    const double *reR = acc + (long)f * mtot, *imR = acc + (long)(F + f) * mtot;
    double sum = 0.0;
    for (long c = threadIdx.x; c < blk; c += 256) {
        const long ca = fd.off[q] + c, cb = ca + blk;
        const double ear = reR[ca], eai = imR[ca], ebr = reR[cb], ebi = imR[cb];
        const double har = reR[mE_pad + ca], hai = imR[mE_pad + ca], hbr = reR[mE_pad + cb], hbi = imR[mE_pad + cb];
        // Re(Ea * conj(Hb)) - Re(Eb * conj(Ha))
        sum += (ear * hbr + eai * hbi) - (ebr * har + ebi * hai);
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q * F + f] = red[0] * dA;
}

// ----------------------------------------------------------------- dense <-> padded field copies

__global__ void k_pack(const float *f, Geom g, float *dense, int to_dense) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= g.NZ) return;
    const long d = ((long)i * g.NY + j) * g.NZ + k;
    const long p = gidx(g, i, j, k);
    if (to_dense) dense[d] = f[p];
    else const_cast<float *>(f)[p] = dense[d];
}

// ----------------------------------------------------------------- discrete adjoint
//
// No reference counterpart.  The forward step is linear in the state for
// fixed densities (TFSF is an external source), so the adjoint applies the
// transposed updates in reverse time.  Weights (derived from the forward
// update equations and their CPML recursions):
//   E side, per E cell with Q = Ceh * Ehat (vacuum) or a0 * Ehat_tot (ADE):
//     Vxy = Q ik_y + c_y (psi^ex_y + Q)   Vxz = Q ik_z - c_z (psi^ex_z - Q)
//     Vyz = Q ik_z + c_z (psi^ey_z + Q)   Vyx = Q ik_x - c_x (psi^ey_x - Q)
//     Vzx = Q ik_x + c_x (psi^ez_x + Q)   Vzy = Q ik_y - c_y (psi^ez_y - Q)
//   H side, per H cell with G = Che * Hhat_tot:
//     Wxz = G ik_z + c_z (psi^hx_z + G)   Wxy = G ik_y - c_y (psi^hx_y - G)
//     Wyx = G ik_x + c_x (psi^hy_x + G)   Wyz = G ik_z - c_z (psi^hy_z - G)
//     Wzy = G ik_y + c_y (psi^hz_y + G)   Wzx = G ik_x - c_x (psi^hz_x - G)
//   Hhat_tot = Hhat' + forward differences of V; Ehat = base + backward
//   differences of W; each psi adjoint becomes b * (its bracket).

__device__ inline bool e_dom(const Geom &g, int c, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0) return false;
    if (c == 0) return i < g.NX - 1 && j > 0 && j < g.NY - 1 && k > 0 && k < g.NZ - 1;
    if (c == 1) return i > 0 && i < g.NX - 1 && j < g.NY - 1 && k > 0 && k < g.NZ - 1;
    return i > 0 && i < g.NX - 1 && j > 0 && j < g.NY - 1 && k < g.NZ - 1;
}
__device__ inline bool h_dom(const Geom &g, int c, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0) return false;
    if (c == 0) return i < g.NX && j < g.NY - 1 && k < g.NZ - 1;
    if (c == 1) return i < g.NX - 1 && j < g.NY && k < g.NZ - 1;
    return i < g.NX - 1 && j < g.NY - 1 && k < g.NZ;
}

__device__ inline long adj_box(const AdjDev &a, int c, int i, int j, int k) {
    const int bi = i - a.blo[0], bj = j - a.blo[1], bk = k - a.blo[2];
    if (bi < 0 || bj < 0 || bk < 0 || bi >= a.bn[0] || bj >= a.bn[1] || bk >= a.bn[2]) return -1;
    const long b = ((long)bi * a.bn[1] + bj) * a.bn[2] + bk;
    return a.bid[c][b] ? b : -1;
}

// E-side weights of component c at (i,j,k) (va, vb in the order of the table
// above) and the base of its E^n adjoint; wr: owner updates psi_E.
__device__ inline void adj_v(const AdjDev &a, int c, int i, int j, int k, float eh, bool wr, float &va, float &vb,
                             float &base) {
    va = vb = base = 0.f;
    const Geom &g = a.g;
    if (!e_dom(g, c, i, j, k)) return;
    const CpmlDev &cx = a.cp[0], &cy = a.cp[1], &cz = a.cp[2];
    if (a.bn[0] > 0) {
        const long b = adj_box(a, c, i, j, k);
        if (b >= 0) {
            const float Q = a.Qb[c][b];
            base = a.Eb[c][b];
            if (c == 0) { va = Q * cy.ik_e[j]; vb = Q * cz.ik_e[k]; }
            else if (c == 1) { va = Q * cz.ik_e[k]; vb = Q * cx.ik_e[i]; }
            else { va = Q * cx.ik_e[i]; vb = Q * cy.ik_e[j]; }
            return;
        }
    }
    const float Q = a.c.Ceh * eh;
    base = a.c.Cee * eh;
    const bool pin = i < g.NX - 1 && j < g.NY - 1 && k < g.NZ - 1;
    const int lx = pin ? e_layer(cx, i) : -1, ly = pin ? e_layer(cy, j) : -1, lz = pin ? e_layer(cz, k) : -1;
    const int Lz = 2 * (cz.Nc + 1), Ly = 2 * (cy.Nc + 1);
    if (c == 0) {
        va = Q * cy.ik_e[j];
        vb = Q * cz.ik_e[k];
        if (ly >= 0) { const long q = yslab(g, Ly, i, ly, k); const float t = a.psiE_in.p[1][q] + Q; va += cy.ce[ly] * t; if (wr) a.psiE_out.p[1][q] = cy.be[ly] * t; }
        if (lz >= 0) { const long q = zslab(g, Lz, i, j, lz); const float t = a.psiE_in.p[0][q] - Q; vb -= cz.ce[lz] * t; if (wr) a.psiE_out.p[0][q] = cz.be[lz] * t; }
    } else if (c == 1) {
        va = Q * cz.ik_e[k];
        vb = Q * cx.ik_e[i];
        if (lz >= 0) { const long q = zslab(g, Lz, i, j, lz); const float t = a.psiE_in.p[2][q] + Q; va += cz.ce[lz] * t; if (wr) a.psiE_out.p[2][q] = cz.be[lz] * t; }
        if (lx >= 0) { const long q = xslab(g, lx, j, k); const float t = a.psiE_in.p[3][q] - Q; vb -= cx.ce[lx] * t; if (wr) a.psiE_out.p[3][q] = cx.be[lx] * t; }
    } else {
        va = Q * cx.ik_e[i];
        vb = Q * cy.ik_e[j];
        if (lx >= 0) { const long q = xslab(g, lx, j, k); const float t = a.psiE_in.p[5][q] + Q; va += cx.ce[lx] * t; if (wr) a.psiE_out.p[5][q] = cx.be[lx] * t; }
        if (ly >= 0) { const long q = yslab(g, Ly, i, ly, k); const float t = a.psiE_in.p[4][q] - Q; vb -= cy.ce[ly] * t; if (wr) a.psiE_out.p[4][q] = cy.be[ly] * t; }
    }
}

// H-side weights of component c at (i,j,k) from Hhat_tot; returns the H^{n-1/2} adjoint.
__device__ inline float adj_w(const AdjDev &a, int c, int i, int j, int k, float ht, bool wr, float &wa, float &wb) {
    wa = wb = 0.f;
    const Geom &g = a.g;
    if (!h_dom(g, c, i, j, k)) return 0.f;
    const CpmlDev &cx = a.cp[0], &cy = a.cp[1], &cz = a.cp[2];
    const float G = a.c.Che * ht;
    const bool pin = i < g.NX - 1 && j < g.NY - 1 && k < g.NZ - 1;
    const int lx = pin ? h_layer(cx, i) : -1, ly = pin ? h_layer(cy, j) : -1, lz = pin ? h_layer(cz, k) : -1;
    const int Lz = 2 * cz.Nc, Ly = 2 * cy.Nc;
    if (c == 0) {
        wa = G * cz.ik_h[k];
        wb = G * cy.ik_h[j];
        if (lz >= 0) { const long q = zslab(g, Lz, i, j, lz); const float t = a.psiH_in.p[0][q] + G; wa += cz.ch[lz] * t; if (wr) a.psiH_out.p[0][q] = cz.bh[lz] * t; }
        if (ly >= 0) { const long q = yslab(g, Ly, i, ly, k); const float t = a.psiH_in.p[1][q] - G; wb -= cy.ch[ly] * t; if (wr) a.psiH_out.p[1][q] = cy.bh[ly] * t; }
    } else if (c == 1) {
        wa = G * cx.ik_h[i];
        wb = G * cz.ik_h[k];
        if (lx >= 0) { const long q = xslab(g, lx, j, k); const float t = a.psiH_in.p[3][q] + G; wa += cx.ch[lx] * t; if (wr) a.psiH_out.p[3][q] = cx.bh[lx] * t; }
        if (lz >= 0) { const long q = zslab(g, Lz, i, j, lz); const float t = a.psiH_in.p[2][q] - G; wb -= cz.ch[lz] * t; if (wr) a.psiH_out.p[2][q] = cz.bh[lz] * t; }
    } else {
        wa = G * cy.ik_h[j];
        wb = G * cx.ik_h[i];
        if (ly >= 0) { const long q = yslab(g, Ly, i, ly, k); const float t = a.psiH_in.p[5][q] + G; wa += cy.ch[ly] * t; if (wr) a.psiH_out.p[5][q] = cy.bh[ly] * t; }
        if (lx >= 0) { const long q = xslab(g, lx, j, k); const float t = a.psiH_in.p[4][q] - G; wb -= cx.ch[lx] * t; if (wr) a.psiH_out.p[4][q] = cx.bh[lx] * t; }
    }
    return a.c.Chh * ht;
}

// Fused transposed step over one to four x-y boxes of cells (all z).  Tile:
// 64 lanes along z (k = kt*62 - 1 + lane; lanes 0 and 63 are halo) by TJ
// rows along y (j = jt*(TJ-1) - 1 + row; row 0 is halo), marching x.  Each
// thread forms the V weights of its own cell one plane ahead and the W
// weights after the H half; V at k+1 and W at k-1 come from the adjacent
// lane (DPP), W at j-1 and V at j+1 from the adjacent row through LDS (V one
// plane ahead, so one barrier per plane serves both; the top row computes V
// at j+1 itself), V at i+1 and W at i-1 stay in registers.  CPML layers are
// hoisted per lane (z), per row (y) and per plane (x); dispersive-box lookups
// happen only where the tile meets the box.
struct AdjV { float xy, xz, yz, yx, zx, zy, bx, by, bz; };
template <int TJ>
__global__ __launch_bounds__(64 * TJ) void k_adj_step(AdjDev a, int ichunk, int nkt, Pieces pc) {
    __shared__ float sW[2][2][TJ][64];   // Wzy, Wxy (E at j+1 takes them as j-1)
    __shared__ float sV[2][2][TJ][64];   // Vxy, Vzy one plane ahead (H at j-1 takes them as j+1)
    const Geom &g = a.g;
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    int jt, kt, ch, ia, ib, ja, jbx;
    {
        int L = blockIdx.x, q = 0;
        while (q + 1 < pc.n && L >= pc.start[q + 1]) q++;
        L -= pc.start[q];
        jt = L % pc.njt[q];
        kt = (L / pc.njt[q]) % nkt;
        ch = L / (pc.njt[q] * nkt);
        ia = pc.ia[q]; ib = pc.ib[q]; ja = pc.ja[q]; jbx = pc.jb[q];
    }
    const int lane = threadIdx.x, row = threadIdx.y;
    const int k = kt * 62 - 1 + lane;
    const int j = __builtin_amdgcn_readfirstlane(ja + jt * (TJ - 1) - 1 + row);
    const int i0 = ia + ch * ichunk, i1 = min(i0 + ichunk, ib);
    const bool own = lane > 0 && lane < 63 && row > 0 && k < NZ && j < jbx;
    const int kc = min(k, NZ - 1), jc = min(j, NY - 1);
    const int kk = k < NZ ? k : NZ, jj = j < NY ? j : NY;   // beyond the grid: outside every domain
    const long sz = g.sz, sxy = g.sxy;
    const long cidx = (long)jc * sz + kc;
    const CpmlDev &cx = a.cp[0], &cy = a.cp[1], &cz = a.cp[2];
    const float Chh = a.c.Chh, Che = a.c.Che, Cee = a.c.Cee, Ceh = a.c.Ceh;

    // (j,k) parts of the component domains (e_dom, h_dom) and of the psi range
    const bool kin = kk >= 0, jin = jj >= 0;
    const bool ejx = jin && kin && jj > 0 && jj < NY - 1 && kk > 0 && kk < NZ - 1;
    const bool ejy = jin && kin && jj < NY - 1 && kk > 0 && kk < NZ - 1;
    const bool ejz = jin && kin && jj > 0 && jj < NY - 1 && kk < NZ - 1;
    const bool hjx = jin && kin && jj < NY - 1 && kk < NZ - 1;
    const bool hjy = jin && kin && jj < NY && kk < NZ - 1;
    const bool hjz = jin && kin && jj < NY - 1 && kk < NZ;
    const bool pjk = jin && kin && jj < NY - 1 && kk < NZ - 1;
    const int kq = max(kc, 0), jq = max(jc, 0);
    const float ikz_e = cz.ik_e[kq], ikz_h = cz.ik_h[kq], iky_e = cy.ik_e[jq], iky_h = cy.ik_h[jq];
    // z slabs per lane, y slabs per row
    const int lze = pjk ? e_layer(cz, kk) : -1, lzh = pjk ? h_layer(cz, kk) : -1;
    const bool zE = __ballot(lze >= 0) != 0, zH = __ballot(lzh >= 0) != 0;
    const float bze = lze >= 0 ? cz.be[lze] : 0.f, cze = lze >= 0 ? cz.ce[lze] : 0.f;
    const float bzh = lzh >= 0 ? cz.bh[lzh] : 0.f, czh = lzh >= 0 ? cz.ch[lzh] : 0.f;
    const int Lze = 2 * (cz.Nc + 1), Lzh = 2 * cz.Nc, Lye = 2 * (cy.Nc + 1), Lyh = 2 * cy.Nc;
    const int lye = pjk ? e_layer(cy, jj) : -1, lyh = pjk ? h_layer(cy, jj) : -1;   // row-uniform
    const float bye = lye >= 0 ? cy.be[lye] : 0.f, cye = lye >= 0 ? cy.ce[lye] : 0.f;
    const float byh = lyh >= 0 ? cy.bh[lyh] : 0.f, cyh = lyh >= 0 ? cy.ch[lyh] : 0.f;
    const int oz = max(kk, 0);
    // dispersive box along (j,k)
    const bool bjk = a.bn[0] > 0 && jj >= a.blo[1] && jj < a.blo[1] + a.bn[1] && kk >= a.blo[2] && kk < a.blo[2] + a.bn[2];
    const bool tbox = __ballot(bjk) != 0;
    const long bjkoff = bjk ? (long)(jj - a.blo[1]) * a.bn[2] + (kk - a.blo[2]) : 0;
    const float *Ex = a.Ein[0], *Ey = a.Ein[1], *Ez = a.Ein[2];

    // V weights and bases of the own cell at plane ip from Ehat (ex, ey, ez); wr: psi_E written
    auto vcell = [&](int ip, float ex, float ey, float ez, bool wr, AdjV &v) {
        const bool dx = ejx && ip >= 0 && ip < NX - 1, dy = ejy && ip > 0 && ip < NX - 1, dz = ejz && ip > 0 && ip < NX - 1;
        const float ikx_e = cx.ik_e[min(max(ip, 0), NX - 1)];
        float Qx = Ceh * ex, Qy = Ceh * ey, Qz = Ceh * ez;
        float bx = Cee * ex, by = Cee * ey, bz = Cee * ez;
        if (tbox && ip >= a.blo[0] && ip < a.blo[0] + a.bn[0]) {
            if (bjk) {
                const long b = (long)(ip - a.blo[0]) * a.bn[1] * a.bn[2] + bjkoff;
                if (a.bid[0][b]) { Qx = a.Qb[0][b]; bx = a.Eb[0][b]; }
                if (a.bid[1][b]) { Qy = a.Qb[1][b]; by = a.Eb[1][b]; }
                if (a.bid[2][b]) { Qz = a.Qb[2][b]; bz = a.Eb[2][b]; }
            }
        }
        v.xy = Qx * iky_e; v.xz = Qx * ikz_e;
        v.yz = Qy * ikz_e; v.yx = Qy * ikx_e;
        v.zx = Qz * ikx_e; v.zy = Qz * iky_e;
        const bool pin = pjk && ip >= 0 && ip < NX - 1;
        // CPML (dispersive cells never lie in a slab)
        if (pin && lye >= 0) {
            const long q = yslab(g, Lye, ip, lye, oz);
            if (dx) { const float t = a.psiE_in.p[1][q] + Qx; v.xy += cye * t; if (wr) a.psiE_out.p[1][q] = bye * t; }
            if (dz) { const float t = a.psiE_in.p[4][q] - Qz; v.zy -= cye * t; if (wr) a.psiE_out.p[4][q] = bye * t; }
        }
        if (pin && zE && lze >= 0) {
            const long q = zslab(g, Lze, ip, jj, lze);
            if (dx) { const float t = a.psiE_in.p[0][q] - Qx; v.xz -= cze * t; if (wr) a.psiE_out.p[0][q] = bze * t; }
            if (dy) { const float t = a.psiE_in.p[2][q] + Qy; v.yz += cze * t; if (wr) a.psiE_out.p[2][q] = bze * t; }
        }
        const int lxe = pin ? e_layer(cx, ip) : -1;
        if (lxe >= 0) {
            const long q = xslab(g, lxe, jj, oz);
            const float bxe = cx.be[lxe], cxe = cx.ce[lxe];
            if (dy) { const float t = a.psiE_in.p[3][q] - Qy; v.yx -= cxe * t; if (wr) a.psiE_out.p[3][q] = bxe * t; }
            if (dz) { const float t = a.psiE_in.p[5][q] + Qz; v.zx += cxe * t; if (wr) a.psiE_out.p[5][q] = bxe * t; }
        }
        if (!dx) { v.xy = 0.f; v.xz = 0.f; bx = 0.f; }
        if (!dy) { v.yz = 0.f; v.yx = 0.f; by = 0.f; }
        if (!dz) { v.zx = 0.f; v.zy = 0.f; bz = 0.f; }
        v.bx = bx; v.by = by; v.bz = bz;
    };
    // V at j+1 (Vxy of ex, Vzy of ez) for the top row, which has no row above in the tile
    auto vtop = [&](int ip, float &xy, float &zy) {
        xy = zy = 0.f;
        if (row == TJ - 1 && j + 1 < NY && k < NZ) {
            const long p = (long)ip * sxy + cidx + sz;
            float t0, t1;
            adj_v(a, 0, ip, j + 1, kk, Ex[p], false, xy, t0, t1);
            adj_v(a, 2, ip, j + 1, kk, Ez[p], false, t0, zy, t1);
        }
    };
    // W weights from Hhat_tot at plane ip; returns the H^{n-1/2} adjoints in hx, hy, hz
    auto wcell = [&](int ip, float &hx, float &hy, float &hz, bool wr, float &wxz, float &wxy, float &wyx, float &wyz,
                     float &wzy, float &wzx) {
        const bool dx = hjx && ip >= 0 && ip < NX, dy = hjy && ip >= 0 && ip < NX - 1, dz = hjz && ip >= 0 && ip < NX - 1;
        const float ikx_h = cx.ik_h[min(max(ip, 0), NX - 1)];
        const float Gx = Che * hx, Gy = Che * hy, Gz = Che * hz;
        wxz = Gx * ikz_h; wxy = Gx * iky_h;
        wyx = Gy * ikx_h; wyz = Gy * ikz_h;
        wzy = Gz * iky_h; wzx = Gz * ikx_h;
        const bool pin = pjk && ip >= 0 && ip < NX - 1;
        if (pin && zH && lzh >= 0) {
            const long q = zslab(g, Lzh, ip, jj, lzh);
            if (dx) { const float t = a.psiH_in.p[0][q] + Gx; wxz += czh * t; if (wr) a.psiH_out.p[0][q] = bzh * t; }
            if (dy) { const float t = a.psiH_in.p[2][q] - Gy; wyz -= czh * t; if (wr) a.psiH_out.p[2][q] = bzh * t; }
        }
        if (pin && lyh >= 0) {
            const long q = yslab(g, Lyh, ip, lyh, oz);
            if (dx) { const float t = a.psiH_in.p[1][q] - Gx; wxy -= cyh * t; if (wr) a.psiH_out.p[1][q] = byh * t; }
            if (dz) { const float t = a.psiH_in.p[5][q] + Gz; wzy += cyh * t; if (wr) a.psiH_out.p[5][q] = byh * t; }
        }
        const int lxh = pin ? h_layer(cx, ip) : -1;
        if (lxh >= 0) {
            const long q = xslab(g, lxh, jj, oz);
            const float bxh = cx.bh[lxh], cxh = cx.ch[lxh];
            if (dy) { const float t = a.psiH_in.p[3][q] + Gy; wyx += cxh * t; if (wr) a.psiH_out.p[3][q] = bxh * t; }
            if (dz) { const float t = a.psiH_in.p[4][q] - Gz; wzx -= cxh * t; if (wr) a.psiH_out.p[4][q] = bxh * t; }
        }
        if (!dx) { wxz = 0.f; wxy = 0.f; }
        if (!dy) { wyx = 0.f; wyz = 0.f; }
        if (!dz) { wzy = 0.f; wzx = 0.f; }
        hx = dx ? Chh * hx : 0.f;
        hy = dy ? Chh * hy : 0.f;
        hz = dz ? Chh * hz : 0.f;
    };
    auto ld3 = [&](int ip, float &x, float &y, float &z) {
        const long p = (long)ip * sxy + cidx;
        x = Ex[p]; y = Ey[p]; z = Ez[p];
    };

    AdjV vc, vn;
    float e0x, e0y, e0z;
    ld3(i0, e0x, e0y, e0z);
    vcell(i0, e0x, e0y, e0z, own, vc);
    float wzxP = 0.f, wyxP = 0.f;
    sV[0][0][row][lane] = vc.xy;
    sV[0][1][row][lane] = vc.zy;
    if (i0 > 0) {
        // Hhat_tot and W of plane i0-1 (values only): its V at j+1 through LDS, at i+1 = vc
        AdjV vp;
        float px, py, pz;
        ld3(i0 - 1, px, py, pz);
        vcell(i0 - 1, px, py, pz, false, vp);
        sV[1][0][row][lane] = vp.xy;
        sV[1][1][row][lane] = vp.zy;
        __syncthreads();
        float xyJ, zyJ;
        if (row < TJ - 1) { xyJ = sV[1][0][row + 1][lane]; zyJ = sV[1][1][row + 1][lane]; }
        else vtop(i0 - 1, xyJ, zyJ);
        const float yzK = from_upper_lane(vp.yz), xzK = from_upper_lane(vp.xz);
        const long p = (long)(i0 - 1) * sxy + cidx;
        float hx = a.Hin[0][p] + vp.yz - yzK - vp.zy + zyJ;
        float hy = a.Hin[1][p] - vp.xz + xzK + vp.zx - vc.zx;
        float hz = a.Hin[2][p] + vp.xy - xyJ - vp.yx + vc.yx;
        float t0, t1, t2, t3;
        wcell(i0 - 1, hx, hy, hz, false, t0, t1, wyxP, t2, t3, wzxP);
    }
    __syncthreads();
    for (int i = i0; i < i1; i++) {
        const int rb = (i - i0) & 1;
        const long p = (long)i * sxy + cidx;
        const float hinx = a.Hin[0][p], hiny = a.Hin[1][p], hinz = a.Hin[2][p];
        float nx_, ny_, nz_;
        ld3(i + 1, nx_, ny_, nz_);
        vcell(i + 1, nx_, ny_, nz_, own && i + 1 < i1, vn);
        float xyJ, zyJ;
        if (row < TJ - 1) { xyJ = sV[rb][0][row + 1][lane]; zyJ = sV[rb][1][row + 1][lane]; }
        else vtop(i, xyJ, zyJ);
        const float yzK = from_upper_lane(vc.yz), xzK = from_upper_lane(vc.xz);
        float hx = hinx + vc.yz - yzK - vc.zy + zyJ;
        float hy = hiny - vc.xz + xzK + vc.zx - vn.zx;
        float hz = hinz + vc.xy - xyJ - vc.yx + vn.yx;
        float wxz, wxy, wyx, wyz, wzy, wzx;
        wcell(i, hx, hy, hz, own, wxz, wxy, wyx, wyz, wzy, wzx);
        if (own) { a.Hout[0][p] = hx; a.Hout[1][p] = hy; a.Hout[2][p] = hz; }
        sW[rb][0][row][lane] = wzy;
        sW[rb][1][row][lane] = wxy;
        sV[rb ^ 1][0][row][lane] = vn.xy;
        sV[rb ^ 1][1][row][lane] = vn.zy;
        const float wyzK = from_lower_lane(wyz), wxzK = from_lower_lane(wxz);
        __syncthreads();
        if (own) {
            const float wzyJ = sW[rb][0][row - 1][lane], wxyJ = sW[rb][1][row - 1][lane];
            const bool dx = ejx && i < NX - 1, dy = ejy && i > 0 && i < NX - 1, dz = ejz && i > 0 && i < NX - 1;
            a.Eout[0][p] = dx ? vc.bx - wyzK + wyz + wzyJ - wzy : 0.f;
            a.Eout[1][p] = dy ? vc.by + wxzK - wxz - wzxP + wzx : 0.f;
            a.Eout[2][p] = dz ? vc.bz - wxyJ + wxy + wyxP - wyx : 0.f;
        }
        wzxP = wzx; wyxP = wyx;
        vc = vn;
    }
}

// Dispersive cells after the adjoint row kernel (variant 3), which applies
// the vacuum weights everywhere.  At a dispersive E cell the curl weight is
// Q = Qb instead of Ceh Ehat and the base Eb instead of Cee Ehat; the step is
// linear, so the outputs take dV = (Qb - Ceh Ehat) 1/(kappa d) through the
// forward differences of the H half, dH, and then Eb - Cee Ehat plus the
// backward differences of dW = Che dH 1/(kappa d) in the E half (both
// differences per box cell formed by k_adj_disp: dq, db).  One thread per cell
// of the box grown by one cell on every side; the region must be free of CPML
// (checked by the host).
__device__ inline long adj_box_any(const AdjDev &a, int i, int j, int k) {
    const int bi = i - a.blo[0], bj = j - a.blo[1], bk = k - a.blo[2];
    if (bi < 0 || bj < 0 || bk < 0 || bi >= a.bn[0] || bj >= a.bn[1] || bk >= a.bn[2]) return -1;
    return ((long)bi * a.bn[1] + bj) * a.bn[2] + bk;
}
__device__ inline float adj_dq(const AdjDev &a, int c, int i, int j, int k) {
    const long b = adj_box_any(a, i, j, k);
    return b < 0 ? 0.f : a.dq[c][b];
}
__device__ inline void adj_dh(const AdjDev &a, int i, int j, int k, float &hx, float &hy, float &hz) {
    const float ix = a.cp[0].ik_e[i], iy = a.cp[1].ik_e[j], iz = a.cp[2].ik_e[k];
    // dV at the cell and its +1 neighbours (1/(kappa d) of the E positions)
    const float yz = adj_dq(a, 1, i, j, k) * iz, yz1 = adj_dq(a, 1, i, j, k + 1) * a.cp[2].ik_e[k + 1];
    const float zy = adj_dq(a, 2, i, j, k) * iy, zy1 = adj_dq(a, 2, i, j + 1, k) * a.cp[1].ik_e[j + 1];
    const float xz = adj_dq(a, 0, i, j, k) * iz, xz1 = adj_dq(a, 0, i, j, k + 1) * a.cp[2].ik_e[k + 1];
    const float zx = adj_dq(a, 2, i, j, k) * ix, zx1 = adj_dq(a, 2, i + 1, j, k) * a.cp[0].ik_e[i + 1];
    const float xy = adj_dq(a, 0, i, j, k) * iy, xy1 = adj_dq(a, 0, i, j + 1, k) * a.cp[1].ik_e[j + 1];
    const float yx = adj_dq(a, 1, i, j, k) * ix, yx1 = adj_dq(a, 1, i + 1, j, k) * a.cp[0].ik_e[i + 1];
    hx = (yz - yz1) - (zy - zy1);
    hy = -(xz - xz1) + (zx - zx1);
    hz = (xy - xy1) - (yx - yx1);
}
// Correction of one cell (i, j, k) of the box grown by one cell: the H and E
// amounts the dispersive cells add to the vacuum transposed step.
__device__ inline void box_corr(const AdjDev &a, int i, int j, int k, long p, float *h, float *e) {
    float hx, hy, hz;
    adj_dh(a, i, j, k, hx, hy, hz);
    float xm, ym, zm, t0, t1;
    const float Che = a.c.Che;
    const float hx_i = a.cp[0].ik_h[i], hy_i = a.cp[1].ik_h[j], hz_i = a.cp[2].ik_h[k];
    // E half: dW at the cell and its -1 neighbours (1/(kappa d) of the H positions)
    float ex = 0.f, ey = 0.f, ez = 0.f;
    adj_dh(a, i, j, k - 1, xm, ym, t0);          // k-1: dWyz, dWxz
    const float ikzm = a.cp[2].ik_h[k - 1];
    ex += Che * (hy * hz_i - ym * ikzm);
    ey -= Che * (hx * hz_i - xm * ikzm);
    adj_dh(a, i, j - 1, k, xm, t0, zm);          // j-1: dWzy, dWxy
    const float ikym = a.cp[1].ik_h[j - 1];
    ex -= Che * (hz * hy_i - zm * ikym);
    ez += Che * (hx * hy_i - xm * ikym);
    adj_dh(a, i - 1, j, k, t0, ym, zm);          // i-1: dWzx, dWyx
    const float ikxm = a.cp[0].ik_h[i - 1];
    ey += Che * (hz * hx_i - zm * ikxm);
    ez -= Che * (hy * hx_i - ym * ikxm);
    (void)t1;
    const long b = adj_box_any(a, i, j, k);
    if (b >= 0) {
        // zero at cells that are not dispersive
        ex += a.db[0][b];
        ey += a.db[1][b];
        ez += a.db[2][b];
    }
    (void)p;
    h[0] = hx; h[1] = hy; h[2] = hz;
    e[0] = ex; e[1] = ey; e[2] = ez;
}
// Adds a cell's correction to the step's outputs, then the injection where it lands.
__device__ inline void box_apply(const AdjDev &a, const AdjInject &inj, long p, const float *h, const float *e) {
    for (int c = 0; c < 3; c++)
        if (h[c] != 0.f) a.Hout[c][p] += a.c.Chh * h[c];
    // the injection lands after this cell's correction, as a separate k_adj_inject would
    const bool at = inj.on && p == inj.p0;
    for (int c = 0; c < 3; c++)
        if (e[c] != 0.f || at) {
            float v = a.Eout[c][p];
            if (e[c] != 0.f) v += e[c];
            if (at) v += inj.a[c];
            a.Eout[c][p] = v;
        }
}
// Cell of thread idx in the grown box, or the injection outside it (idx == n).
__device__ inline bool box_cell(const AdjDev &a, const AdjInject &inj, long idx, int &i, int &j, int &k) {
    const int nx = a.bn[0] + 2, ny = a.bn[1] + 2, nz = a.bn[2] + 2;
    const long n = (long)nx * ny * nz;
    if (idx == n && inj.on && !inj.inside) {
        // injection outside the corrected cells: as k_adj_inject
        for (int c = 0; c < 3; c++) a.Eout[c][inj.p0] += inj.a[c];
    }
    if (idx >= n) return false;
    i = a.blo[0] - 1 + (int)(idx / ((long)ny * nz));
    j = a.blo[1] - 1 + (int)((idx / nz) % ny);
    k = a.blo[2] - 1 + (int)(idx % nz);
    return true;
}
__global__ void k_adj_box_fix(AdjDev a, AdjInject inj) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!box_cell(a, inj, idx, i, j, k)) return;
    const long p = gidx(a.g, i, j, k);
    float h[3], e[3];
    box_corr(a, i, j, k, p, h, e);
    box_apply(a, inj, p, h, e);
}
// The correction in two launches: its amounts (from the step's inputs and the
// transposed dispersive update alone, so beside the row kernel), then their
// addition after it.  corr: 6 x n floats (H then E, component-major).
__global__ void k_adj_box_corr(AdjDev a, float *corr) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const AdjInject none{};
    int i, j, k;
    if (!box_cell(a, none, idx, i, j, k)) return;
    const long n = (long)(a.bn[0] + 2) * (a.bn[1] + 2) * (a.bn[2] + 2);
    float h[3], e[3];
    box_corr(a, i, j, k, gidx(a.g, i, j, k), h, e);
    for (int c = 0; c < 3; c++) { corr[c * n + idx] = h[c]; corr[(3 + c) * n + idx] = e[c]; }
}
__global__ void k_adj_box_add(AdjDev a, AdjInject inj, const float *corr) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!box_cell(a, inj, idx, i, j, k)) return;
    const long n = (long)(a.bn[0] + 2) * (a.bn[1] + 2) * (a.bn[2] + 2);
    float h[3], e[3];
    for (int c = 0; c < 3; c++) { h[c] = corr[c * n + idx]; e[c] = corr[(3 + c) * n + idx]; }
    box_apply(a, inj, gidx(a.g, i, j, k), h, e);
}

// Adjoint of the (P, J) dispersive update at every dispersive box cell:
// given the adjoint of (E^{n+1}, E^n, P^{n+1}, J^{n+1}) returns that of
// (E^n, E^{n-1}, P^n, J^n), the curl weight Q = a0 Ehat_tot for k_adj_step,
// and accumulates dJ/drho = Ehat_tot dE'/drho + sum_q Jhat_tot (p3 E' + p4 E + p5 E1).
__global__ void k_adj_disp(AdjDispDev A) {
    const DispDev &d = A.d;
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb) return;
    const int bk = (int)(idx % d.n[2]);
    const int bj = (int)((idx / d.n[2]) % d.n[1]);
    const int bi = (int)(idx / ((long)d.n[2] * d.n[1]));
    const long p = gidx(d.g, d.lo[0] + bi, d.lo[1] + bj, d.lo[2] + bk);
    double gsum = 0.0;
    for (int c = 0; c < 3; c++) {
        if (!d.id[c][idx]) {
            if (A.dq[c]) { A.dq[c][idx] = 0.f; A.db[c][idx] = 0.f; }
            continue;
        }
        float a0 = d.a0, a1 = d.a1, a2 = d.a2, r = 1.f, inv = 0.f;
        float e0[3] = {d.e0[0], d.e0[1], d.e0[2]}, e1[3] = {d.e1[0], d.e1[1], d.e1[2]};
        if (d.rho) {
            r = d.rho[idx];
            inv = 1.f / (1.f + r * d.u1);
            a0 = d.A0 * inv;
            a1 = (1.f + r * d.u2) * inv;
            a2 = r * d.u3 * inv;
            for (int q = 0; q < 3; q++) { e0[q] = d.G0[q] * inv; e1[q] = -d.K2[q] * inv; }
        }
        float *const *s = A.a[c];
        float Jt[3];
        const float er = A.Ehat[c][p];
        float et = er;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            Jt[q] = s[2 + 2 * q][idx] + s[1 + 2 * q][idx];   // J' feeds P' = P + J'
            et += r * d.p3[q] * Jt[q];
        }
        float eb = a1 * et + s[0][idx], e1n = -a2 * et;
#pragma unroll
        for (int q = 0; q < 3; q++) { eb += r * d.p4[q] * Jt[q]; e1n += r * d.p5[q] * Jt[q]; }
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float ph = s[1 + 2 * q][idx];
            s[1 + 2 * q][idx] = ph + d.g0[q] * Jt[q] - e0[q] * et;
            s[2 + 2 * q][idx] = d.g1[q] * Jt[q] - e1[q] * et;
        }
        s[0][idx] = e1n;
        const float qb = a0 * et;
        A.Qb[c][idx] = qb;
        A.Eb[c][idx] = eb;
        if (A.dq[c]) {
            A.dq[c][idx] = qb - A.Ceh * er;
            A.db[c][idx] = eb - A.Cee * er;
        }
        if (A.grad && d.rho) {
            const float Ep = A.tp1[c][idx], En = A.t0[c][idx], Em = A.tm1[c][idx];
            const float dE = (d.u2 * En - d.u3 * Em - d.u1 * Ep) * inv;
            double gc = (double)et * dE;
#pragma unroll
            for (int q = 0; q < 3; q++) gc += (double)Jt[q] * (d.p3[q] * Ep + d.p4[q] * En + d.p5[q] * Em);
            gsum += gc;
        }
    }
    if (A.grad && d.rho) A.grad[idx] += gsum;
}

__global__ void k_adj_inject(float *ex, float *ey, float *ez, long p0, float ax, float ay, float az) {
    if (threadIdx.x == 0) { ex[p0] += ax; ey[p0] += ay; ez[p0] += az; }
}

__global__ void k_probe(const float *ex, const float *ey, const float *ez, long p0, float *out) {
    if (threadIdx.x == 0) { out[0] = ex[p0]; out[1] = ey[p0]; out[2] = ez[p0]; }
}

}  // namespace

// ================================================================ launch wrappers

hipError_t launch_fused_step(const StepDev &s, int tj, int ichunk, const int (*box)[4], int nbox, hipStream_t st) {
    if (nbox < 1 || nbox > 4 || ichunk < 1) return hipErrorInvalidValue;
    const int nkt = (s.g.NZ + 62) / 63;
    Pieces pc{};
    pc.n = nbox;
    long total = 0;
    for (int q = 0; q < nbox; q++) {
        pc.ia[q] = box[q][0]; pc.ib[q] = box[q][1]; pc.ja[q] = box[q][2]; pc.jb[q] = box[q][3];
        if (pc.ia[q] < 0 || pc.ib[q] > s.g.NX || pc.ia[q] >= pc.ib[q] || pc.ja[q] < 0 || pc.jb[q] > s.g.NY ||
            pc.ja[q] >= pc.jb[q])
            return hipErrorInvalidValue;
        pc.njt[q] = (pc.jb[q] - pc.ja[q] + tj - 2) / (tj - 1);
        pc.start[q] = (int)total;
        total += (long)pc.njt[q] * nkt * ((pc.ib[q] - pc.ia[q] + ichunk - 1) / ichunk);
    }
    if (total > 0x7fffffff) return hipErrorInvalidValue;
    pc.start[nbox] = (int)total;
    const dim3 grid((unsigned)total);
    switch (tj) {
        case 4: hipLaunchKernelGGL(k_fused_step<4>, grid, dim3(64, 4), 0, st, s, ichunk, nkt, pc); break;
        case 8: hipLaunchKernelGGL(k_fused_step<8>, grid, dim3(64, 8), 0, st, s, ichunk, nkt, pc); break;
        case 16: hipLaunchKernelGGL(k_fused_step<16>, grid, dim3(64, 16), 0, st, s, ichunk, nkt, pc); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}


// ----------------------------------------------------------------- batched launches

// Items are read through the constant address space: every block reads its
// item at a uniform address, so the fields land in scalar registers exactly
// as kernel arguments do.
#define CONST_AS __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const T &batch_item(const T *items, unsigned idx) {
    const CONST_AS T *c = (const CONST_AS T *)items;
    return *(const T *)(c + idx);
}
__device__ __forceinline__ unsigned batch_map(const unsigned *map) {
    const CONST_AS unsigned *c = (const CONST_AS unsigned *)map;
    return c[blockIdx.x];
}

template <int TJ>
__global__ __launch_bounds__(64 * TJ) void k_fused_batch(const FusedItem *items, const unsigned *map) {
    const unsigned m = batch_map(map);
    const FusedItem &it = batch_item(items, m >> kBatchLocalBits);
    fused_body<TJ>(it.s, it.ichunk, it.nkt, it.pc, (int)(m & ((1u << kBatchLocalBits) - 1)));
}

template <int W, int R, bool XG, bool YG, int T>
__global__ __launch_bounds__(64 * W * T) __attribute__((amdgpu_waves_per_eu(R == 2 ? 4 : 1))) void k_rows_batch(const RowsArgs *items, const unsigned *map) {
    const unsigned m = batch_map(map);
    rows_body<W, R, XG, YG, false, T>(batch_item(items, m >> kBatchLocalBits), (int)(m & ((1u << kBatchLocalBits) - 1)));
}

// Local block L of a simulation's TFSF fix-up grid ((span + 63) / 64, span, 12) in x-fastest order.
// Four 64-lane rows of the fix-up's planes per block (one per wave).
__global__ __launch_bounds__(256) void k_tfsf_fix_batch(const FusedItem *items, const unsigned *map) {
    const unsigned m = batch_map(map);
    const FusedItem &it = batch_item(items, m >> kBatchLocalBits);
    const TfsfDev &t = it.s.tf;
    int span = 0;
    for (int a = 0; a < 3; a++) span = max(span, t.t1[a] - t.t0[a] + 3);
    const int nbx = (span + 63) / 64;
    const int L = (int)(m & ((1u << kBatchLocalBits) - 1)) * 4 + (int)(threadIdx.x >> 6);
    if (L >= nbx * span * 12) return;
    tfsf_fix_body(it.s, L % nbx, (L / nbx) % span, L / (nbx * span), threadIdx.x & 63);
}

__global__ __launch_bounds__(1024) void k_line_batch(const LineDev *items, const int *dyn, int step_off) {
    const CONST_AS int *d = (const CONST_AS int *)dyn;
    line_body(batch_item(items, blockIdx.x), d[0] + step_off);
}

__global__ __launch_bounds__(256) void k_disp_batch(const DispDev *items, const unsigned *map) {
    const unsigned m = batch_map(map);
    const DispDev &d = batch_item(items, m >> kBatchLocalBits);
    disp_body(d, (long)(m & ((1u << kBatchLocalBits) - 1)) * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_sample_batch(const SampleItem *items, const unsigned *map, const int *dyn, int sidx) {
    const unsigned m = batch_map(map);
    const SampleItem &it = batch_item(items, m >> kBatchLocalBits);
    const CONST_AS int *d = (const CONST_AS int *)dyn;
    float *row = it.X + (size_t)(d[1] + sidx) * it.mtot;
    sample_body(it.g, it.fd, it.E, it.H, row, (long)(m & ((1u << kBatchLocalBits) - 1)) * 256 + threadIdx.x);
}

__global__ void k_batch_advance(int *dyn, int dt, int dslot) {
    if (threadIdx.x == 0) {
        dyn[0] += dt;
        dyn[1] += dslot;
    }
}

long fused_pieces(const Geom &g, int tj, int ichunk, const int (*box)[4], int nbox, Pieces *pc, int *nkt) {
    if ((tj != 4 && tj != 8 && tj != 16) || ichunk < 1 || nbox < 1 || nbox > 4) return -1;
    Pieces q{};
    q.n = nbox;
    const int nk = (g.NZ + 62) / 63;
    long total = 0;
    for (int b = 0; b < nbox; b++) {
        q.ia[b] = box[b][0]; q.ib[b] = box[b][1]; q.ja[b] = box[b][2]; q.jb[b] = box[b][3];
        if (q.ia[b] < 0 || q.ib[b] > g.NX || q.ia[b] >= q.ib[b] || q.ja[b] < 0 || q.jb[b] > g.NY || q.ja[b] >= q.jb[b])
            return -1;
        q.njt[b] = (q.jb[b] - q.ja[b] + tj - 2) / (tj - 1);
        q.start[b] = (int)total;
        total += (long)q.njt[b] * nk * ((q.ib[b] - q.ia[b] + ichunk - 1) / ichunk);
    }
    if (total > 0x7fffffff) return -1;
    q.start[nbox] = (int)total;
    if (pc) *pc = q;
    if (nkt) *nkt = nk;
    return total;
}

long fused_batch_blocks(const Geom &g, int tj, int ichunk, Pieces *pc, int *nkt) {
    const int full[1][4] = {{0, g.NX, 0, g.NY}};
    return fused_pieces(g, tj, ichunk, full, 1, pc, nkt);
}

hipError_t launch_fused_batch(const FusedItem *items, const unsigned *map, int nblocks, int tj, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    switch (tj) {
        case 4: hipLaunchKernelGGL(k_fused_batch<4>, dim3(nblocks), dim3(64, 4), 0, st, items, map); break;
        case 8: hipLaunchKernelGGL(k_fused_batch<8>, dim3(nblocks), dim3(64, 8), 0, st, items, map); break;
        case 16: hipLaunchKernelGGL(k_fused_batch<16>, dim3(nblocks), dim3(64, 16), 0, st, items, map); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int W>
static hipError_t launch_rows_batch_w(const RowsArgs *items, const unsigned *map, int nblocks, int R, int variant,
                                      hipStream_t st) {
    constexpr int T = W <= 5 ? 2 : 1;   // rows_batch_tiles(W)
    const dim3 grid(nblocks), block(64, W * T);
    if (R == 3 && variant == 0) hipLaunchKernelGGL((k_rows_batch<W, 3, false, false, T>), grid, block, 0, st, items, map);
    else if (R == 2 && variant == 0) hipLaunchKernelGGL((k_rows_batch<W, 2, false, false, T>), grid, block, 0, st, items, map);
    else if (R == 2 && variant == 1) hipLaunchKernelGGL((k_rows_batch<W, 2, true, false, T>), grid, block, 0, st, items, map);
    else if (R == 2 && variant == 2) hipLaunchKernelGGL((k_rows_batch<W, 2, false, true, T>), grid, block, 0, st, items, map);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

int rows_batch_tiles(int W) { return W <= 5 ? 2 : 1; }

hipError_t launch_rows_batch(const RowsArgs *items, const unsigned *map, int nblocks, int W, int R, int variant,
                             hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    switch (W) {
        case 2: return launch_rows_batch_w<2>(items, map, nblocks, R, variant, st);
        case 3: return launch_rows_batch_w<3>(items, map, nblocks, R, variant, st);
        case 4: return launch_rows_batch_w<4>(items, map, nblocks, R, variant, st);
        case 5: return launch_rows_batch_w<5>(items, map, nblocks, R, variant, st);
        case 6: return launch_rows_batch_w<6>(items, map, nblocks, R, variant, st);
        case 8: return launch_rows_batch_w<8>(items, map, nblocks, R, variant, st);
        case 12: return launch_rows_batch_w<12>(items, map, nblocks, R, variant, st);
        case 16: return launch_rows_batch_w<16>(items, map, nblocks, R, variant, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_tfsf_fix_batch(const FusedItem *items, const unsigned *map, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tfsf_fix_batch, dim3(nblocks), dim3(256), 0, st, items, map);
    return hipGetLastError();
}

long tfsf_fix_blocks(const StepDev &s) {
    int span = 0;
    for (int a = 0; a < 3; a++) span = max(span, s.tf.t1[a] - s.tf.t0[a] + 3);
    return ((long)((span + 63) / 64) * span * 12 + 3) / 4;
}

hipError_t launch_line_batch(const LineDev *items, int n, const int *dyn, int step_off, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_line_batch, dim3(n), dim3(1024), 0, st, items, dyn, step_off);
    return hipGetLastError();
}

hipError_t launch_disp_batch(const DispDev *items, const unsigned *map, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_disp_batch, dim3(nblocks), dim3(256), 0, st, items, map);
    return hipGetLastError();
}

hipError_t launch_sample_batch(const SampleItem *items, const unsigned *map, int nblocks, const int *dyn, int sidx,
                               hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sample_batch, dim3(nblocks), dim3(256), 0, st, items, map, dyn, sidx);
    return hipGetLastError();
}

hipError_t launch_batch_advance(int *dyn, int dt, int dslot, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, st, dyn, dt, dslot);
    return hipGetLastError();
}

template <int W, bool XG, bool YG, bool ADJ>
static hipError_t launch_rows_w(const RowsArgs &ra, int R, dim3 grid, hipStream_t st) {
    if (ra.tiles == 2) {
        if constexpr (W * 2 <= 16 && !(XG && YG)) {
            switch (R) {
                case 2: hipLaunchKernelGGL((k_step_rows<W, 2, XG, YG, ADJ, 2>), grid, dim3(64, 2 * W), 0, st, ra); break;
                case 3:
                    if constexpr (!XG && !YG && !ADJ) {
                        hipLaunchKernelGGL((k_step_rows<W, 3, false, false, false, 2>), grid, dim3(64, 2 * W), 0, st, ra);
                        break;
                    }
                    return hipErrorInvalidValue;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
    if (ra.tiles == 3) {
        if constexpr (W * 3 <= 16 && !XG && !YG && !ADJ) {
            switch (R) {
                case 2: hipLaunchKernelGGL((k_step_rows<W, 2, false, false, false, 3>), grid, dim3(64, 3 * W), 0, st, ra); break;
                case 3: hipLaunchKernelGGL((k_step_rows<W, 3, false, false, false, 3>), grid, dim3(64, 3 * W), 0, st, ra); break;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
    switch (R) {
        case 2: hipLaunchKernelGGL((k_step_rows<W, 2, XG, YG, ADJ>), grid, dim3(64, W), 0, st, ra); break;
        case 3:
            if constexpr (!(XG && YG) && (!ADJ || (!XG && !YG))) {
                hipLaunchKernelGGL((k_step_rows<W, 3, XG, YG, ADJ>), grid, dim3(64, W), 0, st, ra);
                break;
            }
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool XG, bool YG, bool ADJ = false>
static hipError_t launch_rows_g(const RowsArgs &ra, int R, int W, dim3 grid, hipStream_t st) {
    switch (W) {
        case 2: return launch_rows_w<2, XG, YG, ADJ>(ra, R, grid, st);
        case 3: return launch_rows_w<3, XG, YG, ADJ>(ra, R, grid, st);
        case 4: return launch_rows_w<4, XG, YG, ADJ>(ra, R, grid, st);
        case 5: return launch_rows_w<5, XG, YG, ADJ>(ra, R, grid, st);
        case 6: return launch_rows_w<6, XG, YG, ADJ>(ra, R, grid, st);
        case 8: return launch_rows_w<8, XG, YG, ADJ>(ra, R, grid, st);
        case 12: return launch_rows_w<12, XG, YG, ADJ>(ra, R, grid, st);
        default: return launch_rows_w<16, XG, YG, ADJ>(ra, R, grid, st);
    }
}

int rows_waves(int NZ) {
    static const int ws[] = {2, 3, 4, 5, 6, 8, 12, 16};
    for (int w : ws)
        if (64 * w - 2 >= NZ + 1) return w;   // one segment: lanes k = -1 .. 64w-2 cover the owned 0 .. NZ-1
    return 16;
}

long rows_prepare(RowsArgs &ra, int R, int variant) {
    if (R < 2 || R > 3 || ra.ichunk < 1 || ra.npiece < 1 || ra.npiece > 4 || variant < 0 || variant > 6) return -1;
    if (variant >= 4 && R != 2) return -1;   // the slab and corner variants of the transposed step are built for two rows
    const int W = rows_waves(ra.NZ);
    if (ra.tiles < 1) ra.tiles = 1;
    if (ra.tiles != 1 && (ra.tiles < 2 || ra.tiles > 3 || ra.tiles * W > 16 || (variant != 0 && ra.tiles != 2) ||
                          (variant >= 3 && R != 2)))
        return -1;
    const bool xg = variant == 1 || variant == 4 || variant == 6, yg = variant == 2 || variant == 5 || variant == 6;
    if (variant == 6 && ra.tiles != 1) return -1;   // the corner variant is built for one row tile
    const int T = ra.tiles;
    ra.nseg = (ra.NZ + 1 + 64 * W - 3) / (64 * W - 2);
    long total = 0;
    for (int q = 0; q < ra.npiece; q++) {
        const int *b = ra.box[q];
        if (b[0] < 0 || b[1] > ra.NX || b[0] >= b[1] || b[2] < 0 || b[3] > ra.NY || b[2] >= b[3]) return -1;
        if (!xg && (b[0] < 1 || b[1] > ra.NX - 1)) return -1;   // x must be clean
        if (!yg && (b[2] < 1 || b[3] > ra.NY - 1)) return -1;   // y must be clean
        ra.start[q] = (int)total;
        total += (long)(((b[3] - b[2] + R - 1) / R + T - 1) / T) * ra.nseg * ((b[1] - b[0] + ra.ichunk - 1) / ra.ichunk);
    }
    if (total > 0x7fffffff) return -1;
    ra.start[ra.npiece] = (int)total;
    return total;
}

hipError_t launch_step_rows(RowsArgs ra, int R, int variant, hipStream_t st) {
    const long total = rows_prepare(ra, R, variant);
    if (total < 0) return hipErrorInvalidValue;
    const int W = rows_waves(ra.NZ);
    const dim3 grid((unsigned)total);
    switch (variant) {
        case 0: return launch_rows_g<false, false>(ra, R, W, grid, st);
        case 1: return launch_rows_g<true, false>(ra, R, W, grid, st);
        case 2: return launch_rows_g<false, true>(ra, R, W, grid, st);
        case 3: return launch_rows_g<false, false, true>(ra, R, W, grid, st);
        case 4: return launch_rows_g<true, false, true>(ra, R, W, grid, st);
        case 5: return launch_rows_g<false, true, true>(ra, R, W, grid, st);
        default: return launch_rows_g<true, true, true>(ra, R, W, grid, st);
    }
}

hipError_t launch_adj_step(const AdjDev &a, int tj, int ichunk, const int (*box)[4], int nbox, hipStream_t st) {
    if (nbox < 1 || nbox > 4 || ichunk < 1 || (tj != 4 && tj != 8 && tj != 16)) return hipErrorInvalidValue;
    const int nkt = (a.g.NZ + 61) / 62;
    Pieces pc{};
    pc.n = nbox;
    long total = 0;
    for (int q = 0; q < nbox; q++) {
        pc.ia[q] = box[q][0]; pc.ib[q] = box[q][1]; pc.ja[q] = box[q][2]; pc.jb[q] = box[q][3];
        if (pc.ia[q] < 0 || pc.ib[q] > a.g.NX || pc.ia[q] >= pc.ib[q] || pc.ja[q] < 0 || pc.jb[q] > a.g.NY ||
            pc.ja[q] >= pc.jb[q])
            return hipErrorInvalidValue;
        pc.njt[q] = (pc.jb[q] - pc.ja[q] + tj - 2) / (tj - 1);
        pc.start[q] = (int)total;
        total += (long)pc.njt[q] * nkt * ((pc.ib[q] - pc.ia[q] + ichunk - 1) / ichunk);
    }
    if (total > 0x7fffffff) return hipErrorInvalidValue;
    pc.start[nbox] = (int)total;
    const dim3 grid((unsigned)total);
    switch (tj) {
        case 4: hipLaunchKernelGGL(k_adj_step<4>, grid, dim3(64, 4), 0, st, a, ichunk, nkt, pc); break;
        case 8: hipLaunchKernelGGL(k_adj_step<8>, grid, dim3(64, 8), 0, st, a, ichunk, nkt, pc); break;
        default: hipLaunchKernelGGL(k_adj_step<16>, grid, dim3(64, 16), 0, st, a, ichunk, nkt, pc); break;
    }
    return hipGetLastError();
}

long adj_box_cells(const AdjDev &a) {
    return a.bn[0] <= 0 ? 0 : (long)(a.bn[0] + 2) * (a.bn[1] + 2) * (a.bn[2] + 2);
}

hipError_t launch_adj_box_fix(const AdjDev &a, const AdjInject &inj, hipStream_t st) {
    const long n = adj_box_cells(a);
    if (n <= 0) return inj.on ? hipErrorInvalidValue : hipSuccess;
    // one thread beyond the cells for an injection outside them
    hipLaunchKernelGGL(k_adj_box_fix, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, st, a, inj);
    return hipGetLastError();
}

hipError_t launch_adj_box_corr(const AdjDev &a, float *corr, hipStream_t st) {
    const long n = adj_box_cells(a);
    if (n <= 0 || !corr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_adj_box_corr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, corr);
    return hipGetLastError();
}

hipError_t launch_adj_box_add(const AdjDev &a, const AdjInject &inj, const float *corr, hipStream_t st) {
    const long n = adj_box_cells(a);
    if (n <= 0 || !corr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_adj_box_add, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, st, a, inj, corr);
    return hipGetLastError();
}

hipError_t launch_adj_disp(const AdjDispDev &a, hipStream_t st) {
    const long nb = (long)a.d.n[0] * a.d.n[1] * a.d.n[2];
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_disp, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_adj_inject(float *const E[3], long p0, float ax, float ay, float az, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_inject, dim3(1), dim3(64), 0, st, E[0], E[1], E[2], p0, ax, ay, az);
    return hipGetLastError();
}

hipError_t launch_probe(const float *const E[3], long p0, float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, st, E[0], E[1], E[2], p0, out);
    return hipGetLastError();
}

hipError_t launch_periodic_h(const StepDev &s, hipStream_t st) {
    const Geom &g = s.g;
    const dim3 wg((g.NZ + 63) / 64, g.NX + 1 > g.NY ? g.NX + 1 : g.NY, 2), block(64);
    hipLaunchKernelGGL(k_wrap_e, wg, block, 0, st, const_cast<float *>(s.Ein[0]), const_cast<float *>(s.Ein[1]),
                       const_cast<float *>(s.Ein[2]), g);
    const dim3 grid((g.NZ + 63) / 64, g.NY, g.NX);
    hipLaunchKernelGGL(k_per_h, grid, block, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_periodic_e(const StepDev &s, hipStream_t st) {
    const Geom &g = s.g;
    const dim3 wg((g.NZ + 63) / 64, g.NX + 1 > g.NY ? g.NX + 1 : g.NY, 2), block(64);
    hipLaunchKernelGGL(k_wrap_h, wg, block, 0, st, const_cast<float *>(s.Hin[0]), const_cast<float *>(s.Hin[1]),
                       const_cast<float *>(s.Hin[2]), g);
    const dim3 grid((g.NZ + 63) / 64, g.NY, g.NX);
    hipLaunchKernelGGL(k_per_e, grid, block, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_periodic_tfsf(const StepDev &s, hipStream_t st) {
    const dim3 grid((s.g.NY + 63) / 64, s.g.NX), block(64);
    hipLaunchKernelGGL(k_per_tfsf, grid, block, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_split_h(const StepDev &s, hipStream_t st) {
    dim3 grid((s.g.NZ + 63) / 64, s.g.NY, s.g.NX), block(64);
    hipLaunchKernelGGL(k_split_h, grid, block, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_split_e(const StepDev &s, hipStream_t st) {
    dim3 grid((s.g.NZ + 63) / 64, s.g.NY, s.g.NX), block(64);
    hipLaunchKernelGGL(k_split_e, grid, block, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_tfsf_split(const StepDev &s, hipStream_t st) {
    const TfsfDev &t = s.tf;
    const int nk = t.t1[2] - t.t0[2] + 2, nj = t.t1[1] - t.t0[1] + 2, ni = t.t1[0] - t.t0[0] + 2;
    dim3 grid((nk + 63) / 64, nj, ni), block(64);
    hipLaunchKernelGGL(k_tfsf_split, grid, block, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_tfsf_fix(const StepDev &s, hipStream_t st) {
    const TfsfDev &t = s.tf;
    int span = 0;
    for (int a = 0; a < 3; a++) span = max(span, t.t1[a] - t.t0[a] + 3);
    dim3 grid((span + 63) / 64, span, 12), block(64);   // the body assumes 64-thread blocks
    hipLaunchKernelGGL(k_tfsf_fix, grid, block, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_post(const StepDev &s, const DispDev *d, bool sample, hipStream_t st) {
    static_assert(sizeof(StepDev) + sizeof(DispDev) + 32 <= 4096, "kernel arguments");
    const TfsfDev &t = s.tf;
    int span = 0;
    for (int a = 0; a < 3; a++) span = max(span, t.t1[a] - t.t0[a] + 3);
    const int nbx = (span + 63) / 64;
    const long nb = d ? (long)d->n[0] * d->n[1] * d->n[2] : 0;
    const long nd = (nb + 255) / 256, nt = ((long)nbx * span * 12 + 3) / 4, ns = sample ? (s.fd.mE + 255) / 256 : 0;
    if (nd + nt + ns > 0x7fffffff) return hipErrorInvalidValue;
    const DispDev none{};
    hipLaunchKernelGGL(k_post, dim3((unsigned)(nd + nt + ns)), dim3(256), 0, st, s, d ? *d : none, (int)nd, nbx, span,
                       (int)nt, (int)ns);
    return hipGetLastError();
}

hipError_t launch_line(const LineDev &l, hipStream_t st) {
    hipLaunchKernelGGL(k_line, dim3(1), dim3(1024), 0, st, l);
    return hipGetLastError();
}

hipError_t launch_dispersive(const DispDev &d, hipStream_t st) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dispersive, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_hydro(const DispDev &d, hipStream_t st) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    if (nb <= 0) return hipSuccess;
    const dim3 grid((unsigned)((nb + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_hydro_v_ce0, grid, block, 0, st, d);
    for (int c = 1; c < 3; c++) hipLaunchKernelGGL(k_hydro_ce, grid, block, 0, st, d, c);
    return hipGetLastError();
}

hipError_t launch_nonlocal_gd(const DispDev &d, hipStream_t st) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    if (nb <= 0 || !d.gd[0]) return hipSuccess;
    hipLaunchKernelGGL(k_nonlocal_gd, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_disp_sync_state(const DispDev &d, int from_field, hipStream_t st) {
    const long nb = (long)d.n[0] * d.n[1] * d.n[2];
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_disp_sync, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, d, from_field);
    return hipGetLastError();
}

hipError_t launch_dft_sample_split(const StepDev &s, hipStream_t st) {
    const long n = s.fd.mE;
    hipLaunchKernelGGL(k_dft_sample_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s);
    return hipGetLastError();
}

hipError_t launch_twiddle(double *W, const int *t_slot_host, int K, int nvalid, int F, const double *freq,
                          double dt, hipStream_t st) {
    if (K > 256) return hipErrorInvalidValue;
    SlotTimes ts;
    for (int i = 0; i < 256; i++) ts.t[i] = i < K ? t_slot_host[i] : 0;
    const int n = 2 * K * 2 * F;
    hipLaunchKernelGGL(k_twiddle, dim3((n + 255) / 256), dim3(256), 0, st, W, ts, K, nvalid, F, freq, dt);
    return hipGetLastError();
}

hipError_t launch_dft_contract(const float *X, const double *W, double *acc, long mtot, long mE_pad, int K,
                               int F, hipStream_t st) {
    dim3 grid((unsigned)((mtot + 255) / 256), (2 * F + kRC - 1) / kRC), block(256);
    hipLaunchKernelGGL(k_dft_contract, grid, block, 0, st, X, W, acc, mtot, mE_pad, K, F);
    return hipGetLastError();
}

hipError_t launch_twiddle_frag(void *Wf, const int *t_slot_host, int K, int nvalid, int F, const double *freq,
                               double dt, int parts, hipStream_t st) {
    if (K > 256 || (parts != 2 && parts != 3)) return hipErrorInvalidValue;
    SlotTimes ts;
    for (int i = 0; i < 256; i++) ts.t[i] = i < K ? t_slot_host[i] : 0;
    const int nks = (K + 15) / 16, nmt = (2 * F + 31) / 32;
    const int n = 2 * nks * nmt * 64;
    if (parts == 2)
        hipLaunchKernelGGL(k_twiddle_frag<2>, dim3((n + 255) / 256), dim3(256), 0, st, (bf16x8 *)Wf, ts, nks, nmt, nvalid,
                           F, freq, dt);
    else
        hipLaunchKernelGGL(k_twiddle_frag<3>, dim3((n + 255) / 256), dim3(256), 0, st, (bf16x8 *)Wf, ts, nks, nmt, nvalid,
                           F, freq, dt);
    return hipGetLastError();
}

hipError_t launch_dft_mfma(const float *X, const void *Wf, double *acc, long mtot, long mE_pad, int K, int F, int parts,
                           hipStream_t st) {
    if (parts != 2 && parts != 3) return hipErrorInvalidValue;
    const int nks = (K + 15) / 16, nmt = (2 * F + 31) / 32;
    dim3 grid((unsigned)((mtot + 127) / 128), (nmt + 3) / 4), block(256);
    if (parts == 2)
        hipLaunchKernelGGL(k_dft_mfma<2>, grid, block, 0, st, X, (const bf16x8 *)Wf, acc, mtot, mE_pad, nks, nmt, 2 * F);
    else
        hipLaunchKernelGGL(k_dft_mfma<3>, grid, block, 0, st, X, (const bf16x8 *)Wf, acc, mtot, mE_pad, nks, nmt, 2 * F);
    return hipGetLastError();
}

hipError_t launch_flux_reduce(const double *acc, const FaceDev &fd, long mtot, long mE_pad, int F, double dA,
                              double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_flux, dim3(12, F), dim3(256), 0, st, acc, fd, mtot, mE_pad, F, dA, out);
    return hipGetLastError();
}

hipError_t launch_pack_field(const float *f, const Geom &g, float *dense, int to_dense, hipStream_t st) {
    dim3 grid((g.NZ + 63) / 64, g.NY, g.NX), block(64);
    hipLaunchKernelGGL(k_pack, grid, block, 0, st, f, g, dense, to_dense);
    return hipGetLastError();
}
