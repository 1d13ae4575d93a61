// Internal structures shared by the host side (api.cpp, setup.cpp) and the
// HIP kernels (kernels.hip).  Not part of the public ABI.
#ifndef FDTD_INTERNAL_H_
#define FDTD_INTERNAL_H_

#include <stdint.h>
#include <hip/hip_runtime.h>

// ---------------------------------------------------------------------------
// Field layout in HBM.  Every 3D field is FP32, z fastest, with one ghost
// layer on each side of every axis so stencil reads at -1 and N never need a
// bounds test.  The stored pointer is pre-offset to element (0,0,0):
//     f[i*sxy + j*sz + k],  i in [-1, NX], j in [-1, NY], k in [-1, NZ].
// sz is NZ+2 rounded up to 32 floats (128 B rows); ghosts are zero forever.
// ---------------------------------------------------------------------------
struct Geom {
    int NX, NY, NZ;
    int sz;        // row pitch (floats)
    long sxy;      // plane pitch (floats) = sz * (NY + 2)
};

__host__ __device__ inline long gidx(const Geom &g, int i, int j, int k) {
    return (long)i * g.sxy + (long)j * g.sz + k;
}

// CPML tables per axis, FP32 copies of the FP64 host profiles.
//   ik_e[N], ik_h[N]   : 1/(kappa*d) at E and H positions (ked/khd, setup.cu:518-529)
//   be[2*(Nc+1)], ce[2*(Nc+1)]: E recursion, near slab at [0, Nc], far at Nc+1+k2
//   bh[2*Nc],     ch[2*Nc]    : H recursion, near at [0, Nc), far at Nc+k2
//   ce/ch already divided by the cell size.
struct CpmlDev {
    const float *ik_e, *ik_h;
    const float *be, *ce, *bh, *ch;
    int Nc;     // slab thickness
    int F;      // first cell of the far slab (cpml_F_* in the reference)
};

// psi storage.  Slab arrays hold the near slab then the far slab along the
// slab axis (H: 2*Nc layers, E: 2*(Nc+1) layers), z fastest.
//   z-slabs : [NX][NY][layers]
//   y-slabs : [NX][layers][NZ]
//   x-slabs : [layers][NY][NZ]
// Order of the six arrays per field type (component, slab axis):
//   H: 0 hx_z, 1 hx_y, 2 hy_z, 3 hy_x, 4 hz_x, 5 hz_y
//   E: 0 ex_z, 1 ex_y, 2 ey_z, 3 ey_x, 4 ez_y, 5 ez_x
struct PsiSet {
    float *p[6];
};

// TFSF plane-wave source constants (CORRECT_X/Y/Z, main.cu:369-758).
struct TfsfDev {
    int t0[3], t1[3];        // face planes Ntfsf and N - Ntfsf per axis
    float s[3];              // propagation direction (sin th cos ph, sin th sin ph, cos th)
    float ef[3], hf[3];      // polarisation factors of E_inc and H_inc components
    float cH, cE;            // Che, Ceh
    float inv_d[3];          // 1/dx, 1/dy, 1/dz
    float dscale;            // dx / d_1D
    int m0;
    const double *e_inc;     // e_inc^n
    const double *h_inc;     // h_inc^{n+1/2}
};

// 12 DFT monitor faces; columns of the sample matrix (see dft_* in kernels.hip).
struct FaceDev {
    int plane[12];           // index along the normal axis
    int lo_a[12], lo_b[12];  // face range start along outer/inner in-plane axes
    int na[12], nb[12];      // face extents
    long off[12];            // first column of component 0 of this face (E half)
    long mE;                 // columns in the E half (H half starts at mE_pad)
    long mE_pad;
    unsigned *status;        // set to 1 when a sampled value is NaN or Inf (host-visible word)
};

// Uniform update coefficients (setup.cu:2015-2062 with uniform media).
struct CoefDev {
    float Chh, Che, Cee, Ceh;
};

// Arguments of the fused and split field kernels.
struct StepDev {
    Geom g;
    CoefDev c;
    CpmlDev cp[3];
    TfsfDev tf;
    int tfsf_on;
    const float *Ein[3], *Hin[3];   // E^n, H^{n-1/2}
    float *Eout[3], *Hout[3];       // E^{n+1}, H^{n+1/2}
    PsiSet psiH_in, psiH_out, psiE;
    // DFT sampling
    FaceDev fd;
    float *sample_row;              // row of the sample ring buffer, or null (no sample this step)
};

// Dispersive (ADE) region: a dense box [lo, lo+n) of the grid.
struct DispDev {
    Geom g;
    int lo[3], n[3];
    const uint8_t *id[3];           // per E component: 1 = dispersive, 0 = not
    const float *rho;               // design density per box cell, or null (rho = 1 everywhere)
    // density blend (eps_inf(rho) = 1 + rho (eps_inf - 1), pole drives x rho), all divided by eps0:
    // C_E/eps0 = 1 + rho u1, C_E1/eps0 = 1 + rho u2, C_E2/eps0 = rho u3
    float u1, u2, u3, A0;           // A0 = dt / (z0 eps0)
    float G0[3], K2[3];             // k1 + k2 - 1 and k2 per pole
    float *traj[3];                 // if set: E^{n+1} of every box cell is written here (adjoint recompute)
    float *traj_prev[2][3];         // if set (first recorded step): E^{n-1}, E^n written here
    // state per component, box-dense SoA: 0 E^n, 1 E^{n-1}, then per pole q
    // (0 Drude, 1-2 CP) P~^n at 2+2q and J~^n = P~^n - P~^{n-1} at 3+2q, P~ = P/eps0
    float *st[3][8];
    const float *H[3];              // H^{n+1/2}
    float *E[3];                    // output field (E^{n+1} written at dispersive cells)
    const float *ik_e[3];
    // E^{n+1} = a0*curl + a1*E^n - a2*E^{n-1} - sum_q (e0[q] P^n + e1[q] J^n)
    // J^{n+1} = g1[q] J^n + g0[q] P^n + p3[q] E^{n+1} + p4[q] E^n + p5[q] E^{n-1};  P^{n+1} = P^n + J^{n+1}
    float a0, a1, a2;
    float e0[3], e1[3], g0[3], g1[3], p3[3], p4[3], p5[3];
    // hydrodynamic model (pole 0 becomes the electron velocity v, st[2] = v^n,
    // st[3] = v^n - v^{n-1}; charge stored as n~ = ND / N_EQ); its kernels
    // compute in FP64 with the FP64 coefficients below (ha*, he*, hg*, hp*:
    // those of the E update and CP poles above)
    int hydro;
    float *vnew[3];                 // v^{n+1} from the momentum update
    float *nd[3], *nd_prev[3];      // n~^{n-1} (updated to n~^{n+1}), n~^n
    double h_d1m1, h_d3, h_d4;      // v^{n+1} = v + (d1-1) J + d3 E^n + d4 E^{n-1} + h_nl (gd + gd2) / cbrt(1 + n~)
    double h_nl, h_inv_d, h_cont;   // h_nl = beta^2/dd, 1/dx, 2 dt/dx
    double h_cur;                   // dt e0 N_EQ / C_E: current term of the E update
    double h_inv_neq;               // 1 / N_EQ
    double ha0, ha1, ha2, he0[3], he1[3], hg0[3], hg1[3], hp3[3], hp4[3], hp5[3];
    // NONLOCAL == 1: gd[c] = grad(div P~_d^n) at each dispersive cell of component c
    // (k_nonlocal_gd, before the box update); E^{n+1} -= nl_e gd, J~_d^{n+1} += nl_p gd
    float *gd[3];
    float nl_e, nl_p, nl_i2;        // d_NL eps0 / C_E, d_NL, 1/dx^2
};

// Adjoint step arguments.  The transposed step has the stencil shape of the
// forward step: the adjoint of H^{n+1/2} gathers forward differences of the
// E-side weights V, the adjoint of E^n gathers backward differences of the
// H-side weights W.
struct AdjDev {
    Geom g;
    CoefDev c;
    CpmlDev cp[3];
    const float *Ein[3], *Hin[3];   // adjoint of E^{n+1}, H^{n+1/2}
    float *Eout[3], *Hout[3];       // adjoint of E^n, H^{n-1/2}
    PsiSet psiE_in, psiE_out, psiH_in, psiH_out;
    // dispersive box: curl weight Q and the E^n adjoint base from k_adj_disp
    int blo[3], bn[3];
    const uint8_t *bid[3];
    const float *Qb[3], *Eb[3];
    // per box cell, Qb - Ceh Ehat and Eb - Cee Ehat (zero at cells that are not dispersive)
    const float *dq[3], *db[3];
};

// The objective's injection dJ/dE at one cell, added to the E adjoint by
// k_adj_inject, or folded into the dispersive-cell correction of the step
// before (after its own update of that cell).
struct AdjInject {
    long p0;                        // cell (field layout)
    float a[3];
    int on;
    int inside;                     // p0 lies in the cells of the correction launch
};

// Adjoint of the dispersive update on the box (run before k_adj_step).
struct AdjDispDev {
    DispDev d;                      // forward coefficients, geometry, rho
    const float *Ehat[3];           // adjoint of E^{n+1} (field layout)
    float *a[3][7];                 // adjoint state: 0 E^{n-1}, 1+2q P_q, 2+2q J_q
    float *Qb[3], *Eb[3];
    float *dq[3], *db[3];           // Qb - Ceh Ehat, Eb - Cee Ehat (the correction after the row kernel), or null
    float Ceh, Cee;
    const float *tm1[3], *t0[3], *tp1[3];   // E^{n-1}, E^n, E^{n+1} on the box (gradient)
    double *grad;                   // dJ/drho per box cell, or null
};

// 1D incident line (main.cu:2280-2337, 2831-2859), FP64.
struct LineDev {
    const double *e_old;
    double *e_new;
    double *h;
    double *mur;                    // e1, e2
    int L, src;
    double ch, ce, mur_coef;
    double amp, delay, width, w0dt; // pulse: amp*exp(-((t-delay)/width)^2/2)*sin(w0dt*t)
    int static_field;
    int t;
    // divergence probe (main.cu:2728-2731): ex^n at one cell; status set to 1
    // when it is NaN or Inf (either may be null)
    const float *probe;
    unsigned *status;
};

// ---------------------------------------------------------------------------
// Launch wrappers (kernels.hip).  All return hipError_t of the launch.
// ---------------------------------------------------------------------------
// Sub-boxes of owned cells for the general fused kernel: x in [ia, ib), y in
// [ja, jb), all z; blocks of piece q are [start[q], start[q+1]).
struct Pieces {
    int n;
    int ia[4], ib[4], ja[4], jb[4];
    int njt[4];
    int start[5];
};

// Arguments of the row-blocked fused step (k_step_rows).  Field pointers are
// allocation bases (element (-1,-1,-1)); aux is the table image (CPML tables,
// interior 1/(kappa d), psi pointers; layout in kernels.hip, built by the
// host).  The kernel runs the cells x in [box[q][0], box[q][1]) by y in
// [box[q][2], box[q][3]), all z, of npiece boxes; an axis taken as clean
// (x in variants 0 and 2, y in 0 and 1) must carry no CPML there, have
// 1/(kappa d) = ik0[e][axis] (e: 0 E, 1 H), pass every domain test, and
// have an H-clean plane or row just below the box.  nseg and start are
// filled in by the launcher.
struct RowsArgs {
    const char *Ein[3], *Hin[3];
    char *Eout[3], *Hout[3];
    const unsigned *aux;
    float Chh, Che, Cee, Ceh;
    float ik0[2][2];
    int NX, NY, NZ, sz;
    unsigned sxy4;                      // plane pitch in bytes
    int Nc[3], F[3];
    int pp;                             // psi_H input set
    int pe;                             // adjoint variant: psi_E input set
    int ichunk;
    int xcd;                            // 1: XCD-contiguous block order (j-neighbour tiles share an L2)
    int npiece;
    int box[4][4];
    int nseg;
    int start[5];
    int tiles;                          // row tiles per block: 1, or 2-3 for the clean variant when tiles * W <= 16
};
// Words of the table image and its layout (see kernels.hip).
constexpr int kRowsAuxWords = 1548 + 48;
constexpr int kRowsAuxAxis = 512, kRowsAuxMaxEdge = 64, kRowsAuxMaxLayers = 64;
// Waves per block of the row-blocked kernel for a given NZ (one z segment when NZ < 64*16 - 2).
int rows_waves(int NZ);
// R: own rows per thread (2 or 3); TFSF is left to launch_tfsf_fix.
// variant 0: clean boxes (no x or y CPML); 1: x general (x CPML and x
// boundaries per plane), y clean; 2: y general (per row), x clean; 3: the
// transposed step of the adjoint on clean boxes (coefficients passed as
// Chh = 1, Che = -Ceh, Cee, Ceh = -Che; the table image holds both adjoint
// psi_E sets; dispersive cells are corrected afterwards by launch_adj_box_fix);
// 4, 5: the transposed step on x slabs (as 1) and y slabs (as 2), two rows,
// one row tile, no dispersive cell in the boxes; 6: the transposed step on
// the corner columns (x and y general), two rows, one row tile.
hipError_t launch_step_rows(RowsArgs ra, int R, int variant, hipStream_t st);

// box[q] = {ia, ib, ja, jb}; one to four disjoint boxes.
hipError_t launch_fused_step(const StepDev &s, int tj, int ichunk, const int (*box)[4], int nbox, hipStream_t st);
hipError_t launch_split_h(const StepDev &s, hipStream_t st);
// Periodic x, y: E ghost wrap + H update (with the near z plane of the TFSF
// when s.tfsf_on), H ghost wrap + E update, and the plane corrections alone
// for the split path.  In place: Ein == Eout, Hin == Hout.
hipError_t launch_periodic_h(const StepDev &s, hipStream_t st);
hipError_t launch_periodic_e(const StepDev &s, hipStream_t st);
hipError_t launch_periodic_tfsf(const StepDev &s, hipStream_t st);
hipError_t launch_split_e(const StepDev &s, hipStream_t st);
hipError_t launch_tfsf_split(const StepDev &s, hipStream_t st);
// TFSF terms added after a fused step run with tfsf_on = 0 (H faces, E faces, curl of dH).
hipError_t launch_tfsf_fix(const StepDev &s, hipStream_t st);
hipError_t launch_line(const LineDev &l, hipStream_t st);
hipError_t launch_dispersive(const DispDev &d, hipStream_t st);
// launch_tfsf_fix, launch_dispersive (Drude + CP, when d is given) and, with
// sample, the DFT sample of s.Eout / s.Hout into s.sample_row, in one launch;
// only for cell sets that are pairwise disjoint.
hipError_t launch_post(const StepDev &s, const DispDev *d, bool sample, hipStream_t st);
hipError_t launch_dft_sample_split(const StepDev &s, hipStream_t st);
hipError_t launch_twiddle(double *W, const int *t_slot_host, int K, int nvalid, int F, const double *freq,
                          double dt, hipStream_t st);
hipError_t launch_dft_contract(const float *X, const double *W, double *acc, long mtot, long mE_pad, int K,
                               int F, hipStream_t st);
// Split-BF16 MFMA path in `parts` (2 or 3) BF16 parts: twiddle fragments
// (2 * ceil(K/16) * ceil(2F/32) * parts * 64 * 16 bytes) and contraction.
hipError_t launch_twiddle_frag(void *Wf, const int *t_slot_host, int K, int nvalid, int F, const double *freq,
                               double dt, int parts, hipStream_t st);
hipError_t launch_dft_mfma(const float *X, const void *Wf, double *acc, long mtot, long mE_pad, int K, int F, int parts,
                           hipStream_t st);
hipError_t launch_flux_reduce(const double *acc, const FaceDev &fd, long mtot, long mE_pad, int F, double dA,
                              double *out, hipStream_t st);
hipError_t launch_pack_field(const float *f, const Geom &g, float *dense, int to_dense, hipStream_t st);
hipError_t launch_disp_sync_state(const DispDev &d, int from_field, hipStream_t st);
// Hydrodynamic model: velocity update, then per component (x, y, z) continuity and E update.
hipError_t launch_hydro(const DispDev &d, hipStream_t st);
// NONLOCAL == 1: grad(div) of the Drude polarisation into d.gd (no-op when d.gd[0] is null).
hipError_t launch_nonlocal_gd(const DispDev &d, hipStream_t st);

// ---------------------------------------------------------------------------
// Batched launches: one launch runs the same kernel over many independent
// simulations.  Per-simulation arguments sit in device memory (one item per
// simulation, read with scalar loads); map[b] names the item of block b and
// the block's index within that simulation's cell set: item << 22 | local.
// ---------------------------------------------------------------------------
constexpr int kBatchLocalBits = 22;
struct FusedItem {
    StepDev s;
    Pieces pc;
    int ichunk, nkt;
};
struct SampleItem {
    Geom g;
    FaceDev fd;
    const float *E[3], *H[3];
    float *X;                       // sample ring [K][mtot]
    long mtot;
};
// dyn = {t, slot}: the step index of the first step of this launch sequence
// and the ring row of its first sample; k_batch_advance moves them on.
hipError_t launch_fused_batch(const FusedItem *items, const unsigned *map, int nblocks, int tj, hipStream_t st);
hipError_t launch_line_batch(const LineDev *items, int n, const int *dyn, int step_off, hipStream_t st);
hipError_t launch_disp_batch(const DispDev *items, const unsigned *map, int nblocks, hipStream_t st);
hipError_t launch_sample_batch(const SampleItem *items, const unsigned *map, int nblocks, const int *dyn, int sidx,
                               hipStream_t st);
hipError_t launch_batch_advance(int *dyn, int dt, int dslot, hipStream_t st);
// Row-blocked kernel over many simulations: items are RowsArgs prepared by
// rows_prepare (same W = rows_waves(NZ), R and variant 0-2 for all of them).
hipError_t launch_rows_batch(const RowsArgs *items, const unsigned *map, int nblocks, int W, int R, int variant,
                             hipStream_t st);
// Row tiles per block of launch_rows_batch for W waves per z column (items'
// tiles must be set to it before rows_prepare).
int rows_batch_tiles(int W);
// TFSF fix-up over many simulations (item.s as for launch_tfsf_fix).
hipError_t launch_tfsf_fix_batch(const FusedItem *items, const unsigned *map, int nblocks, hipStream_t st);
// Blocks of launch_tfsf_fix_batch for one simulation (four 64-lane rows each).
long tfsf_fix_blocks(const StepDev &s);
// Fills nseg and start of a RowsArgs for R and variant; returns its block count or -1.
long rows_prepare(RowsArgs &ra, int R, int variant);
// Blocks of one simulation in a batched launch of the general kernel (tile tj x ichunk over the whole grid,
// or over one to four x-y boxes); -1 when the arguments do not fit.
long fused_batch_blocks(const Geom &g, int tj, int ichunk, Pieces *pc, int *nkt);
long fused_pieces(const Geom &g, int tj, int ichunk, const int (*box)[4], int nbox, Pieces *pc, int *nkt);

// Adjoint: inject dJ/dE^{n+1}(r0), transposed dispersive update, fused transposed step.
hipError_t launch_adj_inject(float *const E[3], long p0, float ax, float ay, float az, hipStream_t st);
hipError_t launch_adj_disp(const AdjDispDev &a, hipStream_t st);
// box[q] = {ia, ib, ja, jb}: one to four disjoint x-y boxes of cells (all z).
hipError_t launch_adj_step(const AdjDev &a, int tj, int ichunk, const int (*box)[4], int nbox, hipStream_t st);
// Dispersive-cell correction after the adjoint row kernel (variant 3 of
// launch_step_rows), then the injection inj (if on) into its E output.
hipError_t launch_adj_box_fix(const AdjDev &a, const AdjInject &inj, hipStream_t st);
// The same in two launches: the amounts into corr (6 x adj_box_cells floats;
// reads the step's inputs and the transposed dispersive update only), then
// their addition to the outputs after the row kernel, with the injection.
long adj_box_cells(const AdjDev &a);
hipError_t launch_adj_box_corr(const AdjDev &a, float *corr, hipStream_t st);
hipError_t launch_adj_box_add(const AdjDev &a, const AdjInject &inj, const float *corr, hipStream_t st);
// E at one cell for every step (objective probe).
hipError_t launch_probe(const float *const E[3], long p0, float *out, hipStream_t st);

#endif
