/*
 * fdtd.h — C ABI of the MI355X-native 3D FDTD plasmonics solver.
 *
 * The reference (jbaxt089/FDTD-CUDA) has no library boundary: its time loop
 * in main.cu:2724-3175 launches 15 kernels with 30-70 raw arguments each and
 * keeps all state in ~700 process globals (global_var.h:32-815).  This header
 * replaces that loop body with an opaque handle that keeps the same call
 * structure:
 *
 *     setup                      -> fdtd_create            (SETUP_CONST, MATERIAL_MATRIX,
 *                                                          SETUP_Drude_CP, SETUP_CPML,
 *                                                          SETUP_TFSF, ALLOCATE_MEM)
 *     UPDATE_hx/hy/hz            -> fdtd_update_h          (main.cu:2735-2769)
 *     UPDATE_h_inc, CORRECT_X/Y/Z,
 *     UPDATE_e_inc, Mur, source  -> fdtd_tfsf              (main.cu:2791-2859)
 *     UpdateHydroP*, UPDATE_e*   -> fdtd_update_e          (main.cu:2869-2901)
 *     ScattAbs (t % 4 == 0), t++ -> fdtd_dft_accumulate    (main.cu:2920-2934)
 *     the four above, fused      -> fdtd_step
 *     CalculateAbsScatt          -> fdtd_flux              (Spectral.cu:411-1177)
 *
 * Conventions: every entry point returns FDTD_OK (0) or a negative code and
 * never exits the process.  Host buffers belong to the caller and are copied
 * in or out; the handle owns all device memory and one HIP stream.  Calls on
 * one handle must not overlap; different handles are independent.  Arrays
 * crossing the boundary use the reference's layout: index k + NZ*(j + NY*i)
 * with z fastest (main.cu:16-19), over the augmented grid NX*NY*NZ.
 * Device state is FP32; coefficients are computed on the host in FP64.
 */
#ifndef FDTD_H_
#define FDTD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDTD_OK               0
#define FDTD_ERR_INVALID     -1   /* bad argument or parameter */
#define FDTD_ERR_ALLOC       -2   /* device or host allocation failed */
#define FDTD_ERR_DEVICE      -3   /* HIP runtime error */
#define FDTD_ERR_NONFINITE   -4   /* NaN/Inf found in the fields (divergence) */
#define FDTD_ERR_UNSUPPORTED -5   /* feature not available for this setup */

/* Field component ids for fdtd_get_field / fdtd_set_field. */
enum { FDTD_EX = 0, FDTD_EY, FDTD_EZ, FDTD_HX, FDTD_HY, FDTD_HZ };

/*
 * Simulation parameters.  The first block mirrors the Input.txt tokens that
 * the hot path consumes (setup.cu:17-58); the second exposes values the
 * reference hard-codes (defaults equal the reference's hard-coded values).
 * fdtd_default_params fills every field.
 */
typedef struct fdtd_params {
    /* --- Input.txt mirror --- */
    int    Tend;               /* steps; the reference loop runs t = 0..Tend (main.cu:2724)    */
    double dx, dy, dz;         /* cell size in metres; the solver requires dx == dy == dz      */
    int    NCELLX, NCELLY, NCELLZ;   /* interior cells, before CPML and +1 (setup.cu:377-393)  */
    int    NcpmlX, NcpmlY, NcpmlZ;   /* CPML thickness in cells                                */
    int    nano_sphere;        /* 0: none, 1: staircased sphere at the grid centre (setup.cu:219) */
    double nano_sphere_radius; /* metres                                                       */
    int    material;           /* 2: silver, 3: gold Drude + 2 CP (setup.cu:1164-1211), 7: Lorentz-Drude
                                  (two Lorentz poles + Drude, setup.cu:1256-1275, 1384-1402)  */
    int    first_medium;       /* background material id; must be <= 1                        */
    int    NUM_freq;           /* DFT frequencies                                              */
    int    WL_or_freq;         /* 1: linear in wavelength, 0: linear in frequency (main.cu:2448) */
    double BandWidth;          /* pulse bandwidth (Hz) (tfsf2.cu:37)                           */
    int    NONLOCAL;           /* 1: grad-div term on the Drude pole with velocity hydro_beta
                                  (updatefields.cu:876-891, 1424-1604), 3: hydrodynamic nonlocal
                                  model (setup.cu:995)                                         */
    double MAX_AMP;            /* source amplitude (main.cu:2848)                              */
    double TimeFactor;         /* dt = 0.5*dy/c0/TimeFactor (setup.cu:359)                     */
    double lam_0;              /* pulse centre wavelength (m); f_0 = c0/lam_0 (setup.cu:62)    */
    int    StaticField;        /* 1: ramped static source, no DFT (main.cu:2852-2857)          */
    /* --- values the reference hard-codes --- */
    int    tfsf_inset;         /* TFSF box inset inside the CPML, 10 (setup.cu:428)            */
    int    abs_offset;         /* absorption box inside the TFSF box, 5 (setup.cu:435)         */
    int    sca_offset;         /* scattering box outside the TFSF box, 5 (setup.cu:451)        */
    double inc_theta, inc_phi, polar_psi;  /* incidence angles in degrees, 0 (tfsf2.cu:11-13)  */
    int    cpml_exp;           /* polynomial grading order, 3 (setup.cu:481)                   */
    double kappa_max;          /* 5 (setup.cu:484)                                              */
    double alpha_max;          /* 0.05 (setup.cu:496)                                           */
    double sigma_factor;       /* sigma_max = sigma_factor*(m+1)/(Z0*d), 0.8 (setup.cu:491)     */
    double lam_min, lam_max;   /* DFT wavelength range, 200 nm .. 700 nm (main.cu:2446-2447)    */
    int    dft_stride;         /* sample every dft_stride steps, 4 (main.cu:2920)               */
    int    dft_block;          /* K: samples buffered per MFMA contraction                      */
    int    source_offset;      /* source cell m0 - source_offset, 50 (main.cu:2851)             */
    double hydro_beta;         /* nonlocal velocity beta (m/s) of NONLOCAL 1 and 3,
                                  1.39e6*sqrt(3/4) (setup.cu:980-983)                          */
    double xs_area;            /* cross-section normaliser (m^2); 0 -> pi*r^2 (Spectral.cu:1166) */
    int    dft_exact;          /* DFT contraction: 0 split-BF16 MFMA, hi + lo parts (default), 1 FP64,
                                  2 split-BF16 MFMA in three parts (six products, FP32-level error) */
    int    periodic_xy;        /* 1: x and y periodic (Periodic_XY, main.cu:781-820 and the periodic
                                  branches of every update): NcpmlX = NcpmlY = 0, grid NCELLX by
                                  NCELLY with no +1 (setup.cu:401-405), normal incidence only, the
                                  plane wave enters through the near z plane of the TFSF box alone
                                  (main.cu:698), spectra from the reflection and transmission planes
                                  (fdtd_reflectance)                                              */
} fdtd_params;

/* Derived grid facts (read-only). */
typedef struct fdtd_grid_info {
    int    NX, NY, NZ;         /* augmented grid, interior + 2*Ncpml + 1                       */
    int    Ncpml[3];
    double dt;
    int    tfsf[3];            /* Ntfsf{X,Y,Z}: TFSF box spans [tfsf, N - tfsf]               */
    int    abs_start[3], abs_end[3];   /* absorption monitor box [start, end)                  */
    int    sca_start[3], sca_end[3];   /* scattering monitor box [start, end)                  */
    int    inc_length, m0;     /* 1D incident line (setup.cu:350, tfsf2.cu:29)                */
    double width, delay, f_0;  /* pulse parameters (tfsf2.cu:37-40)                           */
    int    num_freq;
    int    t;                  /* index of the next step                                      */
    int    hydro;              /* 1 when the hydrodynamic model is on                          */
    int    n_dispersive[3];    /* E-component cells on the dispersive (ADE) branch             */
    int    box_lo[3], box_hi[3];       /* bounding box of dispersive cells, [lo, hi)          */
    int    sample_count;       /* DFT samples taken so far                                     */
    int    fused_path;         /* fdtd_step's field update: 1 row-blocked kernel + TFSF fix-up,
                                  0 general kernel with TFSF inside (no reference counterpart) */
    int    clean[4];           /* x [clean0, clean1) by y [clean2, clean3), all z: cells run by the
                                  CPML-free variant of the row-blocked kernel; -1 when none       */
    int    schedule;           /* fdtd_step on two streams: bit 0 the TFSF fix-up, bit 1 the
                                  dispersive box, bit 2 the DFT sample follow the clean box without
                                  waiting for the side pieces; bit 3 the next step's y slabs and
                                  corner columns start at the join, beside the post-update launch
                                  (no reference counterpart)                                      */
    int    tile[7];            /* tiles fdtd_step uses (no reference counterpart): row-blocked kernel
                                  rows per thread, clean-box x chunk, side-slab x chunk, row tiles
                                  per block on the clean box and on the side slabs, tile rows of the
                                  corner columns, streams of the side pieces beside the clean box
                                  (0: none, 1: one, 3: x slabs, y slabs, corners on their own)      */
} fdtd_grid_info;

typedef struct fdtd_sim fdtd_sim;

/* Fills *p with the reference's defaults (a 64^3 gold-sphere run). */
void fdtd_default_params(fdtd_params *p);

/* Setup: SETUP_CONST (setup.cu:347), MATERIAL_MATRIX (setup.cu:135), the
 * frequency list (main.cu:2446-2463), SETUP_Drude_CP (setup.cu:979),
 * ALLOCATE_MEM (setup.cu:2396), SETUP_TFSF (tfsf2.cu:6), the update
 * coefficients (setup.cu:2015-2062), SETUP_CPML (setup.cu:513) and the source
 * spectrum (main.cu:2629-2644).  Fields start at zero. */
int  fdtd_create(const fdtd_params *p, fdtd_sim **out);
int  fdtd_destroy(fdtd_sim *sim);

/* Message for the last failure on this handle (or of the last failed
 * fdtd_create when sim is NULL).  Never NULL. */
const char *fdtd_last_error(const fdtd_sim *sim);

int  fdtd_get_info(const fdtd_sim *sim, fdtd_grid_info *info);

/* Replaces the per-component material masks mat_matrix{X,Y,Z}
 * (setup.cu:135-342).  Each array is NX*NY*NZ ids; an E component takes the
 * dispersive branch where 1 < id < 6 (main.cu:1175) or id == 7 (the
 * Lorentz-Drude material, which the reference's test leaves out).  Dispersive cells must
 * lie strictly inside the absorption box. */
int  fdtd_set_materials(fdtd_sim *sim, const uint8_t *id_x, const uint8_t *id_y,
                        const uint8_t *id_z);
/* Paints a staircased sphere of the configured material into the masks with
 * the reference's rule (setup.cu:219-248): an E component at Yee position x
 * is inside when |x - centre| < (int)(radius/dx), centre in cells.  Used for
 * the dimer and other multi-particle geometries; same validity rules as
 * fdtd_set_materials. */
int  fdtd_add_sphere(fdtd_sim *sim, const double centre[3], double radius);
/* Nanorod (no reference counterpart; parameter-sweep workloads): a cylinder
 * with hemispherical caps, total length `length` and radius `radius`
 * (metres), along axis 0..2 (x..z), centre in cells.  An E component at Yee
 * position x is inside when its distance to the axis segment between the cap
 * centres is < radius/dx; same validity rules as fdtd_set_materials. */
int  fdtd_add_rod(fdtd_sim *sim, const double centre[3], int axis, double length, double radius);
/* Design region (no reference counterpart; inverse-design workloads): the block
 * [lo, lo+n) of cells becomes dispersive for all three E components with the
 * ADE coefficients blended by the density rho in [0, 1] (i-major, k fastest):
 * eps_inf(rho) = 1 + rho (eps_inf - 1) and every pole drive scaled by rho,
 * so rho = 0 is vacuum and rho = 1 the configured metal.  Rebuilding the
 * dispersive box frees the device buffers sized by the previous one (box
 * state and the adjoint's), so a design loop on one handle keeps its device
 * memory flat. */
int  fdtd_set_design(fdtd_sim *sim, const int lo[3], const int n[3], const float *rho);
/* Discrete adjoint (no reference counterpart).  From the current state (t = 0)
 * runs nsteps forward with a full-state checkpoint every ckpt steps, then the
 * transposed step in reverse time, recomputing each segment from its
 * checkpoint (the last segment keeps its box trajectory from the forward pass
 * and is not recomputed).  ckpt <= 0 keeps E on the dispersive box for every step instead
 * (3 floats per box cell per step of device memory) and needs no recompute;
 * both give the same bits.  Objective J = sum_f sum_c |sum_t E_c^{t+1}(r0) e^{i 2 pi f_j (t+1) dt}|^2
 * over the nf frequencies freq (Hz) at cell r0.  Writes J and dJ/drho for the
 * design block (same order as fdtd_set_design).  The field state afterwards is
 * the forward state after nsteps. */
int  fdtd_adjoint(fdtd_sim *sim, int nsteps, int ckpt, const int r0[3], int nf, const double *freq,
                  double *J, double *grad);
/* Copies a mask out (comp 0: mat_matrix, 1..3: X, Y, Z). */
int  fdtd_get_materials(const fdtd_sim *sim, int comp, uint8_t *out);

/* Split step, same order as the reference loop body.  fdtd_dft_accumulate
 * ends the step: it samples the DFT monitors when t % dft_stride == 0 and
 * advances t. */
int  fdtd_update_h(fdtd_sim *sim);
int  fdtd_tfsf(fdtd_sim *sim);
int  fdtd_update_e(fdtd_sim *sim);
int  fdtd_dft_accumulate(fdtd_sim *sim);

/* Fused fast path: nsteps full steps (H + TFSF + E + DFT). */
int  fdtd_step(fdtd_sim *sim, int nsteps);

/* Point monitor (no reference counterpart; the gap field of dimer
 * workloads): records E at cell r after every following step (fdtd_step or
 * the split path), up to `capacity` steps.  fdtd_point_spectrum returns, per
 * frequency (Hz), |E(f)|^2 = sum_c |sum_t E_c^{t+1}(r) e^{i 2 pi f (t+1) dt}|^2
 * in FP64 (the DFT convention of the monitors), and when amp is not NULL the
 * three complex amplitudes (re, im pairs, nf*3 of them). */
int  fdtd_set_point_monitor(fdtd_sim *sim, const int r[3], int capacity);
int  fdtd_point_spectrum(fdtd_sim *sim, int nf, const double *freq, double *e2, double *amp);

/* Blocks until queued work finished; reports device errors, and
 * FDTD_ERR_NONFINITE once a NaN or Inf has been seen.  The check follows the
 * reference's per-step NaN probe (ex at NX/2, NY/2, NZ/2 - NZ/4 of the
 * augmented grid, main.cu:2728-2731), read by the 1D-line kernel every step,
 * and every value the DFT monitors sample; the flag is a host-visible word,
 * so fdtd_step and fdtd_batch_run also return FDTD_ERR_NONFINITE (without
 * stepping) once it is set, and fdtd_flux, fdtd_get_dft and
 * fdtd_point_spectrum report it. */
int  fdtd_sync(fdtd_sim *sim);

/* Field copies, NX*NY*NZ floats in the reference layout. */
int  fdtd_get_field(fdtd_sim *sim, int comp, float *out);
/* One plane of a field component, copied straight from the device (no
 * reference counterpart; SNAPSHOT_2D, snapshot.cu:25-60, reads planes of a
 * full host copy): axis 0 gives the plane i = index as out[j*NZ + k], axis 1
 * the plane j = index as out[i*NZ + k], axis 2 the plane k = index as
 * out[i*NY + j] (augmented grid).  comp 0-5 as fdtd_get_field; comp 6 (with
 * the hydrodynamic model) the density perturbation ND = N_EQ (n~x + n~y +
 * n~z) / 3 in m^-3, zero outside the metal box (the ND snapshots,
 * snapshot.cu:80-133); comp 7-9 the Drude pole of ex, ey, ez (Px_d, Py_d,
 * Pz_d in the reference's units: the polarisation, or with the hydrodynamic
 * model the electron velocity; main.cu:2962-2964), zero outside the metal box. */
int  fdtd_get_slice(fdtd_sim *sim, int comp, int axis, int index, float *out);
int  fdtd_set_field(fdtd_sim *sim, int comp, const float *in);

/* Checkpoint / resume (no reference counterpart: the reference keeps its
 * state in process globals for one run of main(), global_var.h:32-815).
 * fdtd_checkpoint_size gives the bytes of an image of the whole simulation
 * state at the current step: masks and design densities, E and H, the CPML
 * psi slabs, the dispersive (and hydrodynamic) state, the 1D line, the DFT
 * sample ring and accumulators, and the point-monitor samples.
 * fdtd_checkpoint_save writes it into buf (at least that many bytes);
 * fdtd_checkpoint_load restores it into a handle created with the same
 * parameters (other materials are replaced by the saved ones), after which
 * stepping gives the same bits as the saving handle would have.  Loading
 * clears a NaN/Inf flag. */
int  fdtd_checkpoint_size(fdtd_sim *sim, size_t *bytes);
int  fdtd_checkpoint_save(fdtd_sim *sim, void *buf, size_t bytes);
int  fdtd_checkpoint_load(fdtd_sim *sim, const void *buf, size_t bytes);

/* 1D incident line (e_inc, h_inc), inc_length values each. */
int  fdtd_get_line(fdtd_sim *sim, double *e_inc, double *h_inc);

/* Frequencies (Hz) and the source spectrum E_incident (re, im pairs). */
int  fdtd_get_spectrum_setup(const fdtd_sim *sim, double *freq, double *e_incident);

/* Accumulated DFT of one monitor face (0..11: box*6 + axis*2 + side, box 0
 * = scattering, 1 = absorption; axis 0..2 = x..z; side 0 near, 1 far) and
 * one tangential component (0..3 = E_a, E_b, H_a, H_b with (a,b) the cyclic
 * pair after the normal axis).  out holds NUM_freq*na*nb (re, im) pairs,
 * b fastest.  Pending samples are contracted first. */
int  fdtd_get_dft(fdtd_sim *sim, int face, int comp, double *out);

/* CalculateAbsScatt (Spectral.cu:411-1177) in FP64: per frequency the
 * scattering, absorption and extinction cross sections (normalised by
 * xs_area) and the 12 raw face fluxes raw[face*NUM_freq + m].  Any pointer may
 * be NULL. */
int  fdtd_flux(fdtd_sim *sim, double *sca, double *abs_, double *ext, double *raw);

/* fdtd_flux without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (set-up of HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The face sums are
 * reduced in a fixed device array of the library (12 * FDTD_NOALLOC_MAX_FREQ
 * doubles per device, held until the process ends) instead of a buffer of the
 * handle; the results equal fdtd_flux's bit for bit.  raw (12 * NUM_freq
 * doubles) is required; NUM_freq above FDTD_NOALLOC_MAX_FREQ returns
 * FDTD_ERR_INVALID.  Calls that overlap on one device share that array and
 * race: serialise them. */
#define FDTD_NOALLOC_MAX_FREQ 256
int  fdtd_flux_noalloc(fdtd_sim *sim, double *sca, double *abs_, double *ext, double *raw);

/* Reflectance and transmittance of a periodic (periodic_xy) simulation
 * (Reflectance_XZ, main.cu:3195-3330, and the plane DFTs of Fourier_Transform,
 * Spectral.cu:8-130).  The reflection plane is k_r = Ntfsf_z - sca_offset on
 * the scattered-field side of the source plane, the transmission plane k_t =
 * NZ - Ntfsf_z - sca_offset; their tangential E and H are the z faces of the
 * DFT monitors (fdtd_get_dft faces 4 and 5).  Per frequency: R = reflected
 * and T = transmitted power through the planes over the incident power, the
 * incident power taken from the DFT of the 1D line at the source plane with
 * the monitors' stride and phase (physical normalisation); ref_r and ref_t
 * are the reference's printed columns Reflectivity/Amplitude and
 * Transmit/Amplitude: sum Re(Ey Hx* - Ex Hy*) dA / Area over each plane
 * divided by |E_incident|^2 of the source spectrum (main.cu:3026-3034).  Any
 * pointer may be NULL. */
int  fdtd_reflectance(fdtd_sim *sim, double *R, double *T, double *ref_r, double *ref_t);
/* Metal layer for periodic simulations (the inf_disp_slab masks,
 * setup.cu:156-200): every E component at k in [k0, k1) over the whole x-y
 * extent takes the configured material.  The layer must lie between the
 * source plane and the transmission plane (k0 > Ntfsf_z, k1 <= k_t). */
int  fdtd_add_layer(fdtd_sim *sim, int k0, int k1);

/* Device time of the field update of fdtd_step accumulated since the last
 * reset (HIP events on the solver stream; no reference counterpart).  Returns
 * the time and count recorded so far, then resets and sets the mode: 0 off,
 * 1 all field-update kernels of a step, 2 the clean-box launch of the
 * row-blocked kernel alone (the dominant one; same as 1 when there is none),
 * 3 the K-sample DFT contractions (twiddles and contraction kernel). */
int  fdtd_kernel_timing(fdtd_sim *sim, int enable, double *total_ms, long *launches);

/*
 * Batched sweeps (no reference counterpart: the reference runs one
 * simulation per process, main.cu:2724-2934).  A batch advances many
 * independent simulations of any grid sizes together: each kernel of a step
 * runs once for all of them (the blocks of every simulation in one grid), and
 * the launches of one DFT sampling period are captured in a HIP graph and
 * replayed.  The field update is fdtd_step's: the row-blocked kernel with the
 * TFSF fix-up where the simulation's grid allows it, else the general fused
 * kernel; the tile of each grid size's dominant launch is picked by the
 * autotuner (timed on the device at fdtd_batch_create, remembered per process
 * and, when the environment variable FDTD_TUNE_FILE names a file, across
 * processes).
 *
 * The batch does not own the simulations, which must outlive it.  They must
 * be at the same step t, share dft_stride and dft_block, and have no
 * hydrodynamic model.  Between fdtd_batch_run calls each simulation can be
 * read (fdtd_flux, fdtd_get_field ...) or stepped on its own.
 */
typedef struct fdtd_batch fdtd_batch;
#define FDTD_BATCH_NO_GRAPH 1   /* launch every step directly instead of replaying a graph */
#define FDTD_BATCH_NO_TUNE  2   /* the simulations' own tiles (those fdtd_step uses)       */
#define FDTD_BATCH_GENERAL  4   /* general fused kernel for every simulation               */
int  fdtd_batch_create(fdtd_sim *const *sims, int n, int flags, fdtd_batch **out);
/* Advances every simulation by nsteps full steps (as fdtd_step does). */
int  fdtd_batch_run(fdtd_batch *b, int nsteps);
/* Field-update path and tile of simulation i: path 1 = row-blocked kernel
 * (a = rows per thread, c = x chunk of the clean box), 0 = general kernel
 * (a = block rows, tj - 1 owned, c = x chunk). */
int  fdtd_batch_tile(const fdtd_batch *b, int i, int *path, int *a, int *c);
const char *fdtd_batch_last_error(const fdtd_batch *b);
int  fdtd_batch_destroy(fdtd_batch *b);

#ifdef __cplusplus
}
#endif
#endif /* FDTD_H_ */
