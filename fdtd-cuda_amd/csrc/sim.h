// The simulation handle behind the C ABI and the host helpers that build
// kernel arguments from it, shared by the single-simulation API (api.cpp) and
// batched sweeps (batch.cpp).  Not part of the public ABI.
#ifndef FDTD_SIM_H_
#define FDTD_SIM_H_

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fdtd.h"
#include "host.h"
#include "internal.h"

struct fdtd_sim {
    fdtd::Setup S;
    std::string err = "no error";
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    Geom g{};
    // fields: [parity][component]
    float *E[2][3] = {}, *H[2][3] = {};
    int p = 0;                      // current field parity
    float *psiH[2][6] = {}, *psiE[6] = {};
    int pp = 0;                     // current psi_H parity
    // CPML tables (float)
    float *ik_e[3] = {}, *ik_h[3] = {}, *be[3] = {}, *ce[3] = {}, *bh[3] = {}, *ch[3] = {};
    // 1D line
    double *line_e[2] = {}, *line_h = nullptr, *mur = nullptr;
    int le = 0;
    // dispersive box
    int box_lo[3] = {0, 0, 0}, box_n[3] = {0, 0, 0};
    long nbox = 0;
    uint8_t *disp_id[3] = {};
    float *disp_st[3][8] = {};
    float *h_vnew[3] = {}, *h_nd[3] = {}, *h_ndp[3] = {};
    float *nl_gd[3] = {};           // NONLOCAL == 1: grad(div P_d) per box cell (scratch of each step)
    int n_disp[3] = {0, 0, 0};
    // DFT
    FaceDev fd{};
    long mtot = 0;
    float *X = nullptr;             // [K][mtot]
    double *acc = nullptr;          // [2F][mtot]
    double *W = nullptr;            // [2][K][2F] FP64 twiddles (exact mode)
    void *Wf = nullptr;             // split-BF16 twiddle fragments (MFMA mode)
    double *freq_d = nullptr;
    double *flux_d = nullptr;       // [12][F]
    int K = 64, slot = 0;
    std::vector<int> slot_t;
    int samples = 0;
    // design region and adjoint
    int des_lo[3] = {0, 0, 0}, des_n[3] = {0, 0, 0};
    std::vector<float> des_rho;     // design densities (i-major, k fastest)
    float *rho_box = nullptr;       // density per dispersive-box cell (1 outside the design block)
    bool sampling = true;           // DFT sampling on (off while the adjoint recomputes a segment)
    float *probe = nullptr;         // E(r0) per step while the adjoint's forward runs
    float *mon = nullptr;           // point monitor: E at mon_p for every step since mon_t0 (mon_cap steps)
    long mon_p = 0;
    int mon_t0 = 0, mon_cap = 0, mon_cap_alloc = 0;
    long probe_p = 0;
    int probe_t0 = 0;               // step whose E^{n+1} lands in probe[0..2]
    float *traj_step[3] = {};       // dispersive kernel writes E^{n+1} here when set
    float *traj_first[2][3] = {};   // and E^{n-1}, E^n here for the first recorded step
    struct Adj {
        bool ready = false;
        float *E[2][3] = {}, *H[2][3] = {};
        float *psiE[2][6] = {}, *psiH[2][6] = {};
        float *st[3][7] = {};
        float *Qb[3] = {}, *Eb[3] = {};
        float *dq[3] = {}, *db[3] = {};   // Qb - Ceh Ehat, Eb - Cee Ehat per box cell
        double *grad = nullptr;
        bool rows = false;          // transposed step on the clean box by the row-blocked kernel
        bool rows_slabs = false;    // and on the x and y slabs by its slab variants
        float *corr = nullptr;      // amounts of the dispersive-cell correction (6 per cell of the box grown by one)
        unsigned *aux_d = nullptr;  // its table image (adjoint psi pointers)
    } adj;
    // misc
    float *dense = nullptr;         // NX*NY*NZ staging for field copies
    size_t fsz = 0;                 // floats per padded field allocation
    long fbase = 0;                 // offset of element (0,0,0) in it
    size_t hsz[6] = {}, esz[6] = {};
    int t = 0;
    int device = 0;                 // HIP device current at fdtd_create (streams and memory live there)
    // divergence watch: a pinned host word the line probe and the DFT sample
    // kernels set to 1 on a NaN or Inf (status_d: its device address)
    unsigned *status_h = nullptr, *status_d = nullptr;
    long probe_cell = 0;            // ex probe of the line kernel (main.cu:2728: NX/2, NY/2, NZ/2 - NZ/4)
    int tj = 8, ichunk = 32;
    int ctj = 4;                    // tile rows of the general kernel on the corner columns (FDTD_CTJ: 4, 8, 16)
    // fused step: row-blocked kernel over the whole grid + TFSF fix-up, or the general kernel alone
    bool use_rows = false;
    bool has_clean = false;         // clean box non-empty: clean variant there, general variant on the rest
    int cbox[4] = {};               // clean box {ilo, ihi, jlo, jhi}
    int xbox[2][4] = {}, ybox[2][4] = {}, cbox4[4][4] = {};   // x-general, y-general and corner pieces
    int nxbox = 0, nybox = 0, ncorner = 0;
    float ik0[2][2] = {};           // interior 1/(kappa d): [E, H][x, y]
    unsigned *aux_d = nullptr;      // table image of the row-blocked kernel
    std::vector<uint32_t> aux_img;  // host copy of it
    bool concurrent = true;         // side slabs on a second stream beside the clean box (FDTD_CONCURRENT=0 to serialise)
    hipStream_t stream2 = nullptr;
    // ev_line: the 1D line's completion in a step; the transposed step reuses it
    // for the dispersive update or the clean box (FDTD_ADJ_SCHED 2, 1)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_line = nullptr;
    // side pieces of fdtd_step on separate streams (FDTD_SIDE_STREAMS: 1 = all on stream2, 3 = x
    // slabs after the line on stream2, y slabs on stream3, corner columns on stream4)
    int side_streams = 3;
    bool chain_join = true;         // side streams join through stream2 (FDTD_CHAIN_JOIN=0: each joins the main stream)
    bool post_after_join = true;    // with chained joins, the post-update launch after the join (FDTD_POST_AFTER_JOIN)
    int early_side = 2;             // the next step's y slabs and corners (1), and x slabs (2: the line moves to the
                                    // corner stream), start at this step's join (FDTD_EARLY_SIDE)
    bool early_ready = false;       // ev_early holds that point for the next step of this fdtd_step call
    hipEvent_t ev_early = nullptr;
    hipStream_t stream3 = nullptr, stream4 = nullptr;
    hipEvent_t ev_join3 = nullptr, ev_join4 = nullptr;
    bool early = true;              // TFSF fix and dispersive box ahead of the join when their cells allow (FDTD_EARLY=0 to keep them after)
    std::vector<float> t_ike[3], t_ikh[3], t_be[3], t_ce[3], t_bh[3], t_ch[3];   // host copies of the CPML tables
    int rows = 2, cchunk = 0;       // own rows per thread and x chunk of the row-blocked kernel (clean box;
                                    // 0: 32 with two-tile blocks when that still leaves at least 512
                                    // clean-box blocks, else 16; see plan_rows)
    // row tiles per block of the clean-box launch (FDTD_TILES: 1-3) and of the
    // side-slab launches (FDTD_STILES: 1 or 2); 0: two when the z column takes
    // at most 5 waves, else one (see rows_tiles)
    int tiles = 0, stiles = 0;
    int schunk = 16;                // x chunk of the side slabs
    int adj_schunk = 8;             // x chunk of the side slabs in the transposed step, general kernel (FDTD_ADJ_SCHUNK)
    int adj_rchunk = 8;             // x chunk of the side slabs in the transposed step, row-blocked variants (FDTD_ADJ_RCHUNK)
    int adj_sched = 3;              // transposed dispersive update and correction: 0 main stream, 1 corner stream, 2 update there, correction on main,
                                    // 3 update and correction amounts there, their addition on main (FDTD_ADJ_SCHED)
    bool adj_fuse_inject = true;    // the next step's injection in the correction launch (FDTD_ADJ_FUSE_INJECT)
    bool adj_early = false;         // the next transposed step's slabs start at the join (FDTD_ADJ_EARLY; measured slower)
    bool adj_early_allowed = false; // ... for this sweep's objective cell
    bool adj_early_ready = false;   // ev_early holds that point for the next transposed step
    int adj_tiles = 1;              // row tiles per block of the transposed clean box, two rows (FDTD_ADJ_TILES: 1, 2)
    int adj_stiles = 1;             // the same for its x and y slabs (FDTD_ADJ_STILES: 1, 2)
    int adj_rows = 0;               // rows per thread of the transposed clean box (FDTD_ADJ_ROWS; 0: as the forward step)
    int adj_cchunk = 0;             // its x chunk (FDTD_ADJ_CCHUNK; 0: as the forward clean box)
    int adj_ctj = 8, adj_cich = 2;  // tile rows and x chunk of the transposed corner columns on k_adj_step (FDTD_ADJ_CTJ, FDTD_ADJ_CICH)
    bool adj_corner_rows = true;    // transposed corner columns on the row-blocked kernel (FDTD_ADJ_CORNER_ROWS; 0: k_adj_step)
    int adj_crchunk = 4;            // their x chunk (FDTD_ADJ_CRCHUNK)
    int adj_stj = 16;               // tile rows of the side slabs in the transposed step (FDTD_ADJ_STJ; 4, 8, 16 measured ratios 3.14, 3.16, 3.07 over 2000 steps)
    int xcd = 0;                    // XCD-contiguous block order in the row-blocked kernel (FDTD_XCD=1)
    int timing_what = 1;            // 1: all field-update kernels of a step, 2: the row-blocked kernel only
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t ev_used = 0;
};

namespace fdtd_host {

int set_err(fdtd_sim *s, int code, const std::string &m);

#define HIPCHK(sim, expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fdtd_host::set_err(sim, FDTD_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(fdtd_sim *s, T **ptr, size_t count) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        s->err = std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e);
        return FDTD_ERR_ALLOC;
    }
    s->allocs.push_back(p);
    e = hipMemsetAsync(p, 0, std::max<size_t>(count, 1) * sizeof(T), s->stream);
    if (e != hipSuccess) { s->err = hipGetErrorString(e); return FDTD_ERR_DEVICE; }
    *ptr = (T *)p;
    return FDTD_OK;
}

// Frees a device buffer of the handle (one dalloc'd or uploaded) and clears
// the pointer; the caller has synchronised the device.
template <typename T>
void dfree(fdtd_sim *s, T *&ptr) {
    if (!ptr) return;
    for (size_t n = 0; n < s->allocs.size(); n++)
        if (s->allocs[n] == (void *)ptr) {
            s->allocs.erase(s->allocs.begin() + (long)n);
            (void)hipFree((void *)ptr);
            break;
        }
    ptr = nullptr;
}

// Allocation zero-fills run on the solver stream, so uploads are queued on the
// same stream (and waited for, the host vector being temporary) to keep them
// ordered after the fill.
template <typename T>
int upload(fdtd_sim *s, T **ptr, const std::vector<T> &v) {
    int rc = dalloc(s, ptr, v.size());
    if (rc) return rc;
    HIPCHK(s, hipMemcpyAsync(*ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

// Kernel arguments at the handle's current parity.
StepDev make_step(fdtd_sim *s, bool fused);
LineDev make_line(fdtd_sim *s);
DispDev make_disp(fdtd_sim *s, int parity);
// Row-blocked kernel arguments at the handle's parity (clean-box chunk; boxes and R set by the caller).
RowsArgs make_rows(fdtd_sim *s);
// Cell-set tests of the two-stream schedule (the clean box and the TFSF,
// dispersive and monitor cells; see api.cpp).
int rows_tiles(int NZ);
bool tfsf_in_clean(const fdtd_sim *s);
bool disp_in_clean(const fdtd_sim *s);
bool tfsf_off_clean_edges(const fdtd_sim *s);
// nsteps steps keeping an early start from the step before (fdtd_step clears it first)
int step_steps(fdtd_sim *s, int nsteps);
// FDTD_ERR_NONFINITE (with a message) when a kernel has flagged a NaN or Inf.
int check_finite(fdtd_sim *s);
// DFT bookkeeping: a sampling step, the ring slot just written, the K-sample contraction.
bool sampling_step(const fdtd_sim *s);
int after_sample(fdtd_sim *s);
int contract(fdtd_sim *s);
// Face-flux spectra from the 12 x F face sums Q (host, FP64): normalises Q in
// place by |E_incident|^2 and fills sca/abs/ext (any may be null).
void flux_combine(const fdtd_sim *s, double *Q, double *sca, double *abs_, double *ext);

}  // namespace fdtd_host

#endif
