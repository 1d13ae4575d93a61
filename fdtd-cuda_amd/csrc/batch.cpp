// Batched sweeps (fdtd_batch_* in include/fdtd.h): many independent
// simulations of any grid sizes advanced together.  No reference counterpart:
// the reference runs one simulation per process (main.cu:2724-2934); each
// batched step is that loop body for every simulation at once.
//
// Per step every kernel runs ONCE for all simulations: the 1D lines (one
// block each), the general fused field update with TFSF inside (the blocks of
// every simulation in one grid, one launch per tile height), the dispersive
// boxes and, on sampling steps, the monitor samples.  Per-simulation
// arguments live in device memory, one array per time parity, and each block
// finds its simulation in a block map.  The only arguments that change from
// one sampling period to the next (the step index of the source pulse, the
// ring row of the sample) are read from a two-int device counter that the
// last launch of the period advances; so the launches of one period are
// captured once in a HIP graph and replayed.  The K-sample DFT contraction
// runs between replays (every dft_block samples) with the simulation's own
// launches.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <tuple>

#include "sim.h"

using namespace fdtd_host;

struct fdtd_batch {
    std::vector<fdtd_sim *> sims;
    int flags = 0;
    std::string err = "no error";
    hipStream_t stream = nullptr;
    std::vector<int> tj, ichunk;            // tile per simulation
    // per-simulation arguments, [t & 1]
    FusedItem *fused[2] = {};
    LineDev *line[2] = {};
    DispDev *disp[2] = {};
    SampleItem *samp[2] = {};
    int *dyn = nullptr;                     // {t of the next step, ring row of the next sample}
    // block maps: fused per tile height (4, 8, 16), dispersive, samples
    unsigned *fmap[3] = {}, *dmap = nullptr, *smap = nullptr;
    std::vector<unsigned> h_fmap[3], h_dmap, h_smap;
    size_t cap_fmap[3] = {}, cap_dmap = 0, cap_smap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    std::vector<void *> allocs;
};

namespace {

thread_local std::string g_batch_err = "no error";
const int kTileRows[3] = {4, 8, 16};

int berr(fdtd_batch *b, int code, const std::string &m) {
    if (b) b->err = m;
    else g_batch_err = m;
    return code;
}

#define BCHK(b, expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return berr(b, FDTD_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int balloc(fdtd_batch *b, T **p, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
        return berr(b, FDTD_ERR_ALLOC, "hipMalloc failed for batch arguments");
    b->allocs.push_back(q);
    *p = (T *)q;
    return FDTD_OK;
}

int tile_index(int tj) { return tj == 4 ? 0 : (tj == 8 ? 1 : (tj == 16 ? 2 : -1)); }

// ---------------------------------------------------------------- autotuner

// Tile per grid size, remembered for the process and optionally in the text
// file named by FDTD_TUNE_FILE (lines "NX NY NZ tj ichunk").
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_tune;
bool g_tune_loaded = false;

void tune_load() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *f = std::getenv("FDTD_TUNE_FILE");
    if (!f) return;
    std::ifstream in(f);
    int nx, ny, nz, tj, ic;
    while (in >> nx >> ny >> nz >> tj >> ic)
        if (tile_index(tj) >= 0 && ic > 0) g_tune[{nx, ny, nz}] = {tj, ic};
}

void tune_save() {
    const char *f = std::getenv("FDTD_TUNE_FILE");
    if (!f) return;
    std::ofstream out(f);
    for (auto &kv : g_tune)
        out << std::get<0>(kv.first) << ' ' << std::get<1>(kv.first) << ' ' << std::get<2>(kv.first) << ' '
            << kv.second.first << ' ' << kv.second.second << '\n';
}

// Fused-kernel arguments of simulation s for the step whose index has parity
// par (the handle's parities flip with every fused step).
FusedItem fused_item(fdtd_sim *s, int par, int tj, int ichunk) {
    const int flip = (s->t & 1) ^ par;
    const int q = s->p ^ flip, pp = s->pp ^ flip;
    s->le ^= flip;
    StepDev d = make_step(s, true);
    s->le ^= flip;
    for (int c = 0; c < 3; c++) {
        d.Ein[c] = s->E[q][c]; d.Hin[c] = s->H[q][c];
        d.Eout[c] = s->E[1 - q][c]; d.Hout[c] = s->H[1 - q][c];
    }
    for (int a = 0; a < 6; a++) { d.psiH_in.p[a] = s->psiH[pp][a]; d.psiH_out.p[a] = s->psiH[1 - pp][a]; d.psiE.p[a] = s->psiE[a]; }
    FusedItem it{};
    it.s = d;
    it.ichunk = ichunk;
    fused_batch_blocks(s->g, tj, ichunk, &it.pc, &it.nkt);
    return it;
}

// Times every candidate tile on simulation s and keeps the fastest per cell.
// The kernel writes only the handle's spare time level and psi_E (in place),
// so psi_E is saved first and restored after; the live state is untouched.
// Each candidate runs as a batched launch of `rep` copies of the simulation's
// blocks (identical results written by each copy), so that small grids are
// timed on a filled device as they run in a batch.
int autotune(fdtd_batch *b, fdtd_sim *s, int *tj_out, int *ic_out) {
    const auto key = std::make_tuple(s->g.NX, s->g.NY, s->g.NZ);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load();
        auto it = g_tune.find(key);
        if (it != g_tune.end()) { *tj_out = it->second.first; *ic_out = it->second.second; return FDTD_OK; }
    }
    BCHK(b, hipStreamSynchronize(s->stream));
    std::vector<void *> tmp;
    auto cleanup = [&]() { for (void *p : tmp) (void)hipFree(p); };
    float *saved[6] = {};
    for (int a = 0; a < 6; a++) {
        void *p = nullptr;
        if (hipMalloc(&p, s->esz[a] * sizeof(float)) != hipSuccess) { cleanup(); return berr(b, FDTD_ERR_ALLOC, "autotune: hipMalloc failed"); }
        tmp.push_back(p);
        saved[a] = (float *)p;
        if (hipMemcpyAsync(saved[a], s->psiE[a], s->esz[a] * sizeof(float), hipMemcpyDeviceToDevice, b->stream) != hipSuccess) {
            cleanup();
            return berr(b, FDTD_ERR_DEVICE, "autotune: psi copy failed");
        }
    }
    FusedItem *item_d = nullptr;
    unsigned *map_d = nullptr;
    const size_t map_cap = 1u << 20;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FDTD_OK;
    if (hipMalloc((void **)&item_d, sizeof(FusedItem)) != hipSuccess) rc = FDTD_ERR_ALLOC;
    else tmp.push_back(item_d);
    if (!rc && hipMalloc((void **)&map_d, map_cap * sizeof(unsigned)) != hipSuccess) rc = FDTD_ERR_ALLOC;
    else if (!rc) tmp.push_back(map_d);
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = FDTD_ERR_DEVICE;
    double best = 1e30;
    int btj = 8, bic = 32;
    const int chunks[4] = {8, 16, 32, 64};
    for (int ti = 0; ti < 3 && !rc; ti++)
        for (int ci = 0; ci < 4 && !rc; ci++) {
            const int tj = kTileRows[ti], ic = chunks[ci];
            const FusedItem it = fused_item(s, s->t & 1, tj, ic);
            const long nb = it.pc.start[1];
            if (nb <= 0 || nb >= (1L << kBatchLocalBits)) continue;
            const int rep = (int)std::max(1L, std::min(16L, (4096 + nb - 1) / nb));
            std::vector<unsigned> map;
            for (int r = 0; r < rep; r++)
                for (long L = 0; L < nb; L++) map.push_back((unsigned)L);
            if (map.size() > map_cap) continue;
            if (hipStreamSynchronize(b->stream) != hipSuccess ||
                hipMemcpy(item_d, &it, sizeof(it), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(map_d, map.data(), map.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
                rc = FDTD_ERR_DEVICE;
                break;
            }
            const int nblk = (int)map.size();
            bool ok = launch_fused_batch(item_d, map_d, nblk, tj, b->stream) == hipSuccess &&
                      hipEventRecord(e0, b->stream) == hipSuccess;
            for (int r = 0; r < 3 && ok; r++) ok = launch_fused_batch(item_d, map_d, nblk, tj, b->stream) == hipSuccess;
            ok = ok && hipEventRecord(e1, b->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float ms = 0;
            if (!ok || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = FDTD_ERR_DEVICE; break; }
            const double per = ms / (3.0 * rep);
            if (per < best) { best = per; btj = tj; bic = ic; }
        }
    // restore psi_E whatever happened above
    for (int a = 0; a < 6; a++)
        if (hipMemcpyAsync(s->psiE[a], saved[a], s->esz[a] * sizeof(float), hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
            rc = FDTD_ERR_DEVICE;
    if (hipStreamSynchronize(b->stream) != hipSuccess) rc = FDTD_ERR_DEVICE;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    cleanup();
    if (rc) return berr(b, rc, "autotune failed");
    *tj_out = btj;
    *ic_out = bic;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = {btj, bic};
    tune_save();
    return FDTD_OK;
}

// ---------------------------------------------------------------- launches

// Uploads every per-simulation argument for both parities and rebuilds the
// block maps; a changed map invalidates the captured graph.
int upload_items(fdtd_batch *b) {
    const int n = (int)b->sims.size();
    std::vector<unsigned> fm[3], dm, sm;
    for (int par = 0; par < 2; par++) {
        std::vector<FusedItem> fi(n);
        std::vector<LineDev> li(n);
        std::vector<DispDev> di(n);
        std::vector<SampleItem> si(n);
        for (int i = 0; i < n; i++) {
            fdtd_sim *s = b->sims[i];
            const int flip = (s->t & 1) ^ par;
            fi[i] = fused_item(s, par, b->tj[i], b->ichunk[i]);
            const int q = s->p ^ flip;
            s->le ^= flip;
            li[i] = make_line(s);
            s->le ^= flip;
            di[i] = make_disp(s, 1 - q);
            SampleItem &S = si[i];
            S = SampleItem{};
            S.g = s->g;
            S.fd = s->fd;
            for (int c = 0; c < 3; c++) { S.E[c] = s->E[1 - q][c]; S.H[c] = s->H[1 - q][c]; }
            S.X = s->X;
            S.mtot = s->mtot;
            if (par == 0) {
                const long nb = fi[i].pc.start[1];
                const int g = tile_index(b->tj[i]);
                for (long L = 0; L < nb; L++) fm[g].push_back(((unsigned)i << kBatchLocalBits) | (unsigned)L);
                const long nd = (s->nbox + 255) / 256;
                for (long L = 0; L < nd; L++) dm.push_back(((unsigned)i << kBatchLocalBits) | (unsigned)L);
                const long ns = (s->fd.mE + 255) / 256;
                for (long L = 0; L < ns; L++) sm.push_back(((unsigned)i << kBatchLocalBits) | (unsigned)L);
            }
        }
        BCHK(b, hipMemcpy(b->fused[par], fi.data(), n * sizeof(FusedItem), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->line[par], li.data(), n * sizeof(LineDev), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->disp[par], di.data(), n * sizeof(DispDev), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->samp[par], si.data(), n * sizeof(SampleItem), hipMemcpyHostToDevice));
    }
    auto put = [&](std::vector<unsigned> &h, unsigned *&d, size_t &cap, std::vector<unsigned> &v) -> int {
        if (v == h && d) return FDTD_OK;
        b->graph_valid = false;
        if (v.size() > cap || !d) {
            // maps are re-allocated only when they grow; old buffers stay owned by the batch
            if (balloc(b, &d, v.size())) return FDTD_ERR_ALLOC;
            cap = std::max<size_t>(v.size(), 1);
        }
        if (!v.empty()) BCHK(b, hipMemcpy(d, v.data(), v.size() * sizeof(unsigned), hipMemcpyHostToDevice));
        h = v;
        return FDTD_OK;
    };
    for (int g = 0; g < 3; g++)
        if (int rc = put(b->h_fmap[g], b->fmap[g], b->cap_fmap[g], fm[g])) return rc;
    if (int rc = put(b->h_dmap, b->dmap, b->cap_dmap, dm)) return rc;
    if (int rc = put(b->h_smap, b->smap, b->cap_smap, sm)) return rc;
    return FDTD_OK;
}

// One step for every simulation: par = parity of its index, off = its offset
// from dyn's step, sidx = ring row offset of its sample or -1.
int enqueue_step(fdtd_batch *b, int par, int off, int sidx) {
    const int n = (int)b->sims.size();
    BCHK(b, launch_line_batch(b->line[par], n, b->dyn, off, b->stream));
    for (int g = 0; g < 3; g++)
        BCHK(b, launch_fused_batch(b->fused[par], b->fmap[g], (int)b->h_fmap[g].size(), kTileRows[g], b->stream));
    BCHK(b, launch_disp_batch(b->disp[par], b->dmap, (int)b->h_dmap.size(), b->stream));
    if (sidx >= 0) BCHK(b, launch_sample_batch(b->samp[par], b->smap, (int)b->h_smap.size(), b->dyn, sidx, b->stream));
    return FDTD_OK;
}

// Host bookkeeping of one step of every simulation (what fdtd_step does after its launches).
void mirror_step(fdtd_batch *b, bool sampled) {
    for (fdtd_sim *s : b->sims) {
        if (sampled) {
            s->slot_t[s->slot] = s->t;
            s->slot++;
            s->samples++;
        }
        s->p ^= 1;
        s->pp ^= 1;
        s->le ^= 1;
        s->t++;
    }
}

// Contracts every simulation's pending samples on the batch stream and
// rewinds the device ring row.
int contract_all(fdtd_batch *b) {
    const int slot = b->sims[0]->slot;
    if (slot == 0) return FDTD_OK;
    for (fdtd_sim *s : b->sims) {
        hipStream_t own = s->stream;
        s->stream = b->stream;
        const int rc = contract(s);
        s->stream = own;
        if (rc) return berr(b, rc, "contraction failed: " + s->err);
    }
    BCHK(b, launch_batch_advance(b->dyn, 0, -slot, b->stream));
    return FDTD_OK;
}

int period_of(const fdtd_sim *s) {
    const int stride = s->S.P.dft_stride;
    return stride % 2 == 0 ? stride : 2 * stride;
}

// Captures one period of steps (starting at a step index divisible by the
// period, so parity and sampling offsets are fixed) and the counter advance.
int capture(fdtd_batch *b) {
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    const fdtd_sim *s0 = b->sims[0];
    const int P = period_of(s0), stride = s0->S.P.dft_stride;
    const bool samples = s0->S.P.StaticField == 0;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int sidx = 0, rc = FDTD_OK;
    for (int o = 0; o < P && !rc; o++) {
        const bool samp = samples && o % stride == 0;
        rc = enqueue_step(b, o & 1, o, samp ? sidx : -1);
        if (samp) sidx++;
    }
    if (!rc && launch_batch_advance(b->dyn, P, sidx, b->stream) != hipSuccess) rc = FDTD_ERR_DEVICE;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : berr(b, FDTD_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    b->graph = g;
    BCHK(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->graph_valid = true;
    return FDTD_OK;
}

int check_sims(fdtd_batch *b) {
    const fdtd_sim *s0 = b->sims[0];
    for (fdtd_sim *s : b->sims) {
        if (s->t != s0->t) return berr(b, FDTD_ERR_INVALID, "batch: simulations must be at the same step");
        if (s->S.P.dft_stride != s0->S.P.dft_stride || s->K != s0->K || s->S.P.StaticField != s0->S.P.StaticField)
            return berr(b, FDTD_ERR_INVALID, "batch: simulations must share dft_stride, dft_block and StaticField");
        if (s->S.hydro) return berr(b, FDTD_ERR_UNSUPPORTED, "batch: hydrodynamic simulations are not batched");
        if (s->probe || !s->sampling) return berr(b, FDTD_ERR_INVALID, "batch: simulation is inside an adjoint run");
    }
    return FDTD_OK;
}

}  // namespace

int fdtd_batch_create(fdtd_sim *const *sims, int n, int flags, fdtd_batch **out) {
    if (!out) return berr(nullptr, FDTD_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!sims || n < 1 || n >= (1 << (32 - kBatchLocalBits))) return berr(nullptr, FDTD_ERR_INVALID, "batch: 1 to 1023 simulations");
    fdtd_batch *b = new (std::nothrow) fdtd_batch();
    if (!b) return berr(nullptr, FDTD_ERR_ALLOC, "host allocation failed");
    auto fail = [&](int rc) {
        std::string m = b->err;
        fdtd_batch_destroy(b);
        return berr(nullptr, rc, m);
    };
    b->flags = flags;
    for (int i = 0; i < n; i++) {
        if (!sims[i]) return fail(berr(b, FDTD_ERR_INVALID, "batch: null simulation"));
        b->sims.push_back(sims[i]);
    }
    if (int rc = check_sims(b)) return fail(rc);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(berr(b, FDTD_ERR_DEVICE, "hipStreamCreate failed"));
    int rc;
    for (int par = 0; par < 2; par++)
        if ((rc = balloc(b, &b->fused[par], n)) || (rc = balloc(b, &b->line[par], n)) || (rc = balloc(b, &b->disp[par], n)) ||
            (rc = balloc(b, &b->samp[par], n)))
            return fail(rc);
    if ((rc = balloc(b, &b->dyn, 2))) return fail(rc);
    b->tj.assign(n, 8);
    b->ichunk.assign(n, 32);
    if (!(flags & FDTD_BATCH_NO_TUNE))
        for (int i = 0; i < n; i++)
            if ((rc = autotune(b, b->sims[i], &b->tj[i], &b->ichunk[i]))) return fail(rc);
    for (int i = 0; i < n; i++) {
        const long nb = fused_batch_blocks(b->sims[i]->g, b->tj[i], b->ichunk[i], nullptr, nullptr);
        if (nb <= 0 || nb >= (1L << kBatchLocalBits) || (b->sims[i]->nbox + 255) / 256 >= (1L << kBatchLocalBits) ||
            (b->sims[i]->fd.mE + 255) / 256 >= (1L << kBatchLocalBits))
            return fail(berr(b, FDTD_ERR_UNSUPPORTED, "batch: grid too large for the block map"));
    }
    *out = b;
    return FDTD_OK;
}

int fdtd_batch_run(fdtd_batch *b, int nsteps) {
    if (!b || nsteps < 0) return FDTD_ERR_INVALID;
    if (int rc = check_sims(b)) return rc;
    if (nsteps == 0) return FDTD_OK;
    // every simulation's own queue drains first; pending samples are contracted so all rings start empty
    for (fdtd_sim *s : b->sims) {
        if (int rc = contract(s)) return berr(b, rc, "contraction failed: " + s->err);
        BCHK(b, hipStreamSynchronize(s->stream));
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    if (int rc = upload_items(b)) return rc;
    fdtd_sim *s0 = b->sims[0];
    const int dyn0[2] = {s0->t, 0};
    BCHK(b, hipMemcpy(b->dyn, dyn0, sizeof(dyn0), hipMemcpyHostToDevice));
    const int P = period_of(s0), stride = s0->S.P.dft_stride, K = s0->K;
    const bool samples = s0->S.P.StaticField == 0;
    const int per_period = samples ? P / stride : 0;
    const bool use_graph = !(b->flags & FDTD_BATCH_NO_GRAPH);
    int done = 0;
    while (done < nsteps) {
        if (use_graph && s0->t % P == 0 && nsteps - done >= P && per_period <= K) {
            if (s0->slot + per_period > K)
                if (int rc = contract_all(b)) return rc;
            if (!b->graph_valid)
                if (int rc = capture(b)) return rc;
            BCHK(b, hipGraphLaunch(b->exec, b->stream));
            for (int o = 0; o < P; o++) mirror_step(b, samples && o % stride == 0);
            done += P;
        } else {
            const bool samp = samples && s0->t % stride == 0;
            if (samp && s0->slot == K)
                if (int rc = contract_all(b)) return rc;
            if (int rc = enqueue_step(b, s0->t & 1, 0, samp ? 0 : -1)) return rc;
            BCHK(b, launch_batch_advance(b->dyn, 1, samp ? 1 : 0, b->stream));
            mirror_step(b, samp);
            done++;
        }
        if (s0->slot == K)
            if (int rc = contract_all(b)) return rc;
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    return FDTD_OK;
}

int fdtd_batch_tile(const fdtd_batch *b, int i, int *tj, int *ichunk) {
    if (!b || i < 0 || i >= (int)b->sims.size()) return FDTD_ERR_INVALID;
    if (tj) *tj = b->tj[i];
    if (ichunk) *ichunk = b->ichunk[i];
    return FDTD_OK;
}

const char *fdtd_batch_last_error(const fdtd_batch *b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int fdtd_batch_destroy(fdtd_batch *b) {
    if (!b) return FDTD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return FDTD_OK;
}
