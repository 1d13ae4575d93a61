// Batched sweeps (fdtd_batch_* in include/fdtd.h): many independent
// simulations of any grid sizes advanced together.  No reference counterpart:
// the reference runs one simulation per process (main.cu:2724-2934); each
// batched step is that loop body for every simulation at once.
//
// Per step every kernel runs ONCE for all simulations that use it: the 1D
// lines (one block each); the field update of fdtd_step's fast path (the
// row-blocked kernel on clean boxes and side slabs, one launch per variant,
// block height and rows per thread; the general kernel on corner columns, or
// on whole grids where the row path does not apply, one launch per tile
// height); the TFSF fix-up; the dispersive boxes and, on sampling steps, the
// monitor samples.  Per-simulation arguments live in device memory, one array
// per time parity, and each block finds its simulation in a block map.  The
// only arguments that change from one sampling period to the next (the step
// index of the source pulse, the ring row of the sample) are read from a
// two-int device counter that the last launch of the period advances; so the
// launches of one period are captured once in a HIP graph and replayed.  The
// K-sample DFT contraction runs between replays (every dft_block samples)
// with the simulation's own launches.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <tuple>

#include "sim.h"

using namespace fdtd_host;

namespace {

// One batched launch of a step: its kernel, shape and block map.
struct Launch {
    enum Type { Rows = 0, Fused = 1, Tfsf = 2, Disp = 3, Sample = 4 };
    int type = 0;
    int a = 0, b = 0, c = 0;             // Rows: W, R, variant; Fused: tile rows
    std::vector<unsigned> map;
    unsigned *d = nullptr;
    bool same(const Launch &o) const { return type == o.type && a == o.a && b == o.b && c == o.c && map == o.map; }
};

}  // namespace

struct fdtd_batch {
    std::vector<fdtd_sim *> sims;
    int flags = 0;
    std::string err = "no error";
    hipStream_t stream = nullptr;
    // tile per simulation: row path (rows per thread, x chunk of the clean box)
    // or general kernel (tile rows, x chunk)
    std::vector<int> rowpath, ta, tb;
    // per-simulation arguments, [t & 1]
    FusedItem *fused[2] = {};
    RowsArgs *rows[2][3] = {};
    LineDev *line[2] = {};
    DispDev *disp[2] = {};
    SampleItem *samp[2] = {};
    int *dyn = nullptr;                     // {t of the next step, ring row of the next sample}
    std::vector<Launch> launches;           // field update, TFSF fix-up, dispersive, samples (in order)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    std::vector<void *> allocs;
    // second stream for the lines, side slabs and corner columns beside the
    // clean boxes (FDTD_CONCURRENT=0: one stream); the fork and join are graph
    // edges in a captured period
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_line = nullptr;
    // y-slab and general-kernel launches on streams of their own (FDTD_SIDE_STREAMS=3; else all on stream2)
    hipStream_t stream3 = nullptr, stream4 = nullptr;
    hipEvent_t ev_join3 = nullptr, ev_join4 = nullptr;
    // TFSF fix-ups (every row-path simulation's planes in its clean box) and
    // dispersive boxes (every simulation on the row path, every box in its
    // clean box) follow the clean boxes without waiting for the side pieces
    bool tf_early = false, dp_early = false;
    // early side start (FDTD_BATCH_EARLY=1; off by default, the 32-rod sweep
    // ran 1-9 % slower with it): the next step's x and y slabs wait for this
    // step's join (ev_early) instead of its post-update launches; early_ok
    // when every simulation allows it, early_ready when the step enqueued
    // just before, in the same run or capture, recorded ev_early
    hipEvent_t ev_early = nullptr;
    bool early = false, early_ok = false, early_ready = false;
};

namespace {

thread_local std::string g_batch_err = "no error";

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

// Runs f with the handle's parities set as for a step whose index has parity par.
template <typename F>
auto at_parity(fdtd_sim *s, int par, F f) {
    const int flip = (s->t & 1) ^ par;
    s->p ^= flip; s->pp ^= flip; s->le ^= flip;
    auto r = f();
    s->p ^= flip; s->pp ^= flip; s->le ^= flip;
    return r;
}

// General-kernel arguments at the handle's current parity (fdtd_step's StepDev).
StepDev fused_step(fdtd_sim *s, bool tfsf_inside) {
    StepDev d = make_step(s, true);
    d.tfsf_on = tfsf_inside ? 1 : 0;
    const int q = s->p, pp = s->pp;
    for (int c = 0; c < 3; c++) {
        d.Ein[c] = s->E[q][c]; d.Hin[c] = s->H[q][c];
        d.Eout[c] = s->E[1 - q][c]; d.Hout[c] = s->H[1 - q][c];
    }
    for (int a = 0; a < 6; a++) { d.psiH_in.p[a] = s->psiH[pp][a]; d.psiH_out.p[a] = s->psiH[1 - pp][a]; d.psiE.p[a] = s->psiE[a]; }
    return d;
}

// General kernel over the whole grid (TFSF inside), tile rows tj and x chunk ic.
FusedItem fused_full(fdtd_sim *s, int tj, int ic) {
    FusedItem it{};
    it.s = fused_step(s, true);
    it.ichunk = ic;
    fused_batch_blocks(s->g, tj, ic, &it.pc, &it.nkt);
    return it;
}

// Row path: the corner columns on the general kernel (TFSF left to the fix-up).
FusedItem fused_corners(fdtd_sim *s) {
    FusedItem it{};
    it.s = fused_step(s, false);
    it.ichunk = 2;
    if (s->ncorner) fused_pieces(s->g, s->ctj, 2, s->cbox4, s->ncorner, &it.pc, &it.nkt);
    return it;
}

// Row path: the clean box (variant 0, R rows, x chunk ic) or the x / y side slabs (variants 1, 2).
RowsArgs rows_item(fdtd_sim *s, int variant, int R, int ic) {
    RowsArgs r = make_rows(s);
    r.xcd = 0;
    r.tiles = rows_batch_tiles(rows_waves(s->g.NZ));
    if (variant == 0) {
        r.ichunk = ic;
        r.npiece = 1;
        std::memcpy(r.box[0], s->cbox, sizeof(s->cbox));
        rows_prepare(r, R, 0);
    } else {
        const int n = variant == 1 ? s->nxbox : s->nybox;
        r.ichunk = s->schunk;
        r.npiece = n;
        if (n) {
            std::memcpy(r.box, variant == 1 ? s->xbox : s->ybox, sizeof(r.box));
            rows_prepare(r, 2, variant);
        }
    }
    return r;
}

long rows_blocks(const RowsArgs &r) { return r.npiece ? r.start[r.npiece] : 0; }

// ---------------------------------------------------------------- autotuner

// Tile per grid size and path, remembered for the process and optionally in
// the text file named by FDTD_TUNE_FILE (lines "NX NY NZ path a b").
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, std::pair<int, int>> g_tune;
bool g_tune_loaded = false;

void tune_load() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *f = std::getenv("FDTD_TUNE_FILE");
    if (!f) return;
    std::ifstream in(f);
    int nx, ny, nz, path, ta, tb;
    while (in >> nx >> ny >> nz >> path >> ta >> tb)
        if (ta > 0 && tb > 0) g_tune[{nx, ny, nz, path}] = {ta, tb};
}

void tune_save() {
    const char *f = std::getenv("FDTD_TUNE_FILE");
    if (!f) return;
    std::ofstream out(f);
    for (auto &kv : g_tune)
        out << std::get<0>(kv.first) << ' ' << std::get<1>(kv.first) << ' ' << std::get<2>(kv.first) << ' '
            << std::get<3>(kv.first) << ' ' << kv.second.first << ' ' << kv.second.second << '\n';
}

// Times every candidate tile of simulation s's dominant launch and keeps the
// fastest per cell: on the row path the clean box (rows per thread 2 or 3 by
// x chunk 8, 16 or 32), else the general kernel over the grid (tile rows 4, 8
// or 16 by x chunk 8 to 64).  The kernels write only the handle's spare time
// level and psi_E (in place), so psi_E is saved first and restored after; the
// live state is untouched.  Each candidate runs as a batched launch of `rep`
// copies of the simulation's blocks (identical results written by each copy)
// so that small grids are timed on a filled device, as they run in a batch.
int autotune(fdtd_batch *b, fdtd_sim *s, bool rowpath, int *a_out, int *b_out) {
    const auto key = std::make_tuple(s->g.NX, s->g.NY, s->g.NZ, rowpath ? 1 : 0);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load();
        auto it = g_tune.find(key);
        if (it != g_tune.end()) { *a_out = it->second.first; *b_out = it->second.second; return FDTD_OK; }
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
    void *item_d = nullptr;
    unsigned *map_d = nullptr;
    const size_t map_cap = 1u << 20;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FDTD_OK;
    if (hipMalloc(&item_d, std::max(sizeof(FusedItem), sizeof(RowsArgs))) != hipSuccess) rc = FDTD_ERR_ALLOC;
    else tmp.push_back(item_d);
    if (!rc && hipMalloc((void **)&map_d, map_cap * sizeof(unsigned)) != hipSuccess) rc = FDTD_ERR_ALLOC;
    else if (!rc) tmp.push_back(map_d);
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = FDTD_ERR_DEVICE;
    double best = 1e30;
    int ba = rowpath ? s->rows : 8, bb = rowpath ? s->cchunk : 32;
    std::vector<std::pair<int, int>> cands;
    if (rowpath) {
        for (int R : {2, 3})
            for (int ic : {8, 16, 32}) cands.push_back({R, ic});
    } else {
        for (int tj : {4, 8, 16})
            for (int ic : {8, 16, 32, 64}) cands.push_back({tj, ic});
    }
    const int W = rows_waves(s->g.NZ);
    for (size_t ci = 0; ci < cands.size() && !rc; ci++) {
        const int ca = cands[ci].first, cb = cands[ci].second;
        long nb;
        hipError_t he = hipSuccess;
        if (rowpath) {
            const RowsArgs r = rows_item(s, 0, ca, cb);
            nb = rows_blocks(r);
            if (hipStreamSynchronize(b->stream) != hipSuccess || hipMemcpy(item_d, &r, sizeof(r), hipMemcpyHostToDevice) != hipSuccess)
                he = hipErrorUnknown;
        } else {
            const FusedItem it = fused_full(s, ca, cb);
            nb = it.pc.start[1];
            if (hipStreamSynchronize(b->stream) != hipSuccess || hipMemcpy(item_d, &it, sizeof(it), hipMemcpyHostToDevice) != hipSuccess)
                he = hipErrorUnknown;
        }
        if (he != hipSuccess) { rc = FDTD_ERR_DEVICE; break; }
        if (nb <= 0 || nb >= (1L << kBatchLocalBits)) continue;
        const int rep = (int)std::max(1L, std::min(16L, (4096 + nb - 1) / nb));
        std::vector<unsigned> map;
        for (int r = 0; r < rep; r++)
            for (long L = 0; L < nb; L++) map.push_back((unsigned)L);
        if (map.size() > map_cap) continue;
        if (hipMemcpy(map_d, map.data(), map.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) { rc = FDTD_ERR_DEVICE; break; }
        const int nblk = (int)map.size();
        auto launch = [&]() {
            return rowpath ? launch_rows_batch((const RowsArgs *)item_d, map_d, nblk, W, ca, 0, b->stream)
                           : launch_fused_batch((const FusedItem *)item_d, map_d, nblk, ca, b->stream);
        };
        bool ok = launch() == hipSuccess && hipEventRecord(e0, b->stream) == hipSuccess;
        for (int r = 0; r < 3 && ok; r++) ok = launch() == hipSuccess;
        ok = ok && hipEventRecord(e1, b->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float ms = 0;
        if (!ok || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = FDTD_ERR_DEVICE; break; }
        const double per = ms / (3.0 * rep);
        if (per < best) { best = per; ba = ca; bb = cb; }
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
    *a_out = ba;
    *b_out = bb;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = {ba, bb};
    tune_save();
    return FDTD_OK;
}

// ---------------------------------------------------------------- launches

Launch &group(std::vector<Launch> &ls, int type, int a, int b, int c) {
    for (Launch &l : ls)
        if (l.type == type && l.a == a && l.b == b && l.c == c) return l;
    ls.push_back(Launch{});
    Launch &l = ls.back();
    l.type = type; l.a = a; l.b = b; l.c = c;
    return l;
}

void add_blocks(Launch &l, int sim, long n) {
    for (long L = 0; L < n; L++) l.map.push_back(((unsigned)sim << kBatchLocalBits) | (unsigned)L);
}

// Uploads every per-simulation argument for both parities and rebuilds the
// launch list; a changed list invalidates the captured graph.
int upload_items(fdtd_batch *b) {
    const int n = (int)b->sims.size();
    std::vector<Launch> ls;
    for (int par = 0; par < 2; par++) {
        std::vector<FusedItem> fi(n);
        std::vector<RowsArgs> ri[3] = {std::vector<RowsArgs>(n), std::vector<RowsArgs>(n), std::vector<RowsArgs>(n)};
        std::vector<LineDev> li(n);
        std::vector<DispDev> di(n);
        std::vector<SampleItem> si(n);
        for (int i = 0; i < n; i++) {
            fdtd_sim *s = b->sims[i];
            if (b->rowpath[i]) {
                fi[i] = at_parity(s, par, [&] { return fused_corners(s); });
                for (int v = 0; v < 3; v++) ri[v][i] = at_parity(s, par, [&] { return rows_item(s, v, b->ta[i], b->tb[i]); });
            } else {
                fi[i] = at_parity(s, par, [&] { return fused_full(s, b->ta[i], b->tb[i]); });
            }
            li[i] = at_parity(s, par, [&] { return make_line(s); });
            const int q = at_parity(s, par, [&] { return s->p; });
            di[i] = make_disp(s, 1 - q);
            SampleItem &S = si[i];
            S = SampleItem{};
            S.g = s->g;
            S.fd = s->fd;
            for (int c = 0; c < 3; c++) { S.E[c] = s->E[1 - q][c]; S.H[c] = s->H[1 - q][c]; }
            S.X = s->X;
            S.mtot = s->mtot;
        }
        if (par == 0) {
            // field update: row-blocked launches, then the general kernel; TFSF fix-up; dispersive; samples
            for (int i = 0; i < n; i++) {
                if (!b->rowpath[i]) continue;
                const int W = rows_waves(b->sims[i]->g.NZ);
                add_blocks(group(ls, Launch::Rows, W, b->ta[i], 0), i, rows_blocks(ri[0][i]));
                if (rows_blocks(ri[1][i])) add_blocks(group(ls, Launch::Rows, W, 2, 1), i, rows_blocks(ri[1][i]));
                if (rows_blocks(ri[2][i])) add_blocks(group(ls, Launch::Rows, W, 2, 2), i, rows_blocks(ri[2][i]));
            }
            for (int i = 0; i < n; i++) {
                const int tj = b->rowpath[i] ? b->sims[i]->ctj : b->ta[i];
                const long nb = b->rowpath[i] && !b->sims[i]->ncorner ? 0 : fi[i].pc.start[fi[i].pc.n];
                if (nb) add_blocks(group(ls, Launch::Fused, tj, 0, 0), i, nb);
            }
            for (int i = 0; i < n; i++)
                if (b->rowpath[i]) add_blocks(group(ls, Launch::Tfsf, 0, 0, 0), i, tfsf_fix_blocks(fi[i].s));
            for (int i = 0; i < n; i++) add_blocks(group(ls, Launch::Disp, 0, 0, 0), i, (b->sims[i]->nbox + 255) / 256);
            for (int i = 0; i < n; i++) add_blocks(group(ls, Launch::Sample, 0, 0, 0), i, (b->sims[i]->fd.mE + 255) / 256);
        }
        BCHK(b, hipMemcpy(b->fused[par], fi.data(), n * sizeof(FusedItem), hipMemcpyHostToDevice));
        for (int v = 0; v < 3; v++) BCHK(b, hipMemcpy(b->rows[par][v], ri[v].data(), n * sizeof(RowsArgs), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->line[par], li.data(), n * sizeof(LineDev), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->disp[par], di.data(), n * sizeof(DispDev), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->samp[par], si.data(), n * sizeof(SampleItem), hipMemcpyHostToDevice));
    }
    b->tf_early = b->dp_early = true;
    for (int i = 0; i < n; i++) {
        fdtd_sim *s = b->sims[i];
        if (b->rowpath[i] && !tfsf_in_clean(s)) b->tf_early = false;
        if (!b->rowpath[i] || (s->nbox > 0 && !disp_in_clean(s))) b->dp_early = false;
    }
    b->early_ok = b->tf_early && b->dp_early;
    for (int i = 0; i < n; i++)
        if (!tfsf_off_clean_edges(b->sims[i])) b->early_ok = false;
    bool same = ls.size() == b->launches.size();
    for (size_t q = 0; same && q < ls.size(); q++) same = ls[q].same(b->launches[q]);
    if (same) return FDTD_OK;
    b->graph_valid = false;
    for (Launch &l : ls) {
        if (l.map.empty()) continue;
        if (int rc = balloc(b, &l.d, l.map.size())) return rc;
        BCHK(b, hipMemcpy(l.d, l.map.data(), l.map.size() * sizeof(unsigned), hipMemcpyHostToDevice));
    }
    b->launches = std::move(ls);
    return FDTD_OK;
}

// One step for every simulation under the early side start: the x and y
// slabs (and on one side stream the corner columns) wait for the last step's
// join only, when that step was enqueued just before (early_ready), and so
// run beside its post-update launches, which write no cell they read: every
// simulation is on the row path, its TFSF fix-up cells keep off the clean
// box's outer planes and rows, its dispersive box with a 2-cell margin lies
// in the clean box, the samples read that step's outputs, and the side
// pieces write the other ping-pong buffers.  The lines keep the fork (a line
// overwrites the incident field the last step's TFSF fix-ups read): first on
// the corner stream, ahead of the corner columns, with three side streams,
// after the side pieces with one.  Every side stream joins the x-slab stream,
// the main stream waits once, records ev_early and runs the TFSF fix-ups, the
// dispersive boxes and the samples.
int enqueue_step_early(fdtd_batch *b, int par, int off, int sidx) {
    const int n = (int)b->sims.size();
    const bool three = b->stream3 && b->stream4;
    hipStream_t side_y = three ? b->stream3 : b->stream2, side_c = three ? b->stream4 : b->stream2;
    BCHK(b, hipEventRecord(b->ev_fork, b->stream));
    hipEvent_t e = b->early_ready ? b->ev_early : b->ev_fork;
    b->early_ready = false;
    BCHK(b, hipStreamWaitEvent(b->stream2, e, 0));
    if (three) {
        BCHK(b, hipStreamWaitEvent(b->stream3, e, 0));
        BCHK(b, hipStreamWaitEvent(b->stream4, b->ev_fork, 0));
        BCHK(b, launch_line_batch(b->line[par], n, b->dyn, off, b->stream4));
    }
    for (const Launch &l : b->launches) {
        const int nb = (int)l.map.size();
        if (l.type == Launch::Rows && l.c != 0)
            BCHK(b, launch_rows_batch(b->rows[par][l.c], l.d, nb, l.a, l.b, l.c, l.c == 2 ? side_y : b->stream2));
        if (l.type == Launch::Fused) BCHK(b, launch_fused_batch(b->fused[par], l.d, nb, l.a, side_c));
    }
    if (three) {
        BCHK(b, hipEventRecord(b->ev_join3, b->stream3));
        BCHK(b, hipEventRecord(b->ev_join4, b->stream4));
        BCHK(b, hipStreamWaitEvent(b->stream2, b->ev_join3, 0));
        BCHK(b, hipStreamWaitEvent(b->stream2, b->ev_join4, 0));
    } else {
        // one side stream: the lines follow the side pieces there, behind the fork
        if (e != b->ev_fork) BCHK(b, hipStreamWaitEvent(b->stream2, b->ev_fork, 0));
        BCHK(b, launch_line_batch(b->line[par], n, b->dyn, off, b->stream2));
    }
    BCHK(b, hipEventRecord(b->ev_join, b->stream2));
    for (const Launch &l : b->launches)
        if (l.type == Launch::Rows && l.c == 0)
            BCHK(b, launch_rows_batch(b->rows[par][l.c], l.d, (int)l.map.size(), l.a, l.b, l.c, b->stream));
    BCHK(b, hipStreamWaitEvent(b->stream, b->ev_join, 0));
    BCHK(b, hipEventRecord(b->ev_early, b->stream));
    for (const Launch &l : b->launches) {
        const int nb = (int)l.map.size();
        if (l.type == Launch::Tfsf) BCHK(b, launch_tfsf_fix_batch(b->fused[par], l.d, nb, b->stream));
        if (l.type == Launch::Disp) BCHK(b, launch_disp_batch(b->disp[par], l.d, nb, b->stream));
        if (l.type == Launch::Sample && sidx >= 0) BCHK(b, launch_sample_batch(b->samp[par], l.d, nb, b->dyn, sidx, b->stream));
    }
    b->early_ready = true;
    return FDTD_OK;
}

// One step for every simulation: par = parity of its index, off = its offset
// from dyn's step, sidx = ring row offset of its sample or -1.  The lines,
// the side slabs and the general-kernel launches (corner columns, or whole
// grids off the row path) run on the second stream beside the clean boxes
// (with FDTD_SIDE_STREAMS=3 the y slabs and the general kernel on two more):
// each writes only its own cells and reads the previous step's fields.  The
// TFSF fix-ups and the dispersive boxes follow the clean boxes on the main
// stream, ahead of the join where tf_early and dp_early allow (the same
// order as after it); the samples follow the join.
int enqueue_step(fdtd_batch *b, int par, int off, int sidx) {
    const int n = (int)b->sims.size();
    const bool fork = b->stream2 != nullptr;
    const bool three = fork && b->stream3 && b->stream4;
    if (fork && b->early && b->early_ok && b->ev_early) return enqueue_step_early(b, par, off, sidx);
    hipStream_t side = fork ? b->stream2 : b->stream;
    hipStream_t side_y = three ? b->stream3 : side, side_g = three ? b->stream4 : side;
    const bool tf_early = fork && b->tf_early, dp_early = tf_early && b->dp_early;
    if (fork) {
        BCHK(b, hipEventRecord(b->ev_fork, b->stream));
        BCHK(b, hipStreamWaitEvent(b->stream2, b->ev_fork, 0));
        if (three) {
            BCHK(b, hipStreamWaitEvent(b->stream3, b->ev_fork, 0));
            BCHK(b, hipStreamWaitEvent(b->stream4, b->ev_fork, 0));
        }
    }
    BCHK(b, launch_line_batch(b->line[par], n, b->dyn, off, side));
    if (tf_early) BCHK(b, hipEventRecord(b->ev_line, side));
    for (const Launch &l : b->launches) {
        const int nb = (int)l.map.size();
        if (l.type == Launch::Rows && l.c != 0)
            BCHK(b, launch_rows_batch(b->rows[par][l.c], l.d, nb, l.a, l.b, l.c, l.c == 2 ? side_y : side));
        if (l.type == Launch::Fused) BCHK(b, launch_fused_batch(b->fused[par], l.d, nb, l.a, side_g));
    }
    if (three) {
        BCHK(b, hipEventRecord(b->ev_join3, b->stream3));
        BCHK(b, hipEventRecord(b->ev_join4, b->stream4));
    }
    for (const Launch &l : b->launches)
        if (l.type == Launch::Rows && l.c == 0)
            BCHK(b, launch_rows_batch(b->rows[par][l.c], l.d, (int)l.map.size(), l.a, l.b, l.c, b->stream));
    auto post = [&](bool early) -> int {
        for (const Launch &l : b->launches) {
            const int nb = (int)l.map.size();
            if (l.type == Launch::Tfsf && early == tf_early) BCHK(b, launch_tfsf_fix_batch(b->fused[par], l.d, nb, b->stream));
            if (l.type == Launch::Disp && early == dp_early) BCHK(b, launch_disp_batch(b->disp[par], l.d, nb, b->stream));
            if (l.type == Launch::Sample && !early && sidx >= 0)
                BCHK(b, launch_sample_batch(b->samp[par], l.d, nb, b->dyn, sidx, b->stream));
        }
        return FDTD_OK;
    };
    if (tf_early) {
        BCHK(b, hipStreamWaitEvent(b->stream, b->ev_line, 0));
        if (int rc = post(true)) return rc;
    }
    if (fork) {
        BCHK(b, hipEventRecord(b->ev_join, b->stream2));
        BCHK(b, hipStreamWaitEvent(b->stream, b->ev_join, 0));
    }
    if (three) {
        BCHK(b, hipStreamWaitEvent(b->stream, b->ev_join3, 0));
        BCHK(b, hipStreamWaitEvent(b->stream, b->ev_join4, 0));
    }
    return post(false);
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
    // a captured period starts from the plain fork, and no later step may
    // wait on an event recorded only inside the capture
    b->early_ready = false;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int sidx = 0, rc = FDTD_OK;
    for (int o = 0; o < P && !rc; o++) {
        const bool samp = samples && o % stride == 0;
        rc = enqueue_step(b, o & 1, o, samp ? sidx : -1);
        if (samp) sidx++;
    }
    if (!rc && launch_batch_advance(b->dyn, P, sidx, b->stream) != hipSuccess) rc = FDTD_ERR_DEVICE;
    b->early_ready = false;
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
        if (s->S.hydro || s->S.nonlocal1 || s->S.P.periodic_xy)
            return berr(b, FDTD_ERR_UNSUPPORTED, "batch: nonlocal (NONLOCAL 1 or 3) and periodic simulations are not batched");
        if (s->probe || !s->sampling) return berr(b, FDTD_ERR_INVALID, "batch: simulation is inside an adjoint run");
        if (s->mon) return berr(b, FDTD_ERR_UNSUPPORTED, "batch: point monitors are recorded by fdtd_step only");
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
    const char *conc = std::getenv("FDTD_CONCURRENT");
    if (!conc || std::atoi(conc) != 0) {
        if (hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_line, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_early, hipEventDisableTiming) != hipSuccess)
            return fail(berr(b, FDTD_ERR_DEVICE, "second stream or events could not be created"));
        // one side stream by default: the batch's launches already fill the device
        // (32-rod sweep: 51.6-52.0 k Mcell-updates/s with one, 51.0-51.2 k with three)
        const char *ss = std::getenv("FDTD_SIDE_STREAMS");
        if (ss && std::atoi(ss) == 3 &&
            (hipStreamCreateWithFlags(&b->stream3, hipStreamNonBlocking) != hipSuccess ||
             hipStreamCreateWithFlags(&b->stream4, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreateWithFlags(&b->ev_join3, hipEventDisableTiming) != hipSuccess ||
             hipEventCreateWithFlags(&b->ev_join4, hipEventDisableTiming) != hipSuccess))
            return fail(berr(b, FDTD_ERR_DEVICE, "side streams or events could not be created"));
        const char *ee = std::getenv("FDTD_BATCH_EARLY");
        b->early = ee && std::atoi(ee) != 0;
    }
    int rc;
    for (int par = 0; par < 2; par++) {
        if ((rc = balloc(b, &b->fused[par], n)) || (rc = balloc(b, &b->line[par], n)) || (rc = balloc(b, &b->disp[par], n)) ||
            (rc = balloc(b, &b->samp[par], n)))
            return fail(rc);
        for (int v = 0; v < 3; v++)
            if ((rc = balloc(b, &b->rows[par][v], n))) return fail(rc);
    }
    if ((rc = balloc(b, &b->dyn, 2))) return fail(rc);
    b->rowpath.assign(n, 0);
    b->ta.assign(n, 8);
    b->tb.assign(n, 32);
    for (int i = 0; i < n; i++) {
        fdtd_sim *s = b->sims[i];
        b->rowpath[i] = s->use_rows && !(flags & FDTD_BATCH_GENERAL) ? 1 : 0;
        if (b->rowpath[i]) { b->ta[i] = s->rows; b->tb[i] = s->cchunk; }
        if (!(flags & FDTD_BATCH_NO_TUNE))
            if ((rc = autotune(b, s, b->rowpath[i] != 0, &b->ta[i], &b->tb[i]))) return fail(rc);
        long nb[4] = {0, 0, 0, 0};
        if (b->rowpath[i]) {
            for (int v = 0; v < 3; v++) nb[v] = rows_blocks(rows_item(s, v, b->ta[i], b->tb[i]));
            if (nb[0] <= 0) return fail(berr(b, FDTD_ERR_UNSUPPORTED, "batch: row-blocked tile does not fit"));
        } else {
            nb[0] = fused_batch_blocks(s->g, b->ta[i], b->tb[i], nullptr, nullptr);
            if (nb[0] <= 0) return fail(berr(b, FDTD_ERR_UNSUPPORTED, "batch: general tile does not fit"));
        }
        nb[3] = std::max((s->nbox + 255) / 256, (s->fd.mE + 255) / 256);
        for (long v : nb)
            if (v >= (1L << kBatchLocalBits)) return fail(berr(b, FDTD_ERR_UNSUPPORTED, "batch: grid too large for the block map"));
    }
    *out = b;
    return FDTD_OK;
}

int fdtd_batch_run(fdtd_batch *b, int nsteps) {
    if (!b || nsteps < 0) return FDTD_ERR_INVALID;
    if (int rc = check_sims(b)) return rc;
    for (fdtd_sim *s : b->sims)
        if (int rc = check_finite(s)) return berr(b, rc, s->err);
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
    b->early_ready = false;
    int done = 0;
    while (done < nsteps) {
        if (use_graph && s0->t % P == 0 && nsteps - done >= P && per_period <= K) {
            if (s0->slot + per_period > K)
                if (int rc = contract_all(b)) return rc;
            if (!b->graph_valid)
                if (int rc = capture(b)) return rc;
            BCHK(b, hipGraphLaunch(b->exec, b->stream));
            b->early_ready = false;
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

int fdtd_batch_tile(const fdtd_batch *b, int i, int *path, int *a, int *c) {
    if (!b || i < 0 || i >= (int)b->sims.size()) return FDTD_ERR_INVALID;
    if (path) *path = b->rowpath[i];
    if (a) *a = b->ta[i];
    if (c) *c = b->tb[i];
    return FDTD_OK;
}

const char *fdtd_batch_last_error(const fdtd_batch *b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int fdtd_batch_destroy(fdtd_batch *b) {
    if (!b) return FDTD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->stream2) (void)hipStreamSynchronize(b->stream2);
    for (hipStream_t q : {b->stream3, b->stream4})
        if (q) (void)hipStreamSynchronize(q);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->ev_line) (void)hipEventDestroy(b->ev_line);
    if (b->ev_join3) (void)hipEventDestroy(b->ev_join3);
    if (b->ev_join4) (void)hipEventDestroy(b->ev_join4);
    if (b->ev_early) (void)hipEventDestroy(b->ev_early);
    if (b->stream2) (void)hipStreamDestroy(b->stream2);
    if (b->stream3) (void)hipStreamDestroy(b->stream3);
    if (b->stream4) (void)hipStreamDestroy(b->stream4);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return FDTD_OK;
}
