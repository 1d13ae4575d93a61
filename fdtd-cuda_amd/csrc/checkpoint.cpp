// Checkpoint / resume of a whole simulation (fdtd_checkpoint_* in
// include/fdtd.h).  The reference has no restart: its state is ~700 process
// globals (global_var.h:32-815) that live for one run of main().
//
// An image is a header followed by sections, each a 64-bit byte count and
// its bytes:
//   materials  mat_matrix, X, Y, Z masks (4 x NX NY NZ bytes)
//   design     block origin, extent and densities (may be empty)
//   fields     E and H at the current time level, padded device layout
//   psi        psi_H (current set) and psi_E slabs
//   dispersive box state (8 arrays per E component; hydro: n~ pair per component)
//   line       e_inc, h_inc, the two Mur values
//   dft        the first `slot` rows of the sample ring and the FP64 accumulators
//   monitor    point-monitor samples recorded so far (may be empty)
// Raw device buffers are copied as they are, so a restored handle continues
// with the same bits as the one that saved (time-level parities need not
// match: each handle reads its own current level).
#include <cstring>

#include "sim.h"

using namespace fdtd_host;

namespace {

constexpr char kMagic[8] = {'F', 'D', 'T', 'D', 'C', 'K', 'P', '1'};

struct Header {
    char magic[8];
    uint32_t version, header_bytes;
    int32_t N[3], Nc[3], F, K, stride, hydro, nsec;
    int64_t fsz, mtot, nbox;
    double dt;
    int32_t t, samples, slot, mon_cap, mon_t0, des_lo[3], des_n[3];
    int64_t mon_p;
    int32_t slot_t[256];
};

// One section: a list of (pointer, bytes) pieces on the device or the host.
struct Piece {
    void *p;
    size_t n;
    bool dev;
};

struct Layout {
    std::vector<std::vector<Piece>> sec;
    size_t bytes = 0;
};

Layout layout(fdtd_sim *s) {
    Layout L;
    const size_t N = (size_t)s->g.NX * s->g.NY * s->g.NZ;
    std::vector<Piece> mats;
    for (int c = 0; c < 4; c++) mats.push_back({s->S.mat[c].data(), N, false});
    L.sec.push_back(mats);
    L.sec.push_back({{s->des_rho.data(), s->des_rho.size() * sizeof(float), false}});
    std::vector<Piece> f;
    for (int c = 0; c < 3; c++) {
        f.push_back({s->E[s->p][c] - s->fbase, s->fsz * sizeof(float), true});
        f.push_back({s->H[s->p][c] - s->fbase, s->fsz * sizeof(float), true});
    }
    L.sec.push_back(f);
    std::vector<Piece> ps;
    for (int a = 0; a < 6; a++) ps.push_back({s->psiH[s->pp][a], s->hsz[a] * sizeof(float), true});
    for (int a = 0; a < 6; a++) ps.push_back({s->psiE[a], s->esz[a] * sizeof(float), true});
    L.sec.push_back(ps);
    std::vector<Piece> db;
    if (s->nbox > 0)
        for (int c = 0; c < 3; c++) {
            for (int q = 0; q < 8; q++) db.push_back({s->disp_st[c][q], s->nbox * sizeof(float), true});
            if (s->S.hydro) {
                db.push_back({s->h_nd[c], s->nbox * sizeof(float), true});
                db.push_back({s->h_ndp[c], s->nbox * sizeof(float), true});
            }
        }
    L.sec.push_back(db);
    L.sec.push_back({{s->line_e[s->le], s->S.L * sizeof(double), true},
                     {s->line_h, s->S.L * sizeof(double), true},
                     {s->mur, 2 * sizeof(double), true}});
    L.sec.push_back({{s->X, (size_t)s->slot * s->mtot * sizeof(float), true},
                     {s->acc, (size_t)2 * s->S.P.NUM_freq * s->mtot * sizeof(double), true}});
    const int nmon = s->mon ? std::min(std::max(s->t - s->mon_t0, 0), s->mon_cap) : 0;
    L.sec.push_back({{s->mon, (size_t)3 * nmon * sizeof(float), true}});
    L.bytes = sizeof(Header);
    for (auto &sec : L.sec) {
        L.bytes += sizeof(uint64_t);
        for (auto &p : sec) L.bytes += p.n;
    }
    return L;
}

Header make_header(const fdtd_sim *s, int nsec) {
    Header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.header_bytes = sizeof(Header);
    for (int a = 0; a < 3; a++) { h.N[a] = s->S.N[a]; h.Nc[a] = s->S.cp[a].Nc; }
    h.F = s->S.P.NUM_freq;
    h.K = s->K;
    h.stride = s->S.P.dft_stride;
    h.hydro = s->S.hydro ? 1 : 0;
    h.nsec = nsec;
    h.fsz = (int64_t)s->fsz;
    h.mtot = s->mtot;
    h.nbox = s->nbox;
    h.dt = s->S.dt;
    h.t = s->t;
    h.samples = s->samples;
    h.slot = s->slot;
    h.mon_cap = s->mon ? s->mon_cap : 0;
    h.mon_t0 = s->mon_t0;
    h.mon_p = s->mon_p;
    for (int a = 0; a < 3; a++) { h.des_lo[a] = s->des_lo[a]; h.des_n[a] = s->des_n[a]; }
    for (int q = 0; q < s->K && q < 256; q++) h.slot_t[q] = s->slot_t[q];
    return h;
}

}  // namespace

int fdtd_checkpoint_size(fdtd_sim *s, size_t *bytes) {
    if (!s || !bytes) return FDTD_ERR_INVALID;
    *bytes = layout(s).bytes;
    return FDTD_OK;
}

int fdtd_checkpoint_save(fdtd_sim *s, void *buf, size_t bytes) {
    if (!s || !buf) return FDTD_ERR_INVALID;
    const Layout L = layout(s);
    if (bytes < L.bytes) return set_err(s, FDTD_ERR_INVALID, "checkpoint buffer smaller than fdtd_checkpoint_size");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (int rc = check_finite(s)) return rc;
    char *o = (char *)buf;
    const Header h = make_header(s, (int)L.sec.size());
    std::memcpy(o, &h, sizeof(h));
    o += sizeof(h);
    for (auto &sec : L.sec) {
        uint64_t n = 0;
        for (auto &p : sec) n += p.n;
        std::memcpy(o, &n, sizeof(n));
        o += sizeof(n);
        for (auto &p : sec) {
            if (p.n && p.dev) HIPCHK(s, hipMemcpy(o, p.p, p.n, hipMemcpyDeviceToHost));
            else if (p.n) std::memcpy(o, p.p, p.n);
            o += p.n;
        }
    }
    return FDTD_OK;
}

int fdtd_checkpoint_load(fdtd_sim *s, const void *buf, size_t bytes) {
    if (!s || !buf) return FDTD_ERR_INVALID;
    if (bytes < sizeof(Header)) return set_err(s, FDTD_ERR_INVALID, "checkpoint image too short");
    Header h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 1 || h.header_bytes != sizeof(Header))
        return set_err(s, FDTD_ERR_INVALID, "not a checkpoint image of this library version");
    const Header mine = make_header(s, 0);
    bool same = h.F == mine.F && h.K == mine.K && h.stride == mine.stride && h.hydro == mine.hydro &&
                h.fsz == mine.fsz && h.mtot == mine.mtot && h.dt == mine.dt;
    for (int a = 0; a < 3; a++) same = same && h.N[a] == mine.N[a] && h.Nc[a] == mine.Nc[a];
    if (!same) return set_err(s, FDTD_ERR_INVALID, "checkpoint was saved by a simulation with other parameters");
    const char *in = (const char *)buf + sizeof(Header);
    const char *end = (const char *)buf + bytes;
    auto section = [&](uint64_t &n) -> const char * {
        if (end - in < (long)sizeof(uint64_t)) return nullptr;
        std::memcpy(&n, in, sizeof(n));
        in += sizeof(n);
        if ((uint64_t)(end - in) < n) return nullptr;
        const char *p = in;
        in += n;
        return p;
    };
    HIPCHK(s, hipStreamSynchronize(s->stream));
    // materials and design first: they decide the dispersive box and its size
    uint64_t n = 0;
    const char *mats = section(n);
    const size_t N = (size_t)s->g.NX * s->g.NY * s->g.NZ;
    if (!mats || n != 4 * N) return set_err(s, FDTD_ERR_INVALID, "checkpoint: bad materials section");
    const char *des = section(n);
    if (!des || n % sizeof(float)) return set_err(s, FDTD_ERR_INVALID, "checkpoint: bad design section");
    const std::vector<float> rho((const float *)des, (const float *)(des + n));
    bool remap = rho != s->des_rho;
    for (int c = 1; c < 4; c++) remap = remap || std::memcmp(mats + c * N, s->S.mat[c].data(), N) != 0;
    if (remap) {
        for (int a = 0; a < 3; a++) { s->des_lo[a] = h.des_lo[a]; s->des_n[a] = h.des_n[a]; }
        s->des_rho = rho;
        if (int rc = fdtd_set_materials(s, (const uint8_t *)mats + N, (const uint8_t *)mats + 2 * N,
                                        (const uint8_t *)mats + 3 * N))
            return rc;
    }
    std::memcpy(s->S.mat[0].data(), mats, N);
    if (s->nbox != h.nbox) return set_err(s, FDTD_ERR_INVALID, "checkpoint: dispersive box does not match");
    // point monitor: allocate before the layout is taken
    if (h.mon_cap > 0) {
        const int r[3] = {0, 0, 0};
        if (int rc = fdtd_set_point_monitor(s, r, h.mon_cap)) return rc;
        s->mon_p = h.mon_p;
    } else {
        s->mon = nullptr;
        s->mon_cap = 0;
    }
    s->t = h.t;
    s->samples = h.samples;
    s->slot = h.slot;
    s->mon_t0 = h.mon_t0;
    for (int q = 0; q < s->K && q < 256; q++) s->slot_t[q] = h.slot_t[q];
    const Layout L = layout(s);
    for (size_t q = 2; q < L.sec.size(); q++) {
        const char *p = section(n);
        uint64_t want = 0;
        for (auto &pc : L.sec[q]) want += pc.n;
        if (!p || n != want) return set_err(s, FDTD_ERR_INVALID, "checkpoint: section sizes do not match");
        for (auto &pc : L.sec[q]) {
            if (pc.n) HIPCHK(s, hipMemcpy(pc.p, p, pc.n, hipMemcpyHostToDevice));
            p += pc.n;
        }
    }
    *(volatile unsigned *)s->status_h = 0u;
    return FDTD_OK;
}
