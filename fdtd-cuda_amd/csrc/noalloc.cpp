// Face-flux spectra without any device allocation (fdtd_flux_noalloc in
// include/fdtd.h).  fdtd_flux reduces the accumulated monitor DFTs into 12 x F
// face sums held in a buffer the handle allocates; this variant reduces them
// into `scratch`, a file-scope device array of fixed size, so a caller that
// must not allocate device memory (code under HIP graph capture set-up, a
// serving loop, a harness that forbids hipMalloc) can read spectra at any
// time.  The contraction of pending samples and the reduction kernel are the
// ones fdtd_flux launches, so both entry points return the same bits.
//
// Every call on one device shares that device's copy of `scratch` and leaves
// its last contents there: two calls that overlap on one device (two host
// threads, two handles on different streams) race, and must be serialised by
// the caller.  The reduction writes every entry it reads back, so nothing in
// the array has to start at zero.
#include <cmath>
#include <mutex>

#include "sim.h"

using namespace fdtd_host;

namespace {

constexpr int kMaxDevices = 64;

// 12 face sums per frequency for up to FDTD_NOALLOC_MAX_FREQ frequencies
// (24 KiB on every device the library runs on, held until the process ends)
static __device__ __align__(256) double scratch[12 * FDTD_NOALLOC_MAX_FREQ];

std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};   // device address of each device's copy

// The host-side name `scratch` is not a device address: look the address up
// once per device, with that device current, and use only that pointer.
int scratch_on(fdtd_sim *s, double **out) {
    if (s->device < 0 || s->device >= kMaxDevices)
        return set_err(s, FDTD_ERR_DEVICE, "fdtd_flux_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[s->device]) {
        int cur = 0;
        HIPCHK(s, hipGetDevice(&cur));
        if (cur != s->device) HIPCHK(s, hipSetDevice(s->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != s->device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return set_err(s, FDTD_ERR_DEVICE, std::string("fdtd_flux_noalloc: hipGetSymbolAddress failed: ") +
                                                   hipGetErrorString(e));
        g_ptr[s->device] = static_cast<double *>(p);
    }
    *out = g_ptr[s->device];
    return FDTD_OK;
}

}  // namespace

int fdtd_flux_noalloc(fdtd_sim *s, double *sca, double *abs_, double *ext, double *raw) {
    if (!s || !raw) return FDTD_ERR_INVALID;
    const int F = s->S.P.NUM_freq;
    if (F > FDTD_NOALLOC_MAX_FREQ)
        return set_err(s, FDTD_ERR_INVALID, "fdtd_flux_noalloc: NUM_freq " + std::to_string(F) +
                                                " exceeds FDTD_NOALLOC_MAX_FREQ (" +
                                                std::to_string(FDTD_NOALLOC_MAX_FREQ) + ")");
    double *q = nullptr;
    if (int rc = scratch_on(s, &q)) return rc;
    if (int rc = contract(s)) return rc;
    HIPCHK(s, launch_flux_reduce(s->acc, s->fd, s->mtot, s->fd.mE_pad, F, s->S.P.dx * s->S.P.dy, q, s->stream));
    // raw is the caller's: the face sums land there and are combined in place
    HIPCHK(s, hipMemcpyAsync(raw, q, (size_t)12 * F * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (int rc = check_finite(s)) return rc;
    for (long n = 0; n < 12L * F; n++)
        if (!std::isfinite(raw[n])) return set_err(s, FDTD_ERR_NONFINITE, "non-finite face flux (the monitor spectra diverged)");
    flux_combine(s, raw, sca, abs_, ext);
    return FDTD_OK;
}
