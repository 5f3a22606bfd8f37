// Poynting vector and flux through planes on the device (src/flux.jl:15-66):
// fdfd_poynting(_d) and fdfd_flux(_d).  The grid is passed explicitly, so
// post-processing never touches the handle's grid or factors; the handle
// only supplies the device and stream.
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "runtime.h"

using namespace fdfd;

namespace {

// Scratch of fdfd_flux_d_noalloc: the plane table (int4 per plane) followed
// by the per-plane sums (double per plane), for at most FLUX_NOALLOC_MAX
// planes (96 KiB).  Each GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` names a host-side stand-in, so the
// device address of each GPU's copy is looked up once per device
// (hipGetSymbolAddress with that device current, from fdfd_create) and only
// that address is used.
constexpr int FLUX_NOALLOC_MAX = 4096;
constexpr size_t FLUX_SCRATCH_BYTES = (size_t)FLUX_NOALLOC_MAX * (sizeof(int4) + sizeof(double));
static __device__ __align__(256) unsigned char scratch[FLUX_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};
static_assert(sizeof(fdfd_flux_plane) == sizeof(int4), "a plane is copied as one int4");

int check_grid(Handle* h, int pol, int Nx, int Ny) {
    if (pol != FDFD_TM && pol != FDFD_TE) return fail(h, FDFD_ERR_INVALID, "pol must be TM or TE");
    if (Nx < 1 || Ny < 1) return fail(h, FDFD_ERR_INVALID, "Nx and Ny must be >= 1");
    return FDFD_OK;
}

int check_planes(Handle* h, int Nx, int Ny, int nplanes, const fdfd_flux_plane* planes) {
    if (nplanes < 1 || !planes) return fail(h, FDFD_ERR_INVALID, "nplanes >= 1 and planes required");
    for (int p = 0; p < nplanes; ++p) {
        const fdfd_flux_plane& P = planes[p];
        const std::string tag = "flux plane " + std::to_string(p) + ": ";
        if (P.normal != 1 && P.normal != 2) return fail(h, FDFD_ERR_INVALID, tag + "normal must be 1 (x) or 2 (y)");
        const int across = P.normal == 1 ? Nx : Ny, along = P.normal == 1 ? Ny : Nx;
        // the plane's line is averaged with the next one (index + 1)
        if (P.index < 0 || P.index + 1 >= across)
            return fail(h, FDFD_ERR_INVALID, tag + "index must satisfy 0 <= index and index + 1 < the normal extent");
        if (P.lo < 0 || P.hi >= along || P.lo > P.hi)
            return fail(h, FDFD_ERR_INVALID, tag + "need 0 <= lo <= hi < the extent along the plane");
    }
    return FDFD_OK;
}

int run_flux(Handle* h, int pol, int Nx, int Ny, double dx, double dy, const double2* d_f, int nplanes,
             const fdfd_flux_plane* planes, double* flux) {
    std::vector<int4> hp(nplanes);
    for (int p = 0; p < nplanes; ++p) hp[p] = make_int4(planes[p].normal, planes[p].index, planes[p].lo, planes[p].hi);
    char* buf = nullptr;
    const size_t bytes = nplanes * (sizeof(int4) + sizeof(double));
    if (hipMalloc(&buf, bytes) != hipSuccess) return fail(h, FDFD_ERR_OOM, "flux workspace allocation failed");
    int4* d_p = (int4*)buf;
    double* d_o = (double*)(d_p + nplanes);
    int rc = stream_copy(h, d_p, hp.data(), nplanes * sizeof(int4), hipMemcpyHostToDevice);
    if (!rc) {
        launch_flux(h->stream, d_o, d_f, d_p, nplanes, Nx, Ny, pol, dx, dy);
        rc = stream_copy(h, flux, d_o, nplanes * sizeof(double), hipMemcpyDeviceToHost);
    }
    (void)hipFree(buf);
    return rc;
}

}  // namespace

namespace fdfd {
// Device address of this device's flux scratch (called by fdfd_create with
// the device current; once per device).
int rt_flux_scratch_setup(Handle* h) {
    if (h->device < 0 || h->device >= MAX_DEVICES)
        return fail(h, FDFD_ERR_INVALID, "device ordinal beyond the flux scratch table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[h->device]) return FDFD_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(h, FDFD_ERR_HIP, std::string("flux scratch lookup: ") + hipGetErrorString(e));
    scratch_ptr[h->device] = p;
    return FDFD_OK;
}
}  // namespace fdfd

extern "C" int fdfd_poynting_d(fdfd_handle* hh, int pol, int Nx, int Ny, const void* d_fields,
                               void* d_S) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h || !d_fields || !d_S) return FDFD_ERR_INVALID;
    int rc = check_grid(h, pol, Nx, Ny);
    if (rc) return rc;
    (void)hipSetDevice(h->device);
    launch_poynting(h->stream, (double*)d_S, (const double2*)d_fields, Nx, Ny, pol);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FDFD_OK;
}

extern "C" int fdfd_poynting(fdfd_handle* hh, int pol, int Nx, int Ny, const fdfd_complex* fields,
                             double* S) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h || !fields || !S) return FDFD_ERR_INVALID;
    int rc = check_grid(h, pol, Nx, Ny);
    if (rc) return rc;
    (void)hipSetDevice(h->device);
    const long N = (long)Nx * Ny;
    char* buf = nullptr;
    if (hipMalloc(&buf, N * (3 * sizeof(double2) + 2 * sizeof(double))) != hipSuccess)
        return fail(h, FDFD_ERR_OOM, "Poynting workspace allocation failed");
    double2* d_f = (double2*)buf;
    double* d_S = (double*)(d_f + 3 * N);
    rc = stream_copy(h, d_f, fields, 3 * N * sizeof(double2), hipMemcpyHostToDevice);
    if (!rc) {
        launch_poynting(h->stream, d_S, d_f, Nx, Ny, pol);
        rc = stream_copy(h, S, d_S, 2 * N * sizeof(double), hipMemcpyDeviceToHost);
    }
    (void)hipFree(buf);
    return rc;
}

extern "C" int fdfd_flux_d(fdfd_handle* hh, int pol, int Nx, int Ny, double dx, double dy,
                           const void* d_fields, int nplanes, const fdfd_flux_plane* planes,
                           double* flux) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h || !d_fields || !flux) return FDFD_ERR_INVALID;
    int rc = check_grid(h, pol, Nx, Ny);
    if (!rc) rc = check_planes(h, Nx, Ny, nplanes, planes);
    if (rc) return rc;
    (void)hipSetDevice(h->device);
    return run_flux(h, pol, Nx, Ny, dx, dy, (const double2*)d_fields, nplanes, planes, flux);
}

extern "C" int fdfd_flux(fdfd_handle* hh, int pol, int Nx, int Ny, double dx, double dy,
                         const fdfd_complex* fields, int nplanes, const fdfd_flux_plane* planes,
                         double* flux) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h || !fields || !flux) return FDFD_ERR_INVALID;
    int rc = check_grid(h, pol, Nx, Ny);
    if (!rc) rc = check_planes(h, Nx, Ny, nplanes, planes);
    if (rc) return rc;
    (void)hipSetDevice(h->device);
    const long N = (long)Nx * Ny;
    double2* d_f = nullptr;
    if (hipMalloc(&d_f, 3 * N * sizeof(double2)) != hipSuccess)
        return fail(h, FDFD_ERR_OOM, "flux field buffer allocation failed");
    rc = stream_copy(h, d_f, fields, 3 * N * sizeof(double2), hipMemcpyHostToDevice);
    if (!rc) rc = run_flux(h, pol, Nx, Ny, dx, dy, d_f, nplanes, planes, flux);
    (void)hipFree(d_f);
    return rc;
}

// fdfd_flux_d without any device allocation: the plane table and the sums go
// through this device's static scratch instead of a per-call workspace.  At
// most FLUX_NOALLOC_MAX planes.  Calls that overlap on one GPU (several
// handles or threads) share the scratch and race; each call finishes (stream
// synchronized) before it returns.  Every sum the kernel reads back is
// written by the same call, so nothing needs clearing between calls.
extern "C" int fdfd_flux_d_noalloc(fdfd_handle* hh, int pol, int Nx, int Ny, double dx, double dy,
                                   const void* d_fields, int nplanes,
                                   const fdfd_flux_plane* planes, double* flux) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h || !d_fields || !flux) return FDFD_ERR_INVALID;
    int rc = check_grid(h, pol, Nx, Ny);
    if (!rc) rc = check_planes(h, Nx, Ny, nplanes, planes);
    if (rc) return rc;
    if (nplanes > FLUX_NOALLOC_MAX)
        return fail(h, FDFD_ERR_INVALID, "fdfd_flux_d_noalloc takes at most " +
                                             std::to_string(FLUX_NOALLOC_MAX) + " planes");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(scratch_mu);
        base = (h->device >= 0 && h->device < MAX_DEVICES) ? scratch_ptr[h->device] : nullptr;
    }
    if (!base) return fail(h, FDFD_ERR_STATE, "flux scratch of this device was not set up");
    (void)hipSetDevice(h->device);
    int4* d_p = (int4*)base;
    double* d_o = (double*)((char*)base + (size_t)FLUX_NOALLOC_MAX * sizeof(int4));
    HIPCHK(h, hipMemcpyAsync(d_p, planes, sizeof(int4) * nplanes, hipMemcpyHostToDevice, h->stream));
    launch_flux(h->stream, d_o, (const double2*)d_fields, d_p, nplanes, Nx, Ny, pol, dx, dy);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(flux, d_o, sizeof(double) * nplanes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FDFD_OK;
}
