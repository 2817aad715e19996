// capi.hip - the extern "C" surface of libostr_hip.so (include/ostr_hip.h):
// handle lifetime, argument validation, error reporting and dispatch to the
// kernel launchers. No process-global geometry (the reference keeps it in
// __constant__ symbols, Tomo_OSTR.cu:81-115) and no exit(): every failure is
// an error code plus a message.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ostr_internal.h"

#define OSTR_API extern "C" __attribute__((visibility("default")))

struct ostr_handle {
    ostr_geom geom;
    ostr::DevGeom dg;
    int n_views = 0;
    int device = 0;
    std::vector<float> angles;
    std::vector<int32_t> index;
    std::vector<ostr::ViewTab> views_host;  // current slot order
    std::vector<ostr::ViewTab> views_nat;   // identity order
    ostr::ViewTab* views_dev = nullptr;
    ostr::ViewTab* views_nat_dev = nullptr;
    unsigned long long* counter = nullptr;
    int lds_limit = 0;        // LDS bytes per work-group on this device
    void* work = nullptr;     // grow-on-demand device work buffer
    size_t work_bytes = 0;
    double* dwork = nullptr;  // FP64 partial sums of the reductions (+2 results)
    size_t dwork_elems = 0;
    // Set once a call on this handle ran on a capturing stream: a graph then
    // holds work/dwork addresses and may be replayed after any later call, so
    // an outgrown buffer is kept in `retired` (freed by ostr_destroy).
    bool captured = false;
    std::vector<void*> retired;
    std::string err;
};

namespace {

thread_local std::string g_err = "no error";

int fail(ostr_handle* h, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h)
        h->err = buf;
    else
        g_err = buf;
    return code;
}

int hip_fail(ostr_handle* h, hipError_t e, const char* where) {
    return fail(h, OSTR_E_HIP, "%s: HIP error %d (%s)", where, (int)e, hipGetErrorString(e));
}

#define CHECK_HANDLE(h)                                                    \
    do {                                                                   \
        if (!(h)) return fail(nullptr, OSTR_E_ARG, "%s: NULL handle", __func__); \
    } while (0)

int use_device(ostr_handle* h) {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    return OSTR_OK;
}

int check_range(ostr_handle* h, int a0, int a1, const char* fn) {
    if (a0 < 0 || a1 > h->n_views || a0 > a1)
        return fail(h, OSTR_E_RANGE, "%s: view range [%d,%d) outside [0,%d)", fn, a0, a1,
                    h->n_views);
    return OSTR_OK;
}

int check_dtype(ostr_handle* h, int dt, const char* fn) {
    if (dt != OSTR_F32 && dt != OSTR_BF16 && dt != OSTR_F16)
        return fail(h, OSTR_E_UNSUPPORTED, "%s: unknown dtype %d", fn, dt);
    return OSTR_OK;
}

// The backprojectors read a launch's sinogram slots (at most MAX_SLOTS per
// launch) through a buffer resource with 32-bit byte offsets: one launch's
// span must stay below 2^32 bytes.
int check_sino_span(ostr_handle* h, int64_t n_slots, int64_t elem_bytes, const char* fn) {
    const int64_t slots = n_slots < ostr::MAX_SLOTS ? n_slots : ostr::MAX_SLOTS;
    const int64_t bytes = slots * h->geom.nbx * h->geom.nby * elem_bytes;
    if (bytes > (int64_t)0xFFFFFFF0)
        return fail(h, OSTR_E_RANGE,
                    "%s: %lld views of %dx%d bins span %lld bytes per launch; at most 2^32 - 16 "
                    "(use fewer views per call)",
                    fn, (long long)slots, h->geom.nbx, h->geom.nby, (long long)bytes);
    return OSTR_OK;
}

// The bins whose rays may meet the volume: the central projection of the
// volume's 8 corners from the source onto the detector plane bounds the
// shadow of every part of the box on the source's side of the plane; its
// bin range, widened by 2 bins, clamped to the detector. All bins when a
// corner is not below the source.
void view_bbox(const ostr_geom& g, ostr::ViewTab& v) {
    v.i0 = 0;
    v.i1 = g.nbx - 1;
    v.j0 = 0;
    v.j1 = g.nby - 1;
    const double sx = v.x0, sy = g.source_y, sz = v.z0, ez = -(double)g.detector_r;
    const double xs[2] = {g.x_p0, (double)g.x_p0 + (double)g.nx * g.dx};
    const double ys[2] = {g.y_p0, (double)g.y_p0 + (double)g.ny * g.dy};
    const double zs[2] = {g.z_p0, (double)g.z_p0 + (double)g.nz * g.dz};
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int c = 0; c < 8; ++c) {
        const double cx = xs[c & 1], cy = ys[(c >> 1) & 1], cz = zs[c >> 2];
        if (!(cz < sz) || !(ez < sz)) return;
        const double t = (ez - sz) / (cz - sz);
        const double px = sx + t * (cx - sx), py = sy + t * (cy - sy);
        if (!std::isfinite(px) || !std::isfinite(py)) return;
        xmin = std::min(xmin, px);
        xmax = std::max(xmax, px);
        ymin = std::min(ymin, py);
        ymax = std::max(ymax, py);
    }
    const double qi0 = std::floor((xmin - g.x_d0) / g.dbx - 0.5) - 2.0;
    const double qi1 = std::ceil((xmax - g.x_d0) / g.dbx - 0.5) + 2.0;
    const double qj0 = std::floor((ymin - g.y_d0) / g.dby - 0.5) - 2.0;
    const double qj1 = std::ceil((ymax - g.y_d0) / g.dby - 0.5) + 2.0;
    v.i0 = (int)std::max(0.0, std::min(qi0, (double)g.nbx));
    v.i1 = (int)std::max(-1.0, std::min(qi1, (double)g.nbx - 1));
    v.j0 = (int)std::max(0.0, std::min(qj0, (double)g.nby));
    v.j1 = (int)std::max(-1.0, std::min(qj1, (double)g.nby - 1));
}

ostr::ViewTab view_of(const ostr_geom& g, float theta) {
    ostr::ViewTab v;
    v.s = (float)std::sin((double)theta);
    v.c = (float)std::cos((double)theta);
    v.x0 = g.source_r * v.s;
    v.z0 = g.source_r * v.c;
    view_bbox(g, v);
    return v;
}

int upload_views(ostr_handle* h) {
    h->views_host.resize(h->n_views);
    for (int a = 0; a < h->n_views; ++a)
        h->views_host[a] = view_of(h->geom, h->angles[h->index[a]]);
    hipError_t e = hipMemcpy(h->views_dev, h->views_host.data(),
                             sizeof(ostr::ViewTab) * h->n_views, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(h, e, "upload view table");
    return OSTR_OK;
}

// True when `st` is recording a graph. The legacy null stream cannot be
// captured and is not queried (a query of it during another stream's global
// capture would invalidate that capture).
bool capturing(hipStream_t st) {
    if (!st) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Grow one of the handle's device work buffers to at least `bytes`
// (stream-ordered: the old buffer may still be read by queued work on `st`).
// After a capture the old buffer is retired, not freed: a recorded graph may
// still write to it on replay. Growing during capture is an error (neither a
// synchronize nor an allocation may be recorded).
int grow(ostr_handle* h, void** buf, size_t* have, size_t bytes, hipStream_t st,
         const char* fn) {
    const bool cap = capturing(st);
    h->captured = h->captured || cap;
    if (*have >= bytes) return OSTR_OK;
    if (cap)
        return fail(h, OSTR_E_ARG,
                    "%s: work buffer must grow to %zu bytes during stream capture; make the "
                    "same call once before capturing",
                    fn, bytes);
    if (*buf) {
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(h, e, fn);
        if (h->captured)
            h->retired.push_back(*buf);
        else
            (void)hipFree(*buf);
        *buf = nullptr;
        *have = 0;
    }
    hipError_t e = hipMalloc(buf, bytes);
    if (e != hipSuccess) return hip_fail(h, e, fn);
    *have = bytes;
    return OSTR_OK;
}

int ensure_work(ostr_handle* h, size_t bytes, hipStream_t st, const char* fn) {
    return grow(h, &h->work, &h->work_bytes, bytes, st, fn);
}

int ensure_dwork(ostr_handle* h, size_t elems, hipStream_t st, const char* fn) {
    size_t have = h->dwork_elems * sizeof(double);
    void* buf = h->dwork;
    int rc = grow(h, &buf, &have, elems * sizeof(double), st, fn);
    h->dwork = (double*)buf;
    h->dwork_elems = have / sizeof(double);
    return rc;
}

int validate_geom(const ostr_geom* g) {
    if (!g) return fail(nullptr, OSTR_E_ARG, "ostr_create: NULL geometry");
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || g->nbx <= 0 || g->nby <= 0)
        return fail(nullptr, OSTR_E_ARG, "ostr_create: grid sizes must be positive");
    if (g->nx >= 32768 || g->ny >= 32768 || g->nz >= 32768)
        return fail(nullptr, OSTR_E_ARG, "ostr_create: volume dimensions must be < 32768");
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0) || !(g->dbx > 0) || !(g->dby > 0))
        return fail(nullptr, OSTR_E_ARG, "ostr_create: voxel and bin sizes must be positive");
    if ((int64_t)g->nbx * g->nby > (int64_t)1 << 30)
        return fail(nullptr, OSTR_E_ARG, "ostr_create: detector too large");
    // 32-bit byte offsets and buffer-resource sizes: < 2^30 voxels, with a
    // margin of two slices and two rows (the walk's out-of-volume indices
    // one or two planes past an edge must not wrap into the volume's range)
    const int64_t nxy = (int64_t)g->nx * g->ny;
    if (nxy * g->nz + 2 * nxy + 2 * (int64_t)g->nx + 2 > (int64_t)1 << 30)
        return fail(nullptr, OSTR_E_ARG,
                    "ostr_create: volume too large (nx ny nz + 2 nx ny + 2 nx + 2 must be "
                    "<= 2^30)");
    return OSTR_OK;
}

}  // namespace

namespace ostr {

DevGeom make_dev_geom(const ostr_geom& g) {
    DevGeom d;
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz; d.nbx = g.nbx; d.nby = g.nby;
    d.dx = g.dx; d.dy = g.dy; d.dz = g.dz;
    d.x_p0 = g.x_p0; d.y_p0 = g.y_p0; d.z_p0 = g.z_p0;
    d.dbx = g.dbx; d.dby = g.dby; d.x_d0 = g.x_d0; d.y_d0 = g.y_d0;
    d.det_r = g.detector_r; d.src_r = g.source_r; d.src_y = g.source_y;
    d.ext_x = (float)g.nx * g.dx;
    d.ext_y = (float)g.ny * g.dy;
    d.ext_z = (float)g.nz * g.dz;
    d.top_x = std::fma((float)g.nx, g.dx, g.x_p0);
    d.top_y = std::fma((float)g.ny, g.dy, g.y_p0);
    d.top_z = std::fma((float)g.nz, g.dz, g.z_p0);
    d.hp_x = g.dx + g.x_p0;
    d.hp_y = g.dy + g.y_p0;
    d.hp_z = g.dz + g.z_p0;
    d.last_x = std::fma((float)(g.nx - 1), g.dx, g.x_p0);
    d.last_y = std::fma((float)(g.ny - 1), g.dy, g.y_p0);
    d.last_z = std::fma((float)(g.nz - 1), g.dz, g.z_p0);
    return d;
}

}  // namespace ostr

// Scratch of the allocation-free objective (ostr_objective_noalloc): the FP64
// partial sums of the data and penalty terms plus the two results, at most
// NOALLOC_DOUBLES (8 MiB). Each device holds its own copy for the life of the
// process. Host code never uses this name as an address: ostr_create resolves
// the device's copy with hipGetSymbolAddress once per device (g_noalloc_ptr).
constexpr int64_t NOALLOC_DOUBLES = (int64_t)1 << 20;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's copy of `scratch` (resolved on first use per device,
// with that device current: called from ostr_create).
int noalloc_resolve(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return fail(nullptr, OSTR_E_UNSUPPORTED, "ostr_create: device %d >= %d", device,
                    NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return OSTR_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(nullptr, e, "ostr_create: scratch address");
    if (!p) return fail(nullptr, OSTR_E_HIP, "ostr_create: scratch address is NULL");
    g_noalloc_ptr[device] = (double*)p;
    return OSTR_OK;
}

}  // namespace

// ======================================================================
OSTR_API int ostr_create(const ostr_geom* geom, const float* angles_rad, const int32_t* index,
                         int32_t n_views, ostr_handle** out) {
    if (!out) return fail(nullptr, OSTR_E_ARG, "ostr_create: NULL output pointer");
    *out = nullptr;
    int rc = validate_geom(geom);
    if (rc) return rc;
    if (n_views <= 0 || !angles_rad)
        return fail(nullptr, OSTR_E_ARG, "ostr_create: need n_views > 0 and angles");
    for (int i = 0; i < n_views; ++i)
        if (!std::isfinite(angles_rad[i]))
            return fail(nullptr, OSTR_E_ARG, "ostr_create: angle %d is not finite", i);
    if (index)
        for (int i = 0; i < n_views; ++i)
            if (index[i] < 0 || index[i] >= n_views)
                return fail(nullptr, OSTR_E_ARG, "ostr_create: index[%d]=%d out of range", i,
                            index[i]);
    ostr_handle* h = new ostr_handle();
    h->geom = *geom;
    h->dg = ostr::make_dev_geom(*geom);
    h->n_views = n_views;
    h->angles.assign(angles_rad, angles_rad + n_views);
    h->index.resize(n_views);
    for (int i = 0; i < n_views; ++i) h->index[i] = index ? index[i] : i;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess && (rc = noalloc_resolve(h->device)) != OSTR_OK) {
        ostr_destroy(h);
        return rc;
    }
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&h->lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                  h->device);
    if (e == hipSuccess) e = hipMalloc(&h->views_dev, sizeof(ostr::ViewTab) * n_views);
    if (e == hipSuccess) e = hipMalloc(&h->views_nat_dev, sizeof(ostr::ViewTab) * n_views);
    if (e == hipSuccess) e = hipMalloc(&h->counter, sizeof(unsigned long long));
    if (e != hipSuccess) {
        hip_fail(nullptr, e, "ostr_create");
        ostr_destroy(h);
        return OSTR_E_HIP;
    }
    h->views_nat.resize(n_views);
    for (int a = 0; a < n_views; ++a) h->views_nat[a] = view_of(*geom, h->angles[a]);
    e = hipMemcpy(h->views_nat_dev, h->views_nat.data(), sizeof(ostr::ViewTab) * n_views,
                  hipMemcpyHostToDevice);
    if (e != hipSuccess || upload_views(h) != OSTR_OK) {
        if (e != hipSuccess) hip_fail(nullptr, e, "ostr_create");
        else g_err = h->err;
        ostr_destroy(h);
        return OSTR_E_HIP;
    }
    *out = h;
    return OSTR_OK;
}

OSTR_API void ostr_destroy(ostr_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->views_dev) (void)hipFree(h->views_dev);
    if (h->views_nat_dev) (void)hipFree(h->views_nat_dev);
    if (h->counter) (void)hipFree(h->counter);
    if (h->work) (void)hipFree(h->work);
    if (h->dwork) (void)hipFree(h->dwork);
    for (void* p : h->retired) (void)hipFree(p);
    delete h;
}

OSTR_API const char* ostr_last_error(const ostr_handle* h) {
    return h ? h->err.c_str() : g_err.c_str();
}

OSTR_API int ostr_set_index(ostr_handle* h, const int32_t* index) {
    CHECK_HANDLE(h);
    if (!index) return fail(h, OSTR_E_ARG, "ostr_set_index: NULL index");
    for (int i = 0; i < h->n_views; ++i)
        if (index[i] < 0 || index[i] >= h->n_views)
            return fail(h, OSTR_E_ARG, "ostr_set_index: index[%d]=%d out of range", i, index[i]);
    int rc = use_device(h);
    if (rc) return rc;
    // the view table may still be read by queued kernels on any stream
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(h, e, "ostr_set_index");
    h->index.assign(index, index + h->n_views);
    return upload_views(h);
}

OSTR_API int ostr_get_geom(const ostr_handle* h, ostr_geom* out) {
    if (!h || !out) return fail(nullptr, OSTR_E_ARG, "ostr_get_geom: NULL argument");
    *out = h->geom;
    return OSTR_OK;
}

OSTR_API int32_t ostr_n_views(const ostr_handle* h) { return h ? h->n_views : -1; }

// ------------------------------------------------------------ launchers --
OSTR_API int ostr_fproj(ostr_handle* h, const void* img, int32_t img_dtype, void* prj,
                        int32_t prj_dtype, int32_t a0, int32_t a1, int32_t cos_weight,
                        void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_fproj")) || (rc = check_dtype(h, img_dtype, "ostr_fproj")) ||
        (rc = check_dtype(h, prj_dtype, "ostr_fproj")))
        return rc;
    if (!img || !prj) return fail(h, OSTR_E_ARG, "ostr_fproj: NULL buffer");
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_project(h->dg, h->views_dev, a0, a1, 0, cos_weight, img,
                                        img_dtype, prj, prj_dtype, nullptr, nullptr, nullptr,
                                        (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_fproj");
}

OSTR_API int ostr_gamma(ostr_handle* h, float* gamma, int32_t a0, int32_t a1, int32_t mode,
                        void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_gamma"))) return rc;
    if (mode != OSTR_GAMMA_RAY_LENGTH && mode != OSTR_GAMMA_ROW_SUM)
        return fail(h, OSTR_E_ARG, "ostr_gamma: unknown mode %d", mode);
    if (!gamma) return fail(h, OSTR_E_ARG, "ostr_gamma: NULL buffer");
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_project(h->dg, h->views_dev, a0, a1,
                                        mode == OSTR_GAMMA_ROW_SUM ? 2 : 1, 0, nullptr,
                                        OSTR_F32, gamma, OSTR_F32, nullptr, nullptr, nullptr,
                                        (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_gamma");
}

OSTR_API int ostr_bproj_sbp(ostr_handle* h, float* img_accum, const void* prj,
                            int32_t prj_dtype, int32_t a0, int32_t a1, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_bproj_sbp")) ||
        (rc = check_dtype(h, prj_dtype, "ostr_bproj_sbp")))
        return rc;
    if (!img_accum || !prj) return fail(h, OSTR_E_ARG, "ostr_bproj_sbp: NULL buffer");
    if ((rc = check_sino_span(h, a1 - a0, ostr::elem_size(prj_dtype), "ostr_bproj_sbp")))
        return rc;
    if ((rc = use_device(h))) return rc;
    if (a0 == a1) return OSTR_OK;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = (int64_t)(a1 - a0) * h->geom.nbx * h->geom.nby;
    if ((rc = ensure_work(h, n * ostr::elem_size(prj_dtype), st, "ostr_bproj_sbp"))) return rc;
    // the kernel reads detector-row-major bins: transpose the reference layout first
    hipError_t e = ostr::launch_transpose_bins(prj, h->work, a1 - a0, h->geom.nbx,
                                               h->geom.nby, prj_dtype, st);
    if (e == hipSuccess)
        e = ostr::launch_sbp_rows(h->dg, h->views_host.data(), a0, a1, img_accum, h->work,
                                  prj_dtype, st);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_bproj_sbp");
}

// ---- OS-SART ---------------------------------------------------------------

OSTR_API int ostr_fproj_sart(ostr_handle* h, const float* img, float* prj, float* norm,
                             int32_t a0, int32_t a1, int32_t cos_weight, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_fproj_sart"))) return rc;
    if (!img || !prj || !norm) return fail(h, OSTR_E_ARG, "ostr_fproj_sart: NULL buffer");
    if (prj == norm) return fail(h, OSTR_E_ARG, "ostr_fproj_sart: prj and norm alias");
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_sart_project(h->dg, h->views_dev, a0, a1, cos_weight, img, prj,
                                             norm, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_fproj_sart");
}

OSTR_API int ostr_sart_prj_diff(ostr_handle* h, float* diff, const float* p, const float* est,
                                const float* norm, int64_t n, void* stream) {
    CHECK_HANDLE(h);
    if (!diff || !p || !est || !norm || n < 0)
        return fail(h, OSTR_E_ARG, "ostr_sart_prj_diff: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_sart_diff(diff, p, est, norm, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_sart_prj_diff");
}

namespace {
int bproj_sart(ostr_handle* h, float* img, const float* diff, const float* prior, int a0,
               int a1, float lambda, float beta, void* stream, const char* fn) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, fn))) return rc;
    if (!img || !diff) return fail(h, OSTR_E_ARG, "%s: NULL buffer", fn);
    if (a1 - a0 > ostr::MAX_SLOTS)
        return fail(h, OSTR_E_RANGE, "%s: at most %d views per call", fn, ostr::MAX_SLOTS);
    if ((rc = check_sino_span(h, a1 - a0, 4, fn))) return rc;
    if ((rc = use_device(h))) return rc;
    if (a0 == a1) return OSTR_OK;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = (int64_t)(a1 - a0) * h->geom.nbx * h->geom.nby;
    if ((rc = ensure_work(h, n * sizeof(float), st, fn))) return rc;
    hipError_t e = ostr::launch_transpose_bins(diff, h->work, a1 - a0, h->geom.nbx, h->geom.nby,
                                               OSTR_F32, st);
    if (e == hipSuccess)
        e = ostr::launch_sart_update(h->dg, h->views_host.data(), a0, a1, img,
                                     (const float*)h->work, prior, lambda, beta, st);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, fn);
}
}  // namespace

OSTR_API int ostr_bproj_sart(ostr_handle* h, float* img, const float* diff, int32_t a0,
                             int32_t a1, float lambda, void* stream) {
    return bproj_sart(h, img, diff, nullptr, a0, a1, lambda, 0.0f, stream, "ostr_bproj_sart");
}

OSTR_API int ostr_bproj_sart_r(ostr_handle* h, float* img, const float* diff, const float* prior,
                               int32_t a0, int32_t a1, float lambda, float beta, void* stream) {
    if (h && !prior) return fail(h, OSTR_E_ARG, "ostr_bproj_sart_r: NULL prior");
    return bproj_sart(h, img, diff, prior, a0, a1, lambda, beta, stream, "ostr_bproj_sart_r");
}

OSTR_API int ostr_prior_sart(ostr_handle* h, float* prior, const float* img, float delta,
                             void* stream) {
    CHECK_HANDLE(h);
    if (!prior || !img) return fail(h, OSTR_E_ARG, "ostr_prior_sart: NULL buffer");
    if (prior == img) return fail(h, OSTR_E_ARG, "ostr_prior_sart: prior and img alias");
    if (delta == 0.0f) return fail(h, OSTR_E_ARG, "ostr_prior_sart: delta cannot be zero");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_huber_prior(h->dg, prior, img, delta, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_prior_sart");
}

OSTR_API int ostr_line_integrals(ostr_handle* h, float* p, const float* y, const float* b,
                                 int64_t n, void* stream) {
    CHECK_HANDLE(h);
    if (!p || !y || !b || n < 0) return fail(h, OSTR_E_ARG, "ostr_line_integrals: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_line_integrals(p, y, b, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_line_integrals");
}

OSTR_API int ostr_fbp_filter(ostr_handle* h, float* q, const float* p, const float* taps,
                             int32_t ntaps, int32_t a0, int32_t a1, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_fbp_filter"))) return rc;
    if (!q || !p || !taps || q == p) return fail(h, OSTR_E_ARG, "ostr_fbp_filter: bad buffers");
    if (ntaps <= 0 || ntaps % 2 == 0)
        return fail(h, OSTR_E_ARG, "ostr_fbp_filter: ntaps must be odd and positive");
    const int64_t L = (int64_t)1 << ostr::fbp_fft_log2(h->geom.nbx);
    if ((int64_t)(L + L / 2) * 8 > h->lds_limit)
        return fail(h, OSTR_E_RANGE,
                    "ostr_fbp_filter: nbx=%d needs a %lld-point FFT, beyond the %d bytes of LDS "
                    "per work-group",
                    h->geom.nbx, (long long)L, h->lds_limit);
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = ensure_work(h, (size_t)(L + L / 2) * sizeof(float2), st, "ostr_fbp_filter")))
        return rc;
    hipError_t e = ostr::launch_fbp_filter(h->dg, h->views_dev, a0, a1, taps, ntaps, p,
                                           (float2*)h->work, q, h->lds_limit, st);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_fbp_filter (spectrum, FFT filter)");
}

// ---- ReconDBT SART / ML steps ------------------------------------------------

OSTR_API int ostr_rdbt_sart_residual(ostr_handle* h, float* y, const float* p, const float* est,
                                     const float* l, float thresh, int64_t n, void* stream) {
    CHECK_HANDLE(h);
    if (!y || !p || !est || !l || n < 0)
        return fail(h, OSTR_E_ARG, "ostr_rdbt_sart_residual: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_rdbt_sart_residual(y, p, est, l, thresh, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_rdbt_sart_residual");
}

OSTR_API int ostr_rdbt_sart_update(ostr_handle* h, float* x, const float* num, const float* den,
                                   float step, int32_t clamp, int64_t n, void* stream) {
    CHECK_HANDLE(h);
    if (!x || !num || !den || n < 0)
        return fail(h, OSTR_E_ARG, "ostr_rdbt_sart_update: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_rdbt_sart_update(x, num, den, step, clamp, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_rdbt_sart_update");
}

OSTR_API int ostr_rdbt_ml_terms(ostr_handle* h, float* ydif, float* wdrr, const float* drr,
                                const float* counts, const float* i0, float i0_scalar, int64_t n,
                                void* stream) {
    CHECK_HANDLE(h);
    if (!ydif || !wdrr || !drr || !counts || n < 0 || ydif == wdrr)
        return fail(h, OSTR_E_ARG, "ostr_rdbt_ml_terms: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_rdbt_ml_terms(ydif, wdrr, drr, counts, i0, i0_scalar, n,
                                              (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_rdbt_ml_terms");
}

OSTR_API int ostr_rdbt_ml_update(ostr_handle* h, float* x, float* upd, const float* num,
                                 const float* den, float step, int64_t n, void* stream) {
    CHECK_HANDLE(h);
    if (!x || !upd || !num || !den || n < 0)
        return fail(h, OSTR_E_ARG, "ostr_rdbt_ml_update: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_rdbt_ml_update(x, upd, num, den, step, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_rdbt_ml_update");
}

OSTR_API int ostr_fproj_4r(ostr_handle* h, const float* img, float* prj, float* prj_4r,
                           int32_t a0, int32_t a1, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_fproj_4r"))) return rc;
    if (!img || !prj) return fail(h, OSTR_E_ARG, "ostr_fproj_4r: NULL buffer");
    if (prj == prj_4r) return fail(h, OSTR_E_ARG, "ostr_fproj_4r: prj and prj_4r alias");
    if ((int64_t)4 * h->geom.nbx * h->geom.nby >= ((int64_t)1 << 31))
        return fail(h, OSTR_E_RANGE, "ostr_fproj_4r: too many sub-bins per view");
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_project_4r(h->dg, h->views_dev, a0, a1, img, prj, prj_4r,
                                           (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_fproj_4r");
}

OSTR_API int ostr_sart_fproj_diff(ostr_handle* h, const float* img, const float* p_rows,
                                  float* diff_rows, int32_t a0, int32_t a1, int32_t cos_weight,
                                  void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_sart_fproj_diff"))) return rc;
    if (!img || !p_rows || !diff_rows)
        return fail(h, OSTR_E_ARG, "ostr_sart_fproj_diff: NULL buffer");
    if (p_rows == diff_rows) return fail(h, OSTR_E_ARG, "ostr_sart_fproj_diff: p and diff alias");
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_sart_project_diff(h->dg, h->views_dev, a0, a1, cos_weight, img,
                                                  p_rows, diff_rows, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_sart_fproj_diff");
}

OSTR_API int ostr_sart_update_fused(ostr_handle* h, const float* x_in, float* x_out,
                                    const float* diff_rows, int32_t a0, int32_t a1,
                                    float lambda, float beta, float delta, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_sart_update_fused"))) return rc;
    if (!x_in || !x_out || !diff_rows)
        return fail(h, OSTR_E_ARG, "ostr_sart_update_fused: NULL buffer");
    if (x_in == x_out)
        return fail(h, OSTR_E_ARG, "ostr_sart_update_fused: x_in and x_out must differ");
    if (beta != 0.0f && delta == 0.0f)
        return fail(h, OSTR_E_ARG, "ostr_sart_update_fused: delta cannot be zero");
    if (a1 - a0 > ostr::MAX_SLOTS)
        return fail(h, OSTR_E_RANGE, "ostr_sart_update_fused: at most %d views per call",
                    ostr::MAX_SLOTS);
    if ((rc = check_sino_span(h, a1 - a0, 4, "ostr_sart_update_fused"))) return rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_sart_update_fused(h->dg, h->views_host.data(), a0, a1, x_in,
                                                  x_out, diff_rows, lambda, beta, delta,
                                                  (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_sart_update_fused");
}

OSTR_API int ostr_bproj_matched(ostr_handle* h, float* img_accum, const void* prj,
                                int32_t prj_dtype, int32_t a0, int32_t a1, int32_t cos_weight,
                                int32_t rows_layout, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_bproj_matched")) ||
        (rc = check_dtype(h, prj_dtype, "ostr_bproj_matched")))
        return rc;
    if (!img_accum || !prj) return fail(h, OSTR_E_ARG, "ostr_bproj_matched: NULL buffer");
    if ((rc = check_sino_span(h, a1 - a0, ostr::elem_size(prj_dtype), "ostr_bproj_matched")))
        return rc;
    if ((rc = use_device(h))) return rc;
    if (a0 == a1) return OSTR_OK;
    hipStream_t st = (hipStream_t)stream;
    const void* rows = prj;
    hipError_t e = hipSuccess;
    if (!rows_layout) {
        const int64_t n = (int64_t)(a1 - a0) * h->geom.nbx * h->geom.nby;
        if ((rc = ensure_work(h, n * ostr::elem_size(prj_dtype), st, "ostr_bproj_matched")))
            return rc;
        e = ostr::launch_transpose_bins(prj, h->work, a1 - a0, h->geom.nbx, h->geom.nby,
                                        prj_dtype, st);
        rows = h->work;
    }
    if (e == hipSuccess)
        e = ostr::launch_matched_rows(h->dg, h->views_host.data(), a0, a1, cos_weight, img_accum,
                                      rows, prj_dtype, st);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_bproj_matched");
}

OSTR_API int ostr_bp_init(ostr_handle* h, float* img, const float* p_or_y, const float* b,
                          int32_t counts, void* stream) {
    CHECK_HANDLE(h);
    if (!img || !p_or_y || (counts && !b))
        return fail(h, OSTR_E_ARG, "ostr_bp_init: NULL buffer");
    int rc;
    if ((rc = check_sino_span(h, 1, 4, "ostr_bp_init"))) return rc;
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const ostr_geom& G = h->geom;
    const int64_t nbin = (int64_t)G.nbx * G.nby, nvox = (int64_t)G.nx * G.ny * G.nz;
    if ((rc = ensure_work(h, (size_t)(3 * nbin + 2 * nvox) * sizeof(float), st, "ostr_bp_init")))
        return rc;
    float* l = (float*)h->work;
    float* q = l + nbin;
    float* ones = q + nbin;
    float* num = ones + nbin;
    float* den = num + nvox;
    const float thr = std::fabs(G.dz) / 10.0f;
    hipError_t e = ostr::launch_fill(img, 0.0f, nvox, st);
    if (e == hipSuccess) e = ostr::launch_fill(ones, 1.0f, nbin, st);
    // natural view order, identity table (BP.m loops over the views as acquired)
    for (int v = 0; v < h->n_views && e == hipSuccess; ++v) {
        e = ostr::launch_project(h->dg, h->views_nat_dev, v, v + 1, 2, 0, nullptr, OSTR_F32, l,
                                 OSTR_F32, nullptr, nullptr, nullptr, st);
        if (e == hipSuccess)
            e = ostr::launch_bp_data(q, l, p_or_y + v * nbin, counts ? b + v * nbin : nullptr,
                                     counts, h->views_nat[v].c, thr, G.nbx, G.nby, st);
        if (e == hipSuccess) e = ostr::launch_fill(num, 0.0f, 2 * nvox, st);
        if (e == hipSuccess)
            e = ostr::launch_matched_rows(h->dg, h->views_nat.data(), v, v + 1, 0, num, q,
                                          OSTR_F32, st);
        if (e == hipSuccess)
            e = ostr::launch_matched_rows(h->dg, h->views_nat.data(), v, v + 1, 0, den, ones,
                                          OSTR_F32, st);
        if (e == hipSuccess) e = ostr::launch_bp_accum(img, num, den, nvox, st);
    }
    if (e == hipSuccess) e = ostr::launch_div(img, (float)h->n_views, nvox, st);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_bp_init");
}

OSTR_API int ostr_bproj_sbp_rows(ostr_handle* h, float* img_accum, const void* prj,
                                 int32_t prj_dtype, int32_t a0, int32_t a1, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_bproj_sbp_rows")) ||
        (rc = check_dtype(h, prj_dtype, "ostr_bproj_sbp_rows")))
        return rc;
    if (!img_accum || !prj) return fail(h, OSTR_E_ARG, "ostr_bproj_sbp_rows: NULL buffer");
    if ((rc = check_sino_span(h, a1 - a0, ostr::elem_size(prj_dtype), "ostr_bproj_sbp_rows")))
        return rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_sbp_rows(h->dg, h->views_host.data(), a0, a1, img_accum, prj,
                                         prj_dtype, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_bproj_sbp_rows");
}

OSTR_API int ostr_prior(ostr_handle* h, float* rdd, float* rd, const void* img,
                        int32_t img_dtype, int32_t potential, float delta, float z_xy_ratio,
                        void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_dtype(h, img_dtype, "ostr_prior"))) return rc;
    if (potential < OSTR_PRIOR_Q2 || potential > OSTR_PRIOR_HYPERBOLA)
        return fail(h, OSTR_E_ARG, "ostr_prior: unknown potential %d", potential);
    if (potential != OSTR_PRIOR_Q2 && delta == 0.0f)
        return fail(h, OSTR_E_ARG, "ostr_prior: delta cannot be zero");
    if (!(z_xy_ratio > 0.0f)) return fail(h, OSTR_E_ARG, "ostr_prior: z_xy_ratio must be > 0");
    if (!rdd || !rd || !img) return fail(h, OSTR_E_ARG, "ostr_prior: NULL buffer");
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_prior(h->dg, rdd, rd, img, img_dtype, potential, delta,
                                      z_xy_ratio, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_prior");
}

// --------------------------------------------------------- elementwise --
OSTR_API int ostr_yry(ostr_handle* h, float* yry, const float* y, const float* r, int64_t n,
                      void* stream) {
    CHECK_HANDLE(h);
    if (!yry || !y || n < 0) return fail(h, OSTR_E_ARG, "ostr_yry: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_yry(yry, y, r, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_yry");
}

OSTR_API int ostr_mul(ostr_handle* h, float* out, const float* a, const float* b, int64_t n,
                      void* stream) {
    CHECK_HANDLE(h);
    if (!out || !a || !b || n < 0) return fail(h, OSTR_E_ARG, "ostr_mul: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_mul(out, a, b, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_mul");
}

OSTR_API int ostr_compute_h(ostr_handle* h, float* hsub, const float* y, const float* b,
                            const float* l, const float* r, int64_t n, void* stream) {
    CHECK_HANDLE(h);
    if (!hsub || !y || !b || !l || n < 0)
        return fail(h, OSTR_E_ARG, "ostr_compute_h: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_compute_h(hsub, y, b, l, r, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_compute_h");
}

OSTR_API int ostr_update_est(ostr_handle* h, float* est, const float* capL, const float* rd,
                             const float* d, const float* rdd, float beta, float capl_scale,
                             int32_t guard, void* stream) {
    CHECK_HANDLE(h);
    if (!est || !capL || !rd || !d || !rdd)
        return fail(h, OSTR_E_ARG, "ostr_update_est: NULL buffer");
    int rc;
    if ((rc = use_device(h))) return rc;
    int64_t n = (int64_t)h->geom.nx * h->geom.ny * h->geom.nz;
    hipError_t e = ostr::launch_update_est(est, capL, rd, d, rdd, beta, capl_scale, guard, n,
                                           (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_update_est");
}

OSTR_API int ostr_extrapolate(ostr_handle* h, float* x, const float* z, const float* z_prev,
                              float c, void* stream) {
    CHECK_HANDLE(h);
    if (!x || !z || !z_prev) return fail(h, OSTR_E_ARG, "ostr_extrapolate: NULL buffer");
    if (x == z || x == z_prev) return fail(h, OSTR_E_ARG, "ostr_extrapolate: x aliases an input");
    int rc;
    if ((rc = use_device(h))) return rc;
    int64_t n = (int64_t)h->geom.nx * h->geom.ny * h->geom.nz;
    hipError_t e = ostr::launch_extrapolate(x, z, z_prev, c, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_extrapolate");
}

OSTR_API int ostr_restart_dot(ostr_handle* h, const float* x, const float* z,
                              const float* z_prev, double* out, void* stream) {
    CHECK_HANDLE(h);
    if (!x || !z || !z_prev || !out) return fail(h, OSTR_E_ARG, "ostr_restart_dot: NULL buffer");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = ensure_dwork(h, 1024 + 1, st, "ostr_restart_dot"))) return rc;
    int64_t n = (int64_t)h->geom.nx * h->geom.ny * h->geom.nz;
    hipError_t e = ostr::launch_restart_dot(x, z, z_prev, n, h->dwork, h->dwork + 1024, st);
    double host = 0.0;
    if (e == hipSuccess)
        e = hipMemcpyAsync(&host, h->dwork + 1024, sizeof host, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(h, e, "ostr_restart_dot");
    *out = host;
    return OSTR_OK;
}

OSTR_API int ostr_momentum_init(ostr_handle* h, double* state, void* stream) {
    CHECK_HANDLE(h);
    if (!state) return fail(h, OSTR_E_ARG, "ostr_momentum_init: NULL state");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_momentum_init(state, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_momentum_init");
}

OSTR_API int ostr_momentum_step(ostr_handle* h, void* x_next, const void* z,
                                const void* z_prev, const void* x_k, int32_t dtype,
                                double* state, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_dtype(h, dtype, "ostr_momentum_step"))) return rc;
    if (!x_next || !z || !z_prev || !x_k || !state)
        return fail(h, OSTR_E_ARG, "ostr_momentum_step: NULL buffer");
    if (x_next == z || x_next == z_prev || x_next == x_k)
        return fail(h, OSTR_E_ARG, "ostr_momentum_step: x_next aliases an input");
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = ensure_dwork(h, 1024 + 4, st, "ostr_momentum_step"))) return rc;
    int64_t n = (int64_t)h->geom.nx * h->geom.ny * h->geom.nz;
    hipError_t e =
        ostr::launch_momentum_step(x_next, z, z_prev, x_k, dtype, n, h->dwork, state, st);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_momentum_step");
}

OSTR_API int ostr_accum_f64(ostr_handle* h, double* acc, const float* x, int64_t n,
                            void* stream) {
    CHECK_HANDLE(h);
    if (!acc || !x || n < 0) return fail(h, OSTR_E_ARG, "ostr_accum_f64: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_add_f64(acc, x, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_accum_f64");
}

OSTR_API int ostr_f64_to_f32(ostr_handle* h, float* out, const double* acc, int64_t n,
                             void* stream) {
    CHECK_HANDLE(h);
    if (!acc || !out || n < 0) return fail(h, OSTR_E_ARG, "ostr_f64_to_f32: bad argument");
    int rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_f64_to_f32(out, acc, n, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_f64_to_f32");
}

OSTR_API int ostr_regroup(ostr_handle* h, float* dst, const float* src, int32_t n_subsets,
                          int32_t to_rows, void* stream) {
    CHECK_HANDLE(h);
    if (!dst || !src || dst == src) return fail(h, OSTR_E_ARG, "ostr_regroup: bad buffers");
    if (n_subsets <= 0 || h->n_views % n_subsets)
        return fail(h, OSTR_E_ARG, "ostr_regroup: %d views not divisible into %d subsets",
                    h->n_views, n_subsets);
    int rc;
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nbin = (int64_t)h->geom.nbx * h->geom.nby;
    hipError_t e = hipSuccess;
    if (!to_rows) {
        e = ostr::launch_regroup(dst, src, h->n_views, n_subsets, nbin, st);
    } else {
        const int vs = h->n_views / n_subsets;
        for (int k = 0; k < h->n_views && e == hipSuccess; ++k) {
            const int m = k / vs, j = k % vs;
            e = ostr::launch_transpose_bins(src + (int64_t)(j * n_subsets + m) * nbin,
                                            dst + (int64_t)k * nbin, 1, h->geom.nbx,
                                            h->geom.nby, OSTR_F32, st);
        }
    }
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_regroup");
}

OSTR_API int ostr_to_rows(ostr_handle* h, void* dst, const void* src, int32_t n_slots,
                          int32_t dtype, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_dtype(h, dtype, "ostr_to_rows"))) return rc;
    if (!dst || !src || dst == src || n_slots < 0)
        return fail(h, OSTR_E_ARG, "ostr_to_rows: bad buffers");
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_transpose_bins(src, dst, n_slots, h->geom.nbx, h->geom.nby,
                                               dtype, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_to_rows");
}

// ---------------------------------------------------------------- fused --
OSTR_API int ostr_precompute_d(ostr_handle* h, float* d, const float* y, const float* r,
                               int32_t gamma_mode, void* stream) {
    CHECK_HANDLE(h);
    if (!d || !y) return fail(h, OSTR_E_ARG, "ostr_precompute_d: NULL buffer");
    if (gamma_mode != OSTR_GAMMA_RAY_LENGTH && gamma_mode != OSTR_GAMMA_ROW_SUM)
        return fail(h, OSTR_E_ARG, "ostr_precompute_d: unknown gamma mode %d", gamma_mode);
    int rc;
    if ((rc = check_sino_span(h, h->n_views, 4, "ostr_precompute_d"))) return rc;
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nall = (int64_t)h->n_views * h->geom.nbx * h->geom.nby;
    if ((rc = ensure_work(h, nall * sizeof(float), st, "ostr_precompute_d"))) return rc;
    const int64_t nvox = (int64_t)h->geom.nx * h->geom.ny * h->geom.nz;
    // gamma and d use the natural view order (computed before the OS regroup,
    // Tomo_OSTR_gold.cpp:450-538); gamma*yry goes to the work buffer row-major
    hipError_t e = ostr::launch_project(h->dg, h->views_nat_dev, 0, h->n_views,
                                        gamma_mode == OSTR_GAMMA_ROW_SUM ? 5 : 4, 0, nullptr,
                                        OSTR_F32, h->work, OSTR_F32, y, nullptr, r, st);
    if (e == hipSuccess) e = ostr::launch_fill(d, 0.0f, nvox, st);
    if (e == hipSuccess)
        e = ostr::launch_sbp_rows(h->dg, h->views_nat.data(), 0, h->n_views, d, h->work,
                                  OSTR_F32, st);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_precompute_d");
}

OSTR_API int ostr_precompute_d_f64(ostr_handle* h, float* d, const float* y, const float* r,
                                   int32_t gamma_mode, void* stream) {
    CHECK_HANDLE(h);
    if (!d || !y) return fail(h, OSTR_E_ARG, "ostr_precompute_d_f64: NULL buffer");
    if (gamma_mode != OSTR_GAMMA_RAY_LENGTH && gamma_mode != OSTR_GAMMA_ROW_SUM)
        return fail(h, OSTR_E_ARG, "ostr_precompute_d_f64: unknown gamma mode %d", gamma_mode);
    int rc;
    if ((rc = check_sino_span(h, 1, 4, "ostr_precompute_d_f64"))) return rc;
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nbin = (int64_t)h->geom.nbx * h->geom.nby;
    const int64_t nall = (int64_t)h->n_views * nbin;
    const int64_t nvox = (int64_t)h->geom.nx * h->geom.ny * h->geom.nz;
    // work: [acc: nvox doubles][one view's SBP: nvox floats][gamma*yry: nall floats]
    if ((rc = ensure_work(h, (size_t)(nvox * 8 + (nvox + nall) * 4), st,
                          "ostr_precompute_d_f64")))
        return rc;
    double* acc = (double*)h->work;
    float* tmp = (float*)(acc + nvox);
    float* gc = tmp + nvox;
    // gamma*yry row-major in natural view order, as ostr_precompute_d; then the
    // views' SBP terms one view at a time, summed in FP64 in view order
    hipError_t e = ostr::launch_project(h->dg, h->views_nat_dev, 0, h->n_views,
                                        gamma_mode == OSTR_GAMMA_ROW_SUM ? 5 : 4, 0, nullptr,
                                        OSTR_F32, gc, OSTR_F32, y, nullptr, r, st);
    if (e == hipSuccess) e = ostr::launch_fill_f64(acc, 0.0, nvox, st);
    for (int v = 0; v < h->n_views && e == hipSuccess; ++v) {
        e = ostr::launch_fill(tmp, 0.0f, nvox, st);
        if (e == hipSuccess)
            e = ostr::launch_sbp_rows(h->dg, h->views_nat.data(), v, v + 1, tmp, gc + v * nbin,
                                      OSTR_F32, st);
        if (e == hipSuccess) e = ostr::launch_add_f64(acc, tmp, nvox, st);
    }
    if (e == hipSuccess) e = ostr::launch_f64_to_f32(d, acc, nvox, st);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_precompute_d_f64");
}

OSTR_API int ostr_fproj_h(ostr_handle* h, const void* img, int32_t img_dtype, const float* y,
                          const float* b, const float* r, void* hsub, int32_t prj_dtype,
                          int32_t a0, int32_t a1, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_fproj_h")) ||
        (rc = check_dtype(h, img_dtype, "ostr_fproj_h")) ||
        (rc = check_dtype(h, prj_dtype, "ostr_fproj_h")))
        return rc;
    if (!img || !y || !b || !hsub) return fail(h, OSTR_E_ARG, "ostr_fproj_h: NULL buffer");
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_project(h->dg, h->views_dev, a0, a1, 3, 1, img, img_dtype, hsub,
                                        prj_dtype, y, b, r, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_fproj_h");
}

OSTR_API int ostr_update_fused(ostr_handle* h, const void* x_in, void* x_out, int32_t img_dtype,
                               const float* d, const void* hsub, int32_t prj_dtype, int32_t a0,
                               int32_t a1, float beta, float capl_scale, int32_t potential,
                               float delta, float z_xy_ratio, int32_t guard, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_update_fused")) ||
        (rc = check_dtype(h, img_dtype, "ostr_update_fused")) ||
        (rc = check_dtype(h, prj_dtype, "ostr_update_fused")))
        return rc;
    if (!x_in || !x_out || !d || !hsub)
        return fail(h, OSTR_E_ARG, "ostr_update_fused: NULL buffer");
    if (x_in == x_out) return fail(h, OSTR_E_ARG, "ostr_update_fused: x_in and x_out alias");
    if (potential < OSTR_PRIOR_Q2 || potential > OSTR_PRIOR_HYPERBOLA)
        return fail(h, OSTR_E_ARG, "ostr_update_fused: unknown potential %d", potential);
    if (potential != OSTR_PRIOR_Q2 && delta == 0.0f)
        return fail(h, OSTR_E_ARG, "ostr_update_fused: delta cannot be zero");
    if (!(z_xy_ratio > 0.0f))
        return fail(h, OSTR_E_ARG, "ostr_update_fused: z_xy_ratio must be > 0");
    if (a1 - a0 > ostr::MAX_SLOTS)
        return fail(h, OSTR_E_RANGE, "ostr_update_fused: at most %d views per subset",
                    ostr::MAX_SLOTS);
    if ((rc = check_sino_span(h, a1 - a0, ostr::elem_size(prj_dtype), "ostr_update_fused")))
        return rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_update_fused(h->dg, h->views_host.data(), a0, a1, x_in, x_out,
                                             img_dtype, d, hsub, prj_dtype, beta, capl_scale,
                                             potential, delta, z_xy_ratio, guard,
                                             (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_update_fused");
}

OSTR_API int ostr_update_fused_matched(ostr_handle* h, const void* x_in, void* x_out,
                                       int32_t img_dtype, const float* d, const void* hsub,
                                       int32_t prj_dtype, int32_t a0, int32_t a1, float beta,
                                       float capl_scale, int32_t potential, float delta,
                                       float z_xy_ratio, int32_t guard, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_update_fused_matched")) ||
        (rc = check_dtype(h, img_dtype, "ostr_update_fused_matched")) ||
        (rc = check_dtype(h, prj_dtype, "ostr_update_fused_matched")))
        return rc;
    if (!x_in || !x_out || !d || !hsub)
        return fail(h, OSTR_E_ARG, "ostr_update_fused_matched: NULL buffer");
    if (x_in == x_out)
        return fail(h, OSTR_E_ARG, "ostr_update_fused_matched: x_in and x_out alias");
    if (potential < OSTR_PRIOR_Q2 || potential > OSTR_PRIOR_HYPERBOLA)
        return fail(h, OSTR_E_ARG, "ostr_update_fused_matched: unknown potential %d", potential);
    if (potential != OSTR_PRIOR_Q2 && delta == 0.0f)
        return fail(h, OSTR_E_ARG, "ostr_update_fused_matched: delta cannot be zero");
    if (!(z_xy_ratio > 0.0f))
        return fail(h, OSTR_E_ARG, "ostr_update_fused_matched: z_xy_ratio must be > 0");
    if (a1 - a0 > ostr::MAX_SLOTS)
        return fail(h, OSTR_E_RANGE, "ostr_update_fused_matched: at most %d views per subset",
                    ostr::MAX_SLOTS);
    if ((rc = check_sino_span(h, a1 - a0, ostr::elem_size(prj_dtype),
                              "ostr_update_fused_matched")))
        return rc;
    if ((rc = use_device(h))) return rc;
    hipError_t e = ostr::launch_update_fused_matched(
        h->dg, h->views_host.data(), a0, a1, x_in, x_out, img_dtype, d, hsub, prj_dtype, beta,
        capl_scale, potential, delta, z_xy_ratio, guard, (hipStream_t)stream);
    return e == hipSuccess ? OSTR_OK : hip_fail(h, e, "ostr_update_fused_matched");
}

OSTR_API int ostr_count_nnz(ostr_handle* h, int32_t a0, int32_t a1, int64_t* nnz_out,
                            void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_range(h, a0, a1, "ostr_count_nnz"))) return rc;
    if (!nnz_out) return fail(h, OSTR_E_ARG, "ostr_count_nnz: NULL output");
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    unsigned long long host = 0;
    hipError_t e = hipMemcpyAsync(h->counter, &host, sizeof host, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = ostr::launch_count_nnz(h->dg, h->views_dev, a0, a1, h->counter, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&host, h->counter, sizeof host, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(h, e, "ostr_count_nnz");
    *nnz_out = (int64_t)host;
    return OSTR_OK;
}

// ---- FP64 objective --------------------------------------------------------

namespace {

// Validation shared by both objective entry points.
int objective_args(ostr_handle* h, int32_t img_dtype, const void* img, const float* y,
                   const float* b, double* out, int32_t potential, double delta,
                   const char* fn) {
    int rc;
    if ((rc = check_dtype(h, img_dtype, fn))) return rc;
    if (!img || !y || !b || !out) return fail(h, OSTR_E_ARG, "%s: NULL buffer", fn);
    if (potential < OSTR_PRIOR_Q2 || potential > OSTR_PRIOR_HYPERBOLA)
        return fail(h, OSTR_E_ARG, "%s: unknown potential %d", fn, potential);
    if (potential != OSTR_PRIOR_Q2 && delta == 0.0)
        return fail(h, OSTR_E_ARG, "%s: delta cannot be zero", fn);
    return OSTR_OK;
}

// The objective's launches with `work` holding nd + np + 2 doubles (the data
// partials, the penalty partials, the two sums); every partial is written by
// its block, so nothing needs clearing.
int objective_run(ostr_handle* h, const void* img, int32_t img_dtype, const float* y,
                  const float* b, const float* r, int32_t potential, double delta,
                  double z_xy_ratio, double* out, double* work, hipStream_t st,
                  const char* fn) {
    const int64_t nd = ostr::objective_data_partials(h->dg, h->n_views);
    const int64_t np = ostr::penalty_partials(h->dg);
    double* pd = work;
    double* pp = work + nd;
    double* res = work + nd + np;
    int64_t n1 = 0, n2 = 0;
    // all views in natural order: the objective does not depend on the OS regroup
    hipError_t e = ostr::launch_objective_data(h->dg, h->views_nat_dev, h->n_views, img,
                                               img_dtype, y, b, r, pd, &n1, st);
    if (e == hipSuccess)
        e = ostr::launch_penalty(h->dg, img, img_dtype, potential, delta, z_xy_ratio, pp, &n2, st);
    if (e == hipSuccess) e = ostr::launch_reduce_partials(pd, n1, res, st);
    if (e == hipSuccess) e = ostr::launch_reduce_partials(pp, n2, res + 1, st);
    double host[2] = {0.0, 0.0};
    if (e == hipSuccess) e = hipMemcpyAsync(host, res, sizeof host, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(h, e, fn);
    out[0] = host[0];
    out[1] = 0.5 * host[1];
    return OSTR_OK;
}

}  // namespace

OSTR_API int ostr_objective(ostr_handle* h, const void* img, int32_t img_dtype, const float* y,
                            const float* b, const float* r, int32_t potential, double delta,
                            double z_xy_ratio, double* out, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = objective_args(h, img_dtype, img, y, b, out, potential, delta, "ostr_objective")))
        return rc;
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nd = ostr::objective_data_partials(h->dg, h->n_views);
    const int64_t np = ostr::penalty_partials(h->dg);
    if ((rc = ensure_dwork(h, (size_t)(nd + np + 2), st, "ostr_objective"))) return rc;
    return objective_run(h, img, img_dtype, y, b, r, potential, delta, z_xy_ratio, out, h->dwork,
                         st, "ostr_objective");
}

OSTR_API int ostr_objective_noalloc(ostr_handle* h, const void* img, int32_t img_dtype,
                                    const float* y, const float* b, const float* r,
                                    int32_t potential, double delta, double z_xy_ratio,
                                    double* out, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = objective_args(h, img_dtype, img, y, b, out, potential, delta,
                             "ostr_objective_noalloc")))
        return rc;
    const int64_t nd = ostr::objective_data_partials(h->dg, h->n_views);
    const int64_t np = ostr::penalty_partials(h->dg);
    if (nd + np + 2 > NOALLOC_DOUBLES)
        return fail(h, OSTR_E_RANGE,
                    "ostr_objective_noalloc: needs %lld FP64 partials (%lld data + %lld penalty "
                    "+ 2), the static scratch holds %lld; use ostr_objective",
                    (long long)(nd + np + 2), (long long)nd, (long long)np,
                    (long long)NOALLOC_DOUBLES);
    double* work = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        work = g_noalloc_ptr[h->device];
    }
    if (!work)
        return fail(h, OSTR_E_HIP, "ostr_objective_noalloc: scratch of device %d not resolved",
                    h->device);
    if ((rc = use_device(h))) return rc;
    return objective_run(h, img, img_dtype, y, b, r, potential, delta, z_xy_ratio, out, work,
                         (hipStream_t)stream, "ostr_objective_noalloc");
}

OSTR_API int ostr_diff_norms(ostr_handle* h, const void* a, const void* b, int32_t dtype,
                             int64_t n, double* out, void* stream) {
    CHECK_HANDLE(h);
    int rc;
    if ((rc = check_dtype(h, dtype, "ostr_diff_norms"))) return rc;
    if (!a || !b || !out || n < 0) return fail(h, OSTR_E_ARG, "ostr_diff_norms: bad argument");
    if ((rc = use_device(h))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int64_t np = ostr::diff_norms_partials(n);
    if ((rc = ensure_dwork(h, (size_t)(np + 2), st, "ostr_diff_norms"))) return rc;
    int64_t n1 = 0;
    hipError_t e = ostr::launch_diff_norms(a, b, dtype, n, h->dwork, &n1, st);
    if (e == hipSuccess) e = ostr::launch_reduce_partials(h->dwork, n1 / 2, h->dwork + np, st);
    if (e == hipSuccess)
        e = ostr::launch_reduce_partials(h->dwork + n1 / 2, n1 / 2, h->dwork + np + 1, st);
    double host[2] = {0.0, 0.0};
    if (e == hipSuccess)
        e = hipMemcpyAsync(host, h->dwork + np, sizeof host, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(h, e, "ostr_diff_norms");
    out[0] = host[0];
    out[1] = host[1];
    return OSTR_OK;
}

// ---- batched cases ---------------------------------------------------------

// A launch table: the tile schedule (2 * 8 * (n + 1) ints, see BatchWalk) at
// offset 0 and the case entries at `case_off`.
struct ostr_batch {
    static constexpr int RING = 8;  // tables in flight
    int device = 0;
    int max_cases = 0;
    int n_cu = 0;
    size_t case_off = 0, slot_bytes = 0;
    unsigned char* host[RING] = {};  // pinned staging
    unsigned char* dev = nullptr;    // RING tables of slot_bytes
    hipEvent_t ev[RING] = {};
    bool used[RING] = {};
    int pos = 0;
    std::string err;
};

namespace {

int bfail(ostr_batch* bt, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (bt)
        bt->err = buf;
    else
        g_err = buf;
    return code;
}

int bhip(ostr_batch* bt, hipError_t e, const char* where) {
    return bfail(bt, OSTR_E_HIP, "%s: HIP error %d (%s)", where, (int)e, hipGetErrorString(e));
}

void batch_release(ostr_batch* bt) {
    for (int k = 0; k < ostr_batch::RING; ++k) {
        if (bt->ev[k]) (void)hipEventDestroy(bt->ev[k]);
        if (bt->host[k]) (void)hipHostFree(bt->host[k]);
    }
    if (bt->dev) (void)hipFree(bt->dev);
    delete bt;
}

// Validates the cases, fills the next staging table and copies it to the
// device on `st`. kind 0 = forward, 1 = update. On success *table is the
// device table, *total the number of tiles and *slot the ring position.
int batch_grid(const ostr_batch* bt, int total, int groups_per_cu) {
    if (groups_per_cu == 0) return total;
    const int64_t g = (int64_t)bt->n_cu * groups_per_cu;
    return (int)(g < total ? g : total);
}

int batch_stage(ostr_batch* bt, const ostr_batch_case* cases, int n, int prj_dtype,
                int potential, const ostr_tile_cfg* cfg, int kind, hipStream_t st,
                const char* fn, ostr::BatchCase** table, const int** sched, int* total,
                int* grid_out, int* slot, int* single_pass) {
    if (!bt) return bfail(nullptr, OSTR_E_ARG, "%s: NULL batch", fn);
    if (!cases || !cfg) return bfail(bt, OSTR_E_ARG, "%s: NULL argument", fn);
    if (n < 1 || n > bt->max_cases)
        return bfail(bt, OSTR_E_RANGE, "%s: %d cases, batch holds 1..%d", fn, n, bt->max_cases);
    if (prj_dtype != OSTR_F32 && prj_dtype != OSTR_BF16 && prj_dtype != OSTR_F16)
        return bfail(bt, OSTR_E_UNSUPPORTED, "%s: unknown dtype %d", fn, prj_dtype);
    if (kind == 0) {
        const int r = cfg->fwd_rows;
        if (r != 1 && r != 2 && r != 4 && r != 8)
            return bfail(bt, OSTR_E_ARG, "%s: fwd_rows %d not in {1,2,4,8}", fn, r);
        if (cfg->fwd_groups_per_cu < 0 || cfg->fwd_groups_per_cu > 64)
            return bfail(bt, OSTR_E_ARG, "%s: fwd_groups_per_cu outside 0..64", fn);
    } else {
        if (!ostr::batch_update_shape_ok(cfg->upd_rows, cfg->upd_waves))
            return bfail(bt, OSTR_E_ARG, "%s: update shape %dx%d not in {8x4, 8x2, 4x4}", fn,
                         cfg->upd_rows, cfg->upd_waves);
        if (cfg->upd_groups_per_cu < 0 || cfg->upd_groups_per_cu > 64)
            return bfail(bt, OSTR_E_ARG, "%s: upd_groups_per_cu outside 0..64", fn);
        if (potential < OSTR_PRIOR_Q2 || potential > OSTR_PRIOR_HYPERBOLA)
            return bfail(bt, OSTR_E_ARG, "%s: unknown potential %d", fn, potential);
    }
    for (int c = 0; c < n; ++c) {
        const ostr_batch_case& k = cases[c];
        if (!k.h) return bfail(bt, OSTR_E_ARG, "%s: case %d: NULL handle", fn, c);
        if (k.h->device != bt->device)
            return bfail(bt, OSTR_E_ARG, "%s: case %d is on device %d, batch on %d", fn, c,
                         k.h->device, bt->device);
        if (k.a0 < 0 || k.a1 > k.h->n_views || k.a0 > k.a1 || k.a1 - k.a0 > ostr::MAX_SLOTS)
            return bfail(bt, OSTR_E_RANGE, "%s: case %d: slots [%d,%d) invalid (n_views %d, "
                         "at most %d per subset)", fn, c, k.a0, k.a1, k.h->n_views,
                         ostr::MAX_SLOTS);
        if (!k.x_in || !k.hsub || (kind == 0 && (!k.y || !k.b)) ||
            (kind == 1 && (!k.x_out || !k.d)))
            return bfail(bt, OSTR_E_ARG, "%s: case %d: NULL buffer", fn, c);
        if (kind == 1) {
            if (k.x_in == k.x_out)
                return bfail(bt, OSTR_E_ARG, "%s: case %d: x_in and x_out alias", fn, c);
            if ((int64_t)(k.a1 - k.a0) * k.h->geom.nbx * k.h->geom.nby *
                    ostr::elem_size(prj_dtype) > (int64_t)0xFFFFFFF0)
                return bfail(bt, OSTR_E_RANGE, "%s: case %d: the subset's h spans 2^32 bytes "
                             "or more (32-bit offsets)", fn, c);
            if (potential != OSTR_PRIOR_Q2 && k.delta == 0.0f)
                return bfail(bt, OSTR_E_ARG, "%s: case %d: delta cannot be zero", fn, c);
            if (!(k.z_xy_ratio > 0.0f))
                return bfail(bt, OSTR_E_ARG, "%s: case %d: z_xy_ratio must be > 0", fn, c);
        }
    }
    hipError_t e = hipSetDevice(bt->device);
    if (e != hipSuccess) return bhip(bt, e, fn);
    const int k = bt->pos;
    if (bt->used[k]) {  // the table of RING calls ago must be consumed
        e = hipEventSynchronize(bt->ev[k]);
        if (e != hipSuccess) return bhip(bt, e, fn);
    }
    ostr::BatchCase* H = (ostr::BatchCase*)(bt->host[k] + bt->case_off);
    std::vector<int64_t> cnt(n);
    int64_t tiles = 0;
    for (int c = 0; c < n; ++c) {
        const ostr_batch_case& in = cases[c];
        ostr::BatchCase& o = H[c];
        const ostr::DevGeom& g = in.h->dg;
        o.g = g;
        o.n_slots = in.a1 - in.a0;
        cnt[c] = kind == 0 ? ostr::batch_fwd_tiles(g, o.n_slots, cfg->fwd_rows)
                           : ostr::batch_upd_tiles(g, cfg->upd_rows, cfg->upd_waves);
        tiles += cnt[c];
        if (tiles > 0x7fffffff) return bfail(bt, OSTR_E_RANGE, "%s: too many tiles", fn);
        o.guard = in.guard;
        o.beta = in.beta;
        o.capl_scale = in.capl_scale;
        o.inv_delta = (kind == 1 && potential != OSTR_PRIOR_Q2) ? 1.0f / in.delta : 0.0f;
        o.x_in = in.x_in;
        o.x_out = in.x_out;
        o.d = in.d;
        o.y = in.y;
        o.b = in.b;
        o.r = in.r;
        o.h = in.hsub;
        ostr::batch_slots(o.S, o.v, in.h->views_host.data(), in.a0, in.a1, g.src_y);
        if (kind == 1) {
            ostr::PriorW w27;
            ostr::prior_weights(w27.w, in.z_xy_ratio);
            o.W = ostr::prior_params(w27, potential, in.delta);
        }
    }
    const int gpc = kind == 0 ? cfg->fwd_groups_per_cu : cfg->upd_groups_per_cu;
    int grid = batch_grid(bt, (int)tiles, gpc);
    ostr::batch_schedule((int*)bt->host[k], cnt.data(), n, grid);
    // one group per tile: widen the grid to 8 x the largest per-XCD tile count
    // (the schedule is the same for any grid >= 8), so no group needs a second
    // tile and the kernels run their single-pass form
    *single_pass = 0;
    if (gpc == 0 && grid >= 8) {
        int mx = 0;
        for (int x = 0; x < 8; ++x) {
            const int len = ((const int*)bt->host[k])[x * (n + 1) + n];
            mx = len > mx ? len : mx;
        }
        grid = 8 * mx;
        *single_pass = 1;
    }
    unsigned char* D = bt->dev + (size_t)k * bt->slot_bytes;
    e = hipMemcpyAsync(D, bt->host[k], bt->case_off + sizeof(ostr::BatchCase) * n,
                       hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return bhip(bt, e, fn);
    *table = (ostr::BatchCase*)(D + bt->case_off);
    *sched = (const int*)D;
    *total = (int)tiles;
    *grid_out = grid;
    *slot = k;
    return OSTR_OK;
}

int batch_finish(ostr_batch* bt, int slot, hipError_t launch, hipStream_t st, const char* fn) {
    if (launch != hipSuccess) return bhip(bt, launch, fn);
    hipError_t e = hipEventRecord(bt->ev[slot], st);
    if (e != hipSuccess) return bhip(bt, e, fn);
    bt->used[slot] = true;
    bt->pos = (slot + 1) % ostr_batch::RING;
    return OSTR_OK;
}

}  // namespace

OSTR_API int ostr_batch_create(int32_t device, int32_t max_cases, ostr_batch** out) {
    if (!out) return bfail(nullptr, OSTR_E_ARG, "ostr_batch_create: NULL output");
    *out = nullptr;
    if (max_cases < 1 || max_cases > 4096)
        return bfail(nullptr, OSTR_E_ARG, "ostr_batch_create: max_cases outside 1..4096");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return bhip(nullptr, e, "ostr_batch_create");
    if (device < 0 || device >= ndev)
        return bfail(nullptr, OSTR_E_ARG, "ostr_batch_create: no device %d", device);
    ostr_batch* bt = new ostr_batch;
    bt->device = device;
    bt->max_cases = max_cases;
    const size_t sched_bytes = sizeof(int) * 16 * ((size_t)max_cases + 1);
    bt->case_off = (sched_bytes + 255) & ~(size_t)255;
    bt->slot_bytes = bt->case_off + sizeof(ostr::BatchCase) * (size_t)max_cases;
    bt->slot_bytes = (bt->slot_bytes + 255) & ~(size_t)255;
    e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&bt->n_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipMalloc((void**)&bt->dev, bt->slot_bytes * ostr_batch::RING);
    for (int k = 0; k < ostr_batch::RING && e == hipSuccess; ++k) {
        e = hipHostMalloc((void**)&bt->host[k], bt->slot_bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->ev[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        bhip(nullptr, e, "ostr_batch_create");
        batch_release(bt);
        return OSTR_E_HIP;
    }
    *out = bt;
    return OSTR_OK;
}

OSTR_API void ostr_batch_destroy(ostr_batch* bt) {
    if (!bt) return;
    (void)hipSetDevice(bt->device);
    for (int k = 0; k < ostr_batch::RING; ++k)
        if (bt->used[k]) (void)hipEventSynchronize(bt->ev[k]);
    batch_release(bt);
}

OSTR_API const char* ostr_batch_last_error(const ostr_batch* bt) {
    return bt ? bt->err.c_str() : g_err.c_str();
}

OSTR_API int ostr_batch_fproj_h(ostr_batch* bt, const ostr_batch_case* cases, int32_t n_cases,
                                int32_t prj_dtype, const ostr_tile_cfg* cfg, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ostr::BatchCase* table;
    const int* sched;
    int total, grid, slot, single, rc;
    if ((rc = batch_stage(bt, cases, n_cases, prj_dtype, 0, cfg, 0, st, "ostr_batch_fproj_h",
                          &table, &sched, &total, &grid, &slot, &single)))
        return rc;
    hipError_t e = ostr::launch_batch_project(table, sched, n_cases, total, grid,
                                              cfg->fwd_rows, prj_dtype, single, st);
    return batch_finish(bt, slot, e, st, "ostr_batch_fproj_h");
}

OSTR_API int ostr_batch_update_fused(ostr_batch* bt, const ostr_batch_case* cases,
                                     int32_t n_cases, int32_t prj_dtype, int32_t potential,
                                     const ostr_tile_cfg* cfg, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ostr::BatchCase* table;
    const int* sched;
    int total, grid, slot, single, rc;
    if ((rc = batch_stage(bt, cases, n_cases, prj_dtype, potential, cfg, 1, st,
                          "ostr_batch_update_fused", &table, &sched, &total, &grid, &slot,
                          &single)))
        return rc;
    hipError_t e = ostr::launch_batch_update(table, sched, n_cases, total, grid, cfg->upd_rows,
                                             cfg->upd_waves, potential, prj_dtype, single, st);
    return batch_finish(bt, slot, e, st, "ostr_batch_update_fused");
}

// ---- warm-start network kernels (include/ostr_cnn.h) ------------------------

#include "../../include/ostr_cnn.h"

namespace {

int cnn_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int cnn_done(hipError_t e, const char* fn) {
    if (e == hipSuccess) return OSTR_OK;
    return cnn_fail(OSTR_E_HIP, "%s: HIP error %d (%s)", fn, (int)e, hipGetErrorString(e));
}

bool dims_ok(int nz, int ny, int nx) { return nz > 0 && ny > 0 && nx > 0; }

}  // namespace

OSTR_API int32_t ostr_cnn_kpad(int32_t c) { return ostr::cnn_kpad(c); }

OSTR_API int ostr_cnn_pack(const float* w, int32_t cin, int32_t cout, void* wf, void* wd,
                           void* stream) {
    if (!w || !wf || !wd || cin <= 0 || cout <= 0)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_pack: bad argument");
    return cnn_done(ostr::launch_cnn_pack(w, cin, cout, wf, wd, (hipStream_t)stream),
                    "ostr_cnn_pack");
}

OSTR_API int ostr_cnn_conv3d(int32_t cin, int32_t cout, const void* x, int32_t xcs, int32_t xco,
                             const void* wf, const float* bias, void* y, int32_t ycs,
                             int32_t yco, const void* mask, int32_t mcs, int32_t mco,
                             int32_t relu, int32_t nz, int32_t ny, int32_t nx, void* stream) {
    if (!ostr::cnn_conv_supported(cin, cout))
        return cnn_fail(OSTR_E_UNSUPPORTED, "ostr_cnn_conv3d: shape %dx%d not compiled", cin,
                        cout);
    if (!x || !wf || !y || !dims_ok(nz, ny, nx))
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_conv3d: bad argument");
    if (xcs % 8 || xco % 8 || xco + cin > xcs || ycs % 4 || yco % 4 || yco + cout > ycs ||
        (mask && (mcs % 4 || mco % 4 || mco + cout > mcs)))
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_conv3d: channel stride/offset not aligned or "
                                    "out of range");
    if ((uintptr_t)x % 16 || (uintptr_t)bias % 16 || (uintptr_t)y % 8 || (uintptr_t)mask % 8)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_conv3d: x and bias must be 16-byte, y and mask "
                                    "8-byte aligned");
    return cnn_done(ostr::launch_cnn_conv(cin, cout, x, xcs, xco, wf, bias, y, ycs, yco, mask,
                                          mcs, mco, relu, nz, ny, nx, (hipStream_t)stream),
                    "ostr_cnn_conv3d");
}

OSTR_API size_t ostr_cnn_wgrad_workspace(int32_t cin, int32_t cout, int32_t nz, int32_t ny,
                                         int32_t nx, int32_t nsplit) {
    (void)nsplit;  // kept for the interface; the partials follow a fixed grid
    if (!ostr::cnn_wgrad_supported(cin, cout) || !dims_ok(nz, ny, nx)) return 0;
    return ostr::cnn_wgrad_work_bytes(cin, cout, nz, ny, nx);
}

OSTR_API int ostr_cnn_wgrad(int32_t cin, int32_t cout, const void* x, int32_t xcs, int32_t xco,
                            const void* dy, int32_t dcs, int32_t dco, int32_t nz, int32_t ny,
                            int32_t nx, float* dw, float* db, void* work, size_t work_bytes,
                            int32_t nsplit, void* stream) {
    if (!ostr::cnn_wgrad_supported(cin, cout))
        return cnn_fail(OSTR_E_UNSUPPORTED, "ostr_cnn_wgrad: shape %dx%d not compiled", cin,
                        cout);
    if (!x || !dy || !dw || !work || !dims_ok(nz, ny, nx) || nsplit < 1 || nsplit > 1024)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_wgrad: bad argument");
    if (xco + cin > xcs || dco + cout > dcs)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_wgrad: channel view out of range");
    if (xcs % 8 || xco % 8 || dcs % 8 || dco % 8 || (uintptr_t)x % 16 || (uintptr_t)dy % 16)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_wgrad: channel strides/offsets must be multiples "
                                    "of 8 and x, dy 16-byte aligned");
    if (work_bytes < ostr::cnn_wgrad_work_bytes(cin, cout, nz, ny, nx))
        return cnn_fail(OSTR_E_RANGE, "ostr_cnn_wgrad: workspace too small");
    return cnn_done(ostr::launch_cnn_wgrad(cin, cout, x, xcs, xco, dy, dcs, dco, nz, ny, nx, dw,
                                           db, work, (hipStream_t)stream),
                    "ostr_cnn_wgrad");
}

OSTR_API int ostr_cnn_pool2(const void* x, int32_t xcs, int32_t xco, void* y, int32_t ycs,
                            int32_t yco, int32_t C, int32_t nz, int32_t ny, int32_t nx,
                            void* stream) {
    if (!x || !y || C <= 0 || !dims_ok(nz, ny, nx) || nz % 2 || ny % 2 || nx % 2)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_pool2: bad argument (sizes must be even)");
    return cnn_done(ostr::launch_cnn_pool(x, xcs, xco, y, ycs, yco, C, nz, ny, nx,
                                          (hipStream_t)stream),
                    "ostr_cnn_pool2");
}

OSTR_API int ostr_cnn_pool2_bwd(const void* dy, int32_t dcs, int32_t dco, void* dx, int32_t xcs,
                                int32_t xco, const void* mask, int32_t mcs, int32_t mco,
                                int32_t accumulate, int32_t C, int32_t nz, int32_t ny,
                                int32_t nx, void* stream) {
    if (!dy || !dx || C <= 0 || !dims_ok(nz, ny, nx) || nz % 2 || ny % 2 || nx % 2)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_pool2_bwd: bad argument (sizes must be even)");
    return cnn_done(ostr::launch_cnn_pool_bwd(dy, dcs, dco, dx, xcs, xco, mask, mcs, mco,
                                              accumulate, C, nz, ny, nx, (hipStream_t)stream),
                    "ostr_cnn_pool2_bwd");
}

OSTR_API int ostr_cnn_up2(const void* x, int32_t xcs, int32_t xco, void* y, int32_t ycs,
                          int32_t yco, int32_t C, int32_t nz, int32_t ny, int32_t nx,
                          void* stream) {
    if (!x || !y || C <= 0 || !dims_ok(nz, ny, nx) || nz % 2 || ny % 2 || nx % 2)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_up2: bad argument (sizes must be even)");
    return cnn_done(ostr::launch_cnn_up(x, xcs, xco, y, ycs, yco, C, nz, ny, nx,
                                        (hipStream_t)stream),
                    "ostr_cnn_up2");
}

OSTR_API int ostr_cnn_up2_bwd(const void* dy, int32_t dcs, int32_t dco, void* dx, int32_t xcs,
                              int32_t xco, const void* mask, int32_t mcs, int32_t mco, int32_t C,
                              int32_t nz, int32_t ny, int32_t nx, void* stream) {
    if (!dy || !dx || C <= 0 || !dims_ok(nz, ny, nx) || nz % 2 || ny % 2 || nx % 2)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_up2_bwd: bad argument (sizes must be even)");
    return cnn_done(ostr::launch_cnn_up_bwd(dy, dcs, dco, dx, xcs, xco, mask, mcs, mco, C, nz,
                                            ny, nx, (hipStream_t)stream),
                    "ostr_cnn_up2_bwd");
}

OSTR_API int ostr_cnn_input(const float* img, int64_t n, float scale, void* x, void* stream) {
    if (!img || !x || n <= 0) return cnn_fail(OSTR_E_ARG, "ostr_cnn_input: bad argument");
    return cnn_done(ostr::launch_cnn_input(img, n, scale, x, (hipStream_t)stream),
                    "ostr_cnn_input");
}

OSTR_API int ostr_cnn_output(const void* out, int32_t ocs, const float* img, int64_t n,
                             float scale, int32_t clamp, float* pred, void* stream) {
    if (!out || !img || !pred || n <= 0 || ocs <= 0 || scale == 0.0f)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_output: bad argument");
    return cnn_done(ostr::launch_cnn_output(out, ocs, img, n, 1.0f / scale, clamp, pred,
                                            (hipStream_t)stream),
                    "ostr_cnn_output");
}

OSTR_API int ostr_cnn_output_smooth(const void* out, int32_t ocs, const float* img, int32_t nz,
                                    int32_t ny, int32_t nx, float scale, const float* taps,
                                    int32_t radius, int32_t clamp, float* pred, void* stream) {
    if (!out || !img || !pred || !taps || nz <= 0 || ny <= 0 || nx <= 0 || ocs <= 0 ||
        scale == 0.0f)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_output_smooth: bad argument");
    if (radius < 0 || radius > ostr::CNN_SMOOTH_MAX_R)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_output_smooth: radius %d outside 0..%d", radius,
                        ostr::CNN_SMOOTH_MAX_R);
    return cnn_done(ostr::launch_cnn_output_smooth(out, ocs, img, nz, ny, nx, 1.0f / scale, taps,
                                                   radius, clamp, pred, (hipStream_t)stream),
                    "ostr_cnn_output_smooth");
}

OSTR_API int64_t ostr_cnn_mse_workspace(int64_t n) { return ostr::cnn_mse_partials(n) + 1; }

OSTR_API int ostr_cnn_mse(const float* pred, const float* target, int64_t n, float gscale,
                          void* dout, int32_t dcs, double* work, int64_t work_elems,
                          double* loss, void* stream) {
    if (!pred || !target || !work || !loss || n <= 0 || (dout && (dcs <= 0 || dcs % 8)))
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_mse: bad argument");
    const int64_t np = ostr::cnn_mse_partials(n);
    if (work_elems < np + 1) return cnn_fail(OSTR_E_RANGE, "ostr_cnn_mse: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = ostr::launch_cnn_mse(pred, target, n, gscale, dout, dcs, work, work + np, st);
    if (e == hipSuccess) e = hipMemcpyAsync(loss, work + np, sizeof(double),
                                            hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return cnn_done(e, "ostr_cnn_mse");
}

OSTR_API int ostr_cnn_adam(float* p, float* m, float* v, const float* g, int64_t n, float lr,
                           float b1, float b2, float eps, int32_t step, void* stream) {
    if (!p || !m || !v || !g || n <= 0 || step < 1)
        return cnn_fail(OSTR_E_ARG, "ostr_cnn_adam: bad argument");
    return cnn_done(ostr::launch_cnn_adam(p, m, v, g, n, lr, b1, b2, eps, step,
                                          (hipStream_t)stream),
                    "ostr_cnn_adam");
}
