// backproject.hip - voxel-driven bilinear backprojector (SBP) and the fused
// OSTR subset update for gfx950.
//
// SBP follows backprj_gpu_manyviews_SBP (Tomo_OSTR.cu:1553-1734): the voxel
// centre is projected from the source onto the detector plane z = -detectorR;
// the bilinear taps (ilb,jlb) are chosen with the reference's half-bin rule and
// masked one by one to the detector; the interpolated value is scaled by
// Vsize_z * |S|^2/|S-V|^2 and summed over the slots of the launch. The ray
// parameter t and the tap positions use the reference's exact divisions (a
// 1-ulp change of the tap position at bin ~200 is 1.5e-5 of a bin, visible on
// rough data); only the depth-weight division is a hardware reciprocal (a
// factor within ~1 ulp).
//
// The fused update computes, per voxel and in one pass,
//   capL = M * SBP(h)                      (Tomo_OSTR_gold.cpp:759-771)
//   RD, RDD = prior(x_in)                  (:775)
//   x_out = max(0, x - (capL + b RD)/(d + 2 b RDD))   (update_est :265-295)
// reading x_in and writing x_out (double-buffered, so the prior sees the
// pre-update image as the reference's sequential order does).
#include "prior_device.h"

namespace ostr {

namespace {

constexpr int MAX_SLOTS = 64;  // views per launch handled in one pass

struct SlotTab {
    float x0[MAX_SLOTS], z0[MAX_SLOTS];
    float num[MAX_SLOTS];  // |S|^2 evaluated as fma(z0,z0,fma(y0,y0,x0*x0))
};

__device__ __forceinline__ float cell_centre(float origin, int i, float size) {
    return (float)((double)origin + ((double)i + 0.5) * (double)size);
}

// Sum over slots [0, n) of the SBP contribution for the voxel centred at
// (x1, y1, z1). prj points at slot 0 of the launch ([slot][nbx][nby]).
template <typename TP>
__device__ __forceinline__ float sbp_voxel(const DevGeom& g, const SlotTab& S, int n,
                                           const TP* __restrict__ prj, float x1, float y1,
                                           float z1, float inv_dbx, float inv_dby) {
    const int64_t nbin = (int64_t)g.nbx * g.nby;
    const float y0 = g.src_y;
    const float ey = y0 - y1;
    float total = 0.0f;
    for (int a = 0; a < n; ++a) {
        const float x0 = S.x0[a], z0 = S.z0[a];
        const float ex = x0 - x1, ez = z0 - z1;
        const float den = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
        const float dw = S.num[a] * __builtin_amdgcn_rcpf(den);
        if (z1 == z0) continue;
        const float t = (-g.det_r - z0) / (z1 - z0);
        const float x2 = fmaf(x1 - x0, t, x0);
        const float y2 = fmaf(y1 - y0, t, y0);
        const float imb = (x2 - g.x_d0) / g.dbx;
        const float jmb = (y2 - g.y_d0) / g.dby;
        int il = (int)floorf(imb);
        if (imb < (float)il + 0.5f) il -= 1;
        int jl = (int)floorf(jmb);
        if (jmb < (float)jl + 0.5f) jl -= 1;
        const float fi = imb - ((float)il + 0.5f);
        const float fj = jmb - ((float)jl + 0.5f);
        const TP* P = prj + (int64_t)a * nbin;
        const bool i0 = (unsigned)il < (unsigned)g.nbx;
        const bool i1 = (unsigned)(il + 1) < (unsigned)g.nbx;
        const bool j0 = (unsigned)jl < (unsigned)g.nby;
        const bool j1 = (unsigned)(jl + 1) < (unsigned)g.nby;
        const int64_t b00 = (int64_t)il * g.nby + jl;
        float d1 = 0.0f, d2 = 0.0f;
        if (i0 && j0) d1 = (1.0f - fi) * ld<TP>(P, b00);
        if (i1 && j0) d1 = fmaf(fi, ld<TP>(P, b00 + g.nby), d1);
        if (i0 && j1) d2 = (1.0f - fi) * ld<TP>(P, b00 + 1);
        if (i1 && j1) d2 = fmaf(fi, ld<TP>(P, b00 + g.nby + 1), d2);
        float u = fmaf(fj, d2, (1.0f - fj) * d1);
        u = (u * g.dz) * dw;
        total = total + u;
    }
    return total;
}

template <typename TP>
__global__ __launch_bounds__(256) void sbp_kernel(DevGeom g, SlotTab S, int n_slots,
                                                  const TP* __restrict__ prj,
                                                  float* __restrict__ img, float inv_dbx,
                                                  float inv_dby) {
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    if (ix >= g.nx || iy >= g.ny) return;
    const float x1 = cell_centre(g.x_p0, ix, g.dx);
    const float y1 = cell_centre(g.y_p0, iy, g.dy);
    const float z1 = cell_centre(g.z_p0, iz, g.dz);
    const float tot = sbp_voxel<TP>(g, S, n_slots, prj, x1, y1, z1, inv_dbx, inv_dby);
    const int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
    img[vi] = img[vi] + tot;
}

template <int POT, typename TI, typename TP>
__global__ __launch_bounds__(256) void update_fused_kernel(
    DevGeom g, SlotTab S, int n_slots, const TP* __restrict__ hsub, const TI* __restrict__ x_in,
    TI* __restrict__ x_out, const float* __restrict__ d, PriorW W, float inv_delta, float beta,
    float capl_scale, int guard, float inv_dbx, float inv_dby) {
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    if (ix >= g.nx || iy >= g.ny) return;
    const float x1 = cell_centre(g.x_p0, ix, g.dx);
    const float y1 = cell_centre(g.y_p0, iy, g.dy);
    const float z1 = cell_centre(g.z_p0, iz, g.dz);
    const int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
    // capL for this voxel: 0 + total, then * M (float), as the reference host does
    const float capl = capl_scale * (0.0f + sbp_voxel<TP>(g, S, n_slots, hsub, x1, y1, z1,
                                                         inv_dbx, inv_dby));
    float rdd = 0.0f, rd = 0.0f;
    prior_voxel<POT, TI>(g, x_in, ix, iy, iz, W, inv_delta, rdd, rd);
    const float xv = ld<TI>(x_in, vi);
    const float num = capl + beta * rd;
    const float den = d[vi] + (2.0f * beta) * rdd;
    float xn;
    if (guard && den == 0.0f) {
        xn = xv;
    } else {
        xn = xv - num / den;
        xn = xn < 0.0f ? 0.0f : xn;
    }
    st<TI>(x_out, vi, xn);
}

bool fill_slots(SlotTab& S, const ViewTab* host_views, int a0, int a1, float src_y) {
    if (a1 - a0 > MAX_SLOTS) return false;
    for (int a = a0; a < a1; ++a) {
        const ViewTab& v = host_views[a];
        S.x0[a - a0] = v.x0;
        S.z0[a - a0] = v.z0;
        S.num[a - a0] = std::fma(v.z0, v.z0, std::fma(src_y, src_y, v.x0 * v.x0));
    }
    return true;
}

template <typename TP>
hipError_t sbp_t(const DevGeom& g, const SlotTab& S, int n, float* img, const void* prj,
                 hipStream_t st) {
    dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL((sbp_kernel<TP>), grid, dim3(256), 0, st, g, S, n, (const TP*)prj, img,
                       1.0f / g.dbx, 1.0f / g.dby);
    return hipGetLastError();
}

template <int POT, typename TI, typename TP>
hipError_t fused_t(const DevGeom& g, const SlotTab& S, int n, const void* x_in, void* x_out,
                   const float* d, const void* hsub, const PriorW& W, float inv_delta,
                   float beta, float capl_scale, int guard, hipStream_t st) {
    dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL((update_fused_kernel<POT, TI, TP>), grid, dim3(256), 0, st, g, S, n,
                       (const TP*)hsub, (const TI*)x_in, (TI*)x_out, d, W, inv_delta, beta,
                       capl_scale, guard, 1.0f / g.dbx, 1.0f / g.dby);
    return hipGetLastError();
}

template <int POT, typename TI>
hipError_t fused_p(int prj_dtype, const DevGeom& g, const SlotTab& S, int n, const void* x_in,
                   void* x_out, const float* d, const void* hsub, const PriorW& W,
                   float inv_delta, float beta, float capl_scale, int guard, hipStream_t st) {
    switch (prj_dtype) {
        case OSTR_F32:
            return fused_t<POT, TI, float>(g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
        case OSTR_BF16:
            return fused_t<POT, TI, bf16_t>(g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
        case OSTR_F16:
            return fused_t<POT, TI, _Float16>(g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
    }
    return hipErrorInvalidValue;
}

template <int POT>
hipError_t fused_i(int img_dtype, int prj_dtype, const DevGeom& g, const SlotTab& S, int n,
                   const void* x_in, void* x_out, const float* d, const void* hsub,
                   const PriorW& W, float inv_delta, float beta, float capl_scale, int guard,
                   hipStream_t st) {
    switch (img_dtype) {
        case OSTR_F32:
            return fused_p<POT, float>(prj_dtype, g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
        case OSTR_BF16:
            return fused_p<POT, bf16_t>(prj_dtype, g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
        case OSTR_F16:
            return fused_p<POT, _Float16>(prj_dtype, g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

// `views` here is the HOST copy of the slot table (the slot parameters go to
// the kernel by value).
hipError_t launch_sbp(const DevGeom& g, const ViewTab* views, int a0, int a1,
                      float* img_accum, const void* prj, int prj_dtype, hipStream_t st) {
    for (int b = a0; b < a1; b += MAX_SLOTS) {
        int e = b + MAX_SLOTS < a1 ? b + MAX_SLOTS : a1;
        SlotTab S;
        fill_slots(S, views, b, e, g.src_y);
        const int64_t off = (int64_t)(b - a0) * g.nbx * g.nby;
        hipError_t err;
        switch (prj_dtype) {
            case OSTR_F32:
                err = sbp_t<float>(g, S, e - b, img_accum, (const float*)prj + off, st);
                break;
            case OSTR_BF16:
                err = sbp_t<bf16_t>(g, S, e - b, img_accum, (const bf16_t*)prj + off, st);
                break;
            case OSTR_F16:
                err = sbp_t<_Float16>(g, S, e - b, img_accum, (const _Float16*)prj + off, st);
                break;
            default:
                return hipErrorInvalidValue;
        }
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

hipError_t launch_update_fused(const DevGeom& g, const ViewTab* views, int a0, int a1,
                               const void* x_in, void* x_out, int img_dtype, const float* d,
                               const void* hsub, int prj_dtype, float beta, float capl_scale,
                               int potential, float delta, float rho, int guard,
                               hipStream_t st) {
    SlotTab S;
    if (!fill_slots(S, views, a0, a1, g.src_y)) return hipErrorInvalidValue;
    PriorW W;
    prior_weights(W.w, rho);
    float inv_delta = potential == OSTR_PRIOR_Q2 ? 0.0f : 1.0f / delta;
    const int n = a1 - a0;
    switch (potential) {
        case OSTR_PRIOR_Q2:
            return fused_i<OSTR_PRIOR_Q2>(img_dtype, prj_dtype, g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
        case OSTR_PRIOR_LANGE:
            return fused_i<OSTR_PRIOR_LANGE>(img_dtype, prj_dtype, g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
        case OSTR_PRIOR_HYPERBOLA:
            return fused_i<OSTR_PRIOR_HYPERBOLA>(img_dtype, prj_dtype, g, S, n, x_in, x_out, d, hsub, W, inv_delta, beta, capl_scale, guard, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace ostr
