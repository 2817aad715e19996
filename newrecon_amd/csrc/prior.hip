// prior.hip - standalone prior gradient/curvature kernel (accumulating),
// the counterpart of prior_GPU_OSTR / prior_GPU_OSTR_Q2 (Tomo_OSTR.cu:
// 2189-2245; kernels G_Q2_prior :118-184, G_Fessler_prior :187-255, and the
// hyperbola potential of G_Huber_prior :258-325).
//
// One lane per voxel, 64 lanes along x; neighbour reads go through L1/L2.
#include <cmath>

#include "prior_device.h"

namespace ostr {

void prior_weights(float w[27], float rho) {
    int n = 0;
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                float dist = std::sqrt((float)(di * di + dj * dj) + ((float)(dk * dk) * rho) * rho);
                w[n++] = dist == 0.0f ? 0.0f : (float)(1.0 / (double)dist);
            }
}

namespace {

template <int POT, typename TI>
__global__ __launch_bounds__(256) void prior_kernel(DevGeom g, PriorClass W, float inv_delta,
                                                    const TI* __restrict__ x,
                                                    float* __restrict__ rdd,
                                                    float* __restrict__ rd) {
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    if (ix >= g.nx || iy >= g.ny) return;
    const int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
    float a = rdd[vi], b = rd[vi];
    if (W.k >= 0.0f)
        prior_voxel<POT, true, TI>(g, x, ix, iy, iz, W, inv_delta, a, b);
    else
        prior_voxel<POT, false, TI>(g, x, ix, iy, iz, W, inv_delta, a, b);
    rdd[vi] = a;
    rd[vi] = b;
}

template <typename TI>
hipError_t launch_t(const DevGeom& g, float* rdd, float* rd, const void* img, int pot,
                    const PriorClass& W, float inv_delta, hipStream_t st) {
    dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    switch (pot) {
        case OSTR_PRIOR_Q2:
            hipLaunchKernelGGL((prior_kernel<OSTR_PRIOR_Q2, TI>), grid, dim3(256), 0, st, g, W,
                               inv_delta, (const TI*)img, rdd, rd);
            break;
        case OSTR_PRIOR_LANGE:
            hipLaunchKernelGGL((prior_kernel<OSTR_PRIOR_LANGE, TI>), grid, dim3(256), 0, st, g,
                               W, inv_delta, (const TI*)img, rdd, rd);
            break;
        case OSTR_PRIOR_HYPERBOLA:
            hipLaunchKernelGGL((prior_kernel<OSTR_PRIOR_HYPERBOLA, TI>), grid, dim3(256), 0, st,
                               g, W, inv_delta, (const TI*)img, rdd, rd);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_prior(const DevGeom& g, float* rdd, float* rd, const void* img,
                        int img_dtype, int potential, float delta, float rho, hipStream_t st) {
    PriorW W27;
    prior_weights(W27.w, rho);
    const PriorClass W = prior_params(W27, potential, delta);
    float inv_delta = potential == OSTR_PRIOR_Q2 ? 0.0f : 1.0f / delta;
    switch (img_dtype) {
        case OSTR_F32:
            return launch_t<float>(g, rdd, rd, img, potential, W, inv_delta, st);
        case OSTR_BF16:
            return launch_t<bf16_t>(g, rdd, rd, img, potential, W, inv_delta, st);
        case OSTR_F16:
            return launch_t<_Float16>(g, rdd, rd, img, potential, W, inv_delta, st);
    }
    return hipErrorInvalidValue;
}

namespace {

struct PriorW64 {
    double w[27];
};

// psi(t) of the three potentials in FP64; the hyperbola uses the cancellation-
// free form delta^2 q^2 / (sqrt(1 + q^2) + 1) of delta^2 (sqrt(1 + q^2) - 1).
template <int POT>
__device__ __forceinline__ double psi64(double t, double delta) {
    if (POT == OSTR_PRIOR_Q2) return 0.5 * t * t;
    const double q = fabs(t / delta);
    if (POT == OSTR_PRIOR_LANGE) return delta * delta * (q - log1p(q));
    return delta * delta * (q * q / (sqrt(1.0 + q * q) + 1.0));
}

// R(x) = 1/2 sum_j sum_{k in N(j), inside the volume} w_jk psi(x_j - x_k),
// one partial per work-group (the 1/2 is applied by the caller).
template <int POT, typename TI>
__global__ __launch_bounds__(256) void penalty_kernel(DevGeom g, PriorW64 W, double delta,
                                                      const TI* __restrict__ x,
                                                      double* __restrict__ partials) {
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    double acc = 0.0;
    if (ix < g.nx && iy < g.ny) {
        const int64_t plane = (int64_t)g.nx * g.ny;
        const int64_t vi = iz * plane + (int64_t)iy * g.nx + ix;
        const double xc = (double)ld<TI>(x, vi);
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    const int n = (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
                    if (n == 13) continue;
                    const int kx = ix + di, ky = iy + dj, kz = iz + dk;
                    if ((unsigned)kx >= (unsigned)g.nx || (unsigned)ky >= (unsigned)g.ny ||
                        (unsigned)kz >= (unsigned)g.nz)
                        continue;
                    const double t = xc - (double)ld<TI>(x, vi + dk * plane + dj * g.nx + di);
                    acc += W.w[n] * psi64<POT>(t, delta);
                }
    }
    const double sum = block_sum_256(acc);
    if (threadIdx.x == 0)
        partials[((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = sum;
}

template <typename TI>
hipError_t penalty_t(const DevGeom& g, int pot, const PriorW64& W, double delta,
                     const void* img, double* partials, hipStream_t st) {
    dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    switch (pot) {
        case OSTR_PRIOR_Q2:
            hipLaunchKernelGGL((penalty_kernel<OSTR_PRIOR_Q2, TI>), grid, dim3(256), 0, st, g, W,
                               delta, (const TI*)img, partials);
            break;
        case OSTR_PRIOR_LANGE:
            hipLaunchKernelGGL((penalty_kernel<OSTR_PRIOR_LANGE, TI>), grid, dim3(256), 0, st, g,
                               W, delta, (const TI*)img, partials);
            break;
        case OSTR_PRIOR_HYPERBOLA:
            hipLaunchKernelGGL((penalty_kernel<OSTR_PRIOR_HYPERBOLA, TI>), grid, dim3(256), 0, st,
                               g, W, delta, (const TI*)img, partials);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

int64_t penalty_partials(const DevGeom& g) {
    return (int64_t)((g.nx + 63) / 64) * ((g.ny + 3) / 4) * g.nz;
}

hipError_t launch_penalty(const DevGeom& g, const void* img, int img_dtype, int potential,
                          double delta, double rho, double* partials, int64_t* n_partials,
                          hipStream_t st) {
    PriorW64 W;
    int n = 0;
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                double dist = std::sqrt((double)(di * di + dj * dj) + (dk * rho) * (dk * rho));
                W.w[n++] = dist == 0.0 ? 0.0 : 1.0 / dist;
            }
    *n_partials = penalty_partials(g);
    switch (img_dtype) {
        case OSTR_F32:
            return penalty_t<float>(g, potential, W, delta, img, partials, st);
        case OSTR_BF16:
            return penalty_t<bf16_t>(g, potential, W, delta, img, partials, st);
        case OSTR_F16:
            return penalty_t<_Float16>(g, potential, W, delta, img, partials, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace ostr
