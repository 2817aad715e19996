// prior_device.h - 27-neighbour prior terms for one voxel (device code).
//
// For the 26 neighbours k of voxel j (z outer, then y, then x, as the
// reference loops, Tomo_OSTR.cu:227-253), with t = x_j - x_k:
//   rdd_j += w_k * omega(t),   rd_j += w_k * omega(t) * t
// w_k = 1/sqrt(di^2 + dj^2 + (rho dk)^2); a neighbour outside the volume is
// replaced by the voxel itself (t = 0, omega = 1: it still adds w_k to rdd,
// Tomo_OSTR.cu:238-241). omega: Q2 1; Lange 1/(1+|t/delta|); hyperbola
// 1/sqrt(1+(t/delta)^2). The centre term has w = 0 and is skipped.
#pragma once

#include "ostr_internal.h"

namespace ostr {

template <int POT>
__device__ __forceinline__ float prior_omega(float t, float inv_delta) {
    if (POT == OSTR_PRIOR_Q2) return 1.0f;
    float q = t * inv_delta;
    if (POT == OSTR_PRIOR_LANGE) return __builtin_amdgcn_rcpf(1.0f + fabsf(q));
    return __builtin_amdgcn_rsqf(fmaf(q, q, 1.0f));
}

// Per-neighbour factor of the class-summed form of the sliding-window kernels:
// omega(t) = |delta| r(t) with r = 1/sqrt(t^2 + delta^2) (hyperbola) or
// 1/(|delta| + |t|) (Lange), |delta| folded into the class weights (k = delta^2,
// resp. |delta|, see PriorClass); Q2 r = 1. k < 0: r = omega(t) through
// q = t/delta (FOLDED_ONLY callers are never given k < 0).
template <int POT, bool FOLDED_ONLY>
__device__ __forceinline__ float prior_r(float t, float k, float inv_delta) {
    if (POT == OSTR_PRIOR_Q2) return 1.0f;
    if (FOLDED_ONLY || k >= 0.0f) {
        if (POT == OSTR_PRIOR_LANGE) return __builtin_amdgcn_rcpf(k + fabsf(t));
        return __builtin_amdgcn_rsqf(fmaf(t, t, k));
    }
    return prior_omega<POT>(t, inv_delta);
}

// Accumulates the prior terms of voxel (ix,iy,iz) of x into rdd/rd in the
// class-summed form of the sliding-window kernels (per class c:
// a[c] = sum r(t), b[c] = sum r(t) t in neighbour order; then
// rdd += w_c a[c], rd += w_c b[c] for c = 0..4), so that from rdd = rd = 0 the
// result is bit-identical to theirs. FOLDED: W.k >= 0.
template <int POT, bool FOLDED, typename TI>
__device__ __forceinline__ void prior_voxel(const DevGeom& g, const TI* __restrict__ x, int ix,
                                            int iy, int iz, const PriorClass& W,
                                            float inv_delta, float& rdd, float& rd) {
    const int64_t plane = (int64_t)g.nx * g.ny;
    const int64_t vi = (int64_t)iz * plane + (int64_t)iy * g.nx + ix;
    const float xc = ld<TI>(x, vi);
    float a[5], b[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) a[k] = b[k] = 0.0f;
#pragma unroll
    for (int dk = -1; dk <= 1; ++dk) {
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
#pragma unroll
            for (int di = -1; di <= 1; ++di) {
                const int n = (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
                if (n == 13) continue;  // centre: weight 0
                const int kx = ix + di, ky = iy + dj, kz = iz + dk;
                const bool in = (unsigned)kx < (unsigned)g.nx && (unsigned)ky < (unsigned)g.ny &&
                                (unsigned)kz < (unsigned)g.nz;
                const float xk = in ? ld<TI>(x, vi + dk * plane + dj * g.nx + di) : xc;
                const float t = xc - xk;
                const float rt = prior_r<POT, FOLDED>(t, W.k, inv_delta);
                const int k = prior_class(n);
                a[k] = a[k] + rt;
                b[k] = fmaf(rt, t, b[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        rdd = fmaf(W.w[k], a[k], rdd);
        rd = fmaf(W.w[k], b[k], rd);
    }
}

}  // namespace ostr
