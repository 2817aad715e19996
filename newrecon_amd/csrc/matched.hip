// matched.hip - exact transpose of the Siddon forward projector, voxel-driven.
//
// No reference counterpart: the reference backprojects with the unmatched
// bilinear SBP (Tomo_OSTR.cu:1553-1734). This kernel accumulates, for every
// voxel j, x_j += sum_views cos(t) sum_i a_ij p_i (matched_voxel in
// matched_device.h: the candidate bins of the voxel's shadow, each ray clipped
// against the voxel box).
#include "matched_device.h"

namespace ostr {

namespace {

// One voxel per lane: 64 lanes on consecutive voxels along x, one row per
// wave (4 rows per work-group), one slice per grid z. Deterministic: every
// voxel is owned by one lane, no atomics.
template <typename TP>
__global__ __launch_bounds__(256) void matched_kernel(DevGeom g, MatchedSlots S, int n_slots,
                                                      int cos_weight,
                                                      const TP* __restrict__ prj,
                                                      float* __restrict__ img) {
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    if (ix >= g.nx || iy >= g.ny) return;
    const VolReader<TP> P(prj, (uint32_t)n_slots * g.nbx * g.nby);
    const float total = matched_voxel<TP>(g, S, n_slots, cos_weight, P, ix, iy, iz);
    const int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
    img[vi] = img[vi] + total;
}

template <typename TP>
hipError_t matched_t(const DevGeom& g, const MatchedSlots& S, int n, int cos_weight,
                     const void* prj, float* img, hipStream_t st) {
    dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL((matched_kernel<TP>), grid, dim3(256), 0, st, g, S, n, cos_weight,
                       (const TP*)prj, img);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_matched_rows(const DevGeom& g, const ViewTab* views, int a0, int a1,
                               int cos_weight, float* img_accum, const void* prj,
                               int prj_dtype, hipStream_t st) {
    for (int b = a0; b < a1; b += MAX_SLOTS) {
        const int e = b + MAX_SLOTS < a1 ? b + MAX_SLOTS : a1;
        MatchedSlots S;
        for (int a = b; a < e; ++a) {
            S.x0[a - b] = views[a].x0;
            S.z0[a - b] = views[a].z0;
            S.c[a - b] = views[a].c;
        }
        const size_t off = (size_t)(b - a0) * g.nbx * g.nby * elem_size(prj_dtype);
        const void* p = (const char*)prj + off;
        hipError_t err;
        switch (prj_dtype) {
            case OSTR_F32:
                err = matched_t<float>(g, S, e - b, cos_weight, p, img_accum, st);
                break;
            case OSTR_BF16:
                err = matched_t<bf16_t>(g, S, e - b, cos_weight, p, img_accum, st);
                break;
            case OSTR_F16:
                err = matched_t<_Float16>(g, S, e - b, cos_weight, p, img_accum, st);
                break;
            default:
                return hipErrorInvalidValue;
        }
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

}  // namespace ostr
