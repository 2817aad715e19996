// ostr_internal.h - types and launchers shared by the HIP translation units of
// libostr_hip.so. Not part of the public C ABI (see include/ostr_hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ostr_hip.h"

namespace ostr {

// Geometry as kernel argument, plus products the walker needs.
struct DevGeom {
    int nx, ny, nz, nbx, nby;
    float dx, dy, dz, x_p0, y_p0, z_p0;
    float dbx, dby, x_d0, y_d0, det_r, src_r, src_y;
    float ext_x, ext_y, ext_z;   // n*h  (float product, as int*float in the reference)
    float top_x, top_y, top_z;   // fma(n, h, p0): upper face of the volume
    float hp_x, hp_y, hp_z;      // h + p0: upper face of voxel 0
    float last_x, last_y, last_z;// fma(n-1, h, p0): lower face of the last voxel
};

// Per-slot view table entry: source x, source z, cos(theta), sin(theta).
struct ViewTab {
    float x0, z0, c, s;
};

DevGeom make_dev_geom(const ostr_geom& g);

// ---- storage type helpers ------------------------------------------------
struct bf16_t {
    uint16_t bits;
};

template <typename T>
__device__ __forceinline__ float ld(const T* p, int64_t i);
template <>
__device__ __forceinline__ float ld<float>(const float* p, int64_t i) {
    return p[i];
}
template <>
__device__ __forceinline__ float ld<bf16_t>(const bf16_t* p, int64_t i) {
    return __uint_as_float((uint32_t)p[i].bits << 16);
}
template <>
__device__ __forceinline__ float ld<_Float16>(const _Float16* p, int64_t i) {
    return (float)p[i];
}

template <typename T>
__device__ __forceinline__ void st(T* p, int64_t i, float v);
template <>
__device__ __forceinline__ void st<float>(float* p, int64_t i, float v) {
    p[i] = v;
}
template <>
__device__ __forceinline__ void st<bf16_t>(bf16_t* p, int64_t i, float v) {
    uint32_t u = __float_as_uint(v);
    uint16_t b;
    if ((u & 0x7fffffffu) > 0x7f800000u)
        b = 0x7fc0;  // quiet NaN
    else
        b = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
    p[i].bits = b;
}
template <>
__device__ __forceinline__ void st<_Float16>(_Float16* p, int64_t i, float v) {
    p[i] = (_Float16)v;
}

// Volume gathers through a raw buffer resource: 32-bit byte offsets (no 64-bit
// address arithmetic per gather) and hardware range checking against the
// volume size. Word 3 0x00020000 is the raw-buffer descriptor of the gfx9
// family (gfx90a/gfx94x/gfx950).
template <typename T>
struct VolReader {
    __amdgpu_buffer_rsrc_t rsrc;
    __device__ VolReader(const T* p, uint32_t n) {
        rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * sizeof(T)),
                                                 0x00020000);
    }
    __device__ __forceinline__ float operator()(uint32_t i) const;
};
template <>
__device__ __forceinline__ float VolReader<float>::operator()(uint32_t i) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, i * 4u, 0, 0));
}
template <>
__device__ __forceinline__ float VolReader<bf16_t>::operator()(uint32_t i) const {
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rsrc, i * 2u, 0, 0)
                           << 16);
}
template <>
__device__ __forceinline__ float VolReader<_Float16>::operator()(uint32_t i) const {
    unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(rsrc, i * 2u, 0, 0);
    return (float)__builtin_bit_cast(_Float16, b);
}

// Sum of one double per thread over a 256-thread work-group in a fixed order
// (wave tree, then the 4 wave sums left to right); the result is valid in
// thread 0. Deterministic: no atomics.
__device__ __forceinline__ double block_sum_256(double v) {
    __shared__ double part[4];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = ((part[0] + part[1]) + part[2]) + part[3];
    return s;
}

// ---- launchers (return hipError_t of the launch) -------------------------
// Forward family. kind: 0 = projection, 1 = gamma ray length, 2 = gamma row
// sum (reference layout out), 3 = projection + h epilogue (row-major y/b/r/h),
// 4 / 5 = gamma ray length / row sum times yry (row-major out).
hipError_t launch_project(const DevGeom& g, const ViewTab* views, int a0, int a1, int kind,
                          int cos_weight, const void* img, int img_dtype, void* out,
                          int out_dtype, const float* y, const float* b, const float* r,
                          hipStream_t st);
hipError_t launch_count_nnz(const DevGeom& g, const ViewTab* views, int a0, int a1,
                            unsigned long long* counter, hipStream_t st);

// Backprojection family. prj / hsub are detector-row-major [slot][nby][nbx];
// `views` is the host copy of the slot table.
hipError_t launch_sbp_rows(const DevGeom& g, const ViewTab* views, int a0, int a1,
                           float* img_accum, const void* prj, int prj_dtype, hipStream_t st);
// Exact transpose of the forward projector (voxel-driven ray/box clipping),
// ACCUMULATED into img; prj detector-row-major.
hipError_t launch_matched_rows(const DevGeom& g, const ViewTab* views, int a0, int a1,
                               int cos_weight, float* img_accum, const void* prj,
                               int prj_dtype, hipStream_t st);
// [slot][nbx][nby] -> [slot][nby][nbx] (dtype selects 4- or 2-byte elements).
hipError_t launch_transpose_bins(const void* src, void* dst, int n_slots, int nbx, int nby,
                                 int dtype, hipStream_t st);
size_t elem_size(int dtype);
hipError_t launch_update_fused(const DevGeom& g, const ViewTab* views, int a0, int a1,
                               const void* x_in, void* x_out, int img_dtype, const float* d,
                               const void* hsub, int prj_dtype, float beta, float capl_scale,
                               int potential, float delta, float rho, int guard,
                               hipStream_t st);

// Prior.
hipError_t launch_prior(const DevGeom& g, float* rdd, float* rd, const void* img,
                        int img_dtype, int potential, float delta, float rho, hipStream_t st);

// Elementwise.
hipError_t launch_yry(float* out, const float* y, const float* r, int64_t n, hipStream_t st);
hipError_t launch_mul(float* out, const float* a, const float* b, int64_t n, hipStream_t st);
hipError_t launch_compute_h(float* h, const float* y, const float* b, const float* l,
                            const float* r, int64_t n, hipStream_t st);
hipError_t launch_update_est(float* x, const float* capl, const float* rd, const float* d,
                             const float* rdd, float beta, float capl_scale, int guard,
                             int64_t n, hipStream_t st);
hipError_t launch_regroup(float* dst, const float* src, int n_views, int n_subsets,
                          int64_t nbin, hipStream_t st);
hipError_t launch_yry_gamma(float* gc, const float* y, const float* r, int64_t n,
                            hipStream_t st);
// BP.m initial image pieces (one view at a time).
hipError_t launch_bp_data(float* q_rows, const float* l, const float* py, const float* b,
                          int counts, float c, float thr, int nbx, int nby, hipStream_t st);
hipError_t launch_bp_accum(float* img, const float* num, const float* den, int64_t n,
                           hipStream_t st);
hipError_t launch_div(float* p, float v, int64_t n, hipStream_t st);
// Zero/constant fill by a kernel (no hipMemsetAsync, so callers may capture it).
hipError_t launch_fill(float* p, float v, int64_t n, hipStream_t st);

// FP64 objective pieces. Each writes one double per work-group into
// `partials` (its count is returned through *n_partials); reduce_partials
// folds them in a fixed order into *out.
hipError_t launch_objective_data(const DevGeom& g, const ViewTab* views, int n_views,
                                 const void* img, int img_dtype, const float* y,
                                 const float* b, const float* r, double* partials,
                                 int64_t* n_partials, hipStream_t st);
hipError_t launch_penalty(const DevGeom& g, const void* img, int img_dtype, int potential,
                          double delta, double rho, double* partials, int64_t* n_partials,
                          hipStream_t st);
hipError_t launch_diff_norms(const void* a, const void* b, int dtype, int64_t n,
                             double* partials, int64_t* n_partials, hipStream_t st);
hipError_t launch_reduce_partials(const double* partials, int64_t n, double* out,
                                  hipStream_t st);
int64_t objective_data_partials(const DevGeom& g, int n_views);
int64_t penalty_partials(const DevGeom& g);
int64_t diff_norms_partials(int64_t n);

// Prior neighbour weights (host): w[27] in z,y,x order, centre 0.
void prior_weights(float w[27], float rho);

// Work-group to tile mapping that keeps consecutive tiles on one XCD: the
// hardware deals work-groups round-robin over the 8 XCDs, so work-group b
// runs on XCD b % 8; giving XCD k the contiguous tile range
// [k*per, (k+1)*per) keeps neighbouring tiles (and the volume region they
// read) in one L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    const int nxcd = 8;
    int per = (nblocks + nxcd - 1) / nxcd;
    int xcd = b % nxcd, slot = b / nxcd;
    // the last XCDs may own fewer tiles when nblocks % 8 != 0
    int full = nblocks - (per - 1) * nxcd;  // XCDs owning `per` tiles
    int t = xcd < full ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
    return t;
}

}  // namespace ostr
