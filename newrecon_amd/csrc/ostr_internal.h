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

// Per-slot view table entry: source x, source z, cos(theta), sin(theta), and
// the detector bins [i0, i1] x [j0, j1] whose rays may meet the volume (the
// bounding box of its 8 corners projected from the source onto the detector
// plane, widened by 2 bins; rays to bins outside it miss the volume, so the
// ray kernels skip their walk there). An empty box (i0 > i1) is allowed.
struct ViewTab {
    float x0, z0, c, s;
    int i0, i1, j0, j1;
};
__device__ __forceinline__ bool view_may_hit(const ViewTab& v, int i, int j) {
    return i >= v.i0 && i <= v.i1 && j >= v.j0 && j <= v.j1;
}

constexpr int MAX_SLOTS = 64;  // views per backprojection pass

// Slot parameters of the backprojector for the slots of one launch.
struct SlotTab {
    float x0[MAX_SLOTS], z0[MAX_SLOTS];
    float num[MAX_SLOTS];  // |S|^2 evaluated as fma(z0,z0,fma(y0,y0,x0*x0))
};

// 27-neighbour prior weights, index dz*9 + dy*3 + dx (centre 13 is 0).
struct PriorW {
    float w[27];
};

// The same weights by class: with dx == dy a weight depends only on whether
// the neighbour is in the voxel's slice and on how many of its x, y offsets
// are non-zero, and prior_weights computes every member of a class from the
// same operands, so the 5 class values are the 27 weights bit for bit.
// Class of index dz*9+dy*3+dx: in-slice (dz = 1) m-1, else 2+m, where m is the
// number of non-zero x, y offsets.
// k: parameter of the folded form of the sliding-window kernels (prior_r):
// delta^2 (hyperbola) or |delta| (Lange) with |delta| folded into w, 0 for Q2,
// -1 for a delta outside [2^-60, 2^60] (then w are the plain weights and the
// kernels evaluate omega through q = t / delta).
struct PriorClass {
    float w[5];
    float k;
};
__host__ __device__ constexpr int prior_class(int nidx) {
    return (nidx / 9 == 1) ? ((nidx % 3 != 1) + ((nidx / 3) % 3 != 1)) - 1
                           : 2 + ((nidx % 3 != 1) + ((nidx / 3) % 3 != 1));
}
inline PriorClass prior_classes(const PriorW& W) {
    // representatives: in-slice edge 10, in-slice diagonal 9; other slice
    // centre 4, edge 1, corner 0
    return PriorClass{{W.w[10], W.w[9], W.w[4], W.w[1], W.w[0]}, 0.0f};
}
// Class weights and k for a potential and delta (see PriorClass).
inline PriorClass prior_params(const PriorW& W, int potential, float delta) {
    PriorClass P = prior_classes(W);
    if (potential == OSTR_PRIOR_Q2) return P;
    const double ad = delta < 0.0f ? -(double)delta : (double)delta;
    if (!(ad >= 0x1p-60 && ad <= 0x1p60)) {
        P.k = -1.0f;
        return P;
    }
    for (int c = 0; c < 5; ++c) P.w[c] = (float)((double)P.w[c] * ad);
    P.k = potential == OSTR_PRIOR_LANGE ? (float)ad : (float)(ad * ad);
    return P;
}

// One case of a batched launch (device-resident table, rebuilt per call).
struct BatchCase {
    DevGeom g;
    int n_slots;
    int guard;
    float beta, capl_scale, inv_delta;
    int pad_;
    const void* x_in;
    void* x_out;
    const float* d;
    const float* y;
    const float* b;
    const float* r;
    void* h;
    PriorClass W;
    ViewTab v[MAX_SLOTS];  // slot table for the forward walk
    SlotTab S;             // slot table for the backprojector
};

// Wave-uniform copy of a struct of 4-byte words from memory: each word is
// loaded and taken with readfirstlane, so the copy lives in scalar registers
// (as kernel arguments do) and the compiler knows it is uniform.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T& src) {
    static_assert(sizeof(T) % 4 == 0, "struct of 4-byte words");
    T out;
    const int* s = reinterpret_cast<const int*>(&src);
    int* d = reinterpret_cast<int*>(&out);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) d[k] = __builtin_amdgcn_readfirstlane(s[k]);
    return out;
}

template <typename P>
__device__ __forceinline__ P* uniform_ptr(P* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (P*)(((uint64_t)hi << 32) | lo);
}

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
// float -> bf16 bits, round to nearest even, NaN -> quiet NaN
__device__ __forceinline__ uint16_t bf16_bits(float v) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
template <>
__device__ __forceinline__ void st<bf16_t>(bf16_t* p, int64_t i, float v) {
    p[i].bits = bf16_bits(v);
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
    __device__ __forceinline__ float operator()(uint32_t i) const { return at_byte(i * sizeof(T)); }
    // the element at byte offset b (a multiple of sizeof(T))
    __device__ __forceinline__ float at_byte(uint32_t b) const;
    // elements e and e + 1 from one 8-byte load: for 4-byte types at e, for
    // 2-byte types from the 4-byte boundary at or below e, so when e is odd
    // element e + 2 must lie inside the buffer (the caller guarantees it)
    __device__ __forceinline__ void pair(uint32_t e, float& a, float& b) const;
    // pair() of a 4-byte type at byte offset vb + sb, with sb wave-uniform (it
    // goes in the load's scalar offset, so the lane's address costs nothing
    // per row)
    __device__ __forceinline__ void pair_bytes(uint32_t vb, uint32_t sb, float& a,
                                               float& b) const;
};
template <>
__device__ __forceinline__ float VolReader<float>::at_byte(uint32_t b) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, b, 0, 0));
}
template <>
__device__ __forceinline__ float VolReader<bf16_t>::at_byte(uint32_t b) const {
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rsrc, b, 0, 0) << 16);
}
template <>
__device__ __forceinline__ float VolReader<_Float16>::at_byte(uint32_t b) const {
    unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(rsrc, b, 0, 0);
    return (float)__builtin_bit_cast(_Float16, h);
}
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
// the 32 bits holding elements e (low half) and e + 1 (high half) of a
// 2-byte array, from one aligned 8-byte load and a funnel shift
__device__ __forceinline__ uint32_t pair16(__amdgpu_buffer_rsrc_t rsrc, uint32_t e) {
    const u32x2_t w = __builtin_amdgcn_raw_buffer_load_b64(rsrc, (e & ~1u) * 2u, 0, 0);
    return __builtin_amdgcn_alignbit(w.y, w.x, (e & 1u) * 16u);
}
template <>
__device__ __forceinline__ void VolReader<float>::pair(uint32_t e, float& a, float& b) const {
    const u32x2_t w = __builtin_amdgcn_raw_buffer_load_b64(rsrc, e * 4u, 0, 0);
    a = __uint_as_float(w.x);
    b = __uint_as_float(w.y);
}
template <>
__device__ __forceinline__ void VolReader<float>::pair_bytes(uint32_t vb, uint32_t sb, float& a,
                                                             float& b) const {
    const u32x2_t w = __builtin_amdgcn_raw_buffer_load_b64(rsrc, vb, sb, 0);
    a = __uint_as_float(w.x);
    b = __uint_as_float(w.y);
}
template <>
__device__ __forceinline__ void VolReader<bf16_t>::pair(uint32_t e, float& a, float& b) const {
    const uint32_t p = pair16(rsrc, e);
    a = __uint_as_float(p << 16);
    b = __uint_as_float(p & 0xFFFF0000u);
}
template <>
__device__ __forceinline__ void VolReader<_Float16>::pair(uint32_t e, float& a, float& b) const {
    const uint32_t p = pair16(rsrc, e);
    a = (float)__builtin_bit_cast(_Float16, (unsigned short)(p & 0xFFFFu));
    b = (float)__builtin_bit_cast(_Float16, (unsigned short)(p >> 16));
}

// Volume stores through a raw buffer resource (vector buffer stores, 32-bit
// byte offsets), converting from float like st<T>.
template <typename T>
struct VolWriter {
    __amdgpu_buffer_rsrc_t rsrc;
    __device__ VolWriter(T* p, uint32_t n) {
        rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * sizeof(T)),
                                                 0x00020000);
    }
    __device__ __forceinline__ void operator()(uint32_t i, float v) const;
};
template <>
__device__ __forceinline__ void VolWriter<float>::operator()(uint32_t i, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rsrc, i * 4u, 0, 0);
}
template <>
__device__ __forceinline__ void VolWriter<bf16_t>::operator()(uint32_t i, float v) const {
    __builtin_amdgcn_raw_buffer_store_b16(bf16_bits(v), rsrc, i * 2u, 0, 0);
}
template <>
__device__ __forceinline__ void VolWriter<_Float16>::operator()(uint32_t i, float v) const {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)v), rsrc,
                                          i * 2u, 0, 0);
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

// Batched launches over a device table of `n_cases` BatchCase entries with
// `total` tiles; `grid` work-groups (one per tile, or a persistent grid).
// Forward: projection + h (kind 3), FP32 image, `rows` detector rows per group
// (1, 2, 4, 8). Update: FP32 image, `rows` x `waves` shape (see
// batch_update_shape_ok).
int64_t batch_fwd_tiles(const DevGeom& g, int n_slots, int rows);
int64_t batch_upd_tiles(const DevGeom& g, int rows, int waves);
bool batch_update_shape_ok(int rows, int waves);
// Fills the backprojector and walker slot tables for slots [a0, a1) of a
// handle's host view table (a1 - a0 <= MAX_SLOTS).
void batch_slots(SlotTab& S, ViewTab* v, const ViewTab* host_views, int a0, int a1,
                 float src_y);
// single_pass: the grid gives every XCD at least as many groups as it has
// tiles (each group takes at most one); otherwise the groups loop.
hipError_t launch_batch_project(const BatchCase* cases, const int* sched, int n_cases, int total, int grid,
                                int rows, int prj_dtype, int single_pass, hipStream_t st);
hipError_t launch_batch_update(const BatchCase* cases, const int* sched, int n_cases, int total, int grid,
                               int rows, int waves, int potential, int prj_dtype,
                               int single_pass, hipStream_t st);

// Warm-start network (cnn.hip). Channels-last BF16 views (pointer, channel
// stride, channel offset).
bool cnn_conv_supported(int cin, int cout);
bool cnn_wgrad_supported(int cin, int cout);
int cnn_kpad(int c);
size_t cnn_wgrad_work_bytes(int cin, int cout, int nz, int ny, int nx);
int64_t cnn_mse_partials(int64_t n);
hipError_t launch_cnn_conv(int cin, int cout, const void* x, int xcs, int xco, const void* w,
                           const float* bias, void* y, int ycs, int yco, const void* mask,
                           int mcs, int mco, int relu, int nz, int ny, int nx, hipStream_t st);
hipError_t launch_cnn_wgrad(int cin, int cout, const void* x, int xcs, int xco, const void* dy,
                            int dcs, int dco, int nz, int ny, int nx, float* dw, float* db,
                            void* work, hipStream_t st);
hipError_t launch_cnn_pool(const void* x, int xcs, int xco, void* y, int ycs, int yco, int C,
                           int nz, int ny, int nx, hipStream_t st);
hipError_t launch_cnn_pool_bwd(const void* dy, int dcs, int dco, void* dx, int xcs, int xco,
                               const void* mask, int mcs, int mco, int accumulate, int C, int nz,
                               int ny, int nx, hipStream_t st);
hipError_t launch_cnn_up(const void* x, int xcs, int xco, void* y, int ycs, int yco, int C,
                         int nz, int ny, int nx, hipStream_t st);
hipError_t launch_cnn_up_bwd(const void* dy, int dcs, int dco, void* dx, int xcs, int xco,
                             const void* mask, int mcs, int mco, int C, int nz, int ny, int nx,
                             hipStream_t st);
hipError_t launch_cnn_input(const float* img, int64_t n, float scale, void* x, hipStream_t st);
hipError_t launch_cnn_output(const void* out, int ocs, const float* img, int64_t n,
                             float inv_scale, int clamp, float* pred, hipStream_t st);
// Residual output with an in-plane separable filter (taps[2r + 1], r <=
// CNN_SMOOTH_MAX_R, edge values repeated) applied to out before it is added.
constexpr int CNN_SMOOTH_MAX_R = 4;
hipError_t launch_cnn_output_smooth(const void* out, int ocs, const float* img, int nz, int ny,
                                    int nx, float inv_scale, const float* taps, int r, int clamp,
                                    float* pred, hipStream_t st);
hipError_t launch_cnn_mse(const float* pred, const float* target, int64_t n, float gscale,
                          void* dout, int dcs, double* partials, double* result, hipStream_t st);
hipError_t launch_cnn_pack(const float* w, int cin, int cout, void* wf, void* wd,
                           hipStream_t st);
hipError_t launch_cnn_adam(float* p, float* m, float* v, const float* g, int64_t n, float lr,
                           float b1, float b2, float eps, int step, hipStream_t st);

// OS-SART: forward with row sums (reference layout), projection difference,
// backprojection + update in place (diff detector-row-major; prior may be
// NULL), and the Huber-potential prior (accumulating).
hipError_t launch_sart_project(const DevGeom& g, const ViewTab* views, int a0, int a1,
                               int cos_weight, const float* img, float* prj, float* norm,
                               hipStream_t st);
hipError_t launch_sart_diff(float* diff, const float* p, const float* est, const float* norm,
                            int64_t n, hipStream_t st);
hipError_t launch_sart_update(const DevGeom& g, const ViewTab* views, int a0, int a1,
                              float* img, const float* diff_rows, const float* prior,
                              float lambda, float beta, hipStream_t st);
hipError_t launch_huber_prior(const DevGeom& g, float* prior, const float* img, float delta,
                              hipStream_t st);
// Accelerated SQS momentum with the restart decision on the device (no host
// round trip): state = {t, c, restarts, last test value}; partials: 1024 + 4
// doubles (the restart-test partial sums, then a snapshot of state).
hipError_t launch_momentum_init(double* state, hipStream_t st);
// The images in `dtype` storage (F32 / BF16 / F16), arithmetic in FP32 / FP64.
hipError_t launch_momentum_step(void* x_next, const void* z, const void* z_prev,
                                const void* x_k, int dtype, int64_t n, double* partials,
                                double* state, hipStream_t st);
// FP64 accumulation: acc += x; out = (float) acc; p = v.
hipError_t launch_add_f64(double* acc, const float* x, int64_t n, hipStream_t st);
hipError_t launch_f64_to_f32(float* out, const double* acc, int64_t n, hipStream_t st);
hipError_t launch_fill_f64(double* p, double v, int64_t n, hipStream_t st);
// Row filter of the DBT FBP (fbp.hip): cosine weight D / |S - E| and the
// convolution along detector x with the centred taps, by FFT in LDS over
// L = 2^fbp_fft_log2(nbx) >= 2 nbx - 1 points. spec: 3 L / 2 float2 of scratch
// (twiddles and the taps' spectrum); lds_limit: the device's LDS bytes per
// group (hipErrorInvalidValue when one column pair does not fit).
int fbp_fft_log2(int nbx);
hipError_t launch_fbp_filter(const DevGeom& g, const ViewTab* views, int a0, int a1,
                             const float* taps, int ntaps, const float* p, float2* spec,
                             float* q, int lds_limit, hipStream_t st);
hipError_t launch_line_integrals(float* p, const float* y, const float* b, int64_t n,
                                 hipStream_t st);
// Fused OSTR subset update with the exact transpose (hsub detector-row-major;
// image and h each in F32 / BF16 / F16 storage, d FP32).
hipError_t launch_update_fused_matched(const DevGeom& g, const ViewTab* views, int a0, int a1,
                                       const void* x_in, void* x_out, int img_dtype,
                                       const float* d, const void* hsub, int prj_dtype,
                                       float beta, float capl_scale, int potential, float delta,
                                       float rho, int guard, hipStream_t st);
// ReconDBT SART / ML (SART_dbt.m, ML_dbt.m) elementwise steps.
hipError_t launch_rdbt_sart_residual(float* y, const float* p, const float* est,
                                     const float* l, float thresh, int64_t n, hipStream_t st);
hipError_t launch_rdbt_sart_update(float* x, const float* num, const float* den, float step,
                                   int clamp, int64_t n, hipStream_t st);
hipError_t launch_rdbt_ml_terms(float* ydif, float* wdrr, const float* drr, const float* counts,
                                const float* i0, float i0s, int64_t n, hipStream_t st);
hipError_t launch_rdbt_ml_update(float* x, float* upd, const float* num, const float* den,
                                 float step, int64_t n, hipStream_t st);
// 2x2 supersampled forward + merge (prj [slot][nby][nbx]; prj4 [slot][2nby][2nbx]
// or null).
hipError_t launch_project_4r(const DevGeom& g, const ViewTab* views, int a0, int a1,
                             const float* img, float* prj, float* prj4, hipStream_t st);
// Fused OS-SART subset: forward + SART_prj_diff (p and diff detector-row-major),
// then backprojection + Huber prior (beta != 0) + update from x_in into x_out.
hipError_t launch_sart_project_diff(const DevGeom& g, const ViewTab* views, int a0, int a1,
                                    int cos_weight, const float* img, const float* p_rows,
                                    float* diff_rows, hipStream_t st);
hipError_t launch_sart_update_fused(const DevGeom& g, const ViewTab* views, int a0, int a1,
                                    const float* x_in, float* x_out, const float* diff_rows,
                                    float lambda, float beta, float delta, hipStream_t st);

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
hipError_t launch_extrapolate(float* x, const float* z, const float* zp, float c, int64_t n,
                              hipStream_t st);
// out[0] = sum (x - z)(z - zp) in FP64; partials: 1024 doubles
hipError_t launch_restart_dot(const float* x, const float* z, const float* zp, int64_t n,
                              double* partials, double* out, hipStream_t st);

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

// Work-group to tile mapping for kernels whose tiles cost very different
// amounts: of a detector's ray tiles only those in the volume's shadow walk
// (about 15% of the bins at the clinical size, a band across the middle
// columns), so giving each XCD one contiguous range, as xcd_tile does, leaves
// the XCDs that own the outer columns idle. Here tiles are grouped in blocks
// of SPREAD_BLOCK consecutive tiles and the blocks of each run of 8 are dealt
// one per XCD (block B on XCD B % 8): every XCD gets an even share of the
// shadow, and the groups of one XCD still work through neighbouring tiles
// (L2 reuse within a block). Work-group b runs on XCD b % 8; the tiles past
// the last whole run of 8 blocks keep b's own index.
constexpr int SPREAD_BLOCK = 16;
__device__ __forceinline__ int xcd_spread_tile(int b, int nblocks) {
    constexpr int span = 8 * SPREAD_BLOCK;
    const int whole = (nblocks / span) * span;
    if (b >= whole) return b;
    const int r = b % span;
    return (b - r) + (r % 8) * SPREAD_BLOCK + r / 8;
}

// The same spread for a launch that hands each XCD one contiguous range of a
// case's tile indices (BatchWalk): a permutation of [0, n_tiles) whose eighths
// cover every eighth block. The whole blocks of SPREAD_BLOCK tiles are put in
// the order 0, 8, 16, ..., 1, 9, 17, ..., 7, 15, ... (position p -> block);
// a partial last block keeps its place.
__device__ __forceinline__ int spread_block_perm(int t, int n_tiles) {
    const int nb = n_tiles / SPREAD_BLOCK;  // whole blocks
    const int p = t / SPREAD_BLOCK;
    if (p >= nb) return t;
    const int q = nb / 8, rem = nb % 8;    // class k holds q + (k < rem) blocks
    int k, m;
    if (p < rem * (q + 1)) {
        k = p / (q + 1);
        m = p % (q + 1);
    } else {
        const int p2 = p - rem * (q + 1);
        k = rem + p2 / q;
        m = p2 % q;
    }
    return (k + 8 * m) * SPREAD_BLOCK + t % SPREAD_BLOCK;
}

// Tile walk of a batched launch (one-tile-per-group and persistent grids
// alike). Work-group b runs on XCD b % 8. The tiles of each case are cut into
// nxcd = min(8, grid) contiguous chunks and XCD x owns chunk x of every case;
// the groups of an XCD walk its chunks in case order, interleaved (group k of
// n takes local tiles k, k+n, k+2n, ...). Tiles of one case cost about the
// same and tiles of different cases do not (view count, thickness), so every
// XCD gets the same mix of work, and the groups in flight on one XCD work on
// neighbouring tiles of one case (L2 reuse).
//
// The host writes the schedule (batch_schedule): for XCD x, xoff[x][c] is
// the position of case c's chunk in the XCD's list (xoff[x][n_cases] = list
// length) and xlo[x][c] the first tile of that chunk in the case's own tile
// numbering; rows are n_cases + 1 ints, xoff rows first (8 rows), then xlo.
// next() finds the case by binary search, kept while the group stays in it,
// and returns false once the group's share is done; every group reaches that.
struct BatchWalk {
    const int* xoff;
    const int* xlo;
    int n_cases, n, u, len, c, c_begin, c_end, c_lo;

    __device__ int at(const int* p, int i) const { return __builtin_amdgcn_readfirstlane(p[i]); }
    __device__ BatchWalk(const int* sched, int nc) : n_cases(nc) {
        const int G = gridDim.x;
        const int nxcd = G < 8 ? G : 8;
        const int x = blockIdx.x % nxcd;
        n = G / nxcd + (x < G % nxcd ? 1 : 0);
        u = blockIdx.x / nxcd;
        xoff = sched + x * (nc + 1);
        xlo = sched + (8 + x) * (nc + 1);
        len = at(xoff, nc);
        c = 0;
        c_begin = c_end = 0;
    }
    __device__ bool next(int& t, int& case_idx) {
        if (u >= len) return false;
        if (u >= c_end) {  // largest c with xoff[c] <= u (a non-empty chunk)
            int lo = 0, hi = n_cases;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (at(xoff, mid) <= u) lo = mid; else hi = mid;
            }
            c = lo;
            c_begin = at(xoff, c);
            c_end = at(xoff, c + 1);
            c_lo = at(xlo, c);
        }
        t = c_lo + (u - c_begin);
        case_idx = c;
        u += n;
        return true;
    }
};

// Host side of BatchWalk: fills sched (2 * 8 * (n_cases + 1) ints) for tile
// counts `tiles` of the cases and a grid of `grid` work-groups.
inline void batch_schedule(int* sched, const int64_t* tiles, int n_cases, int grid) {
    const int nxcd = grid < 8 ? grid : 8;
    const int row = n_cases + 1;
    for (int x = 0; x < 8; ++x) {
        int off = 0;
        for (int c = 0; c < n_cases; ++c) {
            const int lo = (int)(tiles[c] * x / nxcd), hi = (int)(tiles[c] * (x + 1) / nxcd);
            sched[x * row + c] = off;
            sched[(8 + x) * row + c] = lo;
            off += x < nxcd ? hi - lo : 0;
        }
        sched[x * row + n_cases] = off;
        sched[(8 + x) * row + n_cases] = 0;
    }
}

}  // namespace ostr
