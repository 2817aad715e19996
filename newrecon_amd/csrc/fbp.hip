// fbp.hip - the row filter of the DBT FBP as a GEMM on the matrix cores.
//
// The filter is a convolution along detector x (the tube-motion direction)
// with odd-length centred taps, samples off the detector 0:
//   q[a][i][j] = sum_k taps[k] pw[a][i + k - c][j],  c = (ntaps - 1) / 2,
// that is q[a] = T pw[a] with the nbx x nbx Toeplitz matrix
// T[i][i'] = taps[i' - i + c] (0 outside the taps). With the views in the
// reference layout [slot][nbx][nby] (j fastest), pw[a] read column-major is
// the nby x nbx matrix pw[a]^T, so one strided-batched SGEMM forms
// q[a]^T = pw[a]^T T^T for every view: 2 nby nbx^2 flops per view in FP32 on
// MFMA (rocBLAS), instead of a direct sum whose per-bin loads made it
// gather-bound.
#include <rocblas/rocblas.h>

#include "ostr_internal.h"

namespace ostr {

namespace {

// T^T column-major = T row-major: element o = i * nbx + i' holds T[i][i'].
__global__ __launch_bounds__(256) void toeplitz_kernel(const float* __restrict__ taps, int ntaps,
                                                       int nbx, float* __restrict__ T) {
    const int64_t n = (int64_t)nbx * nbx;
    const int c = (ntaps - 1) / 2;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < n;
         o += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(o / nbx), ip = (int)(o % nbx);
        const int k = ip - i + c;
        T[o] = (k >= 0 && k < ntaps) ? taps[k] : 0.0f;
    }
}

}  // namespace

void fbp_blas_destroy(void* blas) {
    if (blas) (void)rocblas_destroy_handle((rocblas_handle)blas);
}

hipError_t launch_fbp_filter(const DevGeom& g, const ViewTab* views, int a0, int a1,
                             const float* taps, int ntaps, const float* p, float* pw,
                             float* T_buf, float* q, void** blas, hipStream_t st) {
    const int n_slots = a1 - a0;
    if (n_slots <= 0) return hipSuccess;
    hipError_t e = launch_fbp_weight(g, views, a0, a1, p, pw, st);
    if (e != hipSuccess) return e;
    const int64_t nt = (int64_t)g.nbx * g.nbx;
    const int64_t blocks = (nt + 255) / 256 < 8192 ? (nt + 255) / 256 : 8192;
    hipLaunchKernelGGL(toeplitz_kernel, dim3((unsigned)blocks), dim3(256), 0, st, taps, ntaps,
                       g.nbx, T_buf);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (!*blas) {
        rocblas_handle hb = nullptr;
        if (rocblas_create_handle(&hb) != rocblas_status_success) return hipErrorUnknown;
        *blas = hb;
    }
    rocblas_handle hb = (rocblas_handle)*blas;
    if (rocblas_set_stream(hb, st) != rocblas_status_success) return hipErrorUnknown;
    const float one = 1.0f, zero = 0.0f;
    const rocblas_stride sv = (rocblas_stride)g.nbx * g.nby;
    const rocblas_status s = rocblas_sgemm_strided_batched(
        hb, rocblas_operation_none, rocblas_operation_none, g.nby, g.nbx, g.nbx, &one, pw,
        g.nby, sv, T_buf, g.nbx, 0, &zero, q, g.nby, sv, n_slots);
    return s == rocblas_status_success ? hipSuccess : hipErrorUnknown;
}

}  // namespace ostr
