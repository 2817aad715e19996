// elementwise.hip - the reference's host-side CPU loops, as device kernels:
//   compute_yry        Tomo_OSTR_gold.cpp:183-205
//   compute_gamma_yry  Tomo_OSTR_gold.cpp:208-218
//   regroup_prj        Tomo_OSTR_gold.cpp:220-242
//   compute_h          Tomo_OSTR_gold.cpp:245-260
//   capL * M + update_est  Tomo_OSTR_gold.cpp:767-771, 265-295
// Grid-stride loops, 256 threads, 4 elements in flight per lane.
#include "ostr_internal.h"

namespace ostr {

namespace {

inline unsigned grid_for(int64_t n) {
    int64_t blocks = (n + 1023) / 1024;
    if (blocks < 1) blocks = 1;
    if (blocks > 65536) blocks = 65536;
    return (unsigned)blocks;
}

__device__ __forceinline__ float yry_one(float y, float r) {
    if (y == 0.0f) return 0.0f;
    float dif = y - r;
    if (dif <= 0.0f) dif = y;
    return (dif * dif) / y;
}

__global__ __launch_bounds__(256) void yry_kernel(float* __restrict__ out,
                                                  const float* __restrict__ y,
                                                  const float* __restrict__ r, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        out[i] = yry_one(y[i], r ? r[i] : 0.0f);
}

// gc = gamma * yry with gamma already in gc (fused precompute)
__global__ __launch_bounds__(256) void yry_gamma_kernel(float* __restrict__ gc,
                                                        const float* __restrict__ y,
                                                        const float* __restrict__ r, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        gc[i] = yry_one(y[i], r ? r[i] : 0.0f) * gc[i];
}

__global__ __launch_bounds__(256) void mul_kernel(float* __restrict__ out,
                                                  const float* __restrict__ a,
                                                  const float* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        out[i] = a[i] * b[i];
}

__global__ __launch_bounds__(256) void h_kernel(float* __restrict__ h,
                                                const float* __restrict__ y,
                                                const float* __restrict__ b,
                                                const float* __restrict__ l,
                                                const float* __restrict__ r, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        float yt = b[i] * expf(-l[i]);
        float den = r ? yt + r[i] : yt;
        h[i] = (y[i] / den - 1.0f) * yt;
    }
}

__global__ __launch_bounds__(256) void update_kernel(float* __restrict__ x,
                                                     const float* __restrict__ capl,
                                                     const float* __restrict__ rd,
                                                     const float* __restrict__ d,
                                                     const float* __restrict__ rdd, float beta,
                                                     float capl_scale, int guard, int64_t n) {
    const float two_beta = 2.0f * beta;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        float num = capl_scale * capl[i] + beta * rd[i];
        float den = d[i] + two_beta * rdd[i];
        if (guard && den == 0.0f) continue;
        float v = x[i] - num / den;
        x[i] = v < 0.0f ? 0.0f : v;
    }
}

// dst block k = m*vs + j  <-  src block j*M + m, blocks of nbin floats
__global__ __launch_bounds__(256) void regroup_kernel(float* __restrict__ dst,
                                                      const float* __restrict__ src,
                                                      int n_views, int n_subsets, int64_t nbin) {
    const int vs = n_views / n_subsets;
    const int k = blockIdx.y;
    const int m = k / vs, j = k % vs;
    const float* s = src + (int64_t)(j * n_subsets + m) * nbin;
    float* o = dst + (int64_t)k * nbin;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nbin;
         i += (int64_t)gridDim.x * 256)
        o[i] = s[i];
}

// BP.m step for one view: q = p / (c l) where c l > thr, else 0. p is the line
// integral (counts = 0) or log(b / max(y, 1)) of counts y and blank b; l and
// p/y/b in the reference layout [nbx][nby], q written detector-row-major.
__global__ __launch_bounds__(256) void bp_data_kernel(float* __restrict__ q_rows,
                                                      const float* __restrict__ l,
                                                      const float* __restrict__ py,
                                                      const float* __restrict__ b, int counts,
                                                      float c, float thr, int nbx, int nby) {
    const int64_t n = (int64_t)nbx * nby;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256) {
        const int i = (int)(k / nby), j = (int)(k % nby);
        const float lv = c * l[k];
        float p = py[k];
        if (counts) p = logf(b[k] / fmaxf(p, 1.0f));
        q_rows[(int64_t)j * nbx + i] = lv > thr ? p / lv : 0.0f;
    }
}

__global__ __launch_bounds__(256) void bp_accum_kernel(float* __restrict__ img,
                                                       const float* __restrict__ num,
                                                       const float* __restrict__ den,
                                                       int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        if (den[i] != 0.0f) img[i] = img[i] + num[i] / den[i];
}

__global__ __launch_bounds__(256) void div_kernel(float* __restrict__ p, float v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        p[i] = p[i] / v;
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ p, float v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        p[i] = v;
}

}  // namespace

hipError_t launch_bp_data(float* q_rows, const float* l, const float* py, const float* b,
                          int counts, float c, float thr, int nbx, int nby, hipStream_t st) {
    hipLaunchKernelGGL(bp_data_kernel, dim3(grid_for((int64_t)nbx * nby)), dim3(256), 0, st,
                       q_rows, l, py, b, counts, c, thr, nbx, nby);
    return hipGetLastError();
}

hipError_t launch_bp_accum(float* img, const float* num, const float* den, int64_t n,
                           hipStream_t st) {
    hipLaunchKernelGGL(bp_accum_kernel, dim3(grid_for(n)), dim3(256), 0, st, img, num, den, n);
    return hipGetLastError();
}

hipError_t launch_div(float* p, float v, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(div_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

// x = max(0, z + c (z - z_prev)): the momentum extrapolation of accelerated
// SQS / OS-SQS (Nesterov; Kim, Ramani & Fessler 2015), with the non-negativity
// projection of the OSTR update.
__global__ __launch_bounds__(256) void extrapolate_kernel(float* __restrict__ x,
                                                          const float* __restrict__ z,
                                                          const float* __restrict__ zp, float c,
                                                          int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = fmaf(c, z[i] - zp[i], z[i]);
        x[i] = v > 0.0f ? v : 0.0f;
    }
}

// diff = (p - est) / norm where norm != 0, else 0 (SART_prj_diff_kernel,
// Tomo_OSTR.cu:2145-2183).
__global__ __launch_bounds__(256) void sart_diff_kernel(float* __restrict__ diff,
                                                        const float* __restrict__ p,
                                                        const float* __restrict__ est,
                                                        const float* __restrict__ norm,
                                                        int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        diff[i] = norm[i] != 0.0f ? (p[i] - est[i]) / norm[i] : 0.0f;
}

// ReconDBT SART (SART_dbt.m:55-66), one view: y = (p - est) / l where
// l > thresh (the row sum above 5 dz), else y keeps its value (the MATLAB y
// persists from the previous view there).
__global__ __launch_bounds__(256) void rdbt_sart_residual_kernel(
    float* __restrict__ y, const float* __restrict__ p, const float* __restrict__ est,
    const float* __restrict__ l, float thresh, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float li = l[i];
        if (li > thresh) y[i] = (p[i] - est[i]) / li;
    }
}

// ReconDBT SART update (SART_dbt.m:70-76, :98): u = num / den with NaN -> 0,
// x = x + step u; clamp: then x = 0 where x < 0 (after the last view).
__global__ __launch_bounds__(256) void rdbt_sart_update_kernel(
    float* __restrict__ x, const float* __restrict__ num, const float* __restrict__ den,
    float step, int clamp, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float u = num[i] / den[i];
        if (u != u) u = 0.0f;
        float xn = x[i] + step * u;
        if (clamp && xn < 0.0f) xn = 0.0f;
        x[i] = xn;
    }
}

// ReconDBT ML (ML_dbt.m:88-96), one view: yb = I0 exp(-drr), ydif = yb - counts,
// wdrr = drr yb; I0 per bin (i0 != null) or the scalar i0s.
__global__ __launch_bounds__(256) void rdbt_ml_terms_kernel(
    float* __restrict__ ydif, float* __restrict__ wdrr, const float* __restrict__ drr,
    const float* __restrict__ counts, const float* __restrict__ i0, float i0s, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float l = drr[i];
        const float yb = (i0 ? i0[i] : i0s) * expf(-l);
        ydif[i] = yb - counts[i];
        wdrr[i] = l * yb;
    }
}

// ReconDBT ML update (ML_dbt.m:100-105): where den != 0, upd = num / den (upd
// keeps its previous value elsewhere), NaN -> 0; x = x + (x upd) step, NaN -> 0,
// then x = 0 where x < 0.
__global__ __launch_bounds__(256) void rdbt_ml_update_kernel(
    float* __restrict__ x, float* __restrict__ upd, const float* __restrict__ num,
    const float* __restrict__ den, float step, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float u = upd[i];
        if (den[i] != 0.0f) u = num[i] / den[i];
        if (u != u) u = 0.0f;
        upd[i] = u;
        const float xi = x[i];
        float xn = xi + (xi * u) * step;
        if (xn != xn) xn = 0.0f;
        if (xn < 0.0f) xn = 0.0f;
        x[i] = xn;
    }
}

// Line integrals of counts y under blank b: p = log(b / max(y, 1)) (the data
// step of ReconDBT/BP.m:54-60, as bp_data_kernel forms it).
__global__ __launch_bounds__(256) void line_integrals_kernel(float* __restrict__ p,
                                                             const float* __restrict__ y,
                                                             const float* __restrict__ b,
                                                             int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = logf(b[i] / fmaxf(y[i], 1.0f));
}

hipError_t launch_line_integrals(float* p, const float* y, const float* b, int64_t n,
                                 hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(line_integrals_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, y, b, n);
    return hipGetLastError();
}

hipError_t launch_rdbt_sart_residual(float* y, const float* p, const float* est,
                                     const float* l, float thresh, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rdbt_sart_residual_kernel, dim3(grid_for(n)), dim3(256), 0, st, y, p,
                       est, l, thresh, n);
    return hipGetLastError();
}

hipError_t launch_rdbt_sart_update(float* x, const float* num, const float* den, float step,
                                   int clamp, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rdbt_sart_update_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, num,
                       den, step, clamp, n);
    return hipGetLastError();
}

hipError_t launch_rdbt_ml_terms(float* ydif, float* wdrr, const float* drr, const float* counts,
                                const float* i0, float i0s, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rdbt_ml_terms_kernel, dim3(grid_for(n)), dim3(256), 0, st, ydif, wdrr,
                       drr, counts, i0, i0s, n);
    return hipGetLastError();
}

hipError_t launch_rdbt_ml_update(float* x, float* upd, const float* num, const float* den,
                                 float step, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rdbt_ml_update_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, upd, num,
                       den, step, n);
    return hipGetLastError();
}

hipError_t launch_sart_diff(float* diff, const float* p, const float* est, const float* norm,
                            int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sart_diff_kernel, dim3(grid_for(n)), dim3(256), 0, st, diff, p, est, norm,
                       n);
    return hipGetLastError();
}

hipError_t launch_extrapolate(float* x, const float* z, const float* zp, float c, int64_t n,
                              hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(extrapolate_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, z, zp, c, n);
    return hipGetLastError();
}

hipError_t launch_fill(float* p, float v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_yry(float* out, const float* y, const float* r, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(yry_kernel, dim3(grid_for(n)), dim3(256), 0, st, out, y, r, n);
    return hipGetLastError();
}

hipError_t launch_yry_gamma(float* gc, const float* y, const float* r, int64_t n,
                            hipStream_t st) {
    hipLaunchKernelGGL(yry_gamma_kernel, dim3(grid_for(n)), dim3(256), 0, st, gc, y, r, n);
    return hipGetLastError();
}

hipError_t launch_mul(float* out, const float* a, const float* b, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(mul_kernel, dim3(grid_for(n)), dim3(256), 0, st, out, a, b, n);
    return hipGetLastError();
}

hipError_t launch_compute_h(float* h, const float* y, const float* b, const float* l,
                            const float* r, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(h_kernel, dim3(grid_for(n)), dim3(256), 0, st, h, y, b, l, r, n);
    return hipGetLastError();
}

hipError_t launch_update_est(float* x, const float* capl, const float* rd, const float* d,
                             const float* rdd, float beta, float capl_scale, int guard,
                             int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(update_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, capl, rd, d, rdd,
                       beta, capl_scale, guard, n);
    return hipGetLastError();
}

hipError_t launch_regroup(float* dst, const float* src, int n_views, int n_subsets,
                          int64_t nbin, hipStream_t st) {
    int64_t bx = (nbin + 1023) / 1024;
    if (bx > 4096) bx = 4096;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(regroup_kernel, dim3((unsigned)bx, n_views), dim3(256), 0, st, dst, src,
                       n_views, n_subsets, nbin);
    return hipGetLastError();
}

namespace {

constexpr int NORM_BLOCKS = 1024;

// partials[k] = sum over block k's strided elements of (a - b)^2 and
// partials[NORM_BLOCKS + k] = sum of b^2, in FP64 (fixed grid: deterministic).
template <typename T>
__global__ __launch_bounds__(256) void diff_norms_kernel(const T* __restrict__ a,
                                                         const T* __restrict__ b, int64_t n,
                                                         double* __restrict__ partials) {
    double dd = 0.0, bb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)NORM_BLOCKS * 256) {
        const double bv = (double)ld<T>(b, i);
        const double df = (double)ld<T>(a, i) - bv;
        dd += df * df;
        bb += bv * bv;
    }
    const double s0 = block_sum_256(dd);
    __syncthreads();
    const double s1 = block_sum_256(bb);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s0;
        partials[NORM_BLOCKS + blockIdx.x] = s1;
    }
}

// partials[k] = sum over block k's strided elements of (x - z)(z - zp), FP64
// (the gradient-restart test of accelerated SQS; fixed grid: deterministic).
// With state: block 0 also copies the 4 state doubles to snap, which the
// momentum step's second kernel reads while it rewrites state. T: the
// images' storage type (each element is widened exactly to double).
template <typename T>
__global__ __launch_bounds__(256) void restart_dot_kernel(const T* __restrict__ x,
                                                          const T* __restrict__ z,
                                                          const T* __restrict__ zp, int64_t n,
                                                          double* __restrict__ partials,
                                                          const double* __restrict__ state,
                                                          double* __restrict__ snap) {
    if (state && blockIdx.x == 0 && threadIdx.x < 4) snap[threadIdx.x] = state[threadIdx.x];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)NORM_BLOCKS * 256) {
        const double zi = (double)ld<T>(z, i);
        acc += ((double)ld<T>(x, i) - zi) * (zi - (double)ld<T>(zp, i));
    }
    const double s = block_sum_256(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* __restrict__ p,
                                                              int64_t n,
                                                              double* __restrict__ out) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) acc += p[i];
    const double s = block_sum_256(acc);
    if (threadIdx.x == 0) out[0] = s;
}

// Momentum state (4 doubles): t, c, restarts, last restart-test value.
__global__ void momentum_init_kernel(double* __restrict__ state) {
    if (threadIdx.x == 0) {
        state[0] = 1.0;
        state[1] = 0.0;
        state[2] = 0.0;
        state[3] = 0.0;
    }
}

// One block: the restart test dot = sum of the partials (the same fixed order
// as reduce_partials_kernel), then, on the device, the Nesterov coefficient:
// dot > 0 -> restart (t = 1, c = 0); else t' = (1 + sqrt(1 + 4 t^2)) / 2,
// c = (t - 1) / t', t = t' (O'Donoghue & Candes 2015 gradient restart).
__global__ __launch_bounds__(256) void momentum_coef_kernel(const double* __restrict__ p,
                                                            int64_t n,
                                                            double* __restrict__ state) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) acc += p[i];
    const double dot = block_sum_256(acc);
    if (threadIdx.x == 0) {
        const double t = state[0];
        if (dot > 0.0) {
            state[0] = 1.0;
            state[1] = 0.0;
            state[2] = state[2] + 1.0;
        } else {
            const double tn = 0.5 * (1.0 + sqrt(1.0 + 4.0 * t * t));
            state[1] = (t - 1.0) / tn;
            state[0] = tn;
        }
        state[3] = dot;
    }
}

// The momentum step after restart_dot_kernel, in one launch: every block sums
// the n_p partials in the fixed order of momentum_coef_kernel (so each gets
// the same dot) and applies the coefficient rule to the state snapshot snap
// (t, c, restarts, last test value, taken by restart_dot_kernel); block 0
// writes the new state. Then x = max(0, z + c (z - zp)) with c = (float) of
// the new coefficient, as extrapolate_kernel. The grid is at most
// MOMENTUM_BLOCKS blocks, which bounds the repeated partial sums.
constexpr int MOMENTUM_BLOCKS = 512;

template <typename T>
__global__ __launch_bounds__(256) void momentum_extrapolate_kernel(
    T* __restrict__ x, const T* __restrict__ z, const T* __restrict__ zp,
    const double* __restrict__ p, int64_t n_p, const double* __restrict__ snap,
    double* __restrict__ state, int64_t n) {
    __shared__ float c_blk;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n_p; i += 256) acc += p[i];
    const double dot = block_sum_256(acc);
    if (threadIdx.x == 0) {
        const double t = snap[0];
        double tn, c, restarts = snap[2];
        if (dot > 0.0) {
            tn = 1.0;
            c = 0.0;
            restarts = restarts + 1.0;
        } else {
            tn = 0.5 * (1.0 + sqrt(1.0 + 4.0 * t * t));
            c = (t - 1.0) / tn;
        }
        c_blk = (float)c;
        if (blockIdx.x == 0) {
            state[0] = tn;
            state[1] = c;
            state[2] = restarts;
            state[3] = dot;
        }
    }
    __syncthreads();
    const float c = c_blk;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float zi = ld<T>(z, i);
        const float v = fmaf(c, zi - ld<T>(zp, i), zi);
        st<T>(x, i, v > 0.0f ? v : 0.0f);
    }
}

// acc += x (FP64 accumulation of FP32 terms: the denominator precompute
// summed over views in double); out = (float) acc.
__global__ __launch_bounds__(256) void add_f64_kernel(double* __restrict__ acc,
                                                      const float* __restrict__ x, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        acc[i] = acc[i] + (double)x[i];
}

__global__ __launch_bounds__(256) void f64_to_f32_kernel(float* __restrict__ out,
                                                         const double* __restrict__ acc,
                                                         int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)acc[i];
}

__global__ __launch_bounds__(256) void fill_f64_kernel(double* __restrict__ p, double v,
                                                       int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
hipError_t momentum_step_t(T* x_next, const T* z, const T* z_prev, const T* x_k, int64_t n,
                           double* partials, double* state, hipStream_t st) {
    double* snap = partials + NORM_BLOCKS;
    hipLaunchKernelGGL((restart_dot_kernel<T>), dim3(NORM_BLOCKS), dim3(256), 0, st, x_k, z,
                       z_prev, n, partials, (const double*)state, snap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (n <= 0) {
        hipLaunchKernelGGL(momentum_coef_kernel, dim3(1), dim3(256), 0, st, partials,
                           (int64_t)NORM_BLOCKS, state);
        return hipGetLastError();
    }
    const unsigned blocks = grid_for(n) < MOMENTUM_BLOCKS ? grid_for(n) : MOMENTUM_BLOCKS;
    hipLaunchKernelGGL((momentum_extrapolate_kernel<T>), dim3(blocks), dim3(256), 0, st, x_next,
                       z, z_prev, (const double*)partials, (int64_t)NORM_BLOCKS,
                       (const double*)snap, state, n);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_momentum_init(double* state, hipStream_t st) {
    hipLaunchKernelGGL(momentum_init_kernel, dim3(1), dim3(64), 0, st, state);
    return hipGetLastError();
}

hipError_t launch_momentum_step(void* x_next, const void* z, const void* z_prev,
                                const void* x_k, int dtype, int64_t n, double* partials,
                                double* state, hipStream_t st) {
    switch (dtype) {
        case OSTR_F32:
            return momentum_step_t<float>((float*)x_next, (const float*)z, (const float*)z_prev,
                                          (const float*)x_k, n, partials, state, st);
        case OSTR_BF16:
            return momentum_step_t<bf16_t>((bf16_t*)x_next, (const bf16_t*)z,
                                           (const bf16_t*)z_prev, (const bf16_t*)x_k, n,
                                           partials, state, st);
        case OSTR_F16:
            return momentum_step_t<_Float16>((_Float16*)x_next, (const _Float16*)z,
                                             (const _Float16*)z_prev, (const _Float16*)x_k, n,
                                             partials, state, st);
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t launch_add_f64(double* acc, const float* x, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(add_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, acc, x, n);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(float* out, const double* acc, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_for(n)), dim3(256), 0, st, out, acc, n);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double* p, double v, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

int64_t diff_norms_partials(int64_t) { return 2 * NORM_BLOCKS; }

hipError_t launch_diff_norms(const void* a, const void* b, int dtype, int64_t n,
                             double* partials, int64_t* n_partials, hipStream_t st) {
    *n_partials = 2 * NORM_BLOCKS;
    switch (dtype) {
        case OSTR_F32:
            hipLaunchKernelGGL((diff_norms_kernel<float>), dim3(NORM_BLOCKS), dim3(256), 0, st,
                               (const float*)a, (const float*)b, n, partials);
            break;
        case OSTR_BF16:
            hipLaunchKernelGGL((diff_norms_kernel<bf16_t>), dim3(NORM_BLOCKS), dim3(256), 0, st,
                               (const bf16_t*)a, (const bf16_t*)b, n, partials);
            break;
        case OSTR_F16:
            hipLaunchKernelGGL((diff_norms_kernel<_Float16>), dim3(NORM_BLOCKS), dim3(256), 0,
                               st, (const _Float16*)a, (const _Float16*)b, n, partials);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_restart_dot(const float* x, const float* z, const float* zp, int64_t n,
                              double* partials, double* out, hipStream_t st) {
    hipLaunchKernelGGL((restart_dot_kernel<float>), dim3(NORM_BLOCKS), dim3(256), 0, st, x, z,
                       zp, n, partials, (const double*)nullptr, (double*)nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_partials(partials, NORM_BLOCKS, out, st);
}

hipError_t launch_reduce_partials(const double* partials, int64_t n, double* out,
                                  hipStream_t st) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, st, partials, n, out);
    return hipGetLastError();
}

}  // namespace ostr
