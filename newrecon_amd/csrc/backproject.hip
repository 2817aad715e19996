// backproject.hip - voxel-driven bilinear backprojector (SBP) and the fused
// OSTR subset update for gfx950.
//
// SBP follows backprj_gpu_manyviews_SBP (Tomo_OSTR.cu:1553-1734): the voxel
// centre is projected from the source onto the detector plane z = -detectorR;
// the bilinear taps (ilb,jlb) are chosen with the reference's half-bin rule and
// masked one by one to the detector; the interpolated value is scaled by
// Vsize_z * |S|^2/|S-V|^2 and summed over the slots of the launch, in slot
// order. The ray parameter t and the tap positions use the reference's exact
// divisions (a 1-ulp change of a tap position at bin ~200 is 1.5e-5 of a bin,
// visible on rough data); only the depth-weight division is a hardware
// reciprocal (a factor within ~1 ulp).
//
// MI355X mapping. The kernels read the sinogram in detector-row-major layout
// [slot][nby][nbx] (x fastest) and put the 64 lanes of a wave on 64
// consecutive voxels along x, so a wave's taps fall on ~50 consecutive bins of
// one or two detector rows (a few cache lines). Because every voxel of one
// slice sees the same ray parameter t, the tap column (ilb, fracI) depends
// only on (x, slice, view) and the tap row (jlb, fracJ) only on (y, slice,
// view): a lane computes its column terms once per view and reuses them for
// the ROWS rows it owns, and the row terms are wave-uniform. Work-group tiles
// (64 x 16 voxels of one slice) are dealt to XCDs in contiguous ranges.
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
constexpr int ROWS = 8;        // voxel rows (y) per wave
constexpr int WAVES = 4;       // waves per work-group
constexpr int TILE_X = 64, TILE_Y = ROWS * WAVES;

struct SlotTab {
    float x0[MAX_SLOTS], z0[MAX_SLOTS];
    float num[MAX_SLOTS];  // |S|^2 evaluated as fma(z0,z0,fma(y0,y0,x0*x0))
};

__device__ __forceinline__ float cell_centre(float origin, int i, float size) {
    return (float)((double)origin + ((double)i + 0.5) * (double)size);
}

// Half-bin tap rule of Tomo_OSTR.cu:1668-1678.
__device__ __forceinline__ void tap(float m, int& l, float& f) {
    l = (int)floorf(m);
    if (m < (float)l + 0.5f) l -= 1;
    f = m - ((float)l + 0.5f);
}

struct TileCoord {
    int ix, iy0, iz;
};

__device__ __forceinline__ TileCoord tile_coord(const DevGeom& g) {
    const int tiles_x = (g.nx + TILE_X - 1) / TILE_X;
    const int tiles_y = (g.ny + TILE_Y - 1) / TILE_Y;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = t % tiles_x;
    const int rest = t / tiles_x;
    TileCoord c;
    c.ix = bx * TILE_X + (threadIdx.x & 63);
    c.iy0 = (rest % tiles_y) * TILE_Y + (threadIdx.x >> 6) * ROWS;
    c.iz = rest / tiles_y;
    return c;
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// total[r] = sum over slots of the SBP value of voxel (ix, iy0 + r, iz);
// prj is [slot][nby][nbx]. The row terms (jlb, fracJ, y offset, masks) of a
// (row, slot) pair are wave-uniform: lane r + ROWS*k evaluates them for row r
// and slot k of a chunk of VIEW_CHUNK slots, and every lane reads them back
// with readlane, so each exact division runs once per wave.
template <typename TP>
__device__ __forceinline__ void sbp_rows(const DevGeom& g, const SlotTab& S, int n,
                                         const TP* __restrict__ prj, int ix, int iy0, int iz,
                                         float total[ROWS]) {
    constexpr int VIEW_CHUNK = 64 / ROWS;
    const int64_t nbin = (int64_t)g.nbx * g.nby;
    const int lane = threadIdx.x & 63;
    const float y0 = g.src_y;
    const float x1 = cell_centre(g.x_p0, ix, g.dx);
    const float z1 = cell_centre(g.z_p0, iz, g.dz);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) total[r] = 0.0f;
    for (int c0 = 0; c0 < n; c0 += VIEW_CHUNK) {
        // this lane's (row, slot) entry of the chunk
        int rt_jl = 0;
        float rt_fj = 0.0f, rt_ey = 0.0f;
        {
            const int r = lane % ROWS, a = c0 + lane / ROWS;
            if (a < n) {
                const float z0 = S.z0[a];
                const float y1 = cell_centre(g.y_p0, iy0 + r, g.dy);
                rt_ey = y0 - y1;
                if (z1 != z0) {
                    const float t = (-g.det_r - z0) / (z1 - z0);
                    const float y2 = fmaf(y1 - y0, t, y0);
                    tap((y2 - g.y_d0) / g.dby, rt_jl, rt_fj);
                }
            }
        }
        const int na = n - c0 < VIEW_CHUNK ? n - c0 : VIEW_CHUNK;
        for (int k = 0; k < na; ++k) {
            const int a = c0 + k;
            const float x0 = S.x0[a], z0 = S.z0[a], num = S.num[a];
            if (z1 == z0) continue;
            const float ex = x0 - x1, ez = z0 - z1;
            const float exx = ex * ex;
            const float t = (-g.det_r - z0) / (z1 - z0);
            const float x2 = fmaf(x1 - x0, t, x0);
            int il;
            float fi;
            tap((x2 - g.x_d0) / g.dbx, il, fi);
            const bool i0 = (unsigned)il < (unsigned)g.nbx;
            const bool i1 = (unsigned)(il + 1) < (unsigned)g.nbx;
            const float omfi = 1.0f - fi;
            const TP* P = prj + (int64_t)a * nbin + il;
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                const int src = k * ROWS + r;
                const int jl = __builtin_amdgcn_readlane(rt_jl, src);
                const float fj = readlane_f(rt_fj, src);
                const float ey = readlane_f(rt_ey, src);
                const bool j0 = (unsigned)jl < (unsigned)g.nby;
                const bool j1 = (unsigned)(jl + 1) < (unsigned)g.nby;
                const float den = fmaf(ez, ez, fmaf(ey, ey, exx));
                const float dw = num * __builtin_amdgcn_rcpf(den);
                const TP* row0 = P + (int64_t)jl * g.nbx;
                float p00 = 0.0f, p10 = 0.0f, p01 = 0.0f, p11 = 0.0f;
                if (j0) {
                    if (i0) p00 = ld<TP>(row0, 0);
                    if (i1) p10 = ld<TP>(row0, 1);
                }
                if (j1) {
                    if (i0) p01 = ld<TP>(row0, g.nbx);
                    if (i1) p11 = ld<TP>(row0, g.nbx + 1);
                }
                float d1 = omfi * p00;
                d1 = fmaf(fi, p10, d1);
                float d2 = omfi * p01;
                d2 = fmaf(fi, p11, d2);
                float u = fmaf(fj, d2, (1.0f - fj) * d1);
                u = (u * g.dz) * dw;
                total[r] = total[r] + u;
            }
        }
    }
}

template <typename TP>
__global__ __launch_bounds__(256) void sbp_kernel(DevGeom g, SlotTab S, int n_slots,
                                                  const TP* __restrict__ prj,
                                                  float* __restrict__ img) {
    const TileCoord c = tile_coord(g);
    // whole waves past the last row leave; lanes past the last column stay
    // (they evaluate row terms other lanes read) and only skip the stores
    if (c.iy0 >= g.ny) return;
    float total[ROWS];
    sbp_rows<TP>(g, S, n_slots, prj, c.ix, c.iy0, c.iz, total);
    if (c.ix >= g.nx) return;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int iy = c.iy0 + r;
        if (iy >= g.ny) break;
        const int64_t vi = ((int64_t)c.iz * g.ny + iy) * g.nx + c.ix;
        img[vi] = img[vi] + total[r];
    }
}

template <int POT, typename TI, typename TP>
__global__ __launch_bounds__(256) void update_fused_kernel(
    DevGeom g, SlotTab S, int n_slots, const TP* __restrict__ hsub, const TI* __restrict__ x_in,
    TI* __restrict__ x_out, const float* __restrict__ d, PriorW W, float inv_delta, float beta,
    float capl_scale, int guard) {
    const TileCoord c = tile_coord(g);
    if (c.iy0 >= g.ny) return;  // see sbp_kernel
    float total[ROWS];
    sbp_rows<TP>(g, S, n_slots, hsub, c.ix, c.iy0, c.iz, total);
    if (c.ix >= g.nx) return;
    const float two_beta = 2.0f * beta;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int iy = c.iy0 + r;
        if (iy >= g.ny) break;
        const int64_t vi = ((int64_t)c.iz * g.ny + iy) * g.nx + c.ix;
        // capL: 0 + total, then * M (float), as the reference host does
        const float capl = capl_scale * (0.0f + total[r]);
        float rdd = 0.0f, rd = 0.0f;
        prior_voxel<POT, TI>(g, x_in, c.ix, iy, c.iz, W, inv_delta, rdd, rd);
        const float xv = ld<TI>(x_in, vi);
        const float num = capl + beta * rd;
        const float den = d[vi] + two_beta * rdd;
        float xn;
        if (guard && den == 0.0f) {
            xn = xv;
        } else {
            xn = xv - num / den;
            xn = xn < 0.0f ? 0.0f : xn;
        }
        st<TI>(x_out, vi, xn);
    }
}

// [slot][nbx][nby] -> [slot][nby][nbx], 64x64 tiles through LDS
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const T* __restrict__ src,
                                                        T* __restrict__ dst, int nbx, int nby) {
    __shared__ T tile[64][65];
    const int slot = blockIdx.z;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    const T* s = src + (int64_t)slot * nbx * nby;
    T* o = dst + (int64_t)slot * nbx * nby;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = w; k < 64; k += 4) {
        const int i = i0 + k, j = j0 + lane;
        if (i < nbx && j < nby) tile[k][lane] = s[(int64_t)i * nby + j];
    }
    __syncthreads();
    for (int k = w; k < 64; k += 4) {
        const int j = j0 + k, i = i0 + lane;
        if (i < nbx && j < nby) o[(int64_t)j * nbx + i] = tile[lane][k];
    }
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

unsigned volume_tiles(const DevGeom& g) {
    const int64_t tx = (g.nx + TILE_X - 1) / TILE_X, ty = (g.ny + TILE_Y - 1) / TILE_Y;
    return (unsigned)(tx * ty * g.nz);
}

template <typename TP>
hipError_t sbp_t(const DevGeom& g, const SlotTab& S, int n, float* img, const void* prj,
                 hipStream_t st) {
    hipLaunchKernelGGL((sbp_kernel<TP>), dim3(volume_tiles(g)), dim3(256), 0, st, g, S, n,
                       (const TP*)prj, img);
    return hipGetLastError();
}

template <int POT, typename TI, typename TP>
hipError_t fused_t(const DevGeom& g, const SlotTab& S, int n, const void* x_in, void* x_out,
                   const float* d, const void* hsub, const PriorW& W, float inv_delta,
                   float beta, float capl_scale, int guard, hipStream_t st) {
    hipLaunchKernelGGL((update_fused_kernel<POT, TI, TP>), dim3(volume_tiles(g)), dim3(256), 0,
                       st, g, S, n, (const TP*)hsub, (const TI*)x_in, (TI*)x_out, d, W,
                       inv_delta, beta, capl_scale, guard);
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

size_t elem_size(int dtype) { return dtype == OSTR_F32 ? 4 : 2; }

hipError_t launch_transpose_bins(const void* src, void* dst, int n_slots, int nbx, int nby,
                                 int dtype, hipStream_t st) {
    if (n_slots <= 0) return hipSuccess;
    dim3 grid((nby + 63) / 64, (nbx + 63) / 64, n_slots);
    if (dtype == OSTR_F32)
        hipLaunchKernelGGL((transpose_kernel<float>), grid, dim3(256), 0, st, (const float*)src,
                           (float*)dst, nbx, nby);
    else
        hipLaunchKernelGGL((transpose_kernel<uint16_t>), grid, dim3(256), 0, st,
                           (const uint16_t*)src, (uint16_t*)dst, nbx, nby);
    return hipGetLastError();
}

// `views` is the HOST copy of the slot table (slot parameters go to the kernel
// by value). prj is detector-row-major [slot][nby][nbx].
hipError_t launch_sbp_rows(const DevGeom& g, const ViewTab* views, int a0, int a1,
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
