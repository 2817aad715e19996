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
// (64 x 32 voxels of one slice) are dealt to XCDs in contiguous ranges.
//
// The fused update computes, per voxel and in one pass,
//   capL = M * SBP(h)                      (Tomo_OSTR_gold.cpp:759-771)
//   RD, RDD = prior(x_in)                  (:775)
//   x_out = max(0, x - (capL + b RD)/(d + 2 b RDD))   (update_est :265-295)
// reading x_in and writing x_out (double-buffered, so the prior sees the
// pre-update image as the reference's sequential order does).
#include <type_traits>

#include "matched_device.h"
#include "prior_device.h"

namespace ostr {

namespace {

// Tile shape of the single-case kernels: CASE_ROWS voxel rows (y) per wave,
// CASE_WAVES waves per work-group, so a work-group covers 64 x 32 voxels of
// one slice. The batched kernels take the shape as template arguments (R, NW).
constexpr int CASE_ROWS = 8;
constexpr int CASE_WAVES = 4;
constexpr int TILE_X = 64;
// prior_rows modes (by_prior_mode)
constexpr int PRIOR_INTERIOR = 0, PRIOR_EDGE = 1, PRIOR_GENERIC = 2;

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

// Voxel coordinates of this thread in volume tile t (x tiles fastest, then y
// tiles, then slices); the wave owns rows iy0 .. iy0+R-1.
template <int R, int NW>
__device__ __forceinline__ TileCoord tile_coord(const DevGeom& g, int t) {
    constexpr int TILE_Y = R * NW;
    const int tiles_x = (g.nx + TILE_X - 1) / TILE_X;
    const int tiles_y = (g.ny + TILE_Y - 1) / TILE_Y;
    const int bx = t % tiles_x;
    const int rest = t / tiles_x;
    TileCoord c;
    c.ix = bx * TILE_X + (threadIdx.x & 63);
    c.iy0 = (rest % tiles_y) * TILE_Y + (threadIdx.x >> 6) * R;
    c.iz = rest / tiles_y;
    return c;
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// total[r] = sum over slots of the SBP value of voxel (ix, iy0 + r, iz);
// prj is [slot][nby][nbx], read through a buffer resource with 32-bit element
// offsets. The row terms (jlb, fracJ, y offset) of a (row, slot) pair are
// wave-uniform: lane r + ROWS*k evaluates them for row r and slot k of a chunk
// of VIEW_CHUNK slots, and every lane reads them back with readlane, so each
// exact division runs once per wave. When every lane's two tap columns and the
// row's two tap rows lie on the detector (the common case), the four taps are
// loaded without masks; otherwise each tap is masked on its own, as the
// reference does.
// SENS: also sens[r] = sum over slots of the same bilinear weights applied to
// the valid taps' indicator (the OS-SART sensitivity, backprj_OSSART_gpu_
// manyviews, Tomo_OSTR.cu:1856-1888), scaled like the data term.
template <typename TP, int ROWS, bool SENS = false>
__device__ __forceinline__ void sbp_rows(const DevGeom& g, const SlotTab& S, int n,
                                         const VolReader<TP>& P, int ix, int iy0, int iz,
                                         float total[ROWS], float* sens = nullptr) {
    constexpr int VIEW_CHUNK = 64 / ROWS;
    const uint32_t nbin = (uint32_t)g.nbx * (uint32_t)g.nby;
    const uint32_t nbx = (uint32_t)g.nbx;
    const int lane = threadIdx.x & 63;
    const float y0 = g.src_y;
    const float x1 = cell_centre(g.x_p0, ix, g.dx);
    const float z1 = cell_centre(g.z_p0, iz, g.dz);
    // slot parameters: lane a holds slot a (n <= 64); the loops below take
    // them with readlane (uniform) or shfl (per lane), no memory reads
    const float sx0 = lane < n ? S.x0[lane] : 0.0f;
    const float sz0 = lane < n ? S.z0[lane] : 0.0f;
    const float snum = lane < n ? S.num[lane] : 0.0f;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        total[r] = 0.0f;
        if (SENS) sens[r] = 0.0f;
    }
    for (int c0 = 0; c0 < n; c0 += VIEW_CHUNK) {
        int rt_jl = -2;
        float rt_fj = 0.0f, rt_ey = 0.0f, rt_t = 0.0f;
        {
            const int r = lane % ROWS, a = c0 + lane / ROWS;
            const float z0 = __shfl(sz0, a & 63, 64);
            if (a < n) {
                const float y1 = cell_centre(g.y_p0, iy0 + r, g.dy);
                rt_ey = y0 - y1;
                if (z1 != z0) {
                    const float t = (-g.det_r - z0) / (z1 - z0);
                    rt_t = t;
                    const float y2 = fmaf(y1 - y0, t, y0);
                    tap((y2 - g.y_d0) / g.dby, rt_jl, rt_fj);
                }
            }
        }
        const int na = n - c0 < VIEW_CHUNK ? n - c0 : VIEW_CHUNK;
        // y0 - y1 of row r does not depend on the view: read once per chunk
        float eyr[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) eyr[r] = readlane_f(rt_ey, r);
        for (int k = 0; k < na; ++k) {
            const int a = c0 + k;
            const float x0 = readlane_f(sx0, a), z0 = readlane_f(sz0, a);
            const float num = readlane_f(snum, a);
            if (z1 == z0) continue;
            const float ex = x0 - x1, ez = z0 - z1;
            const float exx = ex * ex;
            // the view's ray parameter (the slice's, the same for every lane):
            // row 0's row-term lane divided it already
            const float t = readlane_f(rt_t, k * ROWS);
            const float x2 = fmaf(x1 - x0, t, x0);
            int il;
            float fi;
            tap((x2 - g.x_d0) / g.dbx, il, fi);
            const bool i0 = (unsigned)il < (unsigned)g.nbx;
            const bool i1 = (unsigned)(il + 1) < (unsigned)g.nbx;
            const bool wave_i_ok = __all(i0 && i1);
            const float omfi = 1.0f - fi;
            const uint32_t base = (uint32_t)a * nbin + (uint32_t)il;
            // Common case: every lane's two columns and every row's two rows
            // on the detector. Then the taps of all ROWS rows are loaded in
            // one straight run (2 ROWS loads in flight) before any is used;
            // with a per-row choice of load form each row's loads would wait
            // for the previous row's. Same operations in the same order as
            // the per-row form below, so the same bits.
            bool all_rows = wave_i_ok;
            int jlr[ROWS];  // wave-uniform
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                const int jl = __builtin_amdgcn_readlane(rt_jl, k * ROWS + r);
                all_rows = all_rows && (unsigned)jl < (unsigned)g.nby &&
                           (unsigned)(jl + 1) < (unsigned)g.nby &&
                           (sizeof(TP) == 4 || a + 1 < n || jl + 2 < g.nby);
                jlr[r] = jl;
            }
            if (all_rows) {
                float p[ROWS][4];
                if constexpr (sizeof(TP) == 4) {
                    // the row's offset jl nbx is wave-uniform: it rides in the
                    // scalar offset, the lane's column in the vector offset
                    const uint32_t vb = base * 4u;
#pragma unroll
                    for (int r = 0; r < ROWS; ++r) {
                        const uint32_t jl = (uint32_t)jlr[r];
                        P.pair_bytes(vb, jl * nbx * 4u, p[r][0], p[r][1]);
                        P.pair_bytes(vb, (jl + 1u) * nbx * 4u, p[r][2], p[r][3]);
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < ROWS; ++r) {
                        const uint32_t off = base + (uint32_t)jlr[r] * nbx;
                        P.pair(off, p[r][0], p[r][1]);
                        P.pair(off + nbx, p[r][2], p[r][3]);
                    }
                }
#pragma unroll
                for (int r = 0; r < ROWS; ++r) {
                    const int src = k * ROWS + r;
                    const float fj = readlane_f(rt_fj, src);
                    const float ey = eyr[r];
                    const float den = fmaf(ez, ez, fmaf(ey, ey, exx));
                    const float dw = num * __builtin_amdgcn_rcpf(den);
                    float d1 = omfi * p[r][0];
                    d1 = fmaf(fi, p[r][1], d1);
                    float d2 = omfi * p[r][2];
                    d2 = fmaf(fi, p[r][3], d2);
                    float u = fmaf(fj, d2, (1.0f - fj) * d1);
                    u = (u * g.dz) * dw;
                    total[r] = total[r] + u;
                    if (SENS) {
                        // every tap valid: d1_sen = d2_sen = (1-fI) + fI
                        const float d1s = omfi + fi;
                        float us = fmaf(fj, d1s, (1.0f - fj) * d1s);
                        us = (us * g.dz) * dw;
                        sens[r] = sens[r] + us;
                    }
                }
                continue;
            }
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
                const uint32_t off = base + (uint32_t)jl * nbx;
                float p00, p10, p01, p11;
                // both columns and rows on the detector: the two taps of a
                // row as one pair (one load for 16-bit h; the pair of row
                // jl + 1 may read one element past its end, which exists
                // unless that row is the last of the last slot)
                if (wave_i_ok && j0 && j1 &&
                    (sizeof(TP) == 4 || a + 1 < n || jl + 2 < g.nby)) {
                    P.pair(off, p00, p10);
                    P.pair(off + nbx, p01, p11);
                } else if (wave_i_ok && j0 && j1) {
                    p00 = P(off);
                    p10 = P(off + 1);
                    p01 = P(off + nbx);
                    p11 = P(off + nbx + 1);
                } else {
                    p00 = (i0 && j0) ? P(off) : 0.0f;
                    p10 = (i1 && j0) ? P(off + 1) : 0.0f;
                    p01 = (i0 && j1) ? P(off + nbx) : 0.0f;
                    p11 = (i1 && j1) ? P(off + nbx + 1) : 0.0f;
                }
                float d1 = omfi * p00;
                d1 = fmaf(fi, p10, d1);
                float d2 = omfi * p01;
                d2 = fmaf(fi, p11, d2);
                float u = fmaf(fj, d2, (1.0f - fj) * d1);
                u = (u * g.dz) * dw;
                total[r] = total[r] + u;
                if (SENS) {
                    // d1_sen = (1-fI)[tap 00] + fI[tap 10], likewise d2_sen
                    const float d1s = ((i0 && j0) ? omfi : 0.0f) + ((i1 && j0) ? fi : 0.0f);
                    const float d2s = ((i0 && j1) ? omfi : 0.0f) + ((i1 && j1) ? fi : 0.0f);
                    float us = fmaf(fj, d2s, (1.0f - fj) * d1s);
                    us = (us * g.dz) * dw;
                    sens[r] = sens[r] + us;
                }
            }
        }
    }
}

template <typename TP>
__global__ __launch_bounds__(256) void sbp_kernel(DevGeom g, SlotTab S, int n_slots,
                                                  const TP* __restrict__ prj,
                                                  float* __restrict__ img) {
    const TileCoord c = tile_coord<CASE_ROWS, CASE_WAVES>(g, xcd_tile(blockIdx.x, gridDim.x));
    // whole waves past the last row leave; lanes past the last column stay
    // (they evaluate row terms other lanes read) and only skip the stores
    if (c.iy0 >= g.ny) return;
    const VolReader<TP> P(prj, (uint32_t)n_slots * g.nbx * g.nby);
    float total[CASE_ROWS];
    sbp_rows<TP, CASE_ROWS>(g, S, n_slots, P, c.ix, c.iy0, c.iz, total);
    if (c.ix >= g.nx) return;
#pragma unroll
    for (int r = 0; r < CASE_ROWS; ++r) {
        const int iy = c.iy0 + r;
        if (iy >= g.ny) break;
        const int64_t vi = ((int64_t)c.iz * g.ny + iy) * g.nx + c.ix;
        img[vi] = img[vi] + total[r];
    }
}

// OS-SART backprojection and update (backprj_OSSART_gpu_manyviews and the _R
// variant with a prior, Tomo_OSTR.cu:1737-2143): per voxel, u = sum / sens
// over the launch's slots (0 where sens == 0), then, in place,
//   x = max(0, x + lambda * u)                          (PRIOR false)
//   x = max(0, x + lambda * (u + beta * prior / sens))  (PRIOR true; the sign is
//                                                        the reference's)
template <typename TP, bool PRIOR>
__global__ __launch_bounds__(256) void sart_update_kernel(DevGeom g, SlotTab S, int n_slots,
                                                          const TP* __restrict__ diff,
                                                          float* __restrict__ img,
                                                          const float* __restrict__ prior,
                                                          float lambda, float beta) {
    const TileCoord c = tile_coord<CASE_ROWS, CASE_WAVES>(g, xcd_tile(blockIdx.x, gridDim.x));
    if (c.iy0 >= g.ny) return;  // see sbp_kernel
    const VolReader<TP> P(diff, (uint32_t)n_slots * g.nbx * g.nby);
    float total[CASE_ROWS], sens[CASE_ROWS];
    sbp_rows<TP, CASE_ROWS, true>(g, S, n_slots, P, c.ix, c.iy0, c.iz, total, sens);
    if (c.ix >= g.nx) return;
#pragma unroll
    for (int r = 0; r < CASE_ROWS; ++r) {
        const int iy = c.iy0 + r;
        if (iy >= g.ny) break;
        const int64_t vi = ((int64_t)c.iz * g.ny + iy) * g.nx + c.ix;
        float u = 0.0f, bt = 0.0f;
        if (sens[r] != 0.0f) {
            u = total[r] / sens[r];
            if (PRIOR) bt = (beta * prior[vi]) / sens[r];
        }
        const float xn = fmaf(lambda, PRIOR ? u + bt : u, img[vi]);
        img[vi] = xn < 0.0f ? 0.0f : xn;
    }
}

// Prior terms of the ROWS voxels (ix, iy0 + r, iz), r = 0..ROWS-1, from a
// sliding 3x3x3 window: each step loads the 9 values (3 slices x 3 columns) of
// the next row, so a thread reads 9*(ROWS+2) values instead of 26*ROWS. The
// 26 neighbours are visited in the reference's order (dz outer, dy, dx inner)
// and summed per weight class: a[c] = sum r(t), b[c] = sum r(t) t (prior_r:
// omega = |delta| r with |delta| in the class weights), then
// rdd = sum_c w_c a[c], rd = init(r) + sum_c w_c b[c], the same terms
// w_k omega(t) and w_k omega(t) t as the reference's per-neighbour sums with
// one weight multiply per class instead of one per neighbour and no t/delta;
// then epi(r, vi, x_centre, rdd, rd) is called for each voxel inside the
// volume. MODE (prior_mode): PRIOR_INTERIOR, the window stays inside the
// volume; otherwise a neighbour outside the volume contributes t = 0 (it is
// replaced by the voxel itself, Tomo_OSTR.cu:238-241), and PRIOR_GENERIC also
// takes k < 0.
template <int POT, int MODE, typename TI, int ROWS, typename Init, typename Epi>
__device__ __forceinline__ void prior_rows(const DevGeom& g, const VolReader<TI>& X,
                                           const TileCoord& c, const PriorClass& W,
                                           float inv_delta, Init init, Epi epi) {
    constexpr bool CHECKED = MODE != PRIOR_INTERIOR;
    const int ix = c.ix, iy0 = c.iy0, iz = c.iz;
    const uint32_t nx = g.nx, plane = (uint32_t)g.nx * g.ny;
    bool cx[3], cz[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        cx[q] = (unsigned)(ix + q - 1) < (unsigned)g.nx;
        cz[q] = (unsigned)(iz + q - 1) < (unsigned)g.nz;
    }
    float win[3][3][3];  // [row dy][slice dz][column dx]
    bool ok[3][3][3];
    auto load_row = [&](int slot, int y) {
        const bool cy = (unsigned)y < (unsigned)g.ny;
#pragma unroll
        for (int dz = 0; dz < 3; ++dz) {
            const uint32_t o = (uint32_t)(iz + dz - 1) * plane + (uint32_t)y * nx + (uint32_t)ix;
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                if (CHECKED) {
                    ok[slot][dz][dx] = cy && cz[dz] && cx[dx];
                    win[slot][dz][dx] = ok[slot][dz][dx] ? X(o + dx - 1) : 0.0f;
                } else {
                    win[slot][dz][dx] = X(o + dx - 1);
                }
            }
        }
    };
    load_row(0, iy0 - 1);
    load_row(1, iy0);
    const bool lane_ok = ix < g.nx;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        load_row(2, iy0 + r + 1);
        const float xc = win[1][1][1];
        const int iy = iy0 + r;
        const bool store = lane_ok && iy < g.ny;
        float a[5], b[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) a[k] = b[k] = 0.0f;
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const int nidx = dz * 9 + dy * 3 + dx;
                    if (nidx == 13) continue;
                    float t = xc - win[dy][dz][dx];
                    if (CHECKED) t = ok[dy][dz][dx] ? t : 0.0f;
                    const float rt = prior_r<POT, MODE != PRIOR_GENERIC>(t, W.k, inv_delta);
                    const int k = prior_class(nidx);
                    a[k] = a[k] + rt;
                    b[k] = fmaf(rt, t, b[k]);
                }
        float rdd = 0.0f, rd = init(r, store);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            rdd = fmaf(W.w[k], a[k], rdd);
            rd = fmaf(W.w[k], b[k], rd);
        }
        if (store) epi(r, (uint32_t)iz * plane + (uint32_t)iy * nx + (uint32_t)ix, xc, rdd, rd);
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                win[0][dz][dx] = win[1][dz][dx];
                win[1][dz][dx] = win[2][dz][dx];
                if (CHECKED) {
                    ok[0][dz][dx] = ok[1][dz][dx];
                    ok[1][dz][dx] = ok[2][dz][dx];
                }
            }
    }
}

// Calls f(std::integral_constant<int, MODE>) with the prior_rows mode of the
// wave (wave-uniform): PRIOR_INTERIOR when the 3x3x3 neighbourhoods of its R
// rows stay inside the volume and k >= 0, PRIOR_EDGE when only k >= 0,
// PRIOR_GENERIC otherwise.
template <int R, typename F>
__device__ __forceinline__ void by_prior_mode(const DevGeom& g, const TileCoord& c,
                                              const PriorClass& W, F f) {
    const int x_first = c.ix - (threadIdx.x & 63);
    const bool inside = x_first >= 1 && x_first + 64 < g.nx && c.iy0 >= 1 &&
                        c.iy0 + R < g.ny && c.iz >= 1 && c.iz + 1 < g.nz;
    if (W.k < 0.0f)
        f(std::integral_constant<int, PRIOR_GENERIC>());
    else if (inside)
        f(std::integral_constant<int, PRIOR_INTERIOR>());
    else
        f(std::integral_constant<int, PRIOR_EDGE>());
}

// Prior terms and SQS update of the ROWS voxels (the OSTR epilogue of the
// fused update): capL = M (0 + total), x = max(0, x - (capL + beta rd) /
// (d + 2 beta rdd)); guard: keep x where the denominator is 0.
template <int POT, int MODE, typename TI, int ROWS>
__device__ __forceinline__ void update_rows(const DevGeom& g, const VolReader<TI>& X,
                                            const VolWriter<TI>& XO, const VolReader<float>& D,
                                            const TileCoord& c, const float total[ROWS],
                                            const PriorClass& W, float inv_delta, float beta,
                                            float capl_scale, int guard) {
    const float two_beta = 2.0f * beta;
    prior_rows<POT, MODE, TI, ROWS>(
        g, X, c, W, inv_delta, [](int, bool) { return 0.0f; },
        [&](int r, uint32_t vi, float xc, float rdd, float rd) {
            // capL: 0 + total, then * M (float), as the reference host does
            const float capl = capl_scale * (0.0f + total[r]);
            const float num = capl + beta * rd;
            const float den = D(vi) + two_beta * rdd;
            float xn;
            if (guard && den == 0.0f) {
                xn = xc;
            } else {
                xn = xc - num / den;
                xn = xn < 0.0f ? 0.0f : xn;
            }
            XO(vi, xn);
        });
}

// The SART prior G_Huber_prior (Tomo_OSTR.cu:258-325), accumulating:
//   prior[v] += sum over the 26 neighbours of w d / sqrt(1 + (d/delta)^2),
// d = x[v] - x[neighbour] (0 for neighbours outside the volume), w = 1/distance
// in voxel-index units (no z-size ratio). This is the hyperbola gradient term
// rd of the OSTR prior with rho = 1, evaluated in FP32 (rsq of 1 + q^2, the
// terms added in the reference's order from prior[v]); the reference forms
// each term in double and rounds the sum to float after every term.
__global__ __launch_bounds__(256) void huber_prior_kernel(DevGeom g, PriorClass W,
                                                          float inv_delta,
                                                          const float* __restrict__ x,
                                                          float* __restrict__ prior) {
    const TileCoord c = tile_coord<CASE_ROWS, CASE_WAVES>(g, xcd_tile(blockIdx.x, gridDim.x));
    if (c.iy0 >= g.ny) return;  // see sbp_kernel
    const uint32_t nvox = (uint32_t)g.nx * g.ny * g.nz;
    const VolReader<float> X(x, nvox);
    const VolReader<float> PR(prior, nvox);
    const VolWriter<float> PW(prior, nvox);
    const uint32_t row0 = ((uint32_t)c.iz * g.ny + (uint32_t)c.iy0) * g.nx + (uint32_t)c.ix;
    auto init = [&](int r, bool store) { return store ? PR(row0 + (uint32_t)r * g.nx) : 0.0f; };
    auto epi = [&](int, uint32_t vi, float, float, float rd) { PW(vi, rd); };
    by_prior_mode<CASE_ROWS>(g, c, W, [&](auto m) {
        prior_rows<OSTR_PRIOR_HYPERBOLA, decltype(m)::value, float, CASE_ROWS>(g, X, c, W,
                                                                              inv_delta, init, epi);
    });
}

// Fused OS-SART subset update (the backprj_OSSART[_R] update with the Huber
// prior evaluated in the same pass): per voxel, total and sens as
// sart_update_kernel, prior = 0 + the Huber terms of x_in, then
// x_out = max(0, x_in + lambda (total/sens [+ beta prior/sens])). Reading
// x_in and writing x_out keeps the prior's neighbours at their pre-update
// values, as the separate prior launch does.
template <bool PRIOR>
__global__ __launch_bounds__(256) void sart_update_fused_kernel(
    DevGeom g, SlotTab S, int n_slots, const float* __restrict__ diff,
    const float* __restrict__ x_in, float* __restrict__ x_out, PriorClass W, float inv_delta,
    float lambda, float beta) {
    const TileCoord c = tile_coord<CASE_ROWS, CASE_WAVES>(g, xcd_tile(blockIdx.x, gridDim.x));
    if (c.iy0 >= g.ny) return;  // see sbp_kernel
    const uint32_t nvox = (uint32_t)g.nx * g.ny * g.nz;
    const VolReader<float> P(diff, (uint32_t)n_slots * g.nbx * g.nby);
    const VolReader<float> X(x_in, nvox);
    const VolWriter<float> XO(x_out, nvox);
    float total[CASE_ROWS], sens[CASE_ROWS];
    sbp_rows<float, CASE_ROWS, true>(g, S, n_slots, P, c.ix, c.iy0, c.iz, total, sens);
    auto epi = [&](int r, uint32_t vi, float xc, float, float rd) {
        float u = 0.0f, bt = 0.0f;
        if (sens[r] != 0.0f) {
            u = total[r] / sens[r];
            if (PRIOR) bt = (beta * rd) / sens[r];
        }
        const float xn = fmaf(lambda, PRIOR ? u + bt : u, xc);
        XO(vi, xn < 0.0f ? 0.0f : xn);
    };
    if (PRIOR) {
        auto init = [](int, bool) { return 0.0f; };
        by_prior_mode<CASE_ROWS>(g, c, W, [&](auto m) {
            prior_rows<OSTR_PRIOR_HYPERBOLA, decltype(m)::value, float, CASE_ROWS>(
                g, X, c, W, inv_delta, init, epi);
        });
    } else {
        if (c.ix >= g.nx) return;
        const uint32_t plane = (uint32_t)g.nx * g.ny;
#pragma unroll
        for (int r = 0; r < CASE_ROWS; ++r) {
            const int iy = c.iy0 + r;
            if (iy >= g.ny) break;
            const uint32_t vi = (uint32_t)c.iz * plane + (uint32_t)iy * g.nx + (uint32_t)c.ix;
            epi(r, vi, X(vi), 0.0f, 0.0f);
        }
    }
}

// Fused update of volume tile t (R rows per wave, NW waves per work-group).
template <int POT, typename TI, typename TP, int R, int NW>
__device__ __forceinline__ void update_tile(const DevGeom& g, const SlotTab& S, int n_slots,
                                            const TP* __restrict__ hsub,
                                            const TI* __restrict__ x_in, TI* __restrict__ x_out,
                                            const float* __restrict__ d, const PriorClass& W,
                                            float inv_delta, float beta, float capl_scale,
                                            int guard, int t) {
    const TileCoord c = tile_coord<R, NW>(g, t);
    if (c.iy0 >= g.ny) return;  // see sbp_kernel
    const uint32_t nvox = (uint32_t)g.nx * g.ny * g.nz;
    const VolReader<TP> P(hsub, (uint32_t)n_slots * g.nbx * g.nby);
    const VolReader<TI> X(x_in, nvox);
    const VolWriter<TI> XO(x_out, nvox);
    const VolReader<float> D(d, nvox);
    float total[R];
    sbp_rows<TP, R>(g, S, n_slots, P, c.ix, c.iy0, c.iz, total);
    by_prior_mode<R>(g, c, W, [&](auto m) {
        update_rows<POT, decltype(m)::value, TI, R>(g, X, XO, D, c, total, W, inv_delta, beta,
                                                    capl_scale, guard);
    });
}

// Fused OSTR update with the exact transpose (the matched counterpart of
// update_fused_kernel): per voxel capL = M (0 + sum_views cos sum_i a_ij h_i)
// from matched_voxel, then the prior window on x_in and the SQS update into
// x_out, the same epilogue (update_rows). h detector-row-major. One row per
// lane (MATCHED_ROWS = 1): the gather dominates and is latency-bound, so the
// launch keeps as many waves as the standalone transpose.
constexpr int MATCHED_ROWS = 1;

template <int POT, typename TI, typename TP>
__global__ __launch_bounds__(256) void update_fused_matched_kernel(
    DevGeom g, MatchedSlots S, int n_slots, const TP* __restrict__ hsub,
    const TI* __restrict__ x_in, TI* __restrict__ x_out, const float* __restrict__ d,
    PriorClass W, float inv_delta, float beta, float capl_scale, int guard) {
    constexpr int R = MATCHED_ROWS;
    const TileCoord c = tile_coord<R, CASE_WAVES>(g, xcd_tile(blockIdx.x, gridDim.x));
    if (c.iy0 >= g.ny) return;  // see sbp_kernel
    const uint32_t nvox = (uint32_t)g.nx * g.ny * g.nz;
    const VolReader<TP> P(hsub, (uint32_t)n_slots * g.nbx * g.nby);
    const VolReader<TI> X(x_in, nvox);
    const VolWriter<TI> XO(x_out, nvox);
    const VolReader<float> D(d, nvox);
    float total[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        total[r] = (c.ix < g.nx && c.iy0 + r < g.ny)
                       ? matched_voxel<TP>(g, S, n_slots, 1, P, c.ix, c.iy0 + r, c.iz)
                       : 0.0f;
    by_prior_mode<R>(g, c, W, [&](auto m) {
        update_rows<POT, decltype(m)::value, TI, R>(g, X, XO, D, c, total, W, inv_delta,
                                                    beta, capl_scale, guard);
    });
}

template <int POT, typename TI, typename TP>
__global__ __launch_bounds__(256) void update_fused_kernel(
    DevGeom g, SlotTab S, int n_slots, const TP* __restrict__ hsub, const TI* __restrict__ x_in,
    TI* __restrict__ x_out, const float* __restrict__ d, PriorClass W, float inv_delta,
    float beta, float capl_scale, int guard) {
    update_tile<POT, TI, TP, CASE_ROWS, CASE_WAVES>(g, S, n_slots, hsub, x_in, x_out, d, W,
                                                    inv_delta, beta, capl_scale, guard,
                                                    xcd_tile(blockIdx.x, gridDim.x));
}

// Batched fused update: every case of the table in one launch. LOOP: a
// persistent grid (a fixed number of groups per CU, each looping over its
// share of the tiles, see BatchWalk); otherwise the grid gives every XCD at
// least as many groups as it has tiles and each group takes at most one, so
// the walk's state is not live across the tile (scalar registers the tile
// would otherwise spill). The case's scalars are read from the read-only case
// table with uniform addresses (scalar loads).
template <int POT, typename TI, typename TP, int R, int NW, bool LOOP>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(5))) void batch_update_kernel(
    const BatchCase* __restrict__ cases, const int* __restrict__ sched, int n_cases) {
    BatchWalk walk(sched, n_cases);
    int t, c;
    while (walk.next(t, c)) {
        const BatchCase& bc = cases[c];
        update_tile<POT, TI, TP, R, NW>(
            bc.g, bc.S, bc.n_slots, (const TP*)bc.h, (const TI*)bc.x_in, (TI*)bc.x_out, bc.d,
            bc.W, bc.inv_delta, bc.beta, bc.capl_scale, bc.guard, t);
        if (!LOOP) break;
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

unsigned volume_tiles(const DevGeom& g, int tile_y = CASE_ROWS * CASE_WAVES) {
    const int64_t tx = (g.nx + TILE_X - 1) / TILE_X, ty = (g.ny + tile_y - 1) / tile_y;
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
                   const float* d, const void* hsub, const PriorClass& W, float inv_delta,
                   float beta, float capl_scale, int guard, hipStream_t st) {
    hipLaunchKernelGGL((update_fused_kernel<POT, TI, TP>), dim3(volume_tiles(g)), dim3(256), 0,
                       st, g, S, n, (const TP*)hsub, (const TI*)x_in, (TI*)x_out, d, W,
                       inv_delta, beta, capl_scale, guard);
    return hipGetLastError();
}

template <int POT, typename TI>
hipError_t fused_p(int prj_dtype, const DevGeom& g, const SlotTab& S, int n, const void* x_in,
                   void* x_out, const float* d, const void* hsub, const PriorClass& W,
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
                   const PriorClass& W, float inv_delta, float beta, float capl_scale,
                   int guard, hipStream_t st) {
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

template <int POT, typename TP, int R, int NW>
hipError_t batch_upd_t(const BatchCase* cases, const int* sched, int n_cases, int total, int grid,
                       int single_pass, hipStream_t st) {
    if (single_pass)
        hipLaunchKernelGGL((batch_update_kernel<POT, float, TP, R, NW, false>), dim3(grid),
                           dim3(64 * NW), 0, st, cases, sched, n_cases);
    else
        hipLaunchKernelGGL((batch_update_kernel<POT, float, TP, R, NW, true>), dim3(grid),
                           dim3(64 * NW), 0, st, cases, sched, n_cases);
    return hipGetLastError();
}

template <int POT, typename TP>
hipError_t batch_upd_s(int rows, int waves, const BatchCase* cases, const int* sched, int n_cases, int total,
                       int grid, int single, hipStream_t st) {
    if (rows == 8 && waves == 4) return batch_upd_t<POT, TP, 8, 4>(cases, sched, n_cases, total, grid, single, st);
    if (rows == 8 && waves == 2) return batch_upd_t<POT, TP, 8, 2>(cases, sched, n_cases, total, grid, single, st);
    if (rows == 4 && waves == 4) return batch_upd_t<POT, TP, 4, 4>(cases, sched, n_cases, total, grid, single, st);
    return hipErrorInvalidValue;
}

template <int POT>
hipError_t batch_upd_p(int prj_dtype, int rows, int waves, const BatchCase* cases,
                       const int* sched, int n_cases, int total, int grid, int single,
                       hipStream_t st) {
    switch (prj_dtype) {
        case OSTR_F32: return batch_upd_s<POT, float>(rows, waves, cases, sched, n_cases, total, grid, single, st);
        case OSTR_BF16: return batch_upd_s<POT, bf16_t>(rows, waves, cases, sched, n_cases, total, grid, single, st);
        case OSTR_F16: return batch_upd_s<POT, _Float16>(rows, waves, cases, sched, n_cases, total, grid, single, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

bool batch_update_shape_ok(int rows, int waves) {
    return (rows == 8 && (waves == 4 || waves == 2)) || (rows == 4 && waves == 4);
}

int64_t batch_upd_tiles(const DevGeom& g, int rows, int waves) {
    return volume_tiles(g, rows * waves);
}

void batch_slots(SlotTab& S, ViewTab* v, const ViewTab* host_views, int a0, int a1,
                 float src_y) {
    fill_slots(S, host_views, a0, a1, src_y);
    for (int a = a0; a < a1; ++a) v[a - a0] = host_views[a];
}

hipError_t launch_batch_update(const BatchCase* cases, const int* sched, int n_cases, int total, int grid,
                               int rows, int waves, int potential, int prj_dtype,
                               int single_pass, hipStream_t st) {
    if (total == 0) return hipSuccess;
    switch (potential) {
        case OSTR_PRIOR_Q2:
            return batch_upd_p<OSTR_PRIOR_Q2>(prj_dtype, rows, waves, cases, sched, n_cases, total, grid, single_pass, st);
        case OSTR_PRIOR_LANGE:
            return batch_upd_p<OSTR_PRIOR_LANGE>(prj_dtype, rows, waves, cases, sched, n_cases, total, grid, single_pass, st);
        case OSTR_PRIOR_HYPERBOLA:
            return batch_upd_p<OSTR_PRIOR_HYPERBOLA>(prj_dtype, rows, waves, cases, sched, n_cases, total, grid, single_pass, st);
    }
    return hipErrorInvalidValue;
}

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
hipError_t launch_sart_update(const DevGeom& g, const ViewTab* views, int a0, int a1,
                              float* img, const float* diff_rows, const float* prior,
                              float lambda, float beta, hipStream_t st) {
    SlotTab S;
    if (!fill_slots(S, views, a0, a1, g.src_y)) return hipErrorInvalidValue;
    if (prior)
        hipLaunchKernelGGL((sart_update_kernel<float, true>), dim3(volume_tiles(g)), dim3(256),
                           0, st, g, S, a1 - a0, diff_rows, img, prior, lambda, beta);
    else
        hipLaunchKernelGGL((sart_update_kernel<float, false>), dim3(volume_tiles(g)),
                           dim3(256), 0, st, g, S, a1 - a0, diff_rows, img, nullptr, lambda,
                           0.0f);
    return hipGetLastError();
}

hipError_t launch_huber_prior(const DevGeom& g, float* prior, const float* img, float delta,
                              hipStream_t st) {
    PriorW W27;
    prior_weights(W27.w, 1.0f);
    hipLaunchKernelGGL(huber_prior_kernel, dim3(volume_tiles(g)), dim3(256), 0, st, g,
                       prior_params(W27, OSTR_PRIOR_HYPERBOLA, delta), 1.0f / delta, img,
                       prior);
    return hipGetLastError();
}

hipError_t launch_sart_update_fused(const DevGeom& g, const ViewTab* views, int a0, int a1,
                                    const float* x_in, float* x_out, const float* diff_rows,
                                    float lambda, float beta, float delta, hipStream_t st) {
    SlotTab S;
    if (!fill_slots(S, views, a0, a1, g.src_y)) return hipErrorInvalidValue;
    PriorW W27;
    prior_weights(W27.w, 1.0f);
    const PriorClass W = prior_params(W27, OSTR_PRIOR_HYPERBOLA, delta);
    if (beta != 0.0f)
        hipLaunchKernelGGL(sart_update_fused_kernel<true>, dim3(volume_tiles(g)), dim3(256), 0,
                           st, g, S, a1 - a0, diff_rows, x_in, x_out, W, 1.0f / delta, lambda,
                           beta);
    else
        hipLaunchKernelGGL(sart_update_fused_kernel<false>, dim3(volume_tiles(g)), dim3(256), 0,
                           st, g, S, a1 - a0, diff_rows, x_in, x_out, W, 0.0f, lambda, 0.0f);
    return hipGetLastError();
}

namespace {

template <int POT, typename TI, typename TP>
hipError_t fused_matched_t(const DevGeom& g, const MatchedSlots& S, int n, const void* x_in,
                           void* x_out, const float* d, const void* hsub, const PriorClass& W,
                           float inv_delta, float beta, float capl_scale, int guard,
                           hipStream_t st) {
    const dim3 grid(volume_tiles(g, MATCHED_ROWS * CASE_WAVES));
    hipLaunchKernelGGL((update_fused_matched_kernel<POT, TI, TP>), grid, dim3(256), 0, st, g, S,
                       n, (const TP*)hsub, (const TI*)x_in, (TI*)x_out, d, W, inv_delta, beta,
                       capl_scale, guard);
    return hipGetLastError();
}

template <int POT, typename TI>
hipError_t fused_matched_h(int prj_dtype, const DevGeom& g, const MatchedSlots& S, int n,
                           const void* x_in, void* x_out, const float* d, const void* hsub,
                           const PriorClass& W, float inv_delta, float beta, float capl_scale,
                           int guard, hipStream_t st) {
    switch (prj_dtype) {
        case OSTR_F32:
            return fused_matched_t<POT, TI, float>(g, S, n, x_in, x_out, d, hsub, W, inv_delta,
                                                   beta, capl_scale, guard, st);
        case OSTR_BF16:
            return fused_matched_t<POT, TI, bf16_t>(g, S, n, x_in, x_out, d, hsub, W, inv_delta,
                                                    beta, capl_scale, guard, st);
        case OSTR_F16:
            return fused_matched_t<POT, TI, _Float16>(g, S, n, x_in, x_out, d, hsub, W,
                                                      inv_delta, beta, capl_scale, guard, st);
        default:
            return hipErrorInvalidValue;
    }
}

template <int POT>
hipError_t fused_matched_x(int img_dtype, int prj_dtype, const DevGeom& g,
                           const MatchedSlots& S, int n, const void* x_in, void* x_out,
                           const float* d, const void* hsub, const PriorClass& W,
                           float inv_delta, float beta, float capl_scale, int guard,
                           hipStream_t st) {
    switch (img_dtype) {
        case OSTR_F32:
            return fused_matched_h<POT, float>(prj_dtype, g, S, n, x_in, x_out, d, hsub, W,
                                               inv_delta, beta, capl_scale, guard, st);
        case OSTR_BF16:
            return fused_matched_h<POT, bf16_t>(prj_dtype, g, S, n, x_in, x_out, d, hsub, W,
                                                inv_delta, beta, capl_scale, guard, st);
        case OSTR_F16:
            return fused_matched_h<POT, _Float16>(prj_dtype, g, S, n, x_in, x_out, d, hsub, W,
                                                  inv_delta, beta, capl_scale, guard, st);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_update_fused_matched(const DevGeom& g, const ViewTab* views, int a0, int a1,
                                       const void* x_in, void* x_out, int img_dtype,
                                       const float* d, const void* hsub, int prj_dtype,
                                       float beta, float capl_scale, int potential, float delta,
                                       float rho, int guard, hipStream_t st) {
    if (a1 - a0 > MAX_SLOTS || a1 < a0) return hipErrorInvalidValue;
    MatchedSlots S;
    for (int a = a0; a < a1; ++a) {
        S.x0[a - a0] = views[a].x0;
        S.z0[a - a0] = views[a].z0;
        S.c[a - a0] = views[a].c;
    }
    PriorW W27;
    prior_weights(W27.w, rho);
    const PriorClass W = prior_params(W27, potential, delta);
    const float inv_delta = potential == OSTR_PRIOR_Q2 ? 0.0f : 1.0f / delta;
    const int n = a1 - a0;
    switch (potential) {
        case OSTR_PRIOR_Q2:
            return fused_matched_x<OSTR_PRIOR_Q2>(img_dtype, prj_dtype, g, S, n, x_in, x_out, d,
                                                  hsub, W, inv_delta, beta, capl_scale, guard,
                                                  st);
        case OSTR_PRIOR_LANGE:
            return fused_matched_x<OSTR_PRIOR_LANGE>(img_dtype, prj_dtype, g, S, n, x_in, x_out,
                                                     d, hsub, W, inv_delta, beta, capl_scale,
                                                     guard, st);
        case OSTR_PRIOR_HYPERBOLA:
            return fused_matched_x<OSTR_PRIOR_HYPERBOLA>(img_dtype, prj_dtype, g, S, n, x_in,
                                                         x_out, d, hsub, W, inv_delta, beta,
                                                         capl_scale, guard, st);
        default:
            return hipErrorInvalidValue;
    }
}

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
    PriorW W27;
    prior_weights(W27.w, rho);
    const PriorClass W = prior_params(W27, potential, delta);
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
