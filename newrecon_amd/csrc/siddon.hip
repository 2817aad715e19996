// siddon.hip - ray-driven exact Siddon forward projector family for gfx950.
//
// Semantics follow the reference walker rayTrace3D_GPU_direct_notexturememory_OSTR
// (Tomo_OSTR.cu:612-890) and the kernels ..._OSTR_cos / ..._normprj
// (Tomo_OSTR.cu:1313-1449): slab test, entry voxel, incremental plane
// crossings with exact tie handling, relative stop test and tail segment, the
// same fma placement (expressions `a*b + c` of the reference, which nvcc
// contracts) and no other contraction (built with -ffp-contract=off).
//
// MI355X mapping: one ray per lane, 64 lanes along detector x (the direction
// that is contiguous in the volume, so gathers of a wave fall in a few cache
// lines); a work-group is a 64 x 4 bin tile of one view slot; blocks of 16
// tiles (64 x 64 bins) are dealt round-robin to the XCDs (xcd_spread_tile), so
// each XCD gets an even share of the rays through the volume and neighbouring
// tiles of a block share an L2.
// Per-view trig comes from the handle's table (host FP64, rounded to float).
// The relative stop test (lmax - at)/lmax > 1e-6 is replaced by `at < thr`
// with thr computed once per ray as the exact float boundary of that test, so
// the loop has no division and takes the same steps.
#include "ostr_internal.h"

namespace ostr {

namespace {

constexpr int TILE_I = 64;   // bins along detector x per tile (one per lane)
constexpr int TILE_J = 4;    // bins along detector y per tile (one per wave)
constexpr int WALK_BATCH = 4;  // walk steps per batch (two batches' gathers in flight)

__device__ __forceinline__ float bin_centre(float origin, int i, float size) {
    return (float)((double)origin + ((double)i + 0.5) * (double)size);
}

struct Axis {
    float lo, hi, next, inc;
    int p, step;
};

__device__ __forceinline__ bool axis_limits(Axis& a, float s, float e, float p0, float ext,
                                            float top) {
    float d = e - s;
    if (d != 0.0f) {
        float lo = (p0 - s) / d;
        float hi = lo + ext / d;
        if (lo > hi) {
            a.lo = hi;
            a.hi = lo;
        } else {
            a.lo = lo;
            a.hi = hi;
        }
        return true;
    }
    if (s >= top || s <= p0) return false;
    a.lo = -1000.0f;
    a.hi = 1000.0f;
    return true;
}

__device__ __forceinline__ void axis_entry(Axis& a, float lmin, float s, float e, float p0,
                                           float h, int n, float hp, float last) {
    float d = e - s;
    if (a.lo != lmin) {
        a.p = (int)floorf((fmaf(lmin, d, s) - p0) / h);
        if (s < e)
            a.next = (fmaf((float)(a.p + 1), h, p0) - s) / d;
        else if (s == e)
            a.next = 1000.0f;
        else
            a.next = (fmaf((float)a.p, h, p0) - s) / d;
    } else if (s < e) {
        a.p = 0;
        a.next = (hp - s) / d;
    } else {
        a.p = n - 1;
        a.next = (last - s) / d;
    }
    float inc = h / d;
    // next - inc == next + (-inc) exactly, so a signed increment keeps the
    // reference's two branches in one add.
    a.inc = (s < e) ? inc : -inc;
    a.step = (s < e) ? 1 : -1;
}

__device__ __forceinline__ bool keep_walking(float lmax, float at) {
    return (lmax - at) / lmax > 1e-6f;
}

// Smallest float thr with !keep_walking(lmax, thr); for lmax > 0 the test is
// monotone in `at`, so keep_walking(lmax, at) <=> at < thr.
__device__ __forceinline__ float stop_threshold(float lmax) {
    float thr = lmax - lmax * 1e-6f;
    if (keep_walking(lmax, thr)) {
        do {
            thr = __uint_as_float(__float_as_uint(thr) + 1u);
        } while (keep_walking(lmax, thr));
    } else {
        float prev = __uint_as_float(__float_as_uint(thr) - 1u);
        while (!keep_walking(lmax, prev)) {
            thr = prev;
            prev = __uint_as_float(__float_as_uint(thr) - 1u);
        }
    }
    return thr;
}

// One Siddon ray from the source S to the bin centre E. Returns the weighted
// sum (ONES: the plain row sum); hit tells whether the ray meets the volume.
// COUNT: also counts in-volume segments of positive length into *nseg.
// NORM: also accumulates the plain row sum into *norm (sum_norm of the SART
// walker, Tomo_OSTR.cu:534-539, 601-605: the same segments, added in order,
// so it equals the ONES walk bit for bit).
template <bool ONES, bool COUNT, typename TI, bool NORM = false>
__device__ __forceinline__ float walk(const DevGeom& g, const VolReader<TI>& img, float sx,
                                      float sy, float sz, float ex, float ey, float ez,
                                      bool& hit, float& len_out, int& nseg,
                                      float* norm = nullptr) {
    hit = false;
    if (NORM) *norm = 0.0f;
    float ddx = ex - sx, ddy = ey - sy, ddz = ez - sz;
    float len = sqrtf(fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx)));
    len_out = len;
    Axis X, Y, Z;
    if (!axis_limits(X, sx, ex, g.x_p0, g.ext_x, g.top_x)) return 0.0f;
    if (!axis_limits(Y, sy, ey, g.y_p0, g.ext_y, g.top_y)) return 0.0f;
    if (!axis_limits(Z, sz, ez, g.z_p0, g.ext_z, g.top_z)) return 0.0f;
    float lmax = X.hi;
    if (lmax > Y.hi) lmax = Y.hi;
    if (lmax > Z.hi) lmax = Z.hi;
    float lmin = X.lo;
    if (lmin < Y.lo) lmin = Y.lo;
    if (lmin < Z.lo) lmin = Z.lo;
    if (!(lmin < lmax)) return 0.0f;
    axis_entry(X, lmin, sx, ex, g.x_p0, g.dx, g.nx, g.hp_x, g.last_x);
    axis_entry(Y, lmin, sy, ey, g.y_p0, g.dy, g.ny, g.hp_y, g.last_y);
    axis_entry(Z, lmin, sz, ez, g.z_p0, g.dz, g.nz, g.hp_z, g.last_z);

    const unsigned nx = g.nx, ny = g.ny, nz = g.nz;
    // Each axis index is kept as its byte offset in the volume, px = X.p E,
    // py = Y.p nx E, pz = Z.p nx ny E (E = element size), so one select per
    // crossing updates both the in-volume test and the voxel's byte offset
    // vid = px + py + pz. An index of -1 or -2 wraps to 2^32 - k stride, and
    // the volume size is limited (at handle creation) to
    // (nvox + 2 nx ny + 2 nx + 2) E <= 2^32, so it still compares as outside.
    constexpr uint32_t E = sizeof(TI);
    const uint32_t sxs = (uint32_t)X.step * E, sys = (uint32_t)Y.step * (nx * E),
                   szs = (uint32_t)Z.step * (nx * ny * E);
    const uint32_t bx = nx * E, by = ny * (nx * E), bz = nz * (ny * nx * E);
    uint32_t px = (uint32_t)X.p * E, py = (uint32_t)Y.p * (nx * E),
             pz = (uint32_t)Z.p * (nx * ny * E);
    float sum = 0.0f;
    float cur = lmin;
    // min of the three next crossings; the reference's cascade of `>` tests
    // picks the same value (no NaNs reach here)
    float at = fminf(fminf(X.next, Y.next), Z.next);
    auto inside = [&]() { return px < bx && py < by && pz < bz; };
    auto advance = [&]() {
        if (at == X.next) {
            px += sxs;
            X.next = X.next + X.inc;
        }
        if (at == Y.next) {
            py += sys;
            Y.next = Y.next + Y.inc;
        }
        if (at == Z.next) {
            pz += szs;
            Z.next = Z.next + Z.inc;
        }
        cur = at;
        at = fminf(fminf(X.next, Y.next), Z.next);
    };

    if (lmax > 0.0f) {
        // Steps are taken in batches of WALK_BATCH: the geometry of a batch is
        // advanced first, then its gathers are issued together and the sums
        // accumulated in step order, so a batch pays one memory latency.
        const float thr = stop_threshold(lmax);
        bool active = at < thr;
        // One batch of steps: its geometry into seg / off / ok. Branch-free: a
        // lane that stops inside the batch keeps its state through the
        // remaining steps (every update is a select on act), so the per-step
        // flags stay lane masks and no step needs a branch.
        auto batch = [&](float (&seg)[WALK_BATCH], uint32_t (&off)[WALK_BATCH],
                         bool (&ok)[WALK_BATCH]) {
#pragma unroll
            for (int k = 0; k < WALK_BATCH; ++k) {
                const bool act = active;
                seg[k] = (at - cur) * len;
                // non-short-circuit & (no lazily evaluated operand, so no
                // exec-mask branch per step)
                ok[k] = act & (px < bx) & (py < by) & (pz < bz);
                // a masked step gathers from past the buffer's end, which
                // reads 0, so fma(0, seg, sum) leaves sum as it is (seg is
                // finite and sum never -0): the accumulation needs no select
                off[k] = ok[k] ? px + py + pz : 0xFFFFFFF0u;
                const bool sx = act & (at == X.next);
                const bool sy = act & (at == Y.next);
                const bool sz = act & (at == Z.next);
                px += sx ? sxs : 0u;
                py += sy ? sys : 0u;
                pz += sz ? szs : 0u;
                X.next = sx ? X.next + X.inc : X.next;
                Y.next = sy ? Y.next + Y.inc : Y.next;
                Z.next = sz ? Z.next + Z.inc : Z.next;
                const float nat = fminf(fminf(X.next, Y.next), Z.next);
                cur = act ? at : cur;
                at = act ? nat : at;
                active = act & (at < thr);
            }
        };
        auto gather = [&](const uint32_t (&off)[WALK_BATCH], float (&v)[WALK_BATCH]) {
#pragma unroll
            for (int k = 0; k < WALK_BATCH; ++k) v[k] = ONES ? 1.0f : img.at_byte(off[k]);
        };
        auto accumulate = [&](const float (&v)[WALK_BATCH], const float (&seg)[WALK_BATCH],
                              const bool (&ok)[WALK_BATCH]) {
#pragma unroll
            for (int k = 0; k < WALK_BATCH; ++k) {
                sum = (ONES && !ok[k]) ? sum : fmaf(v[k], seg[k], sum);
                if (NORM) *norm = ok[k] ? fmaf(1.0f, seg[k], *norm) : *norm;
                if (COUNT && ok[k] && seg[k] > 0.0f) ++nseg;
            }
        };
        // Software-pipelined over two sets of batch registers (A, B): the
        // gathers of the next batch are issued before the values of the
        // current one are summed, so their latency overlaps the next batch's
        // geometry (the scheduling barriers keep the compiler from hoisting a
        // sum, and with it the wait for its gathers, above that geometry).
        // The wave runs while any of its lanes walks, two batches per turn; a
        // lane that has stopped takes masked steps (state held, gathers past
        // the buffer's end, which add nothing), so its batches and the order
        // of its sums are those of a plain batch loop.
        if (active) {
            float segA[WALK_BATCH], vA[WALK_BATCH], segB[WALK_BATCH], vB[WALK_BATCH];
            uint32_t offA[WALK_BATCH], offB[WALK_BATCH];
            bool okA[WALK_BATCH], okB[WALK_BATCH];
            batch(segA, offA, okA);
            gather(offA, vA);
            while (true) {
                batch(segB, offB, okB);
                gather(offB, vB);
                __builtin_amdgcn_sched_barrier(0);
                accumulate(vA, segA, okA);
                __builtin_amdgcn_sched_barrier(0);
                batch(segA, offA, okA);
                gather(offA, vA);
                __builtin_amdgcn_sched_barrier(0);
                accumulate(vB, segB, okB);
                if (!__any(active)) break;
            }
            accumulate(vA, segA, okA);
        }
    } else {
        // lmax <= 0 (volume behind the source): the literal relative test
        while (keep_walking(lmax, at)) {
            float seg = (at - cur) * len;
            if (inside()) {
                float v = ONES ? 1.0f : img.at_byte(px + py + pz);
                sum = fmaf(v, seg, sum);
                if (NORM) *norm = fmaf(1.0f, seg, *norm);
                if (COUNT && seg > 0.0f) ++nseg;
            }
            advance();
        }
    }
    float seg = (lmax - cur) * len;
    if (inside()) {
        float v = ONES ? 1.0f : img.at_byte(px + py + pz);
        sum = fmaf(v, seg, sum);
        if (NORM) *norm = fmaf(1.0f, seg, *norm);
        if (COUNT && seg > 0.0f) ++nseg;
    }
    hit = true;
    return sum;
}

// kind: 0 projection, 1 gamma ray length, 2 gamma row sum (reference layout
// [slot][nbx][nby] out); 3 projection + h epilogue (y, b, r and h in
// detector-row-major layout [slot][nby][nbx]); 4 / 5 gamma (ray length / row
// sum) times yry(y, r) with y, r read in the reference layout and the product
// written row-major (the fused precompute).
// Bin (i, j) of launch slot `slot` (view parameters v).
template <int KIND, typename TI, typename TO>
__device__ __forceinline__ void project_bin(const DevGeom& g, const ViewTab& v, int slot, int i,
                                            int j, int cos_weight, const TI* __restrict__ img,
                                            TO* __restrict__ out, const float* __restrict__ y,
                                            const float* __restrict__ b,
                                            const float* __restrict__ r) {
    const int64_t o_ref = ((int64_t)slot * g.nbx + i) * g.nby + j;
    const int64_t o_row = ((int64_t)slot * g.nby + j) * g.nbx + i;
    // the h epilogue's inputs are loaded first: their latency overlaps the
    // ray set-up and the walk (most bins miss the volume and have no walk)
    float yb = 0.0f, bb = 0.0f, rb = 0.0f;
    if (KIND == 3) {
        yb = y[o_row];
        bb = b[o_row];
        if (r) rb = r[o_row];
    }
    const float ex = bin_centre(g.x_d0, i, g.dbx);
    const float ey = bin_centre(g.y_d0, j, g.dby);
    const float ez = -g.det_r;
    bool hit = false;
    float len = 0.0f;
    int nseg = 0;
    float s = 0.0f;
    const VolReader<TI> vol(img, (uint32_t)g.nx * g.ny * g.nz);
    if (view_may_hit(v, i, j)) {  // else the ray misses the volume
        if (KIND == 0 || KIND == 3)
            s = walk<false, false, TI>(g, vol, v.x0, g.src_y, v.z0, ex, ey, ez, hit, len, nseg);
        else if (KIND == 2 || KIND == 5)
            s = walk<true, false, TI>(g, vol, v.x0, g.src_y, v.z0, ex, ey, ez, hit, len, nseg);
        else {
            walk<true, false, TI>(g, vol, v.x0, g.src_y, v.z0, ex, ey, ez, hit, len, nseg);
            s = len;
        }
    }
    float val;
    if (!hit)
        val = 0.0f;
    else if ((KIND == 0 && cos_weight) || KIND == 3)
        val = s * v.c;
    else
        val = s;
    if (KIND == 3) {
        // h epilogue (compute_h, Tomo_OSTR_gold.cpp:255-258)
        float yt = bb * expf(-val);
        float den = r ? yt + rb : yt;
        val = (yb / den - 1.0f) * yt;
    }
    if (KIND == 4 || KIND == 5) {
        // yry * gamma (compute_yry :183-205, compute_gamma_yry :208-218)
        const float yv = y[o_ref];
        float c = 0.0f;
        if (yv != 0.0f) {
            float dif = yv - (r ? r[o_ref] : 0.0f);
            if (dif <= 0.0f) dif = yv;
            c = (dif * dif) / yv;
        }
        val = c * val;
    }
    st<TO>(out, (KIND >= 3) ? o_row : o_ref, val);
}

// Tile t of a launch over n_slots views, TJ detector rows x 64 columns per
// work-group; tile order: j-tile fastest, then view slot, then i-tile (the
// groups working at one time share a view and neighbouring columns).
template <int TJ>
__device__ __forceinline__ bool bin_of_tile(const DevGeom& g, int n_slots, int t, int& slot,
                                            int& i, int& j) {
    const int tiles_j = (g.nby + TJ - 1) / TJ;
    const int tj = t % tiles_j;
    const int rest = t / tiles_j;
    slot = rest % n_slots;
    const int ti = rest / n_slots;
    i = ti * TILE_I + (threadIdx.x & 63);
    j = tj * TJ + (threadIdx.x >> 6);
    return i < g.nbx && j < g.nby;
}

template <int KIND, typename TI, typename TO>
__global__ __launch_bounds__(256) void project_kernel(DevGeom g, const ViewTab* __restrict__ views,
                                                      int a0, int n_slots, int tiles_i,
                                                      int tiles_j, int cos_weight,
                                                      const TI* __restrict__ img,
                                                      TO* __restrict__ out,
                                                      const float* __restrict__ y,
                                                      const float* __restrict__ b,
                                                      const float* __restrict__ r) {
    int slot, i, j;
    if (!bin_of_tile<TILE_J>(g, n_slots, xcd_spread_tile(blockIdx.x, gridDim.x), slot, i,
                             j))
        return;
    project_bin<KIND, TI, TO>(g, views[a0 + slot], slot, i, j, cos_weight, img, out, y, b, r);
}

// SART forward projection (ray_trace_gpu_manyangles_direct_notexturememory[_cos],
// Tomo_OSTR.cu:1175-1311, walker rayTrace3D_GPU_direct_notexturememory
// :327-611): est = (cos t) sum_j a_ij x_j and norm = sum_j a_ij (no cos), both
// from one walk and 0 for rays that miss the volume.
// DIFF = false: est and norm are written in the reference layout [slot][nbx][nby].
// DIFF = true (the fused OS-SART forward): SART_prj_diff (:2145-2183) is
// applied at once, diff = (p - est) / norm where norm != 0, else 0, with p read
// and diff written detector-row-major [slot][nby][nbx], the layout the SART
// backprojection gathers from; est and norm stay in registers.
template <bool DIFF>
__global__ __launch_bounds__(256) void sart_project_kernel(DevGeom g,
                                                           const ViewTab* __restrict__ views,
                                                           int a0, int n_slots, int cos_weight,
                                                           const float* __restrict__ img,
                                                           float* __restrict__ prj,
                                                           float* __restrict__ norm,
                                                           const float* __restrict__ p_rows) {
    int slot, i, j;
    if (!bin_of_tile<TILE_J>(g, n_slots, xcd_spread_tile(blockIdx.x, gridDim.x), slot, i,
                             j))
        return;
    const ViewTab v = views[a0 + slot];
    const float ex = bin_centre(g.x_d0, i, g.dbx);
    const float ey = bin_centre(g.y_d0, j, g.dby);
    const float ez = -g.det_r;
    const VolReader<float> vol(img, (uint32_t)g.nx * g.ny * g.nz);
    bool hit = false;
    float len, nrm = 0.0f, s = 0.0f;
    int nseg = 0;
    if (view_may_hit(v, i, j))
        s = walk<false, false, float, true>(g, vol, v.x0, g.src_y, v.z0, ex, ey, ez, hit, len,
                                            nseg, &nrm);
    const float est = hit ? (cos_weight ? s * v.c : s) : 0.0f;
    nrm = hit ? nrm : 0.0f;
    if (DIFF) {
        const int64_t o = ((int64_t)slot * g.nby + j) * g.nbx + i;
        prj[o] = nrm != 0.0f ? (p_rows[o] - est) / nrm : 0.0f;
    } else {
        const int64_t o = ((int64_t)slot * g.nbx + i) * g.nby + j;
        prj[o] = est;
        norm[o] = nrm;
    }
}

// 2x2 supersampled forward projection with the merge fused
// (ray_trace_4R_gpu_manyangles_direct_notexturememory + merge_4R_to_1R_global,
// Tomo_OSTR.cu:1453-1550, launcher fprojectCB1_4R_GPU_SART :2398-2433): one
// lane per sub-bin of the 2x subdivided detector, sub-bin centre
// x_d0 + (i + 0.5) Bsize_x / 2, plain sum (no cos); a work-group walks a tile of
// 64 x 4 sub-bins, and its first wave averages each 2x2 block from LDS in the
// order s(2i,2j) + s(2i+1,2j) + s(2i,2j+1) + s(2i+1,2j+1), then / 4 (the
// reference's merge reads an uninitialised fourth index, :1545-1547; this is
// the intended average). Both outputs are detector-row-major: prj
// [slot][nby][nbx] and, when prj4 is not null, prj4 [slot][2 nby][2 nbx]. Missed
// sub-bins are 0 (the reference leaves them unwritten).
__global__ __launch_bounds__(256) void project_4r_kernel(DevGeom g,
                                                         const ViewTab* __restrict__ views,
                                                         int a0, int n_slots,
                                                         const float* __restrict__ img,
                                                         float* __restrict__ prj,
                                                         float* __restrict__ prj4) {
    __shared__ float sub[4][64];
    const int nbx2 = 2 * g.nbx, nby2 = 2 * g.nby;
    const int tiles_j = (nby2 + 3) / 4;
    const int t = xcd_spread_tile(blockIdx.x, gridDim.x);
    const int tj = t % tiles_j;
    const int rest = t / tiles_j;
    const int slot = rest % n_slots;
    const int ti = rest / n_slots;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ii = ti * 64 + lane, jj = tj * 4 + w;
    float val = 0.0f;
    if (ii < nbx2 && jj < nby2) {
        const ViewTab v = views[a0 + slot];
        const VolReader<float> vol(img, (uint32_t)g.nx * g.ny * g.nz);
        bool hit = false;
        float len, s = 0.0f;
        int nseg = 0;
        // sub-bin (ii, jj) lies in bin (ii / 2, jj / 2); the box's 2-bin margin covers it
        if (view_may_hit(v, ii >> 1, jj >> 1))
            s = walk<false, false, float>(g, vol, v.x0, g.src_y, v.z0,
                                          bin_centre(g.x_d0, ii, 0.5f * g.dbx),
                                          bin_centre(g.y_d0, jj, 0.5f * g.dby), -g.det_r, hit,
                                          len, nseg);
        val = hit ? s : 0.0f;
        if (prj4) prj4[((int64_t)slot * nby2 + jj) * nbx2 + ii] = val;
    }
    sub[w][lane] = val;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int r = threadIdx.x >> 5, c = threadIdx.x & 31;
        const int i = ti * 32 + c, j = tj * 2 + r;
        if (i < g.nbx && j < g.nby) {
            const float m = ((sub[2 * r][2 * c] + sub[2 * r][2 * c + 1]) + sub[2 * r + 1][2 * c]) +
                            sub[2 * r + 1][2 * c + 1];
            prj[((int64_t)slot * g.nby + j) * g.nbx + i] = m / 4.0f;
        }
    }
}

// Batched projection + h (kind 3) over every case of the table in one launch;
// schedule as batch_update_kernel.
template <typename TI, typename TO, int TJ, bool LOOP>
__global__ __launch_bounds__(64 * TJ) void batch_project_kernel(
    const BatchCase* __restrict__ cases, const int* __restrict__ sched, int n_cases) {
    BatchWalk walk(sched, n_cases);
    int t, c;
    while (walk.next(t, c)) {
        const BatchCase& bc = cases[c];
        const DevGeom g = uniform_copy(bc.g);
        const int n_slots = __builtin_amdgcn_readfirstlane(bc.n_slots);
        // this XCD's contiguous share of the case's tiles, spread over the
        // detector (only the shadow's rays walk)
        const int n_tiles = ((g.nbx + TILE_I - 1) / TILE_I) * ((g.nby + TJ - 1) / TJ) * n_slots;
        int slot, i, j;
        const bool in = bin_of_tile<TJ>(g, n_slots, spread_block_perm(t, n_tiles), slot, i, j);
        const ViewTab v = uniform_copy(bc.v[slot]);  // slot is uniform over the group
        if (in)
            project_bin<3, TI, TO>(g, v, slot, i, j, 1, (const TI*)uniform_ptr(bc.x_in),
                                   (TO*)uniform_ptr(bc.h), uniform_ptr(bc.y), uniform_ptr(bc.b),
                                   uniform_ptr(bc.r));
        if (!LOOP) break;
    }
}

__global__ __launch_bounds__(256) void count_kernel(DevGeom g, const ViewTab* __restrict__ views,
                                                    int a0, int n_slots, int tiles_i,
                                                    int tiles_j,
                                                    unsigned long long* __restrict__ counter) {
    const int t = xcd_spread_tile(blockIdx.x, gridDim.x);
    const int tj = t % tiles_j;
    const int rest = t / tiles_j;
    const int slot = rest % n_slots;
    const int ti = rest / n_slots;
    const int i = ti * TILE_I + (threadIdx.x & 63);
    const int j = tj * TILE_J + (threadIdx.x >> 6);
    int nseg = 0;
    if (i < g.nbx && j < g.nby) {
        const ViewTab v = views[a0 + slot];
        bool hit;
        float len;
        const VolReader<float> none(nullptr, 0u);
        if (view_may_hit(v, i, j))
            walk<true, true, float>(g, none, v.x0, g.src_y, v.z0, bin_centre(g.x_d0, i, g.dbx),
                                    bin_centre(g.y_d0, j, g.dby), -g.det_r, hit, len, nseg);
    }
    // wave reduction, one atomic per wave
    unsigned long long c = (unsigned long long)nseg;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(counter, c);
}

template <int KIND, typename TI, typename TO>
hipError_t launch_kind(const DevGeom& g, const ViewTab* views, int a0, int a1, int cos_weight,
                       const void* img, void* out, const float* y, const float* b,
                       const float* r, hipStream_t st) {
    const int n_slots = a1 - a0;
    const int tiles_i = (g.nbx + TILE_I - 1) / TILE_I;
    const int tiles_j = (g.nby + TILE_J - 1) / TILE_J;
    const int64_t nblocks = (int64_t)tiles_i * tiles_j * n_slots;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((project_kernel<KIND, TI, TO>), dim3((unsigned)nblocks), dim3(256), 0, st,
                       g, views, a0, n_slots, tiles_i, tiles_j, cos_weight,
                       (const TI*)img, (TO*)out, y, b, r);
    return hipGetLastError();
}

template <int KIND, typename TI>
hipError_t launch_out(int out_dtype, const DevGeom& g, const ViewTab* views, int a0, int a1,
                      int cos_weight, const void* img, void* out, const float* y,
                      const float* b, const float* r, hipStream_t st) {
    switch (out_dtype) {
        case OSTR_F32:
            return launch_kind<KIND, TI, float>(g, views, a0, a1, cos_weight, img, out, y, b, r, st);
        case OSTR_BF16:
            return launch_kind<KIND, TI, bf16_t>(g, views, a0, a1, cos_weight, img, out, y, b, r, st);
        case OSTR_F16:
            return launch_kind<KIND, TI, _Float16>(g, views, a0, a1, cos_weight, img, out, y, b, r, st);
    }
    return hipErrorInvalidValue;
}

template <int KIND>
hipError_t launch_in(int img_dtype, int out_dtype, const DevGeom& g, const ViewTab* views,
                     int a0, int a1, int cos_weight, const void* img, void* out,
                     const float* y, const float* b, const float* r, hipStream_t st) {
    switch (img_dtype) {
        case OSTR_F32:
            return launch_out<KIND, float>(out_dtype, g, views, a0, a1, cos_weight, img, out, y, b, r, st);
        case OSTR_BF16:
            return launch_out<KIND, bf16_t>(out_dtype, g, views, a0, a1, cos_weight, img, out, y, b, r, st);
        case OSTR_F16:
            return launch_out<KIND, _Float16>(out_dtype, g, views, a0, a1, cos_weight, img, out, y, b, r, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_project(const DevGeom& g, const ViewTab* views, int a0, int a1, int kind,
                          int cos_weight, const void* img, int img_dtype, void* out,
                          int out_dtype, const float* y, const float* b, const float* r,
                          hipStream_t st) {
    switch (kind) {
        case 0:
            return launch_in<0>(img_dtype, out_dtype, g, views, a0, a1, cos_weight, img, out, y, b, r, st);
        case 1:
            return launch_kind<1, float, float>(g, views, a0, a1, 0, nullptr, out, y, b, r, st);
        case 2:
            return launch_kind<2, float, float>(g, views, a0, a1, 0, nullptr, out, y, b, r, st);
        case 3:
            return launch_in<3>(img_dtype, out_dtype, g, views, a0, a1, 1, img, out, y, b, r, st);
        case 4:
            return launch_kind<4, float, float>(g, views, a0, a1, 0, nullptr, out, y, b, r, st);
        case 5:
            return launch_kind<5, float, float>(g, views, a0, a1, 0, nullptr, out, y, b, r, st);
    }
    return hipErrorInvalidValue;
}

namespace {
template <typename TO, int TJ>
hipError_t batch_fwd_t(const BatchCase* cases, const int* sched, int n_cases, int total, int grid,
                       int single_pass, hipStream_t st) {
    if (single_pass)
        hipLaunchKernelGGL((batch_project_kernel<float, TO, TJ, false>), dim3(grid),
                           dim3(64 * TJ), 0, st, cases, sched, n_cases);
    else
        hipLaunchKernelGGL((batch_project_kernel<float, TO, TJ, true>), dim3(grid),
                           dim3(64 * TJ), 0, st, cases, sched, n_cases);
    return hipGetLastError();
}

template <typename TO>
hipError_t batch_fwd_o(int rows, const BatchCase* cases, const int* sched, int n_cases, int total, int grid,
                       int single, hipStream_t st) {
    switch (rows) {
        case 1: return batch_fwd_t<TO, 1>(cases, sched, n_cases, total, grid, single, st);
        case 2: return batch_fwd_t<TO, 2>(cases, sched, n_cases, total, grid, single, st);
        case 4: return batch_fwd_t<TO, 4>(cases, sched, n_cases, total, grid, single, st);
        case 8: return batch_fwd_t<TO, 8>(cases, sched, n_cases, total, grid, single, st);
    }
    return hipErrorInvalidValue;
}
}  // namespace

int64_t batch_fwd_tiles(const DevGeom& g, int n_slots, int rows) {
    return (int64_t)((g.nbx + TILE_I - 1) / TILE_I) * ((g.nby + rows - 1) / rows) * n_slots;
}

hipError_t launch_batch_project(const BatchCase* cases, const int* sched, int n_cases, int total, int grid,
                                int rows, int prj_dtype, int single_pass, hipStream_t st) {
    if (total == 0) return hipSuccess;
    switch (prj_dtype) {
        case OSTR_F32: return batch_fwd_o<float>(rows, cases, sched, n_cases, total, grid, single_pass, st);
        case OSTR_BF16: return batch_fwd_o<bf16_t>(rows, cases, sched, n_cases, total, grid, single_pass, st);
        case OSTR_F16: return batch_fwd_o<_Float16>(rows, cases, sched, n_cases, total, grid, single_pass, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sart_project(const DevGeom& g, const ViewTab* views, int a0, int a1,
                               int cos_weight, const float* img, float* prj, float* norm,
                               hipStream_t st) {
    const int n_slots = a1 - a0;
    const int64_t nblocks = (int64_t)((g.nbx + TILE_I - 1) / TILE_I) *
                            ((g.nby + TILE_J - 1) / TILE_J) * n_slots;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(sart_project_kernel<false>, dim3((unsigned)nblocks), dim3(256), 0, st, g,
                       views, a0, n_slots, cos_weight, img, prj, norm, nullptr);
    return hipGetLastError();
}

hipError_t launch_sart_project_diff(const DevGeom& g, const ViewTab* views, int a0, int a1,
                                    int cos_weight, const float* img, const float* p_rows,
                                    float* diff_rows, hipStream_t st) {
    const int n_slots = a1 - a0;
    const int64_t nblocks = (int64_t)((g.nbx + TILE_I - 1) / TILE_I) *
                            ((g.nby + TILE_J - 1) / TILE_J) * n_slots;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(sart_project_kernel<true>, dim3((unsigned)nblocks), dim3(256), 0, st, g,
                       views, a0, n_slots, cos_weight, img, diff_rows, nullptr, p_rows);
    return hipGetLastError();
}

hipError_t launch_project_4r(const DevGeom& g, const ViewTab* views, int a0, int a1,
                             const float* img, float* prj, float* prj4, hipStream_t st) {
    const int n_slots = a1 - a0;
    const int64_t nblocks = (int64_t)((2 * g.nbx + 63) / 64) * ((2 * g.nby + 3) / 4) * n_slots;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(project_4r_kernel, dim3((unsigned)nblocks), dim3(256), 0, st, g, views,
                       a0, n_slots, img, prj, prj4);
    return hipGetLastError();
}

hipError_t launch_count_nnz(const DevGeom& g, const ViewTab* views, int a0, int a1,
                            unsigned long long* counter, hipStream_t st) {
    const int n_slots = a1 - a0;
    const int tiles_i = (g.nbx + TILE_I - 1) / TILE_I;
    const int tiles_j = (g.nby + TILE_J - 1) / TILE_J;
    const int64_t nblocks = (int64_t)tiles_i * tiles_j * n_slots;
    if (nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL(count_kernel, dim3((unsigned)nblocks), dim3(256), 0, st, g, views, a0,
                       n_slots, tiles_i, tiles_j, counter);
    return hipGetLastError();
}

namespace {

// FP64 data term of the penalized negative Poisson log-likelihood, one term
// per ray of every view: (yhat + r) - y log(yhat + r), yhat = b exp(-l) with
// l = cos(t) * (Siddon sum, FP32). y, b, r in the reference layout
// [view][nbx][nby] (natural view order). One partial sum per work-group.
template <typename TI>
__global__ __launch_bounds__(256) void objective_data_kernel(
    DevGeom g, const ViewTab* __restrict__ views, int n_slots, int tiles_j,
    const TI* __restrict__ img, const float* __restrict__ y, const float* __restrict__ b,
    const float* __restrict__ r, double* __restrict__ partials) {
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int tj = t % tiles_j;
    const int rest = t / tiles_j;
    const int slot = rest % n_slots;
    const int ti = rest / n_slots;
    const int i = ti * TILE_I + (threadIdx.x & 63);
    const int j = tj * TILE_J + (threadIdx.x >> 6);
    double term = 0.0;
    if (i < g.nbx && j < g.nby) {
        const ViewTab v = views[slot];
        const VolReader<TI> vol(img, (uint32_t)g.nx * g.ny * g.nz);
        bool hit = false;
        float len, s = 0.0f;
        int nseg = 0;
        if (view_may_hit(v, i, j))
            s = walk<false, false, TI>(g, vol, v.x0, g.src_y, v.z0, bin_centre(g.x_d0, i, g.dbx),
                                       bin_centre(g.y_d0, j, g.dby), -g.det_r, hit, len, nseg);
        const float l = hit ? s * v.c : 0.0f;
        const int64_t o = ((int64_t)slot * g.nbx + i) * g.nby + j;
        const double yr = (double)b[o] * exp(-(double)l) + (r ? (double)r[o] : 0.0);
        const double yv = (double)y[o];
        term = yv == 0.0 ? yr : yr - yv * log(yr);
    }
    const double sum = block_sum_256(term);
    if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

}  // namespace

int64_t objective_data_partials(const DevGeom& g, int n_views) {
    return (int64_t)((g.nbx + TILE_I - 1) / TILE_I) * ((g.nby + TILE_J - 1) / TILE_J) * n_views;
}

hipError_t launch_objective_data(const DevGeom& g, const ViewTab* views, int n_views,
                                 const void* img, int img_dtype, const float* y,
                                 const float* b, const float* r, double* partials,
                                 int64_t* n_partials, hipStream_t st) {
    const int tiles_j = (g.nby + TILE_J - 1) / TILE_J;
    const int64_t nblocks = objective_data_partials(g, n_views);
    *n_partials = nblocks;
    if (nblocks == 0) return hipSuccess;
    if (nblocks > 0x7fffffff) return hipErrorInvalidValue;
    switch (img_dtype) {
        case OSTR_F32:
            hipLaunchKernelGGL((objective_data_kernel<float>), dim3((unsigned)nblocks), dim3(256),
                               0, st, g, views, n_views, tiles_j, (const float*)img, y, b, r,
                               partials);
            break;
        case OSTR_BF16:
            hipLaunchKernelGGL((objective_data_kernel<bf16_t>), dim3((unsigned)nblocks),
                               dim3(256), 0, st, g, views, n_views, tiles_j,
                               (const bf16_t*)img, y, b, r, partials);
            break;
        case OSTR_F16:
            hipLaunchKernelGGL((objective_data_kernel<_Float16>), dim3((unsigned)nblocks),
                               dim3(256), 0, st, g, views, n_views, tiles_j,
                               (const _Float16*)img, y, b, r, partials);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ostr
