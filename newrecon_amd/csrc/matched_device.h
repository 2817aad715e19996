// matched_device.h - per-voxel gather of the exact transpose of the Siddon
// forward projector (device code shared by the standalone transpose and the
// fused matched OSTR update).
//
// For voxel j: total = sum over the launch's views of cos(t) sum_i a_ij p_i,
// with a_ij the length of ray i (source -> bin centre, the forward
// projector's ray) inside the voxel box: the entry of the forward walk
// (Tomo_OSTR.cu:612-890) in exact arithmetic. The candidate rays are the bins
// whose centres fall in the bounding box of the voxel's 8 corners projected
// from the source onto the detector plane; each candidate is clipped against
// the voxel box (slab method: the z-slab clip once per slot, the y slab once
// per detector row, the x slab per bin; min / max are order-free, so the
// clip is that of the three slabs taken in any order; the slab reciprocals
// and the ray length use the hardware reciprocal and square root, <= 1 ulp
// each: the gather is bound by its clip arithmetic, and an exact division per
// slab cost a quarter of its time). The sinogram is detector-row-major
// [slot][nby][nbx], so neighbouring lanes (voxels along x) read neighbouring
// bins.
//
// Precision: each ray is parametrised from its detector end, p(t) = E + t (S -
// E). The volume lies next to the detector, so the clip parameters are small
// (t ~ 0.03 instead of ~0.97 from the source) and their FP32 resolution, which
// sets the error of a segment length, is ~16x finer.
#pragma once

#include "ostr_internal.h"

namespace ostr {

struct MatchedSlots {
    float x0[MAX_SLOTS], z0[MAX_SLOTS], c[MAX_SLOTS];
};

// origin + (i + 1/2) size rounded once to float: i + 1/2 is exact in FP32
// (|i| < 2^23), so the fma is the correctly rounded value. The forward walk's
// bin_centre forms the same sum in FP64 and rounds it to float; the product
// is exact in FP64 (at most 48 significant bits) and so is the sum while
// |origin| < 2^28 size (all its bits within 53 places), so the two agree, at
// a quarter of the cost (no FP64 conversions or arithmetic).
__device__ __forceinline__ float matched_bin_centre(float origin, int i, float size) {
    return fmaf((float)i + 0.5f, size, origin);
}

// Parametric overlap of the line s + t (e - s) with [lo, hi] along one axis.
__device__ __forceinline__ void matched_slab(float s, float e, float lo, float hi, float& t0,
                                             float& t1) {
    const float d = e - s;
    if (d != 0.0f) {
        const float inv = __builtin_amdgcn_rcpf(d);
        float a = (lo - s) * inv, b = (hi - s) * inv;
        t0 = fmaxf(t0, fminf(a, b));
        t1 = fminf(t1, fmaxf(a, b));
    } else if (!(s > lo && s < hi)) {
        t1 = -1.0f;  // parallel and outside: empty
        t0 = 1.0f;
    }
}

// sum over slots a of (cos_weight ? cos(t_a) : 1) sum_i a_ij P[a][i] for
// voxel (ix, iy, iz), slot sums accumulated with fma in slot order.
template <typename TP>
__device__ __forceinline__ float matched_voxel(const DevGeom& g, const MatchedSlots& S,
                                               int n_slots, int cos_weight,
                                               const VolReader<TP>& P, int ix, int iy, int iz) {
    const float bx0 = fmaf((float)ix, g.dx, g.x_p0), bx1 = fmaf((float)(ix + 1), g.dx, g.x_p0);
    const float by0 = fmaf((float)iy, g.dy, g.y_p0), by1 = fmaf((float)(iy + 1), g.dy, g.y_p0);
    const float bz0 = fmaf((float)iz, g.dz, g.z_p0), bz1 = fmaf((float)(iz + 1), g.dz, g.z_p0);
    const float sy = g.src_y, ez = -g.det_r;
    const uint32_t nbin = (uint32_t)g.nbx * g.nby;
    const float inv_dbx = 1.0f / g.dbx, inv_dby = 1.0f / g.dby;
    float total = 0.0f;
    for (int a = 0; a < n_slots; ++a) {
        const float sx = S.x0[a], sz = S.z0[a];
        // shadow of the voxel on the detector plane (corners at the two z faces);
        // it only bounds the candidate range, so hardware reciprocals suffice
        const float ta = (ez - sz) * __builtin_amdgcn_rcpf(bz0 - sz);
        const float tb = (ez - sz) * __builtin_amdgcn_rcpf(bz1 - sz);
        const float xa0 = fmaf(bx0 - sx, ta, sx), xa1 = fmaf(bx1 - sx, ta, sx);
        const float xb0 = fmaf(bx0 - sx, tb, sx), xb1 = fmaf(bx1 - sx, tb, sx);
        const float ya0 = fmaf(by0 - sy, ta, sy), ya1 = fmaf(by1 - sy, ta, sy);
        const float yb0 = fmaf(by0 - sy, tb, sy), yb1 = fmaf(by1 - sy, tb, sy);
        const float xmin = fminf(fminf(xa0, xa1), fminf(xb0, xb1));
        const float xmax = fmaxf(fmaxf(xa0, xa1), fmaxf(xb0, xb1));
        const float ymin = fminf(fminf(ya0, ya1), fminf(yb0, yb1));
        const float ymax = fmaxf(fmaxf(ya0, ya1), fmaxf(yb0, yb1));
        // bins whose centres lie in [min, max]; floor/ceil widen the range by
        // up to one bin, so rounding of the shadow, or of the bin coordinate
        // q = (min - d0) / size (taken here as a product with 1 / size), cannot
        // drop a candidate: an error below one bin moves floor(q - 0.5) to at
        // most the first bin whose centre is >= min, and likewise for ceil
        // (extra candidates clip to length 0)
        int i0 = (int)floorf((xmin - g.x_d0) * inv_dbx - 0.5f);
        int i1 = (int)ceilf((xmax - g.x_d0) * inv_dbx - 0.5f);
        int j0 = (int)floorf((ymin - g.y_d0) * inv_dby - 0.5f);
        int j1 = (int)ceilf((ymax - g.y_d0) * inv_dby - 0.5f);
        i0 = max(i0, 0);
        j0 = max(j0, 0);
        i1 = min(i1, g.nbx - 1);
        j1 = min(j1, g.nby - 1);
        // the z slab is the same for every candidate ray of the slot
        float tz0 = -3.0e38f, tz1 = 3.0e38f;
        matched_slab(ez, sz, bz0, bz1, tz0, tz1);
        const float ddz = ez - sz;
        float acc = 0.0f;
        for (int j = j0; j <= j1; ++j) {
            const float ey = matched_bin_centre(g.y_d0, j, g.dby);
            float ty0 = tz0, ty1 = tz1;
            matched_slab(ey, sy, by0, by1, ty0, ty1);
            if (!(ty0 < ty1)) continue;
            const float ddy = ey - sy;
            const float dyz = fmaf(ddz, ddz, ddy * ddy);
            // the column as an FP32 counter (i + 1/2 exact), so the bin centre
            // is matched_bin_centre's fma without a conversion; the x slab of
            // matched_slab inline, its d = sx - ex also giving the ray
            // length (d^2 = (ex - sx)^2 exactly)
            const uint32_t row = (uint32_t)a * nbin + (uint32_t)j * g.nbx;
            float fi = (float)i0 + 0.5f;
            for (int i = i0; i <= i1; ++i, fi += 1.0f) {
                const float ex = fmaf(fi, g.dbx, g.x_d0);
                float t0 = ty0, t1 = ty1;
                const float d = sx - ex;
                if (d != 0.0f) {
                    const float inv = __builtin_amdgcn_rcpf(d);
                    const float c0 = (bx0 - ex) * inv, c1 = (bx1 - ex) * inv;
                    t0 = fmaxf(t0, fminf(c0, c1));
                    t1 = fminf(t1, fmaxf(c0, c1));
                } else if (!(ex > bx0 && ex < bx1)) {
                    t1 = -1.0f;  // parallel and outside: empty
                    t0 = 1.0f;
                }
                if (t0 < t1) {
                    const float len = __builtin_amdgcn_sqrtf(fmaf(d, d, dyz));
                    acc = fmaf((t1 - t0) * len, P(row + (uint32_t)i), acc);
                }
            }
        }
        total = fmaf(cos_weight ? S.c[a] : 1.0f, acc, total);
    }
    return total;
}

}  // namespace ostr
