// ostr_oracle.cpp - CPU restatement of the reference OSTR DBT path.
//
// TEST INFRASTRUCTURE ONLY. This library is the parity checker for
// libostr_hip.so: only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg load it. The product path never calls into it.
//
// One template, two builds: T = float mirrors the reference's FP32 arithmetic
// (the parity oracle); T = double is the FP64 reference used for error budgets
// and the adjoint identity. Where the reference expression is `a*b + c` the
// restatement writes fma(a, b, c): nvcc contracts those expressions by default,
// and the HIP kernels write the same fma, so the two sides evaluate the same
// expression graph. Everything else is compiled with -ffp-contract=off.
//
// Sources followed (paths relative to the reference repository):
//   Siddon walker .................. Tomo_OSTR.cu:612-890
//   forward kernel (view loop, bin centre, cos weight, write-if-hit)
//                                    Tomo_OSTR.cu:1313-1376
//   gamma (normprj): ray length ..... Tomo_OSTR.cu:892-1173 (returns Length, :1172)
//                    row sum ........ ostr.ipynb raw lines 595-815
//   SBP backprojector ............... Tomo_OSTR.cu:1553-1734
//   priors Q2 / Lange / hyperbola ... Tomo_OSTR.cu:118-184 / 187-255 / 258-325
//   yry, gamma*yry, regroup, h, update
//                                    Tomo_OSTR_gold.cpp:183-295
//   OSTR driver ..................... Tomo_OSTR_gold.cpp:301-918 (Lange), 920-1469 (Q2)
//   OS-SART: forward + row sum ...... Tomo_OSTR.cu:2278-2333 (kernels :1313-1376 family)
//            SART_prj_diff .......... Tomo_OSTR.cu:2145-2183
//            backprojection update .. Tomo_OSTR.cu:1737-2143 (plain and _R)
//            Huber prior ............ Tomo_OSTR.cu:258-325
//   2x2 supersampled forward + merge  Tomo_OSTR.cu:1453-1550, 2398-2433
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ostr_hip.h"

namespace {

template <typename T>
struct Geo {
    int nx, ny, nz, nbx, nby;
    T dx, dy, dz, x_p0, y_p0, z_p0, dbx, dby, x_d0, y_d0, det_r, src_r, src_y;
    explicit Geo(const ostr_geom& g)
        : nx(g.nx), ny(g.ny), nz(g.nz), nbx(g.nbx), nby(g.nby), dx(g.dx), dy(g.dy),
          dz(g.dz), x_p0(g.x_p0), y_p0(g.y_p0), z_p0(g.z_p0), dbx(g.dbx), dby(g.dby),
          x_d0(g.x_d0), y_d0(g.y_d0), det_r(g.detector_r), src_r(g.source_r),
          src_y(g.source_y) {}
    int64_t nvox() const { return (int64_t)nx * ny * nz; }
    int64_t nbin() const { return (int64_t)nbx * nby; }
};

inline float fmaT(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double fmaT(double a, double b, double c) { return std::fma(a, b, c); }

// Source position of one view slot. sin/cos are evaluated in double from the
// float32 angle and rounded to T (the reference calls sinf/cosf in the kernel,
// Tomo_OSTR.cu:1340-1345).
template <typename T>
struct ViewPos {
    T x0, z0, c;
};

template <typename T>
ViewPos<T> view_pos(const Geo<T>& g, float theta) {
    ViewPos<T> v;
    T s = (T)std::sin((double)theta);
    v.c = (T)std::cos((double)theta);
    v.x0 = g.src_r * s;
    v.z0 = g.src_r * v.c;
    return v;
}

// Detector bin centre / voxel centre: the reference evaluates (i+0.5)*size in
// double before storing to float (Tomo_OSTR.cu:1351-1352, 1625-1627).
template <typename T>
inline T cell_centre(T origin, int i, T size) {
    return (T)((double)origin + ((double)i + 0.5) * (double)size);
}

// One axis of the Siddon walk.
template <typename T>
struct Ax {
    T lo, hi;   // parametric entry/exit of the slab [p0, p0 + n h]
    T next;     // parameter of the next plane crossing
    T inc;      // h / d
    int p;      // current voxel index along the axis
    bool up;    // s < e
};

// Parametric slab limits along one axis; returns false if the ray is parallel
// to the axis and outside the slab (Tomo_OSTR.cu:641-715).
template <typename T>
inline bool axis_limits(Ax<T>& a, T s, T e, T p0, T h, int n) {
    T d = e - s;
    if (d != (T)0) {
        a.lo = (p0 - s) / d;
        a.hi = a.lo + ((T)n * h) / d;
        if (a.lo > a.hi) std::swap(a.lo, a.hi);
        return true;
    }
    if (s >= fmaT((T)n, h, p0) || s <= p0) return false;
    a.lo = (T)-1000;
    a.hi = (T)1000;
    return true;
}

// Entry voxel index and first crossing along one axis (Tomo_OSTR.cu:726-801).
template <typename T>
inline void axis_entry(Ax<T>& a, T lmin, T s, T e, T p0, T h, int n) {
    T d = e - s;
    a.up = s < e;
    if (a.lo != lmin) {
        a.p = (int)std::floor((fmaT(lmin, d, s) - p0) / h);
        if (s < e)
            a.next = (fmaT((T)(a.p + 1), h, p0) - s) / d;
        else if (s == e)
            a.next = (T)1000;
        else
            a.next = (fmaT((T)a.p, h, p0) - s) / d;
    } else if (s < e) {
        a.p = 0;
        a.next = ((h + p0) - s) / d;
    } else {
        a.p = n - 1;
        a.next = (fmaT((T)(n - 1), h, p0) - s) / d;
    }
    a.inc = h / d;
}

template <typename T>
inline void axis_step(Ax<T>& a, T at) {
    if (at == a.next) {
        if (a.up) {
            a.p += 1;
            a.next = a.next + a.inc;
        } else {
            a.p -= 1;
            a.next = a.next - a.inc;
        }
    }
}

// Siddon ray walk from (sx,sy,sz) to (ex,ey,ez). Returns sum_j img_j a_ij
// (img == nullptr: sum_j a_ij), sets *hit, *length, and adds the number of
// in-volume segments with a_ij > 0 to *nseg when nseg != nullptr.
template <typename T, typename I>
T walk(const Geo<T>& g, const I* img, T sx, T sy, T sz, T ex, T ey, T ez, bool* hit,
       T* length, int64_t* nseg) {
    *hit = false;
    T ddx = ex - sx, ddy = ey - sy, ddz = ez - sz;
    T len = std::sqrt(fmaT(ddz, ddz, fmaT(ddy, ddy, ddx * ddx)));
    *length = len;
    Ax<T> X, Y, Z;
    if (!axis_limits(X, sx, ex, g.x_p0, g.dx, g.nx)) return 0;
    if (!axis_limits(Y, sy, ey, g.y_p0, g.dy, g.ny)) return 0;
    if (!axis_limits(Z, sz, ez, g.z_p0, g.dz, g.nz)) return 0;
    T lmax = X.hi;
    if (lmax > Y.hi) lmax = Y.hi;
    if (lmax > Z.hi) lmax = Z.hi;
    T lmin = X.lo;
    if (lmin < Y.lo) lmin = Y.lo;
    if (lmin < Z.lo) lmin = Z.lo;
    if (lmin >= lmax) return 0;
    axis_entry(X, lmin, sx, ex, g.x_p0, g.dx, g.nx);
    axis_entry(Y, lmin, sy, ey, g.y_p0, g.dy, g.ny);
    axis_entry(Z, lmin, sz, ez, g.z_p0, g.dz, g.nz);

    auto inside = [&]() {
        return X.p >= 0 && X.p < g.nx && Y.p >= 0 && Y.p < g.ny && Z.p >= 0 && Z.p < g.nz;
    };
    auto value = [&]() -> T {
        if (!img) return (T)1;
        return (T)img[((int64_t)Z.p * g.ny + Y.p) * g.nx + X.p];
    };
    T sum = 0;
    T cur = lmin;
    T at = X.next;
    if (at > Y.next) at = Y.next;
    if (at > Z.next) at = Z.next;
    // relative stop test of Tomo_OSTR.cu:815 (1e-6 compared in double there)
    while ((double)((lmax - at) / lmax) > 1e-6) {
        T seg = (at - cur) * len;
        if (inside()) {
            sum = fmaT(value(), seg, sum);
            if (nseg && seg > 0) ++*nseg;
        }
        axis_step(X, at);
        axis_step(Y, at);
        axis_step(Z, at);
        cur = at;
        at = X.next;
        if (at > Y.next) at = Y.next;
        if (at > Z.next) at = Z.next;
    }
    T seg = (lmax - cur) * len;
    if (inside()) {
        sum = fmaT(value(), seg, sum);
        if (nseg && seg > 0) ++*nseg;
    }
    *hit = true;
    return sum;
}

inline int nthreads_or_default(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// ---------------------------------------------------------------- forward --
// kind 0: cos-weighted / plain projection of img; 1: gamma ray length;
// 2: gamma row sum.
template <typename T>
void project(const Geo<T>& g, const float* angles, const int* index, const T* img, T* prj,
             int a0, int a1, int kind, int cos_weight, int nthr) {
    const int nb = g.nbx;
    int rows = (a1 - a0) * nb;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads_or_default(nthr))
    for (int r = 0; r < rows; ++r) {
        int a = a0 + r / nb;
        int i = r % nb;
        ViewPos<T> v = view_pos(g, angles[index[a]]);
        T ex = cell_centre(g.x_d0, i, g.dbx);
        T ez = -g.det_r;
        T* out = prj + ((int64_t)(a - a0) * g.nbx + i) * g.nby;
        for (int j = 0; j < g.nby; ++j) {
            T ey = cell_centre(g.y_d0, j, g.dby);
            bool hit;
            T len;
            T s;
            if (kind == 0)
                s = walk<T, T>(g, img, v.x0, g.src_y, v.z0, ex, ey, ez, &hit, &len, nullptr);
            else if (kind == 2)
                s = walk<T, T>(g, nullptr, v.x0, g.src_y, v.z0, ex, ey, ez, &hit, &len,
                               nullptr);
            else {
                walk<T, T>(g, nullptr, v.x0, g.src_y, v.z0, ex, ey, ez, &hit, &len, nullptr);
                s = len;
            }
            if (!hit)
                out[j] = 0;
            else if (kind == 0 && cos_weight)
                out[j] = s * v.c;
            else
                out[j] = s;
        }
    }
}

template <typename T>
int64_t count_nnz(const Geo<T>& g, const float* angles, const int* index, int a0, int a1,
                  int nthr) {
    int64_t total = 0;
    int rows = (a1 - a0) * g.nbx;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total) \
    num_threads(nthreads_or_default(nthr))
    for (int r = 0; r < rows; ++r) {
        int a = a0 + r / g.nbx;
        int i = r % g.nbx;
        ViewPos<T> v = view_pos(g, angles[index[a]]);
        T ex = cell_centre(g.x_d0, i, g.dbx);
        for (int j = 0; j < g.nby; ++j) {
            T ey = cell_centre(g.y_d0, j, g.dby);
            bool hit;
            T len;
            int64_t n = 0;
            walk<T, T>(g, nullptr, v.x0, g.src_y, v.z0, ex, ey, -g.det_r, &hit, &len, &n);
            total += n;
        }
    }
    return total;
}

// Exact transpose of project(kind 0): img += (cos A)^T prj. Serial, so the sum
// order is fixed (used for the adjoint identity on small cases).
template <typename T>
void backproject_matched(const Geo<T>& g, const float* angles, const int* index, T* img,
                         const T* prj, int a0, int a1, int cos_weight) {
    for (int a = a0; a < a1; ++a) {
        ViewPos<T> v = view_pos(g, angles[index[a]]);
        for (int i = 0; i < g.nbx; ++i) {
            T ex = cell_centre(g.x_d0, i, g.dbx);
            for (int j = 0; j < g.nby; ++j) {
                T p = prj[((int64_t)(a - a0) * g.nbx + i) * g.nby + j];
                if (cos_weight) p = p * v.c;
                T ey = cell_centre(g.y_d0, j, g.dby);
                // Re-walk the ray, scattering p * a_ij.
                T sx = v.x0, sy = g.src_y, sz = v.z0, ez = -g.det_r;
                T ddx = ex - sx, ddy = ey - sy, ddz = ez - sz;
                T len = std::sqrt(fmaT(ddz, ddz, fmaT(ddy, ddy, ddx * ddx)));
                Ax<T> X, Y, Z;
                if (!axis_limits(X, sx, ex, g.x_p0, g.dx, g.nx)) continue;
                if (!axis_limits(Y, sy, ey, g.y_p0, g.dy, g.ny)) continue;
                if (!axis_limits(Z, sz, ez, g.z_p0, g.dz, g.nz)) continue;
                T lmax = std::min(X.hi, std::min(Y.hi, Z.hi));
                T lmin = X.lo;
                if (lmin < Y.lo) lmin = Y.lo;
                if (lmin < Z.lo) lmin = Z.lo;
                if (lmin >= lmax) continue;
                axis_entry(X, lmin, sx, ex, g.x_p0, g.dx, g.nx);
                axis_entry(Y, lmin, sy, ey, g.y_p0, g.dy, g.ny);
                axis_entry(Z, lmin, sz, ez, g.z_p0, g.dz, g.nz);
                auto deposit = [&](T seg) {
                    if (X.p >= 0 && X.p < g.nx && Y.p >= 0 && Y.p < g.ny && Z.p >= 0 &&
                        Z.p < g.nz) {
                        T& dst = img[((int64_t)Z.p * g.ny + Y.p) * g.nx + X.p];
                        dst = fmaT(p, seg, dst);
                    }
                };
                T cur = lmin;
                T at = std::min(X.next, std::min(Y.next, Z.next));
                while ((double)((lmax - at) / lmax) > 1e-6) {
                    deposit((at - cur) * len);
                    axis_step(X, at);
                    axis_step(Y, at);
                    axis_step(Z, at);
                    cur = at;
                    at = X.next;
                    if (at > Y.next) at = Y.next;
                    if (at > Z.next) at = Z.next;
                }
                deposit((lmax - cur) * len);
            }
        }
    }
}

// Exact geometric transpose, voxel-driven: img_j += sum_a c_a sum_i |ray_i n
// voxel_j| prj_i, the candidate rays being the bins whose centres lie in the
// (floor/ceil-widened) bounding box of the voxel's corners projected from the
// source onto the detector. The rays are the forward projector's FP32 rays
// (source and bin centre rounded to float, Tomo_OSTR.cu:1340-1358); their
// lengths inside each voxel are computed in T. It differs from
// backproject_matched only where the walk's relative stop test
// (Tomo_OSTR.cu:815) hands the last < 1e-6*max_l of a ray to the voxel before
// the exit.
template <typename T>
void backproject_clip(const Geo<T>& g, const Geo<float>& gf, const float* angles,
                      const int* index, T* img, const T* prj, int a0, int a1, int cos_weight,
                      int nthr = 1) {
    const T ez = (T)(-gf.det_r), sy = (T)gf.src_y;
    // one thread per slice; each voxel is summed by one thread in a fixed order
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthr > 0 ? nthr : 1)
    for (int iz = 0; iz < g.nz; ++iz)
        for (int iy = 0; iy < g.ny; ++iy)
            for (int ix = 0; ix < g.nx; ++ix) {
                const T bx0 = g.x_p0 + ix * g.dx, bx1 = g.x_p0 + (ix + 1) * g.dx;
                const T by0 = g.y_p0 + iy * g.dy, by1 = g.y_p0 + (iy + 1) * g.dy;
                const T bz0 = g.z_p0 + iz * g.dz, bz1 = g.z_p0 + (iz + 1) * g.dz;
                T total = 0;
                for (int a = a0; a < a1; ++a) {
                    const ViewPos<float> v = view_pos(gf, angles[index[a]]);
                    const T sx = (T)v.x0, sz = (T)v.z0;
                    const T ta = (ez - sz) / (bz0 - sz), tb = (ez - sz) / (bz1 - sz);
                    const T xs[4] = {sx + (bx0 - sx) * ta, sx + (bx1 - sx) * ta,
                                     sx + (bx0 - sx) * tb, sx + (bx1 - sx) * tb};
                    const T ys[4] = {sy + (by0 - sy) * ta, sy + (by1 - sy) * ta,
                                     sy + (by0 - sy) * tb, sy + (by1 - sy) * tb};
                    const T xmin = *std::min_element(xs, xs + 4), xmax = *std::max_element(xs, xs + 4);
                    const T ymin = *std::min_element(ys, ys + 4), ymax = *std::max_element(ys, ys + 4);
                    const int i0 = std::max(0, (int)std::floor((xmin - g.x_d0) / g.dbx - 0.5));
                    const int i1 = std::min(g.nbx - 1, (int)std::ceil((xmax - g.x_d0) / g.dbx - 0.5));
                    const int j0 = std::max(0, (int)std::floor((ymin - g.y_d0) / g.dby - 0.5));
                    const int j1 = std::min(g.nby - 1, (int)std::ceil((ymax - g.y_d0) / g.dby - 0.5));
                    T acc = 0;
                    for (int j = j0; j <= j1; ++j)
                        for (int i = i0; i <= i1; ++i) {
                            const T ex = (T)cell_centre(gf.x_d0, i, gf.dbx);
                            const T ey = (T)cell_centre(gf.y_d0, j, gf.dby);
                            const T e[3] = {ex, ey, ez}, s[3] = {sx, sy, sz};
                            const T lo[3] = {bx0, by0, bz0}, hi[3] = {bx1, by1, bz1};
                            T t0 = -1e30, t1 = 1e30;
                            for (int k = 0; k < 3; ++k) {
                                const T d = s[k] - e[k];
                                if (d != 0) {
                                    const T u = (lo[k] - e[k]) / d, w = (hi[k] - e[k]) / d;
                                    t0 = std::max(t0, std::min(u, w));
                                    t1 = std::min(t1, std::max(u, w));
                                } else if (!(e[k] > lo[k] && e[k] < hi[k])) {
                                    t1 = -1;
                                    t0 = 1;
                                }
                            }
                            if (t0 < t1) {
                                const T len = std::sqrt((ex - sx) * (ex - sx) + (ey - sy) * (ey - sy) +
                                                        (ez - sz) * (ez - sz));
                                acc += (t1 - t0) * len *
                                       prj[((int64_t)(a - a0) * g.nbx + i) * g.nby + j];
                            }
                        }
                    total += (cos_weight ? (T)v.c : (T)1) * acc;
                }
                img[((int64_t)iz * g.ny + iy) * g.nx + ix] += total;
            }
}

// One-step backprojected image of ReconDBT/BP.m:54-82 with G = cos(t) A (the
// OSTR forward model): per view, l = G 1 (row sums), q = p / l where
// l > |dz|/10 (else 0), img += (G^T q) / (G^T 1) (0 where G^T 1 = 0); finally
// img / n_views. p: line integrals [V][nbx][nby], natural view order. The
// per-view cos cancels in G^T q / G^T 1, so both use the plain transpose.
template <typename T>
void bp_init(const Geo<T>& g, const Geo<float>& gf, const float* angles, int n_views,
             const T* p, T* img) {
    const int64_t nbin = g.nbin(), nvox = g.nvox();
    std::vector<int> idx(n_views);
    for (int i = 0; i < n_views; ++i) idx[i] = i;
    std::vector<T> l(nbin), q(nbin), ones(nbin, (T)1), num(nvox), den(nvox);
    std::fill(img, img + nvox, (T)0);
    const T thr = std::fabs(g.dz) / 10;
    for (int v = 0; v < n_views; ++v) {
        project(g, angles, idx.data(), (const T*)nullptr, l.data(), v, v + 1, 2, 0, 0);
        const T c = (T)view_pos(gf, angles[v]).c;
        for (int64_t i = 0; i < nbin; ++i) {
            const T lv = c * l[i];
            q[i] = lv > thr ? p[(int64_t)v * nbin + i] / lv : (T)0;
        }
        std::fill(num.begin(), num.end(), (T)0);
        std::fill(den.begin(), den.end(), (T)0);
        backproject_clip(g, gf, angles, idx.data(), num.data(), q.data(), v, v + 1, 0);
        backproject_clip(g, gf, angles, idx.data(), den.data(), ones.data(), v, v + 1, 0);
        for (int64_t j = 0; j < nvox; ++j)
            if (den[j] != 0) img[j] += num[j] / den[j];
    }
    for (int64_t j = 0; j < nvox; ++j) img[j] /= (T)n_views;
}

// ------------------------------------------------------------- SBP ------
template <typename T>
void backproject_sbp(const Geo<T>& g, const float* angles, const int* index, T* img,
                     const T* prj, int a0, int a1, int nthr) {
    std::vector<ViewPos<T>> vp(a1 - a0);
    for (int a = a0; a < a1; ++a) vp[a - a0] = view_pos(g, angles[index[a]]);
    const int64_t nbin = g.nbin();
    int rows = g.nz * g.ny;
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthr))
    for (int r = 0; r < rows; ++r) {
        int iz = r / g.ny, iy = r % g.ny;
        T y1 = cell_centre(g.y_p0, iy, g.dy);
        T z1 = cell_centre(g.z_p0, iz, g.dz);
        for (int ix = 0; ix < g.nx; ++ix) {
            T x1 = cell_centre(g.x_p0, ix, g.dx);
            T total = 0;
            for (int a = a0; a < a1; ++a) {
                const ViewPos<T>& v = vp[a - a0];
                T x0 = v.x0, y0 = g.src_y, z0 = v.z0;
                T ex = x0 - x1, ey = y0 - y1, ez = z0 - z1;
                T dw = fmaT(z0, z0, fmaT(y0, y0, x0 * x0)) / fmaT(ez, ez, fmaT(ey, ey, ex * ex));
                if (z1 == z0) continue;
                T t = (-g.det_r - z0) / (z1 - z0);
                T x2 = fmaT(x1 - x0, t, x0);
                T y2 = fmaT(y1 - y0, t, y0);
                T imb = (x2 - g.x_d0) / g.dbx;
                T jmb = (y2 - g.y_d0) / g.dby;
                int il = (int)std::floor(imb);
                if (imb < (T)il + (T)0.5) il -= 1;
                int jl = (int)std::floor(jmb);
                if (jmb < (T)jl + (T)0.5) jl -= 1;
                T fi = imb - ((T)il + (T)0.5);
                T fj = jmb - ((T)jl + (T)0.5);
                const T* P = prj + (int64_t)(a - a0) * nbin;
                bool i0 = il >= 0 && il < g.nbx, i1 = il + 1 >= 0 && il + 1 < g.nbx;
                bool j0 = jl >= 0 && jl < g.nby, j1 = jl + 1 >= 0 && jl + 1 < g.nby;
                T d1 = 0, d2 = 0;
                if (i0 && j0) d1 = ((T)1 - fi) * P[(int64_t)il * g.nby + jl];
                if (i1 && j0) d1 = fmaT(fi, P[(int64_t)(il + 1) * g.nby + jl], d1);
                if (i0 && j1) d2 = ((T)1 - fi) * P[(int64_t)il * g.nby + jl + 1];
                if (i1 && j1) d2 = fmaT(fi, P[(int64_t)(il + 1) * g.nby + jl + 1], d2);
                T u = fmaT(fj, d2, ((T)1 - fj) * d1);
                u = (u * g.dz) * dw;
                total = total + u;
            }
            int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
            img[vi] = img[vi] + total;
        }
    }
}

// ----------------------------------------------------------- priors -----
// Neighbour weights w = 1/dist, dist = sqrt(di^2 + dj^2 + (dk rho)^2) in float
// with the integer part summed first (Tomo_OSTR.cu:229); the reciprocal is
// taken in double and rounded (the reference's 1.0/distance). Centre -> 0.
template <typename T>
void prior_weights(T w[27], T rho) {
    int n = 0;
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                T dist = std::sqrt((T)(di * di + dj * dj) + ((T)(dk * dk) * rho) * rho);
                w[n++] = (dist == (T)0) ? (T)0 : (T)(1.0 / (double)dist);
            }
}

template <typename T>
inline T prior_omega(int potential, T t, T inv_delta) {
    if (potential == OSTR_PRIOR_Q2) return (T)1;
    T q = t * inv_delta;
    if (potential == OSTR_PRIOR_LANGE) return (T)1 / ((T)1 + std::fabs(q));
    return (T)1 / std::sqrt(fmaT(q, q, (T)1));
}

template <typename T>
void prior(const Geo<T>& g, T* rdd, T* rd, const T* x, int potential, T delta, T rho,
           int nthr) {
    T w[27];
    prior_weights(w, rho);
    T inv_delta = potential == OSTR_PRIOR_Q2 ? (T)0 : (T)1 / delta;
    int rows = g.nz * g.ny;
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthr))
    for (int r = 0; r < rows; ++r) {
        int iz = r / g.ny, iy = r % g.ny;
        for (int ix = 0; ix < g.nx; ++ix) {
            int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
            T xc = x[vi];
            T acc_dd = rdd[vi], acc_d = rd[vi];
            int n = 0;
            for (int dk = -1; dk <= 1; ++dk)
                for (int dj = -1; dj <= 1; ++dj)
                    for (int di = -1; di <= 1; ++di, ++n) {
                        int kx = ix + di, ky = iy + dj, kz = iz + dk;
                        bool in = kx >= 0 && kx < g.nx && ky >= 0 && ky < g.ny && kz >= 0 &&
                                  kz < g.nz;
                        T t = in ? xc - x[((int64_t)kz * g.ny + ky) * g.nx + kx] : (T)0;
                        T e = w[n] * prior_omega(potential, t, inv_delta);
                        acc_dd = acc_dd + e;
                        acc_d = fmaT(e, t, acc_d);
                    }
            rdd[vi] = acc_dd;
            rd[vi] = acc_d;
        }
    }
}

template <typename T>
inline T prior_psi(int potential, T t, T delta) {
    if (potential == OSTR_PRIOR_Q2) return t * t / 2;
    T q = std::fabs(t / delta);
    if (potential == OSTR_PRIOR_LANGE) return delta * delta * (q - std::log1p(q));
    // delta^2 (sqrt(1 + q^2) - 1) without the cancellation at small q
    return delta * delta * (q * q / (std::sqrt(1 + q * q) + 1));
}

// R(x) = 1/2 sum_j sum_{k in N(j), in volume} w_jk psi(x_j - x_k)
template <typename T>
double penalty(const Geo<T>& g, const T* x, int potential, T delta, T rho) {
    T w[27];
    prior_weights(w, rho);
    double total = 0;
    for (int iz = 0; iz < g.nz; ++iz)
        for (int iy = 0; iy < g.ny; ++iy)
            for (int ix = 0; ix < g.nx; ++ix) {
                int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
                int n = 0;
                for (int dk = -1; dk <= 1; ++dk)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di, ++n) {
                            int kx = ix + di, ky = iy + dj, kz = iz + dk;
                            if (kx < 0 || kx >= g.nx || ky < 0 || ky >= g.ny || kz < 0 ||
                                kz >= g.nz || w[n] == 0)
                                continue;
                            T t = x[vi] - x[((int64_t)kz * g.ny + ky) * g.nx + kx];
                            total += (double)w[n] * (double)prior_psi(potential, t, delta);
                        }
            }
    return 0.5 * total;
}

// ------------------------------------------------------ elementwise -----
template <typename T>
void yry(T* out, const T* y, const T* r, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        if (y[i] == 0) {
            out[i] = 0;
            continue;
        }
        T dif = y[i] - (r ? r[i] : (T)0);
        if (dif <= 0) dif = y[i];
        out[i] = (dif * dif) / y[i];
    }
}

template <typename T>
void compute_h(T* h, const T* y, const T* b, const T* l, const T* r, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        T yt = b[i] * std::exp(-l[i]);
        T den = r ? yt + r[i] : yt;
        h[i] = (y[i] / den - (T)1) * yt;
    }
}

template <typename T>
void update_est(T* x, const T* capl, const T* rd, const T* d, const T* rdd, T beta,
                T capl_scale, int guard, int64_t n) {
    T two_beta = (T)2 * beta;
    for (int64_t i = 0; i < n; ++i) {
        T num = capl_scale * capl[i] + beta * rd[i];
        T den = d[i] + two_beta * rdd[i];
        if (guard && den == 0) continue;
        T v = x[i] - num / den;
        x[i] = v < 0 ? (T)0 : v;
    }
}

template <typename T>
void regroup(T* dst, const T* src, int n_views, int n_subsets, int64_t nbin) {
    int vs = n_views / n_subsets;
    int64_t k = 0;
    for (int m = 0; m < n_subsets; ++m)
        for (int j = 0; j < vs; ++j, k += nbin)
            std::memcpy(dst + k, src + (int64_t)(j * n_subsets + m) * nbin, nbin * sizeof(T));
}

// ------------------------------------------------------------ OS-SART -----
// SART_prj_diff_kernel (Tomo_OSTR.cu:2145-2183).
template <typename T>
void sart_prj_diff(T* diff, const T* p, const T* est, const T* norm, int64_t n) {
    for (int64_t i = 0; i < n; ++i) diff[i] = norm[i] != 0 ? (p[i] - est[i]) / norm[i] : (T)0;
}

// backprj_OSSART_gpu_manyviews / _R (Tomo_OSTR.cu:1737-2143): the SBP gather
// of diff (reference layout) and the sensitivity (the same weights on the
// valid taps' indicator), then img = max(0, img + lambda (sum/sens [+ beta
// prior / sens])) in place; sum/sens taken as 0 where sens == 0.
template <typename T>
void backproject_sart(const Geo<T>& g, const float* angles, const int* index, T* img,
                      const T* diff, const T* prior, int a0, int a1, T lambda, T beta,
                      int nthr) {
    std::vector<ViewPos<T>> vp(a1 - a0);
    for (int a = a0; a < a1; ++a) vp[a - a0] = view_pos(g, angles[index[a]]);
    const int64_t nbin = g.nbin();
    int rows = g.nz * g.ny;
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthr))
    for (int r = 0; r < rows; ++r) {
        int iz = r / g.ny, iy = r % g.ny;
        T y1 = cell_centre(g.y_p0, iy, g.dy);
        T z1 = cell_centre(g.z_p0, iz, g.dz);
        for (int ix = 0; ix < g.nx; ++ix) {
            T x1 = cell_centre(g.x_p0, ix, g.dx);
            T total = 0, sens = 0;
            for (int a = a0; a < a1; ++a) {
                const ViewPos<T>& v = vp[a - a0];
                T x0 = v.x0, y0 = g.src_y, z0 = v.z0;
                T ex = x0 - x1, ey = y0 - y1, ez = z0 - z1;
                T dw = fmaT(z0, z0, fmaT(y0, y0, x0 * x0)) / fmaT(ez, ez, fmaT(ey, ey, ex * ex));
                if (z1 == z0) continue;
                T t = (-g.det_r - z0) / (z1 - z0);
                T x2 = fmaT(x1 - x0, t, x0);
                T y2 = fmaT(y1 - y0, t, y0);
                T imb = (x2 - g.x_d0) / g.dbx;
                T jmb = (y2 - g.y_d0) / g.dby;
                int il = (int)std::floor(imb);
                if (imb < (T)il + (T)0.5) il -= 1;
                int jl = (int)std::floor(jmb);
                if (jmb < (T)jl + (T)0.5) jl -= 1;
                T fi = imb - ((T)il + (T)0.5);
                T fj = jmb - ((T)jl + (T)0.5);
                const T* P = diff + (int64_t)(a - a0) * nbin;
                bool i0 = il >= 0 && il < g.nbx, i1 = il + 1 >= 0 && il + 1 < g.nbx;
                bool j0 = jl >= 0 && jl < g.nby, j1 = jl + 1 >= 0 && jl + 1 < g.nby;
                T d1 = 0, d2 = 0, s1 = 0, s2 = 0;
                if (i0 && j0) { d1 = ((T)1 - fi) * P[(int64_t)il * g.nby + jl]; s1 = (T)1 - fi; }
                if (i1 && j0) { d1 = fmaT(fi, P[(int64_t)(il + 1) * g.nby + jl], d1); s1 = s1 + fi; }
                if (i0 && j1) { d2 = ((T)1 - fi) * P[(int64_t)il * g.nby + jl + 1]; s2 = (T)1 - fi; }
                if (i1 && j1) { d2 = fmaT(fi, P[(int64_t)(il + 1) * g.nby + jl + 1], d2); s2 = s2 + fi; }
                T u = fmaT(fj, d2, ((T)1 - fj) * d1);
                u = (u * g.dz) * dw;
                T us = fmaT(fj, s2, ((T)1 - fj) * s1);
                us = (us * g.dz) * dw;
                total = total + u;
                sens = sens + us;
            }
            int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
            T u = 0, bt = 0;
            if (sens != 0) {
                u = total / sens;
                if (prior) bt = (beta * prior[vi]) / sens;
            }
            T xn = fmaT(lambda, prior ? u + bt : u, img[vi]);
            img[vi] = xn < 0 ? (T)0 : xn;
        }
    }
}

// G_Huber_prior (Tomo_OSTR.cu:258-325), accumulating into prior. The FP32
// mirror keeps the reference's precision: distance = sqrtf of the integer
// squared offset (centre 0 -> 1), diff / delta and its square in float, the
// square root in double rounded to the float denominator, the term in double,
// the running sum rounded to float after each term.
template <typename T>
void prior_huber(const Geo<T>& g, T* prior, const T* x, T delta, int nthr) {
    int rows = g.nz * g.ny;
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthr))
    for (int r = 0; r < rows; ++r) {
        int iz = r / g.ny, iy = r % g.ny;
        for (int ix = 0; ix < g.nx; ++ix) {
            int64_t vi = ((int64_t)iz * g.ny + iy) * g.nx + ix;
            T xc = x[vi];
            T acc = prior[vi];
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        float dist = std::sqrt((float)(dx * dx + dy * dy + dz * dz));
                        if (dist == 0.0f) dist = 1.0f;
                        int nz = iz + dz, ny = iy + dy, nx = ix + dx;
                        T xn = xc;
                        if (nx >= 0 && ny >= 0 && nz >= 0 && nx < g.nx && ny < g.ny && nz < g.nz)
                            xn = x[((int64_t)nz * g.ny + ny) * g.nx + nx];
                        T diff = xc - xn;
                        T q = diff / delta;
                        T den = (T)std::sqrt(1.0 + (double)(q * q));
                        acc = (T)((double)acc + ((1.0 / (double)dist) * (double)diff) / (double)den);
                    }
            prior[vi] = acc;
        }
    }
}

// OS-SART driver (the reference has the launchers but no driver; this is the
// standard loop on them): line integrals p (natural view order), the OSTR
// subset order (slot m*Vs+j = view j*M+m); per subset: SART forward of the
// subset (cos-weighted when cos_weight), diff, [Huber prior when beta != 0],
// backprojection update with relaxation lambda.
template <typename T>
int run_sart(const ostr_geom& gg, const float* angles, int n_views, const T* p_in,
             const T* x_init, int M, int n_iters, T lambda, T beta, T delta, int cos_weight,
             int nthr, T* x_out, T* snapshots) {
    Geo<T> g(gg);
    if (M <= 0 || n_views % M != 0) return OSTR_E_ARG;
    if (M > 0 && n_views / M > 64) return OSTR_E_RANGE;
    if (beta != 0 && delta == 0) return OSTR_E_ARG;
    const int vs = n_views / M;
    const int64_t nbin = g.nbin(), nvox = g.nvox();
    std::vector<int> idx(n_views);
    for (int m = 0, k = 0; m < M; ++m)
        for (int j = 0; j < vs; ++j) idx[k++] = j * M + m;
    std::vector<T> p((size_t)n_views * nbin);
    regroup(p.data(), p_in, n_views, M, nbin);
    std::vector<T> x(nvox), est(vs * nbin), norm(vs * nbin), diff(vs * nbin), pri(nvox);
    if (x_init)
        std::memcpy(x.data(), x_init, nvox * sizeof(T));
    else
        std::fill(x.begin(), x.end(), (T)0);
    for (int it = 0; it < n_iters; ++it) {
        for (int m = 0; m < M; ++m) {
            int a0 = m * vs, a1 = a0 + vs;
            project(g, angles, idx.data(), x.data(), est.data(), a0, a1, 0, cos_weight, nthr);
            project(g, angles, idx.data(), (const T*)nullptr, norm.data(), a0, a1, 2, 0, nthr);
            sart_prj_diff(diff.data(), p.data() + a0 * nbin, est.data(), norm.data(), vs * nbin);
            if (beta != 0) {
                std::fill(pri.begin(), pri.end(), (T)0);
                prior_huber(g, pri.data(), x.data(), delta, nthr);
            }
            backproject_sart(g, angles, idx.data(), x.data(), diff.data(),
                             beta != 0 ? pri.data() : (const T*)nullptr, a0, a1, lambda, beta,
                             nthr);
        }
        if (snapshots) std::memcpy(snapshots + (int64_t)it * nvox, x.data(), nvox * sizeof(T));
    }
    std::memcpy(x_out, x.data(), nvox * sizeof(T));
    return 0;
}

// ------------------------------------------------- 2x2 supersampled forward --
// ray_trace_4R_gpu_manyangles_direct_notexturememory (Tomo_OSTR.cu:1453-1515):
// one ray per sub-bin of a 2x2 subdivided detector, sub-bin centre
// x_d0 + (i + 0.5)/2 * Bsize_x (double, rounded), plain sum (no cos), written
// detector-row-major [slot][2 nby][2 nbx]; the reference leaves missed sub-bins
// unwritten, here they are 0 (equal when the buffer starts zeroed).
// merge_4R_to_1R_global (:1517-1550): prj[slot][j][i] = (s(2i,2j) + s(2i+1,2j)
// + s(2i,2j+1) + s(2i+1,2j+1)) / 4, summed left to right. The reference's merge
// assigns its third index twice and reads an uninitialised fourth
// (:1545-1547); the restatement averages the four sub-bins.
template <typename T>
void project_4r(const Geo<T>& g, const float* angles, const int* index, const T* img, T* prj,
                T* prj4, int a0, int a1, int nthr) {
    const int nbx2 = 2 * g.nbx, nby2 = 2 * g.nby;
    const T hbx = g.dbx * (T)0.5, hby = g.dby * (T)0.5;
    int rows = (a1 - a0) * nby2;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads_or_default(nthr))
    for (int r = 0; r < rows; ++r) {
        int a = a0 + r / nby2;
        int jj = r % nby2;
        ViewPos<T> v = view_pos(g, angles[index[a]]);
        T ey = cell_centre(g.y_d0, jj, hby);
        T* out = prj4 + ((int64_t)(a - a0) * nby2 + jj) * nbx2;
        for (int ii = 0; ii < nbx2; ++ii) {
            T ex = cell_centre(g.x_d0, ii, hbx);
            bool hit;
            T len;
            T s = walk<T, T>(g, img, v.x0, g.src_y, v.z0, ex, ey, -g.det_r, &hit, &len, nullptr);
            out[ii] = hit ? s : (T)0;
        }
    }
    const int64_t n_out = (int64_t)(a1 - a0) * g.nby;
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthr))
    for (int64_t r = 0; r < n_out; ++r) {
        int64_t a = r / g.nby;
        int j = (int)(r % g.nby);
        const T* s0 = prj4 + (a * nby2 + 2 * j) * nbx2;
        const T* s1 = s0 + nbx2;
        T* out = prj + (a * g.nby + j) * g.nbx;
        for (int i = 0; i < g.nbx; ++i)
            out[i] = (((s0[2 * i] + s0[2 * i + 1]) + s1[2 * i]) + s1[2 * i + 1]) / (T)4;
    }
}

// -------------------------------------------------------- objective -----
template <typename T>
double objective(const Geo<T>& g, const float* angles, const int* index, int n_views,
                 const T* x, const T* y, const T* b, const T* r, int potential, T beta,
                 T delta, T rho, int nthr) {
    std::vector<T> l((size_t)(n_views * g.nbin()));
    project(g, angles, index, x, l.data(), 0, n_views, 0, 1, nthr);
    double data = 0;
    for (size_t i = 0; i < l.size(); ++i) {
        double yb = (double)b[i] * std::exp(-(double)l[i]) + (r ? (double)r[i] : 0.0);
        data += y[i] == 0 ? yb : yb - (double)y[i] * std::log(yb);
    }
    return data + (double)beta * penalty(g, x, potential, delta, rho);
}

// ------------------------------------------------------------ driver ----
// backprojector 0: the reference's SBP for d and capL (Tomo_OSTR_gold.cpp
// :497, :759). 1: the exact transpose G^T of the forward model G = cos(t) A
// (backproject_clip), with d = G^T(yry * G 1) - the separable-quadratic-
// surrogate denominator for which the iteration is a descent method on the
// objective (with SBP, B != G^T and it is not); gamma_mode is not used.
struct RunOpts {
    int n_subsets, n_iters, potential, gamma_mode, guard, nthreads;
    double beta, delta;
    int backprojector;
};

template <typename T>
int run(const ostr_geom& gg, const float* angles, int n_views, const T* y_in, const T* b_in,
        const T* r_in, const T* x_init, const RunOpts& o, T* x_out, T* snapshots) {
    Geo<T> g(gg);
    const Geo<float> gf(gg);
    const int M = o.n_subsets;
    if (M <= 0 || n_views % M != 0) return OSTR_E_ARG;
    if (g.dx != g.dy) return OSTR_E_ARG;
    if (o.potential != OSTR_PRIOR_Q2 && o.delta == 0) return OSTR_E_ARG;
    if (o.backprojector != 0 && o.backprojector != 1) return OSTR_E_ARG;
    const bool matched = o.backprojector == 1;
    const int vs = n_views / M;
    const int64_t nbin = g.nbin(), nvox = g.nvox(), nall = nbin * n_views;
    std::vector<int> index(n_views);
    for (int i = 0; i < n_views; ++i) index[i] = i;
    T rho = g.dz / g.dx;
    T beta = (T)o.beta, delta = (T)o.delta;

    // precompute: yry, gamma, gamma*yry, d (Tomo_OSTR_gold.cpp:415-560)
    std::vector<T> c(nall), gam(nall), d(nvox, (T)0);
    yry(c.data(), y_in, r_in, nall);
    if (matched) {  // gamma = G 1 (cos-weighted row sums)
        std::vector<T> ones(nvox, (T)1);
        project(g, angles, index.data(), ones.data(), gam.data(), 0, n_views, 0, 1,
                o.nthreads);
    } else {
        project(g, angles, index.data(), (const T*)nullptr, gam.data(), 0, n_views,
                o.gamma_mode == OSTR_GAMMA_ROW_SUM ? 2 : 1, 0, o.nthreads);
    }
    for (int64_t i = 0; i < nall; ++i) gam[i] = c[i] * gam[i];
    if (matched)
        backproject_clip(g, gf, angles, index.data(), d.data(), gam.data(), 0, n_views, 1,
                         o.nthreads);
    else
        backproject_sbp(g, angles, index.data(), d.data(), gam.data(), 0, n_views, o.nthreads);

    // regroup data and view order (:602-642)
    std::vector<T> y(nall), b(nall), r;
    regroup(y.data(), y_in, n_views, M, nbin);
    regroup(b.data(), b_in, n_views, M, nbin);
    if (r_in) {
        r.resize(nall);
        regroup(r.data(), r_in, n_views, M, nbin);
    }
    std::vector<int> idx(n_views);
    for (int m = 0, k = 0; m < M; ++m)
        for (int j = 0; j < vs; ++j) idx[k++] = j * M + m;

    std::vector<T> x(nvox), l(vs * nbin), h(vs * nbin), capl(nvox), rd(nvox), rdd(nvox);
    if (x_init)
        std::memcpy(x.data(), x_init, nvox * sizeof(T));
    else
        std::fill(x.begin(), x.end(), (T)0);
    for (int it = 0; it < o.n_iters; ++it) {
        for (int m = 0; m < M; ++m) {
            int a0 = m * vs, a1 = (m + 1) * vs;
            project(g, angles, idx.data(), x.data(), l.data(), a0, a1, 0, 1, o.nthreads);
            compute_h(h.data(), y.data() + a0 * nbin, b.data() + a0 * nbin, l.data(),
                      r_in ? r.data() + a0 * nbin : (const T*)nullptr, vs * nbin);
            std::fill(capl.begin(), capl.end(), (T)0);
            if (matched)
                backproject_clip(g, gf, angles, idx.data(), capl.data(), h.data(), a0, a1, 1,
                                 o.nthreads);
            else
                backproject_sbp(g, angles, idx.data(), capl.data(), h.data(), a0, a1,
                                o.nthreads);
            std::fill(rd.begin(), rd.end(), (T)0);
            std::fill(rdd.begin(), rdd.end(), (T)0);
            prior(g, rdd.data(), rd.data(), x.data(), o.potential, delta, rho, o.nthreads);
            update_est(x.data(), capl.data(), rd.data(), d.data(), rdd.data(), beta, (T)M,
                       o.guard, nvox);
        }
        if (snapshots) std::memcpy(snapshots + (int64_t)it * nvox, x.data(), nvox * sizeof(T));
    }
    std::memcpy(x_out, x.data(), nvox * sizeof(T));
    return 0;
}

}  // namespace

// ======================================================= C ABI exports ====
#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

// Forward / gamma -----------------------------------------------------------
ORC_EXPORT int orc_fproj_f32(const ostr_geom* g, const float* ang, const int* idx,
                             const float* img, float* prj, int a0, int a1, int cos_w, int nt) {
    project(Geo<float>(*g), ang, idx, img, prj, a0, a1, 0, cos_w, nt);
    return 0;
}
ORC_EXPORT int orc_fproj_f64(const ostr_geom* g, const float* ang, const int* idx,
                             const double* img, double* prj, int a0, int a1, int cos_w,
                             int nt) {
    project(Geo<double>(*g), ang, idx, img, prj, a0, a1, 0, cos_w, nt);
    return 0;
}
ORC_EXPORT int orc_gamma_f32(const ostr_geom* g, const float* ang, const int* idx,
                             float* out, int a0, int a1, int mode, int nt) {
    project(Geo<float>(*g), ang, idx, (const float*)nullptr, out, a0, a1,
            mode == OSTR_GAMMA_ROW_SUM ? 2 : 1, 0, nt);
    return 0;
}
ORC_EXPORT int orc_gamma_f64(const ostr_geom* g, const float* ang, const int* idx,
                             double* out, int a0, int a1, int mode, int nt) {
    project(Geo<double>(*g), ang, idx, (const double*)nullptr, out, a0, a1,
            mode == OSTR_GAMMA_ROW_SUM ? 2 : 1, 0, nt);
    return 0;
}
ORC_EXPORT int64_t orc_count_nnz(const ostr_geom* g, const float* ang, const int* idx,
                                 int a0, int a1, int nt) {
    return count_nnz(Geo<float>(*g), ang, idx, a0, a1, nt);
}
// Single ray walk (for closed-form tests): returns sum, writes hit and length.
ORC_EXPORT double orc_walk_f64(const ostr_geom* g, const double* img, const double* src,
                               const double* dst, int* hit, double* length) {
    bool h;
    double len;
    double s = walk<double, double>(Geo<double>(*g), img, src[0], src[1], src[2], dst[0],
                                    dst[1], dst[2], &h, &len, nullptr);
    *hit = h;
    *length = len;
    return s;
}
ORC_EXPORT float orc_walk_f32(const ostr_geom* g, const float* img, const float* src,
                              const float* dst, int* hit, float* length) {
    bool h;
    float len;
    float s = walk<float, float>(Geo<float>(*g), img, src[0], src[1], src[2], dst[0], dst[1],
                                 dst[2], &h, &len, nullptr);
    *hit = h;
    *length = len;
    return s;
}

// Backprojectors -------------------------------------------------------------
ORC_EXPORT int orc_bproj_sbp_f32(const ostr_geom* g, const float* ang, const int* idx,
                                 float* img, const float* prj, int a0, int a1, int nt) {
    backproject_sbp(Geo<float>(*g), ang, idx, img, prj, a0, a1, nt);
    return 0;
}
ORC_EXPORT int orc_bproj_sbp_f64(const ostr_geom* g, const float* ang, const int* idx,
                                 double* img, const double* prj, int a0, int a1, int nt) {
    backproject_sbp(Geo<double>(*g), ang, idx, img, prj, a0, a1, nt);
    return 0;
}
ORC_EXPORT int orc_bproj_matched_f32(const ostr_geom* g, const float* ang, const int* idx,
                                     float* img, const float* prj, int a0, int a1, int cw) {
    backproject_matched(Geo<float>(*g), ang, idx, img, prj, a0, a1, cw);
    return 0;
}
ORC_EXPORT int orc_bp_init_f64(const ostr_geom* g, const float* ang, int nv, const double* p,
                               double* img) {
    bp_init(Geo<double>(*g), Geo<float>(*g), ang, nv, p, img);
    return 0;
}
ORC_EXPORT int orc_bproj_clip_f64(const ostr_geom* g, const float* ang, const int* idx,
                                  double* img, const double* prj, int a0, int a1, int cw,
                                  int nt) {
    backproject_clip(Geo<double>(*g), Geo<float>(*g), ang, idx, img, prj, a0, a1, cw, nt);
    return 0;
}
ORC_EXPORT int orc_bproj_matched_f64(const ostr_geom* g, const float* ang, const int* idx,
                                     double* img, const double* prj, int a0, int a1, int cw) {
    backproject_matched(Geo<double>(*g), ang, idx, img, prj, a0, a1, cw);
    return 0;
}

// Prior ---------------------------------------------------------------------
ORC_EXPORT int orc_prior_f32(const ostr_geom* g, float* rdd, float* rd, const float* x,
                             int pot, float delta, float rho, int nt) {
    if (pot != OSTR_PRIOR_Q2 && delta == 0) return OSTR_E_ARG;
    prior(Geo<float>(*g), rdd, rd, x, pot, delta, rho, nt);
    return 0;
}
ORC_EXPORT int orc_prior_f64(const ostr_geom* g, double* rdd, double* rd, const double* x,
                             int pot, double delta, double rho, int nt) {
    if (pot != OSTR_PRIOR_Q2 && delta == 0) return OSTR_E_ARG;
    prior(Geo<double>(*g), rdd, rd, x, pot, delta, rho, nt);
    return 0;
}
ORC_EXPORT double orc_penalty_f64(const ostr_geom* g, const double* x, int pot, double delta,
                                  double rho) {
    return penalty(Geo<double>(*g), x, pot, delta, rho);
}

// Elementwise ---------------------------------------------------------------
ORC_EXPORT void orc_yry_f32(float* o, const float* y, const float* r, int64_t n) {
    yry(o, y, r, n);
}
ORC_EXPORT void orc_yry_f64(double* o, const double* y, const double* r, int64_t n) {
    yry(o, y, r, n);
}
ORC_EXPORT void orc_compute_h_f32(float* h, const float* y, const float* b, const float* l,
                                  const float* r, int64_t n) {
    compute_h(h, y, b, l, r, n);
}
ORC_EXPORT void orc_compute_h_f64(double* h, const double* y, const double* b,
                                  const double* l, const double* r, int64_t n) {
    compute_h(h, y, b, l, r, n);
}
ORC_EXPORT void orc_update_est_f32(float* x, const float* capl, const float* rd,
                                   const float* d, const float* rdd, float beta, float scale,
                                   int guard, int64_t n) {
    update_est(x, capl, rd, d, rdd, beta, scale, guard, n);
}
ORC_EXPORT void orc_update_est_f64(double* x, const double* capl, const double* rd,
                                   const double* d, const double* rdd, double beta,
                                   double scale, int guard, int64_t n) {
    update_est(x, capl, rd, d, rdd, beta, scale, guard, n);
}
ORC_EXPORT void orc_regroup_f32(float* dst, const float* src, int nv, int m, int64_t nbin) {
    regroup(dst, src, nv, m, nbin);
}

// Objective and driver ------------------------------------------------------
ORC_EXPORT double orc_objective_f64(const ostr_geom* g, const float* ang, int nv,
                                    const double* x, const double* y, const double* b,
                                    const double* r, int pot, double beta, double delta,
                                    int nt) {
    std::vector<int> idx(nv);
    for (int i = 0; i < nv; ++i) idx[i] = i;
    Geo<double> G(*g);
    return objective(G, ang, idx.data(), nv, x, y, b, r, pot, beta, delta, G.dz / G.dx, nt);
}

ORC_EXPORT int orc_run_f32(const ostr_geom* g, const float* ang, int nv, const float* y,
                           const float* b, const float* r, const float* x0, int n_subsets,
                           int n_iters, double beta, double delta, int pot, int gamma_mode,
                           int guard, int backprojector, int nt, float* x_out,
                           float* snapshots) {
    RunOpts o{n_subsets, n_iters, pot, gamma_mode, guard, nt, beta, delta, backprojector};
    return run<float>(*g, ang, nv, y, b, r, x0, o, x_out, snapshots);
}
// 2x2 supersampled forward -------------------------------------------------------
ORC_EXPORT int orc_fproj_4r_f32(const ostr_geom* g, const float* ang, const int* idx,
                                const float* img, float* prj, float* prj4, int a0, int a1,
                                int nt) {
    project_4r(Geo<float>(*g), ang, idx, img, prj, prj4, a0, a1, nt);
    return 0;
}
ORC_EXPORT int orc_fproj_4r_f64(const ostr_geom* g, const float* ang, const int* idx,
                                const double* img, double* prj, double* prj4, int a0, int a1,
                                int nt) {
    project_4r(Geo<double>(*g), ang, idx, img, prj, prj4, a0, a1, nt);
    return 0;
}

// OS-SART --------------------------------------------------------------------
ORC_EXPORT void orc_sart_prj_diff_f32(float* diff, const float* p, const float* est,
                                      const float* norm, int64_t n) {
    sart_prj_diff(diff, p, est, norm, n);
}
ORC_EXPORT void orc_sart_prj_diff_f64(double* diff, const double* p, const double* est,
                                      const double* norm, int64_t n) {
    sart_prj_diff(diff, p, est, norm, n);
}
ORC_EXPORT int orc_bproj_sart_f32(const ostr_geom* g, const float* ang, const int* idx,
                                  float* img, const float* diff, const float* prior, int a0,
                                  int a1, float lambda, float beta, int nt) {
    backproject_sart(Geo<float>(*g), ang, idx, img, diff, prior, a0, a1, lambda, beta, nt);
    return 0;
}
ORC_EXPORT int orc_bproj_sart_f64(const ostr_geom* g, const float* ang, const int* idx,
                                  double* img, const double* diff, const double* prior,
                                  int a0, int a1, double lambda, double beta, int nt) {
    backproject_sart(Geo<double>(*g), ang, idx, img, diff, prior, a0, a1, lambda, beta, nt);
    return 0;
}
ORC_EXPORT int orc_prior_huber_f32(const ostr_geom* g, float* prior, const float* x,
                                   float delta, int nt) {
    if (delta == 0) return OSTR_E_ARG;
    prior_huber(Geo<float>(*g), prior, x, delta, nt);
    return 0;
}
ORC_EXPORT int orc_prior_huber_f64(const ostr_geom* g, double* prior, const double* x,
                                   double delta, int nt) {
    if (delta == 0) return OSTR_E_ARG;
    prior_huber(Geo<double>(*g), prior, x, delta, nt);
    return 0;
}
ORC_EXPORT int orc_run_sart_f32(const ostr_geom* g, const float* ang, int nv, const float* p,
                                const float* x0, int n_subsets, int n_iters, float lambda,
                                float beta, float delta, int cos_weight, int nt, float* x_out,
                                float* snapshots) {
    return run_sart<float>(*g, ang, nv, p, x0, n_subsets, n_iters, lambda, beta, delta,
                           cos_weight, nt, x_out, snapshots);
}
ORC_EXPORT int orc_run_sart_f64(const ostr_geom* g, const float* ang, int nv,
                                const double* p, const double* x0, int n_subsets, int n_iters,
                                double lambda, double beta, double delta, int cos_weight,
                                int nt, double* x_out, double* snapshots) {
    return run_sart<double>(*g, ang, nv, p, x0, n_subsets, n_iters, lambda, beta, delta,
                            cos_weight, nt, x_out, snapshots);
}

ORC_EXPORT int orc_run_f64(const ostr_geom* g, const float* ang, int nv, const double* y,
                           const double* b, const double* r, const double* x0, int n_subsets,
                           int n_iters, double beta, double delta, int pot, int gamma_mode,
                           int guard, int backprojector, int nt, double* x_out,
                           double* snapshots) {
    RunOpts o{n_subsets, n_iters, pot, gamma_mode, guard, nt, beta, delta, backprojector};
    return run<double>(*g, ang, nv, y, b, r, x0, o, x_out, snapshots);
}
