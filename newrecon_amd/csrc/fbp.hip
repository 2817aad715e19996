// fbp.hip - the row filter of the DBT FBP: cosine weight and FFT convolution
// in LDS.
//
// The filter is a convolution along detector x (the tube-motion direction)
// with odd-length centred taps, samples off the detector 0:
//   q[a][i][j] = sum_i' taps[i' - i + c] pw[a][i'][j],  c = (ntaps - 1) / 2,
// pw the cosine-weighted views. Only taps with |i' - i| <= nbx - 1 meet data,
// so with L the power of two >= 2 nbx - 1 the sum is exactly the circular
// convolution of the zero-padded column with
//   H[m mod L] = taps[c - m],  |m| <= nbx - 1 (0 elsewhere),
// q = IFFT(FFT(pw column) FFT(H)). Two real columns ride in one complex
// sequence (column j real, j + 1 imaginary): H is real, so the convolution
// keeps them apart. Per column pair a forward and an inverse radix-2 FFT
// (L/2 log2 L butterflies each, two stages per LDS pass) replace the 2 nbx^2
// multiply-adds of the direct sum.
//
// fbp_spectrum_kernel (one group) forms FFT(H) / L and the twiddle table
// exp(-2 pi i k / L), k < L/2, rounded from double; fbp_filter_kernel gives
// each group SEQ column pairs of one view: it loads them weighted and in
// bit-reversed order into LDS, runs the decimation-in-time forward transform
// (natural order out), multiplies by the spectrum, runs the
// decimation-in-frequency inverse with conjugate twiddles (bit-reversed order
// out) and stores rows i < nbx from their bit-reversed positions, so no
// reordering pass is needed.
#include "ostr_internal.h"

namespace ostr {

namespace {

constexpr int FFT_THREADS = 256;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
    return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
    return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ unsigned bitrev(unsigned i, int logl) {
    return __builtin_bitreverse32(i) >> (32 - logl);
}

// Forward transform, bit-reversed order in, natural order out, on the seq
// sequences of L = 2^logl elements at buf + s L. Radix-2 decimation in time,
// two stages at a time: the quadruple (i0, i0 + m, i0 + 2m, i0 + 3m) of stage
// half-size m goes through both radix-2 butterflies in registers (the same
// operations, one LDS round trip and barrier instead of two); an odd first
// stage runs alone. Ends with a barrier.
__device__ void fft_dit(float2* buf, int seq, int logl, const float2* tw) {
    const int L = 1 << logl;
    int lg = 0;
    if (logl & 1) {
        __syncthreads();
        for (int b = threadIdx.x; b < L / 2; b += FFT_THREADS) {
            const int i0 = b << 1;
            for (int s = 0; s < seq; ++s) {
                float2* a = buf + s * L;
                const float2 u = a[i0], o = cmul(a[i0 + 1], tw[0]);
                a[i0] = cadd(u, o);
                a[i0 + 1] = csub(u, o);
            }
        }
        lg = 1;
    }
    for (; lg < logl; lg += 2) {
        const int m = 1 << lg;
        __syncthreads();
        for (int b = threadIdx.x; b < L / 4; b += FFT_THREADS) {
            const int k = b & (m - 1);
            const int i0 = ((b >> lg) << (lg + 2)) | k;
            const float2 w = tw[k << (logl - 1 - lg)];
            const float2 w4a = tw[k << (logl - 2 - lg)];
            const float2 w4b = tw[(k + m) << (logl - 2 - lg)];
            for (int s = 0; s < seq; ++s) {
                float2* a = buf + s * L;
                const float2 e0 = a[i0], e1 = a[i0 + m], e2 = a[i0 + 2 * m], e3 = a[i0 + 3 * m];
                const float2 o1 = cmul(e1, w), o3 = cmul(e3, w);
                const float2 A0 = cadd(e0, o1), A1 = csub(e0, o1);
                const float2 A2 = cadd(e2, o3), A3 = csub(e2, o3);
                const float2 p2 = cmul(A2, w4a), p3 = cmul(A3, w4b);
                a[i0] = cadd(A0, p2);
                a[i0 + 2 * m] = csub(A0, p2);
                a[i0 + m] = cadd(A1, p3);
                a[i0 + 3 * m] = csub(A1, p3);
            }
        }
    }
    __syncthreads();
}

// Inverse transform without the 1/L, natural order in, bit-reversed order out:
// radix-2 decimation in frequency with conjugate twiddles, two stages at a
// time (the quadruple (i0, i0 + M/2, i0 + M, i0 + 3M/2) of stage half-size M
// through both butterflies in registers); an odd last stage runs alone.
__device__ void fft_dif_inverse(float2* buf, int seq, int logl, const float2* tw) {
    const int L = 1 << logl;
    auto ctw = [&](int i) {
        const float2 t = tw[i];
        return make_float2(t.x, -t.y);
    };
    int lg = logl - 1;
    for (; lg >= 1; lg -= 2) {
        const int M = 1 << lg, h = M >> 1;
        __syncthreads();
        for (int b = threadIdx.x; b < L / 4; b += FFT_THREADS) {
            const int k = b & (h - 1);
            const int i0 = ((b >> (lg - 1)) << (lg + 1)) | k;
            const float2 wk = ctw(k << (logl - 1 - lg));
            const float2 wh = ctw((k + h) << (logl - 1 - lg));
            const float2 v = ctw(k << (logl - lg));
            for (int s = 0; s < seq; ++s) {
                float2* a = buf + s * L;
                const float2 e0 = a[i0], e1 = a[i0 + h], e2 = a[i0 + M], e3 = a[i0 + M + h];
                const float2 A0 = cadd(e0, e2), A2 = cmul(csub(e0, e2), wk);
                const float2 A1 = cadd(e1, e3), A3 = cmul(csub(e1, e3), wh);
                a[i0] = cadd(A0, A1);
                a[i0 + h] = cmul(csub(A0, A1), v);
                a[i0 + M] = cadd(A2, A3);
                a[i0 + M + h] = cmul(csub(A2, A3), v);
            }
        }
    }
    if (lg == 0) {
        __syncthreads();
        for (int b = threadIdx.x; b < L / 2; b += FFT_THREADS) {
            const int i0 = b << 1;
            const float2 w = ctw(0);
            for (int s = 0; s < seq; ++s) {
                float2* a = buf + s * L;
                const float2 u = a[i0], v = a[i0 + 1];
                a[i0] = cadd(u, v);
                a[i0 + 1] = cmul(csub(u, v), w);
            }
        }
    }
    __syncthreads();
}

// One group: the twiddles tw (global, L/2) and hs = FFT(H) / L (global, L).
// LDS: L + L/2 float2.
__global__ __launch_bounds__(FFT_THREADS) void fbp_spectrum_kernel(const float* __restrict__ taps,
                                                                   int ntaps, int nbx, int logl,
                                                                   float2* __restrict__ tw_out,
                                                                   float2* __restrict__ hs) {
    extern __shared__ float2 lds[];
    const int L = 1 << logl, c = (ntaps - 1) / 2;
    float2* tw = lds;
    float2* a = lds + L / 2;
    for (int k = threadIdx.x; k < L / 2; k += FFT_THREADS) {
        double sn, cs;
        sincospi(-2.0 * (double)k / (double)L, &sn, &cs);
        const float2 w = make_float2((float)cs, (float)sn);
        tw[k] = w;
        tw_out[k] = w;
    }
    for (int p = threadIdx.x; p < L; p += FFT_THREADS) {
        const int m = p < L / 2 ? p : p - L;
        const int t = c - m;
        const bool on = m >= -(nbx - 1) && m <= nbx - 1 && t >= 0 && t < ntaps;
        a[bitrev((unsigned)p, logl)] = make_float2(on ? taps[t] : 0.0f, 0.0f);
    }
    fft_dit(a, 1, logl, tw);
    const float inv = 1.0f / (float)L;  // a power of two: exact
    for (int p = threadIdx.x; p < L; p += FFT_THREADS)
        hs[p] = make_float2(a[p].x * inv, a[p].y * inv);
}

// Group b of an n-group launch runs on XCD b % 8; hand each XCD one contiguous
// range of task indices so that neighbouring column chunks, which share cache
// lines of every detector row, meet in the same L2.
__device__ __forceinline__ int xcd_contiguous(int b, int n) {
    const int x = b & 7, q = n >> 3, r = n & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Group = SEQ column pairs (2 SEQ columns from j0) of one slot. LDS: L/2 + SEQ L
// float2. p and q are [slot][nbx][nby] from slot a0; vec: nby % 4 == 0 and p,
// q 16-byte aligned (SEQ = 2: float4 row loads and stores).
template <int SEQ>
__global__ __launch_bounds__(FFT_THREADS) void fbp_filter_kernel(
    DevGeom g, const ViewTab* __restrict__ views, int a0, int n_slots, int n_chunks, int logl,
    const float2* __restrict__ tw_g, const float2* __restrict__ hs, const float* __restrict__ p,
    float* __restrict__ q, int vec) {
    extern __shared__ float2 lds[];
    const int L = 1 << logl;
    float2* tw = lds;
    float2* buf = lds + L / 2;
    const int task = xcd_contiguous((int)blockIdx.x, n_chunks * n_slots);
    const int slot = task / n_chunks;
    const int j0 = (task % n_chunks) * 2 * SEQ;
    const int nbx = g.nbx, nby = g.nby;
    const ViewTab v = views[a0 + slot];
    const float D = v.z0 + g.det_r;
    float dy[2 * SEQ];
#pragma unroll
    for (int c = 0; c < 2 * SEQ; ++c) {
        const float ey = (float)((double)g.y_d0 + ((double)(j0 + c) + 0.5) * (double)g.dby);
        dy[c] = ey - g.src_y;
    }
    const int64_t base = (int64_t)slot * nbx * nby + j0;

    for (int k = threadIdx.x; k < L / 2; k += FFT_THREADS) tw[k] = tw_g[k];
    for (int i = threadIdx.x; i < L; i += FFT_THREADS) {
        float val[2 * SEQ];
#pragma unroll
        for (int c = 0; c < 2 * SEQ; ++c) val[c] = 0.0f;
        if (i < nbx) {
            const float ex = (float)((double)g.x_d0 + ((double)i + 0.5) * (double)g.dbx);
            const float dx = ex - v.x0;
            const float* row = p + base + (int64_t)i * nby;
            if (SEQ == 2 && vec) {
                const float4 r4 = *(const float4*)row;
                val[0] = r4.x;
                val[1] = r4.y;
                val[2] = r4.z;
                val[3] = r4.w;
            } else {
#pragma unroll
                for (int c = 0; c < 2 * SEQ; ++c) val[c] = j0 + c < nby ? row[c] : 0.0f;
            }
#pragma unroll
            for (int c = 0; c < 2 * SEQ; ++c)
                val[c] *= D / sqrtf(fmaf(dx, dx, fmaf(dy[c], dy[c], D * D)));
        }
        const unsigned r = bitrev((unsigned)i, logl);
#pragma unroll
        for (int s = 0; s < SEQ; ++s) buf[s * L + r] = make_float2(val[2 * s], val[2 * s + 1]);
    }
    fft_dit(buf, SEQ, logl, tw);
    for (int f = threadIdx.x; f < L; f += FFT_THREADS) {
        const float2 h = hs[f];
#pragma unroll
        for (int s = 0; s < SEQ; ++s) buf[s * L + f] = cmul(buf[s * L + f], h);
    }
    fft_dif_inverse(buf, SEQ, logl, tw);
    for (int i = threadIdx.x; i < nbx; i += FFT_THREADS) {
        const unsigned r = bitrev((unsigned)i, logl);
        float val[2 * SEQ];
#pragma unroll
        for (int s = 0; s < SEQ; ++s) {
            const float2 y = buf[s * L + r];
            val[2 * s] = y.x;
            val[2 * s + 1] = y.y;
        }
        float* row = q + base + (int64_t)i * nby;
        if (SEQ == 2 && vec) {
            *(float4*)row = make_float4(val[0], val[1], val[2], val[3]);
        } else {
#pragma unroll
            for (int c = 0; c < 2 * SEQ; ++c)
                if (j0 + c < nby) row[c] = val[c];
        }
    }
}

// Dynamic LDS beyond 64 KiB per group needs the kernel's opt-in.
hipError_t allow_lds(const void* kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace

int fbp_fft_log2(int nbx) {
    int logl = 1;
    while ((1 << logl) < 2 * nbx - 1) ++logl;
    return logl;
}

hipError_t launch_fbp_filter(const DevGeom& g, const ViewTab* views, int a0, int a1,
                             const float* taps, int ntaps, const float* p, float2* spec,
                             float* q, int lds_limit, hipStream_t st) {
    const int n_slots = a1 - a0;
    if (n_slots <= 0) return hipSuccess;
    const int logl = fbp_fft_log2(g.nbx);
    const int L = 1 << logl;
    float2* tw = spec;        // L/2
    float2* hs = spec + L / 2;  // L
    const size_t spec_lds = (size_t)(L + L / 2) * sizeof(float2);
    const size_t lds2 = (size_t)(L / 2 + 2 * L) * sizeof(float2);
    const size_t lds1 = (size_t)(L / 2 + L) * sizeof(float2);
    if (lds1 > (size_t)lds_limit || spec_lds > (size_t)lds_limit) return hipErrorInvalidValue;
    hipError_t e = allow_lds((const void*)fbp_spectrum_kernel, spec_lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fbp_spectrum_kernel, dim3(1), dim3(FFT_THREADS), spec_lds, st, taps,
                       ntaps, g.nbx, logl, tw, hs);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int vec = (g.nby % 4 == 0) && ((uintptr_t)p % 16 == 0) && ((uintptr_t)q % 16 == 0);
    if (lds2 <= (size_t)lds_limit) {
        if ((e = allow_lds((const void*)fbp_filter_kernel<2>, lds2)) != hipSuccess) return e;
        const int n_chunks = (g.nby + 3) / 4;
        hipLaunchKernelGGL(fbp_filter_kernel<2>, dim3((unsigned)(n_chunks * n_slots)),
                           dim3(FFT_THREADS), lds2, st, g, views, a0, n_slots, n_chunks, logl,
                           tw, hs, p, q, vec);
    } else {
        if ((e = allow_lds((const void*)fbp_filter_kernel<1>, lds1)) != hipSuccess) return e;
        const int n_chunks = (g.nby + 1) / 2;
        hipLaunchKernelGGL(fbp_filter_kernel<1>, dim3((unsigned)(n_chunks * n_slots)),
                           dim3(FFT_THREADS), lds1, st, g, views, a0, n_slots, n_chunks, logl,
                           tw, hs, p, q, 0);
    }
    return hipGetLastError();
}

}  // namespace ostr
