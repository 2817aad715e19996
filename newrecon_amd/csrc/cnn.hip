// cnn.hip - kernels of the warm-start network: 3-D convolution (forward,
// data gradient, weight gradient) on BF16 MFMA with FP32 accumulation,
// 2x average pooling and 2x nearest upsampling with their gradients, the
// input/output/MSE head, weight packing and Adam.
//
// Tensors are channels-last BF16 views (pointer, channel stride, channel
// offset): element (voxel v, channel c) at p[v * cs + co + c], voxel
// v = (z * ny + y) * nx + x. Skip connections are free: a layer writes into
// its channel range of the concatenation buffer the next layer reads.
//
// Convolution (3x3x3, zero padding 1, stride 1) as an implicit GEMM on
// v_mfma_f32_16x16x32_bf16 (A 16x32, B 32x16, D 16x16; lane l holds A row
// l%16 / B column l%16 at k = 8*(l/16)+0..7, and D column l%16 at rows
// 4*(l/16)+0..3):
//   D[co][voxel] += sum_k W[co][k] * P[k][voxel],  k = tap * CIN + ci,
//   tap = (dz+1)*9 + (dy+1)*3 + (dx+1),  P[k][v] = X[v + (dz,dy,dx)][ci].
// A is read as 16 bytes of the packed weights, B as 16 bytes (8 channels of
// one tap) of the input from a ring of input slabs in LDS, and a lane stores 4
// consecutive output channels of one voxel (8 bytes). The data gradient is the
// same kernel on the output gradient with the flipped, transposed weights,
// masked by the layer input's ReLU (dX = 0 where X <= 0).
//
// Weight gradient: dW[co][tap][ci] = sum_v dY[v][co] X[v + off(tap)][ci],
// a GEMM reduced over voxels: A = dY (16 co x 32 voxels), B = X shifted by
// the tap (32 voxels x 16 ci), both read 16 bytes at a time from
// channel-major, row-padded copies [C][nz][ny][nxp] (three x-shifted copies
// of X, so every load is aligned). Each work-group reduces a share of the
// rows for one tap; its waves' sums are combined in LDS in a fixed order and
// written as a partial; a second kernel sums the partials in a fixed order,
// so the gradient is deterministic.
#include <initializer_list>

#include "ostr_internal.h"

namespace ostr {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int kpad(int cin) { return ((27 * cin + 31) / 32) * 32; }

__device__ __forceinline__ u16 to_bf16(float v) { return bf16_bits(v); }
__device__ __forceinline__ float from_bf16(u16 b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 load8(const u16* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}
__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int e = 0; e < 8; ++e) z[e] = (__bf16)0.0f;
    return z;
}

// ---- convolution through an LDS ring of input slabs -----------------------
// Work-group: 4 compute waves and 1-2 loader waves; tile XW voxels along x by 4R
// rows along y, marching in z over the slices [z0, z0 + zc). The input slabs
// of slices z - 1, z, z + 1 (x0 - 1 .. x0 + XW, y0 - 1 .. y0 + 4R, all CIN
// channels, zero outside the volume) sit in a 3-slot LDS ring, slice s in
// slot s mod 3, layout [c8][row][XW + 2] of 16-byte elements (8 channels): the
// 16 lanes of one B operand read 256 contiguous bytes. While the compute
// waves work on slice z, the loader waves hold slice z + 2's slab in
// registers and writes it over slice z - 1's slot between two barriers, so
// the compute waves never wait on the slab loads (a wave's memory counter is
// in order: a wait for its own weight loads would also wait for them).
// Each compute wave owns R output rows. The K sum is grouped by dy: chunk
// (dy, q) holds the pairs p = 4 q + kg -> (dz, dx, c8) = (p / (3 NC8),
// (p / NC8) % 3, p % NC8) (pairs past 9 NC8 get zero weights), so the B
// operand of input row r and chunk q serves output rows r - dyi, dyi = dy + 1,
// of all three dy: (R + 2) LDS reads per 3 R MFMAs per column block and
// output-channel block. A operands are read from the packed weights
// (tap * CIN + 8 c8, 16 bytes): all of them once, into registers, where they
// fit, else one chunk ahead.
// Slab elements per LDS slot, and loader waves: each holds at most 24 of its
// 16-byte elements per lane in flight.
constexpr int conv_slab(int cin, int r, int xw) { return cin / 8 * (4 * r + 2) * (xw + 2); }
constexpr int conv_loaders(int slab) { return (slab + 64 * 24 - 1) / (64 * 24); }

template <int CIN, int COUT, int R, int XW>
__global__ __launch_bounds__(256 + 64 * conv_loaders(conv_slab(CIN, R, XW))) void conv3d_lds_kernel(
    const u16* __restrict__ x, int xcs, int xco, const u16* __restrict__ w,
    const float* __restrict__ bias, u16* __restrict__ y, int ycs, int yco,
    const u16* __restrict__ mask, int mcs, int mco, int relu, int nz, int ny, int nx,
    int tiles_x, int tiles_y, int zc) {
    constexpr int NC8 = CIN / 8, NP = 9 * NC8, NQ = (NP + 3) / 4;
    constexpr int KP = kpad(CIN), NCO = COUT / 16, NVB = XW / 16;
    constexpr int TR = 4 * R + 2, TW = XW + 2, SLAB = conv_slab(CIN, R, XW);
    constexpr int LW = conv_loaders(SLAB), NT = 256 + 64 * LW;
    extern __shared__ uint4 slabs[];  // 3 * SLAB
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, kg = lane >> 4;
    // group b runs on XCD b % 8: give each XCD a contiguous range of tiles so
    // that neighbouring tiles, which share halo rows, meet in one L2
    const int n_groups = (int)gridDim.x;
    int t;
    {
        const int xcd = (int)blockIdx.x & 7, qq = n_groups >> 3, rr = n_groups & 7;
        t = xcd * qq + (xcd < rr ? xcd : rr) + ((int)blockIdx.x >> 3);
    }
    const int tx = t % tiles_x;
    t /= tiles_x;
    const int ty = t % tiles_y, tz = t / tiles_y;
    const int x0 = tx * XW, y0 = ty * 4 * R, z0 = tz * zc;
    const int z1 = z0 + zc < nz ? z0 + zc : nz;

    // slab element e -> (row, xi, c8), c8 fastest (contiguous global reads)
    auto src_of = [&](int e, int zi, bool& ok) -> const uint4* {
        const int c8 = e % NC8, rest = e / NC8;
        const int xi = rest % TW, row = rest / TW;
        const int gx = x0 - 1 + xi, gy = y0 - 1 + row;
        ok = e < SLAB && (unsigned)zi < (unsigned)nz && (unsigned)gx < (unsigned)nx &&
             (unsigned)gy < (unsigned)ny;
        return reinterpret_cast<const uint4*>(x + (((int64_t)zi * ny + gy) * nx + gx) * xcs +
                                              xco + c8 * 8);
    };
    auto dst_of = [&](int e, int zi) -> uint4* {
        const int c8 = e % NC8, rest = e / NC8;
        const int xi = rest % TW, row = rest / TW;
        return slabs + ((zi + 3) % 3) * SLAB + (c8 * TR + row) * TW + xi;
    };

    // the first three slabs, by all waves
    for (int zi = z0 - 1; zi <= z0 + 1; ++zi) {
        constexpr int NA = (SLAB + NT - 1) / NT;
        uint4 v[NA];
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            bool ok;
            const uint4* src = src_of((int)threadIdx.x + NT * i, zi, ok);
            v[i] = ok ? *src : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = (int)threadIdx.x + NT * i;
            if (e < SLAB) *dst_of(e, zi) = v[i];
        }
    }
    __syncthreads();

    if (wave >= 4) {
        // loaders: slab z + 2 in registers during step z, stored between the barriers
        constexpr int NL = (SLAB + 64 * LW - 1) / (64 * LW);
        const int l0 = (int)threadIdx.x - 256;
#pragma unroll 1
        for (int z = z0; z + 1 < z1; ++z) {
            uint4 v[NL];
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                bool ok;
                const uint4* src = src_of(l0 + 64 * LW * i, z + 2, ok);
                v[i] = ok ? *src : make_uint4(0u, 0u, 0u, 0u);
            }
            __syncthreads();  // slice z - 1's slot is read no more
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                const int e = l0 + 64 * LW * i;
                if (e < SLAB) *dst_of(e, z + 2) = v[i];
            }
            __syncthreads();
        }
        return;
    }

    auto load_a = [&](int q, bf16x8 (&a)[3][NCO]) {
        const int p = 4 * q + kg;
        const int c8 = p % NC8, dxi = (p / NC8) % 3, dzi = p / (3 * NC8);
#pragma unroll
        for (int dyi = 0; dyi < 3; ++dyi)
#pragma unroll
            for (int cb = 0; cb < NCO; ++cb)
                a[dyi][cb] = p < NP ? load8(w + (int64_t)(cb * 16 + n) * KP +
                                            (dzi * 9 + dyi * 3 + dxi) * CIN + c8 * 8)
                                    : zero8();
    };
    // B operands: row r of chunk q is read from LDS two row steps before its
    // MFMAs, into slot r of a ring of R + 2 row buffers (the reads of the next
    // chunk's rows 0 and 1 overlap this chunk's last two rows); scheduling
    // barriers keep the compiler from sinking the reads to their use.
    constexpr int RR = R + 2;
    constexpr bool PIPE = R == 1 && NCO == 1;
    auto bptr = [&](int z, int q) -> const uint4* {
        // pairs past the last read a valid slab element (their weights are 0)
        const int p = 4 * q + kg < NP ? 4 * q + kg : NP - 1;
        const int c8 = p % NC8, dxi = (p / NC8) % 3, dzi = p / (3 * NC8);
        return slabs + ((z + dzi + 2) % 3) * SLAB + (c8 * TR + wave * R) * TW + n + dxi;
    };
    bf16x8 ring[PIPE ? RR : 1][NVB];
    auto read_row = [&](const uint4* bp, int r, bf16x8 (&dst)[NVB]) {
#pragma unroll
        for (int nb = 0; nb < NVB; ++nb)
            dst[nb] = *reinterpret_cast<const bf16x8*>(bp + r * TW + 16 * nb);
    };
    // one chunk: the B operand of input row r serves the output rows r - dyi.
    // With one row and one channel block per wave (a few MFMAs per row) the
    // reads run two row steps ahead through the ring; otherwise the compiler's
    // own interleaving of reads and MFMAs measured faster.
    auto chunk_direct = [&](int z, int q, const bf16x8 (&a)[3][NCO],
                            f32x4 (&acc)[R][NVB][NCO]) {
        const uint4* bp = bptr(z, q);
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            bf16x8 bop[NVB];
            read_row(bp, r, bop);
#pragma unroll
            for (int dyi = 0; dyi < 3; ++dyi) {
                const int o = r - dyi;
                if (o < 0 || o >= R) continue;
#pragma unroll
                for (int nb = 0; nb < NVB; ++nb)
#pragma unroll
                    for (int cb = 0; cb < NCO; ++cb)
                        acc[o][nb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            a[dyi][cb], bop[nb], acc[o][nb][cb], 0, 0, 0);
            }
        }
    };
    auto chunk_ring = [&](int z, int q, const uint4*& bq, const bf16x8 (&a)[3][NCO],
                          f32x4 (&acc)[R][NVB][NCO]) {
        const bool more_q = q + 1 < NQ;
        const uint4* bn = bptr(z, more_q ? q + 1 : q);
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            if (r + 2 < RR)
                read_row(bq, r + 2, ring[(r + 2) % (PIPE ? RR : 1)]);
            else if (more_q)
                read_row(bn, r + 2 - RR, ring[(r + 2 - RR) % (PIPE ? RR : 1)]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int dyi = 0; dyi < 3; ++dyi) {
                const int o = r - dyi;
                if (o < 0 || o >= R) continue;
#pragma unroll
                for (int nb = 0; nb < NVB; ++nb)
#pragma unroll
                    for (int cb = 0; cb < NCO; ++cb)
                        acc[o][nb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            a[dyi][cb], ring[r % (PIPE ? RR : 1)][nb], acc[o][nb][cb], 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        bq = bn;
    };
    auto chunk = [&](int z, int q, const uint4*& bq, const bf16x8 (&a)[3][NCO],
                     f32x4 (&acc)[R][NVB][NCO]) {
        if constexpr (PIPE)
            chunk_ring(z, q, bq, a, acc);
        else
            chunk_direct(z, q, a, acc);
    };
    // VGPR budget (256 per lane: a group of 5-6 waves puts two on some SIMD).
    // The weights stay in registers for the whole kernel when they, the
    // accumulators and the row ring fit (in 220 VGPRs; 180 where the compiler
    // schedules the reads); otherwise each chunk's are loaded one chunk ahead,
    // or at the chunk where two chunks' would not fit.
    constexpr int ACC = 4 * R * NVB * NCO + (PIPE ? 4 * RR * NVB : 0);
    constexpr bool AREG = 12 * NQ * NCO + ACC <= (PIPE ? 220 : 180);
    constexpr int NQR = AREG ? NQ : 1, NQA = AREG || 24 * NCO + ACC > 200 ? 1 : 2;
    bf16x8 aw[NQR][3][NCO], ab[NQA][3][NCO];
    if constexpr (AREG) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) load_a(q, aw[q]);
    } else if constexpr (NQA == 2) {
        load_a(0, ab[0]);
    }
#pragma unroll 1
    for (int z = z0; z < z1; ++z) {
        f32x4 acc[R][NVB][NCO];
#pragma unroll
        for (int o = 0; o < R; ++o)
#pragma unroll
            for (int nb = 0; nb < NVB; ++nb)
#pragma unroll
                for (int cb = 0; cb < NCO; ++cb) acc[o][nb][cb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const uint4* bq = bptr(z, 0);
        if constexpr (PIPE) {
            read_row(bq, 0, ring[0]);
            read_row(bq, 1, ring[1]);
        }
        if constexpr (AREG) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) chunk(z, q, bq, aw[q], acc);
        } else {
#pragma unroll 1
            for (int q = 0; q < NQ; ++q) {
                if constexpr (NQA == 2) {
                    // after the last chunk: the next slice's first
                    load_a(q + 1 < NQ ? q + 1 : 0, ab[NQA - 1]);
                    chunk(z, q, bq, ab[0], acc);
#pragma unroll
                    for (int dyi = 0; dyi < 3; ++dyi)
#pragma unroll
                        for (int cb = 0; cb < NCO; ++cb) ab[0][dyi][cb] = ab[NQA - 1][dyi][cb];
                } else {
                    load_a(q, ab[0]);
                    chunk(z, q, bq, ab[0], acc);
                }
            }
        }
        // epilogue: the bias and the ReLU masks of all outputs first (4 channels
        // = 8 bytes), then the stores
        float4 bv[NCO];
#pragma unroll
        for (int cb = 0; cb < NCO; ++cb)
            bv[cb] = bias ? *reinterpret_cast<const float4*>(bias + cb * 16 + kg * 4)
                          : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint2 mk[R][NVB][NCO];
#pragma unroll
        for (int o = 0; o < R; ++o)
#pragma unroll
            for (int nb = 0; nb < NVB; ++nb)
#pragma unroll
                for (int cb = 0; cb < NCO; ++cb) {
                    const int gy = y0 + wave * R + o, gx = x0 + nb * 16 + n;
                    const int64_t v = ((int64_t)z * ny + gy) * nx + gx;
                    mk[o][nb][cb] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
                    if (mask && gy < ny && gx < nx)
                        mk[o][nb][cb] = *reinterpret_cast<const uint2*>(
                            mask + v * mcs + mco + cb * 16 + kg * 4);
                }
#pragma unroll
        for (int o = 0; o < R; ++o) {
            const int gy = y0 + wave * R + o;
            if (gy >= ny) continue;
#pragma unroll
            for (int nb = 0; nb < NVB; ++nb) {
                const int gx = x0 + nb * 16 + n;
                if (gx >= nx) continue;
                const int64_t v = ((int64_t)z * ny + gy) * nx + gx;
#pragma unroll
                for (int cb = 0; cb < NCO; ++cb) {
                    const int c0 = cb * 16 + kg * 4;
                    const uint32_t m2[2] = {mk[o][nb][cb].x, mk[o][nb][cb].y};
                    u16 ov[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        float val = acc[o][nb][cb][e];
                        if (bias) val += e == 0 ? bv[cb].x : e == 1 ? bv[cb].y : e == 2 ? bv[cb].z : bv[cb].w;
                        if (relu) val = val > 0.0f ? val : 0.0f;
                        const u16 mb = (u16)(m2[e >> 1] >> (16 * (e & 1)));
                        if (mask && !(from_bf16(mb) > 0.0f)) val = 0.0f;
                        ov[e] = to_bf16(val);
                    }
                    uint2 packed;
                    packed.x = (uint32_t)ov[0] | ((uint32_t)ov[1] << 16);
                    packed.y = (uint32_t)ov[2] | ((uint32_t)ov[3] << 16);
                    *reinterpret_cast<uint2*>(y + v * ycs + yco + c0) = packed;
                }
            }
        }
        if (z + 1 < z1) {
            __syncthreads();  // the loader overwrites slice z - 1's slot
            __syncthreads();  // slice z + 2 is in place
        }
    }
}

// ---- weight gradient through an LDS ring ------------------------------------
// dW[co][tap][ci] = sum_v dY[v][co] X[v + off(tap)][ci] and db[co] =
// sum_v dY[v][co] reduce over voxels, so both MFMA operands hold 8
// consecutive voxels of one channel (K = 32 voxels of one row). They are read
// from channels-last LDS slabs with the transposing read ds_read_b64_tr_b16:
// within each 16 lanes, lane i receives element i & 3 of the 4 that source
// lanes (i >> 2) + 4 e (e = 0..3) address, so when source lane s addresses
// channels 4 (s & 3) .. + 3 of voxel v0 + (s >> 2), lane i receives voxels
// v0 .. v0 + 3 of channel i (two reads: 8 voxels); any voxel shift works, so
// the 27 taps need no shifted copies. Work-group: an XW x TY (x, y) tile
// marching in z over [z0, z0 + zc) with the X slabs of slices z - 1 .. z + 1
// (halo included, zero outside the volume) in a 3-slot ring and dY of slice z
// in a 2-slot ring, both in the convolution's [c8][row][x] layout, the next
// slices fetched into registers during compute (no global load is waited on
// inside the step). blockIdx.y picks one 16-channel block of the output
// (cob) and one share (tg) of its 27 NCI (tap, 16-channel input block)
// accumulators, which are dealt to the 4 waves round-robin; every wave also
// sums dY against ones for the bias (share 0's wave 0 keeps it). Each
// group writes its partial sums: part[g][co][tap CIN + ci], bpart[g][co].
// Channels past CIN (CIN = 8) read channels 0..7 again; their columns are not
// written.
constexpr int wgrad_xw(int cin) { return cin >= 64 ? 32 : 64; }
constexpr int wgrad_ty(int cin, int cout) { return cin <= 16 && cout <= 16 ? 8 : 4; }
constexpr int wgrad_slab(int cin, int cout) {
    return cin / 8 * (wgrad_ty(cin, cout) + 2) * (wgrad_xw(cin) + 2);
}
constexpr int wgrad_dslab(int cin, int cout) { return 2 * wgrad_ty(cin, cout) * wgrad_xw(cin); }
// accumulator tiles of one output block, their shares (at most 28 per wave)
constexpr int wgrad_tiles(int cin) { return 27 * ((cin + 15) / 16); }
constexpr int wgrad_shares(int cin) { return (wgrad_tiles(cin) + 111) / 112; }

typedef short v4i16 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ bf16x8 tr_read8(const uint4* base, int off0, int off1) {
    typedef __attribute__((address_space(3))) v4i16 lds_v4;
    const char* b = reinterpret_cast<const char*>(base);
    typedef short v8i16 __attribute__((ext_vector_type(8)));
    const v4i16 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4*)(b + off0));
    const v4i16 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4*)(b + off1));
    const v8i16 both = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_bit_cast(bf16x8, both);
}

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void wgrad_lds_kernel(
    const u16* __restrict__ x, int xcs, int xco, const u16* __restrict__ dy, int dcs, int dco,
    float* __restrict__ part, float* __restrict__ bpart, int nz, int ny, int nx, int tiles_x,
    int tiles_y, int zc) {
    constexpr int NC8 = CIN / 8, NCI = (CIN + 15) / 16;
    constexpr int XW = wgrad_xw(CIN), WG_TY = wgrad_ty(CIN, COUT), KPR = XW / 32;
    constexpr int TR = WG_TY + 2, TW = XW + 2;
    constexpr int SLAB = wgrad_slab(CIN, COUT), DSLAB = wgrad_dslab(CIN, COUT);
    constexpr int TC = wgrad_tiles(CIN), TG = wgrad_shares(CIN);
    constexpr int TPW = (TC + TG - 1) / TG;  // tiles per work-group
    constexpr int NTW = (TPW + 3) / 4;       // tiles per wave
    extern __shared__ uint4 lds[];           // 3 X slabs, then 2 dY slabs
    uint4* xs = lds;
    uint4* ds = lds + 3 * SLAB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cob = (int)blockIdx.y / TG, tg = (int)blockIdx.y % TG;
    const int n_groups = (int)gridDim.x;
    int t;
    {
        const int xcd = (int)blockIdx.x & 7, qq = n_groups >> 3, rr = n_groups & 7;
        t = xcd * qq + (xcd < rr ? xcd : rr) + ((int)blockIdx.x >> 3);
    }
    const int group = t;
    const int tx = t % tiles_x;
    t /= tiles_x;
    const int ty = t % tiles_y, tz = t / tiles_y;
    const int x0 = tx * XW, y0 = ty * WG_TY, z0 = tz * zc;
    const int z1 = z0 + zc < nz ? z0 + zc : nz;
    const int dco_b = dco + 16 * cob;

    // This thread's slab elements e = threadIdx.x + 256 i: X slab elements
    // (e < SLAB, [c8][row][x] with the halo) and dY slab elements (SLAB <= e <
    // SLAB + DSLAB, this group's 16 output channels); their offsets inside a
    // slice (-1: outside the volume's x/y range, zero) and their LDS index
    // inside a slot, computed once.
    constexpr int NE = (SLAB + DSLAB + 255) / 256;
    int soff[NE], loff[NE];
#pragma unroll
    for (int i = 0; i < NE; ++i) {
        const int e = (int)threadIdx.x + 256 * i;
        soff[i] = -1;
        loff[i] = 0;
        if (e < SLAB) {
            const int c8 = e % NC8, rest = e / NC8;
            const int xi = rest % TW, row = rest / TW;
            const int gx = x0 - 1 + xi, gy = y0 - 1 + row;
            if ((unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny)
                soff[i] = (gy * nx + gx) * xcs + xco + c8 * 8;
            loff[i] = (c8 * TR + row) * TW + xi;
        } else if (e < SLAB + DSLAB) {
            const int f = e - SLAB;
            const int c8 = f & 1, rest = f >> 1;
            const int xi = rest % XW, row = rest / XW;
            const int gx = x0 + xi, gy = y0 + row;
            if (gx < nx && gy < ny) soff[i] = (gy * nx + gx) * dcs + dco_b + c8 * 8;
            loff[i] = (c8 * WG_TY + row) * XW + xi;
        }
    }
    const int64_t xslice = (int64_t)ny * nx * xcs, dslice = (int64_t)ny * nx * dcs;
    uint4 pre[NE];
    // X slab of slice zx; with dY (zd >= 0) also the dY of slice zd
    auto fetch = [&](int zx, int zd) {
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = (int)threadIdx.x + 256 * i;
            pre[i] = make_uint4(0u, 0u, 0u, 0u);
            if (e < SLAB) {
                if (soff[i] >= 0 && (unsigned)zx < (unsigned)nz)
                    pre[i] = *reinterpret_cast<const uint4*>(x + zx * xslice + soff[i]);
            } else if (e < SLAB + DSLAB) {
                if (soff[i] >= 0 && zd >= 0 && zd < nz)
                    pre[i] = *reinterpret_cast<const uint4*>(dy + zd * dslice + soff[i]);
            }
        }
    };
    auto store = [&](int zx, int zd) {
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = (int)threadIdx.x + 256 * i;
            if (e < SLAB)
                xs[((zx + 3) % 3) * SLAB + loff[i]] = pre[i];
            else if (e < SLAB + DSLAB && zd >= 0)
                ds[((zd + 2) % 2) * DSLAB + loff[i]] = pre[i];
        }
    };
    fetch(z0 - 1, -1);
    store(z0 - 1, -1);
    fetch(z0, -1);
    store(z0, -1);
    fetch(z0 + 1, z0);
    store(z0 + 1, z0);
    __syncthreads();

    // per-lane parts of the transposing reads: source lane s = lane & 15 of
    // group kg addresses channels 4 (s & 3) .. + 3 of voxel 8 kg + (s >> 2)
    const int s = lane & 15, kg = lane >> 4;
    const int pc = (s & 3) >> 1, hb = (s & 1) * 8, vx = 8 * kg + (s >> 2);
    f32x4 acc[NTW], accb = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < NTW; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 ones;
#pragma unroll
    for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.0f;
#pragma unroll 1
    for (int z = z0; z < z1; ++z) {
        const bool more = z + 1 < z1;
        if (more) fetch(z + 2, z + 1);
        const uint4* dslot = ds + ((z + 2) % 2) * DSLAB;
#pragma unroll 1
        for (int kc = 0; kc < KPR * WG_TY; ++kc) {
            const int r = kc / KPR, xh = (kc % KPR) * 32;
            // A: dY of voxels (x0 + xh + 8 kg .. + 7, row r), channels 16 cob .. + 15
            const int aoff = ((pc * WG_TY + r) * XW + xh + vx) * 16 + hb;
            const bf16x8 a = tr_read8(dslot, aoff, aoff + 4 * 16);
            // branch-free, so that the reads of all tiles can be issued ahead:
            // every wave sums the bias, and a wave's slots past its share's
            // last tile repeat a tile (not written)
            accb = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, ones, accb, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < NTW; ++i) {
                const int ul = wave + 4 * i < TPW ? wave + 4 * i : TPW - 1;
                const int u = tg * TPW + ul < TC ? tg * TPW + ul : TC - 1;
                const int tap = u / NCI, cb = u % NCI;
                const int dzi = tap / 9, dyi = (tap / 3) % 3, dxi = tap % 3;
                const int c8 = 2 * cb + pc < NC8 ? 2 * cb + pc : NC8 - 1;
                const uint4* xslot = xs + ((z + dzi + 2) % 3) * SLAB;
                const int boff = ((c8 * TR + r + dyi) * TW + xh + vx + dxi) * 16 + hb;
                const bf16x8 b = tr_read8(xslot, boff, boff + 4 * 16);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
            }
        }
        if (more) {
            __syncthreads();  // slice z - 1's X slot and dY slot are read no more
            store(z + 2, z + 1);
            __syncthreads();
        }
    }
    // D row co = 16 cob + 4 kg + e, column ci = lane & 15
    constexpr int K = 27 * CIN;
    float* out = part + (int64_t)group * COUT * K;
#pragma unroll
    for (int i = 0; i < NTW; ++i) {
        const int ul = wave + 4 * i, u = tg * TPW + ul;
        if (ul >= TPW || u >= TC) continue;
        const int tap = u / NCI, ci = (u % NCI) * 16 + (lane & 15);
        if (ci < CIN) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                out[(int64_t)(16 * cob + 4 * kg + e) * K + tap * CIN + ci] = acc[i][e];
        }
    }
    if (tg == 0 && wave == 0 && (lane & 15) == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) bpart[(int64_t)group * COUT + 16 * cob + 4 * kg + e] = accb[e];
    }
}

// dw[i] = sum_s part[s][i], i < n, in a fixed order: four lanes per output
// sum s = g, g + 4, ... (g = 0..3, s ascending), then (((g0 + g1) + g2) + g3).
// A work-group covers 64 outputs (grid: ceil(n / 64) groups of 256).
__global__ __launch_bounds__(256) void sum_parts_kernel(const float* __restrict__ part,
                                                        int nsplit, int64_t n,
                                                        float* __restrict__ dw) {
    __shared__ float red[4][64];
    const int il = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + il;
    float s = 0.0f;
    if (i < n) {
        // eight partials loaded before they are added (in the same order),
        // so a lane waits for one memory round trip per eight, not per term
        int k = g;
        for (; k + 28 < nsplit; k += 32) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = part[(int64_t)(k + 4 * j) * n + i];
#pragma unroll
            for (int j = 0; j < 8; ++j) s += v[j];
        }
        for (; k < nsplit; k += 4) s += part[(int64_t)k * n + i];
    }
    red[g][il] = s;
    __syncthreads();
    if (g == 0 && i < n) dw[i] = ((red[0][il] + red[1][il]) + red[2][il]) + red[3][il];
}
static unsigned sum_parts_grid(int64_t n) { return (unsigned)((n + 63) / 64); }

// ---- pooling / upsampling (channels-last views) ---------------------------
// y (nz/2, ny/2, nx/2) = mean of the 8 children of x (nz, ny, nx).
__global__ void pool_kernel(const u16* __restrict__ x, int xcs, int xco, u16* __restrict__ y,
                            int ycs, int yco, int C, int nz, int ny, int nx) {
    const int hz = nz / 2, hy = ny / 2, hx = nx / 2;
    const int64_t n = (int64_t)hz * hy * hx * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int px = (int)(p % hx), py = (int)((p / hx) % hy), pz = (int)(p / ((int64_t)hx * hy));
        float s = 0.0f;
        for (int q = 0; q < 8; ++q) {
            const int64_t v = ((int64_t)(2 * pz + (q >> 2)) * ny + 2 * py + ((q >> 1) & 1)) * nx +
                              2 * px + (q & 1);
            s += from_bf16(x[v * xcs + xco + c]);
        }
        y[p * ycs + yco + c] = to_bf16(s * 0.125f);
    }
}

// dx (nz, ny, nx) += dy(parent) / 8 where mask(child) > 0 (else dx unchanged
// plus 0); accumulate = 0 overwrites dx instead.
__global__ void pool_bwd_kernel(const u16* __restrict__ dy, int dcs, int dco,
                                u16* __restrict__ dxp, int xcs, int xco,
                                const u16* __restrict__ mask, int mcs, int mco, int accumulate,
                                int C, int nz, int ny, int nx) {
    const int64_t n = (int64_t)nz * ny * nx * C;
    const int hy = ny / 2, hx = nx / 2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t v = i / C;
        const int xx = (int)(v % nx), yy = (int)((v / nx) % ny), zz = (int)(v / ((int64_t)nx * ny));
        const int64_t p = ((int64_t)(zz / 2) * hy + yy / 2) * hx + xx / 2;
        float g = from_bf16(dy[p * dcs + dco + c]) * 0.125f;
        if (mask && !(from_bf16(mask[v * mcs + mco + c]) > 0.0f)) g = 0.0f;
        const int64_t o = v * xcs + xco + c;
        dxp[o] = to_bf16(accumulate ? from_bf16(dxp[o]) + g : g);
    }
}

// y (nz, ny, nx) = x (parent)
__global__ void up_kernel(const u16* __restrict__ x, int xcs, int xco, u16* __restrict__ y,
                          int ycs, int yco, int C, int nz, int ny, int nx) {
    const int64_t n = (int64_t)nz * ny * nx * C;
    const int hy = ny / 2, hx = nx / 2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t v = i / C;
        const int xx = (int)(v % nx), yy = (int)((v / nx) % ny), zz = (int)(v / ((int64_t)nx * ny));
        const int64_t p = ((int64_t)(zz / 2) * hy + yy / 2) * hx + xx / 2;
        y[v * ycs + yco + c] = x[p * xcs + xco + c];
    }
}

// dx (parent) = sum of dy over the 8 children, masked by mask(parent) > 0
__global__ void up_bwd_kernel(const u16* __restrict__ dy, int dcs, int dco, u16* __restrict__ dx,
                              int xcs, int xco, const u16* __restrict__ mask, int mcs, int mco,
                              int C, int nz, int ny, int nx) {
    const int hz = nz / 2, hy = ny / 2, hx = nx / 2;
    const int64_t n = (int64_t)hz * hy * hx * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int px = (int)(p % hx), py = (int)((p / hx) % hy), pz = (int)(p / ((int64_t)hx * hy));
        float s = 0.0f;
        for (int q = 0; q < 8; ++q) {
            const int64_t v = ((int64_t)(2 * pz + (q >> 2)) * ny + 2 * py + ((q >> 1) & 1)) * nx +
                              2 * px + (q & 1);
            s += from_bf16(dy[v * dcs + dco + c]);
        }
        if (mask && !(from_bf16(mask[p * mcs + mco + c]) > 0.0f)) s = 0.0f;
        dx[p * xcs + xco + c] = to_bf16(s);
    }
}

// The same four operations on 8 channels (16 bytes) per thread with 32-bit
// index arithmetic, each channel's arithmetic in the same order as above (the
// same bits): used when the channel count, strides and offsets are multiples
// of 8, the pointers 16-byte aligned and the element count below 2^31.
__device__ __forceinline__ void unpack8(uint4 v, float f[8]) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[2 * k] = __uint_as_float(w[k] << 16);
        f[2 * k + 1] = __uint_as_float(w[k] & 0xFFFF0000u);
    }
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        w[k] = (uint32_t)to_bf16(f[2 * k]) | ((uint32_t)to_bf16(f[2 * k + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ const uint4& at8(const u16* p, int64_t e) {
    return *reinterpret_cast<const uint4*>(p + e);
}

__global__ __launch_bounds__(256) void pool8_kernel(const u16* __restrict__ x, int xcs, int xco,
                                                    u16* __restrict__ y, int ycs, int yco,
                                                    int C8, int nz, int ny, int nx) {
    const int hy = ny / 2, hx = nx / 2;
    const int n = (nz / 2) * hy * hx * C8;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c8 = i % C8, p = i / C8;
        const int px = p % hx, t = p / hx, py = t % hy, pz = t / hy;
        float s[8], f[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int v = ((2 * pz + (q >> 2)) * ny + 2 * py + ((q >> 1) & 1)) * nx + 2 * px +
                          (q & 1);
            unpack8(at8(x, (int64_t)v * xcs + xco + 8 * c8), f);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += f[k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] *= 0.125f;
        *reinterpret_cast<uint4*>(y + (int64_t)p * ycs + yco + 8 * c8) = pack8(s);
    }
}

__global__ __launch_bounds__(256) void pool_bwd8_kernel(const u16* __restrict__ dy, int dcs,
                                                        int dco, u16* __restrict__ dxp, int xcs,
                                                        int xco, const u16* __restrict__ mask,
                                                        int mcs, int mco, int accumulate, int C8,
                                                        int nz, int ny, int nx) {
    const int n = nz * ny * nx * C8;
    const int hy = ny / 2, hx = nx / 2;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c8 = i % C8, v = i / C8;
        const int xx = v % nx, t = v / nx, yy = t % ny, zz = t / ny;
        const int p = ((zz / 2) * hy + yy / 2) * hx + xx / 2;
        float g[8], m[8], d[8];
        unpack8(at8(dy, (int64_t)p * dcs + dco + 8 * c8), g);
#pragma unroll
        for (int k = 0; k < 8; ++k) g[k] *= 0.125f;
        if (mask) {
            unpack8(at8(mask, (int64_t)v * mcs + mco + 8 * c8), m);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (!(m[k] > 0.0f)) g[k] = 0.0f;
        }
        u16* o = dxp + (int64_t)v * xcs + xco + 8 * c8;
        if (accumulate) {
            unpack8(*reinterpret_cast<const uint4*>(o), d);
#pragma unroll
            for (int k = 0; k < 8; ++k) g[k] = d[k] + g[k];
        }
        *reinterpret_cast<uint4*>(o) = pack8(g);
    }
}

__global__ __launch_bounds__(256) void up8_kernel(const u16* __restrict__ x, int xcs, int xco,
                                                  u16* __restrict__ y, int ycs, int yco, int C8,
                                                  int nz, int ny, int nx) {
    const int n = nz * ny * nx * C8;
    const int hy = ny / 2, hx = nx / 2;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c8 = i % C8, v = i / C8;
        const int xx = v % nx, t = v / nx, yy = t % ny, zz = t / ny;
        const int p = ((zz / 2) * hy + yy / 2) * hx + xx / 2;
        *reinterpret_cast<uint4*>(y + (int64_t)v * ycs + yco + 8 * c8) =
            at8(x, (int64_t)p * xcs + xco + 8 * c8);
    }
}

__global__ __launch_bounds__(256) void up_bwd8_kernel(const u16* __restrict__ dy, int dcs,
                                                      int dco, u16* __restrict__ dx, int xcs,
                                                      int xco, const u16* __restrict__ mask,
                                                      int mcs, int mco, int C8, int nz, int ny,
                                                      int nx) {
    const int hy = ny / 2, hx = nx / 2;
    const int n = (nz / 2) * hy * hx * C8;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c8 = i % C8, p = i / C8;
        const int px = p % hx, t = p / hx, py = t % hy, pz = t / hy;
        float s[8], f[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int v = ((2 * pz + (q >> 2)) * ny + 2 * py + ((q >> 1) & 1)) * nx + 2 * px +
                          (q & 1);
            unpack8(at8(dy, (int64_t)v * dcs + dco + 8 * c8), f);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += f[k];
        }
        if (mask) {
            unpack8(at8(mask, (int64_t)p * mcs + mco + 8 * c8), f);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (!(f[k] > 0.0f)) s[k] = 0.0f;
        }
        *reinterpret_cast<uint4*>(dx + (int64_t)p * xcs + xco + 8 * c8) = pack8(s);
    }
}

// ---- head: input, output, loss -------------------------------------------
// x[v][0..7] = (img[v] * scale, 0, ..., 0)
__global__ void input_kernel(const float* __restrict__ img, int64_t n, float scale,
                             u16* __restrict__ x) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        uint4 o = {0u, 0u, 0u, 0u};
        o.x = to_bf16(img[v] * scale);
        *reinterpret_cast<uint4*>(x + v * 8) = o;
    }
}

// pred[v] = img[v] + out[v][0] / scale (residual); clamp at 0 when clamp != 0
__global__ void output_kernel(const u16* __restrict__ out, int ocs, const float* __restrict__ img,
                              int64_t n, float inv_scale, int clamp, float* __restrict__ pred) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        float p = img[v] + from_bf16(out[v * ocs]) * inv_scale;
        pred[v] = (clamp && p < 0.0f) ? 0.0f : p;
    }
}

// pred[v] = img[v] + (S out)[v] / scale with S the separable in-plane filter
// sum_{j,i} w[j] w[i] out[z][clamp(y + j - r)][clamp(x + i - r)] (taps w of
// length 2r + 1, indices clamped to the volume: the edge value repeats); the
// terms are summed row by row (j outer, i inner) in FP32; clamp at 0 when
// clamp != 0. The warm start's low-passed residual.
struct SmoothTaps {
    float w[2 * CNN_SMOOTH_MAX_R + 1];
};

__global__ __launch_bounds__(256) void output_smooth_kernel(
    const u16* __restrict__ out, int ocs, const float* __restrict__ img, int nz, int ny, int nx,
    float inv_scale, SmoothTaps T, int r, int clamp, float* __restrict__ pred) {
    const int64_t n = (int64_t)nz * ny * nx;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(v % nx);
        const int y = (int)((v / nx) % ny);
        const int64_t plane = v - (int64_t)y * nx - x;
        float acc = 0.0f;
        for (int j = -r; j <= r; ++j) {
            const int yy = min(max(y + j, 0), ny - 1);
            float row = 0.0f;
            for (int i = -r; i <= r; ++i) {
                const int xx = min(max(x + i, 0), nx - 1);
                row = fmaf(T.w[i + r], from_bf16(out[(plane + (int64_t)yy * nx + xx) * ocs]), row);
            }
            acc = fmaf(T.w[j + r], row, acc);
        }
        const float p = img[v] + acc * inv_scale;
        pred[v] = (clamp && p < 0.0f) ? 0.0f : p;
    }
}

// squared-error partial sums (FP64, one per work-group) and the gradient of
// mean((pred - target)^2) with respect to out[v][0] (other channels 0).
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ pred,
                                                  const float* __restrict__ target, int64_t n,
                                                  float gscale, u16* __restrict__ dout, int dcs,
                                                  double* __restrict__ partials) {
    double s = 0.0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const float e = pred[v] - target[v];
        s += (double)e * (double)e;
        if (dout) {
            uint4 z = {0u, 0u, 0u, 0u};
            for (int c = 0; c < dcs; c += 8) *reinterpret_cast<uint4*>(dout + v * dcs + c) = z;
            dout[v * dcs] = to_bf16(gscale * e);
        }
    }
    s = block_sum_256(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ---- parameters ------------------------------------------------------------
// W (fp32 [COUT][27][CIN]) -> forward bf16 [COUT][kpad(CIN)] (k = tap*CIN+ci)
// and data-gradient bf16 [CIN][kpad(COUT)] (k = tap*COUT+co, flipped taps).
__global__ void pack_kernel(const float* __restrict__ wsrc, int cin, int cout,
                            u16* __restrict__ wf, u16* __restrict__ wd) {
    const int kpf = kpad(cin), kpd = kpad(cout);
    const int64_t nf = (int64_t)cout * kpf, nd = (int64_t)cin * kpd;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nf + nd;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < nf) {
            const int co = (int)(i / kpf), k = (int)(i % kpf);
            wf[i] = k < 27 * cin ? to_bf16(wsrc[((int64_t)co * 27 + k / cin) * cin + k % cin]) : 0;
        } else {
            const int64_t e = i - nf;
            const int ci = (int)(e / kpd), k = (int)(e % kpd);
            u16 val = 0;
            if (k < 27 * cout) {
                const int tap = k / cout, co = k % cout;
                val = to_bf16(wsrc[((int64_t)co * 27 + (26 - tap)) * cin + ci]);
            }
            wd[e] = val;
        }
    }
}

// Adam (Kingma & Ba 2015), bias-corrected, FP32 master parameters.
__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

unsigned grid_for(int64_t n, int per = 256) {
    int64_t g = (n + per - 1) / per;
    if (g > 65536) g = 65536;
    return (unsigned)(g < 1 ? 1 : g);
}

// Tile of the LDS convolution: R output rows per wave, XW voxels along x.
template <int CIN, int COUT>
struct ConvTile {
    static constexpr int R = CIN <= 8 ? 4 : (CIN <= 16 ? 2 : 1);
    static constexpr int XW = CIN <= 16 ? 64 : 32;
    static constexpr size_t LDS = (size_t)3 * (CIN / 8) * (4 * R + 2) * (XW + 2) * 16;
};

template <int CIN, int COUT>
hipError_t conv_t(const u16* x, int xcs, int xco, const u16* w, const float* bias, u16* y,
                  int ycs, int yco, const u16* mask, int mcs, int mco, int relu, int nz, int ny,
                  int nx, hipStream_t st) {
    using T = ConvTile<CIN, COUT>;
    const auto kern = conv3d_lds_kernel<CIN, COUT, T::R, T::XW>;
    if (T::LDS > 65536) {
        const hipError_t e = hipFuncSetAttribute((const void*)kern,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)T::LDS);
        if (e != hipSuccess) return e;
    }
    const int tiles_x = (nx + T::XW - 1) / T::XW, tiles_y = (ny + 4 * T::R - 1) / (4 * T::R);
    // about 256 groups (one per CU): z chunks of zc slices, each with a
    // three-slab prologue; longer chunks measured faster at the training size
    const int64_t tiles = (int64_t)tiles_x * tiles_y;
    int64_t nzc = (256 + tiles - 1) / tiles;
    if (nzc > nz) nzc = nz;
    const int zc = (int)((nz + nzc - 1) / nzc);
    nzc = (nz + zc - 1) / zc;
    const int nt = 256 + 64 * conv_loaders(conv_slab(CIN, T::R, T::XW));
    hipLaunchKernelGGL(kern, dim3((unsigned)(tiles * nzc)), dim3(nt), T::LDS, st, x, xcs,
                       xco, w, bias, y, ycs, yco, mask, mcs, mco, relu, nz, ny, nx, tiles_x,
                       tiles_y, zc);
    return hipGetLastError();
}

}  // namespace

// Supported (CIN, COUT) pairs of the forward and data-gradient convolutions.
#define OSTR_CNN_CONV_SHAPES(X) \
    X(8, 16) X(16, 16) X(16, 32) X(32, 32) X(32, 64) X(64, 64) X(96, 32) X(48, 16) \
    X(32, 16) X(64, 32) X(32, 96) X(16, 48)

hipError_t launch_cnn_conv(int cin, int cout, const void* x, int xcs, int xco, const void* w,
                           const float* bias, void* y, int ycs, int yco, const void* mask,
                           int mcs, int mco, int relu, int nz, int ny, int nx, hipStream_t st) {
#define CASE(A, B)                                                                          \
    if (cin == A && cout == B)                                                              \
        return conv_t<A, B>((const u16*)x, xcs, xco, (const u16*)w, bias, (u16*)y, ycs, yco, \
                            (const u16*)mask, mcs, mco, relu, nz, ny, nx, st);
    OSTR_CNN_CONV_SHAPES(CASE)
#undef CASE
    return hipErrorInvalidValue;
}

bool cnn_conv_supported(int cin, int cout) {
#define CASE(A, B) if (cin == A && cout == B) return true;
    OSTR_CNN_CONV_SHAPES(CASE)
#undef CASE
    return false;
}

// The weight gradient (wgrad_lds_kernel): supported (CIN, COUT) pairs (the
// forward shapes), grid and workspace.
#define OSTR_CNN_WGRAD_SHAPES(X) \
    X(8, 16) X(16, 16) X(48, 16) X(16, 32) X(32, 32) X(32, 64) X(64, 64) X(96, 32)
bool cnn_wgrad_supported(int cin, int cout) {
#define CASE(A, B) if (cin == A && cout == B) return true;
    OSTR_CNN_WGRAD_SHAPES(CASE)
#undef CASE
    return false;
}
struct WgradGrid {
    int tiles_x, tiles_y, zc, groups;
};
static WgradGrid wgrad_lds_grid(int cin, int cout, int nz, int ny, int nx) {
    WgradGrid g;
    const int xw = wgrad_xw(cin), ty = wgrad_ty(cin, cout);
    g.tiles_x = (nx + xw - 1) / xw;
    g.tiles_y = (ny + ty - 1) / ty;
    // about 256 spatial groups (one per CU), each marching over zc slices
    const int64_t tiles = (int64_t)g.tiles_x * g.tiles_y;
    int64_t nzc = (256 + tiles - 1) / tiles;
    if (nzc > nz) nzc = nz;
    g.zc = (int)((nz + nzc - 1) / nzc);
    g.groups = (int)(tiles * ((nz + g.zc - 1) / g.zc));
    return g;
}
static size_t wgrad_lds_bytes(int cin, int cout, int nz, int ny, int nx) {
    const WgradGrid g = wgrad_lds_grid(cin, cout, nz, ny, nx);
    return sizeof(float) * (size_t)g.groups * ((size_t)cout * 27 * cin + cout) + 256;
}

size_t cnn_wgrad_work_bytes(int cin, int cout, int nz, int ny, int nx) {
    return wgrad_lds_bytes(cin, cout, nz, ny, nx);
}

template <int CIN, int COUT>
static hipError_t wgrad_lds_t(const void* x, int xcs, int xco, const void* dy, int dcs, int dco,
                              int nz, int ny, int nx, float* dw, float* db, void* work,
                              hipStream_t st) {
    const WgradGrid g = wgrad_lds_grid(CIN, COUT, nz, ny, nx);
    const int64_t n = (int64_t)COUT * 27 * CIN;
    float* part = (float*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    float* bpart = part + (int64_t)g.groups * n;
    const size_t lds = (size_t)(3 * wgrad_slab(CIN, COUT) + 2 * wgrad_dslab(CIN, COUT)) * 16;
    const auto kern = wgrad_lds_kernel<CIN, COUT>;
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(
            (const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const dim3 grid((unsigned)g.groups, (unsigned)(COUT / 16 * wgrad_shares(CIN)));
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, (const u16*)x, xcs, xco, (const u16*)dy,
                       dcs, dco, part, bpart, nz, ny, nx, g.tiles_x, g.tiles_y, g.zc);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(sum_parts_grid(n)), dim3(256), 0, st,
                       (const float*)part, g.groups, n, dw);
    if (db)
        hipLaunchKernelGGL(sum_parts_kernel, dim3(sum_parts_grid(COUT)), dim3(256), 0, st,
                           (const float*)bpart, g.groups, (int64_t)COUT, db);
    return hipGetLastError();
}

hipError_t launch_cnn_wgrad(int cin, int cout, const void* x, int xcs, int xco, const void* dy,
                            int dcs, int dco, int nz, int ny, int nx, float* dw, float* db,
                            void* work, hipStream_t st) {
    if (xcs % 8 || xco % 8 || dcs % 8 || dco % 8 || (uintptr_t)x % 16 || (uintptr_t)dy % 16)
        return hipErrorInvalidValue;
#define CASE(A, B)      \
    if (cin == A && cout == B) \
        return wgrad_lds_t<A, B>(x, xcs, xco, dy, dcs, dco, nz, ny, nx, dw, db, work, st);
    OSTR_CNN_WGRAD_SHAPES(CASE)
#undef CASE
    return hipErrorInvalidValue;
}

// 16-byte access allowed: channel count, strides, offsets and pointers, and
// the 32-bit index arithmetic of the vector kernels (nvox (C / 8) < 2^31 for
// the nvox voxels of the finer level)
static bool vec8(int C, int64_t nvox, std::initializer_list<int> cs_co,
                 std::initializer_list<const void*> ptrs) {
    if (C % 8 || nvox * (C / 8) >= ((int64_t)1 << 31)) return false;
    for (int v : cs_co)
        if (v % 8) return false;
    for (const void* p : ptrs)
        if ((uintptr_t)p % 16) return false;
    return true;
}

hipError_t launch_cnn_pool(const void* x, int xcs, int xco, void* y, int ycs, int yco, int C,
                           int nz, int ny, int nx, hipStream_t st) {
    const int64_t n = (int64_t)(nz / 2) * (ny / 2) * (nx / 2) * C;
    if (vec8(C, (int64_t)nz * ny * nx, {xcs, xco, ycs, yco}, {x, y})) {
        hipLaunchKernelGGL(pool8_kernel, dim3(grid_for(n / 8)), dim3(256), 0, st,
                           (const u16*)x, xcs, xco, (u16*)y, ycs, yco, C / 8, nz, ny, nx);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(pool_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const u16*)x, xcs, xco,
                       (u16*)y, ycs, yco, C, nz, ny, nx);
    return hipGetLastError();
}

hipError_t launch_cnn_pool_bwd(const void* dy, int dcs, int dco, void* dx, int xcs, int xco,
                               const void* mask, int mcs, int mco, int accumulate, int C, int nz,
                               int ny, int nx, hipStream_t st) {
    const int64_t n = (int64_t)nz * ny * nx * C;
    if (vec8(C, (int64_t)nz * ny * nx, {dcs, dco, xcs, xco, mask ? mcs : 0, mask ? mco : 0},
             {dy, dx, mask})) {
        hipLaunchKernelGGL(pool_bwd8_kernel, dim3(grid_for(n / 8)), dim3(256), 0, st,
                           (const u16*)dy, dcs, dco, (u16*)dx, xcs, xco, (const u16*)mask, mcs,
                           mco, accumulate, C / 8, nz, ny, nx);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(pool_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const u16*)dy, dcs,
                       dco, (u16*)dx, xcs, xco, (const u16*)mask, mcs, mco, accumulate, C, nz, ny,
                       nx);
    return hipGetLastError();
}

hipError_t launch_cnn_up(const void* x, int xcs, int xco, void* y, int ycs, int yco, int C,
                         int nz, int ny, int nx, hipStream_t st) {
    const int64_t n = (int64_t)nz * ny * nx * C;
    if (vec8(C, (int64_t)nz * ny * nx, {xcs, xco, ycs, yco}, {x, y})) {
        hipLaunchKernelGGL(up8_kernel, dim3(grid_for(n / 8)), dim3(256), 0, st, (const u16*)x,
                           xcs, xco, (u16*)y, ycs, yco, C / 8, nz, ny, nx);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(up_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const u16*)x, xcs, xco,
                       (u16*)y, ycs, yco, C, nz, ny, nx);
    return hipGetLastError();
}

hipError_t launch_cnn_up_bwd(const void* dy, int dcs, int dco, void* dx, int xcs, int xco,
                             const void* mask, int mcs, int mco, int C, int nz, int ny, int nx,
                             hipStream_t st) {
    const int64_t n = (int64_t)(nz / 2) * (ny / 2) * (nx / 2) * C;
    if (vec8(C, (int64_t)nz * ny * nx, {dcs, dco, xcs, xco, mask ? mcs : 0, mask ? mco : 0},
             {dy, dx, mask})) {
        hipLaunchKernelGGL(up_bwd8_kernel, dim3(grid_for(n / 8)), dim3(256), 0, st,
                           (const u16*)dy, dcs, dco, (u16*)dx, xcs, xco, (const u16*)mask, mcs,
                           mco, C / 8, nz, ny, nx);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(up_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const u16*)dy, dcs,
                       dco, (u16*)dx, xcs, xco, (const u16*)mask, mcs, mco, C, nz, ny, nx);
    return hipGetLastError();
}

hipError_t launch_cnn_input(const float* img, int64_t n, float scale, void* x, hipStream_t st) {
    hipLaunchKernelGGL(input_kernel, dim3(grid_for(n)), dim3(256), 0, st, img, n, scale, (u16*)x);
    return hipGetLastError();
}

hipError_t launch_cnn_output(const void* out, int ocs, const float* img, int64_t n,
                             float inv_scale, int clamp, float* pred, hipStream_t st) {
    hipLaunchKernelGGL(output_kernel, dim3(grid_for(n)), dim3(256), 0, st, (const u16*)out, ocs,
                       img, n, inv_scale, clamp, pred);
    return hipGetLastError();
}

hipError_t launch_cnn_output_smooth(const void* out, int ocs, const float* img, int nz, int ny,
                                    int nx, float inv_scale, const float* taps, int r, int clamp,
                                    float* pred, hipStream_t st) {
    if (r < 0 || r > CNN_SMOOTH_MAX_R) return hipErrorInvalidValue;
    SmoothTaps T = {};
    for (int k = 0; k <= 2 * r; ++k) T.w[k] = taps[k];
    const int64_t n = (int64_t)nz * ny * nx;
    hipLaunchKernelGGL(output_smooth_kernel, dim3(grid_for(n)), dim3(256), 0, st,
                       (const u16*)out, ocs, img, nz, ny, nx, inv_scale, T, r, clamp, pred);
    return hipGetLastError();
}

int64_t cnn_mse_partials(int64_t n) { return grid_for(n, 256 * 8); }

hipError_t launch_cnn_mse(const float* pred, const float* target, int64_t n, float gscale,
                          void* dout, int dcs, double* partials, double* result, hipStream_t st) {
    const unsigned g = (unsigned)cnn_mse_partials(n);
    hipLaunchKernelGGL(mse_kernel, dim3(g), dim3(256), 0, st, pred, target, n, gscale,
                       (u16*)dout, dcs, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_partials(partials, g, result, st);
}

hipError_t launch_cnn_pack(const float* w, int cin, int cout, void* wf, void* wd,
                           hipStream_t st) {
    const int64_t n = (int64_t)cout * kpad(cin) + (int64_t)cin * kpad(cout);
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for(n)), dim3(256), 0, st, w, cin, cout, (u16*)wf,
                       (u16*)wd);
    return hipGetLastError();
}

int cnn_kpad(int c) { return kpad(c); }

hipError_t launch_cnn_adam(float* p, float* m, float* v, const float* g, int64_t n, float lr,
                           float b1, float b2, float eps, int step, hipStream_t st) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, m, v, g, n, lr, b1,
                       b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace ostr
