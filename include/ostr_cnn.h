/*
 * ostr_cnn.h - C ABI of the warm-start network kernels of libostr_hip.so.
 *
 * No counterpart in the reference's OSTR path (its learned model is a Keras
 * network in recon1.ipynb, run outside the reconstruction). These calls are
 * the building blocks of the learned warm start: a 3-D CNN that maps the
 * one-step backprojected image to an estimate of the converged OSTR image,
 * trained and run on the GPU with BF16 operands on MFMA and FP32
 * accumulation.
 *
 * Tensors are channels-last BF16 views: element (voxel v, channel c) at
 * p[v * cs + co + c], v = (z * ny + y) * nx + x, with cs the channel stride and
 * co the channel offset of the view. Conventions as in ostr_hip.h: device
 * pointers owned by the caller, asynchronous on `stream` unless stated, 0 on
 * success or a negative OSTR_E_* code; the message is ostr_last_error(NULL).
 * Nothing here allocates device memory: workspaces are passed in.
 */
#ifndef OSTR_CNN_H
#define OSTR_CNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Padded K of a packed weight matrix: 27 * c rounded up to 32. */
int32_t ostr_cnn_kpad(int32_t c);

/* Packs FP32 master weights w [cout][27][cin] (tap = (dz+1)*9+(dy+1)*3+(dx+1))
 * into the forward matrix wf (bf16 [cout][kpad(cin)], k = tap*cin + ci) and
 * the data-gradient matrix wd (bf16 [cin][kpad(cout)], flipped taps). */
int ostr_cnn_pack(const float* w, int32_t cin, int32_t cout, void* wf, void* wd, void* stream);

/* y = act(conv3x3x3(x; wf) + bias), zero padding, stride 1. relu != 0 clamps
 * at 0; mask != NULL zeroes outputs where mask <= 0 (the ReLU gradient). The
 * data gradient is this call on the output gradient with wd, cin/cout
 * swapped, bias NULL and the layer input as mask. Channel strides/offsets of
 * x multiples of 8, of y and mask multiples of 4; x and bias 16-byte
 * aligned, y and mask 8-byte aligned (OSTR_E_ARG otherwise). (cin, cout) must be one of
 * the compiled shapes (8,16) (16,16) (16,32) (32,32) (32,64) (64,64) (96,32)
 * (48,16) (32,16) (64,32) (32,96) (16,48). */
int ostr_cnn_conv3d(int32_t cin, int32_t cout, const void* x, int32_t xcs, int32_t xco,
                    const void* wf, const float* bias, void* y, int32_t ycs, int32_t yco,
                    const void* mask, int32_t mcs, int32_t mco, int32_t relu, int32_t nz,
                    int32_t ny, int32_t nx, void* stream);

/* Workspace (bytes) of ostr_cnn_wgrad for these sizes (one partial sum per
 * work-group of a fixed grid); nsplit is accepted and not used. 0 for an
 * unsupported shape. */
size_t ostr_cnn_wgrad_workspace(int32_t cin, int32_t cout, int32_t nz, int32_t ny, int32_t nx,
                                int32_t nsplit);

/* dw [cout][27][cin] = gradient of the convolution weights for input x and
 * output gradient dy (FP32, written); db [cout] likewise (may be NULL).
 * Deterministic: one partial sum per work-group of a fixed grid, added in a
 * fixed order (nsplit, 1..1024, is accepted and not used). Channel strides and
 * offsets multiples of 8, x and dy 16-byte aligned (OSTR_E_ARG otherwise).
 * (cin, cout) one of (8,16) (16,16) (48,16) (16,32) (32,32) (32,64) (64,64)
 * (96,32). */
int ostr_cnn_wgrad(int32_t cin, int32_t cout, const void* x, int32_t xcs, int32_t xco,
                   const void* dy, int32_t dcs, int32_t dco, int32_t nz, int32_t ny, int32_t nx,
                   float* dw, float* db, void* work, size_t work_bytes, int32_t nsplit,
                   void* stream);

/* 2x average pooling (even sizes): y (nz/2, ny/2, nx/2, C) = mean of the 8
 * children of x (nz, ny, nx, C). */
int ostr_cnn_pool2(const void* x, int32_t xcs, int32_t xco, void* y, int32_t ycs, int32_t yco,
                   int32_t C, int32_t nz, int32_t ny, int32_t nx, void* stream);
/* Its gradient: dx(child) = dy(parent)/8 (added to dx when accumulate),
 * zero where mask(child) <= 0 (mask may be NULL). */
int ostr_cnn_pool2_bwd(const void* dy, int32_t dcs, int32_t dco, void* dx, int32_t xcs,
                       int32_t xco, const void* mask, int32_t mcs, int32_t mco,
                       int32_t accumulate, int32_t C, int32_t nz, int32_t ny, int32_t nx,
                       void* stream);
/* 2x nearest upsampling: y (nz, ny, nx, C) = x(parent) (sizes of y given). */
int ostr_cnn_up2(const void* x, int32_t xcs, int32_t xco, void* y, int32_t ycs, int32_t yco,
                 int32_t C, int32_t nz, int32_t ny, int32_t nx, void* stream);
/* Its gradient: dx(parent) = sum of dy over the 8 children, zero where
 * mask(parent) <= 0 (mask may be NULL). Sizes are those of dy. */
int ostr_cnn_up2_bwd(const void* dy, int32_t dcs, int32_t dco, void* dx, int32_t xcs,
                     int32_t xco, const void* mask, int32_t mcs, int32_t mco, int32_t C,
                     int32_t nz, int32_t ny, int32_t nx, void* stream);

/* Network input: x[v][0..7] = (img[v] * scale, 0, ..., 0) (bf16, stride 8). */
int ostr_cnn_input(const float* img, int64_t n, float scale, void* x, void* stream);
/* Network output: pred[v] = img[v] + out[v * ocs] / scale (the residual
 * estimate), clamped at 0 when clamp != 0. */
int ostr_cnn_output(const void* out, int32_t ocs, const float* img, int64_t n, float scale,
                    int32_t clamp, float* pred, void* stream);
/* ostr_cnn_output with the residual low-passed in-plane first (the warm
 * start's correction without its finest detail):
 * pred[v] = img[v] + (sum_{j,i} taps[j] taps[i] out[z][y'][x'] * ocs) / scale,
 * y' = clamp(y + j - radius, 0, ny - 1), x' likewise (the edge repeats), taps
 * of length 2 radius + 1, radius <= 4; rows summed in order (j outer, i inner)
 * in FP32; clamped at 0 when clamp != 0. img, pred: [nz][ny][nx] f32. */
int ostr_cnn_output_smooth(const void* out, int32_t ocs, const float* img, int32_t nz,
                           int32_t ny, int32_t nx, float scale, const float* taps,
                           int32_t radius, int32_t clamp, float* pred, void* stream);
/* Size (doubles) of the ostr_cnn_mse workspace for n voxels. */
int64_t ostr_cnn_mse_workspace(int64_t n);
/* Synchronous. *loss = sum (pred - target)^2 (FP64, fixed order); when dout is
 * not NULL, dout[v][0] = gscale * (pred - target) and the other dcs - 1
 * channels 0 (dcs a multiple of 8). */
int ostr_cnn_mse(const float* pred, const float* target, int64_t n, float gscale, void* dout,
                 int32_t dcs, double* work, int64_t work_elems, double* loss, void* stream);

/* One Adam step (bias-corrected; step >= 1) on FP32 parameters p with
 * moments m, v and gradient g. */
int ostr_cnn_adam(float* p, float* m, float* v, const float* g, int64_t n, float lr, float b1,
                  float b2, float eps, int32_t step, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* OSTR_CNN_H */
