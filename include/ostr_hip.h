/*
 * ostr_hip.h - C ABI of libostr_hip.so, the gfx950 HIP implementation of the
 * OSTR (ordered-subsets transmission) DBT reconstruction path.
 *
 * Every entry point takes plain pointers and sizes. Image and sinogram pointers
 * are DEVICE pointers owned by the caller (for example torch.Tensor.data_ptr()
 * on a ROCm build); the library never frees them. `stream` is a hipStream_t
 * passed as void* (0 = the null stream). All calls are asynchronous on that
 * stream unless stated otherwise. Return value: 0 on success, a negative
 * OSTR_E* code on failure; ostr_last_error() gives the message. A handle is not
 * thread-safe; separate handles are independent (no process-global geometry).
 *
 * Layouts (same as the reference):
 *   volume   [nz][ny][nx], x fastest            (Tomo_OSTR.cu:1604)
 *   sinogram [view slot][nbx][nby], nby fastest (Tomo_OSTR.cu:1365)
 * A "slot" is a position in the view ordering table `index` (slot a uses the
 * angle angles[index[a]]), exactly like d_index in the reference launchers.
 */
#ifndef OSTR_HIP_H
#define OSTR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry; mirrors the reference's __constant__ globals (Tomo_OSTR.cu:81-115,
 * datafile keys Tomo_OSTR.cu:2578-2650). Lengths in mm. */
typedef struct ostr_geom {
    int32_t nx, ny, nz;        /* IMGSIZx/y/z                                   */
    float   dx, dy, dz;        /* Vsize_x/y/z                                   */
    float   x_p0, y_p0, z_p0;  /* corner of voxel (0,0,0)                       */
    int32_t nbx, nby;          /* BINSx, BINSy                                  */
    float   dbx, dby;          /* Bsize_x, Bsize_y                              */
    float   x_d0, y_d0;        /* corner of detector bin (0,0)                  */
    float   detector_r;        /* detectorR: flat detector plane z = -detectorR */
    float   source_r;          /* sourceR: source at (R sin t, sourceY, R cos t) */
    float   source_y;          /* sourceY                                       */
} ostr_geom;

typedef struct ostr_handle ostr_handle;

/* error codes */
#define OSTR_OK            0
#define OSTR_E_ARG        -1  /* invalid argument / bad configuration          */
#define OSTR_E_HIP        -2  /* HIP runtime error                             */
#define OSTR_E_RANGE      -3  /* view range or size out of bounds              */
#define OSTR_E_UNSUPPORTED -4 /* dtype / mode not supported                    */

/* storage dtypes for images and sinograms */
#define OSTR_F32  0
#define OSTR_BF16 1
#define OSTR_F16  2

/* gamma (denominator row term) modes */
#define OSTR_GAMMA_RAY_LENGTH 0 /* gamma = |S - D| for rays that meet the box (Tomo_OSTR.cu:1172) */
#define OSTR_GAMMA_ROW_SUM    1 /* gamma = sum_j a_ij (ostr.ipynb normprj, raw line 815)         */

/* prior potentials */
#define OSTR_PRIOR_Q2        0 /* quadratic, G_Q2_prior (Tomo_OSTR.cu:118-184)                    */
#define OSTR_PRIOR_LANGE     1 /* Lange/"fair", G_Fessler_prior (Tomo_OSTR.cu:187-255)            */
#define OSTR_PRIOR_HYPERBOLA 2 /* hyperbola potential of G_Huber_prior (Tomo_OSTR.cu:258-325)
                                  with the curvature and z weighting of G_Fessler_prior          */

/* ---------------------------------------------------------------- handle -- */

/* Create a handle on the current HIP device. `angles_rad` holds n_views
 * float32 angles in radians (the reference's angle file, Tomo_OSTR_gold.cpp:
 * 1474-1488); `index` the initial slot->view table (NULL = identity). Per-view
 * sin/cos are evaluated once on the host in double and rounded to float.
 * Replaces the geometry upload of main() (Tomo_OSTR.cu:2653-2690). Limits
 * (OSTR_E_ARG otherwise): volume dimensions < 32768, nx ny nz + 2 nx ny + 2 nx + 2
 * <= 2^30 (32-bit byte offsets into the volume), nbx nby <= 2^30. The
 * backprojecting calls (SBP, exact transpose, OS-SART and fused updates, the d
 * precompute, BP image) also need min(views, 64) nbx nby elem_bytes < 2^32 - 16
 * per call (32-bit byte offsets into the sinogram; OSTR_E_RANGE otherwise). On
 * a handle used under stream capture, work buffers that later calls outgrow are
 * kept until ostr_destroy (a recorded graph may still write to them), and a
 * call that would have to grow one while capturing fails with OSTR_E_ARG. */
int ostr_create(const ostr_geom* geom, const float* angles_rad, const int32_t* index,
                int32_t n_views, ostr_handle** out);
void ostr_destroy(ostr_handle* h);
/* Message of the last failure on h (h == NULL: last failure of ostr_create or
 * of a call given a NULL handle, per thread). Never NULL. */
const char* ostr_last_error(const ostr_handle* h);
/* Replace the slot->view table (OS regroup, Tomo_OSTR_gold.cpp:628-642).
 * Synchronous. */
int ostr_set_index(ostr_handle* h, const int32_t* index);
int ostr_get_geom(const ostr_handle* h, ostr_geom* out);
int32_t ostr_n_views(const ostr_handle* h);

/* ------------------------------------------- reference launcher mirrors -- */

/* Forward projection of slots [a0,a1): prj[(a-a0)][i][j] = cos_t * sum_j a_ij x_j
 * (cos_weight != 0) or the plain Siddon line integral (cos_weight == 0).
 * Replaces fprojectCB_1R_GPU_OSTR_cos (Tomo_OSTR.cu:2334-2359; kernel
 * :1313-1376, walker :612-890). Difference: rays that miss the volume are
 * written as 0 instead of left untouched (the reference driver zeroes the
 * buffer before every call, Tomo_OSTR_gold.cpp:843, so the result is equal). */
int ostr_fproj(ostr_handle* h, const void* img, int32_t img_dtype, void* prj,
               int32_t prj_dtype, int32_t a0, int32_t a1, int32_t cos_weight, void* stream);

/* Denominator row term gamma for slots [a0,a1) into gamma[(a-a0)][i][j] (f32).
 * Replaces fprojectCB_1R_GPU_OSTR_normprj (Tomo_OSTR.cu:2361-2393, walker
 * :892-1173). mode = OSTR_GAMMA_*. Misses are written as 0. */
int ostr_gamma(ostr_handle* h, float* gamma, int32_t a0, int32_t a1, int32_t mode,
               void* stream);

/* Voxel-driven bilinear backprojection of slots [a0,a1), ACCUMULATED into img
 * (f32). Replaces bprojectCB_GPU_SBP (Tomo_OSTR.cu:2435-2457, kernel
 * :1553-1734). prj is [(a-a0)][nbx][nby] in prj_dtype. */
int ostr_bproj_sbp(ostr_handle* h, float* img_accum, const void* prj, int32_t prj_dtype,
                   int32_t a0, int32_t a1, void* stream);

/* Exact transpose of ostr_fproj, ACCUMULATED into img (f32): img_j +=
 * sum_{a in [a0,a1)} c_a sum_i a_ij prj[a-a0][i] with c_a = cos(t_a) (cos_weight
 * != 0) or 1, a_ij the length of ray i inside voxel j. No reference
 * counterpart (the reference backprojects with the unmatched SBP): it makes
 * <A x, y> = <x, A^T y>. Voxel-driven (ray/box clipping of the rays through
 * each voxel's shadow), deterministic, no atomics. prj in the reference layout
 * [slot][nbx][nby] (rows_layout = 0) or detector-row-major [slot][nby][nbx]. */
int ostr_bproj_matched(ostr_handle* h, float* img_accum, const void* prj, int32_t prj_dtype,
                       int32_t a0, int32_t a1, int32_t cos_weight, int32_t rows_layout,
                       void* stream);

/* One-step backprojected image of ReconDBT/BP.m:54-82, written to img (f32),
 * with G = cos(t) A: per view l = G 1, q = p / l where l > |dz|/10 (else 0),
 * img += (G^T q) / (G^T 1) (0 where G^T 1 = 0), then img / n_views. The
 * reference's OSTR starts from zeros; this is the start the warm-start path
 * refines. p_or_y: line integrals p (counts = 0) or counts y with p = log(b /
 * max(y, 1)) (counts != 0), natural view order, reference layout
 * [V][nbx][nby]; b only with counts. Uses the handle's work buffer. */
int ostr_bp_init(ostr_handle* h, float* img, const float* p_or_y, const float* b,
                 int32_t counts, void* stream);

/* Prior gradient and curvature ACCUMULATED into rd and rdd (f32):
 * rd_j += sum_k w_jk psi'(x_j - x_k), rdd_j += sum_k w_jk omega(x_j - x_k) over
 * the 26 neighbours, w = 1/sqrt(di^2+dj^2+(z_xy_ratio dk)^2), out-of-volume
 * neighbours replaced by the voxel itself. Replaces prior_GPU_OSTR
 * (Tomo_OSTR.cu:2216-2245) and prior_GPU_OSTR_Q2 (:2189-2212). delta must be
 * non-zero for LANGE and HYPERBOLA (the reference exits, :2232-2236). */
int ostr_prior(ostr_handle* h, float* rdd, float* rd, const void* img, int32_t img_dtype,
               int32_t potential, float delta, float z_xy_ratio, void* stream);

/* -------------------------------------- host CPU ops, now on the device -- */

/* yry[i] = (y-r)^2/y with y-r <= 0 -> y, and 0 where y == 0
 * (compute_yry, Tomo_OSTR_gold.cpp:183-205). r may be NULL (r = 0). */
int ostr_yry(ostr_handle* h, float* yry, const float* y, const float* r, int64_t n,
             void* stream);
/* out[i] = a[i] * b[i] (compute_gamma_yry, Tomo_OSTR_gold.cpp:208-218). */
int ostr_mul(ostr_handle* h, float* out, const float* a, const float* b, int64_t n,
             void* stream);
/* hsub[i] = (y/(b e^{-l} + r) - 1) b e^{-l} (compute_h, Tomo_OSTR_gold.cpp:
 * 245-260). r may be NULL. y/b/r/h in f32. */
int ostr_compute_h(ostr_handle* h, float* hsub, const float* y, const float* b,
                   const float* l, const float* r, int64_t n, void* stream);
/* est = max(0, est - (capl_scale*capL + beta*rd)/(d + 2 beta rdd)) over the
 * volume (capL scaling Tomo_OSTR_gold.cpp:767-771 + update_est :265-295).
 * guard != 0: voxels with a zero denominator are left unchanged (the reference
 * produces 0/0 = NaN there, guard == 0 reproduces that). */
int ostr_update_est(ostr_handle* h, float* est, const float* capL, const float* rd,
                    const float* d, const float* rdd, float beta, float capl_scale,
                    int32_t guard, void* stream);
/* Subset-major regroup of a [n_views][nbx][nby] f32 array: view block m*Vs+j
 * of dst = view block j*M+m of src (regroup_prj, Tomo_OSTR_gold.cpp:220-242).
 * to_rows != 0 also transposes every view to the detector-row-major layout
 * [nby][nbx] used by the fused path. dst != src. */
int ostr_regroup(ostr_handle* h, float* dst, const float* src, int32_t n_subsets,
                 int32_t to_rows, void* stream);
/* [n_slots][nbx][nby] -> detector-row-major [n_slots][nby][nbx] (no
 * reference counterpart: layout conversion for the fused path). */
int ostr_to_rows(ostr_handle* h, void* dst, const void* src, int32_t n_slots, int32_t dtype,
                 void* stream);

/* ---------------------------------------------------------- fused path -- */
/* The fused path keeps subset sinograms (y, b, r, h) in the detector-row-major
 * layout [slot][nby][nbx] (x fastest), so that the 64 lanes of a wave, placed
 * on consecutive voxels along x, read consecutive bins. */

/* Same as ostr_bproj_sbp with prj already detector-row-major. */
int ostr_bproj_sbp_rows(ostr_handle* h, float* img_accum, const void* prj, int32_t prj_dtype,
                        int32_t a0, int32_t a1, void* stream);

/* d = SBP_all(gamma * yry) over all n_views slots, written (not accumulated)
 * to d (f32). Fuses ostr_gamma + ostr_yry + ostr_mul + ostr_bproj_sbp of the
 * reference precompute (Tomo_OSTR_gold.cpp:409-560). y, r in natural view
 * order [n_views][nbx][nby] f32 (r may be NULL). Uses a device work buffer of
 * n_views*nbx*nby floats allocated on the first call and kept by the handle. */
int ostr_precompute_d(ostr_handle* h, float* d, const float* y, const float* r,
                      int32_t gamma_mode, void* stream);

/* ostr_precompute_d with the sum over views in FP64 (the denominator
 * reduction in double): each view's SBP term in FP32 as ostr_precompute_d
 * computes it, the terms summed in view order in double, d rounded to f32 once.
 * Work buffer: nx ny nz (8 + 4) + n_views nbx nby 4 bytes. */
int ostr_precompute_d_f64(ostr_handle* h, float* d, const float* y, const float* r,
                          int32_t gamma_mode, void* stream);

/* Forward projection of slots [a0,a1) with the h epilogue: hsub = h(y, b, r,
 * cos_t * A x). y, b, r are the subset's slices, detector-row-major
 * [(a-a0)][nby][nbx] (r may be NULL); hsub (same layout) in prj_dtype; y/b/r
 * in f32. Fuses ostr_fproj + ostr_compute_h. */
int ostr_fproj_h(ostr_handle* h, const void* img, int32_t img_dtype, const float* y,
                 const float* b, const float* r, void* hsub, int32_t prj_dtype,
                 int32_t a0, int32_t a1, void* stream);

/* x_out = max(0, x_in - (M*SBP(hsub) + beta*RD(x_in)) / (d + 2 beta RDD(x_in)))
 * in one pass (hsub detector-row-major, as written by ostr_fproj_h): SBP
 * gather of the subset's h, 27-neighbour prior on x_in
 * (Jacobi: the prior sees the pre-update image, Tomo_OSTR_gold.cpp:775 before
 * :814) and the update. x_in and x_out must not alias. */
int ostr_update_fused(ostr_handle* h, const void* x_in, void* x_out, int32_t img_dtype,
                      const float* d, const void* hsub, int32_t prj_dtype, int32_t a0,
                      int32_t a1, float beta, float capl_scale, int32_t potential,
                      float delta, float z_xy_ratio, int32_t guard, void* stream);

/* ostr_update_fused with the exact transpose in place of SBP (no reference
 * counterpart): x_out = x_in - (capl_scale (0 + G^T h) + beta RD) /
 * (d + 2 beta RDD), G^T the transpose of ostr_fproj (cos-weighted, as
 * ostr_bproj_matched), the prior of x_in, guard as ostr_update_est. Image (x_in,
 * x_out) and h each stored as OSTR_F32 / OSTR_BF16 / OSTR_F16 (arithmetic in
 * FP32, d FP32); h detector-row-major; at most 64 views; x_in and x_out must
 * differ. With FP32 storage it equals ostr_bproj_matched + ostr_prior +
 * ostr_update_est bit for bit. */
int ostr_update_fused_matched(ostr_handle* h, const void* x_in, void* x_out, int32_t img_dtype,
                              const float* d, const void* hsub, int32_t prj_dtype, int32_t a0,
                              int32_t a1, float beta, float capl_scale, int32_t potential,
                              float delta, float z_xy_ratio, int32_t guard, void* stream);

/* ------------------------------------------------------- measurement -- */

/* Number of (ray, voxel) pairs with a_ij > 0 for slots [a0,a1): the exact nnz
 * of the subset's system matrix. Synchronous; result written to *nnz_out. */
int ostr_count_nnz(ostr_handle* h, int32_t a0, int32_t a1, int64_t* nnz_out, void* stream);

/* FP64 objective of img over ALL views (no reference counterpart; the
 * reference has no objective). out[0] = sum_i (yhat_i + r_i) - y_i log(yhat_i
 * + r_i) with yhat = b exp(-cos_t A x) (rays summed in f32, terms in f64;
 * y_i = 0 contributes yhat_i + r_i); out[1] = R(x) = 1/2 sum_j sum_{k in
 * N(j), inside} w_jk psi(x_j - x_k) in f64, psi: Q2 t^2/2, Lange
 * d^2(|t/d| - log(1+|t/d|)), hyperbola d^2(sqrt(1+(t/d)^2) - 1). The
 * penalized objective is out[0] + beta*out[1]. y, b, r: natural view order,
 * reference layout [V][nbx][nby] (r may be NULL). out: host pointer to 2
 * doubles. Synchronous; deterministic (fixed-order reductions, no atomics). */
int ostr_objective(ostr_handle* h, const void* img, int32_t img_dtype, const float* y,
                   const float* b, const float* r, int32_t potential, double delta,
                   double z_xy_ratio, double* out, void* stream);

/* ostr_objective without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc). Its FP64 partial sums go to
 * a static per-device scratch array of 2^20 doubles (8 MiB, resolved by
 * ostr_create, held until the process ends); cases needing more than
 * (nbx/64 rounded up)(nby/4 rounded up) n_views + (nx/64 rounded up)(ny/4
 * rounded up) nz + 2 doubles return OSTR_E_RANGE. Same arguments, results and
 * kernels as ostr_objective, bit for bit. All calls on one device share the
 * array: calls that overlap on one device (other streams, other threads)
 * race. */
int ostr_objective_noalloc(ostr_handle* h, const void* img, int32_t img_dtype, const float* y,
                           const float* b, const float* r, int32_t potential, double delta,
                           double z_xy_ratio, double* out, void* stream);

/* out[0] = sum (a_i - b_i)^2, out[1] = sum b_i^2 over n elements, in f64
 * (convergence test and error against a reference image; no reference
 * counterpart). Synchronous; deterministic. */
int ostr_diff_norms(ostr_handle* h, const void* a, const void* b, int32_t dtype, int64_t n,
                    double* out, void* stream);

/* ---- OS-SART (same projectors; the reference has the launchers but no driver)
 * All f32, reference sinogram layout [slot][nbx][nby]. */

/* fprojectCB_1R_GPU_SART / _cos (Tomo_OSTR.cu:2278, :2306): prj = (cos t) A x
 * for cos_weight != 0, else A x; norm = A 1 (row sums, never cos-weighted);
 * both 0 for rays that miss the volume. */
int ostr_fproj_sart(ostr_handle* h, const float* img, float* prj, float* norm, int32_t a0,
                    int32_t a1, int32_t cos_weight, void* stream);

/* SART_prj_diff (Tomo_OSTR.cu:2537): diff = (p - est) / norm where norm != 0,
 * else 0, over n bins. */
int ostr_sart_prj_diff(ostr_handle* h, float* diff, const float* p, const float* est,
                       const float* norm, int64_t n, void* stream);

/* bprojectCB_4B_GPU_SART (Tomo_OSTR.cu:2460): per voxel, u = SBP(diff) /
 * SBP-sensitivity over slots [a0,a1) (0 where the sensitivity is 0), then in
 * place img = max(0, img + lambda u). At most 64 views per call. */
int ostr_bproj_sart(ostr_handle* h, float* img, const float* diff, int32_t a0, int32_t a1,
                    float lambda, void* stream);

/* bprojectCB_4B_GPU_SART_R (Tomo_OSTR.cu:2497): as ostr_bproj_sart with
 * img = max(0, img + lambda (u + beta prior / sensitivity)) (the reference's
 * sign). */
int ostr_bproj_sart_r(ostr_handle* h, float* img, const float* diff, const float* prior,
                      int32_t a0, int32_t a1, float lambda, float beta, void* stream);

/* prior_GPU_SART (Tomo_OSTR.cu:2249): prior += sum over the 27 neighbours of
 * (1/distance) d / sqrt(1 + (d/delta)^2), d = img[v] - img[neighbour], distance
 * in voxel-index units (G_Huber_prior :258-325), evaluated in FP32 (the
 * reference forms each term in double). delta == 0 and prior == img are
 * errors. */
int ostr_prior_sart(ostr_handle* h, float* prior, const float* img, float delta, void* stream);

/* FDK-style filtering for a DBT FBP (ReconDBT/fbp_dbt.m's role; its IRT
 * feldkamp is not available, so this is the textbook form on the flat
 * detector): over slots [a0,a1) in the reference layout [slot][nbx][nby],
 * q[a][i][j] = sum_k taps[k] cw(i', j) p[a][i'][j], i' = i + k - (ntaps-1)/2,
 * cw = D / |S - E| the cosine of the ray (S the slot's source, E the bin centre,
 * D the source height above the detector plane), samples off the detector 0.
 * ntaps odd; taps in device memory (taps with |i' - i| > nbx - 1 never meet
 * data). The sum is evaluated as an FP32 FFT convolution in LDS over
 * L = 2^ceil(log2(2 nbx - 1)) points (exact in exact arithmetic; deterministic),
 * so nbx is limited by the LDS of one work-group: nbx <= 4096 on MI355X
 * (160 KiB), OSTR_E_RANGE beyond. The handle's work buffer grows to 12 L bytes
 * (the taps' spectrum and the twiddles). */
int ostr_fbp_filter(ostr_handle* h, float* q, const float* p, const float* taps, int32_t ntaps,
                    int32_t a0, int32_t a1, void* stream);

/* ---- ReconDBT SART and ML (ReconDBT/SART_dbt.m, ML_dbt.m) on this build's
 * projector pair: G = ostr_fproj without cos, G^T = ostr_bproj_matched without
 * cos. The MATLAB package applies IRT system matrices; these are its per-view
 * elementwise steps. All f32, n elements, any matching layout. */

/* p = log(b / max(y, 1)) over n elements: the line integrals of counts y
 * under blank b, as ReconDBT/BP.m:54-60 forms them (no reference launcher). */
int ostr_line_integrals(ostr_handle* h, float* p, const float* y, const float* b, int64_t n,
                        void* stream);
/* SART_dbt.m:55-66: y = (p - est) / l where l > thresh (5 dz in SART_dbt.m),
 * else y is left unchanged (the MATLAB y persists across views there). */
int ostr_rdbt_sart_residual(ostr_handle* h, float* y, const float* p, const float* est,
                            const float* l, float thresh, int64_t n, void* stream);
/* SART_dbt.m:70-76, 98: x += step * (num / den) with NaN quotients 0; clamp != 0
 * then sets x = 0 where x < 0 (the end-of-iteration clamp). */
int ostr_rdbt_sart_update(ostr_handle* h, float* x, const float* num, const float* den,
                          float step, int32_t clamp, int64_t n, void* stream);
/* ML_dbt.m:88-96: yb = I0 exp(-drr) (I0 per element, or i0_scalar when i0 is
 * NULL), ydif = yb - counts, wdrr = drr * yb. */
int ostr_rdbt_ml_terms(ostr_handle* h, float* ydif, float* wdrr, const float* drr,
                       const float* counts, const float* i0, float i0_scalar, int64_t n,
                       void* stream);
/* ML_dbt.m:100-105: upd = num / den where den != 0 (kept elsewhere), NaN -> 0;
 * x = x + x * upd * step, NaN -> 0, negative -> 0. */
int ostr_rdbt_ml_update(ostr_handle* h, float* x, float* upd, const float* num,
                        const float* den, float step, int64_t n, void* stream);

/* fprojectCB1_4R_GPU_SART (Tomo_OSTR.cu:2398-2433; kernels :1453-1550): the
 * detector subdivided 2x2, one ray per sub-bin (centre x_d0 + (i + 0.5) Bsize/2),
 * plain sums (no cos), averaged per bin: prj [slot][nby][nbx] (the merge's
 * detector-row-major layout) = (s(2i,2j) + s(2i+1,2j) + s(2i,2j+1) +
 * s(2i+1,2j+1)) / 4; prj_4r [slot][2 nby][2 nbx] receives the sub-bin sums when
 * not NULL. Missed sub-bins count 0 (the reference leaves them unwritten, and
 * its merge reads an uninitialised fourth index, :1545-1547). */
int ostr_fproj_4r(ostr_handle* h, const float* img, float* prj, float* prj_4r, int32_t a0,
                  int32_t a1, void* stream);

/* Fused OS-SART subset, two launches (no reference counterpart: the reference
 * runs the five launchers above one after another).
 * ostr_sart_fproj_diff: diff = SART_prj_diff(p, fproj_sart(img)) over slots
 * [a0,a1) with p read and diff written detector-row-major [slot][nby][nbx].
 * ostr_sart_update_fused: x_out = the bprojectCB_4B_GPU_SART[_R] update of x_in
 * from diff_rows, with the Huber prior of x_in (delta) when beta != 0; x_in and
 * x_out must differ. At most 64 views per call. */
int ostr_sart_fproj_diff(ostr_handle* h, const float* img, const float* p_rows,
                         float* diff_rows, int32_t a0, int32_t a1, int32_t cos_weight,
                         void* stream);
int ostr_sart_update_fused(ostr_handle* h, const float* x_in, float* x_out,
                           const float* diff_rows, int32_t a0, int32_t a1, float lambda,
                           float beta, float delta, void* stream);

/* The momentum step of accelerated OSTR with the restart decision on the
 * device, so an iteration needs no host round trip and can be captured in a
 * HIP graph (no reference counterpart). state: 4 doubles in device memory,
 * {t, c, restarts, last test value}, set to {1, 0, 0, 0} by ostr_momentum_init.
 * ostr_momentum_step: dot = <x_k - z, z - z_prev> (FP64, fixed order); dot > 0
 * restarts (t = 1, c = 0, restarts += 1), else t' = (1 + sqrt(1 + 4 t^2)) / 2,
 * c = (t - 1) / t'; then x_next = max(0, z + c (z - z_prev)) in FP32, stored
 * like the inputs. The four images share one storage type `dtype` (OSTR_F32 /
 * OSTR_BF16 / OSTR_F16); the dot widens each element exactly to double.
 * Asynchronous. */
int ostr_momentum_init(ostr_handle* h, double* state, void* stream);
int ostr_momentum_step(ostr_handle* h, void* x_next, const void* z, const void* z_prev,
                       const void* x_k, int32_t dtype, double* state, void* stream);

/* FP64 accumulation over the n elements (no reference counterpart; the
 * denominator's sum over views in double): ostr_accum_f64 acc += x,
 * ostr_f64_to_f32 out = (float) acc (round to nearest). Asynchronous. */
int ostr_accum_f64(ostr_handle* h, double* acc, const float* x, int64_t n, void* stream);
int ostr_f64_to_f32(ostr_handle* h, float* out, const double* acc, int64_t n, void* stream);

/* x = max(0, z + c (z - z_prev)) over the volume (f32): the momentum step of
 * accelerated OSTR (Nesterov momentum on the OS / SQS update, Kim, Ramani &
 * Fessler, IEEE TMI 2015). No reference counterpart (the reference iterates
 * plain OSTR only). x may alias neither z nor z_prev. */
int ostr_extrapolate(ostr_handle* h, float* x, const float* z, const float* z_prev, float c,
                     void* stream);

/* Synchronous. *out = sum over the volume of (x - z)(z - z_prev) in FP64
 * (deterministic): the gradient-restart test of accelerated SQS
 * (O'Donoghue & Candes 2015): restart the momentum when it is > 0, x being
 * the point the last pass started from and z its result. */
int ostr_restart_dot(ostr_handle* h, const float* x, const float* z, const float* z_prev,
                     double* out, void* stream);

/* ---- batched cases: one launch over several reconstructions ----------------
 * The reference runs one geometry per process (__constant__ state,
 * Tomo_OSTR.cu:81-115) and so one case per launch; these calls have no
 * reference counterpart. They run the two fused OSTR kernels (ostr_fproj_h and
 * ostr_update_fused) for the current subset of every case of a batch in ONE
 * launch each; cases may differ in geometry, thickness and view count. Per
 * case the results are bit-identical to the single-case calls (every bin and
 * voxel is computed by the same code in the same order; the tile shape only
 * changes which work-group computes it). Images are f32. */
typedef struct ostr_batch ostr_batch;

typedef struct ostr_batch_case {
    ostr_handle* h;         /* geometry and slot table of the case             */
    int32_t a0, a1;         /* slots [a0, a1) of the handle's index (<= 64)   */
    const void* x_in;       /* image read (f32)                                */
    void* x_out;            /* image written by ostr_batch_update_fused        */
    const float* d;         /* denominator [nz][ny][nx]                        */
    const float* y;         /* subset counts, detector-row-major [slot][nby][nbx] */
    const float* b;         /* subset blank, same layout                       */
    const float* r;         /* subset scatter, same layout, or NULL            */
    void* hsub;             /* subset h (prj dtype), written by ostr_batch_fproj_h */
    float beta, capl_scale, delta, z_xy_ratio;
    int32_t guard;
} ostr_batch_case;

/* Tile and grid shapes of the batched launches (the autotuner's choices).
 * groups_per_cu = 0: one work-group per tile; k > 0: a persistent grid of k
 * work-groups per compute unit, each looping over its share of the tiles. */
typedef struct ostr_tile_cfg {
    int32_t fwd_rows;           /* detector rows per forward work-group: 1, 2, 4, 8 */
    int32_t fwd_groups_per_cu;
    int32_t upd_rows;           /* voxel rows per wave in the update: 4 or 8        */
    int32_t upd_waves;          /* waves per update work-group: (8,2) (8,4) (4,4)   */
    int32_t upd_groups_per_cu;
} ostr_tile_cfg;

/* A batch object on `device` for up to max_cases cases: owns the device case
 * table and its pinned staging ring. Calls on one batch object are ordered by
 * the caller (one host thread). */
int ostr_batch_create(int32_t device, int32_t max_cases, ostr_batch** out);
void ostr_batch_destroy(ostr_batch* bt);
const char* ostr_batch_last_error(const ostr_batch* bt);

/* ostr_fproj_h for every case: hsub_c = h(y_c, b_c, r_c, cos_t * A_c x_in_c)
 * over the slots [a0_c, a1_c) of case c. */
int ostr_batch_fproj_h(ostr_batch* bt, const ostr_batch_case* cases, int32_t n_cases,
                       int32_t prj_dtype, const ostr_tile_cfg* cfg, void* stream);

/* ostr_update_fused for every case (one prior potential for the batch; beta,
 * delta, z_xy_ratio, capl_scale, guard per case). */
int ostr_batch_update_fused(ostr_batch* bt, const ostr_batch_case* cases, int32_t n_cases,
                            int32_t prj_dtype, int32_t potential, const ostr_tile_cfg* cfg,
                            void* stream);

#ifdef __cplusplus
}
#endif
#endif /* OSTR_HIP_H */
