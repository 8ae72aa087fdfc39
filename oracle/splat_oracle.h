// CPU oracle for the 3D Gaussian Splatting trainer path -- TEST INFRASTRUCTURE ONLY.
//
// A serial C++17 restatement of the reference trainer in engine/src/splat (projection, tile
// binning, forward and backward compositing, L1 + D-SSIM loss, Adam, densification, trainer
// schedule, kNN initialisation). It is the checker the GPU path is compared against and is
// never linked into, called by, or shipped with the product library. Only tests/, the smoke
// check and the CPU-baseline leg of the benchmark load it.
//
// Floating point follows the reference's operation order with contraction disabled
// (-ffp-contract=off), so the FP32 functions reproduce the reference bit for bit when the
// reference runs on one thread (its only order-dependent step is the float atomic merge of
// the backward pass, Rasterizer.cpp:419-425, which is serial here).
//
// Parity status per function is listed in the header comment of splat_oracle.cpp.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int width, height;
    float fx, fy, cx, cy;
    float R[9]; // world -> camera, row-major
    float t[3];
} orc_camera; // RenderCamera, Rasterizer.h:13-21

typedef struct {
    int tile_size;
    float near_plane;
    float eps2d;
    float background[3];
    float min_alpha;
    float min_transmittance;
} orc_settings; // RasterSettings, Rasterizer.h:23-32

typedef struct {
    float *means, *scales, *quats, *opacities, *colors;
    int n;
} orc_model; // GaussianModel SoA, GaussianModel.h:23-37

typedef struct {
    float means, scales, quats, opacities, colors;
} orc_lrs; // LearningRates, Adam.h:8-14

typedef struct {
    int iterations, max_gaussians;
    float ssim_lambda;
    orc_lrs lr;
    float means_lr_final, densify_grad_threshold, percent_dense;
    int refine_start, refine_stop, refine_interval, opacity_reset_interval;
    float prune_opacity;
    unsigned seed;
    orc_settings raster;
} orc_train_options; // TrainOptions, Trainer.h:19-37

typedef struct {
    float total, l1, ssim;
} orc_loss_value; // LossValue, Loss.h:7-11

// ---- scalar helpers
float orc_expf(float x);                                   // host C library expf
void orc_expf_array(const float *x, float *y, long n);
void orc_default_settings(orc_settings *s);
void orc_options_for_iterations(int iterations, orc_train_options *o);

// ---- projection (Rasterizer.cpp:48-110, 137-204). All outputs sized by n; only entries
// with radius > 0 are written besides radius and flags (both zeroed for every entry).
void orc_project(const orc_model *m, const orc_camera *cam, const orc_settings *st,
                 float *mean2d, float *conic, float *cov2d, float *depth, float *rgb,
                 float *alpha, int *radius, int *rect, uint8_t *flags);

// ---- tile binning (Rasterizer.cpp:206-261). Returns K (pairs) or -1 if K > capacity.
long orc_bin_tiles(int n, const int *radius, const int *rect, const float *depth, int tiles_x,
                   int tiles_y, int *tile_offsets, int *tile_list, long capacity);

// ---- rasterizer object (Rasterizer.h:38-90)
typedef struct orc_raster orc_raster;
orc_raster *orc_raster_new(void);
void orc_raster_free(orc_raster *r);
void orc_raster_forward(orc_raster *r, const orc_model *m, const orc_camera *cam,
                        const orc_settings *st, float *rgb);
// Adds into grads (same n) and, if non-null, into mean2d_grad_ndc[n].
void orc_raster_backward(orc_raster *r, const orc_model *m, const float *dl_drgb,
                         orc_model *grads, float *mean2d_grad_ndc);
// SH degree 1..3 variants (extension): rest = 15 x 3 coefficients per Gaussian, coefficient-
// major; grad_rest (may be NULL) is added to.
void orc_raster_forward_sh(orc_raster *r, const orc_model *m, const float *rest, int degree,
                           const orc_camera *cam, const orc_settings *st, float *rgb);
void orc_raster_backward_sh(orc_raster *r, const orc_model *m, const float *rest, int degree,
                            const float *dl_drgb, orc_model *grads, float *grad_rest,
                            float *mean2d_grad_ndc);
// Accessors of the state kept by the last forward.
long orc_raster_num_pairs(const orc_raster *r);
int orc_raster_num_tiles(const orc_raster *r);
void orc_raster_get_state(const orc_raster *r, int *radius, int *tile_offsets, int *tile_list,
                          float *final_t, int *n_contrib);
void orc_raster_get_projection(const orc_raster *r, float *mean2d, float *conic, float *depth,
                               float *rgb, float *alpha, int *rect, uint8_t *flags);
void orc_raster_get_grad2d(const orc_raster *r, float *grad2d); // [9n] after backward
// Projection (Rasterizer.cpp:137-204) in FP64 of double-precision parameters: mean2d [2n],
// conic [3n], clamped rgb [3n], alpha [n] where radius [n] > 0.
void orc_project_f64(int n, const double *means, const double *scales, const double *quats,
                     const double *opacities, const double *colors, const double *rest, int degree,
                     const orc_camera *camera, const orc_settings *st, double *mean2d,
                     double *conic, double *rgb, double *alpha, int *radius);
// The backward's chain-rule stage (Rasterizer.cpp:432-544) in FP64, from the given 2D gradients
// [9n] and the last forward's state; the outputs (zeroed by the caller) are added to. g_rest may
// be NULL; rest is NULL for degree 0.
void orc_chain_rule_f64(const orc_raster *r, const orc_model *m, const float *rest, int degree,
                        const float *grad2d, double *g_means, double *g_scales, double *g_quats,
                        double *g_opacities, double *g_colors, double *g_rest);

// ---- FP64 forward with the same algorithm (finite-difference oracle). Model as doubles.
void orc_render_f64(int n, const double *means, const double *scales, const double *quats,
                    const double *opacities, const double *colors, const orc_camera *cam,
                    const orc_settings *st, double *rgb);

void orc_render_f64_sh(int n, const double *means, const double *scales, const double *quats,
                       const double *opacities, const double *colors, const double *rest,
                       int degree, const orc_camera *cam, const orc_settings *st, double *rgb);
void orc_adam_array(float *p, const float *g, float *m, float *v, long len, float lr, int step);
void orc_sh_basis(float x, float y, float z, int degree, double *out15);

// ---- loss (Loss.cpp:47-185) and psnr (Loss.cpp:187-196)
orc_loss_value orc_photometric_loss(const float *render, const float *target, int w, int h,
                                    float lambda, float *grad);
float orc_psnr(const float *a, const float *b, long count);

// ---- Adam (Adam.cpp:13-28, 68-78): one step over all groups, moments m, v in model layout.
void orc_adam_step(orc_model *params, const orc_model *grads, orc_model *m, orc_model *v,
                   const orc_lrs *lr, int step);

// ---- densification decisions with given statistics and pre-drawn normals
// (Trainer.cpp:105-169). keep[n] (0/1), kind[n] (0 none, 1 clone, 2 split accepted),
// returns the number of appended Gaussians; the appended block (clones and split children,
// in index order) is written to `extra` (capacity extra_cap). `normals` holds 6 draws per
// accepted split in acceptance order; *normals_used receives how many were consumed.
long orc_densify(const orc_model *m, const float *grad_accum, const int *grad_count,
                 float extent, int max_gaussians, float grad_threshold, float percent_dense,
                 float prune_opacity, int opacity_was_reset, const float *normals,
                 long normals_len, uint8_t *keep, uint8_t *kind, orc_model *extra, long extra_cap,
                 long *normals_used);

// ---- kNN-distance initialisation (GaussianModel.cpp:68-144). rgb may be NULL.
void orc_init_from_points(int n, const float *xyz, const float *rgb, int rgb_count,
                          orc_model *out);

// ---- viewer (app/src/SplatView.cpp:117-180, 349-397): rgba [4n] and cov6 [6n] of a model; the
// front-to-back order [n] of centers [3n] for the view matrix's third row row2 [4].
void orc_view_attributes(const orc_model *m, uint8_t *rgba, float *cov6);
void orc_view_depth_order(const float *centers, long n, const float *row2, uint32_t *order);
// ---- trainer (Trainer.cpp:37-197). Images: cams[n_images], rgb8 concatenated HWC uint8.
// orc_trainer_set_sh (extension, call before the first step): SH degree 1..3 coefficients
// starting at rest (NULL: zero), trained with rate lr.colors * lr_scale (0 -> 1/20); orc_trainer_get_sh
// copies them out ([size][15][3]).
typedef struct orc_trainer orc_trainer;
orc_trainer *orc_trainer_new(const orc_model *initial, const orc_camera *cams,
                             const uint8_t *rgb8, int n_images, float extent,
                             const orc_train_options *opt);
void orc_trainer_free(orc_trainer *t);
void orc_trainer_set_sh(orc_trainer *t, int degree, float lr_scale, const float *rest);
void orc_trainer_get_sh(const orc_trainer *t, float *rest);
void orc_trainer_step(orc_trainer *t);
int orc_trainer_done(const orc_trainer *t);
int orc_trainer_iteration(const orc_trainer *t);
int orc_trainer_size(const orc_trainer *t);
void orc_trainer_get_model(const orc_trainer *t, orc_model *out); // out arrays sized by size
void orc_trainer_stats(const orc_trainer *t, int *iteration, float *loss, float *psnr,
                       long *gaussians);
float orc_trainer_evaluate_psnr(orc_trainer *t);
float orc_trainer_evaluate_l1(orc_trainer *t);
void orc_trainer_render(orc_trainer *t, const orc_camera *cam, float *rgb);

#ifdef __cplusplus
}
#endif
