/*
 * u3d_hip.h -- C ABI of the MI355X Gaussian Splatting trainer library (libu3d_hip.so).
 *
 * This is the drop-in boundary for the trainer hot path of engine/src/splat. The reference
 * exposes a C++ class API (no FFI); every entry point below replaces one of its calls and the
 * C++ shims in univeral_3d_reconstruction_system_amd/host/ re-create the reference classes
 * (u3d::Rasterizer, u3d::photometricLoss, u3d::Adam, u3d::Trainer) on top of it.
 *
 * Conventions
 *  - Every function returns U3D_OK (0) or an error code; u3d_last_error() returns a
 *    thread-local message for the last failure. No C++ exception crosses the ABI.
 *  - Array arguments of the compute functions are DEVICE pointers (HIP global memory of the
 *    context's device) unless the name ends in _host. Work is queued on the context's stream;
 *    functions that return values to the host synchronise that stream.
 *  - Layouts are the reference's: Gaussian fields are structure-of-arrays with xyz / wxyz
 *    interleaved per field (GaussianModel.h:23-37), images are height*width*3 interleaved RGB
 *    (Rasterizer.h:44), pixel (x, y) is sampled at the integer coordinate (Rasterizer.h:36-37).
 *  - A context is used from one host thread at a time, like the reference objects.
 */
#ifndef U3D_HIP_H
#define U3D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define U3D_OK 0
#define U3D_ERR_INVALID 1     /* bad argument or size */
#define U3D_ERR_HIP 2         /* HIP runtime failure */
#define U3D_ERR_CAPACITY 3    /* a fixed capacity was exceeded */
#define U3D_ERR_UNSUPPORTED 4 /* valid for the reference, not supported by this build */

typedef struct u3d_ctx u3d_ctx;
typedef struct u3d_raster u3d_raster;
typedef struct u3d_adam u3d_adam;
typedef struct u3d_trainer u3d_trainer;

/* RenderCamera (Rasterizer.h:13-21): float pinhole, R row-major world->camera, OpenCV axes. */
typedef struct {
    int width, height;
    float fx, fy, cx, cy;
    float R[9];
    float t[3];
} u3d_camera;

/* RasterSettings (Rasterizer.h:23-32). Any tile_size > 0; multiples of 16 take the LDS-staged
 * compositing kernels (16: the quadrant forward and the segmented backward), other sizes a
 * per-pixel walk. The batched engine (views_per_step > 1, use_graph, scene sets) needs a
 * multiple of 16 and returns U3D_ERR_UNSUPPORTED otherwise. */
typedef struct {
    int tile_size;
    float near_plane;
    float eps2d;
    float background[3];
    float min_alpha;
    float min_transmittance;
} u3d_raster_settings;

/* GaussianModel (GaussianModel.h:23-37) as device pointers; also used for gradients and
 * Adam moments. colors holds 3 SH-DC coefficients per Gaussian. */
typedef struct {
    float *means;     /* [3n] */
    float *scales;    /* [3n] log standard deviations */
    float *quats;     /* [4n] w x y z, not necessarily normalised */
    float *opacities; /* [n]  logits */
    float *colors;    /* [3n] SH degree-0 coefficients */
    int64_t n;
} u3d_gaussians;

/* LearningRates (Adam.h:8-14). */
typedef struct {
    float means, scales, quats, opacities, colors;
} u3d_learning_rates;

/* LossValue (Loss.h:7-11). */
typedef struct {
    float total, l1, ssim;
} u3d_loss_value;

/* TrainOptions (Trainer.h:19-37). */
typedef struct {
    int iterations;
    int max_gaussians;
    float ssim_lambda;
    u3d_learning_rates lr;
    float means_lr_final;
    float densify_grad_threshold;
    float percent_dense;
    int refine_start, refine_stop, refine_interval, opacity_reset_interval;
    float prune_opacity;
    unsigned seed;
    u3d_raster_settings raster;
    /* Extensions (no reference counterpart; zero keeps the reference behaviour). */
    int sh_degree;       /* 0..3: view-dependent colour, coefficients start at zero */
    float sh_lr_scale;   /* rate of the SH coefficients = lr.colors * scale (0 -> 1/20) */
    int views_per_step;  /* B >= 1 views per optimiser step (0 -> 1) */
    int use_graph;       /* capture each step's device work as a hipGraph */
    /* With B > 1: 1 scales the learning rates by sqrt(B) and the Adam decay rates to beta^B
     * (the batch-size rule of Zhao et al., "On Scaling Up 3D Gaussian Splatting Training",
     * 2024); 0 keeps them. */
    int batch_scaling;
} u3d_train_options;

/* TrainStats (Trainer.h:39-44). */
typedef struct {
    int iteration;
    float loss;
    float psnr;
    int64_t gaussians;
} u3d_train_stats;

/* Device views of the state a forward pass leaves for backward (Rasterizer.h:60-89). */
typedef struct {
    int64_t n;               /* Gaussians of the last forward */
    int64_t num_pairs;       /* K = tile intersections */
    int tiles_x, tiles_y;
    const int *radius;       /* [n], 0 = not visible */
    const int *rect;         /* [4n] tile x0 y0 x1 y1 (exclusive), valid where radius > 0 */
    const float *depth;      /* [n] camera-space z, valid where radius > 0 */
    const float *mean2d;     /* [2n] */
    const float *conic;      /* [3n] A B C of the inverse 2D covariance */
    const float *color;      /* [3n] clamped rgb */
    const float *opacity;    /* [n] sigmoid(opacity logit) */
    const uint8_t *flags;    /* [n] bit0/1 Jacobian x/y clamped, bit2..4 colour clamped */
    const int *tile_offsets; /* [tiles+1] */
    const int *tile_list;    /* [K] Gaussian index per (tile, depth) slot */
    const float *final_t;    /* [H*W] transmittance after compositing */
    const int *n_contrib;    /* [H*W] last contributing list index + 1 */
} u3d_raster_state;

/* ------------------------------------------------------------------------------ context */
const char *u3d_last_error(void);
const char *u3d_version(void);
/* Creates a context on HIP device `device` with its own stream (default priority). */
int u3d_ctx_create(int device, u3d_ctx **out);
void u3d_ctx_destroy(u3d_ctx *ctx);
int u3d_ctx_synchronize(u3d_ctx *ctx);
/* The context's hipStream_t, for callers that enqueue their own work around the library. */
void *u3d_ctx_stream(u3d_ctx *ctx);
int u3d_malloc(u3d_ctx *ctx, size_t bytes, void **out);
int u3d_free(u3d_ctx *ctx, void *ptr);
int u3d_memcpy_h2d(u3d_ctx *ctx, void *dst, const void *src_host, size_t bytes);
int u3d_memcpy_d2h(u3d_ctx *ctx, void *dst_host, const void *src, size_t bytes);
int u3d_memset(u3d_ctx *ctx, void *dst, int value, size_t bytes);
/* Context options (no reference counterpart). U3D_OPT_SEGMENTED_BACKWARD (default 1): with
 * 16-pixel tiles the forward keeps a per-pixel checkpoint (transmittance, front colour) at
 * every 64-record boundary of each tile list, and the backward walks those segments as
 * independent work items instead of each tile's whole list in one workgroup. Images are
 * unchanged; gradients differ from the unsegmented walk by float rounding only. Applies from
 * the next forward. */
#define U3D_OPT_SEGMENTED_BACKWARD 1
/* U3D_OPT_SORT_SPIN_LIMIT (default 0 = a bound no wait that ends comes near): how many lookback
 * re-reads a pass of the depth sort over all Gaussians (used for tile grids beyond the direct
 * binning's, (tiles_x + 1)(tiles_y + 1) > 12288) makes before it gives up a wait and flags the
 * sort as failed; a negative value gives every wait up at once. The kernels stay in bounds
 * either way, and that forward (or the trainer step that ran it) returns U3D_ERR_HIP. A test
 * hook for that error path. */
#define U3D_OPT_SORT_SPIN_LIMIT 2
int u3d_ctx_set_option(u3d_ctx *ctx, int option, int value);

/* ------------------------------------------------------------------- rasterizer (L1 core) */
/* Rasterizer(ThreadPool&) (Rasterizer.h:41). */
int u3d_raster_create(u3d_ctx *ctx, u3d_raster **out);
void u3d_raster_destroy(u3d_raster *r);
/* Rasterizer::forward (Rasterizer.h:44-45, Rasterizer.cpp:263-330): renders into rgb
 * [H*W*3] and keeps the state for the next backward. Bit-identical to the reference. */
int u3d_raster_forward(u3d_raster *r, const u3d_gaussians *model, const u3d_camera *camera,
                       const u3d_raster_settings *settings, float *rgb);
/* Rasterizer::backward (Rasterizer.h:47-51, Rasterizer.cpp:334-545): ADDS the parameter
 * gradients of the last forward into `grads` (same n) and, if non-null, the densification
 * statistic |dL/d mean2d| in NDC into mean2d_grad_ndc [n]. */
int u3d_raster_backward(u3d_raster *r, const u3d_gaussians *model, const float *dl_drgb,
                        u3d_gaussians *grads, float *mean2d_grad_ndc);
/* Spherical-harmonic colour of degree 1..3 (extension; the reference trains DC only):
 * sh_rest holds 15 x 3 coefficients per Gaussian (coefficient-major, [n][15][3]); the first
 * 3, 8 or 15 are used for degree 1, 2, 3. Degree 0 (or sh_rest == NULL) is the plain call. */
int u3d_raster_forward_sh(u3d_raster *r, const u3d_gaussians *model, const float *sh_rest,
                          int sh_degree, const u3d_camera *camera,
                          const u3d_raster_settings *settings, float *rgb);
/* Backward of the last forward_sh: also ADDS dL/d sh_rest into grad_sh_rest [45n] (nullable). */
int u3d_raster_backward_sh(u3d_raster *r, const u3d_gaussians *model, const float *sh_rest,
                           int sh_degree, const float *dl_drgb, u3d_gaussians *grads,
                           float *grad_sh_rest, float *mean2d_grad_ndc);
/* Rasterizer::radii (Rasterizer.h:53-54): device pointer valid until the next forward. */
int u3d_raster_radii(u3d_raster *r, const int **radii, int64_t *n);
/* Test and tooling access to the forward state (Rasterizer.h:60-89). */
int u3d_raster_get_state(u3d_raster *r, u3d_raster_state *out);
/* The per-Gaussian 2D gradients [9n] (mx my A B C r g b alpha) of the last backward. */
int u3d_raster_get_grad2d(u3d_raster *r, const float **grad2d);

/* ------------------------------------------------------------------------------- loss */
/* photometricLoss (Loss.h:18-19, Loss.cpp:47-185): (1-l)*L1 + l*(1-SSIM), 11x11 sigma 1.5
 * window, zero padding; OVERWRITES grad [H*W*3]; the value is returned to the host. */
int u3d_photometric_loss(u3d_ctx *ctx, const float *render, const float *target, int width,
                         int height, float ssim_lambda, float *grad, u3d_loss_value *out_host);
/* No counterpart in the reference: u3d_photometric_loss for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). Same results bit for bit (the
 * loss as one fused launch, where u3d_photometric_loss passes the SSIM partials through its
 * scratch between two launches); its
 * scratch is a static 1 MiB array of the library on each GPU (images up to 14,559 blocks of
 * 32x32 pixels, e.g. 3840x3840; larger ones return U3D_ERR_CAPACITY), so it calls no hipMalloc
 * and never synchronises. value_device (nullable) receives {total, l1, ssim, psnr} as 4 device
 * floats. All calls on one GPU share that scratch: calls that overlap on one GPU race. */
int u3d_photometric_loss_noalloc(u3d_ctx *ctx, const float *render, const float *target, int width,
                                 int height, float ssim_lambda, float *grad, float *value_device);
/* psnr (Loss.h:21, Loss.cpp:187-196) of two device arrays. */
int u3d_psnr(u3d_ctx *ctx, const float *a, const float *b, int64_t count, float *out_host);

/* ------------------------------------------------------------------------------- Adam */
/* Adam (Adam.h:18-33): moments live on the device in GaussianModel layout. */
int u3d_adam_create(u3d_ctx *ctx, u3d_adam **out);
void u3d_adam_destroy(u3d_adam *a);
int u3d_adam_reset(u3d_adam *a, int64_t n);                     /* Adam::reset, Adam.cpp:31-38 */
int u3d_adam_compact(u3d_adam *a, const uint8_t *keep);        /* Adam::compact, :40-44 (device keep[n]) */
int u3d_adam_grow(u3d_adam *a, int64_t count);                 /* Adam::grow, :46-60 */
int u3d_adam_reset_opacity_moments(u3d_adam *a);               /* :62-66 */
/* Adam::step (Adam.cpp:68-78): one step over all five groups with per-group rates. */
int u3d_adam_step(u3d_adam *a, u3d_gaussians *params, const u3d_gaussians *grads,
                  const u3d_learning_rates *lr);
/* Device views of the first and second moments. */
int u3d_adam_moments(u3d_adam *a, u3d_gaussians *m, u3d_gaussians *v, int *step);

/* --------------------------------------------------------------------- densification */
/* Decisions of Trainer::densifyAndPrune (Trainer.cpp:105-169) given the statistics:
 * keep[n] (1 keep / 0 remove), kind[n] (0 none, 1 clone, 2 split), and the counts. Device
 * arrays; counts are returned to the host. budget = max(0, max_gaussians - n). */
int u3d_densify_decide(u3d_ctx *ctx, const u3d_gaussians *model, const float *grad_accum,
                       const int *grad_count, float extent, int max_gaussians,
                       float grad_threshold, float percent_dense, float prune_opacity,
                       int opacity_was_reset, uint8_t *keep, uint8_t *kind,
                       int64_t *n_clone_host, int64_t *n_split_host, int64_t *n_keep_host);
/* Applies the decisions in place: compacts kept entries (order preserved) and appends the
 * clones and split children in index order; `normals` holds 6 N(0,1) draws per accepted
 * split in acceptance order (device, may be NULL if no split). `log_shrink` = log(1.6f).
 * The model arrays must have capacity for the new size, returned in *new_n_host. */
int u3d_densify_apply(u3d_ctx *ctx, u3d_gaussians *model, const uint8_t *keep,
                      const uint8_t *kind, const float *normals, float log_shrink,
                      int64_t *new_n_host);

/* ------------------------------------------------------------------------------ trainer */
/* Trainer (Trainer.h:249-292) with all state resident on the device. Host inputs: the initial
 * model (SoA host arrays), cameras[n_images], images as concatenated uint8 HWC. */
int u3d_trainer_create(u3d_ctx *ctx, const float *means_host, const float *scales_host,
                       const float *quats_host, const float *opacities_host,
                       const float *colors_host, int64_t n, const u3d_camera *cameras,
                       const uint8_t *rgb8_host, int n_images, float scene_extent,
                       const u3d_train_options *options, u3d_trainer **out);
void u3d_trainer_destroy(u3d_trainer *t);
int u3d_trainer_step(u3d_trainer *t);                 /* Trainer::step, Trainer.cpp:53-103 */
int u3d_trainer_done(u3d_trainer *t);                 /* 1 when finished */
int u3d_trainer_iteration(u3d_trainer *t);
int64_t u3d_trainer_size(u3d_trainer *t);
/* Trainer::model: copies the current parameters to host arrays sized by u3d_trainer_size. */
int u3d_trainer_get_model(u3d_trainer *t, float *means_host, float *scales_host,
                          float *quats_host, float *opacities_host, float *colors_host);
/* Adam moments (Adam.h:30-31) as 14 * size floats each: means 3n, scales 3n, quats 4n,
 * opacities n, colours 3n (checkpointing; tests). */
int u3d_trainer_get_moments(u3d_trainer *t, float *m_host, float *v_host);
/* SH coefficients [45 * size] of a trainer created with sh_degree > 0 ([n][15][3], as
 * u3d_raster_forward_sh). set_sh_rest replaces them (the trainer starts them at zero). */
int u3d_trainer_get_sh_rest(u3d_trainer *t, float *sh_rest_host);
int u3d_trainer_set_sh_rest(u3d_trainer *t, const float *sh_rest_host);
/* Adam moments of the SH coefficients, [45 * size] each, same layout (checkpointing; tests). */
int u3d_trainer_get_sh_moments(u3d_trainer *t, float *m_host, float *v_host);
/* The densification statistics since the last densification (Trainer.cpp:76-79, :442-445):
 * accumulated |dL/d mean2d| in NDC [size] and the count of steps each Gaussian was visible. */
int u3d_trainer_get_densify_stats(u3d_trainer *t, float *accum_host, int *count_host);
int u3d_trainer_stats(u3d_trainer *t, u3d_train_stats *out);      /* Trainer::stats */
int u3d_trainer_evaluate_psnr(u3d_trainer *t, float *out_host);   /* Trainer.cpp:180-192 */
/* Mean absolute error over all training views (time-to-target metric). */
int u3d_trainer_evaluate_l1(u3d_trainer *t, float *out_host);
int u3d_trainer_render(u3d_trainer *t, const u3d_camera *camera, float *rgb_host);

/* ------------------------------------------------------------------ seed-point initialisation */
/* initGaussiansFromPoints (GaussianModel.cpp:68-144) on the device: xyz [3n] and rgb
 * [rgb_count] are device arrays (rgb_count < 3n leaves the missing colours at 0.5, as the
 * reference); `out` needs n entries of capacity (out->n >= n). The hash-grid neighbour search
 * reproduces the reference's buckets, so the mean 3-NN squared distance is bit-identical. */
int u3d_points_knn_init(u3d_ctx *ctx, const float *xyz, const float *rgb, int rgb_count,
                        int64_t n, u3d_gaussians *out);

/* Learned initialiser (no reference counterpart): an MLP 80 -> 64 -> 256 -> 256 -> 256 -> 16,
 * GELU, BF16 weights and activations with FP32 accumulation on MFMA, trained on the device.
 * Its 11 outputs correct the kNN initialisation of each point: 3 log-scales, 4 quaternion
 * components, the opacity logit and 3 DC colour coefficients. */
typedef struct u3d_init_net u3d_init_net;
#define U3D_FEATURE_DIM 80
#define U3D_NET_OUT 16
/* Per-point network input [n x U3D_FEATURE_DIM] (device): offsets and distances of the 16
 * nearest neighbours in units of the point's kNN scale, colour - 0.5, and the fraction of the
 * cameras (host array, may be NULL) that see the point. */
int u3d_points_features(u3d_ctx *ctx, const float *xyz, const float *rgb, int rgb_count,
                        int64_t n, const u3d_camera *cameras_host, int n_cameras,
                        float near_plane, float *features);
int u3d_init_net_create(u3d_ctx *ctx, unsigned seed, u3d_init_net **out);
void u3d_init_net_destroy(u3d_init_net *net);
/* Parameters: per layer, W [out x in] row-major then b [out], FP32 master copies. */
int64_t u3d_init_net_param_count(void);
int u3d_init_net_get_params(u3d_init_net *net, float *params_host);
int u3d_init_net_set_params(u3d_init_net *net, const float *params_host);
/* Gradients of the last training step, same layout as the parameters. */
int u3d_init_net_get_grads(u3d_init_net *net, float *grads_host);
/* out [n x U3D_NET_OUT] (device) = network(features). */
int u3d_init_net_forward(u3d_init_net *net, const float *features, int64_t n, float *out);
/* One Adam step (lr; betas 0.9 / 0.999, eps 1e-8) on the mean over the n rows of
 * sum_c w[c] (out[c] - target[c])^2 for the first 11 columns (targets [n x 16] device,
 * weights16 host). *loss_host (nullable) receives the loss before the step. */
int u3d_init_net_train_step(u3d_init_net *net, const float *features, const float *targets,
                            const float *weights16_host, int64_t n, float lr, float *loss_host);
/* kNN initialisation corrected by the network (net == NULL: the kNN initialisation). */
int u3d_points_learned_init(u3d_ctx *ctx, u3d_init_net *net, const float *xyz, const float *rgb,
                            int rgb_count, int64_t n, const u3d_camera *cameras_host,
                            int n_cameras, float near_plane, u3d_gaussians *out);

/* -------------------------------------------------------------------------- model files */
/* writeSplatPly (SplatPly.h:12-13, SplatPly.cpp:261-281) from host arrays: binary little
 * endian x y z nx ny nz f_dc_0..2 [f_rest_*] opacity scale_0..2 rot_0..3. With sh_degree 1..3
 * the SH coefficients sh_rest [n][15][3] are written as f_rest_0 .. f_rest_{3K-1}
 * (channel-major, K = 3, 8, 15), the usual 3DGS extension of the reference's layout. */
int u3d_write_splat_ply(const char *path, const float *means_host, const float *scales_host,
                        const float *quats_host, const float *opacities_host,
                        const float *colors_host, int64_t n, const float *sh_rest_host,
                        int sh_degree);
/* Vertex count of a PLY file, whether it is a Gaussian splat file (it has f_dc_0, opacity,
 * scale_0 and rot_0, as readSplatPly decides, SplatPly.cpp:295) and the SH degree of its
 * f_rest_* properties (0 without). */
int u3d_ply_info(const char *path, int64_t *vertices, int *is_splat, int *sh_degree);
/* readSplatPly for splat files (ASCII or binary of either byte order, other elements skipped,
 * SplatPly.cpp:133-235, 284-342): vertices whose x, y or z is not finite are dropped and
 * *n_out is the count kept; with capacity 0 only *n_out and *sh_degree_out are set, otherwise
 * host arrays of at least *n_out entries are filled (sh_rest [n][15][3] when the file has
 * f_rest_*, missing ones zero). A point-cloud file returns U3D_ERR_INVALID: the reference turns
 * one into Gaussians with initGaussiansFromPoints (u3d_read_points_ply, u3d_points_knn_init). */
int u3d_read_splat_ply(const char *path, int64_t capacity, float *means_host, float *scales_host,
                       float *quats_host, float *opacities_host, float *colors_host,
                       float *sh_rest_host, int64_t *n_out, int *sh_degree_out);
/* writePointsPly (SplatPly.cpp:344-362): x y z float, red green blue uchar (rgb in 0..1,
 * clamped and rounded; 0.8 for points beyond rgb_count / 3). */
int u3d_write_points_ply(const char *path, const float *xyz_host, const float *rgb_host,
                         int64_t rgb_count, int64_t n);
/* readPointsPly (SplatPly.cpp:364-386): rgb in 0..1 (uchar / 255, ushort / 65535; 0.8 when the
 * file has no colours); points with a non-finite coordinate dropped. capacity 0: only *n_out. */
int u3d_read_points_ply(const char *path, int64_t capacity, float *xyz_host, float *rgb_host,
                        int64_t *n_out);
/* writeSplatPly(trainer.model()) (Reconstructor.cpp:361): one copy from the device. */
int u3d_trainer_export_ply(u3d_trainer *t, const char *path);

/* ------------------------------------------------------------------------------- viewer */
/* Display attributes of a loaded model (app/src/SplatView.cpp:117-180), device arrays: rgba
 * [4n] uint8 = round(clamp(0.5 + C0 * colour, 0, 1) * 255) and round(sigmoid(opacity) * 255),
 * cov6 [6n] = the upper triangle xx xy xz yy yz zz of M M^T, M = R(q / (|q| + 1e-12))
 * diag(exp(scale)). Synchronous. */
int u3d_view_splat_attributes(u3d_ctx *ctx, const u3d_gaussians *model, uint8_t *rgba, float *cov6);
/* Front-to-back order for a view (SplatView.cpp:349-397): depth z_i = -(v . (c_i, 1)) with v
 * the third row of the row-major 4x4 view matrix (view_row2 host [4]), 16-bit keys
 * (z_i - min z) * 65535 / (max z - min z) truncated, and order [n] (device) the stable sort of
 * the keys (the reference's counting sort). packed (device, nullable) receives the 40-byte
 * records {float center[3]; uint8 rgba[4]; float cov[6]} in that order from centers [3n],
 * rgba [4n] and cov6 [6n]. Synchronous. */
int u3d_view_depth_order(u3d_ctx *ctx, const float *centers, int64_t n, const float *view_row2,
                         uint32_t *order, const uint8_t *rgba, const float *cov6, void *packed);

/* --------------------------------------------------------------------------- scene sets */
/* Independent scenes trained together (no reference counterpart: the reference trains one
 * scene, one view per step, Trainer.cpp:64). Each step of a set trains every member that is
 * not done on its views_per_step views: one projection pass, one sort and one persistent
 * compositing launch per direction for all of them, then one Adam launch. Members keep their
 * own state, schedule and random stream, so a member trains as it would alone with the
 * batched engine. All members must share u3d_raster_settings, and each member's cameras one
 * image size. */
typedef struct u3d_scene_set u3d_scene_set;
/* hbm_budget_bytes: device memory the members and the engine may hold (0: 90% of the
 * device's free memory at creation). */
int u3d_scene_set_create(u3d_ctx *ctx, int use_graph, size_t hbm_budget_bytes,
                         u3d_scene_set **out);
void u3d_scene_set_destroy(u3d_scene_set *set); /* destroys the members too */
/* Admission by HBM budget: the set takes ownership of `scene` when its resident state plus
 * the engine's working set for it fits the remaining budget; otherwise returns
 * U3D_ERR_CAPACITY and the caller keeps the trainer. */
int u3d_scene_set_add(u3d_scene_set *set, u3d_trainer *scene);
int u3d_scene_set_size(u3d_scene_set *set);
/* Borrowed handle of member k, for the u3d_trainer_* queries (never step or destroy it). */
u3d_trainer *u3d_scene_set_member(u3d_scene_set *set, int k);
int u3d_scene_set_step(u3d_scene_set *set);
int u3d_scene_set_done(u3d_scene_set *set); /* 1 when every member is done */
/* Bytes of the budget in use (members + engine working set). */
size_t u3d_scene_set_bytes(u3d_scene_set *set);

/* Tuning of the batched engine (trainers with views_per_step > 1 or use_graph, scene sets).
 * Every choice gives the same images and gradients up to float summation order, except
 * tile_size, which changes which Gaussians reach a tile near the 3-sigma cut exactly as
 * RasterSettings::tileSize does in the reference. */
typedef struct {
    int heavy_first;        /* persistent compositing takes the longest tile lists first */
    int workgroups_per_cu;  /* persistent workgroups per CU (0: occupancy limit) */
    int lds_batch;          /* tile-list records staged in LDS per round: 64, 128, 256 (0: 256) */
    int sort_items;         /* radix sort keys per lane per block: 4 or 8 (0: by size) */
    int tile_size;          /* 0: keep the raster settings' tile size */
    int atomic_lanes;       /* backward: Gaussians reaching at most this many lanes of a wave
                             * (0, 4, 16) add per lane to LDS instead of a wave reduction */
    int sort_radix;         /* digit bits per radix pass of the step's sorts: 4..8 (0: 8) */
    int pair_sort;          /* how tile lists are built: 0: per-image difference grids in
                             * Gaussian index order, then every list sorted on the depth keys;
                             * 2: a radix sort of every entry's depth key, then per-image
                             * difference grids in that order; (0 and 2 where every image's tile
                             * grid fits the binning's LDS grid) 1: emitted (tile, Gaussian) pairs
                             * radix-sorted on the tile id */
} u3d_batch_tuning;
int u3d_trainer_get_tuning(u3d_trainer *t, u3d_batch_tuning *out);
int u3d_trainer_set_tuning(u3d_trainer *t, const u3d_batch_tuning *tuning);
int u3d_scene_set_set_tuning(u3d_scene_set *set, const u3d_batch_tuning *tuning);
/* Autotuner: times each candidate configuration (LDS batch x block order x workgroups per CU x
 * sort items, and tile sizes 16 / 32 when tune_tile_size; atomic_lanes stays 0, measured
 * slower at 300k / 640x480) over steps_per_candidate of the
 * trainer's next steps (after one warm-up step each; they are real training steps) and keeps
 * the fastest, written to *chosen (nullable). */
int u3d_trainer_autotune(u3d_trainer *t, int steps_per_candidate, int tune_tile_size,
                         u3d_batch_tuning *chosen);
int u3d_scene_set_autotune(u3d_scene_set *set, int steps_per_candidate, int tune_tile_size,
                           u3d_batch_tuning *chosen);

/* ----------------------------------------------------------------------- measurement */
/* Pipeline stages timed with HIP events on the context's stream (no reference counterpart:
 * the reference has no profiling; these feed the roofline report of the benchmark). */
#define U3D_STAGE_PROJECT 0    /* EWA projection + SH-DC colour */
#define U3D_STAGE_SORT 1       /* tile binning and per-tile depth order (or depth sort, pair
                                  emission, tile sort, tile offsets) */
#define U3D_STAGE_RASTER_FWD 2 /* front-to-back compositing */
#define U3D_STAGE_LOSS 3       /* fused L1 + D-SSIM value and gradient */
#define U3D_STAGE_RASTER_BWD 4 /* reverse compositing -> per-Gaussian 2D gradients */
#define U3D_STAGE_CHAIN 5      /* 2D -> 3D chain rule */
#define U3D_STAGE_ADAM 6       /* fused Adam over all parameter groups */
#define U3D_STAGE_DENSIFY 7    /* clone / split / prune / opacity reset */
#define U3D_STAGE_GRAPH 8      /* one replay of a captured whole-step graph (batched engine) */
#define U3D_NUM_STAGES 9
/* Turns event timing on or off (off by default) and clears the accumulated totals. */
int u3d_ctx_set_timing(u3d_ctx *ctx, int enabled);
/* Waits for the recorded events; writes total milliseconds and launch counts per stage. */
int u3d_ctx_stage_times(u3d_ctx *ctx, double total_ms[U3D_NUM_STAGES],
                        int64_t counts[U3D_NUM_STAGES]);
/* Work counters summed over the trainer's steps since creation: steps, Gaussians (per view),
 * pixels, tile intersections K, views, steps whose tile-pair list overflowed its capacity
 * (batched engine; the farthest pairs of that step were dropped), and evaluated (pixel,
 * Gaussian) pairs -- the sum of the contributor counts of every training forward, counted on
 * the device (single-view steps with 16-pixel tiles; 0 otherwise). Synchronises the stream. */
#define U3D_NUM_COUNTERS 7
int u3d_trainer_counters(u3d_trainer *t, int64_t values[U3D_NUM_COUNTERS]);

/* ------------------------------------------------------------------------------ helpers */
/* The device exp (bit-identical to the host expf) evaluated on the host, for tests. */
void u3d_expf_host_array(const float *x, float *y, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* U3D_HIP_H */
