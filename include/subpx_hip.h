/*
 * subpx_hip.h -- C-ABI of the MI355X-native per-feature subpixel refinement path.
 *
 * Every pointer argument marked "device" is a device pointer (e.g. torch.Tensor.data_ptr()
 * of a tensor on the current HIP device); "host" pointers are ordinary process memory.
 * The caller owns every buffer. Calls are ordered on the given stream (NULL = the
 * device's null stream) and return 0 on success or a negative SPX_E* code; the text
 * of the last error on the calling thread is available from spx_last_error().
 * Failed fits never make a call fail: they are reported per feature through a status
 * byte, as the reference reports NaN + ok=False per feature.
 *
 * Image convention (reference subpx/synthetic_arrays.py:22-24): points are [x, y] in
 * pixels, pixel centres at integer coordinates, images row-major (H, W) uint16 with a
 * row pitch in elements.
 */
#ifndef SUBPX_HIP_H
#define SUBPX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* spx_stream_t; /* a hipStream_t; NULL = null stream */

/* ---- error classes -------------------------------------------------------------- */
#define SPX_OK 0
#define SPX_E_ARG (-1)         /* bad argument: shape, size, null pointer, option     */
#define SPX_E_NOMEM (-2)       /* device workspace could not be allocated             */
#define SPX_E_HIP (-3)         /* a HIP runtime call failed                            */
#define SPX_E_UNSUPPORTED (-4) /* valid request the library does not implement        */
#define SPX_E_CAPACITY (-5)    /* more components than the caller's capacity          */

/* ---- per-feature fit status (status_out bytes) ----------------------------------- */
#define SPX_FIT_CONVERGED 0    /* a step below step_tol was reached (and applied)      */
#define SPX_FIT_MAX_ITER 1     /* iteration cap reached; parameters are the best seen  */
#define SPX_FIT_SINGULAR 2     /* the damped normal matrix stayed singular            */
#define SPX_FIT_DIVERGED 3     /* non-finite cost or parameters                       */
#define SPX_FIT_TOO_FEW 4      /* fewer in-image patch pixels than the model needs    */
#define SPX_FIT_OUTSIDE 5      /* the fitted centre left the patch                    */
#define SPX_FIT_BAD_START 6    /* moment start impossible (flat patch)                */

/* ---- fit models (model bytes) ------------------------------------------------------ */
#define SPX_MODEL_ERF_DISK6 0    /* B + A/2 erfc((rho-r)/(sqrt2 s)), rho = |p - c|       */
#define SPX_MODEL_ERF_ELLIPSE7 1 /* as DISK6 with rho = sqrt(dx^2 + (dy/q)^2), q 7th param*/

/* Options of the batched Levenberg-Marquardt erf fits. */
typedef struct SpxFitOptions {
    int32_t max_iter;     /* LM step attempts per feature (30)                         */
    float step_tol;       /* converged when a solved step moves x0,y0,r,s < this (1e-6 px);
                             that step is applied without a trial pass */
    float lambda0;        /* initial Marquardt damping (1e-3)                          */
    float lambda_up;      /* damping factor after a rejected step (10)                 */
    float lambda_down;    /* damping factor after an accepted step (0.1)               */
    float init_sigma;     /* edge sigma of the moment start (1.0 px)                   */
    int32_t tile_waves;   /* waves per workgroup (1..4), 0 = library default (4); the
                             Python layer's autotuner (fit.autotune_launch) picks one   */
    int32_t persistent;   /* 1 = grid of the chip's resident capacity pulling from the
                             work queue, 0 = one wave per feature                        */
    int32_t bucket_mask;  /* bit 2 c + m: launch the (patch class c, model m) kernel;
                             c = 0..4 for at most 4, 8, 16, 32, 64 pixels per lane, from an
                             upper bound of the feature's fitted pixels (its patch's pixels
                             inside the image and, with neighbours, inside its cell) / 64
                             rounded up; m = 1 for ellipses. Features of buckets not
                             launched are not fitted (NaN, status too few). 0 = all ten,
                             so callers that know their patch sizes skip empty launches  */
    int32_t learned_iters; /* > 0: budgeted fit of a learned start (init; without init the moment
                             start, as a control): at most this many step attempts, and a
                             residual judge at the start and after every accepted step accepts
                             the fit when the Gauss-Newton model there predicts the minimum's
                             cost within 1e-6 of the current cost (that step is then applied);
                             a feature not accepted within the budget is fitted again from the
                             moment start (max_iter attempts) by the same wave in the same
                             launch, and iters_out counts the attempts of both.
                             0: plain fit (init, if given, is just the start) */
    int32_t bucket_waves; /* bits 2 b..2 b+1: waves per SIMD the register allocation of bucket
                             b's kernel leaves room for: 0 = library default, 1, 2, 3 = 2, 3, 4
                             waves (more waves hide the FP64 solve's latency, fewer keep the
                             LM state out of scratch); fit.autotune_launch picks them.
                             Results do not depend on it */
} SpxFitOptions;

/* Options of the edge-erf refinement (defaults of the reference's
 * refine_centers_edge_erf_gpu, subpx/_gpu/centers.py:882-900). */
typedef struct SpxEdgeErfOptions {
    int32_t band_rad;       /* 4: half width of the averaged band                        */
    int32_t edge_rad;       /* 10: profile half length (2..15)                            */
    int32_t smooth_passes;  /* 2: [1 4 6 4 1]/16 passes                                   */
    int32_t loc_rad;        /* 3: grad-moment window; the erf polarity uses loc_rad - 1  */
    float grad_power;       /* 8                                                          */
    int32_t iters;          /* 2: outer passes                                            */
    int32_t erf_fit_rad;    /* 5                                                          */
    int32_t erf_gn_iters;   /* 5                                                          */
    float erf_init_sigma;   /* 1.25                                                       */
    float erf_sigma_min;    /* 0.35                                                       */
    float erf_sigma_max;    /* 6                                                          */
    float erf_max_shift;    /* 2.5                                                        */
    float erf_damp;         /* 1e-4                                                       */
    float erf_cond_max;     /* 1e8                                                        */
} SpxEdgeErfOptions;

#define SPX_DTYPE_U16 0
#define SPX_DTYPE_F32 1

/* Options of the window log-quadratic refinement (defaults of the reference's
 * refine_centers_logquad_gpu_match_cpu, subpx/_gpu/centers.py:288-309). */
typedef struct SpxLogquadWindowOptions {
    int32_t win_rad;        /* 4: fit window (2 win_rad + 1)^2                              */
    int32_t bg_win_pad;     /* 2: outer window radius win_rad + bg_win_pad (at most 10)    */
    int32_t bg_dilate;      /* 2: blob dilation before the background median              */
    int32_t fit_dilate;     /* 1: blob dilation of the fit mask                            */
    int32_t min_bg_pts;     /* 12                                                          */
    int32_t min_plane_pts;  /* 20                                                          */
    int32_t plane_bg;       /* 1: plane background, 0: scalar median                       */
    int32_t reserved;       /* 0                                                           */
    double core_frac;       /* 0.2                                                         */
    double weight_power;    /* 1                                                           */
    double bg_tail_frac;    /* 0.05                                                        */
    double max_cond;        /* 1e10                                                        */
    double negdef_eps;      /* 1e-12                                                       */
    double eps;             /* 1e-12                                                       */
} SpxLogquadWindowOptions;

/* Synthetic training tiles (spx_render_tiles). */
typedef struct SpxTileSpec {
    int32_t layout;       /* 0 = square array (8 neighbours), 1 = staggered (6)           */
    int32_t cols;         /* tiles per mosaic row                                         */
    float pitch;          /* neighbour distance, px                                       */
    float jitter;         /* sigma of the centre jitter per axis, px                      */
    float r_lo, r_hi;     /* radius ~ U[r_lo, r_hi]                                       */
    float s_lo, s_hi;     /* edge sigma ~ U[s_lo, s_hi]                                   */
    float ellipse_frac;   /* probability of an ellipse, q ~ U[0.8, 1]                     */
    float amplitude;      /* A                                                            */
    float background;     /* B                                                            */
    float read_sigma;     /* Gaussian read noise, counts                                  */
} SpxTileSpec;

/* ---- lifecycle ---------------------------------------------------------------------- */
/* Select and initialise a HIP device (reference: subpx/backends.py:55-59 gpu_device). */
int spx_init(int device);
/* Free every library-owned workspace on every device (reference: subpx/_gpu/_cuda.py:35-46). */
void spx_release(void);
/* Text of the last error raised on this thread ("" if none). */
const char* spx_last_error(void);
/* Library version string. */
const char* spx_version(void);

/* ---- integer front end ---------------------------------------------------------------- */
/* Otsu threshold of a uint16 image over its 65,536-bin histogram.
 * Replaces cv2.threshold(g, 0, 255, BINARY|OTSU) at subpx/_cpu/centers.py:258 and the
 * global Otsu of the tiled path at subpx/_gpu/centers.py:1789-1796 (step > 1 samples the
 * image as img[::step, ::step]). img: device. thresh_out: host. */
int spx_otsu_u16(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step,
                 int32_t* thresh_out, spx_stream_t stream);

/* Triangle threshold, replacing cv2.threshold(..., THRESH_TRIANGLE) at
 * subpx/_cpu/centers.py:258 (OpenCV defines it for 8-bit images only): img values must be
 * 0..255 (SPX_E_ARG otherwise). thresh_out: host. */
int spx_triangle_u8(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step, int32_t* thresh_out,
                    spx_stream_t stream);

/* Binary mask (invert ? v <= thresh : v > thresh) as uint16 0 / 1, contiguous (H, W); the
 * binarisation of cv2.threshold(..., THRESH_BINARY[_INV]) (subpx/_cpu/centers.py:257-258). */
int spx_threshold_mask(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh, int32_t invert,
                       uint16_t* mask_out, spx_stream_t stream);

/* One erosion (op 0) or dilation (op 1) of a 0 / non-zero uint16 mask by the 3x3 cross,
 * pixels outside the image ignored: one iteration of cv2.morphologyEx with the 3x3
 * MORPH_ELLIPSE element (subpx/_cpu/centers.py:260-265). out: contiguous (H, W), not in. */
int spx_morph_cross(const uint16_t* in, int64_t H, int64_t W, int64_t pitch, int32_t op, uint16_t* out,
                    spx_stream_t stream);

/* Threshold (foreground = invert ? g <= thresh : g > thresh) + connected components with
 * stats, replacing _segment_binary + connected_components_stats_cpu
 * (subpx/_cpu/centers.py:247-279, subpx/_cpu/components.py:18-34) and
 * connected_components_stats_gpu (subpx/_gpu/components.py:64-134).
 * connectivity 4 or 8. Labels: 0 background, 1..n-1 components numbered in the order of
 * their first pixel in a raster scan over 2x2 blocks (8-connectivity) or pixels
 * (4-connectivity). img, labels (H*W int32, row pitch W; NULL: no label image, the stats
 * only), stats ((cap,5) int32 [x, y, w, h, area], row 0 = background), centroids ((cap,2)
 * f64): device.
 * n_out (host) receives the label count including the background; SPX_E_CAPACITY when it
 * exceeds cap (n_out still holds the true count). H and W below 2^24. */
int spx_threshold_cc(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh,
                     int32_t invert, int32_t connectivity, int32_t* labels, int32_t* stats,
                     double* centroids, int32_t cap, int32_t* n_out, spx_stream_t stream);

/* ---- per-feature refinement ------------------------------------------------------------- */
/* CPU log-quadratic refinement restated per label (subpx/_cpu/centers.py:67-107 with the
 * ROI rule of :376-391 and the border median of :31-42). rows: device (K,5) int32
 * [x, y, w, h, label] (bbox from stats); pad: ROI padding. centers_out (K,2) f64 in image
 * coordinates, ok_out (K) uint8: device. */
int spx_refine_logquad(const uint16_t* img, int64_t H, int64_t W, int64_t pitch,
                       const int32_t* labels, const int32_t* rows, int32_t K, int32_t pad,
                       double* centers_out, uint8_t* ok_out, spx_stream_t stream);

/* Window log-quadratic refinement, the GPU backend's "logquad" method, replacing
 * refine_centers_logquad_gpu_match_cpu (subpx/_gpu/centers.py:288-517): plane background
 * from a lower median outside the dilated blob, weighted log-quadratic fit in a
 * (2 win_rad + 1)^2 window at rint(centroid). img (H, W) uint16 / float32 (dtype) with a row
 * pitch in elements; labels (H, W) int32 contiguous; rows (K, 6) f32 [x, y, w, h, cx, cy];
 * lab_ids (K) int32; centers_out (K, 2) f64 (the centroid where refused); ok_out (K) uint8:
 * device. opt: host. */
int spx_refine_logquad_window(const void* img, int32_t dtype, int64_t H, int64_t W, int64_t pitch,
                              const int32_t* labels, const float* rows, const int32_t* lab_ids, int32_t K,
                              int32_t pad, const SpxLogquadWindowOptions* opt, double* centers_out,
                              uint8_t* ok_out, spx_stream_t stream);

/* Weighted centroid of the GPU detect path's "weighted" and "none" methods
 * (subpx/_gpu/centers.py:1687-1710 -> _refine_batch_gpu :931-955 -> _weighted_from_batch_gpu
 * :43-66, border median of subpx/_cpu/centers.py:31-42). Same arguments as
 * spx_refine_logquad; NaN and ok 0 where the centroid leaves the ROI. */
int spx_refine_weighted(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                        const int32_t* rows, int32_t K, int32_t pad, double* centers_out, uint8_t* ok_out,
                        spx_stream_t stream);

/* Edge gradient-moment centre refinement, replacing _refine_edge_moment_core /
 * refine_centers_edge_moment_gpu (subpx/_gpu/centers.py:547-636): the profiles and seeds of
 * spx_edge_erf without the erf fits (band_rad, edge_rad, smooth_passes, loc_rad, grad_power
 * and iters of opt are used). Same arguments as spx_edge_erf. */
int spx_edge_moment(const void* img, int32_t dtype, int64_t H, int64_t W, int64_t pitch, const float* rows,
                    int32_t K, const SpxEdgeErfOptions* opt, double* centers_out, uint8_t* ok_out,
                    spx_stream_t stream);

/* Edge-erf centre refinement, replacing _refine_edge_erf_core / refine_centers_edge_erf_gpu
 * (subpx/_gpu/centers.py:782-929): left/right/top/bottom edge profiles (band means, binomial
 * smoothing), gradient-moment seeds, 5-parameter erf-edge Gauss-Newton fits, centre =
 * midpoints, two passes. img (H, W) uint16 (dtype SPX_DTYPE_U16) or float32 (SPX_DTYPE_F32)
 * with a row pitch in elements; rows (K, 6) f32 [x, y, w, h, cx, cy]; centers_out (K, 2) f64;
 * ok_out (K) uint8: device. opt: host. */
int spx_edge_erf(const void* img, int32_t dtype, int64_t H, int64_t W, int64_t pitch, const float* rows,
                 int32_t K, const SpxEdgeErfOptions* opt, double* centers_out, uint8_t* ok_out,
                 spx_stream_t stream);

/* One erf-edge fit per profile (the reference's _edge_from_profile_erf_gpu,
 * subpx/_gpu/centers.py:665-779; polarity window opt->loc_rad): prof (K, L) f64, base (K)
 * f64 (position of sample 0), seed (K) f64; x_out (K) f64 (NaN when refused), ok_out (K)
 * uint8. All device. */
int spx_edge_profile_erf(const double* prof, int32_t K, int32_t L, const double* base, const double* seed,
                         const SpxEdgeErfOptions* opt, double* x_out, uint8_t* ok_out, spx_stream_t stream);

/* Batched erf-disk / erf-ellipse Levenberg-Marquardt fit of one patch per feature (no
 * reference function: the model comes from the task definition; the wave-per-feature LM
 * follows the reference's lm_fit kernel, subpx/_gpu/deblend.py:55-152, and _fit_disks,
 * subpx/_cpu/invariant_cd.py:165-246). Patch = (2*half+1)^2 pixels centred on
 * rint(seed); pixels outside the image are left out. seeds (K,2) f32 [x, y], half (K)
 * int16 in [2, 31], model (K) uint8 SPX_MODEL_*, init (K,7) f32 or NULL (moment start),
 * params_out (K,7) f32 [x0, y0, r, sigma, A, B, q] (q = 1 for disks), centers_out (K,2)
 * f64 [x0, y0] or NULL (the same centres unrounded: an f32 coordinate of a 16384-px page
 * rounds by up to 1e-3 px, of a 32768-px page by up to 2e-3 px), rms_out (K) f32,
 * iters_out (K) int16, status_out (K) uint8: device. opt: host. */
int spx_fit_batch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                  const int16_t* half, const uint8_t* model, const float* init,
                  const SpxFitOptions* opt, int32_t K, float* params_out, double* centers_out,
                  float* rms_out, int16_t* iters_out, uint8_t* status_out, spx_stream_t stream);

/* spx_fit_batch with a neighbour (Voronoi cell) mask: a patch pixel takes part only when it
 * is nearer to the feature's own seed than to each of its listed neighbour seeds (ties to
 * the lower feature index), and the moment start's background ring adds the cell's
 * boundary pixels. neighbours: device (K, nb) int32 feature ids (ids outside [0, K), e.g. -1,
 * mean none), nb in [1, 8];
 * null = no mask (spx_fit_batch). No reference counterpart: the neighbour policy of
 * patches larger than the hole pitch, after the reference's Voronoi cells
 * (subpx/_cpu/cells.py:8-84). */
int spx_fit_batch_cells(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                        const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions* opt,
                        int32_t K, const int32_t* neighbours, int32_t nb, float* params_out,
                        double* centers_out, float* rms_out, int16_t* iters_out, uint8_t* status_out,
                        spx_stream_t stream);

/* A batched fit recorded as one HIP graph: the bucketing kernels, every bucket's persistent
 * kernel on the four forked streams and the joins, replayed by one hipGraphLaunch. No
 * reference counterpart: the reference launches each CuPy kernel from Python per iteration
 * (subpx/_gpu/deblend.py:243-268 runs lm_fit once per batch of cells); a caller that fits
 * batch after batch of the same size with the same buffers (a page loop writing new seeds,
 * patch sizes or image contents into them) records the batch once and replays it.
 * spx_fit_graph_create takes spx_fit_batch_cells's arguments (neighbours may be NULL)
 * except the stream, records without running anything, and binds the graph to these
 * buffers: a launch reads and writes them as they are when the graph runs. The graph owns
 * its scratch, counters, capture and helper streams, so its launches never touch the
 * workspace of other calls; launches of one graph must not overlap (on one stream they are
 * ordered). Results equal spx_fit_batch_cells's on the same inputs; spx_fit_stats after a
 * launch reads the graph's counters. K >= 1. */
typedef struct SpxFitGraph* spx_fit_graph_t;
int spx_fit_graph_create(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                         const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions* opt,
                         int32_t K, const int32_t* neighbours, int32_t nb, float* params_out, double* centers_out,
                         float* rms_out, int16_t* iters_out, uint8_t* status_out, spx_fit_graph_t* graph_out);
int spx_fit_graph_launch(spx_fit_graph_t graph, spx_stream_t stream);
/* Waits for the graph's launches, then frees it (NULL is ignored). */
int spx_fit_graph_destroy(spx_fit_graph_t graph);

/* Largest batch of spx_fit_batch_noalloc. */
#define SPX_FIT_NOALLOC_MAX_FEATURES (1 << 20)

/* spx_fit_batch_cells (neighbours may be NULL) that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Its bucket counters, work queues,
 * feature lists and work counters live in a fixed device array of 4 MiB + 512 B per GPU, held
 * until the process ends, so K is at most SPX_FIT_NOALLOC_MAX_FEATURES (SPX_E_ARG beyond).
 * Every call on a GPU shares that array and clears what must start at zero: calls that
 * overlap on one GPU (other streams or threads) race and must be serialised by the caller.
 * Results equal spx_fit_batch_cells's; spx_fit_stats reads the array's counters. */
int spx_fit_batch_noalloc(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                          const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions* opt,
                          int32_t K, const int32_t* neighbours, int32_t nb, float* params_out, double* centers_out,
                          float* rms_out, int16_t* iters_out, uint8_t* status_out, spx_stream_t stream);

/* The nb nearest other seeds of every seed within `radius` px (ordered by distance, ties by
 * index; -1 where fewer exist), by a uniform-grid hash on the device. seeds (K, 2) f32 in an
 * H x W image, out (K, nb) int32: device. nb in [1, 8]. A seed with a non-finite coordinate
 * takes no part (it is nobody's neighbour and its own row is all -1). */
int spx_seed_neighbours(const float* seeds, int32_t K, int64_t H, int64_t W, float radius, int32_t nb,
                        int32_t* out, spx_stream_t stream);

/* Work counters of the last spx_fit_batch call issued on this thread (no reference
 * counterpart; used to state algorithmic FLOP rates): out[0] = sum over features of pixel
 * passes x patch pixels for 6-parameter fits, out[1] the same for 7-parameter fits,
 * out[2] = pixel passes, out[3] = LM step attempts. out: host uint64[4]. Waits for the
 * stream of that fit, then copies on `stream` and synchronises it. */
int spx_fit_stats(uint64_t* out, spx_stream_t stream);

/* The first n (1..18) counters of the last fit issued on this thread: those of spx_fit_stats,
 * then out[4] = features whose learned start did not converge within opt->learned_iters and
 * were fitted again from the moment start; out[5] = features with more fitted pixels than
 * their bucket's bound (0 unless the bound is wrong; their extra pixels are left out);
 * out[6..7] = 0; out[8 + b] = features in bucket b = 2 c + m (see bucket_mask), b = 0..9.
 * out: host uint64[n]. */
int spx_fit_counters(uint64_t* out, int32_t n, spx_stream_t stream);

/* ---- learned initial guess (no reference counterpart; task-defined CNN) ------------------ */
/* Parameter count of the CNN: conv3x3 1->16->32->64->64 (stride 2) + fc 256->128->7, each
 * layer a (9 Cin or fan-in) x Cout weight matrix followed by Cout biases, FP32. */
int64_t spx_cnn_param_count(void);

/* C (M x N, row-major FP32) = A B with A(m, k) = A[m sam + k sak] and B(k, n) = B[k sbk +
 * n sbn] BF16, FP32 accumulation on the matrix cores (the GEMM every CNN layer uses).
 * A, B, C: device. */
int spx_gemm_bf16(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                  float* C, int32_t M, int32_t N, int32_t K, spx_stream_t stream);

/* Network input of each feature: its (2 half + 1)^2 fit patch resampled to 32 x 32 and
 * normalised by the ring mean B0 and A0 = max - B0. x_out (K, 32, 32) BF16, norm_out (K, 5)
 * f32 {cx, cy, side, B0, A0}: device. */
int spx_cnn_prepare(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                    const int16_t* half, int32_t K, uint16_t* x_out, float* norm_out, spx_stream_t stream);

/* Fit parameters (K, 7) -> normalised network targets (K, 7) for the given norms. Device. */
int spx_cnn_encode(const float* params, const float* norm, int32_t K, float* t_out, spx_stream_t stream);

/* The learned start in one launch: spx_cnn_prepare + spx_cnn_forward with decoding, the
 * patch resampled straight into LDS (same values as the two-step path). img (H, pitch)
 * uint16, seeds (K, 2) f32, half (K) int16, params_out (K, 7) f32 [x0, y0, r, s, A, B, q]:
 * device. No reference counterpart (the learned start is task-defined). */
int spx_cnn_propose(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds, const int16_t* half,
                    int32_t K, const float* theta, const uint16_t* theta_bf, float* params_out, spx_stream_t stream);

/* Forward pass. theta (FP32, biases read from it) and theta_bf (BF16 copy) hold the
 * parameters; out (K, 7): decoded fit parameters [x0, y0, r, s, A, B, q] when norm is given,
 * raw outputs when norm is NULL. Device. */
int spx_cnn_forward(const uint16_t* x, int32_t K, const float* theta, const uint16_t* theta_bf,
                    const float* norm, float* out, spx_stream_t stream);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8; step counts from 1) on a batch of B inputs
 * against targets (B, 7) with optional per-output weights (B, 7): mean squared error,
 * backward through every layer, update of theta, theta_bf, adam_m, adam_v (all device).
 * loss_out: host float or NULL (NULL keeps the call asynchronous). */
int spx_cnn_train_step(const uint16_t* x, const float* target, const float* weight, int32_t B, float* theta,
                       uint16_t* theta_bf, float* adam_m, float* adam_v, int32_t step, float lr, float* loss_out,
                       spx_stream_t stream);

/* Loss gradient of a batch without updating (the gradient spx_cnn_train_step applies):
 * grad_out (spx_cnn_param_count() floats, device), loss_out host float or NULL. */
int spx_cnn_grad(const uint16_t* x, const float* target, const float* weight, int32_t B, float* theta,
                 uint16_t* theta_bf, float* grad_out, float* loss_out, spx_stream_t stream);

/* Whole synthetic page: holes (n, 6) f32 [x, y, r, s, q, A] in image coordinates (device),
 * each I += A erfc((rho - r) / (sqrt2 s)) / 2 with rho = sqrt(dx^2 + (dy / q)^2), on a constant
 * background, then Poisson(I) photon noise + N(0, read_sigma^2), rounded and
 * clipped to uint16 (device image, row pitch in elements). No reference counterpart: the
 * generator of the large benchmark pages, rendered on the device. */
int spx_render_page(uint64_t seed, const float* holes, int32_t n, int64_t H, int64_t W, float background,
                    float read_sigma, uint16_t* img, int64_t pitch, spx_stream_t stream);

/* Render n synthetic tiles of side `tile` into a mosaic (tile t at column t % cols, row
 * t / cols): one hole at the tile centre plus its array neighbours, erf-disk / ellipse
 * model, Poisson photon noise + read noise, uint16. truth_out (n, 7)
 * [x0, y0, r, s, A, B, q] of the centre hole in mosaic coordinates. img, truth: device.
 * No reference counterpart (training data for the learned start). */
int spx_render_tiles(uint64_t seed, int32_t n, int32_t tile, const SpxTileSpec* spec, uint16_t* img,
                     int64_t pitch, float* truth_out, spx_stream_t stream);

/* theta_bf = BF16(theta), n values, device. */
int spx_cnn_to_bf16(const float* theta, int64_t n, uint16_t* theta_bf, spx_stream_t stream);

/* ---- Outline hole stage (reference subpx/_gpu/outline_onepass.py:252-427, the GPU form of
 * PageFit.fit_holes, subpx/_outline/page.py:587-729) ----
 * Page model P(p) = sum_i a_i [T0 + d_i T1 + d_i^2 T2](p - c_i) + far light, FP64 (H, W)
 * contiguous; hole i fits data(p) = P(p) + level_i on its (2R+1)^2 window. Hole state arrays
 * c (n, 2), a, level, delta (n) are FP64 device arrays indexed by hole. */

/* Template grid of an erf-edged disk 1/2 erfc((rho - r - d) / (sqrt2 s)): N x N fine points,
 * U per pixel, centre at N/2, 8 floats per point [T0, T1, T2, Tx, Ty, 0, 0, 0] with
 * T1 = dT/dd, T2 = (1/2) d2T/dd2 and Tx, Ty the spatial derivatives of T0 w, each times the
 * near window w(rho) = (1 + cos(pi clip((rho - r0) / ramp, 0, 1))) / 2. far (far_box^2, or
 * NULL) = T0 (1 - w) at whole-pixel offsets, centre at far_box / 2. Stands in for the
 * outline templates of set_template (outline_onepass.py:730-793) with a disk outline. */
int spx_hole_templates(double radius, double sigma, int32_t N, int32_t U, double r0, double ramp,
                       int32_t far_box, float* tmpl, float* far, spx_stream_t stream);

typedef struct SpxHoleStepOptions {
    double nv0, nv1;   /* pixel weight 1 / (nv0 + nv1 max(model, 0)) */
    double rscale;     /* Cauchy weight 1 / (1 + (r / rscale)^2); 0 = off */
    double max_move;   /* centre step clip, px */
    double max_delta;  /* edge offset bound, px */
    double a_med;      /* brightness bound 5 a_med */
    double relax;      /* factor on the brightness, level and edge-offset steps */
    int32_t size;      /* 1: the edge offset is fitted (5 parameters), 0: held (4) */
    int32_t reserved;
} SpxHoleStepOptions;

/* One Gauss-Newton step of the holes idx[0..n) (one colour class; windows must not overlap
 * another hole of the class being written) against the page model: normal equations of
 * (x, y, a, level[, delta]) with noise x Cauchy weights, 1e-9 trace ridge, bounded step
 * (outline_onepass.py:252-363: hole_normal + hole_update). image FP32 (H, pitch). Writes
 * c_old (n, 2) = centres before the step and stat (n, 4) = [centre move, edge-offset move,
 * leftover rms, edge-offset standard error]. */
int spx_hole_step(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page,
                  const float* tmpl, int32_t N, int32_t U, int32_t R, const int32_t* idx, int32_t n,
                  double* c, double* a, double* level, double* delta, double* c_old,
                  const SpxHoleStepOptions* opt, double* stat, spx_stream_t stream);

/* Take the holes' cached light (cache (n, (2Rr+1)^2) FP32, at offsets around rint(c_old))
 * out of the page model and put their light at the current state in; offsets with
 * ox^2 + oy^2 > rl2 are skipped (outline_onepass.py:365-397). The changes meet in 64-bit
 * fixed-point integers (2^-32 counts) and each touched pixel takes its total in one FP64
 * addition, so the page is the same bit for bit whatever order the threads run in. */
int spx_hole_render(double* page, int64_t H, int64_t W, float* cache, const float* tmpl, int32_t N,
                    int32_t U, int32_t Rr, double rl2, const int32_t* idx, int32_t n, const double* c_old,
                    const double* c, const double* a, const double* delta, spx_stream_t stream);

/* One Gauss-Seidel sweep over the colour classes: for j = 0 .. ncolour-1, spx_hole_step then
 * spx_hole_render of the holes order[start[j] .. start[j+1]) (start: host array of ncolour + 1
 * offsets; order, c_old (n, 2), cache (n, (2Rr+1)^2) and stat (n, 4): device, in colour order).
 * The reference's per-colour loop of fit_holes (outline_onepass.py:855-870) in one call. */
int spx_hole_sweep(const float* image, int64_t H, int64_t W, int64_t pitch, double* page, const float* tmpl,
                   int32_t N, int32_t U, int32_t R, int32_t Rr, double rl2, const int32_t* order,
                   const int32_t* start, int32_t ncolour, double* c, double* a, double* level, double* delta,
                   double* c_old, float* cache, const SpxHoleStepOptions* opt, double* stat, spx_stream_t stream);

/* Order statistics: out[j] = the ks[j]-th smallest (0-based) of x[cols[j] + i * stride], i < n,
 * for j < nk <= 8 (cols, ks: host arrays, cols may be NULL for 0; x, out: device; no NaNs).
 * Exact selection by an 8-round radix select over the doubles' order-preserving 64-bit keys,
 * one block per rank. Serves the hole pass's median leftover and 99th-percentile moves
 * (cp.median / cp.percentile in outline_onepass.py:871-873). */
int spx_select_kth(const double* x, int64_t n, int64_t stride, const int64_t* cols, const int64_t* ks, int32_t nk,
                   double* out, spx_stream_t stream);

/* Keys splat of a_i (1 + 4 delta_i / cd0) at (c_i + (mx, my)) added to the padded FP32 grid
 * (hp, wp) (outline_onepass.py:399-427); the far light is this grid convolved with the far
 * kernel. The taps meet in 64-bit fixed-point sums (order-independent). */
int spx_far_splat(const double* c, const double* a, const double* delta, double cd0, int32_t n, int32_t mx,
                  int32_t my, int32_t hp, int32_t wp, float* grid, spx_stream_t stream);

/* ---- page outline fit (fit_outline: subpx/outline.py:34-170, subpx/_outline/page.py) ----
 * The template T = pixel box * blur * coverage(outline) lives on an N x N fine grid, U points
 * per pixel; the page model and the hole state are those of the hole stage above. */

/* Pixel coverage of nbatch outlines (model.py:149-159): out (nbatch, N, N) FP64, out[b][iy][ix] =
 * mean over the aa x aa sub-samples s of clamp(0.5 + (edge[b][ang[s]] - rho[s]) * scale, 0, 1),
 * with the sub-samples' radius rho and angle-table index ang (N aa x N aa, row-major)
 * precomputed and edge[b] (nbatch x 8192) outline b's radius on the angle table (scale = U aa:
 * a one-sample ramp). */
int spx_outline_coverage(const double* edge, const double* rho, const int32_t* ang, int32_t N, int32_t aa,
                         double scale, int32_t nbatch, double* out, spx_stream_t stream);

/* Per-window gather of the holes idx[0..n) (page.py:206-216, 294-299, 420-540, 732-766): for
 * window pixel q of hole i (row-major over (2R+1)^2 offsets around rint(c)), at t = i P + q:
 * valid[t] (inside the frame), data[t] (grey level at the clamped pixel), model[t] = page +
 * level (page / level NULL: 0), v0[t] = T0 and own[t] = a (T0 + d T1 + d^2 T2) sampled from
 * tmpl (the hole stage's 8-float layout) at the pixel's fine offset, bin[t] = rint(sy) N +
 * rint(sx) of that offset when valid, inside the grid and (radius_fine > 0) within radius_fine
 * fine units of the centre, else -1. Any output may be NULL. */
int spx_outline_windows(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page,
                        const float* tmpl, int32_t N, int32_t U, int32_t R, const int32_t* idx, int32_t n,
                        const double* c, const double* a, const double* level, const double* delta,
                        double radius_fine, int32_t* bin, uint8_t* valid, double* data, double* model, double* own,
                        double* v0, spx_stream_t stream);

/* Stacked average hole on the fine grid (page.py:488-505 plain, 732-766 neighbour-cleaned): every
 * window pixel of the holes idx[0..n) with good[h] != 0 (good NULL: all) inside the frame adds to
 * bin rint(sy) N + rint(sx) of its fine offset (inside the grid; within radius_fine fine units of
 * the centre when radius_fine > 0): cleaned = 0: sums += grey level, wsum += 1; cleaned = 1:
 * sums += w (data - page - level + own) / a, wsum += w with w = a^2 / (nv0 + nv1 max(page +
 * level, 0)) and own = a (T0 + d T1 + d^2 T2). sums, wsum (N x N FP64) accumulate (zero them
 * first); the pixels' terms are summed as 64-bit fixed-point integers (2^-24 per unit), so the
 * sums do not depend on the order of the atomics. */
int spx_outline_stack(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page, const float* tmpl,
                      int32_t N, int32_t U, int32_t R, const int32_t* idx, int32_t n, const uint8_t* good,
                      const double* c, const double* a, const double* level, const double* delta, double nv0,
                      double nv1, double radius_fine, int32_t cleaned, double* sums, double* wsum,
                      spx_stream_t stream);

/* Keys splat of a_i, a_i d_i and a_i d_i^2 at c_i + (mx, my) added to three padded FP64 grids
 * (grid + k hp wp, k = 0, 1, 2; holes whose taps leave the grid are dropped): the far light is
 * the sum of the three convolved with the far kernels of T0, T1 and T2 (page.py:341-397). The
 * taps meet in 64-bit fixed-point sums (order-independent). */
int spx_outline_far_splat(const double* c, const double* a, const double* delta, int32_t n, int32_t mx,
                          int32_t my, int32_t hp, int32_t wp, double* grid, spx_stream_t stream);

/* Voronoi-cell refinements of the GPU detect path (refine="radial_symmetry" /
 * "isophote_curvature", reference subpx/_gpu/centers.py:1545-1630).
 * spx_voronoi_labels replaces compute_voronoi_labels_gpu (subpx/_gpu/voronoi.py:116-145):
 * labels (H, W) int32 (device, or NULL) = index of the nearest of the n seeds (device (n, 2)
 * f64 [x, y]) by FP64 squared distance, exact ties to the lowest index, -1 everywhere without
 * seeds; areas (n) int32 (device, or NULL) = each seed's pixel count; dist (H, W) f64 (device,
 * or NULL; needs n >= 1) = the distance to the nearest seed (with seeds at pixel centres: SciPy's
 * distance_transform_edt of the seed-free pixels, as the deblending's free-pixel rule takes it,
 * reference subpx/_cpu/deblend.py:445-451). */
int spx_voronoi_labels(const double* seeds, int32_t n, int64_t H, int64_t W, int32_t* labels, int32_t* areas,
                       double* dist, spx_stream_t stream);
/* Replaces _extract_voronoi_rois (subpx/_cpu/cells.py:34-84) on the device: for component row j
 * of rows (device (N, 6) f64 [x, y, w, h, cx, cy]) the ROI bbox +- pad clipped to the image, the
 * cell = ROI pixels whose label is j, fill = 10th percentile (NumPy linear) of the cell's border
 * pixels (on the ROI frame or with a 4-neighbour outside the cell); rois (N, Hm, Wm) f32 = grey
 * levels in the cell, the fill elsewhere (padding too), masks (N, Hm, Wm) u8 = the cell,
 * origins (N, 2) int32 [x0, y0]; Hm, Wm = the largest ROI's size (the caller's). ROI stacks
 * beyond about 12,000 px per ROI: SPX_E_UNSUPPORTED. */
int spx_voronoi_rois(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                     const double* rows, int32_t N, int32_t pad, int32_t Hm, int32_t Wm, float* rois, uint8_t* masks,
                     int32_t* origins, spx_stream_t stream);
/* Replace _radial_symmetry_batch_gpu / _isophote_curvature_batch_gpu (subpx/_gpu/centers.py:
 * 978-1230, 1233-1469): centres of the N ROIs of a stack (device rois (N, Hm, Wm) f32, masks u8)
 * upsampled by upsample_factor (Keys' cubic a = -0.5, edge-aligned; masks nearest), masks eroded
 * by boundary_margin (capped as the reference caps it); centers (N, 2) f64 ROI px [x, y] (NaN
 * outside the ROI), quality (N) f64: the radial-symmetry residual / the isophote votes' spread
 * (inf outside). */
int spx_radial_symmetry(const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                        int32_t upsample_factor, int32_t boundary_margin, double* centers, double* quality,
                        spx_stream_t stream);
int spx_isophote_curvature(const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                           int32_t upsample_factor, int32_t boundary_margin, double pre_smooth_sigma,
                           double* centers, double* quality, spx_stream_t stream);

/* Deblending behind detect_centers(background="model") (reference subpx/centers.py:180-316;
 * subpx/_cpu/deblend.py:48-522, whose GPU twin is subpx/_gpu/deblend.py): the image as a
 * background map plus one 2-D Gaussian per hole [A, x0, y0, u1, u2, u3] (precision L L^T,
 * L = [[e^u1, 0], [u2, e^u3]]), alternated with neighbour-cleaned refits. All pointers device;
 * cells are a (N, Hm, Wm) stack with u8 masks and int32 (N, 2) origins [x0, y0]; images
 * (H, W) f64. Host code (imageprocessor_new_amd/deblend.py) runs the iteration, the status
 * rules and the tile interpolation.
 * spx_deblend_init (:152-171, 467-476): data = image window - the cell border's 10th-percentile
 * grey level on the cell (0 elsewhere) and the start params (N, 6) (peak, seed, core moment).
 * spx_gauss_fit (:174-254): one LM fit per cell of data on its mask from init (max_iter steps,
 * damping lam diag + 1e-12, amplitude in [0, 3 max |data|], log-widths in [-log cap, -log 0.5]).
 * spx_gauss_render (:74-107): model = sum of the Gaussians (global centres) within radius[j] of
 * their centre in x and y (64-bit fixed-point sums, order-independent).
 * spx_deblend_clean (:136-149): window(image) - window(background) - (window(model) - own) on
 * the cell, 0 elsewhere.
 * spx_tile_medians (:280-291; subpx/_gpu/deblend.py:364-404): per tile x tile block the count
 * and median (mean of the middle pair) of image - model over pixels with model < model_below
 * and dist >= dist_atleast; NaN without pixels. counts int64.
 * spx_tile_map (:294-320): the map from the (ny, nx) medians: np.interp along x between the
 * tile centres xc (nx), then linear between the tile rows' centres yc (ny), clamped.
 * spx_cell_refine (subpx/_gpu/centers.py:33-111): every cleaned cell's centre (ROI px) by the
 * weighted centroid (method 0) or the log-quadratic fit (1) above its 2-px border median, on
 * the mask, or with amplitude (N) non-NULL on the mask's pixels above frac x peak (peak = the
 * amplitude when finite and positive, else the cell's maximum; subpx/centers.py:255-260).
 * spx_deblend_window (subpx/centers.py:241-252): cleaned cells inside reach + pad of their
 * seed (ROI px), as an f32 stack and u8 mask for the Voronoi estimators. */
int spx_deblend_init(const double* image, int64_t H, int64_t W, const uint8_t* masks, const int32_t* origins,
                     const double* seeds, int32_t N, int32_t Hm, int32_t Wm, double* data, double* params,
                     spx_stream_t stream);
int spx_gauss_fit(const double* data, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm, const double* init,
                  int32_t max_iter, double* params_out, spx_stream_t stream);
int spx_gauss_render(const double* params_global, const double* radius, int32_t N, int64_t H, int64_t W,
                     double* model, spx_stream_t stream);
int spx_deblend_clean(const double* image, const double* background, const double* model,
                      const double* params_global, const double* radius, const uint8_t* masks, const int32_t* origins,
                      int32_t N, int32_t Hm, int32_t Wm, int64_t H, int64_t W, double* out, spx_stream_t stream);
int spx_tile_medians(const double* image, const double* model, const double* dist, int64_t H, int64_t W, int32_t tile,
                     double model_below, double dist_atleast, double* medians, int64_t* counts, spx_stream_t stream);
int spx_tile_map(const double* medians, const double* yc, const double* xc, int32_t ny, int32_t nx, int64_t H,
                 int64_t W, double* map, spx_stream_t stream);
int spx_cell_refine(const double* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm, int32_t method,
                    const double* amplitude, double frac, double* centers, spx_stream_t stream);
int spx_deblend_window(const double* cleaned, const uint8_t* masks, const double* seeds_local, const double* reach,
                       int32_t N, int32_t Hm, int32_t Wm, int32_t pad, float* rois, uint8_t* masks_out,
                       spx_stream_t stream);

/* On-disk pages. The reference reads stripe pages with tifffile (subpx/mosaic.py:51-79,
 * iter_pages_across_vertical / iter_pages_canonical: every page of every file, in file order,
 * page.asarray()); these entry points replace those reads. spx_tiff_open parses a TIFF
 * (classic or BigTIFF, either byte order) and lists its pages; a page must hold one grey sample
 * per pixel, 8- or 16-bit unsigned, in strips or tiles, uncompressed, PackBits or Deflate, with
 * or without the horizontal-differencing predictor (else SPX_E_UNSUPPORTED naming the tag).
 * spx_tiff_read_page decodes page `page` into host memory dst (rows dst_pitch_bytes apart;
 * pinned memory lets the caller's upload to the device overlap the next read): the stored
 * samples, native byte order, no photometric inversion (as tifffile's asarray). Host only; a
 * handle is used by one thread at a time. */
typedef struct SpxTiff* spx_tiff_t;
int spx_tiff_open(const char* path, spx_tiff_t* out);
int spx_tiff_page_count(spx_tiff_t tiff, int32_t* n_out);
int spx_tiff_page_info(spx_tiff_t tiff, int32_t page, int64_t* H, int64_t* W, int32_t* bits);
int spx_tiff_read_page(spx_tiff_t tiff, int32_t page, void* dst, int64_t dst_pitch_bytes);
int spx_tiff_close(spx_tiff_t tiff);

/* ---- One-pass outline fit (fit_outline_fast: subpx/outline.py:361-493; the device work of
 * subpx/_gpu/outline_onepass.py and subpx/_outline/onepass.py). Lengths in px, FP64 unless
 * stated; complex arrays hold (re, im) pairs. The hole pass reuses spx_hole_sweep,
 * spx_hole_render and spx_far_splat above with this fit's templates. ---- */

/* Separable smoothing of image (H, W) FP32 with the symmetric taps taps[0..radius] and clamped
 * edges, down the columns first (into tmp) then along the rows (into out); each output is
 * taps[0] x + sum_{k=radius..1} taps[k] (x[-k] + x[+k]) (find_seeds' gaussian_filter,
 * outline_onepass.py:474). */
int spx_onepass_smooth(const float* image, int64_t H, int64_t W, const double* taps, int32_t radius, double* tmp,
                       double* out, spx_stream_t stream);
/* Local maxima of sm (H, W): flags (H, W) = not on the frame, > thr and equal to the maximum of
 * the (2 nms + 1)^2 window with clamped edges; row_counts (H) = peaks per row
 * (outline_onepass.py:475-480). */
int spx_onepass_peaks(const double* sm, int64_t H, int64_t W, int32_t nms, double thr, uint8_t* flags,
                      int32_t* row_counts, spx_stream_t stream);
/* The peaks in row-major order from row_offsets (H, exclusive prefix of the row counts):
 * seeds (n, 2) [x, y], each refined by a parabola through log(max(v - base, 1e-6)) of the three
 * samples along each axis, the vertex clipped to +-0.5 px (outline_onepass.py:490-501). */
int spx_onepass_peak_list(const double* sm, const uint8_t* flags, const int32_t* row_offsets, int64_t H, int64_t W,
                          double base, double* seeds, spx_stream_t stream);
/* count (device u64) = pixels of image[::step, ::step] above thr (outline_onepass.py:473). */
int spx_onepass_count_above(const float* image, int64_t H, int64_t W, int32_t step, double thr,
                            unsigned long long* count, spx_stream_t stream);
/* Distance from every hole c (n, 2) to its nearest other hole among the 3 x 3 cells round its
 * own (cells of side cs, ncx x ncy; order / cell_start: the holes sorted by cell and the cells'
 * first positions, ncx ncy + 1 entries); 1e15 when none is there (outline_onepass.py:227-250). */
int spx_onepass_nearest(const double* c, int32_t n, const int32_t* order, const int32_t* cell_start, int32_t ncx,
                        int32_t ncy, double cs, double* out, spx_stream_t stream);
/* Plain stack (outline_onepass.py:81-107, 582-592): every window pixel (half side R) of the
 * holes with use[i] != 0 inside the frame adds its grey level to the bin rint((p - c) U) + Ms / 2
 * of the Ms x Ms grid; stack = mean per bin (NaN where none fell), wsum = pixels per bin. The
 * sums are 64-bit fixed-point integers (2^24 per grey level): the same bits every run. */
int spx_onepass_stack(const float* image, int64_t H, int64_t W, const double* c, int32_t n, const uint8_t* use,
                      int32_t R, int32_t U, int32_t Ms, double* stack, double* wsum, spx_stream_t stream);
/* Every hole against the stack (outline_onepass.py:112-171): out (n, 6) = brightness and level
 * (2 x 2 least squares over the pixels within sqrt(core2) of the hole), leftover rms over the
 * window, one Gauss-Newton step (dx, dy) of the centre over the core pixels, pixels compared. */
int spx_onepass_versus(const float* image, int64_t H, int64_t W, const double* c, int32_t n, const double* stack,
                       int32_t R, int32_t U, int32_t Ms, double core2, double* out, spx_stream_t stream);
/* Pair maps before the FFT (outline_onepass.py:184-224, 663-687): hist (U^2, Ng, Ng); every hole
 * i with use[i] != 0, of sub-pixel class q = (mod(rint(-c_y U), U), mod(rint(-c_x U), U)) and
 * residual s_i = (-c U - rint(-c U)) / U, adds (a_j / amean) / (stacked holes of class q) at
 * c_j - c_i - s_i (cubic B-spline taps, Ug per px, wrapped) to map q for every hole j of the
 * cells within ceil(reach / cs) of its cell with |c_j - c_i| <= reach per axis (j = i included).
 * Fixed-point sums (2^40 per count). class_counts (U^2, may be NULL): stacked holes per class. */
int spx_onepass_pair_hist(const double* c, const double* a, double amean, const uint8_t* use, int32_t n, int32_t U,
                          const int32_t* order, const int32_t* cell_start, int32_t ncx, int32_t ncy, double cs,
                          double reach, int32_t Ug, int32_t Ng, double* hist, double* class_counts,
                          spx_stream_t stream);
/* gamma (nclass, 2 kmax + 1, kmax + 1) complex = spec[q][ky mod Ng][kx] / (centre x fix) for
 * ky = -kmax .. kmax, kx = 0 .. kmax, f = k / L (outline_onepass.py:688-698); spec (nclass, Ng,
 * Ng / 2 + 1) complex, the maps' real FFT. */
int spx_onepass_pair_crop(const double* spec, int32_t nclass, int32_t Ng, int32_t Ug, int32_t L, int32_t kmax,
                          double* gamma, spx_stream_t stream);
/* The outline's transform along its contour and its shape derivatives (onepass.py:129-162,
 * 281-312): geo (4, rows, n) = [x, y, dx/dphi, dy/dphi], gk (rows, n, K) = r dr/dp_k, area (rows);
 * out (rows, F, 1 + K) complex at the frequencies (fx, fy) (F). K <= 16. */
int spx_onepass_outline_hat(const double* geo, const double* gk, const double* area, int32_t rows, int32_t n,
                            int32_t K, double dphi, const double* fx, const double* fy, int32_t F, double* out,
                            spx_stream_t stream);
/* Band coefficients of the template and (tm != NULL) its derivatives along a stage's free
 * directions (onepass.py:314-355): coeff (B, 1 + nf, F) complex from ohat (B, F, 1 + K), the
 * pages' blur parameters blur (B, nb) (kind 0: two Gaussians, nb 5; 1: Gaussian x pupil, nb 3;
 * 2: one Gaussian, nb 2), tm (K + nb, nf) and the pixel box pix (F). nf <= 16. */
int spx_onepass_coeff(const double* ohat, int32_t K, const double* blur, int32_t kind, int32_t nb, const double* tm,
                      int32_t nf, const double* fx, const double* fy, const double* pix, int32_t B, int32_t F,
                      double* coeff, spx_stream_t stream);
/* The grid of starts' coefficients (onepass.py:453-458): coeff (B, ns nw, F) = O_(b,i) OTF_j pix,
 * ohat (B ns, F, 1) complex, blur (nw, nb). */
int spx_onepass_grid_coeff(const double* ohat, int32_t ns, const double* blur, int32_t nw, int32_t kind, int32_t nb,
                           const double* fx, const double* fy, const double* pix, int32_t B, int32_t F, double* coeff,
                           spx_stream_t stream);
/* The stack's model by class (onepass.py:357-369): M (B, C, slots, npts, npts) = Re Ey[qy] G ExT[qx]
 * with G (nky, nkx) = coeff[b][c][fmap] dbl gamma[b][q] (0 where fmap < 0); slots = the U^2
 * classes (cls_sel NULL) or the page's class cls_sel[b]. Ey (U, npts, nky), ExT (U, nkx, npts),
 * gamma (B, U^2, nky, nkx) complex. */
int spx_onepass_synth(const double* coeff, int32_t B, int32_t C, int32_t F, const int32_t* fmap, const double* dbl,
                      const double* gamma, int32_t U, const int32_t* cls_sel, const double* Ey, const double* ExT,
                      int32_t npts, int32_t nky, int32_t nkx, double* M, spx_stream_t stream);
/* Variable projection per page (onepass.py:371-410): out (B, 3 + nf + nf^2) = [cost, scale,
 * const, g (nf), A (nf x nf)] of M (B, C = 1 + nf, nbins) against y, w (B, nbins); scale and
 * constant solved exactly, the derivatives projected off them (nf = 0: [cost, scale, const]). */
int spx_onepass_project(const double* M, const double* y, const double* w, int32_t B, int32_t C, int32_t nbins,
                        double* out, spx_stream_t stream);
/* The grid of starts' costs (onepass.py:459-470): table (B, C), the cost of model c on the bins of
 * the page's class cls_sel[b] with the best scale and constant; M (B, C, nb), y, w (B, ncls, nb). */
int spx_onepass_grid_costs(const double* M, const double* y, const double* w, const int32_t* cls_sel, int32_t B,
                           int32_t C, int32_t ncls, int32_t nb, double* table, spx_stream_t stream);
/* The page fit's Levenberg-Marquardt iterations on the device (onepass.py:504-623): one stage of
 * B pages with nf free parameters of P (nf <= 16). The caller owns every array (device) and
 * fills the stage constants; per iteration it calls spx_onepass_lm_trial (first = 0), evaluates
 * the trial theta (spx_onepass_lm_geometry, spx_onepass_outline_hat, spx_onepass_coeff with the
 * blur rows, spx_onepass_synth, spx_onepass_project into ev) and calls spx_onepass_lm_update;
 * *alldone = 1 once every page's stage has ended. The stage starts with q set, lm_trial(first
 * = 1), the evaluation and spx_onepass_lm_init. */
typedef struct SpxLmStage {
    int32_t B, P, nf, nshape, nb, npr;
    const double* fixed;  /* (B, P) theta entries no free parameter drives */
    const double* tm;     /* (P, nf) free parameter -> theta entries (ties included) */
    const double* lo;     /* (nf) bounds of the free parameters */
    const double* hi;
    const double* PL;     /* (npr, P) prior rows; pm, pw (npr): means and 1 / sd^2 */
    const double* pm;
    const double* pw;
    const double* pj;     /* (npr, nf) = PL tm */
    const double* pa;     /* (nf, nf) = pj^T diag(pw) pj */
    const double* ev;     /* (B, 3 + nf + nf^2): spx_onepass_project's output for the trial */
    double* q;            /* (B, nf) accepted point */
    double* qn;           /* (B, nf) trial point */
    double* theta;        /* (B, P) trial theta */
    double* blur;         /* (B, nb) trial theta's blur part */
    double* cost;         /* (B) with the priors */
    double* scale;
    double* cst;
    double* g;            /* (B, nf) */
    double* A;            /* (B, nf, nf) */
    double* lam;          /* (B) damping */
    double* nu;           /* (B) damping growth */
    int32_t* done;        /* (B) */
    int32_t* alldone;     /* (1) */
    double ftol, xtol;
} SpxLmStage;
int spx_onepass_lm_trial(const SpxLmStage* stage, int32_t first, spx_stream_t stream);
/* The outline of every trial theta (B, P): geo (4, B, n) = [x, y, dx/dphi, dy/dphi] at n equal
 * angles, gk (B, n, 4 + 2 nk) = r dr/d(shape parameter), area (B); ks (host, nk <= 4) the
 * harmonics (onepass.py:246-279). */
int spx_onepass_lm_geometry(const double* theta, int32_t B, int32_t P, int32_t n, int32_t nk, const int32_t* ks,
                            double* geo, double* gk, double* area, spx_stream_t stream);
int spx_onepass_lm_init(const SpxLmStage* stage, spx_stream_t stream);
int spx_onepass_lm_update(const SpxLmStage* stage, spx_stream_t stream);
/* Analytic templates on a band (onepass.py:52-126): for every polygon row of xy (rows, 2, n) and
 * blur row (rows, nb), the polygon's exact transform x OTF x sinc(fx) sinc(fy) [x sinc(fx box)
 * sinc(fy box)] x exp(-2 pi i (fx + fy) (N / 2) / Uu), f = k / L, |kx|, |ky| <= kmax, kx >= 0,
 * written into full (rows, N, N / 2 + 1) complex (N = L Uu; other entries untouched) for the
 * inverse real FFT. */
int spx_onepass_band(const double* xy, int32_t rows, int32_t n, const double* blur, int32_t kind, int32_t nb,
                     int32_t L, int32_t Uu, int32_t kmax, double box_px, double* full, spx_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SUBPX_HIP_H */
