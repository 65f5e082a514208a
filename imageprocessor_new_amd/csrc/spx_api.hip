// C-ABI entry points of libsubpx_hip.so: lifecycle, error text, workspaces and argument
// checks. Kernels live in the other translation units.
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "spx_internal.hpp"

namespace spx {

namespace {
thread_local std::string g_last_error;

struct DeviceWorkspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
// Workspaces and fit counters by (device ordinal, stream).
using StreamKey = std::pair<int, hipStream_t>;
std::mutex g_ws_mutex;
std::map<StreamKey, DeviceWorkspace> g_ws;
std::map<StreamKey, unsigned long long*> g_fit_stats;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int workspace(size_t bytes, void** out, hipStream_t stream) {
    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    DeviceWorkspace& w = g_ws[StreamKey(dev, stream)];
    if (w.bytes < bytes) {
        if (w.ptr != nullptr) {
            // Work already queued on this stream may still read the old buffer.
            SPX_HIP(hipStreamSynchronize(stream));
            SPX_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        size_t grow = bytes + bytes / 4 + 4096;
        if (hipMalloc(&w.ptr, grow) != hipSuccess) {
            (void)hipGetLastError();
            w.ptr = nullptr;
            set_error("workspace: hipMalloc of " + std::to_string(grow) + " bytes failed");
            return SPX_E_NOMEM;
        }
        w.bytes = grow;
    }
    *out = w.ptr;
    return SPX_OK;
}

int fit_stats_buffer(unsigned long long** out, hipStream_t stream) {
    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    unsigned long long*& p = g_fit_stats[StreamKey(dev, stream)];
    if (p == nullptr) {
        void* m = nullptr;
        if (hipMalloc(&m, kFitCounters * sizeof(unsigned long long)) != hipSuccess) {
            (void)hipGetLastError();
            set_error("fit counters: hipMalloc failed");
            return SPX_E_NOMEM;
        }
        p = static_cast<unsigned long long*>(m);
    }
    *out = p;
    return SPX_OK;
}

int fit_batch_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                     const int16_t* half, const uint8_t* model, const float* init,
                     const SpxFitOptions& opt, int32_t K, const int32_t* nbr, int32_t nb, float* params,
                     double* centers, float* rms, int16_t* iters, uint8_t* status, hipStream_t stream);

int fit_stats_copy(uint64_t* out, int n, hipStream_t stream);
int fit_scratch_init(int dev);
void fit_streams_release();
int fit_batch_noalloc_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                             const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions& opt,
                             int32_t K, const int32_t* nbr, int32_t nb, float* params, double* centers, float* rms,
                             int16_t* iters, uint8_t* status, hipStream_t stream);
int fit_graph_create(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                     const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions& opt,
                     int32_t K, const int32_t* nbr, int32_t nb, float* params, double* centers, float* rms,
                     int16_t* iters, uint8_t* status, SpxFitGraph** out);
int fit_graph_launch(SpxFitGraph* g, hipStream_t stream);
int fit_graph_destroy(SpxFitGraph* g);
int voronoi_labels_launch(const double* seeds, int32_t n, int64_t H, int64_t W, int32_t* labels, int32_t* areas,
                          double* dist, hipStream_t stream);
int deblend_init_launch(const double* image, int64_t H, int64_t W, const uint8_t* masks, const int32_t* origins,
                        const double* seeds, int32_t N, int32_t Hm, int32_t Wm, double* data, double* params,
                        hipStream_t stream);
int gauss_fit_launch(const double* data, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm, const double* init,
                     int32_t max_iter, double* out, hipStream_t stream);
int gauss_render_launch(const double* pg, const double* radius, int32_t N, int64_t H, int64_t W, double* model,
                        hipStream_t stream);
int deblend_clean_launch(const double* image, const double* background, const double* model, const double* pg,
                         const double* radius, const uint8_t* masks, const int32_t* origins, int32_t N, int32_t Hm,
                         int32_t Wm, int64_t H, int64_t W, double* out, hipStream_t stream);
int tile_medians_launch(const double* image, const double* model, const double* dist, int64_t H, int64_t W,
                        int32_t tile, double model_below, double dist_atleast, double* medians, int64_t* counts,
                        hipStream_t stream);
int tile_map_launch(const double* medians, const double* yc, const double* xc, int32_t ny, int32_t nx, int64_t H,
                    int64_t W, double* out, hipStream_t stream);
int cell_refine_launch(const double* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm, int32_t method,
                       const double* amplitude, double frac, double* centers, hipStream_t stream);
int deblend_window_launch(const double* cleaned, const uint8_t* masks, const double* seeds_local,
                          const double* reach, int32_t N, int32_t Hm, int32_t Wm, int32_t pad, float* rois,
                          uint8_t* out_masks, hipStream_t stream);
int voronoi_rois_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                        const double* rows, int32_t N, int32_t pad, int32_t Hm, int32_t Wm, float* rois,
                        uint8_t* masks, int32_t* origins, hipStream_t stream);
int radial_symmetry_launch(const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                           int32_t factor, int32_t margin, double* centers, double* quality, hipStream_t stream);
int isophote_curvature_launch(const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                              int32_t factor, int32_t margin, double sigma, double* centers, double* quality,
                              hipStream_t stream);
int otsu_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step,
                int32_t* thresh_out, hipStream_t stream);
int threshold_cc_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh,
                        int32_t invert, int32_t connectivity, int32_t* labels, int32_t* stats,
                        double* cents, int32_t cap, int32_t* n_out, hipStream_t stream);
int logquad_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                   const int32_t* rows, int32_t K, int32_t pad, double* centers, uint8_t* ok,
                   hipStream_t stream);

int64_t cnn_param_count();
int gemm_bf16_launch(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                     float* C, int M, int N, int K, hipStream_t stream);
int cnn_prepare_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                       const int16_t* half, int K, uint16_t* x_out, float* norm, hipStream_t stream);
int cnn_encode_launch(const float* params, const float* norm, int K, float* t, hipStream_t stream);
int cnn_propose_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                       const int16_t* half, int K, const float* theta, const uint16_t* theta_bf, float* params,
                       hipStream_t stream);
int cnn_forward_launch(const uint16_t* x, int K, const float* theta, const uint16_t* theta_bf, const float* norm,
                       float* out, hipStream_t stream);
int cnn_train_step_launch(const uint16_t* x, const float* t, const float* w, int Bn, float* theta,
                          uint16_t* theta_bf, float* adam_m, float* adam_v, int step, float lr, float* loss_out,
                          float* grad_out, hipStream_t stream);
int cnn_to_bf16_launch(const float* theta, int64_t n, uint16_t* theta_bf, hipStream_t stream);
int seed_neighbours_launch(const float* seeds, int32_t K, int64_t H, int64_t W, float radius, int32_t nb,
                           int32_t* out, hipStream_t stream);
int triangle_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step, int32_t* thresh_out,
                    hipStream_t stream);
int threshold_mask_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh, int32_t invert,
                          uint16_t* out, hipStream_t stream);
int morph_cross_launch(const uint16_t* in, int64_t H, int64_t W, int64_t pitch, int32_t op, uint16_t* out,
                       hipStream_t stream);
int logquad_window_launch(const void* img, int dtype, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                          const float* rows, const int32_t* lab_ids, int32_t K, int32_t pad,
                          const SpxLogquadWindowOptions& opt, double* centers, uint8_t* ok, hipStream_t stream);
int weighted_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                    const int32_t* rows, int32_t K, int32_t pad, double* centers, uint8_t* ok, hipStream_t stream);
int edge_moment_launch(const void* img, int dtype, int64_t H, int64_t W, int64_t pitch, const float* rows, int K,
                       const SpxEdgeErfOptions& opt, double* centers, uint8_t* ok, hipStream_t stream);
int edge_erf_launch(const void* img, int dtype, int64_t H, int64_t W, int64_t pitch, const float* rows, int K,
                    const SpxEdgeErfOptions& opt, double* centers, uint8_t* ok, hipStream_t stream);
int edge_profile_launch(const double* prof, int K, int L, const double* base, const double* seed,
                        const SpxEdgeErfOptions& opt, double* x, uint8_t* ok, hipStream_t stream);
int render_page_launch(uint64_t seed, const float* holes, int n, int64_t H, int64_t W, float background,
                       float read_sigma, uint16_t* img, int64_t pitch, hipStream_t stream);
int render_tiles_launch(uint64_t seed, int n, int T, const SpxTileSpec& spec, uint16_t* img, int64_t pitch,
                        float* truth, hipStream_t stream);
int hole_step_launch(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page, const float* tmpl,
                     int N, int U, int R, const int32_t* idx, int n, double* c, double* a, double* level,
                     double* delta, double* c_old, double nv0, double nv1, double rscale, int size, double max_move,
                     double max_delta, double a_med, double relax, double* stat, hipStream_t stream);
int hole_render_launch(double* page, unsigned long long* acc, int64_t H, int64_t W, float* cache, const float* tmpl,
                       int N, int U, int Rr, double rl2, const int32_t* idx, int n, const double* c_old, const double* c,
                       const double* a, const double* delta, hipStream_t stream);
int zeroed_accumulator(int64_t n, unsigned long long** out, hipStream_t stream);
int select_kth_launch(const double* x, int64_t n, int64_t stride, const int64_t* cols, const int64_t* ks, int nk,
                      double* out, hipStream_t stream);
int far_splat_launch(const double* c, const double* a, const double* delta, double cd0, int n, int mx, int my, int hp,
                     int wp, float* grid, hipStream_t stream);
int disk_templates_launch(double radius, double sigma, int N, int U, double r0, double ramp, int B, float* tmpl,
                          float* far, hipStream_t stream);
int outline_coverage_launch(const double* edge, const double* rho, const int32_t* ang, int N, int aa, double scale,
                            int nbatch, double* out, hipStream_t stream);
int outline_window_launch(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page,
                          const float* tmpl, int N, int U, int R, const int32_t* idx, int n, const double* c,
                          const double* a, const double* level, const double* delta, double radius_fine,
                          int32_t* bin, uint8_t* valid, double* data, double* model, double* own, double* v0,
                          hipStream_t stream);
int outline_far_splat_launch(const double* c, const double* a, const double* delta, int n, int mx, int my, int hp,
                             int wp, double* grid, hipStream_t stream);
int outline_stack_launch(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page,
                         const float* tmpl, int N, int U, int R, const int32_t* idx, int n, const uint8_t* good,
                         const double* c, const double* a, const double* level, const double* delta, double nv0,
                         double nv1, double radius_fine, int cleaned, double* sums, double* wsum, hipStream_t stream);

}  // namespace spx

extern "C" {

int64_t spx_cnn_param_count(void) { return spx::cnn_param_count(); }

int spx_gemm_bf16(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                  float* C, int32_t M, int32_t N, int32_t K, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(A != nullptr && B != nullptr && C != nullptr, "spx_gemm_bf16: null pointer");
    SPX_REQUIRE(M > 0 && N > 0 && K > 0, "spx_gemm_bf16: empty shape");
    return spx::gemm_bf16_launch(A, sam, sak, B, sbk, sbn, C, M, N, K, spx::as_stream(stream));
}

int spx_cnn_prepare(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                    const int16_t* half, int32_t K, uint16_t* x_out, float* norm_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(K >= 0, "spx_cnn_prepare: K < 0");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(img != nullptr && seeds != nullptr && half != nullptr && x_out != nullptr && norm_out != nullptr,
                "spx_cnn_prepare: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_cnn_prepare: bad image shape or pitch");
    return spx::cnn_prepare_launch(img, H, W, pitch, seeds, half, K, x_out, norm_out, spx::as_stream(stream));
}

int spx_cnn_encode(const float* params, const float* norm, int32_t K, float* t_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(K >= 0 && (K == 0 || (params && norm && t_out)), "spx_cnn_encode: bad arguments");
    return spx::cnn_encode_launch(params, norm, K, t_out, spx::as_stream(stream));
}

int spx_cnn_propose(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds, const int16_t* half,
                    int32_t K, const float* theta, const uint16_t* theta_bf, float* params_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(K >= 0, "spx_cnn_propose: K < 0");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_cnn_propose: bad image shape or pitch");
    SPX_REQUIRE(K == 0 || (img && seeds && half && theta && theta_bf && params_out), "spx_cnn_propose: null pointer");
    return spx::cnn_propose_launch(img, H, W, pitch, seeds, half, K, theta, theta_bf, params_out,
                                   spx::as_stream(stream));
}

int spx_cnn_forward(const uint16_t* x, int32_t K, const float* theta, const uint16_t* theta_bf,
                    const float* norm, float* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(K >= 0, "spx_cnn_forward: K < 0");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(x != nullptr && theta != nullptr && theta_bf != nullptr && out != nullptr,
                "spx_cnn_forward: null pointer");
    return spx::cnn_forward_launch(x, K, theta, theta_bf, norm, out, spx::as_stream(stream));
}

int spx_cnn_train_step(const uint16_t* x, const float* target, const float* weight, int32_t B, float* theta,
                       uint16_t* theta_bf, float* adam_m, float* adam_v, int32_t step, float lr, float* loss_out,
                       spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(B > 0 && step >= 1 && lr > 0.0f, "spx_cnn_train_step: bad batch, step or lr");
    SPX_REQUIRE(x && target && theta && theta_bf && adam_m && adam_v, "spx_cnn_train_step: null pointer");
    return spx::cnn_train_step_launch(x, target, weight, B, theta, theta_bf, adam_m, adam_v, step, lr, loss_out,
                                      nullptr, spx::as_stream(stream));
}

int spx_cnn_grad(const uint16_t* x, const float* target, const float* weight, int32_t B, float* theta,
                 uint16_t* theta_bf, float* grad_out, float* loss_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(B > 0, "spx_cnn_grad: B <= 0");
    SPX_REQUIRE(x && target && theta && theta_bf && grad_out, "spx_cnn_grad: null pointer");
    return spx::cnn_train_step_launch(x, target, weight, B, theta, theta_bf, nullptr, nullptr, 1, 1.0f, loss_out,
                                      grad_out, spx::as_stream(stream));
}

int spx_edge_erf(const void* img, int32_t dtype, int64_t H, int64_t W, int64_t pitch, const float* rows,
                 int32_t K, const SpxEdgeErfOptions* opt, double* centers_out, uint8_t* ok_out,
                 spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(opt != nullptr && K >= 0, "spx_edge_erf: bad arguments");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(img && rows && centers_out && ok_out, "spx_edge_erf: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_edge_erf: bad image shape or pitch");
    SPX_REQUIRE(opt->iters >= 0 && opt->erf_gn_iters >= 0 && opt->smooth_passes >= 0 && opt->loc_rad >= 1,
                "spx_edge_erf: bad iteration or window options");
    return spx::edge_erf_launch(img, dtype, H, W, pitch, rows, K, *opt, centers_out, ok_out, spx::as_stream(stream));
}

int spx_seed_neighbours(const float* seeds, int32_t K, int64_t H, int64_t W, float radius, int32_t nb,
                        int32_t* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(K >= 0, "spx_seed_neighbours: K < 0");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(seeds != nullptr && out != nullptr, "spx_seed_neighbours: null pointer");
    SPX_REQUIRE(H > 0 && W > 0, "spx_seed_neighbours: bad image shape");
    SPX_REQUIRE(radius > 0.0f && nb >= 1 && nb <= 8, "spx_seed_neighbours: need radius > 0 and nb in [1, 8]");
    return spx::seed_neighbours_launch(seeds, K, H, W, radius, nb, out, spx::as_stream(stream));
}

int spx_triangle_u8(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step, int32_t* thresh_out,
                    spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(img != nullptr && thresh_out != nullptr, "spx_triangle_u8: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W && step >= 1, "spx_triangle_u8: bad image shape, pitch or step");
    return spx::triangle_launch(img, H, W, pitch, step, thresh_out, spx::as_stream(stream));
}

int spx_threshold_mask(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh, int32_t invert,
                       uint16_t* mask_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(img != nullptr && mask_out != nullptr, "spx_threshold_mask: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_threshold_mask: bad image shape or pitch");
    return spx::threshold_mask_launch(img, H, W, pitch, thresh, invert, mask_out, spx::as_stream(stream));
}

int spx_morph_cross(const uint16_t* in, int64_t H, int64_t W, int64_t pitch, int32_t op, uint16_t* out,
                    spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(in != nullptr && out != nullptr && in != out, "spx_morph_cross: null or aliased buffers");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_morph_cross: bad image shape or pitch");
    SPX_REQUIRE(op == 0 || op == 1, "spx_morph_cross: op must be 0 (erode) or 1 (dilate)");
    return spx::morph_cross_launch(in, H, W, pitch, op, out, spx::as_stream(stream));
}

int spx_edge_moment(const void* img, int32_t dtype, int64_t H, int64_t W, int64_t pitch, const float* rows,
                    int32_t K, const SpxEdgeErfOptions* opt, double* centers_out, uint8_t* ok_out,
                    spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(opt != nullptr && K >= 0, "spx_edge_moment: bad arguments");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(img && rows && centers_out && ok_out, "spx_edge_moment: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_edge_moment: bad image shape or pitch");
    SPX_REQUIRE(opt->iters >= 0 && opt->smooth_passes >= 0 && opt->loc_rad >= 0,
                "spx_edge_moment: bad iteration or window options");
    return spx::edge_moment_launch(img, dtype, H, W, pitch, rows, K, *opt, centers_out, ok_out,
                                   spx::as_stream(stream));
}

int spx_refine_logquad_window(const void* img, int32_t dtype, int64_t H, int64_t W, int64_t pitch,
                              const int32_t* labels, const float* rows, const int32_t* lab_ids, int32_t K,
                              int32_t pad, const SpxLogquadWindowOptions* opt, double* centers_out,
                              uint8_t* ok_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(opt != nullptr && K >= 0, "spx_refine_logquad_window: bad arguments");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(img && labels && rows && lab_ids && centers_out && ok_out, "spx_refine_logquad_window: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_refine_logquad_window: bad image shape or pitch");
    SPX_REQUIRE(pad >= 0, "spx_refine_logquad_window: pad < 0");
    return spx::logquad_window_launch(img, dtype, H, W, pitch, labels, rows, lab_ids, K, pad, *opt, centers_out,
                                      ok_out, spx::as_stream(stream));
}

int spx_refine_weighted(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                        const int32_t* rows, int32_t K, int32_t pad, double* centers_out, uint8_t* ok_out,
                        spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(K >= 0, "spx_refine_weighted: K < 0");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(img && labels && rows && centers_out && ok_out, "spx_refine_weighted: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_refine_weighted: bad image shape or pitch");
    SPX_REQUIRE(pad >= 0, "spx_refine_weighted: pad < 0");
    return spx::weighted_launch(static_cast<const uint16_t*>(img), H, W, pitch, labels, rows, K, pad, centers_out,
                                ok_out, spx::as_stream(stream));
}

int spx_edge_profile_erf(const double* prof, int32_t K, int32_t L, const double* base, const double* seed,
                         const SpxEdgeErfOptions* opt, double* x_out, uint8_t* ok_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(opt != nullptr && K >= 0, "spx_edge_profile_erf: bad arguments");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(prof && base && seed && x_out && ok_out, "spx_edge_profile_erf: null pointer");
    return spx::edge_profile_launch(prof, K, L, base, seed, *opt, x_out, ok_out, spx::as_stream(stream));
}

int spx_render_page(uint64_t seed, const float* holes, int32_t n, int64_t H, int64_t W, float background,
                    float read_sigma, uint16_t* img, int64_t pitch, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(n >= 0 && H > 0 && W > 0 && pitch >= W, "spx_render_page: bad sizes");
    SPX_REQUIRE(img != nullptr && (n == 0 || holes != nullptr), "spx_render_page: null pointer");
    SPX_REQUIRE((W + 63) / 64 * ((H + 63) / 64) < (int64_t)0x7fffffff, "spx_render_page: image too large");
    return spx::render_page_launch(seed, holes, n, H, W, background, read_sigma, img, pitch, spx::as_stream(stream));
}

int spx_render_tiles(uint64_t seed, int32_t n, int32_t tile, const SpxTileSpec* spec, uint16_t* img,
                     int64_t pitch, float* truth_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(spec != nullptr && img != nullptr && truth_out != nullptr, "spx_render_tiles: null pointer");
    SPX_REQUIRE(n >= 0 && tile >= 8 && tile <= 256 && spec->cols >= 1, "spx_render_tiles: bad sizes");
    SPX_REQUIRE(pitch >= (int64_t)tile * spec->cols, "spx_render_tiles: pitch smaller than the mosaic width");
    SPX_REQUIRE(spec->r_lo > 0 && spec->r_hi >= spec->r_lo && spec->s_lo > 0 && spec->s_hi >= spec->s_lo,
                "spx_render_tiles: bad radius or sigma range");
    return spx::render_tiles_launch(seed, n, tile, *spec, img, pitch, truth_out, spx::as_stream(stream));
}

int spx_cnn_to_bf16(const float* theta, int64_t n, uint16_t* theta_bf, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(theta && theta_bf && n >= 0, "spx_cnn_to_bf16: bad arguments");
    return spx::cnn_to_bf16_launch(theta, n, theta_bf, spx::as_stream(stream));
}

int spx_fit_stats(uint64_t* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(out != nullptr, "spx_fit_stats: out is NULL");
    return spx::fit_stats_copy(out, 4, spx::as_stream(stream));
}

int spx_fit_counters(uint64_t* out, int32_t n, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(out != nullptr, "spx_fit_counters: out is NULL");
    SPX_REQUIRE(n >= 1 && n <= 18, "spx_fit_counters: n must lie in [1, 18]");
    return spx::fit_stats_copy(out, n, spx::as_stream(stream));
}


int spx_init(int device) {
    spx::clear_error();
    int n = 0;
    SPX_HIP(hipGetDeviceCount(&n));
    SPX_REQUIRE(device >= 0 && device < n, "spx_init: device ordinal out of range");
    SPX_HIP(hipSetDevice(device));
    SPX_HIP(hipFree(nullptr));  // creates the context
    return spx::fit_scratch_init(device);
}

void spx_release(void) {
    std::lock_guard<std::mutex> lock(spx::g_ws_mutex);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return;
    for (auto& kv : spx::g_ws) {
        if (kv.second.ptr == nullptr) continue;
        if (hipSetDevice(kv.first.first) == hipSuccess) {
            (void)hipStreamSynchronize(kv.first.second);
            (void)hipFree(kv.second.ptr);
        }
    }
    spx::g_ws.clear();
    for (auto& kv : spx::g_fit_stats) {
        if (kv.second == nullptr) continue;
        if (hipSetDevice(kv.first.first) == hipSuccess) {
            (void)hipStreamSynchronize(kv.first.second);
            (void)hipFree(kv.second);
        }
    }
    spx::g_fit_stats.clear();
    spx::fit_streams_release();
    (void)hipSetDevice(cur);
}

const char* spx_last_error(void) { return spx::g_last_error.c_str(); }

const char* spx_version(void) { return "subpx-hip 0.1 gfx950"; }

int spx_fit_batch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                  const int16_t* half, const uint8_t* model, const float* init,
                  const SpxFitOptions* opt, int32_t K, float* params_out, double* centers_out,
                  float* rms_out, int16_t* iters_out, uint8_t* status_out, spx_stream_t stream) {
    return spx_fit_batch_cells(img, H, W, pitch, seeds, half, model, init, opt, K, nullptr, 0, params_out,
                               centers_out, rms_out, iters_out, status_out, stream);
}

namespace {
int check_fit_args(const char* name, const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                   const int16_t* half, const uint8_t* model, const SpxFitOptions* opt, int32_t K,
                   const int32_t* neighbours, int32_t nb, const float* params_out, const float* rms_out,
                   const int16_t* iters_out, const uint8_t* status_out) {
    const std::string n(name);
    SPX_REQUIRE(opt != nullptr, n + ": opt is NULL");
    SPX_REQUIRE(K >= 0, n + ": K < 0");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(img != nullptr && seeds != nullptr && half != nullptr && model != nullptr, n + ": null input pointer");
    SPX_REQUIRE(params_out != nullptr && rms_out != nullptr && iters_out != nullptr && status_out != nullptr,
                n + ": null output pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, n + ": bad image shape or pitch");
    SPX_REQUIRE(opt->max_iter >= 0 && opt->max_iter <= 10000, n + ": max_iter out of range");
    SPX_REQUIRE(opt->learned_iters >= 0 && opt->learned_iters <= 1000, n + ": learned_iters out of range");
    SPX_REQUIRE(opt->step_tol >= 0.0f, n + ": step_tol < 0");
    SPX_REQUIRE(opt->lambda0 > 0.0f && opt->lambda_up > 1.0f && opt->lambda_down > 0.0f && opt->lambda_down < 1.0f,
                n + ": damping options out of range");
    SPX_REQUIRE(opt->init_sigma > 0.0f, n + ": init_sigma <= 0");
    SPX_REQUIRE(neighbours == nullptr || (nb >= 1 && nb <= 8), n + ": nb must lie in [1, 8]");
    return SPX_OK;
}
}  // namespace

int spx_fit_batch_cells(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                        const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions* opt,
                        int32_t K, const int32_t* neighbours, int32_t nb, float* params_out,
                        double* centers_out, float* rms_out, int16_t* iters_out, uint8_t* status_out,
                        spx_stream_t stream) {
    spx::clear_error();
    int rc = check_fit_args("spx_fit_batch", img, H, W, pitch, seeds, half, model, opt, K, neighbours, nb, params_out,
                            rms_out, iters_out, status_out);
    if (rc != SPX_OK || K == 0) return rc;
    return spx::fit_batch_launch(img, H, W, pitch, seeds, half, model, init, *opt, K, neighbours,
                                 neighbours ? nb : 0, params_out, centers_out, rms_out, iters_out, status_out,
                                 spx::as_stream(stream));
}

int spx_fit_batch_noalloc(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                          const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions* opt,
                          int32_t K, const int32_t* neighbours, int32_t nb, float* params_out, double* centers_out,
                          float* rms_out, int16_t* iters_out, uint8_t* status_out, spx_stream_t stream) {
    spx::clear_error();
    int rc = check_fit_args("spx_fit_batch_noalloc", img, H, W, pitch, seeds, half, model, opt, K, neighbours, nb,
                            params_out, rms_out, iters_out, status_out);
    if (rc != SPX_OK) return rc;
    return spx::fit_batch_noalloc_launch(img, H, W, pitch, seeds, half, model, init, *opt, K, neighbours,
                                         neighbours ? nb : 0, params_out, centers_out, rms_out, iters_out, status_out,
                                         spx::as_stream(stream));
}

int spx_fit_graph_create(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                         const int16_t* half, const uint8_t* model, const float* init, const SpxFitOptions* opt,
                         int32_t K, const int32_t* neighbours, int32_t nb, float* params_out, double* centers_out,
                         float* rms_out, int16_t* iters_out, uint8_t* status_out, spx_fit_graph_t* graph_out) {
    spx::clear_error();
    SPX_REQUIRE(graph_out != nullptr, "spx_fit_graph_create: graph_out is NULL");
    *graph_out = nullptr;
    SPX_REQUIRE(K >= 1, "spx_fit_graph_create: K must be at least 1");
    int rc = check_fit_args("spx_fit_graph_create", img, H, W, pitch, seeds, half, model, opt, K, neighbours, nb,
                            params_out, rms_out, iters_out, status_out);
    if (rc != SPX_OK) return rc;
    return spx::fit_graph_create(img, H, W, pitch, seeds, half, model, init, *opt, K, neighbours,
                                 neighbours ? nb : 0, params_out, centers_out, rms_out, iters_out, status_out,
                                 graph_out);
}

int spx_fit_graph_launch(spx_fit_graph_t graph, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(graph != nullptr, "spx_fit_graph_launch: graph is NULL");
    return spx::fit_graph_launch(graph, spx::as_stream(stream));
}

int spx_fit_graph_destroy(spx_fit_graph_t graph) {
    spx::clear_error();
    return spx::fit_graph_destroy(graph);
}

int spx_voronoi_labels(const double* seeds, int32_t n, int64_t H, int64_t W, int32_t* labels, int32_t* areas,
                       double* dist, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(n >= 0 && (n == 0 || seeds != nullptr), "spx_voronoi_labels: bad seeds");
    SPX_REQUIRE(H > 0 && W > 0, "spx_voronoi_labels: bad image shape");
    SPX_REQUIRE(labels != nullptr || areas != nullptr || dist != nullptr, "spx_voronoi_labels: no output");
    return spx::voronoi_labels_launch(seeds, n, H, W, labels, areas, dist, spx::as_stream(stream));
}

namespace {
int check_cells(const char* name, int32_t N, int32_t Hm, int32_t Wm) {
    const std::string n(name);
    SPX_REQUIRE(N >= 0, n + ": N < 0");
    SPX_REQUIRE(N == 0 || (Hm >= 1 && Wm >= 1 && (int64_t)Hm * Wm < ((int64_t)1 << 28)), n + ": bad cell stack size");
    return SPX_OK;
}
}  // namespace

int spx_deblend_init(const double* image, int64_t H, int64_t W, const uint8_t* masks, const int32_t* origins,
                     const double* seeds, int32_t N, int32_t Hm, int32_t Wm, double* data, double* params,
                     spx_stream_t stream) {
    spx::clear_error();
    int rc = check_cells("spx_deblend_init", N, Hm, Wm);
    if (rc != SPX_OK || N == 0) return rc;
    SPX_REQUIRE(image != nullptr && masks != nullptr && origins != nullptr && seeds != nullptr && data != nullptr &&
                    params != nullptr, "spx_deblend_init: null pointer");
    SPX_REQUIRE(H > 0 && W > 0, "spx_deblend_init: bad image shape");
    return spx::deblend_init_launch(image, H, W, masks, origins, seeds, N, Hm, Wm, data, params,
                                    spx::as_stream(stream));
}

int spx_gauss_fit(const double* data, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm, const double* init,
                  int32_t max_iter, double* params_out, spx_stream_t stream) {
    spx::clear_error();
    int rc = check_cells("spx_gauss_fit", N, Hm, Wm);
    if (rc != SPX_OK || N == 0) return rc;
    SPX_REQUIRE(data != nullptr && masks != nullptr && init != nullptr && params_out != nullptr,
                "spx_gauss_fit: null pointer");
    SPX_REQUIRE(max_iter >= 0 && max_iter <= 10000, "spx_gauss_fit: max_iter out of range");
    return spx::gauss_fit_launch(data, masks, N, Hm, Wm, init, max_iter, params_out, spx::as_stream(stream));
}

int spx_gauss_render(const double* params_global, const double* radius, int32_t N, int64_t H, int64_t W,
                     double* model, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(N >= 0 && H > 0 && W > 0 && model != nullptr, "spx_gauss_render: bad arguments");
    SPX_REQUIRE(N == 0 || (params_global != nullptr && radius != nullptr), "spx_gauss_render: null pointer");
    return spx::gauss_render_launch(params_global, radius, N, H, W, model, spx::as_stream(stream));
}

int spx_deblend_clean(const double* image, const double* background, const double* model,
                      const double* params_global, const double* radius, const uint8_t* masks, const int32_t* origins,
                      int32_t N, int32_t Hm, int32_t Wm, int64_t H, int64_t W, double* out, spx_stream_t stream) {
    spx::clear_error();
    int rc = check_cells("spx_deblend_clean", N, Hm, Wm);
    if (rc != SPX_OK || N == 0) return rc;
    SPX_REQUIRE(image != nullptr && background != nullptr && model != nullptr && params_global != nullptr &&
                    radius != nullptr && masks != nullptr && origins != nullptr && out != nullptr,
                "spx_deblend_clean: null pointer");
    SPX_REQUIRE(H > 0 && W > 0, "spx_deblend_clean: bad image shape");
    return spx::deblend_clean_launch(image, background, model, params_global, radius, masks, origins, N, Hm, Wm, H, W,
                                     out, spx::as_stream(stream));
}

int spx_tile_medians(const double* image, const double* model, const double* dist, int64_t H, int64_t W, int32_t tile,
                     double model_below, double dist_atleast, double* medians, int64_t* counts, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(image != nullptr && model != nullptr && dist != nullptr && medians != nullptr && counts != nullptr,
                "spx_tile_medians: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && tile >= 1 && tile <= 4096, "spx_tile_medians: bad image shape or tile");
    return spx::tile_medians_launch(image, model, dist, H, W, tile, model_below, dist_atleast, medians, counts,
                                    spx::as_stream(stream));
}

int spx_tile_map(const double* medians, const double* yc, const double* xc, int32_t ny, int32_t nx, int64_t H,
                 int64_t W, double* map, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(medians != nullptr && yc != nullptr && xc != nullptr && map != nullptr, "spx_tile_map: null pointer");
    SPX_REQUIRE(ny >= 1 && nx >= 1 && H > 0 && W > 0, "spx_tile_map: bad sizes");
    return spx::tile_map_launch(medians, yc, xc, ny, nx, H, W, map, spx::as_stream(stream));
}

int spx_cell_refine(const double* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm, int32_t method,
                    const double* amplitude, double frac, double* centers, spx_stream_t stream) {
    spx::clear_error();
    int rc = check_cells("spx_cell_refine", N, Hm, Wm);
    if (rc != SPX_OK || N == 0) return rc;
    SPX_REQUIRE(rois != nullptr && masks != nullptr && centers != nullptr, "spx_cell_refine: null pointer");
    SPX_REQUIRE(method == 0 || method == 1, "spx_cell_refine: method must be 0 (weighted) or 1 (logquad)");
    return spx::cell_refine_launch(rois, masks, N, Hm, Wm, method, amplitude, frac, centers, spx::as_stream(stream));
}

int spx_deblend_window(const double* cleaned, const uint8_t* masks, const double* seeds_local, const double* reach,
                       int32_t N, int32_t Hm, int32_t Wm, int32_t pad, float* rois, uint8_t* masks_out,
                       spx_stream_t stream) {
    spx::clear_error();
    int rc = check_cells("spx_deblend_window", N, Hm, Wm);
    if (rc != SPX_OK || N == 0) return rc;
    SPX_REQUIRE(cleaned != nullptr && masks != nullptr && seeds_local != nullptr && reach != nullptr &&
                    rois != nullptr && masks_out != nullptr, "spx_deblend_window: null pointer");
    return spx::deblend_window_launch(cleaned, masks, seeds_local, reach, N, Hm, Wm, pad, rois, masks_out,
                                      spx::as_stream(stream));
}

int spx_voronoi_rois(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                     const double* rows, int32_t N, int32_t pad, int32_t Hm, int32_t Wm, float* rois, uint8_t* masks,
                     int32_t* origins, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(N >= 0 && pad >= 0, "spx_voronoi_rois: N < 0 or pad < 0");
    if (N == 0) return SPX_OK;
    SPX_REQUIRE(img != nullptr && labels != nullptr && rows != nullptr && rois != nullptr && masks != nullptr &&
                    origins != nullptr, "spx_voronoi_rois: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_voronoi_rois: bad image shape or pitch");
    SPX_REQUIRE(Hm >= 1 && Wm >= 1, "spx_voronoi_rois: empty ROI stack");
    return spx::voronoi_rois_launch(img, H, W, pitch, labels, rows, N, pad, Hm, Wm, rois, masks, origins,
                                    spx::as_stream(stream));
}

namespace {
int check_estimator(const char* name, const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                    int32_t factor, int32_t margin, const double* centers, const double* quality) {
    const std::string n(name);
    SPX_REQUIRE(N >= 0, n + ": N < 0");
    if (N == 0) return SPX_OK;
    SPX_REQUIRE(rois != nullptr && masks != nullptr && centers != nullptr && quality != nullptr, n + ": null pointer");
    SPX_REQUIRE(Hm >= 1 && Wm >= 1, n + ": empty ROIs");
    SPX_REQUIRE(factor >= 1 && factor <= 16, n + ": upsample_factor must lie in [1, 16]");
    SPX_REQUIRE(margin >= 0, n + ": boundary_margin < 0");
    SPX_REQUIRE((int64_t)Hm * factor * Wm * factor < ((int64_t)1 << 30), n + ": upsampled ROI too large");
    return SPX_OK;
}
}  // namespace

int spx_radial_symmetry(const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                        int32_t upsample_factor, int32_t boundary_margin, double* centers, double* quality,
                        spx_stream_t stream) {
    spx::clear_error();
    int rc = check_estimator("spx_radial_symmetry", rois, masks, N, Hm, Wm, upsample_factor, boundary_margin, centers,
                             quality);
    if (rc != SPX_OK || N == 0) return rc;
    return spx::radial_symmetry_launch(rois, masks, N, Hm, Wm, upsample_factor, boundary_margin, centers, quality,
                                       spx::as_stream(stream));
}

int spx_isophote_curvature(const float* rois, const uint8_t* masks, int32_t N, int32_t Hm, int32_t Wm,
                           int32_t upsample_factor, int32_t boundary_margin, double pre_smooth_sigma,
                           double* centers, double* quality, spx_stream_t stream) {
    spx::clear_error();
    int rc = check_estimator("spx_isophote_curvature", rois, masks, N, Hm, Wm, upsample_factor, boundary_margin,
                             centers, quality);
    if (rc != SPX_OK || N == 0) return rc;
    SPX_REQUIRE(pre_smooth_sigma >= 0.0 && pre_smooth_sigma <= 8.0, "spx_isophote_curvature: pre_smooth_sigma "
                "must lie in [0, 8]");
    return spx::isophote_curvature_launch(rois, masks, N, Hm, Wm, upsample_factor, boundary_margin, pre_smooth_sigma,
                                          centers, quality, spx::as_stream(stream));
}

int spx_otsu_u16(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step,
                 int32_t* thresh_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(img != nullptr && thresh_out != nullptr, "spx_otsu_u16: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_otsu_u16: bad image shape or pitch");
    SPX_REQUIRE(step >= 1, "spx_otsu_u16: step < 1");
    return spx::otsu_launch(img, H, W, pitch, step, thresh_out, spx::as_stream(stream));
}

int spx_threshold_cc(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh,
                     int32_t invert, int32_t connectivity, int32_t* labels, int32_t* stats,
                     double* centroids, int32_t cap, int32_t* n_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(img != nullptr && n_out != nullptr, "spx_threshold_cc: null pointer");
    SPX_REQUIRE(cap == 0 || (stats != nullptr && centroids != nullptr), "spx_threshold_cc: null stats");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_threshold_cc: bad image shape or pitch");
    SPX_REQUIRE(H < ((int64_t)1 << 24) && W < ((int64_t)1 << 24), "spx_threshold_cc: image side of 2^24 px or more");
    SPX_REQUIRE(connectivity == 4 || connectivity == 8, "spx_threshold_cc: connectivity must be 4 or 8");
    SPX_REQUIRE(cap >= 0, "spx_threshold_cc: cap < 0");
    return spx::threshold_cc_launch(img, H, W, pitch, thresh, invert, connectivity, labels, stats,
                                    centroids, cap, n_out, spx::as_stream(stream));
}

int spx_refine_logquad(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                       const int32_t* rows, int32_t K, int32_t pad, double* centers_out, uint8_t* ok_out,
                       spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(K >= 0, "spx_refine_logquad: K < 0");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(img != nullptr && labels != nullptr && rows != nullptr && centers_out != nullptr &&
                    ok_out != nullptr,
                "spx_refine_logquad: null pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_refine_logquad: bad image shape or pitch");
    SPX_REQUIRE(pad >= 0, "spx_refine_logquad: pad < 0");
    return spx::logquad_launch(img, H, W, pitch, labels, rows, K, pad, centers_out, ok_out,
                               spx::as_stream(stream));
}

int spx_hole_templates(double radius, double sigma, int32_t N, int32_t U, double r0, double ramp, int32_t far_box,
                       float* tmpl, float* far, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(tmpl != nullptr, "spx_hole_templates: null pointer");
    SPX_REQUIRE(N >= 8 && N <= 8192 && (N % 2) == 0 && U >= 1, "spx_hole_templates: bad grid");
    SPX_REQUIRE(radius > 0 && sigma > 0 && ramp > 0, "spx_hole_templates: radius, sigma and ramp must be > 0");
    SPX_REQUIRE(far == nullptr || (far_box >= 2 && far_box <= 4096), "spx_hole_templates: bad far box");
    return spx::disk_templates_launch(radius, sigma, N, U, r0, ramp, far_box, tmpl, far, spx::as_stream(stream));
}

int spx_hole_step(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page, const float* tmpl,
                  int32_t N, int32_t U, int32_t R, const int32_t* idx, int32_t n, double* c, double* a, double* level,
                  double* delta, double* c_old, const SpxHoleStepOptions* opt, double* stat, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(opt != nullptr, "spx_hole_step: null options");
    SPX_REQUIRE(n >= 0 && H > 0 && W > 0 && pitch >= W, "spx_hole_step: bad sizes");
    SPX_REQUIRE(R >= 1 && R <= 255 && N >= 8 && U >= 1, "spx_hole_step: bad window or grid");
    SPX_REQUIRE(n == 0 || (image && page && tmpl && idx && c && a && level && delta && c_old && stat),
                "spx_hole_step: null pointer");
    SPX_REQUIRE(opt->nv0 > 0 && opt->nv1 >= 0, "spx_hole_step: noise variance line must be positive");
    return spx::hole_step_launch(image, H, W, pitch, page, tmpl, N, U, R, idx, n, c, a, level, delta, c_old, opt->nv0,
                                 opt->nv1, opt->rscale, opt->size ? 1 : 0, opt->max_move, opt->max_delta, opt->a_med,
                                 opt->relax, stat, spx::as_stream(stream));
}

int spx_hole_render(double* page, int64_t H, int64_t W, float* cache, const float* tmpl, int32_t N, int32_t U,
                    int32_t Rr, double rl2, const int32_t* idx, int32_t n, const double* c_old, const double* c,
                    const double* a, const double* delta, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(n >= 0 && H > 0 && W > 0 && Rr >= 0 && Rr <= 255 && N >= 8 && U >= 1, "spx_hole_render: bad sizes");
    SPX_REQUIRE(n == 0 || (page && cache && tmpl && idx && c_old && c && a && delta), "spx_hole_render: null pointer");
    SPX_REQUIRE((int64_t)n * (2 * Rr + 1) * (2 * Rr + 1) < ((int64_t)1 << 40), "spx_hole_render: too many offsets");
    if (n == 0) return SPX_OK;
    unsigned long long* acc = nullptr;
    hipStream_t s = spx::as_stream(stream);
    int rc = spx::zeroed_accumulator(H * W, &acc, s);
    if (rc != SPX_OK) return rc;
    return spx::hole_render_launch(page, acc, H, W, cache, tmpl, N, U, Rr, rl2, idx, n, c_old, c, a, delta, s);
}

int spx_hole_sweep(const float* image, int64_t H, int64_t W, int64_t pitch, double* page, const float* tmpl, int32_t N,
                   int32_t U, int32_t R, int32_t Rr, double rl2, const int32_t* order, const int32_t* start,
                   int32_t ncolour, double* c, double* a, double* level, double* delta, double* c_old, float* cache,
                   const SpxHoleStepOptions* opt, double* stat, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(opt != nullptr && start != nullptr, "spx_hole_sweep: null options or offsets");
    SPX_REQUIRE(ncolour >= 0 && H > 0 && W > 0 && pitch >= W, "spx_hole_sweep: bad sizes");
    SPX_REQUIRE(R >= 1 && R <= 255 && Rr >= 0 && Rr <= 255 && N >= 8 && U >= 1, "spx_hole_sweep: bad window or grid");
    SPX_REQUIRE(opt->nv0 > 0 && opt->nv1 >= 0, "spx_hole_sweep: noise variance line must be positive");
    SPX_REQUIRE(start[0] == 0, "spx_hole_sweep: offsets must start at 0");
    for (int j = 0; j < ncolour; ++j) SPX_REQUIRE(start[j + 1] >= start[j], "spx_hole_sweep: offsets must not decrease");
    const int32_t n = start[ncolour];
    SPX_REQUIRE(n == 0 || (image && page && tmpl && order && c && a && level && delta && c_old && cache && stat),
                "spx_hole_sweep: null pointer");
    const int64_t P = (int64_t)(2 * Rr + 1) * (2 * Rr + 1);
    hipStream_t s = spx::as_stream(stream);
    unsigned long long* acc = nullptr;
    if (n > 0) {
        int rc = spx::zeroed_accumulator(H * W, &acc, s);
        if (rc != SPX_OK) return rc;
    }
    for (int j = 0; j < ncolour; ++j) {
        const int32_t b = start[j], m = start[j + 1] - start[j];
        if (m == 0) continue;
        int rc = spx::hole_step_launch(image, H, W, pitch, page, tmpl, N, U, R, order + b, m, c, a, level, delta,
                                       c_old + 2 * (int64_t)b, opt->nv0, opt->nv1, opt->rscale, opt->size ? 1 : 0,
                                       opt->max_move, opt->max_delta, opt->a_med, opt->relax, stat + 4 * (int64_t)b, s);
        if (rc != SPX_OK) return rc;
        rc = spx::hole_render_launch(page, acc, H, W, cache + P * b, tmpl, N, U, Rr, rl2, order + b, m,
                                     c_old + 2 * (int64_t)b, c, a, delta, s);
        if (rc != SPX_OK) return rc;
    }
    return SPX_OK;
}

int spx_select_kth(const double* x, int64_t n, int64_t stride, const int64_t* cols, const int64_t* ks, int32_t nk,
                   double* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(ks != nullptr && nk >= 0 && nk <= 8, "spx_select_kth: 0..8 ranks");
    SPX_REQUIRE(n > 0 && stride >= 1 && x != nullptr && out != nullptr, "spx_select_kth: bad input");
    for (int j = 0; j < nk; ++j) {
        SPX_REQUIRE(ks[j] >= 0 && ks[j] < n, "spx_select_kth: rank out of range");
        SPX_REQUIRE(cols == nullptr || cols[j] >= 0, "spx_select_kth: negative column offset");
    }
    return spx::select_kth_launch(x, n, stride, cols, ks, nk, out, spx::as_stream(stream));
}

int spx_far_splat(const double* c, const double* a, const double* delta, double cd0, int32_t n, int32_t mx, int32_t my,
                  int32_t hp, int32_t wp, float* grid, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(n >= 0 && hp > 0 && wp > 0 && cd0 > 0, "spx_far_splat: bad sizes");
    SPX_REQUIRE(n == 0 || (c && a && delta && grid), "spx_far_splat: null pointer");
    return spx::far_splat_launch(c, a, delta, cd0, n, mx, my, hp, wp, grid, spx::as_stream(stream));
}

int spx_outline_coverage(const double* edge, const double* rho, const int32_t* ang, int32_t N, int32_t aa,
                         double scale, int32_t nbatch, double* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(N >= 8 && N <= 4096 && aa >= 1 && aa <= 16, "spx_outline_coverage: bad grid");
    SPX_REQUIRE(nbatch >= 0 && nbatch <= 65535, "spx_outline_coverage: 0..65535 outlines per call");
    SPX_REQUIRE(nbatch == 0 || (edge && rho && ang && out), "spx_outline_coverage: null pointer");
    return spx::outline_coverage_launch(edge, rho, ang, N, aa, scale, nbatch, out, spx::as_stream(stream));
}

int spx_outline_windows(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page,
                        const float* tmpl, int32_t N, int32_t U, int32_t R, const int32_t* idx, int32_t n,
                        const double* c, const double* a, const double* level, const double* delta,
                        double radius_fine, int32_t* bin, uint8_t* valid, double* data, double* model, double* own,
                        double* v0, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(n >= 0 && H > 0 && W > 0 && pitch >= W, "spx_outline_windows: bad sizes");
    SPX_REQUIRE(R >= 0 && R <= 255 && N >= 8 && N <= 4096 && U >= 1, "spx_outline_windows: bad window or grid");
    SPX_REQUIRE(n == 0 || (image && idx && c), "spx_outline_windows: null pointer");
    SPX_REQUIRE(!(own || v0) || tmpl, "spx_outline_windows: template values need the template");
    SPX_REQUIRE(!own || (a && delta), "spx_outline_windows: own light needs brightness and edge offset");
    return spx::outline_window_launch(image, H, W, pitch, page, tmpl, N, U, R, idx, n, c, a, level, delta,
                                      radius_fine, bin, valid, data, model, own, v0, spx::as_stream(stream));
}

int spx_outline_far_splat(const double* c, const double* a, const double* delta, int32_t n, int32_t mx, int32_t my,
                          int32_t hp, int32_t wp, double* grid, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(n >= 0 && hp > 0 && wp > 0, "spx_outline_far_splat: bad sizes");
    SPX_REQUIRE(n == 0 || (c && a && delta && grid), "spx_outline_far_splat: null pointer");
    return spx::outline_far_splat_launch(c, a, delta, n, mx, my, hp, wp, grid, spx::as_stream(stream));
}

int spx_outline_stack(const float* image, int64_t H, int64_t W, int64_t pitch, const double* page, const float* tmpl,
                      int32_t N, int32_t U, int32_t R, const int32_t* idx, int32_t n, const uint8_t* good,
                      const double* c, const double* a, const double* level, const double* delta, double nv0,
                      double nv1, double radius_fine, int32_t cleaned, double* sums, double* wsum,
                      spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(n >= 0 && H > 0 && W > 0 && pitch >= W, "spx_outline_stack: bad sizes");
    SPX_REQUIRE(R >= 0 && R <= 255 && N >= 8 && N <= 4096 && U >= 1, "spx_outline_stack: bad window or grid");
    SPX_REQUIRE(n == 0 || (image && idx && c && sums && wsum), "spx_outline_stack: null pointer");
    SPX_REQUIRE(!cleaned || (page && tmpl && a && level && delta), "spx_outline_stack: the cleaned stack needs the "
                "page model, template and hole state");
    SPX_REQUIRE(!cleaned || nv0 > 0, "spx_outline_stack: noise variance line must be positive");
    return spx::outline_stack_launch(image, H, W, pitch, page, tmpl, N, U, R, idx, n, good, c, a, level, delta, nv0,
                                     nv1, radius_fine, cleaned ? 1 : 0, sums, wsum, spx::as_stream(stream));
}

}  // extern "C"
