// C-ABI entry points of libsubpx_hip.so: lifecycle, error text, workspaces and argument
// checks. Kernels live in the other translation units.
#include <mutex>
#include <string>
#include <vector>

#include "spx_internal.hpp"

namespace spx {

namespace {
thread_local std::string g_last_error;

struct DeviceWorkspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::vector<DeviceWorkspace> g_ws;  // indexed by device ordinal
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int workspace(size_t bytes, void** out) {
    int dev = 0;
    SPX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
    DeviceWorkspace& w = g_ws[dev];
    if (w.bytes < bytes) {
        if (w.ptr != nullptr) {
            // Work already queued may still read the old buffer.
            SPX_HIP(hipDeviceSynchronize());
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

int fit_batch_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                     const int16_t* half, const uint8_t* model, const float* init,
                     const SpxFitOptions& opt, int32_t K, float* params, float* rms, int16_t* iters,
                     uint8_t* status, hipStream_t stream);

int fit_stats_copy(uint64_t* out, hipStream_t stream);
int otsu_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t step,
                int32_t* thresh_out, hipStream_t stream);
int threshold_cc_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, int32_t thresh,
                        int32_t invert, int32_t connectivity, int32_t* labels, int32_t* stats,
                        double* cents, int32_t cap, int32_t* n_out, hipStream_t stream);
int logquad_launch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const int32_t* labels,
                   const int32_t* rows, int32_t K, int32_t pad, double* centers, uint8_t* ok,
                   hipStream_t stream);

}  // namespace spx

extern "C" {

int spx_fit_stats(uint64_t* out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(out != nullptr, "spx_fit_stats: out is NULL");
    return spx::fit_stats_copy(out, spx::as_stream(stream));
}


int spx_init(int device) {
    spx::clear_error();
    int n = 0;
    SPX_HIP(hipGetDeviceCount(&n));
    SPX_REQUIRE(device >= 0 && device < n, "spx_init: device ordinal out of range");
    SPX_HIP(hipSetDevice(device));
    SPX_HIP(hipFree(nullptr));  // creates the context
    return SPX_OK;
}

void spx_release(void) {
    std::lock_guard<std::mutex> lock(spx::g_ws_mutex);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return;
    for (size_t d = 0; d < spx::g_ws.size(); ++d) {
        if (spx::g_ws[d].ptr == nullptr) continue;
        if (hipSetDevice((int)d) == hipSuccess) {
            (void)hipDeviceSynchronize();
            (void)hipFree(spx::g_ws[d].ptr);
        }
        spx::g_ws[d] = spx::DeviceWorkspace{};
    }
    (void)hipSetDevice(cur);
}

const char* spx_last_error(void) { return spx::g_last_error.c_str(); }

const char* spx_version(void) { return "subpx-hip 0.1 gfx950"; }

int spx_fit_batch(const uint16_t* img, int64_t H, int64_t W, int64_t pitch, const float* seeds,
                  const int16_t* half, const uint8_t* model, const float* init,
                  const SpxFitOptions* opt, int32_t K, float* params_out, float* rms_out,
                  int16_t* iters_out, uint8_t* status_out, spx_stream_t stream) {
    spx::clear_error();
    SPX_REQUIRE(opt != nullptr, "spx_fit_batch: opt is NULL");
    SPX_REQUIRE(K >= 0, "spx_fit_batch: K < 0");
    if (K == 0) return SPX_OK;
    SPX_REQUIRE(img != nullptr && seeds != nullptr && half != nullptr && model != nullptr,
                "spx_fit_batch: null input pointer");
    SPX_REQUIRE(params_out != nullptr && rms_out != nullptr && iters_out != nullptr &&
                    status_out != nullptr,
                "spx_fit_batch: null output pointer");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_fit_batch: bad image shape or pitch");
    SPX_REQUIRE(opt->max_iter >= 0 && opt->max_iter <= 10000, "spx_fit_batch: max_iter out of range");
    SPX_REQUIRE(opt->step_tol >= 0.0f, "spx_fit_batch: step_tol < 0");
    SPX_REQUIRE(opt->lambda0 > 0.0f && opt->lambda_up > 1.0f && opt->lambda_down > 0.0f &&
                    opt->lambda_down < 1.0f,
                "spx_fit_batch: damping options out of range");
    SPX_REQUIRE(opt->init_sigma > 0.0f, "spx_fit_batch: init_sigma <= 0");
    return spx::fit_batch_launch(img, H, W, pitch, seeds, half, model, init, *opt, K, params_out,
                                 rms_out, iters_out, status_out, spx::as_stream(stream));
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
    SPX_REQUIRE(img != nullptr && labels != nullptr && n_out != nullptr, "spx_threshold_cc: null pointer");
    SPX_REQUIRE(cap == 0 || (stats != nullptr && centroids != nullptr), "spx_threshold_cc: null stats");
    SPX_REQUIRE(H > 0 && W > 0 && pitch >= W, "spx_threshold_cc: bad image shape or pitch");
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

}  // extern "C"
