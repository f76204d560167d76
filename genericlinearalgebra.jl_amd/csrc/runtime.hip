// Handle, workspace and error reporting of libgla_hip.so.
#include <cstdarg>
#include <cstring>

#include "gla_internal.hpp"

namespace gla {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int ensure_workspace(Handle *h, size_t bytes)
{
    if (bytes <= h->ws_bytes) return GLA_OK;
    // Round up to 64 MiB granules so that a sequence of growing calls
    // reallocates rarely.
    const size_t gran = size_t(64) << 20;
    size_t want = (bytes + gran - 1) / gran * gran;
    if (h->ws) {
        // The old buffer may still be read by queued work.
        GLA_HIP_TRY(hipDeviceSynchronize());
        GLA_HIP_TRY(hipFree(h->ws));
        h->ws = nullptr;
        h->ws_bytes = 0;
    }
    GLA_HIP_TRY(hipMalloc(&h->ws, want));
    h->ws_bytes = want;
    return GLA_OK;
}

namespace {
// (Re)creates the look-ahead Cholesky's stream pair: `masked`, limited to
// every CU but the highest-numbered `reserve`, and `chain`, a non-blocking
// stream on all CUs (a CU-masked stream is created blocking, so the chain
// cannot run on the caller's stream beside it).  reserve 0 drops both.
int set_side_cu_reserve(Handle *h, int reserve)
{
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_HIP_TRY(hipDeviceSynchronize());
    if (h->masked) { (void)hipStreamDestroy(h->masked); h->masked = nullptr; }
    if (h->chain) { (void)hipStreamDestroy(h->chain); h->chain = nullptr; }
    h->side_cu_reserve = reserve;
    if (reserve > 0) {
        const int words = (h->num_cus + 31) / 32;
        uint32_t mask[16] = {0};
        for (int c = 0; c < h->num_cus - reserve && c < 32 * 16; ++c)
            mask[c / 32] |= 1u << (c % 32);
        GLA_HIP_TRY(hipExtStreamCreateWithCUMask(&h->masked, (uint32_t)words, mask));
        GLA_HIP_TRY(hipStreamCreateWithFlags(&h->chain, hipStreamNonBlocking));
    }
    return GLA_OK;
}
}  // namespace

}  // namespace gla

extern "C" {

const char *gla_version(void) { return "gla_hip 0.1 (gfx950)"; }

const char *gla_last_error_string(void) { return gla::g_err; }

int gla_set_tuning(gla_handle_t hh, const char *key, int64_t value)
{
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (!key) return -2;
    if (std::strcmp(key, "chol_block") == 0) {
        if (value != 0 && value != -1 && (value < 128 || value % 128 != 0)) {
            gla::set_error("chol_block must be -1, 0 or a multiple of 128");
            return -3;
        }
        h->chol_block = value;
        return GLA_OK;
    }
    if (std::strcmp(key, "ldlt_block") == 0) {
        if (value < 128 || value % 128 != 0 || value > 4096) {
            gla::set_error("ldlt_block must be a multiple of 128 in [128, 4096]");
            return -3;
        }
        h->ldlt_block = value;
        return GLA_OK;
    }
    if (std::strcmp(key, "mixed_block") == 0) {
        if (value < 128 || value % 128 != 0 || value > 4096) {
            gla::set_error("mixed_block must be a multiple of 128 in [128, 4096]");
            return -3;
        }
        h->mixed_block = (int)value;
        return GLA_OK;
    }
    if (std::strcmp(key, "qr_panel") == 0) {
        if (value != 1 && value != 2) {
            gla::set_error("qr_panel must be 1 or 2");
            return -3;
        }
        h->qr_panel = (int)value;
        return GLA_OK;
    }
    if (std::strcmp(key, "qr_inner") == 0) {
        if (value != -1 && value != 0 && value != 32 && value != 64) {
            gla::set_error("qr_inner must be -1, 0, 32 or 64");
            return -3;
        }
        h->qr_inner = (int)value;
        return GLA_OK;
    }
    if (std::strcmp(key, "gemm_dl") == 0) {
        if (value < 0 || value > 3) {
            gla::set_error("gemm_dl must be 0, 1, 2 or 3");
            return -3;
        }
        gla::set_gemm_dl((int)value);
        return GLA_OK;
    }
    if (std::strcmp(key, "gemm_lds1") == 0) {
        if (value != 0 && value != 1) {
            gla::set_error("gemm_lds1 must be 0 or 1");
            return -3;
        }
        gla::set_gemm_lds1((int)value);
        return GLA_OK;
    }
    if (std::strcmp(key, "qr_hr") == 0) {
        if (value < -1) {
            gla::set_error("qr_hr must be -1, 0 or a row count");
            return -3;
        }
        h->qr_hr = value;
        return GLA_OK;
    }
    if (std::strcmp(key, "chol_psolve") == 0) {
        if (value != 0 && value != 1) {
            gla::set_error("chol_psolve must be 0 or 1");
            return -3;
        }
        h->chol_psolve = (int)value;
        return GLA_OK;
    }
    if (std::strcmp(key, "chol_syrk_late") == 0) {
        if (value != 0 && value != 1) {
            gla::set_error("chol_syrk_late must be 0 or 1");
            return -3;
        }
        h->chol_syrk_late = (int)value;
        return GLA_OK;
    }
    if (std::strcmp(key, "qr_lookahead") == 0) {
        if (value != 0 && value != 1) {
            gla::set_error("qr_lookahead must be 0 or 1");
            return -3;
        }
        h->qr_lookahead = (int)value;
        return GLA_OK;
    }
    if (std::strcmp(key, "spin_limit") == 0) {
        if (value < 1 || value > (int64_t)0x7fffffff) {
            gla::set_error("spin_limit must be in [1, 2^31)");
            return -3;
        }
        h->spin_limit = (unsigned)value;
        return GLA_OK;
    }
    if (std::strcmp(key, "side_cu_reserve") == 0) {
        if (value < 0 || value > h->num_cus / 2 || h->num_cus > 32 * 16) {
            gla::set_error("side_cu_reserve must be in [0, CUs/2]");
            return -3;
        }
        return gla::set_side_cu_reserve(h, (int)value);
    }
    if (std::strcmp(key, "qr_wait_test") == 0) {
        if (value != 0 && value != 1) {
            gla::set_error("qr_wait_test must be 0 or 1");
            return -3;
        }
        h->qr_wait_test = (int)value;
        return GLA_OK;
    }
    gla::set_error("unknown tuning key '%s'", key);
    return -2;
}

int gla_get_tuning(gla_handle_t hh, const char *key, int64_t *value)
{
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (!key) return -2;
    if (!value) return -3;
    const struct { const char *name; int64_t v; } knobs[] = {
        {"chol_block", h->chol_block},       {"ldlt_block", h->ldlt_block},
        {"mixed_block", h->mixed_block},     {"qr_panel", h->qr_panel},
        {"qr_inner", h->qr_inner},           {"gemm_dl", gla::get_gemm_dl()},
        {"gemm_lds1", gla::get_gemm_lds1()}, {"qr_hr", h->qr_hr},
        {"chol_syrk_late", h->chol_syrk_late}, {"qr_lookahead", h->qr_lookahead},
        {"chol_psolve", h->chol_psolve},
        {"spin_limit", (int64_t)h->spin_limit}, {"side_cu_reserve", h->side_cu_reserve},
        {"qr_wait_test", h->qr_wait_test},
    };
    for (const auto &k : knobs)
        if (std::strcmp(key, k.name) == 0) { *value = k.v; return GLA_OK; }
    gla::set_error("unknown tuning knob '%s'", key);
    return -2;
}

int gla_handle_create(gla_handle_t *out)
{
    using gla::Handle;
    if (!out) return -1;
    Handle *h = new Handle();
    GLA_HIP_TRY(hipGetDevice(&h->device));
    hipDeviceProp_t prop;
    GLA_HIP_TRY(hipGetDeviceProperties(&prop, h->device));
    h->num_cus = prop.multiProcessorCount;
    GLA_HIP_TRY(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    GLA_HIP_TRY(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    GLA_HIP_TRY(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    GLA_HIP_TRY(hipEventCreateWithFlags(&h->ev_step[0], hipEventDisableTiming));
    GLA_HIP_TRY(hipEventCreateWithFlags(&h->ev_step[1], hipEventDisableTiming));
    GLA_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&h->dinfo), 256));
    GLA_HIP_TRY(hipMemset(h->dinfo, 0, 256));
    GLA_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&h->dbar), 4096));
    GLA_HIP_TRY(hipMemset(h->dbar, 0, 4096));
    // The trailing SYRK of the look-ahead Cholesky leaves an eighth of the
    // CUs (32 of 256) to the chain of small kernels beside it.
    if (h->num_cus <= 32 * 16) {
        const int rc = gla::set_side_cu_reserve(h, h->num_cus / 8);
        if (rc != GLA_OK) return rc;
    }
    GLA_HIP_TRY(hipDeviceSynchronize());
    *out = reinterpret_cast<gla_handle_t>(h);
    return GLA_OK;
}

int gla_handle_destroy(gla_handle_t hh)
{
    using gla::Handle;
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_HIP_TRY(hipDeviceSynchronize());
    if (h->ws) (void)hipFree(h->ws);
    if (h->dinfo) (void)hipFree(h->dinfo);
    if (h->dbar) (void)hipFree(h->dbar);
    (void)hipEventDestroy(h->ev_fork);
    (void)hipEventDestroy(h->ev_join);
    (void)hipEventDestroy(h->ev_step[0]);
    (void)hipEventDestroy(h->ev_step[1]);
    (void)hipStreamDestroy(h->side);
    if (h->masked) (void)hipStreamDestroy(h->masked);
    if (h->chain) (void)hipStreamDestroy(h->chain);
    delete h;
    return GLA_OK;
}

int gla_gemm_d(gla_handle_t hh, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
               double *C, int64_t ldc, void *stream)
{
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) return -1;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0) return -6;
    GLA_HIP_TRY(hipSetDevice(h->device));
    return gla::gemm<double>(reinterpret_cast<hipStream_t>(stream), ta, tb, 0, m, n, k, alpha, A,
                             lda, B, ldb, beta, C, ldc, 0);
}

int gla_gemm_s(gla_handle_t hh, char ta, char tb, int64_t m, int64_t n, int64_t k, float alpha,
               const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
               int64_t ldc, void *stream)
{
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) return -1;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0) return -6;
    GLA_HIP_TRY(hipSetDevice(h->device));
    return gla::gemm<float>(reinterpret_cast<hipStream_t>(stream), ta, tb, 0, m, n, k, alpha, A,
                            lda, B, ldb, beta, C, ldc, 0);
}

}  // extern "C"
