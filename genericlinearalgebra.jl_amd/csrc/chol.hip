// Cholesky on gfx950: cholUnblocked!, cholBlocked!, cholRecursive!
// (src/cholesky.jl:3-55) and the rdiv!/rankUpdate! steps they call.
//
// Kernels
//   chol_panel_kernel  The diagonal block (≤128×128) is factored
//                      register-resident by every workgroup, column by column
//                      exactly as cholUnblocked! does (sqrt, reciprocal scale,
//                      rank-1 update), and in the same sweep each workgroup
//                      solves its own 64 rows of the block column below,
//                      X·L11ᵀ = A21 (the rdiv! at src/cholesky.jl:28).  One
//                      launch therefore does potrf(diag)+trsm(panel) with no
//                      inter-workgroup communication.  With factor = 0 the
//                      diagonal block already holds L and only the solve runs.
//   gemm (gemm.hip)    The trailing rankUpdate!(Hermitian(A22,:L), A21, -1)
//                      as an MFMA SYRK over lower tiles.
//
// Drivers
//   blocked   right-looking, block nb ≤ 128, depth-1 look-ahead: the update
//             of the next block column runs on the caller's stream right
//             after the panel, the rest of the trailing update runs on the
//             handle's side stream and overlaps the next panel.
//   recursive halves n as cholRecursive! does; below a threshold the blocked
//             driver factors the leaf.  rdiv! with a large L11 is itself
//             recursive (two half solves and a GEMM), its leaves are
//             factor = 0 panel launches.
#include <algorithm>

#include "gla_internal.hpp"

namespace gla {

__global__ void k_set_int(int *p, int v) { *p = v; }

namespace {

constexpr int kPanelNB = 128;   // max diagonal block
constexpr int kPanelRB = 64;    // rows of A21 per workgroup
constexpr int kPanelThreads = 512;

template <typename T>
__device__ __forceinline__ T dev_sqrt(T x);
template <>
__device__ __forceinline__ double dev_sqrt<double>(double x) { return __builtin_sqrt(x); }
template <>
__device__ __forceinline__ float dev_sqrt<float>(float x) { return __builtin_sqrtf(x); }

template <typename T, int NB, int RB, int NTH>
__global__ __launch_bounds__(NTH) void chol_panel_kernel(int nb, const T *__restrict__ A11,
                                                         int64_t lda, T *__restrict__ Lw,
                                                         int64_t ldw, T *__restrict__ B,
                                                         int64_t ldb, int64_t nrows, int factor,
                                                         int *info, int64_t info_base)
{
    constexpr int GD = NTH / NB;   // column groups of the diagonal block
    constexpr int QD = NB / GD;    // diagonal elements per thread
    constexpr int GB = NTH / RB;   // column groups of the row block
    constexpr int QB = NB / GB;    // row-block elements per thread
    static_assert(NB % 64 == 0 && RB % 64 == 0, "wave-uniform column groups");

    __shared__ T lbuf[NB];   // column k of L below the diagonal
    __shared__ T xbuf[RB];   // column k of X for this row block
    __shared__ T sd[2];      // L[k][k] and its reciprocal

    const int t = threadIdx.x;
    const int di = t % NB, dg = t / NB;
    const int bi = t % RB, bg = t / RB;
    const int64_t row = (int64_t)blockIdx.x * RB + bi;
    const bool rowok = row < nrows;

    T dr[QD], br[QB];
#pragma unroll
    for (int q = 0; q < QD; ++q) {
        const int dj = dg + GD * q;
        dr[q] = (di < nb && dj <= di) ? A11[di + (int64_t)dj * lda] : T(0);
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int bj = bg + GB * q;
        br[q] = (rowok && bj < nb) ? B[row + (int64_t)bj * ldb] : T(0);
    }

    // Pivot of column 0.
    if (di == 0 && dg == 0) {
        T akk = dr[0];
        T d = akk;
        if (factor) {
            if (akk < T(0) && blockIdx.x == 0 && info) atomicCAS(info, 0, (int)(info_base + 1));
            d = dev_sqrt(akk);
            dr[0] = d;
        }
        sd[0] = d;
        sd[1] = T(1) / d;
    }
    __syncthreads();

    for (int k = 0; k < nb; ++k) {
        const T rinv = sd[1];
        // Column k of L (cholesky.jl:8, rmul! by the reciprocal) and of X.
        if (dg == k % GD && di > k && di < nb) {
            const int q = k / GD;
            T cur = T(0);
#pragma unroll
            for (int qq = 0; qq < QD; ++qq)
                if (qq == q) cur = dr[qq];
            const T l = factor ? cur * rinv : cur;
#pragma unroll
            for (int qq = 0; qq < QD; ++qq)
                if (qq == q) dr[qq] = l;
            lbuf[di] = l;
        }
        if (bg == k % GB) {
            const int q = k / GB;
            T x = T(0);
#pragma unroll
            for (int qq = 0; qq < QB; ++qq)
                if (qq == q) x = br[qq];
            x = x * rinv;
#pragma unroll
            for (int qq = 0; qq < QB; ++qq)
                if (qq == q) br[qq] = x;
            xbuf[bi] = x;
        }
        __syncthreads();

        // Rank-1 updates (juliaBLAS.jl:53-63 for the diagonal block; the
        // right-looking substitution for the rows below).
        if (factor && di > k && di < nb) {
            const T li = lbuf[di];
#pragma unroll
            for (int q = 0; q < QD; ++q) {
                const int dj = dg + GD * q;
                if (dj > k && dj <= di) dr[q] -= li * lbuf[dj];
            }
        }
        {
            const T xi = xbuf[bi];
#pragma unroll
            for (int q = 0; q < QB; ++q) {
                const int bj = bg + GB * q;
                if (bj > k && bj < nb) br[q] -= xi * lbuf[bj];
            }
        }
        // Pivot of column k+1, by its owner, right after its own update.
        if (k + 1 < nb && di == k + 1 && dg == (k + 1) % GD) {
            const int q = (k + 1) / GD;
            T akk = T(0);
#pragma unroll
            for (int qq = 0; qq < QD; ++qq)
                if (qq == q) akk = dr[qq];
            T d = akk;
            if (factor) {
                if (akk < T(0) && blockIdx.x == 0 && info)
                    atomicCAS(info, 0, (int)(info_base + k + 2));
                d = dev_sqrt(akk);
#pragma unroll
                for (int qq = 0; qq < QD; ++qq)
                    if (qq == q) dr[qq] = d;
            }
            sd[0] = d;
            sd[1] = T(1) / d;
        }
        __syncthreads();
    }

    if (factor && blockIdx.x == 0 && Lw) {
#pragma unroll
        for (int q = 0; q < QD; ++q) {
            const int dj = dg + GD * q;
            if (di < nb && dj <= di) Lw[di + (int64_t)dj * ldw] = dr[q];
        }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int bj = bg + GB * q;
        if (rowok && bj < nb) B[row + (int64_t)bj * ldb] = br[q];
    }
}

// Copies the lower triangle (incl. diagonal) of an nb×nb block.
template <typename T>
__global__ void copy_lower_kernel(int nb, const T *__restrict__ src, int64_t lds,
                                  T *__restrict__ dst, int64_t ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < nb && j <= i) dst[i + (int64_t)j * ldd] = src[i + (int64_t)j * lds];
}

// Swaps the strict upper and strict lower triangles of an n×n matrix, in
// place (A <- Aᵀ off the diagonal).  Used for the Val{:U} convention.
template <typename T>
__global__ void swap_triangles_kernel(int64_t n, T *A, int64_t lda)
{
    const int64_t j = (int64_t)blockIdx.y * 16 + threadIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 16 + threadIdx.x;
    if (i < n && j < n && i > j) {
        T a = A[i + j * lda];
        T b = A[j + i * lda];
        A[i + j * lda] = b;
        A[j + i * lda] = a;
    }
}

template <typename T>
int launch_panel(hipStream_t s, int nb, const T *A11, int64_t lda, T *Lw, int64_t ldw, T *B,
                 int64_t ldb, int64_t nrows, int factor, int *info, int64_t info_base)
{
    const int64_t grid = nrows > 0 ? cdiv(nrows, kPanelRB) : 1;
    hipLaunchKernelGGL((chol_panel_kernel<T, kPanelNB, kPanelRB, kPanelThreads>),
                       dim3((unsigned)grid), dim3(kPanelThreads), 0, s, nb, A11, lda, Lw, ldw,
                       B, ldb, nrows, factor, info, info_base);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template <typename T>
int launch_copy_lower(hipStream_t s, int nb, const T *src, int64_t lds, T *dst, int64_t ldd)
{
    hipLaunchKernelGGL((copy_lower_kernel<T>), dim3((unsigned)cdiv(nb, 128), nb), dim3(128), 0,
                       s, nb, src, lds, dst, ldd);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

#define GLA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != GLA_OK) return _rc; \
    } while (0)

// Right-looking blocked Cholesky of the lower triangle with look-ahead.
// Diagonal blocks are factored into the workspace Lw (n × nb) and copied
// into A once every panel that reads them has run.
template <typename T>
int chol_blocked_dev(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int nb, T *Lw,
                     int64_t info_base)
{
    if (n <= 0) return GLA_OK;
    const int64_t nsteps = cdiv(n, nb);
    const bool overlap = nsteps > 2;
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_fork, s));
        GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_fork, 0));
    }
    int64_t k = 0;
    for (int64_t step = 0; step < nsteps; ++step, k += nb) {
        const int kb = (int)std::min<int64_t>(nb, n - k);
        const int64_t below = n - k - kb;
        T *Akk = A + k + k * lda;
        T *A21 = Akk + kb;
        GLA_TRY(launch_panel<T>(s, kb, Akk, lda, Lw + k, n, A21, lda, below, 1, h->dinfo,
                                info_base + k));
        if (below == 0) break;
        const int kb2 = (int)std::min<int64_t>(nb, below);
        const int64_t rest = below - kb2;
        if (overlap && rest > 0) {
            // panel(step) done -> side may update columns beyond the next panel
            GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        }
        if (overlap && step > 0) {
            // the side update of the previous step touched the next panel's columns
            GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
        }
        // Next block column: A[k+kb:n, k+kb:k+kb+kb2] -= A21 * A21[0:kb2]ᵀ (lower part)
        GLA_TRY(gemm<T>(s, 'N', 'T', 1, below, kb2, kb, T(-1), A21, lda, A21, lda, T(1),
                        A21 + kb * lda, lda, 0));
        if (rest > 0) {
            hipStream_t us = overlap ? h->side : s;
            if (overlap) GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_step[0], 0));
            T *X2 = A21 + kb2;
            GLA_TRY(gemm<T>(us, 'N', 'T', 1, rest, rest, kb, T(-1), X2, lda, X2, lda, T(1),
                            X2 + (kb + kb2) * lda, lda, 0));
            if (overlap) GLA_HIP_TRY(hipEventRecord(h->ev_step[1], h->side));
        }
    }
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_join, h->side));
        GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
    }
    // Diagonal blocks from Lw into A.
    k = 0;
    for (int64_t step = 0; step < nsteps; ++step, k += nb) {
        const int kb = (int)std::min<int64_t>(nb, n - k);
        GLA_TRY(launch_copy_lower<T>(s, kb, Lw + k, n, A + k + k * lda, lda));
    }
    return GLA_OK;
}

// B (m×n) <- B L⁻ᵀ, L n×n lower, already factored.  Recursive on L.
template <typename T>
int trsm_rlt_dev(hipStream_t s, int64_t m, int64_t n, const T *L, int64_t ldl, T *B,
                 int64_t ldb)
{
    if (m <= 0 || n <= 0) return GLA_OK;
    if (n <= kPanelNB)
        return launch_panel<T>(s, (int)n, L, ldl, nullptr, 0, B, ldb, m, 0, nullptr, 0);
    // split at a multiple of the panel width
    int64_t n1 = ((n / 2) + kPanelNB - 1) / kPanelNB * kPanelNB;
    if (n1 >= n) n1 = n - kPanelNB;
    GLA_TRY(trsm_rlt_dev<T>(s, m, n1, L, ldl, B, ldb));
    // B2 -= X1 * L21ᵀ, L21 = L[n1:n, 0:n1]
    GLA_TRY(gemm<T>(s, 'N', 'T', 0, m, n - n1, n1, T(-1), B, ldb, L + n1, ldl, T(1),
                    B + n1 * ldb, ldb, 0));
    return trsm_rlt_dev<T>(s, m, n - n1, L + n1 + n1 * ldl, ldl, B + n1 * ldb, ldb);
}

// cholRecursive! (src/cholesky.jl:37-55): halve until n <= leaf, then the
// blocked driver.
template <typename T>
int chol_recursive_dev(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int64_t leaf,
                       T *Lw, int64_t info_base)
{
    if (n <= leaf) return chol_blocked_dev<T>(h, s, n, A, lda, kPanelNB, Lw, info_base);
    int64_t n2 = n / 2;                                              // :44
    if (n2 > kPanelNB) n2 = n2 / kPanelNB * kPanelNB;                // keep blocks aligned
    GLA_TRY(chol_recursive_dev<T>(h, s, n2, A, lda, leaf, Lw, info_base));   // :46
    GLA_TRY(trsm_rlt_dev<T>(s, n - n2, n2, A, lda, A + n2, lda));    // :48 rdiv!
    GLA_TRY(gemm<T>(s, 'N', 'T', 1, n - n2, n - n2, n2, T(-1), A + n2, lda, A + n2, lda, T(1),
                    A + n2 + n2 * lda, lda, 0));                     // :51 rankUpdate!
    return chol_recursive_dev<T>(h, s, n - n2, A + n2 + n2 * lda, lda, leaf, Lw,
                                 info_base + n2);                    // :52
}

template <typename T>
int swap_triangles(hipStream_t s, int64_t n, T *A, int64_t lda)
{
    dim3 grid((unsigned)cdiv(n, 16), (unsigned)cdiv(n, 16));
    hipLaunchKernelGGL((swap_triangles_kernel<T>), grid, dim3(16, 16), 0, s, n, A, lda);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

}  // namespace

// algo: 0 unblocked, 1 blocked (param = blocksize), 2 recursive (param = cutoff)
template <typename T>
int chol_device(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int algo, int64_t param)
{
    if (n == 0) return GLA_OK;
    GLA_TRY(ensure_workspace(h, sizeof(T) * (size_t)n * kPanelNB));
    T *Lw = reinterpret_cast<T *>(h->ws);
    if (algo == 0) {
        // cholUnblocked!: one panel launch when the matrix fits the register
        // resident block, otherwise the same column sweep blocked by 128.
        return chol_blocked_dev<T>(h, s, n, A, lda, kPanelNB, Lw, 0);
    }
    if (algo == 1) {
        int nb = (int)std::min<int64_t>(std::max<int64_t>(param, 1), kPanelNB);
        return chol_blocked_dev<T>(h, s, n, A, lda, nb, Lw, 0);
    }
    // cholRecursive!: the reference recurses to 1×1 (cutoff does not reach
    // the A11 branch, :46); on the device the recursion stops at a leaf of
    // at least 2048 (or `cutoff` if larger) that the blocked driver factors.
    int64_t leaf = std::max<int64_t>(param, 2048);
    return chol_recursive_dev<T>(h, s, n, A, lda, leaf, Lw, 0);
}

template int chol_device<double>(Handle *, hipStream_t, int64_t, double *, int64_t, int, int64_t);
template int chol_device<float>(Handle *, hipStream_t, int64_t, float *, int64_t, int, int64_t);

// ------------------------------------------------------------------ C ABI --
namespace {

template <typename T>
int chol_entry(gla_handle_t hh, char uplo, int64_t n, T *A, int64_t lda, int algo, int64_t param,
               int64_t *info, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (uplo != 'L' && uplo != 'U' && uplo != 'l' && uplo != 'u') {
        set_error("uplo must be 'L' or 'U'");
        return -2;
    }
    if (n < 0) { set_error("n < 0"); return -3; }
    if (!A && n > 0) { set_error("A is null"); return -4; }
    if (lda < std::max<int64_t>(1, n)) { set_error("lda < max(1,n)"); return -5; }
    if (algo == 1 && param < 1) { set_error("blocksize < 1"); return -6; }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, h->dinfo, 0);
    GLA_LAUNCH_CHECK();
    const bool upper = (uplo == 'U' || uplo == 'u');
    if (upper) GLA_TRY(swap_triangles<T>(s, n, A, lda));
    GLA_TRY(chol_device<T>(h, s, n, A, lda, algo, param));
    if (upper) GLA_TRY(swap_triangles<T>(s, n, A, lda));
    if (info) {
        int hv = 0;
        GLA_HIP_TRY(hipMemcpyAsync(&hv, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
        GLA_HIP_TRY(hipStreamSynchronize(s));
        *info = hv;
    }
    return GLA_OK;
}

template <typename T>
int rank_update_entry(gla_handle_t hh, char uplo, int64_t n, int64_t k, T alpha, const T *A,
                      int64_t lda, T *C, int64_t ldc, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (uplo != 'L' && uplo != 'U') { set_error("uplo must be 'L' or 'U'"); return -2; }
    if (n < 0) return -3;
    if (k < 0) return -4;
    if (lda < std::max<int64_t>(1, n)) { set_error("lda < max(1,n)"); return -7; }
    if (ldc < std::max<int64_t>(1, n)) { set_error("ldc < max(1,n)"); return -9; }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    if (uplo == 'L') return gemm<T>(s, 'N', 'T', 1, n, n, k, alpha, A, lda, A, lda, T(1), C, ldc, 0);
    return gemm<T>(s, 'N', 'T', 2, n, n, k, alpha, A, lda, A, lda, T(1), C, ldc, 0);
}

template <typename T>
int trsm_entry(gla_handle_t hh, int64_t m, int64_t n, const T *L, int64_t ldl, T *B, int64_t ldb,
               void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (ldl < std::max<int64_t>(1, n)) { set_error("ldl < max(1,n)"); return -5; }
    if (ldb < std::max<int64_t>(1, m)) { set_error("ldb < max(1,m)"); return -7; }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    return trsm_rlt_dev<T>(s, m, n, L, ldl, B, ldb);
}

}  // namespace
}  // namespace gla

using gla::chol_entry;

extern "C" {

int gla_chol_unblocked_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda,
                         int64_t *info, void *stream)
{
    return chol_entry<double>(h, uplo, n, A, lda, 0, 0, info, stream);
}
int gla_chol_unblocked_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda,
                         int64_t *info, void *stream)
{
    return chol_entry<float>(h, uplo, n, A, lda, 0, 0, info, stream);
}
int gla_chol_blocked_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda,
                       int64_t blocksize, int64_t *info, void *stream)
{
    return chol_entry<double>(h, uplo, n, A, lda, 1, blocksize, info, stream);
}
int gla_chol_blocked_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda,
                       int64_t blocksize, int64_t *info, void *stream)
{
    return chol_entry<float>(h, uplo, n, A, lda, 1, blocksize, info, stream);
}
int gla_chol_recursive_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda,
                         int64_t cutoff, int64_t *info, void *stream)
{
    return chol_entry<double>(h, uplo, n, A, lda, 2, cutoff, info, stream);
}
int gla_chol_recursive_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda,
                         int64_t cutoff, int64_t *info, void *stream)
{
    return chol_entry<float>(h, uplo, n, A, lda, 2, cutoff, info, stream);
}
int gla_rank_update_d(gla_handle_t h, char uplo, int64_t n, int64_t k, double alpha,
                      const double *A, int64_t lda, double *C, int64_t ldc, void *stream)
{
    return gla::rank_update_entry<double>(h, uplo, n, k, alpha, A, lda, C, ldc, stream);
}
int gla_rank_update_s(gla_handle_t h, char uplo, int64_t n, int64_t k, float alpha,
                      const float *A, int64_t lda, float *C, int64_t ldc, void *stream)
{
    return gla::rank_update_entry<float>(h, uplo, n, k, alpha, A, lda, C, ldc, stream);
}
int gla_trsm_rlt_d(gla_handle_t h, int64_t m, int64_t n, const double *L, int64_t ldl,
                   double *B, int64_t ldb, void *stream)
{
    return gla::trsm_entry<double>(h, m, n, L, ldl, B, ldb, stream);
}
int gla_trsm_rlt_s(gla_handle_t h, int64_t m, int64_t n, const float *L, int64_t ldl, float *B,
                   int64_t ldb, void *stream)
{
    return gla::trsm_entry<float>(h, m, n, L, ldl, B, ldb, stream);
}

}  // extern "C"
