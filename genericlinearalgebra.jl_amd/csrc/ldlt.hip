// Blocked LDLᴴ without pivoting for real symmetric matrices (FP64/FP32):
// ldlt!(Hermitian(A, uplo), blocksize), src/ldlt.jl:150-157, which runs
// _ldlt_lower_blocked! (:80-104) or _ldlt_upper_blocked! (:121-148).
//
// Lower (A = L D Lᵀ, L unit lower stored below the diagonal, D on it), one
// step of width kb at column k:
//   - the kb×kb diagonal block: _ldlt_lower! (:16-33) in one workgroup as
//     16×16 tiles (the diagonal tile factored column by column in registers,
//     the rest on MFMA; ldlt_factor_block in chol.hip); the same launch
//     writes the inverses of the 16×16 diagonal tiles of its unit factor;
//   - the panel: W = A21 L11⁻ᵀ (rdiv! by UnitLowerTriangular(A11)', :86) on
//     the Cholesky's MFMA panel solve with those unit tile inverses, then
//     one pass keeping W and writing L21 = W D⁻¹ (rdiv! by Diagonal(d₁),
//     :87);
//   - the trailing update A22 −= L21 D L21ᵀ = W L21ᵀ (the loop at :88-95)
//     on the MFMA GEMM restricted to the lower triangle.
// Large matrices take 512-wide steps with depth-1 look-ahead (the next
// block column on the caller's stream, the rest of the update on the side
// stream), the 512-wide diagonal block itself by 128-wide steps.  The
// factors do not depend on the block width, up to rounding.
//
// Upper (A = Uᵀ D U): the matrix is transposed in place, factored as lower
// and transposed back (U = Lᵀ), the strict lower triangle keeping its input.
// Zero pivots divide by zero and propagate Inf/NaN, as in the reference.
#include <algorithm>

#include "gla_internal.hpp"

namespace gla {

namespace {

#define GLA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != GLA_OK) return _rc; \
    } while (0)

constexpr int kNB = 128;    // largest diagonal block of one factor launch

// W = A21 (the panel solve's output A21 L11⁻ᵀ) and A21 <- W D⁻¹, D the
// diagonal of the factored A11 (:87).
template <typename T>
__global__ void ldlt_scale_kernel(int64_t rows, const T *__restrict__ A11, T *__restrict__ A21,
                                  int64_t lda, T *__restrict__ W, int64_t ldw)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= rows) return;
    const T d = A11[j + j * lda];
    const T w = A21[i + j * lda];
    W[i + j * ldw] = w;
    A21[i + j * lda] = w / d;
}

// A <- Aᵀ off the diagonal (the strict triangles swap places).
template <typename T>
__global__ void transpose_tri_kernel(int64_t n, T *A, int64_t lda)
{
    const int64_t j = (int64_t)blockIdx.y * 16 + threadIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 16 + threadIdx.x;
    if (i < n && j < n && i > j) {
        const T a = A[i + j * lda], b = A[j + i * lda];
        A[i + j * lda] = b;
        A[j + i * lda] = a;
    }
}

// _ldlt_lower! of the diagonal block (:84), nb ≤ 128, and the inverses of
// its unit 16×16 diagonal tiles: the MFMA-tiled kernel beside the Cholesky's
// diagonal factor.
template <typename T>
int launch_factor(hipStream_t s, int nb, T *A, int64_t lda, T *Linv)
{
    return ldlt_factor_block<T>(s, nb, A, lda, Linv);
}

template <typename T>
int launch_scale(hipStream_t s, int64_t rows, int kb, const T *A11, T *A21, int64_t lda, T *W,
                 int64_t ldw)
{
    if (rows <= 0) return GLA_OK;
    hipLaunchKernelGGL((ldlt_scale_kernel<T>), dim3((unsigned)cdiv(rows, 256), (unsigned)kb),
                       dim3(256), 0, s, rows, A11, A21, lda, W, ldw);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// Right-looking LDLᵀ with nb-wide steps (nb ≤ 128).  Step k's tile
// inverses land at Linv + (k/16)·256, so that after the call Linv holds
// those of the whole unit factor (the look-ahead driver's panel solve
// reads them).  W0/W1 (ldw ≥ n rows, nb columns): the panel's W, alternating
// by step so the side stream may still read the previous one.
template <typename T>
int ldlt_blocked_dev(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int nb, T *Linv,
                     T *W0, T *W1, int64_t ldw, bool use_side)
{
    if (n <= 0) return GLA_OK;
    const int64_t nsteps = cdiv(n, nb);
    const bool overlap = use_side && nsteps > 2;
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
        T *Li = Linv + (k / 16) * 256;
        T *Wk = (step & 1) ? W1 : W0;
        GLA_TRY(launch_factor<T>(s, kb, Akk, lda, Li));                         // :84
        if (below == 0) break;
        GLA_TRY(chol_solve_block<T>(s, kb, Akk, lda, Li, A21, lda, below));     // :86
        GLA_TRY(launch_scale<T>(s, below, kb, Akk, A21, lda, Wk, ldw));          // :87
        const int kb2 = (int)std::min<int64_t>(nb, below);
        const int64_t rest = below - kb2;
        if (overlap && rest > 0) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (overlap && step > 0) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
        // :88-95 on the next block column first (look-ahead) ...
        GLA_TRY(gemm<T>(s, 'N', 'T', 1, below, kb2, kb, T(-1), Wk, ldw, A21, lda, T(1),
                        A21 + kb * lda, lda, 0));
        // ... and on the rest of A22 beside the next steps
        if (rest > 0) {
            hipStream_t us = overlap ? h->side : s;
            if (overlap) GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_step[0], 0));
            GLA_TRY(gemm<T>(us, 'N', 'T', 1, rest, rest, kb, T(-1), Wk + kb2, ldw, A21 + kb2, lda,
                            T(1), A21 + kb2 + (kb + kb2) * lda, lda, 0));
            if (overlap) GLA_HIP_TRY(hipEventRecord(h->ev_step[1], h->side));
        }
    }
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_join, h->side));
        GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
    }
    return GLA_OK;
}

// NB-wide steps (NB a multiple of 128) with depth-1 look-ahead; each
// diagonal block by the 128-wide driver on the caller's stream, its panel
// solved by the recursive right-side solve with the unit tile inverses.
template <typename T>
int ldlt_la_dev(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int NB, T *Linv, T *W0,
                T *W1, int64_t ldw, T *Wd0, T *Wd1)
{
    if (n <= NB) return ldlt_blocked_dev<T>(h, s, n, A, lda, kNB, Linv, W0, W1, ldw, true);
    const int64_t nsteps = cdiv(n, NB);
    const bool overlap = nsteps > 2;
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_fork, s));
        GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_fork, 0));
    }
    int64_t k = 0;
    for (int64_t step = 0; step < nsteps; ++step, k += NB) {
        const int64_t kb = std::min<int64_t>(NB, n - k);
        const int64_t below = n - k - kb;
        T *Akk = A + k + k * lda;
        T *A21 = Akk + kb;
        T *Wk = (step & 1) ? W1 : W0;
        GLA_TRY(ldlt_blocked_dev<T>(h, s, kb, Akk, lda, kNB, Linv, Wd0, Wd1, NB, false));
        if (below == 0) break;
        GLA_TRY(trsm_rlt<T>(s, below, kb, Akk, lda, Linv, A21, lda));
        GLA_TRY(launch_scale<T>(s, below, (int)kb, Akk, A21, lda, Wk, ldw));
        const int64_t kb2 = std::min<int64_t>(NB, below);
        const int64_t rest = below - kb2;
        if (overlap && step > 0) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
        GLA_TRY(gemm<T>(s, 'N', 'T', 1, below, kb2, kb, T(-1), Wk, ldw, A21, lda, T(1),
                        A21 + kb * lda, lda, 0));
        if (overlap && rest > 0) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (rest > 0) {
            hipStream_t us = overlap ? h->side : s;
            if (overlap) GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_step[0], 0));
            GLA_TRY(gemm<T>(us, 'N', 'T', 1, rest, rest, kb, T(-1), Wk + kb2, ldw, A21 + kb2, lda,
                            T(1), A21 + kb2 + (kb + kb2) * lda, lda, 0));
            if (overlap) GLA_HIP_TRY(hipEventRecord(h->ev_step[1], h->side));
        }
    }
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_join, h->side));
        GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
    }
    return GLA_OK;
}

template <typename T>
int ldlt_entry(gla_handle_t hh, char uplo, int64_t n, T *A, int64_t lda, int64_t blocksize,
               void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (uplo != 'L' && uplo != 'U') { set_error("uplo must be 'L' or 'U'"); return -2; }
    if (n < 0) { set_error("n < 0"); return -3; }
    if (!A && n > 0) { set_error("A is null"); return -4; }
    if (lda < std::max<int64_t>(1, n)) { set_error("lda < max(1,n)"); return -5; }
    if (blocksize < 1) { set_error("blocksize < 1"); return -6; }
    if (n == 0) return GLA_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    const int NB = (int)h->ldlt_block;
    const int64_t wcols = NB > kNB ? NB : kNB;
    const int64_t ldw = (n + 1) / 2 * 2;
    const size_t linv = (size_t)(cdiv(n, 16) + 8) * 256;
    const size_t elems = linv + 2 * (size_t)ldw * wcols + 2 * (size_t)wcols * kNB + 64;
    GLA_TRY(ensure_workspace(h, elems * sizeof(T)));
    T *Linv = reinterpret_cast<T *>(h->ws);
    T *W0 = Linv + linv + (32 / sizeof(T)), *W1 = W0 + (size_t)ldw * wcols;
    T *Wd0 = W1 + (size_t)ldw * wcols, *Wd1 = Wd0 + (size_t)wcols * kNB;
    const dim3 tg((unsigned)cdiv(n, 16), (unsigned)cdiv(n, 16));
    if (uplo == 'U') {
        hipLaunchKernelGGL((transpose_tri_kernel<T>), tg, dim3(16, 16), 0, s, n, A, lda);
        GLA_LAUNCH_CHECK();
    }
    if (NB > kNB)
        GLA_TRY(ldlt_la_dev<T>(h, s, n, A, lda, NB, Linv, W0, W1, ldw, Wd0, Wd1));
    else
        GLA_TRY(ldlt_blocked_dev<T>(h, s, n, A, lda, kNB, Linv, W0, W1, ldw, true));
    if (uplo == 'U') {
        hipLaunchKernelGGL((transpose_tri_kernel<T>), tg, dim3(16, 16), 0, s, n, A, lda);
        GLA_LAUNCH_CHECK();
    }
    return GLA_OK;
}

}  // namespace
}  // namespace gla

extern "C" {

int gla_ldlt_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda, int64_t blocksize,
               void *stream)
{
    return gla::ldlt_entry<double>(h, uplo, n, A, lda, blocksize, stream);
}
int gla_ldlt_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda, int64_t blocksize,
               void *stream)
{
    return gla::ldlt_entry<float>(h, uplo, n, A, lda, blocksize, stream);
}

}  // extern "C"
