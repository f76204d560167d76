// Complex (ComplexF64 / ComplexF32) Cholesky: cholUnblocked!, cholBlocked!,
// cholRecursive! (src/cholesky.jl:3-55) for Hermitian positive definite A,
// with the rank updates the reference sends to BLAS her!/herk!
// (src/juliaBLAS.jl:41-48,78-86), and rankUpdate! itself for complex C.
//
// Storage: Julia's layout, column-major with interleaved (re, im) pairs.
//
// Device plan, one step of width kb ≤ 64 at column k:
//   - the diagonal block: one workgroup, the lower triangle resident in LDS,
//     column by column as cholUnblocked! does it (complex square root of the
//     pivot, the column scaled by 1/real(pivot), her! on the rest with the
//     diagonal kept real), then L₁₁⁻¹ by forward substitution in LDS;
//   - the panel: A21 <- A21 (L₁₁⁻¹)ᴴ (rdiv!(A21, LowerTriangular(A11)')) as
//     a complex MFMA product;
//   - the trailing update A22 −= A21 A21ᴴ (herk!) on the same complex MFMA
//     kernel restricted to the lower triangle, imaginary parts of the
//     diagonal set to zero as herk! leaves them.
// Complex products run as four real MFMAs (re·re, −im·im, re·im, im·re) on
// re/im planes staged separately in LDS.  The factor agrees with the
// reference's to rounding for any step width.
#include <algorithm>

#include "gla_internal.hpp"

namespace gla {

namespace {

#define GLA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != GLA_OK) return _rc; \
    } while (0)

constexpr int kZB = 64;      // diagonal block and GEMM tile (complex elements)
constexpr int kZK = 16;      // K-slab of the complex GEMM

// Complex GEMM: C = α op(A) op(B) + β C (α, β real), op ∈ {N, C (conjugate
// transpose)}.  MODE 0 all of C, 1 elements with row + diag_off ≥ col,
// 2 elements with row + diag_off ≤ col.  HERM: elements on the diagonal
// (row + diag_off == col) get a zero imaginary part.  64×64 tiles, 256
// threads as 2×2 waves of 32×32 (2×2 blocks of 16×16).
template <typename T, bool TA, bool TB, int MODE, bool HERM>
__global__ __launch_bounds__(256) void zgemm_kernel(int64_t M, int64_t N, int64_t K, T alpha,
                                                    const T *__restrict__ A, int64_t lda,
                                                    const T *__restrict__ B, int64_t ldb, T beta,
                                                    T *__restrict__ C, int64_t ldc,
                                                    int64_t diag_off)
{
    typedef PanelMfma<T> Mf;
    typedef typename Mf::acc_t acc_t;
    constexpr int LD = kZB + 4;
    __shared__ T Ar[kZK * LD], Ai[kZK * LD], Br[kZK * LD], Bi[kZK * LD];
    const int64_t m0 = (int64_t)blockIdx.x * kZB, n0 = (int64_t)blockIdx.y * kZB;
    if (MODE == 1 && m0 + kZB - 1 + diag_off < n0) return;        // tile above the diagonal
    if (MODE == 2 && m0 + diag_off > n0 + kZB - 1) return;        // tile below it
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    acc_t accr[2][2], acci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accr[i][j] = acc_t{0, 0, 0, 0};
            acci[i][j] = acc_t{0, 0, 0, 0};
        }
    for (int64_t k0 = 0; k0 < K; k0 += kZK) {
        T ar[4], ai[4], br[4], bi[4];
        // op(A) (x, k): N — A[x + k·lda]; C — conj(A[k + x·lda])
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            int x, k;
            if (!TA) { x = t & 63; k = (t >> 6) + 4 * l; }
            else { k = t & 15; x = (t >> 4) + 16 * l; }
            const int64_t gx = m0 + x, gk = k0 + k;
            ar[l] = T(0);
            ai[l] = T(0);
            if (gx < M && gk < K) {
                const T *p = A + 2 * (TA ? gk + gx * lda : gx + gk * lda);
                ar[l] = p[0];
                ai[l] = TA ? -p[1] : p[1];
            }
        }
        // op(B) (k, y): N — B[k + y·ldb]; C — conj(B[y + k·ldb])
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            int y, k;
            if (TB) { y = t & 63; k = (t >> 6) + 4 * l; }
            else { k = t & 15; y = (t >> 4) + 16 * l; }
            const int64_t gy = n0 + y, gk = k0 + k;
            br[l] = T(0);
            bi[l] = T(0);
            if (gy < N && gk < K) {
                const T *p = B + 2 * (TB ? gy + gk * ldb : gk + gy * ldb);
                br[l] = p[0];
                bi[l] = TB ? -p[1] : p[1];
            }
        }
        __syncthreads();   // the previous slab is consumed
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            int x, k;
            if (!TA) { x = t & 63; k = (t >> 6) + 4 * l; }
            else { k = t & 15; x = (t >> 4) + 16 * l; }
            Ar[k * LD + x] = ar[l];
            Ai[k * LD + x] = ai[l];
            int y, kb;
            if (TB) { y = t & 63; kb = (t >> 6) + 4 * l; }
            else { kb = t & 15; y = (t >> 4) + 16 * l; }
            Br[kb * LD + y] = br[l];
            Bi[kb * LD + y] = bi[l];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kZK / 4; ++kk) {
            const int kr = kk * 4 + (lane >> 4);
            T a_r[2], a_i[2], b_r[2], b_i[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a_r[i] = Ar[kr * LD + wm * 32 + i * 16 + (lane & 15)];
                a_i[i] = Ai[kr * LD + wm * 32 + i * 16 + (lane & 15)];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                b_r[j] = Br[kr * LD + wn * 32 + j * 16 + (lane & 15)];
                b_i[j] = Bi[kr * LD + wn * 32 + j * 16 + (lane & 15)];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    accr[i][j] = Mf::mma(a_r[i], b_r[j], accr[i][j]);
                    accr[i][j] = Mf::mma(-a_i[i], b_i[j], accr[i][j]);
                    acci[i][j] = Mf::mma(a_r[i], b_i[j], acci[i][j]);
                    acci[i][j] = Mf::mma(a_i[i], b_r[j], acci[i][j]);
                }
        }
    }
    const bool beta0 = beta == T(0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gr = m0 + wm * 32 + i * 16 + Mf::row(lane, r);
                const int64_t gc = n0 + wn * 32 + j * 16 + (lane & 15);
                if (gr >= M || gc >= N) continue;
                if (MODE == 1 && gr + diag_off < gc) continue;
                if (MODE == 2 && gr + diag_off > gc) continue;
                T *c = C + 2 * (gr + gc * ldc);
                T re = alpha * accr[i][j][r], im = alpha * acci[i][j][r];
                if (!beta0) {
                    re += beta * c[0];
                    im += beta * c[1];
                }
                if (HERM && gr + diag_off == gc) im = T(0);
                c[0] = re;
                c[1] = im;
            }
}

template <typename T, bool TA, bool TB, int MODE, bool HERM>
int zgemm_launch(hipStream_t s, int64_t M, int64_t N, int64_t K, T alpha, const T *A, int64_t lda,
                 const T *B, int64_t ldb, T beta, T *C, int64_t ldc, int64_t diag_off)
{
    if (M <= 0 || N <= 0) return GLA_OK;
    const dim3 grid((unsigned)cdiv(M, kZB), (unsigned)cdiv(N, kZB));
    hipLaunchKernelGGL((zgemm_kernel<T, TA, TB, MODE, HERM>), grid, dim3(256), 0, s, M, N, K,
                       alpha, A, lda, B, ldb, beta, C, ldc, diag_off);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// Square root of x + iy with the branch of the complex sqrt (principal
// value, the imaginary part's sign following y): the textbook formula
// with |z| = hypot(x, y).
template <typename T>
__device__ __forceinline__ void csqrt_dev(T x, T y, T &re, T &im)
{
    if (x == T(0) && y == T(0)) {
        re = T(0);
        im = y;
        return;
    }
    const T r = hypot(x, y);
    if (x >= T(0)) {
        const T s = sqrt((r + x) * T(0.5));
        re = s;
        im = y / (T(2) * s);
    } else {
        const T s = sqrt((r - x) * T(0.5));
        re = fabs(y) / (T(2) * s);
        im = copysign(s, y);
    }
}

__device__ __forceinline__ int zpk(int i, int j) { return j * kZB - (j * (j - 1)) / 2 + (i - j); }

// cholUnblocked! of the nb×nb (nb ≤ 64) complex block at A, then
// Linv (64×64, ld 64, zero above the diagonal and past nb) = L⁻¹.
// *info gets info_base + k + 1 for the first pivot with a negative real part.
template <typename T>
__global__ __launch_bounds__(256) void zchol_factor_kernel(int nb, T *__restrict__ A, int64_t lda,
                                                           T *__restrict__ Linv, int *info,
                                                           int64_t info_base)
{
    __shared__ T Sr[kZB * (kZB + 1) / 2], Si[kZB * (kZB + 1) / 2];
    __shared__ T Xr[kZB * kZB], Xi[kZB * kZB];
    const int t = threadIdx.x;
    const int i = t & 63, p = t >> 6;          // row, column class mod 4
    if (i < nb)
        for (int j = p; j <= i; j += 4) {
            const T *a = A + 2 * (i + (int64_t)j * lda);
            Sr[zpk(i, j)] = a[0];
            Si[zpk(i, j)] = a[1];
        }
    for (int k = 0; k < nb; ++k) {
        __syncthreads();
        const T pr = Sr[zpk(k, k)], pi = Si[zpk(k, k)];
        T lr, li;
        csqrt_dev(pr, pi, lr, li);                                   // :5
        __syncthreads();
        if (t == 0) {
            Sr[zpk(k, k)] = lr;
            Si[zpk(k, k)] = li;
            if (pr < T(0) && *info == 0) *info = (int)(info_base + k + 1);
        }
        const T rs = T(1) / lr;                                      // :8 inv(real(A[1,1]))
        if (p == 0 && i > k && i < nb) {
            Sr[zpk(i, k)] *= rs;
            Si[zpk(i, k)] *= rs;
        }
        __syncthreads();
        if (i > k && i < nb) {                                       // :11 her!
            const T air = Sr[zpk(i, k)], aii = Si[zpk(i, k)];
            int j = k + 1;
            while ((j & 3) != p) ++j;
            for (; j <= i; j += 4) {
                const T ajr = Sr[zpk(j, k)], aji = Si[zpk(j, k)];
                // A[i,j] −= a_i conj(a_j)
                Sr[zpk(i, j)] -= air * ajr + aii * aji;
                if (j == i) Si[zpk(i, j)] = T(0);
                else Si[zpk(i, j)] -= aii * ajr - air * aji;
            }
        }
    }
    __syncthreads();
    if (i < nb)
        for (int j = p; j <= i; j += 4) {
            T *a = A + 2 * (i + (int64_t)j * lda);
            a[0] = Sr[zpk(i, j)];
            a[1] = Si[zpk(i, j)];
        }
    // L⁻¹ column c by forward substitution (thread c), in LDS
    if (t < kZB) {
        const int c = t;
        for (int r = 0; r < kZB; ++r) {
            T xr = T(0), xi = T(0);
            if (r >= c && r < nb && c < nb) {
                T sr = r == c ? T(1) : T(0), si = T(0);
                for (int l = c; l < r; ++l) {
                    const T lr_ = Sr[zpk(r, l)], li_ = Si[zpk(r, l)];
                    const T yr = Xr[l * kZB + c], yi = Xi[l * kZB + c];
                    sr -= lr_ * yr - li_ * yi;
                    si -= lr_ * yi + li_ * yr;
                }
                // divide by L[r,r]
                const T dr = Sr[zpk(r, r)], di = Si[zpk(r, r)];
                const T den = dr * dr + di * di;
                xr = (sr * dr + si * di) / den;
                xi = (si * dr - sr * di) / den;
            }
            Xr[r * kZB + c] = xr;
            Xi[r * kZB + c] = xi;
        }
    }
    __syncthreads();
    for (int e = t; e < kZB * kZB; e += 256) {
        const int r = e % kZB, c = e / kZB;
        Linv[2 * e] = Xr[r * kZB + c];
        Linv[2 * e + 1] = Xi[r * kZB + c];
    }
}

template <typename T>
__global__ void zcopy_kernel(int64_t rows, int64_t cols, const T *S, int64_t lds, T *D, int64_t ldd)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i < rows && j < cols) {
        D[2 * (i + j * ldd)] = S[2 * (i + j * lds)];
        D[2 * (i + j * ldd) + 1] = S[2 * (i + j * lds) + 1];
    }
}

// (A_ij, A_ji) <- (conj(A_ji), conj(A_ij)) off the diagonal: the Hermitian
// matrix kept in the other triangle.
template <typename T>
__global__ void zswap_conj_kernel(int64_t n, T *A, int64_t lda)
{
    const int64_t j = (int64_t)blockIdx.y * 16 + threadIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 16 + threadIdx.x;
    if (i < n && j < n && i > j) {
        T *a = A + 2 * (i + j * lda), *b = A + 2 * (j + i * lda);
        const T ar = a[0], ai = a[1];
        a[0] = b[0];
        a[1] = -b[1];
        b[0] = ar;
        b[1] = -ai;
    }
}

// Right-looking complex Cholesky with nb-wide steps (nb ≤ 64).
template <typename T>
int zchol_dev(hipStream_t s, int64_t n, T *A, int64_t lda, int nb, T *Linv, T *W, int64_t ldw,
              int *info)
{
    for (int64_t k = 0; k < n; k += nb) {
        const int kb = (int)std::min<int64_t>(nb, n - k);
        const int64_t below = n - k - kb;
        T *Akk = A + 2 * (k + k * lda);
        T *A21 = Akk + 2 * kb;
        hipLaunchKernelGGL((zchol_factor_kernel<T>), dim3(1), dim3(256), 0, s, kb, Akk, lda, Linv,
                           info, k);
        GLA_LAUNCH_CHECK();
        if (below == 0) break;
        hipLaunchKernelGGL((zcopy_kernel<T>), dim3((unsigned)cdiv(below, 256), (unsigned)kb),
                           dim3(256), 0, s, below, (int64_t)kb, (const T *)A21, lda, W, ldw);
        GLA_LAUNCH_CHECK();
        // A21 = W (L₁₁⁻¹)ᴴ
        GLA_TRY((zgemm_launch<T, false, true, 0, false>(s, below, kb, kb, T(1), W, ldw, Linv, kZB,
                                                        T(0), A21, lda, 0)));
        // A22 −= A21 A21ᴴ (lower, herk!)
        GLA_TRY((zgemm_launch<T, false, true, 1, true>(s, below, below, kb, T(-1), A21, lda, A21,
                                                       lda, T(1), A21 + 2 * kb * lda, lda, 0)));
    }
    return GLA_OK;
}

// algo 0 unblocked, 1 blocked (param = blocksize), 2 recursive (param = cutoff)
template <typename T>
int zchol_entry(gla_handle_t hh, char uplo, int64_t n, void *Av, int64_t lda, int algo,
                int64_t param, int64_t *info, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    T *A = reinterpret_cast<T *>(Av);
    if (!h) { set_error("null handle"); return -1; }
    if (uplo != 'L' && uplo != 'U') { set_error("uplo must be 'L' or 'U'"); return -2; }
    if (n < 0) { set_error("n < 0"); return -3; }
    if (!A && n > 0) { set_error("A is null"); return -4; }
    if (lda < std::max<int64_t>(1, n)) { set_error("lda < max(1,n)"); return -5; }
    if (algo == 1 && param < 1) { set_error("blocksize < 1"); return -6; }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, h->dinfo, 0);
    GLA_LAUNCH_CHECK();
    if (n > 0) {
        const int64_t ldw = std::max<int64_t>(n, 1);
        GLA_TRY(ensure_workspace(h, sizeof(T) * 2 * ((size_t)kZB * kZB + (size_t)ldw * kZB + 64)));
        T *Linv = reinterpret_cast<T *>(h->ws), *W = Linv + 2 * kZB * kZB + 64;
        const int nb = algo == 1 ? (int)std::min<int64_t>(param, kZB) : kZB;
        const dim3 tg((unsigned)cdiv(n, 16), (unsigned)cdiv(n, 16));
        if (uplo == 'U') {
            hipLaunchKernelGGL((zswap_conj_kernel<T>), tg, dim3(16, 16), 0, s, n, A, lda);
            GLA_LAUNCH_CHECK();
        }
        GLA_TRY(zchol_dev<T>(s, n, A, lda, nb, Linv, W, ldw, h->dinfo));
        if (uplo == 'U') {
            hipLaunchKernelGGL((zswap_conj_kernel<T>), tg, dim3(16, 16), 0, s, n, A, lda);
            GLA_LAUNCH_CHECK();
        }
    }
    if (info) {
        int hv = 0;
        GLA_HIP_TRY(hipMemcpyAsync(&hv, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
        GLA_HIP_TRY(hipStreamSynchronize(s));
        *info = hv;
    }
    return GLA_OK;
}

// rankUpdate!(C::Hermitian, A, α) for complex C (herk!): the uplo triangle
// of C += α A Aᴴ, the diagonal kept real.
template <typename T>
int zherk_entry(gla_handle_t hh, char uplo, int64_t n, int64_t k, T alpha, const void *Av,
                int64_t lda, void *Cv, int64_t ldc, void *stream)
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
    const T *A = reinterpret_cast<const T *>(Av);
    T *Cm = reinterpret_cast<T *>(Cv);
    if (uplo == 'L')
        return zgemm_launch<T, false, true, 1, true>(s, n, n, k, alpha, A, lda, A, lda, T(1), Cm,
                                                     ldc, 0);
    return zgemm_launch<T, false, true, 2, true>(s, n, n, k, alpha, A, lda, A, lda, T(1), Cm, ldc,
                                                 0);
}

}  // namespace
}  // namespace gla

extern "C" {

int gla_chol_unblocked_z(gla_handle_t h, char uplo, int64_t n, gla_cdouble *A, int64_t lda,
                         int64_t *info, void *stream)
{
    return gla::zchol_entry<double>(h, uplo, n, A, lda, 0, 0, info, stream);
}
int gla_chol_unblocked_c(gla_handle_t h, char uplo, int64_t n, gla_cfloat *A, int64_t lda,
                         int64_t *info, void *stream)
{
    return gla::zchol_entry<float>(h, uplo, n, A, lda, 0, 0, info, stream);
}
int gla_chol_blocked_z(gla_handle_t h, char uplo, int64_t n, gla_cdouble *A, int64_t lda,
                       int64_t blocksize, int64_t *info, void *stream)
{
    return gla::zchol_entry<double>(h, uplo, n, A, lda, 1, blocksize, info, stream);
}
int gla_chol_blocked_c(gla_handle_t h, char uplo, int64_t n, gla_cfloat *A, int64_t lda,
                       int64_t blocksize, int64_t *info, void *stream)
{
    return gla::zchol_entry<float>(h, uplo, n, A, lda, 1, blocksize, info, stream);
}
int gla_chol_recursive_z(gla_handle_t h, char uplo, int64_t n, gla_cdouble *A, int64_t lda,
                         int64_t cutoff, int64_t *info, void *stream)
{
    return gla::zchol_entry<double>(h, uplo, n, A, lda, 2, cutoff, info, stream);
}
int gla_chol_recursive_c(gla_handle_t h, char uplo, int64_t n, gla_cfloat *A, int64_t lda,
                         int64_t cutoff, int64_t *info, void *stream)
{
    return gla::zchol_entry<float>(h, uplo, n, A, lda, 2, cutoff, info, stream);
}
int gla_rank_update_z(gla_handle_t h, char uplo, int64_t n, int64_t k, double alpha,
                      const gla_cdouble *A, int64_t lda, gla_cdouble *C, int64_t ldc, void *stream)
{
    return gla::zherk_entry<double>(h, uplo, n, k, alpha, A, lda, C, ldc, stream);
}
int gla_rank_update_c(gla_handle_t h, char uplo, int64_t n, int64_t k, float alpha,
                      const gla_cfloat *A, int64_t lda, gla_cfloat *C, int64_t ldc, void *stream)
{
    return gla::zherk_entry<float>(h, uplo, n, k, alpha, A, lda, C, ldc, stream);
}

}  // extern "C"
