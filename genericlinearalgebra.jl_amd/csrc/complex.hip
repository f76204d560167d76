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
                                                    int64_t diag_off, int64_t kchunk,
                                                    T *__restrict__ part)
{
    // split K (grid z): this workgroup contracts [k_begin, k_end) and, with
    // `part`, writes α·(its partial product) to part[z] (M×N, ld M) for
    // zsplit_reduce_kernel instead of touching C
    const int64_t k_begin = (int64_t)blockIdx.z * kchunk;
    const int64_t k_end = k_begin + kchunk < K ? k_begin + kchunk : K;
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
    for (int64_t k0 = k_begin; k0 < k_end; k0 += kZK) {
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
            if (gx < M && gk < k_end) {
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
            if (gy < N && gk < k_end) {
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
                if (part) {
                    T *q = part + 2 * ((int64_t)blockIdx.z * M * N + gr + gc * M);
                    q[0] = alpha * accr[i][j][r];
                    q[1] = alpha * acci[i][j][r];
                    continue;
                }
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

// C = β C + Σ_z part[z] (M×N complex, ld M), the partials in a fixed order.
template <typename T>
__global__ void zsplit_reduce_kernel(int64_t M, int64_t N, int S, const T *__restrict__ part,
                                     T beta, T *__restrict__ C, int64_t ldc)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M * N) return;
    const int64_t r = e % M, c = e / M;
    T re = T(0), im = T(0);
    for (int z = 0; z < S; ++z) {
        re += part[2 * ((int64_t)z * M * N + e)];
        im += part[2 * ((int64_t)z * M * N + e) + 1];
    }
    T *p = C + 2 * (r + c * ldc);
    if (beta != T(0)) {
        re += beta * p[0];
        im += beta * p[1];
    }
    p[0] = re;
    p[1] = im;
}

// ws (ws_elems real elements, may be null): room for split-K partials.  A
// product with few 64×64 tiles and a long K (Vᴴ A in the compact-WY update,
// the Gram products of T) is split along K over grid z into ws and summed
// by zsplit_reduce_kernel, so that it spreads over the chip.
template <typename T, bool TA, bool TB, int MODE, bool HERM>
int zgemm_launch(hipStream_t s, int64_t M, int64_t N, int64_t K, T alpha, const T *A, int64_t lda,
                 const T *B, int64_t ldb, T beta, T *C, int64_t ldc, int64_t diag_off,
                 T *ws = nullptr, int64_t ws_elems = 0)
{
    if (M <= 0 || N <= 0) return GLA_OK;
    const int64_t tiles = cdiv(M, kZB) * cdiv(N, kZB);
    int64_t S = 1;
    if (MODE == 0 && !HERM && ws && tiles < 128 && K >= 1024) {
        S = std::min<int64_t>(std::min<int64_t>(cdiv(K, 256), std::max<int64_t>(1, 256 / tiles)),
                              64);
        while (S > 1 && 2 * S * M * N > ws_elems) --S;
    }
    int64_t kchunk = K > 0 ? K : 1;
    if (S > 1) {
        kchunk = cdiv(cdiv(K, S), kZK) * kZK;
        S = cdiv(K, kchunk);
    }
    const dim3 grid((unsigned)cdiv(M, kZB), (unsigned)cdiv(N, kZB), (unsigned)S);
    hipLaunchKernelGGL((zgemm_kernel<T, TA, TB, MODE, HERM>), grid, dim3(256), 0, s, M, N, K,
                       alpha, A, lda, B, ldb, beta, C, ldc, diag_off, kchunk,
                       S > 1 ? ws : (T *)nullptr);
    GLA_LAUNCH_CHECK();
    if (S > 1) {
        hipLaunchKernelGGL((zsplit_reduce_kernel<T>), dim3((unsigned)cdiv(M * N, 256)), dim3(256),
                           0, s, M, N, (int)S, (const T *)ws, beta, C, ldc);
        GLA_LAUNCH_CHECK();
    }
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
    constexpr int XLD = kZB + 1;               // padded row of X
    __shared__ T Lr[kZB * kZB], Li[kZB * kZB];   // L, dense row-major (for the inverse)
    __shared__ T Xr[kZB * XLD], Xi[kZB * XLD];
    __shared__ T Cr[2][kZB], Ci[2][kZB];       // column k, double-buffered
    const int t = threadIdx.x;
    const int i = t & 63, p = t >> 6;          // row, column class mod 4
    // thread (i, p) holds row i at the columns j = p + 4m in registers
    constexpr int NM = kZB / 4;
    T ar[NM], ai[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const int j = p + 4 * m;
        const bool ok = i < nb && j <= i;
        const int ic = i < nb ? i : nb - 1, jc = j <= ic ? j : ic;
        const T *a = A + 2 * (ic + (int64_t)jc * lda);
        const T vr = a[0], vi = a[1];
        ar[m] = ok ? vr : T(0);
        ai[m] = ok ? vi : T(0);
    }
    // per column k: the owners of column k publish it (one barrier, the
    // buffer alternates), every thread takes the pivot's square root and
    // the scaled column, and updates its own entries j in (k, i]
    for (int k = 0; k < nb; ++k) {
        const int buf = k & 1, mk = k >> 2;
        if (p == (k & 3)) {
            T vr = ar[0], vi = ai[0];
#pragma unroll
            for (int m = 1; m < NM; ++m) {
                vr = m == mk ? ar[m] : vr;
                vi = m == mk ? ai[m] : vi;
            }
            Cr[buf][i] = vr;
            Ci[buf][i] = vi;
        }
        __syncthreads();
        const T pr = Cr[buf][k], pi = Ci[buf][k];
        T lr, li, rs;
        if (pi == T(0) && pr > T(0)) {
            // a real positive pivot (every pivot of a Hermitian positive
            // definite matrix): the complex square root is the real one
            sqrt_and_rsqrt(pr, lr, rs);
            li = T(0);
        } else {
            csqrt_dev(pr, pi, lr, li);                               // :5
            rs = T(1) / lr;                                          // :8 inv(real(A[1,1]))
        }
        const T lir = Cr[buf][i] * rs, lii = Ci[buf][i] * rs;        // L[i, k]
        if (t == 0 && pr < T(0) && *info == 0) *info = (int)(info_base + k + 1);
        const bool row_live = i > k && i < nb;
#pragma unroll
        for (int m = 0; m < NM; ++m) {                               // :11 her!
            const int j = p + 4 * m;
            if (row_live && j > k && j <= i) {
                const T ljr = Cr[buf][j] * rs, lji = Ci[buf][j] * rs;
                // A[i,j] −= L[i,k] conj(L[j,k])
                ar[m] -= lir * ljr + lii * lji;
                ai[m] = j == i ? T(0) : ai[m] - (lii * ljr - lir * lji);
            }
        }
        if (p == (k & 3)) {                                          // column k itself
            const T nr = i == k ? lr : lir, ni = i == k ? li : lii;
            const bool set = i >= k && i < nb;
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                ar[m] = (set && m == mk) ? nr : ar[m];
                ai[m] = (set && m == mk) ? ni : ai[m];
            }
        }
    }
    // L to A and to the packed LDS triangle (for the inverse)
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const int j = p + 4 * m;
        if (i < nb && j <= i) {
            T *a = A + 2 * (i + (int64_t)j * lda);
            a[0] = ar[m];
            a[1] = ai[m];
            Lr[i * kZB + j] = ar[m];
            Li[i * kZB + j] = ai[m];
        }
    }
    // L⁻¹ row by row, every column at once: X[r][c] = (δ_rc − Σ_{c≤l<r}
    // L[r][l] X[l][c]) / L[r][r].  Lanes 4c..4c+3 take column c, lane q the
    // terms l ≡ q (mod 4), summed over the quad; a quad lives in one wave, so
    // each row waits only for the wave's own LDS writes of the rows above.
    __syncthreads();
    {
        const int c = t >> 2, q = t & 3;
        for (int r = 0; r < kZB; ++r) {
            // all 16 terms' LDS reads issued together (indices clamped into
            // the lower triangle, out-of-range terms zeroed)
            T sr = T(0), si = T(0);
#pragma unroll
            for (int u = 0; u < kZB / 4; ++u) {
                const int l = c + q + 4 * u;
                const bool ok = r < nb && c < nb && l < r;
                const int lc = ok ? l : c, rr = ok ? r : lc;
                const T lr_ = Lr[rr * kZB + lc], li_ = Li[rr * kZB + lc];
                const T yr = Xr[lc * XLD + c], yi = Xi[lc * XLD + c];
                sr += ok ? lr_ * yr - li_ * yi : T(0);
                si += ok ? lr_ * yi + li_ * yr : T(0);
            }
            sr = quad_sum(sr);
            si = quad_sum(si);
            if (q == 0) {
                T xr = T(0), xi = T(0);
                if (r >= c && r < nb && c < nb) {
                    const T br = (r == c ? T(1) : T(0)) - sr, bi = -si;
                    const T dr = Lr[r * kZB + r], di = Li[r * kZB + r];
                    const T den = dr * dr + di * di;
                    xr = (br * dr + bi * di) / den;
                    xi = (bi * dr - br * di) / den;
                }
                Xr[r * XLD + c] = xr;
                Xi[r * XLD + c] = xi;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
    }
    __syncthreads();
    for (int e = t; e < kZB * kZB; e += 256) {
        const int r = e % kZB, c = e / kZB;
        Linv[2 * e] = Xr[r * XLD + c];
        Linv[2 * e + 1] = Xi[r * XLD + c];
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

// Right-looking complex Cholesky with nb-wide steps (nb ≤ 64) and depth-1
// look-ahead: each step's herk! updates the next block column first on the
// caller's stream, and the rest of the trailing matrix on the handle's side
// stream, where it overlaps the next step's diagonal factor and panel solve
// (which touch only that block column); the next step's updates wait for it.
// Every element of A22 takes the same K-ordered update as in one product.
template <typename T>
int zchol_dev(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int nb, T *Linv, T *W,
              int64_t ldw, int *info)
{
    bool pending = false;   // a trailing update in flight on the side stream
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
        // A22 −= A21 A21ᴴ (lower, herk!): the previous step's side update
        // first, then the next block column here and the rest beside it
        if (pending) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
        pending = false;
        T *A22 = A21 + 2 * kb * lda;
        const int64_t kb2 = std::min<int64_t>(nb, below);
        GLA_TRY((zgemm_launch<T, false, true, 1, true>(s, below, kb2, kb, T(-1), A21, lda, A21,
                                                       lda, T(1), A22, lda, 0)));
        const int64_t rest = below - kb2;
        if (rest > 0) {
            GLA_HIP_TRY(hipEventRecord(h->ev_fork, s));
            GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_fork, 0));
            GLA_TRY((zgemm_launch<T, false, true, 1, true>(
                h->side, rest, rest, kb, T(-1), A21 + 2 * kb2, lda, A21 + 2 * kb2, lda, T(1),
                A22 + 2 * (kb2 + kb2 * lda), lda, 0)));
            GLA_HIP_TRY(hipEventRecord(h->ev_join, h->side));
            pending = true;
        }
    }
    if (pending) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
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
        GLA_TRY(zchol_dev<T>(h, s, n, A, lda, nb, Linv, W, ldw, h->dinfo));
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


// ------------------------------------------------------------- complex QR --
// qrUnblocked! (src/qr.jl:86-111) with the stdlib's complex reflector! and
// reflectorApply! (β = −copysign(‖x‖, real(x₁)), τ complex, Hᴴ = I − τ̄ v vᴴ
// applied from the left), on one block column: per column, every
// workgroup forms the reflector's scalars from the norm partials, writes
// its rows of v and the partial dots Σ conj(v_i) A[i, q] of the columns
// right of it (zqr_reflect_kernel); then each workgroup sums those dots in
// a fixed order and applies A[i, q] −= v_i τ̄ d_q to its rows, adding the
// next column's norm partials (zqr_apply_kernel).
constexpr int kZQ = 32;     // widest block column of one panel pass

template <typename T>
__device__ __forceinline__ T zwave_sum(T v)
{
    return wave_sum_dpp(v);
}

template <typename T>
__device__ T zblock_sum(const T *part, int cnt, T *sh)
{
    T s = T(0);
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) s += part[i];
    s = zwave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    T r = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return r;
}

// ‖A[j:m, j]‖² partials of the block column's first column.
template <typename T>
__global__ void zqr_norm_kernel(int64_t m, int j, const T *A, int64_t lda, T *part)
{
    __shared__ T sh[4];
    const int64_t i = j + (int64_t)blockIdx.x * 256 + threadIdx.x;
    T s = T(0);
    if (i < m) {
        const T *a = A + 2 * (i + (int64_t)j * lda);
        s = a[0] * a[0] + a[1] * a[1];
    }
    s = zwave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Column j of the m×w block column: the reflector (v into A below the
// diagonal, β on it, τ into tau[j]) and the partials dpart[blk·kZQ + q] =
// Σ_i conj(v_i) A[i, q] over this workgroup's rows, q = j+1..w−1.
template <typename T>
__global__ __launch_bounds__(256) void zqr_reflect_kernel(int64_t m, int w, int j, T *A,
                                                          int64_t lda, const T *part, int pcnt,
                                                          T *tau, T *dpart, T *scal)
{
    __shared__ T sh[4];
    __shared__ T red[4][kZQ][2];
    const T nrm2 = zblock_sum(part, pcnt, sh);
    const T *a1 = A + 2 * (j + (int64_t)j * lda);
    const T x1r = a1[0], x1i = a1[1];
    const T normu = sqrt(nrm2);
    T xir = T(1), xii = T(0), taur = T(0), taui = T(0), beta = x1r;
    bool zero = normu == T(0);
    if (!zero) {
        const T nu = copysign(normu, x1r);
        xir = x1r + nu;
        xii = x1i;
        taur = xir / nu;
        taui = xii / nu;
        beta = -nu;
    }
    const int64_t i = j + (int64_t)blockIdx.x * 256 + threadIdx.x;
    // v_i = x_i / ξ₁ (complex division), v_j = 1
    T vr = T(0), vi = T(0);
    if (i < m) {
        if (i == j) {
            vr = T(1);
        } else {
            const T *a = A + 2 * (i + (int64_t)j * lda);
            if (zero) {
                vr = a[0];
                vi = a[1];
            } else {
                const T den = xir * xir + xii * xii;
                vr = (a[0] * xir + a[1] * xii) / den;
                vi = (a[1] * xir - a[0] * xii) / den;
            }
        }
    }
    // dots with the remaining columns (before writing v: column j is not
    // among them), eight columns' loads issued at a time
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t ic = i < m ? i : m - 1;
    for (int q0 = j + 1; q0 < w; q0 += 8) {
        T br[8], bi[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u < w ? q0 + u : q0;
            const T *b = A + 2 * (ic + (int64_t)q * lda);
            br[u] = b[0];
            bi[u] = b[1];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u;
            if (q < w) {
                const T dr = zwave_sum(vr * br[u] + vi * bi[u]);   // conj(v) · b
                const T di = zwave_sum(vr * bi[u] - vi * br[u]);
                if (lane == 0) {
                    red[wv][q][0] = dr;
                    red[wv][q][1] = di;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < w && (int)threadIdx.x > j) {
        const int q = threadIdx.x;
        dpart[2 * ((int64_t)blockIdx.x * kZQ + q)] = red[0][q][0] + red[1][q][0] + red[2][q][0] + red[3][q][0];
        dpart[2 * ((int64_t)blockIdx.x * kZQ + q) + 1] =
            red[0][q][1] + red[1][q][1] + red[2][q][1] + red[3][q][1];
    }
    // v below the diagonal (each workgroup its own rows); A[j, j] is read
    // by every workgroup as x₁, so β goes there in the apply kernel
    if (i < m && i > j && !zero) {
        T *a = A + 2 * (i + (int64_t)j * lda);
        a[0] = vr;
        a[1] = vi;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        tau[2 * j] = taur;
        tau[2 * j + 1] = taui;
        scal[0] = taur;
        scal[1] = taui;
        scal[2] = beta;
        scal[3] = zero ? x1i : T(0);
    }
}

// A[i, q] −= v_i · conj(τ) d_q for rows i ≥ j, q = j+1..w−1 (d_q summed over
// the reflect kernel's partials in a fixed order); β into A[j, j]; partial
// ‖A[j+1:m, j+1]‖².
template <typename T>
__global__ __launch_bounds__(256) void zqr_apply_kernel(int64_t m, int w, int j, T *A, int64_t lda,
                                                        const T *dpart, int dcnt, const T *scal,
                                                        T *part)
{
    __shared__ T sh[4];
    __shared__ T dq[kZQ][2];
    const T taur = scal[0], taui = scal[1];
    if (threadIdx.x < w && (int)threadIdx.x > j) {
        const int q = threadIdx.x;
        T dr = T(0), di = T(0);
        for (int b = 0; b < dcnt; ++b) {
            dr += dpart[2 * ((int64_t)b * kZQ + q)];
            di += dpart[2 * ((int64_t)b * kZQ + q) + 1];
        }
        // conj(τ) d
        dq[q][0] = taur * dr + taui * di;
        dq[q][1] = taur * di - taui * dr;
    }
    __syncthreads();
    const int64_t i = j + (int64_t)blockIdx.x * 256 + threadIdx.x;
    T s = T(0);
    if (i < m) {
        T vr = T(1), vi = T(0);
        if (i > j) {
            const T *a = A + 2 * (i + (int64_t)j * lda);
            vr = a[0];
            vi = a[1];
        }
        // eight columns at a time: all loads, then the updates and stores
        for (int q0 = j + 1; q0 < w; q0 += 8) {
            T br[8], bi[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = q0 + u < w ? q0 + u : q0;
                const T *b = A + 2 * (i + (int64_t)q * lda);
                br[u] = b[0];
                bi[u] = b[1];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = q0 + u;
                if (q < w) {
                    const T cr = dq[q][0], ci = dq[q][1];
                    const T nr = br[u] - (vr * cr - vi * ci);
                    const T ni = bi[u] - (vr * ci + vi * cr);
                    T *b = A + 2 * (i + (int64_t)q * lda);
                    b[0] = nr;
                    b[1] = ni;
                    if (q == j + 1 && i > j) s = nr * nr + ni * ni;
                }
            }
        }
        if (i == j) {
            T *a = A + 2 * (i + (int64_t)j * lda);
            a[0] = scal[2];
            a[1] = scal[3];
        }
    }
    s = zwave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// qrUnblocked! of the m×w block column at A (w ≤ kZQ).
template <typename T>
int zqr_panel(hipStream_t s, int64_t m, int w, T *A, int64_t lda, T *tau, T *ws)
{
    const int64_t g = cdiv(m, 256);
    T *part = ws, *dpart = part + g + 8, *scal = dpart + 2 * g * kZQ;   // scal: τ, β
    hipLaunchKernelGGL((zqr_norm_kernel<T>), dim3((unsigned)g), dim3(256), 0, s, m, 0,
                       (const T *)A, lda, part);
    GLA_LAUNCH_CHECK();
    int pcnt = (int)g;
    for (int j = 0; j < w && j < m; ++j) {
        const unsigned gj = (unsigned)cdiv(m - j, 256);
        hipLaunchKernelGGL((zqr_reflect_kernel<T>), dim3(gj), dim3(256), 0, s, m, w, j, A, lda,
                           (const T *)part, pcnt, tau, dpart, scal);
        GLA_LAUNCH_CHECK();
        hipLaunchKernelGGL((zqr_apply_kernel<T>), dim3(gj), dim3(256), 0, s, m, w, j, A, lda,
                           (const T *)dpart, (int)gj, (const T *)scal, part);
        GLA_LAUNCH_CHECK();
        pcnt = (int)gj;
    }
    return GLA_OK;
}

// V (m×k, ld ldv): the unit lower reflectors of F explicitly (ones on the
// diagonal, zeros above).
template <typename T>
__global__ void zmake_v_kernel(int64_t m, int64_t k, const T *F, int64_t ldf, T *V, int64_t ldv)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (i >= m || c >= k) return;
    T *v = V + 2 * (i + c * ldv);
    if (i < c) {
        v[0] = T(0);
        v[1] = T(0);
    } else if (i == c) {
        v[0] = T(1);
        v[1] = T(0);
    } else {
        v[0] = F[2 * (i + c * ldf)];
        v[1] = F[2 * (i + c * ldf) + 1];
    }
}

// T (k×k, k ≤ 64) from G = VᴴV (k×k, ld ldg) and τ: S = I + diag(τ)
// striu(G), T = S⁻¹ diag(τ) (the UT transform of src/qr.jl:64-83 with the
// first term of each entry conjugated, so that I − V T Vᴴ = H₁⋯H_k).
// One workgroup: the strict upper part of S is formed once into LDS
// (coalesced reads of G), then thread c computes column c of S⁻¹ by back
// substitution from LDS alone.
template <typename T>
__global__ void zqr_tsmall_kernel(int k, const T *G, int64_t ldg, const T *tau, T *Tm, int64_t ldt)
{
    __shared__ T Xr[kZB * kZB], Xi[kZB * kZB];
    __shared__ T Sr[kZB * kZB], Si[kZB * kZB];   // S[r, l] = τ_r G[r, l], r < l
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        const int r = e % k, l = e / k;
        T vr = T(0), vi = T(0);
        if (r < l) {
            const T gr = G[2 * (r + (int64_t)l * ldg)], gi = G[2 * (r + (int64_t)l * ldg) + 1];
            const T tr = tau[2 * r], ti = tau[2 * r + 1];
            vr = tr * gr - ti * gi;
            vi = tr * gi + ti * gr;
        }
        Sr[r * kZB + l] = vr;
        Si[r * kZB + l] = vi;
    }
    __syncthreads();
    const int c = threadIdx.x;
    if (c < k) {
        // column c of S⁻¹ (unit upper): x_c = 1, x_r = −Σ_{l=r+1..c} S[r,l] x_l for r < c
        for (int r = k - 1; r >= 0; --r) {
            T xr = T(0), xi = T(0);
            if (r == c) xr = T(1);
            else if (r < c) {
                T sr = T(0), si = T(0);
                for (int l = r + 1; l <= c; ++l) {
                    const T s_r = Sr[r * kZB + l], s_i = Si[r * kZB + l];
                    const T yr = Xr[l * kZB + c], yi = Xi[l * kZB + c];
                    sr += s_r * yr - s_i * yi;
                    si += s_r * yi + s_i * yr;
                }
                xr = -sr;
                xi = -si;
            }
            Xr[r * kZB + c] = xr;
            Xi[r * kZB + c] = xi;
        }
    }
    __syncthreads();
    // T[r, c] = (S⁻¹)[r, c] τ_c for r ≤ c, 0 below
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        const int r = e % k, cc = e / k;
        T vr = T(0), vi = T(0);
        if (r <= cc) {
            const T yr = Xr[r * kZB + cc], yi = Xi[r * kZB + cc];
            const T tr = tau[2 * cc], ti = tau[2 * cc + 1];
            vr = yr * tr - yi * ti;
            vi = yr * ti + yi * tr;
        }
        Tm[2 * (r + (int64_t)cc * ldt)] = vr;
        Tm[2 * (r + (int64_t)cc * ldt) + 1] = vi;
    }
}

// S[0:rows, 0:gridDim.y] = 0 (complex, ld lds).
template <typename T>
__global__ void zzero_kernel(int64_t rows, T *S, int64_t lds)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i < rows) {
        S[2 * (i + j * lds)] = T(0);
        S[2 * (i + j * lds) + 1] = T(0);
    }
}

// T (k×k) of the first k reflectors of the explicit V (m×k), recursively:
// T = [T11 T12; 0 T22], T12 = −T11 (V1ᴴ V2) T22 (the blocked recurrence of
// the same UT transform).  scratch: 2·64² + k·k complex elements.
template <typename T>
int zqr_tblock_dev(hipStream_t s, int64_t m, int64_t k, const T *V, int64_t ldv, const T *tau,
                   T *Tm, int64_t ldt, T *scratch, T *part = nullptr, int64_t part_elems = 0)
{
    if (k <= kZB) {
        T *G = scratch;
        GLA_TRY((zgemm_launch<T, true, false, 0, false>(s, k, k, m, T(1), V, ldv, V, ldv, T(0), G,
                                                         k, 0, part, part_elems)));
        hipLaunchKernelGGL((zqr_tsmall_kernel<T>), dim3(1), dim3(kZB), 0, s, (int)k,
                           (const T *)G, k, tau, Tm, ldt);
        GLA_LAUNCH_CHECK();
        return GLA_OK;
    }
    const int64_t k1 = (k / 2 + kZB - 1) / kZB * kZB;
    GLA_TRY(zqr_tblock_dev<T>(s, m, k1, V, ldv, tau, Tm, ldt, scratch, part, part_elems));
    GLA_TRY(zqr_tblock_dev<T>(s, m - k1, k - k1, V + 2 * (k1 + k1 * ldv), ldv, tau + 2 * k1,
                              Tm + 2 * (k1 + k1 * ldt), ldt, scratch, part, part_elems));
    // G = V1ᴴ V2 over rows k1..m (V2 is zero above), P = T11 G, T12 = −P T22
    const int64_t k2 = k - k1;
    T *Gm = scratch, *P = Gm + 2 * k1 * k2;
    GLA_TRY((zgemm_launch<T, true, false, 0, false>(s, k1, k2, m - k1, T(1), V + 2 * k1, ldv,
                                                     V + 2 * (k1 + k1 * ldv), ldv, T(0), Gm, k1,
                                                     0, part, part_elems)));
    GLA_TRY((zgemm_launch<T, false, false, 0, false>(s, k1, k2, k1, T(1), Tm, ldt, Gm, k1, T(0), P,
                                                      k1, 0)));
    GLA_TRY((zgemm_launch<T, false, false, 0, false>(s, k1, k2, k2, T(-1), P, k1,
                                                      Tm + 2 * (k1 + k1 * ldt), ldt, T(0),
                                                      Tm + 2 * k1 * ldt, ldt, 0)));
    // T is upper triangular: the block below T11 is zero
    hipLaunchKernelGGL((zzero_kernel<T>), dim3((unsigned)cdiv(k2, 256), (unsigned)k1), dim3(256),
                       0, s, k2, Tm + 2 * k1, ldt);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// A (mA×nA) <- op(H) A, H = I − V T Vᴴ (V mA×b explicit, T b×b):
// W1 = Vᴴ A, W2 = op(T) W1 (op = ᴴ for trans 'C', none for 'N'), A −= V W2.
template <typename T>
int zwy_apply_dev(hipStream_t s, char trans, int64_t mA, int64_t nA, int64_t b, const T *V,
                  int64_t ldv, const T *Tm, int64_t ldt, T *A, int64_t lda, T *W1, T *W2,
                  T *part = nullptr, int64_t part_elems = 0)
{
    if (mA <= 0 || nA <= 0 || b <= 0) return GLA_OK;
    GLA_TRY((zgemm_launch<T, true, false, 0, false>(s, b, nA, mA, T(1), V, ldv, A, lda, T(0), W1,
                                                     b, 0, part, part_elems)));
    if (trans == 'C')
        GLA_TRY((zgemm_launch<T, true, false, 0, false>(s, b, nA, b, T(1), Tm, ldt, W1, b, T(0),
                                                         W2, b, 0)));
    else
        GLA_TRY((zgemm_launch<T, false, false, 0, false>(s, b, nA, b, T(1), Tm, ldt, W1, b, T(0),
                                                          W2, b, 0)));
    return zgemm_launch<T, false, false, 0, false>(s, mA, nA, b, T(-1), V, ldv, W2, b, T(1), A,
                                                   lda, 0);
}

// Real elements of the split-K partials of the complex QR: 16 partial
// products of a 32-row panel by n columns (or of a 32×32 Gram matrix).
inline int64_t zqr_part_elems(int64_t n) { return 2 * 16 * (int64_t)kZQ * std::max<int64_t>(n, kZQ); }

// qrBlocked! (src/qr.jl:113-146) for complex A: block columns w =
// min(blocksize, 32) wide, each factored by zqr_panel, its T formed and
// Hᴴ applied to the columns right of it.
template <typename T>
int zqr_blocked_dev(hipStream_t s, int64_t m, int64_t n, T *A, int64_t lda, int64_t bs, T *tau,
                    T *ws)
{
    const int64_t k = std::min(m, n);
    const int w0 = (int)std::min<int64_t>(std::max<int64_t>(bs, 1), kZQ);
    const int64_t g = cdiv(m, 256);
    T *pan = ws;                                       // panel scratch
    T *V = pan + 2 * (g + 8 + 2 * g * kZQ + 64);       // m × w0
    T *Tm = V + 2 * m * w0;                            // w0 × w0
    T *sc = Tm + 2 * kZQ * kZQ;                        // tblock scratch
    T *W1 = sc + 2 * (2 * kZB * kZB + kZQ * kZQ), *W2 = W1 + 2 * (int64_t)w0 * n;
    T *part = W2 + 2 * (int64_t)w0 * n;                // split-K partials
    const int64_t part_elems = zqr_part_elems(n);
    for (int64_t off = 0; off < k; off += w0) {
        const int w = (int)std::min<int64_t>(w0, n - off);
        const int64_t mm = m - off;
        T *Ab = A + 2 * (off + off * lda);
        GLA_TRY(zqr_panel<T>(s, mm, w, Ab, lda, tau + 2 * off, pan));
        const int64_t rest = n - off - w;
        if (rest > 0) {
            const int kb = (int)std::min<int64_t>(mm, w);
            hipLaunchKernelGGL((zmake_v_kernel<T>), dim3((unsigned)cdiv(mm, 256), (unsigned)kb),
                               dim3(256), 0, s, mm, (int64_t)kb, (const T *)Ab, lda, V, mm);
            GLA_LAUNCH_CHECK();
            GLA_TRY(zqr_tblock_dev<T>(s, mm, kb, V, mm, tau + 2 * off, Tm, kZQ, sc, part,
                                      part_elems));
            GLA_TRY(zwy_apply_dev<T>(s, 'C', mm, rest, kb, V, mm, Tm, kZQ, Ab + 2 * w * lda, lda,
                                     W1, W2, part, part_elems));
        }
    }
    return GLA_OK;
}

size_t zqr_scratch(int64_t m, int64_t n)
{
    const int64_t g = cdiv(m, 256);
    return 2 * ((size_t)(g + 8 + 2 * g * kZQ + 64) + (size_t)m * kZQ + kZQ * kZQ +
                (2 * kZB * kZB + kZQ * kZQ) + 2 * (size_t)kZQ * n) +
           (size_t)zqr_part_elems(n) + 256;
}

template <typename T>
int zqr_entry(gla_handle_t hh, int64_t m, int64_t n, void *Av, int64_t lda, int64_t bs,
              void *tauv, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (m < 0) return -2;
    if (n < 0) return -3;
    T *A = reinterpret_cast<T *>(Av);
    T *tau = reinterpret_cast<T *>(tauv);
    if (!A && m > 0 && n > 0) { set_error("A is null"); return -4; }
    if (lda < std::max<int64_t>(1, m)) { set_error("lda < max(1,m)"); return -5; }
    if (bs < 1) { set_error("blocksize < 1"); return -6; }
    if (!tau && std::min(m, n) > 0) { set_error("tau is null"); return -7; }
    if (std::min(m, n) == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_TRY(ensure_workspace(h, zqr_scratch(m, n) * sizeof(T)));
    return zqr_blocked_dev<T>(reinterpret_cast<hipStream_t>(stream), m, n, A, lda, bs, tau,
                              reinterpret_cast<T *>(h->ws));
}

template <typename T>
int ztblock_entry(gla_handle_t hh, int64_t m, int64_t n, const void *Fv, int64_t ldf,
                  const void *tauv, void *Tv, int64_t ldt, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    const int64_t k = std::min(m, n);
    if (m < 0 || n < 0) return -2;
    if (ldf < std::max<int64_t>(1, m)) return -5;
    if (ldt < std::max<int64_t>(1, k)) return -8;
    if (k == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    const size_t need = 2 * ((size_t)m * k + 2 * kZB * kZB + (size_t)k * k) + 256;
    GLA_TRY(ensure_workspace(h, need * sizeof(T)));
    T *V = reinterpret_cast<T *>(h->ws), *sc = V + 2 * m * k;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL((zmake_v_kernel<T>), dim3((unsigned)cdiv(m, 256), (unsigned)k), dim3(256),
                       0, s, m, k, reinterpret_cast<const T *>(Fv), ldf, V, m);
    GLA_LAUNCH_CHECK();
    return zqr_tblock_dev<T>(s, m, k, V, m, reinterpret_cast<const T *>(tauv),
                             reinterpret_cast<T *>(Tv), ldt, sc);
}

template <typename T>
int zwy_entry(gla_handle_t hh, char trans, int64_t mA, int64_t nA, const void *Vv, int64_t ldv,
              int64_t nV, const void *Tv, int64_t ldt, void *Av, int64_t lda, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (trans != 'N' && trans != 'C') { set_error("trans must be 'N' or 'C'"); return -2; }
    if (mA < 0) return -3;
    if (nA < 0) return -4;
    if (ldv < std::max<int64_t>(1, mA)) return -6;
    const int64_t b = std::min(mA, nV);
    if (ldt < std::max<int64_t>(1, b)) return -9;
    if (lda < std::max<int64_t>(1, mA)) return -11;
    if (b == 0 || nA == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    const size_t need = 2 * ((size_t)mA * b + 2 * (size_t)b * nA) + 256;
    GLA_TRY(ensure_workspace(h, need * sizeof(T)));
    T *V = reinterpret_cast<T *>(h->ws), *W1 = V + 2 * mA * b, *W2 = W1 + 2 * b * nA;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL((zmake_v_kernel<T>), dim3((unsigned)cdiv(mA, 256), (unsigned)b), dim3(256),
                       0, s, mA, b, reinterpret_cast<const T *>(Vv), ldv, V, mA);
    GLA_LAUNCH_CHECK();
    return zwy_apply_dev<T>(s, trans, mA, nA, b, V, mA, reinterpret_cast<const T *>(Tv), ldt,
                            reinterpret_cast<T *>(Av), lda, W1, W2);
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

int gla_qr_unblocked_z(gla_handle_t h, int64_t m, int64_t n, gla_cdouble *A, int64_t lda,
                       gla_cdouble *tau, void *stream)
{
    return gla::zqr_entry<double>(h, m, n, A, lda, 32, tau, stream);
}
int gla_qr_unblocked_c(gla_handle_t h, int64_t m, int64_t n, gla_cfloat *A, int64_t lda,
                       gla_cfloat *tau, void *stream)
{
    return gla::zqr_entry<float>(h, m, n, A, lda, 32, tau, stream);
}
int gla_qr_blocked_z(gla_handle_t h, int64_t m, int64_t n, gla_cdouble *A, int64_t lda,
                     int64_t blocksize, gla_cdouble *tau, void *stream)
{
    return gla::zqr_entry<double>(h, m, n, A, lda, blocksize, tau, stream);
}
int gla_qr_blocked_c(gla_handle_t h, int64_t m, int64_t n, gla_cfloat *A, int64_t lda,
                     int64_t blocksize, gla_cfloat *tau, void *stream)
{
    return gla::zqr_entry<float>(h, m, n, A, lda, blocksize, tau, stream);
}
int gla_qr_tblock_z(gla_handle_t h, int64_t m, int64_t n, const gla_cdouble *F, int64_t ldf,
                    const gla_cdouble *tau, gla_cdouble *T, int64_t ldt, void *stream)
{
    return gla::ztblock_entry<double>(h, m, n, F, ldf, tau, T, ldt, stream);
}
int gla_qr_tblock_c(gla_handle_t h, int64_t m, int64_t n, const gla_cfloat *F, int64_t ldf,
                    const gla_cfloat *tau, gla_cfloat *T, int64_t ldt, void *stream)
{
    return gla::ztblock_entry<float>(h, m, n, F, ldf, tau, T, ldt, stream);
}
int gla_wy_apply_z(gla_handle_t h, char trans, int64_t mA, int64_t nA, const gla_cdouble *V,
                   int64_t ldv, int64_t nV, const gla_cdouble *T, int64_t ldt, gla_cdouble *A,
                   int64_t lda, void *stream)
{
    return gla::zwy_entry<double>(h, trans, mA, nA, V, ldv, nV, T, ldt, A, lda, stream);
}
int gla_wy_apply_c(gla_handle_t h, char trans, int64_t mA, int64_t nA, const gla_cfloat *V,
                   int64_t ldv, int64_t nV, const gla_cfloat *T, int64_t ldt, gla_cfloat *A,
                   int64_t lda, void *stream)
{
    return gla::zwy_entry<float>(h, trans, mA, nA, V, ldv, nV, T, ldt, A, lda, stream);
}

}  // extern "C"
