// Batched Cholesky factor + solve for many SPD matrices of uneven size
// (n ≤ 256), FP64 and FP32, one right-hand side each, in one persistent
// launch.  No counterpart in the reference (which factors one matrix per
// call, src/cholesky.jl); each matrix gets cholUnblocked!'s factor (to
// rounding), then the solve L Lᵀ x = b.
//
// Plan (host): matrices are sorted by n, largest first.  Those above a tuned
// threshold `wave_max` become one work item each, handled by a whole
// workgroup; the rest are packed four to an item, one wave per matrix.
// Workgroups take items round-robin (item i → workgroup i % grid).
//
// Wave path (any n ≤ 256, the default): see wave_chol_solve — no workgroup
// barrier, the panel in registers, left-looking updates streamed from L.
// Workgroup path: right-looking, 32-column panels factored in LDS with one
// barrier per column, trailing lower triangle updated in 4×4 register blocks
// in memory, solves with the 32×32 diagonal blocks staged in LDS.
#include <algorithm>
#include <vector>

#include "gla_internal.hpp"

namespace gla {

struct BatchPlan {
    int device = 0;
    int64_t batch = 0;
    int nitems = 0;
    int grid = 0;
    int wave_max = 32;
    int *d_n = nullptr;        // n per matrix (device)
    int4 *d_items = nullptr;   // work items (device)
};

namespace {

constexpr int kBThreads = 256;
constexpr int kBMax = 256;     // largest n
constexpr int kPW = 32;        // panel width
constexpr int kPLD = kPW + 1;  // padded panel row
constexpr int kWMax = kBMax;   // the wave path takes any n ≤ 256
constexpr int kWP = 16;        // panel width of the wave path

template <typename T>
__device__ __forceinline__ T bsqrt(T x);
template <>
__device__ __forceinline__ double bsqrt<double>(double x) { return __builtin_sqrt(x); }
template <>
__device__ __forceinline__ float bsqrt<float>(float x) { return __builtin_sqrtf(x); }

__device__ __forceinline__ void border() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// ---------------------------------------------------------------- wave --
// One wave factors and solves one matrix (n ≤ 256) with no workgroup
// barrier.  Lane l holds rows l + 64s (s < 4).  Left-looking by 16-column
// panels: the panel sits in registers; the contributions of the columns
// already factored are applied as a GEMM (L rows streamed from memory, the
// panel's own 16 rows staged in wave-private LDS), then its 16 columns are
// factored in registers (pivot through a scalar register, the scaled column
// shared through LDS), then written back.  The solves keep b in registers:
// forward by column sweeps, backward by per-column wave sums.

template <typename T>
__device__ __forceinline__ T rlane(T v, int l)
{
    if constexpr (sizeof(T) == 8) {
        int2 u = __builtin_bit_cast(int2, v);
        u.x = __builtin_amdgcn_readlane(u.x, l);
        u.y = __builtin_amdgcn_readlane(u.y, l);
        return __builtin_bit_cast(T, u);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    }
}

template <int CTRL, typename T>
__device__ __forceinline__ T dppmov(T v)
{
    if constexpr (sizeof(T) == 8) {
        int2 u = __builtin_bit_cast(int2, v);
        u.x = __builtin_amdgcn_update_dpp(0, u.x, CTRL, 0xF, 0xF, false);
        u.y = __builtin_amdgcn_update_dpp(0, u.y, CTRL, 0xF, 0xF, false);
        return __builtin_bit_cast(T, u);
    } else {
        return __builtin_bit_cast(
            T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
    }
}

// Sum over the 64 lanes, the same value in every lane and a fixed order:
// quads (xor 1, xor 2), half rows, rows by DPP, then the four rows.
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    v += dppmov<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dppmov<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dppmov<0x141>(v);   // row_half_mirror
    v += dppmov<0x140>(v);   // row_mirror
    return (rlane(v, 0) + rlane(v, 16)) + (rlane(v, 32) + rlane(v, 48));
}

template <typename T>
__device__ __forceinline__ T sel4(const T v[4], int s)
{
    T r = v[0];
    r = s == 1 ? v[1] : r;
    r = s == 2 ? v[2] : r;
    r = s == 3 ? v[3] : r;
    return r;
}

template <typename T>
__device__ void wave_chol_solve(int n, T *__restrict__ A, int lda, T *__restrict__ b, int *info,
                                T *Bt)
{
    const int lane = threadIdx.x & 63;
    T *colv = Bt + kWP * kWP;
    int bad = 0;
    for (int p0 = 0; p0 < n; p0 += kWP) {
        const int pw = min(kWP, n - p0);
        T pan[4][kWP];
        // (a) the panel's lower part, rows ≥ p0
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int r = lane + 64 * s;
#pragma unroll
            for (int j = 0; j < kWP; ++j)
                pan[s][j] = (r < n && j < pw && r >= p0 + j) ? A[r + (int64_t)(p0 + j) * lda] : T(0);
        }
        // (b) pan -= L[rows, 0:p0] · L[p0:p0+32, 0:p0]ᵀ, 32 columns of L at a time
        for (int kp = 0; kp < p0; kp += kWP) {
#pragma unroll
            for (int u = 0; u < kWP * kWP / 64; ++u) {
                const int e = lane + 64 * u, j = e & (kWP - 1), kk = e / kWP;
                Bt[kk * kWP + j] = j < pw ? A[(p0 + j) + (int64_t)(kp + kk) * lda] : T(0);
            }
            border();
#pragma unroll 2
            for (int kk = 0; kk < kWP; ++kk) {
                T a[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int r = lane + 64 * s;
                    a[s] = (r < n && r >= p0) ? A[r + (int64_t)(kp + kk) * lda] : T(0);
                }
#pragma unroll
                for (int j = 0; j < kWP; ++j) {
                    const T bj = Bt[kk * kWP + j];
#pragma unroll
                    for (int s = 0; s < 4; ++s) pan[s][j] -= a[s] * bj;
                }
            }
            border();
        }
        // (c) factor the panel's 32 columns (cholUnblocked! on them)
        const int so = p0 >> 6, lo = p0 & 63;     // rows p0..p0+kWP-1: lanes lo.., slot so
#pragma unroll
        for (int c = 0; c < kWP; ++c) {
            if (c < pw) {
            T pv[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) pv[s] = pan[s][c];
            const T piv = rlane(sel4(pv, so), lo + c);
            if (piv < T(0) && !bad) bad = p0 + c + 1;
            const T d = bsqrt(piv);
            const T rr = T(1) / d;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int r = lane + 64 * s;
                pan[s][c] = r > p0 + c ? pan[s][c] * rr : (r == p0 + c ? d : pan[s][c]);
            }
            if (lane >= lo && lane < lo + kWP) {
                T q[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) q[s] = pan[s][c];
                colv[lane - lo] = sel4(q, so);
            }
            border();
#pragma unroll
            for (int j = c + 1; j < kWP; ++j) {
                const T lj = colv[j];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int r = lane + 64 * s;
                    if (j < pw && r >= p0 + j) pan[s][j] -= pan[s][c] * lj;
                }
            }
            border();
            }
        }
        // (d) write the panel back
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int r = lane + 64 * s;
#pragma unroll
            for (int j = 0; j < kWP; ++j)
                if (r < n && j < pw && r >= p0 + j) A[r + (int64_t)(p0 + j) * lda] = pan[s][j];
        }
    }
    // L y = b, then Lᵀ x = y, with b and 1/L_rr in registers
    T bv[4], rdg[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int r = lane + 64 * s;
        bv[s] = r < n ? b[r] : T(0);
        rdg[s] = r < n ? T(1) / A[r + (int64_t)r * lda] : T(0);
    }
    for (int c = 0; c < n; ++c) {
        const int cs = c >> 6, cl = c & 63;
        const T yc = rlane(sel4(bv, cs), cl) * rlane(sel4(rdg, cs), cl);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int r = lane + 64 * s;
            if (s == cs && lane == cl) bv[s] = yc;
            if (r > c && r < n) bv[s] -= A[r + (int64_t)c * lda] * yc;
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        T part = T(0);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int r = lane + 64 * s;
            if (r > c && r < n) part += A[r + (int64_t)c * lda] * bv[s];
        }
        const T sum = wave_sum(part);
        const int cs = c >> 6, cl = c & 63;
        const T xc = (rlane(sel4(bv, cs), cl) - sum) * rlane(sel4(rdg, cs), cl);
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (s == cs && lane == cl) bv[s] = xc;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int r = lane + 64 * s;
        if (r < n) b[r] = bv[s];
    }
    if (lane == 0 && info) *info = bad;
}

// ----------------------------------------------------------- workgroup --
template <typename T>
__device__ void wg_chol_solve(int n, T *__restrict__ A, int lda, T *__restrict__ b, int *info,
                              T *P, T *bs, T *sc)
{
    const int t = threadIdx.x;
    int bad = 0;   // kept by thread 0
    for (int p0 = 0; p0 < n; p0 += kPW) {
        const int pw = min(kPW, n - p0), rows = n - p0;
        T *Ap = A + p0 + (int64_t)p0 * lda;
        // thread t owns panel rows t, t + 256 (n ≤ 256: one row)
        for (int r = t; r < rows; r += kBThreads)
            for (int c = 0; c < pw && c <= r; ++c) P[r * kPLD + c] = Ap[r + (int64_t)c * lda];
        __syncthreads();
        // panel factor, one barrier per column: the panel keeps the Schur
        // complement unscaled (P_ij -= P_ic P_jc / P_cc, every thread reading
        // the pivot itself), and the columns are scaled by 1/√P_cc at the end
        // — the same L as sqrt / reciprocal scale / rank-1 update
        for (int c = 0; c < pw; ++c) {
            const T piv = P[c * kPLD + c];
            if (t == 0 && piv < T(0) && !bad) bad = p0 + c + 1;
            const T rp = T(1) / piv;
            for (int i = c + 1 + t; i < rows; i += kBThreads) {
                const T lic = P[i * kPLD + c] * rp;
                const int je = min(pw - 1, i);
                for (int j = c + 1; j <= je; ++j) P[i * kPLD + j] -= lic * P[j * kPLD + c];
            }
            __syncthreads();
        }
        if (t < pw) {
            const T d = bsqrt(P[t * kPLD + t]);
            sc[t] = d;
            sc[kPW + t] = T(1) / d;
        }
        __syncthreads();
        for (int r = t; r < rows; r += kBThreads)
            for (int c = 0; c < pw && c <= r; ++c)
                P[r * kPLD + c] = c < r ? P[r * kPLD + c] * sc[kPW + c] : sc[c];
        __syncthreads();
        for (int r = t; r < rows; r += kBThreads)
            for (int c = 0; c < pw && c <= r; ++c) Ap[r + (int64_t)c * lda] = P[r * kPLD + c];
        // trailing A22 -= P2 P2ᵀ, lower triangle, 4×4 blocks per thread
        const int r2 = rows - pw;
        if (r2 > 0) {
            const int nb4 = (r2 + 3) / 4;
            const int nblk = nb4 * (nb4 + 1) / 2;
            T *A22 = Ap + pw + (int64_t)pw * lda;
            const T *P2 = P + pw * kPLD;
            for (int blk = t; blk < nblk; blk += kBThreads) {
                // blk -> (bi, bj) with bj <= bi in the packed lower triangle of blocks
                int bi = (int)((__builtin_sqrt(8.0 * blk + 1.0) - 1.0) * 0.5);
                while ((bi + 1) * (bi + 2) / 2 <= blk) ++bi;
                while (bi * (bi + 1) / 2 > blk) --bi;
                const int bj = blk - bi * (bi + 1) / 2;
                T acc[4][4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] = T(0);
                for (int c = 0; c < pw; ++c) {
                    T x[4], y[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int i = bi * 4 + u, j = bj * 4 + u;
                        x[u] = i < r2 ? P2[i * kPLD + c] : T(0);
                        y[u] = j < r2 ? P2[j * kPLD + c] : T(0);
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[u][v] += x[u] * y[v];
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = bj * 4 + v;
                    if (j >= r2) continue;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int i = bi * 4 + u;
                        if (i < r2 && j <= i) A22[i + (int64_t)j * lda] -= acc[u][v];
                    }
                }
            }
        }
        __syncthreads();
    }
    // The solves stage each 32×32 diagonal block of L and the reciprocals of
    // its diagonal in LDS, so the serial 32-step chains touch no global
    // memory; the rows outside the block are updated in parallel.
    T *D = P + kPW * kPLD;                 // rows 32..63 of the panel buffer
    auto stage = [&](int p0, int pw) {
        for (int e = t; e < kPW * kPW; e += kBThreads) {
            const int i = e & (kPW - 1), k = e / kPW;
            if (k <= i && i < pw) D[i * kPLD + k] = A[(p0 + i) + (int64_t)(p0 + k) * lda];
        }
        __syncthreads();
        if (t < pw) sc[kPW + t] = T(1) / D[t * kPLD + t];
        __syncthreads();
    };
    // L y = b, 32 rows at a time
    for (int e = t; e < n; e += kBThreads) bs[e] = b[e];
    __syncthreads();
    for (int p0 = 0; p0 < n; p0 += kPW) {
        const int pw = min(kPW, n - p0);
        stage(p0, pw);
        if (t < 64) {
            for (int k = 0; k < pw; ++k) {
                const T y = bs[p0 + k] * sc[kPW + k];
                border();
                if (t == k) bs[p0 + k] = y;
                if (t > k && t < pw) bs[p0 + t] -= D[t * kPLD + k] * y;
                border();
            }
        }
        __syncthreads();
        for (int i = p0 + pw + t; i < n; i += kBThreads) {
            T s = bs[i];
            for (int k = 0; k < pw; ++k) s -= A[i + (int64_t)(p0 + k) * lda] * bs[p0 + k];
            bs[i] = s;
        }
        __syncthreads();
    }
    // Lᵀ x = y, last block first
    const int nblocks = (n + kPW - 1) / kPW;
    for (int pb = nblocks - 1; pb >= 0; --pb) {
        const int p0 = pb * kPW, pw = min(kPW, n - p0);
        // y_p -= L[p0+pw:n, p]ᵀ x[p0+pw:n]: 8 partial sums per column
        {
            const int c = t % kPW, h = t / kPW;     // 32 columns × 8 groups
            T s = T(0);
            if (c < pw)
                for (int i = p0 + pw + h; i < n; i += kBThreads / kPW)
                    s += A[i + (int64_t)(p0 + c) * lda] * bs[i];
            P[h * kPLD + c] = s;
        }
        __syncthreads();
        if (t < pw) {
            T s = T(0);
#pragma unroll
            for (int h = 0; h < kBThreads / kPW; ++h) s += P[h * kPLD + t];
            bs[p0 + t] -= s;
        }
        stage(p0, pw);
        if (t < 64) {
            for (int k = pw - 1; k >= 0; --k) {
                const T x = bs[p0 + k] * sc[kPW + k];
                border();
                if (t == k) bs[p0 + k] = x;
                if (t < k) bs[p0 + t] -= D[k * kPLD + t] * x;
                border();
            }
        }
        __syncthreads();
    }
    for (int e = t; e < n; e += kBThreads) b[e] = bs[e];
    if (t == 0 && info) *info = bad;
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(kBThreads, 1) void batched_chol_solve_kernel(
    int nitems, const int4 *__restrict__ items, const int *__restrict__ nvec,
    T *const *__restrict__ Aptr, const int *__restrict__ ldav, T *const *__restrict__ bptr,
    int *__restrict__ info)
{
    __shared__ T P[kBMax * kPLD];       // panel (workgroup path) or 4 wave matrices
    __shared__ T bs[kBMax];
    __shared__ T sc[2 * kPW];
    static_assert(4 * (kWP * kWP + kWP) <= kBMax * kPLD, "wave scratch fits the panel buffer");
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int4 item = items[it];
        if (item.y == -2) {                 // one matrix, whole workgroup
            const int i = item.x;
            wg_chol_solve<T>(nvec[i], Aptr[i], ldav[i], bptr[i], info ? info + i : nullptr, P, bs,
                             sc);
        } else {                            // up to four small matrices, one per wave
            const int w = threadIdx.x >> 6;
            const int i = w == 0 ? item.x : (w == 1 ? item.y : (w == 2 ? item.z : item.w));
            if (i >= 0)
                wave_chol_solve<T>(nvec[i], Aptr[i], ldav[i], bptr[i], info ? info + i : nullptr,
                                   P + w * (kWP * kWP + kWP));
            __syncthreads();
        }
    }
}

}  // namespace

template <typename T>
int batched_launch(Handle *h, BatchPlan *pl, T *const *A, const int *lda, T *const *b, int *info,
                   hipStream_t s)
{
    if (pl->nitems == 0) return GLA_OK;
    (void)h;
    hipLaunchKernelGGL((batched_chol_solve_kernel<T>), dim3(pl->grid), dim3(kBThreads), 0, s,
                       pl->nitems, pl->d_items, pl->d_n, A, lda, b, info);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

}  // namespace gla

extern "C" {

int gla_batch_plan_create(gla_handle_t hh, int64_t batch, const int32_t *n_host, int32_t wave_max,
                          int32_t grid_mult, gla_batch_plan_t *out)
{
    using gla::BatchPlan;
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (batch < 0) return -2;
    if (batch > 0 && !n_host) return -3;
    if (wave_max < 0 || wave_max > gla::kWMax) { gla::set_error("wave_max must be in [0, 256]"); return -4; }
    if (grid_mult < 1 || grid_mult > 64) return -5;
    if (!out) return -6;
    for (int64_t i = 0; i < batch; ++i)
        if (n_host[i] < 0 || n_host[i] > gla::kBMax) {
            gla::set_error("matrix %lld: n = %d outside [0, 256]", (long long)i, n_host[i]);
            return GLA_ERR_UNSUPPORTED;
        }
    GLA_HIP_TRY(hipSetDevice(h->device));
    std::vector<int> order((size_t)batch);
    for (int64_t i = 0; i < batch; ++i) order[(size_t)i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n_host[a] > n_host[b]; });
    std::vector<int4> items;
    size_t k = 0;
    for (; k < order.size() && n_host[order[k]] > wave_max; ++k)
        items.push_back(make_int4(order[k], -2, -2, -2));
    for (; k < order.size(); k += 4) {
        int4 it = make_int4(-1, -1, -1, -1);
        it.x = order[k];
        if (k + 1 < order.size()) it.y = order[k + 1];
        if (k + 2 < order.size()) it.z = order[k + 2];
        if (k + 3 < order.size()) it.w = order[k + 3];
        items.push_back(it);
    }
    auto *pl = new BatchPlan();
    pl->device = h->device;
    pl->batch = batch;
    pl->nitems = (int)items.size();
    pl->wave_max = wave_max;
    pl->grid = std::max(1, std::min<int>(pl->nitems, h->num_cus * grid_mult));
    if (batch > 0) {
        GLA_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&pl->d_n), sizeof(int) * batch));
        GLA_HIP_TRY(hipMemcpy(pl->d_n, n_host, sizeof(int) * batch, hipMemcpyHostToDevice));
    }
    if (!items.empty()) {
        GLA_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&pl->d_items), sizeof(int4) * items.size()));
        GLA_HIP_TRY(hipMemcpy(pl->d_items, items.data(), sizeof(int4) * items.size(),
                              hipMemcpyHostToDevice));
    }
    *out = reinterpret_cast<gla_batch_plan_t>(pl);
    return GLA_OK;
}

int gla_batch_plan_destroy(gla_batch_plan_t p)
{
    auto *pl = reinterpret_cast<gla::BatchPlan *>(p);
    if (!pl) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(pl->device));
    GLA_HIP_TRY(hipDeviceSynchronize());
    if (pl->d_n) (void)hipFree(pl->d_n);
    if (pl->d_items) (void)hipFree(pl->d_items);
    delete pl;
    return GLA_OK;
}

int gla_chol_solve_batched_d(gla_handle_t hh, gla_batch_plan_t p, double *const *A,
                             const int32_t *lda, double *const *b, int32_t *info, void *stream)
{
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    auto *pl = reinterpret_cast<gla::BatchPlan *>(p);
    if (!h) return -1;
    if (!pl) return -2;
    if (pl->batch > 0 && (!A || !lda || !b)) return -3;
    return gla::batched_launch<double>(h, pl, A, lda, b, info, reinterpret_cast<hipStream_t>(stream));
}

int gla_chol_solve_batched_s(gla_handle_t hh, gla_batch_plan_t p, float *const *A,
                             const int32_t *lda, float *const *b, int32_t *info, void *stream)
{
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    auto *pl = reinterpret_cast<gla::BatchPlan *>(p);
    if (!h) return -1;
    if (!pl) return -2;
    if (pl->batch > 0 && (!A || !lda || !b)) return -3;
    return gla::batched_launch<float>(h, pl, A, lda, b, info, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
