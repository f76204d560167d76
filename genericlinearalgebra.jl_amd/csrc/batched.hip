// Batched Cholesky factor + solve for many SPD matrices of uneven size
// (n ≤ 256), FP64 and FP32, one right-hand side each, in one persistent
// launch.  No counterpart in the reference (which factors one matrix per
// call, src/cholesky.jl); each matrix gets exactly cholUnblocked!'s result,
// then the solve L Lᵀ x = b.
//
// Plan (host): matrices are sorted by n, largest first.  Those above a tuned
// threshold `wave_max` become one work item each, handled by a whole
// workgroup; the rest are packed four to an item, one wave per matrix.
// Workgroups take items round-robin (item i → workgroup i % grid), so the
// large matrices spread over the chip first and the small ones fill in.
//
// Workgroup path (n ≤ 256): right-looking, 32-column panels.  The panel
// (rows p0..n-1) is factored in LDS column by column (pivot sqrt, reciprocal
// scale, rank-1 update), written back, and the trailing lower triangle takes
// A22 -= P2 P2ᵀ with each thread owning 4×4 blocks (the matrix itself stays
// in L2/HBM).  The solves go 32 rows at a time: a wave solves the 32×32
// triangle in LDS, the workgroup updates the remaining right-hand side.
// Wave path (n ≤ 32): the matrix lives in LDS; one wave factors and solves
// it with wave-synchronous LDS traffic only.
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
constexpr int kWMax = 32;      // largest n of the wave path
constexpr int kWLD = kWMax + 1;

template <typename T>
__device__ __forceinline__ T bsqrt(T x);
template <>
__device__ __forceinline__ double bsqrt<double>(double x) { return __builtin_sqrt(x); }
template <>
__device__ __forceinline__ float bsqrt<float>(float x) { return __builtin_sqrtf(x); }

__device__ __forceinline__ void border() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// ---------------------------------------------------------------- wave --
// One wave: factor and solve an n×n (n ≤ 32) matrix held in S (LDS).
template <typename T>
__device__ void wave_chol_solve(int n, T *__restrict__ A, int lda, T *__restrict__ b, int *info,
                                T *S, T *bs)
{
    const int lane = threadIdx.x & 63;
    for (int e = lane; e < n * n; e += 64) {
        const int i = e % n, j = e / n;
        if (j <= i) S[i * kWLD + j] = A[i + (int64_t)j * lda];
    }
    if (lane < n) bs[lane] = b[lane];
    border();
    int bad = 0;
    for (int k = 0; k < n; ++k) {
        T d = S[k * kWLD + k];
        if (d < T(0) && !bad) bad = k + 1;
        d = bsqrt(d);
        const T r = T(1) / d;
        border();
        if (lane == 0) S[k * kWLD + k] = d;
        if (lane > k && lane < n) S[lane * kWLD + k] *= r;
        border();
        for (int e = lane; e < (n - k - 1) * (n - k - 1); e += 64) {
            const int i = k + 1 + e % (n - k - 1), j = k + 1 + e / (n - k - 1);
            if (j <= i) S[i * kWLD + j] -= S[i * kWLD + k] * S[j * kWLD + k];
        }
        border();
    }
    // L y = b, then Lᵀ x = y (column sweeps)
    for (int k = 0; k < n; ++k) {
        const T y = bs[k] / S[k * kWLD + k];
        border();
        if (lane == k) bs[k] = y;
        if (lane > k && lane < n) bs[lane] -= S[lane * kWLD + k] * y;
        border();
    }
    for (int k = n - 1; k >= 0; --k) {
        const T x = bs[k] / S[k * kWLD + k];
        border();
        if (lane == k) bs[k] = x;
        if (lane < k) bs[lane] -= S[k * kWLD + lane] * x;
        border();
    }
    for (int e = lane; e < n * n; e += 64) {
        const int i = e % n, j = e / n;
        if (j <= i) A[i + (int64_t)j * lda] = S[i * kWLD + j];
    }
    if (lane < n) b[lane] = bs[lane];
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
        for (int e = t; e < rows * pw; e += kBThreads) {
            const int r = e % rows, c = e / rows;
            if (c <= r) P[r * kPLD + c] = Ap[r + (int64_t)c * lda];
        }
        __syncthreads();
        // panel factor: pivot, scale, rank-1 update of the panel columns
        for (int c = 0; c < pw; ++c) {
            if (t == 0) {
                T d = P[c * kPLD + c];
                if (d < T(0) && !bad) bad = p0 + c + 1;
                d = bsqrt(d);
                P[c * kPLD + c] = d;
                sc[0] = T(1) / d;
            }
            __syncthreads();
            const T r = sc[0];
            for (int i = c + 1 + t; i < rows; i += kBThreads) P[i * kPLD + c] *= r;
            __syncthreads();
            for (int e = t; e < (rows - c - 1) * (pw - c - 1); e += kBThreads) {
                const int i = c + 1 + e % (rows - c - 1), j = c + 1 + e / (rows - c - 1);
                if (j <= i) P[i * kPLD + j] -= P[i * kPLD + c] * P[j * kPLD + c];
            }
            __syncthreads();
        }
        for (int e = t; e < rows * pw; e += kBThreads) {
            const int r = e % rows, c = e / rows;
            if (c <= r) Ap[r + (int64_t)c * lda] = P[r * kPLD + c];
        }
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
    // L y = b, 32 rows at a time
    for (int e = t; e < n; e += kBThreads) bs[e] = b[e];
    __syncthreads();
    for (int p0 = 0; p0 < n; p0 += kPW) {
        const int pw = min(kPW, n - p0);
        if (t < 64) {
            for (int k = 0; k < pw; ++k) {
                const T y = bs[p0 + k] / A[(p0 + k) + (int64_t)(p0 + k) * lda];
                border();
                if (t == k) bs[p0 + k] = y;
                if (t > k && t < pw) bs[p0 + t] -= A[(p0 + t) + (int64_t)(p0 + k) * lda] * y;
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
        __syncthreads();
        if (t < 64) {
            for (int k = pw - 1; k >= 0; --k) {
                const T x = bs[p0 + k] / A[(p0 + k) + (int64_t)(p0 + k) * lda];
                border();
                if (t == k) bs[p0 + k] = x;
                if (t < k) bs[p0 + t] -= A[(p0 + k) + (int64_t)(p0 + t) * lda] * x;
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
__global__ __launch_bounds__(kBThreads) void batched_chol_solve_kernel(
    int nitems, const int4 *__restrict__ items, const int *__restrict__ nvec,
    T *const *__restrict__ Aptr, const int *__restrict__ ldav, T *const *__restrict__ bptr,
    int *__restrict__ info)
{
    __shared__ T P[kBMax * kPLD];       // panel (workgroup path) or 4 wave matrices
    __shared__ T bs[kBMax];
    __shared__ T sc[2];
    static_assert(4 * (kWMax * kWLD + kWMax) <= kBMax * kPLD, "wave matrices fit the panel");
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int4 item = items[it];
        if (item.y == -2) {                 // one matrix, whole workgroup
            const int i = item.x;
            wg_chol_solve<T>(nvec[i], Aptr[i], ldav[i], bptr[i], info ? info + i : nullptr, P, bs,
                             sc);
        } else {                            // up to four small matrices, one per wave
            const int w = threadIdx.x >> 6;
            const int i = w == 0 ? item.x : (w == 1 ? item.y : (w == 2 ? item.z : item.w));
            if (i >= 0) {
                T *S = P + w * (kWMax * kWLD + kWMax);
                wave_chol_solve<T>(nvec[i], Aptr[i], ldav[i], bptr[i], info ? info + i : nullptr,
                                   S, S + kWMax * kWLD);
            }
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
    if (wave_max < 0 || wave_max > gla::kWMax) { gla::set_error("wave_max must be in [0, 32]"); return -4; }
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
