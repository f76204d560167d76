// Householder QR on gfx950: qrUnblocked!, qrBlocked!, QR2[Tuple{:QBlocked}]
// and the compact-WY applications lmul!(H, A, M) / lmul!(H', A, M)
// (src/qr.jl:64-146, src/householder.jl:82-159, src/juliaBLAS.jl:154-166).
//
// Panel (qr_panel_kernel).  A block column of width w ≤ 128 is factored
// column by column exactly as qrUnblocked! does — reflector! then the left
// reflectorApply! on the remaining panel columns — with the panel spread over
// P workgroups by rows and held in registers (thread (c, g) owns column c and
// 64 consecutive rows).  Each Householder step needs sums over all rows, so
// each column costs one grid-wide exchange: every workgroup publishes, for
// every panel column c, the partial Σ_{r>j} x_r a_rc (x = column j) and its
// row-j entry a_jc; after a barrier every workgroup adds the P partials in
// the same order, so all of them hold identical totals.  From these:
//   ‖x‖² = x_j² + Σ x_r²,  ν = copysign(‖x‖, x_j),  ξ = x_j + ν,  τ = ξ/ν,
//   v_r = x_r/ξ (r > j),   vᵀa_c = a_jc + (Σ x_r a_rc)/ξ,
//   a_c -= τ (vᵀa_c) v,    and for c < j, v_cᵀv_j = V_jc + (Σ V_rc x_r)/ξ,
// the last giving column j of the compact-WY T by the forward recurrence
// T[0:j, j] = -τ_j T[0:j,0:j] (V[:,0:j]ᵀ v_j), kept in LDS by workgroup 0.
// The kernel also writes V with its unit diagonal and zero upper part
// explicit (Vw), which the WY update reads as a plain matrix.
//
// Tall block columns (qr_panel_hr) are factored instead by two CholeskyQR
// passes and a reconstruction of the same reflectors from the orthonormal
// factor — the signed LU of its top block — in a fixed number of launches
// (hr_gram_kernel … hr_finish_kernel); the column-by-column kernel stays
// behind it as a device-gated fallback for block columns too ill-conditioned
// for CholeskyQR.
//
// Trailing update (src/householder.jl:119-157):  W = Vwᵀ A2 (MFMA GEMM,
// split-K), W = Tᵀ W, A2 -= Vw W — the same three contractions the
// reference makes, with V1 folded into Vw.
#include <algorithm>

#include "gla_coherent.hpp"
#include "gla_internal.hpp"

namespace gla {

template <typename T>
int unit_upper_explicit(hipStream_t s, int64_t n, const T *src, int64_t lds, T *dst, int64_t ldd);

namespace {

constexpr int kQrThreads = 1024;
constexpr int kQrNB = 128;   // max panel width
constexpr int kQrRPT = 16;   // rows per thread (LDS-reduction panel kernel)
constexpr int kQrRPL = 2;    // rows per lane (wave-butterfly panel kernel): 128 rows per workgroup

#define GLA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != GLA_OK) return _rc; \
    } while (0)

template <typename T>
struct QrPanelP {
    int64_t m;
    int w, kmax;
    T *A;
    int64_t lda;
    T *tau;
    T *Vw;
    int64_t ldv;
    T *Tw;
    int64_t ldt;
    T *part;
    unsigned *bar;
    unsigned bar_base;
    // polls before a grid wait gives up (raising *err); added to every wait
    // target when nonzero (a test setting: no wait can then be met)
    unsigned spin_limit;
    unsigned bar_skew;
    int *err;
    // null: always run; otherwise run only if *run_if != 0 (the fallback
    // behind the reconstructed-Householder panel)
    const int *run_if;
};

template <typename T>
__device__ __forceinline__ T dsqrt(T x);
template <>
__device__ __forceinline__ double dsqrt<double>(double x) { return __builtin_sqrt(x); }
template <>
__device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_sqrtf(x); }
template <typename T>
__device__ __forceinline__ T dcopysign(T a, T b);
template <>
__device__ __forceinline__ double dcopysign<double>(double a, double b) { return __builtin_copysign(a, b); }
template <>
__device__ __forceinline__ float dcopysign<float>(float a, float b) { return __builtin_copysignf(a, b); }

// Grid-wide exchange.  Partials are written with agent-scope (sc1) stores,
// each writer waits for its stores to complete (vmcnt(0)), then one thread
// per workgroup bumps the arrival counter.  Readers poll the counter and read
// the partials with agent-scope loads.  No agent-scope fence is used: on a
// multi-XCD gfx950 one would write back and invalidate the whole L2 of the
// XCD (buffer_wbl2 / buffer_inv), stalling the concurrent trailing update.
// Waits are bounded: after `limit` polls *err is raised and the workgroup
// carries on with the partials it can see (every read stays inside the
// exchange buffer), so no launch can hang; the host reports the flag.

__device__ __forceinline__ void grid_arrive(unsigned *bar)
{
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void grid_wait(unsigned *bar, unsigned target, int *err,
                                          unsigned limit)
{
    if (threadIdx.x == 0) {
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > limit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
}

// Column jj of T by the forward recurrence T[0:jj,jj] = -τ T[0:jj,0:jj] g,
// the sum over l split four ways across the workgroup (workgroup 0 only).
template <typename T, int NB, int NTH>
__device__ __forceinline__ void t_column(T *Ts, const T *g, T *tpart, int jj, T tau)
{
    constexpr int H = NTH / NB;
    const int t = threadIdx.x, i = t % NB, h = t / NB;
    if (i < jj) {
        T s = T(0);
        for (int l = i + h; l < jj; l += H) s += Ts[i + l * NB] * g[l];
        tpart[h * NB + i] = s;
    }
    __syncthreads();
    if (t < jj) {
        T s = T(0);
#pragma unroll
        for (int hh = 0; hh < H; ++hh) s += tpart[hh * NB + t];
        Ts[t + jj * NB] = -tau * s;
    }
    if (t == jj) Ts[jj + jj * NB] = tau;
}

// Sum over the wave of v[c], c < 8, by a transposing butterfly: after the
// xor-32/16/8 exchanges each lane keeps one column's partial, the xor-4/2/1
// steps finish it.  Lane l returns the total of column
// c(l) = 4·bit5(l) + 2·bit4(l) + bit3(l).  Fixed order: every run and every
// lane adds the same values in the same order.
template <typename T>
__device__ __forceinline__ T wave_reduce8(const T v[8])
{
    const int lane = threadIdx.x & 63;
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
    T u[4], w[2];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const T send = b5 ? v[k] : v[k + 4];
        u[k] = (b5 ? v[k + 4] : v[k]) + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const T send = b4 ? u[k] : u[k + 2];
        w[k] = (b4 ? u[k + 2] : u[k]) + __shfl_xor(send, 16);
    }
    const T send = b3 ? w[0] : w[1];
    T z = (b3 ? w[1] : w[0]) + __shfl_xor(send, 8);
    z += __shfl_xor(z, 4);
    z += __shfl_xor(z, 2);
    z += __shfl_xor(z, 1);
    return z;
}

// A panel launched as the fallback of the reconstructed-Householder path
// runs only if that path failed (*run_if != 0).  When it does not run it
// still advances the grid-barrier counter by what its column loop adds
// (kmax arrivals of each of the P workgroups), so the barrier targets of
// later panels stay valid.
template <typename T>
__device__ __forceinline__ bool panel_skipped(const QrPanelP<T> &p)
{
    if (!p.run_if || *p.run_if != 0) return false;
    if (gridDim.x > 1 && blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_fetch_add(p.bar, (unsigned)p.kmax * gridDim.x, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Householder panel, one column at a time as qrUnblocked! does
// (src/qr.jl:86-111), over P workgroups that each hold R = 64·RPL rows of
// the m×w panel in registers: wave v holds CPW = NB/(NTH/64) consecutive
// columns, lane l holds rows l + 64q (q < RPL).  Per column j:
//   (1) the wave holding column j puts x = column j in LDS;
//   (2) every lane forms its rows' share of aᶜᵀx (rows > j) for its
//       wave's columns, a wave butterfly sums them, one lane per column publishes
//       the workgroup's partial (device-coherent store); workgroup 0 also
//       publishes row j (it holds rows 0..R-1, and j < w ≤ 128 ≤ R);
//   (3) grid barrier; every workgroup gathers the P partials of each column
//       in a fixed order, so all hold identical totals;
//   (4) Julia's reflector! from the totals (ν = copysign(‖x‖, x_j),
//       ξ = x_j + ν, τ = ξ/ν, v = x/ξ), each lane dividing its own rows;
//   (5) columns c > j take a_c -= τ(a_jc + aᶜᵀx/ξ)·v (reflectorApply!),
//       column j becomes (β = −ν, v).
// Workgroup 0 also builds the compact-WY T column by column from the same
// totals (vᵢᵀv_j = a_ji + aⁱᵀx/ξ), overlapped with the grid barrier.
template <typename T, int NB, int RPL, int NTH>
__global__ __launch_bounds__(NTH) void qr_panel2_kernel(QrPanelP<T> p)
{
    constexpr int R = 64 * RPL;
    constexpr int H = NTH / NB;                  // partial-sum groups of the gather
    constexpr int CPW = NB / (NTH / 64);         // columns per wave
    static_assert(CPW == 8, "eight columns per wave");
    static_assert(R >= NB, "row j of every column lives in workgroup 0");

    __shared__ T xs[2][R];
    __shared__ T red[H][NB];        // gather sub-sums; also the T recurrence's scratch
    __shared__ T tot[NB];           // Σ_{r>j} a_rc x_r over the panel
    __shared__ T rowj[NB];          // a_jc
    __shared__ T gsh[NB];           // v_iᵀ v_j, i < j (workgroup 0)
    __shared__ T Ts[NB * NB];       // T, column-major (workgroup 0)
    T *const tpart = &red[0][0];

    if (panel_skipped(p)) return;
    const int t = threadIdx.x, lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wg = blockIdx.x, P = gridDim.x;
    const int64_t m = p.m, lda = p.lda;
    const int64_t row0 = (int64_t)wg * R;
    const int c0 = CPW * wv;                     // first column of this wave

    T a[CPW][RPL];
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            const int64_t r = row0 + lane + 64 * q;
            a[c][q] = (r < m && c0 + c < p.w) ? p.A[r + (int64_t)(c0 + c) * lda] : T(0);
        }
    if (wg == 0)
        for (int e = t; e < NB * NB; e += NTH) Ts[e] = T(0);
    T tau_prev = T(0);

    for (int j = 0; j < p.kmax; ++j) {
        const int buf = j & 1;
        const int jw = j / CPW, jc = j % CPW;
        // (1) column j to LDS
        if (wv == jw) {
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                T v = a[0][q];
#pragma unroll
                for (int c = 1; c < CPW; ++c) v = (c == jc) ? a[c][q] : v;
                xs[buf][lane + 64 * q] = v;
            }
        }
        __syncthreads();
        T x[RPL];
#pragma unroll
        for (int q = 0; q < RPL; ++q) x[q] = xs[buf][lane + 64 * q];
        // (2) partial aᶜᵀx over rows > j, summed over the wave
        T pr[CPW];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            T s = T(0);
#pragma unroll
            for (int q = 0; q < RPL; ++q) s += (row0 + lane + 64 * q > j) ? a[c][q] * x[q] : T(0);
            pr[c] = s;
        }
        const T z = wave_reduce8(pr);
        const int zc = c0 + 4 * ((lane >> 5) & 1) + 2 * ((lane >> 4) & 1) + ((lane >> 3) & 1);
        const bool zl = (lane & 7) == 0;
        // row j of this wave's columns (workgroup 0 holds rows 0..R-1)
        const bool rl = wg == 0 && lane == (j & 63);
        const int rq = j >> 6;
        if (P > 1) {
            T *part = p.part + (size_t)buf * (P + 1) * NB;
            if (zl) store_coherent(part + (size_t)wg * NB + zc, z);
            if (rl) {
#pragma unroll
                for (int c = 0; c < CPW; ++c) {
                    T v = a[c][0];
#pragma unroll
                    for (int q = 1; q < RPL; ++q) v = (q == rq) ? a[c][q] : v;
                    store_coherent(part + (size_t)P * NB + c0 + c, v);
                }
            }
            stores_done();
            grid_arrive(p.bar);
            if (wg == 0 && j > 0) t_column<T, NB, NTH>(Ts, gsh, tpart, j - 1, tau_prev);
            grid_wait(p.bar, p.bar_base + (unsigned)(j + 1) * (unsigned)P + p.bar_skew, p.err,
                      p.spin_limit);
            // (3) totals in a fixed order: H sub-sums per column over the
            // partials p ≡ h (mod H), eight loads in flight, then the H in
            // order; row j travels in the first batch's round trip
            {
                const int e = t % NB, h = t / NB;
                T s = T(0);
                for (int pp0 = h; pp0 < P; pp0 += 8 * H) {
                    const T *ptr[8];
                    T v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int pp = pp0 + u * H;
                        ptr[u] = part + (size_t)(pp < P ? pp : h) * NB + e;
                    }
                    if (pp0 == h) {
                        T rv;
                        load9_coherent(ptr, v, part + (size_t)P * NB + e, rv);
                        if (t < NB) rowj[t] = rv;
                    } else {
                        load8_coherent(ptr, v);
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        if (pp0 + u * H < P) s += v[u];
                }
                red[h][e] = s;
                __syncthreads();
                if (t < NB) {
                    T s2 = T(0);
#pragma unroll
                    for (int hh = 0; hh < H; ++hh) s2 += red[hh][t];
                    tot[t] = s2;
                }
            }
        } else {
            if (zl) tot[zc] = z;
            if (rl) {
#pragma unroll
                for (int c = 0; c < CPW; ++c) {
                    T v = a[c][0];
#pragma unroll
                    for (int q = 1; q < RPL; ++q) v = (q == rq) ? a[c][q] : v;
                    rowj[c0 + c] = v;
                }
            }
            __syncthreads();   // red is the T recurrence's scratch below
            if (j > 0) t_column<T, NB, NTH>(Ts, gsh, tpart, j - 1, tau_prev);
        }
        __syncthreads();

        // (4) the reflector of column j; the divisions by ξ of reflector!
        // (x[i] /= ξ1) are taken as products with 1/ξ, one rounding apart
        const T xj = rowj[j];
        const T nrm = dsqrt(xj * xj + tot[j]);
        const bool live = nrm != T(0);
        const T nu = dcopysign(nrm, xj);
        const T xi = xj + nu;
        const T tauj = live ? xi / nu : T(0);
        const T rxi = live ? T(1) / xi : T(0);
        if (wg == 0 && t < j) gsh[t] = live ? rowj[t] + tot[t] * rxi : T(0);

        // (5) apply to the remaining columns; column j becomes (β, v)
        if (live) {
            T v[RPL];
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                const int64_t r = row0 + lane + 64 * q;
                v[q] = r > j ? x[q] * rxi : (r == j ? T(1) : T(0));
            }
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int cg = c0 + c;
                if (cg > j && cg < p.w) {
                    const T coef = tauj * (rowj[cg] + tot[cg] * rxi);
#pragma unroll
                    for (int q = 0; q < RPL; ++q) a[c][q] -= coef * v[q];
                } else if (cg == j) {
#pragma unroll
                    for (int q = 0; q < RPL; ++q) {
                        const int64_t r = row0 + lane + 64 * q;
                        a[c][q] = r > j ? v[q] : (r == j ? -nu : a[c][q]);
                    }
                }
            }
        }
        if (wg == 0 && t == 0) p.tau[j] = tauj;
        tau_prev = tauj;
    }
    __syncthreads();
    if (wg == 0 && p.kmax > 0) {
        t_column<T, NB, NTH>(Ts, gsh, tpart, p.kmax - 1, tau_prev);
        __syncthreads();
    }

    // Write back the factored panel and the explicit V.
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int cg = c0 + c;
        if (cg >= p.w) continue;
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            const int64_t r = row0 + lane + 64 * q;
            if (r >= m) continue;
            p.A[r + (int64_t)cg * lda] = a[c][q];
            if (p.Vw && cg < p.kmax)
                p.Vw[r + (int64_t)cg * p.ldv] = r < cg ? T(0) : (r == cg ? T(1) : a[c][q]);
        }
    }
    if (wg == 0 && p.Tw) {
        for (int e = t; e < p.kmax * p.kmax; e += NTH) {
            const int i = e % p.kmax, jj = e / p.kmax;
            p.Tw[i + (int64_t)jj * p.ldt] = Ts[i + jj * NB];
        }
    }
}

template <typename T, int NB, int RPT>
__global__ __launch_bounds__(kQrThreads) void qr_panel_kernel(QrPanelP<T> p)
{
    constexpr int G = kQrThreads / NB;   // row groups
    constexpr int R = G * RPT;           // rows per workgroup
    static_assert(NB % 64 == 0, "wave-uniform row groups");
    static_assert(2 * NB <= kQrThreads, "one total per thread");

    __shared__ T xs[R];             // column j of this workgroup's rows, then v
    __shared__ T red[2][G][NB];     // per-group partial sums; also the split
                                    // sums of the T recurrence and of the gather
    __shared__ T tot[2 * NB];       // grid totals for the current column
    __shared__ T gsh[NB];           // v_iᵀ v_j for i < j (workgroup 0)
    __shared__ T Ts[NB * NB];       // T, column-major (workgroup 0)
    T *const tpart = &red[0][0][0];

    if (panel_skipped(p)) return;
    const int t = threadIdx.x;
    // g is uniform across a wave (NB is a multiple of 64): keep it, and the
    // row indices derived from it, in scalar registers
    const int c = t % NB, g = __builtin_amdgcn_readfirstlane(t / NB);
    const int wg = blockIdx.x, P = gridDim.x;
    const int64_t m = p.m;
    const int64_t row0 = (int64_t)wg * R;             // first row of this workgroup
    const int rb = (int)row0 + g * RPT;               // first row of this thread
    const bool cact = c < p.w;
    const int64_t lda = p.lda;

    T a[RPT];
    {
        const T *col = p.A + (int64_t)c * lda + rb;   // this thread's rows of column c
        const int nr = cact ? (int)std::min<int64_t>(RPT, m - rb) : 0;
#pragma unroll
        for (int q = 0; q < RPT; ++q) a[q] = q < nr ? col[q] : T(0);
    }
    if (wg == 0)
        for (int e = t; e < NB * NB; e += kQrThreads) Ts[e] = T(0);
    T tau_prev = T(0);

    for (int j = 0; j < p.kmax; ++j) {
        // (1) column j of this workgroup's rows to LDS
        if (c == j) {
#pragma unroll
            for (int q = 0; q < RPT; ++q) xs[g * RPT + q] = a[q];
        }
        __syncthreads();

        // (2) partial Σ_{r>j} a_rc x_r and the row-j entry of column c
        T dot = T(0), rowj = T(0);
        {
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int r = rb + q;
                const T pr = a[q] * xs[g * RPT + q];
                dot += (r > j) ? pr : T(0);
                rowj = (r == j) ? a[q] : rowj;
            }
            if (!cact) dot = rowj = T(0);
        }
        red[0][g][c] = dot;
        red[1][g][c] = rowj;
        __syncthreads();

        if (P > 1) {
            T *part = p.part + (size_t)(j & 1) * P * 2 * NB;
            if (t < NB) {
                T s0 = T(0), s1 = T(0);
#pragma unroll
                for (int gg = 0; gg < G; ++gg) {
                    s0 += red[0][gg][t];
                    s1 += red[1][gg][t];
                }
                store_coherent(part + (size_t)wg * 2 * NB + t, s0);
                store_coherent(part + (size_t)wg * 2 * NB + NB + t, s1);
                stores_done();
            }
            grid_arrive(p.bar);
            // workgroup 0 finishes T for the previous column while the grid gathers
            if (wg == 0 && j > 0) t_column<T, NB, kQrThreads>(Ts, gsh, tpart, j - 1, tau_prev);
            grid_wait(p.bar, p.bar_base + (unsigned)(j + 1) * (unsigned)P + p.bar_skew, p.err,
                      p.spin_limit);
            // (3) identical totals in every workgroup: fixed summation order.
            // GS threads per total each add every GS-th partial (independent
            // loads, one round trip each), then the GS sums are added in order.
            {
                constexpr int GS = kQrThreads / (2 * NB);
                const int e = t % (2 * NB), h = t / (2 * NB);
                const T *src = part + e;
                T s = T(0);
                // this thread's partials are pp = h, h+GS, h+2GS, ... < P;
                // eight per round trip, added in increasing pp
                for (int pp0 = h; pp0 < P; pp0 += 8 * GS) {
                    const T *ptr[8];
                    T v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int pp = pp0 + u * GS;
                        ptr[u] = src + (size_t)(pp < P ? pp : h) * 2 * NB;
                    }
                    load8_coherent(ptr, v);
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        if (pp0 + u * GS < P) s += v[u];
                }
                tpart[h * 2 * NB + e] = s;
                __syncthreads();
                if (t < 2 * NB) {
                    T s = T(0);
#pragma unroll
                    for (int hh = 0; hh < GS; ++hh) s += tpart[hh * 2 * NB + t];
                    tot[t] = s;
                }
            }
        } else {
            if (t < NB) {
                T s0 = T(0), s1 = T(0);
#pragma unroll
                for (int gg = 0; gg < G; ++gg) {
                    s0 += red[0][gg][t];
                    s1 += red[1][gg][t];
                }
                tot[t] = s0;
                tot[NB + t] = s1;
            }
            __syncthreads();   // red is reused as tpart below
            if (wg == 0 && j > 0) t_column<T, NB, kQrThreads>(Ts, gsh, tpart, j - 1, tau_prev);
        }
        __syncthreads();

        // (4) the reflector of column j (Julia's reflector!)
        const T xj = tot[NB + j];
        const T nrm = dsqrt(xj * xj + tot[j]);
        const bool live = nrm != T(0);
        const T nu = dcopysign(nrm, xj);
        const T xi = xj + nu;
        const T tauj = live ? xi / nu : T(0);
        // v_r = x_r / ξ below the diagonal (x unchanged when the norm is 0)
        if (live)
            for (int e = t; e < R; e += kQrThreads)
                if (row0 + e > j) xs[e] = xs[e] / xi;
        if (wg == 0 && t < j) gsh[t] = live ? tot[NB + t] + tot[t] / xi : T(0);
        __syncthreads();

        // (5) apply to the remaining columns; column j becomes (β, v)
        if (cact && live) {
            if (c > j) {
                const T coef = tauj * (tot[NB + c] + tot[c] / xi);
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int r = rb + q;
                    const T v = (r == j) ? T(1) : xs[g * RPT + q];
                    const T upd = a[q] - coef * v;
                    a[q] = (r >= j) ? upd : a[q];
                }
            } else if (c == j) {
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int r = rb + q;
                    a[q] = (r > j) ? xs[g * RPT + q] : (r == j ? -nu : a[q]);
                }
            }
        }
        if (wg == 0 && t == 0) p.tau[j] = tauj;
        tau_prev = tauj;
        __syncthreads();
    }
    if (wg == 0 && p.kmax > 0) {
        t_column<T, NB, kQrThreads>(Ts, gsh, tpart, p.kmax - 1, tau_prev);
        __syncthreads();
    }

    // Write back the factored panel and the explicit V.
    {
        T *col = p.A + (int64_t)c * lda + rb;
        const int nr = cact ? (int)std::min<int64_t>(RPT, m - rb) : 0;
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (q < nr) col[q] = a[q];
        if (p.Vw && c < p.kmax) {
            T *vc = p.Vw + (int64_t)c * p.ldv + rb;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int r = rb + q;
                if (q < nr) vc[q] = r < c ? T(0) : (r == c ? T(1) : a[q]);
            }
        }
    }
    if (wg == 0 && p.Tw) {
        for (int e = t; e < p.kmax * p.kmax; e += kQrThreads) {
            const int i = e % p.kmax, jj = e / p.kmax;
            p.Tw[i + (int64_t)jj * p.ldt] = Ts[i + jj * NB];
        }
    }
}

// ---------------------------------------------------------- small kernels --
// dst (m×n) = unit-lower explicit copy of src: 1 on the diagonal, 0 above.
template <typename T>
__global__ void unit_lower_copy_kernel(int64_t m, int64_t n, const T *src, int64_t lds, T *dst,
                                       int64_t ldd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= m) return;
    dst[i + j * ldd] = i < j ? T(0) : (i == j ? T(1) : src[i + j * lds]);
}

// dst (n×n) = upper triangle of src (incl. diagonal), zeros below.
template <typename T>
__global__ void upper_copy_kernel(int64_t n, const T *src, int64_t lds, T *dst, int64_t ldd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    dst[i + j * ldd] = i <= j ? src[i + j * lds] : T(0);
}

template <typename T>
__global__ void copy_kernel(int64_t m, int64_t n, const T *src, int64_t lds, T *dst, int64_t ldd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i < m) dst[i + j * ldd] = src[i + j * lds];
}

// Tm = I + D·striu(G) as a full k×k matrix (src/qr.jl:69-74 before inv!).
template <typename T>
__global__ void tb_init_kernel(int64_t k, const T *Gm, int64_t ldg, const T *tau, T *Tm,
                               int64_t ldt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= k) return;
    Tm[i + j * ldt] = i < j ? tau[i] * Gm[i + j * ldg] : (i == j ? T(1) : T(0));
}

// T[j,j] = τ_j; T[i,j] *= τ_j (src/qr.jl:76-81).
template <typename T>
__global__ void tb_scale_kernel(int64_t k, const T *tau, T *Tm, int64_t ldt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= k) return;
    if (i == j) Tm[i + j * ldt] = tau[j];
    else if (i < j) Tm[i + j * ldt] *= tau[j];
}

// In-place inverse of a unit upper-triangular k×k block (k ≤ 128), one
// workgroup: column j of the inverse is -X[0:j,0:j]·U[0:j,j] with the
// columns before j already inverted (the column sweep of LAPACK trti2).
template <typename T, int NB>
__global__ __launch_bounds__(256) void trtri_unit_upper_leaf(int k, T *U, int64_t ldu)
{
    __shared__ T S[NB * (NB + 1)];
    constexpr int LD = NB + 1;
    for (int e = threadIdx.x; e < k * k; e += 256) {
        const int i = e % k, j = e / k;
        S[i + j * LD] = i < j ? U[i + (int64_t)j * ldu] : (i == j ? T(1) : T(0));
    }
    __syncthreads();
    for (int j = 1; j < k; ++j) {
        T s = T(0);
        const int i = threadIdx.x;
        if (i < j) {
            s = S[i + j * LD];
            // unrolled so that the LDS reads of eight terms are issued
            // together; the sum keeps its order
#pragma unroll 8
            for (int l = i + 1; l < j; ++l) s += S[i + l * LD] * S[l + j * LD];
        }
        __syncthreads();
        if (i < j) S[i + j * LD] = -s;
        __syncthreads();
    }
    for (int e = threadIdx.x; e < k * k; e += 256) {
        const int i = e % k, j = e / k;
        if (i < j) U[i + (int64_t)j * ldu] = S[i + j * LD];
    }
}

template <typename T>
int launch2d(hipStream_t s, void (*kern)(int64_t, int64_t, const T *, int64_t, T *, int64_t),
             int64_t m, int64_t n, const T *src, int64_t lds, T *dst, int64_t ldd)
{
    if (m <= 0 || n <= 0) return GLA_OK;
    hipLaunchKernelGGL(kern, dim3((unsigned)cdiv(m, 256), (unsigned)n), dim3(256), 0, s, m, n, src,
                       lds, dst, ldd);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template <typename T>
int unit_lower_copy(hipStream_t s, int64_t m, int64_t n, const T *src, int64_t lds, T *dst,
                    int64_t ldd)
{
    return launch2d<T>(s, unit_lower_copy_kernel<T>, m, n, src, lds, dst, ldd);
}
template <typename T>
int copy_mat(hipStream_t s, int64_t m, int64_t n, const T *src, int64_t lds, T *dst, int64_t ldd)
{
    return launch2d<T>(s, copy_kernel<T>, m, n, src, lds, dst, ldd);
}
template <typename T>
int upper_copy(hipStream_t s, int64_t n, const T *src, int64_t lds, T *dst, int64_t ldd)
{
    if (n <= 0) return GLA_OK;
    hipLaunchKernelGGL(upper_copy_kernel<T>, dim3((unsigned)cdiv(n, 256), (unsigned)n), dim3(256),
                       0, s, n, src, lds, dst, ldd);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// Unit upper-triangular inverse, recursive: X12 = -X11 (U12 X22).
// ws must hold k*k/4 + 1 elements.
template <typename T>
int trtri_unit_upper(hipStream_t s, int64_t k, T *U, int64_t ldu, T *ws)
{
    if (k <= 0) return GLA_OK;
    if (k <= kQrNB) {
        hipLaunchKernelGGL((trtri_unit_upper_leaf<T, kQrNB>), dim3(1), dim3(256), 0, s, (int)k,
                           U, ldu);
        GLA_LAUNCH_CHECK();
        return GLA_OK;
    }
    int64_t k1 = std::max<int64_t>(kQrNB, (k / 2) / kQrNB * kQrNB);
    int64_t k2 = k - k1;
    T *U12 = U + k1 * ldu, *U22 = U + k1 + k1 * ldu;
    GLA_TRY(trtri_unit_upper<T>(s, k1, U, ldu, ws));
    GLA_TRY(trtri_unit_upper<T>(s, k2, U22, ldu, ws));
    // The leaf inverse keeps U's strict lower part untouched, so the GEMMs
    // read explicit-triangular copies.
    T *X11 = ws;                         // k1×k1 (upper incl. unit diag)
    T *W = ws + k1 * k1;                 // k1×k2
    T *X22 = W + k1 * k2;                // k2×k2
    GLA_TRY(unit_upper_explicit(s, k1, U, ldu, X11, k1));
    GLA_TRY(unit_upper_explicit(s, k2, U22, ldu, X22, k2));
    GLA_TRY(gemm<T>(s, 'N', 'N', 0, k1, k2, k2, T(1), U12, ldu, X22, k2, T(0), W, k1));
    GLA_TRY(gemm<T>(s, 'N', 'N', 0, k1, k2, k1, T(-1), X11, k1, W, k1, T(0), U12, ldu));
    return GLA_OK;
}

template <typename T>
__global__ void unit_upper_copy_kernel(int64_t n, const T *src, int64_t lds, T *dst, int64_t ldd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    dst[i + j * ldd] = i < j ? src[i + j * lds] : (i == j ? T(1) : T(0));
}

// Zero the strict upper triangle of the top k rows of Vw (m×k) and all of
// T (k×k): the sub-panel launches write only their own diagonal blocks.
template <typename T>
__global__ void qr_sub_clear_kernel(int k, T *Vw, int64_t ldv, T *Tw, int64_t ldt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= k) return;
    if (i < j) Vw[i + (int64_t)j * ldv] = T(0);
    Tw[i + (int64_t)j * ldt] = T(0);
}

// Compact-WY T of a block column factored as sub-panels of width ib, from
// the sub-panels' diagonal blocks T_ss (already in T) and G = VᵀV:
//   T[0:s0, blk] = −T[0:s0, 0:s0] · (G[0:s0, blk] · T_ss),   s0 = s·ib,
// the blocked form of the forward recurrence of src/qr.jl:64-83 (LAPACK
// dlarft: T12 = −T11 V1ᵀV2 T22).  One workgroup; T (k ≤ 128) in LDS, the
// product Y kept transposed in T's zero lower part (rows s0.., columns
// 0..s0-1), which no later step reads.
template <typename T>
__global__ __launch_bounds__(1024) void qr_t_merge_kernel(int k, int ib, T *Tw, int64_t ldt,
                                                          const T *G, int64_t ldg)
{
    constexpr int LD = 129;
    __shared__ T Ts[128 * LD];
    const int t = threadIdx.x;
    for (int e = t; e < k * k; e += 1024) {
        const int i = e % k, j = e / k;
        Ts[i + j * LD] = Tw[i + (int64_t)j * ldt];
    }
    __syncthreads();
    for (int s0 = ib; s0 < k; s0 += ib) {
        const int b = min(ib, k - s0);
        // Y (s0×b) = G[0:s0, s0:s0+b] · T_ss, T_ss upper triangular
        for (int e = t; e < s0 * b; e += 1024) {
            const int i = e % s0, c = e / s0;
            T acc = T(0);
            for (int l = 0; l <= c; ++l)
                acc += G[i + (int64_t)(s0 + l) * ldg] * Ts[s0 + l + (s0 + c) * LD];
            Ts[s0 + c + i * LD] = acc;                      // Yᵀ
        }
        __syncthreads();
        // T[0:s0, s0:s0+b] = −T[0:s0, 0:s0] · Y, T[0:s0,0:s0] upper triangular
        for (int e = t; e < s0 * b; e += 1024) {
            const int i = e % s0, c = e / s0;
            T acc = T(0);
            for (int l = i; l < s0; ++l) acc += Ts[i + l * LD] * Ts[s0 + c + l * LD];
            Ts[i + (s0 + c) * LD] = -acc;
        }
        __syncthreads();
    }
    for (int e = t; e < k * k; e += 1024) {
        const int i = e % k, j = e / k;
        if (i < j) Tw[i + (int64_t)j * ldt] = Ts[i + j * LD];
    }
}

// ---------------------------------------------------------------------------
// Reconstructed-Householder panel, for tall block columns.
//
// The Householder QR that qrUnblocked! computes for a panel P (m×k, m ≥ k),
// P = Q_h R_h with Q_h = (I − V T Vᵀ)[I;0], V unit lower trapezoidal and
// V₁ its top k×k block, satisfies
//     [I;0] − Q_h = V · (T V₁ᵀ),
// an LU factorisation without pivoting (V unit lower; T V₁ᵀ upper with
// diagonal τ, since T's diagonal is τ and V₁'s is 1).  Every other thin QR
// P = Q R' has Q = Q_h S for a diagonal sign matrix S, and LU factors are
// unique, so the Householder factors follow from any orthonormal Q:
//   (1) Q, R' by two CholeskyQR passes: G = XᵀX (MFMA GEMM, split-K),
//       G = L Lᵀ (the 128-wide diagonal-block Cholesky kernel), X ← X L⁻ᵀ;
//       R' = L₂ᵀ L₁ᵀ;
//   (2) the LU of S − Q[0:k, :] with S chosen during the elimination:
//       S_jj = copysign(1, −q̂_jj), q̂_jj the updated diagonal entry, which
//       makes the pivot S_jj (1 + |q̂_jj|), i.e. τ_j = 1 + |q̂_jj| ∈ [1, 2],
//       the range reflector! gives (τ = 1 + |ξ₁|/‖x‖); the rows below the
//       top block follow by a right solve, V₂ = −Q[k:m, :] U⁻¹;
//   (3) V = L, τ_j = S_jj U_jj, T = (U S) V₁⁻ᵀ, R_h = S R'.
// These are the reflectors, τ, T and R of the column-by-column algorithm,
// up to rounding, from O(1) grid-wide steps instead of one exchange per
// column.  The path checks itself: a Cholesky breakdown, or a second Gram
// matrix farther than 0.5 from I in the Frobenius norm (the first pass left
// Q with κ² > 3), sets a device flag; the panel is then left untouched and
// the column-by-column panel kernel, launched behind the path and gated on
// the flag, factors it instead.

// Gram matrix G = XᵀX of an m×k panel (k ≤ 128), lower 16×16 tiles only:
// workgroup p sums rows of 32-row chunks p, p + P, … on MFMA (the chunk
// staged in LDS, wave w takes tiles w, w + 4, …) and writes its partial
// tiles to part[p][tile][256] (element (i, j) of a tile at 16 i + j);
// hr_gram_reduce_kernel adds the P partials in a fixed order.  With
// copy_to (the first pass) the chunks are also copied there (Q ← P) and
// the path's flags are cleared (flag[0]: the path failed; flag[1]: the
// second Gram matrix is not within 0.1·√eps of I).
constexpr int kGrRB = 32;
constexpr int kGrLD = 132;
// most workgroups of the Gram pass (one 32-row chunk each up to 8192 rows;
// 64 workgroups of four chunks were slower: 31 + 5 µs against 14 + 10 µs
// for the pass and its reduction at 8192×128)
constexpr int kGrBlocks = 256;

__host__ __device__ inline int hr_ntiles(int k)
{
    const int nt = (k + 15) / 16;
    return nt * (nt + 1) / 2;
}

template <typename T>
__global__ __launch_bounds__(256) void hr_gram_kernel(int64_t m, int k, const T *X, int64_t ldx,
                                                      T *part, T *copy_to, int64_t ldc,
                                                      int *flag)
{
    typedef PanelMfma<T> Mf;
    typedef typename Mf::acc_t acc_t;
    __shared__ T Xs[kGrRB * kGrLD];
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int nt = (k + 15) / 16, ntl = nt * (nt + 1) / 2;
    const int li = lane & 15, lk = lane >> 4;
    if (copy_to && blockIdx.x == 0 && t == 0) {
        flag[0] = 0;
        flag[1] = 0;
    }
    acc_t acc[9];
    int tbi[9], tbj[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        acc[q] = acc_t{T(0), T(0), T(0), T(0)};
        const int ti = w + 4 * q;
        int bi = 0;
        while ((bi + 1) * (bi + 2) / 2 <= ti) ++bi;
        tbi[q] = bi;
        tbj[q] = ti - bi * (bi + 1) / 2;
    }
    for (int64_t r0 = (int64_t)blockIdx.x * kGrRB; r0 < m; r0 += (int64_t)gridDim.x * kGrRB) {
#pragma unroll
        for (int e = t; e < kGrRB * 128; e += 256) {
            const int r = e & (kGrRB - 1), c = e / kGrRB;
            const bool in = r0 + r < m && c < k;
            const T v = in ? X[r0 + r + (int64_t)c * ldx] : T(0);
            Xs[r * kGrLD + c] = v;
            if (copy_to && in) copy_to[r0 + r + (int64_t)c * ldc] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            if (w + 4 * q >= ntl) break;
            const int ca = tbi[q] * 16 + li, cb = tbj[q] * 16 + li;
#pragma unroll
            for (int kq = 0; kq < kGrRB / 4; ++kq) {
                const int r = 4 * kq + lk;
                acc[q] = Mf::mma(Xs[r * kGrLD + ca], Xs[r * kGrLD + cb], acc[q]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int ti = w + 4 * q;
        if (ti >= ntl) break;
        T *dst = part + ((size_t)blockIdx.x * ntl + ti) * 256;
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[Mf::row(lane, r) * 16 + li] = acc[q][r];
    }
}

// G (k×k, ld k, both triangles) = Σ_p part[p], one workgroup per tile, the
// P partials summed in four interleaved groups then across the groups, a
// fixed order.  With check: flags the path (atomicCAS(flag, 0, −1)) if the
// tile's share of ‖G − I‖_F² exceeds 0.25 / ntiles, so an unflagged G has
// ‖G − I‖_F ≤ 0.5; a non-finite G is flagged too.  It also raises flag[1]
// unless the share is within 0.01·eps / ntiles (‖G − I‖_F ≤ 0.1·√eps),
// where the first-order factor of I + E replaces the full Cholesky
// factorisation (chol_factor_block_near_identity).
template <typename T>
__global__ __launch_bounds__(1024) void hr_gram_reduce_kernel(int k, int P, const T *part, T *G,
                                                              int *flag, int check)
{
    __shared__ T red[4][256];
    __shared__ double wsum[4];
    const int t = threadIdx.x, e = t & 255, g = t >> 8;
    const int nt = (k + 15) / 16, ntl = nt * (nt + 1) / 2, ti = blockIdx.x;
    int bi = 0;
    while ((bi + 1) * (bi + 2) / 2 <= ti) ++bi;
    const int bj = ti - bi * (bi + 1) / 2;
    T s = T(0);
#pragma unroll 16
    for (int p = g; p < P; p += 4) s += part[((size_t)p * ntl + ti) * 256 + e];
    red[g][e] = s;
    __syncthreads();
    if (t >= 256) return;
    const T v = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
    const int i = bi * 16 + (e >> 4), j = bj * 16 + (e & 15);
    double c = 0.0;
    if (i < k && j < k) {
        G[i + (int64_t)j * k] = v;
        G[j + (int64_t)i * k] = v;
        const double d = (double)v - (i == j ? 1.0 : 0.0);
        c = i > j ? 2.0 * d * d : (i == j ? d * d : 0.0);
    }
    if (!check) return;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((t & 63) == 0) wsum[t >> 6] = c;
    __syncthreads();
    if (t == 0) {
        const double tot = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
        if (!(tot <= 0.25 / ntl)) atomicCAS(flag, 0, -1);
        const double eps = sizeof(T) == 8 ? 2.220446049250313e-16 : 1.1920928955078125e-7;
        if (!(tot <= 0.01 * eps / ntl)) atomicCAS(flag + 1, 0, 1);
    }
}

// R = R₂ R₁ (k×k, ld k, zero below the diagonal) from the two Cholesky
// factors (R_c = L_cᵀ, L_c lower, ld k; their strict upper parts are
// ignored): one workgroup per upper 16×16 tile, the two strips in LDS.
template <typename T>
__global__ __launch_bounds__(256) void hr_r_kernel(int k, const T *L1, const T *L2, T *R)
{
    const int bi = blockIdx.x, bj = blockIdx.y;
    if (bi > bj) return;
    __shared__ T A[128 * 17], B[128 * 17];   // A[l][ii] = L₂(l, i), B[l][jj] = L₁(j, l)
    const int t = threadIdx.x;
    for (int e = t; e < 128 * 16; e += 256) {
        const int l = e >> 4, c = e & 15;
        const int i = bi * 16 + c, j = bj * 16 + c;
        A[l * 17 + c] = (l < k && i < k && l >= i) ? L2[l + (int64_t)i * k] : T(0);
        B[l * 17 + c] = (l < k && j < k && l <= j) ? L1[j + (int64_t)l * k] : T(0);
    }
    __syncthreads();
    const int ii = t & 15, jj = t >> 4;
    const int i = bi * 16 + ii, j = bj * 16 + jj;
    T acc = T(0);
    for (int l = bi * 16; l < min(k, bj * 16 + 16); ++l) acc += A[l * 17 + ii] * B[l * 17 + jj];
    if (i < k && j < k) R[i + (int64_t)j * k] = i <= j ? acc : T(0);
}

// v from lane `src` (wave-uniform) to every lane, through SGPRs.
template <typename T>
__device__ __forceinline__ T wave_bcast(T v, int src)
{
    if constexpr (sizeof(T) == 8) {
        int2 u = __builtin_bit_cast(int2, v);
        u.x = __builtin_amdgcn_readlane(u.x, src);
        u.y = __builtin_amdgcn_readlane(u.y, src);
        return __builtin_bit_cast(T, u);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
    }
}

constexpr int kHrLD = 129;   // LDS leading dimension of the top block
constexpr int kHrThreads = 256;
constexpr int kHrTL = 17;    // leading dimension of the 16×16 inverses in LDS

__device__ __forceinline__ void hr_wave_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// 1/u for the pivots of the top-block LU (|u| ∈ [1, 2]): the hardware
// reciprocal refined by two Newton steps, r ← r + r(1 − u r), which takes
// it to within an ulp of the quotient.
template <typename T>
__device__ __forceinline__ T recip1(T u)
{
    T r;
    if constexpr (sizeof(T) == 8) r = __builtin_amdgcn_rcp(u);
    else r = __builtin_amdgcn_rcpf(u);
    r = __builtin_fma(__builtin_fma(-u, r, T(1)), r, r);
    r = __builtin_fma(__builtin_fma(-u, r, T(1)), r, r);
    return r;
}

// Column C of the elimination of a 16×16 diagonal block held in row layout
// (lane r of each 16-lane row holds d[c] = M[jb + r][jb + c]): the pivot
// −q̂_CC comes from lane C by DPP broadcast, S_C = copysign(1, −q̂_CC) and
// u_C = −q̂_CC + S_C; rows r > C take l = d[C]/u_C and d[c] −= l·u_(C,c)
// for c > C, the pivot row's entries broadcast from lane C as well (rows at
// or above the pivot run the same FMAs with l = 0).  The steps are
// branch-free, so the 16 columns form one block the compiler can schedule
// across.
template <int C, int J, typename T>
struct HrLuRow {
    static __device__ __forceinline__ void run(T d[16], T l)
    {
        d[J] = fma_t(-l, row_bcast<C>(d[J]), d[J]);
        HrLuRow<C, J + 1, T>::run(d, l);
    }
};
template <int C, typename T>
struct HrLuRow<C, 16, T> {
    static __device__ __forceinline__ void run(T *, T) {}
};
template <int C, typename T>
struct HrLuCols {
    static __device__ __forceinline__ void run(T d[16], T dv[16], T uv[16], int r, int nbk)
    {
        if (C < nbk) {
            const T mjj = row_bcast<C>(d[C]);   // −q̂_CC
            const T dj = dcopysign<T>(T(1), mjj);
            const T uj = mjj + dj;              // |u_C| = 1 + |q̂_CC|
            const T ru = recip1(uj);
            // rows at or above the pivot take l = 0, which leaves them as
            // they are (fma(−0, u, d) = d)
            const bool below = r > C;
            const T l = d[C] * ru;
            HrLuRow<C, C + 1, T>::run(d, below ? l : T(0));
            d[C] = below ? l : (r == C ? uj : d[C]);
            dv[C] = dj;
            uv[C] = uj;
        }
        HrLuCols<C + 1, T>::run(d, dv, uv, r, nbk);
    }
};
template <typename T>
struct HrLuCols<16, T> {
    static __device__ __forceinline__ void run(T *, T *, T *, int, int) {}
};

// Multipliers of the backward elimination for Z = U⁻¹ (rows of U in
// registers, myr = 1/U[r][r] in lane r): m[K] = U[r][K]/U[K][K] for K > r,
// the reciprocal broadcast from lane K; zero otherwise.
template <int K, typename T>
struct HrUpMul {
    static __device__ __forceinline__ void run(const T d[16], T myr, int r, T m[16])
    {
        const T rk = row_bcast<K>(myr);
        m[K] = r < K ? d[K] * rk : T(0);
        HrUpMul<K + 1, T>::run(d, myr, r, m);
    }
};
template <typename T>
struct HrUpMul<16, T> {
    static __device__ __forceinline__ void run(const T *, T, int, T *) {}
};

// Steps (2)-(3) on the top block: one workgroup of 4 waves, k ≤ 128, the
// block in LDS (M, column-major), blocked by 16 columns as a right-looking
// LU without pivoting.  Per block J = jb..jb+15:
//   wave 0 eliminates the 16×16 diagonal block column by column (lane r
//   holds row r; the pivot row moves by DPP broadcast), choosing
//   S_jj = copysign(1, −q̂_jj) from the updated diagonal entry — which
//   depends on this block alone — and forms W_J = L_JJ⁻¹ and Z = U_JJ⁻¹
//   from the rows still in its registers; then, on MFMA over 16×16 tiles:
//   U_J,right = W_J M_J,right,
//   L_below,J = M_below,J Z, and M_below,right −= L_below,J U_J,right.
// Then T = Û V₁⁻ᵀ (Û = U S) in place over U, by column blocks:
//   T[:, J] = Acc[:, J] W_Jᵀ, Acc[:, right] −= T[:, J] V₁[right, J]ᵀ
// (T upper and V₁ strictly lower share M).  Writes V₁ (explicit: unit
// diagonal, zero above) over Q's top block, Un = −Uᵀ (lower, ld k) and the
// inverses of its 16×16 diagonal tiles (UnInv) for the solve of the rows
// below, T (zero below the diagonal), S (Sv) and τ.
template <typename T>
__global__ __launch_bounds__(kHrThreads) void hr_lu_kernel(int k, T *Q, int64_t ldq,
                                                           const int *flag, T *Un, T *UnInv,
                                                           T *Tw, int64_t ldt, T *Sv, T *tau)
{
    typedef PanelMfma<T> Mf;
    typedef typename Mf::acc_t acc_t;
    __shared__ T M[128 * kHrLD];
    __shared__ T Ws[8][16 * kHrTL];
    __shared__ T Zs[16 * kHrTL];
    __shared__ T Dv[128], Uv[128];
    if (*flag != 0) return;
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int li = lane & 15, lk = lane >> 4;
    {
#pragma unroll 1
      for (int q0 = 0; q0 < 32; q0 += 16) {
        // unconditional loads from clamped indices (the top k×k block always
        // exists), so that all 32 stay in flight; out-of-range values are
        // replaced afterwards
        T v[16][2];
#pragma unroll
        for (int qq = 0; qq < 16; ++qq)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int i = lane + 64 * s, l = w + 4 * (q0 + qq);
                const int ic = i < k ? i : k - 1, lcl = l < k ? l : k - 1;
                const T q = Q[ic + (int64_t)lcl * ldq];
                v[qq][s] = (i < k && l < k) ? -q : T(0);
            }
#pragma unroll
        for (int qq = 0; qq < 16; ++qq)
#pragma unroll
            for (int s = 0; s < 2; ++s) M[lane + 64 * s + (w + 4 * (q0 + qq)) * kHrLD] = v[qq][s];
      }
    }
    __syncthreads();

    // one 16×16 tile product on MFMA: acc += A B, A(i, kk) = a(i, kk),
    // B(kk, j) = b(kk, j); element r of lane l of the result sits at
    // (Mf::row(l, r), l % 16)
    // The diagonal block of column block jb: wave 0 eliminates it, forms the
    // inverses W = L⁻¹, Z = U⁻¹ and the tile inverse of Un.
    auto elim = [&](int jb) {
        const int nbk = min(16, k - jb);
        T *Wb = Ws[jb >> 4];
        // lanes 0..15 (row r = lane)
        const int r = li;
        T d[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) d[c] = M[jb + r + (jb + c) * kHrLD];
        // scheduling fences between the chained stages of this wave's
        // work keep their loads from being hoisted into the elimination
        __builtin_amdgcn_sched_barrier(0);
        T dv[16], uv[16];
        HrLuCols<0, T>::run(d, dv, uv, r, nbk);
        __builtin_amdgcn_sched_barrier(0);
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c < nbk) {
                    Dv[jb + c] = dv[c];
                    Uv[jb + c] = uv[c];
                }
        }
        if (lane < 16 && jb + r < 128) {
#pragma unroll
            for (int c = 0; c < 16; ++c) M[jb + r + (jb + c) * kHrLD] = d[c];
        }
        // W = L⁻¹ (unit lower) and Z = U⁻¹ (upper) from the rows in
        // registers, four columns per lane (wz[t] = X[r][g + 4t]); rows and
        // columns past nbk (a partial last block) are the identity
        const int g = lane >> 4;
        T m[16], wv[4], zv[4];
#pragma unroll
        for (int c = 0; c < 16; ++c) m[c] = r > c ? d[c] : T(0);
#pragma unroll
        for (int t = 0; t < 4; ++t) wv[t] = zv[t] = (r == g + 4 * t) ? T(1) : T(0);
        TileInvSteps<0, T>::run(m, wv);
        T udiag = d[0];
#pragma unroll
        for (int c = 1; c < 16; ++c) udiag = r == c ? d[c] : udiag;
        const T myr = r < nbk ? recip1(udiag) : T(1);
        // unscaled rows Y[i] = U_ii Z[i]: Y[i] −= (U[i][k]/U[k][k]) Y[k]
        HrUpMul<0, T>::run(d, myr, r, m);
        TileInvStepsUp<15, T>::run(m, zv);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int j = g + 4 * t;
            zv[t] *= myr;
            Wb[r + j * kHrTL] = wv[t];
            Zs[r + j * kHrTL] = zv[t];
            // the 16×16 diagonal-tile inverse of Un = −Uᵀ, row-major as
            // chol_solve_kernel reads it: (−U_JJᵀ)⁻¹ = −Zᵀ (the padding of
            // a partial last block is the identity)
            UnInv[(jb >> 4) * 256 + j * 16 + r] =
                (j >= nbk || r >= nbk) ? (j == r ? T(1) : T(0)) : -zv[t];
        }
    };
    // One 16×16 tile of the trailing update M_below,right −= L_below,J U_J,right.
    auto trail_tile = [&](int jb, int c0, int nct, int it) {
        const int io = c0 + 16 * (it % nct), jo = c0 + 16 * (it / nct);
        acc_t acc;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = M[io + Mf::row(lane, q) + (jo + li) * kHrLD];
#pragma unroll
        for (int kq = 0; kq < 4; ++kq) {
            const int kk = 4 * kq + lk;
            acc = Mf::mma(-M[io + li + (jb + kk) * kHrLD], M[jb + kk + (jo + li) * kHrLD], acc);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) M[io + Mf::row(lane, q) + (jo + li) * kHrLD] = acc[q];
    };
    if (w == 0) elim(0);
    __syncthreads();
    for (int jb = 0; jb < k; jb += 16) {
        const int nbk = min(16, k - jb), c0 = jb + nbk;
        T *Wb = Ws[jb >> 4];
        if (c0 >= k) break;
        const int nct = (k - c0 + 15) / 16;
        // ---- U_J,right = W M_J,right and L_below,J = M_below,J Z
        for (int it = w; it < 2 * nct; it += kHrThreads / 64) {
            const int o = c0 + 16 * (it >> 1);
            acc_t acc = {T(0), T(0), T(0), T(0)};
            if ((it & 1) == 0) {
#pragma unroll
                for (int kq = 0; kq < 4; ++kq) {
                    const int kk = 4 * kq + lk;
                    acc = Mf::mma(Wb[li + kk * kHrTL], M[jb + kk + (o + li) * kHrLD], acc);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) M[jb + Mf::row(lane, q) + (o + li) * kHrLD] = acc[q];
            } else {
#pragma unroll
                for (int kq = 0; kq < 4; ++kq) {
                    const int kk = 4 * kq + lk;
                    acc = Mf::mma(M[o + li + (jb + kk) * kHrLD], Zs[kk + li * kHrTL], acc);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) M[o + Mf::row(lane, q) + (jb + li) * kHrLD] = acc[q];
            }
        }
        __syncthreads();
        // ---- M_below,right −= L_below,J U_J,right: wave 0 updates the next
        //      diagonal tile and eliminates it while waves 1–3 update the rest
        if (w == 0) {
            trail_tile(jb, c0, nct, 0);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            elim(c0);
        } else {
            for (int it = w; it < nct * nct; it += kHrThreads / 64 - 1) trail_tile(jb, c0, nct, it);
        }
        __syncthreads();
    }

    // ---- outputs of the LU: V₁ over Q's top block, Û = U S in place
    {
#pragma unroll 1
      for (int q0 = 0; q0 < 32; q0 += 8) {
        T v[8][2];
#pragma unroll
        for (int qq = 0; qq < 8; ++qq)
#pragma unroll
            for (int s = 0; s < 2; ++s) v[qq][s] = M[lane + 64 * s + (w + 4 * (q0 + qq)) * kHrLD];
#pragma unroll
        for (int qq = 0; qq < 8; ++qq)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int i = lane + 64 * s, l = w + 4 * (q0 + qq);
                if (i < k && l < k) {
                    Q[i + (int64_t)l * ldq] = i > l ? v[qq][s] : (i == l ? T(1) : T(0));
                    if (i <= l) M[i + l * kHrLD] = v[qq][s] * Dv[l];
                }
            }
      }
    }
    if (t < k) {
        Sv[t] = Dv[t];
        tau[t] = Dv[t] * Uv[t];
    }
    __syncthreads();
    // Un(l, i) = −U(i, l) = −Û(i, l) S_l, written down the columns of Un
    {
#pragma unroll 1
      for (int q0 = 0; q0 < 32; q0 += 8) {
        T v[8][2];
#pragma unroll
        for (int qq = 0; qq < 8; ++qq)
#pragma unroll
            for (int s = 0; s < 2; ++s) v[qq][s] = M[w + 4 * (q0 + qq) + (lane + 64 * s) * kHrLD];
#pragma unroll
        for (int qq = 0; qq < 8; ++qq)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int i = w + 4 * (q0 + qq), l = lane + 64 * s;
                if (i <= l && l < k) Un[l + (int64_t)i * k] = -v[qq][s] * Dv[l];
            }
      }
    }
    __syncthreads();   // the T sweep below rewrites M

    // ---- T = Û V₁⁻ᵀ
    for (int jb = 0; jb < k; jb += 16) {
        const int c0 = min(k, jb + 16);
        const T *Wb = Ws[jb >> 4];
        // T[0:c0, J] = Acc[0:c0, J] W_Jᵀ (Acc(i, c) for i > c is V₁: zero)
        for (int it = w; it <= (jb >> 4); it += kHrThreads / 64) {
            const int io = 16 * it;
            acc_t acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int kq = 0; kq < 4; ++kq) {
                const int kk = 4 * kq + lk;
                const T a = (io + li <= jb + kk) ? M[io + li + (jb + kk) * kHrLD] : T(0);
                acc = Mf::mma(a, Wb[li + kk * kHrTL], acc);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = io + Mf::row(lane, q), c = jb + li;
                if (i <= c) M[i + c * kHrLD] = acc[q];
            }
        }
        __syncthreads();
        if (c0 >= k) break;
        const int nct = (k - c0 + 15) / 16, nrt = (jb >> 4) + 1;
        // Acc[0:c0, right] −= T[0:c0, J] V₁[right, J]ᵀ
        for (int it = w; it < nrt * nct; it += kHrThreads / 64) {
            const int io = 16 * (it % nrt), jo = c0 + 16 * (it / nrt);
            acc_t acc;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = M[io + Mf::row(lane, q) + (jo + li) * kHrLD];
#pragma unroll
            for (int kq = 0; kq < 4; ++kq) {
                const int kk = 4 * kq + lk;
                const T a = (io + li <= jb + kk) ? -M[io + li + (jb + kk) * kHrLD] : T(0);
                acc = Mf::mma(a, M[jo + li + (jb + kk) * kHrLD], acc);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) M[io + Mf::row(lane, q) + (jo + li) * kHrLD] = acc[q];
        }
        __syncthreads();
    }
    {
#pragma unroll 1
      for (int q0 = 0; q0 < 32; q0 += 8) {
        T v[8][2];
#pragma unroll
        for (int qq = 0; qq < 8; ++qq)
#pragma unroll
            for (int s = 0; s < 2; ++s) v[qq][s] = M[lane + 64 * s + (w + 4 * (q0 + qq)) * kHrLD];
#pragma unroll
        for (int qq = 0; qq < 8; ++qq)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int i = lane + 64 * s, l = w + 4 * (q0 + qq);
                if (i < k && l < k) Tw[i + (int64_t)l * ldt] = i <= l ? v[qq][s] : T(0);
            }
      }
    }
}

// The reconstructed factors over the panel, if the path succeeded: rows
// ≥ k take V (from Q's storage, where the solve left it), the top block V₁
// below the diagonal and R_h = S R on and above it.
template <typename T>
__global__ void hr_finish_kernel(int64_t m, int k, const T *Vw, int64_t ldv, const T *R,
                                 const T *Sv, T *A, int64_t lda, const int *flag)
{
    if (*flag != 0) return;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= m) return;
    A[i + (int64_t)j * lda] =
        i > j ? Vw[i + (int64_t)j * ldv] : Sv[i] * R[i + (int64_t)j * k];
}

// Workspace layout of one qrBlocked! call.
template <typename T>
struct QrWs {
    T *Vw, *Tw, *W1, *W2, *part, *split;
    size_t split_elems;
};

// Scratch of the reconstructed-Householder panel (k ≤ nb ≤ 128): the two
// Gram matrices / Cholesky factors, −Uᵀ, R, the 16×16 tile inverses, S,
// and the Gram partials of hr_gram_blocks(m) workgroups.
inline size_t hr_sq(int k) { return (size_t)cdiv((int64_t)k * k, 64) * 64 + 64; }
inline int hr_gram_blocks(int64_t m) { return (int)std::min<int64_t>(kGrBlocks, cdiv(m, kGrRB)); }
inline size_t hr_scratch_elems(int64_t m, int nb)
{
    return 4 * hr_sq(nb) + tile_inverse_elems(kQrNB) + 64 + 128 +
           (size_t)hr_gram_blocks(m) * hr_ntiles(std::min(nb, kQrNB)) * 256;
}

template <typename T>
size_t qr_ws_elems(int64_t m, int64_t n, int nb, int64_t *split_elems)
{
    // partial sums: 2 parities × (P + 1) × 2·NB covers both panel kernels
    // (workgroups of at least 64 rows)
    const int64_t P = std::max<int64_t>(1, cdiv(m, 64));
    const int64_t sp = 8 * (int64_t)nb * std::max<int64_t>(n, 1);
    *split_elems = sp;
    // two panel sets (V, T) and two update sets (W1, W2, split-K partials):
    // the look-ahead driver updates on the side stream with one set while
    // the caller's stream factors the next panel with the other
    return 2 * ((size_t)m * nb + (size_t)nb * nb + 2 * (size_t)nb * std::max<int64_t>(n, 1) +
                (size_t)sp) +
           2 * (size_t)(P + 1) * 2 * kQrNB + 64 + hr_scratch_elems(m, nb) + 128;
}

template <typename T>
int qr_panel(Handle *h, hipStream_t s, int64_t m, int w, T *A, int64_t lda, T *tau, T *Vw,
             int64_t ldv, T *Tw, int64_t ldt, T *part, unsigned &bar_base,
             const int *run_if = nullptr)
{
    // 1: LDS-reduction kernel; 2: wave-butterfly kernel, sized to the panel
    // width: eight columns per wave, so narrower panels run fewer waves per
    // workgroup with more rows per lane (fewer workgroups to synchronise)
    const int panel = h->qr_panel;
    const int wide = w <= 32 ? 32 : (w <= 64 ? 64 : 128);
    const int R = panel == 1 ? (kQrThreads / kQrNB) * kQrRPT
                             : (wide == 32 ? 512 : (wide == 64 ? 256 : 64 * kQrRPL));
    const int64_t P = std::max<int64_t>(1, cdiv(m, R));
    // every workgroup must be resident at once for the grid barrier
    if (P > h->num_cus) {
        set_error("qr panel: a block column of %lld rows exceeds the register-resident panel "
                  "(at most %lld rows at width %d on %d CUs)",
                  (long long)m, (long long)R * h->num_cus, wide, h->num_cus);
        return GLA_ERR_UNSUPPORTED;
    }
    QrPanelP<T> p;
    p.m = m;
    p.w = w;
    p.kmax = (int)std::min<int64_t>(m, w);
    p.A = A;
    p.lda = lda;
    p.tau = tau;
    p.Vw = Vw;
    p.ldv = ldv;
    p.Tw = Tw;
    p.ldt = ldt;
    p.part = part;
    p.bar = h->dbar;
    p.bar_base = bar_base;
    p.spin_limit = h->spin_limit;
    p.bar_skew = h->qr_wait_test ? (1u << 30) : 0u;
    p.err = h->dinfo + 1;
    p.run_if = run_if;
    if (panel == 1)
        hipLaunchKernelGGL((qr_panel_kernel<T, kQrNB, kQrRPT>), dim3((unsigned)P),
                           dim3(kQrThreads), 0, s, p);
    else if (wide == 32)
        hipLaunchKernelGGL((qr_panel2_kernel<T, 32, 8, 256>), dim3((unsigned)P), dim3(256), 0, s,
                           p);
    else if (wide == 64)
        hipLaunchKernelGGL((qr_panel2_kernel<T, 64, 4, 512>), dim3((unsigned)P), dim3(512), 0, s,
                           p);
    else
        hipLaunchKernelGGL((qr_panel2_kernel<T, kQrNB, kQrRPL, kQrThreads>), dim3((unsigned)P),
                           dim3(kQrThreads), 0, s, p);
    GLA_LAUNCH_CHECK();
    if (P > 1) bar_base += (unsigned)(p.kmax * P);
    return GLA_OK;
}

// Start of a factorisation: the grid-barrier counter and the panel's
// wait-timeout flag to zero.
__global__ void qr_reset_kernel(unsigned *bar, int *err)
{
    *bar = 0u;
    *err = 0;
}

// qrBlocked!'s workspace after a step's lmul!(H', A2, M) (src/qr.jl:132,
// src/householder.jl:142-153): M = −V₁ W2 with V₁ the unit-lower top k×k
// block of the explicit V and W2 = Tᵀ VᵀA2 (k×n2, ld k); written to
// work[0:k, c0:c0+n2].  One thread per element.
template <typename T>
__global__ void qr_work_kernel(int k, int64_t n2, const T *Vw, int64_t ldv, const T *W2, T *work,
                               int64_t ldw)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (r >= k || c >= n2) return;
    const T *w = W2 + c * k;
    T acc = w[r];
    for (int i = 0; i < r; ++i) acc += Vw[r + (int64_t)i * ldv] * w[i];
    work[r + c * ldw] = -acc;
}

template <typename T>
int qr_work(hipStream_t s, int k, int64_t n2, const T *Vw, int64_t ldv, const T *W2, T *work,
            int64_t ldw)
{
    if (!work || k <= 0 || n2 <= 0) return GLA_OK;
    for (int64_t c0 = 0; c0 < n2; c0 += 65535) {
        const int64_t nc = std::min<int64_t>(65535, n2 - c0);
        hipLaunchKernelGGL(qr_work_kernel<T>, dim3((unsigned)cdiv(k, 128), (unsigned)nc),
                           dim3(128), 0, s, k, nc, Vw, ldv, W2 + c0 * k, work + c0 * ldw, ldw);
        GLA_LAUNCH_CHECK();
    }
    return GLA_OK;
}

// A2 (mm×n2) <- (I - V T Vᵀ)ᵀ A2 with explicit V (mm×k) and T (k×k upper);
// trans 'T' applies Qᵀ (src/householder.jl:119-157), 'N' applies Q (:82-115).
template <typename T>
int wy_update(hipStream_t s, char trans, int64_t mm, int64_t n2, int64_t k, const T *Vw,
              int64_t ldv, const T *Tw, int64_t ldt, T *A2, int64_t lda, T *W1, T *W2, T *split,
              size_t split_elems)
{
    if (n2 <= 0 || k <= 0 || mm <= 0) return GLA_OK;
    // :142-144  W1 = V1ᵀA1 + V2ᵀA2 = Vᵀ A2
    GLA_TRY(gemm<T>(s, 'T', 'N', 0, k, n2, mm, T(1), Vw, ldv, A2, lda, T(0), W1, k, 0, split,
                    split_elems));
    // :146  W2 = Tᵀ W1  (or T W1 for Q)
    GLA_TRY(gemm<T>(s, trans == 'T' ? 'T' : 'N', 'N', 0, k, n2, k, T(1), Tw, ldt, W1, k, T(0), W2,
                    k));
    // :149-154  A2 -= V W2
    GLA_TRY(gemm<T>(s, 'N', 'N', 0, mm, n2, k, T(-1), Vw, ldv, W2, k, T(1), A2, lda));
    return GLA_OK;
}

// Rows of the register-resident panel kernel's workgroups for width w
// (the fallback of the reconstructed-Householder path).
inline int panel_rows(const Handle *h, int w)
{
    if (h->qr_panel == 1) return (kQrThreads / kQrNB) * kQrRPT;
    return w <= 32 ? 512 : (w <= 64 ? 256 : 64 * kQrRPL);
}

// Whether block column mm×w is factored by the reconstructed-Householder
// path: at least as tall as wide, qr_hr rows or taller (auto: any height
// for w ≥ 64, where it was faster from 128×128 up; 2048 rows for narrower
// block columns), and within the fallback panel's capacity.
inline bool use_hr(const Handle *h, int64_t mm, int w)
{
    const int64_t minrows = h->qr_hr >= 0 ? h->qr_hr : (w >= 64 ? 1 : 2048);
    if (h->qr_hr == 0 || mm < w || mm < minrows || w > kQrNB) return false;
    return cdiv(mm, panel_rows(h, w)) <= h->num_cus;
}

// The reconstructed-Householder panel (see hr_gram_kernel … hr_finish_kernel):
// factors Ab (mm×w, mm ≥ w) in place into R and V, writes τ, the explicit
// V (Vw, ld mm) and T (Tw, ld w).  Stream-ordered launches on s, no host
// synchronisation; the column-by-column panel kernel runs last, gated on
// the device flag, when the path's checks fail.
template <typename T>
int hr_gram(hipStream_t s, int64_t m, int k, const T *X, int64_t ldx, T *G, T *part, int *flag,
            int check, T *copy_to = nullptr, int64_t ldc = 0)
{
    const int P = hr_gram_blocks(m);
    hipLaunchKernelGGL(hr_gram_kernel<T>, dim3((unsigned)P), dim3(256), 0, s, m, k, X, ldx, part,
                       copy_to, ldc, flag);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(hr_gram_reduce_kernel<T>, dim3((unsigned)hr_ntiles(k)), dim3(1024), 0, s, k,
                       P, part, G, flag, check);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template <typename T>
int qr_panel_hr(Handle *h, hipStream_t s, int64_t mm, int w, T *Ab, int64_t lda, T *tau, T *Vw,
                T *Tw, T *hr, T *part, unsigned &bar_base)
{
    const int k = w;
    int *flag = h->dinfo + 16;
    const size_t sq = hr_sq(k);
    T *G1 = hr, *G2 = G1 + sq, *Un = G2 + sq, *Rb = Un + sq, *Linv = Rb + sq;
    T *Sv = Linv + tile_inverse_elems(kQrNB) + 64;
    T *gpart = Sv + 128;
    const dim3 cgrid((unsigned)cdiv(mm, 256), (unsigned)k);
    // (1) two CholeskyQR passes: Vw <- P R₁⁻¹ R₂⁻¹ (the first Gram pass also
    // copies P to Vw and clears the flags)
    GLA_TRY(hr_gram<T>(s, mm, k, Ab, lda, G1, gpart, flag, 0, Vw, mm));
    GLA_TRY(chol_factor_block<T>(s, k, G1, k, flag, 0, Linv));
    GLA_TRY(trsm_rlt<T>(s, mm, k, G1, k, Linv, Vw, mm));
    GLA_TRY(hr_gram<T>(s, mm, k, Vw, mm, G2, gpart, flag, 1));
    // G₂ = I + E: the first-order factor when E is below √eps (the usual
    // case), the full factorisation otherwise (flag[1], from the reduction)
    GLA_TRY(chol_factor_block_near_identity<T>(s, k, G2, k, flag, 0, Linv, flag));
    GLA_TRY(trsm_rlt<T>(s, mm, k, G2, k, Linv, Vw, mm));
    const unsigned nt = (unsigned)cdiv(k, 16);
    hipLaunchKernelGGL(hr_r_kernel<T>, dim3(nt, nt), dim3(256), 0, s, k, G1, G2, Rb);
    GLA_LAUNCH_CHECK();
    // (2)-(3) the signed LU of the top block, T, τ; V₂ = Q₂ (−U)⁻¹
    hipLaunchKernelGGL(hr_lu_kernel<T>, dim3(1), dim3(kHrThreads), 0, s, k, Vw, mm, flag, Un,
                       Linv, Tw, (int64_t)k, Sv, tau);
    GLA_LAUNCH_CHECK();
    GLA_TRY(trsm_rlt<T>(s, mm - k, k, Un, k, Linv, Vw + k, mm));
    hipLaunchKernelGGL(hr_finish_kernel<T>, cgrid, dim3(256), 0, s, mm, k, Vw, mm, Rb, Sv, Ab,
                       lda, flag);
    GLA_LAUNCH_CHECK();
    // fallback: the column-by-column panel over the untouched panel
    return qr_panel<T>(h, s, mm, w, Ab, lda, tau, Vw, mm, Tw, k, part, bar_base, flag);
}

// Factor one block column (mm×w at Ab) into V (explicit, Vw) and T (Tw):
// one panel launch, or sub-panels of width h->qr_inner with compact-WY
// updates between them and the block's T assembled from theirs and G = VᵀV.
template <typename T>
int qr_block_column(Handle *h, hipStream_t s, int64_t mm, int w, T *Ab, int64_t lda, T *tau,
                    T *Vw, T *Tw, T *W1, T *W2, T *split, size_t split_elems, T *part,
                    unsigned &bar_base, T *hr)
{
    const int64_t k = std::min<int64_t>(mm, w);
    if (use_hr(h, mm, w))
        return qr_panel_hr<T>(h, s, mm, w, Ab, lda, tau, Vw, Tw, hr, part, bar_base);
    // auto (-1): sub-panels of 64 for block columns taller than 12288 rows,
    // where the full-width panel spreads over more workgroups than the
    // 64-wide one and its per-column exchange is slower
    int ib = h->qr_inner >= 0 ? h->qr_inner : (mm > 12288 ? 64 : 0);
    // every workgroup of a panel launch must be resident at once (the grid
    // exchange waits for all of them): block columns too tall for their
    // width are factored as narrower sub-panels (64, then 32 wide)
    if (h->qr_panel == 2) {
        int cur = (ib > 0 && k > ib) ? ib : w;
        while (cur > 32 && cdiv(mm, panel_rows(h, cur)) > h->num_cus) cur = cur > 64 ? 64 : 32;
        if (cur < k) ib = cur;
    }
    if (!(h->qr_panel == 2 && ib > 0 && k > ib))
        return qr_panel<T>(h, s, mm, w, Ab, lda, tau, Vw, mm, Tw, k, part, bar_base);
    hipLaunchKernelGGL(qr_sub_clear_kernel<T>, dim3(1, (unsigned)k), dim3(128), 0, s, (int)k, Vw,
                       mm, Tw, k);
    GLA_LAUNCH_CHECK();
    for (int64_t s0 = 0; s0 < k; s0 += ib) {
        const int iw = (int)std::min<int64_t>(ib, w - s0);
        const int64_t ms = mm - s0;
        const int64_t ks = std::min<int64_t>(ms, iw);
        T *Vs = Vw + s0 + s0 * mm;
        T *Tss = Tw + s0 + s0 * k;
        GLA_TRY(qr_panel<T>(h, s, ms, iw, Ab + s0 + s0 * lda, lda, tau + s0, Vs, mm, Tss, k, part,
                            bar_base));
        if (w - s0 - iw > 0)
            GLA_TRY(wy_update<T>(s, 'T', ms, w - s0 - iw, ks, Vs, mm, Tss, k,
                                 Ab + s0 + (s0 + iw) * lda, lda, W1, W2, split, split_elems));
    }
    T *G = W1;   // k×k, k ≤ nb
    GLA_TRY(gemm<T>(s, 'T', 'N', 0, k, k, mm, T(1), Vw, mm, Vw, mm, T(0), G, k, 0, split,
                    split_elems));
    hipLaunchKernelGGL(qr_t_merge_kernel<T>, dim3(1), dim3(1024), 0, s, (int)k, ib, Tw, k, G, k);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// qrBlocked! (src/qr.jl:113-146).  With h->qr_lookahead, depth-1
// look-ahead: after panel k the caller's stream applies its reflectors to
// block column k+1 only and factors panel k+1 at once, while the side
// stream applies them to the rest of the trailing matrix; panels alternate
// between two (V, T) sets and the two streams use separate update buffers.
template <typename T>
int qr_blocked_dev(Handle *h, hipStream_t s, int64_t m, int64_t n, T *A, int64_t lda, int nb,
                   T *tau, T *work, int64_t ldw)
{
    if (m <= 0 || n <= 0) return GLA_OK;
    int64_t split_elems = 0;
    const size_t need = qr_ws_elems<T>(m, n, nb, &split_elems);
    GLA_TRY(ensure_workspace(h, need * sizeof(T)));
    const size_t sp = (size_t)split_elems, nw = (size_t)nb * std::max<int64_t>(n, 1);
    T *p = reinterpret_cast<T *>(h->ws);
    T *Vw[2], *Tw[2], *W1[2], *W2[2], *split[2];
    for (int q = 0; q < 2; ++q) {
        Vw[q] = p;
        Tw[q] = Vw[q] + (size_t)m * nb;
        W1[q] = Tw[q] + (size_t)nb * nb;
        W2[q] = W1[q] + nw;
        split[q] = W2[q] + nw;
        p = split[q] + sp;
    }
    T *part = p;
    const int64_t P = std::max<int64_t>(1, cdiv(m, 64));
    // 64-element aligned
    const size_t hoff = (size_t)(part - reinterpret_cast<T *>(h->ws)) +
                        2 * (size_t)(P + 1) * 2 * kQrNB + 64;
    T *hr = reinterpret_cast<T *>(h->ws) + (hoff + 63) / 64 * 64;
    const int64_t npanels = cdiv(std::min(m, n), (int64_t)nb);
    const bool la = h->qr_lookahead && npanels > 2;

    hipLaunchKernelGGL(qr_reset_kernel, dim3(1), dim3(1), 0, s, h->dbar, h->dinfo + 1);
    GLA_LAUNCH_CHECK();
    // With CUs reserved from the side stream (side_cu_reserve), the panel
    // chain runs on the handle's chain stream with every CU and the trailing
    // updates on the masked one, both forked from (and joined back to) the
    // caller's stream: the one-workgroup steps of the panel then find CUs
    // beside the trailing products.
    const bool masked = la && h->masked != nullptr;
    hipStream_t s0 = s;
    hipStream_t side = masked ? h->masked : h->side;
    if (la) {
        GLA_HIP_TRY(hipEventRecord(h->ev_fork, s0));
        GLA_HIP_TRY(hipStreamWaitEvent(side, h->ev_fork, 0));
        if (masked) {
            GLA_HIP_TRY(hipStreamWaitEvent(h->chain, h->ev_fork, 0));
            s = h->chain;
        }
    }
    unsigned bar_base = 0;
    int64_t step = 0;
    for (int64_t j0 = 0;; j0 += nb, ++step) {
        const int64_t mm = m - j0, nn = n - j0;
        const int w = (int)std::min<int64_t>(nn, nb);                     // src/qr.jl:121
        const int64_t k = std::min<int64_t>(mm, w);
        T *Ab = A + j0 + j0 * lda;
        const int q = (int)(step & 1);
        GLA_TRY(qr_block_column<T>(h, s, mm, w, Ab, lda, tau + j0, Vw[q], Tw[q], W1[0], W2[0],
                                   split[0], sp, part, bar_base, hr));    // :123-125
        if (nn > nb) {                                                    // :126-133
            if (!la) {
                GLA_TRY(wy_update<T>(s, 'T', mm, nn - w, k, Vw[q], mm, Tw[q], k, Ab + w * lda, lda,
                                     W1[0], W2[0], split[0], sp));
                GLA_TRY(qr_work<T>(s, (int)k, nn - w, Vw[q], mm, W2[0], work, ldw));
            } else {
                // block column k+1 here; it was last touched by the side
                // stream's update of the previous step
                const int64_t w2 = std::min<int64_t>(nb, nn - w);
                const int64_t rest = nn - w - w2;
                if (step > 0) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
                GLA_TRY(wy_update<T>(s, 'T', mm, w2, k, Vw[q], mm, Tw[q], k, Ab + w * lda, lda,
                                     W1[0], W2[0], split[0], sp));
                GLA_TRY(qr_work<T>(s, (int)k, w2, Vw[q], mm, W2[0], work, ldw));
                if (rest > 0) {
                    GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
                    GLA_HIP_TRY(hipStreamWaitEvent(side, h->ev_step[0], 0));
                    GLA_TRY(wy_update<T>(side, 'T', mm, rest, k, Vw[q], mm, Tw[q], k,
                                         Ab + (w + w2) * lda, lda, W1[1], W2[1], split[1], sp));
                    // the same stream orders this step's columns after the
                    // previous step's; the chain writes columns [0, w2)
                    // only, disjoint from these
                    GLA_TRY(qr_work<T>(side, (int)k, rest, Vw[q], mm, W2[1],
                                       work ? work + w2 * ldw : nullptr, ldw));
                    GLA_HIP_TRY(hipEventRecord(h->ev_step[1], side));
                }
            }
        }
        if (!(mm > nb && nn > nb)) break;                                 // :136
    }
    if (la) {
        GLA_HIP_TRY(hipEventRecord(h->ev_join, side));
        GLA_HIP_TRY(hipStreamWaitEvent(s0, h->ev_join, 0));
        if (masked) {
            GLA_HIP_TRY(hipEventRecord(h->ev_fork, s));
            GLA_HIP_TRY(hipStreamWaitEvent(s0, h->ev_fork, 0));
        }
    }
    return GLA_OK;
}

}  // namespace

// Building blocks shared with the adjoints: the compact-WY T of the first k
// reflectors stored in F (m×k), and A (mA×nA) <- op(Q) A for Q = I − V T Vᵀ
// with V the first b reflectors of F.
size_t qr_tblock_scratch(int64_t m, int64_t k) { return (size_t)m * k + 2 * (size_t)k * k + 64; }
size_t qr_apply_scratch(int64_t mA, int64_t nA, int64_t b)
{
    return (size_t)mA * b + (size_t)b * b + 2 * (size_t)b * nA + 8 * (size_t)b * nA + 64;
}

template <typename T>
int qr_tblock_dev(hipStream_t s, int64_t m, int64_t k, const T *F, int64_t ldf, const T *tau,
                  T *Tm, int64_t ldt, T *scratch)
{
    if (k <= 0) return GLA_OK;
    // Vw (m×k) explicit, G = VwᵀVw (k×k), inverse scratch k*k
    T *Vw = scratch;
    T *Gm = Vw + (size_t)m * k;
    T *iw = Gm + (size_t)k * k;
    GLA_TRY(unit_lower_copy<T>(s, m, k, F, ldf, Vw, m));
    GLA_TRY(gemm<T>(s, 'T', 'N', 0, k, k, m, T(1), Vw, m, Vw, m, T(0), Gm, k));   // :69-74
    hipLaunchKernelGGL(tb_init_kernel<T>, dim3((unsigned)cdiv(k, 256), (unsigned)k), dim3(256), 0,
                       s, k, Gm, k, tau, Tm, ldt);
    GLA_LAUNCH_CHECK();
    GLA_TRY(trtri_unit_upper<T>(s, k, Tm, ldt, iw));                                // :75
    hipLaunchKernelGGL(tb_scale_kernel<T>, dim3((unsigned)cdiv(k, 256), (unsigned)k), dim3(256),
                       0, s, k, tau, Tm, ldt);                                      // :76-81
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template <typename T>
int qr_apply_dev(hipStream_t s, char trans, int64_t mA, int64_t nA, int64_t b, const T *V,
                 int64_t ldv, const T *Tm, int64_t ldt, T *A, int64_t lda, T *scratch)
{
    if (b <= 0 || nA <= 0) return GLA_OK;
    T *Vw = scratch;
    T *Tw = Vw + (size_t)mA * b;
    T *W1 = Tw + (size_t)b * b;
    T *W2 = W1 + (size_t)b * nA;
    T *split = W2 + (size_t)b * nA;
    GLA_TRY(unit_lower_copy<T>(s, mA, b, V, ldv, Vw, mA));
    GLA_TRY(upper_copy<T>(s, b, Tm, ldt, Tw, b));
    return wy_update<T>(s, trans, mA, nA, b, Vw, mA, Tw, b, A, lda, W1, W2, split,
                        8 * (size_t)b * nA);
}

#define GLA_INST_QR(T)                                                                           \
    template int qr_tblock_dev<T>(hipStream_t, int64_t, int64_t, const T *, int64_t, const T *, \
                                  T *, int64_t, T *);                                            \
    template int qr_apply_dev<T>(hipStream_t, char, int64_t, int64_t, int64_t, const T *,        \
                                 int64_t, const T *, int64_t, T *, int64_t, T *);
GLA_INST_QR(double)
GLA_INST_QR(float)
#undef GLA_INST_QR

template <typename T>
int unit_upper_explicit(hipStream_t s, int64_t n, const T *src, int64_t lds, T *dst, int64_t ldd)
{
    if (n <= 0) return GLA_OK;
    hipLaunchKernelGGL(unit_upper_copy_kernel<T>, dim3((unsigned)cdiv(n, 256), (unsigned)n),
                       dim3(256), 0, s, n, src, lds, dst, ldd);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// ------------------------------------------------------------------ C ABI --
namespace {

template <typename T>
int qr_entry(gla_handle_t hh, int64_t m, int64_t n, T *A, int64_t lda, int64_t bs, T *tau,
             T *work, int64_t ldwork, void *stream, bool unblocked)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (lda < std::max<int64_t>(1, m)) { set_error("lda < max(1,m)"); return -5; }
    if (!unblocked && bs < 1) { set_error("blocksize < 1"); return -6; }
    if (!tau && std::min(m, n) > 0) { set_error("tau is null"); return -7; }
    const int nb = (int)std::min<int64_t>(unblocked ? kQrNB : bs, kQrNB);
    if (work && ldwork < std::max<int64_t>(1, std::min<int64_t>(m, nb))) {
        set_error("ldwork < min(m, blocksize)");
        return -9;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    if (m == 0 || n == 0) return GLA_OK;
    GLA_TRY(qr_blocked_dev<T>(h, s, m, n, A, lda, nb, tau, work, ldwork));
    // a panel whose grid wait gave up has produced meaningless factors
    int timed_out = 0;
    GLA_HIP_TRY(hipMemcpyAsync(&timed_out, h->dinfo + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    GLA_HIP_TRY(hipStreamSynchronize(s));
    if (timed_out) {
        set_error("qr panel: a grid-wide wait gave up after %u polls; A, tau and work do not "
                  "hold a valid factorisation",
                  h->spin_limit);
        return GLA_ERR_TIMEOUT;
    }
    return GLA_OK;
}

template <typename T>
int tblock_entry(gla_handle_t hh, int64_t m, int64_t n, const T *F, int64_t ldf, const T *tau,
                 T *Tm, int64_t ldt, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (ldf < std::max<int64_t>(1, m)) return -5;
    const int64_t k = std::min(m, n);
    if (ldt < std::max<int64_t>(1, k)) return -8;
    if (k == 0) return GLA_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_TRY(ensure_workspace(h, qr_tblock_scratch(m, k) * sizeof(T)));
    return qr_tblock_dev<T>(s, m, k, F, ldf, tau, Tm, ldt, reinterpret_cast<T *>(h->ws));
}

template <typename T>
int wy_entry(gla_handle_t hh, char trans, int64_t mA, int64_t nA, const T *V, int64_t ldv,
             int64_t nV, const T *Tm, int64_t ldt, T *A, int64_t lda, T *M, int64_t ldm,
             void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (trans != 'N' && trans != 'T' && trans != 'C') return -2;
    if (mA < 0) return -3;
    if (nA < 0) return -4;
    if (ldv < std::max<int64_t>(1, mA)) return -6;
    if (nV < 0) return -7;
    const int64_t b = std::min(mA, nV);
    if (ldt < std::max<int64_t>(1, b)) return -9;
    if (lda < std::max<int64_t>(1, mA)) return -11;
    if (M && ldm < std::max<int64_t>(1, b)) return -13;
    if (b == 0 || nA == 0) return GLA_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_TRY(ensure_workspace(h, qr_apply_scratch(mA, nA, b) * sizeof(T)));
    T *ws = reinterpret_cast<T *>(h->ws);
    GLA_TRY(qr_apply_dev<T>(s, trans == 'N' ? 'N' : 'T', mA, nA, b, V, ldv, Tm, ldt, A, lda, ws));
    // W2 = op(T) Vᵀ A, the reference's workspace M after the call
    if (M) GLA_TRY(copy_mat<T>(s, b, nA, ws + (size_t)mA * b + (size_t)b * b + (size_t)b * nA, b,
                               M, ldm));
    return GLA_OK;
}

template <typename T>
int lmul_ul_entry(gla_handle_t hh, int64_t m, int64_t n, const T *L, int64_t ldl, T *B,
                  int64_t ldb, T alpha, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { set_error("null handle"); return -1; }
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (ldl < std::max<int64_t>(1, m)) return -5;
    if (ldb < std::max<int64_t>(1, m)) return -7;
    if (m == 0 || n == 0) return GLA_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_HIP_TRY(hipSetDevice(h->device));
    const size_t need = (size_t)m * m + (size_t)m * n + 64;
    GLA_TRY(ensure_workspace(h, need * sizeof(T)));
    T *Lw = reinterpret_cast<T *>(h->ws);
    T *W = Lw + (size_t)m * m;
    GLA_TRY(unit_lower_copy<T>(s, m, m, L, ldl, Lw, m));
    GLA_TRY(gemm<T>(s, 'N', 'N', 0, m, n, m, alpha, Lw, m, B, ldb, T(0), W, m));
    return copy_mat<T>(s, m, n, W, m, B, ldb);
}

}  // namespace
}  // namespace gla

extern "C" {

int gla_qr_unblocked_d(gla_handle_t h, int64_t m, int64_t n, double *A, int64_t lda, double *tau,
                       void *stream)
{
    return gla::qr_entry<double>(h, m, n, A, lda, 0, tau, nullptr, 0, stream, true);
}
int gla_qr_unblocked_s(gla_handle_t h, int64_t m, int64_t n, float *A, int64_t lda, float *tau,
                       void *stream)
{
    return gla::qr_entry<float>(h, m, n, A, lda, 0, tau, nullptr, 0, stream, true);
}
int gla_qr_blocked_d(gla_handle_t h, int64_t m, int64_t n, double *A, int64_t lda,
                     int64_t blocksize, double *tau, double *work, int64_t ldwork, void *stream)
{
    return gla::qr_entry<double>(h, m, n, A, lda, blocksize, tau, work, ldwork, stream, false);
}
int gla_qr_blocked_s(gla_handle_t h, int64_t m, int64_t n, float *A, int64_t lda,
                     int64_t blocksize, float *tau, float *work, int64_t ldwork, void *stream)
{
    return gla::qr_entry<float>(h, m, n, A, lda, blocksize, tau, work, ldwork, stream, false);
}
int gla_qr_tblock_d(gla_handle_t h, int64_t m, int64_t n, const double *F, int64_t ldf,
                    const double *tau, double *T, int64_t ldt, void *stream)
{
    return gla::tblock_entry<double>(h, m, n, F, ldf, tau, T, ldt, stream);
}
int gla_qr_tblock_s(gla_handle_t h, int64_t m, int64_t n, const float *F, int64_t ldf,
                    const float *tau, float *T, int64_t ldt, void *stream)
{
    return gla::tblock_entry<float>(h, m, n, F, ldf, tau, T, ldt, stream);
}
int gla_wy_apply_d(gla_handle_t h, char trans, int64_t mA, int64_t nA, const double *V,
                   int64_t ldv, int64_t nV, const double *T, int64_t ldt, double *A, int64_t lda,
                   double *M, int64_t ldm, void *stream)
{
    return gla::wy_entry<double>(h, trans, mA, nA, V, ldv, nV, T, ldt, A, lda, M, ldm, stream);
}
int gla_wy_apply_s(gla_handle_t h, char trans, int64_t mA, int64_t nA, const float *V,
                   int64_t ldv, int64_t nV, const float *T, int64_t ldt, float *A, int64_t lda,
                   float *M, int64_t ldm, void *stream)
{
    return gla::wy_entry<float>(h, trans, mA, nA, V, ldv, nV, T, ldt, A, lda, M, ldm, stream);
}
int gla_lmul_unit_lower_d(gla_handle_t h, int64_t m, int64_t n, const double *L, int64_t ldl,
                          double *B, int64_t ldb, double alpha, void *stream)
{
    return gla::lmul_ul_entry<double>(h, m, n, L, ldl, B, ldb, alpha, stream);
}
int gla_lmul_unit_lower_s(gla_handle_t h, int64_t m, int64_t n, const float *L, int64_t ldl,
                          float *B, int64_t ldb, float alpha, void *stream)
{
    return gla::lmul_ul_entry<float>(h, m, n, L, ldl, B, ldb, alpha, stream);
}

}  // extern "C"
