// Batched Cholesky factor + solve for many SPD matrices of uneven size
// (n ≤ 256), FP64 and FP32, one right-hand side each, in one or two
// persistent launches.  No counterpart in the reference (which factors one matrix per
// call, src/cholesky.jl); each matrix gets cholUnblocked!'s factor (to
// rounding), then the solve L Lᵀ x = b.
//
// Plan (host): matrices are sorted by n, largest first.  Those above a tuned
// threshold `wave_max` go to the workgroup kernel, one matrix per workgroup
// (matrix i → workgroup i % grid); the rest go to the wave kernel, one wave
// per matrix, ranks dealt to the waves in alternating order.
//
// Wave path (any n ≤ 256, the default): see wave_chol_solve — 16×16 MFMA
// tiles, no workgroup barrier, left-looking panels streamed from L.
// Workgroup path: right-looking, 32-column panels factored in LDS with one
// barrier per column, trailing lower triangle updated in 4×4 register blocks
// in memory, solves with the 32×32 diagonal blocks staged in LDS.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "gla_internal.hpp"

namespace gla {

struct BatchPlan {
    int device = 0;
    int64_t batch = 0;
    int nwg = 0;               // matrices above wave_max (one workgroup each)
    int nsmall = 0;            // matrices at or below it (one wave each)
    int grid_wg = 0, grid_wave = 0;
    int wave_max = 256;
    int *d_n = nullptr;        // n per matrix (device)
    int *d_wg = nullptr;       // the workgroup matrices, largest first (device)
    int *d_small = nullptr;    // the wave matrices, largest first (device)
};

namespace {

constexpr int kBThreads = 256;
constexpr int kBMax = 256;     // largest n
constexpr int kPW = 32;        // panel width
constexpr int kPLD = kPW + 1;  // padded panel row
constexpr int kWMax = kBMax;   // the wave path takes any n ≤ 256

template <typename T>
__device__ __forceinline__ T bsqrt(T x);
template <>
__device__ __forceinline__ double bsqrt<double>(double x) { return __builtin_sqrt(x); }
template <>
__device__ __forceinline__ float bsqrt<float>(float x) { return __builtin_sqrtf(x); }

__device__ __forceinline__ void border() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// ---------------------------------------------------------------- wave --
// One wave factors and solves one matrix (n ≤ 256) on 16×16 MFMA tiles,
// left-looking by 16-column panels, with no workgroup barrier.  A tile sits
// in the MFMA accumulator transposed: lane (i = l % 16, g = l / 16) holds
// row i of the tile at the columns c_q = Mf::row(l, q), q < 4.  So every
// load and store of a tile moves 16 consecutive rows of a column, and a tile
// in this layout is at once the B operand of a product whose K index runs
// through the c_q (a permutation of K that both operands share).  Per panel
// p0 (columns p0..p0+15), the rows below in groups of kMR tiles:
//   (1) C = A − L[rows, 0:p0] L[p0:p0+16, 0:p0]ᵀ on MFMA, the operands
//       streamed through a per-wave LDS ring (several slabs of K in flight,
//       which is what keeps the wave busy: the columns of L come back from
//       HBM, not from a cache); the first group also carries the diagonal
//       tile C_dd, whose two operands are the same panel block;
//   (2) after the first group's K loop, C_dd is factored in row layout by
//       the Cholesky's tile factor (cholUnblocked!'s column steps,
//       src/cholesky.jl:3-15), L_dd is written back, y_p = L_dd⁻¹ b_p is
//       solved by 16 broadcast steps and W = L_dd⁻¹ formed in registers;
//   (3) each group: X = C L_dd⁻ᵀ as Xᵀ = W Cᵀ on MFMA, written back, and
//       b_rows −= X y_p.
// Then Lᵀ x = y from the last panel: z = y_p − L_belowᵀ x_below (16-lane row
// sums), and a 16-step back substitution on the columns of L_dd.

// A[r][c] for r, c < n, else 0; the load itself is unconditional (clamped
// indices), so that a run of them stays in flight.
// Global-memory view of a pointer read from device memory (the compiler
// cannot infer it, and flat accesses cost more than global ones).
template <typename T>
using gmem = __attribute__((address_space(1))) T;

template <typename T>
__device__ __forceinline__ T ldz(const gmem<T> *__restrict__ A, int lda, int n, int r, int c)
{
    const int rc = r < n ? r : n - 1, cc = c < n ? c : n - 1;
    const T v = A[rc + (int64_t)cc * lda];
    return (r < n && c < n) ? v : T(0);
}

// Forward substitution on the diagonal tile (row layout, a[c] = L[i][c],
// zero above the diagonal; rd[c] = 1/L[c][c] in every lane): z_i ← y_i.
template <int C, typename T>
struct DiagFwd {
    static __device__ __forceinline__ void run(const T a[16], const T rd[16], int i, T &z)
    {
        const T y = row_bcast<C>(z) * rd[C];
        z = i == C ? y : fma_t(-a[C], y, z);
        DiagFwd<C + 1, T>::run(a, rd, i, z);
    }
};
template <typename T>
struct DiagFwd<16, T> {
    static __device__ __forceinline__ void run(const T *, const T *, int, T &) {}
};

// Back substitution on the diagonal tile's transpose (column layout,
// lc[k] = L[k][i], zero above the diagonal; rdi = 1/L[i][i]): z_i ← x_i.
template <int C, typename T>
struct DiagBwd {
    static __device__ __forceinline__ void run(const T lc[16], T rdi, int i, T &z)
    {
        const T x = row_bcast<C>(z * rdi);
        z = i == C ? x : fma_t(-lc[C], x, z);
        DiagBwd<C - 1, T>::run(lc, rdi, i, z);
    }
};
template <typename T>
struct DiagBwd<-1, T> {
    static __device__ __forceinline__ void run(const T *, T, int, T &) {}
};

constexpr int kPW2 = 32;     // panel width of the wave path
constexpr int kMR = 4;       // ring blocks per slab, less one
constexpr int kMTL = 17;     // padded row of the staging tile

// K-loop operand staging: slabs of KB columns of L (16-row blocks, one for
// the panel's rows and one per row tile, each [k][16 rows]), an NS-deep ring
// per wave filled by direct-to-LDS dword loads (element-aligned for any lda,
// and no registers held by loads in flight).
template <typename T>
struct WaveCfg;
template <>
struct WaveCfg<double> {
    static constexpr int KB = 8, NS = 3;
};
template <>
struct WaveCfg<float> {
    static constexpr int KB = 16, NS = 2;
};
template <typename T>
struct WaveLds {
    static constexpr int KB = WaveCfg<T>::KB, NS = WaveCfg<T>::NS;
    static constexpr int kBlock = 16 * KB;                   // elements of one block
    static constexpr int kSlab = (kMR + 1) * kBlock;
    static constexpr int kLoads = kSlab * (int)sizeof(T) / 256;   // dword loads per slab per wave
    static constexpr int kRing = NS * kSlab;
    static constexpr int kElems = kRing + 16 * kMTL + kBMax + 64;
};

template <int N>
__device__ __forceinline__ void wait_vm()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
}

template <typename T>
__device__ void wave_chol_solve(int n, T *Ap, int lda, T *bp, int *info, T *S)
{
    typedef WaveLds<T> WL;
    constexpr int KB = WL::KB, NS = WL::NS;
    gmem<T> *__restrict__ A = (gmem<T> *)Ap;
    gmem<T> *__restrict__ b = (gmem<T> *)bp;
    typedef PanelMfma<T> Mf;
    typedef typename Mf::acc_t acc_t;
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
    T *ring = S;                        // NS slabs
    T *Ts = S + NS * WL::kSlab;         // 16 × kMTL staging tile
    T *bv = Ts + 16 * kMTL;             // b, then y, then x
    T *red = bv + kBMax;                // per-group partial sums
    int cq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) cq[q] = Mf::row(lane, q);
    // this lane's element of each 256-byte load: (row, k) within a block, and
    // the byte within the element (FP64 elements take two lanes)
    constexpr int EPL = 256 / (int)sizeof(T);    // elements per load
    const int e_l = sizeof(T) == 8 ? (lane >> 1) : lane;
    const int lrow = e_l & 15, lk = e_l >> 4;
    const int lbyte = sizeof(T) == 8 ? 4 * (lane & 1) : 0;
    // Streams slabs 0..nslab−1 (KB columns of L each) of NB 16-row blocks
    // (block t from row rb[t]; rows past n are clamped to row n − 1, finite
    // values whose products are never stored) through the ring, as deep as
    // the ring holds slabs of that size; pre() runs once the first slabs are
    // in flight, then body(slab, s) for each slab in order.  Every slab loads
    // all NB blocks, so each wait counts a fixed number of loads.  Load
    // address: a scalar base (the slab's columns) plus this lane's byte
    // offset within the block.
    auto stream = [&](auto nbc, const int *rb, int nslab, auto &&pre, auto &&body) {
        constexpr int NB = decltype(nbc)::value;
        constexpr int LPS = NB * (WL::kBlock / EPL);      // loads per slab
        constexpr int SS = NB * WL::kBlock;               // slab elements
        constexpr int NSB = WL::kRing / SS > 16 ? 16 : WL::kRing / SS;
        static_assert(NSB >= 2 && (NSB - 2) * LPS <= 63, "ring depth");
        unsigned ob[NB];
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            const int rr = min(rb[t] + lrow, n - 1);
            ob[t] = (unsigned)(rr + lk * lda) * (unsigned)sizeof(T) + lbyte;
        }
        auto issue = [&](int s) {
            T *st = ring + (s % NSB) * SS;
#pragma unroll
            for (int u = 0; u < WL::kBlock / EPL; ++u) {
                const char *base = (const char *)(A + (int64_t)(s * KB + u * (EPL / 16)) * lda);
#pragma unroll
                for (int t = 0; t < NB; ++t)
                    __builtin_amdgcn_global_load_lds(
                        (const void *)(base + ob[t]),
                        (__attribute__((address_space(3))) void *)(st + t * WL::kBlock + u * EPL),
                        4, 0, 0);
            }
        };
        for (int s = 0; s < NSB - 1 && s < nslab; ++s) issue(s);
        pre();
        for (int s = 0; s < nslab; ++s) {
            if (s + NSB - 1 < nslab)
                wait_vm<(NSB - 2) * LPS>();
            else
                wait_vm<0>();
            border();   // keeps the LDS reads below the wait
            if (s + NSB - 1 < nslab) issue(s + NSB - 1);
            body((const T *)(ring + (s % NSB) * SS), s);
        }
    };
    for (int e = lane; e < n; e += 64) bv[e] = b[e];
    border();
    int bad = 0;
    auto nothing = [] {};
    // factor + invert + forward-solve one 16×16 diagonal tile given in the
    // accumulator layout: L written back at (row0, row0) for the first nv
    // rows, y = L⁻¹ z at bv[row0..], W = L⁻¹ to w (w[q] = W[i][c_q])
    auto diag_tile = [&](const acc_t &t, int row0, int nv, T zin, T w[4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) Ts[li * kMTL + cq[q]] = t[q];
        border();
        T a[16], rd[16];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            a[j] = li >= nv ? (j == li ? T(1) : T(0)) : (j <= li ? Ts[li * kMTL + j] : T(0));
        border();
        int fneg = 16;
        TileFactorCols<0, T>::run(a, rd, li, fneg);
        if (fneg < nv && !bad) bad = row0 + fneg + 1;
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] = j <= li ? a[j] : T(0);
        if (g == 0 && li < nv) {
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j <= li) A[(row0 + li) + (int64_t)(row0 + j) * lda] = a[j];
        }
        T z = li < nv ? zin : T(0);
        DiagFwd<0, T>::run(a, rd, li, z);
        if (g == 0 && li < nv) bv[row0 + li] = z;
        // W = L⁻¹: a[k] becomes the multiplier L[i][k]/L[k][k] (i > k)
        T ri = rd[0];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            ri = li == k ? rd[k] : ri;
            a[k] = li > k ? a[k] * rd[k] : T(0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = li == cq[q] ? T(1) : T(0);
        TileInvSteps<0, T>::run(a, w);
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] *= ri;
        border();
    };
    // Σ over the tile's 4 columns of this lane (values v, weights y) and
    // the four lane groups: the same sum in every lane of row i
    auto row_dot = [&](const T v[4], const T y[4]) {
        T part = T(0);
#pragma unroll
        for (int q = 0; q < 4; ++q) part = fma_t(v[q], y[q], part);
        red[g * 16 + li] = part;
        border();
        const T sum = (red[li] + red[16 + li]) + (red[32 + li] + red[48 + li]);
        border();
        return sum;
    };
    for (int p0 = 0; p0 < n; p0 += kPW2) {
        // panel of 32 columns as tiles 0 (p0..p0+15) and 1 (p0+16..p0+31)
        const int pw = min(kPW2, n - p0), pw0 = min(16, pw), pw1 = pw - pw0;
        const int nslab = p0 / KB;
        T w0[4], w1[4], y0q[4], y1q[4];
        acc_t l10;
        {
            // the 32×32 diagonal block: T00, T10, T11
            acc_t t00, t10, t11;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                t00[q] = ldz(A, lda, n, p0 + li, p0 + cq[q]);
                t10[q] = ldz(A, lda, n, p0 + 16 + li, p0 + cq[q]);
                t11[q] = ldz(A, lda, n, p0 + 16 + li, p0 + 16 + cq[q]);
            }
            const int rb[2] = {p0, p0 + 16};
            stream(std::integral_constant<int, 2>{}, rb, nslab, nothing, [&](const T *st, int) {
#pragma unroll
                for (int h = 0; h < KB / 4; ++h) {
                    const int o = (4 * h + g) * 16 + li;
                    const T v0 = -st[o], v1 = st[WL::kBlock + o];
                    t00 = Mf::mma(v0, -v0, t00);
                    t10 = Mf::mma(v0, v1, t10);
                    t11 = Mf::mma(-v1, v1, t11);
                }
            });
            diag_tile(t00, p0, pw0, bv[p0 + li], w0);
#pragma unroll
            for (int q = 0; q < 4; ++q) y0q[q] = cq[q] < pw0 ? bv[p0 + cq[q]] : T(0);
            // L10 = T10 L00⁻ᵀ, T11 −= L10 L10ᵀ, z1 = b1 − L10 y0
            l10 = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int q = 0; q < 4; ++q) l10 = Mf::mma(w0[q], t10[q], l10);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (li < pw1 && cq[q] < pw0) A[(p0 + 16 + li) + (int64_t)(p0 + cq[q]) * lda] = l10[q];
                t11 = Mf::mma(-l10[q], l10[q], t11);
            }
            T l10v[4] = {l10[0], l10[1], l10[2], l10[3]};
            const T s1 = row_dot(l10v, y0q);
            diag_tile(t11, p0 + 16, pw1, bv[p0 + 16 + li] - s1, w1);
#pragma unroll
            for (int q = 0; q < 4; ++q) y1q[q] = cq[q] < pw1 ? bv[p0 + 16 + cq[q]] : T(0);
        }
        // the rows below, two tiles at a time: C_j −= R L_jᵀ for the
        // panel's tiles j = 0, 1, then X_0 = C_0 L00⁻ᵀ, C_1 −= X_0 L10ᵀ,
        // X_1 = C_1 L11⁻ᵀ, written back; b_rows −= X_0 y0 + X_1 y1
        for (int r0 = p0 + kPW2; r0 < n; r0 += 32) {
            const int nt = min(2, (n - r0 + 15) >> 4);
            acc_t c[2][2];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        c[t][j][q] = t < nt ? ldz(A, lda, n, r0 + 16 * t + li, p0 + 16 * j + cq[q]) : T(0);
            const int rb[4] = {p0, p0 + 16, r0, r0 + 16};
            auto group = [&](auto nbc) {
                constexpr int NB = decltype(nbc)::value;
                stream(nbc, rb, nslab, nothing, [&](const T *st, int) {
#pragma unroll
                    for (int h = 0; h < KB / 4; ++h) {
                        const int o = (4 * h + g) * 16 + li;
                        const T v0 = -st[o], v1 = -st[WL::kBlock + o];
#pragma unroll
                        for (int t = 0; t < NB - 2; ++t) {
                            const T rv = st[(2 + t) * WL::kBlock + o];
                            c[t][0] = Mf::mma(v0, rv, c[t][0]);
                            c[t][1] = Mf::mma(v1, rv, c[t][1]);
                        }
                    }
                });
            };
            if (nt == 1)
                group(std::integral_constant<int, 3>{});
            else
                group(std::integral_constant<int, 4>{});
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (t >= nt) continue;
                acc_t x0 = {T(0), T(0), T(0), T(0)}, x1 = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int q = 0; q < 4; ++q) x0 = Mf::mma(w0[q], c[t][0][q], x0);
#pragma unroll
                for (int q = 0; q < 4; ++q) c[t][1] = Mf::mma(-l10[q], x0[q], c[t][1]);
#pragma unroll
                for (int q = 0; q < 4; ++q) x1 = Mf::mma(w1[q], c[t][1][q], x1);
                const int r = r0 + 16 * t + li;
                T xv[4], yv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (r < n && cq[q] < pw0) A[r + (int64_t)(p0 + cq[q]) * lda] = x0[q];
                    if (r < n && cq[q] < pw1) A[r + (int64_t)(p0 + 16 + cq[q]) * lda] = x1[q];
                    xv[q] = x0[q];
                    yv[q] = x1[q];
                }
                T part[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) part[q] = fma_t(xv[q], y0q[q], yv[q] * y1q[q]);
                const T one[4] = {T(1), T(1), T(1), T(1)};
                const T sum = row_dot(part, one);
                if (g == 0 && r < n) bv[r] -= sum;
            }
        }
    }
    // Lᵀ x = y, panels from the last: x_p from the 32×32 diagonal block
    // (x1 = L11ᵀ⁻¹ y1, x0 = L00ᵀ⁻¹ (y0 − L10ᵀ x1), 16-step back
    // substitutions on the tiles' columns), then y[0:p0] −= L[p0:p0+32, 0:p0]ᵀ x_p
    // streamed through the ring (16-lane row sums per column)
    auto diag_bwd = [&](const T lt[4], int row0, int nv, T zin) {
#pragma unroll
        for (int q = 0; q < 4; ++q) Ts[li * kMTL + cq[q]] = lt[q];
        border();
        T lc[16];
#pragma unroll
        for (int k = 0; k < 16; ++k)
            lc[k] = (li < nv && k < nv) ? (k >= li ? Ts[k * kMTL + li] : T(0)) : (k == li ? T(1) : T(0));
        const T rdi = li < nv ? T(1) / Ts[li * kMTL + li] : T(1);
        T z = li < nv ? zin : T(0);
        border();
        DiagBwd<15, T>::run(lc, rdi, li, z);
        if (g == 0 && li < nv) bv[row0 + li] = z;
        border();
        return li < nv ? z : T(0);
    };
    for (int p0 = ((n - 1) / kPW2) * kPW2; p0 >= 0; p0 -= kPW2) {
        const int pw = min(kPW2, n - p0), pw0 = min(16, pw), pw1 = pw - pw0;
        const int nslab = p0 / KB;
        T lt00[4], lt10[4], lt11[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            lt00[q] = ldz(A, lda, n, p0 + li, p0 + cq[q]);
            lt10[q] = ldz(A, lda, n, p0 + 16 + li, p0 + cq[q]);
            lt11[q] = ldz(A, lda, n, p0 + 16 + li, p0 + 16 + cq[q]);
        }
        T xr0 = T(0), xr1 = T(0);
        auto solve = [&] {
            xr1 = diag_bwd(lt11, p0 + 16, pw1, bv[p0 + 16 + li]);
            // S[j] = Σ_i L10[i][j] x1[i] for the columns c_q of this lane group
            T s[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) s[q] = row_sum16(lt10[q] * xr1);
            if (li == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) red[cq[q]] = s[q];
            }
            border();
            const T z0 = bv[p0 + li] - red[li];
            border();
            xr0 = diag_bwd(lt00, p0, pw0, z0);
        };
        const int rb[2] = {p0, p0 + 16};
        stream(std::integral_constant<int, 2>{}, rb, nslab, solve, [&](const T *st, int s) {
#pragma unroll
            for (int h = 0; h < KB / 4; ++h) {
                const int o = (4 * h + g) * 16 + li;
                const T v = row_sum16(fma_t(st[o], xr0, st[WL::kBlock + o] * xr1));
                if (li == 0) bv[s * KB + 4 * h + g] -= v;
            }
        });
        border();
    }
    for (int e = lane; e < n; e += 64) b[e] = bv[e];
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

// Workgroup items: one matrix per workgroup.
template <typename T>
__global__ __launch_bounds__(kBThreads, 1) void batched_wg_kernel(
    int nwg, const int *__restrict__ items, const int *__restrict__ nvec,
    T *const *__restrict__ Aptr, const int *__restrict__ ldav, T *const *__restrict__ bptr,
    int *__restrict__ info)
{
    __shared__ T P[kBMax * kPLD];
    __shared__ T bs[kBMax];
    __shared__ T sc[2 * kPW];
    for (int it = blockIdx.x; it < nwg; it += gridDim.x) {
        const int i = items[it];
        wg_chol_solve<T>(nvec[i], Aptr[i], ldav[i], bptr[i], info ? info + i : nullptr, P, bs, sc);
    }
}

// Wave items: the matrices at or below wave_max, largest first; global wave
// gw of GW takes rank gw in even rounds and GW − 1 − gw in odd ones (rounds
// of GW ranks), which evens out the sum of sizes per wave.
template <typename T>
__global__ __launch_bounds__(kBThreads) __attribute__((amdgpu_waves_per_eu(2))) void batched_wave_kernel(
    int nsmall, const int *__restrict__ order, const int *__restrict__ nvec,
    T *const *__restrict__ Aptr, const int *__restrict__ ldav, T *const *__restrict__ bptr,
    int *__restrict__ info)
{
    __shared__ T S[4 * WaveLds<T>::kElems];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int GW = gridDim.x * 4, gw = blockIdx.x * 4 + w;
    for (int base = 0, odd = 0; base < nsmall; base += GW, odd ^= 1) {
        const int j = base + (odd ? GW - 1 - gw : gw);
        if (j >= nsmall) continue;
        const int i = order[j];
        wave_chol_solve<T>(nvec[i], Aptr[i], ldav[i], bptr[i], info ? info + i : nullptr,
                           S + w * WaveLds<T>::kElems);
    }
}

}  // namespace

template <typename T>
int batched_launch(Handle *h, BatchPlan *pl, T *const *A, const int *lda, T *const *b, int *info,
                   hipStream_t s)
{
    (void)h;
    if (pl->nwg > 0) {
        hipLaunchKernelGGL((batched_wg_kernel<T>), dim3(pl->grid_wg), dim3(kBThreads), 0, s, pl->nwg,
                           pl->d_wg, pl->d_n, A, lda, b, info);
        GLA_LAUNCH_CHECK();
    }
    if (pl->nsmall > 0) {
        hipLaunchKernelGGL((batched_wave_kernel<T>), dim3(pl->grid_wave), dim3(kBThreads), 0, s,
                           pl->nsmall, pl->d_small, pl->d_n, A, lda, b, info);
        GLA_LAUNCH_CHECK();
    }
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
    size_t k = 0;
    while (k < order.size() && n_host[order[k]] > wave_max) ++k;
    auto *pl = new BatchPlan();
    pl->device = h->device;
    pl->batch = batch;
    pl->nwg = (int)k;
    pl->nsmall = (int)(order.size() - k);
    pl->wave_max = wave_max;
    pl->grid_wg = std::max(1, std::min<int>(pl->nwg, h->num_cus * grid_mult));
    pl->grid_wave = std::max(1, std::min<int>((pl->nsmall + 3) / 4, h->num_cus * grid_mult));
    if (batch > 0) {
        GLA_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&pl->d_n), sizeof(int) * batch));
        GLA_HIP_TRY(hipMemcpy(pl->d_n, n_host, sizeof(int) * batch, hipMemcpyHostToDevice));
    }
    if (pl->nwg > 0) {
        GLA_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&pl->d_wg), sizeof(int) * pl->nwg));
        GLA_HIP_TRY(hipMemcpy(pl->d_wg, order.data(), sizeof(int) * pl->nwg, hipMemcpyHostToDevice));
    }
    if (pl->nsmall > 0) {
        GLA_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&pl->d_small), sizeof(int) * pl->nsmall));
        GLA_HIP_TRY(hipMemcpy(pl->d_small, order.data() + k, sizeof(int) * pl->nsmall,
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
    if (pl->d_wg) (void)hipFree(pl->d_wg);
    if (pl->d_small) (void)hipFree(pl->d_small);
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
