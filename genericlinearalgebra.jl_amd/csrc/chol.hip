// Cholesky on gfx950: cholUnblocked!, cholBlocked!, cholRecursive!
// (src/cholesky.jl:3-55) and the rdiv!/rankUpdate! steps they call.
//
// Kernels
//   chol_factor_kernel  One workgroup factors a diagonal block (≤128×128) in
//                       place.  The block is held as 16×16 MFMA accumulator
//                       tiles; wave 0 holds the diagonal tiles, waves 1–3
//                       the rest.  For each 16-wide column block the tiles
//                       below the diagonal become X = A·L_kk⁻ᵀ, then wave 0
//                       factors the next diagonal tile in registers column
//                       by column as cholUnblocked! does (sqrt, reciprocal
//                       scale, rank-1 update; DPP broadcasts move the column)
//                       and inverts it while waves 1–3 apply the rank-16
//                       update C -= X_i X_jᵀ on MFMA.  The tile inverses are
//                       kept for the panel solve.
//   chol_solve_kernel   rdiv!(A21, LowerTriangular(A11)') (src/cholesky.jl:28,
//                       48) for a ≤128-wide L11: each workgroup takes 64 rows
//                       of A21, L11's tiles and tile inverses sit in LDS, and
//                       every 16-column step is X = B_kb·L_kk⁻ᵀ then
//                       B_j -= X·L_(j,kb)ᵀ on MFMA.  A batched form (grid y)
//                       inverts many diagonal blocks at once.
//   gemm (gemm.hip)     The trailing rankUpdate!(Hermitian(A22,:L), A21, -1)
//                       as an MFMA SYRK over lower tiles.
//
// Drivers
//   blocked    right-looking, block nb ≤ 128, depth-1 look-ahead: the update
//              of the next block column runs on the caller's stream right
//              after the panel, the rest of the trailing update runs on the
//              handle's side stream and overlaps the next panel.
//   look-ahead the same loop with wide blocks (NB = 512 by default): the
//              diagonal block by the blocked driver, the panel by the
//              recursive rdiv!, the trailing SYRK with K = NB.
//   recursive  halves n as cholRecursive! does; below a threshold the blocked
//              driver factors the leaf.  rdiv! with a large L11 is itself
//              recursive (two half solves and a GEMM); its leaves are
//              chol_solve_kernel launches.
#include <algorithm>
#include <mutex>

#include "gla_internal.hpp"

namespace gla {

__global__ void k_set_int(int *p, int v) { *p = v; }

namespace {

constexpr int kPanelNB = 128;   // max diagonal block
constexpr int kPanelRB = 64;    // rows of A21 per solve workgroup
constexpr int kPanelThreads = 256;
constexpr int kXLD = 17;        // padded row length of the LDS column blocks

#define GLA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != GLA_OK) return _rc; \
    } while (0)



// One wave holds a 16×16 tile with lane (i = l % 16, g = l / 16) owning row
// i, columns g, g+4, g+8, g+12 (v[t] = tile[i][g + 4t]).  Row i of column c
// therefore sits in lane i + 16*(c % 4), slot c / 4.

// v from lane `src` (a wave-uniform constant) to every lane, through SGPRs.
template <typename T>
__device__ __forceinline__ T lane_bcast(T v, int src)
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



template <typename T>
__device__ __forceinline__ void tile_factor_rows(T a[16], T rd[16], int nvalid, int *info,
                                                 int64_t gbase)
{
    const int i = threadIdx.x & 15;
    int fneg = 16;
    TileFactorCols<0, T>::run(a, rd, i, fneg);
    if (fneg < 16 && fneg < nvalid && (threadIdx.x & 63) == 0 && info)
        atomicCAS(info, 0, (int)(gbase + fneg + 1));
#pragma unroll
    for (int j = 0; j < 16; ++j) a[j] = j <= i ? a[j] : T(0);
}

// w <- L⁻¹ for a lower 16×16 tile L (rd = reciprocals of its diagonal), in
// the four-column layout (w[t] = W[i][g + 4t]).  Forward elimination of
// [L | I] with unscaled rows: V[i] = e_i − Σ_{k<i} m_ik V[k] with the
// multipliers m_ik = L[i][k]·rd[k] formed once (zero for i ≤ k, so every
// step is one broadcast and one unconditional FMA per column), then
// W[i] = rd[i]·V[i] (the steps are TileInvSteps).  L comes in row layout
// (li[k] = L[i][k]).
template <typename T>
__device__ __forceinline__ void tile_inverse_rows(const T li[16], const T rd[16], T w[4])
{
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, g = lane >> 4;
    T m[16], ri = rd[0];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        m[k] = i > k ? li[k] * rd[k] : T(0);
        ri = i == k ? rd[k] : ri;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) w[t] = (i == g + 4 * t) ? T(1) : T(0);
    TileInvSteps<0, T>::run(m, w);
#pragma unroll
    for (int t = 0; t < 4; ++t) w[t] *= ri;
}

// The same from L in the four-column layout (v[t] = L[i][g + 4t]).
template <typename T>
__device__ __forceinline__ void tile_inverse(const T v[4], const T rd[16], T w[4])
{
    const int i = threadIdx.x & 15;
    T li[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) li[k] = __shfl(v[k >> 2], i + 16 * (k & 3));
    tile_inverse_rows<T>(li, rd, w);
}

// Orders one wave's LDS accesses, for data exchanged between its lanes.
__device__ __forceinline__ void wave_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Strictly-lower tile rows of waves 1..3: {7, 2}, {6, 3}, {5, 4, 1} (nine,
// nine and ten tiles), so that each 16-wide step spreads its panel tiles
// and its trailing updates evenly.  Slot q -> tile (bi, bj); bi = -1 past
// the wave's last tile.
__host__ __device__ constexpr int oslot_bi(int w, int q)
{
    return w == 1 ? (q < 7 ? 7 : (q < 9 ? 2 : -1))
         : w == 2 ? (q < 6 ? 6 : (q < 9 ? 3 : -1))
                  : (q < 5 ? 5 : (q < 9 ? 4 : (q < 10 ? 1 : -1)));
}
__host__ __device__ constexpr int oslot_bj(int w, int q)
{
    return w == 1 ? (q < 7 ? q : q - 7) : w == 2 ? (q < 6 ? q : q - 6) : (q < 5 ? q : (q < 9 ? q - 5 : q - 9));
}
// Every strictly-lower tile of the 8×8-tile block belongs to exactly one
// slot of waves 1–3, and every slot past a wave's last is marked unused.
constexpr bool oslots_cover()
{
    int seen[8][8] = {};
    for (int w = 1; w <= 3; ++w)
        for (int q = 0; q < 10; ++q) {
            const int bi = oslot_bi(w, q), bj = oslot_bj(w, q);
            if (bi < 0) continue;
            if (bi > 7 || bj < 0 || bj >= bi) return false;
            ++seen[bi][bj];
        }
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < i; ++j)
            if (seen[i][j] != 1) return false;
    return true;
}
static_assert(oslots_cover(), "the tile slots of waves 1-3 must cover the strictly-lower tiles once");

// Cholesky tile factor: slot q of wave W (wave 0: diagonal tile q).
template <int W>
__host__ __device__ constexpr int dslot_bi(int q) { return W == 0 ? q : oslot_bi(W, q); }
template <int W>
__host__ __device__ constexpr int dslot_bj(int q) { return W == 0 ? q : oslot_bj(W, q); }

// d (a diagonal tile, transposed) -= X_j X_jᵀ for X_j = tile row j of X (LDS)
template <typename T>
__device__ __forceinline__ void diag_update(typename PanelMfma<T>::acc_t &d, const T *X, int j)
{
    typedef PanelMfma<T> M;
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15;
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) {
        const int kk = (lane >> 4) + 4 * kq;
        const T x = X[(j * 16 + lc) * kXLD + kk];
        d = M::mma(-x, x, d);
    }
}

// Wave 0: factors diagonal tile j (d, transposed) in row layout through the
// wave's LDS tile D, inverts it, leaves L (transposed) in d, the inverse in Li
// (row-major) and, when given, LinvG / Linv32.
template <typename T>
__device__ __forceinline__ void diag_factor(typename PanelMfma<T>::acc_t &d, int j, int nb,
                                            int *info, int64_t info_base, T *D, T *Li,
                                            T *__restrict__ LinvG, float *__restrict__ Linv32)
{
    typedef PanelMfma<T> M;
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15;
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) D[lc * kXLD + M::row(lane, r)] = d[r];
    wave_order();
    T a[16], w[4], rd[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) a[c] = c <= i ? D[i * kXLD + c] : T(0);
    tile_factor_rows<T>(a, rd, nb - j * 16, info, info_base + j * 16);
    tile_inverse_rows<T>(a, rd, w);
    wave_order();
    if (g == 0)
#pragma unroll
        for (int c = 0; c < 16; ++c) D[i * kXLD + c] = a[c];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        Li[i * kXLD + g + 4 * t] = w[t];
        if (LinvG) LinvG[j * 256 + i * 16 + g + 4 * t] = w[t];
        if (Linv32) Linv32[j * 256 + i * 16 + g + 4 * t] = float(w[t]);
    }
    wave_order();
#pragma unroll
    for (int r = 0; r < 4; ++r) d[r] = D[lc * kXLD + M::row(lane, r)];
}

// One workgroup factors the nb×nb (nb ≤ 128) diagonal block as 16×16 MFMA
// tiles held transposed (lane: row bi·16 + l%16 of the tile, columns
// bj·16 + row(l, r)).  Wave 0 holds the eight diagonal tiles, waves 1–3 the
// strictly-lower ones.  Each 16-wide step kb takes two barriers: (A) waves
// 1–3 form the panel tiles X = A_(i,kb) L_kk⁻ᵀ while wave 0 applies the
// previous step's update to the diagonal tiles beyond kb + 1; (B) wave 0
// updates, factors and inverts diagonal tile kb + 1 while waves 1–3 apply
// the rank-16 update to theirs, so the serial tile factor overlaps the
// trailing update.  Every product is formed from the same operands in the
// same order as the right-looking sweep cholBlocked! describes
// (src/cholesky.jl:17-35).
// TS: storage type of A (T, or FP32 for the mixed-precision driver, which
// factors its FP32 diagonal tiles in FP64); Linv32: optional FP32 copy of
// the tile inverses.
template <typename T, int W, typename TS = T>
__device__ __forceinline__ void factor_body(int nb, TS *__restrict__ A, int64_t lda, int *info,
                                             int64_t info_base, T *__restrict__ LinvG, T *Xb,
                                             T *Db, T *Lib, float *__restrict__ Linv32 = nullptr)
{
    typedef PanelMfma<T> M;
    typedef typename M::acc_t acc_t;
    constexpr int NQ = W == 0 ? 8 : 10;
    constexpr int NBLK = kPanelNB / 16;
    constexpr int XS = kPanelNB * kXLD, LS = 16 * kXLD;
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15;
    const int nkb = (nb + 15) / 16;

    // tiles transposed: lane holds row bi*16 + lc at columns bj*16 + row(lane, r)
    acc_t acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int bi = dslot_bi<W>(q), bj = dslot_bj<W>(q);
        if (bi < 0) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gi = bi * 16 + lc, gj = bj * 16 + M::row(lane, r);
            T v = T(0);
            if (gj <= gi) v = gi < nb ? T(A[gi + (int64_t)gj * lda]) : (gi == gj ? T(1) : T(0));
            acc[q][r] = v;
        }
    }
    if constexpr (W == 0) diag_factor<T>(acc[0], 0, nb, info, info_base, Db, Lib, LinvG, Linv32);
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < NBLK; ++kb) {
        if (kb >= nkb) break;
        T *X = Xb + (kb & 1) * XS;
        // (A) waves 1-3: the panel tiles X = A_(i,kb) L_kk⁻ᵀ (rdiv!), as
        //     Xᵀ = L_kk⁻¹ A_(i,kb)ᵀ with the B operand from registers, then to
        //     LDS (row-major); wave 0: the diagonal tiles beyond kb + 1 take
        //     the previous step's update
        if constexpr (W == 0) {
            if (kb > 0) {
                const T *Xp = Xb + ((kb - 1) & 1) * XS;
#pragma unroll
                for (int j = kb + 1; j < NBLK; ++j) {
                    if (j >= nkb) break;
                    diag_update<T>(acc[j], Xp, j);
                }
            }
        } else {
            const T *Lik = Lib + (kb & 1) * LS;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int bi = dslot_bi<W>(q), bj = dslot_bj<W>(q);
                if (bi < 0 || bj != kb) continue;
                if (bi >= nkb) continue;
                acc_t x = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int kq = 0; kq < 4; ++kq)
                    x = M::mma(Lik[lc * kXLD + M::row(lane, kq)], acc[q][kq], x);
                acc[q] = x;
#pragma unroll
                for (int r = 0; r < 4; ++r) X[(bi * 16 + lc) * kXLD + M::row(lane, r)] = x[r];
            }
        }
        __syncthreads();
        // (B) wave 0: the next diagonal tile takes this step's update and is
        //     factored and inverted; waves 1-3: rank-16 trailing update of
        //     their tiles, C_ijᵀ -= X_j X_iᵀ (rankUpdate!)
        if constexpr (W == 0) {
            if (kb + 1 < nkb) {
                diag_update<T>(acc[kb + 1], X, kb + 1);
                diag_factor<T>(acc[kb + 1], kb + 1, nb, info, info_base, Db,
                               Lib + ((kb + 1) & 1) * LS, LinvG, Linv32);
            }
        } else {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int bi = dslot_bi<W>(q), bj = dslot_bj<W>(q);
                if (bi < 0 || bj <= kb) continue;
                if (bi >= nkb) continue;
#pragma unroll
                for (int kq = 0; kq < 4; ++kq) {
                    const int kk = (lane >> 4) + 4 * kq;
                    const T a = -X[(bj * 16 + lc) * kXLD + kk];
                    const T b = X[(bi * 16 + lc) * kXLD + kk];
                    acc[q] = M::mma(a, b, acc[q]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int bi = dslot_bi<W>(q), bj = dslot_bj<W>(q);
        if (bi < 0) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gi = bi * 16 + lc, gj = bj * 16 + M::row(lane, r);
            if (gi < nb && gj <= gi) A[gi + (int64_t)gj * lda] = TS(acc[q][r]);
        }
    }
}

// Factors the nb×nb (nb ≤ 128) diagonal block at A in place (lower
// triangle; the strict upper triangle is not touched) and stores the
// inverses of its 16×16 diagonal tiles (row-major, 256 elements each) in
// LinvG for the solve that follows.  One workgroup.
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void chol_factor_kernel(int nb, T *__restrict__ A,
                                                                    int64_t lda, int *info,
                                                                    int64_t info_base,
                                                                    T *__restrict__ LinvG)
{
    __shared__ T Xb[2 * kPanelNB * kXLD];
    __shared__ T Db[2 * 16 * kXLD];
    __shared__ T Lib[2 * 16 * kXLD];
    switch (threadIdx.x >> 6) {
    case 0: factor_body<T, 0>(nb, A, lda, info, info_base, LinvG, Xb, Db, Lib); break;
    case 1: factor_body<T, 1>(nb, A, lda, info, info_base, LinvG, Xb, Db, Lib); break;
    case 2: factor_body<T, 2>(nb, A, lda, info, info_base, LinvG, Xb, Db, Lib); break;
    default: factor_body<T, 3>(nb, A, lda, info, info_base, LinvG, Xb, Db, Lib); break;
    }
}

// ---------------------------------------------------------------- LDLᴴ --
// _ldlt_lower! (src/ldlt.jl:16-33) of a 16×16 tile in row layout, column C:
// δ = a_CC (broadcast from lane C), l_i = a_iC / δ (:21), then
// a_ij −= l_i (l_j δ) for C < j ≤ i (:24-26).  Lane C keeps δ on the
// diagonal, rows below take l; entries right of the diagonal are left
// meaningless (cleared by the caller).  dv[C] receives δ.  A zero pivot
// divides by zero and the Inf/NaN propagate, as in the reference.
template <int C, int J, typename T>
struct LdltColUpdate {
    static __device__ __forceinline__ void run(T a[16], T l, T ld)
    {
        a[J] = fma_t(-l, row_bcast<J>(ld), a[J]);
        LdltColUpdate<C, J + 1, T>::run(a, l, ld);
    }
};
template <int C, typename T>
struct LdltColUpdate<C, 16, T> {
    static __device__ __forceinline__ void run(T *, T, T) {}
};
template <int C, typename T>
struct LdltTileCols {
    static __device__ __forceinline__ void run(T a[16], T dv[16], int i)
    {
        const T d = row_bcast<C>(a[C]);
        dv[C] = d;
        const T l = a[C] / d;
        const T ld = l * d;
        LdltColUpdate<C, C + 1, T>::run(a, l, ld);
        a[C] = i == C ? d : l;
        LdltTileCols<C + 1, T>::run(a, dv, i);
    }
};
template <typename T>
struct LdltTileCols<16, T> {
    static __device__ __forceinline__ void run(T *, T *, int) {}
};

// w <- L⁻¹ (four-column layout) for the unit lower tile L in row layout
// (li[k] = L[i][k] for k < i): forward elimination with unit pivots.
template <typename T>
__device__ __forceinline__ void tile_inverse_unit_rows(const T li[16], T w[4])
{
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, g = lane >> 4;
    T m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = i > k ? li[k] : T(0);
#pragma unroll
    for (int t = 0; t < 4; ++t) w[t] = (i == g + 4 * t) ? T(1) : T(0);
    TileInvSteps<0, T>::run(m, w);
}

// Wave 0: LDLᴴ of diagonal tile j (d, transposed) in row layout through
// the wave's LDS tile D (LdltTileCols), the inverse of its unit factor to
// Li (row-major) and LinvG, the pivots to dv; d keeps δ on the diagonal and
// L below it (transposed).
template <typename T>
__device__ __forceinline__ void ldlt_diag_factor(typename PanelMfma<T>::acc_t &d, int j, T *D,
                                                 T *Li, T *dv, T *__restrict__ LinvG)
{
    typedef PanelMfma<T> M;
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15;
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) D[lc * kXLD + M::row(lane, r)] = d[r];
    wave_order();
    T a[16], w[4], dd[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) a[c] = c <= i ? D[i * kXLD + c] : T(0);
    LdltTileCols<0, T>::run(a, dd, i);
#pragma unroll
    for (int c = 0; c < 16; ++c) a[c] = c <= i ? a[c] : T(0);
    tile_inverse_unit_rows<T>(a, w);
    wave_order();
    if (g == 0)
#pragma unroll
        for (int c = 0; c < 16; ++c) D[i * kXLD + c] = a[c];
    if (lane < 16) {
        T dl = dd[0];   // the pivots are the same in every lane: lane c stores δ_c
#pragma unroll
        for (int c = 1; c < 16; ++c) dl = lane == c ? dd[c] : dl;
        dv[lane] = dl;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        Li[i * kXLD + g + 4 * t] = w[t];
        if (LinvG) LinvG[j * 256 + i * 16 + g + 4 * t] = w[t];
    }
    wave_order();
#pragma unroll
    for (int r = 0; r < 4; ++r) d[r] = D[lc * kXLD + M::row(lane, r)];
}

// d (a diagonal tile, transposed) −= W_j L_jᵀ for tile row j of W (X) and
// of L (Y) in LDS
template <typename T>
__device__ __forceinline__ void ldlt_diag_update(typename PanelMfma<T>::acc_t &d, const T *X,
                                                 const T *Y, int j)
{
    typedef PanelMfma<T> M;
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15;
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) {
        const int kk = (lane >> 4) + 4 * kq;
        d = M::mma(-Y[(j * 16 + lc) * kXLD + kk], X[(j * 16 + lc) * kXLD + kk], d);
    }
}

// The LDLᴴ of the nb×nb (nb ≤ 128) diagonal block at A, with the tiles held
// as in factor_body (transposed; wave 0 the diagonal tiles, waves 1–3 the
// rest) and the same two-barrier step: (A) waves 1–3 form W = A L_kk⁻ᵀ
// (rdiv! by UnitLowerTriangular, src/ldlt.jl:86) and L = W D_kk⁻¹ (rdiv! by
// Diagonal, :87, a division) for their panel tiles while wave 0 applies the
// previous step's update to the diagonal tiles beyond kb + 1; (B) wave 0
// updates diagonal tile kb + 1, factors it (LdltTileCols) and inverts its
// unit factor while waves 1–3 apply C −= W Lᵀ (:88-95) on MFMA.  Writes D
// on the diagonal and L below it, and the inverses of the unit 16×16
// diagonal tiles (row-major, 256 elements each) to LinvG.
template <typename T, int W>
__device__ __forceinline__ void ldlt_body(int nb, T *__restrict__ A, int64_t lda,
                                          T *__restrict__ LinvG, T *Xb, T *Yb, T *Db, T *Lib,
                                          T *dvs)
{
    typedef PanelMfma<T> M;
    typedef typename M::acc_t acc_t;
    constexpr int NQ = W == 0 ? 8 : 10;
    constexpr int NBLK = kPanelNB / 16;
    constexpr int XS = kPanelNB * kXLD, LS = 16 * kXLD;
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15;
    const int nkb = (nb + 15) / 16;

    acc_t acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int bi = dslot_bi<W>(q), bj = dslot_bj<W>(q);
        if (bi < 0) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gi = bi * 16 + lc, gj = bj * 16 + M::row(lane, r);
            T v = T(0);
            if (gj <= gi) v = gi < nb ? A[gi + (int64_t)gj * lda] : (gi == gj ? T(1) : T(0));
            acc[q][r] = v;
        }
    }
    if constexpr (W == 0) ldlt_diag_factor<T>(acc[0], 0, Db, Lib, dvs, LinvG);
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < NBLK; ++kb) {
        if (kb >= nkb) break;
        T *X = Xb + (kb & 1) * XS;   // W of column block kb
        T *Y = Yb + (kb & 1) * XS;   // L of column block kb
        // (A)
        if constexpr (W == 0) {
            if (kb > 0) {
                const T *Xp = Xb + ((kb - 1) & 1) * XS, *Yp = Yb + ((kb - 1) & 1) * XS;
#pragma unroll
                for (int j = kb + 1; j < NBLK; ++j) {
                    if (j >= nkb) break;
                    ldlt_diag_update<T>(acc[j], Xp, Yp, j);
                }
            }
        } else {
            const T *Lik = Lib + (kb & 1) * LS;
            const T *dv = dvs + (kb & 1) * 16;
            T dcol[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) dcol[r] = dv[M::row(lane, r)];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int bi = dslot_bi<W>(q), bj = dslot_bj<W>(q);
                if (bi < 0 || bj != kb) continue;
                if (bi >= nkb) continue;
                acc_t x = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int kq = 0; kq < 4; ++kq)
                    x = M::mma(Lik[lc * kXLD + M::row(lane, kq)], acc[q][kq], x);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    acc[q][r] = x[r] / dcol[r];
                    X[(bi * 16 + lc) * kXLD + M::row(lane, r)] = x[r];
                    Y[(bi * 16 + lc) * kXLD + M::row(lane, r)] = acc[q][r];
                }
            }
        }
        __syncthreads();
        // (B)
        if constexpr (W == 0) {
            if (kb + 1 < nkb) {
                ldlt_diag_update<T>(acc[kb + 1], X, Y, kb + 1);
                ldlt_diag_factor<T>(acc[kb + 1], kb + 1, Db, Lib + ((kb + 1) & 1) * LS,
                                    dvs + ((kb + 1) & 1) * 16, LinvG);
            }
        } else {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int bi = dslot_bi<W>(q), bj = dslot_bj<W>(q);
                if (bi < 0 || bj <= kb) continue;
                if (bi >= nkb) continue;
#pragma unroll
                for (int kq = 0; kq < 4; ++kq) {
                    const int kk = (lane >> 4) + 4 * kq;
                    const T a = -Y[(bj * 16 + lc) * kXLD + kk];
                    const T b = X[(bi * 16 + lc) * kXLD + kk];
                    acc[q] = M::mma(a, b, acc[q]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int bi = dslot_bi<W>(q), bj = dslot_bj<W>(q);
        if (bi < 0) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gi = bi * 16 + lc, gj = bj * 16 + M::row(lane, r);
            if (gi < nb && gj <= gi) A[gi + (int64_t)gj * lda] = acc[q][r];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kPanelThreads) void ldlt_tile_factor_kernel(int nb, T *__restrict__ A,
                                                                         int64_t lda,
                                                                         T *__restrict__ LinvG)
{
    __shared__ T Xb[2 * kPanelNB * kXLD];
    __shared__ T Yb[2 * kPanelNB * kXLD];
    __shared__ T Db[2 * 16 * kXLD];
    __shared__ T Lib[2 * 16 * kXLD];
    __shared__ T dvs[2 * 16];
    switch (threadIdx.x >> 6) {
    case 0: ldlt_body<T, 0>(nb, A, lda, LinvG, Xb, Yb, Db, Lib, dvs); break;
    case 1: ldlt_body<T, 1>(nb, A, lda, LinvG, Xb, Yb, Db, Lib, dvs); break;
    case 2: ldlt_body<T, 2>(nb, A, lda, LinvG, Xb, Yb, Db, Lib, dvs); break;
    default: ldlt_body<T, 3>(nb, A, lda, LinvG, Xb, Yb, Db, Lib, dvs); break;
    }
}

// The same for a block that an earlier kernel has classified through
// flags: flags[0] != 0 — nothing to do (the caller's path failed);
// flags[1] == 0 — A = I + E with ‖E‖_F ≤ 0.1·√eps, whose factor is
// L = I + Φ(E), Φ(E) = the strictly lower part of E plus half its diagonal,
// since (I + F)(I + F)ᵀ = I + E gives F = Φ(E − F Fᵀ) and ‖F Fᵀ‖ ≤ ‖E‖² ≤
// 0.01·eps, below the rounding of any entry; likewise the 16×16 tile
// inverses (I + F_tt)⁻¹ = I − F_tt; otherwise the full factorisation.
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void chol_factor_near_identity_kernel(
    int nb, T *__restrict__ A, int64_t lda, int *info, int64_t info_base, T *__restrict__ LinvG,
    const int *flags)
{
    __shared__ T Xb[2 * kPanelNB * kXLD];
    __shared__ T Db[2 * 16 * kXLD];
    __shared__ T Lib[2 * 16 * kXLD];
    if (flags[0] != 0) return;
    if (flags[1] == 0) {
        const int nt = (nb + 15) / 16;
        for (int e = threadIdx.x; e < nt * 256; e += kPanelThreads) {
            const int t = e >> 8, i = (e >> 4) & 15, j = e & 15;
            const int gi = 16 * t + i, gj = 16 * t + j;
            T v;
            if (gi >= nb || gj >= nb) v = gi == gj ? T(1) : T(0);
            else if (gi == gj) v = T(1) - (A[gi + (int64_t)gi * lda] - T(1)) * T(0.5);
            else v = gi > gj ? -A[gi + (int64_t)gj * lda] : T(0);
            LinvG[e] = v;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nb; e += kPanelThreads)
            A[e + (int64_t)e * lda] = T(1) + (A[e + (int64_t)e * lda] - T(1)) * T(0.5);
        return;
    }
    switch (threadIdx.x >> 6) {
    case 0: factor_body<T, 0>(nb, A, lda, info, info_base, LinvG, Xb, Db, Lib); break;
    case 1: factor_body<T, 1>(nb, A, lda, info, info_base, LinvG, Xb, Db, Lib); break;
    case 2: factor_body<T, 2>(nb, A, lda, info, info_base, LinvG, Xb, Db, Lib); break;
    default: factor_body<T, 3>(nb, A, lda, info, info_base, LinvG, Xb, Db, Lib); break;
    }
}

// The same factorisation of an FP32 diagonal block computed in FP64: the
// block is read and written back in FP32, the tile inverses are stored in
// FP32 (Linv32).
__global__ __launch_bounds__(kPanelThreads) void chol_factor_f32io_kernel(int nb,
                                                                          float *__restrict__ A,
                                                                          int64_t lda, int *info,
                                                                          int64_t info_base,
                                                                          float *__restrict__ Linv32)
{
    __shared__ double Xb[2 * kPanelNB * kXLD];
    __shared__ double Db[2 * 16 * kXLD];
    __shared__ double Lib[2 * 16 * kXLD];
    switch (threadIdx.x >> 6) {
    case 0: factor_body<double, 0, float>(nb, A, lda, info, info_base, nullptr, Xb, Db, Lib, Linv32); break;
    case 1: factor_body<double, 1, float>(nb, A, lda, info, info_base, nullptr, Xb, Db, Lib, Linv32); break;
    case 2: factor_body<double, 2, float>(nb, A, lda, info, info_base, nullptr, Xb, Db, Lib, Linv32); break;
    default: factor_body<double, 3, float>(nb, A, lda, info, info_base, nullptr, Xb, Db, Lib, Linv32); break;
    }
}

// A strided batch of diagonal-block factorisations (nb ≤ 128 each), one
// workgroup per matrix: matrix b at A + b·sA, tile inverses at LinvG + b·sLinv,
// its first failing pivot in info[b] (info zeroed by the caller).
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void chol_factor_batched_kernel(int nb, T *A,
                                                                            int64_t lda,
                                                                            int64_t sA, int *info,
                                                                            T *LinvG,
                                                                            int64_t sLinv)
{
    __shared__ T Xb[2 * kPanelNB * kXLD];
    __shared__ T Db[2 * 16 * kXLD];
    __shared__ T Lib[2 * 16 * kXLD];
    const int64_t b = blockIdx.x;
    T *Ab = A + b * sA;
    T *Lg = LinvG + b * sLinv;
    int *ib = info + b;
    switch (threadIdx.x >> 6) {
    case 0: factor_body<T, 0>(nb, Ab, lda, ib, 0, Lg, Xb, Db, Lib); break;
    case 1: factor_body<T, 1>(nb, Ab, lda, ib, 0, Lg, Xb, Db, Lib); break;
    case 2: factor_body<T, 2>(nb, Ab, lda, ib, 0, Lg, Xb, Db, Lib); break;
    default: factor_body<T, 3>(nb, Ab, lda, ib, 0, Lg, Xb, Db, Lib); break;
    }
}

// Inverses of the 16×16 diagonal tiles of a lower n×n L (tiles past n are
// padded with the identity); one wave per tile.
template <typename T>
__global__ __launch_bounds__(256) void tile_inverse_kernel(int64_t n, const T *__restrict__ L,
                                                           int64_t ldl, T *__restrict__ LinvG,
                                                           int64_t sL = 0, int64_t sLinv = 0)
{
    L += blockIdx.y * sL;          // strided batch along grid y
    LinvG += blockIdx.y * sLinv;
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile * 16 >= n) return;
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, g = lane >> 4;
    T v[4], w[4], rd[16];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int j = g + 4 * t;
        const int64_t gi = tile * 16 + i, gj = tile * 16 + j;
        v[t] = (j <= i) ? (gi < n ? L[gi + gj * ldl] : (i == j ? T(1) : T(0))) : T(0);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) rd[k] = T(1) / __shfl(v[k >> 2], k + 16 * (k & 3));
    tile_inverse<T>(v, rd, w);
#pragma unroll
    for (int t = 0; t < 4; ++t) LinvG[tile * 256 + i * 16 + g + 4 * t] = w[t];
}

// Tile q of the strictly-lower tiles (bj, kb), bj > kb, of an 8×8-tile
// block, stored column block after column block: its kb and bj.
__host__ __device__ constexpr int lt_kb(int q)
{
    int kb = 0, base = 0;
    while (base + (7 - kb) <= q) {
        base += 7 - kb;
        ++kb;
    }
    return kb;
}
__host__ __device__ constexpr int lt_bj(int q)
{
    int kb = 0, base = 0;
    while (base + (7 - kb) <= q) {
        base += 7 - kb;
        ++kb;
    }
    return kb + 1 + (q - base);
}

// Index of tile (bj, kb), bj > kb, in Lt: column blocks stored one after
// another.
__device__ __forceinline__ int lt_index(int bj, int kb) { return 7 * kb - kb * (kb - 1) / 2 + (bj - kb - 1); }

// Stages the strictly-lower 16×16 tiles of the nb×nb (nb ≤ 128) lower L and
// the inverses of its diagonal tiles into LDS (row-major tiles, padded rows),
// all loads issued before the first store.  L is read with 16 consecutive
// rows of one column per 16 lanes (thread t: row t % 16, column t / 16 % 16
// of every tile), the inverses as stored (row-major, 256 per tile).
template <typename T>
__device__ __forceinline__ void stage_l_tiles(int nb, const T *__restrict__ L, int64_t ldl,
                                              const T *__restrict__ LinvG, T *Lt, T *Li)
{
    constexpr int NBLK = kPanelNB / 16;
    constexpr int NLT = NBLK * (NBLK - 1) / 2;
    constexpr int TS = 16 * kXLD;
    static_assert(kPanelThreads == 256, "one tile element per thread");
    const int nkb = (nb + 15) / 16;
    const int ri = threadIdx.x & 15, rj = (threadIdx.x >> 4) & 15;
    const int i = (threadIdx.x >> 4) & 15, j = threadIdx.x & 15;
    // unconditional loads from clamped indices, so that all of them stay in
    // flight; out-of-range values replaced afterwards
    T lv[NLT], iv[NBLK];
#pragma unroll
    for (int q = 0; q < NLT; ++q) {
        const int kb = lt_kb(q), bj = lt_bj(q);
        const int gi = bj * 16 + ri, gj = kb * 16 + rj;
        const int gic = gi < nb ? gi : nb - 1, gjc = gj < nb ? gj : nb - 1;
        const T v = L[gic + (int64_t)gjc * ldl];
        lv[q] = (gi < nb && gj < nb) ? v : T(0);
    }
#pragma unroll
    for (int q = 0; q < NBLK; ++q) {
        const T v = LinvG[(q < nkb ? q : nkb - 1) * 256 + threadIdx.x];
        iv[q] = q < nkb ? v : T(0);
    }
#pragma unroll
    for (int q = 0; q < NLT; ++q) Lt[q * TS + ri * kXLD + rj] = lv[q];
#pragma unroll
    for (int q = 0; q < NBLK; ++q) Li[q * TS + i * kXLD + j] = iv[q];
}

// acc (this wave's 16 rows × 128 columns as eight 16×16 tiles, each held
// TRANSPOSED: acc[bj] = B_bjᵀ in accumulator layout) <- acc L⁻ᵀ with L's
// tiles and tile inverses staged in LDS.  The accumulator layout of Bᵀ is
// the MFMA's B-operand layout with K index row(lane, kq) in step kq, so for
// each column block kb the wave forms X_kbᵀ = L_kk⁻¹ B_kbᵀ and updates
// B_jᵀ -= L_(j,kb) X_kbᵀ (j > kb) with the B operands straight from
// registers and only L read from LDS: no LDS round trip and no barrier
// (every wave touches only its own rows).  The column-block loop is
// unrolled, so every tile index is static.
template <typename T>
__device__ __forceinline__ void solve_tiles(int nkb, typename PanelMfma<T>::acc_t acc[8],
                                            const T *Lt, const T *Li)
{
    typedef PanelMfma<T> M;
    typedef typename M::acc_t acc_t;
    constexpr int NBLK = kPanelNB / 16;
    constexpr int TS = 16 * kXLD;
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15;
#pragma unroll
    for (int kb = 0; kb < NBLK; ++kb) {
        if (kb >= nkb) break;
        const T *Lik = Li + kb * TS;
        acc_t x = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int kq = 0; kq < 4; ++kq)
            x = M::mma(Lik[lc * kXLD + M::row(lane, kq)], acc[kb][kq], x);
        acc[kb] = x;
        T nx[4];
#pragma unroll
        for (int kq = 0; kq < 4; ++kq) nx[kq] = -x[kq];
#pragma unroll
        for (int bj = kb + 1; bj < NBLK; ++bj) {
            if (bj >= nkb) break;
            const T *Lt_jk = Lt + lt_index(bj, kb) * TS;
#pragma unroll
            for (int kq = 0; kq < 4; ++kq)
                acc[bj] = M::mma(Lt_jk[lc * kXLD + M::row(lane, kq)], nx[kq], acc[bj]);
        }
    }
}

// B (nrows×nb) <- B L⁻ᵀ for the lower nb×nb L (nb ≤ 128) whose 16×16
// diagonal-tile inverses are in LinvG — rdiv!(A21, LowerTriangular(A11)'),
// src/cholesky.jl:28,48.  Each workgroup takes 64 rows; wave w holds tile
// row w of them (8 tiles).  L's 28 strictly-lower tiles and the 8 inverses
// are read into LDS once, then solve_tiles runs the column sweep.
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void chol_solve_kernel(int nb, const T *__restrict__ L,
                                                                   int64_t ldl,
                                                                   const T *__restrict__ LinvG,
                                                                   T *__restrict__ B, int64_t ldb,
                                                                   int64_t nrows, int64_t sL,
                                                                   int64_t sLinv, int64_t sB)
{
    // batched form: blockIdx.y selects (L, LinvG, B) at the given strides
    L += blockIdx.y * sL;
    LinvG += blockIdx.y * sLinv;
    B += blockIdx.y * sB;
    typedef PanelMfma<T> M;
    typedef typename M::acc_t acc_t;
    constexpr int NBLK = kPanelNB / 16;             // 8
    constexpr int NLT = NBLK * (NBLK - 1) / 2;       // 28 strictly-lower tiles
    constexpr int TS = 16 * kXLD;                    // one padded tile
    __shared__ T Lt[NLT * TS];                       // tile (bj, kb), bj > kb
    __shared__ T Li[NBLK * TS];                      // inverses of the diagonal tiles
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int lc = lane & 15;
    const int64_t row0 = (int64_t)blockIdx.x * kPanelRB;
    const int nkb = (nb + 15) / 16;

    // the wave's rows, tiles transposed (lane: row lc, columns row(lane, r)),
    // loaded first so that their latency overlaps L's
    acc_t acc[NBLK];
    const int64_t gr = row0 + w * 16 + lc;
    const int64_t grc = gr < nrows ? gr : nrows - 1;
#pragma unroll
    for (int bj = 0; bj < NBLK; ++bj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gj = bj * 16 + M::row(lane, r);
            const int gjc = gj < nb ? gj : nb - 1;
            const T v = B[grc + (int64_t)gjc * ldb];
            acc[bj][r] = (gr < nrows && gj < nb) ? v : T(0);
        }
    stage_l_tiles<T>(nb, L, ldl, LinvG, Lt, Li);
    __syncthreads();
    solve_tiles<T>(nkb, acc, Lt, Li);
#pragma unroll
    for (int bj = 0; bj < NBLK; ++bj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gj = bj * 16 + M::row(lane, r);
            if (gr < nrows && gj < nb) B[gr + (int64_t)gj * ldb] = acc[bj][r];
        }
}

// B (nrows×nb) <- B L⁻ᵀ for a wide lower L (nb a multiple of 128, up to
// kPsolveMaxNB), already factored, with its 16×16 diagonal-tile inverses in
// LinvG: rdiv!(A21, LowerTriangular(A11)') (src/cholesky.jl:28) of a whole
// look-ahead panel in one launch.  Each workgroup takes 64 rows and sweeps
// the 128-wide column blocks left to right: block c first takes the
// products of the blocks already solved, B_c -= Σ_{c'<c} X_c' L(c,c')ᵀ, as
// one MFMA product over K = 128c with operands staged through LDS in
// 32-wide slabs, then X_c = B_c L_cc⁻ᵀ by solve_tiles.  A workgroup reads
// only L and its own rows, so no workgroup waits for another.
constexpr int kPsolveMaxNB = 768;   // wider panels: the recursive solve is faster
constexpr int kPsolveKC = 32;                 // K slab of the product
constexpr int kPsolveGLD = kPsolveKC + 1;     // padded LDS row
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void chol_psolve_kernel(int nb, const T *__restrict__ L,
                                                                    int64_t ldl,
                                                                    const T *__restrict__ LinvG,
                                                                    T *__restrict__ B, int64_t ldb,
                                                                    int64_t nrows)
{
    typedef PanelMfma<T> M;
    typedef typename M::acc_t acc_t;
    constexpr int NBLK = kPanelNB / 16;
    constexpr int NLT = NBLK * (NBLK - 1) / 2;
    constexpr int TS = 16 * kXLD;
    __shared__ T Lt[NLT * TS];
    __shared__ T Li[NBLK * TS];
    __shared__ T Gx[kPanelRB * kPsolveGLD];   // this workgroup's solved rows, one K slab
    __shared__ T Gl[kPanelNB * kPsolveGLD];   // rows of block c of L, the same K slab
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int lc = lane & 15;
    const int64_t row0 = (int64_t)blockIdx.x * kPanelRB;
    const int64_t gr = row0 + w * 16 + lc;   // this lane's row in every tile
    const int nc = nb / kPanelNB;
    for (int c = 0; c < nc; ++c) {
        // block c of the wave's rows, tiles transposed as solve_tiles takes them
        acc_t acc[NBLK];
#pragma unroll
        for (int bj = 0; bj < NBLK; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gj = (int64_t)c * kPanelNB + bj * 16 + M::row(lane, r);
                acc[bj][r] = gr < nrows ? B[gr + gj * ldb] : T(0);
            }
        // B_cᵀ -= L[128c : 128c+128, 0:128c] X[:, 0:128c]ᵀ, the global loads
        // of slab s + 1 in flight while slab s is multiplied
        const T *Lc = L + (int64_t)c * kPanelNB;
        constexpr int XE = kPanelRB * kPsolveKC / kPanelThreads;   // 8 per thread
        constexpr int LE = kPanelNB * kPsolveKC / kPanelThreads;   // 16 per thread
        const int kend = c * kPanelNB;
        T xv[XE], lv[LE];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int e = 0; e < XE; ++e) {
                const int idx = threadIdx.x + e * kPanelThreads;
                const int i = idx % kPanelRB, kk = idx / kPanelRB;
                xv[e] = row0 + i < nrows ? B[row0 + i + (int64_t)(k0 + kk) * ldb] : T(0);
            }
#pragma unroll
            for (int e = 0; e < LE; ++e) {
                const int idx = threadIdx.x + e * kPanelThreads;
                const int i = idx % kPanelNB, kk = idx / kPanelNB;
                lv[e] = Lc[i + (int64_t)(k0 + kk) * ldl];
            }
        };
        // block c − 1 was stored by other waves at the end of the last sweep
        __syncthreads();
        if (kend > 0) fetch(0);
        for (int k0 = 0; k0 < kend; k0 += kPsolveKC) {
            __syncthreads();   // the previous slab has been read
#pragma unroll
            for (int e = 0; e < XE; ++e) {
                const int idx = threadIdx.x + e * kPanelThreads;
                Gx[(idx % kPanelRB) * kPsolveGLD + idx / kPanelRB] = xv[e];
            }
#pragma unroll
            for (int e = 0; e < LE; ++e) {
                const int idx = threadIdx.x + e * kPanelThreads;
                Gl[(idx % kPanelNB) * kPsolveGLD + idx / kPanelNB] = lv[e];
            }
            __syncthreads();
            if (k0 + kPsolveKC < kend) fetch(k0 + kPsolveKC);
#pragma unroll
            for (int kq = 0; kq < kPsolveKC / 4; ++kq) {
                const int kk = 4 * kq + (lane >> 4);
                const T xb = -Gx[(w * 16 + lc) * kPsolveGLD + kk];
#pragma unroll
                for (int bj = 0; bj < NBLK; ++bj)
                    acc[bj] = M::mma(Gl[(bj * 16 + lc) * kPsolveGLD + kk], xb, acc[bj]);
            }
        }
        // X_c = B_c L_cc⁻ᵀ
        __syncthreads();   // Lt/Li of the previous block have been read
        stage_l_tiles<T>(kPanelNB, L + (int64_t)c * kPanelNB * (1 + ldl), ldl,
                         LinvG + (int64_t)c * NBLK * 256, Lt, Li);
        __syncthreads();
        solve_tiles<T>(NBLK, acc, Lt, Li);
#pragma unroll
        for (int bj = 0; bj < NBLK; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gj = (int64_t)c * kPanelNB + bj * 16 + M::row(lane, r);
                if (gr < nrows) B[gr + gj * ldb] = acc[bj][r];
            }
        // the stores of block c are read back by this workgroup's later
        // slabs (workgroup-scope visibility at the barrier that opens them)
    }
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
int launch_factor(hipStream_t s, int nb, T *A, int64_t lda, int *info, int64_t info_base,
                  T *LinvG)
{
    hipLaunchKernelGGL((chol_factor_kernel<T>), dim3(1), dim3(kPanelThreads), 0, s, nb, A, lda,
                       info, info_base, LinvG);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template <typename T>
int launch_solve(hipStream_t s, int nb, const T *L, int64_t ldl, const T *LinvG, T *B,
                 int64_t ldb, int64_t nrows)
{
    if (nrows <= 0 || nb <= 0) return GLA_OK;
    hipLaunchKernelGGL((chol_solve_kernel<T>), dim3((unsigned)cdiv(nrows, kPanelRB)),
                       dim3(kPanelThreads), 0, s, nb, L, ldl, LinvG, B, ldb, nrows, 0, 0, 0);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template <typename T>
int launch_psolve(hipStream_t s, int nb, const T *L, int64_t ldl, const T *LinvG, T *B,
                  int64_t ldb, int64_t nrows)
{
    if (nrows <= 0) return GLA_OK;
    hipLaunchKernelGGL((chol_psolve_kernel<T>), dim3((unsigned)cdiv(nrows, kPanelRB)),
                       dim3(kPanelThreads), 0, s, nb, L, ldl, LinvG, B, ldb, nrows);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// `batch` independent solves at strides sL, sLinv, sB (one grid row each).
template <typename T>
int launch_solve_batched(hipStream_t s, int nb, const T *L, int64_t ldl, int64_t sL,
                         const T *LinvG, int64_t sLinv, T *B, int64_t ldb, int64_t sB,
                         int64_t nrows, int64_t batch)
{
    if (nrows <= 0 || batch <= 0 || nb <= 0) return GLA_OK;
    hipLaunchKernelGGL((chol_solve_kernel<T>), dim3((unsigned)cdiv(nrows, kPanelRB), (unsigned)batch),
                       dim3(kPanelThreads), 0, s, nb, L, ldl, LinvG, B, ldb, nrows, sL, sLinv, sB);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template <typename T>
__global__ void identity_blocks_kernel(int64_t nblk, T *D)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nblk * kPanelNB * kPanelNB) {
        const int64_t r = e % kPanelNB, c = (e / kPanelNB) % kPanelNB;
        D[e] = r == c ? T(1) : T(0);
    }
}

template <typename T>
int launch_tile_inverse(hipStream_t s, int64_t n, const T *L, int64_t ldl, T *LinvG)
{
    if (n <= 0) return GLA_OK;
    hipLaunchKernelGGL((tile_inverse_kernel<T>), dim3((unsigned)cdiv(cdiv(n, 16), 4)), dim3(256),
                       0, s, n, L, ldl, LinvG);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// Elements of the per-call workspace: the 16×16 diagonal-tile inverses of
// one 128-wide panel (blocked driver) or of L11 (recursive rdiv!).
inline size_t linv_elems(int64_t n) { return (size_t)(cdiv(n, 16) + 8) * 256; }

// Right-looking blocked Cholesky of the lower triangle with look-ahead.
// keep_linv: every block's 16×16 tile inverses stay in Linv at the block's
// own offset (tile q of the whole matrix at q·256), for a solve that
// follows; otherwise each block reuses the start of Linv.
template <typename T>
int chol_blocked_dev(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int nb,
                     int64_t info_base, T *Linv, bool use_side = true, bool keep_linv = false)
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
        T *Lv = keep_linv ? Linv + (k / 16) * 256 : Linv;
        GLA_TRY(launch_factor<T>(s, kb, Akk, lda, h->dinfo, info_base + k, Lv));    // :25
        GLA_TRY(launch_solve<T>(s, kb, Akk, lda, Lv, A21, lda, below));             // :27-28
        if (below == 0) break;
        if (!overlap) {
            // :30-31 rankUpdate! of all of A22 in one launch: with nothing
            // beside it, splitting off the next block column only adds a
            // launch (same products, same order)
            GLA_TRY(gemm<T>(s, 'N', 'T', 1, below, below, kb, T(-1), A21, lda, A21, lda, T(1),
                            A21 + kb * lda, lda, 0));
            continue;
        }
        const int kb2 = (int)std::min<int64_t>(nb, below);
        const int64_t rest = below - kb2;
        if (overlap && rest > 0) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (overlap && step > 0) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
        // :30-31 rankUpdate! on the next block column first (look-ahead) ...
        GLA_TRY(gemm<T>(s, 'N', 'T', 1, below, kb2, kb, T(-1), A21, lda, A21, lda, T(1),
                        A21 + kb * lda, lda, 0));
        // ... and on the rest of A22 beside the next panels
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
    return GLA_OK;
}

// B (m×n) <- B L⁻ᵀ, L n×n lower, already factored, with the inverses of its
// 16×16 diagonal tiles in Linv.  Recursive on L: two half solves and a GEMM.
template <typename T>
int trsm_rlt_dev(hipStream_t s, int64_t m, int64_t n, const T *L, int64_t ldl, const T *Linv,
                 T *B, int64_t ldb)
{
    if (m <= 0 || n <= 0) return GLA_OK;
    if (n <= kPanelNB) return launch_solve<T>(s, (int)n, L, ldl, Linv, B, ldb, m);
    // split at a multiple of the panel width
    int64_t n1 = ((n / 2) + kPanelNB - 1) / kPanelNB * kPanelNB;
    if (n1 >= n) n1 = n - kPanelNB;
    GLA_TRY(trsm_rlt_dev<T>(s, m, n1, L, ldl, Linv, B, ldb));
    // B2 -= X1 * L21ᵀ, L21 = L[n1:n, 0:n1]
    GLA_TRY(gemm<T>(s, 'N', 'T', 0, m, n - n1, n1, T(-1), B, ldb, L + n1, ldl, T(1),
                    B + n1 * ldb, ldb, 0));
    return trsm_rlt_dev<T>(s, m, n - n1, L + n1 + n1 * ldl, ldl, Linv + (n1 / 16) * 256,
                           B + n1 * ldb, ldb);
}

// Blocked right-looking Cholesky with wide blocks (NB > 128) and depth-1
// look-ahead — cholBlocked!'s loop (src/cholesky.jl:17-35) with the
// diagonal block factored by the 128-wide driver on the caller's stream,
// the panel solved in one chol_psolve_kernel launch (or the recursive
// rdiv!), the next block column updated on
// the caller's stream and the rest of the trailing rankUpdate! (K = NB, an
// efficient MFMA SYRK) on the side stream, where it overlaps the next
// diagonal factor and panel solve.
template <typename T>
int chol_la_dev(Handle *h, hipStream_t s0, int64_t n, T *A, int64_t lda, int NB,
                int64_t info_base, T *Linv)
{
    if (n <= 0) return GLA_OK;
    if (NB <= kPanelNB || n <= NB) return chol_blocked_dev<T>(h, s0, n, A, lda, kPanelNB, info_base, Linv);
    const int64_t nsteps = cdiv(n, NB);
    const bool overlap = nsteps > 2;
    // With CUs reserved from the SYRK, the chain runs on an internal stream
    // with every CU and the SYRK on the masked one, both forked from (and
    // joined back to) the caller's stream.
    const bool masked = overlap && h->masked != nullptr;
    hipStream_t s = masked ? h->chain : s0;
    hipStream_t side = masked ? h->masked : h->side;
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_fork, s0));
        GLA_HIP_TRY(hipStreamWaitEvent(side, h->ev_fork, 0));
        if (masked) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_fork, 0));
    }
    int64_t k = 0;
    for (int64_t step = 0; step < nsteps; ++step, k += NB) {
        const int64_t kb = std::min<int64_t>(NB, n - k);
        const int64_t below = n - k - kb;
        T *Akk = A + k + k * lda;
        T *A21 = Akk + kb;
        // :27-28 rdiv! of the panel: one launch over all of it when the block
        // is a whole number of 128-wide tiles, else the recursive solve
        const bool fused = h->chol_psolve != 0 && kb % kPanelNB == 0 && kb <= kPsolveMaxNB;
        GLA_TRY(chol_blocked_dev<T>(h, s, kb, Akk, lda, kPanelNB, info_base + k, Linv, false,
                                    fused));
        if (below == 0) break;
        if (fused) {
            GLA_TRY(launch_psolve<T>(s, (int)kb, Akk, lda, Linv, A21, lda, below));
        } else {
            GLA_TRY(launch_tile_inverse<T>(s, kb, Akk, lda, Linv));
            GLA_TRY(trsm_rlt_dev<T>(s, below, kb, Akk, lda, Linv, A21, lda));
        }
        const int64_t kb2 = std::min<int64_t>(NB, below);
        const int64_t rest = below - kb2;
        // the side stream's SYRK is released after the next-block-column
        // update (chol_syrk_late, default), which then runs without the
        // SYRK's workgroups crowding it, or right after the panel solve
        const bool late = h->chol_syrk_late != 0;
        if (overlap && rest > 0 && !late) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (overlap && step > 0) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
        GLA_TRY(gemm<T>(s, 'N', 'T', 1, below, kb2, kb, T(-1), A21, lda, A21, lda, T(1),
                        A21 + kb * lda, lda, 0));
        if (overlap && rest > 0 && late) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (rest > 0) {
            hipStream_t us = overlap ? side : s;
            if (overlap) GLA_HIP_TRY(hipStreamWaitEvent(side, h->ev_step[0], 0));
            T *X2 = A21 + kb2;
            GLA_TRY(gemm<T>(us, 'N', 'T', 1, rest, rest, kb, T(-1), X2, lda, X2, lda, T(1),
                            X2 + (kb + kb2) * lda, lda, 0));
            if (overlap) GLA_HIP_TRY(hipEventRecord(h->ev_step[1], side));
        }
    }
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_join, side));
        GLA_HIP_TRY(hipStreamWaitEvent(s0, h->ev_join, 0));
        if (masked) {
            GLA_HIP_TRY(hipEventRecord(h->ev_fork, s));
            GLA_HIP_TRY(hipStreamWaitEvent(s0, h->ev_fork, 0));
        }
    }
    return GLA_OK;
}

// cholRecursive! (src/cholesky.jl:37-55): halve until n <= leaf, then the
// blocked driver.
// use_side = false keeps every launch on the caller's stream (the leaves'
// look-ahead then runs in order: the same arithmetic, no side stream).
template <typename T>
int chol_recursive_dev(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int64_t leaf,
                       int64_t info_base, T *Linv, bool use_side = true)
{
    if (n <= leaf)
        return chol_blocked_dev<T>(h, s, n, A, lda, kPanelNB, info_base, Linv, use_side);
    int64_t n2 = n / 2;                                              // :44
    if (n2 > kPanelNB) n2 = n2 / kPanelNB * kPanelNB;                // keep blocks aligned
    GLA_TRY(chol_recursive_dev<T>(h, s, n2, A, lda, leaf, info_base, Linv, use_side));   // :46
    GLA_TRY(launch_tile_inverse<T>(s, n2, A, lda, Linv));            // :48 rdiv!
    GLA_TRY(trsm_rlt_dev<T>(s, n - n2, n2, A, lda, Linv, A + n2, lda));
    GLA_TRY(gemm<T>(s, 'N', 'T', 1, n - n2, n - n2, n2, T(-1), A + n2, lda, A + n2, lda, T(1),
                    A + n2 + n2 * lda, lda, 0));                     // :51 rankUpdate!
    return chol_recursive_dev<T>(h, s, n - n2, A + n2 + n2 * lda, lda, leaf,
                                 info_base + n2, Linv, use_side);    // :52
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

// Building blocks shared with the mixed-precision driver.
template <typename T>
int chol_factor_block(hipStream_t s, int nb, T *A, int64_t lda, int *info, int64_t info_base,
                      T *LinvG)
{
    return launch_factor<T>(s, nb, A, lda, info, info_base, LinvG);
}
template <typename T>
int ldlt_factor_block(hipStream_t s, int nb, T *A, int64_t lda, T *LinvG)
{
    hipLaunchKernelGGL((ldlt_tile_factor_kernel<T>), dim3(1), dim3(kPanelThreads), 0, s, nb, A,
                       lda, LinvG);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}
template <typename T>
int chol_factor_block_near_identity(hipStream_t s, int nb, T *A, int64_t lda, int *info,
                                    int64_t info_base, T *LinvG, const int *flags)
{
    hipLaunchKernelGGL((chol_factor_near_identity_kernel<T>), dim3(1), dim3(kPanelThreads), 0, s,
                       nb, A, lda, info, info_base, LinvG, flags);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}
int chol_factor_block_f32io(hipStream_t s, int nb, float *A, int64_t lda, int *info,
                            int64_t info_base, float *Linv32)
{
    hipLaunchKernelGGL(chol_factor_f32io_kernel, dim3(1), dim3(kPanelThreads), 0, s, nb, A, lda,
                       info, info_base, Linv32);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}
template <typename T>
int chol_solve_block(hipStream_t s, int nb, const T *L, int64_t ldl, const T *LinvG, T *B,
                     int64_t ldb, int64_t nrows)
{
    return launch_solve<T>(s, nb, L, ldl, LinvG, B, ldb, nrows);
}
template <typename T>
int tile_inverses(hipStream_t s, int64_t n, const T *L, int64_t ldl, T *LinvG)
{
    return launch_tile_inverse<T>(s, n, L, ldl, LinvG);
}
template <typename T>
int trsm_rlt(hipStream_t s, int64_t m, int64_t n, const T *L, int64_t ldl, const T *Linv, T *B,
             int64_t ldb)
{
    return trsm_rlt_dev<T>(s, m, n, L, ldl, Linv, B, ldb);
}
template <typename T>
int chol_factor_batched(hipStream_t s, int nb, T *A, int64_t lda, int64_t sA, int *info, T *LinvG,
                        int64_t sLinv, int64_t batch)
{
    if (batch <= 0 || nb <= 0) return GLA_OK;
    hipLaunchKernelGGL((chol_factor_batched_kernel<T>), dim3((unsigned)batch), dim3(kPanelThreads),
                       0, s, nb, A, lda, sA, info, LinvG, sLinv);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}
template <typename T>
int tile_inverses_batched(hipStream_t s, int64_t n, const T *L, int64_t ldl, int64_t sL, T *LinvG,
                          int64_t sLinv, int64_t batch)
{
    if (n <= 0 || batch <= 0) return GLA_OK;
    hipLaunchKernelGGL((tile_inverse_kernel<T>), dim3((unsigned)cdiv(cdiv(n, 16), 4), (unsigned)batch),
                       dim3(256), 0, s, n, L, ldl, LinvG, sL, sLinv);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}
template <typename T>
int trsm_rlt_batched(hipStream_t s, int64_t m, int n, const T *L, int64_t ldl, int64_t sL,
                     const T *Linv, int64_t sLinv, T *B, int64_t ldb, int64_t sB, int64_t batch)
{
    if (n > kPanelNB) {
        set_error("trsm_rlt_batched: n = %d exceeds %d", n, kPanelNB);
        return GLA_ERR_UNSUPPORTED;
    }
    return launch_solve_batched<T>(s, n, L, ldl, sL, Linv, sLinv, B, ldb, sB, m, batch);
}
size_t tile_inverse_elems(int64_t n) { return linv_elems(n); }

// Inverses of the 128×128 diagonal blocks of a lower n×n L, transposed:
// block b (rows/columns 128b .. 128b+127, the last one possibly partial)
// gets Dt_b = L_bb⁻ᵀ, column-major 128×128 at Dt + b·128².  Linv holds the
// 16×16 diagonal-tile inverses of L (tile_inverses).  Two launches: the
// identity, then every block's rdiv! I·L_bb⁻ᵀ at once.
template <typename T>
int diag_block_inverses(hipStream_t s, int64_t n, const T *L, int64_t ldl, const T *Linv, T *Dt)
{
    if (n <= 0) return GLA_OK;
    const int64_t nblk = cdiv(n, kPanelNB), nfull = n / kPanelNB;
    const int64_t bs = (int64_t)kPanelNB * kPanelNB;
    hipLaunchKernelGGL((identity_blocks_kernel<T>), dim3((unsigned)cdiv(nblk * bs, 256)), dim3(256),
                       0, s, nblk, Dt);
    GLA_LAUNCH_CHECK();
    const int64_t dstride = kPanelNB * (1 + ldl), lstride = (kPanelNB / 16) * 256;
    GLA_TRY(launch_solve_batched<T>(s, kPanelNB, L, ldl, dstride, Linv, lstride, Dt, kPanelNB, bs,
                                    kPanelNB, nfull));
    if (nfull < nblk)
        GLA_TRY(launch_solve_batched<T>(s, (int)(n - nfull * kPanelNB), L + nfull * dstride, ldl, 0,
                                        Linv + nfull * lstride, 0, Dt + nfull * bs, kPanelNB, 0,
                                        kPanelNB, 1));
    return GLA_OK;
}

#define GLA_INST_BLOCKS(T)                                                                    \
    template int chol_factor_block<T>(hipStream_t, int, T *, int64_t, int *, int64_t, T *);  \
    template int ldlt_factor_block<T>(hipStream_t, int, T *, int64_t, T *);                  \
    template int chol_factor_block_near_identity<T>(hipStream_t, int, T *, int64_t, int *,   \
                                                    int64_t, T *, const int *);              \
    template int chol_solve_block<T>(hipStream_t, int, const T *, int64_t, const T *, T *,   \
                                     int64_t, int64_t);                                      \
    template int tile_inverses<T>(hipStream_t, int64_t, const T *, int64_t, T *);            \
    template int trsm_rlt<T>(hipStream_t, int64_t, int64_t, const T *, int64_t, const T *,   \
                             T *, int64_t); \
    template int chol_factor_batched<T>(hipStream_t, int, T *, int64_t, int64_t, int *, T *,  \
                                        int64_t, int64_t);                                    \
    template int tile_inverses_batched<T>(hipStream_t, int64_t, const T *, int64_t, int64_t,  \
                                          T *, int64_t, int64_t);                             \
    template int trsm_rlt_batched<T>(hipStream_t, int64_t, int, const T *, int64_t, int64_t,  \
                                     const T *, int64_t, T *, int64_t, int64_t, int64_t);     \
    template int diag_block_inverses<T>(hipStream_t, int64_t, const T *, int64_t, const T *, T *);

GLA_INST_BLOCKS(double)
GLA_INST_BLOCKS(float)
#undef GLA_INST_BLOCKS

// algo: 0 unblocked, 1 blocked (param = blocksize), 2 recursive (param = cutoff)
// Step width of the look-ahead driver for chol_block −1 (the default), from
// n = 2048…16384 sweeps on MI355X: below 3072 the recursive driver (0), then
// 256-wide steps while the chain of diagonal factor and panel solve sets the
// pace, wider steps as the trailing SYRK grows to dominate.
inline int64_t auto_chol_block(int64_t n)
{
    if (n < 3072) return 0;
    if (n < 10240) return 256;
    if (n < 14336) return 512;
    return 1024;
}

template <typename T>
int chol_device(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int algo, int64_t param,
                T *Linv)
{
    if (n == 0) return GLA_OK;
    if (!Linv) {
        GLA_TRY(ensure_workspace(h, sizeof(T) * linv_elems(n)));
        Linv = reinterpret_cast<T *>(h->ws);
    }
    if (algo == 0) {
        // cholUnblocked!: one factor launch when the matrix fits the 128×128
        // block, otherwise the same column sweep blocked by 128.
        return chol_blocked_dev<T>(h, s, n, A, lda, kPanelNB, 0, Linv);
    }
    if (algo == 1) {
        int nb = (int)std::min<int64_t>(std::max<int64_t>(param, 1), kPanelNB);
        return chol_blocked_dev<T>(h, s, n, A, lda, nb, 0, Linv);
    }
    if (algo == 3) return chol_la_dev<T>(h, s, n, A, lda, (int)param, 0, Linv);
    // cholRecursive!: the reference recurses to 1×1 (cutoff does not reach
    // the A11 branch, :46); on the device the recursion stops at a leaf of
    // at least 2048 (or `cutoff` if larger) that the blocked driver factors.
    const int64_t NB = h->chol_block < 0 ? auto_chol_block(n) : h->chol_block;
    if (NB > kPanelNB) return chol_la_dev<T>(h, s, n, A, lda, (int)NB, 0, Linv);
    int64_t leaf = std::max<int64_t>(param, 2048);
    return chol_recursive_dev<T>(h, s, n, A, lda, leaf, 0, Linv);
}

template int chol_device<double>(Handle *, hipStream_t, int64_t, double *, int64_t, int, int64_t,
                                 double *);
template int chol_device<float>(Handle *, hipStream_t, int64_t, float *, int64_t, int, int64_t,
                                float *);

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
    if (m == 0 || n == 0) return GLA_OK;
    GLA_TRY(ensure_workspace(h, sizeof(T) * linv_elems(n)));
    T *Linv = reinterpret_cast<T *>(h->ws);
    GLA_TRY(launch_tile_inverse<T>(s, n, L, ldl, Linv));
    return trsm_rlt_dev<T>(s, m, n, L, ldl, Linv, B, ldb);
}

}  // namespace

// ------------------------------------------------ no-allocation entry point --
// cholRecursive! for callers that cannot allocate device memory (code under
// HIP graph capture, serving loops): no handle, no hipMalloc.  Its scratch —
// the 16×16 diagonal-tile inverses of the panel solves and the info word —
// is this file-scope array, one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself: its device address on the
// current GPU comes from hipGetSymbolAddress, taken once per device.
constexpr int64_t kNoAllocMaxN = 32768;
constexpr size_t kNoAllocLinvBytes = (size_t)(kNoAllocMaxN / 16 + 8) * 256 * sizeof(double);
constexpr size_t kNoAllocBytes = kNoAllocLinvBytes + 256;   // + the info word
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current GPU.
int scratch_address(void **out)
{
    int dev = 0;
    GLA_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device ordinal %d beyond the %d this library tracks", dev, kMaxDevices);
        return GLA_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        GLA_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GLA_OK;
}

int chol_recursive_noalloc(char uplo, int64_t n, double *A, int64_t lda, int64_t cutoff,
                           int64_t *info, void *stream)
{
    if (uplo != 'L' && uplo != 'U' && uplo != 'l' && uplo != 'u') {
        set_error("uplo must be 'L' or 'U'");
        return -1;
    }
    if (n < 0) { set_error("n < 0"); return -2; }
    if (n > kNoAllocMaxN) {
        set_error("n = %lld exceeds the no-allocation entry point's maximum %lld (its fixed "
                  "scratch holds the tile inverses of n <= %lld); use gla_chol_recursive_d",
                  (long long)n, (long long)kNoAllocMaxN, (long long)kNoAllocMaxN);
        return GLA_ERR_UNSUPPORTED;
    }
    if (!A && n > 0) { set_error("A is null"); return -3; }
    if (lda < std::max<int64_t>(1, n)) { set_error("lda < max(1,n)"); return -4; }
    void *base = nullptr;
    GLA_TRY(scratch_address(&base));
    double *Linv = reinterpret_cast<double *>(base);
    int *dinfo = reinterpret_cast<int *>(static_cast<unsigned char *>(base) + kNoAllocLinvBytes);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, dinfo, 0);   // a kernel, not a memset
    GLA_LAUNCH_CHECK();
    // A handle-shaped view of the scratch: the drivers read only its info
    // word when use_side is false.
    Handle view;
    view.dinfo = dinfo;
    const bool upper = (uplo == 'U' || uplo == 'u');
    if (upper) GLA_TRY(swap_triangles<double>(s, n, A, lda));
    if (n > 0)
        GLA_TRY(chol_recursive_dev<double>(&view, s, n, A, lda, std::max<int64_t>(cutoff, 2048), 0,
                                           Linv, false));
    if (upper) GLA_TRY(swap_triangles<double>(s, n, A, lda));
    if (info) {
        int hv = 0;
        GLA_HIP_TRY(hipMemcpyAsync(&hv, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
        GLA_HIP_TRY(hipStreamSynchronize(s));
        *info = hv;
    }
    return GLA_OK;
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
int gla_chol_recursive_noalloc_d(char uplo, int64_t n, double *A, int64_t lda, int64_t cutoff,
                                 int64_t *info, void *stream)
{
    return gla::chol_recursive_noalloc(uplo, n, A, lda, cutoff, info, stream);
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
