// Internal declarations shared by the HIP translation units of libgla_hip.so.
// gfx950 (MI355X) only.  Matrices are column-major with explicit leading
// dimensions, exactly the Julia layout the reference works on.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "gla_hip.h"

namespace gla {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Thread-local error text returned by gla_last_error_string().
void set_error(const char *fmt, ...);

#define GLA_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            ::gla::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,              \
                             hipGetErrorString(_e));                                   \
            return GLA_ERR_HIP;                                                        \
        }                                                                              \
    } while (0)

#define GLA_LAUNCH_CHECK()                                                             \
    do {                                                                               \
        hipError_t _e = hipGetLastError();                                             \
        if (_e != hipSuccess) {                                                        \
            ::gla::set_error("%s:%d launch failed: %s", __FILE__, __LINE__,            \
                             hipGetErrorString(_e));                                   \
            return GLA_ERR_HIP;                                                        \
        }                                                                              \
    } while (0)

// v from lane K of each 16-lane row to the whole row (DPP row_newbcast:K).
// Every lane reads a valid source, so no "old" value needs setting up; a
// 64-bit value moves in one v_mov_b64_dpp (gfx950 has DPP on 64-bit moves
// for row_newbcast).
template <int K, typename T>
__device__ __forceinline__ T row_bcast(T v)
{
    static_assert(K >= 0 && K < 16, "lane of a 16-lane row");
    return __builtin_amdgcn_update_dpp(v, v, 0x150 + K, 0xF, 0xF, true);
}

// v from the lane DPP control CTRL selects within its row (bound_ctrl off:
// a source outside the row reads 0); 64-bit values move as two halves.
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

// Sum over the 16 lanes of each row, the same bits in every lane of the row:
// pairs, quads, half rows, rows (each step adds two equal-order partials).
template <typename T>
__device__ __forceinline__ T row_sum16(T v)
{
    v += dppmov<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dppmov<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dppmov<0x141>(v);   // row_half_mirror
    v += dppmov<0x140>(v);   // row_mirror
    return v;
}

// Sum over the 4 lanes of each quad, the same bits in every lane of it.
template <typename T>
__device__ __forceinline__ T quad_sum(T v)
{
    v += dppmov<0xB1>(v);
    v += dppmov<0x4E>(v);
    return v;
}

// v of lane `src` (wave-uniform) in every lane.
template <typename T>
__device__ __forceinline__ T lane_value(T v, int src)
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

// Sum over the 64 lanes (all active), the same value in every lane and a
// fixed order: 16-lane row sums by DPP, then the four rows.
template <typename T>
__device__ __forceinline__ T wave_sum_dpp(T v)
{
    v = row_sum16(v);
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// Fused multiply-add in T's own precision.
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Forward elimination steps of a 16×16 lower-triangular inverse in the
// four-column layout (lane (i, g) holds w[t] = V[i][g + 4t]): for K = 0…14,
// V[i] −= m[K]·V[K] with row K of V broadcast from lane K of each 16-lane
// row; m[K] is this lane's multiplier (zero for i ≤ K), so every step is one
// broadcast and one unconditional FMA per column.
template <int K, typename T>
struct TileInvSteps {
    static __device__ __forceinline__ void run(const T m[16], T w[4])
    {
#pragma unroll
        for (int t = 0; t < 4; ++t) w[t] = fma_t(-m[K], row_bcast<K>(w[t]), w[t]);
        TileInvSteps<K + 1, T>::run(m, w);
    }
};
template <typename T>
struct TileInvSteps<15, T> {   // row 15 has no rows below it
    static __device__ __forceinline__ void run(const T *, T *) {}
};

// The same backward, for an upper-triangular inverse: K = 15…1,
// V[i] −= m[K]·V[K] (m[K] zero for i ≥ K).
template <int K, typename T>
struct TileInvStepsUp {
    static __device__ __forceinline__ void run(const T m[16], T w[4])
    {
#pragma unroll
        for (int t = 0; t < 4; ++t) w[t] = fma_t(-m[K], row_bcast<K>(w[t]), w[t]);
        TileInvStepsUp<K - 1, T>::run(m, w);
    }
};
template <typename T>
struct TileInvStepsUp<0, T> {   // row 0 has no rows above it
    static __device__ __forceinline__ void run(const T *, T *) {}
};

// s = √a and r = 1/√a from the hardware reciprocal square root refined by
// Newton steps (two for FP64, one for FP32) and one correction of s = a·r:
// both within an ulp or two of sqrt(a) and 1/sqrt(a), at a fraction of the
// latency of a correctly rounded square root followed by a division, which
// sit on the serial path of every column.  a = 0 gives s = 0, r = ∞ and a
// negative or NaN a gives NaN, as sqrt and 1/sqrt do.
__device__ __forceinline__ void sqrt_and_rsqrt(double a, double &s, double &r)
{
    double y = __builtin_amdgcn_rsq(a);
    const double h = 0.5 * a;
    double e = __builtin_fma(-h * y, y, 0.5);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-h * y, y, 0.5);
    y = __builtin_fma(y, e, y);
    double q = a * y;
    q = __builtin_fma(__builtin_fma(-q, q, a), 0.5 * y, q);
    const bool zero = a == 0.0;
    s = zero ? 0.0 : q;
    r = zero ? __builtin_inf() : y;
}
__device__ __forceinline__ void sqrt_and_rsqrt(float a, float &s, float &r)
{
    float y = __builtin_amdgcn_rsqf(a);
    const float h = 0.5f * a;
    const float e = __builtin_fmaf(-h * y, y, 0.5f);
    y = __builtin_fmaf(y, e, y);
    float q = a * y;
    q = __builtin_fmaf(__builtin_fmaf(-q, q, a), 0.5f * y, q);
    const bool zero = a == 0.0f;
    s = zero ? 0.0f : q;
    r = zero ? __builtin_inff() : y;
}

// Row layout of a 16×16 tile: lane (i = l % 16, g = l / 16) holds all of row
// i in a[0..15], the same in each of the four 16-lane groups, so that every
// value a column step needs moves by a DPP broadcast within the lane's own
// 16-lane row (a few cycles) rather than through the LDS crossbar.

// Column C's rank-1 update of the entries a[J..15]: a[j] -= l_i l_j, with
// l_j = L[j][C] broadcast from lane j.  Entries right of the diagonal (j > i)
// take meaningless values and are cleared by the caller.
template <int C, int J, typename T>
struct TileColUpdate {
    static __device__ __forceinline__ void run(T a[16], T l)
    {
        const T lj = row_bcast<J>(l);
        if constexpr (sizeof(T) == 8)
            a[J] = __builtin_fma(-l, lj, a[J]);
        else
            a[J] = __builtin_fmaf(-l, lj, a[J]);
        TileColUpdate<C, J + 1, T>::run(a, l);
    }
};
template <int C, typename T>
struct TileColUpdate<C, 16, T> {
    static __device__ __forceinline__ void run(T *, T) {}
};

// Factors the symmetric tile held in row layout in place, right-looking,
// column by column as cholUnblocked! does (src/cholesky.jl:3-15): sqrt of
// the pivot, scale by its reciprocal, rank-1 update (the update products are
// (a_ic r)(a_jc r), each factor rounded once, as in the scaled column).
// rd[c] receives 1/L[c][c].  The first pivot whose square root is taken of
// a negative number, among the first `nvalid`, is recorded in *info as
// gbase + c + 1.  Entries right of the diagonal are left meaningless.
// The 16 column steps are one branch-free block (the first negative pivot
// is only noted in `fneg`, its column or 16 if none, and reported after the
// tile), so that the compiler can overlap the trailing updates of column c
// with the square root of c + 1.
template <int C, typename T>
struct TileFactorCols {
    static __device__ __forceinline__ void run(T a[16], T rd[16], int i, int &fneg)
    {
        const T d = row_bcast<C>(a[C]);
        fneg = (d < T(0) && fneg == 16) ? C : fneg;
        T s, r;
        sqrt_and_rsqrt(d, s, r);
        rd[C] = r;
        const T l = a[C] * r;
        a[C] = i == C ? s : l;
        TileColUpdate<C, C + 1, T>::run(a, l);
        TileFactorCols<C + 1, T>::run(a, rd, i, fneg);
    }
};
template <typename T>
struct TileFactorCols<16, T> {
    static __device__ __forceinline__ void run(T *, T *, int, int &) {}
};

// 16×16×4 MFMA with one A and one B element per lane; accumulator element r
// of lane l sits at (row(l, r), l % 16) of the 16×16 tile.
template <typename T>
struct PanelMfma;
template <>
struct PanelMfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct PanelMfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};


// ------------------------------------------------------------------ handle --
// Workspace slots carved out of one device allocation that only grows.
struct Handle {
    int device = 0;
    hipStream_t side = nullptr;       // second stream for look-ahead overlap
    hipEvent_t ev_fork = nullptr;     // main -> side
    hipEvent_t ev_join = nullptr;     // side -> main
    hipEvent_t ev_step[2] = {nullptr, nullptr};
    int *dinfo = nullptr;             // device: first failing pivot (1-based), 0 = ok
    unsigned *dbar = nullptr;         // device: grid-barrier counters
    void *ws = nullptr;               // device workspace
    size_t ws_bytes = 0;
    int num_cus = 256;
    // Tuning: block width of the look-ahead driver behind cholRecursive!
    // (0: the recursive-halving driver; −1: chosen from n).  Set with
    // gla_set_tuning.
    int64_t chol_block = -1;
    // Solve the look-ahead Cholesky's panels in one launch each (1) or by the
    // recursive rdiv! (0).
    int chol_psolve = 1;
    // Release the look-ahead Cholesky's side-stream SYRK after the
    // next-block-column update (1) rather than right after the panel solve.
    int chol_syrk_late = 1;
    // Step width of the LDLᵀ driver (128: 128-wide steps; a larger multiple
    // of 128: that width with depth-1 look-ahead).
    int64_t ldlt_block = 512;
    // Block width of the mixed-precision factorisation's look-ahead driver.
    int mixed_block = 512;
    // QR panel kernel: 1 LDS-reduction form, 2 wave-butterfly form.
    int qr_panel = 2;
    // Width of the sub-panels a QR block column is factored in (0: the whole
    // block column in one panel launch; 32 or 64; -1: by the column height).
    int qr_inner = -1;
    // Depth-1 look-ahead in qrBlocked! (trailing update on the side stream).
    int qr_lookahead = 1;
    // Block columns at least this tall are factored by the
    // reconstructed-Householder panel (0: never; -1: any height for block
    // columns at least 64 wide, 2048 rows for narrower ones).
    int64_t qr_hr = -1;
    // Polls before a grid-wide wait of the QR panel gives up (the call then
    // returns GLA_ERR_TIMEOUT).
    unsigned spin_limit = 1u << 25;
    // Test setting: every grid-wide wait of the QR panel targets an arrival
    // count that never comes, so each gives up after spin_limit polls.
    int qr_wait_test = 0;
    // CUs kept free of the look-ahead Cholesky's trailing SYRK (0: none;
    // a new handle takes num_cus/8).  When set, the driver runs its chain on
    // `chain` (all CUs) and the SYRK on `masked` (the other CUs), both
    // forked from the caller's stream.
    int side_cu_reserve = 0;
    hipStream_t chain = nullptr;
    hipStream_t masked = nullptr;
};

// Ensure the workspace holds at least `bytes` (grows with hipMalloc, which
// synchronises the device; the first call of a given size pays for it).
int ensure_workspace(Handle *h, size_t bytes);

// ------------------------------------------------------------------ GEMM --
// C = alpha*op(A)*op(B) + beta*C  with op(A) M×K, op(B) K×N.
// mode 0: every element; mode 1: only elements with (row + diag_off) >= col
// ("lower" relative to a diagonal), tiles wholly above it are skipped.
// ta/tb: 'N' or 'T'.  Supported (ta,tb,mode): (N,N,0) (N,T,0) (T,N,0) (N,T,1)
// (N,T,2).  kfrom_row = 1: op(A)(i,k) = 0 for k < i (an upper-triangular
// op(A)), so each tile contracts k from its first row on (no split-K).
template <typename T>
int gemm(hipStream_t s, char ta, char tb, int mode, int64_t M, int64_t N, int64_t K, T alpha,
         const T *A, int64_t lda, const T *B, int64_t ldb, T beta, T *C, int64_t ldc,
         int64_t diag_off = 0, T *ws = nullptr, size_t ws_elems = 0, int kfrom_row = 0);
// Selects the FP64 direct-to-LDS GEMM kernel (0: off).
void set_gemm_dl(int v);
void set_gemm_lds1(int v);
int get_gemm_dl();
int get_gemm_lds1();
// The same over a strided batch (grid z): product b reads A + b·sA, B + b·sB
// and writes C + b·sC; no split-K.
template <typename T>
int gemm_batched(hipStream_t s, char ta, char tb, int mode, int64_t M, int64_t N, int64_t K,
                 T alpha, const T *A, int64_t lda, int64_t sA, const T *B, int64_t ldb,
                 int64_t sB, T beta, T *C, int64_t ldc, int64_t sC, int batch, int kfrom_row = 0);

// ------------------------------------------------------------- Cholesky --
// Linv: scratch for the 16×16 diagonal-tile inverses (tile_inverse_elems(n)
// elements); null takes it from the handle's workspace.
template <typename T>
int chol_device(Handle *h, hipStream_t s, int64_t n, T *A, int64_t lda, int algo,
                int64_t param, T *Linv = nullptr);

// Cholesky building blocks (chol.hip): factor one ≤128 diagonal block in
// place and store the inverses of its 16×16 diagonal tiles; solve rows
// below it (B <- B L⁻ᵀ) with those inverses; inverses of all 16×16 diagonal
// tiles of an n×n lower L; the recursive right-side solve B <- B L⁻ᵀ.
// LDLᴴ (_ldlt_lower!, src/ldlt.jl:16-33) of the nb×nb (nb ≤ 128) block at
// A in one workgroup: D on the diagonal, the unit factor's L below it, and
// the inverses of its unit 16×16 diagonal tiles in LinvG.
template <typename T>
int ldlt_factor_block(hipStream_t s, int nb, T *A, int64_t lda, T *LinvG);
template <typename T>
int chol_factor_block(hipStream_t s, int nb, T *A, int64_t lda, int *info, int64_t info_base,
                      T *LinvG);
// The same for a block classified on the device: flags[0] != 0 — skip;
// flags[1] == 0 — A is within 0.1·√eps of I (Frobenius), factored to
// first order (L = I + strictly-lower(A) + (diag(A) − 1)/2); otherwise the
// full factorisation.
template <typename T>
int chol_factor_block_near_identity(hipStream_t s, int nb, T *A, int64_t lda, int *info,
                                    int64_t info_base, T *LinvG, const int *flags);
// The diagonal-block factorisation of an FP32 block (nb ≤ 128) computed in
// FP64: read and written in FP32, FP32 tile inverses in Linv32.
int chol_factor_block_f32io(hipStream_t s, int nb, float *A, int64_t lda, int *info,
                            int64_t info_base, float *Linv32);
template <typename T>
int chol_solve_block(hipStream_t s, int nb, const T *L, int64_t ldl, const T *LinvG, T *B,
                     int64_t ldb, int64_t nrows);
template <typename T>
int tile_inverses(hipStream_t s, int64_t n, const T *L, int64_t ldl, T *LinvG);
// Strided-batch forms for matrices of order ≤ 128 (one diagonal block):
// factorisation with per-matrix info, 16×16 tile inverses, and B <- B L⁻ᵀ.
template <typename T>
int chol_factor_batched(hipStream_t s, int nb, T *A, int64_t lda, int64_t sA, int *info, T *LinvG,
                        int64_t sLinv, int64_t batch);
template <typename T>
int tile_inverses_batched(hipStream_t s, int64_t n, const T *L, int64_t ldl, int64_t sL, T *LinvG,
                          int64_t sLinv, int64_t batch);
template <typename T>
int trsm_rlt_batched(hipStream_t s, int64_t m, int n, const T *L, int64_t ldl, int64_t sL,
                     const T *Linv, int64_t sLinv, T *B, int64_t ldb, int64_t sB, int64_t batch);
template <typename T>
int trsm_rlt(hipStream_t s, int64_t m, int64_t n, const T *L, int64_t ldl, const T *Linv, T *B,
             int64_t ldb);
size_t tile_inverse_elems(int64_t n);
// Dt_b = L_bb⁻ᵀ for every 128×128 diagonal block b of L (column-major
// 128×128 at Dt + b·128²), from the tile inverses Linv.
template <typename T>
int diag_block_inverses(hipStream_t s, int64_t n, const T *L, int64_t ldl, const T *Linv, T *Dt);

// Multi-right-hand-side lower triangular solve (trsv.hip): Y = L⁻¹ R for up
// to 16 right-hand sides held row-major (element (i, k) at [16 i + k]),
// blocked by 128 rows, one launch.  Dt from diag_block_inverses, Mt from
// diag_sub (npad × 128, npad = n rounded up to 128).  rev:
// row i of the solve reads R row n−1−i and adds into X row n−1−i (the
// backward solve Lᵀ z = y run as a lower solve with the reversed matrix
// J Lᵀ J).  If X is given, X[·, k] += Y[·, k] for k < nrhs (FP64,
// column-major).  Y is scratch of the same layout; flags: one int per
// 128-row block, values below `epoch`; err is raised if a wait times out.
template <typename T>
int trsv16(Handle *h, hipStream_t s, int64_t n, const T *L, int64_t ldl, const T *Dt, const T *Mt,
           const T *R, T *Y, int rev, double *X, int64_t ldx, int nrhs, int *flags, int epoch,
           int *err);
// Mt (npad × 128) = L_bb⁻¹ L_{b,b−1} stacked over the 128-row blocks b ≥ 1.
template <typename T>
int diag_sub(hipStream_t s, int64_t n, const T *L, int64_t ldl, const T *Dt, T *Mt);
// R (row-major n×16, type T) = B − sym(A) X for the lower-stored FP64 A and
// nrhs ≤ 16 columns of X, B (column-major); Xt: symv_scratch(n) doubles;
// nrm[k] += ‖r_k‖², nrm[16 + k] += ‖x_k‖² (nrm zeroed by the caller).
template <typename T>
int symv_residual(hipStream_t s, int64_t n, const double *A, int64_t lda, const double *X,
                  int64_t ldx, const double *B, int64_t ldb, int nrhs, double *Xt, T *R,
                  double *nrm);
size_t symv_scratch(int64_t n);

// QR building blocks (qr.hip): T of the first k reflectors of F (m×k), and
// A (mA×nA) <- op(Q) A, trans 'N' or 'T', Q from the first b reflectors.
size_t qr_tblock_scratch(int64_t m, int64_t k);
size_t qr_apply_scratch(int64_t mA, int64_t nA, int64_t b);
template <typename T>
int qr_tblock_dev(hipStream_t s, int64_t m, int64_t k, const T *F, int64_t ldf, const T *tau,
                  T *Tm, int64_t ldt, T *scratch);
template <typename T>
int qr_apply_dev(hipStream_t s, char trans, int64_t mA, int64_t nA, int64_t b, const T *V,
                 int64_t ldv, const T *Tm, int64_t ldt, T *A, int64_t lda, T *scratch);

// Small helpers used across units.
__global__ void k_set_int(int *p, int v);

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace gla
