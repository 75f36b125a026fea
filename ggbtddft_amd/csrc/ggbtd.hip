// ggbtd.hip -- gfx950 kernels and C ABI of the real-time TDDFT propagator.
//
// One time step (t_n -> t_n + dt) is the palindrome
//     Vh . Le . Lo . K . Lo . Le . Vh         (right-most acts first)
// on psi[l][c][i] (complex, u-units); see include/ggbtd.h for layouts.
//   Vh     theta-space phase: Psi = B psi, Psi *= mask e^{-i dV dt/2}, psi = B^T Psi
//   Le/Lo  laser r E cos(theta) for dt/2 as 2x2 rotations on (l, l+1) blocks
//   K      per-l exp(-i H_l dt) = Q diag(e^{-i lam dt}) Q^T on the retained
//          eigenpairs (the reference's per-l cublasZgemmBatched,
//          h_timePropagation.cuh:17-26)
// followed by the potential rebuild from psi_{n+1} (explicit update as in
// h_timePropagation.cuh:97-101) and the observables (dipoleMoment.cuh,
// dipoleAcceleration.cuh) as fused reductions.
//
// Kernels, launch order per step:
//   k_angular<mode 0>   Vh, Le, Lo                       grid (N / 8 NT, C)
//   k_kin_fwd(_tiled)   Y = e^{-i lam dt} Q^T psi        grid (nk/16, 2C/16, n_l) or 64 x 128 tiles
//   k_kin_bwd(_tiled)   psi = Q Y                        grid (N/16,  2C/16, n_l) or 64 x 128 tiles
//   k_angular<mode 1>   Lo, Le, Vh, |B psi|^2, observable partials
//   k_rho               rho_u = sum occ |Psi|^2, v_xc; observable sums
//   k_tile_gemm<Qmom>   multipole moments q_L = P_L^T rho_u
//   k_hartree           V_L by prefix/suffix affine scans  grid (n_har, n_inst)
//   k_tile_gemm<Dv>     dV = sum_L V_L P_L, v_xc - v_gs, phase factors of the next step
//
// The matrix products run on the matrix cores: v_mfma_f64_16x16x4f64, and
// v_mfma_f32_16x16x4f32 for the FP32 products of the mixed form.
// Lane layouts (checked on the device by ggb_selftest_mfma):
//   A (16x4):  lane L holds A[L % 16][L / 16]
//   B (4x16):  lane L holds B[L / 16][L % 16]
//   D (16x16): accumulator item r of lane L holds D[row][L % 16] with
//              row = 4 r + L / 16 for v_mfma_f64_16x16x4f64 (measured on gfx950)
//              row = 4 (L / 16) + r for v_mfma_f32_16x16x4f32
//   (mfma_t<T>::row).

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ggbtd.h"

typedef double d4_t __attribute__((ext_vector_type(4)));

#define WG 256
#define RT 8          // radial nodes per angular tile (16 real columns)
#ifndef ANG_NT_MIXED
#define ANG_NT_MIXED 2  // angular column tiles per work-group in the mixed form (1, 2 or 4)
#endif
#define RHO_MIN 1e-40 // v_xc = 0 at or below this density
#define KIN_NARROW_MAX_C 8  // up to 8 complex columns: one 16-wide MFMA column tile
#define TIMING_STRIDE 16    // timing mode 1 samples every 16th step

static __device__ __forceinline__ d4_t mfma64(double a, double b, d4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

typedef float f4_t __attribute__((ext_vector_type(4)));

// Matrix-core product D += A B of one 16x16x4 tile in the compute type T
// (FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32; same lane layout).
// mixed: FP32 products on an FP64 state that is never rounded to FP32.  The
// matrix-core products only see small quantities: the potential half step is
// applied as psi += B^T ((f - 1) B psi), and the kinetic propagator acts on
// the remainder psi - alpha phi of each orbital after its ground-state
// component alpha phi (an eigenvector of exp(-i H_l dt), advanced exactly in
// FP64 as alpha e^{-i eps dt} phi).  The operators are those of the FP64 form.
template <typename T> struct mfma_t;
template <> struct mfma_t<double> {
    typedef d4_t v4;
    static constexpr bool mixed = false;
    static __device__ __forceinline__ v4 op(double a, double b, v4 c) { return mfma64(a, b, c); }
    // accumulator item r of lane L holds D[row(L / 16, r)][L % 16]
    static __device__ __forceinline__ int row(int kr, int r) { return 4 * r + kr; }
};
template <> struct mfma_t<float> {
    typedef f4_t v4;
    static constexpr bool mixed = true;
    static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kr, int r) { return 4 * kr + r; }
};

// Work-groups of a launch are dispatched round-robin over the 8 XCDs, each
// with its own L2.  xcd_linear renumbers them so that consecutive logical
// work-groups run on the same XCD and share its L2 (a bijection for any grid).
static __device__ __forceinline__ int xcd_linear() {
    const int gx = gridDim.x, gy = gridDim.y;
    const int T = gx * gy * gridDim.z;
    const int p = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int q = T / 8, r = T % 8, x = p % 8;
    return x * q + min(x, r) + p / 8;
}
// the value of lane L ^ 1 (the re/im partner of a real column pair) through
// a DPP quad permutation [1,0,3,2] in the VALU, where __shfl_xor takes an LDS
// round trip (ds_bpermute) per 32 bits; the same value.  Every lane of the
// wave must be active.
static __device__ __forceinline__ float lane_xor1(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
}
static __device__ __forceinline__ double lane_xor1(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, 0xB1, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
struct tile3 {
    int x, y, z;
};
// logical (x, y, z) of this work-group, x fastest
static __device__ __forceinline__ tile3 xcd_tile() {
    const int g = xcd_linear();
    tile3 t;
    t.x = g % gridDim.x;
    t.y = (g / gridDim.x) % gridDim.y;
    t.z = g / (gridDim.x * gridDim.y);
    return t;
}


// component p of e^{-i lam dt} (v_re + i v_im), from this lane's v and its
// partner's vo; f = (cos(lam dt), sin(lam dt)) from the table built at upload
template <typename T>
static __device__ __forceinline__ T kin_factor(double v, double vo, int p, double2 f) {
    const double re = p ? vo : v, im = p ? v : vo;
    return (T)(p ? (f.x * im - f.y * re) : (f.x * re + f.y * im));
}

// Ground-state reference of the mixed form: for column c (orbital k = c mod
// n_orb, block l0[k]) alpha = <phi_k|psi_l0[c]> and beta = alpha e^{-i eps_k dt},
// FP64, fixed-order reduction.  ab[c] = (alpha_re, alpha_im, beta_re, beta_im).
struct KinRef {
    const double *phi;   // [n_orb][N]
    const int32_t *l0;   // [n_orb]
    const double4 *ab;   // [C]
    const int32_t *active;  // [n_inst] instances in use, or null (all)
};

// whether any instance with a column among real columns [n0, n0 + nreal) is
// in use (work-group-uniform; idle serving slots skip their kinetic tiles)
static __device__ __forceinline__ bool kin_cols_active(const KinRef &kr, int n0, int nreal, int C,
                                                       int n_orb) {
    if (!kr.active) return true;
    const int c0 = n0 >> 1, c1 = min(C, (n0 + nreal + 1) >> 1);
    for (int q = c0 / n_orb; q * n_orb < c1; ++q)
        if (kr.active[q]) return true;
    return false;
}
__global__ __launch_bounds__(WG) void k_kin_ref(const double *__restrict__ psi,
                                                const double *__restrict__ phi,
                                                const int32_t *__restrict__ l0,
                                                const double *__restrict__ eps, double4 *ab,
                                                int N, int C, int n_orb, double dt) {
    __shared__ double red[2][WG / 64];
    const int c = blockIdx.x, k = c % n_orb;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const double *ph = phi + (size_t)k * N;
    const double2 *x = (const double2 *)psi + ((size_t)l0[k] * C + c) * N;
    double sr = 0.0, si = 0.0;
    for (int i = t; i < N; i += WG) {
        const double2 v = x[i];
        sr += ph[i] * v.x;
        si += ph[i] * v.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_down(sr, off);
        si += __shfl_down(si, off);
    }
    if (lane == 0) {
        red[0][wave] = sr;
        red[1][wave] = si;
    }
    __syncthreads();
    if (t == 0) {
        double ar = 0.0, ai = 0.0;
        for (int w = 0; w < WG / 64; ++w) {
            ar += red[0][w];
            ai += red[1][w];
        }
        double sn, cs;
        sincos(eps[k] * dt, &sn, &cs);
        ab[c] = make_double4(ar, ai, cs * ar + sn * ai, cs * ai - sn * ar);
    }
}

// FP32 copy of the state for the tiled pass 1, the ground-state component
// alpha phi taken out (mixed form): psi32[l][c][i] = psi - [l = l0] alpha phi.
// only_l0: just the ground-state block of each column (the other blocks were
// written by the angular kernel's first half, which ends every step's input)
__global__ __launch_bounds__(WG) void k_kin_prep(const double2 *__restrict__ psi, KinRef ref,
                                                 float2 *__restrict__ psi32, int N, int C,
                                                 int n_orb, int only_l0) {
    const int i = blockIdx.x * WG + threadIdx.x;
    const int c = blockIdx.y;
    const int k = c % n_orb;
    const int l = only_l0 ? ref.l0[k] : blockIdx.z;
    if (i >= N) return;
    const size_t o = ((size_t)l * C + c) * N + i;
    double2 v = psi[o];
    if (ref.phi && ref.l0[k] == l) {
        const double4 a = ref.ab[c];
        const double f = ref.phi[(size_t)k * N + i];
        v.x -= a.x * f;
        v.y -= a.y * f;
    }
    psi32[o] = make_float2((float)v.x, (float)v.y);
}

// reference component of lane column (c, p) in block l: coefficient (alpha or
// beta, component p) and phi row, or a null row
static __device__ __forceinline__ const double *kin_ref_row(const KinRef &kr, int l, int c, int p,
                                                            int N, int n_orb, bool beta,
                                                            double *coef) {
    const int k = c % n_orb;
    if (!kr.phi || kr.l0[k] != l) return nullptr;
    const double4 a = kr.ab[c];
    *coef = beta ? (p ? a.w : a.z) : (p ? a.y : a.x);
    return kr.phi + (size_t)k * N;
}

// ---------------------------------------------------------------------------
// Kinetic propagator, pass 1: Y[l][c][p][a] = (e^{-i lam_a dt} sum_i Q[l][i][a] psi[l][c][i])_p
// M = eigen index a, N = real column (c, p = re/im), K = radial node i.
// One work-group computes 16 TA x 16 outputs; its NW waves take equal shares
// of K (more loads in flight per CU than one wave per tile), then the partial
// tiles are added in fixed order through LDS.  Y is planar in (re, im)
// so that pass 2 reads contiguous runs of a.  T is the compute type (tables
// and Y in T, state in FP64, phase in FP64).
// ---------------------------------------------------------------------------
#ifndef KIN_KU
#define KIN_KU 2
#endif
#define KU KIN_KU
// waves per work-group of the narrow passes (K split over them): 4 in FP64, 8
// in the mixed form, whose FP32 pass 1 takes 17.8 instead of 19.7 us at argon
// size with 8 (FP64: 22.6 against 22.1 us)
#ifndef KIN_NW
#define KIN_NW 4
#endif
#ifndef KIN_NW32
#define KIN_NW32 8
#endif
// waves 0..3 finish accumulator rows 0..3 of every tile after the K split
static_assert(KIN_NW >= 4 && KIN_NW <= 16 && KIN_NW32 >= 4 && KIN_NW32 <= 16, "4 to 16 waves per narrow work-group");
template <typename T>
struct kin_nw {
    static constexpr int value = mfma_t<T>::mixed ? KIN_NW32 : KIN_NW;
};
#define KIN_WG_MAX (64 * (KIN_NW > KIN_NW32 ? KIN_NW : KIN_NW32))
// Panel form: one work-group takes the 16 TA eigen rows a0 .. a0 + 16 TA - 1 as
// TA tiles with interleaved rows (row ar of tile t is a = a0 + TA ar + t).  A
// lane reads its TA table values of a K row as one vector, the 16 lanes of the
// row read one run of 16 TA values, and each state load feeds TA products.
// Every tile keeps the K split, the product order and the fixed-order LDS sum
// of one tile per work-group, so the result does not depend on TA.  TA = 2
// takes an argon FP64 pass from 25.4 to 22.9 us (FP32 22.8 to 20.7); TA = 4
// is no faster in FP64 and slower in FP32.
#ifndef KIN_TA
#define KIN_TA 2
#endif
static_assert(KIN_TA == 1 || KIN_TA == 2 || KIN_TA == 4, "KIN_TA: 1, 2 or 4 (a divisor of 16)");
template <typename T, int TA>
__global__ __launch_bounds__(KIN_WG_MAX) void k_kin_fwd(const T *__restrict__ Q,
                                                const double2 *__restrict__ kph,
                                                const int32_t *__restrict__ nkeep,
                                                const double *__restrict__ psi, KinRef ref,
                                                T *__restrict__ Y, int N, int nk_max, int C,
                                                int n_orb) {
    typedef typename mfma_t<T>::v4 v4;
    typedef T vta __attribute__((ext_vector_type(TA)));
    constexpr int NW = kin_nw<T>::value;
    __shared__ double red[NW][TA][64][4];
    const tile3 tb = xcd_tile();  // (eigen panel, column tile, l): panels of one l share psi
    const int l = tb.z;
    const int nk = nkeep[l];
    const int a0 = tb.x * 16 * TA;
    if (a0 >= nk) return;  // block-uniform
    if (!kin_cols_active(ref, tb.y * 16, 16, C, n_orb)) return;  // idle instances only
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int col = tb.y * 16 + lc;
    const int c = col >> 1, p = col & 1;
    const bool colok = c < C;
    const T *Ql = Q + (size_t)l * N * nk_max + a0 + TA * lc;
    const bool lv = a0 + TA * lc < nk;  // interleaved rows: this lane's TA rows hold some a < nkeep
    const double *Xl = psi + ((size_t)l * C + (colok ? c : 0)) * N * 2 + p;
    const int kq = (((N + NW - 1) / NW + 3) / 4) * 4;  // K share of a wave, multiple of 4
    const int kb = wave * kq, ke = min(N, kb + kq);
    double rc = 0.0;
    const double *rrow = (mfma_t<T>::mixed && colok) ? kin_ref_row(ref, l, c, p, N, n_orb, false, &rc) : nullptr;
    // wave w < 4 finishes accumulator row w of every tile: its phase factors
    // are loaded ahead of the products
    const int ar = mfma_t<T>::row(kr, wave & 3);
    double2 kp[TA];
    if (wave < 4) {
#pragma unroll
        for (int t = 0; t < TA; ++t) {
            const int a = a0 + TA * ar + t;
            kp[t] = kph[(size_t)l * nk_max + (a < nk ? a : 0)];
        }
    }
    v4 acc[TA];
#pragma unroll
    for (int t = 0; t < TA; ++t) acc[t] = v4{0, 0, 0, 0};
    for (int k0 = kb; k0 < ke; k0 += 4 * KU) {
        T av[KU][TA], bv[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int i = k0 + 4 * u + kr;
            const bool ok = i < ke;
#pragma unroll
            for (int t = 0; t < TA; ++t) av[u][t] = (T)0;
            if (ok && lv) {
                const vta q = *reinterpret_cast<const vta *>(Ql + (size_t)i * nk_max);
#pragma unroll
                for (int t = 0; t < TA; ++t) av[u][t] = q[t];
            }
            double x = (ok && colok) ? Xl[(size_t)i * 2] : 0.0;
            if (rrow && ok) x -= rc * rrow[i];
            bv[u] = (T)x;
        }
#pragma unroll
        for (int u = 0; u < KU; ++u)
#pragma unroll
            for (int t = 0; t < TA; ++t) acc[t] = mfma_t<T>::op(av[u][t], bv[u], acc[t]);
    }
    // partial tiles through LDS (wave 0's in the last slot); wave w < 4 then finishes
    // accumulator row r = w of every tile, so the four epilogues run side by side
#pragma unroll
    for (int t = 0; t < TA; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wave > 0 ? wave - 1 : NW - 1][t][lane][r] = (double)acc[t][r];
    __syncthreads();
    if (wave >= 4) return;
    const int r = wave;
#pragma unroll
    for (int t = 0; t < TA; ++t) {
        double v = red[NW - 1][t][lane][r];
        for (int w = 0; w < NW - 1; ++w) v += red[w][t][lane][r];
        const double vo = lane_xor1(v);
        const int a = a0 + TA * ar + t;
        const T o = kin_factor<T>(v, vo, p, kp[t]);
        if (colok && a < nk) Y[(((size_t)l * C + c) * 2 + p) * nk_max + a] = o;
    }
}

// element (a, i) of block l of the quad-interleaved transposed table of the
// narrow pass 2: ((a / 4) N + i) 4 + a % 4 within the block
static __host__ __device__ __forceinline__ size_t kin_qt_index(size_t l, size_t i, size_t a, size_t N, size_t nk_max) {
    return l * nk_max * N + ((a / 4) * N + i) * 4 + a % 4;
}

// Kinetic propagator, pass 2: psi[l][c][i] = sum_a QT[l][a][i] Y[l][c][a].
// One work-group per 16 x 16 output tile, K = a split over the NW waves as in
// pass 1.  Within a wave the K order is permuted (lane kr of MFMA j in chunk u
// takes a = k0 + 16u + 4kr + j) so that each lane reads 4 consecutive a of Q
// and of planar Y as one vector.  Q and Y are zero beyond nkeep up to nk_max
// (host padding, zeroed Y), so the loop runs to nkeep rounded up to 16
// without per-element guards; rows i >= N and columns c >= C read row/column
// 0 and are not stored.
#ifndef KIN_KU2
#define KIN_KU2 2
#endif
#define KU2 KIN_KU2
template <typename T>
__global__ __launch_bounds__(KIN_WG_MAX) void k_kin_bwd(const T *__restrict__ QT,
                                                const int32_t *__restrict__ nkeep,
                                                const T *__restrict__ Y, KinRef ref,
                                                double *__restrict__ psi, int N, int nk_max,
                                                int C, int n_orb) {
    typedef typename mfma_t<T>::v4 v4;
    constexpr int NW = kin_nw<T>::value;
    __shared__ double red[NW][64][4];
    const tile3 tb = xcd_tile();  // (radial tile, column tile, l): tiles of one l share Y
    const int l = tb.z;
    const int nk = nkeep[l];
    const int kend = min(nk_max, (nk + 15) & ~15);
    if (!kin_cols_active(ref, tb.y * 16, 16, C, n_orb)) return;  // idle instances only
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = tb.x * 16;
    const int kr = lane >> 4, lc = lane & 15;
    const int col = tb.y * 16 + lc;
    const int c = col >> 1, p = col & 1;
    const bool colok = c < C;
    const int im_ = i0 + lc;
    // the transposed table, quad-interleaved (kin_qt_index): a lane's four
    // consecutive a (one K group of its MFMAs) are one vector, and the 16 lanes
    // of a row read 16 consecutive i (a 256- or 512-byte run).  Against the
    // plain transpose (four scalar loads per K group) it takes 21.1 instead of
    // 24.2 us per FP64 argon pass (13.7 instead of 15.3 in FP32).
    const T *Qc = QT + kin_qt_index(l, im_ < N ? im_ : 0, 4 * kr, N, nk_max);
    const T *Yr = Y + (((size_t)l * C + (colok ? c : 0)) * 2 + p) * nk_max + 4 * kr;
    const int kq = (((kend + 15) / 16 + NW - 1) / NW) * 16;  // K share of a wave, multiple of 16
    const int kb = wave * kq, ke = min(kend, kb + kq);
    // wave w < 4 finishes accumulator row w: the mixed form's ground-state term
    // of that element is loaded ahead of the products
    const int io = i0 + mfma_t<T>::row(kr, wave & 3);
    double rc = 0.0, rv = 0.0;
    const double *rrow = (mfma_t<T>::mixed && colok && wave < 4) ? kin_ref_row(ref, l, c, p, N, n_orb, true, &rc) : nullptr;
    if (rrow) rv = rrow[io < N ? io : 0];
    v4 acc = {0, 0, 0, 0};
    for (int k0 = kb; k0 < ke; k0 += 16 * KU2) {
        v4 av[KU2], bv[KU2];
#pragma unroll
        for (int u = 0; u < KU2; ++u) {
            if (k0 + 16 * u < ke) {
                av[u] = *reinterpret_cast<const v4 *>(Qc + (size_t)((k0 + 16 * u) / 4) * N * 4);
                bv[u] = *reinterpret_cast<const v4 *>(Yr + k0 + 16 * u);
            } else {
                av[u] = (v4){0, 0, 0, 0};
                bv[u] = (v4){0, 0, 0, 0};
            }
        }
#pragma unroll
        for (int u = 0; u < KU2; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc = mfma_t<T>::op(av[u][j], bv[u][j], acc);
    }
    // partial tiles through LDS (wave 0's in the last slot); wave w < 4 then finishes
    // accumulator row r = w, so the four epilogues run side by side
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wave > 0 ? wave - 1 : NW - 1][lane][r] = (double)acc[r];
    __syncthreads();
    if (wave >= 4) return;
    double *Xl = psi + (size_t)l * C * N * 2;
    const int r = wave;
    double v = red[NW - 1][lane][r];
    for (int w = 0; w < NW - 1; ++w) v += red[w][lane][r];
    if (colok && io < N) {
        if (rrow) v += rc * rv;
        Xl[((size_t)c * N + io) * 2 + p] = v;
    }
}

// ---------------------------------------------------------------------------
// Tiled kinetic passes for many columns (instances sharing the tables).
// Work-group tile 64 (M) x 128 (real columns); wave w owns the 32 x 64
// sub-tile (w >> 1, w & 1) as 2 x 4 MFMA tiles.  K advances in chunks of
// TKT values (16, or 32 in FP32) staged in LDS, double buffered, the
// next chunk prefetched into registers while the current one is multiplied;
// one barrier per chunk.  Rows of the LDS tiles are padded by 16 values so
// that the four K rows a wave reads per product fall on distinct banks.
// ---------------------------------------------------------------------------
#define TM 64
#define TN 128
#ifndef KIN_DMA
#define KIN_DMA 1  // FP32 tiled pass 1 through direct global-to-LDS copies (k_kin_fwd_dma)
#endif
#define TMP (TM + 16)
#define TNP (TN + 16)

// 8 products of one K step of 4 for the wave's 2 x 4 tiles
template <typename T>
static __device__ __forceinline__ void kin_tile_step(const T (*As)[TMP], const T (*Bs)[TNP], int kk,
                                                     int wm, int wn, int lc,
                                                     typename mfma_t<T>::v4 (&acc)[2][4]) {
    const T a0 = As[kk][wm + lc], a1 = As[kk][wm + 16 + lc];
    T b[4];
#pragma unroll
    for (int y = 0; y < 4; ++y) b[y] = Bs[kk][wn + 16 * y + lc];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        acc[0][y] = mfma_t<T>::op(a0, b[y], acc[0][y]);
        acc[1][y] = mfma_t<T>::op(a1, b[y], acc[1][y]);
    }
}

template <typename T, int TKT>
__global__ __launch_bounds__(WG) void k_kin_fwd_tiled(const T *__restrict__ Q,
                                                      const double2 *__restrict__ kph,
                                                      const int32_t *__restrict__ nkeep,
                                                      const double *__restrict__ psi,
                                                      const float2 *__restrict__ psi32, KinRef ref,
                                                      T *__restrict__ Y, int N, int nk_max,
                                                      int C, int n_orb) {
    // FP32 (mixed): the state comes from psi32 (k_kin_prep), FP64: from psi
    typedef typename mfma_t<T>::v4 v4;
    constexpr int AV = TKT * TM / WG;  // A values per thread (consecutive a)
    constexpr int BV = TKT * 64 / WG;  // complex B values per thread (consecutive i)
    __shared__ T As[2][TKT][TMP];      // [k = i][m = a]
    __shared__ T Bs[2][TKT][TNP];      // [k = i][n = real column]
    const tile3 tb = xcd_tile();       // (eigen tile, column tile, l)
    const int l = tb.z;
    const int nk = nkeep[l];
    const int a0 = tb.x * TM;
    if (a0 >= nk) return;  // block-uniform
    const int n0 = tb.y * TN;
    if (!kin_cols_active(ref, n0, TN, C, n_orb)) return;  // idle instances only
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    const T *Ql = Q + (size_t)l * N * nk_max;
    const double2 *Xl = (const double2 *)psi + (size_t)l * C * N;
    // global -> register mapping
    const int ar = t / (TM / AV), am = (t % (TM / AV)) * AV;  // A: row i, AV consecutive a
    const int bc = t / (TKT / BV), bk = (t % (TKT / BV)) * BV;  // B: complex column, BV consecutive i
    const int cg = (n0 >> 1) + bc;
    const bool cok = cg < C;
    const bool aok = a0 + am < nk_max;  // nk_max is a multiple of 16 >= AV: all or none
    const float2 *Xs = psi32 + ((size_t)l * C + (cok ? cg : 0)) * N;
    T ra[AV], rb[2 * BV];
    auto load = [&](int k0) {
        const int i = k0 + ar;
        const bool ok = aok && i < N;
#pragma unroll
        for (int u = 0; u < AV; ++u) ra[u] = ok ? Ql[(size_t)i * nk_max + a0 + am + u] : (T)0;
#pragma unroll
        for (int u = 0; u < BV; ++u) {
            const int ii = k0 + bk + u;
            const bool okb = cok && ii < N;
            if (mfma_t<T>::mixed) {
                const float2 v = okb ? Xs[ii] : make_float2(0.f, 0.f);
                rb[2 * u] = (T)v.x;
                rb[2 * u + 1] = (T)v.y;
            } else {
                const double2 v = okb ? Xl[(size_t)cg * N + ii] : make_double2(0.0, 0.0);
                rb[2 * u] = (T)v.x;
                rb[2 * u + 1] = (T)v.y;
            }
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int u = 0; u < AV; ++u) As[buf][ar][am + u] = ra[u];
#pragma unroll
        for (int u = 0; u < BV; ++u) {
            Bs[buf][bk + u][2 * bc] = rb[2 * u];
            Bs[buf][bk + u][2 * bc + 1] = rb[2 * u + 1];
        }
    };
    v4 acc[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = (v4){0, 0, 0, 0};
    load(0);
    stash(0);
    __syncthreads();
    // two-level sum: each chunk's products in acc, added to hi after the chunk,
    // so no FP32 accumulation chain is longer than one chunk plus the chunk count
    v4 hi[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) hi[x][y] = (v4){0, 0, 0, 0};
    int buf = 0;
    for (int k0 = 0; k0 < N; k0 += TKT) {
        const bool more = k0 + TKT < N;
        if (more) load(k0 + TKT);
#pragma unroll
        for (int ks = 0; ks < TKT / 4; ++ks) kin_tile_step<T>(As[buf], Bs[buf], ks * 4 + kr, wm, wn, lc, acc);
        if (more) stash(buf ^ 1);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) hi[x][y] += acc[x][y], acc[x][y] = (v4){0, 0, 0, 0};
        __syncthreads();
        buf ^= 1;
    }
    const double2 *kphl = kph + (size_t)l * nk_max;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int col = n0 + wn + 16 * y + lc;
            const int c = col >> 1, p = col & 1;
            const int ab = a0 + wm + 16 * x;
            T o[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = ab + mfma_t<T>::row(kr, r);
                const double v = (double)hi[x][y][r];
                const double vo = lane_xor1(v);
                o[r] = kin_factor<T>(v, vo, p, kphl[a < nk ? a : 0]);
            }
            if (c < C) {
                T *Yo = Y + (((size_t)l * C + c) * 2 + p) * nk_max;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = ab + mfma_t<T>::row(kr, r);
                    if (a < nk) Yo[a] = o[r];
                }
            }
        }
    }
}

// FP32 tiled pass 1 with the tiles copied from global memory straight into
// LDS (global_load_lds_dwordx4: no staging registers, no LDS stores), three
// buffers and two K chunks in flight; one barrier per chunk.  It forms the
// products of k_kin_fwd_tiled<float, 16> in the same order (same chunks, same
// MFMA sequence, same two-level sum), so the stored Y is the same bit for bit.
// Layouts, swizzled so that the operand reads of each half-wave hit 32
// distinct banks (the LDS serves 32 banks per half-wave):
//   A chunk: 16 rows k x 64 a, element (k, m) at k 64 + ((m + 16 (k & 3)) & 63);
//   B chunk: per complex column c (64 of them) a 128-byte block of 8 slots of
//     (k, k+1) x (re, im); slot q = k/2 sits at position (q + c) & 7.
// Rows a >= nk_max and columns beyond C load a valid in-range element instead
// of zeros; they only reach outputs that are not stored.  Needs N % 16 == 0.
typedef __attribute__((address_space(3))) void lds_void_t;
#define KD_WAIT(n) __builtin_amdgcn_s_waitcnt(0xF70 | (n))  // vmcnt(n), other counters untouched
__global__ __launch_bounds__(WG) void k_kin_fwd_dma(const float *__restrict__ Q,
                                                    const double2 *__restrict__ kph,
                                                    const int32_t *__restrict__ nkeep,
                                                    const float2 *__restrict__ psi32, KinRef ref,
                                                    float *__restrict__ Y, int N, int nk_max, int C,
                                                    int n_orb) {
    typedef mfma_t<float>::v4 v4;
    // three buffers of (A chunk: 1024 floats, B chunk: 2048 floats)
    __shared__ __align__(16) float lds[3 * 3072];
    const tile3 tb = xcd_tile();  // (eigen tile, column tile, l)
    const int l = tb.z;
    const int nk = nkeep[l];
    const int a0 = tb.x * TM;
    if (a0 >= nk) return;  // block-uniform
    const int n0 = tb.y * TN;
    if (!kin_cols_active(ref, n0, TN, C, n_orb)) return;  // idle instances only
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    // this lane's copy sources: A rows 4 wave + lane / 16 (one instruction per
    // chunk), B columns 16 wave + 8 j + lane / 8 (two instructions per chunk)
    const int ka = 4 * wave + (lane >> 4);
    const int ma = ((lane & 15) * 4 - 16 * (ka & 3)) & 63;
    const float *Asrc = Q + ((size_t)l * N + ka) * nk_max + (a0 + ma < nk_max ? a0 + ma : 0);
    const float2 *Bsrc[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int cl = 16 * wave + 8 * j + (lane >> 3);
        const int q = ((lane & 7) - cl) & 7;
        const int cg = (n0 >> 1) + cl;
        Bsrc[j] = psi32 + ((size_t)l * C + (cg < C ? cg : 0)) * N + 2 * q;
    }
    auto copy = [&](int c, int buf) {
        const int i0 = 16 * c;
        float *As = lds + 3072 * buf, *Bs = As + 1024;
        __builtin_amdgcn_global_load_lds((const void *)(Asrc + (size_t)i0 * nk_max), (lds_void_t *)(As + 256 * wave), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void *)(Bsrc[0] + i0), (lds_void_t *)(Bs + 512 * wave), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void *)(Bsrc[1] + i0), (lds_void_t *)(Bs + 512 * wave + 256), 16, 0, 0);
    };
    // operand offsets within a buffer, the same for every chunk
    int oa0[4], oa1[4], ob[4][4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        const int kk = 4 * ks + kr;
        oa0[ks] = kk * 64 + ((wm + lc + 16 * (kk & 3)) & 63);
        oa1[ks] = kk * 64 + ((wm + 16 + lc + 16 * (kk & 3)) & 63);
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int n = wn + 16 * y + lc, cc = n >> 1;
            ob[ks][y] = 1024 + cc * 32 + (((kk >> 1) + cc) & 7) * 4 + (kk & 1) * 2 + (n & 1);
        }
    }
    v4 acc[2][4], hi[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = (v4){0, 0, 0, 0}, hi[x][y] = (v4){0, 0, 0, 0};
    const int nch = N / 16;
    // chunk c from buffer buf; the copy of chunk c + 2 goes to buffer nb, the
    // one of chunk c - 1, which every wave has finished with at the barrier
    auto chunk = [&](int c, int buf, int nb) {
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wave's reads of the last chunk are done
        if (c + 1 < nch)
            KD_WAIT(3);  // this wave's copies of chunk c are done (chunk c + 1's may be in flight)
        else
            KD_WAIT(0);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (c + 2 < nch) copy(c + 2, nb);
        const float *L = lds + 3072 * buf;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const float av0 = L[oa0[ks]];
            const float av1 = L[oa1[ks]];
            float b[4];
#pragma unroll
            for (int y = 0; y < 4; ++y) b[y] = L[ob[ks][y]];
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                acc[0][y] = mfma_t<float>::op(av0, b[y], acc[0][y]);
                acc[1][y] = mfma_t<float>::op(av1, b[y], acc[1][y]);
            }
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) hi[x][y] += acc[x][y], acc[x][y] = (v4){0, 0, 0, 0};
    };
    copy(0, 0);
    if (nch > 1) copy(1, 1);
    for (int c = 0; c < nch; c += 3) {
        chunk(c, 0, 2);
        if (c + 1 < nch) chunk(c + 1, 1, 0);
        if (c + 2 < nch) chunk(c + 2, 2, 1);
    }
    const double2 *kphl = kph + (size_t)l * nk_max;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int col = n0 + wn + 16 * y + lc;
            const int c = col >> 1, p = col & 1;
            const int ab = a0 + wm + 16 * x;
            float o[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = ab + mfma_t<float>::row(kr, r);
                const double v = (double)hi[x][y][r];
                const double vo = lane_xor1(v);
                o[r] = kin_factor<float>(v, vo, p, kphl[a < nk ? a : 0]);
            }
            if (c < C) {
                float *Yo = Y + (((size_t)l * C + c) * 2 + p) * nk_max;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = ab + mfma_t<float>::row(kr, r);
                    if (a < nk) Yo[a] = o[r];
                }
            }
        }
    }
}

template <typename T, int TKT>
__global__ __launch_bounds__(WG) void k_kin_bwd_tiled(const T *__restrict__ Q,
                                                      const int32_t *__restrict__ nkeep,
                                                      const T *__restrict__ Y, KinRef ref,
                                                      double *__restrict__ psi, int N, int nk_max,
                                                      int C, int n_orb) {
    typedef typename mfma_t<T>::v4 v4;
    constexpr int AV = TKT * TM / WG;  // A values per thread (consecutive a of one row i)
    constexpr int BV = TKT * TN / WG;  // B values per thread (consecutive a of one column)
    __shared__ T As[2][TKT][TMP];      // [k = a][m = i]
    __shared__ T Bs[2][TKT][TNP];      // [k = a][n = real column]
    const tile3 tb = xcd_tile();       // (radial tile, column tile, l)
    const int l = tb.z;
    const int nk = nkeep[l];
    const int i0 = tb.x * TM;
    const int n0 = tb.y * TN;
    if (!kin_cols_active(ref, n0, TN, C, n_orb)) return;  // idle instances only
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    const T *Ql = Q + (size_t)l * N * nk_max;
    const T *Yl = Y + (size_t)l * C * 2 * nk_max;
    const int am = t / (TKT / AV), ak = (t % (TKT / AV)) * AV;  // A: row i, AV consecutive a
    const int bn = t / (TKT / BV), bk = (t % (TKT / BV)) * BV;  // B: real column, BV consecutive a
    const int ia = i0 + am;
    const int colb = n0 + bn;
    const bool bok = (colb >> 1) < C;
    // the tables and Y are zero from nkeep up to nk_max: runs of AV, BV (<= 16,
    // dividing 16) are wholly inside or outside [0, nk_max)
    T ra[AV], rb[BV];
    auto load = [&](int k0) {
        const bool oka = ia < N && k0 + ak < nk_max;
#pragma unroll
        for (int u = 0; u < AV; ++u) ra[u] = oka ? Ql[(size_t)ia * nk_max + k0 + ak + u] : (T)0;
        const bool okb = bok && k0 + bk < nk_max;
#pragma unroll
        for (int u = 0; u < BV; ++u) rb[u] = okb ? Yl[(size_t)colb * nk_max + k0 + bk + u] : (T)0;
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int u = 0; u < AV; ++u) As[buf][ak + u][am] = ra[u];
#pragma unroll
        for (int u = 0; u < BV; ++u) Bs[buf][bk + u][bn] = rb[u];
    };
    v4 acc[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = (v4){0, 0, 0, 0};
    load(0);
    stash(0);
    __syncthreads();
    // two-level sum: each chunk's products in acc, added to hi after the chunk,
    // so no FP32 accumulation chain is longer than one chunk plus the chunk count
    v4 hi[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) hi[x][y] = (v4){0, 0, 0, 0};
    int buf = 0;
    for (int k0 = 0; k0 < nk; k0 += TKT) {
        const bool more = k0 + TKT < nk;
        if (more) load(k0 + TKT);
#pragma unroll
        for (int ks = 0; ks < TKT / 4; ++ks) kin_tile_step<T>(As[buf], Bs[buf], ks * 4 + kr, wm, wn, lc, acc);
        if (more) stash(buf ^ 1);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) hi[x][y] += acc[x][y], acc[x][y] = (v4){0, 0, 0, 0};
        __syncthreads();
        buf ^= 1;
    }
    double *Xl = psi + (size_t)l * C * N * 2;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int col = n0 + wn + 16 * y + lc;
            const int c = col >> 1, p = col & 1;
            if (c >= C) continue;
            double rc = 0.0;
            const double *rrow = mfma_t<T>::mixed ? kin_ref_row(ref, l, c, p, N, n_orb, true, &rc) : nullptr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wm + 16 * x + mfma_t<T>::row(kr, r);
                if (i < N) {
                    double v = (double)hi[x][y][r];
                    if (rrow) v += rc * rrow[i];
                    Xl[((size_t)c * N + i) * 2 + p] = v;
                }
            }
        }
    }
}

// FP32 tiled pass 2 with direct global-to-LDS copies, as k_kin_fwd_dma: the
// products, chunks and sums of k_kin_bwd_tiled<float, 16> in the same order
// (the same psi bit for bit).  Both operands are K-contiguous in memory (Q
// rows i, Y rows of a column), so the chunk tiles are kept row-major by their
// M or N index: element (row, k) at row 16 + 4 ((k / 4 + (row / 4)) & 3) + k % 4.
// The LDS serves 32 banks per half-wave, and the 16 lanes of a half-wave read
// 16 different 64-byte rows: a 2-way bank conflict that no rotation within a
// row removes.  Rows i >= N and columns beyond C load row/column 0 and are
// not stored.
__global__ __launch_bounds__(WG) void k_kin_bwd_dma(const float *__restrict__ Q,
                                                    const int32_t *__restrict__ nkeep,
                                                    const float *__restrict__ Y, KinRef ref,
                                                    double *__restrict__ psi, int N, int nk_max,
                                                    int C, int n_orb) {
    typedef mfma_t<float>::v4 v4;
    __shared__ __align__(16) float lds[3 * 3072];  // three buffers of (A: 64 x 16, B: 128 x 16)
    const tile3 tb = xcd_tile();  // (radial tile, column tile, l)
    const int l = tb.z;
    const int nk = nkeep[l];
    const int i0 = tb.x * TM;
    const int n0 = tb.y * TN;
    if (!kin_cols_active(ref, n0, TN, C, n_orb)) return;  // idle instances only
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    // copy sources: A rows 16 wave + lane / 4 (one instruction), B columns
    // 32 wave + 16 j + lane / 4 (two instructions); slot lane % 4 of a row holds
    // the K quad (slot - row / 4) & 3
    const int ra = 16 * wave + (lane >> 2);
    const int qa = ((lane & 3) - (ra >> 2)) & 3;
    const float *Asrc = Q + ((size_t)l * N + (i0 + ra < N ? i0 + ra : 0)) * nk_max + 4 * qa;
    const float *Bsrc[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int rb = 32 * wave + 16 * j + (lane >> 2);
        const int qb = ((lane & 3) - (rb >> 2)) & 3;
        const int col = n0 + rb;
        Bsrc[j] = Y + ((size_t)l * C * 2 + ((col >> 1) < C ? col : 0)) * nk_max + 4 * qb;
    }
    auto copy = [&](int c, int buf) {
        const int k0 = 16 * c;
        float *As = lds + 3072 * buf, *Bs = As + 1024;
        __builtin_amdgcn_global_load_lds((const void *)(Asrc + k0), (lds_void_t *)(As + 256 * wave), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void *)(Bsrc[0] + k0), (lds_void_t *)(Bs + 512 * wave), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void *)(Bsrc[1] + k0), (lds_void_t *)(Bs + 512 * wave + 256), 16, 0, 0);
    };
    int oa0[4], oa1[4], ob[4][4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        const int m0 = wm + lc, m1 = wm + 16 + lc;
        oa0[ks] = m0 * 16 + 4 * ((ks + (m0 >> 2)) & 3) + kr;
        oa1[ks] = m1 * 16 + 4 * ((ks + (m1 >> 2)) & 3) + kr;
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int n = wn + 16 * y + lc;
            ob[ks][y] = 1024 + n * 16 + 4 * ((ks + (n >> 2)) & 3) + kr;
        }
    }
    v4 acc[2][4], hi[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = (v4){0, 0, 0, 0}, hi[x][y] = (v4){0, 0, 0, 0};
    const int nch = (nk + 15) / 16;
    auto chunk = [&](int c, int buf, int nb) {
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wave's reads of the last chunk are done
        if (c + 1 < nch)
            KD_WAIT(3);
        else
            KD_WAIT(0);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (c + 2 < nch) copy(c + 2, nb);
        const float *L = lds + 3072 * buf;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const float av0 = L[oa0[ks]];
            const float av1 = L[oa1[ks]];
            float b[4];
#pragma unroll
            for (int y = 0; y < 4; ++y) b[y] = L[ob[ks][y]];
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                acc[0][y] = mfma_t<float>::op(av0, b[y], acc[0][y]);
                acc[1][y] = mfma_t<float>::op(av1, b[y], acc[1][y]);
            }
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) hi[x][y] += acc[x][y], acc[x][y] = (v4){0, 0, 0, 0};
    };
    if (nch > 0) copy(0, 0);
    if (nch > 1) copy(1, 1);
    for (int c = 0; c < nch; c += 3) {
        chunk(c, 0, 2);
        if (c + 1 < nch) chunk(c + 1, 1, 0);
        if (c + 2 < nch) chunk(c + 2, 2, 1);
    }
    double *Xl = psi + (size_t)l * C * N * 2;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int col = n0 + wn + 16 * y + lc;
            const int c = col >> 1, p = col & 1;
            if (c >= C) continue;
            double rc = 0.0;
            const double *rrow = kin_ref_row(ref, l, c, p, N, n_orb, true, &rc);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wm + 16 * x + mfma_t<float>::row(kr, r);
                if (i < N) {
                    double v = (double)hi[x][y][r];
                    if (rrow) v += rc * rrow[i];
                    Xl[((size_t)c * N + i) * 2 + p] = v;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Compensated split-FP32 kinetic passes (precision 2).  Every FP64 operand is
// split into FP32 parts, v = v_hi + v_lo with v_hi = fl32(v), v_lo =
// fl32(v - v_hi), and a product a b is formed as a_hi b_hi + a_hi b_lo +
// a_lo b_hi on the FP32 matrix cores (the dropped a_lo b_lo is 2^-48 of
// a b); the main and correction products go to separate FP32 accumulators,
// whose chunk sums (16 values of K) are added in FP64.  Tiles, staging and
// XCD order as the tiled passes above (64 x 128 work-group tiles, K chunks of
// 16 double buffered in LDS).  Pass 1 splits the FP64 state minus its
// ground-state component itself; Y leaves pass 1 as (Y_hi, Y_lo).
// ---------------------------------------------------------------------------
#define SPK 16
static __device__ __forceinline__ void split32(double v, float &hi, float &lo) {
    hi = (float)v;
    lo = (float)(v - (double)hi);
}

// 3 x 8 products of one K step of 4 for the wave's 2 x 4 tiles
static __device__ __forceinline__ void kin_split_step(const float (*Ah)[TMP], const float (*Al)[TMP],
                                                      const float (*Bh)[TNP], const float (*Bl)[TNP],
                                                      int kk, int wm, int wn, int lc, f4_t (&acc)[2][4],
                                                      f4_t (&cor)[2][4]) {
    const float ah0 = Ah[kk][wm + lc], ah1 = Ah[kk][wm + 16 + lc];
    const float al0 = Al[kk][wm + lc], al1 = Al[kk][wm + 16 + lc];
    float bh[4], bl[4];
#pragma unroll
    for (int y = 0; y < 4; ++y) bh[y] = Bh[kk][wn + 16 * y + lc], bl[y] = Bl[kk][wn + 16 * y + lc];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        acc[0][y] = mfma_t<float>::op(ah0, bh[y], acc[0][y]);
        acc[1][y] = mfma_t<float>::op(ah1, bh[y], acc[1][y]);
        cor[0][y] = mfma_t<float>::op(ah0, bl[y], cor[0][y]);
        cor[1][y] = mfma_t<float>::op(ah1, bl[y], cor[1][y]);
        cor[0][y] = mfma_t<float>::op(al0, bh[y], cor[0][y]);
        cor[1][y] = mfma_t<float>::op(al1, bh[y], cor[1][y]);
    }
}

__global__ __launch_bounds__(WG) void k_kin_fwd_split(const float *__restrict__ Qh,
                                                      const float *__restrict__ Ql,
                                                      const double2 *__restrict__ kph,
                                                      const int32_t *__restrict__ nkeep,
                                                      const double *__restrict__ psi, KinRef ref,
                                                      float *__restrict__ Yh, float *__restrict__ Yl,
                                                      int N, int nk_max, int C, int n_orb) {
    constexpr int AV = SPK * TM / WG;  // A values per thread (consecutive a)
    constexpr int BV = SPK * 64 / WG;  // complex B values per thread (consecutive i)
    __shared__ float Ash[2][SPK][TMP], Asl[2][SPK][TMP];  // [k = i][m = a]
    __shared__ float Bsh[2][SPK][TNP], Bsl[2][SPK][TNP];  // [k = i][n = real column]
    const tile3 tb = xcd_tile();
    const int l = tb.z;
    const int nk = nkeep[l];
    const int a0 = tb.x * TM;
    if (a0 >= nk) return;  // block-uniform
    const int n0 = tb.y * TN;
    if (!kin_cols_active(ref, n0, TN, C, n_orb)) return;  // idle instances only
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    const size_t qoff = (size_t)l * N * nk_max;
    const double2 *Xl = (const double2 *)psi + (size_t)l * C * N;
    const int ar = t / (TM / AV), am = (t % (TM / AV)) * AV;
    const int bc = t / (SPK / BV), bk = (t % (SPK / BV)) * BV;
    const int cg = (n0 >> 1) + bc;
    const bool cok = cg < C;
    const bool aok = a0 + am < nk_max;
    // ground-state component of this thread's column in block l (both parts)
    double cre = 0.0, cim = 0.0;
    const double *rre = cok ? kin_ref_row(ref, l, cg, 0, N, n_orb, false, &cre) : nullptr;
    if (rre) (void)kin_ref_row(ref, l, cg, 1, N, n_orb, false, &cim);
    float rah[AV], ral[AV], rbh[2 * BV], rbl[2 * BV];
    auto load = [&](int k0) {
        const int i = k0 + ar;
        const bool ok = aok && i < N;
#pragma unroll
        for (int u = 0; u < AV; ++u) {
            const size_t o = qoff + (size_t)i * nk_max + a0 + am + u;
            rah[u] = ok ? Qh[o] : 0.f;
            ral[u] = ok ? Ql[o] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < BV; ++u) {
            const int ii = k0 + bk + u;
            double2 v = (cok && ii < N) ? Xl[(size_t)cg * N + ii] : make_double2(0.0, 0.0);
            if (rre && ii < N) {
                v.x -= cre * rre[ii];
                v.y -= cim * rre[ii];
            }
            split32(v.x, rbh[2 * u], rbl[2 * u]);
            split32(v.y, rbh[2 * u + 1], rbl[2 * u + 1]);
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int u = 0; u < AV; ++u) Ash[buf][ar][am + u] = rah[u], Asl[buf][ar][am + u] = ral[u];
#pragma unroll
        for (int u = 0; u < BV; ++u) {
            Bsh[buf][bk + u][2 * bc] = rbh[2 * u];
            Bsh[buf][bk + u][2 * bc + 1] = rbh[2 * u + 1];
            Bsl[buf][bk + u][2 * bc] = rbl[2 * u];
            Bsl[buf][bk + u][2 * bc + 1] = rbl[2 * u + 1];
        }
    };
    f4_t acc[2][4], cor[2][4];
    double hi[2][4][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            acc[x][y] = (f4_t){0, 0, 0, 0};
            cor[x][y] = (f4_t){0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r) hi[x][y][r] = 0.0;
        }
    load(0);
    stash(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < N; k0 += SPK) {
        const bool more = k0 + SPK < N;
        if (more) load(k0 + SPK);
#pragma unroll
        for (int ks = 0; ks < SPK / 4; ++ks)
            kin_split_step(Ash[buf], Asl[buf], Bsh[buf], Bsl[buf], ks * 4 + kr, wm, wn, lc, acc, cor);
        if (more) stash(buf ^ 1);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
#pragma unroll
                for (int r = 0; r < 4; ++r) hi[x][y][r] += (double)acc[x][y][r] + (double)cor[x][y][r];
                acc[x][y] = (f4_t){0, 0, 0, 0};
                cor[x][y] = (f4_t){0, 0, 0, 0};
            }
        __syncthreads();
        buf ^= 1;
    }
    const double2 *kphl = kph + (size_t)l * nk_max;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int col = n0 + wn + 16 * y + lc;
            const int c = col >> 1, p = col & 1;
            const int ab = a0 + wm + 16 * x;
            double o[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = ab + mfma_t<float>::row(kr, r);
                const double v = hi[x][y][r];
                const double vo = lane_xor1(v);
                o[r] = kin_factor<double>(v, vo, p, kphl[a < nk ? a : 0]);
            }
            if (c < C) {
                const size_t yo = (((size_t)l * C + c) * 2 + p) * nk_max;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = ab + mfma_t<float>::row(kr, r);
                    if (a < nk) split32(o[r], Yh[yo + a], Yl[yo + a]);
                }
            }
        }
    }
}

__global__ __launch_bounds__(WG) void k_kin_bwd_split(const float *__restrict__ Qh,
                                                      const float *__restrict__ Ql,
                                                      const int32_t *__restrict__ nkeep,
                                                      const float *__restrict__ Yh,
                                                      const float *__restrict__ Yl, KinRef ref,
                                                      double *__restrict__ psi, int N, int nk_max,
                                                      int C, int n_orb) {
    constexpr int AV = SPK * TM / WG;  // A values per thread (consecutive a of one row i)
    constexpr int BV = SPK * TN / WG;  // B values per thread (consecutive a of one column)
    __shared__ float Ash[2][SPK][TMP], Asl[2][SPK][TMP];  // [k = a][m = i]
    __shared__ float Bsh[2][SPK][TNP], Bsl[2][SPK][TNP];  // [k = a][n = real column]
    const tile3 tb = xcd_tile();
    const int l = tb.z;
    const int nk = nkeep[l];
    const int i0 = tb.x * TM;
    const int n0 = tb.y * TN;
    if (!kin_cols_active(ref, n0, TN, C, n_orb)) return;  // idle instances only
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    const size_t qoff = (size_t)l * N * nk_max;
    const size_t yoff = (size_t)l * C * 2 * nk_max;
    const int am = t / (SPK / AV), ak = (t % (SPK / AV)) * AV;
    const int bn = t / (SPK / BV), bk = (t % (SPK / BV)) * BV;
    const int ia = i0 + am;
    const int colb = n0 + bn;
    const bool bok = (colb >> 1) < C;
    // the tables and Y are zero from nkeep up to nk_max: runs of AV, BV (<= 16,
    // dividing 16) are wholly inside or outside [0, nk_max)
    float rah[AV], ral[AV], rbh[BV], rbl[BV];
    auto load = [&](int k0) {
        const bool oka = ia < N && k0 + ak < nk_max;
#pragma unroll
        for (int u = 0; u < AV; ++u) {
            const size_t o = qoff + (size_t)ia * nk_max + k0 + ak + u;
            rah[u] = oka ? Qh[o] : 0.f;
            ral[u] = oka ? Ql[o] : 0.f;
        }
        const bool okb = bok && k0 + bk < nk_max;
#pragma unroll
        for (int u = 0; u < BV; ++u) {
            const size_t o = yoff + (size_t)colb * nk_max + k0 + bk + u;
            rbh[u] = okb ? Yh[o] : 0.f;
            rbl[u] = okb ? Yl[o] : 0.f;
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int u = 0; u < AV; ++u) Ash[buf][ak + u][am] = rah[u], Asl[buf][ak + u][am] = ral[u];
#pragma unroll
        for (int u = 0; u < BV; ++u) Bsh[buf][bk + u][bn] = rbh[u], Bsl[buf][bk + u][bn] = rbl[u];
    };
    f4_t acc[2][4], cor[2][4];
    double hi[2][4][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            acc[x][y] = (f4_t){0, 0, 0, 0};
            cor[x][y] = (f4_t){0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r) hi[x][y][r] = 0.0;
        }
    load(0);
    stash(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < nk; k0 += SPK) {
        const bool more = k0 + SPK < nk;
        if (more) load(k0 + SPK);
#pragma unroll
        for (int ks = 0; ks < SPK / 4; ++ks)
            kin_split_step(Ash[buf], Asl[buf], Bsh[buf], Bsl[buf], ks * 4 + kr, wm, wn, lc, acc, cor);
        if (more) stash(buf ^ 1);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
#pragma unroll
                for (int r = 0; r < 4; ++r) hi[x][y][r] += (double)acc[x][y][r] + (double)cor[x][y][r];
                acc[x][y] = (f4_t){0, 0, 0, 0};
                cor[x][y] = (f4_t){0, 0, 0, 0};
            }
        __syncthreads();
        buf ^= 1;
    }
    double *Xo = psi + (size_t)l * C * N * 2;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int col = n0 + wn + 16 * y + lc;
            const int c = col >> 1, p = col & 1;
            if (c >= C) continue;
            double rc = 0.0;
            const double *rrow = kin_ref_row(ref, l, c, p, N, n_orb, true, &rc);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wm + 16 * x + mfma_t<float>::row(kr, r);
                if (i < N) {
                    double v = hi[x][y][r];
                    if (rrow) v += rc * rrow[i];
                    Xo[((size_t)c * N + i) * 2 + p] = v;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Angular kernel.  One work-group = one orbital column c, RT radial nodes.
//
// Gauss-Legendre nodes are symmetric, chi_{j'} = -chi_j with j' = Nth-1-j, and
// Theta_lm(-x) = (-1)^{l-m} Theta_lm(x).  Splitting the l of an orbital by the
// parity of l - m (s: l = m + 2k, a: l = m + 1 + 2k) and keeping the first
// h = ceil(Nth/2) nodes, the transform of the tile X[l][col] becomes
//     E = Bs X_s,  O = Ba X_a,   Psi_j = E_j + O_j,  Psi_{j'} = E_j - O_j,
// and the back transform of the theta-space values P
//     X_s = Bs^T S,  X_a = Ba^T D,  S_j = P_j + P_{j'},  D_j = P_j - P_{j'}
// (middle node of odd Nth: S = D = P_mid).  That is half the matrix-core work
// of the dense N_theta x N_l transform, and the two half tables of one m
// (hq x (sS + sA) values) are staged in LDS once per work-group (GLB = true
// reads them from global memory when they do not fit).
// LDS: tables (T) | X[Nl][16] (FP64) | S[hq][16] | D[hq][16] (T), real column
// n = 2 * i_local + re/im; in the mixed form S/D overlay X during the
// potential half step (AngArgs.alias).  The state tile X, the laser rotations, the phase
// factors and the observable sums are FP64 in both forms; the mixed form
// applies the potential half step as X += B^T ((f - 1) B X).  The dipole-type
// sums use sum_j chi_j |Psi_j|^2 = sum_{j<h} chi_j 4 Re(E_j^* O_j), so that the
// large even part of a core orbital never enters them through a difference.
// mode 0: Vh, Le, Lo          (first half of the step)
// mode 1: Lo, Le, Vh, store, then |B psi|^2 + observable partials
// mode 2: |B psi|^2 + observable partials only (initial potential)
// Waves: wave w takes theta tile w of the forward product and output tile w
// of the back product (s tiles first, then a tiles); the launch uses
// max(ceil(h/16), ceil(ns/16) + ceil(na/16)) <= 8 waves.
// ---------------------------------------------------------------------------
#define ANG_MAXW 8
#ifndef ANG_DMA
#define ANG_DMA 1  // tables and l-space tile copied straight into LDS
#endif
#define ANG_MAX_NL 128  // largest N_l ggb_create accepts
struct AngArgs {
    double *psi;
    const void *Bh;        // [n_m][nblk] parity half tables, compute type
    const int32_t *mcol;   // [n_orb]
    const void *phase;     // [n_inst][N][Nth] f = mask e^{-i dV dt/2}: double2 f (FP64 form),
                           // float2 (f - 1) (mixed form)
    const double *r;
    const double *chi;     // [Nth]
    const double *E;       // [n_inst][n_field]
    void *dens;            // [C][N][2][hq] |Psi|^2 in the compute type, folded: [..][0][j] is
                           // node j < h, [..][1][j] its mirror N_theta - 1 - j
    double *obs_part;      // [C][ntiles][3]
    const int *cur;        // graph replay: field step = (cur[0] + step) mod n_field
    float2 *psi32;         // mode 0, FP32 tiled kinetic: FP32 copy of the stored state
    const int32_t *l0;     // blocks left to k_kin_prep (ground-state component), or null
    int N, Nl, Nlx, Nth, h, hq, sS, sA, nblk, C, n_orb, n_field, step;
    int alias;  // mixed form: S/D live in the l-space tile's LDS (see k_angular)
    const int32_t *active;  // [n_inst] 0: instance idle, its work-groups return at once
    double dt;
};

static __device__ __forceinline__ double c_lm(int l, int m) {
    // <l m|cos theta|l+1 m> (internal.py:80-84, angularCoefficients.cuh d_c)
    const double dl = (double)l, dm = (double)m;
    return sqrt(((dl + 1.0) * (dl + 1.0) - dm * dm) / ((2.0 * dl + 1.0) * (2.0 * dl + 3.0)));
}

// cl[l] = c_lm(l, m) of the work-group's m, staged in LDS once; X rows hold
// XS = 16 NT doubles (RT NT radial nodes, re/im interleaved)
template <int NT>
static __device__ void ang_laser(double *X, int parity, int m, int Nl, const double *a_i,
                                 const double *cl) {
    // pairs (l, l+1), l = m + parity + 2 q, l + 1 < Nl; thread -> (q, i_local)
    constexpr int R = RT * NT, XS = 16 * NT;
    const int t = threadIdx.x;
    const int il = t % R;
    for (int q = t / R;; q += blockDim.x / R) {
        const int l = m + parity + 2 * q;
        if (l + 1 >= Nl) break;
        const double th = a_i[il] * cl[l];
        double sn, cs;
        sincos(th, &sn, &cs);
        double *x0 = X + l * XS + 2 * il;
        double *x1 = X + (l + 1) * XS + 2 * il;
        const double r0 = x0[0], i0 = x0[1], r1 = x1[0], i1 = x1[1];
        // [x0', x1'] = [[cs, -i sn], [-i sn, cs]] [x0, x1]
        x0[0] = cs * r0 + sn * i1;
        x0[1] = cs * i0 - sn * r1;
        x1[0] = cs * r1 + sn * i0;
        x1[1] = cs * i1 - sn * r0;
    }
}

// parity sizes of orbital m: ns = #{l = m + 2k <= Nl-1}, na = #{l = m + 1 + 2k <= Nl-1}
struct ang_par {
    int ns, na, ks, ka;
    __device__ ang_par(int Nl, int m) {
        ns = (Nl - 1 - m) / 2 + 1;
        na = (Nl - m) / 2;
        ks = (ns + 3) & ~3;
        ka = (na + 3) & ~3;
    }
};

// forward product of theta tile `tile` (rows j = 16 tile + 0..15 of the first
// half) for the NT column tiles: e[t] = Bs X_s, o[t] = Ba X_a; each table
// value feeds NT products.
template <typename T, int NT>
static __device__ void ang_fwd(const T *Bs, const T *Ba, const double *X, const AngArgs &g,
                               const ang_par &pr, int m, int tile, typename mfma_t<T>::v4 (&e)[NT],
                               typename mfma_t<T>::v4 (&o)[NT]) {
    typedef typename mfma_t<T>::v4 v4;
    constexpr int XS = 16 * NT;
    const int lane = threadIdx.x & 63, kr = lane >> 4, lc = lane & 15;
    // rows j >= hq read row hq - 1: their results are never used (j >= h);
    // table columns k >= ns are zero and X rows up to Nl + 5 exist (zero), so
    // the loops need no per-element guards
    const int j = min(tile * 16 + lc, g.hq - 1);
    const T *As = Bs + (size_t)j * g.sS;
    const T *Aa = Ba + (size_t)j * g.sA;
#pragma unroll
    for (int u = 0; u < NT; ++u) e[u] = (v4){0, 0, 0, 0}, o[u] = (v4){0, 0, 0, 0};
    const int kmax = pr.ks > pr.ka ? pr.ks : pr.ka;
    for (int k0 = 0; k0 < kmax; k0 += 4) {
        const int k = k0 + kr;
        if (k0 < pr.ks) {
            const T a = As[k];
#pragma unroll
            for (int u = 0; u < NT; ++u) e[u] = mfma_t<T>::op(a, (T)X[(m + 2 * k) * XS + 16 * u + lc], e[u]);
        }
        if (k0 < pr.ka) {
            const T a = Aa[k];
#pragma unroll
            for (int u = 0; u < NT; ++u)
                o[u] = mfma_t<T>::op(a, (T)X[(m + 1 + 2 * k) * XS + 16 * u + lc], o[u]);
        }
    }
}

// row of X that output r of this lane's back-product item `item` (s tiles,
// then a tiles) belongs to, or -1 past the parity's last l
template <typename T>
static __device__ int ang_bwd_row(const ang_par &pr, int m, int item, int r) {
    const int kr = (threadIdx.x & 63) >> 4;
    const int nst = (pr.ns + 15) / 16;
    const bool sym = item < nst;
    const int ko = (sym ? item : item - nst) * 16 + mfma_t<T>::row(kr, r);
    return ko < (sym ? pr.ns : pr.na) ? m + (sym ? 0 : 1) + 2 * ko : -1;
}

// back product of output tile `item` (s tiles, then a tiles) for the NT column
// tiles: v[t][r] for the rows ang_bwd_row gives; two accumulator chains per
// column tile over alternating groups of 4 nodes, added at the end.
template <typename T, int NT>
static __device__ void ang_bwd_sum(const T *Bs, const T *Ba, const T *S, const T *D,
                                   const AngArgs &g, const ang_par &pr, int item, double (&v)[NT][4]) {
    typedef typename mfma_t<T>::v4 v4;
    constexpr int XS = 16 * NT;
    const int lane = threadIdx.x & 63, kr = lane >> 4, lc = lane & 15;
    const int nst = (pr.ns + 15) / 16;
    const bool sym = item < nst;
    const int tile = sym ? item : item - nst;
    const T *A = sym ? Bs : Ba;
    const T *Pv = sym ? S : D;
    const int st = sym ? g.sS : g.sA;
    const int kk = sym ? pr.ks : pr.ka;
    // output rows k >= kk read column kk - 1 (in range); they are not stored
    const T *Ac = A + min(tile * 16 + lc, kk - 1);
    v4 a0[NT], a1[NT];
#pragma unroll
    for (int u = 0; u < NT; ++u) a0[u] = (v4){0, 0, 0, 0}, a1[u] = (v4){0, 0, 0, 0};
    int j0 = 0;
    for (; j0 + 8 <= g.hq; j0 += 8) {
        const int ja = j0 + kr, jb = j0 + 4 + kr;
        const T xa = Ac[(size_t)ja * st], xb = Ac[(size_t)jb * st];
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            a0[u] = mfma_t<T>::op(xa, Pv[ja * XS + 16 * u + lc], a0[u]);
            a1[u] = mfma_t<T>::op(xb, Pv[jb * XS + 16 * u + lc], a1[u]);
        }
    }
    if (j0 < g.hq) {
        const int ja = j0 + kr;
        const T xa = Ac[(size_t)ja * st];
#pragma unroll
        for (int u = 0; u < NT; ++u) a0[u] = mfma_t<T>::op(xa, Pv[ja * XS + 16 * u + lc], a0[u]);
    }
#pragma unroll
    for (int u = 0; u < NT; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) v[u][r] = (double)(a0[u][r] + a1[u][r]);
}

template <int MODE, typename T, bool GLB, int NT>
__global__ __launch_bounds__(64 * ANG_MAXW) void k_angular(AngArgs g) {
    typedef typename mfma_t<T>::v4 v4;
    constexpr int R = RT * NT, XS = 16 * NT;  // radial nodes and X row length of the work-group
    extern __shared__ __align__(16) double lds[];
    __shared__ double a_i[R];
    __shared__ double cl[ANG_MAX_NL];
    __shared__ double red[3][ANG_MAXW];

    // logical order (orbital, radial tile, instance), orbital fastest: the
    // orbitals of one instance and radial tile share the phase factors in L2
    const int lg = xcd_linear();
    const int korb = lg % g.n_orb, rest = lg / g.n_orb;
    const int tile_i = rest % gridDim.x, inst = rest / gridDim.x;
    if (g.active && !g.active[inst]) return;  // idle serving slot (block-uniform)
    const int c = inst * g.n_orb + korb;
    const int i0 = tile_i * R;
    const int m = g.mcol[c % g.n_orb];
    const int t = threadIdx.x, nthr = blockDim.x;
    const int wave = t >> 6, lane = t & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int ilc = lc >> 1, p = lc & 1;  // column -> (i_local within the column tile, re/im)
    const ang_par pr(g.Nl, m);
    const int mt = (g.h + 15) / 16;
    const int nbt = (pr.ns + 15) / 16 + (pr.na + 15) / 16;

    const T *Bsrc = reinterpret_cast<const T *>(g.Bh) + (size_t)m * g.nblk;
    T *Tl = reinterpret_cast<T *>(lds);
    double *X = reinterpret_cast<double *>(GLB ? Tl : Tl + g.nblk);  // [Nlx][XS]
    const bool AL = mfma_t<T>::mixed && g.alias;                     // S/D inside X
    T *S = reinterpret_cast<T *>(AL ? X : X + g.Nlx * XS);           // [hq][XS]
    T *D = S + g.hq * XS;                                            // [hq][XS]
    const T *Bs = GLB ? Bsrc : Tl;
    const T *Ba = Bs + (size_t)g.hq * g.sS;

    // radial node of this lane in column tile u: i0 + 8 u + ilc
    int iu[NT];
    bool oku[NT];
#pragma unroll
    for (int u = 0; u < NT; ++u) iu[u] = i0 + RT * u + ilc, oku[u] = iu[u] < g.N;
    if (ANG_DMA) {
        // straight from global memory into LDS (global_load_lds_dwordx4: 16 bytes
        // per lane, 1 KB per wave instruction, no staging registers)
        const int nw = nthr >> 6;
        if (!GLB) {
            // the block of one m is a multiple of 16 bytes (host pads nblk)
            const int n16 = g.nblk * (int)sizeof(T) / 16;
            for (int b = 64 * wave; b < n16; b += 64 * nw)
                if (b + lane < n16)
                    __builtin_amdgcn_global_load_lds((const void *)(reinterpret_cast<const int4 *>(Bsrc) + b + lane),
                                                     (lds_void_t *)(reinterpret_cast<int4 *>(Tl) + b), 16, 0, 0);
        }
        // rows l < N_l of the l-space tile: a row is the R nodes i0.. of psi[l][c]
        // (R x 16 bytes, contiguous in both), 64 / R rows per instruction; nodes
        // past N load node N - 1 (their results are not stored)
        constexpr int RPI = 64 / R;
        const int ii = min(i0 + lane % R, g.N - 1);
        for (int l0 = RPI * wave; l0 < g.Nl; l0 += RPI * nw) {
            const int l = l0 + lane / R;
            if (l < g.Nl)
                __builtin_amdgcn_global_load_lds((const void *)((const double2 *)g.psi + ((size_t)l * g.C + c) * g.N + ii),
                                                 (lds_void_t *)(X + l0 * XS), 16, 0, 0);
        }
        for (int idx = t; idx < (g.Nlx - g.Nl) * XS; idx += nthr) X[g.Nl * XS + idx] = 0.0;
    } else {
        if (!GLB) {
            // the block of one m is a multiple of 16 bytes (host pads nblk)
            const int n16 = g.nblk * (int)sizeof(T) / 16;
            const int4 *s4 = reinterpret_cast<const int4 *>(Bsrc);
            int4 *d4 = reinterpret_cast<int4 *>(Tl);
            for (int idx = t; idx < n16; idx += nthr) d4[idx] = s4[idx];
        }
        // load the l-space tile
        for (int idx = t; idx < g.Nlx * R; idx += nthr) {
            const int l = idx / R, il = idx % R;
            const int ii = i0 + il;
            double2 v = make_double2(0.0, 0.0);
            if (l < g.Nl && ii < g.N) v = ((const double2 *)g.psi)[((size_t)l * g.C + c) * g.N + ii];
            X[l * XS + 2 * il] = v.x;
            X[l * XS + 2 * il + 1] = v.y;
        }
    }
    if (MODE != 2 && t < R) {
        const int i = i0 + t;
        const int fs = g.cur ? (int)(((long)g.cur[0] + g.step) % g.n_field) : g.step;
        const double Ev = g.E[(size_t)inst * g.n_field + fs];
        a_i[t] = (i < g.N) ? 0.5 * g.dt * Ev * g.r[i] : 0.0;
    }
    if (MODE != 2)
        for (int l = t; l < g.Nl; l += nthr) cl[l] = c_lm(l, m);
    __syncthreads();

    if (MODE == 1) {
        ang_laser<NT>(X, 1, m, g.Nl, a_i, cl);
        __syncthreads();
        ang_laser<NT>(X, 0, m, g.Nl, a_i, cl);
        __syncthreads();
    }

    if (MODE != 2) {
        // Vh: theta space, phase, folded sums S/D, back to l space.  With S/D
        // inside the l-space tile (AL): the forward products and this wave's
        // old tile values are taken first, S/D written after a barrier, the
        // back products formed in registers, and the tile rewritten after a
        // barrier (rows l < m and l >= N_l back to zero)
        T svr[NT][4], dvr[NT][4];
        if (wave < mt) {
            // phase factors at this lane's theta rows j and j' = Nth - 1 - j of
            // each column tile, loaded ahead of the forward products that cover
            // their latency: ph[u][r] = (f_j.x, f_j.y, f_j'.x, f_j'.y)
            T ph[NT][4][4];
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                const size_t pbase = ((size_t)inst * g.N + (oku[u] ? iu[u] : 0)) * g.Nth;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = wave * 16 + mfma_t<T>::row(kr, r);
                    const int jq = j < g.h ? j : 0;  // rows past the half: any valid row, unused
                    const int jm = g.Nth - 1 - jq;
                    if (mfma_t<T>::mixed) {
                        // mixed form: the table holds (f - 1), formed in FP64 before the rounding
                        const float2 *pf = reinterpret_cast<const float2 *>(g.phase) + pbase;
                        const float2 fa_ = pf[jq], fb_ = pf[jm];
                        ph[u][r][0] = fa_.x, ph[u][r][1] = fa_.y, ph[u][r][2] = fb_.x, ph[u][r][3] = fb_.y;
                    } else {
                        const double2 *pf = reinterpret_cast<const double2 *>(g.phase) + pbase;
                        const double2 fa_ = pf[jq], fb_ = pf[jm];
                        ph[u][r][0] = (T)fa_.x, ph[u][r][1] = (T)fa_.y;
                        ph[u][r][2] = (T)fb_.x, ph[u][r][3] = (T)fb_.y;
                    }
                }
            }
            v4 e[NT], o[NT];
            ang_fwd<T, NT>(Bs, Ba, X, g, pr, m, wave, e, o);
#pragma unroll
            for (int u = 0; u < NT; ++u) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = wave * 16 + mfma_t<T>::row(kr, r);
                    const int jm = g.Nth - 1 - j;
                    const T pa = e[u][r] + o[u][r], pb = e[u][r] - o[u][r];
                    const T pao = lane_xor1(pa), pbo = lane_xor1(pb);
                    T sv = (T)0, dv = (T)0;
                    if (j < g.h && oku[u]) {
                        // component p of f (x + i y): x fx - y fy (p = 0), y fx + x fy (p = 1);
                        // this lane holds component p, its partner the other
                        const T sg = p ? (T)1 : (T)-1;
                        const T qa = pa * ph[u][r][0] + sg * pao * ph[u][r][1];
                        const T qb = pb * ph[u][r][2] + sg * pbo * ph[u][r][3];
                        sv = (jm == j) ? qa : qa + qb;
                        dv = (jm == j) ? qa : qa - qb;
                    }
                    svr[u][r] = sv;
                    dvr[u][r] = dv;
                    if (!AL && j < g.hq) {
                        S[j * XS + 16 * u + lc] = sv;
                        D[j * XS + 16 * u + lc] = dv;
                    }
                }
            }
        }
        double xo[NT][4];
        if (AL) {
            if (wave < nbt) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int l = ang_bwd_row<T>(pr, m, wave, r);
#pragma unroll
                    for (int u = 0; u < NT; ++u) xo[u][r] = l >= 0 ? X[l * XS + 16 * u + lc] : 0.0;
                }
            }
            __syncthreads();
            if (wave < mt) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = wave * 16 + mfma_t<T>::row(kr, r);
                    if (j < g.hq) {
#pragma unroll
                        for (int u = 0; u < NT; ++u) {
                            S[j * XS + 16 * u + lc] = svr[u][r];
                            D[j * XS + 16 * u + lc] = dvr[u][r];
                        }
                    }
                }
            }
        }
        __syncthreads();
        double v[NT][4];
        if (wave < nbt) ang_bwd_sum<T, NT>(Bs, Ba, S, D, g, pr, wave, v);
        if (AL) {
            __syncthreads();
            const int npad = g.Nlx - g.Nl;
            for (int idx = t; idx < (m + npad) * XS; idx += nthr) {
                const int row = idx / XS < m ? idx / XS : g.Nl + idx / XS - m;
                X[row * XS + idx % XS] = 0.0;
            }
        }
        if (wave < nbt) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int l = ang_bwd_row<T>(pr, m, wave, r);
                if (l >= 0) {
#pragma unroll
                    for (int u = 0; u < NT; ++u) {
                        double &x = X[l * XS + 16 * u + lc];
                        // mixed form: added to the old values (in registers when S/D overlay X)
                        x = !mfma_t<T>::mixed ? v[u][r] : (AL ? xo[u][r] : x) + v[u][r];
                    }
                }
            }
        }
        __syncthreads();
    }

    if (MODE == 0) {
        ang_laser<NT>(X, 0, m, g.Nl, a_i, cl);
        __syncthreads();
        ang_laser<NT>(X, 1, m, g.Nl, a_i, cl);
        __syncthreads();
    }

    // the density phase's r_i and chi_j, loaded ahead of the state stores: one
    // vector-memory counter covers loads and stores, so a load issued after the
    // stores would also wait for all of them
    double rdn[NT], cdn[4];
    if (MODE != 0) {
#pragma unroll
        for (int u = 0; u < NT; ++u) rdn[u] = oku[u] ? g.r[iu[u]] : 1.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = wave * 16 + mfma_t<T>::row(kr, r);
            cdn[r] = g.chi[j < g.h ? j : 0];
        }
    }
    if (MODE != 2) {
        const int lref = (MODE == 0 && g.psi32 && g.l0) ? g.l0[korb] : -1;
        for (int idx = t; idx < g.Nl * R; idx += nthr) {
            const int l = idx / R, il = idx % R;
            const int ii = i0 + il;
            if (ii < g.N) {
                const size_t o = ((size_t)l * g.C + c) * g.N + ii;
                const double xr = X[l * XS + 2 * il], xi = X[l * XS + 2 * il + 1];
                ((double2 *)g.psi)[o] = make_double2(xr, xi);
                if (MODE == 0 && g.psi32 && l != lref) g.psi32[o] = make_float2((float)xr, (float)xi);
            }
        }
    }
    if (MODE == 0) return;

    // |B psi|^2 and observable partials (accumulated in FP64)
    double sd = 0.0, sa = 0.0, sn = 0.0;
    if (wave < mt) {
        v4 e[NT], o[NT];
        ang_fwd<T, NT>(Bs, Ba, X, g, pr, m, wave, e, o);
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const double ri = rdn[u];
            const double ir2 = 1.0 / (ri * ri);
            // folded |Psi|^2 rows: lanes p = 0 store the nodes j, lanes p = 1 their mirrors
            T *dens = reinterpret_cast<T *>(g.dens) + (((size_t)c * g.N + (oku[u] ? iu[u] : 0)) * 2 + p) * g.hq;
            T dq[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = wave * 16 + mfma_t<T>::row(kr, r);
                const int jm = g.Nth - 1 - j;
                const double ed = (double)e[u][r], od = (double)o[u][r];
                const double va = ed + od, vb = ed - od;
                const double s2a = va * va + lane_xor1(va * va);
                const double s2b = vb * vb + lane_xor1(vb * vb);
                const double eo = 4.0 * (ed * od + lane_xor1(ed * od));  // |Psi_j|^2 - |Psi_j'|^2
                // lane p = 1 takes its partner's s2b, so that the mirror value is the
                // one lane p = 0 forms (a contracted multiply-add is not symmetric)
                const double s2bp = lane_xor1(s2b);
                dq[r] = (T)(p ? s2bp : s2a);
                if (p == 0 && oku[u] && j < g.h) {
                    if (jm != j) {
                        const double ca = cdn[r];
                        sd += ri * ca * eo;
                        sa += ca * eo * ir2;
                        sn += s2a + s2b;
                    } else {
                        sn += s2a;  // middle node, chi = 0
                    }
                }
            }
            // rows h .. hq - 1 hold values that k_rho never reads, and so does the
            // mirror half at the middle node of an odd N_theta
            if (oku[u]) {
                if constexpr (mfma_t<T>::mixed) {
                    // this lane's four items are the consecutive rows 4 kr .. 4 kr + 3
                    // of its tile: one 16-byte store (hq is a multiple of 4)
                    const int j0 = wave * 16 + 4 * kr;
                    if (j0 < g.hq) *reinterpret_cast<float4 *>(dens + j0) = make_float4(dq[0], dq[1], dq[2], dq[3]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = wave * 16 + mfma_t<T>::row(kr, r);
                        if (j < g.h) dens[j] = dq[r];
                    }
                }
            }
        }
    }
    // work-group reduction, fixed order
    for (int off = 32; off > 0; off >>= 1) {
        sd += __shfl_down(sd, off);
        sa += __shfl_down(sa, off);
        sn += __shfl_down(sn, off);
    }
    if (lane == 0) {
        red[0][wave] = sd;
        red[1][wave] = sa;
        red[2][wave] = sn;
    }
    // barrier over LDS only: the state and |Psi|^2 stores above may still be in
    // flight (a fenced barrier would wait for every one of them to complete)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wave's partial sums are in LDS
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (t < 3) {
        double s = 0.0;
        for (int w = 0; w < nthr / 64; ++w) s += red[t][w];
        g.obs_part[((size_t)c * gridDim.x + tile_i) * 3 + t] = s;
    }
}

// per-column sums of the observable partials -> obs[step_rel][c][3]
struct ObsArgs {
    const double *part;  // [C][ntiles][3]
    double *obs;         // destination rows, or null (no reduction)
    int C, ntiles;
    const int *cur, *origin;  // graph replay: row = cur[0] - origin[0] + local
    int local;
};
// one wave per (column, observable) pair idx; lane t sums tiles t, t+64, ...
// then a fixed shuffle tree, so the result does not depend on timing
static __device__ void obs_reduce_pair(const ObsArgs &o, int idx) {
    const int c = idx / 3, k = idx % 3;
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int tl = lane; tl < o.ntiles; tl += 64) s += o.part[((size_t)c * o.ntiles + tl) * 3 + k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) {
        const size_t row = o.cur ? (size_t)(o.cur[0] - o.origin[0] + o.local) : 0;
        o.obs[row * o.C * 3 + idx] = s;
    }
}
__global__ __launch_bounds__(64) void k_obs_reduce(ObsArgs o) { obs_reduce_pair(o, blockIdx.x); }

// Per-(column, l) channel dipole of the state after the step,
//   chan[row][c][l] = <x_l | r (c(l, m) x_{l+1} + c(l-1, m) x_{l-1})>
// (conjugate on the left; dipoleMoment.cuh:1-47 reduced as by cublasZdotc,
// :71), one wave per (l, column), fixed-order reduction.  Rows l < m and idle
// instances hold zero.
struct ChanArgs {
    const double *psi, *r;
    const int32_t *mcol, *active;
    double2 *chan;               // [rows][C][Nl]
    int N, Nl, C, n_orb;
    const int *cur, *origin;     // graph replay: row = cur[0] - origin[0] + local
    int local;
};
__global__ __launch_bounds__(64) void k_chan_dip(ChanArgs a) {
    const int l = blockIdx.x, c = blockIdx.y, lane = threadIdx.x;
    const int m = a.mcol[c % a.n_orb];
    const size_t row = a.cur ? (size_t)(a.cur[0] - a.origin[0] + a.local) : (size_t)a.local;
    double sr = 0.0, si = 0.0;
    if (l >= m && !(a.active && !a.active[c / a.n_orb])) {
        const size_t plane = (size_t)a.C * a.N;  // one l of the state, complex values
        const double2 *x = reinterpret_cast<const double2 *>(a.psi) + (size_t)l * plane + (size_t)c * a.N;
        const bool up = l + 1 < a.Nl, dn = l - 1 >= m;
        const double cu = up ? c_lm(l, m) : 0.0, cd = dn ? c_lm(l - 1, m) : 0.0;
        const double2 *xu = up ? x + plane : x;
        const double2 *xd = dn ? x - plane : x;
        for (int i = lane; i < a.N; i += 64) {
            const double2 v = x[i], u = xu[i], d = xd[i];
            const double ri = a.r[i];
            const double zr = ri * (cu * u.x + cd * d.x), zi = ri * (cu * u.y + cd * d.y);
            sr += v.x * zr + v.y * zi;
            si += v.x * zi - v.y * zr;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_down(sr, off);
        si += __shfl_down(si, off);
    }
    if (lane == 0) a.chan[(row * a.C + c) * a.Nl + l] = make_double2(sr, si);
}

// advances the replayed graph's step counter
__global__ void k_bump(int *cur, int by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) cur[0] += by;
}

template <typename T> struct lda_t;
template <> struct lda_t<double> {
    static __device__ double cbrt_(double x) { return cbrt(x); }
    static __device__ double sqrt_(double x) { return sqrt(x); }
    static __device__ double log1p_(double x) { return log1p(x); }
    static constexpr double rho_min = RHO_MIN;
};
template <> struct lda_t<float> {
    static __device__ float cbrt_(float x) { return cbrtf(x); }
    static __device__ float sqrt_(float x) { return sqrtf(x); }
    static __device__ float log1p_(float x) { return log1pf(x); }
    static constexpr float rho_min = 1e-30f;  // FP32: dropped exchange below 1e-10 Hartree
};

// Slater exchange + Perdew-Wang 1992 correlation, spin unpolarised, in T
template <typename T>
static __device__ T lda_vxc(T rho) {
    typedef lda_t<T> F;
    if (!(rho > (T)F::rho_min)) return (T)0;
    const T A = (T)0.031091, a1 = (T)0.21370, b1 = (T)7.5957, b2 = (T)3.5876, b3 = (T)1.6382,
            b4 = (T)0.49294, pi = (T)M_PI;
    const T vx = -F::cbrt_((T)3 * rho / pi);
    const T rs = F::cbrt_((T)3 / ((T)4 * pi * rho));
    const T srs = F::sqrt_(rs);
    const T q0 = (T)-2 * A * ((T)1 + a1 * rs);
    const T q1 = (T)2 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
    const T dq1 = A * (b1 / srs + (T)2 * b2 + (T)3 * b3 * srs + (T)4 * b4 * rs);
    const T lg = F::log1p_((T)1 / q1);
    const T eps = q0 * lg;
    const T deps = (T)-2 * A * a1 * lg - q0 * dq1 / (q1 * q1 + q1);
    return vx + eps - rs / (T)3 * deps;
}

// Density and LDA v_xc (in T) pointwise on the (r_i, theta_j) grid:
// rho_u(i, j) = sum_k occ_k |Psi_k(i, j)|^2 (+ the frozen core), kept for the
// multipole moments, and v_xc of the density rho = rho_u / (2 pi omega_j wr_i r_i^2).
template <typename T>
__global__ __launch_bounds__(WG) void k_rho(const T *__restrict__ dens,
                                            const double *__restrict__ occ,
                                            const double *__restrict__ rho_frozen,
                                            const double *__restrict__ r,
                                            const double *__restrict__ wr,
                                            const double *__restrict__ omega,
                                            double *__restrict__ rho_u, double *__restrict__ vxc,
                                            int N, int Nth, int h, int hq, int n_orb, ObsArgs ob,
                                            const int32_t *__restrict__ active) {
    // the blocks past the (r, theta) grid (instance row 0) reduce the observable
    // partials of the step, one (column, observable) pair per wave
    const int nrho = (int)(((size_t)N * Nth + WG - 1) / WG);
    if ((int)blockIdx.x >= nrho) {
        const int pair = ((int)blockIdx.x - nrho) * (WG / 64) + (threadIdx.x >> 6);
        if (blockIdx.y == 0 && ob.obs && pair < ob.C * 3) obs_reduce_pair(ob, pair);
        return;
    }
    const int inst = blockIdx.y;
    if (active && !active[inst]) return;  // idle serving slot (block-uniform)
    const int npt = N * Nth;  // < 2^31 (ggb_create bounds N and N_theta)
    const int idx = (int)blockIdx.x * WG + (int)threadIdx.x;
    if (idx >= npt) return;
    const int i = idx / Nth, j = idx - i * Nth;
    // the folded |Psi|^2 rows: node j < h in the first half, its mirror in the second
    const size_t nd = (size_t)N * 2 * hq;  // one orbital column
    const T *dk = dens + (size_t)inst * n_orb * nd + ((size_t)i * 2 + (j < h ? 0 : 1)) * hq + (j < h ? j : Nth - 1 - j);
    // four orbitals' loads in flight at a time, summed in orbital order
    double ru = 0.0;
    int k = 0;
    for (; k + 4 <= n_orb; k += 4) {
        const T d0 = dk[k * nd], d1 = dk[(k + 1) * nd];
        const T d2 = dk[(k + 2) * nd], d3 = dk[(k + 3) * nd];
        ru += occ[k] * d0;
        ru += occ[k + 1] * d1;
        ru += occ[k + 2] * d2;
        ru += occ[k + 3] * d3;
    }
    for (; k < n_orb; ++k) ru += occ[k] * dk[k * nd];
    if (rho_frozen) ru += 0.5 * rho_frozen[i] * omega[j];
    const double rho = ru / (2.0 * M_PI * omega[j] * wr[i] * r[i] * r[i]);
    const size_t o = (size_t)inst * N * Nth + idx;
    rho_u[o] = ru;
    vxc[o] = (double)lda_vxc<T>((T)rho);
}

// One 16 x 16 tile of D = A B in FP64 on the matrix cores per work-group
// (tile (blockIdx.y, blockIdx.x), instance blockIdx.z), the K range split
// over the four waves and the partial tiles added in fixed order.
// A[m][k] = A[z saz + m sam + k sak], B[k][n] = B[z sbz + k sbk + n sbn], zero
// outside M x K and K x N.  Wave 0 hands each element to epi(z, m, n, value).
struct GemmOperands {
    const double *A, *B;
    long sam, sak, saz, sbk, sbn, sbz;
    int M, N, K;
    const int32_t *active;  // [z] 0: instance idle, skipped (null: all)
};
#ifndef TG_KU
#define TG_KU 9
#endif
template <class Epi>
__global__ __launch_bounds__(WG) void k_tile_gemm(GemmOperands g, Epi epi) {
    __shared__ double red[4][64][4];
    const tile3 tb = xcd_tile();  // n tiles of one m tile run on one XCD and share A in its L2
    const int z = tb.z;
    if (g.active && !g.active[z]) return;  // block-uniform
    const int m0 = tb.y * 16, n0 = tb.x * 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int m = m0 + lc, n = n0 + lc;
    const bool mok = m < g.M, nok = n < g.N;
    const double *Ar = g.A + z * g.saz + (long)(mok ? m : 0) * g.sam;
    const double *Bc = g.B + z * g.sbz + (long)(nok ? n : 0) * g.sbn;
    const int kq = (((g.K + 3) / 4 + 3) / 4) * 4;  // quarter of K, multiple of 4
    const int kb = wave * kq, ke = min(g.K, kb + kq);
    // wave w finishes accumulator row r = w: the epilogue's other inputs of
    // that element are loaded ahead of the products (in-range indices)
    const int mo = m0 + mfma_t<double>::row(kr, wave);
    const typename Epi::Pre pre = epi.pre(z, mo < g.M ? mo : 0, nok ? n : 0);
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    // TG_KU loads of each operand in flight per lane; 9 covers a quarter of
    // K = 129 (N_theta or the multipole count at L_max = 64) in one round
    for (int k0 = kb; k0 < ke; k0 += 4 * TG_KU) {
        double av[TG_KU], bv[TG_KU];
#pragma unroll
        for (int u = 0; u < TG_KU; ++u) {
            const int k = k0 + 4 * u + kr;
            const bool ok = k < ke;
            av[u] = (ok && mok) ? Ar[(long)k * g.sak] : 0.0;
            bv[u] = (ok && nok) ? Bc[(long)k * g.sbk] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < TG_KU; ++u) acc = mfma64(av[u], bv[u], acc);
    }
    // partial tiles through LDS (wave 0's in the last slot); wave w < 4 then finishes
    // accumulator row r = w, so the four epilogues run side by side
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wave > 0 ? wave - 1 : 3][lane][r] = acc[r];
    __syncthreads();
    const int r = wave;
    const double v = ((red[3][lane][r] + red[0][lane][r]) + red[1][lane][r]) + red[2][lane][r];
    if (mo < g.M && nok) epi.apply(z, mo, n, v, pre);
}

// The same products with one operand panel reused across tiles: a work-group
// takes one 16-wide tile of the narrow dimension and up to TGW_T tiles of the
// wide one (WIDE_M: along M, A differs per tile and B is shared; else along
// N).  The K range is split over the four waves exactly as in k_tile_gemm and
// each wave keeps one accumulator tile per wide tile, so every element is
// formed by the same products in the same order and the partial tiles are
// added in the same order: the results equal k_tile_gemm's bit for bit.  The
// shared operand is loaded once per K step instead of once per tile, and the
// epilogue's other inputs (Epi::pre) are loaded before the products.  Measured
// at 64 argon instances (129 multipoles and theta nodes): 3 tiles per
// work-group beat 2, 4, 5 and 9 (fewer work-groups, more registers, lower
// occupancy), and 3 K steps per round of loads beat 2 and 5.
// grid (narrow tiles, groups of TGW_T wide tiles, z)
#ifndef TGW_T
#define TGW_T 3   // wide tiles per work-group
#endif
#ifndef TGW_U
#define TGW_U 3   // K steps per round of loads
#endif
#ifndef TGW_G
#define TGW_G 3   // tiles per reduction round through LDS
#endif
#ifndef TGW_MIN_INST
#define TGW_MIN_INST 8  // instances from which the potential products use the panel form
#endif
template <class Epi, bool WIDE_M>
__global__ __launch_bounds__(WG) void k_tile_gemm_wide(GemmOperands g, Epi epi) {
    __shared__ double red[4][TGW_G][64][4];
    const int z = blockIdx.z;
    if (g.active && !g.active[z]) return;  // block-uniform
    const int s0 = blockIdx.x * 16;            // narrow tile origin
    const int w0 = blockIdx.y * TGW_T * 16;    // first wide tile origin
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kr = lane >> 4, lc = lane & 15;
    const int Wn = WIDE_M ? g.M : g.N;         // extent of the wide dimension
    const int nt = min(TGW_T, (Wn - w0 + 15) / 16);
    // the narrow operand (B if WIDE_M, else A): one row/column per lane
    const int sidx = s0 + lc;
    const bool sok = sidx < (WIDE_M ? g.N : g.M);
    const double *Sp = WIDE_M ? g.B + z * g.sbz + (long)(sok ? sidx : 0) * g.sbn
                              : g.A + z * g.saz + (long)(sok ? sidx : 0) * g.sam;
    const long ssk = WIDE_M ? g.sbk : g.sak;
    // the wide operand: lane row/column lc of tile t at w0 + 16 t + lc (offsets
    // within one z fit in 32 bits)
    const double *Wp = (WIDE_M ? g.A + z * g.saz : g.B + z * g.sbz) + (long)w0 * (WIDE_M ? g.sam : g.sbn);
    const int wsi = (int)(WIDE_M ? g.sam : g.sbn), wsk = (int)(WIDE_M ? g.sak : g.sbk);
    const int kq = (((g.K + 3) / 4 + 3) / 4) * 4;  // quarter of K, multiple of 4 (as k_tile_gemm)
    const int kb = wave * kq, ke = min(g.K, kb + kq);
    // wave w finishes accumulator row r = w of every tile: element (m, n) of
    // tile t, whose other epilogue inputs are loaded now, under the products
    const int r = wave;
    const int ro = mfma_t<double>::row(kr, r);
    typename Epi::Pre pre[TGW_T];
#pragma unroll
    for (int t = 0; t < TGW_T; ++t) {
        const int m = WIDE_M ? w0 + 16 * t + ro : s0 + ro;
        const int n = WIDE_M ? s0 + lc : w0 + 16 * t + lc;
        if (t < nt && m < g.M && n < g.N) pre[t] = epi.pre(z, m, n);
    }
    d4_t acc[TGW_T];
#pragma unroll
    for (int t = 0; t < TGW_T; ++t) acc[t] = (d4_t){0.0, 0.0, 0.0, 0.0};
    for (int k0 = kb; k0 < ke; k0 += 4 * TGW_U) {
        double sv[TGW_U], wv[TGW_T][TGW_U];
#pragma unroll
        for (int u = 0; u < TGW_U; ++u) {
            const int k = k0 + 4 * u + kr;
            const bool ok = k < ke;
            sv[u] = (ok && sok) ? Sp[(long)k * ssk] : 0.0;
#pragma unroll
            for (int t = 0; t < TGW_T; ++t) {
                const int wl = 16 * t + lc;  // index within the group
                wv[t][u] = (ok && t < nt && w0 + wl < Wn) ? Wp[wl * wsi + k * wsk] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < TGW_U; ++u)
#pragma unroll
            for (int t = 0; t < TGW_T; ++t)
                acc[t] = WIDE_M ? mfma64(wv[t][u], sv[u], acc[t]) : mfma64(sv[u], wv[t][u], acc[t]);
    }
    // partial tiles through LDS, TGW_G tiles per round (wave 0's in the last
    // slot); wave w then finishes accumulator row r = w of each tile
#pragma unroll
    for (int t0 = 0; t0 < TGW_T; t0 += TGW_G) {
        if (t0 >= nt) break;  // block-uniform
#pragma unroll
        for (int q = 0; q < TGW_G; ++q)
            if (t0 + q < TGW_T)
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) red[wave > 0 ? wave - 1 : 3][q][lane][rr] = acc[t0 + q][rr];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < TGW_G; ++q) {
            const int t = t0 + q;
            if (t < TGW_T && t < nt) {
                const double v = ((red[3][q][lane][r] + red[0][q][lane][r]) + red[1][q][lane][r]) + red[2][q][lane][r];
                const int m = WIDE_M ? w0 + 16 * t + ro : s0 + ro;
                const int n = WIDE_M ? s0 + lc : w0 + 16 * t + lc;
                if (m < g.M && n < g.N) epi.apply(z, m, n, v, pre[t]);
            }
        }
        __syncthreads();
    }
}

// multipole moments q_L(i) = sum_j P_L(chi_j) rho_u(i, j): M = L, N = i, K = j
struct QmomEpi {
    double *q;
    int N, n_har;
    struct Pre {};
    __device__ Pre pre(int, int, int) const { return Pre(); }
    __device__ void apply(int z, int L, int i, double v, const Pre &) const {
        q[((size_t)z * n_har + L) * N + i] = v;
    }
    __device__ void operator()(int z, int L, int i, double v) const { apply(z, L, i, v, Pre()); }
};

// ---------------------------------------------------------------------------
// Hartree multipole V_L(r_a) = sum_b q_L(b) r_<^L / r_>^{L+1} by two affine
// scans (ratio form, no overflow):
//   A_a = A_{a-1} (r_{a-1}/r_a)^{L+1} + q_a / r_a
//   B_a = (r_a/r_{a+1})^L (B_{a+1} + q_{a+1} / r_{a+1}),   V = A + B
// with the ratio powers tabulated on the host as exp(n log(r_{a-1}/r_a)).
// One work-group per (L, instance): each thread composes the affine maps of
// a contiguous chunk of radial nodes; the chunk maps are scanned within each
// wave by shuffles and across the four waves through LDS (one barrier), in a
// fixed order; each thread then re-runs its chunk from its start value.
// ---------------------------------------------------------------------------
// inclusive scans of affine maps x -> a x + b over the work-group, both at
// once: (al, be) with the maps of lower threads applied first, (sa, sb) with
// those of higher threads first.  The two share each shuffle round and each
// barrier, and each keeps the operation order of a scan run alone.
static __device__ void affine_scan2(double &al, double &be, double &sa, double &sb, double (*sm)[4]) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const double pa = __shfl_up(al, off), pb = __shfl_up(be, off);
        const double qa = __shfl_down(sa, off), qb = __shfl_down(sb, off);
        if (lane >= off) {
            be = pb * al + be;
            al = pa * al;
        }
        if (lane + off < 64) {
            sb = qb * sa + sb;
            sa = qa * sa;
        }
    }
    if (lane == 63) {
        sm[wave][0] = al;
        sm[wave][1] = be;
    }
    if (lane == 0) {
        sm[wave][2] = sa;
        sm[wave][3] = sb;
    }
    __syncthreads();
    // maps of the whole waves before this one (forward) or after it (backward), nearest last
    double ca = 1.0, cb = 0.0;
    for (int w = 0; w < wave; ++w) {
        cb = cb * sm[w][0] + sm[w][1];
        ca = ca * sm[w][0];
    }
    be = cb * al + be;
    al = ca * al;
    double da = 1.0, db = 0.0;
    for (int w = WG / 64 - 1; w > wave; --w) {
        db = db * sm[w][2] + sm[w][3];
        da = da * sm[w][2];
    }
    sb = db * sa + sb;
    sa = da * sa;
}

// hp: the ratio powers and 1/r, from the host (shared by all instances):
// hp[L][a] = (r_{a-1}/r_a)^{L+1} (0 at a = 0), hp[n_har + L][a] = (r_a/r_{a+1})^L
// (a < N - 1), hp[2 n_har][a] = 1/r_a
// CS: radial nodes per thread at most (N <= CS * WG); the thread's alpha,
// gamma, q and 1/r are loaded once into registers (one round of loads) and
// reused by the four passes, which keep the arithmetic order of the scans.
template <int CS>
__global__ __launch_bounds__(WG) void k_hartree(const double *__restrict__ q,
                                                const double *__restrict__ hp,
                                                double *__restrict__ V, int N, int n_har,
                                                const int32_t *__restrict__ active) {
    __shared__ double sm[WG / 64][4];
    __shared__ double edge[2][WG];
    const int L = blockIdx.x, inst = blockIdx.y;
    if (active && !active[inst]) return;  // idle serving slot (block-uniform)
    const int t = threadIdx.x;
    const double *qL = q + ((size_t)inst * n_har + L) * N;
    double *VL = V + ((size_t)inst * n_har + L) * N;
    const double *alpha = hp + (size_t)L * N;
    const double *gamma = hp + ((size_t)n_har + L) * N;
    const double *ir = hp + (size_t)2 * n_har * N;
    const int cs = (N + WG - 1) / WG;
    const int s0 = min(N, t * cs), s1 = min(N, s0 + cs);
    const int n = s1 - s0;  // nodes of this thread, <= CS
    double ra[CS], rg[CS], rq[CS + 1], ri[CS + 1];
#pragma unroll
    for (int c = 0; c < CS; ++c) {
        const bool in = c < n;
        ra[c] = in ? alpha[s0 + c] : 0.0;
        rg[c] = in ? gamma[s0 + c] : 0.0;
        rq[c] = in ? qL[s0 + c] : 0.0;
        ri[c] = in ? ir[s0 + c] : 0.0;
    }
    // q and 1/r of the node after the chunk (the suffix of its last node)
    const double qn = (n > 0 && s1 < N) ? qL[s1] : 0.0, irn = (n > 0 && s1 < N) ? ir[s1] : 0.0;
    rq[CS] = 0.0;
    ri[CS] = 0.0;
    // ---- prefix: A_a = alpha_a A_{a-1} + beta_a
    double al = 1.0, be = 0.0;
#pragma unroll
    for (int c = 0; c < CS; ++c)
        if (c < n) {
            al *= ra[c];
            be = be * ra[c] + rq[c] * ri[c];
        }
    // ---- suffix: B_a = gamma_a B_{a+1} + gamma_a q_{a+1}/r_{a+1}, B_{N-1} = 0
    double sa = 1.0, sb = 0.0;
#pragma unroll
    for (int c = CS - 1; c >= 0; --c)
        if (c < n) {
            if (s0 + c == N - 1) {
                sa = 0.0;
                sb = 0.0;
                continue;
            }
            const bool nx = c + 1 < n;
            const double q1 = nx ? rq[c + 1] : qn, i1 = nx ? ri[c + 1] : irn;
            const double gam = rg[c];
            sa *= gam;
            sb = sb * gam + gam * q1 * i1;
        }
    affine_scan2(al, be, sa, sb, sm);
    edge[0][t] = be;  // A at the end of this chunk
    edge[1][t] = sb;  // B at the start of this chunk
    __syncthreads();
    double A = t > 0 ? edge[0][t - 1] : 0.0;
    double va[CS];
#pragma unroll
    for (int c = 0; c < CS; ++c) {
        if (c < n) A = A * ra[c] + rq[c] * ri[c];
        va[c] = A;
    }
    double Bv = t < WG - 1 ? edge[1][t + 1] : 0.0;
#pragma unroll
    for (int c = CS - 1; c >= 0; --c)
        if (c < n) {
            if (s0 + c == N - 1) {
                Bv = 0.0;
            } else {
                const bool nx = c + 1 < n;
                const double q1 = nx ? rq[c + 1] : qn, i1 = nx ? ri[c + 1] : irn;
                Bv = rg[c] * (Bv + q1 * i1);
            }
            VL[s0 + c] = va[c] + Bv;
        }
}

// dV = sum_L V_L(r_i) P_L(chi_j) + v_xc - v_gs and phase = mask e^{-i dV dt/2}:
// the Legendre sum as a tile product (M = i, N = j, K = L), this epilogue.
// The mixed form stores float2 (f - 1): the angular kernel applies the
// potential half step as X += B^T ((f - 1) B X) and only needs that difference,
// formed here in FP64 and rounded once.
struct DvEpi {
    const double *vxc, *v_gs, *mask;
    double *dv;
    void *phase;
    int N, Nth, interacting, mixed;
    double dt;
    // the element's other inputs, loadable before its product is done
    struct Pre {
        double x, g, mk;
    };
    __device__ Pre pre(int z, int i, int j) const {
        const size_t o = ((size_t)z * N + i) * Nth + j;
        Pre r;
        r.x = interacting ? vxc[o] : 0.0;
        r.g = v_gs[i];
        r.mk = mask[i];
        return r;
    }
    __device__ void operator()(int z, int i, int j, double v) const { apply(z, i, j, v, pre(z, i, j)); }
    __device__ void apply(int z, int i, int j, double v, const Pre &pr) const {
        const size_t o = ((size_t)z * N + i) * Nth + j;
        const double sv = interacting ? v + pr.x - pr.g : 0.0;
        dv[o] = sv;
        double sn, cs;
        sincos(0.5 * dt * sv, &sn, &cs);
        const double fr = pr.mk * cs, fi = -pr.mk * sn;
        if (mixed)
            reinterpret_cast<float2 *>(phase)[o] = make_float2((float)(fr - 1.0), (float)fi);
        else
            reinterpret_cast<double2 *>(phase)[o] = make_double2(fr, fi);
    }
};

// ---------------------------------------------------------------------------
// HHG spectrum on the device (analysis.py:447-460, internal.py:361-379):
// Kaiser window (sym = False), optional 8th-order central second difference,
// zero padding by n_pad * n on each side (M = (2 n_pad + 1) n points),
// |fftshift(fft(fftshift(x)))|^2, and the super-Gaussian harmonic masks of
// argonInSitu/arg.py:42-67.  M is any length: the DFT is Bluestein's
// algorithm, X_k = w_k sum_j (y_j w_j) conj(w_{k-j}), w_j = e^{-i pi j^2 / M},
// with the convolution done by radix-2 Stockham FFTs of length L >= 2M - 1.
// ---------------------------------------------------------------------------
static __device__ double bessel_i0(double x) {
    // power series sum_k ((x/2)^2)^k / (k!)^2, all terms positive
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 200; ++k) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

// x_j = signal_j * I0(beta sqrt(1 - (2j/n - 1)^2)) / I0(beta)   (scipy kaiser, sym=False)
__global__ __launch_bounds__(WG) void k_spec_window(const double *__restrict__ sig,
                                                    double *__restrict__ x, int n, double beta) {
    const int j = blockIdx.x * WG + threadIdx.x;
    if (j >= n) return;
    const double t = 2.0 * (double)j / (double)n - 1.0;
    x[j] = sig[j] * bessel_i0(beta * sqrt(fmax(0.0, 1.0 - t * t))) / bessel_i0(beta);
}

// 8th-order central second difference, zero on the 4 edge points at each end
__global__ __launch_bounds__(WG) void k_spec_fd8(const double *__restrict__ x,
                                                 double *__restrict__ y, int n, double inv_dt2) {
    const int j = blockIdx.x * WG + threadIdx.x;
    if (j >= n) return;
    double v = 0.0;
    if (n > 8 && j >= 4 && j < n - 4) {
        const double c[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
        v = c[0] * x[j];
        for (int s = 1; s < 5; ++s) v += c[s] * (x[j + s] + x[j - s]);
    }
    y[j] = v * inv_dt2;
}

// e^{sign i pi j^2 / M}, with j^2 reduced mod 2M exactly
static __device__ double2 chirp(long j, long M, double sign) {
    const long r = (j * j) % (2 * M);
    double s, c;
    sincospi(sign * (double)r / (double)M, &s, &c);
    return make_double2(c, s);
}

// a_j = y_j w_j for j < M (y = fftshift of the padded signal), zero up to L
__global__ __launch_bounds__(WG) void k_blue_a(const double *__restrict__ x, double2 *__restrict__ a,
                                               int n, int n_pad, long M, long L) {
    const long j = (long)blockIdx.x * WG + threadIdx.x;
    if (j >= L) return;
    double2 v = make_double2(0.0, 0.0);
    if (j < M) {
        const long i = (j - M / 2 + M) % M;       // index into the padded signal
        const long k = i - (long)n_pad * n;       // index into x
        if (k >= 0 && k < n) {
            const double2 w = chirp(j, M, -1.0);
            v = make_double2(x[k] * w.x, x[k] * w.y);
        }
    }
    a[j] = v;
}

// b_j = conj(w_j) for |j| < M, wrapped to length L
__global__ __launch_bounds__(WG) void k_blue_b(double2 *__restrict__ b, long M, long L) {
    const long j = (long)blockIdx.x * WG + threadIdx.x;
    if (j >= L) return;
    double2 v = make_double2(0.0, 0.0);
    if (j < M) v = chirp(j, M, 1.0);
    else if (j > L - M) v = chirp(L - j, M, 1.0);
    b[j] = v;
}

// one radix-2 Stockham pass (sub-transform size Ns -> 2 Ns), natural order at the end
__global__ __launch_bounds__(WG) void k_fft_pass(const double2 *__restrict__ in,
                                                 double2 *__restrict__ out, long L, long Ns,
                                                 double sign) {
    const long j = (long)blockIdx.x * WG + threadIdx.x;
    if (j >= L / 2) return;
    const long k = j % Ns;
    double s, c;
    sincospi(sign * (double)k / (double)Ns, &s, &c);
    const double2 v0 = in[j], u = in[j + L / 2];
    const double2 v1 = make_double2(u.x * c - u.y * s, u.x * s + u.y * c);
    const long d = (j / Ns) * 2 * Ns + k;
    out[d] = make_double2(v0.x + v1.x, v0.y + v1.y);
    out[d + Ns] = make_double2(v0.x - v1.x, v0.y - v1.y);
}

__global__ __launch_bounds__(WG) void k_cmul(double2 *__restrict__ a, const double2 *__restrict__ b,
                                             long L) {
    const long j = (long)blockIdx.x * WG + threadIdx.x;
    if (j >= L) return;
    const double2 u = a[j], v = b[j];
    a[j] = make_double2(u.x * v.x - u.y * v.y, u.x * v.y + u.y * v.x);
}

// power[q] = |X_{(q - M/2) mod M}|^2, X_k = w_k C_k / L
__global__ __launch_bounds__(WG) void k_blue_power(const double2 *__restrict__ C,
                                                   double *__restrict__ power, long M, long L) {
    const long q = (long)blockIdx.x * WG + threadIdx.x;
    if (q >= M) return;
    const long k = (q - M / 2 + M) % M;
    const double2 w = chirp(k, M, -1.0), cv = C[k];
    const double xr = (w.x * cv.x - w.y * cv.y) / (double)L, xi = (w.x * cv.y + w.y * cv.x) / (double)L;
    power[q] = xr * xr + xi * xi;
}

// intensity of harmonic `order`: sum_q exp(-((q - i0) / (0.25 (ip - im)))^32) S_q, with
// i0, ip, im the grid points nearest to H, H + 1/2, H - 1/2 times omega0 (one work-group
// per order, fixed-order reduction over the band where the mask is not negligible)
__global__ __launch_bounds__(WG) void k_harmonics(const double *__restrict__ power,
                                                  const double *__restrict__ orders,
                                                  double *__restrict__ out, long M, double dt,
                                                  double omega0) {
    __shared__ double red[WG / 64];
    const double H = orders[blockIdx.x];
    const double scale = (double)M * dt / (2.0 * M_PI);  // grid points per unit of omega
    auto nearest = [&](double w) {
        long q = (long)llround(w * scale) + M / 2;
        return q < 0 ? 0 : (q > M - 1 ? M - 1 : q);
    };
    const long i0 = nearest(H * omega0), ip = nearest((H + 0.5) * omega0),
               im = nearest((H - 0.5) * omega0);
    const double width = 0.25 * (double)(ip - im);
    double s = 0.0;
    if (width > 0.0) {
        const long half = (long)(2.0 * width) + 1;  // exp(-2^32) underflows beyond
        const long lo = i0 - half < 0 ? 0 : i0 - half, hi = i0 + half > M - 1 ? M - 1 : i0 + half;
        for (long q = lo + threadIdx.x; q <= hi; q += WG) {
            const double u = ((double)q - (double)i0) / width;
            const double u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u16 = u8 * u8;
            s += exp(-(u16 * u16)) * power[q];
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < WG / 64; ++w) t += red[w];
        out[blockIdx.x] = t;
    }
}

// plain streaming copy of known size (calibrates the HBM byte counters)
__global__ __launch_bounds__(WG) void k_copy(const double2 *__restrict__ src,
                                             double2 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n; i += (size_t)gridDim.x * WG)
        dst[i] = src[i];
}

__global__ void k_selftest_mfma(double *out) {
    const int lane = threadIdx.x;
    const double a = (double)(lane % 16) + 0.25 * (double)(lane / 16);
    const double b = 0.5 * (double)(lane / 16) - (double)(lane % 16);
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    acc = mfma64(a, b, acc);
    for (int r = 0; r < 4; ++r) out[mfma_t<double>::row(lane / 16, r) * 16 + lane % 16] = acc[r];
    f4_t accf = {0.f, 0.f, 0.f, 0.f};
    accf = mfma_t<float>::op((float)a, (float)b, accf);
    for (int r = 0; r < 4; ++r)
        out[256 + mfma_t<float>::row(lane / 16, r) * 16 + lane % 16] = (double)accf[r];
}

// ===========================================================================
// Host side
// ===========================================================================
// Sizes of the parity-split angular tables and the angular kernel's launch.
#define ANG_LDS_MAX 81920  // tables staged in LDS up to 80 KiB (two work-groups per CU on gfx950)
struct ang_geom {
    int h, hq, sS, sA, nblk, waves;
    size_t lds;  // dynamic LDS bytes
    bool glb;    // tables read from global memory
    bool alias;  // mixed form: S/D share the l-space tile's LDS
    int nt;      // column tiles (of RT radial nodes) per work-group
    int ntiles;  // work-groups along r
};
static ang_geom ang_geometry(const ggb_problem *p) {
    ang_geom a;
    const int es = p->precision ? 4 : 8;
    const int L = p->n_l - 1;
    const int ns0 = L / 2 + 1, na0 = (L + 1) / 2;  // m = 0 has the most l of each parity
    a.h = (p->n_theta + 1) / 2;
    a.hq = (a.h + 3) & ~3;
    a.sS = ((ns0 + 3) & ~3) + 1;  // odd row strides spread a tile's rows over the LDS banks
    a.sA = ((na0 + 3) & ~3) + 1;
    const int per16 = 16 / es;
    a.nblk = ((a.hq * (a.sS + a.sA) + per16 - 1) / per16) * per16;
    const int mt = (a.h + 15) / 16, nb = (ns0 + 15) / 16 + (na0 + 15) / 16;
    a.waves = mt > nb ? mt : nb;
    // mixed form: the folded sums S/D (FP32) fit in the FP64 l-space tile's
    // space whenever hq <= N_l + 6; the kernel parks the old tile values it
    // needs in registers meanwhile (one work-group less of LDS per CU)
    // mixed form (many columns): two column tiles per work-group, each table
    // value then feeds two products and the per-work-group work is shared by
    // twice the radial nodes
    a.nt = p->precision ? ANG_NT_MIXED : 1;
    static_assert(ANG_NT_MIXED == 1 || ANG_NT_MIXED == 2 || ANG_NT_MIXED == 4, "column tiles: 1, 2 or 4");
    a.ntiles = (p->n_r + RT * a.nt - 1) / (RT * a.nt);
    const size_t xbytes = (size_t)(p->n_l + 6) * 16 * a.nt * sizeof(double);
    const size_t sdbytes = (size_t)2 * a.hq * 16 * a.nt * es;
    a.alias = p->precision && sdbytes <= xbytes;
    const size_t tiles = a.alias ? xbytes : xbytes + sdbytes;
    const size_t stat = (RT * a.nt + ANG_MAX_NL + 3 * ANG_MAXW) * sizeof(double);
#ifndef ANG_FORCE_GLB
#define ANG_FORCE_GLB 0
#endif
    a.glb = ANG_FORCE_GLB || tiles + (size_t)a.nblk * es + stat > ANG_LDS_MAX;
    a.lds = a.glb ? tiles : tiles + (size_t)a.nblk * es;
    return a;
}

struct ggb_ctx {
    ggb_problem p;
    int C, ntiles;
    ang_geom ang;  // parity-split angular tables and launch
    hipStream_t stream;
    std::string err;
    bool have_static, have_orbitals;
    int n_field;
    bool field_ring;  // ggb_field_ring: field index = step mod n_field
    // device buffers
    double *psi, *Y, *Q, *QT, *B, *PL, *PLT, *chi, *omega, *r, *hpow, *wr, *v_gs, *mask, *occ, *rho_frozen;
    int32_t *nkeep, *mcol;
    int32_t *active;  // [n_inst] instances in use (ggb_set_active), all at creation
    double2 *kph;  // [n_l][nk_max] (cos, sin)(lam dt) of the kinetic propagator
    double *dens, *vxc, *rho_u, *q, *V, *dv, *obs_part, *obs, *E;
    double2 *phase;
    // mixed form: ground-state reference of each orbital and its overlaps
    int kin_variant;  // GGB_KIN_*
    bool have_ref;
    float2 *psi32;  // FP32 copy of the state for the tiled pass 1 (mixed form, C > 8)
    float *Qlo, *Ylo;  // precision 2: low FP32 parts of Q and of Y
    double *phi_ref, *eps;
    int32_t *l0;
    double4 *ab;
    size_t obs_cap;
    // optional per-(column, l) channel dipoles of every step (ggb_set_channels)
    bool channels;
    double2 *chan;   // [rows][C][n_l]
    size_t chan_cap; // complex values
    int chan_rows;   // rows written by the last ggb_run
    // optional hipGraph of graph_steps steps (ggb_set_graph)
    int graph_steps;
    hipGraphExec_t graph_exec;
    int *dcur;  // [0] = step counter of the replay, [1] = first step of the run call
    unsigned long gen;  // incremented whenever a captured step may no longer be valid
    // optional HIP-event timing of the step phases (ggb_set_timing)
    int timing;                     // 0 off, 1 kinetic only, 2 every phase
    std::vector<hipEvent_t> events; // GGB_NPHASE + 1 per step of a run call
    double phase_ms[GGB_NPHASE];
    long phase_n[GGB_NPHASE];
};

static std::string g_create_err;

#define HIPCHK(ctx, call)                                                               \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
            return GGB_E_HIP;                                                           \
        }                                                                               \
    } while (0)

static int fail(ggb_ctx *ctx, int code, const std::string &msg) {
    ctx->err = msg;
    return code;
}

// a captured graph holds raw buffer pointers: drop it whenever one may move
static void drop_graph(ggb_ctx *ctx) {
    ++ctx->gen;  // graphs of groups holding this context are captured again too
    if (ctx->graph_exec) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
}

template <typename T>
static hipError_t dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, n * sizeof(T));
}

extern "C" int ggb_abi_version(void) { return GGB_ABI_VERSION; }

extern "C" const char *ggb_last_error(const ggb_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

extern "C" void ggb_destroy(ggb_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->p.device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    void *bufs[] = {ctx->psi, ctx->Y, ctx->Q, ctx->QT, ctx->kph, ctx->B, ctx->PL, ctx->PLT, ctx->chi, ctx->omega,
                    ctx->r, ctx->hpow, ctx->wr, ctx->v_gs, ctx->mask, ctx->occ, ctx->rho_frozen,
                    ctx->nkeep, ctx->mcol, ctx->dens, ctx->vxc, ctx->rho_u, ctx->q, ctx->V, ctx->dv,
                    ctx->obs_part, ctx->obs, ctx->E, ctx->phase, ctx->phi_ref, ctx->eps,
                    ctx->l0, ctx->ab, ctx->psi32, ctx->active, ctx->chan, ctx->Qlo, ctx->Ylo};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->dcur) (void)hipFree(ctx->dcur);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

static size_t footprint_bytes(const ggb_problem *p, int ntiles) {
    const size_t N = p->n_r, Nl = p->n_l, C = (size_t)p->n_orb * p->n_inst, Nth = p->n_theta;
    const size_t nI = p->n_inst, nk = p->nk_max, nh = p->n_har;
    size_t d = 0;  // doubles
    d += Nl * C * N * 2;                 // psi
    const size_t es = p->precision ? 4 : 8;
    const size_t nblk = ang_geometry(p).nblk;
    // Y, Q, QT (not in precision 2), B; precision 2 adds the low parts of Q and Y
    size_t tbytes = (Nl * C * nk * 2 + (p->precision == 2 ? 1 : 2) * Nl * N * nk + (size_t)p->n_m * nblk) * es;
    if (p->precision == 2) tbytes += (Nl * N * nk + Nl * C * nk * 2) * sizeof(float);
    d += 2 * Nl * nk;                    // e^{-i lam dt}
    d += 2 * Nth * nh + 2 * Nth;         // PL, PLT, chi, omega
    d += 5 * N + p->n_orb;               // r, wr, v_gs, mask, rho_frozen, occ
    d += (2 * nh + 1) * N;               // Hartree ratio powers and 1/r
    d += C * N * 2 * (size_t)ang_geometry(p).hq;  // dens (folded rows)
    d += 3 * nI * N * Nth;               // vxc, rho_u, dv
    d += 2 * nI * N * Nth;               // phase (double2)
    d += 2 * nI * nh * N;                // q, V
    d += C * (size_t)ntiles * 3;         // obs_part
    if (p->precision) {
        d += (size_t)p->n_orb * (N + 1) + 4 * C;                 // phi_ref, eps, ab
        if (p->precision == 1) tbytes += Nl * C * N * 2 * 4;   // psi32
    }
    return d * sizeof(double) + tbytes + (Nl + 2 * p->n_orb) * sizeof(int32_t);
}

extern "C" int ggb_footprint(const ggb_problem *p, uint64_t *bytes) {
    if (!p || !bytes || p->n_r < 1 || p->n_l < 1 || p->n_theta < 1 || p->n_orb < 1 || p->n_inst < 1)
        return GGB_E_ARG;
    *bytes = footprint_bytes(p, (p->n_r + RT - 1) / RT);
    return GGB_OK;
}

extern "C" int ggb_create(const ggb_problem *p, ggb_ctx **out) {
    if (!p || !out) {
        g_create_err = "ggb_create: null argument";
        return GGB_E_ARG;
    }
    *out = nullptr;
    if (p->abi_version != GGB_ABI_VERSION) {
        g_create_err = "ggb_create: abi_version mismatch";
        return GGB_E_ARG;
    }
    if (p->n_r < 4 || p->n_r > 4096 || p->n_l < 2 || p->n_l > 128 || p->n_theta < p->n_l ||
        p->n_theta > 256 || p->n_har < 1 || p->n_har > 256 || p->n_orb < 1 || p->n_inst < 1 ||
        p->n_m < 1 || p->n_m > p->n_l || p->nk_max < 16 || p->nk_max % 16 != 0 ||
        p->nk_max > ((p->n_r + 15) / 16) * 16 || !(p->dt > 0.0) ||
        (p->precision < 0 || p->precision > 2) ||
        (long)p->n_orb * p->n_inst > 65535) {
        g_create_err = "ggb_create: problem sizes out of range";
        return GGB_E_ARG;
    }
    ggb_ctx *ctx = new ggb_ctx();
    memset((void *)&ctx->p, 0, sizeof(ctx->p));
    ctx->p = *p;
    ctx->ang = ang_geometry(p);
    ctx->C = p->n_orb * p->n_inst;
    ctx->ntiles = (p->n_r + RT - 1) / RT;
    ctx->have_static = ctx->have_orbitals = false;
    ctx->n_field = 0;
    ctx->field_ring = false;
    ctx->obs_cap = 0;
    ctx->channels = false;
    ctx->chan = nullptr;
    ctx->chan_cap = 0;
    ctx->chan_rows = 0;
    ctx->kph = nullptr;
    ctx->psi = ctx->Y = ctx->Q = ctx->QT = ctx->B = ctx->PL = ctx->PLT = ctx->chi = ctx->omega = nullptr;
    ctx->r = ctx->hpow = ctx->wr = ctx->v_gs = ctx->mask = ctx->occ = ctx->rho_frozen = nullptr;
    ctx->nkeep = ctx->mcol = ctx->active = nullptr;
    ctx->dens = ctx->vxc = ctx->rho_u = ctx->q = ctx->V = ctx->dv = ctx->obs_part = ctx->obs = ctx->E = nullptr;
    ctx->phase = nullptr;
    ctx->kin_variant = GGB_KIN_DEFAULT;
    ctx->have_ref = false;
    ctx->psi32 = nullptr;
    ctx->Qlo = ctx->Ylo = nullptr;
    ctx->phi_ref = ctx->eps = nullptr;
    ctx->l0 = nullptr;
    ctx->ab = nullptr;
    ctx->stream = nullptr;
    ctx->timing = 0;
    ctx->graph_steps = 0;
    ctx->graph_exec = nullptr;
    ctx->dcur = nullptr;
    ctx->gen = 0;
    for (int k = 0; k < GGB_NPHASE; ++k) {
        ctx->phase_ms[k] = 0.0;
        ctx->phase_n[k] = 0;
    }

    auto bail = [&](const std::string &m, int code) {
        g_create_err = m;
        ggb_destroy(ctx);
        return code;
    };
    if (hipSetDevice(p->device) != hipSuccess) return bail("ggb_create: hipSetDevice failed", GGB_E_HIP);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
        return bail("ggb_create: hipStreamCreate failed", GGB_E_HIP);
    const size_t N = p->n_r, Nl = p->n_l, C = ctx->C, Nth = p->n_theta, nI = p->n_inst;
    bool ok = true;
    ok &= dalloc(&ctx->psi, Nl * C * N * 2) == hipSuccess;
    // Y, Q and the angular tables hold the compute type (8 or 4 bytes per element)
    const size_t es = p->precision ? sizeof(float) : sizeof(double);
    const size_t nY = Nl * C * (size_t)p->nk_max * 2, nQ = Nl * N * (size_t)p->nk_max;
    const size_t nB = (size_t)p->n_m * ctx->ang.nblk;
    ok &= dalloc(&ctx->Y, (nY * es + 7) / 8) == hipSuccess;
    ok &= dalloc(&ctx->Q, (nQ * es + 7) / 8) == hipSuccess;
    if (p->precision != 2)  // [l][a][i] for the narrow pass 2 (precision 2 runs the split passes only)
        ok &= dalloc(&ctx->QT, (nQ * es + 7) / 8) == hipSuccess;
    ok &= dalloc(&ctx->kph, Nl * (size_t)p->nk_max) == hipSuccess;
    ok &= dalloc(&ctx->nkeep, Nl) == hipSuccess;
    ok &= dalloc(&ctx->B, (nB * es + 7) / 8) == hipSuccess;
    ok &= dalloc(&ctx->PL, Nth * (size_t)p->n_har) == hipSuccess;
    ok &= dalloc(&ctx->PLT, Nth * (size_t)p->n_har) == hipSuccess;
    ok &= dalloc(&ctx->chi, Nth) == hipSuccess;
    ok &= dalloc(&ctx->omega, Nth) == hipSuccess;
    ok &= dalloc(&ctx->r, N) == hipSuccess;
    ok &= dalloc(&ctx->hpow, (2 * (size_t)p->n_har + 1) * N) == hipSuccess;  // Hartree ratio powers, 1/r
    ok &= dalloc(&ctx->wr, N) == hipSuccess;
    ok &= dalloc(&ctx->v_gs, N) == hipSuccess;
    ok &= dalloc(&ctx->mask, N) == hipSuccess;
    ok &= dalloc(&ctx->occ, (size_t)p->n_orb) == hipSuccess;
    ok &= dalloc(&ctx->mcol, (size_t)p->n_orb) == hipSuccess;
    ok &= dalloc(&ctx->active, (size_t)p->n_inst) == hipSuccess;
    ok &= dalloc(&ctx->dens, C * N * 2 * (size_t)ctx->ang.hq) == hipSuccess;  // folded |Psi|^2 rows
    ok &= dalloc(&ctx->vxc, nI * N * Nth) == hipSuccess;
    ok &= dalloc(&ctx->rho_u, nI * N * Nth) == hipSuccess;
    ok &= dalloc(&ctx->q, nI * (size_t)p->n_har * N) == hipSuccess;
    ok &= dalloc(&ctx->V, nI * (size_t)p->n_har * N) == hipSuccess;
    ok &= dalloc(&ctx->dv, nI * N * Nth) == hipSuccess;
    ok &= dalloc(&ctx->phase, nI * N * Nth) == hipSuccess;
    ok &= dalloc(&ctx->obs_part, C * (size_t)ctx->ntiles * 3) == hipSuccess;
    if (p->precision) {
        ok &= dalloc(&ctx->phi_ref, (size_t)p->n_orb * N) == hipSuccess;
        ok &= dalloc(&ctx->eps, (size_t)p->n_orb) == hipSuccess;
        ok &= dalloc(&ctx->l0, (size_t)p->n_orb) == hipSuccess;
        ok &= dalloc(&ctx->ab, C) == hipSuccess;
        if (p->precision == 1) ok &= dalloc(&ctx->psi32, Nl * C * N) == hipSuccess;  // tiled pass 1
        if (p->precision == 2) {
            ok &= dalloc(&ctx->Qlo, nQ) == hipSuccess;
            ok &= dalloc(&ctx->Ylo, nY) == hipSuccess;
        }
    }
    if (!ok) return bail("ggb_create: device allocation failed", GGB_E_NOMEM);
    // zero everything that kernels may read before it is written
    if (hipMemset(ctx->B, 0, nB * es) != hipSuccess || hipMemset(ctx->Q, 0, nQ * es) != hipSuccess ||
        (ctx->QT && hipMemset(ctx->QT, 0, nQ * es) != hipSuccess) ||
        (ctx->Qlo && hipMemset(ctx->Qlo, 0, nQ * sizeof(float)) != hipSuccess) ||
        (ctx->Ylo && hipMemset(ctx->Ylo, 0, nY * sizeof(float)) != hipSuccess) ||
        hipMemset(ctx->kph, 0, Nl * (size_t)p->nk_max * sizeof(double2)) != hipSuccess ||
        hipMemset(ctx->Y, 0, nY * es) != hipSuccess ||
        hipMemset(ctx->psi, 0, Nl * C * N * 2 * sizeof(double)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return bail("ggb_create: hipMemset failed", GGB_E_HIP);
    {
        const std::vector<int32_t> all(p->n_inst, 1);  // every instance in use
        if (hipMemcpy(ctx->active, all.data(), all.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return bail("ggb_create: copy of the instance mask failed", GGB_E_HIP);
    }
    *out = ctx;
    return GGB_OK;
}

extern "C" int ggb_upload_static(ggb_ctx *ctx, const ggb_static *s) {
    if (!ctx) return GGB_E_ARG;
    if (!s || !s->r || !s->wr || !s->v_gs || !s->mask || !s->Q || !s->lam || !s->nkeep || !s->B ||
        !s->PL || !s->chi || !s->omega || !s->m || !s->occ)
        return fail(ctx, GGB_E_ARG, "ggb_upload_static: null table");
    const ggb_problem &p = ctx->p;
    for (int l = 0; l < p.n_l; ++l)
        if (s->nkeep[l] < 1 || s->nkeep[l] > p.nk_max)
            return fail(ctx, GGB_E_ARG, "ggb_upload_static: nkeep out of range");
    for (int k = 0; k < p.n_orb; ++k)
        if (s->m[k] < 0 || s->m[k] >= p.n_m || s->m[k] >= p.n_l)
            return fail(ctx, GGB_E_ARG, "ggb_upload_static: orbital m out of range");
    for (int i = 0; i < p.n_r; ++i)
        if (!(s->r[i] > 0.0) || !(s->wr[i] > 0.0))
            return fail(ctx, GGB_E_ARG, "ggb_upload_static: r and wr must be positive");
    HIPCHK(ctx, hipSetDevice(p.device));
    drop_graph(ctx);
    const size_t N = p.n_r, Nl = p.n_l, Nth = p.n_theta;
    hipStream_t st = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(ctx->r, s->r, N * 8, hipMemcpyHostToDevice, st));
    {
        // the Hartree scans' ratio powers as exp(n log(r_{a-1}/r_a)) (no overflow
        // at large L) and 1/r, laid out as k_hartree reads them
        const size_t nh = p.n_har;
        std::vector<double> lrh(N, 0.0), hp((2 * nh + 1) * N, 0.0);
        for (size_t a = 1; a < N; ++a) lrh[a] = std::log(s->r[a - 1] / s->r[a]);
        for (size_t L = 0; L < nh; ++L)
            for (size_t a = 0; a < N; ++a) {
                hp[L * N + a] = a > 0 ? std::exp((double)(L + 1) * lrh[a]) : 0.0;
                hp[(nh + L) * N + a] = a + 1 < N ? std::exp((double)L * lrh[a + 1]) : 0.0;
            }
        for (size_t a = 0; a < N; ++a) hp[2 * nh * N + a] = 1.0 / s->r[a];
        HIPCHK(ctx, hipMemcpy(ctx->hpow, hp.data(), hp.size() * 8, hipMemcpyHostToDevice));
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->wr, s->wr, N * 8, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->v_gs, s->v_gs, N * 8, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->mask, s->mask, N * 8, hipMemcpyHostToDevice, st));
    {
        // Q[l][i][a] and its quad-interleaved transpose QT (read by the narrow pass 2), compute type
        const size_t nk = p.nk_max, nQ = Nl * N * nk;
        if (p.precision == 0) {
            HIPCHK(ctx, hipMemcpy(ctx->Q, s->Q, nQ * 8, hipMemcpyHostToDevice));
            std::vector<double> qt(nQ);
            for (size_t l = 0; l < Nl; ++l)
                for (size_t i = 0; i < N; ++i)
                    for (size_t a = 0; a < nk; ++a) qt[kin_qt_index(l, i, a, N, nk)] = s->Q[(l * N + i) * nk + a];
            HIPCHK(ctx, hipMemcpy(ctx->QT, qt.data(), nQ * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> q32(nQ);
            for (size_t k = 0; k < nQ; ++k) q32[k] = (float)s->Q[k];
            HIPCHK(ctx, hipMemcpy(ctx->Q, q32.data(), nQ * 4, hipMemcpyHostToDevice));
            if (ctx->QT) {
                std::vector<float> qt(nQ);
                for (size_t l = 0; l < Nl; ++l)
                    for (size_t i = 0; i < N; ++i)
                        for (size_t a = 0; a < nk; ++a) qt[kin_qt_index(l, i, a, N, nk)] = q32[(l * N + i) * nk + a];
                HIPCHK(ctx, hipMemcpy(ctx->QT, qt.data(), nQ * 4, hipMemcpyHostToDevice));
            }
            if (ctx->Qlo) {  // low parts: Q - fl32(Q), rounded to FP32
                std::vector<float> lo(nQ);
                for (size_t k = 0; k < nQ; ++k) lo[k] = (float)(s->Q[k] - (double)q32[k]);
                HIPCHK(ctx, hipMemcpy(ctx->Qlo, lo.data(), nQ * 4, hipMemcpyHostToDevice));
            }
        }
    }
    {
        // e^{-i lam dt} of every retained eigenvalue, in FP64 on the host
        std::vector<double2> f(Nl * p.nk_max, make_double2(1.0, 0.0));
        for (size_t l = 0; l < Nl; ++l)
            for (int a = 0; a < s->nkeep[l]; ++a) {
                const double x = s->lam[l * p.nk_max + a] * p.dt;
                f[l * p.nk_max + a] = make_double2(std::cos(x), std::sin(x));
            }
        HIPCHK(ctx, hipMemcpy(ctx->kph, f.data(), f.size() * sizeof(double2), hipMemcpyHostToDevice));
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->nkeep, s->nkeep, Nl * 4, hipMemcpyHostToDevice, st));
    {
        // parity half tables of each m (see the angular kernel): rows j < h of
        // the theta grid, Bs[j][k] = B[m][j][m + 2k], Ba[j][k] = B[m][j][m + 1 + 2k],
        // zero padded.  They stand for the whole grid only if it is symmetric.
        const ang_geom &ag = ctx->ang;
        double bmax = 0.0;
        for (size_t k = 0; k < (size_t)p.n_m * Nth * Nl; ++k) bmax = std::max(bmax, std::fabs(s->B[k]));
        const double tol = 1e-13 * (bmax > 0.0 ? bmax : 1.0);
        for (size_t j = 0; j < Nth; ++j)
            if (std::fabs(s->chi[j] + s->chi[Nth - 1 - j]) > 1e-13 ||
                std::fabs(s->omega[j] - s->omega[Nth - 1 - j]) > 1e-13 * std::fabs(s->omega[j]))
                return fail(ctx, GGB_E_ARG, "ggb_upload_static: theta grid must be symmetric (Gauss-Legendre)");
        for (int mm = 0; mm < p.n_m; ++mm)
            for (size_t j = 0; j < Nth; ++j)
                for (size_t l = 0; l < Nl; ++l) {
                    const double sg = ((l + mm) & 1) ? -1.0 : 1.0;
                    const double a = s->B[((size_t)mm * Nth + j) * Nl + l];
                    const double b = s->B[((size_t)mm * Nth + Nth - 1 - j) * Nl + l];
                    if (std::fabs(a - sg * b) > tol || ((int)l < mm && a != 0.0))
                        return fail(ctx, GGB_E_ARG,
                                    "ggb_upload_static: B must satisfy B[m][N_theta-1-j][l] = (-1)^(l-m) B[m][j][l]"
                                    " and vanish for l < m");
                }
        std::vector<double> bh((size_t)p.n_m * ag.nblk, 0.0);
        for (int mm = 0; mm < p.n_m; ++mm) {
            double *Bs = bh.data() + (size_t)mm * ag.nblk;
            double *Ba = Bs + (size_t)ag.hq * ag.sS;
            for (int j = 0; j < ag.h; ++j) {
                const double *row = s->B + ((size_t)mm * Nth + j) * Nl;
                for (int k = 0; mm + 2 * k < (int)Nl; ++k) Bs[(size_t)j * ag.sS + k] = row[mm + 2 * k];
                for (int k = 0; mm + 1 + 2 * k < (int)Nl; ++k) Ba[(size_t)j * ag.sA + k] = row[mm + 1 + 2 * k];
            }
        }
        if (p.precision == 0) {
            HIPCHK(ctx, hipMemcpy(ctx->B, bh.data(), bh.size() * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> b32(bh.size());
            for (size_t k = 0; k < bh.size(); ++k) b32[k] = (float)bh[k];
            HIPCHK(ctx, hipMemcpy(ctx->B, b32.data(), b32.size() * 4, hipMemcpyHostToDevice));
        }
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->PL, s->PL, Nth * p.n_har * 8, hipMemcpyHostToDevice, st));
    {
        std::vector<double> plt((size_t)Nth * p.n_har);
        for (size_t j = 0; j < Nth; ++j)
            for (int L = 0; L < p.n_har; ++L) plt[(size_t)L * Nth + j] = s->PL[j * p.n_har + L];
        HIPCHK(ctx, hipMemcpy(ctx->PLT, plt.data(), plt.size() * 8, hipMemcpyHostToDevice));
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->chi, s->chi, Nth * 8, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->omega, s->omega, Nth * 8, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->mcol, s->m, p.n_orb * 4, hipMemcpyHostToDevice, st));
    ctx->have_ref = false;
    if (p.precision && s->phi && s->l0 && s->eps) {
        for (int k = 0; k < p.n_orb; ++k)
            if (s->l0[k] < s->m[k] || s->l0[k] >= p.n_l)
                return fail(ctx, GGB_E_ARG, "ggb_upload_static: reference l0 out of range");
        HIPCHK(ctx, hipMemcpy(ctx->phi_ref, s->phi, (size_t)p.n_orb * N * 8, hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(ctx->eps, s->eps, (size_t)p.n_orb * 8, hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(ctx->l0, s->l0, (size_t)p.n_orb * 4, hipMemcpyHostToDevice));
        ctx->have_ref = true;
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->occ, s->occ, p.n_orb * 8, hipMemcpyHostToDevice, st));
    if (s->rho_frozen) {
        if (!ctx->rho_frozen) HIPCHK(ctx, dalloc(&ctx->rho_frozen, N));
        HIPCHK(ctx, hipMemcpyAsync(ctx->rho_frozen, s->rho_frozen, N * 8, hipMemcpyHostToDevice, st));
    } else if (ctx->rho_frozen) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        (void)hipFree(ctx->rho_frozen);
        ctx->rho_frozen = nullptr;
    }
    HIPCHK(ctx, hipStreamSynchronize(st));
    ctx->have_static = true;
    ctx->have_orbitals = false;
    return GGB_OK;
}

static AngArgs ang_args(ggb_ctx *ctx, int step) {
    AngArgs g;
    g.psi = ctx->psi;
    g.Bh = ctx->B;
    g.mcol = ctx->mcol;
    g.phase = ctx->phase;
    g.r = ctx->r;
    g.chi = ctx->chi;
    g.E = ctx->E;
    g.dens = ctx->dens;
    g.obs_part = ctx->obs_part;
    g.N = ctx->p.n_r;
    g.Nl = ctx->p.n_l;
    g.Nlx = ctx->p.n_l + 6;  // rows up to Nl + 5 are read by the parity products (zero)
    g.Nth = ctx->p.n_theta;
    g.h = ctx->ang.h;
    g.hq = ctx->ang.hq;
    g.sS = ctx->ang.sS;
    g.sA = ctx->ang.sA;
    g.nblk = ctx->ang.nblk;
    g.alias = ctx->ang.alias ? 1 : 0;
    g.C = ctx->C;
    g.n_orb = ctx->p.n_orb;
    g.n_field = ctx->n_field;
    g.step = step;
    g.dt = ctx->p.dt;
    g.cur = nullptr;
    g.psi32 = nullptr;
    g.l0 = nullptr;
    g.active = ctx->active;
    return g;
}

template <int MODE, typename T>
static void launch_angular_t(ggb_ctx *ctx, const AngArgs &g) {
    const ang_geom &ag = ctx->ang;
    const dim3 ga(ag.ntiles, ctx->C);
    // the geometry (ag.nt, ag.ntiles, ag.lds) was computed for this tile count
    switch (ag.nt) {
        case 1:
            if (ag.glb)
                k_angular<MODE, T, true, 1><<<ga, 64 * ag.waves, ag.lds, ctx->stream>>>(g);
            else
                k_angular<MODE, T, false, 1><<<ga, 64 * ag.waves, ag.lds, ctx->stream>>>(g);
            break;
        case 2:
            if (ag.glb)
                k_angular<MODE, T, true, 2><<<ga, 64 * ag.waves, ag.lds, ctx->stream>>>(g);
            else
                k_angular<MODE, T, false, 2><<<ga, 64 * ag.waves, ag.lds, ctx->stream>>>(g);
            break;
#if ANG_NT_MIXED == 4
        case 4:
            if (ag.glb)
                k_angular<MODE, T, true, 4><<<ga, 64 * ag.waves, ag.lds, ctx->stream>>>(g);
            else
                k_angular<MODE, T, false, 4><<<ga, 64 * ag.waves, ag.lds, ctx->stream>>>(g);
            break;
#endif
        default:
            break;  // ggb_create accepts only the tile counts above
    }
}

static int kin_variant(const ggb_ctx *ctx);

template <int MODE>
static void launch_angular(ggb_ctx *ctx, int step, const int *cur = nullptr) {
    AngArgs g = ang_args(ctx, step);
    g.cur = cur;
    if (MODE == 0 && ctx->psi32 &&
        (kin_variant(ctx) == GGB_KIN_TILED16 || kin_variant(ctx) == GGB_KIN_TILED32)) {
        g.psi32 = ctx->psi32;  // the kinetic pass 1 that follows reads it
        g.l0 = ctx->have_ref ? ctx->l0 : nullptr;
    }
    if (ctx->p.precision)
        launch_angular_t<MODE, float>(ctx, g);
    else
        launch_angular_t<MODE, double>(ctx, g);
}

// potential rebuild from the |Psi|^2 tiles already in ctx->dens
// ob: the observable reduction of the step, folded into k_rho (interacting) or
// launched on its own
static hipError_t no_mark_k(int) { return hipSuccess; }

// mk(0) after the density kernel, mk(1) after the Hartree scans (timing events)
template <typename MarkK = hipError_t (*)(int)>
static int launch_potential(ggb_ctx *ctx, const ObsArgs *ob = nullptr, MarkK mk = no_mark_k) {
    const ggb_problem &p = ctx->p;
    hipStream_t st = ctx->stream;
    const int N = p.n_r, Nth = p.n_theta, nh = p.n_har;
    ObsArgs none;
    none.obs = nullptr;
    const ObsArgs oa = ob ? *ob : none;
    if (oa.obs && !p.interacting) k_obs_reduce<<<oa.C * 3, 64, 0, st>>>(oa);
    if (p.interacting) {
        const unsigned nrho = (unsigned)(((size_t)N * Nth + WG - 1) / WG);
        const unsigned nobs = oa.obs ? (unsigned)((oa.C * 3 + WG / 64 - 1) / (WG / 64)) : 0u;
        const dim3 gr(nrho + nobs, p.n_inst);
        if (p.precision)
            k_rho<float><<<gr, WG, 0, st>>>(reinterpret_cast<const float *>(ctx->dens), ctx->occ, ctx->rho_frozen, ctx->r, ctx->wr,
                                            ctx->omega, ctx->rho_u, ctx->vxc, N, Nth, ctx->ang.h, ctx->ang.hq, p.n_orb, oa,
                                            ctx->active);
        else
            k_rho<double><<<gr, WG, 0, st>>>(ctx->dens, ctx->occ, ctx->rho_frozen, ctx->r, ctx->wr,
                                             ctx->omega, ctx->rho_u, ctx->vxc, N, Nth, ctx->ang.h, ctx->ang.hq, p.n_orb, oa,
                                             ctx->active);
        GemmOperands gq;  // q[L][i] = sum_j PLT[L][j] rho_u[i][j]
        gq.A = ctx->PLT;
        gq.sam = Nth;
        gq.sak = 1;
        gq.saz = 0;
        gq.B = ctx->rho_u;
        gq.sbk = 1;
        gq.sbn = Nth;
        gq.sbz = (long)N * Nth;
        gq.M = nh;
        gq.active = ctx->active;
        gq.N = N;
        gq.K = Nth;
        QmomEpi eq{ctx->q, N, nh};
        if (p.n_inst >= TGW_MIN_INST) {
            gq.A = ctx->PL;  // the same table as [j][L]: the 16 lanes of a K row read 16 consecutive L
            gq.sam = 1;
            gq.sak = nh;
            const int ntl = (nh + 15) / 16;
            k_tile_gemm_wide<QmomEpi, true><<<dim3((N + 15) / 16, (ntl + TGW_T - 1) / TGW_T, p.n_inst), WG, 0, st>>>(gq, eq);
        } else {
            k_tile_gemm<<<dim3((N + 15) / 16, (nh + 15) / 16, p.n_inst), WG, 0, st>>>(gq, eq);
        }
        HIPCHK(ctx, mk(0));
        const int cs = (N + WG - 1) / WG;  // N <= 4096 (ggb_create): cs <= 16
        auto har = cs <= 2 ? k_hartree<2> : cs <= 4 ? k_hartree<4> : cs <= 8 ? k_hartree<8> : k_hartree<16>;
        har<<<dim3(nh, p.n_inst), WG, 0, st>>>(ctx->q, ctx->hpow, ctx->V, N, nh, ctx->active);
        HIPCHK(ctx, mk(1));
    } else {
        HIPCHK(ctx, mk(0));
        HIPCHK(ctx, mk(1));
    }
    GemmOperands gv;  // V(i, theta_j) = sum_L V[L][i] PLT[L][j]
    gv.A = ctx->V;
    gv.sam = 1;
    gv.sak = N;
    gv.saz = (long)nh * N;
    gv.B = ctx->PLT;
    gv.sbk = Nth;
    gv.sbn = 1;
    gv.sbz = 0;
    gv.M = N;
    gv.active = ctx->active;
    gv.N = Nth;
    gv.K = p.interacting ? nh : 0;
    DvEpi ev{ctx->vxc, ctx->v_gs, ctx->mask, ctx->dv, ctx->phase, N, Nth, p.interacting, p.precision, p.dt};
    if (p.n_inst >= TGW_MIN_INST) {
        const int ntj = (Nth + 15) / 16;
        k_tile_gemm_wide<DvEpi, false><<<dim3((N + 15) / 16, (ntj + TGW_T - 1) / TGW_T, p.n_inst), WG, 0, st>>>(gv, ev);
    } else {
        k_tile_gemm<<<dim3((Nth + 15) / 16, (N + 15) / 16, p.n_inst), WG, 0, st>>>(gv, ev);
    }
    HIPCHK(ctx, hipGetLastError());
    return GGB_OK;
}

static int launch_first_half(ggb_ctx *ctx, int step, const int *cur = nullptr) {
    launch_angular<0>(ctx, step, cur);
    HIPCHK(ctx, hipGetLastError());
    return GGB_OK;
}

// kinetic kernel variant in use: the one set (ggb_set_kinetic_variant /
// ggb_autotune) or the default (narrow up to 8 columns, else tiled, K chunks of 16)
static int kin_variant(const ggb_ctx *ctx) {
    if (ctx->p.precision == 2) return GGB_KIN_SPLIT16;
    if (ctx->kin_variant != GGB_KIN_DEFAULT) return ctx->kin_variant;
    return ctx->C > KIN_NARROW_MAX_C ? GGB_KIN_TILED16 : GGB_KIN_NARROW;
}

// fresh: the angular first half has just written psi32 (all blocks but the
// ground-state ones); otherwise pass 1 converts every block itself
// launch of a kinetic product kernel; with events (timing) through
// hipExtLaunchKernel, which stamps e0/e1 with the kernel's own start and end
template <typename... KArgs, typename... Args>
static void kin_launch(void (*k)(KArgs...), dim3 g, int threads, hipStream_t st, hipEvent_t e0,
                       hipEvent_t e1, Args... args) {
    if (e0)
        hipExtLaunchKernelGGL(k, g, dim3(threads), 0, st, e0, e1, 0, args...);
    else
        k<<<g, threads, 0, st>>>(args...);
}

// e0/e1 (optional): events stamped with the start and end of the pass's
// product kernel itself, for the kernel-duration timing
template <typename T>
static void launch_kinetic_t(ggb_ctx *ctx, int pass, bool fresh, hipEvent_t e0 = nullptr,
                             hipEvent_t e1 = nullptr) {
    const ggb_problem &p = ctx->p;
    const T *Q = reinterpret_cast<const T *>(ctx->Q);
    T *Y = reinterpret_cast<T *>(ctx->Y);
    KinRef ref;
    ref.phi = (mfma_t<T>::mixed && ctx->have_ref) ? ctx->phi_ref : nullptr;
    ref.l0 = ctx->l0;
    ref.ab = ctx->ab;
    ref.active = ctx->active;
    if (pass == 0 && ref.phi)
        k_kin_ref<<<ctx->C, WG, 0, ctx->stream>>>(ctx->psi, ctx->phi_ref, ctx->l0, ctx->eps, ctx->ab,
                                                  p.n_r, ctx->C, p.n_orb, p.dt);
    const int var = kin_variant(ctx);
    if (var == GGB_KIN_SPLIT16) {
        const int nct = (2 * ctx->C + TN - 1) / TN;
        const dim3 gf((p.nk_max + TM - 1) / TM, nct, p.n_l), gb((p.n_r + TM - 1) / TM, nct, p.n_l);
        const float *Qh = reinterpret_cast<const float *>(ctx->Q);
        float *Yh = reinterpret_cast<float *>(ctx->Y);
        const double2 *kph = ctx->kph;
        const int32_t *nkeep = ctx->nkeep;
        const double *psi_in = ctx->psi;
        const float *Ql = ctx->Qlo;
        float *Yl = ctx->Ylo;
        double *psi_out = ctx->psi;
        const int N = p.n_r, nkm = p.nk_max, C = ctx->C, no = p.n_orb;
        if (pass == 0)
            kin_launch(k_kin_fwd_split, gf, WG, ctx->stream, e0, e1, Qh, Ql, kph, nkeep, psi_in, ref, Yh,
                       Yl, N, nkm, C, no);
        else
            kin_launch(k_kin_bwd_split, gb, WG, ctx->stream, e0, e1, Qh, Ql, nkeep, (const float *)Yh,
                       (const float *)Yl, ref, psi_out, N, nkm, C, no);
        return;
    }
    if (var >= GGB_KIN_TILED16) {
        const int nct = (2 * ctx->C + TN - 1) / TN;
        const dim3 gf((p.nk_max + TM - 1) / TM, nct, p.n_l), gb((p.n_r + TM - 1) / TM, nct, p.n_l);
        if (pass == 0 && mfma_t<T>::mixed && !(fresh && !ref.phi))
            k_kin_prep<<<dim3((p.n_r + WG - 1) / WG, ctx->C, fresh ? 1 : p.n_l), WG, 0, ctx->stream>>>(
                (const double2 *)ctx->psi, ref, ctx->psi32, p.n_r, ctx->C, p.n_orb, fresh ? 1 : 0);
        // FP64 chunks stay at 16 (32 would need 114 KB of LDS)
        const bool k32 = var == GGB_KIN_TILED32 && mfma_t<T>::mixed;
        const double2 *kph = ctx->kph;
        const int32_t *nkeep = ctx->nkeep;
        const double *psi_in = ctx->psi;
        double *psi_out = ctx->psi;
        const float2 *psi32 = ctx->psi32;
        const int N = p.n_r, nkm = p.nk_max, C = ctx->C, no = p.n_orb;
        if (pass == 0 && k32)
            kin_launch(k_kin_fwd_tiled<T, 32>, gf, WG, ctx->stream, e0, e1, Q,
                                  kph, nkeep, psi_in, psi32, ref, Y, N, nkm, C, no);
        else if (pass == 0 && mfma_t<T>::mixed && KIN_DMA && N % 16 == 0)
            kin_launch(k_kin_fwd_dma, gf, WG, ctx->stream, e0, e1, reinterpret_cast<const float *>(Q),
                       kph, nkeep, psi32, ref, reinterpret_cast<float *>(Y), N, nkm, C, no);
        else if (pass == 0)
            kin_launch(k_kin_fwd_tiled<T, 16>, gf, WG, ctx->stream, e0, e1, Q,
                                  kph, nkeep, psi_in, psi32, ref, Y, N, nkm, C, no);
        else if (k32)
            kin_launch(k_kin_bwd_tiled<T, 32>, gb, WG, ctx->stream, e0, e1,
                                  (const T *)Q, nkeep, (const T *)Y, ref, psi_out, N, nkm, C, no);
        else if (mfma_t<T>::mixed && KIN_DMA)
            kin_launch(k_kin_bwd_dma, gb, WG, ctx->stream, e0, e1, reinterpret_cast<const float *>(Q), nkeep,
                       reinterpret_cast<const float *>(Y), ref, psi_out, N, nkm, C, no);
        else
            kin_launch(k_kin_bwd_tiled<T, 16>, gb, WG, ctx->stream, e0, e1,
                                  (const T *)Q, nkeep, (const T *)Y, ref, psi_out, N, nkm, C, no);
        return;
    }
    const int ncol = (2 * ctx->C + 15) / 16;
    const double2 *kph = ctx->kph;
    const int32_t *nkeep = ctx->nkeep;
    const int N = p.n_r, nkm = p.nk_max, C = ctx->C, no = p.n_orb;
    if (pass == 0) {
        const double *psi_in = ctx->psi;
        kin_launch(k_kin_fwd<T, KIN_TA>, dim3((p.nk_max / 16 + KIN_TA - 1) / KIN_TA, ncol, p.n_l), 64 * kin_nw<T>::value,
                   ctx->stream,
                   e0, e1, Q, kph, nkeep,
                   psi_in, ref, Y, N, nkm, C, no);
    } else {
        kin_launch(k_kin_bwd<T>, dim3((p.n_r + 15) / 16, ncol, p.n_l), 64 * kin_nw<T>::value, ctx->stream, e0, e1,
                   reinterpret_cast<const T *>(ctx->QT), nkeep, (const T *)Y, ref, ctx->psi, N, nkm,
                   C, no);
    }
}

static int launch_kinetic_pass(ggb_ctx *ctx, int pass, bool fresh = false, hipEvent_t e0 = nullptr,
                               hipEvent_t e1 = nullptr) {
    if (ctx->p.precision)
        launch_kinetic_t<float>(ctx, pass, fresh, e0, e1);
    else
        launch_kinetic_t<double>(ctx, pass, fresh, e0, e1);
    HIPCHK(ctx, hipGetLastError());
    return GGB_OK;
}

static int launch_kinetic(ggb_ctx *ctx, bool fresh = false) {
    int rc = launch_kinetic_pass(ctx, 0, fresh);
    return rc ? rc : launch_kinetic_pass(ctx, 1, fresh);
}

static hipError_t no_mark() { return hipSuccess; }

// channel dipoles of the state after the step into dst (row 0, or the
// replay row cur[0] - cur[1] + step)
static void launch_channels(ggb_ctx *ctx, int step, double2 *dst, const int *cur) {
    ChanArgs a;
    a.psi = ctx->psi;
    a.r = ctx->r;
    a.mcol = ctx->mcol;
    a.active = ctx->active;
    a.chan = dst;
    a.N = ctx->p.n_r;
    a.Nl = ctx->p.n_l;
    a.C = ctx->C;
    a.n_orb = ctx->p.n_orb;
    a.cur = cur;
    a.origin = cur ? cur + 1 : nullptr;
    a.local = cur ? step : 0;
    k_chan_dip<<<dim3(ctx->p.n_l, ctx->C), 64, 0, ctx->stream>>>(a);
}

template <typename Mark = hipError_t (*)(), typename MarkK = hipError_t (*)(int)>
static int launch_second_half(ggb_ctx *ctx, int step, double *obs_dst, const int *cur = nullptr,
                              Mark between = no_mark, double2 *chan_dst = nullptr,
                              MarkK mk = no_mark_k) {
    launch_angular<1>(ctx, step, cur);
    HIPCHK(ctx, hipGetLastError());
    if (chan_dst) launch_channels(ctx, step, chan_dst, cur);
    HIPCHK(ctx, between());  // timing event between the angular half and the rebuild
    ObsArgs ob;
    ob.part = ctx->obs_part;
    ob.obs = obs_dst;
    ob.C = ctx->C;
    ob.ntiles = ctx->ang.ntiles;  // observable partials per column: one per angular work-group along r
    ob.cur = cur;
    ob.origin = cur ? cur + 1 : nullptr;
    ob.local = step;
    return launch_potential(ctx, &ob, mk);
}

extern "C" int ggb_upload_orbitals(ggb_ctx *ctx, const double *psi) {
    if (!ctx) return GGB_E_ARG;
    if (!psi) return fail(ctx, GGB_E_ARG, "ggb_upload_orbitals: null psi");
    if (!ctx->have_static) return fail(ctx, GGB_E_STATE, "ggb_upload_orbitals: static data not uploaded");
    const ggb_problem &p = ctx->p;
    HIPCHK(ctx, hipSetDevice(p.device));
    HIPCHK(ctx, hipMemcpyAsync(ctx->psi, psi, (size_t)p.n_l * ctx->C * p.n_r * 16,
                               hipMemcpyHostToDevice, ctx->stream));
    if (!ctx->E) {
        // a one-sample zero field so that the argument block is valid
        HIPCHK(ctx, dalloc(&ctx->E, (size_t)p.n_inst));
        HIPCHK(ctx, hipMemsetAsync(ctx->E, 0, (size_t)p.n_inst * 8, ctx->stream));
        ctx->n_field = 1;
    }
    launch_angular<2>(ctx, 0);
    HIPCHK(ctx, hipGetLastError());
    int rc = launch_potential(ctx);
    if (rc) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_orbitals = true;
    return GGB_OK;
}

extern "C" int ggb_upload_columns(ggb_ctx *ctx, int32_t c0, int32_t nc, const double *psi) {
    if (!ctx) return GGB_E_ARG;
    if (!psi || c0 < 0 || nc < 1 || c0 + nc > ctx->C)
        return fail(ctx, GGB_E_ARG, "ggb_upload_columns: bad column range");
    if (!ctx->have_orbitals) return fail(ctx, GGB_E_STATE, "ggb_upload_columns: upload all orbitals first");
    const ggb_problem &p = ctx->p;
    HIPCHK(ctx, hipSetDevice(p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const size_t row = (size_t)p.n_r * 16;  // one column of one l
    // stream-ordered copies only: a synchronous copy on the legacy stream
    // fails while another thread captures a graph on its own stream
    HIPCHK(ctx, hipMemcpy2DAsync(ctx->psi + (size_t)c0 * p.n_r * 2, ctx->C * row, psi, nc * row,
                                 nc * row, p.n_l, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GGB_OK;
}

extern "C" int ggb_set_active(ggb_ctx *ctx, const int32_t *active) {
    if (!ctx) return GGB_E_ARG;
    if (!active) return fail(ctx, GGB_E_ARG, "ggb_set_active: null mask");
    const int n = ctx->p.n_inst;
    std::vector<int32_t> m(n);
    for (int k = 0; k < n; ++k) m[k] = active[k] ? 1 : 0;
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    // stream-ordered: work already queued sees the old mask, later work the new
    HIPCHK(ctx, hipMemcpyAsync(ctx->active, m.data(), (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GGB_OK;
}

extern "C" int ggb_rebuild_potential(ggb_ctx *ctx) {
    if (!ctx) return GGB_E_ARG;
    if (!ctx->have_orbitals) return fail(ctx, GGB_E_STATE, "ggb_rebuild_potential: no orbitals");
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    launch_angular<2>(ctx, 0);
    HIPCHK(ctx, hipGetLastError());
    int rc = launch_potential(ctx);
    if (rc) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GGB_OK;
}

extern "C" int ggb_set_field(ggb_ctx *ctx, const double *E, int32_t n_steps) {
    if (!ctx) return GGB_E_ARG;
    if (!E || n_steps < 1) return fail(ctx, GGB_E_ARG, "ggb_set_field: empty field");
    const ggb_problem &p = ctx->p;
    HIPCHK(ctx, hipSetDevice(p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    if (ctx->E) {
        (void)hipFree(ctx->E);
        ctx->E = nullptr;
    }
    HIPCHK(ctx, dalloc(&ctx->E, (size_t)p.n_inst * n_steps));
    HIPCHK(ctx, hipMemcpyAsync(ctx->E, E, (size_t)p.n_inst * n_steps * 8, hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_field = n_steps;
    ctx->field_ring = false;
    return GGB_OK;
}

extern "C" int ggb_field_ring(ggb_ctx *ctx, int32_t capacity) {
    if (!ctx) return GGB_E_ARG;
    if (capacity < 1) return fail(ctx, GGB_E_ARG, "ggb_field_ring: capacity must be positive");
    const ggb_problem &p = ctx->p;
    HIPCHK(ctx, hipSetDevice(p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    if (ctx->E) {
        (void)hipFree(ctx->E);
        ctx->E = nullptr;
    }
    HIPCHK(ctx, dalloc(&ctx->E, (size_t)p.n_inst * capacity));
    HIPCHK(ctx, hipMemsetAsync(ctx->E, 0, (size_t)p.n_inst * capacity * 8, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_field = capacity;
    ctx->field_ring = true;
    return GGB_OK;
}

extern "C" int ggb_set_field_slice(ggb_ctx *ctx, int32_t inst, int32_t start, int32_t n,
                                   const double *E) {
    if (!ctx) return GGB_E_ARG;
    const ggb_problem &p = ctx->p;
    if (!E || inst < 0 || inst >= p.n_inst || start < 0 || n < 0)
        return fail(ctx, GGB_E_ARG, "ggb_set_field_slice: bad instance or range");
    if (!ctx->E || ctx->n_field < 1) return fail(ctx, GGB_E_STATE, "ggb_set_field_slice: no field buffer");
    if (!ctx->field_ring && (long)start + n > ctx->n_field)
        return fail(ctx, GGB_E_ARG, "ggb_set_field_slice: range outside the field");
    if (n > ctx->n_field) return fail(ctx, GGB_E_ARG, "ggb_set_field_slice: longer than the ring");
    HIPCHK(ctx, hipSetDevice(p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    double *base = ctx->E + (size_t)inst * ctx->n_field;
    const int s0 = start % ctx->n_field;
    const int first = (s0 + n <= ctx->n_field) ? n : ctx->n_field - s0;
    HIPCHK(ctx, hipMemcpyAsync(base + s0, E, (size_t)first * 8, hipMemcpyHostToDevice, ctx->stream));
    if (first < n)
        HIPCHK(ctx, hipMemcpyAsync(base, E + first, (size_t)(n - first) * 8, hipMemcpyHostToDevice,
                                   ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GGB_OK;
}

// Capture graph_steps steps once (field index and observable row taken from
// the device counter dcur), then replay it; returns the number of steps done.
static int run_graph(ggb_ctx *ctx, int step0, int nsteps, int *done) {
    const int G = ctx->graph_steps;
    if (!ctx->dcur) HIPCHK(ctx, hipMalloc((void **)&ctx->dcur, 2 * sizeof(int)));
    if (!ctx->graph_exec) {
        hipGraph_t graph;
        HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = GGB_OK;
        for (int s = 0; s < G && !rc; ++s) {
            rc = launch_first_half(ctx, s, ctx->dcur);
            if (!rc) rc = launch_kinetic(ctx, true);
            if (!rc)
                rc = launch_second_half(ctx, s, ctx->obs, ctx->dcur, no_mark,
                                        ctx->channels ? ctx->chan : nullptr);
        }
        if (!rc) k_bump<<<1, 64, 0, ctx->stream>>>(ctx->dcur, G);
        hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (rc) return rc;
        if (e != hipSuccess) return fail(ctx, GGB_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            ctx->graph_exec = nullptr;
            return fail(ctx, GGB_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
    }
    int h[2] = {step0, step0};
    HIPCHK(ctx, hipMemcpyAsync(ctx->dcur, h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    const int reps = nsteps / G;
    for (int k = 0; k < reps; ++k) HIPCHK(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
    *done = reps * G;
    return GGB_OK;
}

extern "C" int ggb_set_graph(ggb_ctx *ctx, int32_t steps) {
    if (!ctx) return GGB_E_ARG;
    if (steps < 0 || steps > 4096) return fail(ctx, GGB_E_ARG, "ggb_set_graph: steps in [0, 4096]");
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->graph_steps = steps;
    return GGB_OK;
}

extern "C" int ggb_run(ggb_ctx *ctx, int32_t step0, int32_t nsteps, double *dip, double *acc,
                       double *norm) {
    if (!ctx) return GGB_E_ARG;
    if (!ctx->have_orbitals) return fail(ctx, GGB_E_STATE, "ggb_run: orbitals not uploaded");
    if (nsteps < 0 || step0 < 0 || (!ctx->field_ring && (long)step0 + nsteps > ctx->n_field))
        return fail(ctx, GGB_E_ARG, "ggb_run: steps outside the uploaded field");
    if (nsteps == 0) return GGB_OK;
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    const size_t need = (size_t)nsteps * ctx->C * 3;
    if (need > ctx->obs_cap) {
        drop_graph(ctx);
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->obs) (void)hipFree(ctx->obs);
        ctx->obs = nullptr;
        ctx->obs_cap = 0;
        HIPCHK(ctx, dalloc(&ctx->obs, need));
        ctx->obs_cap = need;
    }
    const size_t cneed = ctx->channels ? (size_t)nsteps * ctx->C * ctx->p.n_l : 0;
    if (cneed > ctx->chan_cap) {
        drop_graph(ctx);
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->chan) (void)hipFree(ctx->chan);
        ctx->chan = nullptr;
        ctx->chan_cap = 0;
        HIPCHK(ctx, dalloc(&ctx->chan, cneed));
        ctx->chan_cap = cneed;
    }
    ctx->chan_rows = ctx->channels ? nsteps : 0;
    const int ne = GGB_NPHASE + 1;
    if (ctx->timing) {
        const size_t want = (size_t)nsteps * ne;
        while (ctx->events.size() < want) {
            hipEvent_t e;
            HIPCHK(ctx, hipEventCreate(&e));
            ctx->events.push_back(e);
        }
    }
    // mode 2: events between all phases; mode 1: the two kinetic product
    // kernels of every TIMING_STRIDE-th step, stamped by their own dispatch
    // (start/end of the kernel, events 1-2 and 3-4 of the step)
    auto mark = [&](int s, int k) -> hipError_t {
        if (ctx->timing != 2) return hipSuccess;
        return hipEventRecord(ctx->events[(size_t)s * ne + k], ctx->stream);
    };
    auto kev = [&](int s, int k) -> hipEvent_t {
        return (ctx->timing == 1 && s % TIMING_STRIDE == 0) ? ctx->events[(size_t)s * ne + k] : nullptr;
    };
    int s_begin = 0;
    if (ctx->graph_steps > 0 && !ctx->timing && nsteps >= ctx->graph_steps) {
        int rc = run_graph(ctx, step0, nsteps, &s_begin);
        if (rc) return rc;
    }
    for (int s = s_begin; s < nsteps; ++s) {
        const int step = (int)(((long)step0 + s) % ctx->n_field);
        HIPCHK(ctx, mark(s, 0));
        int rc = launch_first_half(ctx, step);
        HIPCHK(ctx, mark(s, 1));
        if (!rc) rc = launch_kinetic_pass(ctx, 0, true, kev(s, 1), kev(s, 2));
        HIPCHK(ctx, mark(s, 2));
        if (!rc) rc = launch_kinetic_pass(ctx, 1, true, kev(s, 3), kev(s, 4));
        HIPCHK(ctx, mark(s, 3));
        if (!rc)
            rc = launch_second_half(ctx, step, ctx->obs + (size_t)s * ctx->C * 3, nullptr,
                                    [&]() { return mark(s, 4); },
                                    ctx->channels ? ctx->chan + (size_t)s * ctx->C * ctx->p.n_l : nullptr,
                                    [&](int k) { return mark(s, 5 + k); });
        HIPCHK(ctx, mark(s, 7));
        if (rc) return rc;
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->timing) {
        for (int s = 0; s < nsteps; ++s) {
            for (int k = 0; k < GGB_NPHASE; ++k) {
                if (ctx->timing == 1 && (k < 1 || k > 2 || s % TIMING_STRIDE)) continue;
                // mode 1: phase 1 = events 1-2 (pass 1 kernel), phase 2 = events 3-4
                const int ka = ctx->timing == 1 ? 2 * k - 1 : k;
                float ms = 0.f;
                HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->events[(size_t)s * ne + ka],
                                                ctx->events[(size_t)s * ne + ka + 1]));
                ctx->phase_ms[k] += ms;
                ctx->phase_n[k] += 1;
            }
        }
    }
    if (dip || acc || norm) {
        double *h = new double[need];
        hipError_t e = hipMemcpyAsync(h, ctx->obs, need * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            delete[] h;
            return fail(ctx, GGB_E_HIP, std::string("ggb_run: copy of observables: ") + hipGetErrorString(e));
        }
        for (size_t k = 0; k < (size_t)nsteps * ctx->C; ++k) {
            if (dip) dip[k] = h[3 * k + 0];
            if (acc) acc[k] = h[3 * k + 1];
            if (norm) norm[k] = h[3 * k + 2];
        }
        delete[] h;
    }
    return GGB_OK;
}

// ---------------------------------------------------------------------------
// Groups: the step chunks of several contexts (e.g. the lanes of a batched
// scan, one context per grid) replayed by one hipGraph launch.  The graph is
// captured on the group's own stream: it forks to every context's stream, runs
// graph_steps steps of each context there (field index and observable row
// from each context's device step counter, as in run_graph), and joins back;
// contexts therefore overlap on the device with one host launch per chunk.
// The graph is captured again whenever one of its contexts changed anything a
// captured step holds (a drop of that context's own graph bumps ctx->gen).
// ---------------------------------------------------------------------------
struct ggb_group {
    std::vector<ggb_ctx *> ctx;
    std::vector<unsigned long> gen;  // generations at capture
    int steps;
    int device;
    hipStream_t lead;
    hipEvent_t fork;
    std::vector<hipEvent_t> join, ready;
    hipGraphExec_t exec;
    size_t obs_rows;  // rows of every context's observable buffer when captured
    std::string err;
};

static int gfail(ggb_group *g, int code, const std::string &m) {
    g->err = m;
    return code;
}

#define GHIP(g, call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return gfail((g), GGB_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" void ggb_group_destroy(ggb_group *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->lead) (void)hipStreamSynchronize(g->lead);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->fork) (void)hipEventDestroy(g->fork);
    for (hipEvent_t e : g->join) (void)hipEventDestroy(e);
    for (hipEvent_t e : g->ready) (void)hipEventDestroy(e);
    if (g->lead) (void)hipStreamDestroy(g->lead);
    delete g;
}

extern "C" int ggb_group_create(ggb_ctx *const *ctxs, int32_t n, int32_t steps, ggb_group **out) {
    if (!out) return GGB_E_ARG;
    *out = nullptr;
    if (!ctxs || n < 1 || n > 64 || steps < 1 || steps > 4096) {
        g_create_err = "ggb_group_create: 1..64 contexts and 1..4096 steps per graph";
        return GGB_E_ARG;
    }
    for (int k = 0; k < n; ++k) {
        if (!ctxs[k] || ctxs[k]->p.device != ctxs[0]->p.device) {
            g_create_err = "ggb_group_create: null context or contexts on different devices";
            return GGB_E_ARG;
        }
        for (int j = 0; j < k; ++j)
            if (ctxs[j] == ctxs[k]) {
                g_create_err = "ggb_group_create: a context appears twice";
                return GGB_E_ARG;
            }
    }
    ggb_group *g = new ggb_group();
    g->ctx.assign(ctxs, ctxs + n);
    g->gen.assign(n, 0);
    g->steps = steps;
    g->device = ctxs[0]->p.device;
    g->lead = nullptr;
    g->fork = nullptr;
    g->exec = nullptr;
    g->obs_rows = 0;
    bool ok = hipSetDevice(g->device) == hipSuccess &&
              hipStreamCreateWithFlags(&g->lead, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < n && ok; ++k) {
        hipEvent_t a = nullptr, b = nullptr;
        ok = hipEventCreateWithFlags(&a, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&b, hipEventDisableTiming) == hipSuccess;
        if (a) g->join.push_back(a);
        if (b) g->ready.push_back(b);
    }
    if (!ok) {
        g_create_err = "ggb_group_create: stream or event creation failed";
        ggb_group_destroy(g);
        return GGB_E_HIP;
    }
    *out = g;
    return GGB_OK;
}

extern "C" const char *ggb_group_last_error(const ggb_group *g) {
    return g ? g->err.c_str() : g_create_err.c_str();
}

static int group_capture(ggb_group *g) {
    if (g->exec) {
        (void)hipGraphExecDestroy(g->exec);
        g->exec = nullptr;
    }
    const int G = g->steps;
    hipGraph_t graph;
    GHIP(g, hipStreamBeginCapture(g->lead, hipStreamCaptureModeThreadLocal));
    int rc = GGB_OK;
    hipError_t e = hipEventRecord(g->fork, g->lead);
    for (size_t k = 0; k < g->ctx.size() && e == hipSuccess && !rc; ++k) {
        ggb_ctx *c = g->ctx[k];
        e = hipStreamWaitEvent(c->stream, g->fork, 0);
        for (int s = 0; s < G && e == hipSuccess && !rc; ++s) {
            rc = launch_first_half(c, s, c->dcur);
            if (!rc) rc = launch_kinetic(c, true);
            if (!rc) rc = launch_second_half(c, s, c->obs, c->dcur, no_mark, c->channels ? c->chan : nullptr);
        }
        if (!rc && e == hipSuccess) {
            k_bump<<<1, 64, 0, c->stream>>>(c->dcur, G);
            e = hipEventRecord(g->join[k], c->stream);
        }
        if (e == hipSuccess) e = hipStreamWaitEvent(g->lead, g->join[k], 0);
        if (rc) e = hipErrorUnknown;
        if (rc) g->err = "group capture: " + c->err;
    }
    hipError_t ee = hipStreamEndCapture(g->lead, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return gfail(g, GGB_E_HIP, std::string("group capture: ") + hipGetErrorString(e));
    if (ee != hipSuccess) return gfail(g, GGB_E_HIP, std::string("group capture end: ") + hipGetErrorString(ee));
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g->exec = nullptr;
        return gfail(g, GGB_E_HIP, std::string("group graph instantiate: ") + hipGetErrorString(e));
    }
    for (size_t k = 0; k < g->ctx.size(); ++k) g->gen[k] = g->ctx[k]->gen;
    return GGB_OK;
}

extern "C" int ggb_group_run(ggb_group *g, const int32_t *step0, int32_t nsteps, double *const *dip,
                             double *const *acc, double *const *norm) {
    if (!g) return GGB_E_ARG;
    if (!step0 || nsteps < 0) return gfail(g, GGB_E_ARG, "ggb_group_run: null step0 or negative steps");
    if (nsteps == 0) return GGB_OK;
    const size_t n = g->ctx.size();
    for (size_t k = 0; k < n; ++k) {
        ggb_ctx *c = g->ctx[k];
        if (!c->have_orbitals) return gfail(g, GGB_E_STATE, "ggb_group_run: a context has no orbitals");
        if (c->timing) return gfail(g, GGB_E_STATE, "ggb_group_run: timing is on in a context");
        if (step0[k] < 0 || (!c->field_ring && (long)step0[k] + nsteps > c->n_field))
            return gfail(g, GGB_E_ARG, "ggb_group_run: steps outside a context's field");
    }
    GHIP(g, hipSetDevice(g->device));
    // buffers the captured steps write: observables (and channels) for nsteps rows
    bool moved = false;
    for (size_t k = 0; k < n; ++k) {
        ggb_ctx *c = g->ctx[k];
        const size_t need = (size_t)nsteps * c->C * 3;
        if (need > c->obs_cap) {
            drop_graph(c);
            GHIP(g, hipStreamSynchronize(c->stream));
            if (c->obs) (void)hipFree(c->obs);
            c->obs = nullptr;
            c->obs_cap = 0;
            GHIP(g, dalloc(&c->obs, need));
            c->obs_cap = need;
            moved = true;
        }
        const size_t cneed = c->channels ? (size_t)nsteps * c->C * c->p.n_l : 0;
        if (cneed > c->chan_cap) {
            drop_graph(c);
            GHIP(g, hipStreamSynchronize(c->stream));
            if (c->chan) (void)hipFree(c->chan);
            c->chan = nullptr;
            c->chan_cap = 0;
            GHIP(g, dalloc(&c->chan, cneed));
            c->chan_cap = cneed;
            moved = true;
        }
        c->chan_rows = c->channels ? nsteps : 0;
        if (!c->dcur) {
            GHIP(g, hipMalloc((void **)&c->dcur, 2 * sizeof(int)));
            moved = true;
        }
    }
    const int G = g->steps;
    const int reps = nsteps / G;
    bool stale = moved || !g->exec;
    for (size_t k = 0; k < n && !stale; ++k) stale = g->gen[k] != g->ctx[k]->gen;
    if (reps > 0 && stale) {
        for (size_t k = 0; k < n; ++k) GHIP(g, hipStreamSynchronize(g->ctx[k]->stream));
        int rc = group_capture(g);
        if (rc) return rc;
    }
    // step counters, then the graph launches on the lead stream after them
    for (size_t k = 0; k < n; ++k) {
        ggb_ctx *c = g->ctx[k];
        int h[2] = {step0[k], step0[k]};
        GHIP(g, hipMemcpyAsync(c->dcur, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
        GHIP(g, hipEventRecord(g->ready[k], c->stream));
        GHIP(g, hipStreamWaitEvent(g->lead, g->ready[k], 0));
    }
    for (int r = 0; r < reps; ++r) GHIP(g, hipGraphLaunch(g->exec, g->lead));
    GHIP(g, hipEventRecord(g->fork, g->lead));
    // the steps after the last whole graph: plain launches on each context's stream
    for (size_t k = 0; k < n; ++k) {
        ggb_ctx *c = g->ctx[k];
        GHIP(g, hipStreamWaitEvent(c->stream, g->fork, 0));
        for (int s = reps * G; s < nsteps; ++s) {
            const int step = (int)(((long)step0[k] + s) % c->n_field);
            int rc = launch_first_half(c, step);
            if (!rc) rc = launch_kinetic(c, true);
            if (!rc)
                rc = launch_second_half(c, step, c->obs + (size_t)s * c->C * 3, nullptr, no_mark,
                                        c->channels ? c->chan + (size_t)s * c->C * c->p.n_l : nullptr);
            if (rc) return gfail(g, rc, "ggb_group_run: " + c->err);
        }
    }
    for (size_t k = 0; k < n; ++k) {
        ggb_ctx *c = g->ctx[k];
        GHIP(g, hipStreamSynchronize(c->stream));
        if (!(dip && dip[k]) && !(acc && acc[k]) && !(norm && norm[k])) continue;
        const size_t need = (size_t)nsteps * c->C * 3;
        std::vector<double> h(need);
        GHIP(g, hipMemcpyAsync(h.data(), c->obs, need * 8, hipMemcpyDeviceToHost, c->stream));
        GHIP(g, hipStreamSynchronize(c->stream));
        for (size_t q = 0; q < (size_t)nsteps * c->C; ++q) {
            if (dip && dip[k]) dip[k][q] = h[3 * q + 0];
            if (acc && acc[k]) acc[k][q] = h[3 * q + 1];
            if (norm && norm[k]) norm[k][q] = h[3 * q + 2];
        }
    }
    GHIP(g, hipStreamSynchronize(g->lead));
    return GGB_OK;
}

extern "C" int ggb_set_channels(ggb_ctx *ctx, int32_t on) {
    if (!ctx) return GGB_E_ARG;
    if (on != 0 && on != 1) return fail(ctx, GGB_E_ARG, "ggb_set_channels: on must be 0 or 1");
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->channels != (on == 1)) drop_graph(ctx);  // the captured step changes
    ctx->channels = on == 1;
    ctx->chan_rows = 0;
    return GGB_OK;
}

extern "C" int ggb_get_channels(ggb_ctx *ctx, int32_t nsteps, double *chan) {
    if (!ctx) return GGB_E_ARG;
    if (!chan) return fail(ctx, GGB_E_ARG, "ggb_get_channels: null buffer");
    if (nsteps != ctx->chan_rows || nsteps < 1)
        return fail(ctx, GGB_E_STATE, "ggb_get_channels: the last ggb_run did not record that many steps");
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(chan, ctx->chan, (size_t)nsteps * ctx->C * ctx->p.n_l * sizeof(double2),
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GGB_OK;
}

extern "C" int ggb_download_orbitals(ggb_ctx *ctx, double *psi) {
    if (!ctx) return GGB_E_ARG;
    if (!psi) return fail(ctx, GGB_E_ARG, "ggb_download_orbitals: null psi");
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(psi, ctx->psi, (size_t)ctx->p.n_l * ctx->C * ctx->p.n_r * 16,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GGB_OK;
}

extern "C" int ggb_download_potential(ggb_ctx *ctx, double *dv) {
    if (!ctx) return GGB_E_ARG;
    if (!dv) return fail(ctx, GGB_E_ARG, "ggb_download_potential: null buffer");
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(dv, ctx->dv, (size_t)ctx->p.n_inst * ctx->p.n_r * ctx->p.n_theta * 8,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GGB_OK;
}

extern "C" int ggb_debug_substep(ggb_ctx *ctx, int32_t which, int32_t step) {
    if (!ctx) return GGB_E_ARG;
    if (!ctx->have_orbitals) return fail(ctx, GGB_E_STATE, "ggb_debug_substep: orbitals not uploaded");
    if (step < 0 || (!ctx->field_ring && step >= ctx->n_field))
        return fail(ctx, GGB_E_ARG, "ggb_debug_substep: step outside field");
    step = step % ctx->n_field;
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    int rc;
    switch (which) {
        case 0: rc = launch_first_half(ctx, step); break;
        case 1: rc = launch_kinetic(ctx); break;
        case 2: {
            launch_angular<1>(ctx, step);
            HIPCHK(ctx, hipGetLastError());
            rc = GGB_OK;
            break;
        }
        case 3: {
            launch_angular<2>(ctx, step);
            HIPCHK(ctx, hipGetLastError());
            rc = launch_potential(ctx);
            break;
        }
        default: return fail(ctx, GGB_E_ARG, "ggb_debug_substep: unknown sub-step");
    }
    if (rc) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GGB_OK;
}

extern "C" int ggb_selftest_mfma(int32_t device, double *out) {
    if (!out) return GGB_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return GGB_E_HIP;
    double *d = nullptr;
    if (hipMalloc((void **)&d, 512 * sizeof(double)) != hipSuccess) return GGB_E_NOMEM;
    k_selftest_mfma<<<1, 64>>>(d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, d, 512 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? GGB_OK : GGB_E_HIP;
}

extern "C" int ggb_set_timing(ggb_ctx *ctx, int32_t mode) {
    if (!ctx) return GGB_E_ARG;
    if (mode < 0 || mode > 2) return fail(ctx, GGB_E_ARG, "ggb_set_timing: mode must be 0, 1 or 2");
    ctx->timing = mode;
    for (int k = 0; k < GGB_NPHASE; ++k) {
        ctx->phase_ms[k] = 0.0;
        ctx->phase_n[k] = 0;
    }
    return GGB_OK;
}

extern "C" int ggb_get_timing(const ggb_ctx *ctx, double *ms_total, int64_t *count) {
    if (!ctx || !ms_total || !count) return GGB_E_ARG;
    for (int k = 0; k < GGB_NPHASE; ++k) {
        ms_total[k] = ctx->phase_ms[k];
        count[k] = ctx->phase_n[k];
    }
    return GGB_OK;
}

// radix-2 Stockham FFT of length L = 2^m in place of buf (tmp: scratch of L)
static hipError_t fft_pow2(double2 *buf, double2 *tmp, long L, double sign, hipStream_t st) {
    double2 *in = buf, *out = tmp;
    const unsigned nb = (unsigned)((L / 2 + WG - 1) / WG);
    for (long Ns = 1; Ns < L; Ns *= 2) {
        k_fft_pass<<<nb, WG, 0, st>>>(in, out, L, Ns, sign);
        double2 *t = in;
        in = out;
        out = t;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && in != buf) e = hipMemcpyAsync(buf, in, L * sizeof(double2), hipMemcpyDeviceToDevice, st);
    return e;
}

// device buffers of one spectrum: the signal and its windowed copy (n), the
// power spectrum (M), the three Bluestein/FFT buffers (L) and the harmonic
// orders and intensities (n_orders)
struct SpecBufs {
    double *sig, *x, *pow, *ord, *inten;
    double2 *A, *B, *T;
};

static bool spec_args_ok(const double *signal, int32_t n, double dt, double beta, int32_t n_pad,
                         double omega0, const double *orders, int32_t n_orders,
                         const double *intensity, const char *who) {
    if (!signal || n < 16 || n_pad < 0 || !(dt > 0.0) || !(beta >= 0.0) || n_orders < 0 ||
        (n_orders > 0 && (!orders || !intensity || !(omega0 > 0.0)))) {
        g_create_err = std::string(who) + ": bad arguments";
        return false;
    }
    if ((2L * n_pad + 1) * n > (1L << 27)) {
        g_create_err = std::string(who) + ": (2 n_pad + 1) n exceeds 2^27";
        return false;
    }
    return true;
}

static long spec_fft_length(long M) {
    long L = 1;
    while (L < 2 * M - 1) L *= 2;
    return L;
}

// the spectrum on stream st into the buffers b; copies in and out, then
// waits for st.  Returns GGB_OK or GGB_E_HIP (message in g_create_err).
static int spectrum_on(const SpecBufs &b, hipStream_t st, const double *signal, int32_t n, double dt,
                       int32_t differentiate, double beta, int32_t n_pad, double omega0,
                       const double *orders, int32_t n_orders, double *power, double *intensity,
                       const char *who) {
    const long M = (2L * n_pad + 1) * n, L = spec_fft_length(M);
    int rc = GGB_OK;
    auto hip = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == GGB_OK) {
            g_create_err = std::string(who) + ": " + what + ": " + hipGetErrorString(e);
            rc = GGB_E_HIP;
        }
        return rc == GGB_OK;
    };
    if (!hip(hipMemcpyAsync(b.sig, signal, (size_t)n * 8, hipMemcpyHostToDevice, st), "upload"))
        return rc;
    const unsigned gn = (unsigned)((n + WG - 1) / WG);
    const unsigned gL = (unsigned)((L + WG - 1) / WG);
    const unsigned gM = (unsigned)((M + WG - 1) / WG);
    k_spec_window<<<gn, WG, 0, st>>>(b.sig, b.x, n, beta);
    double *x = b.x;
    if (differentiate) {
        k_spec_fd8<<<gn, WG, 0, st>>>(b.x, b.sig, n, 1.0 / (dt * dt));
        x = b.sig;
    }
    k_blue_a<<<gL, WG, 0, st>>>(x, b.A, n, n_pad, M, L);
    k_blue_b<<<gL, WG, 0, st>>>(b.B, M, L);
    if (hip(hipGetLastError(), "launch") && hip(fft_pow2(b.A, b.T, L, -1.0, st), "fft") &&
        hip(fft_pow2(b.B, b.T, L, -1.0, st), "fft")) {
        k_cmul<<<gL, WG, 0, st>>>(b.A, b.B, L);
        if (hip(fft_pow2(b.A, b.T, L, 1.0, st), "inverse fft")) {
            k_blue_power<<<gM, WG, 0, st>>>(b.A, b.pow, M, L);
            if (n_orders > 0 &&
                hip(hipMemcpyAsync(b.ord, orders, (size_t)n_orders * 8, hipMemcpyHostToDevice, st), "upload"))
                k_harmonics<<<n_orders, WG, 0, st>>>(b.pow, b.ord, b.inten, M, dt, omega0);
            if (hip(hipGetLastError(), "launch") && power)
                hip(hipMemcpyAsync(power, b.pow, (size_t)M * 8, hipMemcpyDeviceToHost, st), "download");
            if (rc == GGB_OK && n_orders > 0)
                hip(hipMemcpyAsync(intensity, b.inten, (size_t)n_orders * 8, hipMemcpyDeviceToHost, st),
                    "download");
        }
    }
    hip(hipStreamSynchronize(st), "synchronize");
    return rc;
}

extern "C" int ggb_spectrum(int32_t device, const double *signal, int32_t n, double dt,
                            int32_t differentiate, double beta, int32_t n_pad, double omega0,
                            const double *orders, int32_t n_orders, double *power, double *intensity) {
    if (!spec_args_ok(signal, n, dt, beta, n_pad, omega0, orders, n_orders, intensity, "ggb_spectrum"))
        return GGB_E_ARG;
    const long M = (2L * n_pad + 1) * n, L = spec_fft_length(M);
    if (hipSetDevice(device) != hipSuccess) {
        g_create_err = "ggb_spectrum: hipSetDevice failed";
        return GGB_E_HIP;
    }
    hipStream_t st = nullptr;
    SpecBufs b = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc = GGB_E_HIP;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        g_create_err = "ggb_spectrum: stream creation failed";
    } else if (dalloc(&b.sig, (size_t)n) != hipSuccess || dalloc(&b.x, (size_t)n) != hipSuccess ||
               dalloc(&b.pow, (size_t)M) != hipSuccess || dalloc(&b.A, (size_t)L) != hipSuccess ||
               dalloc(&b.B, (size_t)L) != hipSuccess || dalloc(&b.T, (size_t)L) != hipSuccess ||
               dalloc(&b.ord, (size_t)(n_orders > 0 ? n_orders : 1)) != hipSuccess ||
               dalloc(&b.inten, (size_t)(n_orders > 0 ? n_orders : 1)) != hipSuccess) {
        g_create_err = "ggb_spectrum: device allocation failed";
        rc = GGB_E_NOMEM;
    } else {
        rc = spectrum_on(b, st, signal, n, dt, differentiate, beta, n_pad, omega0, orders, n_orders,
                         power, intensity, "ggb_spectrum");
    }
    void *bufs[] = {b.sig, b.x, b.pow, b.ord, b.inten, b.A, b.B, b.T};
    for (void *q : bufs)
        if (q) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free spectrum: the buffers of one spectrum live in a file-scope
// device array of fixed size (each device holds its own copy until the process
// ends).  Host code never passes the array's name to a kernel: its device
// address is looked up once per device (hipGetSymbolAddress) and that address
// is carved into the buffers.  Capacity: FFT length L <= 2^20, so
// M = (2 n_pad + 1) n <= 2^19 (L >= 2M - 1), and at most 256 harmonic orders.
// ---------------------------------------------------------------------------
#define SPEC_NA_L (1L << 20)
#define SPEC_NA_ORDERS 256
#define SPEC_NA_DEVICES 64
static constexpr size_t SPEC_NA_BYTES =
    3 * (size_t)SPEC_NA_L * sizeof(double2)          // A, B, T
    + 3 * (size_t)(SPEC_NA_L / 2) * sizeof(double)   // sig, x (n <= M), pow (M <= L / 2)
    + 2 * (size_t)SPEC_NA_ORDERS * sizeof(double);   // orders, intensities
static __device__ __align__(256) unsigned char scratch[SPEC_NA_BYTES];
static void *g_scratch_ptr[SPEC_NA_DEVICES];
static std::mutex g_scratch_mu;

extern "C" int ggb_spectrum_noalloc(int32_t device, const double *signal, int32_t n, double dt,
                                    int32_t differentiate, double beta, int32_t n_pad, double omega0,
                                    const double *orders, int32_t n_orders, double *power,
                                    double *intensity) {
    if (!spec_args_ok(signal, n, dt, beta, n_pad, omega0, orders, n_orders, intensity,
                      "ggb_spectrum_noalloc"))
        return GGB_E_ARG;
    const long M = (2L * n_pad + 1) * n;
    if (spec_fft_length(M) > SPEC_NA_L || n_orders > SPEC_NA_ORDERS) {
        g_create_err = "ggb_spectrum_noalloc: (2 n_pad + 1) n must be <= 2^19 (FFT length <= 2^20) "
                       "and n_orders <= 256; use ggb_spectrum for larger spectra";
        return GGB_E_ARG;
    }
    if (device < 0 || device >= SPEC_NA_DEVICES) {
        g_create_err = "ggb_spectrum_noalloc: device ordinal out of range";
        return GGB_E_ARG;
    }
    if (hipSetDevice(device) != hipSuccess) {
        g_create_err = "ggb_spectrum_noalloc: hipSetDevice failed";
        return GGB_E_HIP;
    }
    unsigned char *base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void *ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                g_create_err = std::string("ggb_spectrum_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e);
                return GGB_E_HIP;
            }
            g_scratch_ptr[device] = ptr;
        }
        base = static_cast<unsigned char *>(g_scratch_ptr[device]);
    }
    SpecBufs b;
    b.A = reinterpret_cast<double2 *>(base);
    b.B = b.A + SPEC_NA_L;
    b.T = b.B + SPEC_NA_L;
    b.sig = reinterpret_cast<double *>(b.T + SPEC_NA_L);
    b.x = b.sig + SPEC_NA_L / 2;
    b.pow = b.x + SPEC_NA_L / 2;
    b.ord = b.pow + SPEC_NA_L / 2;
    b.inten = b.ord + SPEC_NA_ORDERS;
    // every buffer is written before it is read (window, chirps, FFT passes,
    // power), so nothing needs clearing between calls
    return spectrum_on(b, nullptr, signal, n, dt, differentiate, beta, n_pad, omega0, orders, n_orders,
                       power, intensity, "ggb_spectrum_noalloc");
}

extern "C" int ggb_set_kinetic_variant(ggb_ctx *ctx, int32_t variant) {
    if (!ctx) return GGB_E_ARG;
    const bool split = ctx->p.precision == 2;
    if (variant < GGB_KIN_DEFAULT || variant > GGB_KIN_SPLIT16 ||
        (variant == GGB_KIN_TILED32 && ctx->p.precision == 0) ||
        (split && variant != GGB_KIN_DEFAULT && variant != GGB_KIN_SPLIT16) ||
        (!split && variant == GGB_KIN_SPLIT16))
        return fail(ctx, GGB_E_ARG, "ggb_set_kinetic_variant: unknown or unsupported variant");
    if (variant != ctx->kin_variant) drop_graph(ctx);
    ctx->kin_variant = variant;
    return GGB_OK;
}

extern "C" int ggb_autotune(ggb_ctx *ctx, int32_t reps, int32_t *chosen, double *ms) {
    if (!ctx) return GGB_E_ARG;
    if (reps < 1 || reps > 1000) return fail(ctx, GGB_E_ARG, "ggb_autotune: reps must be 1..1000");
    if (!ctx->have_static || !ctx->have_orbitals)
        return fail(ctx, GGB_E_STATE, "ggb_autotune: tables and orbitals must be uploaded");
    HIPCHK(ctx, hipSetDevice(ctx->p.device));
    drop_graph(ctx);
    const size_t bytes = (size_t)ctx->p.n_l * ctx->C * ctx->p.n_r * 2 * sizeof(double);
    double *saved = nullptr;
    HIPCHK(ctx, hipMalloc((void **)&saved, bytes));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double t[GGB_KIN_NVARIANT] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int best = GGB_KIN_DEFAULT, rc = GGB_OK;
    const int prev = ctx->kin_variant;
    hipError_t e = hipMemcpyAsync(saved, ctx->psi, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const int first = ctx->p.precision == 2 ? GGB_KIN_SPLIT16 : GGB_KIN_NARROW;
    const int last = ctx->p.precision == 2 ? GGB_KIN_SPLIT16 : ctx->p.precision ? GGB_KIN_TILED32 : GGB_KIN_TILED16;
    for (int v = first; v <= last && e == hipSuccess && rc == GGB_OK; ++v) {
        ctx->kin_variant = v;
        rc = launch_kinetic(ctx);  // warm-up
        if (rc == GGB_OK) e = hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps && rc == GGB_OK && e == hipSuccess; ++r) rc = launch_kinetic(ctx);
        if (rc == GGB_OK && e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
        if (rc == GGB_OK && e == hipSuccess) e = hipEventSynchronize(e1);
        float el = 0.f;
        if (rc == GGB_OK && e == hipSuccess) e = hipEventElapsedTime(&el, e0, e1);
        if (rc == GGB_OK && e == hipSuccess) {
            t[v] = (double)el / reps;
            if (best == GGB_KIN_DEFAULT || t[v] < t[best]) best = v;
        }
        if (e == hipSuccess)  // the passes advanced the orbitals: put them back
            e = hipMemcpyAsync(ctx->psi, saved, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(saved);
    if (rc == GGB_OK && e != hipSuccess) rc = fail(ctx, GGB_E_HIP, std::string("ggb_autotune: ") + hipGetErrorString(e));
    ctx->kin_variant = (rc == GGB_OK) ? best : prev;
    if (rc == GGB_OK && chosen) *chosen = best;
    if (rc == GGB_OK && ms)
        for (int v = 0; v < GGB_KIN_NVARIANT; ++v) ms[v] = t[v];
    return rc;
}

// ---------------------------------------------------------------------------
// Kinetic propagator tables on the device: the radial Hamiltonians
// H_l = -T/2 + diag(l(l+1)/2r^2 + v) of every l (column-major, one matrix per
// l), their eigenpairs by one batched divide-and-conquer call (rocSOLVER
// dsyevd, FP64, ascending eigenvalues), then the eigenvectors with eigenvalue
// <= ecut packed as Q[l][i][a] (a < nk_cap, zero beyond nkeep[l]).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_build_hl(const double *__restrict__ T, const double *__restrict__ v,
                                                 const double *__restrict__ r, double *__restrict__ H, int N) {
    const size_t idx = (size_t)blockIdx.x * WG + threadIdx.x;  // element (i, j) at j N + i
    const int l = blockIdx.y;
    if (idx >= (size_t)N * N) return;
    const int j = (int)(idx / N), i = (int)(idx % N);
    double h = -0.5 * T[(size_t)i * N + j];
    if (i == j) h += 0.5 * l * (l + 1.0) / (r[i] * r[i]) + v[i];
    H[(size_t)l * N * N + idx] = h;
}

__global__ __launch_bounds__(WG) void k_pack_q(const double *__restrict__ V, const int32_t *__restrict__ nkeep,
                                               double *__restrict__ Q, int N, int nk_cap) {
    const size_t idx = (size_t)blockIdx.x * WG + threadIdx.x;  // Q[l][i][a], a fastest
    const int l = blockIdx.y;
    if (idx >= (size_t)N * nk_cap) return;
    const int i = (int)(idx / nk_cap), a = (int)(idx % nk_cap);
    Q[(size_t)l * N * nk_cap + idx] = a < nkeep[l] ? V[(size_t)l * N * N + (size_t)a * N + i] : 0.0;
}

extern "C" int ggb_propagator_tables(int32_t device, int32_t n_r, int32_t n_l, const double *r,
                                     const double *T, const double *v, double ecut, int32_t nk_cap,
                                     double *Q, double *lam, int32_t *nkeep, int32_t *nk_needed) {
    auto bad = [](const char *m) {
        g_create_err = m;
        return GGB_E_ARG;
    };
    if (!r || !T || !v || !Q || !lam || !nkeep) return bad("ggb_propagator_tables: null argument");
    if (n_r < 4 || n_r > 4096 || n_l < 1 || n_l > 1024 || nk_cap < 1 || nk_cap > n_r + 15)
        return bad("ggb_propagator_tables: sizes out of range");
    if (hipSetDevice(device) != hipSuccess) {
        g_create_err = "ggb_propagator_tables: hipSetDevice failed";
        return GGB_E_HIP;
    }
    const size_t N = (size_t)n_r, NN = N * N;
    double *dT = nullptr, *dv = nullptr, *dr = nullptr, *dH = nullptr, *dD = nullptr, *dE = nullptr,
           *dQ = nullptr;
    int32_t *dk = nullptr;
    rocblas_int *dinfo = nullptr;
    hipStream_t st = nullptr;
    rocblas_handle hb = nullptr;
    int rc = GGB_OK;
    std::string err;
    auto hipok = [&](hipError_t e, const char *what) {
        if (e == hipSuccess) return true;
        if (rc == GGB_OK) {
            rc = (e == hipErrorOutOfMemory) ? GGB_E_NOMEM : GGB_E_HIP;
            err = std::string("ggb_propagator_tables: ") + what + ": " + hipGetErrorString(e);
        }
        return false;
    };
    std::vector<double> D((size_t)n_l * N);
    std::vector<int32_t> keep(n_l);
    std::vector<rocblas_int> info(n_l, 0);
    if (hipok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream") &&
        hipok(hipMalloc((void **)&dT, NN * 8), "alloc") && hipok(hipMalloc((void **)&dv, N * 8), "alloc") &&
        hipok(hipMalloc((void **)&dr, N * 8), "alloc") &&
        hipok(hipMalloc((void **)&dH, (size_t)n_l * NN * 8), "alloc") &&
        hipok(hipMalloc((void **)&dD, (size_t)n_l * N * 8), "alloc") &&
        hipok(hipMalloc((void **)&dE, (size_t)n_l * N * 8), "alloc") &&
        hipok(hipMalloc((void **)&dinfo, (size_t)n_l * sizeof(rocblas_int)), "alloc") &&
        hipok(hipMemcpyAsync(dT, T, NN * 8, hipMemcpyHostToDevice, st), "copy") &&
        hipok(hipMemcpyAsync(dv, v, N * 8, hipMemcpyHostToDevice, st), "copy") &&
        hipok(hipMemcpyAsync(dr, r, N * 8, hipMemcpyHostToDevice, st), "copy")) {
        k_build_hl<<<dim3((unsigned)((NN + WG - 1) / WG), n_l), WG, 0, st>>>(dT, dv, dr, dH, n_r);
        if (hipok(hipGetLastError(), "build")) {
            rocblas_status bs = rocblas_create_handle(&hb);
            if (bs == rocblas_status_success) bs = rocblas_set_stream(hb, st);
            if (bs == rocblas_status_success)
                bs = rocsolver_dsyevd_strided_batched(hb, rocblas_evect_original, rocblas_fill_upper, n_r, dH,
                                                      n_r, (rocblas_stride)NN, dD, (rocblas_stride)N, dE,
                                                      (rocblas_stride)N, dinfo, n_l);
            if (bs != rocblas_status_success) {
                rc = GGB_E_HIP;
                err = std::string("ggb_propagator_tables: rocsolver_dsyevd: ") + rocblas_status_to_string(bs);
            }
        }
        if (rc == GGB_OK && hipok(hipMemcpyAsync(D.data(), dD, D.size() * 8, hipMemcpyDeviceToHost, st), "copy") &&
            hipok(hipMemcpyAsync(info.data(), dinfo, info.size() * sizeof(rocblas_int), hipMemcpyDeviceToHost, st),
                  "copy") &&
            hipok(hipStreamSynchronize(st), "sync")) {
            int need = 1;
            for (int l = 0; l < n_l && rc == GGB_OK; ++l) {
                if (info[l] != 0) {
                    rc = GGB_E_HIP;
                    err = "ggb_propagator_tables: no convergence for l = " + std::to_string(l);
                    break;
                }
                int k = 0;
                while (k < n_r && D[(size_t)l * N + k] <= ecut) ++k;  // ascending
                keep[l] = k;
                need = std::max(need, k);
            }
            const int need16 = ((need + 15) / 16) * 16;
            if (nk_needed) *nk_needed = need16;
            if (rc == GGB_OK && need > nk_cap) {
                rc = GGB_E_ARG;
                err = "ggb_propagator_tables: nk_cap below the largest kept count (" + std::to_string(need16) + ")";
            }
            if (rc == GGB_OK && hipok(hipMalloc((void **)&dk, (size_t)n_l * 4), "alloc") &&
                hipok(hipMalloc((void **)&dQ, (size_t)n_l * N * nk_cap * 8), "alloc") &&
                hipok(hipMemcpyAsync(dk, keep.data(), (size_t)n_l * 4, hipMemcpyHostToDevice, st), "copy")) {
                k_pack_q<<<dim3((unsigned)((N * nk_cap + WG - 1) / WG), n_l), WG, 0, st>>>(dH, dk, dQ, n_r, nk_cap);
                if (hipok(hipGetLastError(), "pack") &&
                    hipok(hipMemcpyAsync(Q, dQ, (size_t)n_l * N * nk_cap * 8, hipMemcpyDeviceToHost, st), "copy") &&
                    hipok(hipStreamSynchronize(st), "sync")) {
                    for (int l = 0; l < n_l; ++l) {
                        nkeep[l] = keep[l];
                        for (int a = 0; a < nk_cap; ++a) lam[(size_t)l * nk_cap + a] = a < keep[l] ? D[(size_t)l * N + a] : 0.0;
                    }
                }
            }
        }
    }
    if (hb) (void)rocblas_destroy_handle(hb);
    if (st) (void)hipStreamSynchronize(st);
    for (void *q : {(void *)dT, (void *)dv, (void *)dr, (void *)dH, (void *)dD, (void *)dE, (void *)dQ, (void *)dk,
                    (void *)dinfo})
        if (q) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
    if (rc != GGB_OK) g_create_err = err;
    return rc;
}

extern "C" int ggb_selftest_copy(int32_t device, uint64_t bytes, double *ms) {
    if (bytes < 16 || bytes % 16 != 0 || !ms) return GGB_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return GGB_E_HIP;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return GGB_E_NOMEM;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return GGB_E_NOMEM;
    }
    hipError_t e = hipMemset(a, 0, bytes);
    hipEvent_t e0, e1;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const size_t n = bytes / 16;
    if (e == hipSuccess) {
        k_copy<<<2048, WG>>>((const double2 *)a, (double2 *)b, n);  // warm
        (void)hipEventRecord(e0, 0);
        k_copy<<<2048, WG>>>((const double2 *)a, (double2 *)b, n);
        (void)hipEventRecord(e1, 0);
        e = hipEventSynchronize(e1);
        float f = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
        *ms = f;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return e == hipSuccess ? GGB_OK : GGB_E_HIP;
}

// ---------------------------------------------------------------------------
// Ground state on the device: the step before the propagation (a15's
// orbitals and v_gs).  The host iteration it replaces is the imaginary-time
// iteration with the exact exponential of the frozen-potential radial
// Hamiltonian (one dense eigendecomposition per l and iteration); the
// reference computes its ground state by diagonalisation in
// src/scripts/phys/groundState.py:182-457.  Both converge to the same fixed
// point, the self-consistent spherical LDA ground state of the discretised
// problem: one real radial function u_s (u units) per closed (n, l) shell.
// One SCF iteration:
//   k_gs_potential  rho_u = sum_s occ_s u_s^2,  v_out = V_0[rho_u] + v_xc[rho],
//                   v <- v + mix (v_out - v)   (v = v_out at the first)
//   k_gs_sweep      one work-group per l group, `sweeps` passes over its
//                   shells in shell order; for shell s:
//                     x <- u_s made orthogonal to the group's lower shells
//                          (modified Gram-Schmidt) and normalised
//                     r  = H x - theta x,  theta = x.Hx,
//                          H = -T/2 + l(l+1)/2r^2 - Z/r + v
//                     w  = Q0 diag(1/(|lam0 - theta| + gamma)) Q0^T r, with
//                          (Q0, lam0) the eigenpairs of the bare-nucleus H
//                          (v = 0) of this l, orthogonalised to the lower
//                          shells and to x, normalised
//                     u_s <- the lower Ritz vector of span(x, w)
//                   eps_s = u_s.H u_s
// until max_s |eps_s - eps_s(prev)| < tol and max |v_out - v| < 1e3 tol (the
// host iteration's test).  An exact eigenvector has r = 0 and is left as it
// is, so the fixed point is the one of the exact iteration; the spectral
// preconditioner (positive definite) only sets the rate.  All FP64.
// ---------------------------------------------------------------------------
#define GS_WG 1024
#define GS_MAX_N 1536   // k_gs_sweep keeps 5 N-vectors in LDS (60 KiB) plus 8 KiB of scratch
#define GS_MAX_K 8      // shells per l group

struct GsArgs {
    const double *r, *wr, *T, *lam0, *occ;
    const float *Q0, *Q0T;  // bare-nucleus eigenvectors [g][i][k] and transposes (preconditioner)
    const int32_t *gsh;  // [n_group][GS_MAX_K] shells of each l group, shell order, -1 padded
    const int32_t *gl;   // [n_group] l of each group
    double *u;           // [n_shell][N]
    double *eps;         // [n_shell] (first iteration: +inf)
    double *v, *rho_u;   // [N]
    double *res;         // [0] max |v_out - v|, [1 + g] max |eps - eps_prev| of group g
    int N, n_shell, n_group, sweeps, interacting, first, final_;
    double Z, mix, gamma;
};

// sum over the work-group, fixed order (every thread gets the result)
static __device__ double gs_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();  // earlier readers of red are done
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
    return s;
}

static __device__ double gs_max(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off));
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s = fmax(s, red[k]);
    return s;
}

// inclusive scan (dir 0: prefix, 1: suffix) of one value per thread, Hillis-Steele in LDS
static __device__ double gs_scan(double v, int dir, double *buf) {
    const int t = threadIdx.x, n = blockDim.x;
    const int me = dir ? n - 1 - t : t;
    __syncthreads();
    buf[me] = v;
    __syncthreads();
    for (int off = 1; off < n; off <<= 1) {
        const double o = me >= off ? buf[me - off] : 0.0;
        __syncthreads();
        buf[me] += o;
        __syncthreads();
    }
    return buf[me];
}

// rho_u, V_0 + v_xc and the damped update (final_: v <- v_out, no mixing)
__global__ __launch_bounds__(GS_WG) void k_gs_potential(GsArgs a) {
    __shared__ double buf[GS_WG];
    __shared__ double red[GS_WG / 64];
    const int t = threadIdx.x, nt = blockDim.x, N = a.N;
    // thread t owns the radial chunk [i0, i1)
    const int ch = (N + nt - 1) / nt;
    const int i0 = min(N, t * ch), i1 = min(N, i0 + ch);
    double sq = 0.0, sqr = 0.0;
    for (int i = i0; i < i1; ++i) {
        double ru = 0.0;
        for (int s = 0; s < a.n_shell; ++s) {
            const double x = a.u[(size_t)s * N + i];
            ru += a.occ[s] * x * x;
        }
        a.rho_u[i] = ru;
        sq += ru;
        sqr += ru / a.r[i];
    }
    // V_0(r_i) = (1/r_i) sum_{b <= i} q_b + sum_{b > i} q_b / r_b
    const double pre = gs_scan(sq, 0, buf) - sq;    // sum over the earlier chunks
    const double suf = gs_scan(sqr, 1, buf) - sqr;  // sum over the later chunks
    double run = pre, dmax = 0.0;
    for (int i = i0; i < i1; ++i) {
        const double ru = a.rho_u[i], ri = a.r[i];
        run += ru;
        double tail = suf;
        for (int b = i + 1; b < i1; ++b) tail += a.rho_u[b] / a.r[b];
        double vo = 0.0;
        if (a.interacting) {
            const double rho = ru / (4.0 * M_PI * a.wr[i] * ri * ri);
            vo = run / ri + tail + lda_vxc<double>(rho);
        }
        const double vn = (a.final_ || a.first) ? vo : a.v[i] + a.mix * (vo - a.v[i]);
        a.v[i] = vn;
        dmax = fmax(dmax, fabs(vo - vn));
    }
    dmax = gs_max(dmax, red);
    if (t == 0) a.res[0] = dmax;
}

// out[c] = sum_j M[j][c] x[j] (M row-major N x N; threads read along c,
// coalesced).  With N < blockDim the j range is split into P = blockDim / N
// parts (more loads in flight), the parts added in fixed order through scr
// ((P - 1) N <= blockDim doubles); eight partial sums per thread.
template <typename TE>
static __device__ void gs_colsum(const TE *M, const double *x, int N, double *out, double *scr) {
    const int nt = blockDim.x;
    const int P = N < nt ? nt / N : 1;
    const int jl = (N + P - 1) / P;
    for (int idx = threadIdx.x; idx < P * N; idx += nt) {
        const int c = idx % N, part = idx / N;
        const int j0 = part * jl, j1 = min(N, j0 + jl);
        const TE *Mc = M + c;
        double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int j = j0;
        for (; j + 8 <= j1; j += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) s[u] += (double)Mc[(size_t)(j + u) * N] * x[j + u];
        }
        for (; j < j1; ++j) s[0] += (double)Mc[(size_t)j * N] * x[j];
        const double v = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
        if (part == 0)
            out[c] = v;
        else
            scr[(size_t)(part - 1) * N + c] = v;
    }
    __syncthreads();
    if (P > 1) {
        for (int c = threadIdx.x; c < N; c += nt) {
            double v = out[c];
            for (int q = 1; q < P; ++q) v += scr[(size_t)(q - 1) * N + c];
            out[c] = v;
        }
        __syncthreads();
    }
}

// y = H x for the group's l (T symmetric, read by columns)
static __device__ void gs_apply_h(const GsArgs &a, int l, const double *x, double *y, double *scr) {
    const int N = a.N;
    const double cl = 0.5 * (double)l * (double)(l + 1);
    gs_colsum(a.T, x, N, y, scr);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double ri = a.r[i];
        const double d = cl / (ri * ri) - a.Z / ri + a.v[i];
        y[i] = -0.5 * y[i] + d * x[i];
    }
}

// w = Q0 diag(1/(|lam0 - theta| + gamma)) Q0^T r, through tmp (all in LDS); the
// eigenvector tables in FP32 (the preconditioner only sets the rate)
static __device__ void gs_precondition(const GsArgs &a, int g, double theta, const double *rv,
                                       double *tmp, double *w, double *scr) {
    const int N = a.N;
    const float *Q = a.Q0 + (size_t)g * N * N;    // [i][k]
    const float *QT = a.Q0T + (size_t)g * N * N;  // [k][i]
    const double *lam = a.lam0 + (size_t)g * N;
    gs_colsum(Q, rv, N, tmp, scr);
    for (int k = threadIdx.x; k < N; k += blockDim.x) tmp[k] /= fabs(lam[k] - theta) + a.gamma;
    __syncthreads();
    gs_colsum(QT, tmp, N, w, scr);
}

static __device__ double gs_dot(const double *x, const double *y, int N, double *red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) s += x[i] * y[i];
    return gs_sum(s, red);
}

// x -= (q.x) q for each lower shell q of the group, in shell order
static __device__ void gs_deflate(const GsArgs &a, const int32_t *sh, int j, double *x, double *red) {
    const int N = a.N;
    for (int q = 0; q < j; ++q) {
        const double *uq = a.u + (size_t)sh[q] * N;
        const double c = gs_dot(uq, x, N, red);
        for (int i = threadIdx.x; i < N; i += blockDim.x) x[i] -= c * uq[i];
        __syncthreads();
    }
}

// one work-group per l group: `sweeps` preconditioned Rayleigh-Ritz passes
// over the group's shells with the potential a.v; eps and the group's
// largest eps change into a.res[1 + g]
__global__ __launch_bounds__(GS_WG) void k_gs_sweep(GsArgs a) {
    extern __shared__ double gsm[];
    __shared__ double red[GS_WG / 64];
    __shared__ double scr[GS_WG];
    const int g = blockIdx.x, N = a.N, l = a.gl[g];
    const int32_t *sh = a.gsh + (size_t)g * GS_MAX_K;
    int k = 0;
    while (k < GS_MAX_K && sh[k] >= 0) ++k;
    double *x = gsm, *hx = gsm + N, *w = gsm + 2 * N, *hw = gsm + 3 * N, *tmp = gsm + 4 * N;
    double dmax = 0.0;
    for (int sw = 0; sw < a.sweeps; ++sw) {
        for (int j = 0; j < k; ++j) {
            double *us = a.u + (size_t)sh[j] * N;
            for (int i = threadIdx.x; i < N; i += blockDim.x) x[i] = us[i];
            __syncthreads();
            gs_deflate(a, sh, j, x, red);
            const double nx = sqrt(gs_dot(x, x, N, red));
            for (int i = threadIdx.x; i < N; i += blockDim.x) x[i] /= nx;
            __syncthreads();
            gs_apply_h(a, l, x, hx, scr);
            __syncthreads();
            const double theta = gs_dot(x, hx, N, red);
            for (int i = threadIdx.x; i < N; i += blockDim.x) hw[i] = hx[i] - theta * x[i];  // residual
            __syncthreads();
            gs_precondition(a, g, theta, hw, tmp, w, scr);
            gs_deflate(a, sh, j, w, red);
            const double cx = gs_dot(x, w, N, red);
            for (int i = threadIdx.x; i < N; i += blockDim.x) w[i] -= cx * x[i];
            __syncthreads();
            const double nw = sqrt(gs_dot(w, w, N, red));
            // new x = ax x + aw w (and H x alike)
            double ax = 1.0, aw = 0.0;
            if (nw > 1e-300) {
                for (int i = threadIdx.x; i < N; i += blockDim.x) w[i] /= nw;
                __syncthreads();
                gs_apply_h(a, l, w, hw, scr);
                __syncthreads();
                const double b = gs_dot(x, hw, N, red);
                const double c = gs_dot(w, hw, N, red);
                // Jacobi rotation of [[theta, b], [b, c]] (t = tan of the angle):
                // eigenvalues theta - t b with vector (cs, -sn) and c + t b with
                // vector (sn, cs); the lower one is kept
                double tt = 0.0;
                if (b != 0.0) {
                    const double zeta = (c - theta) / (2.0 * b);
                    tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                }
                const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = tt * cs;
                if (c + tt * b < theta - tt * b)
                    ax = sn, aw = cs;
                else
                    ax = cs, aw = -sn;
            }
            for (int i = threadIdx.x; i < N; i += blockDim.x) {
                const double xn = ax * x[i] + aw * w[i];
                x[i] = xn;
                hx[i] = ax * hx[i] + aw * hw[i];  // H of the new vector, by linearity
                us[i] = xn;
            }
            __syncthreads();
            if (sw == a.sweeps - 1) {
                // orbital energy with the final vector (later shells leave it
                // unchanged) and the largest change (+inf at the first iteration)
                const double e = gs_dot(x, hx, N, red);
                dmax = fmax(dmax, fabs(e - a.eps[sh[j]]));
                __syncthreads();
                if (threadIdx.x == 0) a.eps[sh[j]] = e;
            }
        }
    }
    if (threadIdx.x == 0) a.res[1 + g] = dmax;
}

extern "C" int ggb_ground_state(const ggb_gs_problem *p, const int32_t *shell_l, const double *occ,
                                const double *r, const double *wr, const double *T,
                                const double *Q0, const double *lam0, double *u, double *eps,
                                double *v_gs, double *rho_u, int32_t *iterations,
                                int32_t *converged) {
    auto bad = [](const char *msg) {
        g_create_err = std::string("ggb_ground_state: ") + msg;
        return GGB_E_ARG;
    };
    if (!p || !shell_l || !occ || !r || !wr || !T || !Q0 || !lam0 || !u || !eps || !v_gs || !rho_u ||
        !iterations || !converged)
        return bad("null argument");
    const int N = p->n_r, ns = p->n_shell;
    if (N < 4 || N > GS_MAX_N) return bad("n_r outside 4..1536");
    if (ns < 1 || ns > 64) return bad("n_shell outside 1..64");
    if (p->sweeps < 1 || p->max_iter < 1 || !(p->tol > 0.0) || !(p->mix > 0.0 && p->mix <= 1.0) ||
        !(p->gamma > 0.0) || !(p->Z > 0.0))
        return bad("sweeps, max_iter, tol, mix (0, 1], gamma or Z out of range");
    // l groups: the distinct l of the shells, ascending; shells of a group in shell order
    std::vector<int> ls;
    for (int s = 0; s < ns; ++s) {
        if (shell_l[s] < 0 || shell_l[s] > 127 || !(occ[s] >= 0.0)) return bad("shell l or occupation");
        if (std::find(ls.begin(), ls.end(), shell_l[s]) == ls.end()) ls.push_back(shell_l[s]);
    }
    std::sort(ls.begin(), ls.end());
    const int G = (int)ls.size();
    std::vector<int32_t> gsh((size_t)G * GS_MAX_K, -1), gl(G);
    for (int g = 0; g < G; ++g) {
        gl[g] = ls[g];
        int k = 0;
        for (int s = 0; s < ns; ++s)
            if (shell_l[s] == ls[g]) {
                if (k >= GS_MAX_K) return bad("more than 8 shells of one l");
                if (k >= N) return bad("more shells of one l than radial nodes");
                gsh[(size_t)g * GS_MAX_K + k++] = s;
            }
    }
    if (hipSetDevice(p->device) != hipSuccess) {
        g_create_err = "ggb_ground_state: hipSetDevice failed";
        return GGB_E_HIP;
    }
    const size_t NN = (size_t)N * N;
    // transposed eigenvector tables and the start: u_s = the k-th bare-nucleus
    // eigenvector of its l (k = rank of s within its group)
    std::vector<float> q0f((size_t)G * NN), q0t((size_t)G * NN);
    std::vector<double> u0((size_t)ns * N), einit(ns, INFINITY);
    for (int g = 0; g < G; ++g) {
        const double *Qg = Q0 + (size_t)g * NN;
        float *QFg = q0f.data() + (size_t)g * NN, *QTg = q0t.data() + (size_t)g * NN;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < N; ++a) {
                QFg[(size_t)i * N + a] = (float)Qg[(size_t)i * N + a];
                QTg[(size_t)a * N + i] = (float)Qg[(size_t)i * N + a];
            }
        for (int k = 0; k < GS_MAX_K && gsh[(size_t)g * GS_MAX_K + k] >= 0; ++k) {
            const int s = gsh[(size_t)g * GS_MAX_K + k];
            for (int i = 0; i < N; ++i) u0[(size_t)s * N + i] = Qg[(size_t)i * N + k];
        }
    }
    hipStream_t st = nullptr;
    double *dr = nullptr, *dwr = nullptr, *dT = nullptr, *dlam = nullptr, *docc = nullptr, *du = nullptr,
           *deps = nullptr, *dv = nullptr, *drho = nullptr, *dres = nullptr;
    float *dQ = nullptr, *dQT = nullptr;
    int32_t *dgsh = nullptr, *dgl = nullptr;
    int rc = GGB_OK;
    auto hip = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == GGB_OK) {
            g_create_err = std::string("ggb_ground_state: ") + what + ": " + hipGetErrorString(e);
            rc = e == hipErrorOutOfMemory ? GGB_E_NOMEM : GGB_E_HIP;
        }
        return rc == GGB_OK;
    };
    auto up = [&](void *d, const void *h, size_t bytes) {
        return hip(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st), "upload");
    };
    int it = 0;
    bool conv = false;
    if (hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream") && hip(dalloc(&dr, N), "alloc") &&
        hip(dalloc(&dwr, N), "alloc") && hip(dalloc(&dT, NN), "alloc") &&
        hip(dalloc(&dQ, (size_t)G * NN), "alloc") && hip(dalloc(&dQT, (size_t)G * NN), "alloc") &&
        hip(dalloc(&dlam, (size_t)G * N), "alloc") && hip(dalloc(&docc, ns), "alloc") &&
        hip(dalloc(&du, (size_t)ns * N), "alloc") && hip(dalloc(&deps, ns), "alloc") &&
        hip(dalloc(&dv, N), "alloc") && hip(dalloc(&drho, N), "alloc") && hip(dalloc(&dres, 1 + G), "alloc") &&
        hip(dalloc(&dgsh, (size_t)G * GS_MAX_K), "alloc") && hip(dalloc(&dgl, G), "alloc") &&
        up(dr, r, N * 8) && up(dwr, wr, N * 8) && up(dT, T, NN * 8) && up(dQ, q0f.data(), (size_t)G * NN * 4) &&
        up(dQT, q0t.data(), (size_t)G * NN * 4) && up(dlam, lam0, (size_t)G * N * 8) &&
        up(docc, occ, ns * 8) && up(du, u0.data(), (size_t)ns * N * 8) && up(deps, einit.data(), ns * 8) &&
        up(dgsh, gsh.data(), gsh.size() * 4) && up(dgl, gl.data(), G * 4)) {
        GsArgs a;
        a.r = dr, a.wr = dwr, a.T = dT, a.Q0 = dQ, a.Q0T = dQT, a.lam0 = dlam, a.occ = docc;
        a.gsh = dgsh, a.gl = dgl, a.u = du, a.eps = deps, a.v = dv, a.rho_u = drho, a.res = dres;
        a.N = N, a.n_shell = ns, a.n_group = G, a.sweeps = p->sweeps, a.interacting = p->interacting ? 1 : 0;
        a.Z = p->Z, a.mix = p->mix, a.gamma = p->gamma;
        a.final_ = 0;
        const size_t lds = (size_t)5 * N * sizeof(double);
        std::vector<double> res(1 + G);
        for (it = 1; it <= p->max_iter; ++it) {
            a.first = it == 1;
            k_gs_potential<<<1, GS_WG, 0, st>>>(a);
            k_gs_sweep<<<G, GS_WG, lds, st>>>(a);
            if (!hip(hipGetLastError(), "launch") ||
                !hip(hipMemcpyAsync(res.data(), dres, (1 + G) * 8, hipMemcpyDeviceToHost, st), "download") ||
                !hip(hipStreamSynchronize(st), "sync"))
                break;
            double de = 0.0;
            for (int g = 0; g < G; ++g) de = std::max(de, res[1 + g]);
            if (de < p->tol && res[0] < 1e3 * p->tol) {
                conv = true;
                break;
            }
        }
        if (it > p->max_iter) it = p->max_iter;
        if (rc == GGB_OK) {
            a.first = 0, a.final_ = 1;  // v_gs = v_out of the final orbitals
            k_gs_potential<<<1, GS_WG, 0, st>>>(a);
            (void)(hip(hipGetLastError(), "launch") &&
                   hip(hipMemcpyAsync(u, du, (size_t)ns * N * 8, hipMemcpyDeviceToHost, st), "download") &&
                   hip(hipMemcpyAsync(eps, deps, ns * 8, hipMemcpyDeviceToHost, st), "download") &&
                   hip(hipMemcpyAsync(v_gs, dv, N * 8, hipMemcpyDeviceToHost, st), "download") &&
                   hip(hipMemcpyAsync(rho_u, drho, N * 8, hipMemcpyDeviceToHost, st), "download") &&
                   hip(hipStreamSynchronize(st), "sync"));
        }
    }
    for (void *d : {(void *)dr, (void *)dwr, (void *)dT, (void *)dQ, (void *)dQT, (void *)dlam, (void *)docc,
                    (void *)du, (void *)deps, (void *)dv, (void *)drho, (void *)dres, (void *)dgsh, (void *)dgl})
        if (d) (void)hipFree(d);
    if (st) (void)hipStreamDestroy(st);
    *iterations = it;
    *converged = conv ? 1 : 0;
    return rc;
}
