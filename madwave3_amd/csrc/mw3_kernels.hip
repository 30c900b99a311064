// mw3_kernels.hip -- gfx950 kernels of libmw3.
//
// One Chebyshev step of SRC/main_madwave3.f:324-353 is:
//   radial kinetic  (SRC/mod_Hphi_01y2.f:762-832): 2-D DST-I, scale, 2-D DST-I
//                   as three passes over the dense slabs (R rows, r columns
//                   with scaling, R rows);
//   angular term    (SRC/mod_Hphi_01y2.f:859-902) in factor form:
//                   FBR = D^T psi, stencil over Omega with the j^2 / l^2
//                   coefficients, back with D;
//   fused update    V psi (SRC/mod_Hphi_01y2.f:838-853), compressed-grid
//                   projection, damped three-term update
//                   (SRC/main_madwave3.f:332-339), norm partial sums;
//   projections     flux-line samples (SRC/mod_flux_01y2.f:598-628) and
//                   reactant-state amplitudes (SRC/mod_flux_01y2.f:672-702)
//                   into a ring of kRing steps; the energy sums
//                   sum_k C_k(E) x_k (CoefEcheby, SRC/main_madwave3.f:462-490)
//                   run once per ring as a complex GEMM.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "mw3_internal.h"

namespace mw3 {

// ---------------------------------------------------------------------------
// batched GEMM operands: C[m][n] = sum_k A[m][k] B[k][n] per descriptor
// ---------------------------------------------------------------------------
// operand element with the DST fold (see GemmDesc)
template <typename TACC, typename T>
__device__ __forceinline__ TACC ld_fold(const T* base, int64_t idx, int64_t kstride, int k,
                                        int fold_n, int fold_sgn) {
  TACC v = (TACC)base[idx];
  if (fold_n > 0) {
    int mk = fold_n - 1 - k;
    if (mk != k) {
      TACC w = (TACC)base[idx + (int64_t)(mk - k) * kstride];
      v = fold_sgn > 0 ? v + w : v - w;
    }
  }
  return v;
}

struct KinParams {
  const int32_t* nn1;
  const double* p2r1;
  const double* p2r2;
  const double* divi;
  double radcut;
  int n1;
  // angular FBR coefficients (kin_mode 2)
  const double* acoef;
  const double* bcoef;
  int nJ, n2;
};

// radial kinetic factor of mode (k1 = rmul m + radd, k2 = n) of angle kin_ia:
// kin_mode 0: min(E, radcut) divi (Tradial, SRC/mod_Hphi_01y2.f:770-771);
// kin_mode 1: (radcut - E)_+ divi, the low-mode correction of c psi
// kin_mode 2: the diagonal FBR coefficient a_j(r1) + b_j(R) of channel
// kin_ia for output row m = j and column n = r n2 + R (angkinj2 / bfgridl2mat
// in factor form, used when no channel has a Coriolis neighbour)
__device__ __forceinline__ double kin_factor(const GemmDesc& d, const KinParams& kp, int m, int n) {
  if (d.kin_mode == 2) {
    const int r = n / kp.n2, R = n - r * kp.n2;
    return kp.acoef[((int64_t)d.kin_ia * kp.nJ + m) * kp.n1 + r] +
           kp.bcoef[(((int64_t)d.kin_ia * 3 + 1) * kp.nJ + m) * kp.n2 + R];
  }
  int k1 = d.rmul * m + d.radd;
  double e = kp.p2r1[(int64_t)d.kin_ia * kp.n1 + k1] + kp.p2r2[n];
  if (d.kin_mode) {
    e = kp.radcut - e;
    if (e < 0.0) e = 0.0;
  } else if (e > kp.radcut) {
    e = kp.radcut;
  }
  return e * kp.divi[d.kin_ia];
}


// ---------------------------------------------------------------------------
// MFMA GEMM, FP64 (v_mfma_f64_16x16x4f64) or FP32 (v_mfma_f32_16x16x4f32).
// 256 threads = 2 x 2 waves; each wave owns TM x TN MFMA tiles of 16 x 16, so
// the block tile is (32 TM) x (32 TN).  K is staged through LDS BK at a time;
// the next K tile is loaded into registers while the current one is in the
// MFMAs.  Operand loads run along whichever of the operand's two axes is
// contiguous in memory.  ROLE only names the instantiation (one symbol per
// pass of the step, so kernel traces separate them).
// Lane layout (wave64): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k;
// D[i][j] in lane j + 16 (i % 4), register i / 4 (FP64) and in lane
// j + 16 (i / 4), register i % 4 (FP32).
// ---------------------------------------------------------------------------
constexpr int BK = 16;
typedef double mw3_d4 __attribute__((ext_vector_type(4)));
typedef float mw3_f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
  typedef mw3_d4 V;
  static __device__ __forceinline__ V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <>
struct MfmaAcc<float> {
  typedef mw3_f4 V;
  static __device__ __forceinline__ V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// element e of a (rows x BK) A tile or (BK x cols) B tile -> (row/col, k)
template <int ROWS>
__device__ __forceinline__ void tile_pos(int e, bool rowfast, int& rc, int& kk) {
  if (rowfast) {
    rc = e % ROWS;
    kk = e / ROWS;
  } else {
    rc = e / BK;
    kk = e % BK;
  }
}

template <typename T, int TM, int TN, typename TA, typename TB>
__device__ __forceinline__ void mfma_load_tiles(const GemmDesc& d, const GemmStrides& st,
                                                const TA* Ab, const TB* Bb, int m0, int n0, int k0,
                                                bool amf, bool bnf, T (&ra)[2 * TM],
                                                T (&rb)[2 * TN]) {
  constexpr int BMt = 32 * TM, BNt = 32 * TN;
  const bool fa = d.fold_op == 1, fb = d.fold_op == 2;
#pragma unroll
  for (int l = 0; l < 2 * TM; l++) {
    int mm, kk;
    tile_pos<BMt>(threadIdx.x + l * 256, amf, mm, kk);
    int m = m0 + mm, k = k0 + kk;
    T v = 0;
    if (m < d.M && k < d.K) {
      int64_t idx = (int64_t)m * d.lda_m + (int64_t)k * d.lda_k;
      v = fa ? ld_fold<T>(Ab, idx, d.lda_k, k, d.fold_n, d.fold_sgn) : (T)Ab[idx];
    }
    ra[l] = v;
  }
#pragma unroll
  for (int l = 0; l < 2 * TN; l++) {
    int nn, kk;
    tile_pos<BNt>(threadIdx.x + l * 256, bnf, nn, kk);
    int k = k0 + kk, n = n0 + nn;
    T v = 0;
    if (k < d.K && n < d.N) {
      int64_t idx = (int64_t)k * st.ldb_k + (int64_t)n * st.ldb_n;
      v = fb ? ld_fold<T>(Bb, idx, st.ldb_k, k, d.fold_n, d.fold_sgn) : (T)Bb[idx];
    }
    rb[l] = v;
  }
}

template <int ROLE, typename T, int TM, int TN, typename TA, typename TB, typename TC>
__global__ __launch_bounds__(256) void k_gemm_mfma(const TA* __restrict__ A, const TB* __restrict__ B,
                                                   TC* __restrict__ C,
                                                   const GemmDesc* __restrict__ descs,
                                                   GemmStrides st, KinParams kp) {
  typedef MfmaAcc<T> M;
  constexpr int BMt = 32 * TM, BNt = 32 * TN;
  const GemmDesc d = descs[blockIdx.z];
  const int m0 = blockIdx.y * BMt, n0 = blockIdx.x * BNt;
  if (m0 >= d.M || n0 >= d.N) return;
  __shared__ T As[BK][BMt + 4];
  __shared__ T Bs[BK][BNt + 4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const bool amf = d.lda_m == 1, bnf = st.ldb_n == 1;
  typename M::V acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++) acc[a][b] = (typename M::V){0, 0, 0, 0};
  const TA* Ab = A + d.a_off;
  const TB* Bb = B + d.b_off;
  T ra[2 * TM], rb[2 * TN];
  mfma_load_tiles<T, TM, TN>(d, st, Ab, Bb, m0, n0, 0, amf, bnf, ra, rb);
  for (int k0 = 0; k0 < d.K; k0 += BK) {
#pragma unroll
    for (int l = 0; l < 2 * TM; l++) {
      int mm, kk;
      tile_pos<BMt>(threadIdx.x + l * 256, amf, mm, kk);
      As[kk][mm] = ra[l];
    }
#pragma unroll
    for (int l = 0; l < 2 * TN; l++) {
      int nn, kk;
      tile_pos<BNt>(threadIdx.x + l * 256, bnf, nn, kk);
      Bs[kk][nn] = rb[l];
    }
    __syncthreads();
    if (k0 + BK < d.K) mfma_load_tiles<T, TM, TN>(d, st, Ab, Bb, m0, n0, k0 + BK, amf, bnf, ra, rb);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4), c16 = lane & 15;
      T av[TM], bv[TN];
#pragma unroll
      for (int a = 0; a < TM; a++) av[a] = As[kr][wm * 16 * TM + 16 * a + c16];
#pragma unroll
      for (int b = 0; b < TN; b++) bv[b] = Bs[kr][wn * 16 * TN + 16 * b + c16];
#pragma unroll
      for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++) acc[a][b] = M::mma(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
  TC* Cb = C + d.c_off;
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int m = m0 + wm * 16 * TM + a * 16 + M::row(lane, r);
        int n = n0 + wn * 16 * TN + b * 16 + (lane & 15);
        if (m < d.M && n < d.N) {
          double v = (double)acc[a][b][r];
          if (d.kin_ia >= 0) v *= kin_factor(d, kp, m, n);
          Cb[(int64_t)m * st.ldc_m + (int64_t)n * st.ldc_n] = (TC)v;
        }
      }
}

// ---------------------------------------------------------------------------
// Vectorised MFMA GEMM for plans whose operand rows are 16-byte aligned
// (offsets, strides and fold length multiples of 4; checked on the host):
// every operand access is a 4-element vector along the operand's contiguous
// axis.  A is K-contiguous for the R passes and the angular back transform
// (rows of psi / of D) and M-contiguous for the r passes and the angular
// forward transform (columns of the tables); B is always N-contiguous.
// The MFMA k index of lane group g = lane / 16 at step s is 4 g + s, so a
// K-contiguous A row hands each lane its four k values in one LDS read.
// ---------------------------------------------------------------------------
template <typename T, typename TS>
__device__ __forceinline__ void ld4(const TS* p, T (&v)[4]) {
  typedef TS TS4 __attribute__((ext_vector_type(4)));
  TS4 x = *(const TS4*)p;
  v[0] = (T)x.x;
  v[1] = (T)x.y;
  v[2] = (T)x.z;
  v[3] = (T)x.w;
}

template <typename T>
__device__ __forceinline__ void st4(T* p, const T (&v)[4]) {
  typedef T T4 __attribute__((ext_vector_type(4)));
  *(T4*)p = (T4){v[0], v[1], v[2], v[3]};
}

// A row m, elements k .. k+3 (K-contiguous), folded along K when fold_op == 1
template <typename T, typename TA>
__device__ __forceinline__ void ld_a_kc(const GemmDesc& d, const TA* Ab, int m, int k, T (&v)[4]) {
#pragma unroll
  for (int e = 0; e < 4; e++) v[e] = 0;
  if (m < d.M && k < d.K) {
    const TA* row = Ab + (int64_t)m * d.lda_m;
    ld4(row + k, v);
    if (d.fold_op == 1) {
      // element k+e pairs with fold_n-1-k-e = (fold_n-4-k) + (3-e)
      T w[4];
      ld4(row + (d.fold_n - 4 - k), w);
#pragma unroll
      for (int e = 0; e < 4; e++) v[e] = d.fold_sgn > 0 ? v[e] + w[3 - e] : v[e] - w[3 - e];
    }
#pragma unroll
    for (int e = 0; e < 4; e++)
      if (k + e >= d.K) v[e] = 0;
  }
}

// operand row k (stride ldk), elements i .. i+3 of the contiguous axis (extent
// MN), folded along the rows when fold is set
template <typename T, typename TS>
__device__ __forceinline__ void ld_mn(const TS* base, int64_t ldk, int k, int i, int K, int MN,
                                      bool fold, int fold_n, int sgn, T (&v)[4]) {
#pragma unroll
  for (int e = 0; e < 4; e++) v[e] = 0;
  if (k < K && i < MN) {
    const TS* row = base + (int64_t)k * ldk + i;
    const int mk = fold_n - 1 - k;
    const bool f = fold && mk != k;
    const int64_t dp = (int64_t)(mk - k) * ldk;
    if (i + 3 < MN) {
      ld4(row, v);
      if (f) {
        T w[4];
        ld4(row + dp, w);
#pragma unroll
        for (int e = 0; e < 4; e++) v[e] = sgn > 0 ? v[e] + w[e] : v[e] - w[e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (i + e < MN) {
          T x = (T)row[e];
          if (f) {
            T y = (T)row[dp + e];
            x = sgn > 0 ? x + y : x - y;
          }
          v[e] = x;
        }
    }
  }
}

template <int ROLE>
struct RoleAKC {
  static constexpr bool value = ROLE == G_DST_R || ROLE == G_DST_R_INV || ROLE == G_ANG_BWD;
};

// Block -> tile order.  Workgroups go to the 8 XCDs round robin by linear
// id, each XCD with its own L2; hardware id hw is mapped to a logical id so
// that each XCD runs one contiguous run of logical ids, and the logical order
// puts the blocks that read the same operand panel next to each other:
//   share_b = 0 (A panel shared: same descriptor group, same M tile)
//           order [z / group][y][z % group][x];
//   share_b = 1 (B panel shared: same descriptor, same N tile)
//           order [z][x][y].
struct GemmGrid {
  int gx, gy, group, share_b;
};

__device__ __forceinline__ void gemm_block(const GemmGrid& g, int& x, int& y, int& z) {
  const int hw = blockIdx.x, q = (int)(gridDim.x >> 3);
  const int l = hw < 8 * q ? (hw & 7) * q + (hw >> 3) : hw;
  if (g.share_b) {
    y = l % g.gy;
    const int t = l / g.gy;
    x = t % g.gx;
    z = t / g.gx;
  } else {
    x = l % g.gx;
    int t = l / g.gx;
    const int zg = t % g.group;
    t /= g.group;
    y = t % g.gy;
    z = (t / g.gy) * g.group + zg;
  }
}

template <int ROLE, typename T, int TM, int TN, typename TA, typename TB, typename TC>
__global__ __launch_bounds__(256) void k_gemm_vec(const TA* __restrict__ A, const TB* __restrict__ B,
                                                  TC* __restrict__ C,
                                                  const GemmDesc* __restrict__ descs,
                                                  GemmStrides st, KinParams kp, GemmGrid gg) {
  typedef MfmaAcc<T> M;
  constexpr bool AKC = RoleAKC<ROLE>::value;
  constexpr int BMt = 32 * TM, BNt = 32 * TN;
  constexpr int LA = AKC ? BK + 4 : BMt + 4;  // LDS row length of A (K-major or M-major)
  constexpr int LB = BNt + 4;
  int bx, by, bz;
  gemm_block(gg, bx, by, bz);
  const GemmDesc d = descs[bz];
  const int m0 = by * BMt, n0 = bx * BNt;
  if (m0 >= d.M || n0 >= d.N) return;
  __shared__ __align__(16) T As[AKC ? BMt * LA : BK * LA];
  __shared__ __align__(16) T Bs[BK * LB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int g = lane >> 4, c16 = lane & 15;
  const bool fb = d.fold_op == 2;
  typename M::V acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++) acc[a][b] = (typename M::V){0, 0, 0, 0};
  const TA* Ab = A + d.a_off;
  const TB* Bb = B + d.b_off;
  T ra[TM / 2][4], rb[TN / 2][4];
  auto load = [&](int k0) {
#pragma unroll
    for (int l = 0; l < TM / 2; l++) {
      const int v = threadIdx.x + 256 * l;
      if (AKC)
        ld_a_kc<T>(d, Ab, m0 + (v >> 2), k0 + 4 * (v & 3), ra[l]);
      else
        ld_mn<T>(Ab, d.lda_k, k0 + v / (8 * TM), m0 + 4 * (v % (8 * TM)), d.K, d.M, false, 0, 1,
                 ra[l]);
    }
#pragma unroll
    for (int l = 0; l < TN / 2; l++) {
      const int v = threadIdx.x + 256 * l;
      ld_mn<T>(Bb, st.ldb_k, k0 + v / (8 * TN), n0 + 4 * (v % (8 * TN)), d.K, d.N, fb, d.fold_n,
               d.fold_sgn, rb[l]);
    }
  };
  load(0);
  for (int k0 = 0; k0 < d.K; k0 += BK) {
#pragma unroll
    for (int l = 0; l < TM / 2; l++) {
      const int v = threadIdx.x + 256 * l;
      if (AKC)
        st4(&As[(v >> 2) * LA + 4 * (v & 3)], ra[l]);
      else
        st4(&As[(v / (8 * TM)) * LA + 4 * (v % (8 * TM))], ra[l]);
    }
#pragma unroll
    for (int l = 0; l < TN / 2; l++) {
      const int v = threadIdx.x + 256 * l;
      st4(&Bs[(v / (8 * TN)) * LB + 4 * (v % (8 * TN))], rb[l]);
    }
    __syncthreads();
    if (k0 + BK < d.K) load(k0 + BK);
    T av[TM][4], bv[TN][4];
#pragma unroll
    for (int a = 0; a < TM; a++) {
      const int row = wm * 16 * TM + 16 * a + c16;
      if (AKC) {
        ld4(&As[row * LA + 4 * g], av[a]);
      } else {
#pragma unroll
        for (int s = 0; s < 4; s++) av[a][s] = As[(4 * g + s) * LA + row];
      }
    }
#pragma unroll
    for (int b = 0; b < TN; b++) {
      const int col = wn * 16 * TN + 16 * b + c16;
#pragma unroll
      for (int s = 0; s < 4; s++) bv[b][s] = Bs[(4 * g + s) * LB + col];
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++) acc[a][b] = M::mma(av[a][s], bv[b][s], acc[a][b]);
    __syncthreads();
  }
  TC* Cb = C + d.c_off;
  if (d.kin_mode == 2 && kp.n2 % BNt == 0) {
    // angular coefficients with the block's columns inside one r row:
    // r uniform per block, a_j(r) once per output row
    const int r0 = n0 / kp.n2, Rb = n0 - r0 * kp.n2;
    const double* ac = kp.acoef + (int64_t)d.kin_ia * kp.nJ * kp.n1 + r0;
    const double* bc = kp.bcoef + ((int64_t)d.kin_ia * 3 + 1) * kp.nJ * kp.n2 + Rb;
#pragma unroll
    for (int a = 0; a < TM; a++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm * 16 * TM + a * 16 + M::row(lane, r);
        if (m >= d.M) continue;
        const double am = ac[(int64_t)m * kp.n1];
        const double* bm = bc + (int64_t)m * kp.n2;
#pragma unroll
        for (int b = 0; b < TN; b++) {
          const int nl = wn * 16 * TN + b * 16 + (lane & 15);
          if (n0 + nl < d.N)
            Cb[(int64_t)m * st.ldc_m + (int64_t)(n0 + nl) * st.ldc_n] =
                (TC)((double)acc[a][b][r] * (am + bm[nl]));
        }
      }
    return;
  }
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int m = m0 + wm * 16 * TM + a * 16 + M::row(lane, r);
        int n = n0 + wn * 16 * TN + b * 16 + (lane & 15);
        if (m < d.M && n < d.N) {
          double v = (double)acc[a][b][r];
          if (d.kin_ia >= 0) v *= kin_factor(d, kp, m, n);
          Cb[(int64_t)m * st.ldc_m + (int64_t)n * st.ldc_n] = (TC)v;
        }
      }
}

template <int ROLE, int TM, int TN, typename TACC, typename TA, typename TB, typename TC>
static void launch_gemm_vec(hipStream_t s, const TA* A, const TB* B, TC* C, const GemmPlan& p,
                            KinParams kp) {
  GemmGrid gg{(p.maxN + 32 * TN - 1) / (32 * TN), (p.maxM + 32 * TM - 1) / (32 * TM), p.group,
              p.share_b};
  const unsigned nblk = (unsigned)gg.gx * (unsigned)gg.gy * (unsigned)p.nb;
  hipLaunchKernelGGL((k_gemm_vec<ROLE, TACC, TM, TN, TA, TB, TC>), dim3(nblk), dim3(256), 0, s, A,
                     B, C, p.d_descs, p.st, kp, gg);
}

template <int ROLE, typename TA, typename TB, typename TC, typename TACC>
static hipError_t run_gemm(hipStream_t s, const TA* A, const TB* B, TC* C, const GemmPlan& p,
                           KinParams kp = KinParams{}) {
  if (p.nb == 0 || p.maxM == 0 || p.maxN == 0) return hipSuccess;  // nothing to write
  if (!p.vec) {
    dim3 grid((p.maxN + 63) / 64, (p.maxM + 63) / 64, p.nb);
    hipLaunchKernelGGL((k_gemm_mfma<ROLE, TACC, 2, 2, TA, TB, TC>), grid, dim3(256), 0, s, A, B, C,
                       p.d_descs, p.st, kp);
    return hipGetLastError();
  }
  launch_gemm_vec<ROLE, 2, 2, TACC>(s, A, B, C, p, kp);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// A0 <-> dense
// ---------------------------------------------------------------------------
__global__ void k_scatter(const double* __restrict__ a0, const uint32_t* __restrict__ pt,
                          double* __restrict__ dense, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dense[pt[i]] = a0[i];
}

__global__ void k_gather(const double* __restrict__ dense, const uint32_t* __restrict__ pt,
                         double* __restrict__ a0, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a0[i] = dense[pt[i]];
}

__global__ void k_zero(uint32_t* p, int64_t nwords) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// kinetic scaling between the forward and inverse r transforms
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_kin_scale(T* __restrict__ x, const int32_t* __restrict__ nn1,
                            const double* __restrict__ p2r1, const double* __restrict__ p2r2,
                            const double* __restrict__ divi, double radcut, int nc, int na, int n1,
                            int n2) {
  // one block row per (c, ia, k1)
  int64_t row = blockIdx.x;
  int k1 = row % n1;
  int ia = (row / n1) % na;
  if (k1 >= nn1[ia]) return;
  T* xr = x + row * n2;
  double p1 = p2r1[(int64_t)ia * n1 + k1];
  double dv = divi[ia];
  for (int k2 = threadIdx.x; k2 < n2; k2 += blockDim.x) {
    double e = p1 + p2r2[k2];
    if (e > radcut) e = radcut;
    xr[k2] = (T)((double)xr[k2] * e * dv);
  }
}

// ---------------------------------------------------------------------------
// angular stencil in the FBR: phi_t = (a_t(r) + b_t0(R)) f_t + b_t-(R) f_s- + b_t+(R) f_s+
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fbr_coef(const T* __restrict__ fbr, T* __restrict__ phi,
                           const double* __restrict__ acoef, const double* __restrict__ bcoef,
                           const int32_t* __restrict__ nbr, int nc, int nJ, int n1, int n2) {
  int64_t row = blockIdx.x;  // (t, jj, r)
  int r = row % n1;
  int jj = (row / n1) % nJ;
  int t = row / ((int64_t)n1 * nJ);
  int64_t plane = (int64_t)n1 * n2;
  double a = acoef[((int64_t)t * nJ + jj) * n1 + r];
  const double* b0 = bcoef + (((int64_t)t * 3 + 1) * nJ + jj) * n2;
  const double* bm = bcoef + (((int64_t)t * 3 + 0) * nJ + jj) * n2;
  const double* bp = bcoef + (((int64_t)t * 3 + 2) * nJ + jj) * n2;
  int sm = nbr[t * 3 + 0], sp = nbr[t * 3 + 2];
  const T* ft = fbr + ((int64_t)t * nJ + jj) * plane + (int64_t)r * n2;
  const T* fm = sm >= 0 ? fbr + ((int64_t)sm * nJ + jj) * plane + (int64_t)r * n2 : nullptr;
  const T* fp = sp >= 0 ? fbr + ((int64_t)sp * nJ + jj) * plane + (int64_t)r * n2 : nullptr;
  T* out = phi + ((int64_t)t * nJ + jj) * plane + (int64_t)r * n2;
  for (int R = threadIdx.x; R < n2; R += blockDim.x) {
    double v = (a + b0[R]) * (double)ft[R];
    if (fm) v += bm[R] * (double)fm[R];
    if (fp) v += bp[R] * (double)fp[R];
    out[R] = (T)v;
  }
}

// The same coefficients with every channel of one (j, 16-row chunk of r) in
// one block: each thread takes an R column, loads the b_j(R) triples of all
// NC channels once for the chunk, and per row reads the NC FBR values once
// (each FBR value is read once instead of three times, and the coefficient
// rows once per 16 rows instead of once per row).  The row's NC values also
// go to the thread's own LDS column, from which the Omega +- 1 neighbours are
// read by index (selects over every channel pair kept NC^2 compare masks in
// scalar registers, which spilled from NC = 3 on); a thread reads back only
// its own column, so no barrier is needed.
constexpr int FBR_RB = 16, FBR_THREADS = 256;
template <typename T, int NC>
__global__ __launch_bounds__(FBR_THREADS) void k_fbr_coef_all(const T* __restrict__ fbr,
                                                      T* __restrict__ phi,
                                                      const double* __restrict__ acoef,
                                                      const double* __restrict__ bcoef,
                                                      const int32_t* __restrict__ nbr, int nJ,
                                                      int n1, int n2) {
  const int nrc = (n1 + FBR_RB - 1) / FBR_RB;
  const int jj = blockIdx.x / nrc, rc = blockIdx.x - jj * nrc;
  const int64_t plane = (int64_t)n1 * n2, cst = (int64_t)nJ * plane;
  const int r0 = rc * FBR_RB, r1 = r0 + FBR_RB < n1 ? r0 + FBR_RB : n1;
  __shared__ double fl[NC][FBR_THREADS];
  for (int R = threadIdx.x; R < n2; R += blockDim.x) {
    double b0[NC], bm[NC], bp[NC];
#pragma unroll
    for (int t = 0; t < NC; t++) {
      b0[t] = bcoef[(((int64_t)t * 3 + 1) * nJ + jj) * n2 + R];
      bm[t] = bcoef[(((int64_t)t * 3 + 0) * nJ + jj) * n2 + R];
      bp[t] = bcoef[(((int64_t)t * 3 + 2) * nJ + jj) * n2 + R];
    }
    const T* fb = fbr + (int64_t)jj * plane + R;
    T* ob = phi + (int64_t)jj * plane + R;
    for (int r = r0; r < r1; r++) {
      double f[NC];
#pragma unroll
      for (int t = 0; t < NC; t++) {
        f[t] = (double)fb[t * cst + (int64_t)r * n2];
        fl[t][threadIdx.x] = f[t];
      }
#pragma unroll
      for (int t = 0; t < NC; t++) {
        double v = (acoef[((int64_t)t * nJ + jj) * n1 + r] + b0[t]) * f[t];
        const int sm = nbr[t * 3 + 0], sp = nbr[t * 3 + 2];  // uniform
        // an absent neighbour reads the channel's own value: bm / bp are zero there
        const double fm = fl[sm >= 0 ? sm : t][threadIdx.x];
        const double fp = fl[sp >= 0 ? sp : t][threadIdx.x];
        v += bm[t] * fm + bp[t] * fp;
        ob[t * cst + (int64_t)r * n2] = (T)v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Fused V psi + compressed-grid projection + damped update
// (SRC/mod_Hphi_01y2.f:838-853, SRC/main_madwave3.f:245-248, 332-339) with the
// per-step projections of the new psi (SRC/mod_flux_01y2.f:598-628, 672-702).
// Block = one (channel, angle) and 64 consecutive R; its 8 waves take the r
// rows r = w, w+8, ...; lane = R, so every row access is one contiguous run.
// Points outside the compressed grid are neither read nor written: every
// state buffer is zero there and stays zero.  mode 0: hout = H psi;
// mode 1: out = H_norm psi (first step, undamped); mode 2: damped update.
// proj: also sum zAR1flux(r) psi_new(r, R) and take psi_new(ir1flux, R) for the
// flux columns, and psi_new(r, ir2abs) / sqrt(ah2) for the reactant slice.
struct UpdArgs {
  const int32_t* items;  // [nitems][4]: group (c, ia, R chunk), first row entry, count, 0
  const int32_t* urow;   // listed rows r
  const uint64_t* uword; // their mask words
  double* upart;         // [nitems][3][64] flux partials (sum zAR1flux psi: re, im; psi at ir1flux)
  int nitems;
  const int32_t* nn1;
  const double* absfr1;
  const double* absfr2;
  const int32_t* ch_elec;
  const int32_t* ch_omega;
  const int32_t* ch_cpl;  // [nc][4] coupled channels (CPL4 kernels)
  const double* psi0;  // autocorrelation reference (null: none)
  double* part;        // per-block (norm, autocorrelation) partials
  int nc, ne, na, n1, n2;
  double emindlt, rdelta2, radcut;
  int mode, pflux;  // pflux: flux-column projections on
  // projections (slice non-null: reactant slice on)
  const int32_t* fmap;
  const int32_t* ir1flux;
  const double* zar;
  double* zr;
  double* zdr;
  double* slice;
  float* out32;  // FP32 copy of the new psi (mixed precision), or null
  // mixed precision, damped update: psi_k enters only through the operator
  // (its diagonal V, radcut and emindlt terms), so it is read from its FP32
  // copy like every other operator input
  const float* psi32;
  int ir2abs;
  double rsah2;
  // product projection: per item and lane (R), sum over the item's rows of
  // T(c, ia, r, R) psi_new (T = transRgjac, operator precision)
  const void* tprod;
  int64_t tp_cstride;
  double* ppart;  // [nitems][64]
};


// MODE 0: hout = H psi; 1: first step (undamped); 2: damped update.  SINGLE:
// one electronic state (V diagonal; any number of Omega channels).  PROJ: flux columns and reactant
// slice.  AUTO: <psi_0|psi_new> partials.  PRODP: product-projection partials.
// All are template parameters so that no load sits behind a runtime branch
// (the compiler waits for every load at the join of such a branch, which
// serialises the batch).
// CPL4 (several states): at most 4 channels share an Omega; their V and psi
// values are loaded with the row's other loads, from the host-built list.
template <typename TT, typename TV, int MODE, bool SINGLE, bool PROJ, bool AUTO, bool PRODP,
          bool CPL4 = false>
__global__ __launch_bounds__(256) void k_update(const double* __restrict__ psi,
                                               double* __restrict__ out_buf,
                                               const TT* __restrict__ tkin,
                                               const TT* __restrict__ ang,
                                               const TV* __restrict__ vd, UpdArgs a) {
  const int64_t plane = (int64_t)a.n1 * a.n2, slab = (int64_t)a.na * plane;
  const int nrc = (a.n2 + 63) / 64;
  const int lane = threadIdx.x & 63;
  const int item = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (item >= a.nitems) return;  // no block-level synchronisation in this kernel
  const int g = a.items[4 * item], off = a.items[4 * item + 1], cnt = a.items[4 * item + 2];
  const int rc = g % nrc, cia = g / nrc;
  const int ia = cia % a.na, c = cia / a.na;
  const int R = rc * 64 + lane;
  const bool inR = R < a.n2;
  const int Ra = inR ? R : a.n2 - 1;  // address of lanes past the grid (never used)
  const int nn = a.nn1[ia];
  const int et = a.ch_elec[c], om = a.ch_omega[c];
  const double f2 = a.absfr2[Ra];
  const int rflux = PROJ ? a.ir1flux[ia] - 1 : -1;
  const bool cvjcol = PROJ && a.slice && R == a.ir2abs - 1;
  double s_norm = 0.0, s_auto = 0.0, zre = 0.0, zim = 0.0, zval = 0.0, sprod = 0.0;
  const TT* tprod = PRODP ? (const TT*)a.tprod + (int64_t)c * a.tp_cstride : nullptr;
  constexpr bool P32 = MODE == 2 && sizeof(TT) == 4;  // psi_k from its FP32 copy
  // the item's rows, four at a time: the loads of the four rows and their
  // per-row tables are issued before any is used, with no branch around them
  constexpr int NK = CPL4 ? 4 : 1;
  // coupled channels: V and psi offsets and the 0 / 1 weight (entries past
  // the count repeat the channel itself with weight 0), once per wave
  int64_t cvo[NK], cpo[NK];
  double cw[NK];
#pragma unroll
  for (int k = 0; k < NK; k++) {
    const int s2r = CPL4 ? a.ch_cpl[4 * c + k] : c;
    const int s2 = s2r >= 0 ? s2r : c;
    cvo[k] = CPL4 ? ((int64_t)et * a.ne + a.ch_elec[s2]) * slab : 0;
    cpo[k] = (int64_t)s2 * slab;
    cw[k] = s2r >= 0 ? 1.0 : 0.0;
  }
  for (int b = 0; b < cnt; b += 4) {
    bool act[4];
    int rr[4];
    int64_t pl[4];
    double p[4], ob[4], tk[4], an[4], vv[4], f1[4], za[4], zb[4], p0[4], tp[4];
    double cv[4][NK], cp[4][NK];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int idx = b + u < cnt ? b + u : cnt - 1;
      rr[u] = a.urow[off + idx];
      const uint64_t w = a.uword[off + idx];
      const bool kept = (w >> lane) & 1ull;
      act[u] = b + u < cnt && inR && kept;
      // lanes without a kept point read the row's first kept point instead
      // of their own: the row's loads then touch only the cache lines that
      // hold kept points (the values they load are not used)
      const int Ru = kept ? Ra : rc * 64 + (w ? (int)__builtin_ctzll(w) : 0);
      pl[u] = (int64_t)ia * plane + (int64_t)rr[u] * a.n2 + Ru;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = rr[u];
      const int64_t ie = (int64_t)c * slab + pl[u];
      p[u] = P32 ? (double)a.psi32[ie] : psi[ie];
      tk[u] = (double)tkin[ie];
      an[u] = (double)ang[ie];
      vv[u] = SINGLE ? (double)vd[pl[u]] : 0.0;
      ob[u] = MODE == 2 ? out_buf[ie] : 0.0;
      f1[u] = MODE == 2 ? a.absfr1[r] : 0.0;
      za[u] = PROJ ? a.zar[2 * r] : 0.0;
      zb[u] = PROJ ? a.zar[2 * r + 1] : 0.0;
      p0[u] = AUTO ? a.psi0[ie] : 0.0;
      tp[u] = PRODP ? (double)tprod[pl[u]] : 0.0;
      if (CPL4) {
#pragma unroll
        for (int k = 0; k < NK; k++) {
          cv[u][k] = (double)vd[cvo[k] + pl[u]];
          cp[u][k] = P32 ? (double)a.psi32[cpo[k] + pl[u]] : psi[cpo[k] + pl[u]];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (!act[u]) continue;
      const int r = rr[u];
      const int64_t i = (int64_t)c * slab + pl[u];
      // T psi = radcut psi - low-mode correction on r < nn1; psi beyond (Q1)
      double h = (r < nn ? a.radcut * p[u] - tk[u] : p[u]) + an[u];
      if (SINGLE) {
        h += p[u] * vv[u];
      } else if (CPL4) {
#pragma unroll
        for (int k = 0; k < NK; k++) h += cp[u][k] * (cv[u][k] * cw[k]);
      } else {
        for (int s2 = 0; s2 < a.nc; s2++) {
          if (a.ch_omega[s2] != om) continue;
          const int es = a.ch_elec[s2];
          const double v = (double)vd[((int64_t)et * a.ne + es) * slab + pl[u]];
          const int64_t is2 = (int64_t)s2 * slab + pl[u];
          const double ps = (s2 == c) ? p[u] : (P32 ? (double)a.psi32[is2] : psi[is2]);
          h += ps * v;
        }
      }
      if (MODE == 0) {
        out_buf[i] = h;
        continue;
      }
      const double hn = (h - a.emindlt * p[u]) * a.rdelta2;
      double o;
      if (MODE == 1) {
        o = hn;
      } else {
        const double f = f1[u] * f2;
        o = (2.0 * hn - f * ob[u]) * f;
      }
      out_buf[i] = o;
      if (a.out32) a.out32[i] = (float)o;
      s_norm += o * o;
      if (AUTO) s_auto += p0[u] * o;
      if (PRODP) sprod += tp[u] * o;
      if (PROJ) {
        zre += za[u] * o;
        zim += zb[u] * o;
        if (r == rflux) zval = o;
        if (cvjcol) a.slice[((int64_t)c * a.na + ia) * a.n1 + r] = o * a.rsah2;
      }
    }
  }
  if (PROJ) {
    double* up = a.upart + (int64_t)item * 3 * 64;
    up[lane] = zre;
    up[64 + lane] = zim;
    up[128 + lane] = zval;
  }
  if (PRODP) a.ppart[(int64_t)item * 64 + lane] = sprod;
  if (a.part) {
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
      s_norm += __shfl_xor(s_norm, o2);
      s_auto += __shfl_xor(s_auto, o2);
    }
    if (lane == 0) {
      a.part[2 * item] = s_norm;
      a.part[2 * item + 1] = s_auto;
    }
  }
}

// flux columns of each (channel, angle, 64-R chunk) from the partials of its
// work items (in item order); chunks with no listed rows write zeros
// ring slot of a per-step projection kernel: `s`, or, inside a captured step
// (one graph serves every slot), the value a device counter holds when the
// kernel runs; the counter advances after the step's last projection kernel
struct RingSlot {
  const int32_t* d;
  int32_t s;
  __device__ int64_t get() const { return d ? (int64_t)*d : (int64_t)s; }
};

static RingSlot ring_slot(const Ctx* c, int slot) {
  return RingSlot{c->graph_slot_dev ? c->d_slot : nullptr, slot};
}

__global__ void k_slot_set(int32_t* d, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *d = v;
}

__global__ void k_slot_next(int32_t* d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *d = *d + 1;
}

hipError_t launch_slot_set(Ctx* c, int v, hipStream_t st) {
  hipLaunchKernelGGL(k_slot_set, dim3(1), dim3(64), 0, st, c->d_slot, v);
  return hipGetLastError();
}

hipError_t launch_slot_next(Ctx* c, hipStream_t st) {
  hipLaunchKernelGGL(k_slot_next, dim3(1), dim3(64), 0, st, c->d_slot);
  return hipGetLastError();
}

__global__ __launch_bounds__(64) void k_update_flux(const double* __restrict__ upart,
                                                    const int32_t* __restrict__ gfirst,
                                                    const int32_t* __restrict__ fmap,
                                                    double* __restrict__ zr_ring,
                                                    double* __restrict__ zdr_ring, int n2,
                                                    int64_t nflux, RingSlot rs) {
  const int64_t slot = rs.get();
  double* __restrict__ zr = zr_ring + slot * nflux;
  double* __restrict__ zdr = zdr_ring + 2 * slot * nflux;
  const int g = blockIdx.x, lane = threadIdx.x;
  const int nrc = (n2 + 63) / 64;
  const int rc = g % nrc, cia = g / nrc;
  const int R = rc * 64 + lane;
  if (R >= n2) return;
  double re = 0.0, im = 0.0, zv = 0.0;
  const int i0 = gfirst[g], i1 = gfirst[g + 1];
#pragma unroll 4
  for (int it = i0; it < i1; it++) {
    const double* up = upart + (int64_t)it * 3 * 64;
    re += up[lane];
    im += up[64 + lane];
    zv += up[128 + lane];
  }
  const int f = fmap[(int64_t)cia * n2 + R];
  if (f >= 0) {
    zdr[2 * (int64_t)f] = re;
    zdr[2 * (int64_t)f + 1] = im;
    zr[f] = zv;
  }
}

template <typename T, int MODE, bool SINGLE, bool PROJ, bool AUTO, bool PRODP = false,
          bool CPL4 = false>
static void launch_update_kernel(Ctx* c, unsigned nb, const double* psi, double* out,
                                 const UpdArgs& a) {
  hipLaunchKernelGGL((k_update<T, T, MODE, SINGLE, PROJ, AUTO, PRODP, CPL4>), dim3(nb), dim3(256), 0,
                     c->stream, psi, out, (const T*)c->d_t2, (const T*)c->d_ang,
                     (const T*)c->d_v, a);
}

// prodp (product projection) runs with the projection variant (PROJ) only
template <typename T, bool SINGLE, bool CPL4>
static void dispatch_update(Ctx* c, unsigned nb, const double* psi, double* out, const UpdArgs& a,
                            bool proj, bool autoc, bool prodp) {
  if (a.mode == 0)
    launch_update_kernel<T, 0, SINGLE, false, false, false, CPL4>(c, nb, psi, out, a);
  else if (a.mode == 1)
    autoc ? launch_update_kernel<T, 1, SINGLE, false, true, false, CPL4>(c, nb, psi, out, a)
          : launch_update_kernel<T, 1, SINGLE, false, false, false, CPL4>(c, nb, psi, out, a);
  else if (prodp)
    autoc ? launch_update_kernel<T, 2, SINGLE, true, true, true, CPL4>(c, nb, psi, out, a)
          : launch_update_kernel<T, 2, SINGLE, true, false, true, CPL4>(c, nb, psi, out, a);
  else if (proj)
    autoc ? launch_update_kernel<T, 2, SINGLE, true, true, false, CPL4>(c, nb, psi, out, a)
          : launch_update_kernel<T, 2, SINGLE, true, false, false, CPL4>(c, nb, psi, out, a);
  else
    autoc ? launch_update_kernel<T, 2, SINGLE, false, true, false, CPL4>(c, nb, psi, out, a)
          : launch_update_kernel<T, 2, SINGLE, false, false, false, CPL4>(c, nb, psi, out, a);
}

__global__ void k_reduce_parts(const double* __restrict__ part, int nb,
                               double* __restrict__ out_ring, RingSlot rs) {
  double* __restrict__ out = out_ring + 2 * (int64_t)rs.get();
  __shared__ double sh[2][256];
  // four independent partial sums per thread, so that four loads are in
  // flight at a time instead of one
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2* p2 = (const d2*)part;
  d2 s0 = {0, 0}, s1 = {0, 0}, s2 = {0, 0}, s3 = {0, 0};
  int i = threadIdx.x;
  for (; i + 3 * (int)blockDim.x < nb; i += 4 * blockDim.x) {
    s0 += p2[i];
    s1 += p2[i + blockDim.x];
    s2 += p2[i + 2 * blockDim.x];
    s3 += p2[i + 3 * blockDim.x];
  }
  for (; i < nb; i += blockDim.x) s0 += p2[i];
  const d2 st = (s0 + s1) + (s2 + s3);
  double a = st.x, b = st.y;
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sh[0][0];
    out[1] = sh[1][0];
  }
}

// ---------------------------------------------------------------------------
// per-step projections into the ring.  The flux line (every (c, ia, R)
// column: zdR = sum_r zAR1flux(r) psi(c,ia,r,R), the full-state derivative of
// totalflux_k, SRC/mod_flux_01y2.f:598-609, and zR = psi at r = ir1flux(ia),
// stored at flux points f = fmap(c,ia,R) >= 0) and the reactant slice at
// R = ir2abs are taken inside k_update.
// ---------------------------------------------------------------------------
// Cvj(c, j, v) = sum_r fd(r, v, j) sum_ia slice(c, ia, r) D(ia, j); block per (c, jj)
// (Cvjflux_k, SRC/mod_flux_01y2.f:672-702; slice = psi at R = ir2abs / sqrt(ah2)).
// 16 waves: the angle sum is split over the waves (lanes over r), then each
// wave takes some of the v sums.
constexpr int CVJ_WAVES = 16;
__global__ __launch_bounds__(64 * CVJ_WAVES) void k_cvj(const double* __restrict__ slice,
                                                        const double* __restrict__ dcvj,
                                                        const double* __restrict__ fdv,
                                                        double* __restrict__ cvj_ring, int nj,
                                                        int nv, int na, int n1, int64_t ncvj,
                                                        RingSlot rs) {
  double* __restrict__ cvj = cvj_ring + (int64_t)rs.get() * ncvj;
  const int c = blockIdx.x / nj, jj = blockIdx.x % nj;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  extern __shared__ double xj[];  // [n1]
  __shared__ double part[CVJ_WAVES][64];
  const double* dj = dcvj + ((int64_t)c * nj + jj) * na;
  const double* sl = slice + (int64_t)c * na * n1;
  for (int r0 = 0; r0 < n1; r0 += 64) {
    const int r = r0 + lane;
    double s = 0, s1 = 0;
    if (r < n1) {
      int ia = wv;
      for (; ia + CVJ_WAVES < na; ia += 2 * CVJ_WAVES) {  // two loads in flight
        s += sl[(int64_t)ia * n1 + r] * dj[ia];
        s1 += sl[(int64_t)(ia + CVJ_WAVES) * n1 + r] * dj[ia + CVJ_WAVES];
      }
      if (ia < na) s += sl[(int64_t)ia * n1 + r] * dj[ia];
      s += s1;
    }
    part[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && r < n1) {
      double t = 0;
      for (int q = 0; q < CVJ_WAVES; q++) t += part[q][lane];
      xj[r] = t;
    }
    __syncthreads();
  }
  for (int v = wv; v < nv; v += CVJ_WAVES) {
    const double* f = fdv + (((int64_t)c * nj + jj) * nv + v) * n1;
    double s = 0;
    for (int r = lane; r < n1; r += 64) s += xj[r] * f[r];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) cvj[((int64_t)c * nj + jj) * nv + v] = s;
  }
}

// ---------------------------------------------------------------------------
// energy accumulation over a ring of steps k0 .. k0+ns-1
//   acc[x][E] += sum_s C_{k0+s}(E) z[s][x],  C_k(E) = 2 hbr/delta e^{-i k acos Es}/sqrt(1-Es^2)
// ---------------------------------------------------------------------------
__global__ void k_wmat(double* __restrict__ w, const double* __restrict__ etot, int netot, int k0,
                       int ns, double hbr, double emindlt, double delta2) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  int s = blockIdx.y;
  if (e >= netot || s >= ns) return;
  int k = k0 + s;
  double d = (k == 0) ? 1.0 : 2.0;
  double zf = d * hbr / delta2;
  double Es = (etot[e] - emindlt) / delta2;
  double expo = -(double)k * acos(Es);
  double deno = sqrt(1.0 - Es * Es);
  w[2 * ((int64_t)s * netot + e)] = zf * cos(expo) / deno;
  w[2 * ((int64_t)s * netot + e) + 1] = zf * sin(expo) / deno;
}

// real samples z[s][x] (zr, cvj) or complex (zdr); skip: per-x flag of the
// reference state whose steps k < kmin are left out (Cvjflux_k).  A thread
// owns one energy and EACC_X consecutive x, so each C_k(E) it loads serves
// EACC_X samples.  The samples of EACC_S steps are the same for the whole
// block: they are staged in LDS (zero where left out or past nx), so the
// inner loop has no branch and reads them as broadcasts.  Each (x, E) still
// sums its steps in order.
constexpr int EACC_X = 16, EACC_S = 32;
template <bool CPLX>
__global__ __launch_bounds__(128) void k_eacc(const double* __restrict__ w,
                                              const double* __restrict__ z,
                                              double* __restrict__ acc, int64_t nx, int netot,
                                              int ns, int k0, const uint8_t* __restrict__ skip,
                                              int kmin, int64_t xbegin) {
  __shared__ double zs[EACC_S][EACC_X][2];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const bool eok = e < netot;
  const int ec = eok ? e : netot - 1;
  const int64_t x0 = xbegin + (int64_t)blockIdx.y * EACC_X;
  double ar[EACC_X], ai[EACC_X];
#pragma unroll
  for (int j = 0; j < EACC_X; j++) ar[j] = ai[j] = 0.0;
  for (int s0 = 0; s0 < ns; s0 += EACC_S) {
    const int nsl = ns - s0 < EACC_S ? ns - s0 : EACC_S;
    __syncthreads();  // the previous pass no longer reads zs
    for (int q = threadIdx.x; q < EACC_S * EACC_X; q += blockDim.x) {
      const int sl = q / EACC_X, j = q - sl * EACC_X;
      const int64_t x = x0 + j;
      const int s = s0 + sl;
      const bool off = sl >= nsl || x >= nx || (skip && skip[x] && k0 + s < kmin);
      double zr = 0.0, zi = 0.0;
      if (!off) {
        if (CPLX) {
          zr = z[2 * ((int64_t)s * nx + x)];
          zi = z[2 * ((int64_t)s * nx + x) + 1];
        } else {
          zr = z[(int64_t)s * nx + x];
        }
      }
      zs[sl][j][0] = zr;
      zs[sl][j][1] = zi;
    }
    __syncthreads();
    for (int sl = 0; sl < nsl; sl++) {
      const int s = s0 + sl;
      const double wr = w[2 * ((int64_t)s * netot + ec)], wi = w[2 * ((int64_t)s * netot + ec) + 1];
#pragma unroll
      for (int j = 0; j < EACC_X; j++) {
        const double zr = zs[sl][j][0];
        if (CPLX) {
          const double zi = zs[sl][j][1];
          ar[j] += wr * zr - wi * zi;
          ai[j] += wr * zi + wi * zr;
        } else {
          ar[j] += wr * zr;
          ai[j] += wi * zr;
        }
      }
    }
  }
  if (!eok) return;
#pragma unroll
  for (int j = 0; j < EACC_X; j++)
    if (x0 + j < nx) {
      acc[2 * ((x0 + j) * netot + e)] += ar[j];
      acc[2 * ((x0 + j) * netot + e) + 1] += ai[j];
    }
}

// acc[x][E] += sum_s C_{k0+s}(E) z[s][x] over the ring, x in chunks of the grid limit
static hipError_t launch_eacc(hipStream_t st, const double* w, const double* z, bool cplx,
                              double* acc, int64_t nx, int netot, int ns, int k0,
                              const uint8_t* skip, int kmin) {
  const int64_t maxy = 65535LL * EACC_X;  // x per launch (grid.y limit)
  for (int64_t x0 = 0; x0 < nx; x0 += maxy) {
    const int64_t n = std::min<int64_t>(maxy, nx - x0);
    dim3 g((unsigned)((netot + 127) / 128), (unsigned)((n + EACC_X - 1) / EACC_X));
    if (cplx)
      hipLaunchKernelGGL(k_eacc<true>, g, dim3(128), 0, st, w, z, acc, nx, netot, ns, k0, skip,
                         kmin, x0);
    else
      hipLaunchKernelGGL(k_eacc<false>, g, dim3(128), 0, st, w, z, acc, nx, netot, ns, k0, skip,
                         kmin, x0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// psi(E): zfe[e][i] += C_k(E_e) psi[i] on the dense slab (mask implied by psi)
__global__ void k_psie(double* __restrict__ zfe, const double* __restrict__ psi,
                       const double* __restrict__ efe, int nofe, int64_t total, int k, double hbr,
                       double emindlt, double delta2) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  double p = psi[i];
  double d = (k == 0) ? 1.0 : 2.0;
  for (int e = 0; e < nofe; e++) {
    double Es = (efe[e] - emindlt) / delta2;
    double expo = -(double)k * acos(Es);
    double zf = d * hbr / delta2 / sqrt(1.0 - Es * Es);
    zfe[2 * ((int64_t)e * total + i)] += zf * cos(expo) * p;
    zfe[2 * ((int64_t)e * total + i) + 1] += zf * sin(expo) * p;
  }
}

// ---------------------------------------------------------------------------
// spectrum from the autocorrelation (cheby-spectra, SRC/photo/cheby-spectra.f:
// 71-90): spec(E) = [a_0 + sum_k a_k cos(-k acos Es)] / (sqrt(1 - Es^2) delta2)
// with a_0 = auto_0 hbr, a_k = 2 hbr e^{-k xkgam} auto_k; NaN -> 0.  One
// thread per energy; the a_k pass through LDS in chunks shared by the block.
// ---------------------------------------------------------------------------
__global__ void k_spec_coef(const double* __restrict__ autoc, double* __restrict__ a, int kmax,
                            double hbr, double xkgam) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= kmax; k += gridDim.x * blockDim.x)
    a[k] = k == 0 ? autoc[0] * hbr : 2.0 * hbr * exp(-(double)k * xkgam) * autoc[k];
}

constexpr int SPEC_CHUNK = 2048;

__global__ __launch_bounds__(256) void k_spectrum(const double* __restrict__ a, int kmax,
                                                  const double* __restrict__ e, int ne,
                                                  double emindlt, double delta2,
                                                  double* __restrict__ spec) {
  __shared__ double sa[SPEC_CHUNK];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double Es = i < ne ? (e[i] - emindlt) / delta2 : 0.0;
  const double th = acos(Es);
  double sp = a[0];
  for (int k0 = 1; k0 <= kmax; k0 += SPEC_CHUNK) {
    const int nk = min(SPEC_CHUNK, kmax - k0 + 1);
    __syncthreads();
    for (int q = threadIdx.x; q < nk; q += blockDim.x) sa[q] = a[k0 + q];
    __syncthreads();
    for (int q = 0; q < nk; q++) sp += sa[q] * cos(-(double)(k0 + q) * th);
  }
  if (i < ne) {
    sp = sp / (sqrt(1.0 - Es * Es) * delta2);
    spec[i] = isnan(sp) ? 0.0 : sp;
  }
}

hipError_t launch_spectrum(hipStream_t s, const double* d_auto, double* d_a, int kmax,
                           double hbr, double xkgam, const double* d_e, int ne, double emindlt,
                           double delta2, double* d_spec) {
  hipLaunchKernelGGL(k_spec_coef, dim3((unsigned)std::min(1024, (kmax + 256) / 256)), dim3(256),
                     0, s, d_auto, d_a, kmax, hbr, xkgam);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_spectrum, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, d_a, kmax,
                     d_e, ne, emindlt, delta2, d_spec);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_zero(Ctx* c, void* p, size_t bytes) {
  if (!bytes) return hipSuccess;
  int64_t words = (int64_t)(bytes / 4);
  unsigned g = (unsigned)std::min<int64_t>((words + 255) / 256, 65536);
  hipLaunchKernelGGL(k_zero, dim3(g), dim3(256), 0, c->stream, (uint32_t*)p, words);
  return hipGetLastError();
}

hipError_t launch_scatter_a0(Ctx* c, const double* a0, double* dense) {
  hipError_t e = launch_zero(c, dense, sizeof(double) * (size_t)c->t.nc * c->t.nslab);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scatter, dim3(nblk(c->t.ntot, 256)), dim3(256), 0, c->stream, a0,
                     c->d_pt_dense, dense, c->t.ntot);
  return hipGetLastError();
}

hipError_t launch_gather_a0(Ctx* c, const double* dense, double* a0) {
  hipLaunchKernelGGL(k_gather, dim3(nblk(c->t.ntot, 256)), dim3(256), 0, c->stream, dense,
                     c->d_pt_dense, a0, c->t.ntot);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// GEMM plans (shapes are fixed per context)
// ---------------------------------------------------------------------------
static bool plan_vec_ok(const std::vector<GemmDesc>& ds, const GemmStrides& st, bool akc);

static hipError_t upload_plan(GemmPlan* p, const std::vector<GemmDesc>& d, bool akc) {
  p->vec = plan_vec_ok(d, p->st, akc) ? 1 : 0;
  p->nb = (int)d.size();
  p->maxM = p->maxN = 0;
  for (auto& x : d) {
    p->maxM = std::max(p->maxM, x.M);
    p->maxN = std::max(p->maxN, x.N);
  }
  hipError_t e = hipMalloc(&p->d_descs, sizeof(GemmDesc) * d.size());
  if (e != hipSuccess) return e;
  return hipMemcpy(p->d_descs, d.data(), sizeof(GemmDesc) * d.size(), hipMemcpyHostToDevice);
}

// whether every descriptor of a plan meets the vector kernel's alignment
// rules (see k_gemm_vec); akc: A is K-contiguous
static bool plan_vec_ok(const std::vector<GemmDesc>& ds, const GemmStrides& st, bool akc) {
  if (st.ldb_n != 1 || st.ldb_k % 4) return false;
  for (const GemmDesc& d : ds) {
    if (d.a_off % 4 || d.b_off % 4) return false;
    if (akc) {
      if (d.lda_k != 1 || d.lda_m % 4) return false;
      if (d.fold_op == 1 ? (d.fold_n % 4 != 0 || 2 * d.K > d.fold_n)
                         : (d.lda_m < ((d.K + 3) & ~3)))
        return false;
    } else {
      if (d.lda_m != 1 || d.lda_k % 4 || d.fold_op == 1) return false;
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Angular term in one pass, for contexts with no Coriolis neighbour (the FBR
// coefficients are diagonal in j):
//   F[j][p]       = sum_ia D_c(ia, j) psi[c][ia][p]
//   F[j][p]      *= a_j(r1) + b_j(R)                     (p = r1 n2 + R)
//   ang[c][ia][p] = sum_j D_c(ia, j) F[j][p]
// the forward product, the coefficients and the back product of the two-GEMM
// form, with F kept out of HBM.  Persistent blocks of ANG_WAVES waves walk
// tiles of 16 ANG_WAVES columns p; each wave owns 16 columns and every row of
// both products (NT row tiles of 16, zero padded).  D_c stays in LDS while a
// block works on channel c.  F never leaves the registers: the MFMA result
// layout (column = lane % 16, row MfmaAcc::row(lane, r) in register r) is the
// B-operand layout of four k steps in which lane group g takes
// k = row(lane, r) at step r, so the forward accumulators are the back
// product's B operands as they stand.  The forward product runs over ia in
// chunks of 16, lane group g taking ia = 16 c + 4 g + s at step s; the LDS
// row pitch (4 mod 64 words for FP32, 2 mod 32 doubles for FP64) spreads the
// four groups' rows over distinct banks.
// ---------------------------------------------------------------------------
constexpr int ANG_WAVES = 8;

// an opaque copy of a lane value: address arithmetic built on it stays where
// it is written instead of being hoisted to the top of the tile, where its
// results would hold registers through both products
__device__ __forceinline__ int lane_opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <typename T>
struct AngPitch;
template <>
struct AngPitch<float> {
  static constexpr int extra = 4;
};
template <>
struct AngPitch<double> {
  static constexpr int extra = 2;
};

template <typename T, int NT>
__global__ __launch_bounds__(64 * ANG_WAVES) void k_ang_fused(
    const T* __restrict__ psi, const T* __restrict__ Dmat, const double* __restrict__ acoef,
    const double* __restrict__ bcoef, T* __restrict__ ang, const int32_t* __restrict__ tiles,
    int ntiles, int na, int nJ, int nJp, int n1, int n2) {
  typedef MfmaAcc<T> M;
  typedef typename M::V V;
  constexpr int NR = 16 * NT;                 // padded rows of D (ia) and of F (j)
  constexpr int P = NR + AngPitch<T>::extra;  // LDS row pitch
  constexpr int XA = sizeof(T) == 8 ? 2 : 3;  // psi chunks loaded ahead of the MFMAs
  __shared__ __align__(16) T Ds[NR * P];
  const int pl = n1 * n2;  // the host checks that a channel's slab is below 2 GiB
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, i = lane & 15;
  const int tpc = (pl + 16 * ANG_WAVES - 1) / (16 * ANG_WAVES);  // tiles per channel
  // loads take a scalar channel base and a 32-bit per-lane byte offset
  // x[4 c + s] = psi[ch][16 c + 4 g + s][p]: the B operand of forward step
  // 4 c + s; rows past na read row na - 1 (finite), which meets zero rows of D
  T x[4 * NT];
  auto ldx = [&](int c, int tl) {
    const int ch = tl / tpc;
    const char* pb = (const char*)(psi + (int64_t)ch * na * pl);
    int pp = (tl - ch * tpc) * (16 * ANG_WAVES) + 16 * wave + lane_opaque(i);
    pp = pp < pl ? pp : pl - 1;
    const int gg = lane_opaque(g);
#pragma unroll
    for (int s = 0; s < 4; s++) {
      int ia = 16 * c + 4 * gg + s;
      ia = ia < na ? ia : na - 1;
      x[4 * c + s] = *(const T*)(pb + (uint32_t)(ia * pl + pp) * (uint32_t)sizeof(T));
    }
  };
  // the block walks the listed tiles (channel * tiles per channel + tile);
  // tiles with no point of the compressed grid are not listed: psi is zero
  // there and the update reads no angular value there
  int kt = blockIdx.x;
  if (kt >= ntiles) return;
  // entries: tile index | (waves whose 16 columns hold a grid point) << 24
  int tile = tiles[kt] & 0xffffff;
#pragma unroll
  for (int c = 0; c < XA && c < NT; c++) ldx(c, tile);
  int cur = -1;
  for (; kt < ntiles; kt += gridDim.x) {
    const int tent = tiles[kt];
    tile = tent & 0xffffff;
    const int ch = tile / tpc;
    if (ch != cur) {  // block-uniform
      __syncthreads();
      // D_c rows in 4-element vectors (nJp is a multiple of 4), every load
      // issued before the first LDS store
      const T* Dc = Dmat + (int64_t)ch * na * nJp;
      constexpr int NV = (NR * NR / 4 + 64 * ANG_WAVES - 1) / (64 * ANG_WAVES);
      constexpr int NG = NV < 4 ? NV : 4;  // vectors in flight per thread
#pragma unroll
      for (int l0 = 0; l0 < NV; l0 += NG) {
        T dv[NG][4];
#pragma unroll
        for (int l = 0; l < NG; l++) {
          const int v = threadIdx.x + 64 * ANG_WAVES * (l0 + l);
          const int ia = v / (NR / 4), j0 = 4 * (v - ia * (NR / 4));
#pragma unroll
          for (int e = 0; e < 4; e++) dv[l][e] = 0;
          if (l0 + l < NV && ia < na && j0 < nJp) {
            ld4(Dc + (int64_t)ia * nJp + j0, dv[l]);
#pragma unroll
            for (int e = 0; e < 4; e++)
              if (j0 + e >= nJ) dv[l][e] = 0;
          }
        }
#pragma unroll
        for (int l = 0; l < NG; l++) {
          const int v = threadIdx.x + 64 * ANG_WAVES * (l0 + l);
          const int ia = v / (NR / 4), j0 = 4 * (v - ia * (NR / 4));
          if (l0 + l < NV && ia < NR) st4(&Ds[ia * P + j0], dv[l]);
        }
      }
      __syncthreads();
      cur = ch;
    }
    // the next tile of this block (the last one repeats itself), whose first
    // psi chunks load during this tile's back product
    const int tnext =
        kt + (int)gridDim.x < ntiles ? tiles[kt + (int)gridDim.x] & 0xffffff : tile;
    if (!((tent >> (24 + wave)) & 1)) {  // no grid point in this wave's columns
#pragma unroll
      for (int c = 0; c < XA && c < NT; c++) ldx(c, tnext);
      continue;
    }
    const int p = (tile - ch * tpc) * (16 * ANG_WAVES) + 16 * wave + i;
    const int pc = p < pl ? p : pl - 1;
    const int r1 = pc / n2, R = pc - r1 * n2;
    const char* ac = (const char*)(acoef + (int64_t)ch * nJ * n1);
    const char* bc = (const char*)(bcoef + ((int64_t)ch * 3 + 1) * nJ * n2);
    // coefficient pairs of row tile t (rows j >= nJ are zero in F)
    double ca[2][4], cb[2][4];
    auto ldcoef = [&](int t, double (&a)[4], double (&b)[4]) {
      const int ll = lane_opaque(lane), rr = lane_opaque(r1), RR = lane_opaque(R);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int j = 16 * t + M::row(ll, r);
        j = j < nJ ? j : nJ - 1;
        a[r] = *(const double*)(ac + (uint32_t)(j * n1 + rr) * 8u);
        b[r] = *(const double*)(bc + (uint32_t)(j * n2 + RR) * 8u);
      }
    };
    // forward product; each stage reads the next stage's D values from LDS
    // before its own MFMAs (the scheduling barriers keep that order and stop
    // the compiler from hoisting every read of the unrolled loop)
    V f[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) f[t] = (V){0, 0, 0, 0};
    T a1[2][NT];
    auto rd1 = [&](int k, T (&a)[NT]) {
      const T* drow = &Ds[(16 * (k >> 2) + 4 * g + (k & 3)) * P + i];
#pragma unroll
      for (int t = 0; t < NT; t++) a[t] = drow[16 * t];
    };
    rd1(0, a1[0]);
#pragma unroll
    for (int k = 0; k < 4 * NT; k++) {
      if (k + 1 < 4 * NT) rd1(k + 1, a1[(k + 1) & 1]);
      if ((k & 3) == 0 && (k >> 2) + XA < NT) ldx((k >> 2) + XA, tile);
      if (k == (NT > 1 ? 4 * NT - 8 : 0)) ldcoef(0, ca[0], cb[0]);
      if (NT > 1 && k == 4 * NT - 4) ldcoef(1, ca[1], cb[1]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < NT; t++) f[t] = M::mma(a1[k & 1][t], x[k], f[t]);
      __builtin_amdgcn_sched_barrier(0);
    }
    // back product in NH passes over the output tiles (NT / NH accumulators
    // live at a time); row tile t of F, scaled by its coefficients during
    // the first pass, feeds the pass's output tiles.  Stage (h, t, uu) reads
    // a[r] = D(16 u + i, 16 t + row(lane, r)), u = h UN + uu, two stages
    // ahead; halfway through row tile t of the first pass the VALU scales row
    // tile t + 1 (its coefficients loaded a row tile earlier) while the MFMAs
    // run
    constexpr int NH = NT % 2 == 0 ? 2 : 1;
    constexpr int UN = NT / NH;
    constexpr int NQ = NH * NT * UN;  // stages
    auto scale = [&](int t) {
#pragma unroll
      for (int r = 0; r < 4; r++)
        f[t][r] = (T)((double)f[t][r] * (ca[t & 1][r] + cb[t & 1][r]));
    };
    T a2[3][4], a4[4][4];
    auto rd2 = [&](int q, T (&a)[4]) {
      const int h = q / (NT * UN), t = (q - h * NT * UN) / UN;
      const int u = h * UN + (q - h * NT * UN - t * UN);
      const T* drow = &Ds[(16 * u + i) * P + 16 * t];
      if (sizeof(T) == 4) {
        ld4(drow + 4 * g, a);
      } else {
#pragma unroll
        for (int r = 0; r < 4; r++) a[r] = drow[M::row(lane, r)];
      }
    };
    if constexpr (UN % 2 == 0) {
      rd2(0, a4[0]);
      rd2(1, a4[1]);
    } else {
      rd2(0, a2[0]);
      if (NQ > 1) rd2(1, a2[1]);
    }
    scale(0);
    constexpr int UH = UN / 2 > 0 ? UN / 2 - 1 : 0;
    char* ob = (char*)(ang + (int64_t)ch * na * pl);
#pragma unroll
    for (int h = 0; h < NH; h++) {
      V o[UN];
#pragma unroll
      for (int uu = 0; uu < UN; uu++) o[uu] = (V){0, 0, 0, 0};
      if constexpr (UN % 2 == 0) {
        // stages (h, t, uu) and (h, t, uu + 1) together, their MFMAs
        // alternating between the two output tiles: no MFMA waits on the one
        // just before (stage q read into slot q % 4, two stages ahead)
#pragma unroll
        for (int t = 0; t < NT; t++) {
#pragma unroll
          for (int uu = 0; uu < UN; uu += 2) {
            const int q = (h * NT + t) * UN + uu;
            if (q + 2 < NQ) rd2(q + 2, a4[(q + 2) % 4]);
            if (q + 3 < NQ) rd2(q + 3, a4[(q + 3) % 4]);
            if (h == 0 && uu == 0 && t + 2 < NT) ldcoef(t + 2, ca[t & 1], cb[t & 1]);
            if (uu == 0 && t == 0 && h == NH - 1) {
#pragma unroll
              for (int c = 0; c < XA && c < NT; c++) ldx(c, tnext);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; r++) {
              o[uu] = M::mma(a4[q % 4][r], f[t][r], o[uu]);
              o[uu + 1] = M::mma(a4[(q + 1) % 4][r], f[t][r], o[uu + 1]);
            }
            __builtin_amdgcn_sched_barrier(0);
            if (h == 0 && uu <= UH && UH < uu + 2 && t + 1 < NT) scale(t + 1);
          }
        }
      } else {
#pragma unroll
        for (int t = 0; t < NT; t++) {
#pragma unroll
          for (int uu = 0; uu < UN; uu++) {
            const int q = (h * NT + t) * UN + uu;
            if (q + 2 < NQ) rd2(q + 2, a2[(q + 2) % 3]);
            if (h == 0 && uu == 0 && t + 2 < NT) ldcoef(t + 2, ca[t & 1], cb[t & 1]);
            if (uu == 0 && t == 0 && h == NH - 1) {
#pragma unroll
              for (int c = 0; c < XA && c < NT; c++) ldx(c, tnext);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; r++) o[uu] = M::mma(a2[q % 3][r], f[t][r], o[uu]);
            __builtin_amdgcn_sched_barrier(0);
            if (h == 0 && uu == UH && t + 1 < NT) scale(t + 1);
          }
        }
      }
      const int ll = lane_opaque(lane), pp = lane_opaque(p);
#pragma unroll
      for (int uu = 0; uu < UN; uu++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int ia = 16 * (h * UN + uu) + M::row(ll, r);
          if (ia < na && pp < pl)
            *(T*)(ob + (uint32_t)(ia * pl + pp) * (uint32_t)sizeof(T)) = o[uu][r];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// The one-pass angular term folded on the parity of the angular functions.
// For an even number of angles whose functions satisfy D(na-1-ia, j) =
// s_j D(ia, j) with s_j = +-1 (the host checks it to 1e-12 of each
// function's largest value), with h = na / 2, u = psi[ia] + psi[na-1-ia]
// and v = psi[ia] - psi[na-1-ia] (ia < h):
//   F[j] = sum_{ia<h} D(ia, j) u[ia] (s_j = +1) or D(ia, j) v[ia] (s_j = -1)
//   G[j] = (a_j(r1) + b_j(R)) F[j]
//   E[ia] = sum_{s_j=+1} D(ia, j) G[j],   O[ia] = sum_{s_j=-1} D(ia, j) G[j]
//   ang[ia] = E + O,   ang[na-1-ia] = E - O
// which halves both products.  The j axis runs class by class: columns
// 0 .. NR-1 of the LDS image of D hold the s_j = +1 functions and NR ..
// 2 NR-1 the s_j = -1 ones (zero padded), and the coefficient tables follow
// the same order, so row tile T of F (T < 2 NT) is class T / NT.
//
// Work: the 16-column groups of the plane that hold a compressed-grid point
// (the same list for every channel), one group per wave at a time.  The
// grid is nbc blocks per channel; block b of a channel takes D once and its
// waves 2 m and 2 m + 1 walk the list entries 2 (b + nbc m) + {0, 1}, plus
// multiples of 8 nbc: neighbouring groups (64-byte halves of one 128-byte
// line in FP32) are read by one block at one time, the pairs of a last
// partial round go one per block before any block takes a second, and the
// waves of one SIMD differ by at most one group.
// ---------------------------------------------------------------------------
// FP32: at most 128 VGPRs (4 waves per SIMD), so that two blocks share a CU
template <typename T, int NT>
__global__ __launch_bounds__(64 * ANG_WAVES, sizeof(T) == 4 ? 4 : 1) void k_ang_fold(
    const T* __restrict__ psi, const T* __restrict__ Dfold, const T* __restrict__ afold,
    const T* __restrict__ bfold, T* __restrict__ ang, const int32_t* __restrict__ groups,
    int ngroups, int nbc, int na, int n1, int n2) {
  typedef MfmaAcc<T> M;
  typedef typename M::V V;
  constexpr int NR = 16 * NT, NC2 = 2 * NR;    // rows (ia < h) and columns (both classes)
  constexpr int P = NC2 + AngPitch<T>::extra;  // LDS row pitch
  constexpr int XA = 2;                        // psi chunks loaded ahead of the MFMAs
  __shared__ __align__(16) T Ds[NR * P];
  const int pl = n1 * n2, h = na >> 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, i = lane & 15;
  const int ch = blockIdx.x / nbc, bl = blockIdx.x - ch * nbc;
  const int stride = nbc * ANG_WAVES;
  const char* pb = (const char*)(psi + (int64_t)ch * na * pl);
  // xl[4 c + s] = psi[16 c + 4 g + s][p], xh[..] = psi[na-1-(16 c + 4 g + s)][p];
  // angles past h read angle h-1 and its mirror (finite), against zero rows of D
  T xl[4 * NT], xh[4 * NT];
  auto ldx = [&](int c, int col0) {  // col0: first column of the wave's group
    int pp = col0 + lane_opaque(i);
    pp = pp < pl ? pp : pl - 1;
    const int gg = lane_opaque(g);
#pragma unroll
    for (int s = 0; s < 4; s++) {
      int ia = 16 * c + 4 * gg + s;
      ia = ia < h ? ia : h - 1;
      xl[4 * c + s] = *(const T*)(pb + (uint32_t)(ia * pl + pp) * (uint32_t)sizeof(T));
      xh[4 * c + s] = *(const T*)(pb + (uint32_t)((na - 1 - ia) * pl + pp) * (uint32_t)sizeof(T));
    }
  };
  // groups with no point of the compressed grid are not listed: psi is zero
  // there and the update reads no angular value there.  The first group's
  // psi loads go out before the D image's, so that the two latencies overlap
  int gi = 2 * (bl + nbc * (wave >> 1)) + (wave & 1);
  const bool busy = gi < ngroups;
  if (busy) {
#pragma unroll
    for (int c = 0; c < XA && c < NT; c++) ldx(c, 16 * groups[gi]);
  }
  {
      const T* Dc = Dfold + (int64_t)ch * NR * NC2;
      constexpr int NV = (NR * NC2 / 4 + 64 * ANG_WAVES - 1) / (64 * ANG_WAVES);
      constexpr int NG = NV < 4 ? NV : 4;  // vectors in flight per thread
#pragma unroll
      for (int l0 = 0; l0 < NV; l0 += NG) {
        T dv[NG][4];
#pragma unroll
        for (int l = 0; l < NG; l++) {
          int v = threadIdx.x + 64 * ANG_WAVES * (l0 + l);
          v = v < NR * NC2 / 4 ? v : NR * NC2 / 4 - 1;
          ld4(Dc + 4 * (int64_t)v, dv[l]);
        }
#pragma unroll
        for (int l = 0; l < NG; l++) {
          int v = threadIdx.x + 64 * ANG_WAVES * (l0 + l);
          v = v < NR * NC2 / 4 ? v : NR * NC2 / 4 - 1;
          const int row = v / (NC2 / 4), c4 = v - row * (NC2 / 4);
          st4(&Ds[row * P + 4 * c4], dv[l]);
        }
      }
      __syncthreads();
  }
  if (!busy) return;  // no block-wide barrier follows
  const char* ac = (const char*)(afold + (int64_t)ch * NC2 * n1);
  const char* bc = (const char*)(bfold + (int64_t)ch * NC2 * n2);
  char* ob = (char*)(ang + (int64_t)ch * na * pl);
  for (; gi < ngroups; gi += stride) {
    const int col = 16 * groups[gi];
    const int cnext = 16 * groups[gi + stride < ngroups ? gi + stride : gi];
    const int p = col + i;
    const int pc = p < pl ? p : pl - 1;
    const int r1 = pc / n2, R = pc - r1 * n2;
    T ca[2][4], cb[2][4];
    auto ldcoef = [&](int t, T (&a)[4], T (&b)[4]) {
      const int ll = lane_opaque(lane), rr = lane_opaque(r1), RR = lane_opaque(R);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int j = 16 * t + M::row(ll, r);
        a[r] = *(const T*)(ac + (uint32_t)(j * n1 + rr) * (uint32_t)sizeof(T));
        b[r] = *(const T*)(bc + (uint32_t)(j * n2 + RR) * (uint32_t)sizeof(T));
      }
    };
    // forward products: f[t] (t < NT) the s = +1 rows from u, f[NT + t] the
    // s = -1 rows from v
    V f[2 * NT];
#pragma unroll
    for (int t = 0; t < 2 * NT; t++) f[t] = (V){0, 0, 0, 0};
    T a1[2][2 * NT];
    auto rd1 = [&](int k, T (&a)[2 * NT]) {
      const T* drow = &Ds[(16 * (k >> 2) + 4 * g + (k & 3)) * P + i];
#pragma unroll
      for (int t = 0; t < 2 * NT; t++) a[t] = drow[16 * t];
    };
    rd1(0, a1[0]);
#pragma unroll
    for (int k = 0; k < 4 * NT; k++) {
      if (k + 1 < 4 * NT) rd1(k + 1, a1[(k + 1) & 1]);
      if ((k & 3) == 0 && (k >> 2) + XA < NT) ldx((k >> 2) + XA, col);
      if (k == (NT > 1 ? 4 * NT - 8 : 0)) ldcoef(0, ca[0], cb[0]);
      if (k == 4 * NT - 4) ldcoef(1, ca[1], cb[1]);
      __builtin_amdgcn_sched_barrier(0);
      const T u = xl[k] + xh[k], v = xl[k] - xh[k];
#pragma unroll
      for (int t = 0; t < NT; t++) f[t] = M::mma(a1[k & 1][t], u, f[t]);
#pragma unroll
      for (int t = NT; t < 2 * NT; t++) f[t] = M::mma(a1[k & 1][t], v, f[t]);
      __builtin_amdgcn_sched_barrier(0);
    }
    // back products: row tile t of F (class t / NT), scaled by its
    // coefficients, feeds the class's NT output tiles o[(t / NT) NT + u];
    // stage (t, u) reads a[r] = D(16 u + i, 16 t + row(lane, r)) two stages
    // ahead, and halfway through row tile t the VALU scales row tile t + 1
    constexpr int NQ = 2 * NT * NT;
    auto scale = [&](int t) {
#pragma unroll
      for (int r = 0; r < 4; r++)
        f[t][r] = f[t][r] * (ca[t & 1][r] + cb[t & 1][r]);
    };
    V o[2 * NT];
#pragma unroll
    for (int u = 0; u < 2 * NT; u++) o[u] = (V){0, 0, 0, 0};
    T a2[3][4];
    auto rd2 = [&](int q, T (&a)[4]) {
      const int t = q / NT, u = q - t * NT;
      const T* drow = &Ds[(16 * u + i) * P + 16 * t];
      if (sizeof(T) == 4) {
        ld4(drow + 4 * g, a);
      } else {
#pragma unroll
        for (int r = 0; r < 4; r++) a[r] = drow[M::row(lane, r)];
      }
    };
    constexpr int UH = NT / 2 > 0 ? NT / 2 - 1 : 0;
    if constexpr (NT % 2 == 0) {
      // stages (t, u) and (t, u + 1) together, their MFMAs alternating
      // between the two output tiles: no MFMA waits on the one just before
      T a4[4][4];  // stage q in slot q % 4, read two stages ahead
      rd2(0, a4[0]);
      rd2(1, a4[1]);
      scale(0);
#pragma unroll
      for (int t = 0; t < 2 * NT; t++) {
#pragma unroll
        for (int u = 0; u < NT; u += 2) {
          const int q = t * NT + u;
          if (q + 2 < NQ) rd2(q + 2, a4[(q + 2) % 4]);
          if (q + 3 < NQ) rd2(q + 3, a4[(q + 3) % 4]);
          if (u == 0 && t + 2 < 2 * NT) ldcoef(t + 2, ca[t & 1], cb[t & 1]);
          if (u == 0 && t == NT) {
#pragma unroll
            for (int c = 0; c < XA && c < NT; c++) ldx(c, cnext);
          }
          __builtin_amdgcn_sched_barrier(0);
          const int ou = (t / NT) * NT + u;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            o[ou] = M::mma(a4[q % 4][r], f[t][r], o[ou]);
            o[ou + 1] = M::mma(a4[(q + 1) % 4][r], f[t][r], o[ou + 1]);
          }
          __builtin_amdgcn_sched_barrier(0);
          if (u <= UH && UH < u + 2 && t + 1 < 2 * NT) scale(t + 1);
        }
      }
    } else {
      rd2(0, a2[0]);
      if (NQ > 1) rd2(1, a2[1]);
      scale(0);
#pragma unroll
      for (int t = 0; t < 2 * NT; t++) {
#pragma unroll
        for (int u = 0; u < NT; u++) {
          const int q = t * NT + u;
          if (q + 2 < NQ) rd2(q + 2, a2[(q + 2) % 3]);
          if (u == 0 && t + 2 < 2 * NT) ldcoef(t + 2, ca[t & 1], cb[t & 1]);
          if (u == 0 && t == NT) {
#pragma unroll
            for (int c = 0; c < XA && c < NT; c++) ldx(c, cnext);
          }
          __builtin_amdgcn_sched_barrier(0);
          const int ou = (t / NT) * NT + u;
#pragma unroll
          for (int r = 0; r < 4; r++) o[ou] = M::mma(a2[q % 3][r], f[t][r], o[ou]);
          __builtin_amdgcn_sched_barrier(0);
          if (u == UH && t + 1 < 2 * NT) scale(t + 1);
        }
      }
    }
    const int ll = lane_opaque(lane), pp = lane_opaque(p);
#pragma unroll
    for (int u = 0; u < NT; u++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ia = 16 * u + M::row(ll, r);
        if (ia < h && pp < pl) {
          const T e = o[u][r], od = o[NT + u][r];
          *(T*)(ob + (uint32_t)(ia * pl + pp) * (uint32_t)sizeof(T)) = e + od;
          *(T*)(ob + (uint32_t)((na - 1 - ia) * pl + pp) * (uint32_t)sizeof(T)) = e - od;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// k_ang_fold for FP32 data with both products on the FP16 matrix cores,
// compensated as in k_rad_inv_h: operands split into hi = fp16(x) and
// lo = fp16(x - hi), sums of hi hi + hi lo + lo hi in FP32, each product's
// right operand scaled per column by a power of two so that its largest
// entry lies in [2^14, 2^15) and the result scaled back exactly.  The
// depth-32 products take 32 angles (forward) or 32 functions of one class
// (back); for odd NT both depths are padded to 32 ceil(NT / 2) with zero
// rows and columns of the images, and the back products run an even number
// of output tiles per class.
//   forward: F[j][p] = sum_ia D(ia, j) w[ia][p], w = u (class +1) or v;
//     lane (g, i) holds column i and angles 32 m + 8 g .. + 7, and reads the
//     table operand from a j-major image Dj[j][ia] (16-byte reads);
//   back: out[ia][p] = sum_j D(ia, j) G[j][p] per class; lane (g, i) holds
//     rows 16 t + 4 g + r of F, so its 8 depth slots of chunk m are its own
//     rows of tiles 2 m and 2 m + 1 (no exchange), and the angle-major image
//     Di stores each class's columns permuted to that order (16-byte reads).
// Both images are built from the FP32 table once per block.  Per group the
// next group's psi values and coefficients are loaded while this one runs.
// PART 0 is the whole pass.  With Coriolis-coupled channels the two products
// run as two launches: PART 1 stores F, unscaled, into the FBR slab at the
// rows of its functions (jmap: function at each class-order position, -1
// for padding); PART 3 forms G = (a + b) F + b_m F(Omega-1) + b_p F(Omega+1)
// from that slab (bmf, bpf: the neighbour coefficients in class order; nbr:
// the neighbour channels) and runs the back product.
// ---------------------------------------------------------------------------
template <int NT, int PART>
__global__ __launch_bounds__(64 * ANG_WAVES, 2) void k_ang_fold_h(
    const float* __restrict__ psi, const float* __restrict__ Dfold,
    const float* __restrict__ afold, const float* __restrict__ bfold, float* __restrict__ ang,
    const int32_t* __restrict__ groups, int ngroups, int nbc, int na, int n1, int n2,
    float* __restrict__ fbr, const int32_t* __restrict__ jmap, int nJ,
    const float* __restrict__ bmf, const float* __restrict__ bpf,
    const int32_t* __restrict__ nbr) {
  typedef _Float16 H8 __attribute__((ext_vector_type(8)));
  typedef _Float16 H4 __attribute__((ext_vector_type(4)));
  typedef _Float16 H2 __attribute__((ext_vector_type(2)));
  typedef float V __attribute__((ext_vector_type(4)));
  constexpr int NR = 16 * NT, NC2 = 2 * NR, MK = (NT + 1) / 2;
  constexpr int KP = 32 * MK, NTP = 2 * MK;     // padded depth and output tiles per class
  constexpr int PF = KP + 8, PB = 2 * KP + 8;  // image pitches (halves)
  constexpr int SF = NC2 * PF, SB = KP * PB;
  constexpr bool BACK = PART == 3;  // back product only
  __shared__ __align__(16) _Float16 Dj[BACK ? 8 : 2 * SF];  // hi, lo
  __shared__ __align__(16) _Float16 Di[PART == 1 ? 8 : 2 * SB];
  const int pl = n1 * n2, h = na >> 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, i = lane & 15;
  const int ch = blockIdx.x / nbc, bl = blockIdx.x - ch * nbc;
  const int stride = nbc * ANG_WAVES;
  const char* pb = (const char*)(psi + (int64_t)ch * na * pl);
  const char* ac = (const char*)(afold + (int64_t)ch * NC2 * n1);
  const char* bc = (const char*)(bfold + (int64_t)ch * NC2 * n2);
  char* ob = (char*)(ang + (int64_t)ch * na * pl);
  // a group's inputs: psi at angles 32 m + 8 g + e and their mirrors (angles
  // past h read angle h-1, against zero rows of D), the coefficients of rows
  // j = 16 t + 4 g + r
  // (parts 1 and 2: the function j at position 16 t + 4 g + r of this
  // channel's class order, -1 for padding; G read from the coefficient
  // pass's slab at those functions)
  float xl[MK][8], xh[MK][8], ca[2 * NT][4], cb[2 * NT][4], gv[2 * NT][4];
  float fm[2 * NT][4], fp[2 * NT][4], cm[2 * NT][4], cq[2 * NT][4];
  int jm[2 * NT][4];
  // part 3: G = (a + b) F + b_m F(Omega - 1) + b_p F(Omega + 1) formed here
  // from the FBR slab (function order, so a neighbour's row is the same j);
  // the term of an absent neighbour is dropped (its row reads the channel's
  // own, times zero)
  int sm = ch, sp = ch;
  float wm = 0.0f, wp = 0.0f;
  if constexpr (PART == 3) {
    sm = nbr[ch * 3 + 0] >= 0 ? nbr[ch * 3 + 0] : ch;
    sp = nbr[ch * 3 + 2] >= 0 ? nbr[ch * 3 + 2] : ch;
    wm = nbr[ch * 3 + 0] >= 0 ? 1.0f : 0.0f;
    wp = nbr[ch * 3 + 2] >= 0 ? 1.0f : 0.0f;
  }
  const float* fbo = fbr + (int64_t)ch * nJ * pl;
  const float* fbm = fbr + (int64_t)sm * nJ * pl;
  const float* fbp = fbr + (int64_t)sp * nJ * pl;
  const char* mc = (const char*)(bmf + (int64_t)ch * NC2 * n2);
  const char* qc = (const char*)(bpf + (int64_t)ch * NC2 * n2);
  // part 3 keeps the map in LDS (registers) and loads each group's inputs at
  // its start
  __shared__ int jms[PART == 3 ? NC2 : 1];
  if constexpr (PART == 3) {
    for (int v = threadIdx.x; v < NC2; v += 64 * ANG_WAVES) jms[v] = jmap[ch * NC2 + v];
  }
  if constexpr (PART == 1) {
#pragma unroll
    for (int t = 0; t < 2 * NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) jm[t][r] = jmap[ch * NC2 + 16 * t + 4 * g + r];
  }
  auto ldall = [&](int col0) {
    int pp = col0 + lane_opaque(i);
    pp = pp < pl ? pp : pl - 1;
    const int gg = lane_opaque(g), ll = lane_opaque(lane);
    if constexpr (PART == 3) {  // F only: the coefficients load separately
#pragma unroll
      for (int t = 0; t < 2 * NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int jj = jms[16 * t + 4 * gg + r];
          const int64_t o = (int64_t)(jj >= 0 ? jj : 0) * pl + pp;
          gv[t][r] = fbo[o];
          fm[t][r] = fbm[o];
          fp[t][r] = fbp[o];
        }
      return;
    }
#pragma unroll
    for (int m = 0; m < MK; m++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        int ia = 32 * m + 8 * gg + e;
        ia = ia < h ? ia : h - 1;
        xl[m][e] = *(const float*)(pb + (uint32_t)(ia * pl + pp) * 4u);
        xh[m][e] = *(const float*)(pb + (uint32_t)((na - 1 - ia) * pl + pp) * 4u);
      }
    if constexpr (PART == 1) return;
    const int r1 = pp / n2, R = pp - r1 * n2;
#pragma unroll
    for (int t = 0; t < 2 * NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int j = 16 * t + 4 * (ll >> 4) + r;
        ca[t][r] = *(const float*)(ac + (uint32_t)(j * n1 + r1) * 4u);
        cb[t][r] = *(const float*)(bc + (uint32_t)(j * n2 + R) * 4u);
      }
  };
  // part 3: the group's four coefficient rows
  auto ldcoef3 = [&](int col0) {
    int pp = col0 + lane_opaque(i);
    pp = pp < pl ? pp : pl - 1;
    const int ll = lane_opaque(lane), r1 = pp / n2, R = pp - r1 * n2;
#pragma unroll
    for (int t = 0; t < 2 * NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int j = 16 * t + 4 * (ll >> 4) + r;
        ca[t][r] = *(const float*)(ac + (uint32_t)(j * n1 + r1) * 4u);
        cb[t][r] = *(const float*)(bc + (uint32_t)(j * n2 + R) * 4u);
        cm[t][r] = *(const float*)(mc + (uint32_t)(j * n2 + R) * 4u);
        cq[t][r] = *(const float*)(qc + (uint32_t)(j * n2 + R) * 4u);
      }
  };
  int gi = 2 * (bl + nbc * (wave >> 1)) + (wave & 1);
  const bool busy = gi < ngroups;
  if constexpr (PART != 3) {
    if (busy) ldall(16 * groups[gi]);  // in flight during the image copy
  }
  // part 3 prefetches the next group's F up to three tiles (measured: 2%
  // faster at three, 2% slower at four, where it holds 246 VGPRs)
  constexpr bool PF3 = NT <= 3;
  bool first3 = true;
  if constexpr (KP != NR) {  // padded depth: zero images first
    if constexpr (!BACK)
      for (int v = threadIdx.x; v < SF / 4; v += 64 * ANG_WAVES)
        *(H8*)&Dj[8 * v] = (H8){0, 0, 0, 0, 0, 0, 0, 0};
    if constexpr (PART != 1)
      for (int v = threadIdx.x; v < SB / 4; v += 64 * ANG_WAVES)
        *(H8*)&Di[8 * v] = (H8){0, 0, 0, 0, 0, 0, 0, 0};
    __syncthreads();
  }
  {
    const float* Dc = Dfold + (int64_t)ch * NR * NC2;
    constexpr int NV4 = NR * NC2 / 4;
    constexpr int NV = (NV4 + 64 * ANG_WAVES - 1) / (64 * ANG_WAVES);
    constexpr int NG = NV < 4 ? NV : 4;
#pragma unroll
    for (int l0 = 0; l0 < (PART == 1 ? 0 : NV); l0 += NG) {
      float dv[NG][4];
#pragma unroll
      for (int l = 0; l < NG; l++) {
        int v = threadIdx.x + 64 * ANG_WAVES * (l0 + l);
        v = v < NV4 ? v : NV4 - 1;
        ld4(Dc + 4 * (int64_t)v, dv[l]);
      }
#pragma unroll
      for (int l = 0; l < NG; l++) {
        int v = threadIdx.x + 64 * ANG_WAVES * (l0 + l);
        v = v < NV4 ? v : NV4 - 1;
        const int ia = v / (NC2 / 4), j0 = 4 * (v - ia * (NC2 / 4));
        H4 hv, lv;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          hv[r] = (_Float16)dv[l][r];
          lv[r] = (_Float16)(dv[l][r] - (float)hv[r]);
        }
        // class column jl = 16 (2 m + hh) + 4 q + r goes to 32 m + 8 q + 4 hh + r
        const int cls = j0 / NR, jl = j0 - cls * NR;
        const int cp = cls * KP + 32 * (jl >> 5) + 8 * ((jl >> 2) & 3) + 4 * ((jl >> 4) & 1);
        *(H4*)&Di[ia * PB + cp] = hv;
        *(H4*)&Di[SB + ia * PB + cp] = lv;
      }
    }
    // the j-major image in a second pass: a thread takes two consecutive
    // angles of four functions, so that the lanes of a wave store
    // consecutive 4-byte words of one image row
    constexpr int NP = BACK ? 0 : NR / 2 * (NC2 / 4);
    constexpr int NVP = (NP + 64 * ANG_WAVES - 1) / (64 * ANG_WAVES);
    float dp[NVP > 0 ? NVP : 1][2][4];
#pragma unroll
    for (int l = 0; l < NVP; l++) {
      int v = threadIdx.x + 64 * ANG_WAVES * l;
      v = v < NP ? v : NP - 1;
      const int ia = 2 * (v % (NR / 2)), j0 = 4 * (v / (NR / 2));
      ld4(Dc + (int64_t)ia * NC2 + j0, dp[l][0]);
      ld4(Dc + (int64_t)(ia + 1) * NC2 + j0, dp[l][1]);
    }
#pragma unroll
    for (int l = 0; l < NVP; l++) {
      int v = threadIdx.x + 64 * ANG_WAVES * l;
      v = v < NP ? v : NP - 1;
      const int ia = 2 * (v % (NR / 2)), j0 = 4 * (v / (NR / 2));
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const _Float16 h0 = (_Float16)dp[l][0][r], h1 = (_Float16)dp[l][1][r];
        *(H2*)&Dj[(j0 + r) * PF + ia] = (H2){h0, h1};
        *(H2*)&Dj[SF + (j0 + r) * PF + ia] =
            (H2){(_Float16)(dp[l][0][r] - (float)h0), (_Float16)(dp[l][1][r] - (float)h1)};
      }
    }
    __syncthreads();
  }
  if (!busy) return;  // no block-wide barrier follows
  auto maxcol = [](float x) {  // over the four lanes of one column
    x = fmaxf(x, __shfl_xor(x, 16));
    return fmaxf(x, __shfl_xor(x, 32));
  };
  auto expo = [](float mx) {  // mx 2^e in [2^14, 2^15)
    int ex = 0;
    if (mx > 0.0f) {
      frexpf(mx, &ex);
      ex = 15 - ex;
    }
    return ex;
  };
  for (; gi < ngroups; gi += stride) {
    const int col = 16 * groups[gi];
    const int cnext = 16 * groups[gi + stride < ngroups ? gi + stride : gi];
    const int p = col + i;
    V f[2 * NT];
    if constexpr (PART == 3) {
      if (first3 || !PF3) ldall(col);
      first3 = false;
      ldcoef3(col);
#pragma unroll
      for (int t = 0; t < 2 * NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const float gq = (ca[t][r] + cb[t][r]) * gv[t][r] + wm * cm[t][r] * fm[t][r] +
                           wp * cq[t][r] * fp[t][r];
          f[t][r] = jms[16 * t + 4 * g + r] >= 0 ? gq : 0.0f;
        }
      if constexpr (PF3) ldall(cnext);  // the next group's F, during this one
    } else {
    // fold, scale per column and class, split
    float mu = 0.0f, mv = 0.0f;
#pragma unroll
    for (int m = 0; m < MK; m++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float u = xl[m][e] + xh[m][e], v = xl[m][e] - xh[m][e];
        xl[m][e] = u;
        xh[m][e] = v;
        mu = fmaxf(mu, fabsf(u));
        mv = fmaxf(mv, fabsf(v));
      }
    const int eu = expo(maxcol(mu)), ev = expo(maxcol(mv));
    H8 bw[2][MK][2];  // [class][m][hi, lo]
#pragma unroll
    for (int m = 0; m < MK; m++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float su = ldexpf(xl[m][e], eu), sv = ldexpf(xh[m][e], ev);
        const _Float16 hu = (_Float16)su, hv = (_Float16)sv;
        bw[0][m][0][e] = hu;
        bw[0][m][1][e] = (_Float16)(su - (float)hu);
        bw[1][m][0][e] = hv;
        bw[1][m][1][e] = (_Float16)(sv - (float)hv);
      }
    float cs[2 * NT][4];
    if constexpr (PART == 0) {
#pragma unroll
      for (int t = 0; t < 2 * NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) cs[t][r] = ca[t][r] + cb[t][r];
    }
    ldall(cnext);  // the next group's inputs, in flight during this one
    // forward: tile pairs (t, t + 1) of one class, MFMAs alternating between them
#pragma unroll
    for (int t = 0; t < 2 * NT; t++) f[t] = (V){0, 0, 0, 0};
    constexpr int NQF = NT * MK;  // steps: tile pair tp, chunk m
    H8 af[2][2][2];               // [buffer][tile of the pair][hi, lo]
    auto rdf = [&](int q, H8 (&a)[2][2]) {
      const int tp = q / MK, m = q - tp * MK;
#pragma unroll
      for (int d = 0; d < 2; d++) {
        const _Float16* s = &Dj[(16 * (2 * tp + d) + i) * PF + 32 * m + 8 * g];
        a[d][0] = *(const H8*)s;
        a[d][1] = *(const H8*)(s + SF);
      }
    };
    rdf(0, af[0]);
#pragma unroll
    for (int q = 0; q < NQF; q++) {
      if (q + 1 < NQF) rdf(q + 1, af[(q + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
      // the two tiles of a pair may be of two classes (odd NT)
      const int tp = q / MK, m = q - tp * MK, t0 = 2 * tp, c0 = t0 / NT, c1 = (t0 + 1) / NT;
      const auto& A = af[q & 1];
      f[t0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][0], bw[c0][m][0], f[t0], 0, 0, 0);
      f[t0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][0], bw[c1][m][0], f[t0 + 1], 0, 0, 0);
      f[t0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][0], bw[c0][m][1], f[t0], 0, 0, 0);
      f[t0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][0], bw[c1][m][1], f[t0 + 1], 0, 0, 0);
      f[t0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][1], bw[c0][m][0], f[t0], 0, 0, 0);
      f[t0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][1], bw[c1][m][0], f[t0 + 1], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (PART == 1) {
      // F, unscaled, to the FBR slab rows of its functions
      const int pp = lane_opaque(p);
      float* fbc = fbr + (int64_t)ch * nJ * pl;
#pragma unroll
      for (int t = 0; t < 2 * NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (jm[t][r] >= 0 && pp < pl)
            fbc[(int64_t)jm[t][r] * pl + pp] = ldexpf(f[t][r], t < NT ? -eu : -ev);
      continue;
    }
    // G = (a + b) F, unscaled
#pragma unroll
    for (int t = 0; t < 2 * NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) f[t][r] = ldexpf(f[t][r] * cs[t][r], t < NT ? -eu : -ev);
    }
    // G scaled per column and class and split
    float mg[2] = {0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 2 * NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) mg[t / NT] = fmaxf(mg[t / NT], fabsf(f[t][r]));
    const int eg[2] = {expo(maxcol(mg[0])), expo(maxcol(mg[1]))};
    H8 bg[2][MK][2];  // [class][m][hi, lo]: slot 4 hh + r = row r of tile 2 m + hh
#pragma unroll
    for (int c2 = 0; c2 < 2; c2++)
#pragma unroll
      for (int m = 0; m < MK; m++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int tt = 2 * m + (e >> 2);  // past NT: the zero padding of the depth
          const float s = tt < NT ? ldexpf(f[c2 * NT + (tt < NT ? tt : 0)][e & 3], eg[c2]) : 0.0f;
          const _Float16 hs = (_Float16)s;
          bg[c2][m][0][e] = hs;
          bg[c2][m][1][e] = (_Float16)(s - (float)hs);
        }
    // back: output tile pairs (u, u + 1) of one class
    V o[2 * NTP];
#pragma unroll
    for (int u = 0; u < 2 * NTP; u++) o[u] = (V){0, 0, 0, 0};
    constexpr int NQB = 2 * MK * MK;  // steps: class, tile pair, chunk
    H8 ab[2][2][2];
    auto rdb = [&](int q, H8 (&a)[2][2]) {
      const int c2 = q / (NQB / 2), rem = q - c2 * (NQB / 2), up = rem / MK, m = rem - up * MK;
#pragma unroll
      for (int d = 0; d < 2; d++) {
        const _Float16* s = &Di[(16 * (2 * up + d) + i) * PB + c2 * KP + 32 * m + 8 * g];
        a[d][0] = *(const H8*)s;
        a[d][1] = *(const H8*)(s + SB);
      }
    };
    rdb(0, ab[0]);
#pragma unroll
    for (int q = 0; q < NQB; q++) {
      if (q + 1 < NQB) rdb(q + 1, ab[(q + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
      const int c2 = q / (NQB / 2), rem = q - c2 * (NQB / 2), up = rem / MK, m = rem - up * MK;
      const int u0 = c2 * NTP + 2 * up;
      const auto& A = ab[q & 1];
      o[u0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][0], bg[c2][m][0], o[u0], 0, 0, 0);
      o[u0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][0], bg[c2][m][0], o[u0 + 1], 0, 0, 0);
      o[u0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][0], bg[c2][m][1], o[u0], 0, 0, 0);
      o[u0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][0], bg[c2][m][1], o[u0 + 1], 0, 0, 0);
      o[u0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][1], bg[c2][m][0], o[u0], 0, 0, 0);
      o[u0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][1], bg[c2][m][0], o[u0 + 1], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
    }
    const int ll = lane_opaque(lane), pp = lane_opaque(p);
#pragma unroll
    for (int u = 0; u < NT; u++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ia = 16 * u + 4 * (ll >> 4) + r;
        if (ia < h && pp < pl) {
          const float e = ldexpf(o[u][r], -eg[0]), od = ldexpf(o[NTP + u][r], -eg[1]);
          *(float*)(ob + (uint32_t)(ia * pl + pp) * 4u) = e + od;
          *(float*)(ob + (uint32_t)((na - 1 - ia) * pl + pp) * 4u) = e - od;
        }
      }
  }
}

// blocks per channel: the resident blocks shared evenly by the channels, and
// no more than the group list can feed
template <typename K>
static hipError_t ang_fold_grid_k(K kernel, int nc, int ngroups, int* nbc) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 64 * ANG_WAVES, 0);
  if (e != hipSuccess) return e;
  int dev = 0, ncu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  const int per = (nb > 0 ? nb : 1) * ncu / (nc > 0 ? nc : 1);
  const int need = (ngroups + ANG_WAVES - 1) / ANG_WAVES;
  *nbc = std::max(1, std::min(per, need));
  return hipSuccess;
}

template <typename T, int NT>
static hipError_t ang_fold_grid(Ctx* c, int nc, int ngroups, int* nbc) {
  if constexpr (sizeof(T) == 4) {
    if (c->ang_fold_h) {
      hipError_t e = ang_fold_grid_k(k_ang_fold_h<NT, 0>, nc, ngroups, &c->ang_fold_h_nbc);
      if (e != hipSuccess) return e;
    }
  }
  return ang_fold_grid_k(k_ang_fold<T, NT>, nc, ngroups, nbc);
}

template <typename T, int NT>
static void launch_ang_fold_nt(Ctx* c, const T* psi, hipStream_t st) {
  const HostTables& t = c->t;
  if (c->ang_ngroups == 0) return;
  // fewer blocks per channel (the grid option) walk more groups each and
  // leave CUs to the radial passes beside them
  const int gdiv = c->opt_ang_grid > 0 ? c->opt_ang_grid : 1;
  if constexpr (sizeof(T) == 4) {
    if (c->ang_fold_h && c->opt_fp16) {  // products on the FP16 matrix cores, compensated
      const int nbh = std::max(1, c->ang_fold_h_nbc / gdiv);
      hipLaunchKernelGGL((k_ang_fold_h<NT, 0>), dim3((unsigned)(nbh * t.nc)), dim3(64 * ANG_WAVES),
                         0, st, psi, (const float*)c->d_fold_D, (const float*)c->d_fold_a,
                         (const float*)c->d_fold_b, (float*)c->d_ang, c->d_ang_groups,
                         c->ang_ngroups, nbh, t.na, t.n1, t.n2, nullptr, nullptr, 0, nullptr,
                         nullptr, nullptr);
      return;
    }
  }
  const int nbc = std::max(1, c->ang_fold_nbc / gdiv);
  hipLaunchKernelGGL((k_ang_fold<T, NT>), dim3((unsigned)(nbc * t.nc)), dim3(64 * ANG_WAVES), 0,
                     st, psi, (const T*)c->d_fold_D, (const T*)c->d_fold_a,
                     (const T*)c->d_fold_b, (T*)c->d_ang, c->d_ang_groups, c->ang_ngroups, nbc,
                     t.na, t.n1, t.n2);
}

template <typename T>
static hipError_t launch_ang_fold(Ctx* c, const T* psi, hipStream_t st) {
  switch (c->ang_fold) {
    case 1: launch_ang_fold_nt<T, 1>(c, psi, st); break;
    case 2: launch_ang_fold_nt<T, 2>(c, psi, st); break;
    case 3: launch_ang_fold_nt<T, 3>(c, psi, st); break;
    case 4: launch_ang_fold_nt<T, 4>(c, psi, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t ang_fold_setup(Ctx* c, int nt) {
  const int nc = c->t.nc, ng = c->ang_ngroups;
  c->ang_fold_h = sizeof(T) == 4;
  switch (nt) {
    case 1: return ang_fold_grid<T, 1>(c, nc, ng, &c->ang_fold_nbc);
    case 2: return ang_fold_grid<T, 2>(c, nc, ng, &c->ang_fold_nbc);
    case 3: return ang_fold_grid<T, 3>(c, nc, ng, &c->ang_fold_nbc);
    case 4: return ang_fold_grid<T, 4>(c, nc, ng, &c->ang_fold_nbc);
    default: return hipErrorInvalidValue;
  }
}

// the folded parts beside the Coriolis pass (FP32 data): their grids
template <int NT>
static hipError_t ang_cor_grid(Ctx* c) {
  hipError_t e = ang_fold_grid_k(k_ang_fold_h<NT, 1>, c->t.nc, c->ang_ngroups, &c->ang_cor_nbc[0]);
  if (e != hipSuccess) return e;
  return ang_fold_grid_k(k_ang_fold_h<NT, 3>, c->t.nc, c->ang_ngroups, &c->ang_cor_nbc[2]);
}

static hipError_t ang_cor_setup(Ctx* c, int nt) {
  switch (nt) {
    case 1: return ang_cor_grid<1>(c);
    case 2: return ang_cor_grid<2>(c);
    case 3: return ang_cor_grid<3>(c);
    case 4: return ang_cor_grid<4>(c);
    default: return hipErrorInvalidValue;
  }
}

// part 1: psi -> F into the FBR slab (function order); part 3: the FBR slab
// -> the angular term, with the coefficients and the Omega +- 1 terms formed
// in the kernel
template <int NT, int PART>
static void launch_ang_cor_nt(Ctx* c, const float* psi, hipStream_t st) {
  const HostTables& t = c->t;
  const int gdiv = c->opt_ang_grid > 0 ? c->opt_ang_grid : 1;
  const int nbc = std::max(1, c->ang_cor_nbc[PART - 1] / gdiv);
  hipLaunchKernelGGL((k_ang_fold_h<NT, PART>), dim3((unsigned)(nbc * t.nc)), dim3(64 * ANG_WAVES), 0,
                     st, psi, (const float*)c->d_fold_D, (const float*)c->d_fold_a,
                     (const float*)c->d_fold_b, (float*)c->d_ang, c->d_ang_groups, c->ang_ngroups,
                     nbc, t.na, t.n1, t.n2, (float*)c->d_fbr, c->d_fold_jmap, t.nJ, (const float*)c->d_fold_bm, (const float*)c->d_fold_bp,
                     c->d_nbr);
}

template <int PART>
static hipError_t launch_ang_cor(Ctx* c, const float* psi, hipStream_t st) {
  switch (c->ang_fold_cor) {
    case 1: launch_ang_cor_nt<1, PART>(c, psi, st); break;
    case 2: launch_ang_cor_nt<2, PART>(c, psi, st); break;
    case 3: launch_ang_cor_nt<3, PART>(c, psi, st); break;
    case 4: launch_ang_cor_nt<4, PART>(c, psi, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// the folded angular tables when every channel's functions have the parity
// symmetry; returns the kernel's tile count (0: not folded)
static hipError_t build_ang_fold(Ctx* c, int* nt_out) {
  const HostTables& t = c->t;
  *nt_out = 0;
  if (t.na % 2 != 0) return hipSuccess;
  const int na = t.na, h = na / 2, nJ = t.nJ, nc = t.nc;
  std::vector<int> cls((size_t)nc * nJ, 1);
  int nemax = 0, nomax = 0;
  for (int ch = 0; ch < nc; ch++) {
    int ne = 0, no = 0;
    for (int j = 0; j < nJ; j++) {
      auto D = [&](int ia) { return t.Dmat[((size_t)ch * na + ia) * t.nJp + j]; };
      double dmax = 0, dot = 0;
      for (int ia = 0; ia < na; ia++) dmax = std::fmax(dmax, std::fabs(D(ia)));
      for (int ia = 0; ia < h; ia++) dot += D(na - 1 - ia) * D(ia);
      const int sj = dot >= 0 ? 1 : -1;
      for (int ia = 0; ia < h; ia++)
        if (std::fabs(D(na - 1 - ia) - sj * D(ia)) > 1e-12 * dmax) return hipSuccess;
      cls[(size_t)ch * nJ + j] = sj;
      (sj > 0 ? ne : no)++;
    }
    nemax = std::max(nemax, ne);
    nomax = std::max(nomax, no);
  }
  const int nt = (std::max(h, std::max(nemax, nomax)) + 15) / 16;
  if (nt > 4) return hipSuccess;
  const int NR = 16 * nt, NC2 = 2 * NR;
  std::vector<double> Df((size_t)nc * NR * NC2, 0.0), af((size_t)nc * NC2 * t.n1, 0.0),
      bf((size_t)nc * NC2 * t.n2, 0.0);
  std::vector<int32_t> jmap((size_t)nc * NC2, -1);
  std::vector<double> bmf((size_t)nc * NC2 * t.n2, 0.0), bpf((size_t)nc * NC2 * t.n2, 0.0);
  for (int ch = 0; ch < nc; ch++) {
    int slot[2] = {0, NR};
    for (int j = 0; j < nJ; j++) {
      const int col = slot[cls[(size_t)ch * nJ + j] > 0 ? 0 : 1]++;
      jmap[(size_t)ch * NC2 + col] = j;
      for (int ia = 0; ia < h; ia++)
        Df[((size_t)ch * NR + ia) * NC2 + col] = t.Dmat[((size_t)ch * na + ia) * t.nJp + j];
      for (int r = 0; r < t.n1; r++)
        af[((size_t)ch * NC2 + col) * t.n1 + r] = t.acoef[((size_t)ch * nJ + j) * t.n1 + r];
      for (int R = 0; R < t.n2; R++) {
        bf[((size_t)ch * NC2 + col) * t.n2 + R] =
            t.bcoef[(((size_t)ch * 3 + 1) * nJ + j) * t.n2 + R];
        bmf[((size_t)ch * NC2 + col) * t.n2 + R] = t.bcoef[(((size_t)ch * 3 + 0) * nJ + j) * t.n2 + R];
        bpf[((size_t)ch * NC2 + col) * t.n2 + R] = t.bcoef[(((size_t)ch * 3 + 2) * nJ + j) * t.n2 + R];
      }
    }
  }
  hipError_t e;
  if (c->precision == MW3_MIXED) {
    std::vector<float> d32(Df.begin(), Df.end());
    if ((e = hipMalloc(&c->d_fold_D, d32.size() * sizeof(float))) != hipSuccess) return e;
    if ((e = hipMemcpy(c->d_fold_D, d32.data(), d32.size() * sizeof(float), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return e;
  } else {
    if ((e = hipMalloc(&c->d_fold_D, Df.size() * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMemcpy(c->d_fold_D, Df.data(), Df.size() * sizeof(double), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return e;
  }
  // the coefficient tables in the operator precision
  auto put_table = [&](void** dst, const std::vector<double>& v) -> hipError_t {
    if (c->precision == MW3_MIXED) {
      std::vector<float> v32(v.begin(), v.end());
      hipError_t err = hipMalloc(dst, v32.size() * sizeof(float));
      if (err != hipSuccess) return err;
      return hipMemcpy(*dst, v32.data(), v32.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    hipError_t err = hipMalloc(dst, v.size() * sizeof(double));
    if (err != hipSuccess) return err;
    return hipMemcpy(*dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice);
  };
  if ((e = put_table(&c->d_fold_a, af)) != hipSuccess) return e;
  if ((e = put_table(&c->d_fold_b, bf)) != hipSuccess) return e;
  if ((e = put_table(&c->d_fold_bm, bmf)) != hipSuccess) return e;
  if ((e = put_table(&c->d_fold_bp, bpf)) != hipSuccess) return e;
  if ((e = hipMalloc(&c->d_fold_jmap, jmap.size() * sizeof(int32_t))) != hipSuccess) return e;
  if ((e = hipMemcpy(c->d_fold_jmap, jmap.data(), jmap.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice)) != hipSuccess)
    return e;
  *nt_out = nt;
  return hipSuccess;
}

template <typename T, int NT>
static hipError_t ang_fused_grid(int* blocks) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ang_fused<T, NT>,
                                                              64 * ANG_WAVES, 0);
  if (e != hipSuccess) return e;
  int dev = 0, ncu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  *blocks = (nb > 0 ? nb : 1) * ncu;
  return hipSuccess;
}

template <typename T, int NT>
static void launch_ang_fused_nt(Ctx* c, const T* psi, hipStream_t st) {
  const HostTables& t = c->t;
  const int nt = c->ang_ntiles;
  const unsigned nblk = (unsigned)(nt < c->ang_grid ? nt : c->ang_grid);
  if (nblk == 0) return;
  hipLaunchKernelGGL((k_ang_fused<T, NT>), dim3(nblk), dim3(64 * ANG_WAVES), 0, st, psi,
                     (const T*)c->d_Dmat, (const double*)c->d_acoef, (const double*)c->d_bcoef,
                     (T*)c->d_ang, c->d_ang_tiles, nt, t.na, t.nJ, t.nJp, t.n1, t.n2);
}

template <typename T>
static hipError_t launch_ang_fused(Ctx* c, const T* psi, hipStream_t st) {
  switch (c->ang_nt) {
    case 1: launch_ang_fused_nt<T, 1>(c, psi, st); break;
    case 2: launch_ang_fused_nt<T, 2>(c, psi, st); break;
    case 3: launch_ang_fused_nt<T, 3>(c, psi, st); break;
    case 4: launch_ang_fused_nt<T, 4>(c, psi, st); break;
    case 5: launch_ang_fused_nt<T, 5>(c, psi, st); break;
    case 6: launch_ang_fused_nt<T, 6>(c, psi, st); break;
    case 7: launch_ang_fused_nt<T, 7>(c, psi, st); break;
    case 8: launch_ang_fused_nt<T, 8>(c, psi, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t ang_fused_setup(Ctx* c, int nt) {
  switch (nt) {
    case 1: return ang_fused_grid<T, 1>(&c->ang_grid);
    case 2: return ang_fused_grid<T, 2>(&c->ang_grid);
    case 3: return ang_fused_grid<T, 3>(&c->ang_grid);
    case 4: return ang_fused_grid<T, 4>(&c->ang_grid);
    case 5: return ang_fused_grid<T, 5>(&c->ang_grid);
    case 6: return ang_fused_grid<T, 6>(&c->ang_grid);
    case 7: return ang_fused_grid<T, 7>(&c->ang_grid);
    case 8: return ang_fused_grid<T, 8>(&c->ang_grid);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// First and last passes of the radial correction over R, for n2 = 32 HB with
// HB in {8, 10, 12, 16}, on work items of 16 rows r < nn1(ia) of one
// (channel, angle): int4 {channel * na + angle, row tile, nn1, k2lo}.
// With h = n2 / 2 and S(n2-1-R, k) = (-1)^k S(R, k):
//   k_rad_fwd: t1[r][k2] = sum_R S(R, k2) psi[r][R] for k2 < k2lo(ia), as
//     even k2: sum_{j<h} (psi[j] + psi[n2-1-j]) S(j, k2)
//     odd  k2: sum_{j<h} (psi[j] - psi[n2-1-j]) S(j, k2);
//   k_rad_inv: corr[r][R] = E + O and corr[r][n2-1-R] = E - O for R < h,
//     E / O the sums of t1[r][k2] S(R, k2) over even / odd k2 < k2lo(ia)
//     (half the MFMAs of the unfolded product, every output written once).
// Both are persistent: each wave keeps its slice of the sine table in
// registers for the whole launch (the forward pass: one 16-wide k2 tile of
// one parity, all j; the inverse pass: one pair of 16-wide R tiles, all
// k2), and the block walks the work items with the item's 16-row tile staged
// in LDS (the forward pass stages it folded, psi[j] +- psi[n2-1-j]); the
// next items' tiles load into registers during the MFMAs.
// MFMA k index of lane group g at step s: j = 16 kb + 4 g + s (forward) and
// k2 = 32 kb + 8 g + 2 s + parity (inverse), so a lane's operands for four
// steps are one 16-byte read.
// ---------------------------------------------------------------------------
template <typename T>
struct RadPad;  // LDS row padding (elements): rows start 4 banks apart
template <>
struct RadPad<float> {
  static constexpr int v = 4;
};
template <>
struct RadPad<double> {
  static constexpr int v = 2;
};

// waves of k_rad_fwd per (parity, k' tile): two in FP32 when HB is a
// multiple of 4, each over half the j range (the block then has 12 waves for
// three k' tiles, three per SIMD); one otherwise, and in FP64, whose table
// slice leaves no room for more waves
template <typename T, int HB>
struct RadSplit {
  static constexpr int v = sizeof(T) == 4 && HB % 4 == 0 ? 2 : 1;
};

// work items per block of k_rad_fwd / k_rad_inv (the grid is raised to keep
// each block's list within it)
constexpr int kRadMaxItems = 256;

// items of k_rad_fwd whose loads are in flight while a wave multiplies: two
// in FP32; one in FP64, whose registers hold no second set
template <typename T>
struct RadDepth {
  static constexpr int v = sizeof(T) == 4 ? 2 : 1;
};

template <typename T, int NKB, int HB, bool FULL>
__global__ __launch_bounds__(128 * NKB * (RadSplit<T, HB>::v)) void k_rad_fwd(
    const T* __restrict__ psi, const T* __restrict__ Tt, T* __restrict__ t1,
    const int4* __restrict__ items, int nitems, int n1) {
  typedef MfmaAcc<T> M;
  typedef typename M::V V;
  constexpr int KS = RadSplit<T, HB>::v, HBW = HB / KS;  // j blocks per wave (even)
  constexpr int N2 = 32 * HB, H = N2 / 2, NT = 128 * NKB * KS;  // threads
  constexpr int LP = H + RadPad<T>::v;
  // the tile is staged folded: fold[0][row][j] = psi[j] + psi[n2-1-j] and
  // fold[1][row][j] = psi[j] - psi[n2-1-j], j < h; a thread loads the
  // vectors at j and at its mirror and folds them as it stores
  constexpr int NP = 16 * H / 4, PER = (NP + NT - 1) / NT;
  __shared__ __align__(16) T fold[2][2][16 * LP];  // [buffer][parity]
  // partial tiles of both j halves, by buffer (KS = 2)
  __shared__ V red[KS == 2 ? 2 : 1][KS == 2 ? 2 : 1][KS == 2 ? 2 * NKB : 1][64];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, wave = threadIdx.x >> 6;
  const int kh = wave / (2 * NKB), w2 = wave - kh * 2 * NKB;  // j half, tile index
  const int p = w2 / NKB, t = w2 - p * NKB;  // parity and k' tile of this wave
  const int64_t plane = (int64_t)n1 * N2;
  // b[kb][s] = S(16 (kh HBW + kb) + 4 g + s, 2 (16 t + i) + p) over this
  // wave's j blocks
  T b[HBW][4];
  {
    const T* trow = Tt + (int64_t)(2 * (16 * t + i) + p) * H + 16 * kh * HBW + 4 * g;
#pragma unroll
    for (int kb = 0; kb < HBW; kb++) ld4(trow + 16 * kb, b[kb]);
  }
  auto fetch = [&](const int4& it, T(&pf)[PER][2][4]) {
#pragma unroll
    for (int l = 0; l < PER; l++) {
      int v = threadIdx.x + NT * l;
      v = v < NP ? v : NP - 1;  // spare threads repeat the last pair
      const int row = v / (H / 4), c4 = v - row * (H / 4);
      int r = 16 * it.y + row;
      r = r < it.z ? r : it.z - 1;
      const T* src = psi + it.x * plane + (int64_t)r * N2;
      ld4(src + 4 * c4, pf[l][0]);               // psi[j .. j+3], j = 4 c4
      ld4(src + (N2 - 4 - 4 * c4), pf[l][1]);    // psi[n2-4-j .. n2-1-j]
    }
  };
  auto put = [&](const T(&pf)[PER][2][4], int buf) {
#pragma unroll
    for (int l = 0; l < PER; l++) {
      int v = threadIdx.x + NT * l;
      v = v < NP ? v : NP - 1;
      const int row = v / (H / 4), c4 = v - row * (H / 4);
      T e[4], o[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        e[q] = pf[l][0][q] + pf[l][1][3 - q];
        o[q] = pf[l][0][q] - pf[l][1][3 - q];
      }
      st4(&fold[buf][0][row * LP + 4 * c4], e);
      st4(&fold[buf][1][row * LP + 4 * c4], o);
    }
  };
  auto mult = [&](int buf) {
    // two accumulator chains (even and odd j blocks), summed at the end;
    // the operands of the next pair of j blocks are read from LDS before the
    // MFMAs of this pair are issued (the scheduling barrier keeps that
    // order), so the LDS latency hides behind the matrix work
    V acc0 = (V){0, 0, 0, 0}, acc1 = (V){0, 0, 0, 0};
    const T* frow = &fold[buf][p][i * LP + 16 * kh * HBW + 4 * g];
    T x[2][2][4];
    ld4(frow, x[0][0]);
    ld4(frow + 16, x[0][1]);
#pragma unroll
    for (int kb = 0; kb < HBW; kb += 2) {
      const int c = (kb >> 1) & 1;
      if (kb + 2 < HBW) {
        ld4(frow + 16 * (kb + 2), x[c ^ 1][0]);
        ld4(frow + 16 * (kb + 3), x[c ^ 1][1]);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int s = 0; s < 4; s++) {
        acc0 = M::mma(x[c][0][s], b[kb][s], acc0);
        acc1 = M::mma(x[c][1][s], b[kb + 1][s], acc1);
      }
    }
    V acc;
#pragma unroll
    for (int rr = 0; rr < 4; rr++) acc[rr] = acc0[rr] + acc1[rr];
    return acc;
  };
  // With FULL (n1 a multiple of 16) every lane stores, with no branch:
  // rows nn1 .. 16 ceil(nn1 / 16) - 1 (< n1) and columns k2lo .. 32 NKB - 1
  // then hold finite values that no later pass uses (the r-pass tables are
  // zero on those rows, the r-pass factor (radcut - E)_+ is zero on those
  // columns and the inverse pass reads them as zero).  Otherwise rows past
  // n1 are not stored.  With KS = 2 the two halves' tiles meet in LDS (left
  // there by the previous item) and each half stores two of a lane's four
  // rows: every wave runs the same stores, so no branch leaves the compiler
  // unsure how many stores are in flight when it waits for the next loads.
  auto store = [&](const V& acc, const int4& it, int buf) {
    T* orow = t1 + it.x * plane;
    const int k2 = 2 * (16 * t + i) + p;
    if constexpr (KS == 2) {
      const V o = red[buf][kh ^ 1][w2][lane];
      const T v0 = kh ? acc[2] + o[2] : acc[0] + o[0];
      const T v1 = kh ? acc[3] + o[3] : acc[1] + o[1];
      const int r0 = 16 * it.y + M::row(lane, 2 * kh), r1 = 16 * it.y + M::row(lane, 2 * kh + 1);
      if (FULL || r0 < n1) orow[(int64_t)r0 * N2 + k2] = v0;
      if (FULL || r1 < n1) orow[(int64_t)r1 * N2 + k2] = v1;
    } else {
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int row = 16 * it.y + M::row(lane, rr);
        if (FULL || row < n1) orow[(int64_t)row * N2 + k2] = acc[rr];
      }
    }
  };
  auto keep = [&](const V& acc, int buf) {  // this half's partial tile
    if constexpr (KS == 2) red[buf][kh][w2][lane] = acc;
  };
  const int G = (int)gridDim.x;
  if ((int)blockIdx.x >= nitems) return;  // block-uniform
  // this block's items blockIdx.x + k G, k < K, copied to LDS once: read
  // from there they cost an in-order LDS read, where a scalar load would make
  // every later LDS wait a full one
  __shared__ int4 its[kRadMaxItems];
  const int K = (nitems - 1 - (int)blockIdx.x) / G + 1;  // <= kRadMaxItems (host)
  for (int k = threadIdx.x; k < K; k += NT) its[k] = items[blockIdx.x + k * G];
  __syncthreads();
  auto at = [&](int k) { return its[k < K ? k : K - 1]; };
  // Two LDS buffers, one barrier per item: while a wave multiplies item k
  // out of one buffer it has stored item k + 1 (loaded into registers during
  // an earlier item) into the other and issued the loads of a later item.
  // The barrier at the end of an item frees its buffer and completes the
  // other (and the partial tiles of that item).  Each item's results are
  // stored at the start of the next one, before the tile is staged: the wait
  // for the staged loads then leaves those stores in flight.  A prefetch past
  // the last item re-reads the last item, unused; the scheduling barriers
  // keep this order of issue.
  if constexpr (RadDepth<T>::v == 2) {
    // two register sets: the loads of item k are issued two items before its
    // tile is staged, so two items' loads are in flight while a wave
    // multiplies (the loop is unrolled by two so that each set has fixed
    // registers)
    T pa[PER][2][4], pb[PER][2][4];
    fetch(at(0), pa);
    put(pa, 0);
    fetch(at(1), pa);
    fetch(at(2), pb);
    __syncthreads();
    put(pa, 1);
    fetch(at(3), pa);
    __builtin_amdgcn_sched_barrier(0);
    V acc = mult(0);
    keep(acc, 0);
    __syncthreads();
    // item k is staged in buffer b; p holds item k + 1
    auto body = [&](T(&p)[PER][2][4], int k, int b) {
      store(acc, at(k - 1), b ^ 1);
      __builtin_amdgcn_sched_barrier(0);
      put(p, b ^ 1);
      fetch(at(k + 3), p);
      __builtin_amdgcn_sched_barrier(0);
      acc = mult(b);
      keep(acc, b);
      __syncthreads();
    };
    int k = 1, last = 0;  // last: buffer of the last item multiplied
    while (k < K) {
      body(pb, k, 1);
      last = 1;
      if (++k >= K) break;
      body(pa, k, 0);
      last = 0;
      ++k;
    }
    store(acc, at(k - 1), last);
  } else {
    T pf[PER][2][4];
    fetch(at(0), pf);
    put(pf, 0);
    fetch(at(1), pf);
    __syncthreads();
    put(pf, 1);
    fetch(at(2), pf);
    __builtin_amdgcn_sched_barrier(0);
    V acc = mult(0);
    keep(acc, 0);
    __syncthreads();
    int k = 1, buf = 1;
    for (; k < K; k++, buf ^= 1) {
      store(acc, at(k - 1), buf ^ 1);
      __builtin_amdgcn_sched_barrier(0);
      put(pf, buf ^ 1);
      fetch(at(k + 2), pf);
      __builtin_amdgcn_sched_barrier(0);
      acc = mult(buf);
      keep(acc, buf);
      __syncthreads();
    }
    store(acc, at(k - 1), buf ^ 1);
  }
}

// ---------------------------------------------------------------------------
// The forward R pass in FP32 with the sine table in LDS: the same sums as
// k_rad_fwd (t1[r][k2] = sum_{j<h} f_p(r, j) S(j, k2) with the fold f_0 =
// psi[j] + psi[n2-1-j] for even k2, f_1 = psi[j] - psi[n2-1-j] for odd k2),
// but each wave owns whole work items: it streams its 16-row tile from HBM,
// folds it in registers and multiplies it by every low mode, reading the
// table from LDS.  The block copies the table into LDS once (even modes,
// then odd ones, rows padded by 4 floats) and its waves then run with no
// barrier, so one wave's loads overlap another's matrix work.
// Lane (g, i) holds row i of the tile and, at step s of j block kb,
// j = 16 kb + 4 g + s (one 16-byte load at j and one at its mirror); the
// table read for mode tile t of parity p is the 16-byte row segment
// S(16 kb + 4 g .. + 3, 2 (16 t + i) + p).  Loads run RD j blocks ahead;
// the first item's go out before the table copy, and each next item's before
// the stores of the one before.
// Items are spread over the waves round robin by block first, so that every
// SIMD gets about the same number.  Stores as k_rad_fwd: rows past nn1 of
// the last row tile hold finite values no later pass reads (FULL), and the
// two parities of a mode pair are one 8-byte store.
// ---------------------------------------------------------------------------
constexpr int kFwdLdsWaves = 8;

template <int NKB, int HB, bool FULL>
__global__ __launch_bounds__(64 * kFwdLdsWaves) void k_rad_fwd_lds(
    const float* __restrict__ psi, const float* __restrict__ Tt, float* __restrict__ t1,
    const int4* __restrict__ items, int nitems, int n1) {
  typedef MfmaAcc<float> M;
  typedef M::V V;
  constexpr int N2 = 32 * HB, H = N2 / 2, NB = HB;  // j blocks of 16
  constexpr int LP = H + 4, NM = 16 * NKB;           // modes per parity
  constexpr int RD = 4;                              // j blocks loaded ahead
  __shared__ __align__(16) float tab[2][NM * LP];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, w = threadIdx.x >> 6;
  if (nitems <= 0) return;  // grid-uniform
  const int64_t plane = (int64_t)n1 * N2;
  const int nb = (int)gridDim.x, stride = nb * kFwdLdsWaves;
  int k = (int)blockIdx.x + nb * w;
  int4 it = items[k < nitems ? k : nitems - 1];
  const float *src, *mir;
  float a[RD][4], m[RD][4];
  auto start = [&]() {  // the item's first RD j blocks
    int r = 16 * it.y + i;
    r = r < it.z ? r : it.z - 1;
    src = psi + it.x * plane + (int64_t)r * N2 + 4 * g;
    mir = psi + it.x * plane + (int64_t)r * N2 + (N2 - 4 - 4 * g);
#pragma unroll
    for (int d = 0; d < RD; d++) {
      ld4(src + 16 * d, a[d]);
      ld4(mir - 16 * d, m[d]);
    }
  };
  start();  // in flight during the table copy
  {
    // every load of the copy in flight at once (a thread past the end
    // repeats the last vector)
    constexpr int NV = 2 * NM * (H / 4), PT = (NV + 64 * kFwdLdsWaves - 1) / (64 * kFwdLdsWaves);
    float x[PT][4];
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kFwdLdsWaves * l;
      v = v < NV ? v : NV - 1;
      ld4(Tt + (int64_t)(v / (H / 4)) * H + 4 * (v % (H / 4)), x[l]);
    }
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kFwdLdsWaves * l;
      v = v < NV ? v : NV - 1;
      const int row = v / (H / 4), c4 = v % (H / 4);  // row = k2
      st4(&tab[row & 1][(row >> 1) * LP + 4 * c4], x[l]);
    }
  }
  __syncthreads();
  while (k < nitems) {  // wave-uniform
    V acc[2][NKB];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int t = 0; t < NKB; t++) acc[p][t] = (V){0, 0, 0, 0};
    float b[2][2][NKB][4];  // [j block parity][mode parity][tile]
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int t = 0; t < NKB; t++) ld4(&tab[p][(16 * t + i) * LP + 4 * g], b[0][p][t]);
#pragma unroll
    for (int kb = 0; kb < NB; kb++) {
      const int sl = kb % RD, cb = kb & 1;
      float e[4], o[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        e[q] = a[sl][q] + m[sl][3 - q];
        o[q] = a[sl][q] - m[sl][3 - q];
      }
      if (kb + RD < NB) {
        ld4(src + 16 * (kb + RD), a[sl]);
        ld4(mir - 16 * (kb + RD), m[sl]);
      }
      if (kb + 1 < NB) {
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
          for (int t = 0; t < NKB; t++)
            ld4(&tab[p][(16 * t + i) * LP + 16 * (kb + 1) + 4 * g], b[cb ^ 1][p][t]);
      }
      __builtin_amdgcn_sched_barrier(0);
      // one step of every accumulator between two scheduling barriers: no
      // MFMA waits on the one issued just before it
#pragma unroll
      for (int s = 0; s < 4; s++) {
#pragma unroll
        for (int t = 0; t < NKB; t++) {
          acc[0][t] = M::mma(e[s], b[cb][0][t][s], acc[0][t]);
          acc[1][t] = M::mma(o[s], b[cb][1][t][s], acc[1][t]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    float* orow = t1 + it.x * plane;
    const int y0 = 16 * it.y;
    k += stride;
    if (k < nitems) {  // the next item's first loads go out before the stores
      it = items[k];
      start();
    }
#pragma unroll
    for (int t = 0; t < NKB; t++)
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int row = y0 + M::row(lane, rr);
        typedef float F2 __attribute__((ext_vector_type(2)));
        if (FULL || row < n1)
          *(F2*)&orow[(int64_t)row * N2 + 2 * (16 * t + i)] = (F2){acc[0][t][rr], acc[1][t][rr]};
      }
  }
}

// ---------------------------------------------------------------------------
// The inverse R pass in FP32 with the sine table in LDS: the sums of
// k_rad_inv (corr[r][R] = E + O, corr[r][n2-1-R] = E - O for R < h, E / O
// over the even / odd modes k2 < k2lo) with each wave owning whole work
// items: it loads its 16 rows of t1 once (columns k2 >= k2lo read as zero),
// then walks the 16-wide R tiles of the first half, reading the table
// S(R, k2) from LDS one tile ahead, and skips the tile pairs whose rows hold
// no compressed-grid point in those columns or their mirrors (the update
// reads none of that output).  The block copies the table (rows R < h,
// padded by 4 floats) into LDS once; the waves then run with no barrier.
// Lane (g, i) holds row i and modes 32 kb + 8 g + e (e < 8: steps s take
// e = 2 s + parity), and reads the table row R = 16 tq + i at the same modes;
// with the table as the MFMA's A operand, the lane's results are
// R = 16 tq + 4 g .. + 3 of row i: two 16-byte stores per tile.  The first
// item's loads are issued before the table copy, so the two overlap.
// ---------------------------------------------------------------------------
constexpr int kInvLdsWaves = 8;

template <int NKB, int HB, bool FULL>
__global__ __launch_bounds__(64 * kInvLdsWaves) void k_rad_inv_lds(
    const float* __restrict__ t1, const float* __restrict__ Ti, float* __restrict__ corr,
    const int4* __restrict__ items, int nitems, int n1) {
  typedef MfmaAcc<float> M;
  typedef M::V V;
  constexpr int N2 = 32 * HB, H = N2 / 2, NTQ = H / 16;  // R tiles of the first half
  constexpr int K2P = 32 * NKB, LP = K2P + 4;
  __shared__ __align__(16) float tab[H * LP];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, w = threadIdx.x >> 6;
  if (nitems <= 0) return;  // grid-uniform
  const int64_t plane = (int64_t)n1 * N2;
  const int nb = (int)gridDim.x, stride = nb * kInvLdsWaves;
  int k = (int)blockIdx.x + nb * w;
  // the item's 16 rows of t1: lane (g, i) row i, modes 32 kb + 8 g .. + 7
  int4 it = items[k < nitems ? k : nitems - 1];
  float a[NKB][2][4];
  auto load_a = [&]() {
    int r = 16 * it.y + i;
    r = r < it.z ? r : it.z - 1;
    const float* src = t1 + it.x * plane + (int64_t)r * N2 + 8 * g;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
      ld4(src + 32 * kb, a[kb][0]);
      ld4(src + 32 * kb + 4, a[kb][1]);
    }
  };
  load_a();  // in flight during the table copy
  {
    constexpr int NV = H * (K2P / 4), PT = (NV + 64 * kInvLdsWaves - 1) / (64 * kInvLdsWaves);
    float x[PT][4];
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kInvLdsWaves * l;
      v = v < NV ? v : NV - 1;
      ld4(Ti + (int64_t)(v / (K2P / 4)) * K2P + 4 * (v % (K2P / 4)), x[l]);
    }
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kInvLdsWaves * l;
      v = v < NV ? v : NV - 1;
      st4(&tab[(v / (K2P / 4)) * LP + 4 * (v % (K2P / 4))], x[l]);
    }
  }
  __syncthreads();
  float b[2][NKB][2][4];
  auto rdb = [&](int tq, float (&bb)[NKB][2][4]) {
    const float* trow = &tab[(16 * tq + i) * LP + 8 * g];
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
      ld4(trow + 32 * kb, bb[kb][0]);
      ld4(trow + 32 * kb + 4, bb[kb][1]);
    }
  };
  while (k < nitems) {  // wave-uniform
    rdb(0, b[0]);
    const int kl = it.w & 0xffff;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++)
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (32 * kb + 8 * g + e >= kl) a[kb][e >> 2][e & 3] = 0.0f;
    float* orow = corr + it.x * plane + (int64_t)(16 * it.y + i) * N2;
    const bool live = FULL || 16 * it.y + i < it.z;
#pragma unroll
    for (int tq = 0; tq < NTQ; tq++) {
      if (tq + 1 < NTQ) rdb(tq + 1, b[(tq + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);  // one tile ahead, no further
      if ((it.w >> (16 + (tq >> 1))) & 1) {  // wave-uniform
        // table as the A operand: the result's rows are R = 16 tq + 4 g + rr
        // and its columns the t1 rows, so a lane stores four consecutive R
        V E = (V){0, 0, 0, 0}, O = (V){0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < NKB; kb++)
#pragma unroll
          for (int s = 0; s < 4; s++) {
            const int e0 = 2 * s, e1 = 2 * s + 1;
            E = M::mma(b[tq & 1][kb][e0 >> 2][e0 & 3], a[kb][e0 >> 2][e0 & 3], E);
            O = M::mma(b[tq & 1][kb][e1 >> 2][e1 & 3], a[kb][e1 >> 2][e1 & 3], O);
          }
        float ep[4], em[4];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          ep[rr] = E[rr] + O[rr];
          em[3 - rr] = E[rr] - O[rr];
        }
        if (live) {
          st4(orow + 16 * tq + 4 * g, ep);
          st4(orow + (N2 - 4 - 16 * tq - 4 * g), em);
        }
      }
    }
    k += stride;
    if (k < nitems) {
      it = items[k];
      load_a();
    }
  }
}

// ---------------------------------------------------------------------------
// The inverse R pass of k_rad_inv_lds with its products on the FP16 matrix
// cores, compensated: each FP32 operand x is split into x_h = fp16(x) and
// x_l = fp16(x - x_h) (22 significant bits), and the sums take
// A_h B_h + A_h B_l + A_l B_h, accumulated in FP32.  The term dropped,
// A_l B_l, and the splits are each about 2^-22 of |A| |B|, the size of FP32
// rounding over these sums.  The table (|S| <= 2) is split in the LDS copy;
// each t1 row is first scaled by a power of two so that its largest entry
// lies in [2^14, 2^15) (the FP16 range is 2^-24 .. 65504), and the results of
// that row are scaled back exactly.  v_mfma_f32_16x16x32_f16 does 8192
// multiply-adds per issue against 1024 for v_mfma_f32_16x16x4_f32 at
// half the cycles, so the three products cost 3/16 of the FP32 ones.
// Per parity p the modes 2 k' + p, k' < 16 NKB, run as ceil(NKB / 2)
// products of depth 32; lane (g, i) holds k' = 32 m + 8 g .. + 7 of row i
// (the same k' slots in the table operand, whose rows are R = 16 tq + i), so
// a lane's t1 values are the 16 consecutive modes 64 m + 16 g .. + 15.
// Modes at or past k2lo are zero; the table slots past 16 NKB read the last
// 8 of the row instead (finite, times zero).  Stores, tile skipping and item
// order as k_rad_inv_lds.
// ---------------------------------------------------------------------------
template <int NKB, int HB, bool FULL>
__global__ __launch_bounds__(64 * kInvLdsWaves) void k_rad_inv_h(
    const float* __restrict__ t1, const float* __restrict__ Ti, float* __restrict__ corr,
    const int4* __restrict__ items, int nitems, int n1) {
  typedef _Float16 H8 __attribute__((ext_vector_type(8)));
  typedef _Float16 H2 __attribute__((ext_vector_type(2)));
  typedef float V __attribute__((ext_vector_type(4)));
  constexpr int N2 = 32 * HB, H = N2 / 2, NTQ = H / 16;
  constexpr int K2P = 32 * NKB, KH = 16 * NKB, LPH = KH + 8, MK = (NKB + 1) / 2;
  constexpr int PART = H * LPH;  // halves per table part: E hi, E lo, O hi, O lo
  __shared__ __align__(16) _Float16 tab[4 * PART];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, w = threadIdx.x >> 6;
  if (nitems <= 0) return;  // grid-uniform
  const int64_t plane = (int64_t)n1 * N2;
  const int nb = (int)gridDim.x, stride = nb * kInvLdsWaves;
  int k = (int)blockIdx.x + nb * w;
  int4 it = items[k < nitems ? k : nitems - 1];
  float a[MK][16];  // row i of t1, modes 64 m + 16 g + c
  auto load_a = [&]() {
    int r = 16 * it.y + i;
    r = r < it.z ? r : it.z - 1;
    const float* src = t1 + it.x * plane + (int64_t)r * N2 + 16 * g;
#pragma unroll
    for (int m = 0; m < MK; m++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        float x[4];
        ld4(src + 64 * m + 4 * q, x);
#pragma unroll
        for (int c = 0; c < 4; c++) a[m][4 * q + c] = x[c];
      }
  };
  load_a();  // in flight during the table copy
  {
    constexpr int NV = H * (K2P / 4), PT = (NV + 64 * kInvLdsWaves - 1) / (64 * kInvLdsWaves);
    float x[PT][4];
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kInvLdsWaves * l;
      v = v < NV ? v : NV - 1;
      ld4(Ti + (int64_t)(v / (K2P / 4)) * K2P + 4 * (v % (K2P / 4)), x[l]);
    }
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kInvLdsWaves * l;
      v = v < NV ? v : NV - 1;
      // modes 4 c4 .. + 3 of row R: k' = 2 c4, 2 c4 + 1 of each parity
      const int R = v / (K2P / 4), c4 = v % (K2P / 4), o = R * LPH + 2 * c4;
#pragma unroll
      for (int p = 0; p < 2; p++) {
        const float s0 = x[l][p], s1 = x[l][p + 2];
        const _Float16 h0 = (_Float16)s0, h1 = (_Float16)s1;
        *(H2*)&tab[(2 * p) * PART + o] = (H2){h0, h1};
        *(H2*)&tab[(2 * p + 1) * PART + o] =
            (H2){(_Float16)(s0 - (float)h0), (_Float16)(s1 - (float)h1)};
      }
    }
  }
  __syncthreads();
  H8 ta[2][2][MK][2];  // [buffer][parity][m][hi, lo]
  auto rdb = [&](int tq, H8 (&bb)[2][MK][2]) {
    const int row = (16 * tq + i) * LPH;
#pragma unroll
    for (int m = 0; m < MK; m++) {
      int kk = 32 * m + 8 * g;
      kk = kk + 8 <= KH ? kk : KH - 8;
#pragma unroll
      for (int p = 0; p < 2; p++)
#pragma unroll
        for (int h = 0; h < 2; h++) bb[p][m][h] = *(const H8*)&tab[(2 * p + h) * PART + row + kk];
    }
  };
  while (k < nitems) {  // wave-uniform
    rdb(0, ta[0]);
    // zero the modes at or past k2lo, scale the row, split it
    const int kl = it.w & 0xffff;
    float mx = 0.0f;
#pragma unroll
    for (int m = 0; m < MK; m++)
#pragma unroll
      for (int c = 0; c < 16; c++) {
        if (64 * m + 16 * g + c >= kl) a[m][c] = 0.0f;
        mx = fmaxf(mx, fabsf(a[m][c]));
      }
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    int ex = 0;
    if (mx > 0.0f) {
      frexpf(mx, &ex);
      ex = 15 - ex;  // mx 2^ex in [2^14, 2^15)
    }
    H8 tb[2][MK][2];  // [parity][m][hi, lo]
#pragma unroll
    for (int m = 0; m < MK; m++)
#pragma unroll
      for (int p = 0; p < 2; p++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const float xs = ldexpf(a[m][2 * j + p], ex);
          const _Float16 hx = (_Float16)xs;
          tb[p][m][0][j] = hx;
          tb[p][m][1][j] = (_Float16)(xs - (float)hx);
        }
    float* orow = corr + it.x * plane + (int64_t)(16 * it.y + i) * N2;
    const bool live = FULL || 16 * it.y + i < it.z;
#pragma unroll
    for (int tq = 0; tq < NTQ; tq++) {
      if (tq + 1 < NTQ) rdb(tq + 1, ta[(tq + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);  // one tile ahead, no further
      if ((it.w >> (16 + (tq >> 1))) & 1) {  // wave-uniform
        // table as the A operand: rows R = 16 tq + 4 g + rr, columns the t1
        // rows; the cross terms sum apart from the main ones
        V E0 = (V){0, 0, 0, 0}, E1 = E0, O0 = E0, O1 = E0;
        const auto& A = ta[tq & 1];
#pragma unroll
        for (int m = 0; m < MK; m++) {
          E0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][m][0], tb[0][m][0], E0, 0, 0, 0);
          O0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][m][0], tb[1][m][0], O0, 0, 0, 0);
          E1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][m][0], tb[0][m][1], E1, 0, 0, 0);
          O1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][m][0], tb[1][m][1], O1, 0, 0, 0);
          E1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[0][m][1], tb[0][m][0], E1, 0, 0, 0);
          O1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[1][m][1], tb[1][m][0], O1, 0, 0, 0);
        }
        float ep[4], em[4];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          const float e = E0[rr] + E1[rr], o = O0[rr] + O1[rr];
          ep[rr] = ldexpf(e + o, -ex);
          em[3 - rr] = ldexpf(e - o, -ex);
        }
        if (live) {
          st4(orow + 16 * tq + 4 * g, ep);
          st4(orow + (N2 - 4 - 16 * tq - 4 * g), em);
        }
      }
    }
    k += stride;
    if (k < nitems) {
      it = items[k];
      load_a();
    }
  }
}

// ---------------------------------------------------------------------------
// The FP64 R passes with the sine table in LDS.  The FP64 tables (about 200
// KB) do not fit one block's LDS, so each is split over a pair of blocks on
// one XCD (blocks b and b + 8, as workgroups go to the XCDs round robin by
// linear id): in the forward pass the two mode parities, in the inverse pass
// the two halves of the R range.  Both blocks of a pair walk the same work
// items with the same waves, so the second block's reads of an item's rows
// hit the L2 the first one filled.  Otherwise as k_rad_fwd_lds /
// k_rad_inv_lds: the table copied into LDS once, then one wave per item with
// no barrier.  The grid is a multiple of 16 blocks.
// ---------------------------------------------------------------------------
constexpr int kRadLds64Waves = 8;

__device__ __forceinline__ void lds64_pair(int& part, int& pair, int& npairs) {
  const int b = (int)blockIdx.x;
  part = (b >> 3) & 1;
  pair = (b >> 4) * 8 + (b & 7);
  npairs = (int)gridDim.x / 2;
}

template <int NKB, int HB, bool FULL>
__global__ __launch_bounds__(64 * kRadLds64Waves) void k_rad_fwd_lds64(
    const double* __restrict__ psi, const double* __restrict__ Tt, double* __restrict__ t1,
    const int4* __restrict__ items, int nitems, int n1) {
  typedef MfmaAcc<double> M;
  typedef M::V V;
  constexpr int N2 = 32 * HB, H = N2 / 2, NB = HB;  // j blocks of 16
  constexpr int LP = H + 2, NM = 16 * NKB;           // modes of one parity
  constexpr int RD = 3;                              // j blocks loaded ahead
  __shared__ __align__(16) double tab[NM * LP];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, w = threadIdx.x >> 6;
  if (nitems <= 0) return;  // grid-uniform
  int p, pair, npairs;
  lds64_pair(p, pair, npairs);
  const int64_t plane = (int64_t)n1 * N2;
  const int stride = npairs * kRadLds64Waves;
  int k = pair + npairs * w;
  int4 it = items[k < nitems ? k : nitems - 1];
  const double *src, *mir;
  double a[RD][4], m[RD][4];
  auto start = [&]() {
    int r = 16 * it.y + i;
    r = r < it.z ? r : it.z - 1;
    src = psi + it.x * plane + (int64_t)r * N2 + 4 * g;
    mir = psi + it.x * plane + (int64_t)r * N2 + (N2 - 4 - 4 * g);
#pragma unroll
    for (int d = 0; d < RD; d++) {
      ld4(src + 16 * d, a[d]);
      ld4(mir - 16 * d, m[d]);
    }
  };
  start();  // in flight during the table copy
  {
    // rows k2 = 2 m + p of the table, every load in flight at once
    constexpr int NV = NM * (H / 4), PT = (NV + 64 * kRadLds64Waves - 1) / (64 * kRadLds64Waves);
    double x[PT][4];
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kRadLds64Waves * l;
      v = v < NV ? v : NV - 1;
      ld4(Tt + (int64_t)(2 * (v / (H / 4)) + p) * H + 4 * (v % (H / 4)), x[l]);
    }
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kRadLds64Waves * l;
      v = v < NV ? v : NV - 1;
      st4(&tab[(v / (H / 4)) * LP + 4 * (v % (H / 4))], x[l]);
    }
  }
  __syncthreads();
  const double sg = p ? -1.0 : 1.0;
  while (k < nitems) {  // wave-uniform
    V acc[NKB];
#pragma unroll
    for (int t = 0; t < NKB; t++) acc[t] = (V){0, 0, 0, 0};
    double b[2][NKB][4];
#pragma unroll
    for (int t = 0; t < NKB; t++) ld4(&tab[(16 * t + i) * LP + 4 * g], b[0][t]);
#pragma unroll
    for (int kb = 0; kb < NB; kb++) {
      const int sl = kb % RD, cb = kb & 1;
      double f[4];  // psi[j] + psi[n2-1-j] (even modes) or psi[j] - psi[n2-1-j] (odd)
#pragma unroll
      for (int q = 0; q < 4; q++) f[q] = a[sl][q] + sg * m[sl][3 - q];
      if (kb + RD < NB) {
        ld4(src + 16 * (kb + RD), a[sl]);
        ld4(mir - 16 * (kb + RD), m[sl]);
      }
      if (kb + 1 < NB) {
#pragma unroll
        for (int t = 0; t < NKB; t++) ld4(&tab[(16 * t + i) * LP + 16 * (kb + 1) + 4 * g], b[cb ^ 1][t]);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int s = 0; s < 4; s++) {
#pragma unroll
        for (int t = 0; t < NKB; t++) acc[t] = M::mma(f[s], b[cb][t][s], acc[t]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    double* orow = t1 + it.x * plane;
    const int y0 = 16 * it.y;
    k += stride;
    if (k < nitems) {  // the next item's first loads go out before the stores
      it = items[k];
      start();
    }
#pragma unroll
    for (int t = 0; t < NKB; t++)
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int row = y0 + M::row(lane, rr);
        if (FULL || row < n1) orow[(int64_t)row * N2 + 2 * (16 * t + i) + p] = acc[t][rr];
      }
  }
}

template <int NKB, int HB, bool FULL>
__global__ __launch_bounds__(64 * kRadLds64Waves) void k_rad_inv_lds64(
    const double* __restrict__ t1, const double* __restrict__ Ti, double* __restrict__ corr,
    const int4* __restrict__ items, int nitems, int n1) {
  typedef MfmaAcc<double> M;
  typedef M::V V;
  constexpr int N2 = 32 * HB, H = N2 / 2, NTQ = H / 16, NTH = NTQ / 2;  // R tiles per half
  constexpr int K2P = 32 * NKB, LP = K2P + 2;
  __shared__ __align__(16) double tab[16 * NTH * LP];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, w = threadIdx.x >> 6;
  if (nitems <= 0) return;  // grid-uniform
  int h, pair, npairs;
  lds64_pair(h, pair, npairs);
  const int64_t plane = (int64_t)n1 * N2;
  const int stride = npairs * kRadLds64Waves;
  int k = pair + npairs * w;
  int4 it = items[k < nitems ? k : nitems - 1];
  double a[NKB][2][4];
  auto load_a = [&]() {
    int r = 16 * it.y + i;
    r = r < it.z ? r : it.z - 1;
    const double* src = t1 + it.x * plane + (int64_t)r * N2 + 8 * g;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
      ld4(src + 32 * kb, a[kb][0]);
      ld4(src + 32 * kb + 4, a[kb][1]);
    }
  };
  load_a();  // in flight during the table copy
  {
    // rows R = 16 NTH h .. of the table
    constexpr int NV = 16 * NTH * (K2P / 4);
    constexpr int PT = (NV + 64 * kRadLds64Waves - 1) / (64 * kRadLds64Waves);
    double x[PT][4];
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kRadLds64Waves * l;
      v = v < NV ? v : NV - 1;
      ld4(Ti + (int64_t)(16 * NTH * h + v / (K2P / 4)) * K2P + 4 * (v % (K2P / 4)), x[l]);
    }
#pragma unroll
    for (int l = 0; l < PT; l++) {
      int v = threadIdx.x + 64 * kRadLds64Waves * l;
      v = v < NV ? v : NV - 1;
      st4(&tab[(v / (K2P / 4)) * LP + 4 * (v % (K2P / 4))], x[l]);
    }
  }
  __syncthreads();
  while (k < nitems) {  // wave-uniform
    const int kl = it.w & 0xffff;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++)
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (32 * kb + 8 * g + e >= kl) a[kb][e >> 2][e & 3] = 0.0;
    double* ob = corr + it.x * plane;
#pragma unroll
    for (int tl = 0; tl < NTH; tl++) {
      const int tq = NTH * h + tl;
      if ((it.w >> (16 + (tq >> 1))) & 1) {  // wave-uniform
        double b[NKB][2][4];
        const double* trow = &tab[(16 * tl + i) * LP + 8 * g];
#pragma unroll
        for (int kb = 0; kb < NKB; kb++) {
          ld4(trow + 32 * kb, b[kb][0]);
          ld4(trow + 32 * kb + 4, b[kb][1]);
        }
        V E = (V){0, 0, 0, 0}, O = (V){0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < NKB; kb++)
#pragma unroll
          for (int s = 0; s < 4; s++) {
            const int e0 = 2 * s, e1 = 2 * s + 1;
            E = M::mma(a[kb][e0 >> 2][e0 & 3], b[kb][e0 >> 2][e0 & 3], E);
            O = M::mma(a[kb][e1 >> 2][e1 & 3], b[kb][e1 >> 2][e1 & 3], O);
          }
        const int R = 16 * tq + i;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          const int row = 16 * it.y + M::row(lane, rr);
          if (FULL || row < it.z) {
            ob[(int64_t)row * N2 + R] = E[rr] + O[rr];
            ob[(int64_t)row * N2 + (N2 - 1 - R)] = E[rr] - O[rr];
          }
        }
      }
    }
    k += stride;
    if (k < nitems) {
      it = items[k];
      load_a();
    }
  }
}

template <typename T, int NKB, int HB, bool FULL>
__global__ __launch_bounds__(32 * HB) void k_rad_inv(const T* __restrict__ t1,
                                                     const T* __restrict__ Ti,
                                                     T* __restrict__ corr,
                                                     const int4* __restrict__ items, int nitems,
                                                     int n1) {
  typedef MfmaAcc<T> M;
  typedef typename M::V V;
  constexpr int N2 = 32 * HB, NT = 32 * HB;  // one wave per pair of R tiles of the first half
  constexpr int K2P = 32 * NKB, LP = K2P + RadPad<T>::v;
  constexpr int NV4 = 16 * K2P / 4, PER = (NV4 + NT - 1) / NT;
  __shared__ __align__(16) T tile[2][16 * LP];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, pair = threadIdx.x >> 6;
  const int64_t plane = (int64_t)n1 * N2;
  // b[q][kb][e] = S(R, 32 kb + 8 g + e), R = 16 (2 pair + q) + i
  T b[2][NKB][2][4];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const T* trow = Ti + (int64_t)(16 * (2 * pair + q) + i) * K2P + 8 * g;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
      ld4(trow + 32 * kb, b[q][kb][0]);
      ld4(trow + 32 * kb + 4, b[q][kb][1]);
    }
  }
  T pf[PER][4];
  auto fetch = [&](const int4& it) {
#pragma unroll
    for (int l = 0; l < PER; l++) {
      int v = threadIdx.x + NT * l;
      v = v < NV4 ? v : NV4 - 1;
      const int row = v / (K2P / 4), c4 = v - row * (K2P / 4);
      int r = 16 * it.y + row;
      r = r < it.z ? r : it.z - 1;
      ld4(t1 + it.x * plane + (int64_t)r * N2 + 4 * c4, pf[l]);
    }
  };
  auto put = [&](const int4& it, int buf) {  // columns k2 >= k2lo enter as zero
#pragma unroll
    for (int l = 0; l < PER; l++) {
      int v = threadIdx.x + NT * l;
      v = v < NV4 ? v : NV4 - 1;
      const int row = v / (K2P / 4), c4 = v - row * (K2P / 4);
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (4 * c4 + e >= (it.w & 0xffff)) pf[l][e] = 0;
      st4(&tile[buf][row * LP + 4 * c4], pf[l]);
    }
  };
  // pairs whose 16 rows hold no compressed-grid point in their columns or
  // the mirrored ones: the update reads none of their output
  auto active = [&](const int4& it) { return (it.w >> (16 + pair)) & 1; };
  V E[2], O[2];
  auto mult = [&](int buf) {
    T a[NKB][2][4];
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
      ld4(&tile[buf][i * LP + 32 * kb + 8 * g], a[kb][0]);
      ld4(&tile[buf][i * LP + 32 * kb + 8 * g + 4], a[kb][1]);
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      E[q] = (V){0, 0, 0, 0};
      O[q] = (V){0, 0, 0, 0};
#pragma unroll
      for (int kb = 0; kb < NKB; kb++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const int e0 = 2 * s, e1 = 2 * s + 1;
          E[q] = M::mma(a[kb][e0 >> 2][e0 & 3], b[q][kb][e0 >> 2][e0 & 3], E[q]);
          O[q] = M::mma(a[kb][e1 >> 2][e1 & 3], b[q][kb][e1 >> 2][e1 & 3], O[q]);
        }
    }
  };
  // With FULL (n1 a multiple of 16) rows nn1 .. 16 ceil(nn1 / 16) - 1 are
  // stored too (below n1; the update reads no row past nn1), with no
  // per-lane branch; otherwise rows past nn1 are not stored.
  auto store = [&](const int4& it) {
    T* ob = corr + it.x * plane;
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int R = 16 * (2 * pair + q) + i;
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int row = 16 * it.y + M::row(lane, rr);
        if (FULL || row < it.z) {
          ob[(int64_t)row * N2 + R] = E[q][rr] + O[q][rr];
          ob[(int64_t)row * N2 + (N2 - 1 - R)] = E[q][rr] - O[q][rr];
        }
      }
    }
  };
  const int G = (int)gridDim.x;
  if ((int)blockIdx.x >= nitems) return;  // block-uniform
  // this block's items in LDS, as in k_rad_fwd
  __shared__ int4 its[kRadMaxItems];
  const int K = (nitems - 1 - (int)blockIdx.x) / G + 1;  // <= kRadMaxItems (host)
  for (int k = threadIdx.x; k < K; k += NT) its[k] = items[blockIdx.x + k * G];
  __syncthreads();
  auto at = [&](int k) { return its[k < K ? k : K - 1]; };
  // Two LDS buffers, one barrier per item, as in k_rad_fwd: item k is
  // multiplied out of one buffer after item k + 1 has been staged into the
  // other and the loads of item k + 2 issued; its results are stored at the
  // start of the next item, after the staging.  A prefetch past the last
  // item re-reads the last item, unused.
  int4 cur = at(0);
  fetch(cur);
  put(cur, 0);
  int4 nxt = at(1);
  fetch(nxt);
  __syncthreads();
  int4 prev = cur;
  bool have = false;  // results of prev in E, O
  for (int k = 0, buf = 0; k < K; k++, buf ^= 1) {
    put(nxt, buf ^ 1);
    if (have && active(prev)) store(prev);
    const int4 nn = at(k + 2);
    fetch(nn);
    __builtin_amdgcn_sched_barrier(0);  // issue the loads before the MFMAs
    if (active(cur)) mult(buf);
    have = true;
    __syncthreads();
    prev = cur;
    cur = nxt;
    nxt = nn;
  }
  if (active(prev)) store(prev);
}

// ---------------------------------------------------------------------------
// The two r passes of the radial correction in one kernel, one block of four
// waves per work item = (channel-angle, 16 columns k2 of the low R modes):
//   t2[k1][k2] = (radcut - p2r1(k1) - p2r2(k2))_+ divi sum_{r<nn1} S(r, k1) t1[r][k2]
//   t1[r][k2]  = sum_{k1<k1lo} S(r, k1) t2[k1][k2]            (in place)
// S(r, k1) from the low-mode table L[k1][r] of size nn1 (forward) and its
// transpose LT[r][k1] (back; within each 16-column block of FP64 tables the
// columns run k1 = 16 m + 4 rr + g at position 16 m + 4 g + rr, so that a
// lane's four k1 of one step group are one read).  Both tables are padded
// with zeros to 16 NK1 modes and to whole 16-row blocks of r, over every k1
// at or past k1lo and every r at or past nn1: the products then need no
// masks on loaded operands (a select on a freshly loaded value made the
// compiler wait for every outstanding load, the next block's prefetch
// included).  The t1 rows read past nn1 (below n1) are finite: the forward
// R pass never writes them and the slab starts zeroed.  t2 stays in the
// MFMA accumulators: their layout is the back product's B operand (as in
// k_ang_fused).  Items: int4 {channel * na + angle, column block | k2lo << 16,
// nn1 | k1lo << 16, size index}; sizes: int4 {L offset, L row stride,
// LT offset, 0}, offsets into the one table buffer.
// ---------------------------------------------------------------------------
template <typename T, int NK1>
__global__ __launch_bounds__(256) void k_rad_mid(T* __restrict__ t1, const T* __restrict__ tab,
                                                 const T* __restrict__ tabT,
                                                 const int4* __restrict__ items, int nitems,
                                                 const int4* __restrict__ sizes,
                                                 const double* __restrict__ p2r1,
                                                 const double* __restrict__ p2r2,
                                                 const double* __restrict__ divi, double radcut,
                                                 int na, int n1, int n2) {
  typedef MfmaAcc<T> M;
  typedef typename M::V V;
  constexpr int KTP = 16 * NK1;  // row stride of LT
  __shared__ V red[4][NK1][64];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, w = threadIdx.x >> 6;
  // XCD-aware: XCD x (blocks x, x + 8, ..) runs items [x q8, (x + 1) q8),
  // so the items of one table, which are consecutive, share that XCD's L2
  const int hw = blockIdx.x, q8 = (int)(gridDim.x >> 3);
  const int4 it = items[hw < 8 * q8 ? (hw & 7) * q8 + (hw >> 3) : hw];
  const int cg = it.x, cc = it.y & 0xffff, k2l = it.y >> 16, nn = it.z & 0xffff,
            k1l = it.z >> 16;
  const int4 sz = sizes[it.w];
  const int ia = cg % na;
  T* tb = t1 + (int64_t)cg * n1 * n2;
  const T* L = tabT + sz.x;
  const int ld = sz.y;
  const T* LT = tabT + sz.z;
  (void)tab;
  const int col = 16 * cc + i;
  const int colc = col < n2 ? col : n2 - 1;
  const int nrb = (nn + 15) >> 4;
  // forward over r, r blocks rb = w, w + 4, ..: acc[m] rows k1 = 16 m + (A
  // row), columns k2 = col; the four waves' partial tiles are summed in LDS
  V acc[NK1];
#pragma unroll
  for (int m = 0; m < NK1; m++) acc[m] = (V){0, 0, 0, 0};
  T a[2][NK1][4], b[2][4];
  (void)k1l;
  auto load = [&](int rb, int q) {
    const int r0 = 16 * rb + 4 * g;
#pragma unroll
    for (int m = 0; m < NK1; m++) ld4(L + (int64_t)(16 * m + i) * ld + r0, a[q][m]);
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int r = r0 + s;
      b[q][s] = tb[(int64_t)(r < n1 ? r : n1 - 1) * n2 + colc];
    }
  };
  auto compute = [&](int q) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int m = 0; m < NK1; m++) acc[m] = M::mma(a[q][m][s], b[q][s], acc[m]);
  };
  // this wave's blocks w, w + 4, ..; a missing second block repeats the
  // first and is not computed
  const int nmine = nrb > w ? (nrb - w + 3) / 4 : 0;
  if (nmine > 0) {
    load(w, 0);
    for (int m2 = 0; m2 < nmine; m2 += 2) {
      const int rb0 = w + 4 * m2, rb1 = rb0 + 4;
      load(m2 + 1 < nmine ? rb1 : rb0, 1);
      compute(0);
      load(m2 + 2 < nmine ? rb1 + 4 : rb0, 0);
      if (m2 + 1 < nmine) compute(1);
    }
  }
#pragma unroll
  for (int m = 0; m < NK1; m++) red[w][m][lane] = acc[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NK1; m++) {
    acc[m] = red[0][m][lane];
#pragma unroll
    for (int q = 1; q < 4; q++) acc[m] += red[q][m][lane];
  }
  // (radcut - E)_+ divi, rounded to the operator precision like the stored t2
  {
    const double e2 = p2r2[colc], dv = divi[ia];
#pragma unroll
    for (int m = 0; m < NK1; m++)
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        int k1 = 16 * m + M::row(lane, rr);
        k1 = k1 < n1 ? k1 : n1 - 1;
        double wk = radcut - (p2r1[(int64_t)ia * n1 + k1] + e2);
        wk = wk > 0.0 ? wk * dv : 0.0;
        acc[m][rr] = (T)((double)acc[m][rr] * wk);
      }
  }
  // back over r: output tiles mt = w, w + 4, ..; B operand = acc
  for (int mt = w; mt < nrb; mt += 4) {
    T at[NK1][4];
#pragma unroll
    for (int m = 0; m < NK1; m++) ld4(LT + (int64_t)(16 * mt + i) * KTP + 16 * m + 4 * g, at[m]);
    V o = (V){0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < NK1; m++)
#pragma unroll
      for (int rr = 0; rr < 4; rr++) o = M::mma(at[m][rr], acc[m][rr], o);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const int r = 16 * mt + M::row(lane, rr);
      if (r < nn && col < k2l) tb[(int64_t)r * n2 + col] = o[rr];
    }
  }
}

// ---------------------------------------------------------------------------
// The two r passes in FP32, every load of an item issued before its first
// product: one block of four waves per item (channel-angle, 16 columns k2),
// wave w owning the r blocks w, w + 4, .. (RBW of them) in both products.
// An item is two int4: {channel * na + angle, column block | k2lo << 16,
// nn1 | k1lo << 16, L offset}, {L row stride, LT offset, weight offset, 0};
// the weights (radcut - E)_+ divi come from a table W[angle][k2][k1] (k1
// padded to 16 NK1), so the item needs no other read.  Tables and operand
// layouts as k_rad_mid; the back product keeps two accumulators (steps of
// even and odd index), summed at the end.
// ---------------------------------------------------------------------------
template <int NK1, int RBW>
__global__ __launch_bounds__(256) void k_rad_mid_f(float* __restrict__ t1,
                                                   const float* __restrict__ tabT,
                                                   const float* __restrict__ wts,
                                                   const int4* __restrict__ items, int n1, int n2) {
  typedef MfmaAcc<float> M;
  typedef M::V V;
  constexpr int KTP = 16 * NK1;
  __shared__ V red[4][NK1][64];
  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15, w = threadIdx.x >> 6;
  // XCD-aware: XCD x (blocks x, x + 8, ..) runs items [x q8, (x + 1) q8),
  // so the items of one table, which are consecutive, share that XCD's L2
  const int hw = blockIdx.x, q8 = (int)(gridDim.x >> 3);
  const int it = hw < 8 * q8 ? (hw & 7) * q8 + (hw >> 3) : hw;
  const int4 i0 = items[2 * it], i1 = items[2 * it + 1];
  const int cg = i0.x, cc = i0.y & 0xffff, k2l = i0.y >> 16, nn = i0.z & 0xffff;
  float* tb = t1 + (int64_t)cg * n1 * n2;
  const float* L = tabT + i0.w;
  const int ld = i1.x;
  const float* LT = tabT + i1.y;
  const int col = 16 * cc + i;
  const int colc = col < n2 ? col : n2 - 1;
  const int nrb = (nn + 15) >> 4;  // <= 4 RBW (host)
  float a[RBW][NK1][4], b[RBW][4], at[RBW][NK1][4], wk[NK1][4];
#pragma unroll
  for (int q = 0; q < RBW; q++) {
    const int rb = w + 4 * q < nrb ? w + 4 * q : nrb - 1;
    const int r0 = 16 * rb + 4 * g;
#pragma unroll
    for (int m = 0; m < NK1; m++) ld4(L + (int64_t)(16 * m + i) * ld + r0, a[q][m]);
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int r = r0 + s;
      b[q][s] = tb[(int64_t)(r < n1 ? r : n1 - 1) * n2 + colc];
    }
#pragma unroll
    for (int m = 0; m < NK1; m++) ld4(LT + (int64_t)(16 * rb + i) * KTP + 16 * m + 4 * g, at[q][m]);
  }
#pragma unroll
  for (int m = 0; m < NK1; m++) ld4(wts + i1.z + (int64_t)col * KTP + 16 * m + 4 * g, wk[m]);
  V acc[NK1];
#pragma unroll
  for (int m = 0; m < NK1; m++) acc[m] = (V){0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < RBW; q++)
    if (w + 4 * q < nrb)  // wave-uniform
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int m = 0; m < NK1; m++) acc[m] = M::mma(a[q][m][s], b[q][s], acc[m]);
#pragma unroll
  for (int m = 0; m < NK1; m++) red[w][m][lane] = acc[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NK1; m++) {
    acc[m] = red[0][m][lane];
#pragma unroll
    for (int q = 1; q < 4; q++) acc[m] += red[q][m][lane];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) acc[m][rr] *= wk[m][rr];
  }
#pragma unroll
  for (int q = 0; q < RBW; q++) {
    const int mt = w + 4 * q;
    if (mt >= nrb) break;  // wave-uniform
    V o[2] = {(V){0, 0, 0, 0}, (V){0, 0, 0, 0}};
#pragma unroll
    for (int m = 0; m < NK1; m++)
#pragma unroll
      for (int rr = 0; rr < 4; rr++) o[rr & 1] = M::mma(at[q][m][rr], acc[m][rr], o[rr & 1]);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const int r = 16 * mt + M::row(lane, rr);
      if (r < nn && col < k2l) tb[(int64_t)r * n2 + col] = o[0][rr] + o[1][rr];
    }
  }
}

template <int NK1, int RBW>
static hipError_t launch_rad_mid_f(Ctx* c) {
  hipLaunchKernelGGL((k_rad_mid_f<NK1, RBW>), dim3((unsigned)c->mid_nitems), dim3(256), 0,
                     c->stream, (float*)c->d_t1, (const float*)c->d_lowT,
                     (const float*)c->d_mid_wts, (const int4*)c->d_mid_items2, c->t.n1, c->t.n2);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_rad_mid(Ctx* c) {
  if constexpr (sizeof(T) == 4) {
    if (c->d_mid_items2) {
#define MW3_MIDF(NK, RB) \
  if (c->mid_nk1 == NK && c->mid_rbw == RB) return launch_rad_mid_f<NK, RB>(c);
      MW3_MIDF(1, 1) MW3_MIDF(1, 2) MW3_MIDF(1, 4)
      MW3_MIDF(2, 1) MW3_MIDF(2, 2) MW3_MIDF(2, 4)
      MW3_MIDF(3, 1) MW3_MIDF(3, 2) MW3_MIDF(3, 4)
      MW3_MIDF(4, 1) MW3_MIDF(4, 2) MW3_MIDF(4, 4)
#undef MW3_MIDF
      return hipErrorInvalidValue;
    }
  }
  const HostTables& t = c->t;
  const unsigned nblk = (unsigned)c->mid_nitems;  // one block of 4 waves per item
  const T* tab = (const T*)c->d_dst_tables;
  const T* tabT = (const T*)c->d_lowT;
  const int4* items = (const int4*)c->d_mid_items;
  const int4* sizes = (const int4*)c->d_mid_sizes;
  switch (c->mid_nk1) {
#define MW3_MID(NK)                                                                               \
  case NK:                                                                                        \
    hipLaunchKernelGGL((k_rad_mid<T, NK>), dim3(nblk), dim3(256), 0, c->stream, (T*)c->d_t1, tab, \
                       tabT, items, c->mid_nitems, sizes, c->d_p2r1, c->d_p2r2, c->d_divi,        \
                       t.radcut, t.na, t.n1, t.n2);                                              \
    break;
    MW3_MID(1)
    MW3_MID(2)
    MW3_MID(3)
    MW3_MID(4)
#undef MW3_MID
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// items, size records and transposed low tables of k_rad_mid (r sizes only)
static hipError_t build_rad_mid(Ctx* c) {
  const HostTables& t = c->t;
  c->mid_nitems = 0;
  int k1max = 0;
  for (int ia = 0; ia < t.na; ia++) k1max = std::max(k1max, t.k1lo[ia]);
  const int nk1 = (k1max + 15) / 16;
  if (nk1 < 1 || nk1 > 4 || t.n1 >= 65536 || t.n2 >= 65536 * 16) return hipSuccess;
  const int ktp = 16 * nk1;
  const bool f64 = c->precision != MW3_MIXED;
  std::vector<double> lt;
  std::vector<int32_t> sizes;
  std::vector<int> qrec(t.dst_sizes.size(), -1);
  for (int ia = 0; ia < t.na; ia++) {
    int q = 0;
    while (t.dst_sizes[q] != t.nn1[ia]) q++;
    if (qrec[q] >= 0) continue;
    const int n = t.nn1[ia], ld = t.low_ld[q], rows = (n + 15) / 16 * 16;
    int kcap = 0;
    for (int ja = 0; ja < t.na; ja++)
      if (t.nn1[ja] == n) kcap = std::max(kcap, t.k1lo[ja]);
    qrec[q] = (int)sizes.size() / 4;
    // forward table Lp[k1][r], 16 nk1 rows of ldp = 16 ceil(n1/16) columns,
    // zero at k1 >= k1lo of the size and at r >= nn1
    const int ldp = (t.n1 + 15) / 16 * 16;
    sizes.push_back((int32_t)lt.size());
    sizes.push_back(ldp);
    for (int k1 = 0; k1 < ktp; k1++)
      for (int r = 0; r < ldp; r++)
        lt.push_back(r < n && k1 < kcap ? t.dst_tables[(size_t)t.low_off[q] + (size_t)k1 * ld + r]
                                        : 0.0);
    sizes.push_back((int32_t)lt.size());
    sizes.push_back(0);
    for (int r = 0; r < rows; r++)
      for (int p = 0; p < ktp; p++) {
        // position p holds k1 = p (FP32) or the 4 x 4 transpose within its 16-block (FP64)
        const int k1 = f64 ? (p & ~15) + 4 * (p & 3) + ((p >> 2) & 3) : p;
        double v = 0.0;
        if (r < n && k1 < kcap) v = t.dst_tables[(size_t)t.low_off[q] + (size_t)k1 * ld + r];
        lt.push_back(v);
      }
  }
  if (lt.size() >= (size_t)1 << 31 || t.dst_tables.size() >= (size_t)1 << 31) return hipSuccess;
  // items in table order (size, angle, channel, column block): the kernel
  // runs consecutive items on one XCD
  std::vector<int32_t> items;
  for (size_t q = 0; q < t.dst_sizes.size(); q++)
    for (int ia = 0; ia < t.na; ia++) {
      if (t.dst_sizes[q] != t.nn1[ia]) continue;
      for (int ch = 0; ch < t.nc; ch++)
        for (int cc = 0; 16 * cc < t.k2lo[ia]; cc++) {
          items.push_back(ch * t.na + ia);
          items.push_back(cc | (t.k2lo[ia] << 16));
          items.push_back(t.nn1[ia] | (t.k1lo[ia] << 16));
          items.push_back(qrec[q]);
        }
    }
  if (items.empty()) return hipSuccess;
  hipError_t e;
  auto up = [&](void** d, const void* h, size_t bytes) {
    hipError_t r = hipMalloc(d, bytes);
    if (r != hipSuccess) return r;
    return hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice);
  };
  if (f64) {
    e = up(&c->d_lowT, lt.data(), lt.size() * sizeof(double));
  } else {
    std::vector<float> l32(lt.begin(), lt.end());
    e = up(&c->d_lowT, l32.data(), l32.size() * sizeof(float));
  }
  if (e != hipSuccess) return e;
  if ((e = up((void**)&c->d_mid_items, items.data(), items.size() * 4)) != hipSuccess) return e;
  if ((e = up((void**)&c->d_mid_sizes, sizes.data(), sizes.size() * 4)) != hipSuccess) return e;
  c->mid_nitems = (int)(items.size() / 4);
  c->mid_nk1 = nk1;
  // FP32: k_rad_mid_f's packed items and weight table W[angle][k2][k1] =
  // (radcut - p2r1(k1) - p2r2(k2))_+ divi, rounded to FP32 (k1 clamped to
  // n1 - 1 and k2 to n2 - 1 as k_rad_mid does), when every wave's r blocks
  // fit 4 per wave
  const int nrbmax = (t.n1 + 15) / 16;
  const int rbw = nrbmax <= 4 ? 1 : nrbmax <= 8 ? 2 : nrbmax <= 16 ? 4 : 0;
  if (!f64 && rbw) {
    int ncb = 0;
    for (int ia = 0; ia < t.na; ia++) ncb = std::max(ncb, (t.k2lo[ia] + 15) / 16);
    const int ncol = 16 * ncb;
    std::vector<float> wts((size_t)t.na * ncol * ktp);
    for (int ia = 0; ia < t.na; ia++)
      for (int k2 = 0; k2 < ncol; k2++)
        for (int k1 = 0; k1 < ktp; k1++) {
          const double x = t.radcut - (t.p2r1[(size_t)ia * t.n1 + std::min(k1, t.n1 - 1)] +
                                       t.p2r2[std::min(k2, t.n2 - 1)]);
          wts[((size_t)ia * ncol + k2) * ktp + k1] = (float)(x > 0.0 ? x * t.divi[ia] : 0.0);
        }
    std::vector<int32_t> it2;
    const size_t ni = items.size() / 4;
    for (size_t k = 0; k < ni; k++) {
      const int32_t* r = &items[4 * k];
      const int32_t* z = &sizes[4 * (size_t)r[3]];
      const int ia = r[0] % t.na;
      it2.insert(it2.end(), {r[0], r[1], r[2], z[0], z[1], z[2], (int32_t)((size_t)ia * ncol * ktp), 0});
    }
    if (wts.size() >= (size_t)1 << 31) return hipSuccess;
    if ((e = up((void**)&c->d_mid_wts, wts.data(), wts.size() * sizeof(float))) != hipSuccess) return e;
    if ((e = up((void**)&c->d_mid_items2, it2.data(), it2.size() * 4)) != hipSuccess) return e;
    c->mid_rbw = rbw;
  }
  return hipSuccess;
}

template <typename K>
static hipError_t persistent_grid(K kernel, int threads, int* blocks) {
  int nb = 0, dev = 0, ncu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  *blocks = (nb > 0 ? nb : 1) * ncu;
  return hipSuccess;
}

// launch (or, with setup, size the grids of) the R passes for NKB, HB
template <typename T, int NKB, int HB>
static hipError_t rad_nkb_hb(Ctx* c, const T* psi, T* t1, T* corr, int mode) {
  const HostTables& t = c->t;
  const T* tab = (const T*)c->d_dst_tables;
  const int4* items = (const int4*)c->d_rad_items;
  if (mode == 2) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->rad_ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    // FP64 passes with the table split over block pairs (whole 16-block
    // groups; the HB R tiles of the first half split in two)
    c->rad_lds64 = sizeof(T) == 8 && c->rad_ncu >= 16 && HB % 2 == 0;
    c->rad_inv_h = sizeof(T) == 4;
    e = t.n1 % 16 == 0
                       ? persistent_grid(k_rad_fwd<T, NKB, HB, true>, 128 * NKB * RadSplit<T, HB>::v,
                                         &c->rad_grid[0])
                       : persistent_grid(k_rad_fwd<T, NKB, HB, false>, 128 * NKB * RadSplit<T, HB>::v,
                                         &c->rad_grid[0]);
    if (e != hipSuccess) return e;
    return t.n1 % 16 == 0 ? persistent_grid(k_rad_inv<T, NKB, HB, true>, 32 * HB, &c->rad_grid[1])
                          : persistent_grid(k_rad_inv<T, NKB, HB, false>, 32 * HB, &c->rad_grid[1]);
  }
  const int n = c->rad_nitems;
  // the grid: the resident blocks divided by the grid option (each block
  // loads its waves' table slices once, so fewer blocks read fewer table
  // bytes and walk more items each)
  const int gdiv = c->opt_rad_grid > 0 ? c->opt_rad_grid : 1;
  const int gmin = (n + kRadMaxItems - 1) / kRadMaxItems;  // each block's list fits its LDS copy
  const int g0 = std::max(std::max(c->rad_grid[0] / gdiv, 1), gmin);
  const int g1 = std::max(std::max(c->rad_grid[1] / gdiv, 1), gmin);
  if (mode == 0) {
    if constexpr (sizeof(T) == 4) {  // the table in LDS, one block per CU
      const int nb = std::max(1, std::min(c->rad_ncu, n));
      if (t.n1 % 16 == 0)
        hipLaunchKernelGGL((k_rad_fwd_lds<NKB, HB, true>), dim3(nb), dim3(64 * kFwdLdsWaves), 0,
                           c->stream, psi, tab + t.rad_tt_off, t1, items, n, t.n1);
      else
        hipLaunchKernelGGL((k_rad_fwd_lds<NKB, HB, false>), dim3(nb), dim3(64 * kFwdLdsWaves), 0,
                           c->stream, psi, tab + t.rad_tt_off, t1, items, n, t.n1);
      return hipGetLastError();
    }
    if constexpr (sizeof(T) == 8) {
      if (c->rad_lds64) {  // pairs of blocks split the table (16-block groups)
        const int nb = c->rad_ncu / 16 * 16;
        if (t.n1 % 16 == 0)
          hipLaunchKernelGGL((k_rad_fwd_lds64<NKB, HB, true>), dim3(nb), dim3(64 * kRadLds64Waves),
                             0, c->stream, psi, tab + t.rad_tt_off, t1, items, n, t.n1);
        else
          hipLaunchKernelGGL((k_rad_fwd_lds64<NKB, HB, false>), dim3(nb), dim3(64 * kRadLds64Waves),
                             0, c->stream, psi, tab + t.rad_tt_off, t1, items, n, t.n1);
        return hipGetLastError();
      }
    }
    const int nt = 128 * NKB * RadSplit<T, HB>::v;
    if (t.n1 % 16 == 0)
      hipLaunchKernelGGL((k_rad_fwd<T, NKB, HB, true>), dim3(n < g0 ? n : g0), dim3(nt), 0,
                         c->stream, psi, tab + t.rad_tt_off, t1, items, n, t.n1);
    else
      hipLaunchKernelGGL((k_rad_fwd<T, NKB, HB, false>), dim3(n < g0 ? n : g0), dim3(nt), 0,
                         c->stream, psi, tab + t.rad_tt_off, t1, items, n, t.n1);
  } else {
    if constexpr (sizeof(T) == 4) {  // the table in LDS, one block per CU
      const int nb = std::max(1, std::min(c->rad_ncu, n));
      if (c->rad_inv_h && c->opt_fp16) {  // products on the FP16 matrix cores, compensated
        if (t.n1 % 16 == 0)
          hipLaunchKernelGGL((k_rad_inv_h<NKB, HB, true>), dim3(nb), dim3(64 * kInvLdsWaves), 0,
                             c->stream, t1, tab + t.rad_ti_off, corr, items, n, t.n1);
        else
          hipLaunchKernelGGL((k_rad_inv_h<NKB, HB, false>), dim3(nb), dim3(64 * kInvLdsWaves), 0,
                             c->stream, t1, tab + t.rad_ti_off, corr, items, n, t.n1);
        return hipGetLastError();
      }
      if (t.n1 % 16 == 0)
        hipLaunchKernelGGL((k_rad_inv_lds<NKB, HB, true>), dim3(nb), dim3(64 * kInvLdsWaves), 0,
                           c->stream, t1, tab + t.rad_ti_off, corr, items, n, t.n1);
      else
        hipLaunchKernelGGL((k_rad_inv_lds<NKB, HB, false>), dim3(nb), dim3(64 * kInvLdsWaves), 0,
                           c->stream, t1, tab + t.rad_ti_off, corr, items, n, t.n1);
      return hipGetLastError();
    }
    if constexpr (sizeof(T) == 8) {
      if (c->rad_lds64) {
        const int nb = c->rad_ncu / 16 * 16;
        if (t.n1 % 16 == 0)
          hipLaunchKernelGGL((k_rad_inv_lds64<NKB, HB, true>), dim3(nb), dim3(64 * kRadLds64Waves),
                             0, c->stream, t1, tab + t.rad_ti_off, corr, items, n, t.n1);
        else
          hipLaunchKernelGGL((k_rad_inv_lds64<NKB, HB, false>), dim3(nb), dim3(64 * kRadLds64Waves),
                             0, c->stream, t1, tab + t.rad_ti_off, corr, items, n, t.n1);
        return hipGetLastError();
      }
    }
    if (t.n1 % 16 == 0)
      hipLaunchKernelGGL((k_rad_inv<T, NKB, HB, true>), dim3(n < g1 ? n : g1), dim3(32 * HB), 0,
                         c->stream, t1, tab + t.rad_ti_off, corr, items, n, t.n1);
    else
      hipLaunchKernelGGL((k_rad_inv<T, NKB, HB, false>), dim3(n < g1 ? n : g1), dim3(32 * HB), 0,
                         c->stream, t1, tab + t.rad_ti_off, corr, items, n, t.n1);
  }
  return hipGetLastError();
}

template <typename T, int NKB>
static hipError_t rad_nkb(Ctx* c, const T* psi, T* t1, T* corr, int mode) {
  switch (c->t.n2) {
    case 256: return rad_nkb_hb<T, NKB, 8>(c, psi, t1, corr, mode);
    case 320: return rad_nkb_hb<T, NKB, 10>(c, psi, t1, corr, mode);
    case 384: return rad_nkb_hb<T, NKB, 12>(c, psi, t1, corr, mode);
    case 512: return rad_nkb_hb<T, NKB, 16>(c, psi, t1, corr, mode);
    default: return hipErrorInvalidValue;
  }
}

// mode 0: forward R pass (psi -> t1); 1: inverse (t1 -> corr); 2: size the
// persistent grids
template <typename T>
static hipError_t run_rad(Ctx* c, const T* psi, T* t1, T* corr, int mode) {
  switch (c->t.rad_nkb) {
    case 1: return rad_nkb<T, 1>(c, psi, t1, corr, mode);
    case 2: return rad_nkb<T, 2>(c, psi, t1, corr, mode);
    case 3: return rad_nkb<T, 3>(c, psi, t1, corr, mode);
    case 4: return rad_nkb<T, 4>(c, psi, t1, corr, mode);
    default: return hipErrorInvalidValue;
  }
}


// block order of each plan; every vectorised plan runs the 64 x 64 block tile
// (2 x 2 waves of 2 x 2 MFMA tiles), which beat 128 x 64, 64 x 128 and
// 128 x 128 on every pass
static void plan_tiles(Ctx* c) {
  // R pass (1): the two parities of one (channel, angle) fold the same rows
  c->p_dstR.group = 2;
  c->p_dstR.share_b = 0;
  c->p_dstRi.share_b = 0;
  c->p_dstr.share_b = c->p_dstr2.share_b = c->p_angF.share_b = c->p_angB.share_b = 1;
}

static GemmDesc gd() {
  GemmDesc g{};
  g.kin_ia = -1;
  return g;
}

hipError_t build_plans(Ctx* c) {
  const HostTables& t = c->t;
  const int64_t plane = t.nplane, slab = t.nslab;
  hipError_t e;
  auto tab_q = [&](int n) {
    int q = 0;
    while (t.dst_sizes[q] != n) q++;
    return q;
  };
  // Radial kinetic as c psi - S (c - E)_+ divi S psi on the low modes
  // k1 < k1lo(ia), k2 < k2lo(ia) (see the low-mode tables in mw3_setup.cpp).
  // (1) R rows, folded: t1[r][k2] = sum_R S(R, k2) psi[r][R], k2 < k2lo, r < nn1
  {
    const int n = t.n2, h = n / 2, ke = n - h;
    const int q = tab_q(n), kp = t.dst_ld[q];
    const int64_t toff = t.dst_table_off[q];
    std::vector<GemmDesc> d;
    for (int ch = 0; ch < t.nc; ch++)
      for (int ia = 0; ia < t.na; ia++)
        for (int par = 0; par < 2; par++) {
          const int k2 = t.k2lo[ia];
          GemmDesc g = gd();
          g.a_off = (int64_t)ch * slab + (int64_t)ia * plane;
          g.b_off = toff + (par ? (int64_t)kp * kp : 0);
          g.c_off = g.a_off + par;
          g.M = t.nn1[ia];
          g.N = par ? k2 / 2 : (k2 + 1) / 2;  // modes 2k'+par < k2lo
          g.K = par ? h : ke;
          g.lda_m = n;
          g.lda_k = 1;
          g.fold_op = 1;
          g.fold_n = n;
          g.fold_sgn = par ? -1 : 1;
          d.push_back(g);
        }
    c->p_dstR.st = GemmStrides{kp, 1, n, 2};
    if ((e = upload_plan(&c->p_dstR, d, true)) != hipSuccess) return e;
  }
  // (2) r columns, folded, times (c - E)_+ divi:
  //     t2[k1][k2] = f(k1, k2) sum_r S(r, k1) t1[r][k2], k1 < k1lo
  {
    std::vector<GemmDesc> d;
    for (int ch = 0; ch < t.nc; ch++)
      for (int ia = 0; ia < t.na; ia++) {
        const int n = t.nn1[ia], h = n / 2, ke = n - h;
        const int q = tab_q(n), kp = t.dst_ld[q];
        const int64_t toff = t.dst_table_off[q];
        const int k1 = t.k1lo[ia];
        for (int par = 0; par < 2; par++) {
          GemmDesc g = gd();
          // A[k1'][j1] = S[2k1'+par][j1] = S_par[j1][k1']
          g.a_off = toff + (par ? (int64_t)kp * kp : 0);
          g.lda_m = 1;
          g.lda_k = kp;
          g.b_off = (int64_t)ch * slab + (int64_t)ia * plane;
          g.c_off = g.b_off + (int64_t)par * t.n2;
          g.M = par ? k1 / 2 : (k1 + 1) / 2;
          g.N = t.k2lo[ia];
          g.K = par ? h : ke;
          g.fold_op = 2;
          g.fold_n = n;
          g.fold_sgn = par ? -1 : 1;
          g.kin_ia = ia;
          g.rmul = 2;
          g.radd = par;
          g.kin_mode = 1;
          d.push_back(g);
        }
      }
    c->p_dstr.st = GemmStrides{t.n2, 1, 2 * (int64_t)t.n2, 1};
    if ((e = upload_plan(&c->p_dstr, d, false)) != hipSuccess) return e;
  }
  // (3) back over r: t1[r][k2] = sum_{k1 < k1lo} S(r, k1) t2[k1][k2], r < nn1
  {
    std::vector<GemmDesc> d;
    for (int ch = 0; ch < t.nc; ch++)
      for (int ia = 0; ia < t.na; ia++) {
        const int q = tab_q(t.nn1[ia]);
        GemmDesc g = gd();
        g.a_off = t.low_off[q];  // A[r][k1] = L[k1][r]
        g.lda_m = 1;
        g.lda_k = t.low_ld[q];
        g.b_off = (int64_t)ch * slab + (int64_t)ia * plane;
        g.c_off = g.b_off;
        g.M = t.nn1[ia];
        g.N = t.k2lo[ia];
        g.K = t.k1lo[ia];
        d.push_back(g);
      }
    c->p_dstr2.st = GemmStrides{t.n2, 1, t.n2, 1};
    if ((e = upload_plan(&c->p_dstr2, d, false)) != hipSuccess) return e;
  }
  // (4) back over R: t2[r][R] = sum_{k2 < k2lo} t1[r][k2] S(k2, R) (the
  //     correction subtracted from c psi by the update)
  {
    const int q = tab_q(t.n2);
    std::vector<GemmDesc> d;
    for (int ch = 0; ch < t.nc; ch++)
      for (int ia = 0; ia < t.na; ia++) {
        GemmDesc g = gd();
        g.a_off = (int64_t)ch * slab + (int64_t)ia * plane;
        g.lda_m = t.n2;
        g.lda_k = 1;
        g.b_off = t.low_off[q];
        g.c_off = g.a_off;
        g.M = t.nn1[ia];
        g.N = t.n2;
        g.K = t.k2lo[ia];
        d.push_back(g);
      }
    c->p_dstRi.st = GemmStrides{t.low_ld[q], 1, t.n2, 1};
    if ((e = upload_plan(&c->p_dstRi, d, true)) != hipSuccess) return e;
  }
  // angular forward: fbr[c][jj][p] = sum_ia D_c(ia, jj) psi[c][ia][p]; with no
  // Coriolis neighbour in any channel the coefficients go into its epilogue
  c->ang_fused = true;
  for (int ch = 0; ch < t.nc; ch++)
    if (t.nbr[ch * 3 + 0] >= 0 || t.nbr[ch * 3 + 2] >= 0) c->ang_fused = false;
  {
    std::vector<GemmDesc> d;
    for (int ch = 0; ch < t.nc; ch++) {
      GemmDesc g = gd();
      g.a_off = (int64_t)ch * t.na * t.nJp;
      g.b_off = (int64_t)ch * slab;
      g.c_off = (int64_t)ch * t.nJ * plane;
      g.M = t.nJ;
      g.N = (int)plane;
      g.K = t.na;
      g.lda_m = 1;
      g.lda_k = t.nJp;
      if (c->ang_fused) {  // write phi = (a_j + b_j) D^T psi directly
        g.kin_ia = ch;
        g.kin_mode = 2;
      }
      d.push_back(g);
    }
    c->p_angF.st = GemmStrides{plane, 1, plane, 1};
    if ((e = upload_plan(&c->p_angF, d, false)) != hipSuccess) return e;
  }
  // angular back: ang[c][ia][p] = sum_jj D_c(ia, jj) phi[c][jj][p]
  {
    std::vector<GemmDesc> d;
    for (int ch = 0; ch < t.nc; ch++) {
      GemmDesc g = gd();
      g.a_off = (int64_t)ch * t.na * t.nJp;
      g.b_off = (int64_t)ch * t.nJ * plane;
      g.c_off = (int64_t)ch * slab;
      g.M = t.na;
      g.N = (int)plane;
      g.K = t.nJ;
      g.lda_m = t.nJp;
      g.lda_k = 1;
      d.push_back(g);
    }
    c->p_angB.st = GemmStrides{plane, 1, plane, 1};
    if ((e = upload_plan(&c->p_angB, d, true)) != hipSuccess) return e;
  }
  plan_tiles(c);
  // the two r passes in one kernel where the low r modes fit its tiles
  if ((e = build_rad_mid(c)) != hipSuccess) return e;
  // R passes one wave per 16-row tile when their tables were built
  c->rad_nitems = 0;
  if (t.rad_nkb > 0 && (t.n2 == 256 || t.n2 == 320 || t.n2 == 384 || t.n2 == 512)) {
    // items {channel * na + angle, row tile, nn1, k2lo | active R-tile pairs << 16}
    std::vector<int32_t> items;
    for (int ch = 0; ch < t.nc; ch++)
      for (int ia = 0; ia < t.na; ia++)
        for (int rt = 0; 16 * rt < t.nn1[ia]; rt++) {
          // bit 16 + q: R-tile pair q (columns 32 q .. 32 q + 31 and their
          // mirrors) holds a compressed-grid point in the item's rows
          int pm = 0;
          const int rhi = std::min(16 * rt + 16, t.nn1[ia]);
          for (int q = 0; q < t.n2 / 64; q++)
            for (int r = 16 * rt; r < rhi && !((pm >> q) & 1); r++)
              for (int k = 0; k < 32; k++) {
                const uint8_t* mrow = &t.mask[((size_t)ia * t.n1 + r) * t.n2];
                if (mrow[32 * q + k] || mrow[t.n2 - 1 - 32 * q - k]) {
                  pm |= 1 << q;
                  break;
                }
              }
          items.push_back(ch * t.na + ia);
          items.push_back(rt);
          items.push_back(t.nn1[ia]);
          items.push_back(t.k2lo[ia] | (pm << 16));
        }
    if ((e = hipMalloc(&c->d_rad_items, items.size() * sizeof(int32_t))) != hipSuccess) return e;
    if ((e = hipMemcpy(c->d_rad_items, items.data(), items.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice)) != hipSuccess)
      return e;
    c->rad_nitems = (int)(items.size() / 4);
    e = c->precision == MW3_MIXED ? run_rad<float>(c, nullptr, nullptr, nullptr, 2)
                                  : run_rad<double>(c, nullptr, nullptr, nullptr, 2);
    if (e != hipSuccess) return e;
  }
  // one-pass angular kernel when both products fit its register tiles
  c->ang_nt = 0;
  const int nt = ((t.na > t.nJ ? t.na : t.nJ) + 15) / 16;
  const size_t es = c->precision == MW3_MIXED ? sizeof(float) : sizeof(double);
  const bool offs32 = (double)t.na * (double)plane * (double)es < 2147483648.0 &&
                      (double)t.nJ * (double)(t.n1 > t.n2 ? t.n1 : t.n2) * 8.0 < 2147483648.0;
  if (c->ang_fused && nt <= 8 && offs32) {
    e = c->precision == MW3_MIXED ? ang_fused_setup<float>(c, nt) : ang_fused_setup<double>(c, nt);
    if (e != hipSuccess) return e;
    c->ang_nt = nt;
  }
  // all-channel Coriolis coefficient kernel: up to 8 channels, and the
  // q = 0 / 2 coefficient rows zero wherever a channel has no such neighbour
  // (the kernel adds them against a zero stand-in)
  c->fbr_all = t.nc >= 2 && t.nc <= 8;
  for (int ch = 0; ch < t.nc && c->fbr_all; ch++)
    for (int q = 0; q < 3; q += 2) {
      if (t.nbr[ch * 3 + q] >= 0) continue;
      const double* row = &t.bcoef[(((size_t)ch * 3 + q) * t.nJ) * t.n2];
      for (size_t k = 0; k < (size_t)t.nJ * t.n2; k++)
        if (row[k] != 0.0) {
          c->fbr_all = false;
          break;
        }
    }
  // tiles of 16 ANG_WAVES plane columns with a point of the compressed grid at
  // some angle (the mask is the same for every channel)
  const bool tiles_fit = (plane + 16 * ANG_WAVES - 1) / (16 * ANG_WAVES) * t.nc < (1 << 24);
  if (!tiles_fit) c->ang_nt = 0;  // tile entries keep the index in 24 bits
  // (also for the folded parts beside the Coriolis pass: mixed precision)
  const bool cor_try = !c->ang_fused && c->precision == MW3_MIXED && offs32;
  if (tiles_fit && (c->ang_nt || (c->ang_fused && offs32) || cor_try)) {
    const int64_t tw = 16 * ANG_WAVES, tpc = (plane + tw - 1) / tw;
    std::vector<uint8_t> act((size_t)tpc, 0);  // bit w: wave w's 16 columns
    for (int ia = 0; ia < t.na; ia++)
      for (int64_t q = 0; q < plane; q++)
        if (t.mask[(size_t)ia * plane + q])
          act[(size_t)(q / tw)] |= (uint8_t)(1u << ((q % tw) / 16));
    std::vector<int32_t> tl, gl;
    for (int ch = 0; ch < t.nc; ch++)
      for (int64_t k = 0; k < tpc; k++)
        if (act[(size_t)k])
          tl.push_back((int32_t)((uint32_t)(ch * tpc + k) | ((uint32_t)act[(size_t)k] << 24)));
    // the same columns as 16-column groups (group q: columns 16 q .. 16 q + 15);
    // in FP32 a group is half a 128-byte line, so both halves of a line with
    // a point are listed, side by side (2.7% more groups at configuration 2,
    // half the psi reads): the spare half reads zeros and writes zeros that
    // nothing reads
    const int gstep = c->precision == MW3_MIXED ? 2 : 1;
    for (int64_t k = 0; k < tpc; k++)
      for (int w = 0; w < ANG_WAVES; w += gstep) {
        const unsigned bits = (act[(size_t)k] >> w) & (gstep == 2 ? 3u : 1u);
        if (!bits) continue;
        for (int u = 0; u < gstep; u++) gl.push_back((int32_t)(k * ANG_WAVES + w + u));
      }
    c->ang_ntiles = (int)tl.size();
    c->ang_ngroups = (int)gl.size();
    if (!tl.empty()) {
      if ((e = hipMalloc(&c->d_ang_tiles, tl.size() * sizeof(int32_t))) != hipSuccess) return e;
      if ((e = hipMemcpy(c->d_ang_tiles, tl.data(), tl.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice)) != hipSuccess)
        return e;
      if ((e = hipMalloc(&c->d_ang_groups, gl.size() * sizeof(int32_t))) != hipSuccess) return e;
      if ((e = hipMemcpy(c->d_ang_groups, gl.data(), gl.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    }
  }
  // the parity-folded form of the one-pass kernel where the functions allow it
  c->ang_fold = 0;
  if (c->ang_fused && offs32 && tiles_fit) {
    int ntf = 0;
    if ((e = build_ang_fold(c, &ntf)) != hipSuccess) return e;
    if (ntf > 0) {
      e = c->precision == MW3_MIXED ? ang_fold_setup<float>(c, ntf) : ang_fold_setup<double>(c, ntf);
      if (e != hipSuccess) return e;
      c->ang_fold = ntf;
    }
  }
  // with Coriolis coupling (mixed precision): the folded forward part into
  // the FBR slab and the folded back part from the coefficient pass's output,
  // the coefficient pass between them as in the GEMM variant
  c->ang_fold_cor = 0;
  if (cor_try && tiles_fit && c->ang_ngroups > 0) {
    int ntf = 0;
    if ((e = build_ang_fold(c, &ntf)) != hipSuccess) return e;
    if (ntf > 0) {
      if ((e = ang_cor_setup(c, ntf)) != hipSuccess) return e;
      c->ang_fold_cor = ntf;
    }
  }
  return hipSuccess;
}

void free_plans(Ctx* c) {

  GemmPlan* ps[7] = {&c->p_dstR, &c->p_dstr, &c->p_dstr2, &c->p_dstRi, &c->p_angF, &c->p_angB,
                     &c->p_prod};
  for (auto* p : ps)
    if (p->d_descs) {
      (void)hipFree(p->d_descs);
      p->d_descs = nullptr;
    }
}

// H psi terms into work slabs: d_t2 = T psi (rows r < nn1), d_ang = angular
// psi in the operator precision T: the FP32 copy in mixed mode
template <typename T>
static const T* op_psi(Ctx* c, const double* psi);
template <>
const double* op_psi<double>(Ctx*, const double* psi) {
  return psi;
}
template <>
const float* op_psi<float>(Ctx* c, const double* psi) {
  return psi_shadow(c, psi);
}

template <typename T>
static hipError_t hpsi_impl(Ctx* c, const double* psi64) {
  const T* psi = op_psi<T>(c, psi64);
  if (!psi) return hipErrorInvalidValue;
  const HostTables& t = c->t;
  hipStream_t s = c->stream;
  T* t1 = (T*)c->d_t1;
  T* t2 = (T*)c->d_t2;
  const T* tab = (const T*)c->d_dst_tables;
  hipError_t e;
  // mixed precision: the angular term on the third stream, beside the radial
  // passes (it reads psi, written by the previous update on the main stream,
  // and writes only its own slabs, which the next update reads after ev_ang);
  // measured 3.5% faster per step than one stream at configuration 2, while
  // the FP64 pair ran 2% slower side by side (its angular kernel holds a
  // whole CU's LDS)
  hipStream_t sa = c->opt_side ? c->side : s;
  // kernel variants of this context's options (the autotuner's choices)
  const bool use_rad = c->rad_nitems && c->opt_radial;
  const bool use_mid = c->mid_nitems && c->opt_radial;
  const bool use_fold = c->ang_fold && c->opt_angular == 0;
  const bool use_one = c->ang_nt && c->opt_angular <= 1 && !use_fold;
  // folded parts beside the Coriolis pass (mixed precision, FP16 products)
  const bool use_cor = sizeof(T) == 4 && c->ang_fold_cor && c->opt_angular == 0 && c->opt_fp16 &&
                       !use_fold && !use_one && !c->ang_fused;
  // (its back part forms the coefficients and the Omega +- 1 terms itself:
  // no coefficient pass)
  const bool use_cor3 = use_cor;
  // the forward R pass first (measured 1.3% faster per FP64 step than after
  // the angular term)
  timer_begin(c, K_DST_R);
  e = use_rad ? run_rad<T>(c, psi, t1, t2, 0)
              : run_gemm<G_DST_R, T, T, T, T>(s, psi, tab, t1, c->p_dstR);
  if (e != hipSuccess) return e;
  timer_end(c, K_DST_R);
  if (sa != s) {
    if (c->slice_fresh) {  // recorded after the update that wrote psi, nothing writing it since
      if ((e = hipStreamWaitEvent(sa, c->ev_slice, 0)) != hipSuccess) return e;
    } else {
      if ((e = hipEventRecord(c->ev_psi, s)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(sa, c->ev_psi, 0)) != hipSuccess) return e;
    }
  }
  c->slice_fresh = false;
  c->cvj_joined = false;
  timer_begin(c, K_ANG_FWD, sa);
  if (use_fold) {
    if ((e = launch_ang_fold<T>(c, psi, sa)) != hipSuccess) return e;
    timer_end(c, K_ANG_FWD, sa);
  } else if (use_one) {
    if ((e = launch_ang_fused<T>(c, psi, sa)) != hipSuccess) return e;
    timer_end(c, K_ANG_FWD, sa);
  } else if (c->ang_fused) {
    KinParams ka{};
    ka.n1 = t.n1;
    ka.acoef = (const double*)c->d_acoef;
    ka.bcoef = (const double*)c->d_bcoef;
    ka.nJ = t.nJ;
    ka.n2 = t.n2;
    e = run_gemm<G_ANG_FWD, T, T, T, T>(sa, (const T*)c->d_Dmat, psi, (T*)c->d_phi,
                                        c->p_angF, ka);
    if (e != hipSuccess) return e;
    timer_end(c, K_ANG_FWD, sa);
  } else {
    if constexpr (sizeof(T) == 4) {
      if (use_cor) {
        if ((e = launch_ang_cor<1>(c, psi, sa)) != hipSuccess) return e;
      }
    }
    if (!use_cor) {
      e = run_gemm<G_ANG_FWD, T, T, T, T>(sa, (const T*)c->d_Dmat, psi, (T*)c->d_fbr,
                                          c->p_angF);
      if (e != hipSuccess) return e;
    }
    timer_end(c, K_ANG_FWD, sa);
  }
  if (use_cor3) {
    if constexpr (sizeof(T) == 4) {
      timer_begin(c, K_ANG_BWD, sa);
      // (the psi argument is not read by part 3)
      if ((e = launch_ang_cor<3>(c, (const float*)c->d_phi, sa)) != hipSuccess) return e;
      timer_end(c, K_ANG_BWD, sa);
    }
  } else if (!use_fold && !use_one && !c->ang_fused) {
    timer_begin(c, K_CORIOLIS, sa);
    const unsigned nall = (unsigned)((int64_t)t.nJ * ((t.n1 + FBR_RB - 1) / FBR_RB));
    switch (c->fbr_all ? t.nc : 0) {
#define MW3_FBR(NC)                                                                              \
  case NC:                                                                                       \
    hipLaunchKernelGGL((k_fbr_coef_all<T, NC>), dim3(nall), dim3(FBR_THREADS), 0, sa,           \
                       (const T*)c->d_fbr, (T*)c->d_phi, (const double*)c->d_acoef,             \
                       (const double*)c->d_bcoef, c->d_nbr, t.nJ, t.n1, t.n2);                  \
    break;
      MW3_FBR(2)
      MW3_FBR(3)
      MW3_FBR(4)
      MW3_FBR(5)
      MW3_FBR(6)
      MW3_FBR(7)
      MW3_FBR(8)
#undef MW3_FBR
      default:
        hipLaunchKernelGGL(k_fbr_coef<T>, dim3((unsigned)((int64_t)t.nc * t.nJ * t.n1)),
                           dim3(256), 0, sa, (const T*)c->d_fbr, (T*)c->d_phi,
                           (const double*)c->d_acoef, (const double*)c->d_bcoef, c->d_nbr, t.nc,
                           t.nJ, t.n1, t.n2);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    timer_end(c, K_CORIOLIS, sa);
  }
  if (!use_one && !use_fold && !use_cor3) {
    timer_begin(c, K_ANG_BWD, sa);
    e = run_gemm<G_ANG_BWD, T, T, T, T>(sa, (const T*)c->d_Dmat, (const T*)c->d_phi,
                                        (T*)c->d_ang, c->p_angB);
    if (e != hipSuccess) return e;
    timer_end(c, K_ANG_BWD, sa);
  }
  if (sa != s) {
    // the second stream's sums of the last step end before ev_ang too, so
    // that the next update, which rewrites what they read, waits once
    if (c->cvj_pending) {
      if ((e = hipStreamWaitEvent(sa, c->ev_cvj, 0)) != hipSuccess) return e;
      c->cvj_joined = true;
    }
    if ((e = hipEventRecord(c->ev_ang, sa)) != hipSuccess) return e;
  }
  // radial kinetic correction on the low modes: folded DST over R, folded DST
  // over r times (c - E)_+ divi, back over r, back over R into d_t2; the
  // update forms T psi = radcut psi - d_t2 on rows r < nn1
  KinParams kp{c->d_nn1, c->d_p2r1, c->d_p2r2, c->d_divi, t.radcut, t.n1};
  timer_begin(c, K_DST_r);
  if (use_mid) {
    if ((e = launch_rad_mid<T>(c)) != hipSuccess) return e;
  } else {
    e = run_gemm<G_DST_r_KIN, T, T, T, T>(s, tab, t1, t2, c->p_dstr, kp);
    if (e != hipSuccess) return e;
    e = run_gemm<G_DST_r_INV, T, T, T, T>(s, tab, t2, t1, c->p_dstr2);
    if (e != hipSuccess) return e;
  }
  timer_end(c, K_DST_r);
  timer_begin(c, K_DST_RINV);
  e = use_rad ? run_rad<T>(c, psi, t1, t2, 1)
              : run_gemm<G_DST_R_INV, T, T, T, T>(s, t1, tab, t2, c->p_dstRi);
  if (e != hipSuccess) return e;
  timer_end(c, K_DST_RINV);
  return sa != s ? hipStreamWaitEvent(s, c->ev_ang, 0) : hipSuccess;
}

__global__ void k_to_f32(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

float* psi_shadow(Ctx* c, const double* psi) {
  if (psi == c->d_psi) return c->d_psi32;
  if (psi == c->d_psim1) return c->d_psim132;
  if (psi == c->d_scratch) return c->d_scr32;
  return nullptr;
}

hipError_t launch_to_f32(Ctx* c, const double* src, float* dst) {
  if (!dst) return hipSuccess;
  const int64_t n = (int64_t)c->t.nc * c->t.nslab;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)),
                     dim3(256), 0, c->stream, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_hpsi(Ctx* c, const double* psi) {
  if (c->precision == MW3_MIXED) return hpsi_impl<float>(c, psi);
  return hpsi_impl<double>(c, psi);
}

template <typename T>
static hipError_t update_impl(Ctx* c, int mode, int slot, const double* psi, bool proj) {
  // psi_new rows that are never written (outside the compressed grid) keep
  // zeros, so the reactant slice must start cleared only once (at create)
  const HostTables& t = c->t;
  const int nb = c->nnorm_blocks;  // work items
  UpdArgs a{};
  a.items = c->d_uw_item;
  a.urow = c->d_uw_row;
  a.uword = c->d_uw_word;
  a.upart = c->d_uw_part;
  a.nitems = nb;
  a.nn1 = c->d_nn1;
  a.absfr1 = c->d_absfr1;
  a.absfr2 = c->d_absfr2;
  a.ch_elec = c->d_ch_elec;
  a.ch_omega = c->d_ch_omega;
  a.ch_cpl = c->d_ch_cpl;
  a.psi0 = (c->s.iphoto >= 1) ? c->d_psi0 : nullptr;
  a.part = mode != 0 ? c->d_norm_part : nullptr;
  a.nc = t.nc;
  a.ne = t.ne;
  a.na = t.na;
  a.n1 = t.n1;
  a.n2 = t.n2;
  a.emindlt = c->s.emindlt;
  a.rdelta2 = 1.0 / c->s.delta2;
  a.radcut = t.radcut;
  a.mode = mode;
  a.pflux = (proj && t.nflux > 0 && c->d_fmap) ? 1 : 0;
  a.fmap = c->d_fmap;
  a.ir1flux = c->d_ir1flux;
  a.zar = c->d_zar1flux;
  a.zr = c->d_ring_zr + (int64_t)slot * t.nflux;
  a.zdr = c->d_ring_zdr + (int64_t)2 * slot * t.nflux;
  a.slice = (proj && t.ncvj > 0) ? c->d_slice : nullptr;
  a.ir2abs = c->s.ir2abs;
  a.rsah2 = 1.0 / std::sqrt(c->s.ah2);
  a.out32 = mode != 0 ? psi_shadow(c, c->d_psim1) : nullptr;
  a.psi32 = nullptr;
  if (sizeof(T) == 4 && mode == 2) {
    a.psi32 = psi_shadow(c, psi);  // current: H psi has just read it
    if (!a.psi32) return hipErrorInvalidValue;
  }
  const bool prodp = mode == 2 && proj && c->prod_on;
  a.tprod = prodp ? c->d_tprod : nullptr;
  a.tp_cstride = c->pr_tstride;
  a.ppart = c->d_ppart;
  timer_begin(c, K_UPDATE);
  double* out = mode == 0 ? c->d_hout : c->d_psim1;
  const bool projk = a.pflux || a.slice;
  const bool autoc = a.psi0 != nullptr;
  const unsigned nblk4 = (unsigned)((nb + 3) / 4);  // one work item per wave
  if (nblk4 > 0) {
    // one electronic state: V is diagonal (each Omega channel couples only to
    // itself, through the one V slab), whatever the channel count
    if (t.ne == 1)
      dispatch_update<T, true, false>(c, nblk4, psi, out, a, projk, autoc, prodp);
    else if (c->d_ch_cpl)  // several states, at most 4 channels per Omega
      dispatch_update<T, false, true>(c, nblk4, psi, out, a, projk, autoc, prodp);
    else
      dispatch_update<T, false, false>(c, nblk4, psi, out, a, projk, autoc, prodp);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the step loop (mode 2) reduces the partials on the second stream, beside
  // the next step's H psi (launch_projections)
  if (mode == 1) e = launch_update_tail(c, mode, slot, proj, c->stream);
  timer_end(c, K_UPDATE);
  return e;
}

// flux columns and norm of the update's work-item partials, into ring slot
hipError_t launch_update_tail(Ctx* c, int mode, int slot, bool proj, hipStream_t st) {
  const HostTables& t = c->t;
  if (mode == 2 && proj && t.nflux > 0 && c->d_fmap) {
    const int ng = t.nc * t.na * ((t.n2 + 63) / 64);
    hipLaunchKernelGGL(k_update_flux, dim3(ng), dim3(64), 0, st, c->d_uw_part, c->d_uw_gfirst,
                       c->d_fmap, c->d_ring_zr, c->d_ring_zdr, t.n2, t.nflux,
                       ring_slot(c, slot));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (mode != 0)
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, st, c->d_norm_part,
                       c->nnorm_blocks, c->d_ring_norm, ring_slot(c, slot));
  return hipGetLastError();
}

hipError_t launch_update(Ctx* c, int mode, const double* psi, bool proj) {
  int slot = c->ring_fill;
  if (c->precision == MW3_MIXED) return update_impl<float>(c, mode, slot, psi, proj);
  return update_impl<double>(c, mode, slot, psi, proj);
}

// the per-step sums of the update's partials and slice on stream st: flux
// columns and norm, reactant-state and product-state sums, into ring slot
hipError_t launch_tail_sums(Ctx* c, int slot, hipStream_t st) {
  const HostTables& t = c->t;
  hipError_t e;
  timer_begin(c, K_UPDATE_TAIL, st);
  if ((e = launch_update_tail(c, 2, slot, true, st)) != hipSuccess) return e;
  timer_end(c, K_UPDATE_TAIL, st);
  if (t.ncvj > 0) {
    timer_begin(c, K_PROJ, st);
    int nj = c->s.jmax - c->s.jini + 1, nv = c->s.nvmax - c->s.nvini + 1;
    hipLaunchKernelGGL(k_cvj, dim3(t.nc * nj), dim3(64 * CVJ_WAVES), sizeof(double) * t.n1, st,
                       c->d_slice, c->d_dcvj, c->d_fdv, c->d_ring_cvj, nj, nv, t.na, t.n1,
                       (int64_t)t.ncvj, ring_slot(c, slot));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    timer_end(c, K_PROJ, st);
  }
  if (c->prod_on && (e = launch_prod_step(c, slot, st)) != hipSuccess) return e;
  return hipSuccess;
}

// reactant-state amplitudes of the slice the update wrote, into ring slot
hipError_t launch_projections(Ctx* c, int slot) {
  hipError_t e;
  // on the second stream, after the update that wrote the slice and the
  // partials: the flux columns and the norm of the step, then the
  // reactant- and product-state sums
  if ((e = hipEventRecord(c->ev_slice, c->stream)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(c->aux, c->ev_slice, 0)) != hipSuccess) return e;
  if ((e = launch_tail_sums(c, slot, c->aux)) != hipSuccess) return e;
  if ((e = hipEventRecord(c->ev_cvj, c->aux)) != hipSuccess) return e;
  c->cvj_pending = true;
  return hipSuccess;
}

hipError_t launch_energy_accum(Ctx* c) {
  const HostTables& t = c->t;
  int ns = c->ring_fill;
  if (ns == 0 || c->s.netot <= 0) return hipSuccess;
  timer_begin(c, K_ENERGY);
  hipError_t e0 = c->prod_on ? launch_prod_flush(c) : hipSuccess;
  if (e0 != hipSuccess) return e0;
  int netot = c->s.netot;
  hipLaunchKernelGGL(k_wmat, dim3(nblk(netot, 128), ns), dim3(128), 0, c->stream, c->d_wmat,
                     c->d_etot, netot, c->ring_k0, ns, c->s.hbr, c->s.emindlt, c->s.delta2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // totalflux_k zeroes C_k for k < kminelastic when iphoto = 0 and iprod = 1
  int kmin_flux = (c->s.iphoto == 0 && c->s.iprod == 1) ? c->s.kminelastic : -1;
  if (t.nflux > 0) {
    const uint8_t* sk = kmin_flux >= 0 ? c->d_zero_skip : nullptr;
    if ((e = launch_eacc(c->stream, c->d_wmat, c->d_ring_zr, false, c->d_zcr, t.nflux, netot, ns,
                         c->ring_k0, sk, kmin_flux)) != hipSuccess)
      return e;
    if ((e = launch_eacc(c->stream, c->d_wmat, c->d_ring_zdr, true, c->d_zcdr, t.nflux, netot, ns,
                         c->ring_k0, sk, kmin_flux)) != hipSuccess)
      return e;
  }
  if (t.ncvj > 0) {
    int kmin = (c->s.iprod != 1 && c->s.iphoto == 0) ? c->s.kminelastic : -1;
    if ((e = launch_eacc(c->stream, c->d_wmat, c->d_ring_cvj, false, c->d_zs2, (int64_t)t.ncvj,
                         netot, ns, c->ring_k0, kmin >= 0 ? c->d_cvj_skip : nullptr, kmin)) !=
        hipSuccess)
      return e;
  }
  if (c->prod_on) {  // prodcvj: zS2prod += C_k(E) Cvjprod, every step
    if ((e = launch_eacc(c->stream, c->d_wmat, c->d_ring_cvjp, false, c->d_zs2prod,
                         (int64_t)c->pr_nomb * c->pr_nvj, netot, ns, c->ring_k0, nullptr, -1)) !=
        hipSuccess)
      return e;
  }
  timer_end(c, K_ENERGY);
  return hipSuccess;
}

hipError_t launch_psie(Ctx* c, int k) {
  if (c->s.nofe <= 0) return hipSuccess;
  int64_t total = (int64_t)c->t.nc * c->t.nslab;
  hipLaunchKernelGGL(k_psie, dim3(nblk(total, 256)), dim3(256), 0, c->stream, c->d_zfe,
                     c->d_psim1, c->d_efe, c->s.nofe, total, k, c->s.hbr, c->s.emindlt,
                     c->s.delta2);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// product-state projection (prodpaq / prodcvj, SRC/mod_coortrans01_02.f:694-822)
// ---------------------------------------------------------------------------
// funprod(c, ia, R) = sum over the update items of (c, ia, 64-R chunk) of their
// partials (sum_r transRgjac psi_new), then for each product helicity o and
// electronic state e the channel sum ring[o][e][l] = sum_{c: elec(c) = e}
// dtrans[c][o][l] funprod(c, l), l = (ia - nangproj0) n2p + (R - n2prod0).
// Block per (angle of the range, 64-R chunk of the range), lane = R.  Each
// lane reads back only its own column of f, so no barrier is needed; f is
// sized by the launch to the channel count (nc 64-lane rows), so that the
// block holds only the LDS it uses beside the kernels of the other streams.
__global__ __launch_bounds__(64) void k_prod_fun(const double* __restrict__ ppart,
                                                 const int32_t* __restrict__ gfirst,
                                                 const double* __restrict__ dtrans,
                                                 const int32_t* __restrict__ ch_elec,
                                                 double* __restrict__ ring_all, int nc, int ne,
                                                 int na, int n2, int nap0, int n2p0, int n2p,
                                                 int nomb, int rc0, int nrcp, int64_t L,
                                                 RingSlot rs) {
  double* __restrict__ ring = ring_all + (int64_t)rs.get() * nomb * ne * L;
  extern __shared__ double f[];  // [nc][64]
  const int lane = threadIdx.x;
  const int a = blockIdx.x / nrcp, rc = rc0 + blockIdx.x % nrcp;
  const int ia = nap0 + a;
  const int R = rc * 64 + lane;
  const int nrc = (n2 + 63) / 64;
  for (int ch = 0; ch < nc; ch++) {
    const int g = (ch * na + ia) * nrc + rc;
    double s = 0.0;
    for (int it = gfirst[g]; it < gfirst[g + 1]; it++) s += ppart[(int64_t)it * 64 + lane];
    f[ch * 64 + lane] = s;
  }
  if (R < n2p0 || R >= n2p0 + n2p) return;
  const int64_t l = (int64_t)a * n2p + (R - n2p0);
  for (int o = 0; o < nomb; o++)
    for (int e = 0; e < ne; e++) {
      double acc = 0.0;
      for (int ch = 0; ch < nc; ch++)
        if (ch_elec[ch] == e) acc += f[ch * 64 + lane] * dtrans[((int64_t)ch * nomb + o) * L + l];
      ring[((int64_t)o * ne + e) * L + l] = acc;
    }
}

// Cvjprod[s][o][vj] = sum over the K chunks of the flush GEMM's partials
// part[o][chunk][kRing][nvj]; the last step's row is kept for mw3_cvjprod_get
__global__ void k_prod_sum(const double* __restrict__ part, double* __restrict__ cvjp,
                           double* __restrict__ last, int nomb, int nvj, int nchunk, int ns) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (x >= nomb * nvj || s >= ns) return;
  const int o = x / nvj, vj = x % nvj;
  double acc = 0.0;
  for (int ch = 0; ch < nchunk; ch++)
    acc += part[(((int64_t)o * nchunk + ch) * kRing + s) * nvj + vj];
  cvjp[(int64_t)s * nomb * nvj + x] = acc;
  if (s == ns - 1) last[x] = acc;
}

hipError_t launch_prod_step(Ctx* c, int slot, hipStream_t s) {
  const HostTables& t = c->t;
  const int rc0 = c->pr_n2p0 / 64, rc1 = (c->pr_n2p0 + c->pr_n2p - 1) / 64;
  const int nrcp = rc1 - rc0 + 1;
  if (t.nc > kProdMaxCh) return hipErrorInvalidValue;  // mw3_prod_attach rejects it first
  timer_begin(c, K_PROD, s);
  hipLaunchKernelGGL(k_prod_fun, dim3((unsigned)(c->pr_nap * nrcp)), dim3(64),
                     sizeof(double) * 64 * t.nc, s, c->d_ppart,
                     c->d_uw_gfirst, c->d_dtrans, c->d_ch_elec, c->d_ring_prod, t.nc, t.ne,
                     t.na, t.n2, c->pr_nap0, c->pr_n2p0, c->pr_n2p, c->pr_nomb, rc0, nrcp, c->pr_L,
                     ring_slot(c, slot));
  hipError_t e = hipGetLastError();
  timer_end(c, K_PROD, s);
  return e;
}

// (timed with the energy accumulation it belongs to)
hipError_t launch_prod_flush(Ctx* c) {
  const int ns = c->ring_fill;
  hipError_t e = run_gemm<G_PROD, double, double, double, double>(
      c->stream, c->d_ring_prod, c->d_prodwf, c->d_prod_part, c->p_prod);
  if (e == hipSuccess) {
    const int nx = c->pr_nomb * c->pr_nvj;
    hipLaunchKernelGGL(k_prod_sum, dim3(nblk(nx, 128), ns), dim3(128), 0, c->stream,
                       c->d_prod_part, c->d_ring_cvjp, c->d_cvjp_last, c->pr_nomb, c->pr_nvj,
                       c->pr_nchunk, ns);
    e = hipGetLastError();
  }
  return e;
}

// flush GEMM: per (o, K chunk) C[s][vj] = sum_k ring[s][o][k] prodwf[o][vj][k]
// over k = (e, l) of the chunk, into part[o][chunk][kRing][nvj]
hipError_t build_prod_plan(Ctx* c) {
  const int64_t K = (int64_t)c->t.ne * c->pr_L;
  std::vector<GemmDesc> ds;
  for (int o = 0; o < c->pr_nomb; o++)
    for (int ch = 0; ch < c->pr_nchunk; ch++) {
      GemmDesc d = gd();
      d.a_off = (int64_t)o * K + (int64_t)ch * c->pr_kc;
      d.lda_m = (int)(c->pr_nomb * K);
      d.lda_k = 1;
      d.b_off = (int64_t)o * c->pr_nvj * K + (int64_t)ch * c->pr_kc;
      d.c_off = ((int64_t)o * c->pr_nchunk + ch) * kRing * c->pr_nvj;
      d.M = kRing;
      d.N = c->pr_nvj;
      d.K = (int)std::min<int64_t>(c->pr_kc, K - (int64_t)ch * c->pr_kc);
      ds.push_back(d);
    }
  c->p_prod.st.ldb_k = 1;
  c->p_prod.st.ldb_n = K;
  c->p_prod.st.ldc_m = c->pr_nvj;
  c->p_prod.st.ldc_n = 1;
  hipError_t e = upload_plan(&c->p_prod, ds, true);
  c->p_prod.vec = 0;  // generic MFMA kernel: B is K-contiguous here
  return e;
}

// P(E) and per-state sums (totalflux, SRC/mod_flux_01y2.f:778-816):
// out[0][E] = sum_f [E > V_f] Im(conj(zCR) zCdR), out[1+e][E] per state e;
// potf = nullptr drops the E > V_f test.
__global__ void k_flux_sum(const double* __restrict__ zcr, const double* __restrict__ zcdr,
                           const double* __restrict__ potf, const int32_t* __restrict__ felec,
                           const double* __restrict__ etot, double* __restrict__ out,
                           int64_t nflux, int netot, int ne) {
  int e = blockIdx.x;
  extern __shared__ double sh[];  // [(1+ne)][256]
  double acc[9];
  for (int q = 0; q <= ne && q < 9; q++) acc[q] = 0;
  double E = etot[e];
  for (int64_t f = threadIdx.x; f < nflux; f += blockDim.x) {
    if (potf && !(E > potf[f])) continue;
    double ar = zcr[2 * (f * netot + e)], ai = zcr[2 * (f * netot + e) + 1];
    double br = zcdr[2 * (f * netot + e)], bi = zcdr[2 * (f * netot + e) + 1];
    double v = ar * bi - ai * br;
    acc[0] += v;
    acc[1 + felec[f]] += v;
  }
  for (int q = 0; q <= ne; q++) sh[q * 256 + threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int q = 0; q <= ne; q++) sh[q * 256 + threadIdx.x] += sh[q * 256 + threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q <= ne; q++) out[(int64_t)q * netot + e] = sh[q * 256];
}

hipError_t launch_flux_result(Ctx* c, double* d_out) {
  int ne = c->t.ne;
  hipLaunchKernelGGL(k_flux_sum, dim3(c->s.netot), dim3(256), sizeof(double) * 256 * (1 + ne),
                     c->stream, c->d_zcr, c->d_zcdr, c->flux_etest ? c->d_potiflux : nullptr,
                     c->d_flux_elec, c->d_etot,
                     d_out, c->t.nflux, c->s.netot, ne);
  return hipGetLastError();
}

}  // namespace mw3
