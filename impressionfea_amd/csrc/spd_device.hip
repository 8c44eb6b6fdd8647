// Device supernodal Cholesky for MI355X (gfx950): multifrontal, level-scheduled, deterministic.
//
// Replaces the numeric phases the reference delegates to cuDSS (CudssSolver.cs:126-145) and MKL PARDISO
// (MklPardisoSolver.cs:74-99). Per elimination-tree level:
//   extend-add   children's update blocks into the parent front, one child rank per launch (no atomics),
//   panel        FP64 Cholesky of a 64-wide diagonal block in one wave + FP64 TRSM of the rows below,
//   syrk         trailing update of the front with the panel on MFMA (FP64 16x16x4, or FP32 16x16x4 with FP32
//                accumulation in mixed precision).
// Solves run a multifrontal forward sweep (update vectors passed to the parent in fixed child order) and a gather
// backward sweep, in FP64 whatever the factor precision.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "spd_device.h"

namespace ifea {

constexpr int NB = 64;          // panel width
constexpr int OB = 256;         // outer block: the trailing matrix beyond it is updated once per OB columns
constexpr int PANEL_ROWS = 256; // rows per panel task (one per thread)
constexpr int EXT_TARGET = 8192;
constexpr long long ZERO_CHUNK = 1 << 16;
constexpr int BIG_K = 256;      // supernodes wider than this, or with panels larger than BIG_PANEL entries, use the
constexpr long long BIG_PANEL = 32768;  // blocked multi-workgroup triangular sweeps
constexpr int SB = 64;          // diagonal block of the blocked sweeps (= the factor's panel width)
constexpr int SWEEP_ROWS = 256; // rows (forward) per workgroup of a blocked sweep step
constexpr int SWEEP_COLS = 64;  // columns (backward) per workgroup of a blocked sweep step

// ------------------------------------------------------------------------------------------------ MFMA wrappers
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // item r of lane l holds D[row][l % 16]; the row layout differs between the f64 and f32 instructions
  static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

__device__ __forceinline__ int sn_k(const SymDev& sd, int s) { return sd.sn_first[s + 1] - sd.sn_first[s]; }
__device__ __forceinline__ long long sn_m(const SymDev& sd, int s) { return sd.sn_rowptr[s + 1] - sd.sn_rowptr[s]; }

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// ------------------------------------------------------------------------------------------------ kernels
template <typename T>
__global__ void k_scatter(const double* __restrict__ vals, const int* __restrict__ src, const long long* __restrict__ dst,
                          long long cnt, T* __restrict__ L) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k < cnt) L[dst[k]] = (T)vals[src[k]];
}

// task: (supernode s, chunk index): zero one chunk of the supernode's update block before its level starts
template <typename T>
__global__ __launch_bounds__(256) void k_zero_u(const int4* __restrict__ tasks, SymDev sd, T* __restrict__ U) {
  int4 tk = tasks[blockIdx.x];
  int s = tk.x;
  long long mu = sn_m(sd, s) - sn_k(sd, s);
  long long size = mu * mu;
  long long b = (long long)tk.y * ZERO_CHUNK, e = min(size, b + ZERO_CHUNK);
  T* u = U + sd.Uoff[s];
  for (long long i = b + threadIdx.x; i < e; i += blockDim.x) u[i] = T(0);
}

// task: (child c, first column, end column, parent p) of the child's update block
template <typename T>
__global__ __launch_bounds__(256) void k_extend(const int4* __restrict__ tasks, SymDev sd, T* __restrict__ L,
                                                T* __restrict__ U) {
  int4 tk = tasks[blockIdx.x];
  int c = tk.x, p = tk.w;
  int kc = sn_k(sd, c);
  int mu = (int)(sn_m(sd, c) - kc);
  int kp = sn_k(sd, p);
  long long mp = sn_m(sd, p);
  long long mpu = mp - kp;
  const T* Uc = U + sd.Uoff[c];
  const int* rel = sd.rel + sd.rel_off[c];
  T* Lp = L + sd.Loff[p];
  T* Up = U + sd.Uoff[p];
  for (int j = tk.y; j < tk.z; j++) {
    int pc = rel[j];
    T* dcol = pc < kp ? Lp + (long long)pc * mp : Up + (long long)(pc - kp) * mpu - kp;
    const T* scol = Uc + (long long)j * mu;
    for (int i = j + threadIdx.x; i < mu; i += blockDim.x) dcol[rel[i]] += scol[i];
  }
}

// task: (supernode s, first row, end row, 1 + slot of the factored diagonal block in the scratch, or 0).
// c0 = first column of the 64-wide block. Every task of a front re-factors the (unmodified) diagonal block in LDS;
// the designated task stores the factor to the scratch, and k_commit copies it into the panel after the launch, so
// no task can read a block that another task of the same launch has already overwritten.
template <typename T>
__global__ __launch_bounds__(256) void k_panel(const int4* __restrict__ tasks, int c0, SymDev sd, T* __restrict__ L,
                                               double* __restrict__ dscratch, int* __restrict__ err) {
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int w = min(NB, k - c0);
  T* P = L + sd.Loff[s];
  __shared__ double D[NB][NB + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < NB * NB; idx += blockDim.x) {
    int j = idx / NB, i = idx % NB;
    double v;
    if (i < w && j < w) v = (i >= j) ? (double)P[(long long)(c0 + j) * m + c0 + i] : 0.0;
    else v = (i == j) ? 1.0 : 0.0;
    D[i][j] = v;
  }
  __syncthreads();
  if (tid < 64) {
    // one wave factors the block in LDS: lane i updates row i; column j is read back by broadcast
    const int i = tid;
    bool bad = false;
    for (int j = 0; j < NB; j++) {
      double piv = D[j][j];
      bad |= !(piv > 0.0) || !isfinite(piv);
      double d = sqrt(piv > 0.0 ? piv : 1.0);
      double lij = (i > j) ? D[i][j] / d : 0.0;
      wave_sync_lds();
      if (i == j) D[i][j] = d;
      else if (i > j) D[i][j] = lij;
      wave_sync_lds();
      for (int q = j + 1; q <= i; q++) D[i][q] -= lij * D[q][j];
      wave_sync_lds();
    }
    for (int q = i + 1; q < NB; q++) D[i][q] = 0.0;
    if (tk.w) {
      double* dsc = dscratch + (long long)(tk.w - 1) * NB * NB;
      for (int q = 0; q < NB; q++) dsc[q * NB + i] = D[i][q];
      if (bad && i == 0) atomicOr(err, 1);
    }
  }
  __syncthreads();
  // rows below the block: X = A L11^{-T}, one row per thread
  const long long row = (long long)tk.y + tid;
  const bool valid = row < tk.z;
  double x[NB];
#pragma unroll
  for (int q = 0; q < NB; q++) {
    double a = (q < w && valid) ? (double)P[(long long)(c0 + q) * m + row] : 0.0;
#pragma unroll
    for (int p = 0; p < q; p++) a -= x[p] * D[q][p];
    x[q] = a / D[q][q];
    if (q < w && valid) P[(long long)(c0 + q) * m + row] = (T)x[q];
  }
}


// Diagonal block of one 64-column step: FP64 Cholesky in one wave (lane i owns row i, columns broadcast through LDS),
// then the inverse of the factor (lane c solves L x = e_c). The factor goes into the panel; the inverse is kept
// (Lt[c][r] = Linv[r][c]) for the TRSM launch that follows and for the triangular sweeps. task: (supernode s, 0, 0, 0)
template <typename T>
__global__ __launch_bounds__(64) void k_potrf_inv(const int4* __restrict__ tasks, int c0, SymDev sd,
                                                  T* __restrict__ L, double* __restrict__ Linv,
                                                  int* __restrict__ err) {
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int w = min(NB, k - c0);
  T* P = L + sd.Loff[s];
  __shared__ double D[NB][NB + 1];
  const int i = threadIdx.x;
  for (int j = 0; j < NB; j++) {
    double v;
    if (i < w && j < w) v = (i >= j) ? (double)P[(long long)(c0 + j) * m + c0 + i] : 0.0;
    else v = (i == j) ? 1.0 : 0.0;
    D[i][j] = v;
  }
  wave_sync_lds();
  bool bad = false;
  for (int j = 0; j < NB; j++) {
    double piv = D[j][j];
    bad |= !(piv > 0.0) || !isfinite(piv);
    double d = sqrt(piv > 0.0 ? piv : 1.0);
    double lij = (i > j) ? D[i][j] / d : 0.0;
    wave_sync_lds();
    if (i == j) D[i][j] = d;
    else if (i > j) D[i][j] = lij;
    wave_sync_lds();
    for (int q = j + 1; q <= i; q++) D[i][q] -= lij * D[q][j];
    wave_sync_lds();
  }
  for (int q = i + 1; q < NB; q++) D[i][q] = 0.0;
  wave_sync_lds();
  // factor into the panel (column q, row i)
  for (int q = 0; q < w; q++)
    if (i < w && i >= q) P[(long long)(c0 + q) * m + c0 + i] = (T)D[i][q];
  if (bad && i == 0) atomicOr(err, 1);
  // lane c: x = L^{-1} e_c, x[r] = (delta_rc - sum_{p<r} L[r][p] x[p]) / L[r][r]; store Lt[p][q] = Linv[q][p]
  double* Lt = Linv + sd.Ioff[s] + (long long)(c0 / NB) * NB * NB;
  const int c = i;
  double x[NB];
#pragma unroll
  for (int r = 0; r < NB; r++) {
    double a = (r == c) ? 1.0 : 0.0;
#pragma unroll
    for (int p = 0; p < r; p++) a -= D[r][p] * x[p];
    x[r] = a / D[r][r];
  }
#pragma unroll
  for (int r = 0; r < NB; r++) Lt[c * NB + r] = x[r];  // Linv[r][c] -> Lt[c][r]
}

// Rows below the diagonal block: X = A L11^{-T} as a 64x64x64 FP64 MFMA product per task, written back into the
// panel. task: (supernode s, first row, end row (<= first + 64), 0)
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_mfma(const int4* __restrict__ tasks, int c0, SymDev sd,
                                                   T* __restrict__ L, const double* __restrict__ Linv) {
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int w = min(NB, k - c0);
  T* P = L + sd.Loff[s];
  __shared__ double sA[NB][64 + 2];  // sA[p][r] = A(r0 + r, c0 + p)
  __shared__ double sL[NB][64 + 2];  // sL[p][q] = Linv(q, p)
  const long long r0 = tk.y;
  const int nr = tk.z - tk.y;
  const double* Lt = Linv + sd.Ioff[s] + (long long)(c0 / NB) * NB * NB;
  for (int idx = threadIdx.x; idx < NB * 64; idx += blockDim.x) {
    int p = idx >> 6, rr = idx & 63;
    sA[p][rr] = (p < w && rr < nr) ? (double)P[(long long)(c0 + p) * m + r0 + rr] : 0.0;
    // Lt is stored [c][r] = Linv[r][c]; sL[p][q] = Linv[q][p] = Lt[p][q]
    sL[p][rr] = (p < w && rr < w) ? Lt[p * NB + rr] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int qb = (wv >> 1) * 32, rb = (wv & 1) * 32;
  const int l16 = lane & 15;
  typedef Mfma<double>::v4 v4;
  v4 acc[2][2] = {{{0, 0, 0, 0}, {0, 0, 0, 0}}, {{0, 0, 0, 0}, {0, 0, 0, 0}}};
  for (int kk = 0; kk < NB; kk += 4) {
    const int p = kk + (lane >> 4);
    double a0 = sL[p][qb + l16], a1 = sL[p][qb + 16 + l16];
    double b0 = sA[p][rb + l16], b1 = sA[p][rb + 16 + l16];
    acc[0][0] = Mfma<double>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<double>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<double>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<double>::op(a1, b1, acc[1][1]);
  }
  // D[q][r] = X(r0 + r, c0 + q)
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int q = qb + 16 * a + Mfma<double>::row(lane, r);
        int rr = rb + 16 * b + l16;
        if (q < w && rr < nr) P[(long long)(c0 + q) * m + r0 + rr] = (T)acc[a][b][r];
      }
}

// Copies the factored diagonal blocks of one panel launch into their panels (tasks as k_panel; non-designated exit).
template <typename T>
__global__ __launch_bounds__(256) void k_commit(const int4* __restrict__ tasks, int c0, SymDev sd, T* __restrict__ L,
                                                const double* __restrict__ dscratch) {
  const int4 tk = tasks[blockIdx.x];
  if (!tk.w) return;
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int w = min(NB, k - c0);
  T* P = L + sd.Loff[s];
  const double* dsc = dscratch + (long long)(tk.w - 1) * NB * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    int q = idx / NB, i = idx % NB;
    if (i < w && q < w && i >= q) P[(long long)(c0 + q) * m + c0 + i] = (T)dsc[q * NB + i];
  }
}

// Trailing update C -= X X^T of a front, 64x64 output tile per workgroup, MFMA (FP64 16x16x4 or FP32 16x16x4).
// task: (supernode s, row tile I, column tile J, 0), I >= J.
// mode 0 (inner): X = the 64 columns [kb, kb+w) just factored; C = columns [kb+w, min(k, outer_end)) of the same
//                 outer block, all rows below (delayed right-looking update inside a 256-column outer block).
// mode 1 (outer): X = the outer block [kb, kb+min(OB, k-kb)); C = every column to its right, including the update
//                 block U, so the large trailing matrix is read and written once per 256 columns.
template <typename T>
__global__ __launch_bounds__(256) void k_syrk(const int4* __restrict__ tasks, int kb, int mode, SymDev sd,
                                              T* __restrict__ L, T* __restrict__ U) {
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int kw = mode == 0 ? min(NB, k - kb) : min(OB, k - kb);
  const long long rb = kb + kw;
  const long long col_end = mode == 0 ? min((long long)k, (long long)(kb / OB) * OB + OB) : m;
  const T* P = L + sd.Loff[s];
  __shared__ T Xi[NB][64 + 4];
  __shared__ T Xj[NB][64 + 4];
  const long long i0 = rb + 64LL * tk.y, j0 = rb + 64LL * tk.z;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ib = (wv >> 1) * 32, jb = (wv & 1) * 32;
  const int l16 = lane & 15;
  typedef typename Mfma<T>::v4 v4;
  v4 acc00 = {0, 0, 0, 0}, acc01 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};
  for (int kc = 0; kc < kw; kc += NB) {
    const int w = min(NB, kw - kc);
    if (kc) __syncthreads();
    for (int idx = threadIdx.x; idx < NB * 64; idx += blockDim.x) {
      int q = idx >> 6, rr = idx & 63;
      long long ri = i0 + rr, rj = j0 + rr;
      const T* cp = P + (long long)(kb + kc + q) * m;
      Xi[q][rr] = (q < w && ri < m) ? cp[ri] : T(0);
      Xj[q][rr] = (q < w && rj < m) ? cp[rj] : T(0);
    }
    __syncthreads();
    const int kq = (w + 3) & ~3;
    for (int kk = 0; kk < kq; kk += 4) {
      const int q = kk + (lane >> 4);
      T a0 = Xj[q][jb + l16], a1 = Xj[q][jb + 16 + l16];
      T b0 = Xi[q][ib + l16], b1 = Xi[q][ib + 16 + l16];
      acc00 = Mfma<T>::op(a0, b0, acc00);
      acc01 = Mfma<T>::op(a0, b1, acc01);
      acc10 = Mfma<T>::op(a1, b0, acc10);
      acc11 = Mfma<T>::op(a1, b1, acc11);
    }
  }
  // D[a][b] = sum_q Xj[a][q] Xi[b][q] = contribution to C(i = i0+ib+16bi+b, j = j0+jb+16ai+a)
  v4 accs[2][2] = {{acc00, acc01}, {acc10, acc11}};
  T* Ls = L + sd.Loff[s];
  T* Us = U + sd.Uoff[s];
  const long long mu = m - k;
#pragma unroll
  for (int ai = 0; ai < 2; ai++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        long long j = j0 + jb + 16 * ai + Mfma<T>::row(lane, r);
        long long i = i0 + ib + 16 * bi + l16;
        if (i < m && j < col_end && i >= j) {
          T* cp = (j < k) ? (Ls + j * m + i) : (Us + (j - k) * mu + (i - k));
          *cp -= accs[ai][bi][r];
        }
      }
}


// Outer trailing update with 128x128 output tiles (mode 1 of k_syrk): 4 waves, each a 64x64 quadrant as 4x4 MFMA
// blocks. X is streamed in 16-column chunks through a double-buffered LDS tile; the next chunk's global loads are in
// flight (in registers) while the current chunk feeds the MFMAs, one barrier per chunk.
// task: (supernode s, row tile I, column tile J, 0), I >= J.
template <typename T>
__global__ __launch_bounds__(256) void k_syrk_big(const int4* __restrict__ tasks, int kb, SymDev sd,
                                                  T* __restrict__ L, T* __restrict__ U) {
  constexpr int KC = 16;
  constexpr int PER = KC * 128 / 256;  // elements of each operand a thread stages per chunk
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int kw = min(OB, k - kb);
  const long long rb = kb + kw;
  const T* P = L + sd.Loff[s];
  __shared__ T Xi[2][KC][128 + 4];
  __shared__ T Xj[2][KC][128 + 4];
  const long long i0 = rb + 128LL * tk.y, j0 = rb + 128LL * tk.z;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ib = (wv >> 1) * 64, jb = (wv & 1) * 64;
  const int l16 = lane & 15;
  const int nch = (kw + KC - 1) / KC;
  T ri[PER], rj[PER];
  auto fetch = [&](int c) {
#pragma unroll
    for (int t = 0; t < PER; t++) {
      int idx = threadIdx.x + 256 * t;
      int q = idx >> 7, rr = idx & 127;
      int col = c * KC + q;
      const T* cp = P + (long long)(kb + col) * m;
      long long ra = i0 + rr, rc = j0 + rr;
      ri[t] = (col < kw && ra < m) ? cp[ra] : T(0);
      rj[t] = (col < kw && rc < m) ? cp[rc] : T(0);
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int t = 0; t < PER; t++) {
      int idx = threadIdx.x + 256 * t;
      int q = idx >> 7, rr = idx & 127;
      Xi[buf][q][rr] = ri[t];
      Xj[buf][q][rr] = rj[t];
    }
  };
  typedef typename Mfma<T>::v4 v4;
  v4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = v4{0, 0, 0, 0};
  fetch(0);
  stage(0);
  __syncthreads();
  for (int c = 0; c < nch; c++) {
    const int buf = c & 1;
    if (c + 1 < nch) fetch(c + 1);
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int q = kk + (lane >> 4);
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; a++) av[a] = Xj[buf][q][jb + 16 * a + l16];
#pragma unroll
      for (int b = 0; b < 4; b++) bv[b] = Xi[buf][q][ib + 16 * b + l16];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = Mfma<T>::op(av[a], bv[b], acc[a][b]);
    }
    if (c + 1 < nch) stage(buf ^ 1);
    __syncthreads();
  }
  T* Ls = L + sd.Loff[s];
  T* Us = U + sd.Uoff[s];
  const long long mu = m - k;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        long long j = j0 + jb + 16 * a + Mfma<T>::row(lane, r);
        long long i = i0 + ib + 16 * b + l16;
        if (i < m && j < m && i >= j) {
          T* cp = (j < k) ? (Ls + j * m + i) : (Us + (j - k) * mu + (i - k));
          *cp -= acc[a][b][r];
        }
      }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd(const int* __restrict__ sns, SymDev sd, const T* __restrict__ L,
                                             double* __restrict__ y, double* __restrict__ V, double* __restrict__ Wg,
                                             int cap) {
  extern __shared__ double shw[];
  const int s = sns[blockIdx.x];
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  double* w = (m <= cap) ? shw : Wg + sd.Woff[s];
  const int tid = threadIdx.x, bd = blockDim.x;
  for (long long i = tid; i < m; i += bd) w[i] = i < k ? y[f + i] : 0.0;
  __syncthreads();
  for (int cp = sd.child_ptr[s]; cp < sd.child_ptr[s + 1]; cp++) {
    int c = sd.child_list[cp];
    int mu = (int)(sn_m(sd, c) - sn_k(sd, c));
    const int* rel = sd.rel + sd.rel_off[c];
    const double* vc = V + sd.Voff[c];
    for (int i = tid; i < mu; i += bd) w[rel[i]] += vc[i];
    __syncthreads();
  }
  const T* P = L + sd.Loff[s];
  for (int q = 0; q < k; q++) {
    const T* cq = P + (long long)q * m;
    double xq = w[q] / (double)cq[q];
    if (tid == 0) y[f + q] = xq;
    for (long long i = q + 1 + tid; i < m; i += bd) w[i] -= (double)cq[i] * xq;
    __syncthreads();
  }
  double* vs = V + sd.Voff[s];
  for (long long i = k + tid; i < m; i += bd) vs[i - k] = w[i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd(const int* __restrict__ sns, SymDev sd, const T* __restrict__ L,
                                             double* __restrict__ y, double* __restrict__ Wg, int cap) {
  extern __shared__ double shz[];
  const int s = sns[blockIdx.x];
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int* rows = sd.sn_rows + sd.sn_rowptr[s];
  double* z = (k <= cap) ? shz : Wg + sd.Woff[s];
  const T* P = L + sd.Loff[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int q = wv; q < k; q += nw) {
    const T* cq = P + (long long)q * m;
    double acc = 0.0;
    for (long long i = k + lane; i < m; i += 64) acc += (double)cq[i] * y[rows[i]];
    acc = wave_sum(acc);
    if (lane == 0) z[q] = y[f + q] - acc;
  }
  __syncthreads();
  for (int q = k - 1; q >= 0; q--) {
    double xq = z[q] / (double)P[(long long)q * m + q];
    if (threadIdx.x == 0) y[f + q] = xq;
    for (int p = threadIdx.x; p < q; p += blockDim.x) z[p] -= (double)P[(long long)p * m + q] * xq;
    __syncthreads();
  }
}


// ---------------------------------------------------------------------------------------- blocked sweeps (wide fronts)
// Forward: W_s = [y_own ; 0] + children's update vectors (fixed child order).
__global__ __launch_bounds__(256) void k_fwdb_gather(const int* __restrict__ sns, SymDev sd, const double* __restrict__ y,
                                                     const double* __restrict__ V, double* __restrict__ Wg) {
  const int s = sns[blockIdx.x];
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  double* w = Wg + sd.Woff[s];
  for (long long i = threadIdx.x; i < m; i += blockDim.x) w[i] = i < k ? y[f + i] : 0.0;
  __syncthreads();
  for (int cp = sd.child_ptr[s]; cp < sd.child_ptr[s + 1]; cp++) {
    int c = sd.child_list[cp];
    int mu = (int)(sn_m(sd, c) - sn_k(sd, c));
    const int* rel = sd.rel + sd.rel_off[c];
    const double* vc = V + sd.Voff[c];
    for (int i = threadIdx.x; i < mu; i += blockDim.x) w[rel[i]] += vc[i];
    __syncthreads();
  }
}

// One 64-column step of the forward sweep. task: (s, first row, end row, writer). x_t = Linv_t w_t from the kept
// inverse diagonal block (a parallel 64x64 product, no sequential chain), then W(rows) -= L(rows, t) x_t for this
// task's rows >= c0 + w. W(c0 .. c0+w) is read by every task and written by none.
template <typename T>
__global__ __launch_bounds__(256) void k_fwdb_step(const int4* __restrict__ tasks, int c0, SymDev sd,
                                                   const T* __restrict__ L, const double* __restrict__ Linv,
                                                   double* __restrict__ y, double* __restrict__ Wg) {
  __shared__ double ws[NB], xs[NB];
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int w = min(NB, k - c0);
  const T* P = L + sd.Loff[s];
  const double* Lt = Linv + sd.Ioff[s] + (long long)(c0 / NB) * NB * NB;
  double* W = Wg + sd.Woff[s];
  if (threadIdx.x < NB) ws[threadIdx.x] = threadIdx.x < w ? W[c0 + threadIdx.x] : 0.0;
  __syncthreads();
  if (threadIdx.x < NB) {
    const int r = threadIdx.x;
    double x = 0.0;
    for (int c = 0; c < w; c++) x += Lt[c * NB + r] * ws[c];
    xs[r] = r < w ? x : 0.0;
    if (tk.w && r < w) y[f + c0 + r] = x;
  }
  __syncthreads();
  const long long row = (long long)tk.y + threadIdx.x;
  if (row < tk.z) {
    const T* pr = P + (long long)c0 * m + row;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int q = 0;
    for (; q + 4 <= w; q += 4) {
      a0 += (double)pr[(long long)q * m] * xs[q];
      a1 += (double)pr[(long long)(q + 1) * m] * xs[q + 1];
      a2 += (double)pr[(long long)(q + 2) * m] * xs[q + 2];
      a3 += (double)pr[(long long)(q + 3) * m] * xs[q + 3];
    }
    for (; q < w; q++) a0 += (double)pr[(long long)q * m] * xs[q];
    W[row] -= (a0 + a1) + (a2 + a3);
  }
}

// Update vector of a wide front for its parent: V_s = W_s[k..m).
__global__ __launch_bounds__(256) void k_fwdb_finish(const int* __restrict__ sns, SymDev sd, const double* __restrict__ Wg,
                                                     double* __restrict__ V) {
  const int s = sns[blockIdx.x];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const double* w = Wg + sd.Woff[s];
  double* vs = V + sd.Voff[s];
  for (long long i = k + threadIdx.x; i < m; i += blockDim.x) vs[i - k] = w[i];
}

// Backward, first part: the solution at a wide front's lower rows, gathered into one contiguous vector (the front's
// forward update-vector slot, free once the forward sweep is done).
__global__ __launch_bounds__(256) void k_bwdb_xb(const int* __restrict__ sns, SymDev sd, const double* __restrict__ y,
                                                 double* __restrict__ V) {
  const int s = sns[blockIdx.x];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int* rows = sd.sn_rows + sd.sn_rowptr[s];
  double* xb = V + sd.Voff[s];
  for (long long i = k + threadIdx.x; i < m; i += blockDim.x) xb[i - k] = y[rows[i]];
}

// z_q = y_own[q] - L21(:, q)^T xb for every own column q: one wave per column, 4 independent partial sums per lane.
// task: (s, first column, end column, 0)
template <typename T>
__global__ __launch_bounds__(256) void k_bwdb_gemv(const int4* __restrict__ tasks, SymDev sd, const T* __restrict__ L,
                                                   const double* __restrict__ y, const double* __restrict__ V,
                                                   double* __restrict__ Wg) {
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const long long mb = m - k;
  const T* P = L + sd.Loff[s];
  const double* xb = V + sd.Voff[s];
  double* z = Wg + sd.Woff[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int q = tk.y + wv; q < tk.z; q += 4) {
    const T* cq = P + (long long)q * m + k;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    long long i = lane;
    for (; i + 192 < mb; i += 256) {
      a0 += (double)cq[i] * xb[i];
      a1 += (double)cq[i + 64] * xb[i + 64];
      a2 += (double)cq[i + 128] * xb[i + 128];
      a3 += (double)cq[i + 192] * xb[i + 192];
    }
    for (; i < mb; i += 64) a0 += (double)cq[i] * xb[i];
    double acc = wave_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) z[q] = y[f + q] - acc;
  }
}

// One 64-column step (descending) of the backward sweep. task: (s, first column, end column, writer); columns < c0.
// x_t = Linv_t^T z_t from the kept inverse block, then z_j -= L(c0.., j)^T x_t for this task's columns (wave per
// column).
template <typename T>
__global__ __launch_bounds__(256) void k_bwdb_step(const int4* __restrict__ tasks, int c0, SymDev sd,
                                                   const T* __restrict__ L, const double* __restrict__ Linv,
                                                   double* __restrict__ y, double* __restrict__ Wg) {
  __shared__ double zs[NB], xs[NB];
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int w = min(NB, k - c0);
  const T* P = L + sd.Loff[s];
  const double* Lt = Linv + sd.Ioff[s] + (long long)(c0 / NB) * NB * NB;
  double* z = Wg + sd.Woff[s];
  if (threadIdx.x < NB) zs[threadIdx.x] = threadIdx.x < w ? z[c0 + threadIdx.x] : 0.0;
  __syncthreads();
  if (threadIdx.x < NB) {
    // x_c = sum_r Linv[r][c] z_r = sum_r Lt[c][r] z_r
    const int c = threadIdx.x;
    const double* lc = Lt + c * NB;
    double x = 0.0;
    for (int r = 0; r < w; r++) x += lc[r] * zs[r];
    xs[c] = c < w ? x : 0.0;
    if (tk.w && c < w) y[f + c0 + c] = x;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = tk.y + wv; j < tk.z; j += 4) {
    const T* cj = P + (long long)j * m + c0;
    double acc = lane < w ? (double)cj[lane] * xs[lane] : 0.0;
    acc = wave_sum(acc);
    if (lane == 0) z[j] -= acc;
  }
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_perm_in(const double* __restrict__ b, const int* __restrict__ perm, double* __restrict__ y, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k] = b[perm[k]];
}
__global__ void k_perm_out(const double* __restrict__ y, const int* __restrict__ perm, double* __restrict__ x, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[perm[k]] = y[k];
}

// ------------------------------------------------------------------------------------------------ host
SpdDevice::~SpdDevice() {
  drop_graph();
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
}

hipEvent_t SpdDevice::event(size_t i) {
  while (ev_pool_.size() <= i) {
    hipEvent_t e;
    IFEA_HIP(hipEventCreate(&e));
    ev_pool_.push_back(e);
  }
  return ev_pool_[i];
}

size_t SpdDevice::device_bytes() const {
  return L_.bytes() + U_.bytes() + V_.bytes() + W_.bytes() + amap_dst_.bytes() + amap_src_.bytes() +
         sn_rows_.bytes() + ext_tasks_.bytes() + panel_tasks_.bytes() + syrk_tasks_.bytes() + zero_tasks_.bytes();
}

void SpdDevice::analyze(int n, const int* colPtr, const int* rowIdx, const int* perm, const SymbolicOptions& opt) {
  analyzed_ = false;
  factored_ = false;
  drop_graph();
  int rc = symbolic_analyze(n, colPtr, rowIdx, perm, opt, S_);
  if (rc) throw Error(rc == -2 ? -2 : (rc == -1 ? -1 : -3), "symbolic analysis failed: " + S_.error);
  const Symbolic& S = S_;
  const int ns = S.ns;
  // children lists (ascending)
  std::vector<int> cptr(ns + 1, 0), clist;
  for (int s = 0; s < ns; s++) if (S.sn_parent[s] >= 0) cptr[S.sn_parent[s] + 1]++;
  for (int s = 0; s < ns; s++) cptr[s + 1] += cptr[s];
  clist.resize(cptr[ns]);
  {
    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
    for (int s = 0; s < ns; s++) if (S.sn_parent[s] >= 0) clist[fill[S.sn_parent[s]]++] = s;
  }
  std::vector<long long> Woff(ns + 1, 0), Ioff(ns + 1, 0);
  for (int s = 0; s < ns; s++) {
    Woff[s + 1] = Woff[s] + (S.sn_rowptr[s + 1] - S.sn_rowptr[s]);
    Ioff[s + 1] = Ioff[s] + (long long)((S.sn_first[s + 1] - S.sn_first[s] + NB - 1) / NB) * NB * NB;
  }
  // level schedule and task lists
  int nl = S.max_level + 1;
  levels_.assign(nl, Level());
  std::vector<std::vector<int>> bylevel(nl);
  for (int s = 0; s < ns; s++) bylevel[S.sn_level[s]].push_back(s);
  std::vector<int> lvl_sn;
  std::vector<int4> ext, pan, syr, zer, pot;
  int max_slots = 1;
  for (int l = 0; l < nl; l++) {
    Level& L = levels_[l];
    L.sn_begin = (long long)lvl_sn.size();
    for (int s : bylevel[l]) {
      lvl_sn.push_back(s);
      L.max_m = std::max<long long>(L.max_m, S.sn_rowptr[s + 1] - S.sn_rowptr[s]);
    }
    L.sn_end = (long long)lvl_sn.size();
    for (int s : bylevel[l])
      L.solve_bytes += (double)(S.sn_rowptr[s + 1] - S.sn_rowptr[s]) * (S.sn_first[s + 1] - S.sn_first[s]);
    L.zero.begin = (long long)zer.size();
    for (int s : bylevel[l]) {
      long long mu = (S.sn_rowptr[s + 1] - S.sn_rowptr[s]) - (S.sn_first[s + 1] - S.sn_first[s]);
      long long chunks = (mu * mu + ZERO_CHUNK - 1) / ZERO_CHUNK;
      for (long long c = 0; c < chunks; c++) zer.push_back(make_int4(s, (int)c, 0, 0));
    }
    L.zero.end = (long long)zer.size();
    // extend-add rounds: round r = r-th child of every parent at this level
    int maxc = 0;
    for (int s : bylevel[l]) maxc = std::max(maxc, cptr[s + 1] - cptr[s]);
    for (int r = 0; r < maxc; r++) {
      LaunchList ll;
      ll.begin = (long long)ext.size();
      for (int s : bylevel[l]) {
        if (cptr[s + 1] - cptr[s] <= r) continue;
        int c = clist[cptr[s] + r];
        int kc = S.sn_first[c + 1] - S.sn_first[c];
        int mu = (int)(S.sn_rowptr[c + 1] - S.sn_rowptr[c] - kc);
        int j = 0;
        while (j < mu) {
          int j1 = j;
          long long work = 0;
          while (j1 < mu && work < EXT_TARGET) { work += mu - j1; j1++; }
          ext.push_back(make_int4(c, j, j1, s));
          j = j1;
        }
      }
      ll.end = (long long)ext.size();
      if (ll.end > ll.begin) L.ext_rounds.push_back(ll);
    }
    // factor steps: 64-column panels; inner updates stay inside the 256-column outer block, the outer update
    // of everything to its right follows the outer block's last panel
    int maxk = 0;
    for (int s : bylevel[l]) maxk = std::max(maxk, S.sn_first[s + 1] - S.sn_first[s]);
    for (int c0 = 0; c0 < maxk; c0 += NB) {
      LaunchList lp, ls, lo, lpo;
      lp.begin = (long long)pan.size();
      lpo.begin = (long long)pot.size();
      lpo.c0 = c0;
      ls.begin = (long long)syr.size();
      lp.c0 = ls.c0 = c0;
      double fl = 0, flo = 0;
      int nslot = 0;
      const int ob0 = (c0 / OB) * OB;
      const bool outer_now = (c0 + NB) % OB == 0 || c0 + NB >= maxk;
      for (int s : bylevel[l]) {
        int k = S.sn_first[s + 1] - S.sn_first[s];
        if (k <= c0) continue;
        long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
        int w = std::min(NB, k - c0);
        long long rb = c0 + w;
        long long nrows = m - rb;
        pot.push_back(make_int4(s, nslot, 0, 0));
        for (long long r0 = rb; r0 < m; r0 += 64)
          pan.push_back(make_int4(s, (int)r0, (int)std::min<long long>(m, r0 + 64), nslot));
        nslot++;
        // inner update: columns [rb, min(k, ob0 + OB))
        long long cend = std::min<long long>(k, ob0 + OB);
        if (cend > rb) {
          long long ti = (nrows + 63) / 64, tj = (cend - rb + 63) / 64;
          for (long long I = 0; I < ti; I++)
            for (long long J = 0; J <= std::min(I, tj - 1); J++) syr.push_back(make_int4(s, (int)I, (int)J, 0));
          for (long long j = rb; j < cend; j++) fl += 2.0 * w * (double)(m - j);
        }
      }
      lp.end = (long long)pan.size();
      lpo.end = (long long)pot.size();
      ls.end = (long long)syr.size();
      L.panel_steps.push_back(lp);
      L.potrf_steps.push_back(lpo);
      L.syrk_steps.push_back(ls);
      L.syrk_flops.push_back(fl);
      max_slots = std::max(max_slots, nslot);
      lo.begin = (long long)syr.size();
      lo.c0 = ob0;
      if (outer_now) {
        for (int s : bylevel[l]) {
          int k = S.sn_first[s + 1] - S.sn_first[s];
          if (k <= ob0) continue;
          // the outer update of this front happens once, after its last panel inside [ob0, ob0 + OB)
          int last_c0 = ((std::min(k, ob0 + OB) - 1) / NB) * NB;
          if (last_c0 > c0) continue;
          long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
          int kw = std::min(OB, k - ob0);
          long long rb = ob0 + kw;
          long long nrows = m - rb;
          if (nrows <= 0) continue;
          long long tiles = (nrows + 127) / 128;
          for (long long I = 0; I < tiles; I++)
            for (long long J = 0; J <= I; J++) syr.push_back(make_int4(s, (int)I, (int)J, 0));
          flo += (double)kw * (double)nrows * (double)(nrows + 1);
        }
      }
      lo.end = (long long)syr.size();
      L.outer_steps.push_back(lo);
      L.outer_flops.push_back(flo);
    }
  }
  // sweep schedule: narrow fronts (k <= BIG_K) one workgroup each, wide fronts in 128-column steps
  std::vector<int> swp;
  std::vector<int4> fwt, bwt;
  for (int l = 0; l < nl; l++) {
    Level& L = levels_[l];
    std::vector<int> small, big;
    for (int s : bylevel[l]) {
      long long k = S.sn_first[s + 1] - S.sn_first[s], m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
      (k > BIG_K || k * m > BIG_PANEL ? big : small).push_back(s);
    }
    L.small_begin = (long long)swp.size();
    for (int s : small) {
      swp.push_back(s);
      L.small_max_m = std::max<long long>(L.small_max_m, S.sn_rowptr[s + 1] - S.sn_rowptr[s]);
    }
    L.small_end = (long long)swp.size();
    L.big_begin = (long long)swp.size();
    swp.insert(swp.end(), big.begin(), big.end());
    L.big_end = (long long)swp.size();
    int maxk = 0;
    for (int s : big) maxk = std::max(maxk, S.sn_first[s + 1] - S.sn_first[s]);
    for (int c0 = 0; c0 < maxk; c0 += SB) {
      LaunchList ll;
      ll.begin = (long long)fwt.size();
      ll.c0 = c0;
      for (int s : big) {
        int k = S.sn_first[s + 1] - S.sn_first[s];
        if (k <= c0) continue;
        long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
        long long r0 = c0 + std::min(SB, k - c0);
        long long chunks = std::max<long long>(1, (m - r0 + SWEEP_ROWS - 1) / SWEEP_ROWS);
        for (long long c = 0; c < chunks; c++) {
          long long a = r0 + c * SWEEP_ROWS, b = std::min<long long>(m, a + SWEEP_ROWS);
          fwt.push_back(make_int4(s, (int)a, (int)std::max(a, b), c == 0 ? 1 : 0));
        }
      }
      ll.end = (long long)fwt.size();
      L.fwd_steps.push_back(ll);
    }
    L.bwd_gemv.begin = (long long)bwt.size();
    for (int s : big) {
      int k = S.sn_first[s + 1] - S.sn_first[s];
      for (int q = 0; q < k; q += 16) bwt.push_back(make_int4(s, q, std::min(k, q + 16), 0));
    }
    L.bwd_gemv.end = (long long)bwt.size();
    for (int c0 = ((maxk - 1) / SB) * SB; maxk > 0 && c0 >= 0; c0 -= SB) {
      LaunchList ll;
      ll.begin = (long long)bwt.size();
      ll.c0 = c0;
      for (int s : big) {
        int k = S.sn_first[s + 1] - S.sn_first[s];
        if (k <= c0) continue;
        long long chunks = std::max<long long>(1, (c0 + SWEEP_COLS - 1) / SWEEP_COLS);
        for (long long c = 0; c < chunks; c++) {
          long long a = c * SWEEP_COLS, b = std::min<long long>(c0, a + SWEEP_COLS);
          bwt.push_back(make_int4(s, (int)a, (int)std::max(a, b), c == 0 ? 1 : 0));
        }
      }
      ll.end = (long long)bwt.size();
      L.bwd_steps.push_back(ll);
    }
  }
  // upload
  sweep_sn_.upload(swp, stream_);
  if (fwt.empty()) fwd_tasks_.alloc(1); else fwd_tasks_.upload(fwt, stream_);
  if (bwt.empty()) bwd_tasks_.alloc(1); else bwd_tasks_.upload(bwt, stream_);
  perm_.upload(S.perm, stream_);
  sn_first_.upload(S.sn_first, stream_);
  sn_rowptr_.upload(S.sn_rowptr, stream_);
  sn_rows_.upload(S.sn_rows, stream_);
  Loff_.upload(S.Loff, stream_);
  Uoff_.upload(S.Uoff, stream_);
  Voff_.upload(S.Voff, stream_);
  Woff_.upload(Woff, stream_);
  Ioff_.upload(Ioff, stream_);
  Linv_.alloc(std::max<long long>(1, Ioff[ns]));
  rel_off_.upload(S.rel_off, stream_);
  if (S.rel.empty()) rel_.alloc(1); else rel_.upload(S.rel, stream_);
  child_ptr_.upload(cptr, stream_);
  if (clist.empty()) child_list_.alloc(1); else child_list_.upload(clist, stream_);
  amap_src_.upload(S.amap_src, stream_);
  amap_dst_.upload(S.amap_dst, stream_);
  if (ext.empty()) ext_tasks_.alloc(1); else ext_tasks_.upload(ext, stream_);
  if (pan.empty()) panel_tasks_.alloc(1); else panel_tasks_.upload(pan, stream_);
  if (syr.empty()) syrk_tasks_.alloc(1); else syrk_tasks_.upload(syr, stream_);
  if (zer.empty()) zero_tasks_.alloc(1); else zero_tasks_.upload(zer, stream_);
  if (pot.empty()) potrf_tasks_.alloc(1); else potrf_tasks_.upload(pot, stream_);
  level_sn_.upload(lvl_sn, stream_);
  dscratch_.alloc((size_t)max_slots * NB * NB);
  V_.alloc(std::max<long long>(1, S.Vsize));
  W_.alloc(std::max<long long>(1, Woff[ns]));
  y_.alloc(n);
  err_.alloc(1);
  sd_ = SymDev{sn_first_.p, sn_rowptr_.p, sn_rows_.p, Loff_.p, Uoff_.p, Voff_.p, Woff_.p, rel_off_.p, rel_.p,
               child_ptr_.p, child_list_.p, Ioff_.p};
  IFEA_HIP(hipStreamSynchronize(stream_));
  // host copies not needed after upload
  std::vector<int>().swap(S_.amap_src);
  std::vector<long long>().swap(S_.amap_dst);
  analyzed_ = true;
}

template <typename T>
int SpdDevice::factorize_t(const double* d_vals) {
  const Symbolic& S = S_;
  size_t lbytes = (size_t)std::max<long long>(1, S.Lsize) * sizeof(T);
  size_t ubytes = (size_t)std::max<long long>(1, S.Usize) * sizeof(T);
  L_.alloc(lbytes);
  U_.alloc(ubytes);
  T* L = (T*)L_.p;
  T* U = (T*)U_.p;
  IFEA_HIP(hipMemsetAsync(L, 0, lbytes, stream_));
  IFEA_HIP(hipMemsetAsync(err_.p, 0, sizeof(int), stream_));
  size_t nev = 0;
  long long na = (long long)amap_src_.n;
  if (na) hipLaunchKernelGGL(k_scatter<T>, dim3(grid_for(na, 256)), dim3(256), 0, stream_, d_vals, amap_src_.p,
                             amap_dst_.p, na, L);
  for (size_t l = 0; l < levels_.size(); l++) {
    const Level& lv = levels_[l];
    if (lv.zero.end > lv.zero.begin)
      hipLaunchKernelGGL(k_zero_u<T>, dim3((unsigned)(lv.zero.end - lv.zero.begin)), dim3(256), 0, stream_,
                         zero_tasks_.p + lv.zero.begin, sd_, U);
    for (const LaunchList& r : lv.ext_rounds)
      hipLaunchKernelGGL(k_extend<T>, dim3((unsigned)(r.end - r.begin)), dim3(256), 0, stream_,
                         ext_tasks_.p + r.begin, sd_, L, U);
    for (size_t t = 0; t < lv.panel_steps.size(); t++) {
      const LaunchList& po = lv.potrf_steps[t];
      if (po.end > po.begin)
        hipLaunchKernelGGL(k_potrf_inv<T>, dim3((unsigned)(po.end - po.begin)), dim3(64), 0, stream_,
                           potrf_tasks_.p + po.begin, po.c0, sd_, L, Linv_.p, err_.p);
      const LaunchList& p = lv.panel_steps[t];
      if (p.end > p.begin)
        hipLaunchKernelGGL(k_trsm_mfma<T>, dim3((unsigned)(p.end - p.begin)), dim3(256), 0, stream_,
                           panel_tasks_.p + p.begin, p.c0, sd_, L, Linv_.p);
      for (int mode = 0; mode < 2; mode++) {
        const LaunchList& q = mode == 0 ? lv.syrk_steps[t] : lv.outer_steps[t];
        if (q.end <= q.begin) continue;
        if (profiling_) IFEA_HIP(hipEventRecord(event(2 * nev), stream_));
        if (mode == 0)
          hipLaunchKernelGGL(k_syrk<T>, dim3((unsigned)(q.end - q.begin)), dim3(256), 0, stream_,
                             syrk_tasks_.p + q.begin, q.c0, mode, sd_, L, U);
        else
          hipLaunchKernelGGL(k_syrk_big<T>, dim3((unsigned)(q.end - q.begin)), dim3(256), 0, stream_,
                             syrk_tasks_.p + q.begin, q.c0, sd_, L, U);
        if (profiling_) {
          IFEA_HIP(hipEventRecord(event(2 * nev + 1), stream_));
          nev++;
          prof_.syrk_flops += mode == 0 ? lv.syrk_flops[t] : lv.outer_flops[t];
        }
      }
    }
  }
  IFEA_HIP(hipGetLastError());
  int herr = 0;
  IFEA_HIP(hipMemcpyAsync(&herr, err_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
  IFEA_HIP(hipStreamSynchronize(stream_));
  if (profiling_) {
    for (size_t e = 0; e < nev; e++) {
      float ms = 0;
      IFEA_HIP(hipEventElapsedTime(&ms, event(2 * e), event(2 * e + 1)));
      prof_.syrk_ms += ms;
    }
    prof_.syrk_launches += (long long)nev;
    prof_.factorizations++;
    prof_.factor_flops += S.flops;
  }
  return herr ? -4 : 0;
}

int SpdDevice::factorize_device(const double* d_vals, int precision) {
  if (!analyzed_) throw Error(-5, "Analyze must be called first.");
  hipEvent_t e0, e1;
  IFEA_HIP(hipEventCreate(&e0));
  IFEA_HIP(hipEventCreate(&e1));
  IFEA_HIP(hipEventRecord(e0, stream_));
  prec_ = precision;
  solve64_ = false;
  int rc = precision == 1 ? factorize_t<float>(d_vals) : factorize_t<double>(d_vals);
  if (rc == 0 && precision == 2) {
    long long n = std::max<long long>(1, S_.Lsize);
    L32_.alloc(n);
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, 256)), dim3(256), 0, stream_, (const double*)L_.p, L32_.p, n);
    IFEA_HIP(hipGetLastError());
  }
  IFEA_HIP(hipEventRecord(e1, stream_));
  IFEA_HIP(hipEventSynchronize(e1));
  float ms = 0;
  IFEA_HIP(hipEventElapsedTime(&ms, e0, e1));
  last_factor_ms_ = ms;
  if (profiling_) prof_.factor_ms += ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  factored_ = rc == 0;
  return rc;
}

template <typename T>
void SpdDevice::sweeps(const T* L) {
  const int cap = 8192;  // doubles of LDS per forward/backward work vector
  for (size_t l = 0; l < levels_.size(); l++) {
    const Level& lv = levels_[l];
    unsigned cnt = (unsigned)(lv.small_end - lv.small_begin);
    if (cnt) {
      size_t sh = (size_t)std::min<long long>(lv.small_max_m, cap) * sizeof(double);
      hipLaunchKernelGGL(k_fwd<T>, dim3(cnt), dim3(256), sh, stream_, sweep_sn_.p + lv.small_begin, sd_, L, y_.p,
                         V_.p, W_.p, cap);
    }
    unsigned nbig = (unsigned)(lv.big_end - lv.big_begin);
    if (nbig) {
      hipLaunchKernelGGL(k_fwdb_gather, dim3(nbig), dim3(256), 0, stream_, sweep_sn_.p + lv.big_begin, sd_, y_.p,
                         V_.p, W_.p);
      for (const LaunchList& st : lv.fwd_steps)
        hipLaunchKernelGGL(k_fwdb_step<T>, dim3((unsigned)(st.end - st.begin)), dim3(256), 0, stream_,
                           fwd_tasks_.p + st.begin, st.c0, sd_, L, Linv_.p, y_.p, W_.p);
      hipLaunchKernelGGL(k_fwdb_finish, dim3(nbig), dim3(256), 0, stream_, sweep_sn_.p + lv.big_begin, sd_, W_.p,
                         V_.p);
    }
  }
  for (size_t l = levels_.size(); l-- > 0;) {
    const Level& lv = levels_[l];
    unsigned cnt = (unsigned)(lv.small_end - lv.small_begin);
    if (cnt) {
      size_t sh = (size_t)std::min<long long>(lv.small_max_m, cap) * sizeof(double);
      hipLaunchKernelGGL(k_bwd<T>, dim3(cnt), dim3(256), sh, stream_, sweep_sn_.p + lv.small_begin, sd_, L, y_.p,
                         W_.p, cap);
    }
    if (lv.big_end > lv.big_begin) {
      hipLaunchKernelGGL(k_bwdb_xb, dim3((unsigned)(lv.big_end - lv.big_begin)), dim3(256), 0, stream_,
                         sweep_sn_.p + lv.big_begin, sd_, y_.p, V_.p);
      hipLaunchKernelGGL(k_bwdb_gemv<T>, dim3((unsigned)(lv.bwd_gemv.end - lv.bwd_gemv.begin)), dim3(256), 0,
                         stream_, bwd_tasks_.p + lv.bwd_gemv.begin, sd_, L, y_.p, V_.p, W_.p);
      for (const LaunchList& st : lv.bwd_steps)
        hipLaunchKernelGGL(k_bwdb_step<T>, dim3((unsigned)(st.end - st.begin)), dim3(256), 0, stream_,
                           bwd_tasks_.p + st.begin, st.c0, sd_, L, Linv_.p, y_.p, W_.p);
    }
  }
}

void SpdDevice::drop_graph() {
  if (sweep_exec_) (void)hipGraphExecDestroy(sweep_exec_);
  sweep_exec_ = nullptr;
  sweep_key_ = nullptr;
}

template <typename T>
void SpdDevice::solve_t(const T* L, const double* d_b, double* d_x) {
  const int n = S_.n;
  hipLaunchKernelGGL(k_perm_in, dim3(grid_for(n, 256)), dim3(256), 0, stream_, d_b, perm_.p, y_.p, n);
  if (profiling_) IFEA_HIP(hipEventRecord(event(0), stream_));
  static const bool use_graph = [] {
    const char* e = std::getenv("IFEA_SWEEP_GRAPH");
    return !(e && e[0] == '0');
  }();
  if (!use_graph) {
    sweeps<T>(L);
  } else if (sweep_key_ != (const void*)L || sweep_tsize_ != sizeof(T) || !sweep_exec_) {
    drop_graph();
    hipGraph_t g = nullptr;
    IFEA_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    sweeps<T>(L);
    IFEA_HIP(hipStreamEndCapture(stream_, &g));
    IFEA_HIP(hipGraphInstantiate(&sweep_exec_, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    sweep_key_ = (const void*)L;
    sweep_tsize_ = sizeof(T);
  }
  if (use_graph) IFEA_HIP(hipGraphLaunch(sweep_exec_, stream_));
  if (profiling_) {
    IFEA_HIP(hipEventRecord(event(1), stream_));
    IFEA_HIP(hipEventSynchronize(event(1)));
    float ms = 0;
    IFEA_HIP(hipEventElapsedTime(&ms, event(0), event(1)));
    prof_.solve_ms += ms;
    prof_.solves++;
    double b = 0;
    for (const Level& lv : levels_) b += lv.solve_bytes;
    prof_.solve_bytes += 2.0 * b * sizeof(T);
    prof_.solve_launches += 2 * (long long)levels_.size();
  }
  hipLaunchKernelGGL(k_perm_out, dim3(grid_for(n, 256)), dim3(256), 0, stream_, y_.p, perm_.p, d_x, n);
  IFEA_HIP(hipGetLastError());
}

void SpdDevice::solve_device(const double* d_b, double* d_x) {
  if (!factored_) throw Error(-5, "Factorize must be called first.");
  if (prec_ == 1) solve_t<float>((const float*)L_.p, d_b, d_x);
  else if (prec_ == 2 && !solve64_) solve_t<float>(L32_.p, d_b, d_x);
  else solve_t<double>((const double*)L_.p, d_b, d_x);
}

}  // namespace ifea
