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
#include <type_traits>

#include "spd_device.h"

namespace ifea {

constexpr int NB = 64;          // panel width
constexpr int OB = 256;         // inner block: the 64-column panel updates stay inside it
constexpr int OB2 = 512;        // outer block: the trailing matrix beyond it is updated once per OB2 columns; inside
                                // it, a mid update of OB columns brings its second half up to date
constexpr long long OUTER_GROUP = 8;  // outer-update tiles are listed in groups of this many tile rows and columns
constexpr long long OUTER_NEXT_TILES = OB2 / 128;  // tile columns of an outer update that the next block's panels read
constexpr int EXT_TARGET = 8192;
constexpr int STAGE_PER = NB * 64 / 256;  // 64x64-tile elements each of 256 threads stages (TRSM, inner update)
constexpr int BIG_K = 256;      // supernodes wider than this, or with panels larger than BIG_PANEL entries, use the
constexpr long long BIG_PANEL = 32768;  // blocked multi-workgroup triangular sweeps
constexpr int SB = 2048;        // diagonal block of the blocked sweeps (explicit inverse per block)
constexpr int SWEEP_COLS = 16;  // columns per workgroup of a backward panel step
constexpr int GEMV_COLS = 8;    // columns per workgroup of the backward GEMV (four waves, GEMV_COLS / 4 columns each)
constexpr int GEMV_RB = 4;      // quads (four rows) of each column a backward-GEMV lane loads before using them
constexpr int BT = 32;          // loads a lane issues before using them in the sweep kernels
constexpr int DIAG_WAVES = 4;   // waves of a diagonal-step workgroup (one pair of rows or columns each)
constexpr int DB = 8;           // loads of X (and of the vector) a lane issues before using them in a diagonal step
constexpr int RECT_WAVES = 16;  // waves of a forward update-vector or panel workgroup (they split the columns)
constexpr int RECT_ROWS = 256;  // rows of an FP32 forward update-vector task: four consecutive rows per lane
constexpr int RECT_SPLIT = 2048; // widest column range of one forward update-vector task
constexpr int BWD_NC = 8;       // columns a wave reduces at once in the narrow-front backward sweep
constexpr int BP_ROWS = 512;    // rows of a backward panel step loaded per batch (times 4 columns per wave)

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

// value of v in lane l of the wave (l uniform)
__device__ __forceinline__ double lane_read(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// ------------------------------------------------------------------------------------------------ kernels
template <typename T>
__global__ void k_scatter(const double* __restrict__ vals, const int* __restrict__ src, const long long* __restrict__ dst,
                          long long cnt, T* __restrict__ L) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k < cnt) L[dst[k]] = (T)vals[src[k]];
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
  // one wave per child column (four columns in flight per workgroup), each lane with four rows in flight
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = tk.y + wv; j < tk.z; j += 4) {
    const int pc = rel[j];
    T* dcol = pc < kp ? Lp + (long long)pc * mp : Up + (long long)(pc - kp) * mpu - kp;
    const T* scol = Uc + (long long)j * mu;
    for (int i0 = j + lane; i0 < mu; i0 += 256) {
      int d[4];
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int i = i0 + 64 * u;
        d[u] = i < mu ? rel[i] : -1;
        v[u] = i < mu ? scol[i] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (d[u] >= 0) dcol[d[u]] += v[u];
    }
  }
}

// Diagonal block of one 64-column step, FP64, one 256-thread workgroup: a blocked Cholesky in four 16-column panels
// (the panel's 16x16 diagonal block factored and inverted by one wave, lane i owning row i, pivots and column entries
// taken by lane reads; the rows below by forward substitution, a row per thread; the trailing triangle updated by all
// threads),
// then the blocked inverse of the factor from the panels' diagonal inverses (Linv_IJ = -Linv_II sum_K L_IK Linv_KJ,
// one column block per wave). The factor goes into the panel; the inverse is kept (Lt[c][r] = Linv[r][c]) for the
// TRSM launch that follows and for the triangular sweeps. Columns beyond the block width are the identity.
// task: (supernode s, 0, 0, 0)
constexpr int PB = 16;  // panel width of the diagonal-block Cholesky
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_inv(const int4* __restrict__ tasks, int c0, SymDev sd,
                                                   T* __restrict__ L, double* __restrict__ Linv,
                                                   int* __restrict__ err) {
  // the panel chain is the factorization's critical path and shares CUs with the outer updates of the previous
  // block: its waves take issue priority over theirs (a wave priority, not a queue priority)
  __builtin_amdgcn_s_setprio(3);
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int w = min(NB, k - c0);
  T* P = L + sd.Loff[s];
  __shared__ double D[NB][NB + 1];   // the block, then its factor (lower)
  __shared__ double Li[NB][NB + 1];  // the inverse of the factor (lower)
  __shared__ double Tb[3][PB][PB + 1];
  __shared__ double rdg[NB];         // 1 / L[j][j]
  __shared__ int bad_any;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) bad_any = 0;
  // load: thread (wave, lane) reads row `lane` of columns wave, wave + 4, ... (16 loads in flight, coalesced rows)
  {
    double v[NB / 4];
#pragma unroll
    for (int t = 0; t < NB / 4; t++) {
      const int j = wv + 4 * t, i = lane;
      v[t] = (i < w && j < w && i >= j) ? (double)P[(long long)(c0 + j) * m + c0 + i] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < NB / 4; t++) {
      const int j = wv + 4 * t, i = lane;
      D[i][j] = (i < w && j < w) ? v[t] : ((i == j) ? 1.0 : 0.0);
      Li[i][j] = 0.0;
    }
  }
  __syncthreads();
  bool bad = false;
#pragma unroll 1
  for (int jp = 0; jp < NB; jp += PB) {
    // (a) the panel's diagonal block and its inverse, in wave 0
    if (wv == 0) {
      const int i = lane;
      double r[PB];
#pragma unroll
      for (int jj = 0; jj < PB; jj++) r[jj] = (i < PB && jj <= i) ? D[jp + i][jp + jj] : 0.0;
#pragma unroll
      for (int jj = 0; jj < PB; jj++) {
        const double piv = lane_read(r[jj], jj);
        bad |= !(piv > 0.0) || !isfinite(piv);
        const double d = sqrt(piv > 0.0 ? piv : 1.0);
        const double inv = 1.0 / d;
        const double l = (i > jj) ? r[jj] * inv : ((i == jj) ? d : 0.0);
        if (i == jj) rdg[jp + jj] = inv;
        r[jj] = l;
#pragma unroll
        for (int kk = jj + 1; kk < PB; kk++) r[kk] -= l * lane_read(l, kk);
      }
      if (i < PB) {
#pragma unroll
        for (int jj = 0; jj < PB; jj++) D[jp + i][jp + jj] = jj <= i ? r[jj] : 0.0;
      }
      wave_sync_lds();
      if (i < PB) {
        // lane c: x = L_pp^-1 e_c, x[q] = (delta_qc - sum_{p<q} L[q][p] x[p]) / L[q][q]
        const int c = i;
        double x[PB];
#pragma unroll
        for (int q = 0; q < PB; q++) {
          double a0 = (q == c) ? 1.0 : 0.0, a1 = 0.0;
#pragma unroll
          for (int p2 = 0; p2 + 1 < q; p2 += 2) {
            a0 -= D[jp + q][jp + p2] * x[p2];
            a1 -= D[jp + q][jp + p2 + 1] * x[p2 + 1];
          }
          if (q & 1) a0 -= D[jp + q][jp + q - 1] * x[q - 1];
          x[q] = (a0 + a1) / D[jp + q][jp + q];
        }
#pragma unroll
        for (int q = 0; q < PB; q++) Li[jp + q][jp + c] = x[q];
      }
    }
    __syncthreads();
    const int below = NB - jp - PB;  // rows under the panel's diagonal block
    if (below > 0) {
      // (b) the rows below the panel's diagonal block by forward substitution, one row per thread (the accumulation
      // order of a right-looking column step): X[i][c] = (A[i][c] - sum_{q<c} X[i][q] L[c][q]) / L[c][c]
      if (tid < below) {
        const int i = jp + PB + tid;
        double x[PB];
#pragma unroll
        for (int c = 0; c < PB; c++) {
          double a = D[i][jp + c];
#pragma unroll
          for (int q = 0; q < c; q++) a -= x[q] * D[jp + c][jp + q];
          x[c] = a * rdg[jp + c];
        }
#pragma unroll
        for (int c = 0; c < PB; c++) D[i][jp + c] = x[c];
      }
      __syncthreads();
      // (c) trailing triangle: D[i][j] -= sum_q X[i][q] X[j][q], jp + 16 <= j <= i < 64
      const int nt = below * (below + 1) / 2;
      for (int e = tid; e < nt; e += 256) {
        // e -> (a, b) with 0 <= b <= a < below, row-major over the lower triangle
        int a = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while ((a + 1) * (a + 2) / 2 <= e) a++;
        while (a * (a + 1) / 2 > e) a--;
        const int b2 = e - a * (a + 1) / 2;
        const int i = jp + PB + a, j = jp + PB + b2;
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int q = 0; q < PB; q += 2) {
          s0 += D[i][jp + q] * D[j][jp + q];
          s1 += D[i][jp + q + 1] * D[j][jp + q + 1];
        }
        D[i][j] -= s0 + s1;
      }
      __syncthreads();
    }
  }
  if (bad) atomicOr(&bad_any, 1);
  // blocked inverse: wave J fills column block J below the diagonal, one block row at a time
  if (wv < 3) {
    const int J = wv;
#pragma unroll 1
    for (int I = J + 1; I < NB / PB; I++) {
      double tv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = lane + 64 * u, r = e / PB, c = e % PB;
        double a0 = 0.0, a1 = 0.0;
        for (int K = J; K < I; K++) {
#pragma unroll
          for (int q = 0; q < PB; q += 2) {
            a0 += D[PB * I + r][PB * K + q] * Li[PB * K + q][PB * J + c];
            a1 += D[PB * I + r][PB * K + q + 1] * Li[PB * K + q + 1][PB * J + c];
          }
        }
        tv[u] = a0 + a1;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = lane + 64 * u;
        Tb[J][e / PB][e % PB] = tv[u];
      }
      wave_sync_lds();
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = lane + 64 * u, r = e / PB, c = e % PB;
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int q = 0; q < PB; q += 2) {
          if (q <= r) a0 += Li[PB * I + r][PB * I + q] * Tb[J][q][c];
          if (q + 1 <= r) a1 += Li[PB * I + r][PB * I + q + 1] * Tb[J][q + 1][c];
        }
        Li[PB * I + r][PB * J + c] = -(a0 + a1);
      }
      wave_sync_lds();
    }
  }
  __syncthreads();
  if (tid == 0 && bad_any) atomicOr(err, 1);
  // factor into the panel (column q, row i) and the inverse transposed, Lt[c][r] = Linv[r][c], both row-coalesced
  double* Lt = Linv + sd.Ioff[s] + (long long)(c0 / NB) * NB * NB;
#pragma unroll
  for (int t = 0; t < NB / 4; t++) {
    const int q = wv + 4 * t, i = lane;
    if (i < w && q < w && q <= i) P[(long long)(c0 + q) * m + c0 + i] = (T)D[i][q];
    Lt[q * NB + i] = i >= q ? Li[i][q] : 0.0;
  }
}

// Rows below the diagonal block: X = A L11^{-T} as a 64x64x64 FP64 MFMA product per task, written back into the
// panel. task: (supernode s, first row, end row (<= first + 64), 0)
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_mfma(const int4* __restrict__ tasks, int c0, SymDev sd,
                                                   T* __restrict__ L, const double* __restrict__ Linv) {
  __builtin_amdgcn_s_setprio(3);  // a panel-chain kernel: issue priority, as in k_potrf_inv
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
  {
    // all of a thread's loads in flight before its LDS writes (256 threads, STAGE_PER elements of each tile)
    double va[STAGE_PER], vl[STAGE_PER];
#pragma unroll
    for (int t = 0; t < STAGE_PER; t++) {
      const int idx = threadIdx.x + 256 * t, p = idx >> 6, rr = idx & 63;
      va[t] = (p < w && rr < nr) ? (double)P[(long long)(c0 + p) * m + r0 + rr] : 0.0;
      // Lt is stored [c][r] = Linv[r][c]; sL[p][q] = Linv[q][p] = Lt[p][q]
      vl[t] = (p < w && rr < w) ? Lt[p * NB + rr] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < STAGE_PER; t++) {
      const int idx = threadIdx.x + 256 * t, p = idx >> 6, rr = idx & 63;
      sA[p][rr] = va[t];
      sL[p][rr] = vl[t];
    }
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

// Inner trailing update C -= X X^T of a front, 64x64 output tile per workgroup, MFMA (FP64 16x16x4 or FP32
// 16x16x4): X = the 64 columns [kb, kb+w) just factored; C = columns [kb+w, min(k, inner_end)) of the same OB-column
// inner block, all rows below (delayed right-looking update inside the inner block).
// task: (supernode s, row tile I, column tile J, 0), I >= J.
template <typename T>
__global__ __launch_bounds__(256) void k_syrk(const int4* __restrict__ tasks, int kb, SymDev sd,
                                              T* __restrict__ L) {
  __builtin_amdgcn_s_setprio(3);  // a panel-chain kernel: issue priority, as in k_potrf_inv
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int w = min(NB, k - kb);
  const long long rb = kb + w;
  const long long col_end = min((long long)k, (long long)(kb / OB) * OB + OB);
  const T* P = L + sd.Loff[s];
  __shared__ T Xi[NB][64 + 4];
  __shared__ T Xj[NB][64 + 4];
  const long long i0 = rb + 64LL * tk.y, j0 = rb + 64LL * tk.z;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ib = (wv >> 1) * 32, jb = (wv & 1) * 32;
  const int l16 = lane & 15;
  typedef typename Mfma<T>::v4 v4;
  v4 acc00 = {0, 0, 0, 0}, acc01 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};
  // the lane's 16 output entries (all in the front's own columns: col_end <= k) are loaded before the MFMAs, so
  // their latency overlaps the product; each entry belongs to one tile of the launch
  T* Ls = L + sd.Loff[s];
  T cv[2][2][4];
#pragma unroll
  for (int ai = 0; ai < 2; ai++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long long j = j0 + jb + 16 * ai + Mfma<T>::row(lane, r);
        const long long i = i0 + ib + 16 * bi + l16;
        cv[ai][bi][r] = (i < m && j < col_end && i >= j) ? Ls[j * m + i] : T(0);
      }
  {
    // all of a thread's loads in flight before its LDS writes (256 threads, STAGE_PER elements of each operand)
    T xi[STAGE_PER], xj[STAGE_PER];
#pragma unroll
    for (int t = 0; t < STAGE_PER; t++) {
      const int idx = threadIdx.x + 256 * t, q = idx >> 6, rr = idx & 63;
      const long long ri = i0 + rr, rj = j0 + rr;
      const T* cp = P + (long long)(kb + q) * m;
      xi[t] = (q < w && ri < m) ? cp[ri] : T(0);
      xj[t] = (q < w && rj < m) ? cp[rj] : T(0);
    }
#pragma unroll
    for (int t = 0; t < STAGE_PER; t++) {
      const int idx = threadIdx.x + 256 * t, q = idx >> 6, rr = idx & 63;
      Xi[q][rr] = xi[t];
      Xj[q][rr] = xj[t];
    }
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
  // D[a][b] = sum_q Xj[a][q] Xi[b][q] = contribution to C(i = i0+ib+16bi+b, j = j0+jb+16ai+a)
  v4 accs[2][2] = {{acc00, acc01}, {acc10, acc11}};
#pragma unroll
  for (int ai = 0; ai < 2; ai++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long long j = j0 + jb + 16 * ai + Mfma<T>::row(lane, r);
        const long long i = i0 + ib + 16 * bi + l16;
        if (i < m && j < col_end && i >= j) Ls[j * m + i] = cv[ai][bi][r] - accs[ai][bi][r];
      }
}


// Outer and mid trailing updates with 128x128 output tiles: NT/64 waves, two along the rows (64 each) and
// NT/128 along the columns, each wave a block of 16x16 MFMA tiles (64x32 with 8 waves). X is streamed through LDS in
// k-major chunks of 256 rows (the tile's 128 rows, then its 128 columns). FP64: 8-column stages through four LDS
// slots, filled by loads straight into LDS three stages ahead (below). FP32: 16-column chunks through two slots, the
// loads two chunks ahead in registers: those of chunk c + 2 are issued before the MFMAs of chunk c and written to LDS
// after the MFMAs of chunk c + 1, one barrier per chunk. Both are buffer loads: the front's block
// column kb is the descriptor base, a thread's row is its one 32-bit VGPR offset and the chunk's column a wave-uniform
// SGPR offset, so no load needs a 64-bit address; a load that must read zero (a row beyond the front, a column
// beyond the block) takes an offset past the descriptor's range, for which the hardware returns 0. Register use stays
// at or below 128 VGPRs so that two 8-wave workgroups share a CU and one workgroup's MFMAs run while the other waits
// at its barrier; the tile is subtracted by atomic adds in L2 (one update per entry per launch, so the result does
// not depend on their order). task: (supernode s, row tile I, column tile J, 0), I >= J. Fronts are below 2^19 rows
// (checked at analysis), so the byte offsets within a block of at most OB2 columns fit in 31 bits.
// MID = false (outer): X = the outer block [kb, kb + min(OB2, k - kb)), C = every column to its right (L and U).
// MID = true: X = the first OB columns [kb, kb + OB) of an outer block, C = its second half [kb + OB, min(k, kb +
// OB2)), all rows below; the panels of the second half are then factored against a complete update.
template <typename T>
__device__ __forceinline__ T buffer_ld(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}

template <typename T, int NT, int KC, bool MID>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT / 128))) void k_syrk_big(const int4* __restrict__ tasks, int kb, SymDev sd,
                                                 T* __restrict__ L, T* __restrict__ U) {
  static_assert(NT % 256 == 0, "a thread stages one of the 256 rows");
  constexpr int PER = KC * 256 / NT;   // loads of a thread per chunk: one row, every (NT/256)-th column
  constexpr int CS = NT / 256;         // column stride of a thread's loads
  constexpr int WJ = NT / 128;         // waves along the columns
  constexpr int BJ = 128 / WJ / 16;    // 16x16 blocks of a wave along the columns (rows: 4)
  constexpr int LD = 272;              // LDS row stride: consecutive k rows start 32 banks apart (no conflicts)
  constexpr int OOB = (int)0x80000000u;  // offset past any descriptor range: the load returns 0
  // workgroups are dispatched round-robin over the 8 XCDs: give each XCD a contiguous range of the task list (tiles
  // of the same tile row and neighbouring columns), so the panel rows they share stay in that XCD's L2
  const int nt = gridDim.x, xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int per_x = nt >> 3, extra = nt & 7;
  const int4 tk = tasks[xcd * per_x + min(xcd, extra) + slot];
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const int m = (int)sn_m(sd, s);
  const int kw = min(MID ? OB : OB2, k - kb);
  const int rb = kb + kw;
  const int cend = MID ? min(k, kb + OB2) : m;      // columns updated: [rb, cend)
  const T* P = L + sd.Loff[s] + (long long)kb * m;  // column kb of the front
  const int i0 = rb + 128 * tk.y, j0 = rb + 128 * tk.z;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ib = (wv & 1) * 64, jb = 128 + (wv >> 1) * (128 / WJ);
  const int l16 = lane & 15;
  const int nch = (kw + KC - 1) / KC;
  // a thread stages row vr of the chunk (vr < 128: tile row i0 + vr, else tile column j0 + vr - 128) in columns
  // q0, q0 + CS, ...
  const int vr = threadIdx.x & 255, q0 = threadIdx.x >> 8;
  const int grow = vr < 128 ? i0 + vr : j0 + (vr - 128);
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void*)P, 0, kw * m * (int)sizeof(T), 0x00020000);
  const int vo = grow < m ? (q0 * m + grow) * (int)sizeof(T) : OOB;
  typedef typename Mfma<T>::v4 v4;
  v4 acc[BJ][4];
#pragma unroll
  for (int a = 0; a < BJ; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = v4{0, 0, 0, 0};
  auto compute = [&](const T (*Xb)[LD]) {
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int q = kk + (lane >> 4);
      T av[BJ], bv[4];
#pragma unroll
      for (int a = 0; a < BJ; a++) av[a] = Xb[q][jb + 16 * a + l16];
#pragma unroll
      for (int b = 0; b < 4; b++) bv[b] = Xb[q][ib + 16 * b + l16];
#pragma unroll
      for (int a = 0; a < BJ; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = Mfma<T>::op(av[a], bv[b], acc[a][b]);
    }
  };
  // the chunk pipeline, compiled twice: interior tiles (both 128-row panels inside the front, every chunk full) skip
  // the per-column checks; the choice is uniform over the workgroup. Chunks past the last one load zeros (the
  // pipeline's body has no branches; an odd chunk count adds one chunk of zeros)
  if constexpr (sizeof(T) == 8 && NT == 512 && KC == 16) {
    // FP64: 16-byte buffer loads straight into LDS (no staging registers, no LDS stores), through four LDS stages of
    // 8 k-rows. Lane l of a load brings tile rows 2l and 2l + 1 of one segment of a k-row (the tile's 128 rows, or its
    // 128 columns), so one load fills a 1 KB segment; wave wv loads k-row wv of a stage, both segments. The hardware
    // writes zeros for offsets past the descriptor's range: k-rows past the block take such an offset. A stage's
    // operands are read from LDS first, then the loads of the stage three ahead are issued, then its MFMAs run; each
    // stage ends with the wave's loads of the next stage complete (the two stages after it may still be in flight)
    // and one barrier.
    constexpr int KD = 8;
    const int ncd = (kw + KD - 1) / KD;
    __shared__ T X0[KD][LD], X1[KD][LD], X2[KD][LD], X3[KD][LD];
    const int vo0 = (i0 + 2 * lane) * 8, vo1 = (j0 + 2 * lane) * 8;
    const int wvu = __builtin_amdgcn_readfirstlane(wv);
    auto issue = [&](T (*Xb)[LD], int c) {
      const int col = c * KD + wvu;
      const bool live = col < kw;
      const int so = __builtin_amdgcn_readfirstlane(live ? col * m * 8 : 0);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)&Xb[wvu][0], 16,
                                               live ? vo0 : OOB, so, 0, 0);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)&Xb[wvu][128], 16,
                                               live ? vo1 : OOB, so, 0, 0);
    };
    // the wave's loads done but the last four (two stages of two), its LDS reads done, then the barrier; written out
    // because a fenced barrier would wait for every load in flight
    auto stage_ready = [&]() { asm volatile("s_waitcnt vmcnt(4) lgkmcnt(0)\n\ts_barrier" ::: "memory"); };
    auto chunk = [&](const T (*Xb)[LD], T (*Xn)[LD], int cn) {
      T av[KD / 4][BJ], bv[KD / 4][4];
#pragma unroll
      for (int t = 0; t < KD / 4; t++) {
        const int q = 4 * t + (lane >> 4);
#pragma unroll
        for (int a = 0; a < BJ; a++) av[t][a] = Xb[q][jb + 16 * a + l16];
#pragma unroll
        for (int b = 0; b < 4; b++) bv[t][b] = Xb[q][ib + 16 * b + l16];
      }
      issue(Xn, cn);
#pragma unroll
      for (int t = 0; t < KD / 4; t++)
#pragma unroll
        for (int a = 0; a < BJ; a++)
#pragma unroll
          for (int b = 0; b < 4; b++) acc[a][b] = Mfma<T>::op(av[t][a], bv[t][b], acc[a][b]);
      stage_ready();
    };
    issue(X0, 0);
    issue(X1, 1);
    issue(X2, 2);
    stage_ready();
    for (int c = 0; c < ncd; c += 4) {
      chunk(X0, X3, c + 3);
      chunk(X1, X0, c + 4);
      chunk(X2, X1, c + 5);
      chunk(X3, X2, c + 6);
    }
  } else {
  __shared__ T X[2][KC][LD];
  auto pipeline = [&](auto checked) {
    constexpr bool CHK = decltype(checked)::value;
    auto fetch = [&](T* r, int c) {
      const bool live = c < nch;
      const int cb = live ? c * KC * m * (int)sizeof(T) : 0;
#pragma unroll
      for (int t = 0; t < PER; t++) {
        const int v = CHK ? (c * KC + q0 + CS * t < kw ? vo : OOB) : (live ? vo : OOB);
        r[t] = buffer_ld<T>(rs, v, cb + CS * t * m * (int)sizeof(T));
      }
    };
    auto stage = [&](const T* r, int buf) {
#pragma unroll
      for (int t = 0; t < PER; t++) X[buf][q0 + CS * t][vr] = r[t];
    };
    T ra[PER], rc[PER];
    fetch(ra, 0);
    stage(ra, 0);
    fetch(rc, 1);
    __syncthreads();
    for (int c = 0; c < nch; c += 2) {
      fetch(ra, c + 2);
      compute(X[0]);
      stage(rc, 1);
      __syncthreads();
      fetch(rc, c + 3);
      compute(X[1]);
      stage(ra, 0);
      __syncthreads();
    }
  };
  if (i0 + 128 <= m && j0 + 128 <= m && kw % KC == 0) pipeline(std::false_type{});
  else pipeline(std::true_type{});
  }
  T* Ls = L + sd.Loff[s];
  T* Us = U + sd.Uoff[s];
  const long long mu = m - k;
#pragma unroll
  for (int a = 0; a < BJ; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        long long j = j0 + (jb - 128) + 16 * a + Mfma<T>::row(lane, r);
        long long i = i0 + ib + 16 * b + l16;
        if (i < m && j < cend && i >= j) {
          // each entry of the trailing triangle belongs to exactly one tile of the launch: a hardware atomic add in
          // L2 gives the same bits as load, subtract and store, without the tile waiting for the load. The first
          // outer update of a front (kb = 0) is the first write to its update block: a store.
          if (!MID && kb == 0 && j >= k) {
            Us[(j - k) * mu + (i - k)] = -acc[a][b][r];
          } else {
            T* cp = (j < k) ? (Ls + j * m + i) : (Us + (j - k) * mu + (i - k));
            unsafeAtomicAdd(cp, -acc[a][b][r]);
          }
        }
      }
}

// Narrow fronts, one workgroup each, in 64-column steps from the kept inverse diagonal blocks: per step one parallel
// 64x64 product x_t = Linv_t w_t (waves split the columns, partial sums combined in a fixed order) and one parallel
// update of the rows below, so the dependent chain is ceil(k/64) steps long, not k columns.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd(const int* __restrict__ sns, SymDev sd, const T* __restrict__ L,
                                             const double* __restrict__ Linv, double* __restrict__ y,
                                             double* __restrict__ V, double* __restrict__ Wg, int cap) {
  extern __shared__ double shw[];
  __shared__ double part[4][NB], xs[NB];
  const int s = sns[blockIdx.x];
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  double* w = (m <= cap) ? shw : Wg + sd.Woff[s];
  const int tid = threadIdx.x, bd = blockDim.x, lane = tid & 63, wv = tid >> 6;
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
  for (int c0 = 0; c0 < k; c0 += NB) {
    const int wd = min(NB, k - c0);
    const double* Lt = Linv + sd.Ioff[s] + (long long)(c0 / NB) * NB * NB;
    // x_r = sum_c Linv[r][c] w_c = sum_c Lt[c][r] w_c; wave wv takes c in [16 wv, 16 wv + 16)
    double a = 0.0;
#pragma unroll 4
    for (int c = 16 * wv; c < 16 * wv + 16; c++)
      if (c < wd) a += Lt[c * NB + lane] * w[c0 + c];
    part[wv][lane] = a;
    __syncthreads();
    if (tid < NB) {
      const double x = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
      xs[tid] = tid < wd ? x : 0.0;
      if (tid < wd) y[f + c0 + tid] = x;
    }
    __syncthreads();
    for (long long i = c0 + wd + tid; i < m; i += bd) {
      const T* pr = P + (long long)c0 * m + i;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int q = 0;
      for (; q + 4 <= wd; q += 4) {
        a0 += (double)pr[(long long)q * m] * xs[q];
        a1 += (double)pr[(long long)(q + 1) * m] * xs[q + 1];
        a2 += (double)pr[(long long)(q + 2) * m] * xs[q + 2];
        a3 += (double)pr[(long long)(q + 3) * m] * xs[q + 3];
      }
      for (; q < wd; q++) a0 += (double)pr[(long long)q * m] * xs[q];
      w[i] -= (a0 + a1) + (a2 + a3);
    }
    __syncthreads();
  }
  double* vs = V + sd.Voff[s];
  for (long long i = k + tid; i < m; i += bd) vs[i - k] = w[i];
}

// Backward for narrow fronts. w[0, k) = z, w[k, m) = the solution at the front's lower rows (gathered once); then
// z = y_own - L21^T x_below and descending 64-column steps x_t = Linv_t^T z_t, z_j -= L(c0.., j)^T x_t for the
// columns j < c0. Every product is a column dot over lanes; a wave keeps BWD_NC columns (16 for the 64x64 inverse)
// in flight at once and reduces them together, so the dependent chain is a few batches, not one column at a time.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd(const int* __restrict__ sns, SymDev sd, const T* __restrict__ L,
                                             const double* __restrict__ Linv, double* __restrict__ y,
                                             double* __restrict__ Wg, int cap) {
  extern __shared__ double shz[];
  __shared__ double xs[NB];
  constexpr int NC = BWD_NC;
  const int s = sns[blockIdx.x];
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const int m = (int)sn_m(sd, s);  // narrow: k * m <= BIG_PANEL
  const int* rows = sd.sn_rows + sd.sn_rowptr[s];
  double* w = (m <= cap) ? shz : Wg + sd.Woff[s];
  const T* P = L + sd.Loff[s];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = k + tid; i < m; i += 256) w[i] = y[rows[i]];
  __syncthreads();
  for (int qb = NC * wv; qb < k; qb += 4 * NC) {
    const T* cq[NC];
    double acc[NC];
#pragma unroll
    for (int u = 0; u < NC; u++) {
      cq[u] = P + (long long)min(qb + u, k - 1) * m;  // columns past the front repeat its last one, not stored
      acc[u] = 0.0;
    }
    for (int i = k + lane; i < m; i += 64) {
      const double xv = w[i];
#pragma unroll
      for (int u = 0; u < NC; u++) acc[u] += (double)cq[u][i] * xv;
    }
#pragma unroll
    for (int u = 0; u < NC; u++) {
      const double a = wave_sum(acc[u]);
      if (lane == 0 && qb + u < k) w[qb + u] = y[f + qb + u] - a;
    }
  }
  __syncthreads();
  for (int c0 = ((k - 1) / NB) * NB; c0 >= 0; c0 -= NB) {
    const int wd = min(NB, k - c0);
    const double* Lt = Linv + sd.Ioff[s] + (long long)(c0 / NB) * NB * NB;
    const double zl = lane < wd ? w[c0 + lane] : 0.0;
    {
      // x_c = sum_r Linv[r][c] z_r = sum_r Lt[c][r] z_r for the wave's 16 columns c
      double a[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int c = 16 * wv + u;
        a[u] = (c < wd && lane < wd) ? Lt[c * NB + lane] * zl : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const double t = wave_sum(a[u]);
        if (lane == 0) xs[16 * wv + u] = t;
      }
    }
    __syncthreads();
    if (tid < wd) y[f + c0 + tid] = xs[tid];
    const double xl = lane < wd ? xs[lane] : 0.0;
    for (int jb = NC * wv; jb < c0; jb += 4 * NC) {
      double a[NC];
#pragma unroll
      for (int u = 0; u < NC; u++) {
        const int j = min(jb + u, c0 - 1);
        a[u] = lane < wd ? (double)P[(long long)j * m + c0 + lane] * xl : 0.0;
      }
#pragma unroll
      for (int u = 0; u < NC; u++) {
        const double t = wave_sum(a[u]);
        if (lane == 0 && jb + u < c0) w[jb + u] -= t;
      }
    }
    __syncthreads();
  }
}


// ---------------------------------------------------------------------------------------- blocked sweeps (wide fronts)
// Wide fronts are swept in steps of SB = 2048 columns. Each step is two launches: a diagonal launch x_t = X_t w_t with
// X_t the explicit inverse of the step's wd x wd diagonal block (wd = min(SB, k - c0); built once per factorization
// from the kept 64x64 inverses, k_inv_diag), and a panel launch that applies the step to the front's later own rows
// (forward) or earlier own columns (backward) with many workgroups. X_t is stored column by column with leading
// dimension wd, X[c * wd + r] = inv(L_tt)[r][c], and XT holds its transpose; every step but a front's last is full,
// so the step at c0 starts SB * SB * (c0 / SB) entries into the front's blocks.

__device__ __forceinline__ long long x_step_off(const SymDev& sd, int s, int c0) {
  return sd.Xoff[s] + (long long)(c0 / SB) * SB * SB;
}

// Inverse of one SB-column diagonal block, one 64-column strip j per workgroup: X_jj = Linv_j and, going down,
// X_ij = -Linv_i * sum_{p=j}^{i-1} L_ip X_pj. The 64x64x64 products run on FP64 MFMA: wave w owns rows
// [16 w, 16 w + 16) of the product and its four 16-column blocks; the operand tiles are staged in LDS, the next
// pair's loads issued before the current pair's MFMAs. Results go through LDS so that both copies (X column-major,
// XT row-major) are stored row-coalesced. task: (s, c0, j, 0). Strips above the diagonal are stored as zeros.
template <typename T>
__global__ __launch_bounds__(256) void k_inv_diag(const int4* __restrict__ tasks, SymDev sd, const T* __restrict__ L,
                                                const double* __restrict__ Linv, double* __restrict__ X,
                                                double* __restrict__ XT) {
  __shared__ double As[NB][NB + 1], Bs[NB][NB + 1];
  const int4 tk = tasks[blockIdx.x];
  const int s = tk.x, c0 = tk.y, j = tk.z;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int wd = min(SB, k - c0);
  const int nsub = (wd + NB - 1) / NB;
  const T* P = L + sd.Loff[s];
  const long long ld = wd;
  double* Xb = X + x_step_off(sd, s, c0);
  double* XTb = XT + x_step_off(sd, s, c0);
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int lane = r, w = q, l16 = lane & 15, kq = lane >> 4;
  typedef Mfma<double>::v4 v4;
  const double* Ltj = Linv + sd.Ioff[s] + (long long)((c0 + NB * j) / NB) * NB * NB;
  // only entries inside the wd x wd block are stored (the last strip of a step can be narrower than 64)
  for (int cc = 0; cc < 16; cc++) {
    const int c = NB * j + 16 * q + cc;
    if (c >= wd) break;
    double* xc = Xb + (long long)c * ld;
    for (int i = 0; i < j; i++) {
      xc[NB * i + r] = 0.0;
      XTb[(long long)(NB * i + r) * ld + c] = 0.0;
    }
    if (NB * j + r < wd) {
      const double d = Ltj[(c - NB * j) * NB + r];
      xc[NB * j + r] = d;
      XTb[(long long)(NB * j + r) * ld + c] = d;
    }
  }
  __syncthreads();
  // acc += As * Bs over the 64 k of the staged tiles (wave w: rows 16 w .., all four column blocks)
  auto product = [&](v4 (&acc)[4]) {
#pragma unroll 4
    for (int kk = 0; kk < NB; kk += 4) {
      const double a = As[16 * w + l16][kk + kq];
#pragma unroll
      for (int nb = 0; nb < 4; nb++) acc[nb] = Mfma<double>::op(a, Bs[kk + kq][16 * nb + l16], acc[nb]);
    }
  };
  // item e of lane l of block nb: entry (16 w + (l >> 4) + 4 e, 16 nb + (l & 15))
  auto acc_to_lds = [&](const v4 (&acc)[4], double sign, double (*D)[NB + 1]) {
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
#pragma unroll
      for (int e = 0; e < 4; e++) D[16 * w + Mfma<double>::row(lane, e)][16 * nb + l16] = sign * acc[nb][e];
  };
  for (int i = j + 1; i < nsub; i++) {
    v4 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) acc[nb] = v4{0, 0, 0, 0};
    // operand tiles of pair p0: As = L_i,p0 (rows of strip i, columns of strip p0), Bs = X_p0,j; thread t holds
    // entries (t & 63, (t >> 6) + 4 u) of both
    double xv[16], lv[16];
    auto fetch = [&](int p0) {
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int rr = r, cq = q + 4 * u;
        xv[u] = Xb[(long long)(NB * j + cq) * ld + NB * p0 + rr];  // strips p0 < i < nsub and j < i are full
        lv[u] = NB * i + rr < wd ? (double)P[(long long)(c0 + NB * p0 + cq) * m + c0 + NB * i + rr] : 0.0;
      }
    };
    fetch(j);
    for (int p0 = j; p0 < i; p0++) {
#pragma unroll
      for (int u = 0; u < 16; u++) {
        As[r][q + 4 * u] = lv[u];
        Bs[r][q + 4 * u] = xv[u];
      }
      __syncthreads();
      if (p0 + 1 < i) fetch(p0 + 1);
      product(acc);
      __syncthreads();
    }
    // Bs = sum_p L_ip X_pj, As = Linv_i (Lti[c * NB + r] = Linv_i[r][c]); X_ij = -As * Bs
    const double* Lti = Linv + sd.Ioff[s] + (long long)((c0 + NB * i) / NB) * NB * NB;
    acc_to_lds(acc, 1.0, Bs);
#pragma unroll
    for (int u = 0; u < 16; u++) As[r][q + 4 * u] = Lti[(q + 4 * u) * NB + r];
    __syncthreads();
#pragma unroll
    for (int nb = 0; nb < 4; nb++) acc[nb] = v4{0, 0, 0, 0};
    product(acc);
    __syncthreads();
    acc_to_lds(acc, -1.0, As);
    __syncthreads();
    // X (column-major): lanes over rows; XT (row-major): lanes over columns
    if (NB * i + r < wd) {
#pragma unroll
      for (int cc = 0; cc < 16; cc++)
        Xb[(long long)(NB * j + 16 * q + cc) * ld + NB * i + r] = As[r][16 * q + cc];
    }
#pragma unroll 4
    for (int rr = 0; rr < 16; rr++) {
      const int row = 16 * q + rr;
      if (NB * i + row < wd) XTb[(long long)(NB * i + row) * ld + NB * j + r] = As[row][r];
    }
    __syncthreads();
  }
}

// Forward, first part: W_s = [y_own ; 0] plus the children's update vectors, one thread per row; the contributions
// of a row are listed in child order (row-wise inverse of the children's relative maps), so the sum is fixed-order.
// task: (s, first row, end row, 0)
__device__ __forceinline__ void fwdb_gather_item(const int4 tk, SymDev sd,
                                                     const double* __restrict__ y, const double* __restrict__ V,
                                                     double* __restrict__ Wg) {
  const int s = tk.x;
  const long long i = (long long)tk.y + threadIdx.x;
  if (i >= tk.z) return;
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long* gp = sd.gptr + sd.Goff[s];
  double w = i < k ? y[f + i] : 0.0;
  for (long long e = gp[i]; e < gp[i + 1]; e++) w += V[sd.gsrc[e]];
  Wg[sd.Woff[s] + i] = w;
}

// Forward diagonal step: x_r = sum_{c <= r} X_t[r][c] w_c. One wave per pair of rows (p, wd - 1 - p), so every wave
// reads wd + 1 entries, from the row-major copy XT_t (XT[r * wd + c] = X_t[r][c]) with the lanes over c and DB loads
// of X and of w (L1/L2-resident) in flight per lane; the wave sums in a fixed order. A step runs as ~wd / 2 waves
// spread over the whole chip. task: (s, c0, first pair, 0), DIAG_WAVES pairs per workgroup.
template <typename TX>
__device__ __forceinline__ void fwdb_diag_item(const int4 tk, SymDev sd,
                                                   const TX* __restrict__ XT, double* __restrict__ y,
                                                   const double* __restrict__ Wg) {
  const int s = tk.x, c0 = tk.y;
  const int lane = threadIdx.x & 63;
  const int p = tk.z + (threadIdx.x >> 6);
  const int k = sn_k(sd, s);
  const int wd = min(SB, k - c0);
  if (p >= (wd + 1) / 2) return;
  const int f = sd.sn_first[s];
  const TX* Xb = XT + x_step_off(sd, s, c0);
  const double* w = Wg + sd.Woff[s] + c0;
  for (int h = 0; h < 2; h++) {
    const int r = h == 0 ? p : wd - 1 - p;
    if (h == 1 && r == p) break;
    const TX* xr = Xb + (long long)r * wd;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int base = 0; base <= r; base += 64 * DB) {
      TX xv[DB];
      double wv[DB];
#pragma unroll
      for (int u = 0; u < DB; u++) {
        const int c = min(base + lane + 64 * u, r);  // entries past the row repeat its last one, weighted by zero
        xv[u] = xr[c];
        wv[u] = w[c];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < DB; u++) acc[u & 3] += base + lane + 64 * u <= r ? (double)xv[u] * wv[u] : 0.0;
    }
    const double t = wave_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if (lane == 0) y[f + c0 + r] = t;
  }
}

// acc[u % 4] += sum over columns q in [q0, q1) of pr[q * ld] * yo[q]: a lane's row of a column-major panel against
// a vector every lane of the wave shares. Full batches issue BT column loads at 32-bit offsets from a wave-uniform
// base and take yo by wave-uniform (scalar) loads; the last, partial batch clamps its columns in bounds and weights
// the extra ones by zero. q0, q1 are wave-uniform; ld < 2^27 so BT * ld fits in 32 bits.
template <typename T>
__device__ __forceinline__ void sweep_rows_dot(const T* __restrict__ pr, int ld, const double* __restrict__ yo, int q0,
                                               int q1, double acc[4]) {
  int base = q0;
  for (; base + BT <= q1; base += BT) {
    const T* pc = pr + (long long)base * ld;
    T v[BT];
#pragma unroll
    for (int u = 0; u < BT; u++) v[u] = pc[(unsigned)(u * ld)];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < BT; u++) acc[u & 3] = fma((double)v[u], yo[base + u], acc[u & 3]);
  }
  if (base < q1) {
    T v[BT];
#pragma unroll
    for (int u = 0; u < BT; u++) v[u] = pr[(long long)min(base + u, q1 - 1) * ld];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < BT; u++) {
      const double yu = base + u < q1 ? yo[min(base + u, q1 - 1)] : 0.0;
      acc[u & 3] = fma((double)v[u], yu, acc[u & 3]);
    }
  }
}

// Forward panel step: W(row) -= L(row, c0 .. c0+wd) x_t for 64 of the front's own rows below the step per
// workgroup; the waves split the step's columns (loads issued BT at a time, x_t taken by lane reads) and combine
// their partial sums in a fixed order. task: (s, c0, first row, 0)
template <typename T>
__device__ __forceinline__ void fwdb_panel_item(const int4 tk, SymDev sd,
                                                    const T* __restrict__ L, const double* __restrict__ y,
                                                    double* __restrict__ Wg) {
  __shared__ double part[RECT_WAVES][NB];
  const int s = tk.x, c0 = tk.y;
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long ld = sn_m(sd, s);
  const int wd = min(SB, k - c0);
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long row = (long long)tk.z + lane;  // own rows only: < k
  const int per = (((wd + RECT_WAVES - 1) / RECT_WAVES + BT - 1) / BT) * BT;
  const int q0 = c0 + wv * per, q1 = min(c0 + wd, q0 + per);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (q0 < q1) {
    // every lane runs the loop (the lane reads need all lanes' operands); rows past the front's own rows repeat its
    // last own row and are not stored
    const T* pr = L + sd.Loff[s] + min(row, (long long)k - 1);
    sweep_rows_dot<T>(pr, (int)ld, y + f, q0, q1, acc);
  }
  part[wv][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (threadIdx.x < NB && row < k) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < RECT_WAVES; q++) a += part[q][lane];
    Wg[sd.Woff[s] + row] -= a;
  }
}

// Forward, last part: V(row) = W(row) - L(row, 0 .. k) y_own for the update-vector rows (row >= k) of a wide front,
// after all of its diagonal steps: one launch per level with many workgroups, 64 rows each, the waves splitting a
// column range of the front and combining their partial sums in a fixed order. Fronts wider than RECT_SPLIT columns
// are cut into column ranges so that the upper levels (few row blocks, thousands of columns) fill the chip: such a
// task writes its 64 partial sums to Pp[task] and k_fwdb_rect_sum adds the ranges of a row block in range order.
// task: (s, first row, first column, end column); tid: the task's index in the level's task list.
template <typename T>
__device__ __forceinline__ void fwdb_rect_item(const int4 tk, long long tid, SymDev sd, const T* __restrict__ L,
                                                   const double* __restrict__ y, const double* __restrict__ Wg,
                                                   double* __restrict__ V, double* __restrict__ Pp) {
  __shared__ double part[RECT_WAVES][NB];
  const int s = tk.x;
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long row = (long long)tk.y + lane;
  const int qa = tk.z, qb = tk.w;
  const int per = (((qb - qa + RECT_WAVES - 1) / RECT_WAVES + BT - 1) / BT) * BT;
  const int q0 = qa + wv * per, q1 = min(qb, q0 + per);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (q0 < q1) {
    // every lane runs the loop (wave-uniform column range); rows past the front repeat its last row, not stored
    const T* pr = L + sd.Loff[s] + min(row, m - 1);
    sweep_rows_dot<T>(pr, (int)m, y + f, q0, q1, acc);
  }
  part[wv][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (threadIdx.x < NB && row < m) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < RECT_WAVES; q++) a += part[q][lane];
    if (qa == 0 && qb == k) V[sd.Voff[s] + row - k] = Wg[sd.Woff[s] + row] - a;
    else Pp[tid * NB + lane] = a;
  }
}

// Four consecutive entries of a column, loaded as one 16-byte (FP32) or two 16-byte (FP64) accesses at the element
// alignment of the factor (columns start anywhere: the front offsets and leading dimensions are arbitrary).
template <typename T>
struct Quad { T e[4]; };
template <typename T>
__device__ __forceinline__ Quad<T> load_quad(const T* p) {
  Quad<T> q;
  __builtin_memcpy(&q, p, sizeof(q));
  return q;
}

// acc[e][u % 4] += sum over columns q = base + u in [q0, q1) of pr[q * ld + e] * yo[q]: a lane's four consecutive
// rows of a column-major panel against a vector every lane of the wave shares, each row in four interleaved partial
// sums (short rounding chains). Full batches issue BQ quad loads at 32-bit offsets from a wave-uniform base
// (BQ * ld < 2^31: ld < 2^27) and take yo by wave-uniform loads. q0, q1 are wave-uniform.
template <typename T, int BQ>
__device__ __forceinline__ void sweep_rows4_dot(const T* __restrict__ pr, int ld, const double* __restrict__ yo,
                                                int q0, int q1, double acc[4][4]) {
  int base = q0;
  for (; base + BQ <= q1; base += BQ) {
    const T* pc = pr + (long long)base * ld;
    Quad<T> v[BQ];
#pragma unroll
    for (int u = 0; u < BQ; u++) v[u] = load_quad(pc + (unsigned)(u * ld));
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < BQ; u++) {
      const double yu = yo[base + u];
#pragma unroll
      for (int e = 0; e < 4; e++) acc[e][u & 3] = fma((double)v[u].e[e], yu, acc[e][u & 3]);
    }
  }
  for (int u = 0; base < q1; base++, u++) {
    const Quad<T> v = load_quad(pr + (long long)base * ld);
    const double yu = yo[base];
#pragma unroll
    for (int e = 0; e < 4; e++) acc[e][u & 3] = fma((double)v.e[e], yu, acc[e][u & 3]);
  }
}

// The RECT_ROWS rows [r0, min(r0 + RECT_ROWS, rend)) of a task, four per lane: lane l loads rows rc .. rc + 3 with
// rc = min(r0 + 4 l, rend - 4) (rend >= 4, so the loads stay in the column) and owns those of them at or past
// r0 + 4 l; wave wv's partial sums go to part[wv][row - r0].
template <typename T>
__device__ __forceinline__ void sweep_rows4_task(const T* __restrict__ P, long long ld, long long r0, long long rend,
                                                 const double* __restrict__ yo, int qa, int qb,
                                                 double (*part)[RECT_ROWS]) {
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long r = r0 + 4 * lane;
  const long long rc = min(r, rend - 4);
  constexpr int BQ = sizeof(T) == 4 ? 16 : 8;
  const int per = (((qb - qa + RECT_WAVES - 1) / RECT_WAVES + BQ - 1) / BQ) * BQ;  // whole batches but the last
  const int q0 = qa + wv * per, q1 = min(qb, q0 + per);
  double acc[4][4] = {};
  if (q0 < q1) sweep_rows4_dot<T, BQ>(P + rc, (int)ld, yo, q0, q1, acc);
#pragma unroll
  for (int e = 0; e < 4; e++)
    if (rc + e >= r && rc + e < rend) part[wv][rc + e - r0] = (acc[e][0] + acc[e][1]) + (acc[e][2] + acc[e][3]);
}

// The update-vector rows of the FP32 sweeps: fwdb_rect_item on RECT_ROWS rows per task, four consecutive rows per
// lane (one 16-byte load per lane and column; 64 single-row lanes issue 4-byte loads and leave HBM half idle). Same
// column ranges and partial-sum slots (RECT_ROWS per task). task: (s, first row, first column, end column).
template <typename T>
__device__ __forceinline__ void fwdb_rect4_item(const int4 tk, long long tid, SymDev sd, const T* __restrict__ L,
                                                   const double* __restrict__ y, const double* __restrict__ Wg,
                                                   double* __restrict__ V, double* __restrict__ Pp) {
  __shared__ double part[RECT_WAVES][RECT_ROWS];
  const int s = tk.x;
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int qa = tk.z, qb = tk.w;
  // wide fronts have m > 181 rows, so the clamped quads stay inside the column
  sweep_rows4_task<T>(L + sd.Loff[s], m, tk.y, m, y + f, qa, qb, part);
  __syncthreads();
  const long long row = (long long)tk.y + threadIdx.x;
  if (threadIdx.x < RECT_ROWS && row < m) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < RECT_WAVES; q++) a += part[q][threadIdx.x];
    if (qa == 0 && qb == k) V[sd.Voff[s] + row - k] = Wg[sd.Woff[s] + row] - a;
    else Pp[tid * RECT_ROWS + threadIdx.x] = a;
  }
}

// V(row) = W(row) - sum over the row block's column ranges (in range order) of their partial sums; ROWS rows per
// task (NB for the FP64 sweeps, RECT_ROWS for the FP32 ones). task: (s, first row, index of the first range's task,
// number of ranges)
template <int ROWS>
__device__ __forceinline__ void fwdb_rect_sum_item(const int4 tk, SymDev sd, const double* __restrict__ Wg,
                                                       double* __restrict__ V, const double* __restrict__ Pp) {
  const int s = tk.x;
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int lane = threadIdx.x;
  const long long row = (long long)tk.y + lane;
  if (row >= m) return;
  double a = 0.0;
  for (int g = 0; g < tk.w; g++) a += Pp[((long long)tk.z + g) * ROWS + lane];
  V[sd.Voff[s] + row - k] = Wg[sd.Woff[s] + row] - a;
}

// Backward, first part: the solution at a wide front's lower rows, gathered into one contiguous vector (the front's
// forward update-vector slot, free once the forward sweep is done). task: (s, first row, end row, 0), rows >= k.
__device__ __forceinline__ void bwdb_xb_item(const int4 tk, SymDev sd,
                                                 const double* __restrict__ y, double* __restrict__ V) {
  const int s = tk.x;
  const long long i = (long long)tk.y + threadIdx.x;
  if (i >= tk.z) return;
  const int k = sn_k(sd, s);
  V[sd.Voff[s] + i - k] = y[sd.sn_rows[sd.sn_rowptr[s] + i]];
}

// z_q = y_own[q] - L21(:, q)^T xb for the own columns q of this task: each wave takes GEMV_COLS / 4 consecutive
// columns and shares its loads of xb among them; a lane takes four consecutive rows of every column (one quad load
// per column), GEMV_RB such quads of every column in flight. Columns past the task are clamped to its last column and
// their sums dropped. task: (s, first column, end column, 0)
template <typename T>
__device__ __forceinline__ void bwdb_gemv_item(const int4 tk, SymDev sd, const T* __restrict__ L,
                                                   const double* __restrict__ y, const double* __restrict__ V,
                                                   double* __restrict__ Wg) {
  constexpr int NQ = GEMV_COLS / 4;
  const int s = tk.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q0 = tk.y + NQ * wv;
  if (q0 >= tk.z) return;
  const int nq = min(NQ, tk.z - q0);
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const long long mb = m - k;
  const T* cq[NQ];
#pragma unroll
  for (int j = 0; j < NQ; j++) cq[j] = L + sd.Loff[s] + (long long)(q0 + min(j, nq - 1)) * m + k;
  const double* xb = V + sd.Voff[s];
  double acc[NQ][2];
#pragma unroll
  for (int j = 0; j < NQ; j++) acc[j][0] = acc[j][1] = 0.0;
  long long base = 4 * lane;
  for (; base + 256 * (GEMV_RB - 1) + 3 < mb; base += 256 * GEMV_RB) {  // full batches: every load issued before use
    Quad<double> xv[GEMV_RB];
    Quad<T> v[NQ][GEMV_RB];
#pragma unroll
    for (int u = 0; u < GEMV_RB; u++) xv[u] = load_quad(xb + base + 256 * u);
#pragma unroll
    for (int j = 0; j < NQ; j++)
#pragma unroll
      for (int u = 0; u < GEMV_RB; u++) v[j][u] = load_quad(cq[j] + base + 256 * u);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < NQ; j++)
#pragma unroll
      for (int u = 0; u < GEMV_RB; u++)
#pragma unroll
        for (int e = 0; e < 4; e++) acc[j][e & 1] = fma((double)v[j][u].e[e], xv[u].e[e], acc[j][e & 1]);
  }
  for (; base + 3 < mb; base += 256) {
    const Quad<double> x = load_quad(xb + base);
#pragma unroll
    for (int j = 0; j < NQ; j++) {
      const Quad<T> v = load_quad(cq[j] + base);
#pragma unroll
      for (int e = 0; e < 4; e++) acc[j][e & 1] = fma((double)v.e[e], x.e[e], acc[j][e & 1]);
    }
  }
  if (base < mb) {  // the last lane with rows: fewer than four left
    for (long long i = base; i < mb; i++) {
      const double x = xb[i];
#pragma unroll
      for (int j = 0; j < NQ; j++) acc[j][0] = fma((double)cq[j][i], x, acc[j][0]);
    }
  }
#pragma unroll
  for (int j = 0; j < NQ; j++) {
    const double a = wave_sum(acc[j][0] + acc[j][1]);
    if (lane == 0 && j < nq) Wg[sd.Woff[s] + q0 + j] = y[f + q0 + j] - a;
  }
}

// Backward diagonal step: x_c = sum_{r >= c} X_t[r][c] z_r (X_t^T z_t). One wave per pair of columns
// (p, wd - 1 - p), every wave reading wd + 1 entries of the column-major X_t (X[c * wd + r] = X_t[r][c]) with the
// lanes over r, DB loads of X and of z in flight per lane, summed in a fixed order. task: (s, c0, first pair, 0),
// DIAG_WAVES pairs per workgroup.
template <typename TX>
__device__ __forceinline__ void bwdb_diag_item(const int4 tk, SymDev sd,
                                                   const TX* __restrict__ X, double* __restrict__ y,
                                                   const double* __restrict__ Wg) {
  const int s = tk.x, c0 = tk.y;
  const int lane = threadIdx.x & 63;
  const int p = tk.z + (threadIdx.x >> 6);
  const int k = sn_k(sd, s);
  const int wd = min(SB, k - c0);
  if (p >= (wd + 1) / 2) return;
  const int f = sd.sn_first[s];
  const TX* Xb = X + x_step_off(sd, s, c0);
  const double* z = Wg + sd.Woff[s] + c0;
  for (int h = 0; h < 2; h++) {
    const int c = h == 0 ? p : wd - 1 - p;
    if (h == 1 && c == p) break;
    const TX* xc = Xb + (long long)c * wd;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int base = c; base < wd; base += 64 * DB) {
      TX xv[DB];
      double zv[DB];
#pragma unroll
      for (int u = 0; u < DB; u++) {
        const int r = min(base + lane + 64 * u, wd - 1);  // entries past the column repeat its last one, weighted by zero
        xv[u] = xc[r];
        zv[u] = z[r];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < DB; u++) acc[u & 3] += base + lane + 64 * u < wd ? (double)xv[u] * zv[u] : 0.0;
    }
    const double t = wave_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if (lane == 0) y[f + c0 + c] = t;
  }
}

// Backward panel step: z_j -= L(c0 .. c0+wd, j)^T x_t for the columns j of this task (< c0): wave wv takes columns
// j0 + wv + 4u, lanes over the step's rows, BP_ROWS / 64 rows of each of its 4 columns loaded per batch before the
// products. task: (s, c0, first column, end column), at most 16 columns.
template <typename T>
__device__ __forceinline__ void bwdb_panel_item(const int4 tk, SymDev sd,
                                                    const T* __restrict__ L, const double* __restrict__ y,
                                                    double* __restrict__ Wg) {
  constexpr int NCB = BP_ROWS / 64;
  __shared__ double xs[SB + BP_ROWS];
  const int s = tk.x, c0 = tk.y;
  const int f = sd.sn_first[s];
  const int k = sn_k(sd, s);
  const long long m = sn_m(sd, s);
  const int wd = min(SB, k - c0);
  for (int q = threadIdx.x; q < SB + BP_ROWS; q += 256) xs[q] = q < wd ? y[f + c0 + q] : 0.0;
  __syncthreads();
  const T* P = L + sd.Loff[s];
  double* z = Wg + sd.Woff[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // columns past the task and rows past the step are clamped in bounds (weighted by zero or not stored)
  const T* cj[4];
#pragma unroll
  for (int u = 0; u < 4; u++) cj[u] = P + (long long)min(tk.z + wv + 4 * u, tk.w - 1) * m + c0;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int t0 = 0; t0 < wd; t0 += BP_ROWS) {
    T v[4][NCB];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int t = 0; t < NCB; t++) v[u][t] = cj[u][min(t0 + lane + 64 * t, wd - 1)];
    __builtin_amdgcn_sched_barrier(0);  // the batch's loads in flight before the first product
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int t = 0; t < NCB; t++) {
        const int r = t0 + lane + 64 * t;
        a[u] += r < wd ? (double)v[u][t] * xs[r] : 0.0;
      }
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int j = tk.z + wv + 4 * u;
    const double t = wave_sum(a[u]);
    if (lane == 0 && j < tk.w) z[j] -= t;
  }
}

// one launch per step and kind; the stream orders them
__global__ __launch_bounds__(256) void k_fwdb_gather(const int4* __restrict__ tasks, SymDev sd,
                                                     const double* __restrict__ y, const double* __restrict__ V,
                                                     double* __restrict__ Wg) {
  fwdb_gather_item(tasks[blockIdx.x], sd, y, V, Wg);
}

template <typename TX>
__global__ __launch_bounds__(DIAG_WAVES * 64) void k_fwdb_diag(const int4* __restrict__ tasks, SymDev sd,
                                                   const TX* __restrict__ XT, double* __restrict__ y,
                                                   const double* __restrict__ Wg) {
  fwdb_diag_item<TX>(tasks[blockIdx.x], sd, XT, y, Wg);
}

template <typename T>
__global__ __launch_bounds__(RECT_WAVES * 64) void k_fwdb_panel(const int4* __restrict__ tasks, SymDev sd,
                                                    const T* __restrict__ L, const double* __restrict__ y,
                                                    double* __restrict__ Wg) {
  fwdb_panel_item<T>(tasks[blockIdx.x], sd, L, y, Wg);
}

template <typename T>
__global__ __launch_bounds__(RECT_WAVES * 64) void k_fwdb_rect(const int4* __restrict__ tasks, long long tbase,
                                                               SymDev sd, const T* __restrict__ L,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ Wg, double* __restrict__ V,
                                                               double* __restrict__ Pp) {
  if constexpr (sizeof(T) == 4) fwdb_rect4_item<T>(tasks[blockIdx.x], tbase + blockIdx.x, sd, L, y, Wg, V, Pp);
  else fwdb_rect_item<T>(tasks[blockIdx.x], tbase + blockIdx.x, sd, L, y, Wg, V, Pp);
}

template <int ROWS>
__global__ __launch_bounds__(ROWS) void k_fwdb_rect_sum(const int4* __restrict__ tasks, SymDev sd,
                                                        const double* __restrict__ Wg, double* __restrict__ V,
                                                        const double* __restrict__ Pp) {
  fwdb_rect_sum_item<ROWS>(tasks[blockIdx.x], sd, Wg, V, Pp);
}

__global__ __launch_bounds__(256) void k_bwdb_xb(const int4* __restrict__ tasks, SymDev sd,
                                                 const double* __restrict__ y, double* __restrict__ V) {
  bwdb_xb_item(tasks[blockIdx.x], sd, y, V);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwdb_gemv(const int4* __restrict__ tasks, SymDev sd, const T* __restrict__ L,
                                                   const double* __restrict__ y, const double* __restrict__ V,
                                                   double* __restrict__ Wg) {
  bwdb_gemv_item<T>(tasks[blockIdx.x], sd, L, y, V, Wg);
}

template <typename TX>
__global__ __launch_bounds__(DIAG_WAVES * 64) void k_bwdb_diag(const int4* __restrict__ tasks, SymDev sd,
                                                   const TX* __restrict__ X, double* __restrict__ y,
                                                   const double* __restrict__ Wg) {
  bwdb_diag_item<TX>(tasks[blockIdx.x], sd, X, y, Wg);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwdb_panel(const int4* __restrict__ tasks, SymDev sd,
                                                    const T* __restrict__ L, const double* __restrict__ y,
                                                    double* __restrict__ Wg) {
  bwdb_panel_item<T>(tasks[blockIdx.x], sd, L, y, Wg);
}

// grid-stride: the factor can hold more than 2^32 entries (one dispatch covers fewer work-items)
__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) b[i] = (float)a[i];
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
  if (side_) {
    (void)hipStreamSynchronize(side_);
    (void)hipStreamDestroy(side_);
  }
  for (hipEvent_t e : sweep_ev_) (void)hipEventDestroy(e);
  if (chain_) {
    (void)hipStreamSynchronize(chain_);
    (void)hipStreamSynchronize(outer_);
    (void)hipEventDestroy(ev_main_);
    (void)hipEventDestroy(ev_chain_);
    (void)hipStreamDestroy(chain_);
    (void)hipStreamDestroy(outer_);
  }
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
  return L_.bytes() + U_.bytes() + V_.bytes() + W_.bytes() + Pr_.bytes() + amap_dst_.bytes() + amap_src_.bytes() +
         sn_rows_.bytes() + ext_tasks_.bytes() + panel_tasks_.bytes() + syrk_tasks_.bytes() +
         Linv_.bytes() + X_.bytes() + XT_.bytes() + X32_.bytes() + XT32_.bytes() + gptr_.bytes() + gsrc_.bytes() +
         fwd_tasks_.bytes() + bwd_tasks_.bytes();
}

void SpdDevice::analyze(int n, const int* colPtr, const int* rowIdx, const int* perm, const SymbolicOptions& opt) {
  analyzed_ = false;
  factored_ = false;
  drop_graph();
  int rc = symbolic_analyze(n, colPtr, rowIdx, perm, opt, S_);
  if (rc) throw Error(rc == -2 ? -2 : (rc == -1 ? -1 : -3), "symbolic analysis failed: " + S_.error);
  const Symbolic& S = S_;
  const int ns = S.ns;
  // the sweeps address a front's panel with 32-bit element offsets, and the outer update reads a 512-column block of
  // a front through a buffer descriptor with 32-bit byte offsets (512 x 2^19 x 8 B = 2^31; such a front's update
  // block alone would need 2 TB)
  for (int s = 0; s < ns; s++)
    if (S.sn_rowptr[s + 1] - S.sn_rowptr[s] >= (1LL << 19))
      throw Error(-3, "a supernode front has 2^19 rows or more");
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
  std::vector<int4> ext, pan, syr, pot;
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
    // extend-add rounds: round r = r-th child of every parent at this level. A child's update columns that land in
    // the parent's own columns are added before the parent is factored; those that land in the parent's update block
    // are added after it: the parent's first outer update stores (not adds) into every entry of its update block, so
    // the block is never cleared
    int maxc = 0;
    for (int s : bylevel[l]) maxc = std::max(maxc, cptr[s + 1] - cptr[s]);
    for (int part = 0; part < 2; part++)
      for (int r = 0; r < maxc; r++) {
        LaunchList ll;
        ll.begin = (long long)ext.size();
        for (int s : bylevel[l]) {
          if (cptr[s + 1] - cptr[s] <= r) continue;
          int c = clist[cptr[s] + r];
          int kc = S.sn_first[c + 1] - S.sn_first[c];
          int mu = (int)(S.sn_rowptr[c + 1] - S.sn_rowptr[c] - kc);
          // the child's update rows are sorted; those below the parent's last column are its own columns
          const int* ur = S.sn_rows.data() + S.sn_rowptr[c] + kc;
          int js = (int)(std::lower_bound(ur, ur + mu, S.sn_first[s + 1]) - ur);
          int j = part == 0 ? 0 : js, jend = part == 0 ? js : mu;
          while (j < jend) {
            int j1 = j;
            long long work = 0;
            while (j1 < jend && work < EXT_TARGET) { work += mu - j1; j1++; }
            ext.push_back(make_int4(c, j, j1, s));
            j = j1;
          }
        }
        ll.end = (long long)ext.size();
        if (ll.end > ll.begin) (part == 0 ? L.ext_rounds : L.ext_rounds_u).push_back(ll);
      }
    // factor steps: 64-column panels; inner updates stay inside the OB-column inner block; after the first inner
    // block of an OB2-column outer block, a mid update brings the second one up to date; the outer update of
    // everything to its right follows the outer block's last panel
    int maxk = 0;
    for (int s : bylevel[l]) maxk = std::max(maxk, S.sn_first[s + 1] - S.sn_first[s]);
    std::vector<int4> outa, outb;
    for (int c0 = 0; c0 < maxk; c0 += NB) {
      LaunchList lp, ls, lm, lo, lpo;
      lp.begin = (long long)pan.size();
      lpo.begin = (long long)pot.size();
      lpo.c0 = c0;
      ls.begin = (long long)syr.size();
      lp.c0 = ls.c0 = c0;
      double fl = 0, flm = 0, flo = 0, flo_a = 0;
      int nslot = 0;
      const int ib0 = (c0 / OB) * OB, ob0 = (c0 / OB2) * OB2;
      const bool outer_now = (c0 + NB) % OB2 == 0 || c0 + NB >= maxk;
      const bool mid_now = (c0 + NB) % OB2 == OB;
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
        // inner update: columns [rb, min(k, ib0 + OB))
        long long cend = std::min<long long>(k, ib0 + OB);
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
      lm.begin = (long long)syr.size();
      lm.c0 = ob0;
      if (mid_now) {
        for (int s : bylevel[l]) {
          // a front with columns beyond the first inner block has just factored all of its panels in it
          int k = S.sn_first[s + 1] - S.sn_first[s];
          if (k <= c0 + NB) continue;
          long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
          long long rb = c0 + NB, cend = std::min<long long>(k, ob0 + OB2);
          long long ti = (m - rb + 127) / 128, tj = (cend - rb + 127) / 128;
          for (long long I = 0; I < ti; I++)
            for (long long J = 0; J <= std::min(I, tj - 1); J++) syr.push_back(make_int4(s, (int)I, (int)J, 0));
          for (long long j = rb; j < cend; j++) flm += 2.0 * OB * (double)(m - j);
        }
      }
      lm.end = (long long)syr.size();
      L.mid_steps.push_back(lm);
      L.mid_flops.push_back(flm);
      lo.begin = (long long)syr.size();
      lo.c0 = ob0;
      if (outer_now) {
        for (int s : bylevel[l]) {
          int k = S.sn_first[s + 1] - S.sn_first[s];
          if (k <= ob0) continue;
          // the outer update of this front happens once, after its last panel inside [ob0, ob0 + OB2)
          int last_c0 = ((std::min(k, ob0 + OB2) - 1) / NB) * NB;
          if (last_c0 > c0) continue;
          long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
          int kw = std::min(OB2, k - ob0);
          long long rb = ob0 + kw;
          long long nrows = m - rb;
          if (nrows <= 0) continue;
          // tiles in OUTER_GROUP x OUTER_GROUP groups: an XCD runs a contiguous range of the list, so the tiles
          // resident on it at once read few distinct row panels, which then stay in its L2
          long long tiles = (nrows + 127) / 128;
          for (long long IB = 0; IB < tiles; IB += OUTER_GROUP)
            for (long long JB = 0; JB <= IB; JB += OUTER_GROUP)
              for (long long I = IB; I < std::min(tiles, IB + OUTER_GROUP); I++)
                for (long long J = JB; J <= std::min(I, JB + OUTER_GROUP - 1); J++)
                  (J < OUTER_NEXT_TILES ? outa : outb).push_back(make_int4(s, (int)I, (int)J, 0));
          // part A: the tile columns of the next outer block, which its panels need; part B: the rest
          const long long na = std::min(nrows, (long long)OB2);
          flo_a += (double)kw * (double)na * (double)(2 * nrows - na + 1);
          flo += (double)kw * (double)nrows * (double)(nrows + 1);
        }
      }
      syr.insert(syr.end(), outa.begin(), outa.end());
      lo.end = (long long)syr.size();
      LaunchList lob = lo;
      lob.begin = lo.end;
      syr.insert(syr.end(), outb.begin(), outb.end());
      lob.end = (long long)syr.size();
      outa.clear();
      outb.clear();
      L.outer_steps.push_back(lo);
      L.outer_flops.push_back(flo_a);
      L.outer_b_steps.push_back(lob);
      L.outer_b_flops.push_back(flo - flo_a);
    }
  }
  // sweep schedule: narrow fronts one workgroup each; wide fronts in SB-column steps (diagonal + panel launches)
  std::vector<int> swp;
  std::vector<int4> fwt, bwt, invt;
  std::vector<long long> Xoff(ns + 1, 0), Goff(ns + 1, 0), gptr, gsrc;
  auto sn_kk = [&](int s) { return S.sn_first[s + 1] - S.sn_first[s]; };
  auto sn_mm = [&](int s) { return S.sn_rowptr[s + 1] - S.sn_rowptr[s]; };
  std::vector<char> isbig(ns, 0);
  for (int s = 0; s < ns; s++) {
    long long k = sn_kk(s), m = sn_mm(s);
    isbig[s] = k > BIG_K || k * m > BIG_PANEL;
    Xoff[s + 1] = Xoff[s] + (isbig[s] ? (long long)(k / SB) * SB * SB + (long long)(k % SB) * (k % SB) : 0);
    if (!isbig[s]) continue;
    // row-wise inverse of the children's relative maps, contributions of a row in child order
    Goff[s] = (long long)gptr.size();
    std::vector<long long> cnt(m + 1, 0);
    for (int cp = cptr[s]; cp < cptr[s + 1]; cp++) {
      int c = clist[cp];
      long long mu = sn_mm(c) - sn_kk(c);
      const int* rel = S.rel.data() + S.rel_off[c];
      for (long long i = 0; i < mu; i++) cnt[rel[i] + 1]++;
    }
    const long long base = (long long)gsrc.size();
    for (long long i = 0; i < m; i++) cnt[i + 1] += cnt[i];
    for (long long i = 0; i <= m; i++) gptr.push_back(base + cnt[i]);
    gsrc.resize(base + cnt[m]);
    for (int cp = cptr[s]; cp < cptr[s + 1]; cp++) {
      int c = clist[cp];
      long long mu = sn_mm(c) - sn_kk(c);
      const int* rel = S.rel.data() + S.rel_off[c];
      for (long long i = 0; i < mu; i++) gsrc[base + cnt[rel[i]]++] = S.Voff[c] + i;
    }
  }
  long long rect_part_max = 0;  // partial-sum slots of the largest forward rect launch (tasks times rows per task)
  for (int l = 0; l < nl; l++) {
    Level& L = levels_[l];
    std::vector<int> small, big;
    for (int s : bylevel[l]) (isbig[s] ? big : small).push_back(s);
    L.small_begin = (long long)swp.size();
    for (int s : small) {
      swp.push_back(s);
      L.small_max_m = std::max<long long>(L.small_max_m, sn_mm(s));
    }
    L.small_end = (long long)swp.size();
    int maxk = 0;
    for (int s : big) maxk = std::max(maxk, sn_kk(s));
    // forward
    L.fwd_gather.begin = (long long)fwt.size();
    for (int s : big)
      for (long long r0 = 0; r0 < sn_mm(s); r0 += 256)
        fwt.push_back(make_int4(s, (int)r0, (int)std::min<long long>(sn_mm(s), r0 + 256), 0));
    L.fwd_gather.end = (long long)fwt.size();
    for (int c0 = 0; c0 < maxk; c0 += SB) {
      LaunchList d, pn;
      d.begin = (long long)fwt.size();
      for (int s : big) {
        int k = sn_kk(s);
        if (k <= c0) continue;
        int wd = std::min(SB, k - c0);
        for (int pb = 0; pb < (wd + 1) / 2; pb += DIAG_WAVES) fwt.push_back(make_int4(s, c0, pb, 0));
        for (int j = 0; j * NB < wd; j++) invt.push_back(make_int4(s, c0, j, 0));
      }
      d.end = pn.begin = (long long)fwt.size();
      for (int s : big) {
        int k = sn_kk(s);
        if (k <= c0) continue;
        int wd = std::min(SB, k - c0);
        for (int r0 = c0 + wd; r0 < k; r0 += NB) fwt.push_back(make_int4(s, c0, r0, 0));
      }
      pn.end = (long long)fwt.size();
      L.fwd_diag.push_back(d);
      L.fwd_panel.push_back(pn);
    }
    // update-vector rows in row blocks (NB rows for the FP64 sweeps, RECT_ROWS for the FP32 ones); fronts wider than
    // RECT_SPLIT columns in equal column ranges (partial sums added by a second launch, k_fwdb_rect_sum, in range
    // order)
    for (int pass = 0; pass < 2; pass++) {
      const int rows = pass == 0 ? NB : RECT_ROWS;
      LaunchList& lr = pass == 0 ? L.fwd_rect : L.fwd_rect4;
      LaunchList& ls = pass == 0 ? L.fwd_rect_sum : L.fwd_rect4_sum;
      lr.begin = (long long)fwt.size();
      std::vector<int4> sums;
      for (int s : big) {
        const int k = sn_kk(s);
        const int g = (k + RECT_SPLIT - 1) / RECT_SPLIT;
        const int cw = (k + g - 1) / g;
        for (long long r0 = k; r0 < sn_mm(s); r0 += rows) {
          const long long t0 = (long long)fwt.size() - lr.begin;
          int nr = 0;
          for (int q = 0; q < k; q += cw, nr++) fwt.push_back(make_int4(s, (int)r0, q, std::min(k, q + cw)));
          if (nr > 1) sums.push_back(make_int4(s, (int)r0, (int)t0, nr));
        }
      }
      lr.end = (long long)fwt.size();
      rect_part_max = std::max(rect_part_max, (lr.end - lr.begin) * rows);
      ls.begin = (long long)fwt.size();
      fwt.insert(fwt.end(), sums.begin(), sums.end());
      ls.end = (long long)fwt.size();
    }
    // backward
    L.bwd_xb.begin = (long long)bwt.size();
    for (int s : big)
      for (long long r0 = sn_kk(s); r0 < sn_mm(s); r0 += 256)
        bwt.push_back(make_int4(s, (int)r0, (int)std::min<long long>(sn_mm(s), r0 + 256), 0));
    L.bwd_xb.end = L.bwd_gemv.begin = (long long)bwt.size();
    for (int s : big)
      for (int q = 0; q < sn_kk(s); q += GEMV_COLS) bwt.push_back(make_int4(s, q, std::min(sn_kk(s), q + GEMV_COLS), 0));
    L.bwd_gemv.end = (long long)bwt.size();
    for (int c0 = ((maxk - 1) / SB) * SB; maxk > 0 && c0 >= 0; c0 -= SB) {
      LaunchList d, pn;
      d.begin = (long long)bwt.size();
      for (int s : big) {
        int k = sn_kk(s);
        if (k <= c0) continue;
        int wd = std::min(SB, k - c0);
        for (int pb = 0; pb < (wd + 1) / 2; pb += DIAG_WAVES) bwt.push_back(make_int4(s, c0, pb, 0));
      }
      d.end = pn.begin = (long long)bwt.size();
      for (int s : big) {
        if (sn_kk(s) <= c0) continue;
        for (int j = 0; j < c0; j += SWEEP_COLS) bwt.push_back(make_int4(s, c0, j, std::min(c0, j + SWEEP_COLS)));
      }
      pn.end = (long long)bwt.size();
      L.bwd_diag.push_back(d);
      L.bwd_panel.push_back(pn);
    }
  }
  inv_launch_ = LaunchList{0, (long long)invt.size(), 0};
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
  Xoff_.upload(Xoff, stream_);
  X_.alloc(std::max<long long>(1, Xoff[ns]));
  XT_.alloc(std::max<long long>(1, Xoff[ns]));
  Goff_.upload(Goff, stream_);
  if (gptr.empty()) gptr_.alloc(1); else gptr_.upload(gptr, stream_);
  if (gsrc.empty()) gsrc_.alloc(1); else gsrc_.upload(gsrc, stream_);
  if (invt.empty()) inv_tasks_.alloc(1); else inv_tasks_.upload(invt, stream_);

  rel_off_.upload(S.rel_off, stream_);
  if (S.rel.empty()) rel_.alloc(1); else rel_.upload(S.rel, stream_);
  child_ptr_.upload(cptr, stream_);
  if (clist.empty()) child_list_.alloc(1); else child_list_.upload(clist, stream_);
  amap_src_.upload(S.amap_src, stream_);
  amap_dst_.upload(S.amap_dst, stream_);
  if (ext.empty()) ext_tasks_.alloc(1); else ext_tasks_.upload(ext, stream_);
  if (pan.empty()) panel_tasks_.alloc(1); else panel_tasks_.upload(pan, stream_);
  if (syr.empty()) syrk_tasks_.alloc(1); else syrk_tasks_.upload(syr, stream_);
  if (pot.empty()) potrf_tasks_.alloc(1); else potrf_tasks_.upload(pot, stream_);
  level_sn_.upload(lvl_sn, stream_);
  V_.alloc(std::max<long long>(1, S.Vsize));
  W_.alloc(std::max<long long>(1, Woff[ns]));
  Pr_.alloc(std::max<long long>(1, rect_part_max));
  y_.alloc(n);
  err_.alloc(1);
  sd_ = SymDev{sn_first_.p, sn_rowptr_.p, sn_rows_.p, Loff_.p, Uoff_.p, Voff_.p, Woff_.p, rel_off_.p, rel_.p,
               child_ptr_.p, child_list_.p, Ioff_.p, Xoff_.p, Goff_.p, gptr_.p, gsrc_.p};
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
  // Look-ahead over two streams: the main stream runs the extend-adds and the outer updates, the chain stream the
  // panel chain of each outer block (POTRF, TRSM, inner and mid updates). An outer update runs in two launches: part A,
  // the tile columns of the next outer block, then part B, the rest of the trailing matrix; the next block's chain
  // starts after part A and runs while part B does. Writes never overlap: part B touches only columns beyond the next
  // block, and every update of an entry (the atomic epilogues included) stays on one stream in program order.
  hipStream_t Cs = chain_stream(), M = outer_;
  IFEA_HIP(hipEventRecord(ev_main_, stream_));
  IFEA_HIP(hipStreamWaitEvent(M, ev_main_, 0));
  auto launch_update = [&](hipStream_t st, int mode, const LaunchList& q, double flops) {
    if (q.end <= q.begin) return;
    if (profiling_) IFEA_HIP(hipEventRecord(event(2 * nev), st));
    if (mode == 0)
      hipLaunchKernelGGL(k_syrk<T>, dim3((unsigned)(q.end - q.begin)), dim3(256), 0, st, syrk_tasks_.p + q.begin,
                         q.c0, sd_, L);
    else if (mode == 1)
      hipLaunchKernelGGL((k_syrk_big<T, 512, 16, true>), dim3((unsigned)(q.end - q.begin)), dim3(512), 0, st,
                         syrk_tasks_.p + q.begin, q.c0, sd_, L, U);
    else
      hipLaunchKernelGGL((k_syrk_big<T, 512, 16, false>), dim3((unsigned)(q.end - q.begin)), dim3(512), 0, st,
                         syrk_tasks_.p + q.begin, q.c0, sd_, L, U);
    if (profiling_) {
      IFEA_HIP(hipEventRecord(event(2 * nev + 1), st));
      nev++;
      const bool outer = mode == 2;
      ev_mode_.push_back((char)outer);
      if (!outer) prof_.syrk_flops += flops;
      if (outer) prof_.outer_flops += flops;
      if (outer && sizeof(T) == 4) prof_.outer_flops_f32 += flops;
    }
  };
  for (size_t l = 0; l < levels_.size(); l++) {
    const Level& lv = levels_[l];
    for (const LaunchList& r : lv.ext_rounds)
      hipLaunchKernelGGL(k_extend<T>, dim3((unsigned)(r.end - r.begin)), dim3(256), 0, M, ext_tasks_.p + r.begin,
                         sd_, L, U);
    IFEA_HIP(hipEventRecord(ev_main_, M));
    IFEA_HIP(hipStreamWaitEvent(Cs, ev_main_, 0));
    const size_t nsteps = lv.panel_steps.size();
    for (size_t t = 0; t < nsteps; t++) {
      const int c0 = lv.potrf_steps[t].c0;
      // a new outer block's chain needs part A of the previous outer update
      if (t > 0 && c0 % OB2 == 0) IFEA_HIP(hipStreamWaitEvent(Cs, ev_main_, 0));
      const LaunchList& po = lv.potrf_steps[t];
      if (po.end > po.begin)
        hipLaunchKernelGGL(k_potrf_inv<T>, dim3((unsigned)(po.end - po.begin)), dim3(256), 0, Cs,
                           potrf_tasks_.p + po.begin, po.c0, sd_, L, Linv_.p, err_.p);
      const LaunchList& p = lv.panel_steps[t];
      if (p.end > p.begin)
        hipLaunchKernelGGL(k_trsm_mfma<T>, dim3((unsigned)(p.end - p.begin)), dim3(256), 0, Cs,
                           panel_tasks_.p + p.begin, p.c0, sd_, L, Linv_.p);
      // the profile counts inner and mid updates together
      launch_update(Cs, 0, lv.syrk_steps[t], lv.syrk_flops[t]);
      launch_update(Cs, 1, lv.mid_steps[t], lv.mid_flops[t]);
      const bool outer_now = (c0 + NB) % OB2 == 0 || t + 1 == nsteps;
      if (outer_now) {
        IFEA_HIP(hipEventRecord(ev_chain_, Cs));
        IFEA_HIP(hipStreamWaitEvent(M, ev_chain_, 0));
        launch_update(M, 2, lv.outer_steps[t], lv.outer_flops[t]);
        IFEA_HIP(hipEventRecord(ev_main_, M));
        launch_update(M, 2, lv.outer_b_steps[t], lv.outer_b_flops[t]);
      }
    }
    IFEA_HIP(hipEventRecord(ev_chain_, Cs));
    IFEA_HIP(hipStreamWaitEvent(M, ev_chain_, 0));
    for (const LaunchList& r : lv.ext_rounds_u)
      hipLaunchKernelGGL(k_extend<T>, dim3((unsigned)(r.end - r.begin)), dim3(256), 0, M, ext_tasks_.p + r.begin,
                         sd_, L, U);
  }
  IFEA_HIP(hipEventRecord(ev_main_, M));
  IFEA_HIP(hipStreamWaitEvent(stream_, ev_main_, 0));
  if (inv_launch_.end > inv_launch_.begin)
    hipLaunchKernelGGL(k_inv_diag<T>, dim3((unsigned)(inv_launch_.end - inv_launch_.begin)), dim3(256), 0, stream_,
                       inv_tasks_.p, sd_, L, Linv_.p, X_.p, XT_.p);
  IFEA_HIP(hipGetLastError());
  int herr = 0;
  IFEA_HIP(hipMemcpyAsync(&herr, err_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
  IFEA_HIP(hipStreamSynchronize(stream_));
  if (profiling_) {
    for (size_t e = 0; e < nev; e++) {
      float ms = 0;
      IFEA_HIP(hipEventElapsedTime(&ms, event(2 * e), event(2 * e + 1)));
      if (ev_mode_[e] == 0) {
        prof_.syrk_ms += ms;
        prof_.syrk_launches++;
      } else {
        prof_.outer_ms += ms;
        prof_.outer_launches++;
        if (sizeof(T) == 4) {
          prof_.outer_ms_f32 += ms;
          prof_.outer_launches_f32++;
        }
      }
    }
    ev_mode_.clear();
    prof_.factorizations++;
    if (sizeof(T) == 4) prof_.f32_factorizations++;
    prof_.factor_flops += S.flops;
  }
  return herr ? -4 : 0;
}

int SpdDevice::factorize_device(const double* d_vals, int precision) {
  if (!analyzed_) throw Error(-5, "Analyze must be called first.");
  // the factor storage is overwritten from here on: a factorization that throws or fails leaves no usable factor
  factored_ = false;
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
    hipLaunchKernelGGL(k_to_f32, dim3(grid_stride(n, 256)), dim3(256), 0, stream_, (const double*)L_.p, L32_.p, n);
    IFEA_HIP(hipGetLastError());
  }
  if (rc == 0 && precision != 0) {
    // the FP32 sweeps read FP32 copies of the step-block inverses (computed in FP64 by k_inv_diag)
    long long nx = (long long)X_.n;
    X32_.alloc(nx);
    XT32_.alloc(nx);
    hipLaunchKernelGGL(k_to_f32, dim3(grid_stride(nx, 256)), dim3(256), 0, stream_, (const double*)X_.p, X32_.p, nx);
    hipLaunchKernelGGL(k_to_f32, dim3(grid_stride(nx, 256)), dim3(256), 0, stream_, (const double*)XT_.p, XT32_.p,
                       nx);
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
  const int cap = 7680;  // doubles of LDS per forward/backward work vector (plus 2.5 KB of static LDS)
  // explicit inverses of the wide fronts' step blocks in the sweep's precision (FP32 copies made per factorization)
  const T* Xs = sizeof(T) == 4 ? (const T*)X32_.p : (const T*)X_.p;
  const T* XTs = sizeof(T) == 4 ? (const T*)XT32_.p : (const T*)XT_.p;
  auto launch = [&](auto kern, const LaunchList& ll, const int4* tasks, auto... args) {
    if (ll.end > ll.begin)
      hipLaunchKernelGGL(kern, dim3((unsigned)(ll.end - ll.begin)), dim3(256), 0, stream_, tasks + ll.begin, sd_,
                         args...);
  };
  auto launch_n = [&](auto kern, int threads, const LaunchList& ll, const int4* tasks, auto... args) {
    if (ll.end > ll.begin)
      hipLaunchKernelGGL(kern, dim3((unsigned)(ll.end - ll.begin)), dim3(threads), 0, stream_, tasks + ll.begin, sd_,
                         args...);
  };
  // Within a level the narrow fronts (one workgroup each) and the wide fronts' step chain touch disjoint fronts: the
  // narrow ones run on the side stream, forked from the level's start and joined before the next level (in the
  // captured graph, two parallel branches per level)
  auto fork = [&](size_t l) {
    IFEA_HIP(hipEventRecord(sweep_ev_[2 * l], stream_));
    IFEA_HIP(hipStreamWaitEvent(side_, sweep_ev_[2 * l], 0));
  };
  auto join = [&](size_t l) {
    IFEA_HIP(hipEventRecord(sweep_ev_[2 * l + 1], side_));
    IFEA_HIP(hipStreamWaitEvent(stream_, sweep_ev_[2 * l + 1], 0));
  };
  sweep_events();
  for (size_t l = 0; l < levels_.size(); l++) {
    const Level& lv = levels_[l];
    unsigned cnt = (unsigned)(lv.small_end - lv.small_begin);
    const bool both = cnt && lv.fwd_gather.end > lv.fwd_gather.begin;
    if (cnt) {
      size_t sh = (size_t)std::min<long long>(lv.small_max_m, cap) * sizeof(double);
      if (both) fork(l);
      hipLaunchKernelGGL(k_fwd<T>, dim3(cnt), dim3(256), sh, both ? side_ : stream_, sweep_sn_.p + lv.small_begin,
                         sd_, L, Linv_.p, y_.p, V_.p, W_.p, cap);
    }
    launch(k_fwdb_gather, lv.fwd_gather, fwd_tasks_.p, (const double*)y_.p, (const double*)V_.p, W_.p);
    for (size_t t = 0; t < lv.fwd_diag.size(); t++) {
      launch_n(k_fwdb_diag<T>, DIAG_WAVES * 64, lv.fwd_diag[t], fwd_tasks_.p, XTs, y_.p, (const double*)W_.p);
      launch_n(k_fwdb_panel<T>, RECT_WAVES * 64, lv.fwd_panel[t], fwd_tasks_.p, L, (const double*)y_.p, W_.p);
    }
    // the FP32 sweeps take the update-vector rows four per lane (RECT_ROWS-row tasks), the FP64 ones one per lane
    const LaunchList& lr = sizeof(T) == 4 ? lv.fwd_rect4 : lv.fwd_rect;
    if (lr.end > lr.begin)
      hipLaunchKernelGGL(k_fwdb_rect<T>, dim3((unsigned)(lr.end - lr.begin)), dim3(RECT_WAVES * 64), 0, stream_,
                         fwd_tasks_.p + lr.begin, 0LL, sd_, L, (const double*)y_.p, (const double*)W_.p, V_.p, Pr_.p);
    if (sizeof(T) == 4)
      launch_n(k_fwdb_rect_sum<RECT_ROWS>, RECT_ROWS, lv.fwd_rect4_sum, fwd_tasks_.p, (const double*)W_.p, V_.p,
               (const double*)Pr_.p);
    else
      launch_n(k_fwdb_rect_sum<NB>, NB, lv.fwd_rect_sum, fwd_tasks_.p, (const double*)W_.p, V_.p, (const double*)Pr_.p);
    if (both) join(l);
  }
  for (size_t l = levels_.size(); l-- > 0;) {
    const Level& lv = levels_[l];
    unsigned cnt = (unsigned)(lv.small_end - lv.small_begin);
    const bool both = cnt && lv.bwd_gemv.end > lv.bwd_gemv.begin;
    if (cnt) {
      size_t sh = (size_t)std::min<long long>(lv.small_max_m, cap) * sizeof(double);
      if (both) fork(l);
      hipLaunchKernelGGL(k_bwd<T>, dim3(cnt), dim3(256), sh, both ? side_ : stream_, sweep_sn_.p + lv.small_begin,
                         sd_, L, Linv_.p, y_.p, W_.p, cap);
    }
    launch(k_bwdb_xb, lv.bwd_xb, bwd_tasks_.p, (const double*)y_.p, V_.p);
    launch(k_bwdb_gemv<T>, lv.bwd_gemv, bwd_tasks_.p, L, (const double*)y_.p, (const double*)V_.p, W_.p);
    for (size_t t = 0; t < lv.bwd_diag.size(); t++) {
      launch_n(k_bwdb_diag<T>, DIAG_WAVES * 64, lv.bwd_diag[t], bwd_tasks_.p, Xs, y_.p, (const double*)W_.p);
      launch(k_bwdb_panel<T>, lv.bwd_panel[t], bwd_tasks_.p, L, (const double*)y_.p, W_.p);
    }
    if (both) join(l);
  }
}

// The sweeps' side stream (the narrow fronts of a level, beside the wide fronts' chain on the main stream) and one
// fork/join event pair per level, created once (default priority).
void SpdDevice::sweep_events() {
  if (!side_) IFEA_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
  while (sweep_ev_.size() < 2 * levels_.size()) {
    hipEvent_t e;
    IFEA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    sweep_ev_.push_back(e);
  }
}

// The factorization's two streams: the outer updates and extend-adds on `outer_`, the panel chains on `chain_` (both
// at the default priority).
hipStream_t SpdDevice::chain_stream() {
  if (!chain_) {
    IFEA_HIP(hipStreamCreateWithFlags(&chain_, hipStreamNonBlocking));
    IFEA_HIP(hipStreamCreateWithFlags(&outer_, hipStreamNonBlocking));
    IFEA_HIP(hipEventCreateWithFlags(&ev_main_, hipEventDisableTiming));
    IFEA_HIP(hipEventCreateWithFlags(&ev_chain_, hipEventDisableTiming));
  }
  return chain_;
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
    if (sizeof(T) == 4) prof_.f32_solves++;
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
