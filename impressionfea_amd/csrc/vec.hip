#include "vec.h"

namespace ifea {

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
    if (tid < off) sh[tid] += sh[tid + off];
    __syncthreads();
  }
  return sh[0];
}

__global__ __launch_bounds__(RED_THREADS) void k_dot_part(const double* __restrict__ a, const double* __restrict__ b,
                                                          long long n, int mode, double* __restrict__ part) {
  __shared__ double sh[RED_THREADS];
  double s = 0.0;
  const long long stride = (long long)RED_BLOCKS * RED_THREADS;
  for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += stride)
    s += mode == 0 ? a[i] : (mode == 1 ? a[i] * a[i] : a[i] * b[i]);
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(RED_THREADS) void k_dot_final(const double* __restrict__ part, double* __restrict__ slots,
                                                           int slot) {
  __shared__ double sh[RED_THREADS];
  double s = part[threadIdx.x] + part[threadIdx.x + RED_THREADS];
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) slots[slot] = t;
}

Vec::Vec(hipStream_t s) : s_(s) {
  part_.alloc(RED_BLOCKS);
  slots_.alloc(S_COUNT);
  IFEA_HIP(hipMemsetAsync(slots_.p, 0, S_COUNT * sizeof(double), s_));
}

void Vec::dot(const double* a, const double* b, long long n, int slot) {
  int mode = b == nullptr ? 0 : (b == a ? 1 : 2);
  hipLaunchKernelGGL(k_dot_part, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s_, a, b, n, mode, part_.p);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, s_, part_.p, slots_.p, slot);
}

void Vec::read(int first, int count, double* out) {
  IFEA_HIP(hipMemcpyAsync(out, slots_.p + first, count * sizeof(double), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

void Vec::read_u64(int slot, unsigned long long* out) {
  IFEA_HIP(hipMemcpyAsync(out, slots_.p + slot, sizeof(double), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

void Vec::zero_slot(int slot) { IFEA_HIP(hipMemsetAsync(slots_.p + slot, 0, sizeof(double), s_)); }

__global__ void k_axpby(long long n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

void axpby(hipStream_t s, long long n, double a, const double* x, double b, double* y) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256)), dim3(256), 0, s, n, a, x, b, y);
}

// alpha = rz / pq; x += alpha p; r -= alpha q
__global__ void k_pcg_xr(long long n, const double* __restrict__ slots, const double* __restrict__ p,
                         const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double alpha = slots[S_RZ] / slots[S_PQ];
  if (i < n) {
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
  }
}

void pcg_update_x_r(hipStream_t s, long long n, const double* slots, const double* p, const double* q, double* x,
                    double* r) {
  hipLaunchKernelGGL(k_pcg_xr, dim3(grid_for(n, 256)), dim3(256), 0, s, n, slots, p, q, x, r);
}

// beta = rzn / rz; p = z + beta p  (rz <- rzn is done by k_pcg_rz after all blocks read it)
__global__ void k_pcg_p(long long n, const double* __restrict__ slots, const double* __restrict__ z,
                        double* __restrict__ p) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double beta = slots[S_RZN] / slots[S_RZ];
  if (i < n) p[i] = z[i] + beta * p[i];
}
__global__ void k_pcg_rz(double* slots) { slots[S_RZ] = slots[S_RZN]; }

void pcg_update_p(hipStream_t s, long long n, double* slots, const double* z, double* p) {
  hipLaunchKernelGGL(k_pcg_p, dim3(grid_for(n, 256)), dim3(256), 0, s, n, slots, z, p);
  hipLaunchKernelGGL(k_pcg_rz, dim3(1), dim3(1), 0, s, slots);
}

// one wave per row (= column, the pattern is symmetric)
__global__ __launch_bounds__(256) void k_spmv_sym(int n, const int* __restrict__ colPtr, const int* __restrict__ rowIdx,
                                                  const double* __restrict__ val, const double* __restrict__ x,
                                                  double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;
  double s = 0.0;
  for (int p = colPtr[row] + lane; p < colPtr[row + 1]; p += 64) s += val[p] * x[rowIdx[p]];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = s;
}

void spmv_sym(hipStream_t s, int n, const int* colPtr, const int* rowIdx, const double* val, const double* x,
              double* y) {
  hipLaunchKernelGGL(k_spmv_sym, dim3(grid_for(n, 4)), dim3(256), 0, s, n, colPtr, rowIdx, val, x, y);
}

}  // namespace ifea
