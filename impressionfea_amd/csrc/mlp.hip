// Per-node warm-start network on MI355X: layer products on BF16 MFMA with FP32 accumulation, backward, Adam.
// See include/ifea_mlp.h. All reductions over nodes are fixed-order partial sums.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ifea_mlp.h"
#include "common.h"

namespace ifea {

constexpr int ROWS = 64;        // rows (nodes) per workgroup in the layer kernels
constexpr int MAXF = 128;       // widest layer
constexpr int WCHUNKS = 128;    // node chunks of the weight-gradient reduction

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// ---------------------------------------------------------------- the products on BF16 MFMA
// Every product of the network is a tile C[64 x 128] = sum_r A(row, r) B(r, col) with both operands rounded to BF16
// (nearest even) and FP32 accumulation: v_mfma_f32_16x16x16_bf16, four waves of 16 rows x 128 columns (8 blocks of
// 16x16) per workgroup. Each 16-deep step of the reduction is staged in LDS as BF16, A row-major (the 4 consecutive
// r a lane feeds are one 8-byte read) and B column-major (likewise). The operand accessors return 0 outside the
// matrices, so partial tiles and feature counts below 16 (the 11 inputs, the 3 outputs) need no other handling.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int TR = 64, TC = 128, TK = 16;  // tile rows, tile columns, reduction step
constexpr int SA_LD = TK + 4, SB_LD = TK + 4;  // LDS row strides in BF16 elements (8-byte aligned, spread banks)

// lane l of the 16x16x16 instruction feeds A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16], i = 0..3, and
// receives D[4 (l / 16) + i][l % 16]
template <class FA, class FB>
__device__ __forceinline__ void mfma_tile(int R, FA fa, FB fb, unsigned short (*sA)[SA_LD],
                                          unsigned short (*sB)[SB_LD], f32x4 acc[TC / 16]) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int l16 = lane & 15, kq = 4 * (lane >> 4);
#pragma unroll
  for (int c = 0; c < TC / 16; c++) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int r0 = 0; r0 < R; r0 += TK) {
    __syncthreads();  // the previous step's operands are consumed
    // A: 64 x 16 (4 per thread), B: 16 x 128 (8 per thread), all loads before the LDS writes
    float va[TR * TK / 256], vb[TK * TC / 256];
#pragma unroll
    for (int u = 0; u < TR * TK / 256; u++) {
      const int e = tid + 256 * u, row = e / TK, r = e % TK;
      va[u] = fa(row, r0 + r);
    }
#pragma unroll
    for (int u = 0; u < TK * TC / 256; u++) {
      const int e = tid + 256 * u, col = e / TK, r = e % TK;
      vb[u] = fb(r0 + r, col);
    }
#pragma unroll
    for (int u = 0; u < TR * TK / 256; u++) {
      const int e = tid + 256 * u;
      sA[e / TK][e % TK] = f2bf(va[u]);
    }
#pragma unroll
    for (int u = 0; u < TK * TC / 256; u++) {
      const int e = tid + 256 * u;
      sB[e / TK][e % TK] = f2bf(vb[u]);
    }
    __syncthreads();
    bf16x4 av;
    __builtin_memcpy(&av, &sA[16 * wv + l16][kq], 8);
#pragma unroll
    for (int c = 0; c < TC / 16; c++) {
      bf16x4 bv;
      __builtin_memcpy(&bv, &sB[16 * c + l16][kq], 8);
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[c], 0, 0, 0);
    }
  }
}

// Z = A W + b, Aout = gelu(Z) (or Z); A [n x K] FP32, W [K x M] BF16; 64 rows per workgroup
__global__ __launch_bounds__(256) void k_lin_fwd(const float* __restrict__ A, int n, int K,
                                                 const unsigned short* __restrict__ W, const float* __restrict__ b,
                                                 int M, int act, float* __restrict__ Z, float* __restrict__ Aout) {
  __shared__ unsigned short sA[TR][SA_LD], sB[TC][SB_LD];
  const int r0 = blockIdx.x * TR;
  f32x4 acc[TC / 16];
  mfma_tile(
      K, [&](int row, int k) { return r0 + row < n && k < K ? A[(long long)(r0 + row) * K + k] : 0.0f; },
      [&](int k, int col) { return k < K && col < M ? bf2f(W[k * M + col]) : 0.0f; }, sA, sB, acc);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < TC / 16; c++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = r0 + 16 * wv + 4 * (lane >> 4) + i, col = 16 * c + (lane & 15);
      if (row < n && col < M) {
        const float z = acc[c][i] + b[col];
        const long long o = (long long)row * M + col;
        if (Z) Z[o] = z;
        Aout[o] = act ? gelu(z) : z;
      }
    }
}

// dA[n x K] = (dZ[n x M] W^T) * (act ? gelu'(Zp) : 1)
__global__ __launch_bounds__(256) void k_lin_bwd_data(const float* __restrict__ dZ, int n, int M,
                                                      const unsigned short* __restrict__ W, int K,
                                                      const float* __restrict__ Zp, int act, float* __restrict__ dA) {
  __shared__ unsigned short sA[TR][SA_LD], sB[TC][SB_LD];
  const int r0 = blockIdx.x * TR;
  f32x4 acc[TC / 16];
  mfma_tile(
      M, [&](int row, int m) { return r0 + row < n && m < M ? dZ[(long long)(r0 + row) * M + m] : 0.0f; },
      [&](int m, int k) { return m < M && k < K ? bf2f(W[k * M + m]) : 0.0f; }, sA, sB, acc);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < TC / 16; c++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = r0 + 16 * wv + 4 * (lane >> 4) + i, k = 16 * c + (lane & 15);
      if (row < n && k < K) {
        const long long o = (long long)row * K + k;
        dA[o] = act ? acc[c][i] * gelu_d(Zp[o]) : acc[c][i];
      }
    }
}

// part[c][(K+1) x M]: over chunk c's rows, A^T dZ on MFMA (rows of the product: 64 input features per workgroup,
// blockIdx.y the feature tile) and, by the first feature tile, the bias gradient sum(dZ) in FP32, rows ascending
__global__ __launch_bounds__(256) void k_lin_bwd_w(const float* __restrict__ A, const float* __restrict__ dZ, int n,
                                                   int K, int M, int rows_per_chunk, float* __restrict__ part) {
  __shared__ unsigned short sA[TR][SA_LD], sB[TC][SB_LD];
  const int c = blockIdx.x, k0 = blockIdx.y * TR;
  const int r_begin = c * rows_per_chunk, r_end = min(n, r_begin + rows_per_chunk);
  const int nr = r_end - r_begin;
  f32x4 acc[TC / 16];
  mfma_tile(
      nr, [&](int kk, int r) { return k0 + kk < K && r < nr ? A[(long long)(r_begin + r) * K + k0 + kk] : 0.0f; },
      [&](int r, int m) { return r < nr && m < M ? dZ[(long long)(r_begin + r) * M + m] : 0.0f; }, sA, sB, acc);
  float* pc = part + (long long)c * (K + 1) * M;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int cc = 0; cc < TC / 16; cc++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = k0 + 16 * wv + 4 * (lane >> 4) + i, m = 16 * cc + (lane & 15);
      if (k < K && m < M) pc[k * M + m] = acc[cc][i];
    }
  if (blockIdx.y == 0 && threadIdx.x < M) {
    float sb = 0.0f;
    for (int r = r_begin; r < r_end; r++) sb += dZ[(long long)r * M + threadIdx.x];
    pc[K * M + threadIdx.x] = sb;
  }
}

// out[i] = sum_c part[c][i], c ascending
__global__ void k_reduce_part(const float* __restrict__ part, int C, int len, float* __restrict__ wgrad,
                              float* __restrict__ bgrad, int K, int M) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= len) return;
  float s = 0.0f;
  for (int c = 0; c < C; c++) s += part[(long long)c * len + o];
  int k = o / M, m = o % M;
  if (k < K) wgrad[o] = s;
  else bgrad[m] = s;
}

// dY = 2 (yhat - y) / (n * nout); per-block partial of the squared error
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ yh, const float* __restrict__ y, long long len,
                                             float scale, float* __restrict__ dY, double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < len; i += 256LL * gridDim.x) {
    float d = yh[i] - y[i];
    dY[i] = 2.0f * d * scale;
    s += (double)d * d;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_part(const double* __restrict__ part, int np, double scale, float* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < np; i++) s += part[i];
    *out = (float)(s * scale);
  }
}

__global__ void k_adam(long long np, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float c1, float c2, unsigned short* __restrict__ w16) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= np) return;
  float gi = g[i];
  float mi = 0.9f * m[i] + 0.1f * gi;
  float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
  w[i] = wi;
  w16[i] = f2bf(wi);
}

__global__ void k_to_bf16(long long np, const float* __restrict__ w, unsigned short* __restrict__ w16) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < np) w16[i] = f2bf(w[i]);
}

__global__ void k_gather_rows(const float* __restrict__ src, const int* __restrict__ idx, int count, int width,
                              float* __restrict__ dst) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)count * width) return;
  long long r = i / width, c = i % width;
  dst[i] = src[(long long)idx[r] * width + c];
}

struct Mlp {
  hipStream_t s = nullptr;
  int n_in, width, depth, n_out, nl;
  std::vector<int> fin, fout;
  std::vector<long long> woff, boff;
  long long np = 0;
  DBuf<float> w, m, v, g, part, loss1;
  DBuf<unsigned short> w16;
  DBuf<double> lpart;
  std::vector<DBuf<float>*> Z, A;  // per layer, rows x fout
  DBuf<float> X, Y, dA0, dA1, dY, Xd, Yd;  // batch input/target, gradient ping-pong, dataset
  DBuf<int> idx;
  int cap = 0, ndata = 0;
  long long step = 0;

  Mlp(int ni, int wd, int dp, int no) : n_in(ni), width(wd), depth(dp), n_out(no) {
    if (ni <= 0 || ni > MAXF || wd <= 0 || wd > MAXF || dp <= 0 || no <= 0 || no > MAXF)
      throw Error(-1, "network sizes must be in 1..128");
    nl = dp + 1;
    for (int l = 0; l < nl; l++) {
      fin.push_back(l == 0 ? ni : wd);
      fout.push_back(l == nl - 1 ? no : wd);
    }
    for (int l = 0; l < nl; l++) {
      woff.push_back(np);
      np += (long long)fin[l] * fout[l];
      boff.push_back(np);
      np += fout[l];
    }
    IFEA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    w.alloc(np); m.alloc(np); v.alloc(np); g.alloc(np); w16.alloc(np); loss1.alloc(1);
    IFEA_HIP(hipMemsetAsync(m.p, 0, np * 4, s));
    IFEA_HIP(hipMemsetAsync(v.p, 0, np * 4, s));
    IFEA_HIP(hipMemsetAsync(w.p, 0, np * 4, s));
    IFEA_HIP(hipMemsetAsync(w16.p, 0, np * 2, s));
    int maxlen = 0;
    for (int l = 0; l < nl; l++) maxlen = std::max(maxlen, (fin[l] + 1) * fout[l]);
    part.alloc((size_t)WCHUNKS * maxlen);
    lpart.alloc(256);
    for (int l = 0; l < nl; l++) { Z.push_back(new DBuf<float>()); A.push_back(new DBuf<float>()); }
  }
  ~Mlp() {
    for (auto* p : Z) delete p;
    for (auto* p : A) delete p;
    if (s) (void)hipStreamDestroy(s);
  }
  void ensure(int rows) {
    if (rows <= cap) return;
    cap = rows;
    for (int l = 0; l < nl; l++) { Z[l]->alloc((size_t)rows * fout[l]); A[l]->alloc((size_t)rows * fout[l]); }
    X.alloc((size_t)rows * n_in);
    Y.alloc((size_t)rows * n_out);
    dY.alloc((size_t)rows * n_out);
    dA0.alloc((size_t)rows * MAXF);
    dA1.alloc((size_t)rows * MAXF);
    idx.alloc(rows);
  }
  void forward(int n, const float* dX) {
    const float* in = dX;
    for (int l = 0; l < nl; l++) {
      int act = l < nl - 1;
      hipLaunchKernelGGL(k_lin_fwd, dim3((n + ROWS - 1) / ROWS), dim3(256), 0, s, in, n, fin[l], w16.p + woff[l],
                         w.p + boff[l], fout[l], act, Z[l]->p, A[l]->p);
      in = A[l]->p;
    }
  }
  // gradient of mean((yhat - y)^2) into g; loss (mean) into loss1
  void backward(int n, const float* dX, const float* dYt) {
    long long len = (long long)n * n_out;
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, s, A[nl - 1]->p, dYt, len, 1.0f / (float)len, dY.p, lpart.p);
    hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(64), 0, s, lpart.p, 256, 1.0 / (double)len, loss1.p);
    const float* dz = dY.p;
    for (int l = nl - 1; l >= 0; l--) {
      const float* in = l == 0 ? dX : A[l - 1]->p;
      int K = fin[l], M = fout[l];
      int rpc = (n + WCHUNKS - 1) / WCHUNKS;
      int C = (n + rpc - 1) / rpc;
      hipLaunchKernelGGL(k_lin_bwd_w, dim3(C, (K + TR - 1) / TR), dim3(256), 0, s, in, dz, n, K, M, rpc, part.p);
      int len2 = (K + 1) * M;
      hipLaunchKernelGGL(k_reduce_part, dim3((len2 + 255) / 256), dim3(256), 0, s, part.p, C, len2, g.p + woff[l],
                         g.p + boff[l], K, M);
      if (l > 0) {
        float* out = (l % 2) ? dA0.p : dA1.p;
        hipLaunchKernelGGL(k_lin_bwd_data, dim3((n + ROWS - 1) / ROWS), dim3(256), 0, s, dz, n, M, w16.p + woff[l],
                           K, Z[l - 1]->p, 1, out);
        dz = out;
      }
    }
  }
  void adam(float lr) {
    step++;
    float c1 = 1.0f - std::pow(0.9f, (float)step), c2 = 1.0f - std::pow(0.999f, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(np, 256)), dim3(256), 0, s, np, w.p, g.p, m.p, v.p, lr, c1, c2, w16.p);
  }
};

}  // namespace ifea

using namespace ifea;

struct ifea_mlp {
  std::unique_ptr<Mlp> m;
};

#define MLP_GUARD(body)                \
  try {                                \
    body;                              \
    return 0;                          \
  } catch (const ifea::Error& e) {     \
    set_last_error(e.what());          \
    return e.code;                     \
  } catch (const std::exception& e) {  \
    set_last_error(e.what());          \
    return -6;                         \
  }

extern "C" {

int ifea_mlp_create(ifea_mlp_t* h, int n_in, int width, int depth, int n_out) {
  if (!h) { set_last_error("null handle pointer"); return -1; }
  *h = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) { set_last_error("no HIP device available"); return -7; }
  MLP_GUARD({
    auto p = std::make_unique<ifea_mlp>();
    p->m.reset(new Mlp(n_in, width, depth, n_out));
    *h = p.release();
  })
}

long long ifea_mlp_num_params(ifea_mlp_t h) { return h ? h->m->np : 0; }

int ifea_mlp_set_params(ifea_mlp_t h, const float* params) {
  if (!h || !params) { set_last_error("invalid argument"); return -1; }
  MLP_GUARD({
    Mlp& M = *h->m;
    IFEA_HIP(hipMemcpyAsync(M.w.p, params, M.np * 4, hipMemcpyHostToDevice, M.s));
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(M.np, 256)), dim3(256), 0, M.s, M.np, M.w.p, M.w16.p);
    IFEA_HIP(hipMemsetAsync(M.m.p, 0, M.np * 4, M.s));
    IFEA_HIP(hipMemsetAsync(M.v.p, 0, M.np * 4, M.s));
    M.step = 0;
    IFEA_HIP(hipStreamSynchronize(M.s));
  })
}

int ifea_mlp_get_params(ifea_mlp_t h, float* params) {
  if (!h || !params) { set_last_error("invalid argument"); return -1; }
  MLP_GUARD({
    Mlp& M = *h->m;
    IFEA_HIP(hipMemcpyAsync(params, M.w.p, M.np * 4, hipMemcpyDeviceToHost, M.s));
    IFEA_HIP(hipStreamSynchronize(M.s));
  })
}

int ifea_mlp_forward(ifea_mlp_t h, int n, const float* x, float* y) {
  if (!h || n < 0 || (n && (!x || !y))) { set_last_error("invalid argument"); return -1; }
  MLP_GUARD({
    Mlp& M = *h->m;
    if (n == 0) return 0;
    M.ensure(n);
    IFEA_HIP(hipMemcpyAsync(M.X.p, x, (size_t)n * M.n_in * 4, hipMemcpyHostToDevice, M.s));
    M.forward(n, M.X.p);
    IFEA_HIP(hipMemcpyAsync(y, M.A[M.nl - 1]->p, (size_t)n * M.n_out * 4, hipMemcpyDeviceToHost, M.s));
    IFEA_HIP(hipStreamSynchronize(M.s));
  })
}

int ifea_mlp_loss_grad(ifea_mlp_t h, int n, const float* x, const float* y, float* loss, float* grad) {
  if (!h || n <= 0 || !x || !y) { set_last_error("invalid argument"); return -1; }
  MLP_GUARD({
    Mlp& M = *h->m;
    M.ensure(n);
    IFEA_HIP(hipMemcpyAsync(M.X.p, x, (size_t)n * M.n_in * 4, hipMemcpyHostToDevice, M.s));
    IFEA_HIP(hipMemcpyAsync(M.Y.p, y, (size_t)n * M.n_out * 4, hipMemcpyHostToDevice, M.s));
    M.forward(n, M.X.p);
    M.backward(n, M.X.p, M.Y.p);
    if (loss) IFEA_HIP(hipMemcpyAsync(loss, M.loss1.p, 4, hipMemcpyDeviceToHost, M.s));
    if (grad) IFEA_HIP(hipMemcpyAsync(grad, M.g.p, M.np * 4, hipMemcpyDeviceToHost, M.s));
    IFEA_HIP(hipStreamSynchronize(M.s));
  })
}

int ifea_mlp_set_dataset(ifea_mlp_t h, int n, const float* x, const float* y) {
  if (!h || n <= 0 || !x || !y) { set_last_error("invalid argument"); return -1; }
  MLP_GUARD({
    Mlp& M = *h->m;
    M.Xd.upload(x, (size_t)n * M.n_in, M.s);
    M.Yd.upload(y, (size_t)n * M.n_out, M.s);
    M.ndata = n;
    IFEA_HIP(hipStreamSynchronize(M.s));
  })
}

int ifea_mlp_train_epoch(ifea_mlp_t h, const int* perm, int batch, float lr, float* mean_loss) {
  if (!h || !perm || batch <= 0) { set_last_error("invalid argument"); return -1; }
  MLP_GUARD({
    Mlp& M = *h->m;
    if (M.ndata <= 0) throw Error(-5, "no dataset");
    for (int i = 0; i < M.ndata; i++)
      if (perm[i] < 0 || perm[i] >= M.ndata) throw Error(-1, "permutation index out of range");
    M.ensure(batch);
    DBuf<int> dperm;
    dperm.upload(perm, M.ndata, M.s);
    std::vector<float> losses;
    float l;
    for (int b0 = 0; b0 < M.ndata; b0 += batch) {
      int nb = std::min(batch, M.ndata - b0);
      hipLaunchKernelGGL(k_gather_rows, dim3(grid_for((long long)nb * M.n_in, 256)), dim3(256), 0, M.s, M.Xd.p,
                         dperm.p + b0, nb, M.n_in, M.X.p);
      hipLaunchKernelGGL(k_gather_rows, dim3(grid_for((long long)nb * M.n_out, 256)), dim3(256), 0, M.s, M.Yd.p,
                         dperm.p + b0, nb, M.n_out, M.Y.p);
      M.forward(nb, M.X.p);
      M.backward(nb, M.X.p, M.Y.p);
      M.adam(lr);
      IFEA_HIP(hipMemcpyAsync(&l, M.loss1.p, 4, hipMemcpyDeviceToHost, M.s));
      IFEA_HIP(hipStreamSynchronize(M.s));
      losses.push_back(l);
    }
    double s = 0;
    for (float q : losses) s += q;
    if (mean_loss) *mean_loss = (float)(s / std::max<size_t>(1, losses.size()));
  })
}

void ifea_mlp_destroy(ifea_mlp_t h) { delete h; }

}  // extern "C"
