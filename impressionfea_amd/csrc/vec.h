// Deterministic vector kernels: fixed-grid partial sums reduced in a fixed tree (no float atomics), so every
// reduction gives the same bits run to run, like the reference's serial sums (Forging3DSolver.cs:1066-1110).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace ifea {

constexpr int RED_BLOCKS = 512;
constexpr int RED_THREADS = 256;

// Scalar slots kept on the device.
enum Slot : int {
  S_E_NODE = 0, S_E_ELEM, S_P_NODE, S_P_ELEM,   // energy / power partial sums
  S_DOT0, S_DOT1, S_DOT2, S_DOT3,               // generic dot products
  S_RZ, S_RZN, S_PQ, S_ALPHA, S_BETA,           // PCG scalars
  S_REACTION, S_MAXT, S_MAXB, S_MAXDIAG,         // final evaluation / BC
  S_COUNT = 24
};

class Vec {
 public:
  explicit Vec(hipStream_t s);
  hipStream_t stream() const { return s_; }
  double* slots() { return slots_.p; }
  // slot <- sum_i a[i] * b[i] (b may be null: sum of a; b == a: squared norm)
  void dot(const double* a, const double* b, long long n, int slot);
  // read slots [first, first+count) to host (synchronises the stream)
  void read(int first, int count, double* out);
  void read_u64(int slot, unsigned long long* out);
  void zero_slot(int slot);

 private:
  hipStream_t s_;
  DBuf<double> part_, slots_;
};

// y = a*x + b*y
void axpby(hipStream_t s, long long n, double a, const double* x, double b, double* y);
// y = x + beta * y with beta from a device slot ratio num/den (PCG direction update)
void pcg_update_x_r(hipStream_t s, long long n, const double* slots, const double* p, const double* q, double* x,
                    double* r);
void pcg_update_p(hipStream_t s, long long n, double* slots, const double* z, double* p);
// full symmetric CSC product y = A x (row i gathered from column i)
void spmv_sym(hipStream_t s, int n, const int* colPtr, const int* rowIdx, const double* val, const double* x, double* y);

}  // namespace ifea
