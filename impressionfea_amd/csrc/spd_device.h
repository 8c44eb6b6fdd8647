// Device supernodal Cholesky (multifrontal, level scheduled) for the ISpdSolver slot and the device-resident step.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "common.h"
#include "symbolic.h"

namespace ifea {

// Device view of the symbolic structure.
struct SymDev {
  const int* sn_first;
  const long long* sn_rowptr;
  const int* sn_rows;
  const long long* Loff;
  const long long* Uoff;
  const long long* Voff;
  const long long* Woff;
  const long long* rel_off;
  const int* rel;
  const int* child_ptr;
  const int* child_list;
  const long long* Ioff;   // inverse diagonal blocks (64x64 per 64 columns) of each supernode
  const long long* Xoff;   // wide fronts: explicit inverses of the sweep steps' diagonal blocks (SB x SB, last one smaller)
  const long long* Goff;   // wide fronts: gptr + Goff[s] has m + 1 entries into gsrc
  const long long* gptr;   // row-wise child contributions (indices into the update vectors V), in child order
  const long long* gsrc;
};

struct LaunchList {   // tasks of one launch: [begin, end) into a task array
  long long begin = 0, end = 0;
  int c0 = 0;         // factor step column (panel / syrk)
};

class SpdDevice {
 public:
  explicit SpdDevice(hipStream_t stream) : stream_(stream) {}
  ~SpdDevice();

  // Host symbolic + device structures. perm may be null (nested dissection).
  void analyze(int n, const int* colPtr, const int* rowIdx, const int* perm, const SymbolicOptions& opt);
  bool analyzed() const { return analyzed_; }
  const Symbolic& symbolic() const { return S_; }

  // Numeric factorization from full-CSC values already on the device. precision: 0 FP64 factor; 1 FP32 factor
  // (FP32 MFMA updates, FP64 diagonal-block POTRF and TRSM); 2 FP64 factor plus an FP32 copy of L for the solves.
  // Returns 0 or IFEA_ERR_NOT_SPD (-4).
  int factorize_device(const double* d_vals, int precision);
  bool factored() const { return factored_; }
  int precision() const { return prec_; }
  // For precision 2: solve with the FP64 factor instead of its FP32 copy (the refinement fallback).
  void use_fp64_solves(bool on) { solve64_ = on; }

  // x = A^{-1} b with the current factor; vectors in the original (unpermuted) order, on the device.
  // d_x may alias d_b.
  void solve_device(const double* d_b, double* d_x);

  double last_factor_ms() const { return last_factor_ms_; }

  // Optional per-launch event timing of the dominant kernels (adds one event pair per launch).
  struct Profile {
    long long syrk_launches = 0, solve_launches = 0, factorizations = 0, solves = 0;
    long long outer_launches = 0, f32_factorizations = 0, f32_solves = 0;
    double syrk_ms = 0, syrk_flops = 0;     // inner (64-column) and mid (256-column) updates; 2 * width per entry
    double outer_ms = 0, outer_flops = 0;   // outer (512-column) updates, k_syrk_big
    long long outer_launches_f32 = 0;       // the part of the outer updates done on an FP32 factor
    double outer_ms_f32 = 0, outer_flops_f32 = 0;
    double solve_ms = 0, solve_bytes = 0;   // one forward + backward sweep reads every panel once
    double factor_ms = 0, factor_flops = 0;
  };
  void set_profiling(bool on) { profiling_ = on; }
  const Profile& profile() const { return prof_; }
  void reset_profile() { prof_ = Profile(); }
  size_t device_bytes() const;

 private:
  template <typename T> int factorize_t(const double* d_vals);
  template <typename T> void solve_t(const T* L, const double* d_b, double* d_x);
  template <typename T> void sweeps(const T* L);  // forward + backward sweeps on y_ (captured into a HIP graph)

  hipStream_t stream_;
  // the factorization's second stream (the panel chains of the outer blocks) and the events that order the two
  hipStream_t chain_ = nullptr, outer_ = nullptr;
  hipEvent_t ev_main_ = nullptr, ev_chain_ = nullptr;
  hipStream_t chain_stream();
  // the sweeps' side stream and fork/join events (two per level)
  hipStream_t side_ = nullptr;
  std::vector<hipEvent_t> sweep_ev_;
  void sweep_events();
  Symbolic S_;
  bool analyzed_ = false, factored_ = false;
  int prec_ = 0;
  double last_factor_ms_ = 0;
  // device symbolic
  DBuf<int> perm_, sn_first_, sn_rows_, rel_, child_ptr_, child_list_;
  DBuf<long long> sn_rowptr_, Loff_, Uoff_, Voff_, Woff_, rel_off_, Ioff_;
  DBuf<int> amap_src_;
  DBuf<long long> amap_dst_;
  // task arrays (int4: a, b, c, d)
  DBuf<int4> ext_tasks_, panel_tasks_, syrk_tasks_, potrf_tasks_;
  DBuf<int> level_sn_, sweep_sn_;
  DBuf<int4> fwd_tasks_, bwd_tasks_, inv_tasks_;
  DBuf<long long> Xoff_, Goff_, gptr_, gsrc_;
  LaunchList inv_launch_;
  // launch schedule
  struct Level {
    std::vector<LaunchList> ext_rounds;    // children's columns that land in the fronts' own columns (before)
    std::vector<LaunchList> ext_rounds_u;  // children's columns that land in the fronts' update blocks (after)
    std::vector<LaunchList> panel_steps, syrk_steps, potrf_steps;
    long long sn_begin = 0, sn_end = 0;  // into level_sn_
    std::vector<double> syrk_flops;      // algorithmic flops per syrk step
    std::vector<LaunchList> mid_steps;   // mid updates (second inner block of an outer block), after the panel step
    std::vector<double> mid_flops;
    // outer (512-column) updates, launched after the matching panel step in two parts: outer_steps the tile columns
    // of the next outer block (its panel chain waits for them), outer_b_steps the rest of the trailing matrix, which
    // runs while the next block's panels are factored on the second stream
    std::vector<LaunchList> outer_steps, outer_b_steps;
    std::vector<double> outer_flops, outer_b_flops;
    // triangular sweeps: narrow fronts one workgroup each; wide fronts blocked over several launches
    long long small_begin = 0, small_end = 0, big_begin = 0, big_end = 0;  // into sweep_sn_
    long long small_max_m = 0;
    LaunchList fwd_gather, fwd_rect, fwd_rect_sum, fwd_rect4, fwd_rect4_sum, bwd_xb, bwd_gemv;
    std::vector<LaunchList> fwd_diag, fwd_panel, bwd_diag, bwd_panel;  // per SB-column step (backward descending)
    double solve_bytes = 0;              // panel bytes per element of storage (times sizeof(T) at run time)
    long long max_m = 0;
  };
  std::vector<Level> levels_;
  // numeric storage
  DBuf<unsigned char> L_, U_;
  DBuf<float> L32_;
  bool solve64_ = false;
  // the whole triangular-sweep sequence replayed as one HIP graph (rebuilt when the factor storage changes)
  hipGraphExec_t sweep_exec_ = nullptr;
  const void* sweep_key_ = nullptr;
  size_t sweep_tsize_ = 0;
  void drop_graph();
  DBuf<double> V_, W_, y_, Linv_, X_, XT_;
  DBuf<float> X32_, XT32_;
  DBuf<double> Pr_;  // partial sums of the column ranges of the forward update-vector tasks (one level at a time)  // FP32 copies of X_ and XT_ for the FP32 sweeps (precision 1 and 2)
  DBuf<int> err_;
  SymDev sd_{};
  bool profiling_ = false;
  Profile prof_;
  std::vector<hipEvent_t> ev_pool_;
  std::vector<char> ev_mode_;
  hipEvent_t event(size_t i);
};

}  // namespace ifea
