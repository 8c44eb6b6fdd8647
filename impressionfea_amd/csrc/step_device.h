// Device-resident solve step (see include/ifea_step.h).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../include/ifea_step.h"
#include "common.h"
#include "spd_device.h"
#include "vec.h"

namespace ifea {

// An analytic rigid die as the device evaluates it (kind 0 flat, 1 hex cavity, 2 con-rod half; side 0 bottom, 1 top)
struct DieDev {
  int kind = 0, side = 0;
  double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// TryUpdate's collapse check on host arrays through the library's static scratch array: no device allocation.
int quality_noalloc(int nn, int ne, const int* tets, const double* x, double qmin, double* q);

class StepDevice {
 public:
  StepDevice(int nn, int ne, const int* tets, const int* colPtr, const int* rowIdx, int nf, const int* faces,
             const int* face_owner);
  ~StepDevice();
  void set_params(const ifea_step_params& p) { P_ = p; have_params_ = true; spd_->set_profiling(p.profile != 0); }
  void set_perm(const int* perm);
  void set_option(int key, double value);
  void prepare(const double* x, const double* ax, const int* nfix, const double* fixed_val, const int* contact,
               const double* vguess);
  void set_element_state(const double* eps, const double* sigma);
  // PrepareStep with the nodal frames computed on the device from the symmetry flags and dies of set_geometry
  void set_geometry(const unsigned char* on_x0, const unsigned char* on_y0, int bottom_kind, const double* bottom_p,
                    int top_kind, const double* top_p);
  void prepare_device(const double* x, const int* contact, const double* vguess);
  // LoadVelocityGuess with the current frames: v = R vg, fixed DOFs overwritten
  void load_guess(const double* vguess);
  void get_frames(double* ax, int* nfix, int* die_axis, double* fixed_val);
  // mode: 0 raw residual, 1 BC + negated residual (linear system), 2 energy only
  double assemble(bool matrix, bool newton, const double* d_v, int mode, double* power);
  void assemble_host(int matrix, int newton, const double* v, double* energy, double* power, double* r, double* val);
  void solve_velocity(int warm, ifea_step_stats* st);
  void final_eval(const int* die_axis, double* power, double* reaction, double* vg, double* r, unsigned char* release);
  void advance(double dt, double* eps, double* sigma, double* rate);
  // element mean stress of the last FinalEvaluation, accumulated damage, deviatoric stress of the last assembly
  void element_results(double* mean, double* damage, double* dev6);
  void set_damage(const double* damage);
  void get_v(double* v);
  void set_v(const double* v);
  // index of the first tet whose quality at the host positions x is <= qmin, or -1; per-tet values into q if given
  int quality(const double* x, double qmin, double* q);
  void info(int reset, double* out);

 private:
  int solve_linear(double* b, double* x, ifea_step_stats* st);
  int pcg(const double* b, double* x, double tol, int maxIt, ifea_step_stats* st);
  void lin_solve(const double* b, double* x, ifea_step_stats* st);
  int refactor(int precision, ifea_step_stats* st);
  void prepare_rest(const double* x, const double* vguess);

  hipStream_t s_ = nullptr;
  std::unique_ptr<Vec> vec_;
  std::unique_ptr<SpdDevice> spd_;
  ifea_step_params P_{};
  bool have_params_ = false, prepared_ = false;
  int nn_, ne_, ndof_;
  long long nnz_;
  int max_deg_ = 0, max_patch_ = 0, max_elem_ = 0;
  std::vector<int> user_perm_;
  std::vector<int> h_colPtr_, h_rowIdx_;
  std::vector<double> h_coords_;  // 3 per DOF, for the geometric ordering
  bool need_analyze_ = true, have_factor_ = false;
  int pcg_failures_ = 0, pcg_skip_ = 0;
  bool fp32_unstable_ = false;  // an FP32 factorisation was not positive definite: factor in FP64 from now on
  bool fp32_sweeps_unstable_ = false;  // PCG on a fresh FP64 factor failed with FP32 sweeps: FP64 sweeps from now on
  int opt_first_solve_ = 0;            // IFEA_STEP_OPT_FIRST_SOLVE
  bool first_solve_ = false;           // the next solve_linear is the first of an increment
  int first_fail_streak_ = 0;          // increments in a row whose first solve missed the PCG budget
  int first_skips_ = 0;                // first solves refactored directly since the last PCG probe
  // topology (device)
  DBuf<int> tets_, nt_start_, nt_list_, colPtr_, rowIdx_, nbr_start_, nbr_, p_start_, p_node_, diagpos_;
  DBuf<int> nf_start_, nf_list_, faces_, face_owner_;
  // per increment
  DBuf<double> x_, ax_, fixed_val_, grad_, vol_, vnode_, gcoef_, kvol_, afric_, kfric_;
  DBuf<int> nfix_, contact_, bad_, die_axis_;
  DBuf<unsigned char> on_x0_, on_y0_;
  DieDev bottom_, top_;
  bool have_geometry_ = false, device_frames_ = false;
  // state
  DBuf<double> eps_, el_sig_, el_rate_, el_dev_, el_mean_, dmg_, nodal_;
  // iteration
  DBuf<double> v_, r_, dv_, vtrial_, val_, e_wm_, e_wn_, e_g_, e_energy_, e_power_, n_energy_, n_power_;
  DBuf<double> pr_, pz_, pp_, pq_, tmp_;
  DBuf<double> xq_, qual_;  // positions and per-tet values of the quality check
  DBuf<unsigned long long> maxdiag_;
};

}  // namespace ifea
