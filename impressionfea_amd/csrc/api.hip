// C ABI of libifea_amd.so: the ISpdSolver backend slot (include/ifea_spd.h) and the device-resident solve step
// (include/ifea_step.h). Exceptions never cross the boundary: every entry point maps them to a return code and a
// per-thread message, as the reference's backends map native status codes to exceptions (CudssSolver.cs:147-150,
// MklPardisoSolver.cs:92-97).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ifea_spd.h"
#include "../../include/ifea_step.h"
#include "common.h"
#include "spd_device.h"
#include "step_device.h"
#include "vec.h"

namespace ifea {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }
}  // namespace ifea

using namespace ifea;

#define IFEA_GUARD(body)                                             \
  try {                                                              \
    body;                                                            \
    return 0;                                                        \
  } catch (const ifea::Error& e) {                                   \
    set_last_error(e.what());                                        \
    return e.code;                                                   \
  } catch (const std::exception& e) {                                \
    set_last_error(e.what());                                        \
    return IFEA_ERR_DEVICE;                                          \
  }

static int ensure_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    set_last_error("no HIP device available");
    return IFEA_ERR_NO_DEVICE;
  }
  if (hipSetDevice(0) != hipSuccess) {
    set_last_error("hipSetDevice(0) failed");
    return IFEA_ERR_NO_DEVICE;
  }
  return 0;
}

// ------------------------------------------------------------------------------------------------ ISpdSolver slot
struct ifea_spd {
  hipStream_t stream = nullptr;
  std::unique_ptr<SpdDevice> dev;
  std::unique_ptr<Vec> vec;
  SymbolicOptions opt;
  int precision = 0;
  double refine_tol = 1e-10;
  int refine_maxit = 30;
  int n = 0;
  long long nnz = 0;
  DBuf<int> colPtr, rowIdx;
  DBuf<double> vals, b, x, r, d, q;
  int last_refine = 0, fallbacks = 0;
  std::string name;
  std::vector<double> coords;
};

extern "C" {

const char* ifea_spd_last_error(void) { return ifea::last_error(); }

int ifea_spd_create(ifea_spd_t* handle) {
  if (!handle) { set_last_error("null handle pointer"); return IFEA_ERR_ARG; }
  *handle = nullptr;
  int rc = ensure_device();
  if (rc) return rc;
  IFEA_GUARD({
    auto h = std::make_unique<ifea_spd>();
    IFEA_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->dev.reset(new SpdDevice(h->stream));
    h->vec.reset(new Vec(h->stream));
    hipDeviceProp_t prop;
    IFEA_HIP(hipGetDeviceProperties(&prop, 0));
    h->name = std::string("ifea HIP supernodal Cholesky \xC2\xB7 ") + prop.name + " (" + prop.gcnArchName + ")";
    *handle = h.release();
  })
}

int ifea_spd_name(ifea_spd_t h, char* buf, int len) {
  if (!h || !buf || len <= 0) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  std::strncpy(buf, h->name.c_str(), len - 1);
  buf[len - 1] = 0;
  return 0;
}

int ifea_spd_set_coordinates(ifea_spd_t h, int n, const double* coords) {
  if (!h || n < 0) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  if (coords) h->coords.assign(coords, coords + 3 * (size_t)n);
  else h->coords.clear();
  return 0;
}

static int spd_analyze(ifea_spd_t h, int n, const int* colPtr, const int* rowIdx, const int* perm) {
  if (!h || n <= 0 || !colPtr || !rowIdx) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  if (!h->coords.empty() && h->coords.size() != 3 * (size_t)n) { set_last_error("coordinates do not match n"); return IFEA_ERR_ARG; }
  h->opt.coords = h->coords.empty() ? nullptr : h->coords.data();
  IFEA_GUARD({
    h->dev->analyze(n, colPtr, rowIdx, perm, h->opt);
    h->n = n;
    h->nnz = colPtr[n];
    h->colPtr.upload(colPtr, n + 1, h->stream);
    h->rowIdx.upload(rowIdx, h->nnz, h->stream);
    h->vals.alloc(h->nnz);
    h->b.alloc(n); h->x.alloc(n); h->r.alloc(n); h->d.alloc(n); h->q.alloc(n);
    IFEA_HIP(hipStreamSynchronize(h->stream));
  })
}

int ifea_spd_analyze(ifea_spd_t h, int n, const int* colPtr, const int* rowIdx) {
  return spd_analyze(h, n, colPtr, rowIdx, nullptr);
}

int ifea_spd_analyze_perm(ifea_spd_t h, int n, const int* colPtr, const int* rowIdx, const int* perm) {
  if (!perm) { set_last_error("null permutation"); return IFEA_ERR_ARG; }
  return spd_analyze(h, n, colPtr, rowIdx, perm);
}

int ifea_spd_factorize(ifea_spd_t h, const double* values) {
  if (!h || !values) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  if (!h->dev->analyzed()) { set_last_error("Analyze must be called first."); return IFEA_ERR_ORDER; }
  try {
    IFEA_HIP(hipMemcpyAsync(h->vals.p, values, h->nnz * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int rc = h->dev->factorize_device(h->vals.p, h->precision);
    if (rc == IFEA_ERR_NOT_SPD && h->precision == 1) {
      // not positive definite in FP32: factor in FP64 and keep an FP32 copy for the solves
      h->fallbacks++;
      rc = h->dev->factorize_device(h->vals.p, 2);
    }
    if (rc) { set_last_error("matrix is not positive definite"); return rc; }
    return 0;
  } catch (const ifea::Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {  // host allocation failures and the like never cross the C boundary
    set_last_error(e.what());
    return IFEA_ERR_DEVICE;
  }
}

// x = A^{-1} b; mixed precision refines in FP64: x += M^{-1}(b - A x) until |r| <= tol |b|, FP64 refactor on stall
static int spd_solve_impl(ifea_spd_t h) {
  const int n = h->n;
  SpdDevice& D = *h->dev;
  D.solve_device(h->b.p, h->x.p);
  h->last_refine = 0;
  if (D.precision() != 0) {
    h->vec->dot(h->b.p, h->b.p, n, S_DOT0);
    double bn2;
    h->vec->read(S_DOT0, 1, &bn2);
    double bn = std::sqrt(bn2), prev = INFINITY;
    int stall = 0;
    bool ok = false;
    for (int it = 0; it <= h->refine_maxit; it++) {
      spmv_sym(h->stream, n, h->colPtr.p, h->rowIdx.p, h->vals.p, h->x.p, h->q.p);
      IFEA_HIP(hipMemcpyAsync(h->r.p, h->b.p, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
      axpby(h->stream, n, -1.0, h->q.p, 1.0, h->r.p);
      h->vec->dot(h->r.p, h->r.p, n, S_DOT1);
      double rn2;
      h->vec->read(S_DOT1, 1, &rn2);
      double rn = std::sqrt(rn2);
      if (!(rn == rn)) break;
      if (rn <= h->refine_tol * bn) { ok = true; break; }
      if (rn > 0.5 * prev) { if (++stall >= 2) break; } else stall = 0;
      prev = rn;
      D.solve_device(h->r.p, h->d.p);
      axpby(h->stream, n, 1.0, h->d.p, 1.0, h->x.p);
      h->last_refine++;
    }
    if (!ok) {
      h->fallbacks++;
      if (D.precision() == 2) {
        D.use_fp64_solves(true);
      } else {
        int rc = D.factorize_device(h->vals.p, 0);
        if (rc) return rc;
      }
      D.solve_device(h->b.p, h->x.p);
    }
  }
  h->vec->dot(h->x.p, nullptr, n, S_DOT2);
  double chk;
  h->vec->read(S_DOT2, 1, &chk);
  if (!std::isfinite(chk)) { set_last_error("non-finite solution"); return IFEA_ERR_NOT_SPD; }
  return 0;
}

int ifea_spd_solve(ifea_spd_t h, const double* b, double* x) {
  if (!h || !b || !x) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  if (!h->dev->factored()) { set_last_error("Factorize must be called first."); return IFEA_ERR_ORDER; }
  try {
    IFEA_HIP(hipMemcpyAsync(h->b.p, b, h->n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int rc = spd_solve_impl(h);
    if (rc) return rc;
    IFEA_HIP(hipMemcpyAsync(x, h->x.p, h->n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    IFEA_HIP(hipStreamSynchronize(h->stream));
    return 0;
  } catch (const ifea::Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return IFEA_ERR_DEVICE;
  }
}

long long ifea_spd_factor_nnz(ifea_spd_t h) {
  if (!h || !h->dev->analyzed()) return 0;
  return h->dev->symbolic().nnzL;
}

int ifea_spd_set_option(ifea_spd_t h, int key, double value) {
  if (!h) { set_last_error("null handle"); return IFEA_ERR_ARG; }
  switch (key) {
    case IFEA_OPT_PRECISION: h->precision = (int)value >= 0 && (int)value <= 2 ? (int)value : 0; return 0;
    case IFEA_OPT_REFINE_TOL: h->refine_tol = value; return 0;
    case IFEA_OPT_REFINE_MAXIT: h->refine_maxit = (int)value; return 0;
    case IFEA_OPT_LEAF_BLOCKS: h->opt.leaf_blocks = std::max(1, (int)value); return 0;
    default: set_last_error("unknown option key"); return IFEA_ERR_ARG;
  }
}

int ifea_spd_symbolic(ifea_spd_t h, int* perm, int* parent, int* colcount) {
  if (!h) { set_last_error("null handle"); return IFEA_ERR_ARG; }
  if (!h->dev->analyzed()) { set_last_error("Analyze must be called first."); return IFEA_ERR_ORDER; }
  const Symbolic& S = h->dev->symbolic();
  if (perm) std::memcpy(perm, S.perm.data(), S.n * sizeof(int));
  if (parent) std::memcpy(parent, S.parent.data(), S.n * sizeof(int));
  if (colcount) std::memcpy(colcount, S.colcount.data(), S.n * sizeof(int));
  return 0;
}

int ifea_spd_stats(ifea_spd_t h, double* out) {
  if (!h || !out) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  for (int k = 0; k < 8; k++) out[k] = 0;
  if (h->dev->analyzed()) {
    const Symbolic& S = h->dev->symbolic();
    out[0] = S.flops; out[1] = S.ns; out[2] = S.max_level + 1; out[3] = (double)S.Lsize; out[4] = (double)S.Usize;
  }
  out[5] = h->last_refine;
  out[6] = h->fallbacks;
  out[7] = h->dev->last_factor_ms();
  return 0;
}

int ifea_symbolic_analyze(int n, const int* colPtr, const int* rowIdx, const int* perm_in, const double* coords,
                          int* perm_out, int* parent, int* colcount, long long* nnzL, double* stats6) {
  if (n <= 0 || !colPtr || !rowIdx) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  try {
    Symbolic S;
    SymbolicOptions opt;
    opt.coords = coords;
    int rc = symbolic_analyze(n, colPtr, rowIdx, perm_in, opt, S);
    if (rc) { set_last_error(S.error); return rc == -2 ? IFEA_ERR_PERM : (rc == -1 ? IFEA_ERR_ARG : IFEA_ERR_SYMBOLIC); }
    if (perm_out) std::memcpy(perm_out, S.perm.data(), n * sizeof(int));
    if (parent) std::memcpy(parent, S.parent.data(), n * sizeof(int));
    if (colcount) std::memcpy(colcount, S.colcount.data(), n * sizeof(int));
    if (nnzL) *nnzL = S.nnzL;
    if (stats6) {
      stats6[0] = S.flops; stats6[1] = S.ns; stats6[2] = S.max_level + 1; stats6[3] = (double)S.Lsize;
      stats6[4] = (double)S.Usize; stats6[5] = S.nblocks;
    }
    return 0;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return IFEA_ERR_SYMBOLIC;
  }
}

void ifea_spd_destroy(ifea_spd_t h) {
  if (!h) return;
  h->dev.reset();
  h->vec.reset();
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

// ------------------------------------------------------------------------------------------------ step
struct ifea_step {
  std::unique_ptr<StepDevice> dev;
};

int ifea_step_create(ifea_step_t* h, int nn, int ne, const int* tets, const int* colPtr, const int* rowIdx, int nf,
                     const int* faces, const int* face_owner) {
  if (!h) { set_last_error("null handle pointer"); return IFEA_ERR_ARG; }
  *h = nullptr;
  int rc = ensure_device();
  if (rc) return rc;
  IFEA_GUARD({
    auto s = std::make_unique<ifea_step>();
    s->dev.reset(new StepDevice(nn, ne, tets, colPtr, rowIdx, nf, faces, face_owner));
    *h = s.release();
  })
}

int ifea_step_set_params(ifea_step_t h, const ifea_step_params* p) {
  if (!h || !p) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->set_params(*p))
}

int ifea_step_set_option(ifea_step_t h, int key, double value) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->set_option(key, value))
}

int ifea_step_set_perm(ifea_step_t h, const int* perm) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->set_perm(perm))
}

int ifea_step_prepare(ifea_step_t h, const double* x, const double* ax, const int* nfix, const double* fixed_val,
                      const int* contact, const double* vguess) {
  if (!h || !x || !ax || !nfix || !fixed_val || !contact || !vguess) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->prepare(x, ax, nfix, fixed_val, contact, vguess))
}

int ifea_step_set_geometry(ifea_step_t h, const unsigned char* on_x0, const unsigned char* on_y0, int bottom_kind,
                           const double* bottom_p, int top_kind, const double* top_p) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->set_geometry(on_x0, on_y0, bottom_kind, bottom_p, top_kind, top_p))
}

int ifea_step_prepare_device(ifea_step_t h, const double* x, const int* contact, const double* vguess) {
  if (!h || !x || !contact || !vguess) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->prepare_device(x, contact, vguess))
}

int ifea_step_load_guess(ifea_step_t h, const double* vguess) {
  if (!h || !vguess) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->load_guess(vguess))
}

int ifea_step_get_frames(ifea_step_t h, double* ax, int* nfix, int* die_axis, double* fixed_val) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->get_frames(ax, nfix, die_axis, fixed_val))
}

int ifea_step_set_element_state(ifea_step_t h, const double* eps, const double* sigma) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->set_element_state(eps, sigma))
}

int ifea_step_assemble(ifea_step_t h, int matrix, int newton, const double* v, double* energy, double* power,
                       double* r, double* val) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->assemble_host(matrix, newton, v, energy, power, r, val))
}

int ifea_step_solve_velocity(ifea_step_t h, int warm, ifea_step_stats* stats) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->solve_velocity(warm, stats))
}

int ifea_step_final(ifea_step_t h, const int* die_axis, double* power, double* reaction, double* vg, double* r,
                    unsigned char* release) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->final_eval(die_axis, power, reaction, vg, r, release))
}

int ifea_step_advance(ifea_step_t h, double dt, double* eps, double* sigma, double* rate) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->advance(dt, eps, sigma, rate))
}

int ifea_step_element_results(ifea_step_t h, double* mean, double* damage, double* dev6) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->element_results(mean, damage, dev6))
}

int ifea_step_set_damage(ifea_step_t h, const double* damage) {
  if (!h) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->set_damage(damage))
}

int ifea_step_get_v(ifea_step_t h, double* v) {
  if (!h || !v) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->get_v(v))
}

int ifea_step_set_v(ifea_step_t h, const double* v) {
  if (!h || !v) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->set_v(v))
}

int ifea_step_quality(ifea_step_t h, const double* x, double qmin, int* first_bad, double* quality) {
  if (!h || !x || !first_bad) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(*first_bad = h->dev->quality(x, qmin, quality))
}

int ifea_step_quality_noalloc(int nn, int ne, const int* tets, const double* x, double qmin, int* first_bad,
                              double* quality) {
  if (!tets || !x || !first_bad) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  int rc = ensure_device();
  if (rc) return rc;
  IFEA_GUARD(*first_bad = quality_noalloc(nn, ne, tets, x, qmin, quality))
}

int ifea_step_info(ifea_step_t h, int reset, double* out24) {
  if (!h || !out24) { set_last_error("invalid argument"); return IFEA_ERR_ARG; }
  IFEA_GUARD(h->dev->info(reset, out24))
}

void ifea_step_destroy(ifea_step_t h) { delete h; }

}  // extern "C"
