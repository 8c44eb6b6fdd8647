// TEST INFRASTRUCTURE ONLY — Intel MKL PARDISO called through its documented C prototype, loaded with dlopen.
//
// Restates the reference's MklPardisoSolver (src/ImpressionFEA.Core/Solver/Linear/MklPardisoSolver.cs:20-99):
// mtype 2 (real SPD), iparm as set there (METIS nested dissection, no iterative refinement, pivot perturbation 1e-8,
// factor non-zeros reported, two-level parallel factorisation, zero-based indexing), phases 11 / 22 / 33 on the upper
// triangle in CSR order (UpperCsr, Solver/Linear/ISpdSolver.cs:44-94). The library is located through $IFEA_MKL_RT or
// the default conda path; it is only called, never inspected.
#pragma once
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <vector>

struct PardisoDL {
  typedef void (*pardiso_fn)(void*, const int*, const int*, const int*, const int*, const int*, const void*, const int*,
                             const int*, int*, const int*, int*, const int*, void*, void*, int*);
  typedef int (*getmax_fn)();
  typedef void (*setnum_fn)(int);
  pardiso_fn fn = nullptr;
  getmax_fn getmax = nullptr;
  void* pt[64] = {nullptr};
  int iparm[64] = {0};
  int n = 0;
  bool analyzed = false;
  std::vector<int> rowPtr, colIdx, map, perm, pbuf;
  std::vector<double> values;
  long long nnz_l = 0, mflops = 0, threads = 0;
  int want_perm = 0;

  static void* lib() {
    static void* h = nullptr;
    static bool tried = false;
    if (!tried) {
      tried = true;
      const char* env = std::getenv("IFEA_MKL_RT");
      const char* cands[] = {env, "/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so"};
      for (const char* c : cands) {
        if (!c) continue;
        h = dlopen(c, RTLD_NOW | RTLD_GLOBAL);
        if (h) break;
      }
    }
    return h;
  }
  PardisoDL() {
    void* h = lib();
    if (!h) return;
    fn = (pardiso_fn)dlsym(h, "pardiso");
    getmax = (getmax_fn)dlsym(h, "MKL_Get_Max_Threads");
    threads = getmax ? getmax() : 0;
    iparm[0] = 1;    // user-supplied iparm
    iparm[1] = 3;    // parallel METIS nested dissection
    iparm[7] = 0;    // no iterative refinement
    iparm[9] = 8;    // pivot perturbation 1e-8
    iparm[17] = -1;  // report factor non-zeros
    iparm[18] = -1;  // report factorisation MFLOP
    iparm[23] = 1;   // two-level parallel factorisation
    iparm[26] = 0;   // no matrix checker
    iparm[34] = 1;   // zero-based indexing
  }
  ~PardisoDL() { release(); }
  bool ok() const { return fn != nullptr; }
  static void set_threads(int t) {
    void* h = lib();
    if (!h) return;
    setnum_fn f = (setnum_fn)dlsym(h, "MKL_Set_Num_Threads");
    if (f) f(t);
  }
  int run(int phase, const double* b, double* x) {
    int maxfct = 1, mnum = 1, mtype = 2, nrhs = 1, msglvl = 0, error = 0;
    int* pp = (iparm[4] == 1 || iparm[4] == 2) ? pbuf.data() : nullptr;
    fn(pt, &maxfct, &mnum, &mtype, &phase, &n, values.data(), rowPtr.data(), colIdx.data(), pp, &nrhs, iparm, &msglvl,
       (void*)b, (void*)x, &error);
    if (error == -4) return -4;
    return error ? -20 + error : 0;
  }
  void release() {
    if (!analyzed || !fn) return;
    int phase = -1, maxfct = 1, mnum = 1, mtype = 2, nrhs = 1, msglvl = 0, error = 0;
    fn(pt, &maxfct, &mnum, &mtype, &phase, &n, values.data(), rowPtr.data(), colIdx.data(), nullptr, &nrhs, iparm,
       &msglvl, nullptr, nullptr, &error);
    analyzed = false;
    for (auto& q : pt) q = nullptr;
  }
  // UpperCsr (ISpdSolver.cs:46-78): row i of the upper triangle = CSC column i entries with row >= i
  int analyze(int n_, const int* colPtr, const int* rowIdx, const int* userPerm) {
    release();
    n = n_;
    rowPtr.assign(n + 1, 0);
    colIdx.clear();
    map.clear();
    for (int i = 0; i < n; i++) {
      for (int p = colPtr[i]; p < colPtr[i + 1]; p++)
        if (rowIdx[p] >= i) { colIdx.push_back(rowIdx[p]); map.push_back(p); }
      rowPtr[i + 1] = (int)colIdx.size();
    }
    values.assign(colIdx.size(), 0.0);
    perm.assign(n, 0);
    pbuf.assign(n, 0);
    if (userPerm) {
      // PARDISO's perm, given (iparm[4] = 1) or returned (iparm[4] = 2), is the pivot order: entry k is the original
      // row eliminated k-th. Measured on the 3,993-DOF upset: a known nested-dissection order passed this way gives
      // MKL's fill within 3 % of the exact count (relaxed supernodes), its inverse three times the fill.
      for (int k = 0; k < n; k++) pbuf[k] = userPerm[k];
      for (int k = 0; k < n; k++) perm[k] = userPerm[k];
      iparm[4] = 1;
    } else {
      iparm[4] = want_perm ? 2 : 0;
    }
    int rc = run(11, nullptr, nullptr);
    if (rc) return rc;
    analyzed = true;
    if (iparm[4] == 2) {  // the returned permutation is already a pivot order
      for (int i = 0; i < n; i++) perm[i] = pbuf[i];
      iparm[4] = 0;
    }
    return 0;
  }
  int factorize(const double* full) {
    if (!analyzed) return -30;
    for (size_t k = 0; k < map.size(); k++) values[k] = full[map[k]];
    int rc = run(22, nullptr, nullptr);
    nnz_l = iparm[17];
    mflops = iparm[18];
    return rc;
  }
  int solve(const double* b, double* x) { return run(33, b, x); }
};
