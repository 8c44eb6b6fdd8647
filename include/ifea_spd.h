/* ifea_spd.h — MI355X sparse SPD direct solver behind ImpressionFEA's ISpdSolver slot.
 *
 * Each entry point replaces one member of the reference interface
 *   src/ImpressionFEA.Core/Solver/Linear/ISpdSolver.cs:24-41 (ISpdSolver)
 * and follows the call conventions of the backends that fill that slot today
 *   src/ImpressionFEA.Core/Solver/Linear/CudssSolver.cs:70-165,
 *   src/ImpressionFEA.Core/Solver/Linear/MklPardisoSolver.cs:20-111.
 *
 * Matrices are passed as the reference passes them: a FULL symmetric CSC pattern (both triangles, zero-based,
 * sorted row indices) and values in that pattern's order. Arrays are caller-owned and only read during the call;
 * the library copies what it keeps. A handle owns all of its device memory and its own HIP stream; handles are
 * independent and may be used from different threads at the same time (one thread per handle at a time).
 *
 * Return codes: 0 success; IFEA_ERR_NOT_SPD (-4, as PARDISO's -4) for a matrix that is not positive definite or a
 * non-finite solution; IFEA_ERR_ORDER for a call-order violation ("Analyze must be called first."); other negative
 * values for invalid arguments or device failures. ifea_spd_last_error() gives the message.
 */
#ifndef IFEA_SPD_H
#define IFEA_SPD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifea_spd* ifea_spd_t;

#define IFEA_OK 0
#define IFEA_ERR_ARG -1
#define IFEA_ERR_PERM -2
#define IFEA_ERR_SYMBOLIC -3
#define IFEA_ERR_NOT_SPD -4
#define IFEA_ERR_ORDER -5
#define IFEA_ERR_DEVICE -6
#define IFEA_ERR_NO_DEVICE -7

/* Option keys for ifea_spd_set_option. */
#define IFEA_OPT_PRECISION 1      /* 0: FP64 factor (default). 1: mixed: FP32 supernode updates on MFMA with FP64
                                     diagonal-block POTRF; 2: FP64 factor, FP32 copy of L for the solves. In 1 and 2
                                     ifea_spd_solve refines in FP64 (see below). */
#define IFEA_OPT_REFINE_TOL 2     /* relative residual target of the mixed-precision refinement (default 1e-10) */
#define IFEA_OPT_REFINE_MAXIT 3   /* refinement iterations before the automatic FP64 refactorisation (default 30) */
#define IFEA_OPT_LEAF_BLOCKS 4    /* nested-dissection leaf size in column blocks (default 32) */

/* new CudssSolver / MklPardisoSolver (CudssSolver.cs:70-103, MklPardisoSolver.cs:20-31). Selects device 0. */
int ifea_spd_create(ifea_spd_t* handle);

/* ISpdSolver.Name (ISpdSolver.cs:27): backend and device description, e.g. "ifea HIP supernodal · AMD Instinct MI355X". */
int ifea_spd_name(ifea_spd_t handle, char* buf, int len);

/* ISpdSolver.Analyze (ISpdSolver.cs:32; CudssSolver.cs:105-124): ordering, elimination tree, supernodes, device
 * structures, for a full symmetric CSC pattern. */
int ifea_spd_analyze(ifea_spd_t handle, int n, const int* colPtr, const int* rowIdx);

/* Analyze with a caller ordering: pivot k eliminates column perm[k] (PARDISO iparm[4] = 1 counterpart,
 * MklPardisoSolver.cs:20-31). With the same permutation the elimination tree and factor pattern are identical to any
 * other backend's. */
int ifea_spd_analyze_perm(ifea_spd_t handle, int n, const int* colPtr, const int* rowIdx, const int* perm);

/* ISpdSolver.Factorize (ISpdSolver.cs:35; CudssSolver.cs:126-132): numeric factorisation, values in full-CSC order. */
int ifea_spd_factorize(ifea_spd_t handle, const double* values);

/* ISpdSolver.Solve (ISpdSolver.cs:37; CudssSolver.cs:134-145): x = A^{-1} b. In mixed precision the FP32 factor is
 * refined in FP64 against the factorised matrix until the relative residual target, falling back to an FP64
 * refactorisation when refinement stalls. A non-finite result returns IFEA_ERR_NOT_SPD. */
int ifea_spd_solve(ifea_spd_t handle, const double* b, double* x);

/* ISpdSolver.FactorNonZeros (ISpdSolver.cs:40): structural non-zeros of L (diagonal included); 0 before Analyze. */
long long ifea_spd_factor_nnz(ifea_spd_t handle);

/* Knobs that the reference sets through backend constructors (precision, refinement, ordering). */
int ifea_spd_set_option(ifea_spd_t handle, int key, double value);

/* Results of the last Analyze, for parity checks: pivot order perm[n], elimination tree parent[n] (pivot order,
 * -1 for roots), column counts colcount[n] (diagonal included). Any pointer may be NULL. */
int ifea_spd_symbolic(ifea_spd_t handle, int* perm, int* parent, int* colcount);

/* Statistics: [0] factor flops (sum colcount^2), [1] supernodes, [2] levels, [3] L storage words, [4] update
 * storage words, [5] refinement iterations of the last solve, [6] FP64 fallbacks so far, [7] last factor ms. */
int ifea_spd_stats(ifea_spd_t handle, double* out8);

/* Host-only symbolic analysis, no device needed (no counterpart in the reference; exposes Analyze's ordering and
 * structure for parity checks): perm_in may be NULL (nested dissection). Outputs as ifea_spd_symbolic, plus
 * nnzL (exact structural non-zeros of L) and stats6 = [flops, supernodes, levels, L words, update words, blocks].
 * Any output may be NULL. */
int ifea_symbolic_analyze(int n, const int* colPtr, const int* rowIdx, const int* perm_in, const double* coords,
                          int* perm_out, int* parent, int* colcount, long long* nnzL, double* stats6);

/* Optional column coordinates (3 doubles per column, e.g. the node position of each DOF) used by the next Analyze
 * without a permutation: nested dissection then cuts planes through the mesh instead of graph level sets. No
 * counterpart in the reference (its backends see the pattern only); NULL clears them. */
int ifea_spd_set_coordinates(ifea_spd_t handle, int n, const double* coords);

/* Dispose (ISpdSolver : IDisposable; CudssSolver.ReleasePattern CudssSolver.cs:152-165). NULL is ignored. */
void ifea_spd_destroy(ifea_spd_t handle);

/* Message of the last failing call on this thread. */
const char* ifea_spd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
