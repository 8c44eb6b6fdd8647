/* ifea_step.h — device-resident solve step of the 3D rigid-viscoplastic forging solver on MI355X.
 *
 * The reference keeps assembly private to Forging3DSolver (src/ImpressionFEA.Core/ThreeD/Forging3DSolver.cs:1115),
 * so this is a new seam with no counterpart in the reference: the host (C#, or the Python driver in this package)
 * keeps die geometry, nodal frames and contact (PrepareStep's host part, Forging3DSolver.cs:669-727) and the geometry
 * update (TryUpdate, :1455-1537); the library owns everything else of one increment on the GPU:
 *   element geometry and the ANP penalty matrix        (PrepareStep :730-753, BuildAnp :761-823)
 *   friction areas / shear stresses                    (BuildFriction :829-850)
 *   the warm start                                     (LoadVelocityGuess :854-869)
 *   assembly of residual and tangent                   (Assemble/ElementWork/Kernel :1115-1400)
 *   the Picard->Newton loop with line search, PCG with the cached factor and the refactorisation policy
 *                                                      (SolveVelocity :875-948, SolveLinear/Pcg :965-1087)
 *   the final residual, ram power and reaction         (FinalEvaluation :1402-1451) and the strain update (:626)
 * Arrays are caller-owned and read during the call only. Return codes as in ifea_spd.h.
 */
#ifndef IFEA_STEP_H
#define IFEA_STEP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifea_step* ifea_step_t;

typedef struct ifea_step_params {
  double C, n, m, eps0;      /* flow law sigma = C (eps0 + strain)^n rate^m (FlowStressLaw.cs:8-9) */
  double friction;           /* shear friction factor (SimulationConfig3D.FrictionFactor) */
  double ram;                /* ram speed, mm/s */
  double kpen;               /* absolute ANP penalty = PenaltyFactor * sigma_ref / rate_ref (Forging3DSolver.cs:149) */
  double rate_cut;           /* strain-rate cut-off (Forging3DSolver.cs:147) */
  double u0;                 /* friction regularisation velocity (Forging3DSolver.cs:150) */
  double newton_blend;       /* Forging3DSolver.NewtonBlend */
  double tolerance;          /* beta*|dv|/|v| convergence tolerance */
  double pcg_tol;            /* PCG relative residual target (1e-7 in the reference, :984; a looser value makes the
                                Newton iteration inexact, its own convergence test unchanged) */
  int max_iterations;        /* SimulationConfig3D.MaxIterations */
  int pcg_budget;            /* Forging3DSolver.PcgBudget */
  int quadratic_ls;          /* Forging3DSolver.QuadraticLineSearch */
  int adaptive_newton;       /* Forging3DSolver.AdaptiveNewton */
  int precision;             /* 0: FP64 factor. 1: FP32 supernode updates (MFMA) + FP64 POTRF/residual/PCG; an FP32
                                factor that is not positive definite switches to 2 for the rest of the run, and PCG
                                that stalls on a fresh FP32 factor falls back to FP64. 2: FP64 factor with an FP32
                                copy of L for the triangular solves (half the solve bytes), FP64 PCG refinement */
  int newton_limit;          /* Newton iterations before Picard (20, Forging3DSolver.cs:953) */
  int profile;               /* 1: time the factor's MFMA update launches and the triangular sweeps with HIP events */
} ifea_step_params;

/* Statistics of one solve_velocity call. */
typedef struct ifea_step_stats {
  int iterations, converged, newton_iterations;
  int factorizations, fp64_fallbacks, pcg_calls, pcg_iterations, solves, assemblies;
  double last_rel, last_rnorm;
  double ms_assembly, ms_factor, ms_solve, ms_total;
  int first_solve_skips;     /* first solves that refactored without trying PCG (IFEA_STEP_OPT_FIRST_SOLVE) */
} ifea_step_stats;

/* Option keys for ifea_step_set_option. */
#define IFEA_STEP_OPT_FIRST_SOLVE 1  /* 0 (default): every solve with a cached factor tries PCG first, as
                                        SolveLinear does (Forging3DSolver.cs:965-1019). 1: after the first solve of
                                        two increments in a row missed the PCG budget on the previous increment's
                                        factor, an increment's first solve refactors directly (PCG is probed again
                                        after 8 such increments); later solves keep the reference policy. */

/* Topology and pattern, once per mesh (Allocate/BuildPattern, Forging3DSolver.cs:403-521, TetMesh.cs:9-64).
 * tets[4*ne]; colPtr/rowIdx: the full symmetric CSC two-ring pattern (3x3 node blocks, sorted);
 * faces[3*nf] / face_owner[nf]: non-symmetry boundary faces in the reference's order. */
int ifea_step_create(ifea_step_t* h, int nn, int ne, const int* tets, const int* colPtr, const int* rowIdx,
                     int nf, const int* faces, const int* face_owner);
int ifea_step_set_params(ifea_step_t h, const ifea_step_params* p);

/* Solver policy knobs that are not part of the reference's parameter set (IFEA_STEP_OPT_*). */
int ifea_step_set_option(ifea_step_t h, int key, double value);

/* Optional caller ordering for the factor (see ifea_spd_analyze_perm). */
int ifea_step_set_perm(ifea_step_t h, const int* perm);

/* Per increment (PrepareStep): positions x[3nn]; nodal frames ax[9nn] (rows ax0, ax1, ax2 per node); nfix[nn]
 * (leading fixed local DOFs); fixed_val[3nn]; contact[nn] (0 free, 1 bottom die, 2 top die); vguess[3nn] the
 * global-axis start velocity (the warm-start slot, LoadVelocityGuess). Returns IFEA_ERR_ARG with message
 * "Tetrahedron e is inverted" on det <= 0 (MeshDistortionException). */
int ifea_step_prepare(ifea_step_t h, const double* x, const double* ax, const int* nfix, const double* fixed_val,
                      const int* contact, const double* vguess);

/* PrepareStep with the nodal frames on the device (Forging3DSolver.cs:669-727): once per mesh, the symmetry-plane
 * flags on_x0[nn] / on_y0[nn] and the two analytic rigid dies (kind 0: flat plane at z = p[0]; kind 1: top die with a
 * hexagonal cavity; kind 2: one half of a connecting-rod closed die; p[8] their closed-position parameters, as the
 * Python host's DieSpec); then per increment ifea_step_prepare_device(x[3nn], contact[nn], vguess[3nn]) computes the
 * constraint directions (symmetry normals, the die normal of every contact node by central differences of the die
 * height, HeightFieldDie.Project's normal), their Gram-Schmidt with the 0.2 drop threshold, the completed orthonormal
 * frames and the fixed DOFs (normal die velocity, the top die moving with -ram), and continues as ifea_step_prepare.
 * The frames equal the reference's arithmetic bit for bit (no FMA contraction). */
int ifea_step_set_geometry(ifea_step_t h, const unsigned char* on_x0, const unsigned char* on_y0, int bottom_kind,
                           const double* bottom_p, int top_kind, const double* top_p);
int ifea_step_prepare_device(ifea_step_t h, const double* x, const int* contact, const double* vguess);

/* LoadVelocityGuess (:854-869) with the current frames: v = R vg with the fixed DOFs overwritten (a start candidate). */
int ifea_step_load_guess(ifea_step_t h, const double* vguess);

/* The current frames ax[9nn], nfix[nn], die axes die_axis[nn] (only after ifea_step_prepare_device) and fixed values
 * fixed_val[3nn]; any pointer may be NULL. */
int ifea_step_get_frames(ifea_step_t h, double* ax, int* nfix, int* die_axis, double* fixed_val);

/* Element state: accumulated strain eps[ne] and flow stress sigma[ne] (used by BuildFriction). Either may be NULL. */
int ifea_step_set_element_state(ifea_step_t h, const double* eps, const double* sigma);

/* One Assemble call at the given local-frame velocity v[3nn] (or the current one if NULL): returns Pi and power;
 * r[3nn] / val[nnz] are copied out when not NULL (matrix=1 assembles the tangent). Used by parity tests. Like the
 * reference's Kernel (Forging3DSolver.cs:1351-1356) it overwrites the per-element flow stress and strain rate, which
 * the next ifea_step_prepare reads for the friction coefficients: restore them with ifea_step_set_element_state. */
int ifea_step_assemble(ifea_step_t h, int matrix, int newton, const double* v, double* energy, double* power,
                       double* r, double* val);

/* SolveVelocity: warm = the reference's _warm flag (Newton from the first iteration). */
int ifea_step_solve_velocity(ifea_step_t h, int warm, ifea_step_stats* stats);

/* FinalEvaluation: power (ram load = power / ram * symmetry factor), reaction, global nodal velocities vg[3nn],
 * residual r[3nn] (local frames), release[nn] flags. Any output pointer may be NULL. */
int ifea_step_final(ifea_step_t h, const int* die_axis, double* power, double* reaction, double* vg, double* r,
                    unsigned char* release);

/* eps += rate * dt (Forging3DSolver.cs:626) and the damage update (AccumulateDamage, :1844-1855) on the device;
 * copies out eps, sigma, rate when not NULL. */
int ifea_step_advance(ifea_step_t h, double dt, double* eps, double* sigma, double* rate);

/* Element results (FinalEvaluation :1440-1450, AccumulateDamage :1844-1855): mean[ne] the element mean stress
 * 1/4 sum kPen g_n . vg of the last ifea_step_final; damage[ne] the Cockcroft-Latham damage accumulated by
 * ifea_step_advance (D += max(sigma_1, 0) / sigma_bar * rate * dt, Principal.Max3 for sigma_1); dev6[6ne] the
 * deviatoric stress (xx, yy, zz, xy, yz, zx) of the last assembly. Any pointer may be NULL. */
int ifea_step_element_results(ifea_step_t h, double* mean, double* damage, double* dev6);

/* Damage state damage[ne] (a restart), or zero when NULL. */
int ifea_step_set_damage(ifea_step_t h, const double* damage);

/* Current local-frame velocity v[3nn] (out) / set it (in). */
int ifea_step_get_v(ifea_step_t h, double* v);
int ifea_step_set_v(ifea_step_t h, const double* v);

/* The collapse check of TryUpdate (Forging3DSolver.cs:1509-1516): TetMesh.Quality (TetMesh.cs:107-112) of every tet at
 * the node positions x[3nn] (host), first_bad = the smallest index of a tet with quality <= qmin, or -1; quality[ne]
 * receives every tet's value when not NULL. */
int ifea_step_quality(ifea_step_t h, const double* x, double qmin, int* first_bad, double* quality);

/* ifea_step_quality without a handle and without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). tets[4ne] and x[3nn] are host arrays, staged with the first-index counter and the per-tet values
 * in one static 64 MiB device array per GPU (held from library load until the process ends): a mesh needing more
 * (16 ne + 24 nn + 8 ne bytes with quality, rounded up to 256-byte blocks, plus 256) returns IFEA_ERR_ARG. Calls
 * that overlap on one GPU share that array and race. Results as ifea_step_quality. */
int ifea_step_quality_noalloc(int nn, int ne, const int* tets, const double* x, double qmin, int* first_bad,
                              double* quality);

/* Device and factor statistics (24 doubles): [0] nnz(L), [1] factor flops, [2] supernodes, [3] device bytes,
 * [4] last factor ms, [5] L words, [6] update-pool words, [7] levels; with params.profile set, accumulated since the
 * last call with reset: [8] inner (64-column) MFMA update launches, [9] their ms, [10] their algorithmic flops,
 * [11] triangular sweeps (fwd+bwd), [12] their ms, [13] panel bytes they read, [14] factorizations,
 * [15] factorization ms, [16] outer (256-column) MFMA update launches, [17] their ms, [18] their algorithmic flops,
 * [19] factorizations done in FP32, [20] sweeps done with an FP32 factor, [21..23] the FP32 part of [16..18]. */
int ifea_step_info(ifea_step_t h, int reset, double* out24);

void ifea_step_destroy(ifea_step_t h);

#ifdef __cplusplus
}
#endif
#endif
