/*
 * fooska.h -- C ABI of the MI355X finite-volume step (libfooska_hip.so).
 *
 * One context = one device = one HIP stream; calls on a context must be
 * externally synchronised.  The caller owns host buffers, the library owns
 * device buffers.  Every function returns FK_OK (0) or a negative FK_E* code
 * (fooska_mesh.h); fk_last_error() returns the message of the last error of
 * the calling thread.
 *
 * The entry points replace the reference's host "phase" functions, which are
 * declared in mulitphase Solver/doublePrecision/strdata.h:206-247 and called
 * in this order by explicitLoop (explicitLoop.cu:28-64):
 *   per RK stage: computePrimVec -> computeGrad -> computeFaceValues ->
 *                 computeRHS -> [computeTimeStep at stage 0] -> update
 *   per step:     relaxation, consCalc
 * The first three are fused into fk_compute_face_values (pass G) and the
 * last two into fk_compute_rhs_update (pass F).  Face-to-cell accumulation
 * is a deterministic per-cell gather over the cell's faces in ascending face
 * order (the reference uses floating-point atomics), so repeated runs are
 * bitwise identical.
 */
#ifndef FOOSKA_H
#define FOOSKA_H
#include <stdint.h>

#include "fooska_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FK_MODEL_SINGLE 0    /* 5 equations: rho, rho u, rho v, rho w, rho E (SP/) */
#define FK_MODEL_TWO_PHASE 1 /* 10 equations of mulitphase Solver (DP/, MP/) */

#define FK_FP64 0  /* every operation in double precision */
#define FK_MIXED 1 /* FP32 reconstruction, HLLC, WALE and geometry;
                      FP64 state, RK accumulation, EOS inversion, Newton */

#define FK_RELAX_PLAIN 0 /* reference Newton solves (relaxation.cu, thermRlx.cu, chemRlx.cu) */
#define FK_RELAX_MLP 1   /* MLP warm start + FP64 Newton polish, plain Newton as fallback */
#define FK_RELAX_COLD 2  /* FK_RELAX_MLP with the cell state in place of the MLP
                            start, (T_L, max(p_L, 1 kPa)) for a chemical cell
                            (ablation: what the network adds) */

/* Solver and material constants; field names follow the reference INPUT
 * struct (strdata.h:49-98).  Two-phase: index 0 = liquid, 1 = gas.
 * Single-phase uses index 0 only. */
typedef struct fk_params {
  int32_t model;      /* FK_MODEL_* */
  int32_t precision;  /* FK_FP64 / FK_MIXED */
  int32_t rkSteps;    /* 1 (forward Euler) or 3 (Shu-Osher SSP-RK3) */
  int32_t adaptDt;    /* 1: dt from the CFL condition every adaptDtItr steps */
  int32_t adaptDtItr;
  int32_t viscFlag;   /* Newtonian viscous stresses */
  int32_t turbulence; /* WALE eddy viscosity (Cw = 0.325) */
  int32_t heatFlux;   /* Fourier heat flux k = Cp mu_eff / Pr (single-phase; no reference counterpart) */
  int32_t woodsFlag;  /* Wood mixture sound speed where rho Y_v > chemLimit */
  int32_t neglectX, neglectY, neglectZ;
  int32_t chemFlag;   /* classify cells and run the chemical (p-T-g) relaxation */
  int32_t thermItrMax, chemItrMax;
  int32_t hllcEnergy; /* 0: energy star state as in starRegion.h:32, 1: Toro's form */
  int32_t relaxFlag;  /* two-phase: run the per-step phase relaxation */
  int32_t relaxMode;  /* FK_RELAX_PLAIN, FK_RELAX_MLP or FK_RELAX_COLD for the relaxation of fk_run_steps */
  double dt;          /* fixed time step (adaptDt == 0) */
  double cfl, maxDt, avfMin;
  double thermUrf, chemUrf, chemLimit, satTempUrf;
  double polishTol;   /* relative tolerance of the FP64 Newton polish */
  double gamma[2], pInf[2], Cp[2], eta[2], satConst[2], dynVisc[2], Pr[2];
  /* Finite-rate thermal relaxation (INPUT::thermFntRlxFlag / thermRlxRate,
   * strdata.h:72-73; fntPressRlx, thermRlx.cu:42-67): 1 makes the thermal
   * Newton solve target p_L - p_G = dP, T_L - T_G = dT with
   * dP = (p_L - p_G) exp(-|kappa| dt) (dt = the `dt` field above) instead of
   * equilibrium.  The reference's defaults are 0 and 0.5 (buildInput.cu:35-36).
   * Only FK_RELAX_PLAIN supports it. */
  int32_t thermFntRlxFlag;
  double thermRlxRate;
  /* Inlet (boundCond 3) ghost pressure and temperature (bc.h:5-7).  Both 0:
   * the reference's constants, 11.4e5 Pa / 115.3 K (two-phase) and
   * 4.2 * 101300 Pa / 288 K (single-phase). */
  double inletP, inletT;
  /* Reference pressure of the single-phase mixed-precision path (no
   * reference counterpart): its FP32 face pressures hold p - pRef, so that a
   * flow with small pressure differences on a large background (Mach 0.1:
   * p ~ 71, differences ~ 0.06) keeps FP32 relative precision in them; the
   * absolute pressure is formed for sound speeds and the energy flux.  Set it
   * near the mean pressure; it is rounded to the nearest FP32 value.  0 =
   * none.  Ignored by the FP64 path and the two-phase model. */
  double pRef;
} fk_params;

typedef struct fk_ctx fk_ctx;

typedef struct fk_relax_stats {
  int64_t n_therm, n_chem;      /* cells classified thermal / chemical */
  int64_t newton_iters;         /* Newton iterations summed over cells (therm + chem);
                                   a failed warm-started polish counts as 30 */
  int64_t tsat_iters;           /* saturation-temperature iterations summed */
  int64_t n_fallback;           /* warm-started cells that fell back to the plain solve */
  int64_t n_warm;               /* chemical cells started from the MLP prediction (FK_RELAX_MLP) */
} fk_relax_stats;

const char *fk_last_error(void);

/* Context on HIP device `device`. */
int fk_create(int device, const fk_params *p, fk_ctx **out);
int fk_destroy(fk_ctx *ctx);
/* Replace the constants (model and NEQ must stay the same once a mesh is up). */
int fk_set_params(fk_ctx *ctx, const fk_params *p);

/* Mesh in the reference MESH semantics (strdata.h:100-115, one rank):
 * fc2el[nFcs][2] (owner < neighbour, -1 on boundary faces), boundCond[nFcs],
 * n[nFcs][3] unit normal out of the owner, area[nFcs], fcCenter[nFcs][3],
 * cellCenter[nElem][3], volume[nElem]; shift[nFcs][3] (periodic translation
 * of the face as seen from the neighbour) may be NULL.  Cells have four to
 * six faces (tetrahedra, pyramids, prisms, hexahedra, translator.cu:122-293);
 * every per-cell sum runs over the cell's faces in ascending face id.  A
 * mesh of tetrahedra only keeps 4 face slots per cell, any other mesh 6. */
int fk_upload_mesh(fk_ctx *ctx, int64_t nElem, int64_t nFcs, const int32_t *fc2el, const int32_t *boundCond,
                   const double *n, const double *area, const double *fcCenter, const double *cellCenter,
                   const double *volume, const double *shift);

/* Conserved state, reference AoS layout consVec[nElem][NEQ]. */
int fk_upload_state(fk_ctx *ctx, const double *consVec);
int fk_download_state(fk_ctx *ctx, double *consVec);

/* ---- phase functions (reference order) ---- */
/* initRk (update.cu:141): keep the stage-0 state for the SSP-RK3 stages. */
int fk_init_rk(fk_ctx *ctx);
/* computePrimVec + computeGrad + computeFaceValues (computePrimVec.cu:82,
 * computeGrad.cu:156, computeFaceValues.cu:138): first-order face states,
 * Green-Gauss gradients, neighbour min/max, limiter and MUSCL face states. */
int fk_compute_face_values(fk_ctx *ctx);
/* computeRHS (computeRHS.cu:214): HLLC + non-conservative terms + pressure
 * correction (+ viscous/WALE) gathered per cell; kept on the device for
 * fk_update (FP64; FP32 in mixed precision, as the MP reference stores its
 * RHS, promoted to FP64 in fk_update).  Also records the per-cell signal
 * speed for fk_compute_time_step. */
int fk_compute_rhs(fk_ctx *ctx);
/* computeTimeStep (computeTimeStep.cu:172): dt = min(cfl * min_c V^0.3333 /
 * max signal, maxDt) from the face states of the last fk_compute_rhs. */
int fk_compute_time_step(fk_ctx *ctx, double *dt);
/* update (update.cu:105): RK stage irk with the residual of fk_compute_rhs. */
int fk_update(fk_ctx *ctx, int irk, double dt);
/* computeRHS + update fused: same result as fk_compute_rhs + fk_update (in
 * mixed precision up to the FP32 rounding of the stored residual, which the
 * fused form keeps in FP64 registers). */
int fk_compute_rhs_update(fk_ctx *ctx, int irk, double dt);
/* relaxation (relaxation.cu:77): per-cell phase relaxation, mode FK_RELAX_*. */
int fk_relaxation(fk_ctx *ctx, int mode, fk_relax_stats *stats);
/* Counts summed over every relaxation run without a stats pointer (those of
 * fk_run_steps) since the context was created or last reset; reset != 0
 * clears them after reading (stats may then be NULL).  No reference
 * counterpart (diagnostics of the relaxation cost). */
int fk_relax_totals(fk_ctx *ctx, fk_relax_stats *stats, int reset);
/* consCalc (consCalc.cu:36): out[i] = sum_c U_i V_c (i = 0 integrates 1 - alpha). */
int fk_cons_calc(fk_ctx *ctx, double *out);
/* fk_cons_calc without any allocation: no hipMalloc, hipFree or host heap
 * use in the call.  No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  Its
 * partial sums live in a fixed device array of 4112 doubles (32.9 KB) that
 * each device holds until the process ends, set up by fk_create; results are
 * bitwise those of fk_cons_calc.  Every context on one device shares that
 * array: concurrent calls on one device race and must be serialised by the
 * caller. */
int fk_cons_calc_noalloc(fk_ctx *ctx, double *out);

/* Whole time steps in explicitLoop order (dt is taken before the stage-0
 * update of every adaptDtItr-th step), with the relaxation of relaxMode
 * after the stages.  Returns the last dt and the time.  FK_RELAX_MLP
 * without weights fails with FK_ESTATE before any stage runs.  Every pass G
 * checks the state it reads (phase densities, or the density, positive and
 * every primitive finite); if a check failed during the call, fk_run_steps
 * still completes the steps and advances the clock but returns FK_ERANGE,
 * naming the first such step and its lowest cell (fk_health).  The reference has no
 * such check (it runs on, explicitLoop.cu:66-69 reports only launch errors). */
int fk_run_steps(fk_ctx *ctx, int nsteps, double *dt_last, double *phys_time);
/* Health record of pass G since the last reset (fk_run_steps resets it when
 * it starts): the first step with a non-physical state and the lowest such
 * cell in that step, or -1 for both when every check passed; reset != 0
 * clears it. */
int fk_health(fk_ctx *ctx, int64_t *step, int64_t *cell, int reset);

/* Block until the context's stream is idle. */
int fk_sync(fk_ctx *ctx);
/* Time-step state of the loop: last dt, physical time and step counter (the
 * step counter decides when fk_run_steps recomputes dt, every adaptDtItr
 * steps).  With fk_download_state / fk_upload_state they make a bit-exact
 * restart point; the reference's restart (restart.cu:113-152, 424-497)
 * writes physicalTime and consVec as %.10e text. */
int fk_get_clock(fk_ctx *ctx, double *dt, double *time, int64_t *step);
int fk_set_clock(fk_ctx *ctx, double dt, double time, int64_t step);

/* Download an intermediate array for phase-level parity checks.  Names and
 * layouts (FP64 path; the mixed path stores FP32 where marked f):
 *  "grad"   [nElem][NEQ][3]   Green-Gauss gradients (needs fk_set_debug(1))
 *  "theta"  [nElem][NEQ]      limiter (needs fk_set_debug(1))
 *  "face"   [nElem][NS][NEQ]  reconstructed face states per cell face slot (f);
 *                             NS = 4 on tetrahedral meshes, else 6 (slots past
 *                             a cell's last face hold no face state)
 *  "rhs"    [nElem][NEQ]      residual of the last fk_compute_rhs (f); mixed
 *                             precision stores it as FP32 (widened here)
 *  "slots"  [nElem][NS] int32 face id of each cell face slot, -1 past the last
 *  "signal" [nElem]           max signal speed per cell                      */
int fk_download(fk_ctx *ctx, const char *name, void *host, int64_t bytes);
int fk_set_debug(fk_ctx *ctx, int on);
/* Implementation choices that do not change results (no reference
 * counterpart; all give bitwise equal states).
 *  "fuse_update" (default 1): the RK stage update runs inside pass F at every
 *     stage whose dt is known before pass F (not at a stage that takes a new
 *     dt: there pass F stores the residual and the signal speeds, dt is
 *     reduced, and a streaming kernel applies the update); at 48M cells,
 *     mixed: 7.5 ms against 6.7 + 2.5 ms.  0 = always residual + update
 *     kernel.
 *  "stage_f" (default 1): pass F copies the face states and velocity
 *     gradients of its workgroup's cells to LDS and reads in-workgroup mirror
 *     states from there (4 slots per cell: 256 cells, FP64 128; 6 slots:
 *     two-phase FP32, 128 cells, not in the last stage's relaxation pass).
 *  "stage_g" (default 1): pass G writes its workgroup's face records and
 *     velocity gradients from LDS in 16-byte words (4 slots per cell: 256
 *     cells, FP64 128; 6 slots: two-phase FP32, 128 cells).
 *  "prims_w" (default 1): mixed precision only; the kernels that produce a
 *     new state (fused pass F, relaxation) also write its FP32 primitive
 *     variables, which the next pass G gathers instead of converting the FP64
 *     conserved state of the cell and its neighbours.
 *  "fuse_relax" (default 1): in fk_run_steps the last RK stage's fused pass
 *     F also runs the per-cell relaxation pass on the state it forms, instead
 *     of a separate kernel reading it back: stage 1 of the surrogate
 *     relaxation (FK_RELAX_MLP, FK_RELAX_COLD) or the plain Tsat
 *     classification and thermal solve (FK_RELAX_PLAIN, except while it emits
 *     training pairs); the chemical / listed cells follow in their own kernel.
 * Unknown names return FK_EINVAL. */
int fk_set_tuning(fk_ctx *ctx, const char *name, int64_t value);

/* ---- surrogate of the relaxation Newton solve (no counterpart in the
 * reference).  A 3-64-64-64-3 tanh MLP maps the cell features
 * (log rho, e, Y_L) -- mixture density, internal energy per unit mass, liquid
 * mass fraction -- to the equilibrium (T, asinh(p/1e4), alpha_L).
 * FK_RELAX_MLP classifies every cell exactly (sign of the saturation
 * function) and solves thermal cells by analytic-Jacobian Newton from the
 * cell state (at most 4 iterations in a first pass, up to 30 in a second);
 * chemical cells are solved by an FP64 Newton polish in (ln p, T) from the
 * MLP's (T, p) (the alpha_L output is not used: the polish takes alpha_L
 * from the mass balance).  Where a polish fails the plain reference solve
 * runs from the cell state.
 * Training targets are chemical (p-T-g) equilibria of the states' mixture
 * density and internal energy, solved to 1e-10 by the same (ln p, T) Newton
 * from the cell state; NaN where no two-phase equilibrium exists.  The
 * emission labels every emitted solver state; fk_mlp_label labels the
 * chemical-class states it is given (NaN for the others).
 * Parameters are
 * FK_MLP_PARAMS floats in the order W1[64][3] b1 W2[64][64] b2 W3[64][64] b3
 * W4[3][64] b4; norm[12] = feature mean[3], std[3], target mean[3], std[3]. */
#define FK_MLP_PARAMS 8771
int fk_mlp_set_weights(fk_ctx *ctx, const float *w, const float *norm);
int fk_mlp_get_weights(fk_ctx *ctx, float *w, float *norm);
int fk_mlp_set_norm(fk_ctx *ctx, const float *norm);
/* Features and equilibrium targets of host states U[n][10] (targets NaN
 * for thermal-class states and where no equilibrium exists; the solver
 * state is untouched). */
int fk_mlp_label(fk_ctx *ctx, const double *U, int64_t n, float *x, float *y);
/* Training set on the device: replace (append = 0) or extend it. */
int fk_mlp_data(fk_ctx *ctx, const float *x, const float *y, int64_t n, int append);
/* From now on every plain relaxation also adds the (features, targets) of
 * cells phase + k * stride (phase advancing per call) to the training set,
 * up to `capacity` more samples; stride 0 stops emitting. */
int fk_mlp_emit(fk_ctx *ctx, int64_t stride, int64_t capacity);
int fk_mlp_download_data(fk_ctx *ctx, float *x, float *y, int64_t max, int64_t *n);
/* Gradient of the mean squared error (normalised targets) over the samples
 * idx[B] of the training set, without an update (gradient checks). */
int fk_mlp_grad(fk_ctx *ctx, const int64_t *idx, int64_t B, float *grad, float *loss);
/* `steps` Adam steps on random batches (splitmix64(seed)) of the training set. */
int fk_mlp_train(fk_ctx *ctx, int steps, int batch, float lr, uint64_t seed, float *loss_first, float *loss_last);
/* Forward pass on raw features x[n][3] -> raw targets y[n][3]. */
int fk_mlp_infer(fk_ctx *ctx, const float *x, int64_t n, float *y);

/* Plain device-to-device copy of `bytes` (rounded down to 16 B) between two
 * scratch buffers by a streaming kernel ("stream_copy"); returns its duration.
 * Used to calibrate HBM byte counters and as the achievable-bandwidth
 * reference.  No counterpart in the reference. */
int fk_stream_copy(fk_ctx *ctx, int64_t bytes, double *ms);

/* Per-kernel timing with HIP events on the context's stream.  When on, every
 * launch is bracketed by events; fk_kernel_times reports, for each named
 * kernel, the launch count and the summed milliseconds. */
int fk_set_timing(fk_ctx *ctx, int on);
int fk_kernel_times(fk_ctx *ctx, int max, char (*names)[32], int64_t *counts, double *ms);

#ifdef __cplusplus
}
#endif
#endif
