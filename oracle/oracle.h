/*
 * oracle.h -- CPU FP64 restatement of the reference finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or as the timed CPU baseline; the product never calls it.
 *
 * The reference (sDamianos/FooskaGPU) has no CPU path, so this is a serial
 * C++ re-expression of its CUDA kernels' arithmetic, operation order
 * included, with two deliberate restatements shared with the GPU path:
 *   - face-to-cell accumulation is a per-cell sum over the cell's faces in
 *     ascending face id order (the reference uses atomicAdd, whose order is
 *     unspecified);
 *   - rotations to the face frame use only the unit normal:
 *     star velocity u + (S* - u.n) n, slip/symmetry ghost u - 2 (u.n) n,
 *     wall ghost -u (reference: stored tangents t1, t2; equal up to rounding);
 *   - pow(x,1.5|1.25|2.5) in WALE are x*sqrt(x), x*sqrt(sqrt(x)),
 *     x*x*sqrt(x) (equal up to rounding).
 * Built twice: serial (liboracle.so) and with OpenMP (liboracle_omp.so).
 */
#ifndef FOOSKA_ORACLE_H
#define FOOSKA_ORACLE_H
#include <stdint.h>

#include "../include/fooska.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

int orc_create(const fk_params *p, int64_t nElem, int64_t nFcs, const int32_t *fc2el, const int32_t *boundCond,
               const double *n, const double *area, const double *fcCenter, const double *cellCenter,
               const double *volume, const double *shift, orc **out);
void orc_free(orc *o);
int orc_set_params(orc *o, const fk_params *p);
int orc_set_state(orc *o, const double *U);
int orc_get_state(const orc *o, double *U);

/* phase functions, same order and meaning as the reference */
int orc_init_rk(orc *o);
int orc_compute_face_values(orc *o);
int orc_compute_rhs(orc *o);
int orc_compute_time_step(orc *o, double *dt);
int orc_update(orc *o, int irk, double dt);
int orc_relaxation(orc *o, int64_t stats[4]); /* n_therm, n_chem, newton iters, tsat iters */
int orc_cons_calc(const orc *o, double *out);
int orc_run_steps(orc *o, int nsteps, double *dt_last, double *phys_time);

/* Intermediates: "grad" [C][NEQ][3], "theta" [C][NEQ], "wmin"/"wmax" [C][NEQ],
 * "primF" [F][2][NEQ], "face" [C][NS][NEQ] (reconstructed state of each cell
 * face slot, faces in ascending id, 0 in slots past the cell's faces),
 * "rhs" [C][NEQ], "signal" [C], "slots" [C][NS] int32 (-1 past the cell's
 * faces), "nslots" int32 NS: 4 when every cell is a tetrahedron, else 6. */
int orc_get(const orc *o, const char *name, void *out, int64_t bytes);
/* Test hook: overwrite the stored residual used by orc_update. */
int orc_set_rhs(orc *o, const double *rhs);

/* ---- pointwise routines, exposed for unit tests of the oracle itself ---- */
double orc_eqos(const fk_params *p, int imat, int id, double v1, double v2);
void orc_cons2prim(const fk_params *p, const double *U, double *W);
void orc_prim2cons(const fk_params *p, const double *W, double *U);
/* Convective contribution of one face to its owner (side 0) and neighbour
 * (side 1) cells: WL/WR face states, n unit normal, A area, UL/UR the cells'
 * conserved states (two-phase non-conservative terms and pressure
 * correction), bc the boundary code (0 interior: WR used; else WR ignored and
 * the ghost of WL is used).  out[2][NEQ]. */
void orc_face_flux(const fk_params *p, const double *WL, const double *WR, const double *n, double A, const double *UL,
                   const double *UR, int bc, double *out);
/* Ghost state of the boundary code bc for face state W and unit normal n. */
void orc_bc_ghost(const fk_params *p, int bc, const double *n, double *W);
/* WALE eddy viscosity of the velocity gradient g[3][3] (g[i][j] = d u_i / d x_j)
 * and filter volume vol. */
double orc_wale(const double *g, double vol);
/* Saturation temperature (relaxation.cu:4-38) and iteration count. */
double orc_tsat(const fk_params *p, double P, double T0, double urf, int *iters);
/* Shared portable log10 of the relaxation and the Gibbs energy g_k(P, T). */
double orc_log10(double x);
double orc_exp(double x);
/* Finite-rate targets (fntPressRlx, thermRlx.cu:42-67) of phase volume
 * fractions avf[2], densities rho[2] and internal energies ener[2]. */
void orc_fnt_press_rlx(const fk_params *p, const double *avf, const double *rho, const double *ener, float *dP,
                       float *dT);
double orc_gibbs(const fk_params *p, int k, double P, double T);
/* Relax one cell's conserved state in place; returns 0 thermal, 1 chemical;
 * iters = Newton iterations. */
int orc_relax_cell(const fk_params *p, double *U, int *iters, int *tsat_iters);
/* The chemical relaxation (chemRlx) of one cell with its Newton iterations
 * recorded: rows[max_rows][20] as kChemTraceCols in oracle.cpp, *n_rows the
 * iterations recorded, *aborted 1 when the solve ended on the abort path.
 * Returns the iteration count.  For the oracle's own tests. */
int orc_chem_trace(const fk_params *p, double *U, double *rows, int max_rows, int *n_rows, int *aborted);
/* Threads a parallel region of this build uses: omp_get_max_threads() in the
 * OpenMP build, 1 in the serial one. */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
