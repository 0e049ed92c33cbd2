// CPU FP64 restatement of the reference finite-volume step.
// TEST INFRASTRUCTURE ONLY (see oracle.h): the checker of the GPU path and the
// CPU baseline, never part of the product.
//
// Structure mirrors the reference kernels: face loops compute per-face
// quantities exactly as the CUDA kernels do, cell loops then accumulate them
// in ascending face id order (the reference scatters with atomicAdd).
// File:line citations are relative to "mulitphase Solver/doublePrecision/"
// (DP) or "singlePhase/src/" (SP) of the reference.

#include "oracle.h"
#include "../fooskagpu_amd/csrc/fk_math.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <new>
#include <string>
#include <vector>

namespace {

constexpr int MAXEQ = 10;

inline double clampAvf(const fk_params &p, double a) {
  return std::fmin(std::fmax(p.avfMin / 100, a), 1.0 - p.avfMin / 100);
}

// ---- EOS: DP eqos.h:1-23 (two-phase), SP eqos.h:2-23 (single-phase) ----
double eqos2(const fk_params &p, int k, int id, double v1, double v2) {
  const double g = p.gamma[k], pi = p.pInf[k], cp = p.Cp[k], eta = p.eta[k];
  double v3 = 0;
  switch (id) {
    case 0: v3 = (g - 1) * v1 * v2 - g * pi - (g - 1) * eta * v1; break;
    case 1:
      v3 = (g - 1) * v1 * v2 - g * pi - (g - 1) * eta * v1;
      v3 = (v3 + pi) / ((cp / g) * (g - 1) * v1);
      break;
    case 2: v3 = (v1 + pi) / ((cp / g) * v2 * (g - 1)); break;
    case 3: v3 = (v1 + g * pi + (g - 1) * eta * v2) / (v2 * (g - 1)); break;
    case 4: v3 = std::sqrt((g * (v2 + pi)) / v1); break;
  }
  return v3;
}
double eqos1(const fk_params &p, int id, double v1, double v2) {
  const double g = p.gamma[0], pi = p.pInf[0], cp = p.Cp[0];
  double v3 = 0;
  switch (id) {
    case 0: v3 = (g - 1) * v1 * v2 - g * pi; break;
    case 1: v3 = (g / cp) * (v2 - pi / v1); break;
    case 2: v3 = (v1 + pi) / ((cp / g) * v2 * (g - 1)); break;
    case 3: v3 = (v1 + g * pi) / (v2 * (g - 1)); break;
    case 4: v3 = std::sqrt((g * (v2 + pi)) / v1); break;
  }
  return v3;
}

// ---- conversions: DP cons2prim.h:2-24, prim2cons.h:2-21; SP cons2prim.h, prim2cons.h ----
void cons2prim(const fk_params &p, const double *U, double *s) {
  if (p.model == FK_MODEL_SINGLE) {
    s[0] = U[0];
    s[1] = U[1] / s[0];
    s[2] = U[2] / s[0];
    s[3] = U[3] / s[0];
    s[4] = U[4] / s[0] - 0.5 * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3]);
    s[4] = eqos1(p, 0, s[0], s[4]);
    return;
  }
  s[0] = clampAvf(p, U[0]);
  s[1] = U[1] / s[0];
  s[2] = U[2] / (1 - s[0]);
  double rhoMix = s[0] * s[1] + (1 - s[0]) * s[2];
  s[3] = U[3] / rhoMix;
  s[4] = U[4] / rhoMix;
  s[5] = U[5] / rhoMix;
  double eL = U[6] / (s[0] * s[1]);
  double eG = U[7] / ((1 - s[0]) * s[2]);
  s[6] = eqos2(p, 0, 0, s[1], eL);
  s[7] = eqos2(p, 1, 0, s[2], eG);
  s[8] = U[9];
  s[9] = rhoMix;
}

void prim2cons(const fk_params &p, const double *w, double *c) {
  if (p.model == FK_MODEL_SINGLE) {
    c[0] = w[0];
    c[1] = w[0] * w[1];
    c[2] = w[0] * w[2];
    c[3] = w[0] * w[3];
    c[4] = w[0] * (eqos1(p, 3, w[4], w[0]) + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
    return;
  }
  c[0] = w[0];
  c[1] = w[0] * w[1];
  c[2] = (1 - w[0]) * w[2];
  double rho = c[1] + c[2];
  c[3] = rho * w[3];
  c[4] = rho * w[4];
  c[5] = rho * w[5];
  c[6] = c[1] * eqos2(p, 0, 3, w[6], w[1]);
  c[7] = c[2] * eqos2(p, 1, 3, w[7], w[2]);
  c[8] = c[6] + c[7] + 0.5 * rho * (w[3] * w[3] + w[4] * w[4] + w[5] * w[5]);
  c[9] = w[8];
}

// ---- boundary ghosts: DP bc.h:2-84, SP bc.h:2-78 (normal-only rotation) ----
void bc_ghost(const fk_params &p, int bc, const double *n, double *s) {
  const int iu = p.model == FK_MODEL_SINGLE ? 1 : 3;
  if (bc == 1 || bc == 4) {  // symmetry / slip: reflect the normal velocity
    double un = s[iu] * n[0] + s[iu + 1] * n[1] + s[iu + 2] * n[2];
    for (int i = 0; i < 3; ++i) s[iu + i] = s[iu + i] - 2.0 * un * n[i];
  } else if (bc == 5) {  // no-slip wall
    for (int i = 0; i < 3; ++i) s[iu + i] = -s[iu + i];
  } else if (bc == 3) {  // inlet (Dirichlet p, T; velocity kept)
    // fk_params inletP / inletT; both 0 selects the constants of bc.h:5-7
    const bool ref = p.inletP == 0 && p.inletT == 0;
    if (p.model == FK_MODEL_SINGLE) {
      double P = ref ? 4.2 * 101300 : p.inletP, T = ref ? 288 : p.inletT;
      s[0] = eqos1(p, 2, P, T);
      s[4] = P;
    } else {
      double P = ref ? 11.4e5 : p.inletP, T = ref ? 115.3 : p.inletT;
      s[0] = 1.0 - p.avfMin;
      s[1] = eqos2(p, 0, 2, P, T);
      s[2] = eqos2(p, 1, 2, P, T);
      s[6] = P;
      s[7] = P;
      s[8] = 0;
      s[9] = 0;
    }
  }
  // bc == 2 outlet: zero gradient, state unchanged
}

// ---- wave speeds: DP waveSpeeds.h:2-46, SP waveSpeeds.h:1-18 ----
void wave_speeds(const fk_params &p, const double *n, const double *L, const double *R, double c[3], double &uL,
                 double &uR) {
  if (p.model == FK_MODEL_SINGLE) {
    uL = L[1] * n[0] + L[2] * n[1] + L[3] * n[2];
    uR = R[1] * n[0] + R[2] * n[1] + R[3] * n[2];
    double cL = eqos1(p, 4, L[0], L[4]);
    double cR = eqos1(p, 4, R[0], R[4]);
    c[0] = std::fmin(uL - cL, uR - cR);
    c[2] = std::fmax(uL + cL, uR + cR);
    c[1] = (R[4] - L[4] + L[0] * uL * (c[0] - uL) - R[0] * uR * (c[2] - uR)) / (L[0] * (c[0] - uL) - R[0] * (c[2] - uR));
    return;
  }
  uL = L[3] * n[0] + L[4] * n[1] + L[5] * n[2];
  uR = R[3] * n[0] + R[4] * n[1] + R[5] * n[2];
  double cL, cR;
  if ((L[8] + R[8]) / 2 > p.chemLimit && p.woodsFlag == 1) {
    double rhoMix = L[0] * L[1] + (1 - L[0]) * L[2];
    double cLiq = eqos2(p, 0, 4, L[1], L[6]);
    double cGas = eqos2(p, 1, 4, L[2], L[7]);
    cL = 1 / std::sqrt(rhoMix * (L[0] / (L[1] * cLiq * cLiq) + (1 - L[0]) / (L[2] * cGas * cGas)));
    rhoMix = R[0] * R[1] + (1 - R[0]) * R[2];
    cLiq = eqos2(p, 0, 4, R[1], R[6]);
    cGas = eqos2(p, 1, 4, R[2], R[7]);
    cR = 1 / std::sqrt(rhoMix * (R[0] / (R[1] * cLiq * cLiq) + (1 - R[0]) / (R[2] * cGas * cGas)));
  } else {
    double YL = L[0] * L[1] / (L[0] * L[1] + (1 - L[0]) * L[2]);
    double YG = (1 - L[0]) * L[2] / (L[0] * L[1] + (1 - L[0]) * L[2]);
    double a = eqos2(p, 0, 4, L[1], L[6]), b = eqos2(p, 1, 4, L[2], L[7]);
    cL = std::sqrt(YL * (a * a) + YG * (b * b));
    YL = R[0] * R[1] / (R[0] * R[1] + (1 - R[0]) * R[2]);
    YG = (1 - R[0]) * R[2] / (R[0] * R[1] + (1 - R[0]) * R[2]);
    a = eqos2(p, 0, 4, R[1], R[6]);
    b = eqos2(p, 1, 4, R[2], R[7]);
    cR = std::sqrt(YL * (a * a) + YG * (b * b));
  }
  c[0] = std::fmin(uL - cL, uR - cR);
  c[2] = std::fmax(uL + cL, uR + cR);
  double pressL = L[0] * L[6] + (1 - L[0]) * L[7];
  double pressR = R[0] * R[6] + (1 - R[0]) * R[7];
  double rhoL = L[0] * L[1] + (1 - L[0]) * L[2];
  double rhoR = R[0] * R[1] + (1 - R[0]) * R[2];
  c[1] = (pressR - pressL + rhoL * uL * (c[0] - uL) - rhoR * uR * (c[2] - uR)) / (rhoL * (c[0] - uL) - rhoR * (c[2] - uR));
}

// Per-face quantities of computeFlux (DP computeRHS.cu:78-146 / SP computeRHS.cu:23-98)
struct FaceFlux {
  double adv[MAXEQ];   // two-phase: fluxAdv; single-phase: x+y+z sum of fluxAdv
  double hllc[MAXEQ];
  double unc[3];       // velocity of the non-conservative terms
  double pF;           // face pressure of the upwind state
};

void face_flux(const fk_params &p, const double *WL, const double *WRin, const double *n, int bc, FaceFlux &out) {
  const int neq = p.model == FK_MODEL_SINGLE ? 5 : 10;
  double L[MAXEQ], R[MAXEQ];
  for (int i = 0; i < neq; ++i) {
    L[i] = WL[i];
    R[i] = bc == 0 ? WRin[i] : WL[i];
  }
  if (bc != 0) bc_ghost(p, bc, n, R);
  double c[3], uL, uR;
  wave_speeds(p, n, L, R, c, uL, uR);
  for (int i = 0; i < MAXEQ; ++i) out.hllc[i] = 0.0;
  if (p.model == FK_MODEL_SINGLE) {
    int is_bc5 = (bc == 5);
    c[1] = (1 - is_bc5) * c[1];
    bool isLeft = c[1] >= 0;
    bool isStar = c[0] < 0 && c[2] >= 0;
    const double *s = isLeft ? L : R;
    double un = isLeft ? uL : uR;
    double cK = isLeft ? c[0] : c[2];
    if (isStar) {
      double consF[5], vecS[5], vg[3];
      prim2cons(p, s, consF);
      for (int i = 0; i < 3; ++i) vg[i] = s[1 + i] + (c[1] - un) * n[i];
      vecS[0] = s[0] * (cK - un) / (cK - c[1]);
      vecS[1] = vecS[0] * vg[0];
      vecS[2] = vecS[0] * vg[1];
      vecS[3] = vecS[0] * vg[2];
      vecS[4] = vecS[0] * (eqos1(p, 3, s[4], s[0]) + 0.5 * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3]) +
                           (c[1] - un) * (c[1] + s[4] / (s[0] * (cK - un))));
      for (int iv = 0; iv < 5; ++iv) out.hllc[iv] = cK * (vecS[iv] - consF[iv]);
    }
    // SP flux.h:2-23, summed x + y + z as in computeRHS.cu:90-91
    double fa[15];
    fa[0] = s[1] * s[0] * n[0];
    fa[1] = s[2] * s[0] * n[1];
    fa[2] = s[3] * s[0] * n[2];
    fa[3] = (s[0] * s[1] * s[1] + s[4]) * n[0];
    fa[4] = s[0] * s[1] * s[2] * n[1];
    fa[5] = s[0] * s[1] * s[3] * n[2];
    fa[6] = s[0] * s[2] * s[1] * n[0];
    fa[7] = (s[0] * s[2] * s[2] + s[4]) * n[1];
    fa[8] = s[0] * s[2] * s[3] * n[2];
    fa[9] = s[0] * s[3] * s[1] * n[0];
    fa[10] = s[0] * s[3] * s[2] * n[1];
    fa[11] = (s[0] * s[3] * s[3] + s[4]) * n[2];
    double h = s[0] * (eqos1(p, 3, s[4], s[0]) + 0.5 * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3])) + s[4];
    fa[12] = n[0] * s[1] * h;
    fa[13] = n[1] * s[2] * h;
    fa[14] = n[2] * s[3] * h;
    for (int iv = 0; iv < 5; ++iv) out.adv[iv] = fa[iv * 3 + 0] + fa[iv * 3 + 1] + fa[iv * 3 + 2];
    out.unc[0] = out.unc[1] = out.unc[2] = 0.0;
    out.pF = 0.0;
    return;
  }
  bool isLeft = c[1] >= 0;
  bool isStar = c[0] < 0 && c[2] >= 0;
  const double *s = isLeft ? L : R;
  double un = isLeft ? uL : uR;
  double cK = isLeft ? c[0] : c[2];
  double unc[3] = {s[3], s[4], s[5]};
  double consF[10], vecS[10];
  prim2cons(p, s, consF);
  if (isStar) {
    // DP starRegion.h:3-35
    double SK = cK, SS = c[1];
    for (int i = 0; i < 3; ++i) unc[i] = s[3 + i] + (SS - un) * n[i];
    double rhoSL = s[1] * (SK - un) / (SK - SS);
    double rhoSG = s[2] * (SK - un) / (SK - SS);
    vecS[0] = s[0];
    vecS[1] = s[0] * rhoSL;
    vecS[2] = (1 - s[0]) * rhoSG;
    double rhoS = vecS[1] + vecS[2];
    vecS[3] = rhoS * unc[0];
    vecS[4] = rhoS * unc[1];
    vecS[5] = rhoS * unc[2];
    const double g0 = p.gamma[0], g1 = p.gamma[1], pi0 = p.pInf[0], pi1 = p.pInf[1];
    double pSL = (s[6] + pi0) * ((g0 - 1) * s[1] - (g0 + 1) * rhoSL) / ((g0 - 1) * rhoSL - (g0 + 1) * s[1]) - pi0;
    double pSG = (s[7] + pi1) * ((g1 - 1) * s[2] - (g1 + 1) * rhoSG) / ((g1 - 1) * rhoSG - (g1 + 1) * s[2]) - pi1;
    consF[6] = vecS[1] * eqos2(p, 0, 3, pSL, rhoSL);
    consF[7] = vecS[2] * eqos2(p, 1, 3, pSG, rhoSG);
    double press = s[0] * s[6] + (1 - s[0]) * s[7];
    double rho = s[0] * s[1] + (1 - s[0]) * s[2];
    if (p.hllcEnergy == 1)
      vecS[8] = ((SK - un) / (SK - SS)) * (consF[8] + (SS - un) * (rho * SS + press / (SK - un)));
    else
#ifndef ORC_PLANT_ENERGY_P_SIGN
      vecS[8] = ((SK - un) / (SK - SS)) * (consF[8] + (SS - un) * (SS + press / (rho * (SK - un))));
#else
      vecS[8] = ((SK - un) / (SK - SS)) * (consF[8] + (SS - un) * (SS - press / (rho * (SK - un))));
#endif
    vecS[9] = s[8] * (SK - un) / (SK - SS);
    // computeFluxHllc (DP computeRHS.cu:24-36)
    out.hllc[1] = cK * (vecS[1] - consF[1]);
    out.hllc[2] = cK * (vecS[2] - consF[2]);
    out.hllc[3] = cK * (vecS[3] - consF[3]);
    out.hllc[4] = cK * (vecS[4] - consF[4]);
    out.hllc[5] = cK * (vecS[5] - consF[5]);
    out.hllc[8] = cK * (vecS[8] - consF[8]);
    out.hllc[9] = cK * (vecS[9] - consF[9]);
  }
  // DP flux.h:32-91
  double rho = s[0] * s[1] + (1 - s[0]) * s[2];
  double press = s[0] * s[6] + (1 - s[0]) * s[7];
  double rE = s[0] * s[1] * eqos2(p, 0, 3, s[6], s[1]) + (1 - s[0]) * s[2] * eqos2(p, 1, 3, s[7], s[2]) +
              0.5 * rho * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  double *F = out.adv;
  const double areL = consF[6], areG = consF[7];
  F[0] = unc[0] * s[0] * n[0];
  F[0] += unc[1] * s[0] * n[1];
  F[0] += unc[2] * s[0] * n[2];
  F[1] = s[3] * s[0] * s[1] * n[0];
  F[1] += s[4] * s[0] * s[1] * n[1];
  F[1] += s[5] * s[0] * s[1] * n[2];
  F[2] = s[3] * (1 - s[0]) * s[2] * n[0];
  F[2] += s[4] * (1 - s[0]) * s[2] * n[1];
  F[2] += s[5] * (1 - s[0]) * s[2] * n[2];
  F[3] = rho * s[3] * s[3] * n[0];
  F[3] += rho * s[3] * s[4] * n[1];
  F[3] += rho * s[3] * s[5] * n[2];
  F[4] = rho * s[4] * s[3] * n[0];
  F[4] += rho * s[4] * s[4] * n[1];
  F[4] += rho * s[4] * s[5] * n[2];
  F[5] = rho * s[5] * s[3] * n[0];
  F[5] += rho * s[5] * s[4] * n[1];
  F[5] += rho * s[5] * s[5] * n[2];
  F[6] = unc[0] * areL * n[0];
  F[6] += unc[1] * areL * n[1];
  F[6] += unc[2] * areL * n[2];
  F[7] = unc[0] * areG * n[0];
  F[7] += unc[1] * areG * n[1];
  F[7] += unc[2] * areG * n[2];
  F[8] = s[3] * (rE + press) * n[0];
  F[8] += s[4] * (rE + press) * n[1];
  F[8] += s[5] * (rE + press) * n[2];
  F[9] = s[3] * s[8] * n[0];
  F[9] += s[4] * s[8] * n[1];
  F[9] += s[5] * s[8] * n[2];
  for (int i = 0; i < 3; ++i) out.unc[i] = unc[i];
  out.pF = s[0] * s[6] + (1 - s[0]) * s[7];  // pressureCorrection pF (computeRHS.cu:57)
}

// Side contribution of one face to one cell: DP fluxSource (flux.h:1-30),
// computeCenterPress/pressureCorrection (computeRHS.cu:38-59), atomicAddRHS
// (computeRHS.cu:61-75); SP computeRHS.cu:90-91.
void side_contrib(const fk_params &p, const FaceFlux &ff, const double *n, double A, const double *Uc, int side,
                  double *out) {
  double sign = (side == 0) ? -1.0 : 1.0;
  if (p.model == FK_MODEL_SINGLE) {
    for (int iv = 0; iv < 5; ++iv) {
      double t = ff.adv[iv] + ff.hllc[iv];
      out[iv] = side == 0 ? -t * A : t * A;
    }
    return;
  }
  double avf = clampAvf(p, Uc[0]);
  double rhoLiq = Uc[1] / avf;
  double rhoGas = Uc[2] / (1 - avf);
  double eLiq = Uc[6] / (avf * rhoLiq);
  double eGas = Uc[7] / ((1 - avf) * rhoGas);
  double pLiq = eqos2(p, 0, 0, rhoLiq, eLiq);
  double pGas = eqos2(p, 1, 0, rhoGas, eGas);
  const double *u = ff.unc;
  double nc0 = avf * u[0] * n[0] + avf * u[1] * n[1] + avf * u[2] * n[2];
  double nc1 = avf * pLiq * u[0] * n[0] + avf * pLiq * u[1] * n[1] + avf * pLiq * u[2] * n[2];
  double nc2 = (1 - avf) * pGas * u[0] * n[0] + (1 - avf) * pGas * u[1] * n[1] + (1 - avf) * pGas * u[2] * n[2];
  // computeCenterPress: the same clamp/EOS chain on the cell's own state
  double pC = avf * pLiq + (1 - avf) * pGas;
  double P = ff.pF - pC;
  const double *F = ff.adv, *H = ff.hllc;
  out[0] = sign * (F[0] - nc0) * A;
  out[1] = sign * (F[1] + H[1]) * A;
  out[2] = sign * (F[2] + H[2]) * A;
  out[3] = sign * (F[3] + H[3] + P * n[0]) * A;
  out[4] = sign * (F[4] + H[4] + P * n[1]) * A;
  out[5] = sign * (F[5] + H[5] + P * n[2]) * A;
  out[6] = sign * (F[6] + nc1) * A;
  out[7] = sign * (F[7] + nc2) * A;
  out[8] = sign * (F[8] + H[8]) * A;
  out[9] = sign * (F[9] + H[9]) * A;
}

// ---- WALE (DP computeViscousStresses.cu:3-60) ----
double wale(const double g[3][3], double Ls) {
  double sq[3][3], sd[3][3], st[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      sq[i][j] = 0.0;
      for (int k = 0; k < 3; k++) sq[i][j] += g[i][k] * g[k][j];
    }
  double T = 0.0;
  for (int i = 0; i < 3; i++) T += sq[i][i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; ++j) {
      int kron = (i == j);
      sd[i][j] = 0.5 * (sq[i][j] + sq[j][i]) - (1.0 / 3.0) * kron * T;
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) st[i][j] = 0.5 * (g[i][j] + g[j][i]);
  double sum1 = 0, sum2 = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; ++j) {
      sum1 += sd[i][j] * sd[i][j];
      sum2 += st[i][j] * st[i][j];
    }
  double num = sum1 * std::sqrt(sum1);
  double denom = sum1 * std::sqrt(std::sqrt(sum1)) + sum2 * sum2 * std::sqrt(sum2) + 1e-12;
  if (denom < 1e-10) num = 0;
  return Ls * Ls * num / denom;
}

// ---- relaxation: DP relaxation.cu, thermRlx.cu, chemRlx.cu ----
double tsat(const fk_params &p, double P, double temp, double urf, int *iters) {
  double Pinf = p.pInf[0], g1 = p.gamma[0], g2 = p.gamma[1], Cp1 = p.Cp[0], Cp2 = p.Cp[1];
  double Cv1 = Cp1 / g1, Cv2 = Cp2 / g2, eta1 = p.eta[0], eta2 = p.eta[1];
  double etaTone1 = 0;
  double etaTone2 = p.satConst[0] * fkmath::log10(P) + p.satConst[1];
  double As = (Cp1 - Cp2 + etaTone2 - etaTone1) / (Cp2 - Cv2);
  double Bs = (eta1 - eta2) / (Cp2 - Cv2);
  double Cs = (Cp2 - Cp1) / (Cp2 - Cv2);
  double Ds = (Cp1 - Cv1) / (Cp2 - Cv2);
#ifdef ORC_PLANT_TSAT_NO_D
  Ds = 0;
#endif
  double dt = 1.e-3, fun = 1, T = temp;
  int itr = 0;
  const double lpp = fkmath::log10(P + Pinf), lp = fkmath::log10(P);  // loop invariants
  while (std::fabs(fun) > 1.e-2 && itr < 50) {
    fun = As + Bs / T + Cs * fkmath::log10(T) + Ds * lpp - lp;
    double fun_dt = As + Bs / (T + dt) + Cs * fkmath::log10(T + dt) + Ds * lpp - lp;
    double grad_f = (fun_dt - fun) / dt;
    T = T - urf * fun / grad_f;
    itr++;
  }
  if (iters) *iters = itr;
  return T;
}

struct Therm {
  const fk_params &p;
  int primMat;
  double ar[2], enr[2];
  float dP = 0, dT = 0;
  void residuals(double x[4], double *r0, double *r1) const {
    double sum1 = 0, sum2 = 0;
    for (int k = 0; k < 2; k++)
      if (k != primMat) {
        sum1 = ar[k] / x[k];
        sum2 = ar[k] * (x[2 + k] - enr[k]);
      }
    x[primMat] = (1 / (1 - sum1)) * ar[primMat];
    x[2 + primMat] = enr[primMat] - (1 / ar[primMat]) * sum2;
    double pr[2], te[2];
    for (int k = 0; k < 2; k++) {
      pr[k] = eqos2(p, k, 0, x[k], x[k + 2]);
      te[k] = eqos2(p, k, 1, x[k], x[k + 2]);
    }
    *r0 = pr[0] - pr[1] - dP;
    *r1 = te[0] - te[1] - dT;
  }
};

// fntPressRlx (thermRlx.cu:42-67): finite-rate targets dP, dT of the thermal
// residuals.  The reference's types are kept: `float` for the pressures,
// temperatures, derivatives, c, m, theta, dP and dT, `double` for kappa and the
// exponential (exp as fkmath::exp, bitwise the same on host and device).
void fnt_press_rlx(const fk_params &p, const double avf[2], const double rho[2], const double ener[2], float *dP,
                   float *dT) {
  float de_drho[2], de_dP[2], p_I[2], c[2];
  float m = (float)p.thermRlxRate;
  float press[2];
  press[0] = (float)eqos2(p, 0, 0, rho[0], ener[0]);
  press[1] = (float)eqos2(p, 1, 0, rho[1], ener[1]);
  double kappa = 0;
  for (int iv = 0; iv < 2; iv++) {
    de_drho[iv] = (float)(-((p.gamma[iv] * p.pInf[iv] + press[iv]) / ((p.gamma[iv] - 1) * (rho[iv] * rho[iv]))));
    de_dP[iv] = (float)(1 / ((p.gamma[iv] - 1) * rho[iv]));
    p_I[iv] = press[iv];
    c[iv] = (float)((p_I[iv] / (rho[iv] * rho[iv]) - de_drho[iv]) / de_dP[iv]);
    kappa += rho[iv] * c[iv] / avf[iv];
  }
  kappa = kappa * avf[1] * m;
  *dP = (float)((press[0] - press[1]) * fkmath::exp(-std::fabs(kappa) * p.dt));
  float temp[2];
  temp[0] = (float)eqos2(p, 0, 1, rho[0], ener[0]);
  temp[1] = (float)eqos2(p, 1, 1, rho[1], ener[1]);
  float theta = (press[0] - press[1] > 1e-5) ? (*dP) / (press[0] - press[1]) : 0.0f;
  *dT = theta * (temp[0] - temp[1]);
}

// thermalRlx for one cell (thermRlx.cu:189-235 with helpers :4-187)
int therm_cell(const fk_params &p, double *U) {
  const int itrMax = p.thermItrMax;
  const double urf = p.thermUrf;
  double avf[2], rho[2], vel[3], Y[2], x[4], abortX[4];
  Therm t{p, 0, {0, 0}, {0, 0}};
  avf[0] = clampAvf(p, U[0]);
  avf[1] = 1 - avf[0];
  rho[0] = U[1] / avf[0];
  rho[1] = U[2] / avf[1];
  t.ar[0] = avf[0] * rho[0];
  t.ar[1] = avf[1] * rho[1];
  t.enr[0] = U[6] / t.ar[0];
  t.enr[1] = U[7] / t.ar[1];
  double rhoMix = t.ar[0] + t.ar[1];
  vel[0] = U[3] / rhoMix;
  vel[1] = U[4] / rhoMix;
  vel[2] = U[5] / rhoMix;
  Y[0] = t.ar[0] / rhoMix;
  Y[1] = t.ar[1] / rhoMix;
  double rE = U[8];
  x[0] = rho[0];
  x[1] = rho[1];
  x[2] = t.enr[0];
  x[3] = t.enr[1];
  for (int i = 0; i < 4; ++i) abortX[i] = x[i];
  double mx = -1.0;
  for (int k = 0; k < 2; k++)
    if (avf[k] > mx) {
      mx = avf[k];
      t.primMat = k;
    }
  const int pm = t.primMat;
  double de = rE / rhoMix - Y[0] * t.enr[0] - Y[1] * t.enr[1] - 0.5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
  t.enr[0] += de;
  t.enr[1] += de;
  if (p.thermFntRlxFlag == 1) fnt_press_rlx(p, avf, rho, t.enr, &t.dP, &t.dT);  // thermRlx.cu:213-214
  int abortFlag = 0;
  float minRes = 1e9f;
  bool conv = false;
  int itr = 0;
  while (!conv) {
    itr++;
    double res[2], jac[2][2], inv[2][2];
    t.residuals(x, &res[0], &res[1]);
    // computeJacobian
    double e[2], rdf[2][2];
    int iv = 0;
    for (int j = 0; j < 4; j++)
      if (j != pm && j != pm + 2) {
        e[iv] = 1e-4 * x[j];
        x[j] += e[iv];
        t.residuals(x, &rdf[0][iv], &rdf[1][iv]);
        x[j] -= e[iv];
        iv++;
      }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) {
        jac[i][j] = (rdf[i][j] - res[i]) / e[j];
        if (std::isnan(jac[i][j])) abortFlag = 1;
      }
    // inverseJacobian
    double det = jac[0][0] * jac[1][1] - jac[1][0] * jac[0][1];
    inv[0][0] = jac[1][1] / det;
    inv[1][1] = jac[0][0] / det;
    inv[1][0] = -jac[1][0] / det;
    inv[0][1] = -jac[0][1] / det;
    if (det == 0) abortFlag = 1;
    // abortCheckTherm
    if (abortFlag == 0 && x[0] > 0.0 && x[1] > 0.0) {
      float check = (float)std::sqrt(res[0] * res[0] + res[1] * res[1]);
      if (check < minRes) {
        minRes = check;
        for (int i = 0; i < 4; i++) abortX[i] = x[i];
      }
    }
    // updateSolution
    iv = 0;
    for (int i = 0; i < 4; i++)
      if (i != pm && i != pm + 2) {
        double dx = 0;
        for (int j = 0; j < 2; j++) dx += inv[iv][j] * res[j];
        x[i] = x[i] - urf * dx;
        iv++;
      }
    int check = 0;
    for (int j = 0; j < 2; j++)
      if (std::fabs(res[j]) > 1.e-2) check += 1;
    if (check == 0) conv = true;
    if (abortFlag == 1 || itr > itrMax) {
      for (int i = 0; i < 4; i++) x[i] = abortX[i];
      conv = true;
    }
  }
  // updateConsVec
  double sum1 = 0, sum2 = 0;
  for (int k = 0; k < 2; k++)
    if (k != pm) {
      sum1 = t.ar[k] / x[k];
      sum2 = t.ar[k] * (x[2 + k] - t.enr[k]);
    }
  x[pm] = (1 / (1 - sum1)) * t.ar[pm];
  x[2 + pm] = t.enr[pm] - (1 / t.ar[pm]) * sum2;
  U[0] = clampAvf(p, t.ar[0] / x[0]);
  U[6] = t.ar[0] * x[2];
  U[7] = t.ar[1] * x[3];
  return itr;
}

// Gibbs free energy g_k(p, T) of chemRlx.cu.  The entropy term
// log10(T^gamma / (p + pinf)^(gamma - 1)) is evaluated as
// gamma log10(T) - (gamma - 1) log10(p + pinf) (the same quantity, no pow).
double gibbs(const fk_params &p, int k, double P, double T) {
  double etaTone[2] = {0, p.satConst[0] * fkmath::log10(P) + p.satConst[1]};
  double enthalpy = p.Cp[k] * T + p.eta[k];
  double entropy = (p.Cp[k] / p.gamma[k]) * (p.gamma[k] * fkmath::log10(T) - (p.gamma[k] - 1) * fkmath::log10(P + p.pInf[k])) +
                   etaTone[k];
  return enthalpy - T * entropy;
}

struct Chem {
  const fk_params &p;
  int primMat;
  double rhoMix, rEint;
  void residuals(double x[6], double *r0, double *r1, double *r2) const {
    double s1 = 0, s2 = 0, s3 = 0;
    for (int k = 0; k < 2; k++)
      if (k != primMat) {
        s1 = x[k];
        s2 = x[k] * x[k + 2];
        s3 = x[k] * x[k + 2] * x[k + 4];
      }
    x[primMat] = 1 - s1;
    x[primMat + 2] = (rhoMix - s2) / x[primMat];
    x[primMat + 4] = (rEint - s3) / (x[primMat] * x[primMat + 2]);
    double pr[2], te[2], gi[2];
    for (int k = 0; k < 2; k++) {
      pr[k] = eqos2(p, k, 0, x[k + 2], x[k + 4]);
      te[k] = eqos2(p, k, 1, x[k + 2], x[k + 4]);
      gi[k] = gibbs(p, k, pr[k], te[k]);
    }
    *r0 = pr[0] - pr[1];
    *r1 = te[0] - te[1];
    *r2 = gi[0] - gi[1];
  }
};

// checkUpdate (chemRlx.cu:139-162); urf arrives as float as in the reference
void check_update(const fk_params &p, double xOld, double *x, double dx, int i, float urf, int *abortFlag, double rho) {
  if (i == 0 || i == 1) {
#ifndef ORC_PLANT_NO_ALPHA_BOUND
    if (x[i] < 0) {
      while (x[i] < p.chemLimit && urf > 0.0005) {
        urf = urf / 2;
        x[i] = xOld - urf * dx;
      }
    }
    if (x[i] > 1) {
      while (x[i] > 1.0 - p.chemLimit && urf > 0.0005) {
        urf = urf / 2;
        x[i] = xOld - urf * dx;
      }
    }
    if (x[i] < 0) *abortFlag = 1;
    if (x[i] > 1) *abortFlag = 1;
#endif
  } else if (i == 2 || i == 3) {
    double ar = x[i - 2] * x[i];
    if (x[i] < 0) {
      while (x[i] < 0 && urf > 0.0005) {
        urf = urf / 2;
        x[i] = xOld - urf * dx;
      }
    }
#ifndef ORC_PLANT_NO_DENSITY_BOUND
    if (ar > rho) {
      while (ar > rho && urf > 0.0005) {
        urf = urf / 2;
        x[i] = xOld - urf * dx;
        ar = x[i - 2] * x[i];
      }
    }
    if (ar > rho) *abortFlag = 1;
#endif
    if (x[i] < 0) *abortFlag = 1;
  }
}

// Record of the chemical Newton iterations, for the oracle's own tests
// (orc_chem_trace); per iteration kChemTraceCols doubles:
//   [0, 6)   x before the update, [6, 9) its residuals,
//   [9]      abortCheck: 0 x not eligible (abort flag or bounds), 1 eligible,
//            2 eligible and taken as the best so far,
//   [10]     the FP32 residual norm abortCheck compares,
//   [11, 17) x after the update (with checkUpdate's halving),
//   [17]     1 if the update raised the abort flag,
//   [18, 20) the non-primary phase's volume fraction and density after the
//            full Newton step, before checkUpdate.
constexpr int kChemTraceCols = 20;
struct ChemTrace {
  double *rows;
  int max_rows, n_rows;
  int aborted;  // the solve ended on the abort path (abortX restored)
};

// chemRlx for one cell (chemRlx.cu:217-267 with helpers :4-215)
int chem_cell(const fk_params &p, double *U, ChemTrace *tr = nullptr) {
  const int itrMax = p.chemItrMax;
  const double urf = p.chemUrf;
  double x[6], abortX[6];
  Chem ch{p, 0, 0, 0};
  x[0] = clampAvf(p, U[0]);
  x[1] = 1 - x[0];
  x[2] = U[1] / x[0];
  x[3] = U[2] / x[1];
  x[4] = U[6] / (x[0] * x[2]);
  x[5] = U[7] / (x[1] * x[3]);
  for (int i = 0; i < 6; ++i) abortX[i] = x[i];
  ch.rhoMix = x[0] * x[2] + x[1] * x[3];
  double u = U[3] / ch.rhoMix, v = U[4] / ch.rhoMix, w = U[5] / ch.rhoMix;
  ch.rEint = U[8] - 0.5 * ch.rhoMix * (u * u + v * v + w * w);
  double mx = -1.0;
  for (int k = 0; k < 2; k++)
    if (x[k] > mx) {
      mx = x[k];
      ch.primMat = k;
    }
  const int pm = ch.primMat;
  int abortFlag = 0;
  float minRes = 1e9f;
  bool conv = false;
  int itr = 0;
  while (!conv) {
    itr++;
    double res[3], jac[3][3], inv[3][3];
    ch.residuals(x, &res[0], &res[1], &res[2]);
    double e[3], rdf[3][3];
    int iv = 0;
    for (int j = 0; j < 6; j++)
      if (j != pm && j != pm + 2 && j != pm + 4) {
        e[iv] = std::fabs(x[j]) < 1e-9 ? 1e-10 : 1.e-5 * x[j];
        x[j] += e[iv];
        ch.residuals(x, &rdf[0][iv], &rdf[1][iv], &rdf[2][iv]);
        x[j] -= e[iv];
        iv++;
      }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        jac[i][j] = (rdf[i][j] - res[i]) / e[j];
        if (std::isnan(jac[i][j])) abortFlag = 1;
      }
    double det = jac[0][0] * (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1]) -
                 jac[0][1] * (jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]) +
                 jac[0][2] * (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]);
    if (det == 0) abortFlag = 1;
    double id = 1.0 / det;
    inv[0][0] = id * (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1]);
    inv[0][1] = -id * (jac[0][1] * jac[2][2] - jac[0][2] * jac[2][1]);
    inv[0][2] = id * (jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1]);
    inv[1][0] = -id * (jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]);
    inv[1][1] = id * (jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0]);
    inv[1][2] = -id * (jac[0][0] * jac[1][2] - jac[0][2] * jac[1][0]);
    inv[2][0] = id * (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]);
    inv[2][1] = -id * (jac[0][0] * jac[2][1] - jac[0][1] * jac[2][0]);
    inv[2][2] = id * (jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0]);
    // abortCheck (chemRlx.cu:125-137)
    double *row = (tr && tr->n_rows < tr->max_rows) ? tr->rows + (size_t)tr->n_rows++ * kChemTraceCols : nullptr;
    if (row) {
      for (int i = 0; i < 6; i++) row[i] = x[i];
      for (int i = 0; i < 3; i++) row[6 + i] = res[i];
      row[9] = 0;
      row[10] = (float)std::sqrt(res[0] * res[0] + res[1] * res[1] + res[2] * res[2]);
    }
    if (abortFlag == 0 && x[0] > 0.0 && x[0] < 1.0 && x[1] > 0.0 && x[1] < 1.0 && x[2] > 0.0 && x[3] > 0.0) {
      float check = (float)std::sqrt(res[0] * res[0] + res[1] * res[1] + res[2] * res[2]);
      if (row) row[9] = 1;
      if (check < minRes) {
        minRes = check;
        for (int i = 0; i < 6; i++) abortX[i] = x[i];
        if (row) row[9] = 2;
      }
    }
    // updateSolution (chemRlx.cu:164-193): abortFlag is a by-value copy here
    int abortLocal = abortFlag;
    iv = 0;
    for (int i = 0; i < 6; i++)
      if (i != pm && i != pm + 2 && i != pm + 4) {
        double dx = 0;
        for (int j = 0; j < 3; j++) dx += inv[iv][j] * res[j];
        x[i] = x[i] - urf * dx;
        if (row && (i == 0 || i == 1)) row[18] = x[i];
        if (row && (i == 2 || i == 3)) row[19] = x[i];
        check_update(p, x[i] + urf * dx, x, dx, i, (float)urf, &abortLocal, ch.rhoMix);
        iv++;
      }
    if (row) {
      for (int i = 0; i < 6; i++) row[11 + i] = x[i];
      row[17] = abortLocal;
    }
    int check = 0;
    for (int j = 0; j < 3; j++)
      if (std::fabs(res[j]) > 1.e-1) check += 1;
    if (check == 0) conv = true;
    if (abortLocal == 1 || itr > itrMax) {
#ifndef ORC_PLANT_ABORT_KEEPS_LAST
      for (int i = 0; i < 6; i++) x[i] = abortX[i];
#endif
      if (tr) tr->aborted = 1;
      conv = true;
    }
  }
  // updateConsVec (chemRlx.cu:195-215)
  double s1 = 0, s2 = 0, s3 = 0;
  for (int k = 0; k < 2; k++)
    if (k != pm) {
      s1 = x[k];
      s2 = x[k] * x[k + 2];
      s3 = x[k] * x[k + 2] * x[k + 4];
    }
  x[pm] = 1 - s1;
  x[pm + 2] = (ch.rhoMix - s2) / x[pm];
  x[pm + 4] = (ch.rEint - s3) / (x[pm] * x[pm + 2]);
  U[0] = clampAvf(p, x[0]);
  U[1] = x[0] * x[2];
  U[2] = x[1] * x[3];
  U[6] = x[0] * x[2] * x[4];
  U[7] = x[1] * x[3] * x[5];
  U[9] = x[1] * x[3];
  return itr;
}

// partitionRelaxation classification (relaxation.cu:46-75) + the solve
int relax_cell(const fk_params &p, double *U, int *iters, int *tsat_iters) {
  int kind = 0;
  *tsat_iters = 0;
  if (p.chemFlag == 1) {
    double avf = clampAvf(p, U[0]);
    double rho = U[1] / avf;
    double enr = U[6] / (avf * rho);
    double tempL = eqos2(p, 0, 1, rho, enr);
    double pressL = eqos2(p, 0, 0, rho, enr);
    double Ts = tsat(p, pressL, tempL, p.satTempUrf, tsat_iters);
    kind = (tempL > Ts && avf > p.chemLimit) ? 1 : 0;
  }
  *iters = kind ? chem_cell(p, U) : therm_cell(p, U);
  return kind;
}

}  // namespace

struct orc {
  fk_params p{};
  int neq = 5;
  int64_t nc = 0, nf = 0;
  std::vector<int32_t> fc2el, bc;
  std::vector<double> n, area, fcc, ccc, vol, shift, volCfl, waleLs;
  int ns = 4;                // face slots per cell: 4 (all tetrahedra) or 6
  std::vector<int32_t> cf;   // [nc][ns] faces of each cell, ascending, -1 past the last
  std::vector<int8_t> ncf;   // [nc] faces of each cell (4 tet, 5 pyramid / prism, 6 hex)
  std::vector<double> U, U0, primF, grad, wmin, wmax, theta, ext, rhs, signal;
  std::vector<FaceFlux> ff;
  std::vector<double> visc;  // [nf][5]: momentum x,y,z, energy, heat
  std::vector<double> gT;    // [nc][3] temperature gradient (single-phase heat flux)
  int64_t step = 0;
  double time = 0, dt = 0;
};

namespace {

thread_local std::string g_err;
int fail(int code, const char *m) {
  g_err = m;
  return code;
}

void set_params(orc &o, const fk_params &p) {
  o.p = p;
  o.neq = p.model == FK_MODEL_SINGLE ? 5 : 10;
}

}  // namespace

extern "C" {

int orc_create(const fk_params *p, int64_t nElem, int64_t nFcs, const int32_t *fc2el, const int32_t *boundCond,
               const double *n, const double *area, const double *fcCenter, const double *cellCenter, const double *volume,
               const double *shift, orc **out) {
  if (!p || !fc2el || !boundCond || !n || !area || !fcCenter || !cellCenter || !volume || !out)
    return fail(FK_EINVAL, "null argument");
  if (p->model != FK_MODEL_SINGLE && p->model != FK_MODEL_TWO_PHASE) return fail(FK_EINVAL, "bad model");
  orc *o = new (std::nothrow) orc;
  if (!o) return fail(FK_ENOMEM, "oom");
  try {
    set_params(*o, *p);
    o->nc = nElem;
    o->nf = nFcs;
    o->fc2el.assign(fc2el, fc2el + 2 * nFcs);
    o->bc.assign(boundCond, boundCond + nFcs);
    o->n.assign(n, n + 3 * nFcs);
    o->area.assign(area, area + nFcs);
    o->fcc.assign(fcCenter, fcCenter + 3 * nFcs);
    o->ccc.assign(cellCenter, cellCenter + 3 * nElem);
    o->vol.assign(volume, volume + nElem);
    if (shift) o->shift.assign(shift, shift + 3 * nFcs);
    else o->shift.assign((size_t)(3 * nFcs), 0.0);
    std::vector<int> cnt((size_t)nElem, 0);
    for (int64_t f = 0; f < nFcs; ++f)
      for (int s = 0; s < 2; ++s) {
        int32_t c = fc2el[2 * f + s];
        if (c < 0) continue;
        if (c >= nElem || cnt[(size_t)c] >= 6) {
          delete o;
          return fail(FK_EINVAL, "cell with more than six faces or bad fc2el");
        }
        cnt[(size_t)c]++;
      }
    o->ns = 4;
    for (int64_t c = 0; c < nElem; ++c) {
      if (cnt[(size_t)c] < 4) {
        delete o;
        return fail(FK_EINVAL, "every cell needs four to six faces");
      }
      if (cnt[(size_t)c] != 4) o->ns = 6;
    }
    const int ns = o->ns;
    o->cf.assign((size_t)(ns * nElem), -1);
    o->ncf.assign((size_t)nElem, 0);
    for (int64_t f = 0; f < nFcs; ++f)  // ascending face id per cell
      for (int s = 0; s < 2; ++s) {
        int32_t c = fc2el[2 * f + s];
        if (c < 0) continue;
        o->cf[(size_t)(ns * c + o->ncf[(size_t)c]++)] = (int32_t)f;
      }
    o->volCfl.resize((size_t)nElem);
    for (int64_t c = 0; c < nElem; ++c) o->volCfl[(size_t)c] = std::pow(volume[c], 0.3333);
    o->waleLs.resize((size_t)nFcs);
    for (int64_t f = 0; f < nFcs; ++f) {
      int32_t a = fc2el[2 * f], b = fc2el[2 * f + 1] >= 0 ? fc2el[2 * f + 1] : fc2el[2 * f];
      o->waleLs[(size_t)f] = 0.325 * std::pow((volume[a] + volume[b]) / 2, 0.3333333333333);
    }
    const int neq = o->neq;
    o->U.assign((size_t)(nElem * neq), 0.0);
    o->U0 = o->U;
    o->primF.assign((size_t)(nFcs * 2 * neq), 0.0);
    o->ext.assign((size_t)(nFcs * 2 * neq), 0.0);
    o->grad.assign((size_t)(nElem * neq * 3), 0.0);
    o->wmin.assign((size_t)(nElem * neq), 0.0);
    o->wmax = o->wmin;
    o->theta = o->wmin;
    o->rhs = o->wmin;
    o->signal.assign((size_t)nElem, 0.0);
    o->ff.resize((size_t)nFcs);
    o->visc.assign((size_t)(nFcs * 5), 0.0);
    o->gT.assign((size_t)(nElem * 3), 0.0);
  } catch (const std::bad_alloc &) {
    delete o;
    return fail(FK_ENOMEM, "out of host memory");
  }
  *out = o;
  return FK_OK;
}

void orc_free(orc *o) { delete o; }

int orc_set_params(orc *o, const fk_params *p) {
  if (!o || !p) return fail(FK_EINVAL, "null");
  int neq = p->model == FK_MODEL_SINGLE ? 5 : 10;
  if (neq != o->neq) return fail(FK_EINVAL, "model change");
  set_params(*o, *p);
  return FK_OK;
}

int orc_set_state(orc *o, const double *U) {
  if (!o || !U) return fail(FK_EINVAL, "null");
  std::copy(U, U + o->nc * o->neq, o->U.begin());
  return FK_OK;
}
int orc_get_state(const orc *o, double *U) {
  if (!o || !U) return fail(FK_EINVAL, "null");
  std::copy(o->U.begin(), o->U.end(), U);
  return FK_OK;
}

int orc_init_rk(orc *o) {
  o->U0 = o->U;  // rk3stage0 (update.cu:3-17)
  return FK_OK;
}

// computePrimVec (computePrimVec.cu:38-80) + computeGrad (computeGrad.cu:3-91)
// + computeFaceValues (computeFaceValues.cu:4-113)
int orc_compute_face_values(orc *o) {
  const fk_params &p = o->p;
  const int neq = o->neq;
  const int64_t nf = o->nf, nc = o->nc;
  // convert2PrimVec
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < nf; ++f) {
    int bc = o->bc[(size_t)f];
    int id = (bc == 0);
    double *P0 = &o->primF[(size_t)(f * 2 * neq)], *P1 = P0 + neq;
    cons2prim(p, &o->U[(size_t)(o->fc2el[(size_t)(2 * f)] * neq)], P0);
    cons2prim(p, &o->U[(size_t)(o->fc2el[(size_t)(2 * f + id)] * neq)], P1);
    if (bc != 0) bc_ghost(p, bc, &o->n[(size_t)(3 * f)], P1);
  }
  // greenGauss + atomicMinMaxDouble, accumulated per cell in ascending face order
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; ++c) {
    double *g = &o->grad[(size_t)(c * neq * 3)];
    double *mn = &o->wmin[(size_t)(c * neq)], *mx = &o->wmax[(size_t)(c * neq)];
    for (int v = 0; v < neq; ++v) {
      mn[v] = 1e12;
      mx[v] = -1e12;
      g[3 * v] = g[3 * v + 1] = g[3 * v + 2] = 0.0;
    }
    for (int k = 0; k < o->ncf[(size_t)c]; ++k) {
      int64_t f = o->cf[(size_t)(o->ns * c + k)];
      int side = o->fc2el[(size_t)(2 * f)] == c ? 0 : 1;
      const double *P0 = &o->primF[(size_t)(f * 2 * neq)], *P1 = P0 + neq;
      const double *nv = &o->n[(size_t)(3 * f)];
      double areaF = o->area[(size_t)f], V = o->vol[(size_t)c];
      for (int v = 0; v < neq; ++v) {
        double solL = P0[v], solR = P1[v];
        double other = side == 0 ? solR : solL;
        if (!(mx[v] >= other)) mx[v] = other;  // atomicMinMaxDouble update order: max then min
        if (!(mn[v] <= other)) mn[v] = other;
        for (int d = 0; d < 3; ++d) {
          if (side == 0) g[3 * v + d] += ((solL + solR) / 2) * nv[d] * areaF / V;
          else g[3 * v + d] += -((solL + solR) / 2) * nv[d] * areaF / V;
        }
      }
    }
  }
  // computeLimiter: extension per (face, side), theta per cell
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; ++c) {
    double *th = &o->theta[(size_t)(c * neq)];
    for (int v = 0; v < neq; ++v) th[v] = 1e9;
    for (int k = 0; k < o->ncf[(size_t)c]; ++k) {
      int64_t f = o->cf[(size_t)(o->ns * c + k)];
      int id = o->fc2el[(size_t)(2 * f)] == c ? 0 : 1;
      const double *sh = &o->shift[(size_t)(3 * f)];
      double xf = o->fcc[(size_t)(3 * f)], yf = o->fcc[(size_t)(3 * f + 1)], zf = o->fcc[(size_t)(3 * f + 2)];
      if (id == 1) {  // periodic: the face as seen from the neighbour
        xf = xf - sh[0];
        yf = yf - sh[1];
        zf = zf - sh[2];
      }
      double xc = o->ccc[(size_t)(3 * c)], yc = o->ccc[(size_t)(3 * c + 1)], zc = o->ccc[(size_t)(3 * c + 2)];
      for (int v = 0; v < neq; ++v) {
        double sol = o->primF[(size_t)(f * 2 * neq + id * neq + v)];
        const double *g = &o->grad[(size_t)(c * neq * 3 + v * 3)];
        double w_unlim = sol + (xf - xc) * g[0] + (yf - yc) * g[1] + (zf - zc) * g[2];
        o->ext[(size_t)(f * 2 * neq + id * neq + v)] = (xf - xc) * g[0] + (yf - yc) * g[1] + (zf - zc) * g[2];
        double wmax = o->wmax[(size_t)(c * neq + v)], wmin = o->wmin[(size_t)(c * neq + v)];
        double phi;
        if (w_unlim - sol != 0) phi = (w_unlim > sol) ? (wmax - sol) / (2 * (w_unlim - sol)) : (wmin - sol) / (2 * (w_unlim - sol));
        else phi = 1;
        double tt = std::fmax(0, std::fmin(phi, 1));
        if (!(th[v] <= tt)) th[v] = tt;  // atomicMinDouble
      }
    }
  }
  // updateFaces
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < nf; ++f) {
    int bc = o->bc[(size_t)f];
    for (int id = 0; id < 2; ++id) {
      if (!(id != 1 || bc == 0)) continue;
      int64_t c = o->fc2el[(size_t)(2 * f + id)];
      for (int v = 0; v < neq; ++v)
        o->primF[(size_t)(f * 2 * neq + id * neq + v)] += o->theta[(size_t)(c * neq + v)] * o->ext[(size_t)(f * 2 * neq + id * neq + v)];
    }
  }
  return FK_OK;
}

// computeRHS (computeRHS.cu:214-253) + computeViscousStresses
// (computeViscousStresses.cu:66-174) + per-cell signal of computeTimeStep
// (computeTimeStep.cu:15-46, 97-129)
int orc_compute_rhs(orc *o) {
  const fk_params &p = o->p;
  const int neq = o->neq;
  const int64_t nf = o->nf, nc = o->nc;
  const bool single = p.model == FK_MODEL_SINGLE;
  const bool visc = p.viscFlag == 1;
  const int iu = single ? 1 : 3;  // first velocity component of the primitive vector
  // temperature gradients for the single-phase heat flux, from the cell's rho, p
  if (visc && single && p.heatFlux == 1) {
    const double g = p.gamma[0], pi = p.pInf[0], R = p.Cp[0] * (g - 1) / g;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
      double W[5];
      cons2prim(p, &o->U[(size_t)(c * 5)], W);
      const double *gr = &o->grad[(size_t)(c * 15)];
      for (int d = 0; d < 3; ++d) o->gT[(size_t)(c * 3 + d)] = (gr[12 + d] - ((W[4] + pi) / W[0]) * gr[d]) / (W[0] * R);
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < nf; ++f) {
    int bc = o->bc[(size_t)f];
    const double *P0 = &o->primF[(size_t)(f * 2 * neq)], *P1 = P0 + neq;
    const double *nv = &o->n[(size_t)(3 * f)];
    face_flux(p, P0, P1, nv, bc, o->ff[(size_t)f]);
    if (!visc) continue;
    int id = (bc == 0);
    int64_t c0 = o->fc2el[(size_t)(2 * f)], c1 = o->fc2el[(size_t)(2 * f + id)];
    const double *g0 = &o->grad[(size_t)(c0 * neq * 3)], *g1 = &o->grad[(size_t)(c1 * neq * 3)];
    // face velocity gradient: rows iu..iu+2 (SP row bug at computeViscousStresses.cu:94 fixed)
    double gm[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) gm[i][j] = (g0[(iu + i) * 3 + j] + g1[(iu + i) * 3 + j]) / 2;
    double ux = gm[0][0], uy = gm[0][1], uz = gm[0][2], vx = gm[1][0], vy = gm[1][1], vz = gm[1][2];
    double wx = gm[2][0], wy = gm[2][1], wz = gm[2][2];
    double eddy = 0;
    if (p.turbulence == 1) eddy = wale(gm, o->waleLs[(size_t)f]);
    double st[3][3];
    st[0][0] = (2.0 * ux - (2.0 / 3.0) * (ux + vy + wz));
    st[1][1] = (2.0 * vy - (2.0 / 3.0) * (ux + vy + wz));
    st[2][2] = (2.0 * wz - (2.0 / 3.0) * (ux + vy + wz));
    st[0][1] = (uy + vx);
    st[0][2] = (uz + wx);
    st[1][2] = (vz + wy);
    st[1][0] = st[0][1];
    st[2][1] = st[1][2];
    st[2][0] = st[0][2];
    // primVecF side 1 of a boundary face is the unreconstructed ghost
    double R1[MAXEQ];
    if (bc != 0) {
      cons2prim(p, &o->U[(size_t)(c0 * neq)], R1);
      bc_ghost(p, bc, nv, R1);
    } else {
      for (int v = 0; v < neq; ++v) R1[v] = P1[v];
    }
    double mu;
    if (single) {
      double rhoF = (P0[0] + R1[0]) / 2;
      mu = p.dynVisc[0] + rhoF * eddy;
    } else {
      double avf1 = P0[0], rho1 = avf1 * P0[1] + (1 - avf1) * P0[2];
      double avf2 = R1[0], rho2 = avf2 * R1[1] + (1 - avf2) * R1[2];
      double rhoF = (rho1 + rho2) / 2, avfF = (avf1 + avf2) / 2;
      mu = avfF * p.dynVisc[0] + (1 - avfF) * p.dynVisc[1] + rhoF * eddy;
    }
    double vs[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) vs[i][j] = mu * st[i][j];
    double *out = &o->visc[(size_t)(f * 5)];
    for (int i = 0; i < 3; ++i) out[i] = vs[i][0] * nv[0] + vs[i][1] * nv[1] + vs[i][2] * nv[2];
    double velF[3];
    for (int i = 0; i < 3; ++i) velF[i] = (P0[iu + i] + R1[iu + i]) / 2;
    double fe[3];
    for (int i = 0; i < 3; i++) {
      double stv = 0;
      for (int j = 0; j < 3; j++) stv += st[i][j] * velF[j];
      fe[i] = (mu * stv) * nv[i];
    }
    out[3] = fe[0] + fe[1] + fe[2];
    out[4] = 0.0;
    if (single && p.heatFlux == 1) {
      const double *t0 = &o->gT[(size_t)(c0 * 3)], *t1 = &o->gT[(size_t)(c1 * 3)];
      double k = p.Cp[0] * mu / p.Pr[0];
      double gf[3];
      for (int d = 0; d < 3; ++d) gf[d] = (t0[d] + t1[d]) / 2;
      out[4] = k * (gf[0] * nv[0] + gf[1] * nv[1] + gf[2] * nv[2]);
    }
  }
  // per-cell gather in ascending face order; per-cell max signal speed
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; ++c) {
    double *r = &o->rhs[(size_t)(c * neq)];
    for (int v = 0; v < neq; ++v) r[v] = 0.0;
    double sig = -1e9;
    const double *Uc = &o->U[(size_t)(c * neq)];
    for (int k = 0; k < o->ncf[(size_t)c]; ++k) {
      int64_t f = o->cf[(size_t)(o->ns * c + k)];
      int side = o->fc2el[(size_t)(2 * f)] == c ? 0 : 1;
      const double *nv = &o->n[(size_t)(3 * f)];
      double A = o->area[(size_t)f];
      double cont[MAXEQ];
      side_contrib(p, o->ff[(size_t)f], nv, A, Uc, side, cont);
      for (int v = 0; v < neq; ++v) r[v] += cont[v];
      // computeFaceMaxSignal for both face states
      int bc = o->bc[(size_t)f];
      double cm[2] = {0.0, 0.0};
      for (int s = 0; s < (bc == 0 ? 2 : 1); ++s) {
        const double *W = &o->primF[(size_t)(f * 2 * neq + s * neq)];
        double un, cmix;
        if (single) {
          un = W[1] * nv[0] + W[2] * nv[1] + W[3] * nv[2];
          cmix = eqos1(p, 4, W[0], W[4]);
        } else {
          double avf = W[0], rL = W[1], rG = W[2];
          double rhoMix = avf * rL + (1 - avf) * rG;
          double cL = eqos2(p, 0, 4, rL, W[6]), cG = eqos2(p, 1, 4, rG, W[7]);
          if (W[8] > p.chemLimit && p.woodsFlag == 1)
            cmix = 1 / std::sqrt(rhoMix * (avf / (rL * cL * cL) + (1 - avf) / (rG * cG * cG)));
          else {
            double YL = avf * rL / rhoMix, YG = (1 - avf) * rG / rhoMix;
            cmix = std::sqrt(YL * cL * cL + YG * cG * cG);
          }
          un = W[3] * nv[0] + W[4] * nv[1] + W[5] * nv[2];
        }
        cm[s] = std::fabs(un) + cmix;
      }
      double cMax = std::fmax(cm[0], cm[1]);
      if (!(sig >= cMax)) sig = cMax;  // atomicMaxDouble
    }
    o->signal[(size_t)c] = sig;
    // viscous contributions summed separately in ascending face order, then
    // added to the inviscid sums (the reference adds both with atomics, in
    // no fixed order)
    if (visc) {
      double rv[5] = {0, 0, 0, 0, 0};  // momentum x,y,z, energy, heat
      for (int k = 0; k < o->ncf[(size_t)c]; ++k) {
        int64_t f = o->cf[(size_t)(o->ns * c + k)];
        int side = o->fc2el[(size_t)(2 * f)] == c ? 0 : 1;
        double A = o->area[(size_t)f];
        const double *vv = &o->visc[(size_t)(f * 5)];
        for (int i = 0; i < 4; ++i) rv[i] += side == 0 ? vv[i] * A : -(vv[i]) * A;
        if (single && p.heatFlux == 1) rv[4] += side == 0 ? vv[4] * A : -(vv[4]) * A;
      }
      for (int i = 0; i < 3; ++i) r[iu + i] += rv[i];
      r[single ? 4 : 8] += rv[3];
      if (single && p.heatFlux == 1) r[4] += rv[4];
    }
  }
  return FK_OK;
}

int orc_compute_time_step(orc *o, double *dt) {
  double mn = 1e15;
  for (int64_t c = 0; c < o->nc; ++c) mn = std::fmin(mn, o->volCfl[(size_t)c] / o->signal[(size_t)c]);
  o->dt = std::fmin(o->p.cfl * mn, o->p.maxDt);
  if (dt) *dt = o->dt;
  return FK_OK;
}

// rungeKutta / rk3stage1..3 (update.cu:19-103)
int orc_update(orc *o, int irk, double dt) {
  const int neq = o->neq;
  const int iz = o->p.model == FK_MODEL_SINGLE ? 3 : 5;
  const int rk = o->p.rkSteps;
  if (rk != 1 && rk != 3) return fail(FK_EINVAL, "rkSteps must be 1 or 3");
  if (irk < 0 || irk >= rk) return fail(FK_EINVAL, "bad stage");
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < o->nc; ++c) {
    double V = o->vol[(size_t)c];
    for (int v = 0; v < neq; ++v) {
      int neglect = (v == iz && o->p.neglectZ == 1);
      size_t i = (size_t)(c * neq + v);
      double R = o->rhs[i];
      if (rk == 1) o->U[i] = (1 - neglect) * (o->U[i] + dt * (1.0 / V) * R);
      else if (irk == 0) o->U[i] = (1 - neglect) * (o->U0[i] + dt * (1.0 / V) * R);
      else if (irk == 1)
        o->U[i] = (1 - neglect) * ((3.0 / 4.0) * o->U0[i] + (1.0 / 4.0) * o->U[i] + (1.0 / 4.0) * dt * (1.0 / V) * R);
      else
        o->U[i] = (1 - neglect) * ((1.0 / 3.0) * o->U0[i] + (2.0 / 3.0) * o->U[i] + (2.0 / 3.0) * dt * (1.0 / V) * R);
    }
  }
  return FK_OK;
}

int orc_relaxation(orc *o, int64_t stats[4]) {
  if (o->p.model != FK_MODEL_TWO_PHASE) return fail(FK_EINVAL, "relaxation needs the two-phase model");
  int64_t nt = 0, nch = 0, it = 0, ts = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : nt, nch, it, ts)
  for (int64_t c = 0; c < o->nc; ++c) {
    int iters, tsi;
    int kind = relax_cell(o->p, &o->U[(size_t)(c * 10)], &iters, &tsi);
    nch += kind;
    nt += 1 - kind;
    it += iters;
    ts += tsi;
  }
  if (stats) {
    stats[0] = nt;
    stats[1] = nch;
    stats[2] = it;
    stats[3] = ts;
  }
  return FK_OK;
}

// consCalc/integrateConsVar (consCalc.cu:3-34): iVar 0 integrates 1 - alpha
int orc_cons_calc(const orc *o, double *out) {
  const int neq = o->neq;
  for (int v = 0; v < neq; ++v) {
    double s = 0;
    for (int64_t c = 0; c < o->nc; ++c) {
      double u = o->U[(size_t)(c * neq + v)];
      if (v == 0 && neq == 10) u = 1 - u;
      s += u * o->vol[(size_t)c];
    }
    out[v] = s;
  }
  return FK_OK;
}

int orc_run_steps(orc *o, int nsteps, double *dt_last, double *phys_time) {
  const fk_params &p = o->p;
  for (int s = 0; s < nsteps; ++s) {
    if (p.rkSteps > 1) orc_init_rk(o);
    for (int irk = 0; irk < p.rkSteps; ++irk) {
      orc_compute_face_values(o);
      orc_compute_rhs(o);
      if (p.adaptDt == 1 && irk == 0 && o->step % p.adaptDtItr == 0) orc_compute_time_step(o, nullptr);
      else if (p.adaptDt != 1) o->dt = p.dt;
      int rc = orc_update(o, irk, o->dt);
      if (rc) return rc;
    }
    if (p.model == FK_MODEL_TWO_PHASE && p.relaxFlag == 1) orc_relaxation(o, nullptr);
    o->time += o->dt;
    o->step++;
  }
  if (dt_last) *dt_last = o->dt;
  if (phys_time) *phys_time = o->time;
  return FK_OK;
}

int orc_set_rhs(orc *o, const double *rhs) {
  if (!o || !rhs) return fail(FK_EINVAL, "null");
  std::copy(rhs, rhs + o->nc * o->neq, o->rhs.begin());
  return FK_OK;
}

int orc_get(const orc *o, const char *name, void *out, int64_t bytes) {
  if (!o || !name || !out) return fail(FK_EINVAL, "null");
  std::string s(name);
  const int neq = o->neq;
  auto put = [&](const void *src, int64_t nb) {
    if (nb != bytes) return fail(FK_EINVAL, "size mismatch");
    std::memcpy(out, src, (size_t)nb);
    return FK_OK;
  };
  if (s == "grad") return put(o->grad.data(), (int64_t)(o->grad.size() * 8));
  if (s == "theta") return put(o->theta.data(), (int64_t)(o->theta.size() * 8));
  if (s == "wmin") return put(o->wmin.data(), (int64_t)(o->wmin.size() * 8));
  if (s == "wmax") return put(o->wmax.data(), (int64_t)(o->wmax.size() * 8));
  if (s == "primF") return put(o->primF.data(), (int64_t)(o->primF.size() * 8));
  if (s == "rhs") return put(o->rhs.data(), (int64_t)(o->rhs.size() * 8));
  if (s == "signal") return put(o->signal.data(), (int64_t)(o->signal.size() * 8));
  if (s == "slots") return put(o->cf.data(), (int64_t)(o->cf.size() * 4));
  if (s == "nslots") return put(&o->ns, 4);
  if (s == "face") {
    const int ns = o->ns;
    if (bytes != o->nc * ns * neq * 8) return fail(FK_EINVAL, "size mismatch");
    double *d = (double *)out;
    for (int64_t c = 0; c < o->nc; ++c)
      for (int k = 0; k < ns; ++k) {
        int64_t f = o->cf[(size_t)(ns * c + k)];
        if (f < 0) {  // no face in this slot
          for (int v = 0; v < neq; ++v) d[(c * ns + k) * neq + v] = 0.0;
          continue;
        }
        int side = o->fc2el[(size_t)(2 * f)] == c ? 0 : 1;
        for (int v = 0; v < neq; ++v) d[(c * ns + k) * neq + v] = o->primF[(size_t)(f * 2 * neq + side * neq + v)];
      }
    return FK_OK;
  }
  return fail(FK_EINVAL, "unknown array");
}

double orc_eqos(const fk_params *p, int imat, int id, double v1, double v2) {
  return p->model == FK_MODEL_SINGLE ? eqos1(*p, id, v1, v2) : eqos2(*p, imat, id, v1, v2);
}
void orc_cons2prim(const fk_params *p, const double *U, double *W) { cons2prim(*p, U, W); }
void orc_prim2cons(const fk_params *p, const double *W, double *U) { prim2cons(*p, W, U); }
void orc_bc_ghost(const fk_params *p, int bc, const double *n, double *W) { bc_ghost(*p, bc, n, W); }

void orc_face_flux(const fk_params *p, const double *WL, const double *WR, const double *n, double A, const double *UL,
                   const double *UR, int bc, double *out) {
  const int neq = p->model == FK_MODEL_SINGLE ? 5 : 10;
  FaceFlux ff;
  face_flux(*p, WL, WR, n, bc, ff);
  side_contrib(*p, ff, n, A, UL, 0, out);
  if (bc == 0) side_contrib(*p, ff, n, A, UR, 1, out + neq);
  else
    for (int v = 0; v < neq; ++v) out[neq + v] = 0.0;
}

double orc_wale(const double *g, double vol) {
  double gm[3][3];
  for (int i = 0; i < 9; ++i) gm[i / 3][i % 3] = g[i];
  return wale(gm, 0.325 * std::pow(vol, 0.3333333333333));
}

double orc_tsat(const fk_params *p, double P, double T0, double urf, int *iters) { return tsat(*p, P, T0, urf, iters); }
double orc_log10(double x) { return fkmath::log10(x); }
double orc_exp(double x) { return fkmath::exp(x); }
void orc_fnt_press_rlx(const fk_params *p, const double *avf, const double *rho, const double *ener, float *dP,
                       float *dT) {
  fnt_press_rlx(*p, avf, rho, ener, dP, dT);
}
double orc_gibbs(const fk_params *p, int k, double P, double T) { return gibbs(*p, k, P, T); }

int orc_chem_trace(const fk_params *p, double *U, double *rows, int max_rows, int *n_rows, int *aborted) {
  ChemTrace tr{rows, max_rows, 0, 0};
  int it = chem_cell(*p, U, &tr);
  if (n_rows) *n_rows = tr.n_rows;
  if (aborted) *aborted = tr.aborted;
  return it;
}

int orc_relax_cell(const fk_params *p, double *U, int *iters, int *tsat_iters) {
  int it = 0, ts = 0;
  int k = relax_cell(*p, U, &it, &ts);
  if (iters) *iters = it;
  if (tsat_iters) *tsat_iters = ts;
  return k;
}

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
