// Device code of the per-cell phase relaxation (reference Newton solves of
// relaxation.cu, thermRlx.cu and chemRlx.cu), shared by the plain and the
// surrogate-warm-started relaxation kernels.  FP64 throughout.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/fooska.h"
#include "fk_math.h"

namespace fk {

// ---------------- phase relaxation (relaxation.cu, thermRlx.cu, chemRlx.cu) ----------------

FK_HD inline double clampd(const fk_params &P, double a) {
  return fmin(fmax(P.avfMin / 100, a), 1.0 - P.avfMin / 100);
}
FK_HD inline double eos_p(const fk_params &P, int k, double rho, double e) {
  return (P.gamma[k] - 1) * rho * e - P.gamma[k] * P.pInf[k] - (P.gamma[k] - 1) * P.eta[k] * rho;
}
FK_HD inline double eos_T(const fk_params &P, int k, double rho, double e) {
  double p = (P.gamma[k] - 1) * rho * e - P.gamma[k] * P.pInf[k] - (P.gamma[k] - 1) * P.eta[k] * rho;
  return (p + P.pInf[k]) / ((P.Cp[k] / P.gamma[k]) * (P.gamma[k] - 1) * rho);
}

FK_HD inline double tsat_dev(const fk_params &P, double Pr, double temp, double urf, int *iters) {
  const double Pinf = P.pInf[0], g1 = P.gamma[0], g2 = P.gamma[1], Cp1 = P.Cp[0], Cp2 = P.Cp[1];
  const double Cv1 = Cp1 / g1, Cv2 = Cp2 / g2;
  const double et2 = P.satConst[0] * fkmath::log10(Pr) + P.satConst[1];
  const double As = (Cp1 - Cp2 + et2 - 0.0) / (Cp2 - Cv2);
  const double Bs = (P.eta[0] - P.eta[1]) / (Cp2 - Cv2);
  const double Cs = (Cp2 - Cp1) / (Cp2 - Cv2);
  const double Ds = (Cp1 - Cv1) / (Cp2 - Cv2);
  const double dt = 1.e-3;
  // loop-invariant logarithms, evaluated once (same values in every iteration)
  const double lpp = fkmath::log10(Pr + Pinf), lp = fkmath::log10(Pr);
  double fun = 1, T = temp;
  int itr = 0;
  while (fabs(fun) > 1.e-2 && itr < 50) {
    fun = As + Bs / T + Cs * fkmath::log10(T) + Ds * lpp - lp;
    double fd = As + Bs / (T + dt) + Cs * fkmath::log10(T + dt) + Ds * lpp - lp;
    double gr = (fd - fun) / dt;
    T = T - urf * fun / gr;
    itr++;
  }
  *iters = itr;
  return T;
}

// Finite-rate targets of the thermal relaxation (fntPressRlx,
// thermRlx.cu:42-67).  With the model on, the Newton solve drives
// p_L - p_G to dP and T_L - T_G to dT instead of to zero, where
//   dP = (p_L - p_G) exp(-|kappa| dt),  kappa = alpha_G m sum_k rho_k c_k / alpha_k,
//   c_k = (p_k / rho_k^2 - de_k/drho_k) / (de_k/dp_k)   (= gamma_k (p_k + pinf_k) / rho_k,
//         the squared sound speed),
//   dT = theta (T_L - T_G),  theta = dP / (p_L - p_G) if p_L - p_G > 1e-5, else 0,
// all from the cell's state after the energy correction, once per cell.  The
// precisions of the reference are kept: p_k, T_k, the derivatives, c_k, the
// rate m and dP, dT, theta are FP32 values; kappa and the exponential are
// FP64.  dt is fk_params.dt, the configured step (the reference reads
// d_input->dt, which the adaptive time step never updates).  theta's FP32
// quotient is taken as the FP64 quotient rounded to FP32, which is the
// correctly rounded FP32 quotient (53 >= 2 * 24 + 2 bits).
FK_HD inline void fnt_press_rlx(const fk_params &P, const double avf[2], const double rho[2], const double ener[2],
                                float *dP, float *dT) {
  const float m = (float)P.thermRlxRate;
  float press[2], temp[2];
  double kappa = 0;
  for (int k = 0; k < 2; ++k) {
    press[k] = (float)eos_p(P, k, rho[k], ener[k]);
    const double g = P.gamma[k];
    const float de_drho = (float)(-((g * P.pInf[k] + (double)press[k]) / ((g - 1) * (rho[k] * rho[k]))));
    const float de_dp = (float)(1 / ((g - 1) * rho[k]));
    const float c = (float)(((double)press[k] / (rho[k] * rho[k]) - (double)de_drho) / (double)de_dp);
    kappa += rho[k] * (double)c / avf[k];
  }
  kappa = kappa * avf[1] * (double)m;
  const float dp = press[0] - press[1];
  *dP = (float)((double)dp * fkmath::exp(-fabs(kappa) * P.dt));
  for (int k = 0; k < 2; ++k) temp[k] = (float)eos_T(P, k, rho[k], ener[k]);
  const float theta = ((double)dp > 1e-5) ? (float)((double)*dP / (double)dp) : 0.0f;
  *dT = theta * (temp[0] - temp[1]);
}

struct ThermSys {
  int pm;
  double ar[2], enr[2];
  float dP, dT;  // finite-rate targets, 0 for equilibrium
  FK_HD void res(const fk_params &P, double x[4], double *r0, double *r1) const {
    const int k = 1 - pm;
    double s1 = ar[k] / x[k];
    double s2 = ar[k] * (x[2 + k] - enr[k]);
    x[pm] = (1 / (1 - s1)) * ar[pm];
    x[2 + pm] = enr[pm] - (1 / ar[pm]) * s2;
    double p0 = eos_p(P, 0, x[0], x[2]), p1 = eos_p(P, 1, x[1], x[3]);
    double t0 = eos_T(P, 0, x[0], x[2]), t1 = eos_T(P, 1, x[1], x[3]);
    *r0 = p0 - p1 - (double)dP;
    *r1 = t0 - t1 - (double)dT;
  }
};

FK_HD inline int therm_dev(const fk_params &P, double *U) {
  const int itrMax = P.thermItrMax;
  const double urf = P.thermUrf;
  double avf0 = clampd(P, U[0]);
  double avf1 = 1 - avf0;
  double rho0 = U[1] / avf0, rho1 = U[2] / avf1;
  ThermSys t;
  t.ar[0] = avf0 * rho0;
  t.ar[1] = avf1 * rho1;
  t.enr[0] = U[6] / t.ar[0];
  t.enr[1] = U[7] / t.ar[1];
  const double rhoMix = t.ar[0] + t.ar[1];
  const double v0 = U[3] / rhoMix, v1 = U[4] / rhoMix, v2 = U[5] / rhoMix;
  const double Y0 = t.ar[0] / rhoMix, Y1 = t.ar[1] / rhoMix;
  double x[4] = {rho0, rho1, t.enr[0], t.enr[1]};
  double ax[4] = {x[0], x[1], x[2], x[3]};
  t.pm = (avf1 > avf0) ? 1 : 0;  // primary = larger volume fraction (thermRlx.cu:146-152)
  const int pm = t.pm, k = 1 - pm;
  const double de = U[8] / rhoMix - Y0 * t.enr[0] - Y1 * t.enr[1] - 0.5 * (v0 * v0 + v1 * v1 + v2 * v2);
  t.enr[0] += de;
  t.enr[1] += de;
  t.dP = 0.0f;
  t.dT = 0.0f;
  if (P.thermFntRlxFlag == 1) {
    const double avf[2] = {avf0, avf1}, rho[2] = {rho0, rho1};
    fnt_press_rlx(P, avf, rho, t.enr, &t.dP, &t.dT);
  }
  int abortFlag = 0;
  float minRes = 1e9f;
  int itr = 0;
  bool conv = false;
  while (!conv) {
    itr++;
    double r[2];
    t.res(P, x, &r[0], &r[1]);
    double e[2], rd[2][2];
    // perturb x[k] then x[k+2] (the non-primary unknowns in index order)
    e[0] = 1e-4 * x[k];
    x[k] += e[0];
    t.res(P, x, &rd[0][0], &rd[1][0]);
    x[k] -= e[0];
    e[1] = 1e-4 * x[k + 2];
    x[k + 2] += e[1];
    t.res(P, x, &rd[0][1], &rd[1][1]);
    x[k + 2] -= e[1];
    double J[2][2];
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) {
        J[i][j] = (rd[i][j] - r[i]) / e[j];
        if (__builtin_isnan(J[i][j])) abortFlag = 1;
      }
    const double det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
    double I00 = J[1][1] / det, I11 = J[0][0] / det, I10 = -J[1][0] / det, I01 = -J[0][1] / det;
    if (det == 0) abortFlag = 1;
    if (abortFlag == 0 && x[0] > 0.0 && x[1] > 0.0) {
      float chk = (float)sqrt(r[0] * r[0] + r[1] * r[1]);
      if (chk < minRes) {
        minRes = chk;
        for (int i = 0; i < 4; ++i) ax[i] = x[i];
      }
    }
    double dx0 = 0;
    dx0 += I00 * r[0];
    dx0 += I01 * r[1];
    x[k] = x[k] - urf * dx0;
    double dx1 = 0;
    dx1 += I10 * r[0];
    dx1 += I11 * r[1];
    x[k + 2] = x[k + 2] - urf * dx1;
    if (!(fabs(r[0]) > 1.e-2) && !(fabs(r[1]) > 1.e-2)) conv = true;
    if (abortFlag == 1 || itr > itrMax) {
      for (int i = 0; i < 4; ++i) x[i] = ax[i];
      conv = true;
    }
  }
  const double s1 = t.ar[k] / x[k];
  const double s2 = t.ar[k] * (x[2 + k] - t.enr[k]);
  x[pm] = (1 / (1 - s1)) * t.ar[pm];
  x[2 + pm] = t.enr[pm] - (1 / t.ar[pm]) * s2;
  U[0] = clampd(P, t.ar[0] / x[0]);
  U[6] = t.ar[0] * x[2];
  U[7] = t.ar[1] * x[3];
  return itr;
}

// Gibbs energy of chemRlx.cu; log10(T^g / (p + pinf)^(g - 1)) taken as
// g log10(T) - (g - 1) log10(p + pinf).
FK_HD inline double gibbs_dev(const fk_params &P, int k, double pr, double T) {
  const double et = k == 0 ? 0.0 : P.satConst[0] * fkmath::log10(pr) + P.satConst[1];
  const double h = P.Cp[k] * T + P.eta[k];
  const double s = (P.Cp[k] / P.gamma[k]) * (P.gamma[k] * fkmath::log10(T) - (P.gamma[k] - 1) * fkmath::log10(pr + P.pInf[k])) + et;
  return h - T * s;
}

struct ChemSys {
  int pm;
  double rhoMix, rEint;
  FK_HD void res(const fk_params &P, double x[6], double *r0, double *r1, double *r2) const {
    const int k = 1 - pm;
    const double s1 = x[k], s2 = x[k] * x[k + 2], s3 = x[k] * x[k + 2] * x[k + 4];
    x[pm] = 1 - s1;
    x[pm + 2] = (rhoMix - s2) / x[pm];
    x[pm + 4] = (rEint - s3) / (x[pm] * x[pm + 2]);
    double pr[2], te[2], gi[2];
    for (int q = 0; q < 2; q++) {
      pr[q] = eos_p(P, q, x[q + 2], x[q + 4]);
      te[q] = eos_T(P, q, x[q + 2], x[q + 4]);
      gi[q] = gibbs_dev(P, q, pr[q], te[q]);
    }
    *r0 = pr[0] - pr[1];
    *r1 = te[0] - te[1];
    *r2 = gi[0] - gi[1];
  }
};

FK_HD inline void chk_update(const fk_params &P, double xOld, double *x, double dx, int i, float urf, int *abortFlag,
                           double rho) {
  if (i == 0 || i == 1) {
    if (x[i] < 0)
      while (x[i] < P.chemLimit && urf > 0.0005) {
        urf = urf / 2;
        x[i] = xOld - urf * dx;
      }
    if (x[i] > 1)
      while (x[i] > 1.0 - P.chemLimit && urf > 0.0005) {
        urf = urf / 2;
        x[i] = xOld - urf * dx;
      }
    if (x[i] < 0) *abortFlag = 1;
    if (x[i] > 1) *abortFlag = 1;
  } else if (i == 2 || i == 3) {
    double ar = x[i - 2] * x[i];
    if (x[i] < 0)
      while (x[i] < 0 && urf > 0.0005) {
        urf = urf / 2;
        x[i] = xOld - urf * dx;
      }
    if (ar > rho)
      while (ar > rho && urf > 0.0005) {
        urf = urf / 2;
        x[i] = xOld - urf * dx;
        ar = x[i - 2] * x[i];
      }
    if (x[i] < 0) *abortFlag = 1;
    if (ar > rho) *abortFlag = 1;
  }
}

FK_HD inline int chem_dev(const fk_params &P, double *U) {
  const int itrMax = P.chemItrMax;
  const double urf = P.chemUrf;
  double x[6], ax[6];
  x[0] = clampd(P, U[0]);
  x[1] = 1 - x[0];
  x[2] = U[1] / x[0];
  x[3] = U[2] / x[1];
  x[4] = U[6] / (x[0] * x[2]);
  x[5] = U[7] / (x[1] * x[3]);
  for (int i = 0; i < 6; ++i) ax[i] = x[i];
  ChemSys ch;
  ch.rhoMix = x[0] * x[2] + x[1] * x[3];
  const double u = U[3] / ch.rhoMix, v = U[4] / ch.rhoMix, w = U[5] / ch.rhoMix;
  ch.rEint = U[8] - 0.5 * ch.rhoMix * (u * u + v * v + w * w);
  ch.pm = (x[1] > x[0]) ? 1 : 0;
  const int pm = ch.pm;
  int idx[3];  // non-primary unknowns in index order
  {
    int q = 0;
    for (int j = 0; j < 6; ++j)
      if (j != pm && j != pm + 2 && j != pm + 4) idx[q++] = j;
  }
  int abortFlag = 0;
  float minRes = 1e9f;
  int itr = 0;
  bool conv = false;
  while (!conv) {
    itr++;
    double r[3];
    ch.res(P, x, &r[0], &r[1], &r[2]);
    double e[3], rd[3][3];
    for (int q = 0; q < 3; ++q) {
      const int j = idx[q];
      e[q] = fabs(x[j]) < 1e-9 ? 1e-10 : 1.e-5 * x[j];
      x[j] += e[q];
      ch.res(P, x, &rd[0][q], &rd[1][q], &rd[2][q]);
      x[j] -= e[q];
    }
    double J[3][3], I[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        J[i][j] = (rd[i][j] - r[i]) / e[j];
        if (__builtin_isnan(J[i][j])) abortFlag = 1;
      }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (det == 0) abortFlag = 1;
    const double id = 1.0 / det;
    I[0][0] = id * (J[1][1] * J[2][2] - J[1][2] * J[2][1]);
    I[0][1] = -id * (J[0][1] * J[2][2] - J[0][2] * J[2][1]);
    I[0][2] = id * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
    I[1][0] = -id * (J[1][0] * J[2][2] - J[1][2] * J[2][0]);
    I[1][1] = id * (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
    I[1][2] = -id * (J[0][0] * J[1][2] - J[0][2] * J[1][0]);
    I[2][0] = id * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    I[2][1] = -id * (J[0][0] * J[2][1] - J[0][1] * J[2][0]);
    I[2][2] = id * (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
    if (abortFlag == 0 && x[0] > 0.0 && x[0] < 1.0 && x[1] > 0.0 && x[1] < 1.0 && x[2] > 0.0 && x[3] > 0.0) {
      float chk = (float)sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      if (chk < minRes) {
        minRes = chk;
        for (int i = 0; i < 6; ++i) ax[i] = x[i];
      }
    }
    int abortLocal = abortFlag;
    for (int q = 0; q < 3; ++q) {
      const int i = idx[q];
      double dx = 0;
      for (int j = 0; j < 3; j++) dx += I[q][j] * r[j];
      x[i] = x[i] - urf * dx;
      chk_update(P, x[i] + urf * dx, x, dx, i, (float)urf, &abortLocal, ch.rhoMix);
    }
    if (!(fabs(r[0]) > 1.e-1) && !(fabs(r[1]) > 1.e-1) && !(fabs(r[2]) > 1.e-1)) conv = true;
    if (abortLocal == 1 || itr > itrMax) {
      for (int i = 0; i < 6; ++i) x[i] = ax[i];
      conv = true;
    }
  }
  const int k = 1 - pm;
  const double s1 = x[k], s2 = x[k] * x[k + 2], s3 = x[k] * x[k + 2] * x[k + 4];
  x[pm] = 1 - s1;
  x[pm + 2] = (ch.rhoMix - s2) / x[pm];
  x[pm + 4] = (ch.rEint - s3) / (x[pm] * x[pm + 2]);
  U[0] = clampd(P, x[0]);
  U[1] = x[0] * x[2];
  U[2] = x[1] * x[3];
  U[6] = x[0] * x[2] * x[4];
  U[7] = x[1] * x[3] * x[5];
  U[9] = x[1] * x[3];
  return itr;
}


// ---------------- surrogate path: exact classification + Newton polish ----------------
// Saturation function f(T; p) of relaxation.cu:18-31.  With
// dT f = (C T ln10 - B) / (T^2 ln10) > 0 (B <= 0, C >= 0, checked on the
// host) f is increasing in T, so T_L > Tsat(p_L) <=> f(T_L; p_L) > 0: the
// classification of partitionRelaxation without the Tsat iteration.
FK_HD inline double sat_fun(const fk_params &P, double Pr, double T) {
  const double g1 = P.gamma[0], g2 = P.gamma[1], Cp1 = P.Cp[0], Cp2 = P.Cp[1];
  const double Cv1 = Cp1 / g1, Cv2 = Cp2 / g2;
  const double et2 = P.satConst[0] * fkmath::log10(Pr) + P.satConst[1];
  const double As = (Cp1 - Cp2 + et2) / (Cp2 - Cv2);
  const double Bs = (P.eta[0] - P.eta[1]) / (Cp2 - Cv2);
  const double Cs = (Cp2 - Cp1) / (Cp2 - Cv2);
  const double Ds = (Cp1 - Cv1) / (Cp2 - Cv2);
  return As + Bs / T + Cs * fkmath::log10(T) + Ds * fkmath::log10(Pr + P.pInf[0]) - fkmath::log10(Pr);
}

// Sign of f(T; p) (> 0: superheated liquid), evaluated in FP32 first and in
// FP64 only when |f| < 1e-3: the FP32 terms (|term| < 100) carry errors
// below 1e-4, and df/dT ~ 1e-2 / K, so only states within ~0.1 K of
// saturation need the FP64 evaluation.
FK_HD inline bool sat_superheated(const fk_params &P, double Pr, double T) {
  const double g1 = P.gamma[0], g2 = P.gamma[1], Cp1 = P.Cp[0], Cp2 = P.Cp[1];
  const double Cv1 = Cp1 / g1, Cv2 = Cp2 / g2;
  const double iden = 1 / (Cp2 - Cv2);
  const float pr = (float)Pr, tf = (float)T;
  const float et2 = (float)P.satConst[0] * log10f(pr) + (float)P.satConst[1];
  const float As = (float)(Cp1 - Cp2) * (float)iden + et2 * (float)iden;
  const float Bs = (float)((P.eta[0] - P.eta[1]) * iden);
  const float Cs = (float)((Cp2 - Cp1) * iden);
  const float Ds = (float)((Cp1 - Cv1) * iden);
  const float f = As + Bs / tf + Cs * log10f(tf) + Ds * log10f(pr + (float)P.pInf[0]) - log10f(pr);
  if (fabsf(f) > 1e-3f && __builtin_isfinite(f)) return f > 0;
  return sat_fun(P, Pr, T) > 0;
}

// Thermal (p-T) equilibrium problem of one cell as thermalRlx sets it up
// (thermRlx.cu:4-40, 205-210): partial densities ar_k, phasic energies
// shifted by the defect de of the mixture internal energy, primary phase pm
// (larger volume fraction) and the solved-for phase k = 1 - pm.  The unknowns
// are (rho_k, e_k); the primary phase follows from the volume and energy
// closure (thermRlx.cu:75-81).
struct ThermProb {
  double ar[2], enr[2];
  int pm, k;
};

FK_HD inline void therm_setup(const fk_params &P, const double *U, ThermProb &t) {
  const double a0 = clampd(P, U[0]), a1 = 1 - a0;
  t.ar[0] = U[1];
  t.ar[1] = U[2];
  t.enr[0] = U[6] / U[1];
  t.enr[1] = U[7] / U[2];
  const double ir = 1 / (U[1] + U[2]);
  const double ke = 0.5 * (U[3] * U[3] + U[4] * U[4] + U[5] * U[5]) * ir * ir;
  // de = rho E / rho - Y_L e_L - Y_G e_G - |u|^2 / 2 (Y_k e_k = U[6 + k] / rho)
  const double de = (U[8] - U[6] - U[7]) * ir - ke;
  t.enr[0] += de;
  t.enr[1] += de;
  t.pm = (a1 > a0) ? 1 : 0;
  t.k = 1 - t.pm;
}

// Cold start of the thermal problem: the cell's own (rho_k, e_k + de).
FK_HD inline void therm_cold(const fk_params &P, const double *U, const ThermProb &t, double &xr, double &xe) {
  const double a0 = clampd(P, U[0]);
  xr = t.k == 0 ? U[1] / a0 : U[2] / (1 - a0);
  xe = t.enr[t.k];
}

// A phase state inside the stiffened gas's domain: T > 0 and p + pinf > 0
// (a real sound speed).  With T = (p + pinf) / (Cv (g - 1) rho) the two hold
// together exactly when rho > 0 and p + pinf > 0, which is tested (no
// division).  The equilibrium equations also have roots outside the domain;
// the surrogate's solves reject those and leave the cell to the reference
// solve.
FK_HD inline bool eos_domain(const fk_params &P, int k, double rho, double e) {
  return rho > 0 && eos_p(P, k, rho, e) + P.pInf[k] > 0;
}

// Newton iterations with the analytic Jacobian on (rho_k, e_k) from (xr, xe)
// until |d rho_k| <= tol |rho_k| and |d e_k| <= tol |e_k|.  Returns the
// iterations (> 0), or -1 (no convergence within itmax, singular Jacobian,
// an iterate outside rho_k > ar_k, or a root outside either phase's EOS
// domain); xr, xe hold the last iterate.
// Residuals p_k - p_m and T_k - T_m (signs as thermRlx.cu:95-103) with
//   p = (g - 1) rho (e - eta) - g pinf,  T = (e - eta - pinf / rho) / Cv,
// the stiffened-gas EOS of eqos.h rearranged so that an iteration takes
// three divisions.
FK_HD inline int therm_newton(const fk_params &P, const ThermProb &t, double &xr, double &xe, double tol,
                              int itmax) {
  const int pm = t.pm, k = t.k;
  if (!(xr > t.ar[k]) || !__builtin_isfinite(xe)) return -1;
  const double sk = k == 0 ? 1.0 : -1.0;  // sign of phase k in r = q_0 - q_1
  const double gk = P.gamma[k], gm = P.gamma[pm];
  const double ick = gk / P.Cp[k], icm = gm / P.Cp[pm];  // 1 / Cv
  const double ark = t.ar[k], iarm = 1 / t.ar[pm];
  const double ratio = ark * iarm;  // d e_m / d e_k = -ratio
  const double ek0 = t.enr[k], em0 = t.enr[pm];
  const double etk = P.eta[k], etm = P.eta[pm], pik = P.pInf[k], pim = P.pInf[pm];
  for (int it = 1; it <= itmax; ++it) {
    const double ixr = 1 / xr;
    const double q = ark * ixr;  // alpha_k
    const double omq = 1 - q;
    const double rm = t.ar[pm] / omq;
    const double irm = omq * iarm;
    const double em = em0 - ratio * (xe - ek0);
    if (!(rm > 0)) return -1;
    const double pk = (gk - 1) * xr * (xe - etk) - gk * pik;
    const double pmm = (gm - 1) * rm * (em - etm) - gm * pim;
    const double tk = (xe - etk - pik * ixr) * ick;
    const double tm = (em - etm - pim * irm) * icm;
    const double f0 = sk * (pk - pmm), f1 = sk * (tk - tm);
    const double drm = -rm * rm * iarm * q * ixr;  // d rho_m / d rho_k
    const double J00 = sk * ((gk - 1) * (xe - etk) - (gm - 1) * (em - etm) * drm);
    const double J01 = sk * ((gk - 1) * xr + (gm - 1) * rm * ratio);
    const double J10 = sk * (pik * ick * ixr * ixr - pim * icm * irm * irm * drm);
    const double J11 = sk * (ick + ratio * icm);
    const double det = J00 * J11 - J01 * J10;
    if (!(fabs(det) > 0) || !__builtin_isfinite(det)) return -1;
    const double idet = 1 / det;
    const double dr = (J11 * f0 - J01 * f1) * idet;
    const double dee = (J00 * f1 - J10 * f0) * idet;
    xr -= dr;
    xe -= dee;
    if (!(xr > t.ar[k]) || !__builtin_isfinite(xe)) return -1;
    if (fabs(dr) <= tol * fabs(xr) && fabs(dee) <= tol * fabs(xe)) {
      const double rmf = t.ar[pm] * xr / (xr - ark), emf = em0 - ratio * (xe - ek0);  // rho_m, e_m
      return (eos_domain(P, k, xr, xe) && eos_domain(P, pm, rmf, emf)) ? it : -1;
    }
  }
  return -1;
}

// Writes alpha_L, alpha_L rho_L e_L, alpha_G rho_G e_G of the solution
// (updateConsVec, thermRlx.cu:172-187).
FK_HD inline void therm_commit(const fk_params &P, const ThermProb &t, double xr, double xe, double *U) {
  const int pm = t.pm, k = t.k;
  const double s1 = t.ar[k] / xr, s2 = t.ar[k] * (xe - t.enr[k]);
  double x[4];
  x[k] = xr;
  x[2 + k] = xe;
  x[pm] = (1 / (1 - s1)) * t.ar[pm];
  x[2 + pm] = t.enr[pm] - (1 / t.ar[pm]) * s2;
  U[0] = clampd(P, t.ar[0] / x[0]);
  U[6] = t.ar[0] * x[2];
  U[7] = t.ar[1] * x[3];
}

// Thermal equilibrium by Newton polish from the warm start (rho_k, e_k).
// Returns iterations (> 0) or -1 (the caller falls back).
FK_HD inline int therm_polish(const fk_params &P, double *U, double rk0, double ek0, double tol, int itmax) {
  ThermProb t;
  therm_setup(P, U, t);
  double xr = rk0, xe = ek0;
  const int it = therm_newton(P, t, xr, xe, tol, itmax);
  if (it > 0) therm_commit(P, t, xr, xe, U);
  return it;
}

// Stage 1 of the surrogate relaxation for one cell, in place on u: the exact
// classification (chemical when alpha_L > chemLimit and the liquid is
// superheated at its own pressure), and for a thermal cell at most
// kColdIters analytic-Jacobian Newton iterations from the cell's own state,
// committed when converged.  Returns 1 when the cell goes to the work list
// of stage 2 (chemical, or not converged); *chem is the class and *iters the
// iterations (kColdIters for a failed cold start, 0 for a chemical cell).
constexpr int kColdIters = 4;
FK_HD inline int screen_cell(const fk_params &P, double *u, int *chem, int *iters) {
  int ch = 0;
  if (P.chemFlag == 1) {
    const double avf = clampd(P, u[0]);
    const double rho = u[1] / avf, enr = u[6] / (avf * rho);
    const double tL = eos_T(P, 0, rho, enr), pL = eos_p(P, 0, rho, enr);
    ch = avf > P.chemLimit && sat_superheated(P, pL, tL);
  }
  *chem = ch;
  if (ch) {
    *iters = 0;
    return 1;
  }
  ThermProb t;
  therm_setup(P, u, t);
  double xr, xe;
  therm_cold(P, u, t, xr, xe);
  const int it = therm_newton(P, t, xr, xe, P.polishTol, kColdIters);
  if (it > 0) {
    therm_commit(P, t, xr, xe, u);
    *iters = it;
    return 0;
  }
  *iters = kColdIters;
  return 1;
}

// Plain relaxation of one cell, in place on u (relaxation.cu:40-75 with
// thermRlx.cu:189-235): with chemFlag, a cell whose liquid is above Tsat(p_L)
// and alpha_L > chemLimit is chemical and left for the chemical solve
// (returns 1); every other cell gets the reference thermal Newton solve.
// *iters and *tsat_iters count the Newton and Tsat iterations.
FK_HD inline int plain_cell(const fk_params &P, double *u, int *iters, int *tsat_iters) {
  int chem = 0, ts = 0;
  if (P.chemFlag == 1) {
    const double avf = clampd(P, u[0]);
    const double rho = u[1] / avf;
    const double enr = u[6] / (avf * rho);
    const double tL = eos_T(P, 0, rho, enr);
    const double pL = eos_p(P, 0, rho, enr);
    const double Ts = tsat_dev(P, pL, tL, P.satTempUrf, &ts);
    chem = (tL > Ts && avf > P.chemLimit) ? 1 : 0;
  }
  *tsat_iters = ts;
  *iters = chem ? 0 : therm_dev(P, u);
  return chem;
}

// Chemical (p-T-g) equilibrium of chemRlx (chemRlx.cu:52-76, 195-215): the
// same equations (equal phasic pressures, temperatures and Gibbs energies of
// gibbsFunction, mass and internal energy of the cell conserved), solved in
// the unknowns (ln p, T) common to both phases.  Each phase's density and
// energy follow from (p, T) through the stiffened-gas EOS,
//   rho_k = (p + pinf_k) / ((g_k - 1) Cv_k T),
//   rho_k e_k = (p + g_k pinf_k) / (g_k - 1) + eta_k rho_k,
// the volume fraction from the mass balance,
//   alpha_L = (rho - rho_G) / (rho_L - rho_G),
// and the residuals are the internal energy balance
//   R1 = alpha_L rho_L e_L + alpha_G rho_G e_G - rho e
// and the Gibbs difference R2 = g_L(p, T) - g_G(p, T), with the analytic
// Jacobian
//   dg_k/dT = Cp_k - s_k - Cp_k / ln 10,
//   dg_k/dp = T Cv_k (g_k - 1) / ((p + pinf_k) ln 10) - T d(et_k)/dp,
// (s_k the entropy of gibbsFunction, et_G = satConst0 log10 p + satConst1).
// The reference's unknowns (alpha_k, rho_k, e_k) of the secondary phase pass
// a relative error of its energy to the primary phase amplified by the mass
// ratio (10^3 .. 10^5 for liquid in vapour), which makes vapour pressures
// negative and its Newton steps leave the domain from all but very close
// starts; in (ln p, T) both phases stay on their EOS.  Steps are limited to
// |d ln p| <= 1 and |dT| <= 0.2 T.  Converged when |d ln p| <= tol and
// |dT| <= tol T; the state is then committed as chemRlx commits it (the
// secondary phase from (p, T), the primary phase from the mass and energy
// closure, so both are conserved to rounding).  Returns the iterations, or
// -1 (singular or non-finite step, T <= 0, no convergence within itmax,
// alpha_L outside (0, 1) -- no two-phase equilibrium at this (rho, e) -- or a
// phase outside its EOS domain); the caller then falls back to the
// reference solve.
FK_HD inline int chem_polish(const fk_params &P, double *U, double T0, double p0, double tol, int itmax) {
  constexpr double kLn10 = 2.302585092994045684;
  const double a0 = clampd(P, U[0]);
  const double rho = a0 * (U[1] / a0) + (1 - a0) * (U[2] / (1 - a0));
  const double u = U[3] / rho, v = U[4] / rho, w = U[5] / rho;
  const double rE = U[8] - 0.5 * rho * (u * u + v * v + w * w);
  if (!(T0 > 0) || !__builtin_isfinite(T0) || !__builtin_isfinite(p0)) return -1;
  double lp = fkmath::log(fmax(p0, 1.0)), T = T0;
  for (int it = 1; it <= itmax; ++it) {
    const double p = fkmath::exp(lp);
    double r[2], drp[2], drT[2], E[2], dEp[2], dET[2], g[2], gp[2], gT[2];
    for (int q = 0; q < 2; ++q) {
      const double gm = P.gamma[q], pi = P.pInf[q], cv = P.Cp[q] / gm;
      r[q] = (p + pi) / (cv * (gm - 1) * T);
      drp[q] = r[q] / (p + pi);
      drT[q] = -r[q] / T;
      E[q] = (p + gm * pi) / (gm - 1) + P.eta[q] * r[q];
      dEp[q] = 1 / (gm - 1) + P.eta[q] * drp[q];
      dET[q] = P.eta[q] * drT[q];
      const double et = q == 0 ? 0.0 : P.satConst[0] * fkmath::log10(p) + P.satConst[1];
      const double s = cv * (gm * fkmath::log10(T) - (gm - 1) * fkmath::log10(p + pi)) + et;
      g[q] = P.Cp[q] * T + P.eta[q] - T * s;
      gT[q] = P.Cp[q] - s - P.Cp[q] / kLn10;
      const double dsp = -cv * (gm - 1) / ((p + pi) * kLn10) + (q == 0 ? 0.0 : P.satConst[0] / (p * kLn10));
      gp[q] = -T * dsp;
    }
    const double D = r[0] - r[1];
    const double al = (rho - r[1]) / D;
    const double dap = (-drp[1] - al * (drp[0] - drp[1])) / D, daT = (-drT[1] - al * (drT[0] - drT[1])) / D;
    const double dE = E[0] - E[1];
    const double R1 = E[1] + al * dE - rE, R2 = g[0] - g[1];
    const double J00 = p * (dEp[1] + dap * dE + al * (dEp[0] - dEp[1]));  // d R1 / d ln p
    const double J01 = dET[1] + daT * dE + al * (dET[0] - dET[1]);        // d R1 / d T
    const double J10 = p * (gp[0] - gp[1]);
    const double J11 = gT[0] - gT[1];
    const double det = J00 * J11 - J01 * J10;
    if (!(fabs(det) > 0) || !__builtin_isfinite(det)) return -1;
    double dl = (J11 * R1 - J01 * R2) / det, dT = (J00 * R2 - J10 * R1) / det;
    if (!__builtin_isfinite(dl) || !__builtin_isfinite(dT)) return -1;
    const double lim = fmax(fabs(dl), fabs(dT) / (0.2 * T));
    if (lim > 1) {
      dl /= lim;
      dT /= lim;
    }
    lp -= dl;
    T -= dT;
    if (!(T > 0)) return -1;
    if (fabs(dl) <= tol && fabs(dT) <= tol * T) {
      const double pf = fkmath::exp(lp);
      const double rl = (pf + P.pInf[0]) / ((P.Cp[0] / P.gamma[0]) * (P.gamma[0] - 1) * T);
      const double rg = (pf + P.pInf[1]) / ((P.Cp[1] / P.gamma[1]) * (P.gamma[1] - 1) * T);
      const double aL = (rho - rg) / (rl - rg);
      if (!(aL > 0 && aL < 1)) return -1;
      const int pm = (1 - a0 > a0) ? 1 : 0, k = 1 - pm;
      double x[6];
      x[k] = k == 0 ? aL : 1 - aL;
      x[k + 2] = k == 0 ? rl : rg;
      x[k + 4] = ((pf + P.gamma[k] * P.pInf[k]) / (P.gamma[k] - 1) + P.eta[k] * x[k + 2]) / x[k + 2];
      const double s1 = x[k], s2 = x[k] * x[k + 2], s3 = x[k] * x[k + 2] * x[k + 4];
      x[pm] = 1 - s1;
      x[pm + 2] = (rho - s2) / x[pm];
      x[pm + 4] = (rE - s3) / (x[pm] * x[pm + 2]);
      if (!(x[pm + 2] > 0)) return -1;
      if (!eos_domain(P, 0, x[2], x[4]) || !eos_domain(P, 1, x[3], x[5])) return -1;
      U[0] = clampd(P, x[0]);
      U[1] = x[0] * x[2];
      U[2] = x[1] * x[3];
      U[6] = x[0] * x[2] * x[4];
      U[7] = x[1] * x[3] * x[5];
      U[9] = x[1] * x[3];
      return it;
    }
  }
  return -1;
}

// Cold start of the chemical polish: the liquid's own temperature and
// pressure (at least 1 kPa).  Chemical cells are those whose liquid is
// superheated, so (T_L, p_L) lies on the liquid side of the saturation curve
// the solution sits on.
FK_HD inline void chem_cold(const fk_params &P, const double *U, double &T0, double &p0) {
  const double a = clampd(P, U[0]);
  const double rl = U[1] / a, el = U[6] / U[1];
  T0 = eos_T(P, 0, rl, el);
  p0 = fmax(eos_p(P, 0, rl, el), 1e3);
}

#if defined(__HIPCC__)
// Append of cell c to the surrogate relaxation's work list, wave-aggregated:
// chemical cells from the front (their count in cnt[5]), thermal cells from
// the back (count in cnt[6], k-th entry at list[cap - 1 - k]); entries are
// (c << 1) | chem.  Stage 2 then runs the network on the waves of chemical
// entries only.  The list order does not affect results: every cell is
// solved independently.
__device__ inline void list_push(bool push, int chem, int64_t c, uint32_t *__restrict__ list, int64_t cap,
                                 unsigned long long *__restrict__ cnt) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const bool mine = push && (side == 0) == (chem != 0);
    const unsigned long long m = __ballot(mine);
    if (!m) continue;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long b = 0;
    if (l == leader) b = atomicAdd(&cnt[5 + side], (unsigned long long)__popcll(m));
    b = __shfl(b, leader, 64);
    if (mine) {
      const int64_t k = (int64_t)(b + (unsigned long long)__popcll(m & ((1ull << l) - 1ull)));
      list[side == 0 ? k : cap - 1 - k] = (uint32_t)((c << 1) | chem);
    }
  }
}
#endif

// Surrogate features of a conserved state: (log rho, e, Y_L) with e the
// mixture internal energy per unit mass and Y_L the liquid mass fraction.
// (rho, e, Y_L) fix the thermal equilibrium; (rho, e) fix the chemical one.
FK_HD inline void mlp_features(const fk_params &P, const double *U, float *x) {
  (void)P;
  const double rho = U[1] + U[2];
  const double ke = 0.5 * (U[3] * U[3] + U[4] * U[4] + U[5] * U[5]) / rho;
  x[0] = (float)log(rho);
  x[1] = (float)((U[8] - ke) / rho);
  x[2] = (float)(U[1] / rho);
}

// Targets of a relaxed state: (T_L, asinh(p_L / 1e4), alpha_L), or NaN when
// the state is not at p-T equilibrium (|p_L - p_G| > 1e-6 max(|p_L|, 1e3) Pa
// or |T_L - T_G| > 1e-3 K): the plain solve stops at loose absolute
// tolerances or aborts to its best iterate, and such labels are not a
// function of the features.  Returns whether the targets are valid.
FK_HD inline bool mlp_targets(const fk_params &P, const double *U, float *y) {
  const double a = clampd(P, U[0]);
  const double rl = U[1] / a, el = U[6] / U[1];
  const double rg = U[2] / (1 - a), eg = U[7] / U[2];
  const double pl = eos_p(P, 0, rl, el), pg = eos_p(P, 1, rg, eg);
  const double tl = eos_T(P, 0, rl, el), tg = eos_T(P, 1, rg, eg);
  const bool ok = fabs(pl - pg) <= 1e-6 * fmax(fabs(pl), 1e3) && fabs(tl - tg) <= 1e-3;
  y[0] = ok ? (float)tl : NAN;
  y[1] = ok ? (float)asinh(pl / 1e4) : NAN;
  y[2] = ok ? (float)a : NAN;
  return ok;
}

// ---------------- relaxation counters without same-address atomics ----------------
// Every wave adding its counts to one global counter serialises ~10^6
// atomics on one L2 channel (several ms at 48M cells).  Instead each
// 256-thread workgroup writes its K sums to part[block][K] and one small
// kernel adds the column sums to the counters.

template <int K>
__device__ inline void block_partials(unsigned long long (&v)[K], unsigned long long *__restrict__ part) {
  __shared__ unsigned long long sh[4][K];
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[w][k] = v[k];
  __syncthreads();
  if ((int)threadIdx.x < K) {
    const int k = threadIdx.x;
    part[(int64_t)blockIdx.x * K + k] = sh[0][k] + sh[1][k] + sh[2][k] + sh[3][k];
  }
}

// counts[k] += sum over blocks of part[b][k], k < K.  Each 256-thread block
// sums a contiguous chunk of the partials and adds its K totals with one
// atomic each (a few hundred atomics in all; integer sums, so the result
// does not depend on their order).
template <int K>
__global__ void __launch_bounds__(256) reduce_partials(const unsigned long long *__restrict__ part, int64_t nb,
                                                       unsigned long long *__restrict__ counts) {
  __shared__ unsigned long long sh[4][K];
  const int64_t per = (nb + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = (int64_t)blockIdx.x * per, b1 = b0 + per < nb ? b0 + per : nb;
  unsigned long long v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0;
  for (int64_t b = b0 + threadIdx.x; b < b1; b += 256)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += part[b * K + k];
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if ((int)threadIdx.x < K) {
    const unsigned long long t = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
    if (t) atomicAdd(&counts[threadIdx.x], t);
  }
}

// grid of reduce_partials for nb partials
inline unsigned reduce_grid(int64_t nb) {
  const int64_t g = (nb + 1023) / 1024;
  return (unsigned)(g < 1 ? 1 : (g > 512 ? 512 : g));
}

}  // namespace fk
