// Device-side physics of the finite-volume step, templated on the
// arithmetic type T of reconstruction / flux / WALE (double on the FP64
// path, float on the mixed path).  EOS inversion (conserved -> primitive)
// always runs in double.
//
// Every expression keeps the association order of the reference kernels
// (cited per function) so that the FP64 instantiation reproduces a serial
// FP64 evaluation of the same formulas bit for bit; HIP device code is built
// with -ffp-contract=off for that reason.  Face-frame rotations use only the
// unit normal (star velocity u + (S* - u.n) n, slip ghost u - 2 (u.n) n,
// wall ghost -u), which equals the reference's (n, t1, t2) rotation up to
// rounding.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/fooska.h"

namespace fk {

// Material/solver constants in the arithmetic type of the caller.
template <typename T>
struct Mat {
  T g[2], pi[2], cp[2], eta[2];
  T avfMin, chemLimit, mu[2], Pr0;
  T inP, inT;  // inlet ghost pressure and temperature
  // Single-phase, FP32 (mixed) path only: the FP32 face pressures hold
  // p - pr (fk_params pRef), the absolute pressure is formed where the flux
  // needs it.  Never read on the FP64 path.
  T pr;
  int woods, hllcToro;
};

template <typename T>
__host__ __device__ inline Mat<T> make_mat(const fk_params &p) {
  Mat<T> m;
  for (int k = 0; k < 2; ++k) {
    m.g[k] = (T)p.gamma[k];
    m.pi[k] = (T)p.pInf[k];
    m.cp[k] = (T)p.Cp[k];
    m.eta[k] = (T)p.eta[k];
    m.mu[k] = (T)p.dynVisc[k];
  }
  m.avfMin = (T)p.avfMin;
  m.chemLimit = (T)p.chemLimit;
  m.Pr0 = (T)p.Pr[0];
  m.inP = (T)p.inletP;
  m.inT = (T)p.inletT;
  m.pr = (T)p.pRef;
  m.woods = p.woodsFlag;
  m.hllcToro = p.hllcEnergy;
  return m;
}

// ---------------- EOS (DP eqos.h:1-23, SP eqos.h:2-23) ----------------
// two-phase stiffened gas with energy offset eta, phase k
template <typename T>
__device__ __forceinline__ T p_of(const Mat<T> &m, int k, T rho, T e) {
  return (m.g[k] - 1) * rho * e - m.g[k] * m.pi[k] - (m.g[k] - 1) * m.eta[k] * rho;
}
template <typename T>
__device__ __forceinline__ T e_of(const Mat<T> &m, int k, T p, T rho) {
  return (p + m.g[k] * m.pi[k] + (m.g[k] - 1) * m.eta[k] * rho) / (rho * (m.g[k] - 1));
}
template <typename T>
__device__ __forceinline__ T c_of(const Mat<T> &m, int k, T rho, T p) {
  return sqrt((m.g[k] * (p + m.pi[k])) / rho);
}
template <typename T>
__device__ __forceinline__ T T_of(const Mat<T> &m, int k, T rho, T e) {
  T p = (m.g[k] - 1) * rho * e - m.g[k] * m.pi[k] - (m.g[k] - 1) * m.eta[k] * rho;
  return (p + m.pi[k]) / ((m.cp[k] / m.g[k]) * (m.g[k] - 1) * rho);
}
template <typename T>
__device__ __forceinline__ T rho_of(const Mat<T> &m, int k, T p, T temp) {
  return (p + m.pi[k]) / ((m.cp[k] / m.g[k]) * temp * (m.g[k] - 1));
}
// single-phase (no eta)
template <typename T>
__device__ __forceinline__ T p1_of(const Mat<T> &m, T rho, T e) {
  return (m.g[0] - 1) * rho * e - m.g[0] * m.pi[0];
}
template <typename T>
__device__ __forceinline__ T e1_of(const Mat<T> &m, T p, T rho) {
  return (p + m.g[0] * m.pi[0]) / (rho * (m.g[0] - 1));
}

template <typename T>
__device__ __forceinline__ T clamp_avf(const Mat<T> &m, T a) {
  return fmin(fmax(m.avfMin / 100, a), (T)1.0 - m.avfMin / 100);
}

// ---------------- conserved -> primitive (double) ----------------
// Copies the N doubles of one AoS record into registers with 16-byte loads
// when N is even (records of 10 doubles start 16-byte aligned): half the
// load instructions of element-wise loads, and each covers more of the
// record's cache lines while they are resident.
template <int N>
__device__ __forceinline__ void load_rec(const double *__restrict__ p, double (&u)[N]) {
  if constexpr (N % 2 == 0) {
    const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
      const double2 t = q[i];
      u[2 * i] = t.x;
      u[2 * i + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = p[i];
  }
}

// Stores N doubles of one AoS record with 16-byte stores when N is even.
template <int N>
__device__ __forceinline__ void store_rec(double *__restrict__ p, const double (&u)[N]) {
  if constexpr (N % 2 == 0) {
    double2 *q = reinterpret_cast<double2 *>(p);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) q[i] = make_double2(u[2 * i], u[2 * i + 1]);
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) p[i] = u[i];
  }
}

// two-phase: DP cons2prim.h:7-22; single-phase: SP cons2prim.h:2-10
template <int NEQ>
__device__ __forceinline__ void cons2prim(const Mat<double> &m, const double *U, double *s) {
  if constexpr (NEQ == 5) {
    s[0] = U[0];
    s[1] = U[1] / s[0];
    s[2] = U[2] / s[0];
    s[3] = U[3] / s[0];
    s[4] = U[4] / s[0] - 0.5 * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3]);
    s[4] = p1_of(m, s[0], s[4]);
  } else {
    s[0] = clamp_avf(m, U[0]);
    s[1] = U[1] / s[0];
    s[2] = U[2] / (1 - s[0]);
    double rm = s[0] * s[1] + (1 - s[0]) * s[2];
    s[3] = U[3] / rm;
    s[4] = U[4] / rm;
    s[5] = U[5] / rm;
    double eL = U[6] / (s[0] * s[1]);
    double eG = U[7] / ((1 - s[0]) * s[2]);
    s[6] = p_of(m, 0, s[1], eL);
    s[7] = p_of(m, 1, s[2], eG);
    s[8] = U[9];
    s[9] = rm;
  }
}

// FP64 reciprocal: hardware estimate plus one Newton step (relative error
// near 1e-16, not correctly rounded)
__device__ __forceinline__ double rcp_nr(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

// The primitive variables of a state rounded to FP32: the first-order states
// of the mixed path.  cons2prim's quantities from reciprocals (3 instead of 7
// FP64 divisions two-phase, 1 instead of 4 single-phase), still in FP64 (the
// pressures cancel against gamma pInf): they differ from cons2prim by a few
// FP64 ulps, far below the FP32 rounding, and keep the register footprint of
// the kernels that form them small.  rho_m = U1 + U2 (= aL rhoL + aG rhoG) and
// rho_k e_k = U_{6,7} / a_k.
template <int NEQ>
__device__ __forceinline__ void prims_f32(const Mat<double> &m, const double *__restrict__ U, float *s) {
  if constexpr (NEQ == 5) {
    const double ir = rcp_nr(U[0]);
    const double u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
    const double e = U[4] * ir - 0.5 * (u * u + v * v + w * w);
    s[0] = (float)U[0];
    s[1] = (float)u;
    s[2] = (float)v;
    s[3] = (float)w;
    s[4] = (float)(p1_of(m, U[0], e) - m.pr);  // deviation from pRef (exact in FP32)
  } else {
    const double a = clamp_avf(m, U[0]);
    const double ia = rcp_nr(a), ig = rcp_nr(1.0 - a);
    const double rl = U[1] * ia, rg = U[2] * ig;
    const double rm = U[1] + U[2], irm = rcp_nr(rm);
    s[0] = (float)a;
    s[1] = (float)rl;
    s[2] = (float)rg;
    s[3] = (float)(U[3] * irm);
    s[4] = (float)(U[4] * irm);
    s[5] = (float)(U[5] * irm);
    s[6] = (float)((m.g[0] - 1) * (U[6] * ia) - m.g[0] * m.pi[0] - (m.g[0] - 1) * m.eta[0] * rl);
    s[7] = (float)((m.g[1] - 1) * (U[7] * ig) - m.g[1] * m.pi[1] - (m.g[1] - 1) * m.eta[1] * rg);
    s[8] = (float)U[9];
    s[9] = (float)rm;
  }
}

// FP32 primitive record of a state held in registers, written by the kernels
// that produce a new state so that the next pass G gathers 4 NEQ bytes per
// neighbour instead of 8 NEQ (mixed precision); the same values as pass G
// forms from U itself.
template <int NEQ>
__device__ __forceinline__ void store_prims(const Mat<double> &m, const double *u, float *__restrict__ W) {
  float t[NEQ];
  prims_f32<NEQ>(m, u, t);
#pragma unroll
  for (int v = 0; v < NEQ; ++v) W[v] = t[v];
}

// ---------------- boundary ghost (DP bc.h, SP bc.h) ----------------
template <int NEQ, typename T>
__device__ __forceinline__ void ghost(const Mat<T> &m, int bc, const T n[3], T *s) {
  constexpr int iu = NEQ == 5 ? 1 : 3;
  if (bc == 1 || bc == 4) {
    T un = s[iu] * n[0] + s[iu + 1] * n[1] + s[iu + 2] * n[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) s[iu + i] = s[iu + i] - (T)2.0 * un * n[i];
  } else if (bc == 5) {
#pragma unroll
    for (int i = 0; i < 3; ++i) s[iu + i] = -s[iu + i];
  } else if (bc == 3) {
    // inlet pressure and temperature: fk_params inletP / inletT, the
    // reference's constants (bc.h:5-7) filled in when both are 0
    const T P = m.inP, Tt = m.inT;
    if constexpr (NEQ == 5) {
      s[0] = rho_of(m, 0, P, Tt);
      s[4] = P;
      if constexpr (sizeof(T) == 4) s[4] = P - m.pr;  // FP32 path: deviation from pRef
    } else {
      s[0] = (T)1.0 - m.avfMin;
      s[1] = rho_of(m, 0, P, Tt);
      s[2] = rho_of(m, 1, P, Tt);
      s[6] = P;
      s[7] = P;
      s[8] = 0;
      s[9] = 0;
    }
  }
}

// Per-face result of the Riemann solve, shared by the two sides.
template <typename T, int NEQ>
struct FaceRes {
  T f[NEQ];    // F + H per equation (two-phase: eq 0, 6, 7 are F only)
  T unc[3];    // velocity of the non-conservative terms
  T pF;        // upwind face pressure
};

// HLLC face flux.  Two-phase: DP computeRHS.cu:78-146 with waveSpeeds.h,
// starRegion.h, flux.h; single-phase: SP computeRHS.cu:23-98.
// L, R: face states (R already the ghost on boundary faces).
template <int NEQ, typename T>
__device__ __forceinline__ void riemann(const Mat<T> &m, const T *L, const T *R, const T n[3], int is_wall,
                                        FaceRes<T, NEQ> &out) {
  if constexpr (NEQ == 5) {
    // FP32 path: L[4], R[4] are deviations from the reference pressure pr.
    // Pressure differences (S*) and the momentum flux use them as they are
    // (sum_f pr n_f A_f = 0 over a closed cell, so dropping pr there is
    // exact and avoids rounding a large constant in FP32); sound speeds and
    // the energy flux take the absolute pressure.
    T pa_L = L[4], pa_R = R[4];
    if constexpr (sizeof(T) == 4) {
      pa_L = L[4] + m.pr;
      pa_R = R[4] + m.pr;
    }
    T uL = L[1] * n[0] + L[2] * n[1] + L[3] * n[2];
    T uR = R[1] * n[0] + R[2] * n[1] + R[3] * n[2];
    T cL = c_of(m, 0, L[0], pa_L);
    T cR = c_of(m, 0, R[0], pa_R);
    T S0 = fmin(uL - cL, uR - cR);
    T S2 = fmax(uL + cL, uR + cR);
    T S1 = (R[4] - L[4] + L[0] * uL * (S0 - uL) - R[0] * uR * (S2 - uR)) / (L[0] * (S0 - uL) - R[0] * (S2 - uR));
    S1 = (T)(1 - is_wall) * S1;
    const bool left = S1 >= 0;
    const bool star = S0 < 0 && S2 >= 0;
    const T *s = left ? L : R;
    const T un = left ? uL : uR;
    const T SK = left ? S0 : S2;
    const T ps = left ? pa_L : pa_R;  // absolute upwind pressure
    T ke = (T)0.5 * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3]);
    T e = e1_of(m, ps, s[0]);
    T H[5] = {0, 0, 0, 0, 0};
    if (star) {
      T c0 = s[0], c1 = s[0] * s[1], c2 = s[0] * s[2], c3 = s[0] * s[3], c4 = s[0] * (e + ke);
      T v0 = s[0] * (SK - un) / (SK - S1);
      T vg0 = s[1] + (S1 - un) * n[0], vg1 = s[2] + (S1 - un) * n[1], vg2 = s[3] + (S1 - un) * n[2];
      T v4 = v0 * (e + ke + (S1 - un) * (S1 + ps / (s[0] * (SK - un))));
      H[0] = SK * (v0 - c0);
      H[1] = SK * (v0 * vg0 - c1);
      H[2] = SK * (v0 * vg1 - c2);
      H[3] = SK * (v0 * vg2 - c3);
      H[4] = SK * (v4 - c4);
    }
    T h = s[0] * (e + ke) + ps;
    T F0 = s[1] * s[0] * n[0] + s[2] * s[0] * n[1] + s[3] * s[0] * n[2];
    T F1 = (s[0] * s[1] * s[1] + s[4]) * n[0] + s[0] * s[1] * s[2] * n[1] + s[0] * s[1] * s[3] * n[2];
    T F2 = s[0] * s[2] * s[1] * n[0] + (s[0] * s[2] * s[2] + s[4]) * n[1] + s[0] * s[2] * s[3] * n[2];
    T F3 = s[0] * s[3] * s[1] * n[0] + s[0] * s[3] * s[2] * n[1] + (s[0] * s[3] * s[3] + s[4]) * n[2];
    T F4 = n[0] * s[1] * h + n[1] * s[2] * h + n[2] * s[3] * h;
    out.f[0] = F0 + H[0];
    out.f[1] = F1 + H[1];
    out.f[2] = F2 + H[2];
    out.f[3] = F3 + H[3];
    out.f[4] = F4 + H[4];
    out.unc[0] = out.unc[1] = out.unc[2] = 0;
    out.pF = 0;
  } else {
    T uL = L[3] * n[0] + L[4] * n[1] + L[5] * n[2];
    T uR = R[3] * n[0] + R[4] * n[1] + R[5] * n[2];
    T cL, cR;
    if ((L[8] + R[8]) / 2 > m.chemLimit && m.woods == 1) {
      T rm = L[0] * L[1] + (1 - L[0]) * L[2];
      T ca = c_of(m, 0, L[1], L[6]), cb = c_of(m, 1, L[2], L[7]);
      cL = 1 / sqrt(rm * (L[0] / (L[1] * ca * ca) + (1 - L[0]) / (L[2] * cb * cb)));
      rm = R[0] * R[1] + (1 - R[0]) * R[2];
      ca = c_of(m, 0, R[1], R[6]);
      cb = c_of(m, 1, R[2], R[7]);
      cR = 1 / sqrt(rm * (R[0] / (R[1] * ca * ca) + (1 - R[0]) / (R[2] * cb * cb)));
    } else {
      T d = L[0] * L[1] + (1 - L[0]) * L[2];
      T YL = L[0] * L[1] / d, YG = (1 - L[0]) * L[2] / d;
      T ca = c_of(m, 0, L[1], L[6]), cb = c_of(m, 1, L[2], L[7]);
      cL = sqrt(YL * (ca * ca) + YG * (cb * cb));
      d = R[0] * R[1] + (1 - R[0]) * R[2];
      YL = R[0] * R[1] / d;
      YG = (1 - R[0]) * R[2] / d;
      ca = c_of(m, 0, R[1], R[6]);
      cb = c_of(m, 1, R[2], R[7]);
      cR = sqrt(YL * (ca * ca) + YG * (cb * cb));
    }
    T S0 = fmin(uL - cL, uR - cR);
    T S2 = fmax(uL + cL, uR + cR);
    T pL = L[0] * L[6] + (1 - L[0]) * L[7];
    T pR = R[0] * R[6] + (1 - R[0]) * R[7];
    T rL = L[0] * L[1] + (1 - L[0]) * L[2];
    T rR = R[0] * R[1] + (1 - R[0]) * R[2];
    T S1 = (pR - pL + rL * uL * (S0 - uL) - rR * uR * (S2 - uR)) / (rL * (S0 - uL) - rR * (S2 - uR));
    const bool left = S1 >= 0;
    const bool star = S0 < 0 && S2 >= 0;
    const T *s = left ? L : R;
    const T un = left ? uL : uR;
    const T SK = left ? S0 : S2;
    // prim2cons of the upwind state (DP prim2cons.h:6-19)
    T a = s[0], a1 = 1 - a;
    T C1 = a * s[1], C2 = a1 * s[2];
    T rho = C1 + C2;
    T C6 = C1 * e_of(m, 0, s[6], s[1]);
    T C7 = C2 * e_of(m, 1, s[7], s[2]);
    T C8 = C6 + C7 + (T)0.5 * rho * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    T unc0 = s[3], unc1 = s[4], unc2 = s[5];
    T H1 = 0, H2 = 0, H3 = 0, H4 = 0, H5 = 0, H8 = 0, H9 = 0;
    T areL = C6, areG = C7;
    if (star) {
      // DP starRegion.h:3-35
      unc0 = s[3] + (S1 - un) * n[0];
      unc1 = s[4] + (S1 - un) * n[1];
      unc2 = s[5] + (S1 - un) * n[2];
      T rSL = s[1] * (SK - un) / (SK - S1);
      T rSG = s[2] * (SK - un) / (SK - S1);
      T V1 = a * rSL, V2 = a1 * rSG;
      T rS = V1 + V2;
      T pSL = (s[6] + m.pi[0]) * ((m.g[0] - 1) * s[1] - (m.g[0] + 1) * rSL) / ((m.g[0] - 1) * rSL - (m.g[0] + 1) * s[1]) - m.pi[0];
      T pSG = (s[7] + m.pi[1]) * ((m.g[1] - 1) * s[2] - (m.g[1] + 1) * rSG) / ((m.g[1] - 1) * rSG - (m.g[1] + 1) * s[2]) - m.pi[1];
      areL = V1 * e_of(m, 0, pSL, rSL);
      areG = V2 * e_of(m, 1, pSG, rSG);
      T press = s[0] * s[6] + (1 - s[0]) * s[7];
      T rhoK = s[0] * s[1] + (1 - s[0]) * s[2];
      T V8;
      if (m.hllcToro == 1) V8 = ((SK - un) / (SK - S1)) * (C8 + (S1 - un) * (rhoK * S1 + press / (SK - un)));
      else V8 = ((SK - un) / (SK - S1)) * (C8 + (S1 - un) * (S1 + press / (rhoK * (SK - un))));
      T V9 = s[8] * (SK - un) / (SK - S1);
      H1 = SK * (V1 - C1);
      H2 = SK * (V2 - C2);
      H3 = SK * (rS * unc0 - rho * s[3]);
      H4 = SK * (rS * unc1 - rho * s[4]);
      H5 = SK * (rS * unc2 - rho * s[5]);
      H8 = SK * (V8 - C8);
      H9 = SK * (V9 - s[8]);
    }
    // DP flux.h:32-91
    T press = s[0] * s[6] + (1 - s[0]) * s[7];
    T rE = s[0] * s[1] * e_of(m, 0, s[6], s[1]) + (1 - s[0]) * s[2] * e_of(m, 1, s[7], s[2]) +
           (T)0.5 * rho * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    T F0 = unc0 * s[0] * n[0] + unc1 * s[0] * n[1] + unc2 * s[0] * n[2];
    T F1 = s[3] * s[0] * s[1] * n[0] + s[4] * s[0] * s[1] * n[1] + s[5] * s[0] * s[1] * n[2];
    T F2 = s[3] * (1 - s[0]) * s[2] * n[0] + s[4] * (1 - s[0]) * s[2] * n[1] + s[5] * (1 - s[0]) * s[2] * n[2];
    T F3 = rho * s[3] * s[3] * n[0] + rho * s[3] * s[4] * n[1] + rho * s[3] * s[5] * n[2];
    T F4 = rho * s[4] * s[3] * n[0] + rho * s[4] * s[4] * n[1] + rho * s[4] * s[5] * n[2];
    T F5 = rho * s[5] * s[3] * n[0] + rho * s[5] * s[4] * n[1] + rho * s[5] * s[5] * n[2];
    T F6 = unc0 * areL * n[0] + unc1 * areL * n[1] + unc2 * areL * n[2];
    T F7 = unc0 * areG * n[0] + unc1 * areG * n[1] + unc2 * areG * n[2];
    T F8 = s[3] * (rE + press) * n[0] + s[4] * (rE + press) * n[1] + s[5] * (rE + press) * n[2];
    T F9 = s[3] * s[8] * n[0] + s[4] * s[8] * n[1] + s[5] * s[8] * n[2];
    out.f[0] = F0;
    out.f[1] = F1 + H1;
    out.f[2] = F2 + H2;
    out.f[3] = F3 + H3;
    out.f[4] = F4 + H4;
    out.f[5] = F5 + H5;
    out.f[6] = F6;
    out.f[7] = F7;
    out.f[8] = F8 + H8;
    out.f[9] = F9 + H9;
    out.unc[0] = unc0;
    out.unc[1] = unc1;
    out.unc[2] = unc2;
    out.pF = press;
  }
}

// Maximum signal speed |u.n| + c_mix of a face state
// (DP computeTimeStep.cu:15-46; single-phase: |u.n| + c).
template <int NEQ, typename T>
__device__ __forceinline__ T signal(const Mat<T> &m, const T *W, const T n[3]) {
  if constexpr (NEQ == 5) {
    T un = W[1] * n[0] + W[2] * n[1] + W[3] * n[2];
    T pa = W[4];
    if constexpr (sizeof(T) == 4) pa = W[4] + m.pr;  // FP32 path: deviation from pRef
    return fabs(un) + c_of(m, 0, W[0], pa);
  } else {
    T avf = W[0], rL = W[1], rG = W[2];
    T rm = avf * rL + (1 - avf) * rG;
    T cL = c_of(m, 0, rL, W[6]), cG = c_of(m, 1, rG, W[7]);
    T cmix;
    if (W[8] > m.chemLimit && m.woods == 1) cmix = 1 / sqrt(rm * (avf / (rL * cL * cL) + (1 - avf) / (rG * cG * cG)));
    else {
      T YL = avf * rL / rm, YG = (1 - avf) * rG / rm;
      cmix = sqrt(YL * cL * cL + YG * cG * cG);
    }
    T un = W[3] * n[0] + W[4] * n[1] + W[5] * n[2];
    return fabs(un) + cmix;
  }
}

// WALE eddy viscosity (DP computeViscousStresses.cu:3-60) with the powers
// written through sqrt; Ls = 0.325 * filter volume^(1/3).
template <typename T>
__device__ __forceinline__ T wale(const T g[3][3], T Ls) {
  T sq[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T acc = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) acc += g[i][k] * g[k][j];
      sq[i][j] = acc;
    }
  T tr = (T)0.0;
  tr += sq[0][0];
  tr += sq[1][1];
  tr += sq[2][2];
  T sum1 = 0, sum2 = 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T sd = (T)0.5 * (sq[i][j] + sq[j][i]) - ((T)(1.0 / 3.0) * (T)(i == j)) * tr;
      T st = (T)0.5 * (g[i][j] + g[j][i]);
      sum1 += sd * sd;
      sum2 += st * st;
    }
  T num = sum1 * sqrt(sum1);
  T den = sum1 * sqrt(sqrt(sum1)) + sum2 * sum2 * sqrt(sum2) + (T)1e-12;
  if (den < (T)1e-10) num = 0;
  return Ls * Ls * num / den;
}

}  // namespace fk
