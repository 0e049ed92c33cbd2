// MI355X (gfx950) kernels of the explicit unstructured finite-volume step
// and the C ABI of include/fooska.h.
//
// Layout in HBM (one context):
//   U        3 rotating buffers [nElem][NEQ] double (reference AoS consVec)
//   slots    4 per cell, ordered by ascending face id:
//              link[4C]  (mirror slot in the neighbour << 1) | owner flag,
//                        or -bc on boundary faces (owner side); the
//                        neighbour cell is mirror slot / 4
//              geo[4C]   (n.x, n.y, n.z, A) of the face (owner normal)
//              rvec[4C]  (x_f - x_c) of the face seen from this cell, and the
//                        slot link in the 4th word (bits of an int32 in FP32,
//                        its value in FP64) so pass G reads no link array
//              wls[4C]   WALE filter width 0.325 ((V_o+V_n)/2)^0.3333333333333
//   face     [4C][NEQ] T: MUSCL face state of every cell face slot
//   gvel     [C][NG]  T: velocity gradient (+ temperature gradient)
//   V, Vcfl  [C] double: volume and V^0.3333
// T = double (FP64 path) or float (mixed path).
//
// Pass G (fk_compute_face_values) = computePrimVec + computeGrad +
// computeFaceValues; pass F (fk_compute_rhs[_update]) = computeRHS (+ viscous)
// + update.  Both are cell-centric gathers: each thread owns one cell and
// sums its four faces in slot order, so no atomics and a fixed summation
// order.  The Riemann problem of an interior face is solved by both of its
// cells (identically), which trades ALU for not writing/reading per-face
// fluxes through HBM.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/fooska.h"
#include "fk_device.h"
#include "fk_relax.h"
#include "fk_internal.h"

using namespace fkx;

// Scratch of fk_cons_calc_noalloc: the per-block partial sums of the volume
// integrals ([kConsBlocks][NEQ]) followed by the folded sums ([NEQ]).  Each
// device holds its own copy for the life of the process (32.9 KB).  In host
// code `scratch` only names the symbol: kernels and copies take the device
// address from hipGetSymbolAddress, kept per device in g_scratch.
constexpr int kConsBlocks = 256;
constexpr int kScratchDoubles = kConsBlocks * 16 + 16;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device `dev`, looked up once per
// device (called by fk_create with that device current).
int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(FK_EDEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return FK_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(FK_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[dev] = static_cast<double *>(ptr);
  return FK_OK;
}

double *scratch_of(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[dev];
}
}  // namespace

namespace {

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------

template <int NEQ>
struct Model {
  static constexpr int iu = NEQ == 5 ? 1 : 3;  // first velocity slot of W
};

// number of gradient values kept for the viscous pass
template <int NEQ, bool HEAT>
constexpr int ngrad() {
  return 9 + (HEAT ? 3 : 0);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Remap so that XCD x processes the contiguous block range
// [x G/8, (x+1) G/8): neighbouring cells (close in the mesh order) then share
// an L2.  G is a multiple of 8 (padded by grid_xcd).
constexpr int kXcd = 8;
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t b = blockIdx.x, per = gridDim.x / kXcd;
  return (b % kXcd) * per + b / kXcd;
}

template <typename T>
__device__ __forceinline__ V4<T> ld4(const void *p, int64_t i) {
  return reinterpret_cast<const V4<T> *>(p)[i];
}

// Slots per cell: NS = 4 on meshes of tetrahedra, 6 on meshes with other
// elements (pyramids and prisms have 5 faces, hexahedra 6); a cell's slots past
// its last face hold the link kNoFace and zero geometry.
constexpr int kNoFace = -8;

// slot link -> neighbour cell (or -bc) and owner flag
template <int NS>
__device__ __forceinline__ int link_nb(int lk) { return lk >= 0 ? (lk >> 1) / NS : lk; }
__device__ __forceinline__ bool link_own(int lk) { return lk < 0 || (lk & 1); }
template <typename T>
__device__ __forceinline__ int link_of(T w) {
  if constexpr (sizeof(T) == 4) return __float_as_int(w);
  else return (int)w;
}

// Pass G: first-order states, Green-Gauss gradient, neighbour min/max,
// limiter and MUSCL face states (computePrimVec.cu:38-80,
// computeGrad.cu:51-91, computeFaceValues.cu:29-113).  Every cell converts
// its own state and its neighbours' (cons2prim, FP64) and keeps all
// variables in registers; full face records are written once.  FP64 is held
// to 256 VGPRs (2 waves per SIMD, no AGPR spills: 13.0 against 15.9 ms at
// 48M cells); mixed precision runs at the compiler's choice (3 waves per
// SIMD; 2 and 4 measured no faster).
//
// Mixed precision: the first-order states are fk::prims_f32(U).  FROMW: they
// come from the FP32 primitive records Wp written by the producer of U (the
// same values, half the gathered bytes, no FP64 conversions here: 6.45
// against 7.8 ms at 48M cells).
// Floats per cell of pass G's LDS copy of its face records: the record
// length rounded up to 16-byte words, no padding.  Rows of 40, 44 and 52
// floats measured the same at 48M cells (bank conflicts of the row writes
// do not show), and 40 (40 KB per workgroup) lets four workgroups share a
// CU's LDS.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int N>
constexpr int gout_stride() {
  constexpr int w = (N + 3) / 4;
  return 4 * w;
}

// Workgroup of pass G: kBlock cells, or with the LDS store path (SOUT) the
// cells whose face records fill 30-40 KB, so that four workgroups share a
// CU's LDS: 4 slots 256 cells (FP64 128, its occupancy of 2 waves per SIMD),
// 6 slots (FP32) 128.
template <typename T, int NS, bool SOUT>
constexpr int pass_g_block() {
  if (!SOUT) return kBlock;
  if (NS == 4) return sizeof(T) == 8 ? 128 : kBlock;
  return 128;
}

// Record layouts that take the LDS store path: each cell's face records are
// whole 16-byte words (not single-phase FP32 with 6 slots, 120 bytes), and
// not FP64 with 6 slots, whose 480-byte records leave 64-cell workgroups
// (pass G 3.29 against 3.21 ms with direct stores on 200^3 hexahedra; FP32
// with 6 slots: 1.53 against 2.05 ms).
template <typename T, int NS, int NEQ>
constexpr bool pass_g_sout_ok() {
  return NS * NEQ * (int)sizeof(T) % 16 == 0 && (NS == 4 || sizeof(T) == 4);
}

template <int NEQ, typename T, bool VISC, bool HEAT, bool FROMW, int NS, bool SOUT = false>
__global__ void __launch_bounds__(kBlock, (sizeof(T) == 8 ? 2 : 1))
    pass_g(fk_params P, int64_t nc, const double *__restrict__ U, const float *__restrict__ Wp,
           const int32_t *__restrict__ link, const void *__restrict__ geo, const void *__restrict__ rvec,
           const double *__restrict__ vol, T *__restrict__ face,
           T *__restrict__ gvel, double *__restrict__ dgrad, double *__restrict__ dtheta,
           unsigned long long *__restrict__ health, int step) {
  static_assert(!SOUT || pass_g_sout_ok<T, NS, NEQ>(), "face records staged as 16-byte words");
  constexpr int GB = pass_g_block<T, NS, SOUT>();
  // T per cell in the LDS copy of the face records (FP64: unpadded), and floats
  constexpr int SPT = sizeof(T) == 4 ? gout_stride<NS * NEQ>() : NS * NEQ;
  constexpr int SP = SPT * (int)sizeof(T) / 4;
  __shared__ __attribute__((aligned(16))) float sOut[SOUT ? GB * SP : 1];
  const int64_t c0 = xcd_block() * GB;
  const int64_t cl = c0 + threadIdx.x;
  if constexpr (!SOUT) {
    if (cl >= nc) return;
  }
  // SOUT: every thread reaches the block's barrier; those past the last cell
  // compute the last cell again and write only to their own LDS row, which
  // the copy-out does not read
  const bool live = cl < nc;
  const int64_t c = live ? cl : nc - 1;
  const fk::Mat<double> md = fk::make_mat<double>(P);
  const fk::Mat<T> mt = fk::make_mat<T>(P);
  T W[NEQ];
  if constexpr (FROMW) {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) W[v] = (T)Wp[c * NEQ + v];
  } else if constexpr (sizeof(T) == 4) {
    float Wf[NEQ];
    double Uc[NEQ];
    fk::load_rec<NEQ>(U + c * NEQ, Uc);
    fk::prims_f32<NEQ>(md, Uc, Wf);
#pragma unroll
    for (int v = 0; v < NEQ; ++v) W[v] = Wf[v];
  } else {
    double Wd[NEQ];
    fk::cons2prim<NEQ>(md, U + c * NEQ, Wd);
#pragma unroll
    for (int v = 0; v < NEQ; ++v) W[v] = (T)Wd[v];
  }
  {
    // health of the state: phase (or mixture) densities positive and every
    // primitive finite (a NaN fails the comparisons, an overflow makes the
    // sum inf); the smallest (step, cell) that fails is kept
    T sum = 0;
#pragma unroll
    for (int v = 0; v < NEQ; ++v) sum += W[v];
    const bool ok = (NEQ == 10 ? (W[1] > 0 && W[2] > 0) : W[0] > 0) && fabs(sum) < (T)INFINITY;
    if (!ok && live) atomicMin(health, ((unsigned long long)(unsigned)step << 32) | (unsigned long long)c);
  }
  T g[NEQ][3], mn[NEQ], mx[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    g[v][0] = g[v][1] = g[v][2] = 0;
    mn[v] = (T)1e12;
    mx[v] = (T)-1e12;
  }
  const T V = (T)vol[c];
  const T iV = (T)1 / V;  // FP32: one division per cell
  // The NS slot links are loaded before the face loop and rotated through
  // registers, so each iteration's neighbour gather does not wait on its link
  // load (rolled loop: 6.19 against 6.49 ms mixed, 12.8 against 12.9 ms FP64
  // at 48M cells).  Mixed: the link rides in rvec's 4th word (one 16-byte
  // load per slot); FP64 (32-byte rvec records) measured faster with the
  // 4-byte link array.
  int lq[NS];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if constexpr (sizeof(T) == 4) lq[j] = link_of<T>(ld4<T>(rvec, c * NS + j).w);
    else lq[j] = link[c * NS + j];
  }
#pragma unroll 1
  for (int j = 0; j < NS; ++j) {
    const int64_t s = c * NS + j;
    const int lk = lq[0];
#pragma unroll
    for (int q = 0; q + 1 < NS; ++q) lq[q] = lq[q + 1];
    if (NS > 4 && lk == kNoFace) break;  // past the cell's last face
    const int nb = link_nb<NS>(lk);
    const bool own = link_own(lk);
    const V4<T> ga = ld4<T>(geo, s);
    const T n[3] = {ga.x, ga.y, ga.z};
    // FP32: the slot's signed n A / V, so that each gradient term is one FMA
    // instead of three products and a division (pass G is bound by its VALU
    // issue once its stores go through LDS: 4.64 against 5.16 ms at 48M
    // cells).  The same FMA form in the MUSCL sums and a hardware reciprocal
    // in the limiter saved another 0.1 ms but moved the bubble collapse's
    // loss of positivity earlier (48^3 mesh: step 33 instead of after 50;
    // the scheme is not positivity preserving there), so those stay as the
    // reference writes them.
    T nav[3];
    if constexpr (sizeof(T) == 4) {
      const T sa = (own ? ga.w : -ga.w) * iV;
#pragma unroll
      for (int d = 0; d < 3; ++d) nav[d] = n[d] * sa;
    }
    T Wn[NEQ];
    if (nb >= 0) {
      if constexpr (FROMW) {
#pragma unroll
        for (int v = 0; v < NEQ; ++v) Wn[v] = (T)Wp[(int64_t)nb * NEQ + v];
      } else if constexpr (sizeof(T) == 4) {
        float t[NEQ];
        double Un[NEQ];
        fk::load_rec<NEQ>(U + (int64_t)nb * NEQ, Un);
        fk::prims_f32<NEQ>(md, Un, t);
#pragma unroll
        for (int v = 0; v < NEQ; ++v) Wn[v] = t[v];
      } else {
        double t[NEQ];
        fk::cons2prim<NEQ>(md, U + (int64_t)nb * NEQ, t);
#pragma unroll
        for (int v = 0; v < NEQ; ++v) Wn[v] = (T)t[v];
      }
    } else {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) Wn[v] = W[v];
      fk::ghost<NEQ, T>(mt, -nb, n, Wn);
    }
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      const T o = Wn[v];
      if (!(mx[v] >= o)) mx[v] = o;
      if (!(mn[v] <= o)) mn[v] = o;
      const T avg = (W[v] + Wn[v]) / 2;
      if constexpr (sizeof(T) == 8) {
        // owner +, neighbour -: (-a) b = -(a b) exactly, so selecting the
        // sign of the term equals evaluating -avg n A / V (no branch)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const T t = avg * n[d] * ga.w / V;
          g[v][d] += own ? t : -t;
        }
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) g[v][d] = fmaf(avg, nav[d], g[v][d]);
      }
    }
  }
  // limiter: theta = min over faces of clamp(phi, 0, 1)
  T th[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) th[v] = (T)1e9;
  // slots past the last face have r = 0: phi = 1 there, which leaves theta
  // (at most 1 after the first face) unchanged
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const V4<T> r = ld4<T>(rvec, c * NS + j);
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      const T wu = W[v] + r.x * g[v][0] + r.y * g[v][1] + r.z * g[v][2];
      T phi;
      if (wu - W[v] != 0)
        phi = (wu > W[v]) ? (mx[v] - W[v]) / (2 * (wu - W[v])) : (mn[v] - W[v]) / (2 * (wu - W[v]));
      else
        phi = 1;
      const T tt = fmax((T)0, fmin(phi, (T)1));
      if (!(th[v] <= tt)) th[v] = tt;
    }
  }
  // MUSCL face states
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    V4<T> r = ld4<T>(rvec, c * NS + j);
    // r g is also formed by the limiter loop; the empty asm keeps the
    // compiler from holding those 4 NEQ products in registers until here (it
    // recomputes them instead): mixed 127 instead of 166 VGPRs, 4 waves per
    // SIMD, 5.34 against 5.50 ms with the LDS store path (without it, when
    // the stores bounded the kernel, holding them had measured faster).
    asm volatile("" : "+v"(r.x), "+v"(r.y), "+v"(r.z));
    T *out = SOUT ? reinterpret_cast<T *>(sOut) + threadIdx.x * SPT + j * NEQ : face + (c * NS + j) * NEQ;
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      const T ext = r.x * g[v][0] + r.y * g[v][1] + r.z * g[v][2];
      out[v] = W[v] + th[v] * ext;
    }
  }
  if constexpr (SOUT) {
    // the block's face records are one contiguous range: written from LDS in
    // 16-byte words, each wave store covering 1 KB contiguously (a thread's
    // own 160-byte record otherwise goes out as strided partial lines).
    // Non-temporal: the records are read once, by the next pass, and whole
    // lines written past L2 leave it to the neighbour records this pass
    // gathers (mixed pass G 5.27 against 5.49 ms, FP64 10.83 against 11.12 ms
    // at 48M cells; with partial lines, before this path, such stores had
    // measured 3.5 times slower).
    __syncthreads();
    constexpr int WPC = NS * NEQ * (int)sizeof(T) / 16;  // 16-byte words per cell
    const int64_t ncell = (nc - c0) < GB ? (nc - c0) : GB;
    const int nw = (int)ncell * WPC;
    uint4 *dst = reinterpret_cast<uint4 *>(face + c0 * NS * NEQ);
    for (int w = threadIdx.x; w < nw; w += GB) {
      const int cell = w / WPC, part = w - cell * WPC;
      __builtin_nontemporal_store(*reinterpret_cast<const u32x4 *>(sOut + cell * SP + part * 4),
                                  reinterpret_cast<u32x4 *>(&dst[w]));
    }
  }
  if constexpr (VISC) {
    constexpr int iu = Model<NEQ>::iu;
    constexpr int NG = ngrad<NEQ, HEAT>();
    // SOUT: the gradient rows go through the same LDS (after the face copy
    // has read it), stride NG, and out as the block's contiguous range
    if constexpr (SOUT) __syncthreads();
    T *G = SOUT ? reinterpret_cast<T *>(sOut) + threadIdx.x * NG : gvel + c * NG;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) G[i * 3 + d] = g[iu + i][d];
    if constexpr (HEAT) {
      const T gm = (T)P.gamma[0];
      // FP32 path: W[4] is the deviation from pRef, absolute p = W[4] + pRef
      const T pi = sizeof(T) == 4 ? (T)(P.pInf[0] + P.pRef) : (T)P.pInf[0];
      const T R = (T)P.Cp[0] * (gm - 1) / gm;
#pragma unroll
      for (int d = 0; d < 3; ++d) G[9 + d] = (g[4][d] - ((W[4] + pi) / W[0]) * g[0][d]) / (W[0] * R);
    }
    if constexpr (SOUT) {
      __syncthreads();
      const int64_t ncell = (nc - c0) < GB ? (nc - c0) : GB;
      const int nf = (int)ncell * NG;
      T *dg = gvel + c0 * NG;
      for (int f = threadIdx.x; f < nf; f += GB) __builtin_nontemporal_store(reinterpret_cast<const T *>(sOut)[f], &dg[f]);
    }
  }
  if (SOUT && !live) return;
  if (dgrad) {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      dtheta[c * NEQ + v] = (double)th[v];
#pragma unroll
      for (int d = 0; d < 3; ++d) dgrad[(c * NEQ + v) * 3 + d] = (double)g[v][d];
    }
  }
}

// Pass F: Riemann problem on each face of the cell, non-conservative terms,
// pressure correction, viscous/WALE terms, per-cell gather and (MODE 0) the
// RK stage, (MODE 1) residual + signal speed for the time step, or (MODE 2)
// the residual only.
// (computeRHS.cu:78-146, computeViscousStresses.cu:66-174, update.cu:3-103,
// computeTimeStep.cu:97-159)
//
// STAGE: the workgroup first copies the face states of its own cells (4 slots
// each) and their velocity gradients into LDS; the mirror states and
// neighbour gradients of faces inside the workgroup are then read from LDS,
// only those of faces to other workgroups from global memory.
//
// WOUT (MODE 0, mixed): also write the FP32 primitive record of the new state.
// A template flag rather than a null test: the FP64 conversion at the end
// makes the kernel spill at 3 waves per SIMD (8.4 against 7.6 ms), so the
// stages whose records are not read use the variant without it.
//
// RLX (MODE 0, two-phase): the last RK stage also relaxes the state it has
// just formed, so the state is not read back by a separate kernel: 1 runs
// stage 1 of the surrogate relaxation (fk::screen_cell), 2 the plain
// per-cell relaxation (fk::plain_cell).  Cells left for a second kernel go to
// rlist (1: fk::list_push, lengths in rcnt[5], rcnt[6]; 2: chemical cells,
// length in rcnt[5]); each wave stores its counts in rpart[wave]
// ([n_chem, iterations], and Tsat iterations for 2); Wout gets the primitive
// records of the cells relaxed here.
//
// Workgroup size (pass_f_block): 256 cells; with staging, FP64 128 cells,
// whose face records and gradients take 52 KB of LDS per 128 cells (104 KB
// at 256 would leave one workgroup per CU), and 128 cells (37 KB) with 6
// slots.
template <typename T, bool STAGE, int NS = 4>
constexpr int pass_f_block() {
  if (!STAGE) return kBlock;
  if (NS == 4) return sizeof(T) == 8 ? 128 : kBlock;
  return 128;
}

// Layouts that pass F stages: 4 slots; 6 slots for two-phase FP32 records
// (whole 16-byte words per workgroup row, 37 KB per 128 cells).  On 200^3
// hexahedra (8M cells) the fused stage went from 2.61 to 1.80 ms and the dt
// stage from 2.46 to 1.75 ms; the last stage with the per-cell relaxation
// slowed down there (2.74 to 3.01 ms, its registers at the lower occupancy
// of the larger LDS share), so with 6 slots it stays unstaged.
template <typename T, int NS, int NEQ>
constexpr bool pass_f_stage_ok() {
  return NS == 4 || (sizeof(T) == 4 && NEQ == 10);
}

template <int NEQ, typename T, bool VISC, bool HEAT, int MODE, bool STAGE, int NS, bool WOUT = false, int RLX = 0>
__global__ void __launch_bounds__((pass_f_block<T, STAGE, NS>()), ((MODE == 1 || sizeof(T) == 8) ? 2 : 3)) pass_f(fk_params P, int64_t nc, const double *__restrict__ U,
                                                 const double *__restrict__ U0, double *__restrict__ Unew,
                                                 const int32_t *__restrict__ link,
                                                 const void *__restrict__ geo, const T *__restrict__ wls,
                                                 const double *__restrict__ vol, const double *__restrict__ volcfl,
                                                 const T *__restrict__ face, const T *__restrict__ gvel,
                                                 double *__restrict__ rhs_out, double *__restrict__ dtc, int irk,
                                                 double dt_val, const double *__restrict__ dt_ptr,
                                                 float *__restrict__ Wout, uint32_t *__restrict__ rlist,
                                                 unsigned long long *__restrict__ rcnt,
                                                 unsigned long long *__restrict__ rpart) {
  static_assert(!RLX || (MODE == 0 && NEQ == 10), "fused relaxation: two-phase RK stage only");
  constexpr int KR = RLX == 2 ? 3 : 2;  // counts per wave in rpart
  constexpr int NGS = ngrad<NEQ, HEAT>();
  constexpr int BS = pass_f_block<T, STAGE, NS>();
  const int64_t c0 = xcd_block() * BS;
  const int64_t c = c0 + threadIdx.x;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *sF = reinterpret_cast<T *>(lds_raw);
  T *sG = sF + BS * NS * NEQ;
  if constexpr (STAGE) {
    // the workgroup's face records and gradients are contiguous in HBM and in
    // LDS: copied by direct-to-LDS loads (global_load_lds: lane i of a wave
    // fills bytes [s i, s i + s) after the wave's LDS base, s = 16 or 4), no
    // registers involved.  Measured 7.0 against 7.3 ms (fused pass F, 48M
    // cells) for a copy through registers, one cell per thread.
    typedef __attribute__((address_space(3))) void *lds_ptr;
    const int64_t ncell = (nc - c0) < BS ? (nc - c0) : BS;
    const int wbase = __builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~63u));
    {
      constexpr int kW = BS * NS * NEQ * (int)sizeof(T) / 16;  // 16-byte words of a full block
      static_assert(kW % BS == 0, "whole 16-byte words per thread");
      const int nw = (int)(ncell * NS * NEQ * (int64_t)sizeof(T) / 16);
      const uint4 *src = reinterpret_cast<const uint4 *>(face + c0 * NS * NEQ);
      uint4 *dst = reinterpret_cast<uint4 *>(sF);
#pragma unroll
      for (int k = 0; k < kW / BS; ++k) {
        const int i = k * BS + (int)threadIdx.x;
        if (i < nw) __builtin_amdgcn_global_load_lds((const void *)(src + i), (lds_ptr)(dst + k * BS + wbase), 16, 0, 0);
      }
    }
    if constexpr (VISC) {
      constexpr int WPC = NGS * (int)sizeof(T) / 4;  // 4-byte words per cell
      const int nw = (int)(ncell * WPC);
      const float *src = reinterpret_cast<const float *>(gvel + c0 * NGS);
      float *dst = reinterpret_cast<float *>(sG);
#pragma unroll
      for (int k = 0; k < WPC; ++k) {
        const int i = k * BS + (int)threadIdx.x;
        if (i < nw) __builtin_amdgcn_global_load_lds((const void *)(src + i), (lds_ptr)(dst + k * BS + wbase), 4, 0, 0);
      }
    }
    // the direct-to-LDS loads count as vector memory loads: wait for all of
    // them explicitly before the barrier, so other waves' rows are in LDS
    // whatever the compiler's lowering of the barrier
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  if (c >= nc) {
    // a wave with no cell in range (lane 0 holds its lowest cell) still
    // stores its relaxation counts
    if constexpr (RLX != 0)
      if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < KR; ++k) rpart[(c >> 6) * KR + k] = 0;
    return;
  }
  // face state of global slot `sl` of cell `cell`, from LDS when staged
  auto slot_ptr = [&](int64_t cell, int64_t sl) -> const T * {
    if (STAGE && cell >= c0 && cell < c0 + BS) return sF + (sl - c0 * NS) * NEQ;
    return face + sl * NEQ;
  };
  auto grad_ptr = [&](int64_t cell) -> const T * {
    if (STAGE && cell >= c0 && cell < c0 + BS) return sG + (cell - c0) * NGS;
    return gvel + cell * NGS;
  };
  const fk::Mat<double> md = fk::make_mat<double>(P);
  const fk::Mat<T> mt = fk::make_mat<T>(P);
  constexpr int iu = Model<NEQ>::iu;
  constexpr int NG = ngrad<NEQ, HEAT>();
  double Uc[NEQ];
  fk::load_rec<NEQ>(U + c * NEQ, Uc);
  // fused RK stage (MODE 0): the state part of the stage, a U0 + b U, is
  // formed here, so that the load of U0 overlaps the face loop and U need not
  // stay live; the tail adds the residual part in the same order as
  // rk_update (bitwise equal)
  double base[MODE == 0 ? NEQ : 1];
  if constexpr (MODE == 0) {
    if (P.rkSteps == 1) {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) base[v] = Uc[v];
    } else {
      double u0[NEQ];
      fk::load_rec<NEQ>(U0 + c * NEQ, u0);
#pragma unroll
      for (int v = 0; v < NEQ; ++v) {
        if (irk == 0) base[v] = u0[v];
        else if (irk == 1) base[v] = (3.0 / 4.0) * u0[v] + (1.0 / 4.0) * Uc[v];
        else base[v] = (1.0 / 3.0) * u0[v] + (2.0 / 3.0) * Uc[v];
      }
    }
  }
  // side terms of this cell (fluxSource / computeCenterPress, two-phase)
  T avf = 0, pLiq = 0, pGas = 0;
  double pCd = 0;
  if constexpr (NEQ == 10) {
    const double a = fk::clamp_avf(md, Uc[0]);
    const double rl = Uc[1] / a, rg = Uc[2] / (1 - a);
    const double el = Uc[6] / (a * rl), eg = Uc[7] / ((1 - a) * rg);
    const double pl = fk::p_of(md, 0, rl, el), pg = fk::p_of(md, 1, rg, eg);
    pCd = a * pl + (1 - a) * pg;
    avf = (T)a;
    pLiq = (T)pl;
    pGas = (T)pg;
  }
  double rhs[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; ++v) rhs[v] = 0.0;
  T sig = (T)-1e9;
  // viscous contributions are summed apart (rv) and added to the inviscid sums
  // after the face loop: the canonical order shared with the oracle
  double rv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  T Wc1[NEQ];  // first-order state of this cell (viscous boundary ghost)
  if constexpr (VISC) {
    double t[NEQ];
    fk::cons2prim<NEQ>(md, Uc, t);
    if constexpr (NEQ == 5 && sizeof(T) == 4) t[4] -= md.pr;  // FP32 path: pressure as deviation from pRef
#pragma unroll
    for (int v = 0; v < NEQ; ++v) Wc1[v] = (T)t[v];
  }
  // slot links loaded up front and rotated, as in pass G (mixed 6.85 against
  // 7.12 ms for the fused pass, FP64 14.4 against 14.8 ms)
  int mq[NS];
#pragma unroll
  for (int j = 0; j < NS; ++j) mq[j] = link[c * NS + j];
  for (int j = 0; j < NS; ++j) {
    const int64_t s = c * NS + j;
    const int mv = mq[0];
#pragma unroll
    for (int q = 0; q + 1 < NS; ++q) mq[q] = mq[q + 1];
    if (NS > 4 && mv == kNoFace) break;  // past the cell's last face
    const int nb = link_nb<NS>(mv);
    const bool own = link_own(mv);
    const V4<T> ga = ld4<T>(geo, s);
    const T n[3] = {ga.x, ga.y, ga.z};
    const T A = ga.w;
    T L[NEQ], R[NEQ];
    const T *Fs = slot_ptr(c, s);
    if (nb >= 0) {
      const T *Fn = slot_ptr(nb, (int64_t)(mv >> 1));
      if (own) {
#pragma unroll
        for (int v = 0; v < NEQ; ++v) {
          L[v] = Fs[v];
          R[v] = Fn[v];
        }
      } else {
#pragma unroll
        for (int v = 0; v < NEQ; ++v) {
          L[v] = Fn[v];
          R[v] = Fs[v];
        }
      }
    } else {
#pragma unroll
      for (int v = 0; v < NEQ; ++v) L[v] = R[v] = Fs[v];
      fk::ghost<NEQ, T>(mt, -nb, n, R);
    }
    fk::FaceRes<T, NEQ> fr;
    fk::riemann<NEQ, T>(mt, L, R, n, nb == -5, fr);
    if constexpr (NEQ == 5) {
#pragma unroll
      for (int v = 0; v < 5; ++v) rhs[v] += (double)(own ? -fr.f[v] * A : fr.f[v] * A);
    } else {
      const T sign = own ? (T)-1.0 : (T)1.0;
      const T *u = fr.unc;
      const T nc0 = avf * u[0] * n[0] + avf * u[1] * n[1] + avf * u[2] * n[2];
      const T nc1 = avf * pLiq * u[0] * n[0] + avf * pLiq * u[1] * n[1] + avf * pLiq * u[2] * n[2];
      const T nc2 = (1 - avf) * pGas * u[0] * n[0] + (1 - avf) * pGas * u[1] * n[1] + (1 - avf) * pGas * u[2] * n[2];
      T Pc;
      if constexpr (sizeof(T) == 8) Pc = fr.pF - (T)pCd;
      else Pc = (T)((double)fr.pF - pCd);
      rhs[0] += (double)(sign * (fr.f[0] - nc0) * A);
      rhs[1] += (double)(sign * (fr.f[1]) * A);
      rhs[2] += (double)(sign * (fr.f[2]) * A);
      rhs[3] += (double)(sign * (fr.f[3] + Pc * n[0]) * A);
      rhs[4] += (double)(sign * (fr.f[4] + Pc * n[1]) * A);
      rhs[5] += (double)(sign * (fr.f[5] + Pc * n[2]) * A);
      rhs[6] += (double)(sign * (fr.f[6] + nc1) * A);
      rhs[7] += (double)(sign * (fr.f[7] + nc2) * A);
      rhs[8] += (double)(sign * (fr.f[8]) * A);
      rhs[9] += (double)(sign * (fr.f[9]) * A);
    }
    if constexpr (MODE == 1) {
      const T cl = fk::signal<NEQ, T>(mt, L, n);
      const T cr = nb >= 0 ? fk::signal<NEQ, T>(mt, R, n) : (T)0.0;
      const T cm = fmax(cl, cr);
      if (!(sig >= cm)) sig = cm;
    }
    if constexpr (VISC) {
      const T *Gc = grad_ptr(c);
      const T *Gn = nb >= 0 ? grad_ptr(nb) : Gc;
      const T *G0 = own ? Gc : Gn;  // owner's gradient first, as (g_o + g_n)/2
      const T *G1 = own ? Gn : Gc;
      T gm[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) gm[i][d] = (G0[i * 3 + d] + G1[i * 3 + d]) / 2;
      T eddy = 0;
      if (P.turbulence == 1) eddy = fk::wale<T>(gm, wls[s]);
      const T ux = gm[0][0], uy = gm[0][1], uz = gm[0][2], vx = gm[1][0], vy = gm[1][1], vz = gm[1][2];
      const T wx = gm[2][0], wy = gm[2][1], wz = gm[2][2];
      T st[3][3];
      st[0][0] = (2 * ux - (T)(2.0 / 3.0) * (ux + vy + wz));
      st[1][1] = (2 * vy - (T)(2.0 / 3.0) * (ux + vy + wz));
      st[2][2] = (2 * wz - (T)(2.0 / 3.0) * (ux + vy + wz));
      st[0][1] = st[1][0] = (uy + vx);
      st[0][2] = st[2][0] = (uz + wx);
      st[1][2] = st[2][1] = (vz + wy);
      // side-1 state: reconstructed neighbour, or the first-order ghost on boundaries
      T R1[NEQ];
      if (nb >= 0) {
#pragma unroll
        for (int v = 0; v < NEQ; ++v) R1[v] = R[v];
      } else {
#pragma unroll
        for (int v = 0; v < NEQ; ++v) R1[v] = Wc1[v];
        fk::ghost<NEQ, T>(mt, -nb, n, R1);
      }
      T mu;
      if constexpr (NEQ == 5) {
        const T rhoF = (L[0] + R1[0]) / 2;
        mu = mt.mu[0] + rhoF * eddy;
      } else {
        const T a1 = L[0], r1 = a1 * L[1] + (1 - a1) * L[2];
        const T a2 = R1[0], r2 = a2 * R1[1] + (1 - a2) * R1[2];
        const T rhoF = (r1 + r2) / 2, avfF = (a1 + a2) / 2;
        mu = avfF * mt.mu[0] + (1 - avfF) * mt.mu[1] + rhoF * eddy;
      }
      T fm[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) fm[i] = (mu * st[i][0]) * n[0] + (mu * st[i][1]) * n[1] + (mu * st[i][2]) * n[2];
      T vf[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) vf[i] = (L[iu + i] + R1[iu + i]) / 2;
      T fe[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        T stv = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) stv += st[i][d] * vf[d];
        fe[i] = (mu * stv) * n[i];
      }
      const T en = fe[0] + fe[1] + fe[2];
#pragma unroll
      for (int i = 0; i < 3; ++i) rv[i] += (double)(own ? fm[i] * A : -(fm[i]) * A);
      rv[3] += (double)(own ? en * A : -(en) * A);
      if constexpr (HEAT) {
        const T k = mt.cp[0] * mu / mt.Pr0;
        T gf[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) gf[d] = (G0[9 + d] + G1[9 + d]) / 2;
        const T q = k * (gf[0] * n[0] + gf[1] * n[1] + gf[2] * n[2]);
        rv[4] += (double)(own ? q * A : -(q) * A);
      }
    }
  }
  if constexpr (VISC) {
#pragma unroll
    for (int i = 0; i < 3; ++i) rhs[iu + i] += rv[i];
    rhs[NEQ == 5 ? 4 : 8] += rv[3];
    if constexpr (HEAT) rhs[4] += rv[4];
  }
  if constexpr (MODE != 0) {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      // MODE 1 (the dt stage), mixed precision: the stored residual is FP32,
      // promoted in the update (the MP policy); MODE 2 stores FP64 so that
      // the split stages equal the fused ones bit for bit
      if constexpr (MODE == 1 && sizeof(T) == 4) reinterpret_cast<float *>(rhs_out)[c * NEQ + v] = (float)rhs[v];
      else rhs_out[c * NEQ + v] = rhs[v];
    }
    if constexpr (MODE == 1) dtc[c] = volcfl[c] / (double)sig;
  } else {
    const double dt = dt_ptr ? *dt_ptr : dt_val;
    const double Vc = vol[c];
    constexpr int iz = NEQ == 5 ? 3 : 5;
    double un[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; ++v) {
      const int neglect = (v == iz && P.neglectZ == 1);
      double u;
      if (P.rkSteps == 1 || irk == 0) u = (1 - neglect) * (base[v] + dt * (1.0 / Vc) * rhs[v]);
      else if (irk == 1) u = (1 - neglect) * (base[v] + (1.0 / 4.0) * dt * (1.0 / Vc) * rhs[v]);
      else u = (1 - neglect) * (base[v] + (2.0 / 3.0) * dt * (1.0 / Vc) * rhs[v]);
      if constexpr (RLX == 0) Unew[c * NEQ + v] = u;
      un[v] = u;
    }
    if constexpr (RLX == 2) {
      int it = 0, ts = 0;
      const int chem = fk::plain_cell(P, un, &it, &ts);
#pragma unroll
      for (int v = 0; v < NEQ; ++v) Unew[c * NEQ + v] = un[v];
      if (Wout && !chem) {
        fk::Mat<double> mw = md;
        asm volatile("" : "+s"(mw.g[0]), "+s"(mw.g[1]), "+s"(mw.pi[0]), "+s"(mw.pi[1]), "+s"(mw.eta[0]),
                     "+s"(mw.eta[1]), "+s"(mw.avfMin));
        fk::store_prims<NEQ>(mw, un, Wout + c * NEQ);
      }
      // chemical cells to the list of relax_chem, wave-aggregated; counts
      // summed bit by bit from ballots (no atomics on the counters)
      const int l = threadIdx.x & 63;
      const unsigned long long mp = __ballot(chem);
      if (mp) {
        const int leader = __ffsll((long long)mp) - 1;
        unsigned long long b = 0;
        if (l == leader) b = atomicAdd(&rcnt[5], (unsigned long long)__popcll(mp));
        b = __shfl(b, leader, 64);
        if (chem) rlist[b + __popcll(mp & ((1ull << l) - 1ull))] = (uint32_t)c;
      }
      unsigned long long nit = 0, nts = 0;
      for (int k = 0; __ballot((it >> k) != 0); ++k) nit += (unsigned long long)__popcll(__ballot((it >> k) & 1)) << k;
      for (int k = 0; __ballot((ts >> k) != 0); ++k) nts += (unsigned long long)__popcll(__ballot((ts >> k) & 1)) << k;
      if (l == 0) {
        rpart[(c >> 6) * 3] = (unsigned long long)__popcll(mp);
        rpart[(c >> 6) * 3 + 1] = nit;
        rpart[(c >> 6) * 3 + 2] = nts;
      }
    }
    if constexpr (RLX == 1) {
      int chem = 0, it = 0;
      const int push = fk::screen_cell(P, un, &chem, &it);
#pragma unroll
      for (int v = 0; v < NEQ; ++v) Unew[c * NEQ + v] = un[v];
      if (Wout && !push) {
        fk::Mat<double> mw = md;
        asm volatile("" : "+s"(mw.g[0]), "+s"(mw.g[1]), "+s"(mw.pi[0]), "+s"(mw.pi[1]), "+s"(mw.eta[0]),
                     "+s"(mw.eta[1]), "+s"(mw.avfMin));
        fk::store_prims<NEQ>(mw, un, Wout + c * NEQ);
      }
      // wave-aggregated append to the work list, and the wave's counts
      // without atomics: it <= 4, so three ballots give the iteration sum
      const int l = threadIdx.x & 63;
      fk::list_push(push, chem, c, rlist, nc, rcnt);
      static_assert(fk::kColdIters < 8, "iteration counts are summed from three ballots");
      const unsigned long long nch = __popcll(__ballot(chem));
      const unsigned long long nit = __popcll(__ballot(it & 1)) + 2ull * __popcll(__ballot(it & 2)) +
                                     4ull * __popcll(__ballot(it & 4));
      if (l == 0) {
        rpart[(c >> 6) * 2] = nch;
        rpart[(c >> 6) * 2 + 1] = nit;
      }
    }
    // mixed precision: FP32 primitive record of the new state for pass G.  The
    // EOS constants are re-derived from the (scalar) parameters here: shared
    // with the cell pressure above, they would otherwise be held in vector
    // registers across the face loop and spill.
    if constexpr (WOUT && RLX == 0) {
      fk::Mat<double> mw = md;
      asm volatile("" : "+s"(mw.g[0]), "+s"(mw.g[1]), "+s"(mw.pi[0]), "+s"(mw.pi[1]), "+s"(mw.eta[0]), "+s"(mw.eta[1]),
                   "+s"(mw.avfMin));
      fk::store_prims<NEQ>(mw, un, Wout + c * NEQ);
    }
  }
}

// RK stage from a stored residual (update.cu:19-103); RT: its storage type.
template <int NEQ, typename RT>
__global__ void __launch_bounds__(kBlock) rk_update(fk_params P, int64_t nc, const double *__restrict__ U,
                                                    const double *__restrict__ U0, double *__restrict__ Unew,
                                                    const RT *__restrict__ rhs_s, const double *__restrict__ vol,
                                                    int irk, double dt_val, const double *__restrict__ dt_ptr) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nc * NEQ) return;
  const double ri = (double)rhs_s[i];
  const int64_t c = i / NEQ;
  const int v = (int)(i - c * NEQ);
  const double dt = dt_ptr ? *dt_ptr : dt_val;
  const double Vc = vol[c];
  constexpr int iz = NEQ == 5 ? 3 : 5;
  const int neglect = (v == iz && P.neglectZ == 1);
  double u;
  if (P.rkSteps == 1) u = (1 - neglect) * (U[i] + dt * (1.0 / Vc) * ri);
  else if (irk == 0) u = (1 - neglect) * (U0[i] + dt * (1.0 / Vc) * ri);
  else if (irk == 1) u = (1 - neglect) * ((3.0 / 4.0) * U0[i] + (1.0 / 4.0) * U[i] + (1.0 / 4.0) * dt * (1.0 / Vc) * ri);
  else u = (1 - neglect) * ((1.0 / 3.0) * U0[i] + (2.0 / 3.0) * U[i] + (2.0 / 3.0) * dt * (1.0 / Vc) * ri);
  Unew[i] = u;
}

// The same RK stage one cell per thread (16-byte record loads and stores),
// also writing the FP32 primitive record of the new state for pass G (mixed
// precision).  Same expressions as rk_update, so the same bits.
template <int NEQ, typename RT>
__global__ void __launch_bounds__(kBlock) rk_update_w(fk_params P, int64_t nc, const double *__restrict__ U,
                                                      const double *__restrict__ U0, double *__restrict__ Unew,
                                                      const RT *__restrict__ rhs, const double *__restrict__ vol,
                                                      int irk, double dt_val, const double *__restrict__ dt_ptr,
                                                      float *__restrict__ W) {
  // The new state and its primitive record leave through LDS: a thread's
  // 80-byte (and 40-byte) record stored directly goes out as strided partial
  // lines; from LDS the workgroup stores its contiguous ranges in 16-byte
  // (state) and 8-byte (record) words.  Threads past the last cell compute
  // the last cell again and write only their own LDS rows.
  constexpr int SU = NEQ % 2 ? NEQ : NEQ + 1;  // doubles per LDS row: an odd count of 8-byte words
  constexpr int SWR = NEQ % 2 ? NEQ : NEQ + 1; // floats per record row: an odd count
  __shared__ __attribute__((aligned(16))) double sU[kBlock * SU];
  __shared__ __attribute__((aligned(16))) float sWr[kBlock * SWR];
  const int64_t c0 = (int64_t)blockIdx.x * kBlock;
  const int64_t c = c0 + threadIdx.x < nc ? c0 + threadIdx.x : nc - 1;
  const double dt = dt_ptr ? *dt_ptr : dt_val;
  const double Vc = vol[c];
  constexpr int iz = NEQ == 5 ? 3 : 5;
  double r[NEQ], uc[NEQ], u0[NEQ], un[NEQ];
  if constexpr (sizeof(RT) == 8) {
    fk::load_rec<NEQ>(rhs + c * NEQ, r);
  } else {
#pragma unroll
    for (int v = 0; v < NEQ; ++v) r[v] = (double)rhs[c * NEQ + v];
  }
  if (P.rkSteps == 1 || irk != 0) fk::load_rec<NEQ>(U + c * NEQ, uc);
  if (P.rkSteps != 1) fk::load_rec<NEQ>(U0 + c * NEQ, u0);
#pragma unroll
  for (int v = 0; v < NEQ; ++v) {
    const int neglect = (v == iz && P.neglectZ == 1);
    double u;
    if (P.rkSteps == 1) u = (1 - neglect) * (uc[v] + dt * (1.0 / Vc) * r[v]);
    else if (irk == 0) u = (1 - neglect) * (u0[v] + dt * (1.0 / Vc) * r[v]);
    else if (irk == 1) u = (1 - neglect) * ((3.0 / 4.0) * u0[v] + (1.0 / 4.0) * uc[v] + (1.0 / 4.0) * dt * (1.0 / Vc) * r[v]);
    else u = (1 - neglect) * ((1.0 / 3.0) * u0[v] + (2.0 / 3.0) * uc[v] + (2.0 / 3.0) * dt * (1.0 / Vc) * r[v]);
    un[v] = u;
  }
#pragma unroll
  for (int v = 0; v < NEQ; ++v) sU[threadIdx.x * SU + v] = un[v];
  fk::store_prims<NEQ>(fk::make_mat<double>(P), un, sWr + threadIdx.x * SWR);
  __syncthreads();
  const int ncell = (int)((nc - c0) < kBlock ? (nc - c0) : kBlock);
  const int nu = ncell * NEQ;
  double *du = Unew + c0 * NEQ;
  for (int w = threadIdx.x; w < nu; w += kBlock) {
    const int cell = w / NEQ;
    du[w] = sU[cell * SU + (w - cell * NEQ)];
  }
  float *dw = W + c0 * NEQ;
  for (int w = threadIdx.x; w < nu; w += kBlock) {
    const int cell = w / NEQ;
    dw[w] = sWr[cell * SWR + (w - cell * NEQ)];
  }
}

// min over cells of Vcfl/signal: per-block minima, then one block folds them
// and forms dt = min(cfl * min, maxDt) on the device.
__global__ void __launch_bounds__(kBlock) min_blocks(int64_t n, const double *__restrict__ x, double *__restrict__ out) {
  __shared__ double sh[kBlock];
  double m = 1e15;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) m = fmin(m, x[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) dt_final(int nb, const double *__restrict__ part, double cfl, double maxDt,
                                                   double *__restrict__ scal) {
  __shared__ double sh[kBlock];
  double m = 1e15;
  for (int i = threadIdx.x; i < nb; i += kBlock) m = fmin(m, part[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[0] = fmin(cfl * sh[0], maxDt);
}

// physical time += dt, once per step
__global__ void advance_time(double *__restrict__ scal) {
  if (threadIdx.x == 0) scal[1] += scal[0];
}

// relaxation (relaxation.cu:77-114), reference Newton solves, in two passes:
//   relax_therm: partitionRelaxation (Tsat Newton, relaxation.cu:45-75) of
//                every cell; thermal cells run thermalRlx (thermRlx.cu:190)
//                in place, chemical cells go to a work list
//   relax_chem:  chemRlx (chemRlx.cu:217) on the work list
// Each cell is solved independently with the reference arithmetic, so the
// result does not depend on the list order; the split keeps the register
// footprint of the chemical solve out of the thermal pass.
// counts: [n_chem, newton iterations, tsat iterations, -, -, list length].
using fk::clampd;
using fk::eos_p;
using fk::eos_T;
using fk::tsat_dev;
using fk::therm_dev;
using fk::chem_dev;

__global__ void __launch_bounds__(kBlock) relax_therm(fk_params P, int64_t nc, double *__restrict__ U,
                                                      uint32_t *__restrict__ list,
                                                      unsigned long long *__restrict__ counts,
                                                      unsigned long long *__restrict__ part, float *__restrict__ W) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int l = threadIdx.x & 63;
  int chem = 0, it = 0, ts = 0;
  if (c < nc) {
    double u[10];
    fk::load_rec<10>(U + c * 10, u);
    chem = fk::plain_cell(P, u, &it, &ts);
    if (!chem) {
      U[c * 10 + 0] = u[0];
      U[c * 10 + 6] = u[6];
      U[c * 10 + 7] = u[7];
      if (W) fk::store_prims<10>(fk::make_mat<double>(P), u, W + c * 10);  // chemical cells: relax_chem
    }
  }
  const unsigned long long mask = __ballot(chem);
  const int npush = __popcll(mask);
  if (npush) {
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (l == leader) base = atomicAdd(&counts[5], (unsigned long long)npush);
    base = __shfl(base, leader, 64);
    if (chem) list[base + __popcll(mask & ((1ull << l) - 1ull))] = (uint32_t)c;
  }
  unsigned long long v[3] = {(unsigned long long)chem, (unsigned long long)it, (unsigned long long)ts};
  fk::block_partials<3>(v, part);  // [n_chem, newton iterations, tsat iterations]
}

__global__ void __launch_bounds__(kBlock) relax_chem(fk_params P, double *__restrict__ U,
                                                     const uint32_t *__restrict__ list,
                                                     unsigned long long *__restrict__ counts, float *__restrict__ W) {
  const int64_t n = (int64_t)counts[5];
  unsigned long long nit = 0;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    const int64_t c = list[e];
    double u[10];
    fk::load_rec<10>(U + c * 10, u);
    nit += chem_dev(P, u);
    U[c * 10 + 0] = u[0];
    U[c * 10 + 1] = u[1];
    U[c * 10 + 2] = u[2];
    U[c * 10 + 6] = u[6];
    U[c * 10 + 7] = u[7];
    U[c * 10 + 9] = u[9];
    if (W) fk::store_prims<10>(fk::make_mat<double>(P), u, W + c * 10);
  }
  for (int off = 32; off > 0; off >>= 1) nit += __shfl_down(nit, off, 64);
  if ((threadIdx.x & 63) == 0 && nit) atomicAdd(&counts[1], nit);
}

// sum_c U_i V_c per block (consCalc.cu:3-34); i = 0 integrates 1 - alpha.
template <int NEQ>
__global__ void __launch_bounds__(kBlock) integrate(int64_t nc, const double *__restrict__ U,
                                                    const double *__restrict__ vol, double *__restrict__ part) {
  __shared__ double sh[kBlock];
  for (int v = 0; v < NEQ; ++v) {
    double s = 0;
    for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < nc; c += (int64_t)gridDim.x * kBlock) {
      double u = U[c * NEQ + v];
      if (NEQ == 10 && v == 0) u = 1 - u;
      s += u * vol[c];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * NEQ + v] = sh[0];
    __syncthreads();
  }
}

// Folds the per-block partials of integrate in block order, one thread per
// variable: out[v] = ((p[0][v] + p[1][v]) + ...), the order of the host sum
// of fk_cons_calc, so both give the same bits.
template <int NEQ>
__global__ void fold_partials(int nb, const double *__restrict__ part, double *__restrict__ out) {
  const int v = threadIdx.x;
  if (v >= NEQ) return;
  double s = 0;
  for (int b = 0; b < nb; ++b) s += part[b * NEQ + v];
  out[v] = s;
}

__global__ void __launch_bounds__(kBlock) stream_copy(int64_t n16, const uint4 *__restrict__ a, uint4 *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += (int64_t)gridDim.x * kBlock) b[i] = a[i];
}

__global__ void zero_ull(unsigned long long *p, int n) {
  int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// relaxation totals: counts[8 + i] += counts[i], i < 5
__global__ void accum_counts(unsigned long long *p) {
  int i = threadIdx.x;
  if (i < 5) p[8 + i] += p[i];
}

// ------------------------------------------------------------------------
// host helpers
// ------------------------------------------------------------------------

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
// grid of the XCD-remapped kernels: a multiple of 8 blocks
inline unsigned grid_xcd(int64_t n) { return (grid_for(n) + kXcd - 1) / kXcd * kXcd; }
inline size_t tsz(const fk_ctx *c) { return c->p.precision == FK_MIXED ? 4 : 8; }
inline int ng_of(const fk_ctx *c) { return 9 + ((c->p.model == FK_MODEL_SINGLE && c->p.heatFlux == 1) ? 3 : 0); }

int next_buffer(const fk_ctx *c) {
  for (int i = 0; i < 3; ++i)
    if (i != c->ia && i != c->iu0) return i;
  return (c->ia + 1) % 3;
}

int require(const fk_ctx *c) {
  if (!c) return fail(FK_EINVAL, "null context");
  if (!c->have_mesh) return fail(FK_ESTATE, "no mesh uploaded");
  if (!c->have_state) return fail(FK_ESTATE, "no state uploaded");
  return FK_OK;
}

// upload the precision-dependent slot geometry
int upload_geometry(fk_ctx *c) {
  const int64_t ns = c->nc * c->ns;
  if (c->p.precision == FK_MIXED) {
    std::vector<float> g((size_t)ns * 4), r((size_t)ns * 4), l((size_t)ns);
    for (size_t i = 0; i < g.size(); ++i) g[i] = (float)c->h_geo[i];
    for (size_t i = 0; i < r.size(); ++i) {
      if (i % 4 == 3) {  // the slot link, kept bit-exact
        const int32_t lk = (int32_t)c->h_r[i];
        std::memcpy(&r[i], &lk, 4);
      } else {
        r[i] = (float)c->h_r[i];
      }
    }
    for (size_t i = 0; i < l.size(); ++i) l[i] = (float)c->h_ls[i];
    HIPCHK(hipMemcpy(c->geo, g.data(), g.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->rvec, r.data(), r.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->wls, l.data(), l.size() * 4, hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(c->geo, c->h_geo.data(), c->h_geo.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->rvec, c->h_r.data(), c->h_r.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->wls, c->h_ls.data(), c->h_ls.size() * 8, hipMemcpyHostToDevice));
  }
  return FK_OK;
}

void free_all(fk_ctx *c) {
  void *ptrs[] = {c->link, c->geo, c->rvec, c->wls, c->vol, c->volcfl, c->U[0], c->U[1], c->U[2], c->face,
                  c->gvel, c->rhs, c->dtc, c->red, c->dscal, c->dgrad, c->dtheta, c->lcount, c->rlist,
                  c->rfall, c->rstart, c->rpart, c->wprim, c->health};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->link = nullptr;
  c->wprim = nullptr;
  c->health = nullptr;
  c->w_valid = false;
  c->rlist = nullptr;
  c->rfall = nullptr;
  c->rstart = nullptr;
  c->rstart_cap = 0;
  c->rpart = nullptr;
  c->rlist_cap = 0;
  c->geo = c->rvec = c->wls = nullptr;
  c->vol = c->volcfl = nullptr;
  c->U[0] = c->U[1] = c->U[2] = nullptr;
  c->face = c->gvel = nullptr;
  c->rhs = c->dtc = c->red = c->dscal = c->dgrad = c->dtheta = nullptr;
  c->lcount = nullptr;
}

template <typename T, int NS>
int launch_pass_g_t(fk_ctx *c) {
  const unsigned gr = grid_xcd(c->nc);
  const bool visc = c->p.viscFlag == 1;
  const bool heat = visc && c->p.model == FK_MODEL_SINGLE && c->p.heatFlux == 1;
  double *dg = c->debug ? c->dgrad : nullptr;
  double *dth = c->debug ? c->dtheta : nullptr;
  const double *U = c->U[c->ia];
  T *face = (T *)c->face;
  T *gv = (T *)c->gvel;
  const bool fromw = sizeof(T) == 4 && c->prims_w && c->w_valid;
  // face records through LDS (workgroup sizes: pass_g_block)
  const bool sout = c->stage_g;
  constexpr int gbs = pass_g_block<T, NS, true>();
  const unsigned grs = (unsigned)(((c->nc + gbs - 1) / gbs + kXcd - 1) / kXcd * kXcd);
#define PG3(NEQ, VISC, HEAT, FW, SO)                                                                                  \
  LAUNCH_TIMED(c, "pass_g",                                                                                          \
               hipLaunchKernelGGL((pass_g<NEQ, T, VISC, HEAT, FW, NS, SO>), dim3(SO ? grs : gr),                      \
                                  dim3(SO ? gbs : kBlock), 0, c->stream,                                             \
                                  c->p, c->nc, U, c->wprim, c->link, c->geo, c->rvec, c->vol, face, gv, dg, dth,    \
                                  c->health, (int)std::min<int64_t>(c->step, INT32_MAX)))
#define PG2(NEQ, VISC, HEAT, FW)                                                                                     \
  do {                                                                                                               \
    if constexpr (pass_g_sout_ok<T, NS, NEQ>()) {                                                                    \
      if (sout) {                                                                                                    \
        PG3(NEQ, VISC, HEAT, FW, true);                                                                              \
        break;                                                                                                       \
      }                                                                                                              \
    }                                                                                                                \
    PG3(NEQ, VISC, HEAT, FW, false);                                                                                 \
  } while (0)
#define PG(NEQ, VISC, HEAT)                     \
  do {                                          \
    if constexpr (sizeof(T) == 4) {             \
      if (fromw) {                              \
        PG2(NEQ, VISC, HEAT, true);             \
        break;                                  \
      }                                         \
    }                                           \
    PG2(NEQ, VISC, HEAT, false);                \
  } while (0)
  if (c->neq == 5) {
    if (heat) PG(5, true, true);
    else if (visc) PG(5, true, false);
    else PG(5, false, false);
  } else {
    if (visc) PG(10, true, false);
    else PG(10, false, false);
  }
#undef PG
#undef PG2
#undef PG3
  return FK_OK;
}

template <typename T, int MODE, int NS>
int launch_pass_f_t(fk_ctx *c, int irk, double dt, const double *dtp, double *Unew, bool write_w = false,
                    int relax = 0) {
  const unsigned gr = grid_xcd(c->nc);
  const bool visc = c->p.viscFlag == 1;
  const bool heat = visc && c->p.model == FK_MODEL_SINGLE && c->p.heatFlux == 1;
  const double *U = c->U[c->ia], *U0 = c->U[c->iu0];
  const T *face = (const T *)c->face;
  const T *gv = (const T *)c->gvel;
  const T *wl = (const T *)c->wls;
  const char *nm = MODE == 0 ? "pass_f" : (MODE == 1 ? "pass_f_rhs" : "pass_f_res");
  float *Wout = (MODE == 0 && sizeof(T) == 4 && write_w) ? c->wprim : nullptr;
  // fused relaxation (last stage): records of the relaxed state, mixed precision
  float *Wr = (sizeof(T) == 4 && c->prims_w) ? c->wprim : nullptr;
  unsigned long long *cnt = (unsigned long long *)c->lcount;
  // LDS staging of face states and gradients (pass_f_stage_ok,
  // pass_f_block); with 6 slots not for the last stage's relaxation pass
  const bool stage = c->stage_f && (NS == 4 || relax == 0) &&
                     (c->neq == 10 ? pass_f_stage_ok<T, NS, 10>() : pass_f_stage_ok<T, NS, 5>());
  constexpr int bs = pass_f_block<T, true, NS>();
  const unsigned gs = (unsigned)(((c->nc + bs - 1) / bs + kXcd - 1) / kXcd * kXcd);
  const size_t lds = stage ? (size_t)bs * (NS * c->neq + 12) * sizeof(T) : 0;
  // waves of this launch (per-wave relaxation counts of the fused last stage)
  c->pf_waves = (int64_t)(stage ? (int64_t)gs * (bs / 64) : (int64_t)gr * (kBlock / 64));
#define PF_ARGS(W) \
  c->p, c->nc, U, U0, Unew, c->link, c->geo, wl, c->vol, c->volcfl, face, gv, c->rhs, c->dtc, irk, dt, dtp, W
#define PF(NEQ, VISC, HEAT)                                                                                            \
  do {                                                                                                                 \
    if constexpr (MODE == 0 && NEQ == 10) {                                                                            \
      if (relax == 1) {                                                                                                \
        if (stage)                                                                                                     \
          LAUNCH_TIMED(c, "pass_f_relax",                                                                              \
                       hipLaunchKernelGGL((pass_f<NEQ, T, VISC, HEAT, MODE, (NS == 4), NS, false, 1>), dim3(gs), dim3(bs),       \
                                          lds, c->stream, PF_ARGS(Wr), c->rlist, cnt, c->rpart));                      \
        else                                                                                                           \
          LAUNCH_TIMED(c, "pass_f_relax",                                                                              \
                       hipLaunchKernelGGL((pass_f<NEQ, T, VISC, HEAT, MODE, false, NS, false, 1>), dim3(gr), dim3(kBlock),  \
                                          0, c->stream, PF_ARGS(Wr), c->rlist, cnt, c->rpart));                        \
        break;                                                                                                         \
      }                                                                                                                \
      if (relax == 2) {                                                                                                \
        if (stage)                                                                                                     \
          LAUNCH_TIMED(c, "pass_f_relax",                                                                              \
                       hipLaunchKernelGGL((pass_f<NEQ, T, VISC, HEAT, MODE, (NS == 4), NS, false, 2>), dim3(gs), dim3(bs),       \
                                          lds, c->stream, PF_ARGS(Wr), c->rlist, cnt, c->rpart));                      \
        else                                                                                                           \
          LAUNCH_TIMED(c, "pass_f_relax",                                                                              \
                       hipLaunchKernelGGL((pass_f<NEQ, T, VISC, HEAT, MODE, false, NS, false, 2>), dim3(gr), dim3(kBlock),  \
                                          0, c->stream, PF_ARGS(Wr), c->rlist, cnt, c->rpart));                        \
        break;                                                                                                         \
      }                                                                                                                \
    }                                                                                                                  \
    if constexpr (MODE == 0 && sizeof(T) == 4) {                                                                       \
      if (Wout) {                                                                                                      \
        if (stage)                                                                                                     \
          LAUNCH_TIMED(c, nm, hipLaunchKernelGGL((pass_f<NEQ, T, VISC, HEAT, MODE, (pass_f_stage_ok<T, NS, NEQ>()), NS, true>), dim3(gs),             \
                                                 dim3(bs), lds, c->stream, PF_ARGS(Wout), nullptr, nullptr,             \
                                                 nullptr));                                                             \
        else                                                                                                           \
          LAUNCH_TIMED(c, nm, hipLaunchKernelGGL((pass_f<NEQ, T, VISC, HEAT, MODE, false, NS, true>), dim3(gr),            \
                                                 dim3(kBlock), 0, c->stream, PF_ARGS(Wout), nullptr, nullptr,           \
                                                 nullptr));                                                             \
        break;                                                                                                         \
      }                                                                                                                \
    }                                                                                                                  \
    if (stage)                                                                                                         \
      LAUNCH_TIMED(c, nm, hipLaunchKernelGGL((pass_f<NEQ, T, VISC, HEAT, MODE, (pass_f_stage_ok<T, NS, NEQ>()), NS>), dim3(gs), dim3(bs), lds,         \
                                             c->stream, PF_ARGS(Wout), nullptr, nullptr, nullptr));                     \
    else                                                                                                               \
      LAUNCH_TIMED(c, nm, hipLaunchKernelGGL((pass_f<NEQ, T, VISC, HEAT, MODE, false, NS>), dim3(gr), dim3(kBlock), 0,      \
                                             c->stream, PF_ARGS(Wout), nullptr, nullptr, nullptr));                     \
  } while (0)
  if (c->neq == 5) {
    if (heat) PF(5, true, true);
    else if (visc) PF(5, true, false);
    else PF(5, false, false);
  } else {
    if (visc) PF(10, true, false);
    else PF(10, false, false);
  }
#undef PF
#undef PF_ARGS
  return FK_OK;
}

// the kernels for the mesh's slots per cell (4 or 6)
template <typename T>
int launch_pass_g_ns(fk_ctx *c) {
  return c->ns == 4 ? launch_pass_g_t<T, 4>(c) : launch_pass_g_t<T, 6>(c);
}
template <typename T, int MODE>
int launch_pass_f_ns(fk_ctx *c, int irk, double dt, const double *dtp, double *Unew, bool write_w = false,
                     int relax = 0) {
  return c->ns == 4 ? launch_pass_f_t<T, MODE, 4>(c, irk, dt, dtp, Unew, write_w, relax)
                    : launch_pass_f_t<T, MODE, 6>(c, irk, dt, dtp, Unew, write_w, relax);
}

int do_pass_g(fk_ctx *c) {
  return c->p.precision == FK_MIXED ? launch_pass_g_ns<float>(c) : launch_pass_g_ns<double>(c);
}

int do_pass_f_rhs(fk_ctx *c) {
  int rc = c->p.precision == FK_MIXED ? launch_pass_f_ns<float, 1>(c, 0, 0.0, nullptr, nullptr)
                                      : launch_pass_f_ns<double, 1>(c, 0, 0.0, nullptr, nullptr);
  if (rc) return rc;
  c->rhs_f32 = c->p.precision == FK_MIXED;
  c->rhs_valid = true;
  return FK_OK;
}

int do_update(fk_ctx *c, int irk, double dt, const double *dtp, bool want_w);

// RK stage irk from pass F: fused (MODE 0) or residual (MODE 2) + rk_update,
// as c->fuse_update selects; both give the same bits.  want_w: the fused
// kernel also writes the primitive records of the new state (mixed precision).
// relax (fused form only): the kernel also runs stage 1 of the surrogate
// relaxation (fk_relax_mlp_begin before, fk_relax_mlp_finish after).
int do_pass_f_update(fk_ctx *c, int irk, double dt, const double *dtp, bool want_w, int relax = 0) {
  if (relax) {
    const int nb = next_buffer(c);
    int rc = c->p.precision == FK_MIXED ? launch_pass_f_ns<float, 0>(c, irk, dt, dtp, c->U[nb], false, relax)
                                        : launch_pass_f_ns<double, 0>(c, irk, dt, dtp, c->U[nb], false, relax);
    if (rc) return rc;
    c->ia = nb;
    c->w_valid = false;  // until stage 2 has written the listed cells' records
    c->rhs_valid = false;
    return FK_OK;
  }
  if (!c->fuse_update) {
    int rc = c->p.precision == FK_MIXED ? launch_pass_f_ns<float, 2>(c, 0, 0.0, nullptr, nullptr)
                                        : launch_pass_f_ns<double, 2>(c, 0, 0.0, nullptr, nullptr);
    if (rc) return rc;
    c->rhs_f32 = false;
    rc = do_update(c, irk, dt, dtp, want_w);
    c->rhs_valid = false;
    return rc;
  }
  const int nb = next_buffer(c);
  const bool w = want_w && c->p.precision == FK_MIXED && c->prims_w;
  int rc = c->p.precision == FK_MIXED ? launch_pass_f_ns<float, 0>(c, irk, dt, dtp, c->U[nb], w)
                                      : launch_pass_f_ns<double, 0>(c, irk, dt, dtp, c->U[nb], false);
  if (rc) return rc;
  c->ia = nb;
  c->w_valid = w;
  c->rhs_valid = false;
  return FK_OK;
}

// RK stage from the stored residual (FP32 after the mixed dt stage, c->rhs_f32)
template <int NEQ, typename RT>
int launch_update_t(fk_ctx *c, int irk, double dt, const double *dtp, int nb, bool w) {
  const RT *r = (const RT *)c->rhs;
  if (w)
    LAUNCH_TIMED(c, "rk_update", hipLaunchKernelGGL((rk_update_w<NEQ, RT>), dim3(grid_for(c->nc)), dim3(kBlock), 0,
                                                    c->stream, c->p, c->nc, c->U[c->ia], c->U[c->iu0], c->U[nb], r,
                                                    c->vol, irk, dt, dtp, c->wprim));
  else
    LAUNCH_TIMED(c, "rk_update", hipLaunchKernelGGL((rk_update<NEQ, RT>), dim3(grid_for(c->nc * NEQ)), dim3(kBlock), 0,
                                                    c->stream, c->p, c->nc, c->U[c->ia], c->U[c->iu0], c->U[nb], r,
                                                    c->vol, irk, dt, dtp));
  return FK_OK;
}

int do_update(fk_ctx *c, int irk, double dt, const double *dtp, bool want_w) {
  const int nb = next_buffer(c);
  const bool w = want_w && c->p.precision == FK_MIXED && c->prims_w;
  int rc;
  if (c->neq == 5)
    rc = c->rhs_f32 ? launch_update_t<5, float>(c, irk, dt, dtp, nb, w) : launch_update_t<5, double>(c, irk, dt, dtp, nb, w);
  else
    rc = c->rhs_f32 ? launch_update_t<10, float>(c, irk, dt, dtp, nb, w)
                    : launch_update_t<10, double>(c, irk, dt, dtp, nb, w);
  if (rc) return rc;
  c->ia = nb;
  c->w_valid = w;
  return FK_OK;
}

int do_dt_device(fk_ctx *c) {
  const int nb = 1024;
  LAUNCH_TIMED(c, "dt_reduce",
               hipLaunchKernelGGL(min_blocks, dim3(nb), dim3(kBlock), 0, c->stream, c->nc, c->dtc, c->red);
               hipLaunchKernelGGL(dt_final, dim3(1), dim3(kBlock), 0, c->stream, nb, c->red, c->p.cfl, c->p.maxDt,
                                  c->dscal));
  return FK_OK;
}

// stats of the counters at slots [base, base + 5) of c->lcount
int read_counts(fk_ctx *c, int base, fk_relax_stats *st) {
  unsigned long long h[5];
  HIPCHK(hipMemcpyAsync(h, c->lcount + base, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  st->n_chem = (int64_t)h[0];
  st->newton_iters = (int64_t)h[1];
  st->tsat_iters = (int64_t)h[2];
  st->n_fallback = (int64_t)h[3];
  st->n_warm = (int64_t)h[4];
  st->n_therm = 0;  // set by the caller (cells - n_chem, summed over calls)
  return FK_OK;
}

}  // namespace

// work list of the relaxation passes (one uint32 per cell), allocated on first use
// Network starts (float2 per work-list entry) of the surrogate relaxation.
int ensure_rstart(fk_ctx *c) {
  if (c->rstart_cap >= c->nc) return FK_OK;
  if (c->rstart) (void)hipFree(c->rstart);
  c->rstart = nullptr;
  c->rstart_cap = 0;
  HIPCHK(hipMalloc(&c->rstart, (size_t)c->nc * 2 * sizeof(float)));
  c->rstart_cap = c->nc;
  return FK_OK;
}

int ensure_rlist(fk_ctx *c) {
  if (c->nc >= (int64_t(1) << 31)) return fail(FK_ERANGE, "relaxation supports < 2^31 cells");
  if (c->rlist_cap >= c->nc) return FK_OK;
  if (c->rlist) (void)hipFree(c->rlist);
  if (c->rfall) (void)hipFree(c->rfall);
  if (c->rpart) (void)hipFree(c->rpart);
  c->rlist = nullptr;
  c->rfall = nullptr;
  c->rpart = nullptr;
  c->rlist_cap = 0;
  HIPCHK(hipMalloc(&c->rlist, (size_t)c->nc * sizeof(uint32_t)));
  HIPCHK(hipMalloc(&c->rfall, (size_t)c->nc * sizeof(uint32_t)));
  // [blocks][4] partials of the relaxation kernels, or [waves][2 or 3] of the
  // fused pass F (grid_xcd blocks of 4 waves)
  HIPCHK(hipMalloc(&c->rpart, (size_t)grid_xcd(c->nc) * 12 * sizeof(unsigned long long)));
  c->rlist_cap = c->nc;
  return FK_OK;
}

namespace {

// Checks and scratch of a relaxation of mode `mode`, done before a loop of
// steps starts (the same checks run again inside each relaxation).
int relax_preflight(fk_ctx *c, int mode) {
  if (mode == FK_RELAX_MLP && !c->mlp.ready)
    return fail(FK_ESTATE, "surrogate relaxation needs trained or loaded MLP weights");
  if (mode == FK_RELAX_MLP || mode == FK_RELAX_COLD) {
    const fk_params &P = c->p;
    if (P.chemFlag == 1 && !((P.eta[0] - P.eta[1]) <= 0 && (P.Cp[1] - P.Cp[0]) >= 0))
      return fail(FK_EINVAL, "surrogate classification needs eta_L <= eta_G and Cp_G >= Cp_L");
  }
  int rc = ensure_rlist(c);
  if (rc) return rc;
  return mode == FK_RELAX_MLP ? ensure_rstart(c) : FK_OK;
}

int do_relax(fk_ctx *c, fk_relax_stats *st, int mode = FK_RELAX_PLAIN) {
  if (c->p.model != FK_MODEL_TWO_PHASE) return fail(FK_EINVAL, "relaxation needs the two-phase model");
  unsigned long long *cnt = (unsigned long long *)c->lcount;
  if (mode == FK_RELAX_MLP || mode == FK_RELAX_COLD) {
    if (c->p.thermFntRlxFlag == 1) return fail(FK_EINVAL, "finite-rate relaxation needs FK_RELAX_PLAIN");
    int rc = fk_relax_mlp_impl(c, mode == FK_RELAX_MLP);
    if (rc) return rc;
  } else {
    if (!st) {  // plain relaxation that also emits training pairs, when enabled
      bool done = false;
      int rc = fk_relax_emit_impl(c, &done);
      if (rc || done) {
        c->w_valid = false;  // the emitting relaxation writes no primitive records
        return rc;
      }
    }
    int rc = ensure_rlist(c);
    if (rc) return rc;
    LAUNCH_TIMED(c, "zero", hipLaunchKernelGGL(zero_ull, dim3(1), dim3(64), 0, c->stream, cnt, 6));
    // mixed precision: the relaxation also writes the primitive records of pass G
    float *W = (c->p.precision == FK_MIXED && c->prims_w) ? c->wprim : nullptr;
    LAUNCH_TIMED(c, "relax", hipLaunchKernelGGL(relax_therm, dim3(grid_for(c->nc)), dim3(kBlock), 0, c->stream, c->p,
                                                c->nc, c->U[c->ia], c->rlist, cnt, c->rpart, W));
    LAUNCH_TIMED(c, "relax_count", hipLaunchKernelGGL(fk::reduce_partials<3>, dim3(fk::reduce_grid(grid_for(c->nc))),
                                                      dim3(256), 0, c->stream, c->rpart, (int64_t)grid_for(c->nc), cnt));
    const unsigned gch = (unsigned)std::max<int64_t>(1, std::min<int64_t>(grid_for(c->nc), 1024));
    LAUNCH_TIMED(c, "relax_chem", hipLaunchKernelGGL(relax_chem, dim3(gch), dim3(kBlock), 0, c->stream, c->p,
                                                     c->U[c->ia], c->rlist, cnt, W));
    c->w_valid = W != nullptr;
  }
  if (st) {
    int rc = read_counts(c, 0, st);
    if (rc) return rc;
    st->n_therm = c->nc - st->n_chem;
  } else {
    hipLaunchKernelGGL(accum_counts, dim3(1), dim3(64), 0, c->stream, cnt);
    HIPCHK(hipGetLastError());
    ++c->relax_calls;
  }
  return FK_OK;
}

// After a fused pass F ran stage 1 of the surrogate relaxation: its per-wave
// counts, stage 2 and the running totals (as do_relax).
int relax_fused_finish(fk_ctx *c, bool warm) {
  unsigned long long *cnt = (unsigned long long *)c->lcount;
  const int64_t nw = c->pf_waves;
  LAUNCH_TIMED(c, "relax_count", hipLaunchKernelGGL(fk::reduce_partials<2>, dim3(fk::reduce_grid(nw)), dim3(256), 0,
                                                    c->stream, c->rpart, nw, cnt));
  int rc = fk_relax_mlp_finish(c, warm);
  if (rc) return rc;
  hipLaunchKernelGGL(accum_counts, dim3(1), dim3(64), 0, c->stream, cnt);
  HIPCHK(hipGetLastError());
  ++c->relax_calls;
  return FK_OK;
}

// The plain relaxation inside the last fused pass F: zeroed counts before it,
// then its per-wave counts, the chemical solve of the listed cells and the
// running totals (as do_relax).
int relax_plain_begin(fk_ctx *c) {
  int rc = ensure_rlist(c);
  if (rc) return rc;
  unsigned long long *cnt = (unsigned long long *)c->lcount;
  LAUNCH_TIMED(c, "zero", hipLaunchKernelGGL(zero_ull, dim3(1), dim3(64), 0, c->stream, cnt, 6));
  return FK_OK;
}

int relax_plain_fused_finish(fk_ctx *c) {
  unsigned long long *cnt = (unsigned long long *)c->lcount;
  const int64_t nw = c->pf_waves;
  LAUNCH_TIMED(c, "relax_count", hipLaunchKernelGGL(fk::reduce_partials<3>, dim3(fk::reduce_grid(nw)), dim3(256), 0,
                                                    c->stream, c->rpart, nw, cnt));
  float *W = (c->p.precision == FK_MIXED && c->prims_w) ? c->wprim : nullptr;
  const unsigned gch = (unsigned)std::max<int64_t>(1, std::min<int64_t>(grid_for(c->nc), 1024));
  LAUNCH_TIMED(c, "relax_chem", hipLaunchKernelGGL(relax_chem, dim3(gch), dim3(kBlock), 0, c->stream, c->p,
                                                   c->U[c->ia], c->rlist, cnt, W));
  c->w_valid = W != nullptr;
  hipLaunchKernelGGL(accum_counts, dim3(1), dim3(64), 0, c->stream, cnt);
  HIPCHK(hipGetLastError());
  ++c->relax_calls;
  return FK_OK;
}

int check_params(const fk_params *p) {
  if (!p) return fail(FK_EINVAL, "null params");
  if (p->model != FK_MODEL_SINGLE && p->model != FK_MODEL_TWO_PHASE) return fail(FK_EINVAL, "bad model");
  if (p->precision != FK_FP64 && p->precision != FK_MIXED) return fail(FK_EINVAL, "bad precision");
  if (p->rkSteps != 1 && p->rkSteps != 3) return fail(FK_EINVAL, "rkSteps must be 1 or 3");
  if (p->adaptDt == 1 && p->adaptDtItr < 1) return fail(FK_EINVAL, "adaptDtItr must be >= 1");
  if (p->relaxMode < FK_RELAX_PLAIN || p->relaxMode > FK_RELAX_COLD) return fail(FK_EINVAL, "bad relaxMode");
  if (p->thermFntRlxFlag != 0 && p->thermFntRlxFlag != 1) return fail(FK_EINVAL, "thermFntRlxFlag must be 0 or 1");
  if (p->thermFntRlxFlag == 1 && p->relaxMode != FK_RELAX_PLAIN)
    return fail(FK_EINVAL, "finite-rate relaxation needs relaxMode FK_RELAX_PLAIN");
  const bool inlet_default = p->inletP == 0 && p->inletT == 0;
  if (!inlet_default && !(p->inletT > 0 && std::isfinite(p->inletP) && std::isfinite(p->inletT)))
    return fail(FK_EINVAL, "inletT must be > 0 and inletP finite (or both 0 for the default inlet)");
  if (!std::isfinite(p->pRef) || std::fabs(p->pRef) > 3e38) return fail(FK_EINVAL, "pRef must be finite (FP32 range)");
  return FK_OK;
}

// The parameters as the kernels use them: the inlet state filled in with the
// reference's constants (bc.h:5-7 of either model) when both fields are 0.
fk_params resolved(const fk_params &p) {
  fk_params r = p;
  // one value for the FP64 conversions and the FP32 flux code
  r.pRef = (double)(float)p.pRef;
  if (r.inletP == 0 && r.inletT == 0) {
    const bool sp = r.model == FK_MODEL_SINGLE;
    r.inletP = sp ? 4.2 * 101300 : 11.4e5;
    r.inletT = sp ? 288.0 : 115.3;
  }
  return r;
}

}  // namespace

extern "C" {

const char *fk_last_error(void) { return g_err.c_str(); }

int fk_create(int device, const fk_params *p, fk_ctx **out) {
  if (!out) return fail(FK_EINVAL, "null out");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(FK_EDEVICE, "no such HIP device");
  HIPCHK(hipSetDevice(device));
  rc = scratch_setup(device);  // the no-allocation entry points' scratch on this device
  if (rc) return rc;
  fk_ctx *c = new (std::nothrow) fk_ctx;
  if (!c) return fail(FK_ENOMEM, "out of host memory");
  c->device = device;
  c->p = resolved(*p);
  c->neq = p->model == FK_MODEL_SINGLE ? 5 : 10;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(FK_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = c;
  return FK_OK;
}

int fk_destroy(fk_ctx *c) {
  if (!c) return FK_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto &pd : c->tm.pending) {
    (void)hipEventDestroy(pd.a);
    (void)hipEventDestroy(pd.b);
  }
  for (auto e : c->tm.pool) (void)hipEventDestroy(e);
  free_all(c);
  fk_mlp_free(c);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return FK_OK;
}

int fk_set_params(fk_ctx *c, const fk_params *p) {
  if (!c) return fail(FK_EINVAL, "null context");
  int rc = check_params(p);
  if (rc) return rc;
  if (p->model != c->p.model) return fail(FK_EINVAL, "the model cannot change on a context");
  HIPCHK(hipSetDevice(c->device));
  const bool prec_changed = p->precision != c->p.precision;
  const bool heat_changed = (p->heatFlux != c->p.heatFlux) || (p->viscFlag != c->p.viscFlag);
  c->p = resolved(*p);
  c->w_valid = false;
  if (c->have_mesh && (prec_changed || heat_changed)) {
    // the T-typed arrays were sized for double and for the largest NG; only refill geometry
    rc = upload_geometry(c);
    if (rc) return rc;
  }
  c->rhs_valid = false;
  return FK_OK;
}

int fk_upload_mesh(fk_ctx *c, int64_t nElem, int64_t nFcs, const int32_t *fc2el, const int32_t *boundCond,
                   const double *n, const double *area, const double *fcCenter, const double *cellCenter,
                   const double *volume, const double *shift) {
  if (!c || !fc2el || !boundCond || !n || !area || !fcCenter || !cellCenter || !volume)
    return fail(FK_EINVAL, "null argument");
  if (nElem <= 0 || nFcs <= 0) return fail(FK_EINVAL, "empty mesh");
  if (nElem >= INT32_MAX) return fail(FK_ERANGE, "mesh too large for int32 cell ids");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  // faces per cell: 4 (tetrahedron), 5 (pyramid, prism) or 6 (hexahedron)
  std::vector<uint8_t> cnt((size_t)nElem, 0);
  for (int64_t f = 0; f < nFcs; ++f) {
    const int32_t o = fc2el[2 * f], nb = fc2el[2 * f + 1];
    if (o < 0 || o >= nElem || nb >= nElem || nb == o) return fail(FK_EINVAL, "bad fc2el entry");
    const int bc = boundCond[f];
    if ((nb < 0) != (bc != 0)) return fail(FK_EINVAL, "boundCond must be nonzero exactly on boundary faces");
    if (bc < 0 || bc > 5) return fail(FK_EINVAL, "boundCond must be 0..5");
    for (int s = 0; s < 2; ++s) {
      const int32_t cc = fc2el[2 * f + s];
      if (cc < 0) continue;
      if (cnt[(size_t)cc] >= 6) return fail(FK_EINVAL, "cell with more than six faces");
      cnt[(size_t)cc]++;
    }
  }
  int NS = 4;
  for (int64_t cc = 0; cc < nElem; ++cc) {
    if (cnt[(size_t)cc] < 4) return fail(FK_EINVAL, "every cell needs four to six faces");
    if (cnt[(size_t)cc] != 4) NS = 6;
  }
  if (nElem * NS >= (int64_t)1 << 30) return fail(FK_ERANGE, "mesh too large for 30-bit slot ids");
  // cell -> faces in ascending face id, slots past a cell's last face empty
  std::vector<int32_t> sf((size_t)(nElem * NS), -1);
  std::fill(cnt.begin(), cnt.end(), 0);
  for (int64_t f = 0; f < nFcs; ++f)
    for (int s = 0; s < 2; ++s) {
      const int32_t cc = fc2el[2 * f + s];
      if (cc >= 0) sf[(size_t)(cc * NS + cnt[(size_t)cc]++)] = (int32_t)f;
    }
  const int64_t ns = nElem * NS;
  std::vector<int32_t> hlink((size_t)ns, kNoFace);
  c->h_geo.assign((size_t)ns * 4, 0.0);
  c->h_r.assign((size_t)ns * 4, 0.0);
  c->h_ls.assign((size_t)ns, 0.0);
  // face -> slot of the owner side / neighbour side
  std::vector<int32_t> fslot((size_t)(nFcs * 2), -1);
  for (int64_t s = 0; s < ns; ++s) {
    const int32_t f = sf[(size_t)s];
    if (f < 0) continue;
    const int64_t cc = s / NS;
    const int side = fc2el[2 * f] == cc ? 0 : 1;
    fslot[(size_t)(2 * f + side)] = (int32_t)s;
  }
  for (int64_t s = 0; s < ns; ++s) {
    const int32_t f = sf[(size_t)s];
    if (f < 0) {  // no face: zero geometry, the empty-slot link
      c->h_r[(size_t)(s * 4 + 3)] = (double)kNoFace;
      continue;
    }
    const int64_t cc = s / NS;
    const int side = fc2el[2 * f] == cc ? 0 : 1;
    const int32_t other = fc2el[2 * f + (1 - side)];
    if (other >= 0) {
      hlink[(size_t)s] = (fslot[(size_t)(2 * f + 1 - side)] << 1) | (side == 0 ? 1 : 0);
    } else {
      if (boundCond[f] < 1) return fail(FK_EINVAL, "boundCond must be >= 1 on boundary faces");
      hlink[(size_t)s] = -boundCond[f];
    }
    c->h_r[(size_t)(s * 4 + 3)] = (double)hlink[(size_t)s];
    for (int d = 0; d < 3; ++d) c->h_geo[(size_t)(s * 4 + d)] = n[3 * f + d];
    c->h_geo[(size_t)(s * 4 + 3)] = area[f];
    for (int d = 0; d < 3; ++d) {
      double xf = fcCenter[3 * f + d];
      if (side == 1 && shift) xf = xf - shift[3 * f + d];
      c->h_r[(size_t)(s * 4 + d)] = xf - cellCenter[3 * cc + d];
    }
    const int32_t a = fc2el[2 * f];
    const int32_t bb = fc2el[2 * f + 1] >= 0 ? fc2el[2 * f + 1] : a;
    c->h_ls[(size_t)s] = 0.325 * std::pow((volume[a] + volume[bb]) / 2, 0.3333333333333);
  }
  std::vector<double> vcfl((size_t)nElem);
  for (int64_t cc = 0; cc < nElem; ++cc) vcfl[(size_t)cc] = std::pow(volume[cc], 0.3333);
  free_all(c);
  c->have_mesh = c->have_state = false;
  c->nc = nElem;
  c->nf = nFcs;
  c->ns = NS;
  const int neq = c->neq;
  const size_t cs = (size_t)nElem;
  HIPCHK(hipMalloc(&c->link, (size_t)ns * 4));
  HIPCHK(hipMalloc(&c->geo, (size_t)ns * 4 * 8));
  HIPCHK(hipMalloc(&c->rvec, (size_t)ns * 4 * 8));
  HIPCHK(hipMalloc(&c->wls, (size_t)ns * 8));
  HIPCHK(hipMalloc(&c->vol, cs * 8));
  HIPCHK(hipMalloc(&c->volcfl, cs * 8));
  for (int i = 0; i < 3; ++i) HIPCHK(hipMalloc(&c->U[i], cs * neq * 8));
  HIPCHK(hipMalloc(&c->face, (size_t)ns * neq * 8));
  HIPCHK(hipMalloc(&c->gvel, cs * 12 * 8));
  HIPCHK(hipMalloc(&c->rhs, cs * neq * 8));
  HIPCHK(hipMalloc(&c->dtc, cs * 8));
  HIPCHK(hipMalloc(&c->red, 4096 * 8));
  HIPCHK(hipMalloc(&c->dscal, 16 * 8));
  HIPCHK(hipMalloc(&c->lcount, 16 * 8));
  HIPCHK(hipMalloc(&c->health, sizeof(unsigned long long)));
  {
    const unsigned long long none = ~0ull;
    HIPCHK(hipMemcpy(c->health, &none, sizeof none, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMalloc(&c->wprim, cs * neq * sizeof(float)));
  HIPCHK(hipMemset(c->lcount, 0, 16 * 8));
  HIPCHK(hipMemcpy(c->link, hlink.data(), (size_t)ns * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->vol, volume, cs * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->volcfl, vcfl.data(), cs * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->dscal, 0, 16 * 8));
  c->slot_face.swap(sf);
  int rc = upload_geometry(c);
  if (rc) return rc;
  c->have_mesh = true;
  c->rhs_valid = false;
  return FK_OK;
}

int fk_upload_state(fk_ctx *c, const double *U) {
  if (!c || !U) return fail(FK_EINVAL, "null argument");
  if (!c->have_mesh) return fail(FK_ESTATE, "upload the mesh first");
  HIPCHK(hipSetDevice(c->device));
  c->ia = c->iu0 = 0;
  HIPCHK(hipMemcpy(c->U[0], U, (size_t)c->nc * c->neq * 8, hipMemcpyHostToDevice));
  c->have_state = true;
  c->w_valid = false;
  c->rhs_valid = false;
  return FK_OK;
}

int fk_download_state(fk_ctx *c, double *U) {
  int rc = require(c);
  if (rc) return rc;
  if (!U) return fail(FK_EINVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(U, c->U[c->ia], (size_t)c->nc * c->neq * 8, hipMemcpyDeviceToHost));
  return FK_OK;
}

int fk_init_rk(fk_ctx *c) {
  int rc = require(c);
  if (rc) return rc;
  c->iu0 = c->ia;
  return FK_OK;
}

int fk_compute_face_values(fk_ctx *c) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  rc = do_pass_g(c);
  c->rhs_valid = false;
  return rc;
}

int fk_compute_rhs(fk_ctx *c) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  return do_pass_f_rhs(c);
}

int fk_compute_time_step(fk_ctx *c, double *dt) {
  int rc = require(c);
  if (rc) return rc;
  if (!c->rhs_valid) return fail(FK_ESTATE, "fk_compute_rhs must run before fk_compute_time_step");
  HIPCHK(hipSetDevice(c->device));
  rc = do_dt_device(c);
  if (rc) return rc;
  double h[2];
  HIPCHK(hipMemcpyAsync(h, c->dscal, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->dt = h[0];
  if (dt) *dt = h[0];
  return FK_OK;
}

int fk_update(fk_ctx *c, int irk, double dt) {
  int rc = require(c);
  if (rc) return rc;
  if (!c->rhs_valid) return fail(FK_ESTATE, "fk_compute_rhs must run before fk_update");
  if (irk < 0 || irk >= c->p.rkSteps) return fail(FK_EINVAL, "bad RK stage");
  HIPCHK(hipSetDevice(c->device));
  rc = do_update(c, irk, dt, nullptr, true);
  c->rhs_valid = false;
  return rc;
}

int fk_compute_rhs_update(fk_ctx *c, int irk, double dt) {
  int rc = require(c);
  if (rc) return rc;
  if (irk < 0 || irk >= c->p.rkSteps) return fail(FK_EINVAL, "bad RK stage");
  HIPCHK(hipSetDevice(c->device));
  return do_pass_f_update(c, irk, dt, nullptr, true);
}

int fk_relaxation(fk_ctx *c, int mode, fk_relax_stats *st) {
  int rc = require(c);
  if (rc) return rc;
  if (mode != FK_RELAX_PLAIN && mode != FK_RELAX_MLP && mode != FK_RELAX_COLD)
    return fail(FK_EINVAL, "unknown relaxation mode");
  HIPCHK(hipSetDevice(c->device));
  return do_relax(c, st, mode);
}

int fk_set_tuning(fk_ctx *c, const char *name, int64_t value) {
  if (!c || !name) return fail(FK_EINVAL, "null argument");
  if (std::strcmp(name, "fuse_update") == 0) {
    c->fuse_update = value != 0;
    return FK_OK;
  }
  if (std::strcmp(name, "stage_f") == 0) {
    c->stage_f = value != 0;
    return FK_OK;
  }
  if (std::strcmp(name, "prims_w") == 0) {
    c->prims_w = value != 0;
    c->w_valid = false;
    return FK_OK;
  }
  if (std::strcmp(name, "fuse_relax") == 0) {
    c->fuse_relax = value != 0;
    return FK_OK;
  }
  if (std::strcmp(name, "stage_g") == 0) {
    c->stage_g = value != 0;
    return FK_OK;
  }
  return fail(FK_EINVAL, std::string("unknown tuning option ") + name);
}

int fk_relax_totals(fk_ctx *c, fk_relax_stats *st, int reset) {
  int rc = require(c);
  if (rc) return rc;
  if (!st && !reset) return fail(FK_EINVAL, "null stats");
  HIPCHK(hipSetDevice(c->device));
  if (st) {
    if ((rc = read_counts(c, 8, st))) return rc;
    st->n_therm = c->relax_calls * c->nc - st->n_chem;
  }
  if (reset) {
    hipLaunchKernelGGL(zero_ull, dim3(1), dim3(64), 0, c->stream, (unsigned long long *)c->lcount + 8, 5);
    HIPCHK(hipGetLastError());
    c->relax_calls = 0;
  }
  return FK_OK;
}

int fk_cons_calc(fk_ctx *c, double *out) {
  int rc = require(c);
  if (rc) return rc;
  if (!out) return fail(FK_EINVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const int nb = kConsBlocks;
  if (c->neq == 5)
    hipLaunchKernelGGL((integrate<5>), dim3(nb), dim3(kBlock), 0, c->stream, c->nc, c->U[c->ia], c->vol, c->red);
  else
    hipLaunchKernelGGL((integrate<10>), dim3(nb), dim3(kBlock), 0, c->stream, c->nc, c->U[c->ia], c->vol, c->red);
  HIPCHK(hipGetLastError());
  std::vector<double> h((size_t)nb * c->neq);
  HIPCHK(hipMemcpyAsync(h.data(), c->red, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int v = 0; v < c->neq; ++v) {
    double s = 0;
    for (int b = 0; b < nb; ++b) s += h[(size_t)(b * c->neq + v)];
    out[v] = s;
  }
  return FK_OK;
}

int fk_cons_calc_noalloc(fk_ctx *c, double *out) {
  int rc = require(c);
  if (rc) return rc;
  if (!out) return fail(FK_EINVAL, "null argument");
  if ((int64_t)kConsBlocks * c->neq + c->neq > kScratchDoubles)
    return fail(FK_ERANGE, "fk_cons_calc_noalloc: " + std::to_string(c->neq) + " variables exceed the scratch of " +
                               std::to_string(kScratchDoubles) + " doubles");
  HIPCHK(hipSetDevice(c->device));
  double *ptr = scratch_of(c->device);
  if (!ptr) return fail(FK_ESTATE, "scratch not set up on this device");
  // partials [kConsBlocks][NEQ] at ptr, folded sums [NEQ] after them; every
  // element read is written first in this call, so nothing needs clearing
  double *sums = ptr + (int64_t)kConsBlocks * c->neq;
  if (c->neq == 5) {
    hipLaunchKernelGGL((integrate<5>), dim3(kConsBlocks), dim3(kBlock), 0, c->stream, c->nc, c->U[c->ia], c->vol, ptr);
    hipLaunchKernelGGL((fold_partials<5>), dim3(1), dim3(64), 0, c->stream, kConsBlocks, ptr, sums);
  } else {
    hipLaunchKernelGGL((integrate<10>), dim3(kConsBlocks), dim3(kBlock), 0, c->stream, c->nc, c->U[c->ia], c->vol, ptr);
    hipLaunchKernelGGL((fold_partials<10>), dim3(1), dim3(64), 0, c->stream, kConsBlocks, ptr, sums);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, sums, (size_t)c->neq * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FK_OK;
}

int fk_run_steps(fk_ctx *c, int nsteps, double *dt_last, double *phys_time) {
  int rc = require(c);
  if (rc) return rc;
  if (nsteps < 0) return fail(FK_EINVAL, "negative step count");
  HIPCHK(hipSetDevice(c->device));
  const fk_params &p = c->p;
  // every check and allocation of the relaxation happens here, before any
  // stage runs, so a failure leaves the state and the clock untouched
  if (p.model == FK_MODEL_TWO_PHASE && p.relaxFlag == 1 && nsteps > 0 && (rc = relax_preflight(c, p.relaxMode)))
    return rc;
  // the device scalars hold dt and the physical time; the health record
  // starts clean for this call
  double h[2] = {c->dt, c->time};
  HIPCHK(hipMemcpy(c->dscal, h, sizeof h, hipMemcpyHostToDevice));
  unsigned long long hh = ~0ull;
  HIPCHK(hipMemcpy(c->health, &hh, sizeof hh, hipMemcpyHostToDevice));
  for (int s = 0; s < nsteps; ++s) {
    if (p.rkSteps > 1) c->iu0 = c->ia;
    const bool newdt = p.adaptDt == 1 && c->step % p.adaptDtItr == 0;
    // the last stage's state goes to the relaxation, which writes the
    // primitive records itself
    const bool relax_next = p.model == FK_MODEL_TWO_PHASE && p.relaxFlag == 1;
    const bool warm = p.relaxMode == FK_RELAX_MLP;
    // relaxation inside the last fused pass F: 1 surrogate stage 1, 2 plain
    // (not while the plain relaxation emits training pairs)
    const bool emitting = c->mlp.emit_stride > 0 && c->mlp.cap_data > c->mlp.n_data;
    const int relax_kind = !relax_next ? 0 : (p.relaxMode == FK_RELAX_PLAIN ? (emitting ? 0 : 2) : 1);
    int relaxed_in_f = 0;
    for (int irk = 0; irk < p.rkSteps; ++irk) {
      const bool want_w = irk < p.rkSteps - 1 || !relax_next;
      if ((rc = do_pass_g(c))) return rc;
      if (irk == 0 && newdt) {
        if ((rc = do_pass_f_rhs(c))) return rc;
        if ((rc = do_dt_device(c))) return rc;
        if ((rc = do_update(c, irk, 0.0, c->dscal, want_w))) return rc;
      } else {
        const int fr = (c->fuse_relax && c->fuse_update && irk == p.rkSteps - 1) ? relax_kind : 0;
        if (fr == 1 && (rc = fk_relax_mlp_begin(c, warm))) return rc;
        if (fr == 2 && (rc = relax_plain_begin(c))) return rc;
        if (p.adaptDt == 1) rc = do_pass_f_update(c, irk, 0.0, c->dscal, want_w, fr);
        else rc = do_pass_f_update(c, irk, p.dt, nullptr, want_w, fr);
        if (rc) return rc;
        relaxed_in_f = fr;
      }
    }
    if (p.adaptDt == 1)
      LAUNCH_TIMED(c, "advance_time", hipLaunchKernelGGL(advance_time, dim3(1), dim3(64), 0, c->stream, c->dscal));
    else
      c->dt = p.dt;
    if (relaxed_in_f == 1) {
      if ((rc = relax_fused_finish(c, warm))) return rc;
    } else if (relaxed_in_f == 2) {
      if ((rc = relax_plain_fused_finish(c))) return rc;
    } else if (relax_next) {
      if ((rc = do_relax(c, nullptr, p.relaxMode))) return rc;
    }
    c->step++;
  }
  HIPCHK(hipMemcpyAsync(h, c->dscal, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&hh, c->health, sizeof hh, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (p.adaptDt == 1) {
    c->dt = h[0];
    c->time = h[1];
  } else {
    c->time += p.dt * nsteps;
  }
  if (dt_last) *dt_last = c->dt;
  if (phys_time) *phys_time = c->time;
  if (hh != ~0ull)
    return fail(FK_ERANGE, "non-physical state: non-positive density or non-finite value from step " +
                               std::to_string(hh >> 32) + " on (cell " + std::to_string(hh & 0xffffffffull) + ")");
  return FK_OK;
}

int fk_health(fk_ctx *c, int64_t *step, int64_t *cell, int reset) {
  int rc = require(c);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  unsigned long long hh = 0;
  HIPCHK(hipMemcpyAsync(&hh, c->health, sizeof hh, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (step) *step = hh == ~0ull ? -1 : (int64_t)(hh >> 32);
  if (cell) *cell = hh == ~0ull ? -1 : (int64_t)(hh & 0xffffffffull);
  if (reset) {
    const unsigned long long none = ~0ull;
    HIPCHK(hipMemcpy(c->health, &none, sizeof none, hipMemcpyHostToDevice));
  }
  return FK_OK;
}

int fk_get_clock(fk_ctx *c, double *dt, double *time, int64_t *step) {
  if (!c) return fail(FK_EINVAL, "null context");
  if (dt) *dt = c->dt;
  if (time) *time = c->time;
  if (step) *step = c->step;
  return FK_OK;
}

int fk_set_clock(fk_ctx *c, double dt, double time, int64_t step) {
  if (!c) return fail(FK_EINVAL, "null context");
  if (!(dt >= 0) || step < 0) return fail(FK_EINVAL, "dt must be >= 0 and step >= 0");
  c->dt = dt;
  c->time = time;
  c->step = step;
  return FK_OK;
}

int fk_sync(fk_ctx *c) {
  if (!c) return fail(FK_EINVAL, "null context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FK_OK;
}

int fk_set_debug(fk_ctx *c, int on) {
  if (!c) return fail(FK_EINVAL, "null context");
  if (on && !c->have_mesh) return fail(FK_ESTATE, "upload the mesh before enabling debug output");
  if (on && !c->dgrad) {  // gradient / limiter copies, allocated only when asked for
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMalloc(&c->dgrad, (size_t)c->nc * c->neq * 3 * 8));
    HIPCHK(hipMalloc(&c->dtheta, (size_t)c->nc * c->neq * 8));
  }
  c->debug = on != 0;
  return FK_OK;
}

int fk_download(fk_ctx *c, const char *name, void *host, int64_t bytes) {
  int rc = require(c);
  if (rc) return rc;
  if (!name || !host) return fail(FK_EINVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::string s(name);
  const int64_t nc = c->nc, neq = c->neq;
  const size_t ts = tsz(c);
  auto get = [&](const void *src, int64_t nb) -> int {
    if (nb != bytes) return fail(FK_EINVAL, "size mismatch for " + s + ": expected " + std::to_string(nb));
    HIPCHK(hipMemcpy(host, src, (size_t)nb, hipMemcpyDeviceToHost));
    return FK_OK;
  };
  if ((s == "grad" || s == "theta") && !c->dgrad) return fail(FK_ESTATE, "enable fk_set_debug before pass G");
  if (s == "grad") return get(c->dgrad, nc * neq * 3 * 8);
  if (s == "theta") return get(c->dtheta, nc * neq * 8);
  if (s == "face") return get(c->face, nc * c->ns * neq * (int64_t)ts);
  if (s == "rhs") {
    if (!c->rhs_f32) return get(c->rhs, nc * neq * 8);
    // stored as FP32 by the mixed dt stage: widen on the host
    if (bytes != nc * neq * 8) return fail(FK_EINVAL, "size mismatch for rhs");
    std::vector<float> f((size_t)(nc * neq));
    HIPCHK(hipMemcpy(f.data(), c->rhs, f.size() * 4, hipMemcpyDeviceToHost));
    double *o = (double *)host;
    for (size_t i = 0; i < f.size(); ++i) o[i] = f[i];
    return FK_OK;
  }
  if (s == "signal") {
    // dtc holds Vcfl / signal; return the signal itself
    std::vector<double> d((size_t)nc), v((size_t)nc);
    if (bytes != nc * 8) return fail(FK_EINVAL, "size mismatch for signal");
    HIPCHK(hipMemcpy(d.data(), c->dtc, (size_t)nc * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(v.data(), c->volcfl, (size_t)nc * 8, hipMemcpyDeviceToHost));
    double *o = (double *)host;
    for (int64_t i = 0; i < nc; ++i) o[i] = v[(size_t)i] / d[(size_t)i];
    return FK_OK;
  }
  if (s == "slots") {
    if (bytes != nc * c->ns * 4) return fail(FK_EINVAL, "size mismatch for slots");
    std::memcpy(host, c->slot_face.data(), (size_t)bytes);
    return FK_OK;
  }
  return fail(FK_EINVAL, "unknown array " + s);
}

int fk_stream_copy(fk_ctx *c, int64_t bytes, double *ms) {
  if (!c || !ms) return fail(FK_EINVAL, "null argument");
  if (bytes < 16) return fail(FK_EINVAL, "copy needs at least 16 bytes");
  HIPCHK(hipSetDevice(c->device));
  const int64_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, (size_t)n16 * 16));
  hipError_t e = hipMalloc(&b, (size_t)n16 * 16);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return fail(FK_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipMemsetAsync(a, 1, (size_t)n16 * 16, c->stream));
  const unsigned grid = (unsigned)std::min<int64_t>((n16 + kBlock - 1) / kBlock, 65536);
  hipLaunchKernelGGL(stream_copy, dim3(grid), dim3(kBlock), 0, c->stream, n16, (const uint4 *)a, (uint4 *)b);  // warm
  HIPCHK(hipEventRecord(e0, c->stream));
  hipLaunchKernelGGL(stream_copy, dim3(grid), dim3(kBlock), 0, c->stream, n16, (const uint4 *)a, (uint4 *)b);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return FK_OK;
}

int fk_set_timing(fk_ctx *c, int on) {
  if (!c) return fail(FK_EINVAL, "null context");
  c->timing = on != 0;
  return FK_OK;
}

int fk_kernel_times(fk_ctx *c, int max, char (*names)[32], int64_t *counts, double *ms) {
  if (!c) return fail(FK_EINVAL, "null context");
  HIPCHK(hipSetDevice(c->device));
  int rc = timer_flush(c);
  if (rc) return rc;
  const int n = (int)c->tm.names.size();
  for (int i = 0; i < n && i < max; ++i) {
    if (names) {
      std::strncpy(names[i], c->tm.names[(size_t)i].c_str(), 31);
      names[i][31] = 0;
    }
    if (counts) counts[i] = c->tm.counts[(size_t)i];
    if (ms) ms[i] = c->tm.ms[(size_t)i];
  }
  return n;
}

}  // extern "C"
