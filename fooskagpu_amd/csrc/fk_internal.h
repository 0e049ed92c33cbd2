// Private definitions shared by the translation units of libfooska_hip.so:
// the context struct, error reporting and the per-kernel event timer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fooska.h"

namespace fkx {

inline thread_local std::string g_err;
inline int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

constexpr int kBlock = 256;

template <typename T>
struct V4 {
  T x, y, z, w;
};

struct Timer {
  std::vector<std::string> names;
  std::vector<int64_t> counts;
  std::vector<double> ms;
  struct Pending {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
};

// Surrogate (MLP 3-64-64-64-3) state of a context.
struct Mlp {
  float *w = nullptr;      // FP32 master parameters [kMlpParams]
  float *m = nullptr, *v = nullptr;  // Adam moments
  uint16_t *wb = nullptr;  // BF16 packed copies of W2, W3 and their transposes
  float *grad = nullptr;   // reduced gradient
  float *part = nullptr;   // per-wave gradient partials
  int64_t part_cap = 0;    // waves the partial buffer holds
  float norm[12] = {0, 0, 0, 1, 1, 1, 0, 0, 0, 1, 1, 1};  // x mean, x std, y mean, y std
  bool ready = false;      // weights set or trained
  int64_t adam_t = 0;
  // training set on the device
  float *dx = nullptr, *dy = nullptr;
  int64_t n_data = 0, cap_data = 0;
  // emission of solver states during plain relaxations
  int64_t emit_stride = 0, emit_phase = 0;
};

}  // namespace fkx

struct fk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  fk_params p{};
  int neq = 5;
  int64_t nc = 0, nf = 0;
  int ns = 4;  // face slots per cell: 4 (every cell a tetrahedron) or 6
  bool have_mesh = false, have_state = false, debug = false, timing = false;
  // mesh
  int32_t *link = nullptr;  // [4C] slot links (see fk_kernels.hip)
  void *geo = nullptr, *rvec = nullptr, *wls = nullptr;  // T-typed, sized for double
  double *vol = nullptr, *volcfl = nullptr;
  std::vector<int32_t> slot_face;  // host copy of slot -> face id
  // fields
  double *U[3] = {nullptr, nullptr, nullptr};
  int ia = 0, iu0 = 0;
  void *face = nullptr, *gvel = nullptr;
  double *rhs = nullptr, *dtc = nullptr, *red = nullptr, *dscal = nullptr;
  // state health found by pass G: (step << 32) | cell, the smallest of the
  // cells with a non-positive density or a non-finite primitive; ~0 if none
  unsigned long long *health = nullptr;
  double *dgrad = nullptr, *dtheta = nullptr;
  // relaxation counters: [0, 5) of the last relaxation (n_chem, Newton
  // iterations, Tsat iterations, fallbacks, warm-started cells; slot 5 is the
  // surrogate work-list length), [8, 13) summed over the relaxations of
  // fk_run_steps
  long long *lcount = nullptr;
  int64_t relax_calls = 0;  // relaxations summed in lcount[8, 13)
  // relaxation work list (chemical cells / surrogate list), one entry per cell
  uint32_t *rlist = nullptr;
  uint32_t *rfall = nullptr;   // surrogate relaxation: cells whose polish failed
  void *rstart = nullptr;      // surrogate relaxation: network starts (float2) of the chemical entries
  int64_t rstart_cap = 0;
  unsigned long long *rpart = nullptr;  // per-workgroup relaxation counts [blocks][4]
  int64_t rlist_cap = 0;
  double dt = 0, time = 0;
  int64_t step = 0;
  bool rhs_valid = false;
  bool rhs_f32 = false;      // c->rhs holds FP32 values (mixed dt stage, pass F MODE 1)
  bool fuse_update = true;   // RK update inside pass F (else residual + rk_update)
  bool stage_f = true;       // pass F stages the workgroup's face states in LDS (tetrahedral meshes)
  bool prims_w = true;       // mixed: producers of U also write wprim, pass G reads it
  bool fuse_relax = true;    // per-cell relaxation pass inside the last fused pass F
  bool stage_g = true;       // mixed pass G writes its face records through LDS (4 slots per cell)
  int64_t pf_waves = 0;      // waves of the last pass F launch (its per-wave relaxation counts)
  // FP32 primitive records [C][NEQ] of U[ia] (mixed precision), valid while
  // w_valid: set by the kernels that write them with a new state, cleared by
  // every other change of U
  float *wprim = nullptr;
  bool w_valid = false;
  fkx::Timer tm;
  fkx::Mlp mlp;
  // host geometry kept to rebuild the precision-dependent arrays
  std::vector<double> h_geo, h_r, h_ls;
};

// defined in fk_mlp.hip
int fk_relax_mlp_impl(fk_ctx *c, bool warm);
int fk_relax_mlp_begin(fk_ctx *c, bool warm);   // checks, scratch, zeroed counts
int fk_relax_mlp_finish(fk_ctx *c, bool warm);  // stage 2 (work list) after stage 1
int ensure_rlist(fk_ctx *c);  // defined in fk_kernels.hip
int ensure_rstart(fk_ctx *c);  // defined in fk_kernels.hip
int fk_relax_emit_impl(fk_ctx *c, bool *done);
void fk_mlp_free(fk_ctx *c);

namespace fkx {

#define HIPCHK(x)                                                                                    \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return ::fkx::fail(FK_EDEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

inline int timer_id(fk_ctx *c, const char *name) {
  for (size_t i = 0; i < c->tm.names.size(); ++i)
    if (c->tm.names[i] == name) return (int)i;
  c->tm.names.emplace_back(name);
  c->tm.counts.push_back(0);
  c->tm.ms.push_back(0);
  return (int)c->tm.names.size() - 1;
}

inline int timer_begin(fk_ctx *c, const char *name, int *slot) {
  *slot = -1;
  if (!c->timing) return FK_OK;
  Timer::Pending pd;
  pd.id = timer_id(c, name);
  if (c->tm.pool.size() >= 2) {
    pd.b = c->tm.pool.back();
    c->tm.pool.pop_back();
    pd.a = c->tm.pool.back();
    c->tm.pool.pop_back();
  } else {
    HIPCHK(hipEventCreate(&pd.a));
    HIPCHK(hipEventCreate(&pd.b));
  }
  HIPCHK(hipEventRecord(pd.a, c->stream));
  c->tm.pending.push_back(pd);
  *slot = (int)c->tm.pending.size() - 1;
  return FK_OK;
}
inline int timer_end(fk_ctx *c, int slot) {
  if (slot < 0) return FK_OK;
  HIPCHK(hipEventRecord(c->tm.pending[(size_t)slot].b, c->stream));
  return FK_OK;
}
inline int timer_flush(fk_ctx *c) {
  for (auto &pd : c->tm.pending) {
    HIPCHK(hipEventSynchronize(pd.b));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, pd.a, pd.b));
    c->tm.counts[(size_t)pd.id]++;
    c->tm.ms[(size_t)pd.id] += ms;
    c->tm.pool.push_back(pd.a);
    c->tm.pool.push_back(pd.b);
  }
  c->tm.pending.clear();
  return FK_OK;
}

#define LAUNCH_TIMED(ctx, name, ...)                  \
  do {                                                \
    int slot_;                                        \
    int rc_ = timer_begin(ctx, name, &slot_);         \
    if (rc_) return rc_;                              \
    __VA_ARGS__;                                      \
    HIPCHK(hipGetLastError());                        \
    rc_ = timer_end(ctx, slot_);                      \
    if (rc_) return rc_;                              \
  } while (0)


}  // namespace fkx
