// clv_internal.h -- host-side internals shared by the engine (clv_engine.hip)
// and the serving loop (clv_server.hip): shape buckets of chains, the context
// record, and the launch / per-chain helpers.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/clv.h"
#include "clv_device.h"

namespace clvi {

using clv::ChainParams;
using clv::ChainState;

// A bucket: `nch` chains of one shape in one set of device buffers.  Kernel
// launches cover a bucket (or a range of its chains).
struct Bucket {
  clv_shape shape{};
  int X = 0, R = 0, Nt = 0;
  int64_t S = 0, Vs = 0;
  int nch = 0;
  std::vector<int> ids;  // global chain ids (context buckets)
  int rows = 1, ntiles = 1, nthr = 64, tc = 1, nchunks = 1, tcs = 1;
  int n1 = 1;            // row transform split X = n1 * (X / n1)
  size_t esz = 0;        // bytes per complex element
  void *psi[2] = {nullptr, nullptr}, *psib[2] = {nullptr, nullptr};
  ChainParams *d_prm = nullptr;
  ChainState *d_st = nullptr;
  void *d_tw = nullptr;
  double *d_p2 = nullptr;  // -p_x(k)^2 / X
  double *d_part = nullptr, *d_stage = nullptr, *d_spart = nullptr, *d_F = nullptr,
         *d_acc = nullptr;
  int64_t *d_count = nullptr;
  std::vector<int64_t> host_steps;  // steps taken per chain (host mirror)
  std::vector<int> host_eval;       // per-chain evaluation cadence (serving)
  std::vector<int64_t> host_eval_from;
  std::vector<char> host_active;    // per-chain active flag (host mirror)
  int schedule = 1;      // 1 two-pass, 2 fused
  int epoch = 0;         // slot (epoch & 1) of d_st / d_part is current
  bool primed = false;   // current partials describe the current fields
  size_t part_slot = 0;  // doubles per partials slot
  size_t fused_lds = 0;  // LDS bytes of the fused kernel as configured (0: not available)
  int fused_sets = 1;    // LDS line sets of the fused kernel (2 when they fit)
  int fused_sets_max = 1;
  int fused_batch = 1;   // tau slices per batch of next-step K_0 lines (1, 2, 4)
  int fused_blocks = 1;  // tau blocks of the fused march (a divisor of Nt, blocks of >= 2 slices)
  // kin_dims = 3: transverse kinetic term kt (allocated when a chain asks for it)
  struct KtAxis {
    int L = 1, bmul = 0, rstride = 0, nbase = 0, W = 1;
    void *tw = nullptr;
    double *p2 = nullptr;
  };
  std::vector<KtAxis> kt_axes;  // y and/or z, the axes longer than 1
  void *kt[2] = {nullptr, nullptr};
  bool kt_on = false;
  std::vector<void *> allocs;
  int64_t bytes = 0;
};

// The fused schedule needs no transverse term (it would need it of the new fields).
inline bool uses_fused(const Bucket &b) { return b.schedule == 2 && b.fused_lds > 0 && !b.kt_on; }

// Schedule without timing: the fused march has one workgroup per (tile, chain)
// and walks tau serially, so it only pays off with many of them; below that the
// two-pass kernels (one workgroup per tile and tau) fill the GPU better.  The
// threshold is measured on MI355X: at 512 workgroups fused FP32 is 0.80-0.99 of
// the two-pass time across 2-D and 3-D shapes and Nt = 16..128, at 320-384 it
// is 1.05-1.33.  FP64 (two waves/SIMD, no spills): 0.80-0.81 of the two-pass
// step at 512-1024 workgroups (32 and 64 chains of 64x64x1x64, 16 chains of
// 128x128x1x32), equal at 256 (16 chains of 64x64x1x64).  Counted in waves
// (those were 4-wave workgroups), so two-wave workgroups need twice as many.
// Both schedules give identical results, so the choice may change at any step.
constexpr int64_t kFusedMinWaves = 2048;
// Tau blocks of the fused march for nactive chains: the fewest (a power of two
// dividing Nt, blocks of at least kFusedMinBlock slices) that reach
// kFusedMinWaves waves, 0 when none does.
constexpr int kFusedMinBlock = 8;
inline int fused_blocks_for(const Bucket &b, int nactive) {
  const int64_t waves = int64_t(b.ntiles) * nactive * ((b.nthr + 63) / 64);
  for (int nb = 1; nb <= b.Nt / kFusedMinBlock || nb == 1; nb *= 2) {
    if (b.Nt % nb) break;
    if (waves * nb >= kFusedMinWaves) return nb;
  }
  return 0;
}
inline int auto_schedule(const Bucket &b, int nactive) {
  return (b.fused_lds > 0 && !b.kt_on && fused_blocks_for(b, nactive) > 0) ? 2 : 1;
}
// Untimed choice: the schedule and, for the fused one, its tau blocks.
inline void auto_plan(Bucket &b, int nactive) {
  b.schedule = auto_schedule(b, nactive);
  b.fused_blocks = (b.schedule == 2) ? fused_blocks_for(b, nactive) : 1;
}

}  // namespace clvi

struct clv_ctx {
  int device = 0;
  int precision = CLV_FP64;
  hipStream_t stream = nullptr;
  std::vector<clvi::Bucket> buckets;
  std::vector<std::pair<int, int>> where;  // chain -> (bucket, local)
  std::vector<clv_params> params;
  int64_t bytes = 0;
  std::string err;
  // kernel timing with HIP events on `stream`
  bool profiling = false;
  struct Pending { int kind; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  double kt_ms[7] = {0, 0, 0, 0, 0, 0, 0};
  int64_t kt_n[7] = {0, 0, 0, 0, 0, 0, 0};
  // captured step graphs of clv_propagate, keyed by the eval cadence and the
  // buckets' schedules, state-slot parities and cadence phases at capture time
  bool use_graphs = true;
  struct GraphEntry { std::string key; hipGraphExec_t exec; };
  std::vector<GraphEntry> graphs;
};

// Reported once: the runtime's pending error is cleared, so a later launch
// check does not report it again.
#define HIPCHK(ctx, expr)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      (void)hipGetLastError();                                                             \
      return clvi::fail(ctx, CLV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    }                                                                                      \
  } while (0)

namespace clvi {

int fail(clv_ctx *ctx, int code, const std::string &msg);
bool supported_x(int X);
// per-chain device bytes of a bucket of this shape and precision
int64_t chain_bytes(const clv_shape &s, int precision);
int setup_bucket(clv_ctx *ctx, Bucket &b);
// Row-transform split of a bucket (0: the default for its row length).
int set_row_split(clv_ctx *ctx, Bucket &b, int n1);
void free_bucket(clv_ctx *ctx, Bucket &b);
ChainParams to_device_params(const clv_params &p, int eval_every, int active,
                             int64_t eval_from = 0);
// One Langevin step of every active chain of the bucket; eval_every < 0
// evaluates each chain at its own cadence (ChainParams::eval_every).
void step_bucket(clv_ctx *ctx, Bucket &b, int eval_every);
// Launches of one step of bucket b for graph capture; the observable kernels
// (gated per chain on the device) follow when `obs`; no host bookkeeping.
void enqueue_step_gated(clv_ctx *ctx, Bucket &b, int eval_every, bool obs = true);
void drop_graphs(clv_ctx *ctx);
int read_state(clv_ctx *ctx, const Bucket &b, int local, ChainState *out);
// Serving helpers, all ordered on the context's stream.
int chain_admit(clv_ctx *ctx, Bucket &b, int local, const clv_params &p, int eval_every,
                int64_t eval_from, double n0, double var);
// Accumulate the observables of one chain's current fields now.
void chain_evaluate(clv_ctx *ctx, Bucket &b, int local);
int chain_set_active(clv_ctx *ctx, Bucket &b, int local, const clv_params &p, int eval_every,
                     bool active);
// Free the slots `locals` (their kernels exit at once): one batch of copies
// and one synchronisation.
int chains_deactivate(clv_ctx *ctx, Bucket &b, const std::vector<int> &locals);
int chain_read_observables(clv_ctx *ctx, Bucket &b, int local, clv_observables *out, double *Fx,
                           double *Fy, double *Fz);
void stats_from_state(const ChainState &st, clv_step_stats *out);

}  // namespace clvi
