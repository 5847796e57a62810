// lufqhe_serve.hip -- a determinant server for Monte Carlo walkers.
//
// The reference's Monte Carlo loop (LU-fqhe/src/MonteCarloLoop.cu:280-400)
// rebuilds every walker's matrices on the host each iteration, copies them
// to the device, runs one CALL per GPU and waits.  Here the walkers live on
// the device and requests stream in:
//
//  * walker w owns n_w particle positions (walker_z, 64 slots per walker)
//    and a table of P row permutations (its symmetrisation terms);
//  * a request is one Monte Carlo step of one walker: move particle p by a
//    complex Gaussian delta, then evaluate k permutation terms
//    det(P_t A(z')) of the moved configuration z' and their sum psi';
//  * requests arrive as a Poisson process; the host loop admits them while
//    an HBM budget and batch slots allow, packs the admitted ones into a
//    batch and launches it on one of several streams while earlier batches
//    are still running (continuous batching);
//  * each batch is: one prepare kernel (moves particles into per-request
//    configuration slots, commits accepted configurations, expands per-term
//    index tables), one determinant launch per n bucket with the FQHE matrix
//    built in the kernel's load phase, one per-request sum kernel, one small
//    device-to-host copy of the sums;
//  * on completion the host applies the Metropolis rule
//    |psi'|^2 >= u |psi|^2 (u from the request's own random stream).
//
// Every random number is a function of (seed, request index), a walker has
// at most one request in flight, and every kernel computes each request
// from its own inputs only, so every result is bit-identical whatever the
// arrival rate, batch composition or overlap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <new>
#include <thread>
#include <vector>

#include "../../include/lufqhe.h"
#include "lup_dispatch.hpp"
#include "scaled_sum.hpp"

namespace lufqhe {
int32_t hip_fail(hipError_t e);
const KernelEntry& entry_for(int variant, int n);
int64_t plain_blocks(const KernelEntry& e, int64_t count);
}  // namespace lufqhe

using namespace lufqhe;

namespace {

constexpr int kSlotZ = 64;  // complex positions per configuration slot

// Device descriptor of one request in a batch.
struct ReqDesc {
  int32_t walker;
  int32_t n;
  int32_t particle;
  int32_t commit_slot;  // -1, or the slot holding the walker's accepted configuration
  int32_t slot;         // this request's configuration slot
  int32_t k;
  int32_t perm_start;
  int32_t pad;
  int64_t table_pos;  // first index of its terms in its n bucket's tables
  int64_t out_base;   // first det index in the batch output
  double2 delta;
};

// ---------------------------------------------------------------- kernels
// One workgroup per request: the move (and the commit of the walker's last
// accepted configuration), then the per-term index tables.
__global__ void __launch_bounds__(256) serve_prepare_kernel(
    const ReqDesc* __restrict__ reqs, double2* __restrict__ walker_z,
    double2* __restrict__ slot_z, int32_t perms_per_walker, int64_t* __restrict__ offs,
    int64_t* __restrict__ perm_offs, int64_t* __restrict__ out_idx) {
  const ReqDesc r = reqs[blockIdx.x];
  const int t0 = threadIdx.x;
  if (t0 < r.n) {
    const double2 src = r.commit_slot >= 0 ? slot_z[(int64_t)r.commit_slot * kSlotZ + t0]
                                           : walker_z[(int64_t)r.walker * kSlotZ + t0];
    if (r.commit_slot >= 0) walker_z[(int64_t)r.walker * kSlotZ + t0] = src;
    const double2 z = t0 == r.particle ? make_double2(src.x + r.delta.x, src.y + r.delta.y) : src;
    slot_z[(int64_t)r.slot * kSlotZ + t0] = z;
  }
  for (int t = t0; t < r.k; t += blockDim.x) {
    const int64_t i = r.table_pos + t;
    const int64_t p = ((int64_t)r.walker * perms_per_walker + (r.perm_start + t) % perms_per_walker);
    offs[i] = (int64_t)r.slot * kSlotZ;
    perm_offs[i] = p * kSlotZ;
    out_idx[i] = r.out_base + t;
  }
}

// The per-request sums use the (mant, exp) arithmetic of scaled_sum.hpp.

// One 64-thread workgroup per request: lane l adds terms l, l+64, ... in
// order, then a fixed tree over the lanes.  psi' = out_m * 2^out_e.
__global__ void __launch_bounds__(64) serve_sum_kernel(const ReqDesc* __restrict__ reqs,
                                                       const double2* __restrict__ mant,
                                                       const int32_t* __restrict__ expo,
                                                       double2* __restrict__ out_m,
                                                       int32_t* __restrict__ out_e) {
  __shared__ double2 sm[64];
  __shared__ int se[64];
  const ReqDesc r = reqs[blockIdx.x];
  const int l = threadIdx.x;
  Scaled acc{make_double2(0.0, 0.0), 0};
  for (int t = l; t < r.k; t += 64) {
    const int64_t i = r.out_base + t;
    acc = scaled_add(acc, scaled_norm(mant[i], expo[i]));
  }
  sm[l] = acc.m;
  se[l] = acc.e;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (l < w) {
      const Scaled s = scaled_add(Scaled{sm[l], se[l]}, Scaled{sm[l + w], se[l + w]});
      sm[l] = s.m;
      se[l] = s.e;
    }
    __syncthreads();
  }
  if (l == 0) {
    out_m[blockIdx.x] = sm[0];
    out_e[blockIdx.x] = se[0];
  }
}

// ---------------------------------------------------------------- random numbers
// SplitMix64 over (seed, request, stream): a counter-based generator, so a
// request's numbers do not depend on the order requests are processed in.
inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline uint64_t rnd(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed) ^ a) ^ (b * 0xD1B54A32D192ED03ull));
}
inline double unif(uint64_t seed, uint64_t a, uint64_t b) {  // (0, 1)
  return ((double)(rnd(seed, a, b) >> 11) + 0.5) * 0x1.0p-53;
}
inline double gauss(uint64_t seed, uint64_t a, uint64_t b) {  // Box-Muller
  const double u1 = unif(seed, a, 2 * b), u2 = unif(seed, a, 2 * b + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// the ABI structs' layouts, as bindings in other languages declare them
static_assert(sizeof(lufqhe_serve_config) == 56, "lufqhe_serve_config layout");
static_assert(sizeof(lufqhe_serve_stats) == 120, "lufqhe_serve_stats layout");
static_assert(sizeof(lufqhe_serve_record) == 64, "lufqhe_serve_record layout");
static_assert(sizeof(ReqDesc) == 64, "ReqDesc layout");

}  // namespace

// ---------------------------------------------------------------- server
struct lufqhe_server {
  int device = 0;
  int variant = 0;
  int walkers = 0;
  int perms = 0;
  uint64_t seed = 0;
  std::vector<int32_t> wn;            // n per walker
  std::vector<double2> z0;            // initial configurations [walkers][64]
  double2* d_walker_z = nullptr;      // [walkers][64]
  uint8_t* d_perm = nullptr;          // [walkers][perms][64]
  int64_t resident_bytes = 0;

  ~lufqhe_server() {
    if (d_walker_z) (void)hipFree(d_walker_z);
    if (d_perm) (void)hipFree(d_perm);
  }
};

namespace {

// Per-request record of a run.
struct Req {
  int64_t id;
  int walker;
  int n, particle, k, perm_start;
  double2 delta;
  double t_arrive;
  int64_t bytes;
};

// One in-flight batch slot: device buffers sized for the run's limits.
struct BatchSlot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  ReqDesc* h_desc = nullptr;   // pinned
  double2* h_sum_m = nullptr;  // pinned
  int32_t* h_sum_e = nullptr;  // pinned
  ReqDesc* d_desc = nullptr;
  int64_t* d_offs = nullptr;
  int64_t* d_perm_offs = nullptr;
  int64_t* d_out_idx = nullptr;
  double2* d_mant = nullptr;
  int32_t* d_exp = nullptr;
  double2* d_sum_m = nullptr;
  int32_t* d_sum_e = nullptr;
  void* d_block = nullptr;
  bool busy = false;
  std::vector<Req> reqs;
  std::vector<int32_t> slots;  // configuration slots to release on completion
  int64_t bytes = 0;
};

// Device bytes a request holds while in flight: per term its tables (24 B)
// and outputs (mant + exp, 20 B); its descriptor, sum and configuration slot.
inline int64_t request_bytes(int k) {
  return (int64_t)k * (24 + 20) + (int64_t)sizeof(ReqDesc) + 20 + kSlotZ * 16;
}

// Batch buffers of one in-flight slot and the configuration slots: device
// bytes a run allocates up front.
int64_t slot_bytes(int64_t R, int64_t D) {
  return align256(R * (int64_t)sizeof(ReqDesc)) + 3 * align256(D * 8) + align256(D * 16) +
         align256(D * 4) + align256(R * 16) + align256(R * 4);
}
int64_t arena_bytes(const lufqhe_serve_config& c, int walkers) {
  const int64_t slots = (int64_t)c.max_inflight * c.max_batch_requests + walkers;
  return c.max_inflight * slot_bytes(c.max_batch_requests, c.max_batch_dets) +
         slots * kSlotZ * (int64_t)sizeof(double2);
}

struct Run {
  lufqhe_server* s;
  lufqhe_serve_config cfg;
  int64_t arena_bytes = 0;
  std::vector<BatchSlot> bs;
  double2* d_slot_z = nullptr;
  std::vector<int32_t> free_slots;
  // walker state
  std::vector<int32_t> commit;   // slot with the accepted, uncommitted configuration
  std::vector<double> log_psi;   // log|psi| of the accepted configuration
  std::vector<char> busy;
  std::vector<std::deque<int64_t>> queue;  // arrived requests per walker
  std::deque<int> ready;                   // walkers with a queued request, FIFO by arrival

  ~Run() {
    for (auto& b : bs) {
      if (b.done) (void)hipEventDestroy(b.done);
      if (b.stream) (void)hipStreamDestroy(b.stream);
      if (b.h_desc) (void)hipHostFree(b.h_desc);
      if (b.h_sum_m) (void)hipHostFree(b.h_sum_m);
      if (b.h_sum_e) (void)hipHostFree(b.h_sum_e);
      if (b.d_block) (void)hipFree(b.d_block);
    }
    if (d_slot_z) (void)hipFree(d_slot_z);
  }

  hipError_t alloc() {
    const int nb = cfg.max_inflight;
    const int64_t R = cfg.max_batch_requests, D = cfg.max_batch_dets;
    bs.resize(nb);
    hipError_t err = hipSuccess;
    for (auto& b : bs) {
      if (err == hipSuccess) err = hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking);
      if (err == hipSuccess) err = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
      if (err == hipSuccess) err = hipHostMalloc((void**)&b.h_desc, R * sizeof(ReqDesc), 0);
      if (err == hipSuccess) err = hipHostMalloc((void**)&b.h_sum_m, R * sizeof(double2), 0);
      if (err == hipSuccess) err = hipHostMalloc((void**)&b.h_sum_e, R * sizeof(int32_t), 0);
      // one device block: desc | offs | perm_offs | out_idx | mant | exp | sums
      int64_t o = 0;
      const int64_t o_desc = o; o = align256(o + R * (int64_t)sizeof(ReqDesc));
      const int64_t o_offs = o; o = align256(o + D * 8);
      const int64_t o_poffs = o; o = align256(o + D * 8);
      const int64_t o_oidx = o; o = align256(o + D * 8);
      const int64_t o_mant = o; o = align256(o + D * 16);
      const int64_t o_exp = o; o = align256(o + D * 4);
      const int64_t o_sm = o; o = align256(o + R * 16);
      const int64_t o_se = o; o = align256(o + R * 4);
      if (err == hipSuccess) err = hipMalloc(&b.d_block, (size_t)o);
      if (err != hipSuccess) break;
      char* base = static_cast<char*>(b.d_block);
      b.d_desc = reinterpret_cast<ReqDesc*>(base + o_desc);
      b.d_offs = reinterpret_cast<int64_t*>(base + o_offs);
      b.d_perm_offs = reinterpret_cast<int64_t*>(base + o_poffs);
      b.d_out_idx = reinterpret_cast<int64_t*>(base + o_oidx);
      b.d_mant = reinterpret_cast<double2*>(base + o_mant);
      b.d_exp = reinterpret_cast<int32_t*>(base + o_exp);
      b.d_sum_m = reinterpret_cast<double2*>(base + o_sm);
      b.d_sum_e = reinterpret_cast<int32_t*>(base + o_se);
    }
    // configuration slots: every request in flight plus one held per walker
    const int64_t nslots = (int64_t)nb * R + s->walkers;
    if (err == hipSuccess) err = hipMalloc(&d_slot_z, nslots * kSlotZ * sizeof(double2));
    free_slots.clear();
    for (int64_t i = nslots - 1; i >= 0; --i) free_slots.push_back((int32_t)i);
    return err;
  }

  // Launch batch `b` holding b.reqs.
  hipError_t launch(BatchSlot& b) {
    const int nr = (int)b.reqs.size();
    // bucket positions: requests grouped by n in ascending n
    int64_t per_n[kMaxN + 1] = {};
    for (const auto& r : b.reqs) per_n[r.n] += r.k;
    int64_t start[kMaxN + 2] = {};
    for (int n = 1; n <= kMaxN; ++n) start[n + 1] = start[n] + per_n[n];
    int64_t fill[kMaxN + 2];
    std::memcpy(fill, start, sizeof(fill));
    int64_t out = 0;
    b.slots.clear();
    for (int j = 0; j < nr; ++j) {
      const Req& r = b.reqs[j];
      ReqDesc& d = b.h_desc[j];
      d.walker = r.walker;
      d.n = r.n;
      d.particle = r.particle;
      d.commit_slot = commit[r.walker];
      d.slot = free_slots.back();
      free_slots.pop_back();
      d.k = r.k;
      d.perm_start = r.perm_start;
      d.pad = 0;
      d.table_pos = fill[r.n];
      fill[r.n] += r.k;
      d.out_base = out;
      out += r.k;
      d.delta = r.delta;
      b.slots.push_back(d.slot);
    }
    hipError_t err = hipMemcpyAsync(b.d_desc, b.h_desc, nr * sizeof(ReqDesc),
                                    hipMemcpyHostToDevice, b.stream);
    if (err == hipSuccess) {
      hipLaunchKernelGGL(serve_prepare_kernel, dim3(nr), dim3(256), 0, b.stream, b.d_desc,
                         s->d_walker_z, d_slot_z, s->perms, b.d_offs, b.d_perm_offs,
                         b.d_out_idx);
      err = hipGetLastError();
    }
    for (int n = 1; n <= kMaxN && err == hipSuccess; ++n) {
      if (!per_n[n]) continue;
      const KernelEntry& e = entry_for(s->variant, n);
      BatchArgs a{};
      a.A = d_slot_z;
      a.count = per_n[n];
      a.ld = n;
      a.offs = b.d_offs + start[n];
      a.out_idx = b.d_out_idx + start[n];
      a.row_perm = s->d_perm;
      a.perm_offs = b.d_perm_offs + start[n];
      a.src = SRC_FQHE;
      a.out = DetOut{b.d_mant, b.d_exp, nullptr, nullptr, nullptr, nullptr};
      err = e.launch(a, plain_blocks(e, per_n[n]), b.stream);
    }
    if (err == hipSuccess) {
      hipLaunchKernelGGL(serve_sum_kernel, dim3(nr), dim3(64), 0, b.stream, b.d_desc, b.d_mant,
                         b.d_exp, b.d_sum_m, b.d_sum_e);
      err = hipGetLastError();
    }
    if (err == hipSuccess)
      err = hipMemcpyAsync(b.h_sum_m, b.d_sum_m, nr * sizeof(double2), hipMemcpyDeviceToHost,
                           b.stream);
    if (err == hipSuccess)
      err = hipMemcpyAsync(b.h_sum_e, b.d_sum_e, nr * sizeof(int32_t), hipMemcpyDeviceToHost,
                           b.stream);
    if (err == hipSuccess) err = hipEventRecord(b.done, b.stream);
    return err;
  }
};

}  // namespace

extern "C" {

int32_t lufqhe_server_create(int32_t walkers, const int32_t* h_walker_n, int32_t perms_per_walker,
                             uint64_t seed, int32_t policy, lufqhe_server** out) {
  if (!out || walkers < 1 || !h_walker_n || perms_per_walker < 1) return LUFQHE_ERROR_INVALID_VALUE;
  *out = nullptr;
  for (int w = 0; w < walkers; ++w)
    if (h_walker_n[w] < 1 || h_walker_n[w] > kMaxN) return LUFQHE_ERROR_INVALID_VALUE;
  const int v = variant_of(policy, ARITH_FUSED);
  if (v < 0) return LUFQHE_ERROR_UNSUPPORTED;
  for (int w = 0; w < walkers; ++w)
    if (!entry_for(v, h_walker_n[w]).launch) return LUFQHE_ERROR_NOT_BUILT;
  lufqhe_server* s = new (std::nothrow) lufqhe_server();
  if (!s) return LUFQHE_ERROR_HIP;
  s->variant = v;
  s->walkers = walkers;
  s->perms = perms_per_walker;
  s->seed = seed;
  s->wn.assign(h_walker_n, h_walker_n + walkers);
  // initial positions uniform in the disk of radius sqrt(2n); permutation
  // tables by Fisher-Yates, both from the counter-based generator
  s->z0.assign((size_t)walkers * kSlotZ, make_double2(0.0, 0.0));
  std::vector<uint8_t> perm((size_t)walkers * perms_per_walker * kSlotZ, 0);
  for (int w = 0; w < walkers; ++w) {
    const int n = s->wn[w];
    const double R = std::sqrt(2.0 * n);
    for (int i = 0; i < n; ++i) {
      const double u = unif(seed, 0x5EED0000ull + w, 2 * i);
      const double a = 6.283185307179586 * unif(seed, 0x5EED0000ull + w, 2 * i + 1);
      const double r = R * std::sqrt(u);
      s->z0[(size_t)w * kSlotZ + i] = make_double2(r * std::cos(a), r * std::sin(a));
    }
    for (int p = 0; p < perms_per_walker; ++p) {
      uint8_t* q = &perm[((size_t)w * perms_per_walker + p) * kSlotZ];
      for (int i = 0; i < n; ++i) q[i] = (uint8_t)i;
      for (int i = n - 1; i > 0; --i) {
        const int j = (int)(rnd(seed, 0x9E370000ull + w, (uint64_t)p * kSlotZ + i) % (uint64_t)(i + 1));
        std::swap(q[i], q[j]);
      }
    }
  }
  hipError_t err = hipGetDevice(&s->device);
  const size_t zb = s->z0.size() * sizeof(double2);
  if (err == hipSuccess) err = hipMalloc(&s->d_walker_z, zb);
  if (err == hipSuccess) err = hipMalloc(&s->d_perm, perm.size());
  if (err == hipSuccess) err = hipMemcpy(s->d_walker_z, s->z0.data(), zb, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(s->d_perm, perm.data(), perm.size(), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    delete s;
    return hip_fail(err);
  }
  s->resident_bytes = (int64_t)(zb + perm.size());
  *out = s;
  return LUFQHE_SUCCESS;
}

int32_t lufqhe_server_destroy(lufqhe_server* s) {
  delete s;
  return LUFQHE_SUCCESS;
}

int32_t lufqhe_server_reset(lufqhe_server* s) {
  if (!s) return LUFQHE_ERROR_INVALID_VALUE;
  const hipError_t err = hipMemcpy(s->d_walker_z, s->z0.data(), s->z0.size() * sizeof(double2),
                                   hipMemcpyHostToDevice);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

int32_t lufqhe_server_walker(const lufqhe_server* s, int32_t walker, int32_t* n, void* h_z) {
  if (!s || walker < 0 || walker >= s->walkers) return LUFQHE_ERROR_INVALID_VALUE;
  if (n) *n = s->wn[walker];
  if (!h_z) return LUFQHE_SUCCESS;
  const hipError_t err = hipMemcpy(h_z, s->d_walker_z + (size_t)walker * kSlotZ,
                                   s->wn[walker] * sizeof(double2), hipMemcpyDeviceToHost);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

int32_t lufqhe_server_perms(const lufqhe_server* s, int32_t walker, uint8_t* h_perm) {
  if (!s || walker < 0 || walker >= s->walkers || !h_perm) return LUFQHE_ERROR_INVALID_VALUE;
  const hipError_t err =
      hipMemcpy(h_perm, s->d_perm + (size_t)walker * s->perms * kSlotZ,
                (size_t)s->perms * kSlotZ, hipMemcpyDeviceToHost);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

int32_t lufqhe_server_run(lufqhe_server* s, const lufqhe_serve_config* cfg,
                          lufqhe_serve_stats* stats, lufqhe_serve_record* records) {
  if (!s || !cfg || cfg->requests < 0) return LUFQHE_ERROR_INVALID_VALUE;
  if (cfg->k_min < 1 || cfg->k_max < cfg->k_min || cfg->max_inflight < 1 ||
      cfg->max_inflight > LUFQHE_SERVE_MAX_INFLIGHT || cfg->max_batch_requests < 1 ||
      cfg->max_batch_dets < cfg->k_max || cfg->hbm_budget <= 0)
    return LUFQHE_ERROR_INVALID_VALUE;
  const int64_t arena = arena_bytes(*cfg, s->walkers);
  if (s->resident_bytes + arena + request_bytes(cfg->k_max) > cfg->hbm_budget)
    return LUFQHE_ERROR_OUT_OF_BUDGET;
  int cur = 0;
  hipError_t err = hipGetDevice(&cur);
  if (err != hipSuccess) return hip_fail(err);
  if (cur != s->device) return LUFQHE_ERROR_INVALID_VALUE;

  Run run;
  run.s = s;
  run.cfg = *cfg;
  run.arena_bytes = arena;
  err = run.alloc();
  if (err != hipSuccess) return hip_fail(err);
  const int W = s->walkers;
  run.commit.assign(W, -1);
  run.log_psi.assign(W, -INFINITY);
  run.busy.assign(W, 0);
  run.queue.assign(W, {});

  // the request stream: request r belongs to walker r mod W
  const int64_t NR = cfg->requests;
  std::vector<Req> reqs(NR);
  double t = 0.0;
  for (int64_t r = 0; r < NR; ++r) {
    Req& q = reqs[r];
    q.id = r;
    q.walker = (int)(r % W);
    q.n = s->wn[q.walker];
    q.particle = (int)(rnd(s->seed ^ cfg->seed, r, 1) % (uint64_t)q.n);
    q.delta = make_double2(0.3 * gauss(s->seed ^ cfg->seed, r, 2),
                           0.3 * gauss(s->seed ^ cfg->seed, r, 3));
    q.k = cfg->k_min + (int)(rnd(s->seed ^ cfg->seed, r, 4) % (uint64_t)(cfg->k_max - cfg->k_min + 1));
    q.perm_start = (int)(rnd(s->seed ^ cfg->seed, r, 5) % (uint64_t)s->perms);
    if (cfg->rate > 0) t += -std::log(unif(s->seed ^ cfg->seed, r, 6)) / cfg->rate;
    q.t_arrive = t;
    q.bytes = request_bytes(q.k);
  }
  std::vector<double> latency(NR, 0.0);

  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto now = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
  int64_t next_arrival = 0, completed = 0, dets = 0, batches = 0, budget_waits = 0;
  int64_t inflight_bytes = 0, peak_bytes = 0, batch_dets_sum = 0, last_blocked = -1;
  // what the requests in flight may hold: the budget less the resident
  // walker state and the batch buffers
  const int64_t budget = cfg->hbm_budget - s->resident_bytes - run.arena_bytes;
  double t_first = NR ? reqs[0].t_arrive : 0.0, t_last = 0.0;

  while (completed < NR && err == hipSuccess) {
    bool progressed = false;
    const double tn = now();
    // 1. arrivals
    while (next_arrival < NR && reqs[next_arrival].t_arrive <= tn) {
      const Req& q = reqs[next_arrival];
      if (run.queue[q.walker].empty() && !run.busy[q.walker]) run.ready.push_back(q.walker);
      run.queue[q.walker].push_back(next_arrival);
      ++next_arrival;
      progressed = true;
    }
    // 2. completions
    for (auto& b : run.bs) {
      if (!b.busy) continue;
      const hipError_t q = hipEventQuery(b.done);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) {
        err = q;
        break;
      }
      const double tc = now();
      for (size_t j = 0; j < b.reqs.size(); ++j) {
        const Req& r = b.reqs[j];
        const int w = r.walker;
        const double2 m = b.h_sum_m[j];
        const int e = b.h_sum_e[j];
        const double am = std::hypot(m.x, m.y);
        const double lp = am == 0.0 ? -INFINITY : std::log(am) + e * 0.6931471805599453;
        const double u = unif(s->seed ^ cfg->seed, r.id, 7);
        const bool accept = (lp - run.log_psi[w]) * 2.0 >= std::log(u);
        // the configuration this request read from its commit slot is now in walker_z
        if (b.h_desc[j].commit_slot >= 0) run.free_slots.push_back(b.h_desc[j].commit_slot);
        if (accept) {
          run.commit[w] = b.h_desc[j].slot;
          run.log_psi[w] = lp;
        } else {
          run.commit[w] = -1;
          run.free_slots.push_back(b.h_desc[j].slot);
        }
        latency[r.id] = tc - r.t_arrive;
        if (records) {
          lufqhe_serve_record& rec = records[r.id];
          rec.walker = w;
          rec.particle = r.particle;
          rec.k = r.k;
          rec.perm_start = r.perm_start;
          rec.accepted = accept ? 1 : 0;
          rec.sum_exp = e;
          std::memcpy(rec.delta, &r.delta, sizeof(rec.delta));
          std::memcpy(rec.sum_mant, &m, sizeof(rec.sum_mant));
          rec.latency_s = tc - r.t_arrive;
        }
        run.busy[w] = 0;
        if (!run.queue[w].empty()) run.ready.push_back(w);
        ++completed;
      }
      inflight_bytes -= b.bytes;
      b.busy = false;
      b.reqs.clear();
      t_last = tc;
      progressed = true;
    }
    if (err != hipSuccess) break;
    // 3. admission and launch: fill a free batch slot from the ready walkers
    for (auto& b : run.bs) {
      if (b.busy || run.ready.empty()) continue;
      int64_t bdets = 0, bbytes = 0;
      while (!run.ready.empty() && (int64_t)b.reqs.size() < cfg->max_batch_requests) {
        const int w = run.ready.front();
        const Req& q = reqs[run.queue[w].front()];
        if (bdets + q.k > cfg->max_batch_dets) break;
        if (inflight_bytes + bbytes + q.bytes > budget) {
          if (q.id != last_blocked) ++budget_waits;  // requests held back by the budget
          last_blocked = q.id;
          break;
        }
        // a walker whose previous request is still in flight keeps a held
        // slot; one configuration slot per request is always available
        run.ready.pop_front();
        run.queue[w].pop_front();
        run.busy[w] = 1;
        b.reqs.push_back(q);
        bdets += q.k;
        bbytes += q.bytes;
      }
      if (b.reqs.empty()) break;
      b.bytes = bbytes;
      inflight_bytes += bbytes;
      peak_bytes = std::max(peak_bytes, inflight_bytes);
      err = run.launch(b);
      if (err != hipSuccess) break;
      b.busy = true;
      ++batches;
      dets += bdets;
      batch_dets_sum += bdets;
      progressed = true;
    }
    if (!progressed) {
      // idle: wait a little for the next arrival or completion
      const double wait = next_arrival < NR ? reqs[next_arrival].t_arrive - now() : 1e-5;
      if (wait > 2e-5) std::this_thread::sleep_for(std::chrono::microseconds(
          (int64_t)std::min(wait * 1e6 * 0.5, 200.0)));
    }
  }
  if (err != hipSuccess) {
    for (auto& b : run.bs)
      if (b.stream) (void)hipStreamSynchronize(b.stream);
    return hip_fail(err);
  }
  // commit the accepted, uncommitted configurations so the walkers hold
  // their final state
  {
    std::vector<ReqDesc> fin;
    for (int w = 0; w < W; ++w) {
      if (run.commit[w] < 0) continue;
      ReqDesc d{};
      d.walker = w;
      d.n = s->wn[w];
      d.particle = -1;
      d.commit_slot = run.commit[w];
      d.slot = run.commit[w];
      d.k = 0;
      fin.push_back(d);
    }
    if (!fin.empty()) {
      BatchSlot& b = run.bs[0];
      for (size_t i = 0; i < fin.size(); i += cfg->max_batch_requests) {
        const int nr = (int)std::min<size_t>(cfg->max_batch_requests, fin.size() - i);
        err = hipMemcpy(b.d_desc, fin.data() + i, nr * sizeof(ReqDesc), hipMemcpyHostToDevice);
        if (err != hipSuccess) break;
        hipLaunchKernelGGL(serve_prepare_kernel, dim3(nr), dim3(256), 0, b.stream, b.d_desc,
                           s->d_walker_z, run.d_slot_z, s->perms, b.d_offs, b.d_perm_offs,
                           b.d_out_idx);
        err = hipGetLastError();
        if (err == hipSuccess) err = hipStreamSynchronize(b.stream);
        if (err != hipSuccess) break;
      }
    }
  }
  if (err != hipSuccess) return hip_fail(err);

  if (stats) {
    std::vector<double> lat(latency);
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      if (lat.empty()) return 0.0;
      const size_t i = (size_t)std::min<double>(lat.size() - 1, std::floor(p * (lat.size() - 1) + 0.5));
      return lat[i];
    };
    double mean = 0.0;
    for (double v : lat) mean += v;
    stats->requests = NR;
    stats->dets = dets;
    stats->batches = batches;
    stats->wall_s = t_last - t_first;
    stats->dets_per_s = stats->wall_s > 0 ? dets / stats->wall_s : 0.0;
    stats->requests_per_s = stats->wall_s > 0 ? NR / stats->wall_s : 0.0;
    stats->offered_rate = cfg->rate;
    stats->p50_s = pct(0.50);
    stats->p99_s = pct(0.99);
    stats->max_s = lat.empty() ? 0.0 : lat.back();
    stats->mean_s = lat.empty() ? 0.0 : mean / lat.size();
    stats->mean_batch_dets = batches ? (double)batch_dets_sum / batches : 0.0;
    stats->budget_waits = budget_waits;
    stats->peak_inflight_bytes = peak_bytes;
    stats->resident_bytes = s->resident_bytes + run.arena_bytes;
  }
  return LUFQHE_SUCCESS;
}

}  // extern "C"
