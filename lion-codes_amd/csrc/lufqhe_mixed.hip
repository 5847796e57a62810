// lufqhe_mixed.hip -- mixed-size batches (n = 1..64 in one batch).
//
// The reference handles one matrix size per run (RunInfo, LU-fqhe/src/
// RunInfo.cpp:46-78); a batch of mixed sizes there means one CALL per size
// on repacked uniform batches.  Here a plan buckets the batch by n once:
//   * per-bucket tables of element offsets and output slots live on the
//     device (one allocation, made when the plan is created);
//   * each bucket is one launch of that n's tuned kernel with indexed
//     addressing (BatchArgs::offs / out_idx), so nothing is repacked;
//   * buckets are spread over the plan's streams largest-cost-first
//     (longest-processing-time assignment on F(n) * count) and forked from /
//     joined to the caller's stream with events, so small-n buckets run
//     beside the big ones instead of after them;
//   * the whole fork/launch/join sequence is captured once into a hipGraph
//     and replayed: one graph launch per batch.
// Mode PERSISTENT replaces the per-bucket launches by one persistent kernel
// per size class (lup_persistent.hpp): the buckets of a class form one
// device work queue, drained largest n first by resident waves; the queue
// counters are zeroed by a small kernel at the head of the captured graph.
// Mode PLAIN runs the same kernels one size after another on the caller's
// stream: the uniform-launch-per-size form it is compared with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/lufqhe.h"
#include "lup_dispatch.hpp"

namespace lufqhe {
int32_t hip_fail(hipError_t e);
const KernelEntry& entry_for(int variant, int n);
int64_t plain_blocks(const KernelEntry& e, int64_t count);
}  // namespace lufqhe

using namespace lufqhe;

struct lufqhe_mixed_plan {
  struct Bucket {
    int n = 0;
    int64_t start = 0, count = 0;
    double cost = 0.0;
    int stream = 0;
  };
  int device = 0;
  int variant = 0;
  int64_t count = 0;
  int64_t* d_tables = nullptr;  // [count] offsets sorted by bucket, then [count] output slots
  std::vector<Bucket> buckets;  // in launch order
  int nstreams = 1;
  hipStream_t streams[LUFQHE_MIXED_MAX_STREAMS] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[LUFQHE_MIXED_MAX_STREAMS] = {};
  // recorded after every persistent replay: the per-bucket arguments those
  // replays read are rewritten only once it has completed
  hipEvent_t replayed = nullptr;
  bool replay_pending = false;
  // persistent mode: per class, its buckets' queue and launch size
  struct PClass {
    int cls = 0;
    int64_t first = 0;  // first bucket of the class in the device arrays
    int32_t nbuckets = 0;
    int64_t total_tasks = 0;
    int64_t blocks = 0;
    double cost = 0.0;
    int stream = 0;
  };
  std::vector<PClass> pclasses;           // in launch order
  std::vector<BatchArgs> h_bargs;         // per bucket, in class order
  std::vector<int32_t> h_bucket_index;    // bargs slot -> plan bucket
  void* d_persist = nullptr;              // PersistBucket[] | BatchArgs[] | counters
  PersistBucket* d_pbuckets = nullptr;
  BatchArgs* d_bargs = nullptr;
  uint32_t* d_counters = nullptr;
  const void* p_A = nullptr;              // pointers d_bargs holds
  int p_src = -1;
  DetOut p_out{};
  // captured graph and what it was captured for
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* g_A = nullptr;
  int g_src = 0;
  int g_mode = -1;
  DetOut g_out{};

  ~lufqhe_mixed_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int s = 0; s < nstreams; ++s) {
      if (join[s]) (void)hipEventDestroy(join[s]);
      if (streams[s]) (void)hipStreamDestroy(streams[s]);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (replayed) (void)hipEventDestroy(replayed);
    if (d_tables) (void)hipFree(d_tables);
    if (d_persist) (void)hipFree(d_persist);
  }
};

namespace {

double det_flops(int n) {
  return 8.0 * n * n * n / 3.0 - n * n + 13.0 * n / 3.0 + 6.0 * (n - 1);
}

hipError_t launch_bucket(const lufqhe_mixed_plan& p, const lufqhe_mixed_plan::Bucket& b,
                         const void* d_A, int src, const DetOut& out, hipStream_t stream) {
  const KernelEntry& e = entry_for(p.variant, b.n);
  BatchArgs args{};
  args.A = static_cast<const double2*>(d_A);
  args.count = b.count;
  args.ld = b.n;
  args.stride = 0;
  args.offs = p.d_tables + b.start;
  args.out_idx = p.d_tables + p.count + b.start;
  args.piv_stride = 0;
  args.src = src;
  args.out = out;
  return e.launch(args, plain_blocks(e, b.count), stream);
}

// Fork from `stream`, one launch sequence per plan stream, join back.
hipError_t issue_forked(const lufqhe_mixed_plan& p, const void* d_A, int src,
                        const DetOut& out, hipStream_t stream) {
  hipError_t err = hipEventRecord(p.fork, stream);
  for (int s = 0; s < p.nstreams && err == hipSuccess; ++s)
    err = hipStreamWaitEvent(p.streams[s], p.fork, 0);
  for (const auto& b : p.buckets) {
    if (err != hipSuccess) break;
    err = launch_bucket(p, b, d_A, src, out, p.streams[b.stream]);
  }
  for (int s = 0; s < p.nstreams && err == hipSuccess; ++s) {
    err = hipEventRecord(p.join[s], p.streams[s]);
    if (err == hipSuccess) err = hipStreamWaitEvent(stream, p.join[s], 0);
  }
  return err;
}

bool same_out(const DetOut& a, const DetOut& b) {
  return a.mant == b.mant && a.expo == b.expo && a.logabs == b.logabs && a.phase == b.phase &&
         a.info == b.info && a.piv == b.piv;
}

__global__ void __launch_bounds__(64) persist_reset_kernel(uint32_t* __restrict__ counters, int n) {
  if ((int)threadIdx.x < n) counters[threadIdx.x] = 0u;
}

// The class queues of a plan: buckets grouped by size class, largest n
// first within a class, with their task ranges; classes ordered and spread
// over the plan streams by cost.  Device arrays are allocated here; the
// per-bucket launch arguments are written when the pointers are known.
hipError_t build_persistent(lufqhe_mixed_plan& p) {
  if (p.variant >= kPersistVariants) return hipSuccess;  // no persistent kernels for it
  int cus = 0;
  hipError_t err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p.device);
  if (err != hipSuccess) return err;
  std::vector<PersistBucket> hb;
  for (int c = 0; c < kPersistClasses; ++c) {
    std::vector<int> idx;
    for (int i = 0; i < (int)p.buckets.size(); ++i)
      if (persist_class(p.buckets[i].n) == c) idx.push_back(i);
    if (idx.empty()) continue;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return p.buckets[a].n > p.buckets[b].n; });
    lufqhe_mixed_plan::PClass pc;
    pc.cls = c;
    pc.first = (int64_t)hb.size();
    pc.nbuckets = (int32_t)idx.size();
    for (int i : idx) {
      const auto& b = p.buckets[i];
      const int g = persist_group(b.n);
      hb.push_back(PersistBucket{b.n, 0, pc.total_tasks});
      p.h_bucket_index.push_back(i);
      pc.total_tasks += (b.count + g - 1) / g;
      pc.cost += b.cost;
    }
    // the class counter is 32-bit and every wave adds one past the end: a
    // class too large for it gets no persistent launch (the mode is refused)
    if (pc.total_tasks >= (int64_t{1} << 31)) {
      p.pclasses.clear();
      p.h_bucket_index.clear();
      return hipSuccess;
    }
    const PersistEntry& e = g_persist[p.variant][c];
    if (!e.launch) return hipErrorInvalidDeviceFunction;
    int per_cu = 0;
    err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, e.kernel, 64, 0);
    if (err != hipSuccess) return err;
    pc.blocks = std::min<int64_t>(pc.total_tasks, (int64_t)std::max(1, per_cu) * cus);
    p.pclasses.push_back(pc);
  }
  std::sort(p.pclasses.begin(), p.pclasses.end(),
            [](const auto& a, const auto& b) { return a.cost > b.cost; });
  std::vector<double> load(p.nstreams, 0.0);
  for (auto& pc : p.pclasses) {
    const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    pc.stream = s;
    load[s] += pc.cost;
  }
  p.h_bargs.assign(hb.size(), BatchArgs{});
  const size_t nb = hb.size();
  const size_t o_bargs = (nb * sizeof(PersistBucket) + 255) & ~(size_t)255;
  const size_t o_cnt = (o_bargs + nb * sizeof(BatchArgs) + 255) & ~(size_t)255;
  if (nb == 0) return hipSuccess;
  err = hipMalloc(&p.d_persist, o_cnt + kPersistClasses * sizeof(uint32_t));
  if (err != hipSuccess) return err;
  char* base = static_cast<char*>(p.d_persist);
  p.d_pbuckets = reinterpret_cast<PersistBucket*>(base);
  p.d_bargs = reinterpret_cast<BatchArgs*>(base + o_bargs);
  p.d_counters = reinterpret_cast<uint32_t*>(base + o_cnt);
  return hipMemcpy(p.d_pbuckets, hb.data(), nb * sizeof(PersistBucket), hipMemcpyHostToDevice);
}

// Per-bucket launch arguments for these pointers (synchronous, outside any
// capture; the caller has drained `stream`, the only user of the old ones).
hipError_t write_persistent_args(lufqhe_mixed_plan& p, const void* d_A, int src,
                                 const DetOut& out) {
  for (size_t j = 0; j < p.h_bargs.size(); ++j) {
    const auto& b = p.buckets[p.h_bucket_index[j]];
    BatchArgs& a = p.h_bargs[j];
    a = BatchArgs{};
    a.A = static_cast<const double2*>(d_A);
    a.count = b.count;
    a.ld = b.n;
    a.stride = 0;
    a.offs = p.d_tables + b.start;
    a.out_idx = p.d_tables + p.count + b.start;
    a.piv_stride = 0;
    a.src = src;
    a.out = out;
  }
  hipError_t err = hipMemcpy(p.d_bargs, p.h_bargs.data(), p.h_bargs.size() * sizeof(BatchArgs),
                             hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    p.p_A = d_A;
    p.p_src = src;
    p.p_out = out;
  }
  return err;
}

// Reset the class counters, fork, one persistent launch per class, join.
hipError_t issue_persistent(const lufqhe_mixed_plan& p, hipStream_t stream) {
  hipLaunchKernelGGL(persist_reset_kernel, dim3(1), dim3(64), 0, stream, p.d_counters,
                     kPersistClasses);
  hipError_t err = hipGetLastError();
  if (err == hipSuccess) err = hipEventRecord(p.fork, stream);
  for (int s = 0; s < p.nstreams && err == hipSuccess; ++s)
    err = hipStreamWaitEvent(p.streams[s], p.fork, 0);
  for (const auto& pc : p.pclasses) {
    if (err != hipSuccess) break;
    PersistArgs pa{};
    pa.buckets = p.d_pbuckets + pc.first;
    pa.bargs = p.d_bargs + pc.first;
    pa.nbuckets = pc.nbuckets;
    pa.counter = p.d_counters + pc.cls;
    pa.total_tasks = pc.total_tasks;
    err = g_persist[p.variant][pc.cls].launch(pa, pc.blocks, p.streams[pc.stream]);
  }
  for (int s = 0; s < p.nstreams && err == hipSuccess; ++s) {
    err = hipEventRecord(p.join[s], p.streams[s]);
    if (err == hipSuccess) err = hipStreamWaitEvent(stream, p.join[s], 0);
  }
  return err;
}

}  // namespace

extern "C" {

int32_t lufqhe_mixed_plan_create(const int32_t* h_n, const int64_t* h_offs, int64_t count,
                                 int32_t policy, int32_t nstreams, lufqhe_mixed_plan** out) {
  if (!out || count < 0 || (count > 0 && (!h_n || !h_offs))) return LUFQHE_ERROR_INVALID_VALUE;
  if (nstreams < 1 || nstreams > LUFQHE_MIXED_MAX_STREAMS) return LUFQHE_ERROR_INVALID_VALUE;
  *out = nullptr;
  const int v = variant_of(policy, ARITH_FUSED);
  if (v < 0) return LUFQHE_ERROR_UNSUPPORTED;
  std::vector<int64_t> per_n(kMaxN + 1, 0);
  for (int64_t i = 0; i < count; ++i) {
    if (h_n[i] < 1 || h_n[i] > kMaxN || h_offs[i] < 0) return LUFQHE_ERROR_INVALID_VALUE;
    ++per_n[h_n[i]];
  }
  lufqhe_mixed_plan* p = new (std::nothrow) lufqhe_mixed_plan();
  if (!p) return LUFQHE_ERROR_HIP;
  p->variant = v;
  p->count = count;
  p->nstreams = nstreams;
  hipError_t err = hipGetDevice(&p->device);

  // bucket starts and host tables sorted by n
  std::vector<int64_t> start(kMaxN + 2, 0);
  for (int n = 1; n <= kMaxN; ++n) start[n + 1] = start[n] + per_n[n];
  std::vector<int64_t> tables(2 * (size_t)count);
  std::vector<int64_t> fill(start.begin(), start.end());
  for (int64_t i = 0; i < count; ++i) {
    const int64_t slot = fill[h_n[i]]++;
    tables[slot] = h_offs[i];
    tables[count + slot] = i;
  }
  for (int n = 1; n <= kMaxN; ++n) {
    if (!per_n[n]) continue;
    lufqhe_mixed_plan::Bucket b;
    b.n = n;
    b.start = start[n];
    b.count = per_n[n];
    b.cost = det_flops(n) * (double)per_n[n] + 16.0 * n * n * per_n[n];
    p->buckets.push_back(b);
  }
  // longest-processing-time assignment, largest first within each stream
  std::sort(p->buckets.begin(), p->buckets.end(),
            [](const auto& a, const auto& b) { return a.cost > b.cost; });
  std::vector<double> load(nstreams, 0.0);
  for (auto& b : p->buckets) {
    const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b.stream = s;
    load[s] += b.cost;
  }

  if (err == hipSuccess && count > 0) err = hipMalloc(&p->d_tables, tables.size() * sizeof(int64_t));
  if (err == hipSuccess && count > 0)
    err = hipMemcpy(p->d_tables, tables.data(), tables.size() * sizeof(int64_t),
                    hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
  if (err == hipSuccess) err = hipEventCreateWithFlags(&p->replayed, hipEventDisableTiming);
  for (int s = 0; s < nstreams && err == hipSuccess; ++s) {
    err = hipStreamCreateWithFlags(&p->streams[s], hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&p->join[s], hipEventDisableTiming);
  }
  if (err != hipSuccess) {
    delete p;
    return hip_fail(err);
  }
  for (const auto& b : p->buckets)
    if (!entry_for(v, b.n).launch) {
      delete p;
      return LUFQHE_ERROR_NOT_BUILT;
    }
  err = build_persistent(*p);
  if (err != hipSuccess) {
    delete p;
    return hip_fail(err);
  }
  *out = p;
  return LUFQHE_SUCCESS;
}

int32_t lufqhe_mixed_plan_destroy(lufqhe_mixed_plan* plan) {
  delete plan;
  return LUFQHE_SUCCESS;
}

int32_t lufqhe_mixed_plan_buckets(const lufqhe_mixed_plan* plan, int32_t* n_out,
                                  int64_t* count_out, int32_t* stream_out, int32_t capacity) {
  if (!plan) return -1;
  const int nb = (int)plan->buckets.size();
  for (int i = 0; i < nb && i < capacity; ++i) {
    if (n_out) n_out[i] = plan->buckets[i].n;
    if (count_out) count_out[i] = plan->buckets[i].count;
    if (stream_out) stream_out[i] = plan->buckets[i].stream;
  }
  return nb;
}

int32_t lufqhe_mixed_execute(lufqhe_mixed_plan* plan, const void* d_A, void* d_mant,
                             int32_t* d_exp, double* d_logabs, double* d_phase, int32_t* d_info,
                             int32_t source, int32_t mode, hipStream_t stream) {
  if (!plan) return LUFQHE_ERROR_INVALID_VALUE;
  if (source != LUFQHE_SOURCE_MATRIX && source != LUFQHE_SOURCE_FQHE)
    return LUFQHE_ERROR_INVALID_VALUE;
  const int src = source == LUFQHE_SOURCE_FQHE ? SRC_FQHE : SRC_MATRIX;
  if (plan->count == 0) return LUFQHE_SUCCESS;
  if (!d_A || !d_mant) return LUFQHE_ERROR_INVALID_VALUE;
  // the plan's streams, events and tables belong to the device it was made on
  int cur = 0;
  hipError_t derr = hipGetDevice(&cur);
  if (derr != hipSuccess) return hip_fail(derr);
  if (cur != plan->device) return LUFQHE_ERROR_INVALID_VALUE;
  const DetOut out{static_cast<double2*>(d_mant), d_exp, d_logabs, d_phase, d_info, nullptr};
  hipError_t err = hipSuccess;
  if (mode == LUFQHE_MIXED_PLAIN) {
    // one size after another, on the caller's stream, largest n first
    std::vector<lufqhe_mixed_plan::Bucket> order = plan->buckets;
    std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.n > b.n; });
    for (const auto& b : order) {
      err = launch_bucket(*plan, b, d_A, src, out, stream);
      if (err != hipSuccess) break;
    }
  } else if (mode == LUFQHE_MIXED_STREAMS) {
    err = issue_forked(*plan, d_A, src, out, stream);
  } else if (mode == LUFQHE_MIXED_GRAPH || mode == LUFQHE_MIXED_PERSISTENT) {
    if (stream == nullptr) return LUFQHE_ERROR_INVALID_VALUE;  // no capture on stream 0
    const bool persistent = mode == LUFQHE_MIXED_PERSISTENT;
    if (persistent && (plan->variant >= kPersistVariants || plan->pclasses.empty()))
      return LUFQHE_ERROR_UNSUPPORTED;
    if (!plan->exec || plan->g_mode != mode || plan->g_A != d_A || plan->g_src != src ||
        !same_out(plan->g_out, out)) {
      if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
      if (plan->graph) (void)hipGraphDestroy(plan->graph);
      plan->exec = nullptr;
      plan->graph = nullptr;
      plan->g_mode = -1;
      if (persistent && (plan->p_A != d_A || plan->p_src != src || !same_out(plan->p_out, out))) {
        // earlier replays, on whatever stream, may still read the old
        // arguments: wait for the last one
        if (plan->replay_pending) err = hipEventSynchronize(plan->replayed);
        if (err == hipSuccess) err = write_persistent_args(*plan, d_A, src, out);
      }
      if (err == hipSuccess) err = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
      if (err == hipSuccess) {
        const hipError_t e2 = persistent ? issue_persistent(*plan, stream)
                                         : issue_forked(*plan, d_A, src, out, stream);
        err = hipStreamEndCapture(stream, &plan->graph);
        if (e2 != hipSuccess) err = e2;
      }
      if (err == hipSuccess) err = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
      if (err == hipSuccess) {
        plan->g_A = d_A;
        plan->g_src = src;
        plan->g_mode = mode;
        plan->g_out = out;
      }
    }
    if (err == hipSuccess) err = hipGraphLaunch(plan->exec, stream);
    if (err == hipSuccess && persistent) {
      err = hipEventRecord(plan->replayed, stream);
      plan->replay_pending = err == hipSuccess;
    }
  } else {
    return LUFQHE_ERROR_INVALID_VALUE;
  }
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

}  // extern "C"
