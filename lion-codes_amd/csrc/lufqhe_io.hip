// lufqhe_io.hip -- the reference's matrix input file, read onto the device.
//
// LU-fqhe's LU mode reads its matrices from one raw file
// (LU-fqhe/src/PermHandler.cpp:91-170): nElem matrices of side n, each
// element (re, im) as two float32 (double with __DPREC__), row major, no
// padding, no header.  Every MPI rank reads one contiguous run of scalars,
// bounds computed in float arithmetic:
//   start = (int)((float)(rank * total) / nodes)
//   end   = (int)((float)((rank + 1) * total) / nodes) - 1,  total = 2*nElem*n*n
// and the Monte Carlo loop copies that run to the device as it is
// (LU-fqhe/src/MonteCarloLoop.cu:316).
//
// Here the run is streamed through two pinned staging buffers: the next
// block is read from the file while the previous one is copied to the
// device, and complex64 input crosses the bus as float32 (half the bytes)
// and is widened to complex128 by a kernel on the device.
#include <hip/hip_runtime.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>

#include "../../include/lufqhe.h"

namespace lufqhe {
int32_t hip_fail(hipError_t e);
}  // namespace lufqhe

using namespace lufqhe;

namespace {

constexpr int64_t kStageBytes = 32ll << 20;  // per staging buffer

__global__ void __launch_bounds__(256) widen_kernel(const float* __restrict__ in, int64_t count,
                                                    double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (double)in[i];
}

struct Staging {
  void* host[2] = {nullptr, nullptr};
  float* dev[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  int fd = -1;
  ~Staging() {
    for (int b = 0; b < 2; ++b) {
      if (done[b]) {
        (void)hipEventSynchronize(done[b]);
        (void)hipEventDestroy(done[b]);
      }
      if (host[b]) (void)hipHostFree(host[b]);
      if (dev[b]) (void)hipFree(dev[b]);
    }
    if (fd >= 0) close(fd);
  }
};

bool read_full(int fd, void* buf, int64_t bytes, int64_t offset) {
  char* p = static_cast<char*>(buf);
  while (bytes > 0) {
    const ssize_t r = pread(fd, p, (size_t)bytes, (off_t)offset);
    if (r <= 0) return false;
    p += r;
    bytes -= r;
    offset += r;
  }
  return true;
}

}  // namespace

extern "C" {

int32_t lufqhe_file_chunk(int64_t count, int32_t n, int32_t rank, int32_t nodes, int64_t* first,
                          int64_t* length) {
  if (count < 0 || n < 1 || nodes < 1 || rank < 0 || rank >= nodes || !first || !length)
    return LUFQHE_ERROR_INVALID_VALUE;
  const int64_t total = 2 * count * (int64_t)n * n;
  // the reference's float arithmetic, on a 64-bit product
  const int64_t start = (int64_t)((float)(rank * total) / (float)nodes);
  const int64_t end = (int64_t)((float)((rank + 1) * total) / (float)nodes) - 1;
  *first = start;
  *length = std::max<int64_t>(0, end - start + 1);
  return LUFQHE_SUCCESS;
}

int32_t lufqhe_read_matrices(const char* path, int32_t format, int64_t count, int32_t n,
                             int32_t rank, int32_t nodes, void* d_out, int64_t capacity,
                             int64_t* read_out, hipStream_t stream) {
  if (!path || !d_out || capacity < 0 ||
      (format != LUFQHE_FILE_COMPLEX64 && format != LUFQHE_FILE_COMPLEX128))
    return LUFQHE_ERROR_INVALID_VALUE;
  int64_t first = 0, length = 0;
  int32_t st = lufqhe_file_chunk(count, n, rank, nodes, &first, &length);
  if (st != LUFQHE_SUCCESS) return st;
  const int64_t esize = format == LUFQHE_FILE_COMPLEX64 ? 4 : 8;

  Staging s;
  s.fd = open(path, O_RDONLY);
  if (s.fd < 0) return LUFQHE_ERROR_IO;
  struct stat sb;
  if (fstat(s.fd, &sb) != 0) return LUFQHE_ERROR_IO;
  // a run past the end of the file (the float bounds can overshoot) is cut
  const int64_t avail = std::max<int64_t>(0, (int64_t)sb.st_size / esize - first);
  length = std::min(length, avail);
  if (length > capacity) return LUFQHE_ERROR_INVALID_VALUE;
  if (read_out) *read_out = length;
  if (length == 0) return LUFQHE_SUCCESS;

  const int64_t per = kStageBytes / esize;  // scalars per block
  hipError_t err = hipSuccess;
  for (int b = 0; b < 2 && err == hipSuccess; ++b) {
    err = hipHostMalloc(&s.host[b], (size_t)kStageBytes, 0);
    if (err == hipSuccess && format == LUFQHE_FILE_COMPLEX64)
      err = hipMalloc(&s.dev[b], (size_t)kStageBytes);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&s.done[b], hipEventDisableTiming);
  }
  if (err != hipSuccess) return hip_fail(err);

  double* out = static_cast<double*>(d_out);
  int64_t blk = 0;
  for (int64_t off = 0; off < length && err == hipSuccess; off += per, ++blk) {
    const int b = (int)(blk & 1);
    const int64_t m = std::min(per, length - off);
    err = hipEventSynchronize(s.done[b]);  // the buffer's previous block has left
    if (err != hipSuccess) break;
    if (!read_full(s.fd, s.host[b], m * esize, (first + off) * esize)) return LUFQHE_ERROR_IO;
    if (format == LUFQHE_FILE_COMPLEX64) {
      err = hipMemcpyAsync(s.dev[b], s.host[b], (size_t)(m * esize), hipMemcpyHostToDevice, stream);
      if (err == hipSuccess) {
        hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream,
                           s.dev[b], m, out + off);
        err = hipGetLastError();
      }
    } else {
      err = hipMemcpyAsync(out + off, s.host[b], (size_t)(m * esize), hipMemcpyHostToDevice, stream);
    }
    if (err == hipSuccess) err = hipEventRecord(s.done[b], stream);
  }
  if (err == hipSuccess) err = hipStreamSynchronize(stream);
  return err == hipSuccess ? LUFQHE_SUCCESS : hip_fail(err);
}

}  // extern "C"
