// Allocation-free variant of the batched no-pivot tile LU, for callers that
// cannot allocate device memory (code under stream capture, tight serving
// loops).  The dataflow tile LU needs, beyond the caller's tiles, the full
// inverses of both factors (its workgroups exchange them), the per-block flags
// and one info word per tile.  Here these live in one file-scope device array
// per GPU instead of per-call allocations.  Tiles are factorized one after
// another on the stream, so the inverse region is reused by every tile of a
// call; flags are told apart by a per-tile epoch.
//
// The array keeps the last call's contents: every call clears the flags and
// info words with a fill kernel first.  Two calls that overlap on one GPU
// (two streams, two host threads) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/frank_amd.h"
#include "common.hpp"
#include "kernels.hpp"

namespace frk {

namespace {

constexpr int kNoallocMaxN = 512;
constexpr int kNoallocMaxBatch = 1024;
constexpr int kNoallocFlags = 2 * 64 + 64;  // F and I of at most 8 x 8 blocks, error word, padding
constexpr size_t kInvBytes = sizeof(double) * 2 * kNoallocMaxN * kNoallocMaxN;
constexpr size_t kScratchBytes = kInvBytes + sizeof(int32_t) * (kNoallocFlags + kNoallocMaxBatch);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per
// device with that device current; the host-side name is only a stand-in).
unsigned char* scratch_ptr() {
  int dev = 0;
  FRK_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("no-alloc getrf: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("no-alloc getrf: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_ptr[dev] = p;
  }
  return static_cast<unsigned char*>(g_ptr[dev]);
}

__global__ void clear_words_kernel(int32_t* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

// Factorizes `batch` n x n tiles in place, writing the first zero pivot of
// each tile (1-based row, 0 if none) to info_host.  Status as frk_dgetrf_nopiv_batched.
int getrf_nopiv_noalloc(hipStream_t s, double* const* tiles, int batch, int n, int64_t ld, int32_t* info_host,
                        std::string& msg) {
  if (batch > kNoallocMaxBatch || n > kNoallocMaxN) {
    msg = "no-alloc getrf: batch " + std::to_string(batch) + " x n " + std::to_string(n) +
          " exceeds the fixed scratch (batch <= " + std::to_string(kNoallocMaxBatch) +
          ", n <= " + std::to_string(kNoallocMaxN) + ")";
    return FRK_ERR_CAPACITY;
  }
  unsigned char* base = scratch_ptr();
  double* linv = reinterpret_cast<double*>(base);
  double* uinv = linv + (size_t)kNoallocMaxN * kNoallocMaxN;
  int32_t* words = reinterpret_cast<int32_t*>(base + kInvBytes);
  int32_t* flags = words;
  int32_t* err = words + 2 * 64;
  int32_t* info = words + kNoallocFlags;
  const int nwords = kNoallocFlags + batch;
  hipLaunchKernelGGL(clear_words_kernel, dim3((nwords + 255) / 256), dim3(256), 0, s, words, nwords);
  FRK_HIP_CHECK(hipGetLastError());
  for (int b = 0; b < batch; ++b) {
    TileLuItem it{};
    it.a = tiles[b];
    it.linv = linv;
    it.uinv = uinv;
    it.flags = flags;
    it.err = err;
    it.info = info + b;
    it.ld = ld;
    it.ldi = n;
    it.n = n;
    it.epoch = b + 1;  // flags were cleared above: tile b's blocks are final once they read b + 1
    it.spin_limit = int64_t(1) << 22;
    launch_tile_lu(it, s);
  }
  int32_t h[kNoallocMaxBatch + 1];
  FRK_HIP_CHECK(hipMemcpyAsync(h, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  FRK_HIP_CHECK(hipMemcpyAsync(h + 1, info, sizeof(int32_t) * batch, hipMemcpyDeviceToHost, s));
  FRK_HIP_CHECK(hipStreamSynchronize(s));
  if (h[0] != 0) {
    msg = "no-alloc getrf: a block wait of tile " + std::to_string(h[0] - 1) + " timed out";
    return FRK_ERR_HIP;
  }
  int status = FRK_OK;
  for (int b = 0; b < batch; ++b) {
    if (info_host) info_host[b] = h[b + 1];
    if (h[b + 1] != 0 && status == FRK_OK) {
      status = 1;
      msg = "zero pivot in dense getrf";
    }
  }
  return status;
}

}  // namespace frk
