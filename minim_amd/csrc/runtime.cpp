// runtime.cpp -- device buffers, per-device reduction workspace, transfers.
#include "runtime.hpp"

#include <map>
#include <memory>
#include <mutex>

#include "vec.hpp"

namespace minim_amd {

DeviceBuffer::DeviceBuffer(size_t bytes) { reset(bytes); }

DeviceBuffer::~DeviceBuffer() {
  if (p_) (void)hipFree(p_);
}

void DeviceBuffer::reset(size_t bytes) {
  if (p_) {
    MINIM_HIP_CHECK(hipFree(p_));
    p_ = nullptr;
    bytes_ = 0;
  }
  if (bytes) {
    MINIM_HIP_CHECK(hipMalloc(&p_, bytes));
    bytes_ = bytes;
  }
}

namespace {
// One block per reduced value: partials of value k are contiguous.
__global__ __launch_bounds__(kBlock) void finalize_kernel(const double* __restrict__ partials, int nb,
                                                          double* __restrict__ result) {
  __shared__ double lds[kBlock / kWave];
  const int k = blockIdx.x;
  const double* p = partials + (size_t)k * nb;
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += kBlock) v[0] += p[b];
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) result[k] = v[0];
}
}  // namespace

void launch_finalize(RedScratch rs, int K, int nb, hipStream_t s) {
  if (K < 1 || K > kMaxDots || nb < 1 || nb > kMaxRedBlocks)
    throw std::invalid_argument("reduction shape out of range");
  finalize_kernel<<<K, kBlock, 0, s>>>(rs.partials, nb, rs.result);
  MINIM_HIP_CHECK(hipGetLastError());
}

Workspace::Workspace() {
  const size_t partial_bytes = sizeof(double) * (size_t)kMaxRedBlocks * kMaxDots;
  const size_t result_bytes = sizeof(double) * kMaxDots;
  mem_.reset(partial_bytes + result_bytes + 256);
  char* base = static_cast<char*>(mem_.get());
  rs_.partials = reinterpret_cast<double*>(base);
  rs_.result = reinterpret_cast<double*>(base + partial_bytes);
  MINIM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_), result_bytes, hipHostMallocDefault));
}

// One workspace per host thread and device: reductions issued from different
// host threads (each on its own stream) never share scratch memory.
Workspace& Workspace::current() {
  thread_local std::map<int, std::unique_ptr<Workspace>> per_device;
  int dev = 0;
  MINIM_HIP_CHECK(hipGetDevice(&dev));
  auto it = per_device.find(dev);
  if (it == per_device.end()) {
    it = per_device.emplace(dev, std::unique_ptr<Workspace>(new Workspace())).first;
  }
  return *it->second;
}

void Workspace::fetch(hipStream_t s, int k, double* out) {
  MINIM_HIP_CHECK(hipMemcpyAsync(host_, rs_.result, sizeof(double) * k, hipMemcpyDeviceToHost, s));
  MINIM_HIP_CHECK(hipStreamSynchronize(s));
  for (int i = 0; i < k; i++) out[i] = host_[i];
}

void upload_raw(const double* host, size_t n, int prec, void* dst, hipStream_t s) {
  if (n == 0) return;
  if (prec == F64) {
    MINIM_HIP_CHECK(hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, s));
    MINIM_HIP_CHECK(hipStreamSynchronize(s));
  } else {
    DeviceBuffer tmp(n * sizeof(double));
    MINIM_HIP_CHECK(hipMemcpyAsync(tmp.get(), host, n * sizeof(double), hipMemcpyHostToDevice, s));
    launch_convert(F32, dst, F64, tmp.get(), n, s);
    MINIM_HIP_CHECK(hipStreamSynchronize(s));
  }
}

void download_raw(const void* src, int prec, size_t n, double* host, hipStream_t s) {
  if (n == 0) return;
  if (prec == F64) {
    MINIM_HIP_CHECK(hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, s));
    MINIM_HIP_CHECK(hipStreamSynchronize(s));
  } else {
    DeviceBuffer tmp(n * sizeof(double));
    launch_convert(F64, tmp.get(), F32, src, n, s);
    MINIM_HIP_CHECK(hipMemcpyAsync(host, tmp.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
    MINIM_HIP_CHECK(hipStreamSynchronize(s));
  }
}

void upload(const double* host, size_t n, DVec& dst, hipStream_t s) { upload_raw(host, n, dst.prec, dst.data(), s); }

void download(const DVec& src, double* host, hipStream_t s) { download_raw(src.data(), src.prec, src.n, host, s); }

void device_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  MINIM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

}  // namespace minim_amd
