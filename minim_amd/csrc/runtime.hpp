// runtime.hpp -- device memory, streams and reduction scratch for the host
// side of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <utility>
#include <vector>

#include "common.hpp"

namespace minim_amd {

// Owning device allocation (hipMalloc/hipFree), move-only.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes);
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), bytes_(o.bytes_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(bytes_, o.bytes_);
    return *this;
  }
  void* get() const { return p_; }
  size_t bytes() const { return bytes_; }
  void reset(size_t bytes);  // reallocate (contents undefined)

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

// A device vector of n elements stored in precision prec (F64 or F32).
struct DVec {
  DeviceBuffer buf;
  size_t n = 0;
  int prec = F64;
  DVec() = default;
  DVec(size_t n_, int prec_) : buf(n_ * prec_bytes(prec_)), n(n_), prec(prec_) {}
  void* data() const { return buf.get(); }
};

// Reduction scratch plus a pinned host landing area, one per host thread and
// device.  Launches that reduce from one host thread share it (they are
// ordered on that thread's streams); other host threads have their own.
class Workspace {
 public:
  static Workspace& current();  // for the current HIP device
  ~Workspace();
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  RedScratch rs() const { return rs_; }
  // Copy result[0..k) to host (synchronising stream s) into out.
  void fetch(hipStream_t s, int k, double* out);
  double* device_result() const { return rs_.result; }

 private:
  Workspace();
  RedScratch rs_{};
  double* host_ = nullptr;
  DeviceBuffer mem_;
};

// Host <-> device transfer staging, one per host thread and device: two
// pinned host chunks (double-buffered: the host fills or drains one while the
// copy engine moves the other) and device buffers that only grow, reused by
// every transfer and host-boundary evaluation instead of a fresh allocation.
class Staging {
 public:
  static constexpr size_t kChunk = size_t(16) << 20;  // bytes per pinned chunk
  static Staging& current();  // for the current HIP device
  ~Staging();
  Staging(const Staging&) = delete;
  Staging& operator=(const Staging&) = delete;
  void* pinned(int i) const { return pin_[i]; }
  hipEvent_t event(int i) const { return ev_[i]; }
  // Device buffer slot k (0..3) of at least bytes (contents undefined).
  void* device(int k, size_t bytes);

 private:
  Staging();
  void* pin_[2] = {nullptr, nullptr};
  hipEvent_t ev_[2] = {nullptr, nullptr};
  DeviceBuffer dev_[4];
};

// The allocation-free entry point's reduction scratch on the current device
// (a file-scope device array of noalloc.hip): *nPartials partial sums, then
// kMaxDots result slots.
double* noalloc_scratch(size_t* nPartials);

// Host <-> device copies of FP64 host data into a vector of any precision.
void upload(const double* host, size_t n, DVec& dst, hipStream_t s);
void download(const DVec& src, double* host, hipStream_t s);
void upload_raw(const double* host, size_t n, int prec, void* dst, hipStream_t s);
void download_raw(const void* src, int prec, size_t n, double* host, hipStream_t s);
void device_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

}  // namespace minim_amd
