// runtime.cpp -- device buffers, per-device reduction workspace, transfers.
#include "runtime.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

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
                                                          double* __restrict__ result, int dupK, double* dupDst) {
  __shared__ double lds[kBlock / kWave];
  const int k = blockIdx.x;
  const double* p = partials + (size_t)k * nb;
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += kBlock) v[0] += p[b];
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) {
    result[k] = v[0];
    if (k == dupK) *dupDst = v[0];
  }
}
}  // namespace

void launch_finalize(RedScratch rs, int K, int nb, hipStream_t s, int dupK, double* dupDst) {
  if (K < 1 || K > kMaxDots || nb < 1 || nb > kMaxRedBlocks)
    throw std::invalid_argument("reduction shape out of range");
  if (!dupDst) dupK = -1;
  finalize_kernel<<<K, kBlock, 0, s>>>(rs.partials, nb, rs.result, dupK, dupDst);
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

Workspace::~Workspace() {
  if (host_) (void)hipHostFree(host_);
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

namespace {
// Persistent host worker threads for the conversions of the transfer staging
// (created once, never joined: they sleep between jobs and end with the
// process).  One job at a time; the calling thread takes part.
class HostPool {
 public:
  static HostPool& get() {
    static HostPool* p = new HostPool();
    return *p;
  }
  int size() const { return (int)th_.size() + 1; }
  // fn(i) for i in [0, parts), spread over the workers and the caller.
  void run(int parts, const std::function<void(int)>& fn) {
    if (parts <= 1 || th_.empty()) {
      for (int i = 0; i < parts; i++) fn(i);
      return;
    }
    std::lock_guard<std::mutex> one(run_mu_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      next_ = 0;
      total_ = parts;
      remaining_ = parts;
      gen_++;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return remaining_ == 0; });
    job_ = nullptr;
  }

 private:
  HostPool() {
    unsigned hc = std::thread::hardware_concurrency();
    int n = (int)std::min(16u, std::max(1u, hc)) - 1;
    for (int i = 0; i < n; i++) {
      th_.emplace_back([this] { loop(); });
      th_.back().detach();
    }
  }
  void work() {
    std::unique_lock<std::mutex> lk(mu_);
    while (job_ && next_ < total_) {
      const int i = next_++;
      const std::function<void(int)>* f = job_;
      lk.unlock();
      (*f)(i);
      lk.lock();
      if (--remaining_ == 0) done_.notify_all();
    }
  }
  void loop() {
    unsigned seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
      }
      work();
    }
  }
  std::vector<std::thread> th_;
  std::mutex run_mu_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int next_ = 0, total_ = 0, remaining_ = 0;
  unsigned gen_ = 0;
};

// dst[0..k) = src[0..k) converted between FP64 host data and device
// precision prec (FP32: round to nearest, as the device conversion does;
// FP32 to FP64 is exact), in parallel parts of at least 256 KB.
void host_convert(void* dst, int dprec, const void* src, int sprec, size_t k) {
  const size_t minPart = (size_t(256) << 10) / sizeof(double);
  const int parts = (int)std::min<size_t>((size_t)HostPool::get().size(), std::max<size_t>(1, k / minPart));
  const size_t per = (k + parts - 1) / parts;
  HostPool::get().run(parts, [&](int p) {
    const size_t a = (size_t)p * per, b = std::min(k, a + per);
    if (a >= b) return;
    if (dprec == sprec) {
      std::memcpy(static_cast<char*>(dst) + a * prec_bytes(dprec), static_cast<const char*>(src) + a * prec_bytes(sprec),
                  (b - a) * prec_bytes(dprec));
    } else if (dprec == F32) {
      const double* s = static_cast<const double*>(src);
      float* d = static_cast<float*>(dst);
      for (size_t i = a; i < b; i++) d[i] = static_cast<float>(s[i]);
    } else {
      const float* s = static_cast<const float*>(src);
      double* d = static_cast<double*>(dst);
      for (size_t i = a; i < b; i++) d[i] = static_cast<double>(s[i]);
    }
  });
}
}  // namespace

Staging::Staging() {
  for (int i = 0; i < 2; i++) {
    MINIM_HIP_CHECK(hipHostMalloc(&pin_[i], kChunk, hipHostMallocDefault));
    MINIM_HIP_CHECK(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
  }
}

Staging::~Staging() {
  for (int i = 0; i < 2; i++) {
    if (ev_[i]) (void)hipEventDestroy(ev_[i]);
    if (pin_[i]) (void)hipHostFree(pin_[i]);
  }
}

Staging& Staging::current() {
  thread_local std::map<int, std::unique_ptr<Staging>> per_device;
  int dev = 0;
  MINIM_HIP_CHECK(hipGetDevice(&dev));
  auto it = per_device.find(dev);
  if (it == per_device.end()) it = per_device.emplace(dev, std::unique_ptr<Staging>(new Staging())).first;
  return *it->second;
}

void* Staging::device(int k, size_t bytes) {
  if (k < 0 || k > 3) throw std::invalid_argument("staging slot out of range");
  if (dev_[k].bytes() < bytes) dev_[k].reset(bytes);
  return dev_[k].get();
}

// FP64 host data into a device vector of precision prec: the host converts
// each chunk into pinned memory (FP32 storage crosses PCIe at 4 bytes per
// value) while the copy engine moves the previous chunk.
void upload_raw(const double* host, size_t n, int prec, void* dst, hipStream_t s) {
  if (n == 0) return;
  Staging& st = Staging::current();
  const size_t eb = prec_bytes(prec), per = Staging::kChunk / eb;
  int i = 0;
  for (size_t off = 0; off < n; off += per, i ^= 1) {
    const size_t k = std::min(per, n - off);
    MINIM_HIP_CHECK(hipEventSynchronize(st.event(i)));  // the copy out of this chunk has ended
    host_convert(st.pinned(i), prec, host + off, F64, k);
    MINIM_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(dst) + off * eb, st.pinned(i), k * eb, hipMemcpyHostToDevice, s));
    MINIM_HIP_CHECK(hipEventRecord(st.event(i), s));
  }
  MINIM_HIP_CHECK(hipStreamSynchronize(s));
}

// A device vector of precision prec into FP64 host data: chunk j + 1 is in
// flight while the host converts chunk j out of pinned memory.
void download_raw(const void* src, int prec, size_t n, double* host, hipStream_t s) {
  if (n == 0) return;
  Staging& st = Staging::current();
  const size_t eb = prec_bytes(prec), per = Staging::kChunk / eb;
  const size_t nc = (n + per - 1) / per;
  auto issue = [&](size_t j) {
    const size_t off = j * per, k = std::min(per, n - off);
    MINIM_HIP_CHECK(hipMemcpyAsync(st.pinned(j & 1), static_cast<const char*>(src) + off * eb, k * eb,
                                   hipMemcpyDeviceToHost, s));
    MINIM_HIP_CHECK(hipEventRecord(st.event(j & 1), s));
  };
  issue(0);
  for (size_t j = 0; j < nc; j++) {
    if (j + 1 < nc) issue(j + 1);
    MINIM_HIP_CHECK(hipEventSynchronize(st.event(j & 1)));
    const size_t off = j * per, k = std::min(per, n - off);
    host_convert(host + off, F64, st.pinned(j & 1), prec, k);
  }
  MINIM_HIP_CHECK(hipStreamSynchronize(s));
}

void upload(const double* host, size_t n, DVec& dst, hipStream_t s) { upload_raw(host, n, dst.prec, dst.data(), s); }

void download(const DVec& src, double* host, hipStream_t s) { download_raw(src.data(), src.prec, src.n, host, s); }

void device_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  MINIM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

}  // namespace minim_amd
