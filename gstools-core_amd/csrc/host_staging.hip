// Host-side staging of the synchronous entry points (gsc_host.h).
//
// A host array reaches the device in 16 MB chunks through two pinned
// buffers: the pool threads copy (or gather, for strided views) chunk q into
// buffer q % 2 while the DMA engine moves chunk q - 1; results come back the
// same way.  Pageable copies go through the runtime's own staging at about
// 20 GB/s; pinned chunks run at the PCIe rate and the host side of the copy
// is spread over the pool.
#include "gsc_host.h"

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace gsc {
namespace {

// Persistent workers.  One job runs at a time; the caller takes part, and a
// worker that wakes up after the caller closed the job takes no part in it.
class HostPool {
 public:
  explicit HostPool(int n) {
    for (int i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> l(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  int size() const { return static_cast<int>(th_.size()); }

  void run(int64_t n, int threads, const std::function<void(int64_t)>& fn) {
    std::lock_guard<std::mutex> rl(run_mu_);
    const int helpers = std::min<int64_t>(std::min(threads - 1, size()), n - 1);
    if (helpers <= 0) {
      for (int64_t t = 0; t < n; ++t) fn(t);
      return;
    }
    {
      std::lock_guard<std::mutex> l(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      slots_ = helpers;
      ++gen_;
    }
    cv_.notify_all();
    for (int64_t t; (t = next_.fetch_add(1)) < n;) fn(t);
    std::unique_lock<std::mutex> l(mu_);
    slots_ = 0;
    done_.wait(l, [this] { return running_ == 0; });
    fn_ = nullptr;
  }

 private:
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> l(mu_);
      cv_.wait(l, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      if (slots_ <= 0) continue;
      --slots_;
      ++running_;
      const std::function<void(int64_t)>* fn = fn_;
      const int64_t n = n_;
      l.unlock();
      for (int64_t t; (t = next_.fetch_add(1)) < n;) (*fn)(t);
      l.lock();
      if (--running_ == 0) done_.notify_all();
    }
  }

  std::vector<std::thread> th_;
  std::mutex run_mu_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int64_t)>* fn_ = nullptr;
  int64_t n_ = 0;
  std::atomic<int64_t> next_{0};
  int slots_ = 0, running_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Threads of the pool, the caller included: GSC_NUM_THREADS, else
// OMP_NUM_THREADS, else the CPUs this process may run on; at most 64.
int default_threads() {
  for (const char* var : {"GSC_NUM_THREADS", "OMP_NUM_THREADS"}) {
    const char* v = std::getenv(var);
    if (v && std::atoi(v) > 0) return std::min(64, std::atoi(v));
  }
  cpu_set_t set;
  int n = 0;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
  if (n <= 0) n = static_cast<int>(std::thread::hardware_concurrency());
  return std::min(64, std::max(1, n));
}

HostPool& pool() {
  static HostPool p(default_threads() - 1);
  return p;
}

int resolve_threads(int threads) {
  return threads > 0 ? threads : pool().size() + 1;
}

// Copies flat elements [f0, f1) of a (rows, cols) array between a strided
// host view and a dense buffer; dir 0: host -> dense, 1: dense -> host.
// Pieces of kPiece elements go to the pool.
constexpr int64_t kPiece = int64_t(1) << 17;
void copy_span(char* dense, char* host, int64_t f0, int64_t f1, int64_t cols, int64_t sr,
               int64_t sc, size_t elt, int dir, int threads) {
  const int64_t n = f1 - f0;
  const bool flat = sc == 1 && sr == cols;
  const int64_t pieces = (n + kPiece - 1) / kPiece;
  parallel_for(pieces, threads, [&](int64_t p) {
    const int64_t a = f0 + p * kPiece, b = std::min(f1, a + kPiece);
    char* d = dense + (a - f0) * elt;
    if (flat) {
      char* h = host + a * static_cast<int64_t>(elt);
      if (dir == 0) std::memcpy(d, h, static_cast<size_t>(b - a) * elt);
      else std::memcpy(h, d, static_cast<size_t>(b - a) * elt);
      return;
    }
    int64_t r = a / cols, c = a - r * cols;
    for (int64_t f = a; f < b; ++f) {
      char* h = host + (r * sr + c * sc) * static_cast<int64_t>(elt);
      if (elt == 8) {
        if (dir == 0) std::memcpy(d, h, 8);
        else std::memcpy(h, d, 8);
      } else {
        if (dir == 0) std::memcpy(d, h, 4);
        else std::memcpy(h, d, 4);
      }
      d += elt;
      if (++c == cols) {
        c = 0;
        ++r;
      }
    }
  });
}

template <typename T>
__global__ void __launch_bounds__(256)
k_transpose_cn(const T* __restrict__ in, T* __restrict__ out, int ncomp, int64_t N) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (j >= N) return;
  for (int c = 0; c < ncomp; ++c) out[j * ncomp + c] = in[c * N + j];
}

}  // namespace

void parallel_for(int64_t n_tasks, int threads, const std::function<void(int64_t)>& fn) {
  if (n_tasks <= 0) return;
  pool().run(n_tasks, resolve_threads(threads), fn);
}

bool staging_ready(Staging& s) {
  if (s.pin[0]) return true;
  for (int i = 0; i < 2; ++i) {
    if (hipHostMalloc(&s.pin[i], kStageChunk, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s.ev[i], hipEventDisableTiming) != hipSuccess) {
      staging_release(s);
      return false;
    }
  }
  return true;
}

void staging_release(Staging& s) {
  for (int i = 0; i < 2; ++i) {
    if (s.pin[i]) (void)hipHostFree(s.pin[i]);
    if (s.ev[i]) (void)hipEventDestroy(s.ev[i]);
    s.pin[i] = nullptr;
    s.ev[i] = nullptr;
  }
}

hipError_t stage_upload(Staging& s, hipStream_t stream, double* d_dst, const double* src,
                        int64_t rows, int64_t cols, int64_t sr, int64_t sc, int threads) {
  const int64_t total = rows * cols;
  const int64_t per = static_cast<int64_t>(kStageChunk / sizeof(double));
  hipError_t e = hipSuccess;
  for (int64_t q = 0, f0 = 0; f0 < total; ++q, f0 += per) {
    const int64_t f1 = std::min(total, f0 + per);
    const int b = static_cast<int>(q & 1);
    // the buffer's previous DMA (chunk q - 2, or an earlier call's) has run
    if ((e = hipEventSynchronize(s.ev[b])) != hipSuccess) return e;
    copy_span(static_cast<char*>(s.pin[b]),
              const_cast<char*>(reinterpret_cast<const char*>(src)), f0, f1, cols, sr, sc,
              sizeof(double), 0, threads);
    if ((e = hipMemcpyAsync(d_dst + f0, s.pin[b], sizeof(double) * (f1 - f0),
                            hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (e = hipEventRecord(s.ev[b], stream)) != hipSuccess)
      return e;
  }
  return e;
}

hipError_t stage_download(Staging& s, hipStream_t stream, const void* d_src, void* dst,
                          int64_t rows, int64_t cols, int64_t sr, int64_t sc, size_t elt,
                          int threads) {
  const int64_t total = rows * cols;
  const int64_t per = static_cast<int64_t>(kStageChunk / elt);
  const int64_t nq = (total + per - 1) / per;
  const char* src = static_cast<const char*>(d_src);
  hipError_t e = hipSuccess;
  auto issue = [&](int64_t q) {
    const int64_t f0 = q * per, f1 = std::min(total, f0 + per);
    const int b = static_cast<int>(q & 1);
    hipError_t r = hipMemcpyAsync(s.pin[b], src + f0 * static_cast<int64_t>(elt),
                                  elt * static_cast<size_t>(f1 - f0), hipMemcpyDeviceToHost,
                                  stream);
    return r == hipSuccess ? hipEventRecord(s.ev[b], stream) : r;
  };
  for (int64_t q = 0; q < std::min<int64_t>(nq, 2); ++q)
    if ((e = issue(q)) != hipSuccess) return e;
  for (int64_t q = 0; q < nq; ++q) {
    const int b = static_cast<int>(q & 1);
    if ((e = hipEventSynchronize(s.ev[b])) != hipSuccess) return e;
    const int64_t f0 = q * per, f1 = std::min(total, f0 + per);
    copy_span(static_cast<char*>(s.pin[b]), static_cast<char*>(dst), f0, f1, cols, sr, sc, elt,
              1, threads);
    if (q + 2 < nq && (e = issue(q + 2)) != hipSuccess) return e;
  }
  return e;
}

hipError_t launch_transpose_cn(const void* in, void* out, int ncomp, int64_t N, size_t elt,
                               hipStream_t stream) {
  if (N <= 0) return hipSuccess;
  const unsigned g = static_cast<unsigned>((N + 255) / 256);
  if (elt == 8)
    k_transpose_cn<double><<<g, 256, 0, stream>>>(static_cast<const double*>(in),
                                                  static_cast<double*>(out), ncomp, N);
  else
    k_transpose_cn<float><<<g, 256, 0, stream>>>(static_cast<const float*>(in),
                                                 static_cast<float*>(out), ncomp, N);
  return hipGetLastError();
}

}  // namespace gsc
