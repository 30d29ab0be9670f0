#include "host_copy.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace ticit {
namespace {

constexpr size_t kStripe = size_t(1) << 20;        // bytes per work item
constexpr size_t kParallelMin = size_t(4) << 20;   // below this: the caller copies alone
constexpr unsigned kMaxWorkers = 4;

// A fixed set of detached workers that take 1 MiB stripes of the current
// job from a shared cursor; the caller takes stripes too. Created on first
// use and never destroyed (the workers sleep on the condition variable when
// idle, so the process can exit at any time).
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* pool = new CopyPool();
    return *pool;
  }

  void run(const HostCopy* pieces, size_t n) {
    std::lock_guard<std::mutex> serial(run_mu_);
    std::vector<Item> items;
    for (size_t i = 0; i < n; ++i)
      for (size_t off = 0; off < pieces[i].bytes; off += kStripe)
        items.push_back(Item{static_cast<char*>(pieces[i].dst) + off, static_cast<const char*>(pieces[i].src) + off,
                             std::min(kStripe, pieces[i].bytes - off)});
    {
      std::lock_guard<std::mutex> lk(mu_);
      items_ = &items;
      next_ = 0;
      done_ = 0;
      ++generation_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    finished_.wait(lk, [&] { return done_ == items.size(); });
    items_ = nullptr;
  }

 private:
  struct Item {
    char* dst;
    const char* src;
    size_t bytes;
  };

  CopyPool() {
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    unsigned workers = std::min(kMaxWorkers, hw > 1 ? hw - 1 : 0u);
    for (unsigned i = 0; i < workers; ++i) std::thread([this] { loop(); }).detach();
  }

  // Takes stripes until none are left; counts each finished one.
  void work() {
    for (;;) {
      Item it;
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!items_ || next_ >= items_->size()) return;
        it = (*items_)[next_++];
      }
      std::memcpy(it.dst, it.src, it.bytes);
      std::lock_guard<std::mutex> lk(mu_);
      if (++done_ == items_->size()) finished_.notify_all();
    }
  }

  void loop() {
    unsigned long long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return generation_ != seen; });
        seen = generation_;
      }
      work();
    }
  }

  std::mutex run_mu_;  // one job at a time
  std::mutex mu_;
  std::condition_variable cv_, finished_;
  std::vector<Item>* items_ = nullptr;
  size_t next_ = 0, done_ = 0;
  unsigned long long generation_ = 0;
};

}  // namespace

void host_copy(const HostCopy* pieces, size_t n) {
  size_t total = 0;
  for (size_t i = 0; i < n; ++i) total += pieces[i].bytes;
  if (total < kParallelMin) {
    for (size_t i = 0; i < n; ++i)
      if (pieces[i].bytes) std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
    return;
  }
  CopyPool::get().run(pieces, n);
}

}  // namespace ticit
