// Host-memory copies of returned rows spread over a few library-owned
// threads: moving rows out of pinned staging into freshly allocated result
// buffers is bound by first-touch page faults and one core's copy rate.
#pragma once

#include <cstddef>

namespace ticit {

struct HostCopy {
  void* dst;
  const void* src;
  size_t bytes;
};

// Copies every piece; returns when all are done. Small totals are copied on
// the calling thread.
void host_copy(const HostCopy* pieces, size_t n);

}  // namespace ticit
