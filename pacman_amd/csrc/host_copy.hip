// Copies between pageable host arrays and device memory through pinned
// staging buffers.
//
// The drop-in entry points take host arrays. hipMemcpyAsync from pageable
// memory goes through the runtime's own bounce buffers, one chunk at a
// time (about 25 GB/s host-to-device and 18 GB/s device-to-host measured on
// MI355X), against about 57 GB/s from pinned memory. Large copies are
// therefore split over kLanes host threads: each lane moves its contiguous
// range in kChunk pieces through two pinned buffers of its own, filling (or
// draining) one on the CPU while the DMA engine moves the other, on the
// lane's own stream. The pinned buffers (kLanes x 2 x kChunk bytes per
// device) are allocated with hipHostMalloc on first use and kept for the
// process; small copies go straight through hipMemcpyAsync.
//
// Ordering: with `afterStream` the calling thread first waits for the work
// already queued on s (a device-to-host copy then reads what s produced);
// a host wait rather than a wait of the lanes' streams on an event of s:
// such cross-queue waits stalled the copies' enqueue by milliseconds now
// and then, and the copy returns only once done anyway. Without it the copy
// is ordered after nothing, so it can run beside work on s that does not
// touch dst. Both functions return once the data has arrived (on the
// device, or in the host array). Staged copies are serialised per process.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "rbf_internal.hpp"

namespace pacman_amd {

namespace {

constexpr size_t kChunk = (size_t)4 << 20;
constexpr int kLanes = 8;
constexpr size_t kDirectMax = (size_t)1 << 20;

struct Lane {
  void *buf[2] = {nullptr, nullptr};
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

struct Staging {
  std::vector<Lane> lanes;
};

std::mutex g_mu;
std::map<int, Staging> g_staging;

Staging &staging_for(int dev) {
  Staging &st = g_staging[dev];
  if (st.lanes.empty()) {
    std::vector<Lane> lanes(kLanes);
    for (Lane &l : lanes) {
      for (int b = 0; b < 2; ++b) {
        PACMAN_HIP_CHECK(hipHostMalloc(&l.buf[b], kChunk, hipHostMallocDefault));
        PACMAN_HIP_CHECK(
            hipEventCreateWithFlags(&l.ev[b], hipEventDisableTiming));
      }
      PACMAN_HIP_CHECK(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
    }
    st.lanes = std::move(lanes);
  }
  return st;
}

void lane_to_device(Lane &l, char *dst, const char *src, size_t bytes) {
  int c = 0;
  for (size_t off = 0; off < bytes; off += kChunk, ++c) {
    const size_t len = std::min(kChunk, bytes - off);
    const int b = c & 1;
    if (c >= 2)  // the DMA of chunk c - 2 out of this buffer is done
      PACMAN_HIP_CHECK(hipEventSynchronize(l.ev[b]));
    std::memcpy(l.buf[b], src + off, len);
    PACMAN_HIP_CHECK(hipMemcpyAsync(dst + off, l.buf[b], len,
                                    hipMemcpyHostToDevice, l.stream));
    PACMAN_HIP_CHECK(hipEventRecord(l.ev[b], l.stream));
  }
  PACMAN_HIP_CHECK(hipStreamSynchronize(l.stream));
}

void lane_to_host(Lane &l, char *dst, const char *src, size_t bytes) {
  const size_t n = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t c) {
    const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
    PACMAN_HIP_CHECK(hipMemcpyAsync(l.buf[c & 1], src + off, len,
                                    hipMemcpyDeviceToHost, l.stream));
    PACMAN_HIP_CHECK(hipEventRecord(l.ev[c & 1], l.stream));
  };
  for (size_t c = 0; c < std::min<size_t>(n, 2); ++c)
    issue(c);
  for (size_t c = 0; c < n; ++c) {
    const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
    PACMAN_HIP_CHECK(hipEventSynchronize(l.ev[c & 1]));
    std::memcpy(dst + off, l.buf[c & 1], len);
    if (c + 2 < n)
      issue(c + 2);
  }
  PACMAN_HIP_CHECK(hipStreamSynchronize(l.stream));
}

void staged_copy(bool toDevice, char *dst, const char *src, size_t bytes,
                 hipStream_t s, bool afterStream) {
  int dev = 0;
  PACMAN_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  Staging &st = staging_for(dev);
  // the work queued on s first (a host wait: the copy returns only once
  // done anyway, and the lanes' streams then depend on no other queue)
  if (afterStream)
    PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  // contiguous ranges, 64 KB aligned, at least one chunk each
  const size_t nChunks = (bytes + kChunk - 1) / kChunk;
  const int nl = (int)std::min<size_t>(kLanes, nChunks);
  const size_t per = ((bytes + nl - 1) / nl + 65535) & ~(size_t)65535;
  std::vector<std::exception_ptr> err(nl);
  std::vector<std::thread> th;
  auto run = [&, toDevice](int i) {
    const size_t a = std::min(bytes, (size_t)i * per);
    const size_t e = std::min(bytes, a + per);
    try {
      PACMAN_HIP_CHECK(hipSetDevice(dev));  // the device is per thread
      if (e > a) {
        if (toDevice)
          lane_to_device(st.lanes[i], dst + a, src + a, e - a);
        else
          lane_to_host(st.lanes[i], dst + a, src + a, e - a);
      }
    } catch (...) {
      err[i] = std::current_exception();
    }
  };
  for (int i = 1; i < nl; ++i)
    th.emplace_back(run, i);
  run(0);
  for (auto &t : th)
    t.join();
  for (auto &e : err)
    if (e)
      std::rethrow_exception(e);
}

} // namespace

void copy_to_device(void *dst, const void *src, size_t bytes, hipStream_t s,
                    bool afterStream) {
  if (bytes == 0)
    return;
  TraceRange tr("pacman::copy_to_device");
  if (bytes <= kDirectMax) {
    if (afterStream) {
      PACMAN_HIP_CHECK(
          hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    } else {
      PACMAN_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    }
    return;
  }
  staged_copy(true, static_cast<char *>(dst), static_cast<const char *>(src),
              bytes, s, afterStream);
}

void copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes == 0)
    return;
  TraceRange tr("pacman::copy_to_host");
  if (bytes <= kDirectMax) {
    PACMAN_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    PACMAN_HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  staged_copy(false, static_cast<char *>(dst), static_cast<const char *>(src),
              bytes, s, true);
}

} // namespace pacman_amd
