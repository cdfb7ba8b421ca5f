// Owning wrappers over the device library's memory, stream and event calls.
// The host classes reach the GPU only through include/gat_hip.h; these helpers
// give its raw allocations an owner with a well-defined lifetime (the
// reference's arenas are bump-allocated and its leaf cache outlives them,
// Grid.cpp:715-727 / SystemTest.cpp:66-77).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>

#include "gat_hip.h"

// Host-side error reporting of libgat_host.so (thread-local).
void gath_set_error(const char* fmt, ...);
// Copies the device library's last error into the host error, with context.
int gath_fail_hip(int rc, const char* what);
// Device allocations DeviceBuffer::reserve has made in this process (each
// growth is one); the serving loop reports how many happened while it served.
int64_t gath_device_alloc_count();

#define GATH_TRY(expr, what)                       \
  do {                                             \
    int gath_rc_ = (expr);                         \
    if (gath_rc_ != GAT_OK) return gath_fail_hip(gath_rc_, what); \
  } while (0)

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      std::swap(p_, o.p_);
      std::swap(n_, o.n_);
    }
    return *this;
  }
  // Ensures at least `bytes` of storage; contents are not preserved on growth.
  int reserve(size_t bytes);
  void release();
  void* get() const { return p_; }
  size_t size() const { return n_; }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  ~PinnedBuffer() { release(); }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  int reserve(size_t bytes);
  void release();
  void* get() const { return p_; }
  size_t size() const { return n_; }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Pinned host memory for one batch's transfers, handed out in 256-byte
// aligned regions, so each copy is a DMA from or to pinned memory.  begin()
// waits for `stream` (no copy may still use the arena), sizes it for `bytes`
// and starts over; take() returns the next region, or null when the batch asks
// for more than begin() sized.
class PinnedArena {
 public:
  int begin(size_t bytes, void* stream);
  void* take(size_t bytes);
  template <class T>
  T* take_as(size_t count) {
    return static_cast<T*>(take(count * sizeof(T)));
  }

 private:
  PinnedBuffer buf_;
  size_t used_ = 0;
};

class Stream {
 public:
  Stream() = default;
  ~Stream() { release(); }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  int create();
  void release();
  void* get() const { return s_; }
  int sync() const;

 private:
  void* s_ = nullptr;
};

// Brackets device work on one stream with a pair of events.
// A second stream for work that overlaps the owner's stream: fork() makes it
// wait for what the owner's stream has queued so far, join() makes the owner's
// stream wait for what was queued on it since.
class SideStream {
 public:
  SideStream() = default;
  ~SideStream();
  SideStream(const SideStream&) = delete;
  SideStream& operator=(const SideStream&) = delete;
  int fork(void* owner);
  int join(void* owner);
  void* stream() const { return s_; }

 private:
  void* s_ = nullptr;
  void* fork_ = nullptr;
  void* join_ = nullptr;
};

class EventTimer {
 public:
  EventTimer() = default;
  ~EventTimer();
  EventTimer(const EventTimer&) = delete;
  EventTimer& operator=(const EventTimer&) = delete;
  int start(void* stream);
  int stop(void* stream);
  // Waits for the stop event; returns the elapsed milliseconds in *ms.
  int elapsed(float* ms);

 private:
  int ensure();
  void* a_ = nullptr;
  void* b_ = nullptr;
};
