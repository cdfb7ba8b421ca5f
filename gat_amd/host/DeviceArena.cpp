#include "DeviceArena.h"

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>

static thread_local std::string g_host_error;

void gath_set_error(const char* fmt, ...) {
  char buf[768];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_host_error = buf;
}

int gath_fail_hip(int rc, const char* what) {
  gath_set_error("%s: %s", what, gat_last_error());
  return rc;
}

extern "C" __attribute__((visibility("default"))) const char* gath_last_error(void) { return g_host_error.c_str(); }

static std::atomic<int64_t> g_device_allocs{0};

int64_t gath_device_alloc_count() { return g_device_allocs.load(std::memory_order_relaxed); }

int DeviceBuffer::reserve(size_t bytes) {
  if (bytes <= n_) return GAT_OK;
  release();
  void* p = nullptr;
  g_device_allocs.fetch_add(1, std::memory_order_relaxed);
  GATH_TRY(gat_malloc(&p, bytes), "device allocation");
  p_ = p;
  n_ = bytes;
  return GAT_OK;
}

void DeviceBuffer::release() {
  if (p_) gat_free(p_);
  p_ = nullptr;
  n_ = 0;
}

int PinnedBuffer::reserve(size_t bytes) {
  if (bytes <= n_) return GAT_OK;
  release();
  void* p = nullptr;
  GATH_TRY(gat_host_alloc(&p, bytes), "pinned allocation");
  p_ = p;
  n_ = bytes;
  return GAT_OK;
}

void PinnedBuffer::release() {
  if (p_) gat_host_free(p_);
  p_ = nullptr;
  n_ = 0;
}

int PinnedArena::begin(size_t bytes, void* stream) {
  if (stream) GATH_TRY(gat_stream_sync(stream), "pinned arena");
  used_ = 0;
  return buf_.reserve(std::max<size_t>(bytes, 256));
}

void* PinnedArena::take(size_t bytes) {
  const size_t at = (used_ + 255) & ~size_t(255);
  if (at + bytes > buf_.size()) return nullptr;
  used_ = at + bytes;
  return buf_.as<char>() + at;
}

int Stream::create() {
  if (s_) return GAT_OK;
  GATH_TRY(gat_stream_create(&s_), "stream creation");
  return GAT_OK;
}

void Stream::release() {
  if (s_) gat_stream_destroy(s_);
  s_ = nullptr;
}

int Stream::sync() const {
  GATH_TRY(gat_stream_sync(s_), "stream synchronisation");
  return GAT_OK;
}

SideStream::~SideStream() {
  if (fork_) gat_event_destroy(fork_);
  if (join_) gat_event_destroy(join_);
  if (s_) gat_stream_destroy(s_);
}

int SideStream::fork(void* owner) {
  if (!s_) GATH_TRY(gat_stream_create(&s_), "side stream creation");
  if (!fork_) GATH_TRY(gat_event_create(&fork_), "event creation");
  if (!join_) GATH_TRY(gat_event_create(&join_), "event creation");
  GATH_TRY(gat_event_record(fork_, owner), "event record");
  GATH_TRY(gat_stream_wait_event(s_, fork_), "side stream wait");
  return GAT_OK;
}

int SideStream::join(void* owner) {
  GATH_TRY(gat_event_record(join_, s_), "event record");
  GATH_TRY(gat_stream_wait_event(owner, join_), "stream wait");
  return GAT_OK;
}

EventTimer::~EventTimer() {
  if (a_) gat_event_destroy(a_);
  if (b_) gat_event_destroy(b_);
}

int EventTimer::ensure() {
  if (!a_) GATH_TRY(gat_event_create(&a_), "event creation");
  if (!b_) GATH_TRY(gat_event_create(&b_), "event creation");
  return GAT_OK;
}

int EventTimer::start(void* stream) {
  if (int rc = ensure()) return rc;
  GATH_TRY(gat_event_record(a_, stream), "event record");
  return GAT_OK;
}

int EventTimer::stop(void* stream) {
  if (int rc = ensure()) return rc;
  GATH_TRY(gat_event_record(b_, stream), "event record");
  return GAT_OK;
}

int EventTimer::elapsed(float* ms) {
  GATH_TRY(gat_event_sync(b_), "event synchronisation");
  GATH_TRY(gat_event_elapsed_ms(ms, a_, b_), "event elapsed time");
  return GAT_OK;
}
