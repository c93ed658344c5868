// Common definitions for the MI355X (gfx950) iSVD library: error state, HIP checks,
// device-buffer RAII and the per-call execution context.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace isvd {

using i64 = long long;

// ---------------------------------------------------------------------------------------------
// Error handling.  The reference prints and returns for bad selector strings / chars
// (src/lib/libisvd/util/arg.h:40-48, :70-78) and asserts for everything else
// (src/lib/libisvd/def.h:51-68).  Here every failure prints to stderr, records a message
// readable through isvd_GetLastError(), and unwinds the call with an exception that the
// C-ABI layer catches; ISVD_ASSERT_ABORT=1 restores the reference's abort().
// ---------------------------------------------------------------------------------------------
struct Error {
  std::string msg;
};

void set_last_error(const std::string &msg);
[[noreturn]] void fail(const std::string &msg);

#define ISVD_CHECK(cond, ...)                                                   \
  do {                                                                          \
    if (!(cond)) {                                                              \
      char buf_[512];                                                           \
      std::snprintf(buf_, sizeof(buf_), __VA_ARGS__);                           \
      ::isvd::fail(std::string(__func__) + ": " + buf_);                        \
    }                                                                           \
  } while (0)

#define HIP_CHECK(expr)                                                         \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      char buf_[512];                                                           \
      std::snprintf(buf_, sizeof(buf_), "%s failed: %s (%s) at %s:%d", #expr,   \
                    hipGetErrorName(e_), hipGetErrorString(e_), __FILE__,       \
                    __LINE__);                                                  \
      ::isvd::fail(buf_);                                                       \
    }                                                                           \
  } while (0)

// Launch check: catches invalid launch configurations right away.
#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

// ---------------------------------------------------------------------------------------------
// Execution context: one HIP stream per library call (default priority), stream-ordered
// allocations from the device's default memory pool.
// ---------------------------------------------------------------------------------------------
// A fixed device region handed out front to back (256-byte aligned pieces) in place of
// stream-ordered allocations, for entry points that must not allocate.  Nothing is returned
// to it during a call; the call that installed it resets it.
struct Arena {
  unsigned char *base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  void *take(size_t bytes);          // fails with an error when the region is exhausted
  size_t left() const { return cap - used; }
};

struct Ctx {
  hipStream_t stream = nullptr;
  // a second stream (default priority) for a product that runs beside work on `stream`; it
  // is joined to `stream` by the events before anything else reads what it wrote
  hipStream_t aux = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  int device = 0;
  int cus = 256;  // compute units of the device
  Arena *arena = nullptr;  // non-null while a no-allocation entry point runs
  // > 0 while a product runs beside a kernel that needs CUs of its own: the product splits K
  // into no more chunks than keep its workgroups (one per CU) within this count
  int max_wgs = 0;
};


Ctx &ctx();                // library context (created by isvd_Init_gpu or lazily)
void ensure_init();         // selects device 0, creates the stream

// Caps the workgroups of the enclosed products (Ctx::max_wgs) for the scope.
struct GridCap {
  int saved;
  explicit GridCap(int cap) : saved(ctx().max_wgs) { ctx().max_wgs = cap; }
  ~GridCap() { ctx().max_wgs = saved; }
};

// Device buffer owned by the caller's scope, allocated stream-ordered (or taken from the
// active arena, which needs no release).
template <typename T>
struct DBuf {
  T *p = nullptr;
  i64 n = 0;
  bool arena = false;
  DBuf() = default;
  explicit DBuf(i64 count) { alloc(count); }
  void alloc(i64 count) {
    release();
    n = count;
    if (count > 0) {
      if (ctx().arena) {
        p = static_cast<T *>(ctx().arena->take(sizeof(T) * (size_t)count));
        arena = true;
      } else {
        HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&p), sizeof(T) * (size_t)count,
                                 ctx().stream));
      }
    }
  }
  void release() {
    if (p) {
      if (!arena) (void)hipFreeAsync(p, ctx().stream);
      p = nullptr;
      arena = false;
    }
    n = 0;
  }
  ~DBuf() { release(); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), arena(o.arena) {
    o.p = nullptr;
    o.n = 0;
    o.arena = false;
  }
  T *get() const { return p; }
};

// Split-K partial sums: one device buffer per library context, grown on demand and reused by
// every product (all work is ordered on ctx().stream, so consecutive products never overlap).
// Under an arena the partials are taken from it instead.
void *splitk_workspace(size_t bytes);
// Bytes the split-K partials of one product may occupy (the arena's free space under an arena,
// unbounded otherwise).
size_t splitk_budget();

// Host <- device copy of a few values that returns as soon as they land: the copy goes to
// pinned host memory and the host spins on an event instead of a blocking stream wait.
void read_back(void *host, const void *dev, size_t bytes);
// Pinned slots (0..3) for values a kernel hands to the host while the host queues more work:
// pinned_slot gives the device address of a slot (a kernel may store up to 256 bytes there;
// they are seen by the host once the kernel is done), pinned_slot_record records the slot's
// event on the stream after that kernel, and read_back_wait spins until that event and copies
// the bytes out.  A slot holds one pending value at a time.
void *pinned_slot(int slot);
void pinned_slot_record(int slot);
hipEvent_t pinned_slot_event(int slot);  // the event pinned_slot_record records
void read_back_wait(int slot, void *host, size_t bytes);

// True when ptr is device (HBM) memory usable in place by kernels on the current device.
bool is_device_pointer(const void *ptr);

void sync();  // hipStreamSynchronize(ctx().stream) with error check

inline int ceil_div(i64 a, i64 b) { return (int)((a + b - 1) / b); }

}  // namespace isvd
