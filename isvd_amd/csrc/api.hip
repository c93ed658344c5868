// C ABI of libisvd_mi355x.so: the reference's public entry points (include/isvd.h) over the
// device stages.  Host pointers are staged through HBM; device pointers are used in place.
#include <hip/hip_runtime.h>

#include <cctype>
#include <chrono>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/isvd.h"
#include "la.hpp"
#include "fp8.hpp"
#include "mixed.hpp"
#include "stages.hpp"

size_t isvd_gpu_memory_limit = 0;

namespace isvd {

// ---------------------------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------------------------
static thread_local std::string g_last_error;
static int g_error_count = 0;

void set_last_error(const std::string &msg) { g_last_error = msg; }

void fail(const std::string &msg) {
  std::fprintf(stderr, "isvd: %s\n", msg.c_str());
  std::fflush(stderr);
  g_last_error = msg;
  ++g_error_count;
  const char *ab = std::getenv("ISVD_ASSERT_ABORT");
  if (ab && ab[0] == '1') std::abort();
  throw Error{msg};
}

// ---------------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------------
static Ctx g_ctx;
static bool g_init = false;
static std::mutex g_mutex;
static uint64_t g_pool_threshold = 0;  // the default pool's release threshold before init

// Scratch region of isvd_dPostprocessGramianNoAlloc: every temporary of that entry point
// (Z^T, the l x l Gram and its singular vectors, split-K partial sums) is carved from it, so
// the call allocates nothing.  Each GPU the library runs on holds one copy for the life of the
// process.  In host code `scratch` names a host-side stand-in: kernels and copies use the
// device address recorded at initialization (g_scratch_dev), never the symbol itself.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static unsigned char *g_scratch_dev = nullptr;

Ctx &ctx() { return g_ctx; }

void ensure_init() {
  if (g_init) return;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    fail("no HIP device available (the iSVD library runs on an MI355X GPU)");
  }
  HIP_CHECK(hipSetDevice(0));
  g_ctx.device = 0;
  HIP_CHECK(hipDeviceGetAttribute(&g_ctx.cus, hipDeviceAttributeMultiprocessorCount, 0));
  // A blocking stream: work the caller queued on the legacy default stream (e.g. the kernels
  // that produced device-resident inputs) completes before the library's kernels start.
  HIP_CHECK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamDefault));
  HIP_CHECK(hipStreamCreateWithFlags(&g_ctx.aux, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreateWithFlags(&g_ctx.fork, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&g_ctx.join, hipEventDisableTiming));
  // Stream-ordered buffers come from the device's default pool.  Keep freed memory cached in
  // the pool across synchronizations (the integration loops synchronize once per iteration):
  // with the default threshold of 0 every synchronization returns it to the driver and the
  // next allocation maps it again (milliseconds).  isvd_Finalize_gpu trims the pool.
  hipMemPool_t pool = nullptr;
  HIP_CHECK(hipDeviceGetDefaultMemPool(&pool, 0));
  HIP_CHECK(hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &g_pool_threshold));
  uint64_t keep = UINT64_MAX;
  HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  // device address of this GPU's copy of the no-allocation scratch region
  void *sp = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  g_scratch_dev = static_cast<unsigned char *>(sp);
  g_init = true;
}

void sync() { HIP_CHECK(hipStreamSynchronize(g_ctx.stream)); }

void *Arena::take(size_t bytes) {
  const size_t at = (used + 255) / 256 * 256;
  ISVD_CHECK(at <= cap && bytes <= cap - at,
             "the %zu-byte scratch region is exhausted (%zu bytes requested, %zu in use)", cap,
             bytes, at);
  used = at + bytes;
  return base + at;
}

static void *g_ws = nullptr;
static size_t g_ws_bytes = 0;

size_t splitk_budget() {
  if (!g_ctx.arena) return SIZE_MAX;
  const size_t at = (g_ctx.arena->used + 255) / 256 * 256;
  return at < g_ctx.arena->cap ? g_ctx.arena->cap - at : 0;
}

void *splitk_workspace(size_t bytes) {
  if (g_ctx.arena) {
    // the partials of one product live at the top of the arena, above every buffer taken so
    // far; later buffers may reuse that space, since the stream orders them after the product
    ISVD_CHECK(bytes <= splitk_budget(), "split-K partials (%zu bytes) exceed the scratch left",
               bytes);
    return g_ctx.arena->base + (g_ctx.arena->cap - bytes) / 256 * 256;
  }
  if (bytes > g_ws_bytes) {
    if (g_ws) HIP_CHECK(hipFreeAsync(g_ws, g_ctx.stream));
    g_ws = nullptr;
    g_ws_bytes = 0;
    const size_t grown = bytes + bytes / 4;
    HIP_CHECK(hipMallocAsync(&g_ws, grown, g_ctx.stream));
    g_ws_bytes = grown;
  }
  return g_ws;
}

static void *g_pinned = nullptr;
static hipEvent_t g_readback_ev = nullptr;
constexpr size_t kPinnedBytes = 4096;

// pinned bytes: read_back from the start, then kSlots slots of kSlotBytes at the end
constexpr size_t kSlotBytes = 256;
constexpr int kSlots = 4;
static hipEvent_t g_slot_ev[kSlots] = {};

static void ensure_pinned() {
  if (!g_pinned) {
    HIP_CHECK(hipHostMalloc(&g_pinned, kPinnedBytes, hipHostMallocDefault));
    HIP_CHECK(hipEventCreateWithFlags(&g_readback_ev, hipEventDisableTiming));
    for (int i = 0; i < kSlots; ++i)
      HIP_CHECK(hipEventCreateWithFlags(&g_slot_ev[i], hipEventDisableTiming));
  }
}

static void release_pinned() {
  if (!g_pinned) return;
  HIP_CHECK(hipEventDestroy(g_readback_ev));
  for (int i = 0; i < kSlots; ++i) HIP_CHECK(hipEventDestroy(g_slot_ev[i]));
  HIP_CHECK(hipHostFree(g_pinned));
  g_pinned = nullptr;
  g_readback_ev = nullptr;
  for (int i = 0; i < kSlots; ++i) g_slot_ev[i] = nullptr;
}

static void spin_on(hipEvent_t ev) {
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e == hipSuccess) break;
    if (e != hipErrorNotReady) HIP_CHECK(e);
  }
}

void read_back(void *host, const void *dev, size_t bytes) {
  ISVD_CHECK(bytes <= kPinnedBytes - kSlots * kSlotBytes, "read_back: %zu bytes is too many",
             bytes);
  ensure_pinned();
  HIP_CHECK(hipMemcpyAsync(g_pinned, dev, bytes, hipMemcpyDeviceToHost, g_ctx.stream));
  HIP_CHECK(hipEventRecord(g_readback_ev, g_ctx.stream));
  spin_on(g_readback_ev);
  std::memcpy(host, g_pinned, bytes);
}

static unsigned char *slot_ptr(int slot) {
  return static_cast<unsigned char *>(g_pinned) + kPinnedBytes - (size_t)(kSlots - slot) * kSlotBytes;
}

void *pinned_slot(int slot) {
  ISVD_CHECK(slot >= 0 && slot < kSlots, "pinned_slot: bad slot");
  ensure_pinned();
  void *dev = nullptr;
  HIP_CHECK(hipHostGetDevicePointer(&dev, slot_ptr(slot), 0));
  return dev;
}

void pinned_slot_record(int slot) {
  ISVD_CHECK(slot >= 0 && slot < kSlots, "pinned_slot_record: bad slot");
  ensure_pinned();
  HIP_CHECK(hipEventRecord(g_slot_ev[slot], g_ctx.stream));
}

hipEvent_t pinned_slot_event(int slot) {
  ISVD_CHECK(slot >= 0 && slot < kSlots, "pinned_slot_event: bad slot");
  ensure_pinned();
  return g_slot_ev[slot];
}

void read_back_wait(int slot, void *host, size_t bytes) {
  ISVD_CHECK(slot >= 0 && slot < kSlots && bytes <= kSlotBytes, "read_back_wait: bad slot");
  spin_on(g_slot_ev[slot]);
  std::memcpy(host, slot_ptr(slot), bytes);
}

bool is_device_pointer(const void *ptr) {
  if (!ptr) return false;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// A matrix argument: uses the caller's device memory in place, or stages host memory.
template <typename T>
struct Arg {
  T *host = nullptr;
  T *dev = nullptr;
  i64 count = 0;
  bool out = false;
  DBuf<T> buf;
  Arg(const T *p, i64 n, bool in, bool out_) : count(n), out(out_) {
    if (!p || n <= 0) return;
    if (is_device_pointer(p)) {
      dev = const_cast<T *>(p);
    } else {
      host = const_cast<T *>(p);
      buf.alloc(n);
      dev = buf.get();
      if (in)
        HIP_CHECK(hipMemcpyAsync(dev, host, sizeof(T) * n, hipMemcpyHostToDevice, ctx().stream));
    }
  }
  void finish() {
    if (host && out)
      HIP_CHECK(hipMemcpyAsync(host, dev, sizeof(T) * count, hipMemcpyDeviceToHost, ctx().stream));
  }
};

// elements spanned by an m x n column-major matrix with leading dimension ld
inline i64 span(i64 m, i64 n, i64 ld) { return (m <= 0 || n <= 0) ? 0 : (n - 1) * ld + m; }
// e4m3 rows are read and written in whole 16-byte chunks: a row of n codes spans roundup(n, 16)
inline i64 round16(i64 n) { return (n + 15) / 16 * 16; }

// Columns of a host-resident A staged per block under isvd_gpu_memory_limit, as in the
// reference's GPU stages (src/lib/libisvd/gpu/stage/@x@_sketch_gpu.h:63-77,
// @x@_postprocess_gpu.h:73-82): the budget min(free HBM, limit), less the bytes of the device
// buffers that stay resident through the stage, over the bytes of one column of A, in
// multiples of 64 columns.  n (A staged whole) when no limit is set or all of A fits.
static i64 host_block_cols(i64 n, i64 col_bytes, i64 resident_bytes) {
  if (isvd_gpu_memory_limit == 0) return n;
  size_t free_b = 0, total_b = 0;
  HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const size_t budget = free_b < isvd_gpu_memory_limit ? free_b : isvd_gpu_memory_limit;
  ISVD_CHECK(budget >= (size_t)(resident_bytes + col_bytes),
             "not enough GPU memory: at least %lld bytes are needed (%lld resident + one column "
             "of A), %zu are available under isvd_gpu_memory_limit",
             resident_bytes + col_bytes, resident_bytes, budget);
  i64 cols = (i64)((budget - (size_t)resident_bytes) / (size_t)col_bytes);
  if (cols >= n) return n;
  if (cols > 64) cols = cols / 64 * 64;
  return cols;
}

// ---------------------------------------------------------------------------------------------
// argument parsing: isvd_arg2char / isvd_arg2str (src/lib/libisvd/util/arg.h:22-79)
// ---------------------------------------------------------------------------------------------
static char arg2char(const char *name, char arg, const char *opts) {
  for (const char *o = opts; *o; ++o)
    if (std::toupper((unsigned char)arg) == *o) return *o;
  std::fprintf(stderr, "%s ('%c') must be one of the following characters: ", name, arg);
  for (const char *o = opts; *o; ++o) std::fprintf(stderr, o == opts ? "'%c'" : ", '%c'", *o);
  std::fprintf(stderr, ".\n");
  return 0;
}

static const char *arg2str(const char *name, const char *arg, const char *const *opts, int nopts) {
  if (!arg) arg = "";
  for (int i = 0; i < nopts; ++i) {
    const char *o = opts[i];
    if (std::strlen(o) != std::strlen(arg)) continue;
    bool same = true;
    for (size_t q = 0; o[q]; ++q)
      if (std::toupper((unsigned char)o[q]) != std::toupper((unsigned char)arg[q])) same = false;
    if (same) return o;
  }
  std::fprintf(stderr, "%s ('%s') must be one of the following strings: ", name, arg);
  for (int i = 0; i < nopts; ++i) std::fprintf(stderr, i == 0 ? "\"%s\"" : ", \"%s\"", opts[i]);
  std::fprintf(stderr, ".\n");
  return nullptr;
}

static Dims dims_of(const isvd_Param &p) {
  Dims d;
  d.m = p.nrow; d.n = p.ncol; d.k = p.rank; d.l = p.dim_sketch; d.N = p.num_sketch;
  d.Nl = p.dim_sketch_total;
  return d;
}

static void check_param(const isvd_Param &p) {
  ISVD_CHECK(p.nrow > 0 && p.ncol > 0 && p.rank > 0 && p.over_rank >= 0 && p.num_sketch > 0,
             "invalid parameters (m=%lld n=%lld k=%lld p=%lld N=%lld)", (i64)p.nrow,
             (i64)p.ncol, (i64)p.rank, (i64)p.over_rank, (i64)p.num_sketch);
  ISVD_CHECK(p.dim_sketch == p.rank + p.over_rank &&
                 p.dim_sketch_total == p.dim_sketch * p.num_sketch,
             "inconsistent parameters (l=%lld, Nl=%lld)", (i64)p.dim_sketch,
             (i64)p.dim_sketch_total);
  ISVD_CHECK(p.mpi_size == 1, "this library runs on one GPU (mpi_size must be 1)");
}

// Runs `body` under the library lock; converts failures into an error return.
template <typename F>
static bool guarded(F &&body) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error.clear();
  try {
    body();
    return true;
  } catch (const Error &) {
    if (g_init) (void)hipStreamSynchronize(g_ctx.stream);
    return false;
  }
}

// ---------------------------------------------------------------------------------------------
// precision families: s and d (the reference's), b (A and Omega in BF16, the rest FP32)
// ---------------------------------------------------------------------------------------------
struct PrecS {
  using TA = float;
  using T = float;
  static constexpr bool mixed = false;
  static constexpr const char *name = "Single Precision";
};
struct PrecD {
  using TA = double;
  using T = double;
  static constexpr bool mixed = false;
  static constexpr const char *name = "Double Precision";
};
struct PrecB {
  using TA = uint16_t;
  using T = float;
  static constexpr bool mixed = true;
  static constexpr const char *name = "Mixed Precision (BF16 A)";
};

// ---------------------------------------------------------------------------------------------
// stage entry points (templated, wrapped by the extern "C" functions below)
// ---------------------------------------------------------------------------------------------
template <class P>
static void sketch_impl(const isvd_Param &param, char dista, char ordera,
                        const typename P::TA *a, i64 lda, const typename P::TA *omegat, i64 ldo,
                        typename P::T *yst, i64 ldyst, bool use_seed, unsigned long long seed) {
  using T = typename P::T;
  using TA = typename P::TA;
  const char dista_ = arg2char("DISTA", dista, "CR");
  const char ordera_ = arg2char("ORDERA", ordera, "CR");
  if (!dista_ || !ordera_) return;
  check_param(param);
  const Dims d = dims_of(param);
  // @x@_sketch.h:50-55 (block-col) / :115-120 (block-row)
  ISVD_CHECK(lda >= (ordera_ == 'C' ? d.m : d.n), "lda (%lld) too small", lda);
  if (dista_ == 'C')
    ISVD_CHECK(ldyst == d.Nl, "ldyst (%lld) must equal Nl (%lld)", ldyst, d.Nl);
  else
    ISVD_CHECK(ldyst >= d.Nl, "ldyst (%lld) must be >= Nl (%lld)", ldyst, d.Nl);
  ensure_init();
  const i64 a_span = ordera_ == 'C' ? span(d.m, d.n, lda) : span(d.n, d.m, lda);
  // a host-resident A beyond the memory limit is staged by column blocks
  const bool a_host = a != nullptr && !is_device_pointer(a);
  const i64 cols = a_host ? host_block_cols(d.n, d.m * (i64)sizeof(TA),
                                            d.Nl * (d.m * (i64)sizeof(T) + d.n * (i64)sizeof(TA)))
                          : d.n;
  ISVD_CHECK(cols == d.n || !P::mixed,
             "the mixed-precision family stages a host-resident A whole: raise "
             "isvd_gpu_memory_limit or pass device memory");
  Arg<TA> A(cols == d.n ? a : nullptr, a_span, true, false);
  Arg<T> Y(yst, span(d.Nl, d.m, ldyst), false, true);
  DBuf<TA> om;
  const TA *omd;
  if (!use_seed) ISVD_CHECK(ldo >= d.Nl, "ldomegat (%lld) must be >= Nl (%lld)", ldo, d.Nl);
  Arg<TA> O(use_seed ? nullptr : omegat, use_seed ? 0 : span(d.Nl, d.n, ldo), true, false);
  if (use_seed) {
    om.alloc(d.Nl * d.n);
    if constexpr (P::mixed) gaussian_omega_b(d.Nl, d.n, om.get(), d.Nl, seed);
    else gaussian_omega<T>(d.Nl, d.n, om.get(), d.Nl, seed);
    omd = om.get();
    ldo = d.Nl;
  } else {
    omd = O.dev;
  }
  if constexpr (P::mixed) stage_sketch_b(d, ordera_, A.dev, lda, omd, ldo, Y.dev, ldyst);
  else if (cols < d.n) stage_sketch_host_blocks<T>(d, ordera_, a, lda, omd, ldo, Y.dev, ldyst, cols);
  else stage_sketch<T>(d, ordera_, A.dev, lda, omd, ldo, Y.dev, ldyst);
  Y.finish();
  sync();
}

template <class P>
static void orth_impl(const isvd_Param &param, i64 argc, typename P::T *yst, i64 ldyst,
                      bool tsqr) {
  using T = typename P::T;
  if (argc < 0) return;
  check_param(param);
  const Dims d = dims_of(param);
  ISVD_CHECK(ldyst >= d.Nl, "ldyst (%lld) must be >= Nl (%lld)", ldyst, d.Nl);
  ISVD_CHECK(d.l <= 128, "l = k + p must be <= 128 (got %lld)", d.l);
  if (tsqr) ISVD_CHECK(d.m >= d.l, "tall-skinny QR needs m >= l");
  ensure_init();
  Arg<T> Y(yst, span(d.Nl, d.m, ldyst), true, true);
  if constexpr (P::mixed) {
    if (tsqr) stage_orth_tsqr_b(d, Y.dev, ldyst);
    else stage_orth_gramian_b(d, Y.dev, ldyst);
  } else {
    if (tsqr) stage_orth_tsqr<T>(d, Y.dev, ldyst);
    else stage_orth_gramian<T>(d, Y.dev, ldyst);
  }
  Y.finish();
  sync();
}

template <class P>
static void kn_impl(const isvd_Param &param, const typename P::T *argv, i64 argc,
                    typename P::T *retv, i64 retc, const typename P::T *yst, i64 ldyst,
                    typename P::T *qt, i64 ldqt) {
  using T = typename P::T;
  KnArgs ka;
  // query (@x@_integrate_kolmogorov_nagumo.c:56-61)
  if (argc < 0) {
    if (retc > 0) retv[0] = (T)ka.maxit;
    if (retc > 1) retv[1] = (T)ka.tol;
    return;
  }
  if (argc > 0) ISVD_CHECK(argv != nullptr, "argv is NULL");
  if (retc > 0) ISVD_CHECK(retv != nullptr, "retv is NULL");
  if (argc > 0) ka.maxit = (i64)argv[0];
  if (argc > 1) ka.tol = (double)argv[1];
  check_param(param);
  const Dims d = dims_of(param);
  ISVD_CHECK(ldyst >= d.Nl, "ldyst (%lld) must be >= Nl", ldyst);
  ISVD_CHECK(ldqt >= d.l, "ldqt (%lld) must be >= l", ldqt);
  ISVD_CHECK(ka.maxit >= 0 && ka.tol >= 0, "maxit and tol must be >= 0");
  ISVD_CHECK(d.l <= 128, "l = k + p must be <= 128 (got %lld)", d.l);
  ensure_init();
  Arg<T> Y(yst, span(d.Nl, d.m, ldyst), true, false);
  Arg<T> Q(qt, span(d.l, d.m, ldqt), false, true);
  i64 it = 0;
  double err = 0;
  if constexpr (P::mixed) stage_integrate_kn_b(d, ka, Y.dev, ldyst, Q.dev, ldqt, &it, &err);
  else stage_integrate_kn<T>(d, ka, Y.dev, ldyst, Q.dev, ldqt, &it, &err);
  Q.finish();
  sync();
  if (retc > 0) retv[0] = (T)it;
  if (retc > 1) retv[1] = (T)err;
}

template <class P>
static void wy_impl(const isvd_Param &param, const typename P::T *argv, i64 argc,
                    typename P::T *retv, i64 retc, const typename P::T *yst, i64 ldyst,
                    typename P::T *qt, i64 ldqt) {
  using T = typename P::T;
  WyArgs wa;
  const double defaults[9] = {(double)wa.maxit, wa.tol, wa.tau0, wa.taumax, wa.taumin,
                              (double)wa.taumaxit, wa.beta, wa.sigma, wa.eta};
  if (argc < 0) {  // query (@x@_integrate_wen_yin.c:70-82)
    for (i64 i = 0; i < retc && i < 9; ++i) retv[i] = (T)defaults[i];
    return;
  }
  if (argc > 0) ISVD_CHECK(argv != nullptr, "argv is NULL");
  if (retc > 0) ISVD_CHECK(retv != nullptr, "retv is NULL");
  double v[9];
  for (int i = 0; i < 9; ++i) v[i] = (argc > i) ? (double)argv[i] : defaults[i];
  wa.maxit = (i64)v[0]; wa.tol = v[1]; wa.tau0 = v[2]; wa.taumax = v[3]; wa.taumin = v[4];
  wa.taumaxit = (i64)v[5]; wa.beta = v[6]; wa.sigma = v[7]; wa.eta = v[8];
  check_param(param);
  const Dims d = dims_of(param);
  ISVD_CHECK(ldyst >= d.Nl, "ldyst (%lld) must be >= Nl", ldyst);
  ISVD_CHECK(ldqt >= d.l, "ldqt (%lld) must be >= l", ldqt);
  // (@x@_integrate_wen_yin.c:103-113)
  ISVD_CHECK(wa.maxit >= 0 && wa.tol >= 0 && wa.tau0 >= 0 && wa.taumax >= 0 && wa.taumin >= 0 &&
                 wa.taumaxit >= 0,
             "WY arguments must be >= 0");
  ISVD_CHECK(wa.beta > 0 && wa.beta < 1 && wa.sigma > 0 && wa.sigma < 1 && wa.eta > 0 &&
                 wa.eta < 1,
             "WY beta, sigma, eta must lie in (0, 1)");
  ISVD_CHECK(d.l <= 128, "l = k + p must be <= 128 (got %lld)", d.l);
  ensure_init();
  Arg<T> Y(yst, span(d.Nl, d.m, ldyst), true, false);
  Arg<T> Q(qt, span(d.l, d.m, ldqt), false, true);
  i64 it = 0;
  double mu = 0;
  if constexpr (P::mixed) stage_integrate_wy_b(d, wa, Y.dev, ldyst, Q.dev, ldqt, &it, &mu);
  else stage_integrate_wy<T>(d, wa, Y.dev, ldyst, Q.dev, ldqt, &it, &mu);
  Q.finish();
  sync();
  if (retc > 0) retv[0] = (T)it;
  if (retc > 1) retv[1] = (T)mu;
}

// @x@_integrate_hierarchical_reduction.c:36-141 (yst is destroyed, as there)
template <class P>
static void hr_impl(const isvd_Param &param, i64 argc, typename P::T *yst, i64 ldyst,
                    typename P::T *qt, i64 ldqt) {
  using T = typename P::T;
  if (argc < 0) return;
  check_param(param);
  const Dims d = dims_of(param);
  ISVD_CHECK(ldyst >= d.Nl, "ldyst (%lld) must be >= Nl", ldyst);
  ISVD_CHECK(ldqt >= d.l, "ldqt (%lld) must be >= l", ldqt);
  ISVD_CHECK(d.l <= 128, "l = k + p must be <= 128 (got %lld)", d.l);
  ensure_init();
  Arg<T> Y(yst, span(d.Nl, d.m, ldyst), true, false);
  Arg<T> Q(qt, span(d.l, d.m, ldqt), false, true);
  stage_integrate_hr<T>(d, Y.dev, ldyst, Q.dev, ldqt);
  Q.finish();
  sync();
}

template <class P>
static void post_impl(const isvd_Param &param, i64 argc, char dista, char ordera,
                      const typename P::TA *a, i64 lda, const typename P::T *qt, i64 ldqt,
                      typename P::T *s, typename P::T *ut, i64 ldut, typename P::T *vt,
                      i64 ldvt, int ut_root, int vt_root, char kind) {
  using T = typename P::T;
  using TA = typename P::TA;
  const bool tsqr = (kind == 'T');
  if (argc < 0) return;
  const char dista_ = arg2char("DISTA", dista, "CR");
  const char ordera_ = arg2char("ORDERA", ordera, "CR");
  if (!dista_ || !ordera_) return;
  check_param(param);
  const Dims d = dims_of(param);
  // @x@_postprocess_gramian.c:99-119 / @x@_postprocess.h:44-54, :121-131
  if (ut_root >= 0) ISVD_CHECK(ldut == d.l, "ldut (%lld) must equal l (%lld)", ldut, d.l);
  else if (ut_root == -1) ISVD_CHECK(ldut >= d.l, "ldut (%lld) must be >= l", ldut);
  if (vt_root >= 0) ISVD_CHECK(ldvt == d.l, "ldvt (%lld) must equal l (%lld)", ldvt, d.l);
  else if (vt_root == -1) ISVD_CHECK(ldvt >= d.l, "ldvt (%lld) must be >= l", ldvt);
  ISVD_CHECK(lda >= (ordera_ == 'C' ? d.m : d.n), "lda (%lld) too small", lda);
  ISVD_CHECK(ldqt >= d.l, "ldqt (%lld) must be >= l", ldqt);
  ISVD_CHECK(d.l <= 128, "l = k + p must be <= 128 (got %lld)", d.l);
  ISVD_CHECK(s != nullptr, "s is NULL");
  const bool want_u = ut_root >= -1, want_v = vt_root >= -1;
  if (want_u) ISVD_CHECK(ut != nullptr, "ut is NULL");
  if (want_v) ISVD_CHECK(vt != nullptr, "vt is NULL");
  if (kind == 'S') {
    ISVD_CHECK(!P::mixed, "symmetric postprocessing has no mixed-precision form");
    ISVD_CHECK(d.m == d.n, "symmetric postprocessing needs m == n (m = %lld, n = %lld)", d.m,
               d.n);
  }
  ensure_init();
  const i64 a_span = ordera_ == 'C' ? span(d.m, d.n, lda) : span(d.n, d.m, lda);
  // a host-resident A beyond the memory limit is staged by column blocks; Qbar^T, Z^T and the
  // outputs stay resident
  const bool a_host = a != nullptr && !is_device_pointer(a);
  const i64 sz = (i64)sizeof(T);
  const i64 cols =
      a_host ? host_block_cols(d.n, d.m * (i64)sizeof(TA),
                               sz * (d.l * d.m + d.l * d.n + d.k * (d.m + d.n) + 4 * d.l * d.l))
             : d.n;
  ISVD_CHECK(cols == d.n || (!P::mixed && kind != 'S'),
             "this postprocessing stages a host-resident A whole: raise isvd_gpu_memory_limit "
             "or pass device memory");
  Arg<TA> A(cols == d.n ? a : nullptr, a_span, true, false);
  Arg<T> Q(qt, span(d.l, d.m, ldqt), true, false);
  Arg<T> S(s, d.l, false, true);
  if constexpr (!P::mixed) {
    if (kind == 'S') {  // @x@_postprocess_symmetric.c:110-114, :163-171: U goes to vt if ut is skipped
      T *u_ = want_u ? ut : vt;
      const i64 ldu_ = want_u ? ldut : ldvt;
      const bool want = want_u || want_v;
      Arg<T> U(want ? u_ : nullptr, span(d.k, d.m, ldu_), false, true);
      stage_post_symmetric<T>(d, ordera_, A.dev, lda, Q.dev, ldqt, S.dev, U.dev, ldu_, want);
      Arg<T> V(want_u && want_v ? vt : nullptr, span(d.k, d.n, ldvt), false, true);
      if (want_u && want_v) omatcopy<T>('N', d.k, d.m, T(1), U.dev, ldut, V.dev, ldvt);
      S.finish();
      U.finish();
      V.finish();
      sync();
      return;
    }
  }
  Arg<T> U(want_u ? ut : nullptr, span(d.k, d.m, ldut), false, true);
  Arg<T> V(want_v ? vt : nullptr, span(d.k, d.n, ldvt), false, true);
  if constexpr (P::mixed) {
    stage_post_b(d, tsqr, ordera_, A.dev, lda, Q.dev, ldqt, S.dev, U.dev, ldut, V.dev, ldvt,
                 want_u, want_v);
  } else if (cols < d.n) {
    DBuf<T> zt(d.l * d.n);
    project_host_blocks<T>(d, ordera_, a, lda, Q.dev, ldqt, zt.get(), cols);
    if (tsqr)
      post_tsqr_z<T>(d, Q.dev, ldqt, zt.get(), S.dev, U.dev, ldut, V.dev, ldvt, want_u, want_v);
    else
      post_gramian_z<T>(d, Q.dev, ldqt, zt.get(), S.dev, U.dev, ldut, V.dev, ldvt, want_u,
                        want_v);
  } else if (tsqr) {
    stage_post_tsqr<T>(d, ordera_, A.dev, lda, Q.dev, ldqt, S.dev, U.dev, ldut, V.dev, ldvt,
                       want_u, want_v);
  } else {
    stage_post_gramian<T>(d, ordera_, A.dev, lda, Q.dev, ldqt, S.dev, U.dev, ldut, V.dev, ldvt,
                          want_u, want_v);
  }
  S.finish();
  U.finish();
  V.finish();
  sync();
}

// Gramian postprocessing without allocation: the checks and stages of post_impl (kind 'G', FP64)
// with every temporary carved from this GPU's scratch region.  Device pointers only (staging
// host arrays would allocate).  Z^T, W and U_W are taken from the bottom of the region, each
// product's split-K partial sums from the space above them (fewer K chunks when that space is
// short), so the footprint is 8 (l n + 2 l^2) bytes plus alignment.  Nothing in the region needs
// to start at zero: every kernel writes what it later reads.
static void post_gramian_noalloc_impl(const isvd_Param &param, i64 argc, char dista, char ordera,
                                      const double *a, i64 lda, const double *qt, i64 ldqt,
                                      double *s, double *ut, i64 ldut, double *vt, i64 ldvt,
                                      int ut_root, int vt_root) {
  if (argc < 0) return;
  const char dista_ = arg2char("DISTA", dista, "CR");
  const char ordera_ = arg2char("ORDERA", ordera, "CR");
  if (!dista_ || !ordera_) return;
  check_param(param);
  const Dims d = dims_of(param);
  if (ut_root >= 0) ISVD_CHECK(ldut == d.l, "ldut (%lld) must equal l (%lld)", ldut, d.l);
  else if (ut_root == -1) ISVD_CHECK(ldut >= d.l, "ldut (%lld) must be >= l", ldut);
  if (vt_root >= 0) ISVD_CHECK(ldvt == d.l, "ldvt (%lld) must equal l (%lld)", ldvt, d.l);
  else if (vt_root == -1) ISVD_CHECK(ldvt >= d.l, "ldvt (%lld) must be >= l", ldvt);
  ISVD_CHECK(lda >= (ordera_ == 'C' ? d.m : d.n), "lda (%lld) too small", lda);
  ISVD_CHECK(ldqt >= d.l, "ldqt (%lld) must be >= l", ldqt);
  ISVD_CHECK(d.l <= 128, "l = k + p must be <= 128 (got %lld)", d.l);
  const bool want_u = ut_root >= -1, want_v = vt_root >= -1;
  const i64 need = (d.l * d.n * 8 + 255) / 256 * 256 + 2 * ((d.l * d.l * 8 + 255) / 256 * 256);
  ISVD_CHECK(need <= (i64)kScratchBytes,
             "Z^T, W and U_W need %lld bytes (8 (l n + 2 l^2)), more than the %zu-byte scratch "
             "region (l = %lld, n = %lld)",
             need, kScratchBytes, d.l, d.n);
  ensure_init();
  ISVD_CHECK(g_scratch_dev != nullptr, "the scratch region has no device address");
  auto dev = [](const void *p, const char *name) {
    ISVD_CHECK(p != nullptr && is_device_pointer(p),
               "%s must be device memory (staging a host array would allocate)", name);
  };
  dev(a, "a");
  dev(qt, "qt");
  dev(s, "s");
  if (want_u) dev(ut, "ut");
  if (want_v) dev(vt, "vt");
  Arena arena;
  arena.base = g_scratch_dev;
  arena.cap = kScratchBytes;
  struct Install {
    explicit Install(Arena *a) { g_ctx.arena = a; }
    ~Install() { g_ctx.arena = nullptr; }
  } install(&arena);
  stage_post_gramian<double>(d, ordera_, a, lda, qt, ldqt, s, ut, ldut, vt, ldvt, want_u, want_v);
  sync();
}

// ---------------------------------------------------------------------------------------------
// driver (src/lib/libisvd/core/driver/@x@_isvd.c:106-282)
// ---------------------------------------------------------------------------------------------
static const char *const kOptsS[] = {"GP", "GP_gpu"};
static const char *const kOptsO[] = {"TS", "GR"};
static const char *const kOptsI[] = {"KN", "WY", "HR"};
static const char *const kOptsP[] = {"TS", "GR", "SY", "TS_gpu", "GR_gpu", "SY_gpu"};

static const char *alg_name(char kind, const char *alg, const char *prec) {
  static thread_local std::string s;
  std::string name;
  if (kind == 'S') name = "Gaussian Projection Sketching";
  else if (kind == 'O') name = (alg[0] == 'T') ? "Tall-Skinny QR Orthogonalization" : "Gramian Orthogonalization";
  else if (kind == 'I')
    name = (alg[0] == 'K')   ? "Kolmogorov-Nagumo Integration"
           : (alg[0] == 'W') ? "Wen-Yin Integration"
                             : "Hierarchical Reduction Integration";
  else
    name = (alg[0] == 'T')   ? "Tall-Skinny QR Postprocessing"
           : (alg[0] == 'G') ? "Gramian Postprocessing"
                             : "Symmetric Postprocessing";
  s = std::string(prec) + " GPU (MI355X) " + name;
  return s.c_str();
}

// FP8 sketch input of the e family: A as e4m3 codes with per-128 x 128 block scales
struct E8Input {
  const uint8_t *a8;
  i64 lda8;
  const float *scale;
  i64 lds;
};

template <class P>
static void isvd_impl(const isvd_Param &param, const char *alg_s, const char *alg_o,
                      const char *alg_i, const char *alg_p, const typename P::T *argv[4],
                      const isvd_int_t argc[4], typename P::T *retv[4], const isvd_int_t retc[4],
                      double time[4], FILE *stream, char dista, char ordera,
                      const typename P::TA *a, i64 lda, const typename P::TA *omegat, i64 ldo,
                      typename P::T *s, typename P::T *ut, i64 ldut, typename P::T *vt,
                      i64 ldvt, bool use_seed, unsigned long long seed, int ut_root,
                      int vt_root, const E8Input *e8 = nullptr) {
  using T = typename P::T;
  using TA = typename P::TA;
  const char *as = arg2str("ALG_S", alg_s, kOptsS, 2);
  const char *ao = arg2str("ALG_O", alg_o, kOptsO, 2);
  const char *ai = arg2str("ALG_I", alg_i, kOptsI, 3);
  const char *ap = arg2str("ALG_P", alg_p, kOptsP, 6);
  if (!as || !ao || !ai || !ap) return;
  const T *argv_i = argv ? argv[2] : nullptr;
  const i64 argc_s = argv ? argc[0] : 0, argc_o = argv ? argc[1] : 0, argc_i = argv ? argc[2] : 0,
            argc_p = argv ? argc[3] : 0;
  T *retv_i = retv ? retv[2] : nullptr;
  const i64 retc_i = retv ? retc[2] : 0;
  if (argc_i > 0) ISVD_CHECK(argv_i != nullptr, "argv[2] is NULL");
  if (retc_i > 0) ISVD_CHECK(retv_i != nullptr, "retv[2] is NULL");
  const bool kn = (ai[0] == 'K'), hr = (ai[0] == 'H'), sy = (ap[0] == 'S');
  // query mode (@x@_isvd.c:199-218)
  if (argc_s < 0 || argc_o < 0 || argc_i < 0 || argc_p < 0) {
    if (argc_i < 0) {
      if (kn) kn_impl<P>(param, argv_i, argc_i, retv_i, retc_i, nullptr, 1, nullptr, 1);
      else if (!hr) wy_impl<P>(param, argv_i, argc_i, retv_i, retc_i, nullptr, 1, nullptr, 1);
    }
    return;
  }
  const char dista_ = arg2char("DISTA", dista, "CR");
  const char ordera_ = arg2char("ORDERA", ordera, "CR");
  if (!dista_ || !ordera_) return;
  check_param(param);
  const Dims d = dims_of(param);
  const char *prec_name = e8 ? "Mixed Precision (FP8 e4m3 sketch, BF16 A)" : P::name;
  if (stream) {
    std::fprintf(stream, "Using %s.\n", alg_name('S', as, prec_name));
    std::fprintf(stream, "Using %s.\n", alg_name('O', ao, prec_name));
    std::fprintf(stream, "Using %s.\n", alg_name('I', ai, prec_name));
    std::fprintf(stream, "Using %s.\n\n", alg_name('P', ap, prec_name));
  }
  // argument checks of every stage, before any device work
  ISVD_CHECK(lda >= (ordera_ == 'C' ? d.m : d.n), "lda (%lld) too small", lda);
  ISVD_CHECK(d.l <= 128, "l = k + p must be <= 128 (got %lld)", d.l);
  const bool want_u = ut_root >= -1, want_v = vt_root >= -1;
  if (ut_root >= 0) ISVD_CHECK(ldut == d.l, "ldut (%lld) must equal l (%lld)", ldut, d.l);
  else if (ut_root == -1) ISVD_CHECK(ldut >= d.l, "ldut (%lld) must be >= l", ldut);
  if (vt_root >= 0) ISVD_CHECK(ldvt == d.l, "ldvt (%lld) must equal l (%lld)", ldvt, d.l);
  else if (vt_root == -1) ISVD_CHECK(ldvt >= d.l, "ldvt (%lld) must be >= l", ldvt);
  ISVD_CHECK(s != nullptr, "s is NULL");
  if (ao[0] == 'T') ISVD_CHECK(d.m >= d.l, "tall-skinny QR needs m >= l");
  if (ap[0] == 'T') ISVD_CHECK(d.n >= d.l, "tall-skinny QR postprocessing needs n >= l");
  if (!use_seed) ISVD_CHECK(ldo >= d.Nl, "ldomegat (%lld) must be >= Nl", ldo);
  if (P::mixed) ISVD_CHECK(!hr && !sy, "HR integration and SY postprocessing have no mixed-precision form");
  if (sy) ISVD_CHECK(d.m == d.n, "symmetric postprocessing needs m == n");
  if (e8) {
    ISVD_CHECK(e8->lda8 >= d.n && e8->lda8 % 16 == 0,
               "lda8 (%lld) must be >= n and a multiple of 16", e8->lda8);
    ISVD_CHECK(e8->lds >= (d.n + kF8Block - 1) / kF8Block,
               "ldscale (%lld) must be >= ceil(n / 128)", e8->lds);
    ISVD_CHECK(e8->a8 != nullptr && e8->scale != nullptr, "a8 and scale must not be NULL");
  }
  KnArgs ka;
  WyArgs wa;
  if (argc_i > 0 && !hr) {
    if (kn) {
      ka.maxit = (i64)argv_i[0];
      if (argc_i > 1) ka.tol = (double)argv_i[1];
    } else {
      double *f[9] = {nullptr, &wa.tol, &wa.tau0, &wa.taumax, &wa.taumin, nullptr, &wa.beta,
                      &wa.sigma, &wa.eta};
      for (i64 i = 0; i < argc_i && i < 9; ++i) {
        if (i == 0) wa.maxit = (i64)argv_i[0];
        else if (i == 5) wa.taumaxit = (i64)argv_i[5];
        else *f[i] = (double)argv_i[i];
      }
    }
  }
  (void)argc_s; (void)argc_o; (void)argc_p;
  ensure_init();
  const i64 a_span = ordera_ == 'C' ? span(d.m, d.n, lda) : span(d.n, d.m, lda);
  // a host-resident A beyond the memory limit is staged by column blocks, once for the sketch
  // and once for the projection; the sketch, the bases, Omega^T, Z^T and the outputs stay
  // resident
  const bool a_host = a != nullptr && !is_device_pointer(a) && !e8;
  const i64 sz = (i64)sizeof(T);
  const i64 cols = a_host ? host_block_cols(d.n, d.m * (i64)sizeof(TA),
                                            sz * (d.Nl * d.m + d.l * d.m + d.l * d.n +
                                                  d.k * (d.m + d.n)) +
                                                (i64)sizeof(TA) * d.Nl * d.n)
                          : d.n;
  ISVD_CHECK(cols == d.n || (!P::mixed && !sy),
             "this decomposition stages a host-resident A whole: raise isvd_gpu_memory_limit "
             "or pass device memory");
  Arg<TA> A(cols == d.n ? a : nullptr, a_span, true, false);
  Arg<T> S(s, d.l, false, true);
  Arg<T> U(want_u ? ut : nullptr, span(d.k, d.m, ldut), false, true);
  Arg<T> V(want_v ? vt : nullptr, span(d.k, d.n, ldvt), false, true);
  DBuf<T> yst(d.Nl * d.m), qt(d.l * d.m);
  DBuf<TA> om;
  hipEvent_t ev[5];
  for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
  auto stage_msg = [&](const char *msg) {
    if (stream) { std::fprintf(stream, "%s", msg); std::fflush(stream); }
  };
  i64 it = 0;
  double err = 0;
  // sketching (@x@_isvd.c:243-247)
  stage_msg("Sketching ..................... ");
  HIP_CHECK(hipEventRecord(ev[0], ctx().stream));
  bool sketched = false;
  if constexpr (P::mixed) {
    if (e8) {
      // e4m3 A (row-major) and its block scales; Omega drawn as e4m3 with K contiguous
      const i64 nbi = (d.m + kF8Block - 1) / kF8Block, nbj = (d.n + kF8Block - 1) / kF8Block;
      Arg<uint8_t> A8(e8->a8, span(round16(d.n), d.m, e8->lda8), true, false);
      Arg<float> SC(e8->scale, span(nbj, nbi, e8->lds), true, false);
      const i64 ldo8 = (d.n + 15) / 16 * 16;
      DBuf<uint8_t> om8(ldo8 * d.Nl);
      gaussian_omega_e(d.n, d.Nl, om8.get(), ldo8, seed);
      stage_sketch_e(d, A8.dev, e8->lda8, SC.dev, e8->lds, om8.get(), ldo8, yst.get(), d.Nl);
      sketched = true;
    }
  }
  if (!sketched) {
    const TA *omd = omegat;
    Arg<TA> O(use_seed ? nullptr : omegat, use_seed ? 0 : span(d.Nl, d.n, ldo), true, false);
    if (use_seed) {
      om.alloc(d.Nl * d.n);
      if constexpr (P::mixed) gaussian_omega_b(d.Nl, d.n, om.get(), d.Nl, seed);
      else gaussian_omega<T>(d.Nl, d.n, om.get(), d.Nl, seed);
      omd = om.get();
      ldo = d.Nl;
    } else {
      omd = O.dev;
    }
    if constexpr (P::mixed) stage_sketch_b(d, ordera_, A.dev, lda, omd, ldo, yst.get(), d.Nl);
    else if (cols < d.n)
      stage_sketch_host_blocks<T>(d, ordera_, a, lda, omd, ldo, yst.get(), d.Nl, cols);
    else stage_sketch<T>(d, ordera_, A.dev, lda, omd, ldo, yst.get(), d.Nl);
  }
  HIP_CHECK(hipEventRecord(ev[1], ctx().stream));
  stage_msg("done.\nOrthogonalizing ............... ");
  // before KN: KN reads the bases only through their projectors, so CholeskyQR2 may replace
  // either orthogonalization (see stage_orth_tsqr / stage_orth_gramian)
  if constexpr (P::mixed) {
    if (ao[0] == 'T') stage_orth_tsqr_b(d, yst.get(), d.Nl, kn);
    else stage_orth_gramian_b(d, yst.get(), d.Nl, kn);
  } else {
    if (ao[0] == 'T') stage_orth_tsqr<T>(d, yst.get(), d.Nl, kn);
    else stage_orth_gramian<T>(d, yst.get(), d.Nl, kn);
  }
  HIP_CHECK(hipEventRecord(ev[2], ctx().stream));
  stage_msg("done.\nIntegrating ................... ");
  if constexpr (P::mixed) {
    if (kn) stage_integrate_kn_b(d, ka, yst.get(), d.Nl, qt.get(), d.l, &it, &err);
    else stage_integrate_wy_b(d, wa, yst.get(), d.Nl, qt.get(), d.l, &it, &err);
  } else {
    if (kn) stage_integrate_kn<T>(d, ka, yst.get(), d.Nl, qt.get(), d.l, &it, &err);
    else if (hr) stage_integrate_hr<T>(d, yst.get(), d.Nl, qt.get(), d.l);
    else stage_integrate_wy<T>(d, wa, yst.get(), d.Nl, qt.get(), d.l, &it, &err);
  }
  HIP_CHECK(hipEventRecord(ev[3], ctx().stream));
  stage_msg("done.\nPostprocessing ................ ");
  if constexpr (P::mixed) {
    stage_post_b(d, ap[0] == 'T', ordera_, A.dev, lda, qt.get(), d.l, S.dev, U.dev, ldut, V.dev,
                 ldvt, want_u, want_v);
  } else if (sy) {  // U goes to vt when ut is skipped; V = U (@x@_postprocess_symmetric.c)
    const bool want = want_u || want_v;
    stage_post_symmetric<T>(d, ordera_, A.dev, lda, qt.get(), d.l, S.dev,
                            want_u ? U.dev : V.dev, want_u ? ldut : ldvt, want);
    if (want_u && want_v) omatcopy<T>('N', d.k, d.m, T(1), U.dev, ldut, V.dev, ldvt);
  } else if (cols < d.n) {
    DBuf<T> zt(d.l * d.n);
    project_host_blocks<T>(d, ordera_, a, lda, qt.get(), d.l, zt.get(), cols);
    if (ap[0] == 'T')
      post_tsqr_z<T>(d, qt.get(), d.l, zt.get(), S.dev, U.dev, ldut, V.dev, ldvt, want_u, want_v);
    else
      post_gramian_z<T>(d, qt.get(), d.l, zt.get(), S.dev, U.dev, ldut, V.dev, ldvt, want_u,
                        want_v);
  } else if (ap[0] == 'T') {
    stage_post_tsqr<T>(d, ordera_, A.dev, lda, qt.get(), d.l, S.dev, U.dev, ldut, V.dev, ldvt,
                       want_u, want_v);
  } else {
    stage_post_gramian<T>(d, ordera_, A.dev, lda, qt.get(), d.l, S.dev, U.dev, ldut, V.dev, ldvt,
                          want_u, want_v);
  }
  HIP_CHECK(hipEventRecord(ev[4], ctx().stream));
  S.finish();
  U.finish();
  V.finish();
  sync();
  stage_msg("done.\n");
  if (time) {
    for (int q = 0; q < 4; ++q) {
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
      time[q] = ms * 1e-3;
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (retv && retc) {
    if (retc[2] > 0 && retv[2]) retv[2][0] = (T)it;
    if (retc[2] > 1 && retv[2]) retv[2][1] = (T)err;
  }
}

// adjoint of the driver's rank-k output (no counterpart in the reference)
template <class P>
static void backward_impl(const isvd_Param &param, const char *alg_p, char dista, char ordera,
                          const typename P::T *a, i64 lda, const typename P::T *qt, i64 ldqt,
                          const typename P::T *gs, const typename P::T *gut, i64 ldgut,
                          const typename P::T *gvt, i64 ldgvt, typename P::T *da, i64 ldda) {
  using T = typename P::T;
  const char *ap = arg2str("ALG_P", alg_p, kOptsP, 6);
  const char dista_ = arg2char("DISTA", dista, "CR");
  const char ordera_ = arg2char("ORDERA", ordera, "CR");
  if (!ap || !dista_ || !ordera_) return;
  check_param(param);
  const Dims d = dims_of(param);
  ISVD_CHECK(d.l <= 128, "l = k + p must be <= 128 (got %lld)", d.l);
  ISVD_CHECK(ap[0] != 'S', "the adjoint covers the TS and GR postprocessing");
  ISVD_CHECK(d.n >= d.l || ap[0] != 'T', "tall-skinny QR postprocessing needs n >= l");
  const i64 rows = ordera_ == 'C' ? d.m : d.n, cols = ordera_ == 'C' ? d.n : d.m;
  ISVD_CHECK(lda >= rows, "lda (%lld) too small", lda);
  ISVD_CHECK(ldda >= rows, "ldda (%lld) too small", ldda);
  ISVD_CHECK(ldqt >= d.l, "ldqt (%lld) must be >= l", ldqt);
  ISVD_CHECK(a != nullptr && qt != nullptr && da != nullptr, "a, qt and da must not be NULL");
  if (gut) ISVD_CHECK(ldgut >= d.k, "ldgut (%lld) must be >= k", ldgut);
  if (gvt) ISVD_CHECK(ldgvt >= d.k, "ldgvt (%lld) must be >= k", ldgvt);
  ensure_init();
  Arg<T> A(a, span(rows, cols, lda), true, false);
  Arg<T> Q(qt, span(d.l, d.m, ldqt), true, false);
  Arg<T> GS(gs, d.k, true, false);
  Arg<T> GU(gut, gut ? span(d.k, d.m, ldgut) : 0, true, false);
  Arg<T> GV(gvt, gvt ? span(d.k, d.n, ldgvt) : 0, true, false);
  Arg<T> DA(da, span(rows, cols, ldda), false, true);
  // absent gradients are zero
  DBuf<T> zs, zu, zv;
  const T *gs_d = GS.dev, *gu_d = GU.dev, *gv_d = GV.dev;
  if (!gs_d) { zs.alloc(d.k); fill<T>(d.k, 1, T(0), zs.get(), d.k); gs_d = zs.get(); }
  if (!gu_d) { zu.alloc(d.k * d.m); fill<T>(d.k, d.m, T(0), zu.get(), d.k); gu_d = zu.get(); ldgut = d.k; }
  if (!gv_d) { zv.alloc(d.k * d.n); fill<T>(d.k, d.n, T(0), zv.get(), d.k); gv_d = zv.get(); ldgvt = d.k; }
  stage_isvd_backward<T>(d, ap[0] == 'T', ordera_, A.dev, lda, Q.dev, ldqt, gs_d, gu_d, ldgut,
                         gv_d, ldgvt, DA.dev, ldda);
  DA.finish();
  sync();
}

// e family: the FP8 (e4m3) sketch (no counterpart in the reference)
static void sketch_e_impl(const isvd_Param &param, char dista, const uint8_t *a8, i64 lda8,
                          const float *scale, i64 lds, const uint8_t *omega8, i64 ldo,
                          float *yst, i64 ldyst, bool use_seed, unsigned long long seed) {
  const char dista_ = arg2char("DISTA", dista, "CR");
  if (!dista_) return;
  check_param(param);
  const Dims d = dims_of(param);
  if (dista_ == 'C')
    ISVD_CHECK(ldyst == d.Nl, "ldyst (%lld) must equal Nl (%lld)", ldyst, d.Nl);
  else
    ISVD_CHECK(ldyst >= d.Nl, "ldyst (%lld) must be >= Nl (%lld)", ldyst, d.Nl);
  ISVD_CHECK(lda8 >= d.n && lda8 % 16 == 0, "lda8 (%lld) must be >= n and a multiple of 16",
             lda8);
  const i64 nbi = (d.m + kF8Block - 1) / kF8Block, nbj = (d.n + kF8Block - 1) / kF8Block;
  ISVD_CHECK(lds >= nbj, "ldscale (%lld) must be >= ceil(n / 128) (%lld)", lds, nbj);
  if (!use_seed)
    ISVD_CHECK(ldo >= d.n && ldo % 16 == 0,
               "ldomega (%lld) must be >= n and a multiple of 16", ldo);
  ISVD_CHECK(a8 != nullptr && scale != nullptr && yst != nullptr, "NULL argument");
  ensure_init();
  Arg<uint8_t> A8(a8, span(round16(d.n), d.m, lda8), true, false);
  Arg<float> SC(scale, span(nbj, nbi, lds), true, false);
  Arg<float> Y(yst, span(d.Nl, d.m, ldyst), false, true);
  Arg<uint8_t> O(use_seed ? nullptr : omega8, use_seed ? 0 : span(round16(d.n), d.Nl, ldo), true, false);
  DBuf<uint8_t> om;
  const uint8_t *omd = O.dev;
  if (use_seed) {
    ldo = (d.n + 15) / 16 * 16;
    om.alloc(ldo * d.Nl);
    gaussian_omega_e(d.n, d.Nl, om.get(), ldo, seed);
    omd = om.get();
  }
  stage_sketch_e(d, A8.dev, lda8, SC.dev, lds, omd, ldo, Y.dev, ldyst);
  Y.finish();
  sync();
}

static void quantize_impl(char ordera, i64 m, i64 n, const float *a, i64 lda, uint8_t *a8,
                          i64 lda8, float *scale, i64 lds) {
  const char ordera_ = arg2char("ORDERA", ordera, "CR");
  if (!ordera_) return;
  ISVD_CHECK(m > 0 && n > 0, "invalid sizes (m=%lld n=%lld)", m, n);
  ISVD_CHECK(a != nullptr && a8 != nullptr && scale != nullptr, "NULL argument");
  ensure_init();
  const i64 nbi = (m + kF8Block - 1) / kF8Block, nbj = (n + kF8Block - 1) / kF8Block;
  ISVD_CHECK(lds >= nbj, "ldscale (%lld) must be >= ceil(n / 128) (%lld)", lds, nbj);
  Arg<float> A(a, ordera_ == 'C' ? span(m, n, lda) : span(n, m, lda), true, false);
  Arg<uint8_t> A8(a8, span(round16(n), m, lda8), false, true);
  Arg<float> SC(scale, span(nbj, nbi, lds), false, true);
  quantize_e4m3(ordera_, m, n, A.dev, lda, A8.dev, lda8, SC.dev, lds);
  A8.finish();
  SC.finish();
  sync();
}

}  // namespace isvd

using namespace isvd;

// =============================================================================================
// extern "C"
// =============================================================================================
extern "C" {

const char *isvd_GetLastError(void) { return g_last_error.c_str(); }
int isvd_GetErrorCount(void) { return g_error_count; }

isvd_Param isvd_createParam(const isvd_int_t nrow, const isvd_int_t ncol, const isvd_int_t rank,
                            const isvd_int_t over_rank, const isvd_int_t num_sketch,
                            const mpi_int_t mpi_root, const isvd_MpiComm mpi_comm) {
  // src/lib/libisvd/core/param.c:37-80 with P = 1 (the m <= n assert of :43 is not applied)
  if (!(nrow > 0 && ncol > 0 && rank > 0 && over_rank >= 0 && num_sketch > 0)) {
    std::fprintf(stderr, "isvd: isvd_createParam: invalid sizes (m=%d n=%d k=%d p=%d N=%d)\n",
                 (int)nrow, (int)ncol, (int)rank, (int)over_rank, (int)num_sketch);
    g_last_error = "isvd_createParam: invalid sizes";
    ++g_error_count;
    const char *ab = std::getenv("ISVD_ASSERT_ABORT");
    if (ab && ab[0] == '1') std::abort();
  }
  isvd_Param p = {mpi_comm, 1, mpi_root, 0, nrow, ncol, nrow, ncol, nrow, ncol, nrow, ncol, 0,
                  nrow, 0, ncol, rank, over_rank, rank + over_rank,
                  (rank + over_rank) * num_sketch, num_sketch};
  return p;
}

void isvd_Init_cpu(int *, char ***, const isvd_MpiComm) {}
void isvd_Finalize_cpu(void) {}
void isvd_PrintEnvironment_cpu(const isvd_MpiComm) {
  std::printf("iSVD (MI355X native) -- one process, index width %d bits\n",
              (int)(8 * sizeof(isvd_int_t)));
}

void isvd_Init_gpu(int *, char ***, const isvd_MpiComm) {
  guarded([] { ensure_init(); });
}

void isvd_Finalize_gpu(void) {
  guarded([] {
    if (g_init) {
      // everything queued on either stream is done before anything is released; events and
      // pinned memory go before the streams they were recorded or copied on
      HIP_CHECK(hipStreamSynchronize(g_ctx.aux));
      HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
      if (g_ws) HIP_CHECK(hipFreeAsync(g_ws, g_ctx.stream));
      HIP_CHECK(hipStreamSynchronize(g_ctx.stream));
      g_ws = nullptr;
      g_ws_bytes = 0;
      release_pinned();
      HIP_CHECK(hipEventDestroy(g_ctx.fork));
      HIP_CHECK(hipEventDestroy(g_ctx.join));
      HIP_CHECK(hipStreamDestroy(g_ctx.aux));
      HIP_CHECK(hipStreamDestroy(g_ctx.stream));
      g_ctx.stream = nullptr;
      g_ctx.aux = nullptr;
      g_ctx.fork = g_ctx.join = nullptr;
      hipMemPool_t pool = nullptr;
      HIP_CHECK(hipDeviceGetDefaultMemPool(&pool, 0));
      HIP_CHECK(hipMemPoolTrimTo(pool, 0));
      HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &g_pool_threshold));
      g_init = false;
    }
  });
}

void isvd_PrintEnvironment_gpu(const isvd_MpiComm) {
  guarded([] {
    ensure_init();
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, g_ctx.device));
    std::printf("GPU %d: %s (%s), %d CUs, %.1f GiB HBM, LDS %zu KiB per workgroup\n",
                g_ctx.device, prop.name, prop.gcnArchName, prop.multiProcessorCount,
                prop.totalGlobalMem / 1073741824.0, prop.sharedMemPerBlock / 1024);
  });
}

isvd_s_val_t *isvd_sMalloc(const size_t num) {
  void *p = nullptr;
  if (posix_memalign(&p, 64, num * sizeof(isvd_s_val_t) + 1) != 0) return nullptr;
  return static_cast<isvd_s_val_t *>(p);
}
isvd_d_val_t *isvd_dMalloc(const size_t num) {
  void *p = nullptr;
  if (posix_memalign(&p, 64, num * sizeof(isvd_d_val_t) + 1) != 0) return nullptr;
  return static_cast<isvd_d_val_t *>(p);
}
void isvd_Free(void *ptr) { std::free(ptr); }

#define ISVD_DEFINE(x, P, T, TA)                                                               \
  void isvd_##x##Isvd(const isvd_Param param, const char *alg_s, const char *alg_o,           \
                      const char *alg_i, const char *alg_p, const T *argv[4],                 \
                      const isvd_int_t argc[4], T *retv[4], const isvd_int_t retc[4],         \
                      double time[4], FILE *stream, const char dista, const char ordera,      \
                      const TA *a, const isvd_int_t lda, T *s, T *ut, const isvd_int_t ldut,   \
                      T *vt, const isvd_int_t ldvt, const isvd_int_t seed,                    \
                      const mpi_int_t ut_root, const mpi_int_t vt_root) {                     \
    guarded([&] {                                                                             \
      isvd_impl<P>(param, alg_s, alg_o, alg_i, alg_p, argv, argc, retv, retc, time, stream,   \
                   dista, ordera, a, lda, nullptr, 0, s, ut, ldut, vt, ldvt, true,            \
                   (unsigned long long)(long long)seed, ut_root, vt_root);                    \
    });                                                                                       \
  }                                                                                           \
  void isvd_##x##IsvdOmega(const isvd_Param param, const char *alg_s, const char *alg_o,      \
                           const char *alg_i, const char *alg_p, const T *argv[4],            \
                           const isvd_int_t argc[4], T *retv[4], const isvd_int_t retc[4],    \
                           double time[4], FILE *stream, const char dista, const char ordera, \
                           const TA *a, const isvd_int_t lda, const TA *omegat,                 \
                           const isvd_int_t ldomegat, T *s, T *ut, const isvd_int_t ldut,     \
                           T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,             \
                           const mpi_int_t vt_root) {                                         \
    guarded([&] {                                                                             \
      isvd_impl<P>(param, alg_s, alg_o, alg_i, alg_p, argv, argc, retv, retc, time, stream,   \
                   dista, ordera, a, lda, omegat, ldomegat, s, ut, ldut, vt, ldvt, false, 0,  \
                   ut_root, vt_root);                                                         \
    });                                                                                       \
  }                                                                                           \
  void isvd_##x##SketchGaussianProjection(                                                     \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const TA *a,                 \
      const isvd_int_t lda, T *yst, const isvd_int_t ldyst, const isvd_int_t seed,            \
      const mpi_int_t mpi_root) {                                                             \
    (void)argv; (void)retv; (void)retc; (void)mpi_root;                                       \
    if (argc < 0) return;                                                                     \
    guarded([&] {                                                                             \
      sketch_impl<P>(param, dista, ordera, a, lda, nullptr, 0, yst, ldyst, true,              \
                     (unsigned long long)(long long)seed);                                    \
    });                                                                                       \
  }                                                                                           \
  void isvd_##x##SketchGaussianProjectionOmega(                                                \
      const isvd_Param param, const char dista, const char ordera, const TA *a,                \
      const isvd_int_t lda, const TA *omegat, const isvd_int_t ldomegat, T *yst,               \
      const isvd_int_t ldyst) {                                                               \
    guarded([&] {                                                                             \
      sketch_impl<P>(param, dista, ordera, a, lda, omegat, ldomegat, yst, ldyst, false, 0);   \
    });                                                                                       \
  }                                                                                           \
  void isvd_##x##OrthogonalizeTallSkinnyQr(const isvd_Param param, const T *argv,             \
                                           const isvd_int_t argc, T *retv,                    \
                                           const isvd_int_t retc, T *yst,                     \
                                           const isvd_int_t ldyst) {                          \
    (void)argv; (void)retv; (void)retc;                                                       \
    guarded([&] { orth_impl<P>(param, argc, yst, ldyst, true); });                            \
  }                                                                                           \
  void isvd_##x##OrthogonalizeGramian(const isvd_Param param, const T *argv,                  \
                                      const isvd_int_t argc, T *retv, const isvd_int_t retc,  \
                                      T *yst, const isvd_int_t ldyst) {                       \
    (void)argv; (void)retv; (void)retc;                                                       \
    guarded([&] { orth_impl<P>(param, argc, yst, ldyst, false); });                           \
  }                                                                                           \
  void isvd_##x##IntegrateKolmogorovNagumo(const isvd_Param param, const T *argv,             \
                                           const isvd_int_t argc, T *retv,                    \
                                           const isvd_int_t retc, const T *yst,               \
                                           const isvd_int_t ldyst, T *qt,                     \
                                           const isvd_int_t ldqt) {                           \
    guarded([&] { kn_impl<P>(param, argv, argc, retv, retc, yst, ldyst, qt, ldqt); });        \
  }                                                                                           \
  void isvd_##x##IntegrateWenYin(const isvd_Param param, const T *argv, const isvd_int_t argc, \
                                 T *retv, const isvd_int_t retc, const T *yst,                \
                                 const isvd_int_t ldyst, T *qt, const isvd_int_t ldqt) {      \
    guarded([&] { wy_impl<P>(param, argv, argc, retv, retc, yst, ldyst, qt, ldqt); });        \
  }                                                                                           \
  void isvd_##x##PostprocessTallSkinnyQr(                                                      \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const TA *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root) {                                                              \
    (void)argv; (void)retv; (void)retc;                                                       \
    guarded([&] {                                                                             \
      post_impl<P>(param, argc, dista, ordera, a, lda, qt, ldqt, s, ut, ldut, vt, ldvt,       \
                   ut_root, vt_root, 'T');                                                    \
    });                                                                                       \
  }                                                                                           \
  void isvd_##x##PostprocessGramian(                                                           \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const TA *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root) {                                                              \
    (void)argv; (void)retv; (void)retc;                                                       \
    guarded([&] {                                                                             \
      post_impl<P>(param, argc, dista, ordera, a, lda, qt, ldqt, s, ut, ldut, vt, ldvt,       \
                   ut_root, vt_root, 'G');                                                    \
    });                                                                                       \
  }

#define ISVD_DEFINE_BACKWARD(x, P, T)                                                          \
  void isvd_##x##IsvdBackward(const isvd_Param param, const char *alg_p, const char dista,     \
                              const char ordera, const T *a, const isvd_int_t lda,             \
                              const T *qt, const isvd_int_t ldqt, const T *gs, const T *gut,   \
                              const isvd_int_t ldgut, const T *gvt, const isvd_int_t ldgvt,    \
                              T *da, const isvd_int_t ldda) {                                  \
    guarded([&] {                                                                             \
      backward_impl<P>(param, alg_p, dista, ordera, a, lda, qt, ldqt, gs, gut, ldgut, gvt,    \
                       ldgvt, da, ldda);                                                      \
    });                                                                                       \
  }

#define ISVD_DEFINE_HR_SY(x, P, T)                                                             \
  void isvd_##x##IntegrateHierarchicalReduction(const isvd_Param param, const T *argv,         \
                                                const isvd_int_t argc, T *retv,                \
                                                const isvd_int_t retc, T *yst,                 \
                                                const isvd_int_t ldyst, T *qt,                 \
                                                const isvd_int_t ldqt) {                       \
    (void)argv; (void)retv; (void)retc;                                                       \
    guarded([&] { hr_impl<P>(param, argc, yst, ldyst, qt, ldqt); });                          \
  }                                                                                           \
  void isvd_##x##PostprocessSymmetric(                                                         \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root) {                                                              \
    (void)argv; (void)retv; (void)retc;                                                       \
    guarded([&] {                                                                             \
      post_impl<P>(param, argc, dista, ordera, a, lda, qt, ldqt, s, ut, ldut, vt, ldvt,       \
                   ut_root, vt_root, 'S');                                                    \
    });                                                                                       \
  }                                                                                           \
  void isvd_##x##PostprocessSymmetric_gpu(                                                     \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root) {                                                              \
    isvd_##x##PostprocessSymmetric(param, argv, argc, retv, retc, dista, ordera, a, lda, qt,  \
                                   ldqt, s, ut, ldut, vt, ldvt, ut_root, vt_root);            \
  }

#define ISVD_DEFINE_GPU_TWINS(x, T)                                                            \
  void isvd_##x##SketchGaussianProjection_gpu(                                                 \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, T *yst, const isvd_int_t ldyst, const isvd_int_t seed,            \
      const mpi_int_t mpi_root) {                                                             \
    isvd_##x##SketchGaussianProjection(param, argv, argc, retv, retc, dista, ordera, a, lda,  \
                                       yst, ldyst, seed, mpi_root);                           \
  }                                                                                           \
  void isvd_##x##PostprocessTallSkinnyQr_gpu(                                                  \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root) {                                                              \
    isvd_##x##PostprocessTallSkinnyQr(param, argv, argc, retv, retc, dista, ordera, a, lda,   \
                                      qt, ldqt, s, ut, ldut, vt, ldvt, ut_root, vt_root);     \
  }                                                                                           \
  void isvd_##x##PostprocessGramian_gpu(                                                       \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root) {                                                              \
    isvd_##x##PostprocessGramian(param, argv, argc, retv, retc, dista, ordera, a, lda, qt,    \
                                 ldqt, s, ut, ldut, vt, ldvt, ut_root, vt_root);              \
  }


void isvd_dPostprocessGramianNoAlloc(const isvd_Param param, const isvd_d_val_t *argv,
                                     const isvd_int_t argc, isvd_d_val_t *retv,
                                     const isvd_int_t retc, const char dista, const char ordera,
                                     const isvd_d_val_t *a, const isvd_int_t lda,
                                     const isvd_d_val_t *qt, const isvd_int_t ldqt,
                                     isvd_d_val_t *s, isvd_d_val_t *ut, const isvd_int_t ldut,
                                     isvd_d_val_t *vt, const isvd_int_t ldvt,
                                     const mpi_int_t ut_root, const mpi_int_t vt_root) {
  (void)argv; (void)retv; (void)retc;
  guarded([&] {
    post_gramian_noalloc_impl(param, argc, dista, ordera, a, lda, qt, ldqt, s, ut, ldut, vt,
                              ldvt, ut_root, vt_root);
  });
}

ISVD_DEFINE(s, PrecS, isvd_s_val_t, isvd_s_val_t)
ISVD_DEFINE(d, PrecD, isvd_d_val_t, isvd_d_val_t)
ISVD_DEFINE(b, PrecB, float, isvd_b_val_t)
ISVD_DEFINE_GPU_TWINS(s, isvd_s_val_t)
ISVD_DEFINE_GPU_TWINS(d, isvd_d_val_t)
ISVD_DEFINE_HR_SY(s, PrecS, isvd_s_val_t)
ISVD_DEFINE_HR_SY(d, PrecD, isvd_d_val_t)
ISVD_DEFINE_BACKWARD(s, PrecS, isvd_s_val_t)
ISVD_DEFINE_BACKWARD(d, PrecD, isvd_d_val_t)

void isvd_eQuantize(const char ordera, const isvd_int_t m, const isvd_int_t n, const float *a,
                    const isvd_int_t lda, isvd_e_val_t *a8, const isvd_int_t lda8, float *scale,
                    const isvd_int_t ldscale) {
  guarded([&] { quantize_impl(ordera, m, n, a, lda, a8, lda8, scale, ldscale); });
}

void isvd_eSketchGaussianProjection(const isvd_Param param, const float *argv,
                                    const isvd_int_t argc, float *retv, const isvd_int_t retc,
                                    const char dista, const isvd_e_val_t *a8,
                                    const isvd_int_t lda8, const float *scale,
                                    const isvd_int_t ldscale, float *yst,
                                    const isvd_int_t ldyst, const isvd_int_t seed,
                                    const mpi_int_t mpi_root) {
  (void)argv; (void)retv; (void)retc; (void)mpi_root;
  if (argc < 0) return;
  guarded([&] {
    sketch_e_impl(param, dista, a8, lda8, scale, ldscale, nullptr, 0, yst, ldyst, true,
                  (unsigned long long)(long long)seed);
  });
}

void isvd_eSketchGaussianProjectionOmega(const isvd_Param param, const char dista,
                                         const isvd_e_val_t *a8, const isvd_int_t lda8,
                                         const float *scale, const isvd_int_t ldscale,
                                         const isvd_e_val_t *omega, const isvd_int_t ldomega,
                                         float *yst, const isvd_int_t ldyst) {
  guarded([&] {
    sketch_e_impl(param, dista, a8, lda8, scale, ldscale, omega, ldomega, yst, ldyst, false, 0);
  });
}

void isvd_eIsvd(const isvd_Param param, const char *alg_s, const char *alg_o, const char *alg_i,
                const char *alg_p, const float *argv[4], const isvd_int_t argc[4],
                float *retv[4], const isvd_int_t retc[4], double time[4], FILE *stream,
                const char dista, const isvd_e_val_t *a8, const isvd_int_t lda8,
                const float *scale, const isvd_int_t ldscale, const char ordera,
                const isvd_b_val_t *a, const isvd_int_t lda, float *s, float *ut,
                const isvd_int_t ldut, float *vt, const isvd_int_t ldvt, const isvd_int_t seed,
                const mpi_int_t ut_root, const mpi_int_t vt_root) {
  guarded([&] {
    const E8Input e8{a8, lda8, scale, ldscale};
    isvd_impl<PrecB>(param, alg_s, alg_o, alg_i, alg_p, argv, argc, retv, retc, time, stream,
                     dista, ordera, a, lda, nullptr, 0, s, ut, ldut, vt, ldvt, true,
                     (unsigned long long)(long long)seed, ut_root, vt_root, &e8);
  });
}

}  // extern "C"
