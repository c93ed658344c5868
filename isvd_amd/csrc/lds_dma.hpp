// Direct global -> LDS fills (global_load_lds_dwordx4) shared by the MFMA kernels that stage
// their operands through LDS in a multi-stage pipeline.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace isvd {

// s_waitcnt immediate (gfx9 encoding): at most v vector-memory operations outstanding,
// expcnt and lgkmcnt unconstrained
constexpr int waitcnt_vm(int v) { return (v & 15) | (7 << 4) | (15 << 8) | ((v >> 4) << 14); }
// the same, and no LDS (or scalar-memory) operation outstanding
constexpr int waitcnt_vm_lgkm0(int v) { return (v & 15) | (7 << 4) | ((v >> 4) << 14); }

// One 16-byte-per-lane fill of LDS from global memory (global_load_lds_dwordx4): lane l's 16
// bytes land at lds_base + 16 l.  Issued as inline assembly so that the compiler does not
// track it: its automatic waits would otherwise drain every fill in flight before each LDS
// read.  The kernels wait on the vector-memory counter themselves before a stage is read.
// M0 holds the LDS base; one wait state must separate its write from the load.
// M0 is named as clobbered; no other code in the kernels that use this reads M0 (the
// compiler's only M0 uses are LDS DMA, messages and relative indexing, none of which occur).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ inline void lds_fill16(const void *src, unsigned lds_base) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(src), "s"(lds_base)
               : "memory", "m0");
}
#pragma clang diagnostic pop

// LDS byte address of a __shared__ object (wave-uniform)
__device__ inline unsigned lds_addr(const void *p) {
  return __builtin_amdgcn_readfirstlane(
      (unsigned)(uintptr_t)(const __attribute__((address_space(3))) unsigned char *)(
          static_cast<const unsigned char *>(p)));
}

// Workgroup barrier without the compiler's implied waits: every wave has waited for what it
// needs itself (the fills it issued, counted on the vector-memory counter).
__device__ inline void bare_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

}  // namespace isvd
