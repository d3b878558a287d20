// bkx_device.hpp -- device-side helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace bkx {

constexpr int kWave = 64;   // CDNA wavefront width

// Native 16-byte vectors (the nontemporal builtins take clang vector types).
using f64x2 = double __attribute__((ext_vector_type(2)));
using f32x4 = float __attribute__((ext_vector_type(4)));
using f32x2 = float __attribute__((ext_vector_type(2)));

// Orders LDS traffic between the lanes of one wavefront.  A wavefront's LDS
// instructions are executed in issue order, so a lane may read what another
// lane of the same wavefront wrote before this point without an s_barrier;
// the wavefront-scope fences keep the compiler from moving LDS accesses
// across it.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Keeps the computation of x at this point of the program: an empty asm
// that reads and writes x in its VGPRs, so the arithmetic producing x cannot
// be sunk past it (which would keep every layer's operands live at once).
__device__ __forceinline__ void pin(double& x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void pin(float& x) { asm volatile("" : "+v"(x)); }

// The same uniform pointer, opaque to the optimiser: loads through it in
// different program phases are not merged, so a small wavefront-uniform
// matrix is re-read from the scalar cache per phase instead of being kept
// live (in scalar, then vector registers) across the whole kernel.
// The asm is not volatile and has no memory clobber, so it does not count as
// a store that could clobber the loads after it (which must stay scalar
// loads); TAG makes each phase's asm text distinct so that two launders of
// the same pointer are not merged.  An inline-asm result counts as
// divergent, so the pointer is made provably uniform again with
// readfirstlane, and it is returned in the constant address space: the
// matrices are never written by a kernel, and loads from that space stay
// scalar loads even after stores in a loop (through a global pointer they
// would need a proof that no earlier store clobbers them, and through a
// generic pointer they would be vector flat loads).
template <typename T>
using global_ptr = const T __attribute__((address_space(4)))*;

// dep (a wavefront-uniform value) makes the launder depend on it, so that
// in a loop the loads are not hoisted out when dep is the loop counter.
template <int TAG, typename T>
__device__ __forceinline__ global_ptr<T> launder(const T* p, int dep = 0)
{
    asm("; launder %1 %2" : "+s"(p) : "i"(TAG), "s"(dep));
    const uint64_t x = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
    return (global_ptr<T>)(reinterpret_cast<const T*>((uint64_t(hi) << 32) | lo));
}

// Streaming loads and stores: every geometric factor and input value is used
// by one element only, so they are marked non-temporal.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p)
{
    return __builtin_nontemporal_load(p);
}

// A load that is non-temporal when NT is set.
template <bool NT, typename T>
__device__ __forceinline__ T load(const T* p)
{
    if constexpr (NT) {
        return __builtin_nontemporal_load(p);
    } else {
        return *p;
    }
}

template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v)
{
    __builtin_nontemporal_store(v, p);
}

}  // namespace bkx
