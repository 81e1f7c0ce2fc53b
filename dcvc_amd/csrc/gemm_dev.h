// Device helpers shared by the MFMA GEMM kernels (gemm.hip, ffn.hip):
// direct global->LDS loads, the LDS row swizzle, the K-tile barrier and the
// instruction interleave pattern of the K loop.
#pragma once

#include <type_traits>
#include <utility>

#include "common.h"

namespace dcvc {
namespace {

constexpr int kBK = 32;                 // K tile (halves): one 64-byte LDS row
constexpr int kRowBytes = kBK * 2;
constexpr uint32_t kOOB = 0x80000000u;  // buffer offset beyond every range: reads zero
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format

DCVC_DEV __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kRsrcWord3);
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// 16 bytes per lane from buffer offset `off` straight into LDS at
// dst + 16 * lane (dst wave-uniform).
DCVC_DEV void dma16(__amdgpu_buffer_rsrc_t r, char* dst, uint32_t off) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_ptr_t)dst, 16, off, 0, 0, 0);
}

// XOR swizzle of the 16-byte K chunk within a 64-byte LDS row.
DCVC_DEV int swz(int row) { return ((row >> 2) ^ (row >> 1)) & 3; }

// The same for an LDS row of RB bytes (64: 4 chunks, swz; 128: 8 chunks).  A
// 128-byte row spans 32 of the 64 banks, so the 16 rows one quarter-wave
// fragment read touches are spread by giving each row pair its own chunk
// permutation: bank group = (row & 1) * 8 + (chunk ^ ((row >> 1) & 7)).
template <int RB>
DCVC_DEV int swz_rows(int row) {
    if constexpr (RB == 64)
        return swz(row);
    else
        return (row >> 1) & 7;
}

// s_waitcnt vmcnt(VM) [lgkmcnt(0) if LGKM] (gfx9 encoding), then a workgroup
// barrier.  The compiler fences keep LDS accesses on their side of the
// barrier; no memory-model fence is used because it would also wait for the
// K tiles still in flight.
template <int VM, bool LGKM>
DCVC_DEV void wait_tile_and_barrier() {
    static_assert(VM >= 0 && VM < 64, "vmcnt range");
    constexpr int imm = (VM & 15) | ((VM >> 4) << 14) | (7 << 4) | ((LGKM ? 0 : 15) << 8);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(imm);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// f(std::integral_constant<int, I>) for I in the sequence, in order (a loop
// whose index is a template argument).
template <int... I, class F>
DCVC_DEV void static_for(std::integer_sequence<int, I...>, F&& f) {
    (f(std::integral_constant<int, I>{}), ...);
}

// Scheduling pattern for one k16 step: NDS fragment reads, NMF MFMAs and NVM
// DMA issues spread evenly, so the matrix pipe never waits behind a burst of
// memory instructions.
template <int NDS, int NMF, int NVM, int R = 0>
DCVC_DEV void interleave() {
    if constexpr (R < NDS) {
        constexpr int vm = (NVM * (R + 1)) / NDS - (NVM * R) / NDS;
        constexpr int mf = (NMF * (R + 1)) / NDS - (NMF * R) / NDS;
        if constexpr (vm > 0) __builtin_amdgcn_sched_group_barrier(0x020, vm, 0);  // VMEM read
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                        // DS read
        if constexpr (mf > 0) __builtin_amdgcn_sched_group_barrier(0x008, mf, 0);  // MFMA
        interleave<NDS, NMF, NVM, R + 1>();
    }
}

}  // namespace
}  // namespace dcvc
