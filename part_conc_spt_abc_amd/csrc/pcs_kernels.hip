// pcs_kernels.hip -- gfx950 kernels of the simulation hot path.
//
//   walk_kernel      one wave = one tile of 64*PPL particles of one video;
//                    1 lane = PPL particles with all state in VGPRs. Replaces
//                    K_simulate_system_unr (src/gpu/cusim/kernel-simulate-unroll.inl:69-298),
//                    where one 32-wide warp looped over all particles of a video.
//                    Closed tracks are staged raw, (k, de), in a per-wave LDS
//                    buffer, binned (row<<16 | bin) when the buffer is flushed,
//                    and appended to the proposal's HBM record slice with one
//                    atomic per flush (replaces the per-track global atomicInc
//                    of src/gpu/cusim/histogram2d.inl:98-106).
//   distance_kernel  one workgroup per proposal: counting sort of the records
//                    by row (LDS row counters; the records are read once and
//                    held in registers, the sorted 16-bit bins kept in LDS
//                    when they fit), then a
//                    row sweep over the 2D cumulative histogram with the
//                    running column sums kept in registers: a row's records
//                    enter through an LDS histogram and block scan, or, when
//                    few, straight into each thread's columns; exact integer
//                    sum of (S-R)^2 or |S-R| against the observed summed-area
//                    table, whose rows are loaded ahead. Replaces K_distance
//                    (src/gpu/cusim/kernel-distance.inl:81-241), a single 32x32
//                    block that scanned the whole dense histogram in FP32.
//   serve_loop_kernel  device-resident request loop: one admitter wave copies
//                    requests from a ring in pinned host memory into HBM slots,
//                    the other groups take their tiles, walk them and run the
//                    distance of a request once its last tile is done.
//   pmc_sums_kernel  ABC-PMC importance-weight sums in FP64
//                    (src/gpu/detail/simulation_impl.inl:832-871).
//   dense_hist       scatter of records into dense histograms (tests, export).
//   reset_kernel     zeroes per-proposal counters (a kernel rather than a
//                    memset so it can sit inside a replayed hipGraph).
//   selftest_*       exhaustive checks of the device math and a copy kernel
//                    that calibrates the byte counters (test hooks).
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "pcs_device.h"
#include "pcs_internal.h"

namespace pcs {

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Global-memory view of a pointer whose address space the compiler cannot
// infer (one read out of memory): accesses through it compile to global
// instructions, counted by the vector-memory counter alone, so the compiler
// can wait for one load at a time. A flat access might touch LDS and complete
// out of order, and makes it wait for every outstanding access.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gview(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gview(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

// Pointers typed in the LDS (3) and global (1) address spaces, for the
// arguments of a function of its own: a generic pointer argument compiles to
// flat accesses (which take the vector-memory path to LDS as well and are
// waited for on both counters); one converted from a typed argument inside
// the function compiles to LDS or global instructions, also in the functions
// inlined into it.
typedef const __attribute__((address_space(3))) float2* lds_cf2;
typedef __attribute__((address_space(3))) uint32_t* lds_u32;
typedef __attribute__((address_space(1))) uint32_t* glb_u32;

// Orders this wave's LDS accesses across lanes (staging buffer handoff).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------ angle table

// The Box-Muller angle of code i << 8 | r is the sector centre 2 pi (i + 1/2) / 256
// plus beta = 2 pi (r - 127.5) / 2^16. kSinCosTab holds (sin, cos) of the sector
// centres, kBetaTab (sin beta, cos beta - 1); every entry is correctly rounded
// to FP32 (fit_math_coeffs.py). Every workgroup that walks copies both into
// LDS, kBetaTab right after kSinCosTab.
constexpr int kSinCosTabFloats = 512;
constexpr int kAngleTabFloats = 2 * kSinCosTabFloats;
__constant__ float kSinCosTab[kSinCosTabFloats] = {
    0x1.921d2p-7f, 0x1.fff622p-1f, 0x1.2d8658p-5f, 0x1.ffa72ep-1f, 0x1.f656e8p-5f, 0x1.ff0956p-1f, 0x1.5f6dp-4f, 0x1.fe1cbp-1f,
    0x1.c3785cp-4f, 0x1.fce16p-1f, 0x1.139f0cp-3f, 0x1.fb5798p-1f, 0x1.45576cp-3f, 0x1.f97f92p-1f, 0x1.76dd9ep-3f, 0x1.f7599ap-1f,
    0x1.a82a02p-3f, 0x1.f4e604p-1f, 0x1.d934fep-3f, 0x1.f2253p-1f, 0x1.04fb8p-2f, 0x1.ef178ap-1f, 0x1.1d3444p-2f, 0x1.ebbd8cp-1f,
    0x1.35410cp-2f, 0x1.e817bap-1f, 0x1.4d1e24p-2f, 0x1.e426a4p-1f, 0x1.64c7dep-2f, 0x1.dfeae6p-1f, 0x1.7c3a94p-2f, 0x1.db6526p-1f,
    0x1.9372a6p-2f, 0x1.d69618p-1f, 0x1.aa6c82p-2f, 0x1.d17e78p-1f, 0x1.c1249ep-2f, 0x1.cc1f1p-1f, 0x1.d79776p-2f, 0x1.c678b4p-1f,
    0x1.edc196p-2f, 0x1.c08c42p-1f, 0x1.01cfc8p-1f, 0x1.ba5aa6p-1f, 0x1.0c9704p-1f, 0x1.b3e4d4p-1f, 0x1.1734d6p-1f, 0x1.ad2bcap-1f,
    0x1.21a79ap-1f, 0x1.a63092p-1f, 0x1.2bedb2p-1f, 0x1.9ef43ep-1f, 0x1.36058cp-1f, 0x1.9777fp-1f, 0x1.3fed96p-1f, 0x1.8fbccap-1f,
    0x1.49a44ap-1f, 0x1.87c4p-1f, 0x1.53282ap-1f, 0x1.7f8ecep-1f, 0x1.5c77bcp-1f, 0x1.771e76p-1f, 0x1.659192p-1f, 0x1.6e7446p-1f,
    0x1.6e7446p-1f, 0x1.659192p-1f, 0x1.771e76p-1f, 0x1.5c77bcp-1f, 0x1.7f8ecep-1f, 0x1.53282ap-1f, 0x1.87c4p-1f, 0x1.49a44ap-1f,
    0x1.8fbccap-1f, 0x1.3fed96p-1f, 0x1.9777fp-1f, 0x1.36058cp-1f, 0x1.9ef43ep-1f, 0x1.2bedb2p-1f, 0x1.a63092p-1f, 0x1.21a79ap-1f,
    0x1.ad2bcap-1f, 0x1.1734d6p-1f, 0x1.b3e4d4p-1f, 0x1.0c9704p-1f, 0x1.ba5aa6p-1f, 0x1.01cfc8p-1f, 0x1.c08c42p-1f, 0x1.edc196p-2f,
    0x1.c678b4p-1f, 0x1.d79776p-2f, 0x1.cc1f1p-1f, 0x1.c1249ep-2f, 0x1.d17e78p-1f, 0x1.aa6c82p-2f, 0x1.d69618p-1f, 0x1.9372a6p-2f,
    0x1.db6526p-1f, 0x1.7c3a94p-2f, 0x1.dfeae6p-1f, 0x1.64c7dep-2f, 0x1.e426a4p-1f, 0x1.4d1e24p-2f, 0x1.e817bap-1f, 0x1.35410cp-2f,
    0x1.ebbd8cp-1f, 0x1.1d3444p-2f, 0x1.ef178ap-1f, 0x1.04fb8p-2f, 0x1.f2253p-1f, 0x1.d934fep-3f, 0x1.f4e604p-1f, 0x1.a82a02p-3f,
    0x1.f7599ap-1f, 0x1.76dd9ep-3f, 0x1.f97f92p-1f, 0x1.45576cp-3f, 0x1.fb5798p-1f, 0x1.139f0cp-3f, 0x1.fce16p-1f, 0x1.c3785cp-4f,
    0x1.fe1cbp-1f, 0x1.5f6dp-4f, 0x1.ff0956p-1f, 0x1.f656e8p-5f, 0x1.ffa72ep-1f, 0x1.2d8658p-5f, 0x1.fff622p-1f, 0x1.921d2p-7f,
    0x1.fff622p-1f, -0x1.921d2p-7f, 0x1.ffa72ep-1f, -0x1.2d8658p-5f, 0x1.ff0956p-1f, -0x1.f656e8p-5f, 0x1.fe1cbp-1f, -0x1.5f6dp-4f,
    0x1.fce16p-1f, -0x1.c3785cp-4f, 0x1.fb5798p-1f, -0x1.139f0cp-3f, 0x1.f97f92p-1f, -0x1.45576cp-3f, 0x1.f7599ap-1f, -0x1.76dd9ep-3f,
    0x1.f4e604p-1f, -0x1.a82a02p-3f, 0x1.f2253p-1f, -0x1.d934fep-3f, 0x1.ef178ap-1f, -0x1.04fb8p-2f, 0x1.ebbd8cp-1f, -0x1.1d3444p-2f,
    0x1.e817bap-1f, -0x1.35410cp-2f, 0x1.e426a4p-1f, -0x1.4d1e24p-2f, 0x1.dfeae6p-1f, -0x1.64c7dep-2f, 0x1.db6526p-1f, -0x1.7c3a94p-2f,
    0x1.d69618p-1f, -0x1.9372a6p-2f, 0x1.d17e78p-1f, -0x1.aa6c82p-2f, 0x1.cc1f1p-1f, -0x1.c1249ep-2f, 0x1.c678b4p-1f, -0x1.d79776p-2f,
    0x1.c08c42p-1f, -0x1.edc196p-2f, 0x1.ba5aa6p-1f, -0x1.01cfc8p-1f, 0x1.b3e4d4p-1f, -0x1.0c9704p-1f, 0x1.ad2bcap-1f, -0x1.1734d6p-1f,
    0x1.a63092p-1f, -0x1.21a79ap-1f, 0x1.9ef43ep-1f, -0x1.2bedb2p-1f, 0x1.9777fp-1f, -0x1.36058cp-1f, 0x1.8fbccap-1f, -0x1.3fed96p-1f,
    0x1.87c4p-1f, -0x1.49a44ap-1f, 0x1.7f8ecep-1f, -0x1.53282ap-1f, 0x1.771e76p-1f, -0x1.5c77bcp-1f, 0x1.6e7446p-1f, -0x1.659192p-1f,
    0x1.659192p-1f, -0x1.6e7446p-1f, 0x1.5c77bcp-1f, -0x1.771e76p-1f, 0x1.53282ap-1f, -0x1.7f8ecep-1f, 0x1.49a44ap-1f, -0x1.87c4p-1f,
    0x1.3fed96p-1f, -0x1.8fbccap-1f, 0x1.36058cp-1f, -0x1.9777fp-1f, 0x1.2bedb2p-1f, -0x1.9ef43ep-1f, 0x1.21a79ap-1f, -0x1.a63092p-1f,
    0x1.1734d6p-1f, -0x1.ad2bcap-1f, 0x1.0c9704p-1f, -0x1.b3e4d4p-1f, 0x1.01cfc8p-1f, -0x1.ba5aa6p-1f, 0x1.edc196p-2f, -0x1.c08c42p-1f,
    0x1.d79776p-2f, -0x1.c678b4p-1f, 0x1.c1249ep-2f, -0x1.cc1f1p-1f, 0x1.aa6c82p-2f, -0x1.d17e78p-1f, 0x1.9372a6p-2f, -0x1.d69618p-1f,
    0x1.7c3a94p-2f, -0x1.db6526p-1f, 0x1.64c7dep-2f, -0x1.dfeae6p-1f, 0x1.4d1e24p-2f, -0x1.e426a4p-1f, 0x1.35410cp-2f, -0x1.e817bap-1f,
    0x1.1d3444p-2f, -0x1.ebbd8cp-1f, 0x1.04fb8p-2f, -0x1.ef178ap-1f, 0x1.d934fep-3f, -0x1.f2253p-1f, 0x1.a82a02p-3f, -0x1.f4e604p-1f,
    0x1.76dd9ep-3f, -0x1.f7599ap-1f, 0x1.45576cp-3f, -0x1.f97f92p-1f, 0x1.139f0cp-3f, -0x1.fb5798p-1f, 0x1.c3785cp-4f, -0x1.fce16p-1f,
    0x1.5f6dp-4f, -0x1.fe1cbp-1f, 0x1.f656e8p-5f, -0x1.ff0956p-1f, 0x1.2d8658p-5f, -0x1.ffa72ep-1f, 0x1.921d2p-7f, -0x1.fff622p-1f,
    -0x1.921d2p-7f, -0x1.fff622p-1f, -0x1.2d8658p-5f, -0x1.ffa72ep-1f, -0x1.f656e8p-5f, -0x1.ff0956p-1f, -0x1.5f6dp-4f, -0x1.fe1cbp-1f,
    -0x1.c3785cp-4f, -0x1.fce16p-1f, -0x1.139f0cp-3f, -0x1.fb5798p-1f, -0x1.45576cp-3f, -0x1.f97f92p-1f, -0x1.76dd9ep-3f, -0x1.f7599ap-1f,
    -0x1.a82a02p-3f, -0x1.f4e604p-1f, -0x1.d934fep-3f, -0x1.f2253p-1f, -0x1.04fb8p-2f, -0x1.ef178ap-1f, -0x1.1d3444p-2f, -0x1.ebbd8cp-1f,
    -0x1.35410cp-2f, -0x1.e817bap-1f, -0x1.4d1e24p-2f, -0x1.e426a4p-1f, -0x1.64c7dep-2f, -0x1.dfeae6p-1f, -0x1.7c3a94p-2f, -0x1.db6526p-1f,
    -0x1.9372a6p-2f, -0x1.d69618p-1f, -0x1.aa6c82p-2f, -0x1.d17e78p-1f, -0x1.c1249ep-2f, -0x1.cc1f1p-1f, -0x1.d79776p-2f, -0x1.c678b4p-1f,
    -0x1.edc196p-2f, -0x1.c08c42p-1f, -0x1.01cfc8p-1f, -0x1.ba5aa6p-1f, -0x1.0c9704p-1f, -0x1.b3e4d4p-1f, -0x1.1734d6p-1f, -0x1.ad2bcap-1f,
    -0x1.21a79ap-1f, -0x1.a63092p-1f, -0x1.2bedb2p-1f, -0x1.9ef43ep-1f, -0x1.36058cp-1f, -0x1.9777fp-1f, -0x1.3fed96p-1f, -0x1.8fbccap-1f,
    -0x1.49a44ap-1f, -0x1.87c4p-1f, -0x1.53282ap-1f, -0x1.7f8ecep-1f, -0x1.5c77bcp-1f, -0x1.771e76p-1f, -0x1.659192p-1f, -0x1.6e7446p-1f,
    -0x1.6e7446p-1f, -0x1.659192p-1f, -0x1.771e76p-1f, -0x1.5c77bcp-1f, -0x1.7f8ecep-1f, -0x1.53282ap-1f, -0x1.87c4p-1f, -0x1.49a44ap-1f,
    -0x1.8fbccap-1f, -0x1.3fed96p-1f, -0x1.9777fp-1f, -0x1.36058cp-1f, -0x1.9ef43ep-1f, -0x1.2bedb2p-1f, -0x1.a63092p-1f, -0x1.21a79ap-1f,
    -0x1.ad2bcap-1f, -0x1.1734d6p-1f, -0x1.b3e4d4p-1f, -0x1.0c9704p-1f, -0x1.ba5aa6p-1f, -0x1.01cfc8p-1f, -0x1.c08c42p-1f, -0x1.edc196p-2f,
    -0x1.c678b4p-1f, -0x1.d79776p-2f, -0x1.cc1f1p-1f, -0x1.c1249ep-2f, -0x1.d17e78p-1f, -0x1.aa6c82p-2f, -0x1.d69618p-1f, -0x1.9372a6p-2f,
    -0x1.db6526p-1f, -0x1.7c3a94p-2f, -0x1.dfeae6p-1f, -0x1.64c7dep-2f, -0x1.e426a4p-1f, -0x1.4d1e24p-2f, -0x1.e817bap-1f, -0x1.35410cp-2f,
    -0x1.ebbd8cp-1f, -0x1.1d3444p-2f, -0x1.ef178ap-1f, -0x1.04fb8p-2f, -0x1.f2253p-1f, -0x1.d934fep-3f, -0x1.f4e604p-1f, -0x1.a82a02p-3f,
    -0x1.f7599ap-1f, -0x1.76dd9ep-3f, -0x1.f97f92p-1f, -0x1.45576cp-3f, -0x1.fb5798p-1f, -0x1.139f0cp-3f, -0x1.fce16p-1f, -0x1.c3785cp-4f,
    -0x1.fe1cbp-1f, -0x1.5f6dp-4f, -0x1.ff0956p-1f, -0x1.f656e8p-5f, -0x1.ffa72ep-1f, -0x1.2d8658p-5f, -0x1.fff622p-1f, -0x1.921d2p-7f,
    -0x1.fff622p-1f, 0x1.921d2p-7f, -0x1.ffa72ep-1f, 0x1.2d8658p-5f, -0x1.ff0956p-1f, 0x1.f656e8p-5f, -0x1.fe1cbp-1f, 0x1.5f6dp-4f,
    -0x1.fce16p-1f, 0x1.c3785cp-4f, -0x1.fb5798p-1f, 0x1.139f0cp-3f, -0x1.f97f92p-1f, 0x1.45576cp-3f, -0x1.f7599ap-1f, 0x1.76dd9ep-3f,
    -0x1.f4e604p-1f, 0x1.a82a02p-3f, -0x1.f2253p-1f, 0x1.d934fep-3f, -0x1.ef178ap-1f, 0x1.04fb8p-2f, -0x1.ebbd8cp-1f, 0x1.1d3444p-2f,
    -0x1.e817bap-1f, 0x1.35410cp-2f, -0x1.e426a4p-1f, 0x1.4d1e24p-2f, -0x1.dfeae6p-1f, 0x1.64c7dep-2f, -0x1.db6526p-1f, 0x1.7c3a94p-2f,
    -0x1.d69618p-1f, 0x1.9372a6p-2f, -0x1.d17e78p-1f, 0x1.aa6c82p-2f, -0x1.cc1f1p-1f, 0x1.c1249ep-2f, -0x1.c678b4p-1f, 0x1.d79776p-2f,
    -0x1.c08c42p-1f, 0x1.edc196p-2f, -0x1.ba5aa6p-1f, 0x1.01cfc8p-1f, -0x1.b3e4d4p-1f, 0x1.0c9704p-1f, -0x1.ad2bcap-1f, 0x1.1734d6p-1f,
    -0x1.a63092p-1f, 0x1.21a79ap-1f, -0x1.9ef43ep-1f, 0x1.2bedb2p-1f, -0x1.9777fp-1f, 0x1.36058cp-1f, -0x1.8fbccap-1f, 0x1.3fed96p-1f,
    -0x1.87c4p-1f, 0x1.49a44ap-1f, -0x1.7f8ecep-1f, 0x1.53282ap-1f, -0x1.771e76p-1f, 0x1.5c77bcp-1f, -0x1.6e7446p-1f, 0x1.659192p-1f,
    -0x1.659192p-1f, 0x1.6e7446p-1f, -0x1.5c77bcp-1f, 0x1.771e76p-1f, -0x1.53282ap-1f, 0x1.7f8ecep-1f, -0x1.49a44ap-1f, 0x1.87c4p-1f,
    -0x1.3fed96p-1f, 0x1.8fbccap-1f, -0x1.36058cp-1f, 0x1.9777fp-1f, -0x1.2bedb2p-1f, 0x1.9ef43ep-1f, -0x1.21a79ap-1f, 0x1.a63092p-1f,
    -0x1.1734d6p-1f, 0x1.ad2bcap-1f, -0x1.0c9704p-1f, 0x1.b3e4d4p-1f, -0x1.01cfc8p-1f, 0x1.ba5aa6p-1f, -0x1.edc196p-2f, 0x1.c08c42p-1f,
    -0x1.d79776p-2f, 0x1.c678b4p-1f, -0x1.c1249ep-2f, 0x1.cc1f1p-1f, -0x1.aa6c82p-2f, 0x1.d17e78p-1f, -0x1.9372a6p-2f, 0x1.d69618p-1f,
    -0x1.7c3a94p-2f, 0x1.db6526p-1f, -0x1.64c7dep-2f, 0x1.dfeae6p-1f, -0x1.4d1e24p-2f, 0x1.e426a4p-1f, -0x1.35410cp-2f, 0x1.e817bap-1f,
    -0x1.1d3444p-2f, 0x1.ebbd8cp-1f, -0x1.04fb8p-2f, 0x1.ef178ap-1f, -0x1.d934fep-3f, 0x1.f2253p-1f, -0x1.a82a02p-3f, 0x1.f4e604p-1f,
    -0x1.76dd9ep-3f, 0x1.f7599ap-1f, -0x1.45576cp-3f, 0x1.f97f92p-1f, -0x1.139f0cp-3f, 0x1.fb5798p-1f, -0x1.c3785cp-4f, 0x1.fce16p-1f,
    -0x1.5f6dp-4f, 0x1.fe1cbp-1f, -0x1.f656e8p-5f, 0x1.ff0956p-1f, -0x1.2d8658p-5f, 0x1.ffa72ep-1f, -0x1.921d2p-7f, 0x1.fff622p-1f,
};
__constant__ float kBetaTab[kSinCosTabFloats] = {
    -0x1.908b08p-7f, -0x1.395c62p-14f, -0x1.8d66d8p-7f, -0x1.3476f8p-14f, -0x1.8a42a8p-7f, -0x1.2f9b6ep-14f, -0x1.871e76p-7f, -0x1.2ac9c2p-14f,
    -0x1.83fa46p-7f, -0x1.2601f6p-14f, -0x1.80d614p-7f, -0x1.214406p-14f, -0x1.7db1e4p-7f, -0x1.1c8ff6p-14f, -0x1.7a8db2p-7f, -0x1.17e5c4p-14f,
    -0x1.77698p-7f, -0x1.13457p-14f, -0x1.74454ep-7f, -0x1.0eaefcp-14f, -0x1.71211cp-7f, -0x1.0a2264p-14f, -0x1.6dfceap-7f, -0x1.059facp-14f,
    -0x1.6ad8b6p-7f, -0x1.0126d2p-14f, -0x1.67b484p-7f, -0x1.f96faep-15f, -0x1.64905p-7f, -0x1.f0a574p-15f, -0x1.616c1ep-7f, -0x1.e7eef8p-15f,
    -0x1.5e47eap-7f, -0x1.df4c38p-15f, -0x1.5b23b6p-7f, -0x1.d6bd36p-15f, -0x1.57ff82p-7f, -0x1.ce41fp-15f, -0x1.54db4ep-7f, -0x1.c5da68p-15f,
    -0x1.51b71ap-7f, -0x1.bd869cp-15f, -0x1.4e92e4p-7f, -0x1.b5468cp-15f, -0x1.4b6ebp-7f, -0x1.ad1a3ap-15f, -0x1.484a7cp-7f, -0x1.a501a6p-15f,
    -0x1.452646p-7f, -0x1.9cfccep-15f, -0x1.42021p-7f, -0x1.950bb2p-15f, -0x1.3edddap-7f, -0x1.8d2e54p-15f, -0x1.3bb9a6p-7f, -0x1.8564b4p-15f,
    -0x1.38957p-7f, -0x1.7daedp-15f, -0x1.35713ap-7f, -0x1.760ca8p-15f, -0x1.324d02p-7f, -0x1.6e7e3ep-15f, -0x1.2f28ccp-7f, -0x1.670392p-15f,
    -0x1.2c0496p-7f, -0x1.5f9ca2p-15f, -0x1.28e05ep-7f, -0x1.58496ep-15f, -0x1.25bc28p-7f, -0x1.5109f8p-15f, -0x1.2297fp-7f, -0x1.49de4p-15f,
    -0x1.1f73b8p-7f, -0x1.42c644p-15f, -0x1.1c4f82p-7f, -0x1.3bc206p-15f, -0x1.192b4ap-7f, -0x1.34d184p-15f, -0x1.160712p-7f, -0x1.2df4bep-15f,
    -0x1.12e2dap-7f, -0x1.272bb8p-15f, -0x1.0fbea2p-7f, -0x1.20766cp-15f, -0x1.0c9a68p-7f, -0x1.19d4ep-15f, -0x1.09763p-7f, -0x1.13470ep-15f,
    -0x1.0651f8p-7f, -0x1.0cccfcp-15f, -0x1.032dbep-7f, -0x1.0666a6p-15f, -0x1.000986p-7f, -0x1.00140cp-15f, -0x1.f9ca98p-8f, -0x1.f3aa6p-16f,
    -0x1.f38226p-8f, -0x1.e75422p-16f, -0x1.ed39b4p-8f, -0x1.db255ep-16f, -0x1.e6f14p-8f, -0x1.cf1e14p-16f, -0x1.e0a8ccp-8f, -0x1.c33e44p-16f,
    -0x1.da6058p-8f, -0x1.b785fp-16f, -0x1.d417e4p-8f, -0x1.abf514p-16f, -0x1.cdcf7p-8f, -0x1.a08bb4p-16f, -0x1.c786fap-8f, -0x1.9549ccp-16f,
    -0x1.c13e86p-8f, -0x1.8a2f6p-16f, -0x1.baf61p-8f, -0x1.7f3c6ep-16f, -0x1.b4ad9cp-8f, -0x1.7470f6p-16f, -0x1.ae6526p-8f, -0x1.69ccf8p-16f,
    -0x1.a81cbp-8f, -0x1.5f5074p-16f, -0x1.a1d438p-8f, -0x1.54fb6cp-16f, -0x1.9b8bc2p-8f, -0x1.4acddcp-16f, -0x1.95434cp-8f, -0x1.40c7c8p-16f,
    -0x1.8efad4p-8f, -0x1.36e92ep-16f, -0x1.88b25ep-8f, -0x1.2d320ep-16f, -0x1.8269e6p-8f, -0x1.23a268p-16f, -0x1.7c216ep-8f, -0x1.1a3a3cp-16f,
    -0x1.75d8f6p-8f, -0x1.10f98cp-16f, -0x1.6f907ep-8f, -0x1.07e054p-16f, -0x1.694804p-8f, -0x1.fddd3p-17f, -0x1.62ff8cp-8f, -0x1.ec48acp-17f,
    -0x1.5cb714p-8f, -0x1.db031cp-17f, -0x1.566e9ap-8f, -0x1.ca0c8p-17f, -0x1.50262p-8f, -0x1.b964dap-17f, -0x1.49dda8p-8f, -0x1.a90c28p-17f,
    -0x1.43952ep-8f, -0x1.99026cp-17f, -0x1.3d4cb4p-8f, -0x1.8947a4p-17f, -0x1.37043ap-8f, -0x1.79dbdp-17f, -0x1.30bbcp-8f, -0x1.6abefp-17f,
    -0x1.2a7344p-8f, -0x1.5bf106p-17f, -0x1.242acap-8f, -0x1.4d7212p-17f, -0x1.1de25p-8f, -0x1.3f421p-17f, -0x1.1799d4p-8f, -0x1.316104p-17f,
    -0x1.11515ap-8f, -0x1.23ceeep-17f, -0x1.0b08dep-8f, -0x1.168bccp-17f, -0x1.04c062p-8f, -0x1.09979ep-17f, -0x1.fcefcep-9f, -0x1.f9e4cap-18f,
    -0x1.f05ed6p-9f, -0x1.e13842p-18f, -0x1.e3cddep-9f, -0x1.c929a2p-18f, -0x1.d73ce6p-9f, -0x1.b1b8eep-18f, -0x1.caabeep-9f, -0x1.9ae622p-18f,
    -0x1.be1af4p-9f, -0x1.84b142p-18f, -0x1.b189fcp-9f, -0x1.6f1a4ap-18f, -0x1.a4f902p-9f, -0x1.5a213cp-18f, -0x1.986808p-9f, -0x1.45c616p-18f,
    -0x1.8bd71p-9f, -0x1.3208dcp-18f, -0x1.7f4614p-9f, -0x1.1ee98cp-18f, -0x1.72b51ap-9f, -0x1.0c6824p-18f, -0x1.66242p-9f, -0x1.f5094cp-19f,
    -0x1.599326p-9f, -0x1.d27e24p-19f, -0x1.4d022ap-9f, -0x1.b12edp-19f, -0x1.40713p-9f, -0x1.911b5p-19f, -0x1.33e034p-9f, -0x1.7243a4p-19f,
    -0x1.274f38p-9f, -0x1.54a7cap-19f, -0x1.1abe3ep-9f, -0x1.3847c4p-19f, -0x1.0e2d42p-9f, -0x1.1d2392p-19f, -0x1.019c46p-9f, -0x1.033b32p-19f,
    -0x1.ea1692p-10f, -0x1.d51d5p-20f, -0x1.d0f49ap-10f, -0x1.a63be2p-20f, -0x1.b7d2ap-10f, -0x1.79d21ap-20f, -0x1.9eb0a8p-10f, -0x1.4fdffap-20f,
    -0x1.858eaep-10f, -0x1.286582p-20f, -0x1.6c6cb4p-10f, -0x1.0362bp-20f, -0x1.534abap-10f, -0x1.c1af0cp-21f, -0x1.3a28cp-10f, -0x1.818808p-21f,
    -0x1.2106c6p-10f, -0x1.465052p-21f, -0x1.07e4ccp-10f, -0x1.1007ecp-21f, -0x1.dd85a2p-11f, -0x1.bd5daap-22f, -0x1.ab41aep-11f, -0x1.648a18p-22f,
    -0x1.78fdb8p-11f, -0x1.159526p-22f, -0x1.46b9c2p-11f, -0x1.a0fda4p-23f, -0x1.1475ccp-11f, -0x1.2a8e36p-23f, -0x1.c463acp-12f, -0x1.8fb80ep-24f,
    -0x1.5fdbbep-12f, -0x1.e39c52p-25f, -0x1.f6a7a2p-13f, -0x1.ed7afp-26f, -0x1.2d97c8p-13f, -0x1.634e46p-27f, -0x1.921fb6p-15f, -0x1.3bd3ccp-30f,
    0x1.921fb6p-15f, -0x1.3bd3ccp-30f, 0x1.2d97c8p-13f, -0x1.634e46p-27f, 0x1.f6a7a2p-13f, -0x1.ed7afp-26f, 0x1.5fdbbep-12f, -0x1.e39c52p-25f,
    0x1.c463acp-12f, -0x1.8fb80ep-24f, 0x1.1475ccp-11f, -0x1.2a8e36p-23f, 0x1.46b9c2p-11f, -0x1.a0fda4p-23f, 0x1.78fdb8p-11f, -0x1.159526p-22f,
    0x1.ab41aep-11f, -0x1.648a18p-22f, 0x1.dd85a2p-11f, -0x1.bd5daap-22f, 0x1.07e4ccp-10f, -0x1.1007ecp-21f, 0x1.2106c6p-10f, -0x1.465052p-21f,
    0x1.3a28cp-10f, -0x1.818808p-21f, 0x1.534abap-10f, -0x1.c1af0cp-21f, 0x1.6c6cb4p-10f, -0x1.0362bp-20f, 0x1.858eaep-10f, -0x1.286582p-20f,
    0x1.9eb0a8p-10f, -0x1.4fdffap-20f, 0x1.b7d2ap-10f, -0x1.79d21ap-20f, 0x1.d0f49ap-10f, -0x1.a63be2p-20f, 0x1.ea1692p-10f, -0x1.d51d5p-20f,
    0x1.019c46p-9f, -0x1.033b32p-19f, 0x1.0e2d42p-9f, -0x1.1d2392p-19f, 0x1.1abe3ep-9f, -0x1.3847c4p-19f, 0x1.274f38p-9f, -0x1.54a7cap-19f,
    0x1.33e034p-9f, -0x1.7243a4p-19f, 0x1.40713p-9f, -0x1.911b5p-19f, 0x1.4d022ap-9f, -0x1.b12edp-19f, 0x1.599326p-9f, -0x1.d27e24p-19f,
    0x1.66242p-9f, -0x1.f5094cp-19f, 0x1.72b51ap-9f, -0x1.0c6824p-18f, 0x1.7f4614p-9f, -0x1.1ee98cp-18f, 0x1.8bd71p-9f, -0x1.3208dcp-18f,
    0x1.986808p-9f, -0x1.45c616p-18f, 0x1.a4f902p-9f, -0x1.5a213cp-18f, 0x1.b189fcp-9f, -0x1.6f1a4ap-18f, 0x1.be1af4p-9f, -0x1.84b142p-18f,
    0x1.caabeep-9f, -0x1.9ae622p-18f, 0x1.d73ce6p-9f, -0x1.b1b8eep-18f, 0x1.e3cddep-9f, -0x1.c929a2p-18f, 0x1.f05ed6p-9f, -0x1.e13842p-18f,
    0x1.fcefcep-9f, -0x1.f9e4cap-18f, 0x1.04c062p-8f, -0x1.09979ep-17f, 0x1.0b08dep-8f, -0x1.168bccp-17f, 0x1.11515ap-8f, -0x1.23ceeep-17f,
    0x1.1799d4p-8f, -0x1.316104p-17f, 0x1.1de25p-8f, -0x1.3f421p-17f, 0x1.242acap-8f, -0x1.4d7212p-17f, 0x1.2a7344p-8f, -0x1.5bf106p-17f,
    0x1.30bbcp-8f, -0x1.6abefp-17f, 0x1.37043ap-8f, -0x1.79dbdp-17f, 0x1.3d4cb4p-8f, -0x1.8947a4p-17f, 0x1.43952ep-8f, -0x1.99026cp-17f,
    0x1.49dda8p-8f, -0x1.a90c28p-17f, 0x1.50262p-8f, -0x1.b964dap-17f, 0x1.566e9ap-8f, -0x1.ca0c8p-17f, 0x1.5cb714p-8f, -0x1.db031cp-17f,
    0x1.62ff8cp-8f, -0x1.ec48acp-17f, 0x1.694804p-8f, -0x1.fddd3p-17f, 0x1.6f907ep-8f, -0x1.07e054p-16f, 0x1.75d8f6p-8f, -0x1.10f98cp-16f,
    0x1.7c216ep-8f, -0x1.1a3a3cp-16f, 0x1.8269e6p-8f, -0x1.23a268p-16f, 0x1.88b25ep-8f, -0x1.2d320ep-16f, 0x1.8efad4p-8f, -0x1.36e92ep-16f,
    0x1.95434cp-8f, -0x1.40c7c8p-16f, 0x1.9b8bc2p-8f, -0x1.4acddcp-16f, 0x1.a1d438p-8f, -0x1.54fb6cp-16f, 0x1.a81cbp-8f, -0x1.5f5074p-16f,
    0x1.ae6526p-8f, -0x1.69ccf8p-16f, 0x1.b4ad9cp-8f, -0x1.7470f6p-16f, 0x1.baf61p-8f, -0x1.7f3c6ep-16f, 0x1.c13e86p-8f, -0x1.8a2f6p-16f,
    0x1.c786fap-8f, -0x1.9549ccp-16f, 0x1.cdcf7p-8f, -0x1.a08bb4p-16f, 0x1.d417e4p-8f, -0x1.abf514p-16f, 0x1.da6058p-8f, -0x1.b785fp-16f,
    0x1.e0a8ccp-8f, -0x1.c33e44p-16f, 0x1.e6f14p-8f, -0x1.cf1e14p-16f, 0x1.ed39b4p-8f, -0x1.db255ep-16f, 0x1.f38226p-8f, -0x1.e75422p-16f,
    0x1.f9ca98p-8f, -0x1.f3aa6p-16f, 0x1.000986p-7f, -0x1.00140cp-15f, 0x1.032dbep-7f, -0x1.0666a6p-15f, 0x1.0651f8p-7f, -0x1.0cccfcp-15f,
    0x1.09763p-7f, -0x1.13470ep-15f, 0x1.0c9a68p-7f, -0x1.19d4ep-15f, 0x1.0fbea2p-7f, -0x1.20766cp-15f, 0x1.12e2dap-7f, -0x1.272bb8p-15f,
    0x1.160712p-7f, -0x1.2df4bep-15f, 0x1.192b4ap-7f, -0x1.34d184p-15f, 0x1.1c4f82p-7f, -0x1.3bc206p-15f, 0x1.1f73b8p-7f, -0x1.42c644p-15f,
    0x1.2297fp-7f, -0x1.49de4p-15f, 0x1.25bc28p-7f, -0x1.5109f8p-15f, 0x1.28e05ep-7f, -0x1.58496ep-15f, 0x1.2c0496p-7f, -0x1.5f9ca2p-15f,
    0x1.2f28ccp-7f, -0x1.670392p-15f, 0x1.324d02p-7f, -0x1.6e7e3ep-15f, 0x1.35713ap-7f, -0x1.760ca8p-15f, 0x1.38957p-7f, -0x1.7daedp-15f,
    0x1.3bb9a6p-7f, -0x1.8564b4p-15f, 0x1.3edddap-7f, -0x1.8d2e54p-15f, 0x1.42021p-7f, -0x1.950bb2p-15f, 0x1.452646p-7f, -0x1.9cfccep-15f,
    0x1.484a7cp-7f, -0x1.a501a6p-15f, 0x1.4b6ebp-7f, -0x1.ad1a3ap-15f, 0x1.4e92e4p-7f, -0x1.b5468cp-15f, 0x1.51b71ap-7f, -0x1.bd869cp-15f,
    0x1.54db4ep-7f, -0x1.c5da68p-15f, 0x1.57ff82p-7f, -0x1.ce41fp-15f, 0x1.5b23b6p-7f, -0x1.d6bd36p-15f, 0x1.5e47eap-7f, -0x1.df4c38p-15f,
    0x1.616c1ep-7f, -0x1.e7eef8p-15f, 0x1.64905p-7f, -0x1.f0a574p-15f, 0x1.67b484p-7f, -0x1.f96faep-15f, 0x1.6ad8b6p-7f, -0x1.0126d2p-14f,
    0x1.6dfceap-7f, -0x1.059facp-14f, 0x1.71211cp-7f, -0x1.0a2264p-14f, 0x1.74454ep-7f, -0x1.0eaefcp-14f, 0x1.77698p-7f, -0x1.13457p-14f,
    0x1.7a8db2p-7f, -0x1.17e5c4p-14f, 0x1.7db1e4p-7f, -0x1.1c8ff6p-14f, 0x1.80d614p-7f, -0x1.214406p-14f, 0x1.83fa46p-7f, -0x1.2601f6p-14f,
    0x1.871e76p-7f, -0x1.2ac9c2p-14f, 0x1.8a42a8p-7f, -0x1.2f9b6ep-14f, 0x1.8d66d8p-7f, -0x1.3476f8p-14f, 0x1.908b08p-7f, -0x1.395c62p-14f,
};

// Copy both angle tables into LDS (whole workgroup; the caller synchronises).
__device__ __forceinline__ void load_angle_table(float* lds)
{
    for (uint32_t i = threadIdx.x; i < (uint32_t)kSinCosTabFloats; i += blockDim.x) {
        lds[i] = kSinCosTab[i];
        lds[kSinCosTabFloats + i] = kBetaTab[i];
    }
}

// ------------------------------------------------------------- record stage

struct Stage {
    uint32_t* buf;      // this wave's LDS slice
    uint32_t cap;       // entries
    uint32_t cnt;       // wave-uniform fill
    uint32_t* rec;      // proposal's record slice
    uint32_t slice_cap;
    uint32_t* counter;  // proposal record counter
};

__device__ __forceinline__ void stage_flush(Stage& st)
{
    if (st.cnt == 0u) return;
    wave_lds_sync();
    uint32_t base = 0u;
    if (lane_id() == 0u) base = atomicAdd(st.counter, st.cnt);
    base = uniform(base);
    auto* rec = gview(st.rec);  // record slices are in HBM
    for (uint32_t i = lane_id(); i < st.cnt; i += kWave) {
        const uint32_t idx = base + i;
        if (idx < st.slice_cap) rec[idx] = st.buf[i];
    }
    wave_lds_sync();
    st.cnt = 0u;
}

// Track record: de / ((k-1) 4 dt) then bin = trunc(value / dDE), dropped when
// the bin or the length is outside the matrix
// (src/gpu/cusim/kernel-simulate-unroll.inl:187-190, src/gpu/cusim/histogram2d.inl:98-106).
__device__ __forceinline__ bool track_record(uint32_t k, float de, float four_dt, float dde,
                                             uint32_t n_k, uint32_t n_de, uint32_t& out)
{
    const float q = (float)(k - 1u) * four_dt;
    const float v = de / q;
    const float tb = v / dde;
    const bool ok = (k <= n_k) && (tb < (float)n_de);
    out = ok ? (((k - 1u) << 16) | (uint32_t)tb) : 0u;
    return ok;
}

struct BinSpec {
    float four_dt, dde;
    uint32_t n_k, n_de;
};

// The stage holds raw (k, de) pairs; they are binned here, with every lane of
// the wave busy, rather than when a track closes (where a whole wave would run
// the two divisions for one or two closing lanes). The kept records are
// compacted in place: batch i0 reads words [2 i0, 2 i0 + 128) and writes below
// word i0 + 64, so it never overwrites a pair a later batch still reads.
__device__ __forceinline__ void stage_flush_binned(Stage& st, const BinSpec& bs)
{
    if (st.cnt == 0u) return;
    wave_lds_sync();
    const uint32_t n = st.cnt >> 1;
    uint32_t kept = 0u;
    for (uint32_t i0 = 0u; i0 < n; i0 += (uint32_t)kWave) {
        const uint32_t i = i0 + lane_id();
        uint32_t rv = 0u;
        bool ok = false;
        if (i < n) ok = track_record(st.buf[2u * i], __uint_as_float(st.buf[2u * i + 1u]), bs.four_dt, bs.dde,
                                     bs.n_k, bs.n_de, rv);
        const uint64_t m = ballot(ok);
        wave_lds_sync();
        if (ok) st.buf[kept + mbcnt64(m)] = rv;
        kept += (uint32_t)__builtin_popcountll(m);
    }
    st.cnt = kept;
    stage_flush(st);
}

// Two-word records: (k, DE bits) of the track-list form (BIN = false) or the
// raw (k, de bits) pairs that stage_flush_binned bins (BIN = true), one per
// lane of the non-zero wave mask m. The mask is the store predicate as it is,
// so no per-lane flag is rebuilt from it.
// The stage is flushed only when these records do not fit, so it fills up
// whatever its size (at least 2 x 64 words).
template <bool BIN>
__device__ __forceinline__ void stage_push2(Stage& st, uint64_t m, uint32_t a, uint32_t b, const BinSpec& bs)
{
    const uint32_t need = 2u * (uint32_t)__builtin_popcountll(m);
    if (st.cnt + need > st.cap) {
        if (BIN) stage_flush_binned(st, bs);
        else stage_flush(st);
    }
    const uint32_t pos = st.cnt + 2u * mbcnt64(m);
    if (__builtin_amdgcn_inverse_ballot_w64(m)) {
        st.buf[pos] = a;
        st.buf[pos + 1u] = b;
    }
    st.cnt += need;
}

// ------------------------------------------------------------------- walk

// Positions along periodic axes are kept in box units of a frame whose origin
// is the observation window's lower edge: q in [0, 1), inside the window iff
// q <= 2 ha / L, a step is q = fract(q + (sigma / L) n) (the single periodic
// wrap of the reference for any step shorter than the box). x and y are always
// periodic; z is too unless walls reflect it or detection in z is Gaussian,
// and then z stays in um, centred on the box.
// v_fract_f32(v) = min(v - floor(v), 1 - 2^-24), bit for bit on every v in
// [-0.125, 1.125) (checked exhaustively on the GPU), which the oracle restates.
__device__ __forceinline__ float fract_box(float v) { return __builtin_amdgcn_fractf(v); }

// de is the track's sum of squared lateral steps in units of sigma^2 (the sum
// of nx^2 + ny^2 over its unit normals); track_de gives it in um^2. The open
// track's length is kept as the frame it started in, ts, counted from the
// first frame before the current 4-step chunk (rebased after each chunk): in
// that frame numbering the track ending at frame f has k = f - ts + 1, and
// the per-step update is one select instead of an add and a select.
struct Particle {
    float x, y, z, de;
    float sx, sy, sz;  // step per unit normal: sigma / L in box units, sigma for z in um
    float wz;          // z window: 2 haz / Lz in box units, haz for z in um
    float sig2;        // sigma^2
    int32_t ts;        // first frame of the open track (meaningful while in the window)
    uint32_t id;
    bool active;
    bool det0;         // inside the window in the first frame
};

__device__ __forceinline__ float track_de(const Particle& q) { return q.sig2 * q.de; }

// Per-tile constants of the walk: box and window in the units the walk uses
// (z in box units when periodic: measured from the window's lower edge for
// the slab, from the box floor for Gaussian detection (centred distance
// q - 1/2); in um, centred, between reflecting walls).
struct WalkConsts {
    float hz, Lx, Ly, Lz, ox, oy, wx, wy, cg, four_dt;
    uint32_t kmin, C;
    BinSpec bs;
};

template <bool GAUSS, bool ZREF>
__device__ __forceinline__ WalkConsts walk_consts(const DevSys& sy)
{
    constexpr bool ZBOX = !ZREF;
    WalkConsts K;
    K.hz = sy.half_l[2];
    K.Lx = 2.0f * sy.half_l[0];
    K.Ly = 2.0f * sy.half_l[1];
    K.Lz = 2.0f * K.hz;
    const float hax = sy.half_a[0], hay = sy.half_a[1];
    K.ox = hax / K.Lx - 0.5f;  // window origin offsets
    K.oy = hay / K.Ly - 0.5f;
    K.wx = (hax + hax) / K.Lx;  // window widths
    K.wy = (hay + hay) / K.Ly;
    const float sdz = ZBOX ? sy.detect_sigma / K.Lz : sy.detect_sigma;  // Gaussian detection sigma, z's unit
    K.cg = gauss_c(sdz * sdz);
    K.kmin = sy.kmin;
    K.C = sy.n_species;
    K.four_dt = sy.four_dt;
    K.bs = BinSpec{sy.four_dt, sy.de_bin_width, sy.n_k, sy.n_de};
    return K;
}

// Particle q.id of the tile: start position, species (sigma, window depth)
// and whether it starts inside the window, from Philox block 0 (and block
// 0xffffffff for Gaussian detection).
template <bool GAUSS, bool ZREF>
__device__ __forceinline__ void init_particle(Particle& q, uint32_t N, uint32_t video, uint32_t sim, uint32_t k0,
                                              uint32_t k1, const DevProp& pr, const WalkConsts& K)
{
    constexpr bool ZBOX = !ZREF;
    q.active = q.id < N;
    const U4 w = philox10(0u, q.id, video, sim, k0, k1);
    q.x = fract_box(u24(w.x) + K.ox);
    q.y = fract_box(u24(w.y) + K.oy);
    const float us = u24(w.w);
    uint32_t s = 0u;
#pragma unroll
    for (uint32_t c = 0; c + 1u < PCS_MAX_SPECIES; ++c) s += (c + 1u < K.C && us >= pr.cum_prob[c]) ? 1u : 0u;
    float sg = pr.sigma[0], az = pr.half_az[0];
#pragma unroll
    for (uint32_t c = 1; c < PCS_MAX_SPECIES; ++c) {
        sg = (s == c) ? pr.sigma[c] : sg;
        az = (s == c) ? pr.half_az[c] : az;
    }
    q.sx = sg / K.Lx;
    q.sy = sg / K.Ly;
    q.sig2 = sg * sg;
    if (ZBOX && GAUSS) {
        q.sz = sg / K.Lz;
        q.wz = 0.0f;
        q.z = u24(w.z);
    } else if (ZBOX) {
        q.sz = sg / K.Lz;
        q.wz = (az + az) / K.Lz;
        q.z = fract_box(u24(w.z) + (az / K.Lz - 0.5f));
    } else {
        q.sz = sg;
        q.wz = az;
        q.z = K.hz * (2.0f * u24(w.z) - 1.0f);
    }
    bool det;
    if (GAUSS) {
        const U4 wd = philox10(0xffffffffu, q.id, video, sim, k0, k1);
        det = q.x <= K.wx && q.y <= K.wy && gauss_detect(ZBOX ? q.z - 0.5f : q.z, K.cg, wd.x);
    } else if (ZBOX) {
        det = q.x <= K.wx && q.y <= K.wy && q.z <= q.wz;
    } else {
        det = q.x <= K.wx && q.y <= K.wy && __builtin_fabsf(q.z) <= q.wz;
    }
    q.ts = 0;
    q.det0 = det;
    q.de = 0.0f;
}

// Box-Muller pair p of a chunk from its radius word wr and the word wa that
// holds its angle's low byte (byte p % 4): d0 = r sin, d1 = r cos, y = r^2.
__device__ __forceinline__ void bm_pair(uint32_t wr, uint32_t wa, int p, const float2* tab, float& d0, float& d1,
                                        float& y)
{
    y = neg2log_u24(wr);
    const float r = sqrt_rn_domain(y);
    float sn, cs;
    sincos_tab(tab, wr & 0xffu, (wa >> (8 * (p & 3))) & 0xffu, sn, cs);
    d0 = r * sn;
    d1 = r * cs;
}

// Step s of a chunk (to frame s + 1 of the chunk's numbering) of particle q
// with the unit normals (nx, ny, nz), its pair's r^2 and (Gaussian
// detection) the word wdet: move, detect, close and stage a finished track,
// update the open track. m_prev holds the lanes inside the window at the
// frame before.
template <bool GAUSS, bool ZREF, bool TRACKS>
__device__ __forceinline__ void step_particle(Particle& q, uint64_t m_active, uint64_t& m_prev, float nx, float ny,
                                              float nz, float r2, uint32_t wdet, uint32_t s, const WalkConsts& K,
                                              Stage& st)
{
    constexpr bool ZBOX = !ZREF;
    q.x = fract_box(__builtin_fmaf(q.sx, nx, q.x));
    q.y = fract_box(__builtin_fmaf(q.sy, ny, q.y));
    if (ZBOX) {
        q.z = fract_box(__builtin_fmaf(q.sz, nz, q.z));
    } else {
        const float zn = __builtin_fmaf(q.sz, nz, q.z);
        q.z = wrap_reflect(zn, K.hz);
    }
    const bool cx = q.x <= K.wx, cy = q.y <= K.wy;
    const bool cz = GAUSS ? gauss_detect(ZBOX ? q.z - 0.5f : q.z, K.cg, wdet)
                          : (ZBOX ? q.z <= q.wz : __builtin_fabsf(q.z) <= q.wz);
    // Lanes inside the window and lanes closing a track: wave masks formed
    // from the masks of the single comparisons (no per-lane boolean is
    // materialised). This step reaches frame s + 1; the open track ends at
    // frame s with k = s + 1 - ts, at least kmin iff ts <= s + 1 - kmin. The
    // length test runs only in steps where some lane's track ends (about a
    // quarter of them).
    const int32_t f1 = (int32_t)s + 1;
    const uint64_t m_det = ballot(cx) & ballot(cy) & ballot(cz);
    const uint64_t m_end = m_active & m_prev & ~m_det;
    if (m_end != 0ull) {
        const uint64_t m_closing = m_end & ballot(q.ts <= f1 - (int32_t)K.kmin);
        if (m_closing != 0ull) {
            const bool closing = __builtin_amdgcn_inverse_ballot_w64(m_closing);
            const uint32_t k = (uint32_t)(f1 - q.ts);
            float v = track_de(q);
            if (TRACKS && closing) v = v / ((float)(k - 1u) * K.four_dt);
            stage_push2<!TRACKS>(st, m_closing, k, __float_as_uint(v), K.bs);
        }
    }
    // nx^2 + ny^2 of this step is pair s's r^2, the log's result
    const float de_n = q.de + r2;
    q.de = __builtin_amdgcn_inverse_ballot_w64(m_det & m_prev) ? de_n : 0.0f;
    q.ts = __builtin_amdgcn_inverse_ballot_w64(m_det & ~m_prev) ? f1 : q.ts;
    m_prev = m_det;
}

// A full 4-step chunk of particle q, for a wave alone on its SIMD: the four
// steps' moves and detection masks first (they do not depend on the track
// state), then each step's track logic as step_particle runs it. Every
// value and the order of the staged records are step_particle's; the moves
// and comparisons of later steps no longer wait behind each step's branch.
template <bool GAUSS, bool ZREF>
__device__ __forceinline__ void step_chunk4(Particle& q, uint64_t m_active, uint64_t& m_prev, const float (&d)[12],
                                            const float (&r2)[4], const uint32_t (&wdet)[4], const WalkConsts& K,
                                            Stage& st)
{
    constexpr bool ZBOX = !ZREF;
    uint64_t m_det[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        q.x = fract_box(__builtin_fmaf(q.sx, d[2 * s], q.x));
        q.y = fract_box(__builtin_fmaf(q.sy, d[2 * s + 1], q.y));
        if (ZBOX) {
            q.z = fract_box(__builtin_fmaf(q.sz, d[8 + s], q.z));
        } else {
            const float zn = __builtin_fmaf(q.sz, d[8 + s], q.z);
            q.z = wrap_reflect(zn, K.hz);
        }
        const bool cx = q.x <= K.wx, cy = q.y <= K.wy;
        const bool cz = GAUSS ? gauss_detect(ZBOX ? q.z - 0.5f : q.z, K.cg, wdet[s])
                              : (ZBOX ? q.z <= q.wz : __builtin_fabsf(q.z) <= q.wz);
        m_det[s] = ballot(cx) & ballot(cy) & ballot(cz);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int32_t f1 = s + 1;
        const uint64_t m_end = m_active & m_prev & ~m_det[s];
        if (m_end != 0ull) {
            const uint64_t m_closing = m_end & ballot(q.ts <= f1 - (int32_t)K.kmin);
            if (m_closing != 0ull) {
                const uint32_t k = (uint32_t)(f1 - q.ts);
                stage_push2<true>(st, m_closing, k, __float_as_uint(track_de(q)), K.bs);
            }
        }
        const float de_n = q.de + r2[s];
        q.de = __builtin_amdgcn_inverse_ballot_w64(m_det[s] & m_prev) ? de_n : 0.0f;
        q.ts = __builtin_amdgcn_inverse_ballot_w64(m_det[s] & ~m_prev) ? f1 : q.ts;
        m_prev = m_det[s];
    }
}

// Tracks still open in the last frame (src/cpu/simulate.jl:114-121), which is
// frame 0 after the last chunk's rebase: k = 1 - ts.
template <bool TRACKS>
__device__ __forceinline__ void close_open_tracks(Particle& q, uint64_t m_active, uint64_t m_prev, const WalkConsts& K,
                                                  Stage& st)
{
    const uint64_t m_closing = m_active & m_prev & ballot(q.ts <= 1 - (int32_t)K.kmin);
    if (m_closing != 0ull) {
        const bool closing = __builtin_amdgcn_inverse_ballot_w64(m_closing);
        const uint32_t k = (uint32_t)(1 - q.ts);
        float v = track_de(q);
        if (TRACKS && closing) v = v / ((float)(k - 1u) * K.four_dt);
        stage_push2<!TRACKS>(st, m_closing, k, __float_as_uint(v), K.bs);
    }
}

// TRACKS = true: the track-list form of src/cpu/generate_experiment.jl -
// every track with k >= kmin is emitted as (k, de / ((k-1) 4 dt)) without
// binning; otherwise tracks are binned records (row << 16 | bin).
// Walk of one tile: particles [p0, p0 + 64*PPL) of video `video` (F frames,
// N particles) of simulation pr; the caller has pointed the stage at pr's
// record slice and counter.
template <int PPL, bool GAUSS, bool ZREF, bool TRACKS>
__device__ __forceinline__ void walk_tile(const DevProp& pr, const DevSys& sy, uint32_t F, uint32_t N,
                                          uint32_t video, uint32_t p0, Stage& st, const float2* tab,
                                          uint32_t s0, uint32_t s1)
{
    if (p0 >= N) return;  // empty tile (a request without particles)
    // The counter words other than the particle and the key are the same on
    // every lane: held in SGPRs, the Philox rounds that see only them (round 1,
    // and half of rounds 2 and 3) run on the scalar unit.
    video = uniform(video);
    const uint32_t sim = uniform(pr.sim_id);
    const uint32_t k0 = uniform(s0 ^ pr.key0), k1 = uniform(s1 ^ pr.key1);
    const WalkConsts K = walk_consts<GAUSS, ZREF>(sy);
    // Philox blocks per 4-step chunk: 2 for the 12 unit normals (+1 with
    // Gaussian detection for its 4 uniforms).
    constexpr uint32_t B = GAUSS ? 3u : 2u;

    Particle pt[PPL];
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
        pt[i].id = p0 + (uint32_t)(i * kWave) + lane_id();
        init_particle<GAUSS, ZREF>(pt[i], N, video, sim, k0, k1, pr, K);
    }

    // Per particle slot, wave masks of the lanes holding a particle and of the
    // lanes inside the window in the previous frame. Kept as masks (SGPRs), the
    // per-lane selects read them directly (inverse ballot).
    uint64_t m_active[PPL], m_prev[PPL];
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
        m_active[i] = ballot(pt[i].active);
        m_prev[i] = ballot(pt[i].det0);
    }

    const uint32_t steps = F > 0u ? F - 1u : 0u;
    for (uint32_t c = 0; 4u * c < steps; ++c) {
        const uint32_t nst = min(4u, steps - 4u * c);
        float d[PPL][12];   // pair p: d[2p] = r sin, d[2p + 1] = r cos
        float r2[PPL][4];   // r^2 = -2 ln u of pairs 0..3
        uint32_t wdet[PPL][4];
#pragma unroll
        for (int i = 0; i < PPL; ++i) {
            // Words 0..5: radius uniforms (top 24 bits) of pairs 0..5; the angle of
            // pair p has code (low byte of word p) << 8 | byte p % 4 of word 6 + p / 4.
            const U4 b0 = philox10(B * c + 1u, pt[i].id, video, sim, k0, k1);
            const U4 b1 = philox10(B * c + 2u, pt[i].id, video, sim, k0, k1);
            const uint32_t w[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
            for (int p = 0; p < 6; ++p) {
                float y;
                bm_pair(w[p], w[6 + p / 4], p, tab, d[i][2 * p], d[i][2 * p + 1], y);
                if (p < 4) r2[i][p] = y;
            }
            if (GAUSS) {
                const U4 b2 = philox10(B * c + 3u, pt[i].id, video, sim, k0, k1);
                wdet[i][0] = b2.x;
                wdet[i][1] = b2.y;
                wdet[i][2] = b2.z;
                wdet[i][3] = b2.w;
            }
        }
#pragma unroll
        for (uint32_t s = 0; s < 4u; ++s) {
            if (s >= nst) break;
#pragma unroll
            for (int i = 0; i < PPL; ++i) {
                // Unit normals of this step (x and y from pair s, z from pairs 4
                // and 5); the particle moves by sigma * n.
                step_particle<GAUSS, ZREF, TRACKS>(pt[i], m_active[i], m_prev[i], d[i][2 * s], d[i][2 * s + 1],
                                                   d[i][8 + s], r2[i][s], GAUSS ? wdet[i][s] : 0u, s, K, st);
            }
        }
#pragma unroll
        for (int i = 0; i < PPL; ++i) pt[i].ts -= (int32_t)nst;  // the chunk's last frame becomes frame 0
    }
#pragma unroll
    for (int i = 0; i < PPL; ++i) close_open_tracks<TRACKS>(pt[i], m_active[i], m_prev[i], K, st);
    if (TRACKS) stage_flush(st);
    else stage_flush_binned(st, K.bs);
}

// Cooperative walk of one tile by the four waves of a serving group, for a
// request's latency: a lone wave issues one instruction per few cycles, so
// its tile takes as long as its whole instruction stream. The chunk's work is
// pipelined over three stages with one group barrier per round: in round i,
// waves 0, 1 and 2 run Philox blocks 1, 2 and (Gaussian detection) 3 of
// chunk i into an LDS word buffer, one each; waves 0..2 turn chunk i - 1's
// words into its unit normals - wave w the Box-Muller pairs 2w and 2w + 1,
// wave 2 also hands on the detection words - into a normal buffer; wave 3
// holds the particles, reads chunk i - 2's normals and steps chunk i - 3,
// whose normals it read in the round before, and stages its tracks.
// Both buffers are double (chunk parity). The values and their order are
// walk_tile's, so the results are the same. xbuf: kCoopBufWords words
// (lane-minor, so a wave's accesses to one word hit 64 consecutive words).
constexpr int kCoopWords = 20;     // normals per lane and chunk: d[12], r^2 of pairs 0..3, detection words 0..3
constexpr int kCoopRawWords = 12;  // Philox words per lane and chunk: blocks 1, 2 and 3 (or a spare)
constexpr int kCoopBufWords = 2 * (kCoopWords + kCoopRawWords) * kWave;

template <bool GAUSS, bool ZREF>
__device__ __forceinline__ void walk_tile_coop(const DevProp& pr, const DevSys& sy, uint32_t F, uint32_t N,
                                               uint32_t video, uint32_t p0, Stage& st, const float2* tab,
                                               uint32_t* xbuf)
{
    if (p0 >= N) return;  // the same tile on every wave: all return together
    const uint32_t wave = uniform(threadIdx.x / kWave), lane = lane_id();  // branches on it are uniform
    video = uniform(video);
    const uint32_t sim = uniform(pr.sim_id);
    const uint32_t k0 = uniform(pr.key0), k1 = uniform(pr.key1);
    const WalkConsts K = walk_consts<GAUSS, ZREF>(sy);
    constexpr uint32_t B = GAUSS ? 3u : 2u;
    const uint32_t id = p0 + lane;
    const uint32_t steps = F > 0u ? F - 1u : 0u;
    const uint32_t nc = (steps + 3u) / 4u;
    uint32_t* raw = xbuf + 2 * kCoopWords * kWave;

    // The stepping wave and the producers run loops of their own, with one
    // group barrier per round each and the same number of rounds (the branch
    // on the wave index is wave-uniform), so that neither holds the other's
    // registers.
    if (wave == 3u) {
        Particle q;
        q.id = id;
        init_particle<GAUSS, ZREF>(q, N, video, sim, k0, k1, pr, K);
        const uint64_t m_active = ballot(q.active);
        uint64_t m_prev = ballot(q.det0);
        float d[12], r2[4];  // the chunk read a round ahead of its steps
        uint32_t wdet[4];
        for (uint32_t it = 0; it <= nc + 2u; ++it) {
            // Chunk it - 2's values (made in round it - 1) are read at the start
            // of the round; the chunk read in the round before, it - 3, is
            // stepped while they arrive.
            const bool rd = it >= 2u && it - 2u < nc;
            float dq[12], r2q[4];
            uint32_t wq[4];
            if (rd) {
                const uint32_t c = it - 2u;
                const float* xb = reinterpret_cast<const float*>(xbuf) + (c & 1u) * (uint32_t)(kCoopWords * kWave) + lane;
#pragma unroll
                for (int k = 0; k < 12; ++k) dq[k] = xb[k * kWave];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    r2q[k] = xb[(12 + k) * kWave];
                    wq[k] = GAUSS ? __float_as_uint(xb[(16 + k) * kWave]) : 0u;
                }
            }
            if (it >= 3u) {
                const uint32_t c = it - 3u;
                const uint32_t nst = min(4u, steps - 4u * c);
                if (nst == 4u) {
                    step_chunk4<GAUSS, ZREF>(q, m_active, m_prev, d, r2, wdet, K, st);
                } else {
#pragma unroll
                    for (uint32_t s = 0; s < 4u; ++s) {
                        if (s >= nst) break;
                        step_particle<GAUSS, ZREF, false>(q, m_active, m_prev, d[2 * s], d[2 * s + 1], d[8 + s],
                                                          r2[s], wdet[s], s, K, st);
                    }
                }
                q.ts -= (int32_t)nst;  // the chunk's last frame becomes frame 0
            }
            if (rd) {
#pragma unroll
                for (int k = 0; k < 12; ++k) d[k] = dq[k];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    r2[k] = r2q[k];
                    wdet[k] = wq[k];
                }
            }
            __syncthreads();
        }
        close_open_tracks<false>(q, m_active, m_prev, K, st);
        stage_flush_binned(st, K.bs);
    } else {
        for (uint32_t it = 0; it <= nc + 2u; ++it) {
            // Pair p's radius is word p of blocks 1, 2 (words 0..7), its angle
            // word 6 + p / 4. Every stage runs in every round and this wave's
            // Philox block is not under a branch, so that the compiler can
            // interleave it with the wave's Box-Muller pairs in one basic
            // block: in round 0 the pairs take words no round wrote and write
            // normals to the half that chunk 1 overwrites in round 2, before
            // the stepping wave first reads it in round 3; in rounds nc to
            // nc + 2 the blocks and pairs are those of chunks past the last,
            // written to halves whose chunks were read in the round before.
            // Any word gives finite normals and table indices below 256.
            const uint32_t c = it - 1u;  // the chunk whose normals this round makes
            const int pa = 2 * (int)wave;
            const uint32_t* rr = raw + (c & 1u) * (uint32_t)(kCoopRawWords * kWave) + lane;
            const uint32_t wr0 = rr[pa * kWave], wr1 = rr[(pa + 1) * kWave], wa = rr[(6 + pa / 4) * kWave];
            // block 1 + wave of chunk it: with Gaussian detection wave 2's is
            // block 3, the detection words, handed on in the next round;
            // without it, wave 2's block (chunk it + 1's first) goes unused
            const U4 b = philox10(B * it + 1u + wave, id, video, sim, k0, k1);
            uint32_t* rw = raw + (it & 1u) * (uint32_t)(kCoopRawWords * kWave) + 4u * wave * kWave + lane;
            rw[0] = b.x;
            rw[kWave] = b.y;
            rw[2 * kWave] = b.z;
            rw[3 * kWave] = b.w;
            float* xb = reinterpret_cast<float*>(xbuf) + (c & 1u) * (uint32_t)(kCoopWords * kWave) + lane;
            float d0, d1, y0, d2, d3, y1;
            bm_pair(wr0, wa, pa, tab, d0, d1, y0);
            bm_pair(wr1, wa, pa + 1, tab, d2, d3, y1);
            xb[(2 * pa) * kWave] = d0;
            xb[(2 * pa + 1) * kWave] = d1;
            xb[(2 * pa + 2) * kWave] = d2;
            xb[(2 * pa + 3) * kWave] = d3;
            if (wave < 2u) {
                xb[(12 + pa) * kWave] = y0;
                xb[(13 + pa) * kWave] = y1;
            }
            if (GAUSS && wave == 2u) {
#pragma unroll
                for (int j = 0; j < 4; ++j) xb[(16 + j) * kWave] = __uint_as_float(rr[(8 + j) * kWave]);
            }
            __syncthreads();
        }
    }
}

// Tile t of a batch plan: the video owning it is the largest v with
// vtile[v] <= t (wave-uniform binary search).
template <int PPL, bool GAUSS, bool ZREF, bool TRACKS>
__device__ __forceinline__ void process_tile(const WalkArgs& a, uint32_t t, Stage& st, const float2* tab,
                                             uint32_t s0, uint32_t s1)
{
    uint32_t lo = 0u, hi = a.n_videos;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.vtile[mid] <= t) lo = mid; else hi = mid;
    }
    const uint32_t v = lo;
    const uint32_t P = a.vprop[v];
    const DevProp& pr = a.props[P];
    st.rec = pr.rec;
    st.slice_cap = pr.rec_cap;
    st.counter = a.rec_count + P;
    walk_tile<PPL, GAUSS, ZREF, TRACKS>(pr, a.sys[pr.sys_id], a.vframes[v], a.vparticles[v], v - pr.first_video,
                                        (t - a.vtile[v]) * (uint32_t)(kWave * PPL), st, tab, s0, s1);
}

template <int PPL, bool PERSIST, bool GAUSS, bool ZREF, bool TRACKS = false>
__global__ __launch_bounds__(256) void walk_kernel(WalkArgs a)
{
    extern __shared__ uint32_t stage_lds[];
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t wpg = blockDim.x / kWave;
    Stage st;
    st.buf = stage_lds + wave * a.stage_cap;
    st.cap = a.stage_cap;
    st.cnt = 0u;
    float* tabf = reinterpret_cast<float*>(stage_lds + wpg * a.stage_cap);  // after the stages, 8-byte aligned
    load_angle_table(tabf);
    __syncthreads();
    const float2* tab = reinterpret_cast<const float2*>(tabf);
    const uint64_t seed = *a.seed;
    const uint32_t k0 = uniform((uint32_t)seed), k1 = uniform((uint32_t)(seed >> 32));
    if (PERSIST) {
        // tiles planned on the device end the range where the plan says
        const uint32_t hi = a.tile_total ? (uint32_t)min((uint64_t)a.tile_hi, *a.tile_total) : a.tile_hi;
        for (;;) {
            uint32_t t = 0u;
            if (lane_id() == 0u) t = atomicAdd(a.tile_counter, 1u);
            t = uniform(t) + a.tile_lo;
            if (t >= hi) break;
            process_tile<PPL, GAUSS, ZREF, TRACKS>(a, t, st, tab, k0, k1);
        }
    } else {
        const uint32_t t = a.tile_lo + blockIdx.x * wpg + wave;
        if (t < a.tile_hi) process_tile<PPL, GAUSS, ZREF, TRACKS>(a, t, st, tab, k0, k1);
    }
}

// --------------------------------------------------------------- distance

constexpr int kDistThreads = 256;
// Rows with at most this many records skip the row histogram and block scan.
constexpr int kDistSparseRow = 32;

typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
// Record loads each thread keeps in flight while reading a slice.
constexpr int kDistLoads = 8;

// Inclusive sum over the wave by data-parallel-primitive moves (VALU, no
// LDS round trip): shifts by 1, 2, 4, 8 lanes within each 16-lane row (a lane
// without a source adds 0), then lane 15 of rows 0 and 2 added to rows 1 and
// 3, then lane 31 added to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return (uint32_t)x;
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t& total)
{
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const uint32_t x = wave_incl_scan(v);
    if (lane == kWave - 1) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0u, tot = 0u;
#pragma unroll
    for (int w = 0; w < kDistThreads / kWave; ++w) {
        const uint32_t s = wsum[w];
        base += ((uint32_t)w < wave) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + x - v;
}

// LDS working set of the distance of one simulation: row counts / row
// offsets (n_k), one row's column histogram (n_de), the block-scan partials,
// the deepest occupied row, the final reduction (kDistRed words) and room for
// the bins sorted by row (srtl_cap 16-bit entries; a slice with more records
// sorts them into its HBM scratch instead).
struct DistLds {
    uint32_t* rowc;
    uint32_t* colh;
    uint32_t* wsum;
    uint32_t* smax;
    uint64_t* red;
    uint16_t* srtl;
    uint32_t srtl_cap;
};
constexpr int kDistRed = 2 * (kDistThreads / kWave);

// Result of distance_block, valid in thread 0: the exact sum in 128 bits
// (hi:lo), the records and the status.
struct DistOut {
    uint64_t lo, hi;
    uint32_t tracks;
    int32_t status;
#if PCS_TRACE
    uint64_t t_sweep;  // device wall clock when the row sweep began
#endif
};

// 128-bit accumulation of the distance sum (it passes 2^64 when cumulative
// counts approach 2^32): (hi:lo) += (bh:bl).
__device__ __forceinline__ void add128(uint64_t& lo, uint64_t& hi, uint64_t bl, uint64_t bh)
{
    lo += bl;
    hi += bh + (lo < bl ? 1ull : 0ull);
}

// Exact sum saturated at 2^64 - 1, and its log10 from the full 128 bits
// (2^64 hi + lo rounded once to FP64).
__device__ __forceinline__ uint64_t dist_exact(const DistOut& o) { return o.hi ? ~0ull : o.lo; }
__device__ __forceinline__ double dist_log10(const DistOut& o)
{
    return o.status == PCS_OK ? log10(ldexp((double)o.hi, 64) + (double)o.lo) : __builtin_nan("");
}

// Exact distance of one simulation's record slice against the observed
// summed-area table, by a whole 256-thread block. SEG * 256 >= n_de.
//
// The slice is read from HBM once when it holds at most KREG * 256 records:
// record t = tid + 256 u stays in register held[u] between the row count and
// the scatter. Records past that are read a second time by the scatter. The
// scatter writes the 16-bit bins sorted by row into LDS when the slice fits
// L.srtl_cap, else into the slice's HBM scratch. Within a row the order of the
// bins depends on the atomics; the sum depends only on the row's counts.
template <int SEG, int KREG>
__device__ __forceinline__ DistOut distance_block(const DevProp& pr, const DevSys& sy, uint32_t count, const DistLds& L)
{
    uint32_t* rowc = L.rowc;
    uint32_t* colh = L.colh;
    uint32_t* wsum = L.wsum;
    uint64_t* red = L.red;
    const uint32_t tid = threadIdx.x;
    const uint32_t n_k = sy.n_k, n_de = sy.n_de;
    const auto* rec = gview(pr.rec);

    if (count > pr.rec_cap)  // record slice overflowed: counted, not stored
        return DistOut{0ull, 0ull, count, PCS_EOVERFLOW};
    const uint32_t T = count;
    // Issue priority over walking waves that share the SIMDs: a distance is
    // one workgroup on the critical path of its simulation's latency.
    __builtin_amdgcn_s_setprio(3);

    for (uint32_t i = tid; i < n_k; i += kDistThreads) rowc[i] = 0u;
    for (uint32_t i = tid; i < n_de; i += kDistThreads) colh[i] = 0u;
    if (tid == 0) *L.smax = 0u;
    __syncthreads();

    // Row histogram of the records and the deepest occupied row.
    constexpr uint32_t kHeldMax = (uint32_t)KREG * kDistThreads;
    const uint32_t nheld = min(T, kHeldMax);
    // A register past the slice holds kNoRecord (row 0xffff > PCS_MAX_NK).
    // The scatter takes the held words through an empty asm statement, so
    // that nothing the count derived from them (rows, lane masks) is kept
    // alive across the scan in place of being recomputed.
    constexpr uint32_t kNoRecord = ~0u;
    uint32_t held[KREG > 0 ? KREG : 1];
    uint32_t mr = 0u;
    // Loads are issued for every slot at an index clamped into the slice and
    // masked afterwards: a load under a lane mask would be waited for before
    // the next one is issued.
#pragma unroll
    for (int u = 0; u < KREG; ++u) {
        const uint32_t t = tid + (uint32_t)u * kDistThreads;
        held[u] = kNoRecord;
        if ((uint32_t)u * kDistThreads < nheld) {
            const uint32_t v = rec[min(t, nheld - 1u)];
            held[u] = (t < nheld) ? v : kNoRecord;
        }
    }
#pragma unroll
    for (int u = 0; u < KREG; ++u) {
        if (held[u] != kNoRecord) {
            const uint32_t r = held[u] >> 16;
            atomicAdd(&rowc[r], 1u);
            mr = max(mr, r);
        }
    }
    // Past the held records: each thread keeps kLoads loads in flight
    // (unmasked, at indices clamped into the slice); with no records held in
    // registers (a serving request's distance, alone on the GPU) 12, which
    // still fits its registers without spills.
    constexpr int kLoads = KREG == 0 ? 12 : kDistLoads;
    for (uint32_t t0 = nheld + tid; t0 < T; t0 += kLoads * kDistThreads) {
        uint32_t v[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; ++u) v[u] = rec[min(t0 + (uint32_t)u * kDistThreads, T - 1u)];
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
            if (t0 + (uint32_t)u * kDistThreads < T) {
                const uint32_t r = v[u] >> 16;
                atomicAdd(&rowc[r], 1u);
                mr = max(mr, r);
            }
        }
    }
    atomicMax(L.smax, mr);
    __syncthreads();

    // Exclusive scan of the row counts: rowc[r] becomes the first slot of row r.
    {
        const uint32_t ch = (n_k + kDistThreads - 1) / kDistThreads;
        const uint32_t b0 = min(tid * ch, n_k), b1 = min(b0 + ch, n_k);
        uint32_t run = 0u;
        for (uint32_t i = b0; i < b1; ++i) run += rowc[i];
        uint32_t tot;
        uint32_t ex = block_excl_scan(run, wsum, tot);
        for (uint32_t i = b0; i < b1; ++i) {
            const uint32_t c = rowc[i];
            rowc[i] = ex;
            ex += c;
        }
    }
    __syncthreads();

    // Scatter bins by row; afterwards rowc[r] is one past the last slot of row r.
    const bool in_lds = T <= L.srtl_cap;
    uint16_t* srtl = L.srtl;
    auto* srtg = gview(reinterpret_cast<uint16_t*>(pr.srt));
    auto put = [&](uint32_t v) {
        const uint32_t pos = atomicAdd(&rowc[v >> 16], 1u);
        if (in_lds) srtl[pos] = (uint16_t)v;
        else srtg[pos] = (uint16_t)v;
    };
#pragma unroll
    for (int u = 0; u < KREG; ++u) {
        uint32_t v = held[u];
        asm volatile("" : "+v"(v));
        if (v != kNoRecord) put(v);
    }
    for (uint32_t t0 = nheld + tid; t0 < T; t0 += kLoads * kDistThreads) {
        uint32_t v[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; ++u) v[u] = rec[min(t0 + (uint32_t)u * kDistThreads, T - 1u)];
#pragma unroll
        for (int u = 0; u < kLoads; ++u)
            if (t0 + (uint32_t)u * kDistThreads < T) put(v[u]);
    }
    __threadfence_block();
    __syncthreads();
    const uint32_t last = uniform(max(*L.smax, sy.maxrow_obs));  // < n_k
    const uint32_t j0 = tid * SEG;
    uint64_t acc = 0ull, acc_hi = 0ull;
    const bool l1 = sy.dist_kind == PCS_DIST_L1;
    const auto* R = gview(sy.obs_sat);
    const uint32_t maxrow_obs = sy.maxrow_obs;
    const uint32_t lane = tid % kWave;
#if PCS_TRACE
    const uint64_t t_sweep = wall_clock64();
#endif

    // The row sweep, compiled for bins sorted into LDS and for bins in HBM.
    // Every row issues the same global loads on every path (the observed row
    // PF rows ahead and, bins in HBM, the next row's sparse bins): the
    // compiler can then wait for a row's own loads, issued rows earlier,
    // instead of for every load in flight, which made each row of a lone
    // distance cost a memory round trip.
    auto sweep = [&](auto bins_in_lds) {
        constexpr bool IN_LDS = decltype(bins_in_lds)::value;
        auto bin_at = [&](uint32_t t) -> uint32_t {
            if constexpr (IN_LDS) return (uint32_t)srtl[t];
            else return (uint32_t)srtg[t];
        };
        uint32_t Ccol[SEG];
#pragma unroll
        for (int m = 0; m < SEG; ++m) Ccol[m] = 0u;

        // Observed rows are loaded PF rows ahead of their use into a ring of
        // register slots (slot d holds row i0 + d). Past the deepest observed
        // row the table repeats that row, so a refill asks for row
        // min(i + PF, maxrow_obs): no load leaves the table, and the rows past
        // it are all served from one row in cache. Columns past n_de load the
        // last column and are not summed.
        constexpr int PF = 2;
        uint32_t rb[PF][SEG];
        auto load_row = [&](uint32_t (&dst)[SEG], uint32_t row) {
            const auto* Rrow = R + (uint64_t)row * n_de;
#pragma unroll
            for (int m = 0; m < SEG; ++m) dst[m] = Rrow[min(j0 + (uint32_t)m, n_de - 1u)];
        };
#pragma unroll
        for (int d = 0; d < PF; ++d) load_row(rb[d], min((uint32_t)d, maxrow_obs));

        // Row r's slots [s0, s1) and the lane's slot s0 + lane of it (clamped
        // into the slice), read one row ahead of their use, so that a row's
        // bins arrive while the row before is summed. Lanes past the row's
        // records hold some other bin: only lanes below the count are read,
        // and only in a row of at most kDistSparseRow records. No slice, no
        // read (the LDS form may then have no room for bins at all).
        struct RowAhead {
            uint32_t bin, s0, s1;
        };
        auto row_ahead = [&](uint32_t r) -> RowAhead {
            const uint32_t rr = min(r, last);
            const uint32_t a = uniform((rr == 0u) ? 0u : rowc[rr - 1u]), e = uniform(rowc[rr]);
            uint32_t b = 0u;
            if (T != 0u) b = bin_at(min(a + lane, T - 1u));
            return RowAhead{b, a, e};
        };
        // Column pairs (j0 + 2p, j0 + 2p + 1) as 16-bit halves (columns and
        // per-row counts stay below 2^15): jp holds j0 + 1 and j0 + 2, and
        // pair p compares them with the bin less 2p.
        constexpr int NP = (SEG + 1) / 2;
        const u16x2 jp{(uint16_t)(j0 + 1u), (uint16_t)(j0 + 2u)};
        struct RowSum {
            uint64_t lo, hi;
        };
        // Row i with records in slots [s0, s1) (wave-uniform) and this lane's
        // slot of it: the records go into the column counts, then the row's
        // sum against observed row values rv is returned.
        auto do_row = [&](uint32_t s0, uint32_t s1, uint32_t mine, const uint32_t (&rv)[SEG]) -> RowSum {
            const uint32_t nr = s1 - s0;
            if (nr > 0u && nr <= (uint32_t)kDistSparseRow) {
                // A few records: every wave holds them (one per lane) and
                // each thread counts those at or left of its columns, for a
                // column pair at once: bin - (column + 1) is negative, the
                // half's top bit, exactly when the bin is at or left of the
                // column (it lies in [-2^12, 2^12) as a 16-bit difference).
                u16x2 h[NP];
#pragma unroll
                for (int p = 0; p < NP; ++p) h[p] = u16x2{0, 0};
                for (uint32_t q = 0u; q < nr; ++q) {
                    const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)mine, (int)q);
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        const uint16_t bp = (uint16_t)(b - 2u * (uint32_t)p);
                        h[p] += (u16x2)(u16x2{bp, bp} - jp) >> (uint16_t)15;
                    }
                }
#pragma unroll
                for (int m = 0; m < SEG; ++m) Ccol[m] += (m & 1) ? (uint32_t)h[m / 2].y : (uint32_t)h[m / 2].x;
            } else if (nr > 0u) {
                // four bins per thread read (at indices clamped into the row)
                // before their atomics
                for (uint32_t t = s0 + tid; t < s1; t += 4u * kDistThreads) {
                    uint32_t b[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) b[u] = bin_at(min(t + (uint32_t)u * kDistThreads, s1 - 1u));
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (t + (uint32_t)u * kDistThreads < s1) atomicAdd(&colh[b[u]], 1u);
                }
                __syncthreads();
                uint32_t pre[SEG];
                uint32_t run = 0u;
#pragma unroll
                for (int m = 0; m < SEG; ++m) {
                    run += (j0 + m < n_de) ? colh[j0 + m] : 0u;
                    pre[m] = run;
                }
                uint32_t tot;
                const uint32_t ex = block_excl_scan(run, wsum, tot);
#pragma unroll
                for (int m = 0; m < SEG; ++m) Ccol[m] += ex + pre[m];
#pragma unroll
                for (int m = 0; m < SEG; ++m)
                    if (j0 + m < n_de) colh[j0 + m] = 0u;
                __syncthreads();
            }
            uint64_t racc = 0ull, racc_hi = 0ull;
#pragma unroll
            for (int m = 0; m < SEG; ++m) {
                if (j0 + m < n_de) {
                    const uint32_t dd = Ccol[m] > rv[m] ? Ccol[m] - rv[m] : rv[m] - Ccol[m];
                    add128(racc, racc_hi, l1 ? (uint64_t)dd : (uint64_t)dd * dd, 0ull);
                }
            }
            return RowSum{racc, racc_hi};
        };
        // A row's sum counted for w rows: acc += w r in three 32 x 32 -> 64-bit
        // multiply-adds (r.hi < SEG: a row sums SEG squares below 2^64).
        auto add_rows = [&](const RowSum& r, uint32_t w) {
            const uint64_t p0 = (uint64_t)(uint32_t)r.lo * w;
            const uint64_t p1 = (uint64_t)(uint32_t)(r.lo >> 32) * w + (p0 >> 32);
            const uint64_t ph = (uint64_t)(uint32_t)r.hi * w + (p1 >> 32);
            add128(acc, acc_hi, (p1 << 32) | (uint32_t)p0, ph);
        };

        // Rows 0..min(last, maxrow_obs): the observed table changes from row
        // to row, so every row is summed; rows past `last` repeat row `last`
        // in both tables.
        const uint32_t p1 = min(last, maxrow_obs);
        RowSum rs{0ull, 0ull};
        RowAhead next = row_ahead(0u);
        for (uint32_t i0 = 0; i0 <= p1; i0 += PF) {
#pragma unroll
            for (int d = 0; d < PF; ++d) {
                const uint32_t i = i0 + (uint32_t)d;
                if (i > p1) break;
                const RowAhead cur = next;
                next = row_ahead(i + 1u);
                rs = do_row(cur.s0, cur.s1, cur.bin, rb[d]);
                if (i == last) add_rows(rs, n_k - last);
                else add128(acc, acc_hi, rs.lo, rs.hi);
                // From the deepest observed row on, every slot already holds it.
                if (i < maxrow_obs) load_row(rb[d], min(i + (uint32_t)PF, maxrow_obs));
            }
        }
        // Rows maxrow_obs + 1..last: the observed table repeats row maxrow_obs
        // (every ring slot holds it), so a row without records adds what the
        // row above it did. The rows with records are found 64 at a time by
        // one ballot over the row offsets, and each one's slots are read
        // while the one before it is summed. A row's sum is added once the
        // next row with records is known, counted for itself and the rows
        // without records after it (row maxrow_obs, already added once, for
        // those only). Row `last` holds records (it is the deepest record
        // row), so T > 0 here and every search ends at or before it.
        if (last > maxrow_obs) {
            struct RecRow {
                uint32_t i, s0, s1, bin;
            };
            uint32_t base = maxrow_obs + 1u - (uint32_t)kWave, bg = 0u, e = 0u;
            uint64_t mask = 0ull;
            auto next_rec_row = [&]() -> RecRow {
                while (mask == 0ull && base + (uint32_t)kWave <= last) {
                    base += (uint32_t)kWave;
                    const uint32_t r = base + lane, rr = min(r, last);
                    e = rowc[rr];
                    bg = rowc[rr - 1u];
                    mask = ballot(r <= last && e != bg);
                }
                if (mask == 0ull) return RecRow{last, T, T, 0u};  // not reached
                const uint32_t l = (uint32_t)__builtin_ctzll(mask);
                mask &= mask - 1ull;
                const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)bg, (int)l);
                const uint32_t z = (uint32_t)__builtin_amdgcn_readlane((int)e, (int)l);
                return RecRow{base + l, a, z, bin_at(min(a + lane, T - 1u))};
            };
            uint32_t prev = maxrow_obs, already = 1u;
            RecRow cur = next_rec_row();
            for (;;) {
                // the row and its slots are wave-uniform (kept in SGPRs, so
                // that the branches on them stay scalar)
                const uint32_t i = uniform(cur.i), s0 = uniform(cur.s0), s1 = uniform(cur.s1);
                RecRow nxt = cur;
                if (i < last) nxt = next_rec_row();
                add_rows(rs, i - prev - already);
                rs = do_row(s0, s1, cur.bin, rb[0]);
                prev = i;
                already = 0u;
                if (i >= last) break;
                cur = nxt;
            }
            add_rows(rs, n_k - last);
        }
    };
    if (in_lds) sweep(std::true_type{});
    else sweep(std::false_type{});

    // 128-bit sum over the block: across each wave by shuffles, then the
    // waves' partial sums in LDS.
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint64_t l2 = __shfl_xor(acc, o, kWave), h2 = __shfl_xor(acc_hi, o, kWave);
        add128(acc, acc_hi, l2, h2);
    }
    const uint32_t nw = kDistThreads / kWave;
    if (lane == 0u) {
        red[tid / kWave] = acc;
        red[nw + tid / kWave] = acc_hi;
    }
    __syncthreads();
    uint64_t lo = 0ull, hi = 0ull;
    for (uint32_t w = 0; w < nw; ++w) add128(lo, hi, red[w], red[nw + w]);
    __syncthreads();
    __builtin_amdgcn_s_setprio(0);
    DistOut o{lo, hi, T, PCS_OK};
#if PCS_TRACE
    o.t_sweep = t_sweep;
#endif
    return o;
}

// Records a distance workgroup holds in registers between the row count and
// the scatter (per thread; the register peak stays at or below the row
// sweep's, so occupancy is unchanged).
constexpr int kDistHeld = 32;

// Dynamic LDS: row counts (n_k_max words, rounded up to 4), one row's column
// histogram (n_de_max words), then srtl_cap 16-bit sorted bins; sized to the
// models of the launch and to the occupancy the registers allow.
// Occupancy targets per column-segment width: 8 waves per SIMD (64 VGPRs)
// for histograms up to 512 columns and 6 for up to 1,024, which the register
// allocation meets without spills; wider segments take what they need.
template <int SEG>
__global__ __launch_bounds__(kDistThreads) __attribute__((amdgpu_waves_per_eu(SEG <= 2 ? 8 : (SEG == 4 ? 6 : 1))))
void distance_kernel(DistArgs a)
{
    extern __shared__ uint32_t dist_lds[];
    uint32_t* rowc = dist_lds;
    uint32_t* colh = dist_lds + a.colh_off;
    uint16_t* srtl = reinterpret_cast<uint16_t*>(dist_lds + a.srtl_off);
    __shared__ uint32_t wsum[kDistThreads / kWave];
    __shared__ uint32_t smax;
    __shared__ uint64_t red[kDistRed];

    const uint32_t P = a.prop_lo + blockIdx.x;
    const DevProp& pr = a.props[P];
    const DistOut o = distance_block<SEG, kDistHeld>(pr, a.sys[pr.sys_id], a.rec_count[P],
                                                     DistLds{rowc, colh, wsum, &smax, red, srtl, a.srtl_cap});
    if (threadIdx.x == 0) {  // an overflowed slice is rerun by the host
        a.exact[P] = dist_exact(o);
        a.dist[P] = dist_log10(o);
        a.tracks[P] = o.tracks;
        a.status[P] = o.status;
    }
}

// ------------------------------------------------------------ request loop
//
// serve_loop_kernel is the device-resident form of the serving loop: a
// persistent grid of 4-wave groups fed from a ring in pinned host memory.
// Group 0 is the admitter: one wave that copies published requests into
// their HBM slots in request order, ahead of the workers. Every other group is
// a worker: wave 0 claims up to 4 tiles of the oldest request with unclaimed
// tiles, each wave walks one of them, and the group that completes a
// request's last tile computes its distance with the block algorithm above
// and writes the result to host memory. Every wait is bounded: groups leave
// on the quit flag (set by the admitter when the host raises stop, or by a
// worker after idle_ticks in which no group claimed a tile).

enum : uint32_t { kLoopIdle = 0u, kLoopWork = 1u, kLoopExit = 2u };

__device__ __forceinline__ uint64_t bcast64(uint64_t v)
{
    return ((uint64_t)uniform((uint32_t)(v >> 32)) << 32) | (uint64_t)uniform((uint32_t)v);
}

__device__ __forceinline__ uint64_t ld_agent64(const uint64_t* p)
{
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t ld_agent32(const uint32_t* p)
{
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t ld_system64(const uint64_t* p)
{
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ uint32_t ld_system32(const uint32_t* p)
{
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void st_agent32(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_agent64(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool cas_agent64(uint64_t* p, uint64_t expect, uint64_t want)
{
    return __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-uniform copy of *p: vector loads (slot memory is written inside the
// kernel, so it may not go through the scalar cache) broadcast from lane 0.
template <typename T>
__device__ __forceinline__ T load_uniform(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "word-sized struct");
    uint32_t w[sizeof(T) / 4];
    const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; ++i) w[i] = uniform(src[i]);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

// Admitter, wave-wide: copy request q into its slot if the host has published
// it and the slot's previous request (q - R) has finished. Only the admitter
// writes slot contents, state 4q+1 and 4q+2, in request order.
__device__ bool loop_admit(const LoopArgs& a, uint64_t q)
{
    const uint32_t slot = (uint32_t)(q % a.R);
    LoopSlot& S = a.slots[slot];
    const bool l0 = lane_id() == 0u;
    const uint64_t expect = q < a.R ? 0ull : 4ull * (q - a.R) + 3ull;
    const LoopReq& Q = a.ring[slot];
    // The slot's state and the ring entry's sequence number, both loads in
    // flight at once: an idle poll costs one round trip to host memory.
    uint64_t st = 0ull, sq = 0ull;
    if (l0) {
        st = ld_agent64(&S.state);
        sq = ld_system64(&Q.seq);
    }
    if (bcast64(st) != expect || bcast64(sq) != q + 1ull) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the ring entry was written before its seq
    if (l0) st_agent64(&S.state, 4ull * q + 1ull);
    // One round of host reads: the proposal (36 words), n_videos, tiles and,
    // in the wave's other lanes, the first words of the video table (all of
    // it for up to 7 videos); a longer table is read in a second round.
    constexpr uint32_t kPropWords = sizeof(DevProp) / 4u;
    constexpr uint32_t kVid0 = kPropWords + 2u;  // first lane holding a video-table word
    static_assert(kVid0 < (uint32_t)kWave, "proposal words and counts fit one wave");
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&Q.prop);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&S.prop);
    const uint32_t lane = lane_id();
    const uint32_t vcap = 3u * (a.maxv + 1u);  // words of a slot's video table
    const uint32_t* vs = a.ring_vid + (uint64_t)slot * vcap;
    uint32_t* vd = a.slot_vid + (uint64_t)slot * vcap;
    uint32_t w = 0u;
    if (lane < kPropWords) w = src[lane];
    else if (lane == kPropWords) w = Q.n_videos;
    else if (lane == kPropWords + 1u) w = Q.tiles;
    else if (lane - kVid0 < vcap) w = vs[lane - kVid0];
    if (lane < kPropWords) dst[lane] = w;
    const uint32_t nv = min(__shfl(w, (int)kPropWords, kWave), a.maxv);
    const uint32_t tiles = __shfl(w, (int)kPropWords + 1, kWave);
    const uint32_t vw = 3u * (nv + 1u);
    if (lane >= kVid0 && lane - kVid0 < vw) vd[lane - kVid0] = w;
    for (uint32_t i = (uint32_t)kWave - kVid0 + lane; i < vw; i += kWave) vd[i] = vs[i];
    // A request of few tiles admitted onto a quiet loop has each tile walked by
    // a whole group: a lone wave's walk is bound by its own instruction
    // stream, and the group's walk issues more instructions per tile, which a
    // loaded GPU pays for in throughput.
    uint32_t coop = 0u;
    if (l0 && a.coop_max_tiles != 0u && tiles <= a.coop_max_tiles)
        coop = q - ld_agent64(&a.ctl->finished) <= (uint64_t)a.coop_max_open ? 1u : 0u;
    if (l0) {
#if PCS_TRACE
        S.t_admit = wall_clock64();
        S.t_claim = 0ull;
#endif
        S.tiles = tiles;
        S.coop = coop;
        S.n_videos = nv;
        S.tiles_done = 0u;
        S.rec_count = 0u;
        st_agent64(&S.claim, (q & 0xffffffffull) << 32);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (l0) st_agent64(&S.state, 4ull * q + 2ull);
    return true;
}

__device__ void loop_admitter(const LoopArgs& a)
{
    const bool l0 = lane_id() == 0u;
    uint64_t q = bcast64(l0 ? ld_agent64(&a.ctl->admit_next) : 0ull);
    for (;;) {
        uint32_t leave = 0u;
        if (l0) {
            leave = ld_agent32(&a.ctl->quit);
            if (leave == 0u && ld_system32(a.stop) != 0u) {
                st_agent32(&a.ctl->quit, 1u);
                leave = 1u;
            }
        }
        if (uniform(leave)) return;
        uint32_t n = 0u;
        while (n < 64u && loop_admit(a, q)) {
            ++q;
            ++n;
        }
        if (n != 0u) {
            if (l0) st_agent64(&a.ctl->admit_next, q);
        } else {
            __builtin_amdgcn_s_sleep(16);
        }
    }
}

// Worker, wave-wide: take the fetch-add result `old` on a slot's claim word.
// If it names request r (the one the worker looked at), it claims tiles
// [t, t+4) of r when t < T. If the slot has moved on to a newer request q,
// the increment landed on q's word and claims q's tiles: those are resolved
// against q once the admitter has opened q, so that no tile is skipped.
__device__ uint32_t loop_resolve(const LoopArgs& a, LoopSlot& S, uint64_t r, uint32_t T, uint64_t old, uint32_t k,
                                 uint64_t& r_out, uint32_t& t0, uint32_t& nt)
{
    const bool l0 = lane_id() == 0u;
    const uint32_t q32 = (uint32_t)(old >> 32), t = (uint32_t)old;
    if (q32 == (uint32_t)r) {
        if (t >= T) return kLoopIdle;
        r_out = r;
        t0 = t;
        nt = min(k, T - t);
        return kLoopWork;
    }
    for (uint32_t spin = 0; spin < (1u << 20); ++spin) {
        const uint64_t st = bcast64(l0 ? ld_agent64(&S.state) : 0ull);
        const uint64_t q = st >> 2;
        if ((uint32_t)q != q32 || (st & 3ull) == 3ull) return kLoopIdle;  // q finished: the claim was surplus
        if ((st & 3ull) == 2ull) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const uint32_t Tq = uniform(l0 ? ld_agent32(&S.tiles) : 0u);
            if (t >= Tq) return kLoopIdle;
            r_out = q;
            t0 = t;
            nt = min(k, Tq - t);
            return kLoopWork;
        }
        __builtin_amdgcn_s_sleep(2);  // the admitter is copying q in
    }
    if (l0) {
        st_agent32(&a.ctl->error, 1u);
        st_agent32(&a.ctl->quit, 1u);
    }
    return kLoopExit;
}

// Index of this wave's CU: XCC id << 8 | SE, SH, CU id of HW_ID (< 2048).
__device__ __forceinline__ uint32_t hw_cu_slot()
{
    const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);   // HW_ID
    const uint32_t xcc = __builtin_amdgcn_s_getreg((31 << 11) | 20); // XCC_ID
    return (((xcc & 7u) << 8) | ((hw >> 8) & 0xffu)) & (kLoopCuSlots - 1u);
}

// Worker decision (wave 0): kLoopWork with (r, t0, nt) claimed (coop: one
// tile, walked by the whole group), kLoopIdle or kLoopExit.
__device__ uint32_t loop_decide(const LoopArgs& a, uint64_t& r_out, uint32_t& t0, uint32_t& nt, bool& coop)
{
    const bool l0 = lane_id() == 0u;
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (uniform(l0 ? ld_agent32(&a.ctl->quit) : 0u)) return kLoopExit;
        const uint64_t r = bcast64(l0 ? ld_agent64(&a.ctl->cursor) : 0ull);
        LoopSlot& S = a.slots[r % a.R];
        const uint64_t st = bcast64(l0 ? ld_agent64(&S.state) : 0ull);
        if (st == 4ull * r + 2ull) {
            const uint64_t tc = bcast64(l0 ? ld_agent64(reinterpret_cast<uint64_t*>(&S.tiles)) : 0ull);
            const uint32_t T = (uint32_t)tc;
            const uint64_t w = bcast64(l0 ? ld_agent64(&S.claim) : 0ull);
            if ((uint32_t)(w >> 32) == (uint32_t)r && (uint32_t)w < T) {
                // A request the admitter marked is claimed a tile at a time,
                // each walked by the whole group; others kLoopWaves tiles at a
                // time, a wave each.
                const uint32_t k = (tc >> 32) != 0ull ? 1u : (uint32_t)kLoopWaves;
                coop = k == 1u;
                uint64_t old = 0ull;
                if (l0) old = __hip_atomic_fetch_add(&S.claim, (uint64_t)k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                old = bcast64(old);
                const uint32_t kind = loop_resolve(a, S, r, T, old, k, r_out, t0, nt);
                if (kind != kLoopIdle) return kind;
                continue;
            }
        } else if (st != 4ull * r + 3ull) {
            return kLoopIdle;  // r not admitted yet (or the cursor was read stale)
        }
        // Every tile of r is claimed: advance once r + 1 is open.
        const uint64_t st1 = bcast64(l0 ? ld_agent64(&a.slots[(r + 1ull) % a.R].state) : 0ull);
        if (st1 != 4ull * (r + 1ull) + 2ull && st1 != 4ull * (r + 1ull) + 3ull) return kLoopIdle;
        if (l0) cas_agent64(&a.ctl->cursor, r, r + 1ull);
    }
    return kLoopIdle;
}

// The cooperative walk as a function of its own, called from the serving
// loop: its registers are allocated apart from the loop's (inlined, its four
// model classes raised the loop's register pressure and spill traffic on the
// path every request takes). The proposal and the model are read from their
// HBM copies and the stage is passed by value (empty on entry and return), so
// that nothing of the caller's has its address taken. The angle table, the
// exchange buffer and the stage are in LDS, the record slice and its counter
// in HBM.
__attribute__((noinline)) __device__ void walk_tile_coop_call(const DevProp* gp, const DevSys* gs, uint32_t F,
                                                              uint32_t N, uint32_t video, uint32_t p0, Stage st,
                                                              lds_cf2 tab_l, lds_u32 xbuf_l, lds_u32 buf_l,
                                                              glb_u32 rec_g, glb_u32 counter_g)
{
    // Scalar arguments arrive in VGPRs, as if they could differ between lanes;
    // read once, they are wave-uniform again (the chunk loop, the step count
    // and the stage fill then branch on the scalar unit, and the wave masks of
    // the walk stay in SGPRs).
    F = uniform(F);
    N = uniform(N);
    p0 = uniform(p0);
    video = uniform(video);
    st.cap = uniform(st.cap);
    st.cnt = uniform(st.cnt);
    st.slice_cap = uniform(st.slice_cap);
    const float2* tab = (const float2*)tab_l;
    uint32_t* xbuf = (uint32_t*)xbuf_l;
    st.buf = (uint32_t*)buf_l;
    st.rec = (uint32_t*)rec_g;
    st.counter = (uint32_t*)counter_g;
    const DevProp pr = load_uniform(gp);
    const DevSys sy = load_uniform(gs);
    const bool gauss = sy.detect_model == PCS_DETECT_GAUSS_Z;
    const bool zref = sy.z_boundary == PCS_Z_REFLECT;
    if (gauss) {
        if (zref) walk_tile_coop<true, true>(pr, sy, F, N, video, p0, st, tab, xbuf);
        else walk_tile_coop<true, false>(pr, sy, F, N, video, p0, st, tab, xbuf);
    } else {
        if (zref) walk_tile_coop<false, true>(pr, sy, F, N, video, p0, st, tab, xbuf);
        else walk_tile_coop<false, false>(pr, sy, F, N, video, p0, st, tab, xbuf);
    }
}

// A request's distance in the serving loop, as a function of its own: its
// row sweep keeps its registers (inlined, the loop's live state pushed the
// sweep into spill reloads on every row). The working set is rebuilt here
// from the dynamic LDS base, so that its accesses stay LDS accesses; the
// proposal and the model are read from their HBM copies.
template <int SEG>
__attribute__((noinline)) __device__ DistOut serve_distance(const DevProp* gp, const DevSys* gs, uint32_t count,
                                                            uint32_t rowc_off, uint32_t nk_lds, uint32_t nde_lds,
                                                            uint32_t srtl_cap)
{
    extern __shared__ uint32_t loop_lds[];
    __shared__ uint32_t wsum[kDistThreads / kWave];
    __shared__ uint32_t smax;
    __shared__ uint64_t red[kDistRed];
    // wave-uniform again (scalar arguments arrive in VGPRs)
    count = uniform(count);
    rowc_off = uniform(rowc_off);
    nk_lds = uniform(nk_lds);
    nde_lds = uniform(nde_lds);
    srtl_cap = uniform(srtl_cap);
    uint32_t* rowc = loop_lds + rowc_off;
    const DistLds L{rowc, rowc + nk_lds, wsum, &smax, red, reinterpret_cast<uint16_t*>(rowc + nk_lds + nde_lds),
                    srtl_cap};
    const DevProp pr = load_uniform(gp);
    const DevSys sy = load_uniform(gs);
    return distance_block<SEG, 0>(pr, sy, count, L);
}

template <int SEG>
__global__ __launch_bounds__(kLoopWaves * kWave) __attribute__((amdgpu_waves_per_eu(6, 6))) void serve_loop_kernel(LoopArgs a)
{
    extern __shared__ uint32_t loop_lds[];
    __shared__ uint64_t cmd_r;
    __shared__ uint32_t cmd[6];  // kind, t0, nt, last, sleep units, coop

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid / kWave;
    if (blockIdx.x == 0u) {  // the admitter: one wave, no barriers
        if (wave == 0u) loop_admitter(a);
        return;
    }
    Stage st;
    st.buf = loop_lds + wave * a.stage_cap;
    st.cap = a.stage_cap;
    st.cnt = 0u;
    float* tabf = reinterpret_cast<float*>(loop_lds + kLoopWaves * a.stage_cap);
    load_angle_table(tabf);
    __syncthreads();
    const float2* tab = reinterpret_cast<const float2*>(tabf);
    const uint32_t rowc_off = kLoopWaves * a.stage_cap + kAngleTabFloats;
    uint32_t* rowc = loop_lds + rowc_off;  // the distance's LDS, also the group walk's exchange buffers
    const uint32_t vw = 3u * (a.maxv + 1u);
    // Polling: one idle group per CU holds the CU's poller token and polls
    // every sleep unit (~3.4 us); the CU's other idle groups back off from 1
    // to 64 units after their last work. A request that reaches a quiet GPU
    // is therefore claimed first by one group per CU and its tiles run one
    // wave per SIMD, while groups that just finished work stay responsive
    // under load. The token passes on when its holder claims work; any group
    // that finds open tiles claims them.
    const uint32_t cu = hw_cu_slot();
    bool poller = false;
    uint32_t idle_run = 0u;
    uint64_t last_work = wall_clock64();

    for (;;) {
        if (wave == 0u) {
            uint64_t r = 0ull;
            uint32_t t0 = 0u, nt = 0u;
            bool coop = false;
            uint32_t kind = loop_decide(a, r, t0, nt, coop);
            const uint64_t now = wall_clock64();
            uint32_t units = 1u;
            if (kind == kLoopWork) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // slot contents of the admitter
                last_work = now;
                idle_run = 0u;
                if (lane_id() == 0u) {
                    st_agent64(&a.ctl->last_claim, now);
                    if (poller) st_agent32(&a.cu_token[cu], 0u);
                }
                poller = false;
            } else if (kind == kLoopIdle) {
                if (!poller) {
                    uint32_t won = 0u;
                    if (lane_id() == 0u) {
                        uint32_t expect = 0u;
                        won = __hip_atomic_compare_exchange_strong(&a.cu_token[cu], &expect, 1u, __ATOMIC_RELAXED,
                                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 1u : 0u;
                    }
                    poller = uniform(won) != 0u;
                }
                units = poller ? 1u : (1u << min(idle_run, 6u));
                idle_run = min(idle_run + 1u, 6u);
                if (now - last_work > a.idle_ticks) {
                    // idle here; has any group claimed work lately?
                    const uint64_t g = bcast64(lane_id() == 0u ? ld_agent64(&a.ctl->last_claim) : 0ull);
                    if (g > last_work) last_work = g;
                    if (now - last_work > a.idle_ticks) {
                        if (lane_id() == 0u) st_agent32(&a.ctl->quit, 1u);
                        kind = kLoopExit;
                    }
                }
            }
            if (kind == kLoopExit && poller) {
                if (lane_id() == 0u) st_agent32(&a.cu_token[cu], 0u);
                poller = false;
            }
            if (lane_id() == 0u) {
                cmd[0] = kind;
                cmd[1] = t0;
                cmd[2] = nt;
                cmd[4] = units;
                cmd[5] = coop ? 1u : 0u;
                cmd_r = r;
            }
        }
        __syncthreads();
        const uint32_t kind = uniform(cmd[0]);
        if (kind == kLoopExit) break;
        if (kind == kLoopIdle) {
            const uint32_t units = uniform(cmd[4]);
            for (uint32_t i = 0; i < units; ++i) __builtin_amdgcn_s_sleep(127);
            __syncthreads();
            continue;
        }
        const uint64_t r = bcast64(cmd_r);
        const uint32_t t0 = uniform(cmd[1]), nt = uniform(cmd[2]);
        const bool coop = uniform(cmd[5]) != 0u;
        const uint32_t slot = (uint32_t)(r % a.R);
        LoopSlot& S = a.slots[slot];
#if PCS_TRACE
        if (tid == 0u && t0 == 0u) S.t_claim = wall_clock64();
#endif
        if (coop || wave < nt) {
            const uint32_t t = t0 + (coop ? 0u : wave);
            const uint32_t* vid = a.slot_vid + (uint64_t)slot * vw;
            uint32_t lo = 0u, hi = uniform(S.n_videos);
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (uniform(vid[3u * mid + 2u]) <= t) lo = mid; else hi = mid;
            }
            const DevProp pr = load_uniform(&S.prop);
            const DevSys sy = load_uniform(&a.sys[pr.sys_id]);
            st.rec = pr.rec;
            st.slice_cap = pr.rec_cap;
            st.counter = &S.rec_count;
            const uint32_t F = uniform(vid[3u * lo]), N = uniform(vid[3u * lo + 1u]);
            const uint32_t p0 = (t - uniform(vid[3u * lo + 2u])) * (uint32_t)kWave;
            const bool gauss = sy.detect_model == PCS_DETECT_GAUSS_Z;
            const bool zref = sy.z_boundary == PCS_Z_REFLECT;
            if (coop) {  // exchange buffers in the distance's room, idle while a tile is walked
                walk_tile_coop_call(&S.prop, &a.sys[pr.sys_id], F, N, lo, p0, st, (lds_cf2)tab, (lds_u32)rowc,
                                    (lds_u32)st.buf, (glb_u32)st.rec, (glb_u32)st.counter);
            } else if (gauss) {
                if (zref) walk_tile<1, true, true, false>(pr, sy, F, N, lo, p0, st, tab, 0u, 0u);
                else walk_tile<1, true, false, false>(pr, sy, F, N, lo, p0, st, tab, 0u, 0u);
            } else {
                if (zref) walk_tile<1, false, true, false>(pr, sy, F, N, lo, p0, st, tab, 0u, 0u);
                else walk_tile<1, false, false, false>(pr, sy, F, N, lo, p0, st, tab, 0u, 0u);
            }
        }
        __syncthreads();  // every wave's records are stored
        if (tid == 0u) {  // release them at agent scope with the count
            const uint32_t prev = __hip_atomic_fetch_add(&S.tiles_done, nt, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            cmd[3] = (prev + nt == S.tiles) ? 1u : 0u;
        }
        __syncthreads();
        if (uniform(cmd[3]) != 0u) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // records of every group
            const uint32_t sys_id = uniform(S.prop.sys_id);
            const uint32_t count = uniform(ld_agent32(&S.rec_count));
#if PCS_TRACE
            const uint64_t t_dist = wall_clock64();
#endif
            const DistOut o =
                serve_distance<SEG>(&S.prop, &a.sys[sys_id], count, rowc_off, a.nk_lds, a.nde_lds, a.srtl_cap);
            if (tid == 0u) {
                LoopDone& d = a.done[slot];
#if PCS_TRACE
                d.t_admit = S.t_admit;
                d.t_claim = S.t_claim;
                d.t_dist = t_dist;
                d.t_sweep = o.t_sweep;
                d.t_done = wall_clock64();
#endif
                d.exact = dist_exact(o);
                d.dist = dist_log10(o);
                d.tracks = o.tracks;
                d.status = o.status;
                d.group_walked = S.coop;
                __hip_atomic_store(&S.state, 4ull * r + 3ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&d.seq, r + 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_fetch_add(&a.ctl->finished, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------- helpers

__global__ void reset_kernel(uint32_t* rec_count, uint32_t n, uint32_t* tile_counter)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rec_count[i] = 0u;
    if (i == 0 && tile_counter) *tile_counter = 0u;
}

__global__ void dense_hist_kernel(const DevProp* props, uint32_t prop_lo, const uint32_t* rec_count, uint32_t n_k, uint32_t n_de, uint32_t* hist)
{
    const uint32_t P = prop_lo + blockIdx.y;
    const DevProp& pr = props[P];
    const uint32_t T = min(rec_count[P], pr.rec_cap);
    uint32_t* H = hist + (uint64_t)blockIdx.y * n_k * n_de;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
        const uint32_t rv = pr.rec[t];
        atomicAdd(&H[(uint64_t)(rv >> 16) * n_de + (rv & 0xffffu)], 1u);
    }
}

// Exclusive prefix sum over the 1024 threads of a block: within each wave by
// shuffles, then over the 16 wave totals by wave 0 (sh: 32 entries).
constexpr int kPrepThreads = 1024;
constexpr int kPrepWaves = kPrepThreads / kWave;

__device__ uint64_t block_scan_u64(uint64_t v, uint64_t* sh, uint64_t& total)
{
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, kWave);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == kWave - 1) sh[wave] = x;
    __syncthreads();
    if (wave == 0u) {
        uint64_t t = lane < (uint32_t)kPrepWaves ? sh[lane] : 0ull;
#pragma unroll
        for (int o = 1; o < kPrepWaves; o <<= 1) {
            const uint64_t y = __shfl_up(t, o, kWave);
            if (lane >= (uint32_t)o) t += y;
        }
        if (lane < (uint32_t)kPrepWaves) sh[kPrepWaves + lane] = t;
    }
    __syncthreads();
    const uint64_t base = wave ? sh[kPrepWaves + wave - 1] : 0ull;
    total = sh[2 * kPrepWaves - 1];
    __syncthreads();
    return base + x - v;
}

// Device-side plan of a batch (one workgroup): record slices by the
// capacity estimate, prefix-summed into the arena; video -> proposal map;
// tile prefix. Proposals (videos) go in rounds of 1024, thread t taking
// element r * 1024 + t, so that each round's loads and stores are coalesced.
__global__ __launch_bounds__(kPrepThreads) void prepare_kernel(PrepArgs a)
{
    __shared__ uint64_t sh[2 * kPrepWaves];
    const uint32_t tid = threadIdx.x;
    const DevSys& sy = a.sys[0];
    uint64_t base = 0ull;
    for (uint32_t r0 = 0; r0 < a.n; r0 += kPrepThreads) {
        const uint32_t i = r0 + tid;
        uint64_t cap = 0ull;
        if (i < a.n) {
            const pcs_proposal& p = a.in[i];
            cap = slice_capacity(sy, p.half_az, p.n_videos, a.vframes + p.first_video, a.vparticles + p.first_video,
                                 a.cap_factor);
        }
        uint64_t round_total;
        const uint64_t off = base + block_scan_u64(cap, sh, round_total);
        if (i < a.n) {
            const pcs_proposal& p = a.in[i];
            const uint64_t at = off < a.arena_cap ? off : a.arena_cap;
            const uint64_t room = a.arena_cap - at;
            DevProp d;
            d.sim_id = p.sim_id;
            d.first_video = p.first_video;
            d.n_videos = p.n_videos;
            d.sys_id = 0u;
            d.rec = a.rec + at;
            d.srt = a.srt + at;
            d.rec_cap = (uint32_t)(cap < room ? cap : room);
            d.key0 = 0u;
            d.key1 = 0u;
            d.pad = 0u;
            for (int c = 0; c < PCS_MAX_SPECIES; ++c) {
                d.half_az[c] = p.half_az[c];
                d.cum_prob[c] = p.cum_prob[c];
                d.sigma[c] = p.sigma[c];
            }
            a.props[i] = d;
            a.rec_count[i] = 0u;
            for (uint32_t v = p.first_video; v < p.first_video + p.n_videos; ++v) a.vprop[v] = i;
        }
        base += round_total;
    }
    uint64_t tbase = 0ull;
    for (uint32_t r0 = 0; r0 < a.n_videos; r0 += kPrepThreads) {
        const uint32_t v = r0 + tid;
        const uint64_t tiles = v < a.n_videos ? (a.vparticles[v] + a.tile_particles - 1u) / a.tile_particles : 0ull;
        uint64_t round_total;
        const uint64_t t = tbase + block_scan_u64(tiles, sh, round_total);
        if (v < a.n_videos) a.vtile[v] = (uint32_t)t;
        tbase += round_total;
    }
    if (tid == 0) {
        a.vtile[a.n_videos] = (uint32_t)tbase;
        a.totals[0] = base;
        a.totals[1] = tbase;
        *a.tile_counter = 0u;
    }
}

// ------------------------------------------------------- ABC-PMC proposals
//
// propose_kernel: one wave per trial t. Its random words are Philox4x32-10
// keyed by the seed with counter (word, t, slot, kPropTag): the parent from
// (kPropParent, t, 0), parameter p's normal from (p, t, 0) - an FP32
// Box-Muller normal of the walk's math: radius of word x times the cosine of
// the 16-bit angle code in the top of word y - and lane l's Poisson arrivals
// for video v from (b, t, kPropPois | v << 6 | l), b = 0, 1, ...
// (src/gpu/detail/simulation_impl.inl:256-341, src/cpu/estimate.jl:126-150).
constexpr uint32_t kPropTag = 0x70726F70u;
constexpr uint32_t kPropParent = 0xFFFFFFFFu;
constexpr uint32_t kPropPois = 0x40000000u;
constexpr int kPropReflectMax = 64;

__device__ __forceinline__ U4 prop_block(uint32_t w0, uint32_t trial, uint32_t slot, uint32_t k0, uint32_t k1)
{
    return philox10(kPropTag, trial, w0, slot, k0, k1);  // counter (w0, trial, slot, tag)
}

// Reflection into [lo, hi] of displace_ (src/gpu/detail/simulation_impl.inl:66-79),
// at most kPropReflectMax sweeps (a step of over 128 interval widths) and
// then clamped, so that every wave ends.
__device__ __forceinline__ double reflect_closed(double x, double lo, double hi)
{
    for (int i = 0; i < kPropReflectMax && (x < lo || x > hi); ++i) {
        if (x < lo) x += 2.0 * (lo - x);
        if (x > hi) x -= 2.0 * (x - hi);
    }
    return x < lo ? lo : (x > hi ? hi : x);
}

__global__ __launch_bounds__(256) void propose_kernel(ProposeArgs a)
{
    __shared__ float tabf[kAngleTabFloats];
    load_angle_table(tabf);
    __syncthreads();
    const float2* tab = reinterpret_cast<const float2*>(tabf);
    const uint32_t lane = lane_id();
    const uint32_t t = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (t >= a.n_trials) return;  // wave-uniform
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const uint32_t C = a.C, Z = a.Z, P = 2u * C + Z;

    // parent: the first member whose cumulative weight reaches u (wrand_idx_)
    uint32_t parent;
    {
        const U4 w = prop_block(kPropParent, t, 0u, k0, k1);
        const double u = ((double)(w.x >> 5) * 67108864.0 + (double)(w.y >> 6)) * 0x1p-53;
        uint32_t lo = 0u, hi = a.n_pop;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.cum_w[mid] < u) lo = mid + 1u;
            else hi = mid;
        }
        parent = lo < a.n_pop ? lo : a.n_pop - 1u;
    }

    // lane p < P: parameter p displaced by tau_p n and reflected into its bounds
    double x = 0.0;
    if (lane < P) {
        const U4 w = prop_block(lane, t, 0u, k0, k1);
        float sn, cs;
        sincos_tab(tab, (w.y >> 24) & 0xffu, (w.y >> 16) & 0xffu, sn, cs);
        const float n = bm_radius(w.x) * cs;
        double tp = 0.0;
#pragma unroll
        for (uint32_t q = 0; q < 3u * PCS_MAX_SPECIES; ++q) tp = (lane == q) ? a.tau[q] : tp;
        const uint32_t kind = lane < C ? 0u : (lane < 2u * C ? 1u : 2u);
        const double blo = kind == 0u ? a.lo[0] : (kind == 1u ? a.lo[1] : a.lo[2]);
        const double bhi = kind == 0u ? a.hi[0] : (kind == 1u ? a.hi[1] : a.hi[2]);
        x = reflect_closed(a.theta[(uint64_t)lane * a.n_pop + parent] + tp * (double)n, blo, bhi);
    }

    // components sorted by m (stable): lane c < C ranks component c, and each
    // row's lane takes its value from the component of its rank
    uint32_t rank = 0u;
#pragma unroll
    for (uint32_t c2 = 0; c2 < PCS_MAX_SPECIES; ++c2) {
        const double m2 = __shfl(x, (int)c2, kWave);
        if (c2 < C && lane < C && (m2 < x || (m2 == x && c2 < lane))) ++rank;
    }
    uint32_t target = 0xffffffffu, base = 0u;
    if (lane < C) {
        target = lane;
    } else if (lane < 2u * C) {
        target = lane - C;
        base = C;
    } else if (Z > 1u && lane < P) {
        target = lane - 2u * C;
        base = 2u * C;
    }
    uint32_t src = lane;
#pragma unroll
    for (uint32_t c2 = 0; c2 < PCS_MAX_SPECIES; ++c2) {
        const uint32_t r2 = __shfl(rank, (int)c2, kWave);
        if (c2 < C && r2 == target) src = base + c2;
    }
    const double y = __shfl(x, (int)src, kWave);
    if (lane < P) a.theta_out[(uint64_t)lane * a.n_trials + t] = y;

    // per-trial precompute (src/gpu/detail/simulation_impl.inl:306-341)
    double mq[PCS_MAX_SPECIES], cq[PCS_MAX_SPECIES], aq[PCS_MAX_SPECIES];
#pragma unroll
    for (uint32_t c = 0; c < PCS_MAX_SPECIES; ++c) {
        mq[c] = __shfl(y, (int)c, kWave);
        cq[c] = __shfl(y, (int)((C + c) & 63u), kWave);
        aq[c] = __shfl(y, (int)((2u * C + (Z > 1u ? c : 0u)) & 63u), kWave);
    }
    double csum = 0.0;
#pragma unroll
    for (uint32_t c = 0; c < PCS_MAX_SPECIES; ++c) {
        cq[c] = c < C ? exp10_d(cq[c]) : 0.0;
        if (c < C) csum += cq[c];
    }
    // lane w < 28 writes word w of the trial's pcs_proposal record
    uint32_t word = 0u;
    if (lane == 0u) {
        word = a.first_sim_id + t;
    } else if (lane == 1u) {
        word = t * a.J;
    } else if (lane == 2u) {
        word = a.J;
    } else if (lane >= 4u && lane < 28u) {
        const uint32_t f = (lane - 4u) / PCS_MAX_SPECIES, i = (lane - 4u) % PCS_MAX_SPECIES;
        float v = 0.0f;
        double acc = 0.0;
#pragma unroll
        for (uint32_t c = 0; c < PCS_MAX_SPECIES; ++c) {
            if (c < C) acc += cq[c] / csum;
            if (c == i && c < C) {
                if (f == 0u) v = 0.5f * (float)aq[c];
                else if (f == 1u) v = (c + 1u == C) ? 2.0f : (float)acc;
                else v = sqrtf((float)(2.0 * mq[c] * a.deltat));
            }
        }
        word = __float_as_uint(v);
    }
    if (lane < 28u) reinterpret_cast<uint32_t*>(a.props + t)[lane] = word;

    // per video: Poisson(lambda) = sum over the lanes of the arrivals of a
    // unit-rate Poisson process on [0, lambda / 64]
    const double mu = csum * a.volume / 64.0;
    for (uint32_t v = 0; v < a.J; ++v) {
        const uint32_t slot = kPropPois | (v << 6) | lane;
        double s = 0.0;
        uint32_t k = 0u;
        for (uint32_t blk = 0;; ++blk) {
            const U4 w = prop_block(blk, t, slot, k0, k1);
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
            bool done = false;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!done) {
                    s += (double)(0.5f * neg2log_u24(ws[j]));
                    if (!(s <= mu)) done = true;  // also ends on a NaN
                    else ++k;
                }
            }
            if (done) break;
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) k += __shfl_xor(k, o, kWave);
        if (lane == 0u) a.vparticles[(uint64_t)t * a.J + v] = k;
    }
}

// radius of word w[i] and sin/cos of the angle whose 16-bit code is w[i] >> 16
__global__ __launch_bounds__(256) void selftest_math_kernel(const uint32_t* w, float* radius, float* s, float* c,
                                                            uint64_t n)
{
    __shared__ float tabf[kAngleTabFloats];
    load_angle_table(tabf);
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = w[i];
    radius[i] = bm_radius(v);
    float sn, cs;
    sincos_tab(reinterpret_cast<const float2*>(tabf), v >> 24, (v >> 16) & 0xffu, sn, cs);
    s[i] = sn;
    c[i] = cs;
}

// Counts the floats x (every bit pattern in [lo, lo + n)) on which the walk's
// radius square root is not correctly rounded. s is the correctly rounded
// root iff mlo^2 < x < mhi^2 for the midpoints mlo, mhi between s and its two
// neighbours: they have at most 26 significant bits, so their squares and the
// comparisons are exact in FP64.
__global__ __launch_bounds__(256) void selftest_sqrt_kernel(uint32_t lo, uint32_t n, unsigned long long* bad)
{
    uint32_t mine = 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float x = __uint_as_float(lo + i);
        const uint32_t sb = __float_as_uint(sqrt_rn_domain(x));
        bool ok = sb > 0u && sb < 0x7f800000u;
        if (ok) {
            const double sd = (double)__uint_as_float(sb);
            const double mlo = 0.5 * (sd + (double)__uint_as_float(sb - 1u));
            const double mhi = 0.5 * (sd + (double)__uint_as_float(sb + 1u));
            const double xd = (double)x;
            ok = mlo * mlo < xd && xd < mhi * mhi;
        }
        mine += ok ? 0u : 1u;
    }
    if (mine != 0u) atomicAdd(bad, (unsigned long long)mine);
}

__global__ void selftest_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// pmc-standard importance kernel sums (src/gpu/detail/simulation_impl.inl:832-871):
// wsum[j] = sum_i w_old[i] exp(-1/2 sum_p ((x_new[p][j] - x_old[p][i]) / tau_p)^2),
// inputs pre-divided by tau. The old members are split into chunks of
// kPmcChunk (a 2D grid: new members x chunks, so that a large population fills
// the GPU); a block stages 256 old members of its chunk at a time in LDS and
// each thread sums them for its new member; pmc_reduce_kernel adds a new
// member's chunk sums in chunk order. The order of every sum depends only on
// the sizes, so results are reproducible.
constexpr int kPmcMaxP = 3 * PCS_MAX_SPECIES;
constexpr uint32_t kPmcChunk = 4096;

// 2^(j/64), j = 0..63, rounded to nearest FP64 (fit_math_coeffs.py), and ln2/64
// split into a head of 35 significant bits (k * head is exact for |k| < 2^18)
// and the rounded tail.
__constant__ double kExp2Tab64[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};
constexpr double kLn2d64Hi = 0x1.62e42fefc0000p-7, kLn2d64Lo = -0x1.c610ca86c3899p-43;
constexpr double k64dLn2 = 0x1.71547652b82fep+6;

// e^y for y <= 0, tab = kExp2Tab64 in LDS: y = (64 e + j) ln2 / 64 + r with
// k = 64 e + j = rint(64 y / ln2) (Cody-Waite), |r| <= ln2 / 128, e^r by its
// Taylor series to degree 5 (remainder below 4e-17 relative), then
// 2^(j/64) e^r scaled by 2^e; within 3 ulp. y is clamped at -800, where the
// result has long underflowed to 0.
__device__ __forceinline__ double exp_nonpos(double y, const double* tab)
{
    y = __builtin_fmax(y, -800.0);
    const double k = __builtin_rint(y * k64dLn2);
    double r = __builtin_fma(-k, kLn2d64Hi, y);
    r = __builtin_fma(-k, kLn2d64Lo, r);
    double p = __builtin_fma(r, 0x1.1111111111111p-7, 0x1.5555555555555p-5);
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    const int ki = (int)k;
    return __builtin_ldexp(tab[ki & 63] * p, ki >> 6);
}

// P = the number of parameter rows when it is known at compile time (1..8);
// P = 0 reads it from p_rt (up to kPmcMaxP). Dynamic LDS: the exponential's
// table (64 doubles), then the rows of 256 staged old members, then their
// weights.
template <int P>
__global__ __launch_bounds__(256) void pmc_sums_kernel(const double* xn, uint32_t n_new, const double* xo,
                                                       const double* w_old, uint32_t n_old, uint32_t p_rt, double* part)
{
    constexpr int PM = P > 0 ? P : kPmcMaxP;
    extern __shared__ double pmc_lds[];
    const uint32_t np = P > 0 ? (uint32_t)P : p_rt;
    const uint32_t tid = threadIdx.x;
    const uint32_t j = blockIdx.x * 256u + tid;
    const uint32_t i_lo = blockIdx.y * kPmcChunk, i_hi = min(n_old, i_lo + kPmcChunk);
    double* tab = pmc_lds;
    double* sx = pmc_lds + 64;
    double* shw = sx + (size_t)np * 256u;
    if (tid < 64u) tab[tid] = kExp2Tab64[tid];  // first read after the loop's barrier
    double x[PM];
#pragma unroll
    for (int p = 0; p < PM; ++p) x[p] = ((uint32_t)p < np && j < n_new) ? xn[(uint64_t)p * n_new + j] : 0.0;
    double acc = 0.0;
    for (uint32_t t0 = i_lo; t0 < i_hi; t0 += 256u) {
        const uint32_t i = t0 + tid;
        for (uint32_t p = 0; p < np; ++p) sx[p * 256u + tid] = i < i_hi ? xo[(uint64_t)p * n_old + i] : 0.0;
        shw[tid] = i < i_hi ? w_old[i] : 0.0;
        __syncthreads();
        const uint32_t m = min(256u, i_hi - t0);
        for (uint32_t ii = 0; ii < m; ++ii) {
            double s2 = 0.0;
#pragma unroll
            for (int p = 0; p < PM; ++p) {
                if (P > 0 || (uint32_t)p < np) {
                    const double d = x[p] - sx[p * 256u + ii];
                    s2 = __builtin_fma(d, d, s2);
                }
            }
            acc = __builtin_fma(shw[ii], exp_nonpos(-0.5 * s2, tab), acc);
        }
        __syncthreads();
    }
    if (j < n_new) part[(uint64_t)blockIdx.y * n_new + j] = acc;
}

// Cumulative weights in member order: the same FP64 sums as a sequential
// cumsum. One wave loads kCumLoads x 64 weights at a time (coalesced, all in
// flight), then one running sum passes through them lane by lane (readlane of
// each weight, a wave-uniform add), and each lane keeps the sum at its member.
constexpr int kCumLoads = 8;
__global__ __launch_bounds__(kWave) void cumsum_kernel(const double* w, uint32_t n, double* cum)
{
    const uint32_t lane = threadIdx.x;
    double acc = 0.0;
    for (uint32_t base = 0; base < n; base += (uint32_t)(kCumLoads * kWave)) {
        double v[kCumLoads];
#pragma unroll
        for (int u = 0; u < kCumLoads; ++u) {
            const uint32_t i = base + (uint32_t)u * kWave + lane;
            v[u] = i < n ? w[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kCumLoads; ++u) {
            const uint32_t first = base + (uint32_t)u * kWave;
            if (first >= n) break;
            const uint32_t nl = min((uint32_t)kWave, n - first);
            const uint64_t bits = __double_as_longlong(v[u]);
            double mine = 0.0;
            for (uint32_t l = 0; l < nl; ++l) {
                const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bits, (int)l);
                const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(bits >> 32), (int)l);
                acc += __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
                mine = (lane == l) ? acc : mine;
            }
            if (lane < nl) cum[first + lane] = mine;
        }
    }
}

// Parameter rows in units of their displacement scale (the pmc kernel's
// inputs): correctly rounded FP64 division, as on the host.
__global__ void scale_rows_kernel(const double* src, uint32_t n, ScaleRows sr, double* dst)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = blockIdx.y;
    if (j >= n || k >= sr.n_rows) return;
    uint32_t row = 0u;
    double t = 1.0;
#pragma unroll
    for (uint32_t q = 0; q < 3u * PCS_MAX_SPECIES; ++q) {
        row = (k == q) ? sr.rows[q] : row;
        t = (k == q) ? sr.tau[q] : t;
    }
    dst[(uint64_t)k * n + j] = src[(uint64_t)row * n + j] / t;
}

__global__ void pmc_reduce_kernel(const double* part, uint32_t n_new, uint32_t chunks, double* wsum)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_new) return;
    double s = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) s += part[(uint64_t)c * n_new + j];
    wsum[j] = s;
}

// ---------------------------------------------------------------- launchers

uint32_t pmc_chunks(uint32_t n_old) { return n_old ? (n_old + kPmcChunk - 1u) / kPmcChunk : 1u; }

hipError_t launch_pmc_sums(const double* xn, uint32_t n_new, const double* xo, const double* w_old, uint32_t n_old,
                           uint32_t P, double* part, double* wsum, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    if (n_new == 0) return hipSuccess;
    if (P == 0 || P > (uint32_t)kPmcMaxP) return hipErrorInvalidValue;
    const uint32_t chunks = pmc_chunks(n_old);
    const dim3 g((n_new + 255u) / 256u, chunks);
    const size_t lds = (64u + ((size_t)P + 1u) * 256u) * sizeof(double);
    const dim3 b(256);
    switch (P) {
    case 1: hipLaunchKernelGGL(pmc_sums_kernel<1>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    case 2: hipLaunchKernelGGL(pmc_sums_kernel<2>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    case 3: hipLaunchKernelGGL(pmc_sums_kernel<3>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    case 4: hipLaunchKernelGGL(pmc_sums_kernel<4>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    case 5: hipLaunchKernelGGL(pmc_sums_kernel<5>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    case 6: hipLaunchKernelGGL(pmc_sums_kernel<6>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    case 7: hipLaunchKernelGGL(pmc_sums_kernel<7>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    case 8: hipLaunchKernelGGL(pmc_sums_kernel<8>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    default: hipLaunchKernelGGL(pmc_sums_kernel<0>, g, b, lds, s, xn, n_new, xo, w_old, n_old, P, part); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pmc_reduce_kernel, dim3((n_new + 255u) / 256u), dim3(256), 0, s, part, n_new, chunks, wsum);
    return hipGetLastError();
}

hipError_t launch_selftest_copy(const void* src, void* dst, uint64_t bytes, hipStream_t st)
{
    (void)hipGetLastError();  // report only this launch's error
    const uint64_t n = bytes / 16u;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(selftest_copy_kernel, dim3(4096), dim3(256), 0, st, (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}

hipError_t launch_selftest_math(const uint32_t* w, float* radius, float* s, float* c, uint64_t n, hipStream_t st)
{
    (void)hipGetLastError();  // report only this launch's error
    if (n == 0) return hipSuccess;
    const uint64_t nb = (n + 255u) / 256u;
    hipLaunchKernelGGL(selftest_math_kernel, dim3((uint32_t)nb), dim3(256), 0, st, w, radius, s, c, n);
    return hipGetLastError();
}

hipError_t launch_selftest_sqrt(uint32_t lo, uint32_t n, unsigned long long* bad, hipStream_t st)
{
    (void)hipGetLastError();  // report only this launch's error
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(selftest_sqrt_kernel, dim3(8192), dim3(256), 0, st, lo, n, bad);
    return hipGetLastError();
}

hipError_t launch_reset(uint32_t* rec_count, uint32_t n, uint32_t* tile_counter, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    const uint32_t nb = (n + 255u) / 256u;
    hipLaunchKernelGGL(reset_kernel, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, rec_count, n, tile_counter);
    return hipGetLastError();
}

template <int PPL, bool PERSIST, bool GAUSS, bool ZREF>
static hipError_t launch_walk_t(const WalkArgs& a, int wpg, uint32_t grid, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    const size_t shmem = ((size_t)wpg * a.stage_cap + kAngleTabFloats) * sizeof(uint32_t);
    hipLaunchKernelGGL((walk_kernel<PPL, PERSIST, GAUSS, ZREF>), dim3(grid), dim3(wpg * kWave), shmem, s, a);
    return hipGetLastError();
}

template <int PPL, bool PERSIST>
static hipError_t launch_walk_m(const WalkArgs& a, int wpg, uint32_t grid, bool gauss, bool zref, hipStream_t s)
{
    if (gauss) return zref ? launch_walk_t<PPL, PERSIST, true, true>(a, wpg, grid, s)
                           : launch_walk_t<PPL, PERSIST, true, false>(a, wpg, grid, s);
    return zref ? launch_walk_t<PPL, PERSIST, false, true>(a, wpg, grid, s)
                : launch_walk_t<PPL, PERSIST, false, false>(a, wpg, grid, s);
}

hipError_t launch_walk(const WalkArgs& a, int ppl, int wpg, bool persistent, uint32_t grid_groups,
                       bool gauss, bool zref, hipStream_t s)
{
    if (a.tile_hi <= a.tile_lo) return hipSuccess;
    uint32_t grid = grid_groups;
    if (!persistent) {
        const uint32_t n = a.tile_hi - a.tile_lo;
        grid = (n + (uint32_t)wpg - 1u) / (uint32_t)wpg;
    }
    if (ppl == 2)
        return persistent ? launch_walk_m<2, true>(a, wpg, grid, gauss, zref, s)
                          : launch_walk_m<2, false>(a, wpg, grid, gauss, zref, s);
    return persistent ? launch_walk_m<1, true>(a, wpg, grid, gauss, zref, s)
                      : launch_walk_m<1, false>(a, wpg, grid, gauss, zref, s);
}

hipError_t launch_tracks(const WalkArgs& a, bool gauss, bool zref, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    if (a.tile_hi <= a.tile_lo) return hipSuccess;
    const uint32_t grid = (a.tile_hi - a.tile_lo + 3u) / 4u;
    const size_t shmem = ((size_t)4 * a.stage_cap + kAngleTabFloats) * sizeof(uint32_t);
    if (gauss && zref)
        hipLaunchKernelGGL((walk_kernel<1, false, true, true, true>), dim3(grid), dim3(256), shmem, s, a);
    else if (gauss)
        hipLaunchKernelGGL((walk_kernel<1, false, true, false, true>), dim3(grid), dim3(256), shmem, s, a);
    else if (zref)
        hipLaunchKernelGGL((walk_kernel<1, false, false, true, true>), dim3(grid), dim3(256), shmem, s, a);
    else
        hipLaunchKernelGGL((walk_kernel<1, false, false, false, true>), dim3(grid), dim3(256), shmem, s, a);
    return hipGetLastError();
}

// Sorted 16-bit bins that fit in the LDS a workgroup of `fn` may add to its
// `base` bytes of dynamic LDS without lowering the workgroups per CU that its
// registers and `base` allow; at most `most`, a multiple of 64, 0 if none.
static uint32_t lds_bins_free(const void* fn, int threads, size_t base, uint32_t most)
{
    int dev = 0, lds_cu = 0, lds_block = 0, n = 0;
    hipFuncAttributes fa{};
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        hipFuncGetAttributes(&fa, fn) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, base) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return 0u;
    }
    const int64_t per_block = std::min<int64_t>((int64_t)lds_cu / n, lds_block);
    const int64_t room = per_block - (int64_t)base - (int64_t)fa.sharedSizeBytes;
    uint32_t cap = room > 0 ? (uint32_t)std::min<int64_t>(room / 2, most) & ~63u : 0u;
    int n2 = 0;
    while (cap > 0u) {
        const size_t sz = base + ((size_t)cap + 1u) / 2u * sizeof(uint32_t);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n2, fn, threads, sz) == hipSuccess && n2 >= n) break;
        (void)hipGetLastError();
        cap = cap > 1024u ? (cap - 1024u) & ~63u : 0u;
    }
    return cap;
}

// LDS sorted-bin capacity of distance_kernel<SEG> per (device, base size),
// found once: the occupancy queries are host calls that do not touch a stream.
template <int SEG>
static uint32_t dist_lds_bins(size_t base)
{
    static std::mutex mu;
    static std::map<std::pair<int, size_t>, uint32_t> memo;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(dev, base);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    const uint32_t cap = lds_bins_free(reinterpret_cast<const void*>(distance_kernel<SEG>), kDistThreads, base,
                                       (uint32_t)kDistHeld * kDistThreads);
    memo.emplace(key, cap);
    return cap;
}

template <int SEG>
static void launch_distance_t(DistArgs a, size_t base, hipStream_t s)
{
    a.srtl_off = (uint32_t)(base / sizeof(uint32_t));
    a.srtl_cap = dist_lds_bins<SEG>(base);
    const size_t shmem = base + ((size_t)a.srtl_cap + 1u) / 2u * sizeof(uint32_t);
    hipLaunchKernelGGL(distance_kernel<SEG>, dim3(a.n_props), dim3(kDistThreads), shmem, s, a);
}

hipError_t launch_distance(const DistArgs& args, uint32_t n_k_max, uint32_t n_de_max, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    if (args.n_props == 0) return hipSuccess;
    if (n_k_max == 0u || n_k_max > PCS_MAX_NK || n_de_max == 0u || n_de_max > PCS_MAX_NDE)
        return hipErrorInvalidValue;
    DistArgs a = args;
    a.colh_off = (n_k_max + 3u) & ~3u;
    const size_t base = ((size_t)a.colh_off + ((n_de_max + 3u) & ~3u)) * sizeof(uint32_t);
    if (n_de_max <= 256) launch_distance_t<1>(a, base, s);
    else if (n_de_max <= 512) launch_distance_t<2>(a, base, s);
    else if (n_de_max <= 1024) launch_distance_t<4>(a, base, s);
    else if (n_de_max <= 2048) launch_distance_t<8>(a, base, s);
    else launch_distance_t<16>(a, base, s);
    return hipGetLastError();
}

bool serve_loop_coop_fits(const LoopArgs& a)
{
    return (size_t)a.nk_lds + a.nde_lds + a.srtl_cap / 2u >= (size_t)kCoopBufWords;
}

size_t serve_loop_lds(const LoopArgs& a)
{
    return ((size_t)kLoopWaves * a.stage_cap + kAngleTabFloats + a.nk_lds + a.nde_lds + (a.srtl_cap + 1u) / 2u) *
           sizeof(uint32_t);
}

template <int SEG>
static hipError_t loop_occ_t(const LoopArgs& a, int* n)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, serve_loop_kernel<SEG>, kLoopWaves * kWave,
                                                        serve_loop_lds(a));
}

template <int SEG>
static hipError_t loop_fit_t(LoopArgs& a, int* n)
{
    a.srtl_cap = 0u;
    const size_t base = serve_loop_lds(a);
    a.srtl_cap = lds_bins_free(reinterpret_cast<const void*>(serve_loop_kernel<SEG>), kLoopWaves * kWave, base,
                               1u << 16);
    return loop_occ_t<SEG>(a, n);
}

hipError_t serve_loop_fit(LoopArgs& a, uint32_t n_de_max, int* groups_per_cu)
{
    if (n_de_max <= 256) return loop_fit_t<1>(a, groups_per_cu);
    if (n_de_max <= 512) return loop_fit_t<2>(a, groups_per_cu);
    if (n_de_max <= 1024) return loop_fit_t<4>(a, groups_per_cu);
    if (n_de_max <= 2048) return loop_fit_t<8>(a, groups_per_cu);
    return loop_fit_t<16>(a, groups_per_cu);
}

hipError_t serve_loop_occupancy(const LoopArgs& a, uint32_t n_de_max, int* groups_per_cu)
{
    if (n_de_max <= 256) return loop_occ_t<1>(a, groups_per_cu);
    if (n_de_max <= 512) return loop_occ_t<2>(a, groups_per_cu);
    if (n_de_max <= 1024) return loop_occ_t<4>(a, groups_per_cu);
    if (n_de_max <= 2048) return loop_occ_t<8>(a, groups_per_cu);
    return loop_occ_t<16>(a, groups_per_cu);
}

hipError_t launch_serve_loop(const LoopArgs& a, uint32_t grid, uint32_t n_de_max, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    if (grid == 0 || n_de_max > a.nde_lds) return hipErrorInvalidValue;
    const dim3 g(grid), b(kLoopWaves * kWave);
    const size_t sh = serve_loop_lds(a);
    if (n_de_max <= 256) hipLaunchKernelGGL(serve_loop_kernel<1>, g, b, sh, s, a);
    else if (n_de_max <= 512) hipLaunchKernelGGL(serve_loop_kernel<2>, g, b, sh, s, a);
    else if (n_de_max <= 1024) hipLaunchKernelGGL(serve_loop_kernel<4>, g, b, sh, s, a);
    else if (n_de_max <= 2048) hipLaunchKernelGGL(serve_loop_kernel<8>, g, b, sh, s, a);
    else hipLaunchKernelGGL(serve_loop_kernel<16>, g, b, sh, s, a);
    return hipGetLastError();
}

hipError_t launch_prepare(const PrepArgs& a, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    hipLaunchKernelGGL(prepare_kernel, dim3(1), dim3(kPrepThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_propose(const ProposeArgs& a, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    if (a.n_trials == 0) return hipSuccess;
    hipLaunchKernelGGL(propose_kernel, dim3((a.n_trials + 3u) / 4u), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_cumsum(const double* w, uint32_t n, double* cum, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(cumsum_kernel, dim3(1), dim3(kWave), 0, s, w, n, cum);
    return hipGetLastError();
}

hipError_t launch_scale_rows(const double* src, uint32_t n, const ScaleRows& sr, double* dst, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    if (n == 0 || sr.n_rows == 0) return hipSuccess;
    hipLaunchKernelGGL(scale_rows_kernel, dim3((n + 255u) / 256u, sr.n_rows), dim3(256), 0, s, src, n, sr, dst);
    return hipGetLastError();
}

hipError_t launch_dense_hist(const DevProp* props, uint32_t prop_lo, uint32_t n_props, const uint32_t* rec_count, uint32_t max_count, uint32_t n_k, uint32_t n_de,
                             uint32_t* hist, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launch's error
    if (n_props == 0) return hipSuccess;
    uint32_t gx = (max_count + 255u) / 256u;
    gx = gx < 1u ? 1u : (gx > 1024u ? 1024u : gx);
    hipLaunchKernelGGL(dense_hist_kernel, dim3(gx, n_props), dim3(256), 0, s, props, prop_lo, rec_count,
                       n_k, n_de, hist);
    return hipGetLastError();
}

}  // namespace pcs
