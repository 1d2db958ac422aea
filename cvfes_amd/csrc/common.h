// Shared helpers of the gfx950 kernels: error reporting, launch geometry, vector loads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "cvfes_mi355x.h"

namespace cvfes {

// Thread-local text of the last failure, returned by cvfes_last_error().
std::string &last_error();

inline int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    last_error() = buf;
    return code;
}

inline int check_launch(const char *what) {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return fail(CVFES_ERR_HIP, "%s: %s", what, hipGetErrorString(err));
    return CVFES_OK;
}

#define CVFES_REQUIRE(cond, ...)                                  \
    do {                                                          \
        if (!(cond)) return ::cvfes::fail(CVFES_ERR_ARG, __VA_ARGS__); \
    } while (0)

constexpr int kBlock = 256;    // 4 wavefronts of 64 lanes
constexpr int kNumXcd = 8;     // MI355X: 8 XCDs, each with its own L2

// Grid size for `work` threads, padded to a multiple of the XCD count so that the
// XCD-contiguous remapping below is a bijection.
inline unsigned grid_for(int64_t work) {
    int64_t blocks = (work + kBlock - 1) / kBlock;
    blocks = (blocks + kNumXcd - 1) / kNumXcd * kNumXcd;
    if (blocks < kNumXcd) blocks = kNumXcd;
    return static_cast<unsigned>(blocks);
}

// Output arrays of at least this many bytes are written with non-temporal stores: they cannot
// stay in the 256 MiB Infinity Cache for the next launch anyway, and streaming them past L2 leaves
// it to the gathered neighbour data (2 % faster steps on a 315 MB output).  Smaller outputs keep
// ordinary stores so that the next step finds them in cache.
constexpr int64_t kStreamOutBytes = int64_t(256) << 20;

__host__ __device__ inline bool stream_out(int64_t n_nodes, int64_t n_smp) {
    return n_nodes * n_smp * 24 >= kStreamOutBytes;
}

// Unused LDS per 256-thread workgroup (one wave on each SIMD) that caps a kernel at w waves per
// SIMD: a CU's 160 KiB then holds w such workgroups and no more.  0 (no cap) unless 2 <= w <= 7.
inline unsigned occupancy_lds(int64_t w) {
    constexpr int64_t kLdsPerCu = 160 * 1024;
    if (w < 2 || w > 7) return 0;
    return static_cast<unsigned>(kLdsPerCu / (w + 1) + 1024);   // in (kLdsPerCu/(w+1), kLdsPerCu/w]
}

// Workgroups are dispatched round-robin over the XCDs (workgroup b runs on XCD b % 8).  Remap so
// that XCD k processes the k-th contiguous slice of the logical block range: consecutive block
// rows then share one L2, which is where their common neighbour columns are cached.
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t b = blockIdx.x;
    const int64_t per = gridDim.x / kNumXcd;
    return (b % kNumXcd) * per + b / kNumXcd;
}

template <int VW>
struct Vec {
    double v[VW];
};

template <int VW>
__device__ __forceinline__ Vec<VW> load_vec(const double *__restrict__ p) {
    Vec<VW> r;
    if constexpr (VW == 2) {
        const double2 t = *reinterpret_cast<const double2 *>(p);
        r.v[0] = t.x;
        r.v[1] = t.y;
    } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) r.v[k] = p[k];
    }
    return r;
}

template <int VW>
__device__ __forceinline__ void store_vec(double *__restrict__ p, const Vec<VW> &r) {
    if constexpr (VW == 2) {
        *reinterpret_cast<double2 *>(p) = make_double2(r.v[0], r.v[1]);
    } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) p[k] = r.v[k];
    }
}

// Pointers typed into the global address space.  Kernels whose array pointers arrive through
// memory (the batched descriptors) use them so that the compiler emits global loads with scalar
// bases instead of flat loads with 64-bit vector addresses.
#define CVFES_G __attribute__((address_space(1)))

// Pointers typed into the constant address space, for arrays that no kernel writes while the
// launch runs (structure, unit blocks, masses, loads).  A wave-uniform address then becomes a
// scalar load without the compiler having to prove that no store of the kernel can reach it,
// a proof it gives up on in large kernels; a divergent address is still a vector global load.
#define CVFES_K __attribute__((address_space(4)))

typedef double gdbl2_t __attribute__((ext_vector_type(2)));

template <int VW>
__device__ __forceinline__ Vec<VW> load_vec(const CVFES_G double *__restrict__ p) {
    Vec<VW> r;
    if constexpr (VW == 2) {
        const gdbl2_t t = *reinterpret_cast<const CVFES_G gdbl2_t *>(p);
        r.v[0] = t.x;
        r.v[1] = t.y;
    } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) r.v[k] = p[k];
    }
    return r;
}

template <int VW>
__device__ __forceinline__ void store_vec(CVFES_G double *__restrict__ p, const Vec<VW> &r) {
    if constexpr (VW == 2) {
        *reinterpret_cast<CVFES_G gdbl2_t *>(p) = gdbl2_t{r.v[0], r.v[1]};
    } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) p[k] = r.v[k];
    }
}

// Non-temporal stores of VW consecutive samples, as one 16-byte store for VW = 2 (two 8-byte
// non-temporal stores would write half lines past L2).
template <int VW>
__device__ __forceinline__ void store_vec_nt(double *__restrict__ p, const Vec<VW> &r) {
    if constexpr (VW == 2) {
        __builtin_nontemporal_store(gdbl2_t{r.v[0], r.v[1]}, reinterpret_cast<gdbl2_t *>(p));
    } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) __builtin_nontemporal_store(r.v[k], p + k);
    }
}

template <int VW>
__device__ __forceinline__ void store_vec_nt(CVFES_G double *__restrict__ p, const Vec<VW> &r) {
    if constexpr (VW == 2) {
        __builtin_nontemporal_store(gdbl2_t{r.v[0], r.v[1]}, reinterpret_cast<CVFES_G gdbl2_t *>(p));
    } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) __builtin_nontemporal_store(r.v[k], p + k);
    }
}

}  // namespace cvfes
