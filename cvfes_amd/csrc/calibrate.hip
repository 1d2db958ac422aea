// Streaming calibration kernels: the rate the HBM reaches with the load and store forms of the step
// kernels (16-byte lane loads of two consecutive doubles, non-temporal or ordinary; 16-byte
// non-temporal stores).  They give the achievable ceiling the step kernels are compared with, and
// the known byte counts that calibrate the HBM traffic counters.
#include "common.h"

namespace cvfes {
namespace {

constexpr int kUnroll = 4;   // independent 16-byte loads in flight per lane and iteration

template <bool NT>
__device__ __forceinline__ gdbl2_t load2(const gdbl2_t *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// Sum of src[0 .. 2*n2) (n2 double2 items) per workgroup into partial[blockIdx.x]: 8 bytes written
// per workgroup, so the launch is read-only to within 0.01 %.
template <bool NT>
__global__ __launch_bounds__(kBlock) void k_stream_read(const gdbl2_t *__restrict__ src, int64_t n2,
                                                        double *__restrict__ partial) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    int64_t i = xcd_block() * kBlock + threadIdx.x;
    double acc[kUnroll] = {};
    for (; i + (kUnroll - 1) * stride < n2; i += kUnroll * stride) {
        gdbl2_t v[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) v[k] = load2<NT>(src + i + k * stride);
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) acc[k] += v[k].x + v[k].y;
    }
    for (; i < n2; i += stride) {
        const gdbl2_t v = load2<NT>(src + i);
        acc[0] += v.x + v.y;
    }
    double s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    __shared__ double wsum[kBlock / 64];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
        partial[blockIdx.x] = t;
    }
}

// dst = src over n2 double2 items: 16-byte loads (non-temporal when NT) and 16-byte non-temporal
// stores, as the step kernels write their output.
template <bool NT>
__global__ __launch_bounds__(kBlock) void k_stream_copy(const gdbl2_t *__restrict__ src, int64_t n2,
                                                        gdbl2_t *__restrict__ dst) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    int64_t i = xcd_block() * kBlock + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n2; i += kUnroll * stride) {
        gdbl2_t v[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) v[k] = load2<NT>(src + i + k * stride);
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) __builtin_nontemporal_store(v[k], dst + i + k * stride);
    }
    for (; i < n2; i += stride) __builtin_nontemporal_store(load2<NT>(src + i), dst + i);
}

}  // namespace
}  // namespace cvfes

using namespace cvfes;

extern "C" int64_t cvfes_stream_blocks(int64_t n_doubles) {
    if (n_doubles < 2) return kNumXcd;
    // about 16 wavefronts per CU on 256 CUs, fewer for small arrays (at least 4 iterations each)
    const int64_t items = n_doubles / 2;
    int64_t blocks = items / (int64_t(kBlock) * kUnroll * 4);
    blocks = blocks < 1024 ? blocks : 1024;
    blocks = (blocks + kNumXcd - 1) / kNumXcd * kNumXcd;
    return blocks < kNumXcd ? kNumXcd : blocks;
}

extern "C" int cvfes_stream_read(void *stream, const double *src, int64_t n_doubles, int nontemporal,
                                 double *partial) {
    CVFES_REQUIRE(src && partial, "cvfes_stream_read: null pointer");
    CVFES_REQUIRE(n_doubles >= 2 && n_doubles % 2 == 0, "cvfes_stream_read: n_doubles must be even");
    CVFES_REQUIRE(reinterpret_cast<uintptr_t>(src) % 16 == 0, "cvfes_stream_read: src not 16-byte aligned");
    const dim3 grid(static_cast<unsigned>(cvfes_stream_blocks(n_doubles))), block(kBlock);
    const auto *s = reinterpret_cast<const gdbl2_t *>(src);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nontemporal)
        hipLaunchKernelGGL(k_stream_read<true>, grid, block, 0, st, s, n_doubles / 2, partial);
    else
        hipLaunchKernelGGL(k_stream_read<false>, grid, block, 0, st, s, n_doubles / 2, partial);
    return check_launch("k_stream_read");
}

extern "C" int cvfes_stream_copy(void *stream, const double *src, int64_t n_doubles, int nontemporal,
                                 double *dst) {
    CVFES_REQUIRE(src && dst && src != dst, "cvfes_stream_copy: bad pointers");
    CVFES_REQUIRE(n_doubles >= 2 && n_doubles % 2 == 0, "cvfes_stream_copy: n_doubles must be even");
    CVFES_REQUIRE(reinterpret_cast<uintptr_t>(src) % 16 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0,
                  "cvfes_stream_copy: buffers not 16-byte aligned");
    const dim3 grid(static_cast<unsigned>(cvfes_stream_blocks(n_doubles))), block(kBlock);
    const auto *s = reinterpret_cast<const gdbl2_t *>(src);
    auto *d = reinterpret_cast<gdbl2_t *>(dst);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nontemporal)
        hipLaunchKernelGGL(k_stream_copy<true>, grid, block, 0, st, s, n_doubles / 2, d);
    else
        hipLaunchKernelGGL(k_stream_copy<false>, grid, block, 0, st, s, n_doubles / 2, d);
    return check_launch("k_stream_copy");
}
