// Batched ensemble: many meshes of uneven size, each with its own sample count, operator, time
// step and loads, advance together in one persistent launch per step.
//
// Work is cut into wave tasks of 64 lanes of one mesh; a lane owns VW consecutive samples of one
// block row (VW = 2 for even sample counts, else 1).  A mesh whose samples fill at least 3/4 of
// whole wavefronts gets wave-uniform rows (each row padded to whole wavefronts: the mesh
// descriptor, row pointers and column indices come in through scalar loads); otherwise its rows
// are packed back to back so that no lane idles (row data through vector loads), as the
// single-mesh kernels do.  Every mesh's tasks are cut into 8 contiguous slices, one per XCD
// (workgroup b runs on XCD b mod 8); an XCD's waves walk its slice of every mesh, mesh after mesh,
// each wave a contiguous range.  So the XCDs get equal shares of every mesh whatever the meshes'
// sizes and per-task costs, and neighbouring rows (which share gathered columns) run on one XCD
// and its L2.  The row arithmetic is the single-mesh step_row (rowops.h).
//
// Buffer roles rotate with `phase` (0, 1, 2): u = state[phase], u_prev = state[(phase+2)%3],
// out = state[(phase+1)%3].  The step index used for the load tables is *counter + k_local (or
// k_local without a counter).  A captured graph holds a multiple of 3 steps plus one counter
// increment, so a replay leaves the roles and the counter ready for the next replay.
#include "rowops.h"

#include <algorithm>
#include <vector>

namespace cvfes {
namespace {

// The descriptors' pointers arrive through memory, so the compiler cannot tell that they point to
// global memory and would use flat loads with 64-bit vector addresses; the row step gets a view
// with global-address-space pointers instead (global loads with scalar bases, fewer registers).
template <class T>
__device__ __forceinline__ const CVFES_G T *as_global(const T *p) {
    return (const CVFES_G T *)p;
}

// Arrays the step only reads: constant address space, so wave-uniform reads are scalar loads.
template <class T>
__device__ __forceinline__ const CVFES_K T *as_const(const T *p) {
    return (const CVFES_K T *)p;
}

template <int OP, bool BETA, bool DAMP, int VW>
__device__ __forceinline__ void batch_task(const cvfes_batch_mesh_t &d, int phase, int64_t n,
                                           int32_t row, int64_t s, bool active) {
    GStepArgs a;
    a.k = GRowArgs{d.sys.n_nodes, d.sys.n_smp, as_const(d.sys.indptr), as_const(d.sys.indices),
                   as_global(d.sys.values), as_const(d.sys.cptr), as_const(d.sys.celem),
                   as_const(d.sys.khat), as_global(d.sys.coef), as_global(d.sys.uvalues),
                   as_const(d.sys.vblk)};
    a.u = as_global(d.state[phase]);
    a.up = as_global(d.state[(phase + 2) % 3]);
    a.out = (CVFES_G double *)d.state[(phase + 1) % 3];
    a.lm = as_const(d.sys.lm);
    a.lm_sstride = d.sys.lm_sstride;
    a.f1 = as_const(d.sys.f1);
    a.fixed = as_const(d.sys.fixed);
    a.load = GLoad{as_const(d.load.p), as_const(d.load.ga), as_const(d.load.qa),
                   as_const(d.load.gb), as_const(d.load.qb)};
    a.bdy_u = d.sys.bdy_u;
    a.dt2 = d.dt * d.dt;
    a.bdt = d.beta / d.dt;
    a.ddt = d.damp / d.dt;
    a.nt_out = false;   // ordinary stores: the ensemble measured faster without streaming them
    step_row<OP, BETA, DAMP, VW>(a, row, s, active, n);
}

// Path sets: a kernel instance only contains the operators and damping forms that occur in the
// ensemble (cvfes_batch_plan reports them), so it needs no more registers than those paths do.
template <int VW, bool STORED, bool BD>
__device__ __forceinline__ void batch_dispatch(const cvfes_batch_mesh_t &d, int phase, int64_t n,
                                               int32_t row, int64_t s, bool active) {
    const bool mf = !STORED || d.method == CVFES_METHOD_MATFREE;
    const bool beta = BD && d.beta != 0.0, damp = BD && d.damp != 0.0;
    if (mf) {
        if (beta) {
            if (damp) batch_task<kOpMatfree, true, true, VW>(d, phase, n, row, s, active);
            else batch_task<kOpMatfree, true, false, VW>(d, phase, n, row, s, active);
        } else {
            if (damp) batch_task<kOpMatfree, false, true, VW>(d, phase, n, row, s, active);
            else batch_task<kOpMatfree, false, false, VW>(d, phase, n, row, s, active);
        }
    } else if constexpr (STORED) {
        // stored values in full storage when the mesh holds them, else in symmetric storage
        if (d.sys.values) {
            if (beta) {
                if (damp) batch_task<kOpStored, true, true, VW>(d, phase, n, row, s, active);
                else batch_task<kOpStored, true, false, VW>(d, phase, n, row, s, active);
            } else {
                if (damp) batch_task<kOpStored, false, true, VW>(d, phase, n, row, s, active);
                else batch_task<kOpStored, false, false, VW>(d, phase, n, row, s, active);
            }
        } else {
            if (beta) {
                if (damp) batch_task<kOpSym, true, true, VW>(d, phase, n, row, s, active);
                else batch_task<kOpSym, true, false, VW>(d, phase, n, row, s, active);
            } else {
                if (damp) batch_task<kOpSym, false, true, VW>(d, phase, n, row, s, active);
                else batch_task<kOpSym, false, false, VW>(d, phase, n, row, s, active);
            }
        }
    }
}

template <bool ONE, bool U2, bool P2, bool V1, bool STORED, bool BD>
__device__ __forceinline__ void batch_body(const cvfes_batch_mesh_t *__restrict__ meshes, int n_mesh,
                                           int64_t n_tasks, int phase,
                                           const int64_t *__restrict__ counter, int64_t k_local,
                                           int groups_first) {
    constexpr int64_t kWavesPerBlock = kBlock / 64;
    // The grid is a multiple of 8 blocks and block b runs on XCD b mod 8: the XCD's waves walk the
    // XCD's item list (its slice of every mesh, mesh after mesh) in contiguous ranges.
    const int64_t xcd = blockIdx.x % kNumXcd;
    const int64_t waves_x = static_cast<int64_t>(gridDim.x / kNumXcd) * kWavesPerBlock;
    const int64_t lw = __builtin_amdgcn_readfirstlane(
        static_cast<int32_t>((blockIdx.x / kNumXcd) * kWavesPerBlock + threadIdx.x / 64));
    const int lane = threadIdx.x & 63;
    const int64_t items = n_tasks / kNumXcd;
    // ONE: the grid has a wavefront per item (no loop, fewer registers)
    const int64_t per = ONE ? 1 : (items + waves_x - 1) / waves_x;
    const int64_t i0 = lw * per;
    const int64_t i1 = i0 + per < items ? i0 + per : items;
    if (i0 >= i1) return;
    const int64_t n = (counter ? counter[0] : 0) + k_local;
    // Mesh of the first item: the number of meshes with xtask0 <= i0, minus one.  Each lane reads
    // one mesh's xtask0 and a ballot counts them.
    // xtask0 of every mesh, packed after the descriptors (one coalesced load per 64 meshes)
    const int64_t *__restrict__ xt = reinterpret_cast<const int64_t *>(meshes + n_mesh);
    int mi = -1;
    for (int base = 0; base < n_mesh; base += 64) {
        const int m = base + lane;
        const bool le = m < n_mesh && xt[m] <= i0;
        mi += __popcll(__ballot(le));
    }
    for (int64_t i = i0; i < i1; ++i) {
        while (mi + 1 < n_mesh && i >= xt[mi + 1]) ++mi;
        const cvfes_batch_mesh_t &d = meshes[mi];
        const int64_t local = xcd * d.slice + (i - d.xtask0);    // task within the mesh
        if (local >= d.n_mtasks) continue;                       // the last slice is short
        // lane index inside the mesh; 32-bit (cvfes_batch_plan checks n_nodes * lanes < 2^31)
        const uint32_t g = static_cast<uint32_t>(local * 64) + lane;
        const uint32_t lanes = static_cast<uint32_t>(d.lanes);
        const int64_t S = d.sys.n_smp;
        const int vw = static_cast<int>(d.vw);
        if (d.packed) {
            const uint32_t r = g / lanes;
            int64_t s = static_cast<int64_t>(g - r * lanes) * vw;
            const bool in = r < static_cast<uint32_t>(d.sys.n_nodes);
            const bool active = in && s < S;
            const int32_t row = in ? static_cast<int32_t>(r) : static_cast<int32_t>(d.sys.n_nodes - 1);
            if (!active) s = S - vw;
            if (P2 && (!V1 || vw == 2)) batch_dispatch<2, STORED, BD>(d, phase, n, row, s, active);
            else if (V1) batch_dispatch<1, STORED, BD>(d, phase, n, row, s, active);
        } else {
            int32_t row;
            int64_t s;
            if (groups_first && lanes > 64) {
                // sample group first: wave `local` takes group local / n_nodes (64*vw samples) of
                // row local % n_nodes, so the rows in flight carry half (or less) of their samples
                const uint32_t nn = static_cast<uint32_t>(d.sys.n_nodes);
                const uint32_t grp = static_cast<uint32_t>(local) / nn;
                row = __builtin_amdgcn_readfirstlane(
                    static_cast<int32_t>(static_cast<uint32_t>(local) - grp * nn));
                s = static_cast<int64_t>(grp * 64u + static_cast<uint32_t>(lane)) * vw;
            } else {
                row = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(g / lanes));
                s = static_cast<int64_t>(g - static_cast<uint32_t>(row) * lanes) * vw;
            }
            const bool active = s < S;
            if (!active) s = S - vw;
            if (U2 && (!V1 || vw == 2)) batch_dispatch<2, STORED, BD>(d, phase, n, row, s, active);
            else if (V1) batch_dispatch<1, STORED, BD>(d, phase, n, row, s, active);
        }
    }
}

// CAP: at most 64 registers (8 waves per SIMD), with spills where the paths need more; the
// autotuner decides whether the extra waves pay for the spills.
template <bool ONE, bool U2, bool P2, bool V1, bool STORED, bool BD, bool CAP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(CAP ? 8 : 1)))
void k_batch_step(const cvfes_batch_mesh_t *__restrict__ meshes, int n_mesh, int64_t n_tasks,
                  int phase, const int64_t *__restrict__ counter, int64_t k_local, int groups_first) {
    batch_body<ONE, U2, P2, V1, STORED, BD>(meshes, n_mesh, n_tasks, phase, counter, k_local,
                                            groups_first);
}

using BatchKernel = void (*)(const cvfes_batch_mesh_t *, int, int64_t, int, const int64_t *, int64_t,
                             int);

// Smallest instance whose lane paths cover the plan's: (uniform two-sample, packed two-sample,
// one-sample) in {(1,1,1), (1,1,0), (1,0,1), (1,0,0), (0,0,1)}.
template <bool ONE, bool STORED, bool BD, bool CAP>
BatchKernel batch_kernel_lanes(int64_t kinds) {
    const bool u2 = kinds & 1, v1 = kinds & 2, p2 = kinds & 16;
    if (p2) return v1 ? k_batch_step<ONE, true, true, true, STORED, BD, CAP>
                      : k_batch_step<ONE, true, true, false, STORED, BD, CAP>;
    if (u2) return v1 ? k_batch_step<ONE, true, false, true, STORED, BD, CAP>
                      : k_batch_step<ONE, true, false, false, STORED, BD, CAP>;
    return k_batch_step<ONE, false, false, true, STORED, BD, CAP>;
}

template <bool ONE, bool CAP>
BatchKernel batch_kernel_sb(int64_t kinds) {
    const bool st = kinds & 4, bd = kinds & 8;
    if (st) return bd ? batch_kernel_lanes<ONE, true, true, CAP>(kinds) : batch_kernel_lanes<ONE, true, false, CAP>(kinds);
    return bd ? batch_kernel_lanes<ONE, false, true, CAP>(kinds) : batch_kernel_lanes<ONE, false, false, CAP>(kinds);
}

// kinds (from cvfes_batch_plan): bit 0 some wave-uniform mesh with two samples per lane, bit 1
// some mesh with one sample per lane, bit 2 some stored mesh, bit 3 some mesh with beta or damp,
// bit 4 some packed mesh with two samples per lane; bit 5 (set by the caller) caps the registers.
BatchKernel batch_kernel(int64_t kinds, bool one) {
    const bool cap = kinds & 32;
    if (one) return cap ? batch_kernel_sb<true, true>(kinds) : batch_kernel_sb<true, false>(kinds);
    return cap ? batch_kernel_sb<false, true>(kinds) : batch_kernel_sb<false, false>(kinds);
}

// kinds from cvfes_batch_plan (bits 0..4, at least one lane path), the caller's register cap (bit
// 5), waves-per-SIMD cap (bits 6..8: 0 or 2..7) and sample-group-first wave order (bit 9).
bool valid_kinds(int64_t kinds) {
    const int64_t w = (kinds >> 6) & 7;
    return kinds >= 1 && kinds < 1024 && (kinds & 19) && (w == 0 || w >= 2);
}

__global__ void k_counter_add(int64_t *counter, int64_t by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += by;
}

unsigned waves_to_blocks(int64_t n_waves) {
    int64_t b = (n_waves + 3) / 4;
    b = (b + kNumXcd - 1) / kNumXcd * kNumXcd;
    return static_cast<unsigned>(b < kNumXcd ? kNumXcd : b);
}

int default_waves() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    return cus * 16;
}

}  // namespace
}  // namespace cvfes

using namespace cvfes;

extern "C" int64_t cvfes_batch_plan_bytes(int n_mesh) {
    return n_mesh < 1 ? -1 : static_cast<int64_t>(n_mesh) * (sizeof(cvfes_batch_mesh_t) + sizeof(int64_t));
}

extern "C" int cvfes_batch_plan(void *stream, cvfes_batch_mesh_t *meshes, int n_mesh,
                                void *dev_meshes, int64_t *n_tasks, int64_t *kinds) {
    CVFES_REQUIRE(meshes && n_mesh >= 1 && dev_meshes && n_tasks && kinds,
                  "cvfes_batch_plan: bad arguments");
    int64_t t = 0, k = 0, x = 0;
    for (int i = 0; i < n_mesh; ++i) {
        cvfes_batch_mesh_t &m = meshes[i];
        const cvfes_system_t &s = m.sys;
        CVFES_REQUIRE(s.n_nodes >= 1 && s.n_smp >= 1 && s.indptr && s.indices && s.lm && s.f1,
                      "cvfes_batch_plan: mesh %d incomplete", i);
        CVFES_REQUIRE(s.n_nodes * 3 < (int64_t(1) << 31), "cvfes_batch_plan: mesh %d too large", i);
        CVFES_REQUIRE(m.method == CVFES_METHOD_STORED || m.method == CVFES_METHOD_MATFREE,
                      "cvfes_batch_plan: mesh %d unknown method", i);
        CVFES_REQUIRE(m.method == CVFES_METHOD_MATFREE || s.values || (s.uvalues && s.vblk),
                      "cvfes_batch_plan: mesh %d has no stored values", i);
        CVFES_REQUIRE(m.method == CVFES_METHOD_STORED || (s.cptr && s.celem && s.khat && s.coef),
                      "cvfes_batch_plan: mesh %d lacks matrix-free data", i);
        CVFES_REQUIRE(m.state[0] && m.state[1] && m.state[2], "cvfes_batch_plan: mesh %d state", i);
        CVFES_REQUIRE(m.load.p && (!m.load.ga || m.load.qa) && (!m.load.gb || m.load.qb),
                      "cvfes_batch_plan: mesh %d load tables", i);
        CVFES_REQUIRE(m.dt > 0.0, "cvfes_batch_plan: mesh %d dt", i);
        CVFES_REQUIRE(s.lm_sstride == 0 || s.lm_sstride == 1, "cvfes_batch_plan: mesh %d lm_sstride", i);
        CVFES_REQUIRE(m.vw >= 0 && m.vw <= 3, "cvfes_batch_plan: mesh %d vw hint", i);
        // hint: 0 two samples per lane for even counts, 1 one everywhere, 2 two samples only where
        // the rows are wave-uniform with two (packed rows take one)
        const int64_t hint = m.vw;
        m.vw = (hint != 1 && s.n_smp % 2 == 0) ? 2 : 1;
        int64_t l = s.n_smp / m.vw, padded = (l + 63) / 64 * 64;
        m.packed = 4 * l >= 3 * padded ? 0 : 1;
        if (hint == 2 && m.vw == 2 && m.packed) {
            m.vw = 1;
            l = s.n_smp;
            padded = (l + 63) / 64 * 64;
            m.packed = 4 * l >= 3 * padded ? 0 : 1;
        }
        k |= m.vw == 1 ? 2 : (m.packed ? 16 : 1);
        if (m.method == CVFES_METHOD_STORED) k |= 4;
        if (m.beta != 0.0 || m.damp != 0.0) k |= 8;
        m.lanes = m.packed ? l : padded;
        CVFES_REQUIRE(s.n_nodes * m.lanes < (int64_t(1) << 31) - 64,
                      "cvfes_batch_plan: mesh %d has too many row lanes", i);
        m.task0 = t;
        m.n_mtasks = (s.n_nodes * m.lanes + 63) / 64;
        m.slice = (m.n_mtasks + kNumXcd - 1) / kNumXcd;
        m.xtask0 = x;
        t += m.n_mtasks;
        x += m.slice;
    }
    *n_tasks = x * kNumXcd;
    *kinds = k;
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::vector<int64_t> xt(n_mesh);
    for (int i = 0; i < n_mesh; ++i) xt[i] = meshes[i].xtask0;
    hipError_t err = hipMemcpyAsync(dev_meshes, meshes, sizeof(cvfes_batch_mesh_t) * n_mesh,
                                    hipMemcpyHostToDevice, st);
    if (err == hipSuccess)
        err = hipMemcpyAsync(static_cast<char *>(dev_meshes) + sizeof(cvfes_batch_mesh_t) * n_mesh,
                             xt.data(), sizeof(int64_t) * n_mesh, hipMemcpyHostToDevice, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err != hipSuccess) return fail(CVFES_ERR_HIP, "cvfes_batch_plan: %s", hipGetErrorString(err));
    return CVFES_OK;
}

extern "C" int cvfes_batch_step(void *stream, const void *dev_meshes, int n_mesh, int64_t n_tasks,
                                int phase, const int64_t *counter, int64_t k_local,
                                int64_t n_waves, int64_t kinds) {
    CVFES_REQUIRE(dev_meshes && n_mesh >= 1 && n_tasks >= 0, "cvfes_batch_step: bad arguments");
    CVFES_REQUIRE(phase >= 0 && phase < 3, "cvfes_batch_step: phase must be 0, 1 or 2");
    CVFES_REQUIRE(valid_kinds(kinds), "cvfes_batch_step: kinds must come from cvfes_batch_plan");
    if (n_tasks == 0) return CVFES_OK;
    const int64_t w = n_waves > 0 ? n_waves : default_waves();
    hipLaunchKernelGGL(batch_kernel(kinds & 63, w >= n_tasks), dim3(waves_to_blocks(w)), dim3(kBlock),
                       occupancy_lds((kinds >> 6) & 7), static_cast<hipStream_t>(stream),
                       static_cast<const cvfes_batch_mesh_t *>(dev_meshes), n_mesh, n_tasks, phase,
                       counter, k_local, static_cast<int>((kinds >> 9) & 1));
    return check_launch("k_batch_step");
}

extern "C" int cvfes_counter_add(void *stream, int64_t *counter, int64_t by) {
    CVFES_REQUIRE(counter, "cvfes_counter_add: null counter");
    hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                       counter, by);
    return check_launch("k_counter_add");
}

extern "C" int cvfes_batch_graph_create(const void *dev_meshes, int n_mesh, int64_t n_tasks,
                                        int64_t *counter, int64_t block_steps, int phase0,
                                        int64_t n_waves, int64_t kinds, void **graph) {
    CVFES_REQUIRE(valid_kinds(kinds), "cvfes_batch_graph_create: kinds must come from cvfes_batch_plan");
    CVFES_REQUIRE(dev_meshes && n_mesh >= 1 && counter && graph, "cvfes_batch_graph_create: bad arguments");
    CVFES_REQUIRE(block_steps >= 3 && block_steps % 3 == 0,
                  "cvfes_batch_graph_create: block_steps must be a positive multiple of 3");
    CVFES_REQUIRE(phase0 >= 0 && phase0 < 3, "cvfes_batch_graph_create: phase0 must be 0, 1 or 2");
    const int64_t w = n_waves > 0 ? n_waves : default_waves();
    hipStream_t cs;
    hipError_t err = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (err != hipSuccess) return fail(CVFES_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(err));
    err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (err != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return fail(CVFES_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(err));
    }
    for (int64_t k = 0; k < block_steps; ++k)
        hipLaunchKernelGGL(batch_kernel(kinds & 63, w >= n_tasks), dim3(waves_to_blocks(w)), dim3(kBlock),
                           occupancy_lds((kinds >> 6) & 7), cs,
                           static_cast<const cvfes_batch_mesh_t *>(dev_meshes), n_mesh, n_tasks,
                           static_cast<int>((phase0 + k) % 3), counter, k,
                           static_cast<int>((kinds >> 9) & 1));
    hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, cs, counter, block_steps);
    hipGraph_t g = nullptr;
    err = hipStreamEndCapture(cs, &g);
    hipGraphExec_t exec = nullptr;
    if (err == hipSuccess) err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (err != hipSuccess) return fail(CVFES_ERR_HIP, "graph capture: %s", hipGetErrorString(err));
    *graph = exec;
    return CVFES_OK;
}

extern "C" int cvfes_graph_launch(void *graph, void *stream) {
    CVFES_REQUIRE(graph, "cvfes_graph_launch: null graph");
    hipError_t err = hipGraphLaunch(static_cast<hipGraphExec_t>(graph), static_cast<hipStream_t>(stream));
    if (err != hipSuccess) return fail(CVFES_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(err));
    return CVFES_OK;
}

extern "C" int cvfes_graph_destroy(void *graph) {
    if (!graph) return CVFES_OK;
    hipError_t err = hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph));
    if (err != hipSuccess) return fail(CVFES_ERR_HIP, "hipGraphExecDestroy: %s", hipGetErrorString(err));
    return CVFES_OK;
}
