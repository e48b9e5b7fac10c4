// Measurement probe (not part of the library): the cost of a grid-wide barrier inside one
// cooperative launch against the cost of a kernel boundary inside a HIP graph, at the grid sizes
// of the engine's colour passes (one barrier would replace one kernel boundary of an MCS).
// Barriers: cooperative groups' grid.sync(), and a plain counter barrier (one atomic per
// workgroup, spin on an acquire load). Both need a co-resident grid, which the cooperative
// launch guarantees (it refuses a larger grid). Build: make -C celly.jl_amd probe_grid_barrier
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include <cstdio>
namespace cg = cooperative_groups;

__global__ void barriers_kernel(int n, int *sink) {
    cg::grid_group g = cg::this_grid();
    for (int i = 0; i < n; ++i) g.sync();
    if (blockIdx.x == 0 && threadIdx.x == 0) sink[0] = n;
}
// Counter barrier: barrier i completes when every workgroup has arrived i + 1 times.
__global__ void counter_barriers_kernel(int n, unsigned *count, int *sink) {
    for (int i = 0; i < n; ++i) {
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            atomicAdd(count, 1u);
            const unsigned target = (unsigned)(i + 1) * gridDim.x;
            while (__hip_atomic_load(count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target)
                __builtin_amdgcn_s_sleep(1);
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) sink[0] = n;
}
__global__ void empty_kernel(int *sink) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sink[0] += 1;
}
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
int main() {
    int *sink; CK(hipMalloc(&sink, 4));
    hipStream_t s; CK(hipStreamCreate(&s));
    hipEvent_t a, b; CK(hipEventCreate(&a)); CK(hipEventCreate(&b));
    int dev = 0, coop = 0; CK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    int per_cu = 0; CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, barriers_kernel, 256, 0));
    int cus = 0; CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    printf("cooperative %d, blocks per CU %d, CUs %d\n", coop, per_cu, cus);
    for (int grid : {256, 1024, 2048}) {
        if (grid > per_cu * cus) continue;
        for (int n : {8, 64}) {
            void *args[] = {&n, &sink};
            CK(hipLaunchCooperativeKernel((void *)barriers_kernel, dim3(grid), dim3(256), args, 0, s));
            CK(hipStreamSynchronize(s));
            CK(hipEventRecord(a, s));
            for (int r = 0; r < 10; ++r)
                CK(hipLaunchCooperativeKernel((void *)barriers_kernel, dim3(grid), dim3(256), args, 0, s));
            CK(hipEventRecord(b, s)); CK(hipEventSynchronize(b));
            float ms; CK(hipEventElapsedTime(&ms, a, b));
            printf("grid %d x 256: %d barriers %.2f us per launch, %.3f us per barrier (incl. launch share)\n",
                   grid, n, 1e3 * ms / 10, 1e3 * ms / 10 / n);
        }
    }
    unsigned *count; CK(hipMalloc(&count, 4));
    for (int grid : {256, 1024, 2048}) {
        if (grid > per_cu * cus) continue;
        int n = 64;
        CK(hipMemset(count, 0, 4));
        void *args[] = {&n, &count, &sink};
        CK(hipLaunchCooperativeKernel((void *)counter_barriers_kernel, dim3(grid), dim3(256), args, 0, s));
        CK(hipStreamSynchronize(s));
        CK(hipMemset(count, 0, 4));
        CK(hipEventRecord(a, s));
        CK(hipLaunchCooperativeKernel((void *)counter_barriers_kernel, dim3(grid), dim3(256), args, 0, s));
        CK(hipEventRecord(b, s)); CK(hipEventSynchronize(b));
        float ms; CK(hipEventElapsedTime(&ms, a, b));
        printf("grid %d x 256: counter barrier %.3f us per barrier (64 in one launch)\n", grid, 1e3 * ms / n);
    }
    // empty kernels captured in a graph, 64 per graph
    for (int grid : {256, 2048, 4096}) {
        hipGraph_t gph; hipGraphExec_t ex;
        CK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
        for (int i = 0; i < 64; ++i) hipLaunchKernelGGL(empty_kernel, dim3(grid), dim3(256), 0, s, sink);
        CK(hipStreamEndCapture(s, &gph));
        CK(hipGraphInstantiate(&ex, gph, nullptr, nullptr, 0));
        CK(hipGraphLaunch(ex, s)); CK(hipStreamSynchronize(s));
        CK(hipEventRecord(a, s));
        for (int r = 0; r < 10; ++r) CK(hipGraphLaunch(ex, s));
        CK(hipEventRecord(b, s)); CK(hipEventSynchronize(b));
        float ms; CK(hipEventElapsedTime(&ms, a, b));
        printf("graph of 64 empty kernels %d x 256: %.3f us per kernel\n", grid, 1e3 * ms / 10 / 64);
    }
    return 0;
}
