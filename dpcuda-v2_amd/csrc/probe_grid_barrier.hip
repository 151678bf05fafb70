// Measurement behind the batched engine's launch design: the cost of one grid-wide
// barrier of a cooperative (persistent) kernel, at 256..2048 co-resident blocks of 256
// threads, against the cost of one node of a replayed hipGraph (100 one-block-busy
// kernels per graph), each per step of a dependent chain.  Not part of libdpm.so;
// `make probe` builds _build/probe_grid_barrier.  The grids are capped at the
// co-resident block count the occupancy query reports, so every barrier completes.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include <cstdio>
#include <cstdlib>
namespace cg = cooperative_groups;

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    printf("error %s at %d\n", hipGetErrorString(e_), __LINE__); exit(1); } } while (0)

__global__ void k_sync(int iters, int* out) {
    cg::grid_group g = cg::this_grid();
    for (int i = 0; i < iters; ++i) {
        if (threadIdx.x == 0 && blockIdx.x == 0) out[0] += 1;
        g.sync();
    }
}

__global__ void k_node(int* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) out[1] += 1;
}

int main() {
    CK(hipSetDevice(0));
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    int per = 0;
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_sync, 256, 0));
    const int maxg = per * prop.multiProcessorCount;
    printf("CUs %d, co-resident blocks of 256: %d per CU, %d total\n", prop.multiProcessorCount,
           per, maxg);
    int* out;
    CK(hipMalloc(&out, 64));
    CK(hipMemset(out, 0, 64));
    hipStream_t s;
    CK(hipStreamCreate(&s));
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    const int iters = 2000;
    int grids[4] = {256, 512, 1024, 2048};
    for (int gi = 0; gi < 4; ++gi) {
        int grid = grids[gi];
        if (grid > maxg) continue;
        void* args[2] = {(void*)&iters, (void*)&out};
        int warm = 10;
        void* wargs[2] = {(void*)&warm, (void*)&out};
        CK(hipLaunchCooperativeKernel((void*)k_sync, dim3(grid), dim3(256), wargs, 0, s));
        CK(hipEventRecord(a, s));
        CK(hipLaunchCooperativeKernel((void*)k_sync, dim3(grid), dim3(256), args, 0, s));
        CK(hipEventRecord(b, s));
        CK(hipEventSynchronize(b));
        float ms;
        CK(hipEventElapsedTime(&ms, a, b));
        printf("grid barrier, %d blocks: %.3f us per barrier\n", grid, ms * 1e3 / iters);
        fflush(stdout);
    }
    for (int gi = 0; gi < 4; ++gi) {
        int grid = grids[gi];
        hipGraph_t gph;
        hipGraphExec_t ex;
        CK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < 100; ++i) hipLaunchKernelGGL(k_node, dim3(grid), dim3(256), 0, s, out);
        CK(hipStreamEndCapture(s, &gph));
        CK(hipGraphInstantiate(&ex, gph, nullptr, nullptr, 0));
        CK(hipGraphLaunch(ex, s));
        CK(hipStreamSynchronize(s));
        CK(hipEventRecord(a, s));
        for (int r = 0; r < 20; ++r) CK(hipGraphLaunch(ex, s));
        CK(hipEventRecord(b, s));
        CK(hipEventSynchronize(b));
        float ms;
        CK(hipEventElapsedTime(&ms, a, b));
        printf("graph node, %d blocks: %.3f us per node\n", grid, ms * 1e3 / 2000);
        fflush(stdout);
        CK(hipGraphExecDestroy(ex));
        CK(hipGraphDestroy(gph));
    }
    int h[2];
    CK(hipMemcpy(h, out, 8, hipMemcpyDeviceToHost));
    printf("counters %d %d\n", h[0], h[1]);
    return 0;
}
