// copy_probe.hip -- calibration of the HBM byte counters.
//
// A plain streaming copy of a known size, launched five times.  Each launch
// reads exactly `mib` MiB and writes exactly `mib` MiB, so under
//   rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) --kernel-trace -- lib/copy_probe MIB
// the ratio counter / size is the factor that turns the same counter, read on
// the engine's kernels, into HBM bytes.
//
//   usage: copy_probe [MIB]     (default 1024)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHK(x)                                                                          \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "copy_probe: %s: %s\n", #x, hipGetErrorString(e_));    \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

__global__ __launch_bounds__(256) void k_probe_fill(double2 *__restrict__ p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_double2((double)i, -(double)i);
}

__global__ __launch_bounds__(256) void k_copy_probe(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                    size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char **argv)
{
    const long mib = argc > 1 ? std::strtol(argv[1], nullptr, 10) : 1024;
    if (mib < 1 || mib > 65536) {
        std::fprintf(stderr, "copy_probe: MIB must be in 1..65536\n");
        return 2;
    }
    const size_t bytes = (size_t)mib << 20;
    const size_t n = bytes / sizeof(double2);
    double2 *src = nullptr, *dst = nullptr;
    CHK(hipMalloc(&src, bytes));
    CHK(hipMalloc(&dst, bytes));
    k_probe_fill<<<8192, 256>>>(src, n);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    CHK(hipEventRecord(e0, nullptr));
    for (int k = 0; k < 5; ++k) k_copy_probe<<<8192, 256>>>(src, dst, n);
    CHK(hipGetLastError());
    CHK(hipEventRecord(e1, nullptr));
    CHK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("copy_probe: 5 launches x %ld MiB read + %ld MiB written, %.1f GB/s\n", mib, mib,
                10.0 * (double)bytes / (ms * 1e-3) / 1e9);
    CHK(hipFree(src));
    CHK(hipFree(dst));
    return 0;
}
