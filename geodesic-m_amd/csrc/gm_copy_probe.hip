// Plain device-to-device copy of a known size, used only to calibrate the
// FETCH_SIZE / WRITE_SIZE hardware counters against a kernel whose HBM
// traffic is known exactly: each launch reads and writes `bytes` once.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char **argv) {
    const size_t bytes = (argc > 1 ? (size_t)atoll(argv[1]) : (size_t)256) << 20;  // MiB
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    for (int it = 0; it < 5; it++) k_copy<<<4096, 256>>>(a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    printf("copy probe: 5 launches of %zu bytes read + %zu bytes written each\n", bytes, bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
