// efie_noalloc.hip -- the device scratch array of the allocation-free
// dense solve (efie_zgesv_noalloc) and the one kernel only it needs.
//
// `scratch` is a file-scope device array of kNoAllocScratchBytes, one copy
// per GPU, held until the process ends.  Host code never passes the array's
// name to a kernel or copy (in host code it is a stand-in, not a device
// address): noalloc_scratch() looks up the device address of the current
// GPU's copy once with hipGetSymbolAddress, at context set-up, and every
// later use goes through that address.
#include "efie_internal.h"

#include <mutex>

namespace efie {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int device, void** ptr) {
    static std::mutex mu;
    static void* addr[kNoAllocMaxDevices] = {};
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

// In-place transpose of the leading n x n block of a row-major complex
// matrix with leading dimension ld: one block per 32 x 32 tile pair (I <= J),
// both tiles through LDS.
__global__ __launch_bounds__(256) void k_transpose_square(double* __restrict__ A, int n, size_t ld) {
    __shared__ double2 a[32][33], b[32][33];
    const long long id = blockIdx.x;
    int J = (int)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while ((long long)(J + 1) * (J + 2) / 2 <= id) ++J;
    while ((long long)J * (J + 1) / 2 > id) --J;
    const int I = (int)(id - (long long)J * (J + 1) / 2);
    double2* M = reinterpret_cast<double2*>(A);
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int r = e >> 5, c = e & 31;
        const int i1 = 32 * I + r, j1 = 32 * J + c, i2 = 32 * J + r, j2 = 32 * I + c;
        if (i1 < n && j1 < n) a[r][c] = M[(size_t)i1 * ld + j1];
        if (i2 < n && j2 < n) b[r][c] = M[(size_t)i2 * ld + j2];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int r = e >> 5, c = e & 31;
        const int i1 = 32 * I + r, j1 = 32 * J + c, i2 = 32 * J + r, j2 = 32 * I + c;
        // (I, J) tile gets the transpose of the (J, I) tile and vice versa
        if (i1 < n && j1 < n) M[(size_t)i1 * ld + j1] = b[c][r];
        if (I != J && i2 < n && j2 < n) M[(size_t)i2 * ld + j2] = a[c][r];
    }
}

hipError_t launch_transpose_square(double* A, int n, size_t ld, hipStream_t s) {
    const long long T = (n + 31) / 32;
    hipLaunchKernelGGL(k_transpose_square, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, A, n, ld);
    return hipGetLastError();
}

}  // namespace efie
