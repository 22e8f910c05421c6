// efie_mixed.hip -- kernels of the mixed-precision solve: complex FP32 LU
// of the FP64-filled Z, then FP64 iterative refinement
//   r = V - Z x (FP64),  d = (LU_32)^-1 r,  x += d
// until ||r|| <= ir_tol ||V|| (per frequency, decided on the device).
#include "efie_internal.h"

namespace efie {

// A32 = (float) A64 for an n x (n+1) row-major batch with leading dimension lda.
__global__ void k_to_f32(const double* __restrict__ A64, float* __restrict__ A32, size_t total) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < total; i += stride) {
        double2 v = reinterpret_cast<const double2*>(A64)[i];
        reinterpret_cast<float2*>(A32)[i] = make_float2((float)v.x, (float)v.y);
    }
}

// X64[b][i] = (double) A32[b][i][n] (the solved right-hand-side column).
__global__ void k_col_to_vec(const float* __restrict__ A32, size_t strideA, int n, int lda,
                             double* __restrict__ X, int batch) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || b >= batch) return;
    float2 v = reinterpret_cast<const float2*>(A32 + (size_t)b * strideA)[(size_t)i * lda + n];
    reinterpret_cast<double2*>(X + 2 * (size_t)b * n)[i] = make_double2(v.x, v.y);
}

// R = V - A64 X (one wave per row); rn2[b] += |R_i|^2.  rn2 must be zero.
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ A64, size_t strideA,
                                                  int n, int lda, const double* __restrict__ V,
                                                  size_t vStride, const double* __restrict__ X,
                                                  double* __restrict__ R, double* __restrict__ rn2,
                                                  const int* __restrict__ done) {
    const int b = blockIdx.y;
    if (done[b]) return;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double2* a = reinterpret_cast<const double2*>(A64 + (size_t)b * strideA) + (size_t)row * lda;
    const double2* x = reinterpret_cast<const double2*>(X) + (size_t)b * n;
    double sr = 0.0, si = 0.0;
    for (int j = lane; j < n; j += 64) {
        double2 av = a[j], xv = x[j];
        sr += av.x * xv.x - av.y * xv.y;
        si += av.x * xv.y + av.y * xv.x;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o, 64);
        si += __shfl_xor(si, o, 64);
    }
    if (lane == 0) {
        const double2 v = reinterpret_cast<const double2*>(V + (size_t)b * vStride)[row];
        const double rr = v.x - sr, ri = v.y - si;
        reinterpret_cast<double2*>(R)[(size_t)b * n + row] = make_double2(rr, ri);
        atomicAdd(rn2 + b, rr * rr + ri * ri);
    }
}

// ------------------------------------------------ symmetric residual ---
// For Z = Z^T (bit-exact: the record fill mirrors its tiles) R = V - Z x
// reads only the 64 x 64 tiles on and above the diagonal, half the bytes.
// k_symv_strips: one block per (matrix b, 32-row strip S) walks the 64 x 64
// tiles J >= D = S/2 (the diagonal block D is stored whole): it adds
// Z_SJ x_J to the strip's rows (kept in registers, summed over the lanes
// and waves at the end, written to Y) and, for J > D, writes Z_SJ^T x_S,
// the contribution to the rows of block J, as the partial
// P[b][S][64J..64J+63].  k_symv_finish sums Y and the partials of the
// strips above in a fixed order (no atomics: the result does not depend on
// scheduling) and forms R and ||R||^2.  Wave w owns columns 16w..16w+15 of
// every tile and all 32 rows: lane (g, c) = (lane / 16, lane % 16) takes
// column 16w + c and rows g + 4k, k < 8, so one load instruction reads four
// 256-byte row segments, and a column's sum over the strip's rows stays
// within the wave (no barrier per tile).
//
// The kernel runs beside the trailing-update GEMM of the next batch, whose
// waves take 128 registers each (four per SIMD): a wave here must stay
// within the fourth wave's 128 registers, or the GEMM keeps only two waves
// on every SIMD this kernel shares; hence 32-row strips (ROWS = 8 rows per
// thread).  A residual with nothing beside it takes 64-row strips
// (ROWS = 16): half the strips, partials and x reloads.
constexpr int kSymRowsShared = 8, kSymRowsAlone = 16;

template <int ROWS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ROWS == kSymRowsShared ? 4 : 1))) void
k_symv_strips(const double* __restrict__ A64, size_t strideA, int n, int lda, const double* __restrict__ X,
              double* __restrict__ Y, double* __restrict__ P, const int* __restrict__ done) {
    constexpr int kSymRows = ROWS, kSymStrip = 4 * ROWS;
    __shared__ double2 red[4][kSymStrip];
    const int b = blockIdx.x, S = blockIdx.y;
    if (done[b]) return;
    const int nblk = (n + 63) / 64, ns = (n + kSymStrip - 1) / kSymStrip;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int g = lane >> 4, c = lane & 15;
    const double2* Ab = reinterpret_cast<const double2*>(A64 + (size_t)b * strideA);
    const double2* xb = reinterpret_cast<const double2*>(X) + (size_t)b * n;
    const int r0 = kSymStrip * S, D = r0 / 64;
    double2 xr[kSymRows], acc[kSymRows];
#pragma unroll
    for (int k = 0; k < kSymRows; ++k) {
        const int r = r0 + g + 4 * k;
        const double2 v = xb[r < n ? r : n - 1];
        xr[k] = r < n ? v : make_double2(0.0, 0.0);
        acc[k] = make_double2(0.0, 0.0);
    }
    // tile J: acc += Z_SJ x_J; with the transposed product (J > D) also
    // Z_SJ^T x_S into P
    auto tile = [&](int J, bool tr) {
        const int col = 64 * J + 16 * w + c;
        const int cc = col < n ? col : n - 1;
        // all loads unconditional from clamped addresses, masked afterwards
        const double2 xv = xb[cc];
        const double2 xj = col < n ? xv : make_double2(0.0, 0.0);
        double2 z[kSymRows];
        if (r0 + kSymStrip <= n) {
            // whole strip inside: one lane address, rows a uniform stride apart
            const double2* zp = Ab + (size_t)(r0 + g) * lda + cc;
#pragma unroll
            for (int k = 0; k < kSymRows; ++k) z[k] = zp[(size_t)(4 * k) * lda];
        } else {
            // last strip: rows past the matrix read its last row (an opaque
            // zero keeps the clamped offsets out of registers held across
            // the tiles)
            int z0 = 0;
            asm volatile("" : "+s"(z0));
#pragma unroll
            for (int k = 0; k < kSymRows; ++k) {
                const int r = r0 + g + 4 * k + z0;
                z[k] = Ab[(size_t)(r < n ? r : n - 1) * lda + cc];
            }
        }
        double pr = 0.0, pi = 0.0;
#pragma unroll
        for (int k = 0; k < kSymRows; ++k) {
            acc[k].x += z[k].x * xj.x - z[k].y * xj.y;
            acc[k].y += z[k].x * xj.y + z[k].y * xj.x;
            if (tr) {
                pr += z[k].x * xr[k].x - z[k].y * xr[k].y;
                pi += z[k].x * xr[k].y + z[k].y * xr[k].x;
            }
        }
        if (tr) {
            // over the four row groups of the column: (g0 + g1) + (g2 + g3)
            pr += __shfl_xor(pr, 16, 64);
            pi += __shfl_xor(pi, 16, 64);
            pr += __shfl_xor(pr, 32, 64);
            pi += __shfl_xor(pi, 32, 64);
            if (g == 0 && col < n) reinterpret_cast<double2*>(P)[((size_t)b * ns + S) * n + col] = make_double2(pr, pi);
        }
    };
    tile(D, false);
    for (int J = D + 1; J < nblk; ++J) tile(J, true);
    // row sums: over the 16 columns of the wave, then over the waves
#pragma unroll
    for (int k = 0; k < kSymRows; ++k) {
        double sr = acc[k].x, si = acc[k].y;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            sr += __shfl_xor(sr, o, 64);
            si += __shfl_xor(si, o, 64);
        }
        if (c == 0) red[w][g + 4 * k] = make_double2(sr, si);
    }
    __syncthreads();
    if (threadIdx.x < kSymStrip) {
        const int r = r0 + threadIdx.x;
        const double2 a0 = red[0][threadIdx.x], a1 = red[1][threadIdx.x], a2 = red[2][threadIdx.x],
                      a3 = red[3][threadIdx.x];
        if (r < n)
            reinterpret_cast<double2*>(Y)[(size_t)b * n + r] =
                make_double2((a0.x + a1.x) + (a2.x + a3.x), (a0.y + a1.y) + (a2.y + a3.y));
    }
}

// R[b][i] = V[b][i] - (Y[b][i] + sum over the strips S of the blocks above
// row i's block of P[b][S][i]), in place over Y; rn2[b] += ||R||^2 (one
// atomic per block).  xZero: x = 0, so R = V without reading Y or P.
__global__ __launch_bounds__(256) void k_symv_finish(const double* __restrict__ V, size_t vStride, int n,
                                                     const double* __restrict__ P, double* __restrict__ R,
                                                     double* __restrict__ rn2, const int* __restrict__ done,
                                                     int xZero, int kSymStrip) {
    __shared__ double red[4];
    const int b = blockIdx.y;
    if (done[b]) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ns = (n + kSymStrip - 1) / kSymStrip;
    double e = 0.0;
    if (i < n) {
        double2* y = reinterpret_cast<double2*>(R) + (size_t)b * n + i;
        double sx = 0.0, sy = 0.0;
        if (!xZero) {
            const double2 y0 = *y;
            sx = y0.x;
            sy = y0.y;
            const double2* p = reinterpret_cast<const double2*>(P) + (size_t)b * ns * n + i;
            const int nS = (i / 64) * (64 / kSymStrip);
            for (int S = 0; S < nS; ++S) {
                const double2 q = p[(size_t)S * n];
                sx += q.x;
                sy += q.y;
            }
        }
        const double2 v = reinterpret_cast<const double2*>(V + (size_t)b * vStride)[i];
        const double rr = v.x - sx, ri = v.y - sy;
        *y = make_double2(rr, ri);
        e = rr * rr + ri * ri;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(rn2 + b, (red[0] + red[1]) + (red[2] + red[3]));
}

// done[b] once ||R|| <= tol ||V|| or the residual stopped shrinking (less
// than a factor 2 per correction: the FP64 evaluation floor of V - Z x);
// keeps the last residual ratio and counts corrections; clears rn2.
__global__ void k_ir_check(double* __restrict__ rn2, const double* __restrict__ vn2, double tol, int batch,
                           int* __restrict__ done, int* __restrict__ iters, double* __restrict__ resid,
                           int firstPass, int lastPass, int remaining) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    if (!done[b]) {
        const double ratio = sqrt(rn2[b] / vn2[b]);
        const bool stalled = !firstPass && ratio > 0.5 * resid[b];
        // remaining > 0: also stop when this step's reduction, kept up for
        // the remaining steps, would not reach tol (the caller retries)
        const bool hopeless =
            !firstPass && remaining > 0 && ratio * pow(ratio / resid[b], (double)remaining) > tol;
        resid[b] = firstPass ? ratio : fmin(ratio, resid[b]);
        if (ratio <= tol || stalled || hopeless) done[b] = 1;
        else if (!lastPass) iters[b] += 1;
    }
    rn2[b] = 0.0;
}

// R32[b][i] = (float) R[b][perm[b][i]]   (row interchanges of the LU)
__global__ void k_perm_f32(const double* __restrict__ R, const int* __restrict__ perm, int n, int batch,
                           float* __restrict__ R32, const int* __restrict__ done) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || b >= batch || done[b]) return;
    const double2 v = reinterpret_cast<const double2*>(R)[(size_t)b * n + perm[(size_t)b * n + i]];
    reinterpret_cast<float2*>(R32)[(size_t)b * n + i] = make_float2((float)v.x, (float)v.y);
}

// X += (double) D32
__global__ void k_add_f32(double* __restrict__ X, const float* __restrict__ D32, int n, int batch,
                          const int* __restrict__ done) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || b >= batch || done[b]) return;
    double2* x = reinterpret_cast<double2*>(X) + (size_t)b * n + i;
    const float2 d = reinterpret_cast<const float2*>(D32)[(size_t)b * n + i];
    double2 v = *x;
    v.x += (double)d.x;
    v.y += (double)d.y;
    *x = v;
}

__global__ void k_fill_int(int* p, int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_dbl(double* p, double v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Plain streaming copy (counter calibration).
__global__ void k_copy16(const double2* __restrict__ src, double2* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (const double2*)src, (double2*)dst, n);
    return hipGetLastError();
}

hipError_t launch_to_f32(const double* A64, float* A32, size_t complexElems, hipStream_t s) {
    const size_t blocks = std::min<size_t>((complexElems + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, A64, A32, complexElems);
    return hipGetLastError();
}

hipError_t launch_col_to_vec(const float* A32, size_t strideA, int n, int lda, double* X, int batch,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_col_to_vec, dim3((n + 255) / 256, batch), dim3(256), 0, s, A32, strideA, n, lda, X,
                       batch);
    return hipGetLastError();
}

hipError_t launch_residual(const double* A64, size_t strideA, int n, int lda, const double* V,
                           size_t vStride, const double* X, double* R, double* rn2, const int* done,
                           int batch, hipStream_t s) {
    hipLaunchKernelGGL(k_residual, dim3((n + 3) / 4, batch), dim3(256), 0, s, A64, strideA, n, lda, V, vStride,
                       X, R, rn2, done);
    return hipGetLastError();
}

hipError_t launch_residual_sym(const double* A64, size_t strideA, int n, int lda, const double* V, size_t vStride,
                               const double* X, double* R, double* P, double* rn2, const int* done, int batch,
                               int xZero, bool shared, hipStream_t s) {
    const int strip = 4 * (shared ? kSymRowsShared : kSymRowsAlone);
    if (!xZero) {
        if (shared)
            hipLaunchKernelGGL(k_symv_strips<kSymRowsShared>, dim3(batch, (n + strip - 1) / strip), dim3(256), 0, s,
                               A64, strideA, n, lda, X, R, P, done);
        else
            hipLaunchKernelGGL(k_symv_strips<kSymRowsAlone>, dim3(batch, (n + strip - 1) / strip), dim3(256), 0, s,
                               A64, strideA, n, lda, X, R, P, done);
    }
    hipLaunchKernelGGL(k_symv_finish, dim3((n + 255) / 256, batch), dim3(256), 0, s, V, vStride, n, P, R, rn2, done,
                       xZero, strip);
    return hipGetLastError();
}

// sized for the narrower (shared) strips
size_t residual_sym_bytes(int n, int batch) {
    const int strip = 4 * kSymRowsShared;
    return sizeof(double) * 2 * (size_t)batch * ((n + strip - 1) / strip) * n;
}

hipError_t launch_ir_check(double* rn2, const double* vn2, double tol, int batch, int* done, int* iters,
                           double* resid, int firstPass, int lastPass, int remaining, hipStream_t s) {
    hipLaunchKernelGGL(k_ir_check, dim3((batch + 63) / 64), dim3(64), 0, s, rn2, vn2, tol, batch, done, iters,
                       resid, firstPass, lastPass, remaining);
    return hipGetLastError();
}

hipError_t launch_perm_f32(const double* R, const int* perm, int n, int batch, float* R32, const int* done,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_perm_f32, dim3((n + 255) / 256, batch), dim3(256), 0, s, R, perm, n, batch, R32, done);
    return hipGetLastError();
}

hipError_t launch_add_f32(double* X, const float* D32, int n, int batch, const int* done, hipStream_t s) {
    hipLaunchKernelGGL(k_add_f32, dim3((n + 255) / 256, batch), dim3(256), 0, s, X, D32, n, batch, done);
    return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int v, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_dbl(double* p, double v, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_dbl, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}


// Column n of every batch matrix (FP64, and the complex64 copy when Z32 is
// non-null) set to the right-hand side V (matrix b's at V + b*vStride).
__global__ __launch_bounds__(256) void k_set_rhs(const double* __restrict__ V, size_t vStride, double* __restrict__ Z,
                                                 float* __restrict__ Z32, int n, int lda, size_t zElems) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t e = (size_t)blockIdx.y * zElems + (size_t)i * lda + n;
    const double2 v = reinterpret_cast<const double2*>(V + blockIdx.y * vStride)[i];
    reinterpret_cast<double2*>(Z)[e] = v;
    if (Z32) reinterpret_cast<float2*>(Z32)[e] = make_float2((float)v.x, (float)v.y);
}

hipError_t launch_set_rhs(const double* V, double* Z, float* Z32, int n, int lda, size_t zElems, int nb,
                          hipStream_t s, size_t vStride) {
    hipLaunchKernelGGL(k_set_rhs, dim3((n + 255) / 256, nb), dim3(256), 0, s, V, vStride, Z, Z32, n, lda, zElems);
    return hipGetLastError();
}

// ------------------------------------------------------ packed batches ---
// A batch may hold the matrices of several problems at one order np: a
// problem of order n < np fills the leading n x n block and the rest is the
// identity.  k_pad_identity writes that rest: row r < n gets zeros in
// columns n..np-1, row r >= n zeros in columns 0..np-1 but a one at r (FP64
// matrix, and its complex64 copy when Z32 is set).  Every factorisation and
// solve treats the padding exactly: its entries meet the problem's only in
// products with an exact zero, so the leading block factors as the n x n
// matrix would and the padding's part of every solution is exactly zero.
// Grid (np, nf): block (r, f) writes row r of matrix f.
__global__ __launch_bounds__(256) void k_pad_identity(double* __restrict__ Z, float* __restrict__ Z32,
                                                      size_t zElems, int n, int np, int lda) {
    const int r = blockIdx.x;
    const size_t row = (size_t)blockIdx.y * zElems + (size_t)r * lda;
    for (int j = (r < n ? n : 0) + threadIdx.x; j < np; j += 256) {
        const double d = j == r ? 1.0 : 0.0;
        reinterpret_cast<double2*>(Z)[row + j] = make_double2(d, 0.0);
        if (Z32) reinterpret_cast<float2*>(Z32)[row + j] = make_float2((float)d, 0.0f);
    }
}

hipError_t launch_pad_identity(double* Z, float* Z32, size_t zElems, int n, int np, int lda, int nf, hipStream_t s) {
    if (n >= np || nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pad_identity, dim3(np, nf), dim3(256), 0, s, Z, Z32, zElems, n, np, lda);
    return hipGetLastError();
}

// out[f][i] = V[i] for i < n and 0 for n <= i < np, f < nf (interleaved
// complex, out[f] at out + 2 np f): a problem's right-hand side padded to
// the batch order.
__global__ __launch_bounds__(256) void k_pack_rhs(const double* __restrict__ V, int n, int np,
                                                  double* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const double2 v = i < n ? reinterpret_cast<const double2*>(V)[i] : make_double2(0.0, 0.0);
    reinterpret_cast<double2*>(out)[(size_t)blockIdx.y * np + i] = v;
}

hipError_t launch_pack_rhs(const double* V, int n, int np, int nf, double* out, hipStream_t s) {
    if (nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_rhs, dim3((np + 255) / 256, nf), dim3(256), 0, s, V, n, np, out);
    return hipGetLastError();
}

}  // namespace efie
