// efie_mlp.hip -- the near-pair surrogate network and its use in the fill.
//
// Network: 13 -> 64 -> 64 -> 64 -> 16, GELU (exact, erf form) after each
// hidden layer, FP32 weights.  Parameters, one flat float array:
// W0[64][13] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[16][64] b3[16].
//
// Every layer is a GEMM over a block of 64 samples on the matrix cores
// (FP32 MFMA): the forward pass, the backward pass of the mean squared
// error (activations and output gradients staged in LDS, GELU slopes in
// registers), and the weight gradient as dZ^T H over the block's samples.
// Blocks write partial gradients that a second kernel sums in block order,
// so a gradient is bit-reproducible.  Adam updates the parameters.
// Inference: k_sur_records writes the records of 64 near pairs per block,
// the plain product rule plus the network's de-normalised correction.
// Nothing here has a counterpart in the reference.
#include "efie_internal.h"
#include "efie_math.h"

#include <math.h>

namespace efie {

constexpr int kIn = kMlpIn, kH = kMlpHidden, kOut = kMlpOut;
constexpr int oW0 = 0, oB0 = oW0 + kH * kIn, oW1 = oB0 + kH, oB1 = oW1 + kH * kH, oW2 = oB1 + kH,
              oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kOut * kH, kNP = oB3 + kOut;
static_assert(kNP == kMlpParams, "parameter count");
static_assert(kH == 64 && kOut == 16 && kIn <= 16, "tile shapes below assume 13 -> 64 -> 64 -> 64 -> 16");

// A block takes 64 samples; wave w owns samples 16w..16w+15 and runs every
// layer on them as 16x16 MFMA tiles (v_mfma_f32_16x16x4f32: A[i][k] in
// lane i + 16k, B[k][j] in lane j + 16k, lane l element v of the result
// holds D[4(l/16)+v][l%16]).  Weights and activations live in LDS with
// odd row strides so that the operand reads spread over the banks.
constexpr int kTB = 64;
constexpr int PX = 17, PA = 65;  // row strides: inputs (13 + zero pad), 64-wide activations
constexpr int sW0 = 0, sW1 = sW0 + kH * PX, sW2 = sW1 + kH * PA, sW3 = sW2 + kH * PA, sB0 = sW3 + kOut * PA,
              sB1 = sB0 + kH, sB2 = sB1 + kH, sB3 = sB2 + kH, sWEnd = sB3 + kOut;

typedef float f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
// d/dz gelu = Phi(z) + z phi(z)
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Flat parameters -> padded LDS copy (pads zero).
__device__ __forceinline__ void load_weights(const float* __restrict__ p, float* sm) {
    for (int e = threadIdx.x; e < kH * PX; e += blockDim.x) {
        const int r = e / PX, c = e % PX;
        sm[sW0 + e] = c < kIn ? p[oW0 + r * kIn + c] : 0.0f;
    }
    for (int e = threadIdx.x; e < kH * PA; e += blockDim.x) {
        const int r = e / PA, c = e % PA;
        sm[sW1 + e] = c < kH ? p[oW1 + r * kH + c] : 0.0f;
        sm[sW2 + e] = c < kH ? p[oW2 + r * kH + c] : 0.0f;
        if (r < kOut) sm[sW3 + e] = c < kH ? p[oW3 + r * kH + c] : 0.0f;
    }
    for (int e = threadIdx.x; e < 3 * kH + kOut; e += blockDim.x) {
        const int l = e / kH, j = e % kH;  // biases are not contiguous in the flat layout
        sm[sB0 + e] = p[(l == 0 ? oB0 : l == 1 ? oB1 : l == 2 ? oB2 : oB3) + j];
    }
}

// One hidden layer on the wave's rows: out = gelu(in W^T + b) (KP inputs,
// a multiple of 4, row stride PIN); GELU slopes to g when non-null.
template <int KP, int PIN, int PW>
__device__ __forceinline__ void fwd_hidden(const float* W, const float* b, const float* in, float* out, int s0,
                                           f4* g) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int jt = 0; jt < kH / 16; ++jt) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k0 = 0; k0 < KP; k0 += 4)
            acc = mfma4(in[(s0 + li) * PIN + k0 + lg], W[(16 * jt + li) * PW + k0 + lg], acc);
        const float bj = b[16 * jt + li];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float z = acc[v] + bj;
            out[(s0 + 4 * lg + v) * PA + 16 * jt + li] = gelu(z);
            if (g) g[jt][v] = gelu_d(z);
        }
    }
}

// Output layer on the wave's rows: lane element v = y[s0 + 4 lg + v][li].
__device__ __forceinline__ f4 fwd_out(const float* sm, const float* in, int s0) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k0 = 0; k0 < kH; k0 += 4)
        acc = mfma4(in[(s0 + li) * PA + k0 + lg], sm[sW3 + li * PA + k0 + lg], acc);
    const float bj = sm[sB3 + li];
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += bj;
    return acc;
}

// Forward of the wave's 16 samples: x rows in sX (stride PX, columns 13..16
// zero), activations through sA and sB; returns the output tile.
__device__ __forceinline__ f4 mlp_rows(const float* sm, const float* sX, float* sA, float* sB, int s0) {
    fwd_hidden<16, PX, PX>(sm + sW0, sm + sB0, sX, sA, s0, nullptr);
    fwd_hidden<kH, PA, PA>(sm + sW1, sm + sB1, sA, sB, s0, nullptr);
    fwd_hidden<kH, PA, PA>(sm + sW2, sm + sB2, sB, sA, s0, nullptr);
    return fwd_out(sm, sA, s0);
}

// LDS floats of the forward kernels: weights, x, two activation buffers
constexpr int kFwdLds = sWEnd + kTB * PX + 2 * kTB * PA;

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, const float* __restrict__ params,
                                                     float* __restrict__ Y, int N) {
    extern __shared__ float sm[];
    float* sX = sm + sWEnd;
    float* sA = sX + kTB * PX;
    float* sB = sA + kTB * PA;
    load_weights(params, sm);
    const int base = blockIdx.x * kTB;
    for (int e = threadIdx.x; e < kTB * PX; e += blockDim.x) {
        const int r = e / PX, c = e % PX;
        sX[e] = (c < kIn && base + r < N) ? X[(size_t)(base + r) * kIn + c] : 0.0f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4, s0 = 16 * (threadIdx.x >> 6);
    const f4 y = mlp_rows(sm, sX, sA, sB, s0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int s = base + s0 + 4 * lg + v;
        if (s < N) Y[(size_t)s * kOut + li] = y[v];
    }
}

// Backward helpers.  dh = dZ W (dZ rows of the wave, K outputs, W row
// stride PW): lane element v of tile it = dh[s0 + 4 lg + v][16 it + li].
template <int K, int NT, int PW>
__device__ __forceinline__ void bwd_dh(const float* dZ, const float* W, int s0, f4* dh) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int it = 0; it < NT; ++it) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 4)
            acc = mfma4(dZ[(s0 + li) * PA + k0 + lg], W[(k0 + lg) * PW + 16 * it + li], acc);
        dh[it] = acc;
    }
}

// dW[o][i] = sum over the block's 64 samples of dZ[s][o] in[s][i] for the
// output tile ot and input tiles 0..NI-1 (F real inputs, in row stride
// PIN), written to the block's partial gradient.
template <int NI, int F, int PIN>
__device__ __forceinline__ void bwd_dw(const float* dZ, const float* in, int ot, float* __restrict__ part, int wOff) {
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int it = 0; it < NI; ++it) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k0 = 0; k0 < kTB; k0 += 4)
            acc = mfma4(dZ[(k0 + lg) * PA + 16 * ot + li], in[(k0 + lg) * PIN + 16 * it + li], acc);
        const int i = 16 * it + li;
        if (i < F) {
#pragma unroll
            for (int v = 0; v < 4; ++v) part[wOff + (16 * ot + 4 * lg + v) * F + i] = acc[v];
        }
    }
}

// db[o] = sum over the block's samples of dZ[s][o], o < H
__device__ __forceinline__ void bwd_db(const float* dZ, int H, float* __restrict__ part, int bOff) {
    const int t = threadIdx.x;
    if (t < H) {
        float s = 0.0f;
        for (int r = 0; r < kTB; ++r) s += dZ[r * PA + t];
        part[bOff + t] = s;
    }
}

// LDS floats of the training kernel: weights, x, h0, h1, h2, dZ
constexpr int kTrainLds = sWEnd + kTB * PX + 4 * kTB * PA;

// Forward and backward of the loss sum((y - t)^2) / (16 N) for 64 samples
// per block: the block's gradient into part[block][kNP], its loss sum into
// lossPart[block].  Samples past N are zero inputs with zero output
// gradient.
__global__ __launch_bounds__(256) void k_mlp_train(const float* __restrict__ X, const float* __restrict__ T,
                                                   const float* __restrict__ params, int N, float* __restrict__ part,
                                                   float* __restrict__ lossPart) {
    extern __shared__ float sm[];
    __shared__ float red[4];
    float* sX = sm + sWEnd;
    float* sH0 = sX + kTB * PX;
    float* sH1 = sH0 + kTB * PA;
    float* sH2 = sH1 + kTB * PA;
    float* sDZ = sH2 + kTB * PA;
    load_weights(params, sm);
    const int base = blockIdx.x * kTB;
    for (int e = threadIdx.x; e < kTB * PX; e += blockDim.x) {
        const int r = e / PX, c = e % PX;
        sX[e] = (c < kIn && base + r < N) ? X[(size_t)(base + r) * kIn + c] : 0.0f;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4, s0 = 16 * wave;
    float* pb = part + (size_t)blockIdx.x * kNP;

    f4 g0[4], g1[4], g2[4], dh[4];
    fwd_hidden<16, PX, PX>(sm + sW0, sm + sB0, sX, sH0, s0, g0);
    fwd_hidden<kH, PA, PA>(sm + sW1, sm + sB1, sH0, sH1, s0, g1);
    fwd_hidden<kH, PA, PA>(sm + sW2, sm + sB2, sH1, sH2, s0, g2);
    const f4 y = fwd_out(sm, sH2, s0);
    const float sc = 2.0f / ((float)N * (float)kOut);
    float loss = 0.0f;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int s = base + s0 + 4 * lg + v;
        const float e = s < N ? y[v] - T[(size_t)s * kOut + li] : 0.0f;
        loss = fmaf(e, e, loss);
        sDZ[(s0 + 4 * lg + v) * PA + li] = sc * e;
    }
    __syncthreads();
    // layer 3: dW3 (one output tile, input tile = wave), db3, dh2
    bwd_dw<1, kH, PA>(sDZ, sH2 + 16 * wave, 0, pb, oW3 + 16 * wave);
    bwd_db(sDZ, kOut, pb, oB3);
    bwd_dh<kOut, 4, PA>(sDZ, sm + sW3, s0, dh);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int v = 0; v < 4; ++v) sDZ[(s0 + 4 * lg + v) * PA + 16 * it + li] = dh[it][v] * g2[it][v];
    __syncthreads();
    // layer 2
    bwd_dw<4, kH, PA>(sDZ, sH1, wave, pb, oW2);
    bwd_db(sDZ, kH, pb, oB2);
    bwd_dh<kH, 4, PA>(sDZ, sm + sW2, s0, dh);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int v = 0; v < 4; ++v) sDZ[(s0 + 4 * lg + v) * PA + 16 * it + li] = dh[it][v] * g1[it][v];
    __syncthreads();
    // layer 1
    bwd_dw<4, kH, PA>(sDZ, sH0, wave, pb, oW1);
    bwd_db(sDZ, kH, pb, oB1);
    bwd_dh<kH, 4, PA>(sDZ, sm + sW1, s0, dh);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int v = 0; v < 4; ++v) sDZ[(s0 + 4 * lg + v) * PA + 16 * it + li] = dh[it][v] * g0[it][v];
    __syncthreads();
    // layer 0
    bwd_dw<1, kIn, PX>(sDZ, sX, wave, pb, oW0);
    bwd_db(sDZ, kH, pb, oB0);
    // loss
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (lane == 0) red[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) lossPart[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// grad[i] = sum over blocks of part[b][i] in block order; the thread past
// the parameters sums the loss partials.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int nBlocks,
                                                    const float* __restrict__ lossPart, float scale,
                                                    float* __restrict__ grad, float* __restrict__ lossOut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kNP) {
        float s = 0.0f;
        for (int c = 0; c < nBlocks; ++c) s += part[(size_t)c * kNP + i];
        grad[i] = s;
    } else if (i == kNP) {
        double s = 0.0;
        for (int c = 0; c < nBlocks; ++c) s += lossPart[c];
        lossOut[0] = (float)(s * scale);
    }
}

// Adam (Kingma and Ba, 2015) with bias correction
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------------ inference --
// Records of the near pairs inside the network's envelope, 64 pairs per
// block, each written once: the plain product rule of the regular pairs
// (sur_base) plus the network's correction of it.  The network (MFMA
// forward as above, inputs standardised from the pair's features and k d)
// or a label buffer gives T (re, im of S0, S1, S1', S2, normalised, times
// exp(+j k d) 4 pi); with B the product rule in the same frame and
// C = T exp(-j k d) / (4 pi) + B, the record is
//   S0 = L^3 C0,  S1 = L^4 Rm^T C1,  S1' = L^4 Rm^T C1',  S2 = L^5 C2.

// The product rule on one pair in the surrogate's frame: v holds t1's then
// t2's vertices (lengths in units of L), reg the rule's points as
// barycentric weights and weight fractions of twice the area ([Nq][4]), kt
// the wavenumber times L.  Moments as a record's (S1 about t1's centroid,
// S1' about t2's) with G = exp(-j kt R) / (4 pi R).  Thread u of a group of
// four consecutive lanes takes the test points u, u + 4, ...; the group's
// sums meet by shuffles, every thread of the group ends with all 16.
__device__ void sur_base(const double* __restrict__ v, const double* __restrict__ reg, int Nq, double kt, int u,
                         double* res) {
    double c[6], a2[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const double* w = v + 9 * t;
#pragma unroll
        for (int d = 0; d < 3; ++d) c[3 * t + d] = (w[d] + w[3 + d] + w[6 + d]) / 3.0;
        const double ex = w[3] - w[0], ey = w[4] - w[1], ez = w[5] - w[2];
        const double fx = w[6] - w[0], fy = w[7] - w[1], fz = w[8] - w[2];
        const double nx = ey * fz - ez * fy, ny = ez * fx - ex * fz, nz = ex * fy - ey * fx;
        a2[t] = sqrt(nx * nx + ny * ny + nz * nz);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) res[k] = 0.0;
    const double inv4pi = 1.0 / (4.0 * 3.141592653589793238462643383279502884);
    for (int i = u; i < Nq; i += 4) {
        const double* b = reg + 4 * i;
        const double px = b[0] * v[0] + b[1] * v[3] + b[2] * v[6], py = b[0] * v[1] + b[1] * v[4] + b[2] * v[7],
                     pz = b[0] * v[2] + b[1] * v[5] + b[2] * v[8];
        const double wi = b[3] * a2[0];
        double gr = 0, gi = 0, hxr = 0, hxi = 0, hyr = 0, hyi = 0, hzr = 0, hzi = 0;
        for (int j = 0; j < Nq; ++j) {
            const double* q = reg + 4 * j;
            const double qx = q[0] * v[9] + q[1] * v[12] + q[2] * v[15], qy = q[0] * v[10] + q[1] * v[13] + q[2] * v[16],
                         qz = q[0] * v[11] + q[1] * v[14] + q[2] * v[17];
            const double dx = px - qx, dy = py - qy, dz = pz - qz;
            const double R = sqrt(dx * dx + dy * dy + dz * dz);
            double sn, cs;
            sincos_cw(kt * R, &sn, &cs);
            const double f = q[3] * a2[1] * inv4pi / R;
            const double wr = cs * f, wim = -sn * f;
            const double rx = qx - c[3], ry = qy - c[4], rz = qz - c[5];
            gr += wr;
            gi += wim;
            hxr += wr * rx;
            hxi += wim * rx;
            hyr += wr * ry;
            hyi += wim * ry;
            hzr += wr * rz;
            hzi += wim * rz;
        }
        const double rx = px - c[0], ry = py - c[1], rz = pz - c[2];
        res[0] += wi * gr;
        res[1] += wi * gi;
        res[2] += wi * rx * gr;
        res[3] += wi * rx * gi;
        res[4] += wi * ry * gr;
        res[5] += wi * ry * gi;
        res[6] += wi * rz * gr;
        res[7] += wi * rz * gi;
        res[8] += wi * hxr;
        res[9] += wi * hxi;
        res[10] += wi * hyr;
        res[11] += wi * hyi;
        res[12] += wi * hzr;
        res[13] += wi * hzi;
        res[14] += wi * (rx * hxr + ry * hyr + rz * hzr);
        res[15] += wi * (rx * hxi + ry * hyi + rz * hzi);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        res[k] += __shfl_xor(res[k], 1);
        res[k] += __shfl_xor(res[k], 2);
    }
}

template <bool FROM_BUF>
__global__ __launch_bounds__(256) void k_sur_records(const SurPair* __restrict__ pairs, int n,
                                                     const float* __restrict__ model, const double* __restrict__ Tbuf,
                                                     const double* __restrict__ tri18, const double* __restrict__ reg,
                                                     int Nq, const Freq* __restrict__ fr, Rec* __restrict__ recs) {
    extern __shared__ float sm[];
    __shared__ double sBase[kTB][17];
    const int base = blockIdx.x * kTB;
    const double kr = fr->kr;
    // the product rule: four threads per pair
    {
        const int pl = threadIdx.x >> 2, u = threadIdx.x & 3, p = base + pl;
        double v[18];
        const int pc = p < n ? p : n - 1;
#pragma unroll
        for (int i = 0; i < 18; ++i) v[i] = tri18[(size_t)pc * 18 + i];
        double res[16];
        sur_base(v, reg, Nq, kr * pairs[pc].L, u, res);
        if (u == 0) {
#pragma unroll
            for (int k = 0; k < 16; ++k) sBase[pl][k] = res[k];
        }
    }
    const float* xm = model + kNP;   // standardisation after the parameters
    const float* xi = xm + kIn;
    const float* ym = xi + kIn;
    const float* ys = ym + kOut;
    float* sX = sm + sWEnd;
    if (!FROM_BUF) {
        float* sA = sX + kTB * PX;
        float* sB = sA + kTB * PA;
        load_weights(model, sm);
        for (int e = threadIdx.x; e < kTB * PX; e += blockDim.x) {
            const int r = e / PX, c = e % PX, p = base + r;
            float v = 0.0f;
            if (p < n && c < kIn) {
                const float raw = c < 12 ? pairs[p].feat[c] : (float)(kr * pairs[p].d);
                v = (raw - xm[c]) * xi[c];
            }
            sX[e] = v;
        }
        __syncthreads();
        const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4, s0 = 16 * (threadIdx.x >> 6);
        const f4 y = mlp_rows(sm, sX, sA, sB, s0);
        // outputs over the wave's own x rows
#pragma unroll
        for (int v = 0; v < 4; ++v) sX[(s0 + 4 * lg + v) * PX + li] = y[v];
    }
    __syncthreads();
    const int t = threadIdx.x, p = base + t;
    if (t >= kTB || p >= n) return;
    const SurPair& P = pairs[p];
    const double d = P.d, L = P.L;
    double T[kOut];
#pragma unroll
    for (int o = 0; o < kOut; ++o)
        T[o] = FROM_BUF ? Tbuf[(size_t)p * kOut + o] : (double)fmaf(sX[t * PX + o], ys[o], ym[o]);
    double sn, cs;
    sincos_cw(kr * d, &sn, &cs);
    const double er = cs * (1.0 / (4.0 * 3.141592653589793238462643383279502884));
    const double ei = -sn * (1.0 / (4.0 * 3.141592653589793238462643383279502884));
    double C[kOut];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        C[2 * q] = T[2 * q] * er - T[2 * q + 1] * ei + sBase[t][2 * q];
        C[2 * q + 1] = T[2 * q] * ei + T[2 * q + 1] * er + sBase[t][2 * q + 1];
    }
    const double L3 = L * L * L, L4 = L3 * L, L5 = L4 * L;
    double* r = recs[rec_index(P.t1, P.t2)].v;
    r[0] = L3 * C[0];
    r[1] = L3 * C[1];
#pragma unroll
    for (int v = 0; v < 2; ++v) {        // S1 (slots 1-3), S1' (slots 4-6)
#pragma unroll
        for (int a = 0; a < 3; ++a) {    // global component a = sum_b Rm[b][a] local_b
            double gr = 0.0, gi = 0.0;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                gr += P.Rm[3 * b + a] * C[2 * (1 + 3 * v + b)];
                gi += P.Rm[3 * b + a] * C[2 * (1 + 3 * v + b) + 1];
            }
            r[2 * (1 + 3 * v + a)] = L4 * gr;
            r[2 * (1 + 3 * v + a) + 1] = L4 * gi;
        }
    }
    r[14] = L5 * C[14];
    r[15] = L5 * C[15];
}

// ------------------------------------------------------------ launchers --

static int blocks_of(int N) { return (N + kTB - 1) / kTB; }

// Kernels with more than 64 KiB of dynamic LDS declare it once per process.
static hipError_t allow_lds() {
    static hipError_t e = [] {
        hipError_t r = hipFuncSetAttribute((const void*)k_mlp_train, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(sizeof(float) * kTrainLds));
        if (r == hipSuccess)
            r = hipFuncSetAttribute((const void*)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(float) * kFwdLds));
        if (r == hipSuccess)
            r = hipFuncSetAttribute((const void*)k_sur_records<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(float) * kFwdLds));
        return r;
    }();
    return e;
}

hipError_t launch_mlp_forward(const float* X, const float* params, float* Y, int N, hipStream_t s) {
    if (N <= 0) return hipSuccess;
    hipError_t e = allow_lds();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_of(N)), dim3(256), sizeof(float) * kFwdLds, s, X, params, Y, N);
    return hipGetLastError();
}

size_t mlp_grad_work_floats(int N) { return (size_t)blocks_of(N) * (kNP + 1); }

hipError_t launch_mlp_loss_grad(const float* X, const float* T, const float* params, int N, float* work,
                                float* grad, float* lossOut, hipStream_t s) {
    if (N <= 0) return hipErrorInvalidValue;
    hipError_t e = allow_lds();
    if (e != hipSuccess) return e;
    const int nb = blocks_of(N);
    float* part = work;
    float* lossPart = part + (size_t)nb * kNP;
    hipLaunchKernelGGL(k_mlp_train, dim3(nb), dim3(256), sizeof(float) * kTrainLds, s, X, T, params, N, part,
                       lossPart);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kNP + 1 + 255) / 256), dim3(256), 0, s, part, nb, lossPart,
                       1.0f / ((float)N * (float)kOut), grad, lossOut);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, int step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t launch_sur_records(const SurPair* pairs, int n, const float* model, const double* Tbuf,
                              const double* tri18, const double* regBary, int Nq, const Freq* fr, Rec* recs,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (Nq < 1 || Nq > 16) return hipErrorInvalidValue;
    if (Tbuf) {
        hipLaunchKernelGGL(k_sur_records<true>, dim3(blocks_of(n)), dim3(256), 0, s, pairs, n, model, Tbuf, tri18,
                           regBary, Nq, fr, recs);
    } else {
        hipError_t e = allow_lds();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_sur_records<false>, dim3(blocks_of(n)), dim3(256), sizeof(float) * kFwdLds, s, pairs, n,
                           model, Tbuf, tri18, regBary, Nq, fr, recs);
    }
    return hipGetLastError();
}

}  // namespace efie
