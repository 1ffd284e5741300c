// 3-D convolution kernels of the learned preconditioner (gfx950).
//
// Periodic 3x3x3 convolutions on grids (B, C, D, H, W), C-contiguous, the
// layout of a batch of band grids with a channel axis.  No counterpart in
// the reference: these serve the learned correction that stands beside the
// multigrid preconditioner.
//
//   forward:        y[b,o,p] = bias[o] + sum_{i,t} W[o,i,t] x[b,i,p+d_t]
//   backward data:  gx[b,i,p] = sum_{o,t} W[o,i,t] gy[b,o,p-d_t]
//   backward weight gW[o,i,t] = sum_{b,p} gy[b,o,p] x[b,i,p+d_t]
//   backward bias   gb[o]     = sum_{b,p} gy[b,o,p]
// with d_t = (t/9 - 1, (t/3)%3 - 1, t%3 - 1) and periodic wrap.  Weights are
// FP32 (o, i, 27).  Each thread of the data kernels computes one point for
// 16 output channels.  The weight kernel splits the B*P points into chunks:
// one workgroup per (o, i, chunk) reduces its 27 sums (and the bias sum) in
// LDS, then the chunks are summed in order (fixed order, no atomics).  FP32
// with FMA.
#include <algorithm>

#include "common.hpp"

namespace gpd {

constexpr int kCOB = 16;   // output channels per thread

struct ConvDims {
    int B, Cin, Cout, D, H, W;
};

__device__ __forceinline__ int wrapp(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// mode 0: forward (weight index o*Cin*27 + i*27 + t); mode 1: backward data
// (roles of o and i swapped, taps mirrored).
template <int MODE>
__global__ __launch_bounds__(256) void conv3d_kernel(const float* __restrict__ x,
                                                    const float* __restrict__ Wt,
                                                    const float* __restrict__ bias,
                                                    float* __restrict__ y, ConvDims d,
                                                    int accumulate)
{
    const int P = d.D * d.H * d.W;
    // channels seen by this kernel: nin inputs, nout outputs
    const int nin = MODE == 0 ? d.Cin : d.Cout;
    const int nout = MODE == 0 ? d.Cout : d.Cin;
    const int ocb = (nout + kCOB - 1) / kCOB;
    const long w = blockIdx.y;                  // (b, output-channel block)
    const int b = (int)(w / ocb), o0 = (int)(w % ocb) * kCOB;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int px = p % d.W, py = (p / d.W) % d.H, pz = p / (d.W * d.H);
    int nb[27];
#pragma unroll
    for (int t = 0; t < 27; t++) {
        const int s = MODE == 0 ? 1 : -1;
        const int z = wrapp(pz + s * (t / 9 - 1), d.D);
        const int yy = wrapp(py + s * ((t / 3) % 3 - 1), d.H);
        const int xx = wrapp(px + s * (t % 3 - 1), d.W);
        nb[t] = (z * d.H + yy) * d.W + xx;
    }
    float acc[kCOB];
#pragma unroll
    for (int c = 0; c < kCOB; c++) acc[c] = 0.f;
    const float* xb = x + (long)b * nin * P;
    for (int i = 0; i < nin; i++) {
        const float* xi = xb + (long)i * P;
        float v[27];
#pragma unroll
        for (int t = 0; t < 27; t++) v[t] = xi[nb[t]];
#pragma unroll
        for (int c = 0; c < kCOB; c++) {
            const int o = o0 + c;
            if (o >= nout) break;
            const float* wr = MODE == 0 ? Wt + ((long)o * d.Cin + i) * 27
                                        : Wt + ((long)i * d.Cin + o) * 27;
            float s = acc[c];
#pragma unroll
            for (int t = 0; t < 27; t++) s = __builtin_fmaf(wr[t], v[t], s);
            acc[c] = s;
        }
    }
    float* yb = y + (long)b * nout * P;
#pragma unroll
    for (int c = 0; c < kCOB; c++) {
        const int o = o0 + c;
        if (o >= nout) break;
        float r = acc[c] + ((MODE == 0 && bias) ? bias[o] : 0.f);
        float* dst = yb + (long)o * P + p;
        *dst = accumulate ? *dst + r : r;
    }
}

// Partial gW[o,i,t] = sum_{b,p in chunk} gy[b,o,p] x[b,i,p+d_t]: one
// workgroup per (o, i, chunk of the B*P points), 27 sums reduced in LDS in
// a fixed order; part[(chunk * Cout + o) * Cin + i][27].
__global__ __launch_bounds__(256) void conv3d_wgrad_kernel(const float* __restrict__ x,
                                                          const float* __restrict__ gy,
                                                          float* __restrict__ part,
                                                          float* __restrict__ partb, ConvDims d,
                                                          long chunk)
{
    __shared__ float red[256][28];
    const long P = (long)d.D * d.H * d.W;
    const int oi = blockIdx.x;
    const int o = oi / d.Cin, i = oi % d.Cin;
    const long q0 = (long)blockIdx.y * chunk;
    const long q1 = min((long)d.B * P, q0 + chunk);
    float acc[27];
#pragma unroll
    for (int t = 0; t < 27; t++) acc[t] = 0.f;
    float accb = 0.f;   // bias gradient, kept by the i == 0 workgroups
    for (long q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const int b = (int)(q / P);
        const int p = (int)(q % P);
        const float* xi = x + ((long)b * d.Cin + i) * P;
        const float g = gy[((long)b * d.Cout + o) * P + p];
        accb += g;
        const int px = p % d.W, py = (p / d.W) % d.H, pz = p / (d.W * d.H);
#pragma unroll
        for (int t = 0; t < 27; t++) {
            const int z = wrapp(pz + t / 9 - 1, d.D);
            const int yy = wrapp(py + (t / 3) % 3 - 1, d.H);
            const int xx = wrapp(px + t % 3 - 1, d.W);
            acc[t] = __builtin_fmaf(g, xi[(z * d.H + yy) * d.W + xx], acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < 27; t++) red[threadIdx.x][t] = acc[t];
    red[threadIdx.x][27] = accb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int t = 0; t < 28; t++) red[threadIdx.x][t] += red[threadIdx.x + s][t];
        __syncthreads();
    }
    if (threadIdx.x < 27)
        part[((long)blockIdx.y * d.Cout * d.Cin + oi) * 27 + threadIdx.x] = red[0][threadIdx.x];
    if (threadIdx.x == 27 && i == 0 && partb)
        partb[(long)blockIdx.y * d.Cout + o] = red[0][27];
}

// gW = sum over chunks of the partials, in chunk order
__global__ void conv3d_wgrad_reduce_kernel(const float* __restrict__ part, float* __restrict__ gW,
                                           long n, int nchunk)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += part[(long)c * n + k];
    gW[k] = s;
}

// leaky ReLU and its gradient (slope a for x < 0), elementwise
__global__ void leaky_relu_kernel(long n, float a, const float* __restrict__ x,
                                  float* __restrict__ y)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float v = x[i];
        y[i] = v > 0.f ? v : a * v;
    }
}

__global__ void leaky_relu_grad_kernel(long n, float a, const float* __restrict__ x,
                                       const float* __restrict__ gy, float* __restrict__ gx)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gx[i] = x[i] > 0.f ? gy[i] : a * gy[i];
}

// y = alpha x + beta y, elementwise
__global__ void scale_add_kernel(long n, float alpha, const float* __restrict__ x, float beta,
                                 float* __restrict__ y)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = alpha * x[i] + beta * y[i];
}

// Adam step on n parameters (bias-corrected learning rate lr_t)
__global__ void adam_kernel(long n, float lr_t, float b1, float b2, float eps,
                            float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

}  // namespace gpd

using namespace gpd;

static bool conv_dims_ok(int B, int Cin, int Cout, int D, int H, int W)
{
    return B > 0 && Cin > 0 && Cout > 0 && D > 0 && H > 0 && W > 0 && B <= 65535 &&
           (long)D * H * W < (1L << 31);
}

extern "C" int gpd_conv3d_forward(int B, int Cin, int Cout, int D, int H, int W,
                                  const float* d_x, const float* d_w, const float* d_bias,
                                  float* d_y, int accumulate, void* stream)
{
    GPD_REQUIRE(conv_dims_ok(B, Cin, Cout, D, H, W), GPD_ERR_ARG, "bad conv sizes");
    GPD_REQUIRE(d_x && d_w && d_y && d_x != d_y, GPD_ERR_ARG, "bad arrays");
    ConvDims d = {B, Cin, Cout, D, H, W};
    const long P = (long)D * H * W;
    const int ocb = (Cout + kCOB - 1) / kCOB;
    GPD_REQUIRE((long)B * ocb <= 65535, GPD_ERR_ARG, "batch too large");
    hipLaunchKernelGGL(conv3d_kernel<0>, dim3((P + 255) / 256, B * ocb), dim3(256), 0,
                       (hipStream_t)stream, d_x, d_w, d_bias, d_y, d, accumulate);
    return last_error_to_status(hipGetLastError());
}

extern "C" int gpd_conv3d_backward_data(int B, int Cin, int Cout, int D, int H, int W,
                                        const float* d_gy, const float* d_w, float* d_gx,
                                        int accumulate, void* stream)
{
    GPD_REQUIRE(conv_dims_ok(B, Cin, Cout, D, H, W), GPD_ERR_ARG, "bad conv sizes");
    GPD_REQUIRE(d_gy && d_w && d_gx && d_gy != d_gx, GPD_ERR_ARG, "bad arrays");
    ConvDims d = {B, Cin, Cout, D, H, W};
    const long P = (long)D * H * W;
    const int icb = (Cin + kCOB - 1) / kCOB;
    GPD_REQUIRE((long)B * icb <= 65535, GPD_ERR_ARG, "batch too large");
    hipLaunchKernelGGL(conv3d_kernel<1>, dim3((P + 255) / 256, B * icb), dim3(256), 0,
                       (hipStream_t)stream, d_gy, d_w, nullptr, d_gx, d, accumulate);
    return last_error_to_status(hipGetLastError());
}

extern "C" int gpd_conv3d_backward_weight(int B, int Cin, int Cout, int D, int H, int W,
                                          const float* d_x, const float* d_gy, float* d_gw,
                                          float* d_gb, void* stream)
{
    GPD_REQUIRE(conv_dims_ok(B, Cin, Cout, D, H, W), GPD_ERR_ARG, "bad conv sizes");
    GPD_REQUIRE(d_x && d_gy && d_gw, GPD_ERR_ARG, "bad arrays");
    GPD_REQUIRE((long)Cout * Cin <= (1L << 30), GPD_ERR_ARG, "too many channel pairs");
    ConvDims d = {B, Cin, Cout, D, H, W};
    hipStream_t st = (hipStream_t)stream;
    // split the B*P points so that about 2048 workgroups run, chunks of at
    // least 4096 points
    const long npts = (long)B * D * H * W;
    const long pairs = (long)Cout * Cin;
    long nchunk = std::max(1L, std::min((2048 + pairs - 1) / pairs, npts / 4096));
    nchunk = std::min(nchunk, 65535L);
    const long chunk = (npts + nchunk - 1) / nchunk;
    nchunk = (npts + chunk - 1) / chunk;
    const long n = pairs * 27;
    float* part = nullptr;
    hipError_t e = hipMallocAsync((void**)&part, sizeof(float) * (n + Cout) * nchunk, st);
    if (e != hipSuccess) {
        set_last_error(hipGetErrorString(e));
        return GPD_ERR_HIP;
    }
    float* partb = part + n * nchunk;
    hipLaunchKernelGGL(conv3d_wgrad_kernel, dim3((unsigned)pairs, (unsigned)nchunk), dim3(256), 0,
                       st, d_x, d_gy, part, partb, d, chunk);
    int rc = last_error_to_status(hipGetLastError());
    if (!rc) {
        hipLaunchKernelGGL(conv3d_wgrad_reduce_kernel, dim3((unsigned)((n + 255) / 256)),
                           dim3(256), 0, st, part, d_gw, n, (int)nchunk);
        rc = last_error_to_status(hipGetLastError());
    }
    if (!rc && d_gb) {
        hipLaunchKernelGGL(conv3d_wgrad_reduce_kernel, dim3((unsigned)((Cout + 255) / 256)),
                           dim3(256), 0, st, partb, d_gb, (long)Cout, (int)nchunk);
        rc = last_error_to_status(hipGetLastError());
    }
    (void)hipFreeAsync(part, st);
    return rc;
}

extern "C" int gpd_leaky_relu(long n, float slope, const float* d_x, float* d_y, void* stream)
{
    GPD_REQUIRE(n >= 0 && (n == 0 || (d_x && d_y)), GPD_ERR_ARG, "bad arrays");
    if (n == 0) return GPD_OK;
    hipLaunchKernelGGL(leaky_relu_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, n, slope, d_x, d_y);
    return last_error_to_status(hipGetLastError());
}

extern "C" int gpd_leaky_relu_backward(long n, float slope, const float* d_x, const float* d_gy,
                                       float* d_gx, void* stream)
{
    GPD_REQUIRE(n >= 0 && (n == 0 || (d_x && d_gy && d_gx)), GPD_ERR_ARG, "bad arrays");
    if (n == 0) return GPD_OK;
    hipLaunchKernelGGL(leaky_relu_grad_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, n, slope, d_x, d_gy, d_gx);
    return last_error_to_status(hipGetLastError());
}

extern "C" int gpd_adam_step(long n, float lr_t, float beta1, float beta2, float eps, float* d_w,
                             const float* d_g, float* d_m, float* d_v, void* stream)
{
    GPD_REQUIRE(n >= 0 && (n == 0 || (d_w && d_g && d_m && d_v)), GPD_ERR_ARG, "bad arrays");
    if (n == 0) return GPD_OK;
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                       lr_t, beta1, beta2, eps, d_w, d_g, d_m, d_v);
    return last_error_to_status(hipGetLastError());
}

extern "C" int gpd_scale_add(long n, float alpha, const float* d_x, float beta, float* d_y,
                             void* stream)
{
    GPD_REQUIRE(n >= 0 && (n == 0 || (d_x && d_y)), GPD_ERR_ARG, "bad arrays");
    if (n == 0) return GPD_OK;
    hipLaunchKernelGGL(scale_add_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       n, alpha, d_x, beta, d_y);
    return last_error_to_status(hipGetLastError());
}

// ------------------------------------------------------------------ //
// Forward convolution on MFMA with bfloat16 operands (inference).
//
// Implicit GEMM per workgroup tile of 256 output points (one z-plane, TY
// rows, TX columns; TX = 32 when 32 divides W, else 16): D[o][p] =
// sum_{t, c} W[o, c, t] x[c, p + d_t] on v_mfma_f32_32x32x8_bf16 with
// A = weights (rows o), B = gathered inputs (columns p), k = 8 channels at
// one tap.  Input channels are staged 8 at a time: the tile's input points
// with their 1-point halo (3 planes) and the weights of those channels go
// to LDS as bfloat16 (round to nearest even); the FP32 accumulators hold
// all 27 taps and all channels.  Each wave owns 64 points (two 32-point
// column blocks) for NOB blocks of 32 output channels.  Requires Cin % 8 ==
// 0, Cout <= 32 NOB, W % 16 == 0 and H divisible by TY.
namespace gpd {

constexpr int kMCh = 8;   // channels per LDS stage (one MFMA k-step)

template <int TX, int NOB>
__global__ __launch_bounds__(256) void conv3d_fwd_mfma_kernel(const float* __restrict__ x,
                                                             const float* __restrict__ Wt,
                                                             const float* __restrict__ bias,
                                                             float* __restrict__ y, ConvDims d)
{
    constexpr int TY = 256 / TX;
    constexpr int HX = TX + 2, HY = TY + 2;
    constexpr int NPTS = 3 * HY * HX;
    constexpr int NO = 32 * NOB;
    __shared__ __align__(16) short xs[NPTS * kMCh];
    __shared__ __align__(16) short ws[27 * NO * kMCh];

    const int P = d.D * d.H * d.W;
    const int ntx = d.W / TX, nty = d.H / TY;
    const long nwork = (long)d.B * d.D * nty * ntx;
    const long w = xcd_work_index(blockIdx.x, nwork);
    if (w < 0) return;
    const int tx_i = (int)(w % ntx);
    long r = w / ntx;
    const int ty_i = (int)(r % nty);
    r /= nty;
    const int z = (int)(r % d.D);
    const int b = (int)(r / d.D);
    const int x0 = tx_i * TX, y0 = ty_i * TY;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;

    f32x16 acc[2][NOB];
#pragma unroll
    for (int pb = 0; pb < 2; pb++)
#pragma unroll
        for (int ob = 0; ob < NOB; ob++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[pb][ob][v] = 0.f;

    // this lane's output points (column of the B operand) in each block
    int qy[2], qx[2];
#pragma unroll
    for (int pb = 0; pb < 2; pb++) {
        const int q = wave * 64 + pb * 32 + lane % 32;
        qy[pb] = q / TX;
        qx[pb] = q % TX;
    }
    const int cq = 4 * (lane / 32);   // this lane's 4 channels of the k-step

    for (int c0 = 0; c0 < d.Cin; c0 += kMCh) {
        __syncthreads();
        for (int s = threadIdx.x; s < NPTS * kMCh; s += 256) {
            const int c = s / NPTS, pt = s % NPTS;
            const int hz = pt / (HY * HX), rem = pt % (HY * HX);
            const int hy = rem / HX, hx = rem % HX;
            const int gz = wrapp(z - 1 + hz, d.D);
            const int gy = wrapp(y0 - 1 + hy, d.H);
            const int gx = wrapp(x0 - 1 + hx, d.W);
            const float v = x[((long)b * d.Cin + c0 + c) * P + ((long)gz * d.H + gy) * d.W + gx];
            xs[pt * kMCh + c] = (short)bf16_bits(v);
        }
        for (int s = threadIdx.x; s < 27 * NO * kMCh; s += 256) {
            const int c = s % kMCh, rem = s / kMCh;
            const int o = rem % NO, t = rem / NO;
            const float v = o < d.Cout ? Wt[((long)o * d.Cin + c0 + c) * 27 + t] : 0.f;
            ws[(t * NO + o) * kMCh + c] = (short)bf16_bits(v);
        }
        __syncthreads();
#pragma unroll 3
        for (int t = 0; t < 27; t++) {
            const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;
            bf16x4 bop[2];
#pragma unroll
            for (int pb = 0; pb < 2; pb++) {
                const int pt = (dz * HY + qy[pb] + dy) * HX + qx[pb] + dx;
                bop[pb] = *reinterpret_cast<const bf16x4*>(&xs[pt * kMCh + cq]);
            }
#pragma unroll
            for (int ob = 0; ob < NOB; ob++) {
                const int o = ob * 32 + lane % 32;
                const bf16x4 aop = *reinterpret_cast<const bf16x4*>(&ws[(t * NO + o) * kMCh + cq]);
#pragma unroll
                for (int pb = 0; pb < 2; pb++)
                    acc[pb][ob] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(aop, bop[pb],
                                                                          acc[pb][ob], 0, 0, 0);
            }
        }
    }
    // D: column = point lane % 32, row = output channel 8 (v / 4) + 4 (lane / 32) + v % 4
#pragma unroll
    for (int pb = 0; pb < 2; pb++) {
        const long p = ((long)z * d.H + y0 + qy[pb]) * d.W + x0 + qx[pb];
#pragma unroll
        for (int ob = 0; ob < NOB; ob++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int o = ob * 32 + 8 * (v / 4) + 4 * (lane / 32) + v % 4;
                if (o < d.Cout)
                    y[((long)b * d.Cout + o) * P + p] = acc[pb][ob][v] + (bias ? bias[o] : 0.f);
            }
    }
}

}  // namespace gpd

extern "C" int gpd_conv3d_forward_bf16(int B, int Cin, int Cout, int D, int H, int W,
                                       const float* d_x, const float* d_w, const float* d_bias,
                                       float* d_y, void* stream)
{
    GPD_REQUIRE(conv_dims_ok(B, Cin, Cout, D, H, W), GPD_ERR_ARG, "bad conv sizes");
    GPD_REQUIRE(d_x && d_w && d_y && d_x != d_y, GPD_ERR_ARG, "bad arrays");
    const int TX = W % 32 == 0 ? 32 : 16;
    const int TY = 256 / TX;
    if (Cin % kMCh != 0 || Cout > 64 || W % 16 != 0 || H % TY != 0) {
        set_last_error("shape not supported by the bfloat16 MFMA convolution");
        return GPD_ERR_UNSUPPORTED;
    }
    ConvDims d = {B, Cin, Cout, D, H, W};
    const long nwork = (long)B * D * (H / TY) * (W / TX);
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid(xcd_grid_size(nwork)), block(256);
    if (TX == 32) {
        if (Cout <= 32)
            hipLaunchKernelGGL((conv3d_fwd_mfma_kernel<32, 1>), grid, block, 0, st, d_x, d_w,
                               d_bias, d_y, d);
        else
            hipLaunchKernelGGL((conv3d_fwd_mfma_kernel<32, 2>), grid, block, 0, st, d_x, d_w,
                               d_bias, d_y, d);
    } else {
        if (Cout <= 32)
            hipLaunchKernelGGL((conv3d_fwd_mfma_kernel<16, 1>), grid, block, 0, st, d_x, d_w,
                               d_bias, d_y, d);
        else
            hipLaunchKernelGGL((conv3d_fwd_mfma_kernel<16, 2>), grid, block, 0, st, d_x, d_w,
                               d_bias, d_y, d);
    }
    return last_error_to_status(hipGetLastError());
}
