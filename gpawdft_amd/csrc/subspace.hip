// Subspace matrices and rotations on MFMA (gfx950).
//
// Replaces the BLAS-3 calls behind MatrixOperator (gpaw/hs_operators.py:
// 251-313, 417-466 via gpaw/grid_descriptor.py:263-323 and c/blas.c:
// dsyrk/dsyr2k/dgemm):
//  * inner products T = X Y^T over the grid (S: X = Y = psi; H: X = psi,
//    Y = H psi, symmetrised).  FP32 inputs: each value is split into three
//    bfloat16 parts x = hi + mid + lo and the six products whose size is
//    above 2^-24 |x y| run on v_mfma_f32_32x32x8_bf16; the FP32 accumulator is
//    flushed into FP64 registers every 64 values of k, and the K-chunks'
//    partial matrices are summed in FP64 by a second kernel (fixed order).
//    FP64 inputs: v_mfma_f64_16x16x4 (its result row is 4 v + lane / 16,
//    column lane % 16, for accumulator register v).
//  * rotations psi' = C psi (C: FP64 N x N): FP32 psi on
//    v_mfma_f32_32x32x2_f32 with the accumulator flushed into FP64
//    registers every 32 terms; FP64 psi on v_mfma_f64_16x16x4.
//  * the PAW term  A += P2 dA P1^T  (small, FP64).
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace gpd {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// x = hi + mid + lo with bfloat16 parts (the remainder after three
// parts is below 2^-24 |x|; all subtractions are exact in FP32)
__device__ __forceinline__ void split3(float x, short& hi, short& mid, short& lo)
{
    const unsigned short h = bf16_bits(x);
    const float r1 = x - bf16_value(h);
    const unsigned short m = bf16_bits(r1);
    const float r2 = r1 - bf16_value(m);
    hi = (short)h;
    mid = (short)m;
    lo = (short)bf16_bits(r2);
}

// ------------------------------------------------------------------ //
// split-bf16 inner products: partial[s][i][j] = sum_{k in chunk s} X[i,k] Y[j,k]
// Workgroup: 64 x 64 output tile, four waves of 32 x 32.

__global__ __launch_bounds__(256) void inner_bf16x3_kernel(const float* __restrict__ X,
                                                          const float* __restrict__ Y, int N,
                                                          long K, long kchunk, int nsplit,
                                                          double* __restrict__ partial)
{
    const int tiles = (N + 63) / 64;
    const long nwork = (long)tiles * tiles * nsplit;
    const long w = xcd_work_index(blockIdx.x, nwork);
    if (w < 0) return;
    // consecutive work items share a K-chunk (and its rows in L2)
    const int tile = (int)(w % (tiles * tiles));
    const int s = (int)(w / (tiles * tiles));
    const int ti = tile / tiles, tj = tile % tiles;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int i0 = ti * 64 + (wave / 2) * 32, j0 = tj * 64 + (wave % 2) * 32;
    const int row = lane % 32, kq = 4 * (lane / 32);
    const int ia = i0 + row, jb = j0 + row;
    const bool va = ia < N, vb = jb < N;
    const float* xa = X + (long)(va ? ia : 0) * K;
    const float* yb = Y + (long)(vb ? jb : 0) * K;
    const long k_begin = (long)s * kchunk;
    const long k_end = std::min(K, k_begin + kchunk);

    // FP32 MFMA accumulation over 64 values of k at a time, then FP64
    f32x16 acc = {};
    double acc64[16];
#pragma unroll
    for (int v = 0; v < 16; v++) acc64[v] = 0.0;
    int steps = 0;
    for (long k0 = k_begin; k0 < k_end; k0 += 8) {
        bf16x4 ah, am, al, bh, bm, bl;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const long k = k0 + kq + q;
            const bool in = k < k_end;
            const float a = (va && in) ? xa[k] : 0.f;
            const float b = (vb && in) ? yb[k] : 0.f;
            short h, m, l;
            split3(a, h, m, l);
            ah[q] = h;
            am[q] = m;
            al[q] = l;
            split3(b, h, m, l);
            bh[q] = h;
            bm[q] = m;
            bl[q] = l;
        }
        // the six products down to 2^-16 relative of x y; the dropped
        // ones (mid.lo, lo.mid, lo.lo) are below 2^-24
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bm, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(am, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(am, bm, acc, 0, 0, 0);
        if (++steps == 8) {
#pragma unroll
            for (int v = 0; v < 16; v++) {
                acc64[v] += (double)acc[v];
                acc[v] = 0.f;
            }
            steps = 0;
        }
    }
    // D layout: column j0 + lane % 32, rows 8 (v / 4) + 4 (lane / 32) + v % 4
    double* P = partial + (long)s * N * N;
    const int jo = j0 + lane % 32;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int io = i0 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
        if (io < N && jo < N) P[(long)io * N + jo] = acc64[v] + (double)acc[v];
    }
}

// FP64 inner products on v_mfma_f64_16x16x4: wave tile 32 x 32 (2 x 2
// blocks), workgroup 64 x 64.
__global__ __launch_bounds__(256) void inner_f64_kernel(const double* __restrict__ X,
                                                       const double* __restrict__ Y, int N,
                                                       long K, long kchunk, int nsplit,
                                                       double* __restrict__ partial)
{
    const int tiles = (N + 63) / 64;
    const long nwork = (long)tiles * tiles * nsplit;
    const long w = xcd_work_index(blockIdx.x, nwork);
    if (w < 0) return;
    const int tile = (int)(w % (tiles * tiles));
    const int s = (int)(w / (tiles * tiles));
    const int ti = tile / tiles, tj = tile % tiles;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int i0 = ti * 64 + (wave / 2) * 32, j0 = tj * 64 + (wave % 2) * 32;
    const int r16 = lane % 16, kk = lane / 16;
    const long k_begin = (long)s * kchunk;
    const long k_end = std::min(K, k_begin + kchunk);
    const double* xa[2];
    const double* yb[2];
    bool va[2], vb[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
        const int ia = i0 + 16 * b + r16, jb = j0 + 16 * b + r16;
        va[b] = ia < N;
        vb[b] = jb < N;
        xa[b] = X + (long)(va[b] ? ia : 0) * K;
        yb[b] = Y + (long)(vb[b] ? jb : 0) * K;
    }
    f64x4 acc[2][2] = {};
    for (long k0 = k_begin; k0 < k_end; k0 += 4) {
        const long k = k0 + kk;
        const bool in = k < k_end;
        double a[2], b[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            a[q] = (va[q] && in) ? xa[q][k] : 0.0;
            b[q] = (vb[q] && in) ? yb[q][k] : 0.0;
        }
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int q = 0; q < 2; q++)
                acc[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[p], b[q], acc[p][q], 0, 0, 0);
    }
    double* P = partial + (long)s * N * N;
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int jo = j0 + 16 * q + r16;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int io = i0 + 16 * p + 4 * v + kk;   // f64 D: row 4v + lane/16
                if (io < N && jo < N) P[(long)io * N + jo] = acc[p][q][v];
            }
        }
}

// out[i][j] = alpha * sum_s P[s][i][j]  (or 0.5 alpha (T_ij + T_ji))
template <typename TP>
__global__ void reduce_partials_kernel(const TP* __restrict__ P, int N, int nsplit, double alpha,
                                       int hermitian, double* __restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)N * N) return;
    const int i = (int)(idx / N), j = (int)(idx % N);
    const long NN = (long)N * N;
    double t = 0.0;
    for (int s = 0; s < nsplit; s++) t += (double)P[s * NN + idx];
    if (hermitian) {
        double u = 0.0;
        const long tr = (long)j * N + i;
        for (int s = 0; s < nsplit; s++) u += (double)P[s * NN + tr];
        out[idx] = 0.5 * alpha * (t + u);
    } else {
        out[idx] = alpha * t;
    }
}

// ------------------------------------------------------------------ //
// rotations: Y[n][k] = sum_m C[n][m] X[m][k]

// FP32: workgroup = 4 waves, output tile 32 rows x 128 columns (one 32 x 32
// block per wave); m in steps of 2 (v_mfma_f32_32x32x2_f32), accumulator
// flushed into FP64 every 32 steps of m.
__global__ __launch_bounds__(256) void rotate_f32_kernel(const double* __restrict__ C,
                                                        const float* __restrict__ X, int N,
                                                        long K, float* __restrict__ Y)
{
    const long ktiles = (K + 127) / 128;
    const int ntiles = (N + 31) / 32;
    const long nwork = ktiles * ntiles;
    const long w = xcd_work_index(blockIdx.x, nwork);
    if (w < 0) return;
    const int nt = (int)(w % ntiles);     // row tiles of one K tile are neighbours
    const long kt = w / ntiles;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int n0 = nt * 32;
    const long k0 = kt * 128 + wave * 32;
    const int l32 = lane % 32, half = lane / 32;
    const int na = n0 + l32;
    const long kb = k0 + l32;
    const bool va = na < N, vb = kb < K;
    f32x16 acc = {};
    double acc64[16];
#pragma unroll
    for (int v = 0; v < 16; v++) acc64[v] = 0.0;
    int cnt = 0;
    for (int m0 = 0; m0 < N; m0 += 2) {
        const int m = m0 + half;
        const float a = (va && m < N) ? (float)C[(long)na * N + m] : 0.f;
        const float b = (vb && m < N) ? X[(long)m * K + kb] : 0.f;
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        if (++cnt == 32) {
#pragma unroll
            for (int v = 0; v < 16; v++) {
                acc64[v] += (double)acc[v];
                acc[v] = 0.f;
            }
            cnt = 0;
        }
    }
#pragma unroll
    for (int v = 0; v < 16; v++) {
        acc64[v] += (double)acc[v];
        const int no = n0 + 8 * (v / 4) + 4 * half + (v % 4);
        if (no < N && vb) Y[(long)no * K + kb] = (float)acc64[v];
    }
}

// FP64: wave tile 16 rows x 64 columns (4 blocks of 16 x 16), workgroup
// 4 waves over 64 rows.
__global__ __launch_bounds__(256) void rotate_f64_kernel(const double* __restrict__ C,
                                                        const double* __restrict__ X, int N,
                                                        long K, double* __restrict__ Y)
{
    const long ktiles = (K + 63) / 64;
    const int ntiles = (N + 63) / 64;
    const long nwork = ktiles * ntiles;
    const long w = xcd_work_index(blockIdx.x, nwork);
    if (w < 0) return;
    const int nt = (int)(w % ntiles);
    const long kt = w / ntiles;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int n0 = nt * 64 + wave * 16;
    const long k0 = kt * 64;
    const int r16 = lane % 16, kk = lane / 16;
    const int na = n0 + r16;
    f64x4 acc[4] = {};
    for (int m0 = 0; m0 < N; m0 += 4) {
        const int m = m0 + kk;
        const double a = (na < N && m < N) ? C[(long)na * N + m] : 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const long kb = k0 + 16 * q + r16;
            const double b = (kb < K && m < N) ? X[(long)m * K + kb] : 0.0;
            acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const long ko = k0 + 16 * q + r16;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int no = n0 + 4 * v + kk;
            if (no < N && ko < K) Y[(long)no * K + ko] = acc[q][v];
        }
    }
}

// ------------------------------------------------------------------ //
// PAW term: A[i][j] += sum_a sum_{l,k} P2[i][M_a+l] dA_a[l][k] P1[j][M_a+k]

template <typename TP>
__global__ void paw_matrix_kernel(int N, int Mmax, int natoms, const int* __restrict__ M_a,
                                  const int* __restrict__ ni_a, const int* __restrict__ dA_off,
                                  const double* __restrict__ dA, const TP* __restrict__ P1,
                                  const TP* __restrict__ P2, double* __restrict__ A)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)N * N) return;
    const int i = (int)(idx / N), j = (int)(idx % N);
    double s = 0.0;
    for (int a = 0; a < natoms; a++) {
        const int M = M_a[a], ni = ni_a[a];
        const double* D = dA + dA_off[a];
        for (int k = 0; k < ni; k++) {
            double q = 0.0;
            for (int l = 0; l < ni; l++) q += (double)P2[(long)i * Mmax + M + l] * D[l * ni + k];
            s += q * (double)P1[(long)j * Mmax + M + k];
        }
    }
    A[idx] += s;
}

static int pick_nsplit(int N, long K)
{
    const int tiles = (N + 63) / 64;
    long want = std::max<long>(1, 2048 / ((long)tiles * tiles));
    return (int)std::max<long>(1, std::min<long>(want, K / 1024));
}

}  // namespace gpd

using namespace gpd;

extern "C" int gpd_inner_products(int dtype, int N, long K, const void* d_X, const void* d_Y,
                                  double alpha, int hermitian, void* d_work, long work_bytes,
                                  double* d_out, void* stream)
{
    GPD_REQUIRE(N >= 0 && K >= 0, GPD_ERR_ARG, "bad sizes");
    if (N == 0) return GPD_OK;
    GPD_REQUIRE(d_X && d_Y && d_out, GPD_ERR_ARG, "null array");
    hipStream_t st = (hipStream_t)stream;
    const int nsplit = pick_nsplit(N, K);
    const size_t need = (size_t)nsplit * N * N * sizeof(double);
    void* work = d_work;
    bool own = false;
    if (!work || (size_t)work_bytes < need) {
        GPD_CHECK_HIP(hipMallocAsync(&work, need, st));
        own = true;
    }
    long kchunk = (K + nsplit - 1) / nsplit;
    kchunk = ((kchunk + 7) / 8) * 8;
    const int tiles = (N + 63) / 64;
    const long nwork = (long)tiles * tiles * nsplit;
    int rc = GPD_OK;
    if (dtype == GPD_F32) {
        hipLaunchKernelGGL(inner_bf16x3_kernel, dim3(xcd_grid_size(nwork)), dim3(256), 0, st,
                           (const float*)d_X, (const float*)d_Y, N, K, kchunk, nsplit,
                           (double*)work);
        rc = last_error_to_status(hipGetLastError());
        if (!rc) {
            long n = (long)N * N;
            hipLaunchKernelGGL(reduce_partials_kernel<double>, dim3((n + 255) / 256), dim3(256),
                               0, st, (const double*)work, N, nsplit, alpha, hermitian, d_out);
            rc = last_error_to_status(hipGetLastError());
        }
    } else if (dtype == GPD_F64) {
        hipLaunchKernelGGL(inner_f64_kernel, dim3(xcd_grid_size(nwork)), dim3(256), 0, st,
                           (const double*)d_X, (const double*)d_Y, N, K, kchunk, nsplit,
                           (double*)work);
        rc = last_error_to_status(hipGetLastError());
        if (!rc) {
            long n = (long)N * N;
            hipLaunchKernelGGL(reduce_partials_kernel<double>, dim3((n + 255) / 256), dim3(256),
                               0, st, (const double*)work, N, nsplit, alpha, hermitian, d_out);
            rc = last_error_to_status(hipGetLastError());
        }
    } else {
        rc = GPD_ERR_ARG;
        set_last_error("unknown dtype");
    }
    if (own) (void)hipFreeAsync(work, st);
    return rc;
}

extern "C" int gpd_rotate(int dtype, int N, long K, const double* d_C, const void* d_X, void* d_Y,
                          void* stream)
{
    GPD_REQUIRE(N >= 0 && K >= 0, GPD_ERR_ARG, "bad sizes");
    if (N == 0 || K == 0) return GPD_OK;
    GPD_REQUIRE(d_C && d_X && d_Y && d_X != d_Y, GPD_ERR_ARG, "bad arrays (out must differ)");
    hipStream_t st = (hipStream_t)stream;
    if (dtype == GPD_F32) {
        const long nwork = ((K + 127) / 128) * ((N + 31) / 32);
        hipLaunchKernelGGL(rotate_f32_kernel, dim3(xcd_grid_size(nwork)), dim3(256), 0, st, d_C,
                           (const float*)d_X, N, K, (float*)d_Y);
        return last_error_to_status(hipGetLastError());
    }
    if (dtype == GPD_F64) {
        const long nwork = ((K + 63) / 64) * ((N + 63) / 64);
        hipLaunchKernelGGL(rotate_f64_kernel, dim3(xcd_grid_size(nwork)), dim3(256), 0, st, d_C,
                           (const double*)d_X, N, K, (double*)d_Y);
        return last_error_to_status(hipGetLastError());
    }
    GPD_REQUIRE(false, GPD_ERR_ARG, "unknown dtype");
}

extern "C" int gpd_paw_matrix_add(int dtype, int N, int Mmax, int natoms, const int* d_M_a,
                                  const int* d_ni_a, const int* d_dA_off, const double* d_dA,
                                  const void* d_P1, const void* d_P2, double* d_A, void* stream)
{
    GPD_REQUIRE(N >= 0 && natoms >= 0, GPD_ERR_ARG, "bad sizes");
    if (N == 0 || natoms == 0) return GPD_OK;
    GPD_REQUIRE(d_M_a && d_ni_a && d_dA_off && d_dA && d_P1 && d_P2 && d_A, GPD_ERR_ARG,
                "null array");
    hipStream_t st = (hipStream_t)stream;
    long n = (long)N * N;
    if (dtype == GPD_F64)
        hipLaunchKernelGGL(paw_matrix_kernel<double>, dim3((n + 255) / 256), dim3(256), 0, st, N,
                           Mmax, natoms, d_M_a, d_ni_a, d_dA_off, d_dA, (const double*)d_P1,
                           (const double*)d_P2, d_A);
    else if (dtype == GPD_F32)
        hipLaunchKernelGGL(paw_matrix_kernel<float>, dim3((n + 255) / 256), dim3(256), 0, st, N,
                           Mmax, natoms, d_M_a, d_ni_a, d_dA_off, d_dA, (const float*)d_P1,
                           (const float*)d_P2, d_A);
    else
        GPD_REQUIRE(false, GPD_ERR_ARG, "unknown dtype");
    return last_error_to_status(hipGetLastError());
}
