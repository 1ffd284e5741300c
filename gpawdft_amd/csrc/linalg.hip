// Dense FP64 linear algebra of the subspace step, on the GPU (gfx950).
//
// inverse_cholesky (replaces c/lapack.c:324-370, dpotrf("U") + dtrtri on
// the C-ordered matrix): S = L L^T from the lower triangle of S, then
// S <- L^-1 (lower), upper triangle zeroed.  info = k > 0 when the leading
// minor of order k is not positive definite (S is then left partly
// factored, as LAPACK does).  One workgroup of 1024 threads streams the
// right-looking factorisation and the row-wise forward substitution
// through L2.
//
// diagonalize (replaces c/lapack.c:129-160, dsyev("V", "U") on the C view,
// i.e. the lower triangle): cyclic two-sided Jacobi with the round-robin
// ("tournament") pair order; every round rotates n/2 disjoint pairs at
// once -- each thread owns one 2 x 2 block and reads the rotation angles
// of its two pairs from the previous round's matrix (ping-pong buffers),
// so a round is a single launch.  Sweeps repeat until the off-diagonal
// norm is below 1e-14 sqrt(n) of the Frobenius norm.  Eigenvalues ascending in w,
// eigenvectors in the rows of A.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace gpd {

// ------------------------------------------------------------------ //
// Cholesky + inverse

__global__ __launch_bounds__(1024) void inverse_cholesky_kernel(double* __restrict__ A, int n,
                                                               double* __restrict__ X,
                                                               int* __restrict__ info)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    // symmetric input: only the lower triangle is read
    for (int k = 0; k < n; k++) {
        __syncthreads();
        const double d = A[(long)k * n + k];
        if (!(d > 0.0)) {
            if (tid == 0) *info = k + 1;
            return;      // uniform: every thread read the same d
        }
        const double l = sqrt(d);
        for (int i = k + 1 + tid; i < n; i += nt) A[(long)i * n + k] /= l;
        __syncthreads();
        if (tid == 0) A[(long)k * n + k] = l;
        const int m = n - k - 1;
        for (long idx = tid; idx < (long)m * m; idx += nt) {
            const int i = k + 1 + (int)(idx / m), j = k + 1 + (int)(idx % m);
            if (j <= i) A[(long)i * n + j] -= A[(long)i * n + k] * A[(long)j * n + k];
        }
    }
    __syncthreads();
    // X = L^-1 by rows: X[i][j] = -(sum_{k=j}^{i-1} L[i][k] X[k][j]) / L[i][i]
    for (int i = 0; i < n; i++) {
        const double lii = A[(long)i * n + i];
        for (int j = tid; j < n; j += nt) {
            double v;
            if (j > i) {
                v = 0.0;
            } else if (j == i) {
                v = 1.0 / lii;
            } else {
                double s = 0.0;
                for (int k = j; k < i; k++) s += A[(long)i * n + k] * X[(long)k * n + j];
                v = -s / lii;
            }
            X[(long)i * n + j] = v;
        }
        __syncthreads();
    }
    for (long idx = tid; idx < (long)n * n; idx += nt) A[idx] = X[idx];
    if (tid == 0) *info = 0;
}

// ------------------------------------------------------------------ //
// Jacobi eigensolver

// round-robin pairing: position 0 fixed, positions 1..n-1 rotate
__device__ __forceinline__ int player(int pos, int r, int n)
{
    return pos == 0 ? 0 : 1 + (pos - 1 + r) % (n - 1);
}

// rotation (c, s) that zeroes A[p][q] of the symmetric 2 x 2 block
__device__ __forceinline__ void jacobi_angle(double app, double aqq, double apq, double& c,
                                             double& s, bool& rot)
{
    rot = apq != 0.0 && fabs(apq) > 1e-300 &&
          fabs(apq) > 2.220446049250313e-16 * 1e-3 * sqrt(fabs(app * aqq));
    if (!rot) {
        c = 1.0;
        s = 0.0;
        return;
    }
    const double tau = (aqq - app) / (2.0 * apq);
    const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
    c = 1.0 / sqrt(1.0 + t * t);
    s = t * c;
}

// one round: Aout = J^T Ain J for the round's pairs; V <- V J
__global__ void jacobi_round_kernel(const double* __restrict__ Ain, double* __restrict__ Aout,
                                    double* __restrict__ V, int n, int r)
{
    const int h = n / 2;
    const long nblk = (long)h * h;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nblk) {
        const int a = (int)(idx / h), b = (int)(idx % h);
        const int pa = player(a, r, n), qa = player(n - 1 - a, r, n);
        const int pb = player(b, r, n), qb = player(n - 1 - b, r, n);
        double ca, sa, cb, sb;
        bool ra, rb;
        jacobi_angle(Ain[(long)pa * n + pa], Ain[(long)qa * n + qa], Ain[(long)pa * n + qa], ca,
                     sa, ra);
        jacobi_angle(Ain[(long)pb * n + pb], Ain[(long)qb * n + qb], Ain[(long)pb * n + qb], cb,
                     sb, rb);
        const double b00 = Ain[(long)pa * n + pb], b01 = Ain[(long)pa * n + qb];
        const double b10 = Ain[(long)qa * n + pb], b11 = Ain[(long)qa * n + qb];
        // columns: B J_b, J = [[c, s], [-s, c]]
        const double c00 = b00 * cb - b01 * sb, c01 = b00 * sb + b01 * cb;
        const double c10 = b10 * cb - b11 * sb, c11 = b10 * sb + b11 * cb;
        // rows: J_a^T (B J_b)
        double d00 = ca * c00 - sa * c10, d01 = ca * c01 - sa * c11;
        double d10 = sa * c00 + ca * c10, d11 = sa * c01 + ca * c11;
        if (a == b && ra) {
            d01 = 0.0;
            d10 = 0.0;
        }
        Aout[(long)pa * n + pb] = d00;
        Aout[(long)pa * n + qb] = d01;
        Aout[(long)qa * n + pb] = d10;
        Aout[(long)qa * n + qb] = d11;
    } else if (idx < nblk + (long)n * h) {
        const long t = idx - nblk;
        const int row = (int)(t / h), a = (int)(t % h);
        const int pa = player(a, r, n), qa = player(n - 1 - a, r, n);
        double ca, sa;
        bool ra;
        jacobi_angle(Ain[(long)pa * n + pa], Ain[(long)qa * n + qa], Ain[(long)pa * n + qa], ca,
                     sa, ra);
        if (!ra) return;
        const double vp = V[(long)row * n + pa], vq = V[(long)row * n + qa];
        V[(long)row * n + pa] = ca * vp - sa * vq;
        V[(long)row * n + qa] = sa * vp + ca * vq;
    }
}

__global__ void symmetrize_from_lower_kernel(const double* __restrict__ A, int n, int npad,
                                             double* __restrict__ B, double* __restrict__ V)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)npad * npad) return;
    const int i = (int)(idx / npad), j = (int)(idx % npad);
    double v = 0.0;
    if (i < n && j < n) v = i >= j ? A[(long)i * n + j] : A[(long)j * n + i];
    else if (i == j) v = 0.0;   // padding index: isolated zero eigenvalue
    B[idx] = v;
    V[idx] = i == j ? 1.0 : 0.0;
}

// off-diagonal and total squared norms, one workgroup
__global__ __launch_bounds__(1024) void offdiag_norm_kernel(const double* __restrict__ A, int n,
                                                           double* __restrict__ out)
{
    __shared__ double so[1024], sd[1024];
    double o = 0.0, d = 0.0;
    for (long idx = threadIdx.x; idx < (long)n * n; idx += blockDim.x) {
        const double v = A[idx];
        if (idx / n == idx % n) d += v * v;
        else o += v * v;
    }
    so[threadIdx.x] = o;
    sd[threadIdx.x] = d;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            so[threadIdx.x] += so[threadIdx.x + s];
            sd[threadIdx.x] += sd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = so[0];
        out[1] = sd[0];
    }
}

// eigenvalues ascending (ties by index) and eigenvectors into rows
__global__ void jacobi_finish_kernel(const double* __restrict__ B, const double* __restrict__ V,
                                     int n, int npad, double* __restrict__ w,
                                     double* __restrict__ A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double wi = B[(long)i * npad + i];
    int rank = 0;
    for (int j = 0; j < n; j++) {
        const double wj = B[(long)j * npad + j];
        if (wj < wi || (wj == wi && j < i)) rank++;
    }
    w[rank] = wi;
    for (int row = 0; row < n; row++) A[(long)rank * n + row] = V[(long)row * npad + i];
}

}  // namespace gpd

using namespace gpd;

extern "C" int gpd_inverse_cholesky(int n, double* d_A, int* info, void* stream)
{
    GPD_REQUIRE(n >= 0 && d_A && info, GPD_ERR_ARG, "bad arguments");
    *info = 0;
    if (n == 0) return GPD_OK;
    hipStream_t st = (hipStream_t)stream;
    double* X = nullptr;
    int* dinfo = nullptr;
    GPD_CHECK_HIP(hipMallocAsync((void**)&X, sizeof(double) * n * n + sizeof(int) * 4, st));
    dinfo = (int*)(X + (long)n * n);
    hipLaunchKernelGGL(inverse_cholesky_kernel, dim3(1), dim3(1024), 0, st, d_A, n, X, dinfo);
    int rc = last_error_to_status(hipGetLastError());
    int hinfo = 0;
    if (!rc) {
        hipError_t e = hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            set_last_error(hipGetErrorString(e));
            rc = GPD_ERR_HIP;
        }
    }
    (void)hipFreeAsync(X, st);
    *info = hinfo;
    return rc;
}

extern "C" int gpd_diagonalize(int n, double* d_A, double* d_w, int* sweeps, void* stream)
{
    GPD_REQUIRE(n >= 0 && d_A && d_w, GPD_ERR_ARG, "bad arguments");
    if (sweeps) *sweeps = 0;
    if (n == 0) return GPD_OK;
    hipStream_t st = (hipStream_t)stream;
    const int npad = n + (n & 1);   // even
    const long NN = (long)npad * npad;
    double* buf = nullptr;
    GPD_CHECK_HIP(hipMallocAsync((void**)&buf, sizeof(double) * (3 * NN + 2), st));
    double* B0 = buf;
    double* B1 = buf + NN;
    double* V = buf + 2 * NN;
    double* norms = buf + 3 * NN;
    int rc = GPD_OK;
    hipLaunchKernelGGL(symmetrize_from_lower_kernel, dim3((NN + 255) / 256), dim3(256), 0, st,
                       d_A, n, npad, B0, V);
    rc = last_error_to_status(hipGetLastError());
    const int h = npad / 2;
    const long work = (long)h * h + (long)npad * h;
    double* cur = B0;
    double* nxt = B1;
    int sweep = 0;
    for (; !rc && sweep < 40; sweep++) {
        for (int r = 0; r < npad - 1; r++) {
            hipLaunchKernelGGL(jacobi_round_kernel, dim3((work + 255) / 256), dim3(256), 0, st,
                               cur, nxt, V, npad, r);
            std::swap(cur, nxt);
        }
        hipLaunchKernelGGL(offdiag_norm_kernel, dim3(1), dim3(1024), 0, st, cur, npad, norms);
        double hn[2] = {0, 0};
        hipError_t e = hipMemcpyAsync(hn, norms, sizeof(hn), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            set_last_error(hipGetErrorString(e));
            rc = GPD_ERR_HIP;
            break;
        }
        // converged when the off-diagonal norm is at the rounding floor,
        // 1e-14 sqrt(n) of the Frobenius norm
        if (hn[0] <= 1e-28 * npad * (hn[0] + hn[1])) {
            sweep++;
            break;
        }
    }
    if (!rc) {
        hipLaunchKernelGGL(jacobi_finish_kernel, dim3((n + 255) / 256), dim3(256), 0, st, cur, V,
                           n, npad, d_w, d_A);
        rc = last_error_to_status(hipGetLastError());
    }
    (void)hipFreeAsync(buf, st);
    if (sweeps) *sweeps = sweep;
    return rc;
}
