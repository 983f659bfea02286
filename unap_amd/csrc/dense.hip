// dense.hip -- inverse of the dense coarsest-level matrix on the GPU.
//
// The fixed V-cycle solves its coarsest level by one GEMV with the inverse of
// the coarsest Galerkin matrix, formed once at setup.  For a symmetric
// positive definite matrix this is in-place blocked Gauss-Jordan without
// pivoting (stable for SPD matrices), 32 columns a step:
//
//   P = A_kk^-1                       one workgroup, in LDS
//   A_kj = P A_kj          (j != k)   row panel
//   A_ij -= A_ik A_kj      (i, j != k) rank-32 update, FP64 MFMA 16x16x4
//   A_ik = -A_ik P         (i != k)   column panel;  A_kk = P
//
// The update carries 2 n^3 of the 2 n^3 + O(n^2 nb) flops and runs on MFMA;
// the host's unblocked O(n^3) loop it replaces takes seconds at n ~ 2700.
// The matrix is padded to a multiple of 64 with an identity block.
#include "dense.hpp"

#include "unap_hip.h"

#include <hip/hip_runtime.h>

#include <vector>

#include "device.hpp"

namespace unap_hip {

namespace {

constexpr int kNb = 32;    // Gauss-Jordan block
constexpr int kTile = 64;  // update tile: 4 waves of 32 x 32

typedef double double4_t __attribute__((ext_vector_type(4)));

// P = inverse of the kNb x kNb diagonal block k (Gauss-Jordan in LDS, no
// pivoting), written to P[kNb * kNb].
__global__ void __launch_bounds__(1024) kGJBlockInverse(const double* __restrict__ A, int N, int k,
                                                        double* __restrict__ P) {
    __shared__ double a[kNb][kNb + 1];
    __shared__ double inv[kNb][kNb + 1];
    const int r = threadIdx.x / kNb, c = threadIdx.x % kNb;  // 1024 threads: one per entry
    a[r][c] = A[(size_t)(k * kNb + r) * N + k * kNb + c];
    inv[r][c] = (r == c) ? 1.0 : 0.0;
    __syncthreads();
    for (int p = 0; p < kNb; ++p) {
        const double piv = 1.0 / a[p][p];
        __syncthreads();
        if (r == p) {
            a[p][c] *= piv;
            inv[p][c] *= piv;
        }
        __syncthreads();
        const double f = a[r][p];
        __syncthreads();
        if (r != p) {
            a[r][c] -= f * a[p][c];
            inv[r][c] -= f * inv[p][c];
        }
        __syncthreads();
    }
    P[r * kNb + c] = inv[r][c];
}

// A_kj = P A_kj for every column j outside block k: one thread per column.
__global__ void __launch_bounds__(256) kGJRowPanel(double* __restrict__ A, int N, int k, const double* __restrict__ P) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N || (j >= k * kNb && j < (k + 1) * kNb)) return;
    double v[kNb];
#pragma unroll
    for (int m = 0; m < kNb; ++m) v[m] = A[(size_t)(k * kNb + m) * N + j];
    for (int rr = 0; rr < kNb; ++rr) {
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < kNb; ++m) acc += P[rr * kNb + m] * v[m];
        A[(size_t)(k * kNb + rr) * N + j] = acc;
    }
}

// A_ij -= A_ik A_kj on 64 x 64 tiles, rows and columns of block k excluded.
// Wave w computes the 32 x 32 quarter (w / 2, w % 2) as 2 x 2 MFMA blocks
// over K = 32 in 8 steps of 4.  Operand layouts of v_mfma_f64_16x16x4f64:
// a of lane l = A[l & 15][l >> 4], b of lane l = B[l >> 4][l & 15], result
// element e of lane l = D[(l >> 4) + 4 e][l & 15].
__global__ void __launch_bounds__(256) kGJUpdate(double* __restrict__ A, int N, int k) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = blockIdx.y * kTile + (wave >> 1) * 32;
    const int col0 = blockIdx.x * kTile + (wave & 1) * 32;
    const int kb = k * kNb;
    double4_t acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < kNb / 4; ++ks) {
        const int kk = kb + 4 * ks + (lane >> 4);
        double a[2], b[2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) a[bi] = A[(size_t)(row0 + 16 * bi + (lane & 15)) * N + kk];
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) b[bj] = A[(size_t)kk * N + col0 + 16 * bj + (lane & 15)];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = row0 + 16 * bi + (lane >> 4) + 4 * e;
                const int j = col0 + 16 * bj + (lane & 15);
                const bool inK = (i >= kb && i < kb + kNb) || (j >= kb && j < kb + kNb);
                if (!inK) A[(size_t)i * N + j] -= acc[bi][bj][e];
            }
}

// A_ik = -A_ik P for every row i outside block k (one thread per row), and
// A_kk = P.
__global__ void __launch_bounds__(256) kGJColPanel(double* __restrict__ A, int N, int k, const double* __restrict__ P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int kb = k * kNb;
    if (i >= kb && i < kb + kNb) {
        for (int c = 0; c < kNb; ++c) A[(size_t)i * N + kb + c] = P[(i - kb) * kNb + c];
        return;
    }
    double v[kNb];
#pragma unroll
    for (int m = 0; m < kNb; ++m) v[m] = A[(size_t)i * N + kb + m];
    for (int c = 0; c < kNb; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < kNb; ++m) acc += v[m] * P[m * kNb + c];
        A[(size_t)i * N + kb + c] = -acc;
    }
}

// Out[0 .. n) rows of the padded inverse, row length n.
__global__ void kGJUnpad(const double* __restrict__ A, int N, int n, double* __restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const size_t i = t / n, j = t % n;
    out[t] = A[i * N + j];
}

}  // namespace

void denseInverseSPD(const std::vector<double>& a, int n, DBuf<double>& inv, hipStream_t s) {
    if (n <= 0) throw Error(UNAP_HIP_ERR_ARG, "empty dense matrix");
    const int N = (n + kTile - 1) / kTile * kTile;
    std::vector<double> padded((size_t)N * N, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) padded[(size_t)i * N + j] = a[(size_t)i * n + j];
    for (int i = n; i < N; ++i) padded[(size_t)i * N + i] = 1.0;
    DBuf<double> A, P(kNb * kNb);
    A.upload(padded, s);
    const dim3 tiles(N / kTile, N / kTile);
    for (int k = 0; k < N / kNb; ++k) {
        kGJBlockInverse<<<1, kNb * kNb, 0, s>>>(A.get(), N, k, P.get());
        UH_LAUNCHED();
        kGJRowPanel<<<(N + 255) / 256, 256, 0, s>>>(A.get(), N, k, P.get());
        UH_LAUNCHED();
        kGJUpdate<<<tiles, 256, 0, s>>>(A.get(), N, k);
        UH_LAUNCHED();
        kGJColPanel<<<(N + 255) / 256, 256, 0, s>>>(A.get(), N, k, P.get());
        UH_LAUNCHED();
    }
    inv.alloc((size_t)n * n);
    const size_t total = (size_t)n * n;
    kGJUnpad<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(A.get(), N, n, inv.get());
    UH_LAUNCHED();
    UH_CHECK(hipStreamSynchronize(s));
}

}  // namespace unap_hip
