// Training and inference kernels of the learned initial guess's network (FP32).
//
// Every dense-layer product, forward and backward, is one LDS-tiled GEMM kernel with
// general strides (so X^T dZ and dZ W^T read their operands in place, no transposed
// copies): 64 x 64 output tile per 256-lane work-group, 4 x 4 per lane, K in steps of
// 16 through LDS.  The forward epilogue adds the bias, keeps the pre-activation and
// applies SiLU; the weight-gradient product splits its long K (= rows of the batch)
// over blockIdx.z and accumulates with atomics.  FP32 MFMA gives no more than the
// MI355X's FP32 vector rate (157.3 TFLOP/s each), and these products are small next to
// the J/K builds, so the kernel stays on the vector ALU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/tonto_mlp.h"

namespace {

struct GemmArgs {
    const float *A;
    long sam, sak;  // A(m, k) = A[m * sam + k * sak]
    const float *B;
    long sbk, sbn;  // B(k, n) = B[k * sbk + n * sbn]
    float *C;
    long scm;       // C(m, n) = C[m * scm + n]
    long M;
    int N, K, kchunk;
    const float *bias;  // added per column (forward epilogue), or null
    float *Z;           // pre-activation copy (ld scm), or null
    int act;            // 1: SiLU
    int atomic;         // accumulate into C with atomics (split-K; no bias / act)
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }

__device__ __forceinline__ float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

constexpr int TM = 64, TN = 64, TK = 16;

__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g)
{
    __shared__ float As[TK][TM + 4];
    __shared__ float Bs[TK][TN + 4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const long m0 = (long)blockIdx.y * TM;
    const int n0 = blockIdx.x * TN;
    const int k_begin = blockIdx.z * g.kchunk;
    const int k_end = min(g.K, k_begin + g.kchunk);
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    // lanes walk the unit-stride index of each operand, so tile loads are coalesced
    const bool a_m_fast = g.sam == 1, b_n_fast = g.sbn == 1;
    for (int k0 = k_begin; k0 < k_end; k0 += TK) {
        for (int t = threadIdx.x; t < TM * TK; t += 256) {
            const int mm = a_m_fast ? t % TM : t / TK, kk = a_m_fast ? t / TM : t % TK;
            const long m = m0 + mm;
            const int k = k0 + kk;
            As[kk][mm] = (m < g.M && k < k_end) ? g.A[m * g.sam + (long)k * g.sak] : 0.0f;
        }
        for (int t = threadIdx.x; t < TN * TK; t += 256) {
            const int nn = b_n_fast ? t % TN : t / TK, kk = b_n_fast ? t / TN : t % TK;
            const int n = n0 + nn, k = k0 + kk;
            Bs[kk][nn] = (n < g.N && k < k_end) ? g.B[(long)k * g.sbk + (long)n * g.sbn] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long m = m0 + ty * 4 + i;
        if (m >= g.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + tx * 4 + j;
            if (n >= g.N) continue;
            float v = acc[i][j];
            if (g.atomic) {
                atomicAdd(&g.C[m * g.scm + n], v);
                continue;
            }
            if (g.bias) v += g.bias[n];
            if (g.Z) g.Z[m * g.scm + n] = v;
            if (g.act) v = silu(v);
            g.C[m * g.scm + n] = v;
        }
    }
}

void gemm(const GemmArgs &g0, hipStream_t s)
{
    GemmArgs g = g0;
    const long mt = (g.M + TM - 1) / TM;
    const int nt = (g.N + TN - 1) / TN;
    int kz = 1;
    if (g.atomic) {
        // split K so that the grid has about 2048 work-groups
        const long want = std::max(1L, 2048 / std::max(1L, mt * nt));
        kz = (int)std::min<long>(want, (g.K + 255) / 256);
        if (kz < 1) kz = 1;
    }
    g.kchunk = ((g.K + kz - 1) / kz + TK - 1) / TK * TK;
    kz = (g.K + g.kchunk - 1) / g.kchunk;
    if (kz < 1) kz = 1;
    gemm_kernel<<<dim3(nt, (unsigned)mt, kz), 256, 0, s>>>(g);
}

__global__ void fill_kernel(float *p, long n, float v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// dZ = dA * SiLU'(Z)
__global__ void silu_back_kernel(const float *dA, const float *Z, float *dZ, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dZ[i] = dA[i] * silu_grad(Z[i]);
}

// db[n] += sum over this work-group's rows of dZ[m][n] (blockDim.x >= N)
__global__ void colsum_kernel(const float *dZ, long M, int N, int rows, float *db)
{
    const int n = threadIdx.x;
    if (n >= N) return;
    const long r0 = (long)blockIdx.x * rows, r1 = min(M, r0 + rows);
    float acc = 0.0f;
    for (long m = r0; m < r1; ++m) acc += dZ[m * N + n];
    atomicAdd(&db[n], acc);
}

__global__ __launch_bounds__(256) void mse_kernel(const float *Y, const float *T, long n, float scale,
                                                  float *dY, float *loss)
{
    __shared__ float red[256];
    float acc = 0.0f;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dY[i] = 2.0f * scale * d;
        acc += d * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, scale * red[0]);
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, long n, float lr, float b1,
                            float b2, float eps, float c1, float c2)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

int grid_1d(long n) { return (int)std::min<long>((n + 255) / 256, 4096); }

struct Layout {
    long W1, b1, W2, b2, W3, b3, n;
};

Layout layout(int n_in, int nh, int n_out)
{
    Layout L;
    L.W1 = 0;
    L.b1 = L.W1 + (long)n_in * nh;
    L.W2 = L.b1 + nh;
    L.b2 = L.W2 + (long)nh * nh;
    L.W3 = L.b2 + nh;
    L.b3 = L.W3 + (long)nh * n_out;
    L.n = L.b3 + n_out;
    return L;
}

// limits: the column sums run one lane per column (<= 1024), grid.y holds the row tiles
constexpr long kMaxRows = 65535L * TM;

bool bad_dims(int n_in, int nh, int n_out, long M)
{
    return n_in <= 0 || nh <= 0 || n_out <= 0 || M < 0 || M > kMaxRows || nh > 1024 || n_out > 1024 ||
           n_in > (1 << 20);
}

}  // namespace

extern "C" long tmlp_n_params(int n_in, int nh, int n_out) { return layout(n_in, nh, n_out).n; }

extern "C" long tmlp_forward_workspace(long M, int nh) { return 4 * M * (long)nh; }

extern "C" long tmlp_backward_workspace(long M, int nh) { return 2 * M * (long)nh; }

extern "C" int tmlp_forward(const float *params, int n_in, int nh, int n_out, const float *X, long M,
                            float *ws, float *Y, void *stream)
{
    if (!params || !X || !ws || !Y || bad_dims(n_in, nh, n_out, M)) return TMLP_E_ARG;
    if (M == 0) return TMLP_OK;
    hipStream_t s = (hipStream_t)stream;
    const Layout L = layout(n_in, nh, n_out);
    float *Z1 = ws, *A1 = ws + M * nh, *Z2 = ws + 2 * M * nh, *A2 = ws + 3 * M * nh;
    GemmArgs g{};
    g.A = X; g.sam = n_in; g.sak = 1;
    g.B = params + L.W1; g.sbk = nh; g.sbn = 1;
    g.C = A1; g.scm = nh; g.M = M; g.N = nh; g.K = n_in;
    g.bias = params + L.b1; g.Z = Z1; g.act = 1; g.atomic = 0;
    gemm(g, s);
    g.A = A1; g.sam = nh;
    g.B = params + L.W2; g.sbk = nh;
    g.C = A2; g.K = nh; g.bias = params + L.b2; g.Z = Z2;
    gemm(g, s);
    g.A = A2; g.sam = nh;
    g.B = params + L.W3; g.sbk = n_out;
    g.C = Y; g.scm = n_out; g.N = n_out; g.K = nh; g.bias = params + L.b3; g.Z = nullptr; g.act = 0;
    gemm(g, s);
    return hipGetLastError() == hipSuccess ? TMLP_OK : TMLP_E_HIP;
}

extern "C" int tmlp_backward(const float *params, int n_in, int nh, int n_out, const float *X, long M,
                             const float *ws, const float *dY, float *grads, float *ws2, void *stream)
{
    if (!params || !X || !ws || !dY || !grads || !ws2 || bad_dims(n_in, nh, n_out, M)) return TMLP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const Layout L = layout(n_in, nh, n_out);
    fill_kernel<<<grid_1d(L.n), 256, 0, s>>>(grads, L.n, 0.0f);
    if (M == 0) return hipGetLastError() == hipSuccess ? TMLP_OK : TMLP_E_HIP;
    const float *Z1 = ws, *A1 = ws + M * nh, *Z2 = ws + 2 * M * nh, *A2 = ws + 3 * M * nh;
    float *dA = ws2, *dZ = ws2 + M * nh;
    const int rows = 512;
    const unsigned cs_grid = (unsigned)((M + rows - 1) / rows);
    GemmArgs g{};
    // layer 3: dW3 = A2^T dY, db3 = colsum dY, dA2 = dY W3^T
    g.A = A2; g.sam = 1; g.sak = nh;
    g.B = dY; g.sbk = n_out; g.sbn = 1;
    g.C = grads + L.W3; g.scm = n_out; g.M = nh; g.N = n_out; g.K = (int)M; g.atomic = 1;
    gemm(g, s);
    colsum_kernel<<<cs_grid, std::max(64, (n_out + 63) / 64 * 64), 0, s>>>(dY, M, n_out, rows, grads + L.b3);
    g = GemmArgs{};
    g.A = dY; g.sam = n_out; g.sak = 1;
    g.B = params + L.W3; g.sbk = 1; g.sbn = n_out;
    g.C = dA; g.scm = nh; g.M = M; g.N = nh; g.K = n_out;
    gemm(g, s);
    silu_back_kernel<<<grid_1d(M * nh), 256, 0, s>>>(dA, Z2, dZ, M * nh);
    // layer 2: dW2 = A1^T dZ2, db2, dA1 = dZ2 W2^T
    g = GemmArgs{};
    g.A = A1; g.sam = 1; g.sak = nh;
    g.B = dZ; g.sbk = nh; g.sbn = 1;
    g.C = grads + L.W2; g.scm = nh; g.M = nh; g.N = nh; g.K = (int)M; g.atomic = 1;
    gemm(g, s);
    colsum_kernel<<<cs_grid, std::max(64, (nh + 63) / 64 * 64), 0, s>>>(dZ, M, nh, rows, grads + L.b2);
    g = GemmArgs{};
    g.A = dZ; g.sam = nh; g.sak = 1;
    g.B = params + L.W2; g.sbk = 1; g.sbn = nh;
    g.C = dA; g.scm = nh; g.M = M; g.N = nh; g.K = nh;
    gemm(g, s);
    silu_back_kernel<<<grid_1d(M * nh), 256, 0, s>>>(dA, Z1, dZ, M * nh);
    // layer 1: dW1 = X^T dZ1, db1
    g = GemmArgs{};
    g.A = X; g.sam = 1; g.sak = n_in;
    g.B = dZ; g.sbk = nh; g.sbn = 1;
    g.C = grads + L.W1; g.scm = nh; g.M = n_in; g.N = nh; g.K = (int)M; g.atomic = 1;
    gemm(g, s);
    colsum_kernel<<<cs_grid, std::max(64, (nh + 63) / 64 * 64), 0, s>>>(dZ, M, nh, rows, grads + L.b1);
    return hipGetLastError() == hipSuccess ? TMLP_OK : TMLP_E_HIP;
}

extern "C" int tmlp_mse(const float *Y, const float *T, long n, float scale, float *dY, float *loss,
                        void *stream)
{
    if (!Y || !T || !dY || !loss || n < 0) return TMLP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    fill_kernel<<<1, 64, 0, s>>>(loss, 1, 0.0f);
    if (n) mse_kernel<<<grid_1d(n), 256, 0, s>>>(Y, T, n, scale, dY, loss);
    return hipGetLastError() == hipSuccess ? TMLP_OK : TMLP_E_HIP;
}

extern "C" int tmlp_adam(float *params, const float *grads, float *m, float *v, long n, float lr,
                         float beta1, float beta2, float eps, int step, void *stream)
{
    if (!params || !grads || !m || !v || n < 0 || step < 1) return TMLP_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    if (n) adam_kernel<<<grid_1d(n), 256, 0, s>>>(params, grads, m, v, n, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? TMLP_OK : TMLP_E_HIP;
}
