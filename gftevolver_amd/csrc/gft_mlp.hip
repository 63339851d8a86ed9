// gft_mlp.hip -- the surrogate network: an MLP with BF16 weights and
// activations, FP32 accumulation on the MFMA matrix cores, trained on the GPU.
//
//   x[8] -> 4 hidden layers of width 256 with GELU -> H heads, each a linear
//   map 256 -> n_out (the heads share the trunk and are trained on different
//   bootstrap weights, so their spread measures the network's uncertainty).
//
// Every matrix product is one kernel, mlp_gemm: a wave computes a 16x16 tile
// of C = A B with gfx950's v_mfma_f32_16x16x32_bf16 (A, B in BF16, FP32
// accumulate; a 16-deep remainder of K with v_mfma_f32_16x16x16_bf16), A and B
// addressed through (row, column) strides so the transposed products of the
// backward pass need no copies.  Where an operand is contiguous along K (the
// forward products and the products with the weights) a lane reads its 8 K
// values as one 16-byte load; the transposed operands of the weight gradients
// are gathered.  The shapes are small (batch x 256 x 256), so the kernel keeps
// operands in registers without LDS staging.
// Master weights, gradients and Adam moments are FP32; the BF16 copies the
// GEMMs read are refreshed after each update.  All reductions (bias
// gradients, loss) run in a fixed order, so training is deterministic.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace gft {

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef short mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 mlp_bf16x8_t __attribute__((ext_vector_type(8)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short mlp_f2bf(float f) {   // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float mlp_bf2f(unsigned short h) { return __uint_as_float(static_cast<unsigned>(h) << 16); }

// GELU(z) = z Phi(z) and its derivative Phi(z) + z phi(z).
__device__ __forceinline__ float mlp_gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float mlp_gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

enum MlpGemmMode { MLP_STORE = 0, MLP_BIAS_GELU = 1, MLP_BIAS = 2 };

struct MlpGemm {
    const unsigned short* A;   // A(m, k) = A[m * sa_m + k * sa_k]
    long long sa_m, sa_k;
    const unsigned short* B;   // B(k, n) = B[k * sb_k + n * sb_n]
    long long sb_k, sb_n;
    int M, N, K;               // multiples of 16
    int mode;
    float* C;                  // MLP_STORE / MLP_BIAS: C[m * ldc + n]; MLP_BIAS_GELU: the pre-activation
    long long ldc;
    const float* bias;         // [N]
    unsigned short* H;         // MLP_BIAS_GELU: BF16 activation H[m * ldh + n]
    long long ldh;
};

// The 8 K values k0 + 8 q .. k0 + 8 q + 7 of a row (A) or column (B) of an
// operand whose K stride is s: one 16-byte load when VEC (s = 1, 16-byte
// aligned), else gathered.
template <bool VEC>
__device__ __forceinline__ mlp_bf16x8 mlp_load8(const unsigned short* p, long long s, int k) {
    if constexpr (VEC) {
        return *reinterpret_cast<const mlp_bf16x8*>(p + k);
    } else {
        mlp_bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = static_cast<short>(p[(long long)(k + j) * s]);
        return v;
    }
}

// One wave per 16x16 output tile, four waves per block.  For each 32-deep slice
// of K, lane l supplies A(row l%16, k0 + 8(l/16) + j) and B(k0 + 8(l/16) + j,
// col l%16), j = 0..7 (any assignment of the slice's K values to (l/16, j) that
// A and B share gives the same products); a 16-deep remainder uses 4 values
// per lane.  Lane l receives D(4(l/16) + j, l%16), j = 0..3.
template <bool VA, bool VB>
__global__ void __launch_bounds__(256) mlp_gemm(MlpGemm g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = g.N >> 4;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= (g.M >> 4) * tiles_n) return;   // whole waves only: MFMA needs every lane
    const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
    const int r = lane & 15, q = lane >> 4;
    mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const unsigned short* Ar = g.A + (long long)(tm * 16 + r) * g.sa_m;
    const unsigned short* Bc = g.B + (long long)(tn * 16 + r) * g.sb_n;
    const int K32 = g.K & ~31;
    for (int k0 = 0; k0 < K32; k0 += 32) {
        const mlp_bf16x8 a = mlp_load8<VA>(Ar, g.sa_k, k0 + 8 * q);
        const mlp_bf16x8 b = mlp_load8<VB>(Bc, g.sb_k, k0 + 8 * q);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_bf16x8_t, a),
                                                      __builtin_bit_cast(mlp_bf16x8_t, b), acc, 0, 0, 0);
    }
    if (K32 < g.K) {   // K is a multiple of 16: one 16-deep remainder
        mlp_bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long long k = K32 + 4 * q + j;
            a[j] = static_cast<short>(Ar[k * g.sa_k]);
            b[j] = static_cast<short>(Bc[k * g.sb_k]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int n = tn * 16 + r;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int m = tm * 16 + 4 * q + j;
        float v = acc[j];
        if (g.mode != MLP_STORE) v += g.bias[n];
        g.C[(long long)m * g.ldc + n] = v;
        if (g.mode == MLP_BIAS_GELU) g.H[(long long)m * g.ldh + n] = mlp_f2bf(mlp_gelu(v));
    }
}

// dZ = dH * GELU'(Z), to BF16 for the next products.
__global__ void __launch_bounds__(256) mlp_gelu_backward(const float* __restrict__ dH, const float* __restrict__ Z,
                                                         unsigned short* __restrict__ dZ, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = mlp_f2bf(dH[i] * mlp_gelu_grad(Z[i]));
}

// out[c] = sum_r src[r * ld + c] (BF16 source), rows in order.
__global__ void __launch_bounds__(256) mlp_colsum_bf16(const unsigned short* __restrict__ src, int rows, int cols,
                                                       long long ld, float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; r++) s += mlp_bf2f(src[(long long)r * ld + c]);
    out[c] = s;
}

// Masked mean-square loss of every head against the target, and its gradient
// dY = 2 w (Y - T) / (n_out * batch * heads) in BF16.  Y[b][h * n_out + o],
// T[b][o], w[b * heads + h].  One block per sample; partial[b] = its loss
// contribution (summed in order by mlp_sum_partials).
__global__ void __launch_bounds__(256) mlp_loss_grad(const float* __restrict__ Y, const float* __restrict__ T,
                                                     const float* __restrict__ w, int heads, int n_out, int batch,
                                                     unsigned short* __restrict__ dY, float* __restrict__ partial) {
    __shared__ float red[256];
    const int b = blockIdx.x;
    const float scale = 1.0f / ((float)n_out * (float)batch * (float)heads);
    float acc = 0.f;
    for (int i = threadIdx.x; i < heads * n_out; i += blockDim.x) {
        const int h = i / n_out, o = i - h * n_out;
        const float wt = w[b * heads + h];
        const float d = Y[(long long)b * heads * n_out + i] - T[(long long)b * n_out + o];
        acc += wt * d * d;
        dY[(long long)b * heads * n_out + i] = mlp_f2bf(2.0f * wt * d * scale);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[b] = red[0] * scale;
}

__global__ void mlp_sum_partials(const float* __restrict__ partial, int n, float* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.f;
        for (int i = 0; i < n; i++) s += partial[i];
        *out = s;
    }
}

// Adam step on FP32 master parameters; refreshes the BF16 copy the GEMMs read
// (bf_index[i] >= 0 gives its position, -1 for biases, which stay FP32).
__global__ void __launch_bounds__(256) mlp_adam(float* __restrict__ p, const float* __restrict__ g,
                                                float* __restrict__ m, float* __restrict__ v, long long n, float lr,
                                                float b1, float b2, float eps, float c1, float c2,
                                                const long long* __restrict__ bf_index,
                                                unsigned short* __restrict__ bf) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    const long long k = bf_index[i];
    if (k >= 0) bf[k] = mlp_f2bf(pi);
}

__global__ void __launch_bounds__(256) mlp_to_bf16(const float* __restrict__ p, const long long* __restrict__ bf_index,
                                                   unsigned short* __restrict__ bf, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long k = bf_index[i];
    if (k >= 0) bf[k] = mlp_f2bf(p[i]);
}

// Gathers a minibatch: X (FP32, n_in per sample) to BF16 rows of width
// in_pad (zero padded), targets and head weights as FP32 (skipped when null).
// idx[b] < 0 gives an all-zero padding row.
__global__ void __launch_bounds__(256) mlp_gather(const float* __restrict__ X, const float* __restrict__ T,
                                                  const float* __restrict__ W, const int* __restrict__ idx, int batch,
                                                  int n_in, int in_pad, int n_out, int heads,
                                                  unsigned short* __restrict__ xb, float* __restrict__ tb,
                                                  float* __restrict__ wb) {
    const int b = blockIdx.x;
    if (b >= batch) return;
    const long long s = idx[b];
    for (int i = threadIdx.x; i < in_pad; i += blockDim.x)
        xb[(long long)b * in_pad + i] = (s >= 0 && i < n_in) ? mlp_f2bf(X[s * n_in + i]) : 0;
    if (T)
        for (int o = threadIdx.x; o < n_out; o += blockDim.x) tb[(long long)b * n_out + o] = s >= 0 ? T[s * n_out + o] : 0.f;
    if (W)
        for (int h = threadIdx.x; h < heads; h += blockDim.x) wb[b * heads + h] = s >= 0 ? W[s * heads + h] : 0.f;
}

// FP32 gradient of a weight from the GEMM's FP32 output, scattered into the
// flat gradient vector (drops the zero-padded input columns of layer 0).
__global__ void __launch_bounds__(256) mlp_scatter_grad(const float* __restrict__ src, int rows, int cols_src,
                                                        int cols, float* __restrict__ dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    const long long r = i / cols, c = i - r * cols;
    dst[i] = src[r * cols_src + c];
}

}  // namespace gft

// ---- host side: the gft_mlp_* C ABI ---------------------------------------------

struct gft_mlp {
    gft_mlp_desc d{};
    int in_pad = 0;                        // n_in rounded up to 16
    int HO = 0;                            // heads * n_out
    long long P = 0;                       // number of FP32 parameters
    std::vector<long long> woff, boff;     // per layer (hidden..., head): weight / bias offsets in params
    std::vector<long long> bfoff;          // per layer: offset of the BF16 weight copy
    std::vector<int> lin, lin_pad, lout;   // per layer: inputs, padded inputs, outputs
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    long long* bf_index = nullptr;
    unsigned short* wbf = nullptr;
    unsigned short* xb = nullptr;          // [max_batch][in_pad]
    std::vector<float*> Z;                 // [max_batch][width] per hidden layer
    std::vector<unsigned short*> H;        // [max_batch][width] per hidden layer
    float *Y = nullptr, *dH = nullptr, *gtmp = nullptr, *tb = nullptr, *wb = nullptr, *partial = nullptr;
    unsigned short *dY = nullptr, *dZ = nullptr;
    float* loss_dev = nullptr;
    float *xstage = nullptr, *tstage = nullptr, *wstage = nullptr;
    int* idx = nullptr;
    hipStream_t st = nullptr;
    long long adam_t = 0;
    std::vector<void*> owned;
    std::string err;
};

namespace {

std::string g_mlp_create_error;

int mlp_fail(gft_mlp* m, int code, const std::string& msg) {
    if (m) m->err = msg; else g_mlp_create_error = msg;
    return code;
}

template <class T>
bool mlp_alloc(gft_mlp* m, T** p, long long n) {
    void* q = nullptr;
    if (hipMalloc(&q, (size_t)std::max<long long>(n, 1) * sizeof(T)) != hipSuccess) return false;
    m->owned.push_back(q);
    *p = static_cast<T*>(q);
    return true;
}

void mlp_launch_gemm(gft_mlp* m, const unsigned short* A, long long sa_m, long long sa_k, const unsigned short* B,
                     long long sb_k, long long sb_n, int M, int N, int K, int mode, float* C, long long ldc,
                     const float* bias, unsigned short* H, long long ldh) {
    gft::MlpGemm g{A, sa_m, sa_k, B, sb_k, sb_n, M, N, K, mode, C, ldc, bias, H, ldh};
    const long long tiles = (long long)(M / 16) * (N / 16);
    // An operand takes the 16-byte loads when it is contiguous along K and every
    // row's (column's) start is 16-byte aligned.
    auto vec = [](const unsigned short* p, long long s_k, long long s_other) {
        return s_k == 1 && s_other % 8 == 0 && reinterpret_cast<uintptr_t>(p) % 16 == 0;
    };
    const bool va = vec(A, sa_k, sa_m), vb = vec(B, sb_k, sb_n);
    const dim3 grid((unsigned)((tiles + 3) / 4)), block(256);
    if (va && vb) gft::mlp_gemm<true, true><<<grid, block, 0, m->st>>>(g);
    else if (va) gft::mlp_gemm<true, false><<<grid, block, 0, m->st>>>(g);
    else if (vb) gft::mlp_gemm<false, true><<<grid, block, 0, m->st>>>(g);
    else gft::mlp_gemm<false, false><<<grid, block, 0, m->st>>>(g);
}

// Forward pass of a batch already gathered into xb (batch a multiple of 16).
void mlp_forward_dev(gft_mlp* m, int batch) {
    const int L = m->d.n_hidden;
    const unsigned short* prev = m->xb;
    for (int l = 0; l < L; l++) {
        mlp_launch_gemm(m, prev, m->lin_pad[l], 1, m->wbf + m->bfoff[l], 1, m->lin_pad[l], batch, m->d.width,
                        m->lin_pad[l], gft::MLP_BIAS_GELU, m->Z[l], m->d.width, m->params + m->boff[l], m->H[l],
                        m->d.width);
        prev = m->H[l];
    }
    mlp_launch_gemm(m, prev, m->d.width, 1, m->wbf + m->bfoff[L], 1, m->d.width, batch, m->HO, m->d.width, gft::MLP_BIAS,
                    m->Y, m->HO, m->params + m->boff[L], nullptr, 0);
}

// Loss, its gradient and the backward pass into m->grads; loss to m->loss_dev.
void mlp_backward_dev(gft_mlp* m, int batch) {
    const int L = m->d.n_hidden, W = m->d.width, HO = m->HO;
    gft::mlp_loss_grad<<<batch, 256, 0, m->st>>>(m->Y, m->tb, m->wb, m->d.n_heads, m->d.n_out, batch, m->dY,
                                                   m->partial);
    gft::mlp_sum_partials<<<1, 64, 0, m->st>>>(m->partial, batch, m->loss_dev);
    // Head: dW = dY^T H_last, db = colsum dY, dH_last = dY W_head.
    const unsigned short* Hl = m->H[L - 1];
    mlp_launch_gemm(m, m->dY, 1, HO, Hl, W, 1, HO, W, batch, gft::MLP_STORE, m->gtmp, W, nullptr, nullptr, 0);
    gft::mlp_scatter_grad<<<(unsigned)(((long long)HO * W + 255) / 256), 256, 0, m->st>>>(m->gtmp, HO, W, W,
                                                                                        m->grads + m->woff[L]);
    gft::mlp_colsum_bf16<<<(HO + 255) / 256, 256, 0, m->st>>>(m->dY, batch, HO, HO, m->grads + m->boff[L]);
    mlp_launch_gemm(m, m->dY, HO, 1, m->wbf + m->bfoff[L], W, 1, batch, W, HO, gft::MLP_STORE, m->dH, W, nullptr,
                    nullptr, 0);
    for (int l = L - 1; l >= 0; l--) {
        const long long n = (long long)batch * W;
        gft::mlp_gelu_backward<<<(unsigned)((n + 255) / 256), 256, 0, m->st>>>(m->dH, m->Z[l], m->dZ, n);
        const unsigned short* Hin = l > 0 ? m->H[l - 1] : m->xb;
        const int ip = m->lin_pad[l];
        mlp_launch_gemm(m, m->dZ, 1, W, Hin, ip, 1, W, ip, batch, gft::MLP_STORE, m->gtmp, ip, nullptr, nullptr, 0);
        gft::mlp_scatter_grad<<<(unsigned)(((long long)W * m->lin[l] + 255) / 256), 256, 0, m->st>>>(
            m->gtmp, W, ip, m->lin[l], m->grads + m->woff[l]);
        gft::mlp_colsum_bf16<<<(W + 255) / 256, 256, 0, m->st>>>(m->dZ, batch, W, W, m->grads + m->boff[l]);
        if (l > 0)
            mlp_launch_gemm(m, m->dZ, W, 1, m->wbf + m->bfoff[l], ip, 1, batch, ip, W, gft::MLP_STORE, m->dH, W,
                            nullptr, nullptr, 0);
    }
}

}  // namespace

extern "C" {

int gft_mlp_create(const gft_mlp_desc* d, gft_mlp** out) {
    if (!d || !out) return mlp_fail(nullptr, GFT_EINVAL, "gft_mlp_create: null argument");
    *out = nullptr;
    if (d->n_in < 1 || d->n_hidden < 1 || d->n_heads < 1 || d->width < 16 || d->width % 16 || d->n_out < 16 ||
        d->n_out % 16 || d->max_batch < 16 || d->max_batch % 16)
        return mlp_fail(nullptr, GFT_EINVAL,
                        "gft_mlp_create: width, n_out and max_batch must be positive multiples of 16");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return mlp_fail(nullptr, GFT_ENODEV, "no HIP device");
    gft_mlp* m = new gft_mlp();
    m->d = *d;
    m->in_pad = (d->n_in + 15) / 16 * 16;
    m->HO = d->n_heads * d->n_out;
    const int L = d->n_hidden;
    long long off = 0, bf = 0;
    for (int l = 0; l <= L; l++) {
        const int in = l == 0 ? d->n_in : d->width;
        const int inp = l == 0 ? m->in_pad : d->width;
        const int o = l == L ? m->HO : d->width;
        m->lin.push_back(in);
        m->lin_pad.push_back(inp);
        m->lout.push_back(o);
        m->woff.push_back(off);
        off += (long long)o * in;
        m->boff.push_back(off);
        off += o;
        m->bfoff.push_back(bf);
        bf += (long long)o * inp;
    }
    m->P = off;
    const long long B = d->max_batch, W = d->width;
    bool ok = hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) == hipSuccess;
    ok = ok && mlp_alloc(m, &m->params, m->P) && mlp_alloc(m, &m->grads, m->P) && mlp_alloc(m, &m->adam_m, m->P) &&
         mlp_alloc(m, &m->adam_v, m->P) && mlp_alloc(m, &m->bf_index, m->P) && mlp_alloc(m, &m->wbf, bf) &&
         mlp_alloc(m, &m->xb, B * m->in_pad) && mlp_alloc(m, &m->Y, B * m->HO) && mlp_alloc(m, &m->dY, B * m->HO) &&
         mlp_alloc(m, &m->dH, B * W) && mlp_alloc(m, &m->dZ, B * W) &&
         mlp_alloc(m, &m->gtmp, std::max<long long>((long long)m->HO * W, W * W)) && mlp_alloc(m, &m->tb, B * d->n_out) &&
         mlp_alloc(m, &m->wb, B * d->n_heads) && mlp_alloc(m, &m->partial, B) && mlp_alloc(m, &m->loss_dev, 1) &&
         mlp_alloc(m, &m->xstage, B * d->n_in) && mlp_alloc(m, &m->tstage, B * d->n_out) &&
         mlp_alloc(m, &m->wstage, B * d->n_heads) && mlp_alloc(m, &m->idx, B);
    for (int l = 0; ok && l < L; l++) {
        float* z = nullptr;
        unsigned short* h = nullptr;
        ok = mlp_alloc(m, &z, B * W) && mlp_alloc(m, &h, B * W);
        m->Z.push_back(z);
        m->H.push_back(h);
    }
    if (!ok) {
        gft_mlp_destroy(m);
        return mlp_fail(nullptr, GFT_ENOMEM, "gft_mlp_create: device allocation failed");
    }
    // Map each parameter to its BF16 copy (weights) or -1 (biases).
    std::vector<long long> map(m->P, -1);
    for (int l = 0; l <= L; l++)
        for (long long r = 0; r < m->lout[l]; r++)
            for (long long c = 0; c < m->lin[l]; c++) map[m->woff[l] + r * m->lin[l] + c] = m->bfoff[l] + r * m->lin_pad[l] + c;
    ok = hipMemcpy(m->bf_index, map.data(), m->P * sizeof(long long), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(m->wbf, 0, bf * sizeof(unsigned short)) == hipSuccess &&
         hipMemset(m->params, 0, m->P * sizeof(float)) == hipSuccess &&
         hipMemset(m->adam_m, 0, m->P * sizeof(float)) == hipSuccess &&
         hipMemset(m->adam_v, 0, m->P * sizeof(float)) == hipSuccess;
    if (!ok) {
        gft_mlp_destroy(m);
        return mlp_fail(nullptr, GFT_EHIP, "gft_mlp_create: initialisation failed");
    }
    *out = m;
    return GFT_OK;
}

void gft_mlp_destroy(gft_mlp* m) {
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    for (void* p : m->owned) (void)hipFree(p);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

const char* gft_mlp_last_error(const gft_mlp* m) { return m ? m->err.c_str() : g_mlp_create_error.c_str(); }

long long gft_mlp_num_params(const gft_mlp* m) { return m ? m->P : 0; }

int gft_mlp_set_params(gft_mlp* m, const float* params) {
    if (!m || !params) return GFT_EINVAL;
    if (hipMemcpyAsync(m->params, params, m->P * sizeof(float), hipMemcpyHostToDevice, m->st) != hipSuccess)
        return mlp_fail(m, GFT_EHIP, "gft_mlp_set_params: copy failed");
    gft::mlp_to_bf16<<<(unsigned)((m->P + 255) / 256), 256, 0, m->st>>>(m->params, m->bf_index, m->wbf, m->P);
    if (hipMemsetAsync(m->adam_m, 0, m->P * sizeof(float), m->st) != hipSuccess ||
        hipMemsetAsync(m->adam_v, 0, m->P * sizeof(float), m->st) != hipSuccess ||
        hipStreamSynchronize(m->st) != hipSuccess)
        return mlp_fail(m, GFT_EHIP, "gft_mlp_set_params: failed");
    m->adam_t = 0;
    return GFT_OK;
}

int gft_mlp_get_params(gft_mlp* m, float* params) {
    if (!m || !params) return GFT_EINVAL;
    if (hipMemcpyAsync(params, m->params, m->P * sizeof(float), hipMemcpyDeviceToHost, m->st) != hipSuccess ||
        hipStreamSynchronize(m->st) != hipSuccess)
        return mlp_fail(m, GFT_EHIP, "gft_mlp_get_params: copy failed");
    return GFT_OK;
}

// y[s][h][o] for n samples x[s][n_in], in chunks of max_batch.
int gft_mlp_forward(gft_mlp* m, const float* x, int n, float* y) {
    if (!m || !x || !y || n < 0) return GFT_EINVAL;
    const int B = m->d.max_batch;
    std::vector<int> ids(B);
    for (int s0 = 0; s0 < n; s0 += B) {
        const int cnt = std::min(B, n - s0), pad = (cnt + 15) / 16 * 16;
        for (int b = 0; b < pad; b++) ids[b] = b < cnt ? b : -1;
        if (hipMemcpyAsync(m->xstage, x + (long long)s0 * m->d.n_in, (size_t)cnt * m->d.n_in * sizeof(float),
                           hipMemcpyHostToDevice, m->st) != hipSuccess ||
            hipMemcpyAsync(m->idx, ids.data(), pad * sizeof(int), hipMemcpyHostToDevice, m->st) != hipSuccess)
            return mlp_fail(m, GFT_EHIP, "gft_mlp_forward: upload failed");
        gft::mlp_gather<<<pad, 64, 0, m->st>>>(m->xstage, nullptr, nullptr, m->idx, pad, m->d.n_in, m->in_pad,
                                                m->d.n_out, m->d.n_heads, m->xb, m->tb, m->wb);
        mlp_forward_dev(m, pad);
        if (hipMemcpyAsync(y + (long long)s0 * m->HO, m->Y, (size_t)cnt * m->HO * sizeof(float), hipMemcpyDeviceToHost,
                           m->st) != hipSuccess ||
            hipStreamSynchronize(m->st) != hipSuccess)
            return mlp_fail(m, GFT_EHIP, "gft_mlp_forward: failed");
    }
    return GFT_OK;
}

// Loss and FP32 gradient (flat, parameter order) of one batch: t[s][n_out]
// targets, w[s][heads] per-head sample weights.
int gft_mlp_gradients(gft_mlp* m, const float* x, const float* t, const float* w, int batch, float* loss,
                      float* grad) {
    if (!m || !x || !t || !w || !loss || !grad) return GFT_EINVAL;
    if (batch < 16 || batch % 16 || batch > m->d.max_batch)
        return mlp_fail(m, GFT_EINVAL, "gft_mlp_gradients: batch must be a multiple of 16 up to max_batch");
    std::vector<int> ids(batch);
    for (int b = 0; b < batch; b++) ids[b] = b;
    if (hipMemcpyAsync(m->xstage, x, (size_t)batch * m->d.n_in * sizeof(float), hipMemcpyHostToDevice, m->st) ||
        hipMemcpyAsync(m->tstage, t, (size_t)batch * m->d.n_out * sizeof(float), hipMemcpyHostToDevice, m->st) ||
        hipMemcpyAsync(m->wstage, w, (size_t)batch * m->d.n_heads * sizeof(float), hipMemcpyHostToDevice, m->st) ||
        hipMemcpyAsync(m->idx, ids.data(), batch * sizeof(int), hipMemcpyHostToDevice, m->st))
        return mlp_fail(m, GFT_EHIP, "gft_mlp_gradients: upload failed");
    gft::mlp_gather<<<batch, 64, 0, m->st>>>(m->xstage, m->tstage, m->wstage, m->idx, batch, m->d.n_in, m->in_pad,
                                              m->d.n_out, m->d.n_heads, m->xb, m->tb, m->wb);
    mlp_forward_dev(m, batch);
    mlp_backward_dev(m, batch);
    if (hipMemcpyAsync(grad, m->grads, m->P * sizeof(float), hipMemcpyDeviceToHost, m->st) ||
        hipMemcpyAsync(loss, m->loss_dev, sizeof(float), hipMemcpyDeviceToHost, m->st) ||
        hipStreamSynchronize(m->st))
        return mlp_fail(m, GFT_EHIP, "gft_mlp_gradients: failed");
    return GFT_OK;
}

// Adam training on the GPU: n samples (x[n][n_in], t[n][n_out], w[n][heads]),
// minibatches of `batch` (multiple of 16) from a seeded shuffle per epoch
// (the last partial batch is dropped), learning rate lr.  epoch_loss[e] gets
// the mean minibatch loss of epoch e.
int gft_mlp_train(gft_mlp* m, const float* x, const float* t, const float* w, int n, int batch, int epochs, float lr,
                  unsigned seed, float* epoch_loss) {
    if (!m || !x || !t || !w || n < 1 || epochs < 0) return GFT_EINVAL;
    if (batch < 16 || batch % 16 || batch > m->d.max_batch || batch > n)
        return mlp_fail(m, GFT_EINVAL, "gft_mlp_train: batch must be a multiple of 16, at most max_batch and n");
    const int nb = n / batch;
    float *X = nullptr, *T = nullptr, *Wt = nullptr, *hist = nullptr;
    int* perm = nullptr;
    auto cleanup = [&]() {
        for (void* p : {(void*)X, (void*)T, (void*)Wt, (void*)hist, (void*)perm})
            if (p) (void)hipFree(p);
    };
    if (hipMalloc(&X, (size_t)n * m->d.n_in * sizeof(float)) || hipMalloc(&T, (size_t)n * m->d.n_out * sizeof(float)) ||
        hipMalloc(&Wt, (size_t)n * m->d.n_heads * sizeof(float)) ||
        hipMalloc(&hist, (size_t)std::max(1, epochs * nb) * sizeof(float)) || hipMalloc(&perm, (size_t)n * sizeof(int))) {
        cleanup();
        return mlp_fail(m, GFT_ENOMEM, "gft_mlp_train: device allocation failed");
    }
    if (hipMemcpyAsync(X, x, (size_t)n * m->d.n_in * sizeof(float), hipMemcpyHostToDevice, m->st) ||
        hipMemcpyAsync(T, t, (size_t)n * m->d.n_out * sizeof(float), hipMemcpyHostToDevice, m->st) ||
        hipMemcpyAsync(Wt, w, (size_t)n * m->d.n_heads * sizeof(float), hipMemcpyHostToDevice, m->st)) {
        cleanup();
        return mlp_fail(m, GFT_EHIP, "gft_mlp_train: upload failed");
    }
    std::mt19937 rng(seed);
    std::vector<int> order(n);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; e++) {
        for (int i = 0; i < n; i++) order[i] = i;
        std::shuffle(order.begin(), order.end(), rng);
        if (hipMemcpyAsync(perm, order.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, m->st) ||
            hipStreamSynchronize(m->st)) {
            cleanup();
            return mlp_fail(m, GFT_EHIP, "gft_mlp_train: shuffle upload failed");
        }
        for (int b = 0; b < nb; b++) {
            gft::mlp_gather<<<batch, 64, 0, m->st>>>(X, T, Wt, perm + (long long)b * batch, batch, m->d.n_in, m->in_pad,
                                                      m->d.n_out, m->d.n_heads, m->xb, m->tb, m->wb);
            mlp_forward_dev(m, batch);
            mlp_backward_dev(m, batch);
            (void)hipMemcpyAsync(hist + (long long)e * nb + b, m->loss_dev, sizeof(float), hipMemcpyDeviceToDevice, m->st);
            m->adam_t++;
            const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
            gft::mlp_adam<<<(unsigned)((m->P + 255) / 256), 256, 0, m->st>>>(m->params, m->grads, m->adam_m, m->adam_v,
                                                                              m->P, lr, b1, b2, eps, c1, c2,
                                                                              m->bf_index, m->wbf);
        }
    }
    std::vector<float> h((size_t)std::max(1, epochs * nb));
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(h.data(), hist, h.size() * sizeof(float), hipMemcpyDeviceToHost, m->st) == hipSuccess &&
                    hipStreamSynchronize(m->st) == hipSuccess;
    cleanup();
    if (!ok) return mlp_fail(m, GFT_EHIP, "gft_mlp_train: kernel failure");
    if (epoch_loss)
        for (int e = 0; e < epochs; e++) {
            double s = 0;
            for (int b = 0; b < nb; b++) s += h[(size_t)e * nb + b];
            epoch_loss[e] = (float)(s / nb);
        }
    return GFT_OK;
}

}  // extern "C"
