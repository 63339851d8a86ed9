// gft_grf.hip -- Gaussian-random-field initial states generated on the GPU.
//
// Each real field component is white noise from a counter-based generator
// (Philox4x32-10: counter = (site, component, 0, 0), key = the 64-bit seed;
// Box-Muller on two 53-bit uniforms), filtered in Fourier space by
// sqrt(P(k)) = |k|^(slope/2) for 0 < |k| <= k_max (zero elsewhere), brought
// back to real space and normalised to zero mean and unit rms.  The 3-D
// transforms are 1-D mixed-radix (4, 2, 3) Stockham FFTs along each axis, one
// 256-thread block per line with the line in LDS, in FP64.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace gft {

// Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
// SC 2011): ten rounds of two 32x32 -> 64-bit multiplies with the Weyl key
// schedule.
GFT_HD inline void philox4x32_10(const unsigned ctr_in[4], const unsigned key_in[2], unsigned out[4]) {
    unsigned c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    unsigned k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const unsigned long long p0 = 0xD2511F53ull * c0, p1 = 0xCD9E8D57ull * c2;
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
        const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Standard normal of (site, component): u1, u2 in (0, 1) from 53 bits each.
__device__ inline double grf_normal(unsigned long long seed, unsigned site, unsigned comp) {
    const unsigned ctr[4] = {site, comp, 0u, 0u};
    const unsigned key[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
    unsigned o[4];
    philox4x32_10(ctr, key, o);
    const unsigned long long a = ((unsigned long long)o[1] << 32) | o[0];
    const unsigned long long b = ((unsigned long long)o[3] << 32) | o[2];
    const double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = ((double)(b >> 11) + 0.5) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

__global__ void __launch_bounds__(256) grf_noise(double2* __restrict__ w, long long n, unsigned long long seed,
                                                 unsigned comp) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) w[s] = make_double2(grf_normal(seed, (unsigned)s, comp), 0.0);
}

struct FftAxis {
    int n;                 // line length (2^a 3^b, <= 256)
    int nr;                // number of radix passes
    int radix[8];          // 4s, then 2s, then 3s
    long long es;          // element stride along the line
    long long inner;       // lines: base(l) = (l / inner) * outer + (l % inner) * inner_stride
    long long outer;
    long long inner_stride;
};

// In-place 1-D DFT of every line along one axis: X_k = sum_j x_j exp(sign 2 pi i jk / n).
__global__ void __launch_bounds__(256) grf_fft_lines(double2* __restrict__ w, FftAxis ax, int sign) {
    __shared__ double2 buf[2][256];
    const long long l = blockIdx.x;
    const long long base = (l / ax.inner) * ax.outer + (l % ax.inner) * ax.inner_stride;
    const int n = ax.n, t = threadIdx.x;
    if (t < n) buf[0][t] = w[base + t * ax.es];
    __syncthreads();
    int src = 0, Ns = 1;
    for (int p = 0; p < ax.nr; p++) {
        const int R = ax.radix[p], m = n / R;
        if (t < m) {
            const int j = t, jr = j % Ns;
            double2 v[4];
            for (int r = 0; r < R; r++) {
                const double2 x = buf[src][j + r * m];
                // twiddle exp(sign 2 pi i r jr / (Ns R))
                double sn, cs;
                sincospi(sign * 2.0 * (double)(r * jr) / (double)(Ns * R), &sn, &cs);
                v[r] = make_double2(x.x * cs - x.y * sn, x.x * sn + x.y * cs);
            }
            double2 o[4];
            if (R == 2) {
                o[0] = make_double2(v[0].x + v[1].x, v[0].y + v[1].y);
                o[1] = make_double2(v[0].x - v[1].x, v[0].y - v[1].y);
            } else if (R == 4) {
                const double2 a0 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
                const double2 a1 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
                const double2 b0 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
                const double2 b1 = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
                // multiply b1 by sign * i
                const double2 ib1 = make_double2(-sign * b1.y, sign * b1.x);
                o[0] = make_double2(a0.x + b0.x, a0.y + b0.y);
                o[1] = make_double2(a1.x + ib1.x, a1.y + ib1.y);
                o[2] = make_double2(a0.x - b0.x, a0.y - b0.y);
                o[3] = make_double2(a1.x - ib1.x, a1.y - ib1.y);
            } else {   // R == 3: w = exp(sign 2 pi i / 3) = -1/2 + sign i sqrt(3)/2
                const double h = 0.86602540378443864676 * sign;
                const double2 s12 = make_double2(v[1].x + v[2].x, v[1].y + v[2].y);
                const double2 d12 = make_double2(v[1].x - v[2].x, v[1].y - v[2].y);
                const double2 mid = make_double2(v[0].x - 0.5 * s12.x, v[0].y - 0.5 * s12.y);
                o[0] = make_double2(v[0].x + s12.x, v[0].y + s12.y);
                o[1] = make_double2(mid.x - h * d12.y, mid.y + h * d12.x);
                o[2] = make_double2(mid.x + h * d12.y, mid.y - h * d12.x);
            }
            const int d = (j / Ns) * Ns * R + jr;
            for (int r = 0; r < R; r++) buf[1 - src][d + r * Ns] = o[r];
        }
        __syncthreads();
        src = 1 - src;
        Ns *= R;
    }
    if (t < n) w[base + t * ax.es] = buf[src][t];
}

// Multiplies mode (x, y, z) by |k|^(slope/2) for 0 < |k| <= kmax, else 0.
__global__ void __launch_bounds__(256) grf_filter(double2* __restrict__ w, int nx, int ny, int nz, double dx,
                                                  double dy, double dz, double slope, double kmax) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (long long)nx * ny * nz) return;
    const int z = (int)(s % nz), y = (int)((s / nz) % ny), x = (int)(s / ((long long)nz * ny));
    auto freq = [](int i, int n, double d) {
        const int f = i < (n + 1) / 2 ? i : i - n;
        return 6.283185307179586477 * (double)f / (n * d);
    };
    const double kx = freq(x, nx, dx), ky = freq(y, ny, dy), kz = freq(z, nz, dz);
    const double k = sqrt(kx * kx + ky * ky + kz * kz);
    const double a = (k > 0.0 && k <= kmax) ? pow(k, 0.5 * slope) : 0.0;
    w[s] = make_double2(w[s].x * a, w[s].y * a);
}

// Real part into f, and per-block (sum f, sum f^2) partials.
__global__ void __launch_bounds__(256) grf_real_stats(const double2* __restrict__ w, double* __restrict__ f,
                                                      long long n, double* __restrict__ partial) {
    __shared__ double lds[4 * 2];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (s < n) {
        const double r = w[s].x;
        f[s] = r;
        v[0] = r;
        v[1] = r * r;
    }
    block_reduce_store<2, 0>(v, nullptr, lds, partial, blockIdx.x);
}

// f <- (f - mean) / rms from stats[0] = sum f, stats[1] = sum f^2 over n; then
// partials of sum f^2 of the result.
__global__ void __launch_bounds__(256) grf_normalise(double* __restrict__ f, long long n,
                                                     const double* __restrict__ stats, double* __restrict__ partial) {
    __shared__ double lds[4 * 1];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const double mean = stats[0] / (double)n;
    const double var = stats[1] / (double)n - mean * mean;
    const double inv = 1.0 / sqrt(var);
    double v[1] = {0.0};
    if (s < n) {
        const double r = (f[s] - mean) * inv;
        f[s] = r;
        v[0] = r * r;
    }
    block_reduce_store<1, 0>(v, nullptr, lds, partial, blockIdx.x);
}

// Host: the radix plan of a line length (2^a 3^b up to 256), or false.
inline bool grf_plan(int n, FftAxis& ax) {
    if (n < 1 || n > 256) return false;
    int m = n, k = 0;
    while (m % 4 == 0 && k < 8) { ax.radix[k++] = 4; m /= 4; }
    while (m % 2 == 0 && k < 8) { ax.radix[k++] = 2; m /= 2; }
    while (m % 3 == 0 && k < 8) { ax.radix[k++] = 3; m /= 3; }
    if (m != 1) return false;
    ax.n = n;
    ax.nr = k;
    return true;
}

}  // namespace gft
