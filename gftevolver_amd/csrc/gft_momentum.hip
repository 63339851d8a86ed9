// gft_momentum.hip -- the fast path: structure-of-arrays kick/drift form of the
// reference leapfrog, one fused kernel per step.
//
// State per site (SM; the 2HDM doubles phi and P):
//   phi  float4            scalar doublet at t, stored as its difference
//                          from a constant reference doublet   (ping-pong)
//   P    float4            phi(t) - phi(t-dt)                  (in place)
//   y[3] float             U(1) link phases at t, wrapped      (ping-pong)
//   q[3] float4            SU(2) link quaternions at t         (ping-pong)
//   E[3] float4            (R^Y, R^1, R^2, R^3) of the last update (in place)
// = 35 reals read and 35 written per site-update.  With P and E the reference's
// update (Core/Source/Model.cpp:345-372, WilsonLoops/Source/StandardModel.cpp:
// 456-514) becomes
//   kick:  P <- (1 - g dt/2) P + dt^2 (Lap phi - V'(phi))
//          E <- (1 - g dt/2) E + dt^2 (M + g_a^2 J)
//   drift: phi <- phi + P;  y <- y + asin(E^Y);  q <- N[(sqrt(1-|E|^2), E) q]
// which is the reference's two-slot recursion with El = vec(U(t) U(t-dt)^dag)
// and sin(y(t) - y(t-dt)) replaced by the stored E of the previous step
// (identical in exact arithmetic).  phi and the links ping-pong because
// neighbours read them; P and E are read and written only at their own site.
//
// Layout: each field component group is its own array of N elements, sites in
// the reference order s = (x*ny + y)*nz + z, so a wave of 64 threads along z
// issues fully coalesced 16-byte loads per group.  Blocks are (BZ x BY) tiles
// of one x-plane; consecutive logical blocks advance along x inside a (y,z)
// column so the x-neighbour planes just read by the previous blocks are still
// in L2, and logical blocks are dealt to the 8 XCDs in contiguous ranges
// (hardware sends block b to XCD b % 8) so each XCD's L2 sees a compact region.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gft_math.hpp"
#include "gft_reduce.hpp"

namespace gft {

template <class R> struct V4T;
template <> struct V4T<float> { using type = float4; };
template <> struct V4T<double> { using type = double4; };

// Device pointers of one state.
template <class R>
struct MomFields {
    using V4 = typename V4T<R>::type;
    V4* phi[2][2];   // [buffer][doublet]
    V4* P[2];        // [doublet]
    R* y[2][3];      // [buffer][direction]
    V4* q[2][3];     // [buffer][direction]
    V4* E[3];        // [direction]
    // Compensated FP32 storage only (null otherwise): the 4-bit residuals of P
    // and E of a site in one 64-bit word, nibble m of a 16-bit group = value m:
    //   lo.x bits 0..15: P of doublet 0, bits 16..31: E[0];
    //   lo.y bits 0..15: E[1],           bits 16..31: E[2];
    // and those of the second doublet's P (2HDM) in lo_P2.
    uint2* lo;
    unsigned short* lo_P2;
    // Higgs-family gauge field of the 2HDM (null unless gaugeNum is 5 or 7):
    R* fy[2][3];        // gaugeNum 5: family U(1) phases y', wrapped   (ping-pong)
    R* fe[3];           // gaugeNum 5: E' of the last update            (in place)
    V4* fq[2][3];       // gaugeNum 7: family SU(2) quaternions         (ping-pong)
    V4* fE[3];          // gaugeNum 7: (0, E'^1, E'^2, E'^3) of the last update (in place)
    // Compensated storage: the residuals of the family E.  gaugeNum 5: nibble a
    // of flo5 = E'[a]; gaugeNum 7: 16-bit group a of flo7 (x low, x high, y
    // low) holds the residuals of fE[a] in nibbles 1..3 (nibble m = value m).
    unsigned* flo5;
    uint2* flo7;
};

struct MomentumState {
    void* base = nullptr;
    size_t bytes = 0;
    bool f64 = false;
    bool cmp = false;        // FP32 storage with 4-bit residuals of P and E
    // Reference doublets: phi is stored as phi - ref (FP32 forms: the mean of the
    // imported phi rounded to FP32, or the vacuum (0, 0, v, 0) of a generated
    // state; FP64 form: zero), so its FP32 rounding is relative to the
    // fluctuation, not to the vacuum value.
    double ref[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    unsigned epoch = 0;      // counts state replacements (import, generation): kernel arguments
                             // that carry ref are rebuilt when it moves
    int cur = 0;             // buffer holding (phi, U) at the current time
    int prev_valid = 0;      // buffer !cur holds the state one step earlier
    MomFields<float> ff{};
    MomFields<double> fd{};
    int bz = 64, by = 4;     // tile shape of the step kernel
    int order = 3;           // block order (see decode_block)
    int min_waves = 4;       // occupancy target of the FP32 SM step kernel (0 = unconstrained, 3, 4)
    int stage = 2;           // FP32 step kernels: 0 global loads, 1 LDS plane tile, 2 + shared plaquettes
    int4* tiles = nullptr;   // device tile table of the current launch shape
    int tiles_key[3] = {0, 0, -1};
};

// ---- float special functions on the small ranges the evolution produces ----
// Taylor polynomials (coefficients 1/n!) after a quadrant reduction with a
// three-part pi/2; |error| < 1e-9 absolute for |x| <= 16.  Arguments below
// 0.78 < pi/4 -- every plaquette angle and the phases of smooth fields -- have
// quadrant 0, where the reduction returns x itself: they skip it (a branch per
// lane, taken by whole waves in practice) and get bitwise the same result.
__device__ __forceinline__ float sin_poly(float r, float r2) {
    float sp = fmaf(r2, 2.75573192239858907e-6f, -1.98412698412698413e-4f);
    sp = fmaf(r2, sp, 8.33333333333333333e-3f);
    sp = fmaf(r2, sp, -1.66666666666666667e-1f);
    return fmaf(r * r2, sp, r);
}
__device__ __forceinline__ float cos_poly(float r2) {
    float cp = fmaf(r2, -2.75573192239858907e-7f, 2.48015873015873016e-5f);
    cp = fmaf(r2, cp, -1.38888888888888889e-3f);
    cp = fmaf(r2, cp, 4.16666666666666667e-2f);
    cp = fmaf(r2, cp, -0.5f);
    return fmaf(r2, cp, 1.0f);
}
// Quadrant reduction: r = x - q pi/2 (three-part pi/2), k = q mod 4.
__device__ __forceinline__ float quadrant_reduce(float x, int* k) {
    const float q = rintf(x * 0.636619772367581343f);
    float r = fmaf(-q, 1.57079637050628662109375f, x);
    r = fmaf(-q, -4.371138828673793e-8f, r);
    r = fmaf(-q, -1.7151245100058819e-15f, r);
    *k = static_cast<int>(q) & 3;
    return r;
}
__device__ __forceinline__ void sincos_fast(float x, float* s, float* c) {
    if (fabsf(x) < 0.78f) {
        const float r2 = x * x;
        *s = sin_poly(x, r2);
        *c = cos_poly(r2);
        return;
    }
    int k;
    const float r = quadrant_reduce(x, &k);
    const float r2 = r * r;
    const float sr = sin_poly(r, r2), cr = cos_poly(r2);
    const float ss = (k & 1) ? cr : sr;
    const float cc = (k & 1) ? sr : cr;
    *s = (k & 2) ? -ss : ss;
    *c = ((k + 1) & 2) ? -cc : cc;
}

__device__ __forceinline__ float sin_fast(float x) {
    if (fabsf(x) < 0.78f) return sin_poly(x, x * x);
    int k;
    const float r = quadrant_reduce(x, &k);
    const float r2 = r * r;
    const float ss = (k & 1) ? cos_poly(r2) : sin_poly(r, r2);
    return (k & 2) ? -ss : ss;
}

// The same functions without the branch: the reduction always runs (for
// |x| < 0.78 it returns x itself and quadrant 0, so the results are bitwise
// those of sincos_fast / sin_fast), both polynomials are evaluated and the
// quadrant selects.  About a dozen more VALU instructions per call and no
// exec-mask branch, for code that waits on memory rather than on the VALU
// (the Higgs-family phases of the gaugeNum-5 kernels).
__device__ __forceinline__ void sincos_nb(float x, float* s, float* c) {
    int k;
    const float r = quadrant_reduce(x, &k);
    const float r2 = r * r;
    const float sr = sin_poly(r, r2), cr = cos_poly(r2);
    const float ss = (k & 1) ? cr : sr;
    const float cc = (k & 1) ? sr : cr;
    *s = (k & 2) ? -ss : ss;
    *c = ((k + 1) & 2) ? -cc : cc;
}
__device__ __forceinline__ float sin_nb(float x) {
    int k;
    const float r = quadrant_reduce(x, &k);
    const float r2 = r * r;
    const float sr = sin_poly(r, r2), cr = cos_poly(r2);
    const float ss = (k & 1) ? cr : sr;
    return (k & 2) ? -ss : ss;
}
__device__ __forceinline__ void sincos_nb(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ double sin_nb(double x) { return sin(x); }

// asin by its Maclaurin series for |x| <= 1/8 (terms to x^11, error < 1e-13
// relative), the library asinf beyond.
__device__ __forceinline__ float asin_small(float x) {
    if (fabsf(x) > 0.125f) return asinf(x);
    const float x2 = x * x;
    float p = fmaf(x2, 2.23721590909090909e-2f, 3.03819444444444444e-2f);
    p = fmaf(x2, p, 4.46428571428571429e-2f);
    p = fmaf(x2, p, 7.5e-2f);
    p = fmaf(x2, p, 1.66666666666666667e-1f);
    return fmaf(x * x2, p, x);
}

__device__ __forceinline__ void sincos_fast(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ double sin_fast(double x) { return sin(x); }
__device__ __forceinline__ double asin_small(double x) { return asin(x); }
// asin_small without the branch: the library asinf and the series are both
// evaluated and |x| selects (bitwise asin_small's result).  It pays only in
// the family-U(1) kernels (gaugeNum 5: 192^3 step 0.913 -> 0.730 ms); the
// SM kernels lose 1.4% with it, the other 2HDM kernels 0-0.6%.
__device__ __forceinline__ float asin_nb(float x) {
    const float big = asinf(x);
    const float x2 = x * x;
    float p = fmaf(x2, 2.23721590909090909e-2f, 3.03819444444444444e-2f);
    p = fmaf(x2, p, 4.46428571428571429e-2f);
    p = fmaf(x2, p, 7.5e-2f);
    p = fmaf(x2, p, 1.66666666666666667e-1f);
    return fabsf(x) > 0.125f ? big : fmaf(x * x2, p, x);
}
__device__ __forceinline__ double asin_nb(double x) { return asin(x); }
template <bool NB, class R>
__device__ __forceinline__ R asin_k(R x) {
    if constexpr (NB) return asin_nb(x);
    else return asin_small(x);
}

// The step's trigonometric functions: NB selects the branch-free forms.  For
// the U(1) phases they pay only in the family-SU(2) kernels (gaugeNum 7:
// 192^3 step 1.071 -> 1.023 ms); the SM kernels, which are bound by VALU
// issue, lose 4% with them, the SU(2) x U(1) and family-U(1) 2HDM kernels
// 1-4%.  The family-U(1) phases always take them (gaugeNum 5: 1.136 ->
// 0.913 ms at 192^3).  Either form gives bitwise the same fields.
template <bool NB, class R>
__device__ __forceinline__ void sincos_k(R x, R* s, R* c) {
    if constexpr (NB) sincos_nb(x, s, c);
    else sincos_fast(x, s, c);
}
template <bool NB, class R>
__device__ __forceinline__ R sin_k(R x) {
    if constexpr (NB) return sin_nb(x);
    else return sin_fast(x);
}
__device__ __forceinline__ float rsqrt_r(float x) { return __frsqrt_rn(x); }
__device__ __forceinline__ double rsqrt_r(double x) { return 1.0 / sqrt(x); }
// Correctly rounded sqrt of x in [2^-24, 1] (1 - |E|^2 for a float |E|^2 < 1):
// the hardware estimate, then the neighbour whose square brackets x -- the
// library's correction step, without its rescaling of inputs below 2^-96 and
// its special-value fixup, which no input in this range needs (0 and 1 also
// come out exact).  Bitwise sqrtf for every such input.
__device__ __forceinline__ float sqrt_unit(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u), sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float vm = __builtin_fmaf(-sm, s, x), vp = __builtin_fmaf(-sp, s, x);
    const float r = vm <= 0.f ? sm : s;
    return vp > 0.f ? sp : r;
}
__device__ __forceinline__ double sqrt_unit(double x) { return sqrt(x); }

// Streaming (non-temporal) access to own-site data that no neighbour reads.
typedef float f4n __attribute__((ext_vector_type(4)));
typedef double d4n __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt(const float4* p) {
    const f4n v = __builtin_nontemporal_load(reinterpret_cast<const f4n*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double4 ld_nt(const double4* p) {
    const d4n v = __builtin_nontemporal_load(reinterpret_cast<const d4n*>(p));
    return make_double4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_nt(const float4& v, float4* p) {
    const f4n n = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(n, reinterpret_cast<f4n*>(p));
}
__device__ __forceinline__ void st_nt(const double4& v, double4* p) {
    const d4n n = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(n, reinterpret_cast<d4n*>(p));
}
typedef unsigned u4n __attribute__((ext_vector_type(4)));
typedef unsigned u2n __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 ld_nt(const uint4* p) {
    const u4n v = __builtin_nontemporal_load(reinterpret_cast<const u4n*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 ld_nt(const uint2* p) {
    const u2n v = __builtin_nontemporal_load(reinterpret_cast<const u2n*>(p));
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(const uint4& v, uint4* p) {
    const u4n n = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(n, reinterpret_cast<u4n*>(p));
}
__device__ __forceinline__ void st_nt(const uint2& v, uint2* p) {
    const u2n n = {v.x, v.y};
    __builtin_nontemporal_store(n, reinterpret_cast<u2n*>(p));
}
__device__ __forceinline__ void st_nt(float v, float* p) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ unsigned ld_nt(const unsigned* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ld_nt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double ld_nt(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_nt(unsigned v, unsigned* p) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(double v, double* p) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ unsigned short ld_nt(const unsigned short* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_nt(unsigned short v, unsigned short* p) { __builtin_nontemporal_store(v, p); }

template <class R>
__device__ __forceinline__ R wrap_pi(R y) {
    const R twopi = R(6.283185307179586476925286766559);
    const R pi = R(3.1415926535897932384626433832795);
    if (y > pi) y -= twopi;
    if (y < -pi) y += twopi;
    return y;
}

// Element idx of an array addressed by a 32-bit byte offset from its base
// (scalar base + vector offset addressing).
template <class T>
__device__ __forceinline__ const T& at(const T* base, unsigned idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * static_cast<unsigned>(sizeof(T)));
}
template <class T>
__device__ __forceinline__ T* ptr_at(T* base, unsigned idx) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * static_cast<unsigned>(sizeof(T)));
}

template <class R>
struct Qt { R w, x, y, z; };   // quaternion (c0, c1, c2, c3)

template <class V4, class R>
__device__ __forceinline__ Qt<R> toq(const V4& v) { return Qt<R>{v.x, v.y, v.z, v.w}; }

// Quaternion products with a fixed operation order (explicit FMAs, contraction
// off): a product of the same operands rounds the same way in every kernel and
// at every call site, which the plaquette sharing below relies on.
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// a * b
template <class R>
__device__ __forceinline__ Qt<R> qmul(const Qt<R>& a, const Qt<R>& b) {
#pragma clang fp contract(off)
    Qt<R> r;
    r.w = fma_r(a.w, b.w, -fma_r(a.x, b.x, fma_r(a.y, b.y, a.z * b.z)));
    r.x = fma_r(a.w, b.x, fma_r(a.x, b.w, fma_r(-a.y, b.z, a.z * b.y)));
    r.y = fma_r(a.w, b.y, fma_r(a.y, b.w, fma_r(-a.z, b.x, a.x * b.z)));
    r.z = fma_r(a.w, b.z, fma_r(a.z, b.w, fma_r(-a.x, b.y, a.y * b.x)));
    return r;
}
template <class R>
__device__ __forceinline__ Qt<R> qconj(const Qt<R>& a) { return Qt<R>{a.w, -a.x, -a.y, -a.z}; }
// vector part of a * b
template <class R>
__device__ __forceinline__ void qvec(const Qt<R>& a, const Qt<R>& b, R* v) {
#pragma clang fp contract(off)
    v[0] = fma_r(a.w, b.x, fma_r(a.x, b.w, fma_r(-a.y, b.z, a.z * b.y)));
    v[1] = fma_r(a.w, b.y, fma_r(a.y, b.w, fma_r(-a.z, b.x, a.x * b.z)));
    v[2] = fma_r(a.w, b.z, fma_r(a.z, b.w, fma_r(-a.x, b.y, a.y * b.x)));
}
// vec(q v q^dag) for a unit quaternion q and a pure vector v (the product
// convention of qmul): with u = vec(q) and c = v x u, v + 2 (q.w c + c x u).
template <class R>
__device__ __forceinline__ void qrot(const Qt<R>& q, const R* v, R* o) {
#pragma clang fp contract(off)
    const R c0 = fma_r(v[1], q.z, -v[2] * q.y);
    const R c1 = fma_r(v[2], q.x, -v[0] * q.z);
    const R c2 = fma_r(v[0], q.y, -v[1] * q.x);
    const R d0 = fma_r(q.w, c0, fma_r(c1, q.z, -c2 * q.y));
    const R d1 = fma_r(q.w, c1, fma_r(c2, q.x, -c0 * q.z));
    const R d2 = fma_r(q.w, c2, fma_r(c0, q.y, -c1 * q.x));
    o[0] = fma_r(R(2), d0, v[0]);
    o[1] = fma_r(R(2), d1, v[1]);
    o[2] = fma_r(R(2), d2, v[2]);
}

// ---- compensated FP32 storage ------------------------------------------------
// A stored value is hi + lo: hi the FP32 value the stencil reads, lo a 4-bit
// residual kept at the owning site only, in units of hi's ulp:
//   lo = q ulp(hi) / 2^3,  ulp(hi) = 2^(e - 150) for hi's biased exponent e,
// with q an integer in [-8, 7]; |lo| <= ulp(hi) / 2 gives |q| <= 4, so the
// rounding of the residual to q (nearest even) is never clamped and carries no
// bias.  (Units of ulp/2^4 would need q = 8 for residuals near a tie; clamping
// those to 7 biases every update the same way and the Gauss-law violation then
// grows linearly.)  Updates are error-free transformations (TwoSum, exact
// products by FMA), so the only rounding per step is that of the residual, at
// most 2^-4 ulp(hi): the state keeps about 28 significant bits for half a byte
// per value -- below the random walk of the FP32 force arithmetic in the Gauss
// law -- while the kernel's arithmetic and neighbour traffic stay FP32.
__device__ __forceinline__ int res_exp(float hi) { return static_cast<int>((__float_as_uint(hi) >> 23) & 0xffu); }
// q of residual r next to hi (round to nearest even).  r is the exact error of
// the FP32 sum hi, so |r| <= ulp(hi) / 2 and |q| <= 4: no clamp is needed.
// The returned word carries q in its low nibble (two's complement), which is
// all res_pack4 keeps: adding 1.5 * 2^23 to x = r / (ulp(hi) / 2^3) rounds x to
// the nearest integer, ties to even (as rintf), into the last bits of the sum,
// whose bit patterns are consecutive integers around 0x4B400000.
__device__ __forceinline__ int res_q(float r, float hi) {
    return static_cast<int>(__float_as_uint(ldexpf(r, 153 - res_exp(hi)) + 12582912.0f));
}
__device__ __forceinline__ float res_from_q(int q, float hi) { return ldexpf(static_cast<float>(q), res_exp(hi) - 153); }
// Four q in 16 bits (nibble m = value m) and back; nibbles 4..7 of a 32-bit
// word are a second group of four (the P residuals of a doublet word).
__device__ __forceinline__ unsigned res_pack4(int q0, int q1, int q2, int q3) {
    return (static_cast<unsigned>(q0) & 0xfu) | ((static_cast<unsigned>(q1) & 0xfu) << 4) |
           ((static_cast<unsigned>(q2) & 0xfu) << 8) | ((static_cast<unsigned>(q3) & 0xfu) << 12);
}
__device__ __forceinline__ int res_nib(unsigned w, int m) { return static_cast<int>(w << (28 - 4 * m)) >> 28; }
__device__ __forceinline__ float res_val(unsigned w, int m, float hi) { return res_from_q(res_nib(w, m), hi); }
// The 16-bit group of the residual word of a site: g = 0 P (doublet 0), 1..3 E[g - 1].
__device__ __forceinline__ unsigned res_group(const uint2& w, int g) {
    return ((g < 2 ? w.x : w.y) >> (16 * (g & 1))) & 0xffffu;
}

// (h, l) <- keep (h + l) + delta; returns the new hi, *r the exact residual.
__device__ __forceinline__ float cmp_update(float h, float l, float keep, float delta, float* r) {
#pragma clang fp contract(off)
    const float p = keep * h;
    const float pe = __builtin_fmaf(keep, h, -p);
    const float t = __builtin_fmaf(keep, l, delta) + pe;
    const float sum = p + t;
    const float bb = sum - p;
    *r = (p - (sum - bb)) + (t - bb);
    return sum;
}
// The same for keep = 1 (no damping): keep h is h and its error 0, so this
// gives bitwise the values of cmp_update with keep = 1 in fewer operations.
__device__ __forceinline__ float cmp_update1(float h, float l, float delta, float* r) {
#pragma clang fp contract(off)
    const float t = l + delta;
    const float sum = h + t;
    const float bb = sum - h;
    *r = (h - (sum - bb)) + (t - bb);
    return sum;
}
// (h1 + l1) + (h2 + l2) with |l1|, |l2| small; returns the hi, *r the residual.
__device__ __forceinline__ float cmp_add(float h1, float l1, float h2, float l2, float* r) {
#pragma clang fp contract(off)
    const float sum = h1 + h2;
    const float bb = sum - h1;
    const float e = (h1 - (sum - bb)) + (h2 - bb);
    const float t = (l1 + l2) + e;
    const float hi = sum + t;
    *r = t - (hi - sum);
    return hi;
}

// SU(2) action (StandardModel.cpp:94-97) of quaternion u on doublet h, in a
// fixed operation order like qmul.
template <class R>
__device__ __forceinline__ void act(const Qt<R>& u, const R* h, R* o) {
#pragma clang fp contract(off)
    o[0] = fma_r(u.w, h[0], fma_r(-u.z, h[1], fma_r(u.y, h[2], -u.x * h[3])));
    o[1] = fma_r(u.z, h[0], fma_r(u.w, h[1], fma_r(u.x, h[2], u.y * h[3])));
    o[2] = fma_r(-u.y, h[0], fma_r(-u.x, h[1], fma_r(u.w, h[2], u.z * h[3])));
    o[3] = fma_r(u.x, h[0], fma_r(-u.y, h[1], fma_r(-u.z, h[2], u.w * h[3])));
}

// T phi = U^W (e^{iy} phi) for one doublet, with (c, s) = (cos y, sin y);
// conj gives U^W^dag (e^{-iy} phi).
template <class R>
__device__ __forceinline__ void transport(const Qt<R>& u, R c, R s, bool conj, const R* h, R* o) {
#pragma clang fp contract(off)
    const R ss = conj ? -s : s;
    R t[4];
    t[0] = fma_r(c, h[0], -ss * h[1]);
    t[1] = fma_r(c, h[1], ss * h[0]);
    t[2] = fma_r(c, h[2], -ss * h[3]);
    t[3] = fma_r(c, h[3], ss * h[2]);
    act(conj ? qconj(u) : u, t, o);
}

// The Higgs-family part of one link of the 2HDM: gaugeNum 5 a phase y' given
// by (c, s) = (cos y', sin y'), gaugeNum 7 an SU(2) quaternion v.
template <class R>
struct FamLink {
    R c, s;
    Qt<R> v;
};

// T phi for both doublets (8 reals): the U(1) phase and SU(2) of transport on
// each doublet, then the Higgs-family part (Gradients/Source/TwoHDM.cpp:
// 106-141): gaugeNum 5 the phase y' with charge +1 on the first doublet and -1
// on the second, gaugeNum 7 the SU(2) v acting on the doublet index of
// (f_i, f_i+1, f_i+4, f_i+5), i = 0, 2 (the real 4x4 action of act).  conj
// gives T^dag.  The three factors commute.
template <class R, int G>
__device__ __forceinline__ void transport2(const Qt<R>& u, R c, R s, const FamLink<R>& f, bool conj, const R* h,
                                           R* o) {
#pragma clang fp contract(off)
    R t[8];
    transport(u, c, s, conj, h, t);
    transport(u, c, s, conj, h + 4, t + 4);
    if constexpr (G == 5) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const R ss = (conj != (k == 1)) ? -f.s : f.s;
            const R* p = t + 4 * k;
            R* q = o + 4 * k;
            q[0] = fma_r(f.c, p[0], -ss * p[1]);
            q[1] = fma_r(f.c, p[1], ss * p[0]);
            q[2] = fma_r(f.c, p[2], -ss * p[3]);
            q[3] = fma_r(f.c, p[3], ss * p[2]);
        }
    } else {
        const Qt<R> v = conj ? qconj(f.v) : f.v;
#pragma unroll
        for (int i = 0; i < 4; i += 2) {
            const R hi[4] = {t[i], t[i + 1], t[i + 4], t[i + 5]};
            R ho[4];
            act(v, hi, ho);
            o[i] = ho[0];
            o[i + 1] = ho[1];
            o[i + 4] = ho[2];
            o[i + 5] = ho[3];
        }
    }
}

// Higgs-family bilinears of both doublets' t and n (8 reals each): the family
// currents and charges (Gradients/Source/TwoHDM.cpp:401-424, 299-319),
// accumulated into j[0] (gaugeNum 5) or j[0..2] (gaugeNum 7).  Antisymmetric:
// they vanish for n = t, like the bilinears of bilinears_add.
template <class R, int G>
__device__ __forceinline__ void family_bilinears(const R* t, const R* n, R* j) {
#pragma unroll
    for (int i = 0; i < 4; i += 2) {
        if constexpr (G == 5) {
            j[0] += R(0.25) * (t[i] * n[i + 1] - t[i + 1] * n[i] - t[i + 4] * n[i + 5] + t[i + 5] * n[i + 4]);
        } else {
            j[0] += R(0.25) * (t[i] * n[i + 5] - t[i + 1] * n[i + 4] + t[i + 4] * n[i + 1] - t[i + 5] * n[i]);
            j[1] += R(0.25) * (-t[i] * n[i + 4] - t[i + 1] * n[i + 5] + t[i + 4] * n[i] + t[i + 5] * n[i + 1]);
            j[2] += R(0.25) * (t[i] * n[i + 1] - t[i + 1] * n[i] - t[i + 4] * n[i + 5] + t[i + 5] * n[i + 4]);
        }
    }
}

template <class V4, class R>
__device__ __forceinline__ void unpack(const V4& v, R* o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }
template <class V4, class R>
__device__ __forceinline__ V4 pack(const R* i) {
    V4 v;
    v.x = i[0];
    v.y = i[1];
    v.z = i[2];
    v.w = i[3];
    return v;
}

template <class R>
struct StepArgs {
    using V4 = typename V4T<R>::type;
    const V4* phi[2];
    V4* phi_out[2];
    V4* P[2];
    const R* y[3];
    R* y_out[3];
    const V4* q[3];
    V4* q_out[3];
    V4* E[3];
    uint2* lo;               // compensated storage: residuals of P and E (see MomFields)
    unsigned short* lo_P2;
    // Higgs-family gauge field (gaugeNum 5 or 7; see MomFields).
    const R* fy[3];
    R* fy_out[3];
    R* fe[3];
    const V4* fq[3];
    V4* fq_out[3];
    V4* fE[3];
    unsigned* flo5;
    uint2* flo7;
    R gsq_f;                 // g''^2 (the family generators' coupling, rounded through float)
    V4 ref[2];               // reference doublets: phi = stored + ref
    int nx, ny, nz, nbz, nby;
    long long nblocks;
    int bz, by, order, bz_shift;   // bz = 1 << bz_shift
    const int4* tiles;       // per hardware block: (member, x, y-tile, z-tile), built on the host
    R inv_sp[3];
    R dt2, keep;             // dt^2 and 1 - gamma dt / 2
    R gsq[4];
    unsigned* err;
};

// Block orders (the hardware sends block b to XCD b % 8):
//   0  hardware order, x fastest inside (y, z) column tiles;
//   1  as 0, with logical blocks dealt to the XCDs in contiguous ranges;
//   2  contiguous ranges per XCD over y-bands: the y-tiles are cut into 8
//      bands, and inside a band the order is x (slowest), then y-tile, then
//      z-tile.  Each XCD then sweeps x-planes of its own band, so the
//      neighbour planes x-1 and x+1 of a band (tens of rows) stay in its L2;
//   3, 4  as 2 with 16 or 32 bands (two or four per XCD, swept one after the
//      other): a plane of a band holds fewer tiles, so the blocks an XCD runs
//      at once span more x-planes of less data each and the x-1 plane is
//      re-read sooner after it was loaded; the rings of the band edges grow.
GFT_HD long long xcd_contiguous(long long b, long long nb) {
    const long long per = nb / 8, rem = nb % 8, xcd = b % 8, idx = b / 8;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Decodes the logical tile (x, y-tile, z-tile) of hardware block b.
GFT_HD void decode_block(long long b, long long nb, int order, int nx, int nby, int nbz, int* x, int* ty, int* tz) {
    const long long L = order == 0 ? b : xcd_contiguous(b, nb);
    if (order <= 1) {
        *x = static_cast<int>(L % nx);
        const long long rest = L / nx;
        *ty = static_cast<int>(rest % nby);
        *tz = static_cast<int>(rest / nby);
        return;
    }
    const int nbands_req = 8 << (order - 2);
    const int br = (nby + nbands_req - 1) / nbands_req;   // y-tiles per band
    const int nbands = (nby + br - 1) / br;
    const long long per_band = (long long)br * nx * nbz;
    int band = static_cast<int>(L / per_band);
    if (band > nbands - 1) band = nbands - 1;
    const long long loc = L - band * per_band;
    const int rows = (nby - band * br) < br ? (nby - band * br) : br;   // the last band may be shorter
    *tz = static_cast<int>(loc % nbz);
    const long long t = loc / nbz;
    *ty = band * br + static_cast<int>(t % rows);
    *x = static_cast<int>(t / rows);
}

// In-plane staging of a block's tile in LDS: q[3], y[3] and phi[ND] of the
// bz x by sites of plane x plus a one-site ring in y and z, one array per field,
// slot (ty + 1) * w + (tz + 1) for tile coordinates (ty, tz), w = bz + 2.  Every
// neighbour of a site inside its own x-plane (x +- e_y, x +- e_z, x + e_y - e_z,
// x + e_z - e_y) is then an LDS read; neighbours in the planes x +- 1 stay global.
struct PlaneLds {
    float4* q[3];
    float* y[3];
    float4* phi[2];
    int w;        // slots per tile row (bz + 2)
    int r0;       // this thread's slot
};

// Slots per field, rounded up to whole 64-slot chunks (one wave fills one chunk
// of one field per instruction).
__host__ __device__ inline int plane_slots(int bz, int by) { return ((bz + 2) * (by + 2) + 63) / 64 * 64; }
__host__ __device__ inline size_t plane_lds_bytes(int bz, int by, int nd) {
    return (size_t)plane_slots(bz, by) * (3 * 16 + 3 * 4 + nd * 16);
}

template <int ND>
__device__ __forceinline__ PlaneLds plane_layout(char* base, int bz, int by) {
    const int nsp = plane_slots(bz, by);
    PlaneLds p;
    float4* q = reinterpret_cast<float4*>(base);
#pragma unroll
    for (int f = 0; f < 3; f++) p.q[f] = q + f * nsp;
#pragma unroll
    for (int k = 0; k < 2; k++) p.phi[k] = q + (3 + (k < ND ? k : 0)) * nsp;
    float* y = reinterpret_cast<float*>(q + (3 + ND) * nsp);
#pragma unroll
    for (int f = 0; f < 3; f++) p.y[f] = y + f * nsp;
    p.w = bz + 2;
    p.r0 = 0;
    return p;
}

// v mod n for v >= -n; `general` false promises v in [-1, 2n).
__device__ __forceinline__ int wrap_index(int v, int n, bool general) {
    if (general) return (v % n + n) % n;
    v = v < 0 ? v + n : v;
    return v >= n ? v - n : v;
}

// Whole block: gathers the tile and ring of plane x (tile origin y0, z0) straight
// into LDS with global_load_lds (no register round trip), then waits for its own
// transfers and meets the other waves at the barrier.

template <int ND>
__device__ __forceinline__ void plane_fill(const StepArgs<float>& a, const PlaneLds& p, int x, int y0, int z0, int bz,
                                           int by) {
    const int w = bz + 2, ns = w * (by + 2);
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int sx = a.ny * a.nz;
    const float rcp_w = __builtin_amdgcn_rcpf(static_cast<float>(w));
    // Ring indices lie in [-1, n - 1 + b]: one conditional add or subtract wraps
    // them unless the tile is longer than the lattice side (b > n).
    const bool gen_y = by > a.ny, gen_z = bz > a.nz;
    for (int c = wave; c * 64 < ns; c += 4) {
        const int r = min(c * 64 + lane, ns - 1);
        // r / w from the float reciprocal ((r + 1/2) / w is at least 1/(2w) from an
        // integer, far above the rounding error), with a one-step correction.
        int ty = static_cast<int>((static_cast<float>(r) + 0.5f) * rcp_w), tz = r - ty * w;
        if (tz < 0) { ty--; tz += w; }
        if (tz >= w) { ty++; tz -= w; }
        const int yy = wrap_index(y0 + ty - 1, a.ny, gen_y);
        const int zz = wrap_index(z0 + tz - 1, a.nz, gen_z);
        const unsigned s = static_cast<unsigned>(x * sx + yy * a.nz + zz);
#pragma unroll
        for (int f = 0; f < 3; f++)
            __builtin_amdgcn_global_load_lds(&at(a.q[f], s), (__attribute__((address_space(3))) void*)(p.q[f] + c * 64),
                                             16, 0, 0);
#pragma unroll
        for (int f = 0; f < 3; f++)
            __builtin_amdgcn_global_load_lds(&at(a.y[f], s), (__attribute__((address_space(3))) void*)(p.y[f] + c * 64),
                                             4, 0, 0);
#pragma unroll
        for (int k = 0; k < ND; k++)
            __builtin_amdgcn_global_load_lds(&at(a.phi[k], s),
                                             (__attribute__((address_space(3))) void*)(p.phi[k] + c * 64), 16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Kick and drift of site s from its forces: F on the doublet(s), M + g^2 J on
// the links (shared by every step kernel variant).
// The site's own momenta and electric fields (with residuals), read once by
// the kick; loadable ahead of it.
template <class R, int ND>
struct OwnState {
    typename V4T<R>::type P[ND];
    typename V4T<R>::type E[3];
    uint2 lw;              // residual word: P of doublet 0 and E
    unsigned short lp2;    // residuals of the second doublet's P
};
template <class R, int ND, bool CMP>
__device__ __forceinline__ void load_own(const StepArgs<R>& a, unsigned s, OwnState<R, ND>& o) {
#pragma unroll
    for (int k = 0; k < ND; k++) o.P[k] = ld_nt(ptr_at(a.P[k], s));
#pragma unroll
    for (int i = 0; i < 3; i++) o.E[i] = ld_nt(ptr_at(a.E[i], s));
    if constexpr (CMP) {
        o.lw = ld_nt(ptr_at(a.lo, s));
        if constexpr (ND == 2) o.lp2 = ld_nt(ptr_at(a.lo_P2, s));
    }
}

// Full doublet = stored value + reference (one rounding per component, the same
// expression wherever a doublet is read, so every staging mode agrees bitwise).
template <class V4, class R>
__device__ __forceinline__ void add_ref(const V4& ref, const R* raw, R* o) {
    o[0] = raw[0] + ref.x;
    o[1] = raw[1] + ref.y;
    o[2] = raw[2] + ref.z;
    o[3] = raw[3] + ref.w;
}

// The gauge source enters the link force as M_i + g_a^2 J_i.  Every step
// kernel adds it at the same point of the magnetic sum -- after the plaquettes
// based at x and at x - e_0, before those based at x - e_1 and x - e_2 -- so the
// currents need no registers past that point and all staging modes round alike.
template <class R>
__device__ __forceinline__ void fold_current(const StepArgs<R>& a, const R (*J)[4], R (*M)[4]) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int m = 0; m < 4; m++) M[i][m] = fma(a.gsq[m], J[i][m], M[i][m]);
}
// PRE: own state already loaded into o (else loaded here, next to its use).
// U1: no damping (keep = 1), with the kick's keep products dropped (exact).
// praw: the site's stored doublet(s) (phi - ref), which the drift advances.
// M: the link forces with the gauge source already folded in (fold_current).
// NBA: the link update's asin in its branch-free form (asin_nb).
template <class R, int ND, bool CMP, bool PRE = false, bool U1 = false, bool FZ = true, bool NBA = false>
// frozen: a site of a fixed face's shell -- every lane still stores, but a
// frozen lane stores the values it read, so every store instruction writes
// whole cache lines (a divergent copy branch would write partial lines twice).
// FZ false (a lattice without fixed faces): no lane is frozen, and the
// selects between computed and read values are not generated.
__device__ __forceinline__ void kick_drift_t(const StepArgs<R>& a, unsigned s, const Qt<R>* U, const R* yv,
                                             const R* praw, const R* F, const R (*M)[4], const OwnState<R, ND>* o,
                                             bool frozen) {
    using V4 = typename V4T<R>::type;
    if constexpr (!FZ) frozen = false;
    uint2 lw = make_uint2(0u, 0u);
    unsigned lp2 = 0u;
    if constexpr (CMP) {
        if constexpr (PRE) {
            lw = o->lw;
            if constexpr (ND == 2) lp2 = o->lp2;
        } else {
            lw = ld_nt(ptr_at(a.lo, s));
            if constexpr (ND == 2) lp2 = ld_nt(ptr_at(a.lo_P2, s));
        }
    }
    unsigned nwx = 0u, nwy = 0u, nlp2 = 0u;   // the new residual groups
#pragma unroll
    for (int k = 0; k < ND; k++) {
        R p[4];
        if constexpr (PRE) unpack(o->P[k], p);
        else unpack(ld_nt(ptr_at(a.P[k], s)), p);
        V4 pn, fn;
        if constexpr (CMP) {
            const unsigned L = k == 0 ? res_group(lw, 0) : lp2;
            float ph[4], pr[4];
            int qp[4];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                if constexpr (U1) ph[m] = cmp_update1(p[m], res_val(L, m, p[m]), a.dt2 * F[4 * k + m], &pr[m]);
                else ph[m] = cmp_update(p[m], res_val(L, m, p[m]), a.keep, a.dt2 * F[4 * k + m], &pr[m]);
                qp[m] = res_q(pr[m], ph[m]);
            }
            pn = make_float4(ph[0], ph[1], ph[2], ph[3]);
            const unsigned Ln = frozen ? L : res_pack4(qp[0], qp[1], qp[2], qp[3]);
            if (k == 0) nwx = Ln;
            else nlp2 = Ln;
        } else {
            // Explicit FMAs fix which product is fused (exact keep * p for keep = 1).
            const R kp[4] = {U1 ? p[0] : a.keep * p[0], U1 ? p[1] : a.keep * p[1], U1 ? p[2] : a.keep * p[2],
                             U1 ? p[3] : a.keep * p[3]};
            pn.x = fma(a.dt2, F[4 * k + 0], kp[0]);
            pn.y = fma(a.dt2, F[4 * k + 1], kp[1]);
            pn.z = fma(a.dt2, F[4 * k + 2], kp[2]);
            pn.w = fma(a.dt2, F[4 * k + 3], kp[3]);
        }
        // Drift of the stored doublet: (phi - ref)(t+dt) = (phi - ref)(t) + P(t+dt).
        // The rounding is relative to the fluctuation phi - ref; P's residual lies
        // far below its ulp.
        fn.x = praw[4 * k + 0] + pn.x;
        fn.y = praw[4 * k + 1] + pn.y;
        fn.z = praw[4 * k + 2] + pn.z;
        fn.w = praw[4 * k + 3] + pn.w;
        if (frozen) {
            pn = pack<V4, R>(p);
            fn = pack<V4, R>(praw + 4 * k);
        }
        st_nt(pn, ptr_at(a.P[k], s));
        st_nt(fn, ptr_at(a.phi_out[k], s));
    }
    unsigned bad = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        R e[4];
        if constexpr (PRE) unpack(o->E[i], e);
        else unpack(ld_nt(ptr_at(a.E[i], s)), e);
        V4 en;
        if constexpr (CMP) {
            const unsigned L = res_group(lw, 1 + i);
            float eh[4], er[4];
#pragma unroll
            for (int m = 0; m < 4; m++)
                if constexpr (U1) eh[m] = cmp_update1(e[m], res_val(L, m, e[m]), a.dt2 * M[i][m], &er[m]);
                else eh[m] = cmp_update(e[m], res_val(L, m, e[m]), a.keep, a.dt2 * M[i][m], &er[m]);
            en = make_float4(eh[0], eh[1], eh[2], eh[3]);
            const unsigned Ln =
                frozen ? L : res_pack4(res_q(er[0], eh[0]), res_q(er[1], eh[1]), res_q(er[2], eh[2]), res_q(er[3], eh[3]));
            if (i == 0) nwx |= Ln << 16;
            else if (i == 1) nwy = Ln;
            else nwy |= Ln << 16;
        } else {
            en.x = fma(a.dt2, M[i][0], U1 ? e[0] : a.keep * e[0]);
            en.y = fma(a.dt2, M[i][1], U1 ? e[1] : a.keep * e[1]);
            en.z = fma(a.dt2, M[i][2], U1 ? e[2] : a.keep * e[2]);
            en.w = fma(a.dt2, M[i][3], U1 ? e[3] : a.keep * e[3]);
        }
        const R e2 = fma(en.w, en.w, fma(en.z, en.z, en.y * en.y));
        if (!frozen && (!(e2 < R(1)) || !(fabs(en.x) < R(1)))) bad = 1;
        const R yn = wrap_pi(yv[i] + asin_k<NBA>(en.x));
        const Qt<R> u{sqrt_unit(R(1) - e2), en.y, en.z, en.w};
        Qt<R> qn = qmul(u, U[i]);
        const R inv = rsqrt_r(fma(qn.z, qn.z, fma(qn.y, qn.y, fma(qn.x, qn.x, qn.w * qn.w))));
        V4 qo;
        qo.x = qn.w * inv;
        qo.y = qn.x * inv;
        qo.z = qn.y * inv;
        qo.w = qn.z * inv;
        if (frozen) {
            en = pack<V4, R>(e);
            qo.x = U[i].w;
            qo.y = U[i].x;
            qo.z = U[i].y;
            qo.w = U[i].z;
        }
        st_nt(en, ptr_at(a.E[i], s));
        st_nt(frozen ? yv[i] : yn, ptr_at(a.y_out[i], s));
        st_nt(qo, ptr_at(a.q_out[i], s));
    }
    if constexpr (CMP) {
        st_nt(make_uint2(nwx, nwy), ptr_at(a.lo, s));
        if constexpr (ND == 2) st_nt(static_cast<unsigned short>(nlp2), ptr_at(a.lo_P2, s));
    }
    if (bad) atomicOr(a.err, 1u);
}

// The kick with uniform branches on damping (keep = 1 takes the shorter path)
// and on fixed faces (a periodic lattice has no frozen lane).
template <class R, int ND, bool CMP, bool PRE = false, bool NBA = false>
__device__ __forceinline__ void kick_drift(const StepArgs<R>& a, unsigned s, const Qt<R>* U, const R* yv,
                                           const R* praw, const R* F, const R (*M)[4], const OwnState<R, ND>* o,
                                           bool frozen, bool any_fixed) {
    if (!any_fixed) {
        if (a.keep == R(1)) kick_drift_t<R, ND, CMP, PRE, true, false, NBA>(a, s, U, yv, praw, F, M, o, false);
        else kick_drift_t<R, ND, CMP, PRE, false, false, NBA>(a, s, U, yv, praw, F, M, o, false);
    } else {
        if (a.keep == R(1)) kick_drift_t<R, ND, CMP, PRE, true, true, NBA>(a, s, U, yv, praw, F, M, o, frozen);
        else kick_drift_t<R, ND, CMP, PRE, false, true, NBA>(a, s, U, yv, praw, F, M, o, frozen);
    }
}

// Kick and drift of the Higgs-family links of site s (gaugeNum 5: phases y'
// with E' = sin(y'(t+dt) - y'(t)); gaugeNum 7: quaternions with
// E' = vec(V(t+dt) V(t)^dag)), from their forces MF (staples + g''^2 J'), as the
// SU(2) x U(1) links in kick_drift_t (WilsonLoops/Source/TwoHDM.cpp:645-702).
template <class R, int G, bool CMP, bool U1, bool FZ>
__device__ __forceinline__ void kick_drift_family_t(const StepArgs<R>& a, unsigned s, const R* fyv, const Qt<R>* FV,
                                                    const R (*MF)[4], bool frozen) {
    using V4 = typename V4T<R>::type;
    if constexpr (!FZ) frozen = false;
    unsigned bad = 0;
    if constexpr (G == 5) {
        unsigned lw = 0u, nw = 0u;
        if constexpr (CMP) lw = ld_nt(ptr_at(a.flo5, s));
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const R e = ld_nt(ptr_at(a.fe[i], s));
            R en;
            if constexpr (CMP) {
                float r;
                if constexpr (U1) en = cmp_update1(e, res_val(lw, i, e), a.dt2 * MF[i][0], &r);
                else en = cmp_update(e, res_val(lw, i, e), a.keep, a.dt2 * MF[i][0], &r);
                nw |= (static_cast<unsigned>(res_q(r, en)) & 0xfu) << (4 * i);
            } else {
                en = fma(a.dt2, MF[i][0], U1 ? e : a.keep * e);
            }
            if (!frozen && !(fabs(en) < R(1))) bad = 1;
            const R yn = wrap_pi(fyv[i] + asin_k<true>(en));
            st_nt(frozen ? e : en, ptr_at(a.fe[i], s));
            st_nt(frozen ? fyv[i] : yn, ptr_at(a.fy_out[i], s));
        }
        if constexpr (CMP) st_nt(frozen ? lw : nw, ptr_at(a.flo5, s));
    } else {
        uint2 lw = make_uint2(0u, 0u);
        unsigned ng[3] = {0u, 0u, 0u};
        if constexpr (CMP) lw = ld_nt(ptr_at(a.flo7, s));
#pragma unroll
        for (int i = 0; i < 3; i++) {
            R e[4];
            unpack(ld_nt(ptr_at(a.fE[i], s)), e);
            V4 en;
            en.x = R(0);
            if constexpr (CMP) {
                const unsigned L = res_group(lw, i);
                float eh[4], er[4];
                eh[0] = er[0] = 0.f;
#pragma unroll
                for (int m = 1; m < 4; m++) {
                    if constexpr (U1) eh[m] = cmp_update1(e[m], res_val(L, m, e[m]), a.dt2 * MF[i][m], &er[m]);
                    else eh[m] = cmp_update(e[m], res_val(L, m, e[m]), a.keep, a.dt2 * MF[i][m], &er[m]);
                }
                en.y = eh[1];
                en.z = eh[2];
                en.w = eh[3];
                ng[i] = frozen ? L : res_pack4(0, res_q(er[1], eh[1]), res_q(er[2], eh[2]), res_q(er[3], eh[3]));
            } else {
                en.y = fma(a.dt2, MF[i][1], U1 ? e[1] : a.keep * e[1]);
                en.z = fma(a.dt2, MF[i][2], U1 ? e[2] : a.keep * e[2]);
                en.w = fma(a.dt2, MF[i][3], U1 ? e[3] : a.keep * e[3]);
            }
            const R e2 = fma(en.w, en.w, fma(en.z, en.z, en.y * en.y));
            if (!frozen && !(e2 < R(1))) bad = 1;
            const Qt<R> u{sqrt_unit(R(1) - e2), en.y, en.z, en.w};
            const Qt<R> qn = qmul(u, FV[i]);
            const R inv = rsqrt_r(fma(qn.z, qn.z, fma(qn.y, qn.y, fma(qn.x, qn.x, qn.w * qn.w))));
            V4 qo;
            qo.x = qn.w * inv;
            qo.y = qn.x * inv;
            qo.z = qn.y * inv;
            qo.w = qn.z * inv;
            if (frozen) {
                en = pack<V4, R>(e);
                qo.x = FV[i].w;
                qo.y = FV[i].x;
                qo.z = FV[i].y;
                qo.w = FV[i].z;
            }
            st_nt(en, ptr_at(a.fE[i], s));
            st_nt(qo, ptr_at(a.fq_out[i], s));
        }
        if constexpr (CMP) st_nt(make_uint2(ng[0] | (ng[1] << 16), ng[2]), ptr_at(a.flo7, s));
    }
    if (bad) atomicOr(a.err, 1u);
}
template <class R, int G, bool CMP>
__device__ __forceinline__ void kick_drift_family(const StepArgs<R>& a, unsigned s, const R* fyv, const Qt<R>* FV,
                                                  const R (*MF)[4], bool frozen, bool any_fixed) {
    if (!any_fixed) {
        if (a.keep == R(1)) kick_drift_family_t<R, G, CMP, true, false>(a, s, fyv, FV, MF, false);
        else kick_drift_family_t<R, G, CMP, false, false>(a, s, fyv, FV, MF, false);
    } else {
        if (a.keep == R(1)) kick_drift_family_t<R, G, CMP, true, true>(a, s, fyv, FV, MF, frozen);
        else kick_drift_family_t<R, G, CMP, false, true>(a, s, fyv, FV, MF, frozen);
    }
}

// Higgs-family link forces (WilsonLoops/Source/TwoHDM.cpp:496-552): the
// staples in the pattern of the U(1) (G = 5) or SU(2) (G = 7) staples of
// mom_site, each plaquette's contributions from one rounded angle or
// quaternion, then the family source g''^2 J'.  FY(f, i, si, j, sj) / FQ(...)
// read the family phase / quaternion of direction f at x + si e_i + sj e_j;
// fyv, FV are the site's own.  MF[i][0] (G = 5) or MF[i][1..3] (G = 7).
template <class R, int G, class FYF, class FQF>
__device__ __forceinline__ void family_force(const StepArgs<R>& a, FYF FY, FQF FQ, const R* fyv, const Qt<R>* FV,
                                             const R (*JF)[3], R (*MF)[4]) {
#pragma unroll
    for (int i = 0; i < 3; i++) MF[i][0] = MF[i][1] = MF[i][2] = MF[i][3] = R(0);
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            if constexpr (G == 5) {
                const R th = (FY(i, j, 1, 0, 0) - fyv[i]) + (fyv[j] - FY(j, i, 1, 0, 0));
                const R st = sin_nb(th);
                MF[i][0] += a.inv_sp[j] * st;
                MF[j][0] -= a.inv_sp[i] * st;
            } else {
                const Qt<R> A = qmul(qconj(FQ(j, i, 1, 0, 0)), FQ(i, j, 1, 0, 0));
                const Qt<R> C = qmul(FV[j], qconj(FV[i]));
                R v[3], w[3];
                qvec(A, C, v);
                qrot(C, v, w);
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    MF[i][1 + m] += a.inv_sp[j] * v[m];
                    MF[j][1 + m] -= a.inv_sp[i] * w[m];
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i == j) continue;
            if constexpr (G == 5) {
                const R yim = FY(i, j, -1, 0, 0), yjm = FY(j, j, -1, 0, 0), yjd = FY(j, i, 1, j, -1);
                if (i < j) MF[i][0] -= a.inv_sp[j] * sin_nb((fyv[i] - yim) + (yjm - yjd));
                else MF[i][0] += a.inv_sp[j] * sin_nb((yjd - yjm) + (yim - fyv[i]));
            } else {
                const Qt<R> Vim = FQ(i, j, -1, 0, 0), Vjm = FQ(j, j, -1, 0, 0), Vjd = FQ(j, i, 1, j, -1);
                R v[3], w[3];
                if (i < j) {
                    qvec(qmul(qconj(Vjd), FV[i]), qmul(Vjm, qconj(Vim)), v);
                    qrot(Vjd, v, w);
#pragma unroll
                    for (int m = 0; m < 3; m++) MF[i][1 + m] -= a.inv_sp[j] * w[m];
                } else {
                    qvec(qmul(qconj(FV[i]), Vjd), qmul(Vim, qconj(Vjm)), v);
                    qrot(FV[i], v, w);
#pragma unroll
                    for (int m = 0; m < 3; m++) MF[i][1 + m] += a.inv_sp[j] * w[m];
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if constexpr (G == 5) MF[i][0] = fma(a.gsq_f, JF[i][0], MF[i][0]);
        else
#pragma unroll
            for (int m = 0; m < 3; m++) MF[i][1 + m] = fma(a.gsq_f, JF[i][m], MF[i][1 + m]);
    }
}

// One site-update at (x, y, z): forces, kick and drift (shared by the
// single-lattice kernel and the ensemble kernel).  PL: in-plane neighbours and
// own q, y, phi come from the block's LDS tile `pl` (FP32 only).
// G: the gauge group (4: SU(2) x U(1); 5, 7: the 2HDM with the Higgs-family
// U(1) or SU(2), whose links are read from global memory in every staging mode).
template <class R, int ND, bool CMP, bool PL = false, int G = 4>
__device__ __forceinline__ void mom_site(const StepArgs<R>& a, const Coef& c, int x, int y, int z,
                                         const PlaneLds& pl = PlaneLds{}) {
    using V4 = typename V4T<R>::type;
    static_assert(G == 4 || ND == 2, "a Higgs-family field acts on two doublets");
    constexpr bool NB = G == 7;

    // 32-bit site offsets (the host limits a lattice to 2^28 sites): addresses
    // become a scalar base plus one 32-bit vector offset, with no 64-bit math.
    const int sy = a.nz, sx = a.ny * a.nz;
    const unsigned s = static_cast<unsigned>(x * sx + y * sy + z);
    // Shell of a fixed face: computed like any site, then carried over unchanged.
    const bool frozen = site_frozen(c, x, y, z);
    // Offsets to the +e_a and -e_a neighbours with periodic wrap.
    int dp[3], dm[3];
    dp[0] = (x == a.nx - 1) ? -(a.nx - 1) * sx : sx;
    dm[0] = (x == 0) ? (a.nx - 1) * sx : -sx;
    dp[1] = (y == a.ny - 1) ? -(a.ny - 1) * sy : sy;
    dm[1] = (y == 0) ? (a.ny - 1) * sy : -sy;
    dp[2] = (z == a.nz - 1) ? -(a.nz - 1) : 1;
    dm[2] = (z == 0) ? (a.nz - 1) : -1;

    // Site x + si e_i + sj e_j (si, sj in {-1, 0, 1}; all compile-time after unrolling).
    auto gsite = [&](int i, int si, int j, int sj) -> unsigned {
        int o = 0;
        if (si > 0) o += dp[i];
        else if (si < 0) o += dm[i];
        if (sj > 0) o += dp[j];
        else if (sj < 0) o += dm[j];
        return s + o;
    };
    auto in_plane = [&](int i, int si, int j, int sj) { return PL && !(si != 0 && i == 0) && !(sj != 0 && j == 0); };
    auto slot = [&](int i, int si, int j, int sj) {
        const int step[3] = {0, pl.w, 1};
        return pl.r0 + si * step[i] + sj * step[j];
    };
    auto Q = [&](int f, int i, int si, int j, int sj) -> Qt<R> {
        if constexpr (PL) {
            if (in_plane(i, si, j, sj)) return toq<float4, R>(pl.q[f][slot(i, si, j, sj)]);
        }
        return toq<V4, R>(at(a.q[f], gsite(i, si, j, sj)));
    };
    auto Y = [&](int f, int i, int si, int j, int sj) -> R {
        if constexpr (PL) {
            if (in_plane(i, si, j, sj)) return pl.y[f][slot(i, si, j, sj)];
        }
        return at(a.y[f], gsite(i, si, j, sj));
    };
    // Stored doublet k at x + si e_i, and the full doublet (stored + reference).
    auto PHIRAW = [&](int k, int i, int si, R* o) {
        if constexpr (PL) {
            if (in_plane(i, si, 0, 0)) { unpack(pl.phi[k][slot(i, si, 0, 0)], o); return; }
        }
        unpack(at(a.phi[k], gsite(i, si, 0, 0)), o);
    };
    auto PHI = [&](int k, int i, int si, R* o) {
        R raw[4];
        PHIRAW(k, i, si, raw);
        add_ref(a.ref[k], raw, o);
    };

    // Own links and their phases.
    Qt<R> U[3];
    R cy[3], sy_[3], yv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        U[i] = Q(i, 0, 0, 0, 0);
        yv[i] = Y(i, 0, 0, 0, 0);
        sincos_k<NB>(yv[i], &sy_[i], &cy[i]);
    }

    R phi[4 * ND], praw[4 * ND];
#pragma unroll
    for (int k = 0; k < ND; k++) {
        PHIRAW(k, 0, 0, praw + 4 * k);
        add_ref(a.ref[k], praw + 4 * k, phi + 4 * k);
    }

    // Higgs-family links (global memory): phase y' or quaternion of direction f
    // at x + si e_i + sj e_j, and the site's own.
    auto FY = [&](int f, int i, int si, int j, int sj) -> R { return at(a.fy[f], gsite(i, si, j, sj)); };
    auto FQ = [&](int f, int i, int si, int j, int sj) -> Qt<R> {
        return toq<V4, R>(at(a.fq[f], gsite(i, si, j, sj)));
    };
    auto FAM = [&](int f, int i, int si) -> FamLink<R> {
        FamLink<R> l{};
        if constexpr (G == 5) sincos_nb(FY(f, i, si, 0, 0), &l.s, &l.c);
        else if constexpr (G == 7) l.v = FQ(f, i, si, 0, 0);
        return l;
    };
    R fyv[3] = {R(0), R(0), R(0)};
    Qt<R> FV[3];
    FamLink<R> FL[3];
    if constexpr (G != 4) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            FL[i] = FAM(i, 0, 0);
            if constexpr (G == 5) fyv[i] = FY(i, 0, 0, 0, 0);
            else FV[i] = FL[i].v;
        }
    }

    // Scalar force and currents.
    R F[4 * ND];
    potential_derivative<R>(c, phi, F);
#pragma unroll
    for (int k = 0; k < 4 * ND; k++) F[k] = -F[k];
    R J[3][4];
    R JF[3][3];   // Higgs-family currents (G 5: [i][0]; 7: [i][0..2])
    if constexpr (G != 4) {
        // Both doublets at once: the family part of T mixes them for G = 7.
        // The same covariant differences as below carry the Laplacian and all
        // currents (every bilinear vanishes on equal arguments).
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const Qt<R> Um = Q(i, i, -1, 0, 0);
            R cm, smn;
            sincos_k<NB>(Y(i, i, -1, 0, 0), &smn, &cm);
            const FamLink<R> Fm = FAM(i, i, -1);
            J[i][0] = J[i][1] = J[i][2] = J[i][3] = R(0);
            JF[i][0] = JF[i][1] = JF[i][2] = R(0);
            R fp[8], fm[8], tx[8], tm[8], d[8], dmv[8], dpv[8];
#pragma unroll
            for (int k = 0; k < 2; k++) {
                PHI(k, i, 1, fp + 4 * k);
                PHI(k, i, -1, fm + 4 * k);
            }
            transport2<R, G>(U[i], cy[i], sy_[i], FL[i], false, phi, tx);   // T_i(x) phi(x)
            transport2<R, G>(Um, cm, smn, Fm, false, fm, tm);                // T_i(x-i) phi(x-i)
#pragma unroll
            for (int m = 0; m < 8; m++) {
                d[m] = fp[m] - tx[m];
                dmv[m] = phi[m] - tm[m];
            }
            transport2<R, G>(U[i], cy[i], sy_[i], FL[i], true, d, dpv);      // T_i(x)^dag d
#pragma unroll
            for (int m = 0; m < 8; m++) F[m] += a.inv_sp[i] * (dpv[m] - dmv[m]);
            bilinears_add(tx, d, J[i]);
            bilinears_add(tx + 4, d + 4, J[i]);
            family_bilinears<R, G>(tx, d, JF[i]);
        }
    } else {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const Qt<R> Um = Q(i, i, -1, 0, 0);
        R cm, smn;
        sincos_k<NB>(Y(i, i, -1, 0, 0), &smn, &cm);
        J[i][0] = J[i][1] = J[i][2] = J[i][3] = R(0);
#pragma unroll
        for (int k = 0; k < ND; k++) {
            // The covariant differences along link i at x and at x-i,
            //   d = phi(x+i) - T_i(x) phi(x),  dm = phi(x) - T_i(x-i) phi(x-i),
            // carry the Laplacian and the current: T^dag phi(x+i) + T phi(x-i) - 2 phi
            // = T_i(x)^dag d - dm and B(T phi, phi(x+i)) = B(T phi, d) (the
            // bilinears vanish on equal arguments).  dm is bitwise the d of site x-i,
            // so the current on a link and the charges it moves at its two ends
            // are computed from the same small differences, not from O(1) terms.
            R fp[4], fm[4], tx[4], tm[4], d[4], dmv[4], dpv[4];
            PHI(k, i, 1, fp);
            PHI(k, i, -1, fm);
            transport(U[i], cy[i], sy_[i], false, phi + 4 * k, tx);  // T_i(x) phi(x)
            transport(Um, cm, smn, false, fm, tm);                    // T_i(x-i) phi(x-i)
#pragma unroll
            for (int m = 0; m < 4; m++) {
                d[m] = fp[m] - tx[m];
                dmv[m] = phi[4 * k + m] - tm[m];
            }
            transport(U[i], cy[i], sy_[i], true, d, dpv);             // T_i(x)^dag d
#pragma unroll
            for (int m = 0; m < 4; m++) F[4 * k + m] += a.inv_sp[i] * (dpv[m] - dmv[m]);
            bilinears_add(tx, d, J[i]);
        }
    }
    }

    // Magnetic force (staples, WilsonLoops/Source/StandardModel.cpp:324-409).
    // Each plaquette of the plane (lo, hi), lo < hi, based at corner b enters the
    // forces on its four links through one rounded quaternion and one rounded
    // angle, computed in the same operation order wherever it is needed:
    //   K_b = A_b C_b,  A_b = U_hi(b+lo)^dag U_lo(b+hi),  C_b = U_hi(b) U_lo(b)^dag,
    //   theta_b = (y_lo(b+hi) - y_lo(b)) + (y_hi(b) - y_hi(b+lo)).
    // With v = vec(K_b) the four contributions are
    //   link lo at b:      +vec K_b            link hi at b:      -rot(C_b) v
    //   link lo at b+hi:   -rot(U_hi(b+lo)) v  link hi at b+lo:   +rot(U_hi(b+lo)) v
    // (rot(q) v = vec(q v q^dag)), with sin(theta_b) for the U(1) part in the same
    // pattern; in exact arithmetic these are the reference's staple sums.  Sharing
    // the rounded K_b and theta_b makes the contributions of one plaquette cancel in
    // the Gauss law up to the rounding of the rotations, which is relative to the
    // plaquette's small deviation from 1 instead of to the size of the links.
    R M[3][4];
#pragma unroll
    for (int i = 0; i < 3; i++) M[i][0] = M[i][1] = M[i][2] = M[i][3] = R(0);
    // Plaquettes based at x (b = x, lo = i, hi = j).
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            const Qt<R> Ujxpi = Q(j, i, 1, 0, 0);
            const Qt<R> Uixpj = Q(i, j, 1, 0, 0);
            const R th = (Y(i, j, 1, 0, 0) - yv[i]) + (yv[j] - Y(j, i, 1, 0, 0));
            const Qt<R> A = qmul(qconj(Ujxpi), Uixpj);
            const Qt<R> C = qmul(U[j], qconj(U[i]));
            R v[3], w[3];
            qvec(A, C, v);
            qrot(C, v, w);
            const R st = sin_k<NB>(th);
            M[i][0] += a.inv_sp[j] * st;
            M[j][0] -= a.inv_sp[i] * st;
#pragma unroll
            for (int m = 0; m < 3; m++) {
                M[i][1 + m] += a.inv_sp[j] * v[m];
                M[j][1 + m] -= a.inv_sp[i] * w[m];
            }
        }
    }
    // Plaquettes based at b = x - j, for link i at x: i < j makes x = b + hi
    // (link lo at b+hi), i > j makes x = b + lo (link hi at b+lo).
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const Qt<R> Ujm = Q(j, j, -1, 0, 0);
        const R yjm = Y(j, j, -1, 0, 0);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i == j) continue;
            const Qt<R> Uim = Q(i, j, -1, 0, 0);
            const Qt<R> Ujd = Q(j, i, 1, j, -1);
            const R yim = Y(i, j, -1, 0, 0), yjd = Y(j, i, 1, j, -1);
            R v[3], w[3];
            if (i < j) {
                // lo = i, hi = j: A_b = U_j(x+i-j)^dag U_i(x), C_b = U_j(x-j) U_i(x-j)^dag.
                const R th = (yv[i] - yim) + (yjm - yjd);
                qvec(qmul(qconj(Ujd), U[i]), qmul(Ujm, qconj(Uim)), v);
                qrot(Ujd, v, w);
                M[i][0] -= a.inv_sp[j] * sin_k<NB>(th);
#pragma unroll
                for (int m = 0; m < 3; m++) M[i][1 + m] -= a.inv_sp[j] * w[m];
            } else {
                // lo = j, hi = i: A_b = U_i(x)^dag U_j(x+i-j), C_b = U_i(x-j) U_j(x-j)^dag.
                const R th = (yjd - yjm) + (yim - yv[i]);
                qvec(qmul(qconj(U[i]), Ujd), qmul(Uim, qconj(Ujm)), v);
                qrot(U[i], v, w);
                M[i][0] += a.inv_sp[j] * sin_k<NB>(th);
#pragma unroll
                for (int m = 0; m < 3; m++) M[i][1 + m] += a.inv_sp[j] * w[m];
            }
        }
        if (j == 0) fold_current(a, J, M);
    }

    kick_drift<R, ND, CMP, false, G == 5>(a, s, U, yv, praw, F, M, nullptr, frozen, c.fixed != 0);

    if constexpr (G != 4) {
        // Higgs-family staples, source and kick.
        R MF[3][4];
        family_force<R, G>(a, FY, FQ, fyv, FV, JF, MF);
        kick_drift_family<R, G, CMP>(a, s, fyv, FV, MF, frozen, c.fixed != 0);
    }
}

// ---- shared plaquettes and covariant differences (FP32 SM, staged) ----------
// Every plaquette is needed by the three sites whose links it touches and every
// covariant difference d_i(x) = phi(x+i) - T_i(x) phi(x) by two sites.  Here a
// block evaluates each in-plane one once: in phase A every site computes its
// own three plaquettes (base b = x) and its own d_i, and the sites of the ring
// row y0 - 1 and ring column z0 - 1 compute theirs; after a barrier they go to
// LDS (over the plane tile, no longer read) and in phase B each site reads the
// plaquettes based at x - e_y and x - e_z and the differences d_y(x - e_y),
// d_z(x - e_z) instead of recomputing them.  Plaquettes based at x - e_x and
// d_x(x - e_x) lie in the plane x - 1 and are still computed per site.  Every
// shared value is produced by the same expressions, in the same operation
// order, as the per-site evaluation in mom_site, and the forces accumulate in
// the same order: the fields are bitwise those of mom_site.

// One plaquette of the plane (lo, hi) based at b from its four links and
// phases (U_hi(b+lo) = Uhi_lo, U_lo(b+hi) = Ulo_hi):
//   K_b = A_b C_b, A_b = U_hi(b+lo)^dag U_lo(b+hi), C_b = U_hi(b) U_lo(b)^dag,
//   theta_b = (y_lo(b+hi) - y_lo(b)) + (y_hi(b) - y_hi(b+lo));
// v = vec K_b, wc = rot(C_b) v, wu = rot(U_hi(b+lo)) v, st = sin theta_b.
template <class R>
struct Plaq {
    R v[3], wc[3], wu[3], st;
};
template <bool NB, class R>
__device__ __forceinline__ void plaquette(const Qt<R>& Ulo, const Qt<R>& Uhi, const Qt<R>& Uhi_lo, const Qt<R>& Ulo_hi,
                                          R ylo, R yhi, R yhi_lo, R ylo_hi, Plaq<R>& p) {
    const R th = (ylo_hi - ylo) + (yhi - yhi_lo);
    const Qt<R> A = qmul(qconj(Uhi_lo), Ulo_hi);
    const Qt<R> C = qmul(Uhi, qconj(Ulo));
    qvec(A, C, p.v);
    qrot(C, p.v, p.wc);
    qrot(Uhi_lo, p.v, p.wu);
    p.st = sin_k<NB>(th);
}

// Shared values in LDS, one float4 per site of the tile plus ring row and ring
// column: index (ty + 1) * (bz + 1) + (tz + 1) for ty in [-1, by), tz in [-1, bz).
//   d1, d2   covariant differences d_y, d_z at the site
//   p01, p02, p12   (rot(U_hi(b+lo)) v, sin theta) of the plaquette based there
struct ShareLds {
    float4* d1;
    float4* d2;
    float4* p01;
    float4* p02;
    float4* p12;
    float4* d1b;   // the second doublet's d_y, d_z (two doublets only)
    float4* d2b;
    int w;   // bz + 1
};
__host__ __device__ inline size_t share_lds_bytes(int bz, int by, int nd) {
    return (size_t)(bz + 1) * (by + 1) * (nd == 2 ? 7 : 5) * 16;
}
__device__ __forceinline__ ShareLds share_layout(char* base, int bz, int by) {
    const int n = (bz + 1) * (by + 1);
    float4* f = reinterpret_cast<float4*>(base);
    return ShareLds{f, f + n, f + 2 * n, f + 3 * n, f + 4 * n, f + 5 * n, f + 6 * n, bz + 1};
}
__device__ __forceinline__ float4 pack_wu(const Plaq<float>& p) { return make_float4(p.wu[0], p.wu[1], p.wu[2], p.st); }

// Whole block, after plane_fill: one site-update per valid thread (tile
// coordinates ty, tz; site (x, y0 + ty, z0 + tz)) with the in-plane plaquettes
// and differences shared.  Needs bz + by <= 256 (one ring item per thread).
// EARLY: the site's own P, E and residuals are loaded at the start of phase A,
// where their latency overlaps the phase (at the cost of registers held across
// the exchange barriers); otherwise after the exchange, next to their use.
// ND: doublets (the 2HDM shares both doublets' differences through LDS as
// well).  G 5 / 7 (two doublets): the differences carry the Higgs-family part
// of the transport (the family links read from global memory), the family
// staples are evaluated per site from global memory, and the family kick
// follows the SU(2) x U(1) one.
template <int ND, bool CMP, bool EARLY = true, int G = 4>
__device__ __forceinline__ void mom_block_shared(const StepArgs<float>& a, const Coef& c, int x, int y0, int z0, int ty,
                                                 int tz, bool valid, const PlaneLds& pl, char* lds) {
    using R = float;
    constexpr bool NB = G == 7;
    const int bz = a.bz, W = pl.w;
    const int sx = a.ny * a.nz;
    const int xp = (x == a.nx - 1) ? 0 : x + 1, xm = (x == 0) ? a.nx - 1 : x - 1;

    // ---- phase A: own site ----
    const int y = y0 + ty, z = z0 + tz;
    const unsigned s = static_cast<unsigned>(x * sx + y * a.nz + z);
    const int r0 = pl.r0;
    const unsigned sm0 = static_cast<unsigned>(xm * sx + y * a.nz + z);   // x - e_0
    Qt<R> U[3];
    R cy[3], sy_[3], yv[3], phi[4 * ND], praw[4 * ND], F[4 * ND], J[3][4], M[3][4], dpv[3][4 * ND], dsh[3][4 * ND];
    Plaq<R> P[3];   // at x: (0,1), (0,2), (1,2)
    static_assert(G == 4 || ND == 2, "a Higgs-family field acts on two doublets");
    // Higgs-family links: direction f at x + si e_i + sj e_j (global memory).
    const int nz = a.nz;
    int dp[3], dm[3];
    dp[0] = (x == a.nx - 1) ? -(a.nx - 1) * sx : sx;
    dm[0] = (x == 0) ? (a.nx - 1) * sx : -sx;
    dp[1] = (y == a.ny - 1) ? -(a.ny - 1) * nz : nz;
    dm[1] = (y == 0) ? (a.ny - 1) * nz : -nz;
    dp[2] = (z == a.nz - 1) ? -(a.nz - 1) : 1;
    dm[2] = (z == 0) ? (a.nz - 1) : -1;
    auto fsite = [&](int i, int si, int j, int sj) -> unsigned {
        int o = 0;
        if (si > 0) o += dp[i];
        else if (si < 0) o += dm[i];
        if (sj > 0) o += dp[j];
        else if (sj < 0) o += dm[j];
        return s + o;
    };
    auto FY = [&](int f, int i, int si, int j, int sj) -> R { return at(a.fy[f], fsite(i, si, j, sj)); };
    auto FQ = [&](int f, int i, int si, int j, int sj) -> Qt<R> { return toq<float4, R>(at(a.fq[f], fsite(i, si, j, sj))); };
    auto FAMAT = [&](int f, unsigned site) -> FamLink<R> {
        FamLink<R> l{};
        if constexpr (G == 5) sincos_nb(at(a.fy[f], site), &l.s, &l.c);
        else if constexpr (G == 7) l.v = toq<float4, R>(at(a.fq[f], site));
        return l;
    };
    R fyv[3] = {R(0), R(0), R(0)}, JF[3][3], MF[3][4];
    Qt<R> FV[3];
    FamLink<R> FL[3];
    OwnState<R, ND> own;
    if (EARLY && valid) load_own<R, ND, CMP>(a, s, own);   // issued early: its latency overlaps phase A
    // Back to the default priority (mom_step raised it for the loads), in every
    // wave: a wave with no valid lane must not keep the raised priority.
    __builtin_amdgcn_s_setprio(0);
    if (valid) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            U[i] = toq<float4, R>(pl.q[i][r0]);
            yv[i] = pl.y[i][r0];
            sincos_k<NB>(yv[i], &sy_[i], &cy[i]);
        }
#pragma unroll
        for (int k = 0; k < ND; k++) {
            unpack(pl.phi[k][r0], praw + 4 * k);
            add_ref(a.ref[k], praw + 4 * k, phi + 4 * k);
        }
        potential_derivative<R>(c, phi, F);
#pragma unroll
        for (int k = 0; k < 4 * ND; k++) F[k] = -F[k];
        // Neighbours x + e_i: plane x + 1 (global) for i = 0, LDS for i = 1, 2.
        const unsigned sp0 = static_cast<unsigned>(xp * sx + y * a.nz + z);
        const int step[3] = {0, W, 1};
        if constexpr (G != 4) {
#pragma unroll
            for (int i = 0; i < 3; i++) {
                FL[i] = FAMAT(i, s);
                if constexpr (G == 5) fyv[i] = at(a.fy[i], s);
                else FV[i] = FL[i].v;
            }
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            J[i][0] = J[i][1] = J[i][2] = J[i][3] = R(0);
            if constexpr (G != 4) {
                // Both doublets at once: the family part of T mixes them (G = 7).
                JF[i][0] = JF[i][1] = JF[i][2] = R(0);
                R fp[8], tx[8];
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    R fr[4];
                    if (i == 0) unpack(at(a.phi[k], sp0), fr);
                    else unpack(pl.phi[k][r0 + step[i]], fr);
                    add_ref(a.ref[k], fr, fp + 4 * k);
                }
                transport2<R, G>(U[i], cy[i], sy_[i], FL[i], false, phi, tx);
#pragma unroll
                for (int m = 0; m < 8; m++) dsh[i][m] = fp[m] - tx[m];
                transport2<R, G>(U[i], cy[i], sy_[i], FL[i], true, dsh[i], dpv[i]);
                bilinears_add(tx, dsh[i], J[i]);
                bilinears_add(tx + 4, dsh[i] + 4, J[i]);
                family_bilinears<R, G>(tx, dsh[i], JF[i]);
                continue;
            }
#pragma unroll
            for (int k = 0; k < ND; k++) {
                R fr[4], fp[4], tx[4];
                if (i == 0) unpack(at(a.phi[k], sp0), fr);
                else unpack(pl.phi[k][r0 + step[i]], fr);
                add_ref(a.ref[k], fr, fp);
                transport(U[i], cy[i], sy_[i], false, phi + 4 * k, tx);
#pragma unroll
                for (int m = 0; m < 4; m++) dsh[i][4 * k + m] = fp[m] - tx[m];
                transport(U[i], cy[i], sy_[i], true, dsh[i] + 4 * k, dpv[i] + 4 * k);
                bilinears_add(tx, dsh[i] + 4 * k, J[i]);
            }
        }
        // d_x(x - e_x) from the plane x - 1; the scalar force's i = 0 term.
        {
            const Qt<R> Um = toq<float4, R>(at(a.q[0], sm0));
            R cm, smn;
            sincos_k<NB>(at(a.y[0], sm0), &smn, &cm);
            if constexpr (G != 4) {
                R fm[8], tm[8];
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    R fr[4];
                    unpack(at(a.phi[k], sm0), fr);
                    add_ref(a.ref[k], fr, fm + 4 * k);
                }
                transport2<R, G>(Um, cm, smn, FAMAT(0, sm0), false, fm, tm);
#pragma unroll
                for (int m = 0; m < 8; m++) F[m] += a.inv_sp[0] * (dpv[0][m] - (phi[m] - tm[m]));
            } else {
#pragma unroll
                for (int k = 0; k < ND; k++) {
                    R fr[4], fm[4], tm[4], dmv0[4];
                    unpack(at(a.phi[k], sm0), fr);
                    add_ref(a.ref[k], fr, fm);
                    transport(Um, cm, smn, false, fm, tm);
#pragma unroll
                    for (int m = 0; m < 4; m++) dmv0[m] = phi[4 * k + m] - tm[m];
#pragma unroll
                    for (int m = 0; m < 4; m++) F[4 * k + m] += a.inv_sp[0] * (dpv[0][4 * k + m] - dmv0[m]);
                }
            }
        }
        // Plaquettes based at x and their contributions to the links at x.
#pragma unroll
        for (int i = 0; i < 3; i++) M[i][0] = M[i][1] = M[i][2] = M[i][3] = R(0);
#pragma unroll
        for (int i = 0; i < 3; i++) {
#pragma unroll
            for (int j = i + 1; j < 3; j++) {
                Plaq<R>& pq = P[i + j - 1];
                Qt<R> Uhi_lo, Ulo_hi;
                R yhi_lo, ylo_hi;
                if (i == 0) {   // U_j(x + e_0): plane x + 1
                    Uhi_lo = toq<float4, R>(at(a.q[j], sp0));
                    yhi_lo = at(a.y[j], sp0);
                } else {
                    Uhi_lo = toq<float4, R>(pl.q[j][r0 + step[i]]);
                    yhi_lo = pl.y[j][r0 + step[i]];
                }
                Ulo_hi = toq<float4, R>(pl.q[i][r0 + step[j]]);
                ylo_hi = pl.y[i][r0 + step[j]];
                plaquette<NB>(U[i], U[j], Uhi_lo, Ulo_hi, yv[i], yv[j], yhi_lo, ylo_hi, pq);
                M[i][0] += a.inv_sp[j] * pq.st;
                M[j][0] -= a.inv_sp[i] * pq.st;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    M[i][1 + m] += a.inv_sp[j] * pq.v[m];
                    M[j][1 + m] -= a.inv_sp[i] * pq.wc[m];
                }
            }
        }
        // Plaquettes based at x - e_0 (plane x - 1): the j = 0 terms of links 1
        // and 2, which follow the plaquettes at x in their force sums.
        {
            const Qt<R> U0m = toq<float4, R>(at(a.q[0], sm0));
            const R y0m = at(a.y[0], sm0);
            const unsigned sm0p1 = static_cast<unsigned>(xm * sx + (y == a.ny - 1 ? 0 : y + 1) * a.nz + z);
            const unsigned sm0p2 = static_cast<unsigned>(xm * sx + y * a.nz + (z == a.nz - 1 ? 0 : z + 1));
#pragma unroll
            for (int i = 1; i < 3; i++) {
                const unsigned sd = i == 1 ? sm0p1 : sm0p2;       // x + e_i - e_0
                Plaq<R> pq;
                plaquette<NB>(U0m, toq<float4, R>(at(a.q[i], sm0)), U[i], toq<float4, R>(at(a.q[0], sd)), y0m,
                          at(a.y[i], sm0), yv[i], at(a.y[0], sd), pq);
                M[i][0] += a.inv_sp[0] * pq.st;
#pragma unroll
                for (int m = 0; m < 3; m++) M[i][1 + m] += a.inv_sp[0] * pq.wu[m];
            }
        }
        fold_current(a, J, M);
        if constexpr (G != 4) family_force<R, G>(a, FY, FQ, fyv, FV, JF, MF);
    }

    // ---- phase A: ring item (y-ring site for k < bz, z-ring site after) ----
    const int k = static_cast<int>(threadIdx.x);
    const bool ring = k < bz + a.by;
    const int rdir = k < bz ? 1 : 2;                      // the ring's direction
    const int rt = k - bz;                                // z-ring: tile row
    const bool ring_valid = ring && (rdir == 1 ? z0 + k < a.nz : y0 + rt < a.ny);
    float4 ring_d[ND], ring_p0, ring_p12;
    if (ring_valid) {
        const int rb = rdir == 1 ? k + 1 : (rt + 1) * W;      // plane slot of the ring site b
        const int rbr = rdir == 1 ? rb + W : rb + 1;          // b + e_r (a tile site)
        const int yy = rdir == 1 ? (y0 == 0 ? a.ny - 1 : y0 - 1) : y0 + rt;
        const int zz = rdir == 1 ? z0 + k : (z0 == 0 ? a.nz - 1 : z0 - 1);
        const unsigned sbp0 = static_cast<unsigned>(xp * sx + yy * a.nz + zz);   // b + e_0
        Qt<R> Ub[3];
        R yb[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            Ub[i] = toq<float4, R>(pl.q[i][rb]);
            yb[i] = pl.y[i][rb];
        }
        // d_r(b) = phi(b + e_r) - T_r(b) phi(b).
        {
            const Qt<R> Ur = rdir == 1 ? Ub[1] : Ub[2];
            const R yr = rdir == 1 ? yb[1] : yb[2];
            R cr, sr;
            sincos_k<NB>(yr, &sr, &cr);
            if constexpr (G != 4) {
                const unsigned sb = static_cast<unsigned>(x * sx + yy * a.nz + zz);   // the ring site b
                R fb[8], fn[8], tx[8];
#pragma unroll
                for (int kd = 0; kd < 2; kd++) {
                    R rawb[4], rawn[4];
                    unpack(pl.phi[kd][rb], rawb);
                    unpack(pl.phi[kd][rbr], rawn);
                    add_ref(a.ref[kd], rawb, fb + 4 * kd);
                    add_ref(a.ref[kd], rawn, fn + 4 * kd);
                }
                transport2<R, G>(Ur, cr, sr, FAMAT(rdir, sb), false, fb, tx);
#pragma unroll
                for (int kd = 0; kd < 2; kd++)
                    ring_d[kd] = make_float4(fn[4 * kd] - tx[4 * kd], fn[4 * kd + 1] - tx[4 * kd + 1],
                                             fn[4 * kd + 2] - tx[4 * kd + 2], fn[4 * kd + 3] - tx[4 * kd + 3]);
            } else {
#pragma unroll
                for (int kd = 0; kd < ND; kd++) {
                    R rawb[4], rawn[4], fb[4], fn[4], tx[4], d[4];
                    unpack(pl.phi[kd][rb], rawb);
                    unpack(pl.phi[kd][rbr], rawn);
                    add_ref(a.ref[kd], rawb, fb);
                    add_ref(a.ref[kd], rawn, fn);
                    transport(Ur, cr, sr, false, fb, tx);
#pragma unroll
                    for (int m = 0; m < 4; m++) d[m] = fn[m] - tx[m];
                    ring_d[kd] = make_float4(d[0], d[1], d[2], d[3]);
                }
            }
        }
        // Plaquette (0, r) based at b: U_r(b + e_0) from the plane x + 1.
        {
            Plaq<R> pq;
            const Qt<R> Ur = rdir == 1 ? Ub[1] : Ub[2];
            const R yr = rdir == 1 ? yb[1] : yb[2];
            const float4 qrn = rdir == 1 ? at(a.q[1], sbp0) : at(a.q[2], sbp0);
            const R yrn = rdir == 1 ? at(a.y[1], sbp0) : at(a.y[2], sbp0);
            plaquette<NB>(Ub[0], Ur, toq<float4, R>(qrn), toq<float4, R>(pl.q[0][rbr]), yb[0], yr, yrn, pl.y[0][rbr], pq);
            ring_p0 = pack_wu(pq);
        }
        // Plaquette (1, 2) based at b: U_2(b + e_1), U_1(b + e_2) in the plane.
        {
            Plaq<R> pq;
            plaquette<NB>(Ub[1], Ub[2], toq<float4, R>(pl.q[2][rb + W]), toq<float4, R>(pl.q[1][rb + 1]), yb[1], yb[2],
                      pl.y[2][rb + W], pl.y[1][rb + 1], pq);
            ring_p12 = pack_wu(pq);
        }
    }

    // ---- exchange through LDS (over the plane tile) ----
    __syncthreads();
    const ShareLds sh = share_layout(lds, bz, a.by);
    const int si = (ty + 1) * sh.w + tz + 1;
    if (valid) {
        sh.d1[si] = make_float4(dsh[1][0], dsh[1][1], dsh[1][2], dsh[1][3]);
        sh.d2[si] = make_float4(dsh[2][0], dsh[2][1], dsh[2][2], dsh[2][3]);
        if constexpr (ND == 2) {
            sh.d1b[si] = make_float4(dsh[1][4], dsh[1][5], dsh[1][6], dsh[1][7]);
            sh.d2b[si] = make_float4(dsh[2][4], dsh[2][5], dsh[2][6], dsh[2][7]);
        }
        sh.p01[si] = pack_wu(P[0]);
        sh.p02[si] = pack_wu(P[1]);
        sh.p12[si] = pack_wu(P[2]);
    }
    if (ring_valid) {
        const int ri = rdir == 1 ? k + 1 : (rt + 1) * sh.w;
        if (rdir == 1) {
            sh.d1[ri] = ring_d[0];
            if constexpr (ND == 2) sh.d1b[ri] = ring_d[1];
            sh.p01[ri] = ring_p0;
        } else {
            sh.d2[ri] = ring_d[0];
            if constexpr (ND == 2) sh.d2b[ri] = ring_d[1];
            sh.p02[ri] = ring_p0;
        }
        sh.p12[ri] = ring_p12;
    }
    __syncthreads();
    if (!valid) return;
    if constexpr (!EARLY) load_own<R, ND, CMP>(a, s, own);
    // Shell of a fixed face: its plaquettes and differences served the
    // neighbours above; the site is computed like any other, then carried over
    // unchanged by the kick's stores.
    const bool frozen = site_frozen(c, x, y, z);

    // ---- phase B: the scalar force's i = 1, 2 terms ----
#pragma unroll
    for (int kd = 0; kd < ND; kd++) {
        const float4 d1m = kd ? sh.d1b[si - sh.w] : sh.d1[si - sh.w], d2m = kd ? sh.d2b[si - 1] : sh.d2[si - 1];
        const R dm1[4] = {d1m.x, d1m.y, d1m.z, d1m.w}, dm2[4] = {d2m.x, d2m.y, d2m.z, d2m.w};
#pragma unroll
        for (int m = 0; m < 4; m++) F[4 * kd + m] += a.inv_sp[1] * (dpv[1][4 * kd + m] - dm1[m]);
#pragma unroll
        for (int m = 0; m < 4; m++) F[4 * kd + m] += a.inv_sp[2] * (dpv[2][4 * kd + m] - dm2[m]);
    }
    // Plaquettes based at x - e_1 and x - e_2 from LDS, in the order j = 1, 2.
    {
        const float4 p01 = sh.p01[si - sh.w], p12a = sh.p12[si - sh.w];   // based at x - e_1
        const float4 p02 = sh.p02[si - 1], p12b = sh.p12[si - 1];         // based at x - e_2
        // j = 1: link 0 (i < j), link 2 (i > j).
        M[0][0] -= a.inv_sp[1] * p01.w;
        M[0][1] -= a.inv_sp[1] * p01.x;
        M[0][2] -= a.inv_sp[1] * p01.y;
        M[0][3] -= a.inv_sp[1] * p01.z;
        M[2][0] += a.inv_sp[1] * p12a.w;
        M[2][1] += a.inv_sp[1] * p12a.x;
        M[2][2] += a.inv_sp[1] * p12a.y;
        M[2][3] += a.inv_sp[1] * p12a.z;
        // j = 2: links 0 and 1 (i < j).
        M[0][0] -= a.inv_sp[2] * p02.w;
        M[0][1] -= a.inv_sp[2] * p02.x;
        M[0][2] -= a.inv_sp[2] * p02.y;
        M[0][3] -= a.inv_sp[2] * p02.z;
        M[1][0] -= a.inv_sp[2] * p12b.w;
        M[1][1] -= a.inv_sp[2] * p12b.x;
        M[1][2] -= a.inv_sp[2] * p12b.y;
        M[1][3] -= a.inv_sp[2] * p12b.z;
    }
    kick_drift<R, ND, CMP, true, G == 5>(a, s, U, yv, praw, F, M, &own, frozen, c.fixed != 0);
    if constexpr (G != 4) kick_drift_family<R, G, CMP>(a, s, fyv, FV, MF, frozen, c.fixed != 0);
}

// MINW: minimum waves per SIMD the register allocation must allow (0: the
// compiler's choice).  Higher occupancy hides more memory latency but may spill.
// PL: 0 every neighbour from global memory; 1 in-plane neighbours from the
// block's LDS plane tile; 2 as 1, with the in-plane plaquettes and covariant
// differences shared (FP32 SM, bz + by <= 256).  All give bitwise the same fields.
// G: gauge group (4, or 5 / 7 for the 2HDM with a Higgs-family field; PL <= 1).
template <class R, int ND, int MINW, bool CMP, int PL = 0, int G = 4>
__global__ void __launch_bounds__(256, MINW > 0 ? MINW : 1) mom_step(StepArgs<R> a, Coef c) {
    const int4 t = a.tiles[blockIdx.x];
    const int tz = static_cast<int>(threadIdx.x) & (a.bz - 1), ty = static_cast<int>(threadIdx.x) >> a.bz_shift;
    const int z = t.w * a.bz + tz;
    const int y = t.z * a.by + ty;
    const bool valid = z < a.nz && y < a.ny;
    if constexpr (PL > 0) {
        static_assert(sizeof(R) == 4, "plane staging is FP32 only");
        static_assert(PL == 1 || G == 4 || ND == 2, "a Higgs-family field acts on two doublets");
        extern __shared__ float4 gft_plane_lds[];
        char* lds = reinterpret_cast<char*>(gft_plane_lds);
        PlaneLds pl = plane_layout<ND>(lds, a.bz, a.by);
        // Stage 2: the wave issues at raised priority until its own-state loads
        // are out (mom_block_shared drops it), so a wave starting a tile gets
        // its memory requests in ahead of the waves that are computing (1.5%
        // faster at 256^3; the ensemble kernel, with 16 bands per member, 3.5%
        // faster).  Priority changes issue order only, not results.
        if constexpr (PL == 2) __builtin_amdgcn_s_setprio(3);
        plane_fill<ND>(a, pl, t.y, t.z * a.by, t.w * a.bz, a.bz, a.by);
        pl.r0 = (ty + 1) * pl.w + tz + 1;
        if constexpr (PL == 2) {
            // At 4 waves per SIMD (and with two doublets) the registers do not hold
            // the own state across the exchange barriers: it is loaded after them.
            mom_block_shared<ND, CMP, ND == 1 && MINW != 4, G>(a, c, t.y, t.z * a.by, t.w * a.bz, ty, tz, valid,
                                                               pl, lds);
        } else {
            if (!valid) return;
            mom_site<R, ND, CMP, true, G>(a, c, t.y, y, z, pl);
        }
    } else {
        if (!valid) return;
        mom_site<R, ND, CMP, false, G>(a, c, t.y, y, z);
    }
}

// The default single-doublet FP32 step (4 waves per SIMD, shared staging,
// compensated and plain storage) is compiled in its own translation unit,
// gft_step_sm.hip, under the instruction-level-parallelism scheduler: the
// same instructions in another order, 0.6% faster at 256^3 (1.0015 ->
// 0.9958 ms, 127 VGPRs, still 4 waves), where that scheduler makes the
// two-doublet kernels slower (they lose occupancy).  These declarations keep
// every other translation unit from instantiating them.
extern template __global__ void mom_step<float, 1, 4, true, 2, 4>(StepArgs<float>, Coef);
extern template __global__ void mom_step<float, 1, 4, false, 2, 4>(StepArgs<float>, Coef);

// One member of a batched ensemble: its step arguments and coefficients.
struct EnsMember {
    StepArgs<float> a;
    Coef c;
};

// One step of every member of an ensemble in one launch (MINW waves per SIMD:
// at 4 the own state is loaded after the exchange barriers, as in mom_step).
// The work list holds (member, x, y-tile, z-tile) per 256-site tile, members
// one after another; XCD k (hardware blocks b with b % 8 == k) takes the k-th
// eighth of the list, one tile per block in list order, so each XCD works on a
// compact region of a few members.  One block per tile (the grid is 8 x the
// longest range) rather than blocks striding through a range: the hardware
// starts blocks in order as earlier ones finish, which keeps the tiles an XCD
// holds at once contiguous in the list, where strided blocks drift apart and
// spread the x-planes in flight beyond what the XCD's L2 holds (measured: 20%
// fewer bytes fetched for 96^3 and 128^3 members, 3% faster on the mixed
// 64-member ensemble).  A block past its XCD's range exits.
// PL: each tile's plane is staged in LDS as in mom_step (the dynamic LDS holds
// the largest member tile).
template <int MINW, int PL>
__global__ void __launch_bounds__(256, MINW) ens_step(const EnsMember* __restrict__ members, const int4* __restrict__ work,
                                                      long long nwork) {
    const long long xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    const long long lo = xcd * nwork / 8, hi = (xcd + 1) * nwork / 8;
    {
        const long long w = lo + idx;
        if (w >= hi) return;
        const int4 t = work[w];
        // Member arguments copied before any store, as a single-lattice launch
        // holds them in its kernel arguments: the per-site arithmetic then
        // compiles to the same instructions as mom_step's.
        const StepArgs<float> a = members[t.x].a;
        const Coef c = members[t.x].c;
        const int tz = static_cast<int>(threadIdx.x) & (a.bz - 1), ty = static_cast<int>(threadIdx.x) >> a.bz_shift;
        const int z = t.w * a.bz + tz;
        const int y = t.z * a.by + ty;
        const bool valid = z < a.nz && y < a.ny;
        if constexpr (PL > 0) {
            extern __shared__ float4 gft_plane_lds[];
            char* lds = reinterpret_cast<char*>(gft_plane_lds);
            PlaneLds pl = plane_layout<1>(lds, a.bz, a.by);
            // Raised issue priority until the own-state loads are out, as in
            // mom_step (mom_block_shared drops it in every wave).
            if constexpr (PL == 2) __builtin_amdgcn_s_setprio(3);
            plane_fill<1>(a, pl, t.y, t.z * a.by, t.w * a.bz, a.bz, a.by);
            pl.r0 = (ty + 1) * pl.w + tz + 1;
            if constexpr (PL == 2)
                mom_block_shared<1, true, MINW != 4>(a, c, t.y, t.z * a.by, t.w * a.bz, ty, tz, valid, pl, lds);
            else if (valid) mom_site<float, 1, true, true>(a, c, t.y, y, z, pl);
        } else {
            if (valid) mom_site<float, 1, true>(a, c, t.y, y, z);
        }
    }
}

// ---- FP64 observables of the momentum form ------------------------------------

template <class R>
struct ObsArgs {
    using V4 = typename V4T<R>::type;
    const V4* phi[2];       // current
    const V4* phi_old[2];   // one step earlier (Gauss only)
    const V4* P[2];
    const R* y[3];
    const V4* q[3];
    const V4* q_old[3];     // links one step earlier (Gauss only)
    const V4* E[3];
    const uint2* lo;                 // residuals of compensated storage (null otherwise)
    const unsigned short* lo_P2;
    // Higgs-family field (gaugeNum 5 or 7; see MomFields): current, and the
    // family SU(2) one step earlier (Gauss only).
    const R* fy[3];
    const R* fe[3];
    const V4* fq[3];
    const V4* fq_old[3];
    const V4* fE[3];
    const unsigned* flo5;
    const uint2* flo7;
    double ref[2][4];                // reference doublets: phi = stored + ref (exact in FP64)
    int nx, ny, nz;
};

__device__ __forceinline__ long long nb_index(int nx, int ny, int nz, long long s, int a, int delta) {
    long long z = s % nz, y = (s / nz) % ny, x = s / ((long long)nz * ny);
    if (a == 0) x = (x + delta + nx) % nx;
    else if (a == 1) y = (y + delta + ny) % ny;
    else z = (z + delta + nz) % nz;
    return (x * ny + y) * nz + z;
}

template <class V4>
__device__ __forceinline__ void load_d(const V4& v, double* o) {
    o[0] = (double)v.x; o[1] = (double)v.y; o[2] = (double)v.z; o[3] = (double)v.w;
}

// Full-precision values of the compensated FP32 storage (hi + 4-bit residual;
// exact in FP64), or the stored values when there are no residuals.  Nibble m
// of the 16-bit group w belongs to value m.
template <class V4>
__device__ __forceinline__ void obs_add_res(const V4& hi, unsigned w, double* o) {
    const float h[4] = {(float)hi.x, (float)hi.y, (float)hi.z, (float)hi.w};
    for (int m = 0; m < 4; m++) o[m] += ldexp((double)res_nib(w, m), res_exp(h[m]) - 153);
}
// The doublet k at site s of buffer `f` (current or previous): stored + ref.
template <class R, class V4>
__device__ __forceinline__ void obs_phi_of(const ObsArgs<R>& a, const V4* f, int k, long long s, double* o) {
    load_d(f[s], o);
    for (int m = 0; m < 4; m++) o[m] += a.ref[k][m];
}
template <class R>
__device__ __forceinline__ void obs_phi(const ObsArgs<R>& a, int k, long long s, double* o) {
    obs_phi_of(a, a.phi[k], k, s, o);
}
template <class R>
__device__ __forceinline__ void obs_P(const ObsArgs<R>& a, int k, long long s, double* o) {
    const auto hi = a.P[k][s];
    load_d(hi, o);
    if (a.lo) obs_add_res(hi, k == 0 ? res_group(a.lo[s], 0) : (unsigned)a.lo_P2[s], o);
}
template <class R>
__device__ __forceinline__ void obs_E(const ObsArgs<R>& a, int i, long long s, double* o) {
    const auto hi = a.E[i][s];
    load_d(hi, o);
    if (a.lo) obs_add_res(hi, res_group(a.lo[s], 1 + i), o);
}

// The Higgs-family electric field at site s: E' (gaugeNum 5, o[0]) or E'^1..3
// (gaugeNum 7, o[1..3]), with its residuals.
template <class R, int G>
__device__ __forceinline__ void obs_fE(const ObsArgs<R>& a, int i, long long s, double* o) {
    if constexpr (G == 5) {
        const R hi = a.fe[i][s];
        o[0] = (double)hi;
        if (a.flo5) o[0] += ldexp((double)res_nib(a.flo5[s], i), res_exp((float)hi) - 153);
    } else {
        const auto hi = a.fE[i][s];
        load_d(hi, o);
        if (a.flo7) obs_add_res(hi, res_group(a.flo7[s], i), o);
        o[0] = 0.0;
    }
}
// The family part of the link of direction i at site s, in FP64.
template <class R, int G>
__device__ __forceinline__ FamLink<double> obs_fam(const ObsArgs<R>& a, const R* const* fy, const typename V4T<R>::type* const* fq,
                                                   int i, long long s) {
    FamLink<double> l{};
    if constexpr (G == 5) {
        const double y = (double)fy[i][s];
        l.c = cos(y);
        l.s = sin(y);
    } else if constexpr (G == 7) {
        double t[4];
        load_d(fq[i][s], t);
        l.v = Qt<double>{t[0], t[1], t[2], t[3]};
    }
    return l;
}

// 1 - sqrt(1 - x) without cancellation.
__device__ __forceinline__ double one_minus_sqrt1m(double x) { return x / (1.0 + sqrt(1.0 - x)); }

// 1 - P0 / |P| of a plaquette quaternion P without cancellation for P0 > 0.
__device__ __forceinline__ double one_minus_p0(const Qt<double>& P1) {
    const double n = sqrt(P1.w * P1.w + P1.x * P1.x + P1.y * P1.y + P1.z * P1.z);
    const double v2 = P1.x * P1.x + P1.y * P1.y + P1.z * P1.z;
    return P1.w > 0 ? v2 / (n * (n + P1.w)) : 1.0 - P1.w / n;
}

// Energy densities of the current state (potential, gradient, kinetic, magnetic,
// electric) with the reference's definitions (Analysers/Source/Energy.cpp:126-132;
// kinetic and electric from the last update, as the reference's two slots give).
// G 5 / 7: with the Higgs-family terms (WilsonLoops/Source/TwoHDM.cpp:496-552,
// 600-634): the family part of the covariant gradient, (1 - cos theta') / g''^2
// or (1 - P'0) / g''^2 per plaquette, (1 - cos) / g''^2 of the family electric field.
template <class R, int ND, int G = 4>
__global__ void __launch_bounds__(256) mom_energy(ObsArgs<R> a, Coef c, double* __restrict__ partial) {
    __shared__ double lds[4 * 5];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double d[5] = {0, 0, 0, 0, 0};
    if (s < c.nsites && !site_frozen(c, s)) {
        double phi[8];
        for (int k = 0; k < ND; k++) obs_phi(a, k, s, phi + 4 * k);
        d[0] = potential_energy(c, phi);
        double ge = 0.0;
        Qt<double> U[3];
        double yv[3];
        for (int i = 0; i < 3; i++) {
            double t[4];
            load_d(a.q[i][s], t);
            U[i] = Qt<double>{t[0], t[1], t[2], t[3]};
            yv[i] = (double)a.y[i][s];
        }
        for (int i = 0; i < 3; i++) {
            const long long sp = nb_index(a.nx, a.ny, a.nz, s, i, 1);
            if constexpr (G != 4) {
                double fp[8], tx[8];
                for (int k = 0; k < 2; k++) obs_phi(a, k, sp, fp + 4 * k);
                transport2<double, G>(U[i], cos(yv[i]), sin(yv[i]), obs_fam<R, G>(a, a.fy, a.fq, i, s), false, phi, tx);
                for (int m = 0; m < 8; m++) {
                    const double dd = (fp[m] - tx[m]) * c.est[i];
                    ge += dd * dd;
                }
                continue;
            }
            for (int k = 0; k < ND; k++) {
                double fp[4], tx[4];
                obs_phi(a, k, sp, fp);
                transport(U[i], cos(yv[i]), sin(yv[i]), false, phi + 4 * k, tx);
                for (int m = 0; m < 4; m++) {
                    const double dd = (fp[m] - tx[m]) * c.est[i];
                    ge += dd * dd;
                }
            }
        }
        d[1] = 0.5 * ge;
        double ke = 0.0;
        for (int k = 0; k < ND; k++) {
            double p[4];
            obs_P(a, k, s, p);
            for (int m = 0; m < 4; m++) ke += p[m] * p[m];
        }
        d[2] = 0.5 * ke * c.inv_sqr_dt;
        double me = 0.0;
        for (int i = 0; i < 3; i++) {
            for (int j = i + 1; j < 3; j++) {
                const long long spi = nb_index(a.nx, a.ny, a.nz, s, i, 1);
                const long long spj = nb_index(a.nx, a.ny, a.nz, s, j, 1);
                double t1[4], t2[4];
                load_d(a.q[j][spi], t1);
                load_d(a.q[i][spj], t2);
                const Qt<double> A = qmul(qconj(Qt<double>{t1[0], t1[1], t1[2], t1[3]}),
                                          Qt<double>{t2[0], t2[1], t2[2], t2[3]});
                const Qt<double> P1 = qmul(A, qmul(U[j], qconj(U[i])));
                const double n = sqrt(P1.w * P1.w + P1.x * P1.x + P1.y * P1.y + P1.z * P1.z);
                const double v2 = P1.x * P1.x + P1.y * P1.y + P1.z * P1.z;
                // 1 - P0/|P| = (|P| - P0)/|P| = |v|^2 / (|P| (|P| + P0)) for P0 > 0.
                const double one_m_p0 = P1.w > 0 ? v2 / (n * (n + P1.w)) : 1.0 - P1.w / n;
                const double th = -(double)a.y[j][spi] + (double)a.y[i][spj] + yv[j] - yv[i];
                const double sh = sin(0.5 * th);
                // Both orientations (i,j) and (j,i) of the plane contribute equally.
                me += 2.0 * 2.0 * c.inv_sp[i] * c.inv_sp[j] * (c.inv_g_sqr * one_m_p0 + c.inv_gp_sqr * 2.0 * sh * sh);
                if constexpr (G == 5) {
                    const double tf = -(double)a.fy[j][spi] + (double)a.fy[i][spj] + (double)a.fy[j][s] -
                                      (double)a.fy[i][s];
                    const double shf = sin(0.5 * tf);
                    me += 2.0 * 2.0 * c.inv_sp[i] * c.inv_sp[j] * c.inv_gpp_sqr * 2.0 * shf * shf;
                } else if constexpr (G == 7) {
                    double v1[4], v2[4], vi[4], vj[4];
                    load_d(a.fq[j][spi], v1);
                    load_d(a.fq[i][spj], v2);
                    load_d(a.fq[i][s], vi);
                    load_d(a.fq[j][s], vj);
                    const Qt<double> Af = qmul(qconj(Qt<double>{v1[0], v1[1], v1[2], v1[3]}),
                                               Qt<double>{v2[0], v2[1], v2[2], v2[3]});
                    const Qt<double> Pf = qmul(Af, qmul(Qt<double>{vj[0], vj[1], vj[2], vj[3]},
                                                        qconj(Qt<double>{vi[0], vi[1], vi[2], vi[3]})));
                    me += 2.0 * 2.0 * c.inv_sp[i] * c.inv_sp[j] * c.inv_gpp_sqr * one_minus_p0(Pf);
                }
            }
        }
        d[3] = me;
        double ee = 0.0;
        for (int i = 0; i < 3; i++) {
            double e[4];
            obs_E(a, i, s, e);
            const double ea2 = e[1] * e[1] + e[2] * e[2] + e[3] * e[3];
            ee += 4.0 * c.inv_sp[i] * (c.inv_g_sqr * one_minus_sqrt1m(ea2) + c.inv_gp_sqr * one_minus_sqrt1m(e[0] * e[0]));
            if constexpr (G != 4) {
                double f[4] = {0, 0, 0, 0};
                obs_fE<R, G>(a, i, s, f);
                const double f2 = G == 5 ? f[0] * f[0] : f[1] * f[1] + f[2] * f[2] + f[3] * f[3];
                ee += 4.0 * c.inv_sp[i] * c.inv_gpp_sqr * one_minus_sqrt1m(f2);
            }
        }
        d[4] = c.inv_sqr_dt * ee;
    }
    block_reduce_store<5, 0>(d, nullptr, lds, partial, blockIdx.x);
}

// Gauss law after a step, from the stored E (the momentum form's electric
// field): G^Y = sum_i [E^Y_i(x) - E^Y_i(x-i)]/dx_i^2 - g'^2 rho^Y,
// G^a = sum_i [vec(U_i(x,t)^dag E_i(x) U_i(x,t)) - E_i(x-i)]/dx_i^2 - g^2 rho^a,
// rho = 1/4 bilinears(phi(t), phi(t+dt)) evaluated as 1/4 bilinears(phi(t), P).  This equals the reference's
// link-difference form (WilsonLoops/Source/StandardModel.cpp:274-322) in exact
// arithmetic, since U(t+dt) U(t)^dag = (sqrt(1-|E|^2), E).
// G 5 / 7: G generators, the Higgs-family ones from the family electric field
// and charges (WilsonLoops/Source/TwoHDM.cpp:352-386, Gradients/Source/
// TwoHDM.cpp:299-319) in the same form as the U(1) (G = 5) or SU(2) (G = 7) ones.
template <class R, int ND, int G = 4>
__global__ void __launch_bounds__(256) mom_gauss(ObsArgs<R> a, Coef c, double* __restrict__ partial) {
    __shared__ double lds[4 * 3 * G];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double sums[2 * G], maxs[G];
    for (int k = 0; k < G; k++) sums[k] = sums[G + k] = maxs[k] = 0.0;
    if (s < c.nsites && !site_frozen(c, s)) {
        // rho = 1/4 B(phi(t), phi(t+dt)) = 1/4 B(phi(t), P): the bilinears are
        // antisymmetric, so B(phi, phi) = 0.  Using the stored P avoids the
        // rounding of phi(t+dt) = phi(t) + P, which is large next to the charge.
        double rho[G];
        for (int k = 0; k < G; k++) rho[k] = 0.0;
        double t8[8], n8[8];
        for (int k = 0; k < ND; k++) {
            obs_P(a, k, s, n8 + 4 * k);
            obs_phi_of(a, a.phi_old[k], k, s, t8 + 4 * k);   // phi(t) as stored in the previous buffer
            bilinears_add(t8 + 4 * k, n8 + 4 * k, rho);
        }
        if constexpr (G != 4) family_bilinears<double, G>(t8, n8, rho + 4);
        double div[G], mag[G];
        for (int k = 0; k < G; k++) div[k] = mag[k] = 0.0;
        for (int i = 0; i < 3; i++) {
            if constexpr (G != 4) {
                const long long smf = nb_index(a.nx, a.ny, a.nz, s, i, -1);
                double f[4] = {0, 0, 0, 0}, fm[4] = {0, 0, 0, 0};
                obs_fE<R, G>(a, i, s, f);
                obs_fE<R, G>(a, i, smf, fm);
                if constexpr (G == 5) {
                    div[4] += (f[0] - fm[0]) * c.inv_sp[i];
                    mag[4] += (fabs(f[0]) + fabs(fm[0])) * c.inv_sp[i];
                } else {
                    double v[4];
                    load_d(a.fq_old[i][s], v);
                    const Qt<double> Vt{v[0], v[1], v[2], v[3]};
                    const Qt<double> rot = qmul(qconj(Vt), qmul(Qt<double>{0.0, f[1], f[2], f[3]}, Vt));
                    const double r3[3] = {rot.x, rot.y, rot.z};
                    for (int m = 0; m < 3; m++) {
                        div[4 + m] += (r3[m] - fm[1 + m]) * c.inv_sp[i];
                        mag[4 + m] += (fabs(r3[m]) + fabs(fm[1 + m])) * c.inv_sp[i];
                    }
                }
            }
        }
        for (int i = 0; i < 3; i++) {
            const long long sm = nb_index(a.nx, a.ny, a.nz, s, i, -1);
            double e[4], em[4], u[4];
            obs_E(a, i, s, e);
            obs_E(a, i, sm, em);
            load_d(a.q_old[i][s], u);
            const Qt<double> Ut{u[0], u[1], u[2], u[3]};
            const Qt<double> Ev{0.0, e[1], e[2], e[3]};
            const Qt<double> rot = qmul(qconj(Ut), qmul(Ev, Ut));
            div[0] += (e[0] - em[0]) * c.inv_sp[i];
            div[1] += (rot.x - em[1]) * c.inv_sp[i];
            div[2] += (rot.y - em[2]) * c.inv_sp[i];
            div[3] += (rot.z - em[3]) * c.inv_sp[i];
            mag[0] += (fabs(e[0]) + fabs(em[0])) * c.inv_sp[i];
            mag[1] += (fabs(rot.x) + fabs(em[1])) * c.inv_sp[i];
            mag[2] += (fabs(rot.y) + fabs(em[2])) * c.inv_sp[i];
            mag[3] += (fabs(rot.z) + fabs(em[3])) * c.inv_sp[i];
        }
        for (int k = 0; k < G; k++) {
            const double g = div[k] - c.gsq[k] * rho[k];
            sums[k] = fabs(g);
            sums[G + k] = mag[k] + fabs(c.gsq[k] * rho[k]);
            maxs[k] = fabs(g);
        }
    }
    block_reduce_store<2 * G, G>(sums, maxs, lds, partial, blockIdx.x);
}

// Wilson loops: mean 1/2 tr and mean cos of the 1x1 and 2x2 loops over the three
// planes at every site, and <|phi|^2>.  The SU(2) trace is the c0 of the
// product of the loop's links; the U(1) loop is the cosine of the summed phases.
template <class R, int ND>
__global__ void __launch_bounds__(256) mom_loops(ObsArgs<R> a, Coef c, double* __restrict__ partial) {
    __shared__ double lds[4 * 5];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double d[5] = {0, 0, 0, 0, 0};
    if (s < c.nsites) {
        for (int k = 0; k < ND; k++) {
            double f[4];
            obs_phi(a, k, s, f);
            d[4] += f[0] * f[0] + f[1] * f[1] + f[2] * f[2] + f[3] * f[3];
        }
        for (int i = 0; i < 3; i++) {
            for (int j = i + 1; j < 3; j++) {
                for (int L = 1; L <= 2; L++) {
                    // Loop x -> x+L i -> x+L i+L j -> x+L j -> x as a product of
                    // transporters (right to left): U_i's along i, then U_j's, then back.
                    Qt<double> W{1.0, 0.0, 0.0, 0.0};
                    double ph = 0.0;
                    long long p = s;
                    for (int st = 0; st < L; st++) {
                        double t[4];
                        load_d(a.q[i][p], t);
                        W = qmul(Qt<double>{t[0], t[1], t[2], t[3]}, W);
                        ph += (double)a.y[i][p];
                        p = nb_index(a.nx, a.ny, a.nz, p, i, 1);
                    }
                    for (int st = 0; st < L; st++) {
                        double t[4];
                        load_d(a.q[j][p], t);
                        W = qmul(Qt<double>{t[0], t[1], t[2], t[3]}, W);
                        ph += (double)a.y[j][p];
                        p = nb_index(a.nx, a.ny, a.nz, p, j, 1);
                    }
                    for (int st = 0; st < L; st++) {
                        p = nb_index(a.nx, a.ny, a.nz, p, i, -1);
                        double t[4];
                        load_d(a.q[i][p], t);
                        W = qmul(qconj(Qt<double>{t[0], t[1], t[2], t[3]}), W);
                        ph -= (double)a.y[i][p];
                    }
                    for (int st = 0; st < L; st++) {
                        p = nb_index(a.nx, a.ny, a.nz, p, j, -1);
                        double t[4];
                        load_d(a.q[j][p], t);
                        W = qmul(qconj(Qt<double>{t[0], t[1], t[2], t[3]}), W);
                        ph -= (double)a.y[j][p];
                    }
                    const double n = sqrt(W.w * W.w + W.x * W.x + W.y * W.y + W.z * W.z);
                    d[(L - 1) * 2] += W.w / n;
                    d[(L - 1) * 2 + 1] += cos(ph);
                }
            }
        }
    }
    block_reduce_store<5, 0>(d, nullptr, lds, partial, blockIdx.x);
}

// ---- host side ----------------------------------------------------------------
// (Left out of a translation unit that defines GFT_KERNELS_ONLY to compile
// selected kernels alone: the launches below would make it emit every kernel.)
#ifndef GFT_KERNELS_ONLY

template <class R>
inline size_t carve(MomFields<R>* f, char* base, long long N, bool cmp, int G) {
    using V4 = typename V4T<R>::type;
    size_t off = 0;
    auto take = [&](size_t elems, size_t esz) {
        const size_t at = off;
        off += (elems * esz + 255) & ~size_t(255);
        return base ? base + at : nullptr;
    };
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < 2; k++) { char* p = take(N, sizeof(V4)); if (f) f->phi[b][k] = reinterpret_cast<V4*>(p); }
    for (int k = 0; k < 2; k++) { char* p = take(N, sizeof(V4)); if (f) f->P[k] = reinterpret_cast<V4*>(p); }
    for (int b = 0; b < 2; b++)
        for (int i = 0; i < 3; i++) { char* p = take(N, sizeof(R)); if (f) f->y[b][i] = reinterpret_cast<R*>(p); }
    for (int b = 0; b < 2; b++)
        for (int i = 0; i < 3; i++) { char* p = take(N, sizeof(V4)); if (f) f->q[b][i] = reinterpret_cast<V4*>(p); }
    for (int i = 0; i < 3; i++) { char* p = take(N, sizeof(V4)); if (f) f->E[i] = reinterpret_cast<V4*>(p); }
    { char* p = cmp ? take(N, sizeof(uint2)) : nullptr; if (f) f->lo = reinterpret_cast<uint2*>(p); }
    { char* p = cmp ? take(N, sizeof(unsigned short)) : nullptr; if (f) f->lo_P2 = reinterpret_cast<unsigned short*>(p); }
    // The Higgs-family field, only for gaugeNum 5 or 7.
    for (int b = 0; b < 2; b++)
        for (int i = 0; i < 3; i++) {
            char* p = G == 5 ? take(N, sizeof(R)) : nullptr;
            if (f) f->fy[b][i] = reinterpret_cast<R*>(p);
            char* pq = G == 7 ? take(N, sizeof(V4)) : nullptr;
            if (f) f->fq[b][i] = reinterpret_cast<V4*>(pq);
        }
    for (int i = 0; i < 3; i++) {
        char* p = G == 5 ? take(N, sizeof(R)) : nullptr;
        if (f) f->fe[i] = reinterpret_cast<R*>(p);
        char* pe = G == 7 ? take(N, sizeof(V4)) : nullptr;
        if (f) f->fE[i] = reinterpret_cast<V4*>(pe);
    }
    { char* p = cmp && G == 5 ? take(N, sizeof(unsigned)) : nullptr; if (f) f->flo5 = reinterpret_cast<unsigned*>(p); }
    { char* p = cmp && G == 7 ? take(N, sizeof(uint2)) : nullptr; if (f) f->flo7 = reinterpret_cast<uint2*>(p); }
    return off;
}

inline int momentum_alloc(MomentumState& m, const Coef& c, bool f64, bool cmp) {
    m.f64 = f64;
    m.cmp = cmp && !f64;
    const long long N = c.nsites;
    // The second-doublet arrays are carved for the SM too (unused there).
    const int G = c.G == 5 || c.G == 7 ? c.G : 4;
    m.bytes = f64 ? carve<double>(nullptr, nullptr, N, false, G) : carve<float>(nullptr, nullptr, N, m.cmp, G);
    if (hipMalloc(&m.base, m.bytes) != hipSuccess) { m.base = nullptr; return GFT_ENOMEM; }
    if (f64) carve<double>(&m.fd, static_cast<char*>(m.base), N, false, G);
    else carve<float>(&m.ff, static_cast<char*>(m.base), N, m.cmp, G);
    m.cur = 0;
    m.prev_valid = 0;
    return GFT_OK;
}

inline void momentum_free(MomentumState& m) {
    if (m.base) (void)hipFree(m.base);
    if (m.tiles) (void)hipFree(m.tiles);
    m.base = nullptr;
    m.tiles = nullptr;
}

// Host-built table mapping hardware block -> (member, x, y-tile, z-tile).
inline std::vector<int4> build_tiles(int member, int nx, int ny, int nz, int bz, int by, int order) {
    const int nbz = (nz + bz - 1) / bz, nby = (ny + by - 1) / by;
    const long long nb = (long long)nbz * nby * nx;
    std::vector<int4> t(nb);
    for (long long b = 0; b < nb; b++) {
        int x, ty, tz;
        decode_block(b, nb, order, nx, nby, nbz, &x, &ty, &tz);
        t[b] = make_int4(member, x, ty, tz);
    }
    return t;
}

// Tiles of one member in band-major order: 16 bands of y-tiles, and inside a
// band x-planes in order with the band's (y, z) tiles inside a plane.  Split
// into 8 contiguous ranges (the ensemble kernel's XCD assignment), a large
// member gives each XCD two y-bands to sweep along x one after the other, as
// block order 3 does for a single lattice, so the x +- 1 planes a tile reads
// stay in that XCD's L2.  A member whose tiles span its z-rows in one or two
// tiles and whose x-plane (148 stored bytes read per site) fits in 2.6 MB is
// one band: its XCD sweeps whole planes, so the y-neighbours of a tile row are
// read from the plane in flight rather than from a band swept long before
// (ns per site-update, homogeneous ensembles: 32^3 0.0604 -> 0.0586, 64^3
// 0.0603 -> 0.0588, 128^3 0.0654 -> 0.0640; the 48^3 and 96^3 members, three
// z-tiles per row, measured no gain and keep their bands; the mixed
// 64-member workload 0.906 -> 0.896 ms).
inline std::vector<int4> build_tiles_banded(int member, int nx, int ny, int nz, int bz, int by) {
    const int nbz = (nz + bz - 1) / bz, nby = (ny + by - 1) / by;
    const bool one_band = nbz <= 2 && static_cast<double>(ny) * nz * 148.0 <= 2.6e6;
    const int br = one_band ? nby : (nby + 15) / 16;
    std::vector<int4> t;
    t.reserve((size_t)nbz * nby * nx);
    for (int b0 = 0; b0 < nby; b0 += br)
        for (int x = 0; x < nx; x++)
            for (int ty = b0; ty < std::min(nby, b0 + br); ty++)
                for (int tz = 0; tz < nbz; tz++) t.push_back(make_int4(member, x, ty, tz));
    return t;
}

inline int ensure_tiles(MomentumState& m, const Coef& c, hipStream_t st, std::string& msg) {
    if (m.tiles && m.tiles_key[0] == m.bz && m.tiles_key[1] == m.by && m.tiles_key[2] == m.order) return GFT_OK;
    const std::vector<int4> t = build_tiles(0, c.nx, c.ny, c.nz, m.bz, m.by, m.order);
    if (m.tiles) (void)hipFree(m.tiles);
    m.tiles = nullptr;
    if (hipMalloc(&m.tiles, t.size() * sizeof(int4)) != hipSuccess) { msg = "tile table allocation failed"; return GFT_ENOMEM; }
    if (hipMemcpyAsync(m.tiles, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice, st) != hipSuccess
        || hipStreamSynchronize(st) != hipSuccess) {
        msg = "tile table upload failed";
        return GFT_EHIP;
    }
    m.tiles_key[0] = m.bz; m.tiles_key[1] = m.by; m.tiles_key[2] = m.order;
    return GFT_OK;
}

inline long long mom_nblocks(const MomentumState& m, const Coef& c) {
    const long long nbz = (c.nz + m.bz - 1) / m.bz, nby = (c.ny + m.by - 1) / m.by;
    return nbz * nby * c.nx;
}

inline double wrap_host(double y) {
    const double twopi = 6.283185307179586476925286766559;
    y = std::fmod(y, twopi);
    if (y >= 3.1415926535897932384626433832795) y -= twopi;
    if (y < -3.1415926535897932384626433832795) y += twopi;
    return y;
}

// 4-bit residuals on the host, as the kernels encode them: q = r 2^(153 - e)
// rounded to nearest even (|q| <= 4; clamped to [-8, 7] as a guard), e the
// biased exponent of the FP32 hi; value q 2^(e - 153); nibble m of a word holds
// value m.
inline int res_exp_host(float hi) {
    unsigned u;
    std::memcpy(&u, &hi, 4);
    return static_cast<int>((u >> 23) & 0xffu);
}
inline unsigned res_q_host(double r, float hi) {
    const double q = std::nearbyint(std::ldexp(r, 153 - res_exp_host(hi)));
    return static_cast<unsigned>(static_cast<int>(std::min(7.0, std::max(-8.0, q)))) & 0xfu;
}
inline double res_val_host(unsigned w, int m, float hi) {
    int q = static_cast<int>((w >> (4 * m)) & 0xfu);
    if (q >= 8) q -= 16;
    return std::ldexp(static_cast<double>(q), res_exp_host(hi) - 153);
}

// Two-slot (reference layout) -> momentum form, on the host in FP64.  The FP32
// forms store each doublet as its difference from a reference doublet (the
// mean of the current slot, rounded to FP32); with compensated storage P and E
// are split into their FP32 hi and 4-bit residuals.
template <class R, class Src>
int momentum_import_t(MomentumState& m, MomFields<R>& f, const Coef& c, const Src* S, const Src* V,
                      unsigned time_iter, hipStream_t st, std::string& err) {
    const long long N = c.nsites;
    const int Ns = c.ns, nd = c.nd;
    const unsigned now = (time_iter + 1) % 2, past = !now;
    std::vector<R> phi(4 * N * 2, R(0)), P(4 * N * 2, R(0)), y(3 * N), q(12 * N), E(12 * N);
    const bool cmp = f.lo != nullptr;
    std::vector<uint2> lo(cmp ? N : 0, make_uint2(0u, 0u));
    std::vector<unsigned short> lo_P2(cmp ? N : 0, 0);
    // Link reals per direction: [y, c0..c3, Higgs-family part (y' or c0'..c3')].
    const int G = c.G == 5 || c.G == 7 ? c.G : 4, VD = 5 + (G == 5 ? 1 : G == 7 ? 4 : 0);
    std::vector<R> fy(G == 5 ? 3 * N : 0), fe(G == 5 ? 3 * N : 0), fq(G == 7 ? 12 * N : 0), fE(G == 7 ? 12 * N : 0);
    std::vector<unsigned> flo5(cmp && G == 5 ? N : 0, 0u);
    std::vector<uint2> flo7(cmp && G == 7 ? N : 0, make_uint2(0u, 0u));
    for (int k = 0; k < 2; k++)
        for (int m4 = 0; m4 < 4; m4++) m.ref[k][m4] = 0.0;
    if (sizeof(R) == 4) {
        for (int k = 0; k < nd; k++)
            for (int m4 = 0; m4 < 4; m4++) {
                double sum = 0.0;
                for (long long s = 0; s < N; s++) sum += (double)S[(now * N + s) * Ns + 4 * k + m4];
                m.ref[k][m4] = (double)(float)(sum / (double)N);
            }
    }
    for (long long s = 0; s < N; s++) {
        for (int k = 0; k < nd; k++) {
            unsigned lp = 0;
            for (int m4 = 0; m4 < 4; m4++) {
                const double a = (double)S[(now * N + s) * Ns + 4 * k + m4];
                const double b = (double)S[(past * N + s) * Ns + 4 * k + m4];
                phi[(k * N + s) * 4 + m4] = (R)(a - m.ref[k][m4]);
                P[(k * N + s) * 4 + m4] = (R)(a - b);
                lp |= res_q_host((a - b) - (double)(float)(a - b), (float)(a - b)) << (4 * m4);
            }
            if (cmp) {
                if (k == 0) lo[s].x = lp;
                else lo_P2[s] = static_cast<unsigned short>(lp);
            }
        }
        for (int i = 0; i < 3; i++) {
            const Src* ln = V + (now * N + s) * (3 * VD) + VD * i;
            const Src* lp = V + (past * N + s) * (3 * VD) + VD * i;
            if (G == 5) {   // E' = sin(y'(t) - y'(t - dt))
                const double ev = std::sin((double)ln[5] - (double)lp[5]);
                fy[i * N + s] = (R)wrap_host((double)ln[5]);
                fe[i * N + s] = (R)ev;
                if (cmp) flo5[s] |= res_q_host(ev - (double)(float)ev, (float)ev) << (4 * i);
            } else if (G == 7) {   // E' = vec(V(t) V(t - dt)^dag)
                const double vn[4] = {(double)ln[5], (double)ln[6], (double)ln[7], (double)ln[8]};
                const double vp[4] = {(double)lp[5], -(double)lp[6], -(double)lp[7], -(double)lp[8]};
                double r[4];
                qmul_full(vn, vp, r);
                r[0] = 0.0;
                unsigned g = 0;
                for (int k = 0; k < 4; k++) {
                    fq[(i * N + s) * 4 + k] = (R)vn[k];
                    fE[(i * N + s) * 4 + k] = (R)r[k];
                    if (k) g |= res_q_host(r[k] - (double)(float)r[k], (float)r[k]) << (4 * k);
                }
                if (cmp) {
                    if (i == 0) flo7[s].x = g;
                    else if (i == 1) flo7[s].x |= g << 16;
                    else flo7[s].y = g;
                }
            }
            y[i * N + s] = (R)wrap_host((double)ln[0]);
            const double un[4] = {(double)ln[1], (double)ln[2], (double)ln[3], (double)ln[4]};
            const double up[4] = {(double)lp[1], -(double)lp[2], -(double)lp[3], -(double)lp[4]};
            for (int k = 0; k < 4; k++) q[(i * N + s) * 4 + k] = (R)un[k];
            double r[4];
            qmul_full(un, up, r);
            const double ev[4] = {std::sin((double)ln[0] - (double)lp[0]), r[1], r[2], r[3]};
            unsigned le = 0;
            for (int k = 0; k < 4; k++) {
                E[(i * N + s) * 4 + k] = (R)ev[k];
                le |= res_q_host(ev[k] - (double)(float)ev[k], (float)ev[k]) << (4 * k);
            }
            if (cmp) {
                if (i == 0) lo[s].x |= le << 16;
                else if (i == 1) lo[s].y = le;
                else lo[s].y |= le << 16;
            }
        }
    }
    m.cur = 0;
    m.prev_valid = 0;
    m.epoch++;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    bool ok = true;
    for (int k = 0; k < nd; k++) {
        ok &= cp(f.phi[0][k], &phi[k * N * 4], N * 4 * sizeof(R));
        ok &= cp(f.P[k], &P[k * N * 4], N * 4 * sizeof(R));
    }
    for (int i = 0; i < 3; i++) {
        ok &= cp(f.y[0][i], &y[i * N], N * sizeof(R));
        ok &= cp(f.q[0][i], &q[i * N * 4], N * 4 * sizeof(R));
        ok &= cp(f.E[i], &E[i * N * 4], N * 4 * sizeof(R));
    }
    if (cmp) {
        ok &= cp(f.lo, lo.data(), N * sizeof(uint2));
        ok &= cp(f.lo_P2, lo_P2.data(), N * sizeof(unsigned short));
    }
    for (int i = 0; i < 3; i++) {
        if (G == 5) {
            ok &= cp(f.fy[0][i], &fy[i * N], N * sizeof(R));
            ok &= cp(f.fe[i], &fe[i * N], N * sizeof(R));
        } else if (G == 7) {
            ok &= cp(f.fq[0][i], &fq[i * N * 4], N * 4 * sizeof(R));
            ok &= cp(f.fE[i], &fE[i * N * 4], N * 4 * sizeof(R));
        }
    }
    if (cmp && G == 5) ok &= cp(f.flo5, flo5.data(), N * sizeof(unsigned));
    if (cmp && G == 7) ok &= cp(f.flo7, flo7.data(), N * sizeof(uint2));
    ok &= hipStreamSynchronize(st) == hipSuccess;
    if (!ok) { err = "momentum import: copy failed"; return GFT_EHIP; }
    return GFT_OK;
}

// Momentum form -> two-slot: the current slot gets (phi, y, q); the other slot
// gets phi - P, y - asin(E^Y), (sqrt(1-|E|^2), E)^dag q.
template <class R, class Dst>
int momentum_export_t(MomentumState& m, MomFields<R>& f, const Coef& c, Dst* S, Dst* V, unsigned time_iter,
                      hipStream_t st, std::string& err) {
    const long long N = c.nsites;
    const int Ns = c.ns, nd = c.nd;
    const unsigned now = (time_iter + 1) % 2, past = !now;
    std::vector<R> phi(4 * N * 2), P(4 * N * 2), y(3 * N), q(12 * N), E(12 * N);
    const bool cmp = f.lo != nullptr;
    std::vector<uint2> lo(cmp ? N : 0);
    std::vector<unsigned short> lo_P2(cmp ? N : 0);
    const int G = c.G == 5 || c.G == 7 ? c.G : 4, VD = 5 + (G == 5 ? 1 : G == 7 ? 4 : 0);
    std::vector<R> fy(G == 5 ? 3 * N : 0), fe(G == 5 ? 3 * N : 0), fq(G == 7 ? 12 * N : 0), fE(G == 7 ? 12 * N : 0);
    std::vector<unsigned> flo5(cmp && G == 5 ? N : 0);
    std::vector<uint2> flo7(cmp && G == 7 ? N : 0);
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
    };
    bool ok = true;
    for (int i = 0; i < 3; i++) {
        if (G == 5) {
            ok &= cp(&fy[i * N], f.fy[m.cur][i], N * sizeof(R));
            ok &= cp(&fe[i * N], f.fe[i], N * sizeof(R));
        } else if (G == 7) {
            ok &= cp(&fq[i * N * 4], f.fq[m.cur][i], N * 4 * sizeof(R));
            ok &= cp(&fE[i * N * 4], f.fE[i], N * 4 * sizeof(R));
        }
    }
    if (cmp && G == 5) ok &= cp(flo5.data(), f.flo5, N * sizeof(unsigned));
    if (cmp && G == 7) ok &= cp(flo7.data(), f.flo7, N * sizeof(uint2));
    for (int k = 0; k < nd; k++) {
        ok &= cp(&phi[k * N * 4], f.phi[m.cur][k], N * 4 * sizeof(R));
        ok &= cp(&P[k * N * 4], f.P[k], N * 4 * sizeof(R));
    }
    for (int i = 0; i < 3; i++) {
        ok &= cp(&y[i * N], f.y[m.cur][i], N * sizeof(R));
        ok &= cp(&q[i * N * 4], f.q[m.cur][i], N * 4 * sizeof(R));
        ok &= cp(&E[i * N * 4], f.E[i], N * 4 * sizeof(R));
    }
    if (cmp) {
        ok &= cp(lo.data(), f.lo, N * sizeof(uint2));
        ok &= cp(lo_P2.data(), f.lo_P2, N * sizeof(unsigned short));
    }
    ok &= hipStreamSynchronize(st) == hipSuccess;
    if (!ok) { err = "momentum export: copy failed"; return GFT_EHIP; }
    auto group = [&](long long s, int g) -> unsigned {
        const unsigned w = g < 2 ? lo[s].x : lo[s].y;
        return (w >> (16 * (g & 1))) & 0xffffu;
    };
    for (long long s = 0; s < N; s++) {
        for (int k = 0; k < nd; k++)
            for (int m4 = 0; m4 < 4; m4++) {
                const double a = (double)phi[(k * N + s) * 4 + m4] + m.ref[k][m4];
                double pp = (double)P[(k * N + s) * 4 + m4];
                if (cmp) pp += res_val_host(k == 0 ? group(s, 0) : lo_P2[s], m4, (float)P[(k * N + s) * 4 + m4]);
                S[(now * N + s) * Ns + 4 * k + m4] = (Dst)a;
                S[(past * N + s) * Ns + 4 * k + m4] = (Dst)(a - pp);
            }
        for (int i = 0; i < 3; i++) {
            Dst* ln = V + (now * N + s) * (3 * VD) + VD * i;
            Dst* lp = V + (past * N + s) * (3 * VD) + VD * i;
            if (G == 5) {   // y'(t - dt) = y'(t) - asin(E')
                double e = (double)fe[i * N + s];
                if (cmp) e += res_val_host(flo5[s], i, (float)fe[i * N + s]);
                const double yf = (double)fy[i * N + s];
                ln[5] = (Dst)yf;
                lp[5] = (Dst)(yf - std::asin(e));
            } else if (G == 7) {   // V(t - dt) = (sqrt(1 - |E'|^2), E')^dag V(t)
                double e[4];
                const unsigned g = !cmp ? 0u : i == 0 ? flo7[s].x & 0xffffu : i == 1 ? flo7[s].x >> 16
                                                                                   : flo7[s].y & 0xffffu;
                for (int k = 0; k < 4; k++) {
                    e[k] = (double)fE[(i * N + s) * 4 + k];
                    if (cmp && k) e[k] += res_val_host(g, k, (float)fE[(i * N + s) * 4 + k]);
                }
                const double vn[4] = {(double)fq[(i * N + s) * 4], (double)fq[(i * N + s) * 4 + 1],
                                      (double)fq[(i * N + s) * 4 + 2], (double)fq[(i * N + s) * 4 + 3]};
                const double ud[4] = {std::sqrt(std::max(0.0, 1.0 - e[1] * e[1] - e[2] * e[2] - e[3] * e[3])), -e[1],
                                      -e[2], -e[3]};
                double vp[4];
                qmul_full(ud, vn, vp);
                for (int k = 0; k < 4; k++) {
                    ln[5 + k] = (Dst)vn[k];
                    lp[5 + k] = (Dst)vp[k];
                }
            }
            const double yy = (double)y[i * N + s];
            double e[4] = {(double)E[(i * N + s) * 4], (double)E[(i * N + s) * 4 + 1],
                           (double)E[(i * N + s) * 4 + 2], (double)E[(i * N + s) * 4 + 3]};
            if (cmp)
                for (int k = 0; k < 4; k++) e[k] += res_val_host(group(s, 1 + i), k, (float)E[(i * N + s) * 4 + k]);
            const double un[4] = {(double)q[(i * N + s) * 4], (double)q[(i * N + s) * 4 + 1],
                                  (double)q[(i * N + s) * 4 + 2], (double)q[(i * N + s) * 4 + 3]};
            const double ud[4] = {std::sqrt(std::max(0.0, 1.0 - e[1] * e[1] - e[2] * e[2] - e[3] * e[3])), -e[1], -e[2],
                                  -e[3]};
            double up[4];
            qmul_full(ud, un, up);
            ln[0] = (Dst)yy;
            lp[0] = (Dst)(yy - std::asin(e[0]));
            for (int k = 0; k < 4; k++) {
                ln[1 + k] = (Dst)un[k];
                lp[1 + k] = (Dst)up[k];
            }
        }
    }
    return GFT_OK;
}

template <class Src>
inline int momentum_import(MomentumState& m, const Coef& c, const Src* S, const Src* V, unsigned it, hipStream_t st,
                           std::string& err) {
    return m.f64 ? momentum_import_t<double>(m, m.fd, c, S, V, it, st, err)
                 : momentum_import_t<float>(m, m.ff, c, S, V, it, st, err);
}

template <class Dst>
inline int momentum_export(MomentumState& m, const Coef& c, Dst* S, Dst* V, unsigned it, hipStream_t st,
                           std::string& err) {
    return m.f64 ? momentum_export_t<double>(m, m.fd, c, S, V, it, st, err)
                 : momentum_export_t<float>(m, m.ff, c, S, V, it, st, err);
}

template <class R>
StepArgs<R> make_step_args(MomentumState& m, MomFields<R>& f, const Coef& c, double damping_step, unsigned* err) {
    StepArgs<R> a{};
    const int cur = m.cur, nxt = !m.cur;
    for (int k = 0; k < 2; k++) {
        a.phi[k] = f.phi[cur][k];
        a.phi_out[k] = f.phi[nxt][k];
        a.P[k] = f.P[k];
    }
    for (int i = 0; i < 3; i++) {
        a.y[i] = f.y[cur][i];
        a.y_out[i] = f.y[nxt][i];
        a.q[i] = f.q[cur][i];
        a.q_out[i] = f.q[nxt][i];
        a.E[i] = f.E[i];
        a.inv_sp[i] = (R)c.inv_sp[i];
    }
    a.lo = f.lo;
    a.lo_P2 = f.lo_P2;
    for (int i = 0; i < 3; i++) {
        a.fy[i] = f.fy[cur][i];
        a.fy_out[i] = f.fy[nxt][i];
        a.fe[i] = f.fe[i];
        a.fq[i] = f.fq[cur][i];
        a.fq_out[i] = f.fq[nxt][i];
        a.fE[i] = f.fE[i];
    }
    a.flo5 = f.flo5;
    a.flo7 = f.flo7;
    a.gsq_f = (R)c.gsq[4];
    for (int k = 0; k < 2; k++) {
        a.ref[k].x = (R)m.ref[k][0];
        a.ref[k].y = (R)m.ref[k][1];
        a.ref[k].z = (R)m.ref[k][2];
        a.ref[k].w = (R)m.ref[k][3];
    }
    a.nx = c.nx; a.ny = c.ny; a.nz = c.nz;
    a.bz = m.bz; a.by = m.by;
    a.nbz = (c.nz + m.bz - 1) / m.bz;
    a.nby = (c.ny + m.by - 1) / m.by;
    a.nblocks = (long long)a.nbz * a.nby * c.nx;
    a.order = m.order;
    a.tiles = m.tiles;
    a.bz_shift = 0;
    while ((1 << a.bz_shift) < m.bz) a.bz_shift++;
    a.dt2 = (R)(c.dt * c.dt);
    a.keep = (R)(1.0 - damping_step);
    for (int k = 0; k < 4; k++) a.gsq[k] = (R)c.gsq[k];
    a.err = err;
    return a;
}

// Largest LDS tile a step block may request (the fixed LDS limit of a gfx9 workgroup).
constexpr size_t kPlaneLdsMax = 65536;

// LDS bytes of a step block with staging mode pl (0, 1, 2).
inline size_t step_lds_bytes(int bz, int by, int nd, int pl) {
    if (pl == 0) return 0;
    const size_t b = plane_lds_bytes(bz, by, nd);
    return pl == 2 ? std::max(b, share_lds_bytes(bz, by, nd)) : b;
}

// The staging mode a launch runs with: the requested one where it applies.
// Stage 2 needs bz + by <= 256 (one ring item per thread).
inline int step_stage_mode(int stage, int bz, int by, int nd, int G = 4) {
    (void)G;
    if (stage >= 2 && (bz + by > 256 || step_lds_bytes(bz, by, nd, 2) > kPlaneLdsMax)) stage = 1;
    if (stage >= 1 && step_lds_bytes(bz, by, nd, 1) > kPlaneLdsMax) stage = 0;
    return stage;
}

template <bool CMP, int PL>
void launch_step_f32(const StepArgs<float>& a, const Coef& c, int min_waves, long long nb, hipStream_t st) {
    const size_t lds = step_lds_bytes(a.bz, a.by, c.nd, PL);
    if constexpr (PL == 2) {
        if (c.nd == 2 && c.G == 5) mom_step<float, 2, 0, CMP, 2, 5><<<nb, 256, lds, st>>>(a, c);
        else if (c.nd == 2 && c.G == 7) mom_step<float, 2, 0, CMP, 2, 7><<<nb, 256, lds, st>>>(a, c);
        else if (c.nd == 2) mom_step<float, 2, 0, CMP, 2><<<nb, 256, lds, st>>>(a, c);
        else if (min_waves == 3) mom_step<float, 1, 3, CMP, 2><<<nb, 256, lds, st>>>(a, c);
        else if (min_waves == 4) mom_step<float, 1, 4, CMP, 2><<<nb, 256, lds, st>>>(a, c);
        else mom_step<float, 1, 0, CMP, 2><<<nb, 256, lds, st>>>(a, c);
    } else {
        if (c.nd == 2 && c.G == 5) mom_step<float, 2, 0, CMP, PL, 5><<<nb, 256, lds, st>>>(a, c);
        else if (c.nd == 2 && c.G == 7) mom_step<float, 2, 0, CMP, PL, 7><<<nb, 256, lds, st>>>(a, c);
        else if (c.nd == 2) mom_step<float, 2, 0, CMP, PL><<<nb, 256, lds, st>>>(a, c);
        else if (min_waves == 3) mom_step<float, 1, 3, CMP, PL><<<nb, 256, lds, st>>>(a, c);
        else if (min_waves == 4) mom_step<float, 1, 4, CMP, PL><<<nb, 256, lds, st>>>(a, c);
        else mom_step<float, 1, 0, CMP, PL><<<nb, 256, lds, st>>>(a, c);
    }
}

inline int momentum_step(MomentumState& m, const Coef& c, double damping_step, hipStream_t st, unsigned* err,
                         std::string& msg) {
    if (m.bz * m.by != 256 || m.bz < 1 || m.by < 1) { msg = "momentum step: tile must hold 256 sites"; return GFT_EINVAL; }
    if (int rc = ensure_tiles(m, c, st, msg)) return rc;
    const long long nb = mom_nblocks(m, c);
    if (m.f64) {
        auto a = make_step_args<double>(m, m.fd, c, damping_step, err);
        if (c.nd == 1) mom_step<double, 1, 0, false><<<nb, 256, 0, st>>>(a, c);
        else if (c.G == 5) mom_step<double, 2, 0, false, 0, 5><<<nb, 256, 0, st>>>(a, c);
        else if (c.G == 7) mom_step<double, 2, 0, false, 0, 7><<<nb, 256, 0, st>>>(a, c);
        else mom_step<double, 2, 0, false><<<nb, 256, 0, st>>>(a, c);
    } else {
        auto a = make_step_args<float>(m, m.ff, c, damping_step, err);
        // Every staging mode gives bitwise the same fields; a launch runs the
        // requested mode where it applies (see step_stage_mode), else a lower one.
        const int pl = step_stage_mode(m.stage, m.bz, m.by, c.nd, c.G);
        if (m.cmp) {
            if (pl == 2) launch_step_f32<true, 2>(a, c, m.min_waves, nb, st);
            else if (pl == 1) launch_step_f32<true, 1>(a, c, m.min_waves, nb, st);
            else launch_step_f32<true, 0>(a, c, m.min_waves, nb, st);
        } else {
            if (pl == 2) launch_step_f32<false, 2>(a, c, m.min_waves, nb, st);
            else if (pl == 1) launch_step_f32<false, 1>(a, c, m.min_waves, nb, st);
            else launch_step_f32<false, 0>(a, c, m.min_waves, nb, st);
        }
    }
    if (hipGetLastError() != hipSuccess) { msg = "momentum step launch failed"; return GFT_EHIP; }
    m.cur = !m.cur;
    m.prev_valid = 1;
    return GFT_OK;
}

template <class R>
ObsArgs<R> make_obs_args(const MomentumState& m, const MomFields<R>& f, const Coef& c) {
    ObsArgs<R> a{};
    const int cur = m.cur, old = !m.cur;
    for (int k = 0; k < 2; k++) {
        a.phi[k] = f.phi[cur][k];
        a.phi_old[k] = f.phi[old][k];
        a.P[k] = f.P[k];
    }
    for (int i = 0; i < 3; i++) {
        a.y[i] = f.y[cur][i];
        a.q[i] = f.q[cur][i];
        a.q_old[i] = f.q[old][i];
        a.E[i] = f.E[i];
    }
    a.lo = f.lo;
    a.lo_P2 = f.lo_P2;
    for (int i = 0; i < 3; i++) {
        a.fy[i] = f.fy[cur][i];
        a.fe[i] = f.fe[i];
        a.fq[i] = f.fq[cur][i];
        a.fq_old[i] = f.fq[old][i];
        a.fE[i] = f.fE[i];
    }
    a.flo5 = f.flo5;
    a.flo7 = f.flo7;
    for (int k = 0; k < 2; k++)
        for (int m4 = 0; m4 < 4; m4++) a.ref[k][m4] = m.ref[k][m4];
    a.nx = c.nx; a.ny = c.ny; a.nz = c.nz;
    return a;
}

enum class MomObs { Energy, Gauss, Loops };

inline int momentum_observe(const MomentumState& m, const Coef& c, MomObs what, double* partial, long long nblocks,
                            hipStream_t st, std::string& msg) {
    if (what == MomObs::Gauss && !m.prev_valid) {
        msg = "Gauss law needs the state of the previous step: run at least one step after import";
        return GFT_EINVAL;
    }
// Energy and Gauss law by gauge group (the Higgs-family terms for G 5, 7); the
// Wilson loops are those of the SU(2) x U(1) links in every group.
#define GFT_MOM_LAUNCH_R(KERNEL, R, FLD)                                                                 \
    {                                                                                                    \
        auto a = make_obs_args<R>(m, FLD, c);                                                            \
        if (c.nd == 1) KERNEL<R, 1><<<nblocks, 256, 0, st>>>(a, c, partial);                             \
        else if (c.G == 5) KERNEL<R, 2, 5><<<nblocks, 256, 0, st>>>(a, c, partial);                      \
        else if (c.G == 7) KERNEL<R, 2, 7><<<nblocks, 256, 0, st>>>(a, c, partial);                      \
        else KERNEL<R, 2><<<nblocks, 256, 0, st>>>(a, c, partial);                                       \
    }
#define GFT_MOM_LAUNCH(KERNEL)                                                                           \
    if (m.f64) GFT_MOM_LAUNCH_R(KERNEL, double, m.fd)                                                    \
    else GFT_MOM_LAUNCH_R(KERNEL, float, m.ff)
    if (what == MomObs::Energy) { GFT_MOM_LAUNCH(mom_energy) }
    else if (what == MomObs::Gauss) { GFT_MOM_LAUNCH(mom_gauss) }
    else if (m.f64) {
        auto a = make_obs_args<double>(m, m.fd, c);
        if (c.nd == 1) mom_loops<double, 1><<<nblocks, 256, 0, st>>>(a, c, partial);
        else mom_loops<double, 2><<<nblocks, 256, 0, st>>>(a, c, partial);
    } else {
        auto a = make_obs_args<float>(m, m.ff, c);
        if (c.nd == 1) mom_loops<float, 1><<<nblocks, 256, 0, st>>>(a, c, partial);
        else mom_loops<float, 2><<<nblocks, 256, 0, st>>>(a, c, partial);
    }
#undef GFT_MOM_LAUNCH
#undef GFT_MOM_LAUNCH_R
    if (hipGetLastError() != hipSuccess) { msg = "momentum observable launch failed"; return GFT_EHIP; }
    return GFT_OK;
}

#endif  // GFT_KERNELS_ONLY

}  // namespace gft
