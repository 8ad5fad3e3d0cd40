// lattice.hpp -- D2Q9 / D3Q19 velocity sets as compile-time functions, and
// the population storage formats.
//
// The velocity order, weights, opposite and free-slip reflection tables are
// the reference's (CoupLB/couplb_lattice.h:24-75): populations are exchanged
// with the host in that index order, so the tables have to be identical.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace couplb {

template <int D> struct Lat;

template <> struct Lat<2> {
  static constexpr int Q = 9;
  __host__ __device__ static constexpr int ex(int q) {
    constexpr int t[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
    return t[q];
  }
  __host__ __device__ static constexpr int ey(int q) {
    constexpr int t[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
    return t[q];
  }
  __host__ __device__ static constexpr int ez(int) { return 0; }
  __host__ __device__ static constexpr double w(int q) {
    return q == 0 ? 4.0 / 9.0 : (q < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
  }
  __host__ __device__ static constexpr int opp(int q) {
    constexpr int t[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};
    return t[q];
  }
  // index with component d of e_q negated (free-slip specular reflection)
  __host__ __device__ static constexpr int reflect(int d, int q) {
    constexpr int t[3][9] = {{0, 2, 1, 3, 4, 8, 7, 6, 5},
                             {0, 1, 2, 4, 3, 7, 8, 5, 6},
                             {0, 1, 2, 3, 4, 5, 6, 7, 8}};
    return t[d][q];
  }
};

template <> struct Lat<3> {
  static constexpr int Q = 19;
  __host__ __device__ static constexpr int ex(int q) {
    constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    return t[q];
  }
  __host__ __device__ static constexpr int ey(int q) {
    constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    return t[q];
  }
  __host__ __device__ static constexpr int ez(int q) {
    constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    return t[q];
  }
  __host__ __device__ static constexpr double w(int q) {
    return q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
  }
  __host__ __device__ static constexpr int opp(int q) {
    constexpr int t[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    return t[q];
  }
  __host__ __device__ static constexpr int reflect(int d, int q) {
    constexpr int t[3][19] = {
        {0, 2, 1, 3, 4, 5, 6, 10, 9, 8, 7, 14, 13, 12, 11, 15, 16, 17, 18},
        {0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 18, 17, 16, 15},
        {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16}};
    return t[d][q];
  }
};

// ---------------------------------------------------------------------------
// Correctly rounded division by the constants cs^2 = fl(1/3) and
// cs^4 = fl(fl(1/3)^2) in three FP64 operations (Markstein: q = x*r with
// r = RN(1/c), remainder by FMA, one correction FMA), so the FP64 path
// rounds exactly like the reference's `x / cs2` and `x / cs4`
// (couplb_collision.h:51-52, couplb_lattice.h:143) at a fraction of the cost
// of a general division.
// ---------------------------------------------------------------------------
__device__ inline double div_cs2(double x) {
  const double c = 1.0 / 3.0;
  const double q = x * 3.0;
  const double e = __builtin_fma(-q, c, x);
  return __builtin_fma(e, 3.0, q);
}
__device__ inline double div_cs4(double x) {
  const double c = (1.0 / 3.0) * (1.0 / 3.0);
  const double q = x * 9.0;
  const double e = __builtin_fma(-q, c, x);
  return __builtin_fma(e, 9.0, q);
}

// ---------------------------------------------------------------------------
// Storage formats.  FP64 stores f_i itself (the reference's representation).
// The reduced formats store the deviation (f_i - w_i*rho0)*scale and compute
// in float; scale is a power of two (exact) that lifts small deviations out
// of the half-precision subnormal range, or sets the 16-bit fixed point's
// resolution.  round(v) is the value v takes once stored.
// ---------------------------------------------------------------------------
struct bf16_t { uint16_t bits; };

template <typename S> struct Store;

template <> struct Store<double> {
  using T = double;                    // arithmetic type
  static constexpr bool deviation = false;
  __device__ static inline T round(T v) { return v; }
  __device__ static inline T load(const double* p) { return __builtin_nontemporal_load(p); }
  __device__ static inline void store(double* p, T v) { __builtin_nontemporal_store(v, p); }
  __device__ static inline T decode(double v) { return v; }
  __device__ static inline double encode(T v) { return v; }
};

template <> struct Store<float> {
  using T = float;
  static constexpr bool deviation = true;
  __device__ static inline T round(T v) { return v; }
  __device__ static inline T load(const float* p) { return __builtin_nontemporal_load(p); }
  __device__ static inline void store(float* p, T v) { __builtin_nontemporal_store(v, p); }
  __device__ static inline T decode(float v) { return v; }
  __device__ static inline float encode(T v) { return v; }
};

template <> struct Store<_Float16> {
  using T = float;
  static constexpr bool deviation = true;
  __device__ static inline T round(T v) { return (float)(_Float16)v; }
  __device__ static inline T load(const _Float16* p) {
    return (float)__builtin_nontemporal_load(p);
  }
  __device__ static inline void store(_Float16* p, T v) {
    __builtin_nontemporal_store((_Float16)v, p);
  }
  __device__ static inline T decode(_Float16 v) { return (float)v; }
  __device__ static inline _Float16 encode(T v) { return (_Float16)v; }
};

template <> struct Store<bf16_t> {
  using T = float;
  static constexpr bool deviation = true;
  __device__ static inline T load(const bf16_t* p) {
    const uint16_t b = __builtin_nontemporal_load(&p->bits);
    return __uint_as_float((uint32_t)b << 16);
  }
  // round to nearest even; inputs are finite in normal operation
  __device__ static inline void store(bf16_t* p, T v) { __builtin_nontemporal_store(encode(v).bits, &p->bits); }
  __device__ static inline T decode(bf16_t v) { return __uint_as_float((uint32_t)v.bits << 16); }
  __device__ static inline bf16_t encode(T v) {
    const uint32_t u = __float_as_uint(v);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return bf16_t{(uint16_t)(r >> 16)};
  }
  __device__ static inline T round(T v) { return decode(encode(v)); }
};

// 16-bit fixed point: the deviation times scale (a power of two), rounded
// to the nearest integer in [-32767, 32767].  In the stored scale every
// value is an integer, so sums of populations are exact in float and the
// step can keep each node's mass exactly.  Values beyond the range are
// clamped to it and counted in i16_saturations (per GPU, since load).
struct i16_t { int16_t v; };
extern __device__ unsigned long long i16_saturations;

template <> struct Store<i16_t> {
  using T = float;
  static constexpr bool deviation = true;
  __device__ static inline T load(const i16_t* p) {
    return (float)__builtin_nontemporal_load(&p->v);
  }
  __device__ static inline void store(i16_t* p, T v) { __builtin_nontemporal_store(encode(v).v, &p->v); }
  __device__ static inline T decode(i16_t v) { return (float)v.v; }
  __device__ static inline i16_t encode(T v) {
    float r = __builtin_rintf(v);  // round to nearest, ties to even
    if (!(r >= -32767.0f && r <= 32767.0f)) {
      atomicAdd(&i16_saturations, 1ull);
      r = (r > 0.0f) ? 32767.0f : -32767.0f;
    }
    return i16_t{(int16_t)(int)r};
  }
  __device__ static inline T round(T v) { return decode(encode(v)); }
};

}  // namespace couplb
