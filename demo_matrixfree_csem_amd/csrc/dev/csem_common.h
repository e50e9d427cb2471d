// csem_common.h -- shared device helpers for the MI355X solve path.
//
// Complex arithmetic is written out in real arithmetic with the exact
// expression shapes gcc uses for C99 complex on x86-64 (no FMA contraction;
// every file including this one is compiled with -ffp-contract=off):
//   a*b   -> (ar*br - ai*bi, ar*bi + ai*br)
//   a/r   -> (ar/r, ai/r)               (real divisor)
//   r*a   -> (r*ar, r*ai)               (real factor)
//   a/b   -> Smith's algorithm as in libgcc's __divdc3 (normal range)
// so device results equal the reference's bit for bit on the same inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double2 z2;
struct z2x4 { z2 a, b, c, d; };

__host__ __device__ __forceinline__ z2 zmk(double x, double y) { return make_double2(x, y); }
__host__ __device__ __forceinline__ z2 zadd(z2 a, z2 b) { return zmk(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ z2 zsub(z2 a, z2 b) { return zmk(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ z2 zmul(z2 a, z2 b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ __forceinline__ z2 zrmul(double r, z2 a) { return zmk(r * a.x, r * a.y); }
__host__ __device__ __forceinline__ z2 zmulr(z2 a, double r) { return zmk(a.x * r, a.y * r); }
__host__ __device__ __forceinline__ z2 zdivr(z2 a, double r) { return zmk(a.x / r, a.y / r); }
// complex minus real, real part only
__host__ __device__ __forceinline__ z2 zsubr(z2 a, double r) { return zmk(a.x - r, a.y); }

// Smith division, libgcc2.c __divdc3 for operands in the normal range.
__host__ __device__ __forceinline__ z2 zdiv(z2 n, z2 q)
{
  double a = n.x, b = n.y, c = q.x, d = q.y, ratio, denom;
  if (fabs(c) < fabs(d)) {
    ratio = c / d;
    denom = (c * ratio) + d;
    return zmk(((a * ratio) + b) / denom, ((b * ratio) - a) / denom);
  }
  ratio = d / c;
  denom = (d * ratio) + c;
  return zmk(((b * ratio) + a) / denom, (b - (a * ratio)) / denom);
}

// x / d for d > 0 given rd = RN(1/d): q = RN(x*rd) is within an ulp of
// x/d, the residual d*q - x is exact in an FMA, and one correction step
// RN(q - rn*rd) is the correctly rounded quotient (Markstein's theorem; it
// is the final step of the hardware division sequence too).  With the
// residual taken as d*q - x, a zero numerator keeps its sign (-0 -> -0,
// +0 -> +0).  Equal to the IEEE quotient for finite operands whose residual
// does not underflow (|x| above about 1e-290).
__host__ __device__ __forceinline__ double ddiv(double x, double d, double rd)
{
  const double q = x * rd;
  const double rn = __builtin_fma(d, q, -x);
  return __builtin_fma(-rn, rd, q);
}
__host__ __device__ __forceinline__ z2 zdivq(z2 a, double d, double rd) { return zmk(ddiv(a.x, d, rd), ddiv(a.y, d, rd)); }

// Whole-wavefront lane shifts (DPP wave_shl:1 / wave_shr:1; a lane with no
// source lane reads 0).  Every lane of the wave must be active.
#define DPP_WAVE_SHL1 0x130  // lane l reads lane l+1
#define DPP_WAVE_SHR1 0x138  // lane l reads lane l-1
#define DPP_WAVE_ROR1 0x13C  // lane l reads lane l-1, lane 0 reads lane 63
#define DPP_WAVE_ROL1 0x134  // lane l reads lane l+1, lane 63 reads lane 0
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int CTRL>
__device__ __forceinline__ z2 zdpp(z2 v) { return zmk(dpp_d<CTRL>(v.x), dpp_d<CTRL>(v.y)); }

// Single precision counterparts for the mixed-precision inner solver.  A
// plain aligned struct rather than HIP's float2 (a union-based vector
// class): selects and conditional copies of it stay in registers.
struct __align__(8) c2 {
  float x, y;
};
__host__ __device__ __forceinline__ c2 cmk(float x, float y) { return c2{x, y}; }
template <int CTRL>
__device__ __forceinline__ c2 cdpp(c2 v) { return c2{dpp_f<CTRL>(v.x), dpp_f<CTRL>(v.y)}; }
// The same shift, but a lane with no source lane keeps `old` (its lane of
// the first argument): one instruction per component shifts a wave's values
// and feeds its edge lane from elsewhere.
template <int CTRL>
__device__ __forceinline__ c2 cdpp_keep(c2 old, c2 v)
{
  return c2{__int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old.x), __float_as_int(v.x), CTRL, 0xf, 0xf, false)),
            __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old.y), __float_as_int(v.y), CTRL, 0xf, 0xf, false))};
}
__host__ __device__ __forceinline__ c2 cadd(c2 a, c2 b) { return cmk(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ c2 csub(c2 a, c2 b) { return cmk(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ c2 cmul(c2 a, c2 b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ __forceinline__ c2 crmul(float r, c2 a) { return cmk(r * a.x, r * a.y); }
__host__ __device__ __forceinline__ c2 cdivr(c2 a, float r) { return cmk(a.x / r, a.y / r); }
__host__ __device__ __forceinline__ c2 cdiv(c2 n, c2 q)
{
  float a = n.x, b = n.y, c = q.x, d = q.y, ratio, denom;
  if (fabsf(c) < fabsf(d)) {
    ratio = c / d;
    denom = (c * ratio) + d;
    return cmk(((a * ratio) + b) / denom, ((b * ratio) - a) / denom);
  }
  ratio = d / c;
  denom = (d * ratio) + c;
  return cmk(((b * ratio) + a) / denom, (b - (a * ratio)) / denom);
}

// Deterministic reduction layout shared by every kernel that produces dot
// product partials: chunks of RED_CHUNK = RED_T * RED_E consecutive
// elements, thread t of a 256-thread block sums elements
// chunk*RED_CHUNK + m*RED_T + t for m = 0..RED_E-1 in order, then the
// block folds v[t] += v[t+h] for h = 128, 64, ..., 1.
#define RED_T 256
#define RED_E 16
#define RED_CHUNK (RED_T * RED_E)

// Block fold of one double per thread (blockDim == 256), result in lane 0
// of wave 0.  `sh` must hold 256 doubles.
__device__ __forceinline__ double block_fold(double v, double *sh)
{
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  if (t < 128) sh[t] = sh[t] + sh[t + 128];
  __syncthreads();
  if (t < 64) {
    v = sh[t] + sh[t + 64];
    // within wave 0: v[t] += v[t+h], h = 32..1
    for (int h = 32; h >= 1; h >>= 1) v = v + __shfl_down(v, h, 64);
  }
  __syncthreads();
  return v;
}

struct Dims {
  int n1, n2, n3;      // operator grid (n3 = operator depth)
  long lEx, lEy, lEz;  // operator component lengths
};

__host__ __device__ __forceinline__ long id1(int n1, int n2, int k, int j, int i) { return i + (long)n1 * (j + (long)(n2 + 1) * k); }
__host__ __device__ __forceinline__ long id2(int n1, int n2, int k, int j, int i) { return i + (long)(n1 + 1) * (j + (long)n2 * k); }
__host__ __device__ __forceinline__ long id3(int n1, int n2, int k, int j, int i) { return i + (long)(n1 + 1) * (j + (long)(n2 + 1) * k); }

#define HIP_OK(call)                                                       \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) return csem_fail(-1, "%s: %s (%s:%d)", #call,    \
                                           hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

extern "C" int csem_fail(int code, const char *fmt, ...);
