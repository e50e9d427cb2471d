// csem_kernels.h -- HIP kernels of the FP64 solve path (gfx950).
//
// Each kernel restates one reference loop (file:line in its comment) with
// the same per-element expression tree, so results equal the oracle bit for
// bit; what changes is the order in which independent elements are visited.
#pragma once
#include "csem_common.h"

// ---------------------------------------------------------------- operator

struct ApplyArgs {
  const z2 *x;
  z2 *y;
  const double *s11, *s22, *s33;  // operator k=0 plane (full grid offset by kstart)
  const double *d1, *d2, *d3, *d1s, *d2s, *d3s;  // d3, d3s offset by kstart
  const z2 *hxg, *hyg;  // ghost planes under the airwave boundary
  z2 iwmu, iomu;
  int n1, n2, n3, fvm, airbc, kc;
  long lEx, lEy;
};

// y = A x (fdfd.c:86-205), one thread per (i, j) column marching over kc
// planes.  H on the faces around each output is recomputed from E with the
// expression of fdfd.c:133-145, so no H array is stored; Hx and Hy of the
// previous plane are carried in registers.
template <int TI, int TJ>
__global__ __launch_bounds__(TI *TJ) void k_apply(ApplyArgs a)
{
  const int i = blockIdx.x * TI + threadIdx.x;
  const int j = blockIdx.y * TJ + threadIdx.y;
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  if (i > n1 || j > n2) return;
  const int k0 = blockIdx.z * a.kc;
  const int k1 = min(k0 + a.kc, n3 + 1);
  const z2 *Ex = a.x, *Ey = a.x + a.lEx, *Ez = a.x + a.lEx + a.lEy;
  z2 *bx = a.y, *by = a.y + a.lEx, *bz = a.y + a.lEx + a.lEy;
  const double *d1 = a.d1, *d2 = a.d2, *d3 = a.d3, *d1s = a.d1s, *d2s = a.d2s, *d3s = a.d3s;
  const z2 iwmu = a.iwmu;

#define HX(k, j, i)                                                                              \
  zmul(zsub(zdivr(zsub(Ez[id3(n1, n2, k, (j) + 1, i)], Ez[id3(n1, n2, k, j, i)]), d2s[j]),       \
            zdivr(zsub(Ey[id2(n1, n2, (k) + 1, j, i)], Ey[id2(n1, n2, k, j, i)]), d3s[k])),      \
       iwmu)
#define HY(k, j, i)                                                                              \
  zmul(zsub(zdivr(zsub(Ex[id1(n1, n2, (k) + 1, j, i)], Ex[id1(n1, n2, k, j, i)]), d3s[k]),       \
            zdivr(zsub(Ez[id3(n1, n2, k, j, (i) + 1)], Ez[id3(n1, n2, k, j, i)]), d1s[i])),      \
       iwmu)
#define HZ(k, j, i)                                                                              \
  zmul(zsub(zdivr(zsub(Ey[id2(n1, n2, k, j, (i) + 1)], Ey[id2(n1, n2, k, j, i)]), d1s[i]),       \
            zdivr(zsub(Ex[id1(n1, n2, k, (j) + 1, i)], Ex[id1(n1, n2, k, j, i)]), d2s[j])),      \
       iwmu)

  const bool iin = (i >= 1 && i <= n1 - 1), jin = (j >= 1 && j <= n2 - 1);
  // Hx is needed where by/bz are formed (1<=i<=n1-1, j<n2), Hy where bx/bz
  // are (i<n1, 1<=j<=n2-1); the previous plane's values ride in registers.
  const bool need_hx = iin && j < n2, need_hy = jin && i < n1;
  z2 hx_prev = zmk(0, 0), hy_prev = zmk(0, 0);
  if (k0 > 0) {
    if (need_hx) hx_prev = HX(k0 - 1, j, i);
    if (need_hy) hy_prev = HY(k0 - 1, j, i);
  }
  for (int k = k0; k < k1; k++) {
    const bool kin = (k >= 1 && k <= n3 - 1) || (a.airbc && k == 0);
    const bool bx_on = i < n1 && jin && kin;
    const bool by_on = j < n2 && iin && kin;
    const bool bz_on = k < n3 && iin && jin;
    z2 hx_c = zmk(0, 0), hy_c = zmk(0, 0), hz_c = zmk(0, 0);
    if (k < n3) {
      if (need_hx) hx_c = HX(k, j, i);
      if (need_hy) hy_c = HY(k, j, i);
    }
    if (bx_on || by_on) hz_c = HZ(k, j, i);
    if (i < n1) {  // bx: i < n1, j <= n2, k <= n3
      z2 v = zmk(0, 0);
      if (bx_on) {
        z2 t1 = zdivr(zsub(hz_c, HZ(k, j - 1, i)), d2[j]);
        z2 hym = (k == 0) ? a.hyg[i + (long)n1 * j] : hy_prev;
        z2 t2 = zdivr(zsub(hy_c, hym), d3[k]);
        long q = id1(n1, n2, k, j, i);
        v = zsub(zsub(t1, t2), zrmul(a.s11[q], Ex[q]));
        if (a.fvm) v = zmul(v, zrmul(d1s[i] * d2[j] * d3[k], a.iomu));
      }
      bx[id1(n1, n2, k, j, i)] = v;
    }
    if (j < n2) {  // by: i <= n1, j < n2, k <= n3
      z2 v = zmk(0, 0);
      if (by_on) {
        z2 hxm = (k == 0) ? a.hxg[i + (long)(n1 + 1) * j] : hx_prev;
        z2 t1 = zdivr(zsub(hx_c, hxm), d3[k]);
        z2 t2 = zdivr(zsub(hz_c, HZ(k, j, i - 1)), d1[i]);
        long q = id2(n1, n2, k, j, i);
        v = zsub(zsub(t1, t2), zrmul(a.s22[q], Ey[q]));
        if (a.fvm) v = zmul(v, zrmul(d1[i] * d2s[j] * d3[k], a.iomu));
      }
      by[id2(n1, n2, k, j, i)] = v;
    }
    if (k < n3) {  // bz: i <= n1, j <= n2, k < n3
      z2 v = zmk(0, 0);
      if (bz_on) {
        z2 t1 = zdivr(zsub(hy_c, HY(k, j, i - 1)), d1[i]);
        z2 t2 = zdivr(zsub(hx_c, HX(k, j - 1, i)), d2[j]);
        long q = id3(n1, n2, k, j, i);
        v = zsub(zsub(t1, t2), zrmul(a.s33[q], Ez[q]));
        if (a.fvm) v = zmul(v, zrmul(d1[i] * d2[j] * d3s[k], a.iomu));
      }
      bz[id3(n1, n2, k, j, i)] = v;
    }
    hx_prev = hx_c;
    hy_prev = hy_c;
  }
#undef HX
#undef HY
#undef HZ
}

// H = curl E / (i w mu0) on all faces (fdfd.c:125-149), Hx | Hy | Hz.
__global__ void k_curl_h(ApplyArgs a, z2 *H, long lHx, long lHy)
{
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long tot = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  if (t >= tot) return;
  int i = t % (n1 + 1);
  int j = (t / (n1 + 1)) % (n2 + 1);
  int k = t / ((long)(n1 + 1) * (n2 + 1));
  const z2 *Ex = a.x, *Ey = a.x + a.lEx, *Ez = a.x + a.lEx + a.lEy;
  const double *d1s = a.d1s, *d2s = a.d2s, *d3s = a.d3s;
  const z2 iwmu = a.iwmu;
  if (j < n2 && k < n3)
    H[i + (long)(n1 + 1) * (j + (long)n2 * k)] =
        zmul(zsub(zdivr(zsub(Ez[id3(n1, n2, k, j + 1, i)], Ez[id3(n1, n2, k, j, i)]), d2s[j]),
                  zdivr(zsub(Ey[id2(n1, n2, k + 1, j, i)], Ey[id2(n1, n2, k, j, i)]), d3s[k])), iwmu);
  if (i < n1 && k < n3)
    H[lHx + i + (long)n1 * (j + (long)(n2 + 1) * k)] =
        zmul(zsub(zdivr(zsub(Ex[id1(n1, n2, k + 1, j, i)], Ex[id1(n1, n2, k, j, i)]), d3s[k]),
                  zdivr(zsub(Ez[id3(n1, n2, k, j, i + 1)], Ez[id3(n1, n2, k, j, i)]), d1s[i])), iwmu);
  if (i < n1 && j < n2)
    H[lHx + lHy + i + (long)n1 * (j + (long)n2 * k)] =
        zmul(zsub(zdivr(zsub(Ey[id2(n1, n2, k, j, i + 1)], Ey[id2(n1, n2, k, j, i)]), d1s[i]),
                  zdivr(zsub(Ex[id1(n1, n2, k, j + 1, i)], Ex[id1(n1, n2, k, j, i)]), d2s[j])), iwmu);
}

// Hz on the plane k = 0 only (the airwave input), idz(0,j,i) = i + n1*j.
__global__ void k_hz0(ApplyArgs a, z2 *hz)
{
  const int n1 = a.n1, n2 = a.n2;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n1 * n2) return;
  int i = t % n1, j = t / n1;
  const z2 *Ex = a.x, *Ey = a.x + a.lEx;
  hz[t] = zmul(zsub(zdivr(zsub(Ey[id2(n1, n2, 0, j, i + 1)], Ey[id2(n1, n2, 0, j, i)]), a.d1s[i]),
                    zdivr(zsub(Ex[id1(n1, n2, 0, j + 1, i)], Ex[id1(n1, n2, 0, j, i)]), a.d2s[j])),
               a.iwmu);
}

// ----------------------------------------------------------------- LU-SGS

struct LuArgs {
  const z2 *x;  // right-hand side, operator grid
  z2 *W;        // sweep state on the full grid, Ex | Ey | Ez
  z2 *y;        // result, operator grid
  const double *s11, *s22, *s33;
  const double *d1, *d2, *d3, *d1s, *d2s, *d3s;
  z2 iwmu, iomu;
  int fvm, n1, n2, n3f, kofs, n3op;
  long fEx, fEy, lEx, lEy;
};

// Cell updates of fdfd.c:357-369 (Ex), 383-395 (Ey), 409-421 (Ez): the cell
// is first set to 0, so every "- E(k,j,i)" term subtracts zero.
__device__ __forceinline__ z2 cell_ex(z2 b, double sig, double d1s_i, double d2s_j, double d2s_jm,
                                      double d2_j, double d3s_k, double d3s_km, double d3_k,
                                      z2 ey_ip, z2 ey_i, z2 ey_jm_ip, z2 ey_jm_i,
                                      z2 ex_jp, z2 ex_jm, z2 ex_kp, z2 ex_km,
                                      z2 ez_ip, z2 ez_i, z2 ez_km_ip, z2 ez_km_i,
                                      z2 iwmu, int fvm, z2 iomu, double vol)
{
  const z2 z0 = zmk(0.0, 0.0);
  z2 t1 = zsub(zdivr(zsub(ey_ip, ey_i), d1s_i), zdivr(zsub(ex_jp, z0), d2s_j));
  z2 t2 = zsub(zdivr(zsub(ey_jm_ip, ey_jm_i), d1s_i), zdivr(zsub(z0, ex_jm), d2s_jm));
  z2 t3 = zsub(zdivr(zsub(ex_kp, z0), d3s_k), zdivr(zsub(ez_ip, ez_i), d1s_i));
  z2 t4 = zsub(zdivr(zsub(z0, ex_km), d3s_km), zdivr(zsub(ez_km_ip, ez_km_i), d1s_i));
  z2 aij = zsub(zmul(zsub(zdivr(zsub(t1, t2), d2_j), zdivr(zsub(t3, t4), d3_k)), iwmu), zrmul(sig, z0));
  double c = (1. / d2s_j + 1. / d2s_jm) / d2_j + (1. / d3s_k + 1. / d3s_km) / d3_k;
  z2 aii = zsubr(zrmul(c, iwmu), sig);
  if (fvm) {
    z2 f = zrmul(vol, iomu);
    aij = zmul(aij, f);
    aii = zmul(aii, f);
  }
  return zdiv(zsub(b, aij), aii);
}

__device__ __forceinline__ z2 cell_ey(z2 b, double sig, double d1s_i, double d1s_im, double d1_i,
                                      double d2s_j, double d3s_k, double d3s_km, double d3_k,
                                      z2 ez_jp, z2 ez_j, z2 ez_km_jp, z2 ez_km_j,
                                      z2 ey_kp, z2 ey_km, z2 ey_ip, z2 ey_im,
                                      z2 ex_jp, z2 ex_j, z2 ex_jp_im, z2 ex_j_im,
                                      z2 iwmu, int fvm, z2 iomu, double vol)
{
  const z2 z0 = zmk(0.0, 0.0);
  z2 t1 = zsub(zdivr(zsub(ez_jp, ez_j), d2s_j), zdivr(zsub(ey_kp, z0), d3s_k));
  z2 t2 = zsub(zdivr(zsub(ez_km_jp, ez_km_j), d2s_j), zdivr(zsub(z0, ey_km), d3s_km));
  z2 t3 = zsub(zdivr(zsub(ey_ip, z0), d1s_i), zdivr(zsub(ex_jp, ex_j), d2s_j));
  z2 t4 = zsub(zdivr(zsub(z0, ey_im), d1s_im), zdivr(zsub(ex_jp_im, ex_j_im), d2s_j));
  z2 aij = zsub(zmul(zsub(zdivr(zsub(t1, t2), d3_k), zdivr(zsub(t3, t4), d1_i)), iwmu), zrmul(sig, z0));
  double c = (1. / d3s_k + 1. / d3s_km) / d3_k + (1. / d1s_i + 1. / d1s_im) / d1_i;
  z2 aii = zsubr(zrmul(c, iwmu), sig);
  if (fvm) {
    z2 f = zrmul(vol, iomu);
    aij = zmul(aij, f);
    aii = zmul(aii, f);
  }
  return zdiv(zsub(b, aij), aii);
}

__device__ __forceinline__ z2 cell_ez(z2 b, double sig, double d1s_i, double d1s_im, double d1_i,
                                      double d2s_j, double d2s_jm, double d2_j, double d3s_k,
                                      z2 ex_kp, z2 ex_k, z2 ex_kp_im, z2 ex_k_im,
                                      z2 ez_ip, z2 ez_im, z2 ez_jp, z2 ez_jm,
                                      z2 ey_kp, z2 ey_k, z2 ey_kp_jm, z2 ey_k_jm,
                                      z2 iwmu, int fvm, z2 iomu, double vol)
{
  const z2 z0 = zmk(0.0, 0.0);
  z2 t1 = zsub(zdivr(zsub(ex_kp, ex_k), d3s_k), zdivr(zsub(ez_ip, z0), d1s_i));
  z2 t2 = zsub(zdivr(zsub(ex_kp_im, ex_k_im), d3s_k), zdivr(zsub(z0, ez_im), d1s_im));
  z2 t3 = zsub(zdivr(zsub(ez_jp, z0), d2s_j), zdivr(zsub(ey_kp, ey_k), d3s_k));
  z2 t4 = zsub(zdivr(zsub(z0, ez_jm), d2s_jm), zdivr(zsub(ey_kp_jm, ey_k_jm), d3s_k));
  z2 aij = zsub(zmul(zsub(zdivr(zsub(t1, t2), d1_i), zdivr(zsub(t3, t4), d2_j)), iwmu), zrmul(sig, z0));
  double c = (1. / d1s_i + 1. / d1s_im) / d1_i + (1. / d2s_j + 1. / d2s_jm) / d2_j;
  z2 aii = zsubr(zrmul(c, iwmu), sig);
  if (fvm) {
    z2 f = zrmul(vol, iomu);
    aij = zmul(aij, f);
    aii = zmul(aii, f);
  }
  return zdiv(zsub(b, aij), aii);
}

// Each component sweep decouples into independent 2-D Gauss-Seidel problems
// along the component's own axis: Ex couples to Ex only through (k,j+-1)
// and (k+-1,j) (fdfd.c:358-361), Ey through (k+-1,i) and (k,i+-1), Ez
// through (j+-1,i) and (j,i+-1).  One workgroup owns a slice; thread t owns
// slow index s = 1+t and visits fast index f = d - s on diagonal d, so
// every cell sees exactly the neighbours the lexicographic loop gives it:
// the new values of (s,f-1) (own register) and (s-1,f) (thread t-1, one
// diagonal earlier, through LDS), and the old values of (s,f+1), (s+1,f)
// -- zero in a forward sweep, the forward result in a backward one (then
// the roles of +1 and -1 swap).

// Every sweep kernel software-pipelines its global loads: the inputs of a
// cell that do not depend on the sweep's own progress (right-hand side,
// conductivity, the other components, old values) are loaded P diagonals
// ahead into a register ring, so the dependent chain of a diagonal waits
// only for LDS and arithmetic, not for memory.  The 1-D spacings of the fast
// axis are staged in LDS once.  Dynamic LDS: 2*blockDim complex exchange
// slots, then two arrays of (nfast+1) doubles.
#define SGS_P 4
#define SGS_MAXT 256  // rows per slice: n3f-1 and n2-1 must not exceed this

// Ex sweep, slice i = blockIdx.x, s = k, f = j.
template <bool FWD>
__global__ __launch_bounds__(SGS_MAXT) void k_sgs_ex(LuArgs a)
{
  extern __shared__ z2 xch[];
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3f;
  const int i = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int k = 1 + t;
  const bool krow = k <= n3 - 1;
  const z2 *Wx = a.W, *Wy = a.W + a.fEx, *Wz = a.W + a.fEx + a.fEy;
  z2 *Wxw = a.W;
  const z2 *bx = a.x;
  double *cd2s = (double *)(xch + 2 * nt), *cd2 = cd2s + (n2 + 1);
  for (int q = t; q <= n2; q += nt) { cd2s[q] = a.d2s[q]; cd2[q] = a.d2[q]; }
  const double d1s_i = a.d1s[i];
  double d3s_k = 0, d3s_km = 0, d3_k = 0;
  if (krow) { d3s_k = a.d3s[k]; d3s_km = a.d3s[k - 1]; d3_k = a.d3[k]; }
  const int dlo = 2, dhi = (n3 - 1) + (n2 - 1), nst = dhi - dlo + 1;
  const int kop = k - a.kofs;
  struct St { z2 b, ey_ip, ey_i, ey_jm_ip, ey_jm_i, ez_ip, ez_i, ez_km_ip, ez_km_i, ex_jm, ex_km; double sig; };
  auto load = [&](St &c, int st) {
    const int j = (FWD ? dlo + st : dhi - st) - k;
    if (!(st < nst && krow && j >= 1 && j <= n2 - 1)) return;
    c.b = (kop >= 0) ? bx[id1(n1, n2, kop, j, i)] : zmk(0, 0);
    c.sig = a.s11[id1(n1, n2, k, j, i)];
    if (!FWD) {
      c.ex_jm = Wx[id1(n1, n2, k, j - 1, i)];
      c.ex_km = Wx[id1(n1, n2, k - 1, j, i)];
      c.ey_ip = Wy[id2(n1, n2, k, j, i + 1)]; c.ey_i = Wy[id2(n1, n2, k, j, i)];
      c.ey_jm_ip = Wy[id2(n1, n2, k, j - 1, i + 1)]; c.ey_jm_i = Wy[id2(n1, n2, k, j - 1, i)];
      c.ez_ip = Wz[id3(n1, n2, k, j, i + 1)]; c.ez_i = Wz[id3(n1, n2, k, j, i)];
      c.ez_km_ip = Wz[id3(n1, n2, k - 1, j, i + 1)]; c.ez_km_i = Wz[id3(n1, n2, k - 1, j, i)];
    }
  };
  St ring[SGS_P];
  const z2 z0 = zmk(0, 0);
#pragma unroll
  for (int q = 0; q < SGS_P; q++) {
    ring[q].b = z0; ring[q].sig = 0;
    ring[q].ey_ip = ring[q].ey_i = ring[q].ey_jm_ip = ring[q].ey_jm_i = z0;
    ring[q].ez_ip = ring[q].ez_i = ring[q].ez_km_ip = ring[q].ez_km_i = z0;
    ring[q].ex_jm = ring[q].ex_km = z0;
    load(ring[q], q);
  }
  __syncthreads();
  z2 prev = z0;
  for (int st0 = 0; st0 < nst; st0 += SGS_P) {
#pragma unroll
    for (int q = 0; q < SGS_P; q++) {
      const int st = st0 + q;
      if (st < nst) {
        const int j = (FWD ? dlo + st : dhi - st) - k;
        z2 nbv = z0;
        if (st > 0) {
          if (FWD) { if (t > 0) nbv = xch[((st - 1) & 1) * nt + t - 1]; }
          else { if (k + 1 <= n3 - 1) nbv = xch[((st - 1) & 1) * nt + t + 1]; }
        }
        if (krow && j >= 1 && j <= n2 - 1) {
          St &c = ring[q];
          const z2 ex_jp = FWD ? z0 : prev, ex_jm = FWD ? prev : c.ex_jm;
          const z2 ex_kp = FWD ? z0 : nbv, ex_km = FWD ? nbv : c.ex_km;
          const double d2s_j = cd2s[j], d2s_jm = cd2s[j - 1], d2_j = cd2[j];
          const double vol = a.fvm ? d1s_i * d2_j * d3_k : 0.0;
          z2 v = cell_ex(c.b, c.sig, d1s_i, d2s_j, d2s_jm, d2_j, d3s_k, d3s_km, d3_k, c.ey_ip, c.ey_i, c.ey_jm_ip,
                         c.ey_jm_i, ex_jp, ex_jm, ex_kp, ex_km, c.ez_ip, c.ez_i, c.ez_km_ip, c.ez_km_i, a.iwmu, a.fvm,
                         a.iomu, vol);
          if (FWD) Wxw[id1(n1, n2, k, j, i)] = v;
          else if (kop >= 0) a.y[id1(n1, n2, kop, j, i)] = v;
          prev = v;
        }
        load(ring[q], st + SGS_P);
        xch[(st & 1) * nt + t] = prev;
        __syncthreads();
      }
    }
  }
  if (!FWD) {  // cells of this slice outside the sweep range are zero rows
    const int nk = a.n3op + 1;
    for (int q = t; q < nk * (n2 + 1); q += nt) {
      int kq = q / (n2 + 1), j = q % (n2 + 1), kf = kq + a.kofs;
      if (!(kf >= 1 && kf <= n3 - 1 && j >= 1 && j <= n2 - 1)) a.y[id1(n1, n2, kq, j, i)] = zmk(0, 0);
    }
  }
}

// Ey sweep, slice j = blockIdx.x, s = k, f = i.
template <bool FWD>
__global__ __launch_bounds__(SGS_MAXT) void k_sgs_ey(LuArgs a)
{
  extern __shared__ z2 xch[];
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3f;
  const int j = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int k = 1 + t;
  const bool krow = k <= n3 - 1;
  const z2 *Wx = a.W, *Wy = a.W + a.fEx, *Wz = a.W + a.fEx + a.fEy;
  z2 *Wyw = a.W + a.fEx;
  const z2 *by = a.x + a.lEx;
  z2 *yy = a.y + a.lEx;
  double *cd1s = (double *)(xch + 2 * nt), *cd1 = cd1s + (n1 + 1);
  for (int q = t; q <= n1; q += nt) { cd1s[q] = a.d1s[q]; cd1[q] = a.d1[q]; }
  const double d2s_j = a.d2s[j], d2_j = a.d2[j];
  double d3s_k = 0, d3s_km = 0, d3_k = 0;
  if (krow) { d3s_k = a.d3s[k]; d3s_km = a.d3s[k - 1]; d3_k = a.d3[k]; }
  const int dlo = 2, dhi = (n3 - 1) + (n1 - 1), nst = dhi - dlo + 1;
  const int kop = k - a.kofs;
  struct St { z2 b, ez_jp, ez_j, ez_km_jp, ez_km_j, ey_im, ey_km, ex_jp, ex_j, ex_jp_im, ex_j_im; double sig; };
  auto load = [&](St &c, int st) {
    const int i = (FWD ? dlo + st : dhi - st) - k;
    if (!(st < nst && krow && i >= 1 && i <= n1 - 1)) return;
    c.b = (kop >= 0) ? by[id2(n1, n2, kop, j, i)] : zmk(0, 0);
    c.sig = a.s22[id2(n1, n2, k, j, i)];
    if (!FWD) {
      c.ey_im = Wy[id2(n1, n2, k, j, i - 1)];
      c.ey_km = Wy[id2(n1, n2, k - 1, j, i)];
      c.ez_jp = Wz[id3(n1, n2, k, j + 1, i)]; c.ez_j = Wz[id3(n1, n2, k, j, i)];
      c.ez_km_jp = Wz[id3(n1, n2, k - 1, j + 1, i)]; c.ez_km_j = Wz[id3(n1, n2, k - 1, j, i)];
    }
    c.ex_jp = Wx[id1(n1, n2, k, j + 1, i)]; c.ex_j = Wx[id1(n1, n2, k, j, i)];
    c.ex_jp_im = Wx[id1(n1, n2, k, j + 1, i - 1)]; c.ex_j_im = Wx[id1(n1, n2, k, j, i - 1)];
  };
  St ring[SGS_P];
  const z2 z0 = zmk(0, 0);
#pragma unroll
  for (int q = 0; q < SGS_P; q++) {
    ring[q].b = z0; ring[q].sig = 0;
    ring[q].ez_jp = ring[q].ez_j = ring[q].ez_km_jp = ring[q].ez_km_j = z0;
    ring[q].ey_im = ring[q].ey_km = z0;
    ring[q].ex_jp = ring[q].ex_j = ring[q].ex_jp_im = ring[q].ex_j_im = z0;
    load(ring[q], q);
  }
  __syncthreads();
  z2 prev = z0;
  for (int st0 = 0; st0 < nst; st0 += SGS_P) {
#pragma unroll
    for (int q = 0; q < SGS_P; q++) {
      const int st = st0 + q;
      if (st < nst) {
        const int i = (FWD ? dlo + st : dhi - st) - k;
        z2 nbv = z0;
        if (st > 0) {
          if (FWD) { if (t > 0) nbv = xch[((st - 1) & 1) * nt + t - 1]; }
          else { if (k + 1 <= n3 - 1) nbv = xch[((st - 1) & 1) * nt + t + 1]; }
        }
        if (krow && i >= 1 && i <= n1 - 1) {
          St &c = ring[q];
          const z2 ey_ip = FWD ? z0 : prev, ey_im = FWD ? prev : c.ey_im;
          const z2 ey_kp = FWD ? z0 : nbv, ey_km = FWD ? nbv : c.ey_km;
          const double d1s_i = cd1s[i], d1s_im = cd1s[i - 1], d1_i = cd1[i];
          const double vol = a.fvm ? d1s_i * d2_j * d3_k : 0.0;
          z2 v = cell_ey(c.b, c.sig, d1s_i, d1s_im, d1_i, d2s_j, d3s_k, d3s_km, d3_k, c.ez_jp, c.ez_j, c.ez_km_jp,
                         c.ez_km_j, ey_kp, ey_km, ey_ip, ey_im, c.ex_jp, c.ex_j, c.ex_jp_im, c.ex_j_im, a.iwmu,
                         a.fvm, a.iomu, vol);
          Wyw[id2(n1, n2, k, j, i)] = v;
          if (!FWD && kop >= 0) yy[id2(n1, n2, kop, j, i)] = v;
          prev = v;
        }
        load(ring[q], st + SGS_P);
        xch[(st & 1) * nt + t] = prev;
        __syncthreads();
      }
    }
  }
  if (!FWD) {
    const int nk = a.n3op + 1;
    for (int q = t; q < nk * (n1 + 1); q += nt) {
      int kq = q / (n1 + 1), i = q % (n1 + 1), kf = kq + a.kofs;
      if (!(kf >= 1 && kf <= n3 - 1 && i >= 1 && i <= n1 - 1)) yy[id2(n1, n2, kq, j, i)] = zmk(0, 0);
    }
  }
}

// Ez sweep, slice k = blockIdx.x, s = j, f = i.
template <bool FWD>
__global__ __launch_bounds__(SGS_MAXT) void k_sgs_ez(LuArgs a)
{
  extern __shared__ z2 xch[];
  const int n1 = a.n1, n2 = a.n2;
  const int k = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int j = 1 + t;
  const bool jrow = j <= n2 - 1;
  const z2 *Wx = a.W, *Wy = a.W + a.fEx, *Wz = a.W + a.fEx + a.fEy;
  z2 *Wzw = a.W + a.fEx + a.fEy;
  const int kop = k - a.kofs;
  const z2 *bz = a.x + a.lEx + a.lEy;
  z2 *yz = a.y + a.lEx + a.lEy;
  double *cd1s = (double *)(xch + 2 * nt), *cd1 = cd1s + (n1 + 1);
  for (int q = t; q <= n1; q += nt) { cd1s[q] = a.d1s[q]; cd1[q] = a.d1[q]; }
  const double d3s_k = a.d3s[k], d3_k = a.d3[k];
  double d2s_j = 0, d2s_jm = 0, d2_j = 0;
  if (jrow) { d2s_j = a.d2s[j]; d2s_jm = a.d2s[j - 1]; d2_j = a.d2[j]; }
  const int dlo = 2, dhi = (n2 - 1) + (n1 - 1), nst = dhi - dlo + 1;
  struct St { z2 b, ez_im, ez_jm, ex_kp, ex_k, ex_kp_im, ex_k_im, ey_kp, ey_k, ey_kp_jm, ey_k_jm; double sig; };
  auto load = [&](St &c, int st) {
    const int i = (FWD ? dlo + st : dhi - st) - j;
    if (!(st < nst && jrow && i >= 1 && i <= n1 - 1)) return;
    c.b = (kop >= 0) ? bz[id3(n1, n2, kop, j, i)] : zmk(0, 0);
    c.sig = a.s33[id3(n1, n2, k, j, i)];
    if (!FWD) {
      c.ez_im = Wz[id3(n1, n2, k, j, i - 1)];
      c.ez_jm = Wz[id3(n1, n2, k, j - 1, i)];
    }
    c.ex_kp = Wx[id1(n1, n2, k + 1, j, i)]; c.ex_k = Wx[id1(n1, n2, k, j, i)];
    c.ex_kp_im = Wx[id1(n1, n2, k + 1, j, i - 1)]; c.ex_k_im = Wx[id1(n1, n2, k, j, i - 1)];
    c.ey_kp = Wy[id2(n1, n2, k + 1, j, i)]; c.ey_k = Wy[id2(n1, n2, k, j, i)];
    c.ey_kp_jm = Wy[id2(n1, n2, k + 1, j - 1, i)]; c.ey_k_jm = Wy[id2(n1, n2, k, j - 1, i)];
  };
  St ring[SGS_P];
  const z2 z0 = zmk(0, 0);
#pragma unroll
  for (int q = 0; q < SGS_P; q++) {
    ring[q].b = z0; ring[q].sig = 0;
    ring[q].ez_im = ring[q].ez_jm = z0;
    ring[q].ex_kp = ring[q].ex_k = ring[q].ex_kp_im = ring[q].ex_k_im = z0;
    ring[q].ey_kp = ring[q].ey_k = ring[q].ey_kp_jm = ring[q].ey_k_jm = z0;
    load(ring[q], q);
  }
  __syncthreads();
  z2 prev = z0;
  for (int st0 = 0; st0 < nst; st0 += SGS_P) {
#pragma unroll
    for (int q = 0; q < SGS_P; q++) {
      const int st = st0 + q;
      if (st < nst) {
        const int i = (FWD ? dlo + st : dhi - st) - j;
        z2 nbv = z0;
        if (st > 0) {
          if (FWD) { if (t > 0) nbv = xch[((st - 1) & 1) * nt + t - 1]; }
          else { if (j + 1 <= n2 - 1) nbv = xch[((st - 1) & 1) * nt + t + 1]; }
        }
        if (jrow && i >= 1 && i <= n1 - 1) {
          St &c = ring[q];
          const z2 ez_ip = FWD ? z0 : prev, ez_im = FWD ? prev : c.ez_im;
          const z2 ez_jp = FWD ? z0 : nbv, ez_jm = FWD ? nbv : c.ez_jm;
          const double d1s_i = cd1s[i], d1s_im = cd1s[i - 1], d1_i = cd1[i];
          const double vol = a.fvm ? d1s_i * d2_j * d3_k : 0.0;
          z2 v = cell_ez(c.b, c.sig, d1s_i, d1s_im, d1_i, d2s_j, d2s_jm, d2_j, d3s_k, c.ex_kp, c.ex_k, c.ex_kp_im,
                         c.ex_k_im, ez_ip, ez_im, ez_jp, ez_jm, c.ey_kp, c.ey_k, c.ey_kp_jm, c.ey_k_jm, a.iwmu,
                         a.fvm, a.iomu, vol);
          Wzw[id3(n1, n2, k, j, i)] = v;
          if (!FWD && kop >= 0) yz[id3(n1, n2, kop, j, i)] = v;
          prev = v;
        }
        load(ring[q], st + SGS_P);
        xch[(st & 1) * nt + t] = prev;
        __syncthreads();
      }
    }
  }
  if (!FWD && kop >= 0) {
    for (int q = t; q < (n2 + 1) * (n1 + 1); q += nt) {
      int jj = q / (n1 + 1), i = q % (n1 + 1);
      if (!(jj >= 1 && jj <= n2 - 1 && i >= 1 && i <= n1 - 1)) yz[id3(n1, n2, kop, jj, i)] = zmk(0, 0);
    }
  }
}

// ----------------------------------------------------------- airwave FFT

// Stockham autosort pass structure shared with the oracle: radix-4 passes
// with span Ns = 1, 4, 16, ..., then one radix-2 pass if log2(N) is odd.
// Pass (R, Ns): for j in [0, N/R): k = j % Ns, inputs u[r] = a[j + r N/R]
// times tw[r k N/(R Ns)] (conjugated for the inverse; skipped when r*k==0),
// R-point DFT, outputs b[(j/Ns) R Ns + k + r Ns].
// `a` and `b` are LDS buffers of `cols` interleaved transforms of length N
// (element e of column c at e*cols + c).
template <int R>
__device__ __forceinline__ void fft_pass(const z2 *a, z2 *b, int N, int Ns, int cols, const z2 *tw, int dir)
{
  const int t = threadIdx.x, nt = blockDim.x;
  const int nb = N / R;
  for (int w = t; w < nb * cols; w += nt) {
    const int c = w % cols, j = w / cols;
    const int k = j % Ns;
    const int o = (j / Ns) * R * Ns + k;
    z2 u[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      u[r] = a[(j + r * nb) * cols + c];
      if (r != 0 && k != 0) {
        z2 tv = tw[(long)r * k * (N / (R * Ns))];
        if (dir > 0) tv.y = -tv.y;
        u[r] = zmul(u[r], tv);
      }
    }
    if (R == 4) {
      z2 a0 = zadd(u[0], u[2]), a1 = zsub(u[0], u[2]), a2 = zadd(u[1], u[R - 1]), a3 = zsub(u[1], u[R - 1]);
      z2 bb = dir < 0 ? zmk(a3.y, -a3.x) : zmk(-a3.y, a3.x);
      b[(o)*cols + c] = zadd(a0, a2);
      b[(o + Ns) * cols + c] = zadd(a1, bb);
      b[(o + 2 * Ns) * cols + c] = zsub(a0, a2);
      b[(o + 3 * Ns) * cols + c] = zsub(a1, bb);
    } else {
      b[(o)*cols + c] = zadd(u[0], u[1]);
      b[(o + Ns) * cols + c] = zsub(u[0], u[1]);
    }
  }
}

// Stockham autosort pass structure shared with the oracle: radix-4 passes
// with span Ns = 1, 4, 16, ..., then one radix-2 pass if log2(N) is odd.
// Pass (R, Ns): for j in [0, N/R): k = j % Ns, inputs u[r] = a[j + r N/R]
// times tw[r k N/(R Ns)] (conjugated for the inverse; skipped when r*k==0),
// R-point DFT, outputs b[(j/Ns) R Ns + k + r Ns].
// `a` and `b` are LDS buffers of `cols` interleaved transforms of length N
// (element e of column c at e*cols + c).  Returns the buffer holding the
// result; every thread of the block must call it.
__device__ __forceinline__ z2 *fft_lds(z2 *a, z2 *b, int N, int cols, const z2 *tw, int dir)
{
  for (int Ns = 1; Ns < N;) {
    if ((N / Ns) % 4 == 0) {
      fft_pass<4>(a, b, N, Ns, cols, tw, dir);
      Ns *= 4;
    } else {
      fft_pass<2>(a, b, N, Ns, cols, tw, dir);
      Ns *= 2;
    }
    __syncthreads();
    z2 *tmp = a; a = b; b = tmp;
  }
  return a;
}

struct AirArgs {
  int N, n1, n2;
  const z2 *hz;            // Hz(z=0), n1*n2
  z2 *buf, *bufx, *bufy;   // N*N spectra / inverse buffers
  const z2 *sH1, *sH2, *tw;
  const int *jx, *jy;
  const double *wx, *wy;
  const int *rows;         // inverse-row work list: rows of bufx then bufy
  int nrowx, nrowy;
  const int *gxj1, *gxj2, *gyj1, *gyj2;
  const double *gxw1, *gxw2, *gyw1, *gyw2;
  z2 *hxg, *hyg;
};

// Bilinear Hz(z=0) -> uniform row i2, then the forward transform of the row
// (airbc_init_free.c:137-153, rows first).
__global__ void k_air_rows_fwd(AirArgs a)
{
  extern __shared__ z2 sm[];
  const int N = a.N, i2 = blockIdx.x, n1 = a.n1;
  z2 *A = sm, *B = sm + N;
  const int j2 = a.jy[i2];
  const double w2 = a.wy[i2];
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) {
    const int j1 = a.jx[i1];
    z2 v = zmk(0, 0);
    if (j1 >= 0 && j2 >= 0) {
      const double w1 = a.wx[i1];
      const z2 *h = a.hz;
      v = zadd(zadd(zadd(zrmul((1. - w1) * (1. - w2), h[j1 + (long)n1 * j2]),
                         zrmul(w1 * (1. - w2), h[j1 + 1 + (long)n1 * j2])),
                    zrmul((1. - w1) * w2, h[j1 + (long)n1 * (j2 + 1)])),
               zrmul(w1 * w2, h[j1 + 1 + (long)n1 * (j2 + 1)]));
    }
    A[i1] = v;
  }
  __syncthreads();
  z2 *res = fft_lds(A, B, N, 1, a.tw, -1);
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) a.buf[i1 + (long)N * i2] = res[i1];
}

// Column pass: forward transform of columns c0..c0+CW-1 of buf, then for
// each multiplier (sH1 -> bufx, sH2 -> bufy) the product and the inverse
// column transform (airbc_init_free.c:153-161, 182-188).
template <int CW>
__global__ void k_air_cols(AirArgs a)
{
  extern __shared__ z2 sm[];
  const int N = a.N, c0 = blockIdx.x * CW;
  z2 *S = sm, *A = sm + N * CW, *B = sm + 2 * N * CW;
  for (int e = threadIdx.x; e < N * CW; e += blockDim.x) {
    int c = e % CW, i2 = e / CW;
    A[e] = a.buf[c0 + c + (long)N * i2];
  }
  __syncthreads();
  z2 *res = fft_lds(A, B, N, CW, a.tw, -1);
  for (int e = threadIdx.x; e < N * CW; e += blockDim.x) S[e] = res[e];
  __syncthreads();
  for (int q = 0; q < 2; q++) {
    const z2 *mul = q == 0 ? a.sH1 : a.sH2;
    z2 *dst = q == 0 ? a.bufx : a.bufy;
    for (int e = threadIdx.x; e < N * CW; e += blockDim.x) {
      int c = e % CW, i2 = e / CW;
      A[e] = zmul(S[e], mul[c0 + c + (long)N * i2]);
    }
    __syncthreads();
    res = fft_lds(A, B, N, CW, a.tw, +1);
    for (int e = threadIdx.x; e < N * CW; e += blockDim.x) {
      int c = e % CW, i2 = e / CW;
      dst[c0 + c + (long)N * i2] = res[e];
    }
    __syncthreads();
  }
}

// Inverse transform of the rows the back-interpolation reads (the rest of
// the inverse rows are never consumed).
__global__ void k_air_rows_inv(AirArgs a)
{
  extern __shared__ z2 sm[];
  const int N = a.N, w = blockIdx.x;
  z2 *buf = (w < a.nrowx) ? a.bufx : a.bufy;
  const long row = a.rows[w];
  z2 *A = sm, *B = sm + N;
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) A[i1] = buf[i1 + N * row];
  __syncthreads();
  z2 *res = fft_lds(A, B, N, 1, a.tw, +1);
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) buf[i1 + N * row] = res[i1];
}

// Normalise by N*N and interpolate back to the ghost Hx at (x1[i1], x2s[i2])
// and ghost Hy at (x1s[i1], x2[i2]) (airbc_init_free.c:162-207).
__global__ void k_air_ghost(AirArgs a)
{
  const int N = a.N, n1 = a.n1, n2 = a.n2;
  const double nn = (double)(N * N);
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = (n1 + 1) * n2, ny = n1 * (n2 + 1);
  if (t < nx) {
    int i1 = t % (n1 + 1), i2 = t / (n1 + 1);
    z2 v = zmk(0, 0);
    if (i1 < n1) {
      int j1 = a.gxj1[i1], j2 = a.gxj2[i2];
      double w1 = a.gxw1[i1], w2 = a.gxw2[i2];
      const z2 *b = a.bufx;
      v = zadd(zadd(zadd(zrmul((1. - w1) * (1. - w2), zdivr(b[j1 + (long)N * j2], nn)),
                         zrmul(w1 * (1. - w2), zdivr(b[j1 + 1 + (long)N * j2], nn))),
                    zrmul((1. - w1) * w2, zdivr(b[j1 + (long)N * (j2 + 1)], nn))),
               zrmul(w1 * w2, zdivr(b[j1 + 1 + (long)N * (j2 + 1)], nn)));
    }
    a.hxg[t] = v;
  } else if (t < nx + ny) {
    t -= nx;
    int i1 = t % n1, i2 = t / n1;
    z2 v = zmk(0, 0);
    if (i2 < n2) {
      int j1 = a.gyj1[i1], j2 = a.gyj2[i2];
      double w1 = a.gyw1[i1], w2 = a.gyw2[i2];
      const z2 *b = a.bufy;
      v = zadd(zadd(zadd(zrmul((1. - w1) * (1. - w2), zdivr(b[j1 + (long)N * j2], nn)),
                         zrmul(w1 * (1. - w2), zdivr(b[j1 + 1 + (long)N * j2], nn))),
                    zrmul((1. - w1) * w2, zdivr(b[j1 + (long)N * (j2 + 1)], nn))),
               zrmul(w1 * w2, zdivr(b[j1 + 1 + (long)N * (j2 + 1)], nn)));
    }
    a.hyg[t] = v;
  }
}

// Generic 2-D transform of an N x N buffer (tests and the public fft2).
__global__ void k_fft_rows(z2 *data, int N, const z2 *tw, int dir)
{
  extern __shared__ z2 sm[];
  z2 *A = sm, *B = sm + N;
  const long row = blockIdx.x;
  for (int i = threadIdx.x; i < N; i += blockDim.x) A[i] = data[i + N * row];
  __syncthreads();
  z2 *res = fft_lds(A, B, N, 1, tw, dir);
  for (int i = threadIdx.x; i < N; i += blockDim.x) data[i + N * row] = res[i];
}

template <int CW>
__global__ void k_fft_cols(z2 *data, int N, const z2 *tw, int dir)
{
  extern __shared__ z2 sm[];
  z2 *A = sm, *B = sm + N * CW;
  const int c0 = blockIdx.x * CW;
  for (int e = threadIdx.x; e < N * CW; e += blockDim.x) A[e] = data[c0 + e % CW + (long)N * (e / CW)];
  __syncthreads();
  z2 *res = fft_lds(A, B, N, CW, tw, dir);
  for (int e = threadIdx.x; e < N * CW; e += blockDim.x) data[c0 + e % CW + (long)N * (e / CW)] = res[e];
}

// -------------------------------------------------------------- reductions

// Partial sums of NP dot products sum a_q conj(b_q) over one chunk.
// term = (ar*br - ai*(-bi), ar*(-bi) + ai*br), as gcc forms a*conj(b).
template <int NP>
__device__ __forceinline__ void dot_terms(const z2 *const *A, const z2 *const *B, long e, double *acc)
{
#pragma unroll
  for (int q = 0; q < NP; q++) {
    z2 u = A[q][e], w = B[q][e];
    double c = w.x, d = -w.y;
    acc[2 * q] = acc[2 * q] + (u.x * c - u.y * d);
    acc[2 * q + 1] = acc[2 * q + 1] + (u.x * d + u.y * c);
  }
}

template <int NP>
__device__ __forceinline__ void chunk_fold_store(double *acc, double *partials, double *sh)
{
#pragma unroll
  for (int q = 0; q < 2 * NP; q++) {
    double v = block_fold(acc[q], sh);
    if (threadIdx.x == 0) partials[blockIdx.x * (2 * NP) + q] = v;
  }
}

struct DotArgs {
  long n;
  const z2 *a[2];
  const z2 *b[2];
  double *partials;
};

template <int NP>
__global__ __launch_bounds__(RED_T) void k_dot_partial(DotArgs g)
{
  __shared__ double sh[RED_T];
  double acc[2 * NP];
#pragma unroll
  for (int q = 0; q < 2 * NP; q++) acc[q] = 0.0;
  const long base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    long e = base + (long)m * RED_T;
    if (e < g.n) dot_terms<NP>(g.a, g.b, e, acc);
  }
  chunk_fold_store<NP>(acc, g.partials, sh);
}

// Second stage over chunk partials: thread t sums partials t, t+256, ...
// then the block fold.  `out` receives 2*NP doubles.
template <int NP>
__device__ __forceinline__ void final_fold(const double *partials, long nch, double *out, double *sh)
{
#pragma unroll
  for (int q = 0; q < 2 * NP; q++) {
    double s = 0.0;
    for (long c = threadIdx.x; c < nch; c += RED_T) s = s + partials[c * (2 * NP) + q];
    double v = block_fold(s, sh);
    if (threadIdx.x == 0) out[q] = v;
  }
}

template <int NP>
__global__ __launch_bounds__(RED_T) void k_dot_final(const double *partials, long nch, double *out)
{
  __shared__ double sh[RED_T];
  final_fold<NP>(partials, nch, out, sh);
}
