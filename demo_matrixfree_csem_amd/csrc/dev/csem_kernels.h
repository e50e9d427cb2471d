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
  const double *r1, *r2, *r3, *r1s, *r2s, *r3s;  // RN(1/d) of the same arrays
  const z2 *hxg, *hyg;  // ghost planes under the airwave boundary
  z2 iwmu, iomu;
  int n1, n2, n3, fvm, airbc;
  int nti, ntj, ntk;  // operator tiles along i, j, k
  long lEx, lEy;
};

// y = A x (fdfd.c:86-205).  A block is AP_NW waves over a tile of rows
// and columns and marches over KC planes: wave w covers row j = j0-1+w,
// lane l column i = i0-1+l.  Waves 0 and AP_NW-1 and lanes 0 and 63 are
// halos: they load E and compute H for their neighbours but write no
// output.  Per plane each lane loads its own E two planes ahead (and sigma
// one plane ahead), so every E value is read once per tile; the j+1
// neighbours of E and the j-1 neighbours of H come through LDS, the i+1
// and i-1 neighbours through whole-wave lane shifts.  H = curl E/(i w mu0)
// on the faces at (k, j, i) (fdfd.c:133-145) is evaluated once per face and
// carried to plane k+1 in registers.  The k-march is fully unrolled (a loop
// back edge would make the compiler wait for the prefetched loads).  Tiles
// are numbered columns fastest, then rows, then planes, and consecutive
// tiles run on one XCD, so halo rows and planes shared by neighbouring
// tiles are L2 hits.  Every division is by a 1-D spacing and uses its
// precomputed reciprocal with one exact correction (ddiv), so each value
// rounds as the reference's division does.
#define AP_NW 8   // waves per block
#define AP_OI 62  // output columns per block
#define AP_OJ (AP_NW - 2)  // output rows per block
template <int KC>
__global__ __launch_bounds__(64 * AP_NW) void k_apply(ApplyArgs a)
{
  __shared__ z2 sEx[AP_NW][64], sEz[AP_NW][64], sHx[AP_NW][64], sHz[AP_NW][64];
  __shared__ double sD[KC][4];  // d3s, 1/d3s, d3, 1/d3 of the tile's planes
  const int ntiles = a.nti * a.ntj * a.ntk, per = (ntiles + 7) / 8;
  const int T = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (T >= ntiles) return;
  const int ti = T % a.nti, tj = (T / a.nti) % a.ntj, tk = T / (a.nti * a.ntj);
  const int l = threadIdx.x, w = threadIdx.y;
  const int i = ti * AP_OI - 1 + l, j = tj * AP_OJ - 1 + w;
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  const int k0 = tk * KC, k1 = min(k0 + KC, n3 + 1);
  const bool live = i >= 0 && i <= n1 && j >= 0 && j <= n2;
  const bool out = live && l >= 1 && l <= AP_OI && w >= 1 && w <= AP_OJ;
  const bool hx_ok = live && j < n2, hy_ok = live && i < n1, hz_ok = hy_ok && j < n2;
  const bool iin = i >= 1 && i <= n1 - 1, jin = j >= 1 && j <= n2 - 1;
  const z2 *Ex = a.x, *Ey = a.x + a.lEx, *Ez = a.x + a.lEx + a.lEy;
  z2 *bx = a.y, *by = a.y + a.lEx, *bz = a.y + a.lEx + a.lEy;
  const z2 iwmu = a.iwmu, z0 = zmk(0, 0);
  // 1-D coefficients of this column and row (clamped reads off the grid)
  const int ic = min(max(i, 0), n1), jc = min(max(j, 0), n2);
  const double d1s_i = a.d1s[ic], r1s_i = a.r1s[ic], d1_i = a.d1[ic], r1_i = a.r1[ic];
  const double d2s_j = a.d2s[jc], r2s_j = a.r2s[jc], d2_j = a.d2[jc], r2_j = a.r2[jc];
  // plane strides and (clamped) in-plane offsets of the three layouts; for
  // every lane that stores, the clamped offset is its own
  const long px = (long)n1 * (n2 + 1), py = (long)(n1 + 1) * n2, pz = (long)(n1 + 1) * (n2 + 1);
  const long ox = min(ic, n1 - 1) + (long)n1 * jc, oy = ic + (long)(n1 + 1) * min(jc, n2 - 1),
             oz = ic + (long)(n1 + 1) * jc;
  struct EV { z2 x, y, z; };
  struct SV { double x, y, z; };
  auto ldE = [&](int k) {
    const long kc = max(k, 0);
    EV e;
    e.x = Ex[min(kc, (long)n3) * px + ox];
    e.y = Ey[min(kc, (long)n3) * py + oy];
    e.z = Ez[min(kc, (long)n3 - 1) * pz + oz];
    return e;
  };
  auto ldS = [&](int k) {
    SV v;
    v.x = a.s11[min(k, n3) * px + ox];
    v.y = a.s22[min(k, n3) * py + oy];
    v.z = a.s33[min(k, n3 - 1) * pz + oz];
    return v;
  };
  const int wu = min(w + 1, AP_NW - 1), wd = max(w - 1, 0);
  // the planes' spacings through LDS: read per plane as uniform values
  // they would be vector loads that the compiler waits for together with
  // the prefetched planes
  if (w == 0 && l < 4 * KC) {
    const int kq = min(k0 + l / 4, n3), c = l % 4;
    const double *src = c == 0 ? a.d3s : c == 1 ? a.r3s : c == 2 ? a.d3 : a.r3;
    sD[l / 4][c] = src[kq];
  }
  EV e[3];
  SV sg[2];
  e[2] = ldE(k0 - 1);
  e[0] = ldE(k0);
  e[1] = ldE(k0 + 1);
  sg[0] = ldS(k0);
  // H of the plane below k0; below k = 0 that is the airwave ghost plane
  z2 hx_prev = z0, hy_prev = z0;
  if (k0 == 0) {
    if (a.airbc) {
      if (hx_ok) hx_prev = a.hxg[ic + (long)(n1 + 1) * jc];
      if (hy_ok) hy_prev = a.hyg[ic + (long)n1 * jc];
    }
  } else {
    const int k = k0 - 1;
    sEz[w][l] = e[2].z;
    __syncthreads();
    const z2 ez_up = sEz[wu][l], ez_r = zdpp<DPP_WAVE_SHL1>(e[2].z);
    const double d3s_k = a.d3s[k], r3s_k = a.r3s[k];
    const z2 hx = zmul(zsub(zdivq(zsub(ez_up, e[2].z), d2s_j, r2s_j), zdivq(zsub(e[0].y, e[2].y), d3s_k, r3s_k)), iwmu);
    const z2 hy = zmul(zsub(zdivq(zsub(e[0].x, e[2].x), d3s_k, r3s_k), zdivq(zsub(ez_r, e[2].z), d1s_i, r1s_i)), iwmu);
    hx_prev = hx_ok ? hx : z0;
    hy_prev = hy_ok ? hy : z0;
    __syncthreads();
  }
  sEx[w][l] = e[0].x;
  sEz[w][l] = e[0].z;
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < KC; kk++) {
    const int k = k0 + kk;
    const EV cur = e[kk % 3], nxt = e[(kk + 1) % 3];
    e[(kk + 2) % 3] = ldE(k + 2);
    sg[(kk + 1) % 2] = ldS(k + 1);
    // H on the faces of plane k
    const double d3s_k = sD[kk][0], r3s_k = sD[kk][1], d3_k = sD[kk][2], r3_k = sD[kk][3];
    const z2 ez_up = sEz[wu][l], ex_up = sEx[wu][l];
    const z2 ey_r = zdpp<DPP_WAVE_SHL1>(cur.y), ez_r = zdpp<DPP_WAVE_SHL1>(cur.z);
    const z2 hx0 = zmul(zsub(zdivq(zsub(ez_up, cur.z), d2s_j, r2s_j), zdivq(zsub(nxt.y, cur.y), d3s_k, r3s_k)), iwmu);
    const z2 hy0 = zmul(zsub(zdivq(zsub(nxt.x, cur.x), d3s_k, r3s_k), zdivq(zsub(ez_r, cur.z), d1s_i, r1s_i)), iwmu);
    const z2 hz0 = zmul(zsub(zdivq(zsub(ey_r, cur.y), d1s_i, r1s_i), zdivq(zsub(ex_up, cur.x), d2s_j, r2s_j)), iwmu);
    const z2 hx = (hx_ok && k < n3) ? hx0 : z0;
    const z2 hy = (hy_ok && k < n3) ? hy0 : z0;
    const z2 hz = hz_ok ? hz0 : z0;
    __syncthreads();
    sHx[w][l] = hx;
    sHz[w][l] = hz;
    sEx[w][l] = nxt.x;
    sEz[w][l] = nxt.z;
    __syncthreads();
    const z2 hz_dn = sHz[wd][l], hx_dn = sHx[wd][l];
    const z2 hy_l = zdpp<DPP_WAVE_SHR1>(hy), hz_l = zdpp<DPP_WAVE_SHR1>(hz);
    if (out && k < k1) {
      const SV sv = sg[kk % 2];
      const bool kin = (k >= 1 && k <= n3 - 1) || (a.airbc && k == 0);
      if (i < n1) {  // bx: i < n1, j <= n2, k <= n3
        z2 v = z0;
        if (jin && kin) {
          const z2 t1 = zdivq(zsub(hz, hz_dn), d2_j, r2_j);
          const z2 t2 = zdivq(zsub(hy, hy_prev), d3_k, r3_k);
          v = zsub(zsub(t1, t2), zrmul(sv.x, cur.x));
          if (a.fvm) v = zmul(v, zrmul(d1s_i * d2_j * d3_k, a.iomu));
        }
        bx[k * px + ox] = v;
      }
      if (j < n2) {  // by: i <= n1, j < n2, k <= n3
        z2 v = z0;
        if (iin && kin) {
          const z2 t1 = zdivq(zsub(hx, hx_prev), d3_k, r3_k);
          const z2 t2 = zdivq(zsub(hz, hz_l), d1_i, r1_i);
          v = zsub(zsub(t1, t2), zrmul(sv.y, cur.y));
          if (a.fvm) v = zmul(v, zrmul(d1_i * d2s_j * d3_k, a.iomu));
        }
        by[k * py + oy] = v;
      }
      if (k < n3) {  // bz: i <= n1, j <= n2, k < n3
        z2 v = z0;
        if (iin && jin) {
          const z2 t1 = zdivq(zsub(hy, hy_l), d1_i, r1_i);
          const z2 t2 = zdivq(zsub(hx, hx_dn), d2_j, r2_j);
          v = zsub(zsub(t1, t2), zrmul(sv.z, cur.z));
          if (a.fvm) v = zmul(v, zrmul(d1_i * d2_j * d3s_k, a.iomu));
        }
        bz[k * pz + oz] = v;
      }
    }
    hx_prev = hx;
    hy_prev = hy;
  }
}

// H = curl E / (i w mu0) on all faces (fdfd.c:125-149), Hx | Hy | Hz.
#ifndef CSEM_TYPES_ONLY
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
#endif

// Hz on the plane k = 0 only (the airwave input), idz(0,j,i) = i + n1*j.
#ifndef CSEM_TYPES_ONLY
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
#endif

// ----------------------------------------------------------------- LU-SGS
//
// lusgs_apply (fdfd.c:287-510): forward Gauss-Seidel over Ex, Ey, Ez in
// lexicographic order, then backward over Ez, Ey, Ex in reverse order, from
// a zero state, on the airwave-extended grid (n3f = n3 + nb planes).
//
// Each component sweep decouples into independent 2-D problems (slices)
// along the component's own axis: Ex couples to Ex only through (k, j+-1)
// and (k+-1, j) (fdfd.c:358-361), Ey through (k+-1, i) and (k, i+-1), Ez
// through (j+-1, i) and (j, i+-1).  One workgroup owns a slice; thread t
// owns slow index s = 1+t and visits fast index f = d - s on diagonal d, so
// every cell sees exactly the neighbours the lexicographic loop gives it:
// the new value of (s, f-1) from its own register and of (s-1, f) from
// thread t-1 one diagonal earlier, through LDS (the roles of +1 and -1
// swap in a backward sweep).
//
// Every cell update has the shape (fdfd.c:357-369, 383-395, 409-421)
//   t1 = P1 - Q1, t2 = P2 - Q2, t3 = P3 - Q3, t4 = P4 - Q4,
//   aij = ((t1 - t2)/dA - (t3 - t4)/dB) * iwmu - sigma*0,
//   aii = c*iwmu - sigma,   E = (b - aij)/aii,
// where the component's own neighbours along axis A sit in Q1 (+) and Q2
// (-), along axis B in P3 (+) and P4 (-), and the two other components in
// P1, P2, Q3, Q4:
//   Ex: A = j (fast), B = k (slow);  P1, P2 from Ey;  Q3, Q4 from Ez
//   Ey: A = k (slow), B = i (fast);  P1, P2 from Ez;  Q3, Q4 from Ex
//   Ez: A = i (fast), B = j (slow);  P1, P2 from Ex;  Q3, Q4 from Ey
// A forward sweep sees 0 for the + neighbours and new values for the -
// ones; a backward sweep sees forward values for the - neighbours and new
// values for the + ones.  Everything except the new values is known when a
// sweep starts, so a parallel pre-pass evaluates it with the same
// expressions into a record per cell,
//   forward:  R = (t1, P2, t3, Q4)        backward: R = (P1, t2, Q3, t4),
// and the diagonal's half of Smith's division (ratio, denominator, its
// reciprocal, branch) is evaluated once per frequency.  The wavefront then
// computes only the dependent remainder, in the reference's order, so each
// intermediate rounds as in the reference.  sigma*0 is +0 for the finite,
// non-negative conductivities the model accepts and x - (+0) == x, so that
// term is dropped.  Divisions by spacings use ddiv.
//
// Layout: a sweep's per-cell inputs (record, right-hand side, diagonal)
// are stored diagonal-major ("skewed") in 16-byte planes: slice q, then
// one plane per 16-byte unit of the record (r0..r3, b) or of the Smith part
// (two), then diagonal d = s + f, then row s, AP slots per diagonal (row s
// at slot s - 1).  The lanes of a wave (consecutive rows of one diagonal)
// load each unit as one contiguous run, every unit of a step from the same
// lane offset, and a lane's next diagonal is a fixed stride away; the
// slots of rows not on a diagonal are never written or read.  The state W (forward
// values, then the backward values of Ey and Ez) and the result y keep the
// reference layout; the pre-passes gather their neighbours from W.

// Sweep geometry of component C: slice q, slow row s in 1..A, fast f in 1..B.
//   Ex: q = i (n1 slices), s = k (A = n3f-1), f = j (B = n2-1)
//   Ey: q = j (n2 slices), s = k (A = n3f-1), f = i (B = n1-1)
//   Ez: q = k (n3f slices), s = j (A = n2-1), f = i (B = n1-1)
struct SkewGeo {
  int A[3], B[3], ns[3];
  long base[3];  // component offsets in the diagonal cache (the Smith part)
  int AP[3];     // FP64 sweep slots per diagonal: A rounded up to a whole wave
  long SL[3];    // FP64 sweep slots per slice: (A + B - 1) diagonals of AP
};

__host__ __device__ __forceinline__ int sk_tri(int m) { return m > 0 ? m * (m + 1) / 2 : 0; }
// cells on diagonals 2..d-1 of an A x B slice
__host__ __device__ __forceinline__ int sk_off(int d, int A, int B)
{
  const int n = d - 2;
  return sk_tri(n) - sk_tri(n - A) - sk_tri(n - B);
}

struct LuRec { z2x4 R; z2 b; };  // 80 B per swept cell

// The record planes of sweep (C, FWD): NP per slice, unit u in plane u
// (Ex forward: b alone, in plane 0); W01: whether units 0 and 1 are written
// (Ey forward sees Ez = 0, so its t1 and the known part of t2 are +0 and
// its sweep does not read them).
template <int C, bool FWD> struct LuPlanes {
  static constexpr int NP = (C == 0 && FWD) ? 1 : 5;
  static constexpr bool W01 = !FWD || C == 2;
};

struct LuArgs {
  const z2 *x;  // right-hand side, operator grid
  z2 *W;        // sweep state on the full grid, Ex | Ey | Ez
  z2 *y;        // result, operator grid
  LuRec *rec;   // pre-pass records of the current sweep, skewed
  const double4 *smith;  // (ratio, |denom|, |1/denom|, code), skewed Ex | Ey | Ez
  const double *s11, *s22, *s33;
  const double *d1, *d2, *d3, *d1s, *d2s, *d3s;
  const double *r1, *r2, *r3, *r1s, *r2s, *r3s;
  z2 iwmu, iomu;
  int fvm, n1, n2, n3f, kofs, n3op;
  long fEx, fEy, lEx, lEy;
  SkewGeo g;
};

template <int C>
__device__ __forceinline__ void lu_kji(int q, int s, int f, int &k, int &j, int &i)
{
  if (C == 0) { i = q; k = s; j = f; }
  else if (C == 1) { j = q; k = s; i = f; }
  else { k = q; j = s; i = f; }
}

// Thread -> (slice q, skewed position p) -> cell; false past the slice.
template <int C>
__device__ __forceinline__ bool lu_decode(const LuArgs &a, int &q, int &s, int &f, int &p)
{
  const int A = a.g.A[C], B = a.g.B[C];
  q = blockIdx.y;
  p = blockIdx.x * 256 + threadIdx.x;
  if (p >= A * B) return false;
  int lo = 2, hi = A + B;  // largest d with sk_off(d) <= p
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sk_off(mid, A, B) <= p) lo = mid; else hi = mid - 1;
  }
  s = max(1, lo - B) + (p - sk_off(lo, A, B));
  f = lo - s;
  return true;
}

// Per-frequency diagonal of component C (the aii lines fdfd.c:363, 389,
// 415 with the fvm scaling of 364-368) reduced to Smith's first half
// (zdiv in csem_common.h): ratio, |denom|, |1/denom| and a code whose sign
// is the sign of denom and whose magnitude is 2 for the |c| < |d| branch.
template <int C>
__global__ __launch_bounds__(256) void k_lu_smith(LuArgs a)
{
  int q, s, f, k, j, i, p;
  if (!lu_decode<C>(a, q, s, f, p)) return;
  lu_kji<C>(q, s, f, k, j, i);
  const double *d1 = a.d1, *d2 = a.d2, *d3 = a.d3, *d1s = a.d1s, *d2s = a.d2s, *d3s = a.d3s;
  const int n1 = a.n1, n2 = a.n2;
  double c, sig;
  if (C == 0) {
    c = (1. / d2s[j] + 1. / d2s[j - 1]) / d2[j] + (1. / d3s[k] + 1. / d3s[k - 1]) / d3[k];
    sig = a.s11[id1(n1, n2, k, j, i)];
  } else if (C == 1) {
    c = (1. / d3s[k] + 1. / d3s[k - 1]) / d3[k] + (1. / d1s[i] + 1. / d1s[i - 1]) / d1[i];
    sig = a.s22[id2(n1, n2, k, j, i)];
  } else {
    c = (1. / d1s[i] + 1. / d1s[i - 1]) / d1[i] + (1. / d2s[j] + 1. / d2s[j - 1]) / d2[j];
    sig = a.s33[id3(n1, n2, k, j, i)];
  }
  z2 aii = zsubr(zrmul(c, a.iwmu), sig);
  if (a.fvm) aii = zmul(aii, zrmul(d1s[i] * d2[j] * d3[k], a.iomu));
  const double cc = aii.x, dd = aii.y;
  double ratio, denom, code;
  if (fabs(cc) < fabs(dd)) { ratio = cc / dd; denom = (cc * ratio) + dd; code = 2; }
  else { ratio = dd / cc; denom = (dd * ratio) + cc; code = 1; }
  if (denom < 0) code = -code;
  const long SL = a.g.SL[C];
  z2 *S = (z2 *)a.smith + 2 * (a.g.base[C] + (long)q * SL) + (long)(s + f - 2) * a.g.AP[C] + (s - 1);
  S[0] = zmk(ratio, fabs(denom));
  S[SL] = zmk(fabs(1.0 / denom), code);
}

// Second half of zdiv: n / aii from the stored diagonal.  Both branches
// are formed and selected, so a wave never splits; x/denom is evaluated as
// (sign*x)/|denom|, which is the same IEEE quotient.
__device__ __forceinline__ z2 smith_finish(z2 n, double4 S)
{
  const bool f = fabs(S.w) > 1.5;
  const double sg = S.w < 0 ? -1.0 : 1.0;
  const double u = f ? n.x : n.y, w = f ? n.y : n.x;
  const double p = w * S.x;
  const double re = (u * S.x) + w;
  const double im = f ? p - u : u - p;
  return zmk(ddiv(re * sg, S.y, S.z), ddiv(im * sg, S.y, S.z));
}

// Value of component C at (k, j, i) of the full-grid state.
#define WX(k, j, i) a.W[id1(n1, n2, k, j, i)]
#define WY(k, j, i) a.W[a.fEx + id2(n1, n2, k, j, i)]
#define WZ(k, j, i) a.W[a.fEx + a.fEy + id3(n1, n2, k, j, i)]

// Pre-pass of sweep (C, FWD): the record R and the right-hand side of every
// swept cell.  A block computes a tile of cells in the reference layout
// (x fastest, so the neighbour gathers from W coalesce), places each record
// in LDS at its position in the tile's skewed order together with its
// global index, and stores the tile's records as consecutive 16-byte units
// (each diagonal's run of the tile is contiguous in the skewed layout).  In
// the forward pass, components swept later are still zero (fdfd.c:338-343);
// W holds forward values, overwritten by the backward values of Ez and Ey
// as those sweeps finish.  Tiles: Ex 16 slices x 4 rows x 8 columns (its
// slices run along x), Ey and Ez 8 rows x 64 columns of one slice.
template <int C> struct LuPreTile { static constexpr int TQ = 1, TS = 8, TF = 64; };
template <> struct LuPreTile<0> { static constexpr int TQ = 16, TS = 4, TF = 8; };
template <int C, bool FWD>
__global__ __launch_bounds__(256) void k_lu_pre(LuArgs a)
{
  using TL = LuPreTile<C>;
  constexpr int TQ = TL::TQ, TS = TL::TS, TF = TL::TF, NC = TQ * TS * TF, NU = 5;  // 16-byte units per record
  __shared__ z2 tile[NC * NU];
  __shared__ long pidx[NC];  // global record index of each slot, -1 off the slices
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3f;
  const int A = a.g.A[C], B = a.g.B[C], ns = a.g.ns[C], AP = a.g.AP[C];
  const long SL = a.g.SL[C];
  const int t = threadIdx.x;
  int q0, s0, f0;
  if (C == 0) { q0 = blockIdx.x * TQ; f0 = 1 + blockIdx.y * TF; s0 = 1 + blockIdx.z * TS; }
  else { f0 = 1 + blockIdx.x * TF; s0 = 1 + blockIdx.y * TS; q0 = blockIdx.z; }
  const z2 z0 = zmk(0, 0);
#pragma unroll
  for (int it = 0; it < NC / 256; it++) {
    int lq, ls, lf;
    if (C == 0) { lq = t % TQ; lf = (t / TQ) % TF; ls = t / (TQ * TF) + it * (256 / (TQ * TF)); }
    else { lq = 0; lf = t % TF; ls = t / TF + it * (256 / TF); }
    const int q = q0 + lq, s = s0 + ls, f = f0 + lf;
    const int dl = ls + lf + 2;
    const int cell = lq * TS * TF + sk_off(dl, TS, TF) + (ls + 1 - max(1, dl - TF));
    if (q >= ns || s > A || f > B) {
      pidx[cell] = -1;
      continue;
    }
    pidx[cell] = q * LuPlanes<C, FWD>::NP * SL + (long)(s + f - 2) * AP + (s - 1);
    int k, j, i;
    lu_kji<C>(q, s, f, k, j, i);
    const int kop = k - a.kofs;
    const long xo = (C == 0) ? id1(n1, n2, kop, j, i) : (C == 1) ? a.lEx + id2(n1, n2, kop, j, i)
                                                                   : a.lEx + a.lEy + id3(n1, n2, kop, j, i);
    z2 *out = tile + cell * NU;
    out[4] = kop >= 0 ? a.x[xo] : z0;
    if (C == 0 && FWD) continue;  // Ex forward: Ey = Ez = 0, every record term is +0
    z2 P1, P2, Q2, Q3, P4, Q4;
    if (C == 0) {  // fdfd.c:483-486: Ey, Ez backward, Ex forward
      const double d1s_i = a.d1s[i], r1s_i = a.r1s[i];
      P1 = zdivq(zsub(WY(k, j, i + 1), WY(k, j, i)), d1s_i, r1s_i);
      P2 = zdivq(zsub(WY(k, j - 1, i + 1), WY(k, j - 1, i)), d1s_i, r1s_i);
      Q2 = zdivq(zsub(z0, WX(k, j - 1, i)), a.d2s[j - 1], a.r2s[j - 1]);
      Q3 = zdivq(zsub(WZ(k, j, i + 1), WZ(k, j, i)), d1s_i, r1s_i);
      P4 = zdivq(zsub(z0, WX(k - 1, j, i)), a.d3s[k - 1], a.r3s[k - 1]);
      Q4 = zdivq(zsub(WZ(k - 1, j, i + 1), WZ(k - 1, j, i)), d1s_i, r1s_i);
    } else if (C == 1) {  // fdfd.c:384-387 / 458-461: Ez (zero forward, backward after), Ex forward
      const double d2s_j = a.d2s[j], r2s_j = a.r2s[j];
      const z2 ez_jp = FWD ? z0 : WZ(k, j + 1, i), ez_j = FWD ? z0 : WZ(k, j, i);
      const z2 ez_km_jp = FWD ? z0 : WZ(k - 1, j + 1, i), ez_km_j = FWD ? z0 : WZ(k - 1, j, i);
      P1 = zdivq(zsub(ez_jp, ez_j), d2s_j, r2s_j);
      P2 = zdivq(zsub(ez_km_jp, ez_km_j), d2s_j, r2s_j);
      Q3 = zdivq(zsub(WX(k, j + 1, i), WX(k, j, i)), d2s_j, r2s_j);
      Q4 = zdivq(zsub(WX(k, j + 1, i - 1), WX(k, j, i - 1)), d2s_j, r2s_j);
      if (!FWD) {
        Q2 = zdivq(zsub(z0, WY(k - 1, j, i)), a.d3s[k - 1], a.r3s[k - 1]);
        P4 = zdivq(zsub(z0, WY(k, j, i - 1)), a.d1s[i - 1], a.r1s[i - 1]);
      }
    } else {  // fdfd.c:410-413 / 435-438: Ex, Ey forward
      const double d3s_k = a.d3s[k], r3s_k = a.r3s[k];
      P1 = zdivq(zsub(WX(k + 1, j, i), WX(k, j, i)), d3s_k, r3s_k);
      P2 = zdivq(zsub(WX(k + 1, j, i - 1), WX(k, j, i - 1)), d3s_k, r3s_k);
      Q3 = zdivq(zsub(WY(k + 1, j, i), WY(k, j, i)), d3s_k, r3s_k);
      Q4 = zdivq(zsub(WY(k + 1, j - 1, i), WY(k, j - 1, i)), d3s_k, r3s_k);
      if (!FWD) {
        Q2 = zdivq(zsub(z0, WZ(k, j, i - 1)), a.d1s[i - 1], a.r1s[i - 1]);
        P4 = zdivq(zsub(z0, WZ(k, j - 1, i)), a.d2s[j - 1], a.r2s[j - 1]);
      }
    }
    if (FWD) {  // Q1 = (0 - 0)/d = +0 and P3 = +0
      out[0] = zsub(P1, z0); out[1] = P2; out[2] = zsub(z0, Q3); out[3] = Q4;
    } else {
      out[0] = P1; out[1] = zsub(P2, Q2); out[2] = Q3; out[3] = zsub(P4, Q4);
    }
  }
  __syncthreads();
  // store the tile's records plane by plane, each diagonal's run of the
  // tile contiguous (Ex forward: b alone; Ey forward: not the +0 units its
  // sweep does not read)
  z2 *rec = (z2 *)a.rec;
  using LP = LuPlanes<C, FWD>;
#pragma unroll
  for (int c = 0; c < NU; c++) {
    if ((C == 0 && FWD && c != 4) || (!LP::W01 && c < 2)) continue;
    const long po = (LP::NP == 1) ? 0 : (long)c * SL;
#pragma unroll
    for (int it = 0; it < NC / 256; it++) {
      const int e = t + it * 256;
      const long p = pidx[e];
      if (p >= 0) rec[p + po] = tile[e * NU + c];
    }
  }
}
#undef WX
#undef WY
#undef WZ

#define SGS_MAXT 256  // rows per slice: n3f-1 and n2-1 must not exceed this

// The sweep of one slice.  One thread per row s = 1 + t visits the fast
// index f = d - s on diagonal d (the roles of +1 and -1 swap in a backward
// sweep); per step it forms the cell's row from the record, the new value
// of (s, f-1) in its own register and the new value of (s-1, f) of the
// neighbouring row (fdfd.c:357-369, 383-395, 409-421, in that expression
// order).  A diagonal costs no LDS round trip and no barrier:
//   - the neighbour row's new value (s - 1, f) (forward; (s + 1, f)
//     backward) comes from the neighbouring lane by a whole-wave DPP lane
//     shift of its previous value; the shift's edge lane keeps the value
//     handed over by the neighbouring wave;
//   - consecutive waves run SG2_L diagonals apart (in the sweep's direction
//     the wave whose rows come first leads); the row at a wave's edge goes
//     through a small LDS ring, written at the step it is computed and read
//     SG2_PD steps before it is used, and a barrier every SG2_BP steps
//     orders every write before its read (SG2_L >= SG2_BP + SG2_PD - 1; a
//     slot is rewritten only SG2_RING steps later, while the barriers keep
//     every wave within SG2_BP steps of the others);
//   - each step's record and diagonal are loaded SG2_P steps ahead into
//     registers by raw buffer loads at a fixed lane offset plus a fixed
//     stride per step (a lane whose cell is off the slice gets an offset
//     past the buffer: the load reads 0 and issues no traffic); its
//     fast-axis spacings SG2_PC steps ahead from an LDS table padded with
//     zeros, so that no index is clamped;
//   - a lane off the slice computes with zeros: before its row starts every
//     input it sees (record, diagonal, spacings or its own ones, the values
//     of its row and of the neighbouring row, which has not started
//     either) is zero and every operation of the update maps +0 operands to
//     +0, so the row's boundary value stays the reference's +0 without a
//     mask; after its row ends, or on a row past A in a forward sweep, it
//     computes values no cell of the slice reads (in a backward sweep the
//     rows past A, whose values row A reads, see only zeros);
//   - the values do not go to memory from the computing waves: a wave's
//     vector-memory counter retires loads and stores in issue order, so
//     waiting for a prefetched record would also wait for every earlier
//     store to complete.  Each step's values go to an LDS staging area
//     instead (two halves of SG2_BP steps), and one more wave of the block
//     stores each half to W and y (raw buffer stores, dropped for lanes off
//     the slice) while the computing waves fill the other half.
// All waves run the same number of steps and meet at the same barriers.
#define SG2_BP 8
#define SG2_PD 4
#define SG2_L (SG2_BP + SG2_PD - 1)
#define SG2_RING 16
#define SG2_PC 4
#define SG2_U 48
#define SG2_OOB ((int)0x80000000)  // a buffer offset past every buffer (all shorter than 2 GiB)
#define SG2_SP 66  // staging slots per step: 64 rows, padded so that the store wave's reads of eight
                   // steps of eight rows (Ey, Ez) meet no LDS bank conflict
static_assert(SG2_RING > SG2_L - SG2_PD + SG2_BP, "edge ring slot rewritten before it is read");
static_assert(SG2_U % SG2_RING == 0 && SG2_U % SG2_BP == 0 && SG2_U % SG2_PD == 0 && SG2_U % SG2_PC == 0,
              "sweep unroll");
// Record units a sweep loads per step (b and the Smith part always) and
// how many steps ahead it loads them.  Ex forward sees Ey = Ez = 0, so its
// record terms are +0 and only b is read; Ey forward sees Ez = 0, so its
// first two terms (t1 and the known part of t2) are +0.  The fewer
// registers a step's loads take, the deeper the prefetch.
template <int C, bool FWD> struct Sg2Load {
  static constexpr bool R01 = !FWD || C == 2, R23 = !FWD || C != 0;
  static constexpr int P = (R01 && R23) ? 6 : 8;
  static constexpr int NP = LuPlanes<C, FWD>::NP, BP = (NP == 1) ? 0 : 4;  // planes per slice, b's plane
  static_assert(R01 == LuPlanes<C, FWD>::W01 || !R01, "the sweep reads only units its pre-pass writes");
  static_assert(SG2_U % P == 0, "sweep unroll");
};
typedef unsigned sg2_v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t sg2_rsrc(const void *p, long len)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)len, 0x00020000);
}
__device__ __forceinline__ z2 sg2_ld(__amdgpu_buffer_rsrc_t r, int ofs, int sofs = 0)
{
  return __builtin_bit_cast(z2, __builtin_amdgcn_raw_buffer_load_b128(r, ofs, sofs, 0));
}
__device__ __forceinline__ void sg2_st(z2 v, __amdgpu_buffer_rsrc_t r, int ofs)
{
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(sg2_v4u, v), r, ofs, 0, 0);
}
// lane shift of a complex double; lanes with no source lane keep `old`
template <int CTRL>
__device__ __forceinline__ double sg2_dpp_keep(double old, double v)
{
  const unsigned long long o = (unsigned long long)__double_as_longlong(old);
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32), CTRL,
                                                            0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int CTRL>
__device__ __forceinline__ z2 sg2_zdpp_keep(z2 old, z2 v)
{
  return zmk(sg2_dpp_keep<CTRL>(old.x, v.x), sg2_dpp_keep<CTRL>(old.y, v.y));
}

// The fast-axis table: entry u holds column f = u - sg2_toff(nw) (zeros
// off [1, B]); the offset and length cover every column a lane of any wave
// reads from its first step to the last prefetch of the last loop
// iteration, forward and backward.
__host__ __device__ __forceinline__ int sg2_toff(int nw)
{
  return 64 * nw + (nw - 1) * SG2_L + SG2_U + SG2_PC + SG2_BP + 4;
}
__host__ __device__ __forceinline__ int sg2_tlen(int A, int B, int nw)
{
  return sg2_toff(nw) + A + B + (nw - 1) * SG2_L + SG2_BP + SG2_U + SG2_PC + 4;
}
// LDS of a block of nw computing waves: the edge rings and the junk slots
// the other lanes' ring stores land in, the value staging (2 halves x nw
// waves x SG2_BP steps x SG2_SP slots) and the fast table (double4 spacings
// and a double cell length per entry)
__host__ __device__ __forceinline__ long sg2_lds_bytes(int A, int B, int nw)
{
  return (long)nw * (2 * SG2_RING + 64) * 16 + 2L * nw * SG2_BP * SG2_SP * 16 + (long)sg2_tlen(A, B, nw) * 40;
}

// Backward sweeps: the operator-grid cells of slice q outside the sweep
// range are zero rows of the result.
template <int C>
__device__ void sg2_zero_frame(const LuArgs &a, int q, int t, int nt)
{
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3f;
  const z2 z0 = zmk(0, 0);
  if (C == 0) {
    const int nk = a.n3op + 1;
    for (int w = t; w < nk * (n2 + 1); w += nt) {
      const int kq = w / (n2 + 1), j = w % (n2 + 1), kf = kq + a.kofs;
      if (!(kf >= 1 && kf <= n3 - 1 && j >= 1 && j <= n2 - 1)) a.y[id1(n1, n2, kq, j, q)] = z0;
    }
  } else if (C == 1) {
    const int nk = a.n3op + 1;
    for (int w = t; w < nk * (n1 + 1); w += nt) {
      const int kq = w / (n1 + 1), i = w % (n1 + 1), kf = kq + a.kofs;
      if (!(kf >= 1 && kf <= n3 - 1 && i >= 1 && i <= n1 - 1)) a.y[a.lEx + id2(n1, n2, kq, q, i)] = z0;
    }
  } else if (q - a.kofs >= 0) {
    for (int w = t; w < (n2 + 1) * (n1 + 1); w += nt) {
      const int jj = w / (n1 + 1), i = w % (n1 + 1);
      if (!(jj >= 1 && jj <= n2 - 1 && i >= 1 && i <= n1 - 1))
        a.y[a.lEx + a.lEy + id3(n1, n2, q - a.kofs, jj, i)] = z0;
    }
  }
}

template <int C, bool FWD, bool FVM, int NW>
__global__ __launch_bounds__(64 * (NW + 1)) void k_sgs2(LuArgs a)
{
  extern __shared__ z2 xs2[];
  const int nslice = a.g.ns[C];
  const int per = (nslice + 7) / 8;
  const int q = (blockIdx.x % 8) * per + blockIdx.x / 8;  // consecutive slices on one XCD
  if (q >= nslice) return;
  const int n1 = a.n1, n2 = a.n2;
  constexpr int nt = 64 * (NW + 1), lagmax = (NW - 1) * SG2_L;  // wave NW stores
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int A = a.g.A[C], B = a.g.B[C];
  const int nst = A + B - 1;
  const int s = 1 + t, sc = min(s, A);
  constexpr int AP = 64 * NW;  // slots per diagonal (the host sets a.g.AP[C] to this)
  const int OFF = sg2_toff(NW), TL = sg2_tlen(A, B, NW);
  z2 *ring = xs2;                               // [NW][SG2_RING]
  z2 *junk = xs2 + NW * SG2_RING;               // [NW][SG2_RING + 64]
  z2 *stg = junk + NW * (SG2_RING + 64);        // [2][NW][SG2_BP][SG2_SP]
  double4 *ftab = (double4 *)(stg + 2 * NW * SG2_BP * SG2_SP);  // [TL]
  double *vtab = (double *)(ftab + TL);                     // [TL]
  {
    const double *g0 = (C == 0) ? a.d2s : a.d1s, *g1 = (C == 0) ? a.d2 : a.d1;
    const double *g2 = (C == 0) ? a.r2s : a.r1s, *g3 = (C == 0) ? a.r2 : a.r1;
    for (int u = t; u < TL; u += nt) {
      const int f = u - OFF, fs = FWD ? f - 1 : f;
      const bool in = f >= 1 && f <= B;
      ftab[u] = in ? make_double4(g0[fs], g2[fs], g1[f], g3[f]) : make_double4(0, 0, 0, 0);
      vtab[u] = in ? ((C == 0) ? a.d2[f] : a.d1s[f]) : 0.0;
    }
    for (int u = t; u < NW * SG2_RING; u += nt) ring[u] = zmk(0, 0);
  }
  const double *h0 = (C == 2) ? a.d2s : a.d3s, *h1 = (C == 2) ? a.d2 : a.d3;
  const double *h2 = (C == 2) ? a.r2s : a.r3s, *h3 = (C == 2) ? a.r2 : a.r3;
  const int ssi = FWD ? sc - 1 : sc;  // slow-axis d*s of the swept neighbour
  const double ssx = h0[ssi], rssx = h2[ssi], sd = h1[sc], rsd = h3[sc];
  const double vA = (C == 0) ? a.d1s[q] : (C == 1) ? a.d2[q] : a.d2[sc];
  const double vB = (C == 2) ? a.d3[q] : a.d3[sc];
  // this slice's records and diagonals as raw buffers
  const long SL = a.g.SL[C];
  using LD = Sg2Load<C, FWD>;
  const __amdgpu_buffer_rsrc_t Rb = sg2_rsrc((const z2 *)a.rec + (long)q * LD::NP * SL, LD::NP * SL * 16L);
  const __amdgpu_buffer_rsrc_t Sb = sg2_rsrc((const z2 *)a.smith + 2 * (a.g.base[C] + (long)q * SL), 2 * SL * 16L);
  const int PS = (int)SL * 16;  // bytes per plane
  // natural offset of cell (q, s, f) = nb + f*FS within the component
  const long PL = (C == 0) ? (long)n1 * (n2 + 1) : (C == 1) ? (long)(n1 + 1) * n2 : (long)(n1 + 1) * (n2 + 1);
  const long SS = (C == 2) ? (long)(n1 + 1) : PL;
  const int FS = (C == 0) ? n1 : 1;
  const long nb = ((C == 0) ? (long)q : (C == 1) ? (long)(n1 + 1) * q : PL * q) + sc * SS;
  const long wofs = (C == 0) ? 0 : (C == 1) ? a.fEx : a.fEx + a.fEy;
  const long xofs = (C == 0) ? 0 : (C == 1) ? a.lEx : a.lEx + a.lEy;
  const long wlen = (C == 0) ? a.fEx : (C == 1) ? a.fEy : (long)(n1 + 1) * (n2 + 1) * a.n3f;
  const long ylen = (C == 0) ? a.lEx : (C == 1) ? a.lEy : (long)(n1 + 1) * (n2 + 1) * a.n3op;
  const __amdgpu_buffer_rsrc_t Wb = sg2_rsrc(a.W + wofs, wlen * (long)sizeof(z2));
  const __amdgpu_buffer_rsrc_t Yb = sg2_rsrc(a.y + xofs, ylen * (long)sizeof(z2));
  const int kop = ((C == 2) ? q : sc) - a.kofs;
  const long ynb = nb - (long)a.kofs * PL;
  const int lag = (FWD ? wv : NW - 1 - wv) * SG2_L;
  const int nex = ((nst + lagmax + SG2_BP - 1) / SG2_BP) * SG2_BP;  // steps run
  const int ntk = ((nex + SG2_U - 1) / SG2_U) * SG2_U;
  const z2 z0 = zmk(0, 0);
  if (wv == NW) {
    // the store wave: after the barrier that ends half h (steps
    // SG2_BP h .. SG2_BP h + SG2_BP - 1), store that half's values while the
    // computing waves fill the other half
    __syncthreads();  // the tables and rings are ready
    const int nhalf = nex / SG2_BP;
    // Offsets are precomputed per lane (f - 1 of the lane's cell at half 0,
    // past any column on a row past A; the cell's byte offsets in W and y,
    // 1 GiB where the row has no y plane) and blended per store: a store
    // wave that issued as many instructions as a computing wave would hold
    // every barrier.
    if (C != 0) {
      // Ey, Ez: rows along x, a row's cells on consecutive diagonals are
      // consecutive in the grid.  Lane l takes step l % 8 of row 8 k + l / 8
      // of a wave, so eight lanes write one row's run of eight consecutive
      // cells and a store covers eight rows' runs (about eight cache lines,
      // against 64 for a step of 64 rows).  Half h moves f by SG2_BP.
      static_assert(SG2_BP == 8, "eight lanes per row, eight rows per store");
      const int u = lane % SG2_BP, rl = lane / SG2_BP;
      const long qb = (C == 1) ? (long)(n1 + 1) * q : PL * q;
      int fm1[NW][8], wo[NW][8], yo[NW][8];
#pragma unroll
      for (int w = 0; w < NW; w++) {
        const int lw = (FWD ? w : NW - 1 - w) * SG2_L;
        const int d = FWD ? 2 + u - lw : A + B - u + lw;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int sw = 1 + 64 * w + 8 * k + rl, scw = min(sw, A), f = d - sw;
          const long cell = qb + scw * SS + f;
          const int kopw = ((C == 2) ? q : scw) - a.kofs;
          fm1[w][k] = sw <= A ? f - 1 : (1 << 30);
          wo[w][k] = (int)cell * (int)sizeof(z2);
          yo[w][k] = (FWD || kopw < 0) ? (1 << 30) : (int)(cell - (long)a.kofs * PL) * (int)sizeof(z2);
        }
      }
      constexpr int FST = FWD ? SG2_BP : -SG2_BP, OST = FST * (int)sizeof(z2);
      for (int h = 0; h < nhalf; h++) {
        __syncthreads();
        const z2 *src = stg + (h & 1) * NW * SG2_BP * SG2_SP + u * SG2_SP + rl;
#pragma unroll
        for (int w = 0; w < NW; w++) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const int m = -(int)((unsigned)(fm1[w][k] + h * FST) < (unsigned)B);
            const z2 v = src[w * SG2_BP * SG2_SP + 8 * k];
            sg2_st(v, Wb, ((wo[w][k] + h * OST) & m) | (SG2_OOB & ~m));
            if (!FWD) sg2_st(v, Yb, ((yo[w][k] + h * OST) & m) | (SG2_OOB & ~m));
          }
        }
      }
      if (!FWD) sg2_zero_frame<C>(a, q, t, nt);
      return;
    }
    // Ex: rows a plane apart, columns a row apart: one step of 64 rows per
    // store; step tau moves f by one, the offsets by FS cells
    int fm1[NW], wo[NW], yo[NW];
#pragma unroll
    for (int w = 0; w < NW; w++) {
      const int sw = 1 + 64 * w + lane, scw = min(sw, A);
      const int lw = (FWD ? w : NW - 1 - w) * SG2_L;
      const long nbw = (long)q + scw * SS;
      const int kopw = scw - a.kofs, f = (FWD ? 2 - lw : A + B + lw) - sw;
      fm1[w] = sw <= A ? f - 1 : (1 << 30);
      wo[w] = (int)(nbw + (long)f * FS) * (int)sizeof(z2);
      yo[w] = kopw < 0 ? (1 << 30) : (int)(nbw - (long)a.kofs * PL + (long)f * FS) * (int)sizeof(z2);
    }
    const int TST = (FWD ? 1 : -1) * FS * (int)sizeof(z2);  // bytes per step
    for (int h = 0; h < nhalf; h++) {
      __syncthreads();
      const z2 *src = stg + (h & 1) * NW * SG2_BP * SG2_SP + lane;
#pragma unroll
      for (int w = 0; w < NW; w++) {
#pragma unroll
        for (int u = 0; u < SG2_BP; u++) {
          const int tau = h * SG2_BP + u;
          const int m = -(int)((unsigned)(fm1[w] + (FWD ? tau : -tau)) < (unsigned)B);
          const z2 v = src[(w * SG2_BP + u) * SG2_SP];
          if (FWD) sg2_st(v, Wb, ((wo[w] + tau * TST) & m) | (SG2_OOB & ~m));
          else sg2_st(v, Yb, ((yo[w] + tau * TST) & m) | (SG2_OOB & ~m));
        }
      }
    }
    if (!FWD) sg2_zero_frame<C>(a, q, t, nt);
    return;
  }
  const bool send = FWD ? (lane == 63 && wv < NW - 1) : (lane == 0 && wv > 0);
  const int zl = __builtin_amdgcn_mbcnt_lo(0u, 0u);  // a lane-varying zero: the ring address stays per lane
  z2 *ring_in = ring + wv * SG2_RING + zl;
  z2 *ring_st = send ? ring + (FWD ? wv + 1 : wv - 1) * SG2_RING : junk + wv * (SG2_RING + 64) + lane;
  z2 *stg_w = stg + wv * SG2_BP * SG2_SP + lane;  // this lane's staging slot of step u of half 0
  constexpr int SG2_P = LD::P;
  // step tau is diagonal d = 2 + tau - lag (forward) or A + B - tau + lag
  // (backward), this lane's cell (s, f = d - s) and its slot (d - 2) AP +
  // s - 1: the slot's byte offset in every plane is rv0 + tau RSTEP,
  // the fast table entry ib + tau (forward) or ib - tau, and the cell is on
  // the slice when f - 1 = fm0 + tau (forward) or fm0 - tau is in [0, B)
  const int d0 = FWD ? 2 - lag : A + B + lag, sl0 = (d0 - 2) * AP + s - 1;
  constexpr int RSTEP = (FWD ? 1 : -1) * AP * 16;
  const int rv0 = sl0 * 16;
  const int ib = d0 - s + OFF;
  const int fm0 = s <= A ? d0 - s - 1 : (1 << 30);
  struct Rg { z2 r0, r1, r2, r3, b; double4 S; };
  auto load = [&](Rg &g, int tau) {
    const bool ok = (unsigned)(FWD ? fm0 + tau : fm0 - tau) < (unsigned)B;
    const int ro = ok ? rv0 + tau * RSTEP : SG2_OOB;  // every plane's lane offset
    if (LD::R01) {
      g.r0 = sg2_ld(Rb, ro);
      g.r1 = sg2_ld(Rb, ro, PS);
    }
    if (LD::R23) {
      g.r2 = sg2_ld(Rb, ro, 2 * PS);
      g.r3 = sg2_ld(Rb, ro, 3 * PS);
    }
    g.b = sg2_ld(Rb, ro, LD::BP * PS);
    const z2 s01 = sg2_ld(Sb, ro), s23 = sg2_ld(Sb, ro, PS);
    g.S = make_double4(s01.x, s01.y, s23.x, s23.y);
  };
  __syncthreads();  // the tables and rings are ready
  Rg rg[SG2_P];
#pragma unroll
  for (int u = 0; u < SG2_P; u++) {
    if (!LD::R01) rg[u].r0 = rg[u].r1 = z0;
    if (!LD::R23) rg[u].r2 = rg[u].r3 = z0;
    load(rg[u], u);
  }
  double4 ftr[SG2_PC];
  double vtr[SG2_PC];
  z2 edg[SG2_PD];
#pragma unroll
  for (int u = 0; u < SG2_PC; u++) {
    ftr[u] = ftab[FWD ? ib + u : ib - u];
    vtr[u] = FVM ? vtab[FWD ? ib + u : ib - u] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < SG2_PD; u++) edg[u] = ring_in[(u - SG2_L - 1) & (SG2_RING - 1)];
  z2 prev = z0;
  const bool afast = (C != 1);
  for (int tk0 = 0; tk0 < ntk; tk0 += SG2_U) {
    // the fast table from step tk0 + SG2_PC on at literal offsets
    const double4 *ct = ftab + (FWD ? ib + tk0 : ib - tk0 - SG2_U - SG2_PC);
    const double *cv = vtab + (FWD ? ib + tk0 : ib - tk0 - SG2_U - SG2_PC);
#pragma unroll
    for (int uu = 0; uu < SG2_U; uu++) {
      const int tau = tk0 + uu;
      const Rg &c = rg[uu % SG2_P];
      const double4 ft = ftr[uu % SG2_PC];
      // new value of (s-1, f) (forward) or (s+1, f) (backward)
      const z2 nbv = FWD ? sg2_zdpp_keep<DPP_WAVE_SHR1>(edg[uu % SG2_PD], prev)
                         : sg2_zdpp_keep<DPP_WAVE_SHL1>(edg[uu % SG2_PD], prev);
      const double fsx = ft.x, rfsx = ft.y, fdf = ft.z, rfdf = ft.w;
      z2 t1, t2, t3, t4;
      if (FWD) {
        const z2 qa = afast ? zdivq(zsub(z0, prev), fsx, rfsx) : zdivq(zsub(z0, nbv), ssx, rssx);
        const z2 pb = afast ? zdivq(zsub(z0, nbv), ssx, rssx) : zdivq(zsub(z0, prev), fsx, rfsx);
        t1 = c.r0; t2 = zsub(c.r1, qa); t3 = c.r2; t4 = zsub(pb, c.r3);
      } else {
        const z2 qa = afast ? zdivq(zsub(prev, z0), fsx, rfsx) : zdivq(zsub(nbv, z0), ssx, rssx);
        const z2 pb = afast ? zdivq(zsub(nbv, z0), ssx, rssx) : zdivq(zsub(prev, z0), fsx, rfsx);
        t1 = zsub(c.r0, qa); t2 = c.r1; t3 = zsub(pb, c.r2); t4 = c.r3;
      }
      const z2 tA = afast ? zdivq(zsub(t1, t2), fdf, rfdf) : zdivq(zsub(t1, t2), sd, rsd);
      const z2 tB = afast ? zdivq(zsub(t3, t4), sd, rsd) : zdivq(zsub(t3, t4), fdf, rfdf);
      z2 aij = zmul(zsub(tA, tB), a.iwmu);
      if (FVM) aij = zmul(aij, zrmul((vtr[uu % SG2_PC] * vA) * vB, a.iomu));
      const z2 v = smith_finish(zsub(c.b, aij), c.S);
      stg_w[((uu / SG2_BP) & 1) * NW * SG2_BP * SG2_SP + (uu % SG2_BP) * SG2_SP] = v;
      prev = v;
      ring_st[uu & (SG2_RING - 1)] = prev;
      if (uu % SG2_BP == SG2_BP - 1) {
        __syncthreads();
        if (tau + 1 >= nex) goto swept;  // the same step for every wave
      }
      edg[uu % SG2_PD] = ring_in[(uu + SG2_PD - SG2_L - 1) & (SG2_RING - 1)];
      ftr[uu % SG2_PC] = ct[FWD ? uu + SG2_PC : SG2_U - uu];
      if (FVM) vtr[uu % SG2_PC] = cv[FWD ? uu + SG2_PC : SG2_U - uu];
      load(rg[uu % SG2_P], tau + SG2_P);
    }
  }
swept:
  if (!FWD) sg2_zero_frame<C>(a, q, t, nt);
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
#ifndef CSEM_TYPES_ONLY
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
#endif

// Column pass: forward transform of columns c0..c0+CW-1 of buf, then for
// each multiplier (sH1 -> bufx, sH2 -> bufy) the product and the inverse
// column transform (airbc_init_free.c:153-161, 182-188).
template <int CW>
__global__ void k_air_cols(AirArgs a)
{
  extern __shared__ z2 sm[];
  // column groups: consecutive groups on one XCD, so the neighbouring
  // strips of a row's cache lines are read and written through its L2
  const int ng = a.N / CW, per = (ng + 7) / 8, g = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (g >= ng) return;
  const int N = a.N, c0 = g * CW;
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
#ifndef CSEM_TYPES_ONLY
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
#endif

// Normalise by N*N and interpolate back to the ghost Hx at (x1[i1], x2s[i2])
// and ghost Hy at (x1s[i1], x2[i2]) (airbc_init_free.c:162-207).
#ifndef CSEM_TYPES_ONLY
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
#endif

// Generic 2-D transform of an N x N buffer (tests and the public fft2).
#ifndef CSEM_TYPES_ONLY
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
#endif

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
