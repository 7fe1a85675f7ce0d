// Device building blocks for the VPFP kernels (gfx950, wave64).
//
// Everything here is written against a small block context `Ctx` that
// provides `tid`, `nthr` (thread index and count in the workgroup), `bid`,
// `nblk` (workgroup index and count along x), `bidy` (workgroup index along
// y: the instance of a batched scan), `sync()` (workgroup barrier),
// `shfl_xor(v, mask, width)` (lane exchange inside a wavefront) and, for the
// cross-workgroup hand-off of the field kernel, `fence()` (device-scope
// fence), `arrive(counter)` (device-scope atomic increment, returns the old
// value), `load_coherent(p)` and `reset(counter)`.  The HIP kernels in
// vpfp.hip instantiate it with the hardware primitives.
//
// Contents
//   * cx<T>        : complex value in registers / LDS (T = double or float)
//   * dft<R>       : R-point DFT in registers (radix-2 network, R <= 32)
//   * fft<M,TPT>   : M-point complex FFT of one sequence held by TPT threads,
//                    Q = M/TPT values per thread in registers, exchanged
//                    through a padded LDS buffer between Stockham passes
//   * couple()     : half-length real-FFT unpack, Fourier multiplier, repack
//   * row_sum<>    : sum over the TPT threads that own one sequence
//   * tridiag      : partition solver for one diagonally dominant
//                    tridiagonal system held by TPT threads (chunk sweeps,
//                    elimination of the even interface unknowns, parallel
//                    cyclic reduction of the TPT odd ones in LDS, Thomas on
//                    the chunk interiors)
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define VP_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define VP_HD inline
#endif

namespace vp {

// ---------------------------------------------------------------------------
// complex arithmetic
// ---------------------------------------------------------------------------
template <class T>
struct cx {
  T re, im;
};

template <class T>
VP_HD cx<T> mk(T r, T i) {
  cx<T> z;
  z.re = r;
  z.im = i;
  return z;
}
template <class T>
VP_HD cx<T> add(cx<T> a, cx<T> b) { return mk<T>(a.re + b.re, a.im + b.im); }
template <class T>
VP_HD cx<T> sub(cx<T> a, cx<T> b) { return mk<T>(a.re - b.re, a.im - b.im); }
template <class T>
VP_HD cx<T> mul(cx<T> a, cx<T> b) {
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a * conj(b)
template <class T>
VP_HD cx<T> mulc(cx<T> a, cx<T> b) {
  return mk<T>(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}

// Padded LDS index: one spare slot every 16 elements keeps the strided
// writes of the first Stockham pass and the chunk reads of the tridiagonal
// solver spread over the LDS banks.
VP_HD constexpr int pidx(int i) { return i + (i >> 4); }
template <int M>
struct Padded {
  static constexpr int size = M + ((M - 1) >> 4) + 1;
};

// ---------------------------------------------------------------------------
// 1/x to double precision: hardware reciprocal estimate refined by two
// Newton steps (within an ulp of the correctly rounded quotient; cheaper
// than the IEEE division sequence in the tridiagonal sweeps)
// ---------------------------------------------------------------------------
VP_HD double rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
#else
  return 1.0 / x;
#endif
}
// The same with one Newton step: the estimate's ~23 correct bits doubled to
// ~46 (relative error ~2^-46), for the FP64 elimination of FP32-stored f,
// whose own rounding (2^-24) is ten orders of magnitude larger.
VP_HD double rcp46(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(x);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
#else
  return 1.0 / x;
#endif
}
template <bool FULL>
VP_HD double rcpp(double x) {
  if constexpr (FULL)
    return rcp(x);
  else
    return rcp46(x);
}

// ---------------------------------------------------------------------------
// exp(-i theta) in the working precision
// ---------------------------------------------------------------------------
// sin and cos of |y| <= pi/2 in FP32 from their Taylor series, Horner in y^2
// with fma: sin to y^13, cos to y^14 (the first terms left out are below
// 1e-9 at pi/2); absolute error ~1 ulp of 1, and sin keeps its relative
// precision for small y (the y term is added last).  Cheaper than a general
// sincosf, which reduces the argument again and checks for large ones.
VP_HD void sincos_half(float y, float& s, float& c) {
  const float y2 = y * y;
  const float ps = fmaf(
      y2, fmaf(y2, fmaf(y2, fmaf(y2, fmaf(y2, 1.605904384e-10f, -2.505210839e-08f), 2.755731922e-06f), -1.984126984e-04f),
               8.333333333e-03f),
      -1.666666667e-01f);
  s = fmaf(y * y2, ps, y);
  const float pc = fmaf(
      y2,
      fmaf(y2, fmaf(y2, fmaf(y2, fmaf(y2, fmaf(y2, -1.147074560e-11f, 2.087675699e-09f), -2.755731922e-07f), 2.480158730e-05f),
                    -1.388888889e-03f),
           4.166666667e-02f),
      -0.5f);
  c = fmaf(y2, pc, 1.f);
}

VP_HD void cis_neg(double th, cx<float>& out) {
  // reduce in double so that large arguments keep float accuracy; then the
  // half angle (|r / 2| <= pi / 2): cos r = 1 - 2 s^2, sin r = 2 s c
  const double r = th - 6.283185307179586 * rint(th * 0.15915494309189535);
  float s, c;
  sincos_half((float)(0.5 * r), s, c);
  out = mk<float>(fmaf(-2.f * s, s, 1.f), -2.f * s * c);
}

// Natural log in FP64 for the f log f moment: x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f / (2 + f) (|s| <= 0.1716),
// log(1 + f) = 2 atanh(s) = f - s (f - R), R = sum_k 2 s^(2k) / (2k + 1) for
// k = 1..11 (the first term left out is below 2e-19 relative), and
// log x = e ln 2 + log(1 + f) with ln 2 split in two; ~2 ulp.  x = 0 gives
// -inf, x < 0 and NaN give NaN, +inf gives +inf, as log does.  About half the
// instructions of the library log, which also covers the correctly rounded
// cases this moment does not need.
VP_HD double log_f64(double x) {
  int e;
  double m = frexp(x, &e);
  const bool lo = m < 0.70710678118654752;
  m = lo ? 2.0 * m : m;
  e -= lo ? 1 : 0;
  const double f = m - 1.0;
  const double s = f * rcp(2.0 + f);
  const double z = s * s;
  double r = 0.086956521739130432;  // 2/23
  r = fma(z, r, 0.095238095238095233);
  r = fma(z, r, 0.10526315789473684);
  r = fma(z, r, 0.11764705882352941);
  r = fma(z, r, 0.13333333333333333);
  r = fma(z, r, 0.15384615384615385);
  r = fma(z, r, 0.18181818181818182);
  r = fma(z, r, 0.22222222222222221);
  r = fma(z, r, 0.2857142857142857);
  r = fma(z, r, 0.40000000000000002);
  r = fma(z, r, 0.66666666666666663);  // 2/3
  r *= z;
  const double l1p = fma(-s, f - r, f);
  const double de = (double)e;
  const double v = fma(de, 0.6931471805599453, fma(de, 2.3190468138462996e-17, l1p));
  return x > 0.0 ? (x < __builtin_huge_val() ? v : x) : (x == 0.0 ? -__builtin_huge_val() : __builtin_nan(""));
}

// theta - 2 pi n with n = rint(theta / 2 pi), two-term Cody-Waite with fma
// (2 pi = C1 + C2, C1 its double rounding): |r| <= pi to a few ulp for the
// arguments of the kicks and drifts
VP_HD double reduce_2pi(double th) {
  const double n = rint(th * 0.15915494309189535);
  return fma(-n, 2.4492935982947064e-16, fma(-n, 6.283185307179586, th));
}

// sin and cos of |y| <= pi/2 in FP64 from their Taylor series, Horner in y^2
// with fma: sin to y^23, cos to y^24 (the first terms left out are below
// 1e-20 at pi/2); ~1 ulp, and sin keeps its relative precision for small y
VP_HD void sincos_half(double y, double& s, double& c) {
  const double y2 = y * y;
  double p = -3.86817017063068413e-23;
  p = fma(y2, p, 1.95729410633912626e-20);
  p = fma(y2, p, -8.22063524662432950e-18);
  p = fma(y2, p, 2.81145725434552060e-15);
  p = fma(y2, p, -7.64716373181981641e-13);
  p = fma(y2, p, 1.60590438368216133e-10);
  p = fma(y2, p, -2.50521083854417202e-08);
  p = fma(y2, p, 2.75573192239858925e-06);
  p = fma(y2, p, -1.98412698412698413e-04);
  p = fma(y2, p, 8.33333333333333322e-03);
  p = fma(y2, p, -1.66666666666666657e-01);
  s = fma(y * y2, p, y);
  double q = 1.61173757109611839e-24;
  q = fma(y2, q, -8.89679139245057408e-22);
  q = fma(y2, q, 4.11031762331216484e-19);
  q = fma(y2, q, -1.56192069685862253e-16);
  q = fma(y2, q, 4.77947733238738525e-14);
  q = fma(y2, q, -1.14707455977297245e-11);
  q = fma(y2, q, 2.08767569878681002e-09);
  q = fma(y2, q, -2.75573192239858883e-07);
  q = fma(y2, q, 2.48015873015873016e-05);
  q = fma(y2, q, -1.38888888888888894e-03);
  q = fma(y2, q, 4.16666666666666644e-02);
  q = fma(y2, q, -0.5);
  c = fma(y2, q, 1.0);
}

VP_HD void cis_neg(double th, cx<double>& out) {
  double s, c;
  sincos_half(0.5 * reduce_2pi(th), s, c);
  out = mk<double>(fma(-2.0 * s, s, 1.0), -2.0 * s * c);
}

// ---------------------------------------------------------------------------
// R-point DFT in registers (radix-2 decimation in time, R <= 32)
// ---------------------------------------------------------------------------
// cos(2 pi j / 32) for j = 0..8, correctly rounded
VP_HD constexpr double cos32q(int j) {
  return j == 0   ? 1.0
         : j == 1 ? 0.9807852804032304
         : j == 2 ? 0.9238795325112867
         : j == 3 ? 0.8314696123025452
         : j == 4 ? 0.7071067811865476
         : j == 5 ? 0.5555702330196022
         : j == 6 ? 0.3826834323650898
         : j == 7 ? 0.19509032201612828
                  : 0.0;
}
VP_HD constexpr double cos32(int J) {
  return (J & 31) <= 8    ? cos32q(J & 31)
         : (J & 31) <= 16 ? -cos32q(16 - (J & 31))
         : (J & 31) <= 24 ? -cos32q((J & 31) - 16)
                          : cos32q(32 - (J & 31));
}
VP_HD constexpr double sin32(int J) { return cos32(J - 8); }

// b * W_32^J  (forward: W = exp(-2 pi i J / 32); inverse: conjugate)
template <bool INV, class T>
VP_HD cx<T> tw32(cx<T> b, int J) {
  J &= 31;
  if (J == 0) return b;
  if (J == 16) return mk<T>(-b.re, -b.im);
  if (J == 8) return INV ? mk<T>(-b.im, b.re) : mk<T>(b.im, -b.re);
  if (J == 24) return INV ? mk<T>(b.im, -b.re) : mk<T>(-b.im, b.re);
  const T c = (T)cos32(J), s = (T)sin32(J);
  return INV ? mk<T>(b.re * c - b.im * s, b.im * c + b.re * s)
             : mk<T>(b.re * c + b.im * s, b.im * c - b.re * s);
}

template <int R>
VP_HD constexpr int brev(int i) {
  int r = 0;
  for (int b = 1; b < R; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

template <int R, bool INV, class T>
VP_HD void dft(cx<T>* x) {
  static_assert(R >= 1 && R <= 32 && (R & (R - 1)) == 0, "radix must be a power of two <= 32");
  if constexpr (R == 1) return;
  cx<T> y[R];
#pragma unroll
  for (int i = 0; i < R; ++i) y[brev<R>(i)] = x[i];
#pragma unroll
  for (int h = 1; h < R; h <<= 1) {
#pragma unroll
    for (int g = 0; g < R; g += 2 * h) {
#pragma unroll
      for (int j = 0; j < h; ++j) {
        const cx<T> a = y[g + j];
        const cx<T> b = tw32<INV>(y[g + j + h], j * (16 / h));
        y[g + j] = add(a, b);
        y[g + j + h] = sub(a, b);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) x[i] = y[i];
}

// ---------------------------------------------------------------------------
// Twiddles W_N^i = exp(-2 pi i i / N) staged in LDS.  Large N: two small
// tables, W_N^i = hi[i >> 6] * lo[i & 63], hi[h] = W_N^{64 h}, lo[j] = W_N^j,
// global layout [lo: 64 entries | hi: max(1, N/64) entries].
// ---------------------------------------------------------------------------
// For N <= kTwFull the whole table W_N^i, i < N, is staged instead: one LDS
// read per twiddle and no complex product.  Larger full tables cost more than
// they save: every workgroup stages its table from L2, and a full FP32 table
// for N = 2048 (16 KiB per one-row workgroup) doubled the row kernel's L2
// reads and measured 6 % slower at nv = 2048 than the two-level one.
constexpr int kTwFull = 1024;
VP_HD constexpr bool tw_full(int N, int bytes) { return bytes > 0 && N <= kTwFull; }
VP_HD constexpr int tw_entries(int N, int bytes) {
  return tw_full(N, bytes) ? N : 64 + (N / 64 > 1 ? N / 64 : 1);
}

// B: bytes of one complex entry (sizeof(cx<T>))
template <int N, int B>
struct TwSize {
  static constexpr bool full = tw_full(N, B);
  static constexpr int total = tw_entries(N, B);
};

template <class T, bool FULL = false>
struct Tw {
  const cx<T>* t;  // LDS copy: W_N^i (FULL) or [lo | hi]
  VP_HD cx<T> operator()(int i) const {
    if constexpr (FULL) {
      return t[i];
    } else {
      const cx<T> lo = t[i & 63];
      if (i < 64) return lo;
      return mul(t[64 + (i >> 6)], lo);
    }
  }
};

// all threads copy the table to LDS (no barrier here)
template <class T, int N, class Ctx>
VP_HD void tw_stage(Ctx& c, cx<T>* dst, const cx<T>* src) {
  for (int i = c.tid; i < TwSize<N, (int)sizeof(cx<T>)>::total; i += c.nthr) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// M-point complex FFT held by TPT threads (Stockham autosort, radix <= 16)
//
// Register layout: thread t holds element t + i*TPT in v[i], i < Q = M/TPT.
// A pass of radix R with NS = product of the earlier radices takes the
// butterfly j = t + q*TPT (inputs j + r*M/R = v[q + r*Q/R]), twiddles input r
// by W_{NS*R}^{r*(j mod NS)}, does an R-point DFT and writes output r to
// position (j - j mod NS)*R + j mod NS + r*NS.  After the last pass the data
// are in natural order, both in v and in `buf`.
//
// tw(n) = exp(-2 pi i n / (2M)); the M-point twiddle W_M^n is tw(2n).
// ---------------------------------------------------------------------------
// `act` false: the thread only takes part in the barriers (no LDS writes).
template <class T, int M, int TPT, int R, int NS, bool INV, class TW, class Ctx>
VP_HD void st_pass(Ctx& c, cx<T>* v, cx<T>* buf, const TW& tw, int t, bool act) {
  constexpr int Q = M / TPT;
  constexpr int NB = Q / R;
  static_assert(NB * R == Q, "radix must divide the per-thread count");
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int j = t + q * TPT;
    const int k = j & (NS - 1);
    cx<T> u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[q + r * NB];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cx<T> w = tw(2 * (r * k * (M / (NS * R))));
        u[r] = INV ? mulc(u[r], w) : mul(u[r], w);
      }
    }
    dft<R, INV>(u);
    const int base = (j - k) * R + k;
    if (act) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[pidx(base + r * NS)] = u[r];
    }
  }
  c.sync();
#pragma unroll
  for (int i = 0; i < Q; ++i) v[i] = buf[pidx(t + i * TPT)];
  c.sync();
}

template <class T, int M, int TPT, bool INV, int NS = 1, class TW, class Ctx>
VP_HD void fft(Ctx& c, cx<T>* v, cx<T>* buf, const TW& tw, int t, bool act = true) {
  constexpr int Q = M / TPT;
  static_assert(Q >= 2 && Q * TPT == M, "each thread must hold at least two values");
  constexpr int RMAX = Q < 16 ? Q : 16;
  constexpr int R = (M / NS) < RMAX ? (M / NS) : RMAX;
  st_pass<T, M, TPT, R, NS, INV>(c, v, buf, tw, t, act);
  if constexpr (NS * R < M) fft<T, M, TPT, INV, NS * R>(c, v, buf, tw, t, act);
}

// ---------------------------------------------------------------------------
// Fourier multipliers for the half-length real transform
//
// For own bin k in [0, M) `pair` returns P_k and P_{M-k}; for k == 0 it
// returns the real factors applied to the DC and Nyquist bins (the real part
// of the multiplier, which is what taking Re() of the complex inverse leaves).
// ---------------------------------------------------------------------------

// Phase shift exp(-i theta_k) of the exponential pushers, theta_k = k phi
// with phi = (dk * dt) * s: dk is the first wavenumber of an fftfreq table
// (checked on the host to be linear in k to rounding), s the advection speed
// (e(x) for the v-kick, v for the x-drift).  Bins are visited in the order
// k = t, t + TPT, t + 2 TPT, ... so P_k is a running product; the partner
// phase is P_{M-k} = exp(-i theta_M) conj(P_k).
template <class T>
struct PhaseMul {
  static constexpr bool kBg = false;
  cx<T> cur, step, pm;
  VP_HD PhaseMul(double phi, int M, int t, int TPT) {
    cis_neg((double)M * phi, pm);
    cis_neg((double)t * phi, cur);
    cis_neg((double)TPT * phi, step);
  }
  VP_HD void pair(int k, cx<T>& pk, cx<T>& pmk) {
    if (k == 0) {
      pk = mk<T>(T(1), T(0));
      pmk = mk<T>(pm.re, T(0));
    } else {
      pk = cur;
      pmk = mulc(pm, pk);
    }
    cur = mul(cur, step);
  }
};

// Poisson multiplier i * one_over_kx (zero real part at DC and Nyquist);
// okx is an LDS copy of the one_over_kx table.
// exp(-i theta) - 1 without cancellation: (-2 sin^2(theta/2), -2 sin(theta/2)
// cos(theta/2)), from one sincos of the half angle after reduction in double
// (FP32 sincos for FP32 results: both parts keep FP32 relative precision
// also for tiny theta)
template <class T>
VP_HD cx<T> cis_neg_m1(double th) {
  const double r = th - 6.283185307179586 * rint(th * 0.15915494309189535);
  if constexpr (sizeof(T) == 4) {
    float s, c;
    sincos_half((float)(0.5 * r), s, c);
    return mk<T>(-2.0f * s * s, -2.0f * s * c);
  } else {
    double s, c;
    sincos_half(0.5 * reduce_2pi(th), s, c);
    return mk<T>(-2.0 * s * s, -2.0 * s * c);
  }
}

// PhaseMul on a stored deviation f - B from a fixed background B whose
// unnormalised spectrum is bhat [N] (storage precision): the shift of B itself
// enters as F'_k += bhat_k (P_k - 1), F'_{k+M} += bhat_{k+M} (conj P_{M-k} - 1).
// c_k = P_k - 1 is carried by c_{k+s} = c_k + c_s + c_k c_s (no cancellation
// for small shifts) and the phases themselves are P_k = 1 + c_k,
// P_M = conj(1 + cpm) with cpm = exp(+i M phi) - 1.  couple() calls init()
// once, then pair(k, ...) and bg_add(k, ...) for each k in turn; all of the
// state is formed in init(), after the forward transform.
template <class T>
struct BgPhaseMul {
  static constexpr bool kBg = true;
  double phi;
  int M, t, TPT;
  const cx<T>* bhat;
  cx<T> cm, cs, cpm;  // P_k - 1, P_TPT - 1, exp(+i M phi) - 1
  VP_HD BgPhaseMul(double phi_, int M_, int t_, int TPT_, const cx<T>* bhat_)
      : phi(phi_), M(M_), t(t_), TPT(TPT_), bhat(bhat_) {}
  VP_HD void init() {
    cm = cis_neg_m1<T>((double)t * phi);
    cs = cis_neg_m1<T>((double)TPT * phi);
    cpm = cis_neg_m1<T>(-(double)M * phi);
  }
  VP_HD void pair(int k, cx<T>& pk, cx<T>& pmk) {
    const cx<T> pm = mk<T>(T(1) + cpm.re, -cpm.im);
    if (k == 0) {
      pk = mk<T>(T(1), T(0));
      pmk = mk<T>(pm.re, T(0));
    } else {
      pk = mk<T>(T(1) + cm.re, cm.im);
      pmk = mulc(pm, pk);
    }
  }
  VP_HD void bg_add(int k, cx<T>& G1, cx<T>& G2) {
    const cx<T> B1 = bhat[k], B2 = bhat[k + M];
    if (k == 0) {
      // DC unchanged; Nyquist multiplier Re(P_M) = cos(M phi)
      G2 = add(G2, mk<T>(B2.re * cpm.re, B2.im * cpm.re));
    } else {
      G1 = add(G1, mul(B1, cm));
      const cx<T> q = add(add(cpm, cm), mul(cpm, cm));
      G2 = add(G2, mul(B2, q));
    }
    cm = add(add(cm, cs), mul(cm, cs));
  }
};

template <class T>
struct PoissonMul {
  static constexpr bool kBg = false;
  const double* okx;
  int M;
  VP_HD void pair(int k, cx<T>& pk, cx<T>& pmk) {
    if (k == 0) {
      pk = mk<T>(T(0), T(0));
      pmk = mk<T>(T(0), T(0));
      return;
    }
    pk = mk<T>(T(0), (T)okx[k]);
    pmk = mk<T>(T(0), (T)okx[M - k]);
  }
};

// z_j = f_{2j} + i f_{2j+1} was transformed to Z (M points).  With
// E_k = (Z_k + conj Z_{M-k})/2, O_k = (Z_k - conj Z_{M-k})/(2i) and
// W = exp(-2 pi i k / N), the real N-point spectrum is F_k = E_k + W O_k and
// F_{k+M} = E_k - W O_k.  After F'_k = P_k F_k, F'_{k+M} = conj(P_{M-k}) F_{k+M}
// the repacked spectrum Z'_k = (F'_k + F'_{k+M} + i (F'_k - F'_{k+M}) conj W)/N
// inverse-transforms (unnormalised M-point) to f'_{2j} + i f'_{2j+1}.
// v[i] holds Z_k and pz[i] holds Z_{M-k} for k = t + i*TPT; tw(k) = W^k.
template <class T, int M, int TPT, class Mul, class TW>
VP_HD void couple(cx<T>* v, const cx<T>* pz, const TW& tw, int t, Mul& mult) {
  constexpr int Q = M / TPT;
  const T invN = T(1) / T(2 * M);
  const T h = T(0.5);
  if constexpr (Mul::kBg) mult.init();
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int k = t + i * TPT;
    const cx<T> a = v[i], p = pz[i];
    const cx<T> E = mk<T>((a.re + p.re) * h, (a.im - p.im) * h);
    const cx<T> O = mk<T>((a.im + p.im) * h, (p.re - a.re) * h);
    const cx<T> W = tw(k);
    const cx<T> WO = mul(W, O);
    const cx<T> Fk = add(E, WO), Fkm = sub(E, WO);
    cx<T> pk, pmk;
    mult.pair(k, pk, pmk);
    cx<T> G1 = mul(pk, Fk);
    cx<T> G2 = mulc(Fkm, pmk);
    if constexpr (Mul::kBg) mult.bg_add(k, G1, G2);
    const cx<T> S = add(G1, G2);
    const cx<T> D = mulc(sub(G1, G2), W);
    v[i] = mk<T>((S.re - D.im) * invN, (S.im + D.re) * invN);
  }
}

// ---------------------------------------------------------------------------
// sums over the TPT threads that own one sequence (all of them get the sum)
// scratch: (TPT/64)*NV doubles per sequence when TPT > 64 (one sequence per
// workgroup in that case)
// ---------------------------------------------------------------------------
template <int TPT, int NV, class Ctx>
VP_HD void row_sum(Ctx& c, double* vals, double* scratch) {
  constexpr int W = TPT < 64 ? TPT : 64;
#pragma unroll
  for (int m = 1; m < W; m <<= 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) vals[q] += c.shfl_xor(vals[q], m, W);
  }
  if constexpr (TPT > 64) {
    constexpr int NW = TPT / 64;
    const int w = c.tid >> 6;
    if ((c.tid & 63) == 0) {
#pragma unroll
      for (int q = 0; q < NV; ++q) scratch[w * NV + q] = vals[q];
    }
    c.sync();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) s += scratch[ww * NV + q];
      vals[q] = s;
    }
    c.sync();
  }
}

// ---------------------------------------------------------------------------
// Partition solver for one tridiagonal system of N = TPT*L unknowns.
//
// Row m reads  l(m) x_{m-1} + b(m) x_m + u(m) x_{m+1} = d_m  with l(0) = 0 and
// u(N-1) = 0.  Thread t owns rows m0 = t*L .. m0+L-1 with d in `d[L]`; on
// return `d` holds the solution.  Coef::view<L>(m0, first, last) gives the
// thread's coefficients l(i), u(i), b(i) of row m0 + i (zero l(0) on the
// first thread and u(L-1) on the last).
// red: 5*TPT doubles of LDS for the interface system (one row per thread
// after the even interface unknowns are eliminated).
// ---------------------------------------------------------------------------
// FULL: reciprocals to full double precision (FP64 f); otherwise to ~2^-46
// (FP32 f).
template <int TPT, int L, class Coef, class Ctx, bool FULL = true>
VP_HD void tridiag(Ctx& c, double* d, const Coef& cf, int t, double* red) {
  static_assert(L >= 2, "chunk must hold at least two rows");
  const int m0 = t * L;
  // this thread's rows m0 .. m0+L-1: coefficients by chunk index i (the
  // first thread's l(0) and the last thread's u(L-1) are zero)
  const auto cv = cf.template view<L>(m0, t == 0, t == TPT - 1);
  // Two independent sweeps, interleaved so that their dependency chains
  // overlap:
  //   down (i = 1 .. L-1), x_{m0} kept free:   x_i = p  + q  x_{m0}     - r  x_{i+1}
  //   up   (i = L-2 .. 0), x_{m0+L-1} kept free: x_i = p2 + q2 x_{m0+L-1} - r2 x_{i-1}
  double p = 0, q = 0, r = 0, p2 = 0, q2 = 0, r2 = 0;
#pragma unroll
  for (int s = 0; s < L - 1; ++s) {
    {
      const int i = 1 + s;
      const double lm = cv.l(i), um = cv.u(i), bm = cv.b(i);
      const double den = (i == 1) ? bm : bm - lm * r;
      const double inv = rcpp<FULL>(den);
      p = (i == 1) ? d[i] * inv : (d[i] - lm * p) * inv;
      q = (i == 1) ? -lm * inv : -lm * q * inv;
      r = um * inv;
    }
    {
      const int i = L - 2 - s;
      const double lm = cv.l(i), um = cv.u(i), bm = cv.b(i);
      const double den = (i == L - 2) ? bm : bm - um * r2;
      const double inv = rcpp<FULL>(den);
      p2 = (i == L - 2) ? d[i] * inv : (d[i] - um * p2) * inv;
      q2 = (i == L - 2) ? -um * inv : -um * q2 * inv;
      r2 = lm * inv;
    }
  }
  // interface unknowns y_{2t} = x_{m0}, y_{2t+1} = x_{m0+L-1}:
  //   row 2t  :  r2 y_{2t-1} + y_{2t} - q2 y_{2t+1} = p2
  //   row 2t+1: -q  y_{2t}   + y_{2t+1} + r  y_{2t+2} = p
  // Eliminating the even unknowns (y_{2t} by row 2t, y_{2t+2} by row 2t+2 of
  // the next thread) leaves one row per thread in z_t = y_{2t+1}:
  //   A z_{t-1} + B z_t + C z_{t+1} = D,
  //   A = q r2, B = 1 - q q2 - r r2', C = r q2', D = p + q p2 - r p2'
  // (primes: thread t+1's values; r = 0 on the last thread, r2 = 0 on the
  // first).  It is a Schur complement of the diagonally dominant system,
  // hence diagonally dominant itself, and is solved by parallel cyclic
  // reduction, one row per thread.
  double* A = red;
  double* B = red + TPT;
  double* C = red + 2 * TPT;
  double* D = red + 3 * TPT;
  A[t] = p2;
  B[t] = q2;
  C[t] = r2;
  c.sync();
  const bool last = t == TPT - 1;
  const double p2n = last ? 0.0 : A[t + 1], q2n = last ? 0.0 : B[t + 1], r2n = last ? 0.0 : C[t + 1];
  c.sync();
  double ra = q * r2, rb = 1.0 - q * q2 - r * r2n, rc = r * q2n, rd = p + q * p2 - r * p2n;
  // each thread publishes 1/B of its row with it (one reciprocal per level)
  double* BI = red + 4 * TPT;
  for (int s = 1; s < TPT; s <<= 1) {
    A[t] = ra;
    BI[t] = rcpp<FULL>(rb);
    C[t] = rc;
    D[t] = rd;
    c.sync();
    const int im = t - s, ip = t + s;
    double na = 0.0, nb = rb, nc = 0.0, nd = rd;
    if (im >= 0) {
      const double al = -ra * BI[im];
      na = al * A[im];
      nb += al * C[im];
      nd += al * D[im];
    }
    if (ip < TPT) {
      const double ga = -rc * BI[ip];
      nc = ga * C[ip];
      nb += ga * A[ip];
      nd += ga * D[ip];
    }
    ra = na;
    rb = nb;
    rc = nc;
    rd = nd;
    c.sync();
  }
  const double xe = rd * rcpp<FULL>(rb);
  // y_{2t} = p2 - r2 z_{t-1} + q2 z_t
  A[t] = xe;
  c.sync();
  const double zm = t > 0 ? A[t - 1] : 0.0;
  c.sync();
  const double xs = p2 - r2 * zm + q2 * xe;
  // interior rows m0+1 .. m0+L-2 with both chunk ends known (Thomas)
  double rr[L];
  d[0] = xs;
#pragma unroll
  for (int i = 1; i < L - 1; ++i) {
    const double lm = cv.l(i), um = cv.u(i), bm = cv.b(i);
    const double den = (i == 1) ? bm : bm - lm * rr[i - 1];
    const double inv = rcpp<FULL>(den);
    d[i] = (i == 1) ? (d[i] - lm * xs) * inv : (d[i] - lm * d[i - 1]) * inv;
    rr[i] = um * inv;
  }
  d[L - 1] = xe;
#pragma unroll
  for (int i = L - 2; i >= 1; --i) d[i] = d[i] - rr[i] * d[i + 1];
}

}  // namespace vp
