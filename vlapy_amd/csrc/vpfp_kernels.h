// Kernel bodies of the VPFP step.  Each body is written for one workgroup
// and a block context `Ctx` (tid, nthr, bid, sync, shfl_xor); vpfp.hip wraps
// them in __global__ kernels.
//
// Data layout (HBM): f is (nx, nv) row-major, v fastest, T = double or float.
// A row (fixed x) is contiguous; a column (fixed v) has stride nv.
//
//   row_body    one or more x-rows per workgroup.  Optional stages, in order:
//               kick A (v-advection by e(x) dt_a), collision (implicit LB/DG
//               solve along v), storage moments, kick B (v-advection by e(x)
//               dt_b).  One HBM read and one write of the row whatever the
//               stages.   [vlasov.py:113-140, collisions.py:26-160,222-265,
//               step.py:143-228]
//   drift_body  C v-columns per workgroup: x-advection of each column and the
//               per-column-tile partial trapz sums sum_v w_v f(x, v) feeding the
//               Poisson solve.   [vlasov.py:83-110, field.py:27-36]
//   field_body  one workgroup: density from the partial sums, pseudo-spectral
//               Poisson solve, driver added, E written (and <E^2>, <D^2>).
//               [field.py:39-88, step.py:215-216]
#pragma once
#include "vpfp_device.h"

namespace vp {

enum RowFlags { RK_KICK_A = 1, RK_COLLIDE = 2, RK_MOMENTS = 4, RK_KICK_B = 8 };
constexpr int kMoments = 8;  // n, j, T, q, fv4, vN, f^2, f log f

// trapz weight of sample m of an N-point row
VP_HD double trapw(int m, int N) { return (m == 0 || m == N - 1) ? 0.5 : 1.0; }

// ---------------------------------------------------------------------------
// spectral multiply along a length-2M real sequence held as M complex values
// ---------------------------------------------------------------------------
template <class T, int M, int TPT, class Mul, class TW, class Ctx>
VP_HD void spectral_apply(Ctx& c, cx<T>* v, cx<T>* buf, const TW& tw, int t, Mul mult, bool act = true) {
  constexpr int Q = M / TPT;
  fft<T, M, TPT, false>(c, v, buf, tw, t, act);
  cx<T> pz[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) pz[i] = buf[pidx((M - (t + i * TPT)) & (M - 1))];
  c.sync();
  couple<T, M, TPT>(v, pz, tw, t, mult);
  fft<T, M, TPT, true>(c, v, buf, tw, t, act);
}

// ---------------------------------------------------------------------------
// collision coefficients (collisions.py:67-79 LB, :138-156 DG)
//   l(m) = nu dt (-T/dv^2 + (v_{m-1} - vbar)/(2 dv)),  m >= 1
//   u(m) = nu dt (-T/dv^2 - (v_{m+1} - vbar)/(2 dv)),  m <= N-2
//   b    = 1 + nu dt (2 T / dv^2)
// ---------------------------------------------------------------------------
// With v_m = v0 + m dvv (the host checks that the grid is uniform to
// rounding) both off-diagonals are affine in m:
//   l(m) = nu dt (A0 + (v_{m-1} - vbar) H) = l0 + ls m
//   u(m) = nu dt (A0 - (v_{m+1} - vbar) H) = u0 - ls m
// with A0 = -T/dv^2, H = 1/(2 dv); one FMA per coefficient.
struct CollCoef {
  int N;
  double l0, u0, ls, bd;
  VP_HD void set(double ndt, double A0, double H, double v0, double dvv, double vbar, double bd_) {
    ls = ndt * (H * dvv);
    l0 = ndt * (A0 + ((v0 - dvv) - vbar) * H);
    u0 = ndt * (A0 - ((v0 + dvv) - vbar) * H);
    bd = bd_;
  }
  // rows m0 + i of one thread: l and u affine in the chunk index i (one FMA
  // each), the ends of the whole row zero
  template <int L>
  struct View {
    double lb, ub, ls, bd;
    bool first, last;
    VP_HD double l(int i) const { return (i == 0 && first) ? 0.0 : lb + ls * (double)i; }
    VP_HD double u(int i) const { return (i == L - 1 && last) ? 0.0 : ub - ls * (double)i; }
    VP_HD double b(int) const { return bd; }
  };
  template <int L>
  VP_HD View<L> view(int m0, bool first, bool last) const {
    return View<L>{l0 + ls * (double)m0, u0 - ls * (double)m0, ls, bd, first, last};
  }
};

// explicit diagonals a (nv-1), b (nv), c (nv-1) of one row
struct ArrCoef {
  const double* a;
  const double* bb;
  const double* cc;
  int N;
  template <int L>
  struct View {
    const double *a, *bb, *cc;
    int m0, N;
    bool first, last;
    // indices clamped into the arrays: the row ends are never read
    VP_HD double l(int i) const {
      const int k = m0 + i - 1;
      return (i == 0 && first) ? 0.0 : a[k > 0 ? k : 0];
    }
    VP_HD double u(int i) const {
      const int k = m0 + i;
      return (i == L - 1 && last) ? 0.0 : cc[k < N - 2 ? k : N - 2];
    }
    VP_HD double b(int i) const { return bb[m0 + i]; }
  };
  template <int L>
  VP_HD View<L> view(int m0, bool first, bool last) const {
    return View<L>{a, bb, cc, m0, N, first, last};
  }
};

// drift velocity and temperature of one row (un-normalised, trapz), chunk layout
// trapz weight of chunk element i of thread t (rows t L .. t L + L - 1 of a
// row of TPT L samples): 1/2 at the two ends of the row
template <int TPT, int L>
VP_HD double chunk_w(int i, int t) {
  return ((i == 0 && t == 0) || (i == L - 1 && t == TPT - 1)) ? 0.5 : 1.0;
}

template <int TPT, int L, class Ctx>
VP_HD void coll_moments(Ctx& c, const double* d, double v0, double dvv, int N, double dv, bool dg, int t,
                        double* rs, double& vbar, double& temp) {
  const int m0 = t * L;
  // v_{m0+i} = vb + i dvv
  const double vb = v0 + (double)m0 * dvv;
  double s[1] = {0.0};
  vbar = 0.0;
  if (dg) {
#pragma unroll
    for (int i = 0; i < L; ++i) s[0] += chunk_w<TPT, L>(i, t) * d[i] * (vb + (double)i * dvv);
    row_sum<TPT, 1>(c, s, rs);
    vbar = dv * s[0];
    s[0] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const double w = (vb + (double)i * dvv) - vbar;
    s[0] += chunk_w<TPT, L>(i, t) * d[i] * (w * w);
  }
  row_sum<TPT, 1>(c, s, rs);
  temp = dv * s[0];
}

// ---------------------------------------------------------------------------
// field kernel: n(x) = scale * sum_tiles part[tile][x];  E = Re ifft(i/kx fft(1 - n)) + driver
// ---------------------------------------------------------------------------
struct FieldArgs {
  const double* part;
  int ntiles;
  int nx;
  double scale;
  double* rho_g;          // [nx] global scratch when the grid has several workgroups
  unsigned int* counter;  // arrival counter (zero between launches)
  const double* okx;      // [nx]
  const cx<double>* tw;   // pow2 kernel: twiddle tables [lo | hi]; DFT kernel: exp(-2 pi i n / nx), n < nx
  const double* driver;   // [nx] or null
  double* e_out;          // [nx]
  double* e_out2;         // [nx] or null
  double* series;         // [2] or null: mean(E^2), mean(driver^2)
  int series_stride;      // batched scan: doubles between consecutive instances' series entries
  // single grid (vpfp_step_n_store): workgroups xm_blk0 and up of the launch
  // form the first xm_n x-modes of xm_f [nx][xm_nv] (FP32 if xm_f32, else
  // FP64) into xm_out [xm_n][xm_nv] (xmodes_body); xm_tw[r] = exp(-2 pi i r / nx)
  const void* xm_f;
  int xm_f32, xm_n, xm_nv, xm_blk0;
  const cx<double>* xm_tw;
  cx<double>* xm_out;
};

// The view of instance inst of a batched field solve: every array is
// instance-major (part [inst][ntiles][nx], okx / driver / E [inst][nx],
// series [inst * series_stride], one 64-byte counter slot per instance).
// Instance 0 is the single-grid case itself.
VP_HD FieldArgs field_instance(const FieldArgs& a, int inst) {
  FieldArgs r = a;
  if (inst == 0) return r;
  const size_t xo = (size_t)inst * a.nx;
  r.part += (size_t)inst * a.ntiles * a.nx;
  if (r.rho_g) r.rho_g += xo;
  if (r.counter) r.counter += inst * 16;
  r.okx += xo;
  if (r.driver) r.driver += xo;
  r.e_out += xo;
  if (r.e_out2) r.e_out2 += xo;
  if (r.series) r.series += (size_t)inst * a.series_stride;
  return r;
}

// rho[x] = 1 - scale * sum_tiles part[tile][x] for x in [x0, x0 + XS)
// (all FT threads; XS divides FT or FT divides XS)
template <int FT, bool COH = false, class Ctx>
VP_HD void field_density(Ctx& c, const FieldArgs& a, int x0, int XS, double* rho, double* tmp) {
  const int nx = a.nx;
  const int G = XS >= FT ? 1 : FT / XS;
  for (int l0 = c.tid; l0 < XS * G; l0 += FT) {
    const int xl = l0 % XS, g = l0 / XS;
    const double* p = a.part + x0 + xl;
    auto ld = [&](int tl) { return COH ? c.load_coherent(p + (size_t)tl * nx) : p[(size_t)tl * nx]; };
    // eight independent partial sums keep eight loads in flight per thread
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int tile = g;
    for (; tile + 7 * G < a.ntiles; tile += 8 * G) {
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += ld(tile + u * G);
    }
    for (int u = 0; tile < a.ntiles; tile += G, ++u) acc[u & 7] += ld(tile);
    tmp[l0] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  }
  c.sync();
  for (int xl = c.tid; xl < XS; xl += FT) {
    double s = 0.0;
    for (int g = 0; g < G; ++g) s += tmp[g * XS + xl];
    rho[xl] = 1.0 - a.scale * s;
  }
  c.sync();
}

// Field at a workgroup's own rows from the drift's partial density rows, by
// the direct DFT of field_dft_body (small single grids, nx <= kProNx):
//   E(x) = Re sum_m F_m exp(2 pi i m x / nx) / nx + driver(x),
//   F_m = i g_m sum_j rho_j exp(-2 pi i m j / nx), g_m = (okx_m - okx_{nx-m}) / 2.
// Every workgroup forms all nx values of E (each sum split over P = FT/nx
// adjacent lanes, joined by shuffles), writes those of rows [row0, row0 +
// nrows) to e_out (and e_out2); the workgroup holding row 0 writes the
// series.  All FT threads call it; erow (= lds + ProCfg::erow_off) holds E.
// lds: tw [kProNx] cx | okx [kProNx] | rho [kProNx] | F [kProNx] cx | tmp [tmp_n] | E [kProNx]
constexpr int kProNx = 128;
template <int FT>
struct ProCfg {
  static constexpr int tmp_n = 2 * kProNx > FT ? 2 * kProNx : FT;
  static constexpr int tw_off = 0;
  static constexpr int okx_off = tw_off + kProNx * 16;
  static constexpr int rho_off = okx_off + kProNx * 8;
  static constexpr int f_off = rho_off + kProNx * 8;
  static constexpr int tmp_off = f_off + kProNx * 16;
  static constexpr int erow_off = tmp_off + tmp_n * 8;
  static constexpr int bytes = erow_off + kProNx * 8;
};

template <int FT, class Ctx>
VP_HD void field_rows(Ctx& c, const FieldArgs& a, unsigned char* lds, int row0, int nrows) {
  using P_ = ProCfg<FT>;
  const int nx = a.nx;
  cx<double>* twl = reinterpret_cast<cx<double>*>(lds + P_::tw_off);
  double* okx = reinterpret_cast<double*>(lds + P_::okx_off);
  double* rho = reinterpret_cast<double*>(lds + P_::rho_off);
  cx<double>* F = reinterpret_cast<cx<double>*>(lds + P_::f_off);
  double* tmp = reinterpret_cast<double*>(lds + P_::tmp_off);
  double* E = reinterpret_cast<double*>(lds + P_::erow_off);
  for (int i = c.tid; i < nx; i += FT) {
    twl[i] = a.tw[i];
    okx[i] = a.okx[i];
  }
  field_density<FT>(c, a, 0, nx, rho, tmp);  // ends with a barrier
  // P lanes per output: the largest power of two <= FT / nx, at most 64
  int P = 1;
  while (2 * P * nx <= FT && P < 64) P *= 2;
  const int part = c.tid % P;
  const int G = FT / P;                  // outputs per round
  const int rounds = (nx + G - 1) / G;   // the same for every thread
  for (int rd = 0; rd < rounds; ++rd) {
    const int m0 = c.tid / P + rd * G;
    const int m = m0 < nx ? m0 : nx - 1;
    cx<double> s = mk<double>(0.0, 0.0);
    for (int j = part; j < nx; j += P) {
      const cx<double> w = twl[(int)(((long long)m * j) % nx)];
      s.re += rho[j] * w.re;
      s.im += rho[j] * w.im;
    }
    for (int o = 1; o < P; o <<= 1) {
      s.re += c.shfl_xor(s.re, o, P);
      s.im += c.shfl_xor(s.im, o, P);
    }
    if (part == 0 && m0 < nx) {
      const double g = 0.5 * (okx[m] - okx[(nx - m) % nx]);
      F[m] = mk<double>(-g * s.im, g * s.re);
    }
  }
  c.sync();
  for (int rd = 0; rd < rounds; ++rd) {
    const int x0 = c.tid / P + rd * G;
    const int x = x0 < nx ? x0 : nx - 1;
    double s = 0.0;
    for (int m = part; m < nx; m += P) {
      const cx<double> w = twl[(int)(((long long)m * x) % nx)];
      s += F[m].re * w.re + F[m].im * w.im;
    }
    for (int o = 1; o < P; o <<= 1) s += c.shfl_xor(s, o, P);
    if (part == 0 && x0 < nx) E[x] = s / nx + (a.driver ? a.driver[x] : 0.0);
  }
  c.sync();
  for (int r = c.tid; r < nrows; r += FT) {
    a.e_out[row0 + r] = E[row0 + r];
    if (a.e_out2) a.e_out2[row0 + r] = E[row0 + r];
  }
  if (row0 == 0 && a.series) {
    double acc[2] = {0.0, 0.0};
    for (int x = c.tid; x < nx; x += FT) {
      const double dr = a.driver ? a.driver[x] : 0.0;
      acc[0] += E[x] * E[x];
      acc[1] += dr * dr;
    }
    row_sum<FT, 2>(c, acc, tmp);
    if (c.tid == 0) {
      a.series[0] = acc[0] / nx;
      a.series[1] = acc[1] / nx;
    }
  }
  c.sync();
}

// ---------------------------------------------------------------------------
// row kernel
// ---------------------------------------------------------------------------
template <class T>
struct RowArgs {
  const T* f_in;
  T* f_out;          // null: f is not written back (moments only)
  int nx, nv;
  const double* e;  // [nx] field for the kicks
  double dt_a, dt_b;
  double dk;         // kv[1] of the fftfreq table
  double v0, dvv;    // v_m = v0 + m dvv
  double dv;
  const cx<T>* tw;   // twiddle tables of exp(-2 pi i n / nv), [lo | hi]
  double nu_dt;
  int op_dg;
  double* moments;  // [kMoments][nx]
  T* f_copy;        // [nx][nv] or null: f after the collision and moments (f itself under delta-f storage)
  int flags;
  const double* nu_dt_inst;  // batched scan: nu*dt per instance (rows are instance-major), or null
  int rows_per_inst;         // x rows per instance in a batch
  // field in the prologue (field_rows): E of the workgroup's rows from the
  // drift's partial densities instead of e (single grid, nx <= kProNx)
  int field_pro;
  FieldArgs fld;
  // delta-f storage: f holds f - B(v) for a fixed background B (bg [nv],
  // FP64; unnormalised spectrum bg_hat [nv] in the storage precision); the
  // kicks shift B spectrally as part of the deviation's multiply (BgPhaseMul),
  // the collision and moments see B + (f - B) in FP64
  const double* bg;
  const cx<T>* bg_hat;
};

template <class T, int M, int Q>
struct RowCfg {
  static constexpr int TPT = M / Q;
  static constexpr int L = 2 * Q;
  static constexpr int tw_bytes = (TwSize<2 * M, (int)sizeof(cx<T>)>::total * (int)sizeof(cx<T>) + 15) & ~15;
  static constexpr int fft_bytes = Padded<M>::size * (int)sizeof(cx<T>);
  static constexpr int red_bytes = 8 * TPT * 8;
  static constexpr int region = ((fft_bytes > red_bytes ? fft_bytes : red_bytes) + 15) & ~15;
  static constexpr int wg = TPT >= 64 ? TPT : 64;
  static constexpr int rows = wg / TPT;
  static constexpr int sum_bytes = (TPT > 64 ? TPT / 64 : 1) * kMoments * 8;
  // the field prologue (field_rows) runs before the twiddles and row buffers
  // are staged and overlays them from offset 0
  static constexpr int pro_bytes = ProCfg<wg>::bytes;
  static constexpr int main_bytes = tw_bytes + rows * region + sum_bytes;
  static constexpr int lds = main_bytes > pro_bytes ? main_bytes : pro_bytes;
  // minimum waves per SIMD asked of the register allocator: 3 (168 VGPRs,
  // small spills) measured faster than 2 for FP32 rows with M <= 1024 (the
  // scan's classes) and slower for FP64 and for M = 4096
  static constexpr int min_waves = (Q == 8 && sizeof(T) == 4 && M <= 1024) ? 3 : 1;
};

// BG: delta-f storage (a.bg, a.bg_hat set); a kernel of its own so that the
// plain step's register allocation is unchanged
template <class T, int M, int Q, class Ctx, bool BG = false>
VP_HD void row_body(Ctx& c, unsigned char* lds, const RowArgs<T>& a) {
  using Cfg = RowCfg<T, M, Q>;
  constexpr int TPT = Cfg::TPT;
  constexpr int L = Cfg::L;
  const int t = c.tid % TPT;
  const int rl = c.tid / TPT;
  const int row = c.bid * Cfg::rows + rl;
  const bool valid = row < a.nx;
  const int rowc = valid ? row : a.nx - 1;
  cx<T>* twl = reinterpret_cast<cx<T>*>(lds);
  unsigned char* base = lds + Cfg::tw_bytes;
  cx<T>* buf = reinterpret_cast<cx<T>*>(base + rl * Cfg::region);
  double* red = reinterpret_cast<double*>(base + rl * Cfg::region);
  double* rs = reinterpret_cast<double*>(base + Cfg::rows * Cfg::region);

  const cx<T>* src = reinterpret_cast<const cx<T>*>(a.f_in + (size_t)rowc * a.nv);
  cx<T> v[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) v[i] = src[t + i * TPT];
  double e = 0.0;
  if (a.field_pro) {
    const int row0 = c.bid * Cfg::rows;
    const int nr = a.nx - row0 < Cfg::rows ? a.nx - row0 : Cfg::rows;
    field_rows<Cfg::wg>(c, a.fld, lds, row0, nr);
    e = reinterpret_cast<const double*>(lds + ProCfg<Cfg::wg>::erow_off)[valid ? row : 0];
    c.sync();  // every thread has its E before the twiddles overwrite the prologue's LDS
  } else if (a.flags & (RK_KICK_A | RK_KICK_B)) {
    e = a.e[rowc];
  }
  tw_stage<T, 2 * M>(c, twl, a.tw);
  c.sync();
  const Tw<T, TwSize<2 * M, (int)sizeof(cx<T>)>::full> tw{twl};

  // kick by shift a: f(v) -> f(v - a), spectrally (the background's shift too
  // under delta-f storage)
  auto kick = [&](double shift) {
    if constexpr (BG)
      spectral_apply<T, M, TPT>(c, v, buf, tw, t, BgPhaseMul<T>(a.dk * shift, M, t, TPT, a.bg_hat));
    else
      spectral_apply<T, M, TPT>(c, v, buf, tw, t, PhaseMul<T>(a.dk * shift, M, t, TPT));
  };
  const bool chunk = (a.flags & (RK_COLLIDE | RK_MOMENTS)) != 0;
  // two half kicks by the same field with nothing between them (no
  // collision, moments or f store: the collisionless fused step) are one
  // kick by the sum of the shifts: the phase factors multiply exactly, so
  // one transform pair replaces two
  const bool merged = (a.flags & RK_KICK_A) && (a.flags & RK_KICK_B) && !chunk && !a.f_copy;
  if (a.flags & RK_KICK_A) kick((merged ? a.dt_a + a.dt_b : a.dt_a) * e);

  if (a.f_copy && !chunk && valid) {
    // copy only: natural order, one complex pair per value (coalesced)
    cx<T>* dst = reinterpret_cast<cx<T>*>(a.f_copy + (size_t)row * a.nv);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      cx<T> z = v[i];
      if constexpr (BG) {
        const int j = t + i * TPT;
        z = mk<T>((T)((double)z.re + a.bg[2 * j]), (T)((double)z.im + a.bg[2 * j + 1]));
      }
      dst[t + i * TPT] = z;
    }
  }
  if (chunk) {
    // natural order -> contiguous chunk of L samples per thread; after a
    // kick buf already holds the row in natural order (the inverse
    // transform's last pass ends by reading v from it)
    if (!(a.flags & RK_KICK_A)) {
#pragma unroll
      for (int i = 0; i < Q; ++i) buf[pidx(t + i * TPT)] = v[i];
      c.sync();
    }
    double d[L];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const cx<T> z = buf[pidx(t * Q + i)];
      d[2 * i] = (double)z.re;
      d[2 * i + 1] = (double)z.im;
    }
    c.sync();
    const int N = a.nv;
    const int m0 = t * L;
    if constexpr (BG) {
#pragma unroll
      for (int i = 0; i < L; ++i) d[i] += a.bg[m0 + i];
    }
    const double nu_dt = a.nu_dt_inst ? a.nu_dt_inst[rowc / a.rows_per_inst] : a.nu_dt;
    if ((a.flags & RK_COLLIDE) && nu_dt > 0.0) {
      double vbar, temp;
      coll_moments<TPT, L>(c, d, a.v0, a.dvv, N, a.dv, a.op_dg != 0, t, rs, vbar, temp);
      const double dv2 = a.dv * a.dv;
      CollCoef cf;
      cf.N = N;
      cf.set(nu_dt, -temp / dv2, 0.5 / a.dv, a.v0, a.dvv, vbar, 1.0 + nu_dt * (2.0 * temp / dv2));
      tridiag<TPT, L, CollCoef, Ctx, sizeof(T) == 8>(c, d, cf, t, red);
    }
    if (a.flags & RK_MOMENTS) {
      double s[kMoments];
#pragma unroll
      for (int q = 0; q < kMoments; ++q) s[q] = 0.0;
      const double vb = a.v0 + (double)m0 * a.dvv;
#pragma unroll
      for (int i = 0; i < L; ++i) {
        const double wf = chunk_w<TPT, L>(i, t) * d[i];
        const double vm = vb + (double)i * a.dvv;
        double p = wf;
        s[0] += p;
        p *= vm;
        s[1] += p;
        p *= vm;
        s[2] += p;
        p *= vm;
        s[3] += p;
        p *= vm;
        s[4] += p;
        p *= vm;
        s[5] += p;
        s[6] += wf * d[i];
        // f log f: in the storage precision (FP32 log on the FP32 path)
        s[7] += wf * (sizeof(T) == 8 ? log_f64(d[i]) : (double)logf((float)d[i]));
      }
      row_sum<TPT, kMoments>(c, s, rs);
      if (valid && t == 0) {
#pragma unroll
        for (int q = 0; q < kMoments; ++q) a.moments[(size_t)q * a.nx + row] = a.dv * s[q];
      }
    }
    if (a.f_copy && valid) {
      // f after the collision (the reference's per-step f store); d holds f
      // itself here also under delta-f storage
      T* dst = a.f_copy + (size_t)row * a.nv + m0;
#pragma unroll
      for (int i = 0; i < L; ++i) dst[i] = (T)d[i];
    }
    if constexpr (BG) {
#pragma unroll
      for (int i = 0; i < L; ++i) d[i] -= a.bg[m0 + i];
    }
    // chunk -> natural order
#pragma unroll
    for (int i = 0; i < Q; ++i) buf[pidx(t * Q + i)] = mk<T>((T)d[2 * i], (T)d[2 * i + 1]);
    c.sync();
#pragma unroll
    for (int i = 0; i < Q; ++i) v[i] = buf[pidx(t + i * TPT)];
    c.sync();
  }

  if ((a.flags & RK_KICK_B) && !merged) kick(a.dt_b * e);

  if (valid && a.f_out) {
    cx<T>* dst = reinterpret_cast<cx<T>*>(a.f_out + (size_t)row * a.nv);
#pragma unroll
    for (int i = 0; i < Q; ++i) dst[t + i * TPT] = v[i];
  }
}

// Several workgroups: each reduces an x slice into rho_g; the last one to
// arrive (device-scope counter) continues with the whole rho.  Returns false
// in the workgroups that are done.
template <int FT, class Ctx>
VP_HD bool field_gather(Ctx& c, const FieldArgs& a, double* rho, double* tmp, int* flag) {
  const int nx = a.nx;
  if (c.nblk == 1) {
    field_density<FT>(c, a, 0, nx, rho, tmp);
    return true;
  }
  const int XS = nx / c.nblk;
  const int x0 = c.bid * XS;
  field_density<FT>(c, a, x0, XS, rho, tmp);
  for (int xl = c.tid; xl < XS; xl += FT) a.rho_g[x0 + xl] = rho[xl];
  c.fence();
  c.sync();
  if (c.tid == 0) *flag = (c.arrive(a.counter) == (unsigned)(c.nblk - 1)) ? 1 : 0;
  c.sync();
  if (!*flag) return false;
  c.fence();
  for (int x = c.tid; x < nx; x += FT) rho[x] = a.rho_g[x];
  if (c.tid == 0) c.reset(a.counter);
  c.sync();
  return true;
}

template <int FT, class Ctx>
VP_HD void field_finish(Ctx& c, const FieldArgs& a, int x, double ex, double* rs, double* acc) {
  const double dr = a.driver ? a.driver[x] : 0.0;
  const double e = ex + dr;
  a.e_out[x] = e;
  if (a.e_out2) a.e_out2[x] = e;
  acc[0] += e * e;
  acc[1] += dr * dr;
}

template <int M, int Q, int FT>
struct FieldCfg {
  static constexpr int TPT = M / Q;
  static constexpr int nx = 2 * M;
  static constexpr int tmp_n = (nx > FT ? nx : FT);
  static constexpr int tw_bytes = TwSize<nx, 16>::total * 16;
  static constexpr int work = (Padded<M>::size * 16 > tmp_n * 8) ? Padded<M>::size * 16 : tmp_n * 8;
  // [twiddles | okx (nx) | rho (nx) | fft buffer / reduction scratch | row_sum scratch | flag]
  static constexpr int lds = tw_bytes + 2 * nx * 8 + work + 64 + 16;
};

// TAIL: called by the last workgroup of the drift kernel, which reads every
// partial density row itself (coherent loads) instead of the slice hand-off
template <int M, int Q, int FT, bool TAIL = false, class Ctx>
VP_HD void field_body(Ctx& c, unsigned char* lds, const FieldArgs& a) {
  using Cfg = FieldCfg<M, Q, FT>;
  constexpr int TPT = Cfg::TPT;
  static_assert(TPT <= FT, "field transform needs at most FT threads");
  cx<double>* twl = reinterpret_cast<cx<double>*>(lds);
  double* okx = reinterpret_cast<double*>(lds + Cfg::tw_bytes);
  double* rho = okx + Cfg::nx;
  cx<double>* buf = reinterpret_cast<cx<double>*>(rho + Cfg::nx);
  double* tmp = reinterpret_cast<double*>(buf);
  double* rs = reinterpret_cast<double*>(lds + Cfg::lds - 80);
  int* flag = reinterpret_cast<int*>(lds + Cfg::lds - 16);
  if constexpr (TAIL) {
    // the caller's acquire fence makes the other workgroups' rows visible
    field_density<FT, false>(c, a, 0, Cfg::nx, rho, tmp);
  } else {
    if (!field_gather<FT>(c, a, rho, tmp, flag)) return;
  }
  tw_stage<double, Cfg::nx>(c, twl, a.tw);
  for (int x = c.tid; x < Cfg::nx; x += FT) okx[x] = a.okx[x];
  c.sync();
  const Tw<double, TwSize<Cfg::nx, 16>::full> tw{twl};
  const bool act = c.tid < TPT;
  const int t = act ? c.tid : 0;
  cx<double> v[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int j = t + i * TPT;
    v[i] = mk<double>(rho[2 * j], rho[2 * j + 1]);
  }
  PoissonMul<double> mul;
  mul.okx = okx;
  mul.M = M;
  spectral_apply<double, M, TPT>(c, v, buf, tw, t, mul, act);
  double acc[2] = {0.0, 0.0};
  if (act) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int j = t + i * TPT;
      field_finish<FT>(c, a, 2 * j, v[i].re, rs, acc);
      field_finish<FT>(c, a, 2 * j + 1, v[i].im, rs, acc);
    }
  }
  if (a.series) {
    row_sum<FT, 2>(c, acc, rs);
    if (c.tid == 0) {
      a.series[0] = acc[0] / a.nx;
      a.series[1] = acc[1] / a.nx;
    }
  }
}

// direct-DFT field solve for any nx (used when nx is not a power of two)
template <int FT, class Ctx>
VP_HD void field_dft_body(Ctx& c, unsigned char* lds, const FieldArgs& a) {
  const int nx = a.nx;
  double* rho = reinterpret_cast<double*>(lds);
  cx<double>* F = reinterpret_cast<cx<double>*>(lds + ((nx * 8 + 15) & ~15));
  double* tmp = reinterpret_cast<double*>(F);
  double* rs = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(F) + (size_t)(nx > FT ? nx : FT) * 16);
  field_density<FT>(c, a, 0, nx, rho, tmp);
  for (int m = c.tid; m < nx; m += FT) {
    cx<double> s = mk<double>(0.0, 0.0);
    for (int j = 0; j < nx; ++j) {
      const cx<double> w = a.tw[(int)(((long long)m * j) % nx)];
      s.re += rho[j] * w.re;
      s.im += rho[j] * w.im;
    }
    // effective multiplier Re-projection: (G_m + conj G_{nx-m}) / 2, G = i okx
    const double g = 0.5 * (a.okx[m] - a.okx[(nx - m) % nx]);
    F[m] = mk<double>(-g * s.im, g * s.re);
  }
  c.sync();
  double acc[2] = {0.0, 0.0};
  for (int x = c.tid; x < nx; x += FT) {
    double s = 0.0;
    for (int m = 0; m < nx; ++m) {
      const cx<double> w = a.tw[(int)(((long long)m * x) % nx)];
      // Re(F_m conj(w))
      s += F[m].re * w.re + F[m].im * w.im;
    }
    field_finish<FT>(c, a, x, s / nx, rs, acc);
  }
  if (a.series) {
    row_sum<FT, 2>(c, acc, rs);
    if (c.tid == 0) {
      a.series[0] = acc[0] / a.nx;
      a.series[1] = acc[1] / a.nx;
    }
  }
}

// ---------------------------------------------------------------------------
// drift kernel (x-advection of C columns) + partial density sums
// ---------------------------------------------------------------------------
template <class T>
struct DriftArgs {
  const T* f_in;
  T* f_out;
  int nx, nv;
  double dt;
  double dk;         // kx[1] of the fftfreq table
  const double* v;   // [nv]
  const cx<T>* tw;   // twiddle tables of exp(-2 pi i n / nx), [lo | hi]
  double* dens;      // [gridDim][nx] partial sums (unscaled), or null
  int fuse_field;    // 1: the last workgroup to finish solves the field (fld)
  FieldArgs fld;     // field arguments for the fused tail (part = dens)
  // batched scan: grid y = instance; f and dens are offset per instance
  // (the fused field tail takes field_instance of fld); dk_inst[inst]
  // replaces dk
  const double* dk_inst;
  int xcd_order;  // 1: consecutive column tiles go to workgroups on one XCD
  // FP32, nx in {32, 64}: the x-transforms as dense real-DFT products on
  // MFMA (vpfp_drift.hip k_drift_dft) with these matrices, [2][nx][nx]:
  // forward rows k <= nx/2 cos(2 pi k x / nx), rows nx/2 + j -sin(2 pi j x / nx)
  // (0 < j < nx/2); inverse [x][k] the matching real synthesis / nx.
  // Null: the Stockham transform (drift_body).
  const float* dft;
};

// Workgroup b of n runs on XCD b % 8 (round-robin dispatch).  With
// xcd_order the workgroups of one XCD take consecutive tiles, so the two
// halves of a 128-byte line read by neighbouring tiles meet in one L2.
VP_HD int xcd_block(int b, int n) { return (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3); }

// values per thread of the x-advection transforms (8 for M = 32 and 64,
// two Stockham passes instead of three, measured no faster on the scan)
VP_HD constexpr int drift_q(int M) { return M <= 64 ? (M < 4 ? M : 4) : 8; }

template <class T, int M, int Q, int C>
struct DriftCfg {
  static constexpr int TPT = M / Q;
  static constexpr int wg = TPT * C;
  // odd stride (in complex elements): the C column buffers of a tile start
  // on different LDS banks (an even multiple of 64 banks made the columns a
  // wave touches collide)
  static constexpr int stride = Padded<M>::size | 1;
  static constexpr int tw_bytes = (TwSize<2 * M, (int)sizeof(cx<T>)>::total * (int)sizeof(cx<T>) + 15) & ~15;
  static constexpr int own = tw_bytes + C * stride * (int)sizeof(cx<T>);
  // fused field tail (small grids): M-point Poisson transform on the same workgroup
  static constexpr int fq = M < 8 ? M : 8;
  static constexpr bool tail = M <= 128 && (M / fq) <= wg;
  static constexpr int field_lds = tail ? FieldCfg<M, fq, wg>::lds : 0;
  static constexpr int lds = own > field_lds ? own : field_lds;
};

template <class T, int M, int Q, int C, class Ctx>
VP_HD void drift_body(Ctx& c, unsigned char* lds, const DriftArgs<T>& a) {
  using Cfg = DriftCfg<T, M, Q, C>;
  constexpr int TPT = Cfg::TPT;
  const int cc = c.tid % C;
  const int t = c.tid / C;
  cx<T>* twl = reinterpret_cast<cx<T>*>(lds);
  cx<T>* buf = reinterpret_cast<cx<T>*>(lds + Cfg::tw_bytes) + cc * Cfg::stride;
  const size_t nv = (size_t)a.nv;
  const int inst = c.bidy;
  const size_t foff = (size_t)inst * (size_t)a.nx * nv;
  const double dk = a.dk_inst ? a.dk_inst[inst] : a.dk;
  tw_stage<T, 2 * M>(c, twl, a.tw);
  const Tw<T, TwSize<2 * M, (int)sizeof(cx<T>)>::full> tw{twl};
  double acc[2 * Q];
#pragma unroll
  for (int i = 0; i < 2 * Q; ++i) acc[i] = 0.0;
  const int ntiles = a.nv / C;
  // workgroup b takes the run of `per` consecutive column tiles starting at
  // b per (per = ntiles / workgroups rounded up): neighbouring narrow tiles
  // share 128-byte lines, which one workgroup then fetches into one L2;
  // it accumulates the density of its tiles into one partial row dens[b][x]
  const int b0 = a.xcd_order ? xcd_block(c.bid, c.nblk) : c.bid;
  const int per = (ntiles + c.nblk - 1) / c.nblk;
  const int t0 = b0 * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
  // the next tile's columns are loaded before this tile's transforms, so
  // its reads are in flight during them
  cx<T> v[Q], nxt[Q];
  double vc_next = 0.0;
  auto load = [&](int tile, cx<T>* dst, double& vcol) {
    const T* fi = a.f_in + foff + tile * C + cc;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int j = t + i * TPT;
      dst[i] = mk<T>(fi[(size_t)(2 * j) * nv], fi[(size_t)(2 * j + 1) * nv]);
    }
    vcol = a.v[tile * C + cc];
  };
  if (t0 < t1) load(t0, nxt, vc_next);
  for (int tile = t0; tile < t1; ++tile) {
    const int col = tile * C + cc;
#pragma unroll
    for (int i = 0; i < Q; ++i) v[i] = nxt[i];
    const double vc = vc_next;
    if (tile + 1 < t1) load(tile + 1, nxt, vc_next);
    c.sync();  // twiddle tables staged (first tile) / buffer free (later tiles)
    spectral_apply<T, M, TPT>(c, v, buf, tw, t, PhaseMul<T>((dk * a.dt) * vc, M, t, TPT));
    T* fo = a.f_out + foff + col;
    const double w = trapw(col, a.nv);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int j = t + i * TPT;
      fo[(size_t)(2 * j) * nv] = v[i].re;
      fo[(size_t)(2 * j + 1) * nv] = v[i].im;
      acc[2 * i] += w * (double)v[i].re;
      acc[2 * i + 1] += w * (double)v[i].im;
    }
  }
  if (a.dens) {
    double* out = a.dens + ((size_t)inst * c.nblk + c.bid) * a.nx;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      double s2[2] = {acc[2 * i], acc[2 * i + 1]};
#pragma unroll
      for (int m = 1; m < C; m <<= 1) {
        s2[0] += c.shfl_xor(s2[0], m, C);
        s2[1] += c.shfl_xor(s2[1], m, C);
      }
      if (cc == 0) {
        const int j = t + i * TPT;
        out[2 * j] = s2[0];
        out[2 * j + 1] = s2[1];
      }
    }
  }
  if constexpr (Cfg::tail) {
    if (a.fuse_field) {
      // last workgroup to arrive reduces every partial row and solves Poisson
      int* flag = reinterpret_cast<int*>(lds + Cfg::lds - 16);
      const FieldArgs fa = field_instance(a.fld, inst);
      c.fence();
      c.sync();
      if (c.tid == 0) *flag = (c.arrive(fa.counter) == (unsigned)(c.nblk - 1)) ? 1 : 0;
      c.sync();
      if (!*flag) return;
      c.fence();
      field_body<M, Cfg::fq, Cfg::wg, true>(c, lds, fa);
      if (c.tid == 0) c.reset(fa.counter);
    }
  }
}

// ---------------------------------------------------------------------------
// x-advection of FP32 columns as two dense real-DFT products on MFMA
// (nx = NX in {32, 64}).  Each wave takes a tile of 16 columns:
//   Y = F X      (F: [NX][NX] forward real DFT, X: [NX][16] the columns)
//   phase        (Re_k, Im_k) of column c times exp(-i k phi_c)
//   X' = G Y'    (G: [NX][NX] inverse real DFT)
// through the context's 16 x 16 x 4 FP32 matrix product (c.mfma_16x16x4:
// D = A B + C, A 16x4, B 4x16; lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[4 (l / 16) + r][l % 16], r < 4).  Rows of Y:
// Re_k for k <= NX/2 (row NX/2 the Nyquist bin), Im_k at row NX/2 + k
// (0 < k < NX/2), so a lane holds Re_k and Im_k of one column in tiles t and
// t + NX/32 at the same position.  The phase is the Stockham drift's:
// phi = (dk dt) v_c, exp(-i k phi) for 0 < k < NX/2 and cos((NX/2) phi) on
// the Nyquist bin (the real part the reference keeps).  The B operand of the
// second product is re-laid through a per-wave LDS tile.  W waves per
// workgroup; the workgroups stride over groups of W tiles (the tile count a
// multiple of W, so every wave makes the same number of passes).
// ---------------------------------------------------------------------------
struct f4 {
  float v[4];
};

template <int NX, int W>
struct DftCfg {
  static constexpr int T = NX / 16;   // 16-row tiles of the products
  static constexpr int KS = NX / 4;   // k-steps of 4
  static constexpr int GP = 17;       // pitch (floats) of a wave's Y' tile
  static constexpr int mat = 2 * NX * NX * 4;
  static constexpr int gt = NX * GP * 4;
  static constexpr int lds = mat + W * gt + W * NX * 8;
};

template <int NX, int W, class Ctx>
VP_HD void drift_dft_body(Ctx& c, unsigned char* lds, const DriftArgs<float>& a) {
  using Cfg = DftCfg<NX, W>;
  constexpr int T = Cfg::T, KS = Cfg::KS, GP = Cfg::GP;
  float* Fw = reinterpret_cast<float*>(lds);
  float* Iv = Fw + NX * NX;
  const int lane = c.tid & 63, w = c.tid >> 6;
  float* Y = reinterpret_cast<float*>(lds + Cfg::mat) + w * NX * GP;
  double* red = reinterpret_cast<double*>(lds + Cfg::mat + W * Cfg::gt);
  for (int i = c.tid; i < 2 * NX * NX; i += c.nthr) Fw[i] = a.dft[i];
  const int li = lane & 15, lk = lane >> 4;
  const size_t nv = (size_t)a.nv;
  const int inst = c.bidy;
  const size_t foff = (size_t)inst * NX * nv;
  const double dk = a.dk_inst ? a.dk_inst[inst] : a.dk;
  double dacc[T][4];
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) dacc[t][r] = 0.0;
  // workgroup b takes a run of consecutive passes (neighbouring 16-column
  // tiles share 128-byte lines)
  const int groups = (a.nv / 16) / W;
  const int per = (groups + c.nblk - 1) / c.nblk;
  const int g0 = c.bid * per, g1 = g0 + per < groups ? g0 + per : groups;
  c.sync();  // matrices staged
  for (int gi = g0; gi < g1; ++gi) {
    const int col = (gi * W + w) * 16 + li;
    const float* fi = a.f_in + foff + col;
    f4 acc[T];
#pragma unroll
    for (int t = 0; t < T; ++t) acc[t] = f4{{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const float xb = fi[(size_t)(4 * s + lk) * nv];
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = c.mfma_16x16x4(Fw[(16 * t + li) * NX + 4 * s + lk], xb, acc[t]);
    }
    // phase: bins k = 16 t + 4 lk + r (t < T/2) by a running product,
    // exp(-i k phi) = exp(-i 4 lk phi) exp(-i r phi) exp(-i 16 t phi)
    const double phi = (dk * a.dt) * a.v[col];
    cx<float> p0, p1, p16;
    cis_neg((double)(4 * lk) * phi, p0);
    cis_neg(phi, p1);
    cis_neg(16.0 * phi, p16);
    cx<float> pt = p0;
#pragma unroll
    for (int t = 0; t < T / 2; ++t) {
      cx<float> pk = pt;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float re = acc[t].v[r], im = acc[t + T / 2].v[r];
        if (t == 0 && r == 0 && lk == 0) {
          // DC unchanged; the partner row holds the real Nyquist bin
          cx<float> pn;
          cis_neg((double)(NX / 2) * phi, pn);
          acc[T / 2].v[0] = pn.re * im;
        } else {
          acc[t].v[r] = re * pk.re - im * pk.im;
          acc[t + T / 2].v[r] = im * pk.re + re * pk.im;
        }
        pk = mul(pk, p1);
      }
      pt = mul(pt, p16);
    }
    // Y' to the wave's LDS tile, rows 16 t + 4 lk + r, column li
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) Y[(16 * t + 4 * lk + r) * GP + li] = acc[t].v[r];
    c.sync();
    f4 out[T];
#pragma unroll
    for (int t = 0; t < T; ++t) out[t] = f4{{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const float yb = Y[(4 * s + lk) * GP + li];
#pragma unroll
      for (int t = 0; t < T; ++t) out[t] = c.mfma_16x16x4(Iv[(16 * t + li) * NX + 4 * s + lk], yb, out[t]);
    }
    c.sync();  // Y read by every lane before the next pass writes it
    float* fo = a.f_out + foff + col;
    const double wc = (col == 0 || col == a.nv - 1) ? 0.5 : 1.0;
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        fo[(size_t)(16 * t + 4 * lk + r) * nv] = out[t].v[r];
        dacc[t][r] += wc * (double)out[t].v[r];
      }
  }
  if (!a.dens) return;
  // partial density of this workgroup: sum over the 16 columns of a lane
  // group, then over the waves
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double v = dacc[t][r];
#pragma unroll
      for (int m = 1; m < 16; m <<= 1) v += c.shfl_xor(v, m, 16);
      if (li == 0) red[w * NX + 16 * t + 4 * lk + r] = v;
    }
  c.sync();
  double* out = a.dens + ((size_t)inst * c.nblk + c.bid) * NX;
  for (int x = c.tid; x < NX; x += c.nthr) {
    double sum = 0.0;
#pragma unroll
    for (int ww = 0; ww < W; ++ww) sum += red[ww * NX + x];
    out[x] = sum;
  }
}

// ---------------------------------------------------------------------------
// density rows: part[x] = scale * sum_v w_v f(x, v)   (one wavefront per row)
// ---------------------------------------------------------------------------
template <class T, class Ctx>
VP_HD void density_body(Ctx& c, const T* f, int nx, int nv, double scale, double* part) {
  const int lane = c.tid & 63;
  const int row = c.bid * (c.nthr / 64) + (c.tid >> 6);
  const int rowc = row < nx ? row : nx - 1;
  const T* fr = f + (size_t)rowc * nv;
  double s[1] = {0.0};
  for (int m = lane; m < nv; m += 64) s[0] += trapw(m, nv) * (double)fr[m];
#pragma unroll
  for (int k = 1; k < 64; k <<= 1) s[0] += c.shfl_xor(s[0], k, 64);
  if (row < nx && lane == 0) part[row] = scale * s[0];
}

// ---------------------------------------------------------------------------
// first nm Fourier modes along x of every v column:
//   out[m][v] = sum_x f(x, v) exp(-2 pi i m x / nx)     (step.py:130-135)
// one thread per (m, v); the x sum is sequential in x order
// ---------------------------------------------------------------------------
template <class T, class Ctx>
VP_HD void xmodes_body(Ctx& c, const T* f, int nx, int nv, int nm, const cx<double>* tw, cx<double>* out) {
  const int idx = c.bid * c.nthr + c.tid;
  if (idx >= nm * nv) return;
  const int m = idx / nv, col = idx % nv;
  double re = 0.0, im = 0.0;
  // r = m x mod nx by increments: the loads of f and of the twiddles do not
  // wait on an index division, so consecutive iterations overlap
  int r = 0;
#pragma unroll 16
  for (int x = 0; x < nx; ++x) {
    const cx<double> w = tw[r];
    const double fv = (double)f[(size_t)x * nv + col];
    re += fv * w.re;
    im += fv * w.im;
    r += m;
    if (r >= nx) r -= nx;
  }
  out[idx] = mk<double>(re, im);
}

// ---------------------------------------------------------------------------
// collision diagonals a, b, c of every row (one wavefront per row)
// ---------------------------------------------------------------------------
template <class T, class Ctx>
VP_HD void diag_body(Ctx& c, const T* f, int nx, int nv, const double* v, double dv, double nu_dt, int op_dg,
                     double* A, double* B, double* Cc) {
  const int lane = c.tid & 63;
  const int row = c.bid * (c.nthr / 64) + (c.tid >> 6);
  const int rowc = row < nx ? row : nx - 1;
  const T* fr = f + (size_t)rowc * nv;
  double vbar = 0.0;
  double s[1] = {0.0};
  if (op_dg) {
    for (int m = lane; m < nv; m += 64) s[0] += trapw(m, nv) * (double)fr[m] * v[m];
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) s[0] += c.shfl_xor(s[0], k, 64);
    vbar = dv * s[0];
    s[0] = 0.0;
  }
  for (int m = lane; m < nv; m += 64) {
    const double w = v[m] - vbar;
    s[0] += trapw(m, nv) * (double)fr[m] * (w * w);
  }
#pragma unroll
  for (int k = 1; k < 64; k <<= 1) s[0] += c.shfl_xor(s[0], k, 64);
  const double temp = dv * s[0];
  if (row >= nx) return;
  const double dv2 = dv * dv;
  const double A0 = -temp / dv2, H = 0.5 / dv;
  const double bd = 1.0 + nu_dt * (2.0 * temp / dv2);
  double* ar = A + (size_t)row * (nv - 1);
  double* cr = Cc + (size_t)row * (nv - 1);
  double* br = B + (size_t)row * nv;
  for (int m = lane; m < nv; m += 64) {
    br[m] = bd;
    if (m < nv - 1) {
      ar[m] = nu_dt * (A0 + (v[m] - vbar) * H);
      cr[m] = nu_dt * (A0 - (v[m + 1] - vbar) * H);
    }
  }
}

// ---------------------------------------------------------------------------
// tridiagonal solve with explicit diagonals (batched over rows)
// ---------------------------------------------------------------------------
template <class T>
struct TriArgs {
  const double* a;  // [nx][nv-1]
  const double* b;  // [nx][nv]
  const double* cc; // [nx][nv-1]
  const T* d;       // [nx][nv]
  T* x;             // [nx][nv]
  int nx, nv;
};

template <class T, int M, int Q, class Ctx>
VP_HD void tri_body(Ctx& c, unsigned char* lds, const TriArgs<T>& a) {
  using Cfg = RowCfg<T, M, Q>;
  constexpr int TPT = Cfg::TPT;
  constexpr int L = Cfg::L;
  const int t = c.tid % TPT;
  const int rl = c.tid / TPT;
  const int row = c.bid * Cfg::rows + rl;
  const bool valid = row < a.nx;
  const int rowc = valid ? row : a.nx - 1;
  double* red = reinterpret_cast<double*>(lds + rl * Cfg::region);
  const int N = a.nv;
  const T* dr = a.d + (size_t)rowc * N;
  double d[L];
#pragma unroll
  for (int i = 0; i < L; ++i) d[i] = (double)dr[t * L + i];
  ArrCoef cf;
  cf.a = a.a + (size_t)rowc * (N - 1);
  cf.bb = a.b + (size_t)rowc * N;
  cf.cc = a.cc + (size_t)rowc * (N - 1);
  cf.N = N;
  tridiag<TPT, L, ArrCoef, Ctx, sizeof(T) == 8>(c, d, cf, t, red);
  if (valid) {
    T* xr = a.x + (size_t)row * N;
#pragma unroll
    for (int i = 0; i < L; ++i) xr[t * L + i] = (T)d[i];
  }
}

// ---------------------------------------------------------------------------
// several batches in one launch (vpfp_group_*): each kernel takes a table of
// segments, one per batch, and its workgroups are numbered across them;
// segment s owns the workgroups [blk0, blk0 of segment s+1).  The wrapper
// picks the workgroup's segment, rebases the arguments and runs the body
// with the workgroup index local to the segment.
// ---------------------------------------------------------------------------
constexpr int kMaxSeg = 8;

template <class T>
struct RowSeg {
  T* f;  // read and written in place
  const double* e;
  const double* nu_dt;  // per instance
  int rows;             // x rows of the batch (instances * nx)
  int rows_per_inst;
  int blk0;
};
template <class T>
struct RowMulti {
  RowArgs<T> base;  // shared: flags, dt, v grid, tables (one nv)
  int nseg;
  RowSeg<T> seg[kMaxSeg];
};

template <class T>
struct DriftSeg {
  T* f;
  const double* v;
  const double* dk;  // per instance
  double* dens;
  int nv;
  int gx;            // workgroups per instance
  int n_inst;
  int blk0;
};
template <class T>
struct DriftMulti {
  DriftArgs<T> base;  // shared: nx, dt, twiddles
  int nseg;
  DriftSeg<T> seg[kMaxSeg];
};

struct FieldSeg {
  const double* part;
  double scale;  // dv of the batch's v grid
  int ntiles;
  const double* okx;
  double* e_out;
  double* series;
  int n_inst;
  int blk0;
};
struct FieldMulti {
  FieldArgs base;  // shared: nx, scale, twiddles
  int nseg;
  FieldSeg seg[kMaxSeg];
};

// segment of workgroup b (blk0 ascending, seg[0].blk0 == 0)
template <class S>
VP_HD int seg_of(const S* seg, int nseg, int b) {
  int s = 0;
  for (int i = 1; i < nseg; ++i)
    if (b >= seg[i].blk0) s = i;
  return s;
}

}  // namespace vp
