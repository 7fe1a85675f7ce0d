// HIP kernels and the C ABI (include/vpfp.h) of the VPFP step on gfx950.
//
// This file holds the field, density, diagonal, x-mode, background and
// tridiagonal kernels, the plan and batch objects, the autotuner and the C
// ABI; the row kernels are in vpfp_row.hip and the drift kernels in
// vpfp_drift.hip (one object per storage type, compiled in parallel).
//
// Launch shapes (wave64):
//   row kernel   M = nv/2 complex points per row, Q = min(8, M) per thread,
//                TPT = M/Q threads per row, max(1, 64/TPT) rows per workgroup
//   drift kernel M = nx/2, Q = min(8, M), TPT = M/Q threads per column,
//                C columns per workgroup (16 for nx <= 256, 8 for nx = 512,
//                4 above), one partial density row per workgroup
//   field kernel one workgroup of 256 threads
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vpfc.h"
#include "vpfp.h"
#include "vpfp_launch.h"

using vp::cx;

namespace vpfp_impl {

namespace {
thread_local std::string g_err;
}

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(VPFP_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return VPFP_OK;
}

const char* last_error() { return g_err.c_str(); }

}  // namespace vpfp_impl

using namespace vpfp_impl;

namespace {

#define VPFP_HIP(x)                                                                 \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return fail(VPFP_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ kernels
constexpr int kFT = 256;

// grid (nblk, 1): one field solve over nblk workgroups; grid (1, n): the n
// instances of a batch, one workgroup each
template <int M, int Q>
__global__ __launch_bounds__(kFT) void k_field(vp::FieldArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  VPFP_CTX;
  if (a.xm_out) {
    // the previous step's f store beside the solve: workgroups from
    // xm_blk0 on take one (mode, v) pair per thread
    if (c.bid >= a.xm_blk0) {
      GpuCtx x{c.tid, c.nthr, c.bid - a.xm_blk0, c.nblk - a.xm_blk0, 0};
      if (a.xm_f32)
        vp::xmodes_body<float>(x, static_cast<const float*>(a.xm_f), a.nx, a.xm_nv, a.xm_n, a.xm_tw, a.xm_out);
      else
        vp::xmodes_body<double>(x, static_cast<const double*>(a.xm_f), a.nx, a.xm_nv, a.xm_n, a.xm_tw, a.xm_out);
      return;
    }
    c.nblk = a.xm_blk0;
  }
  vp::field_body<M, Q, kFT>(c, lds, vp::field_instance(a, c.bidy));
}

// several batches of one nx in one launch (vpfp_group_step_n): one
// workgroup per instance, the segment giving the batch's arrays
template <int M, int Q>
__global__ __launch_bounds__(kFT) void k_field_multi(vp::FieldMulti m) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int s = vp::seg_of(m.seg, m.nseg, (int)blockIdx.x);
  const int inst = (int)blockIdx.x - m.seg[s].blk0;
  vp::FieldArgs a = m.base;
  a.part = m.seg[s].part;
  a.scale = m.seg[s].scale;
  a.ntiles = m.seg[s].ntiles;
  a.okx = m.seg[s].okx;
  a.e_out = m.seg[s].e_out;
  a.series = m.seg[s].series;
  GpuCtx c{(int)threadIdx.x, (int)blockDim.x, 0, 1, inst};
  vp::field_body<M, Q, kFT>(c, lds, vp::field_instance(a, inst));
}

__global__ __launch_bounds__(kFT) void k_field_dft(vp::FieldArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  VPFP_CTX;
  vp::field_dft_body<kFT>(c, lds, a);
}

// f[r][m] += sign * bg[m] on every row (delta-f storage conversion)
template <class T>
__global__ __launch_bounds__(256) void k_background(T* f, size_t n, int nv, const double* bg, double sign) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = (int)(i % (size_t)nv);
  f[i] = (T)((double)f[i] + sign * bg[m]);
}

template <class T>
__global__ __launch_bounds__(256) void k_density(const T* f, int nx, int nv, double scale, double* part) {
  VPFP_CTX;
  vp::density_body<T>(c, f, nx, nv, scale, part);
}

template <class T>
__global__ __launch_bounds__(256) void k_diag(const T* f, int nx, int nv, const double* v, double dv, double nu_dt,
                                              int op_dg, double* A, double* B, double* Cc) {
  VPFP_CTX;
  vp::diag_body<T>(c, f, nx, nv, v, dv, nu_dt, op_dg, A, B, Cc);
}

template <class T>
__global__ __launch_bounds__(256) void k_xmodes(const T* f, int nx, int nv, int nm, const cx<double>* tw,
                                                cx<double>* out) {
  VPFP_CTX;
  vp::xmodes_body<T>(c, f, nx, nv, nm, tw, out);
}

template <class T, int M, int Q>
__global__ __launch_bounds__((vp::RowCfg<T, M, Q>::wg)) void k_tri(vp::TriArgs<T> a) {
  extern __shared__ __align__(16) unsigned char lds[];
  VPFP_CTX;
  vp::tri_body<T, M, Q>(c, lds, a);
}

// Twiddles exp(-2 pi i n / N) built on the device (vpfp_field_noalloc):
// split = 0: the full table w[n], n < N; split = 1: [lo | hi] with
// lo[j] = W_N^j (j < 64), hi[h] = W_N^{64 h} (h < N / 64), the layout of
// split_twiddles for N > vp::kTwFull
__global__ __launch_bounds__(256) void k_twiddles(cx<double>* w, int N, int split) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_entries = split ? 64 + (N / 64 > 1 ? N / 64 : 1) : N;
  if (i >= n_entries) return;
  const int n = split ? (i < 64 ? i : 64 * (i - 64)) : i;
  double sn, cs;
  sincospi(2.0 * (double)n / (double)N, &sn, &cs);
  w[i] = vp::mk<double>(cs, -sn);
}

// ------------------------------------------------------------------ launch helpers
// drift workgroups: one per `per` column tiles (at most 1024 for a single
// grid, which then strides); each workgroup writes one partial density row
constexpr int kDriftMaxWG = 1024;
int drift_grid(int ntiles) { return ntiles < kDriftMaxWG ? ntiles : kDriftMaxWG; }
int drift_wgs(int nv, int C, int per, bool batch) {
  const int tiles = nv / C, w = (tiles + per - 1) / (per > 0 ? per : 1);
  return batch ? w : drift_grid(w);
}

// field workgroups: one per 64 x values (the last to finish solves Poisson)
int field_grid(int nx) { return nx >= 128 ? nx / 64 : 1; }

template <class T, int M>
int launch_tri(const vp::TriArgs<T>& a, hipStream_t s) {
  constexpr int Q = qfor(M);
  using Cfg = vp::RowCfg<T, M, Q>;
  auto kern = k_tri<T, M, Q>;
  if (int rc = set_lds(kern, Cfg::lds)) return rc;
  const int grid = (a.nx + Cfg::rows - 1) / Cfg::rows;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(Cfg::wg), Cfg::lds, s, a);
  return launched();
}

template <int M>
int launch_field(const vp::FieldArgs& a, hipStream_t s, int n_inst) {
  constexpr int Q = (M / kFT) > 8 ? (M / kFT) : (M < 8 ? M : 8);
  using Cfg = vp::FieldCfg<M, Q, kFT>;
  auto kern = k_field<M, Q>;
  if (int rc = set_lds(kern, Cfg::lds)) return rc;
  vp::FieldArgs b = a;
  int xm_blocks = 0;
  if (b.xm_out) {
    if (n_inst > 0) return fail(VPFP_EBADARG, "x-modes ride only on a single-grid field launch");
    b.xm_blk0 = field_grid(a.nx);
    xm_blocks = (b.xm_n * b.xm_nv + kFT - 1) / kFT;
  }
  const dim3 grid = n_inst > 0 ? dim3(1, n_inst) : dim3(field_grid(a.nx) + xm_blocks);
  hipLaunchKernelGGL(kern, grid, dim3(kFT), Cfg::lds, s, b);
  return launched();
}

template <class T>
int row_dispatch(const vp::RowArgs<T>& a, hipStream_t s, int q = 0) {
  return a.bg_hat ? row_dispatch_bg(a, s, q) : row_dispatch_plain(a, s, q);
}

template <class T>
int tri_dispatch(const vp::TriArgs<T>& a, hipStream_t s) {
  const int M = a.nv / 2;
#define CALL(m) launch_tri<T, m>(a, s)
  VPFP_SWITCH_M(M, CALL, 4096)
#undef CALL
  return fail(VPFP_EUNSUPPORTED, "nv must be a power of two in [4, 8192]");
}

template <class T>
int drift_dispatch(const vp::DriftArgs<T>& a, int C, int gx, hipStream_t s, int gy = 0) {
  return vpfp_impl::drift_dispatch(a, C, gx, s, gy);
}

template <int M>
int launch_field_multi(vp::FieldMulti m, hipStream_t s) {
  constexpr int Q = (M / kFT) > 8 ? (M / kFT) : (M < 8 ? M : 8);
  using Cfg = vp::FieldCfg<M, Q, kFT>;
  auto kern = k_field_multi<M, Q>;
  if (int rc = set_lds(kern, Cfg::lds)) return rc;
  long long blocks = 0;
  for (int i = 0; i < m.nseg; ++i) {
    m.seg[i].blk0 = (int)blocks;
    blocks += m.seg[i].n_inst;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kFT), Cfg::lds, s, m);
  return launched();
}

int field_multi_dispatch(const vp::FieldMulti& m, hipStream_t s) {
  const int M = m.base.nx / 2;
#define CALL(mm) launch_field_multi<mm>(m, s)
  VPFP_SWITCH_M(M, CALL, 2048)
#undef CALL
  return fail(VPFP_EUNSUPPORTED, "grouped field solve needs nx a power of two in [4, 4096]");
}

template <class T>
int row_multi_dispatch(const vp::RowMulti<T>& m, hipStream_t s, int q) {
  return m.base.bg_hat ? row_multi_dispatch_bg(m, s, q) : row_multi_dispatch_plain(m, s, q);
}

// n_inst > 0: a batch of n_inst instance-major field solves (pow2 only)
int field_dispatch(const vp::FieldArgs& a, bool pow2, hipStream_t s, int n_inst = 0) {
  if (pow2) {
    const int M = a.nx / 2;
#define CALL(m) launch_field<m>(a, s, n_inst)
    VPFP_SWITCH_M(M, CALL, 2048)
#undef CALL
  }
  if (n_inst > 0) return fail(VPFP_EUNSUPPORTED, "batched field solve needs nx a power of two");
  if (a.nx > 4096) return fail(VPFP_EUNSUPPORTED, "nx must be at most 4096 for the field solve");
  const int lds = ((a.nx * 8 + 15) & ~15) + (a.nx > kFT ? a.nx : kFT) * 16 + 64;
  if (int rc = set_lds(k_field_dft, lds)) return rc;
  hipLaunchKernelGGL(k_field_dft, dim3(1), dim3(kFT), lds, s, a);
  return launched();
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// fftfreq table check: k[0] == 0, k[N-m] == -k[m], and k[m] == m k[1] to
// rounding for m <= N/2 (k[N/2] negative): the kernels form the phase
// k[m] * dt * s as m * (k[1] * dt * s)
bool fftfreq_like(const double* k, int n) {
  if (!is_pow2(n) || n < 4) return true;  // not used by the spectral kernels
  const int M = n / 2;
  if (k[0] != 0.0) return false;
  const double tol = 64.0 * 2.220446049250313e-16 * std::fabs(M * k[1]);
  if (std::fabs(k[M] + M * k[1]) > tol) return false;
  for (int m = 1; m < M; ++m) {
    if (k[n - m] != -k[m]) return false;
    if (std::fabs(k[m] - m * k[1]) > tol) return false;
  }
  return true;
}

// v_m = v0 + m dvv to rounding (the collision kernels form v_m that way)
bool uniform_grid(const double* v, int n, double* v0, double* dvv) {
  *v0 = v[0];
  *dvv = n > 1 ? (v[n - 1] - v[0]) / (n - 1) : 0.0;
  double vmax = 0.0;
  for (int m = 0; m < n; ++m) vmax = std::fmax(vmax, std::fabs(v[m]));
  for (int m = 0; m < n; ++m)
    if (std::fabs(*v0 + m * *dvv - v[m]) > 16.0 * 2.220446049250313e-16 * vmax) return false;
  return true;
}

}  // namespace

// Per-kernel timing (vpfp_plan_set_profiling, vpfp_batch_set_profiling):
// with `on`, every launch of a step is bracketed by two HIP events on its
// stream; collect() waits for the last one and adds the elapsed times per
// kernel class: [0] row, [1] drift, [2] field, [3] other.
// A launch is marked with an id = kind + 4 * index (index: which launch of
// that kind in a grouped step); ms / n sum over ids by kind, ms_id / n_id
// keep each id.
struct Prof {
  bool on = false;
  double ms[4] = {0, 0, 0, 0};
  long long n[4] = {0, 0, 0, 0};
  std::vector<double> ms_id;
  std::vector<long long> n_id;
  std::vector<hipEvent_t> ev;  // pool, 2 per timed launch
  std::vector<int> kind;
  size_t used = 0;
  void reset(bool enable) {
    on = enable;
    for (int k = 0; k < 4; ++k) {
      ms[k] = 0.0;
      n[k] = 0;
    }
    ms_id.clear();
    n_id.clear();
    used = 0;
  }
  ~Prof() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

// Streams an object enqueued work on.  Destroying the object waits for
// them (its buffers must outlive that work); after more than kMax distinct
// streams it waits for the whole device instead.  The streams must still
// exist when the object is destroyed.
struct StreamLog {
  static constexpr int kMax = 8;
  hipStream_t s[kMax] = {};
  int n = 0;
  bool overflow = false;
  hipStream_t use(void* st) {
    hipStream_t h = reinterpret_cast<hipStream_t>(st);
    for (int i = 0; i < n; ++i)
      if (s[i] == h) return h;
    if (n < kMax)
      s[n++] = h;
    else
      overflow = true;
    return h;
  }
  void drain() {
    if (overflow) {
      (void)hipDeviceSynchronize();
      return;
    }
    for (int i = 0; i < n; ++i) (void)hipStreamSynchronize(s[i]);
  }
};

struct vpfp_plan_s {
  int nx = 0, nv = 0, prec = VPFP_F64, op = VPFP_OP_LB;
  double dv = 0, nu = 0, dt = 0;
  double dk_x = 0, dk_v = 0, v0 = 0, dvv = 0;
  bool nx_pow2 = false, nv_pow2 = false;
  int C = 0, ntiles = 0;
  int row_q = 0;   // values per thread of the row kernels (0: row-count rule)
  int fuse = -1;   // field solved by the last drift workgroup: 1 yes, 0 no, -1 size rule
  int xcd = 1;     // drift tile order (xcd_block); never slower in the measured shapes
  int dtiles = 1;  // column tiles per drift workgroup
  int dft = 0;                // x-advection as real-DFT products on MFMA (FP32, nx 32 or 64)
  float* dft_mat = nullptr;   // [2][nx][nx] forward and inverse real-DFT matrices
  double tuned_us = 0.0;
  int delta = 0;                 // f stored as f - M(v)
  double* bg = nullptr;          // M(v_m) [nv] (FP64) and its DFT [nv] (storage precision)
  void* bg_hat = nullptr;
  double *kx = nullptr, *okx = nullptr, *v = nullptr, *kv = nullptr;
  void* tw_v = nullptr;          // cx<T> [lo | hi] tables for nv
  void* tw_x = nullptr;          // cx<T> [lo | hi] tables for nx
  cx<double>* tw_xs = nullptr;   // cx<double> [lo | hi] tables for nx (field)
  cx<double>* tw_xd = nullptr;   // cx<double>[nx] full table (direct-DFT field)
  void* fcopy = nullptr;         // [nx][nv] storage type: f after a step (vpfp_step_n_store's x-modes)
  double* dens = nullptr;        // [max(ntiles,1)][nx]
  double* ering = nullptr;       // [2][nx]
  double* rho_g = nullptr;       // [nx] field reduction scratch
  unsigned int* counter = nullptr;
  double* series = nullptr;      // [2]
  Prof prof;  // per-kernel event timing (vpfp_plan_set_profiling)
  StreamLog streams;
  std::vector<void*> allocs;
  ~vpfp_plan_s() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

namespace {

template <class P>
int dalloc(vpfp_plan_s* p, P** out, size_t bytes) {
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, bytes > 0 ? bytes : 16);
  if (e != hipSuccess) return fail(VPFP_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  p->allocs.push_back(ptr);
  *out = reinterpret_cast<P*>(ptr);
  return VPFP_OK;
}

int upload(void* dst, const void* src, size_t bytes) {
  VPFP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return VPFP_OK;
}

// exp(-2 pi i n / N) for n < N in long double, rounded
template <class T>
std::vector<cx<T>> twiddles(int N) {
  std::vector<cx<T>> w(N);
  const long double two_pi = 6.283185307179586476925286766559L;
  for (int n = 0; n < N; ++n) {
    const long double a = two_pi * (long double)n / (long double)N;
    w[n] = vp::mk<T>((T)(double)cosl(a), (T)(double)(-sinl(a)));
  }
  return w;
}

// Twiddle tables as the kernels stage them: the full W_N^i (i < N) when
// N <= vp::kTwFull, else [lo | hi]: lo[j] = W_N^j (j < 64), hi[h] = W_N^{64 h}
template <class T>
std::vector<cx<T>> split_twiddles(int N) {
  if (vp::tw_full(N, (int)sizeof(cx<T>))) return twiddles<T>(N);
  const int nhi = N / 64 > 1 ? N / 64 : 1;
  std::vector<cx<T>> w(64 + nhi);
  const long double two_pi = 6.283185307179586476925286766559L;
  for (int j = 0; j < 64 + nhi; ++j) {
    const long double n = j < 64 ? (long double)j : (long double)(64 * (j - 64));
    const long double a = two_pi * n / (long double)N;
    w[j] = vp::mk<T>((T)(double)cosl(a), (T)(double)(-sinl(a)));
  }
  return w;
}

// Real-DFT matrices of the MFMA drift, [2][N][N] FP32 (rounded from long
// double): forward F[k][x] = cos(2 pi k x / N) for k <= N/2 and
// F[N/2 + j][x] = -sin(2 pi j x / N) for 0 < j < N/2 (rows Re F_k, Im F_k);
// inverse G[x][k] = c_k cos(2 pi k x / N) / N for k <= N/2 (c = 1 at 0 and
// N/2, else 2) and G[x][N/2 + j] = -2 sin(2 pi j x / N) / N, the real
// synthesis from those rows.
std::vector<float> dft_matrices(int N) {
  std::vector<float> m(2 * (size_t)N * N);
  const long double tp = 6.283185307179586476925286766559L;
  auto ang = [&](long long k, int x) { return tp * (long double)((k * x) % N) / (long double)N; };
  for (int row = 0; row < N; ++row)
    for (int x = 0; x < N; ++x)
      m[(size_t)row * N + x] = (float)(row <= N / 2 ? cosl(ang(row, x)) : -sinl(ang(row - N / 2, x)));
  float* G = m.data() + (size_t)N * N;
  for (int x = 0; x < N; ++x)
    for (int col = 0; col < N; ++col) {
      long double v;
      if (col <= N / 2)
        v = ((col == 0 || col == N / 2) ? 1.0L : 2.0L) * cosl(ang(col, x)) / N;
      else
        v = -2.0L * sinl(ang(col - N / 2, x)) / N;
      G[(size_t)x * N + col] = (float)v;
    }
  return m;
}

bool dft_size(int prec, int nx) { return prec == VPFP_F32 && (nx == 32 || nx == 64); }

template <class T, class P>
int put(vpfp_plan_s* p, P** dst, const std::vector<T>& src) {
  if (int rc = dalloc(p, dst, sizeof(T) * src.size())) return rc;
  return upload(*dst, src.data(), sizeof(T) * src.size());
}

template <class T>
int make_tables(vpfp_plan_s* p) {
  cx<T>* dv_ = nullptr;
  cx<T>* dx_ = nullptr;
  if (int rc = put(p, &dv_, split_twiddles<T>(p->nv))) return rc;
  if (int rc = put(p, &dx_, split_twiddles<T>(p->nx))) return rc;
  if (int rc = put(p, &p->tw_xs, split_twiddles<double>(p->nx))) return rc;
  if (int rc = put(p, &p->tw_xd, twiddles<double>(p->nx))) return rc;
  p->tw_v = dv_;
  p->tw_x = dx_;
  return VPFP_OK;
}

int drift_tile(int nx, int nv) {
  const int M = nx / 2;
  int C = M <= 256 ? 16 : (M == 512 ? 8 : 4);
  while (C > nv) C /= 2;
  return C >= 4 ? C : 0;
}

// Field solve of the fused step: as a tail of the drift (its last workgroup)
// or as a launch of its own.  The tail costs every drift workgroup a
// device-scope release fence, which on this multi-XCD part writes back and
// invalidates the L2 (buffer_wbl2 / buffer_inv); measured at nx = 64,
// nv = 1024 the separate launch is faster (23.9 vs 26.0 us per step), so the
// tail is left to the autotuner.
bool default_fuse(int) { return false; }

enum KernelKind { KK_ROW = 0, KK_DRIFT = 1, KK_FIELD = 2, KK_OTHER = 3 };

// In profiling mode every launch of the fused step is bracketed by two
// events; vpfp_step_n reads them back once its stream has drained.
int prof_mark(Prof& pr, int kind, hipStream_t s) {
  if (!pr.on) return VPFP_OK;
  if (pr.used + 1 >= pr.ev.size()) {
    const size_t grow = pr.ev.size() + 256;
    while (pr.ev.size() < grow) {
      hipEvent_t e;
      VPFP_HIP(hipEventCreate(&e));
      pr.ev.push_back(e);
    }
    pr.kind.resize(pr.ev.size() / 2 + 1);
  }
  if ((pr.used & 1) == 0) pr.kind[pr.used / 2] = kind;
  VPFP_HIP(hipEventRecord(pr.ev[pr.used], s));
  ++pr.used;
  return VPFP_OK;
}

int prof_collect(Prof& pr) {
  if (!pr.on || pr.used == 0) return VPFP_OK;
  VPFP_HIP(hipEventSynchronize(pr.ev[pr.used - 1]));
  for (size_t i = 0; i + 1 < pr.used; i += 2) {
    float ms = 0.f;
    VPFP_HIP(hipEventElapsedTime(&ms, pr.ev[i], pr.ev[i + 1]));
    const int id = pr.kind[i / 2];
    pr.ms[id % 4] += ms;
    pr.n[id % 4] += 1;
    if ((size_t)id >= pr.ms_id.size()) {
      pr.ms_id.resize((size_t)id + 1, 0.0);
      pr.n_id.resize((size_t)id + 1, 0);
    }
    pr.ms_id[id] += ms;
    pr.n_id[id] += 1;
  }
  pr.used = 0;
  return VPFP_OK;
}

#define PROF(kind, call)                                  \
  do {                                                    \
    if (int rc_ = prof_mark(p->prof, kind, s)) return rc_; \
    if (int rc_ = (call)) return rc_;                      \
    if (int rc_ = prof_mark(p->prof, kind, s)) return rc_; \
  } while (0)

template <class T>
vp::RowArgs<T> row_args(vpfp_plan_t p, const void* fin, void* fout) {
  vp::RowArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.f_in = reinterpret_cast<const T*>(fin);
  a.f_out = reinterpret_cast<T*>(fout);
  a.nx = p->nx;
  a.nv = p->nv;
  a.dk = p->dk_v;
  a.v0 = p->v0;
  a.dvv = p->dvv;
  a.dv = p->dv;
  a.tw = reinterpret_cast<const cx<T>*>(p->tw_v);
  a.nu_dt = p->nu * p->dt;
  a.op_dg = p->op == VPFP_OP_DG;
  if (p->delta) {
    a.bg = p->bg;
    a.bg_hat = reinterpret_cast<const cx<T>*>(p->bg_hat);
  }
  return a;
}

template <class T>
vp::DriftArgs<T> drift_args(vpfp_plan_t p, const void* fin, void* fout, double dt, double* dens) {
  vp::DriftArgs<T> a;
  a.f_in = reinterpret_cast<const T*>(fin);
  a.f_out = reinterpret_cast<T*>(fout);
  a.nx = p->nx;
  a.nv = p->nv;
  a.dt = dt;
  a.dk = p->dk_x;
  a.v = p->v;
  a.tw = reinterpret_cast<const cx<T>*>(p->tw_x);
  a.dens = dens;
  a.fuse_field = 0;
  std::memset(&a.fld, 0, sizeof(a.fld));
  a.dk_inst = nullptr;
  a.xcd_order = p->xcd;
  a.dft = p->dft ? p->dft_mat : nullptr;
  return a;
}

vp::FieldArgs field_args(vpfp_plan_t p, const double* part, int ntiles, double scale, const double* driver,
                         double* e_out, double* e_out2, double* series, bool pow2) {
  vp::FieldArgs a;
  std::memset(&a, 0, sizeof(a));
  a.part = part;
  a.ntiles = ntiles;
  a.rho_g = p->rho_g;
  a.counter = p->counter;
  a.nx = p->nx;
  a.scale = scale;
  a.okx = p->okx;
  a.tw = pow2 ? p->tw_xs : p->tw_xd;
  a.driver = driver;
  a.e_out = e_out;
  a.e_out2 = e_out2;
  a.series = series;
  a.series_stride = 0;
  return a;
}

int check_plan(vpfp_plan_t p) {
  if (!p) return fail(VPFP_EBADARG, "null plan");
  return VPFP_OK;
}

int need_spectral_v(vpfp_plan_t p) {
  if (!p->nv_pow2 || p->nv < 4 || p->nv > 8192)
    return fail(VPFP_EUNSUPPORTED, "the v-direction kernels need nv a power of two in [4, 8192], got " +
                                       std::to_string(p->nv));
  return VPFP_OK;
}

int need_spectral_x(vpfp_plan_t p) {
  if (!p->nx_pow2 || p->nx < 4 || p->nx > 2048 || p->C == 0)
    return fail(VPFP_EUNSUPPORTED, "the x-advection needs nx a power of two in [4, 2048] and nv >= 4, got nx=" +
                                       std::to_string(p->nx));
  return VPFP_OK;
}

template <class T>
int do_edfdv(vpfp_plan_t p, const void* fin, void* fout, const double* e, double dt, hipStream_t s) {
  auto a = row_args<T>(p, fin, fout);
  a.e = e;
  a.dt_a = dt;
  a.flags = vp::RK_KICK_A;
  return row_dispatch<T>(a, s, p->row_q);
}

template <class T>
int do_vdfdx(vpfp_plan_t p, const void* fin, void* fout, double dt, double* dens, hipStream_t s) {
  return drift_dispatch<T>(drift_args<T>(p, fin, fout, dt, dens), p->C, p->ntiles, s);
}

template <class T>
int do_density(vpfp_plan_t p, const void* f, double scale, double* part, hipStream_t s) {
  const int rows = 4;
  hipLaunchKernelGGL(k_density<T>, dim3((p->nx + rows - 1) / rows), dim3(64 * rows), 0, s,
                     reinterpret_cast<const T*>(f), p->nx, p->nv, scale, part);
  return launched();
}

template <class T>
int do_collide(vpfp_plan_t p, const void* fin, void* fout, hipStream_t s) {
  auto a = row_args<T>(p, fin, fout);
  a.flags = vp::RK_COLLIDE;
  return row_dispatch<T>(a, s, p->row_q);
}

template <class T>
int do_diag(vpfp_plan_t p, const void* f, double* A, double* B, double* Cc, hipStream_t s) {
  const int rows = 4;
  hipLaunchKernelGGL(k_diag<T>, dim3((p->nx + rows - 1) / rows), dim3(64 * rows), 0, s,
                     reinterpret_cast<const T*>(f), p->nx, p->nv, p->v, p->dv, p->nu * p->dt,
                     (int)(p->op == VPFP_OP_DG), A, B, Cc);
  return launched();
}

template <class T>
int do_tri(vpfp_plan_t p, const double* A, const double* B, const double* Cc, const void* d, void* x,
           hipStream_t s) {
  vp::TriArgs<T> a;
  a.a = A;
  a.b = B;
  a.cc = Cc;
  a.d = reinterpret_cast<const T*>(d);
  a.x = reinterpret_cast<T*>(x);
  a.nx = p->nx;
  a.nv = p->nv;
  return tri_dispatch<T>(a, s);
}

// f_store (optional): per-step f store, written by the row kernel between
// the collision and the next half kick; n_modes == 0: [n][nx][nv] copies of
// f (type T), n_modes > 0: [n][n_modes][nv] complex128 x-modes of f, formed
// from the row kernel's copy in p->fcopy by extra workgroups of the next
// step's field launch (one k_xmodes launch after the last step, and per
// step when the field is not a launch of its own)
template <class T>
int do_step_n(vpfp_plan_t p, void* f, double* e, int n, const double* driver, double* e_hist, double* moments,
              double* series, hipStream_t s, void* f_store = nullptr, int n_modes = 0) {
  const bool coll = p->nu > 0.0;
  const double half = 0.5 * p->dt;
  auto ra = row_args<T>(p, f, f);
  ra.e = e;
  ra.dt_a = half;
  ra.flags = vp::RK_KICK_A;
  if (n > 0) PROF(KK_ROW, row_dispatch<T>(ra, s, p->row_q));
  const double* e_cur = e;
  // field: 0 own launch, 1 tail of the drift, 2 prologue of the row kernel
  const int mode = p->fuse >= 0 ? p->fuse : (default_fuse(p->nx) ? 1 : 0);
  if (mode == 2 && p->nx > vp::kProNx)
    return fail(VPFP_EUNSUPPORTED, "field in the row kernel needs nx <= " + std::to_string(vp::kProNx));
  for (int i = 0; i < n; ++i) {
    double* ei = e_hist ? e_hist + (size_t)i * p->nx : p->ering + (size_t)(i & 1) * p->nx;
    const double* drv = driver ? driver + (size_t)i * p->nx : nullptr;
    double* ser = series ? series + 2 * (size_t)i : nullptr;
    auto fa = field_args(p, p->dens, p->ntiles, p->dv, drv, ei, nullptr, ser, true);
    auto* modes = reinterpret_cast<cx<double>*>(f_store);
    if (f_store && n_modes > 0 && mode == 0 && i > 0) {
      fa.xm_f = p->fcopy;
      fa.xm_f32 = sizeof(T) == 4;
      fa.xm_n = n_modes;
      fa.xm_nv = p->nv;
      fa.xm_tw = p->tw_xd;
      fa.xm_out = modes + (size_t)(i - 1) * n_modes * p->nv;
    }
    if (mode == 1) {
      auto da = drift_args<T>(p, f, f, p->dt, p->dens);
      da.fuse_field = 1;
      da.fld = fa;
      PROF(KK_DRIFT, drift_dispatch<T>(da, p->C, p->ntiles, s));
    } else {
      PROF(KK_DRIFT, (do_vdfdx<T>(p, f, f, p->dt, p->dens, s)));
      if (mode == 0) PROF(KK_FIELD, field_dispatch(fa, true, s));
    }
    e_cur = ei;
    auto rb = row_args<T>(p, f, f);
    rb.e = e_cur;
    if (mode == 2) {
      rb.field_pro = 1;
      rb.fld = field_args(p, p->dens, p->ntiles, p->dv, drv, ei, nullptr, ser, false);
    }
    rb.dt_a = half;
    rb.dt_b = half;
    rb.moments = moments ? moments + (size_t)i * vp::kMoments * p->nx : nullptr;
    rb.flags = vp::RK_KICK_A | (coll ? vp::RK_COLLIDE : 0) | (moments ? vp::RK_MOMENTS : 0) |
               (i + 1 < n ? vp::RK_KICK_B : 0);
    if (f_store)
      rb.f_copy = n_modes > 0 ? reinterpret_cast<T*>(p->fcopy)
                              : reinterpret_cast<T*>(f_store) + (size_t)i * p->nx * p->nv;
    PROF(KK_ROW, row_dispatch<T>(rb, s, p->row_q));
    if (f_store && n_modes > 0 && (mode != 0 || i + 1 == n)) {
      const int total = n_modes * p->nv;
      if (int rc = prof_mark(p->prof, KK_OTHER, s)) return rc;
      hipLaunchKernelGGL(k_xmodes<T>, dim3((total + 255) / 256), dim3(256), 0, s,
                         reinterpret_cast<const T*>(p->fcopy), p->nx, p->nv, n_modes, p->tw_xd,
                         modes + (size_t)i * n_modes * p->nv);
      if (int rc = launched()) return rc;
      if (int rc = prof_mark(p->prof, KK_OTHER, s)) return rc;
    }
  }
  if (n > 0) VPFP_HIP(hipMemcpyAsync(e, e_cur, sizeof(double) * p->nx, hipMemcpyDeviceToDevice, s));
  return prof_collect(p->prof);
}

constexpr int kSchemeFields = VPFP_SCHEME_FIELDS;  // field solves per PEFRL or 6th-order step

// One step of a higher-order splitting as a list of stages: a kick by c dt
// with the latest field, or a drift by c dt followed by a field solve.
struct Stage {
  bool kick;
  double c;
};

// PEFRL (vlasov_poisson.py:99-154, Omelyan, Mryglod & Folk 2002) and the
// 6th-order splitting of Casas et al. 2017 (vlasov_poisson.py:191-232):
// published method constants
std::vector<Stage> scheme_stages(int scheme, double dt) {
  if (scheme == VPFP_SCHEME_PEFRL) {
    const double xsi = 0.1786178958448091, lam = -0.2123418310626054, chi = -0.6626458266981849e-1;
    const double kw0 = 0.5 * (1.0 - 2.0 * lam);
    return {{false, xsi}, {true, kw0}, {false, chi}, {true, lam}, {false, 1.0 - 2.0 * (chi + xsi)},
            {true, lam},  {false, chi}, {true, kw0}, {false, xsi}};
  }
  const double a1 = 0.168735950563437422448196, a2 = 0.377851589220928303880766,
               a3 = -0.093175079568731452657924, b1 = 0.049086460976116245491441,
               b2 = 0.264177609888976700200146, b3 = 0.186735929134907054308413,
               c1 = -0.000069728715055305084099, c2 = -0.000625704827430047189169,
               c3 = -0.002213085124045325561636, d2 = -2.916600457689847816445691e-6,
               d3 = 3.048480261700038788680723e-5, e3 = 4.985549387875068121593988e-7;
  const double k1 = b1 + 2.0 * c1 * std::pow(dt, 2.0);
  const double k2 = b2 + 2.0 * c2 * std::pow(dt, 2.0) + 4.0 * d2 * std::pow(dt, 4.0);
  const double k3 = b3 + 2.0 * c3 * std::pow(dt, 2.0) + 4.0 * d3 * std::pow(dt, 4.0) - 8.0 * e3 * std::pow(dt, 6.0);
  return {{true, k1},  {false, a1}, {true, k2}, {false, a2}, {true, k3},  {false, a3},
          {true, k3},  {false, a2}, {true, k2}, {false, a1}, {true, k1}};
}

// n steps of a scheme, each followed by the collision step, on the device:
// drifts with their field solve, kicks through the row kernel; the last
// kick of a step, the collision, the moments, the f store and the next
// step's first kick share one row launch.  driver [n][kSchemeFields][nx].
template <class T>
int do_step_scheme(vpfp_plan_t p, int scheme, void* f, double* e, int n, const double* driver, double* e_hist,
                   double* moments, double* series, hipStream_t s, void* f_store, int n_modes) {
  const std::vector<Stage> st = scheme_stages(scheme, p->dt);
  const int ns = (int)st.size();
  const bool coll = p->nu > 0.0;
  const bool first_kick = st[0].kick, last_kick = st[ns - 1].kick;
  auto* modes = reinterpret_cast<cx<double>*>(f_store);
  const double* e_cur = e;
  long long field_no = 0;
  if (n > 0 && first_kick) {
    auto ra = row_args<T>(p, f, f);
    ra.e = e;
    ra.dt_a = st[0].c * p->dt;
    ra.flags = vp::RK_KICK_A;
    PROF(KK_ROW, row_dispatch<T>(ra, s, p->row_q));
  }
  for (int i = 0; i < n; ++i) {
    int j = 0;  // field solve of this step
    for (int k = 0; k < ns; ++k) {
      const Stage& g = st[k];
      if (g.kick) {
        if (k == 0 || k == ns - 1) continue;  // fused into the row launch that ends a step
        auto rk = row_args<T>(p, f, f);
        rk.e = e_cur;
        rk.dt_a = g.c * p->dt;
        rk.flags = vp::RK_KICK_A;
        PROF(KK_ROW, row_dispatch<T>(rk, s, p->row_q));
        continue;
      }
      PROF(KK_DRIFT, (do_vdfdx<T>(p, f, f, g.c * p->dt, p->dens, s)));
      const bool last = j == kSchemeFields - 1;
      double* eo = (last && e_hist) ? e_hist + (size_t)i * p->nx : p->ering + (size_t)(field_no & 1) * p->nx;
      const double* drv = driver ? driver + ((size_t)i * kSchemeFields + j) * p->nx : nullptr;
      double* ser = (last && series) ? series + 2 * (size_t)i : nullptr;
      auto fa = field_args(p, p->dens, p->ntiles, p->dv, drv, eo, nullptr, ser, true);
      PROF(KK_FIELD, field_dispatch(fa, true, s));
      e_cur = eo;
      ++j;
      ++field_no;
    }
    auto rb = row_args<T>(p, f, f);
    rb.e = e_cur;
    rb.dt_a = last_kick ? st[ns - 1].c * p->dt : 0.0;
    rb.dt_b = st[0].c * p->dt;
    rb.moments = moments ? moments + (size_t)i * vp::kMoments * p->nx : nullptr;
    rb.flags = (last_kick ? vp::RK_KICK_A : 0) | (coll ? vp::RK_COLLIDE : 0) | (moments ? vp::RK_MOMENTS : 0) |
               (first_kick && i + 1 < n ? vp::RK_KICK_B : 0);
    if (f_store)
      rb.f_copy = n_modes > 0 ? reinterpret_cast<T*>(p->fcopy)
                              : reinterpret_cast<T*>(f_store) + (size_t)i * p->nx * p->nv;
    if (rb.flags || rb.f_copy) PROF(KK_ROW, row_dispatch<T>(rb, s, p->row_q));
    if (f_store && n_modes > 0) {
      const int total = n_modes * p->nv;
      if (int rc = prof_mark(p->prof, KK_OTHER, s)) return rc;
      hipLaunchKernelGGL(k_xmodes<T>, dim3((total + 255) / 256), dim3(256), 0, s,
                         reinterpret_cast<const T*>(p->fcopy), p->nx, p->nv, n_modes, p->tw_xd,
                         modes + (size_t)i * n_modes * p->nv);
      if (int rc = launched()) return rc;
      if (int rc = prof_mark(p->prof, KK_OTHER, s)) return rc;
    }
  }
  if (n > 0) VPFP_HIP(hipMemcpyAsync(e, e_cur, sizeof(double) * p->nx, hipMemcpyDeviceToDevice, s));
  return prof_collect(p->prof);
}

}  // namespace

// ====================================================================== batched scan
struct vpfp_batch_s {
  int n = 0, nx = 0, nv = 0, prec = VPFP_F64, op = VPFP_OP_DG, C = 0, gx = 0;
  int row_q = 0;
  int xcd = 1;
  int dtiles = 1;
  int dft = 0;
  float* dft_mat = nullptr;
  double tuned_us = 0.0;
  vpfc_net_t closure = nullptr;  // learned closure after every step (FP32, nv 256)
  int groups = 0;                // groups holding this batch (vpfp_group_create)
  int delta = 0;                 // f stored as f - M(v)
  double* bg = nullptr;          // M(v_m) [nv] and its unnormalised DFT [nv]
  void* bg_hat = nullptr;        // (cx<double> or cx<float>: the storage precision)
  double dv = 0, dt = 0, v0 = 0, dvv = 0, dk_v = 0;
  double *dk_x = nullptr, *okx = nullptr, *nu_dt = nullptr, *v = nullptr;
  void *tw_v = nullptr, *tw_x = nullptr;
  cx<double>* tw_xs = nullptr;
  double* dens = nullptr;
  unsigned int* counters = nullptr;
  double* ering = nullptr;
  Prof prof;  // per-kernel event timing (vpfp_batch_set_profiling)
  StreamLog streams;
  std::vector<void*> allocs;
  ~vpfp_batch_s() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

namespace {

template <class P>
int balloc(vpfp_batch_s* b, P** out, size_t bytes) {
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, bytes > 0 ? bytes : 16);
  if (e != hipSuccess) return fail(VPFP_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  b->allocs.push_back(ptr);
  *out = reinterpret_cast<P*>(ptr);
  return VPFP_OK;
}

template <class T, class P>
int bput(vpfp_batch_s* b, P** dst, const std::vector<T>& src) {
  if (int rc = balloc(b, dst, sizeof(T) * src.size())) return rc;
  return upload(*dst, src.data(), sizeof(T) * src.size());
}

template <class T>
int batch_tables(vpfp_batch_s* b) {
  cx<T>* tv = nullptr;
  cx<T>* tx = nullptr;
  if (int rc = bput(b, &tv, split_twiddles<T>(b->nv))) return rc;
  if (int rc = bput(b, &tx, split_twiddles<T>(b->nx))) return rc;
  if (int rc = bput(b, &b->tw_xs, split_twiddles<double>(b->nx))) return rc;
  b->tw_v = tv;
  b->tw_x = tx;
  return VPFP_OK;
}

template <class T>
int batch_step_n(vpfp_batch_s* b, void* f, double* e, int n, double* series, hipStream_t s) {
  vp::RowArgs<T> ra;
  std::memset(&ra, 0, sizeof(ra));
  ra.f_in = reinterpret_cast<const T*>(f);
  ra.f_out = reinterpret_cast<T*>(f);
  ra.nx = b->n * b->nx;
  ra.nv = b->nv;
  ra.dk = b->dk_v;
  ra.v0 = b->v0;
  ra.dvv = b->dvv;
  ra.dv = b->dv;
  ra.tw = reinterpret_cast<const cx<T>*>(b->tw_v);
  ra.op_dg = b->op == VPFP_OP_DG;
  ra.nu_dt_inst = b->nu_dt;
  ra.rows_per_inst = b->nx;
  if (b->delta) {
    ra.bg = b->bg;
    ra.bg_hat = reinterpret_cast<const cx<T>*>(b->bg_hat);
  }
  ra.dt_a = ra.dt_b = 0.5 * b->dt;
  ra.e = e;
  ra.flags = vp::RK_KICK_A;
  Prof& pr = b->prof;
#define BPROF(kind, call)                                 \
  do {                                                    \
    if (int rc_ = prof_mark(pr, kind, s)) return rc_;     \
    if (int rc_ = (call)) return rc_;                     \
    if (int rc_ = prof_mark(pr, kind, s)) return rc_;     \
  } while (0)
  if (n > 0) BPROF(KK_ROW, row_dispatch<T>(ra, s, b->row_q));
  const double* e_cur = e;
  const bool closure = b->closure != nullptr;
  for (int i = 0; i < n; ++i) {
    if (closure && i > 0) {
      // first half kick of this step (not fused: the closure ran in between)
      ra.e = e_cur;
      ra.flags = vp::RK_KICK_A;
      BPROF(KK_ROW, row_dispatch<T>(ra, s, b->row_q));
    }
    double* ei = b->ering + (size_t)(i & 1) * b->n * b->nx;
    vp::DriftArgs<T> da;
    std::memset(&da, 0, sizeof(da));
    da.f_in = reinterpret_cast<const T*>(f);
    da.f_out = reinterpret_cast<T*>(f);
    da.nx = b->nx;
    da.nv = b->nv;
    da.dt = b->dt;
    da.v = b->v;
    da.tw = reinterpret_cast<const cx<T>*>(b->tw_x);
    da.dens = b->dens;
    da.dk_inst = b->dk_x;
    da.xcd_order = b->xcd;
    da.dft = b->dft ? b->dft_mat : nullptr;
    // the field is a launch of its own: a last-workgroup hand-off inside the
    // drift would need a device-scope fence (an L2 write-back and invalidate)
    // in each of the n * gx drift workgroups
    vp::FieldArgs fa;
    std::memset(&fa, 0, sizeof(fa));
    fa.part = b->dens;
    fa.ntiles = b->gx;
    fa.nx = b->nx;
    fa.scale = b->dv;
    fa.counter = b->counters;
    fa.okx = b->okx;
    fa.tw = b->tw_xs;
    fa.e_out = ei;
    fa.series = series ? series + (size_t)i * b->n * 2 : nullptr;
    fa.series_stride = 2;
    BPROF(KK_DRIFT, drift_dispatch<T>(da, b->C, b->gx, s, b->n));
    BPROF(KK_FIELD, field_dispatch(fa, true, s, b->n));
    e_cur = ei;
    ra.e = e_cur;
    ra.flags = vp::RK_KICK_A | vp::RK_COLLIDE | (i + 1 < n && !closure ? vp::RK_KICK_B : 0);
    BPROF(KK_ROW, row_dispatch<T>(ra, s, b->row_q));
    if (closure) {
      if (int rc = prof_mark(pr, KK_OTHER, s)) return rc;
      if (int rc = vpfc_apply(b->closure, reinterpret_cast<float*>(f), b->n * b->nx, s)) {
        return fail(rc, std::string("closure: ") + vpfc_last_error());
      }
      if (int rc = prof_mark(pr, KK_OTHER, s)) return rc;
    }
  }
#undef BPROF
  if (n > 0) VPFP_HIP(hipMemcpyAsync(e, e_cur, sizeof(double) * b->n * b->nx, hipMemcpyDeviceToDevice, s));
  return prof_collect(pr);
}

// ---------------------------------------------------------------- autotuner
// Times every valid launch shape of the step on scratch state and keeps the
// fastest: values per thread of the row kernels (4 or 8, where both exist),
// columns per drift workgroup (4, 8, 16) and, for a single grid with
// nx <= 256, the field solve as a drift tail or as its own launch.

// (1 + 0.01 cos(2 pi x / nx)) exp(-v^2 / 2) / sqrt(2 pi) on every row
template <class T>
std::vector<T> tune_state(size_t rows, int nx, int nv, double v0, double dvv) {
  std::vector<T> h(rows * (size_t)nv);
  const double two_pi = 6.283185307179586;
  for (size_t r = 0; r < rows; ++r) {
    const double n = 1.0 + 0.01 * std::cos(two_pi * (double)(r % nx) / nx);
    for (int m = 0; m < nv; ++m) {
      const double vm = v0 + m * dvv;
      h[r * nv + m] = (T)(n * std::exp(-0.5 * vm * vm) / std::sqrt(two_pi));
    }
  }
  return h;
}

struct TuneScratch {
  void* f = nullptr;
  double* e = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  ~TuneScratch() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (f) (void)hipFree(f);
    if (e) (void)hipFree(e);
  }
  template <class T>
  int init(const std::vector<T>& h, size_t ne) {
    VPFP_HIP(hipMalloc(&f, sizeof(T) * h.size()));
    VPFP_HIP(hipMalloc((void**)&e, sizeof(double) * ne));
    VPFP_HIP(hipMemcpy(f, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    VPFP_HIP(hipMemset(e, 0, sizeof(double) * ne));
    VPFP_HIP(hipEventCreate(&a));
    VPFP_HIP(hipEventCreate(&b));
    return VPFP_OK;
  }
  // microseconds per step of run(n_steps) after run(2) as warm-up
  template <class Run>
  int time(Run run, int n_steps, hipStream_t s, double* us) {
    if (int rc = run(2)) return rc;
    VPFP_HIP(hipEventRecord(a, s));
    if (int rc = run(n_steps)) return rc;
    VPFP_HIP(hipEventRecord(b, s));
    VPFP_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    VPFP_HIP(hipEventElapsedTime(&ms, a, b));
    *us = 1e3 * (double)ms / n_steps;
    return VPFP_OK;
  }
};

// values per thread that exist for this row length (delta-f rows have no
// 16-value form: q = 16 would run the 8-value kernel again)
std::vector<int> row_q_candidates(int nv, bool delta = false) {
  const int M = nv / 2;
  std::vector<int> q;
  if (M >= 16 && M <= 2048) q.push_back(4);
  q.push_back(8);
  if (M >= 256 && M <= 4096 && !delta) q.push_back(16);
  return q;
}

// the launch shape under test / kept
struct Shape {
  int q, C, fuse, xcd, dt, dft;
};

// Times every candidate shape with run(n) after `apply(shape)`; keeps the
// fastest (restores `keep` when none ran).  A shape that does not fit the
// grid returns VPFP_EUNSUPPORTED and is skipped; any other error ends the
// search and is returned.
template <class Apply, class Run>
int tune_search(const std::vector<Shape>& cands, const Shape& keep, Apply apply, Run run, TuneScratch& w,
                int n_steps, hipStream_t s, double* best_us) {
  double best = 1e300;
  Shape bs = keep;
  for (const Shape& c : cands) {
    apply(c);
    double us = 0.0;
    const int rc = w.time(run, n_steps, s, &us);
    if (rc == VPFP_EUNSUPPORTED) continue;
    if (rc) {
      apply(keep);
      return rc;
    }
    if (us < best) {
      best = us;
      bs = c;
    }
  }
  apply(best < 1e300 ? bs : keep);
  if (!(best < 1e300)) return fail(VPFP_EHIP, "autotune: no launch shape ran");
  *best_us = best;
  return VPFP_OK;
}

void plan_shape(vpfp_plan_t p, const Shape& c) {
  p->dft = c.dft;
  p->row_q = c.q;
  p->C = c.C;
  p->fuse = c.fuse;
  p->xcd = c.xcd;
  p->dtiles = c.dt;
  p->ntiles = drift_wgs(p->nv, p->C, p->dtiles, false);
}

void batch_shape(vpfp_batch_s* b, const Shape& c) {
  b->dft = c.dft;
  b->row_q = c.q;
  b->C = c.C;
  b->xcd = c.xcd;
  b->dtiles = c.dt;
  b->gx = drift_wgs(b->nv, b->C, b->dtiles, true);
}

template <class T>
int autotune_plan(vpfp_plan_t p, int n_steps, hipStream_t s) {
  TuneScratch w;
  if (int rc = w.init(tune_state<T>(p->nx, p->nx, p->nv, p->v0, p->dvv), 2 * (size_t)p->nx)) return rc;
  std::vector<Shape> cands;
  for (int q : row_q_candidates(p->nv, p->delta != 0))
    for (int C : {4, 8, 16, 32}) {
      if (C > p->nv) continue;
      for (int fuse : {0, 1, 2}) {
        if (fuse == 1 && p->nx > 256) continue;
        if (fuse == 2 && p->nx > vp::kProNx) continue;
        for (int xo : {0, 1})
          for (int dt : {1, 2, 4}) {
            // more tiles per workgroup only where it changes the grid
            if (dt > 1 && p->nv / C / dt < 8) continue;
            cands.push_back({q, C, fuse, xo, dt, 0});
            // the MFMA drift (no drift tail)
            if (p->dft_mat && (C == 16 || C == 32) && fuse != 1) cands.push_back({q, C, fuse, xo, dt, 1});
          }
      }
    }
  const Shape keep{p->row_q, p->C, p->fuse, p->xcd, p->dtiles, p->dft};
  const bool prof = p->prof.on;
  p->prof.on = false;
  double us = 0.0;
  const int rc = tune_search(
      cands, keep, [&](const Shape& c) { plan_shape(p, c); },
      [&](int n) { return do_step_n<T>(p, w.f, w.e, n, nullptr, nullptr, nullptr, nullptr, s); }, w, n_steps, s,
      &us);
  p->prof.on = prof;
  if (rc) return rc;
  p->tuned_us = us;
  return VPFP_OK;
}

template <class T>
int autotune_batch(vpfp_batch_s* b, int n_steps, hipStream_t s) {
  TuneScratch w;
  const size_t rows = (size_t)b->n * b->nx;
  if (int rc = w.init(tune_state<T>(rows, b->nx, b->nv, b->v0, b->dvv), rows)) return rc;
  std::vector<Shape> cands;
  for (int q : row_q_candidates(b->nv, b->delta != 0))
    for (int C : {4, 8, 16, 32}) {
      if (C > b->nv) continue;
      for (int xo : {0, 1})
        for (int dt : {1, 2, 4}) {
          if (dt > 1 && b->nv / C / dt < 2) continue;
          cands.push_back({q, C, 0, xo, dt, 0});
          if (b->dft_mat && (C == 16 || C == 32)) cands.push_back({q, C, 0, xo, dt, 1});
        }
    }
  const Shape keep{b->row_q, b->C, 0, b->xcd, b->dtiles, b->dft};
  const bool prof = b->prof.on;
  b->prof.on = false;
  double us = 0.0;
  const int rc = tune_search(
      cands, keep, [&](const Shape& c) { batch_shape(b, c); },
      [&](int n) { return batch_step_n<T>(b, w.f, w.e, n, nullptr, s); }, w, n_steps, s, &us);
  b->prof.on = prof;
  if (rc) return rc;
  b->tuned_us = us;
  return VPFP_OK;
}

int check_tuning(const vpfp_tuning* t, int nv) {
  if (!t) return fail(VPFP_EBADARG, "null tuning");
  if (t->row_q != 0 && t->row_q != 4 && t->row_q != 8 && t->row_q != 16)
    return fail(VPFP_EBADARG, "row_q must be 0, 4, 8 or 16");
  if (t->drift_tiles < 0 || t->drift_tiles > 64) return fail(VPFP_EBADARG, "drift_tiles must be in [0, 64]");
  if (t->drift_tile != 0 && t->drift_tile != 4 && t->drift_tile != 8 && t->drift_tile != 16 && t->drift_tile != 32)
    return fail(VPFP_EBADARG, "drift_tile must be 0, 4, 8, 16 or 32");
  if (t->drift_order != 0 && t->drift_order != 1) return fail(VPFP_EBADARG, "drift_order must be 0 or 1");
  if (t->drift_tile > nv) return fail(VPFP_EBADARG, "drift_tile exceeds nv");
  if (t->fuse_field < -1 || t->fuse_field > 2) return fail(VPFP_EBADARG, "fuse_field must be -1, 0, 1 or 2");
  if (t->drift_dft != 0 && t->drift_dft != 1) return fail(VPFP_EBADARG, "drift_dft must be 0 or 1");
  return VPFP_OK;
}

}  // namespace

// ====================================================================== grouped scan step
// Several batches (size classes of one scan) stepped together: one row
// launch per distinct nv, one drift and one field launch per distinct nx,
// each over the segments of the batches concerned.
struct vpfp_group_s {
  std::vector<vpfp_batch_s*> b;
  int prec = VPFP_F64, delta = 0;
  struct VGroup {
    int nv;
    std::vector<int> members;  // batch indices
    int q = 0;                 // row values per thread (0: row-count rule)
  };
  struct XGroup {
    int nx;
    std::vector<int> members;
    int C = 16;                // drift columns per tile
    int per = 1;               // column tiles per drift workgroup
    int dft = 0;               // MFMA drift (FP32, nx 32 or 64)
  };
  std::vector<VGroup> vg;
  std::vector<XGroup> xg;
  double tuned_us = 0.0;
  Prof prof;
  StreamLog streams;
};

namespace {

int group_gx(const vpfp_group_s::XGroup& x, const vpfp_batch_s* b) { return drift_wgs(b->nv, x.C, x.per, true); }

template <class T>
vp::RowMulti<T> group_rows(const vpfp_group_s* g, const vpfp_group_s::VGroup& v, void* const* f,
                           const double* const* e, int flags) {
  vp::RowMulti<T> m;
  std::memset(&m, 0, sizeof(m));
  const vpfp_batch_s* b0 = g->b[v.members[0]];
  vp::RowArgs<T>& a = m.base;
  a.nv = b0->nv;
  a.dk = b0->dk_v;
  a.v0 = b0->v0;
  a.dvv = b0->dvv;
  a.dv = b0->dv;
  a.tw = reinterpret_cast<const cx<T>*>(b0->tw_v);
  a.op_dg = b0->op == VPFP_OP_DG;
  a.dt_a = a.dt_b = 0.5 * b0->dt;
  a.flags = flags;
  if (g->delta) {
    a.bg = b0->bg;
    a.bg_hat = reinterpret_cast<const cx<T>*>(b0->bg_hat);
  }
  m.nseg = (int)v.members.size();
  for (int i = 0; i < m.nseg; ++i) {
    const int k = v.members[i];
    const vpfp_batch_s* b = g->b[k];
    m.seg[i].f = reinterpret_cast<T*>(f[k]);
    m.seg[i].e = e[k];
    m.seg[i].nu_dt = b->nu_dt;
    m.seg[i].rows = b->n * b->nx;
    m.seg[i].rows_per_inst = b->nx;
  }
  return m;
}

template <class T>
vp::DriftMulti<T> group_drift(const vpfp_group_s* g, const vpfp_group_s::XGroup& x, void* const* f) {
  vp::DriftMulti<T> m;
  std::memset(&m, 0, sizeof(m));
  const vpfp_batch_s* b0 = g->b[x.members[0]];
  m.base.nx = b0->nx;
  m.base.dt = b0->dt;
  m.base.tw = reinterpret_cast<const cx<T>*>(b0->tw_x);
  m.base.dft = x.dft ? b0->dft_mat : nullptr;
  m.nseg = (int)x.members.size();
  for (int i = 0; i < m.nseg; ++i) {
    const int k = x.members[i];
    const vpfp_batch_s* b = g->b[k];
    m.seg[i].f = reinterpret_cast<T*>(f[k]);
    m.seg[i].v = b->v;
    m.seg[i].dk = b->dk_x;
    m.seg[i].dens = b->dens;
    m.seg[i].nv = b->nv;
    m.seg[i].gx = group_gx(x, b);
    m.seg[i].n_inst = b->n;
  }
  return m;
}

vp::FieldMulti group_field(const vpfp_group_s* g, const vpfp_group_s::XGroup& x, double* const* e_out,
                           double* const* series) {
  vp::FieldMulti m;
  std::memset(&m, 0, sizeof(m));
  const vpfp_batch_s* b0 = g->b[x.members[0]];
  m.base.nx = b0->nx;
  m.base.scale = b0->dv;
  m.base.tw = b0->tw_xs;
  m.base.series_stride = 2;
  m.nseg = (int)x.members.size();
  for (int i = 0; i < m.nseg; ++i) {
    const int k = x.members[i];
    const vpfp_batch_s* b = g->b[k];
    m.seg[i].part = b->dens;
    m.seg[i].scale = b->dv;
    m.seg[i].ntiles = group_gx(x, b);
    m.seg[i].okx = b->okx;
    m.seg[i].e_out = e_out[k];
    m.seg[i].series = series ? series[k] : nullptr;
    m.seg[i].n_inst = b->n;
  }
  return m;
}

#define BPROF_G(kind, call)                               \
  do {                                                    \
    if (int rc_ = prof_mark(pr, kind, s)) return rc_;     \
    if (int rc_ = (call)) return rc_;                     \
    if (int rc_ = prof_mark(pr, kind, s)) return rc_;     \
  } while (0)

template <class T>
int group_step_n(vpfp_group_s* g, void* const* f, double* const* e, int n, double* const* series, hipStream_t s) {
  const size_t nb = g->b.size();
  std::vector<const double*> e_cur(e, e + nb);
  std::vector<double*> ei(nb), ser(nb);
  Prof& pr = g->prof;
  const int nvg = (int)g->vg.size(), nxg = (int)g->xg.size();
  if (n > 0)
    for (int j = 0; j < nvg; ++j)
      BPROF_G(KK_ROW + 4 * j, row_multi_dispatch<T>(group_rows<T>(g, g->vg[j], f, e_cur.data(), vp::RK_KICK_A), s,
                                                    g->vg[j].q));
  for (int i = 0; i < n; ++i) {
    for (size_t k = 0; k < nb; ++k) {
      const vpfp_batch_s* b = g->b[k];
      ei[k] = b->ering + (size_t)(i & 1) * b->n * b->nx;
      ser[k] = series && series[k] ? series[k] + (size_t)i * b->n * 2 : nullptr;
    }
    for (int j = 0; j < nxg; ++j)
      BPROF_G(KK_DRIFT + 4 * j, drift_multi_dispatch(group_drift<T>(g, g->xg[j], f), g->xg[j].C, s));
    for (int j = 0; j < nxg; ++j)
      BPROF_G(KK_FIELD + 4 * j, field_multi_dispatch(group_field(g, g->xg[j], ei.data(), ser.data()), s));
    for (size_t k = 0; k < nb; ++k) e_cur[k] = ei[k];
    const int flags = vp::RK_KICK_A | vp::RK_COLLIDE | (i + 1 < n ? vp::RK_KICK_B : 0);
    for (int j = 0; j < nvg; ++j)
      BPROF_G(KK_ROW + 4 * j, row_multi_dispatch<T>(group_rows<T>(g, g->vg[j], f, e_cur.data(), flags), s, g->vg[j].q));
  }
  if (n > 0)
    for (size_t k = 0; k < nb; ++k) {
      const vpfp_batch_s* b = g->b[k];
      VPFP_HIP(hipMemcpyAsync(e[k], e_cur[k], sizeof(double) * b->n * b->nx, hipMemcpyDeviceToDevice, s));
    }
  return prof_collect(pr);
}
#undef BPROF_G

// Launch shapes of a group: each row group's values per thread and each x
// group's drift tile (columns, tiles per workgroup) timed on scratch state,
// one kernel at a time, then the whole step with the chosen shapes.
template <class T>
int autotune_group(vpfp_group_s* g, int n_steps, hipStream_t s) {
  const size_t nb = g->b.size();
  struct Scratch {
    std::vector<void*> f;
    std::vector<double*> e;
    hipEvent_t a = nullptr, b = nullptr;
    ~Scratch() {
      for (void* p : f) (void)hipFree(p);
      for (double* p : e) (void)hipFree(p);
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } w;
  for (size_t k = 0; k < nb; ++k) {
    const vpfp_batch_s* b = g->b[k];
    const size_t rows = (size_t)b->n * b->nx;
    auto h = tune_state<T>(rows, b->nx, b->nv, b->v0, b->dvv);
    if (g->delta)
      for (size_t r = 0; r < rows; ++r)
        for (int m = 0; m < b->nv; ++m) {
          const double vm = b->v0 + m * b->dvv;
          h[r * b->nv + m] = (T)((double)h[r * b->nv + m] - std::exp(-0.5 * vm * vm) / std::sqrt(6.283185307179586));
        }
    void* fp = nullptr;
    double* ep = nullptr;
    VPFP_HIP(hipMalloc(&fp, sizeof(T) * h.size()));
    w.f.push_back(fp);
    VPFP_HIP(hipMalloc((void**)&ep, sizeof(double) * rows));
    w.e.push_back(ep);
    VPFP_HIP(hipMemcpy(fp, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    VPFP_HIP(hipMemset(ep, 0, sizeof(double) * rows));
  }
  VPFP_HIP(hipEventCreate(&w.a));
  VPFP_HIP(hipEventCreate(&w.b));
  const bool prof = g->prof.on;
  g->prof.on = false;
  // microseconds per call of run() over n_steps calls (after one untimed call);
  // VPFP_EUNSUPPORTED (a shape that does not fit) is passed through
  auto time = [&](auto run, double* us) -> int {
    if (int rc = run()) return rc;
    VPFP_HIP(hipEventRecord(w.a, s));
    for (int i = 0; i < n_steps; ++i)
      if (int rc = run()) return rc;
    VPFP_HIP(hipEventRecord(w.b, s));
    VPFP_HIP(hipEventSynchronize(w.b));
    float ms = 0.f;
    VPFP_HIP(hipEventElapsedTime(&ms, w.a, w.b));
    *us = 1e3 * (double)ms / n_steps;
    return VPFP_OK;
  };
  int rc = VPFP_OK;
  const int flags = vp::RK_KICK_A | vp::RK_COLLIDE | vp::RK_KICK_B;
  for (auto& v : g->vg) {
    double best = 1e300;
    int bq = v.q;
    for (int q : row_q_candidates(v.nv, g->b[v.members[0]]->delta != 0)) {
      double us = 0.0;
      const int r = time([&]() { return row_multi_dispatch<T>(group_rows<T>(g, v, w.f.data(), w.e.data(), flags), s, q); },
                         &us);
      if (r == VPFP_EUNSUPPORTED) continue;
      if (r) {
        rc = r;
        break;
      }
      if (us < best) {
        best = us;
        bq = q;
      }
    }
    if (rc) break;
    v.q = bq;
  }
  std::vector<double*> eo(w.e.begin(), w.e.end());
  for (auto& x : g->xg) {
    if (rc) break;
    double best = 1e300;
    int bC = x.C, bper = x.per, bdft = x.dft;
    const bool can_dft = g->b[x.members[0]]->dft_mat != nullptr;
    for (int dft : {0, 1})
    for (int C : {4, 8, 16, 32})
      for (int per : {1, 2, 4}) {
        if (dft && (!can_dft || (C != 16 && C != 32))) continue;
        bool fits = true;
        for (int k : x.members)
          if (C > g->b[k]->nv || (per > 1 && g->b[k]->nv / C / per < 2)) fits = false;
        if (!fits) continue;
        x.C = C;
        x.per = per;
        x.dft = dft;
        double us = 0.0;
        const int r = time(
            [&]() -> int {
              if (int rr = drift_multi_dispatch(group_drift<T>(g, x, w.f.data()), x.C, s)) return rr;
              return field_multi_dispatch(group_field(g, x, eo.data(), nullptr), s);
            },
            &us);
        if (r == VPFP_EUNSUPPORTED) continue;
        if (r) {
          rc = r;
          break;
        }
        if (us < best) {
          best = us;
          bC = C;
          bper = per;
          bdft = dft;
        }
      }
    x.C = bC;
    x.per = bper;
    x.dft = bdft;
  }
  if (!rc) {
    double us = 0.0;
    rc = time([&]() { return group_step_n<T>(g, w.f.data(), eo.data(), 1, nullptr, s); }, &us);
    if (!rc) g->tuned_us = us;
  }
  g->prof.on = prof;
  return rc;
}

}  // namespace

// ====================================================================== C ABI
namespace {

// unit Maxwellian on a v grid (v_m = v0 + m dvv) and its DFT
// sum_m M_m exp(-2 pi i k m / N) (direct sum in FP64 with an exact twiddle index)
void background_tables(int N, double v0, double dvv, std::vector<double>& m, std::vector<cx<double>>& hat) {
  m.assign(N, 0.0);
  for (int i = 0; i < N; ++i) {
    const double v = v0 + i * dvv;
    m[i] = std::exp(-0.5 * v * v) / std::sqrt(6.283185307179586);
  }
  std::vector<double> wr(N), wi(N);
  for (int j = 0; j < N; ++j) {
    wr[j] = std::cos(6.283185307179586 * j / N);
    wi[j] = -std::sin(6.283185307179586 * j / N);
  }
  hat.assign(N, vp::mk<double>(0.0, 0.0));
  for (int k = 0; k < N; ++k) {
    double re = 0.0, im = 0.0;
    for (int i = 0; i < N; ++i) {
      const int j = (int)(((long long)k * i) % N);
      re += m[i] * wr[j];
      im += m[i] * wi[j];
    }
    hat[k] = vp::mk<double>(re, im);
  }
}

std::vector<cx<float>> to_float(const std::vector<cx<double>>& h) {
  std::vector<cx<float>> out(h.size());
  for (size_t k = 0; k < h.size(); ++k) out[k] = vp::mk<float>((float)h[k].re, (float)h[k].im);
  return out;
}

// once per batch
int batch_background(vpfp_batch_s* b) {
  if (b->bg) return VPFP_OK;
  const int N = b->nv;
  std::vector<double> m;
  std::vector<cx<double>> hat;
  background_tables(N, b->v0, b->dvv, m, hat);
  if (int rc = bput(b, &b->bg, m)) return rc;
  if (b->prec == VPFP_F64) {
    cx<double>* d = nullptr;
    if (int rc = bput(b, &d, hat)) return rc;
    b->bg_hat = d;
  } else {
    cx<float>* d = nullptr;
    if (int rc = bput(b, &d, to_float(hat))) return rc;
    b->bg_hat = d;
  }
  return VPFP_OK;
}

// once per plan
int plan_background(vpfp_plan_s* p) {
  if (p->bg) return VPFP_OK;
  std::vector<double> m;
  std::vector<cx<double>> hat;
  background_tables(p->nv, p->v0, p->dvv, m, hat);
  if (int rc = put(p, &p->bg, m)) return rc;
  if (p->prec == VPFP_F64) {
    cx<double>* d = nullptr;
    if (int rc = put(p, &d, hat)) return rc;
    p->bg_hat = d;
  } else {
    cx<float>* d = nullptr;
    if (int rc = put(p, &d, to_float(hat))) return rc;
    p->bg_hat = d;
  }
  return VPFP_OK;
}

int no_delta(vpfp_plan_s* p, const char* what) {
  if (p->delta) return fail(VPFP_EUNSUPPORTED, std::string(what) + " reads f itself: not with delta-f storage");
  return VPFP_OK;
}

}  // namespace

// ====================================================================== allocation-free field solve
namespace {

// Scratch of vpfp_field_noalloc: the density n(x) (VPFP_NOALLOC_MAX_NX
// doubles) then the twiddle table (VPFP_NOALLOC_MAX_NX complex doubles).
// Each GPU holds one copy until the process ends and every call on that GPU
// shares it.  Host code never passes `scratch` itself to a kernel: its
// device address comes from hipGetSymbolAddress, once per device.
constexpr size_t kNoAllocTwOffset = (size_t)VPFP_NOALLOC_MAX_NX * sizeof(double);
static_assert(kNoAllocTwOffset + (size_t)VPFP_NOALLOC_MAX_NX * sizeof(cx<double>) <= VPFP_NOALLOC_SCRATCH_BYTES,
              "scratch layout");
__device__ __align__(256) unsigned char scratch[VPFP_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch;  // device address of `scratch`, by device ordinal

// address of `scratch` on the current device (looked up on the first call
// on that device)
int scratch_address(unsigned char** out) {
  int dev = 0;
  VPFP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)dev >= g_scratch.size()) g_scratch.resize((size_t)dev + 1, nullptr);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    VPFP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[dev] = reinterpret_cast<unsigned char*>(ptr);
  }
  *out = g_scratch[dev];
  return VPFP_OK;
}

}  // namespace

extern "C" {

int vpfp_version(void) { return 1; }

const char* vpfp_last_error(void) { return vpfp_impl::last_error(); }

int vpfp_plan_create(const vpfp_grid_desc* d, vpfp_plan_t* out) {
  if (!d || !out) return fail(VPFP_EBADARG, "null argument");
  *out = nullptr;
  if (d->nx < 2 || d->nv < 2) return fail(VPFP_EBADARG, "nx and nv must be at least 2");
  if (d->precision != VPFP_F64 && d->precision != VPFP_F32) return fail(VPFP_EBADARG, "unknown precision");
  if (d->op != VPFP_OP_LB && d->op != VPFP_OP_DG)
    return fail(VPFP_EUNSUPPORTED, "collision operator must be lb or dg");
  if (d->nu < 0.0) return fail(VPFP_EUNSUPPORTED, "negative collision frequency");
  if (!d->kx || !d->one_over_kx || !d->v || !d->kv) return fail(VPFP_EBADARG, "null grid array");
  if (!(d->dv > 0.0)) return fail(VPFP_EBADARG, "dv must be positive");
  if (!fftfreq_like(d->kx, d->nx)) return fail(VPFP_EUNSUPPORTED, "kx is not an fftfreq wavenumber table");
  if (!fftfreq_like(d->kv, d->nv)) return fail(VPFP_EUNSUPPORTED, "kv is not an fftfreq wavenumber table");
  double v0 = 0.0, dvv = 0.0;
  if (!uniform_grid(d->v, d->nv, &v0, &dvv)) return fail(VPFP_EUNSUPPORTED, "v is not a uniform grid");

  auto* p = new vpfp_plan_s();
  p->nx = d->nx;
  p->nv = d->nv;
  p->prec = d->precision;
  p->op = d->op;
  p->dv = d->dv;
  p->nu = d->nu;
  p->dt = d->dt;
  p->dk_x = d->nx > 1 ? d->kx[1] : 0.0;
  p->dk_v = d->nv > 1 ? d->kv[1] : 0.0;
  p->v0 = v0;
  p->dvv = dvv;
  p->nx_pow2 = is_pow2(d->nx);
  p->nv_pow2 = is_pow2(d->nv);
  p->C = (p->nx_pow2 && p->nv_pow2 && p->nx >= 4) ? drift_tile(p->nx, p->nv) : 0;
  p->ntiles = p->C ? drift_wgs(p->nv, p->C, p->dtiles, false) : 0;
  int rc = VPFP_OK;
  const size_t nxb = sizeof(double) * d->nx, nvb = sizeof(double) * d->nv;
  if (!rc) rc = dalloc(p, &p->kx, nxb);
  if (!rc) rc = dalloc(p, &p->okx, nxb);
  if (!rc) rc = dalloc(p, &p->v, nvb);
  if (!rc) rc = dalloc(p, &p->kv, nvb);
  if (!rc) rc = upload(p->kx, d->kx, nxb);
  if (!rc) rc = upload(p->okx, d->one_over_kx, nxb);
  if (!rc) rc = upload(p->v, d->v, nvb);
  if (!rc) rc = upload(p->kv, d->kv, nvb);
  if (!rc) rc = p->prec == VPFP_F64 ? make_tables<double>(p) : make_tables<float>(p);
  if (!rc && dft_size(p->prec, p->nx)) rc = put(p, &p->dft_mat, dft_matrices(p->nx));
  // partial density rows for the narrowest drift tile any tuning may pick
  if (!rc) rc = dalloc(p, &p->dens, nxb * (size_t)(p->C ? drift_grid(p->nv / 4) : 1));
  if (!rc) rc = dalloc(p, &p->ering, 2 * nxb);
  if (!rc) rc = dalloc(p, &p->series, 2 * sizeof(double));
  if (!rc) rc = dalloc(p, &p->rho_g, nxb);
  if (!rc) rc = dalloc(p, &p->counter, 64);
  if (!rc) {
    hipError_t e = hipMemset(p->counter, 0, 64);
    if (e != hipSuccess) rc = fail(VPFP_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
  }
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return VPFP_OK;
}

int vpfp_plan_destroy(vpfp_plan_t p) {
  if (!p) return VPFP_OK;
  p->streams.drain();
  delete p;
  return VPFP_OK;
}

int vpfp_plan_set_delta(vpfp_plan_t p, int on) {
  if (int rc = check_plan(p)) return rc;
  if (on != 0 && on != 1) return fail(VPFP_EBADARG, "delta must be 0 or 1");
  if (on && !p->nv_pow2) return fail(VPFP_EUNSUPPORTED, "delta-f storage needs a power-of-two nv");
  if (on)
    if (int rc = plan_background(p)) return rc;
  p->delta = on;
  return VPFP_OK;
}

int vpfp_plan_background(vpfp_plan_t p, void* f, int sign, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!f) return fail(VPFP_EBADARG, "null array");
  if (sign != 1 && sign != -1) return fail(VPFP_EBADARG, "sign must be 1 or -1");
  if (int rc = plan_background(p)) return rc;
  const size_t n = (size_t)p->nx * p->nv;
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (p->prec == VPFP_F64)
    hipLaunchKernelGGL(k_background<double>, dim3(grid), dim3(256), 0, p->streams.use(stream), reinterpret_cast<double*>(f), n,
                       p->nv, p->bg, (double)sign);
  else
    hipLaunchKernelGGL(k_background<float>, dim3(grid), dim3(256), 0, p->streams.use(stream), reinterpret_cast<float*>(f), n,
                       p->nv, p->bg, (double)sign);
  VPFP_HIP(hipGetLastError());
  return VPFP_OK;
}

int vpfp_vdfdx(vpfp_plan_t p, const void* fin, void* fout, double dt, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = need_spectral_x(p)) return rc;
  if (!fin || !fout) return fail(VPFP_EBADARG, "null array");
  return p->prec == VPFP_F64 ? do_vdfdx<double>(p, fin, fout, dt, nullptr, p->streams.use(stream))
                             : do_vdfdx<float>(p, fin, fout, dt, nullptr, p->streams.use(stream));
}

int vpfp_edfdv(vpfp_plan_t p, const void* fin, void* fout, const double* e, double dt, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = need_spectral_v(p)) return rc;
  if (!fin || !fout || !e) return fail(VPFP_EBADARG, "null array");
  return p->prec == VPFP_F64 ? do_edfdv<double>(p, fin, fout, e, dt, p->streams.use(stream))
                             : do_edfdv<float>(p, fin, fout, e, dt, p->streams.use(stream));
}

int vpfp_compute_charges(vpfp_plan_t p, const void* f, double* n_out, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = no_delta(p, "vpfp_compute_charges")) return rc;
  if (!f || !n_out) return fail(VPFP_EBADARG, "null array");
  return p->prec == VPFP_F64 ? do_density<double>(p, f, p->dv, n_out, p->streams.use(stream))
                             : do_density<float>(p, f, p->dv, n_out, p->streams.use(stream));
}

int vpfp_solve_for_field(vpfp_plan_t p, const double* n, double* e_out, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!n || !e_out) return fail(VPFP_EBADARG, "null array");
  if (p->nx < 2) return fail(VPFP_EUNSUPPORTED, "nx must be at least 2");
  const bool pow2 = p->nx_pow2 && p->nx >= 4 && p->nx <= 4096;
  auto fa = field_args(p, n, 1, 1.0, nullptr, e_out, nullptr, nullptr, pow2);
  return field_dispatch(fa, pow2, p->streams.use(stream));
}

int vpfp_field(vpfp_plan_t p, const void* f, const double* driver, double* e_out, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!f || !e_out) return fail(VPFP_EBADARG, "null array");
  int rc = p->prec == VPFP_F64 ? do_density<double>(p, f, p->dv, p->dens, p->streams.use(stream))
                               : do_density<float>(p, f, p->dv, p->dens, p->streams.use(stream));
  if (rc) return rc;
  const bool pow2 = p->nx_pow2 && p->nx >= 4 && p->nx <= 4096;
  auto fa = field_args(p, p->dens, 1, 1.0, driver, e_out, nullptr, nullptr, pow2);
  return field_dispatch(fa, pow2, p->streams.use(stream));
}

int vpfp_field_noalloc(int nx, int nv, int precision, double dv, const double* one_over_kx, const void* f,
                       const double* driver, double* e_out, void* stream) {
  if (nx < 2 || nv < 2) return fail(VPFP_EBADARG, "nx and nv must be at least 2");
  if (nx > VPFP_NOALLOC_MAX_NX)
    return fail(VPFP_EBADARG, "nx = " + std::to_string(nx) + " exceeds the scratch of vpfp_field_noalloc (nx <= " +
                                  std::to_string(VPFP_NOALLOC_MAX_NX) + ")");
  if (precision != VPFP_F64 && precision != VPFP_F32) return fail(VPFP_EBADARG, "unknown precision");
  if (!(dv > 0.0)) return fail(VPFP_EBADARG, "dv must be positive");
  if (!one_over_kx || !f || !e_out) return fail(VPFP_EBADARG, "null array");
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  double* dens = reinterpret_cast<double*>(base);
  cx<double>* tw = reinterpret_cast<cx<double>*>(base + kNoAllocTwOffset);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // power-of-two nx: the FFT field kernel in one workgroup (tables as
  // split_twiddles lays them out); any other nx: the direct DFT (full table)
  const bool pow2 = is_pow2(nx) && nx >= 4;
  const int split = (pow2 && !vp::tw_full(nx, (int)sizeof(cx<double>))) ? 1 : 0;
  const int entries = split ? 64 + (nx / 64 > 1 ? nx / 64 : 1) : nx;
  hipLaunchKernelGGL(k_twiddles, dim3((entries + 255) / 256), dim3(256), 0, s, tw, nx, split);
  if (int rc = launched()) return rc;
  // n(x) = dv * trapz row sums: every entry is written before it is read
  const int rows = 4;
  if (precision == VPFP_F64)
    hipLaunchKernelGGL(k_density<double>, dim3((nx + rows - 1) / rows), dim3(64 * rows), 0, s,
                       reinterpret_cast<const double*>(f), nx, nv, dv, dens);
  else
    hipLaunchKernelGGL(k_density<float>, dim3((nx + rows - 1) / rows), dim3(64 * rows), 0, s,
                       reinterpret_cast<const float*>(f), nx, nv, dv, dens);
  if (int rc = launched()) return rc;
  vp::FieldArgs a;
  std::memset(&a, 0, sizeof(a));
  a.part = dens;
  a.ntiles = 1;
  a.nx = nx;
  a.scale = 1.0;
  a.okx = one_over_kx;
  a.tw = tw;
  a.driver = driver;
  a.e_out = e_out;
  // one workgroup (batched form with one instance): no cross-workgroup
  // counter or reduction buffer
  return field_dispatch(a, pow2, s, pow2 ? 1 : 0);
}

int vpfp_collision_diagonals(vpfp_plan_t p, const void* f, double* a, double* b, double* c, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = no_delta(p, "vpfp_collision_diagonals")) return rc;
  if (!f || !a || !b || !c) return fail(VPFP_EBADARG, "null array");
  return p->prec == VPFP_F64 ? do_diag<double>(p, f, a, b, c, p->streams.use(stream)) : do_diag<float>(p, f, a, b, c, p->streams.use(stream));
}

int vpfp_tridiag_solve(vpfp_plan_t p, const double* a, const double* b, const double* c, const void* d, void* x,
                       void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = no_delta(p, "vpfp_tridiag_solve")) return rc;
  if (int rc = need_spectral_v(p)) return rc;
  if (!a || !b || !c || !d || !x) return fail(VPFP_EBADARG, "null array");
  return p->prec == VPFP_F64 ? do_tri<double>(p, a, b, c, d, x, p->streams.use(stream))
                             : do_tri<float>(p, a, b, c, d, x, p->streams.use(stream));
}

int vpfp_collide(vpfp_plan_t p, const void* fin, void* fout, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!fin || !fout) return fail(VPFP_EBADARG, "null array");
  if (p->nu == 0.0) {
    if (fin != fout) {
      const size_t bytes = (size_t)p->nx * p->nv * (p->prec == VPFP_F64 ? 8 : 4);
      VPFP_HIP(hipMemcpyAsync(fout, fin, bytes, hipMemcpyDeviceToDevice, p->streams.use(stream)));
    }
    return VPFP_OK;
  }
  if (int rc = need_spectral_v(p)) return rc;
  return p->prec == VPFP_F64 ? do_collide<double>(p, fin, fout, p->streams.use(stream)) : do_collide<float>(p, fin, fout, p->streams.use(stream));
}

int vpfp_moments(vpfp_plan_t p, const void* f, double* moments, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!f || !moments) return fail(VPFP_EBADARG, "null array");
  if (int rc = need_spectral_v(p)) return rc;
  if (p->prec == VPFP_F64) {
    auto a = row_args<double>(p, f, nullptr);
    a.moments = moments;
    a.flags = vp::RK_MOMENTS;
    return row_dispatch<double>(a, p->streams.use(stream), p->row_q);
  }
  auto a = row_args<float>(p, f, nullptr);
  a.moments = moments;
  a.flags = vp::RK_MOMENTS;
  return row_dispatch<float>(a, p->streams.use(stream), p->row_q);
}

int vpfp_x_modes(vpfp_plan_t p, const void* f, int n_modes, double* out, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = no_delta(p, "vpfp_x_modes")) return rc;
  if (!f || !out) return fail(VPFP_EBADARG, "null array");
  if (n_modes < 1 || n_modes > p->nx) return fail(VPFP_EBADARG, "n_modes must be in [1, nx]");
  const int total = n_modes * p->nv;
  const dim3 grid((total + 255) / 256);
  auto* o = reinterpret_cast<cx<double>*>(out);
  if (p->prec == VPFP_F64)
    hipLaunchKernelGGL(k_xmodes<double>, grid, dim3(256), 0, p->streams.use(stream), reinterpret_cast<const double*>(f), p->nx,
                       p->nv, n_modes, p->tw_xd, o);
  else
    hipLaunchKernelGGL(k_xmodes<float>, grid, dim3(256), 0, p->streams.use(stream), reinterpret_cast<const float*>(f), p->nx,
                       p->nv, n_modes, p->tw_xd, o);
  return launched();
}

int vpfp_batch_create(const vpfp_batch_desc* d, vpfp_batch_t* out) {
  if (!d || !out) return fail(VPFP_EBADARG, "null argument");
  *out = nullptr;
  if (d->n_inst < 1 || d->n_inst > 65535) return fail(VPFP_EBADARG, "n_inst must be in [1, 65535]");
  if (!is_pow2(d->nx) || d->nx < 4 || d->nx > 256)
    return fail(VPFP_EUNSUPPORTED, "batched scans need nx a power of two in [4, 256]");
  if (!is_pow2(d->nv) || d->nv < 16 || d->nv > 8192)
    return fail(VPFP_EUNSUPPORTED, "batched scans need nv a power of two in [16, 8192]");
  if (d->precision != VPFP_F64 && d->precision != VPFP_F32) return fail(VPFP_EBADARG, "unknown precision");
  if (d->op != VPFP_OP_LB && d->op != VPFP_OP_DG) return fail(VPFP_EUNSUPPORTED, "collision operator must be lb or dg");
  if (!d->kx || !d->one_over_kx || !d->v || !d->kv || !d->nu) return fail(VPFP_EBADARG, "null array");
  if (!(d->dv > 0.0)) return fail(VPFP_EBADARG, "dv must be positive");
  if (!fftfreq_like(d->kv, d->nv)) return fail(VPFP_EUNSUPPORTED, "kv is not an fftfreq wavenumber table");
  double v0 = 0.0, dvv = 0.0;
  if (!uniform_grid(d->v, d->nv, &v0, &dvv)) return fail(VPFP_EUNSUPPORTED, "v is not a uniform grid");
  std::vector<double> dk(d->n_inst), nudt(d->n_inst);
  for (int i = 0; i < d->n_inst; ++i) {
    if (!fftfreq_like(d->kx + (size_t)i * d->nx, d->nx))
      return fail(VPFP_EUNSUPPORTED, "kx of instance " + std::to_string(i) + " is not an fftfreq table");
    if (d->nu[i] < 0.0) return fail(VPFP_EUNSUPPORTED, "negative collision frequency");
    dk[i] = d->kx[(size_t)i * d->nx + 1];
    nudt[i] = d->nu[i] * d->dt;
  }
  auto* b = new vpfp_batch_s();
  b->n = d->n_inst;
  b->nx = d->nx;
  b->nv = d->nv;
  b->prec = d->precision;
  b->op = d->op;
  b->dv = d->dv;
  b->dt = d->dt;
  b->v0 = v0;
  b->dvv = dvv;
  b->dk_v = d->kv[1];
  b->C = drift_tile(d->nx, d->nv);
  b->gx = drift_wgs(d->nv, b->C, b->dtiles, true);
  std::vector<double> vv(d->v, d->v + d->nv);
  std::vector<double> okx(d->one_over_kx, d->one_over_kx + (size_t)d->n_inst * d->nx);
  int rc = VPFP_OK;
  if (!rc) rc = bput(b, &b->dk_x, dk);
  if (!rc) rc = bput(b, &b->nu_dt, nudt);
  if (!rc) rc = bput(b, &b->v, vv);
  if (!rc) rc = bput(b, &b->okx, okx);
  if (!rc) rc = b->prec == VPFP_F64 ? batch_tables<double>(b) : batch_tables<float>(b);
  if (!rc && dft_size(b->prec, b->nx)) rc = bput(b, &b->dft_mat, dft_matrices(b->nx));
  // partial density rows for the narrowest drift tile any tuning may pick
  if (!rc) rc = balloc(b, &b->dens, sizeof(double) * (size_t)b->n * (b->nv / 4) * b->nx);
  if (!rc) rc = balloc(b, &b->counters, sizeof(unsigned) * 16 * (size_t)b->n);
  if (!rc) rc = balloc(b, &b->ering, sizeof(double) * 2 * (size_t)b->n * b->nx);
  if (!rc) {
    hipError_t e = hipMemset(b->counters, 0, sizeof(unsigned) * 16 * (size_t)b->n);
    if (e != hipSuccess) rc = fail(VPFP_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
  }
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return VPFP_OK;
}

int vpfp_batch_set_closure(vpfp_batch_t b, vpfc_net_t net) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  if (b->groups) return fail(VPFP_EUNSUPPORTED, "the batch belongs to a group (grouped steps run no closure)");
  if (net && (b->prec != VPFP_F32 || b->nv != VPFC_NV))
    return fail(VPFP_EUNSUPPORTED, "the closure runs on FP32 batches with nv = " + std::to_string(VPFC_NV));
  if (net && b->delta) return fail(VPFP_EUNSUPPORTED, "the closure reads f itself: not with delta-f storage");
  b->closure = net;
  return VPFP_OK;
}


int vpfp_batch_set_delta(vpfp_batch_t b, int on) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  if (b->groups && on != b->delta) return fail(VPFP_EUNSUPPORTED, "the batch belongs to a group: its storage is fixed");
  if (on != 0 && on != 1) return fail(VPFP_EBADARG, "delta must be 0 or 1");
  if (on && b->closure) return fail(VPFP_EUNSUPPORTED, "the closure reads f itself: not with delta-f storage");
  if (on)
    if (int rc = batch_background(b)) return rc;
  b->delta = on;
  return VPFP_OK;
}

int vpfp_batch_background(vpfp_batch_t b, void* f, int sign, void* stream) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  if (!f) return fail(VPFP_EBADARG, "null array");
  if (sign != 1 && sign != -1) return fail(VPFP_EBADARG, "sign must be 1 or -1");
  if (int rc = batch_background(b)) return rc;
  const size_t n = (size_t)b->n * b->nx * b->nv;
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (b->prec == VPFP_F64)
    hipLaunchKernelGGL(k_background<double>, dim3(grid), dim3(256), 0, b->streams.use(stream), reinterpret_cast<double*>(f), n,
                       b->nv, b->bg, (double)sign);
  else
    hipLaunchKernelGGL(k_background<float>, dim3(grid), dim3(256), 0, b->streams.use(stream), reinterpret_cast<float*>(f), n,
                       b->nv, b->bg, (double)sign);
  VPFP_HIP(hipGetLastError());
  return VPFP_OK;
}

int vpfp_batch_destroy(vpfp_batch_t b) {
  if (!b) return VPFP_OK;
  b->streams.drain();
  delete b;
  return VPFP_OK;
}

int vpfp_batch_step_n(vpfp_batch_t b, void* f, double* e, int n_steps, double* series, void* stream) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  if (!f || !e) return fail(VPFP_EBADARG, "null array");
  if (n_steps < 0) return fail(VPFP_EBADARG, "negative step count");
  return b->prec == VPFP_F64 ? batch_step_n<double>(b, f, e, n_steps, series, b->streams.use(stream))
                             : batch_step_n<float>(b, f, e, n_steps, series, b->streams.use(stream));
}

int vpfp_plan_get_tuning(vpfp_plan_t p, vpfp_tuning* t) {
  if (int rc = check_plan(p)) return rc;
  if (!t) return fail(VPFP_EBADARG, "null tuning");
  t->row_q = p->row_q;
  t->drift_tile = p->C;
  t->fuse_field = p->fuse;
  t->drift_order = p->xcd;
  t->drift_tiles = p->dtiles;
  t->us_per_step = p->tuned_us;
  t->drift_dft = p->dft;
  return VPFP_OK;
}

int vpfp_plan_set_tuning(vpfp_plan_t p, const vpfp_tuning* t) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = check_tuning(t, p->nv)) return rc;
  if (t->drift_dft && !p->dft_mat) return fail(VPFP_EUNSUPPORTED, "the MFMA drift needs FP32 and nx 32 or 64");
  p->dft = t->drift_dft;
  p->row_q = t->row_q;
  if (t->drift_tiles) p->dtiles = t->drift_tiles;
  if (p->C && t->drift_tile) p->C = t->drift_tile;
  if (p->C) p->ntiles = drift_wgs(p->nv, p->C, p->dtiles, false);
  p->fuse = t->fuse_field;
  p->xcd = t->drift_order;
  return VPFP_OK;
}

int vpfp_plan_autotune(vpfp_plan_t p, int n_steps, void* stream, vpfp_tuning* chosen) {
  if (int rc = check_plan(p)) return rc;
  if (n_steps < 1) return fail(VPFP_EBADARG, "n_steps must be positive");
  if (int rc = need_spectral_v(p)) return rc;
  if (int rc = need_spectral_x(p)) return rc;
  const int rc = p->prec == VPFP_F64 ? autotune_plan<double>(p, n_steps, p->streams.use(stream))
                                     : autotune_plan<float>(p, n_steps, p->streams.use(stream));
  if (rc) return rc;
  return chosen ? vpfp_plan_get_tuning(p, chosen) : VPFP_OK;
}

int vpfp_batch_get_tuning(vpfp_batch_t b, vpfp_tuning* t) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  if (!t) return fail(VPFP_EBADARG, "null tuning");
  t->row_q = b->row_q;
  t->drift_tile = b->C;
  t->fuse_field = 0;
  t->drift_order = b->xcd;
  t->drift_tiles = b->dtiles;
  t->us_per_step = b->tuned_us;
  t->drift_dft = b->dft;
  return VPFP_OK;
}

int vpfp_batch_set_tuning(vpfp_batch_t b, const vpfp_tuning* t) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  if (int rc = check_tuning(t, b->nv)) return rc;
  if (t->drift_dft && !b->dft_mat) return fail(VPFP_EUNSUPPORTED, "the MFMA drift needs FP32 and nx 32 or 64");
  b->dft = t->drift_dft;
  b->row_q = t->row_q;
  if (t->drift_tiles) b->dtiles = t->drift_tiles;
  if (t->drift_tile) b->C = t->drift_tile;
  b->gx = drift_wgs(b->nv, b->C, b->dtiles, true);
  b->xcd = t->drift_order;
  return VPFP_OK;
}

int vpfp_batch_autotune(vpfp_batch_t b, int n_steps, void* stream, vpfp_tuning* chosen) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  if (n_steps < 1) return fail(VPFP_EBADARG, "n_steps must be positive");
  const int rc = b->prec == VPFP_F64 ? autotune_batch<double>(b, n_steps, b->streams.use(stream))
                                     : autotune_batch<float>(b, n_steps, b->streams.use(stream));
  if (rc) return rc;
  return chosen ? vpfp_batch_get_tuning(b, chosen) : VPFP_OK;
}

int vpfp_plan_set_profiling(vpfp_plan_t p, int on) {
  if (int rc = check_plan(p)) return rc;
  p->prof.reset(on != 0);
  return VPFP_OK;
}

int vpfp_plan_kernel_times(vpfp_plan_t p, double* ms, long long* counts) {
  if (int rc = check_plan(p)) return rc;
  if (!ms || !counts) return fail(VPFP_EBADARG, "null array");
  for (int k = 0; k < 4; ++k) {
    ms[k] = p->prof.ms[k];
    counts[k] = p->prof.n[k];
  }
  return VPFP_OK;
}

int vpfp_batch_set_profiling(vpfp_batch_t b, int on) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  b->prof.reset(on != 0);
  return VPFP_OK;
}

int vpfp_batch_kernel_times(vpfp_batch_t b, double* ms, long long* counts) {
  if (!b) return fail(VPFP_EBADARG, "null batch");
  if (!ms || !counts) return fail(VPFP_EBADARG, "null array");
  for (int k = 0; k < 4; ++k) {
    ms[k] = b->prof.ms[k];
    counts[k] = b->prof.n[k];
  }
  return VPFP_OK;
}

int vpfp_step_n(vpfp_plan_t p, void* f, double* e, int n, const double* driver, double* e_hist, double* moments,
                double* series, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!f || !e) return fail(VPFP_EBADARG, "null array");
  if (n < 0) return fail(VPFP_EBADARG, "negative step count");
  if (int rc = need_spectral_v(p)) return rc;
  if (int rc = need_spectral_x(p)) return rc;
  if (p->nx > 2048) return fail(VPFP_EUNSUPPORTED, "nx too large for the fused step");
  return p->prec == VPFP_F64 ? do_step_n<double>(p, f, e, n, driver, e_hist, moments, series, p->streams.use(stream))
                             : do_step_n<float>(p, f, e, n, driver, e_hist, moments, series, p->streams.use(stream));
}

int vpfp_group_create(vpfp_batch_t const* batches, int n_batches, vpfp_group_t* out) {
  if (!batches || !out) return fail(VPFP_EBADARG, "null argument");
  *out = nullptr;
  if (n_batches < 1 || n_batches > 64) return fail(VPFP_EBADARG, "n_batches must be in [1, 64]");
  auto* g = new vpfp_group_s();
  for (int k = 0; k < n_batches; ++k) {
    vpfp_batch_s* b = batches[k];
    if (!b) {
      delete g;
      return fail(VPFP_EBADARG, "null batch");
    }
    const vpfp_batch_s* b0 = batches[0];
    if (b->prec != b0->prec || b->op != b0->op || b->dt != b0->dt || b->delta != b0->delta) {
      delete g;
      return fail(VPFP_EUNSUPPORTED, "the batches of a group must share precision, operator, dt and delta-f storage");
    }
    if (b->closure) {
      delete g;
      return fail(VPFP_EUNSUPPORTED, "a batch with a closure cannot join a group");
    }
    g->b.push_back(b);
  }
  g->prec = batches[0]->prec;
  g->delta = batches[0]->delta;
  // groups in order of first appearance; defaults from the largest member
  for (int k = 0; k < n_batches; ++k) {
    const vpfp_batch_s* b = g->b[k];
    auto v = std::find_if(g->vg.begin(), g->vg.end(), [&](const vpfp_group_s::VGroup& q) { return q.nv == b->nv; });
    if (v == g->vg.end()) {
      g->vg.push_back({b->nv, {}, b->row_q});
      v = g->vg.end() - 1;
    }
    const vpfp_batch_s* m0 = v->members.empty() ? nullptr : g->b[v->members[0]];
    if (m0 && (m0->v0 != b->v0 || m0->dvv != b->dvv || m0->dk_v != b->dk_v)) {
      delete g;
      return fail(VPFP_EUNSUPPORTED, "batches of one nv must share the v grid");
    }
    v->members.push_back(k);
    auto x = std::find_if(g->xg.begin(), g->xg.end(), [&](const vpfp_group_s::XGroup& q) { return q.nx == b->nx; });
    if (x == g->xg.end()) {
      g->xg.push_back({b->nx, {}, b->C, b->dtiles, b->dft_mat ? b->dft : 0});
      x = g->xg.end() - 1;
    }
    x->members.push_back(k);
  }
  for (const auto& v : g->vg)
    if (v.members.size() > (size_t)vp::kMaxSeg) {
      delete g;
      return fail(VPFP_EUNSUPPORTED, "at most " + std::to_string(vp::kMaxSeg) + " batches per nv in a group");
    }
  for (auto& x : g->xg) {
    if (x.members.size() > (size_t)vp::kMaxSeg) {
      delete g;
      return fail(VPFP_EUNSUPPORTED, "at most " + std::to_string(vp::kMaxSeg) + " batches per nx in a group");
    }
    int minnv = 1 << 30;
    for (int k : x.members) minnv = std::min(minnv, g->b[k]->nv);
    while (x.C > minnv) x.C /= 2;
  }
  for (vpfp_batch_s* b : g->b) ++b->groups;
  *out = g;
  return VPFP_OK;
}

int vpfp_group_destroy(vpfp_group_t g) {
  if (!g) return VPFP_OK;
  g->streams.drain();
  for (vpfp_batch_s* b : g->b) --b->groups;
  delete g;
  return VPFP_OK;
}

int vpfp_group_step_n(vpfp_group_t g, void* const* f, double* const* e, int n_steps, double* const* series,
                      void* stream) {
  if (!g) return fail(VPFP_EBADARG, "null group");
  if (!f || !e) return fail(VPFP_EBADARG, "null array");
  for (size_t k = 0; k < g->b.size(); ++k)
    if (!f[k] || !e[k]) return fail(VPFP_EBADARG, "null array of batch " + std::to_string(k));
  if (n_steps < 0) return fail(VPFP_EBADARG, "negative step count");
  hipStream_t s = g->streams.use(stream);
  return g->prec == VPFP_F64 ? group_step_n<double>(g, f, e, n_steps, series, s)
                             : group_step_n<float>(g, f, e, n_steps, series, s);
}

int vpfp_group_autotune(vpfp_group_t g, int n_steps, void* stream, double* us_per_step) {
  if (!g) return fail(VPFP_EBADARG, "null group");
  if (n_steps < 1) return fail(VPFP_EBADARG, "n_steps must be positive");
  hipStream_t s = g->streams.use(stream);
  const int rc = g->prec == VPFP_F64 ? autotune_group<double>(g, n_steps, s) : autotune_group<float>(g, n_steps, s);
  if (rc) return rc;
  if (us_per_step) *us_per_step = g->tuned_us;
  return VPFP_OK;
}

int vpfp_group_get_shapes(vpfp_group_t g, int* row_q, int* drift_tile, int* drift_tiles, int* drift_dft) {
  if (!g) return fail(VPFP_EBADARG, "null group");
  if (!row_q || !drift_tile || !drift_tiles || !drift_dft) return fail(VPFP_EBADARG, "null array");
  for (size_t k = 0; k < g->b.size(); ++k) {
    for (const auto& v : g->vg)
      if (std::find(v.members.begin(), v.members.end(), (int)k) != v.members.end()) row_q[k] = v.q;
    for (const auto& x : g->xg)
      if (std::find(x.members.begin(), x.members.end(), (int)k) != x.members.end()) {
        drift_tile[k] = x.C;
        drift_tiles[k] = x.per;
        drift_dft[k] = x.dft;
      }
  }
  return VPFP_OK;
}

int vpfp_group_set_shapes(vpfp_group_t g, const int* row_q, const int* drift_tile, const int* drift_tiles,
                          const int* drift_dft) {
  if (!g) return fail(VPFP_EBADARG, "null group");
  if (!row_q || !drift_tile || !drift_tiles || !drift_dft) return fail(VPFP_EBADARG, "null array");
  for (const auto& v : g->vg) {
    const int q = row_q[v.members[0]];
    if (q != 0 && q != 4 && q != 8 && q != 16) return fail(VPFP_EBADARG, "row_q must be 0, 4, 8 or 16");
  }
  for (const auto& x : g->xg) {
    const int C = drift_tile[x.members[0]], per = drift_tiles[x.members[0]];
    if (C != 4 && C != 8 && C != 16 && C != 32) return fail(VPFP_EBADARG, "drift_tile must be 4, 8, 16 or 32");
    if (per < 1 || per > 64) return fail(VPFP_EBADARG, "drift_tiles must be in [1, 64]");
    const int dft = drift_dft[x.members[0]];
    if (dft != 0 && dft != 1) return fail(VPFP_EBADARG, "drift_dft must be 0 or 1");
    if (dft && !g->b[x.members[0]]->dft_mat) return fail(VPFP_EUNSUPPORTED, "the MFMA drift needs FP32 and nx 32 or 64");
    for (int k : x.members)
      if (C > g->b[k]->nv) return fail(VPFP_EBADARG, "drift_tile exceeds nv");
  }
  for (auto& v : g->vg) v.q = row_q[v.members[0]];
  for (auto& x : g->xg) {
    x.C = drift_tile[x.members[0]];
    x.per = drift_tiles[x.members[0]];
    x.dft = drift_dft[x.members[0]];
  }
  return VPFP_OK;
}

int vpfp_group_launch_times(vpfp_group_t g, int kind, int index, double* ms, long long* count) {
  if (!g) return fail(VPFP_EBADARG, "null group");
  if (kind < 0 || kind > 3 || index < 0 || !ms || !count) return fail(VPFP_EBADARG, "bad argument");
  const size_t id = (size_t)(kind + 4 * index);
  *ms = id < g->prof.ms_id.size() ? g->prof.ms_id[id] : 0.0;
  *count = id < g->prof.n_id.size() ? g->prof.n_id[id] : 0;
  return VPFP_OK;
}

int vpfp_group_set_profiling(vpfp_group_t g, int on) {
  if (!g) return fail(VPFP_EBADARG, "null group");
  g->prof.reset(on != 0);
  return VPFP_OK;
}

int vpfp_group_kernel_times(vpfp_group_t g, double* ms, long long* counts) {
  if (!g) return fail(VPFP_EBADARG, "null group");
  if (!ms || !counts) return fail(VPFP_EBADARG, "null array");
  for (int k = 0; k < 4; ++k) {
    ms[k] = g->prof.ms[k];
    counts[k] = g->prof.n[k];
  }
  return VPFP_OK;
}

int vpfp_step_n_scheme(vpfp_plan_t p, int scheme, void* f, double* e, int n, const double* driver, double* e_hist,
                       double* moments, double* series, void* f_store, int n_modes, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (scheme != VPFP_SCHEME_PEFRL && scheme != VPFP_SCHEME_SIXTH)
    return fail(VPFP_EBADARG, "scheme must be VPFP_SCHEME_PEFRL or VPFP_SCHEME_SIXTH");
  if (!f || !e) return fail(VPFP_EBADARG, "null array");
  if (n < 0) return fail(VPFP_EBADARG, "negative step count");
  if (n_modes < 0 || n_modes > p->nx) return fail(VPFP_EBADARG, "n_modes must be in [0, nx]");
  if (n_modes > 0 && !f_store) return fail(VPFP_EBADARG, "n_modes > 0 needs f_store");
  if (int rc = need_spectral_v(p)) return rc;
  if (int rc = need_spectral_x(p)) return rc;
  if (p->nx > 2048) return fail(VPFP_EUNSUPPORTED, "nx too large for the fused step");
  if (n_modes > 0 && !p->fcopy) {
    if (int rc = dalloc(p, &p->fcopy, (size_t)p->nx * p->nv * (p->prec == VPFP_F64 ? 8 : 4))) return rc;
  }
  return p->prec == VPFP_F64
             ? do_step_scheme<double>(p, scheme, f, e, n, driver, e_hist, moments, series, p->streams.use(stream),
                                      f_store, n_modes)
             : do_step_scheme<float>(p, scheme, f, e, n, driver, e_hist, moments, series, p->streams.use(stream),
                                     f_store, n_modes);
}

int vpfp_step_n_store(vpfp_plan_t p, void* f, double* e, int n, const double* driver, double* e_hist, double* moments,
                      double* series, void* f_store, int n_modes, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!f || !e || !f_store) return fail(VPFP_EBADARG, "null array");
  if (n < 0) return fail(VPFP_EBADARG, "negative step count");
  if (n_modes < 0 || n_modes > p->nx) return fail(VPFP_EBADARG, "n_modes must be in [0, nx]");
  if (int rc = need_spectral_v(p)) return rc;
  if (int rc = need_spectral_x(p)) return rc;
  if (p->nx > 2048) return fail(VPFP_EUNSUPPORTED, "nx too large for the fused step");
  if (n_modes > 0 && !p->fcopy) {
    if (int rc = dalloc(p, &p->fcopy, (size_t)p->nx * p->nv * (p->prec == VPFP_F64 ? 8 : 4))) return rc;
  }
  return p->prec == VPFP_F64
             ? do_step_n<double>(p, f, e, n, driver, e_hist, moments, series, p->streams.use(stream), f_store, n_modes)
             : do_step_n<float>(p, f, e, n, driver, e_hist, moments, series, p->streams.use(stream), f_store, n_modes);
}

}  // extern "C"
