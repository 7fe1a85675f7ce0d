// Drift kernels of the VPFP step (x-advection of column tiles with the
// partial trapz densities; vpfp_kernels.h drift_body) and their dispatch, for
// one storage type per object: compiled with -DVPFP_DRIFT_T=double|float.
//
// Launch shape: M = nx/2 complex points per column, Q = min(8, M) (4 for
// M <= 64), TPT = M/Q threads per column, C columns per workgroup.
#include <type_traits>

#include "vpfp_launch.h"

#ifndef VPFP_DRIFT_T
#error "compile with -DVPFP_DRIFT_T=double|float"
#endif

namespace vpfp_impl {
namespace {

// ---------------------------------------------------------------------------
// x-advection of FP32 columns as two dense real-DFT products on MFMA
// (nx = NX in {32, 64}).  Each wave takes a tile of 16 columns:
//   Y = F X      (F: [NX][NX] forward real DFT, X: [NX][16] the columns)
//   phase        (Re_k, Im_k) of column c times exp(-i k phi_c)
//   X' = G Y'    (G: [NX][NX] inverse real DFT)
// on v_mfma_f32_16x16x4_f32 (D = A B + C; A 16x4, B 4x16): lane l supplies
// A[l % 16][l / 16], B[l / 16][l % 16] and holds D[4 (l / 16) + r][l % 16],
// r < 4.  Rows of Y: Re_k for k <= NX/2 (row NX/2 the Nyquist bin), Im_k at
// row NX/2 + k (0 < k < NX/2), so a lane holds Re_k and Im_k of one column
// in tiles t and t + NX/32 at the same position.  The phase is that of the
// Stockham drift: phi = (dk dt) v_c, exp(-i k phi) for 0 < k < NX/2 and
// cos((NX/2) phi) on the Nyquist bin (the real part the reference keeps).
// The B operand of the second product is re-laid through a per-wave LDS
// tile.  W waves per workgroup; the workgroups stride over groups of W tiles
// (ntiles a multiple of W, so every wave makes the same number of passes).
// ---------------------------------------------------------------------------
template <int NX, int W>
__global__ __launch_bounds__(64 * W) void k_drift_dft(vp::DriftArgs<float> a) {
  extern __shared__ __align__(16) unsigned char lds[];
  VPFP_CTX;
  vp::drift_dft_body<NX, W>(c, lds, a);
}

template <int NX, int W>
__global__ __launch_bounds__(64 * W) void k_drift_multi_dft(vp::DriftMulti<float> m) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int s = vp::seg_of(m.seg, m.nseg, (int)blockIdx.x);
  const int local = (int)blockIdx.x - m.seg[s].blk0;
  const int gx = m.seg[s].gx;
  vp::DriftArgs<float> a = m.base;
  a.f_in = m.seg[s].f;
  a.f_out = m.seg[s].f;
  a.v = m.seg[s].v;
  a.dk_inst = m.seg[s].dk;
  a.dens = m.seg[s].dens;
  a.nv = m.seg[s].nv;
  GpuCtx c{(int)threadIdx.x, (int)blockDim.x, local % gx, gx, local / gx};
  vp::drift_dft_body<NX, W>(c, lds, a);
}

template <int NX, int W>
int launch_dft_one(const vp::DriftArgs<float>& a, hipStream_t s, dim3 grid) {
  constexpr int lds = vp::DftCfg<NX, W>::lds;
  auto kern = k_drift_dft<NX, W>;
  if (int rc = set_lds(kern, lds)) return rc;
  hipLaunchKernelGGL(kern, grid, dim3(64 * W), lds, s, a);
  return launched();
}

template <int NX, int W>
int launch_dft_multi_one(const vp::DriftMulti<float>& m, hipStream_t s, unsigned blocks) {
  constexpr int lds = vp::DftCfg<NX, W>::lds;
  auto kern = k_drift_multi_dft<NX, W>;
  if (int rc = set_lds(kern, lds)) return rc;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * W), lds, s, m);
  return launched();
}

// columns per workgroup pass C = 16 W (16 or 32)
template <class T>
int launch_drift_dft(const vp::DriftArgs<T>& a, int C, hipStream_t s, int gx, int gy) {
  if constexpr (!std::is_same<T, float>::value) {
    return fail(VPFP_EUNSUPPORTED, "the MFMA drift is FP32 only");
  } else {
    if ((C != 16 && C != 32) || (a.nv / 16) % (C / 16) != 0)
      return fail(VPFP_EUNSUPPORTED, "MFMA drift: 16 or 32 columns per pass dividing nv");
    if (a.fuse_field) return fail(VPFP_EUNSUPPORTED, "the MFMA drift has no field tail");
    const dim3 grid = gy > 0 ? dim3(gx, gy) : dim3(gx);
    if (a.nx == 32) return C == 16 ? launch_dft_one<32, 1>(a, s, grid) : launch_dft_one<32, 2>(a, s, grid);
    if (a.nx == 64) return C == 16 ? launch_dft_one<64, 1>(a, s, grid) : launch_dft_one<64, 2>(a, s, grid);
    return fail(VPFP_EUNSUPPORTED, "MFMA drift: nx must be 32 or 64");
  }
}

template <class T>
int launch_drift_multi_dft(vp::DriftMulti<T> m, int C, hipStream_t s) {
  if constexpr (!std::is_same<T, float>::value) {
    return fail(VPFP_EUNSUPPORTED, "the MFMA drift is FP32 only");
  } else {
    if (C != 16 && C != 32) return fail(VPFP_EUNSUPPORTED, "MFMA drift: 16 or 32 columns per pass");
    long long blocks = 0;
    for (int i = 0; i < m.nseg; ++i) {
      if ((m.seg[i].nv / 16) % (C / 16) != 0) return fail(VPFP_EUNSUPPORTED, "MFMA drift: pass width must divide nv");
      m.seg[i].blk0 = (int)blocks;
      blocks += (long long)m.seg[i].gx * m.seg[i].n_inst;
    }
    if (blocks > 0x7fffffffLL) return fail(VPFP_EUNSUPPORTED, "too many drift workgroups for one launch");
    const unsigned nb = (unsigned)blocks;
    if (m.base.nx == 32) return C == 16 ? launch_dft_multi_one<32, 1>(m, s, nb) : launch_dft_multi_one<32, 2>(m, s, nb);
    if (m.base.nx == 64) return C == 16 ? launch_dft_multi_one<64, 1>(m, s, nb) : launch_dft_multi_one<64, 2>(m, s, nb);
    return fail(VPFP_EUNSUPPORTED, "MFMA drift: nx must be 32 or 64");
  }
}


template <class T, int M, int Q, int C>
__global__ __launch_bounds__((vp::DriftCfg<T, M, Q, C>::wg)) void k_drift(vp::DriftArgs<T> a) {
  extern __shared__ __align__(16) unsigned char lds[];
  VPFP_CTX;
  vp::drift_body<T, M, Q, C>(c, lds, a);
}

template <class T, int M, int C>
int launch_drift_c(const vp::DriftArgs<T>& a, hipStream_t s, int gx, int gy) {
  constexpr int Q = vp::drift_q(M);
  using Cfg = vp::DriftCfg<T, M, Q, C>;
  auto kern = k_drift<T, M, Q, C>;
  if (Cfg::wg > 1024 || Cfg::lds > kMaxLds)
    return fail(VPFP_EUNSUPPORTED, "drift tile of " + std::to_string(C) + " columns does not fit a workgroup");
  if (a.fuse_field && !Cfg::tail) return fail(VPFP_EUNSUPPORTED, "field tail not available at this size");
  if (int rc = set_lds(kern, Cfg::lds)) return rc;
  const dim3 grid = gy > 0 ? dim3(gx, gy) : dim3(gx);
  hipLaunchKernelGGL(kern, grid, dim3(Cfg::wg), Cfg::lds, s, a);
  return launched();
}

template <class T, int M>
int launch_drift(const vp::DriftArgs<T>& a, int C, hipStream_t s, int gx, int gy) {
  switch (C) {
    case 4: return launch_drift_c<T, M, 4>(a, s, gx, gy);
    case 8: return launch_drift_c<T, M, 8>(a, s, gx, gy);
    case 16: return launch_drift_c<T, M, 16>(a, s, gx, gy);
    case 32: return launch_drift_c<T, M, 32>(a, s, gx, gy);
  }
  return fail(VPFP_EUNSUPPORTED, "drift tile width");
}

// several batches of one nx in one launch (vpfp_group_step_n): workgroup
// b of segment s advects column tiles of instance (b - blk0) / gx of that
// batch; tiles are taken in workgroup order (no XCD remapping: the
// workgroups of an instance are not a contiguous block of the XCD cycle)
template <class T, int M, int Q, int C>
__global__ __launch_bounds__((vp::DriftCfg<T, M, Q, C>::wg)) void k_drift_multi(vp::DriftMulti<T> m) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int s = vp::seg_of(m.seg, m.nseg, (int)blockIdx.x);
  const int local = (int)blockIdx.x - m.seg[s].blk0;
  const int gx = m.seg[s].gx;
  vp::DriftArgs<T> a = m.base;
  a.f_in = m.seg[s].f;
  a.f_out = m.seg[s].f;
  a.v = m.seg[s].v;
  a.dk_inst = m.seg[s].dk;
  a.dens = m.seg[s].dens;
  a.nv = m.seg[s].nv;
  a.xcd_order = 0;
  a.fuse_field = 0;
  GpuCtx c{(int)threadIdx.x, (int)blockDim.x, local % gx, gx, local / gx};
  vp::drift_body<T, M, Q, C>(c, lds, a);
}

template <class T, int M, int C>
int launch_drift_multi_c(vp::DriftMulti<T> m, hipStream_t s) {
  constexpr int Q = vp::drift_q(M);
  using Cfg = vp::DriftCfg<T, M, Q, C>;
  auto kern = k_drift_multi<T, M, Q, C>;
  if (Cfg::wg > 1024 || Cfg::lds > kMaxLds)
    return fail(VPFP_EUNSUPPORTED, "drift tile of " + std::to_string(C) + " columns does not fit a workgroup");
  for (int i = 0; i < m.nseg; ++i)
    if (m.seg[i].nv < C) return fail(VPFP_EUNSUPPORTED, "drift tile wider than nv");
  if (int rc = set_lds(kern, Cfg::lds)) return rc;
  long long blocks = 0;
  for (int i = 0; i < m.nseg; ++i) {
    m.seg[i].blk0 = (int)blocks;
    blocks += (long long)m.seg[i].gx * m.seg[i].n_inst;
  }
  if (blocks > 0x7fffffffLL) return fail(VPFP_EUNSUPPORTED, "too many drift workgroups for one launch");
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(Cfg::wg), Cfg::lds, s, m);
  return launched();
}

template <class T, int M>
int launch_drift_multi(const vp::DriftMulti<T>& m, int C, hipStream_t s) {
  switch (C) {
    case 4: return launch_drift_multi_c<T, M, 4>(m, s);
    case 8: return launch_drift_multi_c<T, M, 8>(m, s);
    case 16: return launch_drift_multi_c<T, M, 16>(m, s);
    case 32: return launch_drift_multi_c<T, M, 32>(m, s);
  }
  return fail(VPFP_EUNSUPPORTED, "drift tile width");
}

}  // namespace

int drift_multi_dispatch(const vp::DriftMulti<VPFP_DRIFT_T>& m, int C, hipStream_t s) {
  if (m.base.dft) return launch_drift_multi_dft(m, C, s);
  const int M = m.base.nx / 2;
#define CALL(mm) launch_drift_multi<VPFP_DRIFT_T, ((mm) <= 1024 ? (mm) : 2)>(m, C, s)
  VPFP_SWITCH_M(M, CALL, 1024)
#undef CALL
  return fail(VPFP_EUNSUPPORTED, "nx must be a power of two in [4, 2048] for the x-advection");
}

int drift_dispatch(const vp::DriftArgs<VPFP_DRIFT_T>& a, int C, int gx, hipStream_t s, int gy) {
  if (a.dft) return launch_drift_dft(a, C, s, gx, gy);
  const int M = a.nx / 2;
#define CALL(m) launch_drift<VPFP_DRIFT_T, ((m) <= 1024 ? (m) : 2)>(a, C, s, gx, gy)
  VPFP_SWITCH_M(M, CALL, 1024)
#undef CALL
  return fail(VPFP_EUNSUPPORTED, "nx must be a power of two in [4, 2048] for the x-advection");
}

}  // namespace vpfp_impl
