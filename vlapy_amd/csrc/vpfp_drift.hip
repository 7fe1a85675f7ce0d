// Drift kernels of the VPFP step (x-advection of column tiles with the
// partial trapz densities; vpfp_kernels.h drift_body) and their dispatch, for
// one storage type per object: compiled with -DVPFP_DRIFT_T=double|float.
//
// Launch shape: M = nx/2 complex points per column, Q = min(8, M) (4 for
// M <= 64), TPT = M/Q threads per column, C columns per workgroup.
#include "vpfp_launch.h"

#ifndef VPFP_DRIFT_T
#error "compile with -DVPFP_DRIFT_T=double|float"
#endif

namespace vpfp_impl {
namespace {

template <class T, int M, int Q, int C>
__global__ __launch_bounds__((vp::DriftCfg<T, M, Q, C>::wg)) void k_drift(vp::DriftArgs<T> a) {
  extern __shared__ __align__(16) unsigned char lds[];
  VPFP_CTX;
  vp::drift_body<T, M, Q, C>(c, lds, a);
}

template <class T, int M, int C>
int launch_drift_c(const vp::DriftArgs<T>& a, hipStream_t s, int gx, int gy) {
  constexpr int Q = M <= 64 ? (M < 4 ? M : 4) : 8;
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
  constexpr int Q = M <= 64 ? (M < 4 ? M : 4) : 8;
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
  const int M = m.base.nx / 2;
#define CALL(mm) launch_drift_multi<VPFP_DRIFT_T, ((mm) <= 1024 ? (mm) : 2)>(m, C, s)
  VPFP_SWITCH_M(M, CALL, 1024)
#undef CALL
  return fail(VPFP_EUNSUPPORTED, "nx must be a power of two in [4, 2048] for the x-advection");
}

int drift_dispatch(const vp::DriftArgs<VPFP_DRIFT_T>& a, int C, int gx, hipStream_t s, int gy) {
  const int M = a.nx / 2;
#define CALL(m) launch_drift<VPFP_DRIFT_T, ((m) <= 1024 ? (m) : 2)>(a, C, s, gx, gy)
  VPFP_SWITCH_M(M, CALL, 1024)
#undef CALL
  return fail(VPFP_EUNSUPPORTED, "nx must be a power of two in [4, 2048] for the x-advection");
}

}  // namespace vpfp_impl
