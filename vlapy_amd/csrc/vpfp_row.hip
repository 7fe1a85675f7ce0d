// Row kernels of the VPFP step (kick, collision, storage moments along v;
// vpfp_kernels.h row_body) and their dispatch, for one storage type and one
// row variant per object: compiled with -DVPFP_ROW_T=double|float and
// -DVPFP_ROW_BG=0 (f stored) | 1 (f - M stored, delta-f).
//
// Launch shape: M = nv/2 complex points per row, Q values per thread,
// TPT = M/Q threads per row, max(1, 64/TPT) rows per workgroup.
#include "vpfp_launch.h"

#if !defined(VPFP_ROW_T) || !defined(VPFP_ROW_BG)
#error "compile with -DVPFP_ROW_T=double|float -DVPFP_ROW_BG=0|1"
#endif

namespace vpfp_impl {
namespace {

template <class T, int M, int Q, bool BG>
__global__ __launch_bounds__((vp::RowCfg<T, M, Q>::wg), (vp::RowCfg<T, M, Q>::min_waves)) void k_row(vp::RowArgs<T> a) {
  extern __shared__ __align__(16) unsigned char lds[];
  VPFP_CTX;
  vp::row_body<T, M, Q, GpuCtx, BG>(c, lds, a);
}

template <class T, int M, int Q, bool BG>
int launch_row_q(const vp::RowArgs<T>& a, hipStream_t s) {
  using Cfg = vp::RowCfg<T, M, Q>;
  auto kern = k_row<T, M, Q, BG>;
  if (int rc = set_lds(kern, Cfg::lds)) return rc;
  const int grid = (a.nx + Cfg::rows - 1) / Cfg::rows;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(Cfg::wg), Cfg::lds, s, a);
  return launched();
}

// Q = 8 values per thread fills the GPU when there are many rows; with few
// rows (fewer than ~16k threads at Q = 8) Q = 4 spreads each row over twice
// as many threads and shortens the per-row latency chain.
// q: 4, 8 or 16 forces that many values per thread (where it exists: 4 for
// 16 <= M <= 2048, 16 for M >= 256; delta-f rows: 4 or 8, q = 16 runs at
// 8), 0 = the row-count rule above
template <class T, int M, bool BG>
int launch_row(const vp::RowArgs<T>& a, hipStream_t s, int q) {
  if constexpr (!BG && M >= 256 && M <= 4096) {
    if (q == 16) return launch_row_q<T, M, 16, BG>(a, s);
  }
  if constexpr (M >= 16 && M <= 2048) {
    if (q == 4 || (q == 0 && (long long)a.nx * (M / 8) < 16384)) return launch_row_q<T, M, 4, BG>(a, s);
  }
  return launch_row_q<T, M, qfor(M), BG>(a, s);
}

// several batches of one nv in one launch (vpfp_group_step_n): the
// workgroup's segment supplies f, e, nu dt and the row count
template <class T, int M, int Q, bool BG>
__global__ __launch_bounds__((vp::RowCfg<T, M, Q>::wg), (vp::RowCfg<T, M, Q>::min_waves)) void k_row_multi(
    vp::RowMulti<T> m) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int s = vp::seg_of(m.seg, m.nseg, (int)blockIdx.x);
  vp::RowArgs<T> a = m.base;
  a.f_in = m.seg[s].f;
  a.f_out = m.seg[s].f;
  a.e = m.seg[s].e;
  a.nu_dt_inst = m.seg[s].nu_dt;
  a.nx = m.seg[s].rows;
  a.rows_per_inst = m.seg[s].rows_per_inst;
  GpuCtx c{(int)threadIdx.x, (int)blockDim.x, (int)blockIdx.x - m.seg[s].blk0, 0, 0};
  vp::row_body<T, M, Q, GpuCtx, BG>(c, lds, a);
}

template <class T, int M, int Q, bool BG>
int launch_row_multi_q(vp::RowMulti<T> m, hipStream_t s) {
  using Cfg = vp::RowCfg<T, M, Q>;
  auto kern = k_row_multi<T, M, Q, BG>;
  if (int rc = set_lds(kern, Cfg::lds)) return rc;
  long long blocks = 0;
  for (int i = 0; i < m.nseg; ++i) {
    m.seg[i].blk0 = (int)blocks;
    blocks += (m.seg[i].rows + Cfg::rows - 1) / Cfg::rows;
  }
  if (blocks > 0x7fffffffLL) return fail(VPFP_EUNSUPPORTED, "too many rows for one launch");
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(Cfg::wg), Cfg::lds, s, m);
  return launched();
}

template <class T, int M, bool BG>
int launch_row_multi(const vp::RowMulti<T>& m, hipStream_t s, int q) {
  long long rows = 0;
  for (int i = 0; i < m.nseg; ++i) rows += m.seg[i].rows;
  if constexpr (!BG && M >= 256 && M <= 4096) {
    if (q == 16) return launch_row_multi_q<T, M, 16, BG>(m, s);
  }
  if constexpr (M >= 16 && M <= 2048) {
    if (q == 4 || (q == 0 && rows * (M / 8) < 16384)) return launch_row_multi_q<T, M, 4, BG>(m, s);
  }
  return launch_row_multi_q<T, M, qfor(M), BG>(m, s);
}

template <class T, bool BG>
int dispatch_multi(const vp::RowMulti<T>& m, hipStream_t s, int q) {
  const int M = m.base.nv / 2;
#define CALL(mm) launch_row_multi<T, ((mm) <= 4096 ? (mm) : 2), BG>(m, s, q)
  VPFP_SWITCH_M(M, CALL, 4096)
#undef CALL
  return fail(VPFP_EUNSUPPORTED, "nv must be a power of two in [4, 8192]");
}

template <class T, bool BG>
int dispatch(const vp::RowArgs<T>& a, hipStream_t s, int q) {
  const int M = a.nv / 2;
#define CALL(m) launch_row<T, ((m) <= 4096 ? (m) : 2), BG>(a, s, q)
  VPFP_SWITCH_M(M, CALL, 4096)
#undef CALL
  return fail(VPFP_EUNSUPPORTED, "nv must be a power of two in [4, 8192]");
}

}  // namespace

#if VPFP_ROW_BG
int row_dispatch_bg(const vp::RowArgs<VPFP_ROW_T>& a, hipStream_t s, int q) { return dispatch<VPFP_ROW_T, true>(a, s, q); }
int row_multi_dispatch_bg(const vp::RowMulti<VPFP_ROW_T>& m, hipStream_t s, int q) {
  return dispatch_multi<VPFP_ROW_T, true>(m, s, q);
}
#else
int row_dispatch_plain(const vp::RowArgs<VPFP_ROW_T>& a, hipStream_t s, int q) {
  return dispatch<VPFP_ROW_T, false>(a, s, q);
}
int row_multi_dispatch_plain(const vp::RowMulti<VPFP_ROW_T>& m, hipStream_t s, int q) {
  return dispatch_multi<VPFP_ROW_T, false>(m, s, q);
}
#endif

}  // namespace vpfp_impl
