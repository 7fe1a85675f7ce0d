// Internal declarations shared by the translation units of libvpfp.so: the
// error state, the hardware block context the kernel bodies run on, launch
// helpers, and the row and drift dispatchers.  Each dispatcher is compiled in
// a source file of its own per storage type (vpfp_row.hip, vpfp_drift.hip),
// so that the template instantiations build in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "vpfp.h"
#include "vpfp_kernels.h"

namespace vpfp_impl {

// last error message of this thread (vpfp_last_error)
int fail(int code, const std::string& msg);

// VPFP_EHIP with the launch error, if the last launch failed
int launched();

struct GpuCtx {
  int tid, nthr, bid, nblk, bidy;
  __device__ __forceinline__ void sync() { __syncthreads(); }
  __device__ __forceinline__ void fence() { __threadfence(); }
  __device__ __forceinline__ unsigned arrive(unsigned* c) { return atomicAdd(c, 1u); }
  __device__ __forceinline__ void reset(unsigned* c) { __hip_atomic_store(c, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
  __device__ __forceinline__ double load_coherent(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ __forceinline__ double shfl_xor(double v, int m, int w) { return __shfl_xor(v, m, w); }
  __device__ __forceinline__ float shfl_xor(float v, int m, int w) { return __shfl_xor(v, m, w); }
  // v_mfma_f32_16x16x4_f32: D = A B + C (layout at vp::drift_dft_body)
  __device__ __forceinline__ vp::f4 mfma_16x16x4(float a, float b, vp::f4 acc) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    f4v cv = {acc.v[0], acc.v[1], acc.v[2], acc.v[3]};
    cv = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, cv, 0, 0, 0);
    return vp::f4{{cv[0], cv[1], cv[2], cv[3]}};
  }
};

#define VPFP_CTX \
  vpfp_impl::GpuCtx c{(int)threadIdx.x, (int)blockDim.x, (int)blockIdx.x, (int)gridDim.x, (int)blockIdx.y}

// LDS per workgroup on gfx950
constexpr int kMaxLds = 160 * 1024;

// dynamic LDS above 64 KiB must be enabled per kernel
template <class K>
int set_lds(K kern, int bytes) {
  if (bytes > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return fail(VPFP_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
  return VPFP_OK;
}

constexpr int qfor(int M) { return M < 8 ? M : 8; }

#define VPFP_SWITCH_M(M, CALL, MAXM)                 \
  switch (M) {                                       \
    case 2: return CALL(2);                          \
    case 4: return CALL(4);                          \
    case 8: return CALL(8);                          \
    case 16: return CALL(16);                        \
    case 32: return CALL(32);                        \
    case 64: return CALL(64);                        \
    case 128: return CALL(128);                      \
    case 256: return CALL(256);                      \
    case 512: return CALL(512);                      \
    case 1024: if (MAXM >= 1024) return CALL(1024);  \
      break;                                         \
    case 2048: if (MAXM >= 2048) return CALL(2048);  \
      break;                                         \
    case 4096: if (MAXM >= 4096) return CALL(4096);  \
      break;                                         \
  }

// Row kernels (vpfp_row.hip): q = values per thread (0: row-count rule).
// The plain kernels take f itself, the _bg kernels delta-f storage
// (a.bg and a.bg_hat set).
int row_dispatch_plain(const vp::RowArgs<double>& a, hipStream_t s, int q);
int row_dispatch_plain(const vp::RowArgs<float>& a, hipStream_t s, int q);
int row_dispatch_bg(const vp::RowArgs<double>& a, hipStream_t s, int q);
int row_dispatch_bg(const vp::RowArgs<float>& a, hipStream_t s, int q);

// The same over the segments of several batches of one nv (vpfp_group_*);
// the dispatch fills in the segments' first workgroups
int row_multi_dispatch_plain(const vp::RowMulti<double>& m, hipStream_t s, int q);
int row_multi_dispatch_plain(const vp::RowMulti<float>& m, hipStream_t s, int q);
int row_multi_dispatch_bg(const vp::RowMulti<double>& m, hipStream_t s, int q);
int row_multi_dispatch_bg(const vp::RowMulti<float>& m, hipStream_t s, int q);

// Drift kernels (vpfp_drift.hip): C columns per tile, gx workgroups striding
// over the column tiles; gy > 0: a batch of gy instances (grid y = instance)
int drift_dispatch(const vp::DriftArgs<double>& a, int C, int gx, hipStream_t s, int gy);
int drift_dispatch(const vp::DriftArgs<float>& a, int C, int gx, hipStream_t s, int gy);
// several batches of one nx (vpfp_group_*): segment s has n_inst * gx workgroups
int drift_multi_dispatch(const vp::DriftMulti<double>& m, int C, hipStream_t s);
int drift_multi_dispatch(const vp::DriftMulti<float>& m, int C, hipStream_t s);

}  // namespace vpfp_impl
