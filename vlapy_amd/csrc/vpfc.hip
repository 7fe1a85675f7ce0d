// Learned coarse-velocity closure: BF16 MFMA convolution network, forward,
// backward and Adam, for gfx950.  Interface and arithmetic: include/vpfc.h.
//
// Data layouts
//   rows of f, x, y, target, dy: [rows][256] FP32
//   stored activations a1..a3 and pre-activation gradients dz1..dz3:
//     [rows][64][256] BF16, channel-major (v contiguous)
//   LDS "position-major" tile of one row: [260][kP] BF16, row p+2 holds the
//     64 channels at position p (rows 0, 1, 258, 259 are the zero padding),
//     kP = 80 keeps every 8-channel group 16-byte aligned and the MFMA
//     operand reads (16 positions x 16 bytes per lane group) free of LDS
//     bank conflicts: with a 144-byte pitch (kP = 72) they measured 2-way
//     conflicted on gfx950, with 160 bytes conflict-free
//   LDS "channel-major" tile: [64][kQ] BF16, element p+2 of channel c at
//     c*kQ + p + 2 (zero padding at 0, 1, 258, 259)
//
// MFMA v_mfma_f32_16x16x32_bf16 (D = A B + C, A 16x32, B 32x16): lane l
// supplies A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16] for j < 8 and holds
// D[4(l/16)+j][l%16] for j < 4.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vpfc.h"
#include "vpfp.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define FC_HIP(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) return fail(VPFP_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kNV = VPFC_NV, kCH = VPFC_CHANNELS, kKS = VPFC_KERNEL, kK = kCH * kKS;
constexpr int kP = 80;          // position-major pitch (BF16)
constexpr int kQ = 272;         // channel-major pitch (BF16): 16-byte rows, reads up to p + 2 + 15 < 272
constexpr int kQs = 266;        // channel-major pitch for 2-byte column reads (133 dwords: odd)
constexpr int kRowsP = kNV + 4;
constexpr int oW1 = 0, oB1 = oW1 + kCH * kKS, oW2 = oB1 + kCH, oB2 = oW2 + kCH * kK, oW3 = oB2 + kCH,
              oB3 = oW3 + kCH * kK, oW4 = oB3 + kCH, oB4 = oW4 + kK, kParams = oB4 + 1;
static_assert(kParams == VPFC_NPARAMS, "parameter layout");
constexpr int kPartWG = 256;    // workgroups of the gradient kernels (rows strided over them)

__device__ __forceinline__ float bfr(float x) { return (float)(bf16)x; }

__device__ __forceinline__ floatx4 mfma(const bf16x8& a, const bf16x8& b, const floatx4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ----------------------------------------------------------------------------
// 64 -> 64 convolution of one row on MFMA
//   out[co][p] = sum_{k, ci} W[co][k*64 + ci] in[ci][p + k - 2]
// in: position-major LDS tile; A: this wave's fragments of W [64][320] BF16
// (load_a).  Wave w computes
// co in [16w, 16w + 16) for all 256 positions and hands acc[nt][j] =
// out[16w + 4(l/16) + j][16 nt + l%16] to `epi`.
// ----------------------------------------------------------------------------
// A fragments of this wave's 16 output channels for the 10 K-chunks
__device__ __forceinline__ void load_a(const bf16* W, bf16x8 (&A)[10]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bf16* wr = W + (16 * wave + (lane & 15)) * kK + 8 * (lane >> 4);
#pragma unroll
  for (int c = 0; c < 10; ++c) A[c] = *reinterpret_cast<const bf16x8*>(wr + 32 * c);
}

template <class Epi>
__device__ __forceinline__ void mconv(const bf16* in, const bf16x8 (&A)[10], Epi epi) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  floatx4 acc[16];
#pragma unroll
  for (int nt = 0; nt < 16; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 10; ++c) {
    const int tap = c >> 1;
    const int ci0 = (c & 1) * 32 + 8 * g;
#pragma unroll
    for (int nt = 0; nt < 16; ++nt) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(in + (16 * nt + m + tap) * kP + ci0);
      acc[nt] = mfma(A[c], b, acc[nt]);
    }
  }
  epi(acc, 16 * wave + 4 * g, m);
}

// zero the padding rows of a position-major tile
__device__ __forceinline__ void zero_pad_p(bf16* P) {
  for (int i = threadIdx.x; i < 4 * kP; i += blockDim.x) {
    const int r = i / kP, row = r < 2 ? r : kNV + r;
    P[row * kP + i % kP] = (bf16)0.f;
  }
}

// global channel-major BF16 row [64][256] -> position-major LDS tile
__device__ __forceinline__ void stage_p(bf16* P, const bf16* src) {
  const int p = threadIdx.x;  // 256 threads, one position each
#pragma unroll
  for (int c0 = 0; c0 < kCH; c0 += 8) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = src[(c0 + j) * kNV + p];
    *reinterpret_cast<bf16x8*>(P + (p + 2) * kP + c0) = v;
  }
}

// ----------------------------------------------------------------------------
// forward: one workgroup (256 threads, thread = position) per row, rows strided
//
// All four layers run on MFMA over one position-major LDS tile P:
//  * layer 1 (1 -> 64, K = 5): row p+2 of P keeps, in its pad columns 64..71,
//    the stencil record x[p-2..p+2] (BF16, zero outside the row, three zero
//    slots); wave w multiplies it by W1[16w + i][k] (A fragment with K < 5);
//  * layers 2 and 3 (64 -> 64): mconv, then in place over P once every wave
//    holds its sums (the records in columns 64..71 are not touched);
//  * layer 4 (64 -> 1, K = 320): W4 in every row of the A operand, so each
//    lane's sums are y at its own position; wave w takes positions
//    64w..64w+63 and hands y to the position's thread through LDS.
// The activations get tanh, rounded to BF16, as they leave the accumulators.
// ----------------------------------------------------------------------------
struct FwdArgs {
  const float* f;       // [rows][256]
  int rows;
  float in_scale, out_scale;
  const float* bg;      // [256] background profile, or null: the network then sees
                        // (f - bg) / s with s the row's rms of f - bg, and its output and
                        // target are in units of s
  const float* cons;    // [256 * 4 + 9] conservation projection (v, w, bg, unused, G^-1)
                        // applied to the correction, or null
  const float* p;       // FP32 parameters (W1, b1, b2, b3, W4, b4 are read here)
  const bf16* w2;       // [64][320] BF16 copies
  const bf16* w3;
  float* f_out;         // f + out_scale * y (may alias f) or null
  float* y_out;         // y or null
  // training (null otherwise)
  bf16* xin;            // [rows][256] L1 input
  bf16* a1;
  bf16* a2;
  bf16* a3;
  const float* target;  // [rows][256]
  float gscale;         // dy = gscale * (y - target)
  float* dy;            // [rows][256]
  float* loss_rows;     // [rows]: sum_v (y - target)^2
};

// the position-major tile (with the layer-1 stencil records in its pad
// columns) + reduction slots + BF16 W4 + layer-1 A rows + the layer-4
// outputs: 44352 bytes, so three workgroups share a CU's 160 KiB
constexpr int kFwdLds = kRowsP * kP * 2 + 64 + kK * 2 + kCH * 16 + kNV * 4;

// sum over the 256 threads of a workgroup (all threads get it); red: 4 floats
__device__ __forceinline__ float block_sum(float v, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// tanh for activations that are rounded to BF16 next, as x P(x^2) / Q(x^2)
// with P, Q of degree 4 in x^2 (coefficients fitted for the least maximum
// relative error on |x| <= 9 by vlapy_amd/csrc/tanh_fit.py) and x clamped to
// +-9 (tanh(9) = 1 - 3e-8).  FP32 evaluation: relative error <= 4e-7, so GPU
// and oracle activations seldom round to different BF16 values (a one-ulp
// flip moves the outputs it feeds by ~1e-3 of max |y|).  One reciprocal and
// no branch: 13 VALU instructions, against 14 with two transcendentals for
// sign(x) (1 - e^{-2|x|}) / (1 + e^{-2|x|}) with a Taylor series near 0.
__device__ __forceinline__ float tanh_act(float x) {
  const float c = __builtin_amdgcn_fmed3f(x, -9.f, 9.f);
  const float y = c * c;
  const float p =
      fmaf(fmaf(fmaf(fmaf(1.335543409e-08f, y, 2.060965016e-05f), y, 3.495623358e-03f), y, 1.338105500e-01f), y, 1.f);
  const float q =
      fmaf(fmaf(fmaf(fmaf(7.776877737e-07f, y, 3.285691782e-04f), y, 2.587711625e-02f), y, 4.671437144e-01f), y, 1.f);
  return (c * p) * __builtin_amdgcn_rcpf(q);
}

// tanh(acc + b) of accumulator tile nt (channels co0..co0+3 of lane (m, g),
// position 16 nt + m) into P at pr = P + (m + 2) kP + co0 and, when training,
// into the stored activations at gr = gout + (row kCH + co0) kNV + m
template <int NT>
__device__ __forceinline__ void act_tile(const floatx4& acc, const float (&b)[4], bf16* pr, bf16* gr) {
  bf16x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) o[j] = (bf16)tanh_act(acc[j] + b[j]);
  *reinterpret_cast<bf16x4*>(pr + 16 * NT * kP) = o;
  if (gr) {
#pragma unroll
    for (int j = 0; j < 4; ++j) gr[j * kNV + 16 * NT] = o[j];
  }
}

// act_tile over the 16 tiles of a 64 -> 64 layer
template <int NT = 0>
__device__ __forceinline__ void act_tiles(const floatx4* acc, const float (&b)[4], bf16* pr, bf16* gr) {
  if constexpr (NT < 16) {
    act_tile<NT>(acc[NT], b, pr, gr);
    act_tiles<NT + 1>(acc, b, pr, gr);
  }
}

// three waves per SIMD (<= 168 registers) to match the three workgroups the
// LDS allows; left to itself the compiler takes 260 registers (one wave per
// SIMD), which measured 2.4x slower at 262144 rows
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_fwd(FwdArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  bf16* P = reinterpret_cast<bf16*>(lds);
  float* red = reinterpret_cast<float*>(P + kRowsP * kP);
  bf16* w4b = reinterpret_cast<bf16*>(red + 16);  // BF16 W4, K order k*64 + ci
  bf16* w1b = w4b + kK;                            // [64][8]: W1[co][k], k < 5, then zeros
  float* ys = reinterpret_cast<float*>(w1b + 8 * kCH);  // layer-4 output by position
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, m = lane & 15, g = lane >> 4;
  bf16x8 z8;
#pragma unroll
  for (int j = 0; j < 8; ++j) z8[j] = (bf16)0.f;
  zero_pad_p(P);
  *reinterpret_cast<bf16x8*>(P + (t + 2) * kP + kCH) = z8;  // stencil records start at zero
  for (int i = t; i < kK; i += 256) w4b[i] = (bf16)a.p[oW4 + i];
  for (int i = t; i < 8 * kCH; i += 256) w1b[i] = (bf16)((i & 7) < kKS ? a.p[oW1 + (i >> 3) * kKS + (i & 7)] : 0.f);
  const int co0 = 16 * wave + 4 * g;
  bf16* const pr = P + (m + 2) * kP + co0;  // this lane's activation slots (tile 0)
  __syncthreads();  // records zeroed before any thread writes into them
  for (int row = blockIdx.x; row < a.rows; row += gridDim.x) {
    const size_t ro = (size_t)row * kNV;
    const float fv = a.f[ro + t];
    // row scale: rms of the deviation from the background (1 without one)
    float d = fv, srow = 1.f;
    if (a.bg) {
      d = fv - a.bg[t];
      srow = sqrtf(block_sum(d * d, red) * (1.f / kNV) + 1e-30f);
    }
    const bf16 xb = (bf16)((a.in_scale / srow) * d);
    if (a.xin) a.xin[ro + t] = xb;
    // x[t] is entry k of the record of position t + 2 - k
#pragma unroll
    for (int k = 0; k < kKS; ++k) {
      const int p = t + 2 - k;
      if (p >= 0 && p < kNV) P[(p + 2) * kP + kCH + k] = xb;
    }
    __syncthreads();
    // layer 1: 1 -> 64, tanh (writes columns 0..63 only, while the records are read)
    {
      // A[i][k] = W1[16 wave + i][k] for k < 5, zero beyond (lanes g > 0 read
      // the zero padding row of P)
      const bf16x8 A1 = *reinterpret_cast<const bf16x8*>(g == 0 ? w1b + 8 * (16 * wave + m) : P);
      const float b[4] = {a.p[oB1 + co0], a.p[oB1 + co0 + 1], a.p[oB1 + co0 + 2], a.p[oB1 + co0 + 3]};
      bf16* gr = a.a1 ? a.a1 + ((size_t)row * kCH + co0) * kNV + m : nullptr;
      const bf16* rec = P + (m + 2) * kP + kCH;
      floatx4 acc[16];
#pragma unroll
      for (int nt = 0; nt < 16; ++nt)
        acc[nt] = mfma(A1, *reinterpret_cast<const bf16x8*>(rec + 16 * nt * kP), floatx4{0.f, 0.f, 0.f, 0.f});
      act_tiles(acc, b, pr, gr);
    }
    __syncthreads();
    // layers 2 and 3: 64 -> 64, tanh, in place: every wave has its sums in
    // registers before any wave overwrites the tile.  The weight fragments are
    // reloaded per row (L2-resident): keeping both layers' in registers
    // measured slower (register pressure)
    {
      bf16x8 A[10];
      load_a(a.w2, A);
      mconv(P, A, [&](floatx4* acc, int c0, int mm) {
        const float b[4] = {a.p[oB2 + c0], a.p[oB2 + c0 + 1], a.p[oB2 + c0 + 2], a.p[oB2 + c0 + 3]};
        bf16* gr = a.a2 ? a.a2 + ((size_t)row * kCH + c0) * kNV + mm : nullptr;
        __syncthreads();
        act_tiles(acc, b, pr, gr);
      });
    }
    __syncthreads();
    {
      bf16x8 A[10];
      load_a(a.w3, A);
      mconv(P, A, [&](floatx4* acc, int c0, int mm) {
        const float b[4] = {a.p[oB3 + c0], a.p[oB3 + c0 + 1], a.p[oB3 + c0 + 2], a.p[oB3 + c0 + 3]};
        bf16* gr = a.a3 ? a.a3 + ((size_t)row * kCH + c0) * kNV + mm : nullptr;
        __syncthreads();
        act_tiles(acc, b, pr, gr);
      });
    }
    __syncthreads();
    // layer 4: 64 -> 1, linear; every row of A is W4, so lane (m, g) sums
    // y at position 16 nt + m in all four accumulator slots
    {
      floatx4 acc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 10; ++c) {
        const int tap = c >> 1, ci0 = (c & 1) * 32 + 8 * g;
        const bf16x8 a4 = *reinterpret_cast<const bf16x8*>(w4b + 32 * c + 8 * g);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const bf16x8 b = *reinterpret_cast<const bf16x8*>(P + (16 * (4 * wave + q) + m + tap) * kP + ci0);
          acc[q] = mfma(a4, b, acc[q]);
        }
      }
      if (g == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) ys[16 * (4 * wave + q) + m] = acc[q][0];
      }
    }
    __syncthreads();
    float y = a.p[oB4] + ys[t];
    if (a.cons) {
      // remove bg (c0 + c1 v + c2 v^2): trapz moments 1, v, v^2 of y vanish
      const float vv = a.cons[t], w = a.cons[kNV + t], b = a.cons[2 * kNV + t];
      const float m0 = block_sum(w * y, red);
      const float m1 = block_sum(w * vv * y, red);
      const float m2 = block_sum(w * vv * vv * y, red);
      const float* gi = a.cons + 4 * kNV;
      const float c0 = gi[0] * m0 + gi[1] * m1 + gi[2] * m2;
      const float c1 = gi[3] * m0 + gi[4] * m1 + gi[5] * m2;
      const float c2 = gi[6] * m0 + gi[7] * m1 + gi[8] * m2;
      y -= b * (c0 + vv * (c1 + vv * c2));
    }
    if (a.y_out) a.y_out[ro + t] = y;
    if (a.f_out) a.f_out[ro + t] = fv + (a.out_scale * srow) * y;
    if (a.dy) {
      const float e = y - a.target[ro + t] / srow;
      a.dy[ro + t] = a.gscale * e;
      const float s = block_sum(e * e, red);
      if (t == 0) a.loss_rows[row] = s;
    }
    __syncthreads();
  }
}
// ----------------------------------------------------------------------------
// backward through a 64 -> 64 layer (input gradient):
//   dzp[ci][p] = bf16( (sum_{k', co} Wt[ci][k'*64 + co] dz[co][p + k' - 2]) (1 - ap[ci][p]^2) )
// with Wt[ci][k'*64 + co] = W[co][(4 - k')*64 + ci] (flipped, transposed)
// ----------------------------------------------------------------------------
constexpr int kBwdLds = kRowsP * kP * 2;

__global__ __launch_bounds__(256) void k_bwd_data(const bf16* dz, const bf16* wt, const bf16* ap, bf16* dzp,
                                                   int rows) {
  extern __shared__ __align__(16) unsigned char lds[];
  bf16* P = reinterpret_cast<bf16*>(lds);
  zero_pad_p(P);
  bf16x8 A[10];
  load_a(wt, A);
  for (int row = blockIdx.x; row < rows; row += gridDim.x) {
    const size_t ro = (size_t)row * kCH * kNV;
    stage_p(P, dz + ro);
    __syncthreads();
    mconv(P, A, [&](floatx4* acc, int ci0, int m) {
#pragma unroll
      for (int nt = 0; nt < 16; ++nt) {
        const int pos = 16 * nt + m;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const size_t i = ro + (size_t)(ci0 + j) * kNV + pos;
          const float av = (float)ap[i];
          dzp[i] = (bf16)(acc[nt][j] * (1.f - av * av));
        }
      }
    });
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// weight gradient of a 64 -> 64 layer on MFMA (rows strided over the grid):
//   part[wg][oW + co*320 + k*64 + ci] = sum_{rows, p} dz[co][p] x[ci][p + k - 2]
//   part[wg][oB + co]                 = sum_{rows, p} dz[co][p]
// GEMM M = co, N = k*64 + ci (20 tiles of 16), K = positions (8 chunks of 32).
// ----------------------------------------------------------------------------
constexpr int kWgtLds = 2 * kCH * kQ * 2;

// 8 consecutive BF16 of an LDS row starting at element base + tap (base a
// multiple of 8, tap in [0, 4]): two aligned 16-byte reads and a byte shift
template <int TAP>
__device__ __forceinline__ bf16x8 load_shift(const bf16* row, int base) {
  const uint4 lo = *reinterpret_cast<const uint4*>(row + base);
  if constexpr (TAP == 0) {
    return *reinterpret_cast<const bf16x8*>(&lo);
  } else {
    const uint4 hi = *reinterpret_cast<const uint4*>(row + base + 8);
    const unsigned w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    unsigned r[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if constexpr ((TAP & 1) == 0) {
        r[q] = w[TAP / 2 + q];
      } else {
        r[q] = __builtin_amdgcn_alignbyte(w[(TAP + 1) / 2 + q], w[(TAP - 1) / 2 + q], 2);
      }
    }
    bf16x8 out;
    __builtin_memcpy(&out, r, 16);
    return out;
  }
}

__global__ __launch_bounds__(256) void k_wgrad(const bf16* dz, const bf16* x, int rows, float* part, int oW, int oB) {
  extern __shared__ __align__(16) unsigned char lds[];
  bf16* D = reinterpret_cast<bf16*>(lds);   // [64][kQ], element p at c*kQ + p
  bf16* X = D + kCH * kQ;                   // [64][kQ], element p at c*kQ + p + 2
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = lane & 15, g = lane >> 4;
  for (int i = t; i < kCH * (kQ - kNV); i += 256) {
    const int c = i / (kQ - kNV), r = i % (kQ - kNV);
    D[c * kQ + kNV + r] = (bf16)0.f;
    const int xr = r < 2 ? r : kNV + r;  // X padding: 0, 1 and 258 .. kQ-1
    X[c * kQ + xr] = (bf16)0.f;
  }
  floatx4 acc[20];
#pragma unroll
  for (int nt = 0; nt < 20; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
  float db = 0.f;
  for (int row = blockIdx.x; row < rows; row += gridDim.x) {
    const size_t ro = (size_t)row * kCH * kNV;
    // stage: 16 KB of each operand, 8 BF16 per thread per pass
    for (int i = t; i < kCH * kNV / 8; i += 256) {
      const int c = i / (kNV / 8), p = (i % (kNV / 8)) * 8;
      const bf16x8 dv = *reinterpret_cast<const bf16x8*>(dz + ro + c * kNV + p);
      const bf16x8 xv = *reinterpret_cast<const bf16x8*>(x + ro + c * kNV + p);
      *reinterpret_cast<bf16x8*>(D + c * kQ + p) = dv;
#pragma unroll
      for (int j = 0; j < 8; ++j) X[c * kQ + p + 2 + j] = xv[j];
    }
    __syncthreads();
    // bias: thread t sums channel t/4 over a quarter of the positions
    {
      const int c = t >> 2, q = t & 3;
      float s = 0.f;
      for (int p = q * 64; p < q * 64 + 64; ++p) s += (float)D[c * kQ + p];
      db += s;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int base = 32 * c + 8 * g;
      const bf16x8 A = *reinterpret_cast<const bf16x8*>(D + (16 * wave + m) * kQ + base);
#pragma unroll
      for (int nt = 0; nt < 20; ++nt) {
        const int tap = nt >> 2;
        const bf16* xr = X + (16 * (nt & 3) + m) * kQ;
        bf16x8 B;
        switch (tap) {
          case 0: B = load_shift<0>(xr, base); break;
          case 1: B = load_shift<1>(xr, base); break;
          case 2: B = load_shift<2>(xr, base); break;
          case 3: B = load_shift<3>(xr, base); break;
          default: B = load_shift<4>(xr, base); break;
        }
        acc[nt] = mfma(A, B, acc[nt]);
      }
    }
    __syncthreads();
  }
  float* out = part + (size_t)blockIdx.x * kParams;
#pragma unroll
  for (int nt = 0; nt < 20; ++nt) {
#pragma unroll
    for (int j = 0; j < 4; ++j) out[oW + (16 * wave + 4 * g + j) * kK + 16 * nt + m] = acc[nt][j];
  }
  db += __shfl_xor(db, 1, 64);
  db += __shfl_xor(db, 2, 64);
  if ((t & 3) == 0) out[oB + (t >> 2)] = db;
}

// ----------------------------------------------------------------------------
// output layer backward (64 -> 1):
//   dz3[ci][p] = bf16( (sum_k W4[k*64 + ci] bf(dy[p - k + 2])) (1 - a3[ci][p]^2) )
//   part[oW4 + k*64 + ci] = sum bf(dy[p]) a3[ci][p + k - 2],  part[oB4] = sum dy
// ----------------------------------------------------------------------------
constexpr int kOutLds = kCH * kQs * 2 + kRowsP * 4 + 64;

__global__ __launch_bounds__(256) void k_out_bwd(const float* dy, const bf16* a3, const float* p, bf16* dz3,
                                                 int rows, float* part) {
  extern __shared__ __align__(16) unsigned char lds[];
  bf16* X = reinterpret_cast<bf16*>(lds);                    // a3 [64][kQs] at p + 2
  float* ds = reinterpret_cast<float*>(lds + kCH * kQs * 2);  // bf(dy) at p + 2
  const int t = threadIdx.x;
  for (int i = t; i < kCH * 4; i += 256) {
    const int c = i >> 2, r = i & 3;
    X[c * kQs + (r < 2 ? r : kNV + r)] = (bf16)0.f;
  }
  if (t < 4) ds[t < 2 ? t : kNV + t] = 0.f;
  float gw[2] = {0.f, 0.f};  // outputs o = t and o = t + 256 (o < 320)
  float gb = 0.f;
  for (int row = blockIdx.x; row < rows; row += gridDim.x) {
    const size_t ro = (size_t)row * kNV;
    const size_t ra = (size_t)row * kCH * kNV;
    const float d = bfr(dy[ro + t]);
    ds[t + 2] = d;
    gb += dy[ro + t];
    for (int c = 0; c < kCH; ++c) X[c * kQs + t + 2] = a3[ra + (size_t)c * kNV + t];
    __syncthreads();
    // input gradient, thread = position
    for (int ci = 0; ci < kCH; ++ci) {
      float s = 0.f;
#pragma unroll
      for (int k = 0; k < kKS; ++k) s = fmaf(bfr(p[oW4 + k * kCH + ci]), ds[t - k + 4], s);
      const float av = (float)X[ci * kQs + t + 2];
      dz3[ra + (size_t)ci * kNV + t] = (bf16)(s * (1.f - av * av));
    }
    // weight gradient, thread = output
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int o = t + 256 * h;
      if (o < kK) {
        const int k = o / kCH, ci = o % kCH;
        const bf16* xr = X + ci * kQs + k;
        float s = 0.f;
        for (int q = 0; q < kNV; ++q) s = fmaf(ds[q + 2], (float)xr[q], s);
        gw[h] += s;
      }
    }
    __syncthreads();
  }
  float* out = part + (size_t)blockIdx.x * kParams;
  out[oW4 + t] = gw[0];
  if (t + 256 < kK) out[oW4 + t + 256] = gw[1];
  // bias: block sum of gb
  float s = gb;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ float red[4];
  if ((t & 63) == 0) red[t >> 6] = s;
  __syncthreads();
  if (t == 0) out[oB4] = (red[0] + red[1]) + (red[2] + red[3]);
}

// ----------------------------------------------------------------------------
// first layer weight gradient (1 -> 64):
//   part[oW1 + c*5 + k] = sum dz1[c][p] x[p + k - 2],  part[oB1 + c] = sum dz1[c][p]
// ----------------------------------------------------------------------------
constexpr int kInLds = kCH * kQs * 2 + kRowsP * 4 + 64;

__global__ __launch_bounds__(256) void k_in_wgrad(const bf16* dz1, const bf16* xin, int rows, float* part) {
  extern __shared__ __align__(16) unsigned char lds[];
  bf16* D = reinterpret_cast<bf16*>(lds);                     // dz1 [64][kQs]
  float* xs = reinterpret_cast<float*>(lds + kCH * kQs * 2);  // x at p + 2
  const int t = threadIdx.x;
  if (t < 4) xs[t < 2 ? t : kNV + t] = 0.f;
  // thread t: channel c = t / 4, outputs (k = 0..4, bias) over positions q*64 .. q*64+63
  const int c = t >> 2, q = t & 3;
  float gw[kKS + 1] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int row = blockIdx.x; row < rows; row += gridDim.x) {
    const size_t ra = (size_t)row * kCH * kNV;
    xs[t + 2] = (float)xin[(size_t)row * kNV + t];
    for (int cc = 0; cc < kCH; ++cc) D[cc * kQs + t] = dz1[ra + (size_t)cc * kNV + t];
    __syncthreads();
    for (int pp = q * 64; pp < q * 64 + 64; ++pp) {
      const float dv = (float)D[c * kQs + pp];
#pragma unroll
      for (int k = 0; k < kKS; ++k) gw[k] = fmaf(dv, xs[pp + k], gw[k]);
      gw[kKS] += dv;
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k <= kKS; ++k) {
    gw[k] += __shfl_xor(gw[k], 1, 64);
    gw[k] += __shfl_xor(gw[k], 2, 64);
  }
  if (q == 0) {
    float* out = part + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int k = 0; k < kKS; ++k) out[oW1 + c * kKS + k] = gw[k];
    out[oB1 + c] = gw[kKS];
  }
}

// grad[i] = sum over workgroups of part[wg][i]; loss = sum of loss_rows * lscale
__global__ __launch_bounds__(256) void k_reduce(const float* part, int nwg, float* grad, const float* loss_rows,
                                                int rows, float lscale, float* loss) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < kParams) {
    float s = 0.f;
    for (int w = 0; w < nwg; ++w) s += part[(size_t)w * kParams + i];
    grad[i] = s;
  }
  if (blockIdx.x == 0 && loss) {
    __shared__ float red[4];
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += 256) s += loss_rows[r];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *loss = ((red[0] + red[1]) + (red[2] + red[3])) * lscale;
  }
}

// BF16 copies of W2, W3 ([co][k*64+ci]) and their flipped transposes
// ([ci][k'*64+co] = W[co][(4-k')*64+ci])
__device__ __forceinline__ void write_copies(int i, float v, bf16* w2, bf16* w3, bf16* w2t, bf16* w3t) {
  if (i >= oW2 && i < oB2) {
    const int j = i - oW2, co = j / kK, kk = j % kK, k = kk / kCH, ci = kk % kCH;
    w2[j] = (bf16)v;
    w2t[ci * kK + (4 - k) * kCH + co] = (bf16)v;
  } else if (i >= oW3 && i < oB3) {
    const int j = i - oW3, co = j / kK, kk = j % kK, k = kk / kCH, ci = kk % kCH;
    w3[j] = (bf16)v;
    w3t[ci * kK + (4 - k) * kCH + co] = (bf16)v;
  }
}

__global__ __launch_bounds__(256) void k_refresh(const float* p, bf16* w2, bf16* w3, bf16* w2t, bf16* w3t) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < kParams) write_copies(i, p[i], w2, w3, w2t, w3t);
}

__global__ __launch_bounds__(256) void k_adam(float* p, float* mo, float* ve, const float* grad, float lr, float b1,
                                              float b2, float eps, float c1, float c2, bf16* w2, bf16* w3, bf16* w2t,
                                              bf16* w3t) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kParams) return;
  const float gi = grad[i];
  const float m = b1 * mo[i] + (1.f - b1) * gi;
  const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
  mo[i] = m;
  ve[i] = v;
  const float np = p[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  p[i] = np;
  write_copies(i, np, w2, w3, w2t, w3t);
}

}  // namespace

struct vpfc_net_s {
  int max_rows = 0;
  float in_scale = 1.f, out_scale = 1.f;
  float* bg = nullptr;       // [256] background profile (device)
  bool has_bg = false;
  float* cons = nullptr;     // [256 * 4 + 9] conservation projection (device)
  bool has_cons = false;
  int adam_t = 0;
  float* p = nullptr;
  float* mo = nullptr;
  float* ve = nullptr;
  float* grad = nullptr;
  float* part = nullptr;
  bf16 *w2 = nullptr, *w3 = nullptr, *w2t = nullptr, *w3t = nullptr;
  bf16 *xin = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *dz1 = nullptr, *dz2 = nullptr, *dz3 = nullptr;
  float* dy = nullptr;
  float* loss_rows = nullptr;
  float* loss = nullptr;
  std::vector<void*> allocs;
  ~vpfc_net_s() {
    for (void* q : allocs) (void)hipFree(q);
  }
};

namespace {

template <class P>
int nalloc(vpfc_net_s* n, P** out, size_t bytes) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes > 0 ? bytes : 16);
  if (e != hipSuccess) return fail(VPFP_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  n->allocs.push_back(q);
  *out = reinterpret_cast<P*>(q);
  return VPFP_OK;
}

int launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(VPFP_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return VPFP_OK;
}

int set_lds_all() {
  static bool done = false;
  if (done) return VPFP_OK;
  FC_HIP(hipFuncSetAttribute((const void*)k_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, kFwdLds));
  FC_HIP(hipFuncSetAttribute((const void*)k_wgrad, hipFuncAttributeMaxDynamicSharedMemorySize, kWgtLds));
  done = true;
  return VPFP_OK;
}

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// rows are strided over the workgroups: 3072 = 4 x the 3 x 256 that are
// resident at once (measured 5 % faster than 4096, which leaves a partial wave)
int fwd_grid(int rows) { return rows < 3072 ? rows : 3072; }

int check_net(vpfc_net_t n) { return n ? VPFP_OK : fail(VPFP_EBADARG, "null network"); }

FwdArgs fwd_args(vpfc_net_t n, const float* x, int rows) {
  FwdArgs a;
  std::memset(&a, 0, sizeof(a));
  a.f = x;
  a.rows = rows;
  a.in_scale = n->in_scale;
  a.out_scale = n->out_scale;
  a.p = n->p;
  a.w2 = n->w2;
  a.w3 = n->w3;
  a.bg = n->has_bg ? n->bg : nullptr;
  a.cons = n->has_cons ? n->cons : nullptr;
  return a;
}

int run_fwd(const FwdArgs& a, hipStream_t s) {
  if (a.rows == 0) return VPFP_OK;
  if (int rc = set_lds_all()) return rc;
  hipLaunchKernelGGL(k_fwd, dim3(fwd_grid(a.rows)), dim3(256), kFwdLds, s, a);
  return launched();
}

int gradients(vpfc_net_t n, const float* x, const float* target, int rows, float* grad, float* loss,
              hipStream_t s) {
  FwdArgs a = fwd_args(n, x, rows);
  a.cons = nullptr;  // the backward kernels differentiate the unprojected output
  a.xin = n->xin;
  a.a1 = n->a1;
  a.a2 = n->a2;
  a.a3 = n->a3;
  a.target = target;
  const float count = (float)rows * (float)kNV;
  a.gscale = 2.f / count;
  a.dy = n->dy;
  a.loss_rows = n->loss_rows;
  if (int rc = run_fwd(a, s)) return rc;
  const int G = kPartWG;
  hipLaunchKernelGGL(k_out_bwd, dim3(G), dim3(256), kOutLds, s, n->dy, n->a3, n->p, n->dz3, rows, n->part);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(k_wgrad, dim3(G), dim3(256), kWgtLds, s, n->dz3, n->a2, rows, n->part, oW3, oB3);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(k_bwd_data, dim3(fwd_grid(rows)), dim3(256), kBwdLds, s, n->dz3, n->w3t, n->a2, n->dz2, rows);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(k_wgrad, dim3(G), dim3(256), kWgtLds, s, n->dz2, n->a1, rows, n->part, oW2, oB2);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(k_bwd_data, dim3(fwd_grid(rows)), dim3(256), kBwdLds, s, n->dz2, n->w2t, n->a1, n->dz1, rows);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(k_in_wgrad, dim3(G), dim3(256), kInLds, s, n->dz1, n->xin, rows, n->part);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(k_reduce, dim3((kParams + 255) / 256), dim3(256), 0, s, n->part, G, grad, n->loss_rows, rows,
                     1.f / count, loss);
  return launched();
}

}  // namespace

extern "C" {

const char* vpfc_last_error(void) { return g_err.c_str(); }

int vpfc_create(int max_rows, vpfc_net_t* out) {
  if (!out) return fail(VPFP_EBADARG, "null argument");
  *out = nullptr;
  if (max_rows < 1 || max_rows > (1 << 20)) return fail(VPFP_EBADARG, "max_rows must be in [1, 2^20]");
  auto* n = new vpfc_net_s();
  n->max_rows = max_rows;
  const size_t R = (size_t)max_rows, act = R * kCH * kNV * sizeof(bf16);
  int rc = VPFP_OK;
  if (!rc) rc = nalloc(n, &n->p, sizeof(float) * kParams);
  if (!rc) rc = nalloc(n, &n->mo, sizeof(float) * kParams);
  if (!rc) rc = nalloc(n, &n->ve, sizeof(float) * kParams);
  if (!rc) rc = nalloc(n, &n->grad, sizeof(float) * kParams);
  if (!rc) rc = nalloc(n, &n->part, sizeof(float) * kParams * kPartWG);
  if (!rc) rc = nalloc(n, &n->w2, sizeof(bf16) * kCH * kK);
  if (!rc) rc = nalloc(n, &n->w3, sizeof(bf16) * kCH * kK);
  if (!rc) rc = nalloc(n, &n->w2t, sizeof(bf16) * kCH * kK);
  if (!rc) rc = nalloc(n, &n->w3t, sizeof(bf16) * kCH * kK);
  if (!rc) rc = nalloc(n, &n->xin, R * kNV * sizeof(bf16));
  if (!rc) rc = nalloc(n, &n->a1, act);
  if (!rc) rc = nalloc(n, &n->a2, act);
  if (!rc) rc = nalloc(n, &n->a3, act);
  if (!rc) rc = nalloc(n, &n->dz1, act);
  if (!rc) rc = nalloc(n, &n->dz2, act);
  if (!rc) rc = nalloc(n, &n->dz3, act);
  if (!rc) rc = nalloc(n, &n->dy, R * kNV * sizeof(float));
  if (!rc) rc = nalloc(n, &n->loss_rows, R * sizeof(float));
  if (!rc) rc = nalloc(n, &n->loss, sizeof(float));
  if (!rc) rc = nalloc(n, &n->bg, sizeof(float) * kNV);
  if (!rc) rc = nalloc(n, &n->cons, sizeof(float) * (4 * kNV + 9));
  if (!rc) {
    std::vector<float> z(kParams, 0.f);
    hipError_t e = hipMemcpy(n->p, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(n->mo, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(n->ve, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = fail(VPFP_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  if (!rc) {
    hipLaunchKernelGGL(k_refresh, dim3((kParams + 255) / 256), dim3(256), 0, 0, n->p, n->w2, n->w3, n->w2t, n->w3t);
    rc = launched();
    if (!rc) {
      hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) rc = fail(VPFP_EHIP, std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
    }
  }
  if (rc) {
    delete n;
    return rc;
  }
  *out = n;
  return VPFP_OK;
}

int vpfc_destroy(vpfc_net_t n) {
  if (!n) return VPFP_OK;
  (void)hipDeviceSynchronize();
  delete n;
  return VPFP_OK;
}

int vpfc_set_params(vpfc_net_t n, const float* host) {
  if (int rc = check_net(n)) return rc;
  if (!host) return fail(VPFP_EBADARG, "null parameters");
  std::vector<float> z(kParams, 0.f);
  FC_HIP(hipDeviceSynchronize());
  FC_HIP(hipMemcpy(n->p, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
  FC_HIP(hipMemcpy(n->mo, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  FC_HIP(hipMemcpy(n->ve, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  n->adam_t = 0;
  hipLaunchKernelGGL(k_refresh, dim3((kParams + 255) / 256), dim3(256), 0, 0, n->p, n->w2, n->w3, n->w2t, n->w3t);
  if (int rc = launched()) return rc;
  FC_HIP(hipDeviceSynchronize());
  return VPFP_OK;
}

int vpfc_get_params(vpfc_net_t n, float* host) {
  if (int rc = check_net(n)) return rc;
  if (!host) return fail(VPFP_EBADARG, "null parameters");
  FC_HIP(hipDeviceSynchronize());
  FC_HIP(hipMemcpy(host, n->p, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return VPFP_OK;
}

int vpfc_set_scales(vpfc_net_t n, float in_scale, float out_scale) {
  if (int rc = check_net(n)) return rc;
  if (!std::isfinite(in_scale) || !std::isfinite(out_scale)) return fail(VPFP_EBADARG, "non-finite scale");
  n->in_scale = in_scale;
  n->out_scale = out_scale;
  return VPFP_OK;
}

int vpfc_set_background(vpfc_net_t n, const float* host_bg) {
  if (int rc = check_net(n)) return rc;
  FC_HIP(hipDeviceSynchronize());
  if (host_bg) FC_HIP(hipMemcpy(n->bg, host_bg, sizeof(float) * kNV, hipMemcpyHostToDevice));
  n->has_bg = host_bg != nullptr;
  return VPFP_OK;
}

int vpfc_set_conservation(vpfc_net_t n, const double* v, const double* bg) {
  if (int rc = check_net(n)) return rc;
  FC_HIP(hipDeviceSynchronize());
  if (!v || !bg) {
    n->has_cons = false;
    return VPFP_OK;
  }
  // G[l][k] = sum_v w v^(l+k) bg(v) with trapz weights (1, ends 1/2)
  double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  std::vector<float> h(4 * kNV + 9, 0.f);
  for (int j = 0; j < kNV; ++j) {
    const double w = (j == 0 || j == kNV - 1) ? 0.5 : 1.0;
    double pw[5] = {1, v[j], v[j] * v[j], v[j] * v[j] * v[j], v[j] * v[j] * v[j] * v[j]};
    for (int l = 0; l < 3; ++l)
      for (int k = 0; k < 3; ++k) G[l][k] += w * pw[l + k] * bg[j];
    h[j] = (float)v[j];
    h[kNV + j] = (float)w;
    h[2 * kNV + j] = (float)bg[j];
  }
  // inverse of the symmetric 3x3 by cofactors
  const double c00 = G[1][1] * G[2][2] - G[1][2] * G[2][1], c01 = G[1][2] * G[2][0] - G[1][0] * G[2][2],
               c02 = G[1][0] * G[2][1] - G[1][1] * G[2][0];
  const double det = G[0][0] * c00 + G[0][1] * c01 + G[0][2] * c02;
  if (!(std::fabs(det) > 0.0)) return fail(VPFP_EBADARG, "singular moment matrix of the background");
  const double inv[9] = {c00 / det,
                         (G[0][2] * G[2][1] - G[0][1] * G[2][2]) / det,
                         (G[0][1] * G[1][2] - G[0][2] * G[1][1]) / det,
                         c01 / det,
                         (G[0][0] * G[2][2] - G[0][2] * G[2][0]) / det,
                         (G[0][2] * G[1][0] - G[0][0] * G[1][2]) / det,
                         c02 / det,
                         (G[0][1] * G[2][0] - G[0][0] * G[2][1]) / det,
                         (G[0][0] * G[1][1] - G[0][1] * G[1][0]) / det};
  for (int i = 0; i < 9; ++i) h[4 * kNV + i] = (float)inv[i];
  FC_HIP(hipMemcpy(n->cons, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
  n->has_cons = true;
  return VPFP_OK;
}

int vpfc_forward(vpfc_net_t n, const float* x, int rows, float* y, void* stream) {
  if (int rc = check_net(n)) return rc;
  if (rows < 0) return fail(VPFP_EBADARG, "negative row count");
  if (rows && (!x || !y)) return fail(VPFP_EBADARG, "null array");
  FwdArgs a = fwd_args(n, x, rows);
  a.y_out = y;
  return run_fwd(a, S(stream));
}

int vpfc_apply(vpfc_net_t n, float* f, int rows, void* stream) {
  if (int rc = check_net(n)) return rc;
  if (rows < 0) return fail(VPFP_EBADARG, "negative row count");
  if (rows && !f) return fail(VPFP_EBADARG, "null array");
  FwdArgs a = fwd_args(n, f, rows);
  a.f_out = f;
  return run_fwd(a, S(stream));
}

int vpfc_gradients(vpfc_net_t n, const float* x, const float* target, int rows, float* grad, float* loss,
                   void* stream) {
  if (int rc = check_net(n)) return rc;
  if (rows < 1 || rows > n->max_rows) return fail(VPFP_EBADARG, "rows must be in [1, max_rows]");
  if (!x || !target || !grad) return fail(VPFP_EBADARG, "null array");
  return gradients(n, x, target, rows, grad, loss, S(stream));
}

int vpfc_train_step(vpfc_net_t n, const float* x, const float* target, int rows, float lr, float beta1,
                    float beta2, float eps, float* loss, void* stream) {
  if (int rc = check_net(n)) return rc;
  if (rows < 1 || rows > n->max_rows) return fail(VPFP_EBADARG, "rows must be in [1, max_rows]");
  if (!x || !target) return fail(VPFP_EBADARG, "null array");
  if (!(beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps > 0.f))
    return fail(VPFP_EBADARG, "Adam needs 0 <= beta < 1 and eps > 0");
  const hipStream_t s = S(stream);
  if (int rc = gradients(n, x, target, rows, n->grad, loss ? loss : n->loss, s)) return rc;
  n->adam_t += 1;
  const float c1 = 1.f - std::pow(beta1, (float)n->adam_t), c2 = 1.f - std::pow(beta2, (float)n->adam_t);
  hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, n->p, n->mo, n->ve, n->grad, lr, beta1,
                     beta2, eps, c1, c2, n->w2, n->w3, n->w2t, n->w3t);
  return launched();
}

}  // extern "C"
