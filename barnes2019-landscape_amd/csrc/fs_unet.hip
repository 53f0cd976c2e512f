// fs_unet.hip -- the learned proposal's network: a 4-level convolutional
// U-Net in BF16 on the MFMA units (v_mfma_f32_16x16x32_bf16, FP32
// accumulation) with hand-written forward and backward kernels and Adam on
// FP32 master weights.  No reference counterpart: the network ranks candidate
// (K, U, m) parameter sets for an inverse problem so that only the best few
// go to the exact time-step.
//
// Network (input H x W, H and W multiples of 32):
//   input  8 channels: initial terrain, 3 parameter channels, 4 zero
//   enc L  conv3x3+ReLU, conv3x3+ReLU at widths 32, 64, 128, 160 for levels
//          0..3, 2x2 average pooling between levels
//   dec L  nearest 2x upsampling of level L+1, concatenated with level L's
//          encoder output, then conv3x3+ReLU twice (widths 128, 64, 32)
//   head   1x1 conv to one channel (FP32)
// Activations are NHWC BF16; weights are FP32 masters with two BF16 copies:
// [tap][out][in] for the forward conv and [flipped tap][in][out] for the
// data gradient, which is then the same implicit GEMM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "fs_internal.h"
#include "fs_hostutil.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kInChannels = FS_UNET_IN_CHANNELS;
constexpr int kLevels = 4;
constexpr int kWidth[kLevels] = {32, 64, 128, 160};

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

__device__ __forceinline__ f32x4 mfma16(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// 3x3 convolution, stride 1, zero padding, as an implicit GEMM:
//   out[p][n] = bias[n] + sum_{t, c} in[p + off_t][c] * wt[t][n][c]
// M = pixels (64 per workgroup, 16 per wave), N = NT*16 outputs per
// workgroup, K = 9 taps x cin in 32-channel steps.  MFMA operand A is the
// input patch (row = pixel lane%16, 8 channels per lane group lane/16),
// operand B the weights (column = output lane%16), D[4*(lane/16)+i][lane%16].
// Pixels at or beyond M (the last workgroup's tail) read zeros and store
// nothing.  Needs nout % (NT*16) == 0 and cin % 8 == 0.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t *__restrict__ in, int cin,
                                                 const uint16_t *__restrict__ wt,
                                                 const float *__restrict__ bias,
                                                 uint16_t *__restrict__ out, int nout, int H,
                                                 int W, int relu, long M) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const long m0 = static_cast<long>(blockIdx.x) * 64 + wave * 16;
  const int n0 = blockIdx.y * (NT * 16);
  const long p = m0 + r;
  const int x = static_cast<int>(p % W);
  const long q = p / W;
  const int y = static_cast<int>(q % H);
  const long b = q / H;
  f32x4 acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < 9; t++) {
    const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    const bool ok = p < M && yy >= 0 && yy < H && xx >= 0 && xx < W;
    const uint16_t *src = in + ((b * H + (ok ? yy : 0)) * W + (ok ? xx : 0)) * cin;
    const uint16_t *wrow = wt + (static_cast<long>(t) * nout + n0 + r) * cin;
    for (int c0 = 0; c0 < cin; c0 += 32) {
      const int c = c0 + 8 * g;
      const bool cok = c < cin;
      uint4 av = make_uint4(0u, 0u, 0u, 0u);
      if (ok && cok) av = *reinterpret_cast<const uint4 *>(src + c);
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        uint4 bv = make_uint4(0u, 0u, 0u, 0u);
        if (cok) bv = *reinterpret_cast<const uint4 *>(wrow + static_cast<long>(nt) * 16 * cin + c);
        acc[nt] = mfma16(av, bv, acc[nt]);
      }
    }
  }
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int n = n0 + nt * 16 + r;
    const float bb = bias ? bias[n] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float v = acc[nt][i] + bb;
      if (relu) v = fmaxf(v, 0.f);
      if (m0 + 4 * g + i < M) out[(m0 + 4 * g + i) * nout + n] = f2bf(v);
    }
  }
}

// ---------------------------------------------------------------------------
// The same convolution with operands staged in LDS: a workgroup computes an
// 8 x 16 pixel tile (one image) x NB outputs.  For each 32-channel step the
// 10 x 18 input patch (tile + one-pixel halo, zero padded) and the 9 taps'
// weights go to LDS once and every tap's MFMA operands are read from there,
// so each input value is fetched once per tile instead of once per tap and
// wave.  Wave w owns tile rows 2w and 2w+1 (two 16-pixel MFMA rows) and all
// NB outputs.  Rows are padded to 40 bf16 (80 bytes) against bank conflicts.
// Needs nout % NB == 0 and cin % 8 == 0.
// ---------------------------------------------------------------------------
constexpr int kTPH = 8, kTPW = 16;               // pixel tile
constexpr int kPatchH = kTPH + 2, kPatchW = kTPW + 2;
constexpr int kKC = 32;                           // input channels per step
constexpr int kCS = kKC + 8;                      // padded channel stride in LDS

template <int NB>
__global__ __launch_bounds__(256) void k_conv3x3_t(const uint16_t *__restrict__ in, int cin,
                                                   const uint16_t *__restrict__ wt,
                                                   const float *__restrict__ bias,
                                                   uint16_t *__restrict__ out, int nout, int H,
                                                   int W, int relu, int tiles_y) {
  constexpr int NT = NB / 16;
  __shared__ __attribute__((aligned(16))) uint16_t patch[kPatchH * kPatchW * kCS];
  __shared__ __attribute__((aligned(16))) uint16_t ws[9 * NB * kCS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int x0 = blockIdx.x * kTPW;
  const int ty = blockIdx.y % tiles_y;
  const long b = blockIdx.y / tiles_y;
  const int y0 = ty * kTPH;
  const int n0 = blockIdx.z * NB;
  f32x4 acc[2][NT];
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  // the next step's patch and weights are loaded into registers while the
  // current step is computed from LDS
  constexpr int kPV = (kPatchH * kPatchW * 4 + 255) / 256;
  constexpr int kWV = (9 * NB * 4 + 255) / 256;
  uint4 pv[kPV], wv[kWV];
  auto load = [&](int c0) {
#pragma unroll
    for (int k = 0; k < kPV; k++) {
      const int v = k * 256 + threadIdx.x;
      const int pix = v >> 2, c8 = v & 3;
      const int py = pix / kPatchW, px = pix % kPatchW;
      const int yy = y0 - 1 + py, xx = x0 - 1 + px;
      const int c = c0 + 8 * c8;
      pv[k] = make_uint4(0u, 0u, 0u, 0u);
      if (v < kPatchH * kPatchW * 4 && yy >= 0 && yy < H && xx >= 0 && xx < W && c < cin)
        pv[k] = *reinterpret_cast<const uint4 *>(in + ((b * H + yy) * W + xx) * cin + c);
    }
#pragma unroll
    for (int k = 0; k < kWV; k++) {
      const int v = k * 256 + threadIdx.x;
      const int row = v >> 2, c8 = v & 3;           // row = tap * NB + output
      const int t = row / NB, o = row % NB;
      const int c = c0 + 8 * c8;
      wv[k] = make_uint4(0u, 0u, 0u, 0u);
      if (v < 9 * NB * 4 && c < cin)
        wv[k] = *reinterpret_cast<const uint4 *>(wt + (static_cast<long>(t) * nout + n0 + o) * cin + c);
    }
  };
  load(0);
  for (int c0 = 0; c0 < cin; c0 += kKC) {
#pragma unroll
    for (int k = 0; k < kPV; k++) {
      const int v = k * 256 + threadIdx.x;
      if (v < kPatchH * kPatchW * 4)
        *reinterpret_cast<uint4 *>(&patch[(v >> 2) * kCS + 8 * (v & 3)]) = pv[k];
    }
#pragma unroll
    for (int k = 0; k < kWV; k++) {
      const int v = k * 256 + threadIdx.x;
      if (v < 9 * NB * 4) *reinterpret_cast<uint4 *>(&ws[(v >> 2) * kCS + 8 * (v & 3)]) = wv[k];
    }
    __syncthreads();
    if (c0 + kKC < cin) load(c0 + kKC);
#pragma unroll
    for (int t = 0; t < 9; t++) {
      const int dy = t / 3, dx = t % 3;
      uint4 av[2];
#pragma unroll
      for (int mt = 0; mt < 2; mt++) {
        const int pr = 2 * wave + mt;
        av[mt] = *reinterpret_cast<const uint4 *>(&patch[((pr + dy) * kPatchW + r + dx) * kCS + 8 * g]);
      }
#pragma unroll
      for (int nt = 0; nt < NT; nt++) {
        const uint4 bv = *reinterpret_cast<const uint4 *>(&ws[(t * NB + 16 * nt + r) * kCS + 8 * g]);
#pragma unroll
        for (int mt = 0; mt < 2; mt++) acc[mt][nt] = mfma16(av[mt], bv, acc[mt][nt]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int mt = 0; mt < 2; mt++) {
    const int yy = y0 + 2 * wave + mt;
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const int n = n0 + 16 * nt + r;
      const float bb = bias ? bias[n] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int xx = x0 + 4 * g + i;
        if (yy >= H || xx >= W) continue;
        float v = acc[mt][nt][i] + bb;
        if (relu) v = fmaxf(v, 0.f);
        out[((b * H + yy) * W + xx) * nout + n] = f2bf(v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient of a 3x3 convolution, split over pixel chunks:
//   part[chunk][t][co][ci] = sum_{p in chunk} dz[p][co] * x[p + off_t][ci]
// GEMM with M = co, N = ci (64 x 64 per workgroup, 32 x 32 per wave),
// K = pixels in steps of 32, both operands staged through LDS transposed so
// that each lane reads 8 consecutive pixels of one channel.
// ---------------------------------------------------------------------------
constexpr int kDwStride = 40;   // LDS row: 32 pixels + 8 pad (80 bytes)

__global__ __launch_bounds__(256) void k_conv3x3_dw(const uint16_t *__restrict__ dz, int cout,
                                                    const uint16_t *__restrict__ x, int cin,
                                                    float *__restrict__ part, int H, int W,
                                                    long chunk_pix, long M) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64 * kDwStride];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * kDwStride];
  const int tiles_ci = (cin + 63) / 64;
  const int co0 = (blockIdx.x / tiles_ci) * 64, ci0 = (blockIdx.x % tiles_ci) * 64;
  const int t = blockIdx.y, chunk = blockIdx.z;
  const int dy = t / 3 - 1, dx = t % 3 - 1;
  const long k0 = chunk * chunk_pix, k1 = min(M, k0 + chunk_pix);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int rb = (wave >> 1) * 32, cb = (wave & 1) * 32;
  // loader: pixel lp = tid/8, channels 8*(tid%8) .. +7
  const int lp = threadIdx.x >> 3, lc = (threadIdx.x & 7) * 8;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int c = 0; c < 2; c++) acc[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long kp = k0; kp < k1; kp += 32) {
    const long p = kp + lp;
    uint4 zv = make_uint4(0u, 0u, 0u, 0u), xv = make_uint4(0u, 0u, 0u, 0u);
    if (p < k1) {
      if (co0 + lc < cout) zv = *reinterpret_cast<const uint4 *>(dz + p * cout + co0 + lc);
      const int px = static_cast<int>(p % W);
      const long pq = p / W;
      const int py = static_cast<int>(pq % H);
      const long pb = pq / H;
      const int yy = py + dy, xx = px + dx;
      if (ci0 + lc < cin && yy >= 0 && yy < H && xx >= 0 && xx < W)
        xv = *reinterpret_cast<const uint4 *>(x + ((pb * H + yy) * W + xx) * cin + ci0 + lc);
    }
    const uint16_t *zs = reinterpret_cast<const uint16_t *>(&zv);
    const uint16_t *xs = reinterpret_cast<const uint16_t *>(&xv);
#pragma unroll
    for (int j = 0; j < 8; j++) {
      As[(lc + j) * kDwStride + lp] = zs[j];
      Bs[(lc + j) * kDwStride + lp] = xs[j];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const uint4 av = *reinterpret_cast<const uint4 *>(&As[(rb + a * 16 + r) * kDwStride + 8 * g]);
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const uint4 bv = *reinterpret_cast<const uint4 *>(&Bs[(cb + c * 16 + r) * kDwStride + 8 * g]);
        acc[a][c] = mfma16(av, bv, acc[a][c]);
      }
    }
    __syncthreads();
  }
  float *dst = part + (static_cast<long>(chunk) * 9 + t) * cout * cin;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int co = co0 + rb + a * 16 + 4 * g + i, ci = ci0 + cb + c * 16 + r;
        if (co < cout && ci < cin) dst[static_cast<long>(co) * cin + ci] = acc[a][c][i];
      }
}

// dw[i] = sum over chunks of part[chunk][i]
__global__ __launch_bounds__(256) void k_reduce_chunks(const float *__restrict__ part,
                                                       float *__restrict__ dw, long n,
                                                       int nchunks) {
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += gridDim.x * 256l) {
    float s = 0.f;
    for (int c = 0; c < nchunks; c++) s += part[c * n + i];
    dw[i] = s;
  }
}

// ---------------------------------------------------------------------------
// Weight gradient with all nine taps per workgroup:
//   part[chunk][t][co][ci] = sum over the chunk's pixel tiles of
//                            dz[p][co] * x[p + off_t][ci]
// A workgroup owns CB outputs x 32 inputs and a run of 8 x 16 pixel tiles of
// one image level; wave w handles the three taps of row dy = w - 1.  Per tile
// the dz tile goes to LDS as [output][pixel] and three column-shifted copies
// of the input patch (tile + one-pixel halo, zero padded) as
// [input][patch row][16 columns], both transposed in registers in 8 x 8
// blocks.  Every MFMA operand is then one aligned 16-byte LDS read: A = 8
// consecutive pixels of one output, B = the same 8 pixels of one input moved
// by the tap's offset.  With db, the workgroups of the first input block
// also add the bias gradient sum_p dz[p][co] of their pixels to db.  Needs
// H % 8 == 0, W % 16 == 0, cout % CB == 0, cin % 8 == 0.
// ---------------------------------------------------------------------------
constexpr int kDwTH = 8, kDwTW = 16, kDwTP = kDwTH * kDwTW;   // 128 pixels per tile
constexpr int kDwAS = kDwTP + 8;            // A row stride in LDS (pixels + pad)
constexpr int kDwBR = kDwTH + 2;            // patch rows
constexpr int kDwBS = kDwBR * kDwTW + 8;    // B row stride per input channel
constexpr int kDwThreads = 192;             // three waves: tap rows dy = -1, 0, 1

// o[j] = element j of each of v[0..7]: an 8 x 8 transpose of 16-bit values.
__device__ __forceinline__ void transpose8(const uint4 (&v)[8], uint4 (&o)[8]) {
  const uint32_t *w[8];
#pragma unroll
  for (int i = 0; i < 8; i++) w[i] = reinterpret_cast<const uint32_t *>(&v[i]);
#pragma unroll
  for (int j = 0; j < 8; j++) {
    uint32_t *d = reinterpret_cast<uint32_t *>(&o[j]);
    const uint32_t sel = (j & 1) ? 0x07060302u : 0x05040100u;   // high or low halves
#pragma unroll
    for (int k = 0; k < 4; k++) d[k] = __builtin_amdgcn_perm(w[2 * k + 1][j >> 1], w[2 * k][j >> 1], sel);
  }
}

template <int CB>
__global__ __launch_bounds__(kDwThreads) void k_conv3x3_dw_t(
    const uint16_t *__restrict__ dz, int cout, const uint16_t *__restrict__ x, int cin,
    float *__restrict__ part, float *__restrict__ db, int H, int W, int tiles_img, int tiles_x,
    long ntiles, int tiles_per_chunk, int ci_blocks) {
  constexpr int NA = CB / 16;
  constexpr int kABlocks = (kDwTP / 8) * (CB / 8);
  static_assert(kABlocks <= kDwThreads, "one dz block per thread");
  __shared__ __attribute__((aligned(16))) uint16_t As[CB * kDwAS];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[3 * 32 * kDwBS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int co0 = (blockIdx.x / ci_blocks) * CB, ci0 = (blockIdx.x % ci_blocks) * 32;
  const long t0 = static_cast<long>(blockIdx.y) * tiles_per_chunk;
  const long t1 = min(ntiles, t0 + tiles_per_chunk);
  f32x4 acc[3][NA][2];
#pragma unroll
  for (int dx = 0; dx < 3; dx++)
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[dx][a][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bool bias = db != nullptr && blockIdx.x % ci_blocks == 0;
  float bsum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};   // outputs of this thread's dz block
  for (long tile = t0; tile < t1; tile++) {
    const long b = tile / tiles_img;
    const int ti = static_cast<int>(tile % tiles_img);
    const int y0 = (ti / tiles_x) * kDwTH, x0 = (ti % tiles_x) * kDwTW;
    // dz: blocks of 8 pixels (half a tile row) x 8 outputs, one per thread
    if (static_cast<int>(threadIdx.x) < kABlocks) {
      const int blk = threadIdx.x;
      const int pb = blk % (kDwTP / 8), cg = blk / (kDwTP / 8);
      const int py = pb >> 1, px = (pb & 1) * 8;
      const uint16_t *src = dz + ((b * H + y0 + py) * W + x0 + px) * cout + co0 + 8 * cg;
      uint4 v[8], o[8];
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = *reinterpret_cast<const uint4 *>(src + static_cast<long>(i) * cout);
      if (bias) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const uint32_t *w = reinterpret_cast<const uint32_t *>(&v[i]);
#pragma unroll
          for (int k = 0; k < 4; k++) {
            bsum[2 * k] += __uint_as_float(w[k] << 16);
            bsum[2 * k + 1] += __uint_as_float(w[k] & 0xFFFF0000u);
          }
        }
      }
      transpose8(v, o);
#pragma unroll
      for (int j = 0; j < 8; j++)
        *reinterpret_cast<uint4 *>(&As[(8 * cg + j) * kDwAS + py * kDwTW + px]) = o[j];
    }
    // input patch: for each column shift dx, patch row, half row and group of
    // 8 inputs, 8 pixels x 8 inputs
    constexpr int kBBlocks = 3 * kDwBR * 2 * 4;
    for (int blk = threadIdx.x; blk < kBBlocks; blk += kDwThreads) {
      const int cg = blk & 3, h = (blk >> 2) & 1, pr = (blk >> 3) % kDwBR, dx = blk / (8 * kDwBR);
      const int yy = y0 - 1 + pr, xs = x0 - 1 + dx + 8 * h;
      const bool row_ok = yy >= 0 && yy < H && ci0 + 8 * cg < cin;
      const uint16_t *src = x + ((b * H + (row_ok ? yy : 0)) * W) * cin + ci0 + 8 * cg;
      uint4 v[8], o[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int xx = xs + i;
        v[i] = make_uint4(0u, 0u, 0u, 0u);
        if (row_ok && xx >= 0 && xx < W) v[i] = *reinterpret_cast<const uint4 *>(src + static_cast<long>(xx) * cin);
      }
      transpose8(v, o);
#pragma unroll
      for (int j = 0; j < 8; j++)
        *reinterpret_cast<uint4 *>(&Bs[(dx * 32 + 8 * cg + j) * kDwBS + pr * kDwTW + 8 * h]) = o[j];
    }
    __syncthreads();
    // k-steps of 32 pixels: tile rows 2s and 2s + 1
#pragma unroll
    for (int s = 0; s < kDwTH / 2; s++) {
      uint4 av[NA];
#pragma unroll
      for (int a = 0; a < NA; a++)
        av[a] = *reinterpret_cast<const uint4 *>(&As[(16 * a + r) * kDwAS + 32 * s + 8 * g]);
      const int prow = 2 * s + (g >> 1) + wave;   // patch row of this lane's pixels
#pragma unroll
      for (int dx = 0; dx < 3; dx++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          const uint4 bv = *reinterpret_cast<const uint4 *>(
              &Bs[(dx * 32 + 16 * j + r) * kDwBS + prow * kDwTW + 8 * (g & 1)]);
#pragma unroll
          for (int a = 0; a < NA; a++) acc[dx][a][j] = mfma16(av[a], bv, acc[dx][a][j]);
        }
    }
    __syncthreads();
  }
  if (bias) {
    // the 16 pixel blocks of each output group, summed in LDS (Bs is free:
    // the last tile ended with a barrier)
    float *red = reinterpret_cast<float *>(Bs);
    if (static_cast<int>(threadIdx.x) < kABlocks) {
      const int pb = threadIdx.x % (kDwTP / 8), cg = threadIdx.x / (kDwTP / 8);
#pragma unroll
      for (int e = 0; e < 8; e++) red[pb * CB + 8 * cg + e] = bsum[e];
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < CB) {
      float sum = 0.f;
      for (int pb = 0; pb < kDwTP / 8; pb++) sum += red[pb * CB + threadIdx.x];
      atomicAdd(&db[co0 + threadIdx.x], sum);
    }
  }
  float *dst = part + static_cast<long>(blockIdx.y) * 9 * cout * cin;
#pragma unroll
  for (int dx = 0; dx < 3; dx++) {
    const int t = 3 * wave + dx;
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int co = co0 + 16 * a + 4 * g + i, ci = ci0 + 16 * j + r;
          if (ci < cin) dst[(static_cast<long>(t) * cout + co) * cin + ci] = acc[dx][a][j][i];
        }
  }
}

// dw[i] = sum over chunks of part[chunk][i]; 64 elements x 16 chunk lanes per
// workgroup, the chunk lanes combined in LDS.
__global__ __launch_bounds__(1024) void k_reduce_chunks_wide(const float *__restrict__ part,
                                                             float *__restrict__ dw, long n,
                                                             int nchunks) {
  __shared__ float red[16][64];
  const int e = threadIdx.x & 63, cl = threadIdx.x >> 6;
  const long i = static_cast<long>(blockIdx.x) * 64 + e;
  float s = 0.f;
  if (i < n)
    for (int c = cl; c < nchunks; c += 16) s += part[static_cast<long>(c) * n + i];
  red[cl][e] = s;
  __syncthreads();
  if (cl == 0 && i < n) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 16; k++) t += red[k][e];
    dw[i] = t;
  }
}

// db[c] += sum_p dz[p][c] over the block's pixel range; thread t sums the
// 8-channel group t % C8 over every (256 / C8)-th pixel, C % 8 == 0, C <= 512
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *__restrict__ dz, int C,
                                                   float *__restrict__ db, long M,
                                                   long chunk_pix) {
  __shared__ float red[512];
  for (int c = threadIdx.x; c < C; c += 256) red[c] = 0.f;
  __syncthreads();
  const int C8 = C / 8, lanes = 256 / C8;
  const long p0 = blockIdx.x * chunk_pix, p1 = min(M, p0 + chunk_pix);
  if (static_cast<int>(threadIdx.x) < lanes * C8) {
    const int c8 = threadIdx.x % C8, pl = threadIdx.x / C8;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long p = p0 + pl; p < p1; p += lanes) {
      const uint4 v = *reinterpret_cast<const uint4 *>(dz + p * C + 8 * c8);
      const uint16_t *h = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
      for (int k = 0; k < 8; k++) s[k] += bf2f(h[k]);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) atomicAdd(&red[8 * c8 + k], s[k]);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += 256) atomicAdd(&db[c], red[c]);
}

// dz = dy where y > 0 (ReLU gradient from the stored output), 8 per thread
__global__ __launch_bounds__(256) void k_relu_grad(const uint16_t *__restrict__ dy,
                                                   const uint16_t *__restrict__ y,
                                                   uint16_t *__restrict__ dz, long n8) {
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n8; i += gridDim.x * 256l) {
    const uint4 g = reinterpret_cast<const uint4 *>(dy)[i];
    const uint4 v = reinterpret_cast<const uint4 *>(y)[i];
    const uint32_t gw[4] = {g.x, g.y, g.z, g.w}, vw[4] = {v.x, v.y, v.z, v.w};
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t lo = vw[k] & 0xFFFFu, hi = vw[k] >> 16;
      const bool plo = lo != 0 && !(lo & 0x8000u), phi = hi != 0 && !(hi & 0x8000u);
      o[k] = (plo ? (gw[k] & 0xFFFFu) : 0u) | (phi ? (gw[k] & 0xFFFF0000u) : 0u);
    }
    reinterpret_cast<uint4 *>(dz)[i] = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// 2x2 average pooling, NHWC, C % 8 == 0
__global__ __launch_bounds__(256) void k_pool(const uint16_t *__restrict__ in,
                                              uint16_t *__restrict__ out, int B, int H, int W,
                                              int C) {
  const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
  const long n = static_cast<long>(B) * Ho * Wo * C8;
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += gridDim.x * 256l) {
    const int c8 = static_cast<int>(i % C8);
    const long q = i / C8;
    const int xo = static_cast<int>(q % Wo);
    const long q2 = q / Wo;
    const int yo = static_cast<int>(q2 % Ho);
    const long b = q2 / Ho;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const long src = ((b * H + 2 * yo + d / 2) * W + 2 * xo + d % 2) * C + 8 * c8;
      const uint4 v = *reinterpret_cast<const uint4 *>(in + src);
      const uint16_t *h = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
      for (int k = 0; k < 8; k++) s[k] += bf2f(h[k]);
    }
    uint16_t o[8];
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = f2bf(0.25f * s[k]);
    *reinterpret_cast<uint4 *>(out + q * C + 8 * c8) = *reinterpret_cast<const uint4 *>(o);
  }
}

// gradient of the pooling: dx[b,y,x,c] (+)= dy[b,y/2,x/2,c] / 4
__global__ __launch_bounds__(256) void k_pool_grad(const uint16_t *__restrict__ dy,
                                                   uint16_t *__restrict__ dx, int B, int H,
                                                   int W, int C, int accumulate) {
  const int C8 = C / 8;
  const long n = static_cast<long>(B) * H * W * C8;
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += gridDim.x * 256l) {
    const int c8 = static_cast<int>(i % C8);
    const long q = i / C8;
    const int x = static_cast<int>(q % W);
    const long q2 = q / W;
    const int y = static_cast<int>(q2 % H);
    const long b = q2 / H;
    const long src = ((b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + 8 * c8;
    const uint4 v = *reinterpret_cast<const uint4 *>(dy + src);
    const uint16_t *h = reinterpret_cast<const uint16_t *>(&v);
    uint16_t *dst = dx + q * C + 8 * c8;
    uint16_t o[8];
    if (accumulate) {
      const uint4 w = *reinterpret_cast<const uint4 *>(dst);
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&w);
#pragma unroll
      for (int k = 0; k < 8; k++) o[k] = f2bf(bf2f(e[k]) + 0.25f * bf2f(h[k]));
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) o[k] = f2bf(0.25f * bf2f(h[k]));
    }
    *reinterpret_cast<uint4 *>(dst) = *reinterpret_cast<const uint4 *>(o);
  }
}

// out[b,y,x] = [lo[b,y/2,x/2, 0:c1], skip[b,y,x, 0:c2]]
__global__ __launch_bounds__(256) void k_upcat(const uint16_t *__restrict__ lo, int c1,
                                               const uint16_t *__restrict__ skip, int c2,
                                               uint16_t *__restrict__ out, int B, int H, int W) {
  const int C = c1 + c2, C8 = C / 8;
  const long n = static_cast<long>(B) * H * W * C8;
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += gridDim.x * 256l) {
    const int c = static_cast<int>(i % C8) * 8;
    const long q = i / C8;
    const int x = static_cast<int>(q % W);
    const long q2 = q / W;
    const int y = static_cast<int>(q2 % H);
    const long b = q2 / H;
    uint4 v;
    if (c < c1) v = *reinterpret_cast<const uint4 *>(lo + ((b * (H / 2) + y / 2) * (W / 2) + x / 2) * c1 + c);
    else v = *reinterpret_cast<const uint4 *>(skip + q * c2 + (c - c1));
    *reinterpret_cast<uint4 *>(out + q * C + c) = v;
  }
}

// gradient of k_upcat: dlo[b,Y,X,c] = sum of the 2x2 block's dout[.., c];
// dskip[b,y,x,c] = dout[b,y,x,c1+c]
__global__ __launch_bounds__(256) void k_upcat_grad(const uint16_t *__restrict__ dout, int c1,
                                                    int c2, uint16_t *__restrict__ dlo,
                                                    uint16_t *__restrict__ dskip, int B, int H,
                                                    int W) {
  const int C = c1 + c2;
  const int Hl = H / 2, Wl = W / 2;
  const long nlo = static_cast<long>(B) * Hl * Wl * (c1 / 8);
  const long nsk = static_cast<long>(B) * H * W * (c2 / 8);
  for (long i = blockIdx.x * 256l + threadIdx.x; i < nlo + nsk; i += gridDim.x * 256l) {
    if (i < nlo) {
      const int c = static_cast<int>(i % (c1 / 8)) * 8;
      const long q = i / (c1 / 8);
      const int X = static_cast<int>(q % Wl);
      const long q2 = q / Wl;
      const int Y = static_cast<int>(q2 % Hl);
      const long b = q2 / Hl;
      float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int d = 0; d < 4; d++) {
        const long src = ((b * H + 2 * Y + d / 2) * W + 2 * X + d % 2) * C + c;
        const uint4 v = *reinterpret_cast<const uint4 *>(dout + src);
        const uint16_t *h = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
        for (int k = 0; k < 8; k++) s[k] += bf2f(h[k]);
      }
      uint16_t o[8];
#pragma unroll
      for (int k = 0; k < 8; k++) o[k] = f2bf(s[k]);
      *reinterpret_cast<uint4 *>(dlo + q * c1 + c) = *reinterpret_cast<const uint4 *>(o);
    } else {
      const long j = i - nlo;
      const int c = static_cast<int>(j % (c2 / 8)) * 8;
      const long q = j / (c2 / 8);
      *reinterpret_cast<uint4 *>(dskip + q * c2 + c) =
          *reinterpret_cast<const uint4 *>(dout + q * C + c1 + c);
    }
  }
}

// network input: [terrain, p0, p1, p2, 0, 0, 0, 0] per pixel
// (terrain_stride 0: one terrain shared by the batch)
__global__ __launch_bounds__(256) void k_pack_input(const float *__restrict__ terrain,
                                                    const float *__restrict__ params,
                                                    uint16_t *__restrict__ x, int B, long HW,
                                                    long terrain_stride) {
  for (long i = blockIdx.x * 256l + threadIdx.x; i < B * HW; i += gridDim.x * 256l) {
    const long b = i / HW;
    uint16_t o[8] = {f2bf(terrain[b * terrain_stride + i % HW]), f2bf(params[3 * b]),
                     f2bf(params[3 * b + 1]),
                     f2bf(params[3 * b + 2]), 0, 0, 0, 0};
    *reinterpret_cast<uint4 *>(x + i * 8) = *reinterpret_cast<const uint4 *>(o);
  }
}

// sq[b] += sum over the interior (rings 0, 1 excluded) of (yhat - t)^2 for
// the batch's samples, against one target t of H x W
__global__ __launch_bounds__(256) void k_sq_misfit(const float *__restrict__ yhat,
                                                   const float *__restrict__ t,
                                                   float *__restrict__ sq, int H, int W) {
  __shared__ float red;
  if (threadIdx.x == 0) red = 0.f;
  __syncthreads();
  const long HW = static_cast<long>(H) * W;
  const int b = blockIdx.y;
  float s = 0.f;
  for (long i = blockIdx.x * 256l + threadIdx.x; i < HW; i += gridDim.x * 256l) {
    const int x = static_cast<int>(i % W), y = static_cast<int>(i / W);
    if (x < 2 || y < 2 || x >= W - 2 || y >= H - 2) continue;
    const float d = yhat[b * HW + i] - t[i];
    s += d * d;
  }
  atomicAdd(&red, s);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&sq[b], red);
}

// head: yhat[p] = hb + sum_c hw[c] * f[p][c], C = 32
__global__ __launch_bounds__(256) void k_head(const uint16_t *__restrict__ f,
                                              const float *__restrict__ hw,
                                              const float *__restrict__ hb,
                                              float *__restrict__ yhat, long M) {
  for (long p = blockIdx.x * 256l + threadIdx.x; p < M; p += gridDim.x * 256l) {
    float s = hb[0];
#pragma unroll
    for (int c8 = 0; c8 < 4; c8++) {
      const uint4 v = *reinterpret_cast<const uint4 *>(f + p * 32 + 8 * c8);
      const uint16_t *h = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
      for (int k = 0; k < 8; k++) s += hw[8 * c8 + k] * bf2f(h[k]);
    }
    yhat[p] = s;
  }
}

// loss = mean (yhat - t)^2 over M pixels; gy = 2 (yhat - t) / M; head
// gradients; df[p][c] = gy[p] * hw[c].  acc[0] += loss, acc[1..32] += dhw,
// acc[33] += dhb.
__global__ __launch_bounds__(256) void k_head_grad(const uint16_t *__restrict__ f,
                                                   const float *__restrict__ hw,
                                                   const float *__restrict__ yhat,
                                                   const float *__restrict__ tgt,
                                                   uint16_t *__restrict__ df,
                                                   float *__restrict__ acc, long M) {
  __shared__ float red[34];
  if (threadIdx.x < 34) red[threadIdx.x] = 0.f;
  __syncthreads();
  float l = 0.f, gb = 0.f, gw[32];
#pragma unroll
  for (int c = 0; c < 32; c++) gw[c] = 0.f;
  const float inv = 1.0f / static_cast<float>(M);
  for (long p = blockIdx.x * 256l + threadIdx.x; p < M; p += gridDim.x * 256l) {
    const float e = yhat[p] - tgt[p];
    l += e * e * inv;
    const float gy = 2.0f * e * inv;
    gb += gy;
#pragma unroll
    for (int c8 = 0; c8 < 4; c8++) {
      const uint4 v = *reinterpret_cast<const uint4 *>(f + p * 32 + 8 * c8);
      const uint16_t *h = reinterpret_cast<const uint16_t *>(&v);
      uint16_t o[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        gw[8 * c8 + k] += gy * bf2f(h[k]);
        o[k] = f2bf(gy * hw[8 * c8 + k]);
      }
      *reinterpret_cast<uint4 *>(df + p * 32 + 8 * c8) = *reinterpret_cast<const uint4 *>(o);
    }
  }
  atomicAdd(&red[0], l);
  atomicAdd(&red[33], gb);
#pragma unroll
  for (int c = 0; c < 32; c++) atomicAdd(&red[1 + c], gw[c]);
  __syncthreads();
  if (threadIdx.x < 34) atomicAdd(&acc[threadIdx.x], red[threadIdx.x]);
}

// Adam (Kingma & Ba 2015) on the FP32 masters
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v,
                                              long n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += gridDim.x * 256l) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// BF16 copies of one conv layer's weights W[t][co][ci]:
//   wf[t][co][ci] = W[t][co][ci];   wb[t][ci][co] = W[8-t][co][ci]
__global__ __launch_bounds__(256) void k_pack_weights(const float *__restrict__ w,
                                                      uint16_t *__restrict__ wf,
                                                      uint16_t *__restrict__ wb, int cout,
                                                      int cin) {
  const long n = 9l * cout * cin;
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += gridDim.x * 256l) {
    const int ci = static_cast<int>(i % cin);
    const long q = i / cin;
    const int co = static_cast<int>(q % cout);
    const int t = static_cast<int>(q / cout);
    const uint16_t v = f2bf(w[i]);
    wf[i] = v;
    wb[(static_cast<long>(8 - t) * cin + ci) * cout + co] = v;
  }
}

__global__ __launch_bounds__(256) void k_fill_f32(float *p, long n, float v) {
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += gridDim.x * 256l) p[i] = v;
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
namespace {

struct Conv {
  int cin, cout, level;        // level: resolution H>>level
  size_t woff, boff;           // offsets into the parameter vector
  uint16_t *wf = nullptr, *wb = nullptr;
};

struct Tensor {
  uint16_t *d = nullptr;   // device, max_batch x (H>>level) x (W>>level) x C
  int C = 0, level = 0;
};

}  // namespace

struct fs_unet {
  int H = 0, W = 0, max_batch = 0, device = 0;
  hipStream_t stream = nullptr;
  std::vector<Conv> conv;   // 14 layers
  size_t nparam = 0, head_w = 0, head_b = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // FP32 params, grads, Adam
  float *part = nullptr;
  size_t part_n = 0;
  // activations and their gradients
  Tensor x0, e_a[kLevels], e[kLevels], p[kLevels - 1], u[kLevels - 1], d_a[kLevels - 1],
      d[kLevels - 1];
  Tensor ge_a[kLevels], ge[kLevels], gp[kLevels - 1], gu[kLevels - 1], gd_a[kLevels - 1],
      gd[kLevels - 1], gz;   // gz: ReLU-masked gradient scratch (largest tensor)
  float *yhat = nullptr, *tgt = nullptr, *head_acc = nullptr, *terrain = nullptr,
        *params_in = nullptr;
  // resident training set
  float *data_terrain = nullptr, *data_params = nullptr, *data_tgt = nullptr;
  int data_n = 0;
  long adam_t = 0;
  double train_ms = 0, flops_per_sample = 0;
  uint64_t samples_trained = 0;
};

namespace {

int check_launch_u() {
  FS_HIP_TRY(hipGetLastError());
  return FS_OK;
}

int grid_for(long n) { return static_cast<int>(std::min<long>((n + 255) / 256, 4096)); }

int alloc_tensor(fs_unet *u, Tensor &t, int C, int level) {
  t.C = C;
  t.level = level;
  const size_t n = static_cast<size_t>(u->max_batch) * (u->H >> level) * (u->W >> level) * C;
  return fs::dev_alloc(&t.d, n);
}

long pixels(const fs_unet *u, int level, int B) {
  return static_cast<long>(B) * (u->H >> level) * (u->W >> level);
}

// out = conv3x3(in) with weights wt [tap][nout][cin]: the LDS-tiled kernel
// for images of at least one 8 x 16 tile, the direct one for the small
// deep levels.
int launch_conv(fs_unet *u, const uint16_t *in, int cin, const uint16_t *wt, const float *bias,
                uint16_t *out, int nout, int Hl, int Wl, int relu, int B) {
  const long M = static_cast<long>(B) * Hl * Wl;
  if (Hl >= kTPH && Wl >= kTPW) {
    const int tx = (Wl + kTPW - 1) / kTPW, ty = (Hl + kTPH - 1) / kTPH;
    if (nout % 64 == 0)
      hipLaunchKernelGGL(k_conv3x3_t<64>, dim3(tx, ty * B, nout / 64), dim3(256), 0, u->stream,
                         in, cin, wt, bias, out, nout, Hl, Wl, relu, ty);
    else
      hipLaunchKernelGGL(k_conv3x3_t<32>, dim3(tx, ty * B, nout / 32), dim3(256), 0, u->stream,
                         in, cin, wt, bias, out, nout, Hl, Wl, relu, ty);
  } else if (nout % 64 == 0) {
    hipLaunchKernelGGL(k_conv3x3<4>, dim3((M + 63) / 64, nout / 64), dim3(256), 0, u->stream, in,
                       cin, wt, bias, out, nout, Hl, Wl, relu, M);
  } else {
    hipLaunchKernelGGL(k_conv3x3<2>, dim3((M + 63) / 64, nout / 32), dim3(256), 0, u->stream, in,
                       cin, wt, bias, out, nout, Hl, Wl, relu, M);
  }
  return check_launch_u();
}

int conv_fwd(fs_unet *u, const Conv &c, const Tensor &in, Tensor &out, int B) {
  const int Hl = u->H >> c.level, Wl = u->W >> c.level;
  return launch_conv(u, in.d, c.cin, c.wf, u->w + c.boff, out.d, c.cout, Hl, Wl, 1, B);
}

// Work split of the nine-tap weight-gradient kernel: about kDwTargetGroups
// workgroups per layer, whole pixel tiles per chunk.  Levels whose size is
// not a multiple of the 8 x 16 tile use k_conv3x3_dw.
constexpr long kDwTargetGroups = 768;
struct DwPlan {
  bool tiled = false;
  int cb = 0, ci_blocks = 0, blocks = 0, tiles_x = 0, tiles_img = 0, tpc = 0, nchunks = 0;
  long ntiles = 0, max_chunks = 0;
};
DwPlan dw_plan(int cout, int cin, int Hl, int Wl, int B) {
  DwPlan p;
  p.tiled = Hl % kDwTH == 0 && Wl % kDwTW == 0 && cin % 8 == 0 && cout % 32 == 0;
  if (!p.tiled) return p;
  p.cb = cout % 64 == 0 ? 64 : 32;
  p.ci_blocks = (cin + 31) / 32;
  p.blocks = cout / p.cb * p.ci_blocks;
  p.tiles_x = Wl / kDwTW;
  p.tiles_img = (Hl / kDwTH) * p.tiles_x;
  p.ntiles = static_cast<long>(B) * p.tiles_img;
  const long want = (kDwTargetGroups + p.blocks - 1) / p.blocks;
  p.max_chunks = std::min(want, p.ntiles);
  p.tpc = static_cast<int>((p.ntiles + p.max_chunks - 1) / p.max_chunks);
  p.nchunks = static_cast<int>((p.ntiles + p.tpc - 1) / p.tpc);
  return p;
}

// Gradients of one conv+ReLU layer from gout (gradient of its output):
// weight and bias gradients, and the input gradient into gin (if given).
int conv_bwd(fs_unet *u, const Conv &c, const Tensor &in, const Tensor &out, const Tensor &gout,
             Tensor *gin, int B) {
  const int Hl = u->H >> c.level, Wl = u->W >> c.level;
  const long M = pixels(u, c.level, B);
  hipLaunchKernelGGL(k_relu_grad, dim3(grid_for(M * c.cout / 8)), dim3(256), 0, u->stream,
                     gout.d, out.d, u->gz.d, M * c.cout / 8);
  FS_TRY(check_launch_u());
  // weight gradient, split over pixel chunks
  const size_t wn = 9ul * c.cout * c.cin;
  const DwPlan dp = dw_plan(c.cout, c.cin, Hl, Wl, B);
  if (dp.tiled) {
    if (static_cast<size_t>(dp.nchunks) * wn > u->part_n) return FS_ERR_INTERNAL;
    // the bias gradient is summed by the same kernel
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, u->stream, u->g + c.boff,
                       static_cast<long>(c.cout), 0.f);
    if (dp.cb == 64)
      hipLaunchKernelGGL(k_conv3x3_dw_t<64>, dim3(dp.blocks, dp.nchunks), dim3(kDwThreads), 0,
                         u->stream, u->gz.d, c.cout, in.d, c.cin, u->part, u->g + c.boff, Hl, Wl,
                         dp.tiles_img, dp.tiles_x, dp.ntiles, dp.tpc, dp.ci_blocks);
    else
      hipLaunchKernelGGL(k_conv3x3_dw_t<32>, dim3(dp.blocks, dp.nchunks), dim3(kDwThreads), 0,
                         u->stream, u->gz.d, c.cout, in.d, c.cin, u->part, u->g + c.boff, Hl, Wl,
                         dp.tiles_img, dp.tiles_x, dp.ntiles, dp.tpc, dp.ci_blocks);
    FS_TRY(check_launch_u());
    hipLaunchKernelGGL(k_reduce_chunks_wide, dim3(static_cast<unsigned>((wn + 63) / 64)),
                       dim3(1024), 0, u->stream, u->part, u->g + c.woff, static_cast<long>(wn),
                       dp.nchunks);
    FS_TRY(check_launch_u());
  } else {
    const long chunk = std::max<long>(2048, (M + 63) / 64 / 32 * 32);
    const int nchunks = static_cast<int>((M + chunk - 1) / chunk);
    if (static_cast<size_t>(nchunks) * wn > u->part_n) return FS_ERR_INTERNAL;
    const int tiles = ((c.cout + 63) / 64) * ((c.cin + 63) / 64);
    hipLaunchKernelGGL(k_conv3x3_dw, dim3(tiles, 9, nchunks), dim3(256), 0, u->stream, u->gz.d,
                       c.cout, in.d, c.cin, u->part, Hl, Wl, chunk, M);
    FS_TRY(check_launch_u());
    hipLaunchKernelGGL(k_reduce_chunks, dim3(grid_for(static_cast<long>(wn))), dim3(256), 0,
                       u->stream, u->part, u->g + c.woff, static_cast<long>(wn), nchunks);
    FS_TRY(check_launch_u());
  }
  if (!dp.tiled) {
    const long bchunk = 4096;
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, u->stream, u->g + c.boff,
                       static_cast<long>(c.cout), 0.f);
    hipLaunchKernelGGL(k_bias_grad, dim3((M + bchunk - 1) / bchunk), dim3(256), 0, u->stream,
                       u->gz.d, c.cout, u->g + c.boff, M, bchunk);
    FS_TRY(check_launch_u());
  }
  if (gin) FS_TRY(launch_conv(u, u->gz.d, c.cout, c.wb, nullptr, gin->d, c.cin, Hl, Wl, 0, B));
  return FS_OK;
}

int pack_weights(fs_unet *u) {
  for (const Conv &c : u->conv) {
    hipLaunchKernelGGL(k_pack_weights, dim3(grid_for(9l * c.cout * c.cin)), dim3(256), 0,
                       u->stream, u->w + c.woff, c.wf, c.wb, c.cout, c.cin);
    FS_TRY(check_launch_u());
  }
  return FS_OK;
}

// conv layer indices: encoder 2L, 2L+1 (L = 0..3); decoder 8 + 2j, 9 + 2j for
// levels 2, 1, 0 (j = 0, 1, 2)
int enc(int L, int k) { return 2 * L + k; }
int dec(int L, int k) { return 8 + 2 * (2 - L) + k; }

int forward(fs_unet *u, const float *terrain_d, const float *params_d, int B,
            bool shared_terrain = false) {
  const long HW = static_cast<long>(u->H) * u->W;
  hipLaunchKernelGGL(k_pack_input, dim3(grid_for(B * HW)), dim3(256), 0, u->stream, terrain_d,
                     params_d, u->x0.d, B, HW, shared_terrain ? 0l : HW);
  FS_TRY(check_launch_u());
  const Tensor *in = &u->x0;
  for (int L = 0; L < kLevels; L++) {
    FS_TRY(conv_fwd(u, u->conv[enc(L, 0)], *in, u->e_a[L], B));
    FS_TRY(conv_fwd(u, u->conv[enc(L, 1)], u->e_a[L], u->e[L], B));
    if (L < kLevels - 1) {
      hipLaunchKernelGGL(k_pool, dim3(grid_for(pixels(u, L + 1, B) * u->e[L].C / 8)), dim3(256),
                         0, u->stream, u->e[L].d, u->p[L].d, B, u->H >> L, u->W >> L, u->e[L].C);
      FS_TRY(check_launch_u());
      in = &u->p[L];
    }
  }
  const Tensor *lo = &u->e[kLevels - 1];
  for (int L = kLevels - 2; L >= 0; L--) {
    hipLaunchKernelGGL(k_upcat, dim3(grid_for(pixels(u, L, B) * u->u[L].C / 8)), dim3(256), 0,
                       u->stream, lo->d, lo->C, u->e[L].d, u->e[L].C, u->u[L].d, B, u->H >> L,
                       u->W >> L);
    FS_TRY(check_launch_u());
    FS_TRY(conv_fwd(u, u->conv[dec(L, 0)], u->u[L], u->d_a[L], B));
    FS_TRY(conv_fwd(u, u->conv[dec(L, 1)], u->d_a[L], u->d[L], B));
    lo = &u->d[L];
  }
  const long M = pixels(u, 0, B);
  hipLaunchKernelGGL(k_head, dim3(grid_for(M)), dim3(256), 0, u->stream, u->d[0].d,
                     u->w + u->head_w, u->w + u->head_b, u->yhat, M);
  return check_launch_u();
}

// gradients of the loss mean((yhat - tgt)^2) w.r.t. every parameter into g;
// head_acc[0] receives the loss
int backward(fs_unet *u, int B) {
  const long M = pixels(u, 0, B);
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, u->stream, u->head_acc, 34l, 0.f);
  hipLaunchKernelGGL(k_head_grad, dim3(std::min<long>((M + 255) / 256, 1024)), dim3(256), 0,
                     u->stream, u->d[0].d, u->w + u->head_w, u->yhat, u->tgt, u->gd[0].d,
                     u->head_acc, M);
  FS_TRY(check_launch_u());
  FS_HIP_TRY(hipMemcpyAsync(u->g + u->head_w, u->head_acc + 1, 32 * sizeof(float),
                            hipMemcpyDeviceToDevice, u->stream));
  FS_HIP_TRY(hipMemcpyAsync(u->g + u->head_b, u->head_acc + 33, sizeof(float),
                            hipMemcpyDeviceToDevice, u->stream));
  // decoder, top level first
  for (int L = 0; L <= kLevels - 2; L++) {
    FS_TRY(conv_bwd(u, u->conv[dec(L, 1)], u->d_a[L], u->d[L], u->gd[L], &u->gd_a[L], B));
    FS_TRY(conv_bwd(u, u->conv[dec(L, 0)], u->u[L], u->d_a[L], u->gd_a[L], &u->gu[L], B));
    Tensor &glo = (L + 1 <= kLevels - 2) ? u->gd[L + 1] : u->ge[kLevels - 1];
    hipLaunchKernelGGL(k_upcat_grad, dim3(grid_for(pixels(u, L, B) * u->u[L].C / 8)), dim3(256),
                       0, u->stream, u->gu[L].d, glo.C, u->ge[L].C, glo.d, u->ge[L].d, B,
                       u->H >> L, u->W >> L);
    FS_TRY(check_launch_u());
  }
  // encoder, bottom level first; ge[L] holds the skip gradient, the pooled
  // path's gradient is added
  for (int L = kLevels - 1; L >= 0; L--) {
    if (L < kLevels - 1) {
      hipLaunchKernelGGL(k_pool_grad, dim3(grid_for(pixels(u, L, B) * u->ge[L].C / 8)),
                         dim3(256), 0, u->stream, u->gp[L].d, u->ge[L].d, B, u->H >> L,
                         u->W >> L, u->ge[L].C, 1);
      FS_TRY(check_launch_u());
    }
    FS_TRY(conv_bwd(u, u->conv[enc(L, 1)], u->e_a[L], u->e[L], u->ge[L], &u->ge_a[L], B));
    const Tensor &in = L == 0 ? u->x0 : u->p[L - 1];
    FS_TRY(conv_bwd(u, u->conv[enc(L, 0)], in, u->e_a[L], u->ge_a[L],
                    L == 0 ? nullptr : &u->gp[L - 1], B));
  }
  return FS_OK;
}

void free_unet(fs_unet *u) {
  if (!u) return;
  (void)hipSetDevice(u->device);
  if (u->stream) (void)hipStreamSynchronize(u->stream);
  std::vector<void *> ptrs = {u->w, u->g, u->m, u->v, u->part, u->yhat, u->tgt, u->head_acc,
                              u->terrain, u->params_in, u->data_terrain, u->data_params,
                              u->data_tgt, u->x0.d, u->gz.d};
  for (int L = 0; L < kLevels; L++) {
    ptrs.push_back(u->e_a[L].d);
    ptrs.push_back(u->e[L].d);
    ptrs.push_back(u->ge_a[L].d);
    ptrs.push_back(u->ge[L].d);
  }
  for (int L = 0; L < kLevels - 1; L++)
    for (Tensor *t : {&u->p[L], &u->u[L], &u->d_a[L], &u->d[L], &u->gp[L], &u->gu[L],
                      &u->gd_a[L], &u->gd[L]})
      ptrs.push_back(t->d);
  for (Conv &c : u->conv) {
    ptrs.push_back(c.wf);
    ptrs.push_back(c.wb);
  }
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (u->stream) (void)hipStreamDestroy(u->stream);
  delete u;
}

int check_unet(fs_unet *u) {
  if (!u) return FS_ERR_INVALID_ARGUMENT;
  FS_HIP_TRY(hipSetDevice(u->device));
  return FS_OK;
}

int upload_batch(fs_unet *u, const float *terrain, const float *params, const float *tgt,
                 int B) {
  const size_t HW = static_cast<size_t>(u->H) * u->W;
  FS_HIP_TRY(hipMemcpyAsync(u->terrain, terrain, B * HW * sizeof(float), hipMemcpyHostToDevice,
                            u->stream));
  FS_HIP_TRY(hipMemcpyAsync(u->params_in, params, B * 3 * sizeof(float), hipMemcpyHostToDevice,
                            u->stream));
  if (tgt)
    FS_HIP_TRY(hipMemcpyAsync(u->tgt, tgt, B * HW * sizeof(float), hipMemcpyHostToDevice,
                              u->stream));
  return FS_OK;
}

}  // namespace

extern "C" {

int fs_unet_create(int height, int width, int max_batch, uint64_t seed, int device,
                   fs_unet **out) {
  if (!out) return FS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (height < 32 || width < 32 || height % 32 || width % 32 || max_batch < 1 ||
      static_cast<long long>(height) * width * max_batch >= (1ll << 31))
    return FS_ERR_DIMENSIONS;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
    fs::set_last_error("no HIP device " + std::to_string(device));
    return FS_ERR_NO_DEVICE;
  }
  fs_unet *u = new (std::nothrow) fs_unet();
  if (!u) return FS_ERR_OUT_OF_MEMORY;
  struct Guard {
    fs_unet *u;
    ~Guard() { if (u) free_unet(u); }
  } guard{u};
  u->H = height;
  u->W = width;
  u->max_batch = max_batch;
  u->device = device;
  FS_HIP_TRY(hipSetDevice(device));
  FS_HIP_TRY(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
  // layers
  auto add = [&](int cin, int cout, int level) {
    Conv c;
    c.cin = cin;
    c.cout = cout;
    c.level = level;
    u->conv.push_back(c);
  };
  for (int L = 0; L < kLevels; L++) {
    add(L == 0 ? kInChannels : kWidth[L - 1], kWidth[L], L);
    add(kWidth[L], kWidth[L], L);
  }
  for (int L = kLevels - 2; L >= 0; L--) {
    const int below = L == kLevels - 2 ? kWidth[kLevels - 1] : kWidth[L + 1];
    add(below + kWidth[L], kWidth[L], L);
    add(kWidth[L], kWidth[L], L);
  }
  size_t off = 0;
  for (Conv &c : u->conv) {
    c.woff = off;
    off += 9ul * c.cout * c.cin;
    c.boff = off;
    off += c.cout;
  }
  u->head_w = off;
  off += kWidth[0];
  u->head_b = off;
  off += 1;
  u->nparam = off;
  // flops of one forward pass per sample: 2 per multiply-add
  double fl = 0;
  for (const Conv &c : u->conv)
    fl += 2.0 * 9 * c.cin * c.cout * (static_cast<double>(height >> c.level) * (width >> c.level));
  u->flops_per_sample = fl;

  FS_TRY(fs::dev_alloc(&u->w, off));
  FS_TRY(fs::dev_alloc(&u->g, off));
  FS_TRY(fs::dev_alloc(&u->m, off));
  FS_TRY(fs::dev_alloc(&u->v, off));
  for (Conv &c : u->conv) {
    FS_TRY(fs::dev_alloc(&c.wf, 9ul * c.cout * c.cin));
    FS_TRY(fs::dev_alloc(&c.wb, 9ul * c.cout * c.cin));
  }
  size_t part = 0;
  for (const Conv &c : u->conv) {
    const long M = pixels(u, c.level, max_batch);
    const long chunk = std::max<long>(2048, (M + 63) / 64 / 32 * 32);
    part = std::max(part, static_cast<size_t>((M + chunk - 1) / chunk) * 9 * c.cout * c.cin);
    const DwPlan dp = dw_plan(c.cout, c.cin, height >> c.level, width >> c.level, max_batch);
    if (dp.tiled) part = std::max(part, static_cast<size_t>(dp.max_chunks) * 9 * c.cout * c.cin);
  }
  u->part_n = part;
  FS_TRY(fs::dev_alloc(&u->part, part));
  FS_TRY(alloc_tensor(u, u->x0, kInChannels, 0));
  for (int L = 0; L < kLevels; L++) {
    FS_TRY(alloc_tensor(u, u->e_a[L], kWidth[L], L));
    FS_TRY(alloc_tensor(u, u->e[L], kWidth[L], L));
    FS_TRY(alloc_tensor(u, u->ge_a[L], kWidth[L], L));
    FS_TRY(alloc_tensor(u, u->ge[L], kWidth[L], L));
  }
  for (int L = 0; L < kLevels - 1; L++) {
    const int below = L == kLevels - 2 ? kWidth[kLevels - 1] : kWidth[L + 1];
    FS_TRY(alloc_tensor(u, u->p[L], kWidth[L], L + 1));
    FS_TRY(alloc_tensor(u, u->gp[L], kWidth[L], L + 1));
    FS_TRY(alloc_tensor(u, u->u[L], below + kWidth[L], L));
    FS_TRY(alloc_tensor(u, u->gu[L], below + kWidth[L], L));
    FS_TRY(alloc_tensor(u, u->d_a[L], kWidth[L], L));
    FS_TRY(alloc_tensor(u, u->gd_a[L], kWidth[L], L));
    FS_TRY(alloc_tensor(u, u->d[L], kWidth[L], L));
    FS_TRY(alloc_tensor(u, u->gd[L], kWidth[L], L));
  }
  // gz holds the largest conv output
  size_t gzn = 0;
  for (const Conv &c : u->conv)
    gzn = std::max(gzn, static_cast<size_t>(pixels(u, c.level, max_batch)) * c.cout);
  FS_TRY(fs::dev_alloc(&u->gz.d, gzn));
  const size_t HWB = static_cast<size_t>(max_batch) * height * width;
  FS_TRY(fs::dev_alloc(&u->yhat, HWB));
  FS_TRY(fs::dev_alloc(&u->tgt, HWB));
  FS_TRY(fs::dev_alloc(&u->terrain, HWB));
  FS_TRY(fs::dev_alloc(&u->params_in, 3ul * max_batch));
  FS_TRY(fs::dev_alloc(&u->head_acc, 64));
  // He-normal initial weights, zero biases (host RNG, reproducible per seed)
  std::vector<float> w(off, 0.f);
  std::mt19937_64 rng(seed);
  for (const Conv &c : u->conv) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.0f / (9.0f * c.cin)));
    for (size_t i = 0; i < 9ul * c.cout * c.cin; i++) w[c.woff + i] = nd(rng);
  }
  {
    std::normal_distribution<float> nd(0.f, std::sqrt(1.0f / kWidth[0]));
    for (int i = 0; i < kWidth[0]; i++) w[u->head_w + i] = nd(rng);
  }
  FS_HIP_TRY(hipMemcpy(u->w, w.data(), off * sizeof(float), hipMemcpyHostToDevice));
  FS_HIP_TRY(hipMemset(u->m, 0, off * sizeof(float)));
  FS_HIP_TRY(hipMemset(u->v, 0, off * sizeof(float)));
  FS_HIP_TRY(hipMemset(u->g, 0, off * sizeof(float)));
  FS_TRY(pack_weights(u));
  FS_HIP_TRY(hipStreamSynchronize(u->stream));
  guard.u = nullptr;
  *out = u;
  return FS_OK;
}

int fs_unet_destroy(fs_unet *u) {
  if (!u) return FS_ERR_INVALID_ARGUMENT;
  free_unet(u);
  return FS_OK;
}

int fs_unet_param_count(fs_unet *u, size_t *n) {
  if (!u || !n) return FS_ERR_INVALID_ARGUMENT;
  *n = u->nparam;
  return FS_OK;
}

int fs_unet_get_params(fs_unet *u, float *p, size_t n) {
  FS_TRY(check_unet(u));
  if (!p) return FS_ERR_INVALID_ARGUMENT;
  if (n != u->nparam) return FS_ERR_BUFFER_SIZE;
  FS_HIP_TRY(hipStreamSynchronize(u->stream));
  FS_HIP_TRY(hipMemcpy(p, u->w, n * sizeof(float), hipMemcpyDeviceToHost));
  return FS_OK;
}

int fs_unet_set_params(fs_unet *u, const float *p, size_t n) {
  FS_TRY(check_unet(u));
  if (!p) return FS_ERR_INVALID_ARGUMENT;
  if (n != u->nparam) return FS_ERR_BUFFER_SIZE;
  FS_HIP_TRY(hipStreamSynchronize(u->stream));
  FS_HIP_TRY(hipMemcpy(u->w, p, n * sizeof(float), hipMemcpyHostToDevice));
  FS_TRY(pack_weights(u));
  FS_HIP_TRY(hipStreamSynchronize(u->stream));
  return FS_OK;
}

int fs_unet_layer_shapes(fs_unet *u, int *cin, int *cout, int *level, int capacity, int *count) {
  if (!u || !count) return FS_ERR_INVALID_ARGUMENT;
  *count = static_cast<int>(u->conv.size());
  if (!cin && !cout && !level) return FS_OK;
  if (capacity < *count) return FS_ERR_BUFFER_SIZE;
  for (int i = 0; i < *count; i++) {
    if (cin) cin[i] = u->conv[i].cin;
    if (cout) cout[i] = u->conv[i].cout;
    if (level) level[i] = u->conv[i].level;
  }
  return FS_OK;
}

int fs_unet_forward(fs_unet *u, const float *terrain, const float *params, int batch, float *y) {
  FS_TRY(check_unet(u));
  if (!terrain || !params || !y) return FS_ERR_INVALID_ARGUMENT;
  const size_t HW = static_cast<size_t>(u->H) * u->W;
  for (int b0 = 0; b0 < batch; b0 += u->max_batch) {
    const int B = std::min(u->max_batch, batch - b0);
    FS_TRY(upload_batch(u, terrain + b0 * HW, params + 3ul * b0, nullptr, B));
    FS_TRY(forward(u, u->terrain, u->params_in, B));
    FS_HIP_TRY(hipMemcpyAsync(y + b0 * HW, u->yhat, B * HW * sizeof(float),
                              hipMemcpyDeviceToHost, u->stream));
  }
  FS_HIP_TRY(hipStreamSynchronize(u->stream));
  return FS_OK;
}

int fs_unet_loss_grad(fs_unet *u, const float *terrain, const float *params, const float *tgt,
                      int batch, float *loss, float *grad, size_t n) {
  FS_TRY(check_unet(u));
  if (!terrain || !params || !tgt || !loss) return FS_ERR_INVALID_ARGUMENT;
  if (batch < 1 || batch > u->max_batch) return FS_ERR_INVALID_ARGUMENT;
  if (grad && n != u->nparam) return FS_ERR_BUFFER_SIZE;
  FS_TRY(upload_batch(u, terrain, params, tgt, batch));
  FS_TRY(forward(u, u->terrain, u->params_in, batch));
  FS_TRY(backward(u, batch));
  FS_HIP_TRY(hipMemcpyAsync(loss, u->head_acc, sizeof(float), hipMemcpyDeviceToHost, u->stream));
  if (grad)
    FS_HIP_TRY(hipMemcpyAsync(grad, u->g, n * sizeof(float), hipMemcpyDeviceToHost, u->stream));
  FS_HIP_TRY(hipStreamSynchronize(u->stream));
  return FS_OK;
}

int fs_unet_set_data(fs_unet *u, const float *terrain, const float *params, const float *tgt,
                     int count) {
  FS_TRY(check_unet(u));
  if (!terrain || !params || !tgt || count < 1) return FS_ERR_INVALID_ARGUMENT;
  const size_t HW = static_cast<size_t>(u->H) * u->W;
  FS_HIP_TRY(hipStreamSynchronize(u->stream));
  for (float *p : {u->data_terrain, u->data_params, u->data_tgt})
    if (p) (void)hipFree(p);
  u->data_terrain = u->data_params = u->data_tgt = nullptr;
  u->data_n = 0;
  FS_TRY(fs::dev_alloc(&u->data_terrain, count * HW));
  FS_TRY(fs::dev_alloc(&u->data_params, 3ul * count));
  FS_TRY(fs::dev_alloc(&u->data_tgt, count * HW));
  FS_HIP_TRY(hipMemcpy(u->data_terrain, terrain, count * HW * sizeof(float), hipMemcpyHostToDevice));
  FS_HIP_TRY(hipMemcpy(u->data_params, params, 3ul * count * sizeof(float), hipMemcpyHostToDevice));
  FS_HIP_TRY(hipMemcpy(u->data_tgt, tgt, count * HW * sizeof(float), hipMemcpyHostToDevice));
  u->data_n = count;
  return FS_OK;
}

int fs_unet_train(fs_unet *u, int epochs, int batch, float lr, uint64_t seed, float *epoch_loss) {
  FS_TRY(check_unet(u));
  if (epochs < 1 || batch < 1 || batch > u->max_batch || !(lr > 0)) return FS_ERR_INVALID_ARGUMENT;
  if (u->data_n < batch) return FS_ERR_INVALID_ARGUMENT;
  const size_t HW = static_cast<size_t>(u->H) * u->W;
  const int per_epoch = u->data_n / batch;
  const long total = static_cast<long>(epochs) * per_epoch;
  std::mt19937_64 rng(seed);
  std::vector<int> order(u->data_n);
  for (int i = 0; i < u->data_n; i++) order[i] = i;
  std::vector<float> losses(per_epoch);
  float *loss_d = nullptr;
  FS_TRY(fs::dev_alloc(&loss_d, static_cast<size_t>(per_epoch)));
  struct Free {
    float *p;
    ~Free() { if (p) (void)hipFree(p); }
  } fr{loss_d};
  const auto t0 = std::chrono::steady_clock::now();
  long step = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(order.begin(), order.end(), rng);
    for (int s = 0; s < per_epoch; s++, step++) {
      // gather the minibatch device-to-device
      for (int b = 0; b < batch; b++) {
        const int i = order[s * batch + b];
        FS_HIP_TRY(hipMemcpyAsync(u->terrain + b * HW, u->data_terrain + i * HW, HW * sizeof(float),
                                  hipMemcpyDeviceToDevice, u->stream));
        FS_HIP_TRY(hipMemcpyAsync(u->tgt + b * HW, u->data_tgt + i * HW, HW * sizeof(float),
                                  hipMemcpyDeviceToDevice, u->stream));
        FS_HIP_TRY(hipMemcpyAsync(u->params_in + 3 * b, u->data_params + 3l * i,
                                  3 * sizeof(float), hipMemcpyDeviceToDevice, u->stream));
      }
      FS_TRY(forward(u, u->terrain, u->params_in, batch));
      FS_TRY(backward(u, batch));
      FS_HIP_TRY(hipMemcpyAsync(loss_d + s, u->head_acc, sizeof(float), hipMemcpyDeviceToDevice,
                                u->stream));
      // cosine learning-rate decay to lr/20
      const double frac = total > 1 ? static_cast<double>(step) / (total - 1) : 1.0;
      const float lr_t = static_cast<float>(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac))));
      u->adam_t++;
      const float c1 = 1.f - std::pow(0.9f, static_cast<float>(u->adam_t));
      const float c2 = 1.f - std::pow(0.999f, static_cast<float>(u->adam_t));
      hipLaunchKernelGGL(k_adam, dim3(grid_for(static_cast<long>(u->nparam))), dim3(256), 0,
                         u->stream, u->w, u->g, u->m, u->v, static_cast<long>(u->nparam), lr_t,
                         0.9f, 0.999f, 1e-8f, c1, c2);
      FS_TRY(check_launch_u());
      FS_TRY(pack_weights(u));
    }
    FS_HIP_TRY(hipMemcpyAsync(losses.data(), loss_d, per_epoch * sizeof(float),
                              hipMemcpyDeviceToHost, u->stream));
    FS_HIP_TRY(hipStreamSynchronize(u->stream));
    double s = 0;
    for (float l : losses) s += l;
    if (epoch_loss) epoch_loss[ep] = static_cast<float>(s / per_epoch);
  }
  u->train_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  u->samples_trained += static_cast<uint64_t>(total) * batch;
  return FS_OK;
}

int fs_unet_rank(fs_unet *u, const float *terrain, const float *params, int count,
                 const float *target, float *misfit) {
  FS_TRY(check_unet(u));
  if (!terrain || !params || !target || !misfit || count < 1) return FS_ERR_INVALID_ARGUMENT;
  const size_t HW = static_cast<size_t>(u->H) * u->W;
  float *p_d = nullptr, *sq_d = nullptr;
  FS_TRY(fs::dev_alloc(&p_d, 3ul * count));
  struct Free {
    float *a, *b;
    ~Free() {
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
    }
  } fr{p_d, nullptr};
  FS_TRY(fs::dev_alloc(&sq_d, static_cast<size_t>(count)));
  fr.b = sq_d;
  FS_HIP_TRY(hipMemcpyAsync(p_d, params, 3ul * count * sizeof(float), hipMemcpyHostToDevice,
                            u->stream));
  // terrain in u->terrain[0], target in u->tgt[0]
  FS_HIP_TRY(hipMemcpyAsync(u->terrain, terrain, HW * sizeof(float), hipMemcpyHostToDevice,
                            u->stream));
  FS_HIP_TRY(hipMemcpyAsync(u->tgt, target, HW * sizeof(float), hipMemcpyHostToDevice, u->stream));
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(count)), dim3(256), 0, u->stream, sq_d,
                     static_cast<long>(count), 0.f);
  for (int b0 = 0; b0 < count; b0 += u->max_batch) {
    const int B = std::min(u->max_batch, count - b0);
    FS_TRY(forward(u, u->terrain, p_d + 3l * b0, B, true));
    hipLaunchKernelGGL(k_sq_misfit, dim3(static_cast<unsigned>(std::min<size_t>((HW + 255) / 256, 64)), B),
                       dim3(256), 0, u->stream, u->yhat, u->tgt, sq_d + b0, u->H, u->W);
    FS_TRY(check_launch_u());
  }
  std::vector<float> sq(count);
  FS_HIP_TRY(hipMemcpyAsync(sq.data(), sq_d, count * sizeof(float), hipMemcpyDeviceToHost,
                            u->stream));
  FS_HIP_TRY(hipStreamSynchronize(u->stream));
  const double ninterior = static_cast<double>(u->H - 4) * (u->W - 4);
  for (int i = 0; i < count; i++) misfit[i] = static_cast<float>(std::sqrt(sq[i] / ninterior));
  return FS_OK;
}

int fs_unet_stats(fs_unet *u, double *train_ms, double *flops_per_sample, uint64_t *samples) {
  if (!u) return FS_ERR_INVALID_ARGUMENT;
  if (train_ms) *train_ms = u->train_ms;
  if (flops_per_sample) *flops_per_sample = u->flops_per_sample;
  if (samples) *samples = u->samples_trained;
  return FS_OK;
}

}  // extern "C"
