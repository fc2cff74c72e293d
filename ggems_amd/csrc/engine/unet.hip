// 3x3 convolutions of the scatter surrogate's U-Net on the matrix cores:
// BF16 operands, FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Layouts: activations NHWC (channels fastest), weights [Cout][9][Cin]
// (tap = 3 * (dy + 1) + (dx + 1)), so that the GEMM's reduction index
// k = tap * Cin + ci runs over contiguous memory in both operands.  Cin and
// Cout are multiples of 8 and 16 (the first layer pads its two input
// channels); W is a multiple of 16 and N H W of 32.
//
// Forward (implicit GEMM, M = pixels, N = Cout, K = 9 Cin): a wave owns
// 32 consecutive pixels (two runs of 16 within image rows) and 64 output
// channels, i.e. 2 x 4 accumulator tiles of 16 x 16; each K step of 32 loads
// one 16-byte A fragment per pixel tile from memory and one B fragment per
// channel tile from the block's LDS copy of the weights, and issues 8 MFMAs.
// Bias and ReLU are applied in the epilogue.  The data gradient
// is the same kernel on the flipped, transposed weights.
//
// Weight gradient (M = Cout, N = 9 Cin, K = pixels): both operands have
// pixels as their strided index.  A block stages 32 pixels of dY and of the
// shifted input in LDS as they lie in memory (channels contiguous, 16-byte
// stores) and the waves read MFMA fragments with gfx950's transposing LDS
// read; blocks split the pixels and add their partial sums atomically in
// FP32.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../../include/ggems_engine.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

thread_local std::string g_unet_error;

int ufail(int code, const std::string& msg) {
  g_unet_error = msg;
  return code;
}

constexpr int kWaves = 4;
constexpr int kFwdPix = 32;  // pixels per wave
constexpr int kFwdCh = 64;   // output channels per wave

__device__ inline bf16x8 zero8() {
  bf16x8 z;
  for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
  return z;
}

// A fragment: 8 consecutive k of pixel (y, x) for reduction index k0
// (k0 multiple of 8; all 8 lie in one tap because Cin is a multiple of 8)
__device__ inline bf16x8 load_a(const __bf16* x, int img, int y, int xx, int k0, int K, int Cin, int H, int W) {
  if (k0 >= K) return zero8();
  const int tap = k0 / Cin, ci = k0 - tap * Cin;
  const int yy = y + tap / 3 - 1, xs = xx + tap % 3 - 1;
  if (yy < 0 || yy >= H || xs < 0 || xs >= W) return zero8();
  return *reinterpret_cast<const bf16x8*>(x + (((size_t)img * H + yy) * W + xs) * Cin + ci);
}

// The blocks go to the eight XCDs in turn (block b to XCD b % 8), and each
// XCD has its own L2.  xcd_tile() numbers the tiles so that each XCD takes a
// contiguous run of them: neighbouring pixel tiles, which read the same
// image rows through the 3x3 stencil, then share one L2.  Blocks past the
// last tile return at once.
constexpr int kXcds = 8;

__device__ inline int xcd_tile(int n_tiles) {
  const int per = (int)((gridDim.x + kXcds - 1) / kXcds);
  return (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
}

// The block's four waves share one tile of 16 NT output channels and take
// four consecutive 32-pixel tiles; each K step's weight chunk (16 NT
// channels x 32 k) is staged once per block in LDS, double-buffered, instead
// of being fetched from L2 by every wave.  NT is 4 for 64 or more output
// channels and 1 or 2 for the narrow layers, which would otherwise multiply
// zero weights.
constexpr int kBRow = 32 + 8;  // bf16 per staged weight row (padded against bank conflicts)

template <int NT>
__global__ __launch_bounds__(64 * kWaves) void conv_fwd_kernel(const __bf16* __restrict__ x,
                                                               const __bf16* __restrict__ w,
                                                               const float* __restrict__ bias,
                                                               __bf16* __restrict__ y, int N, int H, int W, int Cin,
                                                               int Cout, int relu) {
  constexpr int kCh = 16 * NT;
  __shared__ __bf16 sB[2][kCh * kBRow];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int pix_tiles = (N * H * W) / kFwdPix;
  const int ch_tiles = (Cout + kCh - 1) / kCh;
  const int block_tiles = ((pix_tiles + kWaves - 1) / kWaves) * ch_tiles;
  const int bt = xcd_tile(block_tiles);
  if (bt >= block_tiles) return;  // the whole block: no barrier is skipped
  const int ct = bt % ch_tiles;
  const int pt = (bt / ch_tiles) * kWaves + wave;
  const bool active = pt < pix_tiles;  // every wave stages and synchronises; only active ones compute
  const int p0 = (active ? pt : 0) * kFwdPix;
  // each run of 16 pixels lies in one image row (W is a multiple of 16)
  int img[2], yrow[2], x0[2];
  for (int i = 0; i < 2; ++i) {
    const int q = p0 + 16 * i, rem = q % (H * W);
    img[i] = q / (H * W);
    yrow[i] = rem / W;
    x0[i] = rem - yrow[i] * W;
  }
  const int c0 = ct * kCh;
  const int K = 9 * Cin;
  const int r = lane & 15, g = lane >> 4;  // fragment row / k group
  // staging: thread t < 4 kCh copies 8 k of channel c0 + t / 4
  const int s_co = threadIdx.x >> 2, s_kq = (threadIdx.x & 3) * 8;
  const bool stager = s_co < kCh;
  auto stage = [&](int buf, int k0) {
    if (!stager) return;
    const int co = c0 + s_co, k = k0 + s_kq;
    bf16x8 v = (co < Cout && k < K) ? *reinterpret_cast<const bf16x8*>(w + (size_t)co * K + k) : zero8();
    *reinterpret_cast<bf16x8*>(&sB[buf][s_co * kBRow + s_kq]) = v;
  };
  f32x4 acc[2][NT];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  stage(0, 0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += 32) {
    if (k0 + 32 < K) stage(buf ^ 1, k0 + 32);  // the next chunk, while this one is used
    const int kk = k0 + 8 * g;
    bf16x8 a[2], b[NT];
    for (int i = 0; i < 2; ++i) a[i] = load_a(x, img[i], yrow[i], x0[i] + r, kk, K, Cin, H, W);
    for (int j = 0; j < NT; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&sB[buf][(16 * j + r) * kBRow + 8 * g]);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < NT; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
    buf ^= 1;
  }
  if (!active) return;
  // C fragment: column r (output channel), rows 4 g .. 4 g + 3 (pixels)
  for (int j = 0; j < NT; ++j) {
    const int co = c0 + 16 * j + r;
    if (co >= Cout) continue;
    const float bv = bias ? bias[co] : 0.0f;
    for (int i = 0; i < 2; ++i)
      for (int e = 0; e < 4; ++e) {
        float v = acc[i][j][e] + bv;
        if (relu) v = v > 0.0f ? v : 0.0f;
        const int px = p0 + 16 * i + 4 * g + e;
        y[(size_t)px * Cout + co] = (__bf16)v;
      }
  }
}

// Split-K variant for deep layers (K = 9 Cin of many 32-steps): the block's
// four waves share one 32-pixel x 64-channel output tile and take every
// fourth K chunk; a group of four chunks is staged in LDS at a time and the
// four partial tiles are summed through LDS at the end.  A layer with few
// pixel tiles gets four times the waves this way.
__global__ __launch_bounds__(64 * kWaves) void conv_fwd_splitk_kernel(const __bf16* __restrict__ x,
                                                                      const __bf16* __restrict__ w,
                                                                      const float* __restrict__ bias,
                                                                      __bf16* __restrict__ y, int N, int H, int W,
                                                                      int Cin, int Cout, int relu) {
  __shared__ __bf16 sB[kWaves][kFwdCh * kBRow];
  __shared__ float sP[kWaves][32][64 + 1];  // partial tiles: [wave][element][lane]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ch_tiles = (Cout + kFwdCh - 1) / kFwdCh;
  const int n_tiles = ((N * H * W) / kFwdPix) * ch_tiles;
  const int bt = xcd_tile(n_tiles);
  if (bt >= n_tiles) return;  // the whole block: no barrier is skipped
  const int ct = bt % ch_tiles, pt = bt / ch_tiles;
  const int p0 = pt * kFwdPix;
  int img[2], yrow[2], x0[2];
  for (int i = 0; i < 2; ++i) {
    const int q = p0 + 16 * i, rem = q % (H * W);
    img[i] = q / (H * W);
    yrow[i] = rem / W;
    x0[i] = rem - yrow[i] * W;
  }
  const int c0 = ct * kFwdCh;
  const int K = 9 * Cin;
  const int r = lane & 15, g = lane >> 4;
  const int s_co = threadIdx.x >> 2, s_kq = (threadIdx.x & 3) * 8;
  f32x4 acc[2][4];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kg = 0; kg < K; kg += 32 * kWaves) {
    // stage the group's four chunks: thread t copies 8 k of channel c0 + t / 4 of each
    for (int c = 0; c < kWaves; ++c) {
      const int co = c0 + s_co, k = kg + 32 * c + s_kq;
      bf16x8 v = (co < Cout && k < K) ? *reinterpret_cast<const bf16x8*>(w + (size_t)co * K + k) : zero8();
      *reinterpret_cast<bf16x8*>(&sB[c][s_co * kBRow + s_kq]) = v;
    }
    __syncthreads();
    const int k0 = kg + 32 * wave;
    if (k0 < K) {
      const int kk = k0 + 8 * g;
      bf16x8 a[2], b[4];
      for (int i = 0; i < 2; ++i) a[i] = load_a(x, img[i], yrow[i], x0[i] + r, kk, K, Cin, H, W);
      for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&sB[wave][(16 * j + r) * kBRow + 8 * g]);
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 4; ++j)
      for (int e = 0; e < 4; ++e) sP[wave][(i * 4 + j) * 4 + e][lane] = acc[i][j][e];
  __syncthreads();
  // wave w sums and stores elements 8 w .. 8 w + 7 of every lane's tile
  for (int q = 0; q < 8; ++q) {
    const int idx = 8 * wave + q, i = idx >> 4, j = (idx >> 2) & 3, e = idx & 3;
    const float sum = sP[0][idx][lane] + sP[1][idx][lane] + sP[2][idx][lane] + sP[3][idx][lane];
    const int co = c0 + 16 * j + r;
    if (co >= Cout) continue;
    float v = sum + (bias ? bias[co] : 0.0f);
    if (relu) v = v > 0.0f ? v : 0.0f;
    const int px = p0 + 16 * i + 4 * g + e;
    y[(size_t)px * Cout + co] = (__bf16)v;
  }
}

constexpr int kWgPix = 32;   // pixels per K step
constexpr int kWgN = 64;     // (tap, ci) columns per block
constexpr int kWgMaxM = 128; // output channels per block (all of them)
// staged rows: [pixel][channel], padded by 32 bytes so that the 16 rows one
// transposing read touches per half-wave start 8 banks apart
constexpr int kWgDyRow = kWgMaxM + 16;
constexpr int kWgXRow = kWgN + 16;

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4;

// ds_read_b64_tr_b16 (gfx950): within each group of 16 lanes, element e of
// lane i's result is 16-bit element i % 4 of the 8 bytes that lane
// 4 e + i / 4 of the group addressed (checked on the hardware).  Lane (g, j)
// of an MFMA fragment (j = lane % 16, g = lane / 16) therefore addresses
// row 4 g + j / 4 (+ 16 for the second half), column col0 + 4 (j % 4), and
// receives column col0 + j of rows 4 g .. 4 g + 3 and 16 + 4 g .. 16 + 4 g
// + 3: the fragment's k = 8 g + t is staged pixel 16 (t / 4) + 4 g + t % 4.
// Both operands use this order, so the sum over pixels is unchanged.
__device__ inline bf16x8 frag_tr(const __bf16* tile, int row_len, int col0, int g, int j) {
  const __bf16* p0 = tile + (4 * g + (j >> 2)) * row_len + col0 + 4 * (j & 3);
  const __bf16* p1 = p0 + 16 * row_len;
  const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4*)p0);
  const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4*)p1);
  const s16x8 v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, v);
}

__global__ __launch_bounds__(64 * kWaves) void conv_wgrad_kernel(const __bf16* __restrict__ x,
                                                                 const __bf16* __restrict__ dy,
                                                                 float* __restrict__ dw, int N, int H, int W,
                                                                 int Cin, int Cout, int pix_per_split) {
  // double-buffered: step k + 1's global loads are in flight while step k's
  // MFMAs run, and land in the other buffer
  __shared__ __attribute__((aligned(16))) __bf16 s_dy[2][kWgPix * kWgDyRow];  // [buf][pixel][co]
  __shared__ __attribute__((aligned(16))) __bf16 s_x[2][kWgPix * kWgXRow];    // [buf][pixel][n]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int K9 = 9 * Cin;
  const int n0 = blockIdx.x * kWgN;
  const int P = N * H * W;
  const int pbeg = blockIdx.y * pix_per_split;
  const int pend = min(P, pbeg + pix_per_split);
  const int r = lane & 15, g = lane >> 4;
  const int m_tiles = Cout / 16;
  // per thread: up to two 8-channel pieces of dY (Cout <= 128: 32 x 16 pieces)
  // and one 8-column piece of the shifted input (32 x 8 pieces)
  constexpr int kDyPieces = kWgPix * (kWgMaxM / 8) / (64 * kWaves);
  const int n_dy = kWgPix * (Cout / 8);
  bf16x8 vdy[kDyPieces], vx;
  auto load = [&](int pc) {
    for (int q = 0; q < kDyPieces; ++q) {
      const int t = threadIdx.x + q * 64 * kWaves;
      vdy[q] = zero8();
      if (t < n_dy) {
        const int p = t / (Cout / 8), c8 = (t - p * (Cout / 8)) * 8;
        const int px = pc + p;
        if (px < pend) vdy[q] = *reinterpret_cast<const bf16x8*>(dy + (size_t)px * Cout + c8);
      }
    }
    {
      const int t = threadIdx.x;  // kWgPix * kWgN / 8 == 64 * kWaves pieces
      const int p = t / (kWgN / 8), n8 = (t - p * (kWgN / 8)) * 8;
      const int n = n0 + n8, px = pc + p;
      vx = zero8();
      if (px < pend && n < K9) {
        const int tap = n / Cin, ci = n - tap * Cin;
        const int img = px / (H * W), rem = px - img * H * W, yy = rem / W + tap / 3 - 1, xs = rem % W + tap % 3 - 1;
        if (yy >= 0 && yy < H && xs >= 0 && xs < W)
          vx = *reinterpret_cast<const bf16x8*>(x + (((size_t)img * H + yy) * W + xs) * Cin + ci);
      }
    }
  };
  auto store = [&](int buf) {  // as loaded: channels contiguous, one 16-byte store per piece
    for (int q = 0; q < kDyPieces; ++q) {
      const int t = threadIdx.x + q * 64 * kWaves;
      if (t < n_dy) {
        const int p = t / (Cout / 8), c8 = (t - p * (Cout / 8)) * 8;
        *reinterpret_cast<bf16x8*>(&s_dy[buf][p * kWgDyRow + c8]) = vdy[q];
      }
    }
    const int t = threadIdx.x;
    const int p = t / (kWgN / 8), n8 = (t - p * (kWgN / 8)) * 8;
    *reinterpret_cast<bf16x8*>(&s_x[buf][p * kWgXRow + n8]) = vx;
  };
  // wave w owns column tile w (16 of the block's 64 columns) and every row tile
  f32x4 acc[kWgMaxM / 16];
  for (int i = 0; i < kWgMaxM / 16; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  int buf = 0;
  if (pbeg < pend) {
    load(pbeg);
    store(0);
  }
  __syncthreads();
  for (int pc = pbeg; pc < pend; pc += kWgPix) {
    const bool more = pc + kWgPix < pend;
    if (more) load(pc + kWgPix);
    const bf16x8 b = frag_tr(s_x[buf], kWgXRow, 16 * wave, g, r);
#pragma unroll
    for (int i = 0; i < kWgMaxM / 16; ++i) {
      if (i >= m_tiles) break;
      const bf16x8 a = frag_tr(s_dy[buf], kWgDyRow, 16 * i, g, r);
      acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
    }
    if (more) store(buf ^ 1);  // the other buffer: nobody reads it during this step
    __syncthreads();
    buf ^= 1;
  }
  // C fragment: column n = n0 + 16 wave + r, rows co = 16 i + 4 g + e
  const int n = n0 + 16 * wave + r;
  if (n >= K9) return;
#pragma unroll
  for (int i = 0; i < kWgMaxM / 16; ++i) {
    if (i >= m_tiles) break;
    for (int e = 0; e < 4; ++e) {
      const int co = 16 * i + 4 * g + e;
      unsafeAtomicAdd(&dw[(size_t)co * K9 + n], acc[i][e]);
    }
  }
}

// The backward of a layer's ReLU and bias in one pass: dy_out = dy where
// the layer's output y > 0 (everywhere without ReLU), as BF16, and
// db[c] += sum over pixels of dy_out[., c] in FP32.  A thread takes 8
// channels of a pixel; the grid's thread count is a multiple of C / 8, so a
// thread keeps the same 8 channels over its whole grid-stride loop and sums
// them in registers; the wave folds those sums onto its first C / 8 lanes,
// which add them into the block's LDS sums, and one global atomic per
// channel and block finishes.
constexpr int kMaskBlock = 256;

template <bool kF32>
__global__ __launch_bounds__(kMaskBlock) void relu_bias_backward_kernel(const void* __restrict__ dy_in,
                                                                        const __bf16* __restrict__ y, int relu,
                                                                        __bf16* __restrict__ dy_out,
                                                                        float* __restrict__ db, long long pieces,
                                                                        int C) {
  __shared__ float s_db[kWgMaxM];
  for (int i = threadIdx.x; i < C; i += kMaskBlock) s_db[i] = 0.0f;
  __syncthreads();
  const int groups = C / 8;
  const long long stride = (long long)gridDim.x * kMaskBlock;
  long long q = (long long)blockIdx.x * kMaskBlock + threadIdx.x;
  const int c8 = (int)(q % groups) * 8;  // the same for every q of this thread: stride % groups == 0
  float acc[8];
  for (int e = 0; e < 8; ++e) acc[e] = 0.0f;
  for (; q < pieces; q += stride) {
    float v[8];
    if (kF32) {
      const float4* src = reinterpret_cast<const float4*>(static_cast<const float*>(dy_in) + q * 8);
      const float4 a = src[0], b = src[1];
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
      const bf16x8 d = reinterpret_cast<const bf16x8*>(dy_in)[q];
      for (int e = 0; e < 8; ++e) v[e] = (float)d[e];
    }
    bf16x8 o;
    if (relu) {
      const bf16x8 yy = reinterpret_cast<const bf16x8*>(y)[q];
      for (int e = 0; e < 8; ++e) o[e] = (float)yy[e] > 0.0f ? (__bf16)v[e] : (__bf16)0.0f;
    } else {
      for (int e = 0; e < 8; ++e) o[e] = (__bf16)v[e];
    }
    reinterpret_cast<bf16x8*>(dy_out)[q] = o;
    for (int e = 0; e < 8; ++e) acc[e] += (float)o[e];
  }
  // lanes l and l xor off (off a multiple of C / 8, a power of two) hold the
  // same channels: fold the wave onto its first C / 8 lanes, which then add
  // into the block's sums
  for (int off = 32; off >= groups; off >>= 1)
    for (int e = 0; e < 8; ++e) acc[e] += __shfl_xor(acc[e], off, 64);
  if ((int)(threadIdx.x & 63) < groups)
    for (int e = 0; e < 8; ++e) atomicAdd(&s_db[c8 + e], acc[e]);
  __syncthreads();
  for (int i = threadIdx.x; i < C; i += kMaskBlock) unsafeAtomicAdd(&db[i], s_db[i]);
}

// The U-Net's 2 x 2 max pooling and its nearest-neighbour upsampling with
// the skip concatenation, NHWC BF16, one thread per 8 channels of a pixel:
// bandwidth kernels in place of the strided copies, reductions and masks
// that the same steps take as tensor operations.  The pooling keeps, per
// channel, which input of the window (row-major: 0 top left .. 3 bottom
// right) held the first maximum, and its backward routes the gradient
// there, as max_pool2d does.
constexpr int kGlueBlock = 256;

unsigned glue_grid(long long pieces) {
  const long long b = (pieces + kGlueBlock - 1) / kGlueBlock;
  return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

__global__ __launch_bounds__(kGlueBlock) void pool2x2_kernel(const __bf16* __restrict__ x, __bf16* __restrict__ y,
                                                             uint8_t* __restrict__ arg, int n, int h, int w, int c) {
  const int ho = h / 2, wo = w / 2, g = c / 8;
  const long long total = (long long)n * ho * wo * g;
  for (long long q = blockIdx.x * (long long)kGlueBlock + threadIdx.x; q < total;
       q += (long long)gridDim.x * kGlueBlock) {
    const int c8 = (int)(q % g) * 8;
    long long p = q / g;  // output pixel
    const int xo = (int)(p % wo);
    p /= wo;
    const int yo = (int)(p % ho);
    const long long img = p / ho;
    const __bf16* base = x + ((img * h + 2 * yo) * (long long)w + 2 * xo) * c + c8;
    bf16x8 v[4];
    v[0] = *reinterpret_cast<const bf16x8*>(base);
    v[1] = *reinterpret_cast<const bf16x8*>(base + c);
    v[2] = *reinterpret_cast<const bf16x8*>(base + (long long)w * c);
    v[3] = *reinterpret_cast<const bf16x8*>(base + (long long)w * c + c);
    bf16x8 m;
    uint64_t a = 0;
    for (int e = 0; e < 8; ++e) {
      float best = (float)v[0][e];
      int bi = 0;
      for (int k = 1; k < 4; ++k) {
        const float t = (float)v[k][e];
        if (t > best) {
          best = t;
          bi = k;
        }
      }
      m[e] = v[bi][e];
      a |= (uint64_t)bi << (8 * e);
    }
    *reinterpret_cast<bf16x8*>(y + q * 8) = m;  // output pixel p, channels c8.. : element p c + c8 = 8 q
    *reinterpret_cast<uint64_t*>(arg + q * 8) = a;
  }
}

__global__ __launch_bounds__(kGlueBlock) void pool2x2_backward_kernel(const __bf16* __restrict__ dy,
                                                                      const uint8_t* __restrict__ arg,
                                                                      __bf16* __restrict__ dx, int n, int h, int w,
                                                                      int c) {
  const int ho = h / 2, wo = w / 2, g = c / 8;
  const long long total = (long long)n * ho * wo * g;
  for (long long q = blockIdx.x * (long long)kGlueBlock + threadIdx.x; q < total;
       q += (long long)gridDim.x * kGlueBlock) {
    const int c8 = (int)(q % g) * 8;
    long long p = q / g;
    const int xo = (int)(p % wo);
    p /= wo;
    const int yo = (int)(p % ho);
    const long long img = p / ho;
    const bf16x8 d = *reinterpret_cast<const bf16x8*>(dy + q * 8);
    const uint64_t a = *reinterpret_cast<const uint64_t*>(arg + q * 8);
    __bf16* base = dx + ((img * h + 2 * yo) * (long long)w + 2 * xo) * c + c8;
    __bf16* out[4] = {base, base + c, base + (long long)w * c, base + (long long)w * c + c};
    for (int k = 0; k < 4; ++k) {
      bf16x8 v;
      for (int e = 0; e < 8; ++e) v[e] = ((a >> (8 * e)) & 0xffu) == (uint64_t)k ? d[e] : (__bf16)0.0f;
      *reinterpret_cast<bf16x8*>(out[k]) = v;
    }
  }
}

// out [n][h][w][cz + cs] = (z upsampled 2x by repetition, s): the decoder's
// input, channels of z first
__global__ __launch_bounds__(kGlueBlock) void upcat_kernel(const __bf16* __restrict__ z, const __bf16* __restrict__ s,
                                                           __bf16* __restrict__ out, int n, int h, int w, int cz,
                                                           int cs) {
  const int g = (cz + cs) / 8;
  const long long total = (long long)n * h * w * g;
  for (long long q = blockIdx.x * (long long)kGlueBlock + threadIdx.x; q < total;
       q += (long long)gridDim.x * kGlueBlock) {
    const int c8 = (int)(q % g) * 8;
    long long p = q / g;
    const int xx = (int)(p % w);
    p /= w;
    const int yy = (int)(p % h);
    const long long img = p / h;
    const bf16x8 v = c8 < cz
        ? *reinterpret_cast<const bf16x8*>(z + ((img * (h / 2) + yy / 2) * (long long)(w / 2) + xx / 2) * cz + c8)
        : *reinterpret_cast<const bf16x8*>(s + ((img * h + yy) * (long long)w + xx) * cs + (c8 - cz));
    *reinterpret_cast<bf16x8*>(out + q * 8) = v;
  }
}

// backward of upcat: dz = the sum of dout's z channels over each 2 x 2 block
// (FP32, one rounding), ds = dout's s channels
__global__ __launch_bounds__(kGlueBlock) void upcat_backward_kernel(const __bf16* __restrict__ dout,
                                                                    __bf16* __restrict__ dz, __bf16* __restrict__ ds,
                                                                    int n, int h, int w, int cz, int cs) {
  const int ct = cz + cs, gz = cz / 8, gs = cs / 8;
  const long long nz = (long long)n * (h / 2) * (w / 2) * gz, ns = (long long)n * h * w * gs;
  for (long long q = blockIdx.x * (long long)kGlueBlock + threadIdx.x; q < nz + ns;
       q += (long long)gridDim.x * kGlueBlock) {
    if (q < nz) {
      const int c8 = (int)(q % gz) * 8;
      long long p = q / gz;
      const int xo = (int)(p % (w / 2));
      p /= (w / 2);
      const int yo = (int)(p % (h / 2));
      const long long img = p / (h / 2);
      const __bf16* base = dout + ((img * h + 2 * yo) * (long long)w + 2 * xo) * ct + c8;
      const bf16x8 v0 = *reinterpret_cast<const bf16x8*>(base);
      const bf16x8 v1 = *reinterpret_cast<const bf16x8*>(base + ct);
      const bf16x8 v2 = *reinterpret_cast<const bf16x8*>(base + (long long)w * ct);
      const bf16x8 v3 = *reinterpret_cast<const bf16x8*>(base + (long long)w * ct + ct);
      bf16x8 o;
      for (int e = 0; e < 8; ++e) o[e] = (__bf16)((((float)v0[e] + (float)v1[e]) + (float)v2[e]) + (float)v3[e]);
      *reinterpret_cast<bf16x8*>(dz + q * 8) = o;
    } else {
      const long long r = q - nz;
      const int c8 = (int)(r % gs) * 8;
      const long long p = r / gs;
      *reinterpret_cast<bf16x8*>(ds + r * 8) = *reinterpret_cast<const bf16x8*>(dout + p * ct + cz + c8);
    }
  }
}

// The BF16 operands of a layer's forward and data-gradient convolutions from
// its FP32 master weights, in one pass: w [cout][9][cin] rounded to nearest
// even, and the data gradient's weights [cin][9][cout] with the taps
// reversed (tap 8 - t), in place of a cast, a flip and a transposing copy.
constexpr int kPrepBlock = 256;
__global__ __launch_bounds__(kPrepBlock) void weights_bf16_kernel(const float* __restrict__ w, __bf16* __restrict__ wb,
                                                                   __bf16* __restrict__ wt, float* __restrict__ dw,
                                                                   float* __restrict__ db, int cout, int cin) {
  const int n = cout * 9 * cin;
  for (int i = blockIdx.x * kPrepBlock + threadIdx.x; i < n; i += gridDim.x * kPrepBlock) {
    const int ci = i % cin, t = (i / cin) % 9, co = i / (9 * cin);
    const __bf16 v = (__bf16)w[i];
    wb[i] = v;
    wt[((size_t)ci * 9 + (8 - t)) * cout + co] = v;
    // the layer's gradient accumulators, zeroed here instead of by fills of their own
    if (dw) dw[i] = 0.0f;
    if (db && i < cout) db[i] = 0.0f;
  }
}

bool shapes_ok(int N, int H, int W, int Cin, int Cout) {
  return N > 0 && H > 0 && W > 0 && W % 16 == 0 && ((long long)N * H * W) % kFwdPix == 0 && Cin > 0 &&
         Cin % 8 == 0 && Cout > 0 && Cout % 16 == 0 && (long long)N * H * W * (Cin > Cout ? Cin : Cout) < (1ll << 31);
}

}  // namespace

extern "C" {

const char* ggh_unet_last_error(void) { return g_unet_error.c_str(); }

int ggh_conv3x3_forward(const void* x, const void* w, const float* bias, void* y, int32_t n, int32_t h, int32_t wd,
                        int32_t cin, int32_t cout, int32_t relu, void* stream) {
  if (!x || !w || !y) return ufail(-1, "null pointer");
  if (!shapes_ok(n, h, wd, cin, cout)) return ufail(-2, "unsupported shape (W % 16, N H W % 32, Cin % 8, Cout % 16)");
  const int pix_tiles = n * h * wd / kFwdPix, ch_tiles = (cout + kFwdCh - 1) / kFwdCh;
  // deep layers with few output tiles: split K over the block's waves
  const bool split = 9 * cin >= 8 * 32 * kWaves && pix_tiles * ch_tiles < 1024;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const __bf16 *xb = static_cast<const __bf16*>(x), *wb = static_cast<const __bf16*>(w);
  __bf16* yb = static_cast<__bf16*>(y);
  if (split) {
    const dim3 grid((unsigned)((pix_tiles * ch_tiles + kXcds - 1) / kXcds * kXcds));
    hipLaunchKernelGGL(conv_fwd_splitk_kernel, grid, dim3(64 * kWaves), 0, st, xb, wb, bias, yb, n, h, wd, cin, cout,
                       relu);
  } else {
    // output channels per wave: 64, or 16 / 32 for the narrow layers
    const int nt = cout >= 64 ? 4 : (cout >= 32 ? 2 : 1);
    const int block_tiles = ((pix_tiles + kWaves - 1) / kWaves) * ((cout + 16 * nt - 1) / (16 * nt));
    const dim3 grid((unsigned)((block_tiles + kXcds - 1) / kXcds * kXcds));
    if (nt == 4)
      hipLaunchKernelGGL(conv_fwd_kernel<4>, grid, dim3(64 * kWaves), 0, st, xb, wb, bias, yb, n, h, wd, cin, cout, relu);
    else if (nt == 2)
      hipLaunchKernelGGL(conv_fwd_kernel<2>, grid, dim3(64 * kWaves), 0, st, xb, wb, bias, yb, n, h, wd, cin, cout, relu);
    else
      hipLaunchKernelGGL(conv_fwd_kernel<1>, grid, dim3(64 * kWaves), 0, st, xb, wb, bias, yb, n, h, wd, cin, cout, relu);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ufail(-100 - (int)e, std::string("conv forward: ") + hipGetErrorString(e));
  return 0;
}

int ggh_relu_bias_backward(const void* dy, int32_t dy_is_f32, const void* y, int32_t relu, void* dy_out, float* db,
                           int64_t n_pixels, int32_t c, void* stream) {
  if (!dy || !dy_out || !db || (relu && !y)) return ufail(-1, "null pointer");
  // C / 8 must divide the block size (a thread keeps its 8 channels)
  if (n_pixels <= 0 || (c != 8 && c != 16 && c != 32 && c != 64 && c != 128))
    return ufail(-2, "unsupported shape (C one of 8, 16, 32, 64, 128)");
  const long long pieces = (long long)n_pixels * (c / 8);
  const long long want = (pieces + kMaskBlock - 1) / kMaskBlock;
  // one block per compute unit: every block ends with one global atomic per
  // channel onto the same C addresses, and more blocks queue up there
  // (measured over a training run: 92.7 ms at 256 blocks, 100 at 512, 143 at
  // 1,024, 294 at 4,096, 125 at 128)
  const unsigned blocks = (unsigned)(want < 256 ? want : 256);
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (dy_is_f32)
    hipLaunchKernelGGL(relu_bias_backward_kernel<true>, dim3(blocks), dim3(kMaskBlock), 0, st, dy,
                       static_cast<const __bf16*>(y), relu, static_cast<__bf16*>(dy_out), db, pieces, c);
  else
    hipLaunchKernelGGL(relu_bias_backward_kernel<false>, dim3(blocks), dim3(kMaskBlock), 0, st, dy,
                       static_cast<const __bf16*>(y), relu, static_cast<__bf16*>(dy_out), db, pieces, c);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ufail(-100 - (int)e, std::string("ReLU / bias backward: ") + hipGetErrorString(e));
  return 0;
}

int ggh_conv3x3_weights_bf16(const float* w, void* w_bf16, void* wt_bf16, float* dw_zero, float* db_zero,
                             int32_t cout, int32_t cin, void* stream) {
  if (!w || !w_bf16 || !wt_bf16) return ufail(-1, "null pointer");
  if (cout < 1 || cin < 1 || (long long)cout * 9 * cin >= (1ll << 31)) return ufail(-2, "unsupported shape");
  const long long n = (long long)cout * 9 * cin;
  const unsigned blocks = (unsigned)((n + kPrepBlock - 1) / kPrepBlock < 4096 ? (n + kPrepBlock - 1) / kPrepBlock : 4096);
  hipLaunchKernelGGL(weights_bf16_kernel, dim3(blocks), dim3(kPrepBlock), 0, static_cast<hipStream_t>(stream), w,
                     static_cast<__bf16*>(w_bf16), static_cast<__bf16*>(wt_bf16), dw_zero, db_zero, cout, cin);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ufail(-100 - (int)e, std::string("BF16 weights: ") + hipGetErrorString(e));
  return 0;
}

int ggh_conv3x3_weight_grad(const void* x, const void* dy, float* dw, int32_t n, int32_t h, int32_t wd, int32_t cin,
                            int32_t cout, int32_t splits, void* stream) {
  if (!x || !dy || !dw) return ufail(-1, "null pointer");
  if (!shapes_ok(n, h, wd, cin, cout) || cout > kWgMaxM) return ufail(-2, "unsupported shape (Cout <= 128)");
  const int P = n * h * wd;
  if (splits < 1) splits = 1;
  int per = (P + splits - 1) / splits;
  per = (per + kWgPix - 1) / kWgPix * kWgPix;
  splits = (P + per - 1) / per;
  const int cols = (9 * cin + kWgN - 1) / kWgN;
  hipLaunchKernelGGL(conv_wgrad_kernel, dim3(cols, splits), dim3(64 * kWaves), 0, static_cast<hipStream_t>(stream),
                     static_cast<const __bf16*>(x), static_cast<const __bf16*>(dy), dw, n, h, wd, cin, cout, per);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ufail(-100 - (int)e, std::string("conv weight gradient: ") + hipGetErrorString(e));
  return 0;
}

int ggh_pool2x2_forward(const void* x, void* y, uint8_t* arg, int32_t n, int32_t h, int32_t wd, int32_t c,
                        void* stream) {
  if (!x || !y || !arg) return ufail(-1, "null pointer");
  if (n <= 0 || h <= 0 || wd <= 0 || h % 2 || wd % 2 || c <= 0 || c % 8) return ufail(-2, "unsupported shape");
  const long long pieces = (long long)n * (h / 2) * (wd / 2) * (c / 8);
  hipLaunchKernelGGL(pool2x2_kernel, dim3(glue_grid(pieces)), dim3(kGlueBlock), 0, static_cast<hipStream_t>(stream),
                     static_cast<const __bf16*>(x), static_cast<__bf16*>(y), arg, n, h, wd, c);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ufail(-100 - (int)e, std::string("pooling: ") + hipGetErrorString(e));
  return 0;
}

int ggh_pool2x2_backward(const void* dy, const uint8_t* arg, void* dx, int32_t n, int32_t h, int32_t wd, int32_t c,
                         void* stream) {
  if (!dy || !arg || !dx) return ufail(-1, "null pointer");
  if (n <= 0 || h <= 0 || wd <= 0 || h % 2 || wd % 2 || c <= 0 || c % 8) return ufail(-2, "unsupported shape");
  const long long pieces = (long long)n * (h / 2) * (wd / 2) * (c / 8);
  hipLaunchKernelGGL(pool2x2_backward_kernel, dim3(glue_grid(pieces)), dim3(kGlueBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<const __bf16*>(dy), arg, static_cast<__bf16*>(dx), n,
                     h, wd, c);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ufail(-100 - (int)e, std::string("pooling backward: ") + hipGetErrorString(e));
  return 0;
}

int ggh_upcat_forward(const void* z, const void* s, void* out, int32_t n, int32_t h, int32_t wd, int32_t cz,
                      int32_t cs, void* stream) {
  if (!z || !s || !out) return ufail(-1, "null pointer");
  if (n <= 0 || h <= 0 || wd <= 0 || h % 2 || wd % 2 || cz <= 0 || cz % 8 || cs <= 0 || cs % 8)
    return ufail(-2, "unsupported shape");
  const long long pieces = (long long)n * h * wd * ((cz + cs) / 8);
  hipLaunchKernelGGL(upcat_kernel, dim3(glue_grid(pieces)), dim3(kGlueBlock), 0, static_cast<hipStream_t>(stream),
                     static_cast<const __bf16*>(z), static_cast<const __bf16*>(s), static_cast<__bf16*>(out), n, h, wd,
                     cz, cs);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ufail(-100 - (int)e, std::string("upsampling: ") + hipGetErrorString(e));
  return 0;
}

int ggh_upcat_backward(const void* dout, void* dz, void* ds, int32_t n, int32_t h, int32_t wd, int32_t cz, int32_t cs,
                       void* stream) {
  if (!dout || !dz || !ds) return ufail(-1, "null pointer");
  if (n <= 0 || h <= 0 || wd <= 0 || h % 2 || wd % 2 || cz <= 0 || cz % 8 || cs <= 0 || cs % 8)
    return ufail(-2, "unsupported shape");
  const long long pieces = (long long)n * (h / 2) * (wd / 2) * (cz / 8) + (long long)n * h * wd * (cs / 8);
  hipLaunchKernelGGL(upcat_backward_kernel, dim3(glue_grid(pieces)), dim3(kGlueBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<const __bf16*>(dout), static_cast<__bf16*>(dz),
                     static_cast<__bf16*>(ds), n, h, wd, cz, cs);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ufail(-100 - (int)e, std::string("upsampling backward: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
