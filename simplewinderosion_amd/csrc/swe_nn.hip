// swe_nn.hip -- kernels of the learned surrogate (a periodic convolutional
// encoder-decoder that predicts K map-cycles of erosion at once): 3x3
// convolutions on BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation)
// forward, data gradient (the forward kernel on flipped weights) and weight
// gradient, plus the elementwise, pooling, loss, optimiser and fixed-point
// mass-projection kernels around them.
//
// Activations are NHWC (channels contiguous), bf16; weights [Cout][9][Cin]
// bf16 with tap t = 3*(dy+1) + (dx+1); padding is periodic, like the maps.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/swe_nn.h"

static_assert(sizeof(swe_nn_adam_seg) == 72, "swe_nn_adam_seg: six pointers, two int64, two int32");

namespace swenn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f2bf(float v) { return (bf16)v; }  // round to nearest even

__device__ __forceinline__ int32_t wrapi(int32_t v, int32_t n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// three offsets picked by a wave-uniform index in 0..2
struct Off3 {
  uint32_t v0, v1, v2;
  __device__ __forceinline__ uint32_t at(int32_t i) const { return i == 0 ? v0 : (i == 1 ? v1 : v2); }
};

// ---------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM: rows = pixels (the batch's pixels
// flattened, n-major), columns = output channels, reduction k = tap * Cin +
// ci in steps of 32 (one tap, 32 input channels; channels at or beyond Cin
// read as zero, so any Cin % 8 == 0 works).  A workgroup (4 waves) owns 128
// pixels x NT output channels.  Each step's A block [128][32] (the pixels
// shifted by the tap, periodic) and B block [NT][32] are staged in LDS,
// double buffered with rows padded to 40 elements: every thread loads its
// 16-byte runs of the next step into registers while the waves run the
// current step's MFMAs, then stores them into the other buffer (one barrier
// per step).  Waves tile the block 2 x 2 (NT >= 64) or 4 x 1; per
// v_mfma_f32_16x16x32_bf16, lane l supplies A row l%16 and B row l%16, both
// at k = 8*(l/16) .. +7, and holds D rows 4*(l/16) .. +3 of column l%16.
// A step issues only 8 MFMAs a wave, so its loop, address and barrier work
// is what the step costs: no step divides, and addresses are 32-bit.
// ---------------------------------------------------------------------------
// ACT 0: none, 1: ReLU, 2: zero where mask <= 0 (the ReLU gradient of a
// data gradient through the layer below, mask = that layer's output).
// With part, workgroup z of gridDim.z takes the k-steps [z K / Z, (z + 1) K /
// Z) and stores its raw sums in part[z][pixel][channel] (k_conv_reduce adds
// the Z partials in order and applies the epilogue): split-K for layers with
// few pixel blocks, whose workgroups would otherwise each walk every k-step.
template <int NT, int ACT>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16* __restrict__ x, const bf16* __restrict__ w,
                                                 const float* __restrict__ bias, const bf16* __restrict__ mask,
                                                 bf16* __restrict__ y16, float* __restrict__ y32, int32_t N,
                                                 int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                                                 float* __restrict__ part) {
  constexpr int WM = NT >= 64 ? 2 : 4;  // waves along the pixels
  constexpr int WN = 4 / WM;            // ... and the channels
  constexpr int TM = 128 / WM, TN = NT / WN;
  constexpr int BM = TM / 16, BN = TN / 16;
  constexpr int KC = 32, LD = KC + 8;
  __shared__ __attribute__((aligned(16))) bf16 sA[2][128 * LD];
  __shared__ __attribute__((aligned(16))) bf16 sB[2][NT * LD];
  const int32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t wm = wave / WN, wn = wave % WN;
  const int64_t P = (int64_t)N * H * W;
  const int64_t p0 = (int64_t)blockIdx.x * 128;
  const int32_t co0 = blockIdx.y * NT;
  const int32_t cchunks = (Cin + KC - 1) / KC, ksteps = 9 * cchunks;
  const int32_t g = tid & 3, r0 = tid >> 2;  // loader: row r0 (+64), channel group g

  // Loads are raw buffer loads at 32-bit byte offsets (the host keeps x and
  // w under 2 GiB): an offset at or above kBad lies beyond the buffer and
  // reads zero, which stands for pixels past P, output channels past Cout and
  // input channels past Cin.  Each thread keeps the byte offsets of its two
  // A pixels' three (periodic) rows and three columns, so a step's address
  // is two selects and an add rather than 64-bit index arithmetic.
  // Descriptors: stride 0 (raw: offsets are bytes, checked against
  // num_records = the buffer's size), flags word 0x00020000 = the 32-bit
  // data format the gfx9 family's untyped buffer loads need.
  constexpr uint32_t kBad = 0x80000000u;
  constexpr int kRawFlags = 0x00020000;
  const __amdgpu_buffer_rsrc_t xr =
      __builtin_amdgcn_make_buffer_rsrc((void*)x, (short)0, (int)(P * Cin * 2), kRawFlags);
  const __amdgpu_buffer_rsrc_t wr =
      __builtin_amdgcn_make_buffer_rsrc((void*)w, (short)0, (int)((int64_t)Cout * 9 * Cin * 2), kRawFlags);
  Off3 rowo[2], colo[2];  // named fields, not arrays: a select between them stays in registers
  uint32_t wo[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int64_t p = p0 + r0 + 64 * h;
    const bool ok = p < P;
    const int64_t pc = ok ? p : 0;
    const int32_t ax = (int32_t)(pc % W), ay = (int32_t)((pc / W) % H), ai = (int32_t)(pc / ((int64_t)W * H));
    auto row = [&](int32_t d) { return ok ? (uint32_t)(((int64_t)ai * H + wrapi(ay + d, H)) * W * Cin * 2) : kBad; };
    auto col = [&](int32_t d) { return (uint32_t)(wrapi(ax + d, W) * Cin * 2 + 16 * g); };
    rowo[h] = Off3{row(-1), row(0), row(1)};
    colo[h] = Off3{col(-1), col(0), col(1)};
    const int32_t co = co0 + r0 + 64 * h;
    wo[h] = co < Cout ? (uint32_t)((int64_t)co * 9 * Cin * 2 + 16 * g) : kBad;
  }
  const int32_t kz = (int32_t)gridDim.z, z = (int32_t)blockIdx.z;
  const int32_t s0 = (int32_t)((int64_t)z * ksteps / kz), s1 = (int32_t)((int64_t)(z + 1) * ksteps / kz);
  // the next step to load, kept as tap (row ty, column tx) and chunk c so
  // that no step divides; abase: the A byte offsets of that tap, recomputed
  // only when the tap changes, so a step adds the chunk and selects kBad for
  // channels past Cin
  int32_t ltap = s0 / cchunks, lc = s0 - ltap * cchunks;
  int32_t lty = ltap / 3, ltx = ltap - 3 * lty;
  uint32_t abase[2];
  auto settap = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < 2; h++) abase[h] = rowo[h].at(lty) + colo[h].at(ltx);
  };
  settap();
  const int32_t cin_g = Cin - 8 * g;  // a lane's channels at k are valid while k < cin_g
  bf16x8 ra[2], rb[2];
  auto load = [&]() __attribute__((always_inline)) {
    const int32_t k = lc * KC;
    const bool cok = k < cin_g;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint32_t o = abase[h] + (uint32_t)(k * 2);
      ra[h] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(xr, (int)(cok ? o : kBad), 0, 0));
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int32_t row = r0 + 64 * h;
      if (row < NT && (h == 0 || NT > 64)) {
        const uint32_t o = wo[h] + (uint32_t)((ltap * Cin + k) * 2);
        rb[h] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(wr, (int)(cok ? o : kBad), 0, 0));
      }
    }
    if (++lc == cchunks) {
      lc = 0;
      ++ltap;
      if (++ltx == 3) {
        ltx = 0;
        ++lty;
      }
      settap();
    }
  };
  auto store = [&](int32_t buf) __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < 2; h++) *(bf16x8*)&sA[buf][(r0 + 64 * h) * LD + 8 * g] = ra[h];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int32_t row = r0 + 64 * h;
      if (row < NT && (h == 0 || NT > 64)) *(bf16x8*)&sB[buf][row * LD + 8 * g] = rb[h];
    }
  };

  f32x4 acc[BM][BN];
#pragma unroll
  for (int i = 0; i < BM; i++)
#pragma unroll
    for (int j = 0; j < BN; j++) acc[i][j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};

  load();
  store(0);
  __syncthreads();
  const int32_t kq = 8 * (lane >> 4), rl = lane & 15;
  for (int32_t s = s0; s < s1; s++) {
    const int32_t buf = (s - s0) & 1;
    if (s + 1 < s1) load();
    bf16x8 a[BM], b[BN];
#pragma unroll
    for (int i = 0; i < BM; i++) a[i] = *(const bf16x8*)&sA[buf][(wm * TM + 16 * i + rl) * LD + kq];
#pragma unroll
    for (int j = 0; j < BN; j++) b[j] = *(const bf16x8*)&sB[buf][(wn * TN + 16 * j + rl) * LD + kq];
#pragma unroll
    for (int i = 0; i < BM; i++)
#pragma unroll
      for (int j = 0; j < BN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (s + 1 < s1) store(buf ^ 1);
    __syncthreads();
  }
  if (part) {
    float* pz = part + (int64_t)z * P * Cout;
#pragma unroll
    for (int j = 0; j < BN; j++) {
      const int32_t co = co0 + wn * TN + 16 * j + rl;
      if (co >= Cout) continue;
#pragma unroll
      for (int i = 0; i < BM; i++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int64_t p = p0 + wm * TM + 16 * i + 4 * (lane >> 4) + e;
          if (p < P) pz[p * Cout + co] = acc[i][j][e];
        }
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < BN; j++) {
    const int32_t co = co0 + wn * TN + 16 * j + rl;
    if (co >= Cout) continue;
    const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int i = 0; i < BM; i++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int64_t p = p0 + wm * TM + 16 * i + 4 * (lane >> 4) + e;
        if (p >= P) continue;
        float v = acc[i][j][e] + bv;
        if (ACT == 1) v = v > 0.0f ? v : 0.0f;
        if (ACT == 2) v = bf2f(mask[p * Cout + co]) > 0.0f ? v : 0.0f;
        if (y16) y16[p * Cout + co] = f2bf(v);
        if (y32) y32[p * Cout + co] = v;
      }
  }
}

// ---------------------------------------------------------------------------
// 3x3 convolution over row strips (W % 32 == 0, Cin <= CI, CI = 32 or 64):
// a workgroup owns 32 output channels and walks R rows of one 32-pixel
// column strip of one image.  Its weights [32][9][CI] stay in LDS for the
// whole strip; input rows y - 1, y, y + 1 (34 pixels with the periodic
// halo, all input channels) sit in a four-row LDS ring, the next row loaded
// while the current one is computed, so every input value is read from
// memory once per strip and serves all nine taps.  NHWC rows are the MFMA's
// A layout (lane: one pixel, 8 consecutive channels) and the weights' its B
// layout, so both are copied as 16-byte runs; the tap's shift is the ring
// row and the pixel row within it.  Waves tile the 32 x 32 output block of a
// row 2 x 2 (16 pixels x 16 channels each) and write it before the next row.
// Epilogue as k_conv3x3 (bias; ACT 0 none, 1 ReLU, 2 mask).
// ---------------------------------------------------------------------------
template <int CI, int ACT>
__global__ __launch_bounds__(256) void k_conv_rows(const bf16* __restrict__ x, const bf16* __restrict__ w,
                                                   const float* __restrict__ bias, const bf16* __restrict__ mask,
                                                   bf16* __restrict__ y16, float* __restrict__ y32, int32_t N,
                                                   int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t R) {
  constexpr int LDX = CI + 8;          // ring row stride (elements): 16-byte aligned, shifted banks
  constexpr int LDW = 9 * CI + 8;      // weight row stride
  constexpr int XG = CI / 8;           // 16-byte groups of a pixel
  constexpr int XITEMS = 34 * XG;      // loads of one input row
  constexpr int XPER = (XITEMS + 255) / 256;
  __shared__ __attribute__((aligned(16))) bf16 s_w[32 * LDW];
  __shared__ __attribute__((aligned(16))) bf16 s_x[4][34 * LDX];
  const int32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t segs = W / 32, cps = (H + R - 1) / R;
  const int32_t chunk = blockIdx.x;
  const int32_t rc = chunk % cps, sg = (chunk / cps) % segs, n = chunk / (cps * segs);
  const int32_t co0 = blockIdx.y * 32;
  const int32_t y0 = rc * R, rows = R < H - y0 ? R : H - y0, x0 = sg * 32;
  const int64_t img = (int64_t)n * H;
  const bf16x8 zero = {};

  // weights of the 32 output channels: [co][tap][ci], channels >= Cin zero
  for (int32_t u = tid; u < 32 * 9 * XG; u += 256) {
    const int32_t g = u % XG, t = (u / XG) % 9, col = u / (9 * XG);
    const int32_t co = co0 + col, ci = 8 * g;
    bf16x8 v = zero;
    if (co < Cout && ci < Cin) v = *(const bf16x8*)(w + ((int64_t)co * 9 + t) * Cin + ci);
    *(bf16x8*)&s_w[col * LDW + t * CI + ci] = v;
  }
  // input row loads: items (pixel column c = 0 .. 33, group g)
  int32_t xc[XPER], xg[XPER];
  bool xitem[XPER], xok[XPER];
  int64_t xcol[XPER];
#pragma unroll
  for (int h = 0; h < XPER; h++) {
    const int32_t u = tid + 256 * h;
    xitem[h] = u < XITEMS;
    xc[h] = u / XG;
    xg[h] = u % XG;
    xok[h] = xitem[h] && 8 * xg[h] < Cin;
    xcol[h] = (int64_t)wrapi(x0 + xc[h] - 1, W) * Cin + 8 * xg[h];
  }
  struct Row {
    bf16x8 v[XPER];
  };
  auto load = [&](int32_t k) __attribute__((always_inline)) {  // image row y0 - 1 + k
    Row r;
    const int64_t xrow = (img + wrapi(y0 - 1 + k, H)) * W * Cin;
#pragma unroll
    for (int h = 0; h < XPER; h++) r.v[h] = xok[h] ? *(const bf16x8*)(x + xrow + xcol[h]) : zero;
    return r;
  };
  auto store = [&](const Row& r, int32_t slot) __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < XPER; h++)
      if (xitem[h]) *(bf16x8*)&s_x[slot][xc[h] * LDX + 8 * xg[h]] = r.v[h];
  };
  // ring rows k = 0, 1, 2 in LDS; row k = 3 in registers (two rows ahead
  // of the first step's need)
  Row ra, rb;
  {
    const Row r0 = load(0), r1 = load(1), r2 = load(2);
    if (rows > 1) ra = load(3);
    store(r0, 0);
    store(r1, 1);
    store(r2, 2);
  }
  __syncthreads();

  const int32_t pt = wave & 1, ct = wave >> 1;  // 16 pixels x 16 channels of the 32 x 32 block
  const int32_t kq = 8 * (lane >> 4), rl = lane & 15;
  const int32_t co_b = 16 * ct + rl;            // B row (output channel within the block)
  const int32_t co = co0 + co_b;
  const float bv = (bias && co < Cout) ? bias[co] : 0.0f;
  // A wave's B fragments (its 16 output channels, every tap and 32-channel
  // slice: 9 or 18) stay in registers for the whole strip, so a row step
  // reads only A from LDS (measured faster than reading B per row also at
  // 64 input channels, where they take 72 registers)
  constexpr int BREG = 9 * (CI / 32);
  bf16x8 breg[BREG];
#pragma unroll
  for (int u = 0; u < BREG; u++)
    breg[u] = *(const bf16x8*)(s_w + co_b * LDW + (u / (CI / 32)) * CI + 32 * (u % (CI / 32)) + kq);
  // step s: LDS holds rows s .. s + 2, `held` row s + 3 (needed from step
  // s + 1 on); row s + 4 is loaded into `next`, and `held` is stored into
  // the slot of row s - 1, free since the barrier that ended step s - 1
  auto step = [&](int32_t s, Row& held, Row& next) __attribute__((always_inline)) {
    if (s + 2 < rows) next = load(s + 4);
    // D rows (pixels) 4 (lane / 16) + e of the wave's 16, column = channel co
    const int64_t prow = (img + y0 + s) * W + x0 + 16 * pt + 4 * (lane >> 4);
    float mk[4] = {1.0f, 1.0f, 1.0f, 1.0f};  // the mask is requested before the MFMAs
    if (ACT == 2 && co < Cout) {
#pragma unroll
      for (int e = 0; e < 4; e++) mk[e] = bf2f(mask[(prow + e) * Cout + co]);
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int dyi = 0; dyi < 3; dyi++) {
      const bf16* xr = s_x[(s + dyi) & 3];
#pragma unroll
      for (int dxi = 0; dxi < 3; dxi++) {
        const int32_t t = 3 * dyi + dxi;
        // A: pixel 16 pt + rl of the row, shifted by dx (column + dxi)
        const bf16* ap = xr + (16 * pt + rl + dxi) * LDX + kq;
#pragma unroll
        for (int kc = 0; kc < CI / 32; kc++) {
          const bf16x8 a = *(const bf16x8*)(ap + 32 * kc);
          const bf16x8 b = breg[t * (CI / 32) + kc];
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
      }
    }
    if (co < Cout) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int64_t p = prow + e;
        float v = acc[e] + bv;
        if (ACT == 1) v = v > 0.0f ? v : 0.0f;
        if (ACT == 2) v = mk[e] > 0.0f ? v : 0.0f;
        if (y16) y16[p * Cout + co] = f2bf(v);
        if (y32) y32[p * Cout + co] = v;
      }
    }
    if (s + 1 < rows) store(held, (s + 3) & 3);
    __syncthreads();
  };
  for (int32_t s = 0; s < rows; s += 2) {
    step(s, ra, rb);
    if (s + 1 < rows) step(s + 1, rb, ra);
  }
}

// y = the Z split-K partials added in order + bias, then ACT (as k_conv3x3's
// epilogue); n = pixels x Cout
template <int ACT>
__global__ void k_conv_reduce(const float* __restrict__ part, int32_t kz, int64_t n, int32_t Cout,
                              const float* __restrict__ bias, const bf16* __restrict__ mask, bf16* __restrict__ y16,
                              float* __restrict__ y32) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float a = part[i];
  for (int32_t z = 1; z < kz; z++) a += part[(int64_t)z * n + i];
  // 32-bit remainder where n allows (a 64-bit one per element is slower than the loads)
  const int32_t co = n <= 0x7fffffffLL ? (int32_t)((uint32_t)i % (uint32_t)Cout) : (int32_t)(i % Cout);
  float v = a + (bias ? bias[co] : 0.0f);
  if (ACT == 1) v = v > 0.0f ? v : 0.0f;
  if (ACT == 2) v = bf2f(mask[i]) > 0.0f ? v : 0.0f;
  if (y16) y16[i] = f2bf(v);
  if (y32) y32[i] = v;
}

// wt[ci][t'][co] = w[co][8 - t][ci]: the data gradient of a periodic 3x3
// convolution is the convolution of the output gradient with these weights
__global__ void k_flip(const bf16* __restrict__ w, bf16* __restrict__ wt, int32_t Cin, int32_t Cout) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)Cin * 9 * Cout) return;
  const int32_t co = (int32_t)(i % Cout), t = (int32_t)((i / Cout) % 9), ci = (int32_t)(i / (9 * (int64_t)Cout));
  wt[i] = w[((int64_t)co * 9 + (8 - t)) * Cin + ci];
}

// ---------------------------------------------------------------------------
// Weight and bias gradients: dW[co][t][ci] = sum_p dY[p][co] X[p + s_t][ci],
// db[co] = sum_p dY[p][co].  Rows = output channels, columns = input
// channels of one tap, reduction over pixels.  A workgroup (4 waves) owns a
// 64 (co) x 64 (ci) block of one tap and a chunk of the flattened batch's
// pixels; per step of 32 pixels every thread loads one 16-byte run of dY
// (8 channels of a pixel) and one of the tap-shifted X, and stores them
// channel-major into LDS (double buffered: the next step's runs are loaded
// while the waves run the current MFMAs); each wave accumulates a 32 x 32
// sub-block.  The workgroups of tap 4 (no shift) and the first ci block
// also sum their dY runs for the bias.  Each chunk writes its partial
// block to `part` ([chunk][Cout * 9 * Cin + Cout]); k_reduce_parts adds the
// chunks in a fixed order, so the result is deterministic.  Cin % 8 == Cout % 8 == 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wgrad(const bf16* __restrict__ x, const bf16* __restrict__ dy,
                                               float* __restrict__ part, int32_t N, int32_t H, int32_t W,
                                               int32_t Cin, int32_t Cout, int32_t chunk) {
  constexpr int LD = 40;
  __shared__ __attribute__((aligned(16))) bf16 s_dy[2][64 * LD];  // [co][pixel]
  __shared__ __attribute__((aligned(16))) bf16 s_x[2][64 * LD];   // [ci][pixel]
  __shared__ float s_b[32][65];                                    // bias partials [pixel lane][co]
  const int32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t cit = (Cin + 63) / 64;
  int32_t bq = blockIdx.y;
  const int32_t t = bq % 9;
  bq /= 9;
  const int32_t ciB = (bq % cit) * 64, coB = (bq / cit) * 64;
  const bool do_bias = t == 4 && ciB == 0;
  const int64_t npix = (int64_t)N * H * W;
  const int64_t p0 = (int64_t)blockIdx.x * chunk;
  const int64_t pend = p0 + chunk < npix ? p0 + chunk : npix;
  const int32_t dyo = t / 3 - 1, dxo = t % 3 - 1;
  // loader: pixel q of the step, channel group gg (8 channels)
  const int32_t q = tid >> 3, gg = tid & 7;
  const int32_t co_ld = coB + 8 * gg, ci_ld = ciB + 8 * gg;
  const bool co_ok = co_ld < Cout, ci_ok = ci_ld < Cin;
  const bf16x8 zero = {};
  bf16x8 rdy, rx;
  float bsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  auto load = [&](int64_t q0) __attribute__((always_inline)) {
    const int64_t p = q0 + q;
    rdy = zero;
    rx = zero;
    if (p < pend) {
      if (co_ok) rdy = *(const bf16x8*)(dy + p * Cout + co_ld);
      if (ci_ok) {
        const int32_t xx = (int32_t)(p % W);
        const int64_t r = p / W;
        const int32_t yy = (int32_t)(r % H);
        const int64_t im = r / H;
        const int32_t sy = wrapi(yy + dyo, H), sx = wrapi(xx + dxo, W);
        rx = *(const bf16x8*)(x + ((im * H + sy) * W + sx) * Cin + ci_ld);
      }
    }
  };
  auto store = [&](int32_t buf) __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      s_dy[buf][(8 * gg + e) * LD + q] = rdy[e];
      s_x[buf][(8 * gg + e) * LD + q] = rx[e];
    }
    if (do_bias) {
#pragma unroll
      for (int e = 0; e < 8; e++) bsum[e] += bf2f(rdy[e]);
    }
  };
  // consumer: wave (wave / 2, wave % 2) owns co rows 32 * (wave / 2) .., ci 32 * (wave % 2) ..
  const int32_t wr = 32 * (wave >> 1), wc = 32 * (wave & 1);
  const int32_t kq = 8 * (lane >> 4), rl = lane & 15;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
  load(p0);
  store(0);
  __syncthreads();
  int32_t buf = 0;
  for (int64_t q0 = p0; q0 < pend; q0 += 32) {
    const bool more = q0 + 32 < pend;
    if (more) load(q0 + 32);
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) a[i] = *(const bf16x8*)&s_dy[buf][(wr + 16 * i + rl) * LD + kq];
#pragma unroll
    for (int j = 0; j < 2; j++) b[j] = *(const bf16x8*)&s_x[buf][(wc + 16 * j + rl) * LD + kq];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  const int64_t nw = (int64_t)Cout * 9 * Cin;
  float* out = part + (int64_t)blockIdx.x * (nw + Cout);
  // D rows (co) = wr + 16 i + 4 (lane / 16) + e, column (ci) = wc + 16 j + lane % 16
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int32_t ci = ciB + wc + 16 * j + rl;
    if (ci >= Cin) continue;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int32_t co = coB + wr + 16 * i + 4 * (lane >> 4) + e;
        if (co < Cout) out[((int64_t)co * 9 + t) * Cin + ci] = acc[i][j][e];
      }
  }
  if (do_bias) {  // the 32 pixel lanes' sums per channel, added in a fixed order
#pragma unroll
    for (int e = 0; e < 8; e++) s_b[q][8 * gg + e] = bsum[e];
    __syncthreads();
    if (tid < 64) {
      float sb = 0.0f;
      for (int k = 0; k < 32; k++) sb += s_b[k][tid];
      if (coB + tid < Cout) out[nw + coB + tid] = sb;
    }
  }
}

// ---------------------------------------------------------------------------
// Weight and bias gradients over row strips (W % 32 == 0): a workgroup owns a
// CB (co) x IB (ci) block of all nine taps and walks R rows of one 32-pixel
// column strip of one image.  Per row y it loads the dY row (32 pixels x CB)
// and the X row y + 1 (34 pixels, x0 - 1 .. x0 + 32, periodic); X rows y - 1,
// y, y + 1 stay in a four-row LDS ring, so every X and dY value is loaded
// once per strip and serves all nine taps.  LDS rows are channel-major
// (pixels contiguous, the MFMA's k), 40 elements a row, with the 8-pixel
// blocks of columns 0..31 permuted per row (block ^ (row / 8) % 4) so that
// the transposing two-byte stores of eight channel rows fall in different
// banks.  The tap shift dx = -1, 0, +1 is a B operand read at column
// offset 0, 1, 2: one aligned 16-byte read plus the next dword, shifted by
// one element with v_alignbit for the odd offset.  Waves tile the block
// 2 x 2; each holds 9 x (CB/32) x (IB/32) accumulators.  The workgroups of
// the first ci block also sum the dY rows for the bias.  Partials go to
// `part` ([chunk][Cout * 9 * Cin + Cout]) like k_wgrad's.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t wg_col(int32_t row, int32_t c) { return c < 32 ? c ^ (8 * ((row >> 3) & 3)) : c; }

template <int CB, int IB>
__global__ __launch_bounds__(256) void k_wgrad_rows(const bf16* __restrict__ x, const bf16* __restrict__ dy,
                                                    float* __restrict__ part, int32_t N, int32_t H, int32_t W,
                                                    int32_t Cin, int32_t Cout, int32_t R) {
  constexpr int LD = 40;
  constexpr int TI = CB / 32, TJ = IB / 32;  // 16 x 16 tiles per wave along co and ci
  constexpr int DG = CB / 8, XG = IB / 8;    // 16-byte channel groups of a pixel
  constexpr int XITEMS = 34 * XG;            // X row loads (pixel, group)
  __shared__ __attribute__((aligned(16))) bf16 s_x[4][IB * LD];
  __shared__ __attribute__((aligned(16))) bf16 s_dy[2][CB * LD];
  __shared__ float s_b[32][CB + 1];
  const int32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t segs = W / 32, cps = (H + R - 1) / R;
  const int32_t chunk = blockIdx.x;
  const int32_t rc = chunk % cps, sg = (chunk / cps) % segs, n = chunk / (cps * segs);
  const int32_t nci = (Cin + IB - 1) / IB;
  const int32_t ciB = (int32_t)(blockIdx.y % nci) * IB, coB = (int32_t)(blockIdx.y / nci) * CB;
  const int32_t y0 = rc * R, rows = R < H - y0 ? R : H - y0, x0 = sg * 32;
  const bool do_bias = ciB == 0;
  const bf16x8 zero = {};

  // dY loads: thread -> (pixel dpx, group dg), groups fastest (coalesced)
  const bool dy_item = tid < 32 * DG;
  const int32_t dpx = tid / DG, dg = tid % DG;
  const bool dy_ok = dy_item && coB + 8 * dg < Cout;
  const int64_t dy_col = (int64_t)(x0 + dpx) * Cout + coB + 8 * dg;
  // X loads: items tid and tid + 256 -> (column c, group g)
  int32_t xc[2], xg[2];
  bool xok[2], xitem[2];
  int64_t xcol[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int32_t u = tid + 256 * h;
    xitem[h] = u < XITEMS;
    xc[h] = u / XG;
    xg[h] = u % XG;
    xok[h] = xitem[h] && ciB + 8 * xg[h] < Cin;
    xcol[h] = (int64_t)wrapi(x0 + xc[h] - 1, W) * Cin + ciB + 8 * xg[h];
  }
  const int64_t img = (int64_t)n * H;

  struct Regs {
    bf16x8 d, xa, xb;
  };
  // the loads of X row k (image row y0 - 1 + k) and dY row y
  auto load = [&](int32_t k, int32_t y, bool with_dy) __attribute__((always_inline)) {
    Regs r;
    const int64_t xrow = (img + wrapi(y0 - 1 + k, H)) * W * Cin;
    r.xa = xok[0] ? *(const bf16x8*)(x + xrow + xcol[0]) : zero;
    r.xb = xok[1] ? *(const bf16x8*)(x + xrow + xcol[1]) : zero;
    r.d = (with_dy && dy_ok) ? *(const bf16x8*)(dy + (img + y) * W * Cout + dy_col) : zero;
    return r;
  };
  float bsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  auto store_x = [&](const Regs& r, int32_t slot) __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (!xitem[h]) continue;
      const bf16x8 v = h == 0 ? r.xa : r.xb;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int32_t row = 8 * xg[h] + e;
        s_x[slot][row * LD + wg_col(row, xc[h])] = v[e];
      }
    }
  };
  auto store_dy = [&](const Regs& r, int32_t slot) __attribute__((always_inline)) {
    if (!dy_item) return;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int32_t row = 8 * dg + e;
      s_dy[slot][row * LD + wg_col(row, dpx)] = r.d[e];
    }
    if (do_bias) {
#pragma unroll
      for (int e = 0; e < 8; e++) bsum[e] += bf2f(r.d[e]);
    }
  };

  const int32_t wr = (CB / 2) * (wave >> 1), wc = (IB / 2) * (wave & 1);
  const int32_t kq = 8 * (lane >> 4), rl = lane & 15;
  f32x4 acc[9][TI][TJ];
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
      for (int j = 0; j < TJ; j++) acc[t][i][j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};

  // prologue: X rows y0 - 1, y0, y0 + 1 and dY row y0, all loads in flight together
  {
    const Regs r0 = load(0, y0, true), r1 = load(1, y0, false), r2 = load(2, y0, false);
    store_x(r0, 0);
    store_x(r1, 1);
    store_x(r2, 2);
    store_dy(r0, 0);
  }
  __syncthreads();
  for (int32_t s = 0; s < rows; s++) {
    const bool more = s + 1 < rows;
    Regs nx;
    if (more) nx = load(s + 3, y0 + s + 1, true);
    const int32_t ds = s & 1;
    bf16x8 a[TI];
#pragma unroll
    for (int i = 0; i < TI; i++) {
      const int32_t row = wr + 16 * i + rl;
      a[i] = *(const bf16x8*)&s_dy[ds][row * LD + wg_col(row, kq)];
    }
#pragma unroll
    for (int dyi = 0; dyi < 3; dyi++) {
      const bf16* xr = s_x[(s + dyi) & 3];
#pragma unroll
      for (int j = 0; j < TJ; j++) {
        const int32_t row = wc + 16 * j + rl;
        const u32x4 v0 = *(const u32x4*)&xr[row * LD + wg_col(row, kq)];
        const uint32_t v1 = *(const uint32_t*)&xr[row * LD + wg_col(row, kq + 8)];
        const u32x4 vm = v0;  // columns kq .. kq + 7: pixel x + dx for dx = -1
        const u32x4 vz = {__builtin_amdgcn_alignbit(v0[1], v0[0], 16), __builtin_amdgcn_alignbit(v0[2], v0[1], 16),
                          __builtin_amdgcn_alignbit(v0[3], v0[2], 16), __builtin_amdgcn_alignbit(v1, v0[3], 16)};
        const u32x4 vp = {v0[1], v0[2], v0[3], v1};
        const bf16x8 bm = __builtin_bit_cast(bf16x8, vm), bz = __builtin_bit_cast(bf16x8, vz),
                     bp = __builtin_bit_cast(bf16x8, vp);
#pragma unroll
        for (int i = 0; i < TI; i++) {
          acc[3 * dyi + 0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bm, acc[3 * dyi + 0][i][j], 0, 0, 0);
          acc[3 * dyi + 1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bz, acc[3 * dyi + 1][i][j], 0, 0, 0);
          acc[3 * dyi + 2][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bp, acc[3 * dyi + 2][i][j], 0, 0, 0);
        }
      }
    }
    if (more) {
      store_x(nx, (s + 3) & 3);
      store_dy(nx, ds ^ 1);
    }
    __syncthreads();
  }

  const int64_t nw = (int64_t)Cout * 9 * Cin;
  float* out = part + (int64_t)chunk * (nw + Cout);
  // D rows (co) = wr + 16 i + 4 (lane / 16) + e, column (ci) = wc + 16 j + lane % 16
#pragma unroll
  for (int j = 0; j < TJ; j++) {
    const int32_t ci = ciB + wc + 16 * j + rl;
    if (ci >= Cin) continue;
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int32_t co = coB + wr + 16 * i + 4 * (lane >> 4) + e;
        if (co >= Cout) continue;
#pragma unroll
        for (int t = 0; t < 9; t++) out[((int64_t)co * 9 + t) * Cin + ci] = acc[t][i][j][e];
      }
  }
  if (do_bias) {  // the 32 pixel columns' sums per channel, added in a fixed order
    if (dy_item) {
#pragma unroll
      for (int e = 0; e < 8; e++) s_b[dpx][8 * dg + e] = bsum[e];
    }
    __syncthreads();
    if (tid < CB) {
      float sb = 0.0f;
      for (int k = 0; k < 32; k++) sb += s_b[k][tid];
      if (coB + tid < Cout) out[nw + coB + tid] = sb;
    }
  }
}

// dw[i] (i < nw) and db[i - nw] (nw <= i < nw + Cout, when db is given):
// the chunks' partials added in a fixed order.  A workgroup takes 64
// consecutive entries; its 16 groups of 16 threads read them as float4 runs
// from every 16th chunk (group g: chunks g, g + 16, ...), and the groups'
// sums are added in group order.  Rows are nw + Cout floats, a multiple of
// 8 (Cin % 8 == Cout % 8 == 0), so the runs are aligned.
__global__ __launch_bounds__(256) void k_reduce_parts(const float* __restrict__ part, float* __restrict__ dw,
                                                      float* __restrict__ db, int64_t nw, int32_t Cout,
                                                      int32_t nparts) {
  __shared__ float s_sum[16][65];
  const int32_t tid = threadIdx.x, c4 = tid & 15, g = tid >> 4;
  const int64_t stride = nw + Cout;
  const int64_t n = db ? stride : nw;
  const int64_t i0 = (int64_t)blockIdx.x * 64 + 4 * c4;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  if (i0 < n) {
    const float* p = part + i0;
    int32_t k = g;
    // four chunks' loads in flight per step; the adds stay in chunk order
    for (; k + 48 < nparts; k += 64) {
      const f32x4 v0 = *(const f32x4*)(p + (int64_t)k * stride);
      const f32x4 v1 = *(const f32x4*)(p + (int64_t)(k + 16) * stride);
      const f32x4 v2 = *(const f32x4*)(p + (int64_t)(k + 32) * stride);
      const f32x4 v3 = *(const f32x4*)(p + (int64_t)(k + 48) * stride);
      acc += v0;
      acc += v1;
      acc += v2;
      acc += v3;
    }
    for (; k < nparts; k += 16) acc += *(const f32x4*)(p + (int64_t)k * stride);
  }
#pragma unroll
  for (int e = 0; e < 4; e++) s_sum[g][4 * c4 + e] = acc[e];
  __syncthreads();
  if (tid < 64) {
    const int64_t i = (int64_t)blockIdx.x * 64 + tid;
    if (i < n) {
      float s = 0.0f;
#pragma unroll
      for (int q = 0; q < 16; q++) s += s_sum[q][tid];
      if (i < nw) dw[i] = s;
      else db[i - nw] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Elementwise, pooling, resampling
// ---------------------------------------------------------------------------
// g16 = bf16(g32 * (y > 0)) (ReLU gradient through the layer output), or
// just the conversion when y is null
__global__ void k_relu_grad(const float* __restrict__ g32, const bf16* __restrict__ y, bf16* __restrict__ g16,
                            int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float g = g32[i];
  if (y && !(bf2f(y[i]) > 0.0f)) g = 0.0f;
  g16[i] = f2bf(g);
}

// g16 *= (y > 0), in place
__global__ void k_relu_mask(bf16* __restrict__ g16, const bf16* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (!(bf2f(y[i]) > 0.0f)) g16[i] = f2bf(0.0f);
}

__global__ void k_add(const bf16* __restrict__ a, const bf16* __restrict__ b, bf16* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  out[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// 2x2 average pool (H, W even), NHWC
__global__ void k_pool2(const bf16* __restrict__ x, bf16* __restrict__ y, int32_t N, int32_t H, int32_t W, int32_t C) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int32_t h = H / 2, w = W / 2;
  if (i >= (int64_t)N * h * w * C) return;
  const int32_t c = (int32_t)(i % C);
  const int64_t q = i / C;
  const int32_t xx = (int32_t)(q % w), yy = (int32_t)((q / w) % h);
  const int64_t im = q / ((int64_t)w * h);
  const int64_t b = ((im * H + 2 * yy) * W + 2 * xx) * C + c;
  const float s = bf2f(x[b]) + bf2f(x[b + C]) + bf2f(x[b + (int64_t)W * C]) + bf2f(x[b + (int64_t)W * C + C]);
  y[i] = f2bf(0.25f * s);
}

// gradient of the pool: every input pixel gets a quarter of its output's
__global__ void k_pool2_grad(const bf16* __restrict__ dy, bf16* __restrict__ dx, int32_t N, int32_t H, int32_t W,
                             int32_t C) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)N * H * W * C) return;
  const int32_t c = (int32_t)(i % C);
  const int64_t q = i / C;
  const int32_t xx = (int32_t)(q % W), yy = (int32_t)((q / W) % H);
  const int64_t im = q / ((int64_t)W * H);
  dx[i] = f2bf(0.25f * bf2f(dy[((im * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c]));
}

// nearest 2x upsampling of x [N][h][w][C] into [N][2h][2w][C]
__global__ void k_up2(const bf16* __restrict__ x, bf16* __restrict__ y, int32_t N, int32_t h, int32_t w, int32_t C) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int32_t H = 2 * h, W = 2 * w;
  if (i >= (int64_t)N * H * W * C) return;
  const int32_t c = (int32_t)(i % C);
  const int64_t q = i / C;
  const int32_t xx = (int32_t)(q % W), yy = (int32_t)((q / W) % H);
  const int64_t im = q / ((int64_t)W * H);
  y[i] = x[((im * h + yy / 2) * w + xx / 2) * C + c];
}

// its gradient: the sum over each 2x2 block
__global__ void k_up2_grad(const bf16* __restrict__ dy, bf16* __restrict__ dx, int32_t N, int32_t h, int32_t w,
                           int32_t C) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)N * h * w * C) return;
  const int32_t c = (int32_t)(i % C);
  const int64_t q = i / C;
  const int32_t xx = (int32_t)(q % w), yy = (int32_t)((q / w) % h);
  const int64_t im = q / ((int64_t)w * h);
  const int32_t W = 2 * w;
  const int64_t b = ((im * 2 * h + 2 * yy) * W + 2 * xx) * C + c;
  const float s = bf2f(dy[b]) + bf2f(dy[b + C]) + bf2f(dy[b + (int64_t)W * C]) + bf2f(dy[b + (int64_t)W * C + C]);
  dx[i] = f2bf(s);
}

// ---------------------------------------------------------------------------
// Fused passes of the skip connections, 8 channels (16 bytes) a thread
// (C % 8 == 0); each rounds to bf16 where the unfused sequence of the
// kernels above does, so the results are the same bits.
// ---------------------------------------------------------------------------
// out = bf16(up2(x) + skip): x [N][h][w][C] upsampled (nearest) and added
// to skip [N][2h][2w][C] (k_up2 then k_add)
__global__ void k_up2_add(const bf16* __restrict__ x, const bf16* __restrict__ skip, bf16* __restrict__ out,
                          int32_t N, int32_t h, int32_t w, int32_t C) {
  const int32_t C8 = C / 8, H = 2 * h, W = 2 * w;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)N * H * W * C8) return;
  const int32_t c = 8 * (int32_t)(i % C8);
  const int64_t q = i / C8;
  const int32_t xx = (int32_t)(q % W), yy = (int32_t)((q / W) % H);
  const int64_t im = q / ((int64_t)W * H);
  const bf16x8 a = *(const bf16x8*)(x + ((im * h + yy / 2) * w + xx / 2) * C + c);
  const bf16x8 b = *(const bf16x8*)(skip + q * C + c);
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; e++) r[e] = f2bf(bf2f(a[e]) + bf2f(b[e]));
  *(bf16x8*)(out + q * C + c) = r;
}

// dx = bf16(bf16(dy / 4 upsampled) + skip), zero where mask <= 0 (mask
// optional): the pool's gradient plus the skip connection's, through the
// ReLU of the pooled layer (k_pool2_grad, k_add, k_relu_mask)
__global__ void k_pool2_grad_add(const bf16* __restrict__ dy, const bf16* __restrict__ skip,
                                 const bf16* __restrict__ mask, bf16* __restrict__ dx, int32_t N, int32_t H,
                                 int32_t W, int32_t C) {
  const int32_t C8 = C / 8;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)N * H * W * C8) return;
  const int32_t c = 8 * (int32_t)(i % C8);
  const int64_t q = i / C8;
  const int32_t xx = (int32_t)(q % W), yy = (int32_t)((q / W) % H);
  const int64_t im = q / ((int64_t)W * H);
  const bf16x8 d = *(const bf16x8*)(dy + ((im * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c);
  const bf16x8 sk = *(const bf16x8*)(skip + q * C + c);
  bf16x8 m = {};
  if (mask) m = *(const bf16x8*)(mask + q * C + c);
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; e++) {
    bf16 v = f2bf(bf2f(f2bf(0.25f * bf2f(d[e]))) + bf2f(sk[e]));
    if (mask && !(bf2f(m[e]) > 0.0f)) v = f2bf(0.0f);
    r[e] = v;
  }
  *(bf16x8*)(dx + q * C + c) = r;
}

// dx = bf16(sum of dy over each 2x2 block), zero where mask <= 0 (mask
// optional): up2's gradient through the ReLU of the layer below
// (k_up2_grad, k_relu_mask)
__global__ void k_up2_grad_mask(const bf16* __restrict__ dy, const bf16* __restrict__ mask, bf16* __restrict__ dx,
                                int32_t N, int32_t h, int32_t w, int32_t C) {
  const int32_t C8 = C / 8;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)N * h * w * C8) return;
  const int32_t c = 8 * (int32_t)(i % C8);
  const int64_t q = i / C8;
  const int32_t xx = (int32_t)(q % w), yy = (int32_t)((q / w) % h);
  const int64_t im = q / ((int64_t)w * h);
  const int32_t W = 2 * w;
  const int64_t b = ((im * 2 * h + 2 * yy) * W + 2 * xx) * C + c;
  const bf16x8 d0 = *(const bf16x8*)(dy + b), d1 = *(const bf16x8*)(dy + b + C);
  const bf16x8 d2 = *(const bf16x8*)(dy + b + (int64_t)W * C), d3 = *(const bf16x8*)(dy + b + (int64_t)W * C + C);
  bf16x8 m = {};
  if (mask) m = *(const bf16x8*)(mask + q * C + c);
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; e++) {
    bf16 v = f2bf(bf2f(d0[e]) + bf2f(d1[e]) + bf2f(d2[e]) + bf2f(d3[e]));
    if (mask && !(bf2f(m[e]) > 0.0f)) v = f2bf(0.0f);
    r[e] = v;
  }
  *(bf16x8*)(dx + q * C + c) = r;
}

// 2x2 average pool, 8 channels a thread (k_pool2's arithmetic)
__global__ void k_pool2_v8(const bf16* __restrict__ x, bf16* __restrict__ y, int32_t N, int32_t H, int32_t W,
                           int32_t C) {
  const int32_t C8 = C / 8, h = H / 2, w = W / 2;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)N * h * w * C8) return;
  const int32_t c = 8 * (int32_t)(i % C8);
  const int64_t q = i / C8;
  const int32_t xx = (int32_t)(q % w), yy = (int32_t)((q / w) % h);
  const int64_t im = q / ((int64_t)w * h);
  const int64_t b = ((im * H + 2 * yy) * W + 2 * xx) * C + c;
  const bf16x8 a0 = *(const bf16x8*)(x + b), a1 = *(const bf16x8*)(x + b + C);
  const bf16x8 a2 = *(const bf16x8*)(x + b + (int64_t)W * C), a3 = *(const bf16x8*)(x + b + (int64_t)W * C + C);
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; e++) r[e] = f2bf(0.25f * (bf2f(a0[e]) + bf2f(a1[e]) + bf2f(a2[e]) + bf2f(a3[e])));
  *(bf16x8*)(y + q * C + c) = r;
}

// loss = sum over pixels and the first `used` of C channels of
// (pred - target)^2 / count; grad = 2 (pred - target) / count (0 elsewhere).
// Per-block partial sums of the loss.
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target, float* __restrict__ grad,
                      bf16* __restrict__ grad16, float* __restrict__ part, int64_t npix, int32_t C, int32_t used,
                      float inv_count) {
  __shared__ float s[256];
  const int32_t t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * 256;
  const int64_t i = base + t;
  // pixel and channel of i from the block's (one wave-uniform 64-bit divide)
  // and 32-bit arithmetic
  const int64_t pb = base / C;
  const int32_t r = (int32_t)(base - pb * C) + t;
  float l = 0.0f;
  if (i < npix * C) {
    const int32_t c = r % C;
    float g = 0.0f;
    if (c < used) {
      const float d = pred[i] - target[(pb + r / C) * used + c];
      l = d * d * inv_count;
      g = 2.0f * d * inv_count;
    }
    if (grad) grad[i] = g;
    if (grad16) grad16[i] = f2bf(g);
  }
  // the block's sum as a binary tree over its 256 threads (thread t adds
  // t + k for k = 128, 64, ..., 1); the levels below 64 run in wave 0 with
  // shuffles, which pair the same lanes, so the sum is the same bits
  s[t] = l;
  __syncthreads();
  if (t < 128) s[t] += s[t + 128];
  __syncthreads();
  if (t < 64) {
    float v = s[t] + s[t + 64];
#pragma unroll
    for (int k = 32; k > 0; k >>= 1) v += __shfl_down(v, k, 64);
    if (t == 0) part[blockIdx.x] = v;
  }
}

// Adam on FP32 master weights, with the BF16 copy the convolutions read
__global__ void k_adam(float* __restrict__ w, bf16* __restrict__ w16, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = f2bf(wi);
}

// Adam over many tensors in one launch: element i of the concatenation lies
// in the segment k with seg[k].off <= i < seg[k].off + seg[k].n (binary
// search over the nseg offsets).  The arithmetic is k_adam's; a segment
// with wt16 is a convolution's [Cout][9][Cin] weights whose flipped bf16
// copy (k_flip's layout) is written beside w16.
// hyper (optional): lr, c1, c2 read from device memory instead of the
// arguments, so that a captured graph serves every step
__global__ void k_adam_multi(const swe_nn_adam_seg* __restrict__ seg, int32_t nseg, int64_t total, float lr,
                             float b1, float b2, float eps, float c1, float c2, const float* __restrict__ hyper) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  if (hyper) {
    lr = hyper[0];
    c1 = hyper[1];
    c2 = hyper[2];
  }
  // the segment of the block's first element (a wave-uniform search), then
  // advanced for lanes past its end (a block meets few segments)
  const int64_t i0 = (int64_t)blockIdx.x * 256;
  int32_t lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (seg[mid].off <= i0) lo = mid;
    else hi = mid - 1;
  }
  while (lo + 1 < nseg && seg[lo + 1].off <= i) ++lo;
  const swe_nn_adam_seg& sg = seg[lo];
  const int64_t j = i - sg.off;
  const float gi = sg.g[j];
  const float mi = b1 * sg.m[j] + (1.0f - b1) * gi;
  const float vi = b2 * sg.v[j] + (1.0f - b2) * gi * gi;
  sg.m[j] = mi;
  sg.v[j] = vi;
  const float wi = sg.w[j] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  sg.w[j] = wi;
  if (sg.w16) ((bf16*)sg.w16)[j] = f2bf(wi);
  if (sg.wt16) {
    // j = (co * 9 + t) * cin + ci; 32-bit division where the segment allows
    // (a 64-bit divide per element cost more than the update's memory traffic)
    int32_t ci, t, co;
    if (sg.n <= 0x7fffffffLL) {
      const uint32_t ju = (uint32_t)j, cin = (uint32_t)sg.cin, q = ju / cin;
      ci = (int32_t)(ju - q * cin);
      t = (int32_t)(q % 9u);
      co = (int32_t)(q / 9u);
    } else {
      ci = (int32_t)(j % sg.cin);
      t = (int32_t)((j / sg.cin) % 9);
      co = (int32_t)(j / (9 * (int64_t)sg.cin));
    }
    ((bf16*)sg.wt16)[((int64_t)ci * 9 + (8 - t)) * sg.cout + co] = f2bf(wi);
  }
}

// ---------------------------------------------------------------------------
// Input features of the surrogate, computed in float from the float maps and
// rounded to bf16 only at the end (the map's slopes are far below the bf16
// resolution of the heights themselves): x[n][y][x][0..7] = height anomaly,
// sediment, cos and sin of the wind azimuth, periodic central differences
// along the two map axes, the periodic 5-point Laplacian, 0.
// ---------------------------------------------------------------------------
__global__ void k_features(const float* __restrict__ h, const float* __restrict__ s, const float* __restrict__ mean,
                           const float* __restrict__ az, int32_t N, int32_t H, int32_t W, float h_scale,
                           float s_scale, float g_scale, float l_scale, bf16* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)N * H * W) return;
  const int32_t xx = (int32_t)(i % W), yy = (int32_t)((i / W) % H);
  const int64_t n = i / ((int64_t)W * H);
  const float* hn = h + n * H * W;
  const float c = hn[(int64_t)yy * W + xx];
  const float up = hn[(int64_t)wrapi(yy + 1, H) * W + xx], dn = hn[(int64_t)wrapi(yy - 1, H) * W + xx];
  const float rt = hn[(int64_t)yy * W + wrapi(xx + 1, W)], lt = hn[(int64_t)yy * W + wrapi(xx - 1, W)];
  bf16x8 v;
  v[0] = f2bf((c - mean[n]) / h_scale);
  v[1] = f2bf(s ? s[i] / s_scale : 0.0f);
  v[2] = f2bf(cosf(az[n]));
  v[3] = f2bf(sinf(az[n]));
  v[4] = f2bf(0.5f * (up - dn) / g_scale);
  v[5] = f2bf(0.5f * (rt - lt) / g_scale);
  v[6] = f2bf(((up + dn) + (rt + lt) - 4.0f * c) / l_scale);
  v[7] = f2bf(0.0f);
  *(bf16x8*)(x + i * 8) = v;
}

// ---------------------------------------------------------------------------
// Fixed-point mass projection: the predicted change d (float, per cell) is
// turned into int64 units of 2^-32 whose sum is exactly `target`: q_i =
// round(d_i * 2^32 * scale) and the residual target - sum(q) spread as evenly
// as integers allow (the first r cells get one unit more), then added to the
// master.
// ---------------------------------------------------------------------------
// one cell's predicted change in units of 2^-32, |q| < 2^62 (NaN -> 0)
__device__ __forceinline__ int64_t fx_units(float d, float scale) {
  float v = rintf(d * scale * 4294967296.0f);
  v = v > 4.0e18f ? 4.0e18f : (v < -4.0e18f ? -4.0e18f : v);
  return v == v ? (int64_t)v : 0;
}

__global__ void k_fx_sum(const float* __restrict__ d, int64_t n, float scale, int64_t* __restrict__ part) {
  __shared__ int64_t s[256];
  int64_t acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    acc += fx_units(d[i], scale);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_fx_apply(const float* __restrict__ d, int64_t n, float scale, const int64_t* __restrict__ part,
                           int32_t nparts, int64_t target, int64_t* __restrict__ master) {
  int64_t sum = 0;
  for (int32_t k = 0; k < nparts; k++) sum += part[k];
  const int64_t resid = target - sum;
  // resid = qq * n + rr with 0 <= rr < n
  int64_t qq = resid / n, rr = resid % n;
  if (rr < 0) {
    rr += n;
    qq -= 1;
  }
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    master[i] += fx_units(d[i], scale) + qq + (i < rr ? 1 : 0);
}

}  // namespace swenn

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
using namespace swenn;

static inline hipStream_t S(void* s) { return (hipStream_t)s; }
static inline unsigned G(int64_t n) { return (unsigned)((n + 255) / 256); }
#define SWENN_LAUNCH_CHECK() return hipGetLastError() == hipSuccess ? 0 : -2

extern "C" {

// Maps whose rows are whole 32-pixel strips take k_conv_rows for layers of
// at most 32 input channels, and of at most 64 when the output has at most
// 2^SWENN_CONV_ROWS64_LOG2 values (measured on the surrogate's layers,
// MI355X: 32-channel layers at 256^2 45 -> 30 us; with 64 input channels
// the 57 KB of LDS leave two workgroups per CU and every 32-channel output
// block re-reads the rows, which the register-held weights now outweigh up
// to the 64 -> 64 layers at 128^2, 2^23 values): 32 output channels a
// workgroup, each strip cut into chunks of R rows, enough chunks for
// ~SWENN_CONV_WGS workgroups but at most ceil(H / SWENN_CONV_MINROWS) a
// strip.  Others take k_conv3x3, whose output channels per workgroup are the
// layer's (up to 64), 128 where that still leaves two workgroups per CU of
// the chip's 256.
#ifndef SWENN_CONV_ROWS
#define SWENN_CONV_ROWS 1
#endif
#ifndef SWENN_CONV_WGS
#define SWENN_CONV_WGS 2048
#endif
#ifndef SWENN_CONV_MINROWS
#define SWENN_CONV_MINROWS 8
#endif
// split-K factor of k_conv3x3 for a grid of mt x ncob workgroups: layers
// with fewer workgroups than the chip has CUs and a long reduction are split
// so that ~768 workgroups take >= 8 k-steps each (at most 8 parts)
static int32_t conv_split(int64_t mt, int64_t ncob, int32_t ksteps) {
  const int64_t wgs = mt * ncob;
  if (wgs >= 256 || ksteps < 24) return 1;
  int64_t z = (768 + wgs - 1) / wgs;
  if (z > ksteps / 8) z = ksteps / 8;
  if (z > 8) z = 8;
  return z < 2 ? 1 : (int32_t)z;
}

static bool conv_rows_ok(int32_t W, int32_t Cin, int32_t Cout, int64_t P);

// k_conv3x3's grid: pixel blocks, output-channel blocks of nt, split parts
static void conv_grid(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int64_t* mt, int32_t* nt,
                      int32_t* kz) {
  const int64_t P = (int64_t)N * H * W;
  *mt = (P + 127) / 128;
  *nt = Cout <= 16 ? 16 : (Cout <= 32 ? 32 : (Cout <= 64 || *mt * ((Cout + 127) / 128) < 512 ? 64 : 128));
  *kz = conv_split(*mt, (Cout + *nt - 1) / *nt, 9 * ((Cin + 31) / 32));
}

static int conv_launch(const void* x, int32_t N, int32_t H, int32_t W, int32_t Cin, const void* w, const float* bias,
                       int32_t Cout, int act, const void* mask, void* y16, float* y32, void* stream,
                       float* work = nullptr, int64_t work_len = 0) {
  if (!x || !w || N < 1 || H < 1 || W < 1 || Cin < 8 || Cin % 8 || Cout < 1 || (!y16 && !y32)) return -1;
  const int64_t P = (int64_t)N * H * W;
  // k_conv3x3's buffer loads take 32-bit byte offsets below 2^31
  if (P * Cin * 2 >= ((int64_t)1 << 31) || (int64_t)Cout * 9 * Cin * 2 >= ((int64_t)1 << 31)) return -1;
  const bf16* xx = (const bf16*)x;
  const bf16* ww = (const bf16*)w;
  const bf16* mm = (const bf16*)mask;
  bf16* yy = (bf16*)y16;
  if (conv_rows_ok(W, Cin, Cout, P)) {
    const int64_t strips = (int64_t)N * (W / 32), cob = (Cout + 31) / 32;
    int64_t cps = (SWENN_CONV_WGS + strips * cob - 1) / (strips * cob);
    const int64_t cmax = (H + SWENN_CONV_MINROWS - 1) / SWENN_CONV_MINROWS;
    if (cps > cmax) cps = cmax;
    if (cps < 1) cps = 1;
    const int32_t R = (int32_t)((H + cps - 1) / cps);
    const int64_t chunks = strips * ((H + R - 1) / R);
    if (chunks > 0x7fffffffLL || cob > 65535) return -1;
    const dim3 g((unsigned)chunks, (unsigned)cob);
#define SWENN_CONV_R(CI_, AC)                                                                                   \
  hipLaunchKernelGGL((k_conv_rows<CI_, AC>), g, dim3(256), 0, S(stream), xx, ww, bias, mm, yy, y32, N, H, W, Cin, \
                     Cout, R)
#define SWENN_CONV_R_CI(CI_)                \
  if (act == 1) SWENN_CONV_R(CI_, 1);      \
  else if (act == 2) SWENN_CONV_R(CI_, 2); \
  else SWENN_CONV_R(CI_, 0)
    if (Cin <= 32) { SWENN_CONV_R_CI(32); }
    else { SWENN_CONV_R_CI(64); }
#undef SWENN_CONV_R_CI
#undef SWENN_CONV_R
    SWENN_LAUNCH_CHECK();
  }
  int64_t mt;
  int32_t nt, kz;
  conv_grid(N, H, W, Cin, Cout, &mt, &nt, &kz);
  if (mt > 0x7fffffffLL) return -1;
  if (kz > 1 && (!work || work_len < (int64_t)kz * P * Cout)) kz = 1;  // no room for the partials: unsplit
  float* part = kz > 1 ? work : nullptr;
  const dim3 grid((unsigned)mt, (unsigned)((Cout + nt - 1) / nt), (unsigned)kz);
  // split: the parts hold raw sums (the kernel's epilogue is not run)
  const int ka = kz > 1 ? 0 : act;
#define SWENN_CONV(NT_, AC)                                                                                       \
  hipLaunchKernelGGL((k_conv3x3<NT_, AC>), grid, dim3(256), 0, S(stream), xx, ww, bias, mm, yy, y32, N, H, W, Cin, \
                     Cout, part)
#define SWENN_CONV_NT(NT_)            \
  if (ka == 1) SWENN_CONV(NT_, 1);  \
  else if (ka == 2) SWENN_CONV(NT_, 2); \
  else SWENN_CONV(NT_, 0)
  if (nt == 16) { SWENN_CONV_NT(16); }
  else if (nt == 32) { SWENN_CONV_NT(32); }
  else if (nt == 64) { SWENN_CONV_NT(64); }
  else { SWENN_CONV_NT(128); }
#undef SWENN_CONV_NT
#undef SWENN_CONV
  if (kz > 1) {
    const int64_t n = P * Cout;
    if (act == 1)
      hipLaunchKernelGGL(k_conv_reduce<1>, dim3(G(n)), dim3(256), 0, S(stream), part, kz, n, Cout, bias, mm, yy, y32);
    else if (act == 2)
      hipLaunchKernelGGL(k_conv_reduce<2>, dim3(G(n)), dim3(256), 0, S(stream), part, kz, n, Cout, bias, mm, yy, y32);
    else
      hipLaunchKernelGGL(k_conv_reduce<0>, dim3(G(n)), dim3(256), 0, S(stream), part, kz, n, Cout, bias, mm, yy, y32);
  }
  SWENN_LAUNCH_CHECK();
}

#ifndef SWENN_CONV_ROWS64_LOG2
#define SWENN_CONV_ROWS64_LOG2 23
#endif
static bool conv_rows_ok(int32_t W, int32_t Cin, int32_t Cout, int64_t P) {
  return SWENN_CONV_ROWS && W % 32 == 0 &&
         (Cin <= 32 || (Cin <= 64 && (int64_t)Cout * P <= ((int64_t)1 << SWENN_CONV_ROWS64_LOG2)));
}

int64_t swe_nn_conv3x3_work(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout) {
  if (N < 1 || H < 1 || W < 1 || Cin < 1 || Cout < 1) return 0;
  const int64_t P = (int64_t)N * H * W;
  if (conv_rows_ok(W, Cin, Cout, P)) return 0;
  int64_t mt;
  int32_t nt, kz;
  conv_grid(N, H, W, Cin, Cout, &mt, &nt, &kz);
  return kz > 1 ? (int64_t)kz * P * Cout : 0;
}

int swe_nn_conv3x3_ws(const void* x, int32_t N, int32_t H, int32_t W, int32_t Cin, const void* w, const float* bias,
                      int32_t Cout, int32_t act, const void* mask, void* y16, float* y32, float* work,
                      int64_t work_len, void* stream) {
  if (act < 0 || act > 2 || (act == 2 && !mask)) return -1;
  return conv_launch(x, N, H, W, Cin, w, bias, Cout, act, mask, y16, y32, stream, work, work_len);
}

int swe_nn_conv3x3(const void* x, int32_t N, int32_t H, int32_t W, int32_t Cin, const void* w, const float* bias,
                   int32_t Cout, int32_t relu, void* y16, float* y32, void* stream) {
  return conv_launch(x, N, H, W, Cin, w, bias, Cout, relu ? 1 : 0, nullptr, y16, y32, stream);
}

int swe_nn_conv3x3_masked(const void* x, int32_t N, int32_t H, int32_t W, int32_t Cin, const void* w,
                          const float* bias, int32_t Cout, const void* mask, void* y16, void* stream) {
  if (!y16) return -1;
  return conv_launch(x, N, H, W, Cin, w, bias, Cout, mask ? 2 : 0, mask, y16, nullptr, stream);
}

int swe_nn_flip(const void* w, int32_t Cin, int32_t Cout, void* wt, void* stream) {
  if (!w || !wt) return -1;
  hipLaunchKernelGGL(k_flip, dim3(G((int64_t)Cin * 9 * Cout)), dim3(256), 0, S(stream), (const bf16*)w, (bf16*)wt,
                     Cin, Cout);
  SWENN_LAUNCH_CHECK();
}

// How a weight gradient is split.  Maps whose rows are whole 32-pixel
// strips take k_wgrad_rows: CB x IB channel blocks (32; 64 where the layer
// has that many channels when SWENN_WGRAD_BLOCK is 64), each strip cut into
// chunks of R rows, enough chunks for ~SWENN_WGRAD_WGS workgroups but at
// most ceil(H / 8) a strip.  Measured on the surrogate's training layers
// (MI355X, batch 8 at 256^2): 32 x 32 blocks and 1024 workgroups are the
// fastest (64 x 64 blocks hold 356 registers a lane, one wave per SIMD).  Others take k_wgrad: 64 x 64 blocks of
// one tap over chunks of `chunk` pixels (~1024 workgroups, >= 256 pixels, a
// multiple of 32).  Either way the partials are [chunks][Cout*9*Cin + Cout].
#ifndef SWENN_WGRAD_BLOCK
#define SWENN_WGRAD_BLOCK 32
#endif
#ifndef SWENN_WGRAD_WGS
#define SWENN_WGRAD_WGS 1024
#endif
#ifndef SWENN_WGRAD_ROWS
#define SWENN_WGRAD_ROWS 1
#endif
struct WgradPlan {
  bool rows;
  int32_t cb, ib, R;
  int64_t chunk, chunks, blocks;
};

static WgradPlan wgrad_plan(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout) {
  WgradPlan p{};
  const int64_t npix = (int64_t)N * H * W;
  if (SWENN_WGRAD_ROWS && W % 32 == 0) {
    p.rows = true;
    p.cb = Cout >= 64 && SWENN_WGRAD_BLOCK >= 64 ? 64 : 32;
    p.ib = Cin >= 64 && SWENN_WGRAD_BLOCK >= 64 ? 64 : 32;
    p.blocks = (int64_t)((Cout + p.cb - 1) / p.cb) * ((Cin + p.ib - 1) / p.ib);
    const int64_t strips = (int64_t)N * (W / 32);
    int64_t cps = (SWENN_WGRAD_WGS + strips * p.blocks - 1) / (strips * p.blocks);
    const int64_t cmax = (H + 7) / 8;
    if (cps > cmax) cps = cmax;
    if (cps < 1) cps = 1;
    p.R = (int32_t)((H + cps - 1) / cps);
    p.chunks = strips * ((H + p.R - 1) / p.R);
  } else {
    p.blocks = (int64_t)((Cout + 63) / 64) * 9 * ((Cin + 63) / 64);
    int64_t c = (npix * p.blocks + 1023) / 1024;
    c = (c + 31) / 32 * 32;
    p.chunk = c < 256 ? 256 : c;
    p.chunks = (npix + p.chunk - 1) / p.chunk;
  }
  return p;
}

int64_t swe_nn_wgrad_work(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout) {
  if (N < 1 || H < 1 || W < 1 || Cin < 1 || Cout < 1) return 0;
  return wgrad_plan(N, H, W, Cin, Cout).chunks * ((int64_t)Cout * 9 * Cin + Cout);
}

int swe_nn_wgrad(const void* x, const void* dy, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                 float* dw, float* db, float* work, int64_t work_len, void* stream) {
  if (!x || !dy || !dw || !work || (uintptr_t)work % 16) return -1;
  if (N < 1 || H < 1 || W < 1 || Cin < 8 || Cout < 8 || Cin % 8 || Cout % 8) return -1;
  if (work_len < swe_nn_wgrad_work(N, H, W, Cin, Cout)) return -1;
  const WgradPlan p = wgrad_plan(N, H, W, Cin, Cout);
  if (p.chunks > 0x7fffffffLL || p.chunk > 0x7fffffffLL || p.blocks > 65535) return -1;
  const int64_t nw = (int64_t)Cout * 9 * Cin;
  const dim3 grid((unsigned)p.chunks, (unsigned)p.blocks);
  const bf16* xx = (const bf16*)x;
  const bf16* dd = (const bf16*)dy;
  if (p.rows) {
#define SWENN_ROWS_KERNEL(CB_, IB_)                                                                         \
  hipLaunchKernelGGL((k_wgrad_rows<CB_, IB_>), grid, dim3(256), 0, S(stream), xx, dd, work, N, H, W, Cin, Cout, \
                     p.R)
    if (p.cb == 64 && p.ib == 64) SWENN_ROWS_KERNEL(64, 64);
    else if (p.cb == 64) SWENN_ROWS_KERNEL(64, 32);
    else if (p.ib == 64) SWENN_ROWS_KERNEL(32, 64);
    else SWENN_ROWS_KERNEL(32, 32);
#undef SWENN_ROWS_KERNEL
  } else {
    hipLaunchKernelGGL(k_wgrad, grid, dim3(256), 0, S(stream), xx, dd, work, N, H, W, Cin, Cout, (int32_t)p.chunk);
  }
  // one pass adds the chunks' weight and bias partials (contiguous per chunk)
  const int64_t nred = db ? nw + Cout : nw;
  hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((nred + 63) / 64)), dim3(256), 0, S(stream), work, dw, db, nw,
                     Cout, (int32_t)p.chunks);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_relu_grad(const float* g32, const void* y, void* g16, int64_t n, void* stream) {
  if (!g32 || !g16) return -1;
  hipLaunchKernelGGL(k_relu_grad, dim3(G(n)), dim3(256), 0, S(stream), g32, (const bf16*)y, (bf16*)g16, n);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_relu_mask(void* g16, const void* y, int64_t n, void* stream) {
  if (!g16 || !y) return -1;
  hipLaunchKernelGGL(k_relu_mask, dim3(G(n)), dim3(256), 0, S(stream), (bf16*)g16, (const bf16*)y, n);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_add(const void* a, const void* b, void* out, int64_t n, void* stream) {
  hipLaunchKernelGGL(k_add, dim3(G(n)), dim3(256), 0, S(stream), (const bf16*)a, (const bf16*)b, (bf16*)out, n);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_pool2(const void* x, void* y, int32_t N, int32_t H, int32_t W, int32_t C, void* stream) {
  if (!x || !y || H % 2 || W % 2) return -1;
  if (C % 8 == 0)
    hipLaunchKernelGGL(k_pool2_v8, dim3(G((int64_t)N * (H / 2) * (W / 2) * (C / 8))), dim3(256), 0, S(stream),
                       (const bf16*)x, (bf16*)y, N, H, W, C);
  else
    hipLaunchKernelGGL(k_pool2, dim3(G((int64_t)N * (H / 2) * (W / 2) * C)), dim3(256), 0, S(stream), (const bf16*)x,
                       (bf16*)y, N, H, W, C);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_up2_add(const void* x, const void* skip, void* out, int32_t N, int32_t h, int32_t w, int32_t C,
                   void* stream) {
  if (!x || !skip || !out || N < 1 || h < 1 || w < 1 || C < 8 || C % 8) return -1;
  hipLaunchKernelGGL(k_up2_add, dim3(G((int64_t)N * 4 * h * w * (C / 8))), dim3(256), 0, S(stream), (const bf16*)x,
                     (const bf16*)skip, (bf16*)out, N, h, w, C);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_pool2_grad_add(const void* dy, const void* skip, const void* mask, void* dx, int32_t N, int32_t H,
                          int32_t W, int32_t C, void* stream) {
  if (!dy || !skip || !dx || N < 1 || H < 2 || W < 2 || H % 2 || W % 2 || C < 8 || C % 8) return -1;
  hipLaunchKernelGGL(k_pool2_grad_add, dim3(G((int64_t)N * H * W * (C / 8))), dim3(256), 0, S(stream),
                     (const bf16*)dy, (const bf16*)skip, (const bf16*)mask, (bf16*)dx, N, H, W, C);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_up2_grad_mask(const void* dy, const void* mask, void* dx, int32_t N, int32_t h, int32_t w, int32_t C,
                         void* stream) {
  if (!dy || !dx || N < 1 || h < 1 || w < 1 || C < 8 || C % 8) return -1;
  hipLaunchKernelGGL(k_up2_grad_mask, dim3(G((int64_t)N * h * w * (C / 8))), dim3(256), 0, S(stream),
                     (const bf16*)dy, (const bf16*)mask, (bf16*)dx, N, h, w, C);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_pool2_grad(const void* dy, void* dx, int32_t N, int32_t H, int32_t W, int32_t C, void* stream) {
  if (H % 2 || W % 2) return -1;
  hipLaunchKernelGGL(k_pool2_grad, dim3(G((int64_t)N * H * W * C)), dim3(256), 0, S(stream), (const bf16*)dy,
                     (bf16*)dx, N, H, W, C);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_up2(const void* x, void* y, int32_t N, int32_t h, int32_t w, int32_t C, void* stream) {
  hipLaunchKernelGGL(k_up2, dim3(G((int64_t)N * 4 * h * w * C)), dim3(256), 0, S(stream), (const bf16*)x, (bf16*)y,
                     N, h, w, C);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_up2_grad(const void* dy, void* dx, int32_t N, int32_t h, int32_t w, int32_t C, void* stream) {
  hipLaunchKernelGGL(k_up2_grad, dim3(G((int64_t)N * h * w * C)), dim3(256), 0, S(stream), (const bf16*)dy, (bf16*)dx,
                     N, h, w, C);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_mse(const float* pred, const float* target, float* grad, void* grad16, float* part, int64_t npix,
               int32_t C, int32_t used, void* stream) {
  if (!pred || !target || !part || (!grad && !grad16) || npix < 1 || used < 1 || used > C) return -1;
  const float inv = 1.0f / (float)((double)npix * used);
  hipLaunchKernelGGL(k_mse, dim3(G(npix * C)), dim3(256), 0, S(stream), pred, target, grad, (bf16*)grad16, part, npix,
                     C, used, inv);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_adam(float* w, void* w16, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                float eps, int32_t step, void* stream) {
  if (step < 1) return -1;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(G(n)), dim3(256), 0, S(stream), w, (bf16*)w16, g, m, v, n, lr, b1, b2, eps, c1, c2);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_adam_multi(const swe_nn_adam_seg* seg, int32_t nseg, int64_t total, float lr, float b1, float b2,
                      float eps, int32_t step, void* stream) {
  if (!seg || nseg < 1 || total < 0 || step < 1) return -1;
  if (total == 0) return 0;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam_multi, dim3(G(total)), dim3(256), 0, S(stream), seg, nseg, total, lr, b1, b2, eps, c1, c2,
                     (const float*)nullptr);
  SWENN_LAUNCH_CHECK();
}

void swe_nn_adam_hyper(float lr, float b1, float b2, int32_t step, float* out) {
  out[0] = lr;
  out[1] = 1.0f - powf(b1, (float)step);
  out[2] = 1.0f - powf(b2, (float)step);
}

int swe_nn_adam_multi_dev(const swe_nn_adam_seg* seg, int32_t nseg, int64_t total, const float* hyper, float b1,
                          float b2, float eps, void* stream) {
  if (!seg || !hyper || nseg < 1 || total < 0) return -1;
  if (total == 0) return 0;
  hipLaunchKernelGGL(k_adam_multi, dim3(G(total)), dim3(256), 0, S(stream), seg, nseg, total, 0.0f, b1, b2, eps, 0.0f,
                     0.0f, hyper);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_features(const float* h, const float* s, const float* mean, const float* az, int32_t N, int32_t H,
                    int32_t W, float h_scale, float s_scale, float g_scale, float l_scale, void* x, void* stream) {
  if (!h || !mean || !az || !x || N < 1 || H < 1 || W < 1) return -1;
  hipLaunchKernelGGL(k_features, dim3(G((int64_t)N * H * W)), dim3(256), 0, S(stream), h, s, mean, az, N, H, W,
                     h_scale, s_scale, g_scale, l_scale, (bf16*)x);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_project_mass(const float* d, int64_t n, float scale, int64_t target, int64_t* master, int64_t* work,
                        void* stream) {
  if (!d || !master || !work || n < 1) return -1;
  const int32_t nparts = 256;
  hipLaunchKernelGGL(k_fx_sum, dim3(nparts), dim3(256), 0, S(stream), d, n, scale, work);
  hipLaunchKernelGGL(k_fx_apply, dim3(nparts), dim3(256), 0, S(stream), d, n, scale, (const int64_t*)work, nparts,
                     target, master);
  SWENN_LAUNCH_CHECK();
}

}  // extern "C"
