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

namespace swenn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f2bf(float v) { return (bf16)v; }  // round to nearest even

__device__ __forceinline__ int32_t wrapi(int32_t v, int32_t n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

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
// ---------------------------------------------------------------------------
template <int NT, int ACT>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16* __restrict__ x, const bf16* __restrict__ w,
                                                 const float* __restrict__ bias, bf16* __restrict__ y16,
                                                 float* __restrict__ y32, int32_t N, int32_t H, int32_t W,
                                                 int32_t Cin, int32_t Cout) {
  constexpr int WM = NT >= 64 ? 2 : 4;  // waves along the pixels
  constexpr int WN = 4 / WM;            // ... and the channels
  constexpr int TM = 128 / WM, TN = NT / WN;
  constexpr int BM = TM / 16, BN = TN / 16;
  constexpr int LD = 40;
  constexpr int BROWS = NT < 64 ? NT : 64;  // B rows a pass of 64 threads' row index covers
  __shared__ __attribute__((aligned(16))) bf16 sA[2][128 * LD];
  __shared__ __attribute__((aligned(16))) bf16 sB[2][NT * LD];
  const int32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t wm = wave / WN, wn = wave % WN;
  const int64_t P = (int64_t)N * H * W;
  const int64_t p0 = (int64_t)blockIdx.x * 128;
  const int32_t co0 = blockIdx.y * NT;
  const int32_t cchunks = (Cin + 31) / 32, ksteps = 9 * cchunks;
  const int32_t g = tid & 3, r0 = tid >> 2;  // loader: row r0 (+64), channel group g

  // the two A rows of this thread: image, row, column
  int32_t ai[2], ay[2], ax[2];
  bool aok[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int64_t p = p0 + r0 + 64 * h;
    aok[h] = p < P;
    const int64_t pc = aok[h] ? p : 0;
    ax[h] = (int32_t)(pc % W);
    ay[h] = (int32_t)((pc / W) % H);
    ai[h] = (int32_t)(pc / ((int64_t)W * H));
  }
  const bf16x8 zero = {};
  bf16x8 ra[2], rb[2];
  auto load = [&](int32_t s) __attribute__((always_inline)) {
    const int32_t tap = s / cchunks, ci = (s - tap * cchunks) * 32 + 8 * g;
    const int32_t dy = tap / 3 - 1, dx = tap % 3 - 1;
    const bool cok = ci < Cin;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      ra[h] = zero;
      if (aok[h] && cok) {
        const int32_t yy = wrapi(ay[h] + dy, H), xx = wrapi(ax[h] + dx, W);
        ra[h] = *(const bf16x8*)(x + (((int64_t)ai[h] * H + yy) * W + xx) * Cin + ci);
      }
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      rb[h] = zero;
      const int32_t row = r0 + 64 * h;
      const int32_t co = co0 + row;
      if (row < NT && (h == 0 || NT > 64) && co < Cout && cok) rb[h] = *(const bf16x8*)(w + ((int64_t)co * 9 + tap) * Cin + ci);
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
  (void)BROWS;

  f32x4 acc[BM][BN];
#pragma unroll
  for (int i = 0; i < BM; i++)
#pragma unroll
    for (int j = 0; j < BN; j++) acc[i][j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};

  load(0);
  store(0);
  __syncthreads();
  const int32_t kq = 8 * (lane >> 4), rl = lane & 15;
  for (int32_t s = 0; s < ksteps; s++) {
    const int32_t buf = s & 1;
    if (s + 1 < ksteps) load(s + 1);
    bf16x8 a[BM], b[BN];
#pragma unroll
    for (int i = 0; i < BM; i++) a[i] = *(const bf16x8*)&sA[buf][(wm * TM + 16 * i + rl) * LD + kq];
#pragma unroll
    for (int j = 0; j < BN; j++) b[j] = *(const bf16x8*)&sB[buf][(wn * TN + 16 * j + rl) * LD + kq];
#pragma unroll
    for (int i = 0; i < BM; i++)
#pragma unroll
      for (int j = 0; j < BN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (s + 1 < ksteps) store(buf ^ 1);
    __syncthreads();
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
        if (y16) y16[p * Cout + co] = f2bf(v);
        if (y32) y32[p * Cout + co] = v;
      }
  }
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
// chunks in order, so the result is deterministic.  Cin % 8 == Cout % 8 == 0.
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

// dw[i] (i < nw) and db[i - nw] (nw <= i < nw + Cout, when db is given):
// the chunks' partials added in chunk order
__global__ void k_reduce_parts(const float* __restrict__ part, float* __restrict__ dw, float* __restrict__ db,
                               int64_t nw, int32_t Cout, int32_t nparts) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t stride = nw + Cout;
  if (i >= (db ? stride : nw)) return;
  float s = 0.0f;
  for (int32_t k = 0; k < nparts; k++) s += part[(int64_t)k * stride + i];
  if (i < nw) dw[i] = s;
  else db[i - nw] = s;
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

// loss = sum over pixels and the first `used` of C channels of
// (pred - target)^2 / count; grad = 2 (pred - target) / count (0 elsewhere).
// Per-block partial sums of the loss.
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target, float* __restrict__ grad,
                      float* __restrict__ part, int64_t npix, int32_t C, int32_t used, float inv_count) {
  __shared__ float s[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float l = 0.0f;
  if (i < npix * C) {
    const int32_t c = (int32_t)(i % C);
    float g = 0.0f;
    if (c < used) {
      const float d = pred[i] - target[(i / C) * used + c];
      l = d * d * inv_count;
      g = 2.0f * d * inv_count;
    }
    grad[i] = g;
  }
  s[threadIdx.x] = l;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
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

int swe_nn_conv3x3(const void* x, int32_t N, int32_t H, int32_t W, int32_t Cin, const void* w, const float* bias,
                   int32_t Cout, int32_t relu, void* y16, float* y32, void* stream) {
  if (!x || !w || N < 1 || H < 1 || W < 1 || Cin < 8 || Cin % 8 || Cout < 1 || (!y16 && !y32)) return -1;
  const int64_t P = (int64_t)N * H * W;
  const int64_t mt = (P + 127) / 128;
  if (mt > 0x7fffffffLL) return -1;
  // output channels per workgroup: the layer's (up to 64), 128 where that
  // still leaves two workgroups per CU of the chip's 256
  const int32_t nt = Cout <= 16 ? 16 : (Cout <= 32 ? 32 : (Cout <= 64 || mt * ((Cout + 127) / 128) < 512 ? 64 : 128));
  const dim3 grid((unsigned)mt, (unsigned)((Cout + nt - 1) / nt));
  const bf16* xx = (const bf16*)x;
  const bf16* ww = (const bf16*)w;
  bf16* yy = (bf16*)y16;
#define SWENN_CONV(NT_, AC)                                                                                   \
  hipLaunchKernelGGL((k_conv3x3<NT_, AC>), grid, dim3(256), 0, S(stream), xx, ww, bias, yy, y32, N, H, W, Cin, \
                     Cout)
#define SWENN_CONV_NT(NT_) \
  if (relu) SWENN_CONV(NT_, 1); else SWENN_CONV(NT_, 0)
  if (nt == 16) { SWENN_CONV_NT(16); }
  else if (nt == 32) { SWENN_CONV_NT(32); }
  else if (nt == 64) { SWENN_CONV_NT(64); }
  else { SWENN_CONV_NT(128); }
#undef SWENN_CONV_NT
#undef SWENN_CONV
  SWENN_LAUNCH_CHECK();
}

int swe_nn_flip(const void* w, int32_t Cin, int32_t Cout, void* wt, void* stream) {
  if (!w || !wt) return -1;
  hipLaunchKernelGGL(k_flip, dim3(G((int64_t)Cin * 9 * Cout)), dim3(256), 0, S(stream), (const bf16*)w, (bf16*)wt,
                     Cin, Cout);
  SWENN_LAUNCH_CHECK();
}

// pixels per workgroup of the weight gradient: enough chunks for ~1024
// workgroups, at least 256 pixels, a multiple of 32
static int64_t wgrad_chunk(int64_t npix, int32_t Cin, int32_t Cout) {
  const int64_t tiles = (int64_t)((Cout + 63) / 64) * 9 * ((Cin + 63) / 64);
  int64_t c = (npix * tiles + 1023) / 1024;
  c = (c + 31) / 32 * 32;
  return c < 256 ? 256 : c;
}

int64_t swe_nn_wgrad_work(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout) {
  const int64_t npix = (int64_t)N * H * W;
  const int64_t chunk = wgrad_chunk(npix, Cin, Cout);
  const int64_t chunks = (npix + chunk - 1) / chunk;
  return chunks * ((int64_t)Cout * 9 * Cin + Cout);
}

int swe_nn_wgrad(const void* x, const void* dy, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                 float* dw, float* db, float* work, int64_t work_len, void* stream) {
  if (!x || !dy || !dw || !work) return -1;
  if (Cin < 8 || Cout < 8 || Cin % 8 || Cout % 8) return -1;
  if (work_len < swe_nn_wgrad_work(N, H, W, Cin, Cout)) return -1;
  const int64_t npix = (int64_t)N * H * W;
  const int64_t chunk = wgrad_chunk(npix, Cin, Cout);
  const int64_t chunks = (npix + chunk - 1) / chunk;
  if (chunks > 0x7fffffffLL || chunk > 0x7fffffffLL) return -1;
  const int64_t nw = (int64_t)Cout * 9 * Cin;
  const dim3 grid((unsigned)chunks, (unsigned)(9 * ((Cin + 63) / 64) * ((Cout + 63) / 64)));
  hipLaunchKernelGGL(k_wgrad, grid, dim3(256), 0, S(stream), (const bf16*)x, (const bf16*)dy, work, N, H, W, Cin,
                     Cout, (int32_t)chunk);
  // one pass adds the chunks' weight and bias partials (contiguous per chunk)
  hipLaunchKernelGGL(k_reduce_parts, dim3(G(db ? nw + Cout : nw)), dim3(256), 0, S(stream), work, dw, db, nw, Cout,
                     (int32_t)chunks);
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
  if (H % 2 || W % 2) return -1;
  hipLaunchKernelGGL(k_pool2, dim3(G((int64_t)N * (H / 2) * (W / 2) * C)), dim3(256), 0, S(stream), (const bf16*)x,
                     (bf16*)y, N, H, W, C);
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

int swe_nn_mse(const float* pred, const float* target, float* grad, float* part, int64_t npix, int32_t C,
               int32_t used, void* stream) {
  if (used < 1 || used > C) return -1;
  const float inv = 1.0f / (float)((double)npix * used);
  hipLaunchKernelGGL(k_mse, dim3(G(npix * C)), dim3(256), 0, S(stream), pred, target, grad, part, npix, C, used, inv);
  SWENN_LAUNCH_CHECK();
}

int swe_nn_adam(float* w, void* w16, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                float eps, int32_t step, void* stream) {
  if (step < 1) return -1;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(G(n)), dim3(256), 0, S(stream), w, (bf16*)w16, g, m, v, n, lr, b1, b2, eps, c1, c2);
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
