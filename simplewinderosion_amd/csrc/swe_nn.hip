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
// 3x3 convolution, implicit GEMM: rows = pixels, columns = output channels,
// reduction k = tap * Cin + ci (Cin % 8 == 0, so a lane's 8 consecutive k
// stay within one tap).  A workgroup (4 waves) takes 64 consecutive pixels
// of one image row; wave w owns pixels 16w .. 16w+15 and 16*NACC output
// channels starting at blockIdx.y * 16 * NACC.  Per 16x16x32 MFMA, lane l
// supplies A row l%16 (a pixel) and B column l%16 (an output channel), both
// at k = 8*(l/16) .. +7, and receives D rows 4*(l/16) .. +3 of column l%16.
// ---------------------------------------------------------------------------
template <int NACC, int ACT>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16* __restrict__ x, const bf16* __restrict__ w,
                                                 const float* __restrict__ bias, bf16* __restrict__ y16,
                                                 float* __restrict__ y32, int32_t H, int32_t W, int32_t Cin,
                                                 int32_t Cout) {
  const int32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int32_t segs = (W + 63) / 64;
  const int32_t seg = blockIdx.x % segs;
  const int32_t row = (blockIdx.x / segs) % H;
  const int32_t img = blockIdx.x / (segs * H);
  const int32_t co0 = blockIdx.y * 16 * NACC;
  const int32_t px = seg * 64 + wave * 16 + (lane & 15);  // this lane's A row (pixel)
  const int32_t kq = (lane >> 4) * 8;                     // this lane's k offset in a step
  const int32_t K = 9 * Cin;
  const int32_t pxc = px < W ? px : W - 1;                // a valid pixel for the loads
  const bf16x8 zero = {};

  f32x4 acc[NACC];
#pragma unroll
  for (int j = 0; j < NACC; j++) acc[j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};

  for (int32_t k0 = 0; k0 < K; k0 += 32) {
    const int32_t k = k0 + kq;
    bf16x8 a = zero;
    int32_t tap = 0, ci = 0;
    const bool kin = k < K;
    if (kin) {
      tap = k / Cin;
      ci = k - tap * Cin;
      const int32_t yy = wrapi(row + tap / 3 - 1, H), xx = wrapi(pxc + tap % 3 - 1, W);
      a = *(const bf16x8*)(x + (((int64_t)img * H + yy) * W + xx) * Cin + ci);
    }
#pragma unroll
    for (int j = 0; j < NACC; j++) {
      const int32_t co = co0 + 16 * j + (lane & 15);
      bf16x8 b = zero;
      if (kin && co < Cout) b = *(const bf16x8*)(w + ((int64_t)co * 9 + tap) * Cin + ci);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    }
  }
  // D rows 4*(lane/16) + i of this wave's 16 pixels, column lane%16
#pragma unroll
  for (int j = 0; j < NACC; j++) {
    const int32_t co = co0 + 16 * j + (lane & 15);
    if (co >= Cout) continue;
    const float b = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int32_t p = seg * 64 + wave * 16 + (lane >> 4) * 4 + i;
      if (p >= W) continue;
      float v = acc[j][i] + b;
      if (ACT == 1) v = v > 0.0f ? v : 0.0f;
      const int64_t o = (((int64_t)img * H + row) * W + p) * Cout + co;
      if (y16) y16[o] = f2bf(v);
      if (y32) y32[o] = v;
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
// Weight gradient: dW[co][t][ci] = sum_p dY[p][co] * X[p shifted by t][ci].
// Rows = output channels, columns = input channels, reduction over pixels.
// Workgroup (4 waves) = one tap, a 32 (co) x 32 (ci) block, and a chunk of
// `chunk` pixels of the flattened batch.  Per step of 32 pixels, half the
// threads load dY and half the shifted X as 16-byte runs of 8 channels
// (coalesced NHWC rows) and store them channel-major into LDS; each wave
// then reads its MFMA operands -- 8 consecutive pixels of one channel -- as
// 16-byte LDS reads and accumulates its 16x16 sub-block (w / 2, w % 2).
// Partial sums per chunk go to `part`, reduced by k_reduce_parts
// (deterministic order).  Cin % 8 == 0 and Cout % 8 == 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wgrad(const bf16* __restrict__ x, const bf16* __restrict__ dy,
                                               float* __restrict__ part, int32_t N, int32_t H, int32_t W,
                                               int32_t Cin, int32_t Cout, int32_t chunk) {
  __shared__ __attribute__((aligned(16))) bf16 s_dy[32][40];  // [co][pixel], rows padded against bank conflicts
  __shared__ __attribute__((aligned(16))) bf16 s_x[32][40];   // [ci][pixel]
  const int32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t cib = (Cin + 31) / 32;
  int32_t b = blockIdx.y;
  const int32_t t = b % 9;
  b /= 9;
  const int32_t ciB = (b % cib) * 32, coB = (b / cib) * 32;
  const int32_t npix = N * H * W;  // < 2^31 (checked by the launcher)
  const int32_t p0 = (int32_t)blockIdx.x * chunk;
  const int32_t pend = p0 + chunk < npix ? p0 + chunk : npix;
  const int32_t dyo = t / 3 - 1, dxo = t % 3 - 1;
  // loader role: threads 0..127 load dY, 128..255 load X; (pixel q, group g)
  const bool ldy = tid < 128;
  const int32_t q = (tid & 127) >> 2, g = tid & 3;
  const int32_t cl = (ldy ? coB : ciB) + 8 * g;  // first channel of this thread's run
  const bool cl_ok = cl < (ldy ? Cout : Cin);
  // consumer role
  const int32_t co_l = (wave >> 1) * 16 + (lane & 15), ci_l = (wave & 1) * 16 + (lane & 15);
  const int32_t kq = (lane >> 4) * 8;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int32_t q0 = p0; q0 < pend; q0 += 32) {
    const int32_t p = q0 + q;
    bf16x8 v = {};
    if (p < pend && cl_ok) {
      if (ldy) {
        v = *(const bf16x8*)(dy + (int64_t)p * Cout + cl);
      } else {
        const int32_t xx = p % W, r = p / W;
        const int32_t yy = r % H, im = r / H;
        const int32_t sy = wrapi(yy + dyo, H), sx = wrapi(xx + dxo, W);
        v = *(const bf16x8*)(x + (((int64_t)im * H + sy) * W + sx) * Cin + cl);
      }
    }
    bf16(*dst)[40] = ldy ? s_dy : s_x;
#pragma unroll
    for (int e = 0; e < 8; e++) dst[8 * g + e][q] = v[e];
    __syncthreads();
    const bf16x8 a = *(const bf16x8*)&s_dy[co_l][kq];
    const bf16x8 bb = *(const bf16x8*)&s_x[ci_l][kq];
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc, 0, 0, 0);
    __syncthreads();
  }
  // D rows = coB + 16*(wave/2) + 4*(lane/16) + i, column = ciB + 16*(wave%2) + lane%16
  const int32_t ci = ciB + ci_l;
  if (ci >= Cin) return;
  const int64_t nw = (int64_t)Cout * 9 * Cin;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int32_t c = coB + (wave >> 1) * 16 + (lane >> 4) * 4 + i;
    if (c < Cout) part[(int64_t)blockIdx.x * nw + ((int64_t)c * 9 + t) * Cin + ci] = acc[i];
  }
}

__global__ void k_reduce_parts(const float* __restrict__ part, float* __restrict__ out, int64_t n, int32_t nparts) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int32_t k = 0; k < nparts; k++) s += part[(int64_t)k * n + i];
  out[i] = s;
}

// bias gradient partials: part[chunk][c] = sum over the chunk's pixels;
// the 256 threads form 256 / C rows of C channels, each row summing every
// (256 / C)-th pixel, then the rows are added in LDS (fixed order)
__global__ __launch_bounds__(256) void k_bgrad(const bf16* __restrict__ dy, float* __restrict__ part, int64_t npix,
                                               int32_t C, int32_t chunk) {
  __shared__ float s_sum[256];
  const int64_t p0 = (int64_t)blockIdx.x * chunk;
  const int64_t pend = p0 + chunk < npix ? p0 + chunk : npix;
  if (C > 256) {  // one thread per channel, several channels per thread
    for (int32_t c = threadIdx.x; c < C; c += 256) {
      float s = 0.0f;
      for (int64_t p = p0; p < pend; p++) s += bf2f(dy[p * C + c]);
      part[(int64_t)blockIdx.x * C + c] = s;
    }
    return;
  }
  const int32_t rows = 256 / C;
  const int32_t c = threadIdx.x % C, r = threadIdx.x / C;
  float s = 0.0f;
  if (r < rows)
    for (int64_t p = p0 + r; p < pend; p += rows) s += bf2f(dy[p * C + c]);
  s_sum[threadIdx.x] = s;
  __syncthreads();
  if (r == 0) {
    float tot = 0.0f;
    for (int32_t k = 0; k < rows; k++) tot += s_sum[k * C + c];
    part[(int64_t)blockIdx.x * C + c] = tot;
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
  const int32_t nacc = Cout >= 64 ? 4 : (Cout >= 32 ? 2 : 1);
  const dim3 grid((unsigned)(((W + 63) / 64) * H * N), (unsigned)((Cout + 16 * nacc - 1) / (16 * nacc)));
  const bf16* xx = (const bf16*)x;
  const bf16* ww = (const bf16*)w;
  bf16* yy = (bf16*)y16;
#define SWENN_CONV(NA, AC) \
  hipLaunchKernelGGL((k_conv3x3<NA, AC>), grid, dim3(256), 0, S(stream), xx, ww, bias, yy, y32, H, W, Cin, Cout)
  if (nacc == 4) {
    if (relu) SWENN_CONV(4, 1); else SWENN_CONV(4, 0);
  } else if (nacc == 2) {
    if (relu) SWENN_CONV(2, 1); else SWENN_CONV(2, 0);
  } else {
    if (relu) SWENN_CONV(1, 1); else SWENN_CONV(1, 0);
  }
#undef SWENN_CONV
  SWENN_LAUNCH_CHECK();
}

int swe_nn_flip(const void* w, int32_t Cin, int32_t Cout, void* wt, void* stream) {
  if (!w || !wt) return -1;
  hipLaunchKernelGGL(k_flip, dim3(G((int64_t)Cin * 9 * Cout)), dim3(256), 0, S(stream), (const bf16*)w, (bf16*)wt,
                     Cin, Cout);
  SWENN_LAUNCH_CHECK();
}

int64_t swe_nn_wgrad_work(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout) {
  const int64_t npix = (int64_t)N * H * W;
  const int64_t chunks = (npix + kSweNnChunk - 1) / kSweNnChunk;
  const int64_t nw = (int64_t)Cout * 9 * Cin;
  return chunks * (nw > Cout ? nw : Cout);
}

int swe_nn_wgrad(const void* x, const void* dy, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                 float* dw, float* db, float* work, int64_t work_len, void* stream) {
  if (!x || !dy || !dw || !work) return -1;
  if (Cin < 8 || Cout < 8 || Cin % 8 || Cout % 8) return -1;
  if ((int64_t)N * H * W >= ((int64_t)1 << 31) - kSweNnChunk) return -1;
  if (work_len < swe_nn_wgrad_work(N, H, W, Cin, Cout)) return -1;
  const int64_t npix = (int64_t)N * H * W;
  const int32_t chunks = (int32_t)((npix + kSweNnChunk - 1) / kSweNnChunk);
  const int64_t nw = (int64_t)Cout * 9 * Cin;
  const dim3 grid((unsigned)chunks, (unsigned)(9 * ((Cin + 31) / 32) * ((Cout + 31) / 32)));
  hipLaunchKernelGGL(k_wgrad, grid, dim3(256), 0, S(stream), (const bf16*)x, (const bf16*)dy, work, N, H, W, Cin,
                     Cout, kSweNnChunk);
  hipLaunchKernelGGL(k_reduce_parts, dim3(G(nw)), dim3(256), 0, S(stream), work, dw, nw, chunks);
  if (db) {
    // smaller pixel chunks for the bias: chunks * (kSweNnChunk / kBiasChunk)
    // partials of Cout floats fit the work size (9 * Cin >= 16)
    constexpr int32_t kBiasChunk = kSweNnChunk / 16;
    const int32_t bchunks = (int32_t)((npix + kBiasChunk - 1) / kBiasChunk);
    hipLaunchKernelGGL(k_bgrad, dim3((unsigned)bchunks), dim3(256), 0, S(stream), (const bf16*)dy, work, npix, Cout,
                       kBiasChunk);
    hipLaunchKernelGGL(k_reduce_parts, dim3(G(Cout)), dim3(256), 0, S(stream), work, db, (int64_t)Cout, bchunks);
  }
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
