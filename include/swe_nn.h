/*
 * swe_nn.h -- C-ABI of libswe_nn.so: the kernels of the learned surrogate
 * of the erosion (no counterpart in the reference, which has no surrogate;
 * the surrogate predicts K map-cycles of World::erode, world.hpp:36-77, at
 * once).  All pointers are device pointers; `stream` is a hipStream_t (NULL
 * = the default stream).  Activations are NHWC bf16, weights [Cout][9][Cin]
 * bf16 (tap t = 3*(dy+1) + (dx+1)), padding periodic.  Every call returns 0,
 * -1 for bad arguments or -2 for a launch error.
 */
#ifndef SWE_NN_H
#define SWE_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y = conv3x3(x, w) + bias, optional ReLU; y16 (bf16) and/or y32 (float).
 * Cin % 8 == 0, x and w each under 2 GiB (else -1); BF16 MFMA with FP32
 * accumulation. */
int swe_nn_conv3x3(const void* x, int32_t N, int32_t H, int32_t W, int32_t Cin, const void* w, const float* bias,
                   int32_t Cout, int32_t relu, void* y16, float* y32, void* stream);
/* y16 = bf16(conv3x3(x, w) + bias), zero where mask <= 0 (mask [N][H][W][Cout]
 * bf16, optional; bias optional): a data gradient through the ReLU of the
 * layer that produced mask, in one pass.  The same bits as swe_nn_conv3x3
 * into y32 followed by swe_nn_relu_grad. */
int swe_nn_conv3x3_masked(const void* x, int32_t N, int32_t H, int32_t W, int32_t Cin, const void* w,
                          const float* bias, int32_t Cout, const void* mask, void* y16, void* stream);
/* The same convolution with a caller's workspace: act 0 none, 1 ReLU, 2
 * zero where mask <= 0.  Layers with few pixel blocks and a long reduction
 * (low resolutions, many channels) split the reduction over up to 8
 * workgroups per block, whose partial sums (work: swe_nn_conv3x3_work
 * floats, 0 when the layer is not split) are added in a fixed order; with
 * less work than that the layer runs unsplit. */
int64_t swe_nn_conv3x3_work(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout);
int swe_nn_conv3x3_ws(const void* x, int32_t N, int32_t H, int32_t W, int32_t Cin, const void* w, const float* bias,
                      int32_t Cout, int32_t act, const void* mask, void* y16, float* y32, float* work,
                      int64_t work_len, void* stream);
/* wt[ci][t][co] = w[co][8-t][ci]: the weights whose forward convolution of
 * dy is the data gradient dx (Cin and Cout swap). */
int swe_nn_flip(const void* w, int32_t Cin, int32_t Cout, void* wt, void* stream);
/* floats of work memory swe_nn_wgrad needs */
int64_t swe_nn_wgrad_work(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout);
/* dw = sum_p dy[p] x[p + t]^T (float [Cout][9][Cin]); db = sum_p dy[p]
 * (float [Cout], optional); Cin % 8 == Cout % 8 == 0; deterministic
 * (per-chunk partials in `work`, 16-byte aligned, added in a fixed order). */
int swe_nn_wgrad(const void* x, const void* dy, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                 float* dw, float* db, float* work, int64_t work_len, void* stream);
/* g16 = bf16(g32 * (y > 0)); y may be NULL (conversion only) */
int swe_nn_relu_grad(const float* g32, const void* y, void* g16, int64_t n, void* stream);
/* g16 *= (y > 0) in place (bf16) */
int swe_nn_relu_mask(void* g16, const void* y, int64_t n, void* stream);
int swe_nn_add(const void* a, const void* b, void* out, int64_t n, void* stream);
/* 2x2 average pooling and its gradient; nearest 2x upsampling and its gradient */
int swe_nn_pool2(const void* x, void* y, int32_t N, int32_t H, int32_t W, int32_t C, void* stream);
int swe_nn_pool2_grad(const void* dy, void* dx, int32_t N, int32_t H, int32_t W, int32_t C, void* stream);
int swe_nn_up2(const void* x, void* y, int32_t N, int32_t h, int32_t w, int32_t C, void* stream);
int swe_nn_up2_grad(const void* dy, void* dx, int32_t N, int32_t h, int32_t w, int32_t C, void* stream);
/* Fused skip-connection passes (C % 8 == 0), the same bits as the unfused
 * calls: out = up2(x) + skip (swe_nn_up2, swe_nn_add); dx = pool2_grad(dy)
 * + skip, zero where mask <= 0 (swe_nn_pool2_grad, swe_nn_add,
 * swe_nn_relu_mask; mask optional); dx = up2_grad(dy), zero where mask <= 0
 * (swe_nn_up2_grad, swe_nn_relu_mask; mask optional).  h, w: the low
 * resolution; H, W: the high one. */
int swe_nn_up2_add(const void* x, const void* skip, void* out, int32_t N, int32_t h, int32_t w, int32_t C,
                   void* stream);
int swe_nn_pool2_grad_add(const void* dy, const void* skip, const void* mask, void* dx, int32_t N, int32_t H,
                          int32_t W, int32_t C, void* stream);
int swe_nn_up2_grad_mask(const void* dy, const void* mask, void* dx, int32_t N, int32_t h, int32_t w, int32_t C,
                         void* stream);
/* mean squared error over the first `used` of C channels (target [npix][used]);
 * its gradient by pred [npix][C] as float (grad) and/or bf16 (grad16, the
 * bf16 rounding of grad), at least one of them; part[ceil(npix*C/256)]
 * partial loss sums */
int swe_nn_mse(const float* pred, const float* target, float* grad, void* grad16, float* part, int64_t npix,
               int32_t C, int32_t used, void* stream);
/* Adam on float master weights w (+ bf16 copy w16, optional); step >= 1 */
int swe_nn_adam(float* w, void* w16, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                float eps, int32_t step, void* stream);
/* One segment of swe_nn_adam_multi: n float master values w (elements off ..
 * off + n - 1 of the concatenation), their gradient g and moments m, v; the
 * bf16 copy w16 (optional); for convolution weights [cout][9][cin], the
 * flipped bf16 copy wt16 [cin][9][cout] of swe_nn_flip (optional). */
typedef struct {
  float* w;
  void* w16;
  void* wt16;
  const float* g;
  float* m;
  float* v;
  int64_t off, n;
  int32_t cin, cout;
} swe_nn_adam_seg;
/* Adam (swe_nn_adam's arithmetic) over every segment in one launch; seg is
 * a device array of nseg segments in order of off, total = sum of n. */
int swe_nn_adam_multi(const swe_nn_adam_seg* seg, int32_t nseg, int64_t total, float lr, float b1, float b2,
                      float eps, int32_t step, void* stream);
/* The same with lr and the bias corrections read from device memory at
 * hyper[0..2] (for a step captured once in a graph); swe_nn_adam_hyper
 * fills a host array with (lr, 1 - b1^step, 1 - b2^step) exactly as
 * swe_nn_adam_multi computes them. */
int swe_nn_adam_multi_dev(const swe_nn_adam_seg* seg, int32_t nseg, int64_t total, const float* hyper, float b1,
                          float b2, float eps, void* stream);
void swe_nn_adam_hyper(float lr, float b1, float b2, int32_t step, float* out);
/* Surrogate input features x [N][H][W][8] bf16 from float maps h, s [N][H][W]
 * (s may be NULL), per-map mean heights and wind azimuths [N]: (h - mean) /
 * h_scale, s / s_scale, cos az, sin az, periodic central differences along
 * H and W over g_scale, periodic 5-point Laplacian over l_scale, 0. */
int swe_nn_features(const float* h, const float* s, const float* mean, const float* az, int32_t N, int32_t H,
                    int32_t W, float h_scale, float s_scale, float g_scale, float l_scale, void* x, void* stream);
/* master[i] += q_i with q = round(d * scale * 2^32) corrected so that
 * sum(q) == target exactly (fixed-point mass projection); work: 256 int64 */
int swe_nn_project_mass(const float* d, int64_t n, float scale, int64_t target, int64_t* master, int64_t* work,
                        void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SWE_NN_H */
