/*
 * hpt_nn.h — kernels of the learned denoiser (no counterpart in the
 * reference, whose denoise.py shells out to Intel OIDN): a U-Net over the
 * renderer's colour, albedo and normal images, trained and run on MI355X.
 *
 * Tensors are NHWC in device memory.  Activations are bf16 (uint16 bit
 * patterns); gradients, weights' masters, biases and optimizer state are
 * float32.  A convolution weight tensor is [Cout][k*k][Cin] (Cin innermost),
 * the layout the MFMA fragments read; Cin and the padded Cout are multiples
 * of 16.  Every launcher queues work on `stream` (a hipStream_t, 0 = the
 * null stream) and returns 0, or a non-zero HIP error code for a bad launch.
 */
#ifndef HPT_NN_H
#define HPT_NN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPT_NN_API __attribute__((visibility("default")))

/* y = act(conv_k(x, w) + b): k in {1, 3} (zero padding k/2), x bf16
 * [n][h][w][cin], w bf16 [cout_pad][k*k][cin], y bf16 [n][h][w][cout] with
 * cout <= cout_pad; relu != 0 applies max(., 0).  If y_f32 is non-null the
 * pre-rounding values are also written there as float32 [n][h][w][cout]
 * (the network output); y may then be null. */
HPT_NN_API int hpt_nn_conv_forward(void* stream, const uint16_t* x, const uint16_t* w, const float* b,
                                   uint16_t* y, float* y_f32, int n, int h, int wd, int cin, int cout,
                                   int cout_pad, int k, int relu);

/* hpt_nn_conv_forward (3x3) that also writes the 2 x 2 average pooling of y
 * (the next encoder level's input, as hpt_nn_pool_forward computes it) into
 * y_pool (n x h/2 x w/2 x cout bf16); h and w even.  Same bits as the two
 * calls; the tiled kernel writes both from one pass over the output. */
HPT_NN_API int hpt_nn_conv_forward_pool(void* stream, const uint16_t* x, const uint16_t* w, const float* b,
                                        uint16_t* y, uint16_t* y_pool, int n, int h, int wd, int cin, int cout,
                                        int cout_pad, int relu);

/* hpt_nn_conv_forward with k = 3 on the input concat(upsample2x(low), skip)
 * without materialising it: `low` bf16 [n][h/2][wd/2][c_low] read with
 * nearest upsampling, `skip` bf16 [n][h][wd][c_skip] (h and wd even, channel
 * counts multiples of 8); the weights see input channels [0, c_low) from
 * `low` and [c_low, c_low + c_skip) from `skip` -- the channel order of
 * hpt_nn_upsample_forward followed by hpt_nn_concat_forward(up, skip), whose
 * result this equals bit for bit.  The U-Net decoder's first convolution of
 * each level at inference. */
HPT_NN_API int hpt_nn_conv_forward_upcat(void* stream, const uint16_t* low, const uint16_t* skip,
                                         const uint16_t* w, const float* b, uint16_t* y, float* y_f32, int n,
                                         int h, int wd, int c_low, int c_skip, int cout, int cout_pad, int relu);

/* dx = conv_k^T(dy, w): dy float32 [n][h][w][cout_pad] (channels >= cout
 * zero, already multiplied by the activation's derivative), wt bf16
 * [cin][k*k][cout_pad] = w transposed and spatially flipped
 * (hpt_nn_prepare_weights), dx float32 [n][h][w][cin]. */
HPT_NN_API int hpt_nn_conv_backward_data(void* stream, const float* dy, const uint16_t* wt, float* dx, int n,
                                         int h, int wd, int cin, int cout, int cout_pad, int k);

/* dw[co][t][ci] = sum over pixels of dy[p][co] * x[p + offset(t)][ci] and
 * db[co] = sum of dy[p][co] for co < cout, float32, written (not
 * accumulated); dy as for hpt_nn_conv_backward_data.  `partial` is scratch
 * of hpt_nn_wgrad_scratch_floats(...) floats. */
HPT_NN_API int hpt_nn_conv_backward_weights(void* stream, const float* dy, const uint16_t* x, float* dw, float* db,
                                            float* partial, int n, int h, int wd, int cin, int cout, int cout_pad,
                                            int k);
HPT_NN_API size_t hpt_nn_wgrad_scratch_floats(int n, int h, int wd, int cin, int cout_pad, int k);

/* Master float32 weights [cout][k*k][cin] -> bf16 w [cout_pad][k*k][cin]
 * (rows >= cout zero) and bf16 wt [cin][k*k][cout_pad] (flipped taps). */
HPT_NN_API int hpt_nn_prepare_weights(void* stream, const float* master, uint16_t* w, uint16_t* wt, int cin,
                                      int cout, int cout_pad, int k);

/* dy *= (y > 0) for a ReLU output y (bf16); count elements. */
HPT_NN_API int hpt_nn_relu_backward(void* stream, float* dy, const uint16_t* y, size_t count);

/* 2x2 average pooling and its gradient; nearest 2x upsampling and its
 * gradient (sum of the four children).  (h, wd) is always the size of the
 * forward input x: pooling maps h x wd (even) to h/2 x wd/2, upsampling h x
 * wd to 2h x 2wd. */
HPT_NN_API int hpt_nn_pool_forward(void* stream, const uint16_t* x, uint16_t* y, int n, int h, int wd, int c);
HPT_NN_API int hpt_nn_pool_backward(void* stream, const float* dy, float* dx, int n, int h, int wd, int c);
HPT_NN_API int hpt_nn_upsample_forward(void* stream, const uint16_t* x, uint16_t* y, int n, int h, int wd, int c);
HPT_NN_API int hpt_nn_upsample_backward(void* stream, const float* dy, float* dx, int n, int h, int wd, int c);

/* y[p] = concat(a[p] (ca channels), b[p] (cb channels)); the gradient split. */
HPT_NN_API int hpt_nn_concat_forward(void* stream, const uint16_t* a, const uint16_t* b, uint16_t* y,
                                     size_t pixels, int ca, int cb);
HPT_NN_API int hpt_nn_concat_backward(void* stream, const float* dy, float* da, float* db, size_t pixels, int ca,
                                      int cb);

/* dx += dy elementwise (a skip path's gradient joining the main path). */
HPT_NN_API int hpt_nn_accumulate(void* stream, float* dx, const float* dy, size_t count);

/* Network input of one image from renderer images (float4 per pixel,
 * src_w x src_h): the h x wd window at (x0, y0), zero outside the source,
 * channels log(1 + max(colour, 0)), albedo, normal * 0.5 + 0.5, `aux` (a
 * per-image constant such as the sample count's feature; 0 outside the
 * source), zeros up to 16, as bf16 into x [h][wd][16] (one image of a
 * batch).  `mirror` bit 0 flips the window left-right, bit 1 top-bottom
 * (training augmentation). */
HPT_NN_API int hpt_nn_make_input(void* stream, const float* color, const float* albedo, const float* normal,
                                 int src_w, int src_h, int x0, int y0, uint16_t* x, int h, int wd, int mirror,
                                 float aux);

/* Training target of one image: log(1 + max(colour, 0)) of the same window
 * (and mirror) of a float4 reference image into out float32 [h][wd][3]. */
HPT_NN_API int hpt_nn_make_target(void* stream, const float* color, int src_w, int src_h, int x0, int y0,
                                  float* out, int h, int wd, int mirror);

/* out[p][c] += x[p][c] for c < 3: the network input's log colour (bf16,
 * x_channels per pixel) added to the output, so the network predicts a
 * correction of its input (no gradient flows to x). */
HPT_NN_API int hpt_nn_add_residual(void* stream, float* out, const uint16_t* x, size_t pixels, int x_channels);

/* L1 loss on log(1 + colour): out = network output float32 [p][3],
 * target_log = log(1 + reference colour) float32 [p][3]; dy [p][cout_pad]
 * = sign(out - target) / (3 * pixels) (channels >= 3 zero); the mean |.| is
 * added to *loss_sum (float32, device). */
HPT_NN_API int hpt_nn_l1_loss(void* stream, const float* out, const float* target_log, float* dy,
                              float* loss_sum, size_t pixels, int cout_pad);

/* Adam (Kingma & Ba 2015) on float32: step t >= 1. */
HPT_NN_API int hpt_nn_adam(void* stream, float* p, const float* g, float* m, float* v, size_t count, float lr,
                           float beta1, float beta2, float eps, int t);

/* Denoised colour: max(expm1(network output), 0) of [p][3] written into a
 * float4 image of dst_w x dst_h at (x0, y0) for the h x wd region (pixels
 * outside the image skipped), alpha 1. */
HPT_NN_API int hpt_nn_write_output(void* stream, const float* out, float* dst, int dst_w, int dst_h, int x0,
                                   int y0, int h, int wd);

/* Device memory, copies and timing for the denoiser's host driver, on the
 * HIP runtime this library links (a process must not mix in a second HIP
 * runtime, such as the one a PyTorch wheel bundles, for these buffers).
 * hpt_nn_alloc returns zeroed memory on the current device. */
HPT_NN_API int hpt_nn_alloc(size_t bytes, void** ptr);
HPT_NN_API int hpt_nn_free(void* ptr);
/* hipMemcpyAsync with the direction inferred from the addresses. */
HPT_NN_API int hpt_nn_copy(void* stream, void* dst, const void* src, size_t bytes);
/* p[i] = value for count floats. */
HPT_NN_API int hpt_nn_fill(void* stream, float* p, float value, size_t count);
HPT_NN_API int hpt_nn_stream_synchronize(void* stream);
/* A non-blocking stream for the denoiser's work (never the legacy null
 * stream, whose work would wait for every blocking stream and which HIP
 * refuses while another thread captures a graph).  Lives until the process
 * ends. */
HPT_NN_API int hpt_nn_stream_create(void** stream);
HPT_NN_API int hpt_nn_event_create(void** event);
HPT_NN_API int hpt_nn_event_destroy(void* event);
HPT_NN_API int hpt_nn_event_record(void* event, void* stream);
/* Milliseconds between two recorded events (synchronises on the second). */
HPT_NN_API int hpt_nn_event_elapsed(void* start, void* stop, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* HPT_NN_H */
