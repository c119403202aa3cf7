/* libwavefd_lcodec: learned snapshot codec for the Born/RTM gradient's forward snapshots.
 *
 * No counterpart in the reference: its snapshot store is exact or CvxCompress
 * (src/compressedio.jl:1-148; CvxCompress is un-vendored), and exact storage stays the
 * default.  The codec: a fixed quantiser (8^3 blocks, bfloat16 block scale rounded up,
 * 11 levels, 9 base-11 digits per 32-bit word, one choice bit per block: 3.59 bits per value,
 * 8.9x smaller than float32)
 * and a learned decoder, out = q_hat + g f(q_hat / g), f(q) = (F(q) - F(-q)) / 2 with F a
 * bias-free 3-layer 3x3x3 convolutional network (16 channels, bfloat16 weights and activations,
 * float32 accumulation on the matrix cores), g the code's largest block scale, the correction
 * clamped to half a quantisation step (the exact value's cell).
 *
 *   fields     (n1 fastest: nz, ny, nx) float32, device memory
 *   volumes    npatch volumes of X x Y x Z (z fastest), device memory
 *   parameters per layer W[t][co][c] (27 x 16 x 16, t = (dx+1) 9 + (dy+1) 3 + (dz+1)) then 16
 *              biases, three layers (conv1 reads channel 0, conv3 writes channel 0)
 *
 * Work is ordered on `stream` (NULL: the null stream).  Functions return 0 on success and -1
 * on error (message: wavefd_lcodec_LastError), unless stated otherwise.
 */
#ifndef WAVEFD_LCODEC_H
#define WAVEFD_LCODEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* quantiser */
long wavefd_lcodec_code_words(long n1, long n2, long n3);  /* code size in 32-bit words */
long wavefd_lcodec_encode(const float *x, long n1, long n2, long n3, uint32_t *code, void *stream);
long wavefd_lcodec_dequantize(const uint32_t *code, float *out, void *stream);  /* plain decode */
/* the plain decode of a code of an n1 x n2 x n3 field, without reading its header (no host wait) */
long wavefd_lcodec_dequantize_n(const uint32_t *code, long n1, long n2, long n3, float *out, void *stream);
/* the same quantiser with levels = 11, 5 or 3 (q = rint((levels - 1) x / (2 s)); 9, 13 or 20 digits
 * per word: 3.56, 2.5 or 1.63 bits per value); wavefd_lcodec_dequantize and _halfstep read the
 * levels from the code's header */
long wavefd_lcodec_code_words_l(long n1, long n2, long n3, long levels);
long wavefd_lcodec_encode_l(const float *x, long n1, long n2, long n3, long levels, uint32_t *code, void *stream);
long wavefd_lcodec_dequantize_nl(const uint32_t *code, long n1, long n2, long n3, long levels, float *out,
                                 void *stream);
float wavefd_lcodec_norm(const uint32_t *code);  /* the code's largest block scale; -1 on error */
/* half a quantisation step of every value's block, s / (levels - 1): |x - x_hat| is at most that */
long wavefd_lcodec_halfstep(const uint32_t *code, float *out, void *stream);
/* out = ca a + cb b + cc c over n floats (device; b and c may be NULL), evaluated in that order:
 * the predictions, residuals and reconstructions of the Born driver's predictive snapshot store,
 * the same call on the encoding and the decoding side; no counterpart in the reference */
long wavefd_lcodec_combine(long n, float ca, const float *a, float cb, const float *b, float cc, const float *c,
                           float *out, void *stream);
/* the same store's encoder pass (the values the combine calls above would give): pred = ca a + cb b
 * + cc c and res = x - pred in one pass */
long wavefd_lcodec_predict_residual(long n, float ca, const float *a, float cb, const float *b, float cc,
                                    const float *c, const float *x, float *pred, float *res, void *stream);

/* the decoder network */
long wavefd_lcodec_nparams(void);
void *wavefd_lcodec_net_create(void);  /* NULL on failure */
void wavefd_lcodec_net_free(void *net);
long wavefd_lcodec_net_get(void *net, float *params);        /* host array of nparams */
long wavefd_lcodec_net_set(void *net, const float *params);  /* host; resets the optimiser */
long wavefd_lcodec_net_init(void *net, unsigned long seed);  /* He-normal init, last layer 0 */
/* out = q + norm f(q / norm), the correction clamped to +-hs (one half step per voxel) when hs
 * is not NULL (hs may be out) */
long wavefd_lcodec_forward(void *net, const float *q, long npatch, long X, long Y, long Z, float norm, const float *hs,
                           float *out, void *stream);
/* mean squared error of the forward output against target, its parameter gradient (kept for
 * wavefd_lcodec_adam; copied to the host array grad when not NULL; no gradient flows where the
 * clamp binds) */
long wavefd_lcodec_grad(void *net, const float *q, const float *target, long npatch, long X, long Y, long Z,
                        float norm, const float *hs, float *out, float *loss, float *grad, void *stream);
long wavefd_lcodec_adam(void *net, float lr, void *stream);  /* one Adam step, beta 0.9 / 0.999 */
/* encode with the encoder's per-block choice: quantise, run the learned decode, set the blocks'
 * bits where it is closer to x than the plain dequantisation (scratch1/2: the field's size) */
long wavefd_lcodec_encode_learned(void *net, const float *x, long n1, long n2, long n3, uint32_t *code, float *scratch1,
                                  float *scratch2, void *stream);
/* the same with levels = 11, 5 or 3 (wavefd_lcodec_encode_l's quantiser) */
long wavefd_lcodec_encode_learned_l(void *net, const float *x, long n1, long n2, long n3, long levels, uint32_t *code,
                                    float *scratch1, float *scratch2, void *stream);
/* learned decode of a code (scratch: a device array of the field's size); blocks whose bit is
 * clear take the plain dequantisation */
long wavefd_lcodec_decode(void *net, const uint32_t *code, float *scratch, float *out, void *stream);
/* training patches: dst[i] = scale x the P^3 cube of src at origins[3 i .. 3 i + 2] (n1, n2, n3
 * order, host array) */
long wavefd_lcodec_gather(const float *src, long n1, long n2, long n3, const long *origins, long npatch, long P,
                          float scale, float *dst, void *stream);

const char *wavefd_lcodec_LastError(void);

#ifdef __cplusplus
}
#endif
#endif
