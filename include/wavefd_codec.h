/*
 * libwavefd_codec -- lossy block codec for wavefield snapshots.  No counterpart in the
 * reference's C sources: it stands in for the un-vendored CvxCompress codec behind the
 * reference's Compressor UInt32 mode (src/compressedio.jl:1-148), for callers that store
 * forward snapshots for Born/RTM imaging.  Exact storage (the Compressor's raw modes) is
 * unaffected.
 *
 * A volume n1*n2*n3 (n1 fastest) is cut into b1*b2*b3 blocks; each block stores its max |x|
 * (float32) and every value quantised to `bits` (8 or 16) signed levels of that maximum, so
 * |x - x'| <= max_block / (2 (2^(bits-1) - 1)).  Arrays may be host or device memory.
 */
#ifndef WAVEFD_CODEC_H
#define WAVEFD_CODEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 32-bit words of a compressed volume (buffer size for _compress); -1 on bad geometry */
long wavefd_codec_words(const long *n, const long *b, long bits);
/* compress; returns the words written, -1 on error */
long wavefd_codec_compress(const float *field, const long *n, const long *b, long bits, uint32_t *out);
/* decompress nwords words into field; 0 on success, -1 on error */
long wavefd_codec_decompress(const uint32_t *in, long nwords, float *field);
/* message of the last failure ("" if none) */
const char *wavefd_codec_LastError(void);

#ifdef __cplusplus
}
#endif
#endif
