/* libwavefd_ginsu: ginsu windows and the shared image of batched Born imaging.
 *
 * No counterpart in the reference: WaveFD leaves windowing to its callers (src/WaveFD.jl:52,
 * README.md:20), and its Born callers (JetPackWaveFD, out of tree) cut and paste model windows
 * on the host.  These entry points do it on the device, with integer-exact window rules:
 *
 *   arrays     (nz, ny, nx) float32, Julia column-major, k = (ix*ny + iy)*nz + iz
 *   window     [oz, oz+snz) x [oy, oy+sny) x [ox, ox+snx), wholly inside the big grid
 *   extract    sub[i] = big[window origin + i], bit-exact copy
 *   accumulate big[window origin + i] += sub[i] for i in [lz0,lz1) x [ly0,ly1) x [lx0,lx1),
 *              one float add per cell; a fixed call order gives a bit-reproducible image
 *
 * Device pointers only; work is ordered on `stream` (a propagator's GetStream, or NULL).
 * Functions return 0 on success and -1 on error (message: wavefd_ginsu_LastError).
 */
#ifndef WAVEFD_GINSU_H
#define WAVEFD_GINSU_H

#ifdef __cplusplus
extern "C" {
#endif

long wavefd_ginsu_extract(const float *big, long bnz, long bny, long bnx, long oz, long oy, long ox, float *sub,
                          long snz, long sny, long snx, void *stream);
long wavefd_ginsu_accumulate(float *big, long bnz, long bny, long bnx, long oz, long oy, long ox, const float *sub,
                             long snz, long sny, long snx, long lz0, long lz1, long ly0, long ly1, long lx0, long lx1,
                             void *stream);

/* device helpers for drivers without an allocator of their own (the Born driver) */
void *wavefd_device_alloc(long bytes);                                  /* NULL on failure */
void wavefd_device_free(void *p);
long wavefd_device_copy(void *dst, const void *src, long bytes, void *stream); /* any direction */
long wavefd_device_zero(void *dst, long bytes, void *stream);
long wavefd_device_sync(void *stream);                                  /* NULL: whole device */

const char *wavefd_ginsu_LastError(void);

#ifdef __cplusplus
}
#endif
#endif
