/*
 * libspacetime -- C-ABI of the MI355X build of the reference's src/spacetime.cc (target in
 * src/CMakeLists.txt:13): source injection, receiver extraction and 8-tap time
 * interpolation, with the reference's names, argument order and C types.
 *
 * Injection / extraction run where the field lives: host arrays on the host (the reference's
 * semantics), device arrays (hipMalloc'd, e.g. from Prop3D*_GetDevicePtr) on the GPU; points
 * outside a device field are rejected (see spacetime_LastError).  Data arrays are Julia
 * (nt, nr) column-major: data[(ir-1)*nt + it-1]; it, iu and ir are 1-based.
 */
#ifndef WAVEFD_AMD_SPACETIME_H
#define WAVEFD_AMD_SPACETIME_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/spacetime.cc:132-136 == src/spacetime.jl:793-797 (24 bytes) */
struct SourcePoint32 {
    long iu;
    long ir;
    float c;
};
/* src/spacetime.cc:138-142 == src/spacetime.jl:799-803 */
struct SourcePoint64 {
    long iu;
    long ir;
    double c;
};

/* src/spacetime.cc:144-162 -- field[iu-1] += c * data[(ir-1)*nt + it-1] */
void injectdata_float(float *field, float *data, size_t it, size_t nt, struct SourcePoint32 **partitions,
                      size_t npartitions, long *npoints_per_partition, size_t nthreads);
/* No counterpart in the reference.  injectdata_float for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): the same accumulation, bit
 * for bit, without any device or host allocation.  field and data must be device memory of the
 * current device; the points are staged in a fixed device array of at most 1,048,576 points in at
 * most 4,096 partitions (larger inputs are rejected, see spacetime_LastError).  Every call on a GPU
 * uses the same array: calls that overlap on one GPU race. */
void injectdata_float_noalloc(float *field, float *data, size_t it, size_t nt, struct SourcePoint32 **partitions,
                              size_t npartitions, long *npoints_per_partition, size_t nthreads);
/* src/spacetime.cc:164-182 */
void injectdata_double(double *field, double *data, size_t it, size_t nt, struct SourcePoint64 **partitions,
                       size_t npartitions, long *npoints_per_partition, size_t nthreads);
/* src/spacetime.cc:184-204 -- data[(ir-1)*nt + it-1] += c * field[iu-1] (accumulates) */
void extractdata_float(float *data, float *field, struct SourcePoint32 **partitions, size_t npartitions,
                       long *npoints_per_partition, size_t it, size_t nt, size_t nthreads);
/* src/spacetime.cc:206-226 */
void extractdata_double(double *data, double *field, struct SourcePoint64 **partitions, size_t npartitions,
                        long *npoints_per_partition, size_t it, size_t nt, size_t nthreads);
/* src/spacetime.cc:228-237 (coarse d -> fine m) */
void interpadjoint_1d_float(float **h, float *m, float *d, size_t nh, size_t nm, size_t nd);
/* src/spacetime.cc:239-248 */
void interpadjoint_1d_double(double **h, double *m, double *d, size_t nh, size_t nm, size_t nd);
/* src/spacetime.cc:250-259 (fine m -> coarse d) */
void interpforward_1d_float(float **h, float *d, float *m, size_t nh, size_t nd, size_t nm);
/* src/spacetime.cc:261-270 */
void interpforward_1d_double(double **h, double *d, double *m, size_t nh, size_t nd, size_t nm);
/* src/spacetime.cc:272-283 */
void interpadjoint_nd_float(float **h, float *m, float *d, size_t nh, size_t nm, size_t nd, size_t nslow,
                            size_t nthreads);
/* src/spacetime.cc:285-296 */
void interpadjoint_nd_double(double **h, double *m, double *d, size_t nh, size_t nm, size_t nd, size_t nslow,
                             size_t nthreads);
/* src/spacetime.cc:298-309 */
void interpforward_nd_float(float **h, float *d, float *m, size_t nh, size_t nd, size_t nm, size_t nslow,
                            size_t nthreads);
/* src/spacetime.cc:311-322 */
void interpforward_nd_double(double **h, double *d, double *m, size_t nh, size_t nd, size_t nm, size_t nslow,
                             size_t nthreads);

/* No counterpart in the reference: message of the last device-path failure ("" if none). */
const char *spacetime_LastError(void);

#ifdef __cplusplus
}
#endif
#endif
