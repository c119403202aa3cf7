// Test infrastructure: compiles the reference's Q-profile setup (src/absorb.h, included from where
// it lies under the reference tree; see the Makefile's `ref` target) into oracle/_ref, so that the
// oracle's wo_setup_q / wo2_setup_q can be checked against it bit for bit.  Nothing of the product
// links this.  The two wrappers only give the header's functions C linkage; freqQ >= FLT_EPSILON is
// the caller's duty (the reference exits the process below it).
#include "absorb.h"

extern "C" {
void ref_setup_q_3d(long fs, long nx, long ny, long nz, long nsponge, float dt, float freqQ, float qMin,
                    float qInterior, float *q) {
    setupDtOmegaInvQ_3D(fs, nx, ny, nz, nsponge, 1, dt, freqQ, qMin, qInterior, q);
}
void ref_setup_q_2d(long fs, long nx, long nz, long nsponge, float dt, float freqQ, float qMin, float qInterior,
                    float *q) {
    setupDtOmegaInvQ_2D(fs, nx, nz, nsponge, 1, dt, freqQ, qMin, qInterior, q);
}
}
