// TEST INFRASTRUCTURE -- builds the reference's own CPU half of FidelityFX CAS
// (shaders/ffx_a.h + shaders/ffx_cas.h with A_CPU) into oracle/_ref/libffx_ref.so so the tests
// can check the oracle's CasSetup restatement against it. The headers are read from
// /root/reference/shaders at build time; nothing of them is copied into this repository.
#include <stdint.h>
#define A_CPU 1
#define A_GCC 1
#include "ffx_a.h"
#include "ffx_cas.h"

extern "C" void ref_cas_setup(uint32_t c0[4], uint32_t c1[4], float sharpness, float inX, float inY,
                              float outX, float outY) {
    CasSetup(c0, c1, sharpness, inX, inY, outX, outY);
}

extern "C" uint32_t ref_half_bits(float f) { return AU1_AH1_AF1(f); }
