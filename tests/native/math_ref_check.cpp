// Host check of the acos / atan / atan2 restatement in rp3d_math.h against the C library's acosf /
// atanf / atan2f (the functions the reference build calls): every float in [-1, 1] for acos at a
// stride, a stride over all finite floats for atan, and xorshift-random pairs for atan2, both in the
// hinge's range (sin >= 0, cos in [-1, 1]) and over random bit patterns.  Prints the mismatch counts.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "rp3d_math.h"

using rp3dg::fbits_as;
using rp3dg::fbits_of;

static uint64_t st = 88172645463325252ull;
static uint32_t rnd() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 16); }
static bool same(float a, float b) { return fbits_of(a) == fbits_of(b) || (isnan(a) && isnan(b)); }

int main(int argc, char** argv) {
    const uint32_t stride = argc > 1 ? (uint32_t)atoi(argv[1]) : 1;
    long bad_acos = 0, bad_atan = 0, bad_atan2 = 0, n_acos = 0, n_atan = 0, n_atan2 = 0;
    for (uint64_t u = 0; u <= 0x3f800001u; u += stride)
        for (uint32_t sg = 0; sg < 2; sg++) {
            const float x = fbits_as((int32_t)((uint32_t)u | (sg << 31)));
            n_acos++;
            if (!same(rp3dg::ref_acosf(x), acosf(x))) bad_acos++;
        }
    for (uint64_t u = 0; u <= 0x7f800001u; u += 7ull * stride)
        for (uint32_t sg = 0; sg < 2; sg++) {
            const float x = fbits_as((int32_t)((uint32_t)u | (sg << 31)));
            n_atan++;
            if (!same(rp3dg::ref_atanf(x), atanf(x))) bad_atan++;
        }
    const long pairs = 40000000 / (long)stride;
    for (long i = 0; i < pairs; i++) {
        float y, x;
        if (i % 3 == 0) {
            y = fbits_as((int32_t)rnd());
            x = fbits_as((int32_t)rnd());
        } else {
            y = (float)(rnd() & 0xffffff) / 16777216.0f;
            x = (float)(rnd() & 0xffffff) / 8388608.0f - 1.0f;
        }
        n_atan2++;
        if (!same(rp3dg::ref_atan2f(y, x), atan2f(y, x))) bad_atan2++;
    }
    printf("acosf %ld %ld\natanf %ld %ld\natan2f %ld %ld\n", bad_acos, n_acos, bad_atan, n_atan, bad_atan2, n_atan2);
    return (bad_acos || bad_atan || bad_atan2) ? 1 : 0;
}
