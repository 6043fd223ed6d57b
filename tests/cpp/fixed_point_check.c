/* CPU emulation of the fixed-point conversion of P2G and the G2P adjoint
 * (nuclearmpm_amd/csrc/mpm_device.hpp, fixed_point): t = double(y) + 1.5 * 2^52 (round to nearest
 * even, the GPU's default mode); the result is the bit pattern of t with 0x43380000 subtracted
 * from its high word.
 *
 * Checks the conversion against llrint(y) (round to nearest, ties to even) over float bit patterns
 * with |y| < 2^51: every pattern that is a multiple of `stride`, plus every float in [-4, 4] and
 * every float in [2^50, 2^51) with both signs at a stride of 7.  Prints "checked N mismatch M" and
 * each y whose result differs (as "y=<hex float> got <int> want <int>"), exit status 0. */
#include <fenv.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int64_t fixed_point(float y) {
    const double t = (double)y + 6755399441055744.0;
    uint64_t b;
    memcpy(&b, &t, 8);
    const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32) - 0x43380000u;
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

static uint64_t checked, bad;

static void check(uint32_t bits) {
    float y;
    memcpy(&y, &bits, 4);
    if (!(fabsf(y) < 2251799813685248.0f)) return; /* 2^51 */
    ++checked;
    const int64_t want = llrint((double)y), got = fixed_point(y);
    if (got != want) {
        ++bad;
        if (bad < 20) printf("y=%a got %lld want %lld\n", (double)y, (long long)got, (long long)want);
    }
}

static void range(float a, float b, uint32_t step) { /* a, b same sign, |a| <= |b| */
    uint32_t lo, hi;
    memcpy(&lo, &a, 4);
    memcpy(&hi, &b, 4);
    for (uint32_t x = lo; x <= hi && x >= lo; x += step) check(x);
}

int main(int argc, char** argv) {
    fesetround(FE_TONEAREST);
    const uint64_t stride = argc > 1 ? strtoull(argv[1], NULL, 10) : 61;
    for (uint64_t b = 0; b < (1ull << 32); b += stride) check((uint32_t)b);
    range(0.0f, 4.0f, 1);
    range(-0.0f, -4.0f, 1);
    range(1125899906842624.0f, 2251799813685247.9f, 7);   /* [2^50, 2^51) */
    range(-1125899906842624.0f, -2251799813685247.9f, 7);
    printf("checked %llu mismatch %llu\n", (unsigned long long)checked, (unsigned long long)bad);
    return 0;
}
