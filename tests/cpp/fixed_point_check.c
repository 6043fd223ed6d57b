/* CPU emulation of P2G's fixed-point conversion (nuclearmpm_amd/csrc/mpm_device.hpp,
 * fixed_point): r = double(y); hi = floor(r 2^-32); lo = fma(hi, -2^32, r); the result is
 * (u32)lo | (i32)hi << 32, where the GPU's conversions truncate toward zero and saturate
 * (v_cvt_u32_f64 to [0, 2^32 - 1], v_cvt_i32_f64 to [-2^31, 2^31 - 1]).
 *
 * Checks the conversion against floor(y) over float bit patterns with |y| < 2^62: every pattern
 * that is a multiple of `stride`, plus every float in (-4, -1] and in [-2^-20, -2^-30).  Prints
 * "checked N off_by_one M" and each y whose result is not floor(y) (as "y=<hex float> got <int>"),
 * exit status 0.  */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t cvt_u32(double v) { return v != v || v <= 0.0 ? 0u : v >= 4294967295.0 ? 4294967295u : (uint32_t)v; }
static int32_t cvt_i32(double v) {
    return v != v ? 0 : v <= -2147483648.0 ? INT32_MIN : v >= 2147483647.0 ? INT32_MAX : (int32_t)v;
}

static int64_t fixed_point(float y) {
    const double r = (double)y;
    const double hi = floor(r * 2.3283064365386963e-10);
    const double lo = fma(hi, -4294967296.0, r);
    return (int64_t)((uint64_t)cvt_u32(lo) | ((uint64_t)(uint32_t)cvt_i32(hi) << 32));
}

static uint64_t checked, odd;

static void check(uint32_t bits) {
    float y;
    memcpy(&y, &bits, 4);
    if (!(fabsf(y) < 4.6116860e18f)) return; /* 2^62 */
    ++checked;
    const int64_t want = (int64_t)floorl((long double)y), got = fixed_point(y);
    if (got != want) {
        ++odd;
        printf("y=%a got %lld floor %lld\n", (double)y, (long long)got, (long long)want);
    }
}

int main(int argc, char** argv) {
    const uint64_t stride = argc > 1 ? strtoull(argv[1], NULL, 10) : 61;
    for (uint64_t b = 0; b < (1ull << 32); b += stride) check((uint32_t)b);
    /* negative floats: the larger the magnitude, the larger the bit pattern */
    const float ranges[2][2] = {{-1.0f, -4.0f}, {-9.313225746154785e-10f, -9.5367431640625e-07f}};
    for (int k = 0; k < 2; ++k) {
        uint32_t lo, hi;
        memcpy(&lo, &ranges[k][0], 4);
        memcpy(&hi, &ranges[k][1], 4);
        for (uint32_t b = lo; b <= hi; ++b) check(b);
    }
    printf("checked %llu off_by_one %llu\n", (unsigned long long)checked, (unsigned long long)odd);
    return 0;
}
