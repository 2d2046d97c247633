/* CPU check of two FP32 identities the kernels use (luvas.io-3d_amd/csrc/fp_exact.h), built with
 * -ffp-contract=off so that every operation rounds once, as on the GPU:
 *
 *   div_rn      q0 = RN(a y), q1 = RN(q0 + RN(a - b q0) y), q = RN(q1 + RN(a - b q1) y) with
 *               y = RN(1 / b) equals the IEEE quotient RN(a / b);
 *   sqrt_normal the neighbour selection by two FMA residuals returns RN(sqrt(x)) from every seed
 *               within 1 ulp of the root (the hardware's v_sqrt_f32 is within 1 ulp).
 *
 * Operands: b and a with exponents in [-60, 60] (mantissas uniform, one in eight with the top
 * 15 bits set, one in eight with only low bits), quotients kept when in [2^-100, 2^100); x with
 * exponents in [-96, 119].  Prints the sample count and the mismatch counts.
 * Usage: exact_arith_check <samples> <seed>
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float from_bits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static uint32_t to_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static uint64_t state;

static uint64_t next(void) {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}

static float div_rn(float a, float b, float y) {
    const float q0 = a * y;
    const float q1 = fmaf(fmaf(-b, q0, a), y, q0);
    return fmaf(fmaf(-b, q1, a), y, q1);
}

static float sqrt_select(float x, float s) {
    const float below = from_bits(to_bits(s) - 1), above = from_bits(to_bits(s) + 1);
    const float r = fmaf(-below, s, x) <= 0.0f ? below : s;
    return fmaf(-above, s, x) > 0.0f ? above : r;
}

static float operand(uint64_t r) {
    const int e = (int)(r % 121) - 60;
    uint32_t m = (uint32_t)(r >> 8) & 0x7fffffu;
    if (((r >> 40) & 7) == 0) m |= 0x7fff00u;
    if (((r >> 44) & 7) == 0) m &= 0xffu;
    const float v = from_bits(((uint32_t)(e + 127) << 23) | m);
    return (r >> 63) ? -v : v;
}

int main(int argc, char **argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000L;
    state = argc > 2 ? strtoull(argv[2], NULL, 10) : 88172645463325252ull;
    if (state == 0) state = 1;
    long quotients = 0, bad_div = 0, bad_sqrt = 0;
    for (long i = 0; i < n; i++) {
        const float b = operand(next());
        const float a = operand(next());
        const float y = 1.0f / b;
        const float q = a / b;
        if (fabsf(q) >= 0x1p-100f && fabsf(q) < 0x1p100f) {
            quotients++;
            if (to_bits(div_rn(a, b, y)) != to_bits(q)) bad_div++;
        }
        const uint64_t r = next();
        const float x = from_bits(((uint32_t)((int)(r % 216) - 96 + 127) << 23) | ((uint32_t)(r >> 20) & 0x7fffffu));
        const float root = sqrtf(x);
        for (int k = -1; k <= 1; k++)
            if (to_bits(sqrt_select(x, from_bits(to_bits(root) + (uint32_t)k))) != to_bits(root)) bad_sqrt++;
    }
    printf("samples %ld quotients %ld div_rn_mismatches %ld sqrt_mismatches %ld\n", n, quotients, bad_div, bad_sqrt);
    return 0;
}
