/*
 * Fixture generator: records the host CPU's RSQRTPS (_mm_rsqrt_ps) outputs.
 *
 * Viewer/SponzaScene.cpp:67 computes 1/distance with _mm256_rsqrt_ps, an
 * approximation whose exact bits are implementation-defined. This program
 * observes the instruction on the build host (an Intel Xeon) and checks the
 * table model: for a normal x = 2^E * f, f in [1, 2), rsqrt(x) =
 * T[E & 1][top 11 bits of f] * 2^-(E - (E & 1)) / 2, the table holding
 * rsqrt of [1, 2) (even E) and of [2, 4) (odd E). It is checked
 * exhaustively over all 2^23 mantissas of both ranges, on 2M random normal
 * inputs across every exponent, and on the special inputs (+-0, denormals,
 * +inf, negatives), then the 4096-entry table is printed (even range first),
 * one hex word per line.
 *
 * Build and run:  gcc -O2 -msse2 gen_rsqrtps_table.c -o /tmp/g && /tmp/g > rsqrtps_table.txt
 * Exit status is non-zero if the host violates the table model.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint32_t bits_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float float_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static float host_rsqrt(float x) { return _mm_cvtss_f32(_mm_rsqrt_ps(_mm_set1_ps(x))); }

static uint32_t table[2][2048];

/* The model's prediction for a positive normal input (biased exponent eb). */
static uint32_t predict(uint32_t u)
{
    const int eb = (int)((u >> 23) & 255u), e = eb - 127, p = e & 1;
    const uint32_t t = table[p][(u >> 12) & 2047u];
    const int re = (int)((t >> 23) & 255u) - (e - p) / 2;
    return ((uint32_t)re << 23) | (t & 0x7FFFFFu);
}

int main(void)
{
    for (uint32_t p = 0; p < 2; ++p)
        for (uint32_t i = 0; i < 2048; ++i) table[p][i] = bits_of(host_rsqrt(float_of(((127u + p) << 23) | (i << 12))));

    long bad = 0;
    for (uint32_t p = 0; p < 2; ++p)
        for (uint32_t m = 0; m < (1u << 23); ++m)
            if (bits_of(host_rsqrt(float_of(((127u + p) << 23) | m))) != table[p][m >> 12]) ++bad;

    uint32_t seed = 0x5EED0001u;
    for (uint32_t e = 1; e < 255; ++e)
        for (int t = 0; t < 8192; ++t) {
            seed = seed * 1664525u + 1013904223u;
            const uint32_t u = (e << 23) | (seed >> 9);
            if (bits_of(host_rsqrt(float_of(u))) != predict(u)) ++bad;
        }

    /* specials: +-0 and denormals -> +-inf, +inf -> +0, negative numbers and -inf -> the
       default NaN, a NaN -> that NaN quieted */
    if (bits_of(host_rsqrt(0.0f)) != 0x7F800000u || bits_of(host_rsqrt(-0.0f)) != 0xFF800000u ||
        bits_of(host_rsqrt(float_of(0x00000100u))) != 0x7F800000u ||
        bits_of(host_rsqrt(float_of(0x80000100u))) != 0xFF800000u ||
        bits_of(host_rsqrt(float_of(0x7F800000u))) != 0u || bits_of(host_rsqrt(-1.0f)) != 0xFFC00000u ||
        bits_of(host_rsqrt(float_of(0xFF800000u))) != 0xFFC00000u ||
        bits_of(host_rsqrt(float_of(0x7F800001u))) != 0x7FC00001u ||
        bits_of(host_rsqrt(float_of(0xFFA00000u))) != 0xFFE00000u)
        ++bad;

    for (uint32_t p = 0; p < 2; ++p)
        for (uint32_t i = 0; i < 2048; ++i) printf("%08x\n", table[p][i]);
    fprintf(stderr, "rsqrtps table model violations: %ld\n", bad);
    return bad != 0;
}
