/*
 * Fixture generator: records the host CPU's RCPPS (_mm_rcp_ps) outputs.
 *
 * SoftRast/Rasterizer.cpp:375-376 computes the neighbour 1/w values used for
 * the texture-coordinate derivatives with _mm256_rcp_ps, an approximation
 * whose exact bits are implementation-defined. This program observes the
 * instruction on the build host (an Intel Xeon), checks that its output is a
 * pure function of the top 11 mantissa bits scaled exactly by the exponent
 * (exhaustively over all 2^23 mantissas of [1,2), and on 2M random inputs
 * across every exponent and both signs), and prints the 2048-entry table of
 * rcp(1 + i/2048) bit patterns, one hex word per line.
 *
 * Build and run:  gcc -O2 -msse2 gen_rcpps_table.c -o /tmp/g && /tmp/g > rcpps_table.txt
 * Exit status is non-zero if the host violates the table model.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint32_t bits_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float float_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static float host_rcp(float x) { return _mm_cvtss_f32(_mm_rcp_ps(_mm_set1_ps(x))); }

int main(void)
{
    static uint32_t table[2048];
    for (uint32_t i = 0; i < 2048; ++i)
        table[i] = bits_of(host_rcp(float_of(0x3F800000u | (i << 12))));

    long bad = 0;
    for (uint32_t m = 0; m < (1u << 23); ++m)
        if (bits_of(host_rcp(float_of(0x3F800000u | m))) != table[m >> 12]) ++bad;

    uint32_t seed = 0x5EED0000u;
    for (uint32_t e = 1; e < 255; ++e)
        for (uint32_t s = 0; s < 2; ++s)
            for (int t = 0; t < 4096; ++t) {
                seed = seed * 1664525u + 1013904223u;
                uint32_t m = seed >> 9;
                uint32_t u = (s << 31) | (e << 23) | m;
                uint32_t tb = table[m >> 12];
                int ne = (int)((tb >> 23) & 255u) - ((int)e - 127);
                uint32_t pred = ne >= 255 ? 0x7F800000u : ne <= 0 ? 0u : (((uint32_t)ne << 23) | (tb & 0x7FFFFFu));
                pred |= s << 31;
                if (bits_of(host_rcp(float_of(u))) != pred) ++bad;
            }

    if (bits_of(host_rcp(0.0f)) != 0x7F800000u || bits_of(host_rcp(-0.0f)) != 0xFF800000u ||
        bits_of(host_rcp(float_of(0x00000100u))) != 0x7F800000u ||
        bits_of(host_rcp(float_of(0x7F800000u))) != 0u)
        ++bad;

    for (uint32_t i = 0; i < 2048; ++i) printf("%08x\n", table[i]);
    fprintf(stderr, "rcpps table model violations: %ld\n", bad);
    return bad != 0;
}
