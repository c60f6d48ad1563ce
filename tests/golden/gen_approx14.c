/*
 * Generates and verifies approx14.bin: the lookup tables that reproduce the
 * x86 AVX-512 approximation instructions VRCP14PS and VRSQRT14PS bit for bit.
 *
 * The reference renderer calls these instructions explicitly
 * (SIMD.h:294-295 approx_rsqrt/approx_rcp) and, being compiled with
 * -ffast-math (src/SwRast/CMakeLists.txt:6), also gets them from every vector
 * float division, e.g. SIMD.h:474 perspective_div and Rasterizer.cpp:320
 * rcpArea. Snapped vertex positions, depth planes and resolve barycentrics
 * therefore depend on their exact output bits.
 *
 * Model (verified against the instruction over all 2^32 inputs by --verify):
 *   rcp14:   for a normal or normalised-denormal input 1.m * 2^e with m != 0,
 *            the result is 1.q * 2^(-e-1) where q (16 bits) = RCP[m >> 7].
 *            m == 0 gives the exact power of two. Results below the normal
 *            range are produced as denormals by truncating the shifted
 *            significand; zero -> inf, inf -> zero, NaN -> quiet NaN.
 *   rsqrt14: for input 1.m * 2^(2h+o), o in {0,1}: result 1.q * 2^(-h-1) with
 *            q = RSQRT[o][m >> 8]; exact power of four gives 2^-h; negative
 *            non-zero inputs give the default NaN 0xFFC00000.
 *
 * File layout (little endian u16): RCP[65536] then RSQRT[2][32768].
 *
 * Build and run on an AVX-512 host:
 *   gcc -O2 -mavx512f tests/golden/gen_approx14.c -o gen_approx14
 *   ./gen_approx14 glimpsw_amd/data/approx14.bin
 *   ./gen_approx14 --verify glimpsw_amd/data/approx14.bin   (exhaustive, ~20 s)
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint16_t RCP[65536];
static uint16_t RSQ[2][32768];

static uint32_t model_rcp14(uint32_t x) {
    uint32_t s = x & 0x80000000u;
    int e = (int)((x >> 23) & 255);
    uint32_t m = x & 0x7FFFFFu;
    if (e == 255) return m ? (x | 0x400000u) : s;
    if (e == 0) {
        if (m == 0) return s | 0x7F800000u;
        while (!(m & 0x800000u)) { m <<= 1; e--; }
        m &= 0x7FFFFFu;
        e += 1;
    }
    int re;
    uint32_t mant;
    if (m == 0) { re = 254 - e; mant = 0; }
    else { re = 253 - e; mant = (uint32_t)RCP[m >> 7] << 7; }
    if (re >= 255) return s | 0x7F800000u;
    if (re <= 0) {
        int sh = 1 - re;
        if (sh > 24) return s;
        return s | ((mant | 0x800000u) >> sh);
    }
    return s | ((uint32_t)re << 23) | mant;
}

static uint32_t model_rsqrt14(uint32_t x) {
    uint32_t s = x & 0x80000000u;
    int e = (int)((x >> 23) & 255);
    uint32_t m = x & 0x7FFFFFu;
    if (e == 255) { if (m) return x | 0x400000u; return s ? 0xFFC00000u : 0; }
    if (e == 0) {
        if (m == 0) return s | 0x7F800000u;
        while (!(m & 0x800000u)) { m <<= 1; e--; }
        m &= 0x7FFFFFu;
        e += 1;
    }
    if (s) return 0xFFC00000u;
    int ue = e - 127;
    int oc = ue & 1;
    int half = (ue - oc) / 2;
    if (m == 0 && oc == 0) return (uint32_t)(127 - half) << 23;
    return ((uint32_t)(126 - half) << 23) | ((uint32_t)RSQ[oc][m >> 8] << 7);
}

static void build_tables(void) {
    for (uint32_t i = 0; i < 65536; i += 16) {
        uint32_t in[16], out[16];
        for (int k = 0; k < 16; k++) in[k] = 0x3F800000u | ((i + k) << 7) | 1u;
        __m512 v; memcpy(&v, in, 64);
        __m512 r = _mm512_rcp14_ps(v);
        memcpy(out, &r, 64);
        for (int k = 0; k < 16; k++) RCP[i + k] = (uint16_t)((out[k] >> 7) & 0xFFFF);
    }
    for (int oc = 0; oc < 2; oc++) {
        for (uint32_t i = 0; i < 32768; i += 16) {
            uint32_t in[16], out[16];
            for (int k = 0; k < 16; k++) in[k] = ((127u + (uint32_t)oc) << 23) | ((i + k) << 8) | 1u;
            __m512 v; memcpy(&v, in, 64);
            __m512 r = _mm512_rsqrt14_ps(v);
            memcpy(out, &r, 64);
            for (int k = 0; k < 16; k++) RSQ[oc][i + k] = (uint16_t)((out[k] >> 7) & 0xFFFF);
        }
    }
}

static int verify(void) {
    uint64_t bad_rcp = 0, bad_rsq = 0;
    for (uint64_t b = 0; b < (1ull << 32); b += 16) {
        uint32_t in[16], o1[16], o2[16];
        for (int k = 0; k < 16; k++) in[k] = (uint32_t)(b + (uint64_t)k);
        __m512 v; memcpy(&v, in, 64);
        __m512 r1 = _mm512_rcp14_ps(v), r2 = _mm512_rsqrt14_ps(v);
        memcpy(o1, &r1, 64); memcpy(o2, &r2, 64);
        for (int k = 0; k < 16; k++) {
            bad_rcp += model_rcp14(in[k]) != o1[k];
            bad_rsq += model_rsqrt14(in[k]) != o2[k];
        }
    }
    printf("rcp14 mismatches %llu, rsqrt14 mismatches %llu\n",
           (unsigned long long)bad_rcp, (unsigned long long)bad_rsq);
    return (bad_rcp || bad_rsq) ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc == 3 && strcmp(argv[1], "--verify") == 0) {
        FILE* f = fopen(argv[2], "rb");
        if (!f) { perror("open"); return 2; }
        if (fread(RCP, 2, 65536, f) != 65536 || fread(RSQ, 2, 65536, f) != 65536) { fprintf(stderr, "short read\n"); return 2; }
        fclose(f);
        return verify();
    }
    if (argc != 2) { fprintf(stderr, "usage: %s out.bin | --verify in.bin\n", argv[0]); return 2; }
    build_tables();
    FILE* f = fopen(argv[1], "wb");
    if (!f) { perror("open"); return 2; }
    fwrite(RCP, 2, 65536, f);
    fwrite(RSQ, 2, 65536, f);
    fclose(f);
    return 0;
}
