/*
 * Deterministic synthetic SIFT-like descriptor sets for the putative matcher.
 *
 * This is the input generator of the benchmarks and tests (the reference ships no
 * feature data).  A descriptor is built as
 *   1. 128 draws g_k ~ Gamma(0.5, 1)                (Marsaglia & Tsang 2000, with the
 *                                                     U^(1/a) boost for a < 1)
 *   2. x = g / ||g||_2, x_k = min(x_k, 0.2), x = x / ||x||_2
 *   3. q_k = min(255, floor(512 x_k + 0.5))         (uint8 storage)
 *   4. dequantisation as SIFT_Feature::normalizeDescriptor
 *      (V3D/Base/v3d_feature.h:30-38): var = sum q_k^2 in float, d_k = q_k * (1/sqrtf(var)).
 * The right set of a pair holds noisy copies of distinct, randomly chosen left
 * descriptors (dequantised left + N(0, sigma^2) per component, clamped at 0,
 * renormalised, re-quantised through steps 3-4; no 0.2 clip) and fresh distractors,
 * in random order.
 *
 * Random numbers: SplitMix64 (Steele, Lea & Flood 2014) expands a 64-bit seed into the
 * state of xoshiro256** (Blackman & Vigna 2018).  Seeds of independent streams are
 * derived by hashing (base, config, pair, side) through SplitMix64.
 * Normals use the Marsaglia polar method.  Everything is double precision with libm
 * log/sqrt/exp/pow; results are deterministic for a given libm.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define D 128

typedef struct { uint64_t s[4]; int have_spare; double spare; } rng_t;

static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t xoshiro_next(rng_t* r)
{
    uint64_t* s = r->s;
    uint64_t const result = rotl(s[1] * 5, 7) * 9;
    uint64_t const t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static void rng_init(rng_t* r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
    r->have_spare = 0;
    r->spare = 0.0;
}

/* uniform in [0, 1) with 53 random bits */
static double unif(rng_t* r) { return (double)(xoshiro_next(r) >> 11) * 0x1.0p-53; }

/* uniform in (0, 1) */
static double unif_open(rng_t* r)
{
    double u;
    do { u = unif(r); } while (u <= 0.0);
    return u;
}

static double normal(rng_t* r)
{
    if (r->have_spare) { r->have_spare = 0; return r->spare; }
    double u, v, s;
    do {
        u = 2.0 * unif(r) - 1.0;
        v = 2.0 * unif(r) - 1.0;
        s = u * u + v * v;
    } while (s >= 1.0 || s == 0.0);
    double const m = sqrt(-2.0 * log(s) / s);
    r->spare = v * m;
    r->have_spare = 1;
    return u * m;
}

/* Gamma(a, 1), Marsaglia & Tsang; for a < 1 draw Gamma(a+1) and scale by U^(1/a). */
static double gamma_draw(rng_t* r, double a)
{
    double const aa = (a < 1.0) ? a + 1.0 : a;
    double const d = aa - 1.0 / 3.0;
    double const c = 1.0 / sqrt(9.0 * d);
    double g;
    for (;;) {
        double x, v;
        do {
            x = normal(r);
            v = 1.0 + c * x;
        } while (v <= 0.0);
        v = v * v * v;
        double const u = unif_open(r);
        if (u < 1.0 - 0.0331 * (x * x) * (x * x)) { g = d * v; break; }
        if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) { g = d * v; break; }
    }
    if (a < 1.0) g *= pow(unif_open(r), 1.0 / a);
    return g;
}

static void quantise(double const* x, uint8_t* q)
{
    for (int k = 0; k < D; ++k) {
        double const t = floor(512.0 * x[k] + 0.5);
        q[k] = (uint8_t)(t > 255.0 ? 255.0 : (t < 0.0 ? 0.0 : t));
    }
}

static void normalise(double* x)
{
    double s = 0.0;
    for (int k = 0; k < D; ++k) s += x[k] * x[k];
    double const inv = 1.0 / sqrt(s);
    for (int k = 0; k < D; ++k) x[k] *= inv;
}

static void fresh_descriptor(rng_t* r, uint8_t* q)
{
    double x[D];
    for (int k = 0; k < D; ++k) x[k] = gamma_draw(r, 0.5);
    normalise(x);
    for (int k = 0; k < D; ++k) x[k] = (x[k] > 0.2) ? 0.2 : x[k];
    normalise(x);
    quantise(x, q);
}

/* ---------------------------------------------------------------------------------- */

uint64_t v3ds_seed(uint64_t base, uint32_t config, uint32_t pair, uint32_t side)
{
    uint64_t x = base;
    uint64_t h = splitmix64(&x);
    x = h ^ ((uint64_t)config << 32 | (uint64_t)pair);
    h = splitmix64(&x);
    x = h ^ (uint64_t)side * 0xD6E8FEB86659FD93ull;
    return splitmix64(&x);
}

/* Dequantise n uint8 descriptors exactly as SIFT_Feature::normalizeDescriptor does. */
void v3ds_dequantise(uint8_t const* q, int n, float* out)
{
    for (int i = 0; i < n; ++i) {
        uint8_t const* qi = q + (size_t)i * D;
        float* di = out + (size_t)i * D;
        float var = 0.0f;
        for (int k = 0; k < D; ++k) {
            float const v = (float)qi[k];
            var += v * v;
        }
        float const denom = 1.0f / sqrtf(var);
        for (int k = 0; k < D; ++k) di[k] = (float)qi[k] * denom;
    }
}

/* n fresh descriptors from the stream seeded by `seed`. */
void v3ds_fresh(uint64_t seed, int n, uint8_t* q)
{
    rng_t r;
    rng_init(&r, seed);
    for (int i = 0; i < n; ++i) fresh_descriptor(&r, q + (size_t)i * D);
}

/*
 * One image pair: nL fresh left descriptors; nR right descriptors of which
 * nCopy = round(copyFrac * min(nL, nR)) are noisy copies of distinct left descriptors.
 * truth[j] = index of the left descriptor right j was copied from, or -1.
 * Returns nCopy.
 */
int v3ds_make_pair(uint64_t base, uint32_t config, uint32_t pair, int nL, int nR,
                   double copyFrac, double sigma, uint8_t* qL, uint8_t* qR, int32_t* truth)
{
    if (nL < 0 || nR < 0) return -1;
    v3ds_fresh(v3ds_seed(base, config, pair, 0), nL, qL);

    rng_t r;
    rng_init(&r, v3ds_seed(base, config, pair, 1));

    int const nMin = nL < nR ? nL : nR;
    int nCopy = (int)floor(copyFrac * (double)nMin + 0.5);
    if (nCopy > nMin) nCopy = nMin;
    if (nCopy < 0) nCopy = 0;

    /* distinct source indices: partial Fisher-Yates over the left set */
    int32_t* src = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nL > 0 ? nL : 1));
    for (int i = 0; i < nL; ++i) src[i] = i;
    for (int i = 0; i < nCopy; ++i) {
        int const j = i + (int)(unif(&r) * (double)(nL - i));
        int32_t const t = src[i]; src[i] = src[j]; src[j] = t;
    }

    float* dl = (float*)malloc(sizeof(float) * D);
    uint8_t* tmp = (uint8_t*)malloc((size_t)(nR > 0 ? nR : 1) * D);
    int32_t* tr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nR > 0 ? nR : 1));
    for (int j = 0; j < nR; ++j) {
        if (j < nCopy) {
            v3ds_dequantise(qL + (size_t)src[j] * D, 1, dl);
            double x[D];
            for (int k = 0; k < D; ++k) {
                double v = (double)dl[k] + sigma * normal(&r);
                x[k] = v < 0.0 ? 0.0 : v;
            }
            normalise(x);
            quantise(x, tmp + (size_t)j * D);
            tr[j] = src[j];
        } else {
            fresh_descriptor(&r, tmp + (size_t)j * D);
            tr[j] = -1;
        }
    }
    /* random order of the right set */
    int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nR > 0 ? nR : 1));
    for (int j = 0; j < nR; ++j) perm[j] = j;
    for (int j = nR - 1; j > 0; --j) {
        int const t = (int)(unif(&r) * (double)(j + 1));
        int32_t const s = perm[j]; perm[j] = perm[t]; perm[t] = s;
    }
    for (int j = 0; j < nR; ++j) {
        memcpy(qR + (size_t)j * D, tmp + (size_t)perm[j] * D, D);
        if (truth) truth[j] = tr[perm[j]];
    }
    free(perm); free(tr); free(tmp); free(dl); free(src);
    return nCopy;
}

/*
 * Sizes of a batch of uneven pairs: per image n = floor(exp(U(ln lo, ln hi))),
 * per pair a copy fraction U(fracLo, fracHi).
 */
void v3ds_batch_sizes(uint64_t base, uint32_t config, int nPairs, int lo, int hi,
                      double fracLo, double fracHi, int32_t* nL, int32_t* nR, double* frac)
{
    rng_t r;
    rng_init(&r, v3ds_seed(base, config, 0xFFFFFFFFu, 2));
    double const a = log((double)lo), b = log((double)hi);
    for (int p = 0; p < nPairs; ++p) {
        int l = (int)floor(exp(a + (b - a) * unif(&r)));
        int m = (int)floor(exp(a + (b - a) * unif(&r)));
        nL[p] = l < lo ? lo : (l > hi ? hi : l);
        nR[p] = m < lo ? lo : (m > hi ? hi : m);
        frac[p] = fracLo + (fracHi - fracLo) * unif(&r);
    }
}
