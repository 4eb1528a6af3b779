/*
 * synth.c -- deterministic synthetic luma frames for the ME workloads.
 *
 * Frames are a sum of six low-frequency 2-D sinusoids plus noise, sampled
 * analytically at motion-shifted coordinates (so motion needs no
 * interpolation).  Everything is computed with IEEE double +, -, *, / and
 * sqrt (all correctly rounded; no other libm call, no FMA contraction: build
 * with -ffp-contract=off), so a seed gives the same samples on any host.
 *
 *   random stream  splitmix64, counter based, keyed by (seed, stream, frame,
 *                  purpose)
 *   noise          standard normal by the Box-Muller transform,
 *                  sqrt(-2 ln u1) cos(2 pi u2), with this file's own ln and
 *                  sin (series evaluated in IEEE double)
 *   sinusoid k     A_k in [20,60] (x4 at 10 bit), f_k in [1/256, 1/32]
 *                  cycles/px, orientation in [0,pi), phase in [0,2pi)
 *   sample         clip(floor(DC + sum_k A_k sin(2 pi f_k (x c_k + y s_k)
 *                  + phi_k) + offset + sigma * noise + 0.5)), DC = 2^(bd-1)
 *   motion         the sample at (x, y) evaluates the field at
 *                  (x + sx + rx, y + sy + ry) where (sx, sy) is the global
 *                  shift in pels and (rx, ry) a per-64x64-region integer
 *                  quarter-pel offset in [-region_qpel, region_qpel] / 4.
 */
#include <stdint.h>
#include <string.h>

typedef uint64_t u64;

#define EXPORT __attribute__((visibility("default")))

static u64 mix(u64 x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

EXPORT u64 me_synth_key(u64 seed, int stream, int frame, int purpose)
{
    u64 k = mix(seed);
    k = mix(k ^ (u64)(unsigned)stream);
    k = mix(k ^ ((u64)(unsigned)frame << 20));
    return mix(k ^ ((u64)(unsigned)purpose << 40));
}

static double uni(u64 key, u64 i) { return (double)(mix(key + i * 0xD1B54A32D192ED03ull) >> 11) * (1.0 / 9007199254740992.0); }

EXPORT double me_synth_uniform(u64 key, u64 i) { return uni(key, i); }

static const double PI = 3.141592653589793115997963468544185161590576171875;
static const double TWO_PI = 6.28318530717958623199592693708837032318115234375;
static const double LN2 = 0.69314718055994528622676398299518041312694549560546875;
static const double SQRT2 = 1.4142135623730951454746218587388284504413604736328125;

/* sin on [-pi/2, pi/2]: Taylor series to x^17 (truncation < 1e-11). */
static double sin_poly(double x)
{
    double x2 = x * x;
    double t = 1.0 / 355687428096000.0;                    /* 1/17! */
    t = -1.0 / 1307674368000.0 + x2 * t;                   /* 1/15! */
    t = 1.0 / 6227020800.0 + x2 * t;                       /* 1/13! */
    t = -1.0 / 39916800.0 + x2 * t;                        /* 1/11! */
    t = 1.0 / 362880.0 + x2 * t;                           /* 1/9!  */
    t = -1.0 / 5040.0 + x2 * t;                            /* 1/7!  */
    t = 1.0 / 120.0 + x2 * t;                              /* 1/5!  */
    t = -1.0 / 6.0 + x2 * t;                               /* 1/3!  */
    return x + x * x2 * t;
}

EXPORT double me_synth_sin(double x)
{
    /* reduce to [-pi, pi] with n = floor(x / 2pi + 0.5) computed without libm */
    double q = x / TWO_PI + 0.5;
    long long n = (long long)q;
    if ((double)n > q) n--;
    double r = x - (double)n * TWO_PI;
    if (r > PI / 2) r = PI - r;
    else if (r < -PI / 2) r = -PI - r;
    return sin_poly(r);
}

static double cos_(double x) { return me_synth_sin(x + PI / 2); }

/* ln x for a normal double x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2))
 * (exponent and mantissa taken from the bits, exactly), then
 * ln m = 2 atanh(t), t = (m - 1) / (m + 1), |t| < 0.172, as the odd series
 * to t^29 (truncation below 1e-20). */
EXPORT double me_synth_ln(double x)
{
    u64 b;
    memcpy(&b, &x, sizeof(b));
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    b = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    memcpy(&m, &b, sizeof(m));
    if (m > SQRT2) {
        m *= 0.5;
        e++;
    }
    double t = (m - 1.0) / (m + 1.0), t2 = t * t;
    double s = 1.0 / 29.0;
    for (int k = 27; k >= 1; k -= 2) s = 1.0 / k + t2 * s;
    return (double)e * LN2 + 2.0 * t * s;
}

/* Standard normal sample i of stream `key` (Box-Muller; u1 in (0, 1]). */
EXPORT double me_synth_gauss(u64 key, u64 i)
{
    double u1 = 1.0 - uni(key, 2 * i), u2 = uni(key, 2 * i + 1);
    return __builtin_sqrt(-2.0 * me_synth_ln(u1)) * cos_(TWO_PI * u2);
}

#define NS 6
typedef struct {
    double a[NS], kx[NS], ky[NS], ph[NS];
    double dc;
} field_t;

static void field_init(field_t *f, u64 seed, int stream, int bd)
{
    u64 k = me_synth_key(seed, stream, 0, 1);
    double scale = bd == 8 ? 1.0 : 4.0;
    for (int i = 0; i < NS; i++) {
        double amp = (20.0 + 40.0 * uni(k, 4 * i + 0)) * scale;
        double fr = 1.0 / 256.0 + uni(k, 4 * i + 1) * (1.0 / 32.0 - 1.0 / 256.0);
        double th = PI * uni(k, 4 * i + 2);
        double ph = TWO_PI * uni(k, 4 * i + 3);
        f->a[i] = amp;
        f->kx[i] = TWO_PI * fr * cos_(th);
        f->ky[i] = TWO_PI * fr * me_synth_sin(th);
        f->ph[i] = ph;
    }
    f->dc = (double)(1 << (bd - 1));
}

static double field_at(const field_t *f, double x, double y)
{
    double v = f->dc;
    for (int i = 0; i < NS; i++) v += f->a[i] * me_synth_sin(f->kx[i] * x + f->ky[i] * y + f->ph[i]);
    return v;
}

/* One frame.  out has h rows of w samples (uint8 when bd == 8, else uint16).
 *   seed, stream  pick the field (the scene)
 *   frame, noise_purpose  pick the noise realisation
 *   sx, sy        global shift in pels
 *   region_qpel   max |per-64x64 offset| in quarter pels (0: none); offsets
 *                 are keyed by (seed, stream, frame, purpose 2)
 *   sigma         noise standard deviation
 *   offset        brightness offset added before rounding */
EXPORT void me_synth_frame(u64 seed, int stream, int frame, int w, int h, int bd, double sx, double sy,
                           int region_qpel, double sigma, int noise_purpose, double offset, void *out)
{
    field_t f;
    field_init(&f, seed, stream, bd);
    u64 kn = me_synth_key(seed, stream, frame, noise_purpose);
    u64 kr = me_synth_key(seed, stream, frame, 2);
    int rcols = (w + 63) / 64;
    int maxv = (1 << bd) - 1;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double rx = 0, ry = 0;
            if (region_qpel) {
                u64 ri = (u64)((y / 64) * rcols + (x / 64));
                int span = 2 * region_qpel + 1;
                int ox = (int)(uni(kr, 2 * ri) * span) - region_qpel;
                int oy = (int)(uni(kr, 2 * ri + 1) * span) - region_qpel;
                rx = ox / 4.0;
                ry = oy / 4.0;
            }
            double v = field_at(&f, x + sx + rx, y + sy + ry) + offset;
            if (sigma != 0.0) v += sigma * me_synth_gauss(kn, (u64)y * (u64)w + (u64)x);
            double fl = v + 0.5;
            long long iv = (long long)fl;
            if ((double)iv > fl) iv--;
            int s = iv < 0 ? 0 : (iv > maxv ? maxv : (int)iv);
            if (bd == 8) ((uint8_t *)out)[(size_t)y * w + x] = (uint8_t)s;
            else ((uint16_t *)out)[(size_t)y * w + x] = (uint16_t)s;
        }
}

/* Per-64x64-region quarter-pel offset actually used by me_synth_frame
 * (for checking the recovered motion). */
EXPORT void me_synth_region_offset(u64 seed, int stream, int frame, int w, int region_qpel, int x, int y, int out[2])
{
    u64 kr = me_synth_key(seed, stream, frame, 2);
    int rcols = (w + 63) / 64;
    u64 ri = (u64)((y / 64) * rcols + (x / 64));
    int span = 2 * region_qpel + 1;
    out[0] = (int)(uni(kr, 2 * ri) * span) - region_qpel;
    out[1] = (int)(uni(kr, 2 * ri + 1) * span) - region_qpel;
}
