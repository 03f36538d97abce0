/*
 * Seeded synthetic systems of the shape the tdse hot path consumes:
 * energies E_L ~ U[e_lo, e_hi] sorted ascending per block (so c_0 = 1 is the
 * lowest L = 0 level, as tdse_main.cpp:148-150 assumes) and dipole blocks
 * D_L with i.i.d. N(0, sigma^2), sigma = 1 / sqrt(s_{L+1}).
 *
 * The generator is splitmix64 + Box-Muller so that the same seed gives the
 * same bits on every machine with the same libm; block L's energies use
 * stream 2L and pair L's dipoles stream 2L + 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "synth.h"

static uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

typedef struct {
  uint64_t s;
} sm64;

static void sm_init(sm64 *g, uint64_t seed, uint64_t stream) {
  g->s = mix(seed * 0x9E3779B97F4A7C15ULL + stream * 0xD1B54A32D192ED03ULL +
             0x632BE59BD9B4E019ULL);
}

static uint64_t sm_next(sm64 *g) {
  g->s += 0x9E3779B97F4A7C15ULL;
  return mix(g->s);
}

/* uniform in [0, 1) with 53 random bits */
static double sm_uniform(sm64 *g) {
  return (double)(sm_next(g) >> 11) * (1.0 / 9007199254740992.0);
}

static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

void mpci_synth_uniform(uint64_t seed, uint64_t stream, long long n,
                        double lo, double hi, double *out) {
  sm64 g;
  sm_init(&g, seed, stream);
  for (long long i = 0; i < n; ++i) out[i] = lo + (hi - lo) * sm_uniform(&g);
}

void mpci_synth_normal(uint64_t seed, uint64_t stream, long long n,
                       double sigma, double *out) {
  sm64 g;
  sm_init(&g, seed, stream);
  const double two_pi = 6.283185307179586476925286766559;
  for (long long i = 0; i < n; i += 2) {
    double u1 = 1.0 - sm_uniform(&g); /* (0, 1] */
    double u2 = sm_uniform(&g);
    double r = sqrt(-2.0 * log(u1));
    out[i] = sigma * r * cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = sigma * r * sin(two_pi * u2);
  }
}

int mpci_synth_system(uint64_t seed, int nb, const int *s, double e_lo,
                      double e_hi, double **E, double **D) {
  if (nb < 1 || !s || !E) return -1;
  for (int L = 0; L < nb; ++L) {
    if (s[L] < 1 || !E[L]) return -1;
    mpci_synth_uniform(seed, 2 * (uint64_t)L, s[L], e_lo, e_hi, E[L]);
    qsort(E[L], (size_t)s[L], sizeof(double), cmp_double);
  }
  for (int L = 0; L + 1 < nb; ++L) {
    if (!D || !D[L]) return -1;
    mpci_synth_normal(seed, 2 * (uint64_t)L + 1, (long long)s[L] * s[L + 1],
                      1.0 / sqrt((double)s[L + 1]), D[L]);
  }
  return 0;
}
