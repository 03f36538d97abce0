/* Seeded synthetic inputs (energies and dipole blocks) for the tdse path. */
#ifndef MPCI_SYNTH_H_
#define MPCI_SYNTH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void mpci_synth_uniform(uint64_t seed, uint64_t stream, long long n,
                        double lo, double hi, double *out);
void mpci_synth_normal(uint64_t seed, uint64_t stream, long long n,
                       double sigma, double *out);
/* E[L]: s[L] doubles (sorted ascending); D[L]: s[L] * s[L+1] row-major. */
int mpci_synth_system(uint64_t seed, int nb, const int *s, double e_lo,
                      double e_hi, double **E, double **D);

#ifdef __cplusplus
}
#endif
#endif
