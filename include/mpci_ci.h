/*
 * mpci_ci.h -- C ABI of the MI355X producers of the tdse dipole inputs: the
 * configuration-interaction (CI) basis of MP-CITDSE's `w2e` stage, which
 * turns the two-electron product basis into the CI eigenbasis and the CI
 * dipole blocks that `tdse -e 2` reads (En_CI, CI_dip), and the one-electron
 * dipole blocks of its `d1e` stage that `tdse -e 1` reads (d_if).
 *
 * Plain C types only; every function returns MPCI_OK (0) or a negative
 * MPCI_ERR_* code (the codes of mpci_tdse.h) and mpci_ci_last_error() has
 * the message.  Host buffers are copied; none is retained.  All FP64.
 *
 * Layouts follow the files w2e writes:
 *   vecs (n x n): row k is the k-th eigenvector (the column-major eigenvector
 *                 matrix LAPACK returns, stored as w2e stores CI_vecs);
 *   dipoles:      row-major, n_L x n_{L+1}.
 * Eigenvector signs are fixed by a convention of this library: each
 * eigenvector's component of largest magnitude (the first one on a tie) is
 * positive.  LAPACK leaves the sign to the implementation, so the reference
 * has no canonical sign; physical quantities (energies, populations,
 * |dipole|) do not depend on it.
 */
#ifndef MPCI_CI_H_
#define MPCI_CI_H_

#ifdef __cplusplus
extern "C" {
#endif

const char *mpci_ci_last_error(void);

/* w2e::formCIh0 for one total angular momentum L (src/w2e/w2e.cpp:17-82):
 * H = diag(ens) + V12, with V12 given as the packed upper triangle the
 * reference reads from <pot>V12_<L>.h5:/V_12 (row i holds columns i..n-1,
 * element (i, j) at (2n - i - 1) i / 2 + j), diagonalised (LAPACKE_dsyevd
 * there, rocSOLVER dsyevd here).  eig: n eigenvalues, ascending; vecs: n x n
 * as above. */
int mpci_ci_basis(int n, const double *ens, const double *v12_packed,
                  double *eig, double *vecs);

/* w2e::formCIDipoles for one pair (src/w2e/w2e.cpp:84-125):
 * out = V_L^T D V_{L+1} (row-major n0 x n1), D = d_if (row-major n0 x n1),
 * V_L / V_{L+1} the eigenvector matrices of mpci_ci_basis (rows =
 * eigenvectors).  The products are formed in the reference's order,
 * (V_L^T D) then times V_{L+1}. */
int mpci_ci_dipoles(int n0, int n1, const double *D, const double *vecs0,
                    const double *vecs1, double *out);

/* dmx1e::genDipole for one l (src/d1e/dmx1e.cpp:24-139 with
 * dmx_int::tlenGL / tvelGL, src/d1e/integrator.cpp:3-31): the one-electron
 * dipole block between angular momenta l and l + 1 from radial functions
 * sampled on the outer quadrature (qsz nodes qx, weights qw):
 *   out[n1][n2] = k_l sum_i qw_i P_{n1,l}(x_i) g_{n2}(x_i),
 *   k_l = sqrt((l+1)^2 / (4 (l+1)^2 - 1)),
 *   gauge 'l': g = x P_{n2,l+1};
 *   gauge 'v': g = P'_{n2,l+1} - v2 P_{n2,l+1} / x,
 *              v2 = (l (l+1) - (l+1)(l+2)) / 2.
 * wfn_l, wfn_l1, wfnp_l1: nen x qsz row-major (w1e's Pr_o of l and l + 1
 * and Pr_p of l + 1; wfnp_l1 may be NULL for 'l').  out: nen x nen
 * row-major (the d_if dataset). */
int mpci_d1e_dipoles(int nen, int qsz, int l, char gauge, const double *qx,
                     const double *qw, const double *wfn_l,
                     const double *wfn_l1, const double *wfnp_l1,
                     double *out);

#ifdef __cplusplus
}
#endif
#endif /* MPCI_CI_H_ */
