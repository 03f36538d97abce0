// Producers of the tdse dipole inputs on the GPU.  The w2e stage:
// symmetric eigendecomposition of the two-electron Hamiltonian per total
// angular momentum (rocSOLVER dsyevd) and the projection of the
// product-basis dipoles onto the CI eigenbasis (two rocBLAS dgemm, in the
// reference's order).  The d1e stage: one-electron dipole blocks by
// quadrature, a scaled product of the radial-function samples (one fused
// scaling kernel, one rocBLAS dgemm).  C ABI in include/mpci_ci.h.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/mpci_ci.h"
#include "../../include/mpci_tdse.h"

namespace {

thread_local std::string g_ci_err;

int ci_fail(int code, const std::string &msg) {
  g_ci_err = msg;
  return code;
}

#define CI_HIP(x)                                                         \
  do {                                                                    \
    hipError_t e_ = (x);                                                  \
    if (e_ != hipSuccess)                                                 \
      return ci_fail(MPCI_ERR_HIP, std::string(#x ": ") +                 \
                                       hipGetErrorString(e_));            \
  } while (0)
#define CI_BLAS(x)                                                        \
  do {                                                                    \
    rocblas_status s_ = (x);                                              \
    if (s_ != rocblas_status_success)                                     \
      return ci_fail(MPCI_ERR_HIP, std::string(#x ": ") +                 \
                                       rocblas_status_to_string(s_));     \
  } while (0)

template <typename T>
struct DevMem {
  T *p = nullptr;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, sizeof(T) * (n ? n : 1)); }
};

struct Blas {
  rocblas_handle h = nullptr;
  ~Blas() {
    if (h) (void)rocblas_destroy_handle(h);
  }
};

// H = diag(ens) + V12 in full (column-major, both triangles), from the
// packed upper triangle: element (i, j), i <= j, at (2n - i - 1) i / 2 + j.
__global__ void k_ci_assemble(int n, const double *__restrict__ ens,
                              const double *__restrict__ v12,
                              double *__restrict__ A) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int c = (int)(e / n), r = (int)(e - (long long)c * n);
  const long long i = r < c ? r : c, j = r < c ? c : r;
  double v = v12[(2LL * n - i - 1) * i / 2 + j];
  if (i == j) v += ens[i];
  A[e] = v;
}

// Sign convention: in each eigenvector (column k, contiguous), the component
// of largest magnitude -- the first one on a tie -- is made positive.  One
// workgroup per eigenvector; the reduction order is fixed.
__global__ __launch_bounds__(256) void k_ci_fix_signs(int n,
                                                      double *__restrict__ A) {
  const int k = blockIdx.x, t = threadIdx.x;
  double *v = A + (long long)k * n;
  double best = -1.0;
  int bi = n;
  for (int r = t; r < n; r += 256) {
    const double a = fabs(v[r]);
    if (a > best) {  // strided scan: the first index wins within a lane
      best = a;
      bi = r;
    }
  }
  __shared__ double sb[256];
  __shared__ int si[256];
  sb[t] = best;
  si[t] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      const double b2 = sb[t + w];
      const int i2 = si[t + w];
      if (b2 > sb[t] || (b2 == sb[t] && i2 < si[t])) {
        sb[t] = b2;
        si[t] = i2;
      }
    }
    __syncthreads();
  }
  const bool flip = si[0] < n && v[si[0]] < 0.0;
  __syncthreads();
  if (flip)
    for (int r = t; r < n; r += 256) v[r] = -v[r];
}

// The d1e integrand's second factor times the quadrature weight, per
// (state, node): B[n][i] = qw_i x_i P(x_i) (length) or
// qw_i (P'(x_i) - v2 P(x_i) / x_i) (velocity).
__global__ void k_d1e_weights(int nen, int qsz, int velocity, double v2,
                              const double *__restrict__ qx,
                              const double *__restrict__ qw,
                              const double *__restrict__ w,
                              const double *__restrict__ wp,
                              double *__restrict__ B) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nen * qsz) return;
  const int i = (int)(e % qsz);
  const double x = qx[i];
  B[e] = velocity ? qw[i] * (wp[e] - v2 * w[e] / x) : qw[i] * x * w[e];
}

}  // namespace

extern "C" const char *mpci_ci_last_error(void) { return g_ci_err.c_str(); }

extern "C" int mpci_d1e_dipoles(int nen, int qsz, int l, char gauge,
                                const double *qx, const double *qw,
                                const double *wfn_l, const double *wfn_l1,
                                const double *wfnp_l1, double *out) {
  const bool vel = gauge == 'v';
  if (nen < 1 || qsz < 1 || l < 0 || (gauge != 'v' && gauge != 'l') || !qx ||
      !qw || !wfn_l || !wfn_l1 || (vel && !wfnp_l1) || !out)
    return ci_fail(MPCI_ERR_ARG, "mpci_d1e_dipoles: bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return ci_fail(MPCI_ERR_HIP, "no GPU: the 1e dipoles are computed on the "
                                 "GPU only");
  const long long nq = (long long)nen * qsz;
  DevMem<double> dx, dw, dW1, dW2, dWp, dB, dO;
  CI_HIP(dx.alloc(qsz));
  CI_HIP(dw.alloc(qsz));
  CI_HIP(dW1.alloc(nq));
  CI_HIP(dW2.alloc(nq));
  CI_HIP(dB.alloc(nq));
  CI_HIP(dO.alloc((long long)nen * nen));
  CI_HIP(hipMemcpy(dx.p, qx, sizeof(double) * qsz, hipMemcpyHostToDevice));
  CI_HIP(hipMemcpy(dw.p, qw, sizeof(double) * qsz, hipMemcpyHostToDevice));
  CI_HIP(hipMemcpy(dW1.p, wfn_l, sizeof(double) * nq, hipMemcpyHostToDevice));
  CI_HIP(hipMemcpy(dW2.p, wfn_l1, sizeof(double) * nq,
                   hipMemcpyHostToDevice));
  if (vel) {
    CI_HIP(dWp.alloc(nq));
    CI_HIP(hipMemcpy(dWp.p, wfnp_l1, sizeof(double) * nq,
                     hipMemcpyHostToDevice));
  }
  // the reference's integer expressions (integrator.cpp:9, dmx1e.cpp:72)
  const int lp1 = l + 1;
  const double v2 = (l * (l + 1) - lp1 * (lp1 + 1)) * 0.5;
  const double kl = sqrt((double)(lp1 * lp1) / (4.0 * lp1 * lp1 - 1));
  hipLaunchKernelGGL(k_d1e_weights, dim3((unsigned)((nq + 255) / 256)),
                     dim3(256), 0, 0, nen, qsz, vel ? 1 : 0, v2, dx.p, dw.p,
                     dW2.p, vel ? dWp.p : dW2.p, dB.p);
  CI_HIP(hipGetLastError());
  Blas b;
  CI_BLAS(rocblas_create_handle(&b.h));
  const double zero = 0.0;
  // Column-major views of the row-major nen x qsz arrays are their
  // transposes: out^T (nen x nen) = B W_l^T, i.e. out = k_l W_l B^T.
  CI_BLAS(rocblas_dgemm(b.h, rocblas_operation_transpose,
                        rocblas_operation_none, nen, nen, qsz, &kl, dB.p, qsz,
                        dW1.p, qsz, &zero, dO.p, nen));
  CI_HIP(hipDeviceSynchronize());
  CI_HIP(hipMemcpy(out, dO.p, sizeof(double) * nen * nen,
                   hipMemcpyDeviceToHost));
  return MPCI_OK;
}

extern "C" int mpci_ci_basis(int n, const double *ens, const double *v12_packed,
                             double *eig, double *vecs) {
  if (n < 1 || !ens || !v12_packed || !eig || !vecs)
    return ci_fail(MPCI_ERR_ARG, "mpci_ci_basis: bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return ci_fail(MPCI_ERR_HIP, "no GPU: the CI basis is computed on the GPU "
                                 "only");
  const long long nn = (long long)n * n;
  const long long np = (long long)n * (n + 1) / 2;
  DevMem<double> dA, dEns, dV, dD, dE;
  DevMem<rocblas_int> dInfo;
  CI_HIP(dA.alloc(nn));
  CI_HIP(dEns.alloc(n));
  CI_HIP(dV.alloc(np));
  CI_HIP(dD.alloc(n));
  CI_HIP(dE.alloc(n));
  CI_HIP(dInfo.alloc(1));
  CI_HIP(hipMemcpy(dEns.p, ens, sizeof(double) * n, hipMemcpyHostToDevice));
  CI_HIP(hipMemcpy(dV.p, v12_packed, sizeof(double) * np,
                   hipMemcpyHostToDevice));
  const unsigned nblk = (unsigned)((nn + 255) / 256);
  hipLaunchKernelGGL(k_ci_assemble, dim3(nblk), dim3(256), 0, 0, n, dEns.p,
                     dV.p, dA.p);
  CI_HIP(hipGetLastError());
  Blas b;
  CI_BLAS(rocblas_create_handle(&b.h));
  CI_BLAS(rocsolver_dsyevd(b.h, rocblas_evect_original, rocblas_fill_lower, n,
                           dA.p, n, dD.p, dE.p, dInfo.p));
  hipLaunchKernelGGL(k_ci_fix_signs, dim3((unsigned)n), dim3(256), 0, 0, n,
                     dA.p);
  CI_HIP(hipGetLastError());
  CI_HIP(hipDeviceSynchronize());
  rocblas_int info = 0;
  CI_HIP(hipMemcpy(&info, dInfo.p, sizeof(info), hipMemcpyDeviceToHost));
  if (info != 0)
    return ci_fail(MPCI_ERR_STATE, "dsyevd did not converge (info " +
                                       std::to_string(info) + ")");
  CI_HIP(hipMemcpy(eig, dD.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  CI_HIP(hipMemcpy(vecs, dA.p, sizeof(double) * nn, hipMemcpyDeviceToHost));
  return MPCI_OK;
}

extern "C" int mpci_ci_dipoles(int n0, int n1, const double *D,
                               const double *vecs0, const double *vecs1,
                               double *out) {
  if (n0 < 1 || n1 < 1 || !D || !vecs0 || !vecs1 || !out)
    return ci_fail(MPCI_ERR_ARG, "mpci_ci_dipoles: bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return ci_fail(MPCI_ERR_HIP, "no GPU: the CI dipoles are computed on the "
                                 "GPU only");
  const long long nd = (long long)n0 * n1;
  DevMem<double> dD, dV0, dV1, dT;
  CI_HIP(dD.alloc(nd));
  CI_HIP(dV0.alloc((long long)n0 * n0));
  CI_HIP(dV1.alloc((long long)n1 * n1));
  CI_HIP(dT.alloc(nd));
  CI_HIP(hipMemcpy(dD.p, D, sizeof(double) * nd, hipMemcpyHostToDevice));
  CI_HIP(hipMemcpy(dV0.p, vecs0, sizeof(double) * n0 * n0,
                   hipMemcpyHostToDevice));
  CI_HIP(hipMemcpy(dV1.p, vecs1, sizeof(double) * n1 * n1,
                   hipMemcpyHostToDevice));
  Blas b;
  CI_BLAS(rocblas_create_handle(&b.h));
  const double one = 1.0, zero = 0.0;
  // Column-major views: the row-major D (n0 x n1) is D^T (n1 x n0); the
  // eigenvector files are the matrices whose columns are eigenvectors.
  // T = D^T V_L, then out^T = V_{L+1}^T T, i.e. out = V_L^T D V_{L+1}.
  CI_BLAS(rocblas_dgemm(b.h, rocblas_operation_none, rocblas_operation_none,
                        n1, n0, n0, &one, dD.p, n1, dV0.p, n0, &zero, dT.p,
                        n1));
  CI_BLAS(rocblas_dgemm(b.h, rocblas_operation_transpose,
                        rocblas_operation_none, n1, n0, n1, &one, dV1.p, n1,
                        dT.p, n1, &zero, dD.p, n1));
  CI_HIP(hipDeviceSynchronize());
  CI_HIP(hipMemcpy(out, dD.p, sizeof(double) * nd, hipMemcpyDeviceToHost));
  return MPCI_OK;
}
