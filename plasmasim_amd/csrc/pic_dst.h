// Batched DST-I for the Dirichlet Poisson corrections: the 2D variant's
// renew_phi (PIC2D :2589-2767) and the 3D renew_phi3 (:8315-8665), which
// apply the Numerical-Recipes sine transform sinft (:8182-8215 in 3D) line
// by line.  Both sizes are powers of two.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

namespace pic {

template <typename T>
struct Cplx {
  T re, im;
};

// DST-I of two lines per workgroup through one complex FFT of length 2n:
// z = a + i b, odd-extended (z_0 = z_n = 0, z_{2n-j} = -z_j), so that
// FFT(z)_k = -2i S_a(k) + 2 S_b(k) with S(k) = sum_{j=1}^{n-1} y_j
// sin(pi j k / n).  Line l starts at (l / ninner) * souter + (l % ninner) *
// sline and its element j is j * selem further; y_0 is the Dirichlet point
// and is taken as zero (sinft sets y[1] = 0).  Radix-2 decimation-in-time in
// LDS (4n values of T), input in bit-reversed order.
//   mode 0: out = S                        (forward)
//   mode 1: out = S / A(l, k), 0 if A = 0   (forward, divided by the symbol
//           A = -(sym_l[l] + sym_k[k]) / dq of the standard Laplacian)
//   mode 2: out = 2 S / n                  (inverse)
template <typename T>
__global__ __launch_bounds__(256) void k_dst(T *__restrict__ data, int n, int logL,
                                             int nlines, int ninner, int64_t sline,
                                             int64_t souter, int64_t selem,
                                             const Cplx<T> *__restrict__ tw, int mode,
                                             const double *__restrict__ sym_l,
                                             const double *__restrict__ sym_k,
                                             double dq) {
  extern __shared__ unsigned char lds_raw[];
  T *re = (T *)lds_raw;
  const int L = 2 * n;
  T *im = re + L;
  const int l0 = 2 * blockIdx.x;
  const bool has_b = l0 + 1 < nlines;
  const int64_t base_a = (int64_t)(l0 / ninner) * souter + (int64_t)(l0 % ninner) * sline;
  const int64_t base_b =
      (int64_t)((l0 + 1) / ninner) * souter + (int64_t)((l0 + 1) % ninner) * sline;
  const T *A = data + base_a;
  const T *B = data + base_b;
  for (int j = threadIdx.x; j < L; j += blockDim.x) {
    T a = 0, b = 0;
    if (j != 0 && j != n) {
      const int s = j < n ? j : L - j;
      a = A[(int64_t)s * selem];
      b = has_b ? B[(int64_t)s * selem] : (T)0;
      if (j > n) {
        a = -a;
        b = -b;
      }
    }
    const int r = (int)(__brev((unsigned)j) >> (32 - logL));
    re[r] = a;
    im[r] = b;
  }
  __syncthreads();
  for (int h = 1; h < L; h <<= 1) {
    const int tstride = L / (2 * h);
    for (int t = threadIdx.x; t < L / 2; t += blockDim.x) {
      const int j = t & (h - 1);
      const int i0 = ((t - j) << 1) + j, i1 = i0 + h;
      const Cplx<T> w = tw[j * tstride];
      const T vr = re[i1] * w.re - im[i1] * w.im;
      const T vi = re[i1] * w.im + im[i1] * w.re;
      const T ur = re[i0], ui = im[i0];
      re[i0] = ur + vr;
      im[i0] = ui + vi;
      re[i1] = ur - vr;
      im[i1] = ui - vi;
    }
    __syncthreads();
  }
  T *OA = data + base_a;
  T *OB = data + base_b;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    T sa = -im[k] / (T)2, sb = re[k] / (T)2;
    if (mode == 1) {
      const double Aa = -(sym_l[l0] + sym_k[k]) / dq;
      sa = Aa == 0. ? (T)0 : (T)(sa / (T)Aa);
      if (has_b) {
        const double Ab = -(sym_l[l0 + 1] + sym_k[k]) / dq;
        sb = Ab == 0. ? (T)0 : (T)(sb / (T)Ab);
      }
    } else if (mode == 2) {
      sa = (T)2 * sa / (T)n;
      sb = (T)2 * sb / (T)n;
    }
    OA[(int64_t)k * selem] = sa;
    if (has_b) OB[(int64_t)k * selem] = sb;
  }
}

inline int ilog2_ceil(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// host tables of one axis of length N: L/2 = N complex twiddles
// exp(-2 pi i t / 2N) (in T) followed by N doubles sin^2(pi k / (2N)), the
// reference's symbol terms (:8502-8504; PIC2D :2679-2686); returns a
// malloc'd buffer of *bytes bytes
template <typename T>
unsigned char *dst_axis_tables(int N, size_t *bytes, size_t *tw_bytes) {
  const double Pi = 3.14159265358979323846;
  const int L = 2 * N;
  *tw_bytes = (size_t)(L / 2) * sizeof(Cplx<T>);
  *bytes = *tw_bytes + (size_t)N * sizeof(double);
  unsigned char *h = (unsigned char *)malloc(*bytes);
  if (!h) return nullptr;
  Cplx<T> *tw = (Cplx<T> *)h;
  for (int t = 0; t < L / 2; ++t) {
    const double a = 2.0 * Pi * (double)t / (double)L;
    tw[t].re = (T)cos(a);
    tw[t].im = (T)-sin(a);
  }
  double *sym = (double *)(h + *tw_bytes);
  for (int k = 0; k < N; ++k) sym[k] = pow(sin(Pi * (double)k / (double)N / 2.), 2);
  return h;
}

}  // namespace pic
