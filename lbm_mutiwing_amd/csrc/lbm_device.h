// lbm_device.h - device-side arithmetic shared by the gfx950 kernels.
//
// Every expression keeps the evaluation order of the reference Fortran
// (a*b*c = (a*b)*c, left-to-right sums) and the library is compiled with
// -ffp-contract=off, so the FP64 lattice arithmetic is bit-identical to a
// sequential evaluation of the reference formulas.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

// D2Q9 velocity set (D2Q9model.f90:5-22), 0-based k.
__host__ __device__ constexpr int EX(int k) {
  return k == 1 || k == 5 || k == 8 ? 1 : (k == 3 || k == 6 || k == 7 ? -1 : 0);
}
__host__ __device__ constexpr int EY(int k) {
  return k == 2 || k == 5 || k == 6 ? 1 : (k == 4 || k == 7 || k == 8 ? -1 : 0);
}

// MRT parameters passed by value to every kernel (kernarg -> SGPRs).
struct MrtParams {
  double s[9];      // relaxation rates (ini_mesh.f90:59-67)
  double minv[81];  // inverse of M as computed by nizhen (solution.f90:1474-1529)
  double cs;        // sqrt(3.0)/3 with a single-precision sqrt (boundary_condition.f90:304)
  double w[9];      // lattice weights (single-precision literals, D2Q9model.f90:21-29)
  // FP32-shifted storage: populations are kept as d = f - w in FP32 and the
  // bulk collision runs on d in FP64 with these constants.  With
  // c = (1/9, 1/36, 1/36, 1/6, 1/12, 1/6, 1/12, 1/4, 1/4) (M^-1 = M^T diag c):
  double ra[9];     // c_r (1 - s_r)
  double rb[9];     // c_r s_r
  double rk0;       // -c_0 s_0 sum w: the density's equilibrium moment is 0 (solution.f90:613)
  double eq1, eq2;  // -2 sum w - (M w)_1 and sum w - (M w)_2: equilibrium moments minus those of w
  double wsum;      // sum of the weights (1 + 7.45e-9 with single-precision literals)
};

// Cache policies of the lattice kernels' population accesses.
enum { CP_PLAIN = 0, CP_NT_STORE = 1, CP_NT = 2 };

// Lattice geometry shared by the kernels.
struct Grid {
  int nx, ny;       // nodes
  long long ps;     // plane stride (elements) between populations
};

// m = M f with M of solution.f90:17-25, summed in column order as matmulvoc
// (boundary_condition.f90:1738-1755) does; exact-zero terms are skipped,
// which leaves every partial sum unchanged.
__device__ __forceinline__ void moments_mf(const double f[9], double m[9]) {
  m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  m[1] = -4.0 * f[0] - f[1] - f[2] - f[3] - f[4] + 2.0 * f[5] + 2.0 * f[6] + 2.0 * f[7] + 2.0 * f[8];
  m[2] = 4.0 * f[0] - 2.0 * f[1] - 2.0 * f[2] - 2.0 * f[3] - 2.0 * f[4] + f[5] + f[6] + f[7] + f[8];
  m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
  m[4] = -2.0 * f[1] + 2.0 * f[3] + f[5] - f[6] - f[7] + f[8];
  m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
  m[6] = -2.0 * f[2] + 2.0 * f[4] + f[5] + f[6] - f[7] - f[8];
  m[7] = f[1] - f[2] + f[3] - f[4];
  m[8] = f[5] - f[6] + f[7] - f[8];
}

// f = Minv m* (matmulvoc with the nizhen inverse).  The zero pattern of the
// inverse is structural: row 0 uses m0..m2, rows 1-4 use m0..m2 and one of
// (m3,m4) / (m5,m6) plus m7, rows 5-8 use m0..m6 and m8.
__device__ __forceinline__ void inverse_mf(const double* __restrict__ mi, const double m[9], double f[9]) {
  f[0] = mi[0] * m[0] + mi[1] * m[1] + mi[2] * m[2];
#pragma unroll
  for (int r = 1; r <= 3; r += 2) {
    const double* a = mi + 9 * r;
    f[r] = a[0] * m[0] + a[1] * m[1] + a[2] * m[2] + a[3] * m[3] + a[4] * m[4] + a[7] * m[7];
  }
#pragma unroll
  for (int r = 2; r <= 4; r += 2) {
    const double* a = mi + 9 * r;
    f[r] = a[0] * m[0] + a[1] * m[1] + a[2] * m[2] + a[5] * m[5] + a[6] * m[6] + a[7] * m[7];
  }
#pragma unroll
  for (int r = 5; r <= 8; r++) {
    const double* a = mi + 9 * r;
    f[r] = a[0] * m[0] + a[1] * m[1] + a[2] * m[2] + a[3] * m[3] + a[4] * m[4] + a[5] * m[5] +
           a[6] * m[6] + a[8] * m[8];
  }
}

// MRT collision of one node (colission, solution.f90:599-650): m = M f,
// incompressible m_eq, Guo forcing in moment space, relaxation, f = Minv m*.
__device__ __forceinline__ void collide_node(const MrtParams& P, const double f[9], double rt, double ut,
                                             double vt, double gfx, double gfy, double out[9]) {
  double m[9], t[9], fp[9];
  moments_mf(f, m);
  const double* s = P.s;
  t[0] = 0.0;
  t[1] = -2.0 * rt + 3.0 * (rt * ut * rt * ut + rt * vt * rt * vt);
  t[2] = rt - 3.0 * (rt * ut * rt * ut + rt * vt * rt * vt);
  t[3] = 0.0;
  t[4] = -(rt * ut);
  t[5] = 0.0;
  t[6] = -(rt * vt);
  t[7] = (rt * ut * rt * ut - rt * vt * rt * vt);
  t[8] = (rt * ut * rt * vt);
  fp[0] = 0.0;
  fp[1] = 6.0 * (1.0 - 0.5 * s[1]) * (ut * gfx + vt * gfy);
  fp[2] = -(6.0 * (1.0 - 0.5 * s[2]) * (ut * gfx + vt * gfy));
  fp[3] = gfx;
  fp[4] = -((1.0 - 0.5 * s[4]) * gfx);
  fp[5] = gfy;
  fp[6] = -((1.0 - 0.5 * s[6]) * gfy);
  fp[7] = 2.0 * (1.0 - 0.5 * s[7]) * (ut * gfx - vt * gfy);
  fp[8] = (1.0 - 0.5 * s[8]) * (ut * gfy + vt * gfx);
#pragma unroll
  for (int k = 0; k < 9; k++) t[k] = m[k] + (t[k] - m[k]) * s[k] + fp[k];
  inverse_mf(P.minv, t, out);
}

// Compressible moment equilibrium of the far-field BC
// (boundary_condition.f90:444-455).
__device__ __forceinline__ void meq_bc(double rh, double uu, double vv, double t[9]) {
  const double jx = rh * uu, jy = rh * vv;
  t[0] = rh;
  t[1] = -2.0 * rh + 3.0 * (jx * jx + jy * jy) / rh;
  t[2] = rh - 3.0 * (jx * jx + jy * jy) / rh;
  t[3] = jx;
  t[4] = -jx;
  t[5] = jy;
  t[6] = -jy;
  t[7] = (jx * jx - jy * jy) / rh;
  t[8] = (jx * jy) / rh;
}

// f(b) = Minv [ m_eq(b) + (1-s)(M f0(n) - m_eq(n)) ] (boundary_condition.f90:443-474).
__device__ __forceinline__ void bc_extrapolate(const MrtParams& P, double rb, double ub, double vb,
                                               double rn, double un, double vn, const double f0n[9],
                                               double out[9]) {
  double t1[9], t2[9], t3[9];
  meq_bc(rb, ub, vb, t1);
  meq_bc(rn, un, vn, t2);
  moments_mf(f0n, t3);
#pragma unroll
  for (int k = 0; k < 9; k++) t3[k] = t1[k] + (1.0 - P.s[k]) * (t3[k] - t2[k]);
  inverse_mf(P.minv, t3, out);
}

// Pull the 9 streamed populations of interior node (i, j) (0-based) from f
// (streaming, solution.f90:450-466) and form rho, u, v as physical_variables
// (solution.f90:523-549): rho = m0, jx = m3, jy = m5 of the same sums.
__device__ __forceinline__ void pull_node(const double* __restrict__ f, const Grid& g, int i, int j,
                                          double fo[9]) {
#pragma unroll
  for (int k = 0; k < 9; k++)
    fo[k] = f[(long long)k * g.ps + (long long)(j - EY(k)) * g.nx + (i - EX(k))];
}

// Storage of the populations: FP64 values, or FP32 deviations d = f - w_k
// from the rest equilibrium.  Loads return the FP64 value of f.
template <typename S>
struct Store;
template <>
struct Store<double> {
  static __device__ __forceinline__ double ld(const double* __restrict__ f, long long q, double) { return f[q]; }
  static __device__ __forceinline__ void st(double* __restrict__ f, long long q, double v, double) { f[q] = v; }
};
template <>
struct Store<float> {
  static __device__ __forceinline__ double ld(const float* __restrict__ f, long long q, double wk) {
    return wk + (double)f[q];
  }
  static __device__ __forceinline__ void st(float* __restrict__ f, long long q, double v, double wk) {
    f[q] = (float)(v - wk);
  }
};

template <typename S>
__device__ __forceinline__ void pull_node_s(const S* __restrict__ f, const Grid& g, int i, int j, const double* w,
                                            double fo[9]) {
#pragma unroll
  for (int k = 0; k < 9; k++)
    fo[k] = Store<S>::ld(f, (long long)k * g.ps + (long long)(j - EY(k)) * g.nx + (i - EX(k)), w[k]);
}

__device__ __forceinline__ void node_moments(const double fo[9], double& rho, double& u, double& v) {
  const double r = fo[0] + fo[1] + fo[2] + fo[3] + fo[4] + fo[5] + fo[6] + fo[7] + fo[8];
  const double jx = fo[1] - fo[3] + fo[5] - fo[6] - fo[7] + fo[8];
  const double jy = fo[2] - fo[4] + fo[5] + fo[6] - fo[7] - fo[8];
  rho = r;
  u = jx / r;
  v = jy / r;
}

// Smoothed delta kernel, smooth=1 branch (boundary_condition.f90:1581-1600),
// device asin/sqrt.
__device__ __forceinline__ double delta_kernel(double disr) {
  const double pai = 3.1415926535897932384626433832795;
  if (fabs(disr) <= 0.5) return 0.375 + 0.03125 * pai - 0.25 * disr * disr;
  if (fabs(disr) <= 1.5 && fabs(disr) >= 0.5)
    return 0.25 + 0.125 * (1.0 - disr) * sqrt(-2.0 + 8.0 * disr - 4.0 * disr * disr) -
           0.125 * asin(sqrt(2.0) * (disr - 1.0));
  if (fabs(disr) <= 2.5 && fabs(disr) >= 1.5)
    return 1.0625 - 0.015625 * pai - 0.75 * disr + 0.125 * disr * disr +
           0.0625 * (disr - 2.0) * sqrt(-14.0 + 16.0 * disr - 4.0 * disr * disr) +
           0.0625 * asin(sqrt(2.0) * (disr - 2.0));
  return 0.0;
}

}  // namespace lbm
