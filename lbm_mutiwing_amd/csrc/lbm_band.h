// lbm_band.h - solve S y = c for one right-hand side by one wave, with the
// factors S = L D L^T of the banded IBM system (row-major band, row stride
// BW + 1: band[m*(BW+1) + d] = L(m, m-d) for 1 <= d <= bw, zero for
// bw < d <= BW and in the BW rows past the last; invd[m] = 1/D_m).
//
// Column-oriented substitution with a window of 64 rows that moves down one
// row per step: after the shift of step j, lane d holds the running value of
// row j + d.  The finished row j is read from its lane, lanes 1..bw subtract
// L(j+d, j) times it, and every lane takes the value of the lane above it
// (DPP wave_shl:1), the new row entering lane 63 as zero.  A step's
// dependency chain is one fused multiply-add and one lane read (the shift
// runs beside the read); the lane's band entry has a fixed offset per lane
// and is loaded RING steps ahead.  The right-hand side of the rows that
// entered during a block of 32 steps is added at the block's end, before any
// product reaches them (a row entering at lane 63 meets its first nonzero
// entry at lane bw <= 24, 39 steps or more later), so every row is summed in
// the order of a plain column-oriented substitution.
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

// v of lane l + 1 in lane l; 0 in lane 63 (DPP wave_shl:1, zero fill)
__device__ __forceinline__ double from_lane_above(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xF, 0xF, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v of lane l (uniform l)
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// p[idx] when ok, else 0; the load reads p[0] when not ok, so that it needs
// no branch
__device__ __forceinline__ double load_or_zero(const double* p, long idx, bool ok) {
  const double v = p[ok ? idx : 0];
  return ok ? v : 0.0;
}

constexpr int BAND_RING = 16;   // steps a band entry is loaded ahead of its use
constexpr int BAND_BLOCK = 32;  // steps between additions of right-hand sides
// doubles readable before and after a band (rows -BAND_PAD_ROWS.. and past
// the last): the look-ahead loads run past both ends unclamped
constexpr int BAND_PAD_ROWS = BAND_BLOCK + BAND_RING + 64;

// Forward pass L z = c over the first `steps` rows of an nrows-row band:
// z_j to y[ST*j] for j < steps; returns each lane's running value, so that
// afterwards lane l holds row steps - 1 + l, reduced by the products of
// every row before `steps` (the rows past it are those a twisted
// factorisation leaves to its middle block).  c and y have stride ST (2 or
// -2: one component of an interleaved pair, rows taken in the storage's
// order or in reverse).  band must have BAND_PAD_ROWS * (BW + 1) readable
// doubles after its row steps + BW.
template <int BW, int ST>
__device__ __forceinline__ double band_forward_wave(const double* __restrict__ band, int nrows, int steps,
                                                    const double* __restrict__ c, double* __restrict__ y) {
  static_assert(BW + 1 + BAND_BLOCK <= 64, "rows must meet their first product after their right-hand side");
  static_assert(BAND_BLOCK % BAND_RING == 0, "ring slots repeat every block");
  constexpr int Dd = BW + 1;
  const int lane = threadIdx.x & 63;
  const double mk = lane >= 1 && lane <= BW ? 1.0 : 0.0;    // lanes with a band entry
  const int dl = lane < 1 ? 1 : (lane > BW ? BW : lane);     // their entry; others read one and drop it
  double ring[BAND_RING];
  // before the shift of step j, lane l holds row j - 1 + l; the lane's entry
  // of step j is band[j*Dd + dl*(Dd + 1)] = L(j + dl, j)
  const unsigned off = (unsigned)(dl * (Dd + 1));
#pragma unroll
  for (int q = 0; q < BAND_RING; q++) ring[q] = band[(long)q * Dd + off];
  double f = load_or_zero(c, (long)ST * (lane - 1), lane >= 1 && lane - 1 < nrows);
  double cb = load_or_zero(c, (long)ST * (31 + lane), lane >= 32 && 31 + lane < nrows);
  int j = 0;
  for (; j + BAND_BLOCK <= steps; j += BAND_BLOCK) {
    const int r2 = j + BAND_BLOCK + 31 + lane;   // rows entering in the next block
    const double cb2 = load_or_zero(c, (long)ST * r2, lane >= 32 && r2 < nrows);
    const double* rows = band + (long)(j + BAND_RING) * Dd;
    double* yj = y + (long)ST * j;
#pragma unroll
    for (int s = 0; s < BAND_BLOCK; s++) {
      const double e = ring[s % BAND_RING] * mk;
      ring[s % BAND_RING] = rows[s * Dd + off];
      const double z = lane_value(f, 1);
      f = __builtin_fma(-e, z, from_lane_above(f));
      yj[ST * s] = z;
    }
    f += cb;   // rows j+63 .. j+94, in lanes 32..63
    cb = cb2;
  }
  const double* rows = band + (long)(j + BAND_RING) * Dd;
#pragma unroll
  for (int s = 0; s < BAND_BLOCK; s++) {
    if (j + s < steps) {
      const double e = ring[s % BAND_RING] * mk;
      ring[s % BAND_RING] = rows[s * Dd + off];
      const double z = lane_value(f, 1);
      f = __builtin_fma(-e, z, from_lane_above(f));
      y[(long)ST * (j + s)] = z;
    }
  }
  return f;
}

// Backward pass L^T x = D^-1 z over rows nrows-1 .. 0: y holds z (and, for
// rows whose 1/D is 1 and whose band entries among themselves are zero, the
// finished values of a twisted factorisation's middle block) and receives x.
// band must have BAND_PAD_ROWS * (BW + 1) readable doubles before it.
template <int BW, int ST>
__device__ __forceinline__ void band_backward_wave(const double* __restrict__ band, const double* __restrict__ invd,
                                                   int nrows, double* __restrict__ y) {
  constexpr int Dd = BW + 1;
  const int lane = threadIdx.x & 63;
  const double mk = lane >= 1 && lane <= BW ? 1.0 : 0.0;
  const int dl = lane < 1 ? 1 : (lane > BW ? BW : lane);
  double ring[BAND_RING];
  // before the shift of step k (descending), lane l holds row k + 1 - l; a
  // row enters with t = z / D.  The lane's entry of step k is
  // band[k*Dd + dl] = L(k, k - dl).
  const unsigned off = (unsigned)dl;
  const int top = nrows - 1;
#pragma unroll
  for (int q = 0; q < BAND_RING; q++) ring[q] = band[(long)(top - q) * Dd + off];
  const int m0 = nrows - lane;
  double f = m0 >= 0 && lane >= 1 ? y[(long)ST * m0] * invd[m0] : 0.0;
  int mb = top - 31 - lane;   // rows entering in the first block (lanes 32..63)
  double cb = lane >= 32 && mb >= 0 ? y[(long)ST * mb] * invd[mb] : 0.0;
  int k = top;
  for (; k - (BAND_BLOCK - 1) >= 0; k -= BAND_BLOCK) {
    const int m2 = mb - BAND_BLOCK;
    const double cb2 = lane >= 32 && m2 >= 0 ? y[(long)ST * m2] * invd[m2] : 0.0;
    const double* rows = band + (long)(k - BAND_RING) * Dd;
    double* yk = y + (long)ST * k;
#pragma unroll
    for (int s = 0; s < BAND_BLOCK; s++) {
      const double e = ring[s % BAND_RING] * mk;
      ring[s % BAND_RING] = rows[-s * Dd + (int)off];
      const double x = lane_value(f, 1);
      f = __builtin_fma(-e, x, from_lane_above(f));
      yk[-ST * s] = x;
    }
    f += cb;
    cb = cb2;
    mb = m2;
  }
  const double* rows = band + (long)(k - BAND_RING) * Dd;
#pragma unroll
  for (int s = 0; s < BAND_BLOCK; s++) {
    if (k - s >= 0) {
      const double e = ring[s % BAND_RING] * mk;
      ring[s % BAND_RING] = rows[-s * Dd + (int)off];
      const double x = lane_value(f, 1);
      f = __builtin_fma(-e, x, from_lane_above(f));
      y[(long)ST * (k - s)] = x;
    }
  }
}

// y[2m] = (S^-1 c)_m for m < n, c and y with stride 2 (one component of an
// interleaved pair); y holds z = L^-1 c between the two passes.  Every lane
// stores the value finished at a step (the same value to the same address).
// band must have BAND_PAD_ROWS * (BW + 1) readable doubles before it and
// after its last row (n + BW - 1).
template <int BW>
__device__ __forceinline__ void band_solve_wave(const double* __restrict__ band, const double* __restrict__ invd,
                                                int n, const double* __restrict__ c, double* __restrict__ y) {
  (void)band_forward_wave<BW, 2>(band, n, n, c, y);
  band_backward_wave<BW, 2>(band, invd, n, y);
}

}  // namespace lbm
