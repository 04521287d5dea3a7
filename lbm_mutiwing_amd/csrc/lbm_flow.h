// lbm_flow.h - flow part of the implicit immersed boundary, one step:
// device bodies of the IBM work items of the step kernel (lbm_lattice.hip).
//
//   rhs     c = U_b - D u* (boundary_condition.f90:1609-1629), one wave per
//           segment, u* pulled from the step's source populations
//   matvec  y = S^-1 c, one wave per row of the explicit inverse
//   spread  g = D^T y at every adjoint node (boundary_condition.f90:1639-1651),
//           staged per block of segments for the adjoint items
// The geometry they read (boxes, delta weights, adjoint lists, S^-1) is
// computed ahead of the flow by the geometry kernels (lbm_ibm.hip).
#pragma once

#include <climits>

#include "lbm_internal.h"

namespace lbm {

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// c = U_b - D u* (boundary_condition.f90:1609-1629): one wave per segment,
// lane = box cell; u* of the cell is pulled from the previous step's
// populations with the lattice kernel's arithmetic.
template <typename S>
__device__ __forceinline__ void ibm_rhs_body(const IbmParams& p, const MrtParams& P, const GeoBuffers& G, int slot,
                                             const FlowBuffers& F, const Grid& g, const S* __restrict__ src, int blk) {
  const int gk = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (gk >= p.nseg) return;
  const size_t e = (size_t)slot * p.nseg + gk;
  const int* bx = G.box + e * 4;
  double su = 0.0, sv = 0.0;
  if (lane < BOXC) {
    const int ci = lane / BOXW, cj = lane - ci * BOXW;
    const int i = bx[0] + ci, j = bx[2] + cj;
    if (i <= bx[1] && j <= bx[3]) {
      double f[9];
      pull_node_s(src, g, i, j, P.w, f);
      double rho, u, v;
      node_moments(f, rho, u, v);
      const double phi = G.phx[e * BOXW + ci] * G.phy[e * BOXW + cj];
      su = phi * u;
      sv = phi * v;
    }
  }
  su = wave_sum(su);
  sv = wave_sum(sv);
  if (lane == 0) {
    const int w = gk / p.SPW, k = gk - w * p.SPW;
    const int gp = w * p.SPW + pos_of_seg(k, p.SPW);
    F.cvec[2 * gp] = G.ub[e] - su;
    F.cvec[2 * gp + 1] = G.vb[e] - sv;
  }
}

// y = S^-1 c, one wave per row of the inverse.
__device__ __forceinline__ void ibm_matvec_body(const IbmParams& p, const GeoBuffers& G, int slot,
                                                const FlowBuffers& F, int blk) {
  const int gp = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (gp >= p.nseg) return;
  const int n = p.SPW, w = gp / n, r = gp - w * n;
  const double* row = G.sinv + (((size_t)slot * p.W + w) * n + r) * (size_t)n;
  const double* c = F.cvec + (size_t)w * n * 2;
  double a0 = 0.0, a1 = 0.0;
#pragma unroll 8
  for (int q = lane; q < n; q += 64) {
    const double s = row[q];
    a0 += s * c[2 * q];
    a1 += s * c[2 * q + 1];
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  if (lane == 0) {
    F.yvec[2 * gp] = a0;
    F.yvec[2 * gp + 1] = a1;
  }
}

// g = D^T y at the adjoint nodes (boundary_condition.f90:1639-1651), per
// block of SPREAD_SEG consecutive segments of a wing: every segment whose box
// can hold a node claimed by the block lies within +-bw band positions of
// the claiming segment, and those positions (boxes, 1-D delta weights, y)
// are staged in LDS once for the block.
#ifndef LBM_SPREAD_SEG
#define LBM_SPREAD_SEG 16
#endif
constexpr int SPREAD_SEG = LBM_SPREAD_SEG;
constexpr int SPREAD_STAGE = 2 * SPREAD_SEG + 2 * BMAX + 2;

struct SpreadStage {
  int box[SPREAD_STAGE][4];
  double ph[SPREAD_STAGE][2 * BOXW];
  double y[SPREAD_STAGE][2];
};

struct SpreadBlock {
  int w, k0, k1, plo, bw;
};

// Stage block sb of wing w (all NT threads; the caller synchronises).
template <int NT>
__device__ __forceinline__ SpreadBlock spread_stage(const IbmParams& p, const GeoBuffers& G, int slot,
                                                    const FlowBuffers& F, int w, int sb, SpreadStage& st) {
  const int n = p.SPW, t = threadIdx.x;
  SpreadBlock B;
  B.w = w;
  B.k0 = sb * SPREAD_SEG;
  B.k1 = min(B.k0 + SPREAD_SEG, n);
  const GeoStep* gs = G.step + slot;
  B.bw = min(gs->bw[w], BMAX);
  int plo = INT_MAX, phi = -1;
  for (int k = B.k0; k < B.k1; k++) {
    const int q = pos_of_seg(k, n);
    plo = min(plo, q);
    phi = max(phi, q);
  }
  plo = max(0, plo - B.bw);
  phi = min(n - 1, phi + B.bw);
  B.plo = plo;
  const int ns = min(phi - plo + 1, SPREAD_STAGE);
  const size_t base = (size_t)slot * p.nseg + (size_t)w * n;
  for (int e = t; e < ns; e += NT) {
    const int q = plo + e, k = seg_of_pos(q, n);
    const int* bx = G.box + (base + k) * 4;
#pragma unroll
    for (int c = 0; c < 4; c++) st.box[e][c] = bx[c];
#pragma unroll
    for (int c = 0; c < BOXW; c++) {
      st.ph[e][c] = G.phx[(base + k) * BOXW + c];
      st.ph[e][BOXW + c] = G.phy[(base + k) * BOXW + c];
    }
    st.y[e][0] = F.yvec[2 * ((size_t)w * n + q)];
    st.y[e][1] = F.yvec[2 * ((size_t)w * n + q) + 1];
  }
  return B;
}

// g at adjoint node (i, j) claimed by wing-local segment kown.
__device__ __forceinline__ void spread_g(const SpreadBlock& B, const SpreadStage& st, int n, int i, int j, int kown,
                                         double& gx, double& gy) {
  const int p0 = pos_of_seg(kown, n);
  const int qa = max(0, p0 - B.bw), qb = min(n - 1, p0 + B.bw);
  gx = 0.0;
  gy = 0.0;
  for (int q = qa; q <= qb; q++) {
    const int e = q - B.plo;
    if (i < st.box[e][0] || i > st.box[e][1] || j < st.box[e][2] || j > st.box[e][3]) continue;
    const double phi_ = st.ph[e][i - st.box[e][0]] * st.ph[e][BOXW + j - st.box[e][2]];
    gx = gx + phi_ * st.y[e][0];
    gy = gy + phi_ * st.y[e][1];
  }
}

// Blocks of SPREAD_SEG segments per wing (one for an instance without wings,
// whose block only writes the force row).
__host__ __device__ __forceinline__ int adj_blocks_per_wing(const IbmParams& p) {
  return p.W > 0 ? (p.SPW + SPREAD_SEG - 1) / SPREAD_SEG : 1;
}



}  // namespace lbm
