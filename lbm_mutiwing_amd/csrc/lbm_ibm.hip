// lbm_ibm.hip - implicit velocity-correction immersed boundary on gfx950.
//
// The reference (immersed_boundary_Wu, boundary_condition.f90:1512-1692)
// builds dense nseg x nadj delta matrices D and E = D^T diag(arc), forms
// A = D E with DGEMM and solves A x = U_b - D u* by LU (cuSOLVER) every
// step.  Here the same system is solved without any dense matrix:
//   * A = S diag(arc) with S = D D^T symmetric positive definite, and the
//     delta kernel is separable, so S(k,l) = Sx(k,l) * Sy(k,l) with 1-D sums
//     over the overlap of the two 6-node support boxes;
//   * x = diag(arc)^-1 S^-1 c, hence the spread g = E x = D^T (S^-1 c) never
//     needs arc;
//   * each wing's segments are ordered zig-zag from the first marker
//     (1, n, 2, n-1, ...), which puts upper/lower-surface neighbours next to
//     each other: S becomes a narrow band (half-width about 12) and is
//     factored by a banded Cholesky held in LDS, one workgroup per wing.
// The adjoint node list keeps the reference's order and ownership (segment,
// then i, then j; first claiming segment wins, airfoil.f90:286-335), built in
// parallel with an owner map and a prefix sum.
#include <algorithm>
#include <climits>

#include "lbm_internal.h"

namespace lbm {

__device__ __forceinline__ void raise_error(DevStep* st, int code, int info) {
  if (atomicCAS(&st->error, 0, code) == 0) st->error_info = info;
}

__device__ __forceinline__ int load_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// --------------------------------------------------------------- motion

// smooth_r / smooth_h / smooth_d and interpolate_motion_triplet (mode.f90:220-285).
__device__ double smooth_r(double x) {
  const double pai = 3.1415926535897932384626433832795;
  if (x < 0.0) return 0.0;
  if (x <= 1.0) return 0.5 * x * x + (cos(2.0 * pai * x) - 1.0) / (4.0 * pai * pai);
  return x - 0.5;
}
__device__ double smooth_h(double x) {
  const double pai = 3.1415926535897932384626433832795;
  if (x < 0.0) return 0.0;
  if (x <= 1.0) return x - sin(2.0 * pai * x) / (2.0 * pai);
  return 1.0;
}
__device__ double smooth_d(double x) {
  const double pai = 3.1415926535897932384626433832795;
  if (x >= 0.0 && x <= 1.0) return 1.0 - cos(2.0 * pai * x);
  return 0.0;
}
__device__ void interp_triplet(double s0, double v0, double v1, double dt, double lt, double* s, double* v,
                               double* a) {
  const double dv = v1 - v0;
  if (dt <= 1.0e-14) {
    *s = (v0 + dv) * lt + s0;
    *v = v0 + dv;
    *a = 0.0;
    return;
  }
  const double x = lt / dt;
  *s = dv * dt * smooth_r(x) + v0 * lt + s0;
  *v = dv * smooth_h(x) + v0;
  *a = dv / dt * smooth_d(x);
}

// mode (mode.f90:451-471) for one wing, apply_motion_row (mode.f90:180-218).
__device__ WingMotionState wing_motion(const IbmParams& p, const IbmBuffers& B, int w, double time, bool* ok) {
  double r[9];
  const int col = 1 + w * 9;
  *ok = true;
  if (p.motion_source == LBM_MOTION_STREAM) {
    // lookup_stream_motion_row (mode.f90:414-449)
    const double* S = B.stream_rows;
    const double* T = B.stream_rows + LBM_MAX_MOTION_COLS;
    const double t1 = S[0], t2 = T[0], dt = t2 - t1;
    if (time <= t1 + 1.0e-12) {
      for (int c = 0; c < 9; c++) r[c] = S[col + c];
    } else {
      double lt = time - t1;
      if (lt < 0.0) lt = 0.0;
      const double dmax = dt > 0.0 ? dt : 0.0;
      if (lt > dmax) lt = dmax;
      interp_triplet(S[col], S[col + 1], T[col + 1], dt, lt, &r[0], &r[1], &r[2]);
      interp_triplet(S[col + 3], S[col + 4], T[col + 4], dt, lt, &r[3], &r[4], &r[5]);
      interp_triplet(S[col + 6], S[col + 7], T[col + 7], dt, lt, &r[6], &r[7], &r[8]);
    }
  } else {
    if (p.step_count < 1 || B.step_table == nullptr) {
      *ok = false;
      for (int c = 0; c < 9; c++) r[c] = 0.0;
    } else {
      long long idx = llround(time / p.motion_dt) + 1;
      if (idx > p.step_count) idx = p.step_count;
      if (idx < 1) idx = 1;
      const double* row = B.step_table + (idx - 1) * (long long)p.ncols;
      for (int c = 0; c < 9; c++) r[c] = row[col + c];
      r[3] = r[3] - (p.wing_x0 + (double)w * p.wing_dx);
      r[6] = r[6] - p.wing_y0;
    }
  }
  WingMotionState m;
  m.angle = r[0]; m.omega = r[1]; m.omgac = r[2];
  m.sa = r[3]; m.sau = r[4]; m.saac = r[5];
  m.sb = r[6]; m.sbv = r[7]; m.sbac = r[8];
  return m;
}

// Motion arrays, marker transform and midpoints (move_boundary,
// airfoil.f90:254-272), rigid-body velocity at the midpoints
// (boundary_condition.f90:1540-1555), support boxes (airfoil.f90:288-321) and
// the 1-D delta weights of every box column / row (boundary_condition.f90:
// 1558-1600).  A workgroup takes SEG_PER_BLOCK segments: one thread per
// segment transforms the marker that starts it (and the closing marker of its
// wing) and scans its box, then one thread per (segment, box column or row)
// evaluates a delta weight.
constexpr int SEG_PER_BLOCK = 21;   // 21 segments x 12 weights = 252 threads

__global__ __launch_bounds__(256) void k_markers(IbmParams p, IbmBuffers B, int n) {
  __shared__ double sX[SEG_PER_BLOCK], sY[SEG_PER_BLOCK];
  __shared__ int sb[SEG_PER_BLOCK][4];
  DevStep* st = B.st;
  const int t = threadIdx.x;
  const int k0 = blockIdx.x * SEG_PER_BLOCK;
  const double time = (double)(n - 1) * p.dt;
  const double scale = p.scale, UU = p.uu, ox = p.ox, oy = p.oy;
  if (t < SEG_PER_BLOCK && k0 + t < p.nseg) {
    const int k = k0 + t;
    const int w = k / p.SPW, local = k - w * p.SPW;
    bool ok = true;
    const WingMotionState m = wing_motion(p, B, w, time, &ok);
    if (!ok) raise_error(st, DERR_MOTION, w);
    if (local == 0) {
      st->wing[w] = m;
      if (w == 0) {
        st->time = time;
        st->n = n;
      }
    }
    const double ca = cos(m.angle), sn = sin(m.angle);
    const int pa = w * p.P + local;
    double xe[2], ye[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const double x0 = B.xb0[pa + e], y0 = B.yb0[pa + e];
      xe[e] = ca * x0 + sn * y0 + p.wing_x0 + (double)w * p.wing_dx + m.sa + p.xc +
              p.body_extra_s[w] / (1.0 / scale) - (double)p.movenum * scale;
      ye[e] = -sn * x0 + ca * y0 + p.wing_y0 + m.sb + p.yc;
    }
    B.xb1[pa] = xe[0];
    B.yb1[pa] = ye[0];
    if (local == p.SPW - 1) {
      B.xb1[pa + 1] = xe[1];
      B.yb1[pa + 1] = ye[1];
    }
    const double X = 0.5 * (xe[1] + xe[0]);
    const double Y = 0.5 * (ye[1] + ye[0]);
    B.xb2[k] = X;
    B.yb2[k] = Y;
    const double xs = B.xb0[pa + 1] + B.xb0[pa], ys = B.yb0[pa + 1] + B.yb0[pa];
    const double om = m.omega, an = m.angle;
    B.ub[k] = UU * (-om * sin(an) * 0.5 * xs + om * cos(an) * 0.5 * ys + m.sau) + p.body_extra_v[w];
    B.vb[k] = UU * (-om * cos(an) * 0.5 * xs - om * sin(an) * 0.5 * ys + m.sbv);
    // support box, 1-based scan exactly as the reference
    int i1 = (int)((X - ox) / scale + 1.0), i2 = i1;
    int j1 = (int)((Y - oy) / scale + 1.0), j2 = j1;
    for (int i = i1; i >= 1; i--)
      if (fabs((ox + (double)(i - 1) * scale - X) / scale) > 2.5) { i1 = i + 1; break; }
    for (int i = i2; i <= p.nx; i++)
      if (fabs((ox + (double)(i - 1) * scale - X) / scale) > 2.5) { i2 = i - 1; break; }
    for (int i = j2; i <= p.ny; i++)
      if (fabs((oy + (double)(i - 1) * scale - Y) / scale) > 2.5) { j2 = i - 1; break; }
    for (int i = j1; i >= 1; i--)
      if (fabs((oy + (double)(i - 1) * scale - Y) / scale) > 2.5) { j1 = i + 1; break; }
    if (i1 < 4 || j1 < 4 || i2 > p.nx - 3 || j2 > p.ny - 3 || i2 - i1 >= BOXW || j2 - j1 >= BOXW) {
      raise_error(st, DERR_BOX_EDGE, k);
      i1 = 4; i2 = 3; j1 = 4; j2 = 3;  // empty box
    }
    B.box[4 * k + 0] = i1 - 1;
    B.box[4 * k + 1] = i2 - 1;
    B.box[4 * k + 2] = j1 - 1;
    B.box[4 * k + 3] = j2 - 1;
    sX[t] = X;
    sY[t] = Y;
    sb[t][0] = i1; sb[t][1] = i2; sb[t][2] = j1; sb[t][3] = j2;
  }
  __syncthreads();
  const int ls = t / (2 * BOXW), c = t - ls * 2 * BOXW;
  if (ls < SEG_PER_BLOCK && k0 + ls < p.nseg) {
    const int k = k0 + ls;
    if (c < BOXW) {
      const int i = sb[ls][0] + c;
      const double x0 = ox + scale * (double)(i - 1);
      B.phx[BOXW * k + c] = i <= sb[ls][1] ? delta_kernel(fabs((x0 - sX[ls]) / scale)) : 0.0;
    } else {
      const int j = sb[ls][2] + (c - BOXW);
      const double y0 = oy + scale * (double)(j - 1);
      B.phy[BOXW * k + (c - BOXW)] = j <= sb[ls][3] ? delta_kernel(fabs((y0 - sY[ls]) / scale)) : 0.0;
    }
  }
}

// With no wings only the step time is needed by the force row.
__global__ void k_step_time(IbmParams p, IbmBuffers B, int n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    B.st->time = (double)(n - 1) * p.dt;
    B.st->n = n;
  }
}

void launch_motion_markers(hipStream_t s, const IbmParams& p, const IbmBuffers& B, int n) {
  if (p.nseg > 0)
    hipLaunchKernelGGL(k_markers, dim3((p.nseg + SEG_PER_BLOCK - 1) / SEG_PER_BLOCK), dim3(256), 0, s, p, B, n);
  else
    hipLaunchKernelGGL(k_step_time, dim3(1), dim3(64), 0, s, p, B, n);
}

// ---------------------------------------------------------- adjoint list
//
// The adjoint node list in the reference's order (airfoil.f90:275-335), built
// by five grid-wide passes: release the previous list, mark every node with
// its first (lowest) claiming segment, count the claims per segment, prefix
// sum, and write each segment's claims in (i, j) order.  Every node also
// records the lowest and highest band key (wing, zig-zag position) of the
// segments whose support contains it: two segments are coupled in S exactly
// when they share a node, so the band half-width is the largest key spread
// of any node, and a node whose keys belong to two wings couples the wings.

__global__ __launch_bounds__(256) void k_adj_release(IbmParams p, IbmBuffers B) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  DevStep* st = B.st;
  if (a == 0) st->bw = 0;
  if (a >= st->adj_n) return;
  const long long node = (long long)B.adj_ij[2 * a + 1] * p.nx + B.adj_ij[2 * a];
  B.owner[node] = INT_MAX;
  B.aidx[node] = -1;
  B.kmin[node] = INT_MAX;
  B.kmax[node] = INT_MIN;
}

__global__ __launch_bounds__(256) void k_adj_claim(IbmParams p, IbmBuffers B) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (B.st->error != 0 || q >= p.nseg * BOXW * BOXW) return;
  const int k = q / (BOXW * BOXW), c = q - k * BOXW * BOXW;
  const int ci = c / BOXW, cj = c - ci * BOXW;
  const int* bx = B.box + 4 * k;
  const int i = bx[0] + ci, j = bx[2] + cj;
  if (i > bx[1] || j > bx[3]) return;
  const long long node = (long long)j * p.nx + i;
  const int key = (k / p.SPW) * 65536 + B.pos_of_seg[k];
  atomicMin(&B.owner[node], k);
  atomicMin(&B.kmin[node], key);
  atomicMax(&B.kmax[node], key);
}

// One wave per segment: lane c < 36 is box cell (c/6, c%6) in the
// reference's (i, j) claim order; a ballot gives the claim count and each
// claim's rank.
__device__ __forceinline__ bool cell_claimed(const IbmParams& p, const IbmBuffers& B, int k, int lane, int* ni,
                                             int* nj) {
  const int* bx = B.box + 4 * k;
  const int i = bx[0] + lane / BOXW, j = bx[2] + lane % BOXW;
  *ni = i;
  *nj = j;
  return lane < BOXW * BOXW && i <= bx[1] && j <= bx[3] && B.owner[(long long)j * p.nx + i] == k;
}

__global__ __launch_bounds__(256) void k_adj_count(IbmParams p, IbmBuffers B) {
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= p.nseg) return;
  int i, j;
  const bool c = B.st->error == 0 && cell_claimed(p, B, k, lane, &i, &j);
  const unsigned long long m = __ballot(c);
  if (lane == 0) B.seg_cnt[k] = __popcll(m);
}

constexpr int SCAN_THREADS = 1024;

__global__ __launch_bounds__(SCAN_THREADS) void k_adj_scan(IbmParams p, IbmBuffers B) {
  __shared__ int s_part[SCAN_THREADS];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int chunk = (p.nseg + nt - 1) / nt;
  const int c0 = tid * chunk, c1 = min(c0 + chunk, p.nseg);
  int local = 0;
  for (int k = c0; k < c1; k++) local += B.seg_cnt[k];
  s_part[tid] = local;
  __syncthreads();
  for (int d = 1; d < nt; d <<= 1) {
    const int v = tid >= d ? s_part[tid - d] : 0;
    __syncthreads();
    s_part[tid] += v;
    __syncthreads();
  }
  int run = s_part[tid] - local;
  for (int k = c0; k < c1; k++) {
    B.seg_off[k] = run;
    run += B.seg_cnt[k];
  }
  if (tid == nt - 1) {
    const int total = s_part[nt - 1];
    B.seg_off[p.nseg] = total;
    DevStep* st = B.st;
    if (total > p.cap) raise_error(st, DERR_ADJ_OVERFLOW, total);
    st->adj_n = (st->error != 0) ? 0 : total;
  }
}

__global__ __launch_bounds__(256) void k_adj_write(IbmParams p, IbmBuffers B) {
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  DevStep* st = B.st;
  if (st->error != 0 || k >= p.nseg) return;
  int i, j;
  const bool c = cell_claimed(p, B, k, lane, &i, &j);
  const unsigned long long m = __ballot(c);
  int bw = 0;
  if (c) {
    const int a = B.seg_off[k] + __popcll(m & ((1ull << lane) - 1ull));
    const long long node = (long long)j * p.nx + i;
    B.adj_ij[2 * a] = i;
    B.adj_ij[2 * a + 1] = j;
    B.adj_seg[a] = k;
    B.aidx[node] = a;
    const int lo = B.kmin[node], hi = B.kmax[node];
    if ((lo >> 16) != (hi >> 16)) raise_error(st, DERR_WINGS_COUPLED, k);
    bw = (hi & 0xffff) - (lo & 0xffff);
  }
  for (int off = 32; off > 0; off >>= 1) bw = max(bw, __shfl_xor(bw, off));
  if (lane == 0) {
    if (bw > 0) atomicMax(&st->bw, bw);
    if (bw > BMAX) raise_error(st, DERR_BAND, bw);
  }
}

void launch_adjoint_build(hipStream_t s, const IbmParams& p, const IbmBuffers& B) {
  const int nseg = p.nseg;
  hipLaunchKernelGGL(k_adj_release, dim3((p.cap + 255) / 256), dim3(256), 0, s, p, B);
  hipLaunchKernelGGL(k_adj_claim, dim3((nseg * BOXW * BOXW + 255) / 256), dim3(256), 0, s, p, B);
  hipLaunchKernelGGL(k_adj_count, dim3((nseg + 3) / 4), dim3(256), 0, s, p, B);
  hipLaunchKernelGGL(k_adj_scan, dim3(1), dim3(SCAN_THREADS), 0, s, p, B);
  hipLaunchKernelGGL(k_adj_write, dim3((nseg + 3) / 4), dim3(256), 0, s, p, B);
}

// Episode reset: release the lattice nodes of the current adjoint list
// (move_boundary's body reset, airfoil.f90:275-284).
__global__ void k_clear_adjoint(IbmParams p, IbmBuffers B) {
  const int n = B.st->adj_n;
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
    const long long node = (long long)B.adj_ij[2 * a + 1] * p.nx + B.adj_ij[2 * a];
    B.owner[node] = INT_MAX;
    B.aidx[node] = -1;
    B.kmin[node] = INT_MAX;
    B.kmax[node] = INT_MIN;
  }
}

void launch_clear_adjoint(hipStream_t s, const IbmParams& p, const IbmBuffers& B) {
  hipLaunchKernelGGL(k_clear_adjoint, dim3(64), dim3(256), 0, s, p, B);
}

// ------------------------------------------------------- gather / assemble

// u* and rho at the adjoint nodes (boundary_condition.f90:1609-1616): the
// same pull and moment arithmetic as the lattice kernel.
__global__ __launch_bounds__(256) void k_ibm_gather(IbmBuffers B, Grid g, const double* __restrict__ src) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (B.st->error != 0 || a >= B.st->adj_n) return;
  double f[9];
  pull_node(src, g, B.adj_ij[2 * a], B.adj_ij[2 * a + 1], f);
  double rho, u, v;
  node_moments(f, rho, u, v);
  B.adj_rho[a] = rho;
  B.adj_us[a] = u;
  B.adj_vs[a] = v;
}

void launch_ibm_gather(hipStream_t s, const IbmParams& p, const IbmBuffers& B, const Grid& g, const double* src) {
  if (p.cap <= 0) return;
  hipLaunchKernelGGL(k_ibm_gather, dim3((p.cap + 255) / 256), dim3(256), 0, s, B, g, src);
}

// ------------------------------------------------------------ wing solve

constexpr int NB = 16;                 // block size of the block-tridiagonal form (> band half-width)
constexpr int NB2 = NB * NB;
constexpr int BCR_MAXBLK = 32;         // up to 512 segments per wing in LDS
constexpr int WING_WAVES = 8;
constexpr int WING_THREADS = 64 * WING_WAVES;
constexpr int BCR_WS = 2 * NB2 + 2 * NB;   // per block: X_L, X_U (16x16), x_r (16x2) in global memory

constexpr size_t BCR_WING = (size_t)2 * BCR_MAXBLK * NB2 + (size_t)BCR_MAXBLK * NB * 2;  // assembled D, L, R

// assembled systems (BCR_WING per wing) followed by the reduction workspace
size_t ibm_bcr_ws_doubles(int W) { return (size_t)W * (BCR_WING + (size_t)BCR_MAXBLK * BCR_WS); }

size_t ibm_solve_lds_bytes(int spw) {
  const size_t bcr = (size_t)(2 * BCR_MAXBLK * NB2 + BCR_MAXBLK * NB * 2) * sizeof(double);
  const size_t seq = (size_t)spw * (BMAX + 1 + 3) * sizeof(double);
  const size_t red = 3 * WING_THREADS * sizeof(double);
  return std::max(std::max(bcr, seq), red);
}

// Pointer into LDS (address space 3), so out-of-line solver routines use
// ds_read/ds_write instead of flat accesses.
typedef __attribute__((address_space(3))) double lds_d;
typedef __attribute__((address_space(1))) double glb_d;
#define LDS(p) ((lds_d*)(p))
#define GLB(p) ((glb_d*)(p))

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// In-place lower Cholesky of a 16x16 SPD block (row-major) by one wave.
__device__ __noinline__ void chol16(lds_d* A, int lane, DevStep* st, int info) {
  for (int k = 0; k < NB; k++) {
    if (lane == 0) {
      const double d = A[k * (NB + 1)];
      if (!(d > 0.0)) raise_error(st, DERR_NOT_SPD, info);
      A[k * (NB + 1)] = sqrt(d > 0.0 ? d : 1.0);
    }
    wave_sync();
    const double rdk = 1.0 / A[k * (NB + 1)];
    if (lane > k && lane < NB) A[lane * NB + k] *= rdk;
    wave_sync();
    for (int idx = lane; idx < NB2; idx += 64) {
      const int r = idx >> 4, c = idx & 15;
      if (c > k && r >= c) A[idx] -= A[r * NB + k] * A[c * NB + k];
    }
    wave_sync();
  }
}

// x <- (L L^T)^-1 x for one right-hand side held in registers (L in LDS).
__device__ __forceinline__ void chol16_solve(const lds_d* A, double x[NB]) {
#pragma unroll
  for (int k = 0; k < NB; k++) {
    x[k] = x[k] / A[k * (NB + 1)];
#pragma unroll
    for (int r = k + 1; r < NB; r++) x[r] -= A[r * NB + k] * x[k];
    if ((k & 3) == 3) __asm__ volatile("" ::: "memory");
  }
  // reload the factor for the backward pass instead of keeping all 136
  // entries live from the forward pass
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int k = NB - 1; k >= 0; k--) {
    x[k] = x[k] / A[k * (NB + 1)];
#pragma unroll
    for (int r = 0; r < k; r++) x[r] -= A[k * NB + r] * x[k];
    if ((k & 3) == 0) __asm__ volatile("" ::: "memory");
  }
}

// Block cyclic reduction of the SPD block-tridiagonal system
//   L_i x_{i-1} + D_i x_i + L_{i+1}^T x_{i+1} = r_i,  i = 0..nblk-1,
// with two right-hand sides.  Level l eliminates the blocks i = s, 3s, 5s,
// ... (s = 2^l): each is factored and solved against [L_i, L_{i+s}^T, r_i]
// (X_L, X_U, x_r kept in global memory), then every surviving neighbour j
// takes the Schur complement D_j -= L_j X_U[j-s] + L_{j+s}^T X_L[j+s] and
// couples to j-2s through L_j <- -L_j X_L[j-s].  Each level is a Schur
// complement of an SPD matrix, so the reduction is as stable as Cholesky.
// The solution overwrites R.
__device__ __noinline__ void bcr_solve(lds_d* D, lds_d* L, lds_d* R, glb_d* ws, int nblk, int wave, int lane,
                                       DevStep* st, int wing) {
  int top = 0;
  while ((1 << top) < nblk) top++;
  for (int l = 0; l < top; l++) {
    const int s = 1 << l;
    for (int i = s + 2 * s * wave; i < nblk; i += 2 * s * WING_WAVES) {
      lds_d* Di = D + i * NB2;
      chol16(Di, lane, st, wing * 65536 + i);
      double x[NB];
      const bool has_u = i + s < nblk;
      const lds_d* Lu = L + (i + s) * NB2;
      const lds_d* Li = L + i * NB2;
#pragma unroll
      for (int r = 0; r < NB; r++) {
        double v = 0.0;
        if (lane < NB) v = Li[r * NB + lane];
        else if (lane < 2 * NB) v = has_u ? Lu[(lane - NB) * NB + r] : 0.0;
        else if (lane < 2 * NB + 2) v = R[i * 2 * NB + r * 2 + (lane - 2 * NB)];
        x[r] = v;
      }
      chol16_solve(Di, x);
      glb_d* wsi = ws + i * BCR_WS;
#pragma unroll
      for (int r = 0; r < NB; r++) {
        if (lane < NB) wsi[r * NB + lane] = x[r];                              // X_L
        else if (lane < 2 * NB) wsi[NB2 + r * NB + (lane - NB)] = x[r];        // X_U
        else if (lane < 2 * NB + 2) wsi[2 * NB2 + r * 2 + (lane - 2 * NB)] = x[r];  // x_r
      }
      wave_sync();
    }
    __threadfence_block();
    __syncthreads();
    for (int j = 2 * s * wave; j < nblk; j += 2 * s * WING_WAVES) {
      const int a = j - s, b = j + s;
      const bool ha = a >= 0, hb = b < nblk, hl = j - 2 * s >= 0;
      const lds_d* Lj = L + j * NB2;
      const lds_d* Lb = L + b * NB2;
      const glb_d* wa = ws + (ha ? a : 0) * BCR_WS;
      const glb_d* wb = ws + (hb ? b : 0) * BCR_WS;
      double dnew[4], lnew[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        __builtin_amdgcn_sched_barrier(0);
        const int idx = lane + 64 * q, r = idx >> 4, c = idx & 15;
        double dsum = D[j * NB2 + idx], lsum = 0.0;
        if (ha) {
          double t1 = 0.0, t2 = 0.0;
#pragma unroll 4
          for (int m = 0; m < NB; m++) {
            const double ljm = Lj[r * NB + m];
            t1 += ljm * wa[NB2 + m * NB + c];   // L_j X_U[a]
            t2 += ljm * wa[m * NB + c];         // L_j X_L[a]
          }
          dsum -= t1;
          lsum = hl ? -t2 : 0.0;
        }
        if (hb) {
          double t3 = 0.0;
#pragma unroll 4
          for (int m = 0; m < NB; m++) t3 += Lb[m * NB + r] * wb[m * NB + c];  // L_b^T X_L[b]
          dsum -= t3;
        }
        dnew[q] = dsum;
        lnew[q] = lsum;
      }
      double rnew = 0.0;
      if (lane < 2 * NB) {
        const int r = lane >> 1, qq = lane & 1;
        rnew = R[j * 2 * NB + lane];
        if (ha)
          for (int m = 0; m < NB; m++) rnew -= Lj[r * NB + m] * wa[2 * NB2 + m * 2 + qq];
        if (hb)
          for (int m = 0; m < NB; m++) rnew -= Lb[m * NB + r] * wb[2 * NB2 + m * 2 + qq];
      }
      wave_sync();
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int idx = lane + 64 * q;
        D[j * NB2 + idx] = dnew[q];
        if (ha) L[j * NB2 + idx] = lnew[q];
      }
      if (lane < 2 * NB) R[j * 2 * NB + lane] = rnew;
      wave_sync();
    }
    __syncthreads();
  }
  if (wave == 0) {
    chol16(D, lane, st, wing * 65536);
    double x[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) x[r] = lane < 2 ? R[r * 2 + lane] : 0.0;
    chol16_solve(D, x);
    if (lane < 2)
#pragma unroll
      for (int r = 0; r < NB; r++) R[r * 2 + lane] = x[r];
  }
  __threadfence_block();
  __syncthreads();
  for (int l = top - 1; l >= 0; l--) {
    const int s = 1 << l;
    for (int i = s + 2 * s * wave; i < nblk; i += 2 * s * WING_WAVES) {
      if (lane < 2 * NB) {
        const int r = lane >> 1, qq = lane & 1;
        const glb_d* wsi = ws + i * BCR_WS;
        double x = wsi[2 * NB2 + lane];
        for (int m = 0; m < NB; m++) x -= wsi[r * NB + m] * R[(i - s) * 2 * NB + m * 2 + qq];
        if (i + s < nblk)
          for (int m = 0; m < NB; m++) x -= wsi[NB2 + r * NB + m] * R[(i + s) * 2 * NB + m * 2 + qq];
        R[i * 2 * NB + lane] = x;
      }
    }
    __syncthreads();
  }
}

// Wider bands or more than 512 segments: LDL^T of the band in LDS by one wave.
__device__ __noinline__ void band_solve_lds(lds_d* band, lds_d* ru, lds_d* rv, lds_d* invd, int n, int bw, int lane,
                                            DevStep* st, int wing) {
  constexpr int Dd = BMAX + 1;
  for (int k = 0; k < n; k++) {
    const double d = band[k * Dd];
    if (lane == 0 && !(d > 0.0)) raise_error(st, DERR_NOT_SPD, wing * 65536 + k);
    const double rd = 1.0 / (d > 0.0 ? d : 1.0);
    const int i = lane + 1;
    double li = 0.0;
    if (i <= bw && k + i < n) li = band[(k + i) * Dd + i] * rd;
    wave_sync();
    if (i <= bw && k + i < n)
      for (int j = 1; j <= i; j++) band[(k + i) * Dd + (i - j)] -= li * band[(k + j) * Dd + j];
    wave_sync();
    if (i <= bw && k + i < n) band[(k + i) * Dd + i] = li;
    if (lane == 0) invd[k] = rd;
    wave_sync();
  }
  const int side = lane >> 5, dd = (lane & 31) + 1;
  lds_d* rr = side == 0 ? ru : rv;
  for (int q = 0; q < n; q++) {
    const double y = rr[q];
    wave_sync();
    if (dd <= bw && q + dd < n) rr[q + dd] -= band[(q + dd) * Dd + dd] * y;
    wave_sync();
  }
  for (int q = lane; q < 2 * n; q += 64) (q < n ? ru : rv)[q % n] *= invd[q % n];
  wave_sync();
  for (int q = n - 1; q >= 0; q--) {
    const double x = rr[q];
    wave_sync();
    if (dd <= bw && q - dd >= 0) rr[q - dd] -= band[q * Dd + dd] * x;
    wave_sync();
  }
}

__device__ __forceinline__ double s_entry(const IbmBuffers& B, int k, int l) {
  const int* bk = B.box + 4 * k;
  const int* bl = B.box + 4 * l;
  const int ia = max(bk[0], bl[0]), ib = min(bk[1], bl[1]);
  const int ja = max(bk[2], bl[2]), jb = min(bk[3], bl[3]);
  if (ia > ib || ja > jb) return 0.0;
  const double* pxk = B.phx + BOXW * k;
  const double* pyk = B.phy + BOXW * k;
  const double* pxl = B.phx + BOXW * l;
  const double* pyl = B.phy + BOXW * l;
  double sx = 0.0, sy = 0.0;
  for (int i = ia; i <= ib; i++) sx += pxk[i - bk[0]] * pxl[i - bl[0]];
  for (int j = ja; j <= jb; j++) sy += pyk[j - bk[2]] * pyl[j - bl[2]];
  return sx * sy;
}

// Block-tridiagonal form of S (or, for wide bands, the plain band) and the
// right-hand side U_b - D u* (boundary_condition.f90:1558-1629), one thread
// per matrix entry or segment, written to global memory for k_ibm_wing.
// Block layout per wing: D_i, L_i (16x16 each, row-major) for i < nblk, then
// R (nblk*16 x 2).  Padding rows beyond the wing's segments get D = I.
__global__ __launch_bounds__(256) void k_ibm_assemble(IbmParams p, IbmBuffers B) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  DevStep* st = B.st;
  if (st->error != 0) return;
  const int n = p.SPW, bw = st->bw;
  const int nblk = (n + NB - 1) / NB;
  const bool bcr = bw < NB && n <= BCR_MAXBLK * NB;
  if (t < (long long)p.nseg) {  // right-hand side of segment at band position
    const int gp = (int)t, w = gp / n, pos = gp - w * n;
    const int k = B.seg_of_pos[gp];
    const int* bk = B.box + 4 * k;
    const double* pxk = B.phx + BOXW * k;
    const double* pyk = B.phy + BOXW * k;
    int ai[BOXW * BOXW];
#pragma unroll
    for (int c = 0; c < BOXW * BOXW; c++) {
      const int i = bk[0] + c / BOXW, j = bk[2] + c % BOXW;
      ai[c] = (i <= bk[1] && j <= bk[3]) ? B.aidx[(long long)j * p.nx + i] : -1;
    }
    double su = 0.0, sv = 0.0;
#pragma unroll
    for (int c = 0; c < BOXW * BOXW; c++) {
      if (ai[c] < 0) continue;
      const double phi = pxk[c / BOXW] * pyk[c % BOXW];
      su = su + phi * B.adj_us[ai[c]];
      sv = sv + phi * B.adj_vs[ai[c]];
    }
    if (bcr) {
      double* R = B.bcr_ws + (size_t)w * BCR_WING + 2 * BCR_MAXBLK * NB2;
      R[pos * 2] = B.ub[k] - su;
      R[pos * 2 + 1] = B.vb[k] - sv;
    } else {
      B.rhs[gp] = B.ub[k] - su;
      B.rhs[p.nseg + gp] = B.vb[k] - sv;
    }
    return;
  }
  const long long u = t - p.nseg;
  if (bcr) {
    const long long per = (long long)nblk * 2 * NB2 + 2 * NB;  // D,L entries + padding rhs rows
    const int w = (int)(u / per);
    if (w >= p.W) return;
    const int e = (int)(u - (long long)w * per);
    double* base = B.bcr_ws + (size_t)w * BCR_WING;
    if (e >= nblk * 2 * NB2) {  // padding rows of the right-hand side
      const int r = e - nblk * 2 * NB2;   // 0 .. 2*NB-1
      const int pos = n + r / 2;
      if (pos < nblk * NB) base[2 * BCR_MAXBLK * NB2 + pos * 2 + (r & 1)] = 0.0;
      return;
    }
    const int isL = e / (nblk * NB2), e2 = e - isL * nblk * NB2;
    const int blk = e2 / NB2, r = (e2 - blk * NB2) >> 4, c = (e2 - blk * NB2) & 15;
    const int pos = blk * NB + r;
    const int q = (isL ? blk - 1 : blk) * NB + c;
    double v = 0.0;
    if (pos >= n || q >= n || q < 0) {
      v = (!isL && r == c) ? 1.0 : 0.0;
      if (pos < n && q < n) v = 0.0;
    } else if (abs(pos - q) <= bw) {
      const int kk = B.seg_of_pos[w * n + max(pos, q)], ll = B.seg_of_pos[w * n + min(pos, q)];
      v = s_entry(B, kk, ll);
    }
    base[(isL ? BCR_MAXBLK * NB2 : 0) + blk * NB2 + (r << 4) + c] = v;
  } else {
    const int D = BMAX + 1;
    if (u >= (long long)p.nseg * D) return;
    const int gp = (int)(u / D), d = (int)(u - (long long)gp * D);
    const int w = gp / n, pos = gp - w * n;
    double v = 0.0;
    if (d <= bw && pos - d >= 0) v = s_entry(B, B.seg_of_pos[gp], B.seg_of_pos[gp - d]);
    B.band[(long long)gp * D + d] = v;
  }
}

void launch_ibm_assemble(hipStream_t s, const IbmParams& p, const IbmBuffers& B) {
  if (p.W <= 0) return;
  const int nblk = (p.SPW + NB - 1) / NB;
  const long long bcr = (long long)p.W * ((long long)nblk * 2 * NB2 + 2 * NB);
  const long long band = (long long)p.nseg * (BMAX + 1);
  const long long total = p.nseg + std::max(bcr, band);
  hipLaunchKernelGGL(k_ibm_assemble, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, p, B);
}

// One workgroup per wing: load the assembled system into LDS, solve S y = c
// (block cyclic reduction, or the band LDL^T for wide bands), then spread
// g = D^T y to the wing's adjoint nodes (boundary_condition.f90:1639-1659)
// and reduce the wing's force and moment in a fixed order
// (boundary_condition.f90:1662-1675).  The segment boxes and delta weights
// used by the spread are staged in LDS once the solver storage is free.
__global__ __launch_bounds__(WING_THREADS) void k_ibm_wing(IbmParams p, IbmBuffers B) {
  extern __shared__ double lds[];
  DevStep* st = B.st;
  const int w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = p.SPW;
  if (st->error != 0) return;
  const int bw = st->bw;
  const int base = w * n;
  const bool bcr = bw < NB && n <= BCR_MAXBLK * NB;
  const int nblk = (n + NB - 1) / NB;
  double* D = lds;
  double* L = D + BCR_MAXBLK * NB2;
  double* R = L + BCR_MAXBLK * NB2;
  constexpr int Dd = BMAX + 1;
  double* band = lds;
  double* ru = lds + n * Dd;
  double* rv = ru + n;
  double* invd = rv + n;
  double* solu;  // solution by band position, stride 2 (BCR) or 1
  int sstride;
  if (bcr) {
    const double* g = B.bcr_ws + (size_t)w * BCR_WING;
    for (int t = tid; t < nblk * NB2; t += WING_THREADS) {
      D[t] = g[t];
      L[t] = g[BCR_MAXBLK * NB2 + t];
    }
    for (int t = tid; t < nblk * NB * 2; t += WING_THREADS) R[t] = g[2 * BCR_MAXBLK * NB2 + t];
    __syncthreads();
    bcr_solve(LDS(D), LDS(L), LDS(R), GLB(B.bcr_ws + (size_t)p.W * BCR_WING + (size_t)w * BCR_MAXBLK * BCR_WS),
              nblk, wave, lane, st, w);
    solu = R;
    sstride = 2;
  } else {
    for (int t = tid; t < n * Dd; t += WING_THREADS) band[t] = B.band[(long long)base * Dd + t];
    for (int t = tid; t < n; t += WING_THREADS) {
      ru[t] = B.rhs[base + t];
      rv[t] = B.rhs[p.nseg + base + t];
    }
    __syncthreads();
    if (wave == 0) band_solve_lds(LDS(band), LDS(ru), LDS(rv), LDS(invd), n, bw, lane, st, w);
    solu = ru;
    sstride = 1;
  }
  __syncthreads();
  // stage box and delta weights of the wing's segments by band position in
  // the free D/L area (BCR layout); the wide-band path reads them in place
  double* sph = D;                                         // n * 12
  int* sbox = reinterpret_cast<int*>(sph + n * 2 * BOXW);  // n * 4
  if (bcr) {
    for (int t = tid; t < n; t += WING_THREADS) {
      const int k = B.seg_of_pos[base + t];
      double v[2 * BOXW];
      int bx[4];
#pragma unroll
      for (int c = 0; c < BOXW; c++) {
        v[c] = B.phx[BOXW * k + c];
        v[BOXW + c] = B.phy[BOXW * k + c];
      }
#pragma unroll
      for (int c = 0; c < 4; c++) bx[c] = B.box[4 * k + c];
#pragma unroll
      for (int c = 0; c < 2 * BOXW; c++) sph[t * 2 * BOXW + c] = v[c];
#pragma unroll
      for (int c = 0; c < 4; c++) sbox[t * 4 + c] = bx[c];
    }
    __syncthreads();
  }
  const int a0 = B.seg_off[w * n], a1 = B.seg_off[(w + 1) * n];
  const WingMotionState m = st->wing[w];
  double fx = 0.0, fy = 0.0, fm = 0.0;
  const double* yv_base = bcr ? solu + 1 : rv;
  const int vstride = bcr ? 2 : 1;
  for (int a = a0 + tid; a < a1; a += WING_THREADS) {
    const int i = B.adj_ij[2 * a], j = B.adj_ij[2 * a + 1];
    const int p0 = B.pos_of_seg[B.adj_seg[a]];
    double gx = 0.0, gy = 0.0;
    const int qa = max(0, p0 - bw), qb = min(n - 1, p0 + bw);
    for (int q = qa; q <= qb; q++) {
      const int kq = bcr ? 0 : B.seg_of_pos[base + q];
      const int* bk = bcr ? sbox + 4 * q : B.box + 4 * kq;
      if (i < bk[0] || i > bk[1] || j < bk[2] || j > bk[3]) continue;
      const double* px = bcr ? sph + q * 2 * BOXW : B.phx + BOXW * kq;
      const double* py = bcr ? sph + q * 2 * BOXW + BOXW : B.phy + BOXW * kq;
      const double phi = px[i - bk[0]] * py[j - bk[2]];
      gx = gx + phi * solu[q * sstride];
      gy = gy + phi * yv_base[q * vstride];
    }
    const double rho = B.adj_rho[a];
    B.adj_u[a] = B.adj_us[a] + gx;
    B.adj_v[a] = B.adj_vs[a] + gy;
    const double t1 = 2.0 * rho * gx, t2 = 2.0 * rho * gy;
    B.adj_gx[a] = t1;
    B.adj_gy[a] = t2;
    const double x0 = (p.ox + p.scale * (double)i - (p.xc + p.wing_x0 + (double)w * p.wing_dx + m.sa)) / p.scale -
                      p.body_extra_pos[w];
    const double y0 = (p.oy + p.scale * (double)j - (p.yc + p.wing_y0 + m.sb)) / p.scale;
    fx = fx - t1;
    fy = fy - t2;
    fm = fm + (-x0 * t2 + y0 * t1);
  }
  __syncthreads();
  double* red = lds;  // the solver and staging storage is free now
  red[tid] = fx;
  red[WING_THREADS + tid] = fy;
  red[2 * WING_THREADS + tid] = fm;
  __syncthreads();
  for (int h = WING_THREADS / 2; h > 0; h >>= 1) {
    if (tid < h) {
      red[tid] += red[tid + h];
      red[WING_THREADS + tid] += red[WING_THREADS + tid + h];
      red[2 * WING_THREADS + tid] += red[2 * WING_THREADS + tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    st->raw[w][0] = red[0];
    st->raw[w][1] = red[WING_THREADS];
    st->raw[w][2] = red[2 * WING_THREADS];
  }
}

void launch_ibm_wing(hipStream_t s, const IbmParams& p, const IbmBuffers& B) {
  if (p.W <= 0) return;
  hipLaunchKernelGGL(k_ibm_wing, dim3(p.W), dim3(WING_THREADS), ibm_solve_lds_bytes(p.SPW), s, p, B);
}

// ---------------------------------------------------------------- forces

// Normalisation and force row (solution.f90:330-378), the airfoil_motion row
// (solution.f90:390-401) and the running row sum (solution.f90:379-385).
__global__ void k_forces(IbmParams p, IbmBuffers B, int slot, int row_len) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DevStep* st = B.st;
  const bool ok = st->error == 0;
  const double UU = p.uu, scale = p.scale;
  double wf[MAXW][2], wmom[MAXW];
  double F0 = 0.0, F1 = 0.0, Mt = 0.0;
  for (int w = 0; w < p.W; w++) {
    wf[w][0] = ok ? 2.0 * st->raw[w][0] / (UU * UU * (1.0 / scale)) : 0.0;
    wf[w][1] = ok ? 2.0 * st->raw[w][1] / (UU * UU * (1.0 / scale)) : 0.0;
    wmom[w] = ok ? -2.0 * st->raw[w][2] / (UU * UU * (1.0 / (scale * scale))) : 0.0;
    st->force[w][0] = wf[w][0];
    st->force[w][1] = wf[w][1];
    st->moment[w] = wmom[w];
  }
  for (int w = 0; w < p.W; w++) {
    F0 = F0 + wf[w][0];
    F1 = F1 + wf[w][1];
  }
  for (int w = 0; w < p.W; w++) Mt = Mt + wmom[w];
  double* row = B.rows + (long long)slot * row_len;
  int col = 0;
  row[col++] = st->time;
  for (int w = 0; w < p.W; w++) {
    row[col++] = wf[w][0];
    row[col++] = wf[w][1];
    row[col++] = wmom[w];
  }
  row[col++] = F0;
  row[col++] = F1;
  row[col++] = Mt;
  for (int w = 0; w < p.W; w++) {
    const WingMotionState& m = st->wing[w];
    row[col++] = p.wing_x0 + (double)w * p.wing_dx + m.sa + p.body_extra_s[w] / (1.0 / scale);
    row[col++] = m.sau;
    row[col++] = m.saac;
    row[col++] = p.wing_y0 + m.sb;
    row[col++] = m.sbv;
    row[col++] = m.sbac;
    row[col++] = m.angle;
    row[col++] = m.omega;
    row[col++] = m.omgac;
  }
  for (int c = 0; c < row_len; c++) B.row_sum[c] = B.row_sum[c] + row[c];
  double* mrow = B.mrows + (long long)slot * (1 + 3 * p.W);
  int c = 0;
  mrow[c++] = st->time;
  for (int w = 0; w < p.W; w++) {
    const WingMotionState& m = st->wing[w];
    mrow[c++] = 0.0;
    mrow[c++] = p.body_extra_v[w] / UU;
    mrow[c++] = p.wing_x0 + (double)w * p.wing_dx + m.sa + p.body_extra_s[w] / (1.0 / scale);
  }
}

void launch_forces(hipStream_t s, const IbmParams& p, const IbmBuffers& B, int slot, int row_len) {
  hipLaunchKernelGGL(k_forces, dim3(1), dim3(64), 0, s, p, B, slot, row_len);
}

}  // namespace lbm
