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
// the 1-D delta weights of every box column / row (boundary_condition.f90:1558-1600).
__global__ __launch_bounds__(256) void k_motion_markers(IbmParams p, IbmBuffers B, int n) {
  __shared__ WingMotionState wm[MAXW];
  DevStep* st = B.st;
  const double time = (double)(n - 1) * p.dt;
  const int tid = threadIdx.x;
  if (tid < MAXW) {
    bool ok = true;
    WingMotionState m = {};
    if (tid < p.W) m = wing_motion(p, B, tid, time, &ok);
    if (!ok) raise_error(st, DERR_MOTION, tid);
    wm[tid] = m;
    st->wing[tid] = m;
  }
  if (tid == 0) {
    st->time = time;
    st->n = n;
  }
  __syncthreads();
  const double scale = p.scale;
  for (int q = tid; q < p.npts; q += blockDim.x) {
    const int w = q / p.P;
    const double ca = cos(wm[w].angle), sn = sin(wm[w].angle);
    const double x0 = B.xb0[q], y0 = B.yb0[q];
    B.xb1[q] = ca * x0 + sn * y0 + p.wing_x0 + (double)w * p.wing_dx + wm[w].sa + p.xc +
               p.body_extra_s[w] / (1.0 / scale) - (double)p.movenum * scale;
    B.yb1[q] = -sn * x0 + ca * y0 + p.wing_y0 + wm[w].sb + p.yc;
  }
  __syncthreads();
  const double UU = p.uu, ox = p.ox, oy = p.oy;
  for (int k = tid; k < p.nseg; k += blockDim.x) {
    const int w = k / p.SPW;
    const int pa = w * p.P + (k - w * p.SPW);
    const double X = 0.5 * (B.xb1[pa + 1] + B.xb1[pa]);
    const double Y = 0.5 * (B.yb1[pa + 1] + B.yb1[pa]);
    B.xb2[k] = X;
    B.yb2[k] = Y;
    const double xs = B.xb0[pa + 1] + B.xb0[pa], ys = B.yb0[pa + 1] + B.yb0[pa];
    const double om = wm[w].omega, an = wm[w].angle;
    B.ub[k] = UU * (-om * sin(an) * 0.5 * xs + om * cos(an) * 0.5 * ys + wm[w].sau) + p.body_extra_v[w];
    B.vb[k] = UU * (-om * cos(an) * 0.5 * xs - om * sin(an) * 0.5 * ys + wm[w].sbv);
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
    for (int c = 0; c < BOXW; c++) {
      const int i = i1 + c, j = j1 + c;
      const double x0 = ox + scale * (double)(i - 1);
      const double y0 = oy + scale * (double)(j - 1);
      B.phx[BOXW * k + c] = i <= i2 ? delta_kernel(fabs((x0 - X) / scale)) : 0.0;
      B.phy[BOXW * k + c] = j <= j2 ? delta_kernel(fabs((y0 - Y) / scale)) : 0.0;
    }
  }
}

void launch_motion_markers(hipStream_t s, const IbmParams& p, const IbmBuffers& B, int n) {
  hipLaunchKernelGGL(k_motion_markers, dim3(1), dim3(256), 0, s, p, B, n);
}

// ---------------------------------------------------------- adjoint list

constexpr int BUILD_THREADS = 1024;

// Adjoint node list in the reference's order (airfoil.f90:275-335): clear the
// previous list, mark each node with its first (lowest) claiming segment,
// count and prefix-sum the claims per segment, then write every segment's
// claims in (i, j) order.  Also: band half-width of the zig-zag ordering and
// a check that no two wings share a support node.
__global__ __launch_bounds__(BUILD_THREADS) void k_adjoint_build(IbmParams p, IbmBuffers B) {
  __shared__ int s_part[BUILD_THREADS];
  __shared__ int s_bw, s_total;
  __shared__ int s_wbox[MAXW][4];
  DevStep* st = B.st;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nx = p.nx;
  const int prev_n = st->adj_n;
  // 0. clear the maps at the previous list (move_boundary resets body there)
  for (int a = tid; a < prev_n; a += nt) {
    const long long node = (long long)B.adj_ij[2 * a + 1] * nx + B.adj_ij[2 * a];
    B.owner[node] = INT_MAX;
    B.aidx[node] = -1;
  }
  if (tid == 0) {
    s_bw = 0;
    s_total = 0;
  }
  if (tid < MAXW) {
    s_wbox[tid][0] = INT_MAX; s_wbox[tid][1] = INT_MIN;
    s_wbox[tid][2] = INT_MAX; s_wbox[tid][3] = INT_MIN;
  }
  __threadfence();
  __syncthreads();
  const bool failed = st->error != 0;
  // 1. first claimant of every node
  if (!failed)
    for (int q = tid; q < p.nseg * BOXW * BOXW; q += nt) {
      const int k = q / (BOXW * BOXW), c = q - k * BOXW * BOXW;
      const int ci = c / BOXW, cj = c - ci * BOXW;
      const int* bx = B.box + 4 * k;
      const int i = bx[0] + ci, j = bx[2] + cj;
      if (i <= bx[1] && j <= bx[3]) atomicMin(&B.owner[(long long)j * nx + i], k);
    }
  __threadfence();
  __syncthreads();
  // 2. claims per segment
  for (int k = tid; k < p.nseg; k += nt) {
    const int* bx = B.box + 4 * k;
    int cnt = 0;
    if (!failed)
      for (int i = bx[0]; i <= bx[1]; i++)
        for (int j = bx[2]; j <= bx[3]; j++) cnt += load_agent(&B.owner[(long long)j * nx + i]) == k;
    B.seg_cnt[k] = cnt;
  }
  __threadfence();
  __syncthreads();
  // 3. exclusive prefix sum over segments (chunked + Hillis-Steele in LDS)
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
    B.seg_off[p.nseg] = s_part[nt - 1];
    s_total = s_part[nt - 1];
  }
  __threadfence();
  __syncthreads();
  const int total = s_total;
  const bool overflow = total > p.cap;
  if (overflow && tid == 0) raise_error(st, DERR_ADJ_OVERFLOW, total);
  // 4. write the claims in (segment, i, j) order
  if (!failed && !overflow)
    for (int k = tid; k < p.nseg; k += nt) {
      const int* bx = B.box + 4 * k;
      int a = B.seg_off[k];
      for (int i = bx[0]; i <= bx[1]; i++)
        for (int j = bx[2]; j <= bx[3]; j++) {
          const long long node = (long long)j * nx + i;
          if (load_agent(&B.owner[node]) == k) {
            B.adj_ij[2 * a] = i;
            B.adj_ij[2 * a + 1] = j;
            B.adj_seg[a] = k;
            B.aidx[node] = a;
            a++;
          }
        }
    }
  // 5. band half-width (zig-zag positions) and wing bounding boxes
  for (int k = tid; k < p.nseg; k += nt) {
    const int w = k / p.SPW;
    const int* bk = B.box + 4 * k;
    const int pk = B.pos_of_seg[k];
    int bw = 0;
    for (int l = w * p.SPW; l < (w + 1) * p.SPW; l++) {
      const int* bl = B.box + 4 * l;
      if (bk[1] < bl[0] || bl[1] < bk[0] || bk[3] < bl[2] || bl[3] < bk[2]) continue;
      bw = max(bw, abs(pk - B.pos_of_seg[l]));
    }
    atomicMax(&s_bw, bw);
    atomicMin(&s_wbox[w][0], bk[0]);
    atomicMax(&s_wbox[w][1], bk[1]);
    atomicMin(&s_wbox[w][2], bk[2]);
    atomicMax(&s_wbox[w][3], bk[3]);
  }
  __syncthreads();
  // wings whose bounding boxes overlap must not share a support node
  for (int k = tid; k < p.nseg; k += nt) {
    const int w = k / p.SPW;
    const int* bk = B.box + 4 * k;
    for (int w2 = w + 1; w2 < p.W; w2++) {
      if (s_wbox[w][1] < s_wbox[w2][0] || s_wbox[w2][1] < s_wbox[w][0] || s_wbox[w][3] < s_wbox[w2][2] ||
          s_wbox[w2][3] < s_wbox[w][2])
        continue;
      for (int l = w2 * p.SPW; l < (w2 + 1) * p.SPW; l++) {
        const int* bl = B.box + 4 * l;
        if (bk[1] < bl[0] || bl[1] < bk[0] || bk[3] < bl[2] || bl[3] < bk[2]) continue;
        raise_error(st, DERR_WINGS_COUPLED, k);
      }
    }
  }
  if (tid == 0) {
    st->adj_n = (failed || overflow) ? 0 : total;
    st->bw = s_bw;
    if (s_bw > BMAX) raise_error(st, DERR_BAND, s_bw);
  }
}

void launch_adjoint_build(hipStream_t s, const IbmParams& p, const IbmBuffers& B) {
  hipLaunchKernelGGL(k_adjoint_build, dim3(1), dim3(BUILD_THREADS), 0, s, p, B);
}

// Episode reset: release the lattice nodes of the current adjoint list
// (move_boundary's body reset, airfoil.f90:275-284).
__global__ void k_clear_adjoint(IbmParams p, IbmBuffers B) {
  const int n = B.st->adj_n;
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
    const long long node = (long long)B.adj_ij[2 * a + 1] * p.nx + B.adj_ij[2 * a];
    B.owner[node] = INT_MAX;
    B.aidx[node] = -1;
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

// One thread per band entry (position p, offset d) of each wing:
// S(k, l) = Sx(k,l) Sy(k,l) for k = seg(p), l = seg(p-d); threads with d == 0
// also form the right-hand side U_b - D u* (boundary_condition.f90:1620-1629).
__global__ __launch_bounds__(256) void k_ibm_assemble(IbmParams p, IbmBuffers B) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int D = BMAX + 1;
  if (t >= p.nseg * D) return;
  if (B.st->error != 0) return;
  const int gp = t / D, d = t - gp * D;  // gp = w*SPW + position
  const int w = gp / p.SPW, pos = gp - w * p.SPW;
  const int k = B.seg_of_pos[gp];
  const int* bk = B.box + 4 * k;
  const double* pxk = B.phx + BOXW * k;
  const double* pyk = B.phy + BOXW * k;
  double sval = 0.0;
  if (d <= B.st->bw && pos - d >= 0) {
    const int l = B.seg_of_pos[gp - d];
    const int* bl = B.box + 4 * l;
    const int ia = max(bk[0], bl[0]), ib = min(bk[1], bl[1]);
    const int ja = max(bk[2], bl[2]), jb = min(bk[3], bl[3]);
    if (ia <= ib && ja <= jb) {
      const double* pxl = B.phx + BOXW * l;
      const double* pyl = B.phy + BOXW * l;
      double sx = 0.0, sy = 0.0;
      for (int i = ia; i <= ib; i++) sx += pxk[i - bk[0]] * pxl[i - bl[0]];
      for (int j = ja; j <= jb; j++) sy += pyk[j - bk[2]] * pyl[j - bl[2]];
      sval = sx * sy;
    }
  }
  B.band[(long long)gp * D + d] = sval;
  if (d == 0) {
    double su = 0.0, sv = 0.0;
    for (int i = bk[0]; i <= bk[1]; i++)
      for (int j = bk[2]; j <= bk[3]; j++) {
        const int a = B.aidx[(long long)j * p.nx + i];
        const double phi = pxk[i - bk[0]] * pyk[j - bk[2]];
        su = su + phi * B.adj_us[a];
        sv = sv + phi * B.adj_vs[a];
      }
    B.rhs[gp] = B.ub[k] - su;
    B.rhs[p.nseg + gp] = B.vb[k] - sv;
  }
}

void launch_ibm_assemble(hipStream_t s, const IbmParams& p, const IbmBuffers& B) {
  const int n = p.nseg * (BMAX + 1);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_ibm_assemble, dim3((n + 255) / 256), dim3(256), 0, s, p, B);
}

// ----------------------------------------------------------------- solve

size_t ibm_solve_lds_bytes(int spw) { return (size_t)spw * (BMAX + 1 + 2) * sizeof(double); }

// Banded Cholesky S = L L^T and two triangular solves per wing, one wave per
// wing, the band (stride BMAX+1, entry (p,d) = S(p, p-d)) and both
// right-hand sides resident in LDS.
__global__ __launch_bounds__(64) void k_ibm_solve(IbmParams p, IbmBuffers B) {
  extern __shared__ double lds[];
  const int w = blockIdx.x, lane = threadIdx.x;
  const int n = p.SPW, D = BMAX + 1;
  double* L = lds;            // n * D
  double* r = lds + n * D;    // 2 * n
  DevStep* st = B.st;
  if (st->error != 0) return;
  const int bw = st->bw;
  const double* gband = B.band + (long long)w * n * D;
  for (int q = lane; q < n * D; q += 64) L[q] = gband[q];
  for (int q = lane; q < n; q += 64) {
    r[q] = B.rhs[w * n + q];
    r[n + q] = B.rhs[p.nseg + w * n + q];
  }
  __syncthreads();
  // factorisation, right-looking, lane i owns row k+i of the trailing window
  for (int k = 0; k < n; k++) {
    if (lane == 0) {
      const double d0 = L[k * D];
      if (!(d0 > 0.0)) raise_error(st, DERR_NOT_SPD, w * n + k);
      L[k * D] = sqrt(d0 > 0.0 ? d0 : 1.0);
    }
    __syncthreads();
    const double dk = L[k * D];
    const int i = lane + 1;
    if (i <= bw && k + i < n) L[(k + i) * D + i] = L[(k + i) * D + i] / dk;
    __syncthreads();
    if (i <= bw && k + i < n) {
      const double lik = L[(k + i) * D + i];
      for (int j = 1; j <= i; j++) L[(k + i) * D + (i - j)] -= lik * L[(k + j) * D + j];
    }
    __syncthreads();
  }
  // forward solve L z = r, column oriented; lanes 0-31 first rhs, 32-63 second
  const int side = lane >> 5, dd = (lane & 31) + 1;
  for (int q = 0; q < n; q++) {
    if ((lane & 31) == 0) r[side * n + q] = r[side * n + q] / L[q * D];
    __syncthreads();
    if (dd <= bw && q + dd < n) r[side * n + q + dd] -= L[(q + dd) * D + dd] * r[side * n + q];
    __syncthreads();
  }
  // backward solve L^T y = z
  for (int q = n - 1; q >= 0; q--) {
    if ((lane & 31) == 0) r[side * n + q] = r[side * n + q] / L[q * D];
    __syncthreads();
    if (dd <= bw && q - dd >= 0) r[side * n + q - dd] -= L[q * D + dd] * r[side * n + q];
    __syncthreads();
  }
  for (int q = lane; q < n; q += 64) {
    B.rhs[w * n + q] = r[q];
    B.rhs[p.nseg + w * n + q] = r[n + q];
  }
}

void launch_ibm_solve(hipStream_t s, const IbmParams& p, const IbmBuffers& B) {
  if (p.W <= 0) return;
  hipLaunchKernelGGL(k_ibm_solve, dim3(p.W), dim3(64), ibm_solve_lds_bytes(p.SPW), s, p, B);
}

// ---------------------------------------------------------------- spread

// g = D^T y at every adjoint node (boundary_condition.f90:1639-1640), summed
// over the segments whose support contains the node, all of which lie within
// the band around the claiming segment; u = u* + g, body force 2 rho g and
// the moment contribution (boundary_condition.f90:1645-1675).
__global__ __launch_bounds__(256) void k_ibm_spread(IbmParams p, IbmBuffers B) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  DevStep* st = B.st;
  if (st->error != 0 || a >= st->adj_n) return;
  const int i = B.adj_ij[2 * a], j = B.adj_ij[2 * a + 1];
  const int k0 = B.adj_seg[a];
  const int w = k0 / p.SPW;
  const int p0 = B.pos_of_seg[k0];
  const int bw = st->bw;
  const int base = w * p.SPW;
  const double* yu = B.rhs;
  const double* yv = B.rhs + p.nseg;
  double gx = 0.0, gy = 0.0;
  const int qa = max(0, p0 - bw), qb = min(p.SPW - 1, p0 + bw);
  for (int q = qa; q <= qb; q++) {
    const int k = B.seg_of_pos[base + q];
    const int* bk = B.box + 4 * k;
    if (i < bk[0] || i > bk[1] || j < bk[2] || j > bk[3]) continue;
    const double phi = B.phx[BOXW * k + (i - bk[0])] * B.phy[BOXW * k + (j - bk[2])];
    gx = gx + phi * yu[base + q];
    gy = gy + phi * yv[base + q];
  }
  const double rho = B.adj_rho[a];
  B.adj_u[a] = B.adj_us[a] + gx;
  B.adj_v[a] = B.adj_vs[a] + gy;
  const double t1 = 2.0 * rho * gx, t2 = 2.0 * rho * gy;
  B.adj_gx[a] = t1;
  B.adj_gy[a] = t2;
  const WingMotionState& m = st->wing[w];
  const double x0 = (p.ox + p.scale * (double)i - (p.xc + p.wing_x0 + (double)w * p.wing_dx + m.sa)) / p.scale -
                    p.body_extra_pos[w];
  const double y0 = (p.oy + p.scale * (double)j - (p.yc + p.wing_y0 + m.sb)) / p.scale;
  B.adj_mom[a] = -x0 * t2 + y0 * t1;
}

void launch_ibm_spread(hipStream_t s, const IbmParams& p, const IbmBuffers& B) {
  if (p.cap <= 0) return;
  hipLaunchKernelGGL(k_ibm_spread, dim3((p.cap + 255) / 256), dim3(256), 0, s, p, B);
}

// ---------------------------------------------------------------- forces

// Per-wing force and moment (boundary_condition.f90:1662-1675) as fixed-order
// tree reductions over the wing's contiguous range of the adjoint list, the
// normalisation and force row (solution.f90:330-378), the airfoil_motion row
// (solution.f90:390-401) and the running row sum (solution.f90:379-385).
__global__ __launch_bounds__(256) void k_forces(IbmParams p, IbmBuffers B, int slot, int row_len) {
  __shared__ double sx[256], sy[256], sm[256];
  __shared__ double wf[MAXW][2], wmom[MAXW];
  DevStep* st = B.st;
  const int tid = threadIdx.x;
  const bool ok = st->error == 0;
  for (int w = 0; w < p.W; w++) {
    double ax = 0.0, ay = 0.0, am = 0.0;
    if (ok) {
      const int a0 = B.seg_off[w * p.SPW], a1 = B.seg_off[(w + 1) * p.SPW];
      for (int a = a0 + tid; a < a1; a += blockDim.x) {
        ax = ax - B.adj_gx[a];
        ay = ay - B.adj_gy[a];
        am = am + B.adj_mom[a];
      }
    }
    sx[tid] = ax;
    sy[tid] = ay;
    sm[tid] = am;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (tid < h) {
        sx[tid] += sx[tid + h];
        sy[tid] += sy[tid + h];
        sm[tid] += sm[tid + h];
      }
      __syncthreads();
    }
    if (tid == 0) {
      wf[w][0] = sx[0];
      wf[w][1] = sy[0];
      wmom[w] = sm[0];
    }
    __syncthreads();
  }
  if (tid != 0) return;
  const double UU = p.uu, scale = p.scale;
  double F0 = 0.0, F1 = 0.0, Mt = 0.0;
  for (int w = 0; w < p.W; w++) {
    wf[w][0] = 2.0 * wf[w][0] / (UU * UU * (1.0 / scale));
    wf[w][1] = 2.0 * wf[w][1] / (UU * UU * (1.0 / scale));
    wmom[w] = -2.0 * wmom[w] / (UU * UU * (1.0 / (scale * scale)));
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
  hipLaunchKernelGGL(k_forces, dim3(1), dim3(256), 0, s, p, B, slot, row_len);
}

}  // namespace lbm
