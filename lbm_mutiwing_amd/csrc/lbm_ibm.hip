// lbm_ibm.hip - implicit velocity-correction immersed boundary on gfx950.
//
// The reference (immersed_boundary_Wu, boundary_condition.f90:1512-1692)
// builds dense nseg x nadj delta matrices D and E = D^T diag(arc), forms
// A = D E with DGEMM and solves A x = U_b - D u* by LU (cuSOLVER) every
// step.  The same system is solved here in two parts.
//
// Geometry (depends on the motion only, so it is computed for a batch of
// future steps at once, on its own stream, ahead of the flow):
//   * markers, midpoints, support boxes and 1-D delta weights
//     (move_boundary, airfoil.f90:244-338);
//   * the adjoint node list in the reference's order (segment, then i, then
//     j; the first claiming segment owns a node, airfoil.f90:286-335), built
//     per wing with a claim map over the wing's support rectangle;
//   * S = D D^T.  A = S diag(arc) and the spread g = E x = D^T (S^-1 c), so
//     arc cancels.  The delta kernel is separable: S(k,l) = Sx(k,l) Sy(k,l)
//     with 1-D sums over the overlap of the two support boxes.  Segments are
//     ordered zig-zag (0, n-1, 1, n-2, ...), which puts segments that face
//     each other on the two surfaces next to each other and makes S a narrow
//     band; the band is factored S = L D L^T and inverted explicitly.
// Flow (per step, three short kernels overlapping the lattice kernel):
//   c = U_b - D u*, with u* pulled from the previous step's populations at
//   every box cell; y = S^-1 c as a dense mat-vec; g = D^T y at every
//   adjoint node.  The adjoint-node re-collision (lbm_lattice.hip) applies g,
//   the Guo forcing 2 rho g and reduces the forces.
#include <algorithm>
#include <climits>

#include "lbm_band.h"
#include "lbm_internal.h"

namespace lbm {

__device__ __forceinline__ void raise_error(DevStep* st, int code, int info) {
  if (atomicCAS(&st->error, 0, code) == 0) st->error_info = info;
}

// Load that bypasses the CU's vector cache (values written by atomics or by
// other workgroups of the same launch).
__device__ __forceinline__ int load_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
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
__device__ WingMotionState wing_motion(const IbmParams& p, const GeoBuffers& G, int w, double time, bool* ok) {
  double r[9];
  const int col = 1 + w * 9;
  *ok = true;
  if (p.motion_source == LBM_MOTION_STREAM) {
    // lookup_stream_motion_row (mode.f90:414-449)
    const double* S = G.stream_rows;
    const double* T = G.stream_rows + LBM_MAX_MOTION_COLS;
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
    if (p.step_count < 1 || G.step_table == nullptr) {
      *ok = false;
      for (int c = 0; c < 9; c++) r[c] = 0.0;
    } else {
      long long idx = llround(time / p.motion_dt) + 1;
      if (idx > p.step_count) idx = p.step_count;
      if (idx < 1) idx = 1;
      const double* row = G.step_table + (idx - 1) * (long long)p.ncols;
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

// ------------------------------------------------ geometry: markers, list

// dst[e] = src[e] for e < count by NT threads, LOAD_BATCH loads per thread in
// flight at a time: a plain copy loop waits for each load before the next
// (the store needs its value), one memory latency per element per thread
constexpr int LOAD_BATCH = 8;
template <int NT, typename T>
__device__ __forceinline__ void copy_batched(T* __restrict__ dst, const T* __restrict__ src, int count) {
  for (int e0 = (int)threadIdx.x; e0 < count; e0 += LOAD_BATCH * NT) {
    T v[LOAD_BATCH];
#pragma unroll
    for (int u = 0; u < LOAD_BATCH; u++) {
      const int e = e0 + u * NT;
      v[u] = src[e < count ? e : 0];
    }
#pragma unroll
    for (int u = 0; u < LOAD_BATCH; u++)
      if (e0 + u * NT < count) dst[e0 + u * NT] = v[u];
  }
}

// Threads of a geometry workgroup (marker pass and band factor).  A single
// instance issues few (step, wing) workgroups per geometry batch, so each is
// wide; a batch of episodes issues thousands, which fill the GPU better as
// narrower workgroups (measured on 64 episodes: 512 -> 256 -> 128 threads
// 74.3 -> 79.5 -> 80.9 env-steps/s; a single 4096x2048 instance with two
// wings loses 10 % at 256).
constexpr int GEO_THREADS = 512;
constexpr int GEO_THREADS_BATCH = 128;
constexpr int CLAIM_UNROLL = 12;

size_t geo_wing_lds_bytes(int spw) {
  return (size_t)spw * sizeof(unsigned long long) + (size_t)spw * 4 * sizeof(int) + (size_t)(spw + 1) * sizeof(int);
}

// One workgroup per (step, wing).  Motion and marker update, midpoints,
// rigid-body velocity at the midpoints (boundary_condition.f90:1540-1555),
// support boxes (airfoil.f90:288-321), delta weights of every box column and
// row (boundary_condition.f90:1558-1600), then the adjoint list: every box
// cell claims its node in the wing's claim map with its segment index
// (atomicMin: lowest segment wins, which is the reference's first claimant).
// The band half-width of S follows in the solve, from the boxes.
template <int NT>
__device__ __forceinline__ void geo_wing_body(const IbmParams& p, const GeoBuffers& G, DevStep* st, int slot0,
                                              int n0, int fidx0, int b, int w) {
  extern __shared__ unsigned long long lds_geo[];
  __shared__ WingMotionState s_m;
  __shared__ int s_rect[4];
  __shared__ int s_total;
  __shared__ int s_part[NT];
  const int t = threadIdx.x;
  const int slot = slot0 + b, n = n0 + b, SPW = p.SPW;
  unsigned long long* smask = lds_geo;                       // SPW claim masks (bit = box cell)
  int* sbox = reinterpret_cast<int*>(smask + SPW);           // SPW * 4
  int* scnt = sbox + 4 * SPW;                                // SPW + 1
  const Frame fr = G.frame[fidx0 + b];
  const double time = (double)(n - 1) * p.dt;
  const double scale = p.scale, UU = p.uu, ox = p.ox, oy = p.oy;
  if (t == 0) {
    bool ok = true;
    s_m = wing_motion(p, G, w, time, &ok);
    if (!ok) raise_error(st, DERR_MOTION, w);
    s_rect[0] = INT_MAX; s_rect[1] = INT_MIN; s_rect[2] = INT_MAX; s_rect[3] = INT_MIN;
  }
  __syncthreads();
  const WingMotionState m = s_m;
  const double ca = cos(m.angle), sn = sin(m.angle);
  const size_t sp = (size_t)slot * p.npts, ss = (size_t)slot * p.nseg;
  for (int q = t; q < p.P; q += NT) {
    const int pa = w * p.P + q;
    const double x0 = G.xb0[pa], y0 = G.yb0[pa];
    G.xb1[sp + pa] = ca * x0 + sn * y0 + p.wing_x0 + (double)w * p.wing_dx + m.sa + p.xc +
                     fr.body_extra_s[w] / (1.0 / scale) - (double)fr.movenum * scale;
    G.yb1[sp + pa] = -sn * x0 + ca * y0 + p.wing_y0 + m.sb + p.yc;
  }
  int ri0 = INT_MAX, ri1 = INT_MIN, rj0 = INT_MAX, rj1 = INT_MIN;
  for (int k = t; k < SPW; k += NT) {
    const int gk = w * SPW + k, pa = w * p.P + k;
    double xe[2], ye[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const double x0 = G.xb0[pa + e], y0 = G.yb0[pa + e];
      xe[e] = ca * x0 + sn * y0 + p.wing_x0 + (double)w * p.wing_dx + m.sa + p.xc +
              fr.body_extra_s[w] / (1.0 / scale) - (double)fr.movenum * scale;
      ye[e] = -sn * x0 + ca * y0 + p.wing_y0 + m.sb + p.yc;
    }
    const double X = 0.5 * (xe[1] + xe[0]);
    const double Y = 0.5 * (ye[1] + ye[0]);
    G.xb2[ss + gk] = X;
    G.yb2[ss + gk] = Y;
    const double xs = G.xb0[pa + 1] + G.xb0[pa], ys = G.yb0[pa + 1] + G.yb0[pa];
    const double om = m.omega, an = m.angle;
    G.ub[ss + gk] = UU * (-om * sin(an) * 0.5 * xs + om * cos(an) * 0.5 * ys + m.sau) + fr.body_extra_v[w];
    G.vb[ss + gk] = UU * (-om * cos(an) * 0.5 * xs - om * sin(an) * 0.5 * ys + m.sbv);
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
      raise_error(st, DERR_BOX_EDGE, gk);
      i1 = 4; i2 = 3; j1 = 4; j2 = 3;  // empty box
    }
    int* gb = G.box + (ss + gk) * 4;
    gb[0] = sbox[4 * k + 0] = i1 - 1;
    gb[1] = sbox[4 * k + 1] = i2 - 1;
    gb[2] = sbox[4 * k + 2] = j1 - 1;
    gb[3] = sbox[4 * k + 3] = j2 - 1;
    double* px = G.phx + (ss + gk) * BOXW;
    double* py = G.phy + (ss + gk) * BOXW;
#pragma unroll
    for (int c = 0; c < BOXW; c++) {
      const int i = i1 + c, j = j1 + c;
      const double x0 = ox + scale * (double)(i - 1);
      const double y0 = oy + scale * (double)(j - 1);
      px[c] = i <= i2 ? delta_kernel(fabs((x0 - X) / scale)) : 0.0;
      py[c] = j <= j2 ? delta_kernel(fabs((y0 - Y) / scale)) : 0.0;
    }
    smask[k] = 0ull;
    if (i2 >= i1 && j2 >= j1) {
      ri0 = min(ri0, i1 - 1); ri1 = max(ri1, i2 - 1);
      rj0 = min(rj0, j1 - 1); rj1 = max(rj1, j2 - 1);
    }
  }
  if (ri1 >= ri0) {
    atomicMin(&s_rect[0], ri0); atomicMax(&s_rect[1], ri1);
    atomicMin(&s_rect[2], rj0); atomicMax(&s_rect[3], rj1);
  }
  __syncthreads();
  const int i0 = s_rect[0], j0 = s_rect[2];
  const int wr = s_rect[1] - i0 + 1, hr = s_rect[3] - j0 + 1;
  const bool empty = s_rect[1] < s_rect[0];
  const long long cells = empty ? 0 : (long long)wr * hr;
  GeoStep* gs = G.step + slot;
  int* soff = G.seg_off + ((size_t)slot * p.W + w) * (SPW + 1);
  if (empty || cells > p.map_cap) {
    if (!empty && t == 0) raise_error(st, DERR_MAP, w);
    for (int k = t; k <= SPW; k += NT) soff[k] = 0;
    if (t == 0) {
      gs->wing[w] = m;
      gs->cnt[w] = 0;
      gs->bw[w] = 0;
      gs->rect[w][0] = 0; gs->rect[w][1] = -1; gs->rect[w][2] = 0; gs->rect[w][3] = -1;
      if (w == 0) { gs->time = time; gs->n = n; gs->frame = fr; }
    }
    return;
  }
  // The claim map of a (step, wing) is written and read by this workgroup
  // alone within the launch, so the barriers' workgroup-scope ordering is
  // enough (the claims are read past the CU's cache, by load_agent).  An
  // agent-scope fence here compiles to a write-back and an invalidation of the
  // XCD's whole L2 (buffer_wbl2 / buffer_inv), twice per workgroup: beside
  // the thousands of workgroups of a batched geometry batch (64 episodes) the
  // lattice pass took 1.41 ms with it and 1.00 ms without (0.67 ms in the
  // steps without geometry), and the marker pass itself 0.68 against 0.27 ms.
  int* M = G.map + ((size_t)slot * p.W + w) * (size_t)p.map_cap;  // owner segment per cell
  for (long long c = t; c < cells; c += NT) M[c] = INT_MAX;
  __syncthreads();
  for (int e = t; e < SPW * BOXC; e += NT) {
    const int k = e / BOXC, c = e - k * BOXC;
    const int* bx = sbox + 4 * k;
    const int i = bx[0] + c / BOXW, j = bx[2] + c % BOXW;
    if (i > bx[1] || j > bx[3]) continue;
    atomicMin(&M[(i - i0) + wr * (j - j0)], k);
  }
  __syncthreads();
  for (int e0 = t; e0 < SPW * BOXC; e0 += CLAIM_UNROLL * NT) {
    int own[CLAIM_UNROLL];
#pragma unroll
    for (int u = 0; u < CLAIM_UNROLL; u++) {
      const int e = e0 + u * NT;
      own[u] = -1;
      if (e < SPW * BOXC) {
        const int k = e / BOXC, c = e - k * BOXC;
        const int* bx = sbox + 4 * k;
        const int i = bx[0] + c / BOXW, j = bx[2] + c % BOXW;
        if (i <= bx[1] && j <= bx[3]) own[u] = load_agent(&M[(i - i0) + wr * (j - j0)]);
      }
    }
#pragma unroll
    for (int u = 0; u < CLAIM_UNROLL; u++) {
      const int e = e0 + u * NT;
      if (own[u] < 0) continue;
      const int k = e / BOXC, c = e - k * BOXC;
      if (own[u] == k) atomicOr(&smask[k], 1ull << c);
    }
  }
  __syncthreads();
  // per-segment claim counts, exclusive scan, list in (segment, i, j) order
  for (int k = t; k < SPW; k += NT) scnt[k] = __popcll(smask[k]);
  __syncthreads();
  const int chunk = (SPW + NT - 1) / NT;
  const int c0 = min(t * chunk, SPW), c1 = min(c0 + chunk, SPW);
  int local = 0;
  for (int k = c0; k < c1; k++) local += scnt[k];
  s_part[t] = local;
  __syncthreads();
  for (int d = 1; d < NT; d <<= 1) {
    const int v = t >= d ? s_part[t - d] : 0;
    __syncthreads();
    s_part[t] += v;
    __syncthreads();
  }
  int run = s_part[t] - local;
  for (int k = c0; k < c1; k++) {
    const int cnt = scnt[k];
    scnt[k] = run;
    soff[k] = run;
    run += cnt;
  }
  if (t == NT - 1) {
    s_total = s_part[t];
    soff[SPW] = s_part[t];
  }
  __syncthreads();
  int* aij = G.adj_ij + ((size_t)slot * p.cap + (size_t)w * p.capw) * 2;
  int* aseg = G.adj_seg + (size_t)slot * p.cap + (size_t)w * p.capw;
  for (int k = t; k < SPW; k += NT) {
    unsigned long long mk = smask[k];
    int a = scnt[k];
    const int* bx = sbox + 4 * k;
    while (mk) {
      const int c = __ffsll((long long)mk) - 1;
      mk &= mk - 1ull;
      aij[2 * a] = bx[0] + c / BOXW;
      aij[2 * a + 1] = bx[2] + c % BOXW;
      aseg[a] = k;
      a++;
    }
  }
  if (t == 0) {
    gs->wing[w] = m;
    gs->cnt[w] = s_total;
    gs->rect[w][0] = i0; gs->rect[w][1] = i0 + wr - 1; gs->rect[w][2] = j0; gs->rect[w][3] = j0 + hr - 1;
    if (w == 0) { gs->time = time; gs->n = n; gs->frame = fr; }
  }
}

__global__ __launch_bounds__(GEO_THREADS) void k_geo_wing(IbmParams p, GeoBuffers G, DevStep* st, int slot0, int n0,
                                                          int fidx0) {
  geo_wing_body<GEO_THREADS>(p, G, st, slot0, n0, fidx0, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------- geometry: S and its inverse

__device__ __forceinline__ double s_entry(const int* box, const double* phx, const double* phy, int k, int l) {
  const int* bk = box + 4 * k;
  const int* bl = box + 4 * l;
  const int ia = max(bk[0], bl[0]), ib = min(bk[1], bl[1]);
  const int ja = max(bk[2], bl[2]), jb = min(bk[3], bl[3]);
  if (ia > ib || ja > jb) return 0.0;
  const double* pxk = phx + BOXW * k;
  const double* pyk = phy + BOXW * k;
  const double* pxl = phx + BOXW * l;
  const double* pyl = phy + BOXW * l;
  double sx = 0.0, sy = 0.0;
  for (int i = ia; i <= ib; i++) sx += pxk[i - bk[0]] * pxl[i - bl[0]];
  for (int j = ja; j <= jb; j++) sy += pyk[j - bk[2]] * pyl[j - bl[2]];
  return sx * sy;
}

// s_entry with every load unconditional (box from LDS, the six weights of
// each segment at clamped offsets) and the same sums in the same order, so
// that the loads of several entries can be issued together
__device__ __forceinline__ double s_entry_flat(const int* box, const double* phx, const double* phy, int k, int l) {
  const int* bk = box + 4 * k;
  const int* bl = box + 4 * l;
  const int ia = max(bk[0], bl[0]), ib = min(bk[1], bl[1]);
  const int ja = max(bk[2], bl[2]), jb = min(bk[3], bl[3]);
  const double* pxk = phx + BOXW * k;
  const double* pyk = phy + BOXW * k;
  const double* pxl = phx + BOXW * l;
  const double* pyl = phy + BOXW * l;
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int c = 0; c < BOXW; c++) {   // i = bk[0] + c covers ia..ib (boxes are at most BOXW wide)
    const int i = bk[0] + c;
    const bool in = i >= ia && i <= ib;
    const double a = pxk[c], b = pxl[in ? i - bl[0] : 0];
    if (in) sx += a * b;
  }
#pragma unroll
  for (int c = 0; c < BOXW; c++) {
    const int j = bk[2] + c;
    const bool in = j >= ja && j <= jb;
    const double a = pyk[c], b = pyl[in ? j - bl[2] : 0];
    if (in) sy += a * b;
  }
  return ia > ib || ja > jb ? 0.0 : sx * sy;
}

// dst[e] = S(k, l) for e < count, where entry(e, k, l) names the segments of
// entry e (false: the entry is zero), AB entries per thread at a time with
// all their loads issued before their sums.  AB = 4 for the wide workgroups
// of a single instance, whose few solves wait on memory latency; 1 for the
// batch, whose many solves share the CUs (AB = 4 takes 231 VGPRs a lane).
template <int NT, typename Entry>
__device__ __forceinline__ void assemble_s(double* dst, int count, const int* sbox, const double* phx,
                                           const double* phy, Entry entry) {
  constexpr int AB = NT >= GEO_THREADS ? 4 : 1;
  for (int e0 = (int)threadIdx.x; e0 < count; e0 += AB * NT) {
    double v[AB];
#pragma unroll
    for (int u = 0; u < AB; u++) {
      int k = 0, l = 0;
      const bool ok = entry(e0 + u * NT, k, l);
      const double x = s_entry_flat(sbox, phx, phy, ok ? k : 0, ok ? l : 0);
      v[u] = ok ? x : 0.0;
    }
#pragma unroll
    for (int u = 0; u < AB; u++)
      if (e0 + u * NT < count) dst[e0 + u * NT] = v[u];
  }
}

// In-place S = L D L^T of the lower band (band[r*Dd + d] = S(r, r-d), row
// stride Dd > bw) by one wave: afterwards band[k*Dd] = D_k,
// band[(k+i)*Dd + i] = L(k+i, k), invd[k] = 1/D_k.  Step k updates
// S(k+i, k+j) -= (S(k+i, k) / D_k) S(k+j, k) for 1 <= j <= i <= bw with the
// column k not yet scaled: the bw(bw+1)/2 updates write distinct entries and
// read only column k, so each lane takes up to NPL (i, j) pairs of its own
// and a step costs one round of LDS loads and stores.  The loads of a step do
// not wait for its pivot: each lane loads its pairs' operands, and every lane
// the entries of the next pivot, as the step starts, and forms the next
// pivot D_{k+1} itself with the arithmetic of the (1, 1) update, so that the
// division of step k+1 overlaps the stores of step k.  LDS operations of one
// wave complete in order, so a step's loads see the previous step's stores.
// Column k is scaled by 1/D_k after the last step (it is not read again).
constexpr int LDLT_PAIRS_PER_LANE = (BMAX * (BMAX + 1) / 2 + 63) / 64;

template <int NPL>
__device__ void band_ldlt_t(double* band, double* invd, int n, int npiv, int bw, int Dd, int lane, DevStep* st,
                            int info) {
  const int npairs = bw * (bw + 1) / 2;
  // offsets from row k of a pair's target S(k+i, k+j), and of its operands
  // S(k+i, k) and S(k+j, k); pairs past the last row or past npairs write
  // nothing (their loads read row k)
  int ot[NPL], oa[NPL], oc[NPL], pi[NPL];
#pragma unroll
  for (int u = 0; u < NPL; u++) {
    const int q = lane + 64 * u;
    int i = 0, j = 0;
    if (q < npairs) {
      i = 1;
      while ((i + 1) * i / 2 <= q) i++;
      j = q - i * (i - 1) / 2 + 1;
    }
    pi[u] = i;
    ot[u] = i > 0 ? i * Dd + (i - j) : 0;
    oa[u] = i > 0 ? i * Dd + i : 0;
    oc[u] = i > 0 ? j * Dd + j : 0;
  }
  double d = band[0];
  int bad = -1;
  double* sink = invd + n;   // an unused entry: the store of a pair with nothing to update this step
  for (int k = 0; k < npiv; k++) {
    double* bk = band + (size_t)k * Dd;
    double vt[NPL], va[NPL], vc[NPL];
#pragma unroll
    for (int u = 0; u < NPL; u++) {
      vt[u] = bk[ot[u]];
      va[u] = bk[oa[u]];
      vc[u] = bk[oc[u]];
    }
    const double dn = bk[Dd], ln = bk[Dd + 1];   // S(k+1, k+1), S(k+1, k): row k+1 exists (padding rows)
    if (!(d > 0.0) && bad < 0) bad = k;
    const double rd = 1.0 / (d > 0.0 ? d : 1.0);
    d = dn - ln * rd * ln;   // the (1, 1) update: D_{k+1} (entries beyond bw are zero)
    // every lane stores (to the sink when its pair has no update), so that
    // the loads above stay ahead of the division instead of sinking into a
    // conditional store; 1/D_k is the same in every lane
#pragma unroll
    for (int u = 0; u < NPL; u++) {
      double* dst = pi[u] > 0 && k + pi[u] < n ? bk + ot[u] : sink;
      *dst = vt[u] - va[u] * rd * vc[u];
    }
    invd[k] = rd;
    // no wait for these stores to complete: the next step's loads may alias
    // them, so the compiler keeps them after the stores, and the LDS executes
    // a wave's operations in program order
  }
  wave_sync();
  if (lane == 0 && bad >= 0) raise_error(st, DERR_NOT_SPD, info + bad);
  // L(k + i, k) = S(k + i, k) / D_k
  for (int e = lane; e < npiv * bw; e += 64) {
    const int k = e / bw, i = 1 + e - (e / bw) * bw;
    if (k + i < n) band[(k + i) * Dd + i] = band[(k + i) * Dd + i] * invd[k];
  }
  wave_sync();
}

// L D L^T of the first npiv pivots of an n-row band (npiv = n: the whole
// factorisation; npiv < n leaves rows npiv.. holding the Schur complement of
// the eliminated block, their entries in columns < npiv scaled to L)
__device__ void band_ldlt(double* band, double* invd, int n, int npiv, int bw, int Dd, int lane, DevStep* st,
                          int info) {
  const int npl = (bw * (bw + 1) / 2 + 63) / 64;
  if (npl <= 1)
    band_ldlt_t<1>(band, invd, n, npiv, bw, Dd, lane, st, info);
  else if (npl == 2)
    band_ldlt_t<2>(band, invd, n, npiv, bw, Dd, lane, st, info);
  else if (npl == 3)
    band_ldlt_t<3>(band, invd, n, npiv, bw, Dd, lane, st, info);
  else
    band_ldlt_t<LDLT_PAIRS_PER_LANE>(band, invd, n, npiv, bw, Dd, lane, st, info);
}

// One row of a band substitution: a0 minus the BW products of the row's
// entries l[1..BW] with the ring of the last BW values (slot r - d for d).
template <int BW>
__device__ __forceinline__ double band_dot(const double (&l)[BW + 1], const double (&w)[BW], int r, double a0) {
  double a1 = 0.0, a2 = 0.0;
#pragma unroll
  for (int d = BW; d >= 1; d--) {   // the newest value (d = 1) last: it ends the row's dependency chain
    const double x = w[(r - d + BW) % BW];
    if (d % 3 == 1)
      a0 = __builtin_fma(-l[d], x, a0);
    else if (d % 3 == 2)
      a1 = __builtin_fma(-l[d], x, a1);
    else
      a2 = __builtin_fma(-l[d], x, a2);
  }
  return (a1 + a2) + a0;
}

// the BW entries L(m, m-d) (step 1) or L(m+d, m) (step Dd+1) of a band row or
// column in LDS, d = 1..BW, into l[1..BW] (and l[0] = *extra when extra is
// given: 1/D of the row, loaded in the same group because LDS loads complete
// in order)
template <int BW>
__device__ __forceinline__ void band_load(const double* p, int step, double (&l)[BW + 1],
                                          const double* extra = nullptr) {
  if (extra) l[0] = *extra;
#pragma unroll
  for (int d = 1; d <= BW; d++) l[d] = p[d * step];
  // keep these loads ahead of the row computed after them (the scheduler
  // would otherwise sink them to their use in the next row)
  __builtin_amdgcn_sched_barrier(0);
}

// One forward row m (ring slot r): z_m = e_col(m) - sum_d L(m, m-d) z_{m-d}
// from the row's entries in cur (cur[0] = 1/D_m), while the next row's
// entries load into nxt.
template <int BW>
__device__ __forceinline__ void forward_row(const double* band, const double* invd, int n, int col, int m, int r,
                                            double (&w)[BW], const double (&cur)[BW + 1], double (&nxt)[BW + 1],
                                            double* S) {
  band_load<BW>(band + (m + 1) * (BW + 1), 1, nxt, invd + m + 1);
  const double z = band_dot<BW>(cur, w, r, m == col ? 1.0 : 0.0);
  w[r] = z;
  if (m >= col) S[(size_t)m * n + col] = z * cur[0];
}

// One backward row m (ring slot r): x_m = t - sum_d L(m+d, m) x_{m+d}, while
// the entries of row m - 1 load into nxt.
template <int BW>
__device__ __forceinline__ void backward_row(const double* band, int n, int col, int m, int r, double t,
                                             double (&w)[BW], const double (&cur)[BW + 1], double (&nxt)[BW + 1],
                                             double* S) {
  if (m >= 1) band_load<BW>(band + (m - 1) * (BW + 1), BW + 2, nxt);
  const double x = band_dot<BW>(cur, w, r, t);
  w[r] = x;
  if (col < n) S[(size_t)m * n + col] = x;
}

// Column col of S^-1 = L^-T D^-1 L^-1 e_col (col >= n: a lane with no column,
// which runs along and stores nothing) from the factors in LDS: forward
// substitution, t = D^-1 z written to the output column, then backward
// substitution over the stored t.  Every lane of the wave walks the same rows
// at the same time (a lane whose column starts later carries exact zeros until
// it does), so the band loads are broadcasts and the lanes' stores of a row
// are contiguous.  The rows are taken in blocks of BW with the last BW values
// in a register ring indexed at compile time, each row's BW products summed
// by fused multiply-adds into three partial sums, and each row's entries
// loaded while the row before it computes.  The band is stored with the
// compile-time row stride BW + 1 (BW >= bw), its entries beyond bw and BW rows
// past the last one zero, so the loads of a row have constant offsets and no
// guards (a term with a zero L entry adds an exact zero).
template <int BW>
__device__ __forceinline__ void inverse_column(const double* band, const double* invd, int n, int col, int m0, double* S) {
  static_assert(BW % 2 == 0, "the row entries alternate between two register sets");
  constexpr int Dd = BW + 1;
  double w[BW];
#pragma unroll
  for (int d = 0; d < BW; d++) w[d] = 0.0;
  // forward, row m in w[(m - m0) % BW]; rows above the wave's first column
  // are zero in every lane.  Row entries (and 1/D) are loaded one row ahead,
  // alternately into la and lb (rows up to n + BW - 1 and n + 1 entries of
  // 1/D exist); invd has n + 1 entries, the last one unused.
  double la[BW + 1], lb[BW + 1];
  band_load<BW>(band + m0 * Dd, 1, la, invd + m0);
  int mb = m0;
  for (; mb + BW <= n; mb += BW) {
#pragma unroll
    for (int r = 0; r < BW; r += 2) {
      forward_row<BW>(band, invd, n, col, mb + r, r, w, la, lb, S);
      forward_row<BW>(band, invd, n, col, mb + r + 1, r + 1, w, lb, la, S);
    }
  }
#pragma unroll
  for (int r = 0; r < BW; r += 2) {
    if (mb + r < n) forward_row<BW>(band, invd, n, col, mb + r, r, w, la, lb, S);
    if (mb + r + 1 < n) forward_row<BW>(band, invd, n, col, mb + r + 1, r + 1, w, lb, la, S);
  }
  // backward from m = n-1 down, row m = n-1-q in w[q % BW];
  // L(m+d, m) = band[(m+d)*Dd + d], step Dd+1 in d
#pragma unroll
  for (int d = 0; d < BW; d++) w[d] = 0.0;
  double tv[BW];
#pragma unroll
  for (int r = 0; r < BW; r++) {
    const int m = n - 1 - r;
    tv[r] = (m >= 0 && m >= col) ? S[(size_t)m * n + col] : 0.0;
  }
  band_load<BW>(band + (n - 1) * Dd, Dd + 1, la);
  // the next block's t is loaded while a block computes (up to BW = 16; the
  // widest band has no registers to spare for it)
  constexpr bool AHEAD = BW <= 16;
  for (int qb = 0; qb < n; qb += BW) {
    double tn[BW];
#pragma unroll
    for (int r = 0; r < BW; r++) {
      const int m = n - 1 - ((AHEAD ? qb + BW : qb) + r);
      tn[r] = (m >= 0 && m >= col) ? S[(size_t)m * n + col] : 0.0;
    }
    if (!AHEAD) {
#pragma unroll
      for (int r = 0; r < BW; r++) tv[r] = tn[r];
    }
#pragma unroll
    for (int r = 0; r < BW; r += 2) {
      const int m = n - 1 - (qb + r);
      if (m >= 0) backward_row<BW>(band, n, col, m, r, tv[r], w, la, lb, S);
      if (m - 1 >= 0) backward_row<BW>(band, n, col, m - 1, r + 1, tv[r + 1], w, lb, la, S);
    }
    if (AHEAD) {
#pragma unroll
      for (int r = 0; r < BW; r++) tv[r] = tn[r];
    }
  }
}

// The twisted factors of one (step, wing) (Twist): T and R assembled from
// the boxes in LDS, the top and bottom ends eliminated by waves 0 and 1 at
// once, then the middle block (T's and R's Schur-complemented copies less S
// counted twice) factored densely by wave 0; the middle rows' entries among
// themselves are then cleared and their 1/D set to 1, so that the band
// substitutions leave them as the middle solve sets them.  All of it, in the
// layout of twist_doubles, to bg.
template <int NT>
__device__ void geo_twisted_factor(double* lds, const int* sbox, const double* phx, const double* phy, int n, int bw,
                                   int BWc, DevStep* st, int w, double* bg) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int Dd = BWc + 1;
  const Twist tw = twist_split(n, bw);
  const int m = tw.m;
  double* T = lds;                               // nT + BWc rows
  double* R = T + (size_t)(tw.nT + BWc) * Dd;    // nR + BWc rows
  double* iT = R + (size_t)(tw.nR + BWc) * Dd;   // nT + 1
  double* iR = iT + tw.nT + 1;                   // nR + 1
  double* mid = iR + tw.nR + 1;                  // m x m: the middle block, then its L
  double* midd = mid + m * m;                    // m: its 1/D
  double* smid = midd + m;                       // m x m: S of the middle block
  assemble_s<NT>(T, (tw.nT + BWc) * Dd, sbox, phx, phy, [&](int e, int& k, int& l) {
    const int r = e / Dd, d = e - r * Dd;
    k = seg_of_pos(r, n);
    l = seg_of_pos(r - d, n);
    return r < tw.nT && d <= bw && r - d >= 0;
  });
  // R(r, r-d) = S(n-1-r+d, n-1-r): the higher band position first, as in T
  assemble_s<NT>(R, (tw.nR + BWc) * Dd, sbox, phx, phy, [&](int e, int& k, int& l) {
    const int r = e / Dd, d = e - r * Dd;
    k = seg_of_pos(n - 1 - r + d, n);
    l = seg_of_pos(n - 1 - r, n);
    return r < tw.nR && d <= bw && r - d >= 0;
  });
  __syncthreads();
  for (int e = t; e < m * m; e += NT) {
    const int i = e / m, j = e - i * m;
    smid[e] = j <= i ? T[(size_t)(tw.p + i) * Dd + (i - j)] : 0.0;
  }
  __syncthreads();
  if (wave == 0) band_ldlt(T, iT, tw.nT, tw.p, bw, Dd, lane, st, w * 65536);
  else if (wave == 1) band_ldlt(R, iR, tw.nR, tw.q, bw, Dd, lane, st, w * 65536 + 32768);
  __syncthreads();
  if (wave == 0) {
    // S(p+i, p+j), j <= i: T's copy holds it less the top end's complement,
    // R's (row q+m-1-j, offset i-j) less the bottom end's
    for (int e = lane; e < m * m; e += 64) {
      const int i = e / m, j = e - i * m;
      mid[e] = j <= i ? T[(size_t)(tw.p + i) * Dd + (i - j)] + R[(size_t)(tw.q + m - 1 - j) * Dd + (i - j)] - smid[e]
                      : 0.0;
    }
    wave_sync();
    int bad = -1;
    for (int k = 0; k < m; k++) {
      const double dk = mid[k * m + k];
      if (!(dk > 0.0) && bad < 0) bad = k;
      const double rd = 1.0 / (dk > 0.0 ? dk : 1.0);
      for (int e = lane; e < m * m; e += 64) {
        const int i = e / m, j = e - i * m;
        if (j > k && i >= j) mid[e] = mid[e] - mid[i * m + k] * rd * mid[j * m + k];
      }
      if (lane == 0) midd[k] = rd;
      wave_sync();
    }
    for (int e = lane; e < m * m; e += 64) {   // L(i, k) = M(i, k) / D_k below the diagonal
      const int i = e / m, k = e - i * m;
      if (k < i) mid[e] = mid[e] * midd[k];
    }
    if (lane == 0 && bad >= 0) raise_error(st, DERR_NOT_SPD, w * 65536 + tw.p + bad);
    for (int e = lane; e < m * m; e += 64) {
      const int i = e / m, j = e - i * m;
      if (j <= i) {
        T[(size_t)(tw.p + i) * Dd + (i - j)] = 0.0;
        R[(size_t)(tw.q + i) * Dd + (i - j)] = 0.0;
      }
    }
    for (int i = lane; i < m; i += 64) {
      iT[tw.p + i] = 1.0;
      iR[tw.q + i] = 1.0;
    }
  }
  __syncthreads();
  const int tot = (int)twist_doubles(n, bw, BWc);
  for (int e = t; e < tot; e += NT) bg[e] = T[e];
}

// One workgroup per (step, wing): adjoint total, band assembly of S,
// L D L^T and the explicit inverse (one column per thread).
template <int NT>
__device__ __forceinline__ void geo_solve_body(const IbmParams& p, const GeoBuffers& G, DevStep* st, int slot0, int b,
                                               int w, int direct, int apart) {
  extern __shared__ double lds_s[];
  __shared__ int s_skip, s_bw, s_meet;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int slot = slot0 + b, n = p.SPW;
  GeoStep* gs = G.step + slot;
  if (t == 0) {
    s_skip = load_agent(&st->error) != 0;
    // the coupling of wings (k_geo_couple) reads the inverse; it needs wings
    // that share a node, so support rectangles that meet
    int meet = 0, gap = INT_MAX;
    for (int q = 0; q < p.W; q++)
      for (int r = q + 1; r < p.W; r++) {
        const int* a = gs->rect[q];
        const int* c = gs->rect[r];
        if (a[1] < a[0] || c[1] < c[0]) continue;   // a wing with no nodes
        meet |= a[0] <= c[1] && c[0] <= a[1] && a[2] <= c[3] && c[2] <= a[3];
        gap = min(gap, max(max(a[0] - c[1], c[0] - a[1]), max(a[2] - c[3], c[2] - a[3])));
      }
    s_meet = meet;
    if (w == 0) {
      int tot = 0, b0 = INT_MAX, b1 = INT_MIN;
      for (int q = 0; q < p.W; q++) {
        tot += gs->cnt[q];
        if (gs->rect[q][1] >= gs->rect[q][0]) {
          b0 = min(b0, gs->rect[q][2]);
          b1 = max(b1, gs->rect[q][3]);
        }
      }
      gs->adj_n = tot;
      gs->band[0] = b0;
      gs->band[1] = b1;
      gs->inv_on = !direct || meet;
      st->rect_gap = gap;
      if (apart && meet) raise_error(st, DERR_WINGS_MET, gs->n);
    }
  }
  __syncthreads();
  if (s_skip || gs->cnt[w] == 0) {
    if (t == 0) gs->bw[w] = 0;
    return;
  }
  const size_t base = (size_t)slot * p.nseg + (size_t)w * n;
  const int* box = G.box + base * 4;
  // band half-width: the largest spread in band positions of two segments
  // whose support boxes share a node (checked one past BMAX, so that a wider
  // band is reported)
  {
    int* sb = reinterpret_cast<int*>(lds_s);
    copy_batched<NT>(sb, box, 4 * n);
    if (t == 0) s_bw = 0;
    __syncthreads();
    constexpr int DMAX = BMAX + 1;
    int bwl = 0;
    for (int e = t; e < n * DMAX; e += NT) {
      const int pp = e / DMAX, d = 1 + e - pp * DMAX;
      if (pp - d < 0) continue;
      const int* bk = sb + 4 * seg_of_pos(pp, n);
      const int* bl = sb + 4 * seg_of_pos(pp - d, n);
      if (max(bk[0], bl[0]) <= min(bk[1], bl[1]) && max(bk[2], bl[2]) <= min(bk[3], bl[3])) bwl = max(bwl, d);
    }
    for (int off = 32; off > 0; off >>= 1) bwl = max(bwl, __shfl_xor(bwl, off));
    if (lane == 0) atomicMax(&s_bw, bwl);
    __syncthreads();
  }
  const int bw = s_bw;
  if (t == 0) gs->bw[w] = bw;
  // the LDS of the launch holds bands up to the class of the instance's bound
  if (bw > BMAX || band_class(bw) > p.band_cls) {
    if (t == 0) raise_error(st, DERR_BAND, bw);
    return;
  }
  // row stride: the compile-time band bound of the inverse below
  const int BWc = band_class(bw);
  const int Dd = BWc + 1;
  const double* phx = G.phx + base * BOXW;
  const double* phy = G.phy + base * BOXW;
  double* bg = G.bandg + ((size_t)slot * p.W + w) * band_cap(p);
  const int* sbox = reinterpret_cast<const int*>(lds_s);   // the boxes, staged above
  double* fac = lds_s + 2 * (size_t)n;                      // the factors after them
  // factors for a coupled step's band substitution (on its critical path): twisted
  if (direct && !s_meet && uses_twist(p, bw)) {
    geo_twisted_factor<NT>(fac, sbox, phx, phy, n, bw, BWc, st, w, bg);
    return;
  }
  double* band = fac;                            // (n + BWc) rows of Dd
  double* invd = band + (size_t)(n + BWc) * Dd;
  assemble_s<NT>(band, (n + BWc) * Dd, sbox, phx, phy, [&](int e, int& k, int& l) {
    const int pp = e / Dd, d = e - pp * Dd;
    k = seg_of_pos(pp, n);
    l = seg_of_pos(pp - d, n);
    return pp < n && d <= bw && pp - d >= 0;
  });
  __syncthreads();
  if (wave == 0) band_ldlt(band, invd, n, n, bw, Dd, lane, st, w * 65536);
  __syncthreads();
  // the factors (band, then 1/D, contiguous) to global memory for the inverse
  const int tot = (n + BWc) * Dd + n + 1;
  for (int e = t; e < tot; e += NT) bg[e] = band[e];
}

// The explicit inverse S^-1 of one (step, wing), INV_THREADS columns per
// workgroup (one wave, so that the LDS broadcasts of the band serve one wave
// per workgroup), from the factors the solve left in global memory, first
// copied to LDS.
constexpr int INV_THREADS = 64;

__device__ __forceinline__ void geo_inverse_body(const IbmParams& p, const GeoBuffers& G, DevStep* st, int slot0,
                                                 int b, int w, int chunk) {
  extern __shared__ double lds_i[];
  const int slot = slot0 + b, n = p.SPW;
  const int col0 = chunk * INV_THREADS;
  if (col0 >= n) return;
  const GeoStep* gs = G.step + slot;
  // the solve of this (step, wing) ran only without an error before it and
  // with adjoint nodes and an admissible band
  if (load_agent(&st->error) != 0 || gs->cnt[w] == 0 || !gs->inv_on) return;
  const int bw = gs->bw[w];
  if (bw > BMAX || band_class(bw) > p.band_cls) return;
  const int BWc = band_class(bw);
  const double* bg = G.bandg + ((size_t)slot * p.W + w) * band_cap(p);
  const int tot = (n + BWc) * (BWc + 1) + n + 1;
  copy_batched<INV_THREADS>(lds_i, bg, tot);
  __syncthreads();
  const double* band = lds_i;
  const double* invd = band + (size_t)(n + BWc) * (BWc + 1);
  double* S = G.sinv + ((size_t)slot * p.W + w) * (size_t)n * n;
  const int col = col0 + (int)threadIdx.x;
  const int c = col < n ? col : n + INV_THREADS;   // lanes past the last column run along and store nothing
  if (BWc == 8)
    inverse_column<8>(band, invd, n, c, col0, S);
  else if (BWc == 12)
    inverse_column<12>(band, invd, n, c, col0, S);
  else if (BWc == 16)
    inverse_column<16>(band, invd, n, c, col0, S);
  else
    inverse_column<BMAX>(band, invd, n, c, col0, S);
}

__global__ __launch_bounds__(INV_THREADS) void k_geo_inverse(IbmParams p, GeoBuffers G, DevStep* st, int slot0,
                                                             int chunks) {
  const int b = blockIdx.x / chunks, chunk = blockIdx.x - b * chunks;
  geo_inverse_body(p, G, st, slot0, b, blockIdx.y, chunk);
}

int inverse_chunks(int spw) { return (spw + INV_THREADS - 1) / INV_THREADS; }

__global__ __launch_bounds__(GEO_THREADS) void k_geo_solve(IbmParams p, GeoBuffers G, DevStep* st, int slot0,
                                                           int direct, int apart) {
  geo_solve_body<GEO_THREADS>(p, G, st, slot0, blockIdx.x, blockIdx.y, direct, apart);
}

// band of (SPW + c) rows of stride c + 1, c = the band class of the
// instance's bound, and 1/D (SPW + 1 entries: the inverse loads one row ahead)
size_t geo_factor_lds_bytes(const IbmParams& p) {
  const size_t c = (size_t)p.band_cls;
  return ((p.SPW + c) * (c + 1) + p.SPW + 1) * sizeof(double);
}

// the segments' boxes (4 ints each), then the plain factors or, for solves
// that may form twisted ones (direct), the larger of those and the twisted
// factors (rows of T and R, their 1/D, the middle block twice and its 1/D)
size_t geo_solve_lds_bytes(const IbmParams& p, bool direct) {
  const size_t c = (size_t)p.band_cls;
  const size_t twisted = ((p.SPW + 3 * c) * (c + 1) + p.SPW + c + 2 + 2 * c * c + c) * sizeof(double);
  return 2 * (size_t)p.SPW * sizeof(double) + std::max(geo_factor_lds_bytes(p), direct && p.dyn_on ? twisted : 0);
}

// the band substitution's LDS: a wing's factors (plain or twisted, with
// BAND_RING rows before and after), its right-hand sides and its solution
size_t bsolve_lds_bytes(const IbmParams& p) {
  const size_t c = (size_t)p.band_cls;
  const size_t plain = (p.SPW + c) * (c + 1) + p.SPW + 1;
  const size_t twisted = p.dyn_on ? (p.SPW + 3 * c) * (c + 1) + p.SPW + 2 * c + 2 + c * c + c : 0;
  return (std::max(plain, twisted) + 2 * BAND_RING * (c + 1) + 4 * (size_t)p.SPW) * sizeof(double);
}

// ------------------------------------------------------------ coupled wings
//
// When the support boxes of wings share lattice nodes, the IBM system is the
// one over all the wings' segments (boundary_condition.f90:1558-1640 over all
// segments): S = B + P O P^T, with B = blockdiag(S_ww) the per-wing systems,
// P the selection of the coupling segments K (the segments whose boxes meet a
// box of another wing) and O (|K| x |K|) the cross-wing entries among them
// (zero within a wing).  A shared node belongs to the lowest of its wings
// (its segments claim first, airfoil.f90:288-333).  Woodbury in the form that
// needs no inverse of O gives, with u = B^-1 c (the per-wing solves),
//   y = u - B^-1 P M P^T u,   M = (I + O G)^-1 O,   G = P^T B^-1 P
// (G block-diagonal: G_ww = S_ww^-1[K_w, K_w]).  The geometry pass finds K,
// forms M by Gauss-Jordan in LDS (k_geo_couple) and drops each wing's nodes
// that a lower wing owns from its adjoint list; the flow pass applies the
// correction after the per-wing mat-vecs (k_ibm_couple), and the spread adds
// the coupling segments of higher wings at a wing's nodes.
constexpr int CPL_THREADS = 256;
constexpr int CPL_MAX_SPW = 2048;   // segments per wing the coupling pass handles

__device__ __forceinline__ bool boxes_meet(const int* a, const int* b) {
  return a[0] <= b[1] && b[0] <= a[1] && a[2] <= b[3] && b[2] <= a[3];
}

// S(k, l) for segments of two different wings: s_entry with the boxes and
// weights of each segment given separately.
__device__ double s_cross(const int* bk, const double* pxk, const double* pyk, const int* bl, const double* pxl,
                          const double* pyl) {
  const int ia = max(bk[0], bl[0]), ib = min(bk[1], bl[1]);
  const int ja = max(bk[2], bl[2]), jb = min(bk[3], bl[3]);
  if (ia > ib || ja > jb) return 0.0;
  double sx = 0.0, sy = 0.0;
  for (int i = ia; i <= ib; i++) sx += pxk[i - bk[0]] * pxl[i - bl[0]];
  for (int j = ja; j <= jb; j++) sy += pyk[j - bk[2]] * pyl[j - bl[2]];
  return sx * sy;
}

// exclusive prefix sum of v over the workgroup (CPL_THREADS entries of s);
// returns the total
__device__ int block_scan(int* s, int v) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int d = 1; d < CPL_THREADS; d <<= 1) {
    const int x = t >= d ? s[t - d] : 0;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  const int total = s[CPL_THREADS - 1];
  const int incl = s[t];
  __syncthreads();
  s[t] = incl - v;
  __syncthreads();
  return total;
}

// Dynamic LDS of the coupling pass: G, then [I + O G | O] as two K x K blocks.
size_t geo_couple_lds_bytes() { return (size_t)3 * CPL_KT * CPL_KT * sizeof(double); }

__global__ __launch_bounds__(CPL_THREADS) void k_geo_couple(IbmParams p, GeoBuffers G, DevStep* st, int slot0);

// Static plus dynamic LDS of a coupling workgroup (0 if the runtime cannot say).
size_t geo_couple_lds_total() {
  hipFuncAttributes a{};
  if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_geo_couple)) != hipSuccess) return 0;
  return a.sharedSizeBytes + geo_couple_lds_bytes();
}

// One workgroup per step, after the per-wing solves of the step.
__device__ __forceinline__ void geo_couple_body(const IbmParams& p, const GeoBuffers& G, DevStep* st, int slot) {
  extern __shared__ double lds_c[];
  double* sG = lds_c;                           // G (K x K, block-diagonal)
  double* sL = sG + CPL_KT * CPL_KT;            // I + O G, row stride K
  double* sR = sL + CPL_KT * CPL_KT;            // O, then M
  __shared__ int s_flag[MAXW * MAXW];
  __shared__ int s_cnt[MAXW], s_skip, s_piv;
  __shared__ int s_kw[MAXW][CPL_K];
  __shared__ int s_off[MAXW + 1];
  __shared__ int s_k[CPL_KT], s_w[CPL_KT], s_pos[CPL_KT];
  __shared__ double s_f[CPL_KT];
  __shared__ int s_scan[CPL_THREADS];
  __shared__ int s_seg[CPL_MAX_SPW + 1];
  const int t = threadIdx.x;
  const int W = p.W, n = p.SPW;
  GeoStep* gs = G.step + slot;
  for (int q = t; q < MAXW * MAXW; q += CPL_THREADS) s_flag[q] = 0;
  if (t < MAXW) s_cnt[t] = 0;
  if (t == 0) {
    gs->cpl_on = 0;
    gs->cpl_n = 0;
    s_skip = load_agent(&st->error) != 0;
  }
  __syncthreads();
  if (s_skip) return;
  // pairs of wings whose claim maps share a cell
  for (int w = 0; w < W; w++)
    for (int w2 = w + 1; w2 < W; w2++) {
      const int* r1 = gs->rect[w];
      const int* r2 = gs->rect[w2];
      const int ia = max(r1[0], r2[0]), ib = min(r1[1], r2[1]);
      const int ja = max(r1[2], r2[2]), jb = min(r1[3], r2[3]);
      if (ia > ib || ja > jb) continue;
      const int* M1 = G.map + ((size_t)slot * W + w) * (size_t)p.map_cap;
      const int* M2 = G.map + ((size_t)slot * W + w2) * (size_t)p.map_cap;
      const int w1w = r1[1] - r1[0] + 1, w2w = r2[1] - r2[0] + 1;
      const int ow = ib - ia + 1, oc = ow * (jb - ja + 1);
      for (int c = t; c < oc; c += CPL_THREADS) {
        const int i = ia + c % ow, j = ja + c / ow;
        if (M1[(i - r1[0]) + w1w * (j - r1[2])] != INT_MAX && M2[(i - r2[0]) + w2w * (j - r2[2])] != INT_MAX)
          s_flag[w * MAXW + w2] = 1;
      }
    }
  __syncthreads();
  bool any = false;
  for (int q = 0; q < MAXW * MAXW; q++) any = any || s_flag[q] != 0;
  if (!any) return;
  if (n > CPL_MAX_SPW) {
    if (t == 0) raise_error(st, DERR_WINGS_COUPLED, n);
    return;
  }
  // the coupling segments of each wing: boxes that meet a box of a wing it shares nodes with
  for (int w = 0; w < W; w++) {
    const size_t s0 = (size_t)slot * p.nseg + (size_t)w * n;
    for (int k = t; k < n; k += CPL_THREADS) {
      const int* bk = G.box + (s0 + k) * 4;
      bool meet = false;
      for (int w2 = 0; w2 < W && !meet; w2++) {
        if (w2 == w || !s_flag[min(w, w2) * MAXW + max(w, w2)] || !boxes_meet(bk, gs->rect[w2])) continue;
        const int* ob = G.box + ((size_t)slot * p.nseg + (size_t)w2 * n) * 4;
        for (int l = 0; l < n && !meet; l++) meet = boxes_meet(bk, ob + 4 * l);
      }
      if (meet) {
        const int idx = atomicAdd(&s_cnt[w], 1);
        if (idx < CPL_K) s_kw[w][idx] = k;
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    int run = 0, over = 0;
    for (int w = 0; w < W; w++) {
      s_off[w] = run;
      if (s_cnt[w] > CPL_K) over = max(over, s_cnt[w]);
      run += min(s_cnt[w], CPL_K);
    }
    s_off[W] = run;
    if (over || run > CPL_KT) s_piv = -2 - max(over, run);
    else s_piv = 0;
  }
  __syncthreads();
  if (s_piv < 0) {
    if (t == 0) raise_error(st, DERR_WINGS_COUPLED, 2000 - s_piv - 2);
    return;
  }
  if (t < W) {  // ascending segment order per wing, then the flat list
    int* ks = s_kw[t];
    const int m = s_cnt[t];
    for (int x = 1; x < m; x++) {
      const int v = ks[x];
      int y = x - 1;
      while (y >= 0 && ks[y] > v) {
        ks[y + 1] = ks[y];
        y--;
      }
      ks[y + 1] = v;
    }
    for (int x = 0; x < m; x++) {
      s_k[s_off[t] + x] = ks[x];
      s_w[s_off[t] + x] = t;
      s_pos[s_off[t] + x] = pos_of_seg(ks[x], n);
    }
  }
  __syncthreads();
  const int K = s_off[W];
  // G (from the per-wing inverses) and O (cross-wing entries of S)
  for (int e = t; e < K * K; e += CPL_THREADS) {
    const int i = e / K, j = e - i * K;
    const int wi = s_w[i], wj = s_w[j];
    double gv = 0.0, ov = 0.0;
    if (wi == wj) {
      gv = G.sinv[((size_t)slot * W + wi) * (size_t)n * n + (size_t)s_pos[i] * n + s_pos[j]];
    } else {
      const size_t si = (size_t)slot * p.nseg + (size_t)wi * n + s_k[i];
      const size_t sj = (size_t)slot * p.nseg + (size_t)wj * n + s_k[j];
      ov = s_cross(G.box + si * 4, G.phx + si * BOXW, G.phy + si * BOXW, G.box + sj * 4, G.phx + sj * BOXW,
                   G.phy + sj * BOXW);
    }
    sG[e] = gv;
    sR[e] = ov;
  }
  __syncthreads();
  for (int e = t; e < K * K; e += CPL_THREADS) {   // L = I + O G
    const int i = e / K, j = e - i * K;
    double acc = i == j ? 1.0 : 0.0;
    for (int m = 0; m < K; m++) acc += sR[i * K + m] * sG[m * K + j];
    sL[e] = acc;
  }
  __syncthreads();
  // Gauss-Jordan with partial pivoting on [L | R]: R becomes M = L^-1 O
  for (int c = 0; c < K; c++) {
    if (t == 0) {
      int r = c;
      for (int x = c + 1; x < K; x++)
        if (fabs(sL[x * K + c]) > fabs(sL[r * K + c])) r = x;
      s_piv = sL[r * K + c] != 0.0 ? r : -1;
    }
    __syncthreads();
    const int r = s_piv;
    if (r < 0) {
      if (t == 0) raise_error(st, DERR_WINGS_COUPLED, 3000 + c);
      return;
    }
    if (r != c)
      for (int e = t; e < 2 * K; e += CPL_THREADS) {
        double* A = e < K ? sL : sR;
        const int cc = e < K ? e : e - K;
        const double x = A[c * K + cc];
        A[c * K + cc] = A[r * K + cc];
        A[r * K + cc] = x;
      }
    __syncthreads();
    const double inv = 1.0 / sL[c * K + c];
    __syncthreads();
    for (int e = t; e < 2 * K; e += CPL_THREADS) {
      if (e < K) sL[c * K + e] *= inv;
      else sR[c * K + e - K] *= inv;
    }
    if (t < K) s_f[t] = t == c ? 0.0 : sL[t * K + c];
    __syncthreads();
    for (int e = t; e < 2 * K * K; e += CPL_THREADS) {
      const int rr = e / (2 * K), cc = e - rr * 2 * K;
      if (rr == c) continue;
      if (cc < K) sL[rr * K + cc] -= s_f[rr] * sL[c * K + cc];
      else sR[rr * K + cc - K] -= s_f[rr] * sR[c * K + cc - K];
    }
    __syncthreads();
  }
  double* Mo = G.cpl + (size_t)slot * CPL_KT * CPL_KT;
  for (int e = t; e < K * K; e += CPL_THREADS) Mo[e] = sR[e];
  // each wing's adjoint list without the nodes a lower wing owns, in order;
  // new segment offsets
  for (int b = 1; b < W; b++) {
    bool lower = false;
    for (int a = 0; a < b; a++) lower = lower || s_flag[a * MAXW + b];
    if (!lower) continue;
    for (int k = t; k <= n; k += CPL_THREADS) s_seg[k] = 0;
    __syncthreads();
    const size_t lb = (size_t)slot * p.cap + (size_t)b * p.capw;
    int* aij = G.adj_ij + lb * 2;
    int* aseg = G.adj_seg + lb;
    const int cntb = gs->cnt[b];
    int kept = 0;
    for (int c0 = 0; c0 < cntb; c0 += CPL_THREADS) {
      const int e = c0 + t;
      int keep = 0, i = 0, j = 0, k = 0;
      if (e < cntb) {
        i = aij[2 * e];
        j = aij[2 * e + 1];
        k = aseg[e];
        bool owned = false;
        for (int a = 0; a < b && !owned; a++) {
          if (!s_flag[a * MAXW + b]) continue;
          const int* ra = gs->rect[a];
          const int* MA = G.map + ((size_t)slot * W + a) * (size_t)p.map_cap;
          owned = i >= ra[0] && i <= ra[1] && j >= ra[2] && j <= ra[3] &&
                  MA[(i - ra[0]) + (ra[1] - ra[0] + 1) * (j - ra[2])] != INT_MAX;
        }
        keep = !owned;
      }
      const int total = block_scan(s_scan, keep);
      if (keep) {
        const int dst = kept + s_scan[t];
        aij[2 * dst] = i;
        aij[2 * dst + 1] = j;
        aseg[dst] = k;
        atomicAdd(&s_seg[k], 1);
      }
      kept += total;
      __syncthreads();
    }
    int* soff = G.seg_off + ((size_t)slot * W + b) * (n + 1);
    if (t == 0) {
      int run = 0;
      for (int k = 0; k < n; k++) {
        const int c = s_seg[k];
        soff[k] = run;
        run += c;
      }
      soff[n] = run;
      gs->adj_n -= cntb - kept;
      gs->cnt[b] = kept;
    }
    __syncthreads();
  }
  if (t == 0) {
    gs->cpl_n = K;
    for (int w = 0; w <= W; w++) gs->cpl_off[w] = s_off[w];
    for (int x = 0; x < K; x++) {
      gs->cpl_k[x] = s_k[x];
      gs->cpl_w[x] = s_w[x];
    }
    gs->cpl_on = 1;
  }
}

__global__ __launch_bounds__(CPL_THREADS) void k_geo_couple(IbmParams p, GeoBuffers G, DevStep* st, int slot0) {
  geo_couple_body(p, G, st, slot0 + blockIdx.x);
}

// The coupling correction of the flow pass (one workgroup), after the
// per-wing mat-vecs: F.yvec holds u = blockdiag(S_ww)^-1 c on entry and the
// solution of the coupled system on exit (both components):
// z = u[K], r = M z, y_w = u_w - S_ww^-1[:, K_w] r_w.
__device__ __forceinline__ void ibm_couple_body(const IbmParams& p, const GeoBuffers& G, int slot,
                                                const FlowBuffers& F) {
  const GeoStep* gs = G.step + slot;
  if (!gs->cpl_on) return;
  __shared__ int s_pos[CPL_KT], s_w[CPL_KT];
  __shared__ double s_z[2 * CPL_KT], s_r[2 * CPL_KT];
  const int t = threadIdx.x, n = p.SPW, K = gs->cpl_n;
  const double* Mm = G.cpl + (size_t)slot * CPL_KT * CPL_KT;
  if (t < K) {
    s_pos[t] = pos_of_seg(gs->cpl_k[t], n);
    s_w[t] = gs->cpl_w[t];
  }
  __syncthreads();
  if (t < 2 * K) {
    const int i = t >> 1, c = t & 1;
    s_z[t] = F.yvec[2 * ((size_t)s_w[i] * n + s_pos[i]) + c];
  }
  __syncthreads();
  if (t < 2 * K) {
    const int i = t >> 1, c = t & 1;
    double acc = 0.0;
    for (int j = 0; j < K; j++) acc += Mm[i * K + j] * s_z[2 * j + c];
    s_r[t] = acc;
  }
  __syncthreads();
  for (int w = 0; w < p.W; w++) {
    const int k0 = gs->cpl_off[w], k1 = gs->cpl_off[w + 1];
    if (k0 == k1) continue;
    const double* SW = G.sinv + ((size_t)slot * p.W + w) * (size_t)n * n;
    double* yw = F.yvec + 2 * (size_t)w * n;
    for (int e = t; e < 2 * n; e += CPL_THREADS) {
      const int r = e >> 1, c = e & 1;
      const double* row = SW + (size_t)r * n;
      double acc = yw[e];
      for (int x = k0; x < k1; x++) acc -= row[s_pos[x]] * s_r[2 * x + c];
      yw[e] = acc;
    }
  }
}

__global__ __launch_bounds__(CPL_THREADS) void k_ibm_couple(IbmParams p, GeoBuffers G, int slot, FlowBuffers F) {
  ibm_couple_body(p, G, slot, F);
}

void launch_geometry(hipStream_t s, const IbmParams& p, const GeoBuffers& G, DevStep* st, int slot0, int n0,
                     int fidx0, int count, bool with_inverse, bool direct, bool apart) {
  if (p.W <= 0 || count <= 0) return;
  dim3 grid(count, p.W);
  hipLaunchKernelGGL(k_geo_wing, grid, dim3(GEO_THREADS), geo_wing_lds_bytes(p.SPW), s, p, G, st, slot0, n0, fidx0);
  if (with_inverse) {
    hipLaunchKernelGGL(k_geo_solve, grid, dim3(GEO_THREADS), geo_solve_lds_bytes(p, direct), s, p, G, st, slot0,
                       direct ? 1 : 0, direct && apart ? 1 : 0);
    if (direct && apart) return;   // no step of the launch forms the inverse or couples wings
    const int ch = inverse_chunks(p.SPW);
    hipLaunchKernelGGL(k_geo_inverse, dim3(count * ch, p.W), dim3(INV_THREADS), geo_factor_lds_bytes(p), s, p, G,
                       st, slot0, ch);
    if (p.W >= 2)
      hipLaunchKernelGGL(k_geo_couple, dim3(count), dim3(CPL_THREADS), geo_couple_lds_bytes(), s, p, G, st, slot0);
  }
}

// ------------------------------------------------------------- flow part

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

template <typename S>
__global__ __launch_bounds__(256) void k_ibm_rhs(IbmParams p, MrtParams P, GeoBuffers G, int slot, FlowBuffers F,
                                                 Grid g, const S* __restrict__ src) {
  ibm_rhs_body<S>(p, P, G, slot, F, g, src, blockIdx.x);
}

// y = S^-1 c, one wave per row of the inverse.
__device__ __forceinline__ void ibm_matvec_body(const IbmParams& p, const GeoBuffers& G, int slot,
                                                const FlowBuffers& F, int blk) {
  const int gp = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (gp >= p.nseg || !G.step[slot].inv_on) return;   // no inverse: the band substitution solves
  const int n = p.SPW, w = gp / n, r = gp - w * n;
  const double* row = G.sinv + (((size_t)slot * p.W + w) * n + r) * (size_t)n;
  const double* c = F.cvec + (size_t)w * n * 2;
  double a0 = 0.0, a1 = 0.0;
#pragma unroll 4
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

// y = S^-1 c, one wave per row of the inverse, on steps that formed it.
__global__ __launch_bounds__(256) void k_ibm_matvec(IbmParams p, GeoBuffers G, int slot, FlowBuffers F) {
  if (!G.step[slot].inv_on) return;
  ibm_matvec_body(p, G, slot, F, blockIdx.x);
}

// y = S^-1 c by band substitution (lbm_band.h) on steps without the inverse:
// workgroup w solves wing w; wave c component c (two more waves for the
// bottom ends of twisted factors).
constexpr int BSOLVE_THREADS = 256;   // two components x two ends of a twisted solve

// Twisted substitutions of one wing (four waves: wave & 1 the component,
// wave >> 1 the end): forward from both ends to the middle block, the middle
// block's right-hand sides (each end's products removed, S's own c once),
// its dense solve by the top end's waves, then backward from the middle out.
// band, c and y in LDS (c and y interleaved by component, c and y pointing
// at the wave's component).
template <int BW>
__device__ void twisted_solve(const double* band, int n, int bw, const double* c, double* y) {
  __shared__ double s_cm[2][2][BMAX];   // [end][component][middle row]: c less that end's products
  const int wave = threadIdx.x >> 6, comp = wave & 1, end = wave >> 1, lane = threadIdx.x & 63;
  constexpr int Dd = BW + 1;
  const Twist tw = twist_split(n, bw);
  const int m = tw.m;
  const double* T = band;
  const double* R = T + (size_t)(tw.nT + BW) * Dd;
  const double* iT = R + (size_t)(tw.nR + BW) * Dd;
  const double* iR = iT + tw.nT + 1;
  const double* mid = iR + tw.nR + 1;
  const double* midd = mid + m * m;
  const double* crev = c + 2L * (n - 1);   // the rows of R are those of S in reverse order
  double* yrev = y + 2L * (n - 1);
  // after the forward pass lane l holds middle row l - 1 of its end's order
  const double f = end == 0 ? band_forward_wave<BW, 2>(T, tw.nT, tw.p, c, y)
                            : band_forward_wave<BW, -2>(R, tw.nR, tw.q, crev, yrev);
  if (lane >= 1 && lane <= m) s_cm[end][comp][lane - 1] = f;
  __syncthreads();
  if (end == 0) {
    // lane i: middle row p + i (R holds it as its row q + m - 1 - i)
    double v = lane < m ? s_cm[0][comp][lane] + s_cm[1][comp][m - 1 - lane] - c[2L * (tw.p + lane)] : 0.0;
    for (int j = 0; j < m; j++) {   // L z = v
      const double zj = lane_value(v, j);
      if (lane > j && lane < m) v = v - mid[lane * m + j] * zj;
    }
    if (lane < m) v = v * midd[lane];
    for (int j = m - 1; j >= 0; j--) {   // L^T x = D^-1 z
      const double xj = lane_value(v, j);
      if (lane < j) v = v - mid[j * m + lane] * xj;
    }
    if (lane < m) y[2L * (tw.p + lane)] = v;
  }
  __syncthreads();   // the middle values in y, for both ends
  if (end == 0)
    band_backward_wave<BW, 2>(T, iT, tw.nT, y);
  else
    band_backward_wave<BW, -2>(R, iR, tw.nR, yrev);
}

// doubles of a wing's factors in bandg (twisted or plain, by the band)
__device__ __forceinline__ size_t factor_doubles(int n, int bw, int BWc, bool twisted) {
  return twisted ? twist_doubles(n, bw, BWc) : (size_t)(n + BWc) * (BWc + 1) + n + 1;
}

// The band substitutions run from LDS: a wing's factors (with BAND_RING rows
// of whatever precedes and follows them, which the passes' look-ahead loads
// read), its right-hand sides and its solution are staged there first, so
// that each step of the dependent chain waits for an LDS load rather than
// for global memory (a global load per step, even issued BAND_RING steps
// ahead, left the chain waiting on memory latency).
__device__ __forceinline__ void ibm_bsolve_body(const IbmParams& p, const GeoBuffers& G, int slot,
                                                const FlowBuffers& F, int w) {
  extern __shared__ double lds_b[];
  const GeoStep* gs = G.step + slot;
  const int t = threadIdx.x, wave = t >> 6, comp = wave & 1;
  if (gs->inv_on) return;
  const int n = p.SPW, bw = gs->bw[w];
  double* yg = F.yvec + 2 * (size_t)w * n;
  if (gs->cnt[w] == 0 || bw > BMAX || band_class(bw) > p.band_cls) {
    // no factors (the wing claims no node, or the solve raised an error)
    for (int e = t; e < 2 * n; e += BSOLVE_THREADS) yg[e] = 0.0;
    return;
  }
  const int BWc = band_class(bw);
  const double* gband = G.bandg + ((size_t)slot * p.W + w) * band_cap(p);
  const int pad = BAND_RING * (BWc + 1);
  const bool twisted = uses_twist(p, bw);
  const int nb = (int)factor_doubles(n, bw, BWc, twisted) + 2 * pad;
  double* band = lds_b + pad;
  double* c = lds_b + nb;
  double* y = c + 2 * n;
  copy_batched<BSOLVE_THREADS>(lds_b, gband - pad, nb);
  copy_batched<BSOLVE_THREADS>(c, F.cvec + 2 * (size_t)w * n, 2 * n);
  __syncthreads();
  if (twisted) {   // the geometry solve left twisted factors (a coupled step that forms no inverse)
    if (BWc == 8)
      twisted_solve<8>(band, n, bw, c + comp, y + comp);
    else if (BWc == 12)
      twisted_solve<12>(band, n, bw, c + comp, y + comp);
    else if (BWc == 16)
      twisted_solve<16>(band, n, bw, c + comp, y + comp);
    else
      twisted_solve<BMAX>(band, n, bw, c + comp, y + comp);
  } else if (wave < 2) {   // one wave per component
    const double* invd = band + (size_t)(n + BWc) * (BWc + 1);
    if (BWc == 8)
      band_solve_wave<8>(band, invd, n, c + comp, y + comp);
    else if (BWc == 12)
      band_solve_wave<12>(band, invd, n, c + comp, y + comp);
    else if (BWc == 16)
      band_solve_wave<16>(band, invd, n, c + comp, y + comp);
    else
      band_solve_wave<BMAX>(band, invd, n, c + comp, y + comp);
  }
  __syncthreads();
  for (int e = t; e < 2 * n; e += BSOLVE_THREADS) yg[e] = y[e];
}

__global__ __launch_bounds__(BSOLVE_THREADS) void k_ibm_bsolve(IbmParams p, GeoBuffers G, int slot,
                                                               FlowBuffers F) {
  ibm_bsolve_body(p, G, slot, F, blockIdx.x);
}

// g = D^T y at the adjoint nodes (boundary_condition.f90:1639-1651).  A
// workgroup takes the nodes claimed by SPREAD_SEG consecutive segments; every
// segment whose box can hold them lies within +-bw band positions of the
// claiming segment, and those positions are staged in LDS.
constexpr int SPREAD_SEG = 16;
constexpr int SPREAD_THREADS = 128;
constexpr int SPREAD_STAGE = 2 * SPREAD_SEG + 2 * BMAX + 2;

__device__ __forceinline__ void ibm_spread_body(const IbmParams& p, const GeoBuffers& G, int slot,
                                                const FlowBuffers& F, int blk) {
  __shared__ int s_box[SPREAD_STAGE][4];
  __shared__ double s_ph[SPREAD_STAGE][2 * BOXW];
  __shared__ double s_y[SPREAD_STAGE][2];
  const int n = p.SPW, t = threadIdx.x;
  const int nb = (n + SPREAD_SEG - 1) / SPREAD_SEG;
  const int w = blk / nb, sb = blk - w * nb;
  const int k0 = sb * SPREAD_SEG, k1 = min(k0 + SPREAD_SEG, n);
  const GeoStep* gs = G.step + slot;
  const int bw = min(gs->bw[w], BMAX);
  int plo = INT_MAX, phi = -1;
  for (int k = k0; k < k1; k++) {
    const int q = pos_of_seg(k, n);
    plo = min(plo, q);
    phi = max(phi, q);
  }
  plo = max(0, plo - bw);
  phi = min(n - 1, phi + bw);
  const int ns = min(phi - plo + 1, SPREAD_STAGE);
  const size_t base = (size_t)slot * p.nseg + (size_t)w * n;
  for (int e = t; e < ns; e += SPREAD_THREADS) {
    const int q = plo + e, k = seg_of_pos(q, n);
    const int* bx = G.box + (base + k) * 4;
#pragma unroll
    for (int c = 0; c < 4; c++) s_box[e][c] = bx[c];
#pragma unroll
    for (int c = 0; c < BOXW; c++) {
      s_ph[e][c] = G.phx[(base + k) * BOXW + c];
      s_ph[e][BOXW + c] = G.phy[(base + k) * BOXW + c];
    }
    s_y[e][0] = F.yvec[2 * ((size_t)w * n + q)];
    s_y[e][1] = F.yvec[2 * ((size_t)w * n + q) + 1];
  }
  __syncthreads();
  const int* soff = G.seg_off + ((size_t)slot * p.W + w) * (n + 1);
  const int a0 = soff[k0], a1 = soff[k1];
  const int cpl0 = gs->cpl_on ? gs->cpl_off[w + 1] : 0, cpl1 = gs->cpl_on ? gs->cpl_n : 0;
  const size_t lb = (size_t)slot * p.cap + (size_t)w * p.capw;
  const int* aij = G.adj_ij + lb * 2;
  const int* aseg = G.adj_seg + lb;
  double* gout = F.adj_g + ((size_t)w * p.capw) * 2;
  for (int a = a0 + t; a < a1; a += SPREAD_THREADS) {
    const int i = aij[2 * a], j = aij[2 * a + 1];
    const int p0 = pos_of_seg(aseg[a], n);
    const int qa = max(0, p0 - bw), qb = min(n - 1, p0 + bw);
    double gx = 0.0, gy = 0.0;
    for (int q = qa; q <= qb; q++) {
      const int e = q - plo;
      if (i < s_box[e][0] || i > s_box[e][1] || j < s_box[e][2] || j > s_box[e][3]) continue;
      const double phi_ = s_ph[e][i - s_box[e][0]] * s_ph[e][BOXW + j - s_box[e][2]];
      gx = gx + phi_ * s_y[e][0];
      gy = gy + phi_ * s_y[e][1];
    }
    // the coupling segments of higher wings whose boxes hold the node
    for (int x = cpl0; x < cpl1; x++) {
      const int l = gs->cpl_k[x], wb = gs->cpl_w[x];
      const size_t sl = (size_t)slot * p.nseg + (size_t)wb * n + l;
      const int* bl = G.box + sl * 4;
      if (i < bl[0] || i > bl[1] || j < bl[2] || j > bl[3]) continue;
      const double phi_ = G.phx[sl * BOXW + (i - bl[0])] * G.phy[sl * BOXW + (j - bl[2])];
      const size_t yq = (size_t)wb * n + pos_of_seg(l, n);
      gx = gx + phi_ * F.yvec[2 * yq];
      gy = gy + phi_ * F.yvec[2 * yq + 1];
    }
    gout[2 * a] = gx;
    gout[2 * a + 1] = gy;
  }
}

__global__ __launch_bounds__(SPREAD_THREADS) void k_ibm_spread(IbmParams p, GeoBuffers G, int slot, FlowBuffers F) {
  ibm_spread_body(p, G, slot, F, blockIdx.x);
}

void launch_ibm_flow(hipStream_t s, const IbmParams& p, const MrtParams& P, const GeoBuffers& G, int slot,
                     const FlowBuffers& F, const Grid& g, const void* src, bool f32, bool direct, bool apart) {
  if (p.W <= 0) return;
  const int blk = (p.nseg + 3) / 4;
  const bool sub_only = direct && apart;   // the geometry formed no inverse and no wings couple
  if (f32)
    hipLaunchKernelGGL(k_ibm_rhs<float>, dim3(blk), dim3(256), 0, s, p, P, G, slot, F, g, (const float*)src);
  else
    hipLaunchKernelGGL(k_ibm_rhs<double>, dim3(blk), dim3(256), 0, s, p, P, G, slot, F, g, (const double*)src);
  if (!sub_only) hipLaunchKernelGGL(k_ibm_matvec, dim3(blk), dim3(256), 0, s, p, G, slot, F);
  if (direct) hipLaunchKernelGGL(k_ibm_bsolve, dim3(p.W), dim3(BSOLVE_THREADS), bsolve_lds_bytes(p), s, p, G, slot, F);
  if (p.W >= 2 && !sub_only) hipLaunchKernelGGL(k_ibm_couple, dim3(1), dim3(CPL_THREADS), 0, s, p, G, slot, F);
  const int nb = (p.SPW + SPREAD_SEG - 1) / SPREAD_SEG;
  hipLaunchKernelGGL(k_ibm_spread, dim3(p.W * nb), dim3(SPREAD_THREADS), 0, s, p, G, slot, F);
}

// ------------------------------------------------------- batched launches

int spread_blocks(const IbmParams& p) { return p.W * ((p.SPW + SPREAD_SEG - 1) / SPREAD_SEG); }

// The marker pass and the band factor of a batch loop over their (episode,
// step, wing) items from a grid of one workgroup per item or fewer (the
// barrier at the top of each item keeps the previous item's readers of the
// LDS and of the shared flags behind its writers; every return in the
// bodies is taken by the whole workgroup).
__global__ __launch_bounds__(GEO_THREADS_BATCH) void kb_geo_wing(const EpDesc* __restrict__ eps,
                                                           const int2* __restrict__ items,
                                                           const int* __restrict__ kb, int kk0, int n_items) {
  for (int q = blockIdx.x; q < n_items; q += gridDim.x) {
    __syncthreads();
    const int2 it = items[q];
    const EpDesc& D = eps[it.x];
    const int k0 = kb[0] + kk0;
    const int b = it.y / D.p.W, w = it.y - b * D.p.W;
    geo_wing_body<GEO_THREADS_BATCH>(D.p, D.G, D.st, ep_slot(D, k0), D.n0 + k0, k0, b, w);
  }
}

__global__ __launch_bounds__(GEO_THREADS_BATCH) void kb_geo_solve(const EpDesc* __restrict__ eps,
                                                            const int2* __restrict__ items,
                                                            const int* __restrict__ kb, int kk0, int n_items) {
  for (int q = blockIdx.x; q < n_items; q += gridDim.x) {
    __syncthreads();
    const int2 it = items[q];
    const EpDesc& D = eps[it.x];
    const int k0 = kb[0] + kk0;
    const int b = it.y / D.p.W, w = it.y - b * D.p.W;
    geo_solve_body<GEO_THREADS_BATCH>(D.p, D.G, D.st, ep_slot(D, k0), b, w, D.direct, 0);
  }
}

// inverse columns chunk*64 .. of each (step, wing) item (chunks past the
// episode's last column return)
__global__ __launch_bounds__(INV_THREADS) void kb_geo_inverse(const EpDesc* __restrict__ eps,
                                                              const int2* __restrict__ items,
                                                              const int* __restrict__ kb, int kk0) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  const int k0 = kb[0] + kk0;
  const int b = it.y / D.p.W, w = it.y - b * D.p.W;
  geo_inverse_body(D.p, D.G, D.st, ep_slot(D, k0), b, w, blockIdx.y);
}

// coupling pass of step b of each episode with two or more wings: the
// workgroup of its (b, wing 0) item
__global__ __launch_bounds__(CPL_THREADS) void kb_geo_couple(const EpDesc* __restrict__ eps,
                                                             const int2* __restrict__ items,
                                                             const int* __restrict__ kb, int kk0) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  if (D.p.W < 2 || it.y % D.p.W != 0) return;
  const int k0 = kb[0] + kk0;
  geo_couple_body(D.p, D.G, D.st, ep_slot(D, k0) + it.y / D.p.W);
}

void launch_batch_geometry(hipStream_t s, const EpDesc* eps, const int* kb, const int2* items, int n_items,
                           int k0, size_t lds_wing, size_t lds_solve, size_t lds_inv, int inv_chunks, int grid) {
  if (n_items <= 0) return;
  const int g = grid > 0 && grid < n_items ? grid : n_items;
  hipLaunchKernelGGL(kb_geo_wing, dim3(g), dim3(GEO_THREADS_BATCH), lds_wing, s, eps, items, kb, k0, n_items);
  hipLaunchKernelGGL(kb_geo_solve, dim3(g), dim3(GEO_THREADS_BATCH), lds_solve, s, eps, items, kb, k0, n_items);
  hipLaunchKernelGGL(kb_geo_inverse, dim3(n_items, inv_chunks), dim3(INV_THREADS), lds_inv, s, eps, items, kb, k0);
  hipLaunchKernelGGL(kb_geo_couple, dim3(n_items), dim3(CPL_THREADS), geo_couple_lds_bytes(), s, eps, items, kb, k0);
}

template <typename S>
__global__ __launch_bounds__(256) void kb_ibm_rhs(const EpDesc* __restrict__ eps, const int2* __restrict__ items,
                                                  const int* __restrict__ kb, int kk) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  const int k = kb[0] + kk;
  const int cur = (D.cur0 + k) & 1;
  ibm_rhs_body<S>(D.p, D.P, D.G, ep_slot(D, k), D.F, D.g, (const S*)D.fbuf[cur], it.y);
}

__global__ __launch_bounds__(256) void kb_ibm_matvec(const EpDesc* __restrict__ eps, const int2* __restrict__ items,
                                                     const int* __restrict__ kb, int kk) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  const int k = kb[0] + kk;
  ibm_matvec_body(D.p, D.G, ep_slot(D, k), D.F, it.y);
}

__global__ __launch_bounds__(SPREAD_THREADS) void kb_ibm_spread(const EpDesc* __restrict__ eps,
                                                                const int2* __restrict__ items,
                                                                const int* __restrict__ kb, int kk) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  const int k = kb[0] + kk;
  ibm_spread_body(D.p, D.G, ep_slot(D, k), D.F, it.y);
}

// coupling correction of each episode with two or more wings: the
// workgroup of its first segment item
__global__ __launch_bounds__(CPL_THREADS) void kb_ibm_couple(const EpDesc* __restrict__ eps,
                                                             const int2* __restrict__ items,
                                                             const int* __restrict__ kb, int kk) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  if (it.y != 0 || D.p.W < 2) return;
  const int k = kb[0] + kk;
  ibm_couple_body(D.p, D.G, ep_slot(D, k), D.F);
}

// y = S^-1 c of every (episode, wing) of the wing list by band substitution,
// on the steps whose geometry formed no inverse
__global__ __launch_bounds__(BSOLVE_THREADS) void kb_ibm_bsolve(const EpDesc* __restrict__ eps,
                                                                const int2* __restrict__ items,
                                                                const int* __restrict__ kb, int kk) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  const int k = kb[0] + kk;
  ibm_bsolve_body(D.p, D.G, ep_slot(D, k), D.F, it.y);
}

void launch_batch_ibm_flow(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32,
                           size_t bsolve_lds) {
  if (it.n_seg <= 0) return;
  if (f32)
    hipLaunchKernelGGL(kb_ibm_rhs<float>, dim3(it.n_seg), dim3(256), 0, s, eps, it.seg, it.kbase, k);
  else
    hipLaunchKernelGGL(kb_ibm_rhs<double>, dim3(it.n_seg), dim3(256), 0, s, eps, it.seg, it.kbase, k);
  hipLaunchKernelGGL(kb_ibm_matvec, dim3(it.n_seg), dim3(256), 0, s, eps, it.seg, it.kbase, k);
  if (bsolve_lds > 0 && it.n_wing > 0)
    hipLaunchKernelGGL(kb_ibm_bsolve, dim3(it.n_wing), dim3(BSOLVE_THREADS), bsolve_lds, s, eps, it.wing, it.kbase,
                       k);
  hipLaunchKernelGGL(kb_ibm_couple, dim3(it.n_seg), dim3(CPL_THREADS), 0, s, eps, it.seg, it.kbase, k);
  hipLaunchKernelGGL(kb_ibm_spread, dim3(it.n_spread), dim3(SPREAD_THREADS), 0, s, eps, it.spread, it.kbase, k);
}

}  // namespace lbm
