// lbm_lattice.hip - full-lattice kernels of the time step on gfx950.
//
// k_lattice fuses streaming (solution.f90:440-466), physical_variables
// (solution.f90:483-554) and colission (solution.f90:585-655) for every
// interior node: 9 pulled loads and 9 stores of one SoA plane each, i.e. the
// 144 B/node floor of a two-buffer FP64 D2Q9 step (72 B/node with FP32
// deviation storage).  The nodes that the immersed boundary touches this step
// (the adjoint nodes) are skipped: k_adj_collide collides them with the IBM
// correction, on the IBM stream while the full-lattice pass runs, so neither
// pass waits for the other.
//
// Every kernel that touches populations is a template on the storage type S:
// double (FP64 values) or float (FP32 deviations d = f - w_k).  Every
// kernel computes in FP64: the bulk collision of the FP32 storage on the
// loaded deviations, the boundary ring, the adjoint nodes and all reductions
// on loaded f = w_k + d.
#include <algorithm>
#include <mutex>

#include <hip/hip_ext.h>

#include "lbm_internal.h"

namespace lbm {

// Cache policy of the population loads and stores (CP): each population of
// a step is read once and written once, so the accesses may be marked
// non-temporal.  CP_PLAIN: ordinary accesses; CP_NT_STORE: non-temporal
// stores; CP_NT: non-temporal loads and stores.  Which is fastest depends on
// the lattice and the wave cap, so the autotuner times them.
template <int CP, typename T>
__device__ __forceinline__ T ld_pop(const T* __restrict__ p) {
  if constexpr (CP == CP_NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <int CP, typename T>
__device__ __forceinline__ void st_pop(T* __restrict__ p, T v) {
  if constexpr (CP == CP_NT_STORE || CP == CP_NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// The FP64 step of interior node (i, j) (the ring columns i = 0 and nx-1 and
// the step's adjoint nodes excepted: k_bc_ring and k_adj_collide write those).
template <bool FIELDS, int CP>
__device__ __forceinline__ void lattice_node(const Grid& g, const MrtParams& P, const double* __restrict__ src,
                                             double* __restrict__ dst, const FieldSlot& fs, int i, int j,
                                             const AdjMask& am) {
  if (i < 1 || i > g.nx - 2) return;
  double f[9];
#pragma unroll
  for (int k = 0; k < 9; k++) f[k] = ld_pop<CP>(src + (long long)k * g.ps + (long long)(j - EY(k)) * g.nx + (i - EX(k)));
  // after the loads, so that they do not wait for the mask's own loads
  if (is_adjoint(am, i, j)) return;
  double rho, u, v;
  node_moments(f, rho, u, v);
  double out[9];
  collide_node(P, f, rho, u, v, 0.0, 0.0, out);
  const long long o = (long long)j * g.nx + i;
#pragma unroll
  for (int k = 0; k < 9; k++) st_pop<CP>(dst + (long long)k * g.ps + o, out[k]);
  if (FIELDS) {
    if (fs.rho) fs.rho[o] = rho;   // null: velocities only (erranalysis)
    fs.u[o] = u;
    fs.v[o] = v;
  }
}

// One thread per node of a row tile of BX nodes; tiles start at i = 0 so
// that every wavefront stores whole, aligned spans of each population plane
// (the ring columns i = 0 and i = nx-1 are skipped and left to k_bc_ring).
// Workgroups follow the hardware's round-robin placement over the XCDs, which
// measured faster than remapping tiles to XCD-contiguous runs.  The tile
// width is chosen per instance by timing (tune_instance).
template <bool FIELDS, int BX, int CP>
__global__ __launch_bounds__(BX) void k_lattice(Grid g, MrtParams P, const double* __restrict__ src,
                                                double* __restrict__ dst, FieldSlot fs, int tiles_x, AdjMask am) {
  const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - ty * tiles_x;
  lattice_node<FIELDS, CP>(g, P, src, dst, fs, tx * BX + (int)threadIdx.x, 1 + ty, am);
}

// Node of lane t in tile `blk` of BX consecutive nodes of the rows j = 1..ny-2
// taken as one run (row-major, rows back to back): the batch covers each
// lattice's interior rows with whole tiles and no partial tile at each row
// end.  The run starts at node offset flat_base, the start of row 1 rounded
// down to a multiple of 32 nodes, so that every tile and wavefront (BX a
// multiple of 64) begins on a 128-byte line of the unshifted loads and of
// every store whatever the lattice's width (the plane stride is a multiple of
// 64 nodes); the lanes before row 1 get j = 0.  (Started at row 1
// itself, 20 wingless 2040x1020 episodes ran their step at 0.51 of 8 TB/s,
// 2016x1008 ones at 0.61.)  A tile wraps into the next row at most once when BX <= nx
// (every lattice of the batched workloads), more often on narrower lattices;
// a lane past the last interior row gets j > ny - 2.
__host__ __device__ __forceinline__ int flat_base(const Grid& g) { return g.nx & ~31; }

template <int BX>
__device__ __forceinline__ void flat_node(const Grid& g, int blk, int t, int& i, int& j) {
  const int q0 = flat_base(g) + blk * BX;  // node offset of the tile's first lane
  j = q0 / g.nx;                           // uniform over the tile
  i = q0 - j * g.nx + t;
  while (i >= g.nx) {
    i -= g.nx;
    j += 1;
  }
}

// Bulk step of the FP32-shifted storage: deviations d = f - w from the rest
// equilibrium, loaded as FP32 and collided in FP64, so that the only rounding
// a step adds beyond FP64's is that of the nine FP32 stores.  (Colliding in
// FP32 as well was 7 % faster but its rounding is biased: at 4001x2001, Re
// 1000, the window-averaged forces after 4000 steps moved by up to 8.5e-3 of
// a value, against 9e-5 with the FP64 collision.)  The moments m' = M d come
// from shared partial sums of the nine loads (M of solution.f90:17-25); the
// weights carry no momentum, so j = sum e d, and with the incompressible
// equilibrium (solution.f90:613-624) rho u = j, so the equilibrium moments
// need no division: e_eq = -2 rho + 3 j^2, eps_eq = rho - 3 j^2, q_eq = -j,
// p_eq = (jx^2 - jy^2, jx jy), each shifted by (M w) (rho = sum w + sum d;
// MrtParams eq1, eq2).  Relaxation and the inverse transform are linear: with
// M^-1 = M^T diag(c), c = (1/9, 1/36, 1/36, 1/6, 1/12, 1/6, 1/12, 1/4, 1/4),
// the scaled relaxed moments are tt_r = c_r (1 - s_r) m_r + c_r s_r m_eq,r
// (MrtParams ra, rb, rk0) and d* = M^T tt, again from shared sums.  About 90
// FP64 operations per node.
template <bool FIELDS, int CP>
__device__ __forceinline__ void lattice_f32_node(const Grid& g, const MrtParams& P, const float* __restrict__ src,
                                                 float* __restrict__ dst, const FieldSlot& fs, int i, int j,
                                                 const AdjMask& am) {
  if (i < 1 || i > g.nx - 2) return;
  double d[9];
#pragma unroll
  for (int k = 0; k < 9; k++)
    d[k] = (double)ld_pop<CP>(src + (long long)k * g.ps + (long long)(j - EY(k)) * g.nx + (i - EX(k)));
  if (is_adjoint(am, i, j)) return;
  // moments of d
  const double h13 = d[1] + d[3], h24 = d[2] + d[4];     // axis pairs
  const double h57 = d[5] + d[7], h68 = d[6] + d[8];     // diagonal pairs
  const double s14 = h13 + h24, s58 = h57 + h68;
  const double rp = d[0] + s14 + s58;                    // m0 = rho - sum w
  const double a13 = d[1] - d[3], a24 = d[2] - d[4];
  const double x58 = (d[5] + d[8]) - (d[6] + d[7]);
  const double y58 = (d[5] + d[6]) - (d[7] + d[8]);
  const double jx = a13 + x58, jy = a24 + y58;
  const double m1 = (s58 + s58) - s14 - 4.0 * d[0];
  const double m2 = 4.0 * d[0] - (s14 + s14) + s58;
  const double m4 = x58 - (a13 + a13);
  const double m6 = y58 - (a24 + a24);
  const double m7 = h13 - h24;
  const double m8 = h57 - h68;
  // relaxed moments, scaled by c
  const double xx = jx * jx, yy = jy * jy, jj = xx + yy;
  const double* A = P.ra;
  const double* B = P.rb;
  const double t0 = fma(A[0], rp, P.rk0);
  const double t1 = fma(A[1], m1, B[1] * (3.0 * jj - (rp + rp) + P.eq1));
  const double t2 = fma(A[2], m2, B[2] * (rp - 3.0 * jj + P.eq2));
  const double t3 = A[3] * jx;
  const double t4 = fma(A[4], m4, -B[4] * jx);
  const double t5 = A[5] * jy;
  const double t6 = fma(A[6], m6, -B[6] * jy);
  const double t7 = fma(A[7], m7, B[7] * (xx - yy));
  const double t8 = fma(A[8], m8, B[8] * (jx * jy));
  // d* = M^T tt
  const double pa = t0 - t1 - (t2 + t2);                 // axis directions
  const double qd = t0 + (t1 + t1) + t2;                 // diagonals
  const double ex = t3 - (t4 + t4), ey = t5 - (t6 + t6);
  const double dx = t3 + t4, dy = t5 + t6;
  float out[9];   // the stored deviations
  out[1] = (float)(pa + ex + t7);
  out[3] = (float)(pa - ex + t7);
  out[2] = (float)(pa + ey - t7);
  out[4] = (float)(pa - ey - t7);
  out[5] = (float)(qd + (dx + dy) + t8);
  out[7] = (float)(qd - (dx + dy) + t8);
  out[6] = (float)(qd + (dy - dx) - t8);
  out[8] = (float)(qd - (dy - dx) - t8);
  // with s_0 = 0 (rk0 = 0: the collision conserves the density moment, the
  // reference's rates) the rest population is the remainder of the mass,
  // t0 + 4 (t2 - t1) in exact arithmetic, taken against the eight rounded
  // deviations: the nine stored values then sum to rp up to the one rounding
  // of out[0], so the FP32 stores add no systematic mass drift
  const double o13 = (double)out[1] + (double)out[3], o24 = (double)out[2] + (double)out[4];
  const double o57 = (double)out[5] + (double)out[7], o68 = (double)out[6] + (double)out[8];
  out[0] = (float)(P.rk0 == 0.0 ? rp - ((o13 + o24) + (o57 + o68)) : t0 + 4.0 * (t2 - t1));
  const long long o = (long long)j * g.nx + i;
#pragma unroll
  for (int k = 0; k < 9; k++) st_pop<CP>(dst + (long long)k * g.ps + o, out[k]);
  if (FIELDS) {
    const double rd = P.wsum + rp;
    if (fs.rho) fs.rho[o] = rd;
    fs.u[o] = jx / rd;
    fs.v[o] = jy / rd;
  }
}

template <bool FIELDS, int BX, int CP>
__global__ __launch_bounds__(BX) void k_lattice_f32(Grid g, MrtParams P, const float* __restrict__ src,
                                                    float* __restrict__ dst, FieldSlot fs, int tiles_x, AdjMask am) {
  const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - ty * tiles_x;
  lattice_f32_node<FIELDS, CP>(g, P, src, dst, fs, tx * BX + (int)threadIdx.x, 1 + ty, am);
}

// The kernels of a coupled step (fields written, every interior node), held
// to two waves per SIMD (8 per CU) by their register count instead of by
// unused LDS: an empty asm statement that clobbers v255 makes the kernel
// declare 256 VGPRs, so no third wave of it fits a SIMD, and the CU's LDS
// stays free for the chain beside it, whose geometry solve and band
// substitution stage 50-100 KB each.  (A grid of 8 waves' worth of
// workgroups per CU striding over the tiles held the same cap with nothing
// reserved, but its long-lived waves starved the chain: the marker pass took
// 277 instead of 85 us, the coupled step 0.50 instead of 0.34-0.36 ms.)
template <int BX, int CP>
__global__ __launch_bounds__(BX) void k_lattice_narrow(Grid g, MrtParams P, const double* __restrict__ src,
                                                       double* __restrict__ dst, FieldSlot fs, int tiles_x,
                                                       AdjMask am) {
  asm volatile("" ::: "v255");
  const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - ty * tiles_x;
  lattice_node<true, CP>(g, P, src, dst, fs, tx * BX + (int)threadIdx.x, 1 + ty, am);
}

template <int BX, int CP>
__global__ __launch_bounds__(BX) void k_lattice_f32_narrow(Grid g, MrtParams P, const float* __restrict__ src,
                                                           float* __restrict__ dst, FieldSlot fs, int tiles_x,
                                                           AdjMask am) {
  asm volatile("" ::: "v255");
  const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - ty * tiles_x;
  lattice_f32_node<true, CP>(g, P, src, dst, fs, tx * BX + (int)threadIdx.x, 1 + ty, am);
}

int lattice_tiles_x(const Grid& g, int bx) { return (g.nx + bx - 1) / bx; }
int lattice_tiles_flat(const Grid& g, int bx) {
  return (int)(((long long)(g.ny - 1) * g.nx - flat_base(g) + bx - 1) / bx);
}

template <int BX, int CP>
void launch_lattice_cp(hipStream_t s, const Grid& g, const MrtParams& P, const void* src, void* dst,
                       const FieldSlot* fields, bool f32, const AdjMask& am, hipEvent_t t0, hipEvent_t t1,
                       size_t lds) {
  const int tiles_x = lattice_tiles_x(g, BX);
  dim3 grid(tiles_x * (g.ny - 2)), block(BX);
  const FieldSlot fs = fields ? *fields : FieldSlot{};
  if (f32) {
    const float* a = (const float*)src;
    float* b = (float*)dst;
    if (fields)
      hipExtLaunchKernelGGL((k_lattice_f32<true, BX, CP>), grid, block, lds, s, t0, t1, 0, g, P, a, b, fs, tiles_x, am);
    else
      hipExtLaunchKernelGGL((k_lattice_f32<false, BX, CP>), grid, block, lds, s, t0, t1, 0, g, P, a, b, fs, tiles_x,
                            am);
  } else {
    const double* a = (const double*)src;
    double* b = (double*)dst;
    if (fields)
      hipExtLaunchKernelGGL((k_lattice<true, BX, CP>), grid, block, lds, s, t0, t1, 0, g, P, a, b, fs, tiles_x, am);
    else
      hipExtLaunchKernelGGL((k_lattice<false, BX, CP>), grid, block, lds, s, t0, t1, 0, g, P, a, b, fs, tiles_x, am);
  }
}

template <int BX, int CP>
void launch_narrow_cp(hipStream_t s, const Grid& g, const MrtParams& P, const void* src, void* dst,
                      const FieldSlot& fs, bool f32, const AdjMask& am) {
  const int tiles_x = lattice_tiles_x(g, BX);
  dim3 grid(tiles_x * (g.ny - 2)), block(BX);
  if (f32)
    hipLaunchKernelGGL((k_lattice_f32_narrow<BX, CP>), grid, block, 0, s, g, P, (const float*)src, (float*)dst, fs,
                       tiles_x, am);
  else
    hipLaunchKernelGGL((k_lattice_narrow<BX, CP>), grid, block, 0, s, g, P, (const double*)src, (double*)dst, fs,
                       tiles_x, am);
}

template <int BX>
void launch_narrow_bx(hipStream_t s, const Grid& g, const MrtParams& P, const void* src, void* dst,
                      const FieldSlot& fs, bool f32, int cp, const AdjMask& am) {
  if (cp == CP_NT)
    launch_narrow_cp<BX, CP_NT>(s, g, P, src, dst, fs, f32, am);
  else if (cp == CP_NT_STORE)
    launch_narrow_cp<BX, CP_NT_STORE>(s, g, P, src, dst, fs, f32, am);
  else
    launch_narrow_cp<BX, CP_PLAIN>(s, g, P, src, dst, fs, f32, am);
}

void launch_lattice_narrow(hipStream_t s, const Grid& g, const MrtParams& P, const void* src, void* dst,
                           const FieldSlot& fields, bool f32, int bx, int cp, const AdjMask& mask) {
  if (bx == 128)
    launch_narrow_bx<128>(s, g, P, src, dst, fields, f32, cp, mask);
  else if (bx == 512)
    launch_narrow_bx<512>(s, g, P, src, dst, fields, f32, cp, mask);
  else
    launch_narrow_bx<256>(s, g, P, src, dst, fields, f32, cp, mask);
}

template <int BX>
void launch_lattice_bx(hipStream_t s, const Grid& g, const MrtParams& P, const void* src, void* dst,
                       const FieldSlot* fields, bool f32, int cp, const AdjMask& am, hipEvent_t t0, hipEvent_t t1,
                       size_t lds) {
  if (cp == CP_NT)
    launch_lattice_cp<BX, CP_NT>(s, g, P, src, dst, fields, f32, am, t0, t1, lds);
  else if (cp == CP_NT_STORE)
    launch_lattice_cp<BX, CP_NT_STORE>(s, g, P, src, dst, fields, f32, am, t0, t1, lds);
  else
    launch_lattice_cp<BX, CP_PLAIN>(s, g, P, src, dst, fields, f32, am, t0, t1, lds);
}

void launch_lattice(hipStream_t s, const Grid& g, const MrtParams& P, const void* src, void* dst, FieldSlot* fields,
                    bool f32, int bx, int cp, const AdjMask& mask, hipEvent_t t0, hipEvent_t t1, size_t lds) {
  if (bx == 128)
    launch_lattice_bx<128>(s, g, P, src, dst, fields, f32, cp, mask, t0, t1, lds);
  else if (bx == 512)
    launch_lattice_bx<512>(s, g, P, src, dst, fields, f32, cp, mask, t0, t1, lds);
  else
    launch_lattice_bx<256>(s, g, P, src, dst, fields, f32, cp, mask, t0, t1, lds);
}

// Re-collide the adjoint nodes with the IBM-corrected velocity u = u* + g and
// the body force 2 rho g (boundary_condition.f90:1645-1659, then
// solution.f90:599-650), and reduce each wing's force and moment
// (boundary_condition.f90:1662-1675).  Blocks are (wing, node block); each
// writes its partial sums.  k_force_row adds them in block order and writes
// the normalised force row (solution.f90:330-401); it runs off the lattice
// stream's critical path, at the start of the next step's IBM chain.
constexpr int ADJ_THREADS = 256;
constexpr int ADJ_BLOCKS_PER_WING = 16;   // grid-stride over a wing's adjoint nodes

int adj_collide_blocks_per_wing(const IbmParams& p) {
  return p.W > 0 ? std::min(ADJ_BLOCKS_PER_WING, (p.capw + ADJ_THREADS - 1) / ADJ_THREADS) : 1;
}

// Rigid-body update of wing 1 from the step's raw force (boundary_condition.
// f90:1694-1719): fx = x force summed over all wings, mom = moment of wing 1.
// RK4 (:3168-3202) with constant acceleration over the step, and the pitch
// equation theta'' = p1 (theta + alphad) + p2 mom; or explicit Euler.
__device__ void body_step(const IbmParams& p, double fx, double mom, double (*rel)[MAXW]) {
  double* a = &rel[REL_A][0];
  rel[REL_A_PREV][0] = *a;
  *a = fx / p.dyn_mass;
  const double v1 = rel[REL_V][0], s1 = rel[REL_S][0];
  const double th1 = rel[REL_THETA][0], om1 = rel[REL_TEMP_THETA][0];
  rel[REL_V_PREV][0] = v1;
  rel[REL_S_PREV][0] = s1;
  rel[REL_THETA_PREV][0] = th1;
  rel[REL_TEMP_THETA_PREV][0] = om1;
  const double acc = *a, p1 = p.dyn_p1, p2 = p.dyn_p2, ad = p.dyn_alphad;
  if (!p.dyn_rk4) {
    rel[REL_V][0] = acc + v1;
    rel[REL_S][0] = s1 + v1;
    rel[REL_POS][0] = rel[REL_POS][0] + v1;
    const double th = th1 + om1;
    rel[REL_THETA][0] = th;
    rel[REL_TEMP_THETA][0] = om1 + p1 * (th + ad) + p2 * mom;
    return;
  }
  // stage slopes of (s, v): (v1, a), (v1 + a/2, a), (v1 + a/2, a), (v1 + a, a)
  const double ds = (v1 + 2.0 * (v1 + 0.5 * acc) + 2.0 * (v1 + 0.5 * acc) + (v1 + acc)) / 6.0;
  rel[REL_S][0] = s1 + ds;
  rel[REL_V][0] = v1 + (acc + 2.0 * acc + 2.0 * acc + acc) / 6.0;
  rel[REL_POS][0] = rel[REL_POS][0] + ds;
  // (theta, omega): q1 = omega slope at the start, r_i = omega' at stage i
  const double pm = p2 * mom;
  const double r1 = p1 * (th1 + ad) + pm;
  const double q2 = om1 + 0.5 * r1;
  const double r2 = p1 * (th1 + ad + 0.5 * om1) + pm;
  const double q3 = om1 + 0.5 * r2;
  const double r3 = p1 * (th1 + ad + 0.5 * q2) + pm;
  const double q4 = om1 + r3;
  const double r4 = p1 * (th1 + ad + q3) + pm;
  rel[REL_THETA][0] = th1 + (om1 + 2.0 * q2 + 2.0 * q3 + q4) / 6.0;
  rel[REL_TEMP_THETA][0] = om1 + (r1 + 2.0 * r2 + 2.0 * r3 + r4) / 6.0;
}

// Normalised force row and airfoil_motion row of one step into LDS arrays
// (one thread); the per-wing force and moment of the step go to st.  With
// rigid-body coupling the step's force first advances st->rel, and the
// motion columns carry the body state after that update, as the reference's
// row does (its IBM routine updates body_extra_* before the row is written).
__device__ void compute_rows(const IbmParams& p, const GeoStep* gs, DevStep* st, const double (*raw)[3], int n,
                             double* row, double* mrow) {
  const bool ok = st->error == 0;
  const double UU = p.uu, scale = p.scale;
  double wf[MAXW][2], wmom[MAXW];
  double F0 = 0.0, F1 = 0.0, Mt = 0.0;
  for (int w = 0; w < p.W; w++) {
    const double* r = raw[w];
    wf[w][0] = ok ? 2.0 * r[0] / (UU * UU * (1.0 / scale)) : 0.0;
    wf[w][1] = ok ? 2.0 * r[1] / (UU * UU * (1.0 / scale)) : 0.0;
    wmom[w] = ok ? -2.0 * r[2] / (UU * UU * (1.0 / (scale * scale))) : 0.0;
    st->force[w][0] = wf[w][0];
    st->force[w][1] = wf[w][1];
    st->moment[w] = wmom[w];
  }
  for (int w = 0; w < p.W; w++) {
    F0 = F0 + wf[w][0];
    F1 = F1 + wf[w][1];
  }
  for (int w = 0; w < p.W; w++) Mt = Mt + wmom[w];
  double be_v[MAXW], be_s[MAXW];
  for (int w = 0; w < p.W; w++) {
    be_v[w] = gs->frame.body_extra_v[w];
    be_s[w] = gs->frame.body_extra_s[w];
  }
  if (p.dyn_on && ok && p.W > 0) {
    double fx = 0.0;
    for (int w = 0; w < p.W; w++) fx = fx + raw[w][0];
    body_step(p, fx, raw[0][2], st->rel);
    double av = 0.0, as = 0.0;   // update_body_extra_from_rel (mode.f90:473-495)
    for (int w = 0; w < p.W; w++) {
      av = av + st->rel[REL_V][w];
      as = as + st->rel[REL_S][w];
      be_v[w] = av;
      be_s[w] = as;
    }
  }
  const double time = p.W > 0 ? gs->time : (double)(n - 1) * p.dt;
  int col = 0;
  row[col++] = time;
  for (int w = 0; w < p.W; w++) {
    row[col++] = wf[w][0];
    row[col++] = wf[w][1];
    row[col++] = wmom[w];
  }
  row[col++] = F0;
  row[col++] = F1;
  row[col++] = Mt;
  for (int w = 0; w < p.W; w++) {
    const WingMotionState& m = gs->wing[w];
    row[col++] = p.wing_x0 + (double)w * p.wing_dx + m.sa + be_s[w] / (1.0 / scale);
    row[col++] = m.sau;
    row[col++] = m.saac;
    row[col++] = p.wing_y0 + m.sb;
    row[col++] = m.sbv;
    row[col++] = m.sbac;
    row[col++] = m.angle;
    row[col++] = m.omega;
    row[col++] = m.omgac;
  }
  int c = 0;
  mrow[c++] = time;
  for (int w = 0; w < p.W; w++) {
    const WingMotionState& m = gs->wing[w];
    mrow[c++] = 0.0;
    mrow[c++] = be_v[w] / UU;
    mrow[c++] = p.wing_x0 + (double)w * p.wing_dx + m.sa + be_s[w] / (1.0 / scale);
  }
}

// blk in [0, W*bpw) is (wing, node block).
template <typename S, bool FIELDS>
__device__ __forceinline__ void adj_collide_body(const Grid& g, const MrtParams& P, const IbmParams& p,
                                                 const GeoBuffers& G, int slot, const FlowBuffers& F, int bpw, int n,
                                                 const S* __restrict__ src, S* __restrict__ dst, const FieldSlot& fs,
                                                 int blk) {
  __shared__ double red[3][ADJ_THREADS];
  const int t = threadIdx.x;
  const int w = blk / bpw, b = blk - w * bpw;
  const GeoStep* gs = G.step + slot;
  double fx = 0.0, fy = 0.0, fm = 0.0;
  const int cnt = gs->cnt[w];
  for (int a = b * ADJ_THREADS + t; a < cnt; a += bpw * ADJ_THREADS) {
    const size_t la = (size_t)slot * p.cap + (size_t)w * p.capw + a;
    const int i = G.adj_ij[2 * la], j = G.adj_ij[2 * la + 1];
    double f[9];
    pull_node_s(src, g, i, j, P.w, f);
    double rho, us, vs;
    node_moments(f, rho, us, vs);
    const size_t ga = (size_t)w * p.capw + a;
    const double gx = F.adj_g[2 * ga], gy = F.adj_g[2 * ga + 1];
    const double u = us + gx, v = vs + gy;
    const double t1 = 2.0 * rho * gx, t2 = 2.0 * rho * gy;
    double out[9];
    collide_node(P, f, rho, u, v, t1, t2, out);
    const long long o = (long long)j * g.nx + i;
#pragma unroll
    for (int k = 0; k < 9; k++) Store<S>::st(dst, (long long)k * g.ps + o, out[k], P.w[k]);
    if (FIELDS) {
      if (fs.rho) fs.rho[o] = rho;
      fs.u[o] = u;
      fs.v[o] = v;
    }
    const WingMotionState& m = gs->wing[w];
    const double x0 = (p.ox + p.scale * (double)i - (p.xc + p.wing_x0 + (double)w * p.wing_dx + m.sa)) / p.scale -
                      gs->frame.body_extra_pos[w];
    const double y0 = (p.oy + p.scale * (double)j - (p.yc + p.wing_y0 + m.sb)) / p.scale;
    fx = fx - t1;
    fy = fy - t2;
    fm = fm + (-x0 * t2 + y0 * t1);
  }
  red[0][t] = fx;
  red[1][t] = fy;
  red[2][t] = fm;
  __syncthreads();
  for (int h = ADJ_THREADS / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int c = 0; c < 3; c++) red[c][t] += red[c][t + h];
    __syncthreads();
  }
  if (t == 0) {
    double* part = F.partial + (size_t)(n & 1) * MAXW * bpw * 3;
    for (int c = 0; c < 3; c++) part[((size_t)w * bpw + b) * 3 + c] = red[c][0];
  }
}

template <typename S, bool FIELDS>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_collide(Grid g, MrtParams P, IbmParams p, GeoBuffers G, int slot,
                                                             FlowBuffers F, int bpw, int n, const S* __restrict__ src,
                                                             S* __restrict__ dst, FieldSlot fs) {
  adj_collide_body<S, FIELDS>(g, P, p, G, slot, F, bpw, n, src, dst, fs, blockIdx.x);
}

// Per-wing sums of the adjoint-collide partials in block order, then the
// force row of step n into row slot_row (one 64-thread workgroup).
__device__ __forceinline__ void force_row_body(const IbmParams& p, const GeoBuffers& G, int slot,
                                               const FlowBuffers& F, DevStep* st, int n, int slot_row, int row_len,
                                               int bpw) {
  __shared__ double s_raw[MAXW][3];
  __shared__ double s_row[LBM_MAX_FORCE_ROW];
  __shared__ double s_mrow[1 + 3 * MAXW];
  const int t = threadIdx.x;
  const double* part = F.partial + (size_t)(n & 1) * MAXW * bpw * 3;
  if (t < 3 * p.W) {
    const int ww = t / 3, c = t - 3 * ww;
    double r = 0.0;
    for (int q = 0; q < bpw; q++) r = r + part[((size_t)ww * bpw + q) * 3 + c];
    s_raw[ww][c] = r;
  }
  __syncthreads();
  if (t == 0) {
    compute_rows(p, G.step + slot, st, s_raw, n, s_row, s_mrow);
    if (p.dyn_on && F.frame_out && st->error == 0) {
      // the next step's frame from the updated body state
      // (update_body_extra_from_rel, mode.f90:473-495); movenum changes only
      // with a move, which the host applies
      Frame f = G.step[slot].frame;
      double av = 0.0, as = 0.0, ap = 0.0;
      for (int w = 0; w < p.W; w++) {
        av = av + st->rel[REL_V][w];
        as = as + st->rel[REL_S][w];
        ap = ap + st->rel[REL_POS][w];
        f.body_extra_v[w] = av;
        f.body_extra_s[w] = as;
        f.body_extra_pos[w] = ap;
      }
      *F.frame_out = f;
    }
  }
  __syncthreads();
  double* row = F.rows + (long long)slot_row * row_len;
  for (int c = t; c < row_len; c += blockDim.x) {
    row[c] = s_row[c];
    F.row_sum[c] = F.row_sum[c] + s_row[c];
  }
  double* mrow = F.mrows + (long long)slot_row * (1 + 3 * p.W);
  for (int c = t; c < 1 + 3 * p.W; c += blockDim.x) mrow[c] = s_mrow[c];
}

__global__ __launch_bounds__(64) void k_force_row(IbmParams p, GeoBuffers G, int slot, FlowBuffers F, DevStep* st,
                                                  int n, int slot_row, int row_len, int bpw) {
  force_row_body(p, G, slot, F, st, n, slot_row, row_len, bpw);
}

void launch_force_row(hipStream_t s, const IbmParams& p, const GeoBuffers& G, int slot, const FlowBuffers& F,
                      DevStep* st, int n, int row, int row_len) {
  hipLaunchKernelGGL(k_force_row, dim3(1), dim3(64), 0, s, p, G, slot, F, st, n, row, row_len,
                     adj_collide_blocks_per_wing(p));
}

template <typename S>
void launch_adj_collide_t(hipStream_t s, const Grid& g, const MrtParams& P, const IbmParams& p, const GeoBuffers& G,
                          int slot, const FlowBuffers& F, int n, const S* src, S* dst, FieldSlot* fields) {
  if (p.W <= 0) return;
  const int bpw = adj_collide_blocks_per_wing(p);
  dim3 grid(p.W * bpw), block(ADJ_THREADS);
  if (fields)
    hipLaunchKernelGGL((k_adj_collide<S, true>), grid, block, 0, s, g, P, p, G, slot, F, bpw, n, src, dst, *fields);
  else
    hipLaunchKernelGGL((k_adj_collide<S, false>), grid, block, 0, s, g, P, p, G, slot, F, bpw, n, src, dst,
                       FieldSlot{});
}

void launch_adj_collide(hipStream_t s, const Grid& g, const MrtParams& P, const IbmParams& p, const GeoBuffers& G,
                        int slot, const FlowBuffers& F, int n, const void* src, void* dst, FieldSlot* fields,
                        bool f32) {
  if (f32)
    launch_adj_collide_t(s, g, P, p, G, slot, F, n, (const float*)src, (float*)dst, fields);
  else
    launch_adj_collide_t(s, g, P, p, G, slot, F, n, (const double*)src, (double*)dst, fields);
}

// ------------------------------------------------------------------ far field

struct NodeVals {
  double rho, u, v;
};

template <typename S>
__device__ __forceinline__ NodeVals interior_vals(const S* __restrict__ src, const Grid& g, const double* w, int i,
                                                  int j, double f[9]) {
  pull_node_s(src, g, i, j, w, f);
  NodeVals r;
  node_moments(f, r.rho, r.u, r.v);
  return r;
}

// Characteristic update of left-column node (0, j) (boundary_condition.f90:418-439).
template <typename S>
__device__ NodeVals lodi_left(const S* __restrict__ src, const Grid& g, const double* w, const RingBuffers& R,
                              double cs, int j) {
  const int ny = g.ny;
  const double* mr = R.mom;  // rho[0..ny), u[ny..2ny), v[2ny..3ny) for the left column
  const double trj = mr[j], tuj = mr[ny + j], tvj = mr[2 * ny + j];
  double f2[9], f3[9];
  const NodeVals n2 = interior_vals(src, g, w, 1, j, f2);
  const NodeVals n3 = interior_vals(src, g, w, 2, j, f3);
  const double rhox = (-3.0 * trj + 4.0 * n2.rho - n3.rho) * 0.5;
  const double ux = (-3.0 * tuj + 4.0 * n2.u - n3.u) * 0.5;
  const double vx = (-3.0 * tvj + 4.0 * n2.v - n3.v) * 0.5;
  const double lx1 = (tuj - cs) * (cs * cs * rhox - cs * trj * ux);
  const double lx2 = tuj < 0.0 ? tuj * vx : 0.0;
  const double lx3 = 0.0;
  NodeVals r;
  r.rho = trj - (lx1 + lx3) * 1.5;
  r.u = tuj + (lx1 - lx3) / (2.0 * cs * r.rho);
  r.v = tvj - lx2;
  return r;
}

// Characteristic update of right-column node (nx-1, j) (boundary_condition.f90:500-521).
template <typename S>
__device__ NodeVals lodi_right(const S* __restrict__ src, const Grid& g, const double* w, const RingBuffers& R,
                               double cs, int j) {
  const int ny = g.ny, nx = g.nx;
  const double* mr = R.mom + 3 * ny;
  const double trj = mr[j], tuj = mr[ny + j], tvj = mr[2 * ny + j];
  double f2[9], f3[9];
  const NodeVals n2 = interior_vals(src, g, w, nx - 2, j, f2);
  const NodeVals n3 = interior_vals(src, g, w, nx - 3, j, f3);
  const double rhox = (3.0 * trj - 4.0 * n2.rho + n3.rho) * 0.5;
  const double ux = (3.0 * tuj - 4.0 * n2.u + n3.u) * 0.5;
  const double vx = (3.0 * tvj - 4.0 * n2.v + n3.v) * 0.5;
  const double lx1 = 0.0;
  const double lx2 = tuj >= 0.0 ? tuj * vx : 0.0;
  const double lx3 = (tuj + cs) * (cs * cs * rhox + cs * trj * ux);
  NodeVals r;
  r.rho = trj - (lx1 + lx3) * 1.5;
  r.u = tuj + (lx1 - lx3) / (2.0 * cs * r.rho);
  r.v = tvj - lx2;
  return r;
}

// One thread per boundary-ring node: left and right columns (j = 1..ny-2),
// then the top and bottom walls (i = 0..nx-1), in the order and with the
// overwrites of boundary_condition1 (boundary_condition.f90:403-853).
template <typename S, bool FIELDS>
__device__ __forceinline__ void bc_ring_body(const Grid& g, const MrtParams& P, const RingBuffers& R,
                                             const S* __restrict__ src, S* __restrict__ dst, const FieldSlot& fs,
                                             int blk) {
  const int t = blk * 256 + threadIdx.x;
  const int nx = g.nx, ny = g.ny, nside = ny - 2;
  const int total = 2 * nside + 2 * nx;
  if (t >= total) return;
  const double cs = P.cs;
  const double* w = P.w;
  double out[9], f0n[9];
  int bi, bj;
  NodeVals b, nv;
  if (t < nside) {  // left (boundary_condition.f90:413-483)
    bj = 1 + t;
    bi = 0;
    b = lodi_left(src, g, w, R, cs, bj);
    nv = interior_vals(src, g, w, 1, bj, f0n);
  } else if (t < 2 * nside) {  // right (boundary_condition.f90:495-567)
    bj = 1 + t - nside;
    bi = nx - 1;
    b = lodi_right(src, g, w, R, cs, bj);
    nv = interior_vals(src, g, w, nx - 2, bj, f0n);
  } else {  // top (boundary_condition.f90:657-723) / bottom (784-850) walls
    const int tt = t - 2 * nside;
    const bool top = tt < nx;
    bi = top ? tt : tt - nx;
    bj = top ? ny - 1 : 0;
    const int nj = top ? ny - 2 : 1;  // node inside the wall
    if (bi == 0) {
      nv = lodi_left(src, g, w, R, cs, nj);
#pragma unroll
      for (int k = 0; k < 9; k++) f0n[k] = R.left[(long long)k * ny + nj];
    } else if (bi == nx - 1) {
      nv = lodi_right(src, g, w, R, cs, nj);
#pragma unroll
      for (int k = 0; k < 9; k++) f0n[k] = R.right[(long long)k * ny + nj];
    } else {
      nv = interior_vals(src, g, w, bi, nj, f0n);
    }
    b.rho = nv.rho;
    b.u = 0.0;
    b.v = 0.0;
  }
  bc_extrapolate(P, b.rho, b.u, b.v, nv.rho, nv.u, nv.v, f0n, out);
  const long long o = (long long)bj * nx + bi;
#pragma unroll
  for (int k = 0; k < 9; k++) Store<S>::st(dst, (long long)k * g.ps + o, out[k], w[k]);
  if (FIELDS) {
    if (fs.rho) fs.rho[o] = b.rho;
    fs.u[o] = b.u;
    fs.v[o] = b.v;
  }
}

template <typename S, bool FIELDS>
__global__ __launch_bounds__(256) void k_bc_ring(Grid g, MrtParams P, RingBuffers R, const S* __restrict__ src,
                                                 S* __restrict__ dst, FieldSlot fs) {
  bc_ring_body<S, FIELDS>(g, P, R, src, dst, fs, blockIdx.x);
}

void launch_bc_ring(hipStream_t s, const Grid& g, const MrtParams& P, const RingBuffers& R, const void* src,
                    void* dst, FieldSlot* fields, bool f32) {
  const int total = 2 * (g.ny - 2) + 2 * g.nx;
  dim3 grid((total + 255) / 256), block(256);
  const FieldSlot fs = fields ? *fields : FieldSlot{};
  if (f32) {
    if (fields)
      hipLaunchKernelGGL((k_bc_ring<float, true>), grid, block, 0, s, g, P, R, (const float*)src, (float*)dst, fs);
    else
      hipLaunchKernelGGL((k_bc_ring<float, false>), grid, block, 0, s, g, P, R, (const float*)src, (float*)dst, fs);
  } else {
    if (fields)
      hipLaunchKernelGGL((k_bc_ring<double, true>), grid, block, 0, s, g, P, R, (const double*)src, (double*)dst, fs);
    else
      hipLaunchKernelGGL((k_bc_ring<double, false>), grid, block, 0, s, g, P, R, (const double*)src, (double*)dst,
                         fs);
  }
}

// Moments of the persistent f0 ring columns (physical_variables on the ring,
// solution.f90:528-532); they only change on reset, restart or move_mesh.
__global__ void k_ring_moments(Grid g, RingBuffers R) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ny = g.ny;
  if (t >= 2 * ny) return;
  const int side = t / ny, j = t - side * ny;
  const double* col = side == 0 ? R.left : R.right;
  double f[9];
#pragma unroll
  for (int k = 0; k < 9; k++) f[k] = col[(long long)k * ny + j];
  double rho, u, v;
  node_moments(f, rho, u, v);
  double* m = R.mom + 3 * ny * side;
  m[j] = rho;
  m[ny + j] = u;
  m[2 * ny + j] = v;
}

void launch_ring_moments(hipStream_t s, const Grid& g, const RingBuffers& R) {
  hipLaunchKernelGGL(k_ring_moments, dim3((2 * g.ny + 255) / 256), dim3(256), 0, s, g, R);
}

// init (solution.f90:33-63): f = f0 = f_eq(1, UU, 0) everywhere, u = UU.
struct Feq9 {
  double v[9];
};

template <typename S>
__global__ void k_init_fields(Grid g, MrtParams P, double uu, Feq9 feq, S* f, RingBuffers R, FieldSlot slot) {
  const long long n = (long long)g.nx * g.ny;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 9; k++) Store<S>::st(f, (long long)k * g.ps + q, feq.v[k], P.w[k]);
    slot.rho[q] = 1.0;
    slot.u[q] = uu;
    slot.v[q] = 0.0;
    if (q < g.ny) {
#pragma unroll
      for (int k = 0; k < 9; k++) {
        R.left[(long long)k * g.ny + q] = feq.v[k];
        R.right[(long long)k * g.ny + q] = feq.v[k];
      }
    }
    if (q < g.nx) {
#pragma unroll
      for (int k = 0; k < 9; k++) {
        R.bottom[(long long)k * g.nx + q] = feq.v[k];
        R.top[(long long)k * g.nx + q] = feq.v[k];
      }
    }
  }
}

void launch_init_fields(hipStream_t s, const Grid& g, const MrtParams& P, double uu, const double feq[9], void* f,
                        const RingBuffers& R, FieldSlot slot, bool f32) {
  Feq9 fe;
  for (int k = 0; k < 9; k++) fe.v[k] = feq[k];
  if (f32)
    hipLaunchKernelGGL(k_init_fields<float>, dim3(2048), dim3(256), 0, s, g, P, uu, fe, (float*)f, R, slot);
  else
    hipLaunchKernelGGL(k_init_fields<double>, dim3(2048), dim3(256), 0, s, g, P, uu, fe, (double*)f, R, slot);
}

template <typename S>
__global__ void k_moments_all(Grid g, MrtParams P, const S* __restrict__ f, FieldSlot slot) {
  const long long n = (long long)g.nx * g.ny;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    double fo[9];
#pragma unroll
    for (int k = 0; k < 9; k++) fo[k] = Store<S>::ld(f, (long long)k * g.ps + q, P.w[k]);
    double rho, u, v;
    node_moments(fo, rho, u, v);
    slot.rho[q] = rho;
    slot.u[q] = u;
    slot.v[q] = v;
  }
}

void launch_moments_all(hipStream_t s, const Grid& g, const MrtParams& P, const void* f, FieldSlot slot, bool f32) {
  if (f32)
    hipLaunchKernelGGL(k_moments_all<float>, dim3(2048), dim3(256), 0, s, g, P, (const float*)f, slot);
  else
    hipLaunchKernelGGL(k_moments_all<double>, dim3(2048), dim3(256), 0, s, g, P, (const double*)f, slot);
}

// erranalysis (solution.f90:77-130): sqrt(sum (u-u0)^2 + (v-v0)^2) /
// (sqrt(sum u^2 + v^2) + 1e-30) over the interior.  One workgroup per
// interior row writes the row's two sums (threads take every
// ERR_THREADS-th node, then a fixed tree in LDS); k_err_final adds the rows
// the same way.  The order never changes, so the value is reproducible.
constexpr int ERR_THREADS = 256;

int err_blocks(const Grid& g) { return g.ny - 2; }

__global__ __launch_bounds__(ERR_THREADS) void k_err_partial(Grid g, FieldSlot cur, FieldSlot prev, double* partial) {
  __shared__ double s1[ERR_THREADS], s2[ERR_THREADS];
  double a1 = 0.0, a2 = 0.0;
  const int j = 1 + blockIdx.x;
  constexpr int U = 16;  // nodes per thread whose loads are issued together (a whole row up to nx = 4098)
  for (int i0 = 1 + threadIdx.x; i0 <= g.nx - 2; i0 += U * ERR_THREADS) {
    double cu[U], cv[U], pu[U], pv[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      const int i = i0 + q * ERR_THREADS;
      const long long o = (long long)j * g.nx + i;
      const bool in = i <= g.nx - 2;
      cu[q] = in ? cur.u[o] : 0.0;
      cv[q] = in ? cur.v[o] : 0.0;
      pu[q] = in ? prev.u[o] : 0.0;
      pv[q] = in ? prev.v[o] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < U; q++) {
      const double du = cu[q] - pu[q], dv = cv[q] - pv[q];
      a1 = a1 + du * du + dv * dv;
      a2 = a2 + cu[q] * cu[q] + cv[q] * cv[q];
    }
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int w = ERR_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s1[0];
    partial[2 * blockIdx.x + 1] = s2[0];
  }
}

void launch_err_partial(hipStream_t s, const Grid& g, FieldSlot cur, FieldSlot prev, double* partial, int nblocks) {
  hipLaunchKernelGGL(k_err_partial, dim3(nblocks), dim3(ERR_THREADS), 0, s, g, cur, prev, partial);
}

__global__ __launch_bounds__(ERR_THREADS) void k_err_final(const double* partial, int nblocks, double* out) {
  __shared__ double s1[ERR_THREADS], s2[ERR_THREADS];
  double t1 = 0.0, t2 = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += ERR_THREADS) {
    t1 += partial[2 * b];
    t2 += partial[2 * b + 1];
  }
  s1[threadIdx.x] = t1;
  s2[threadIdx.x] = t2;
  __syncthreads();
  for (int w = ERR_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sqrt(s1[0]) / (sqrt(s2[0]) + 1e-30);
}

void launch_err_final(hipStream_t s, const double* partial, int nblocks, double* out) {
  hipLaunchKernelGGL(k_err_final, dim3(1), dim3(ERR_THREADS), 0, s, partial, nblocks, out);
}

// Lattice mass (the FP64 mass-conservation check; the reference has none, the
// reduction follows the pattern of erranalysis, solution.f90:77-130): the sum
// over every node of its nine populations, in FP64 whatever the storage.  One
// workgroup per lattice row writes the row's sum (threads take every
// MASS_THREADS-th node, then a fixed tree in LDS); k_mass_final adds the rows
// the same way.  The order never changes, so the value is reproducible.
constexpr int MASS_THREADS = 256;

__device__ __forceinline__ double block_sum_fixed(double a, double* s) {
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = MASS_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  return s[0];
}

template <typename S>
__global__ __launch_bounds__(MASS_THREADS) void k_mass_rows(Grid g, MrtParams P, const S* __restrict__ f,
                                                            double* __restrict__ rows) {
  __shared__ double s[MASS_THREADS];
  const int j = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < g.nx; i += MASS_THREADS) {
    const long long o = (long long)j * g.nx + i;
    double node = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) node += Store<S>::ld(f, (long long)k * g.ps + o, P.w[k]);
    a += node;
  }
  const double t = block_sum_fixed(a, s);
  if (threadIdx.x == 0) rows[j] = t;
}

__global__ __launch_bounds__(MASS_THREADS) void k_mass_final(const double* __restrict__ rows, int ny,
                                                             double* __restrict__ out) {
  __shared__ double s[MASS_THREADS];
  double a = 0.0;
  for (int j = threadIdx.x; j < ny; j += MASS_THREADS) a += rows[j];
  const double t = block_sum_fixed(a, s);
  if (threadIdx.x == 0) out[0] = t;
}

void launch_mass(hipStream_t s, const Grid& g, const MrtParams& P, const void* f, bool f32, double* rows,
                 double* out) {
  if (f32)
    hipLaunchKernelGGL(k_mass_rows<float>, dim3(g.ny), dim3(MASS_THREADS), 0, s, g, P, (const float*)f, rows);
  else
    hipLaunchKernelGGL(k_mass_rows<double>, dim3(g.ny), dim3(MASS_THREADS), 0, s, g, P, (const double*)f, rows);
  hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(MASS_THREADS), 0, s, rows, g.ny, out);
}

// The lattice mass of caller-owned populations with no allocation: the row
// sums live in a file-scope device array, one copy per GPU, whose device
// address is looked up once per device (a host reference to the array's name
// is not a device address).  Every call on a GPU shares that GPU's array, so
// calls that overlap on one GPU race.
constexpr int MASS_SCRATCH_ROWS = 32768;
static __device__ __align__(256) double scratch[MASS_SCRATCH_ROWS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

int mass_scratch_rows() { return MASS_SCRATCH_ROWS; }

hipError_t mass_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

// move_mesh column shift (move_mesh.f90:29-57 dr=+1, 185-213 dr=-1): out(i) =
// in(i+dr) where that column exists, else the column is kept.  Populations
// and fields are shifted as stored (the FP32 deviation of a shifted value is
// the shifted deviation).
template <typename T>
__global__ void k_shift_x(Grid g, int dr, const T* __restrict__ in, T* __restrict__ out, int planes) {
  const long long n = (long long)g.nx * g.ny;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n * planes;
       q += (long long)gridDim.x * blockDim.x) {
    const long long p = q / n, r = q - p * n;
    const int i = (int)(r % g.nx);
    const int si = i + dr;
    const bool keep = si < 0 || si >= g.nx;
    const long long pstride = planes == 9 ? g.ps : n;
    out[p * pstride + r] = keep ? in[p * pstride + r] : in[p * pstride + r + dr];
  }
}

void launch_shift_x(hipStream_t s, const Grid& g, int dr, const void* f, void* fshift, int planes, bool f32) {
  if (f32)
    hipLaunchKernelGGL(k_shift_x<float>, dim3(4096), dim3(256), 0, s, g, dr, (const float*)f, (float*)fshift, planes);
  else
    hipLaunchKernelGGL(k_shift_x<double>, dim3(4096), dim3(256), 0, s, g, dr, (const double*)f, (double*)fshift,
                       planes);
}

template <typename S>
__global__ void k_copy_ring_column(Grid g, MrtParams P, const S* __restrict__ f, int col, double* ring) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 9 * g.ny) return;
  const int k = t / g.ny, j = t - k * g.ny;
  ring[(long long)k * g.ny + j] = Store<S>::ld(f, (long long)k * g.ps + (long long)j * g.nx + col, P.w[k]);
}

void launch_copy_ring_column(hipStream_t s, const Grid& g, const MrtParams& P, const void* f, int col, double* ring,
                             bool f32) {
  dim3 grid((9 * g.ny + 255) / 256), block(256);
  if (f32)
    hipLaunchKernelGGL(k_copy_ring_column<float>, grid, block, 0, s, g, P, (const float*)f, col, ring);
  else
    hipLaunchKernelGGL(k_copy_ring_column<double>, grid, block, 0, s, g, P, (const double*)f, col, ring);
}

template <typename S>
__global__ void k_copy_ring_row(Grid g, MrtParams P, const S* __restrict__ f, int row, int keep, double* ring) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 9 * g.nx) return;
  const int k = t / g.nx, i = t - k * g.nx;
  if (i == keep) return;
  ring[(long long)k * g.nx + i] = Store<S>::ld(f, (long long)k * g.ps + (long long)row * g.nx + i, P.w[k]);
}

void launch_copy_ring_row(hipStream_t s, const Grid& g, const MrtParams& P, const void* f, int row, int keep,
                          double* ring, bool f32) {
  dim3 grid((9 * g.nx + 255) / 256), block(256);
  if (f32)
    hipLaunchKernelGGL(k_copy_ring_row<float>, grid, block, 0, s, g, P, (const float*)f, row, keep, ring);
  else
    hipLaunchKernelGGL(k_copy_ring_row<double>, grid, block, 0, s, g, P, (const double*)f, row, keep, ring);
}

// ------------------------------------------------------- batched launches

int bc_ring_blocks(const Grid& g) { return (2 * (g.ny - 2) + 2 * g.nx + 255) / 256; }

// masked = 0 (tile timing before any geometry exists): every interior node.
// A work item is a group of ROWS consecutive tiles of one episode (in
// row-major tile order), so that the chain of dependent loads a workgroup
// needs before its first population load (item, then descriptor, then
// buffer pointers) is paid once per ROWS tiles.
//
// Workgroups are placed on the 8 XCDs round-robin (workgroup b on XCD b % 8),
// and each XCD has its own L2.  Workgroup b takes item run(b % 8) + b / 8,
// where XCD x's run is the x-th eighth of the list in order, so that the
// rows a tile pulls from (j - 1 and j + 1) were loaded by tiles of the same
// XCD shortly before, whatever the lattice's width.  (In list order a row's
// neighbour rows fall on the same XCD only when a row holds a multiple of 8
// tiles; on 64 uneven episodes the mapping took the kernel from 705 to
// 611 us at the median.)
template <bool F32, bool FIELDS, int BX, int ROWS, int CP>
__global__ __launch_bounds__(BX) void kb_lattice(const EpDesc* __restrict__ eps, const int2* __restrict__ items,
                                                 const int* __restrict__ kb, int kk, int masked) {
  const int nb = (int)gridDim.x, per = nb >> 3, extra = nb & 7;
  const int x = (int)blockIdx.x & 7;
  const int2 it = items[x * per + min(x, extra) + ((int)blockIdx.x >> 3)];
  const EpDesc& D = eps[it.x];
  const int k = kb[0] + kk;
  const int n = D.n0 + k, cur = (D.cur0 + k) & 1;
  AdjMask am = adj_mask(D.p, D.G, ep_slot(D, k));
  if (!masked) am.W = 0;
  const Grid g = D.g;
  const FieldSlot fs = D.fslot[n % 3];
#pragma unroll 1
  for (int r = 0; r < ROWS; r++) {
    const int blk = it.y * ROWS + r;
    if (blk >= D.tiles) break;
    int i, j;
    flat_node<BX>(g, blk, (int)threadIdx.x, i, j);
    if (j > g.ny - 2) break;
    if (j < 1) continue;
    if (F32)
      lattice_f32_node<FIELDS, CP>(g, D.P, (const float*)D.fbuf[cur], (float*)D.fbuf[cur ^ 1], fs, i, j, am);
    else
      lattice_node<FIELDS, CP>(g, D.P, (const double*)D.fbuf[cur], (double*)D.fbuf[cur ^ 1], fs, i, j, am);
  }
}

template <typename S, bool FIELDS>
__global__ __launch_bounds__(256) void kb_bc_ring(const EpDesc* __restrict__ eps, const int2* __restrict__ items,
                                                  const int* __restrict__ kb, int kk) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  const int k = kb[0] + kk;
  const int n = D.n0 + k, cur = (D.cur0 + k) & 1;
  bc_ring_body<S, FIELDS>(D.g, D.P, D.R, (const S*)D.fbuf[cur], (S*)D.fbuf[cur ^ 1], D.fslot[n % 3], it.y);
}

template <typename S, bool FIELDS>
__global__ __launch_bounds__(ADJ_THREADS) void kb_adj_collide(const EpDesc* __restrict__ eps,
                                                              const int2* __restrict__ items,
                                                              const int* __restrict__ kb, int kk) {
  const int2 it = items[blockIdx.x];
  const EpDesc& D = eps[it.x];
  const int k = kb[0] + kk;
  const int n = D.n0 + k, cur = (D.cur0 + k) & 1;
  adj_collide_body<S, FIELDS>(D.g, D.P, D.p, D.G, ep_slot(D, k), D.F, D.bpw, n, (const S*)D.fbuf[cur],
                              (S*)D.fbuf[cur ^ 1], D.fslot[n % 3], it.y);
}

// force row of substep k of every episode: one workgroup per episode
__global__ __launch_bounds__(64) void kb_force_row(const EpDesc* __restrict__ eps, const int* __restrict__ kb,
                                                   int kk) {
  const EpDesc& D = eps[blockIdx.x];
  const int k = kb[0] + kk;
  force_row_body(D.p, D.G, ep_slot(D, k), D.F, D.st, D.n0 + k, k, 4 + 12 * D.p.W, D.bpw);
}

void launch_batch_force_row(hipStream_t s, const EpDesc* eps, const int* kb, int count, int k) {
  hipLaunchKernelGGL(kb_force_row, dim3(count), dim3(64), 0, s, eps, kb, k);
}

__global__ __launch_bounds__(64) void k_step_base(int* __restrict__ p, int v, int add) {
  if (threadIdx.x == 0) p[0] = add ? p[0] + v : v;
}

void launch_step_base(hipStream_t s, int* p, int v, bool add) {
  hipLaunchKernelGGL(k_step_base, dim3(1), dim3(64), 0, s, p, v, add ? 1 : 0);
}

template <int BX, int ROWS, int CP>
void launch_batch_lattice_cp(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields,
                             int masked, hipEvent_t t1, size_t lds) {
  dim3 grid(it.n_lattice), block(BX);
  if (f32) {
    if (fields)
      hipExtLaunchKernelGGL((kb_lattice<true, true, BX, ROWS, CP>), grid, block, lds, s, nullptr, t1, 0, eps,
                            it.lattice, it.kbase, k, masked);
    else
      hipExtLaunchKernelGGL((kb_lattice<true, false, BX, ROWS, CP>), grid, block, lds, s, nullptr, t1, 0, eps,
                            it.lattice, it.kbase, k, masked);
  } else {
    if (fields)
      hipExtLaunchKernelGGL((kb_lattice<false, true, BX, ROWS, CP>), grid, block, lds, s, nullptr, t1, 0, eps,
                            it.lattice, it.kbase, k, masked);
    else
      hipExtLaunchKernelGGL((kb_lattice<false, false, BX, ROWS, CP>), grid, block, lds, s, nullptr, t1, 0, eps,
                            it.lattice, it.kbase, k, masked);
  }
}

template <int BX, int ROWS>
void launch_batch_lattice_bx(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields,
                             int cp, int masked, hipEvent_t t1, size_t lds) {
  if (cp == CP_NT)
    launch_batch_lattice_cp<BX, ROWS, CP_NT>(s, eps, it, k, f32, fields, masked, t1, lds);
  else if (cp == CP_NT_STORE)
    launch_batch_lattice_cp<BX, ROWS, CP_NT_STORE>(s, eps, it, k, f32, fields, masked, t1, lds);
  else
    launch_batch_lattice_cp<BX, ROWS, CP_PLAIN>(s, eps, it, k, f32, fields, masked, t1, lds);
}

template <int BX>
void launch_batch_lattice_rows(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields,
                               int rows, int cp, int masked, hipEvent_t t1, size_t lds) {
  if (rows == 4)
    launch_batch_lattice_bx<BX, 4>(s, eps, it, k, f32, fields, cp, masked, t1, lds);
  else if (rows == 2)
    launch_batch_lattice_bx<BX, 2>(s, eps, it, k, f32, fields, cp, masked, t1, lds);
  else
    launch_batch_lattice_bx<BX, 1>(s, eps, it, k, f32, fields, cp, masked, t1, lds);
}

void launch_batch_lattice(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields,
                          int bx, int rows, int cp, bool masked, hipEvent_t t1, size_t lds) {
  if (it.n_lattice <= 0) return;
  if (bx == 128)
    launch_batch_lattice_rows<128>(s, eps, it, k, f32, fields, rows, cp, masked, t1, lds);
  else if (bx == 512)
    launch_batch_lattice_rows<512>(s, eps, it, k, f32, fields, rows, cp, masked, t1, lds);
  else
    launch_batch_lattice_rows<256>(s, eps, it, k, f32, fields, rows, cp, masked, t1, lds);
}

void launch_batch_bc_ring(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields) {
  if (it.n_ring <= 0) return;
  dim3 grid(it.n_ring), block(256);
  if (f32) {
    if (fields) hipLaunchKernelGGL((kb_bc_ring<float, true>), grid, block, 0, s, eps, it.ring, it.kbase, k);
    else hipLaunchKernelGGL((kb_bc_ring<float, false>), grid, block, 0, s, eps, it.ring, it.kbase, k);
  } else {
    if (fields) hipLaunchKernelGGL((kb_bc_ring<double, true>), grid, block, 0, s, eps, it.ring, it.kbase, k);
    else hipLaunchKernelGGL((kb_bc_ring<double, false>), grid, block, 0, s, eps, it.ring, it.kbase, k);
  }
}

void launch_batch_adj_collide(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields) {
  if (it.n_adj <= 0) return;
  dim3 grid(it.n_adj), block(ADJ_THREADS);
  if (f32) {
    if (fields) hipLaunchKernelGGL((kb_adj_collide<float, true>), grid, block, 0, s, eps, it.adj, it.kbase, k);
    else hipLaunchKernelGGL((kb_adj_collide<float, false>), grid, block, 0, s, eps, it.adj, it.kbase, k);
  } else {
    if (fields) hipLaunchKernelGGL((kb_adj_collide<double, true>), grid, block, 0, s, eps, it.adj, it.kbase, k);
    else hipLaunchKernelGGL((kb_adj_collide<double, false>), grid, block, 0, s, eps, it.adj, it.kbase, k);
  }
}

// The lattice kernel's memory pattern alone: every interior node pulls its
// nine populations from the neighbours of the source planes and writes them
// to the destination planes, with no moment or collision arithmetic, under
// the lattice kernel's launch options (cache policy, wave cap).  Its time is
// the ceiling of that access pattern on the device.
template <typename T, int BX, int CP>
__global__ __launch_bounds__(BX) void k_pull_copy(const T* __restrict__ src, T* __restrict__ dst, int nx, int ny) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y;
  if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2) return;
  const size_t ps = (size_t)nx * ny;
  T f[9];
#pragma unroll
  for (int k = 0; k < 9; k++) f[k] = ld_pop<CP>(src + k * ps + (size_t)(j - EY(k)) * nx + (i - EX(k)));
#pragma unroll
  for (int k = 0; k < 9; k++) st_pop<CP>(dst + k * ps + (size_t)j * nx + i, f[k]);
}

template <int CP>
void launch_pull_copy(bool f32, const void* s, void* d, int nx, int ny, size_t lds) {
  const dim3 grid((nx + 255) / 256, ny), block(256);
  if (f32)
    hipLaunchKernelGGL((k_pull_copy<float, 256, CP>), grid, block, lds, 0, (const float*)s, (float*)d, nx, ny);
  else
    hipLaunchKernelGGL((k_pull_copy<double, 256, CP>), grid, block, lds, 0, (const double*)s, (double*)d, nx, ny);
}

// The fastest average over the cache policies and the wave caps the lattice
// kernel's autotuner tries (each variant timed over `launches` launches, the
// best of three rounds).
int measure_pull_copy(int nx, int ny, bool f32, int launches, double* avg_us) {
  const size_t ps = (size_t)nx * ny, bytes = 9 * ps * (f32 ? 4 : 8);
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0, dev = 0, lds_cu = 0, lds_blk = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    lds_cu = lds_blk = 0;
  auto launch = [&](int cp, size_t lds, const void* s, void* d) {
    if (cp == CP_NT)
      launch_pull_copy<CP_NT>(f32, s, d, nx, ny, lds);
    else if (cp == CP_NT_STORE)
      launch_pull_copy<CP_NT_STORE>(f32, s, d, nx, ny, lds);
    else
      launch_pull_copy<CP_PLAIN>(f32, s, d, nx, ny, lds);
  };
  double best = 1e300;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess || hipMemset(a, 0, bytes) != hipSuccess ||
      hipMemset(b, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = LBM_E_DEVICE;
  } else {
    for (int round = 0; round < 3 && rc == 0; round++)
      for (int cp : {CP_PLAIN, CP_NT_STORE, CP_NT})
        for (int wv : {0, 24, 16, 12, 8}) {
          const long long lds = lattice_lds_for_waves(256, wv, lds_cu, lds_blk);
          if (lds < 0 || (wv > 0 && lds_cu <= 0)) continue;
          for (int r = 0; r < 2; r++) launch(cp, (size_t)lds, r & 1 ? b : a, r & 1 ? a : b);
          float ms = 0.f;
          (void)hipEventRecord(e0, 0);
          for (int r = 0; r < launches; r++) launch(cp, (size_t)lds, r & 1 ? b : a, r & 1 ? a : b);
          (void)hipEventRecord(e1, 0);
          if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
              hipGetLastError() != hipSuccess) {
            rc = LBM_E_DEVICE;
            break;
          }
          best = std::min(best, 1000.0 * ms / launches);
        }
    if (rc == 0) *avg_us = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

}  // namespace lbm
