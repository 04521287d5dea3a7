// lbm_internal.h - device state and kernel launch wrappers (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "../../include/lbm_hip.h"
#include "lbm_band.h"
#include "lbm_device.h"

namespace lbm {

constexpr int MAXW = LBM_MAX_WINGS;
constexpr int BMAX = 24;          // largest IBM band half-width the solver accepts
constexpr int BOXW = 6;           // nodes per axis of a delta support box (|r| <= 2.5)
constexpr int BOXC = BOXW * BOXW; // cells of a support box
constexpr int CPL_K = 32;         // most coupling segments of one wing
constexpr int CPL_KT = 64;        // most coupling segments over all wings

// Error codes raised by kernels into DevStep::error (first one wins).
enum DevError : int {
  DERR_NONE = 0,
  DERR_BOX_EDGE = 1,       // support box within 3 nodes of the lattice edge
  DERR_MAP = 2,            // a wing's support rectangle exceeds its claim map
  DERR_BAND = 3,           // IBM band wider than BMAX
  DERR_WINGS_COUPLED = 4,  // more coupling segments than CPL_K of a wing / CPL_KT in all
  DERR_NOT_SPD = 5,        // non-positive pivot in the band factorisation
  DERR_MOTION = 6,         // motion source not ready
  DERR_WINGS_MET = 7,      // wings' support rectangles met on a step launched without the coupling pass
};

struct WingMotionState {
  double angle, omega, omgac, sa, sau, saac, sb, sbv, sbac;
};

// Fields of the rigid-body state body_rel_*, in restart-file order
// (solution.f90:1663-1665).
enum BodyRel : int {
  REL_A, REL_A_PREV, REL_V, REL_V_PREV, REL_S, REL_S_PREV, REL_POS, REL_THETA_PREV, REL_THETA,
  REL_TEMP_THETA_PREV, REL_TEMP_THETA, REL_FIELDS
};

// Error state, last-step forces and, with rigid-body coupling, the body
// state that each step's force row advances (one instance per solver).
struct DevStep {
  int error;
  int error_info;
  int rect_gap;            // smallest gap in nodes between two wings' support rectangles of the last solved step
  int pad;
  double force[MAXW][2];   // normalised per-wing force of the last step
  double moment[MAXW];
  double rel[REL_FIELDS][MAXW];
};

// Rigid-frame state of one step, known on the host before the step runs
// (movenum and update_body_extra_from_rel, main.f90:393-411, mode.f90:473-495).
struct Frame {
  int movenum;
  int pad;
  double body_extra_v[MAXW], body_extra_s[MAXW], body_extra_pos[MAXW];
};

// Geometry of one step (motion, markers, adjoint list, inverse IBM matrix),
// computed ahead of the flow by the geometry kernels.
struct GeoStep {
  WingMotionState wing[MAXW];
  double time;
  int n;
  int adj_n;               // adjoint nodes over all wings
  int cnt[MAXW];           // adjoint nodes per wing
  int bw[MAXW];            // band half-width of S per wing (zig-zag order)
  int rect[MAXW][4];       // i0, i1, j0, j1 (0-based) of the union of the wing's boxes
  int band[2];             // rows j0..j1 spanned by all the rectangles (j0 > j1: none)
  // Coupled wings (cpl_on = 1): wings whose support boxes share lattice
  // nodes.  cpl_k[cpl_off[w] .. cpl_off[w+1]) are the wing-local segments of
  // wing w whose boxes meet a box of another wing (ascending), cpl_w their
  // wing; cpl_n = cpl_off[W] in all.  A shared node belongs to the lowest of
  // its wings, whose segments claim it first.
  int cpl_on, cpl_n;
  int cpl_off[MAXW + 1];
  int cpl_k[CPL_KT], cpl_w[CPL_KT];
  // 1: the step's system is solved with the explicit inverse S^-1 (dense
  // mat-vec); 0: by substitution with the band factors (launches asked for
  // it and no two wings' support rectangles meet)
  int inv_on;
  Frame frame;
};

// Host-known constants of the IBM / motion kernels.
struct IbmParams {
  int W, P, SPW, nseg, npts;
  int capw, cap;           // adjoint capacity per wing (36*SPW) and in total
  int map_cap;             // cells of one claim map
  int nx, ny;
  double scale, ox, oy, uu, dt;
  double wing_x0, wing_dx, wing_y0, xc, yc;
  int motion_source;       // LBM_MOTION_FILE / LBM_MOTION_STREAM
  int step_count;          // rows of the file-mode step table
  int ncols;
  int band_cls;            // band_class() of a bound on every step's band half-width
  double motion_dt;
  // rigid-body coupling of wing 1 (boundary_condition.f90:1694-1719)
  int dyn_on, dyn_rk4;
  double dyn_mass;         // solid_den*area/scale^2 (closed boundary) or solid_den/scale^2
  double dyn_p1, dyn_p2;   // par1*UU^2*scale^2, par2*scale^4
  double dyn_alphad;
};

// Per-slot geometry arrays; slot s of an array X of per-slot length L starts
// at X + s*L.  Adjoint entries of wing w start at w*capw inside a slot.
struct GeoBuffers {
  GeoStep* step;           // [slots]
  const double* xb0;       // body-frame markers (npts), static
  const double* yb0;
  double* xb1; double* yb1;     // [slots][npts] moved markers
  double* xb2; double* yb2;     // [slots][nseg] segment midpoints
  double* ub; double* vb;       // [slots][nseg] rigid-body velocity at the midpoints
  int* box;                     // [slots][nseg][4] i1,i2,j1,j2 (0-based)
  double* phx; double* phy;     // [slots][nseg][BOXW] 1-D delta weights per box column / row
  int* seg_off;                 // [slots][W][SPW+1] wing-local adjoint offsets per segment
  int* adj_ij;                  // [slots][cap][2] (i, j) 0-based
  int* adj_seg;                 // [slots][cap] claiming segment (wing-local)
  double* sinv;                 // [slots][W][SPW][SPW] inverse of S = D D^T, zig-zag order
  double* bandg;                // [slots][W][band_cap] the factors L D L^T of S (band rows, then 1/D)
  double* cpl;                  // [slots][CPL_KT][CPL_KT] coupling matrix M (cpl_n x cpl_n)
  int* map;                     // [slots][W][map_cap] claim maps (owner segment per node)
  const Frame* frame;           // per step of the current call
  const double* step_table;     // file mode: step_count * ncols
  const double* stream_rows;    // stream mode: start row, target row (2*LBM_MAX_MOTION_COLS)
};

// Twisted L D L^T of a band S of n rows and half-width bw, for the band
// substitution at flow time: the first p pivots are eliminated from the top
// (band T: rows 0 .. p+m-1 of S) and the last q from the bottom (band R: the
// rows of S in reverse order, rows 0 .. q+m-1 of J S J), by two waves at
// once; the m = bw rows between them form a dense middle block, S's own
// entries there plus the Schur complements of both ends (the ends share no
// entry: they are more than bw rows apart).  Its substitutions likewise run
// from both ends towards the middle, solve the middle block, and run back
// out.  twist_ok: the bands long enough for it (both ends at least bw + 16
// pivots).
struct Twist {
  int p, q, m;   // pivots of the top end, of the bottom end, middle rows
  int nT, nR;    // rows of T (p + m) and of R (q + m)
};
__host__ __device__ __forceinline__ bool twist_ok(int n, int bw) { return n >= 4 * bw + 32; }
__host__ __device__ __forceinline__ Twist twist_split(int n, int bw) {
  Twist t;
  t.m = bw;
  t.p = (n - bw + 1) / 2;
  t.q = n - t.p - bw;
  t.nT = t.p + bw;
  t.nR = t.q + bw;
  return t;
}
// The twisted factors serve the steps with rigid-body coupling (p.dyn_on),
// whose solve is on the step's critical path; with prescribed motion the
// solve runs beside the lattice kernel, and the plain factors take half the
// LDS (which the batched geometry of many episodes shares).
__host__ __device__ __forceinline__ bool uses_twist(const IbmParams& p, int bw) {
  return p.dyn_on && twist_ok(p.SPW, bw);
}
// doubles of the twisted factors (row stride BWc + 1): T with BWc zero rows
// after it, R likewise, 1/D of T (nT + 1 entries), of R (nR + 1), the middle
// block's unit lower L (m x m, row-major) and its 1/D (m)
__host__ __device__ __forceinline__ size_t twist_doubles(int n, int bw, int BWc) {
  const Twist t = twist_split(n, bw);
  return (size_t)(t.nT + BWc + t.nR + BWc) * (BWc + 1) + t.nT + 1 + t.nR + 1 + (size_t)bw * bw + bw;
}

// doubles per (slot, wing) of GeoBuffers::bandg: the factors L D L^T, either
// (SPW + BMAX) band rows of stride BMAX + 1 at most then SPW + 1 entries of
// 1/D, or the twisted factors (at most SPW + 3 BMAX rows, SPW + 2 BMAX + 2
// entries of 1/D, BMAX^2 + BMAX for the middle block)
__host__ __device__ __forceinline__ size_t band_cap(const IbmParams& p) {
  return (size_t)(p.SPW + 3 * BMAX) * (BMAX + 1) + p.SPW + 2 * BMAX + 2 + BMAX * BMAX + BMAX;
}

// The adjoint nodes of one step, which the full-lattice kernel leaves to
// k_adj_collide (their only writer): the wings' support rectangles and claim
// maps in the step's geometry slot.  W = 0 selects no node.
struct AdjMask {
  const GeoStep* gs;
  const int* map;          // claim maps of the slot, [W][map_cap]; owner INT_MAX = free
  int W, map_cap;
};

__host__ __device__ __forceinline__ AdjMask adj_mask(const IbmParams& p, const GeoBuffers& G, int slot) {
  return AdjMask{G.step + slot, G.map + (size_t)slot * p.W * (size_t)p.map_cap, p.W, p.map_cap};
}

// Is (i, j) (0-based) an adjoint node of the mask's step?  Rows outside the
// band of the wings' rectangles cost one scalar compare.
__device__ __forceinline__ bool is_adjoint(const AdjMask& am, int i, int j) {
  if (am.W == 0 || j < am.gs->band[0] || j > am.gs->band[1]) return false;
  for (int w = 0; w < am.W; w++) {
    const int* r = am.gs->rect[w];
    const int i0 = r[0], i1 = r[1], j0 = r[2], j1 = r[3];
    if (j < j0 || j > j1 || i < i0 || i > i1) continue;
    const int* M = am.map + (size_t)w * am.map_cap;
    if (M[(i - i0) + (i1 - i0 + 1) * (j - j0)] != INT_MAX) return true;
  }
  return false;
}

// Flow-dependent IBM buffers of the step being run.
struct FlowBuffers {
  double* cvec;            // [nseg][2] right-hand side U_b - D u* by global band position
  double* yvec;            // [nseg][2] S^-1 c
  double* adj_g;           // [cap][2] velocity correction g at the adjoint nodes
  double* partial;         // [2][MAXW][blocks per wing][3] force partial sums by step parity
  double* rows;            // force rows of the current call
  double* mrows;           // airfoil_motion rows of the current call
  double* row_sum;         // sum of force rows
  Frame* frame_out;        // rigid-body coupling: the frame of the next step, or null
};

struct RingBuffers {
  double* left;     // f0 ring column i=0, 9 * ny
  double* right;    // f0 ring column i=nx-1, 9 * ny
  double* mom;      // rho,u,v of left column then right column: 6 * ny
  // f0 ring rows j=0 and j=ny-1, 9 * nx each: no step reads them (the wall
  // rows' moments are overwritten by the wall BC), but move_mesh shifts them
  // (move_mesh.f90:30-34) and the checkpoint saves them (solution.f90:1673-1684)
  double* bottom;
  double* top;
};

struct FieldSlot {
  double* rho; double* u; double* v;
};

// One episode of a batched advance: everything a kernel needs to run the
// episode's part of a substep k of the call.  With ns steps per geometry
// batch, substep k runs step n0 + k from fbuf[(cur0 + k) & 1] into the other
// buffer with geometry slot ((bc0 + k/ns) & 1)*ns + k%ns.
struct EpDesc {
  Grid g;
  MrtParams P;
  IbmParams p;
  GeoBuffers G;
  FlowBuffers F;
  RingBuffers R;
  DevStep* st;
  void* fbuf[2];
  FieldSlot fslot[3];      // field slot of step n is fslot[n % 3]
  int cur0, n0, ns, bc0;
  int tiles;               // lattice tiles of the batch's tile width over the interior rows taken as one run
  int bpw;                 // adjoint-collide blocks per wing
  int direct;              // 1: band substitution instead of the inverse on steps whose wings do not meet
  int pad1;
};

__host__ __device__ __forceinline__ int ep_slot(const EpDesc& D, int k) {
  return ((D.bc0 + k / D.ns) & 1) * D.ns + k % D.ns;
}

// Work items of the batched launches: (episode, block of that episode's own
// single-instance grid).
struct BatchItems {
  const int* kbase;                  // device step base: a launch for step k passes k - *kbase
  const int2* lattice; int n_lattice;
  const int2* ring; int n_ring;
  const int2* adj; int n_adj;
  const int2* seg; int n_seg;        // k_ibm_rhs / k_ibm_matvec blocks
  const int2* spread; int n_spread;
  const int2* wing; int n_wing;      // (episode, wing): band substitution of a wing's system
};

// Band position <-> segment in a wing of n segments: positions 0,1,2,3,...
// hold segments 0, n-1, 1, n-2, ... so that segments facing each other on
// the upper and lower surface are neighbours and S is a narrow band.
// Row stride class of a band of half-width bw (the band kernels are compiled
// for these strides; BMAX is the widest band accepted).
__host__ __device__ __forceinline__ int band_class(int bw) {
  return bw <= 8 ? 8 : (bw <= 12 ? 12 : (bw <= 16 ? 16 : BMAX));
}

__host__ __device__ __forceinline__ int seg_of_pos(int q, int n) { return (q & 1) ? n - 1 - (q >> 1) : (q >> 1); }
__host__ __device__ __forceinline__ int pos_of_seg(int k, int n) {
  return k < (n + 1) / 2 ? 2 * k : 2 * (n - 1 - k) + 1;
}

// lattice kernels (lbm_lattice.hip); populations are double (f32 = false) or
// FP32 deviations from the weights (f32 = true)
// bx: lattice tile width (one of kLatticeTiles)
constexpr int kLatticeTiles[3] = {128, 256, 512};
// mask: adjoint nodes of the step, skipped (written by k_adj_collide)
// t0 / t1 (optional): events bound to the launch itself, completing at its
// start and at its end (no separate event packets in the stream); lds:
// dynamic LDS per workgroup, which the kernel does not use: it caps the
// workgroups per CU (lattice_lds_for_waves)
// cp: cache policy of the population accesses (CP_PLAIN, CP_NT_STORE, CP_NT)
void launch_lattice(hipStream_t s, const Grid& g, const MrtParams& P, const void* src, void* dst, FieldSlot* fields,
                    bool f32, int bx, int cp, const AdjMask& mask, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr,
                    size_t lds = 0);
// The lattice step of a coupled step (fields written, every interior node):
// at most 8 waves per CU, held there by the kernel's register count, so that
// no LDS is taken from the step's chain
void launch_lattice_narrow(hipStream_t s, const Grid& g, const MrtParams& P, const void* src, void* dst,
                           const FieldSlot& fields, bool f32, int bx, int cp, const AdjMask& mask);
// Dynamic LDS that lets at most `waves` waves of bx-thread workgroups share a
// CU of lds_cu bytes of LDS: 0 for waves = 0 (no cap), -1 when no size does.
// No size does when it would exceed lds_blk, the device's LDS per workgroup.
inline long long lattice_lds_for_waves(int bx, int waves, int lds_cu, int lds_blk) {
  if (waves <= 0) return 0;
  const long long wg = waves / (bx / 64) > 0 ? waves / (bx / 64) : 1;
  const long long bytes = ((lds_cu / (wg + 1) + 1) + 511) / 512 * 512;
  return wg * bytes <= lds_cu && bytes <= lds_blk ? bytes : -1;
}
int lattice_tiles_x(const Grid& g, int bx);      // row tiles of one row (single instance)
int lattice_tiles_flat(const Grid& g, int bx);   // tiles of the interior rows taken as one run (batch)
void launch_adj_collide(hipStream_t s, const Grid& g, const MrtParams& P, const IbmParams& p, const GeoBuffers& G,
                        int slot, const FlowBuffers& F, int n, const void* src, void* dst, FieldSlot* fields,
                        bool f32);
// force row of step n (geometry slot, row index row) from the adjoint-collide partials
void launch_force_row(hipStream_t s, const IbmParams& p, const GeoBuffers& G, int slot, const FlowBuffers& F,
                      DevStep* st, int n, int row, int row_len);
void launch_bc_ring(hipStream_t s, const Grid& g, const MrtParams& P, const RingBuffers& R, const void* src,
                    void* dst, FieldSlot* fields, bool f32);
void launch_ring_moments(hipStream_t s, const Grid& g, const RingBuffers& R);
void launch_init_fields(hipStream_t s, const Grid& g, const MrtParams& P, double uu, const double feq[9], void* f,
                        const RingBuffers& R, FieldSlot slot, bool f32);
void launch_moments_all(hipStream_t s, const Grid& g, const MrtParams& P, const void* f, FieldSlot slot, bool f32);
void launch_err_partial(hipStream_t s, const Grid& g, FieldSlot cur, FieldSlot prev, double* partial,
                        int nblocks);
void launch_err_final(hipStream_t s, const double* partial, int nblocks, double* out);
int err_blocks(const Grid& g);
// sum over every node of its nine populations (FP64): rows (ny doubles) gets
// the per-row sums, out[0] the total; nothing else is written
void launch_mass(hipStream_t s, const Grid& g, const MrtParams& P, const void* f, bool f32, double* rows,
                 double* out);
// device address of this GPU's file-scope scratch array of the allocation-free
// mass (room for mass_scratch_rows() row sums), looked up once per device
hipError_t mass_scratch(double** out);
int mass_scratch_rows();
int measure_pull_copy(int nx, int ny, bool f32, int launches, double* avg_us);
void launch_shift_x(hipStream_t s, const Grid& g, int dr, const void* f, void* fshift, int planes, bool f32);
void launch_copy_ring_column(hipStream_t s, const Grid& g, const MrtParams& P, const void* f, int col, double* ring,
                             bool f32);
// f0 ring row `row` (0 or ny-1) from the populations f, every column but `keep`
void launch_copy_ring_row(hipStream_t s, const Grid& g, const MrtParams& P, const void* f, int row, int keep,
                          double* ring, bool f32);
int adj_collide_blocks_per_wing(const IbmParams& p);

// IBM geometry kernels (lbm_ibm.hip), one launch covers `count` consecutive
// steps n0.. into slots slot0..; frames are indexed from fidx0.
// direct: the flow solves by band substitution when the step allows it (no
// explicit inverse); otherwise the inverse is formed
// apart: the caller knows that no two wings' support rectangles meet at these
// steps (their gap was more than APART_GAP nodes a step or two before, and
// wings move less than a node per step): the inverse and the coupling pass
// are not launched, and a step whose rectangles meet raises DERR_WINGS_MET
void launch_geometry(hipStream_t s, const IbmParams& p, const GeoBuffers& G, DevStep* st, int slot0, int n0,
                     int fidx0, int count, bool with_inverse, bool direct = false, bool apart = false);
constexpr int APART_GAP = 4;
size_t geo_solve_lds_bytes(const IbmParams& p, bool direct);   // geometry solve (direct: twisted factors possible)
size_t geo_factor_lds_bytes(const IbmParams& p);   // the plain factors alone (the inverse)
size_t bsolve_lds_bytes(const IbmParams& p);       // band substitution of one wing
size_t geo_couple_lds_total();   // static + dynamic LDS of the coupling workgroup
// Batched launches over the episodes of eps (device array).  The substep of a
// launch is *it.kbase + k (the base lives on the device, so that a captured
// graph of one geometry batch replays for every batch of a long advance).
// masked: skip each episode's adjoint nodes of the substep (its geometry must exist)
// rows: tiles per work item (1, 2 or 4; the items were built for it);
// t1 (optional): an event bound to the launch, completing at its end
void launch_batch_lattice(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields,
                          int bx, int rows, int cp, bool masked, hipEvent_t t1 = nullptr, size_t lds = 0);
void launch_batch_bc_ring(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields);
void launch_batch_adj_collide(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32, bool fields);
// bsolve_lds: LDS of the band substitution (bsolve_lds_bytes, the largest of
// the episodes that solve by it: EpDesc::direct), 0 when none does
void launch_batch_ibm_flow(hipStream_t s, const EpDesc* eps, const BatchItems& it, int k, bool f32,
                           size_t bsolve_lds);
void launch_batch_force_row(hipStream_t s, const EpDesc* eps, const int* kb, int count, int k);
// *p = v (add = false) or *p += v (add = true), one thread, in stream order
void launch_step_base(hipStream_t s, int* p, int v, bool add);
// Geometry of substeps *kb + k0 .. *kb + k0 + count - 1 of every episode with
// wings; items are (episode, b*W + w) for b < count.
// grid: workgroups of the marker pass and the band factor, each looping over
// the items (0: one per item)
void launch_batch_geometry(hipStream_t s, const EpDesc* eps, const int* kb, const int2* items, int n_items,
                           int k0, size_t lds_wing, size_t lds_solve, size_t lds_inv, int inv_chunks, int grid);
// workgroups of the inverse per (step, wing) of a wing of spw segments
int inverse_chunks(int spw);
size_t geo_wing_lds_bytes(int spw);
int bc_ring_blocks(const Grid& g);
int spread_blocks(const IbmParams& p);
// IBM flow kernels of one step (direct: the step's geometry was launched
// with direct, so that its system may be solved by band substitution)
void launch_ibm_flow(hipStream_t s, const IbmParams& p, const MrtParams& P, const GeoBuffers& G, int slot,
                     const FlowBuffers& F, const Grid& g, const void* src, bool f32, bool direct = false,
                     bool apart = false);

}  // namespace lbm
