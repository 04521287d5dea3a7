// lbm_internal.h - device state and kernel launch wrappers (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_hip.h"
#include "lbm_device.h"

namespace lbm {

constexpr int MAXW = LBM_MAX_WINGS;
constexpr int BMAX = 24;          // largest IBM band half-width the solvers accept
constexpr int BOXW = 6;           // nodes per axis of a delta support box (|r| <= 2.5)

// Error codes raised by kernels into DevStep::error (first one wins).
enum DevError : int {
  DERR_NONE = 0,
  DERR_BOX_EDGE = 1,       // support box within 3 nodes of the lattice edge
  DERR_ADJ_OVERFLOW = 2,   // adjoint list capacity exceeded
  DERR_BAND = 3,           // IBM band wider than BMAX
  DERR_WINGS_COUPLED = 4,  // supports of two wings overlap
  DERR_NOT_SPD = 5,        // non-positive pivot in the banded Cholesky
  DERR_MOTION = 6,         // motion source not ready
};

struct WingMotionState {
  double angle, omega, omgac, sa, sau, saac, sb, sbv, sbac;
};

// Per-step state written by the device (one instance per solver).
struct DevStep {
  WingMotionState wing[MAXW];
  double time;
  int n;
  int adj_n;
  int bw;
  int error;
  int error_info;
  int pad;
  double force[MAXW][2];   // normalised per-wing force of the last step
  double moment[MAXW];
  double raw[MAXW][3];     // un-normalised Fx, Fy, M of the last step
};

// Host-known constants of the IBM / motion kernels.
struct IbmParams {
  int W, P, SPW, nseg, npts, cap;
  int nx, ny;
  double scale, ox, oy, uu, dt;
  double wing_x0, wing_dx, wing_y0, xc, yc;
  int motion_source;       // LBM_MOTION_FILE / LBM_MOTION_STREAM
  int movenum;
  int step_count;          // rows of the file-mode step table
  int ncols;
  double motion_dt;
  double body_extra_v[MAXW], body_extra_s[MAXW], body_extra_pos[MAXW];
};

// Device buffers of the IBM path.
struct IbmBuffers {
  const double* xb0; const double* yb0;   // body-frame markers (npts)
  double* xb1; double* yb1;               // moved markers (npts)
  double* xb2; double* yb2;               // segment midpoints (nseg)
  const double* arc;                      // segment lengths / scale (nseg)
  double* ub; double* vb;                 // rigid-body velocity at midpoints (nseg)
  int* box;                               // i1,i2,j1,j2 per segment (0-based, nseg*4)
  double* phx; double* phy;               // delta weights per box column/row (nseg*BOXW)
  int* seg_cnt; int* seg_off;             // adjoint nodes claimed per segment / offsets (nseg+1)
  int* owner;                             // lattice map: claiming segment (nx*ny, INT_MAX = none)
  int* aidx;                              // lattice map: adjoint index (nx*ny, -1 = none)
  int* kmin; int* kmax;                   // lattice maps: lowest / highest band key of the
                                          // segments whose support holds the node
  int* adj_ij;                            // (i, j) 0-based per adjoint node (cap*2)
  int* adj_seg;                           // claiming segment per adjoint node (cap)
  double* adj_rho; double* adj_us; double* adj_vs;   // moments at adjoint nodes (cap)
  double* adj_u; double* adj_v;           // corrected velocity (cap)
  double* adj_gx; double* adj_gy;         // body force density 2 rho g (cap)
  double* adj_mom;                        // moment contribution -x0 t2 + y0 t1 (cap)
  int* prev_ij; int* prev_n;              // adjoint list of the previous step (cleanup)
  const int* seg_of_pos;                  // zig-zag order: segment at band position (nseg)
  const int* pos_of_seg;                  // band position of each segment (nseg)
  double* band;                           // band matrix, W * SPW * (BMAX+1)
  double* bcr_ws;                         // block cyclic reduction workspace per wing
  double* rhs;                            // 2 * nseg right-hand side, solution in place
  const double* step_table;               // file mode: step_count * ncols
  const double* stream_rows;              // stream mode: start row, target row (2*LBM_MAX_MOTION_COLS)
  DevStep* st;
  double* rows;                           // force rows of the current call
  double* mrows;                          // airfoil_motion rows of the current call
  double* row_sum;                        // sum of force rows
};

struct RingBuffers {
  double* left;     // f0 ring column i=0, 9 * ny
  double* right;    // f0 ring column i=nx-1, 9 * ny
  double* mom;      // rho,u,v of left column then right column: 6 * ny
};

struct FieldSlot {
  double* rho; double* u; double* v;
};

// lattice kernels (lbm_lattice.hip)
void launch_lattice(hipStream_t s, const Grid& g, const MrtParams& P, const double* src, double* dst,
                    FieldSlot* fields);
void launch_adj_collide(hipStream_t s, const Grid& g, const MrtParams& P, const IbmBuffers& B, int cap,
                        const double* src, double* dst, FieldSlot* fields);
void launch_bc_ring(hipStream_t s, const Grid& g, const MrtParams& P, const RingBuffers& R,
                    const double* src, double* dst, FieldSlot* fields);
void launch_ring_moments(hipStream_t s, const Grid& g, const RingBuffers& R);
void launch_init_fields(hipStream_t s, const Grid& g, double uu, const double feq[9], double* f,
                        const RingBuffers& R, FieldSlot slot);
void launch_moments_all(hipStream_t s, const Grid& g, const double* f, FieldSlot slot);
void launch_err_partial(hipStream_t s, const Grid& g, FieldSlot cur, FieldSlot prev, double* partial,
                        int nblocks);
void launch_err_final(hipStream_t s, const double* partial, int nblocks, double* out);
int err_blocks(const Grid& g);
void launch_shift_x(hipStream_t s, const Grid& g, int dr, const double* f, double* fshift, int planes);
void launch_copy_ring_column(hipStream_t s, const Grid& g, const double* f, int col, double* ring);

// IBM kernels (lbm_ibm.hip)
void launch_motion_markers(hipStream_t s, const IbmParams& p, const IbmBuffers& B, int n);
void launch_adjoint_build(hipStream_t s, const IbmParams& p, const IbmBuffers& B);
void launch_clear_adjoint(hipStream_t s, const IbmParams& p, const IbmBuffers& B);
void launch_ibm_gather(hipStream_t s, const IbmParams& p, const IbmBuffers& B, const Grid& g,
                       const double* src);
void launch_ibm_assemble(hipStream_t s, const IbmParams& p, const IbmBuffers& B);
void launch_ibm_wing(hipStream_t s, const IbmParams& p, const IbmBuffers& B);
void launch_forces(hipStream_t s, const IbmParams& p, const IbmBuffers& B, int slot, int row_len);
size_t ibm_solve_lds_bytes(int spw);
size_t ibm_bcr_ws_doubles(int W);

}  // namespace lbm
