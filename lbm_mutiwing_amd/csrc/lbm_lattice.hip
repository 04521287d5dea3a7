// lbm_lattice.hip - full-lattice kernels of the FP64 time step on gfx950.
//
// k_lattice fuses streaming (solution.f90:440-466), physical_variables
// (solution.f90:483-554) and colission (solution.f90:585-655) for every
// interior node: 9 pulled loads and 9 stores of one SoA plane each, i.e. the
// 144 B/node floor of a two-buffer FP64 D2Q9 step.  Nodes that the immersed
// boundary touches this step are collided here with zero body force and
// collided again by k_adj_collide once the IBM correction is known, so the
// full-lattice pass never waits for the IBM chain.
#include "lbm_internal.h"

namespace lbm {

constexpr int LAT_BLOCK = 256;

// One thread per node of a row tile; tiles start at i = 0 so that every
// wavefront stores whole, aligned 512-byte spans of each population plane
// (the ring columns i = 0 and i = nx-1 are skipped and left to k_bc_ring).
// Workgroups follow the hardware's round-robin placement over the XCDs,
// which measured faster than remapping tiles to XCD-contiguous runs.
template <bool FIELDS>
__global__ __launch_bounds__(LAT_BLOCK) void k_lattice(Grid g, MrtParams P, const double* __restrict__ src,
                                                       double* __restrict__ dst, FieldSlot fs, int tiles_x) {
  const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - ty * tiles_x;
  const int i = tx * LAT_BLOCK + (int)threadIdx.x;
  const int j = 1 + ty;
  if (i < 1 || i > g.nx - 2) return;
  double f[9];
  pull_node(src, g, i, j, f);
  double rho, u, v;
  node_moments(f, rho, u, v);
  double out[9];
  collide_node(P, f, rho, u, v, 0.0, 0.0, out);
  const long long o = (long long)j * g.nx + i;
#pragma unroll
  for (int k = 0; k < 9; k++) dst[(long long)k * g.ps + o] = out[k];
  if (FIELDS) {
    fs.rho[o] = rho;
    fs.u[o] = u;
    fs.v[o] = v;
  }
}

void launch_lattice(hipStream_t s, const Grid& g, const MrtParams& P, const double* src, double* dst,
                    FieldSlot* fields) {
  const int tiles_x = (g.nx + LAT_BLOCK - 1) / LAT_BLOCK;
  dim3 grid(tiles_x * (g.ny - 2)), block(LAT_BLOCK);
  if (fields)
    hipLaunchKernelGGL(k_lattice<true>, grid, block, 0, s, g, P, src, dst, *fields, tiles_x);
  else
    hipLaunchKernelGGL(k_lattice<false>, grid, block, 0, s, g, P, src, dst, FieldSlot{}, tiles_x);
}

// Re-collide the adjoint nodes with the IBM-corrected velocity and the body
// force 2 rho g (boundary_condition.f90:1645-1659 then solution.f90:599-650).
template <bool FIELDS>
__global__ __launch_bounds__(256) void k_adj_collide(Grid g, MrtParams P, IbmBuffers B, const double* __restrict__ src,
                                                     double* __restrict__ dst, FieldSlot fs) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (B.st->error != 0 || a >= B.st->adj_n) return;
  const int i = B.adj_ij[2 * a], j = B.adj_ij[2 * a + 1];
  double f[9];
  pull_node(src, g, i, j, f);
  double rho, u0, v0;
  node_moments(f, rho, u0, v0);
  const double u = B.adj_u[a], v = B.adj_v[a];
  double out[9];
  collide_node(P, f, rho, u, v, B.adj_gx[a], B.adj_gy[a], out);
  const long long o = (long long)j * g.nx + i;
#pragma unroll
  for (int k = 0; k < 9; k++) dst[(long long)k * g.ps + o] = out[k];
  if (FIELDS) {
    fs.u[o] = u;
    fs.v[o] = v;
  }
}

void launch_adj_collide(hipStream_t s, const Grid& g, const MrtParams& P, const IbmBuffers& B, int cap,
                        const double* src, double* dst, FieldSlot* fields) {
  if (cap <= 0) return;
  dim3 grid((cap + 255) / 256), block(256);
  if (fields)
    hipLaunchKernelGGL(k_adj_collide<true>, grid, block, 0, s, g, P, B, src, dst, *fields);
  else
    hipLaunchKernelGGL(k_adj_collide<false>, grid, block, 0, s, g, P, B, src, dst, FieldSlot{});
}

// ------------------------------------------------------------------ far field

struct NodeVals {
  double rho, u, v;
};

__device__ __forceinline__ NodeVals interior_vals(const double* __restrict__ src, const Grid& g, int i, int j,
                                                  double f[9]) {
  pull_node(src, g, i, j, f);
  NodeVals r;
  node_moments(f, r.rho, r.u, r.v);
  return r;
}

// Characteristic update of left-column node (0, j) (boundary_condition.f90:418-439).
__device__ NodeVals lodi_left(const double* __restrict__ src, const Grid& g, const RingBuffers& R, double cs,
                              int j) {
  const int ny = g.ny;
  const double* mr = R.mom;  // rho[0..ny), u[ny..2ny), v[2ny..3ny) for the left column
  const double trj = mr[j], tuj = mr[ny + j], tvj = mr[2 * ny + j];
  double f2[9], f3[9];
  const NodeVals n2 = interior_vals(src, g, 1, j, f2);
  const NodeVals n3 = interior_vals(src, g, 2, j, f3);
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
__device__ NodeVals lodi_right(const double* __restrict__ src, const Grid& g, const RingBuffers& R, double cs,
                               int j) {
  const int ny = g.ny, nx = g.nx;
  const double* mr = R.mom + 3 * ny;
  const double trj = mr[j], tuj = mr[ny + j], tvj = mr[2 * ny + j];
  double f2[9], f3[9];
  const NodeVals n2 = interior_vals(src, g, nx - 2, j, f2);
  const NodeVals n3 = interior_vals(src, g, nx - 3, j, f3);
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
template <bool FIELDS>
__global__ __launch_bounds__(256) void k_bc_ring(Grid g, MrtParams P, RingBuffers R, const double* __restrict__ src,
                                                 double* __restrict__ dst, FieldSlot fs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = g.nx, ny = g.ny, nside = ny - 2;
  const int total = 2 * nside + 2 * nx;
  if (t >= total) return;
  const double cs = P.cs;
  double out[9], f0n[9];
  int bi, bj;
  NodeVals b, nv;
  if (t < nside) {  // left (boundary_condition.f90:413-483)
    bj = 1 + t;
    bi = 0;
    b = lodi_left(src, g, R, cs, bj);
    nv = interior_vals(src, g, 1, bj, f0n);
  } else if (t < 2 * nside) {  // right (boundary_condition.f90:495-567)
    bj = 1 + t - nside;
    bi = nx - 1;
    b = lodi_right(src, g, R, cs, bj);
    nv = interior_vals(src, g, nx - 2, bj, f0n);
  } else {  // top (boundary_condition.f90:657-723) / bottom (784-850) walls
    const int tt = t - 2 * nside;
    const bool top = tt < nx;
    bi = top ? tt : tt - nx;
    bj = top ? ny - 1 : 0;
    const int nj = top ? ny - 2 : 1;  // node inside the wall
    if (bi == 0) {
      nv = lodi_left(src, g, R, cs, nj);
#pragma unroll
      for (int k = 0; k < 9; k++) f0n[k] = R.left[(long long)k * ny + nj];
    } else if (bi == nx - 1) {
      nv = lodi_right(src, g, R, cs, nj);
#pragma unroll
      for (int k = 0; k < 9; k++) f0n[k] = R.right[(long long)k * ny + nj];
    } else {
      nv = interior_vals(src, g, bi, nj, f0n);
    }
    b.rho = nv.rho;
    b.u = 0.0;
    b.v = 0.0;
  }
  bc_extrapolate(P, b.rho, b.u, b.v, nv.rho, nv.u, nv.v, f0n, out);
  const long long o = (long long)bj * nx + bi;
#pragma unroll
  for (int k = 0; k < 9; k++) dst[(long long)k * g.ps + o] = out[k];
  if (FIELDS) {
    fs.rho[o] = b.rho;
    fs.u[o] = b.u;
    fs.v[o] = b.v;
  }
}

void launch_bc_ring(hipStream_t s, const Grid& g, const MrtParams& P, const RingBuffers& R, const double* src,
                    double* dst, FieldSlot* fields) {
  const int total = 2 * (g.ny - 2) + 2 * g.nx;
  dim3 grid((total + 255) / 256), block(256);
  if (fields)
    hipLaunchKernelGGL(k_bc_ring<true>, grid, block, 0, s, g, P, R, src, dst, *fields);
  else
    hipLaunchKernelGGL(k_bc_ring<false>, grid, block, 0, s, g, P, R, src, dst, FieldSlot{});
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

__global__ void k_init_fields(Grid g, double uu, Feq9 feq, double* f, RingBuffers R, FieldSlot slot) {
  const long long n = (long long)g.nx * g.ny;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 9; k++) f[(long long)k * g.ps + q] = feq.v[k];
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
  }
}

void launch_init_fields(hipStream_t s, const Grid& g, double uu, const double feq[9], double* f,
                        const RingBuffers& R, FieldSlot slot) {
  Feq9 fe;
  for (int k = 0; k < 9; k++) fe.v[k] = feq[k];
  hipLaunchKernelGGL(k_init_fields, dim3(2048), dim3(256), 0, s, g, uu, fe, f, R, slot);
}

__global__ void k_moments_all(Grid g, const double* __restrict__ f, FieldSlot slot) {
  const long long n = (long long)g.nx * g.ny;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    double fo[9];
#pragma unroll
    for (int k = 0; k < 9; k++) fo[k] = f[(long long)k * g.ps + q];
    double rho, u, v;
    node_moments(fo, rho, u, v);
    slot.rho[q] = rho;
    slot.u[q] = u;
    slot.v[q] = v;
  }
}

void launch_moments_all(hipStream_t s, const Grid& g, const double* f, FieldSlot slot) {
  hipLaunchKernelGGL(k_moments_all, dim3(2048), dim3(256), 0, s, g, f, slot);
}

// erranalysis (solution.f90:77-130): two fixed-order reductions over the
// interior; one partial per block, summed in block order by k_err_final.
constexpr int ERR_ROWS_PER_BLOCK = 8;

int err_blocks(const Grid& g) { return (g.ny - 2 + ERR_ROWS_PER_BLOCK - 1) / ERR_ROWS_PER_BLOCK; }

__global__ __launch_bounds__(256) void k_err_partial(Grid g, FieldSlot cur, FieldSlot prev, double* partial) {
  __shared__ double s1[256], s2[256];
  double a1 = 0.0, a2 = 0.0;
  const int j0 = 1 + blockIdx.x * ERR_ROWS_PER_BLOCK;
  for (int j = j0; j < j0 + ERR_ROWS_PER_BLOCK && j <= g.ny - 2; j++)
    for (int i = 1 + threadIdx.x; i <= g.nx - 2; i += blockDim.x) {
      const long long o = (long long)j * g.nx + i;
      const double u = cur.u[o], v = cur.v[o];
      const double du = u - prev.u[o], dv = v - prev.v[o];
      a1 = a1 + du * du + dv * dv;
      a2 = a2 + u * u + v * v;
    }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
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
  hipLaunchKernelGGL(k_err_partial, dim3(nblocks), dim3(256), 0, s, g, cur, prev, partial);
}

__global__ void k_err_final(const double* partial, int nblocks, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t1 = 0.0, t2 = 0.0;
  for (int b = 0; b < nblocks; b++) {
    t1 += partial[2 * b];
    t2 += partial[2 * b + 1];
  }
  out[0] = sqrt(t1) / (sqrt(t2) + 1e-30);
}

void launch_err_final(hipStream_t s, const double* partial, int nblocks, double* out) {
  hipLaunchKernelGGL(k_err_final, dim3(1), dim3(64), 0, s, partial, nblocks, out);
}

// move_mesh column shift (move_mesh.f90:29-57 dr=+1, 185-213 dr=-1): out(i) =
// in(i+dr) where that column exists, else the column is kept.
__global__ void k_shift_x(Grid g, int dr, const double* __restrict__ in, double* __restrict__ out, int planes) {
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

void launch_shift_x(hipStream_t s, const Grid& g, int dr, const double* f, double* fshift, int planes) {
  hipLaunchKernelGGL(k_shift_x, dim3(4096), dim3(256), 0, s, g, dr, f, fshift, planes);
}

__global__ void k_copy_ring_column(Grid g, const double* __restrict__ f, int col, double* ring) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 9 * g.ny) return;
  const int k = t / g.ny, j = t - k * g.ny;
  ring[(long long)k * g.ny + j] = f[(long long)k * g.ps + (long long)j * g.nx + col];
}

void launch_copy_ring_column(hipStream_t s, const Grid& g, const double* f, int col, double* ring) {
  hipLaunchKernelGGL(k_copy_ring_column, dim3((9 * g.ny + 255) / 256), dim3(256), 0, s, g, f, col, ring);
}

}  // namespace lbm
