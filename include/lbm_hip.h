/*
 * lbm_hip.h - C ABI of the MI355X (gfx950) time step of the D2Q9 MRT
 * lattice-Boltzmann + implicit immersed-boundary multi-airfoil solver.
 *
 * The reference program (Fortran 90 + OpenACC) has no library boundary: its
 * step loop calls plain subroutines that share module globals.  Each entry
 * point below replaces one of those call sites; the file:line citations name
 * the reference code whose behaviour the entry point reproduces.  The Fortran
 * host binds these symbols with iso_c_binding (bind(C, name=...)), in the
 * same style as its existing libzmq binding (zmq_comm.f90:23-81).
 *
 * Conventions
 *   - Every function returns 0 on success and a non-zero LBM_E_* code on
 *     failure; lbm_last_error() returns the message of the last failure on a
 *     handle.  Host buffers are owned by the caller and only read or written
 *     during the call.  A handle owns all device memory it allocates.
 *   - One host thread per handle.  All work of a handle is ordered on its own
 *     HIP stream(s); every call returns after that work has completed.
 *   - Lattice arrays are column-major like the reference's mesh arrays
 *     (main.f90:167): element (i, j) of an Nx x Ny field is at i + Nx*j
 *     (0-based), population k (0..8) of f at (i + Nx*j) + Nx*Ny*k.
 *   - Indices given to or returned by the API are 1-based where they mirror a
 *     reference quantity (step number n, adjoint node i/j), as in Fortran.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_MAX_WINGS 8                 /* main.f90:17 max_wings_supported        */
#define LBM_MOTION_COLS_PER_WING 9      /* main.f90:19 motion_columns_per_wing    */
#define LBM_MAX_MOTION_COLS (1 + LBM_MOTION_COLS_PER_WING * LBM_MAX_WINGS)
#define LBM_MAX_FORCE_ROW (4 + 12 * LBM_MAX_WINGS) /* main.f90:71               */

#define LBM_MOTION_FILE 1               /* main.f90:88 motion_source_file          */
#define LBM_MOTION_STREAM 2             /* main.f90:89 motion_source_zmq           */

#define LBM_PRECISION_FP64 0            /* populations stored and collided in FP64 */
#define LBM_PRECISION_FP32_SHIFTED 1    /* FP32 storage of f - w_i                 */

#define LBM_OK 0
#define LBM_E_INVALID 1                 /* bad argument / configuration            */
#define LBM_E_STATE 2                   /* call not valid in the current state     */
#define LBM_E_DEVICE 3                  /* HIP runtime error                       */
#define LBM_E_RANGE 4                   /* geometry left the lattice / capacity    */
#define LBM_E_NUMERIC 5                 /* IBM system not positive definite        */

/*
 * Run-time configuration.  In the reference all of these are compile-time
 * parameters of module com_date (main.f90:4-159) or derived in ini_mesh
 * (ini_mesh.f90:15-67).
 */
typedef struct lbm_config {
    int32_t nx, ny;            /* lattice nodes Nx, Ny (ini_mesh.f90:22-23)              */
    int32_t wing_count;        /* 1..LBM_MAX_WINGS (main.f90:18)                          */
    int32_t points_per_wing;   /* markers per wing, closed contour (main.f90:21)          */
    double scale;              /* lattice step in chord units (main.f90:32)              */
    double ox, oy;             /* coordinates of node (1,1) (ini_mesh.f90:36-37)          */
    double uu;                 /* lattice inflow velocity UU (main.f90:12)               */
    double re;                 /* Reynolds number (main.f90:106); nu = UU/scale/Re        */
    double s[9];               /* MRT rates (ini_mesh.f90:59-67); s[7], s[8] <= 0 means
                                  1/tau with tau = 3 nu + 0.5 (ini_mesh.f90:30-31)         */
    double wing_x0;            /* wing_initial_x_origin (main.f90:34)                     */
    double wing_dx;            /* wing_initial_x_spacing (main.f90:35)                    */
    double wing_y0;            /* wing_initial_y_origin (main.f90:36)                     */
    double xc, yc;             /* flow-centre offset (main.f90:107)                       */
    double xp, yp;             /* rotation-axis offset of the marker files (main.f90:108) */
    int32_t motion_source;     /* LBM_MOTION_FILE or LBM_MOTION_STREAM (driver.f90:9-31)  */
    int32_t precision;         /* LBM_PRECISION_*                                         */
    const double *xb0;         /* wing_count*points_per_wing marker x (chord units), as
                                  read by read_boundary (airfoil.f90:229)                  */
    const double *yb0;         /* same for y                                              */
} lbm_config;

typedef struct lbm_handle lbm_handle;

/* Allocate a solver instance on the current HIP device.  Replaces ini_mesh +
 * the allocations of init (ini_mesh.f90:1-80, solution.f90:67-69).  The
 * marker arrays are copied.  The instance starts in the RESET state. */
int lbm_create(const lbm_config *cfg, lbm_handle **out);
int lbm_destroy(lbm_handle *h);

/* Episode reset: init + read_boundary + zeroed rigid-body state
 * (solution.f90:1-75, airfoil.f90:207-242, driver.f90:205-258).  The next
 * step to run is n = 1.  In stream mode the motion segment becomes the zero
 * row at time 0 (driver.f90:120-124). */
int lbm_reset(lbm_handle *h);

/* File-mode motion: the raw motion_params rows (row-major, nrows x ncols with
 * ncols = 1 + 9*wing_count), validated and resampled on the solver step
 * dt = scale*UU (mode.f90:104-178, 327-412).  end_time receives
 * motion_end_time (mode.f90:163). */
int lbm_set_motion_table(lbm_handle *h, const double *rows, int32_t nrows,
                         int32_t ncols, double *end_time);

/* Stream-mode motion segment: start row and target row of 1 + 9*wing_count
 * columns, evaluated per step by lookup_stream_motion_row
 * (mode.f90:414-449).  Built by the host from a STEP command with
 * lbm_build_stream_target_row. */
int lbm_set_stream_segment(lbm_handle *h, const double *start_row,
                           const double *target_row, int32_t ncols);

/* build_stream_target_row (mode.f90:287-325): target row for a STEP with
 * per-wing (domega, dxv, dyv) increments in delta_row[1..3*wing_count]
 * (delta_row[0] is ignored).  Host-only arithmetic, no device work.
 * Returns LBM_E_INVALID when target_time is before the segment start. */
int lbm_build_stream_target_row(int32_t wing_count, const double *start_row,
                                double target_time, const double *delta_row,
                                double *target_row);

/* Run steps n .. n_last (inclusive), where n is the next step of the
 * instance: per step erranalysis (n >= 2), evolution and the move_mesh
 * check (main.f90:374-412, driver.f90:362-408, solution.f90:272-422).
 *   force_rows   optional, (n_last-n+1) x row_len doubles: one force.dat row
 *                per step (solution.f90:352-378)
 *   motion_rows  optional, (n_last-n+1) x (1+3*wing_count): one
 *                airfoil_motion.dat row per step (solution.f90:390-401)
 *   row_sum      optional, row_len doubles: sum of the force rows of this call
 *                (latest_force_sum_row, solution.f90:379-385)
 *   last_row     optional, row_len doubles: force row of step n_last
 *   err          optional: erranalysis value at the last step (solution.f90:77-130)
 * row_len = 4 + 12*wing_count.  n_last < n runs nothing. */
int lbm_advance(lbm_handle *h, int32_t n_last, double *force_rows,
                double *motion_rows, double *row_sum, double *last_row,
                double *err);

/* State counters: next step n, movenum (main.f90:143), flag2 (post-move_mesh
 * window, main.f90:401-412), force row length, adjoint node count. */
int lbm_get_state(lbm_handle *h, int32_t *n_next, int32_t *movenum,
                  int32_t *flag2, int32_t *row_len, int32_t *adjoint_n);

/* Macroscopic fields after the last step, Nx*Ny each (any may be NULL):
 * rho, u, v, pre = rho/3 (solution.f90:546-549, IBM-corrected u/v at the
 * adjoint nodes, boundary_condition.f90:1650-1651, far-field values on the
 * boundary ring, boundary_condition.f90:403-853). */
int lbm_get_fields(lbm_handle *h, double *rho, double *u, double *v,
                   double *pre);

/* Populations f (9*Nx*Ny, FP64 values; FP32-shifted instances return
 * w_k + stored deviation) and the persistent f0 boundary ring of the
 * streaming buffer (same layout; columns 1 and Nx and rows 1 and Ny, as
 * move_mesh leaves them, move_mesh.f90:30-34; interior values are
 * unspecified). */
int lbm_get_populations(lbm_handle *h, double *f, double *f0);
/* Overwrite f (interior and ring) and the f0 ring; u/v of the previous step
 * become the moments of f (test and restart hook, solution.f90:1721-1851). */
int lbm_set_populations(lbm_handle *h, const double *f, const double *f0);

/* Mass-conservation check (no counterpart in the reference, which checks no
 * conservation; the lattice-wide FP64 reduction follows the pattern of
 * erranalysis, solution.f90:77-130): *mass = the sum over every node of the
 * lattice of its nine populations f after the last step (FP32-shifted
 * instances: w_k + stored deviation, summed in FP64).  Reduced on the device
 * in a fixed order, so the value is identical run to run.  Reads the whole
 * population buffer once (604 MB at 4096 x 2048 in FP64). */
int lbm_get_mass(lbm_handle *h, double *mass);

/* Allocation-free variant of lbm_get_mass, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  f points to caller-owned device
 * populations in compact SoA planes (9 planes of nx*ny values, population
 * k of node (i, j) at i + nx*j + nx*ny*k; the library's own buffers space
 * their planes nx*ny rounded up to a multiple of 64 apart): FP64 values for
 * LBM_PRECISION_FP64, FP32 deviations f - w_k for
 * LBM_PRECISION_FP32_SHIFTED.  The kernels are those of lbm_get_mass (same
 * value, bit for bit), enqueued on `stream` (a hipStream_t; NULL is the null
 * stream) with no synchronisation: *mass is a device address, written when
 * the stream reaches the call.  The per-row sums go to a fixed device array
 * of the library (one per GPU, 32768 doubles, held until the process ends),
 * so ny may be at most 32768 (LBM_E_RANGE otherwise), nothing is allocated
 * or freed, and calls that overlap on one GPU race.  Errors are reported by
 * lbm_last_error(NULL, ...). */
int lbm_get_mass_noalloc(const void *f, int32_t nx, int32_t ny, int32_t precision,
                         double *mass, void *stream);

/* Markers after the last step: xb1/yb1 (wing_count*points_per_wing) and
 * segment midpoints xb2/yb2 (airfoil.f90:261-270); any may be NULL. */
int lbm_get_markers(lbm_handle *h, double *xb1, double *yb1, double *xb2,
                    double *yb2);

/* IBM data of the last step: adjoint_p (1-based i, j; 2*capacity int32,
 * i first), adjoint_wing_id (1-based), per-wing normalised force (2*8) and
 * moment (8) (boundary_condition.f90:1645-1692, solution.f90:330-333). */
int lbm_get_ibm(lbm_handle *h, int32_t capacity, int32_t *adjoint_ij,
                int32_t *wing_id, double *wing_force, double *wing_moment);

/* Restart hook (readflowfield, solution.f90:1721-1851, and the restart branch
 * of main.f90:343-347): the next step is n_next and the frame shift is
 * movenum; the field history of erranalysis is dropped (err reads 0 for the
 * next two steps) and flag2 is 1, as after init.  Call it before
 * lbm_set_populations, which records the moments of f as the fields of step
 * n_next - 1. */
int lbm_set_step(lbm_handle *h, int32_t n_next, int32_t movenum);

/* Rigid-body displacement hook: body_rel_pos(wing) (main.f90:77).  In the
 * reference only the (disabled) dynamics solver moves it; setting |pos| >= 1
 * for wing 1 triggers move_mesh after the next step (main.f90:393-411). */
int lbm_set_body_rel_pos(lbm_handle *h, int32_t wing, double pos);

/* Rigid-body state body_rel_* of all wings, 11*wing_count doubles in the
 * order of the restart file (solution.f90:1663-1665): a, a_prev, v, v_prev,
 * s, s_prev, pos, theta_prev, theta, temp_theta_prev, temp_theta, each
 * wing_count long.  lbm_reset zeroes it (driver.f90:218-228); setting it
 * recomputes body_extra_* (mode.f90:473-495), as readflowfield does. */
#define LBM_BODY_REL_FIELDS 11
int lbm_get_body_rel(lbm_handle *h, double *rel);
int lbm_set_body_rel(lbm_handle *h, const double *rel);

/* Rigid-body coupling of wing 1 (boundary_condition.f90:1694-1719, rk4
 * :3168-3202; switches and constants main.f90:98, 122-125, 141-151).  When
 * enabled, each step's IBM force drives wing 1 along x: body_rel_a(1) =
 * Fx / m with Fx the raw x force summed over all wings and m = solid_den *
 * area / scale^2 (closed_boundary) or solid_den / scale^2; v, s and pos
 * follow by RK4 (rk4 != 0, the reference's if_rk4) or explicit Euler, and
 * theta by the pitch equation with p1 = par1*UU^2*scale^2, p2 = par2*scale^4
 * and the moment of wing 1.  body_extra_* follow, so the next step's markers,
 * their velocity and the force row's motion columns move with the body, and
 * |body_rel_pos(1)| >= 1 triggers move_mesh.  A step then depends on the
 * previous step's force: lbm_advance runs such an instance one step at a
 * time (geometry not ahead of the flow), and lbm_batch_advance rejects it.
 * enable = 0 (the reference's setting) restores the prescribed-motion path. */
typedef struct lbm_dynamics {
    int32_t enable;            /* solve_dynamic_equations (main.f90:124-125)   */
    int32_t closed_boundary;   /* main.f90:122-123                             */
    int32_t rk4;               /* if_rk4 (boundary_condition.f90:1527-1529)    */
    int32_t reserved;          /* 0                                            */
    double solid_den;          /* main.f90:145                                 */
    double area;               /* area_airfoil (main.f90:147)                  */
    double par1, par2;         /* main.f90:150-151                             */
    double alphad;             /* radians (main.f90:98)                        */
} lbm_dynamics;
int lbm_set_dynamics(lbm_handle *h, const lbm_dynamics *d);

/* How the IBM system S y = c of a step is solved at flow time (no
 * counterpart in the reference, which solves its dense system by LU,
 * boundary_condition.f90:1618-1640; every mode gives its solution to
 * rounding): from the explicit inverse S^-1, formed by the geometry ahead of
 * the flow (LBM_IBM_SOLVE_INVERSE); by band substitution with the factors of
 * S in steps with rigid-body coupling, whose geometry is on the step's
 * critical path, and from the inverse otherwise (LBM_IBM_SOLVE_COUPLED, the
 * default); or by band substitution in every step (LBM_IBM_SOLVE_SUBSTITUTION:
 * no inverse is formed, which suits instances advanced together by
 * lbm_batch_advance, where the geometry of all episodes competes with the
 * lattice kernel; a single instance with prescribed motion is faster with
 * the inverse).  Steps whose wings' supports meet use the inverse in every
 * mode.  The batched advance solves each instance in its own mode, so an
 * instance gives the same bits alone and in a batch.  LBM_DIRECT_SOLVE=<mode>
 * in the environment sets the default at lbm_create. */
#define LBM_IBM_SOLVE_INVERSE 0
#define LBM_IBM_SOLVE_COUPLED 1
#define LBM_IBM_SOLVE_SUBSTITUTION 2
int lbm_set_ibm_solve(lbm_handle *h, int32_t mode);

/* Measurement hook (no counterpart in the reference): when enabled, every
 * launch of the fused lattice kernel is bracketed by HIP events on the
 * instance's stream; lbm_get_timing returns the summed kernel time (ms) and
 * the launch count since the last enable, and resets nothing. */
int lbm_set_timing(lbm_handle *h, int32_t enable);
int lbm_get_timing(lbm_handle *h, double *lattice_ms, int32_t *launches);

/* Measurement hook (no counterpart in the reference): the average time of
 * `launches` launches of the lattice kernel's memory pattern alone (each
 * interior node of an nx x ny lattice pulls its nine populations from its
 * neighbours and writes them, in the storage of `precision`, with no
 * arithmetic), on the current device, in its own buffers, which it
 * allocates and frees: the fastest over the cache policies and wave caps the
 * lattice kernel's autotuner tries.  The ceiling the lattice kernel's time
 * is compared with. */
int lbm_measure_pull_copy(int32_t nx, int32_t ny, int32_t precision, int32_t launches, double *avg_us);

/* Batched advance (no counterpart in the reference, which runs one episode
 * per process; for hosts that serve many episodes on one GPU).  Advances each
 * of the count instances in hs by `steps` steps from its own next step, with
 * one launch per kernel phase over all instances: every launch covers a work
 * list of (instance, block) items, so many small, uneven lattices fill the GPU
 * together.  Each instance ends in exactly the state lbm_advance would leave
 * it in.  All instances must share the device and the precision, and none may
 * have a pending move_mesh (|body_rel_pos(1)| >= 1) or the rigid-body
 * coupling switched on (lbm_set_dynamics; such an instance steps alone).
 *   row_sums, last_rows  optional, count x LBM_MAX_FORCE_ROW doubles: the
 *                        row_sum / last_row of lbm_advance for instance e at
 *                        e*LBM_MAX_FORCE_ROW (its row_len entries, rest zero)
 *   errs                 optional, count doubles
 *   status               optional, count int32: 0, or the LBM_E_* code of an
 *                        instance whose step failed (lbm_last_error on it)
 * The outputs of the instances without error are filled in either case; the
 * return value is 0 or the code of the first failing instance (also named by
 * lbm_batch_last_error).
 * Launch graphs: an advance of at most 64 steps, and with wings at most one
 * geometry batch (the 1-16 steps whose markers, adjoint lists and inverse
 * IBM matrices are computed together), whose (instances, steps, work lists)
 * were seen once before on this batch object is captured as a HIP graph on
 * its second occurrence and replayed from then on (up to eight graphs are
 * kept, least recently used evicted); other advances and first occurrences
 * are issued launch by launch.  LBM_BATCH_GRAPHS in the environment, read
 * by lbm_batch_create: 0 turns capture off; 2 also replays, in every longer
 * advance, all geometry batches but the first and the last two from one
 * captured graph of one geometry batch (16 steps without wings) per
 * composition; 1 is the default.  Results are identical in every mode. */
typedef struct lbm_batch lbm_batch;
int lbm_batch_create(lbm_batch **out);
int lbm_batch_destroy(lbm_batch *b);
int lbm_batch_advance(lbm_batch *b, lbm_handle *const *hs, int32_t count,
                      int32_t steps, double *row_sums, double *last_rows,
                      double *errs, int32_t *status);
int lbm_batch_last_error(const lbm_batch *b, char *buf, size_t n);

/* Autotuner result (no counterpart in the reference): the tile width of the
 * fused lattice kernel (128, 256 or 512 nodes per workgroup) and its cap on
 * resident waves per CU (0 none, 24, 16, 12 or 8; set through unused
 * dynamic LDS),
 * chosen by timing every pair on the instance's own lattice at its first
 * lbm_advance (0 before that; LBM_LATTICE_TILE=<width> and
 * LBM_LATTICE_WAVES=<n> in the environment fix them).  lbm_batch_get_tuning
 * gives the width and the number of tiles per workgroup (1, 2 or 4) a batch
 * chose together on its first lbm_batch_advance (LBM_BATCH_ROWS=<n> fixes
 * the latter), then the wave cap for that shape.  Any pointer may be NULL. */
int lbm_get_tuning(const lbm_handle *h, int32_t *lattice_tile,
                   int32_t *waves_per_cu);
int lbm_batch_get_tuning(const lbm_batch *b, int32_t *lattice_tile,
                         int32_t *rows_per_item, int32_t *waves_per_cu);

/* Cache policy of the lattice kernel's population loads and stores, chosen
 * by the same timing (no counterpart in the reference): 0 ordinary, 1
 * non-temporal stores, 2 non-temporal loads and stores (a batch times all
 * three with its wave cap).  LBM_LATTICE_CP=<0|1|2> in the environment fixes it. */
#define LBM_CACHE_PLAIN 0
#define LBM_CACHE_NT_STORE 1
#define LBM_CACHE_NT 2
int lbm_get_lattice_policy(const lbm_handle *h, int32_t *cache_policy);
int lbm_batch_get_lattice_policy(const lbm_batch *b, int32_t *cache_policy);

/* Message of the last failed call on h (or of the last failed lbm_create when
 * h is NULL).  Always NUL-terminates when n > 0. */
int lbm_last_error(const lbm_handle *h, char *buf, size_t n);

/* Device name / library build string, for logs. */
const char *lbm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
