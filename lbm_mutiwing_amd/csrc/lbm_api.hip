// lbm_api.hip - C ABI of include/lbm_hip.h: handle lifecycle, host-side
// motion tables, and the per-step launch sequence of the FP64 path.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <vector>

#include "lbm_internal.h"

using namespace lbm;

namespace {

char g_create_error[512] = "";

// nizhen (solution.f90:1474-1529): Gauss-Jordan inverse of the MRT matrix,
// operation for operation, so the device uses the reference's values.
void gauss_jordan_inverse(const double AA[9][9], double b1[9][9]) {
  const int Nm = 9;
  double a[9][9];
  std::memcpy(a, AA, sizeof(a));
  for (int i = 0; i < Nm; i++)
    for (int j = 0; j < Nm; j++) b1[i][j] = i == j ? 1.0 : 0.0;
  for (int k = 0; k < Nm - 1; k++) {
    int ik = k;
    double temp = std::fabs(a[ik][k]);
    for (int i = k; i < Nm; i++)
      if (std::fabs(a[i][k]) > std::fabs(temp)) {
        temp = std::fabs(a[i][k]);
        ik = i;
      }
    for (int j = 0; j < Nm; j++) {
      std::swap(a[ik][j], a[k][j]);
      std::swap(b1[ik][j], b1[k][j]);
    }
    for (int i = 0; i < Nm; i++) {
      const double mik = a[i][k] / a[k][k];
      temp = a[k][k];
      for (int j = k; j < Nm; j++)
        if (i != k) a[i][j] = a[i][j] - mik * a[k][j];
      for (int j = k; j < Nm; j++) a[k][j] = a[k][j] / temp;
      for (int j = 0; j < Nm; j++)
        if (i != k) b1[i][j] = b1[i][j] - mik * b1[k][j];
      for (int j = 0; j < Nm; j++) b1[k][j] = b1[k][j] / temp;
    }
  }
  const double temp = a[Nm - 1][Nm - 1];
  for (int i = 0; i < Nm - 1; i++) {
    const double mik = a[i][Nm - 1] / a[Nm - 1][Nm - 1];
    a[i][Nm - 1] = 0.0;
    for (int j = 0; j < Nm; j++) b1[i][j] = b1[i][j] - mik * b1[Nm - 1][j];
  }
  for (int j = 0; j < Nm; j++) b1[Nm - 1][j] = b1[Nm - 1][j] / temp;
}

constexpr double kMatM[9][9] = {  // solution.f90:17-25
    {1, 1, 1, 1, 1, 1, 1, 1, 1},    {-4, -1, -1, -1, -1, 2, 2, 2, 2}, {4, -2, -2, -2, -2, 1, 1, 1, 1},
    {0, 1, 0, -1, 0, 1, -1, -1, 1}, {0, -2, 0, 2, 0, 1, -1, -1, 1},   {0, 0, 1, 0, -1, 1, 1, -1, -1},
    {0, 0, -2, 0, 2, 1, 1, -1, -1}, {0, 1, -1, 1, -1, 0, 0, 0, 0},    {0, 0, 0, 0, 0, 1, -1, 1, -1}};

// interpolate_motion_triplet and the smooth ramps (mode.f90:220-285), host side
// for build_stream_target_row.
double h_smooth_r(double x) {
  const double pai = 3.1415926535897932384626433832795;
  if (x < 0.0) return 0.0;
  if (x <= 1.0) return 0.5 * x * x + (std::cos(2.0 * pai * x) - 1.0) / (4.0 * pai * pai);
  return x - 0.5;
}
double h_smooth_h(double x) {
  const double pai = 3.1415926535897932384626433832795;
  if (x < 0.0) return 0.0;
  if (x <= 1.0) return x - std::sin(2.0 * pai * x) / (2.0 * pai);
  return 1.0;
}
double h_smooth_d(double x) {
  const double pai = 3.1415926535897932384626433832795;
  if (x >= 0.0 && x <= 1.0) return 1.0 - std::cos(2.0 * pai * x);
  return 0.0;
}
void h_interp_triplet(double s0, double v0, double v1, double dt, double lt, double* s, double* v, double* a) {
  const double dv = v1 - v0;
  if (dt <= 1.0e-14) {
    *s = (v0 + dv) * lt + s0;
    *v = v0 + dv;
    *a = 0.0;
    return;
  }
  const double x = lt / dt;
  *s = dv * dt * h_smooth_r(x) + v0 * lt + s0;
  *v = dv * h_smooth_h(x) + v0;
  *a = dv / dt * h_smooth_d(x);
}

}  // namespace

struct lbm_handle {
  lbm_config cfg;
  Grid g;
  MrtParams mrt;
  IbmParams ip;
  GeoBuffers G;
  FlowBuffers F;
  RingBuffers R;
  hipStream_t stream = nullptr;      // lattice kernels, copies
  hipStream_t stream_ibm = nullptr;  // flow part of the IBM, overlapped with the lattice kernel
  hipStream_t stream_geo = nullptr;  // IBM geometry of the coming steps
  hipStream_t stream_bc = nullptr;   // boundary ring, concurrent with the IBM flow chain
  hipEvent_t ev_ibm = nullptr, ev_step = nullptr, ev_bc = nullptr;
  hipEvent_t ev_state = nullptr;   // coupled steps: body state on the host
  hipEvent_t ev_geo[2] = {nullptr, nullptr};  // geometry of a slot set is ready
  hipEvent_t ev_rel[2] = {nullptr, nullptr};  // last step using a slot set has finished
  int device = 0;
  // band substitution in the flow instead of the explicit inverse: 0 never,
  // 1 on steps with rigid-body coupling (their geometry is on the critical
  // path), 2 on every step
  int direct = 1;
  double feq[9];
  double w[9];
  // populations
  void* fbuf[2] = {nullptr, nullptr};   // double, or float deviations when f32
  bool f32 = false;
  size_t fsize = sizeof(double);        // bytes per stored population value
  int cur = 0;
  // field slots (rho,u,v of the last steps), tagged with their step number
  FieldSlot slot[3];
  FieldSlot spare;  // target of the move_mesh field shift
  int slot_step[3] = {-1, -1, -1};
  double* err_partial = nullptr;
  double* err_out = nullptr;
  double* mass_rows = nullptr;       // per-row sums of lbm_get_mass (ny), then the total

  // IBM / motion device memory
  std::vector<void*> ibm_allocs;     // geometry slots, flow buffers, markers
  Frame* d_frame = nullptr;
  size_t frame_cap = 0;
  double* d_table = nullptr;
  size_t table_cap = 0;
  double* d_stream = nullptr;
  DevStep* d_st = nullptr;
  double* d_rows = nullptr;
  size_t rows_cap = 0;
  double* d_mrows = nullptr;
  double* d_rowsum = nullptr;
  int nslot = 1;                     // steps per geometry batch (slots per set)
  unsigned batch_count = 0;          // geometry batches issued (selects the slot set)
  int last_slot = -1;                // slot of the last step run (-1: none since reset)
  // Geometry of the steps after the last call: the rest of that call's last
  // geometry batch, launched after its own steps with its own error state.
  // The next call uses it when it starts at n_first, nothing the geometry
  // depends on has changed (epoch) and its frames are the predicted ones.
  struct GeoCarry {
    bool valid = false;
    int n_first = 0;                 // first step covered
    int offset = 0;                  // slot of n_first inside its slot set
    int count = 0;                   // steps covered
    unsigned epoch = 0;
    std::vector<Frame> frames;       // frames the geometry was computed with
  } carry;
  unsigned geo_epoch = 0;            // bumped by every change of the motion source or of the body state
  bool carry_on = true;              // LBM_GEO_CARRY=0 in the environment turns the carry off
  DevStep* d_st_carry = nullptr;     // error state of the carried geometry
  // host state
  int W, P, SPW, nseg, npts, row_len;
  int n = 1;
  int movenum = 0;
  int flag2 = 1;
  int last_move_step = -1;           // step after which move_mesh last ran
  double rel[REL_FIELDS][MAXW] = {};  // body_rel_* (restart-file order)
  lbm_dynamics dyn{};                // rigid-body coupling switch and constants
  DevStep* h_st = nullptr;           // pinned host copy of DevStep (coupled steps)
  // pinned host memory the end of a call copies its outputs into (one
  // stream synchronisation per call), and the staging of the frames
  struct HostOut {
    DevStep st;
    double err;
    double last_row[LBM_MAX_FORCE_ROW];
    double row_sum[LBM_MAX_FORCE_ROW];
    int carry_err;                   // error state of the carried geometry
  };
  HostOut* h_out = nullptr;
  Frame* h_frames = nullptr;         // frame_cap entries
  Frame fr{};                        // movenum and body_extra_* of the next step
  bool motion_ready = false;
  double motion_end_time = 0.0;
  std::vector<Frame> frames;         // per step of the current call
  int lat_bx = 0;                    // lattice tile width (0: not tuned yet)
  int lat_waves = 0;                 // lattice waves per CU cap (0: none)
  int lat_cp = 0;                    // cache policy of the lattice kernel's population accesses
  size_t lat_lds = 0;                // dynamic LDS per lattice workgroup that sets the cap
  size_t dyn_lds = 0;                // the same for the cap beside a coupled step's geometry
  bool dyn_narrow = false;           // that cap held by the register-capped kernel instead
  // lattice-kernel timing (lbm_set_timing)
  bool timing = false;
  std::vector<hipEvent_t> ev;        // start/stop pairs of the current call
  double lattice_ms = 0.0;
  int lattice_launches = 0;
  char msg[512] = "";
};

namespace {

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

int fail(lbm_handle* h, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(h ? h->msg : g_create_error, 512, fmt, ap);
  va_end(ap);
  return code;
}

// Host exceptions (a failed host allocation) never leave the C ABI: the
// entry point returns LBM_E_INVALID with the exception's text.
#define LBM_CATCH(h, name)                                                                 \
  catch (const std::exception& e_) {                                                         \
    return fail((h), LBM_E_INVALID, "%s: host exception: %s", name, e_.what());             \
  }                                                                                          \
  catch (...) {                                                                              \
    return fail((h), LBM_E_INVALID, "%s: unknown host exception", name);                    \
  }

#define HIPCHECK(h, expr)                                                                         \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return fail((h), LBM_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(lbm_handle* h, T** p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess)
    return fail(h, LBM_E_DEVICE, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
  return 0;
}

// IBM allocation owned by h->ibm_allocs.
template <typename T>
int ialloc(lbm_handle* h, T** p, size_t count) {
  int rc = dalloc(h, p, count);
  if (rc == 0) h->ibm_allocs.push_back((void*)*p);
  return rc;
}

void update_body_extra(lbm_handle* h) {  // update_body_extra_from_rel (mode.f90:473-495)
  double av = 0.0, as = 0.0, ap = 0.0;
  h->fr.movenum = h->movenum;
  for (int w = 0; w < MAXW; w++) {
    h->fr.body_extra_v[w] = 0.0;
    h->fr.body_extra_s[w] = 0.0;
    h->fr.body_extra_pos[w] = 0.0;
  }
  for (int w = 0; w < h->W; w++) {
    av = av + h->rel[REL_V][w];
    as = as + h->rel[REL_S][w];
    ap = ap + h->rel[REL_POS][w];
    h->fr.body_extra_v[w] = av;
    h->fr.body_extra_s[w] = as;
    h->fr.body_extra_pos[w] = ap;
  }
}

const char* dev_error_text(int code) {
  switch (code) {
    case DERR_BOX_EDGE: return "IBM support box of a segment reaches within 3 nodes of the lattice edge";
    case DERR_MAP: return "support rectangle of a wing exceeds its claim map";
    case DERR_BAND: return "IBM band half-width exceeds the solver limit";
    case DERR_WINGS_MET: return "wings' support rectangles met on a step launched as apart";
    case DERR_WINGS_COUPLED:
      return "more than 32 segments of one wing, or 64 in all, couple wings whose supports share lattice nodes";
    case DERR_NOT_SPD: return "IBM matrix lost positive definiteness";
    case DERR_MOTION: return "motion source not initialised (call lbm_set_motion_table)";
    default: return "unknown device error";
  }
}

void free_all(lbm_handle* h) {
  dfree(h->fbuf[0]);
  dfree(h->fbuf[1]);
  for (auto& s : h->slot) {
    dfree(s.rho);
    dfree(s.u);
    dfree(s.v);
  }
  dfree(h->spare.rho);
  dfree(h->spare.u);
  dfree(h->spare.v);
  dfree(h->err_partial);
  dfree(h->mass_rows);
  dfree(h->d_st_carry);
  dfree(h->err_out);
  dfree(h->R.left);
  dfree(h->R.right);
  dfree(h->R.mom);
  dfree(h->R.bottom);
  dfree(h->R.top);
  for (void* p : h->ibm_allocs) dfree(p);
  h->ibm_allocs.clear();
  dfree(h->d_frame);
  if (h->h_st) (void)hipHostFree(h->h_st);
  h->h_st = nullptr;
  if (h->h_out) (void)hipHostFree(h->h_out);
  h->h_out = nullptr;
  if (h->h_frames) (void)hipHostFree(h->h_frames);
  h->h_frames = nullptr;
  dfree(h->d_table);
  dfree(h->d_stream);
  dfree(h->d_st);
  dfree(h->d_rows);
  dfree(h->d_mrows);
  dfree(h->d_rowsum);
  if (h->ev_ibm) (void)hipEventDestroy(h->ev_ibm);
  if (h->ev_step) (void)hipEventDestroy(h->ev_step);
  for (int s = 0; s < 2; s++) {
    if (h->ev_geo[s]) (void)hipEventDestroy(h->ev_geo[s]);
    if (h->ev_rel[s]) (void)hipEventDestroy(h->ev_rel[s]);
  }
  if (h->stream_ibm) (void)hipStreamDestroy(h->stream_ibm);
  if (h->stream_geo) (void)hipStreamDestroy(h->stream_geo);
  if (h->stream_bc) (void)hipStreamDestroy(h->stream_bc);
  if (h->ev_bc) (void)hipEventDestroy(h->ev_bc);
  if (h->ev_state) (void)hipEventDestroy(h->ev_state);
  for (auto e : h->ev) (void)hipEventDestroy(e);
  h->ev.clear();
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

int ensure_rows(lbm_handle* h, size_t steps) {
  if (steps <= h->rows_cap) return 0;
  dfree(h->d_rows);
  dfree(h->d_mrows);
  h->d_rows = h->d_mrows = nullptr;
  size_t cap = std::max<size_t>(steps, 1024);
  int rc = dalloc(h, &h->d_rows, cap * h->row_len);
  if (rc) return rc;
  rc = dalloc(h, &h->d_mrows, cap * (1 + 3 * h->W));
  if (rc) return rc;
  h->rows_cap = cap;
  h->F.rows = h->d_rows;
  h->F.mrows = h->d_mrows;
  return 0;
}

// Device frames of a call's steps and their pinned staging; the caller has
// synchronised the geometry stream, the only reader of both.
int ensure_frames(lbm_handle* h, size_t steps) {
  if (steps <= h->frame_cap) return 0;
  dfree(h->d_frame);
  h->d_frame = nullptr;
  if (h->h_frames) (void)hipHostFree(h->h_frames);
  h->h_frames = nullptr;
  h->frame_cap = 0;
  size_t cap = std::max<size_t>(steps, 1024);
  int rc = dalloc(h, &h->d_frame, cap);
  if (rc) return rc;
  HIPCHECK(h, hipHostMalloc((void**)&h->h_frames, sizeof(Frame) * cap, hipHostMallocDefault));
  h->frame_cap = cap;
  h->G.frame = h->d_frame;
  return 0;
}

// Lattice-kernel autotuner: every candidate launch configuration (cands[0]
// is the default) is timed on launch() in three interleaved rounds (per
// round and candidate: one warm-up launch, then three timed launches between
// HIP events on stream s); each candidate keeps its fastest round, and the
// fastest candidate is returned, the default unless another is more than
// 1 % faster.
template <typename F>
int autotune(hipStream_t s, const std::vector<int>& cands, F launch) {
  if (cands.size() == 1) return cands[0];
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return cands[0];
  std::vector<float> best_ms(cands.size(), 1e30f);
  for (int round = 0; round < 3; round++)
    for (size_t q = 0; q < cands.size(); q++) {
      launch(cands[q]);
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < 3; r++) launch(cands[q]);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) best_ms[q] = std::min(best_ms[q], ms);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int best = cands[0];
  float best_t = best_ms[0];
  for (size_t q = 1; q < cands.size(); q++)
    if (best_ms[q] < 0.99f * best_ms[0] && best_ms[q] < best_t) {
      best_t = best_ms[q];
      best = cands[q];
    }
  return best;
}

// LBM_LATTICE_TILE=<width> in the environment fixes the tile width (one of
// kLatticeTiles), 0 otherwise.
int env_tile() {
  if (const char* e = getenv("LBM_LATTICE_TILE")) {
    const int v = atoi(e);
    for (int bx : kLatticeTiles)
      if (bx == v) return v;
  }
  return 0;
}

// LBM_LATTICE_WAVES=<n> in the environment fixes the cap on lattice waves
// per CU (0: none), -1 otherwise.
int env_waves() {
  if (const char* e = getenv("LBM_LATTICE_WAVES")) return std::max(0, std::min(63, atoi(e)));
  return -1;
}

// A bound on the half-width of S = D D^T in band positions at any pose.  Two
// segments couple only when their support boxes (the nodes within 2.5
// spacings of the midpoints) share a node, so when the midpoints are at most
// 5 spacings apart in x and in y, hence at most 5 sqrt(2) spacings apart in
// any frame.  The distance between two midpoints of a wing is fixed under its
// rigid motion, so the largest band distance of the pairs within that radius
// bounds every step's band.
int band_bound(const double* xb, const double* yb, int W, int P, double scale) {
  if (W <= 0 || P < 2) return 0;   // no wings: no band
  const int n = P - 1;
  const double r = 5.0 * std::sqrt(2.0) * scale * (1.0 + 1e-9);
  int bw = 0;
  std::vector<double> mx(n), my(n);
  for (int w = 0; w < W; w++) {
    for (int k = 0; k < n; k++) {
      mx[k] = 0.5 * (xb[w * P + k] + xb[w * P + k + 1]);
      my[k] = 0.5 * (yb[w * P + k] + yb[w * P + k + 1]);
    }
    for (int k = 0; k < n; k++)
      for (int l = k + 1; l < n; l++) {
        const double dx = mx[k] - mx[l], dy = my[k] - my[l];
        if (dx * dx + dy * dy <= r * r) bw = std::max(bw, std::abs(pos_of_seg(k, n) - pos_of_seg(l, n)));
      }
  }
  return bw;
}

int lds_per_cu(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess) v = 0;
  return v > 0 ? v : 65536;
}

int lds_per_block(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) v = 0;
  return v > 0 ? v : 65536;
}

// Lattice waves per CU beside the geometry of a step with rigid-body
// coupling (two 256-wide workgroups per CU): the geometry workgroups on the
// step's critical path need LDS and wave slots.  Measured at 4096x2048 with
// two wings: caps 4, 6, 8 and 12 give 0.386, 0.385, 0.381 and 0.446 ms per
// coupled step.
constexpr int DYN_LATTICE_WAVES = 8;

// Caps on the lattice waves per CU the autotuner tries (0: none): fewer
// resident lattice waves leave room for the IBM and ring kernels beside it.
constexpr int kLatticeWaves[5] = {0, 24, 16, 12, 8};

// LBM_LATTICE_CP=<0|1|2> in the environment fixes the cache policy of the
// lattice kernel (CP_PLAIN, CP_NT_STORE, CP_NT), -1 otherwise.
int env_cp() {
  if (const char* e = getenv("LBM_LATTICE_CP")) {
    const int v = atoi(e);
    if (v >= CP_PLAIN && v <= CP_NT) return v;
  }
  return -1;
}

// Tile width, wave cap and cache policy of one instance, timed on its own
// buffers: the launches write interior nodes of the idle population buffer,
// which the next step overwrites entirely.  Candidates are
// cp << 16 | bx << 6 | waves.  Also the LDS of the tighter cap beside the
// geometry of a coupled step.
void tune_instance(lbm_handle* h) {
  if (h->lat_bx) return;
  const void* src = h->fbuf[h->cur];
  void* dst = h->fbuf[1 - h->cur];
  const AdjMask none{nullptr, nullptr, 0, 0};
  const int fixed = env_tile(), fw = env_waves(), fcp = env_cp();
  const int lds_cu = lds_per_cu(h->device), lds_blk = lds_per_block(h->device);
  std::vector<int> cands;
  for (int cp : {CP_PLAIN, CP_NT_STORE, CP_NT})
    for (int bx : {256, 128, 512})
      for (int wv : kLatticeWaves)
        if ((fcp < 0 || cp == fcp) && (!fixed || bx == fixed) && (fw < 0 || wv == fw) &&
            lattice_lds_for_waves(bx, wv, lds_cu, lds_blk) >= 0)
          cands.push_back(cp << 16 | bx << 6 | wv);
  if (cands.empty()) cands.push_back((fcp < 0 ? 0 : fcp) << 16 | (fixed ? fixed : 256) << 6);
  const int best = autotune(h->stream, cands, [&](int c) {
    const int bx = (c >> 6) & 1023, wv = c & 63;
    launch_lattice(h->stream, h->g, h->mrt, src, dst, nullptr, h->f32, bx, c >> 16, none, nullptr, nullptr,
                   (size_t)lattice_lds_for_waves(bx, wv, lds_cu, lds_blk));
  });
  h->lat_bx = (best >> 6) & 1023;
  h->lat_waves = best & 63;
  h->lat_cp = best >> 16;
  h->lat_lds = (size_t)lattice_lds_for_waves(h->lat_bx, h->lat_waves, lds_cu, lds_blk);
  int dyn_waves = DYN_LATTICE_WAVES;
  if (const char* e = getenv("LBM_DYN_WAVES")) dyn_waves = std::max(0, atoi(e));   // 0: no cap
  const long long dyn = lattice_lds_for_waves(h->lat_bx, dyn_waves, lds_cu, lds_blk);
  h->dyn_lds = dyn >= 0 ? (size_t)dyn : h->lat_lds;
  // the default cap of 8 waves per CU by register count, with no LDS taken
  // (LBM_DYN_NARROW=0: by unused LDS, as for the other caps)
  const char* nw = getenv("LBM_DYN_NARROW");
  h->dyn_narrow = dyn_waves == 8 && !(nw && nw[0] == '0');
}

// Per-slot geometry arrays and flow buffers of the IBM.  The number of steps
// per geometry batch follows from the bytes per step (dominated by the
// inverse matrices) with a budget of 256 MB per slot set.
int alloc_ibm(lbm_handle* h, const std::vector<double>& hx, const std::vector<double>& hy) {
  IbmParams& p = h->ip;
  GeoBuffers& G = h->G;
  FlowBuffers& F = h->F;
  const size_t npts = std::max(h->npts, 1), nseg = std::max(h->nseg, 1), cap = std::max(p.cap, 1);
  const size_t W = std::max(h->W, 1), spw = std::max(h->SPW, 1);
  // claim map: square of the largest marker radius about the rotation origin
  double r2 = 0.0;
  for (int q = 0; q < h->npts; q++) r2 = std::max(r2, hx[q] * hx[q] + hy[q] * hy[q]);
  const int side = (int)std::ceil(2.0 * std::sqrt(r2) / h->cfg.scale) + 12;
  p.map_cap = h->W > 0 ? side * side : 1;
  const size_t per_slot = sizeof(GeoStep) + 2 * npts * 8 + 4 * nseg * 8 + 4 * nseg * 4 + 2 * BOXW * nseg * 8 +
                          W * (spw + 1) * 4 + 3 * cap * 4 + W * spw * spw * 8 + W * (size_t)p.map_cap * 4 +
                          (size_t)CPL_KT * CPL_KT * 8 + W * band_cap(p) * 8;
  h->nslot = h->W > 0 ? (int)std::max<size_t>(1, std::min<size_t>(16, (256u << 20) / per_slot)) : 1;
  const size_t ns = 2 * (size_t)h->nslot;
  int rc = 0;
  double* m0;
  // the band substitution loads ahead past both ends of a band
  const size_t band_pad = (size_t)BAND_PAD_ROWS * (BMAX + 1);
  double* bandg = nullptr;
  if ((rc = ialloc(h, &m0, 2 * npts))) return rc;
  G.xb0 = m0;
  G.yb0 = m0 + npts;
  if ((rc = ialloc(h, &G.step, ns)) || (rc = ialloc(h, &G.xb1, ns * npts)) || (rc = ialloc(h, &G.yb1, ns * npts)) ||
      (rc = ialloc(h, &G.xb2, ns * nseg)) || (rc = ialloc(h, &G.yb2, ns * nseg)) ||
      (rc = ialloc(h, &G.ub, ns * nseg)) || (rc = ialloc(h, &G.vb, ns * nseg)) ||
      (rc = ialloc(h, &G.box, ns * nseg * 4)) || (rc = ialloc(h, &G.phx, ns * nseg * BOXW)) ||
      (rc = ialloc(h, &G.phy, ns * nseg * BOXW)) || (rc = ialloc(h, &G.seg_off, ns * W * (spw + 1))) ||
      (rc = ialloc(h, &G.adj_ij, ns * cap * 2)) || (rc = ialloc(h, &G.adj_seg, ns * cap)) ||
      (rc = ialloc(h, &G.sinv, h->W > 0 ? ns * W * spw * spw : 1)) ||
      (rc = ialloc(h, &bandg, (h->W > 0 ? ns * W * band_cap(p) : 1) + 2 * band_pad)) ||
      (rc = ialloc(h, &G.map, h->W > 0 ? ns * W * (size_t)p.map_cap : 1)) ||
      (rc = ialloc(h, &G.cpl, h->W > 1 ? ns * CPL_KT * CPL_KT : 1)))
    return rc;
  G.bandg = bandg + band_pad;
  HIPCHECK(h, hipMemset(bandg, 0, sizeof(double) * ((h->W > 0 ? ns * W * band_cap(p) : 1) + 2 * band_pad)));
  const int bpw = adj_collide_blocks_per_wing(p);
  if ((rc = ialloc(h, &F.cvec, 2 * nseg)) || (rc = ialloc(h, &F.yvec, 2 * nseg)) ||
      (rc = ialloc(h, &F.adj_g, 2 * cap)) || (rc = ialloc(h, &F.partial, 2 * 3 * MAXW * (size_t)bpw)))
    return rc;
  G.step_table = nullptr;
  G.stream_rows = h->d_stream;
  G.frame = nullptr;
  F.row_sum = h->d_rowsum;
  F.frame_out = nullptr;
  HIPCHECK(h, hipMemset(G.step, 0, ns * sizeof(GeoStep)));
  if (h->npts > 0) {
    HIPCHECK(h, hipMemcpy(m0, hx.data(), sizeof(double) * h->npts, hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(m0 + npts, hy.data(), sizeof(double) * h->npts, hipMemcpyHostToDevice));
  }
  return 0;
}

}  // namespace

extern "C" {

const char* lbm_version(void) { return "lbm_mutiwing_amd gfx950 step 0.3 (FP64, FP32-shifted)"; }

int lbm_last_error(const lbm_handle* h, char* buf, size_t n) {
  if (!buf || n == 0) return LBM_E_INVALID;
  snprintf(buf, n, "%s", h ? h->msg : g_create_error);
  return 0;
}

int lbm_create(const lbm_config* cfg, lbm_handle** out) try {
  if (!cfg || !out) return fail(nullptr, LBM_E_INVALID, "null argument");
  *out = nullptr;
  if (cfg->nx < 8 || cfg->ny < 8) return fail(nullptr, LBM_E_INVALID, "lattice must be at least 8x8");
  if (cfg->wing_count < 0 || cfg->wing_count > MAXW) return fail(nullptr, LBM_E_INVALID, "wing_count out of range");
  if (cfg->wing_count > 0 && (cfg->points_per_wing < 3 || !cfg->xb0 || !cfg->yb0))
    return fail(nullptr, LBM_E_INVALID, "marker arrays missing");
  if (!(cfg->scale > 0.0) || !(cfg->uu > 0.0) || !(cfg->re > 0.0))
    return fail(nullptr, LBM_E_INVALID, "scale, uu and re must be positive");
  if (cfg->motion_source != LBM_MOTION_FILE && cfg->motion_source != LBM_MOTION_STREAM)
    return fail(nullptr, LBM_E_INVALID, "motion_source must be 1 (file) or 2 (stream)");
  if (cfg->precision != LBM_PRECISION_FP64 && cfg->precision != LBM_PRECISION_FP32_SHIFTED)
    return fail(nullptr, LBM_E_INVALID, "precision must be LBM_PRECISION_FP64 or LBM_PRECISION_FP32_SHIFTED");
  lbm_handle* h = new lbm_handle();
  h->cfg = *cfg;
  h->cfg.xb0 = h->cfg.yb0 = nullptr;
  h->f32 = cfg->precision == LBM_PRECISION_FP32_SHIFTED;
  h->fsize = h->f32 ? sizeof(float) : sizeof(double);
  h->W = cfg->wing_count;
  h->P = h->W > 0 ? cfg->points_per_wing : 0;
  h->SPW = h->P > 0 ? h->P - 1 : 0;
  h->nseg = h->W * h->SPW;
  h->npts = h->W * h->P;
  h->row_len = 4 + 12 * h->W;
  (void)hipGetDevice(&h->device);
  // model constants (D2Q9model.f90, ini_mesh.f90:30-31, 59-67, boundary_condition.f90:304)
  const double niu = cfg->uu * (1.0 / cfg->scale) / cfg->re;
  const double tau = 3.0 * niu + 0.5;
  for (int k = 0; k < 9; k++) h->mrt.s[k] = cfg->s[k];
  if (!(h->mrt.s[7] > 0.0)) h->mrt.s[7] = 1.0 / tau;
  if (!(h->mrt.s[8] > 0.0)) h->mrt.s[8] = 1.0 / tau;
  double mi[9][9];
  gauss_jordan_inverse(kMatM, mi);
  for (int r = 0; r < 9; r++)
    for (int c = 0; c < 9; c++) h->mrt.minv[9 * r + c] = mi[r][c];
  h->mrt.cs = (double)sqrtf(3.0f) / 3.0;
  h->w[0] = (double)(4.0f / 9.0f);
  for (int k = 1; k < 5; k++) h->w[k] = (double)(1.0f / 9.0f);
  for (int k = 5; k < 9; k++) h->w[k] = (double)(1.0f / 36.0f);
  for (int k = 0; k < 9; k++) {  // D2Q9feq(1, UU, 0) (D2Q9model.f90:35-46)
    const double u1 = cfg->uu, v1 = 0.0;
    const double eu = (double)EX(k) * u1 + (double)EY(k) * v1;
    const double utwo = u1 * u1 + v1 * v1;
    h->feq[k] = h->w[k] * 1.0 * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * utwo);
  }
  // constants of the FP32-shifted storage: weights, scaled rates, and the
  // equilibrium moments of the rest state M w
  {
    static const double kC[9] = {1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 6, 1.0 / 12, 1.0 / 6, 1.0 / 12, 0.25, 0.25};
    double mw[9], ws = 0.0;
    for (int k = 0; k < 9; k++) {
      h->mrt.w[k] = h->w[k];
      ws += h->w[k];
      h->mrt.ra[k] = kC[k] * (1.0 - h->mrt.s[k]);
      h->mrt.rb[k] = kC[k] * h->mrt.s[k];
    }
    for (int r = 0; r < 9; r++) {
      mw[r] = 0.0;
      for (int c = 0; c < 9; c++) mw[r] += kMatM[r][c] * h->w[c];
    }
    h->mrt.wsum = ws;
    h->mrt.rk0 = -kC[0] * h->mrt.s[0] * ws;
    h->mrt.eq1 = -2.0 * ws - mw[1];
    h->mrt.eq2 = ws - mw[2];
  }
  h->g.nx = cfg->nx;
  h->g.ny = cfg->ny;
  const long long np = (long long)cfg->nx * cfg->ny;
  // Population planes are np nodes apart rounded up to a multiple of 64, so
  // that every plane starts on a 256-byte boundary and a 128-byte line of one
  // plane is one of every plane; the gaps are never read or written.
  h->g.ps = (np + 63) & ~63LL;
  // IBM params
  IbmParams& ip = h->ip;
  ip.W = h->W; ip.P = h->P; ip.SPW = h->SPW; ip.nseg = h->nseg; ip.npts = h->npts;
  ip.capw = BOXC * h->SPW;
  ip.cap = h->W * ip.capw;
  ip.nx = cfg->nx; ip.ny = cfg->ny;
  ip.scale = cfg->scale; ip.ox = cfg->ox; ip.oy = cfg->oy; ip.uu = cfg->uu;
  ip.dt = cfg->scale * cfg->uu;
  ip.wing_x0 = cfg->wing_x0; ip.wing_dx = cfg->wing_dx; ip.wing_y0 = cfg->wing_y0;
  ip.xc = cfg->xc; ip.yc = cfg->yc;
  ip.motion_source = cfg->motion_source;
  ip.step_count = 0; ip.ncols = 1 + LBM_MOTION_COLS_PER_WING * h->W;
  ip.motion_dt = ip.dt;
  ip.band_cls = band_class(band_bound(cfg->xb0, cfg->yb0, h->W, h->P, cfg->scale));
  int lds_max = 0;
  (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device);
  const size_t lds_need = std::max(geo_solve_lds_bytes(ip, true), bsolve_lds_bytes(ip));
  if (h->W > 0 && lds_need > (size_t)lds_max) {
    int rc = fail(nullptr, LBM_E_INVALID, "%d segments per wing need %zu B of LDS, device allows %d", h->SPW,
                  lds_need, lds_max);
    delete h;
    return rc;
  }
  if (h->W > 1 && geo_couple_lds_total() > (size_t)lds_max) {
    int rc = fail(nullptr, LBM_E_INVALID, "the coupling of wings needs %zu B of LDS per workgroup, device allows %d",
                  geo_couple_lds_total(), lds_max);
    delete h;
    return rc;
  }
  update_body_extra(h);
  {
    const char* e = std::getenv("LBM_GEO_CARRY");
    h->carry_on = !(e && e[0] == '0');
    const char* d = std::getenv("LBM_DIRECT_SOLVE");
    h->direct = d ? std::max(0, std::min(2, atoi(d))) : LBM_IBM_SOLVE_COUPLED;
  }

  int rc = 0;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream_ibm, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream_geo, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream_bc, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_bc, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_ibm, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_step, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_state, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_geo[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_geo[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_rel[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_rel[1], hipEventDisableTiming) != hipSuccess) {
    free_all(h);
    delete h;
    return fail(nullptr, LBM_E_DEVICE, "stream/event creation failed");
  }
#define ALLOC(ptr, n)                   \
  if ((rc = dalloc(h, &(ptr), (n)))) { \
    snprintf(g_create_error, 512, "%s", h->msg); \
    free_all(h);                        \
    delete h;                           \
    return rc;                          \
  }
  if (h->f32) {
    float *a = nullptr, *b = nullptr;
    ALLOC(a, 9 * (size_t)h->g.ps);
    h->fbuf[0] = a;
    ALLOC(b, 9 * (size_t)h->g.ps);
    h->fbuf[1] = b;
  } else {
    double *a = nullptr, *b = nullptr;
    ALLOC(a, 9 * (size_t)h->g.ps);
    h->fbuf[0] = a;
    ALLOC(b, 9 * (size_t)h->g.ps);
    h->fbuf[1] = b;
  }
  for (auto& s : h->slot) {
    ALLOC(s.rho, np);
    ALLOC(s.u, np);
    ALLOC(s.v, np);
  }
  ALLOC(h->spare.rho, np);
  ALLOC(h->spare.u, np);
  ALLOC(h->spare.v, np);
  ALLOC(h->err_partial, 2 * (size_t)err_blocks(h->g));
  ALLOC(h->err_out, 1);
  ALLOC(h->mass_rows, (size_t)cfg->ny + 1);
  ALLOC(h->R.left, 9 * (size_t)cfg->ny);
  ALLOC(h->R.right, 9 * (size_t)cfg->ny);
  ALLOC(h->R.mom, 6 * (size_t)cfg->ny);
  ALLOC(h->R.bottom, 9 * (size_t)cfg->nx);
  ALLOC(h->R.top, 9 * (size_t)cfg->nx);
  ALLOC(h->d_stream, 2 * LBM_MAX_MOTION_COLS);
  ALLOC(h->d_st, 1);
  ALLOC(h->d_st_carry, 1);
  if (hipHostMalloc((void**)&h->h_out, sizeof(lbm_handle::HostOut), hipHostMallocDefault) != hipSuccess) {
    free_all(h);
    delete h;
    return fail(nullptr, LBM_E_DEVICE, "pinned host allocation failed");
  }
  std::memset(h->h_out, 0, sizeof(lbm_handle::HostOut));
  ALLOC(h->d_rowsum, h->row_len);
#undef ALLOC
  // markers: read_boundary (airfoil.f90:229-241) plus the flow-centre offset
  std::vector<double> hx(h->npts), hy(h->npts);
  for (int q = 0; q < h->npts; q++) {
    hx[q] = cfg->xb0[q] + (cfg->xc - cfg->xp);
    hy[q] = cfg->yb0[q] + (cfg->yc - cfg->yp);
  }
  rc = alloc_ibm(h, hx, hy);
  if (rc == 0) rc = ensure_frames(h, 1024);
  if (rc == 0) rc = lbm_reset(h);
  if (rc) {
    snprintf(g_create_error, 512, "%s", h->msg);
    free_all(h);
    delete h;
    return rc;
  }
  *out = h;
  return 0;
} LBM_CATCH(nullptr, "lbm_create")

int lbm_destroy(lbm_handle* h) {
  if (!h) return 0;
  (void)hipStreamSynchronize(h->stream);
  (void)hipStreamSynchronize(h->stream_geo);
  free_all(h);
  delete h;
  return 0;
}

int lbm_reset(lbm_handle* h) try {
  if (!h) return LBM_E_INVALID;
  HIPCHECK(h, hipStreamSynchronize(h->stream_geo));
  HIPCHECK(h, hipStreamSynchronize(h->stream_ibm));
  h->cur = 0;
  launch_init_fields(h->stream, h->g, h->mrt, h->cfg.uu, h->feq, h->fbuf[0], h->R, h->slot[0], h->f32);
  launch_ring_moments(h->stream, h->g, h->R);
  for (int s = 0; s < 3; s++) h->slot_step[s] = -1;
  h->slot_step[0] = 0;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  HIPCHECK(h, hipMemset(h->d_st, 0, sizeof(DevStep)));
  if (h->cfg.motion_source == LBM_MOTION_STREAM) {  // driver.f90:120-124
    std::vector<double> z(2 * LBM_MAX_MOTION_COLS, 0.0);
    HIPCHECK(h, hipMemcpy(h->d_stream, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice));
    h->motion_ready = true;
  }
  for (int q = 0; q < REL_FIELDS; q++)
    for (int w = 0; w < MAXW; w++) h->rel[q][w] = 0.0;
  h->n = 1;
  h->movenum = 0;
  h->flag2 = 1;
  h->last_move_step = -1;
  h->last_slot = -1;
  h->geo_epoch++;
  update_body_extra(h);
  return 0;
} LBM_CATCH(h, "lbm_reset")

// load_motion_table checks + build_motion_step_table (mode.f90:104-178, 327-412).
int lbm_set_motion_table(lbm_handle* h, const double* rows, int32_t nrows, int32_t ncols, double* end_time) try {
  if (!h || !rows) return fail(h, LBM_E_INVALID, "null argument");
  const int nc = h->ip.ncols;
  if (ncols != nc) return fail(h, LBM_E_INVALID, "motion table must have %d columns, got %d", nc, ncols);
  if (nrows < 1) return fail(h, LBM_E_INVALID, "motion table is empty");
  for (int i = 1; i < nrows; i++)
    if (rows[(size_t)i * nc] < rows[(size_t)(i - 1) * nc])
      return fail(h, LBM_E_INVALID, "motion table time column must be non-decreasing");
  const double end = rows[(size_t)(nrows - 1) * nc];
  const double dt = h->ip.dt;
  const long long count = (long long)((end + 1.0e-12) / dt) + 1;
  std::vector<double> tab((size_t)count * nc);
  auto T = [&](int r, int c) { return rows[(size_t)(r - 1) * nc + c]; };
  int m = 1;
  for (long long s = 1; s <= count; s++) {
    const double tt = (double)(s - 1) * dt;
    double* row = &tab[(size_t)(s - 1) * nc];
    if (tt <= T(1, 0) + 1.0e-12) {
      m = 1;
      for (int c = 0; c < nc; c++) row[c] = T(1, c);
    } else if (tt >= T(nrows, 0) - 1.0e-12) {
      m = nrows;
      for (int c = 0; c < nc; c++) row[c] = T(nrows, c);
    } else {
      m = std::max(1, std::min(m, nrows - 1));
      while (m > 1 && tt <= T(m, 0) + 1.0e-12) m--;
      while (m < nrows - 1 && tt > T(m + 1, 0) + 1.0e-12) m++;
      const double t1 = T(m, 0), t2 = T(m + 1, 0), d = t2 - t1;
      if (std::fabs(d) <= 1.0e-14) {
        for (int c = 0; c < nc; c++) row[c] = T(m + 1, c);
      } else {
        double ratio = (tt - t1) / d;
        ratio = std::min(std::max(ratio, 0.0), 1.0);
        for (int c = 0; c < nc; c++) row[c] = T(m, c) + ratio * (T(m + 1, c) - T(m, c));
      }
    }
    row[0] = tt;
  }
  HIPCHECK(h, hipStreamSynchronize(h->stream_geo));
  if (tab.size() > h->table_cap) {
    dfree(h->d_table);
    h->d_table = nullptr;
    int rc = dalloc(h, &h->d_table, tab.size());
    if (rc) return rc;
    h->table_cap = tab.size();
  }
  HIPCHECK(h, hipMemcpy(h->d_table, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  h->G.step_table = h->d_table;
  h->ip.step_count = (int)count;
  h->geo_epoch++;
  h->motion_end_time = end;
  h->motion_ready = true;
  if (end_time) *end_time = end;
  return 0;
} LBM_CATCH(h, "lbm_set_motion_table")

int lbm_set_stream_segment(lbm_handle* h, const double* start_row, const double* target_row, int32_t ncols) try {
  if (!h || !start_row || !target_row) return fail(h, LBM_E_INVALID, "null argument");
  if (ncols != h->ip.ncols) return fail(h, LBM_E_INVALID, "stream rows must have %d columns", h->ip.ncols);
  std::vector<double> z(2 * LBM_MAX_MOTION_COLS, 0.0);
  for (int c = 0; c < ncols; c++) {
    z[c] = start_row[c];
    z[LBM_MAX_MOTION_COLS + c] = target_row[c];
  }
  HIPCHECK(h, hipStreamSynchronize(h->stream_geo));
  HIPCHECK(h, hipMemcpy(h->d_stream, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice));
  h->motion_ready = true;
  h->geo_epoch++;
  return 0;
} LBM_CATCH(h, "lbm_set_stream_segment")

int lbm_build_stream_target_row(int32_t W, const double* start, double target_time, const double* delta,
                                double* target) {
  if (W < 0 || W > MAXW || !start || !delta || !target) return LBM_E_INVALID;
  const int nc = 1 + LBM_MOTION_COLS_PER_WING * W;
  const double dt = target_time - start[0];
  if (dt < -1.0e-12) return LBM_E_INVALID;
  for (int c = 0; c < nc; c++) target[c] = start[c];
  target[0] = target_time;
  const double lt = std::max(dt, 0.0);
  for (int w = 0; w < W; w++) {
    const int col = 1 + w * 9, dc = 1 + w * 3;
    h_interp_triplet(start[col], start[col + 1], start[col + 1] + delta[dc], dt, lt, &target[col],
                     &target[col + 1], &target[col + 2]);
    h_interp_triplet(start[col + 3], start[col + 4], start[col + 4] + delta[dc + 1], dt, lt, &target[col + 3],
                     &target[col + 4], &target[col + 5]);
    h_interp_triplet(start[col + 6], start[col + 7], start[col + 7] + delta[dc + 2], dt, lt, &target[col + 6],
                     &target[col + 7], &target[col + 8]);
  }
  return 0;
}

}  // extern "C"

namespace {

// move_mesh (move_mesh.f90:17-58, 171-214) after step m, whose fields are in
// slot s; the next step reads the shifted lattice after ev_step.
int apply_move(lbm_handle* h, int dr, int s, int m) {
  launch_shift_x(h->stream, h->g, dr, h->fbuf[h->cur], h->fbuf[1 - h->cur], 9, h->f32);
  h->cur = 1 - h->cur;
  launch_copy_ring_column(h->stream, h->g, h->mrt, h->fbuf[h->cur], dr == 1 ? 0 : h->g.nx - 1,
                          dr == 1 ? h->R.left : h->R.right, h->f32);
  // f0 = the shifted f in every column but the kept one, the wall rows included
  const int keep = dr == 1 ? h->g.nx - 1 : 0;
  launch_copy_ring_row(h->stream, h->g, h->mrt, h->fbuf[h->cur], 0, keep, h->R.bottom, h->f32);
  launch_copy_ring_row(h->stream, h->g, h->mrt, h->fbuf[h->cur], h->g.ny - 1, keep, h->R.top, h->f32);
  launch_ring_moments(h->stream, h->g, h->R);
  FieldSlot& cs = h->slot[s];
  launch_shift_x(h->stream, h->g, dr, cs.rho, h->spare.rho, 1, false);
  launch_shift_x(h->stream, h->g, dr, cs.u, h->spare.u, 1, false);
  launch_shift_x(h->stream, h->g, dr, cs.v, h->spare.v, 1, false);
  std::swap(h->slot[s], h->spare);
  h->rel[REL_POS][0] = h->rel[REL_POS][0] - (double)dr;
  h->movenum += dr;
  update_body_extra(h);
  h->flag2 = 3;
  h->last_move_step = m;
  HIPCHECK(h, hipEventRecord(h->ev_step, h->stream));
  return 0;
}

int step_error(lbm_handle* h, const DevStep& st) {
  return fail(h, st.error == DERR_MOTION ? LBM_E_STATE : (st.error == DERR_NOT_SPD ? LBM_E_NUMERIC : LBM_E_RANGE),
              "step failed: %s (info %d)", dev_error_text(st.error), st.error_info);
}

int finish_advance(lbm_handle* h, int nsteps, int n_last, bool timed, double* force_rows, double* motion_rows,
                   double* row_sum, double* last_row, double* err);

// Advance of an instance with rigid-body coupling: the geometry of step m
// needs the body state left by the force of step m-1, and the move_mesh check
// after step m reads that state on the host.  Each step runs the same kernels
// as lbm_advance (fields written every step, so that a move can shift them),
// with its force row, and the body update in it, right after the adjoint
// collision.  The step's chain (geometry, IBM flow, adjoint collision, force
// row) runs in order on one stream, with no event between its kernels; the
// lattice kernel does not wait for the step's geometry: it writes every
// interior node, beside the chain, and the adjoint collision overwrites the
// adjoint nodes (populations and fields) after it.  The steps are pipelined:
// everything of step m+1 but its force row is issued before the host reads
// the body state of step m; the force row of m+1 is issued once that state
// shows no move.  A move (|body_rel_pos(1)| >= 1) waits for the work of step
// m+1 in flight, shifts the lattice, uploads the new frame and issues step
// m+1 again: its kernels overwrite everything the first issue wrote.  While
// the wings' support rectangles stay more than APART_GAP nodes apart, a step
// launches neither the inverse nor the coupling pass.
int advance_dynamic(lbm_handle* h, int32_t n_last, double* force_rows, double* motion_rows, double* row_sum,
                    double* last_row, double* err) {
  const int n0 = h->n;
  const int nsteps = n_last >= n0 ? n_last - n0 + 1 : 0;
  HIPCHECK(h, hipStreamSynchronize(h->stream_geo));   // geometry carried by a previous call reads the frames
  int rc = ensure_rows(h, (size_t)std::max(nsteps, 1));
  if (rc) return rc;
  rc = ensure_frames(h, 1);
  if (rc) return rc;
  if (nsteps > 0) tune_instance(h);
  HIPCHECK(h, hipMemsetAsync(h->d_rowsum, 0, sizeof(double) * h->row_len, h->stream));
  HIPCHECK(h, hipMemsetAsync(&h->d_st->error, 0, sizeof(int), h->stream));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  if (!h->h_st) HIPCHECK(h, hipHostMalloc((void**)&h->h_st, sizeof(DevStep), hipHostMallocDefault));
  h->carry.valid = false;   // these steps reuse the slot sets
  const int ns = h->nslot;
  hipStream_t chain = h->stream_geo;   // geometry, IBM flow, adjoint collision, force row
  // the force row of each step writes the next step's frame and body state
  // on the device; the host uploads them at the start and after a move
  FlowBuffers F = h->F;
  F.frame_out = h->d_frame;
  auto upload = [&]() -> int {
    HIPCHECK(h, hipMemcpy(h->d_frame, &h->fr, sizeof(Frame), hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(h->d_st->rel, h->rel, sizeof(h->rel), hipMemcpyHostToDevice));
    return 0;
  };
  auto sync_all = [&]() -> int {
    for (hipStream_t q : {h->stream_geo, h->stream_bc, h->stream_ibm, h->stream}) HIPCHECK(h, hipStreamSynchronize(q));
    return 0;
  };
  // the smallest gap between the wings' support rectangles last read on the
  // host (two steps before the one being issued, or one after a move)
  int gap = h->W < 2 ? INT_MAX : INT_MIN;
  // Everything of step m but its force row, from buffer `cur` into the other
  // one with geometry slot set `set`.  first: no earlier step is in flight
  // (the host has synchronised), so the lattice and ring wait for nothing.
  auto issue_body = [&](int m, int cur, int set, bool first) -> int {
    const int slot = set * ns, s = m % 3;
    const bool apart = h->W < 2 || gap > APART_GAP;
    launch_geometry(chain, h->ip, h->G, h->d_st, slot, m, 0, 1, true, h->direct >= 1, apart);
    FieldSlot* fs = &h->slot[s];
    const void* src = h->fbuf[cur];
    void* dst = h->fbuf[1 - cur];
    AdjMask every_node = adj_mask(h->ip, h->G, slot);
    every_node.W = 0;
    if (!first) {   // the populations of step m-1 are complete: adjoint nodes and ring
      for (hipStream_t q : {h->stream, h->stream_bc}) {
        HIPCHECK(h, hipStreamWaitEvent(q, h->ev_ibm, 0));
        HIPCHECK(h, hipStreamWaitEvent(q, h->ev_bc, 0));
      }
    }
    // a tighter wave cap than the tuned one: the step's chain beside the
    // lattice kernel is the critical path here
    if (h->dyn_narrow)
      launch_lattice_narrow(h->stream, h->g, h->mrt, src, dst, *fs, h->f32, h->lat_bx, h->lat_cp, every_node);
    else
      launch_lattice(h->stream, h->g, h->mrt, src, dst, fs, h->f32, h->lat_bx, h->lat_cp, every_node, nullptr,
                     nullptr, h->dyn_lds);
    HIPCHECK(h, hipGetLastError());
    HIPCHECK(h, hipEventRecord(h->ev_step, h->stream));
    launch_bc_ring(h->stream_bc, h->g, h->mrt, h->R, src, dst, fs, h->f32);
    HIPCHECK(h, hipEventRecord(h->ev_bc, h->stream_bc));
    launch_ibm_flow(chain, h->ip, h->mrt, h->G, slot, h->F, h->g, src, h->f32, h->direct >= 1, apart);
    HIPCHECK(h, hipStreamWaitEvent(chain, h->ev_step, 0));   // the lattice kernel's writes first
    launch_adj_collide(chain, h->g, h->mrt, h->ip, h->G, slot, h->F, m, src, dst, fs, h->f32);
    HIPCHECK(h, hipEventRecord(h->ev_ibm, chain));   // step m's adjoint nodes are complete (the ring: ev_bc)
    return 0;
  };
  // the force row of step m (body update, next frame), then its body state
  // to the host from another stream, so that the chain goes on to the next
  // step's geometry
  auto issue_force = [&](int m, int set) -> int {
    launch_force_row(chain, h->ip, h->G, set * ns, F, h->d_st, m, m - n0, h->row_len);
    HIPCHECK(h, hipEventRecord(h->ev_state, chain));
    HIPCHECK(h, hipStreamWaitEvent(h->stream_ibm, h->ev_state, 0));
    HIPCHECK(h, hipMemcpyAsync(h->h_st, h->d_st, sizeof(DevStep), hipMemcpyDeviceToHost, h->stream_ibm));
    HIPCHECK(h, hipEventRecord(h->ev_state, h->stream_ibm));
    return 0;
  };
  if (nsteps > 0) {
    if ((rc = upload())) return rc;
    const int set0 = (int)(h->batch_count & 1u);
    h->batch_count++;
    if ((rc = issue_body(n0, h->cur, set0, true))) return rc;
    if ((rc = issue_force(n0, set0))) return rc;
    int set = set0;
    for (int m = n0; m <= n_last; m++) {
      const int s = m % 3;
      // step m+1 on the device while the host waits for the state of step m
      int set_next = -1;
      const int cur_next = 1 - h->cur;
      if (m < n_last) {
        set_next = (int)(h->batch_count & 1u);
        h->batch_count++;
        if ((rc = issue_body(m + 1, cur_next, set_next, false))) return rc;
      }
      HIPCHECK(h, hipEventSynchronize(h->ev_state));
      h->slot_step[s] = m;
      h->cur = cur_next;
      h->last_slot = set * ns;
      const DevStep& st = *h->h_st;
      if (st.error != 0) {
        if ((rc = sync_all())) return rc;
        h->n = m + 1;
        return step_error(h, st);
      }
      if (h->W >= 2) gap = st.rect_gap;
      for (int q = 0; q < REL_FIELDS; q++)
        for (int w = 0; w < MAXW; w++) h->rel[q][w] = st.rel[q][w];
      update_body_extra(h);
      // the move_mesh check of the run loop (main.f90:393-412)
      const double brp = h->rel[REL_POS][0];
      const int dr = brp >= 1.0 ? 1 : (brp <= -1.0 ? -1 : 0);
      if (dr != 0) {
        if ((rc = sync_all())) return rc;             // step m+1 as issued is discarded
        rc = apply_move(h, dr, s, m);
        if (rc) return rc;
        HIPCHECK(h, hipStreamSynchronize(h->stream));  // the next step's ring and IBM streams read the shifted lattice
        if ((rc = upload())) return rc;                // pos and movenum changed on the host
        // an error the discarded issue raised (its geometry used the frame
        // before the move) is not the step's
        HIPCHECK(h, hipMemset(&h->d_st->error, 0, sizeof(int)));
        if (m < n_last && (rc = issue_body(m + 1, h->cur, set_next, true))) return rc;
      }
      if (h->flag2 > 0) h->flag2 = h->flag2 - 1;
      if (m < n_last) {
        if ((rc = issue_force(m + 1, set_next))) return rc;
        set = set_next;
      }
    }
  }
  h->n = std::max(n0, n_last + 1);
  return finish_advance(h, nsteps, n_last, false, force_rows, motion_rows, row_sum, last_row, err);
}

}  // namespace

extern "C" {

int lbm_advance(lbm_handle* h, int32_t n_last, double* force_rows, double* motion_rows, double* row_sum,
                double* last_row, double* err) try {
  if (!h) return LBM_E_INVALID;
  if (h->cfg.motion_source == LBM_MOTION_FILE && !h->motion_ready && h->W > 0)
    return fail(h, LBM_E_STATE, "motion table not loaded");
  if (h->dyn.enable && h->W > 0) return advance_dynamic(h, n_last, force_rows, motion_rows, row_sum, last_row, err);
  const int n0 = h->n;
  const int nsteps = n_last >= n0 ? n_last - n0 + 1 : 0;
  const int ns = h->nslot;
  // the geometry stream may still run the previous call's carried geometry,
  // which reads the frames
  HIPCHECK(h, hipStreamSynchronize(h->stream_geo));
  int rc = ensure_rows(h, (size_t)std::max(nsteps, 1));
  if (rc) return rc;
  // frames of the call's steps and of the rest of its last geometry batch
  const int span = h->W > 0 ? std::max(nsteps, 1) + ns : std::max(nsteps, 1);
  rc = ensure_frames(h, (size_t)span);
  if (rc) return rc;
  if (h->timing)
    while (h->ev.size() < 2 * (size_t)nsteps) {
      hipEvent_t e;
      HIPCHECK(h, hipEventCreate(&e));
      h->ev.push_back(e);
    }
  // The frame of every step is known now: moves are fully determined by
  // body_rel_pos(1), which only this host changes (main.f90:393-411).  The
  // prediction runs on past the call's last step for the geometry computed
  // ahead of the next call.
  bool carried = false;
  int r = 0;  // slot offset of step n0 inside its slot set
  if (nsteps > 0) {
    Frame* fr = h->h_frames;
    double brp0 = h->W > 0 ? h->rel[REL_POS][0] : 0.0;
    Frame f = h->fr;
    for (int k = 0; k < span; k++) {
      fr[k] = f;
      const int dr = brp0 >= 1.0 ? 1 : (brp0 <= -1.0 ? -1 : 0);
      if (dr != 0) {
        brp0 = brp0 - (double)dr;
        f.movenum += dr;
        double ap = brp0;
        f.body_extra_pos[0] = ap;
        for (int w = 1; w < h->W; w++) {
          ap = ap + h->rel[REL_POS][w];
          f.body_extra_pos[w] = ap;
        }
      }
    }
    const lbm_handle::GeoCarry& c = h->carry;
    if (h->W > 0 && c.valid && c.epoch == h->geo_epoch && c.n_first == n0 && h->h_out->carry_err == 0) {
      const int m = std::min(c.count, nsteps);
      carried = std::memcmp(c.frames.data(), fr, sizeof(Frame) * m) == 0;
      if (carried) r = c.offset;
    }
    h->carry.valid = false;
    HIPCHECK(h, hipMemcpyAsync(h->d_frame, fr, sizeof(Frame) * span, hipMemcpyHostToDevice, h->stream_geo));
  }
  if (nsteps > 0) tune_instance(h);
  // the call's counters are cleared on the lattice stream before ev_step,
  // which every other stream of the call waits for before its first kernel
  HIPCHECK(h, hipMemsetAsync(h->d_rowsum, 0, sizeof(double) * h->row_len, h->stream));
  HIPCHECK(h, hipMemsetAsync(&h->d_st->error, 0, sizeof(int), h->stream));
  HIPCHECK(h, hipEventRecord(h->ev_step, h->stream));
  HIPCHECK(h, hipStreamWaitEvent(h->stream_geo, h->ev_step, 0));
  // Geometry batches of ns steps alternate between two slot sets.  Batch b
  // of the call is computed on the geometry stream one batch ahead of the
  // flow: it is issued when batch b-1 starts, right after its slot set has
  // been released by the force row of the last step of batch b-2, so it runs
  // while batch b-1 steps.  Step k of the call is at kv = k + r of the
  // batch sequence: r = 0, or the offset of the call's first step in the
  // batch carried from the previous call, which is then batch 0 and in the
  // slot set that call used last.  Only the first batch of a call, when not
  // carried, is waited for.  The last batch is computed in full: its steps
  // after n_last (with their own error state, so that a step the caller has
  // not asked for raises nothing) become the carry of the next call.
  const unsigned bc0 = carried ? h->batch_count - 1u : h->batch_count;
  const int nbatch = h->W > 0 ? (nsteps + r + ns - 1) / ns : 0;
  bool carry_pending = false;
  auto set_of = [&](int b) { return (int)((bc0 + (unsigned)b) & 1u); };
  auto issue_geometry = [&](int b) -> int {
    if (b >= nbatch) return 0;
    const int gs = set_of(b), k0 = b * ns - r;   // first step of the batch, call-relative (b >= 1 when r > 0)
    const int own = std::min(ns, nsteps - k0);
    HIPCHECK(h, hipStreamWaitEvent(h->stream_geo, h->ev_rel[gs], 0));
    launch_geometry(h->stream_geo, h->ip, h->G, h->d_st, gs * ns, n0 + k0, k0, own, true, h->direct == 2);
    HIPCHECK(h, hipEventRecord(h->ev_geo[gs], h->stream_geo));
    const int rest = ns - own;
    if (b == nbatch - 1 && rest > 0 && h->carry_on) {
      HIPCHECK(h, hipMemsetAsync(&h->d_st_carry->error, 0, sizeof(int), h->stream_geo));
      launch_geometry(h->stream_geo, h->ip, h->G, h->d_st_carry, gs * ns + own, n0 + k0 + own, k0 + own, rest,
                      true, h->direct == 2);
      HIPCHECK(h, hipMemcpyAsync(&h->h_out->carry_err, &h->d_st_carry->error, sizeof(int), hipMemcpyDeviceToHost,
                                 h->stream_geo));
      lbm_handle::GeoCarry& c = h->carry;
      c.n_first = n0 + k0 + own;
      c.offset = own;
      c.count = rest;
      c.epoch = h->geo_epoch;
      c.frames.assign(h->h_frames + (k0 + own), h->h_frames + (k0 + own + rest));
      carry_pending = true;   // valid once the call has completed
    }
    return 0;
  };
  if (!carried) {
    rc = issue_geometry(0);
    if (rc) return rc;
  }
  int set = 0, slot = 0, prev_set = 0, prev_slot = 0;
  // step_ev completes when the lattice stream has finished the previous
  // step: the call's start (recorded above), then the stop event bound to
  // each lattice launch (no event packet between two lattice kernels), or a
  // recorded event after a move_mesh
  hipEvent_t step_ev = h->ev_step;
  for (int m = n0; m <= n_last; m++) {
    const int k = m - n0, kv = k + r;
    const bool batch_start = h->W > 0 && (k == 0 || kv % ns == 0);
    if (batch_start) {
      prev_set = set;
      set = set_of(kv / ns);
      HIPCHECK(h, hipStreamWaitEvent(h->stream_ibm, h->ev_geo[set], 0));
      HIPCHECK(h, hipStreamWaitEvent(h->stream, h->ev_geo[set], 0));  // the lattice kernel reads the claim maps
    }
    slot = set * ns + kv % ns;
    const int s = m % 3;
    double brp = h->W > 0 ? h->rel[REL_POS][0] : 0.0;
    const int dr = brp >= 1.0 ? 1 : (brp <= -1.0 ? -1 : 0);
    // fields: rho, u, v of the last step (and of a step before a move, which
    // shifts them); u, v of the two steps before it, for erranalysis
    const bool want_fields = (m >= n_last - 2) || dr != 0;
    FieldSlot vel = h->slot[s];
    if (!(m == n_last || dr != 0)) vel.rho = nullptr;
    FieldSlot* fs = want_fields ? &vel : nullptr;
    const void* src = h->fbuf[h->cur];
    void* dst = h->fbuf[1 - h->cur];
    // Every node of dst has exactly one writer this step: the full-lattice
    // kernel (lattice stream) writes the interior nodes other than the
    // adjoint nodes; the IBM flow kernels and the adjoint re-collision
    // (second stream) write the adjoint nodes; the boundary ring (third
    // stream) writes the ring.  All three read only src and run
    // concurrently once the previous step has finished; the next step's
    // lattice kernel waits for the other two through ev_ibm.  The previous
    // step's force row follows on the IBM stream; a geometry slot set is
    // released once the force row of its last step is written.  The host
    // issues the lattice kernel first (the waits on step_ev before it see
    // the previous step's binding of the event) and the geometry of the
    // next batch last, so that the step's longest kernel is queued first.
    HIPCHECK(h, hipStreamWaitEvent(h->stream_ibm, step_ev, 0));
    HIPCHECK(h, hipStreamWaitEvent(h->stream_bc, step_ev, 0));
    step_ev = h->timing ? h->ev[2 * k + 1] : h->ev_step;
    launch_lattice(h->stream, h->g, h->mrt, src, dst, fs, h->f32, h->lat_bx, h->lat_cp, adj_mask(h->ip, h->G, slot),
                   h->timing ? h->ev[2 * k] : nullptr, step_ev, h->lat_lds);
    HIPCHECK(h, hipGetLastError());
    launch_bc_ring(h->stream_bc, h->g, h->mrt, h->R, src, dst, fs, h->f32);
    HIPCHECK(h, hipEventRecord(h->ev_bc, h->stream_bc));
    // the previous step's force row (its partial sums are kept by step parity)
    if (k > 0) {
      launch_force_row(h->stream_ibm, h->ip, h->G, prev_slot, h->F, h->d_st, m - 1, k - 1, h->row_len);
      if (batch_start) HIPCHECK(h, hipEventRecord(h->ev_rel[prev_set], h->stream_ibm));
    }
    launch_ibm_flow(h->stream_ibm, h->ip, h->mrt, h->G, slot, h->F, h->g, src, h->f32, h->direct == 2);
    launch_adj_collide(h->stream_ibm, h->g, h->mrt, h->ip, h->G, slot, h->F, m, src, dst, fs, h->f32);
    HIPCHECK(h, hipStreamWaitEvent(h->stream_ibm, h->ev_bc, 0));
    HIPCHECK(h, hipEventRecord(h->ev_ibm, h->stream_ibm));
    HIPCHECK(h, hipStreamWaitEvent(h->stream, h->ev_ibm, 0));
    if (batch_start) {
      rc = issue_geometry(kv / ns + 1);
      if (rc) return rc;
    }
    prev_slot = slot;
    if (want_fields) h->slot_step[s] = m;
    h->cur = 1 - h->cur;
    if (dr != 0) {
      rc = apply_move(h, dr, s, m);
      if (rc) return rc;
      step_ev = h->ev_step;
    }
    if (h->flag2 > 0) h->flag2 = h->flag2 - 1;
  }
  if (nsteps > 0) {
    HIPCHECK(h, hipEventRecord(h->ev_ibm, h->stream_ibm));   // the previous force row is done
    HIPCHECK(h, hipStreamWaitEvent(h->stream, h->ev_ibm, 0));
    launch_force_row(h->stream, h->ip, h->G, slot, h->F, h->d_st, n_last, nsteps - 1, h->row_len);
    if (h->W > 0) HIPCHECK(h, hipEventRecord(h->ev_rel[set], h->stream));
  }
  h->batch_count = bc0 + (unsigned)nbatch;
  if (nsteps > 0 && h->W > 0) h->last_slot = slot;
  h->n = std::max(n0, n_last + 1);
  rc = finish_advance(h, nsteps, n_last, h->timing, force_rows, motion_rows, row_sum, last_row, err);
  if (rc == 0 && carry_pending) h->carry.valid = true;
  return rc;
} LBM_CATCH(h, "lbm_advance")

}  // extern "C"

namespace {

// Common end of an advance over steps n_last-nsteps+1 .. n_last: the
// erranalysis of the last step, the device error check and the outputs.
int finish_advance(lbm_handle* h, int nsteps, int n_last, bool timed, double* force_rows, double* motion_rows,
                   double* row_sum, double* last_row, double* err) {
  // erranalysis of the last step (solution.f90:77-130): u of step n-1 against step n-2
  double errv = 0.0;
  const int ml = n_last;
  bool have_err = false;
  if (nsteps > 0 && ml >= 2) {
    if (h->last_move_step == ml - 1) {
      have_err = true;  // move_mesh copies u into u0: the interior difference is zero
    } else {
      const int s1 = (ml - 1) % 3, s2 = (ml - 2) % 3;
      if (h->slot_step[s1] == ml - 1 && h->slot_step[s2] == ml - 2) {
        const int nb = err_blocks(h->g);
        launch_err_partial(h->stream, h->g, h->slot[s1], h->slot[s2], h->err_partial, nb);
        launch_err_final(h->stream, h->err_partial, nb, h->err_out);
        have_err = true;
      }
    }
  }
  HIPCHECK(h, hipGetLastError());
  // the outputs into pinned memory behind the step, then one synchronisation
  lbm_handle::HostOut* ho = h->h_out;
  HIPCHECK(h, hipMemcpyAsync(&ho->st, h->d_st, sizeof(DevStep), hipMemcpyDeviceToHost, h->stream));
  const bool read_err = have_err && h->last_move_step != ml - 1;
  if (read_err) HIPCHECK(h, hipMemcpyAsync(&ho->err, h->err_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if (nsteps > 0) {
    if (last_row)
      HIPCHECK(h, hipMemcpyAsync(ho->last_row, h->d_rows + (size_t)(nsteps - 1) * h->row_len,
                                 sizeof(double) * h->row_len, hipMemcpyDeviceToHost, h->stream));
    if (row_sum)
      HIPCHECK(h, hipMemcpyAsync(ho->row_sum, h->d_rowsum, sizeof(double) * h->row_len, hipMemcpyDeviceToHost,
                                 h->stream));
  }
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  if (ho->st.error != 0) return step_error(h, ho->st);
  if (timed)
    for (int i = 0; i < nsteps; i++) {
      float ms = 0.f;
      HIPCHECK(h, hipEventElapsedTime(&ms, h->ev[2 * i], h->ev[2 * i + 1]));
      h->lattice_ms += ms;
      h->lattice_launches++;
    }
  if (read_err) errv = ho->err;
  if (err) *err = errv;
  if (nsteps > 0) {
    if (force_rows)
      HIPCHECK(h, hipMemcpy(force_rows, h->d_rows, sizeof(double) * h->row_len * nsteps, hipMemcpyDeviceToHost));
    if (motion_rows)
      HIPCHECK(h, hipMemcpy(motion_rows, h->d_mrows, sizeof(double) * (1 + 3 * h->W) * nsteps, hipMemcpyDeviceToHost));
    if (last_row) std::memcpy(last_row, ho->last_row, sizeof(double) * h->row_len);
    if (row_sum) std::memcpy(row_sum, ho->row_sum, sizeof(double) * h->row_len);
  } else if (row_sum) {
    for (int c = 0; c < h->row_len; c++) row_sum[c] = 0.0;
  }
  return 0;
}

}  // namespace

extern "C" {

int lbm_get_state(lbm_handle* h, int32_t* n_next, int32_t* movenum, int32_t* flag2, int32_t* row_len,
                  int32_t* adjoint_n) {
  if (!h) return LBM_E_INVALID;
  if (n_next) *n_next = h->n;
  if (movenum) *movenum = h->movenum;
  if (flag2) *flag2 = h->flag2;
  if (row_len) *row_len = h->row_len;
  if (adjoint_n) {
    *adjoint_n = 0;
    if (h->last_slot >= 0) {
      GeoStep gs;
      HIPCHECK(h, hipMemcpy(&gs, h->G.step + h->last_slot, sizeof(GeoStep), hipMemcpyDeviceToHost));
      *adjoint_n = gs.adj_n;
    }
  }
  return 0;
}

int lbm_get_fields(lbm_handle* h, double* rho, double* u, double* v, double* pre) try {
  if (!h) return LBM_E_INVALID;
  const int m = h->n - 1;
  const int s = ((m % 3) + 3) % 3;
  if (h->slot_step[s] != m) return fail(h, LBM_E_STATE, "fields of step %d are not resident", m);
  const size_t np = (size_t)h->g.nx * h->g.ny;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  if (rho) HIPCHECK(h, hipMemcpy(rho, h->slot[s].rho, np * sizeof(double), hipMemcpyDeviceToHost));
  if (u) HIPCHECK(h, hipMemcpy(u, h->slot[s].u, np * sizeof(double), hipMemcpyDeviceToHost));
  if (v) HIPCHECK(h, hipMemcpy(v, h->slot[s].v, np * sizeof(double), hipMemcpyDeviceToHost));
  if (pre) {
    std::vector<double> r(np);
    HIPCHECK(h, hipMemcpy(r.data(), h->slot[s].rho, np * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < np; q++) pre[q] = r[q] / 3.0;
  }
  return 0;
} LBM_CATCH(h, "lbm_get_fields")

int lbm_get_mass(lbm_handle* h, double* mass) {
  if (!h || !mass) return LBM_E_INVALID;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  double* total = h->mass_rows + h->g.ny;
  launch_mass(h->stream, h->g, h->mrt, h->fbuf[h->cur], h->f32, h->mass_rows, total);
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  HIPCHECK(h, hipMemcpy(mass, total, sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int lbm_get_mass_noalloc(const void* f, int32_t nx, int32_t ny, int32_t precision, double* mass, void* stream) {
  if (!f || !mass) return LBM_E_INVALID;
  if (nx < 1 || ny < 1) return LBM_E_INVALID;
  if (precision != LBM_PRECISION_FP64 && precision != LBM_PRECISION_FP32_SHIFTED) return LBM_E_INVALID;
  if (ny > mass_scratch_rows()) {
    snprintf(g_create_error, sizeof(g_create_error), "lbm_get_mass_noalloc: ny = %d exceeds the scratch of %d rows",
             ny, mass_scratch_rows());
    return LBM_E_RANGE;
  }
  double* rows = nullptr;
  hipError_t e = mass_scratch(&rows);
  if (e != hipSuccess) {
    snprintf(g_create_error, sizeof(g_create_error), "lbm_get_mass_noalloc: %s", hipGetErrorString(e));
    return LBM_E_DEVICE;
  }
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.ps = (long long)nx * ny;
  MrtParams P{};
  P.w[0] = (double)(4.0f / 9.0f);  // the weights of lbm_create (D2Q9model.f90:24-32)
  for (int k = 1; k < 5; k++) P.w[k] = (double)(1.0f / 9.0f);
  for (int k = 5; k < 9; k++) P.w[k] = (double)(1.0f / 36.0f);
  launch_mass((hipStream_t)stream, g, P, f, precision == LBM_PRECISION_FP32_SHIFTED, rows, mass);
  e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(g_create_error, sizeof(g_create_error), "lbm_get_mass_noalloc: %s", hipGetErrorString(e));
    return LBM_E_DEVICE;
  }
  return 0;
}

int lbm_get_populations(lbm_handle* h, double* f, double* f0) try {
  if (!h) return LBM_E_INVALID;
  const size_t np = (size_t)h->g.nx * h->g.ny;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  if (f) {
    std::vector<float> tmp(h->f32 ? np : 0);
    for (int k = 0; k < 9; k++) {
      const char* plane = (const char*)h->fbuf[h->cur] + (size_t)k * h->g.ps * h->fsize;
      if (h->f32) {
        HIPCHECK(h, hipMemcpy(tmp.data(), plane, np * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < np; q++) f[k * np + q] = h->w[k] + (double)tmp[q];
      } else {
        HIPCHECK(h, hipMemcpy(f + k * np, plane, np * sizeof(double), hipMemcpyDeviceToHost));
      }
    }
  }
  if (f0) {
    std::memset(f0, 0, 9 * np * sizeof(double));
    const size_t nx = h->g.nx, ny = h->g.ny;
    std::vector<double> L(9 * ny), Rr(9 * ny), B(9 * nx), T(9 * nx);
    HIPCHECK(h, hipMemcpy(L.data(), h->R.left, L.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(h, hipMemcpy(Rr.data(), h->R.right, Rr.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(h, hipMemcpy(B.data(), h->R.bottom, B.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(h, hipMemcpy(T.data(), h->R.top, T.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < 9; k++) {
      for (size_t i = 0; i < nx; i++) {
        f0[k * np + i] = B[k * nx + i];
        f0[k * np + (ny - 1) * nx + i] = T[k * nx + i];
      }
      for (size_t j = 0; j < ny; j++) {
        f0[k * np + j * nx] = L[k * ny + j];
        f0[k * np + j * nx + nx - 1] = Rr[k * ny + j];
      }
    }
  }
  return 0;
} LBM_CATCH(h, "lbm_get_populations")

int lbm_set_populations(lbm_handle* h, const double* f, const double* f0) try {
  if (!h) return LBM_E_INVALID;
  const size_t np = (size_t)h->g.nx * h->g.ny;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  if (f) {
    std::vector<float> tmp(h->f32 ? np : 0);
    for (int k = 0; k < 9; k++) {
      char* plane = (char*)h->fbuf[h->cur] + (size_t)k * h->g.ps * h->fsize;
      if (h->f32) {
        for (size_t q = 0; q < np; q++) tmp[q] = (float)(f[k * np + q] - h->w[k]);
        HIPCHECK(h, hipMemcpy(plane, tmp.data(), np * sizeof(float), hipMemcpyHostToDevice));
      } else {
        HIPCHECK(h, hipMemcpy(plane, f + k * np, np * sizeof(double), hipMemcpyHostToDevice));
      }
    }
    const int m = h->n - 1;
    const int s = ((m % 3) + 3) % 3;
    launch_moments_all(h->stream, h->g, h->mrt, h->fbuf[h->cur], h->slot[s], h->f32);
    h->slot_step[s] = m;
  }
  if (f0) {
    const size_t nx = h->g.nx, ny = h->g.ny;
    std::vector<double> L(9 * ny), Rr(9 * ny), B(9 * nx), T(9 * nx);
    for (int k = 0; k < 9; k++) {
      for (size_t j = 0; j < ny; j++) {
        L[k * ny + j] = f0[k * np + j * nx];
        Rr[k * ny + j] = f0[k * np + j * nx + nx - 1];
      }
      for (size_t i = 0; i < nx; i++) {
        B[k * nx + i] = f0[k * np + i];
        T[k * nx + i] = f0[k * np + (ny - 1) * nx + i];
      }
    }
    HIPCHECK(h, hipMemcpy(h->R.left, L.data(), L.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(h->R.right, Rr.data(), Rr.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(h->R.bottom, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(h, hipMemcpy(h->R.top, T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice));
    launch_ring_moments(h->stream, h->g, h->R);
  }
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return 0;
} LBM_CATCH(h, "lbm_set_populations")

int lbm_get_markers(lbm_handle* h, double* xb1, double* yb1, double* xb2, double* yb2) {
  if (!h) return LBM_E_INVALID;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  const int s = h->last_slot;
  auto get = [&](double* out, const double* base, int count) -> int {
    if (!out || count <= 0) return 0;
    if (s < 0) {
      std::memset(out, 0, sizeof(double) * count);
      return 0;
    }
    HIPCHECK(h, hipMemcpy(out, base + (size_t)s * count, sizeof(double) * count, hipMemcpyDeviceToHost));
    return 0;
  };
  int rc = get(xb1, h->G.xb1, h->npts);
  if (!rc) rc = get(yb1, h->G.yb1, h->npts);
  if (!rc) rc = get(xb2, h->G.xb2, h->nseg);
  if (!rc) rc = get(yb2, h->G.yb2, h->nseg);
  return rc;
}

int lbm_get_ibm(lbm_handle* h, int32_t capacity, int32_t* ij, int32_t* wing_id, double* wing_force,
                double* wing_moment) try {
  if (!h) return LBM_E_INVALID;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  DevStep st;
  HIPCHECK(h, hipMemcpy(&st, h->d_st, sizeof(DevStep), hipMemcpyDeviceToHost));
  if (h->last_slot >= 0 && (ij || wing_id)) {
    GeoStep gs;
    HIPCHECK(h, hipMemcpy(&gs, h->G.step + h->last_slot, sizeof(GeoStep), hipMemcpyDeviceToHost));
    int a = 0;
    for (int w = 0; w < h->W && a < capacity; w++) {
      const int n = std::min(gs.cnt[w], capacity - a);
      if (n <= 0) continue;
      std::vector<int> hij(2 * (size_t)n);
      const size_t off = (size_t)h->last_slot * h->ip.cap + (size_t)w * h->ip.capw;
      HIPCHECK(h, hipMemcpy(hij.data(), h->G.adj_ij + 2 * off, sizeof(int) * 2 * n, hipMemcpyDeviceToHost));
      for (int q = 0; q < n; q++, a++) {
        if (ij) {
          ij[2 * a] = hij[2 * q] + 1;
          ij[2 * a + 1] = hij[2 * q + 1] + 1;
        }
        if (wing_id) wing_id[a] = w + 1;
      }
    }
  }
  for (int w = 0; w < MAXW; w++) {
    if (wing_force) {
      wing_force[2 * w] = w < h->W ? st.force[w][0] : 0.0;
      wing_force[2 * w + 1] = w < h->W ? st.force[w][1] : 0.0;
    }
    if (wing_moment) wing_moment[w] = w < h->W ? st.moment[w] : 0.0;
  }
  return 0;
} LBM_CATCH(h, "lbm_get_ibm")

int lbm_set_timing(lbm_handle* h, int32_t enable) {
  if (!h) return LBM_E_INVALID;
  h->timing = enable != 0;
  h->lattice_ms = 0.0;
  h->lattice_launches = 0;
  return 0;
}

int lbm_get_timing(lbm_handle* h, double* lattice_ms, int32_t* launches) {
  if (!h) return LBM_E_INVALID;
  if (lattice_ms) *lattice_ms = h->lattice_ms;
  if (launches) *launches = h->lattice_launches;
  return 0;
}

int lbm_measure_pull_copy(int32_t nx, int32_t ny, int32_t precision, int32_t launches, double* avg_us) {
  if (nx < 8 || ny < 8 || launches < 1 || !avg_us ||
      (precision != LBM_PRECISION_FP64 && precision != LBM_PRECISION_FP32_SHIFTED))
    return LBM_E_INVALID;
  return measure_pull_copy(nx, ny, precision == LBM_PRECISION_FP32_SHIFTED, launches, avg_us);
}

int lbm_get_tuning(const lbm_handle* h, int32_t* lattice_tile, int32_t* waves_per_cu) {
  if (!h) return LBM_E_INVALID;
  if (lattice_tile) *lattice_tile = h->lat_bx;
  if (waves_per_cu) *waves_per_cu = h->lat_waves;
  return 0;
}

int lbm_get_lattice_policy(const lbm_handle* h, int32_t* cache_policy) {
  if (!h || !cache_policy) return LBM_E_INVALID;
  *cache_policy = h->lat_cp;
  return 0;
}

int lbm_set_step(lbm_handle* h, int32_t n_next, int32_t movenum) {
  if (!h) return LBM_E_INVALID;
  if (n_next < 1) return fail(h, LBM_E_INVALID, "n_next must be at least 1");
  HIPCHECK(h, hipStreamSynchronize(h->stream_geo));
  HIPCHECK(h, hipStreamSynchronize(h->stream_ibm));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  h->n = n_next;
  h->movenum = movenum;
  h->flag2 = 1;  // main.f90:349 holds for a restart too
  h->last_move_step = -1;
  h->last_slot = -1;
  for (int s = 0; s < 3; s++) h->slot_step[s] = -1;
  h->geo_epoch++;
  update_body_extra(h);
  return 0;
}

int lbm_set_body_rel_pos(lbm_handle* h, int32_t wing, double pos) {
  if (!h || wing < 1 || wing > h->W) return fail(h, LBM_E_INVALID, "wing out of range");
  h->rel[REL_POS][wing - 1] = pos;
  h->geo_epoch++;
  update_body_extra(h);
  return 0;
}

int lbm_get_body_rel(lbm_handle* h, double* rel) {
  if (!h || !rel) return fail(h, LBM_E_INVALID, "null argument");
  for (int q = 0; q < REL_FIELDS; q++)
    for (int w = 0; w < h->W; w++) rel[q * h->W + w] = h->rel[q][w];
  return 0;
}

int lbm_set_body_rel(lbm_handle* h, const double* rel) {
  if (!h || !rel) return fail(h, LBM_E_INVALID, "null argument");
  for (int q = 0; q < REL_FIELDS; q++)
    for (int w = 0; w < h->W; w++) h->rel[q][w] = rel[q * h->W + w];
  h->geo_epoch++;
  update_body_extra(h);
  return 0;
}

int lbm_set_dynamics(lbm_handle* h, const lbm_dynamics* d) {
  if (!h || !d) return fail(h, LBM_E_INVALID, "null argument");
  if (d->enable && h->W == 0) return fail(h, LBM_E_INVALID, "rigid-body coupling needs at least one wing");
  if (d->enable && !(d->solid_den > 0.0 && (!d->closed_boundary || d->area > 0.0)))
    return fail(h, LBM_E_INVALID, "solid_den and area must be positive");
  h->dyn = *d;
  h->geo_epoch++;
  IbmParams& p = h->ip;
  const double sc = h->cfg.scale, uu = h->cfg.uu;
  p.dyn_on = d->enable ? 1 : 0;
  p.dyn_rk4 = d->rk4 ? 1 : 0;
  p.dyn_mass = d->closed_boundary ? d->solid_den * d->area / (sc * sc) : d->solid_den / (sc * sc);
  p.dyn_p1 = d->par1 * uu * uu * sc * sc;
  p.dyn_p2 = d->par2 * (sc * sc * sc * sc);
  p.dyn_alphad = d->alphad;
  return 0;
}

int lbm_set_ibm_solve(lbm_handle* h, int32_t mode) {
  if (!h) return LBM_E_INVALID;
  if (mode < LBM_IBM_SOLVE_INVERSE || mode > LBM_IBM_SOLVE_SUBSTITUTION)
    return fail(h, LBM_E_INVALID, "IBM solve mode must be 0, 1 or 2");
  h->direct = mode;
  h->geo_epoch++;   // geometry computed ahead for the other mode is not carried over
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------ batched advance

// A captured batched launch sequence: a whole short advance of one
// (composition, steps) pair, or one geometry batch of a long advance of one
// composition, replayed when the same key comes again.
// Longest advance that is captured whole; it must also stay within one
// geometry batch.  A capture whose geometry stream waits on slot-release
// events recorded inside the same capture (a whole advance over several
// geometry batches) ended the process with a host-side segmentation fault
// inside this call, for a reason not found; no capture here has that form.
constexpr int kMaxGraphSteps = 64;

struct BatchGraph {
  std::vector<long long> key;
  int seen = 0;                      // eager runs of this key so far
  hipGraphExec_t exec = nullptr;
  unsigned long long last_use = 0;
};

struct lbm_batch {
  hipStream_t stream = nullptr, stream_ibm = nullptr, stream_geo = nullptr, stream_bc = nullptr;
  hipEvent_t ev_ibm = nullptr, ev_step = nullptr, ev_bc = nullptr, ev_start = nullptr;
  std::vector<BatchGraph> graphs;
  unsigned long long calls = 0;
  int graph_mode = 1;                // 0 none, 1 short advances, 2 also the middle batches of long ones
  hipEvent_t ev_geo[2] = {nullptr, nullptr}, ev_rel[2] = {nullptr, nullptr};
  int device = 0;
  EpDesc* d_desc = nullptr;
  size_t desc_cap = 0;
  int2* d_items = nullptr;
  size_t items_cap = 0;
  int* d_kbase = nullptr;            // device step base of the batched launches
  int lat_bx = 0;                    // lattice tile width (tuned on the first call)
  int lat_rows = 1;                  // lattice tiles per work item (tuned with the width)
  int lat_waves = 0;                 // lattice waves per CU cap (0: none), tuned last
  int lat_cp = 0;                    // cache policy of the lattice kernel, tuned with the cap
  size_t lat_lds = 0;
  bool direct = false;               // some episode of the call solves by band substitution
  size_t bsolve_lds = 0;             // LDS of its band substitution (the largest episode's)
  char msg[512] = "";
};

namespace {

int bfail(lbm_batch* b, int code, const char* fmt, ...) {
  if (!b) return code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b->msg, sizeof(b->msg), fmt, ap);
  va_end(ap);
  return code;
}

#define LBM_BATCH_CATCH(b, name)                                                           \
  catch (const std::exception& e_) {                                                         \
    return bfail((b), LBM_E_INVALID, "%s: host exception: %s", name, e_.what());            \
  }                                                                                          \
  catch (...) {                                                                              \
    return bfail((b), LBM_E_INVALID, "%s: unknown host exception", name);                   \
  }

#define BCHECK(b, expr)                                                                            \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return bfail((b), LBM_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

void batch_free(lbm_batch* b) {
  for (auto& g : b->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  b->graphs.clear();
  if (b->ev_start) (void)hipEventDestroy(b->ev_start);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_kbase) (void)hipFree(b->d_kbase);
  for (hipEvent_t e : {b->ev_ibm, b->ev_step, b->ev_bc, b->ev_geo[0], b->ev_geo[1], b->ev_rel[0], b->ev_rel[1]})
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t s : {b->stream, b->stream_ibm, b->stream_geo, b->stream_bc})
    if (s) (void)hipStreamDestroy(s);
}

// erranalysis of step ml after a call that ran steps up to ml (solution.f90:77-130)
bool issue_err(lbm_handle* h, int ml, int nsteps, hipStream_t s) {
  if (nsteps <= 0 || ml < 2) return false;
  if (h->last_move_step == ml - 1) return true;
  const int s1 = (ml - 1) % 3, s2 = (ml - 2) % 3;
  if (h->slot_step[s1] != ml - 1 || h->slot_step[s2] != ml - 2) return false;
  const int nb = err_blocks(h->g);
  launch_err_partial(s, h->g, h->slot[s1], h->slot[s2], h->err_partial, nb);
  launch_err_final(s, h->err_partial, nb, h->err_out);
  return true;
}

}  // namespace

extern "C" {

int lbm_batch_create(lbm_batch** out) {
  if (!out) return LBM_E_INVALID;
  lbm_batch* b = new lbm_batch();
  (void)hipGetDevice(&b->device);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream_ibm, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream_geo, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream_bc, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_bc, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_start, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_ibm, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_step, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_geo[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_geo[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_rel[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_rel[1], hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void**)&b->d_kbase, 256) != hipSuccess) {
    batch_free(b);
    delete b;
    return LBM_E_DEVICE;
  }
  if (const char* e = getenv("LBM_BATCH_GRAPHS")) b->graph_mode = std::max(0, std::min(2, atoi(e)));
  *out = b;
  return 0;
}

int lbm_batch_destroy(lbm_batch* b) {
  if (!b) return 0;
  (void)hipStreamSynchronize(b->stream);
  (void)hipStreamSynchronize(b->stream_geo);
  batch_free(b);
  delete b;
  return 0;
}

int lbm_batch_get_tuning(const lbm_batch* b, int32_t* lattice_tile, int32_t* rows_per_item, int32_t* waves_per_cu) {
  if (!b) return LBM_E_INVALID;
  if (lattice_tile) *lattice_tile = b->lat_bx;
  if (rows_per_item) *rows_per_item = b->lat_rows;
  if (waves_per_cu) *waves_per_cu = b->lat_waves;
  return 0;
}

int lbm_batch_get_lattice_policy(const lbm_batch* b, int32_t* cache_policy) {
  if (!b || !cache_policy) return LBM_E_INVALID;
  *cache_policy = b->lat_cp;
  return 0;
}

int lbm_batch_last_error(const lbm_batch* b, char* buf, size_t n) {
  if (!b || !buf || n == 0) return LBM_E_INVALID;
  snprintf(buf, n, "%s", b->msg);
  return 0;
}

int lbm_batch_advance(lbm_batch* b, lbm_handle* const* hs, int32_t count, int32_t steps, double* row_sums,
                      double* last_rows, double* errs, int32_t* status) try {
  if (!b) return LBM_E_INVALID;
  if (!hs || count < 1 || steps < 0) return bfail(b, LBM_E_INVALID, "need at least one instance and steps >= 0");
  const bool f32 = hs[0] && hs[0]->f32;
  int ns = 1 << 30;
  bool wings = false;
  for (int e = 0; e < count; e++) {
    lbm_handle* h = hs[e];
    if (!h) return bfail(b, LBM_E_INVALID, "instance %d is null", e);
    for (int q = 0; q < e; q++)
      if (hs[q] == h) return bfail(b, LBM_E_INVALID, "instance %d appears twice", e);
    if (h->f32 != f32) return bfail(b, LBM_E_INVALID, "instance %d has a different precision", e);
    if (h->device != b->device) return bfail(b, LBM_E_INVALID, "instance %d lives on another device", e);
    if (h->cfg.motion_source == LBM_MOTION_FILE && !h->motion_ready && h->W > 0)
      return bfail(b, LBM_E_STATE, "instance %d: motion table not loaded", e);
    if (h->W > 0 && (h->rel[REL_POS][0] >= 1.0 || h->rel[REL_POS][0] <= -1.0))
      return bfail(b, LBM_E_STATE, "instance %d has a pending move_mesh; advance it alone", e);
    if (h->dyn.enable)
      return bfail(b, LBM_E_STATE, "instance %d has rigid-body coupling on; advance it alone", e);
    if (h->W > 0) {
      wings = true;
      ns = std::min(ns, h->nslot);
    }
  }
  if (!wings) ns = 1;
  if (status)
    for (int e = 0; e < count; e++) status[e] = 0;
  if (steps == 0) {
    for (int e = 0; e < count; e++) {
      if (row_sums) std::fill(row_sums + (size_t)e * LBM_MAX_FORCE_ROW, row_sums + (size_t)(e + 1) * LBM_MAX_FORCE_ROW, 0.0);
      if (last_rows) std::fill(last_rows + (size_t)e * LBM_MAX_FORCE_ROW, last_rows + (size_t)(e + 1) * LBM_MAX_FORCE_ROW, 0.0);
      if (errs) errs[e] = 0.0;
    }
    return 0;
  }
  // per-instance call set-up: rows, frames (constant without moves), counters
  std::vector<EpDesc> desc(count);
  std::vector<int2> items;
  BatchItems it{};
  size_t lds_wing = 0, lds_solve = 0, lds_inv = 0;
  b->direct = false;
  b->bsolve_lds = 0;
  int inv_chunks = 1;
  for (int e = 0; e < count; e++) {
    lbm_handle* h = hs[e];
    BCHECK(b, hipStreamSynchronize(h->stream_geo));
    int rc = ensure_rows(h, (size_t)steps);
    if (!rc) rc = ensure_frames(h, (size_t)steps);
    if (rc) return bfail(b, rc, "instance %d: %s", e, h->msg);
    h->carry.valid = false;   // the batched steps reuse the slot sets
    h->frames.assign((size_t)steps, h->fr);
    BCHECK(b, hipMemcpy(h->d_frame, h->frames.data(), sizeof(Frame) * steps, hipMemcpyHostToDevice));
    BCHECK(b, hipMemset(h->d_rowsum, 0, sizeof(double) * h->row_len));
    BCHECK(b, hipMemset(&h->d_st->error, 0, sizeof(int)));
    EpDesc& D = desc[e];
    D.g = h->g;
    D.P = h->mrt;
    D.p = h->ip;
    D.G = h->G;
    D.F = h->F;
    D.R = h->R;
    D.st = h->d_st;
    D.fbuf[0] = h->fbuf[0];
    D.fbuf[1] = h->fbuf[1];
    for (int q = 0; q < 3; q++) D.fslot[q] = h->slot[q];
    D.cur0 = h->cur;
    D.n0 = h->n;
    D.ns = ns;
    D.bc0 = (int)(h->batch_count & 1u);
    D.tiles = 0;  // set once the tile width is known
    D.bpw = adj_collide_blocks_per_wing(h->ip);
    D.direct = h->direct == 2 ? 1 : 0;   // as the instance's own advance solves its prescribed steps
    D.pad1 = 0;
    if (D.direct && h->W > 0) {
      b->direct = true;
      b->bsolve_lds = std::max(b->bsolve_lds, bsolve_lds_bytes(h->ip));
    }
    if (h->W > 0) {
      lds_wing = std::max(lds_wing, geo_wing_lds_bytes(h->SPW));
      lds_solve = std::max(lds_solve, geo_solve_lds_bytes(h->ip, h->direct == 2));
      lds_inv = std::max(lds_inv, geo_factor_lds_bytes(h->ip));
      inv_chunks = std::max(inv_chunks, inverse_chunks(h->SPW));
    }
  }
  auto add = [&](int e, int nblk) {
    for (int q = 0; q < nblk; q++) items.push_back(make_int2(e, q));
  };
  auto upload = [&]() -> int {
    if ((size_t)count > b->desc_cap) {
      if (b->d_desc) (void)hipFree(b->d_desc);
      b->d_desc = nullptr;
      BCHECK(b, hipMalloc((void**)&b->d_desc, sizeof(EpDesc) * count));
      b->desc_cap = count;
    }
    if (items.size() > b->items_cap) {
      if (b->d_items) (void)hipFree(b->d_items);
      b->d_items = nullptr;
      BCHECK(b, hipMalloc((void**)&b->d_items, sizeof(int2) * items.size()));
      b->items_cap = items.size();
    }
    BCHECK(b, hipStreamSynchronize(b->stream));
    BCHECK(b, hipMemcpy(b->d_desc, desc.data(), sizeof(EpDesc) * count, hipMemcpyHostToDevice));
    BCHECK(b, hipMemcpy(b->d_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice));
    return 0;
  };
  if (!b->lat_bx) {  // tile width and tiles per item of the batched lattice launch, timed on this composition
    int rc = 0;
    launch_step_base(b->stream, b->d_kbase, 0, false);  // the timed launches run substep 0
    const int fixed_bx = env_tile();
    int fixed_rows = 0;
    if (const char* e = getenv("LBM_BATCH_ROWS")) fixed_rows = atoi(e);
    std::vector<int> cands;  // bx * 8 + rows
    for (int bx : {256, 128, 512})
      for (int rows : {1, 2, 4})
        if ((!fixed_bx || bx == fixed_bx) && (fixed_rows < 1 || rows == fixed_rows)) cands.push_back(bx * 8 + rows);
    if (cands.empty()) cands.push_back((fixed_bx ? fixed_bx : 256) * 8 + 1);
    const int best = autotune(b->stream, cands, [&](int c) {
      if (rc) return;
      const int bx = c / 8, rows = c % 8;
      items.clear();
      for (int e = 0; e < count; e++) {
        desc[e].tiles = lattice_tiles_flat(desc[e].g, bx);
        add(e, (desc[e].tiles + rows - 1) / rows);
      }
      if ((rc = upload())) return;
      BatchItems lt{};
      lt.kbase = b->d_kbase;
      lt.lattice = b->d_items;
      lt.n_lattice = (int)items.size();
      launch_batch_lattice(b->stream, b->d_desc, lt, 0, f32, false, bx, rows, CP_PLAIN, false);
    });
    if (rc) return rc;
    b->lat_bx = best / 8;
    b->lat_rows = best % 8;
    // then the cache policy for that shape (cp * 64 + waves).  No wave cap
    // unless LBM_LATTICE_WAVES asks for one: timed alone, a cap can win by a
    // few per cent, but in the batch's steps, beside the IBM kernels and the
    // geometry of many episodes, the uncapped kernel was faster (64 uneven
    // episodes: tile 512, one tile per item, cap none 83.9, cap 24 83.6,
    // cap 16 81.2 env-steps/s)
    const int fw = env_waves(), fcp = env_cp(), lds_cu = lds_per_cu(b->device), lds_blk = lds_per_block(b->device);
    std::vector<int> wc;
    for (int cp : {CP_PLAIN, CP_NT_STORE, CP_NT})
      for (int wv : kLatticeWaves)
        if ((fcp < 0 || cp == fcp) && (fw < 0 ? wv == 0 : wv == fw) &&
            lattice_lds_for_waves(b->lat_bx, wv, lds_cu, lds_blk) >= 0)
          wc.push_back(cp * 64 + wv);
    if (wc.empty()) wc.push_back(fcp > 0 ? fcp * 64 : 0);
    items.clear();
    for (int e = 0; e < count; e++) {
      desc[e].tiles = lattice_tiles_flat(desc[e].g, b->lat_bx);
      add(e, (desc[e].tiles + b->lat_rows - 1) / b->lat_rows);
    }
    if ((rc = upload())) return rc;
    const int bw = autotune(b->stream, wc, [&](int c) {
      BatchItems lt{};
      lt.kbase = b->d_kbase;
      lt.lattice = b->d_items;
      lt.n_lattice = (int)items.size();
      launch_batch_lattice(b->stream, b->d_desc, lt, 0, f32, false, b->lat_bx, b->lat_rows, c / 64, false, nullptr,
                           (size_t)lattice_lds_for_waves(b->lat_bx, c % 64, lds_cu, lds_blk));
    });
    b->lat_waves = bw % 64;
    b->lat_cp = bw / 64;
    b->lat_lds = (size_t)lattice_lds_for_waves(b->lat_bx, b->lat_waves, lds_cu, lds_blk);
    items.clear();
  }
  for (int e = 0; e < count; e++) desc[e].tiles = lattice_tiles_flat(desc[e].g, b->lat_bx);
  // work lists: lattice, ring, adjoint, segment, spread, then the geometry
  // lists of a full and of the last (shorter) geometry batch
  size_t off[9];
  off[0] = items.size();
  for (int e = 0; e < count; e++) add(e, (desc[e].tiles + b->lat_rows - 1) / b->lat_rows);
  off[1] = items.size();
  for (int e = 0; e < count; e++) add(e, bc_ring_blocks(desc[e].g));
  off[2] = items.size();
  for (int e = 0; e < count; e++) add(e, desc[e].p.W * desc[e].bpw);
  off[3] = items.size();
  for (int e = 0; e < count; e++)
    if (desc[e].p.W > 0) add(e, (desc[e].p.nseg + 3) / 4);
  off[4] = items.size();
  for (int e = 0; e < count; e++)
    if (desc[e].p.W > 0) add(e, spread_blocks(desc[e].p));
  off[5] = items.size();
  const int tail = steps % ns == 0 ? ns : steps % ns;
  for (int e = 0; e < count; e++)
    if (desc[e].p.W > 0) add(e, std::min(ns, steps) * desc[e].p.W);
  off[6] = items.size();
  for (int e = 0; e < count; e++)
    if (desc[e].p.W > 0) add(e, tail * desc[e].p.W);
  off[7] = items.size();
  for (int e = 0; e < count; e++)
    if (desc[e].p.W > 0) add(e, desc[e].p.W);
  off[8] = items.size();
  {
    const int rc = upload();
    if (rc) return rc;
  }
  const int2* di = b->d_items;
  it.lattice = di + off[0]; it.n_lattice = (int)(off[1] - off[0]);
  it.ring = di + off[1]; it.n_ring = (int)(off[2] - off[1]);
  it.adj = di + off[2]; it.n_adj = (int)(off[3] - off[2]);
  it.seg = di + off[3]; it.n_seg = (int)(off[4] - off[3]);
  it.spread = di + off[4]; it.n_spread = (int)(off[5] - off[4]);
  it.wing = di + off[7]; it.n_wing = (int)(off[8] - off[7]);
  it.kbase = b->d_kbase;
  // The launch sequence of the call, in segments of whole geometry batches
  // (pl steps; 16 without wings), each ending with the force row of its last
  // step.  At the start of a segment every earlier launch has joined the
  // lattice stream, so the segment's first two geometry batches wait on
  // ev_start, recorded there; later ones wait on the release of their slot
  // set by the force row of the last step of batch bb-2.  Geometry batch bb
  // is issued one batch ahead of the flow (at the first step of batch bb-1).
  // Launches take the substep relative to the device step base kb0.
  const int pl = wings ? ns : 16;
  // A geometry batch issued one batch ahead of the flow has ns steps before
  // the flow needs it: its marker pass and band factor run from a grid of
  // kGeoAheadGrid workgroups (about one per CU) that loop over the items,
  // instead of one workgroup per item (thousands, 8 per CU), which halves
  // what they take from the lattice pass running beside them.  On 64 uneven
  // episodes the lattice pass's mean exceeded its median by 16 us per step
  // with one workgroup per item and by 8 us with 256 (the geometry itself
  // taking 0.75 and 2.3 ms of the batch's 10 ms).  The first batch of a
  // call, which the flow waits for at once, keeps one workgroup per item.
  constexpr int kGeoAheadGrid = 256, kGeoAheadMinSteps = 8;
  auto issue_geometry = [&](int bb, int seg_bb, int kb0) -> int {
    const int k0 = bb * ns;
    if (!wings || k0 >= steps) return 0;
    const int es = bb & 1;
    BCHECK(b, hipStreamWaitEvent(b->stream_geo, bb < seg_bb + 2 ? b->ev_start : b->ev_rel[es], 0));
    const int q = k0 + ns >= steps ? 6 : 5;  // the last (shorter) batch has its own list
    const int grid = bb > 0 && ns >= kGeoAheadMinSteps ? kGeoAheadGrid : 0;
    launch_batch_geometry(b->stream_geo, b->d_desc, b->d_kbase, di + off[q], (int)(off[q + 1] - off[q]), k0 - kb0,
                          lds_wing, lds_solve, lds_inv, inv_chunks, grid);
    BCHECK(b, hipEventRecord(b->ev_geo[es], b->stream_geo));
    return 0;
  };
  // bind: the lattice launches carry the step event as their own completion
  // event (not inside a graph capture, where steps record it explicitly)
  auto issue_steps = [&](int k_begin, int k_end, int kb0, bool bind) -> int {
    const int seg_bb = k_begin / pl;
    BCHECK(b, hipEventRecord(b->ev_start, b->stream));
    if (k_begin == 0) {
      const int rc = issue_geometry(0, 0, kb0);
      if (rc) return rc;
    }
    for (int k = k_begin; k < k_end; k++) {
      const int kk = k - kb0;
      // the geometry of a batch starting at k_begin > 0 joined before the segment
      if (wings && k % ns == 0 && (k > k_begin || k == 0)) {
        const int es = (k / ns) & 1;
        BCHECK(b, hipStreamWaitEvent(b->stream_ibm, b->ev_geo[es], 0));
        BCHECK(b, hipStreamWaitEvent(b->stream, b->ev_geo[es], 0));  // lattice reads the claim maps
      }
      // one writer per node, as in lbm_advance: lattice (all but the adjoint
      // nodes), IBM stream (adjoint nodes), ring stream (boundary ring)
      const bool fields = k >= steps - 3;
      // ev_step: the lattice stream has finished the previous step
      if (!bind || k == k_begin) BCHECK(b, hipEventRecord(b->ev_step, b->stream));
      BCHECK(b, hipStreamWaitEvent(b->stream_ibm, b->ev_step, 0));
      BCHECK(b, hipStreamWaitEvent(b->stream_bc, b->ev_step, 0));
      launch_batch_bc_ring(b->stream_bc, b->d_desc, it, kk, f32, fields);
      BCHECK(b, hipEventRecord(b->ev_bc, b->stream_bc));
      if (k > k_begin) {
        launch_batch_force_row(b->stream_ibm, b->d_desc, b->d_kbase, count, kk - 1);
        if (wings && k % ns == 0) BCHECK(b, hipEventRecord(b->ev_rel[((k - 1) / ns) & 1], b->stream_ibm));
      }
      if (wings && k % ns == 0) {
        const int rc = issue_geometry(k / ns + 1, seg_bb, kb0);
        if (rc) return rc;
      }
      launch_batch_ibm_flow(b->stream_ibm, b->d_desc, it, kk, f32, b->direct ? b->bsolve_lds : 0);
      launch_batch_adj_collide(b->stream_ibm, b->d_desc, it, kk, f32, fields);
      BCHECK(b, hipStreamWaitEvent(b->stream_ibm, b->ev_bc, 0));
      BCHECK(b, hipEventRecord(b->ev_ibm, b->stream_ibm));
      launch_batch_lattice(b->stream, b->d_desc, it, kk, f32, fields, b->lat_bx, b->lat_rows, b->lat_cp, true,
                           bind ? b->ev_step : nullptr, b->lat_lds);
      BCHECK(b, hipGetLastError());
      BCHECK(b, hipStreamWaitEvent(b->stream, b->ev_ibm, 0));
    }
    // the geometry of the batch after the segment was issued inside it
    if (wings && k_end < steps) BCHECK(b, hipStreamWaitEvent(b->stream, b->ev_geo[(k_end / ns) & 1], 0));
    BCHECK(b, hipEventRecord(b->ev_ibm, b->stream_ibm));
    BCHECK(b, hipStreamWaitEvent(b->stream, b->ev_ibm, 0));
    launch_batch_force_row(b->stream, b->d_desc, b->d_kbase, count, k_end - 1 - kb0);
    return 0;
  };
  // Captured graphs.  A call of at most kMaxGraphSteps steps within one
  // geometry batch is captured whole (on the second call of the same
  // composition and length).  With graph_mode 2 a longer call runs its first
  // batch and its last two eagerly, and every batch in between (no field
  // writes, the next batch a full one) by replaying one graph of pl steps
  // that also computes the geometry of the next batch and ends by advancing
  // the device step base by pl; the graph depends only on the composition.
  // Off by default: the graph executor spreads the nodes over its own queues,
  // and 64 uneven episodes ran 9 % slower that way than from the four streams
  // launch by launch.
  // The key: the composition, the device lists and, for a whole advance,
  // its length and the last geometry list (a period graph uses the full
  // list only).
  std::vector<long long> key = {count, ns, f32, wings, b->lat_bx, b->lat_rows, b->lat_waves, b->lat_cp, b->direct,
                                (long long)(size_t)b->d_desc, (long long)(size_t)b->d_items, (long long)lds_wing,
                                (long long)lds_solve};
  for (int q = 0; q < 7; q++) key.push_back((long long)off[q]);
  key.push_back((long long)off[8]);
  for (int e = 0; e < count; e++) key.push_back((long long)(size_t)hs[e]);
  auto find_graph = [&](long long tag) -> BatchGraph* {
    std::vector<long long> kq = key;
    kq.push_back(tag);
    if (tag > 0) kq.push_back((long long)off[7]);
    for (auto& g : b->graphs)
      if (g.key == kq) {
        g.last_use = ++b->calls;
        return &g;
      }
    if (b->graphs.size() >= 8) {  // evict the least recently used
      auto lru = std::min_element(b->graphs.begin(), b->graphs.end(), [](const BatchGraph& x, const BatchGraph& y) {
        return x.last_use < y.last_use;
      });
      if (lru->exec) (void)hipGraphExecDestroy(lru->exec);
      b->graphs.erase(lru);
    }
    b->graphs.push_back(BatchGraph{kq, 0, nullptr, ++b->calls});
    return &b->graphs.back();
  };
  auto capture = [&](BatchGraph* gr, const std::function<int()>& body) -> int {
    hipGraph_t graph = nullptr;
    BCHECK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    const hipError_t ec = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ec != hipSuccess) return bfail(b, LBM_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&gr->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      gr->exec = nullptr;
      return bfail(b, LBM_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    }
    return 0;
  };
  // periods replayed from the period graph: 1 .. J
  int J = 0;
  if (b->graph_mode >= 2 && steps > pl) {
    J = (steps - 1) / pl - 2;
    while (J > 0 && (J + 1) * pl > steps - 3) J--;
  }
  if (J >= 1) {
    const int kc = (J + 1) * pl;
    launch_step_base(b->stream, b->d_kbase, 0, false);
    int rc = issue_steps(0, pl, 0, true);
    if (rc) return rc;
    launch_step_base(b->stream, b->d_kbase, pl, true);
    BatchGraph* gr = find_graph(-pl);
    if (!gr->exec) {
      rc = capture(gr, [&]() -> int {
        const int r = issue_steps(pl, 2 * pl, pl, false);
        if (!r) launch_step_base(b->stream, b->d_kbase, pl, true);
        return r;
      });
      if (rc) return rc;
    }
    for (int j = 1; j <= J; j++) BCHECK(b, hipGraphLaunch(gr->exec, b->stream));
    gr->seen += J;
    rc = issue_steps(kc, steps, kc, true);
    if (rc) return rc;
  } else {
    auto whole = [&](bool bind) -> int {
      launch_step_base(b->stream, b->d_kbase, 0, false);
      return issue_steps(0, steps, 0, bind);
    };
    BatchGraph* gr = nullptr;
    if (b->graph_mode >= 1 && steps <= kMaxGraphSteps && (!wings || steps <= ns)) gr = find_graph(steps);
    if (gr && !gr->exec && gr->seen >= 1) {
      const int rc = capture(gr, [&]() { return whole(false); });
      if (rc) return rc;
    }
    if (gr && gr->exec) {
      BCHECK(b, hipGraphLaunch(gr->exec, b->stream));
    } else {
      const int rc = whole(true);
      if (rc) return rc;
      if (gr) gr->seen++;
    }
  }
  // host state of every instance as if it had run the steps alone
  std::vector<char> have_err(count, 0);
  for (int e = 0; e < count; e++) {
    lbm_handle* h = hs[e];
    const int n0 = h->n, n_last = n0 + steps - 1;
    for (int m = std::max(n0, n_last - 2); m <= n_last; m++) h->slot_step[m % 3] = m;
    h->cur = (h->cur + steps) & 1;
    if (h->W > 0) {
      h->last_slot = ep_slot(desc[e], steps - 1);
      h->batch_count += (unsigned)((steps + ns - 1) / ns);
    }
    h->n = n_last + 1;
    h->flag2 = std::max(0, h->flag2 - steps);
    have_err[e] = issue_err(h, n_last, steps, b->stream);
  }
  BCHECK(b, hipGetLastError());
  BCHECK(b, hipStreamSynchronize(b->stream));
  int first_bad = -1, first_code = 0;
  for (int e = 0; e < count; e++) {
    lbm_handle* h = hs[e];
    DevStep st_host;
    BCHECK(b, hipMemcpy(&st_host, h->d_st, sizeof(DevStep), hipMemcpyDeviceToHost));
    if (st_host.error != 0) {
      const int code = fail(h, st_host.error == DERR_MOTION ? LBM_E_STATE
                                   : (st_host.error == DERR_NOT_SPD ? LBM_E_NUMERIC : LBM_E_RANGE),
                            "step failed: %s (info %d)", dev_error_text(st_host.error), st_host.error_info);
      if (status) status[e] = code;
      if (first_bad < 0) {
        first_bad = e;
        first_code = code;
      }
      continue;
    }
    double errv = 0.0;
    if (have_err[e] && h->last_move_step != h->n - 2)
      BCHECK(b, hipMemcpy(&errv, h->err_out, sizeof(double), hipMemcpyDeviceToHost));
    if (errs) errs[e] = errv;
    if (row_sums) {
      double* r = row_sums + (size_t)e * LBM_MAX_FORCE_ROW;
      std::fill(r, r + LBM_MAX_FORCE_ROW, 0.0);
      BCHECK(b, hipMemcpy(r, h->d_rowsum, sizeof(double) * h->row_len, hipMemcpyDeviceToHost));
    }
    if (last_rows) {
      double* r = last_rows + (size_t)e * LBM_MAX_FORCE_ROW;
      std::fill(r, r + LBM_MAX_FORCE_ROW, 0.0);
      BCHECK(b, hipMemcpy(r, h->d_rows + (size_t)(steps - 1) * h->row_len, sizeof(double) * h->row_len,
                          hipMemcpyDeviceToHost));
    }
  }
  if (first_bad >= 0) return bfail(b, first_code, "instance %d: %s", first_bad, hs[first_bad]->msg);
  return 0;
} LBM_BATCH_CATCH(b, "lbm_batch_advance")

}  // extern "C"
