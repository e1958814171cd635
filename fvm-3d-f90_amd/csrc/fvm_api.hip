// fvm_api.hip -- C ABI of the MI355X residual + RK path (include/fvm_gpu.h):
// contexts, blocks and their connections, host transfers, step orchestration,
// the autotuner and the small auxiliary kernels (ghost fill, CFL reduction,
// dt update).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/fvm_gpu.h"
#include "fvm_launch.hpp"
#include "fvm_stage.hpp"

using namespace fvm;
using namespace fvm::exact;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(FVM_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                       \
  } while (0)

constexpr int kGhost = 3;   // ghost layers per side (src/base/defs_ghost_cell.f90:42)
// Row layout of a component array: real vertex 0 of every row starts a
// 128-byte line (16 doubles; the i- ghosts sit in the padding before it) and
// the row pitch is a whole number of lines, so a tile row of 16 or 32
// vertices is one or two whole lines for the loads and stores of every
// plane (unaligned rows straddle a line more: 150.9 -> 145.4 us per 128^3
// vortex launch).
constexpr int kRowAlign = 16;
constexpr int kOx = kRowAlign;   // array offset of real vertex 0 along i
int64_t row_pitch(int64_t ni) { return (kOx + ni + kGhost + kRowAlign - 1) / kRowAlign * kRowAlign; }
constexpr int kSnapSlots = 4;   // snapshots in flight per context
constexpr int kSnapMaxVars = 10;
const SmaxWays kZeroSmax{};     // uploaded to clear a wave-speed maximum

}  // namespace

struct fvm_block {
  fvm_ctx* ctx = nullptr;
  int dim = 3, nc = 5;
  BlockGeom g{};
  double* q[4] = {nullptr, nullptr, nullptr, nullptr};   // u^n, stage a, stage b, RK4 acc
  size_t state_elems = 0;                                // nc * cs
  bool uploaded = false;
  int nbr[6] = {-1, -1, -1, -1, -1, -1};                // neighbour block index (context order)
  int nbr_shared[6] = {0, 0, 0, 0, 0, 0};
  // launch shape
  int tile = 0, chunk = 16;
  int tiles_x = 0, tiles_y = 0, tiles_z = 0, ntiles = 0, tile_base = 0;
  // independent solve (serving): own dt and CFL accumulator on the device
  const double* dt_dev = nullptr;
  SmaxWays* smax_dev = nullptr;
  const int* halt_dev = nullptr;
  double* rv = nullptr;                                  // viscous residual (nc - 1 components)
  // the four state arrays are one stream-ordered allocation starting at q[0]
  // (serving); otherwise each is its own hipMalloc
  bool pooled = false;
};

struct fvm_ctx {
  int device = 0;
  int num_cu = 256;
  hipStream_t stream = nullptr;
  bool scheme_set = false;
  bool fast = false;       // FVM_NUMERICS_FAST: FMA/reciprocal kernel family
  int launch_mode = FVM_LAUNCH_BATCHED;
  int recon = 0, flux = 0, rk = 0;
  Params P{};
  std::vector<fvm_block*> blocks;
  // 1-to-1 connection patches (fvm_conn_add): host list (block pointers),
  // device table rebuilt with the descriptor table
  struct Patch {
    fvm_block* recv;
    fvm_block* donor;
    int face;
    int64_t rmin[3], rmax[3], dmin[3];
    int map_dir[3], map_sign[3];
    int bc;              // 0: a connection; else the boundary kind that made it
                         // (wall / symmetry: a mirror onto the block itself;
                         // inflow / farfield: a fixed state)
    bool fixed;
    double prim[5];      // fixed state, primitive (rho, u, v, w, p)
  };
  std::vector<Patch> patches;
  // asynchronous snapshots (fvm_snapshot_*): slots of device and pinned host
  // memory; `id` 0 marks a free slot
  struct Snap {
    int64_t id = 0;
    int nvars = 0;
    int which[kSnapMaxVars] = {};     // snapshot_kernel selector per variable
    int codes[kSnapMaxVars] = {};     // the caller's variable codes
    double* d_buf = nullptr;
    double* h_buf = nullptr;
    size_t cap = 0;                   // doubles in d_buf and h_buf
    hipEvent_t conv = nullptr;        // snapshot_kernel has written d_buf
    hipEvent_t done = nullptr;        // the copy into h_buf has landed
    double time = 0.0;
    int64_t step = 0;
    struct Part { fvm_block* blk; int64_t off, n[3]; double h[3]; };
    std::vector<Part> parts;
    std::thread writer;               // fvm_snapshot_save
    int write_rc = 0;
    std::string write_err;
  } snaps[kSnapSlots];
  hipStream_t copy_stream = nullptr;
  int64_t next_snap = 1;
  PatchDesc* d_patch = nullptr;
  size_t patch_cap = 0;
  int64_t patch_points = 0;        // ghost points over all patches
  BlockDesc* d_desc = nullptr;           // descriptor table on the device
  size_t desc_cap = 0;
  bool desc_dirty = true;                // table must be rebuilt before a launch
  bool ghosts_dirty = true;              // u^n ghost layers must be refilled
  bool ghosts_stale = false;             // steps ran: own ghost layers (not read by the
                                         // stages) are refilled before a download with ghosts
  int total_tiles = 0;
  bool has_conn = false;                 // some block face is connected to another block
  bool has_wall = false;                 // some block face is a wall (slip or no-slip)
  bool has_noslip = false;               // some block face is a no-slip wall
  // central viscous terms (fvm_viscous_set); mu_ref <= 0: inviscid
  double visc_mu_ref = 0.0, visc_t_ref = 1.0, visc_s = 0.0, visc_pr = 0.72, visc_r = 1.0;
  SolveSlot* d_slot = nullptr;            // device time-stepping state (CFL mode)
  SolveSlot h_slot{};                     // host staging of d_slot
  double* d_dt = nullptr;                 // &d_slot->dt: CFL dt of the current step
  SmaxWays* d_smax = nullptr;             // wave-speed max of the current state (CFL mode)
  double time = 0.0;                      // solution time after the last step
  double t_end = HUGE_VAL;                // CFL steps stop here (halt), +inf: never
  int64_t steps_done = 0;
  // fvm_stage: the next stage of the step in progress (1: none in progress)
  // and that step's dt
  int stage_next = 1;
  double stage_dt = 0.0;
  // hipGraph replay of whole time steps (fvm_set_graph): one graph per
  // (norm partials wanted) flavour, re-captured when `gen` or the key moves
  bool use_graph = false;
  uint64_t gen = 1;                       // bumped by every change of launch configuration
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    uint64_t gen = 0;
    bool cfl_mode = false;
  } graphs[3];   // per norm mode of launch_step
  double* d_norm = nullptr;               // per-tile partial sums of rhs^2
  size_t norm_cap = 0;                    // doubles
  std::vector<double> h_norm;
  double norms[6] = {0, 0, 0, 0, 0, 0};
  int nan_flag = 0;
  bool have_norms = false;
  // stage-kernel timing: HIP events around the stage launches of every
  // timing_every-th step (0: off); timing_step counts steps since enabled
  int timing_every = 0;
  int64_t timing_step = 0;
  std::vector<hipEvent_t> ev;
  double stage_ms = 0.0;
  int64_t stage_launches = 0;
  double cell_updates = 0.0;
};

// ============================================================================
// auxiliary kernels
// ============================================================================
namespace {

// Ghost layers of one face region of block `bi`, state array `buf`:
// extrapolation copies the boundary vertex (src/base/defs_block_data.f90:
// 226-256), periodic the vertex n periods away, a connected face the donor
// vertex of the neighbour block.
__global__ void ghost_fill_kernel(const BlockDesc* __restrict__ tab, int bi, int buf, int nc, int d,
                                  int side) {
  const BlockDesc& B = tab[bi];
  const BlockGeom& g = B.g;
  double* q = B.q[buf];
  const int n[3] = {g.nx, g.ny, g.nz};
  const int a_dim = d == 0 ? 1 : 0, b_dim = d == 2 ? 1 : 2;
  const int na = n[a_dim], nb = n[b_dim];
  const int64_t total = (int64_t)na * nb * g.ng;
  const int kind = g.bc[2 * d + side];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int aa = (int)(e % na);
    const int bb = (int)((e / na) % nb);
    const int l = 1 + (int)(e / ((int64_t)na * nb));
    int x[3], s[3];
    x[a_dim] = aa; x[b_dim] = bb;
    s[a_dim] = aa; s[b_dim] = bb;
    x[d] = side == 0 ? -l : n[d] - 1 + l;
    const double* src = q;
    int64_t cs = g.cs;
    int64_t si;
    if (kind == BC_CONN) {
      const int nbk = B.nbr[2 * d + side];
      const BlockGeom& h = tab[nbk].g;
      const int hn[3] = {h.nx, h.ny, h.nz};
      const int sh = B.nbr_shared[2 * d + side];
      // low face: ghost -l is the neighbour's vertex n-1-l (shared) or n-l;
      // high face: ghost n-1+l is the neighbour's vertex l (shared) or l-1
      s[d] = side == 0 ? (sh ? hn[d] - 1 - l : hn[d] - l) : (sh ? l : l - 1);
      src = tab[nbk].q[buf];
      cs = h.cs;
      si = (int64_t)(s[2] + h.oz) * h.pxy + (int64_t)(s[1] + h.oy) * h.px + (s[0] + h.ox);
    } else {
      if (kind == BC_PERIODIC) s[d] = side == 0 ? n[d] - l : l - 1;
      else s[d] = side == 0 ? 0 : n[d] - 1;
      si = gidx(g, s[0], s[1], s[2]);
    }
    const int64_t di = gidx(g, x[0], x[1], x[2]);
    for (int c = 0; c < nc; ++c) q[di + c * g.cs] = src[si + c * cs];
  }
}

// Ghost layers of the 1-to-1 connection patches, state array `buf`: point e
// of the launch's list belongs to the patch whose range contains it (patches
// listed by their first point, `start`); the receiver's ghost index maps
// linearly onto the donor's interior (src/base/defs_face_geom.f90:361-403:
// donor ranges and direction transform).
__global__ void patch_copy_kernel(const BlockDesc* __restrict__ tab, const PatchDesc* __restrict__ pd,
                                  int npatch, int64_t total, int buf) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    // the last patch starting at or before e (starts ascend from 0)
    int lo = 0, hi = npatch - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pd[mid].start <= e) lo = mid;
      else hi = mid - 1;
    }
    const PatchDesc& P = pd[lo];
    const int64_t loc = e - P.start;
    const int ng = tab[P.recv].g.ng;
    const int a = (int)(loc % P.na);
    const int b = (int)((loc / P.na) % P.nb);
    const int l = 1 + (int)(loc / ((int64_t)P.na * P.nb));
    if (l > ng) continue;
    int64_t r[3] = {P.rmin[0], P.rmin[1], P.rmin[2]};
    r[P.ta] += a;
    r[P.tb] += b;
    r[P.dn] += P.side ? l : -l;
    int64_t d[3];
#pragma unroll
    for (int ex = 0; ex < 3; ++ex) d[ex] = 0;
#pragma unroll
    for (int ex = 0; ex < 3; ++ex)
      d[P.map_dir[ex]] = P.dmin[P.map_dir[ex]] + P.map_sign[ex] * (r[ex] - P.rmin[ex]);
    const BlockGeom& gr = tab[P.recv].g;
    double* dst = tab[P.recv].q[buf];
    if (P.fixed_state) {
      const int64_t ri = gidx(gr, (int)r[0], (int)r[1], (int)r[2]);
      for (int c = 0; c < P.nc; ++c) dst[ri + c * gr.cs] = P.fixed[c];
      continue;
    }
    const BlockGeom& gd = tab[P.donor].g;
    const double* src = tab[P.donor].q[buf];
    const int64_t ri = gidx(gr, (int)r[0], (int)r[1], (int)r[2]);
    const int64_t di = gidx(gd, (int)d[0], (int)d[1], (int)d[2]);
    // momentum is stored along each block's own index directions: the
    // receiver's component along direction e is the donor's along
    // map_dir[e], times map_sign[e]
    for (int c = 0; c < P.nc; ++c) {
      int sc = c;
      double sg = 1.0;
      if (c >= 1 && c <= P.nc - 2) {
        sc = 1 + P.map_dir[c - 1];
        sg = P.neg_all ? -1.0 : P.map_sign[c - 1];
      }
      dst[ri + c * gr.cs] = sg * src[di + sc * gd.cs];
    }
  }
}

// Compact state (nc components of nx*ny*nz real vertices, i fastest) <->
// the ghost-padded block layout.  `to_block`: src compact, dst block arrays;
// otherwise the reverse.
__global__ __launch_bounds__(256) void compact_kernel(const double* __restrict__ src,
                                                     double* __restrict__ dst, BlockGeom g, int nc,
                                                     int to_block) {
  const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nc * ncell;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / ncell, r = e % ncell;
    const int i = (int)(r % g.nx);
    const int j = (int)((r / g.nx) % g.ny);
    const int k = (int)(r / ((int64_t)g.nx * g.ny));
    const int64_t bi = c * g.cs + gidx(g, i, j, k);
    if (to_block) dst[bi] = src[e];
    else dst[e] = src[bi];
  }
}

// Snapshot variables of one block's u^n (real vertices) into `out`, one
// contiguous array per variable, i fastest: the reference's output array
// vdata(ni, nj, nk, nvars) (src/solve/class_output.f90:377-410).  Selector
// 0..4: density, velocity x/y/z, static pressure (q -> p as
// src/solve/class_flow_field.f90:383-396); 5..9: the conserved components
// rho, rho u, rho v, rho w, rho E (2-D: w = rho w = 0).
struct SnapSel {
  int n;
  int which[kSnapMaxVars];
};

template <int NC>
__global__ __launch_bounds__(256) void snapshot_kernel(const double* __restrict__ q, BlockGeom g,
                                                      double gamma, SnapSel sel,
                                                      double* __restrict__ out) {
  const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ncell;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.nx);
    const int j = (int)((e / g.nx) % g.ny);
    const int k = (int)(e / ((int64_t)g.nx * g.ny));
    const int64_t ci = gidx(g, i, j, k);
    double v[NC], pr[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = q[ci + c * g.cs];
    cons_to_prim<NC>(v, pr, gamma);
    double all[10];
    if constexpr (NC == 5) {
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        all[c] = pr[c];
        all[5 + c] = v[c];
      }
    } else {
      all[0] = pr[0]; all[1] = pr[1]; all[2] = pr[2]; all[3] = 0.0; all[4] = pr[3];
      all[5] = v[0]; all[6] = v[1]; all[7] = v[2]; all[8] = 0.0; all[9] = v[3];
    }
    for (int n = 0; n < sel.n; ++n) out[n * ncell + e] = all[sel.which[n]];
  }
}

// ---- initial states computed on the device ---------------------------------
constexpr int kInitMaxRegions = 16;
static_assert(11 * kInitMaxRegions == FVM_IC_MAX_PARAMS, "region parameter capacity");
struct InitSpec {
  int kind, nregions;
  double origin[3];
  int at_index;                    // 1: x = dx (first[0] + i), ... (origin unused)
  long long first[3];
  double gamma;
  // regions: per region xmin, xmax, ymin, ymax, zmin, zmax, rho, u, v, w, p;
  // vortex: length, beta, xc, yc, u_inf, v_inf, w_inf; TGV: Mach number
  double v[11 * kInitMaxRegions];
};

// Python's float modulo for a positive period: the result takes the sign of
// the period
__device__ __forceinline__ double floor_mod(double a, double period) {
  double r = fmod(a, period);
  if (r != 0.0) {
    if (r < 0.0) r += period;
  } else {
    r = 0.0;
  }
  return r;
}

// Primitive state of one vertex at (x, y, z) for the given initial state.
__device__ void init_state(const InitSpec& s, double x, double y, double z, double p[5]) {
  const double pi = 3.141592653589793;
  if (s.kind == FVM_IC_REGIONS) {
    // src/base/defs_block_data.f90:166-201: regions in order, inclusive bounds
    p[0] = p[1] = p[2] = p[3] = p[4] = 0.0;
    for (int r = 0; r < s.nregions; ++r) {
      const double* b = s.v + 11 * r;
      if (x >= b[0] && x <= b[1] && y >= b[2] && y <= b[3] && z >= b[4] && z <= b[5])
        for (int c = 0; c < 5; ++c) p[c] = b[6 + c];
    }
  } else if (s.kind == FVM_IC_VORTEX) {
    // Yee's isentropic vortex about (xc, yc), periodic in x and y with period
    // `length`; free stream rho = p = 1
    const double L = s.v[0], beta = s.v[1], g = s.gamma;
    const double dxv = floor_mod(x - s.v[2] + 0.5 * L, L) - 0.5 * L;
    const double dyv = floor_mod(y - s.v[3] + 0.5 * L, L) - 0.5 * L;
    const double r2 = dxv * dxv + dyv * dyv;
    const double e = exp(0.5 * (1.0 - r2));
    const double t = 1.0 - (g - 1.0) * beta * beta / (8.0 * g * pi * pi) * e * e;
    const double rho = pow(t, 1.0 / (g - 1.0));
    p[0] = rho;
    p[1] = s.v[4] - beta / (2.0 * pi) * e * dyv;
    p[2] = s.v[5] + beta / (2.0 * pi) * e * dxv;
    p[3] = s.v[6];
    p[4] = pow(rho, g);
  } else {
    // Taylor-Green vortex, rho = p / (R T0) with R = 1 / (gamma Ma^2), T0 = 1
    const double r = 1.0 / (s.gamma * s.v[0] * s.v[0]);
    const double pr = r + (cos(2.0 * x) + cos(2.0 * y)) * (cos(2.0 * z) + 2.0) / 16.0;
    p[0] = pr / r;
    p[1] = sin(x) * cos(y) * cos(z);
    p[2] = -cos(x) * sin(y) * cos(z);
    p[3] = 0.0;
    p[4] = pr;
  }
}

// u^n of the real vertices of one block from an initial-state spec; p -> q
// in the order of src/solve/class_flow_field.f90:353-377.
__global__ __launch_bounds__(256) void init_kernel(double* __restrict__ q, BlockGeom g, int nc,
                                                  InitSpec s) {
  const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ncell;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.nx);
    const int j = (int)((e / g.nx) % g.ny);
    const int k = (int)(e / ((int64_t)g.nx * g.ny));
    double p[5];
    if (s.at_index)   // one product, as the reference's grid x = dx (i - 1)
      init_state(s, g.dx * (double)(s.first[0] + i), g.dy * (double)(s.first[1] + j),
                 g.dz * (double)(s.first[2] + k), p);
    else
      init_state(s, s.origin[0] + g.dx * (double)i, s.origin[1] + g.dy * (double)j,
                 s.origin[2] + g.dz * (double)k, p);
    const double rho = p[0], u = p[1], v = p[2], w = p[3];
    const double en = p[4] / (s.gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
    const int64_t ci = gidx(g, i, j, k);
    q[ci] = rho;
    q[ci + g.cs] = rho * u;
    q[ci + 2 * g.cs] = rho * v;
    if (nc == 5) {
      q[ci + 3 * g.cs] = rho * w;
      q[ci + 4 * g.cs] = en;
    } else {
      q[ci + 3 * g.cs] = en;
    }
  }
}

// Max over real vertices of sum_d (|u_d| + c) / dr_d.
template <int NC>
__global__ __launch_bounds__(256) void cfl_max_kernel(const double* __restrict__ q, BlockGeom g,
                                                     double gamma, SmaxWays* smax) {
  __shared__ double red[4];
  const int64_t total = (int64_t)g.nx * g.ny * g.nz;
  double s = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.nx);
    const int j = (int)((e / g.nx) % g.ny);
    const int k = (int)(e / ((int64_t)g.nx * g.ny));
    const int64_t ci = gidx(g, i, j, k);
    double v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = q[ci + c * g.cs];
    const double w = wave_speed<NC>(v, g, gamma);
    s = nan_max(s, w);
  }
  s = wg_max<256>(s, red);
  if (threadIdx.x == 0) smax_add(smax, s);
}

// Time step of one served solve for the step about to start: dt =
// min(cfl / smax, t_end - t), t += dt; clears the wave-speed accumulator so
// the step's last stage collects the next maximum.  Once t reaches t_end the
// slot halts and the step's stage launches skip the solve.  (A plain
// context's stages do the same work themselves: stage_clock.)
__device__ __forceinline__ void advance_slot(SolveSlot& s, SmaxWays* smax) {
  if (s.halt) return;
  // (an infinite t_end never ends: inf - tol * inf would be NaN)
  if (isfinite(s.t_end) && !(s.t < s.t_end - kEndTol * fabs(s.t_end))) { s.halt = 1; return; }
  const double sm = smax_get(smax);
  double dt = s.cfl / sm;
  if (!(dt > 0.0) || !(dt < 1.0e300)) s.nan = 1;
  const double rest = s.t_end - s.t;
  if (dt > rest) dt = rest;
  s.dt = dt;
  s.t = s.t + dt;
  smax_clear(smax);
  s.steps = s.steps + 1;
}

// Fixed time step: the stage kernels read dt from the slot in either mode.
__global__ void set_dt_kernel(SolveSlot* slot, double dt) { slot->dt = dt; }

// The same for the solves of a serving group (slot ids listed in ids).  A
// solve whose dt is NaN, zero or infinite halts at once: its state stays the
// last one computed, and no later launch writes it while it is retired.
__global__ void solve_dt_kernel(SolveSlot* slots, SmaxWays* smax, const int* ids, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    SolveSlot& sl = slots[ids[i]];
    advance_slot(sl, smax + ids[i]);
    if (sl.nan) sl.halt = 1;
  }
}

}  // namespace

// ============================================================================
// helpers
// ============================================================================
namespace {

int alloc_rv(fvm_block* b);

int check_ctx(fvm_ctx* ctx) {
  if (!ctx) return fail(FVM_EINVAL, "null context");
  return FVM_OK;
}

int set_device(fvm_ctx* ctx) {
  HIP_TRY(hipSetDevice(ctx->device));
  return FVM_OK;
}

int block_index(fvm_ctx* ctx, fvm_block* b) {
  for (size_t i = 0; i < ctx->blocks.size(); ++i)
    if (ctx->blocks[i] == b) return (int)i;
  return -1;
}

// Launch shape of a block.  3-D: tile (TX x TY columns, kTile3d) marching
// over `chunk` planes -- 32 x 8 unless 64 wastes fewer columns, the chunk is
// the longest (<= 32 planes) that still gives >= 512 tiles.  2-D (kTile2d):
// shape 0, rows of 62 columns marching over `chunk` rows, or a TX x TY tile of
// the plane.  FVM_TILE / FVM_TILE2 / FVM_CHUNK override.
void plan_launch(fvm_block* b, int force_tile = -1, int force_chunk = -1) {
  const BlockGeom& g = b->g;
  const char* et = getenv(b->dim == 2 ? "FVM_TILE2" : "FVM_TILE");
  const char* ec = getenv("FVM_CHUNK");
  const int ntile_shapes = b->dim == 2 ? kNumTiles2d : kNumTiles3d;
  int tile_env = et ? std::min(std::max(atoi(et), 0), ntile_shapes - 1) : -1;
  int chunk_env = ec ? std::max(1, atoi(ec)) : -1;
  if (force_tile >= 0) tile_env = force_tile;
  if (force_chunk > 0) chunk_env = force_chunk;
  if (b->dim == 2) {
    const int tile = tile_env >= 0 ? tile_env : 0;
    b->tile = tile;
    if (tile == 0) {
      b->tiles_x = (g.nx + 61) / 62;
      b->tiles_y = 1;
      int chunk = 16;
      while (chunk > 2 && (int64_t)b->tiles_x * ((g.ny + chunk - 1) / chunk) < 2048) chunk /= 2;
      if (chunk_env > 0) chunk = chunk_env;
      b->chunk = chunk;
      b->tiles_z = (g.ny + chunk - 1) / chunk;
    } else {
      b->tiles_x = (g.nx + kTile2d[tile][0] - 1) / kTile2d[tile][0];
      b->tiles_y = (g.ny + kTile2d[tile][1] - 1) / kTile2d[tile][1];
      b->chunk = 1;
      b->tiles_z = 1;
    }
  } else {
    const int w64 = (g.nx + 63) / 64 * 64 - g.nx, w32 = (g.nx + 31) / 32 * 32 - g.nx;
    int tile = w32 <= w64 ? 1 : 0;
    if (tile_env >= 0 && (!b->ctx || tile3d_ok(b->ctx->recon, tile_env))) tile = tile_env;
    b->tile = tile;
    const int tx = kTile3d[tile][0], ty = kTile3d[tile][1];
    b->tiles_x = (g.nx + tx - 1) / tx;
    b->tiles_y = (g.ny + ty - 1) / ty;
    int chunk = 32;
    while (chunk > 8 && (int64_t)b->tiles_x * b->tiles_y * ((g.nz + chunk - 1) / chunk) < 512)
      chunk /= 2;
    if (chunk_env > 0) chunk = chunk_env;
    b->chunk = chunk;
    b->tiles_z = (g.nz + chunk - 1) / chunk;
  }
  b->ntiles = b->tiles_x * b->tiles_y * b->tiles_z;
  if (b->ctx) b->ctx->desc_dirty = true;
}

// Release a block's device arrays: stream-ordered on `stream` for a pooled
// block (null: free and wait), hipFree otherwise.
void free_block_state(fvm_block* b, hipStream_t stream) {
  if (b->pooled) {
    if (b->q[0]) {
      if (stream) {
        (void)hipFreeAsync(b->q[0], stream);
      } else {
        (void)hipFreeAsync(b->q[0], b->ctx->stream);
        (void)hipStreamSynchronize(b->ctx->stream);
      }
    }
  } else {
    for (double* p : b->q) if (p) (void)hipFree(p);
  }
  for (double*& p : b->q) p = nullptr;
  if (b->rv) (void)hipFree(b->rv);
  b->rv = nullptr;
}

// Viscous residual buffer of a block: nc - 1 components (momentum, energy).
int alloc_rv(fvm_block* b) {
  if (b->rv) return FVM_OK;
  const size_t bytes = (size_t)(b->nc - 1) * b->g.cs * sizeof(double);
  if (hipMalloc(&b->rv, bytes) != hipSuccess) {
    b->rv = nullptr;
    return fail(FVM_ENOMEM, "cannot allocate %zu bytes of viscous residual", bytes);
  }
  HIP_TRY(hipMemsetAsync(b->rv, 0, bytes, b->ctx->stream));
  return FVM_OK;
}

// Block creation.  `alloc_stream` null: four hipMalloc'd state arrays, zeroed
// on the context's stream.  Otherwise (serving) one stream-ordered allocation
// on `alloc_stream` holds the four arrays, zeroed on that stream: no device-wide
// synchronisation, and the memory returns to the device pool when freed.
// Extents and spacing a block accepts (also checked when a served request is
// submitted, so admission cannot meet a block it refuses).
int check_block_dims(int64_t ni, int64_t nj, int64_t nk, double dx, double dy, double dz) {
  if (ni < 4 || nj < 4 || nk < 1)
    return fail(FVM_EINVAL, "block needs ni, nj >= 4 and nk >= 1 (got %lld %lld %lld)",
                (long long)ni, (long long)nj, (long long)nk);
  if (ni > (1 << 24) || nj > (1 << 24) || nk > (1 << 24))
    return fail(FVM_EINVAL, "block extent too large");
  const int dim = nk == 1 ? 2 : 3;
  if (dim == 3 && nk < 4) return fail(FVM_EINVAL, "3-D block needs nk >= 4 (got %lld)", (long long)nk);
  if (!(dx > 0) || !(dy > 0) || (dim == 3 && !(dz > 0)))
    return fail(FVM_EINVAL, "vertex spacing must be positive");
  // kernels address a component with a 32-bit byte offset
  const int64_t cs = row_pitch(ni) * (nj + 2 * kGhost) * (dim == 3 ? nk + 2 * kGhost : 1);
  if (cs > (int64_t(1) << 29) - 64)
    return fail(FVM_EINVAL, "block of %lld vertices per component (with ghosts) exceeds 2^29; "
                "split the grid into connected blocks", (long long)cs);
  return FVM_OK;
}

// Scheme codes and gamma fvm_scheme_set accepts (also checked at submit).
int check_scheme(int recon, int limiter, int flux, int rk, double gamma) {
  if (recon != RECON_SIMPLE && recon != RECON_MUSCL && recon != RECON_WENO)
    return fail(FVM_EINVAL, "unknown reconstruction code %d", recon);
  if (recon == RECON_MUSCL && limiter != LIM_NONE && limiter != LIM_MINMOD &&
      limiter != LIM_VANLEER && limiter != LIM_VANALBADA && limiter != -1)
    return fail(FVM_EINVAL, "unknown limiter code %d", limiter);
  if (flux != FLUX_ROE && flux != FLUX_VANLEER && flux != FLUX_AUSM && flux != FLUX_AUSMPLUS)
    return fail(FVM_EINVAL, "unknown flux code %d", flux);
  if (rk != RK_SSP3 && rk != RK4) return fail(FVM_EINVAL, "unknown time scheme code %d", rk);
  if (!(gamma > 1.0)) return fail(FVM_EINVAL, "gamma must exceed 1 (got %g)", gamma);
  return FVM_OK;
}

int block_create_impl(fvm_ctx* ctx, int64_t ni, int64_t nj, int64_t nk, double dx, double dy,
                      double dz, hipStream_t alloc_stream, fvm_block** out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!out) return fail(FVM_EINVAL, "null output pointer");
  if (int rc = check_block_dims(ni, nj, nk, dx, dy, dz)) return rc;
  const int dim = nk == 1 ? 2 : 3;
  if (!ctx->blocks.empty() && ctx->blocks[0]->dim != dim)
    return fail(FVM_EINVAL, "all blocks of a context must have the same dimension");
  if (int rc = set_device(ctx)) return rc;
  fvm_block* b = new fvm_block();
  b->ctx = ctx;
  b->dim = dim;
  b->nc = dim == 3 ? 5 : 4;
  BlockGeom& g = b->g;
  g.nx = (int)ni; g.ny = (int)nj; g.nz = (int)nk;
  g.ox = kOx; g.oy = kGhost; g.oz = dim == 3 ? kGhost : 0;
  g.px = row_pitch(ni);
  g.pxy = g.px * (nj + 2 * kGhost);
  const int64_t planes = dim == 3 ? nk + 2 * kGhost : 1;
  g.cs = (g.pxy * planes + 63) / 64 * 64;   // 512-byte aligned components
  g.dx = dx; g.dy = dy; g.dz = dim == 3 ? dz : 1.0;
  g.rdx = 1.0 / g.dx; g.rdy = 1.0 / g.dy; g.rdz = 1.0 / g.dz;
  for (int f = 0; f < 6; ++f) g.bc[f] = BC_EXTRAP;
  g.ng = kGhost;
  b->state_elems = (size_t)b->nc * g.cs;
  const size_t bytes = b->state_elems * sizeof(double);
  if (alloc_stream) {
    void* base = nullptr;
    if (hipMallocAsync(&base, 4 * bytes, alloc_stream) != hipSuccess) {
      delete b;
      return fail(FVM_ENOMEM, "cannot allocate %zu bytes of block state", 4 * bytes);
    }
    b->pooled = true;
    for (int k = 0; k < 4; ++k) b->q[k] = (double*)base + k * b->state_elems;
    HIP_TRY(hipMemsetAsync(base, 0, 4 * bytes, alloc_stream));
  } else {
    for (int k = 0; k < 4; ++k) {
      if (hipMalloc(&b->q[k], bytes) != hipSuccess) {
        for (double* p : b->q) if (p) (void)hipFree(p);
        delete b;
        return fail(FVM_ENOMEM, "cannot allocate %zu bytes of block state", bytes);
      }
      HIP_TRY(hipMemsetAsync(b->q[k], 0, bytes, ctx->stream));
    }
  }
  if (ctx->visc_mu_ref > 0.0)
    if (int rc = alloc_rv(b)) { free_block_state(b, nullptr); delete b; return rc; }
  plan_launch(b);
  ctx->blocks.push_back(b);
  ctx->desc_dirty = true;
  ctx->ghosts_dirty = true;
  *out = b;
  return FVM_OK;
}

// Host layout (5 components, host_ng ghost layers) -> the block's device
// layout in `stage` (state_elems doubles).  A 2-D block reads one k-plane:
// host_nk == 0, the host array has that one plane and no k ghosts; host_nk
// >= 1, it holds host_nk real planes between host_ng ghost planes on each
// side (the reference's own allocation) and real plane k_plane (1-based) is
// read.
int pack_state(const fvm_block* b, const double* q, int host_ng, std::vector<double>& stage,
               int64_t host_nk = 0, int64_t k_plane = 1) {
  const BlockGeom& g = b->g;
  const int64_t hx = g.nx + 2 * host_ng, hy = g.ny + 2 * host_ng;
  const int64_t hz = b->dim == 3 ? g.nz + 2 * host_ng : (host_nk > 0 ? host_nk + 2 * host_ng : 1);
  const int64_t hzoff = b->dim == 3 ? host_ng : (host_nk > 0 ? host_ng + k_plane - 1 : 0);
  const int64_t hcs = hx * hy * hz;
  stage.assign((size_t)b->state_elems, 0.0);
  for (int c = 0; c < 5; ++c) {
    int dc = c;
    if (b->dim == 2) {
      if (c == 3) {
        for (int64_t j = 0; j < g.ny; ++j)
          for (int64_t i = 0; i < g.nx; ++i) {
            const double v = q[c * hcs + hzoff * hx * hy + (j + host_ng) * hx + i + host_ng];
            if (v != 0.0) return fail(FVM_EINVAL, "2-D block: rho*w must be zero (vertex %lld,%lld)", (long long)i, (long long)j);
          }
        continue;
      }
      if (c == 4) dc = 3;
    }
    for (int64_t k = 0; k < g.nz; ++k)
      for (int64_t j = 0; j < g.ny; ++j) {
        const double* src = q + c * hcs + (k + hzoff) * hx * hy + (j + host_ng) * hx + host_ng;
        double* dst = stage.data() + dc * g.cs + (k + g.oz) * g.pxy + (j + g.oy) * g.px + g.ox;
        std::memcpy(dst, src, g.nx * sizeof(double));
      }
  }
  return FVM_OK;
}

// The block's device layout in `stage` -> host layout (5 components, host_ng
// ghost layers; face ghosts only, edges and corners zero).  A 2-D block with
// host_nk >= 1 (pack_state) fills every real k-plane with its state and each
// ghost k-plane with the real vertices of its neighbouring real plane, as
// the reference's ghost update does along k (src/base/defs_block_data.f90:
// 226-256; the i/j ghosts of a ghost plane are edges, never written there).
void unpack_state(const fvm_block* b, const double* stage, double* q, int host_ng,
                  int64_t host_nk = 0) {
  const BlockGeom& g = b->g;
  const int gl = std::min(host_ng, g.ng);
  const int64_t hx = g.nx + 2 * host_ng, hy = g.ny + 2 * host_ng;
  const bool planes = b->dim == 2 && host_nk > 0;
  const int64_t hz = b->dim == 3 ? g.nz + 2 * host_ng : (planes ? host_nk + 2 * host_ng : 1);
  const int64_t hzoff = b->dim == 3 || planes ? host_ng : 0;
  const int64_t hcs = hx * hy * hz;
  const int gzl = b->dim == 3 ? gl : 0;
  std::memset(q, 0, sizeof(double) * 5 * hcs);
  for (int c = 0; c < 5; ++c) {
    if (b->dim == 2 && c == 3) continue;   // rho*w == 0
    const int dc = (b->dim == 2 && c == 4) ? 3 : c;
    for (int64_t k = -gzl; k < g.nz + gzl; ++k)
      for (int64_t j = -gl; j < g.ny + gl; ++j)
        for (int64_t i = -gl; i < g.nx + gl; ++i) {
          // faces only: at most one index outside the real range
          const int outside = (i < 0 || i >= g.nx) + (j < 0 || j >= g.ny) + (k < 0 || k >= g.nz);
          if (outside > 1) continue;
          q[c * hcs + (k + hzoff) * hx * hy + (j + host_ng) * hx + (i + host_ng)] =
              stage[dc * g.cs + (k + g.oz) * g.pxy + (j + g.oy) * g.px + (i + g.ox)];
        }
  }
  if (!planes) return;
  // the first real plane (written above) into the other real planes, its
  // real vertices into the ghost planes
  const int64_t pl = hx * hy, src = host_ng;
  for (int c = 0; c < 5; ++c) {
    double* qc = q + c * hcs;
    for (int64_t kk = 0; kk < hz; ++kk) {
      if (kk == src) continue;
      if (kk >= host_ng && kk < host_ng + host_nk) {
        std::memcpy(qc + kk * pl, qc + src * pl, pl * sizeof(double));
      } else {
        for (int64_t j = 0; j < g.ny; ++j)
          std::memcpy(qc + kk * pl + (j + host_ng) * hx + host_ng,
                      qc + src * pl + (j + host_ng) * hx + host_ng, g.nx * sizeof(double));
      }
    }
  }
}

// The wall kind of a boundary kind whose face mirrors the block onto itself.
int wall_kind(int bc) {
  switch (bc) {
    case FVM_BC_WALL: case FVM_BC_WALL_INVISCID: case FVM_BC_SYMMETRY: return WALL_SLIP;
    case FVM_BC_WALL_VISCOUS: return WALL_NOSLIP;
  }
  return WALL_NONE;
}

// Descriptor entries of the context's blocks into h[0 .. blocks), tile bases
// and the tile total (blocks in context order, tiles numbered block after
// block).
void fill_desc(fvm_ctx* ctx, BlockDesc* h, const double* dt_ptr = nullptr) {
  if (!dt_ptr) dt_ptr = ctx->d_dt;
  int base = 0;
  ctx->has_conn = false;
  ctx->has_wall = false;
  ctx->has_noslip = false;
  for (size_t i = 0; i < ctx->blocks.size(); ++i) {
    fvm_block* b = ctx->blocks[i];
    BlockDesc& d = h[i];
    d = BlockDesc{};
    for (int f = 0; f < 6; ++f) ctx->has_conn = ctx->has_conn || b->g.bc[f] == BC_CONN;
    for (int k = 0; k < 4; ++k) d.q[k] = b->q[k];
    d.g = b->g;
    for (int f = 0; f < 6; ++f) { d.nbr[f] = b->nbr[f]; d.nbr_shared[f] = b->nbr_shared[f]; }
    d.tiles_x = b->tiles_x; d.tiles_y = b->tiles_y; d.tiles_z = b->tiles_z; d.chunk = b->chunk;
    d.tile_base = base;
    d.dt_dev = b->dt_dev ? b->dt_dev : dt_ptr;
    d.smax = b->smax_dev;
    d.halt = b->halt_dev;
    d.rv = ctx->visc_mu_ref > 0.0 ? b->rv : nullptr;
    for (const fvm_ctx::Patch& q : ctx->patches) {
      if (q.recv != b || q.donor != b) continue;
      const int w = wall_kind(q.bc);
      if (w == WALL_NONE) continue;
      d.wall[q.face] = w;
      if (w == WALL_NOSLIP) d.noslip |= 1 << q.face;
      ctx->has_wall = true;
      ctx->has_noslip = ctx->has_noslip || w == WALL_NOSLIP;
    }
    b->tile_base = base;
    base += b->ntiles;
  }
  ctx->total_tiles = base;
  ++ctx->gen;
}

// Device table of the connection patches (block indices as in the
// descriptor table); checks that every BC_PATCH face is covered exactly once.
int sync_patches(fvm_ctx* ctx) {
  ctx->patch_points = 0;
  if (ctx->patches.empty()) return FVM_OK;
  std::vector<PatchDesc> h;
  int64_t start = 0;
  for (const fvm_ctx::Patch& q : ctx->patches) {
    PatchDesc d{};
    d.recv = block_index(ctx, q.recv);
    d.donor = block_index(ctx, q.donor);
    d.dn = q.face / 2;
    d.side = q.face % 2;
    d.ta = d.dn == 0 ? 1 : 0;
    d.tb = d.dn == 2 ? 1 : 2;
    d.na = (int)(q.rmax[d.ta] - q.rmin[d.ta] + 1);
    d.nb = (int)(q.rmax[d.tb] - q.rmin[d.tb] + 1);
    d.nc = q.recv->nc;
    for (int e = 0; e < 3; ++e) {
      d.map_dir[e] = q.map_dir[e];
      d.map_sign[e] = q.map_sign[e];
      d.rmin[e] = q.rmin[e];
      d.dmin[e] = q.dmin[e];
    }
    d.start = start;
    d.fixed_state = q.fixed ? 1 : 0;
    d.neg_all = wall_kind(q.bc) == WALL_NOSLIP ? 1 : 0;
    if (q.fixed) {
      // p -> q in the order of src/solve/class_flow_field.f90:353-377
      const double rho = q.prim[0], u = q.prim[1], v = q.prim[2], w = q.prim[3], pr = q.prim[4];
      const double e = pr / (ctx->P.gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
      d.fixed[0] = rho;
      d.fixed[1] = rho * u;
      d.fixed[2] = rho * v;
      if (q.recv->nc == 5) {
        d.fixed[3] = rho * w;
        d.fixed[4] = e;
      } else {
        d.fixed[3] = e;
      }
    }
    start += (int64_t)d.na * d.nb * q.recv->g.ng;
    h.push_back(d);
  }
  // coverage: the patches of a face tile it without overlap
  for (fvm_block* b : ctx->blocks)
    for (int f = 0; f < 2 * b->dim; ++f) {
      if (b->g.bc[f] != BC_PATCH) continue;
      const int dn = f / 2, ta = dn == 0 ? 1 : 0, tb = dn == 2 ? 1 : 2;
      const int n[3] = {b->g.nx, b->g.ny, b->g.nz};
      int64_t area = 0;
      std::vector<const fvm_ctx::Patch*> on;
      for (const fvm_ctx::Patch& q : ctx->patches)
        if (q.recv == b && q.face == f) {
          for (const fvm_ctx::Patch* o : on) {
            bool apart = false;
            for (int e : {ta, tb}) apart = apart || q.rmax[e] < o->rmin[e] || o->rmax[e] < q.rmin[e];
            if (!apart) return fail(FVM_EINVAL, "block %d face %d: connection patches overlap",
                                    block_index(ctx, b), f);
          }
          on.push_back(&q);
          area += (q.rmax[ta] - q.rmin[ta] + 1) * (q.rmax[tb] - q.rmin[tb] + 1);
        }
      if (area != (int64_t)n[ta] * n[tb])
        return fail(FVM_EINVAL, "block %d face %d: connection patches cover %lld of %lld vertices",
                    block_index(ctx, b), f, (long long)area, (long long)n[ta] * n[tb]);
    }
  if (h.size() > ctx->patch_cap) {
    if (ctx->d_patch) HIP_TRY(hipFree(ctx->d_patch));
    ctx->d_patch = nullptr;
    if (hipMalloc(&ctx->d_patch, h.size() * sizeof(PatchDesc)) != hipSuccess)
      return fail(FVM_ENOMEM, "cannot allocate the patch table");
    ctx->patch_cap = h.size();
  }
  HIP_TRY(hipMemcpyAsync(ctx->d_patch, h.data(), h.size() * sizeof(PatchDesc), hipMemcpyHostToDevice,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->patch_points = start;
  return FVM_OK;
}

// Fill the connection-patch ghosts of state array `buf` (after a stage, or
// after the face fills of an upload).
int launch_patches(fvm_ctx* ctx, int buf) {
  if (ctx->patch_points <= 0) return FVM_OK;
  const int64_t blocks = std::min<int64_t>(1024, (ctx->patch_points + 255) / 256);
  hipLaunchKernelGGL(patch_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, ctx->d_desc,
                     ctx->d_patch, (int)ctx->patches.size(), ctx->patch_points, buf);
  HIP_TRY(hipGetLastError());
  return FVM_OK;
}

int sync_desc(fvm_ctx* ctx) {
  if (!ctx->desc_dirty) return FVM_OK;
  if (int rc = sync_patches(ctx)) return rc;
  std::vector<BlockDesc> h(ctx->blocks.size());
  fill_desc(ctx, h.data());
  if (h.size() > ctx->desc_cap) {
    if (ctx->d_desc) HIP_TRY(hipFree(ctx->d_desc));
    ctx->d_desc = nullptr;
    const size_t cap = std::max<size_t>(h.size(), 16);
    if (hipMalloc(&ctx->d_desc, cap * sizeof(BlockDesc)) != hipSuccess)
      return fail(FVM_ENOMEM, "cannot allocate the block table");
    ctx->desc_cap = cap;
  }
  if (!h.empty())
    HIP_TRY(hipMemcpyAsync(ctx->d_desc, h.data(), h.size() * sizeof(BlockDesc), hipMemcpyHostToDevice,
                           ctx->stream));
  // the staging vector is released at return: wait for the copy
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->desc_dirty = false;
  return FVM_OK;
}

int ensure_norm_capacity(fvm_ctx* ctx, size_t doubles) {
  if (doubles <= ctx->norm_cap) return FVM_OK;
  if (ctx->d_norm) HIP_TRY(hipFree(ctx->d_norm));
  ctx->d_norm = nullptr;
  if (hipMalloc(&ctx->d_norm, doubles * sizeof(double)) != hipSuccess)
    return fail(FVM_ENOMEM, "cannot allocate %zu norm partials", doubles);
  ctx->norm_cap = doubles;
  ++ctx->gen;
  return FVM_OK;
}

// Refill the face ghost layers of state array `buf` of every block, reading
// the block table `tab` (the context's, or the no-allocation scratch's).
int fill_ghosts_tab(fvm_ctx* ctx, const BlockDesc* tab, int buf) {
  for (size_t i = 0; i < ctx->blocks.size(); ++i) {
    fvm_block* b = ctx->blocks[i];
    for (int d = 0; d < b->dim; ++d)
      for (int side = 0; side < 2; ++side) {
        hipLaunchKernelGGL(ghost_fill_kernel, dim3(256), dim3(256), 0, ctx->stream, tab, (int)i,
                           buf, b->nc, d, side);
        HIP_TRY(hipGetLastError());
      }
  }
  return FVM_OK;
}

// Refill the ghost layers of state array `buf` of every block.
int fill_ghosts_all(fvm_ctx* ctx, int buf) {
  if (int rc = sync_desc(ctx)) return rc;
  if (int rc = fill_ghosts_tab(ctx, ctx->d_desc, buf)) return rc;
  return launch_patches(ctx, buf);
}

// The stage kernels' CONN variant writes connected faces' ghosts and applies
// the no-slip wall condition; the other variant carries neither.
bool conn_variant(const fvm_ctx* ctx) { return ctx->has_conn || ctx->has_noslip; }

template <class Fam>
hipError_t launch_family(fvm_ctx* ctx, int dim, int tile, const StageArgs& a, int grid) {
  if (dim == 2) {
    switch (ctx->recon) {
      case RECON_SIMPLE: return Fam::s2(ctx->flux, tile, conn_variant(ctx), a, grid, ctx->stream);
      case RECON_MUSCL: return Fam::m2(ctx->flux, tile, conn_variant(ctx), a, grid, ctx->stream);
      case RECON_WENO: return Fam::w2(ctx->flux, tile, conn_variant(ctx), a, grid, ctx->stream);
    }
  } else {
    switch (ctx->recon) {
      case RECON_SIMPLE: return Fam::s3(ctx->flux, tile, conn_variant(ctx), a, grid, ctx->stream);
      case RECON_MUSCL: return Fam::m3(ctx->flux, tile, conn_variant(ctx), a, grid, ctx->stream);
      case RECON_WENO: return Fam::w3(ctx->flux, tile, conn_variant(ctx), a, grid, ctx->stream);
    }
  }
  return hipErrorInvalidValue;
}

struct ExactFam {
  static constexpr auto s2 = fvm::exact::launch_stage2d_simple;
  static constexpr auto m2 = fvm::exact::launch_stage2d_muscl;
  static constexpr auto w2 = fvm::exact::launch_stage2d_weno;
  static constexpr auto s3 = fvm::exact::launch_stage3d_simple;
  static constexpr auto m3 = fvm::exact::launch_stage3d_muscl;
  static constexpr auto w3 = fvm::exact::launch_stage3d_weno;
};
struct FastFam {
  static constexpr auto s2 = fvm::fast::launch_stage2d_simple;
  static constexpr auto m2 = fvm::fast::launch_stage2d_muscl;
  static constexpr auto w2 = fvm::fast::launch_stage2d_weno;
  static constexpr auto s3 = fvm::fast::launch_stage3d_simple;
  static constexpr auto m3 = fvm::fast::launch_stage3d_muscl;
  static constexpr auto w3 = fvm::fast::launch_stage3d_weno;
};

int launch_one(fvm_ctx* ctx, int dim, int tile, const StageArgs& a, int grid) {
  if (grid <= 0 || a.ntiles <= 0) return FVM_OK;
  const hipError_t e = ctx->fast ? launch_family<FastFam>(ctx, dim, tile, a, grid)
                                 : launch_family<ExactFam>(ctx, dim, tile, a, grid);
  if (e != hipSuccess) return fail(FVM_EHIP, "stage launch: %s", hipGetErrorString(e));
  return FVM_OK;
}

// Can all blocks go in one launch (same kernel instantiation)?
bool batchable(fvm_ctx* ctx) {
  for (fvm_block* b : ctx->blocks)
    if (b->dim != ctx->blocks[0]->dim || b->tile != ctx->blocks[0]->tile) return false;
  return true;
}

// Grid of a launch over `ntiles` tiles: one workgroup per tile, or in
// persistent mode at most two per CU (a multiple of 8, so each workgroup
// stays on one XCD for every tile it loops over).
int grid_for(fvm_ctx* ctx, int ntiles) {
  if (ctx->launch_mode != FVM_LAUNCH_PERSISTENT) return ntiles;
  return std::min(ntiles, 2 * ctx->num_cu);
}

// The viscous residual of every block's stage input (one launch per block);
// with connected faces the kernels read across them through the table.
int launch_viscous_all(fvm_ctx* ctx, int in_buf, const BlockDesc* tab) {
  ViscArgs v{};
  v.V.mu_ref = ctx->visc_mu_ref;
  v.V.t_ref = ctx->visc_t_ref;
  v.V.s = ctx->visc_s;
  v.V.r_gas = ctx->visc_r;
  v.V.gamma = ctx->P.gamma;
  v.V.cpr = ctx->P.gamma * ctx->visc_r / ((ctx->P.gamma - 1.0) * ctx->visc_pr);
  v.V.rr_gas = 1.0 / ctx->visc_r;
  v.V.rt_ref = 1.0 / ctx->visc_t_ref;
  v.V.sfac = ctx->visc_t_ref + ctx->visc_s;
  v.tab = ctx->has_conn || ctx->has_wall || !ctx->patches.empty() ? tab : nullptr;
  v.in_buf = in_buf;
  v.patches = ctx->d_patch;
  v.npatch = (int)ctx->patches.size();
  for (size_t bi = 0; bi < ctx->blocks.size(); ++bi) {
    fvm_block* b = ctx->blocks[bi];
    v.self = (int)bi;
    v.g = b->g;
    v.q = b->q[in_buf];
    v.rv = b->rv;
    v.tiles_x = (b->g.nx + kViscTile[0] - 1) / kViscTile[0];
    v.tiles_y = (b->g.ny + kViscTile[1] - 1) / kViscTile[1];
    v.chunk = 16;
    const int tz = b->dim == 3 ? (b->g.nz + v.chunk - 1) / v.chunk : 1;
    const int grid = v.tiles_x * v.tiles_y * tz;
    const hipError_t e = ctx->fast ? fvm::fast::launch_viscous(b->dim, v, grid, ctx->stream)
                                   : fvm::exact::launch_viscous(b->dim, v, grid, ctx->stream);
    if (e != hipSuccess) return fail(FVM_EHIP, "viscous launch: %s", hipGetErrorString(e));
  }
  return FVM_OK;
}

// One RK stage over every block: a single launch over the whole tile list
// (batched or persistent), or one launch per block.
int launch_stage_all(fvm_ctx* ctx, StageArgs a, const BlockDesc* tab = nullptr) {
  a.blocks = tab ? tab : ctx->d_desc;
  if (ctx->visc_mu_ref > 0.0)
    if (int rc = launch_viscous_all(ctx, a.in_buf, a.blocks)) return rc;
  a.nblocks = (int)ctx->blocks.size();
  if (ctx->launch_mode != FVM_LAUNCH_PER_BLOCK && batchable(ctx)) {
    a.tile0 = 0;
    a.ntiles = ctx->total_tiles;
    return launch_one(ctx, ctx->blocks[0]->dim, ctx->blocks[0]->tile, a, grid_for(ctx, a.ntiles));
  }
  for (fvm_block* b : ctx->blocks) {
    a.tile0 = b->tile_base;
    a.ntiles = b->ntiles;
    if (int rc = launch_one(ctx, b->dim, b->tile, a, grid_for(ctx, b->ntiles))) return rc;
  }
  return FVM_OK;
}

// Device memory a time step works in: the block table, the per-tile norm
// partials and the time slot -- the context's own, or the static scratch of
// the no-allocation entry point.
struct StepMem {
  const BlockDesc* desc;
  double* norm;
  SolveSlot* slot;
  SmaxWays* smax;
};

StepMem ctx_mem(fvm_ctx* ctx) { return StepMem{ctx->d_desc, ctx->d_norm, ctx->d_slot, ctx->d_smax}; }

// Stage st (1-based) of one time step: its buffers, combine, norm and CFL
// roles, then the launch over every block and the patch refill.
// `norm_mode`: 0 no residual norms, 1 the last stage leaves its norm
// partials, 2 only if the step reached t_end.
int launch_one_stage(fvm_ctx* ctx, int st, bool cfl_mode, int norm_mode, const StepMem& mem) {
  const int nstage = ctx->rk == RK4 ? 4 : 3;
  {
    const bool last = st == nstage;
    StageArgs a{};
    a.P = ctx->P;
    a.q0_buf = 0;
    a.use_acc = 0;
    if (ctx->rk == RK4) {
      a.use_acc = 1;
      switch (st) {
        case 1: a.in_buf = 0; a.out_buf = 1; a.stage = ST_RK4_1; break;
        case 2: a.in_buf = 1; a.out_buf = 2; a.stage = ST_RK4_2; break;
        case 3: a.in_buf = 2; a.out_buf = 1; a.stage = ST_RK4_3; break;
        default: a.in_buf = 1; a.out_buf = 0; a.stage = ST_RK4_4; break;
      }
    } else {
      switch (st) {
        case 1: a.in_buf = 0; a.out_buf = 1; a.stage = ST_SSP1; break;
        case 2: a.in_buf = 1; a.out_buf = 2; a.stage = ST_SSP2; break;
        default: a.in_buf = 2; a.out_buf = 0; a.stage = ST_SSP3; break;
      }
    }
    a.norm_part = (last && norm_mode) ? mem.norm : nullptr;
    a.norm_if_end = last && norm_mode == 2 ? 1 : 0;
    a.smax = (last && cfl_mode) ? mem.smax : nullptr;
    a.halt = cfl_mode ? &mem.slot->halt : nullptr;
    a.slot = cfl_mode ? mem.slot : nullptr;
    a.slot_smax = cfl_mode ? mem.smax : nullptr;
    a.cfl_dt = cfl_mode && st == 1 ? 1 : 0;
    a.commit = cfl_mode && st == 2 ? 1 : 0;
    if (int rc = launch_stage_all(ctx, a, mem.desc)) return rc;
    if (int rc = launch_patches(ctx, a.out_buf)) return rc;
  }
  return FVM_OK;
}

// Per-block sums of the per-tile norm partials in ctx->h_norm, in tile order
// (block_sums: nc per block, context order).
void sum_norm_partials(fvm_ctx* ctx, int nc, double* block_sums) {
  for (size_t bi = 0; bi < ctx->blocks.size(); ++bi) {
    fvm_block* b = ctx->blocks[bi];
    double* sums = block_sums + bi * nc;
    for (int c = 0; c < nc; ++c) sums[c] = 0.0;
    for (int w = 0; w < b->ntiles; ++w)
      for (int c = 0; c < nc; ++c) sums[c] += ctx->h_norm[(size_t)(b->tile_base + w) * nc + c];
  }
}

// Residual norms of a step's last stage (class_iteration.f90:210-248) from
// the per-block sums of rhs^2; blocks combined as the reference accumulates
// them, R = sqrt(R_prev^2 + R_blk^2).  Fills ctx->norms; returns the NaN /
// divergence flag.
int combine_norms(fvm_ctx* ctx, int nc, const double* block_sums) {
  double res_abs[6] = {0, 0, 0, 0, 0, 0};
  for (size_t bi = 0; bi < ctx->blocks.size(); ++bi) {
    fvm_block* b = ctx->blocks[bi];
    const double* sums = block_sums + bi * nc;
    double res2[5] = {0, 0, 0, 0, 0};
    const double npts = (double)b->g.nx * b->g.ny * b->g.nz;
    for (int c = 0; c < nc; ++c) {
      const int e = (nc == 4 && c == 3) ? 4 : c;   // 2-D: energy is equation 5
      res2[e] = sums[c] / npts;
    }
    double res2t = 0.0;
    for (int e = 0; e < 5; ++e) res2t += res2[e];
    res2t = res2t / 5.0;
    for (int e = 0; e < 5; ++e) res_abs[e] = std::sqrt(res_abs[e] * res_abs[e] + res2[e]);
    res_abs[5] = std::sqrt(res_abs[5] * res_abs[5] + res2t);
  }
  int nan = 0;
  for (int e = 0; e < 6; ++e) {
    ctx->norms[e] = res_abs[e];
    if (std::isnan(res_abs[e])) nan = 1;
  }
  if (res_abs[5] > 1.0e10) nan = 1;
  return nan;
}

// One time step of every block: the RK stages (in CFL mode the first one
// computes dt from the device time slot and the second commits the step).
// With timing on (nev non-null), one event pair brackets the step's stage
// launches (ev index *nev): events between launches would stall the stream
// for microseconds each, a large share of a short stage.
int launch_step(fvm_ctx* ctx, bool cfl_mode, int norm_mode, size_t* nev, const StepMem& mem) {
  const int nstage = ctx->rk == RK4 ? 4 : 3;
  if (nev) HIP_TRY(hipEventRecord(ctx->ev[(*nev)++], ctx->stream));
  for (int st = 1; st <= nstage; ++st)
    if (int rc = launch_one_stage(ctx, st, cfl_mode, norm_mode, mem)) return rc;
  if (nev) HIP_TRY(hipEventRecord(ctx->ev[(*nev)++], ctx->stream));
  return FVM_OK;
}

// The instantiated graph of one step, captured from launch_step on the
// context's stream the first time it is needed and again whenever the launch
// configuration changed.  No memset or copy is captured: the step is kernel
// launches only, reading dt, the CFL maximum and the halt flag from device
// memory.
int step_graph(fvm_ctx* ctx, bool cfl_mode, int norm_mode, hipGraphExec_t* out) {
  fvm_ctx::StepGraph& sg = ctx->graphs[norm_mode];
  if (sg.exec && sg.gen == ctx->gen && sg.cfl_mode == cfl_mode) {
    *out = sg.exec;
    return FVM_OK;
  }
  if (sg.exec) {
    HIP_TRY(hipGraphExecDestroy(sg.exec));
    sg.exec = nullptr;
  }
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = launch_step(ctx, cfl_mode, norm_mode, nullptr, ctx_mem(ctx));
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess) return fail(FVM_EHIP, "step capture: %s", hipGetErrorString(ec));
  const hipError_t ei = hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    sg.exec = nullptr;
    return fail(FVM_EHIP, "step graph instantiate: %s", hipGetErrorString(ei));
  }
  sg.gen = ctx->gen;
  sg.cfl_mode = cfl_mode;
  *out = sg.exec;
  return FVM_OK;
}

// ---- memory of the no-allocation step (fvm_step_noalloc) -------------------
// One static device array per GPU holds what fvm_step keeps in context
// allocations: the time slot, the block table, per-block norm sums and the
// per-tile norm partials.  Host code never uses the array's name as an
// address: fvm_ctx_create resolves its address on each device once
// (hipGetSymbolAddress) and the entry point passes that pointer on.
constexpr int kScratchBlocks = 64;          // blocks of a context
constexpr int kScratchTiles = 65536;        // stage tiles over all its blocks
constexpr size_t scratch_round(size_t n) { return (n + 255) / 256 * 256; }
constexpr size_t kScratchSmaxOff = scratch_round(sizeof(SolveSlot));
constexpr size_t kScratchDescOff = kScratchSmaxOff + scratch_round(sizeof(SmaxWays));
constexpr size_t kScratchSumOff = kScratchDescOff + scratch_round(kScratchBlocks * sizeof(BlockDesc));
constexpr size_t kScratchNormOff = kScratchSumOff + scratch_round(kScratchBlocks * 5 * sizeof(double));
constexpr size_t kScratchBytes = kScratchNormOff + (size_t)kScratchTiles * 5 * sizeof(double);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kScratchDevices] = {};

// With `device` current: the device address of its copy of `scratch`.
int scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kScratchDevices || g_scratch_ptr[device]) return FVM_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(FVM_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch_ptr[device] = (unsigned char*)p;
  return FVM_OK;
}

unsigned char* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kScratchDevices ? g_scratch_ptr[device] : nullptr;
}

// Per-block sums of the per-tile norm partials, tiles in order, one thread
// per (block, component): the additions sum_norm_partials makes on the host.
__global__ void norm_block_sum_kernel(const BlockDesc* __restrict__ tab, int nblocks, int total_tiles,
                                      int nc, const double* __restrict__ part, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nblocks * nc) return;
  const int b = e / nc, c = e % nc;
  const int t0 = tab[b].tile_base, t1 = b + 1 < nblocks ? tab[b + 1].tile_base : total_tiles;
  double s = 0.0;
  for (int w = t0; w < t1; ++w) s += part[(int64_t)w * nc + c];
  out[e] = s;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char* fvm_last_error(void) { return g_err.c_str(); }

const char* fvm_version(void) { return "fvm-mi355x 0.2.0 (gfx950, fp64)"; }

int fvm_ctx_create(int device, fvm_ctx** out) {
  if (!out) return fail(FVM_EINVAL, "null output pointer");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(FVM_EINVAL, "device %d out of range (%d devices)", device, ndev);
  fvm_ctx* ctx = new fvm_ctx();
  ctx->device = device;
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->num_cu = prop.multiProcessorCount;
  if (int rc = scratch_setup(device)) {
    delete ctx;
    return rc;
  }
  HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIP_TRY(hipMalloc(&ctx->d_slot, sizeof(SolveSlot)));
  HIP_TRY(hipMemset(ctx->d_slot, 0, sizeof(SolveSlot)));
  HIP_TRY(hipMalloc(&ctx->d_smax, sizeof(SmaxWays)));
  HIP_TRY(hipMemset(ctx->d_smax, 0, sizeof(SmaxWays)));
  ctx->d_dt = &ctx->d_slot->dt;
  *out = ctx;
  g_err.clear();
  return FVM_OK;
}

int fvm_ctx_destroy(fvm_ctx* ctx) {
  if (!ctx) return FVM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (fvm_block* b : ctx->blocks) {
    free_block_state(b, nullptr);
    delete b;
  }
  for (auto& sn : ctx->snaps) {
    if (sn.writer.joinable()) sn.writer.join();
    if (sn.done) (void)hipEventSynchronize(sn.done);
    (void)hipFree(sn.d_buf);
    (void)hipHostFree(sn.h_buf);
    if (sn.done) (void)hipEventDestroy(sn.done);
    if (sn.conv) (void)hipEventDestroy(sn.conv);
  }
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
  for (auto& sg : ctx->graphs)
    if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
  (void)hipFree(ctx->d_desc);
  (void)hipFree(ctx->d_slot);
  (void)hipFree(ctx->d_smax);
  (void)hipFree(ctx->d_norm);
  (void)hipFree(ctx->d_patch);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FVM_OK;
}

int fvm_scheme_set(fvm_ctx* ctx, int recon, int limiter, int flux, int rk, double kappa,
                   double entropy_fix, double gamma, int ref_compat) {
  if (int rc = check_ctx(ctx)) return rc;
  if (int rc = check_scheme(recon, limiter, flux, rk, gamma)) return rc;
  ctx->recon = recon;
  ctx->flux = flux;
  ctx->rk = rk;
  ctx->P.gamma = gamma;
  ctx->P.kappa = kappa;
  ctx->P.entfix = entropy_fix;
  ctx->P.g_gm1 = gamma / (gamma - 1.0);
  ctx->P.limiter = limiter;
  ctx->P.weno_textbook = ref_compat ? 0 : 1;
  ctx->scheme_set = true;
  // a launch shape the new reconstruction cannot use falls back to the default
  for (fvm_block* b : ctx->blocks)
    if (b->dim == 3 && !tile3d_ok(recon, b->tile)) plan_launch(b);
  ++ctx->gen;
  return FVM_OK;
}

int fvm_set_numerics(fvm_ctx* ctx, int mode) {
  if (int rc = check_ctx(ctx)) return rc;
  if (mode != FVM_NUMERICS_EXACT && mode != FVM_NUMERICS_FAST)
    return fail(FVM_EINVAL, "unknown numerics mode %d", mode);
  ctx->fast = mode == FVM_NUMERICS_FAST;
  ++ctx->gen;
  return FVM_OK;
}

int fvm_set_launch_mode(fvm_ctx* ctx, int mode) {
  if (int rc = check_ctx(ctx)) return rc;
  if (mode != FVM_LAUNCH_PER_BLOCK && mode != FVM_LAUNCH_BATCHED && mode != FVM_LAUNCH_PERSISTENT)
    return fail(FVM_EINVAL, "unknown launch mode %d", mode);
  ctx->launch_mode = mode;
  ++ctx->gen;
  return FVM_OK;
}

int fvm_block_create(fvm_ctx* ctx, int64_t ni, int64_t nj, int64_t nk, double dx, double dy,
                     double dz, fvm_block** out) {
  return block_create_impl(ctx, ni, nj, nk, dx, dy, dz, nullptr, out);
}

int fvm_block_destroy(fvm_block* b) {
  if (!b) return FVM_OK;
  fvm_ctx* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  const int idx = block_index(ctx, b);
  // drop connections to this block and renumber the rest
  for (fvm_block* o : ctx->blocks)
    for (int f = 0; f < 6; ++f) {
      if (o->nbr[f] == idx) { o->nbr[f] = -1; o->g.bc[f] = BC_EXTRAP; }
      else if (o->nbr[f] > idx) --o->nbr[f];
    }
  // a face with a connection patch from the block loses all its patches and
  // falls back to extrapolation
  for (const fvm_ctx::Patch& q : ctx->patches)
    if (q.donor == b) q.recv->g.bc[q.face] = BC_EXTRAP;
  std::vector<fvm_ctx::Patch> keep;
  for (const fvm_ctx::Patch& q : ctx->patches)
    if (q.recv != b && q.recv->g.bc[q.face] == BC_PATCH) keep.push_back(q);
  ctx->patches.swap(keep);
  free_block_state(b, nullptr);
  ctx->blocks.erase(std::remove(ctx->blocks.begin(), ctx->blocks.end(), b), ctx->blocks.end());
  delete b;
  ctx->desc_dirty = true;
  ctx->ghosts_dirty = true;
  return FVM_OK;
}

}  // extern "C"

namespace {

// Patches on one face of a block: drop them all.
void drop_face_patches(fvm_ctx* ctx, fvm_block* b, int face) {
  std::vector<fvm_ctx::Patch> keep;
  for (const fvm_ctx::Patch& q : ctx->patches)
    if (!(q.recv == b && q.face == face)) keep.push_back(q);
  ctx->patches.swap(keep);
}

// The face's ghosts come from another block (a connection or connection
// patches), not from a boundary kind.
bool face_connected(fvm_ctx* ctx, fvm_block* b, int face) {
  if (b->g.bc[face] == BC_CONN) return true;
  if (b->g.bc[face] != BC_PATCH) return false;
  for (const fvm_ctx::Patch& q : ctx->patches)
    if (q.recv == b && q.face == face && q.bc == 0) return true;
  return false;
}

enum BcClass { BCC_UNKNOWN, BCC_EXTRAP, BCC_PERIODIC, BCC_MIRROR, BCC_STATE, BCC_NOSLIP };

BcClass bc_class(int kind) {
  switch (kind) {
    case FVM_BC_EXTRAP: case FVM_BC_OUTFLOW: case FVM_BC_OUTFLOW_SUBSONIC:
    case FVM_BC_OUTFLOW_SUPERSONIC: case FVM_BC_EXTRAPOLATE: return BCC_EXTRAP;
    case FVM_BC_PERIODIC: return BCC_PERIODIC;
    case FVM_BC_WALL: case FVM_BC_WALL_INVISCID: case FVM_BC_SYMMETRY: return BCC_MIRROR;
    case FVM_BC_INFLOW: case FVM_BC_INFLOW_SUBSONIC: case FVM_BC_INFLOW_SUPERSONIC:
    case FVM_BC_FARFIELD: return BCC_STATE;
    case 32: case 33: return BCC_NOSLIP;   // WallViscousIsothermal, WallViscous
  }
  return BCC_UNKNOWN;
}

int set_face_kind(fvm_block* b, int face, int kind, const double* prim) {
  if (!b) return fail(FVM_EINVAL, "null block");
  if (face < 0 || face > 5) return fail(FVM_EINVAL, "face %d out of range 0..5", face);
  if (b->dim == 2 && face >= 4) return fail(FVM_EINVAL, "2-D block has no k faces");
  const BcClass cls = bc_class(kind);
  if (cls == BCC_UNKNOWN) return fail(FVM_EINVAL, "unknown BC kind %d", kind);
  if (kind == FVM_BC_WALL_VISCOUS_ISOTHERMAL)
    return fail(FVM_EINVAL, "isothermal no-slip walls (BC kind %d) are not supported; kind %d is the "
                "adiabatic no-slip wall", kind, FVM_BC_WALL_VISCOUS);
  if (cls == BCC_STATE && !prim)
    return fail(FVM_EINVAL, "BC kind %d needs a state: fvm_block_set_bc_state", kind);
  if (cls != BCC_STATE && prim) return fail(FVM_EINVAL, "BC kind %d takes no state", kind);
  fvm_ctx* ctx = b->ctx;
  const int d = face / 2, other = 2 * d + (1 - face % 2);
  if (face_connected(ctx, b, face)) return fail(FVM_EINVAL, "face %d is connected to another block", face);
  if (cls == BCC_PERIODIC && face_connected(ctx, b, other))
    return fail(FVM_EINVAL, "face %d is connected to another block", other);
  if (cls == BCC_STATE && ctx->visc_mu_ref > 0.0)
    return fail(FVM_EINVAL, "inflow and farfield faces are not supported with viscous terms");
  const int n[3] = {b->g.nx, b->g.ny, b->g.nz};
  if ((cls == BCC_MIRROR || cls == BCC_NOSLIP) && n[d] < kGhost + 1)
    return fail(FVM_EINVAL, "a mirrored face needs >= %d vertices across the block", kGhost + 1);
  if (cls == BCC_STATE) {
    for (int c = 0; c < 5; ++c)
      if (!std::isfinite(prim[c])) return fail(FVM_EINVAL, "non-finite boundary state");
    if (!(prim[0] > 0.0) || !(prim[4] > 0.0))
      return fail(FVM_EINVAL, "boundary state needs rho > 0 and p > 0");
    if (b->dim == 2 && prim[3] != 0.0) return fail(FVM_EINVAL, "2-D block: w of the boundary state must be 0");
  }
  drop_face_patches(ctx, b, face);
  if (cls == BCC_PERIODIC) drop_face_patches(ctx, b, other);
  if (b->g.bc[other] == BC_PERIODIC && cls != BCC_PERIODIC) b->g.bc[other] = BC_EXTRAP;
  switch (cls) {
    case BCC_EXTRAP:
      b->g.bc[face] = BC_EXTRAP;
      break;
    case BCC_PERIODIC:
      b->g.bc[face] = b->g.bc[other] = BC_PERIODIC;    // periodic on both faces
      break;
    default: {
      // the whole face as one patch onto the block itself: a mirror about the
      // boundary plane (direction d reversed) or a fixed state
      fvm_ctx::Patch q{};
      q.recv = q.donor = b;
      q.face = face;
      q.bc = kind;
      for (int e = 0; e < 3; ++e) {
        q.rmin[e] = 0;
        q.rmax[e] = n[e] - 1;
        q.map_dir[e] = e;
        q.map_sign[e] = 1;
      }
      q.rmin[d] = q.rmax[d] = face % 2 ? n[d] - 1 : 0;
      for (int e = 0; e < 3; ++e) q.dmin[e] = q.rmin[e];
      if (cls == BCC_STATE) {
        q.fixed = true;
        for (int c = 0; c < 5; ++c) q.prim[c] = prim[c];
      } else {
        q.map_sign[d] = -1;
      }
      ctx->patches.push_back(q);
      b->g.bc[face] = BC_PATCH;
    }
  }
  ctx->desc_dirty = true;
  ctx->ghosts_dirty = true;
  return FVM_OK;
}

}  // namespace

extern "C" {

int fvm_block_set_bc(fvm_block* b, int face, int kind) { return set_face_kind(b, face, kind, nullptr); }

int fvm_block_set_bc_state(fvm_block* b, int face, int kind, const double prim[5]) {
  if (!prim) return fail(FVM_EINVAL, "null boundary state");
  return set_face_kind(b, face, kind, prim);
}

int fvm_block_connect(fvm_block* a, int face, fvm_block* b, int shared) {
  if (!a || !b) return fail(FVM_EINVAL, "null block");
  if (a->ctx != b->ctx) return fail(FVM_EINVAL, "blocks belong to different contexts");
  if (face < 0 || face > 5) return fail(FVM_EINVAL, "face %d out of range 0..5", face);
  if (a->dim == 2 && face >= 4) return fail(FVM_EINVAL, "2-D block has no k faces");
  const int d = face / 2, bface = face ^ 1;
  const int na[3] = {a->g.nx, a->g.ny, a->g.nz}, nb[3] = {b->g.nx, b->g.ny, b->g.nz};
  for (int e = 0; e < 3; ++e)
    if (e != d && na[e] != nb[e])
      return fail(FVM_EINVAL, "faces do not conform: extent %d is %d vs %d", e, na[e], nb[e]);
  if (na[d] < 2 * kGhost || nb[d] < 2 * kGhost)
    return fail(FVM_EINVAL, "connected blocks need >= %d vertices across the face", 2 * kGhost);
  if (a == b) {
    if (!shared) return fvm_block_set_bc(a, face, FVM_BC_PERIODIC);
    return fail(FVM_EINVAL, "a block cannot share a vertex plane with itself");
  }
  fvm_ctx* ctx = a->ctx;
  if ((a->g.bc[face] == BC_PATCH && face_connected(ctx, a, face)) ||
      (b->g.bc[bface] == BC_PATCH && face_connected(ctx, b, bface)))
    return fail(FVM_EINVAL, "a face already has connection patches");
  drop_face_patches(ctx, a, face);
  drop_face_patches(ctx, b, bface);
  // a periodic direction cut open by this connection: the opposite face of
  // either block no longer wraps onto this one (as set_face_kind does when a
  // periodic face gets another kind)
  if (a->g.bc[face ^ 1] == BC_PERIODIC) a->g.bc[face ^ 1] = BC_EXTRAP;
  if (b->g.bc[bface ^ 1] == BC_PERIODIC) b->g.bc[bface ^ 1] = BC_EXTRAP;
  const int ia = block_index(ctx, a), ib = block_index(ctx, b);
  a->g.bc[face] = BC_CONN; a->nbr[face] = ib; a->nbr_shared[face] = shared ? 1 : 0;
  b->g.bc[bface] = BC_CONN; b->nbr[bface] = ia; b->nbr_shared[bface] = shared ? 1 : 0;
  ctx->desc_dirty = true;
  ctx->ghosts_dirty = true;
  return FVM_OK;
}

int fvm_cut_uniform(int64_t range_min, int64_t range_max, int ncut, int64_t* pos) {
  if (!pos) return fail(FVM_EINVAL, "null position array");
  if (ncut < 1 || ncut > range_max - range_min)
    return fail(FVM_EINVAL, "ncut %d out of range 1..%lld cells", ncut, (long long)(range_max - range_min));
  // each cut takes the remaining cells divided by the remaining pieces
  pos[0] = range_min;
  for (int c = 1; c < ncut; ++c) pos[c] = pos[c - 1] + (range_max - pos[c - 1]) / (ncut - c + 1);
  pos[ncut] = range_max;
  return FVM_OK;
}

int fvm_conn_add(fvm_block* recv, const int64_t range_min[3], const int64_t range_max[3],
                 fvm_block* donor, const int64_t donor_range_min[3], const int64_t donor_range_max[3],
                 const int donor_type[3]) {
  if (!recv || !donor) return fail(FVM_EINVAL, "null block");
  if (!range_min || !range_max || !donor_range_min || !donor_range_max || !donor_type)
    return fail(FVM_EINVAL, "null range or transform");
  if (recv->ctx != donor->ctx) return fail(FVM_EINVAL, "blocks belong to different contexts");
  // (a context holds blocks of one dimension; kept for the patch kernel's sake)
  if (recv->dim != donor->dim) return fail(FVM_EINVAL, "a 2-D and a 3-D block cannot be connected");
  fvm_ctx* ctx = recv->ctx;
  const int dim = recv->dim;
  const int nr[3] = {recv->g.nx, recv->g.ny, recv->g.nz};
  const int nd[3] = {donor->g.nx, donor->g.ny, donor->g.nz};
  // the transform: a signed permutation (2-D: k maps to k)
  int used[3] = {0, 0, 0};
  fvm_ctx::Patch q{};
  for (int e = 0; e < 3; ++e) {
    const int t = donor_type[e];
    if (t == 0 || t < -3 || t > 3 || used[std::abs(t) - 1]++)
      return fail(FVM_EINVAL, "donor_type must be a signed permutation of 1, 2, 3");
    q.map_dir[e] = std::abs(t) - 1;
    q.map_sign[e] = t > 0 ? 1 : -1;
  }
  if (dim == 2 && donor_type[2] != 3) return fail(FVM_EINVAL, "2-D blocks: donor_type[2] must be 3");
  // the receiving range: a rectangle on a boundary face (1-based, inclusive);
  // the donor range corresponds point for point through the transform
  if (dim == 2 && (nr[2] != 1 || range_min[2] != 1 || range_max[2] != 1))
    return fail(FVM_EINVAL, "2-D blocks: the k range must be 1..1");
  for (int e = 0; e < 3; ++e) {
    if (range_min[e] < 1 || range_max[e] > nr[e] || range_min[e] > range_max[e])
      return fail(FVM_EINVAL, "range out of the receiving block in direction %d", e);
    const int b = q.map_dir[e];
    if (donor_range_max[e] - donor_range_min[e] != q.map_sign[e] * (range_max[e] - range_min[e]))
      return fail(FVM_EINVAL, "donor range does not match the range in direction %d", e);
    const int64_t lo = std::min(donor_range_min[e], donor_range_max[e]);
    const int64_t hi = std::max(donor_range_min[e], donor_range_max[e]);
    if (lo < 1 || hi > nd[b]) return fail(FVM_EINVAL, "donor range out of the donor block");
    q.rmin[e] = range_min[e] - 1;
    q.rmax[e] = range_max[e] - 1;
    q.dmin[b] = donor_range_min[e] - 1;
  }
  // the face: a direction where the range is one vertex plane on a boundary
  // of the receiver, mapped onto a boundary plane of the donor so that the
  // receiver's ghosts land in the donor's interior (the interface vertex
  // plane belongs to both blocks)
  int face = -1;
  for (int e = 0; e < dim && face < 0; ++e) {
    if (range_min[e] != range_max[e] || nr[e] < 2) continue;
    const int side = range_min[e] == nr[e] ? 1 : (range_min[e] == 1 ? 0 : -1);
    if (side < 0) continue;
    const int bn = q.map_dir[e];
    const int64_t dface = q.dmin[bn];
    if (dface != 0 && dface != nd[bn] - 1) continue;
    const int inward = dface == 0 ? 1 : -1;
    if ((side ? 1 : -1) * q.map_sign[e] != inward) continue;
    if (nd[bn] < 2 * kGhost) return fail(FVM_EINVAL, "the donor needs >= %d vertices across the face", 2 * kGhost);
    face = 2 * e + side;
  }
  if (face < 0)
    return fail(FVM_EINVAL, "the ranges do not describe a receiver face onto a donor face with ghosts "
                            "mapped into the donor");
  if (recv->g.bc[face] == BC_CONN || recv->g.bc[face] == BC_PERIODIC)
    return fail(FVM_EINVAL, "face %d of the receiving block already has a connection or periodic BC", face);
  if (recv->g.bc[face] == BC_PATCH && !face_connected(ctx, recv, face))
    drop_face_patches(ctx, recv, face);   // a wall / symmetry / state face becomes connected
  q.recv = recv;
  q.donor = donor;
  q.face = face;
  ctx->patches.push_back(q);
  recv->g.bc[face] = BC_PATCH;
  ctx->desc_dirty = true;
  ctx->ghosts_dirty = true;
  return FVM_OK;
}

// host (Fortran order, 5 equations, host_ng ghosts per swept side) -> device
int fvm_upload(fvm_block* b, const double* q, int host_ng) {
  if (!b || !q) return fail(FVM_EINVAL, "null block or host array");
  if (host_ng < 0 || host_ng > 16) return fail(FVM_EINVAL, "host_ng %d out of range", host_ng);
  fvm_ctx* ctx = b->ctx;
  if (int rc = set_device(ctx)) return rc;
  std::vector<double> stage;
  if (int rc = pack_state(b, q, host_ng, stage)) return rc;
  HIP_TRY(hipMemcpyAsync(b->q[0], stage.data(), b->state_elems * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  b->uploaded = true;
  ctx->ghosts_dirty = true;
  ctx->stage_next = 1;
  return FVM_OK;
}

int fvm_upload_planes(fvm_block* b, const double* q, int host_ng, int64_t host_nk, int64_t k_plane) {
  if (!b || !q) return fail(FVM_EINVAL, "null block or host array");
  if (host_ng < 0 || host_ng > 16) return fail(FVM_EINVAL, "host_ng %d out of range", host_ng);
  if (b->dim == 3) {
    if (host_nk != b->g.nz)
      return fail(FVM_EINVAL, "3-D block: host_nk %lld must equal nk %d", (long long)host_nk, b->g.nz);
    return fvm_upload(b, q, host_ng);
  }
  if (host_nk < 1 || host_nk > (1 << 20)) return fail(FVM_EINVAL, "host_nk %lld out of range", (long long)host_nk);
  if (k_plane < 1 || k_plane > host_nk)
    return fail(FVM_EINVAL, "k_plane %lld outside 1..%lld", (long long)k_plane, (long long)host_nk);
  fvm_ctx* ctx = b->ctx;
  if (int rc = set_device(ctx)) return rc;
  std::vector<double> stage;
  if (int rc = pack_state(b, q, host_ng, stage, host_nk, k_plane)) return rc;
  HIP_TRY(hipMemcpyAsync(b->q[0], stage.data(), b->state_elems * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  b->uploaded = true;
  ctx->ghosts_dirty = true;
  ctx->stage_next = 1;
  return FVM_OK;
}

int fvm_download(fvm_block* b, double* q, int host_ng) { return fvm_download_planes(b, q, host_ng, 0); }

int fvm_download_planes(fvm_block* b, double* q, int host_ng, int64_t host_nk) {
  if (!b || !q) return fail(FVM_EINVAL, "null block or host array");
  if (host_ng < 0 || host_ng > 16) return fail(FVM_EINVAL, "host_ng %d out of range", host_ng);
  if (b->dim == 3 && host_nk != 0 && host_nk != b->g.nz)
    return fail(FVM_EINVAL, "3-D block: host_nk %lld must equal nk %d", (long long)host_nk, b->g.nz);
  if (b->dim == 2 && (host_nk < 0 || host_nk > (1 << 20)))
    return fail(FVM_EINVAL, "host_nk %lld out of range", (long long)host_nk);
  fvm_ctx* ctx = b->ctx;
  if (int rc = set_device(ctx)) return rc;
  if ((ctx->ghosts_dirty || (ctx->ghosts_stale && host_ng > 0)) && b->uploaded) {
    bool all = true;
    for (fvm_block* o : ctx->blocks) all = all && o->uploaded;
    if (all) {
      if (int rc = sync_desc(ctx)) return rc;
      if (int rc = fill_ghosts_all(ctx, 0)) return rc;
      ctx->ghosts_dirty = false;
      ctx->ghosts_stale = false;
    }
  }
  std::vector<double> stage((size_t)b->state_elems);
  HIP_TRY(hipMemcpyAsync(stage.data(), b->q[0], b->state_elems * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  unpack_state(b, stage.data(), q, host_ng, b->dim == 2 ? host_nk : 0);
  return FVM_OK;
}

int fvm_step(fvm_ctx* ctx, int nsteps, double dt, double cfl, double* dt_last) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->scheme_set) return fail(FVM_ESTATE, "fvm_scheme_set has not been called");
  if (ctx->blocks.empty()) return fail(FVM_ESTATE, "no blocks");
  if (nsteps < 0) return fail(FVM_EINVAL, "nsteps must be >= 0");
  const bool cfl_mode = !(dt > 0.0);
  if (cfl_mode && !(cfl > 0.0)) return fail(FVM_EINVAL, "dt <= 0 needs cfl > 0");
  if (ctx->stage_next != 1)
    return fail(FVM_ESTATE, "a step begun with fvm_stage is at stage %d", ctx->stage_next);
  for (fvm_block* b : ctx->blocks)
    if (!b->uploaded) return fail(FVM_ESTATE, "block state not uploaded");
  if (int rc = set_device(ctx)) return rc;
  ctx->have_norms = false;
  if (nsteps == 0) return FVM_OK;
  if (int rc = sync_desc(ctx)) return rc;
  if (ctx->ghosts_dirty) {
    if (int rc = fill_ghosts_all(ctx, 0)) return rc;
    ctx->ghosts_dirty = false;
  }

  // norm partials: NC doubles per tile
  const int nc = ctx->blocks[0]->nc;
  const size_t need = (size_t)ctx->total_tiles * nc;
  if (int rc = ensure_norm_capacity(ctx, need)) return rc;

  if (cfl_mode) {
    // the device slot restarts from the host's time; the wave-speed maximum
    // of u^n is taken afresh (later steps take it from their last stage)
    ctx->h_slot = SolveSlot{};
    ctx->h_slot.t = ctx->time;
    ctx->h_slot.t_end = ctx->t_end;
    ctx->h_slot.cfl = cfl;
    HIP_TRY(hipMemcpyAsync(ctx->d_slot, &ctx->h_slot, sizeof(SolveSlot), hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->d_smax, &kZeroSmax, sizeof(SmaxWays), hipMemcpyHostToDevice,
                           ctx->stream));
    for (fvm_block* b : ctx->blocks) {
      if (b->nc == 5)
        hipLaunchKernelGGL(cfl_max_kernel<5>, dim3(512), dim3(256), 0, ctx->stream, b->q[0], b->g, ctx->P.gamma, ctx->d_smax);
      else
        hipLaunchKernelGGL(cfl_max_kernel<4>, dim3(512), dim3(256), 0, ctx->stream, b->q[0], b->g, ctx->P.gamma, ctx->d_smax);
      HIP_TRY(hipGetLastError());
    }
  }
  // with an end time a step may be skipped on the device, so every step
  // leaves its norm partials (the last step that ran is the one reported)
  const bool may_halt = cfl_mode && std::isfinite(ctx->t_end);
  if (!cfl_mode) {
    hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, ctx->stream, ctx->d_slot, dt);
    HIP_TRY(hipGetLastError());
  }

  const int nstage = ctx->rk == RK4 ? 4 : 3;
  size_t nev = 0;
  std::vector<int> ev_step;   // step of each timed event pair
  if (ctx->timing_every) {
    const size_t need_ev = 2 * (size_t)((nsteps + ctx->timing_every - 1) / ctx->timing_every + 1);
    while (ctx->ev.size() < need_ev) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      ctx->ev.push_back(e);
    }
  }
  ctx->ghosts_stale = true;
  for (int s = 0; s < nsteps; ++s) {
    // the call's last step reports its norms; an earlier one does if it
    // ends the run (the steps after it are skipped on the device)
    const int norm_mode = s == nsteps - 1 ? 1 : (may_halt ? 2 : 0);
    const bool timed = ctx->timing_every && (ctx->timing_step + s) % ctx->timing_every == 0;
    if (ctx->use_graph && !timed) {
      hipGraphExec_t exec = nullptr;
      if (int rc = step_graph(ctx, cfl_mode, norm_mode, &exec)) return rc;
      HIP_TRY(hipGraphLaunch(exec, ctx->stream));
    } else {
      const size_t before = nev;
      if (int rc = launch_step(ctx, cfl_mode, norm_mode, timed ? &nev : nullptr, ctx_mem(ctx))) return rc;
      for (size_t e = before; e < nev; e += 2) ev_step.push_back(s);
    }
  }

  // residual norms of the last stage (class_iteration.f90:210-248): per-tile
  // partials summed in tile order per block; blocks combined as the
  // reference accumulates them, R = sqrt(R_prev^2 + R_blk^2)
  ctx->h_norm.resize(need);
  HIP_TRY(hipMemcpyAsync(ctx->h_norm.data(), ctx->d_norm, need * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  if (cfl_mode)
    HIP_TRY(hipMemcpyAsync(&ctx->h_slot, ctx->d_slot, sizeof(SolveSlot), hipMemcpyDeviceToHost,
                           ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  double dt_host = dt;
  int ran = nsteps;
  if (cfl_mode) {
    dt_host = ctx->h_slot.dt;
    ran = ctx->h_slot.steps;
    ctx->time = ctx->h_slot.t;
  } else {
    for (int s = 0; s < nsteps; ++s) ctx->time = ctx->time + dt;
  }
  ctx->steps_done += ran;
  if (dt_last) *dt_last = dt_host;
  if (ctx->timing_every) {
    // launches of steps skipped on the device (past t_end) did no work
    double cells = 0.0;
    for (fvm_block* b : ctx->blocks) cells += (double)b->g.nx * b->g.ny * b->g.nz;
    for (size_t e = 0; e < nev; e += 2) {
      if (ev_step[e / 2] >= ran) continue;
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[e], ctx->ev[e + 1]));
      ctx->stage_ms += ms;
      ctx->stage_launches += nstage;
      ctx->cell_updates += cells * nstage;
    }
    ctx->timing_step += nsteps;
  }
  std::vector<double> block_sums(ctx->blocks.size() * nc);
  sum_norm_partials(ctx, nc, block_sums.data());
  int nan = combine_norms(ctx, nc, block_sums.data());
  if (cfl_mode && ctx->h_slot.nan) nan = 1;
  if (ran == 0) {
    // the run had already reached its end time: nothing new to report
    ctx->have_norms = false;
    return FVM_OK;
  }
  ctx->nan_flag = nan;
  ctx->have_norms = true;
  if (cfl_mode && ctx->h_slot.nan) return fail(FVM_ENAN, "CFL wave speed is NaN or zero");
  return FVM_OK;
}

// One Runge-Kutta stage, the body of the stage loop of
// ClassTimeDiscret%Runge_Kutta_sub (src/solve/class_time_discret.f90:
// 172-195): residual of the stage input, combine, ghosts.  Stage 1 takes the
// current u^n as the step's base state (reset_q0, :170); the last stage
// writes u^(n+1), advances the solution time by dt and leaves the norms.
int fvm_stage(fvm_ctx* ctx, int s, double dt) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->scheme_set) return fail(FVM_ESTATE, "fvm_scheme_set has not been called");
  if (ctx->blocks.empty()) return fail(FVM_ESTATE, "no blocks");
  const int nstage = ctx->rk == RK4 ? 4 : 3;
  if (s < 1 || s > nstage) return fail(FVM_EINVAL, "stage %d out of range 1..%d", s, nstage);
  if (s != ctx->stage_next) return fail(FVM_ESTATE, "stage %d called, stage %d expected", s, ctx->stage_next);
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(FVM_EINVAL, "fvm_stage needs a fixed dt > 0");
  if (s > 1 && dt != ctx->stage_dt)
    return fail(FVM_EINVAL, "dt %.17g differs from the step's dt %.17g", dt, ctx->stage_dt);
  for (fvm_block* b : ctx->blocks)
    if (!b->uploaded) return fail(FVM_ESTATE, "block state not uploaded");
  if (int rc = set_device(ctx)) return rc;
  const int nc = ctx->blocks[0]->nc;
  const size_t need = (size_t)ctx->total_tiles * nc;
  if (s == 1) {
    ctx->have_norms = false;
    if (int rc = sync_desc(ctx)) return rc;
    if (ctx->ghosts_dirty) {
      if (int rc = fill_ghosts_all(ctx, 0)) return rc;
      ctx->ghosts_dirty = false;
    }
    if (int rc = ensure_norm_capacity(ctx, (size_t)ctx->total_tiles * nc)) return rc;
    hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, ctx->stream, ctx->d_slot, dt);
    HIP_TRY(hipGetLastError());
    ctx->stage_dt = dt;
    ctx->ghosts_stale = true;
  } else if (ctx->desc_dirty) {
    return fail(FVM_ESTATE, "blocks or launch shapes changed within a step");
  }
  if (int rc = launch_one_stage(ctx, s, false, s == nstage ? 1 : 0, ctx_mem(ctx))) return rc;
  if (s < nstage) {
    ctx->stage_next = s + 1;
    HIP_TRY(hipGetLastError());
    return FVM_OK;
  }
  ctx->stage_next = 1;
  ctx->h_norm.resize(need);
  HIP_TRY(hipMemcpyAsync(ctx->h_norm.data(), ctx->d_norm, need * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->time = ctx->time + dt;
  ctx->steps_done += 1;
  std::vector<double> block_sums(ctx->blocks.size() * nc);
  sum_norm_partials(ctx, nc, block_sums.data());
  ctx->nan_flag = combine_norms(ctx, nc, block_sums.data());
  ctx->have_norms = true;
  return FVM_OK;
}

// fvm_step without any allocation: the step's table, time slot and norm
// partials live in the device's static scratch array (see its header entry).
int fvm_step_noalloc(fvm_ctx* ctx, int nsteps, double dt, double cfl, double* dt_last) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->scheme_set) return fail(FVM_ESTATE, "fvm_scheme_set has not been called");
  if (ctx->blocks.empty()) return fail(FVM_ESTATE, "no blocks");
  if (nsteps < 0) return fail(FVM_EINVAL, "nsteps must be >= 0");
  const bool cfl_mode = !(dt > 0.0);
  if (cfl_mode && !(cfl > 0.0)) return fail(FVM_EINVAL, "dt <= 0 needs cfl > 0");
  if (ctx->stage_next != 1)
    return fail(FVM_ESTATE, "a step begun with fvm_stage is at stage %d", ctx->stage_next);
  if ((int)ctx->blocks.size() > kScratchBlocks)
    return fail(FVM_EINVAL, "fvm_step_noalloc takes at most %d blocks (context has %zu)", kScratchBlocks,
                ctx->blocks.size());
  if (!ctx->patches.empty())
    return fail(FVM_EINVAL, "fvm_step_noalloc does not take connection patches or wall, symmetry, "
                "inflow or farfield faces");
  for (fvm_block* b : ctx->blocks)
    if (!b->uploaded) return fail(FVM_ESTATE, "block state not uploaded");
  unsigned char* base = scratch_ptr(ctx->device);
  if (!base) return fail(FVM_ESTATE, "no scratch memory on device %d", ctx->device);
  if (int rc = set_device(ctx)) return rc;
  ctx->have_norms = false;
  if (nsteps == 0) return FVM_OK;
  const StepMem mem{(const BlockDesc*)(base + kScratchDescOff), (double*)(base + kScratchNormOff),
                    (SolveSlot*)base, (SmaxWays*)(base + kScratchSmaxOff)};
  double* d_sums = (double*)(base + kScratchSumOff);
  // the block table, staged on the stack (the copy is complete before return)
  BlockDesc h_desc[kScratchBlocks];
  fill_desc(ctx, h_desc, &mem.slot->dt);
  if (ctx->total_tiles > kScratchTiles)
    return fail(FVM_EINVAL, "fvm_step_noalloc takes at most %d stage tiles (the blocks have %d)",
                kScratchTiles, ctx->total_tiles);
  const int nb = (int)ctx->blocks.size();
  const int nc = ctx->blocks[0]->nc;
  HIP_TRY(hipMemcpyAsync((void*)mem.desc, h_desc, nb * sizeof(BlockDesc), hipMemcpyHostToDevice,
                         ctx->stream));
  if (ctx->ghosts_dirty) {
    if (int rc = fill_ghosts_tab(ctx, mem.desc, 0)) return rc;
    ctx->ghosts_dirty = false;
  }
  if (cfl_mode) {
    ctx->h_slot = SolveSlot{};
    ctx->h_slot.t = ctx->time;
    ctx->h_slot.t_end = ctx->t_end;
    ctx->h_slot.cfl = cfl;
    HIP_TRY(hipMemcpyAsync(mem.slot, &ctx->h_slot, sizeof(SolveSlot), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(mem.smax, &kZeroSmax, sizeof(SmaxWays), hipMemcpyHostToDevice, ctx->stream));
    for (fvm_block* b : ctx->blocks) {
      if (b->nc == 5)
        hipLaunchKernelGGL(cfl_max_kernel<5>, dim3(512), dim3(256), 0, ctx->stream, b->q[0], b->g,
                           ctx->P.gamma, mem.smax);
      else
        hipLaunchKernelGGL(cfl_max_kernel<4>, dim3(512), dim3(256), 0, ctx->stream, b->q[0], b->g,
                           ctx->P.gamma, mem.smax);
      HIP_TRY(hipGetLastError());
    }
  } else {
    hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, ctx->stream, mem.slot, dt);
    HIP_TRY(hipGetLastError());
  }
  const bool may_halt = cfl_mode && std::isfinite(ctx->t_end);
  ctx->ghosts_stale = true;
  for (int st = 0; st < nsteps; ++st) {
    const int norm_mode = st == nsteps - 1 ? 1 : (may_halt ? 2 : 0);
    if (int rc = launch_step(ctx, cfl_mode, norm_mode, nullptr, mem)) return rc;
  }
  hipLaunchKernelGGL(norm_block_sum_kernel, dim3((nb * nc + 63) / 64), dim3(64), 0, ctx->stream, mem.desc,
                     nb, ctx->total_tiles, nc, mem.norm, d_sums);
  HIP_TRY(hipGetLastError());
  double h_sums[kScratchBlocks * 5];
  HIP_TRY(hipMemcpyAsync(h_sums, d_sums, nb * nc * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (cfl_mode)
    HIP_TRY(hipMemcpyAsync(&ctx->h_slot, mem.slot, sizeof(SolveSlot), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  double dt_host = dt;
  int ran = nsteps;
  if (cfl_mode) {
    dt_host = ctx->h_slot.dt;
    ran = ctx->h_slot.steps;
    ctx->time = ctx->h_slot.t;
  } else {
    for (int st = 0; st < nsteps; ++st) ctx->time = ctx->time + dt;
  }
  ctx->steps_done += ran;
  if (dt_last) *dt_last = dt_host;
  int nan = combine_norms(ctx, nc, h_sums);
  if (cfl_mode && ctx->h_slot.nan) nan = 1;
  if (ran == 0) return FVM_OK;
  ctx->nan_flag = nan;
  ctx->have_norms = true;
  if (cfl_mode && ctx->h_slot.nan) return fail(FVM_ENAN, "CFL wave speed is NaN or zero");
  return FVM_OK;
}

int fvm_viscous_set(fvm_ctx* ctx, double mu_ref, double t_ref, double sutherland_s, double prandtl,
                    double r_gas) {
  if (int rc = check_ctx(ctx)) return rc;
  if (mu_ref <= 0.0) {
    ctx->visc_mu_ref = 0.0;
    ctx->desc_dirty = true;
    ++ctx->gen;
    return FVM_OK;
  }
  if (!(t_ref > 0.0) || !(prandtl > 0.0) || !(r_gas > 0.0) || !(sutherland_s >= 0.0) ||
      !std::isfinite(mu_ref))
    return fail(FVM_EINVAL, "viscous parameters need t_ref, prandtl, r_gas > 0 and s >= 0");
  for (const fvm_ctx::Patch& q : ctx->patches)
    if (q.fixed)
      return fail(FVM_EINVAL, "viscous terms on blocks with inflow or farfield faces are not supported");
  if (int rc = set_device(ctx)) return rc;
  for (fvm_block* b : ctx->blocks)
    if (int rc = alloc_rv(b)) return rc;
  ctx->visc_mu_ref = mu_ref;
  ctx->visc_t_ref = t_ref;
  ctx->visc_s = sutherland_s;
  ctx->visc_pr = prandtl;
  ctx->visc_r = r_gas;
  ctx->desc_dirty = true;
  ++ctx->gen;
  return FVM_OK;
}

int fvm_set_graph(fvm_ctx* ctx, int enable) {
  if (int rc = check_ctx(ctx)) return rc;
  ctx->use_graph = enable != 0;
  return FVM_OK;
}

int fvm_set_time(fvm_ctx* ctx, double t, double t_end) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!std::isfinite(t)) return fail(FVM_EINVAL, "time must be finite");
  ctx->time = t;
  ctx->t_end = (t_end > t && std::isfinite(t_end)) ? t_end : HUGE_VAL;
  return FVM_OK;
}

int fvm_time(fvm_ctx* ctx, double* t, int64_t* steps) {
  if (int rc = check_ctx(ctx)) return rc;
  if (t) *t = ctx->time;
  if (steps) *steps = ctx->steps_done;
  return FVM_OK;
}

int fvm_norms(fvm_ctx* ctx, double out[6], int* nan_flag) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->have_norms) return fail(FVM_ESTATE, "no step has run");
  if (out) for (int e = 0; e < 6; ++e) out[e] = ctx->norms[e];
  if (nan_flag) *nan_flag = ctx->nan_flag;
  return FVM_OK;
}

// Pick the launch shape by timing launches of the first RK stage per
// candidate (3-D: every compiled tile x chunk in {4, 8, 16, 32}; 2-D: chunk in
// {2, 4, 8, 16}).  Batched modes choose one shape for all blocks (they share a
// launch); per-block mode tunes each block alone.  The probe reads u^n and
// writes only stage buffers, so the state is untouched.
int fvm_autotune(fvm_ctx* ctx, int reps) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->scheme_set) return fail(FVM_ESTATE, "fvm_scheme_set has not been called");
  if (ctx->stage_next != 1)
    return fail(FVM_ESTATE, "a step begun with fvm_stage is at stage %d", ctx->stage_next);
  if (ctx->blocks.empty()) return fail(FVM_ESTATE, "no blocks");
  for (fvm_block* b : ctx->blocks)
    if (!b->uploaded) return fail(FVM_ESTATE, "block state not uploaded");
  if (reps < 1) reps = 1;
  if (int rc = set_device(ctx)) return rc;
  if (int rc = sync_desc(ctx)) return rc;
  if (ctx->ghosts_dirty) {
    if (int rc = fill_ghosts_all(ctx, 0)) return rc;
    ctx->ghosts_dirty = false;
  }
  // scratch word and events, released on every return path
  struct TuneRes {
    hipStream_t s;
    SmaxWays* smax = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~TuneRes() {
      (void)hipStreamSynchronize(s);
      if (smax) (void)hipFree(smax);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } res{ctx->stream};
  if (hipMalloc(&res.smax, sizeof(SmaxWays)) != hipSuccess) {
    res.smax = nullptr;
    return fail(FVM_ENOMEM, "cannot allocate the autotuner's scratch maximum");
  }
  SmaxWays* tune_smax = res.smax;
  HIP_TRY(hipMemsetAsync(tune_smax, 0, sizeof(SmaxWays), ctx->stream));
  HIP_TRY(hipEventCreate(&res.e0));
  HIP_TRY(hipEventCreate(&res.e1));
  hipEvent_t e0 = res.e0, e1 = res.e1;
  hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, ctx->stream, ctx->d_slot, 1.0e-6);
  HIP_TRY(hipGetLastError());
  // 3-D k-chunks: besides the powers of two, 6, 10 and 12 (on blocks of
  // 30-90 vertices a chunk that divides the extents more evenly can save a
  // few per cent: the 27-block blob runs 2.5 % faster at 10 than at 8)
  const int chunks3[] = {4, 6, 8, 10, 12, 16, 32}, chunks2[] = {2, 4, 8, 16};
  const int dim = ctx->blocks[0]->dim;
  const int nchunk = dim == 3 ? 7 : 4;
  const bool together = ctx->launch_mode != FVM_LAUNCH_PER_BLOCK;
  const size_t groups = together ? 1 : ctx->blocks.size();
  for (size_t gi = 0; gi < groups; ++gi) {
    std::vector<fvm_block*> sel;
    if (together) sel = ctx->blocks; else sel.push_back(ctx->blocks[gi]);
    float best = 1e30f;
    int best_tile = sel[0]->tile, best_chunk = sel[0]->chunk;
    const int ntiles = dim == 3 ? kNumTiles3d : kNumTiles2d;
    // two passes over the candidates, each keeping its faster time: the
    // first pass also brings the GPU's clocks up, so the shapes timed first
    // are not timed on a cold device
    for (int pass = 0; pass < 2; ++pass)
    for (int tile = 0; tile < ntiles; ++tile)
      for (int ci = 0; ci < nchunk; ++ci) {
        if (dim == 2 && tile > 0 && ci > 0) break;   // a 2-D tile does not march
        if (dim == 3 && !tile3d_ok(ctx->recon, tile)) break;
        const int chunk = dim == 3 ? chunks3[ci] : chunks2[ci];
        for (fvm_block* b : sel) plan_launch(b, tile, chunk);
        if (int rc = sync_desc(ctx)) return rc;
        // every stage kind of the scheme, the last one with the CFL maximum
        // (into a scratch word); outputs alternate between buffers 1 and 2
        // so the state u^n (buffer 0) is left as it was
        const int nstage = ctx->rk == RK4 ? 4 : 3;
        auto sweep = [&]() -> int {
          for (int st = 1; st <= nstage; ++st) {
            StageArgs a{};
            a.P = ctx->P;
            a.q0_buf = 0;
            a.in_buf = st == 1 ? 0 : (st % 2 == 0 ? 1 : 2);
            a.out_buf = st % 2 == 1 ? 1 : 2;
            a.use_acc = ctx->rk == RK4 ? 1 : 0;
            a.stage = ctx->rk == RK4 ? ST_RK4_1 + (st - 1) : ST_SSP1 + (st - 1);
            a.smax = st == nstage ? tune_smax : nullptr;
            a.blocks = ctx->d_desc; a.nblocks = (int)ctx->blocks.size();
            a.tile0 = together ? 0 : sel[0]->tile_base;
            a.ntiles = together ? ctx->total_tiles : sel[0]->ntiles;
            if (int rc = launch_one(ctx, dim, sel[0]->tile, a, grid_for(ctx, a.ntiles))) return rc;
          }
          return FVM_OK;
        };
        if (int rc = sweep()) return rc;   // warm
        HIP_TRY(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps; ++r)
          if (int rc = sweep()) return rc;
        HIP_TRY(hipEventRecord(e1, ctx->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) { best = ms; best_tile = sel[0]->tile; best_chunk = sel[0]->chunk; }
      }
    for (fvm_block* b : sel) plan_launch(b, best_tile, best_chunk);
  }
  if (int rc = sync_desc(ctx)) return rc;
  return FVM_OK;
}

int fvm_block_launch_tile(fvm_block* b, int* tile_id) {
  if (!b || !tile_id) return fail(FVM_EINVAL, "null block or output");
  *tile_id = b->tile;
  return FVM_OK;
}

int fvm_block_launch_shape(fvm_block* b, int* tile_x, int* tile_y, int* chunk) {
  if (!b) return fail(FVM_EINVAL, "null block");
  if (tile_x) *tile_x = b->dim == 3 ? kTile3d[b->tile][0] : kTile2d[b->tile][0];
  if (tile_y) *tile_y = b->dim == 3 ? kTile3d[b->tile][1] : kTile2d[b->tile][1];
  if (chunk) *chunk = b->chunk;
  return FVM_OK;
}

int fvm_timing_enable(fvm_ctx* ctx, int enable) {
  if (int rc = check_ctx(ctx)) return rc;
  if (enable < 0) return fail(FVM_EINVAL, "timing interval %d must be >= 0", enable);
  ctx->timing_every = enable;
  ctx->timing_step = 0;
  ctx->stage_ms = 0.0;
  ctx->stage_launches = 0;
  ctx->cell_updates = 0.0;
  return FVM_OK;
}

int fvm_timing_read(fvm_ctx* ctx, double* stage_ms, int64_t* launches, double* cell_updates) {
  if (int rc = check_ctx(ctx)) return rc;
  if (stage_ms) *stage_ms = ctx->stage_ms;
  if (launches) *launches = ctx->stage_launches;
  if (cell_updates) *cell_updates = ctx->cell_updates;
  return FVM_OK;
}

int fvm_sync(fvm_ctx* ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  if (int rc = set_device(ctx)) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}

}  // extern "C"

extern "C" {

}  // extern "C"

namespace {

// Validate an initial-state spec for a block of dimension `dim` and fill `s`.
int make_init_spec(int dim, int kind, int nparams, const double* params, const double* origin,
                   double gamma, InitSpec* s) {
  if (nparams < 0 || (nparams > 0 && !params)) return fail(FVM_EINVAL, "null parameters");
  *s = InitSpec{};
  s->kind = kind;
  s->gamma = gamma;
  for (int e = 0; e < 3; ++e) s->origin[e] = origin ? origin[e] : 0.0;
  for (int e = 0; e < 3; ++e)
    if (!std::isfinite(s->origin[e])) return fail(FVM_EINVAL, "non-finite origin");
  for (int c = 0; c < nparams; ++c)
    if (std::isnan(params[c])) return fail(FVM_EINVAL, "NaN parameter %d", c);
  switch (kind) {
    case FVM_IC_REGIONS:
      if (nparams < 11 || nparams % 11 || nparams / 11 > kInitMaxRegions)
        return fail(FVM_EINVAL, "regions: 11 parameters per region, 1..%d regions", kInitMaxRegions);
      s->nregions = nparams / 11;
      for (int r = 0; r < s->nregions; ++r) {
        const double* st = params + 11 * r + 6;
        if (!(st[0] > 0.0) || !(st[4] > 0.0)) return fail(FVM_EINVAL, "region %d: rho and p must be > 0", r);
        if (dim == 2 && st[3] != 0.0) return fail(FVM_EINVAL, "2-D block: region %d has w != 0", r);
      }
      break;
    case FVM_IC_VORTEX:
      if (nparams != 7) return fail(FVM_EINVAL, "vortex: 7 parameters (length, beta, xc, yc, u, v, w)");
      if (!(params[0] > 0.0) || !std::isfinite(params[0])) return fail(FVM_EINVAL, "vortex: length must be > 0");
      if (dim == 2 && params[6] != 0.0) return fail(FVM_EINVAL, "2-D block: w must be 0");
      break;
    case FVM_IC_TGV:
      if (nparams != 1 || !(params[0] > 0.0)) return fail(FVM_EINVAL, "Taylor-Green: 1 parameter, Mach > 0");
      if (dim == 2) return fail(FVM_EINVAL, "Taylor-Green needs a 3-D block");
      break;
    default:
      return fail(FVM_EINVAL, "unknown initial state kind %d", kind);
  }
  for (int c = 0; c < nparams; ++c) s->v[c] = params[c];
  return FVM_OK;
}

void launch_init(const fvm_block* b, const InitSpec& s, hipStream_t stream) {
  const int64_t ncell = (int64_t)b->g.nx * b->g.ny * b->g.nz;
  const unsigned grid = (unsigned)std::min<int64_t>(2048, (ncell + 255) / 256);
  hipLaunchKernelGGL(init_kernel, dim3(grid), dim3(256), 0, stream, b->q[0], b->g, b->nc, s);
}

}  // namespace

extern "C" {

int fvm_block_init_at(fvm_block* b, int kind, int nparams, const double* params,
                      const int64_t first[3]) {
  if (!b) return fail(FVM_EINVAL, "null block");
  fvm_ctx* ctx = b->ctx;
  if (!ctx->scheme_set) return fail(FVM_ESTATE, "fvm_scheme_set has not been called (gamma)");
  InitSpec s;
  if (int rc = make_init_spec(b->dim, kind, nparams, params, nullptr, ctx->P.gamma, &s)) return rc;
  if (first) {
    for (int e = 0; e < 3; ++e)
      if (first[e] < 0 || first[e] > (int64_t(1) << 40)) return fail(FVM_EINVAL, "first vertex out of range");
    s.at_index = 1;
    for (int e = 0; e < 3; ++e) s.first[e] = first[e];
  }
  if (int rc = set_device(ctx)) return rc;
  launch_init(b, s, ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  b->uploaded = true;
  ctx->ghosts_dirty = true;
  ctx->stage_next = 1;
  return FVM_OK;
}

int fvm_block_init(fvm_block* b, int kind, int nparams, const double* params, const double origin[3]) {
  if (!b) return fail(FVM_EINVAL, "null block");
  fvm_ctx* ctx = b->ctx;
  if (!ctx->scheme_set) return fail(FVM_ESTATE, "fvm_scheme_set has not been called (gamma)");
  InitSpec s;
  if (int rc = make_init_spec(b->dim, kind, nparams, params, origin, ctx->P.gamma, &s)) return rc;
  if (int rc = set_device(ctx)) return rc;
  launch_init(b, s, ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  b->uploaded = true;
  ctx->ghosts_dirty = true;
  ctx->stage_next = 1;
  return FVM_OK;
}

}  // extern "C"

// ---- asynchronous snapshots and restart files -------------------------------

namespace {

// variable code -> snapshot_kernel selector
int snap_selector(int code) {
  switch (code) {
    case FVM_VAR_DENSITY: return 0;
    case FVM_VAR_VELOCITY_X: return 1;
    case FVM_VAR_VELOCITY_Y: return 2;
    case FVM_VAR_VELOCITY_Z: return 3;
    case FVM_VAR_PRESSURE: return 4;
    case FVM_VAR_RHO: return 5;
    case FVM_VAR_RHOU: return 6;
    case FVM_VAR_RHOV: return 7;
    case FVM_VAR_RHOW: return 8;
    case FVM_VAR_RHOE: return 9;
  }
  return -1;
}

fvm_ctx::Snap* find_snap(fvm_ctx* ctx, int64_t id) {
  for (auto& sn : ctx->snaps)
    if (id > 0 && sn.id == id) return &sn;
  return nullptr;
}

constexpr char kSnapMagic[8] = {'F', 'V', 'M', 'S', 'N', 'A', 'P', '1'};

// Snapshot file (little endian): magic "FVMSNAP1", int32 nblocks, int32
// nvars, int32 codes[nvars], float64 time, int64 step; per block int64
// ni, nj, nk and float64 dx, dy, dz; then per block nvars arrays of
// ni*nj*nk float64, i fastest.
void snap_write(fvm_ctx::Snap* sn, int device, std::string path) {
  sn->write_rc = FVM_OK;
  if (hipSetDevice(device) != hipSuccess || hipEventSynchronize(sn->done) != hipSuccess) {
    sn->write_rc = FVM_EHIP;
    sn->write_err = "snapshot copy failed";
    return;
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    sn->write_rc = FVM_EINVAL;
    sn->write_err = "cannot open " + path + " for writing";
    return;
  }
  bool ok = std::fwrite(kSnapMagic, 1, 8, f) == 8;
  const int32_t nb = (int32_t)sn->parts.size(), nv = sn->nvars;
  ok = ok && std::fwrite(&nb, 4, 1, f) == 1 && std::fwrite(&nv, 4, 1, f) == 1;
  for (int v = 0; v < nv && ok; ++v) {
    const int32_t c = sn->codes[v];
    ok = std::fwrite(&c, 4, 1, f) == 1;
  }
  ok = ok && std::fwrite(&sn->time, 8, 1, f) == 1 && std::fwrite(&sn->step, 8, 1, f) == 1;
  for (const auto& pt : sn->parts)
    ok = ok && std::fwrite(pt.n, 8, 3, f) == 3 && std::fwrite(pt.h, 8, 3, f) == 3;
  size_t total = 0;
  for (const auto& pt : sn->parts) total += (size_t)nv * pt.n[0] * pt.n[1] * pt.n[2];
  ok = ok && std::fwrite(sn->h_buf, 8, total, f) == total;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    sn->write_rc = FVM_EINVAL;
    sn->write_err = "write to " + path + " failed";
  }
}

}  // namespace

extern "C" {

int fvm_snapshot_begin(fvm_ctx* ctx, int nvars, const int* vars, int64_t* id) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!id || !vars) return fail(FVM_EINVAL, "null variable list or id");
  if (nvars < 1 || nvars > kSnapMaxVars) return fail(FVM_EINVAL, "nvars %d out of range 1..%d", nvars, kSnapMaxVars);
  if (ctx->blocks.empty()) return fail(FVM_ESTATE, "no blocks");
  for (fvm_block* b : ctx->blocks)
    if (!b->uploaded) return fail(FVM_ESTATE, "block state not uploaded");
  SnapSel sel{};
  sel.n = nvars;
  for (int v = 0; v < nvars; ++v) {
    sel.which[v] = snap_selector(vars[v]);
    if (sel.which[v] < 0) return fail(FVM_EINVAL, "unknown variable code %d", vars[v]);
  }
  fvm_ctx::Snap* sn = nullptr;
  for (auto& c : ctx->snaps)
    if (c.id == 0) { sn = &c; break; }
  if (!sn) return fail(FVM_ESTATE, "all %d snapshot slots are in use: release one", kSnapSlots);
  if (int rc = set_device(ctx)) return rc;
  if (!ctx->copy_stream) HIP_TRY(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
  if (!sn->done) {   // recorded once at creation, so every later wait on it is defined
    HIP_TRY(hipEventCreateWithFlags(&sn->done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(sn->done, ctx->copy_stream));
  }
  if (!sn->conv) HIP_TRY(hipEventCreateWithFlags(&sn->conv, hipEventDisableTiming));
  sn->parts.clear();
  size_t total = 0;
  for (fvm_block* b : ctx->blocks) {
    fvm_ctx::Snap::Part pt{b, (int64_t)total, {b->g.nx, b->g.ny, b->g.nz}, {b->g.dx, b->g.dy, b->g.dz}};
    sn->parts.push_back(pt);
    total += (size_t)nvars * b->g.nx * b->g.ny * b->g.nz;
  }
  if (total > sn->cap) {
    // the slot's last copy (its snapshot was released) must land first
    HIP_TRY(hipEventSynchronize(sn->done));
    (void)hipFree(sn->d_buf);
    (void)hipHostFree(sn->h_buf);
    sn->d_buf = nullptr;
    sn->h_buf = nullptr;
    sn->cap = 0;
    if (hipMalloc(&sn->d_buf, total * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&sn->h_buf, total * sizeof(double)) != hipSuccess) {
      (void)hipFree(sn->d_buf);
      (void)hipHostFree(sn->h_buf);
      sn->d_buf = nullptr;
      sn->h_buf = nullptr;
      return fail(FVM_ENOMEM, "cannot allocate a %zu-byte snapshot slot", total * sizeof(double));
    }
    sn->cap = total;
  }
  // the kernel overwrites d_buf only after the slot's previous copy read it
  HIP_TRY(hipStreamWaitEvent(ctx->stream, sn->done, 0));
  for (size_t i = 0; i < ctx->blocks.size(); ++i) {
    fvm_block* b = ctx->blocks[i];
    const int64_t ncell = (int64_t)b->g.nx * b->g.ny * b->g.nz;
    const unsigned grid = (unsigned)std::min<int64_t>(1024, (ncell + 255) / 256);
    double* out = sn->d_buf + sn->parts[i].off;
    if (b->nc == 5)
      hipLaunchKernelGGL(snapshot_kernel<5>, dim3(grid), dim3(256), 0, ctx->stream, b->q[0], b->g,
                         ctx->P.gamma, sel, out);
    else
      hipLaunchKernelGGL(snapshot_kernel<4>, dim3(grid), dim3(256), 0, ctx->stream, b->q[0], b->g,
                         ctx->P.gamma, sel, out);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(sn->conv, ctx->stream));
  HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, sn->conv, 0));
  HIP_TRY(hipMemcpyAsync(sn->h_buf, sn->d_buf, total * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->copy_stream));
  HIP_TRY(hipEventRecord(sn->done, ctx->copy_stream));
  sn->nvars = nvars;
  for (int v = 0; v < nvars; ++v) {
    sn->which[v] = sel.which[v];
    sn->codes[v] = vars[v];
  }
  sn->time = ctx->time;
  sn->step = ctx->steps_done;
  sn->id = ctx->next_snap++;
  *id = sn->id;
  return FVM_OK;
}

int fvm_snapshot_ready(fvm_ctx* ctx, int64_t id, int* ready) {
  if (int rc = check_ctx(ctx)) return rc;
  fvm_ctx::Snap* sn = find_snap(ctx, id);
  if (!sn || !ready) return fail(FVM_EINVAL, "unknown snapshot %lld", (long long)id);
  if (int rc = set_device(ctx)) return rc;
  const hipError_t e = hipEventQuery(sn->done);
  if (e == hipErrorNotReady) {
    *ready = 0;
    return FVM_OK;
  }
  HIP_TRY(e);
  *ready = 1;
  return FVM_OK;
}

int fvm_snapshot_read(fvm_ctx* ctx, int64_t id, fvm_block* blk, double* out, double* time,
                      int64_t* step) {
  if (int rc = check_ctx(ctx)) return rc;
  fvm_ctx::Snap* sn = find_snap(ctx, id);
  if (!sn) return fail(FVM_EINVAL, "unknown snapshot %lld", (long long)id);
  if (int rc = set_device(ctx)) return rc;
  HIP_TRY(hipEventSynchronize(sn->done));
  if (time) *time = sn->time;
  if (step) *step = sn->step;
  if (!blk) return FVM_OK;
  if (!out) return fail(FVM_EINVAL, "null output array");
  for (const auto& pt : sn->parts)
    if (pt.blk == blk) {
      const size_t n = (size_t)sn->nvars * pt.n[0] * pt.n[1] * pt.n[2];
      std::memcpy(out, sn->h_buf + pt.off, n * sizeof(double));
      return FVM_OK;
    }
  return fail(FVM_EINVAL, "the block is not in snapshot %lld", (long long)id);
}

int fvm_snapshot_save(fvm_ctx* ctx, int64_t id, const char* path) {
  if (int rc = check_ctx(ctx)) return rc;
  fvm_ctx::Snap* sn = find_snap(ctx, id);
  if (!sn) return fail(FVM_EINVAL, "unknown snapshot %lld", (long long)id);
  if (!path) return fail(FVM_EINVAL, "null path");
  if (sn->writer.joinable()) return fail(FVM_ESTATE, "snapshot %lld is already being saved", (long long)id);
  sn->write_rc = FVM_OK;
  sn->write_err.clear();
  sn->writer = std::thread(snap_write, sn, ctx->device, std::string(path));
  return FVM_OK;
}

int fvm_snapshot_release(fvm_ctx* ctx, int64_t id) {
  if (int rc = check_ctx(ctx)) return rc;
  fvm_ctx::Snap* sn = find_snap(ctx, id);
  if (!sn) return fail(FVM_EINVAL, "unknown snapshot %lld", (long long)id);
  int rc = FVM_OK;
  if (sn->writer.joinable()) {
    sn->writer.join();
    rc = sn->write_rc;
    if (rc) fail(rc, "%s", sn->write_err.c_str());
  }
  sn->id = 0;
  sn->parts.clear();
  return rc;
}

int fvm_restart_load(fvm_ctx* ctx, const char* path) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!path) return fail(FVM_EINVAL, "null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(FVM_EINVAL, "cannot open %s", path);
  struct Closer { FILE* f; ~Closer() { std::fclose(f); } } closer{f};
  char magic[8];
  int32_t nb = 0, nv = 0;
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kSnapMagic, 8) != 0 ||
      std::fread(&nb, 4, 1, f) != 1 || std::fread(&nv, 4, 1, f) != 1 || nv < 1 || nv > kSnapMaxVars)
    return fail(FVM_EINVAL, "%s is not a snapshot file", path);
  int32_t codes[kSnapMaxVars];
  if (std::fread(codes, 4, nv, f) != (size_t)nv) return fail(FVM_EINVAL, "%s: truncated header", path);
  static const int kCons[5] = {FVM_VAR_RHO, FVM_VAR_RHOU, FVM_VAR_RHOV, FVM_VAR_RHOW, FVM_VAR_RHOE};
  if (nv != 5 || !std::equal(codes, codes + 5, kCons))
    return fail(FVM_EINVAL, "%s does not hold the conserved state (rho, rho u, rho v, rho w, rho E)", path);
  double time = 0.0;
  int64_t step = 0;
  if (std::fread(&time, 8, 1, f) != 1 || std::fread(&step, 8, 1, f) != 1)
    return fail(FVM_EINVAL, "%s: truncated header", path);
  if (nb != (int32_t)ctx->blocks.size())
    return fail(FVM_EINVAL, "%s holds %d blocks, the context %zu", path, nb, ctx->blocks.size());
  for (int32_t i = 0; i < nb; ++i) {
    int64_t n[3];
    double h[3];
    if (std::fread(n, 8, 3, f) != 3 || std::fread(h, 8, 3, f) != 3)
      return fail(FVM_EINVAL, "%s: truncated header", path);
    const BlockGeom& g = ctx->blocks[i]->g;
    if (n[0] != g.nx || n[1] != g.ny || n[2] != g.nz)
      return fail(FVM_EINVAL, "%s: block %d is %lldx%lldx%lld, the context's %dx%dx%d", path, i,
                  (long long)n[0], (long long)n[1], (long long)n[2], g.nx, g.ny, g.nz);
    if (h[0] != g.dx || h[1] != g.dy || h[2] != g.dz)
      return fail(FVM_EINVAL, "%s: block %d spacing differs from the context's", path, i);
  }
  std::vector<double> buf;
  for (int32_t i = 0; i < nb; ++i) {
    fvm_block* b = ctx->blocks[i];
    buf.resize((size_t)5 * b->g.nx * b->g.ny * b->g.nz);
    if (std::fread(buf.data(), 8, buf.size(), f) != buf.size())
      return fail(FVM_EINVAL, "%s: truncated data of block %d", path, i);
    if (int rc = fvm_upload(b, buf.data(), 0)) return rc;
  }
  ctx->time = time;
  ctx->steps_done = step;
  return FVM_OK;
}

}  // extern "C"

// ============================================================================
// Serving loop: continuous batching of independent solves
// ============================================================================
namespace {

using Clock = std::chrono::steady_clock;

struct Job {
  int64_t id = 0;
  fvm_request req{};
  const double* src = nullptr;     // borrowed initial state, copied by the staging thread
  double* q0 = nullptr;            // pinned: initial state (layout of q0 in fvm_request), kept
  size_t q0_cap = 0;               // until a batch behind its upload has completed
  std::atomic<int> staged{0};      // 1: q0 holds the initial state; -1: staging failed
  double* res = nullptr;           // pinned: final state as downloaded (rho w of 2-D not written)
  size_t res_cap = 0;
  hipEvent_t res_ev = nullptr;     // download of res done
  std::vector<double> result;      // final state, filled from res on the first fvm_server_result
  double t_submit = 0, t_admit = -1, t_done = -1;
  int steps = 0;
  double t_final = 0;
  double cell_updates = 0;
  int status = 0;                  // 0 pending, 1 active, 2 done, -1 diverged, -2 not staged,
                                   // -3 refused at admission (err says why)
  std::string err;
  int slot = -1;
  int64_t first_batch = 0;         // first batch that advanced this solve (0: none yet)
  fvm_ctx* group = nullptr;
  fvm_block* blk = nullptr;
  size_t bytes = 0;
};

// Kernel instantiation and scheme parameters: solves that agree on all of
// them share one launch per stage.
using GroupKey = std::tuple<int, int, int, int, int, int, double, double, double, int>;

// Per group: end-of-batch event, pinned descriptor staging and pinned slot
// snapshots, one table of max_active entries each per batch parity.  The
// snapshot is copied on the group's stream between its batches, so it holds
// the slots exactly as the batch left them.
struct GroupRes {
  hipEvent_t ev = nullptr;
  BlockDesc* h_desc = nullptr;
  SolveSlot* h_snap = nullptr;
};

// A retired solve's slot: reusable once every batch that may still reference
// it (up to release_after) has completed.
struct SlotHold {
  int slot;
  int64_t release_after;
};

}  // namespace

struct fvm_server {
  int device = 0;
  double budget = 0;        // HBM bytes solves may hold
  double used = 0;
  int max_active = 0;
  Clock::time_point t0;
  int64_t next_id = 1;
  std::deque<Job*> pending;
  std::vector<Job*> active;
  std::map<int64_t, Job*> jobs;
  std::map<GroupKey, fvm_ctx*> groups;
  std::map<fvm_ctx*, GroupRes> gres;
  SolveSlot* d_slots = nullptr;
  SmaxWays* d_smax = nullptr;     // [max_active]: each slot's wave-speed maximum
  std::vector<int> free_slots;
  std::vector<SlotHold> held;
  int* d_ids = nullptr;           // [2][max_active] slot ids of a batch, per group segment
  int* h_ids = nullptr;           // pinned staging of d_ids
  SolveSlot* h_init = nullptr;    // pinned [max_active]: initial slot contents
  SmaxWays* h_zero_smax = nullptr;   // pinned zeros: clears an admitted solve's maximum
  hipStream_t cstream = nullptr;  // uploads, downloads, allocations of solve state
  hipStream_t sstream = nullptr;  // joins the groups' streams at the end of a batch
  hipEvent_t up_ev = nullptr;     // uploads of the latest admissions
  hipEvent_t bev[2] = {nullptr, nullptr};   // every group's part of a batch done (per parity)
  std::deque<int64_t> inflight;   // batches queued and not yet completed (at most 2)
  int64_t batch_seq = 0;
  int chunk3 = 8, chunk2 = 8;     // k-chunk (3-D) and row chunk (2-D) of served blocks
  double total_cell_updates = 0;
  // pinned host buffers (capacity in doubles) and events kept for reuse, so
  // that uploads and downloads are asynchronous without pinning memory per
  // request
  std::multimap<size_t, double*> pin_free;
  size_t pin_free_doubles = 0;
  std::mutex pin_mu;              // pin_free is shared with the staging thread
  std::vector<hipEvent_t> ev_free;
  // staging thread: copies borrowed initial states into pinned buffers
  std::thread stager;
  std::mutex st_mu;
  std::condition_variable st_cv;
  std::deque<Job*> to_stage;
  bool st_stop = false;
};

namespace {

// A pinned buffer of at least n doubles: a free one no more than twice that
// large, else a new one.
double* pin_get(fvm_server* s, size_t n, size_t* cap) {
  std::lock_guard<std::mutex> lk(s->pin_mu);
  auto it = s->pin_free.lower_bound(n);
  if (it != s->pin_free.end() && it->first <= 2 * n) {
    double* p = it->second;
    *cap = it->first;
    s->pin_free_doubles -= it->first;
    s->pin_free.erase(it);
    return p;
  }
  void* p = nullptr;
  if (hipHostMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
  *cap = n;
  return (double*)p;
}

// Keep a pinned buffer for reuse (up to 2^28 doubles, 2 GiB, in the free list).
void pin_put(fvm_server* s, double* p, size_t cap) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(s->pin_mu);
  s->pin_free.insert({cap, p});
  s->pin_free_doubles += cap;
  while (s->pin_free_doubles > (size_t(1) << 28) && !s->pin_free.empty()) {
    auto last = std::prev(s->pin_free.end());
    s->pin_free_doubles -= last->first;
    (void)hipHostFree(last->second);
    s->pin_free.erase(last);
  }
}

hipEvent_t ev_get(fvm_server* s) {
  if (!s->ev_free.empty()) {
    hipEvent_t e = s->ev_free.back();
    s->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

double now_s(fvm_server* s) {
  return std::chrono::duration<double>(Clock::now() - s->t0).count();
}

// Staging thread: copy each borrowed initial state into a pinned buffer, in
// submission order, so admission finds it ready and uploads it asynchronously.
void stager_main(fvm_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    Job* j = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->st_mu);
      s->st_cv.wait(lk, [s] { return s->st_stop || !s->to_stage.empty(); });
      if (s->st_stop) return;
      j = s->to_stage.front();
      s->to_stage.pop_front();
    }
    const fvm_request& r = j->req;
    const size_t n = (size_t)5 * r.n[0] * r.n[1] * r.n[2];
    double* p = pin_get(s, n, &j->q0_cap);
    if (p) std::memcpy(p, j->src, n * sizeof(double));
    j->q0 = p;
    j->staged.store(p ? 1 : -1, std::memory_order_release);
  }
}

int server_group(fvm_server* s, const fvm_request& r, fvm_ctx** out) {
  const int dim = r.n[2] == 1 ? 2 : 3;
  GroupKey key{dim, r.recon, r.limiter, r.flux, r.rk, r.numerics, r.kappa, r.entropy_fix, r.gamma,
               r.ref_compat};
  auto it = s->groups.find(key);
  if (it != s->groups.end()) { *out = it->second; return FVM_OK; }
  fvm_ctx* ctx = nullptr;
  if (int rc = fvm_ctx_create(s->device, &ctx)) return rc;
  if (int rc = fvm_scheme_set(ctx, r.recon, r.limiter, r.flux, r.rk, r.kappa, r.entropy_fix, r.gamma,
                              r.ref_compat)) { fvm_ctx_destroy(ctx); return rc; }
  if (int rc = fvm_set_numerics(ctx, r.numerics)) { fvm_ctx_destroy(ctx); return rc; }
  // a table for every slot, so it never grows while batches are queued
  GroupRes gr;
  if (hipMalloc(&ctx->d_desc, (size_t)s->max_active * sizeof(BlockDesc)) != hipSuccess ||
      hipHostMalloc((void**)&gr.h_desc, 2 * (size_t)s->max_active * sizeof(BlockDesc)) != hipSuccess ||
      hipHostMalloc((void**)&gr.h_snap, 2 * (size_t)s->max_active * sizeof(SolveSlot)) != hipSuccess ||
      hipEventCreateWithFlags(&gr.ev, hipEventDisableTiming) != hipSuccess) {
    if (gr.h_desc) (void)hipHostFree(gr.h_desc);
    if (gr.h_snap) (void)hipHostFree(gr.h_snap);
    fvm_ctx_destroy(ctx);
    return fail(FVM_ENOMEM, "cannot allocate a serving group's block table");
  }
  ctx->desc_cap = (size_t)s->max_active;
  s->groups[key] = ctx;
  s->gres[ctx] = gr;
  *out = ctx;
  return FVM_OK;
}

// Upload a group's block table for the next batch: staged in the pinned table
// of this batch's parity, copied on the group's stream behind the batches
// already queued there (which still read the previous table).
int server_sync_desc(fvm_server* s, fvm_ctx* g, int parity) {
  BlockDesc* h = s->gres[g].h_desc + (size_t)parity * s->max_active;
  fill_desc(g, h);
  if (!g->blocks.empty())
    HIP_TRY(hipMemcpyAsync(g->d_desc, h, g->blocks.size() * sizeof(BlockDesc), hipMemcpyHostToDevice,
                           g->stream));
  g->desc_dirty = false;
  return FVM_OK;
}

// One request's admission: its group, its block (state from the device pool
// on the copy stream), the initial state, a slot.  On failure nothing of it
// is left in the group and the slot is not taken.
int server_admit_one(fvm_server* s, Job* j) {
  fvm_ctx* g = nullptr;
  if (int rc = server_group(s, j->req, &g)) return rc;
  fvm_block* b = nullptr;
  const fvm_request& r = j->req;
  if (int rc = block_create_impl(g, r.n[0], r.n[1], r.n[2], r.spacing[0], r.spacing[1], r.spacing[2],
                                 s->cstream, &b))
    return rc;
  auto drop = [&](int rc) {
    g->blocks.erase(std::remove(g->blocks.begin(), g->blocks.end(), b), g->blocks.end());
    g->desc_dirty = true;
    free_block_state(b, s->cstream);
    delete b;
    return rc;
  };
  // one tile shape per group so its blocks share a launch (WENO: 16 x 16,
  // the shape the autotuner picks for it on 32-128 vertex blocks; otherwise
  // 32 x 8); long k-marches (the group's launch has tiles enough without
  // splitting each block)
  if (b->dim == 3) plan_launch(b, r.recon == RECON_WENO ? 3 : 1, s->chunk3);
  else plan_launch(b, -1, s->chunk2);
  for (int f = 0; f < (r.n[2] == 1 ? 4 : 6); ++f)
    if (r.bc[f] == FVM_BC_PERIODIC)
      if (int rc = fvm_block_set_bc(b, f, FVM_BC_PERIODIC)) return drop(rc);
  // the compact state goes to the stage array q[1] (unused until the first
  // step) and a kernel scatters it into q[0]; the host copies nothing.  An
  // initial-state spec is computed into q[0] instead.
  const int64_t ncell = r.n[0] * r.n[1] * r.n[2];
  hipError_t e = hipSuccess;
  if (r.ic_kind != 0) {
    InitSpec spec;
    if (int rc = make_init_spec(b->dim, r.ic_kind, r.ic_nparams, r.ic_params, r.ic_origin, r.gamma, &spec))
      return drop(rc);
    launch_init(b, spec, s->cstream);
  } else if (b->dim == 3) {
    e = hipMemcpyAsync(b->q[1], j->q0, 5 * ncell * sizeof(double), hipMemcpyHostToDevice, s->cstream);
  } else {   // 2-D: rho, rho u, rho v, E (rho w is zero)
    e = hipMemcpyAsync(b->q[1], j->q0, 3 * ncell * sizeof(double), hipMemcpyHostToDevice, s->cstream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->q[1] + 3 * ncell, j->q0 + 4 * ncell, ncell * sizeof(double),
                         hipMemcpyHostToDevice, s->cstream);
  }
  if (e == hipSuccess && r.ic_kind == 0)
    hipLaunchKernelGGL(compact_kernel, dim3(1024), dim3(256), 0, s->cstream, b->q[1], b->q[0], b->g,
                       b->nc, 1);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) return drop(fail(FVM_EHIP, "staging a served state: %s", hipGetErrorString(e)));
  b->uploaded = true;
  const int slot = s->free_slots.back();
  s->free_slots.pop_back();
  SolveSlot& init = s->h_init[slot];
  init = SolveSlot{};
  init.t_end = r.t_end;
  init.cfl = r.cfl;
  b->dt_dev = &s->d_slots[slot].dt;
  b->smax_dev = s->d_smax + slot;
  b->halt_dev = &s->d_slots[slot].halt;
  g->desc_dirty = true;
  j->slot = slot;
  j->group = g;
  j->blk = b;
  j->status = 1;
  j->t_admit = now_s(s);
  s->used += (double)j->bytes;
  s->active.push_back(j);
  return FVM_OK;
}

// Admit pending jobs in arrival order while a slot and the HBM budget allow.
// Allocation, zeroing and upload of a solve's state run on the copy stream;
// its group's stream waits for them before the solve's first step.
int server_admit(fvm_server* s) {
  std::vector<Job*> admitted;
  while (!s->pending.empty()) {
    Job* j = s->pending.front();
    const int st = j->staged.load(std::memory_order_acquire);
    if (st == 0) break;   // still being copied: later requests wait their turn
    if (st < 0) {         // no pinned memory for it
      s->pending.pop_front();
      j->status = -2;
      j->t_done = now_s(s);
      continue;
    }
    if (s->free_slots.empty() || s->used + (double)j->bytes > s->budget) break;
    s->pending.pop_front();
    if (server_admit_one(s, j) != FVM_OK) {
      // this request alone fails (its group or block could not be made: out
      // of device memory, say); the ones admitted before it still start
      j->status = -3;
      j->err = g_err;
      j->t_done = now_s(s);
      if (j->q0) {
        pin_put(s, j->q0, j->q0_cap);
        j->q0 = nullptr;
      }
      continue;
    }
    admitted.push_back(j);
  }
  if (admitted.empty()) return FVM_OK;
  HIP_TRY(hipEventRecord(s->up_ev, s->cstream));
  std::vector<fvm_ctx*> waited;
  for (Job* j : admitted) {
    fvm_ctx* g = j->group;
    if (std::find(waited.begin(), waited.end(), g) == waited.end()) {
      HIP_TRY(hipStreamWaitEvent(g->stream, s->up_ev, 0));
      waited.push_back(g);
    }
    HIP_TRY(hipMemcpyAsync(s->d_slots + j->slot, s->h_init + j->slot, sizeof(SolveSlot),
                           hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(s->d_smax + j->slot, s->h_zero_smax, sizeof(SmaxWays), hipMemcpyHostToDevice,
                           g->stream));
    fvm_block* b = j->blk;
    if (b->nc == 5)
      hipLaunchKernelGGL(cfl_max_kernel<5>, dim3(256), dim3(256), 0, g->stream, b->q[0], b->g, g->P.gamma, b->smax_dev);
    else
      hipLaunchKernelGGL(cfl_max_kernel<4>, dim3(256), dim3(256), 0, g->stream, b->q[0], b->g, g->P.gamma, b->smax_dev);
    HIP_TRY(hipGetLastError());
  }
  return FVM_OK;
}

// Queue one batch: `nsteps` time steps of every active solve -- per group,
// the per-solve dt update and one batched launch per RK stage, on the group's
// stream -- then a copy of the slot table to pinned memory once every group's
// part has run.  The host does not wait; a solve that reaches its end time
// halts on the device and its tiles are skipped until it is retired.
int server_enqueue(fvm_server* s, int nsteps) {
  std::map<fvm_ctx*, std::vector<int>> ids;
  for (Job* j : s->active) ids[j->group].push_back(j->slot);
  if (ids.empty()) return FVM_OK;
  const int64_t bid = ++s->batch_seq;
  const int par = (int)(bid & 1);
  int* hid = s->h_ids + (size_t)par * s->max_active;
  int* did = s->d_ids + (size_t)par * s->max_active;
  int off = 0;
  for (auto& kv : ids) {
    fvm_ctx* g = kv.first;
    const int n = (int)kv.second.size();
    if (g->desc_dirty)
      if (int rc = server_sync_desc(s, g, par)) return rc;
    std::memcpy(hid + off, kv.second.data(), n * sizeof(int));
    HIP_TRY(hipMemcpyAsync(did + off, hid + off, n * sizeof(int), hipMemcpyHostToDevice, g->stream));
    const int nstage = g->rk == RK4 ? 4 : 3;
    for (int step = 0; step < nsteps; ++step) {
      hipLaunchKernelGGL(solve_dt_kernel, dim3((n + 63) / 64), dim3(64), 0, g->stream, s->d_slots,
                         s->d_smax, did + off, n);
      HIP_TRY(hipGetLastError());
      for (int st = 1; st <= nstage; ++st) {
        StageArgs a{};
        a.P = g->P;
        a.use_acc = g->rk == RK4 ? 1 : 0;
        if (g->rk == RK4) {
          const int ib[4] = {0, 1, 2, 1}, ob[4] = {1, 2, 1, 0};
          const int sk[4] = {ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4};
          a.in_buf = ib[st - 1]; a.out_buf = ob[st - 1]; a.stage = sk[st - 1];
        } else {
          const int ib[3] = {0, 1, 2}, ob[3] = {1, 2, 0};
          const int sk[3] = {ST_SSP1, ST_SSP2, ST_SSP3};
          a.in_buf = ib[st - 1]; a.out_buf = ob[st - 1]; a.stage = sk[st - 1];
        }
        a.block_smax = st == nstage ? 1 : 0;
        if (int rc = launch_stage_all(g, a)) return rc;
      }
    }
    // the slots as this batch leaves them, before the next batch (already
    // queued behind it on this stream) moves them on
    GroupRes& gr = s->gres[g];
    HIP_TRY(hipMemcpyAsync(gr.h_snap + (size_t)par * s->max_active, s->d_slots,
                           (size_t)s->max_active * sizeof(SolveSlot), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipEventRecord(gr.ev, g->stream));
    HIP_TRY(hipStreamWaitEvent(s->sstream, gr.ev, 0));
    off += n;
  }
  HIP_TRY(hipEventRecord(s->bev[par], s->sstream));
  for (Job* j : s->active)
    if (!j->first_batch) j->first_batch = bid;
  s->inflight.push_back(bid);
  return FVM_OK;
}

// Complete the oldest queued batch: wait for its slot snapshot, free the
// slots no queued batch references any more, and retire the solves it
// finished -- download the state (the solve is halted, so the batches still
// queued leave it alone) and hand its arrays back to the device pool behind
// those batches on the group's stream.
int server_complete(fvm_server* s) {
  const int64_t bid = s->inflight.front();
  s->inflight.pop_front();
  const int par = (int)(bid & 1);
  HIP_TRY(hipEventSynchronize(s->bev[par]));
  std::vector<SlotHold> still;
  for (const SlotHold& h : s->held) {
    if (h.release_after <= bid) s->free_slots.push_back(h.slot);
    else still.push_back(h);
  }
  s->held.swap(still);
  std::vector<Job*> keep;
  for (Job* j : s->active) {
    if (!j->first_batch || j->first_batch > bid) { keep.push_back(j); continue; }
    if (j->q0) {   // the batch ran behind the upload
      pin_put(s, j->q0, j->q0_cap);
      j->q0 = nullptr;
    }
    const SolveSlot& sl = s->gres[j->group].h_snap[(size_t)par * s->max_active + j->slot];
    const bool done = sl.halt || sl.nan || !(sl.t < j->req.t_end - kEndTol * std::fabs(j->req.t_end));
    if (!done) { keep.push_back(j); continue; }
    const fvm_request& r = j->req;
    const int64_t ncell = r.n[0] * r.n[1] * r.n[2];
    const int nstage = r.rk == FVM_RK4 ? 4 : 3;
    j->steps = sl.steps;
    j->t_final = sl.t;
    j->cell_updates = (double)ncell * nstage * sl.steps;
    s->total_cell_updates += j->cell_updates;
    // gather the real vertices into the (idle) stage array q[1], then one
    // asynchronous copy to pinned memory; the block's arrays are released
    // behind it
    fvm_block* b = j->blk;
    j->res = pin_get(s, (size_t)5 * ncell, &j->res_cap);
    j->res_ev = ev_get(s);
    if (!j->res || !j->res_ev) return fail(FVM_ENOMEM, "cannot allocate pinned result memory");
    hipLaunchKernelGGL(compact_kernel, dim3(1024), dim3(256), 0, s->cstream, b->q[0], b->q[1], b->g,
                       b->nc, 0);
    HIP_TRY(hipGetLastError());
    if (b->dim == 3) {
      HIP_TRY(hipMemcpyAsync(j->res, b->q[1], 5 * ncell * sizeof(double), hipMemcpyDeviceToHost,
                             s->cstream));
    } else {
      HIP_TRY(hipMemcpyAsync(j->res, b->q[1], 3 * ncell * sizeof(double), hipMemcpyDeviceToHost,
                             s->cstream));
      HIP_TRY(hipMemcpyAsync(j->res + 4 * ncell, b->q[1] + 3 * ncell, ncell * sizeof(double),
                             hipMemcpyDeviceToHost, s->cstream));
    }
    HIP_TRY(hipEventRecord(j->res_ev, s->cstream));
    fvm_ctx* g = j->group;
    g->blocks.erase(std::remove(g->blocks.begin(), g->blocks.end(), b), g->blocks.end());
    g->desc_dirty = true;
    HIP_TRY(hipStreamWaitEvent(g->stream, j->res_ev, 0));
    free_block_state(b, g->stream);
    delete b;
    j->blk = nullptr;
    s->held.push_back({j->slot, s->batch_seq});
    s->used -= (double)j->bytes;
    j->status = sl.nan ? -1 : 2;
    j->t_done = now_s(s);
  }
  s->active.swap(keep);
  return FVM_OK;
}

}  // namespace

extern "C" {

int fvm_server_create(int device, double hbm_budget_bytes, int max_active, fvm_server** out) {
  if (!out) return fail(FVM_EINVAL, "null output pointer");
  if (max_active < 1 || max_active > 4096) return fail(FVM_EINVAL, "max_active %d out of range", max_active);
  HIP_TRY(hipSetDevice(device));
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  fvm_server* s = new fvm_server();
  s->device = device;
  s->budget = hbm_budget_bytes > 0 ? std::min(hbm_budget_bytes, 0.9 * (double)free_b) : 0.8 * (double)free_b;
  s->max_active = max_active;
  s->t0 = Clock::now();
  if (const char* e = getenv("FVM_SERVE_CHUNK3")) s->chunk3 = std::max(1, atoi(e));
  if (const char* e = getenv("FVM_SERVE_CHUNK2")) s->chunk2 = std::max(1, atoi(e));
  for (int i = max_active - 1; i >= 0; --i) s->free_slots.push_back(i);
  // keep freed solve state in the device pool for the next admission
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
    uint64_t keep_all = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep_all);
    // no hidden waits: memory freed behind queued batches is reused only
    // once those batches have run
    int no = 0;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &no);
  }
  const size_t ns = (size_t)max_active;
  if (hipMalloc(&s->d_slots, ns * sizeof(SolveSlot)) != hipSuccess ||
      hipMalloc(&s->d_smax, ns * sizeof(SmaxWays)) != hipSuccess ||
      hipMalloc(&s->d_ids, 2 * ns * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void**)&s->h_ids, 2 * ns * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void**)&s->h_init, ns * sizeof(SolveSlot)) != hipSuccess ||
      hipHostMalloc((void**)&s->h_zero_smax, sizeof(SmaxWays)) != hipSuccess ||
      hipStreamCreateWithFlags(&s->cstream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->sstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->up_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->bev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->bev[1], hipEventDisableTiming) != hipSuccess) {
    fvm_server_destroy(s);
    return fail(FVM_ENOMEM, "cannot allocate the server's slot tables, streams or events");
  }
  HIP_TRY(hipMemset(s->d_slots, 0, ns * sizeof(SolveSlot)));
  HIP_TRY(hipMemset(s->d_smax, 0, ns * sizeof(SmaxWays)));
  std::memset(s->h_zero_smax, 0, sizeof(SmaxWays));
  s->stager = std::thread(stager_main, s);
  *out = s;
  return FVM_OK;
}

int fvm_server_destroy(fvm_server* s) {
  if (!s) return FVM_OK;
  if (s->stager.joinable()) {
    {
      std::lock_guard<std::mutex> lk(s->st_mu);
      s->st_stop = true;
    }
    s->st_cv.notify_all();
    s->stager.join();
  }
  (void)hipSetDevice(s->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : s->groups) fvm_ctx_destroy(kv.second);
  for (auto& kv : s->gres) {
    if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
    if (kv.second.h_desc) (void)hipHostFree(kv.second.h_desc);
    if (kv.second.h_snap) (void)hipHostFree(kv.second.h_snap);
  }
  for (auto& kv : s->jobs) {
    Job* j = kv.second;
    if (j->q0) (void)hipHostFree(j->q0);
    if (j->res) (void)hipHostFree(j->res);
    if (j->res_ev) (void)hipEventDestroy(j->res_ev);
    delete j;
  }
  for (auto& kv : s->pin_free) (void)hipHostFree(kv.second);
  for (hipEvent_t e : s->ev_free) (void)hipEventDestroy(e);
  (void)hipFree(s->d_slots);
  (void)hipFree(s->d_smax);
  (void)hipFree(s->d_ids);
  if (s->h_ids) (void)hipHostFree(s->h_ids);
  if (s->h_init) (void)hipHostFree(s->h_init);
  if (s->h_zero_smax) (void)hipHostFree(s->h_zero_smax);
  if (s->up_ev) (void)hipEventDestroy(s->up_ev);
  for (hipEvent_t e : s->bev) if (e) (void)hipEventDestroy(e);
  if (s->cstream) (void)hipStreamDestroy(s->cstream);
  if (s->sstream) (void)hipStreamDestroy(s->sstream);
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, s->device) == hipSuccess && pool) (void)hipMemPoolTrimTo(pool, 0);
  delete s;
  return FVM_OK;
}

int fvm_server_submit(fvm_server* s, const fvm_request* r, int64_t* id) {
  if (!s || !r) return fail(FVM_EINVAL, "null server or request");
  if (!r->q0 && r->ic_kind == 0) return fail(FVM_EINVAL, "null initial state");
  // everything admission checks, so a request it would refuse is refused here
  if (int rc = check_block_dims(r->n[0], r->n[1], r->n[2], r->spacing[0], r->spacing[1], r->spacing[2]))
    return rc;
  if (int rc = check_scheme(r->recon, r->limiter, r->flux, r->rk, r->gamma)) return rc;
  if (r->numerics != FVM_NUMERICS_EXACT && r->numerics != FVM_NUMERICS_FAST)
    return fail(FVM_EINVAL, "unknown numerics mode %d", r->numerics);
  if (!(r->t_end > 0.0) || !(r->cfl > 0.0)) return fail(FVM_EINVAL, "t_end and cfl must be positive");
  for (int f = 0; f < 6; ++f) {
    const BcClass cls = bc_class(r->bc[f]);
    if (cls != BCC_EXTRAP && cls != BCC_PERIODIC)
      return fail(FVM_EINVAL, "served solves take extrapolated or periodic faces (face %d: kind %d)", f,
                  r->bc[f]);
  }
  const int dim = r->n[2] == 1 ? 2 : 3;
  const int nc = dim == 3 ? 5 : 4;
  const double cs = (double)row_pitch(r->n[0]) * (r->n[1] + 2 * kGhost) *
                    (dim == 3 ? r->n[2] + 2 * kGhost : 1);
  const size_t ncell = (size_t)r->n[0] * r->n[1] * r->n[2];
  if (r->ic_kind != 0) {
    InitSpec spec;
    if (r->ic_nparams > FVM_IC_MAX_PARAMS) return fail(FVM_EINVAL, "ic_nparams above %d", FVM_IC_MAX_PARAMS);
    if (int rc = make_init_spec(dim, r->ic_kind, r->ic_nparams, r->ic_params, r->ic_origin, r->gamma, &spec))
      return rc;
  } else if (dim == 2) {
    for (size_t e = 0; e < ncell; ++e)
      if (r->q0[3 * ncell + e] != 0.0) return fail(FVM_EINVAL, "2-D request: rho*w must be zero");
  }
  Job* j = new Job();
  j->id = s->next_id++;
  j->req = *r;
  j->req.q0 = nullptr;
  if (r->ic_kind != 0) {   // computed on the device at admission: nothing to stage
    j->staged.store(1, std::memory_order_release);
  } else if (r->borrow) {
    j->src = r->q0;
  } else {
    j->q0 = pin_get(s, 5 * ncell, &j->q0_cap);
    if (!j->q0) {
      delete j;
      return fail(FVM_ENOMEM, "cannot allocate pinned memory for the request");
    }
    std::memcpy(j->q0, r->q0, 5 * ncell * sizeof(double));
    j->staged.store(1, std::memory_order_release);
  }
  j->bytes = (size_t)(4.0 * nc * cs * sizeof(double));
  if ((double)j->bytes > s->budget) {
    delete j;
    return fail(FVM_ENOMEM, "request needs %.3g bytes, budget is %.3g", (double)j->bytes, s->budget);
  }
  j->t_submit = now_s(s);
  s->jobs[j->id] = j;
  s->pending.push_back(j);
  if (r->borrow && r->ic_kind == 0) {
    {
      std::lock_guard<std::mutex> lk(s->st_mu);
      s->to_stage.push_back(j);
    }
    s->st_cv.notify_one();
  }
  if (id) *id = j->id;
  return FVM_OK;
}

int fvm_server_poll(fvm_server* s, int iterations, int steps_per_sync, int* n_pending, int* n_active,
                    int* n_done) {
  if (!s) return fail(FVM_EINVAL, "null server");
  if (steps_per_sync < 1 || steps_per_sync > 1024)
    return fail(FVM_EINVAL, "steps_per_sync %d out of range [1, 1024]", steps_per_sync);
  HIP_TRY(hipSetDevice(s->device));
  for (int it = 0; it < iterations; ++it) {
    if (int rc = server_admit(s)) return rc;
    if (!s->active.empty())
      if (int rc = server_enqueue(s, steps_per_sync)) return rc;
    // keep the newest batch queued while the host works; with nothing left
    // to step, finish everything
    while (s->inflight.size() > 1 || (s->active.empty() && !s->inflight.empty()))
      if (int rc = server_complete(s)) return rc;
    if (s->active.empty() && s->inflight.empty()) break;
  }
  int done = 0;
  for (auto& kv : s->jobs) done += kv.second->status == 2 || kv.second->status < 0;
  if (n_pending) *n_pending = (int)s->pending.size();
  if (n_active) *n_active = (int)s->active.size();
  if (n_done) *n_done = done;
  return FVM_OK;
}

int fvm_server_result(fvm_server* s, int64_t id, double* q_out, double info[8]) {
  if (!s) return fail(FVM_EINVAL, "null server");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return fail(FVM_EINVAL, "unknown request id %lld", (long long)id);
  Job* j = it->second;
  if (info) {
    info[0] = j->t_submit; info[1] = j->t_admit; info[2] = j->t_done;
    info[3] = j->steps; info[4] = j->cell_updates; info[5] = j->t_final;
    info[6] = j->status; info[7] = (double)j->bytes;
  }
  if (j->status == 0 || j->status == 1) return fail(FVM_ESTATE, "request %lld not finished", (long long)id);
  if (j->status == -2)
    return fail(FVM_ENOMEM, "request %lld: no pinned host memory to stage its state", (long long)id);
  if (j->status == -3)
    return fail(FVM_EINVAL, "request %lld refused at admission: %s", (long long)id, j->err.c_str());
  if (j->status < 0) return fail(FVM_ENAN, "request %lld diverged (NaN or non-positive dt)", (long long)id);
  if (j->res) {   // first call: wait for the download, keep the state, release the pinned copy
    HIP_TRY(hipEventSynchronize(j->res_ev));
    const fvm_request& r = j->req;
    const size_t ncell = (size_t)(r.n[0] * r.n[1] * r.n[2]);
    j->result.assign(j->res, j->res + 5 * ncell);
    if (r.n[2] == 1) std::fill(j->result.begin() + 3 * ncell, j->result.begin() + 4 * ncell, 0.0);
    pin_put(s, j->res, j->res_cap);
    j->res = nullptr;
    s->ev_free.push_back(j->res_ev);
    j->res_ev = nullptr;
  }
  if (q_out) std::memcpy(q_out, j->result.data(), j->result.size() * sizeof(double));
  return FVM_OK;
}

}  // extern "C"
