// Host engine behind the C ABI of include/lbmx.h. One engine = one lattice on the current
// HIP device: state buffers in HBM, boundary pieces, functional, design field, adjoint and
// the checkpoint schedule. Setup calls are host-only; device memory is allocated on the
// first call that computes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "checkpoint.hpp"
#include "host_common.hpp"
#include "lbmx.h"
#include "lbmx_common.h"
#include "lbmx_device.hpp"
#include "lbmx_launch.h"

namespace lbmx {

static size_t prec_state_bytes(int prec) {
  switch (prec) {
    case LBMX_FP64: return 8;
    case LBMX_FP32: return 4;
    default: return 2;
  }
}
static size_t prec_lam_bytes(int prec) { return prec == LBMX_FP64 ? 8 : 4; }
static size_t prec_gamma_bytes(int prec) { return prec == LBMX_FP64 ? 8 : 4; }

// zero fill by a kernel (also valid inside a captured graph)
static void zero_async(void* p, size_t bytes, hipStream_t s) { HIP_OK(fill_zero(p, bytes, s)); }

// The sweeps of one design-gradient pass, over opaque state buffers (a single lattice, a batch
// of lattices in lockstep, or labelled host buffers that check the schedule).
struct SweepOps {
  virtual ~SweepOps() = default;
  virtual void init(void* dst) = 0;                                              // p_0
  virtual void fwd(const void* src, void* dst, int t, bool with_obj) = 0;       // p_t from p_{t-1}
  virtual void adj(const void* fsrc, const void* lin, void* lout, int t) = 0;  // nu_{t-1} from nu_t
};

struct PassResult {
  long long recompute = 0;  // forward sweeps re-run in the reverse phase
  int max_slots = 0;        // slots in use at most
};

// Forward phase storing checkpoints in `slotbuf` (Wang dynamic checkpointing, or every state
// when naive), then the reverse phase: the adjoint sweep at step t needs p_{t-1}, restored from
// the nearest checkpoint and re-advanced (LBDynamicCheckpointing, src/topOpt/LBCheckpointing.cc:42-48;
// WangCheckpointing::getForwardSolution, src/topOpt/WangCheckpointing.cc:36-126). Storing a
// checkpoint swaps buffer handles (no copy). On return, lin holds the final adjoint nu_0.
static PassResult checkpointed_pass(SweepOps& ops, int nsteps, bool naive, std::vector<void*>& slotbuf,
                                    void* work0, void* work1, void*& lin, void*& lout, bool with_obj,
                                    int mode = LBMX_CHECKPOINT_DYNAMIC) {
  const int slots = (int)slotbuf.size();
  void* work[2] = {work0, work1};
  auto other_work = [&](const void* src) -> void*& { return src == work[0] ? work[1] : work[0]; };
  std::unique_ptr<WangSchedule> sched;
  if (!naive) {
    sched.reset(new WangSchedule(slots));
    if (mode == LBMX_CHECKPOINT_STATIC) sched->plan_static(nsteps - 1);
    if (mode == LBMX_CHECKPOINT_BINOMIAL) sched->plan_binomial(nsteps);
  }
  PassResult r;
  auto used = [&]() { r.max_slots = std::max(r.max_slots, naive ? slots : slots - sched->free_slots()); };
  ops.init(work[0]);
  auto store = [&](int step, void*& holder) -> void* {
    int s = naive ? step : sched->make_checkpoint(step);
    if (s < 0) return holder;       // a static schedule does not keep this step
    std::swap(slotbuf[s], holder);  // the slot now owns the state; holder gets a spare buffer
    return slotbuf[s];
  };
  void* src = store(0, work[0]);
  used();
  const void* last_input = nullptr;  // p_{nsteps-1}, the input of the last forward sweep
  for (int k = 1; k <= nsteps; ++k) {
    void*& dst = other_work(src);
    if (k == nsteps) last_input = src;
    ops.fwd(src, dst, k, with_obj);
    if (k < nsteps) src = store(k, dst);
    else src = dst;
    used();
  }
  for (int step = nsteps; step >= 1; --step) {
    const int index = step - 1;
    const void* fsrc = nullptr;
    if (naive) {
      fsrc = slotbuf[index];
    } else {
      WangSchedule::Plan plan = sched->get_forward_solution(index);
      if (plan.stored()) {
        // a binomial plan may hand the first reverse step the forward pass's last input
        fsrc = plan.from_slot >= 0 ? slotbuf[plan.from_slot] : last_input;
      } else {
        const void* s2 = slotbuf[plan.from_slot];
        size_t next = 0;
        for (int c = plan.from_step + 1; c <= plan.index; ++c) {
          void*& dst = other_work(s2);
          ops.fwd(s2, dst, c, false);
          ++r.recompute;
          if (next < plan.stores.size() && plan.stores[next].first == c) {
            int s = plan.stores[next].second;
            std::swap(slotbuf[s], dst);
            s2 = slotbuf[s];
            ++next;
          } else {
            s2 = dst;
          }
        }
        fsrc = s2;
      }
      used();
    }
    ops.adj(fsrc, lin, lout, step);
    std::swap(lin, lout);
    if (!naive) sched->release();
  }
  return r;
}

struct Engine {
  lbmx_desc d{};
  int Q = 0, D = 0;
  long long N = 0;
  bool inc = true;
  hipStream_t stream = nullptr;

  std::vector<DevBC> bcs;           // pieces in insertion order
  std::vector<std::shared_ptr<DeviceBuffer>> bc_profiles;  // per-node tables of the pieces (shared by split pieces)
  std::vector<std::shared_ptr<std::vector<double>>> bc_time;  // per-piece time-factor tables (null: none)
  std::vector<DevTerm> terms;       // functional terms
  Layout ls{}, ll{};                // padded layouts of state and adjoint fields
  std::vector<double> gamma_host;   // design field
  bool gamma_dirty = true;          // host copy newer than the device copy
  bool gamma_host_stale = false;    // device copy newer than the host copy (set on the device)
  ObjArgs obj{};                    // acc pointer filled at launch
  bool obj_set = false;

  // device
  DeviceBuffer gamma64_dev;  // FP64 copy of a design set on the device, for the FP32-design forms
  DeviceBuffer gamma_dev, bc_dev, acc_dev, red_dev, sens_dev, lam_a, lam_b, raw_aos;
  DeviceBuffer nside_dev;  // Neumann pieces: adjoint transfer coefficients of the last adjoint sweep
  DeviceBuffer bcrow_dev;  // boundary pieces meeting each 64-node row segment (bc_row_index)
  int nside_count = 0;     // doubles in nside_dev
  std::vector<std::unique_ptr<DeviceBuffer>> pool;  // state buffers: [0],[1] working; rest slots
  bool bc_dirty = true;
  void* cur = nullptr;   // p_t
  void* prev = nullptr;  // p_{t-1} (valid when state_kind == 0)
  int t = 0;
  int state_kind = -1;   // -1 none, 0 p_{t-1} valid, 1 uniform equilibrium, 2 raw AoS in raw_aos
  double init_rho = 1, init_u[3] = {0, 0, 0}, init_T = 0;
  int nslots = 64;
  int ckpt_mode = LBMX_CHECKPOINT_DYNAMIC;  // placement of the checkpoints (lbmx_checkpoint_mode_t)
  bool sens_valid = false;  // sens_dev holds dJ/dgamma of the last gradient or reverse steps
  bool adj_valid = false;   // lam_a holds the adjoint of the last gradient or reverse step
  int adj_t = -1;        // step of the next per-step reverse sweep (-1: lbmx_adjoint_reset not called)
  lbmx_stats stats{};

  explicit Engine(const lbmx_desc& desc) : d(desc) {
    if (d.nx < 1 || d.ny < 1 || d.nz < 1) throw Error{LBMX_ERR_ARG, "grid dimensions must be positive"};
    switch (d.lattice) {
      case LBMX_D2Q9: Q = 9; D = 2; break;
      case LBMX_D3Q19: Q = 19; D = 3; break;
      case LBMX_D2Q9_D2Q5: Q = 14; D = 2; break;
      default: throw Error{LBMX_ERR_ARG, "unknown lattice"};
    }
    if (D == 2 && d.nz != 1) throw Error{LBMX_ERR_ARG, "2D lattices need nz = 1"};
    if (d.precision < LBMX_FP64 || d.precision > LBMX_BF16_SHIFTED) throw Error{LBMX_ERR_ARG, "unknown precision"};
    if (d.design < LBMX_DESIGN_NONE || d.design > LBMX_DESIGN_DARCY) throw Error{LBMX_ERR_ARG, "unknown design coupling"};
    bool ok = false;
    switch (d.collision) {
      case LBMX_BGK: case LBMX_MRT: ok = true; break;
      case LBMX_BGK_STANDARD: ok = d.lattice != LBMX_D2Q9_D2Q5; break;
      case LBMX_MRT_STANDARD: ok = d.lattice == LBMX_D3Q19; break;
      case LBMX_CASCADED: ok = d.lattice == LBMX_D2Q9 && d.design != LBMX_DESIGN_BRINKMAN; break;
      default: ok = false;
    }
    if (!ok) throw Error{LBMX_ERR_SUP, "collision operator not available for this lattice"};
    if (d.design == LBMX_DESIGN_DARCY && !(d.lattice == LBMX_D2Q9 && d.collision == LBMX_MRT))
      throw Error{LBMX_ERR_SUP, "the DarcyMRT2d coupling is defined for D2Q9 MRT"};
    inc = !(d.collision == LBMX_MRT_STANDARD || d.collision == LBMX_BGK_STANDARD);
    N = (long long)d.nx * d.ny * d.nz;
    if (N >= (1LL << 31)) throw Error{LBMX_ERR_SUP, "grids of 2^31 nodes or more are not supported"};
    gamma_host.assign(N, 1.0);
    ls = layout_for((int)prec_state_bytes(d.precision));
    ll = layout_for((int)prec_lam_bytes(d.precision));
    if ((unsigned long long)ls.plane * prec_state_bytes(d.precision) >= (1ULL << 32) ||
        (unsigned long long)ll.plane * prec_lam_bytes(d.precision) >= (1ULL << 32))
      throw Error{LBMX_ERR_SUP, "a population plane must stay below 4 GiB (32-bit lane offsets)"};
    obj.nterm = 0;
  }

  // layout of a field of this lattice with the given element size
  Layout layout_for(int elem_bytes) const {
    switch (d.lattice) {
      case LBMX_D2Q9: return make_layout(d.nx, d.ny, d.nz, D, elem_bytes, Q, LatD2Q9::ex, LatD2Q9::ey, LatD2Q9::ez);
      case LBMX_D3Q19:
        return make_layout(d.nx, d.ny, d.nz, D, elem_bytes, Q, LatD3Q19::ex, LatD3Q19::ey, LatD3Q19::ez);
      default:
        return make_layout(d.nx, d.ny, d.nz, D, elem_bytes, Q, LatD2Q9D2Q5::ex, LatD2Q9D2Q5::ey, LatD2Q9D2Q5::ez);
    }
  }

  ~Engine() {
    if (touched) (void)hipStreamSynchronize(stream);
  }
  bool touched = false;

  Geometry geom() const {
    Geometry g{};
    g.nx = d.nx; g.ny = d.ny; g.nz = d.nz;
    g.nblk_x = (d.nx + kTileX - 1) / kTileX;
    g.nblk_y = (d.ny + kTileY - 1) / kTileY;
    g.nblocks = g.nblk_x * g.nblk_y * d.nz;
    g.N = N;
    return g;
  }

  ModelParams params() const {
    ModelParams m{};
    m.omega = d.omega;
    m.alpha_max = d.alpha_max;
    m.interp_q = d.interp_q;
    m.body[0] = d.body_force[0]; m.body[1] = d.body_force[1]; m.body[2] = d.body_force[2];
    m.tdiff = d.thermal_diffusivity;
    m.ramp_p = d.ramp_p;
    m.ramp_ck = d.ramp_ck;
    m.design = d.design;
    m.omega_f = (float)m.omega;
    m.alpha_max_f = (float)m.alpha_max;
    m.interp_q_f = (float)m.interp_q;
    for (int k = 0; k < 3; ++k) m.body_f[k] = (float)m.body[k];
    m.tdiff_f = (float)m.tdiff;
    m.ramp_p_f = (float)m.ramp_p;
    m.ramp_ck_f = (float)m.ramp_ck;
    return m;
  }

  size_t state_bytes() const { return (size_t)ls.plane * Q * prec_state_bytes(d.precision); }
  size_t lam_bytes() const { return (size_t)ll.plane * Q * prec_lam_bytes(d.precision); }
  int grid_blocks() const { return geom().nblocks; }
  // 64-node row segments: one objective accumulator each (the sweeps' per-wave sums)
  size_t nseg() const { return (size_t)((d.nx + kTileX - 1) / kTileX) * d.ny * d.nz; }

  // ---- boundary pieces: NewLBSolver2d::addBoundaryCondition (src/latticeBoltzmann/NewLBSolver2d.cc:106-155),
  //      getBoxBoundaries (src/core/geometry2d.cc:31-53, src/core/geometry3d.cc:36-85) ----
  struct Piece {
    int box[6];
    int n[3];
  };
  std::vector<Piece> pieces() const {
    const int X = d.nx - 1, Y = d.ny - 1, Z = d.nz - 1;
    std::vector<Piece> p;
    auto add = [&](int x0, int x1, int y0, int y1, int z0, int z1, int a, int b, int c) {
      p.push_back(Piece{{x0, x1, y0, y1, z0, z1}, {a, b, c}});
    };
    if (D == 2) {
      add(1, X - 1, Y, Y, 0, 0, 0, 1, 0);
      add(1, X - 1, 0, 0, 0, 0, 0, -1, 0);
      add(0, 0, 1, Y - 1, 0, 0, -1, 0, 0);
      add(X, X, 1, Y - 1, 0, 0, 1, 0, 0);
      add(0, 0, Y, Y, 0, 0, -1, 1, 0);
      add(X, X, Y, Y, 0, 0, 1, 1, 0);
      add(0, 0, 0, 0, 0, 0, -1, -1, 0);
      add(X, X, 0, 0, 0, 0, 1, -1, 0);
    } else {
      add(1, X - 1, Y, Y, 1, Z - 1, 0, 1, 0);
      add(1, X - 1, 0, 0, 1, Z - 1, 0, -1, 0);
      add(0, 0, 1, Y - 1, 1, Z - 1, -1, 0, 0);
      add(X, X, 1, Y - 1, 1, Z - 1, 1, 0, 0);
      add(1, X - 1, 1, Y - 1, Z, Z, 0, 0, 1);
      add(1, X - 1, 1, Y - 1, 0, 0, 0, 0, -1);
      add(0, 0, Y, Y, 1, Z - 1, -1, 1, 0);
      add(X, X, Y, Y, 1, Z - 1, 1, 1, 0);
      add(1, X - 1, Y, Y, Z, Z, 0, 1, 1);
      add(1, X - 1, Y, Y, 0, 0, 0, 1, -1);
      add(0, 0, 0, 0, 1, Z - 1, -1, -1, 0);
      add(X, X, 0, 0, 1, Z - 1, 1, -1, 0);
      add(1, X - 1, 0, 0, Z, Z, 0, -1, 1);
      add(1, X - 1, 0, 0, 0, 0, 0, -1, -1);
      add(0, 0, 1, Y - 1, Z, Z, -1, 0, 1);
      add(0, 0, 1, Y - 1, 0, 0, -1, 0, -1);
      add(X, X, 1, Y - 1, Z, Z, 1, 0, 1);
      add(X, X, 1, Y - 1, 0, 0, 1, 0, -1);
      add(0, 0, Y, Y, 0, 0, -1, 1, -1);
      add(X, X, Y, Y, 0, 0, 1, 1, -1);
      add(0, 0, Y, Y, Z, Z, -1, 1, 1);
      add(X, X, Y, Y, Z, Z, 1, 1, 1);
      add(0, 0, 0, 0, 0, 0, -1, -1, -1);
      add(X, X, 0, 0, 0, 0, 1, -1, -1);
      add(0, 0, 0, 0, Z, Z, -1, -1, 1);
      add(X, X, 0, 0, Z, Z, 1, -1, 1);
    }
    return p;
  }

  bool supported(int type, const int* n) const {
    int nn = std::abs(n[0]) + std::abs(n[1]) + std::abs(n[2]);
    switch (type) {
      case LBMX_BC_BOUNCEBACK: return true;
      case LBMX_BC_ZOUHE_VELOCITY:
        if (D != 2 || nn != 1) return false;
        return inc ? (n[0] == -1 || n[1] == 1 || n[0] == 1) : n[0] == -1;
      case LBMX_BC_ZOUHE_PRESSURE: return D == 2 && nn == 1 && n[1] == 0;
      // the two nodes inwards must be interior nodes
      case LBMX_BC_ZOUHE_NEUMANN: return D == 2 && nn == 1 && n[1] == 0 && inc && d.nx >= 4;
      case LBMX_BC_THERMAL_NEUMANN: return d.lattice == LBMX_D2Q9_D2Q5 && nn == 1 && n[1] == 0 && d.nx >= 4;
      case LBMX_BC_HECHT_VELOCITY: return D == 3 && nn == 1 && n[0] == -1;
      case LBMX_BC_HECHT_PRESSURE: return D == 3 && nn == 1 && n[0] == 1;
      case LBMX_BC_THERMAL_BOUNCEBACK: return d.lattice == LBMX_D2Q9_D2Q5;
      case LBMX_BC_THERMAL_DIRICHLET: return d.lattice == LBMX_D2Q9_D2Q5 && nn == 1;
    }
    return false;
  }

  static bool intersect(const int* a, const int* b, int* o) {
    for (int k = 0; k < 3; ++k) {
      o[2 * k] = std::max(a[2 * k], b[2 * k]);
      o[2 * k + 1] = std::min(a[2 * k + 1], b[2 * k + 1]);
      if (o[2 * k + 1] < o[2 * k]) return false;
    }
    return true;
  }

  void add_boundary(const lbmx_box& box, const lbmx_bc_desc& bc) { add_boundary_table(box, bc, nullptr, 0, 0, nullptr, 0); }

  // profile: ncomp values per node of `box` in natural order (null: bc.profile); tfac: b(t) for
  // t = 0..nt-1 (null: bc.time_mode)
  void add_boundary_table(const lbmx_box& box, const lbmx_bc_desc& bc, const double* profile, size_t nprofile,
                          int ncomp, const double* tfac, size_t nt) {
    if (bc.type < LBMX_BC_BOUNCEBACK || bc.type > LBMX_BC_THERMAL_NEUMANN)
      throw Error{LBMX_ERR_ARG, "unknown boundary type"};
    const bool valued = bc.type == LBMX_BC_ZOUHE_VELOCITY || bc.type == LBMX_BC_HECHT_VELOCITY ||
                        bc.type == LBMX_BC_ZOUHE_PRESSURE || bc.type == LBMX_BC_HECHT_PRESSURE ||
                        bc.type == LBMX_BC_THERMAL_DIRICHLET;
    if ((profile || tfac) && !valued)
      throw Error{LBMX_ERR_ARG, "value tables apply to velocity, pressure and thermal Dirichlet conditions"};
    if (!profile && bc.profile == LBMX_PROFILE_TABLE) throw Error{LBMX_ERR_ARG, "table profile without a table"};
    if (!tfac && bc.time_mode == LBMX_TIME_TABLE) throw Error{LBMX_ERR_ARG, "table time mode without a table"};
    const long long bnx = (long long)box.x1 - box.x0 + 1, bny = (long long)box.y1 - box.y0 + 1,
                    bnz = (long long)box.z1 - box.z0 + 1;
    std::shared_ptr<DeviceBuffer> prof;
    if (profile) {
      const bool vel = bc.type == LBMX_BC_ZOUHE_VELOCITY || bc.type == LBMX_BC_HECHT_VELOCITY;
      const int D = d.lattice == LBMX_D3Q19 ? 3 : 2;
      if (!(ncomp == 1 || (vel && ncomp == D)))
        throw Error{LBMX_ERR_ARG, "profile components must be 1 (or the lattice dimension for a velocity)"};
      if (bnx < 1 || bny < 1 || bnz < 1 || (long long)nprofile != bnx * bny * bnz * ncomp)
        throw Error{LBMX_ERR_ARG, "profile length != nodes of the box x components"};
      for (size_t k = 0; k < nprofile; ++k)
        if (!std::isfinite(profile[k])) throw Error{LBMX_ERR_ARG, "profile values must be finite"};
      prof = std::make_shared<DeviceBuffer>();
      prof->ensure(nprofile * sizeof(double));
      HIP_OK(hipMemcpy(prof->p, profile, nprofile * sizeof(double), hipMemcpyHostToDevice));
    }
    std::shared_ptr<std::vector<double>> ttab;
    if (tfac) {
      if (nt < 1) throw Error{LBMX_ERR_ARG, "empty time-factor table"};
      for (size_t k = 0; k < nt; ++k)
        if (!std::isfinite(tfac[k])) throw Error{LBMX_ERR_ARG, "time factors must be finite"};
      ttab = std::make_shared<std::vector<double>>(tfac, tfac + nt);
    }
    int dom[6] = {0, d.nx - 1, 0, d.ny - 1, 0, d.nz - 1};
    int ub[6] = {box.x0, box.x1, box.y0, box.y1, box.z0, box.z1};
    int b[6];
    if (!intersect(ub, dom, b)) return;
    auto ps = pieces();
    for (auto& pc : ps) {
      int o[6];
      if (intersect(b, pc.box, o) && !supported(bc.type, pc.n))
        throw Error{LBMX_ERR_SUP, "boundary type " + std::to_string(bc.type) +
                                      " is not implemented for orientation (" + std::to_string(pc.n[0]) + "," +
                                      std::to_string(pc.n[1]) + "," + std::to_string(pc.n[2]) + ")"};
    }
    if ((bc.type == LBMX_BC_ZOUHE_VELOCITY || bc.type == LBMX_BC_HECHT_VELOCITY) && !profile &&
        bc.profile != LBMX_PROFILE_UNIFORM && (bc.radius[0] == 0.0))
      throw Error{LBMX_ERR_ARG, "profile radius must be non-zero"};
    if (!tfac && bc.time_mode == LBMX_TIME_SINE && bc.t_period == 0.0)
      throw Error{LBMX_ERR_ARG, "sine period must be non-zero"};
    int npieces = 0;
    for (auto& pc : ps) {
      int o[6];
      if (intersect(b, pc.box, o)) ++npieces;
    }
    if ((int)bcs.size() + npieces > kMaxBC)
      throw Error{LBMX_ERR_SUP, "more than " + std::to_string(kMaxBC) + " boundary pieces"};
    for (auto& pc : ps) {
      int o[6];
      if (!intersect(b, pc.box, o)) continue;
      DevBC e{};
      e.type = bc.type;
      for (int k = 0; k < 3; ++k) e.nrm[k] = pc.n[k];
      for (int k = 0; k < 6; ++k) e.box[k] = o[k];
      e.profile = bc.profile;
      e.time_mode = bc.time_mode;
      e.inc = inc ? 1 : 0;
      e.value = bc.value;
      for (int k = 0; k < 3; ++k) e.dir[k] = bc.dir[k];
      e.center[0] = bc.center[0]; e.center[1] = bc.center[1];
      e.radius[0] = bc.radius[0]; e.radius[1] = bc.radius[1];
      e.t_period = bc.t_period; e.t_amp = bc.t_amp; e.t_offset = bc.t_offset; e.t_phase = bc.t_phase;
      if (profile) {
        e.profile = LBMX_PROFILE_TABLE;
        e.table = static_cast<const double*>(prof->p);
        e.ncomp = ncomp;
        for (int k = 0; k < 6; ++k) e.tbox[k] = ub[k];
      }
      if (tfac) e.time_mode = LBMX_TIME_TABLE;
      bc_profiles.push_back(prof);
      bc_time.push_back(ttab);
      if (is_neumann(e.type)) {  // two adjoint transfer coefficients per node of the piece
        e.aux = nside_count;
        nside_count += 2 * (o[3] - o[2] + 1);
      }
      bcs.push_back(e);
    }
    bc_dirty = true;
  }

  // ---- functional terms (src/functionals/*.hh) ----
  void set_objective(const lbmx_objective& o) {
    ObjArgs a{};
    std::vector<DevTerm> tv;
    a.time_const = o.time_const;
    auto size = [](const lbmx_box& b) {
      return (double)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1) * (b.z1 - b.z0 + 1);
    };
    auto term = [&](const lbmx_box& b, double w, int kind) {
      if (b.x1 < b.x0 || b.y1 < b.y0 || b.z1 < b.z0) throw Error{LBMX_ERR_ARG, "empty objective box"};
      if (b.x0 < 0 || b.y0 < 0 || b.z0 < 0 || b.x1 >= d.nx || b.y1 >= d.ny || b.z1 >= d.nz)
        throw Error{LBMX_ERR_ARG, "objective box outside the grid"};
      tv.emplace_back();
      a.nterm = (int)tv.size();
      DevTerm& t = tv.back();
      t.box[0] = b.x0; t.box[1] = b.x1; t.box[2] = b.y0; t.box[3] = b.y1; t.box[4] = b.z0; t.box[5] = b.z1;
      t.weight = w;
      t.kind = kind;
    };
    switch (o.kind) {
      case LBMX_OBJ_NONE: break;
      case LBMX_OBJ_PRESSURE_DROP:
        term(o.box_a, 1.0 / size(o.box_a), 0);
        term(o.box_b, -1.0 / size(o.box_b), 0);
        break;
      case LBMX_OBJ_OUTFLOW: term(o.box_a, -1.0 / size(o.box_a), 1); break;
      case LBMX_OBJ_THROUGHFLOW:  // ThroughFlow2d (src/functionals/throughFlow2d.hh:16-88)
        term(o.box_a, -1.0 / size(o.box_a), 1);
        term(o.box_b, -1.0 / size(o.box_b), 1);
        break;
      case LBMX_OBJ_DISSIPATED_POWER:
        term(o.box_a, 1.0, 2);
        term(o.box_b, -1.0, 2);
        break;
      case LBMX_OBJ_HEAT_FLUX:
        if (d.lattice != LBMX_D2Q9_D2Q5) throw Error{LBMX_ERR_SUP, "heat flux needs the thermal lattice"};
        term(o.box_a, -1.0, 3);
        break;
      default: throw Error{LBMX_ERR_ARG, "unknown objective"};
    }
    install_terms(a, tv);
  }

  // explicit terms (lbmx_set_objective_terms): the same representation
  void set_objective_terms(const lbmx_term* t, int nterms, double time_const) {
    if (nterms <= 0 || nterms > kMaxTerms) throw Error{LBMX_ERR_ARG, "1 to 4 functional terms"};
    ObjArgs a{};
    a.time_const = time_const;
    std::vector<DevTerm> tv;
    for (int m = 0; m < nterms; ++m) {
      const lbmx_box& b = t[m].box;
      if (t[m].kind < LBMX_TERM_DENSITY || t[m].kind > LBMX_TERM_TEMPERATURE)
        throw Error{LBMX_ERR_ARG, "unknown functional term"};
      if ((t[m].kind == LBMX_TERM_HEAT_FLUX_X || t[m].kind == LBMX_TERM_TEMPERATURE) && d.lattice != LBMX_D2Q9_D2Q5)
        throw Error{LBMX_ERR_SUP, "temperature terms need the thermal lattice"};
      if (!std::isfinite(t[m].weight)) throw Error{LBMX_ERR_ARG, "functional term weight not finite"};
      if (b.x1 < b.x0 || b.y1 < b.y0 || b.z1 < b.z0) throw Error{LBMX_ERR_ARG, "empty objective box"};
      if (b.x0 < 0 || b.y0 < 0 || b.z0 < 0 || b.x1 >= d.nx || b.y1 >= d.ny || b.z1 >= d.nz)
        throw Error{LBMX_ERR_ARG, "objective box outside the grid"};
      DevTerm dt{};
      dt.box[0] = b.x0; dt.box[1] = b.x1; dt.box[2] = b.y0; dt.box[3] = b.y1; dt.box[4] = b.z0; dt.box[5] = b.z1;
      dt.weight = t[m].weight;
      dt.kind = t[m].kind;
      tv.push_back(dt);
    }
    a.nterm = nterms;
    install_terms(a, tv);
  }

  void install_terms(ObjArgs& a, const std::vector<DevTerm>& tv) {
    for (int m = 0; m < a.nterm; ++m) {
      for (int k = 0; k < 6; ++k) a.tbox[m][k] = tv[m].box[k];
      a.tkind[m] = tv[m].kind;
      a.tweight[m] = tv[m].weight;
    }
    obj = a;
    terms = tv;
    obj_set = true;
  }

  void set_design(const double* g, size_t n) {
    if ((long long)n != N) throw Error{LBMX_ERR_ARG, "design vector length != number of nodes"};
    std::copy(g, g + n, gamma_host.begin());
    gamma_dirty = true;
    gamma_host_stale = false;
  }

  // design from a device FP64 array (natural order), ordered on this engine's stream
  void set_design_device(const double* dev) {
    touched = true;
    gamma_dev.ensure((size_t)N * prec_gamma_bytes(d.precision));
    if (d.precision == LBMX_FP64) {
      HIP_OK(hipMemcpyAsync(gamma_dev.p, dev, (size_t)N * 8, hipMemcpyDeviceToDevice, stream));
    } else {
      gamma64_dev.ensure((size_t)N * 8);
      HIP_OK(hipMemcpyAsync(gamma64_dev.p, dev, (size_t)N * 8, hipMemcpyDeviceToDevice, stream));
      HIP_OK(convert_f64_f32(static_cast<float*>(gamma_dev.p), dev, N, stream));
    }
    gamma_dirty = false;
    gamma_host_stale = true;
  }

  void get_design(double* g, size_t n) {
    if ((long long)n != N) throw Error{LBMX_ERR_ARG, "design vector length != number of nodes"};
    if (gamma_host_stale) {
      const void* src = d.precision == LBMX_FP64 ? gamma_dev.p : gamma64_dev.p;
      HIP_OK(hipStreamSynchronize(stream));
      HIP_OK(hipMemcpy(gamma_host.data(), src, (size_t)N * 8, hipMemcpyDeviceToHost));
      gamma_host_stale = false;
    }
    std::copy(gamma_host.begin(), gamma_host.end(), g);
  }

  // ---- device preparation ----
  void upload_static() {
    if (gamma_dirty) {
      gamma_dev.ensure((size_t)N * prec_gamma_bytes(d.precision));
      if (d.precision == LBMX_FP64) {
        HIP_OK(hipMemcpyAsync(gamma_dev.p, gamma_host.data(), N * 8, hipMemcpyHostToDevice, stream));
      } else {
        std::vector<float> gf(gamma_host.begin(), gamma_host.end());
        HIP_OK(hipMemcpyAsync(gamma_dev.p, gf.data(), N * 4, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
      }
      gamma_dirty = false;
    }
    if (bc_dirty) {
      nside_dev.ensure(std::max<size_t>(1, (size_t)nside_count) * sizeof(double));
      bc_dev.ensure(std::max<size_t>(1, bcs.size()) * sizeof(DevBC));
      if (!bcs.empty())
        HIP_OK(hipMemcpyAsync(bc_dev.p, bcs.data(), bcs.size() * sizeof(DevBC), hipMemcpyHostToDevice, stream));
      // the pieces meeting each 64-node row segment (a sweep wave visits only those)
      const long long nbx = (d.nx + 63) / 64;
      std::vector<unsigned long long> rows((size_t)(nbx * d.ny * d.nz), 0ull);
      for (size_t b = 0; b < bcs.size(); ++b) {
        const int* bx = bcs[b].box;
        for (int z = std::max(bx[4], 0); z <= std::min(bx[5], d.nz - 1); ++z)
          for (int y = std::max(bx[2], 0); y <= std::min(bx[3], d.ny - 1); ++y)
            for (int sx = std::max(bx[0], 0) / 64; sx <= std::min(bx[1], d.nx - 1) / 64; ++sx)
              rows[(size_t)bc_row_index(d.nx, d.ny, sx * 64, y, z)] |= 1ull << b;
      }
      bcrow_dev.ensure(rows.size() * sizeof(unsigned long long));
      HIP_OK(hipMemcpyAsync(bcrow_dev.p, rows.data(), rows.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                            stream));
      HIP_OK(hipStreamSynchronize(stream));
      bc_dirty = false;
    }
  }

  void ensure_pool(size_t count) {
    while (pool.size() < count) {
      pool.emplace_back(new DeviceBuffer());
      pool.back()->ensure(state_bytes());
    }
  }

  // time factor of each piece at step t (the functor's t-dependence, e.g. the sinusoidal
  // inlet modulation of applications/RobustPump2d/robustPump2d.cc:14-19)
  void time_factors(int tstep, double* tf) const {
    for (size_t b = 0; b < bcs.size(); ++b) {
      const DevBC& e = bcs[b];
      if (e.time_mode == LBMX_TIME_TABLE) {  // steps past the table are refused up front (check_time_tables)
        const std::vector<double>& v = *bc_time[b];
        tf[b] = v[std::min<size_t>(std::max(tstep, 0), v.size() - 1)];
        continue;
      }
      tf[b] = e.time_mode == LBMX_TIME_SINE
                  ? e.t_offset + e.t_amp * std::sin(2.0 * M_PI * tstep / e.t_period + e.t_phase)
                  : 1.0;
    }
  }

  // a run that applies the boundary conditions up to step `last` needs every time table to reach it
  void check_time_tables(long long last) const {
    for (size_t b = 0; b < bcs.size(); ++b)
      if (bcs[b].time_mode == LBMX_TIME_TABLE && (long long)bc_time[b]->size() <= last)
        throw Error{LBMX_ERR_ARG, "boundary time-factor table holds steps 0.." + std::to_string(bc_time[b]->size() - 1) +
                                      ", the run reaches step " + std::to_string(last)};
  }

  FwdArgs fwd_args(const void* src, void* dst, int tnew, bool with_obj) {
    FwdArgs a{};
    time_factors(tnew, a.tfac);
    a.src = src;
    a.dst = dst;
    a.gamma = gamma_dev.p;
    a.bcs = static_cast<const DevBC*>(bc_dev.p);
    a.bcrows = static_cast<const unsigned long long*>(bcrow_dev.p);
    a.nbc = (int)bcs.size();
    a.t = tnew;
    a.mode = 0;
    a.g = geom();
    a.ls = ls;
    a.m = params();
    a.obj = obj;
    a.obj.acc = (with_obj && obj_set && obj.nterm > 0) ? static_cast<double*>(acc_dev.p) : nullptr;
    return a;
  }

#define LBMX_PREC_SWITCH(LAT, CALL)                        \
  switch (d.precision) {                                   \
    case LBMX_FP64: return CALL(LAT##_f64);                \
    case LBMX_FP32: return CALL(LAT##_f32);                \
    case LBMX_FP16_SHIFTED: return CALL(LAT##_f16);        \
    default: return CALL(LAT##_bf16);                      \
  }
  hipError_t run_fwd(const FwdArgs& a) {
#define CALL(S) launch_fwd_##S(d.collision, d.design, a, stream)
    switch (d.lattice) {
      case LBMX_D2Q9: LBMX_PREC_SWITCH(d2q9, CALL)
      case LBMX_D3Q19: LBMX_PREC_SWITCH(d3q19, CALL)
      default: LBMX_PREC_SWITCH(thermal, CALL)
    }
#undef CALL
  }
  hipError_t run_init(const FwdArgs& a, const double* macros) {
#define CALL(S) launch_init_##S(d.collision, d.design, a, macros, init_u, init_rho, init_T, stream)
    switch (d.lattice) {
      case LBMX_D2Q9: LBMX_PREC_SWITCH(d2q9, CALL)
      case LBMX_D3Q19: LBMX_PREC_SWITCH(d3q19, CALL)
      default: LBMX_PREC_SWITCH(thermal, CALL)
    }
#undef CALL
  }
  static bool is_neumann(int type) { return type == LBMX_BC_ZOUHE_NEUMANN || type == LBMX_BC_THERMAL_NEUMANN; }

  // an adjoint sweep, then the Neumann pieces' transfers into their inward neighbours
  hipError_t run_adj(const AdjArgs& a) {
    hipError_t e = run_adj_sweep(a);
    if (e != hipSuccess || nside_count == 0) return e;
    for (const DevBC& b : bcs) {
      if (!is_neumann(b.type)) continue;
      e = run_neumann_fix(b, a.lam_out);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  hipError_t run_neumann_fix(const DevBC& b, void* lam) {
    const double* ns = static_cast<const double*>(nside_dev.p);
#define CALL(S) launch_neumann_fix_##S(b, lam, ll, ns, stream)
    switch (d.lattice) {
      case LBMX_D2Q9: LBMX_PREC_SWITCH(d2q9, CALL)
      case LBMX_D3Q19: LBMX_PREC_SWITCH(d3q19, CALL)
      default: LBMX_PREC_SWITCH(thermal, CALL)
    }
#undef CALL
  }
  hipError_t run_adj_sweep(const AdjArgs& a) {
#define CALL(S) launch_adj_##S(d.collision, d.design, a, stream)
    switch (d.lattice) {
      case LBMX_D2Q9: LBMX_PREC_SWITCH(d2q9, CALL)
      case LBMX_D3Q19: LBMX_PREC_SWITCH(d3q19, CALL)
      default: LBMX_PREC_SWITCH(thermal, CALL)
    }
#undef CALL
  }
#undef LBMX_PREC_SWITCH

  AdjArgs adj_args(const void* fsrc, const void* lin, void* lout, int tstep, const ObjArgs& oa) {
    AdjArgs a{};
    a.fwd_src = fsrc;
    a.lam_in = lin;
    a.lam_out = lout;
    a.gamma = gamma_dev.p;
    a.sens = static_cast<double*>(sens_dev.p);
    a.nside = static_cast<double*>(nside_dev.p);
    a.bcs = static_cast<const DevBC*>(bc_dev.p);
    a.bcrows = static_cast<const unsigned long long*>(bcrow_dev.p);
    a.nbc = (int)bcs.size();
    a.t = tstep;
    time_factors(tstep, a.tfac);
    a.g = geom();
    a.ls = ls;
    a.ll = ll;
    a.m = params();
    a.obj = oa;
    a.obj.acc = nullptr;
    return a;
  }

  // design, boundary and functional arrays on the device (no state buffers)
  void prepare_static() {
    touched = true;
    upload_static();
    red_dev.ensure(sizeof(double));
  }

  void prepare() {
    touched = true;
    upload_static();
    ensure_pool(2);
    acc_dev.ensure(nseg() * sizeof(double));
    red_dev.ensure(sizeof(double));
  }

  // ---- forward API ----
  void init_equilibrium(double rho0, const double* u0, double T0, const double* macros_host, size_t n) {
    prepare();
    init_rho = rho0;
    for (int k = 0; k < 3; ++k) init_u[k] = u0 ? u0[k] : 0.0;
    init_T = T0;
    DeviceBuffer mac;
    const double* mac_dev = nullptr;
    if (macros_host) {
      if ((long long)n != 5 * N) throw Error{LBMX_ERR_ARG, "macro field must hold 5 values per node"};
      mac.ensure(n * sizeof(double));
      HIP_OK(hipMemcpyAsync(mac.p, macros_host, n * 8, hipMemcpyHostToDevice, stream));
      mac_dev = static_cast<const double*>(mac.p);
      raw_aos.ensure((size_t)N * Q * 8);
    }
    cur = pool[0]->p;
    prev = pool[1]->p;
    FwdArgs a = fwd_args(nullptr, cur, 0, false);
    a.aos_out = macros_host ? raw_aos.p : nullptr;
    HIP_OK(run_init(a, mac_dev));
    if (macros_host) HIP_OK(hipStreamSynchronize(stream));
    t = 0;
    state_kind = macros_host ? 2 : 1;
  }

  void require_state() const {
    if (state_kind < 0) throw Error{LBMX_ERR_STATE, "state not initialised (call lbmx_init_equilibrium)"};
  }

  void zero_acc() { zero_async(acc_dev.p, nseg() * sizeof(double), stream); }

  double reduce_acc() {
    std::vector<double> h(nseg());
    HIP_OK(hipMemcpyAsync(h.data(), acc_dev.p, h.size() * 8, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    double s = 0;
    for (double v : h) s += v;  // fixed order: deterministic
    return s;
  }

  void forward(int nsteps, double* J) {
    require_state();
    if (nsteps < 0) throw Error{LBMX_ERR_ARG, "nsteps must be >= 0"};
    check_time_tables((long long)t + nsteps);
    prepare();
    bool with_obj = J && obj_set && obj.nterm > 0;
    if (with_obj) zero_acc();
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsteps; ++s) {
      FwdArgs a = fwd_args(cur, prev, t + 1, with_obj);
      HIP_OK(run_fwd(a));
      std::swap(cur, prev);
      ++t;
      state_kind = 0;
    }
    stats.fwd_sweeps += nsteps;
    double Jv = with_obj ? reduce_acc() : 0.0;
    HIP_OK(hipStreamSynchronize(stream));
    stats.forward_phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (J) *J = Jv;
  }

  void get_state(double* f, size_t n) {
    require_state();
    if ((long long)n != N * Q) throw Error{LBMX_ERR_ARG, "state length != nodes * Q"};
    prepare();
    if (state_kind == 2) {
      HIP_OK(hipMemcpy(f, raw_aos.p, n * 8, hipMemcpyDeviceToHost));
      return;
    }
    DeviceBuffer out;
    out.ensure(n * 8);
    if (state_kind == 1) {
      FwdArgs a = fwd_args(nullptr, nullptr, 0, false);
      a.aos_out = out.p;
      HIP_OK(run_init(a, nullptr));
    } else {
      FwdArgs a = fwd_args(prev, nullptr, t, false);
      a.mode = 1;
      a.aos_out = out.p;
      HIP_OK(run_fwd(a));
    }
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipMemcpy(f, out.p, n * 8, hipMemcpyDeviceToHost));
  }

  void set_state(const double* f, size_t n, int timestep) {
    if ((long long)n != N * Q) throw Error{LBMX_ERR_ARG, "state length != nodes * Q"};
    prepare();
    raw_aos.ensure(n * 8);
    HIP_OK(hipMemcpyAsync(raw_aos.p, f, n * 8, hipMemcpyHostToDevice, stream));
    cur = pool[0]->p;
    prev = pool[1]->p;
    FwdArgs a = fwd_args(raw_aos.p, cur, timestep, false);
    a.mode = 2;
    HIP_OK(run_fwd(a));
    HIP_OK(hipStreamSynchronize(stream));
    t = timestep;
    state_kind = 2;
  }

  // ---- gradient: forward with checkpoints, reverse sweep with recomputation ----
  void gradient(int nsteps, double rho0, const double* u0, double T0, double* J, double* sens, size_t n) {
    if (!inc) throw Error{LBMX_ERR_SUP, "no adjoint for the standard (compressible) operators"};
    if (d.design == LBMX_DESIGN_DARCY)
      throw Error{LBMX_ERR_SUP, "DarcyMRT2d has no reverse operator (ReverseDarcyMRT2d is empty)"};
    if (nsteps < 1) throw Error{LBMX_ERR_ARG, "nsteps must be >= 1"};
    if (sens && (long long)n != N) throw Error{LBMX_ERR_ARG, "sensitivity length != nodes"};
    check_time_tables(nsteps);
    prepare();
    const bool naive = nslots <= 0;
    const int slots = naive ? nsteps : std::min(nslots, std::max(2, nsteps));
    ensure_pool(2 + slots);
    std::vector<void*> slotbuf(slots);
    for (int s = 0; s < slots; ++s) slotbuf[s] = pool[2 + s]->p;

    lam_a.ensure(lam_bytes());
    lam_b.ensure(lam_bytes());
    sens_dev.ensure((size_t)N * 8);
    init_rho = rho0;
    for (int k = 0; k < 3; ++k) init_u[k] = u0 ? u0[k] : 0.0;
    init_T = T0;
    const bool with_obj = obj_set && obj.nterm > 0;
    ObjArgs oa = obj;
    oa.acc = nullptr;
    if (!with_obj) oa.nterm = 0;

    struct Ops : SweepOps {
      Engine& e;
      const ObjArgs& oa;
      bool reverse_started = false;
      std::chrono::steady_clock::time_point t_fwd_end;
      Ops(Engine& en, const ObjArgs& o) : e(en), oa(o) {}
      void init(void* dst) override {
        FwdArgs a = e.fwd_args(nullptr, dst, 0, false);
        HIP_OK(e.run_init(a, nullptr));
        if (e.obj_set && e.obj.nterm > 0) e.zero_acc();
      }
      void fwd(const void* src, void* dst, int t, bool with) override {
        FwdArgs a = e.fwd_args(src, dst, t, with);
        HIP_OK(e.run_fwd(a));
      }
      void adj(const void* fsrc, const void* lin, void* lout, int t) override {
        if (!reverse_started) {
          reverse_started = true;
          HIP_OK(hipStreamSynchronize(e.stream));  // end of the forward phase
          t_fwd_end = std::chrono::steady_clock::now();
          zero_async(e.lam_a.p, e.lam_bytes(), e.stream);
          zero_async(e.sens_dev.p, (size_t)e.N * 8, e.stream);
        }
        AdjArgs a = e.adj_args(fsrc, lin, lout, t, oa);
        HIP_OK(e.run_adj(a));
      }
    } ops(*this, oa);

    auto tf0 = std::chrono::steady_clock::now();
    void* lin = lam_a.p;
    void* lout = lam_b.p;
    PassResult pr =
        checkpointed_pass(ops, nsteps, naive, slotbuf, pool[0]->p, pool[1]->p, lin, lout, with_obj, ckpt_mode);
    HIP_OK(hipStreamSynchronize(stream));
    auto tr1 = std::chrono::steady_clock::now();
    const double Jv = with_obj ? reduce_acc() : 0.0;
    stats.fwd_sweeps += nsteps + pr.recompute;
    stats.adj_sweeps += nsteps;
    stats.recompute_sweeps += pr.recompute;
    stats.forward_phase_ms = std::chrono::duration<double, std::milli>(ops.t_fwd_end - tf0).count();
    stats.reverse_phase_ms = std::chrono::duration<double, std::milli>(tr1 - ops.t_fwd_end).count();
    if (lin != lam_a.p) std::swap(lam_a.p, lam_b.p);  // final adjoint in lam_a
    sens_valid = adj_valid = true;
    adj_t = -1;
    state_kind = -1;  // the forward buffers were reused
    if (J) *J = Jv;
    if (sens) HIP_OK(hipMemcpy(sens, sens_dev.p, (size_t)N * 8, hipMemcpyDeviceToHost));
  }

  // ---- per-step reverse loop on a caller-driven forward state ----
  // resetAdjoints + setCurrentTimestep (src/adjointLatticeBoltzmann/adjointLBSolver.cc:138-145)
  void adjoint_reset(int timestep) {
    if (!inc) throw Error{LBMX_ERR_SUP, "no adjoint for the standard (compressible) operators"};
    if (d.design == LBMX_DESIGN_DARCY)
      throw Error{LBMX_ERR_SUP, "DarcyMRT2d has no reverse operator (ReverseDarcyMRT2d is empty)"};
    if (timestep < 0) throw Error{LBMX_ERR_ARG, "timestep must be >= 0"};
    prepare();
    lam_a.ensure(lam_bytes());
    lam_b.ensure(lam_bytes());
    sens_dev.ensure((size_t)N * 8);
    zero_async(lam_a.p, lam_bytes(), stream);
    zero_async(sens_dev.p, (size_t)N * 8, stream);
    HIP_OK(hipStreamSynchronize(stream));
    adj_t = timestep;
    sens_valid = adj_valid = true;
  }

  // AdjointLBSolver::setCurrentTimestep (src/adjointLatticeBoltzmann/adjointLBSolver.cc:147-151)
  void adjoint_set_timestep(int timestep) {
    if (adj_t < 0) throw Error{LBMX_ERR_STATE, "adjoint not reset (call lbmx_adjoint_reset)"};
    if (timestep < 0) throw Error{LBMX_ERR_ARG, "timestep must be >= 0"};
    adj_t = timestep;
  }

  // One reverse step t = adj_t (adjointCollide, computeCollideSource, adjointStream:
  // src/adjointLatticeBoltzmann/adjointLBSolver.cc:28-102, applications/RobustPump2d/robustPump2d.cc:268-276)
  // on the current forward state s_t: recomputed from p_{t-1} after lbmx_forward, or read as
  // the AoS state given to lbmx_set_state.
  void adjoint_step(int with_source) {
    if (adj_t < 0) throw Error{LBMX_ERR_STATE, "adjoint not reset (call lbmx_adjoint_reset)"};
    if (adj_t < 1) throw Error{LBMX_ERR_STATE, "the reverse loop ends at step 1 (step 0 is the initial state)"};
    require_state();
    if (t != adj_t)
      throw Error{LBMX_ERR_STATE, "forward state is at step " + std::to_string(t) + ", the reverse step is " +
                                      std::to_string(adj_t)};
    if (state_kind != 0 && state_kind != 2) throw Error{LBMX_ERR_STATE, "no forward state at this step"};
    check_time_tables(adj_t);
    prepare();
    ObjArgs oa = obj;
    oa.acc = nullptr;
    if (!with_source || !obj_set) oa.nterm = 0;
    AdjArgs a = adj_args(state_kind == 0 ? prev : nullptr, lam_a.p, lam_b.p, adj_t, oa);
    if (state_kind == 2) a.fwd_aos = static_cast<const double*>(raw_aos.p);
    HIP_OK(run_adj(a));
    std::swap(lam_a.p, lam_b.p);
    HIP_OK(hipStreamSynchronize(stream));
    stats.adj_sweeps += 1;
    adj_t -= 1;
  }

  void get_sensitivity(double* s, size_t n) {
    if (!sens_valid) throw Error{LBMX_ERR_STATE, "no gradient computed"};
    if ((long long)n != N) throw Error{LBMX_ERR_ARG, "sensitivity length != nodes"};
    HIP_OK(hipMemcpy(s, sens_dev.p, n * 8, hipMemcpyDeviceToHost));
  }

  void get_adjoint(double* out, size_t n) {
    if (!adj_valid) throw Error{LBMX_ERR_STATE, "no adjoint computed by this solver (a batch keeps its own)"};
    if ((long long)n != N * Q) throw Error{LBMX_ERR_ARG, "adjoint length != nodes * Q"};
    const size_t eb = prec_lam_bytes(d.precision);
    std::vector<char> h(lam_bytes());
    HIP_OK(hipMemcpy(h.data(), lam_a.p, h.size(), hipMemcpyDeviceToHost));
    auto val = [&](long long e) -> double {
      return eb == 8 ? reinterpret_cast<const double*>(h.data())[e] : reinterpret_cast<const float*>(h.data())[e];
    };
    const int* ex = lat_ex();
    const int* ey = lat_ey();
    const int* ez = lat_ez();
    const int* op = lat_opp();
    for (int z = 0; z < d.nz; ++z)
      for (int y = 0; y < d.ny; ++y)
        for (int x = 0; x < d.nx; ++x) {
          const long long node = ((long long)z * d.ny + y) * d.nx + x;
          const long long off = ((long long)(z + ll.oz) * ll.py + (y + 1)) * ll.px + ll.ax + x;
          for (int q = 0; q < Q; ++q) {
            const int xo = x + ex[q], yo = y + ey[q], zo = z + ez[q];
            const bool out_dom = (ex[q] || ey[q] || ez[q]) &&
                                 (xo < 0 || xo >= d.nx || yo < 0 || yo >= d.ny || zo < 0 || zo >= d.nz);
            // the buffer holds m_1 at own nodes: nu_0(x)[q] = m_1(x + c_q)[q], or the mid-way
            // reflection m_1(x)[opp q] when x + c_q is outside
            const long long e = out_dom ? (long long)op[q] * ll.plane + off
                                        : (long long)q * ll.plane + off +
                                              (ex[q] + (long long)ey[q] * ll.px + (long long)ez[q] * ll.px * ll.py);
            out[node * Q + q] = val(e);
          }
        }
  }

  // lattice velocity tables on the host (same order as the device lattices)
  const int* lat_ex() const {
    static const int d2[14] = {0, 1, 0, -1, -1, -1, 0, 1, 1, 0, 0, -1, 0, 1};
    static const int d3[19] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
    return D == 3 ? d3 : d2;
  }
  const int* lat_ey() const {
    static const int d2[14] = {0, -1, -1, -1, 0, 1, 1, 1, 0, 0, -1, 0, 1, 0};
    static const int d3[19] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
    return D == 3 ? d3 : d2;
  }
  const int* lat_ez() const {
    static const int d2[14] = {0};
    static const int d3[19] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
    return D == 3 ? d3 : d2;
  }
  const int* lat_opp() const {
    static const int d2[14] = {0, 5, 6, 7, 8, 1, 2, 3, 4, 9, 12, 13, 10, 11};
    static const int d3[19] = {0, 10, 11, 12, 13, 14, 15, 16, 17, 18, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    return D == 3 ? d3 : d2;
  }

  double time_sweeps(int kind, int nsweeps) {
    require_state();
    if (nsweeps < 1) throw Error{LBMX_ERR_ARG, "nsweeps must be >= 1"};
    if (kind == 1 && (!inc || d.design == LBMX_DESIGN_DARCY))
      throw Error{LBMX_ERR_SUP, "no adjoint sweep for this operator"};
    check_time_tables(kind == 0 ? (long long)t + nsweeps : (long long)t);
    prepare();
    EventPair ev;
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    float ms = 0;
    if (kind == 0) {
      const bool with_obj = obj_set && obj.nterm > 0;
      if (with_obj) zero_acc();
      HIP_OK(hipEventRecord(e0, stream));
      for (int k = 0; k < nsweeps; ++k) {
        FwdArgs a = fwd_args(cur, prev, t + 1, with_obj);
        HIP_OK(run_fwd(a));
        std::swap(cur, prev);
        ++t;
      }
      HIP_OK(hipEventRecord(e1, stream));
      state_kind = 0;
      stats.fwd_sweeps += nsweeps;
    } else {
      lam_a.ensure(lam_bytes());
      lam_b.ensure(lam_bytes());
      sens_dev.ensure((size_t)N * 8);
      zero_async(lam_a.p, lam_bytes(), stream);
      zero_async(lam_b.p, lam_bytes(), stream);
      zero_async(sens_dev.p, (size_t)N * 8, stream);
      ObjArgs oa = obj;
      oa.acc = nullptr;
      if (!obj_set) oa.nterm = 0;
      void* lin = lam_a.p;
      void* lout = lam_b.p;
      HIP_OK(hipEventRecord(e0, stream));
      for (int k = 0; k < nsweeps; ++k) {
        AdjArgs a = adj_args(cur, lin, lout, std::max(1, t), oa);
        HIP_OK(run_adj(a));
        std::swap(lin, lout);
      }
      HIP_OK(hipEventRecord(e1, stream));
      stats.adj_sweeps += nsweeps;
    }
    HIP_OK(hipEventSynchronize(e1));
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    if (kind == 0) stats.fwd_kernel_ms = ms; else stats.adj_kernel_ms = ms;
    return ms;
  }

  size_t device_bytes() const {
    size_t b = gamma64_dev.bytes + gamma_dev.bytes + bc_dev.bytes + acc_dev.bytes + red_dev.bytes + sens_dev.bytes + lam_a.bytes +
               lam_b.bytes + raw_aos.bytes + nside_dev.bytes + bcrow_dev.bytes;
    for (auto& p : pool) b += p->bytes;
    return b;
  }
};

#define LBMX_BATCH_PREC(LAT, CALL)                  \
  switch (d.precision) {                             \
    case LBMX_FP64: return CALL(LAT##_f64);          \
    case LBMX_FP32: return CALL(LAT##_f32);          \
    case LBMX_FP16_SHIFTED: return CALL(LAT##_f16);  \
    default: return CALL(LAT##_bf16);                \
  }

// ---- batch: several lattices of one configuration advanced in lockstep ----
// Every sweep of a design-gradient pass is one launch over all members (the realisations and
// design candidates that the reference spreads over MPISplit communicators,
// src/core/MPISplit.cc:150-234). All members share the checkpoint schedule, so a checkpoint
// slot is one batch buffer holding every member's state at its own offset.
static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Batch {
  std::vector<Engine*> m;  // members (owned by their handles)
  hipStream_t stream = nullptr;
  int ty = kTileY;         // rows of the workgroup tile
  bool pair = true;        // forward of a 2- or 4-row tile: two nodes per thread
  int nslots = 64;
  int ckpt_mode = LBMX_CHECKPOINT_DYNAMIC;
  int nblocks = 0;
  int tf_stride = 1;
  size_t state_total = 0, lam_total = 0;
  std::vector<BatchInst> inst;
  std::vector<int> blk_inst;
  DeviceBuffer inst_dev, blk_dev, tfac_dev, lam_a, lam_b;
  std::vector<std::unique_ptr<DeviceBuffer>> pool;
  lbmx_stats stats{};
  std::vector<hipStream_t> member_streams;  // restored when the batch goes away
  bool use_graph = false;                   // replay the captured gradient pass
  hipGraphExec_t gexec = nullptr;
  std::vector<long long> gkey;              // what the captured graph was built for
  long long graph_recompute = 0;

  explicit Batch(std::vector<Engine*> members) : m(std::move(members)) {
    if (m.empty()) throw Error{LBMX_ERR_ARG, "empty batch"};
    const lbmx_desc& d0 = m[0]->d;
    for (size_t i = 0; i < m.size(); ++i) {
      const lbmx_desc& d = m[i]->d;
      if (d.lattice != LBMX_D2Q9 && d.lattice != LBMX_D3Q19 && d.lattice != LBMX_D2Q9_D2Q5)
        throw Error{LBMX_ERR_SUP, "unknown lattice"};
      if (d.lattice != d0.lattice) throw Error{LBMX_ERR_ARG, "batch members must share the lattice"};
      if (m[i]->nside_count > 0) throw Error{LBMX_ERR_SUP, "batched sweeps do not take Zou-He Neumann pieces"};
      if (d.collision != d0.collision || d.design != d0.design || d.precision != d0.precision)
        throw Error{LBMX_ERR_ARG, "batch members must share operator, design coupling and precision"};
      for (size_t k = 0; k < i; ++k)
        if (m[k] == m[i]) throw Error{LBMX_ERR_ARG, "a solver appears twice in the batch"};
    }
    if (d0.design == LBMX_DESIGN_DARCY) throw Error{LBMX_ERR_SUP, "DarcyMRT2d has no reverse operator"};
    if (d0.collision != LBMX_BGK && d0.collision != LBMX_MRT && d0.collision != LBMX_CASCADED)
      throw Error{LBMX_ERR_SUP, "batched sweeps are built for incompressible BGK, MRT and cascaded"};
    if (d0.lattice != LBMX_D2Q9 && d0.collision == LBMX_CASCADED)
      throw Error{LBMX_ERR_SUP, "the cascaded operator is defined for D2Q9"};
    // a stream of its own (default priority): graph capture needs a non-null stream
    HIP_OK(hipStreamCreate(&stream));
    for (Engine* e : m) member_streams.push_back(e->stream);
  }

  ~Batch() {
    (void)hipStreamSynchronize(stream);
    drop_graph();
    for (size_t i = 0; i < m.size(); ++i) m[i]->stream = member_streams[i];
    (void)hipStreamDestroy(stream);
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gkey.clear();
  }

  // every address and size the captured launch sequence depends on
  std::vector<long long> graph_key(int nsteps, int slots, bool naive, double rho0, const double* u0, double T0) const {
    std::vector<long long> k = {nsteps, slots, naive ? 1 : 0, ckpt_mode, ty, pair ? 1 : 0, nblocks, tf_stride};
    auto bits = [&](double v) {
      long long b;
      std::memcpy(&b, &v, sizeof b);
      k.push_back(b);
    };
    bits(rho0); bits(T0);  // the initial state is part of the captured launches
    for (int a = 0; a < 3; ++a) bits(u0 ? u0[a] : 0.0);
    auto ptr = [&](const void* p) { k.push_back((long long)(uintptr_t)p); };
    for (auto& p : pool) ptr(p->p);
    ptr(lam_a.p); ptr(lam_b.p); ptr(inst_dev.p); ptr(blk_dev.p); ptr(tfac_dev.p);
    for (size_t i = 0; i < m.size(); ++i) {
      ptr(inst[i].obj.acc); ptr(inst[i].sens); ptr(inst[i].gamma); ptr(inst[i].bcs);
      k.push_back(inst[i].state_off);
      k.push_back(inst[i].lam_off);
    }
    return k;
  }

  Geometry geom_for(const Engine& e) const {
    Geometry g{};
    g.nx = e.d.nx; g.ny = e.d.ny; g.nz = e.d.nz;
    g.nblk_x = (e.d.nx + kTileX - 1) / kTileX;
    g.nblk_y = (e.d.ny + ty - 1) / ty;
    g.nblocks = g.nblk_x * g.nblk_y * e.d.nz;
    g.N = e.N;
    return g;
  }

  // instance table, workgroup map and member device arrays for the current tile height
  void build() {
    inst.clear();
    blk_inst.clear();
    nblocks = 0;
    tf_stride = 1;
    size_t st = 0, lt = 0;
    for (size_t i = 0; i < m.size(); ++i) {
      Engine& e = *m[i];
      e.stream = stream;
      e.prepare_static();
      const Geometry g = geom_for(e);
      e.acc_dev.ensure(e.nseg() * sizeof(double));
      e.sens_dev.ensure((size_t)e.N * 8);
      BatchInst I{};
      I.g = g;
      I.ls = e.ls;
      I.ll = e.ll;
      I.m = e.params();
      I.obj = e.obj;
      if (!e.obj_set) I.obj.nterm = 0;
      I.obj.acc = I.obj.nterm > 0 ? static_cast<double*>(e.acc_dev.p) : nullptr;
      I.gamma = e.gamma_dev.p;
      I.sens = static_cast<double*>(e.sens_dev.p);
      I.bcs = static_cast<const DevBC*>(e.bc_dev.p);
      I.bcrows = static_cast<const unsigned long long*>(e.bcrow_dev.p);
      I.nbc = (int)e.bcs.size();
      I.blk_begin = nblocks;
      I.state_off = (long long)st;
      I.lam_off = (long long)lt;
      st += align256(e.state_bytes());
      lt += align256(e.lam_bytes());
      for (int b = 0; b < g.nblocks; ++b) blk_inst.push_back((int)i);
      nblocks += g.nblocks;
      tf_stride = std::max(tf_stride, I.nbc);
      inst.push_back(I);
    }
    if (st != state_total) pool.clear();
    state_total = st;
    lam_total = lt;
    inst_dev.ensure(inst.size() * sizeof(BatchInst));
    HIP_OK(hipMemcpyAsync(inst_dev.p, inst.data(), inst.size() * sizeof(BatchInst), hipMemcpyHostToDevice, stream));
    blk_dev.ensure(blk_inst.size() * sizeof(int));
    HIP_OK(hipMemcpyAsync(blk_dev.p, blk_inst.data(), blk_inst.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    HIP_OK(hipStreamSynchronize(stream));
  }

  void ensure_pool(size_t count) {
    while (pool.size() < count) {
      pool.emplace_back(new DeviceBuffer());
      pool.back()->ensure(state_total);
    }
  }

  // time factors of every member's pieces for t = 0..nsteps
  void upload_tfac(int nsteps) {
    const size_t per = (size_t)m.size() * tf_stride;
    std::vector<double> h((size_t)(nsteps + 1) * per, 1.0);
    for (int t = 0; t <= nsteps; ++t)
      for (size_t i = 0; i < m.size(); ++i) m[i]->time_factors(t, &h[t * per + i * tf_stride]);
    tfac_dev.ensure(h.size() * sizeof(double));
    HIP_OK(hipMemcpyAsync(tfac_dev.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HIP_OK(hipStreamSynchronize(stream));
  }

  hipError_t launch(int kind, const void* src, void* dst, const void* lin, void* lout, int t, bool with_obj) {
    BatchStep st{};
    st.src = src;
    st.dst = dst;
    st.lam_in = lin;
    st.lam_out = lout;
    st.inst = static_cast<const BatchInst*>(inst_dev.p);
    st.blk_inst = static_cast<const int*>(blk_dev.p);
    st.tfac = static_cast<const double*>(tfac_dev.p) + (size_t)t * m.size() * tf_stride;
    st.tf_stride = tf_stride;
    st.nblocks = nblocks;
    st.t = t;
    st.with_obj = with_obj ? 1 : 0;
    const lbmx_desc& d = m[0]->d;
#define CALL(S) launch_batch_##S(d.collision, d.design, ty, pair, kind, st, stream)
    switch (d.lattice) {
      case LBMX_D2Q9: LBMX_BATCH_PREC(d2q9, CALL)
      case LBMX_D3Q19: LBMX_BATCH_PREC(d3q19, CALL)
      default: LBMX_BATCH_PREC(thermal, CALL)
    }
#undef CALL
  }

  void init_all(void* dst, double rho0, const double* u0, double T0) {
    for (size_t i = 0; i < m.size(); ++i) {
      Engine& e = *m[i];
      e.init_rho = rho0;
      for (int k = 0; k < 3; ++k) e.init_u[k] = u0 ? u0[k] : 0.0;
      e.init_T = T0;
      FwdArgs a = e.fwd_args(nullptr, static_cast<char*>(dst) + inst[i].state_off, 0, false);
      HIP_OK(e.run_init(a, nullptr));
    }
  }

  void zero_accs() {
    for (size_t i = 0; i < m.size(); ++i)
      if (inst[i].obj.acc) zero_async(inst[i].obj.acc, m[i]->nseg() * sizeof(double), stream);
  }

  double reduce_acc(size_t i) {
    if (!inst[i].obj.acc) return 0.0;
    std::vector<double> h(m[i]->nseg());
    HIP_OK(hipMemcpyAsync(h.data(), inst[i].obj.acc, h.size() * 8, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    double s = 0;
    for (double v : h) s += v;  // fixed order: deterministic
    return s;
  }

  void gradient(int nsteps, double rho0, const double* u0, double T0, double* J, double* const* sens) {
    if (nsteps < 1) throw Error{LBMX_ERR_ARG, "nsteps must be >= 1"};
    for (Engine* e : m) e->check_time_tables(nsteps);
    build();
    const bool naive = nslots <= 0;
    const int slots = naive ? nsteps : std::min(nslots, std::max(2, nsteps));
    ensure_pool(2 + slots);
    std::vector<void*> slotbuf(slots);
    for (int k = 0; k < slots; ++k) slotbuf[k] = pool[2 + k]->p;
    lam_a.ensure(lam_total);
    lam_b.ensure(lam_total);
    upload_tfac(nsteps);

    struct Ops : SweepOps {
      Batch& b;
      double rho0, T0;
      const double* u0;
      bool capturing = false;
      bool reverse_started = false;
      std::chrono::steady_clock::time_point t_fwd_end;
      Ops(Batch& bb, double r, const double* u, double T) : b(bb), rho0(r), T0(T), u0(u) {}
      void init(void* dst) override {
        b.init_all(dst, rho0, u0, T0);
        b.zero_accs();
      }
      void fwd(const void* src, void* dst, int t, bool with) override {
        HIP_OK(b.launch(0, src, dst, nullptr, nullptr, t, with));
      }
      void adj(const void* fsrc, const void* lin, void* lout, int t) override {
        if (!reverse_started) {
          reverse_started = true;
          if (!capturing) HIP_OK(hipStreamSynchronize(b.stream));  // end of the forward phase
          t_fwd_end = std::chrono::steady_clock::now();
          zero_async(b.lam_a.p, b.lam_total, b.stream);
          for (auto& I : b.inst) zero_async(I.sens, (size_t)I.g.N * 8, b.stream);
        }
        HIP_OK(b.launch(1, fsrc, nullptr, lin, lout, t, true));
      }
    } ops(*this, rho0, u0, T0);

    auto tf0 = std::chrono::steady_clock::now();
    void* lin = lam_a.p;
    void* lout = lam_b.p;
    PassResult pr;
    if (!use_graph) {
      pr = checkpointed_pass(ops, nsteps, naive, slotbuf, pool[0]->p, pool[1]->p, lin, lout, true, ckpt_mode);
    } else {
      // the launch sequence of a pass is fixed by (steps, slots, tile height) and the buffer
      // addresses: capture it once, replay it on later calls
      std::vector<long long> key = graph_key(nsteps, slots, naive, rho0, u0, T0);
      if (!gexec || key != gkey) {
        drop_graph();
        ops.capturing = true;
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipGraph_t graph = nullptr;
        try {
          pr = checkpointed_pass(ops, nsteps, naive, slotbuf, pool[0]->p, pool[1]->p, lin, lout, true, ckpt_mode);
        } catch (...) {
          (void)hipStreamEndCapture(stream, &graph);
          if (graph) (void)hipGraphDestroy(graph);
          throw;
        }
        HIP_OK(hipStreamEndCapture(stream, &graph));
        hipError_t ie = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
          gexec = nullptr;
          throw Error{LBMX_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie)};
        }
        gkey = key;
        graph_recompute = pr.recompute;
      }
      pr.recompute = graph_recompute;
      tf0 = std::chrono::steady_clock::now();
      HIP_OK(hipGraphLaunch(gexec, stream));
      ops.t_fwd_end = tf0;  // phases are not separated inside a replay
    }
    HIP_OK(hipStreamSynchronize(stream));
    auto tr1 = std::chrono::steady_clock::now();
    long long nodes = 0;
    for (size_t i = 0; i < m.size(); ++i) {
      if (J) J[i] = reduce_acc(i);
      if (sens && sens[i]) HIP_OK(hipMemcpy(sens[i], inst[i].sens, (size_t)m[i]->N * 8, hipMemcpyDeviceToHost));
      m[i]->sens_valid = true;    // the batch accumulated into the member's own dJ/dgamma array
      m[i]->adj_valid = false;    // but the adjoint lives in the batch's buffers
      m[i]->state_kind = -1;
      nodes += m[i]->N;
    }
    (void)nodes;
    stats.fwd_sweeps += nsteps + pr.recompute;
    stats.adj_sweeps += nsteps;
    stats.recompute_sweeps += pr.recompute;
    stats.forward_phase_ms = std::chrono::duration<double, std::milli>(ops.t_fwd_end - tf0).count();
    stats.reverse_phase_ms = std::chrono::duration<double, std::milli>(tr1 - ops.t_fwd_end).count();
  }

  // Times nsweeps batched forward and adjoint sweeps for each tile shape (height 1, 2 or 4;
  // the forward of 2 and 4 rows paired or not) and keeps the fastest. ms_out[h] (heights 1, 2,
  // 4) receives the better pairing of each height.
  int autotune(int nsweeps, double* ms_out) {
    if (nsweeps < 1) throw Error{LBMX_ERR_ARG, "nsweeps must be >= 1"};
    const int cand_ty[5] = {1, 2, 4, 2, 4};
    const bool cand_pair[5] = {true, true, true, false, false};
    double best = 1e300;
    int best_ty = ty;
    bool best_pair = pair;
    if (ms_out)
      for (int h = 0; h < 3; ++h) ms_out[h] = 1e300;
    EventPair ev;
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    for (int c = 0; c < 5; ++c) {
      ty = cand_ty[c];
      pair = cand_pair[c];
      build();
      ensure_pool(2);
      lam_a.ensure(lam_total);
      lam_b.ensure(lam_total);
      upload_tfac(1);
      const double u0[3] = {0, 0, 0};
      init_all(pool[0]->p, 1.0, u0, 0.0);
      zero_async(lam_a.p, lam_total, stream);
      zero_async(lam_b.p, lam_total, stream);
      for (auto& I : inst) zero_async(I.sens, (size_t)I.g.N * 8, stream);
      void* a = pool[0]->p;
      void* b = pool[1]->p;
      HIP_OK(launch(0, a, b, nullptr, nullptr, 1, false));  // warm-up
      HIP_OK(hipEventRecord(e0, stream));
      for (int k = 0; k < nsweeps; ++k) {
        HIP_OK(launch(0, a, b, nullptr, nullptr, 1, false));
        std::swap(a, b);
      }
      void* li = lam_a.p;
      void* lo = lam_b.p;
      for (int k = 0; k < nsweeps; ++k) {
        HIP_OK(launch(1, a, nullptr, li, lo, 1, true));
        std::swap(li, lo);
      }
      HIP_OK(hipEventRecord(e1, stream));
      HIP_OK(hipEventSynchronize(e1));
      float ms = 0;
      HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      const int h = ty == 1 ? 0 : ty == 2 ? 1 : 2;
      if (ms_out) ms_out[h] = std::min(ms_out[h], double(ms) / nsweeps);
      if (ms < best) {
        best = ms;
        best_ty = ty;
        best_pair = pair;
      }
    }
    ty = best_ty;
    pair = best_pair;
    // the timed sweeps overwrote every member's state and accumulated into its dJ/dgamma array
    for (Engine* e : m) {
      e->state_kind = -1;
      e->sens_valid = false;
      e->adj_valid = false;
    }
    return ty;
  }

  size_t device_bytes() const {
    size_t b = inst_dev.bytes + blk_dev.bytes + tfac_dev.bytes + lam_a.bytes + lam_b.bytes;
    for (auto& p : pool) b += p->bytes;
    return b;
  }
};

}  // namespace lbmx

using lbmx::Engine;
using lbmx::Error;

struct lbmx_solver {
  std::unique_ptr<Engine> e;
};

struct lbmx_batch {
  std::unique_ptr<lbmx::Batch> b;
};

using lbmx::guarded;

namespace lbmx {

static Engine& engine_of(lbmx_solver* s) {
  if (!s || !s->e) throw Error{LBMX_ERR_ARG, "null solver handle"};
  return *s->e;
}

hipStream_t solver_stream(lbmx_solver* s) { return engine_of(s).stream; }

int solver_lattice(lbmx_solver* s) { return engine_of(s).d.lattice; }

long long solver_nodes(lbmx_solver* s, int dims[3]) {
  Engine& e = engine_of(s);
  dims[0] = e.d.nx;
  dims[1] = e.d.ny;
  dims[2] = e.d.nz;
  return e.N;
}

void solver_set_design_device(lbmx_solver* s, const double* dev) { engine_of(s).set_design_device(dev); }

const double* solver_sensitivity_device(lbmx_solver* s) {
  Engine& e = engine_of(s);
  if (!e.sens_valid) throw Error{LBMX_ERR_STATE, "no gradient computed"};
  return static_cast<const double*>(e.sens_dev.p);
}

}  // namespace lbmx

#define CHECK_HANDLE(s)                                      \
  if (!(s) || !(s)->e) {                                     \
    lbmx::g_last_error = "null solver handle";               \
    return LBMX_ERR_ARG;                                     \
  }

extern "C" {

lbmx_status lbmx_create(const lbmx_desc* desc, lbmx_solver** out) {
  if (!desc || !out) {
    lbmx::g_last_error = "null argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    auto* s = new lbmx_solver;
    try {
      s->e.reset(new Engine(*desc));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

lbmx_status lbmx_destroy(lbmx_solver* s) {
  if (!s) return LBMX_OK;
  delete s;
  return LBMX_OK;
}

lbmx_status lbmx_set_stream(lbmx_solver* s, void* stream) {
  CHECK_HANDLE(s);
  s->e->stream = static_cast<hipStream_t>(stream);
  return LBMX_OK;
}

const char* lbmx_last_error(void) { return lbmx::g_last_error.c_str(); }

lbmx_status lbmx_add_boundary(lbmx_solver* s, const lbmx_box* box, const lbmx_bc_desc* bc) {
  CHECK_HANDLE(s);
  if (!box || !bc) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->add_boundary(*box, *bc); });
}

lbmx_status lbmx_add_boundary_table(lbmx_solver* s, const lbmx_box* box, const lbmx_bc_desc* bc, const double* profile,
                                    size_t nprofile, int ncomp, const double* time_factors, size_t nt) {
  CHECK_HANDLE(s);
  if (!box || !bc) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->add_boundary_table(*box, *bc, profile, nprofile, ncomp, time_factors, nt); });
}

lbmx_status lbmx_set_design(lbmx_solver* s, const double* gamma, size_t n) {
  CHECK_HANDLE(s);
  if (!gamma) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->set_design(gamma, n); });
}

lbmx_status lbmx_get_design(lbmx_solver* s, double* gamma, size_t n) {
  CHECK_HANDLE(s);
  if (!gamma) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->get_design(gamma, n); });
}

lbmx_status lbmx_set_objective(lbmx_solver* s, const lbmx_objective* obj) {
  CHECK_HANDLE(s);
  if (!obj) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->set_objective(*obj); });
}
lbmx_status lbmx_set_objective_terms(lbmx_solver* s, const lbmx_term* terms, int nterms, double time_const) {
  CHECK_HANDLE(s);
  if (!terms) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->set_objective_terms(terms, nterms, time_const); });
}

lbmx_status lbmx_init_equilibrium(lbmx_solver* s, double rho0, const double u0[3], double T0) {
  CHECK_HANDLE(s);
  return guarded([&] { s->e->init_equilibrium(rho0, u0, T0, nullptr, 0); });
}

lbmx_status lbmx_init_equilibrium_field(lbmx_solver* s, const double* macros, size_t n) {
  CHECK_HANDLE(s);
  if (!macros) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->init_equilibrium(1.0, nullptr, 0.0, macros, n); });
}

lbmx_status lbmx_forward(lbmx_solver* s, int nsteps, double* J) {
  CHECK_HANDLE(s);
  return guarded([&] { s->e->forward(nsteps, J); });
}

lbmx_status lbmx_get_state(lbmx_solver* s, double* f, size_t n) {
  CHECK_HANDLE(s);
  if (!f) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->get_state(f, n); });
}

lbmx_status lbmx_set_state(lbmx_solver* s, const double* f, size_t n, int timestep) {
  CHECK_HANDLE(s);
  if (!f) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->set_state(f, n, timestep); });
}

lbmx_status lbmx_get_macros(lbmx_solver* s, double* macros, size_t n) {
  CHECK_HANDLE(s);
  if (!macros) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    Engine& e = *s->e;
    if ((long long)n != 5 * e.N) throw Error{LBMX_ERR_ARG, "macro array must hold 5 values per node"};
    std::vector<double> f((size_t)e.N * e.Q);
    e.get_state(f.data(), f.size());
    static const int ex9[9] = {0, 1, 0, -1, -1, -1, 0, 1, 1}, ey9[9] = {0, -1, -1, -1, 0, 1, 1, 1, 0};
    static const int ex19[19] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
    static const int ey19[19] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
    static const int ez19[19] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
    const int QF = e.D == 3 ? 19 : 9;
    for (long long node = 0; node < e.N; ++node) {
      const double* fn = &f[node * e.Q];
      double r = 0, jx = 0, jy = 0, jz = 0, T = 0;
      for (int q = 0; q < QF; ++q) {
        r += fn[q];
        jx += (e.D == 3 ? ex19[q] : ex9[q]) * fn[q];
        jy += (e.D == 3 ? ey19[q] : ey9[q]) * fn[q];
        if (e.D == 3) jz += ez19[q] * fn[q];
      }
      for (int q = QF; q < e.Q; ++q) T += fn[q];
      if (!e.inc) { jx /= r; jy /= r; jz /= r; }
      macros[5 * node] = r; macros[5 * node + 1] = jx; macros[5 * node + 2] = jy;
      macros[5 * node + 3] = jz; macros[5 * node + 4] = T;
    }
  });
}

lbmx_status lbmx_get_timestep(lbmx_solver* s, int* t) {
  CHECK_HANDLE(s);
  if (!t) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  *t = s->e->t;
  return LBMX_OK;
}

lbmx_status lbmx_checkpoint_config(lbmx_solver* s, int nslots) {
  CHECK_HANDLE(s);
  if (nslots == 1) { lbmx::g_last_error = "dynamic checkpointing needs at least 2 slots"; return LBMX_ERR_ARG; }
  s->e->nslots = nslots;
  return LBMX_OK;
}

lbmx_status lbmx_gradient(lbmx_solver* s, int nsteps, double rho0, const double u0[3], double T0, double* J,
                          double* sens, size_t n) {
  CHECK_HANDLE(s);
  return guarded([&] { s->e->gradient(nsteps, rho0, u0, T0, J, sens, n); });
}

lbmx_status lbmx_adjoint_reset(lbmx_solver* s, int timestep) {
  CHECK_HANDLE(s);
  return guarded([&] { s->e->adjoint_reset(timestep); });
}

lbmx_status lbmx_adjoint_set_timestep(lbmx_solver* s, int timestep) {
  CHECK_HANDLE(s);
  return guarded([&] { s->e->adjoint_set_timestep(timestep); });
}

lbmx_status lbmx_adjoint_step(lbmx_solver* s, int with_source) {
  CHECK_HANDLE(s);
  return guarded([&] { s->e->adjoint_step(with_source); });
}

lbmx_status lbmx_set_design_device(lbmx_solver* s, const double* gamma_dev, size_t n) {
  CHECK_HANDLE(s);
  if (!gamma_dev) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    lbmx::Engine& e = *s->e;
    if ((long long)n != e.N) throw lbmx::Error{LBMX_ERR_ARG, "design vector length != number of nodes"};
    e.set_design_device(gamma_dev);
    HIP_OK(hipStreamSynchronize(e.stream));
  });
}

lbmx_status lbmx_get_sensitivity_device(lbmx_solver* s, const double** sens_dev, size_t* n) {
  CHECK_HANDLE(s);
  if (!sens_dev) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    *sens_dev = lbmx::solver_sensitivity_device(s);
    if (n) *n = (size_t)s->e->N;
  });
}

lbmx_status lbmx_get_sensitivity(lbmx_solver* s, double* sens, size_t n) {
  CHECK_HANDLE(s);
  if (!sens) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->get_sensitivity(sens, n); });
}

lbmx_status lbmx_get_adjoint(lbmx_solver* s, double* adj, size_t n) {
  CHECK_HANDLE(s);
  if (!adj) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { s->e->get_adjoint(adj, n); });
}

lbmx_status lbmx_describe_boundaries(lbmx_solver* s, int* out, int max_pieces, int* count) {
  CHECK_HANDLE(s);
  if (!count) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  const auto& b = s->e->bcs;
  *count = (int)b.size();
  for (int k = 0; k < (int)b.size() && k < max_pieces && out; ++k) {
    int* o = out + 10 * k;
    o[0] = b[k].type;
    for (int c = 0; c < 3; ++c) o[1 + c] = b[k].nrm[c];
    for (int c = 0; c < 6; ++c) o[4 + c] = b[k].box[c];
  }
  return LBMX_OK;
}

// Runs the gradient's checkpointed pass (the same code as lbmx_gradient) on labelled host
// buffers: a buffer holds the step of the state it contains; every forward sweep must read
// p_{t-1} and every adjoint sweep must receive p_{t-1}.
lbmx_status lbmx_checkpoint_mode(lbmx_solver* s, int mode) {
  CHECK_HANDLE(s);
  if (mode != LBMX_CHECKPOINT_DYNAMIC && mode != LBMX_CHECKPOINT_STATIC && mode != LBMX_CHECKPOINT_BINOMIAL) {
    lbmx::g_last_error = "unknown checkpoint mode";
    return LBMX_ERR_ARG;
  }
  s->e->ckpt_mode = mode;
  return LBMX_OK;
}

lbmx_status lbmx_checkpoint_simulate(int nslots, int nsteps, long long* advances, int* max_slots) {
  return lbmx_checkpoint_simulate_mode(nslots, nsteps, LBMX_CHECKPOINT_DYNAMIC, advances, max_slots);
}

lbmx_status lbmx_checkpoint_simulate_mode(int nslots, int nsteps, int mode, long long* advances, int* max_slots) {
  if (mode != LBMX_CHECKPOINT_DYNAMIC && mode != LBMX_CHECKPOINT_STATIC && mode != LBMX_CHECKPOINT_BINOMIAL) {
    lbmx::g_last_error = "unknown checkpoint mode";
    return LBMX_ERR_ARG;
  }
  if (!advances || !max_slots || nsteps < 1 || nslots < 2) {
    lbmx::g_last_error = "bad argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    struct Labels : lbmx::SweepOps {
      std::unordered_map<const void*, int> label;
      int get(const void* p) const {
        auto it = label.find(p);
        return it == label.end() ? -1 : it->second;
      }
      void init(void* dst) override { label[dst] = 0; }
      void fwd(const void* src, void* dst, int t, bool) override {
        if (get(src) != t - 1) throw Error{LBMX_ERR_STATE, "forward sweep to step " + std::to_string(t) + " read a wrong state"};
        label[dst] = t;
      }
      void adj(const void* fsrc, const void*, void*, int t) override {
        if (get(fsrc) != t - 1) throw Error{LBMX_ERR_STATE, "delivered the wrong state at step " + std::to_string(t)};
      }
    } ops;
    const int slots = std::min(nslots, std::max(2, nsteps));
    std::vector<char> storage(slots + 4);  // distinct addresses stand for buffers
    std::vector<void*> slotbuf(slots);
    for (int k = 0; k < slots; ++k) slotbuf[k] = &storage[k];
    void* lin = &storage[slots + 2];
    void* lout = &storage[slots + 3];
    lbmx::PassResult r =
        lbmx::checkpointed_pass(ops, nsteps, false, slotbuf, &storage[slots], &storage[slots + 1], lin, lout, false,
                                mode);
    *advances = r.recompute;
    *max_slots = r.max_slots;
  });
}

lbmx_status lbmx_batch_create(lbmx_solver* const* members, int count, lbmx_batch** out) {
  if (!members || !out || count < 1) {
    lbmx::g_last_error = "bad argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    std::vector<Engine*> v;
    for (int i = 0; i < count; ++i) {
      if (!members[i] || !members[i]->e) throw Error{LBMX_ERR_ARG, "null solver handle in batch"};
      v.push_back(members[i]->e.get());
    }
    auto* b = new lbmx_batch;
    try {
      b->b.reset(new lbmx::Batch(std::move(v)));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

lbmx_status lbmx_batch_destroy(lbmx_batch* b) {
  delete b;
  return LBMX_OK;
}

#define CHECK_BATCH(b)                                       \
  if (!(b) || !(b)->b) {                                     \
    lbmx::g_last_error = "null batch handle";                \
    return LBMX_ERR_ARG;                                     \
  }

lbmx_status lbmx_batch_checkpoint_mode(lbmx_batch* b, int mode) {
  if (!b || !b->b) { lbmx::g_last_error = "null batch handle"; return LBMX_ERR_ARG; }
  if (mode != LBMX_CHECKPOINT_DYNAMIC && mode != LBMX_CHECKPOINT_STATIC && mode != LBMX_CHECKPOINT_BINOMIAL) {
    lbmx::g_last_error = "unknown checkpoint mode";
    return LBMX_ERR_ARG;
  }
  b->b->ckpt_mode = mode;
  return LBMX_OK;
}

lbmx_status lbmx_batch_checkpoint_config(lbmx_batch* b, int nslots) {
  CHECK_BATCH(b);
  b->b->nslots = nslots;
  return LBMX_OK;
}

lbmx_status lbmx_batch_set_tile_rows(lbmx_batch* b, int rows) {
  CHECK_BATCH(b);
  if (rows != 1 && rows != 2 && rows != 4) {
    lbmx::g_last_error = "tile rows must be 1, 2 or 4";
    return LBMX_ERR_ARG;
  }
  b->b->ty = rows;
  b->b->pair = true;
  return LBMX_OK;
}

lbmx_status lbmx_batch_set_tile(lbmx_batch* b, int rows, int paired) {
  CHECK_BATCH(b);
  if (rows != 1 && rows != 2 && rows != 4) {
    lbmx::g_last_error = "tile rows must be 1, 2 or 4";
    return LBMX_ERR_ARG;
  }
  b->b->ty = rows;
  b->b->pair = paired != 0;
  return LBMX_OK;
}

lbmx_status lbmx_batch_get_tile(lbmx_batch* b, int* rows, int* paired) {
  CHECK_BATCH(b);
  if (rows) *rows = b->b->ty;
  if (paired) *paired = (b->b->ty > 1 && b->b->pair) ? 1 : 0;
  return LBMX_OK;
}

lbmx_status lbmx_batch_set_graph(lbmx_batch* b, int enable) {
  CHECK_BATCH(b);
  b->b->use_graph = enable != 0;
  if (!enable) b->b->drop_graph();
  return LBMX_OK;
}

lbmx_status lbmx_batch_autotune(lbmx_batch* b, int nsweeps, int* rows, double* ms) {
  CHECK_BATCH(b);
  return guarded([&] {
    int r = b->b->autotune(nsweeps, ms);
    if (rows) *rows = r;
  });
}

lbmx_status lbmx_batch_gradient(lbmx_batch* b, int nsteps, double rho0, const double u0[3], double T0, double* J,
                                double* const* sens) {
  CHECK_BATCH(b);
  return guarded([&] { b->b->gradient(nsteps, rho0, u0, T0, J, sens); });
}

lbmx_status lbmx_batch_get_stats(lbmx_batch* b, lbmx_stats* st) {
  CHECK_BATCH(b);
  if (!st) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  *st = b->b->stats;
  return LBMX_OK;
}

lbmx_status lbmx_batch_device_bytes(lbmx_batch* b, size_t* bytes) {
  CHECK_BATCH(b);
  if (!bytes) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  *bytes = b->b->device_bytes();
  return LBMX_OK;
}

lbmx_status lbmx_get_stats(lbmx_solver* s, lbmx_stats* st) {
  CHECK_HANDLE(s);
  if (!st) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  *st = s->e->stats;
  return LBMX_OK;
}

lbmx_status lbmx_time_sweeps(lbmx_solver* s, int kind, int nsweeps, double* ms) {
  CHECK_HANDLE(s);
  if (!ms) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { *ms = s->e->time_sweeps(kind, nsweeps); });
}

lbmx_status lbmx_device_bytes(lbmx_solver* s, size_t* bytes) {
  CHECK_HANDLE(s);
  if (!bytes) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  *bytes = s->e->device_bytes();
  return LBMX_OK;
}

}  // extern "C"
