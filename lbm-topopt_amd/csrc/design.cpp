// Design pipeline behind the lbmx_filter_* entry points: the design-box mapping, the chain of
// density/projection stages and their sensitivity map, on the device. One handle = one chain
// (a DensityFilter2d/3d, a ProjectionFilter2d/3d or a DoubleProjectionFilter2d of the
// reference) on one full grid.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <vector>

#include "host_common.hpp"
#include "lbmx.h"
#include "lbmx_launch.h"

namespace lbmx {

struct FilterStage {
  int R = 1;
  int project = 0;
  double beta = 1.0, eta = 0.5;
  DeviceBuffer wt, rx;   // cone weights and per-(dz, dy) x extents
  DeviceBuffer den;      // W 1 (rows outside the box: 1)
  DeviceBuffer dens;     // density-filtered field of the last design call
  DeviceBuffer proj;     // projected field of the last design call (projecting stages)
  DeviceBuffer eta_dev;  // per-node thresholds (VariableProjectionFilter2d), when set
  bool var_eta = false;
  const double* eta_field() const { return var_eta ? eta_dev.as<double>() : nullptr; }
  const double* output() const { return static_cast<const double*>(project ? proj.p : dens.p); }
};

struct Filter {
  lbmx_filter_desc d{};
  long long N = 0, ND = 0;
  int fbox[6], dbox[6];
  hipStream_t own = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  DeviceBuffer background, design_in, full0, tmp, acc, design_out;
  std::vector<std::unique_ptr<FilterStage>> stages;
  bool have_design = false;
  double ms[2] = {0, 0};

  explicit Filter(const lbmx_filter_desc& desc) : d(desc) {
    if (d.nx < 1 || d.ny < 1 || d.nz < 1) throw Error{LBMX_ERR_ARG, "grid dimensions must be positive"};
    if (d.nstages < 1 || d.nstages > 2) throw Error{LBMX_ERR_ARG, "nstages must be 1 or 2"};
    for (int k = 0; k < d.nstages; ++k)
      if (d.radius[k] < 1 || d.radius[k] > LBMX_FILTER_MAX_RADIUS)
        throw Error{LBMX_ERR_ARG, "filter radius out of range 1..LBMX_FILTER_MAX_RADIUS"};
    const int dims[3] = {d.nx, d.ny, d.nz};
    const lbmx_box& db = d.design_box;
    const int dv[6] = {db.x0, db.x1, db.y0, db.y1, db.z0, db.z1};
    for (int a = 0; a < 3; ++a) {
      if (dv[2 * a] < 0 || dv[2 * a + 1] >= dims[a] || dv[2 * a] > dv[2 * a + 1])
        throw Error{LBMX_ERR_ARG, "design box must be a non-empty box inside the grid"};
      dbox[2 * a] = dv[2 * a];
      dbox[2 * a + 1] = dv[2 * a + 1];
    }
    // the filter domain is intersected with the grid (boxIntersection with the bounding box)
    const lbmx_box& fb = d.filter_box;
    const int fv[6] = {fb.x0, fb.x1, fb.y0, fb.y1, fb.z0, fb.z1};
    for (int a = 0; a < 3; ++a) {
      fbox[2 * a] = std::max(fv[2 * a], 0);
      fbox[2 * a + 1] = std::min(fv[2 * a + 1], dims[a] - 1);
    }
    N = (long long)d.nx * d.ny * d.nz;
    ND = (long long)(dbox[1] - dbox[0] + 1) * (dbox[3] - dbox[2] + 1) * (dbox[5] - dbox[4] + 1);
    HIP_OK(hipStreamCreate(&own));
    HIP_OK(hipEventCreate(&ev[0]));
    HIP_OK(hipEventCreate(&ev[1]));
    const size_t nb = (size_t)N * 8;
    background.ensure(nb);
    full0.ensure(nb);
    tmp.ensure(nb);
    acc.ensure(nb);
    design_in.ensure((size_t)ND * 8);
    design_out.ensure((size_t)ND * 8);
    std::vector<double> ones((size_t)N, 1.0);
    HIP_OK(hipMemcpy(background.p, ones.data(), nb, hipMemcpyHostToDevice));
    for (int k = 0; k < d.nstages; ++k) {
      auto st = std::make_unique<FilterStage>();
      st->R = d.radius[k];
      st->project = d.project[k] != 0;
      st->beta = d.beta[k];
      st->eta = d.eta[k];
      build_cone(*st);
      st->den.ensure(nb);
      st->dens.ensure(nb);
      if (st->project) st->proj.ensure(nb);
      // denominators: the filter applied to a field of ones, without the division
      ConeArgs a = cone_args(*st, static_cast<const double*>(background.p), st->den.as<double>(), false);
      a.den = nullptr;
      HIP_OK(launch_cone(a, own));
      stages.push_back(std::move(st));
    }
    HIP_OK(hipStreamSynchronize(own));
  }

  ~Filter() {
    if (own) {
      (void)hipStreamSynchronize(own);
      (void)hipStreamDestroy(own);
    }
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }

  // weights R - |d| for |d| <= R (DensityFilter2d::setUp, src/topOpt/densityFilter2d.cc:103-131)
  void build_cone(FilterStage& st) {
    const int R = st.R, S = 2 * R + 1;
    std::vector<double> wt((size_t)S * S * S, 0.0);
    std::vector<int> rx((size_t)S * S, -1);
    const double fr = (double)R;
    for (int dz = -R; dz <= R; ++dz)
      for (int dy = -R; dy <= R; ++dy)
        for (int dx = -R; dx <= R; ++dx) {
          const double dist = std::sqrt((double)(dx * dx + dy * dy + dz * dz));
          if (dist <= fr) {
            wt[((size_t)(dz + R) * S + (dy + R)) * S + (dx + R)] = fr - dist;
            int& r = rx[(size_t)(dz + R) * S + (dy + R)];
            r = std::max(r, std::abs(dx));
          }
        }
    st.wt.ensure(wt.size() * 8);
    st.rx.ensure(rx.size() * sizeof(int));
    HIP_OK(hipMemcpy(st.wt.p, wt.data(), wt.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(st.rx.p, rx.data(), rx.size() * sizeof(int), hipMemcpyHostToDevice));
  }

  ConeArgs cone_args(const FilterStage& st, const double* in, double* out, bool transposed) const {
    ConeArgs a{};
    a.in = in;
    a.out = out;
    a.den = st.den.as<double>();
    a.wt = st.wt.as<double>();
    a.rx = st.rx.as<int>();
    a.nx = d.nx;
    a.ny = d.ny;
    a.nz = d.nz;
    a.R = st.R;
    for (int k = 0; k < 6; ++k) a.box[k] = fbox[k];
    a.transposed = transposed ? 1 : 0;
    return a;
  }

  // tanh(beta eta) and tanh(beta eta) + tanh(beta (1 - eta)) of a projecting stage
  static void proj_consts(const FilterStage& st, double& tb, double& tsum) {
    tb = std::tanh(st.beta * st.eta);
    tsum = tb + std::tanh(st.beta * (1.0 - st.eta));
  }

  // mapDesignToFull + filterDesign of every stage (the last stage's output is the physical
  // field); the design is a host array, or a device array read in place when dev is set
  const double* run_design(const double* design, bool dev, hipStream_t s) {
    const double* din = design;
    if (!dev) {
      HIP_OK(hipMemcpyAsync(design_in.p, design, (size_t)ND * 8, hipMemcpyHostToDevice, s));
      din = design_in.as<double>();
    }
    HIP_OK(hipEventRecord(ev[0], s));
    HIP_OK(launch_map_to_full(full0.as<double>(), din, background.as<double>(), d.nx, d.ny, d.nz, dbox, s));
    const double* x = full0.as<double>();
    for (auto& st : stages) {
      HIP_OK(launch_cone(cone_args(*st, x, st->dens.as<double>(), false), s));
      if (st->project) {
        double tb, tsum;
        proj_consts(*st, tb, tsum);
        HIP_OK(launch_project(st->proj.as<double>(), st->dens.as<double>(), d.nx, d.ny, d.nz, fbox, st->beta,
                              st->eta, tb, tsum, st->eta_field(), s));
      }
      x = st->output();
    }
    HIP_OK(hipEventRecord(ev[1], s));
    have_design = true;
    return x;
  }

  // filterSensitivities (stages in reverse) + mapFullToDesign; sens_dev is only read. The
  // result goes to a host array, or straight into a device array when out_dev is set.
  void run_sensitivities(const double* sens_dev, double* out, bool out_dev, hipStream_t s) {
    for (auto& st : stages)
      if (st->project && !have_design)
        throw Error{LBMX_ERR_STATE, "a projecting filter needs lbmx_filter_design before its sensitivities"};
    HIP_OK(hipEventRecord(ev[0], s));
    const double* cur = sens_dev;
    for (int k = (int)stages.size() - 1; k >= 0; --k) {
      FilterStage& st = *stages[k];
      double tb = 0, tsum = 1;
      if (st.project) proj_consts(st, tb, tsum);
      HIP_OK(launch_sens_scale(tmp.as<double>(), cur, st.dens.as<double>(), st.den.as<double>(), d.nx, d.ny, d.nz,
                               fbox, st.project, st.beta, st.eta, tsum, st.eta_field(), s));
      HIP_OK(launch_cone(cone_args(st, tmp.as<double>(), acc.as<double>(), true), s));
      cur = acc.as<double>();
    }
    HIP_OK(launch_extract_box(out_dev ? out : design_out.as<double>(), cur, d.nx, d.ny, d.nz, dbox, s));
    HIP_OK(hipEventRecord(ev[1], s));
    if (!out_dev) HIP_OK(hipMemcpyAsync(out, design_out.p, (size_t)ND * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    float t = 0;
    HIP_OK(hipEventElapsedTime(&t, ev[0], ev[1]));
    ms[1] = t;
  }

  void finish_design_timing() {
    float t = 0;
    HIP_OK(hipEventElapsedTime(&t, ev[0], ev[1]));
    ms[0] = t;
  }

  void check_sizes(size_t nd, size_t nf, bool need_full) const {
    if ((long long)nd != ND) throw Error{LBMX_ERR_ARG, "design vector length != design-box nodes"};
    if (need_full && (long long)nf != N) throw Error{LBMX_ERR_ARG, "full field length != grid nodes"};
  }
};

}  // namespace lbmx

struct lbmx_filter {
  std::unique_ptr<lbmx::Filter> f;
};

using lbmx::Error;
using lbmx::guarded;

#define CHECK_FILTER(h)                                    \
  if (!(h) || !(h)->f) {                                   \
    lbmx::g_last_error = "null filter handle";             \
    return LBMX_ERR_ARG;                                   \
  }

extern "C" {

lbmx_status lbmx_filter_create(const lbmx_filter_desc* desc, lbmx_filter** out) {
  if (!desc || !out) {
    lbmx::g_last_error = "null argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    auto h = std::make_unique<lbmx_filter>();
    h->f = std::make_unique<lbmx::Filter>(*desc);
    *out = h.release();
  });
}

lbmx_status lbmx_filter_destroy(lbmx_filter* f) {
  delete f;
  return LBMX_OK;
}

lbmx_status lbmx_filter_set_projection(lbmx_filter* h, int stage, double beta, double eta) {
  CHECK_FILTER(h);
  return guarded([&] {
    if (stage < 0 || stage >= (int)h->f->stages.size()) throw Error{LBMX_ERR_ARG, "no such stage"};
    h->f->stages[stage]->beta = beta;
    h->f->stages[stage]->eta = eta;
  });
}

lbmx_status lbmx_filter_set_eta_field(lbmx_filter* h, int stage, const double* eta, size_t n) {
  CHECK_FILTER(h);
  return guarded([&] {
    lbmx::Filter& F = *h->f;
    if (stage < 0 || stage >= (int)F.stages.size()) throw Error{LBMX_ERR_ARG, "no such stage"};
    lbmx::FilterStage& st = *F.stages[stage];
    if (!eta) {
      st.var_eta = false;
      return;
    }
    if ((long long)n != F.N) throw Error{LBMX_ERR_ARG, "eta field length != grid nodes"};
    st.eta_dev.ensure(n * 8);
    HIP_OK(hipMemcpy(st.eta_dev.p, eta, n * 8, hipMemcpyHostToDevice));
    st.var_eta = true;
  });
}

lbmx_status lbmx_filter_increase_beta(lbmx_filter* h, double factor) {
  CHECK_FILTER(h);
  for (auto& st : h->f->stages) st->beta *= factor;
  return LBMX_OK;
}

lbmx_status lbmx_filter_set_background(lbmx_filter* h, const double* full, size_t n) {
  CHECK_FILTER(h);
  if (!full) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    if ((long long)n != h->f->N) throw Error{LBMX_ERR_ARG, "full field length != grid nodes"};
    HIP_OK(hipMemcpy(h->f->background.p, full, n * 8, hipMemcpyHostToDevice));
  });
}

}  // extern "C"

namespace {

lbmx_status filter_design(lbmx_filter* h, const double* design, size_t nd, double* physical, size_t nf, bool dev) {
  CHECK_FILTER(h);
  if (!design) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    lbmx::Filter& F = *h->f;
    F.check_sizes(nd, nf, physical != nullptr);
    const double* x = F.run_design(design, dev, F.own);
    if (physical)
      HIP_OK(hipMemcpyAsync(physical, x, (size_t)F.N * 8, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                            F.own));
    HIP_OK(hipStreamSynchronize(F.own));
    F.finish_design_timing();
  });
}

void check_solver_grid(const lbmx::Filter& F, lbmx_solver* s) {
  int dims[3];
  lbmx::solver_nodes(s, dims);
  if (dims[0] != F.d.nx || dims[1] != F.d.ny || dims[2] != F.d.nz)
    throw Error{LBMX_ERR_ARG, "solver grid differs from the filter grid"};
}

lbmx_status filter_design_to_solver(lbmx_filter* h, const double* design, size_t nd, lbmx_solver* s, bool dev) {
  CHECK_FILTER(h);
  if (!design || !s) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    lbmx::Filter& F = *h->f;
    F.check_sizes(nd, 0, false);
    check_solver_grid(F, s);
    hipStream_t st = lbmx::solver_stream(s);
    const double* x = F.run_design(design, dev, st);
    lbmx::solver_set_design_device(s, x);
    HIP_OK(hipStreamSynchronize(st));
    F.finish_design_timing();
  });
}

lbmx_status filter_sensitivities(lbmx_filter* h, const double* full_sens, size_t nf, double* design_sens, size_t nd,
                                 bool dev) {
  CHECK_FILTER(h);
  if (!full_sens || !design_sens) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    lbmx::Filter& F = *h->f;
    F.check_sizes(nd, nf, true);
    const double* in = full_sens;
    if (!dev) {
      // the host input is staged in the accumulation buffer, which the first stage only
      // reads before the transposed filter overwrites it
      HIP_OK(hipMemcpyAsync(F.acc.p, full_sens, nf * 8, hipMemcpyHostToDevice, F.own));
      in = F.acc.as<double>();
    }
    F.run_sensitivities(in, design_sens, dev, F.own);
  });
}

lbmx_status filter_solver_sensitivities(lbmx_filter* h, lbmx_solver* s, double* design_sens, size_t nd, bool dev) {
  CHECK_FILTER(h);
  if (!s || !design_sens) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    lbmx::Filter& F = *h->f;
    F.check_sizes(nd, 0, false);
    check_solver_grid(F, s);
    const double* sd = lbmx::solver_sensitivity_device(s);
    F.run_sensitivities(sd, design_sens, dev, lbmx::solver_stream(s));
  });
}

}  // namespace

extern "C" {

lbmx_status lbmx_filter_design(lbmx_filter* h, const double* design, size_t nd, double* physical, size_t nf) {
  return filter_design(h, design, nd, physical, nf, false);
}
lbmx_status lbmx_filter_design_device(lbmx_filter* h, const double* design, size_t nd, double* physical, size_t nf) {
  return filter_design(h, design, nd, physical, nf, true);
}
lbmx_status lbmx_filter_design_to_solver(lbmx_filter* h, const double* design, size_t nd, lbmx_solver* s) {
  return filter_design_to_solver(h, design, nd, s, false);
}
lbmx_status lbmx_filter_design_to_solver_device(lbmx_filter* h, const double* design, size_t nd, lbmx_solver* s) {
  return filter_design_to_solver(h, design, nd, s, true);
}
lbmx_status lbmx_filter_sensitivities(lbmx_filter* h, const double* full_sens, size_t nf, double* design_sens,
                                      size_t nd) {
  return filter_sensitivities(h, full_sens, nf, design_sens, nd, false);
}
lbmx_status lbmx_filter_sensitivities_device(lbmx_filter* h, const double* full_sens, size_t nf, double* design_sens,
                                             size_t nd) {
  return filter_sensitivities(h, full_sens, nf, design_sens, nd, true);
}
lbmx_status lbmx_filter_solver_sensitivities(lbmx_filter* h, lbmx_solver* s, double* design_sens, size_t nd) {
  return filter_solver_sensitivities(h, s, design_sens, nd, false);
}
lbmx_status lbmx_filter_solver_sensitivities_device(lbmx_filter* h, lbmx_solver* s, double* design_sens, size_t nd) {
  return filter_solver_sensitivities(h, s, design_sens, nd, true);
}

lbmx_status lbmx_filter_get_field(lbmx_filter* h, int which, double* out, size_t n) {
  CHECK_FILTER(h);
  if (!out) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] {
    lbmx::Filter& F = *h->f;
    if ((long long)n != F.N) throw Error{LBMX_ERR_ARG, "full field length != grid nodes"};
    const int stage = which / 2;
    if (which < 0 || stage >= (int)F.stages.size()) throw Error{LBMX_ERR_ARG, "no such field"};
    if (!F.have_design) throw Error{LBMX_ERR_STATE, "no design filtered yet"};
    const lbmx::FilterStage& st = *F.stages[stage];
    const void* src = (which % 2 == 0) ? st.dens.p : static_cast<const void*>(st.output());
    HIP_OK(hipMemcpy(out, src, n * 8, hipMemcpyDeviceToHost));
  });
}

lbmx_status lbmx_filter_timing(lbmx_filter* h, double ms[2]) {
  CHECK_FILTER(h);
  if (!ms) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  ms[0] = h->f->ms[0];
  ms[1] = h->f->ms[1];
  return LBMX_OK;
}

}  // extern "C"
