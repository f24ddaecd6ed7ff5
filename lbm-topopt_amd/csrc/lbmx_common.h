// Structures shared by the host engine and the HIP kernels (no device code here).
#pragma once
#include <cstddef>
#include <cstdint>

namespace lbmx {

constexpr int kMaxTerms = 4;
constexpr int kMaxQ = 19;   // populations per node (D3Q19)
constexpr int kMaxBC = 64;  // boundary pieces per lattice

// Boundary pieces by row segment: bit b of bcrows[bc_row_index(...)] is set when piece b
// meets the 64 nodes x0..x0+63 (x0 = 64 floor(x/64)) of row (y, z). A sweep's wave covers
// exactly one such segment, so it visits only the pieces that can hold its nodes.
inline constexpr long long bc_row_index(int nx, int ny, int x, int y, int z) {
  return ((long long)z * ny + y) * ((nx + 63) / 64) + x / 64;
}
constexpr int kTileX = 64;   // threads along x per workgroup (one wave)
constexpr int kTileY = 4;    // waves per workgroup, stacked along y

// One boundary piece (a box on one domain face/edge/corner with its outward normal),
// applied on the device in insertion order.
struct DevBC {
  int type;
  int nrm[3];
  int box[6];  // x0 x1 y0 y1 z0 z1 (inclusive)
  int profile;
  int time_mode;
  int inc;     // 1: incompressible variant of the condition
  int aux;     // Neumann pieces: first slot of the piece in the adjoint side buffer (2 per node)
  double value;
  double dir[3];
  double center[2];
  double radius[2];
  double t_period, t_amp, t_offset, t_phase;
  // LBMX_PROFILE_TABLE: per-node values over the box tbox of the lbmx_add_boundary_table call,
  // natural order, ncomp per node (1: a scalar shape; D: one factor per velocity component)
  const double* table;
  int tbox[6];
  int ncomp;
  int pad0;
};

// One functional term: weight * sum over box of g_kind(state); kinds:
// 0 rho, 1 u_x, 2 total-pressure flux (p + |u|^2/2) u_x, 3 u_x T.
struct DevTerm {
  int box[6];
  int kind;
  int pad;
  double weight;
};

// Model constants passed by value to every kernel.
struct ModelParams {
  double omega;
  double alpha_max;
  double interp_q;
  double body[3];
  double tdiff, ramp_p, ramp_ck;
  int design;
  int pad;
  // the same values rounded to FP32 on the host, read by the FP32-arithmetic forms (no
  // per-wave conversions in the sweeps)
  float omega_f, alpha_max_f, interp_q_f, body_f[3], tdiff_f, ramp_p_f, ramp_ck_f, pad_f;
};

// Padded structure-of-arrays layout of one population field in HBM. Interior node (x,y,z)
// lives at element  q*plane + ((z+oz)*py + y+1)*px + ax + x ; rows are 128-byte aligned
// (ax = 128 bytes / element size) and a one-cell halo surrounds the domain. The halo holds
// the mid-way bounce-back reflections written by the sweep that produced the field.
// The byte offsets of each plane's bases are computed once on the host (a sweep's waves would
// otherwise each rebuild 38-57 64-bit addresses on the scalar unit): the own base q*plane, the
// pull base (q*plane - shift_q), read at node x - c_q, and the gather base (q*plane + shift_q),
// read at node x + c_q, with shift_q = c_qx + c_qy px + c_qz px py, times the element size.
struct Layout {
  int px, py;        // padded row length (elements), padded rows per slice
  int ax, oz;        // x offset of interior x = 0; z offset (1 in 3D, 0 in 2D)
  long long plane;   // elements per population plane
  long long own_b[kMaxQ], pull_b[kMaxQ], gath_b[kMaxQ];
};

struct Geometry {
  int nx, ny, nz;
  int nblk_x, nblk_y;  // workgroup tiles along x, y
  int nblocks;         // total workgroups
  long long N;         // interior nodes
};

struct ObjArgs {
  int nterm;
  int tbox[kMaxTerms][6];  // the terms by value (boxes, kinds, weights): no memory loads
  int tkind[kMaxTerms];
  double tweight[kMaxTerms];
  double time_const;
  double* acc;           // running sums per 64-node row segment (FP64), null = off
};

// Forward sweep: dst = C(B_t(S(src))) ; objective of B_t(S(src)) accumulated.
struct FwdArgs {
  const void* src;
  void* dst;
  const void* gamma;
  const DevBC* bcs;
  const unsigned long long* bcrows;  // pieces meeting each 64-node row segment (bc_row_index)
  int nbc;
  int t;              // time of the new state (boundary values)
  double tfac[kMaxBC];  // time factor of each boundary piece at t (computed on the host)
  int mode;           // 0 full sweep, 1 stream+BC only (AoS FP64 out), 2 collide AoS FP64 in
  int pad;
  void* aos_out;      // mode 1: pre-collision state written AoS FP64
  Geometry g;
  Layout ls;          // layout of src/dst
  ModelParams m;
  ObjArgs obj;
};

// Adjoint sweep at reverse step t: reads p_{t-1} (fwd_src), nu_t (lam_in), writes nu_{t-1}.
// With fwd_aos set, the pre-collision state s_t is read from that AoS FP64 array instead of
// being recomputed from p_{t-1} (a caller-held forward state).
struct AdjArgs {
  const void* fwd_src;
  const double* fwd_aos;
  const void* lam_in;
  void* lam_out;
  const void* gamma;
  double* sens;
  double* nside;      // Neumann pieces: the two coefficients each node sends to its inward neighbours
  const DevBC* bcs;
  const unsigned long long* bcrows;  // pieces meeting each 64-node row segment (bc_row_index)
  int nbc;
  int t;
  double tfac[kMaxBC];
  Geometry g;
  Layout ls;          // layout of the forward state
  Layout ll;          // layout of the adjoint fields
  ModelParams m;
  ObjArgs obj;
};

// One instance of a batch: its static description for the batched sweeps.
struct BatchInst {
  Geometry g;             // grid and tiles (for the batch's tile height)
  Layout ls, ll;          // layouts of its state and adjoint fields
  ModelParams m;
  ObjArgs obj;            // functional terms and per-tile accumulators of this instance
  const void* gamma;      // design field (element type of the precision form)
  double* sens;           // dJ/dgamma accumulator
  const DevBC* bcs;
  const unsigned long long* bcrows;  // pieces meeting each 64-node row segment (bc_row_index)
  int nbc;
  int blk_begin;          // first global workgroup of this instance
  long long state_off;    // byte offset of the instance in the batch state buffers
  long long lam_off;      // byte offset in the batch adjoint buffers
};

// Per-launch arguments of a batched sweep.
struct BatchStep {
  const void* src;        // forward: p_{t-1}; adjoint: the forward state p_{t-1}
  void* dst;              // forward: p_t
  const void* lam_in;     // adjoint: nu_t
  void* lam_out;          // adjoint: nu_{t-1}
  const BatchInst* inst;  // device table of instances
  const int* blk_inst;    // instance of each (XCD-remapped) global workgroup
  const double* tfac;     // time factors of this step: [instance][tf_stride]
  int tf_stride;
  int nblocks;            // total workgroups
  int t;                  // time of the new state (forward) / reverse step (adjoint)
  int with_obj;           // forward: accumulate the functional; adjoint: add its source
};

// Q directions c_q = (ex[q], ey[q], ez[q]) of the field's lattice
inline Layout make_layout(int nx, int ny, int nz, int dim, int elem_bytes, int Q, const int* ex, const int* ey,
                          const int* ez) {
  Layout l{};
  l.ax = 128 / elem_bytes;
  int rowlen = ((nx + 1 + l.ax - 1) / l.ax) * l.ax;
  l.px = l.ax + rowlen;
  l.py = ny + 2;
  l.oz = dim == 3 ? 1 : 0;
  long long pz = dim == 3 ? nz + 2 : 1;
  l.plane = (long long)l.px * l.py * pz;
  for (int q = 0; q < Q && q < kMaxQ; ++q) {
    const long long sh = (long long)ex[q] + (long long)ey[q] * l.px + (long long)ez[q] * l.px * l.py;
    l.own_b[q] = (long long)q * l.plane * elem_bytes;
    l.pull_b[q] = ((long long)q * l.plane - sh) * elem_bytes;
    l.gath_b[q] = ((long long)q * l.plane + sh) * elem_bytes;
  }
  return l;
}

}  // namespace lbmx
