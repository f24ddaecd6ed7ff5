// Sweep kernels over the padded structure-of-arrays layout (see Layout in lbmx_common.h).
//
// Addressing: the source of population q for node x is x - c_q (pull). Its element index is
// q*plane - shift_q + off(x), shift_q = cx + cy*px + cz*px*py. The per-plane part is uniform
// (scalar registers) and off(x) is one 32-bit lane offset shared by all planes, so each load
// is a single global_load with a scalar base and no per-lane address arithmetic. Domain
// edges need no special case: the sweep that produced a field also wrote, for every
// boundary node, the mid-way bounce-back reflection of each outgoing population into the
// halo cell it would have streamed to.
//
//  k_fwd : pull p_{t-1}, boundary pieces (insertion order), functional of s_t, collision,
//          store p_t (aligned) + halo reflections at boundary nodes.
//  k_adj : recompute s_t = B_t S p_{t-1}, gather nu_t = S^T m_t from the previous adjoint
//          sweep's m_t (at x + c_q; the own node's opposite plane where x + c_q is outside),
//          reverse collision, functional source, transposed boundaries (reverse order), store
//          m_{t-1} at the own node; dJ/dgamma accumulated in FP64.
//  k_fwd_x2 : k_fwd for the 16-bit storage forms on large grids, two x-adjacent nodes per
//          thread (same results).
// Workgroups are 64 (x) x 4 (y) threads; the linear workgroup id is remapped so that each
// XCD (workgroups are dealt round-robin to the 8 XCDs) receives a contiguous run of tiles and
// x-neighbouring tiles share that XCD's L2 for the lines straddling their edge.
#pragma once
#include "lbmx_collision.hpp"

namespace lbmx {

template <class S>
__device__ __forceinline__ S ld(const char* base, unsigned int off_bytes) {
  return *reinterpret_cast<const S*>(base + off_bytes);
}
template <class S>
__device__ __forceinline__ void st(char* base, unsigned int off_bytes, S v) {
  *reinterpret_cast<S*>(base + off_bytes) = v;
}
// Stores of whole aligned rows that nothing reads before the next sweep bypass the caches'
// retention (non-temporal): on MI355X this leaves L2 to the lines that shifted pulls share
// between neighbouring workgroups. Every sweep writes only its own nodes, so all its stores
// are whole rows; the shifts are all on the read side.
template <class S>
__device__ __forceinline__ void st_nt(char* base, unsigned int off_bytes, S v) {
  __builtin_nontemporal_store(v, reinterpret_cast<S*>(base + off_bytes));
}

// Plane bases from the layout's byte offsets (built on the host for the field's element size,
// which S names here: sizeof(S) of the field's storage type)
// pull address base of plane q: start + (q*plane - shift_q) elements
template <class L, int q, class S>
__device__ __forceinline__ const char* pull_base(const void* start, const Layout& l) {
  static_assert(q < kMaxQ, "plane index");
  return static_cast<const char*>(start) + l.pull_b[q];
}
template <class L, int q, class S>
__device__ __forceinline__ char* pull_base_w(void* start, const Layout& l) {
  return static_cast<char*>(start) + l.pull_b[q];
}
// gather address base of plane q: start + (q*plane + shift_q) elements (node x + c_q)
template <class L, int q, class S>
__device__ __forceinline__ const char* gather_base(const void* start, const Layout& l) {
  return static_cast<const char*>(start) + l.gath_b[q];
}
template <int q, class S>
__device__ __forceinline__ char* own_base_w(void* start, const Layout& l) {
  return static_cast<char*>(start) + l.own_b[q];
}
template <int q, class S>
__device__ __forceinline__ const char* own_base(const void* start, const Layout& l) {
  return static_cast<const char*>(start) + l.own_b[q];
}

__device__ __forceinline__ unsigned int layout_off(const Layout& l, int x, int y, int z) {
  return (unsigned int)(((z + l.oz) * l.py + (y + 1)) * l.px + l.ax + x);
}

// node coordinates of this thread from the XCD-remapped workgroup id
__device__ __forceinline__ void tile_coords(const Geometry& g, int& x, int& y, int& z, int& tile) {
  const int b = blockIdx.x;
  const int nb = g.nblocks;
  const int qd = nb >> 3, r = nb & 7;
  const int xcd = b & 7, k = b >> 3;
  tile = xcd < r ? xcd * (qd + 1) + k : r * (qd + 1) + (xcd - r) * qd + k;
  const int bx = tile % g.nblk_x;
  const int rest = tile / g.nblk_x;
  const int by = rest % g.nblk_y;
  z = rest / g.nblk_y;
  x = bx * kTileX + threadIdx.x;
  y = by * kTileY + threadIdx.y;
}

// true when x + c_q leaves the domain (q is an outgoing direction at a boundary node)
template <class L, int q>
__device__ __forceinline__ bool leaves(int x, int y, int z, int nx, int ny, int nz) {
  bool o = false;
  if constexpr (L::ex[q] == 1) o = o || x == nx - 1;
  if constexpr (L::ex[q] == -1) o = o || x == 0;
  if constexpr (L::ey[q] == 1) o = o || y == ny - 1;
  if constexpr (L::ey[q] == -1) o = o || y == 0;
  if constexpr (L::ez[q] == 1) o = o || z == nz - 1;
  if constexpr (L::ez[q] == -1) o = o || z == 0;
  return o;
}

template <class L>
__device__ __forceinline__ bool on_boundary(int x, int y, int z, int nx, int ny, int nz) {
  bool b = x == 0 || x == nx - 1 || y == 0 || y == ny - 1;
  if constexpr (L::D == 3) b = b || z == 0 || z == nz - 1;
  return b;
}

// Pull of the populations of one node. NT selects the planes read with non-temporal loads:
// 0 none, 1 those without an x shift, 2 all. Each element of the source field is read once in
// a sweep; only x-shifted rows share lines with the neighbouring workgroup. Measured on the
// D3Q19 MRT sweeps: the forward gains with all planes non-temporal, the adjoint (whose λ
// gathers compete for L2) with the unshifted planes only.
template <class L, class P, int NT>
__device__ __forceinline__ void pull_node(const void* src, const Layout& l, unsigned int off,
                                          typename P::T (&f)[L::Q]) {
  using S = typename P::S;
  const unsigned int ob = off * (unsigned int)sizeof(S);
  static_for<0, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (NT == 2 || (NT == 1 && L::ex[q] == 0))
      f[q] = P::template dec<L, q>(__builtin_nontemporal_load(reinterpret_cast<const S*>(pull_base<L, q, S>(src, l) + ob)));
    else
      f[q] = P::template dec<L, q>(ld<S>(pull_base<L, q, S>(src, l), ob));
  });
}

// Halo reflections of a boundary node: p[q] of the node is what the pull of plane opp(q) at
// this node reads from x + c_q, for every q that leaves the domain. Opposite directions share a
// weight and a storage shift, so the node's stored encoding of q is that value. enc(q): the
// node's stored encoding of direction q.
template <class L, class S, class Enc>
__device__ __forceinline__ void store_halo(void* dst, const Layout& l, unsigned int ob, int x, int y, int z, int nx,
                                           int ny, int nz, Enc&& enc) {
  static_for<1, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (L::ex[q] != 0 || L::ey[q] != 0 || L::ez[q] != 0) {
      static_assert(L::w[q] == L::w[L::opp[q]] && store_shift<L, q>() == store_shift<L, L::opp[q]>(),
                    "opposite directions must share a weight and a storage shift");
      if (leaves<L, q>(x, y, z, nx, ny, nz)) st<S>(pull_base_w<L, L::opp[q], S>(dst, l), ob, enc(qc));
    }
  });
}

// The 16-bit forms take the moment-space collisions' doubled populations (collide<..., true>):
// their encoders fold the factor 1/2 into 1/w exactly, so the stored bits are unchanged.
template <class P, int COLL, int DES>
inline constexpr bool kStoreTwice = P::shifted && kCollideTwice<COLL, DES>;

// The node's two hashed words for the stochastic rounding: the node offset keyed by the step
// (uniform per launch)
__device__ __forceinline__ void sr_words(unsigned int off, int tstep, unsigned int& r0, unsigned int& r1) {
  const unsigned int tkey = mix32((unsigned int)tstep * 0x9e3779b9u + 0x632be5abu);
  r0 = (off ^ tkey) * 0x9e3779b1u;
  r0 ^= r0 >> 15;
  r1 = r0 * 0x85ebca6bu;
  r1 ^= r1 >> 13;
}

// store post-collision populations and, at boundary nodes, the halo reflections. tstep seeds
// the stochastic rounding of the reduced storage forms (with the node offset). TW: f holds the
// collision's doubled populations (collide<..., true>), which only the 16-bit codecs take.
template <class L, class P, bool TW = false>
__device__ __forceinline__ void store_node(void* dst, const Layout& l, unsigned int off, const typename P::T (&f)[L::Q],
                                           bool bnd, int x, int y, int z, int nx, int ny, int nz, int tstep) {
  using S = typename P::S;
  const unsigned int ob = off * (unsigned int)sizeof(S);
  if constexpr (P::stochastic) {
    // two directions per 32-bit word: direction q in the low half, q + 1 in the high half
    unsigned int r0, r1;
    sr_words(off, tstep, r0, r1);
    unsigned int e[(L::Q + 1) / 2];
    static_for<0, (L::Q + 1) / 2>([&](auto pc) {
      constexpr int k = decltype(pc)::value, q = 2 * k;
      if constexpr (q + 1 < L::Q)
        e[k] = P::template enc_pair<L, q, q + 1, TW>(f[q], f[q + 1], sr_seed<q, P::kRandomBits>(r0, r1),
                                                     sr_seed<q + 1, P::kRandomBits>(r0, r1));
      else
        e[k] = P::template enc_sr<L, q, false, TW>(0u, f[q], sr_seed<q, P::kRandomBits>(r0, r1));
    });
    auto enc = [&](auto qc) -> S {
      constexpr int q = decltype(qc)::value;
      return (q & 1) ? P::template half_of<true>(e[q >> 1]) : P::template half_of<false>(e[q >> 1]);
    };
    static_for<0, L::Q>([&](auto qc) { st_nt<S>(own_base_w<decltype(qc)::value, S>(dst, l), ob, enc(qc)); });
    if (bnd) store_halo<L, S>(dst, l, ob, x, y, z, nx, ny, nz, enc);
  } else {
    static_assert(!TW, "doubled populations go to the 16-bit codecs only");
    auto enc = [&](auto qc) -> S { return P::template enc<L, decltype(qc)::value>(f[decltype(qc)::value]); };
    static_for<0, L::Q>([&](auto qc) { st_nt<S>(own_base_w<decltype(qc)::value, S>(dst, l), ob, enc(qc)); });
    if (bnd) store_halo<L, S>(dst, l, ob, x, y, z, nx, ny, nz, enc);
  }
}

// Zou-He Neumann conditions (2D, east/west): the node's unknowns follow from the moment of
// the two nodes inwards extrapolated to the boundary, (4 m_1 - m_2) / 3, with m = u_x for the
// flow (IncompressibleZouHeNeumann{East,West}2d<D2Q9>, src/boundaryConditions/ZouHe2d.hh:440-527;
// the unknowns are those of the Zou-He velocity condition with u_y = 0) and m = T for the
// D2Q5 temperature (ThermalZouHeNeumann{East,West}2d<.,D2Q5>, thermalZouHe2d.hh:229-300). The
// grid has nx >= 4, so both neighbours are interior nodes and their post-stream values are
// plain pulls of the source field.
template <class L, class P>
__device__ __forceinline__ void neumann_node(typename P::T (&fio)[L::Q], const DevBC& e, const void* src,
                                             const Layout& ls, int x, int y, int z) {
  using T = typename P::T;
  using S = typename P::S;
  const int dir = -e.nrm[0];
  const unsigned int o1 = layout_off(ls, x + dir, y, z) * (unsigned int)sizeof(S);
  const unsigned int o2 = layout_off(ls, x + 2 * dir, y, z) * (unsigned int)sizeof(S);
  auto g = [&](auto qc, unsigned int ob) -> T {
    constexpr int q = decltype(qc)::value;
    return P::template dec<L, q>(ld<S>(pull_base<L, q, S>(src, ls), ob));
  };
  T f[L::Q];
  static_for<0, L::Q>([&](auto qc) { f[decltype(qc)::value] = fio[decltype(qc)::value]; });
  if (e.type == LBMX_BC_ZOUHE_NEUMANN) {
    const T u1 = g(IC<1>{}, o1) + g(IC<7>{}, o1) + g(IC<8>{}, o1) - (g(IC<3>{}, o1) + g(IC<4>{}, o1) + g(IC<5>{}, o1));
    const T u2 = g(IC<1>{}, o2) + g(IC<7>{}, o2) + g(IC<8>{}, o2) - (g(IC<3>{}, o2) + g(IC<4>{}, o2) + g(IC<5>{}, o2));
    zh_velocity<T>(f, e.nrm, true, (T(4) * u1 - u2) / T(3), T(0));
  } else if constexpr (L::QT == 5) {
    constexpr int o = L::QF;
    const T T1 = g(IC<o>{}, o1) + g(IC<o + 1>{}, o1) + g(IC<o + 2>{}, o1) + g(IC<o + 3>{}, o1) + g(IC<o + 4>{}, o1);
    const T T2 = g(IC<o>{}, o2) + g(IC<o + 1>{}, o2) + g(IC<o + 2>{}, o2) + g(IC<o + 3>{}, o2) + g(IC<o + 4>{}, o2);
    const T Tv = (T(4) * T1 - T2) / T(3);
    const bool east = e.nrm[0] == 1;
    const T g2 = Tv - f[o] - f[o + 1] - f[o + 3] - f[o + 4];
    const T g4 = Tv - f[o] - f[o + 1] - f[o + 2] - f[o + 3];
    f[o + 2] = bsel(east, g2, f[o + 2]);
    f[o + 4] = bsel(east, f[o + 4], g4);
  }
  static_for<0, L::Q>([&](auto qc) { fio[decltype(qc)::value] = f[decltype(qc)::value]; });
}

template <class L>
__device__ __forceinline__ bool is_neumann(const DevBC& e) {
  if constexpr (L::D == 2) return e.type == LBMX_BC_ZOUHE_NEUMANN || e.type == LBMX_BC_THERMAL_NEUMANN;
  else return false;
}

// Boundary pieces in insertion order at node (x,y,z) of the post-stream state; src/ls: the
// field the state was pulled from (the Neumann conditions read two inward neighbours).
// The boundary pieces that meet this wave's row segment. Every lane of a wave has the same y
// and z and x in the same 64-node segment (workgroups are 64 threads wide in x, one wave per
// row), so the first active lane's word is the wave's. The sweeps load the word with their
// population loads (bc_row_word) and make it uniform only in the boundary branch (bc_pieces),
// so its latency overlaps theirs.
static_assert(kTileX == 64, "a wave must cover one 64-node row segment");
__device__ __forceinline__ unsigned long long bc_row_word(const unsigned long long* rows, int nx, int ny, int x, int y,
                                                          int z) {
  return rows[bc_row_index(nx, ny, x, y, z)];
}
__device__ __forceinline__ unsigned long long bc_pieces(unsigned long long m) {
  const unsigned int lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(m));
  const unsigned int hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(m >> 32));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// The boundary pieces of a wave's row segment (descriptors and this step's time factors),
// copied into LDS by the wave itself as soon as its row word has arrived, while its population
// loads are still in flight. Its boundary lanes then read them from LDS instead of through a
// chain of scalar and global loads that would start only after the populations. A wave reads
// back only the slots it wrote itself, so no workgroup barrier is needed; two waves of a
// workgroup may write the same slot, with the same words. Dynamic LDS of the launch:
// pieces_lds_bytes(nbc).
extern __shared__ __align__(16) unsigned char pieces_lds[];

__host__ __device__ constexpr size_t pieces_lds_bytes(int nbc) {
  return (size_t)(nbc > 0 ? nbc : 0) * (sizeof(DevBC) + sizeof(double));
}
__device__ __forceinline__ const DevBC* staged_pieces() { return reinterpret_cast<const DevBC*>(pieces_lds); }
__device__ __forceinline__ const double* staged_tfac(int nbc) {
  return reinterpret_cast<const double*>(pieces_lds + (size_t)nbc * sizeof(DevBC));
}

// Called by every lane of the wave with the same mask.
__device__ __forceinline__ void stage_wave_pieces(const DevBC* __restrict__ bcs, const double* __restrict__ tfac, int nbc,
                                                  unsigned long long mask) {
  constexpr int W = (int)(sizeof(DevBC) / 4);
  static_assert(sizeof(DevBC) % 4 == 0 && W < kTileX, "a descriptor is copied one 32-bit word per lane");
  const int lane = (int)threadIdx.x;
  const unsigned int* s = reinterpret_cast<const unsigned int*>(bcs);
  unsigned int* d = reinterpret_cast<unsigned int*>(pieces_lds);
  double* tf = reinterpret_cast<double*>(pieces_lds + (size_t)nbc * sizeof(DevBC));
  for (unsigned long long m = mask; m != 0; m &= m - 1) {
    const int b = __builtin_ctzll(m);
    if (lane < W) d[b * W + lane] = s[b * W + lane];
    else if (lane == W) tf[b] = tfac[b];
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// The wave's row segment (x0 .. x0+63 of row y, slice z) holds a boundary node.
template <class L>
__device__ __forceinline__ bool segment_on_boundary(int x0, int y, int z, int nx, int ny, int nz) {
  bool b = x0 == 0 || x0 + kTileX >= nx || y == 0 || y == ny - 1;
  if constexpr (L::D == 3) b = b || z == 0 || z == nz - 1;
  return b;
}

template <class L, class P>
__device__ __forceinline__ void apply_bcs(typename P::T (&f)[L::Q], const DevBC* __restrict__ bcs,
                                          unsigned long long pieces, const double* tfac, int x, int y, int z,
                                          const void* src, const Layout& ls) {
  using T = typename P::T;
  for (unsigned long long m = pieces; m != 0; m &= m - 1) {  // increasing piece index
    const int b = __builtin_ctzll(m);
    const DevBC& e = bcs[b];
    if (!in_box(e.box, x, y, z)) continue;
    if (is_neumann<L>(e)) {
      if constexpr (L::D == 2) neumann_node<L, P>(f, e, src, ls, x, y, z);
    } else {
      bc_node<L, T>(f, e, T(tfac[b]), x, y, z);
    }
  }
}

// Transposes in reverse order. A Neumann piece also feeds its inward neighbours: the node
// writes the two coefficients of that transfer to nside (slot e.aux + 2 (y - y0)), which
// k_neumann_fix adds into the neighbours' stored adjoints after the sweep
// (AdjointIncompressibleZouHeNeumann{East,West}2d, src/adjointBoundaryConditions/adjointZouHe2d.cc:114-172;
// AdjointThermalZouHeNeumann{East,West}2d, adjointThermalZouHe2d.hh:161-227).
template <class L, class T>
__device__ __forceinline__ void apply_bcs_adj(T (&l)[L::Q], const DevBC* __restrict__ bcs, unsigned long long pieces,
                                              int x, int y, int z, double* nside) {
  for (unsigned long long m = pieces; m != 0;) {  // decreasing piece index
    const int b = 63 - __builtin_clzll(m);
    m &= ~(1ull << b);
    const DevBC& e = bcs[b];
    if (!in_box(e.box, x, y, z)) continue;
    if constexpr (L::D == 2) {
      if (is_neumann<L>(e)) {
        const bool east = e.nrm[0] == 1;
        T c1, c2;
        if (e.type == LBMX_BC_ZOUHE_NEUMANN) {
          // u_x of neighbour k enters the unknowns with weights -(1/6, 2/3, 1/6) (east) times
          // d u_x / d u_k = 4/3, -1/3; neighbour k's u_x is sum_q c_qx f_q
          const T a = east ? l[3] + l[5] : l[1] + l[7];
          const T m = east ? l[4] : l[8];
          const T s1 = T(2.0 / 9.0) * a + T(8.0 / 9.0) * m;
          const T s2 = T(1.0 / 18.0) * a + T(2.0 / 9.0) * m;
          c1 = east ? -s1 : s1;
          c2 = east ? s2 : -s2;
        } else {
          if constexpr (L::QT == 5) {
            constexpr int o = L::QF;
            const T lu = east ? l[o + 2] : l[o + 4];
            c1 = T(4.0 / 3.0) * lu;
            c2 = -(T(1.0 / 3.0) * lu);
          }
        }
        const int slot = e.aux + 2 * (y - e.box[2]);
        nside[slot] = double(c1);
        nside[slot + 1] = double(c2);
        bc_node_adj<L, T>(l, e);  // the node's own block: that of the Zou-He velocity / temperature condition
        continue;
      }
    }
    bc_node_adj<L, T>(l, e);
  }
}

// After an adjoint sweep: the Neumann transfers of one piece added into the stored adjoints
// m of the two nodes inwards of each of its nodes (flow: c_qx times the coefficient on every
// direction; temperature: the coefficient on every D2Q5 direction). One thread per node of
// the piece; the neighbours of different nodes are different, and pieces run one after another.
template <class L, class P>
__global__ __launch_bounds__(64) void k_neumann_fix(void* lam, Layout ll, DevBC e, const double* nside) {
  using LS = typename P::LS;
  const int y = e.box[2] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (y > e.box[3]) return;
  const int slot = e.aux + 2 * (y - e.box[2]);
  const int dir = -e.nrm[0];
  const bool flow = e.type == LBMX_BC_ZOUHE_NEUMANN;
  for (int k = 1; k <= 2; ++k) {
    const LS c = LS(nside[slot + k - 1]);
    const unsigned int ob = layout_off(ll, e.box[0] + k * dir, y, e.box[4]) * (unsigned int)sizeof(LS);
    static_for<0, L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      char* base = own_base_w<q, LS>(lam, ll);
      if constexpr (q < L::QF && L::ex[q] != 0) {
        if (flow) st<LS>(base, ob, ld<LS>(base, ob) + LS(L::ex[q]) * c);
      } else if constexpr (q >= L::QF) {
        if (!flow) st<LS>(base, ob, ld<LS>(base, ob) + c);
      }
    });
  }
}

// The objective of a sweep is summed per 64-node row segment: the wave of a segment that
// meets a functional term reduces its nodes' values across the wave and adds the sum to the
// segment's accumulator (one writer per segment and sweep, sweeps in stream order), and the host
// adds the accumulators in segment order. The result does not depend on the workgroup tile, and
// no workgroup barrier is needed.
// The tests below combine with bitwise, not short-circuit, operators: every wave evaluates them
// before its pulls, and a short-circuit chain would read the term boxes from the argument
// segment one after another, each read waiting for the previous test. Their coordinates are
// the wave's (segment_coords), so that they run on the scalar unit.
// first x, y and z of the calling wave's row segment, in scalar registers (a wave covers one
// 64-node row segment)
__device__ __forceinline__ void segment_coords(int x, int y, int z, int& x0, int& ys, int& zs) {
  x0 = __builtin_amdgcn_readfirstlane(x - (int)threadIdx.x);
  ys = __builtin_amdgcn_readfirstlane(y);
  zs = __builtin_amdgcn_readfirstlane(z);
}
__device__ __forceinline__ bool segment_meets_box(const int* bb, int x0, int y, int z) {
  return (x0 + kTileX - 1 >= bb[0]) & (x0 <= bb[1]) & (y >= bb[2]) & (y <= bb[3]) & (z >= bb[4]) & (z <= bb[5]);
}
// (the term loops are unrolled over kMaxTerms: a runtime index into the by-value box array
// would place the argument structure in scratch memory)
__device__ __forceinline__ bool segment_meets_terms(const ObjArgs& o, int x0, int y, int z) {
  bool meets = false;
  static_for<0, kMaxTerms>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    meets = meets | ((m < o.nterm) & segment_meets_box(o.tbox[m], x0, y, z));
  });
  return meets;
}
// Per node, on the vector unit. UNIFORM_COUNT: each term is tested under the uniform condition
// m < nterm, so a functional of one or two terms does not pay for the unused slots (the
// single-instance sweeps, whose term count arrives with their arguments). Otherwise every slot
// is tested branch-free, which the batched sweeps keep: their term count comes from the member
// table, and branching on it made the robust ensemble's gradient 1.1% slower.
template <bool UNIFORM_COUNT>
__device__ __forceinline__ bool in_some_term(const ObjArgs& o, int x, int y, int z) {
  bool in = false;
  static_for<0, kMaxTerms>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if constexpr (UNIFORM_COUNT) {
      if (m < o.nterm) in = in | in_box(o.tbox[m], x, y, z);
    } else {
      in = in | ((m < o.nterm) & in_box(o.tbox[m], x, y, z));
    }
  });
  return in;
}
// Wave sum of v over the lanes whose node lies in a term (`contrib`; v is 0 elsewhere), added
// to acc[seg] by one lane. A segment crossed by a face term has one such lane, which adds its
// value itself: the tree sum would return that value unchanged, since the other lanes add
// exact zeros. The addition is an atomic without return, so the wave does not wait for the
// accumulator's read; each segment has one adder per sweep and sweeps run in stream order, so
// the additions into acc[seg] happen in sweep order. Every lane of the wave calls it.
// Summation order: x-adjacent node pairs (2i, 2i+1) first, then a tree over the 32 pair sums
// (pair i with pair i + 16, and so on), the order in which the two-node sweep sums them.
__device__ __forceinline__ void add_segment_sum(double* acc, long long seg, double v, bool contrib) {
  LBMX_CONTRACT_IN_STATEMENT
  const unsigned long long who = __ballot(contrib);
  if (who == 0ull) return;
  if ((who & (who - 1)) == 0ull) {
    if (contrib) __hip_atomic_fetch_add(acc + seg, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  v += __shfl_xor(v, 1, 64);
  for (int off = 32; off > 1; off >>= 1) v += __shfl_down(v, off, 64);
  if (threadIdx.x == 0) __hip_atomic_fetch_add(acc + seg, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The same for a wave of the two-node sweep: lane i holds the pair sum v of nodes 2i, 2i+1 of
// a 128-node stretch, lanes 0-31 for the segment seg0, lanes 32-63 for seg1 (inA/inB: the
// lane's nodes that lie in a term). Each half sums as add_segment_sum does.
__device__ __forceinline__ void add_segment_sums_x2(double* acc, long long seg0, long long seg1, double v, bool inA,
                                                    bool inB) {
  LBMX_CONTRACT_IN_STATEMENT
  const unsigned long long wa = __ballot(inA), wb = __ballot(inB);
  if ((wa | wb) == 0ull) return;
  const int n0 = __builtin_popcountll(wa & 0xffffffffull) + __builtin_popcountll(wb & 0xffffffffull);
  const int n1 = __builtin_popcountll(wa >> 32) + __builtin_popcountll(wb >> 32);
  const bool hi = threadIdx.x >= 32;
  const int n = hi ? n1 : n0;
  const long long seg = hi ? seg1 : seg0;
  // a half with one contributing node: its lane adds v (the other node of the lane adds 0)
  if (n == 1 && (inA | inB)) __hip_atomic_fetch_add(acc + seg, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if ((n0 > 1) | (n1 > 1)) {  // (uniform)
    for (int off = 16; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (n > 1 && (threadIdx.x & 31) == 0) __hip_atomic_fetch_add(acc + seg, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The node's functional value from its pre-collision macros (the collision returns rho and
// j, so they are not summed again); the sweeps add the segment's sum after the node's stores.
// (UNIFORM_COUNT as for in_some_term)
template <int COLL, class T, bool UNIFORM_COUNT>
__device__ __forceinline__ T node_objective(const ObjArgs& o, T rho, const T (&j)[3], T temp, int x, int y, int z) {
  LBMX_CONTRACT_IN_STATEMENT
  T acc = 0;
  static_for<0, kMaxTerms>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if constexpr (UNIFORM_COUNT) {
      if (m < o.nterm) {
        if (in_box(o.tbox[m], x, y, z)) acc += T(o.tweight[m]) * term_value<COLL, T>(o.tkind[m], rho, j, temp);
      }
    } else {
      if ((m < o.nterm) & in_box(o.tbox[m], x, y, z))
        acc += T(o.tweight[m]) * term_value<COLL, T>(o.tkind[m], rho, j, temp);
    }
  });
  return acc * T(o.time_const);
}

// One forward sweep over the node (x,y,z) of workgroup tile `tile`; called by every thread of
// the workgroup (the objective reduction synchronises it). tfac: time factor per boundary piece.
// BATCH: called by the batched sweep (member table; see in_some_term)
template <class L, int COLL, int DES, class P, int TY, bool BATCH = false>
__device__ __forceinline__ void fwd_node(const FwdArgs& a, const double* tfac, int x, int y, int z, int tile) {
  using T = typename P::T;
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const bool active = x < nx && y < ny;
  int x0, ys, zs;
  segment_coords(x, y, z, x0, ys, zs);
  const bool objseg = a.obj.acc != nullptr && a.mode == 0 && (ys < ny) & segment_meets_terms(a.obj, x0, ys, zs);
  const bool bseg = a.nbc > 0 && a.mode != 2 && segment_on_boundary<L>(x0, ys, zs, nx, ny, nz);
  double objv = 0.0;
  // (objseg is the wave's: the per-node test is skipped by whole waves)
  const bool objnode = objseg && (active & in_some_term<!BATCH>(a.obj, x, y, z));
  const unsigned int off = active ? layout_off(a.ls, x, y, z) : 0u;
  const int n = (z * ny + y) * nx + x;
  const bool bnd = active && on_boundary<L>(x, y, z, nx, ny, nz);
  unsigned long long rowword = 0;
  T f[L::Q];
  T gam = T(0);
  if (active) {
    // the row word and the design value are issued with the pulls, so their latency overlaps
    rowword = bc_row_word(a.bcrows, nx, ny, x, y, z);
    gam = T(static_cast<const typename P::G*>(a.gamma)[n]);
    if (a.mode == 2) {
      const double* s = static_cast<const double*>(a.src);
      static_for<0, L::Q>([&](auto qc) { f[decltype(qc)::value] = T(s[(long long)n * L::Q + decltype(qc)::value]); });
    } else {
      pull_node<L, P, 2>(a.src, a.ls, off, f);
    }
  }
  const unsigned long long pieces = bseg ? bc_pieces(rowword) : 0ull;
  if (bseg) stage_wave_pieces(a.bcs, tfac, a.nbc, pieces);
  if (active) {
    if (bnd && a.mode != 2) apply_bcs<L, P>(f, staged_pieces(), pieces, staged_tfac(a.nbc), x, y, z, a.src, a.ls);
    if (a.mode == 1) {
      double* out = static_cast<double*>(a.aos_out);
      static_for<0, L::Q>([&](auto qc) { out[(long long)n * L::Q + decltype(qc)::value] = double(f[decltype(qc)::value]); });
    }
  }
  if (active && a.mode != 1) {
    const T temp = objnode ? node_temperature<L>(f) : T(0);
    T rho, j[3];
    constexpr bool tw = kStoreTwice<P, COLL, DES>;
    collide<L, COLL, DES, T, tw>(f, gam, a.m, rho, j);
    if (objnode) objv = double(node_objective<COLL, T, !BATCH>(a.obj, rho, j, temp, x, y, z));
    store_node<L, P, tw>(a.dst, a.ls, off, f, bnd, x, y, z, nx, ny, nz, a.t);
  }
  if (objseg) add_segment_sum(a.obj.acc, bc_row_index(nx, ny, x0, y, z), objv, objnode);
}

template <class L, int COLL, int DES, class P>
__global__ __launch_bounds__(kTileX* kTileY) void k_fwd(FwdArgs a) {
  int x, y, z, tile;
  tile_coords(a.g, x, y, z, tile);
  fwd_node<L, COLL, DES, P, kTileY>(a, a.tfac, x, y, z, tile);
}

// ---- forward sweep of the 16-bit storage forms, two nodes per lane ----
// With 2 bytes per population, a wave's load or store of one node per lane carries 128 bytes,
// and the work done once per wave (tile coordinates, functional and boundary tests, plane
// bases) is spread over 64 nodes. Here lane i takes the x-adjacent nodes 2i and 2i + 1 of a
// 128-node row stretch (two row segments). x = 0 sits at an even element offset and rows have
// an even length, so the pair's elements of a plane form one aligned 32-bit word: the planes
// without an x shift are read, and all planes written, as such words; the x-shifted planes
// are read as two 16-bit halves (which the compiler merges into one unaligned 32-bit load).
// The word of each plane holds the first node's value in its low half and the second node's
// in its high half, and each node's result replaces its own half once that node has been
// decoded, so one register per plane serves both nodes. Each
// node's arithmetic is that of fwd_node (same boundary pieces, collision, functional value and
// rounding seeds), so the stored state is bitwise the one-node sweep's, and the functional's
// segment sums add in the same order. (A boundary node's halo reflections go to cells outside
// the domain, which no thread's own stores touch, so the first node may write them before
// the pair's words are stored.)
template <class L, int COLL, int DES, class P, bool HI>
__device__ __forceinline__ void fwd_half_node(const FwdArgs& a, unsigned int (&w)[L::Q], typename P::T gam, bool bnd,
                                              unsigned long long pieces, bool objnode, double& objv, int x, int y,
                                              int z, unsigned int off) {
  using T = typename P::T;
  T f[L::Q];
  static_for<0, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[q] = P::template dec<L, q>(P::template half_of<HI>(w[q]));
  });
  if (bnd) apply_bcs<L, P>(f, staged_pieces(), pieces, staged_tfac(a.nbc), x, y, z, a.src, a.ls);
  const T temp = objnode ? node_temperature<L>(f) : T(0);
  T rho, j[3];
  constexpr bool tw = kStoreTwice<P, COLL, DES>;
  collide<L, COLL, DES, T, tw>(f, gam, a.m, rho, j);
  if (objnode) objv = double(node_objective<COLL, T, true>(a.obj, rho, j, temp, x, y, z));
  unsigned int r0, r1;
  sr_words(off, a.t, r0, r1);
  static_for<0, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    w[q] = P::template enc_sr<L, q, HI, tw>(w[q], f[q], sr_seed<q, P::kRandomBits>(r0, r1));
  });
  if (bnd) {
    using S = typename P::S;
    auto enc = [&](auto qc) -> S { return P::template half_of<HI>(w[decltype(qc)::value]); };
    store_halo<L, S>(a.dst, a.ls, off * (unsigned int)sizeof(S), x, y, z, a.g.nx, a.g.ny, a.g.nz, enc);
  }
}

// What a two-node lane loads before it computes: each plane's word for its pair, the pair's
// design values and its segment's boundary row word.
template <class L, class P>
struct X2Loads {
  unsigned int w[L::Q];
  typename P::T gA, gB;
  unsigned long long rowword;
};

template <class L, class P>
__device__ __forceinline__ void x2_load(const FwdArgs& a, int xa, int y, int z, X2Loads<L, P>& ld) {
  using T = typename P::T;
  using S = typename P::S;
  const int nx = a.g.nx, ny = a.g.ny;
  const bool actA = xa < nx && y < ny, actB = xa + 1 < nx && y < ny;
  ld.rowword = 0;
  ld.gA = ld.gB = T(0);
  if (actA) {
    const unsigned int ob = layout_off(a.ls, xa, y, z) * (unsigned int)sizeof(S);
    ld.rowword = bc_row_word(a.bcrows, nx, ny, xa, y, z);
    const typename P::G* gp = static_cast<const typename P::G*>(a.gamma) + ((z * ny + y) * nx + xa);
    ld.gA = T(gp[0]);
    if (actB) ld.gB = T(gp[1]);
    static_for<0, L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const char* base = pull_base<L, q, S>(a.src, a.ls) + ob;
      if constexpr (L::ex[q] == 0) {
        ld.w[q] = __builtin_nontemporal_load(reinterpret_cast<const unsigned int*>(base));
      } else {
        const unsigned int lo = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(base));
        const unsigned int up = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(base) + 1);
        ld.w[q] = lo | (up << 16);
      }
    });
  }
}

// Everything after the loads: boundary pieces, both nodes, the stores and the functional's
// segment sums.
template <class L, int COLL, int DES, class P>
__device__ __forceinline__ void x2_finish(const FwdArgs& a, int xa, int y, int z, X2Loads<L, P>& ld) {
  using S = typename P::S;
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const int xb = xa + 1;
  const bool actA = xa < nx && y < ny, actB = xb < nx && y < ny;
  // the wave's two row segments start at x0 and x1 = x0 + 64 (lanes 0-31 and 32-63)
  const int x0 = __builtin_amdgcn_readfirstlane(xa - 2 * (int)threadIdx.x);
  const int ys = __builtin_amdgcn_readfirstlane(y), zs = __builtin_amdgcn_readfirstlane(z);
  const int x1 = x0 + kTileX;
  const bool hi = threadIdx.x >= 32;
  const bool objseg0 = a.obj.acc != nullptr && (ys < ny) & segment_meets_terms(a.obj, x0, ys, zs);
  const bool objseg1 = a.obj.acc != nullptr && (ys < ny) & (x1 < nx) & segment_meets_terms(a.obj, x1, ys, zs);
  const bool bseg0 = a.nbc > 0 && segment_on_boundary<L>(x0, ys, zs, nx, ny, nz);
  const bool bseg1 = a.nbc > 0 && (x1 < nx) & segment_on_boundary<L>(x1, ys, zs, nx, ny, nz);
  const bool objseg = hi ? objseg1 : objseg0;
  const bool objA = objseg && (actA & in_some_term<true>(a.obj, xa, y, z));
  const bool objB = objseg && (actB & in_some_term<true>(a.obj, xb, y, z));
  const unsigned int off = actA ? layout_off(a.ls, xa, y, z) : 0u;
  const unsigned int ob = off * (unsigned int)sizeof(S);
  const bool bndA = actA && on_boundary<L>(xa, y, z, nx, ny, nz);
  const bool bndB = actB && on_boundary<L>(xb, y, z, nx, ny, nz);
  // the pieces of both segments (a piece that meets only the other segment holds neither node)
  unsigned long long pieces = 0;
  if (bseg0 | bseg1) {
    const unsigned int l0 = __builtin_amdgcn_readlane(static_cast<unsigned int>(ld.rowword), 0);
    const unsigned int h0 = __builtin_amdgcn_readlane(static_cast<unsigned int>(ld.rowword >> 32), 0);
    const unsigned int l1 = __builtin_amdgcn_readlane(static_cast<unsigned int>(ld.rowword), 32);
    const unsigned int h1 = __builtin_amdgcn_readlane(static_cast<unsigned int>(ld.rowword >> 32), 32);
    const unsigned long long p0 = bseg0 ? ((static_cast<unsigned long long>(h0) << 32) | l0) : 0ull;
    const unsigned long long p1 = bseg1 ? ((static_cast<unsigned long long>(h1) << 32) | l1) : 0ull;
    pieces = p0 | p1;
    stage_wave_pieces(a.bcs, a.tfac, a.nbc, pieces);
  }
  double vA = 0.0, vB = 0.0;
  if (actA) fwd_half_node<L, COLL, DES, P, false>(a, ld.w, ld.gA, bndA, pieces, objA, vA, xa, y, z, off);
  if (actB) fwd_half_node<L, COLL, DES, P, true>(a, ld.w, ld.gB, bndB, pieces, objB, vB, xb, y, z, off + 1u);
  if (actA) {
    if (actB) {
      static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        st_nt<unsigned int>(own_base_w<q, S>(a.dst, a.ls), ob, ld.w[q]);
      });
    } else {
      static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        st_nt<S>(own_base_w<q, S>(a.dst, a.ls), ob, P::template half_of<false>(ld.w[q]));
      });
    }
  }
  if (objseg0 | objseg1)
    add_segment_sums_x2(a.obj.acc, bc_row_index(nx, ny, x0, ys, zs), bc_row_index(nx, ny, x1, ys, zs), vA + vB, objA,
                        objB);
}

template <class L, int COLL, int DES, class P>
__device__ __forceinline__ void fwd_node_x2(const FwdArgs& a, int xa, int y, int z) {
  static_assert(sizeof(typename P::S) == 2 && P::stochastic, "two-node sweep of the 16-bit storage forms");
  X2Loads<L, P> ld;
  x2_load<L, P>(a, xa, y, z, ld);
  x2_finish<L, COLL, DES, P>(a, xa, y, z, ld);
}

template <class L, int COLL, int DES, class P>
__global__ __launch_bounds__(kTileX* kTileY) void k_fwd_x2(FwdArgs a) {
  int x, y, z, tile;
  tile_coords(a.g, x, y, z, tile);
  // tile_coords places lane i at x = 64 bx + i; this sweep's tiles are 128 nodes wide
  const int bx = (x - (int)threadIdx.x) / kTileX;
  fwd_node_x2<L, COLL, DES, P>(a, 2 * kTileX * bx + 2 * (int)threadIdx.x, y, z);
}

// Initial state: s_0 = f_eq(rho, u[, T]) from uniform or per-node macros; writes p_0 = C(s_0)
// (IsoEquilibriumInitializationLoop2d, src/dynamics/isothermalLBLoops2d.hh:264-285).
template <class L, int COLL, int DES, class P>
__global__ __launch_bounds__(kTileX* kTileY) void k_init(FwdArgs a, const double* macros, double rho0,
                                                         double ux0, double uy0, double uz0, double T0) {
  using T = typename P::T;
  int x, y, z, tile;
  tile_coords(a.g, x, y, z, tile);
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  if (x >= nx || y >= ny) return;
  const int n = (z * ny + y) * nx + x;
  double r = rho0, u[3] = {ux0, uy0, uz0}, tmp = T0;
  if (macros) {
    r = macros[5 * (long long)n]; u[0] = macros[5 * (long long)n + 1]; u[1] = macros[5 * (long long)n + 2];
    u[2] = macros[5 * (long long)n + 3]; tmp = macros[5 * (long long)n + 4];
  }
  constexpr bool inc = CollTraits<COLL>::inc;
  double fd[L::Q];
  const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  static_for<0, L::QF>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    double cu = cdot<L, q>(u);
    double poly = 3.0 * cu + 4.5 * cu * cu - 1.5 * u2;
    fd[q] = inc ? L::w[q] * (r + poly) : L::w[q] * r * (1.0 + poly);
  });
  if constexpr (L::QT > 0) {
    // ThermalMRTFeq2d (src/dynamics/thermalEquilibriumDistributions2d.hh:28-47): f = M^-1 m_eq
    const double meq[5] = {tmp, u[0] * tmp, u[1] * tmp, -2.0 * tmp, 0.0};
    static_for<0, 5>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      double s = 0;
      static_for<0, 5>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (kB5.M[k][q] != 0) s += kB5.M[k][q] * kB5.invn2[k] * meq[k];
      });
      fd[L::QF + q] = s;
    });
  }
  if (a.aos_out) {
    double* out = static_cast<double*>(a.aos_out);
    static_for<0, L::Q>([&](auto qc) { out[(long long)n * L::Q + decltype(qc)::value] = fd[decltype(qc)::value]; });
  }
  if (a.dst) {
    T f[L::Q];
    static_for<0, L::Q>([&](auto qc) { f[decltype(qc)::value] = T(fd[decltype(qc)::value]); });
    const T gam = T(static_cast<const typename P::G*>(a.gamma)[n]);
    T rho, j[3];
    constexpr bool tw = kStoreTwice<P, COLL, DES>;
    collide<L, COLL, DES, T, tw>(f, gam, a.m, rho, j);
    store_node<L, P, tw>(a.dst, a.ls, layout_off(a.ls, x, y, z), f, on_boundary<L>(x, y, z, nx, ny, nz), x, y, z, nx,
                     ny, nz, 0);
  }
}

// One adjoint sweep over node (x,y,z); tfac: time factor per boundary piece at step a.t.
// BATCH: called by the batched sweep (member table; see in_some_term)
template <class L, int COLL, int DES, class P, int TY, bool BATCH = false>
__device__ __forceinline__ void adj_node(const AdjArgs& a, const double* tfac, int x, int y, int z) {
  using T = typename P::T;
  using LS = typename P::LS;
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const bool active = x < nx && y < ny;
  int x0, ys, zs;
  segment_coords(x, y, z, x0, ys, zs);
  const bool bseg = a.nbc > 0 && segment_on_boundary<L>(x0, ys, zs, nx, ny, nz);
  const bool objseg = a.obj.nterm > 0 && segment_meets_terms(a.obj, x0, ys, zs);
  const int n = (z * ny + y) * nx + x;
  const bool bnd = active && on_boundary<L>(x, y, z, nx, ny, nz);
  unsigned long long rowword = 0;
  T f[L::Q];
  if (active) {
    rowword = bc_row_word(a.bcrows, nx, ny, x, y, z);
    if (a.fwd_aos) {
      static_for<0, L::Q>([&](auto qc) { f[decltype(qc)::value] = T(a.fwd_aos[(long long)n * L::Q + decltype(qc)::value]); });
    } else {
      pull_node<L, P, 1>(a.fwd_src, a.ls, layout_off(a.ls, x, y, z), f);
    }
  }
  const unsigned long long pieces = bseg ? bc_pieces(rowword) : 0ull;
  if (bseg) stage_wave_pieces(a.bcs, tfac, a.nbc, pieces);
  if (!active) return;
  const DevBC* const sbcs = staged_pieces();
  if (bnd && !a.fwd_aos) apply_bcs<L, P>(f, sbcs, pieces, staged_tfac(a.nbc), x, y, z, a.fwd_src, a.ls);
  const unsigned int offl = layout_off(a.ll, x, y, z) * (unsigned int)sizeof(LS);
  // The adjoint field between sweeps is m_t, the boundary-corrected adjoint of s_t before the
  // transposed stream, stored at its own node. The transposed stream is gathered here:
  // nu_t(x)[q] = m_t(x + c_q)[q], or the mid-way reflection m_t(x)[opp q] when x + c_q lies
  // outside the domain (then the gathered halo value is replaced).
  T lam[L::Q];
  static_for<0, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    lam[q] = T(ld<LS>(gather_base<L, q, LS>(a.lam_in, a.ll), offl));
  });
  if (bnd) {
    static_for<1, L::Q>([&](auto qc) {
      constexpr int r = decltype(qc)::value;
      if constexpr (L::ex[r] != 0 || L::ey[r] != 0 || L::ez[r] != 0) {
        if (leaves<L, r>(x, y, z, nx, ny, nz)) lam[r] = T(ld<LS>(own_base<L::opp[r], LS>(a.lam_in, a.ll), offl));
      }
    });
  }
  T srcA = 0, srcB[3] = {0, 0, 0}, srcAT = 0;
  if (objseg && in_some_term<!BATCH>(a.obj, x, y, z)) {
    T rho, j[3];
    flow_macros<L>(f, rho, j);
    T temp = 0;
    if constexpr (L::QT > 0) temp = f[L::QF] + f[L::QF + 1] + f[L::QF + 2] + f[L::QF + 3] + f[L::QF + 4];
    static_for<0, kMaxTerms>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      if ((m < a.obj.nterm) & in_box(a.obj.tbox[m], x, y, z))
        term_source<T>(a.obj.tkind[m], rho, j, temp, T(a.obj.time_const * a.obj.tweight[m]), srcA, srcB,
                       srcAT);
    });
  }
  const T gam = T(static_cast<const typename P::G*>(a.gamma)[n]);
  T dg = 0;
  reverse_collide<L, COLL, DES, T>(f, gam, a.m, lam, dg, srcA, srcB, srcAT);
  __builtin_nontemporal_store(__builtin_nontemporal_load(a.sens + n) + double(dg), a.sens + n);
  if (bnd) apply_bcs_adj<L, T>(lam, sbcs, pieces, x, y, z, a.nside);
  // m_{t-1} at the own node (whole rows); the next adjoint sweep gathers its transposed stream
  static_for<0, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    st_nt<LS>(own_base_w<q, LS>(a.lam_out, a.ll), offl, LS(lam[q]));
  });
}

template <class L, int COLL, int DES, class P>
__global__ __launch_bounds__(kTileX* kTileY) void k_adj(AdjArgs a) {
  int x, y, z, tile;
  tile_coords(a.g, x, y, z, tile);
  adj_node<L, COLL, DES, P, kTileY>(a, a.tfac, x, y, z);
}

// ---- batched sweeps: one launch covers every instance of a batch ------------------------
// The global workgroup id is remapped across XCDs as for a single grid, then mapped to
// (instance, local tile) through blk_inst; each instance's static description (grid,
// layouts, offsets into the batch buffers, design, sensitivity, boundary and functional
// arrays) comes from a device table, the time factors from a per-step table.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int qd = nb >> 3, r = nb & 7;
  const int xcd = b & 7, k = b >> 3;
  return xcd < r ? xcd * (qd + 1) + k : r * (qd + 1) + (xcd - r) * qd + k;
}

// The member table is written by the host before the batch runs and is read-only for every
// kernel, so it is read through the constant address space: its fields are then scalar
// loads wherever the compiler places them (a read of the global table after the sweep's first
// store could otherwise be a vector load, as the store might alias it).
using ConstInst = const __attribute__((address_space(4))) BatchInst;
// a structure of the member table, copied out of the constant address space
template <class T>
__device__ __forceinline__ T from_const(const __attribute__((address_space(4))) T& src) {
  T v;
  __builtin_memcpy(&v, (const T*)&src, sizeof(T));
  return v;
}

template <int TY>
__device__ __forceinline__ ConstInst& batch_locate(const BatchStep& st, int& inst, int& x, int& y, int& z,
                                                   int& tile) {
  const int b = xcd_remap(blockIdx.x, st.nblocks);
  inst = st.blk_inst[b];
  ConstInst& I = *((ConstInst*)(st.inst) + inst);
  tile = b - I.blk_begin;
  const int bx = tile % I.g.nblk_x;
  const int rest = tile / I.g.nblk_x;
  const int by = rest % I.g.nblk_y;
  z = rest / I.g.nblk_y;
  x = bx * kTileX + threadIdx.x;
  y = by * TY + threadIdx.y;
  return I;
}

template <class L, int COLL, int DES, class P, int TY>
__global__ __launch_bounds__(kTileX* TY) void k_fwd_batch(BatchStep st) {
  int i, x, y, z, tile;
  ConstInst& I = batch_locate<TY>(st, i, x, y, z, tile);
  FwdArgs a;
  a.src = static_cast<const char*>(st.src) + I.state_off;
  a.dst = static_cast<char*>(st.dst) + I.state_off;
  a.gamma = I.gamma;
  a.bcs = I.bcs;
  a.bcrows = I.bcrows;
  a.nbc = I.nbc;
  a.t = st.t;
  a.mode = 0;
  a.aos_out = nullptr;
  a.g = from_const(I.g);
  a.ls = from_const(I.ls);
  a.m = from_const(I.m);
  a.obj = from_const(I.obj);
  if (!st.with_obj) a.obj.acc = nullptr;
  fwd_node<L, COLL, DES, P, TY, true>(a, st.tfac + (long long)i * st.tf_stride, x, y, z, tile);
}

template <class L, int COLL, int DES, class P, int TY>
__global__ __launch_bounds__(kTileX* TY) void k_adj_batch(BatchStep st) {
  int i, x, y, z, tile;
  ConstInst& I = batch_locate<TY>(st, i, x, y, z, tile);
  AdjArgs a;
  a.fwd_src = static_cast<const char*>(st.src) + I.state_off;
  a.fwd_aos = nullptr;
  a.nside = nullptr;
  a.lam_in = static_cast<const char*>(st.lam_in) + I.lam_off;
  a.lam_out = static_cast<char*>(st.lam_out) + I.lam_off;
  a.gamma = I.gamma;
  a.sens = I.sens;
  a.bcs = I.bcs;
  a.bcrows = I.bcrows;
  a.nbc = I.nbc;
  a.t = st.t;
  a.g = from_const(I.g);
  a.ls = from_const(I.ls);
  a.ll = from_const(I.ll);
  a.m = from_const(I.m);
  a.obj = from_const(I.obj);
  a.obj.acc = nullptr;
  if (!st.with_obj) a.obj.nterm = 0;
  adj_node<L, COLL, DES, P, TY, true>(a, st.tfac + (long long)i * st.tf_stride, x, y, z);
}

// Batched forward sweep with two nodes per thread: a tile of 2H rows is swept by H waves, each
// thread taking rows y and y + H. Both nodes' loads (pulls, boundary row word, design) are
// issued before either node collides, which doubles the bytes in flight per wave at the same
// occupancy; each node's arithmetic is that of fwd_node, in the same order. (The adjoint,
// which already has three fields in flight per node, keeps one node per thread: paired, it
// was slower.)
// Rows y and y + H of the node's tile (2H rows, H waves).
template <class L, int COLL, int DES, class P, int H>
__device__ __forceinline__ void fwd_pair(const FwdArgs& a, const double* tfac, int x, int y, int z, int tile) {
  using T = typename P::T;
  const int nx = a.g.nx, ny = a.g.ny;
  const int y2 = y + H;
  const bool act1 = x < nx && y < ny, act2 = x < nx && y2 < ny;
  int x0, ys, zs;
  segment_coords(x, y, z, x0, ys, zs);
  const int ys2 = ys + H;
  const bool objseg1 = a.obj.acc != nullptr && (ys < ny) & segment_meets_terms(a.obj, x0, ys, zs);
  const bool objseg2 = a.obj.acc != nullptr && (ys2 < ny) & segment_meets_terms(a.obj, x0, ys2, zs);
  const bool on1 = objseg1 && (act1 & in_some_term<false>(a.obj, x, y, z));
  const bool on2 = objseg2 && (act2 & in_some_term<false>(a.obj, x, y2, z));
  const typename P::G* gp = static_cast<const typename P::G*>(a.gamma);
  T f1[L::Q], f2[L::Q], g1 = 0, g2 = 0;
  unsigned long long rw1 = 0, rw2 = 0;
  unsigned int off1 = 0, off2 = 0;
  if (act1) {
    off1 = layout_off(a.ls, x, y, z);
    rw1 = bc_row_word(a.bcrows, nx, ny, x, y, z);
    pull_node<L, P, 2>(a.src, a.ls, off1, f1);
    g1 = T(gp[(z * ny + y) * nx + x]);
  }
  if (act2) {
    off2 = layout_off(a.ls, x, y2, z);
    rw2 = bc_row_word(a.bcrows, nx, ny, x, y2, z);
    pull_node<L, P, 2>(a.src, a.ls, off2, f2);
    g2 = T(gp[(z * ny + y2) * nx + x]);
  }
  const bool bseg1 = a.nbc > 0 && segment_on_boundary<L>(x0, ys, zs, nx, ny, a.g.nz);
  const bool bseg2 = a.nbc > 0 && segment_on_boundary<L>(x0, ys2, zs, nx, ny, a.g.nz);
  const unsigned long long p1 = bseg1 ? bc_pieces(rw1) : 0ull, p2 = bseg2 ? bc_pieces(rw2) : 0ull;
  if (bseg1 || bseg2) stage_wave_pieces(a.bcs, tfac, a.nbc, p1 | p2);
  double obj1 = 0.0, obj2 = 0.0;
  const bool bnd1 = act1 && on_boundary<L>(x, y, z, nx, ny, a.g.nz);
  const bool bnd2 = act2 && on_boundary<L>(x, y2, z, nx, ny, a.g.nz);
  if (bnd1) apply_bcs<L, P>(f1, staged_pieces(), p1, staged_tfac(a.nbc), x, y, z, a.src, a.ls);
  if (bnd2) apply_bcs<L, P>(f2, staged_pieces(), p2, staged_tfac(a.nbc), x, y2, z, a.src, a.ls);
  T rho, j[3];
  if (act1) {
    const T temp = on1 ? node_temperature<L>(f1) : T(0);
    collide<L, COLL, DES, T, kStoreTwice<P, COLL, DES>>(f1, g1, a.m, rho, j);
    if (on1) obj1 = double(node_objective<COLL, T, false>(a.obj, rho, j, temp, x, y, z));
    store_node<L, P, kStoreTwice<P, COLL, DES>>(a.dst, a.ls, off1, f1, bnd1, x, y, z, nx, ny, a.g.nz, a.t);
  }
  if (act2) {
    const T temp = on2 ? node_temperature<L>(f2) : T(0);
    collide<L, COLL, DES, T, kStoreTwice<P, COLL, DES>>(f2, g2, a.m, rho, j);
    if (on2) obj2 = double(node_objective<COLL, T, false>(a.obj, rho, j, temp, x, y2, z));
    store_node<L, P, kStoreTwice<P, COLL, DES>>(a.dst, a.ls, off2, f2, bnd2, x, y2, z, nx, ny, a.g.nz, a.t);
  }
  if (objseg1) add_segment_sum(a.obj.acc, bc_row_index(nx, ny, x0, y, z), obj1, on1);
  if (objseg2) add_segment_sum(a.obj.acc, bc_row_index(nx, ny, x0, y2, z), obj2, on2);
}

template <class L, int COLL, int DES, class P, int H>
__global__ __launch_bounds__(kTileX* H) void k_fwd_batch_pair(BatchStep st) {
  int i, x, y, z, tile;
  ConstInst& I = batch_locate<2 * H>(st, i, x, y, z, tile);
  FwdArgs a;
  a.src = static_cast<const char*>(st.src) + I.state_off;
  a.dst = static_cast<char*>(st.dst) + I.state_off;
  a.gamma = I.gamma;
  a.bcs = I.bcs;
  a.bcrows = I.bcrows;
  a.nbc = I.nbc;
  a.t = st.t;
  a.mode = 0;
  a.aos_out = nullptr;
  a.g = from_const(I.g);
  a.ls = from_const(I.ls);
  a.m = from_const(I.m);
  a.obj = from_const(I.obj);
  if (!st.with_obj) a.obj.acc = nullptr;
  fwd_pair<L, COLL, DES, P, H>(a, st.tfac + (long long)i * st.tf_stride, x, y, z, tile);
}

// ---- host-side launchers ----
// The FP64 forward sweeps run best at 4 waves per SIMD: the 384 x 192 x 192 D3Q19 sweep takes
// 858 us there against 882 us at the 7 its registers allow (3 is slower again), and the 2D
// sweeps are unchanged. Each workgroup (one wave per SIMD) reserves a quarter of the CU's
// 160 KB of LDS, which caps the CU at 4 workgroups.
constexpr size_t kFwdF64Lds = 40 * 1024;

// The 16-bit forms sweep forward with two nodes per lane (k_fwd_x2) on grids of at least this
// many one-node workgroups (4 per CU). Smaller grids do not fill the GPU even with one node per
// lane: halving their workgroups made the 256 x 128 channel's FP16 forward sweep 17% slower
// (7.9 against 6.7 us), while the 1024 x 512 channel (2048 workgroups) gains 14%. Both forms
// store the same bits.
constexpr int kFwdX2MinBlocks = 1024;

template <class L, int COLL, int DES, class P>
hipError_t launch_fwd_t(const FwdArgs& a, hipStream_t s) {
  size_t lds = pieces_lds_bytes(a.nbc);
  if constexpr (sizeof(typename P::T) == 8) lds = lds > kFwdF64Lds ? lds : kFwdF64Lds;
  if constexpr (sizeof(typename P::S) == 2) {
    if (a.mode == 0 && a.g.nblocks >= kFwdX2MinBlocks) {
      // tiles of 128 x kTileY nodes
      FwdArgs b = a;
      b.g.nblk_x = (a.g.nx + 2 * kTileX - 1) / (2 * kTileX);
      b.g.nblocks = b.g.nblk_x * b.g.nblk_y * a.g.nz;
      hipLaunchKernelGGL((k_fwd_x2<L, COLL, DES, P>), dim3(b.g.nblocks), dim3(kTileX, kTileY, 1), lds, s, b);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_fwd<L, COLL, DES, P>), dim3(a.g.nblocks), dim3(kTileX, kTileY, 1), lds, s, a);
  return hipGetLastError();
}

template <class L, int COLL, int DES, class P>
hipError_t launch_init_t(const FwdArgs& a, const double* macros, const double* u, double rho0, double T0,
                         hipStream_t s) {
  hipLaunchKernelGGL((k_init<L, COLL, DES, P>), dim3(a.g.nblocks), dim3(kTileX, kTileY, 1), 0, s, a, macros, rho0,
                     u[0], u[1], u[2], T0);
  return hipGetLastError();
}

template <class L, int COLL, int DES, class P>
hipError_t launch_adj_t(const AdjArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_adj<L, COLL, DES, P>), dim3(a.g.nblocks), dim3(kTileX, kTileY, 1), pieces_lds_bytes(a.nbc), s,
                     a);
  return hipGetLastError();
}

template <class L, class P>
hipError_t launch_neumann_fix_t(const DevBC& e, void* lam, const Layout& ll, const double* nside, hipStream_t s) {
  if constexpr (L::D != 2) {
    return hipErrorInvalidValue;
  } else {
    const int n = e.box[3] - e.box[2] + 1;
    if (n < 1) return hipSuccess;
    hipLaunchKernelGGL((k_neumann_fix<L, P>), dim3((n + 63) / 64), dim3(64, 1, 1), 0, s, lam, ll, e, nside);
    return hipGetLastError();
  }
}

// A tile of TY rows. The forward takes one node per thread (TY waves) for TY = 1 or when
// `pair` is off, and two per thread (TY / 2 waves) otherwise; the adjoint takes one node per
// thread (TY waves). Both forward forms cover the same tiles with the same node arithmetic.
template <class L, int COLL, int DES, class P, int TY>
hipError_t launch_batch_t(int kind, bool pair, const BatchStep& st, hipStream_t s) {
  const size_t lds = pieces_lds_bytes(st.tf_stride);  // tf_stride: the most pieces of any member
  if (kind == 0) {
    if constexpr (TY == 1) {
      hipLaunchKernelGGL((k_fwd_batch<L, COLL, DES, P, 1>), dim3(st.nblocks), dim3(kTileX, 1, 1), lds, s, st);
    } else {
      static_assert(TY % 2 == 0, "paired tiles have an even number of rows");
      if (pair)
        hipLaunchKernelGGL((k_fwd_batch_pair<L, COLL, DES, P, TY / 2>), dim3(st.nblocks), dim3(kTileX, TY / 2, 1), lds,
                           s, st);
      else
        hipLaunchKernelGGL((k_fwd_batch<L, COLL, DES, P, TY>), dim3(st.nblocks), dim3(kTileX, TY, 1), lds, s, st);
    }
  } else {
    hipLaunchKernelGGL((k_adj_batch<L, COLL, DES, P, TY>), dim3(st.nblocks), dim3(kTileX, TY, 1), lds, s, st);
  }
  return hipGetLastError();
}

}  // namespace lbmx
