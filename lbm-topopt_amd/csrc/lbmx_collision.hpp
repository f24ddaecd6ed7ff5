// Collision operators and their exact reverse operators, specialised at compile time on the
// lattice, the operator (BGK / MRT, incompressible or standard), the design coupling
// (none / partial bounce-back / Brinkman) and the arithmetic type.
//
// The flow populations are handled in "pair space": f_0, s_d = f_d + f_d' and a_d = f_d - f_d'
// for the opposite pairs (d, d' = d + HALF). Every moment-basis row is even or odd in c, so
// moments read only s (even rows) or a (odd rows) and the back transform of a row updates
// only s (by -2 M_k dm_k) or only a: half the multiply-adds of a direct M f / M^T dm. Partial bounce-back only scales
// a by (1 - 2 imult). The macros are rho = f_0 + sum s_d, j = sum c_d a_d. The moments of the
// Guo source w_q[3(c-v).F + 9(c.v)(c.F)] are contractions with compile-time tables.
#pragma once
#include <type_traits>

#include "lbmx_device.hpp"

namespace lbmx {

enum : int { D_NONE = LBMX_DESIGN_NONE, D_PBB = LBMX_DESIGN_PBB, D_BRINK = LBMX_DESIGN_BRINKMAN,
             D_DARCY = LBMX_DESIGN_DARCY };

#define LBMX_BAS BasisOf<L::QF>::B


template <class T> __device__ __forceinline__ T recip(T x) { return T(1) / x; }
// a model parameter in the arithmetic type (FP32 forms read the copy rounded on the host)
template <class T> __device__ __forceinline__ T mpar(double d, float f) {
  if constexpr (sizeof(T) == 4) return f;
  else return T(d);
}
template <> __device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// BPInterpolation (src/core/interpolationFunctions.hh:18-25)
template <class T> __device__ __forceinline__ T bp_val(T q, T s) { return T(1) - s * (T(1) + q) * recip<T>(q + s); }
template <class T> __device__ __forceinline__ T bp_der(T q, T s) {
  T r = recip<T>(q + s);
  return -(T(1) + q) * r + s * (T(1) + q) * (r * r);
}

template <class L, class T>
struct Pairs {
  T f0;
  T s[L::HALF + 1];
  T a[L::HALF + 1];
};

template <class L, class T>
__device__ __forceinline__ void to_pairs(const T (&f)[L::Q], Pairs<L, T>& p) {
  p.f0 = f[0];
  p.s[0] = p.a[0] = T(0);
  static_for<1, L::HALF + 1>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    p.s[d] = f[d] + f[d + L::HALF];
    p.a[d] = f[d] - f[d + L::HALF];
  });
}

// TWICE: hand back 2 f_q for q = 1 .. QF - 1 (the caller folds the exact factor 1/2 into its
// next multiplication)
template <class L, class T, bool TWICE = false>
__device__ __forceinline__ void from_pairs(const Pairs<L, T>& p, T (&f)[L::Q]) {
  f[0] = p.f0;
  static_for<1, L::HALF + 1>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    if constexpr (TWICE) {
      f[d] = p.s[d] + p.a[d];
      f[d + L::HALF] = p.s[d] - p.a[d];
    } else {
      f[d] = T(0.5) * (p.s[d] + p.a[d]);
      f[d + L::HALF] = T(0.5) * (p.s[d] - p.a[d]);
    }
  });
}

template <class L, class T>
__device__ __forceinline__ void pair_macros(const Pairs<L, T>& p, T& rho, T (&j)[3]) {
  T r = p.f0, jx = 0, jy = 0, jz = 0;
  static_for<1, L::HALF + 1>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    r += p.s[d];
    if constexpr (L::ex[d] == 1) jx += p.a[d];
    if constexpr (L::ex[d] == -1) jx -= p.a[d];
    if constexpr (L::ey[d] == 1) jy += p.a[d];
    if constexpr (L::ey[d] == -1) jy -= p.a[d];
    if constexpr (L::ez[d] == 1) jz += p.a[d];
    if constexpr (L::ez[d] == -1) jz -= p.a[d];
  });
  rho = r; j[0] = jx; j[1] = jy; j[2] = jz;
}

constexpr double abs_c(double x) { return x < 0 ? -x : x; }

// Row k carries a Brinkman source unless its polynomial P_k(v) = 3 v.G1_k + 9 v^T G2_k v
// - 3 |v|^2 G0_k vanishes identically (the density row: the Guo source conserves mass).
template <class L, int k>
constexpr bool row_has_src() {
  for (int a = 0; a < 3; ++a) {
    if (abs_c(LBMX_BAS.G1[k][a]) > 1e-12) return true;
    for (int b = 0; b < 3; ++b)
      if (abs_c(9.0 * LBMX_BAS.G2[k][a][b] - (a == b ? 3.0 * LBMX_BAS.G0[k] : 0.0)) > 1e-12) return true;
  }
  return false;
}

template <class L, int DES, int k>
constexpr bool row_active() {
  return LBMX_BAS.rate[k] != 0 || (DES == D_BRINK && row_has_src<L, k>());
}

// row k of M applied in pair space
template <class L, int k, class T>
__device__ __forceinline__ T pair_moment(const Pairs<L, T>& p) {
  static_assert(LBMX_BAS.par[k] != 0, "basis row without parity");
  T m = 0;
  if constexpr (LBMX_BAS.par[k] == 1) {
    if constexpr (LBMX_BAS.M[k][0] != 0) {
      constexpr T kc = T(LBMX_BAS.M[k][0]);
      m += kc * p.f0;
    }
    static_for<1, L::HALF + 1>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      if constexpr (LBMX_BAS.M[k][d] != 0) {
        constexpr T kc = T(LBMX_BAS.M[k][d]);
        m += kc * p.s[d];
      }
    });
  } else {
    static_for<1, L::HALF + 1>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      if constexpr (LBMX_BAS.M[k][d] != 0) {
        constexpr T kc = T(LBMX_BAS.M[k][d]);
        m += kc * p.a[d];
      }
    });
  }
  return m;
}

// x -= M_k^T c for one row k, in pair space
template <class L, int k, class T>
__device__ __forceinline__ void pair_row_update(Pairs<L, T>& x, T c) {
  if constexpr (LBMX_BAS.par[k] == 1) {
    if constexpr (LBMX_BAS.M[k][0] != 0) {
      constexpr T kc = T(LBMX_BAS.M[k][0]);
      x.f0 -= kc * c;
    }
    static_for<1, L::HALF + 1>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      if constexpr (LBMX_BAS.M[k][d] != 0) {
        constexpr T kc = T(2.0 * LBMX_BAS.M[k][d]);
        x.s[d] -= kc * c;
      }
    });
  } else {
    static_for<1, L::HALF + 1>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      if constexpr (LBMX_BAS.M[k][d] != 0) {
        constexpr T kc = T(2.0 * LBMX_BAS.M[k][d]);
        x.a[d] -= kc * c;
      }
    });
  }
}

// m_eq of row k (src/dynamics/incompressibleMRT2d.hh:62-67, src/dynamics/MRT3d.hh:92-109)
template <class L, int k, class T>
__device__ __forceinline__ T meq_row(T rho, const T (&v)[3], T inv_rho_s) {
  LBMX_CONTRACT_IN_STATEMENT
  T m = 0;
  if constexpr (LBMX_BAS.a[k] != 0) {
    constexpr T kc = T(LBMX_BAS.a[k]);
    m += kc * rho;
  }
  static_for<0, 3>([&](auto ac) {
    constexpr int a = decltype(ac)::value;
    if constexpr (LBMX_BAS.b[k][a] != 0) {
      constexpr T kc = T(LBMX_BAS.b[k][a]);
      m += kc * v[a];
    }
  });
  T quad = 0;
  static_for<0, 3>([&](auto ac) {
    constexpr int a = decltype(ac)::value;
    static_for<0, 3>([&](auto bc) {
      constexpr int b = decltype(bc)::value;
      if constexpr (LBMX_BAS.C[k][a][b] != 0) {
        constexpr T kc = T(LBMX_BAS.C[k][a][b]);
        quad += kc * v[a] * v[b];
      }
    });
  });
  return m + quad * inv_rho_s;
}

// P_k(v) = 3 v.G1_k + 9 v^T G2_k v - 3 |v|^2 G0_k ; source moment = -alpha P_k
template <class L, int k, class T>
__device__ __forceinline__ T src_row(const T (&v)[3], T u2) {
  LBMX_CONTRACT_IN_STATEMENT
  T p = 0;
  if constexpr (LBMX_BAS.G0[k] != 0) {
    constexpr T kc = T(3.0 * LBMX_BAS.G0[k]);
    p -= kc * u2;
  }
  static_for<0, 3>([&](auto ac) {
    constexpr int a = decltype(ac)::value;
    if constexpr (LBMX_BAS.G1[k][a] != 0) {
      constexpr T kc = T(3.0 * LBMX_BAS.G1[k][a]);
      p += kc * v[a];
    }
    static_for<0, 3>([&](auto bc) {
      constexpr int b = decltype(bc)::value;
      if constexpr (LBMX_BAS.G2[k][a][b] != 0) {
        constexpr T kc = T(9.0 * LBMX_BAS.G2[k][a][b]);
        p += kc * v[a] * v[b];
      }
    });
  });
  return p;
}

// ---- D3Q19 moments grouped by lattice symmetry --------------------------------------------
// In pair space the D3Q19 basis rows (d'Humieres et al. 2002) read the pairs through a few
// shared sums: the three axis pairs (A_x, A_y, A_z) and, for each coordinate plane, the sum
// and difference of its two diagonal pairs (D, E); the odd rows read, per axis a, the axis
// pair (K_a) and the parts of the diagonal pairs in the two planes through a (H_a). Building
// the rows from these sums, and the back transform from the matching combinations of the
// row updates, takes about half the operations of the row-by-row form. Pair d of the
// lattice order: 1 -z, 2 -y, 3 -x, 4/5 the yz plane (c_y c_z = +1/-1), 6/7 xz, 8/9 xy.
namespace d3q19 {
constexpr bool pair_order_ok() {
  using L = LatD3Q19;
  return L::ez[1] == -1 && L::ey[2] == -1 && L::ex[3] == -1 && L::ey[4] * L::ez[4] == 1 &&
         L::ey[5] * L::ez[5] == -1 && L::ex[4] == 0 && L::ex[5] == 0 && L::ex[6] * L::ez[6] == 1 &&
         L::ex[7] * L::ez[7] == -1 && L::ey[6] == 0 && L::ey[7] == 0 && L::ex[8] * L::ey[8] == 1 &&
         L::ex[9] * L::ey[9] == -1 && L::ez[8] == 0 && L::ez[9] == 0 && L::ex[8] == -1 && L::ex[9] == 1 &&
         L::ex[6] == -1 && L::ex[7] == 1 && L::ey[4] == -1 && L::ey[5] == 1 && L::ez[4] == -1 &&
         L::ez[6] == -1 && L::ey[8] == -1 && L::ey[9] == -1 && L::ez[5] == -1 && L::ez[7] == -1;
}
static_assert(pair_order_ok(), "grouped D3Q19 transform assumes the lattice's pair order");
// the coefficients used below, against the basis table (row k, pair d; d = 0 is f_0)
constexpr bool rows_ok() {
  // even rows 4, 5, 9..15 at f_0, the three axis pairs (z, y, x) and the six diagonal pairs
  constexpr int ev[9] = {4, 5, 9, 10, 11, 12, 13, 14, 15};
  constexpr double c[9][10] = {
      {-30, -11, -11, -11, 8, 8, 8, 8, 8, 8}, {12, -4, -4, -4, 1, 1, 1, 1, 1, 1},
      {0, -1, -1, 2, -2, -2, 1, 1, 1, 1},     {0, 2, 2, -4, -2, -2, 1, 1, 1, 1},
      {0, -1, 1, 0, 0, 0, -1, -1, 1, 1},      {0, 2, -2, 0, 0, 0, -1, -1, 1, 1},
      {0, 0, 0, 0, 0, 0, 0, 0, 1, -1},        {0, 0, 0, 0, 1, -1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 1, -1, 0, 0}};
  for (int i = 0; i < 9; ++i)
    for (int d = 0; d < 10; ++d)
      if (kB19.M[ev[i]][d] != c[i][d] || kB19.par[ev[i]] != 1) return false;
  // odd rows 1, 2, 3, 6, 7, 8, 16, 17, 18 at the nine pairs
  constexpr int od[9] = {1, 2, 3, 6, 7, 8, 16, 17, 18};
  constexpr double o[9][9] = {
      {0, 0, -1, 0, 0, -1, 1, -1, 1},  {0, -1, 0, -1, 1, 0, 0, -1, -1}, {-1, 0, 0, -1, -1, -1, -1, 0, 0},
      {0, 0, 4, 0, 0, -1, 1, -1, 1},   {0, 4, 0, -1, 1, 0, 0, -1, -1},  {4, 0, 0, -1, -1, -1, -1, 0, 0},
      {0, 0, 0, 0, 0, 1, -1, -1, 1},   {0, 0, 0, -1, 1, 0, 0, 1, 1},    {0, 0, 0, 1, 1, -1, -1, 0, 0}};
  for (int i = 0; i < 9; ++i)
    for (int d = 0; d < 9; ++d)
      if (kB19.M[od[i]][d + 1] != o[i][d] || kB19.par[od[i]] != -1) return false;
  return kB19.par[0] == 1;
}
static_assert(rows_ok(), "grouped D3Q19 transform does not match the basis table");

// density and momentum from the pairs
template <class T>
__device__ __forceinline__ void macros(const Pairs<LatD3Q19, T>& p, T& rho, T (&j)[3]) {
  LBMX_CONTRACT_IN_STATEMENT
  rho = p.f0 + ((p.s[1] + p.s[2]) + p.s[3]) + (((p.s[4] + p.s[5]) + (p.s[6] + p.s[7])) + (p.s[8] + p.s[9]));
  j[0] = ((p.a[9] - p.a[8]) + (p.a[7] - p.a[6])) - p.a[3];
  j[1] = (p.a[5] - p.a[4]) - (p.a[8] + p.a[9]) - p.a[2];
  j[2] = -((p.a[6] + p.a[7]) + (p.a[4] + p.a[5])) - p.a[1];
}

// The relaxation in two independent halves: the odd rows read and update only a, the even
// rows only f_0 and s. row(k, m_k) returns C_k = 2 dm_k / |M_k|^2 for row k with moment m_k
// (the moment argument is ignored for the momentum rows, which carry the source alone).
template <class T, class Row>
__device__ __forceinline__ void relax_odd(Pairs<LatD3Q19, T>& p, Row&& row) {
  LBMX_CONTRACT_IN_STATEMENT
  const T Hx_xy = p.a[9] - p.a[8], Hx_xz = p.a[7] - p.a[6];
  const T Hy_xy = -(p.a[8] + p.a[9]), Hy_yz = p.a[5] - p.a[4];
  const T Hz_xz = -(p.a[6] + p.a[7]), Hz_yz = -(p.a[4] + p.a[5]);
  const T Hx = Hx_xy + Hx_xz, Hy = Hy_xy + Hy_yz, Hz = Hz_xz + Hz_yz;
  const T C1 = row(IC<1>{}, T(0)), C2 = row(IC<2>{}, T(0)), C3 = row(IC<3>{}, T(0));
  const T C6 = row(IC<6>{}, Hx + T(4) * p.a[3]);
  const T C7 = row(IC<7>{}, Hy + T(4) * p.a[2]);
  const T C8 = row(IC<8>{}, Hz + T(4) * p.a[1]);
  const T C16 = row(IC<16>{}, Hx_xy - Hx_xz);
  const T C17 = row(IC<17>{}, Hy_yz - Hy_xy);
  const T C18 = row(IC<18>{}, Hz_xz - Hz_yz);
  // p.a -= M^T C over the odd rows
  const T Gx = C1 + C6, Gy = C2 + C7, Gz = C3 + C8;
  const T Xy = Gx + C16, Xz = Gx - C16, Yx = Gy - C17, Yz = Gy + C17, Zx = Gz + C18, Zy = Gz - C18;
  p.a[3] += C1 - T(4) * C6;
  p.a[2] += C2 - T(4) * C7;
  p.a[1] += C3 - T(4) * C8;
  p.a[8] += Xy + Yx;
  p.a[9] -= Xy - Yx;
  p.a[6] += Xz + Zx;
  p.a[7] -= Xz - Zx;
  p.a[4] += Yz + Zy;
  p.a[5] -= Yz - Zy;
}

template <class T, class Row>
__device__ __forceinline__ void relax_even(Pairs<LatD3Q19, T>& p, Row&& row) {
  LBMX_CONTRACT_IN_STATEMENT
  const T Az = p.s[1], Ay = p.s[2], Ax = p.s[3];
  const T Dyz = p.s[4] + p.s[5], Dxz = p.s[6] + p.s[7], Dxy = p.s[8] + p.s[9];
  const T SA = Ax + Ay + Az, SD = Dxy + Dxz + Dyz;
  const T t1 = T(3) * Ax - SA, t2 = SD - T(3) * Dyz, t3 = Ay - Az, t4 = Dxy - Dxz;
  const T C4 = row(IC<4>{}, T(8) * SD - T(11) * SA - T(30) * p.f0);
  const T C5 = row(IC<5>{}, SD - T(4) * SA + T(12) * p.f0);
  const T C9 = row(IC<9>{}, t1 + t2);
  const T C10 = row(IC<10>{}, t2 - T(2) * t1);
  const T C11 = row(IC<11>{}, t3 + t4);
  const T C12 = row(IC<12>{}, t4 - T(2) * t3);
  const T C13 = row(IC<13>{}, p.s[8] - p.s[9]);
  const T C14 = row(IC<14>{}, p.s[4] - p.s[5]);
  const T C15 = row(IC<15>{}, p.s[6] - p.s[7]);
  // f_0 -= M^T C / 2 and s -= M^T C over the even rows
  p.f0 -= T(6) * C5 - T(15) * C4;
  const T ua = -(T(11) * C4 + T(4) * C5), P = C9 - T(2) * C10, W = C11 - T(2) * C12;
  p.s[3] -= ua + T(2) * P;
  p.s[2] -= (ua - P) + W;
  p.s[1] -= (ua - P) - W;
  const T ud = T(8) * C4 + C5, R = C9 + C10, V = C11 + C12;
  const T bxy = (ud + R) + V, bxz = (ud + R) - V, byz = ud - T(2) * R;
  p.s[8] -= bxy + C13;
  p.s[9] -= bxy - C13;
  p.s[6] -= bxz + C15;
  p.s[7] -= bxz - C15;
  p.s[4] -= byz + C14;
  p.s[5] -= byz - C14;
}
}  // namespace d3q19

// ---- forward (pair space) ---------------------------------------------------------------
template <class L, int COLL, int DES, class T>
__device__ __forceinline__ void base_relax(Pairs<L, T>& p, T rho, const T (&v)[3], T omega, T alpha) {
  constexpr bool inc = CollTraits<COLL>::inc;
  const T u2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  if constexpr (!CollTraits<COLL>::mrt) {
    // BGK: f' = f (1 - omega) + omega f_eq  (src/dynamics/incompressibleBGK2d.hh:38-47);
    // even/odd parts of f_eq: 2 w (rho + 4.5 cu^2 - 1.5 u^2), 2 w 3 cu (incompressible)
    const T one_m = T(1) - omega;
    const T kap_a = (T(1) - T(0.5) * omega) * alpha;
    const T rs = inc ? T(1) : rho;  // standard: f_eq = w rho (1 + ...)
    {
      T feq0 = inc ? T(L::w[0]) * (rho - T(1.5) * u2) : T(L::w[0]) * rho * (T(1) - T(1.5) * u2);
      T out = p.f0 * one_m + omega * feq0;
      if constexpr (DES == D_BRINK) out += kap_a * T(L::w[0]) * T(3) * u2;
      p.f0 = out;
    }
    static_for<1, L::HALF + 1>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      const T cu = cdot<L, d>(v);
      const T w2 = T(2.0 * L::w[d]);
      T ev = inc ? w2 * (rho + T(4.5) * cu * cu - T(1.5) * u2) : w2 * rho * (T(1) + T(4.5) * cu * cu - T(1.5) * u2);
      T od = w2 * rs * T(3) * cu;
      T s = p.s[d] * one_m + omega * ev;
      T a = p.a[d] * one_m + omega * od;
      if constexpr (DES == D_BRINK) {
        s -= kap_a * w2 * (T(9) * cu * cu - T(3) * u2);
        a -= kap_a * w2 * T(3) * cu;
      }
      p.s[d] = s;
      p.a[d] = a;
    });
  } else {
    // MRT: f' = f - M^-1 [S (M f - m_eq) - (I - S/2) M src]. The rows of M are orthogonal,
    // so relaxing them one after another, each as the rank-1 update f -= M_k^T dm_k / |M_k|^2,
    // leaves the moments of the remaining rows unchanged and gives the simultaneous result;
    // only the populations stay live (no vector of moments).
    constexpr int Q = L::QF;
    const T inv_rho_s = inc ? T(1) : recip<T>(rho);
    // equilibrium moments take the momentum: j (incompressible) or rho v (standard, whose
    // quadratic terms are j_a j_b / rho: src/dynamics/MRT3d.hh:270-285)
    T vm[3];
    static_for<0, 3>([&](auto ac) {
      constexpr int a = decltype(ac)::value;
      vm[a] = inc ? v[a] : rho * v[a];
    });
    static_for<0, Q>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if constexpr (LBMX_BAS.rate[k] != 0) {
        T mk = pair_moment<L, k>(p);
        constexpr T s_k = T(LBMX_BAS.s[k]), inv_n2 = T(LBMX_BAS.invn2[k]);
        constexpr bool own_rate = LBMX_BAS.rate[k] == 1;
        T sk = own_rate ? omega : s_k;
        T val = sk * (mk - meq_row<L, k>(rho, vm, inv_rho_s));
        if constexpr (DES == D_BRINK && row_has_src<L, k>())
          val += (T(1) - T(0.5) * sk) * alpha * src_row<L, k>(v, u2);
        pair_row_update<L, k>(p, val * inv_n2);
      } else if constexpr (DES == D_BRINK && row_has_src<L, k>()) {
        constexpr T inv_n2 = T(LBMX_BAS.invn2[k]);
        pair_row_update<L, k>(p, alpha * src_row<L, k>(v, u2) * inv_n2);
      }
    });
  }
}

// D3Q19 MRT relaxation with the grouped transforms: the row values of base_relax, each row
// update applied at once (the rows are orthogonal, so the simultaneous update equals the
// row-by-row one).
template <int COLL, int DES, class T>
__device__ __forceinline__ void relax_d3q19(Pairs<LatD3Q19, T>& p, T rho, const T (&v)[3], T omega, T alpha) {
  LBMX_CONTRACT_IN_STATEMENT
  using L = LatD3Q19;
  constexpr bool inc = CollTraits<COLL>::inc;
  static_assert(CollTraits<COLL>::mrt, "MRT operators only");
  static_assert(LBMX_BAS.rate[0] == 0 && !row_has_src<L, 0>(), "the density row is neither relaxed nor forced");
  const T u2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const T inv_rho_s = inc ? T(1) : recip<T>(rho);
  T vm[3];
  static_for<0, 3>([&](auto ac) {
    constexpr int a = decltype(ac)::value;
    vm[a] = inc ? v[a] : rho * v[a];
  });
  auto row = [&](auto kc, T mk) -> T {
    constexpr int k = decltype(kc)::value;
    constexpr T two_inv_n2 = T(2.0 * LBMX_BAS.invn2[k]);
    if constexpr (LBMX_BAS.rate[k] != 0) {
      constexpr T s_k = T(LBMX_BAS.s[k]);
      const T sk = LBMX_BAS.rate[k] == 1 ? omega : s_k;
      T val = sk * (mk - meq_row<L, k>(rho, vm, inv_rho_s));
      if constexpr (DES == D_BRINK && row_has_src<L, k>()) val += (T(1) - T(0.5) * sk) * alpha * src_row<L, k>(v, u2);
      return val * two_inv_n2;
    } else if constexpr (DES == D_BRINK && row_has_src<L, k>()) {
      return alpha * src_row<L, k>(v, u2) * two_inv_n2;
    } else {
      return T(0);
    }
  };
  d3q19::relax_odd<T>(p, row);
  d3q19::relax_even<T>(p, row);
}

// thermal part: ThermalObstacleCollision<ThermalMRT2d<D2Q5_TMRT>,Ramp>
// (src/dynamics/thermalObstacleCollision.hh:47-53, src/dynamics/thermalMRT2d.hh:76-101)
template <class T>
__device__ __forceinline__ T thermal_omega(const ModelParams& mp, T gamma) {
  T diff = mpar<T>(mp.tdiff, mp.tdiff_f) * ramp_interp<T>(mpar<T>(mp.ramp_p, mp.ramp_p_f), mpar<T>(mp.ramp_ck, mp.ramp_ck_f), gamma);
  return T(1) / (T(5) * diff + T(0.5));
}

template <class L, class T>
__device__ __forceinline__ void thermal_relax(T (&f)[L::Q], const T (&j)[3], T gamma, const ModelParams& mp) {
  constexpr int o = L::QF;
  T temp = f[o] + f[o + 1] + f[o + 2] + f[o + 3] + f[o + 4];
  T omT = thermal_omega<T>(mp, gamma);
  T dm[5];
  static_for<1, 5>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    T mk = 0;
    static_for<0, 5>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (kB5.M[k][q] != 0) mk += T(kB5.M[k][q]) * f[o + q];
    });
    T meq = temp * (T(kB5.a[k]) + T(kB5.b[k][0]) * j[0] + T(kB5.b[k][1]) * j[1]);
    T sk = kB5.rate[k] == 1 ? omT : T(kB5.s[k]);
    dm[k] = sk * (mk - meq) * T(kB5.invn2[k]);
  });
  static_for<0, 5>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    T acc = f[o + q];
    static_for<1, 5>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if constexpr (kB5.M[k][q] != 0) acc -= T(kB5.M[k][q]) * dm[k];
    });
    f[o + q] = acc;
  });
}

template <class L, int COLL, int DES, class T>
__device__ __forceinline__ void eq_velocity(T rho, const T (&j)[3], T alpha, T (&v)[3], T& beta) {
  constexpr bool inc = CollTraits<COLL>::inc;
  if constexpr (DES == D_BRINK) {
    beta = inc ? recip<T>(T(1) + T(0.5) * alpha) : recip<T>(rho * (T(1) + T(0.5) * alpha));
    v[0] = beta * j[0]; v[1] = beta * j[1]; v[2] = beta * j[2];
  } else if constexpr (inc) {
    beta = T(1);
    v[0] = j[0]; v[1] = j[1]; v[2] = j[2];
  } else {
    beta = recip<T>(rho);
    v[0] = j[0] * beta; v[1] = j[1] * beta; v[2] = j[2] * beta;
  }
}

// ---- cascaded (D2Q9) -------------------------------------------------------------------
// IncompressibleCascaded2d<D2Q9> (src/dynamics/cascaded2d.hh:40-95): k = K(f, rho, u) with the
// incompressible macros (u = j), then f' = f + T k (momentToDistribution). The shear moments
// k1, k2 relax with omega, the rest with rate 1.
template <class T>
__device__ __forceinline__ void casc_k(const T (&f)[9], T r, T x, T y, T om, T (&k)[6]) {
  const T u2 = x * x + y * y;
  k[0] = (r * u2 - f[8] - f[6] - f[2] - f[4] - T(2) * (f[1] + f[3] + f[7] + f[5] - r / T(3))) / T(12);
  k[1] = om * (f[6] + f[2] - f[8] - f[4] + r * (x * x - y * y)) / T(4);
  k[2] = om * ((f[7] + f[3] - f[5] - f[1]) - r * x * y) / T(4);
  k[3] = (-((f[1] + f[3] - f[7] - f[5] - T(2) * r * x * x * y + y * (r - f[6] - f[2] - f[0])) / T(4) +
            T(0.5) * x * (f[7] - f[5] - f[1] + f[3]))) +
         T(0.5) * y * (-T(3) * k[0] - k[1]) + T(2) * x * k[2];
  k[4] = (-((f[3] + f[5] - f[1] - f[7] - T(2) * r * y * y * x + x * (r - f[4] - f[8] - f[0])) / T(4) +
            T(0.5) * y * (f[7] + f[3] - f[1] - f[5]))) +
         T(0.5) * x * (-T(3) * k[0] + k[1]) + T(2) * y * k[2];
  k[5] = (T(0.25) * (r / T(9) - f[7] - f[5] - f[1] - f[3] +
                     T(2) * (x * (f[7] - f[5] + f[1] - f[3]) + y * (f[7] + f[5] - f[1] - f[3])) +
                     T(4) * x * y * (f[5] - f[7] + f[1] - f[3]) -
                     x * x * (f[6] + f[7] + f[5] + f[2] + f[1] + f[3]) +
                     y * y * (T(3) * r * x * x - f[8] - f[7] - f[5] - f[1] - f[3] - f[4]))) -
         T(2) * k[0] - T(2) * x * k[4] - T(2) * y * k[3] + T(4) * x * y * k[2] - T(1.5) * k[0] * u2 +
         T(0.5) * k[1] * (x * x - y * y);
}

// g = f + T k
template <class T>
__device__ __forceinline__ void casc_apply(const T (&f)[9], const T (&k)[6], T (&g)[9]) {
  g[0] = f[0] - T(4) * k[0] + T(4) * k[5];
  g[1] = f[1] + T(2) * k[0] + k[2] + k[3] - k[4] + k[5];
  g[2] = f[2] - k[0] - k[1] - T(2) * k[3] - T(2) * k[5];
  g[3] = f[3] + T(2) * k[0] - k[2] + k[3] + k[4] + k[5];
  g[4] = f[4] - k[0] + k[1] - T(2) * k[4] - T(2) * k[5];
  g[5] = f[5] + T(2) * k[0] + k[2] - k[3] + k[4] + k[5];
  g[6] = f[6] - k[0] - k[1] + T(2) * k[3] - T(2) * k[5];
  g[7] = f[7] + T(2) * k[0] - k[2] - k[3] - k[4] + k[5];
  g[8] = f[8] - k[0] + k[1] + T(2) * k[4] - T(2) * k[5];
}

// lam <- lam + (dk/df)^T T^T lam for the explicit f terms (reverse-mode derivative of casc_k);
// the parts through rho and u are added to rb, xb, yb.
template <class T>
__device__ __forceinline__ void casc_reverse(const T (&f)[9], T r, T x, T y, T om, T (&lam)[9], T& rb, T& xb,
                                             T& yb) {
  T k[6];
  casc_k<T>(f, r, x, y, om, k);
  const T* l = lam;
  T kb0 = -T(4) * l[0] + T(2) * (l[1] + l[3] + l[5] + l[7]) - (l[2] + l[4] + l[6] + l[8]);
  T kb1 = -l[2] + l[4] - l[6] + l[8];
  T kb2 = l[1] - l[3] + l[5] - l[7];
  T kb3 = l[1] - T(2) * l[2] + l[3] - l[5] + T(2) * l[6] - l[7];
  T kb4 = -l[1] + l[3] - T(2) * l[4] + l[5] - l[7] + T(2) * l[8];
  const T kb5 = T(4) * l[0] + (l[1] + l[3] + l[5] + l[7]) - T(2) * (l[2] + l[4] + l[6] + l[8]);
  T fb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  T r_b = 0, x_b = 0, y_b = 0;
  const T u2 = x * x + y * y;
  {  // k5
    const T b = kb5, Sb = T(0.25) * b;
    const T A5 = f[7] - f[5] + f[1] - f[3], B5 = f[7] + f[5] - f[1] - f[3], C5 = f[5] - f[7] + f[1] - f[3];
    const T E5 = f[6] + f[7] + f[5] + f[2] + f[1] + f[3], F5 = f[8] + f[7] + f[5] + f[1] + f[3] + f[4];
    r_b += Sb * (T(1) / T(9) + T(3) * x * x * y * y);
    x_b += Sb * (T(2) * A5 + T(4) * y * C5 - T(2) * x * E5 + T(6) * r * x * y * y);
    y_b += Sb * (T(2) * B5 + T(4) * x * C5 + T(2) * y * (T(3) * r * x * x - F5));
    const T Ab = Sb * T(2) * x, Bb = Sb * T(2) * y, Cb = Sb * T(4) * x * y, Eb = -Sb * x * x, Fb = -Sb * y * y;
    fb[7] += -Sb + Ab + Bb - Cb + Eb + Fb;
    fb[5] += -Sb - Ab + Bb + Cb + Eb + Fb;
    fb[1] += -Sb + Ab - Bb + Cb + Eb + Fb;
    fb[3] += -Sb - Ab - Bb - Cb + Eb + Fb;
    fb[6] += Eb; fb[2] += Eb; fb[8] += Fb; fb[4] += Fb;
    kb0 += b * (-T(2) - T(1.5) * u2);
    kb4 += b * (-T(2) * x);
    kb3 += b * (-T(2) * y);
    kb2 += b * (T(4) * x * y);
    kb1 += b * (T(0.5) * (x * x - y * y));
    x_b += b * (-T(2) * k[4] + T(4) * y * k[2] - T(3) * k[0] * x + k[1] * x);
    y_b += b * (-T(2) * k[3] + T(4) * x * k[2] - T(3) * k[0] * y - k[1] * y);
  }
  {  // k4
    const T b = kb4, Pb = -T(0.25) * b, Qb = -T(0.5) * y * b;
    const T Q4 = f[7] + f[3] - f[1] - f[5];
    fb[3] += Pb; fb[5] += Pb; fb[1] -= Pb; fb[7] -= Pb;
    r_b += Pb * (-T(2) * y * y * x + x);
    y_b += Pb * (-T(4) * r * x * y);
    x_b += Pb * (-T(2) * r * y * y + (r - f[4] - f[8] - f[0]));
    fb[4] -= Pb * x; fb[8] -= Pb * x; fb[0] -= Pb * x;
    fb[7] += Qb; fb[3] += Qb; fb[1] -= Qb; fb[5] -= Qb;
    y_b += b * (-T(0.5) * Q4 + T(2) * k[2]);
    x_b += b * (T(0.5) * (-T(3) * k[0] + k[1]));
    kb0 += b * (-T(1.5) * x);
    kb1 += b * (T(0.5) * x);
    kb2 += b * (T(2) * y);
  }
  {  // k3
    const T b = kb3, Pb = -T(0.25) * b, Qb = -T(0.5) * x * b;
    const T Q3 = f[7] - f[5] - f[1] + f[3];
    fb[1] += Pb; fb[3] += Pb; fb[7] -= Pb; fb[5] -= Pb;
    r_b += Pb * (-T(2) * x * x * y + y);
    x_b += Pb * (-T(4) * r * x * y);
    y_b += Pb * (-T(2) * r * x * x + (r - f[6] - f[2] - f[0]));
    fb[6] -= Pb * y; fb[2] -= Pb * y; fb[0] -= Pb * y;
    fb[7] += Qb; fb[5] -= Qb; fb[1] -= Qb; fb[3] += Qb;
    x_b += b * (-T(0.5) * Q3 + T(2) * k[2]);
    y_b += b * (T(0.5) * (-T(3) * k[0] - k[1]));
    kb0 += b * (-T(1.5) * y);
    kb1 += b * (-T(0.5) * y);
    kb2 += b * (T(2) * x);
  }
  {  // k2
    const T b = T(0.25) * om * kb2;
    fb[7] += b; fb[3] += b; fb[5] -= b; fb[1] -= b;
    r_b -= b * x * y; x_b -= b * r * y; y_b -= b * r * x;
  }
  {  // k1
    const T b = T(0.25) * om * kb1;
    fb[6] += b; fb[2] += b; fb[8] -= b; fb[4] -= b;
    r_b += b * (x * x - y * y); x_b += b * T(2) * r * x; y_b -= b * T(2) * r * y;
  }
  {  // k0
    const T b = kb0 / T(12);
    fb[8] -= b; fb[6] -= b; fb[2] -= b; fb[4] -= b;
    fb[1] -= T(2) * b; fb[3] -= T(2) * b; fb[7] -= T(2) * b; fb[5] -= T(2) * b;
    r_b += b * (u2 + T(2) / T(3)); x_b += b * T(2) * r * x; y_b += b * T(2) * r * y;
  }
  static_for<0, 9>([&](auto qc) { lam[decltype(qc)::value] += fb[decltype(qc)::value]; });
  rb += r_b;
  xb += x_b;
  yb += y_b;
}

template <class L, int DES, class T>
__device__ __forceinline__ void collide_cascaded(T (&f)[L::Q], T gamma, const ModelParams& mp, T& rho, T (&j)[3]) {
  static_assert(L::Q == 9 && L::QF == 9, "the cascaded operator is defined for D2Q9");
  flow_macros<L>(f, rho, j);
  T k[6];
  casc_k<T>(f, rho, j[0], j[1], mpar<T>(mp.omega, mp.omega_f), k);
  T g[9];
  casc_apply<T>(f, k, g);
  if constexpr (DES == D_PBB) {
    const T imult = T(0.5) * mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_val<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
    const T sc = T(1) - T(2) * imult;
    static_for<1, 5>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      const T sd = g[d] + g[d + 4], ad = (g[d] - g[d + 4]) * sc;
      g[d] = T(0.5) * (sd + ad);
      g[d + 4] = T(0.5) * (sd - ad);
    });
  }
  static_for<0, 9>([&](auto qc) { f[decltype(qc)::value] = g[decltype(qc)::value]; });
}

template <class L, int DES, class T>
__device__ __forceinline__ void reverse_cascaded(const T (&f)[L::Q], T gamma, const ModelParams& mp, T (&lam)[L::Q],
                                                 T& dgam, T srcA, const T (&srcB)[3]) {
  static_assert(L::Q == 9 && L::QF == 9, "the cascaded operator is defined for D2Q9");
  T rho, j[3];
  flow_macros<L>(f, rho, j);
  const T om = mpar<T>(mp.omega, mp.omega_f);
  if constexpr (DES == D_PBB) {
    T k[6], g[9];
    casc_k<T>(f, rho, j[0], j[1], om, k);
    casc_apply<T>(f, k, g);
    const T imult = T(0.5) * mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_val<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
    const T sc = T(1) - T(2) * imult;
    T mbar = 0;
    static_for<1, 5>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      const T ls = lam[d] + lam[d + 4], la = lam[d] - lam[d + 4];
      mbar -= la * (g[d] - g[d + 4]);
      const T las = la * sc;
      lam[d] = T(0.5) * (ls + las);
      lam[d + 4] = T(0.5) * (ls - las);
    });
    dgam += mbar * T(0.5) * mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_der<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
  }
  T rb = srcA, xb = srcB[0], yb = srcB[1];
  casc_reverse<T>(f, rho, j[0], j[1], om, lam, rb, xb, yb);
  static_for<0, 9>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    T v = lam[q] + rb;
    if constexpr (L::ex[q] != 0) v += T(L::ex[q]) * xb;
    if constexpr (L::ey[q] != 0) v += T(L::ey[q]) * yb;
    lam[q] = v;
  });
}

// DarcyMRT2d forcing (src/dynamics/darcyMRT2d.hh:33-59), added to the MRT output: drag
// F = -amax BP_q(gamma) rho u + body force with the unshifted incompressible macros u = j,
// through the operator's closed-form polynomials (MRT rates s1 = s2 = 1.4).
template <class T>
__device__ __forceinline__ void darcy_force(T (&f)[9], T rho, const T (&j)[3], T gamma, const ModelParams& mp) {
  const T s1 = T(1.4), s2 = T(1.4), om = mpar<T>(mp.omega, mp.omega_f);
  const T ux = j[0], uy = j[1];
  const T im = -mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_val<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma) * rho;
  const T Fx = im * ux + mpar<T>(mp.body[0], mp.body_f[0]), Fy = im * uy + mpar<T>(mp.body[1], mp.body_f[1]);
  const T o3 = T(1) / T(3), o6 = T(1) / T(6), o12 = T(1) / T(12), o24 = T(1) / T(24);
  const T ax = (-T(4) * s1 + T(2) * s2 + T(4)) * ux;           // common x-part of the diagonal terms
  const T by = (s1 - T(0.5) * s2 - T(1)) * uy;
  const T cy = (s1 - T(2) * s2 - T(3) * om + T(8));
  const T dd = (s1 - T(2) * s2 + T(3) * om - T(4));
  f[0] += o3 * (s1 + s2 - T(4)) * (Fx * ux + Fy * uy);
  f[1] += o24 * (ax + T(2) + (T(3) * om - T(6)) * uy) * Fx - o6 * Fy * ((-T(0.75) * om + T(1.5)) * ux + T(0.5) + by);
  f[2] += o12 * (-T(4) + cy * uy) * Fy + o12 * Fx * ux * dd;
  f[3] += o24 * (ax - T(2) + (-T(3) * om + T(6)) * uy) * Fx - o6 * Fy * ((T(0.75) * om - T(1.5)) * ux + T(0.5) + by);
  f[4] += o12 * (-T(4) + cy * ux) * Fx + o12 * Fy * uy * dd;
  f[5] += o24 * (ax - T(2) + (T(3) * om - T(6)) * uy) * Fx - o6 * Fy * ((-T(0.75) * om + T(1.5)) * ux - T(0.5) + by);
  f[6] += o12 * (T(4) + cy * uy) * Fy + o12 * Fx * ux * dd;
  f[7] += o24 * (ax + T(2) + (-T(3) * om + T(6)) * uy) * Fx - o6 * Fy * ((T(0.75) * om - T(1.5)) * ux - T(0.5) + by);
  f[8] += o12 * (T(4) + cy * ux) * Fx + o12 * Fy * uy * dd;
}

// Moment-space operators (BGK, MRT): node collision in place; rho, j are the pre-collision macros.
template <class L, int COLL, int DES, class T, bool TWICE = false>
__device__ __forceinline__ void collide_moment(T (&f)[L::Q], T gamma, const ModelParams& mp, T& rho, T (&j)[3]) {
  static_assert(!(TWICE && DES == D_DARCY), "the Darcy force acts on the populations themselves");
  Pairs<L, T> p;
  to_pairs<L>(f, p);
  // (FP32 arithmetic only: the FP64 sweeps are bound by HBM, and the grouped form holds more
  // values live, which would cost the FP64 kernels occupancy)
  constexpr bool grouped = std::is_same_v<L, LatD3Q19> && CollTraits<COLL>::mrt && sizeof(T) == 4;
  if constexpr (grouped) d3q19::macros(p, rho, j);
  else pair_macros<L>(p, rho, j);
  const T omega = mpar<T>(mp.omega, mp.omega_f);
  T alpha = 0, beta, v[3];
  if constexpr (DES == D_BRINK) alpha = mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_val<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
  eq_velocity<L, COLL, DES>(rho, j, alpha, v, beta);
  if constexpr (grouped) relax_d3q19<COLL, DES>(p, rho, v, omega, alpha);
  else base_relax<L, COLL, DES>(p, rho, v, omega, alpha);
  if constexpr (DES == D_PBB) {
    // PartialBouncebackCollision (src/dynamics/partialBouncebackCollision.hh:53-65): the
    // pair blend keeps f_d + f_d' and scales f_d - f_d' by 1 - 2 imult
    const T imult = T(0.5) * mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_val<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
    const T sc = T(1) - T(2) * imult;
    static_for<1, L::HALF + 1>([&](auto dc) { p.a[decltype(dc)::value] *= sc; });
  }
  from_pairs<L, T, TWICE>(p, f);
  if constexpr (DES == D_DARCY) {
    static_assert(L::Q == 9 && CollTraits<COLL>::mrt && CollTraits<COLL>::inc, "DarcyMRT2d is D2Q9 MRT");
    darcy_force<T>(f, rho, j, gamma, mp);
  }
  if constexpr (L::QT > 0) thermal_relax<L>(f, j, gamma, mp);
}

// ---- reverse (pair space) ---------------------------------------------------------------
// lam <- (dC/df)^T lam + (A + c.B) [+ AT on the thermal populations],
// dgam += (dC/dgamma)^T lam. (A, B, AT) carry a functional source that is affine in c, so it
// is added with the macro-space adjoints at no extra cost. The forward quantities needed are
// recomputed from f, as the reference's reverse operators do
// (src/adjointDynamics/reversePartialBounceback.hh:18-41, reverseIncompressibleBGK2d.hh:64-121,
// reverseIncompressibleMRT2d.hh:29-83, reverseMRT3d.hh:29-365).
template <class L, int COLL, int DES, class T>
__device__ __forceinline__ void reverse_moment(const T (&f)[L::Q], T gamma, const ModelParams& mp, T (&lam)[L::Q],
                                               T& dgam, T srcA, const T (&srcB)[3], T srcAT) {
  static_assert(CollTraits<COLL>::inc, "reverse operators exist for the incompressible operators");
  static_assert(DES != D_DARCY, "DarcyMRT2d has no reverse operator");
  constexpr int Q = L::QF, H = L::HALF;
  Pairs<L, T> pf;
  to_pairs<L>(f, pf);
  T rho, j[3];
  pair_macros<L>(pf, rho, j);
  const T omega = mpar<T>(mp.omega, mp.omega_f);
  T alpha = 0, beta, v[3];
  if constexpr (DES == D_BRINK) alpha = mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_val<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
  eq_velocity<L, COLL, DES>(rho, j, alpha, v, beta);
  const T u2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  T gbar = 0;
  T lamT[L::QT > 0 ? 5 : 1];
  if constexpr (L::QT > 0) {
    static_for<0, 5>([&](auto qc) { lamT[decltype(qc)::value] = lam[Q + decltype(qc)::value]; });
  }
  Pairs<L, T> pl;
  to_pairs<L>(lam, pl);

  if constexpr (DES == D_PBB) {
    // odd part g_d - g_d' of the base collision output
    T gd[H + 1];
    if constexpr (!CollTraits<COLL>::mrt) {
      static_for<1, H + 1>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        gd[d] = pf.a[d] * (T(1) - omega) + omega * T(6.0 * L::w[d]) * cdot<L, d>(v);
      });
    } else {
      Pairs<L, T> g;
      static_for<1, H + 1>([&](auto dc) { g.a[decltype(dc)::value] = pf.a[decltype(dc)::value]; });
      static_for<0, Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (LBMX_BAS.rate[k] != 0 && LBMX_BAS.par[k] == -1) {
          constexpr T s_k = T(LBMX_BAS.s[k]), inv_n2 = T(LBMX_BAS.invn2[k]);
          constexpr bool own_rate = LBMX_BAS.rate[k] == 1;
        T sk = own_rate ? omega : s_k;
          T dm = sk * (pair_moment<L, k>(g) - meq_row<L, k>(rho, v, T(1))) * inv_n2;
          pair_row_update<L, k>(g, dm);
        }
      });
      static_for<1, H + 1>([&](auto dc) { gd[decltype(dc)::value] = g.a[decltype(dc)::value]; });
    }
    const T imult = T(0.5) * mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_val<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
    const T sc = T(1) - T(2) * imult;
    T mbar = 0;
    static_for<1, H + 1>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      mbar -= pl.a[d] * gd[d];
      pl.a[d] *= sc;
    });
    gbar += mbar * T(0.5) * mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_der<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
  }

  T rho_bar = 0, vb[3] = {0, 0, 0}, alpha_bar = 0;
  if constexpr (!CollTraits<COLL>::mrt) {
    const T kap = T(1) - T(0.5) * omega;
    T S0 = T(L::w[0]) * pl.f0;
    T ce[3] = {0, 0, 0};  // sum_d w_d (3 c_d la_d + 9 cu_d c_d ls_d)
    T cs[3] = {0, 0, 0};  // sum_d w_d (3 c_d la_d + 18 cu_d c_d ls_d) (source)
    T asum = 0;           // sum_d w_d (3 cu_d la_d + (9 cu_d^2 - 3 u2) ls_d) + w0 l0 (-3 u2)
    if constexpr (DES == D_BRINK) asum = T(L::w[0]) * pl.f0 * (-T(3) * u2);
    static_for<1, H + 1>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      const T cu = cdot<L, d>(v);
      const T wd = T(L::w[d]);
      S0 += wd * pl.s[d];
      const T t1 = wd * (T(3) * pl.a[d] + T(9) * cu * pl.s[d]);
      if constexpr (L::ex[d] != 0) ce[0] += T(L::ex[d]) * t1;
      if constexpr (L::ey[d] != 0) ce[1] += T(L::ey[d]) * t1;
      if constexpr (L::ez[d] != 0) ce[2] += T(L::ez[d]) * t1;
      if constexpr (DES == D_BRINK) {
        const T t2 = wd * (T(3) * pl.a[d] + T(18) * cu * pl.s[d]);
        if constexpr (L::ex[d] != 0) cs[0] += T(L::ex[d]) * t2;
        if constexpr (L::ey[d] != 0) cs[1] += T(L::ey[d]) * t2;
        if constexpr (L::ez[d] != 0) cs[2] += T(L::ez[d]) * t2;
        asum += wd * (T(3) * cu * pl.a[d] + (T(9) * cu * cu - T(3) * u2) * pl.s[d]);
      }
    });
    rho_bar = omega * S0;
    static_for<0, 3>([&](auto ac) {
      constexpr int a = decltype(ac)::value;
      vb[a] = omega * (ce[a] - T(3) * v[a] * S0);
      if constexpr (DES == D_BRINK) vb[a] -= alpha * kap * (cs[a] - T(6) * v[a] * S0);
    });
    if constexpr (DES == D_BRINK) alpha_bar = -kap * asum;
    const T one_m = T(1) - omega;
    pl.f0 *= one_m;
    static_for<1, H + 1>([&](auto dc) {
      pl.s[decltype(dc)::value] *= one_m;
      pl.a[decltype(dc)::value] *= one_m;
    });
  } else {
    // rows in sequence as rank-1 updates of lam (orthogonal rows: the same as all at once)
    static_for<0, Q>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if constexpr (row_active<L, DES, k>()) {
        constexpr T s_k = T(LBMX_BAS.s[k]), inv_n2 = T(LBMX_BAS.invn2[k]);
        const T ak = pair_moment<L, k>(pl) * inv_n2;
        constexpr int rate = LBMX_BAS.rate[k];
        const T sk = rate == 0 ? T(0) : (rate == 1 ? omega : s_k);
        if constexpr (LBMX_BAS.rate[k] != 0) {
          const T b = sk * ak;
          pair_row_update<L, k>(pl, b);
          if constexpr (LBMX_BAS.a[k] != 0) {
            constexpr T kc = T(LBMX_BAS.a[k]);
            rho_bar += b * kc;
          }
          static_for<0, 3>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            if constexpr (LBMX_BAS.b[k][a] != 0) {
              constexpr T kc = T(LBMX_BAS.b[k][a]);
              vb[a] += b * kc;
            }
            static_for<0, 3>([&](auto cc) {
              constexpr int c = decltype(cc)::value;
              constexpr double csum = LBMX_BAS.C[k][a][c] + LBMX_BAS.C[k][c][a];
              if constexpr (csum != 0) vb[a] += b * T(csum) * v[c];
            });
          });
        }
        if constexpr (DES == D_BRINK && row_has_src<L, k>()) {
          // the source moment alpha P_k enters dm_k with (1 - s_k/2); out -= M^T N dm
          const T pbar = -(T(1) - T(0.5) * sk) * ak;  // adjoint of alpha P_k
          alpha_bar += pbar * src_row<L, k>(v, u2);
          const T pa = pbar * alpha;
          if constexpr (LBMX_BAS.G0[k] != 0) {
            constexpr T kc = T(6.0 * LBMX_BAS.G0[k]);
            vb[0] -= pa * kc * v[0];
            vb[1] -= pa * kc * v[1];
            vb[2] -= pa * kc * v[2];
          }
          static_for<0, 3>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            if constexpr (LBMX_BAS.G1[k][a] != 0) {
              constexpr T kc = T(3.0 * LBMX_BAS.G1[k][a]);
              vb[a] += pa * kc;
            }
            static_for<0, 3>([&](auto cc) {
              constexpr int c = decltype(cc)::value;
              constexpr double gs = LBMX_BAS.G2[k][a][c] + LBMX_BAS.G2[k][c][a];
              if constexpr (gs != 0) vb[a] += pa * T(9.0 * gs) * v[c];
            });
          });
        }
      }
    });
  }
  T jb[3] = {beta * vb[0] + srcB[0], beta * vb[1] + srcB[1], beta * vb[2] + srcB[2]};
  rho_bar += srcA;
  if constexpr (DES == D_BRINK) {
    alpha_bar -= T(0.5) * beta * beta * (vb[0] * j[0] + vb[1] * j[1] + vb[2] * j[2]);
    gbar += alpha_bar * mpar<T>(mp.alpha_max, mp.alpha_max_f) * bp_der<T>(mpar<T>(mp.interp_q, mp.interp_q_f), gamma);
  }

  if constexpr (L::QT > 0) {
    constexpr int o = Q;
    T temp = f[o] + f[o + 1] + f[o + 2] + f[o + 3] + f[o + 4];
    T omT = thermal_omega<T>(mp, gamma);
    T ak[5];
    static_for<1, 5>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      T s = 0;
      static_for<0, 5>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (kB5.M[k][q] != 0) s += T(kB5.M[k][q]) * lamT[q];
      });
      ak[k] = s * T(kB5.invn2[k]);
    });
    T Tbar = srcAT, ombar = 0;
    static_for<1, 5>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      T sk = kB5.rate[k] == 1 ? omT : T(kB5.s[k]);
      T b = sk * ak[k];
      Tbar += b * (T(kB5.a[k]) + T(kB5.b[k][0]) * j[0] + T(kB5.b[k][1]) * j[1]);
      if constexpr (kB5.b[k][0] != 0) jb[0] += b * temp * T(kB5.b[k][0]);
      if constexpr (kB5.b[k][1] != 0) jb[1] += b * temp * T(kB5.b[k][1]);
      if constexpr (kB5.rate[k] == 1) {
        T mk = 0;
        static_for<0, 5>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          if constexpr (kB5.M[k][q] != 0) mk += T(kB5.M[k][q]) * f[o + q];
        });
        T meq = temp * (T(kB5.a[k]) + T(kB5.b[k][0]) * j[0] + T(kB5.b[k][1]) * j[1]);
        ombar -= ak[k] * (mk - meq);
      }
    });
    static_for<0, 5>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      T acc = lamT[q];
      static_for<1, 5>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (kB5.M[k][q] != 0) {
          T sk = kB5.rate[k] == 1 ? omT : T(kB5.s[k]);
          acc -= T(kB5.M[k][q]) * sk * ak[k];
        }
      });
      lam[o + q] = acc + Tbar;
    });
    T dR = ramp_diff<T>(mpar<T>(mp.ramp_p, mp.ramp_p_f), mpar<T>(mp.ramp_ck, mp.ramp_ck_f), gamma);
    gbar += ombar * (-omT * omT * T(5) * mpar<T>(mp.tdiff, mp.tdiff_f) * dR);
  }

  // lam_pre = lam_f + rho_bar + c.j_bar (pair space: f0 += rb, s += 2 rb, a += 2 c.jb)
  pl.f0 += rho_bar;
  static_for<1, H + 1>([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    pl.s[d] += T(2) * rho_bar;
    pl.a[d] += T(2) * cdot<L, d>(jb);
  });
  from_pairs<L>(pl, lam);
  dgam += gbar;
}

#undef LBMX_BAS

// ---- entry points -----------------------------------------------------------------------
// The moment-space operators end by halving pair sums, which a caller that multiplies the
// result by constants anyway (the 16-bit storage codecs) can fold into its own factor.
template <int COLL, int DES>
inline constexpr bool kCollideTwice = COLL != C_CASC && DES != D_DARCY;

// Node collision in place; rho, j receive the pre-collision macros. TWICE (only where
// kCollideTwice): f_q comes back doubled for q = 1 .. QF - 1.
template <class L, int COLL, int DES, class T, bool TWICE = false>
__device__ __forceinline__ void collide(T (&f)[L::Q], T gamma, const ModelParams& mp, T& rho, T (&j)[3]) {
  static_assert(!TWICE || kCollideTwice<COLL, DES>, "operator without a final halving");
  if constexpr (COLL == C_CASC) collide_cascaded<L, DES>(f, gamma, mp, rho, j);
  else collide_moment<L, COLL, DES, T, TWICE>(f, gamma, mp, rho, j);
}

// lam <- (dC/df)^T lam + functional source; dgam += (dC/dgamma)^T lam.
template <class L, int COLL, int DES, class T>
__device__ __forceinline__ void reverse_collide(const T (&f)[L::Q], T gamma, const ModelParams& mp, T (&lam)[L::Q],
                                                T& dgam, T srcA, const T (&srcB)[3], T srcAT) {
  if constexpr (COLL == C_CASC) reverse_cascaded<L, DES>(f, gamma, mp, lam, dgam, srcA, srcB);
  else reverse_moment<L, COLL, DES>(f, gamma, mp, lam, dgam, srcA, srcB, srcAT);
}

}  // namespace lbmx
