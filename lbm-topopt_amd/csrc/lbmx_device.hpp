// Device-side lattice Boltzmann operators for CDNA4 (gfx950): lattices, storage codecs,
// collision operators with their exact reverse (transpose) operators, boundary maps with
// their transposes, and functional terms. Everything is header-only and compile-time
// specialised: moment matrices are constexpr tables, so after unrolling every zero entry
// disappears and the work equals a hand-sparse implementation.
#pragma once
#include <hip/hip_runtime.h>

#include "lbmx_common.h"
#include "lbmx.h"

namespace lbmx {

// Multiply-adds are fused only within one statement in the functions marked with this, so that
// their rounding does not depend on the kernel they are inlined into (cross-statement fusion
// follows the basic blocks and the uses of each product there). Instances of one operator
// or functional in different kernels (single lattice, batch, tile shapes) then give bitwise
// equal results.
#define LBMX_CONTRACT_IN_STATEMENT _Pragma("clang fp contract(on)")

// ---------------------------------------------------------------------------------------
// Lattices (velocity order of src/latticeBoltzmann/lattices2d.hh:16-84 and
// lattices3d.hh:34-53: direction q and q+HALF are opposite).
// ---------------------------------------------------------------------------------------
struct LatD2Q9 {
  static constexpr int D = 2, Q = 9, QF = 9, QT = 0, HALF = 4;
  static constexpr int ex[Q] = {0, 1, 0, -1, -1, -1, 0, 1, 1};
  static constexpr int ey[Q] = {0, -1, -1, -1, 0, 1, 1, 1, 0};
  static constexpr int ez[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[Q] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
  static constexpr double w[Q] = {4. / 9., 1. / 36., 1. / 9., 1. / 36., 1. / 9.,
                                  1. / 36., 1. / 9., 1. / 36., 1. / 9.};
};

struct LatD3Q19 {
  static constexpr int D = 3, Q = 19, QF = 19, QT = 0, HALF = 9;
  static constexpr int ex[Q] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
  static constexpr int ey[Q] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
  static constexpr int ez[Q] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
  static constexpr int opp[Q] = {0, 10, 11, 12, 13, 14, 15, 16, 17, 18, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  static constexpr double w[Q] = {1. / 3., 1. / 18., 1. / 18., 1. / 18., 1. / 36., 1. / 36., 1. / 36.,
                                  1. / 36., 1. / 36., 1. / 36., 1. / 18., 1. / 18., 1. / 18.,
                                  1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

// D2Q9 flow populations followed by the D2Q5 temperature populations of one node, the
// interleave of ThermalDualCollision (src/dynamics/thermalDualCollision.hh:57-64).
struct LatD2Q9D2Q5 {
  static constexpr int D = 2, Q = 14, QF = 9, QT = 5, HALF = 4;
  static constexpr int ex[Q] = {0, 1, 0, -1, -1, -1, 0, 1, 1, 0, 0, -1, 0, 1};
  static constexpr int ey[Q] = {0, -1, -1, -1, 0, 1, 1, 1, 0, 0, -1, 0, 1, 0};
  static constexpr int ez[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[Q] = {0, 5, 6, 7, 8, 1, 2, 3, 4, 9, 12, 13, 10, 11};
  static constexpr double w[Q] = {4. / 9., 1. / 36., 1. / 9., 1. / 36., 1. / 9., 1. / 36., 1. / 9.,
                                  1. / 36., 1. / 9., 1. / 3., 1. / 6., 1. / 6., 1. / 6., 1. / 6.};
};

// ---------------------------------------------------------------------------------------
// Moment bases. Row k: polynomial of the velocity, relaxation kind (0 conserved, 1 the
// viscous rate omega, 2 a constant s), and equilibrium m_eq = a rho + b.v + v^T C v / rho_s.
// D2Q9: Lallemand & Luo (2000) with the rates of src/dynamics/incompressibleMRT2d.hh:88-90.
// D3Q19: d'Humieres et al. (2002) with MRT3d_Compile_constants (src/dynamics/MRT3d.hh:9-22).
// D2Q5: ThermalMRT2d<D2Q5_TMRT> (src/dynamics/thermalMRT2d.hh:50-101).
// ---------------------------------------------------------------------------------------
template <int Q>
struct Basis {
  double M[Q][Q];
  double invn2[Q];
  int rate[Q];
  double s[Q];
  double a[Q];
  double b[Q][3];
  double C[Q][3][3];
  int par[Q];         // +1: row even in c (depends on f_d + f_opp), -1: odd
  double G0[Q];       // sum_q M_kq w_q
  double G1[Q][3];    // sum_q M_kq w_q c_qa
  double G2[Q][3][3]; // sum_q M_kq w_q c_qa c_qb
  constexpr Basis() : M{}, invn2{}, rate{}, s{}, a{}, b{}, C{}, par{}, G0{}, G1{}, G2{} {}
  constexpr void finish(const int* ex, const int* ey, const int* ez, const double* w, const int* opp) {
    for (int k = 0; k < Q; ++k) {
      double n = 0;
      for (int q = 0; q < Q; ++q) n += M[k][q] * M[k][q];
      invn2[k] = 1.0 / n;
      bool even = true, odd = true;
      for (int q = 0; q < Q; ++q) {
        if (M[k][q] != M[k][opp[q]]) even = false;
        if (M[k][q] != -M[k][opp[q]]) odd = false;
      }
      par[k] = even ? 1 : (odd ? -1 : 0);
      for (int q = 0; q < Q; ++q) {
        const double c[3] = {double(ex[q]), double(ey[q]), double(ez[q])};
        G0[k] += M[k][q] * w[q];
        for (int x = 0; x < 3; ++x) {
          G1[k][x] += M[k][q] * w[q] * c[x];
          for (int y = 0; y < 3; ++y) G2[k][x][y] += M[k][q] * w[q] * c[x] * c[y];
        }
      }
    }
  }
};

constexpr Basis<9> make_basis_d2q9() {
  Basis<9> B;
  for (int q = 0; q < 9; ++q) {
    double cx = LatD2Q9::ex[q], cy = LatD2Q9::ey[q], c2 = cx * cx + cy * cy;
    B.M[0][q] = 1;
    B.M[1][q] = cx;
    B.M[2][q] = cy;
    B.M[3][q] = -4 + 3 * c2;
    B.M[4][q] = 4 - 10.5 * c2 + 4.5 * c2 * c2;
    B.M[5][q] = (-5 + 3 * c2) * cx;
    B.M[6][q] = (-5 + 3 * c2) * cy;
    B.M[7][q] = cx * cx - cy * cy;
    B.M[8][q] = cx * cy;
  }
  B.a[0] = 1;
  B.b[1][0] = 1;
  B.b[2][1] = 1;
  B.a[3] = -2; B.C[3][0][0] = 3; B.C[3][1][1] = 3; B.rate[3] = 2; B.s[3] = 1.4;
  B.a[4] = 1; B.C[4][0][0] = -3; B.C[4][1][1] = -3; B.rate[4] = 2; B.s[4] = 1.4;
  B.b[5][0] = -1; B.rate[5] = 2; B.s[5] = 1.2;
  B.b[6][1] = -1; B.rate[6] = 2; B.s[6] = 1.2;
  B.C[7][0][0] = 1; B.C[7][1][1] = -1; B.rate[7] = 1;
  B.C[8][0][1] = 0.5; B.C[8][1][0] = 0.5; B.rate[8] = 1;
  B.finish(LatD2Q9::ex, LatD2Q9::ey, LatD2Q9::ez, LatD2Q9::w, LatD2Q9::opp);
  return B;
}

constexpr Basis<19> make_basis_d3q19() {
  Basis<19> B;
  for (int q = 0; q < 19; ++q) {
    double cx = LatD3Q19::ex[q], cy = LatD3Q19::ey[q], cz = LatD3Q19::ez[q];
    double c2 = cx * cx + cy * cy + cz * cz;
    B.M[0][q] = 1;
    B.M[1][q] = cx;
    B.M[2][q] = cy;
    B.M[3][q] = cz;
    B.M[4][q] = 19 * c2 - 30;
    B.M[5][q] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
    B.M[6][q] = (5 * c2 - 9) * cx;
    B.M[7][q] = (5 * c2 - 9) * cy;
    B.M[8][q] = (5 * c2 - 9) * cz;
    B.M[9][q] = 3 * cx * cx - c2;
    B.M[10][q] = (3 * c2 - 5) * (3 * cx * cx - c2);
    B.M[11][q] = cy * cy - cz * cz;
    B.M[12][q] = (3 * c2 - 5) * (cy * cy - cz * cz);
    B.M[13][q] = cx * cy;
    B.M[14][q] = cy * cz;
    B.M[15][q] = cx * cz;
    B.M[16][q] = (cy * cy - cz * cz) * cx;
    B.M[17][q] = (cz * cz - cx * cx) * cy;
    B.M[18][q] = (cx * cx - cy * cy) * cz;
  }
  const double wxx = -0.5;
  B.a[0] = 1;
  for (int d = 0; d < 3; ++d) B.b[1 + d][d] = 1;
  B.a[4] = -11; for (int d = 0; d < 3; ++d) B.C[4][d][d] = 19; B.rate[4] = 2; B.s[4] = 1.19;
  B.a[5] = 3.; for (int d = 0; d < 3; ++d) B.C[5][d][d] = -5.5; B.rate[5] = 2; B.s[5] = 1.4;
  for (int d = 0; d < 3; ++d) { B.b[6 + d][d] = -2. / 3.; B.rate[6 + d] = 2; B.s[6 + d] = 1.2; }
  B.C[9][0][0] = 2; B.C[9][1][1] = -1; B.C[9][2][2] = -1; B.rate[9] = 1;
  B.C[10][0][0] = 2 * wxx; B.C[10][1][1] = -wxx; B.C[10][2][2] = -wxx; B.rate[10] = 2; B.s[10] = 1.4;
  B.C[11][1][1] = 1; B.C[11][2][2] = -1; B.rate[11] = 1;
  B.C[12][1][1] = wxx; B.C[12][2][2] = -wxx; B.rate[12] = 2; B.s[12] = 1.4;
  B.C[13][0][1] = 0.5; B.C[13][1][0] = 0.5; B.rate[13] = 1;
  B.C[14][1][2] = 0.5; B.C[14][2][1] = 0.5; B.rate[14] = 1;
  B.C[15][0][2] = 0.5; B.C[15][2][0] = 0.5; B.rate[15] = 1;
  for (int k = 16; k < 19; ++k) { B.rate[k] = 2; B.s[k] = 1.4; }
  B.finish(LatD3Q19::ex, LatD3Q19::ey, LatD3Q19::ez, LatD3Q19::w, LatD3Q19::opp);
  return B;
}

// thermal rows: T, q_x, q_y, e, p_xx ; m_eq = T (a + b.u)
constexpr Basis<5> make_basis_d2q5() {
  Basis<5> B;
  constexpr int ex[5] = {0, 0, -1, 0, 1};
  constexpr int ey[5] = {0, -1, 0, 1, 0};
  for (int q = 0; q < 5; ++q) {
    double cx = ex[q], cy = ey[q], c2 = cx * cx + cy * cy;
    B.M[0][q] = 1;
    B.M[1][q] = cx;
    B.M[2][q] = cy;
    B.M[3][q] = 5 * c2 - 4;
    B.M[4][q] = cx * cx - cy * cy;
  }
  B.a[0] = 1;
  B.b[1][0] = 1; B.rate[1] = 1;
  B.b[2][1] = 1; B.rate[2] = 1;
  B.a[3] = -2; B.rate[3] = 2; B.s[3] = 1.5;
  B.rate[4] = 2; B.s[4] = 1.5;
  constexpr int ez[5] = {0, 0, 0, 0, 0};
  constexpr double w[5] = {1. / 3., 1. / 6., 1. / 6., 1. / 6., 1. / 6.};
  constexpr int opp[5] = {0, 3, 4, 1, 2};
  B.finish(ex, ey, ez, w, opp);
  return B;
}

static constexpr Basis<9> kB9 = make_basis_d2q9();
static constexpr Basis<19> kB19 = make_basis_d3q19();
static constexpr Basis<5> kB5 = make_basis_d2q5();

template <int Q> struct BasisOf;
template <> struct BasisOf<9> { static constexpr const Basis<9>& B = kB9; };
template <> struct BasisOf<19> { static constexpr const Basis<19>& B = kB19; };

// ---------------------------------------------------------------------------------------
// Storage codecs. Plain forms store f; shifted forms store g = f/w_q - c_q in 16 bits, decoded
// as f = w_q g + w_q c_q. Flow directions shift by the rest density, c_q = 1, so that
// near-equilibrium values use the full mantissa. The D2Q5 temperature directions carry
// f = w_q T with T in the normalised range [0, 1]; they shift by its middle, c_q = 1/2, which
// halves the largest |g|, and with it the rounding step, against a shift of 1.
// ---------------------------------------------------------------------------------------
template <class L, int q>
__host__ __device__ constexpr double store_shift() {
  return q < L::QF ? 1.0 : 0.5;
}
// factor of the value handed to an encoder: 1/2 where the collision returns 2 f_q (TW)
template <class L, int q, bool TW>
__host__ __device__ constexpr double enc_scale() {
  return TW && q >= 1 && q < L::QF ? 0.5 : 1.0;
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
  return __uint_as_float(static_cast<unsigned int>(h) << 16);
}
__device__ __forceinline__ unsigned short f32_to_bf16(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return static_cast<unsigned short>(u >> 16);
}

// Stochastic rounding of the reduced storage forms. Rounding to nearest commits the same
// error every step on a nearly steady state, a persistent bias that the conserved modes do
// not damp; rounding up or down with probability given by the discarded bits makes the error
// zero-mean. The random bits are a hash of (node, step) spread over the directions, so a
// recomputed sweep reproduces the stored state exactly.
__device__ __forceinline__ unsigned int mix32(unsigned int x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// The random bits of direction q: bits sh .. sh + B - 1 of one of the node's two hashed words
// (q & 1), sh = ((q >> 1) * 3) % 17, moved to the top of the seed, where the 16-bit codecs read
// their B random bits (13 for FP16, 16 for BF16).
template <int q, int B>
__device__ __forceinline__ unsigned int sr_seed(unsigned int r0, unsigned int r1) {
  constexpr int sh = ((q >> 1) * 3) % 17;
  static_assert(sh + B <= 32, "window inside the word");
  return ((q & 1) ? r1 : r0) << (32 - B - sh);
}

struct PrecF64 {
  using S = double;   // state storage
  using T = double;   // arithmetic
  using G = double;   // design field storage
  using LS = double;  // adjoint storage
  static constexpr bool shifted = false;
  static constexpr bool stochastic = false;
  template <class L, int q> __device__ static T dec(S v) { return v; }
  template <class L, int q> __device__ static S enc(T v) { return v; }
};
struct PrecF32 {
  using S = float; using T = float; using G = float; using LS = float;
  static constexpr bool shifted = false;
  static constexpr bool stochastic = false;
  template <class L, int q> __device__ static T dec(S v) { return v; }
  template <class L, int q> __device__ static S enc(T v) { return v; }
};
struct PrecF16 {
  using S = _Float16; using T = float; using G = float; using LS = float;
  static constexpr bool shifted = true;
  template <class L, int q> __device__ static T dec(S v) {
    constexpr float w = static_cast<float>(L::w[q]);
    constexpr float wc = static_cast<float>(L::w[q] * store_shift<L, q>());
    return __builtin_fmaf(w, static_cast<float>(v), wc);
  }
  template <class L, int q> __device__ static S enc(T v) {
    constexpr float iw = static_cast<float>(1.0 / L::w[q]);
    constexpr float c = static_cast<float>(store_shift<L, q>());
    return static_cast<_Float16>(__builtin_fmaf(v, iw, -c));
  }
  static constexpr bool stochastic = true;
  static constexpr int kRandomBits = 13;
  // g = f/w - c rounded stochastically: the seed's top 13 bits are added to the FP32 pattern
  // below the FP16 fraction, and a round-toward-zero conversion drops them. Exact stochastic
  // rounding for FP16-normal values; below 2^-14 the conversion truncates to the subnormal grid.
  // (gfx950's stochastic conversion, which rounds stochastically there as well, changed the
  // 4000-step FP16 error of the 2048 x 1024 thermal channel from 7e-4 to 2e-3 against FP64, so
  // the FP16 form keeps this one.)
  // (TW: v is 2 f_q for q = 1 .. QF - 1; the factor 1/2 joins 1/w exactly)
  template <class L, int q, bool TW = false> __device__ static float sr_ready(T v, unsigned int seed) {
    constexpr float iw = static_cast<float>(enc_scale<L, q, TW>() / L::w[q]);
    constexpr float c = static_cast<float>(store_shift<L, q>());
    return __uint_as_float(__float_as_uint(__builtin_fmaf(v, iw, -c)) + (seed >> 19));
  }
  // directions q0 (low half) and q1 (high half) of one word: one conversion for both
  template <class L, int q0, int q1, bool TW = false>
  __device__ static unsigned int enc_pair(T v0, T v1, unsigned int seed0, unsigned int seed1) {
    const float t0 = sr_ready<L, q0, TW>(v0, seed0), t1 = sr_ready<L, q1, TW>(v1, seed1);
    return __builtin_bit_cast(unsigned int, __builtin_amdgcn_cvt_pkrtz(t0, t1));
  }
  // into half HI of `word`, the other half kept: gfx950's stochastic conversion with a zero
  // seed is the round-toward-zero conversion (equal on 4M samples over all FP32 exponents from
  // 2^-40 to 2^15, zeros included) and writes one half
  template <class L, int q, bool HI, bool TW = false>
  __device__ static unsigned int enc_sr(unsigned int word, T v, unsigned int seed) {
    typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
    const half2_t r =
        __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(half2_t, word), sr_ready<L, q, TW>(v, seed), 0u, HI);
    return __builtin_bit_cast(unsigned int, r);
  }
  template <bool HI> __device__ static S half_of(unsigned int word) {
    return __builtin_bit_cast(_Float16, static_cast<unsigned short>(HI ? word >> 16 : word & 0xffffu));
  }
};
struct PrecBF16 {
  using S = unsigned short; using T = float; using G = float; using LS = float;
  static constexpr bool shifted = true;
  template <class L, int q> __device__ static T dec(S v) {
    constexpr float w = static_cast<float>(L::w[q]);
    constexpr float wc = static_cast<float>(L::w[q] * store_shift<L, q>());
    return __builtin_fmaf(w, bf16_to_f32(v), wc);
  }
  template <class L, int q> __device__ static S enc(T v) {
    constexpr float iw = static_cast<float>(1.0 / L::w[q]);
    constexpr float c = static_cast<float>(store_shift<L, q>());
    return f32_to_bf16(__builtin_fmaf(v, iw, -c));
  }
  static constexpr bool stochastic = true;
  static constexpr int kRandomBits = 16;
  // g = f/w - c rounded stochastically into half HI of `word` (the other half is kept): the
  // gfx950 conversion adds the seed's top 16 bits to the FP32 pattern and truncates, as
  // rounding the BF16 way in software would (checked equal on a million samples of both signs)
  template <class L, int q, bool HI, bool TW = false>
  __device__ static unsigned int enc_sr(unsigned int word, T v, unsigned int seed) {
    typedef short short2_t __attribute__((ext_vector_type(2)));
    constexpr float iw = static_cast<float>(enc_scale<L, q, TW>() / L::w[q]);
    constexpr float c = static_cast<float>(store_shift<L, q>());
    const short2_t r =
        __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(short2_t, word), __builtin_fmaf(v, iw, -c), seed, HI);
    return __builtin_bit_cast(unsigned int, r);
  }
  template <class L, int q0, int q1, bool TW = false>
  __device__ static unsigned int enc_pair(T v0, T v1, unsigned int seed0, unsigned int seed1) {
    return enc_sr<L, q1, true, TW>(enc_sr<L, q0, false, TW>(0u, v0, seed0), v1, seed1);
  }
  template <bool HI> __device__ static S half_of(unsigned int word) {
    return static_cast<unsigned short>(HI ? word >> 16 : word & 0xffffu);
  }
};

// ---------------------------------------------------------------------------------------
// Compile-time loop helper: body(std::integral_constant-like) for q = 0..N-1.
// ---------------------------------------------------------------------------------------
template <int I> struct IC { static constexpr int value = I; };
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(IC<B>{});
    static_for<B + 1, E>(f);
  }
}

// ---------------------------------------------------------------------------------------
// Design interpolations (src/core/interpolationFunctions.hh:7-54)
// ---------------------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T bp_interp(T q, T s) { return T(1) - s * (T(1) + q) / (q + s); }
template <class T> __device__ __forceinline__ T bp_diff(T q, T s) {
  T d = q + s;
  return -(T(1) + q) / d + s * (T(1) + q) / (d * d);
}
template <class T> __device__ __forceinline__ T ramp_interp(T p, T ck, T s) {
  return (s * (ck * (T(1) + p) - T(1)) + T(1)) / (ck * (T(1) + p * s));
}
template <class T> __device__ __forceinline__ T ramp_diff(T p, T ck, T s) {
  T den = ck * (T(1) + p * s);
  return (ck * (T(1) + p) - T(1)) / den - ((s * (ck * (T(1) + p) - T(1)) + T(1)) * p) / (den * (T(1) + p * s));
}

// ---------------------------------------------------------------------------------------
// Boundary maps. Each piece is node-local; pieces run in insertion order after streaming
// (src/latticeBoltzmann/LBSolver.cc:288-293). Adjoints run the transposes in reverse order.
// A map whose written elements depend on a runtime orientation works on a local copy and
// writes every element back: stores to different elements on different branches would
// otherwise be merged into one store at a runtime index, which keeps the caller's population
// array out of registers (in scratch memory).
// ---------------------------------------------------------------------------------------
// prescribed value of a piece at node (x,y,z): value * profile(x,y,z) * tf, where tf is the
// piece's time factor at the current step (computed on the host, see tfac in the sweep args)
// (LBMX_PROFILE_TABLE: the factor comes from the piece's per-node table; comp picks the
// component of a table with one factor per velocity component)
template <class T>
__device__ __forceinline__ T bc_value(const DevBC& b, int D, T tf, int x, int y, int z, int comp = 0) {
  if (b.profile == LBMX_PROFILE_TABLE) {
    const long long i = ((long long)(z - b.tbox[4]) * (b.tbox[3] - b.tbox[2] + 1) + (y - b.tbox[2])) *
                            (b.tbox[1] - b.tbox[0] + 1) + (x - b.tbox[0]);
    return T(b.value) * T(b.table[i * b.ncomp + (b.ncomp > 1 ? comp : 0)]) * tf;
  }
  T s, tt;
  if (D == 2) { s = T(b.nrm[0] != 0 ? y : x); tt = 0; }
  else if (b.nrm[0] != 0) { s = T(y); tt = T(z); }
  else if (b.nrm[1] != 0) { s = T(x); tt = T(z); }
  else { s = T(x); tt = T(y); }
  T shape = T(1);
  if (b.profile == LBMX_PROFILE_PARABOLIC) {
    const T R = T(b.radius[0]), d = s - T(b.center[0]);
    shape = (R * R - d * d) / (R * R);
  } else if (b.profile == LBMX_PROFILE_PARABOLIC_DISK) {
    const T R = T(b.radius[0]), d0 = s - T(b.center[0]), d1 = tt - T(b.center[1]);
    const T r2 = d0 * d0 + d1 * d1;
    shape = r2 <= R * R ? T(1) - r2 / (R * R) : T(0);
  } else if (b.profile == LBMX_PROFILE_PARABOLIC_DUCT) {
    const T d0 = (s - T(b.center[0])) / T(b.radius[0]), d1 = (tt - T(b.center[1])) / T(b.radius[1]);
    shape = (T(1) - d0 * d0) * (T(1) - d1 * d1);
  }
  return T(b.value) * shape * tf;
}

// (bitwise, not short-circuit: all six bounds are read at once rather than one after another)
__device__ __forceinline__ bool in_box(const int* bx, int x, int y, int z) {
  return (x >= bx[0]) & (x <= bx[1]) & (y >= bx[2]) & (y <= bx[3]) & (z >= bx[4]) & (z <= bx[5]);
}

// on-grid bounce-back: q takes opp(q) when c_q.n <= 0 (incoming or tangential), per
// OnGridBounceback2d/3d (src/boundaryConditions/bounceback2d.cc:7-193, bounceback3d.cc:8-511);
// the (x0,y0,z0) corner of bounceback3d.cc:434-451 keeps f13 and f9 and sets f18 = f9.
// c ? a : b on the bit patterns. Selecting between two array elements with a runtime
// condition otherwise lets the optimiser turn the pair of loads into one load through a
// selected address, which forces the population array into private (scratch) memory.
__device__ __forceinline__ double bsel(bool c, double a, double b) {
  const unsigned long long m = 0ull - (unsigned long long)c;
  return __builtin_bit_cast(double, (__builtin_bit_cast(unsigned long long, a) & m) |
                                        (__builtin_bit_cast(unsigned long long, b) & ~m));
}
__device__ __forceinline__ float bsel(bool c, float a, float b) {
  const unsigned m = 0u - (unsigned)c;
  return __builtin_bit_cast(float, (__builtin_bit_cast(unsigned, a) & m) | (__builtin_bit_cast(unsigned, b) & ~m));
}

template <class L, int q>
__device__ __forceinline__ bool bb_from_opp(const int* n) {
  if constexpr (q == 0 || q >= L::QF) return false;
  else {
    int cn = L::ex[q] * n[0] + L::ey[q] * n[1] + L::ez[q] * n[2];
    bool r = cn <= 0;
    if constexpr (L::D == 3 && (q == 13 || q == 9)) {
      if (n[0] == -1 && n[1] == -1 && n[2] == -1) r = false;
    }
    return r;
  }
}

template <class L, class T>
__device__ __forceinline__ void bc_bounceback(T (&f)[L::Q], const int* n) {
  T o[L::QF];
  static_for<0, L::QF>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    o[q] = bsel(bb_from_opp<L, q>(n), f[L::opp[q]], f[q]);
  });
  static_for<0, L::QF>([&](auto qc) { f[decltype(qc)::value] = o[decltype(qc)::value]; });
}

template <class L, class T>
__device__ __forceinline__ void bc_bounceback_adj(T (&l)[L::Q], const int* n) {
  T o[L::QF];
  static_for<0, L::QF>([&](auto qc) {
    constexpr int r = decltype(qc)::value;
    T acc = bsel(bb_from_opp<L, r>(n), T(0), l[r]);
    if constexpr (r != 0) acc += bsel(bb_from_opp<L, L::opp[r]>(n), l[L::opp[r]], T(0));
    o[r] = acc;
  });
  static_for<0, L::QF>([&](auto qc) { l[decltype(qc)::value] = o[decltype(qc)::value]; });
}

// D2Q9 Zou-He (src/boundaryConditions/ZouHe2d.hh): incompressible velocity W/N/E,
// standard velocity W, pressure E/W (both variants).
template <class T>
__device__ __forceinline__ void zh_velocity(T* f, const int* n, bool inc, T ux, T uy) {
  T g[9];
  for (int q = 0; q < 9; ++q) g[q] = f[q];
  if (n[0] == -1) {
    if (inc) {
      g[1] = T(0.5) * (g[6] - g[2]) + g[5] + ux / T(6) - T(0.5) * uy;
      g[7] = T(0.5) * (g[2] - g[6]) + g[3] + ux / T(6) + T(0.5) * uy;
      g[8] = g[4] + T(2) / T(3) * ux;
    } else {
      T rho = (g[0] + g[2] + g[6] + T(2) * (g[3] + g[4] + g[5])) / (T(1) - ux);
      g[1] = g[5] + T(0.5) * (g[6] - g[2]) + rho * ux / T(6) - T(0.5) * rho * uy;
      g[7] = g[3] + T(0.5) * (g[2] - g[6]) + rho * ux / T(6) + T(0.5) * rho * uy;
      g[8] = g[4] + T(2) / T(3) * rho * ux;
    }
  } else if (n[1] == 1) {
    g[1] = T(0.5) * (g[4] - g[8]) + g[5] - uy / T(6) + T(0.5) * ux;
    g[2] = g[6] - T(2) / T(3) * uy;
    g[3] = T(0.5) * (g[8] - g[4]) + g[7] - uy / T(6) - T(0.5) * ux;
  } else {
    g[3] = g[7] + T(0.5) * (g[6] - g[2]) - ux / T(6) - T(0.5) * uy;
    g[4] = g[8] - T(2) / T(3) * ux;
    g[5] = g[1] + T(0.5) * (g[2] - g[6]) - ux / T(6) + T(0.5) * uy;
  }
  for (int q = 0; q < 9; ++q) f[q] = g[q];
}

template <class T>
__device__ __forceinline__ void zh_velocity_adj(T* l, const int* n) {
  T g[9];
  for (int q = 0; q < 9; ++q) g[q] = l[q];
  if (n[0] == -1) {
    g[2] += T(0.5) * (g[7] - g[1]);
    g[6] += T(0.5) * (g[1] - g[7]);
    g[5] += g[1]; g[3] += g[7]; g[4] += g[8];
    g[1] = 0; g[7] = 0; g[8] = 0;
  } else if (n[1] == 1) {
    g[4] += T(0.5) * (g[1] - g[3]);
    g[8] += T(0.5) * (g[3] - g[1]);
    g[5] += g[1]; g[6] += g[2]; g[7] += g[3];
    g[1] = 0; g[2] = 0; g[3] = 0;
  } else {
    g[7] += g[3];
    g[6] += T(0.5) * (g[3] - g[5]);
    g[2] += T(0.5) * (g[5] - g[3]);
    g[8] += g[4]; g[1] += g[5];
    g[3] = 0; g[4] = 0; g[5] = 0;
  }
  for (int q = 0; q < 9; ++q) l[q] = g[q];
}

template <class T>
__device__ __forceinline__ void zh_pressure(T* f, const int* n, bool inc, T r0) {
  T g[9];
  for (int q = 0; q < 9; ++q) g[q] = f[q];
  if (n[0] == 1) {
    T s = g[0] + g[2] + g[6] + T(2) * (g[1] + g[7] + g[8]);
    T u0 = inc ? s - r0 : T(-1) + s / r0;
    T sc = inc ? T(1) : r0;
    g[4] = g[8] - T(2) / T(3) * sc * u0;
    g[3] = g[7] - sc * u0 / T(6) + T(0.5) * (g[6] - g[2]);
    g[5] = g[1] - sc * u0 / T(6) + T(0.5) * (g[2] - g[6]);
  } else {
    T s = g[0] + g[2] + g[6] + T(2) * (g[3] + g[4] + g[5]);
    T u0 = inc ? r0 - s : T(1) - s / r0;
    T sc = inc ? T(1) : r0;
    g[1] = g[5] + sc * u0 / T(6) + T(0.5) * (g[6] - g[2]);
    g[7] = g[3] + sc * u0 / T(6) + T(0.5) * (g[2] - g[6]);
    g[8] = g[4] + T(2) / T(3) * sc * u0;
  }
  for (int q = 0; q < 9; ++q) f[q] = g[q];
}

template <class T>
__device__ __forceinline__ void zh_pressure_adj(T* l, const int* n) {
  T g[9];
  for (int q = 0; q < 9; ++q) g[q] = l[q];
  if (n[0] == 1) {
    T ub = -(T(2) / T(3)) * g[4] - g[3] / T(6) - g[5] / T(6);
    g[8] += g[4]; g[7] += g[3]; g[1] += g[5];
    T d = T(0.5) * (g[3] - g[5]);
    g[6] += d; g[2] -= d;
    g[0] += ub; g[2] += ub; g[6] += ub;
    g[1] += T(2) * ub; g[7] += T(2) * ub; g[8] += T(2) * ub;
    g[3] = 0; g[4] = 0; g[5] = 0;
  } else {
    T sb = -(g[1] / T(6) + g[7] / T(6) + (T(2) / T(3)) * g[8]);
    g[5] += g[1]; g[3] += g[7]; g[4] += g[8];
    T d = T(0.5) * (g[1] - g[7]);
    g[6] += d; g[2] -= d;
    g[0] += sb; g[2] += sb; g[6] += sb;
    g[3] += T(2) * sb; g[4] += T(2) * sb; g[5] += T(2) * sb;
    g[1] = 0; g[7] = 0; g[8] = 0;
  }
  for (int q = 0; q < 9; ++q) l[q] = g[q];
}

// D3Q19 Hecht & Harting (src/boundaryConditions/Hecht3d.hh): velocity west, pressure east.
template <class T>
__device__ __forceinline__ void hecht_velocity(T* f, bool inc, T ux, T uy, T uz) {
  T rho = T(1);
  if (!inc)
    rho = (f[0] + f[1] + f[2] + f[4] + f[5] + f[10] + f[11] + f[13] + f[14] +
           T(2) * (f[3] + f[6] + f[8] + f[16] + f[18])) / (T(1) - ux);
  T Nyx = T(0.5) * (f[5] + f[11] + f[13] - (f[2] + f[4] + f[14])) - rho * uy / T(3);
  T Nzx = T(0.5) * (f[10] + f[13] + f[14] - (f[1] + f[4] + f[5])) - rho * uz / T(3);
  f[7] = f[16] + rho * (ux - uz) / T(6) + Nzx;
  f[9] = f[18] + rho * (ux - uy) / T(6) + Nyx;
  f[12] = f[3] + rho * ux / T(3);
  f[15] = f[6] + rho * (ux + uz) / T(6) - Nzx;
  f[17] = f[8] + rho * (ux + uy) / T(6) - Nyx;
}

template <class T>
__device__ __forceinline__ void hecht_nbar(T* l, T nyx, T nzx) {
  T hy = T(0.5) * nyx, hz = T(0.5) * nzx;
  l[5] += hy; l[11] += hy; l[13] += hy; l[2] -= hy; l[4] -= hy; l[14] -= hy;
  l[10] += hz; l[13] += hz; l[14] += hz; l[1] -= hz; l[4] -= hz; l[5] -= hz;
}

template <class T>
__device__ __forceinline__ void hecht_velocity_adj(T* l) {
  T nzx = l[7] - l[15], nyx = l[9] - l[17];
  l[16] += l[7]; l[18] += l[9]; l[3] += l[12]; l[6] += l[15]; l[8] += l[17];
  hecht_nbar(l, nyx, nzx);
  l[7] = 0; l[9] = 0; l[12] = 0; l[15] = 0; l[17] = 0;
}

template <class T>
__device__ __forceinline__ void hecht_pressure(T* f, bool inc, T r0) {
  T s = f[0] + f[1] + f[2] + f[4] + f[5] + f[10] + f[11] + f[13] + f[14] +
        T(2) * (f[7] + f[9] + f[12] + f[15] + f[17]);
  T ux = inc ? s - r0 : T(-1) + s / r0;
  T sc = inc ? T(1) : r0;
  T Nyx = T(0.5) * (f[5] + f[11] + f[13] - (f[2] + f[4] + f[14]));
  T Nzx = T(0.5) * (f[10] + f[13] + f[14] - (f[1] + f[4] + f[5]));
  f[3] = f[12] - sc * ux / T(3);
  f[6] = f[15] - sc * ux / T(6) + Nzx;
  f[8] = f[17] - sc * ux / T(6) + Nyx;
  f[16] = f[7] - sc * ux / T(6) - Nzx;
  f[18] = f[9] - sc * ux / T(6) - Nyx;
}

template <class T>
__device__ __forceinline__ void hecht_pressure_adj(T* l) {
  T ub = -l[3] / T(3) - (l[6] + l[8] + l[16] + l[18]) / T(6);
  T nzx = l[6] - l[16], nyx = l[8] - l[18];
  l[12] += l[3]; l[15] += l[6]; l[17] += l[8]; l[7] += l[16]; l[9] += l[18];
  l[0] += ub; l[1] += ub; l[2] += ub; l[4] += ub; l[5] += ub;
  l[10] += ub; l[11] += ub; l[13] += ub; l[14] += ub;
  T u2 = T(2) * ub;
  l[7] += u2; l[9] += u2; l[12] += u2; l[15] += u2; l[17] += u2;
  hecht_nbar(l, nyx, nzx);
  l[3] = 0; l[6] = 0; l[8] = 0; l[16] = 0; l[18] = 0;
}

// D2Q5 part (offset QF): adiabatic bounce-back (ThermalBounceback2d, thermalBounceback2d.hh:273-490)
// and Dirichlet temperature (ThermalZouHeWest2d, thermalZouHe2d.hh:62-93).
__device__ __forceinline__ int d2q5_incoming_mask(const int* n) {
  // bit q set when c_q . n < 0 (q = 1..4 of D2Q5: (0,-1), (-1,0), (0,1), (1,0))
  int m = 0;
  if (-n[1] < 0) m |= 1 << 1;
  if (-n[0] < 0) m |= 1 << 2;
  if (n[1] < 0) m |= 1 << 3;
  if (n[0] < 0) m |= 1 << 4;
  return m;
}

template <class L, class T>
__device__ __forceinline__ void bc_node(T (&fio)[L::Q], const DevBC& b, T tf, int x, int y, int z) {
  T f[L::Q];
  static_for<0, L::Q>([&](auto qc) { f[decltype(qc)::value] = fio[decltype(qc)::value]; });
  switch (b.type) {
    case LBMX_BC_BOUNCEBACK: bc_bounceback<L>(f, b.nrm); break;
    case LBMX_BC_ZOUHE_VELOCITY:
      if constexpr (L::D == 2) {
        const T ux = bc_value<T>(b, 2, tf, x, y, z, 0) * T(b.dir[0]);
        const T uy = bc_value<T>(b, 2, tf, x, y, z, 1) * T(b.dir[1]);
        zh_velocity<T>(f, b.nrm, b.inc, ux, uy);
      }
      break;
    case LBMX_BC_ZOUHE_PRESSURE:
      if constexpr (L::D == 2) zh_pressure<T>(f, b.nrm, b.inc, bc_value<T>(b, 2, tf, x, y, z));
      break;
    case LBMX_BC_HECHT_VELOCITY:
      if constexpr (L::D == 3) {
        const T ux = bc_value<T>(b, 3, tf, x, y, z, 0) * T(b.dir[0]);
        const T uy = bc_value<T>(b, 3, tf, x, y, z, 1) * T(b.dir[1]);
        const T uz = bc_value<T>(b, 3, tf, x, y, z, 2) * T(b.dir[2]);
        hecht_velocity<T>(f, b.inc, ux, uy, uz);
      }
      break;
    case LBMX_BC_HECHT_PRESSURE:
      if constexpr (L::D == 3) hecht_pressure<T>(f, b.inc, bc_value<T>(b, 3, tf, x, y, z));
      break;
    case LBMX_BC_THERMAL_BOUNCEBACK:
      if constexpr (L::QT == 5) {
        int m = d2q5_incoming_mask(b.nrm);
        constexpr int o = L::QF;
        T g1 = f[o + 1], g2 = f[o + 2], g3 = f[o + 3], g4 = f[o + 4];
        if (m & 2) f[o + 1] = g3;
        if (m & 4) f[o + 2] = g4;
        if (m & 8) f[o + 3] = g1;
        if (m & 16) f[o + 4] = g2;
      }
      break;
    case LBMX_BC_THERMAL_DIRICHLET:
      if constexpr (L::QT == 5) {
        constexpr int o = L::QF;
        T Tv = bc_value<T>(b, 2, tf, x, y, z);
        T s = f[o] + f[o + 1] + f[o + 2] + f[o + 3] + f[o + 4];
        int m = d2q5_incoming_mask(b.nrm);
        if (m & 2) f[o + 1] = Tv - (s - f[o + 1]);
        else if (m & 4) f[o + 2] = Tv - (s - f[o + 2]);
        else if (m & 8) f[o + 3] = Tv - (s - f[o + 3]);
        else if (m & 16) f[o + 4] = Tv - (s - f[o + 4]);
      }
      break;
    default: break;
  }
  static_for<0, L::Q>([&](auto qc) { fio[decltype(qc)::value] = f[decltype(qc)::value]; });
}

template <class L, class T>
__device__ __forceinline__ void bc_node_adj(T (&lio)[L::Q], const DevBC& b) {
  T l[L::Q];
  static_for<0, L::Q>([&](auto qc) { l[decltype(qc)::value] = lio[decltype(qc)::value]; });
  switch (b.type) {
    case LBMX_BC_BOUNCEBACK: bc_bounceback_adj<L>(l, b.nrm); break;
    case LBMX_BC_ZOUHE_VELOCITY:
    case LBMX_BC_ZOUHE_NEUMANN:  // own block (the neighbours' part is sent separately)
      if constexpr (L::D == 2) zh_velocity_adj<T>(l, b.nrm);
      break;
    case LBMX_BC_ZOUHE_PRESSURE:
      if constexpr (L::D == 2) zh_pressure_adj<T>(l, b.nrm);
      break;
    case LBMX_BC_HECHT_VELOCITY:
      if constexpr (L::D == 3) hecht_velocity_adj<T>(l);
      break;
    case LBMX_BC_HECHT_PRESSURE:
      if constexpr (L::D == 3) hecht_pressure_adj<T>(l);
      break;
    case LBMX_BC_THERMAL_BOUNCEBACK:
      if constexpr (L::QT == 5) {
        int m = d2q5_incoming_mask(b.nrm);
        constexpr int o = L::QF;
        if (m & 2) { l[o + 3] += l[o + 1]; l[o + 1] = 0; }
        if (m & 4) { l[o + 4] += l[o + 2]; l[o + 2] = 0; }
        if (m & 8) { l[o + 1] += l[o + 3]; l[o + 3] = 0; }
        if (m & 16) { l[o + 2] += l[o + 4]; l[o + 4] = 0; }
      }
      break;
    case LBMX_BC_THERMAL_DIRICHLET:
    case LBMX_BC_THERMAL_NEUMANN:  // own block (the neighbours' part is sent separately)
      if constexpr (L::QT == 5) {
        constexpr int o = L::QF;
        int m = d2q5_incoming_mask(b.nrm);
        int u = (m & 2) ? 1 : (m & 4) ? 2 : (m & 8) ? 3 : 4;
        T lu = u == 1 ? l[o + 1] : u == 2 ? l[o + 2] : u == 3 ? l[o + 3] : l[o + 4];
        l[o] -= lu;
        if (u != 1) l[o + 1] -= lu; else l[o + 1] = 0;
        if (u != 2) l[o + 2] -= lu; else l[o + 2] = 0;
        if (u != 3) l[o + 3] -= lu; else l[o + 3] = 0;
        if (u != 4) l[o + 4] -= lu; else l[o + 4] = 0;
      }
      break;
    default: break;
  }
  static_for<0, L::Q>([&](auto qc) { lio[decltype(qc)::value] = l[decltype(qc)::value]; });
}

template <class L, int q, class T>
__device__ __forceinline__ T cdot(const T (&v)[3]) {
  T s = 0;
  if constexpr (L::ex[q] != 0) s += T(L::ex[q]) * v[0];
  if constexpr (L::ey[q] != 0) s += T(L::ey[q]) * v[1];
  if constexpr (L::ez[q] != 0) s += T(L::ez[q]) * v[2];
  return s;
}

template <class L, class T>
__device__ __forceinline__ void flow_macros(const T (&f)[L::Q], T& rho, T (&j)[3]) {
  T r = 0, jx = 0, jy = 0, jz = 0;
  static_for<0, L::QF>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    r += f[q];
    if constexpr (L::ex[q] == 1) jx += f[q];
    if constexpr (L::ex[q] == -1) jx -= f[q];
    if constexpr (L::ey[q] == 1) jy += f[q];
    if constexpr (L::ey[q] == -1) jy -= f[q];
    if constexpr (L::ez[q] == 1) jz += f[q];
    if constexpr (L::ez[q] == -1) jz -= f[q];
  });
  rho = r; j[0] = jx; j[1] = jy; j[2] = jz;
}

enum : int { C_BGK = LBMX_BGK, C_MRT = LBMX_MRT, C_CASC = LBMX_CASCADED,
             C_MRT_STD = LBMX_MRT_STANDARD, C_BGK_STD = LBMX_BGK_STANDARD };

template <int COLL> struct CollTraits {
  static constexpr bool inc = !(COLL == C_MRT_STD || COLL == C_BGK_STD);
  static constexpr bool mrt = (COLL == C_MRT || COLL == C_MRT_STD);
};

// ---------------------------------------------------------------------------------------
// Functional terms: value and gradient w.r.t. the node's pre-collision populations.
// ---------------------------------------------------------------------------------------
// value of a term from the node's pre-collision density, momentum and temperature (the sum of
// the D2Q5 populations; 0 on the flow lattices)
template <int COLL, class T>
__device__ __forceinline__ T term_value(int kind, T rho, const T (&j)[3], T temp) {
  LBMX_CONTRACT_IN_STATEMENT
  constexpr bool inc = CollTraits<COLL>::inc;
  switch (kind) {
    case 0: return rho;
    case 1: return inc ? j[0] : j[0] / rho;
    case 2: return (rho / T(3) + T(0.5) * (j[0] * j[0] + j[1] * j[1] + j[2] * j[2])) * j[0];
    case 3: return j[0] * temp;
    case 4: return temp;
  }
  return T(0);
}
template <class L, class T>
__device__ __forceinline__ T node_temperature(const T (&f)[L::Q]) {
  if constexpr (L::QT > 0) return f[L::QF] + f[L::QF + 1] + f[L::QF + 2] + f[L::QF + 3] + f[L::QF + 4];
  else return T(0);
}

// Gradient of a term in the form A + c_q.B on the flow populations (+ AT on the thermal ones).
template <class T>
__device__ __forceinline__ void term_source(int kind, T rho, const T (&j)[3], T temp, T wgt, T& A, T (&B)[3], T& AT) {
  switch (kind) {
    case 0: A += wgt; break;
    case 1: B[0] += wgt; break;
    case 2: {
      const T pt = rho / T(3) + T(0.5) * (j[0] * j[0] + j[1] * j[1] + j[2] * j[2]);
      A += wgt * j[0] / T(3);
      B[0] += wgt * (j[0] * j[0] + pt);
      B[1] += wgt * j[0] * j[1];
      B[2] += wgt * j[0] * j[2];
    } break;
    case 3:
      B[0] += wgt * temp;
      AT += wgt * j[0];
      break;
    case 4: AT += wgt; break;
  }
}


}  // namespace lbmx
