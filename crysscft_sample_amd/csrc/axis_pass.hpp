// One axis pass of the diffusion operator on gfx950.
//
// A reduced field of X*Y*Z values (row-major, z fastest, the layout of
// crysSCFT_sample.cpp:228-229 / :251-252) is viewed as n/M "lines" of length
// M along one axis.  A pass replaces every line v by At^T v, where At is the
// M x M per-axis operator of operators.hpp, and can fuse the pointwise
// E = exp(-w ds/2) multiply on the way in (pre) or out (post).  Three passes
// (z, y, x) make one diffusion(); with pre on the first and post on the last
// they make one contour step E D E of solve_diffusion_equation
// (crysSCFT_sample.cpp:387-395).
//
// Each workgroup (256 threads, 4 wave64s) owns TL consecutive lines: it
// stages the M x TL tile in LDS with coalesced loads, multiplies it by the
// operator held in LDS (FP64/FP32 FMA, RJ x RT register micro-tile per
// thread), writes the result back through LDS and stores it coalesced.
// Tiles are disjoint, so a pass may run in place (out == in).
//
// blockIdx.y selects one of up to two independent propagator chains that
// advance in the same launch (the forward and the backward chain of one SCFT
// iteration).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace crys {

template <typename T>
struct PassChain {
  const T* in;
  T* out;
  const T* pre;   // may be nullptr
  const T* post;  // may be nullptr
};

template <typename T>
struct PassParams {
  PassChain<T> c[2];
  const T* mat;                 // At, M x M, At[i*M + j]
  const unsigned short* mat_lo; // FP32 only: bf16 bits of At - (float)At, or nullptr
  int X, Y, Z;
  int cen;                      // 1: mat is the half-size centrosymmetric operator
  int shift;                    // with cen: half-period pairing (see cen_partner)
  int stream;                   // 1: non-temporal field traffic (field_load)
};

// Pairing of the half-size (centrosymmetric) products.  A line is split into
// u_i = x_i + x_{p(i)} and v_i = x_i - x_{p(i)} for i < M/2, where the
// partner p(i) is the mirror M-1-i (PmmmDCT, P1: D[M-1-j][M-1-k] = D[j][k])
// or, with `shift`, the half-period translation i + M/2 (3m with zero
// translational parts: D[j+M/2][k+M/2] = D[j][k]).  The products run in the
// paired basis, whose position M-1-i holds the element p(i); cen_index maps
// a position of that basis to its index in memory.
template <int MP>
__device__ __forceinline__ int cen_partner(int i, bool shift) {
  return shift ? i + MP / 2 : MP - 1 - i;
}
template <int MP>
__device__ __forceinline__ int cen_index(int i, bool shift) {
  return (shift && i >= MP / 2) ? (3 * MP / 2 - 1) - i : i;
}

// Field loads and stores of a pass: non-temporal when NT (the launch streams
// fields that the next launch reads from other XCDs, as in a propagation),
// ordinary otherwise (a lone diffusion() re-reads one field that stays in
// L2).
template <bool NT, typename T>
__device__ __forceinline__ T field_load(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void field_store(T v, T* p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Register micro-tile width along the lines for a padded length MP.
template <int MP>
struct PassShape {
  static constexpr int RJ = MP / 16;                                     // rows per thread
  static constexpr int RT = MP <= 32 ? 4 : (MP <= 64 ? 2 : 1);           // lines per thread
  static constexpr int TL = 16 * RT;                                     // lines per workgroup
  static constexpr int LDT = TL + 1;                                     // LDS row pitch
  static constexpr bool MAT_LDS = MP <= 64;
};

template <typename T, int AXIS, int MP>
__global__ __launch_bounds__(256) void axis_pass_kernel(PassParams<T> p) {
  using S = PassShape<MP>;
  constexpr int NT = 256;
  constexpr int RJ = S::RJ, RT = S::RT, TL = S::TL, LDT = S::LDT;
  constexpr int KT = MP * TL / NT;             // tile elements per thread
  constexpr int KM = (MP * MP + NT - 1) / NT;  // operator elements per thread
  static_assert(MP * TL % NT == 0, "tile must split evenly over the workgroup");
  __shared__ T tile[MP * LDT];
  __shared__ T mat[S::MAT_LDS ? MP * MP : 1];

  const PassChain<T> ch = p.c[blockIdx.y];
  const int X = p.X, Y = p.Y, Z = p.Z;
  const int M = AXIS == 0 ? X : (AXIS == 1 ? Y : Z);
  const int YZ = Y * Z;
  const int nlines = X * YZ / M;
  const int l0 = blockIdx.x * TL;
  const int nl = nlines - l0 < TL ? nlines - l0 : TL;  // valid lines in this tile
  const int tid = threadIdx.x;
  // a y-pass tile stays inside one x plane when TL divides Z
  const bool yplane = AXIS == 1 && Z % TL == 0;
  const int x0 = yplane ? l0 / Z : 0;
  const int z0 = yplane ? l0 - x0 * Z : 0;

  // ---- element (row i, line ll) handled by slot k of this thread ----
  int gi[KT], lo[KT];
  bool ok[KT];
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int f = tid + NT * k;
    int i, ll;
    if constexpr (AXIS == 2) {
      ll = f / MP;
      i = f - ll * MP;
    } else {
      i = f / TL;
      ll = f - i * TL;
    }
    ok[k] = i < M && ll < nl;
    lo[k] = i * LDT + ll;
    int g;
    if constexpr (AXIS == 2) {
      g = (l0 + ll) * Z + i;
    } else if constexpr (AXIS == 1) {
      if (yplane) {
        g = x0 * YZ + i * Z + z0 + ll;
      } else {
        const int l = l0 + ll, x = l / Z;
        g = x * YZ + i * Z + (l - x * Z);
      }
    } else {
      g = i * YZ + l0 + ll;
    }
    gi[k] = ok[k] ? g : 0;
  }

  // ---- issue every global load first, then fill LDS ----
  T v[KT];
#pragma unroll
  for (int k = 0; k < KT; ++k) v[k] = ok[k] ? ch.in[gi[k]] : T(0);
  if (ch.pre) {
    T e[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) e[k] = ok[k] ? ch.pre[gi[k]] : T(0);
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] *= e[k];
  }
  if constexpr (S::MAT_LDS) {
    T m[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const int f = tid + NT * k;
      const int i = f / MP, j = f - (f / MP) * MP;
      m[k] = (f < MP * MP && i < M && j < M) ? p.mat[i * M + j] : T(0);
    }
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const int f = tid + NT * k;
      if (f < MP * MP) mat[f] = m[k];
    }
  }
#pragma unroll
  for (int k = 0; k < KT; ++k) tile[lo[k]] = v[k];
  __syncthreads();

  // ---- out[j][t] = sum_i At[i][j] * tile[i][t] ----
  const int tx = tid & 15, ty = tid >> 4;
  T acc[RJ][RT];
#pragma unroll
  for (int r = 0; r < RJ; ++r)
#pragma unroll
    for (int c = 0; c < RT; ++c) acc[r][c] = T(0);

  auto body = [&](int i) {
    T a[RJ], b[RT];
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
      const int j = ty * RJ + r;
      if constexpr (S::MAT_LDS) a[r] = mat[i * MP + j];
      else a[r] = j < M ? p.mat[i * M + j] : T(0);
    }
#pragma unroll
    for (int c = 0; c < RT; ++c) b[c] = tile[i * LDT + tx * RT + c];
#pragma unroll
    for (int r = 0; r < RJ; ++r)
#pragma unroll
      for (int c = 0; c < RT; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
  };
  if (M == MP) {
#pragma unroll 8
    for (int i = 0; i < MP; ++i) body(i);
  } else {
#pragma unroll 2
    for (int i = 0; i < M; ++i) body(i);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RJ; ++r)
#pragma unroll
    for (int c = 0; c < RT; ++c) tile[(ty * RJ + r) * LDT + tx * RT + c] = acc[r][c];
  __syncthreads();

  // ---- coalesced store, with the optional post multiply ----
#pragma unroll
  for (int k = 0; k < KT; ++k) v[k] = tile[lo[k]];
  if (ch.post) {
    T e[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) e[k] = ok[k] ? ch.post[gi[k]] : T(0);
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] *= e[k];
  }
#pragma unroll
  for (int k = 0; k < KT; ++k)
    if (ok[k]) ch.out[gi[k]] = v[k];
}

// ---------------------------------------------------------------------------
// Matrix-core variant (MP <= 64): the same pass with the product on the
// 16x16x4 MFMA of the element type (v_mfma_f64_16x16x4_f64 /
// v_mfma_f32_16x16x4_f32).  Out(MP x TL) = D(MP x MP) * Tile(MP x TL): wave w
// owns 16 output rows (j tile w % JT) and TPW column tiles of 16 lines; its
// A operands (D[j][i] = At[i][j], 16 x 4 per step) stay in registers for the
// whole pass, the B operands (4 x 16 tile slices) come from LDS.
// Lane layout (16x16x4, one block): A lane l = A[l%16][l/16],
// B lane l = B[l/16][l%16]; the result register r of lane l holds
// D[4(l/16)+r][l%16] for f32 and D[4r+l/16][l%16] for f64 (checked on gfx950
// with known operands).
// ---------------------------------------------------------------------------
template <typename T>
struct MfmaOp;
template <>
struct MfmaOp<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <>
struct MfmaOp<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef short s4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // 16x16x16 bf16 product into the same f32 accumulator; lane l holds
  // A[l%16][4(l/16)+i] and B[4(l/16)+i][l%16] (same result layout as f32).
  static __device__ inline v4 mma_bf16(s4 a, s4 b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Writes one accumulator element to LDS.  FP32 writes are never merged into
// ds_write2_b32 pairs: hipcc (ROCm 7.2, gfx950) rebases such a pair when its
// offsets exceed the 8-bit dword range and has emitted a wrong offset for one
// element of the pair (mfma_pass_tile<float, *, 32, true>: byte 240 written
// instead of 960).  An empty asm statement with a memory clobber after each
// store keeps the stores apart.  (A volatile pointer kept them apart too, but
// compiled to a flat store of system scope followed by a wait for every
// outstanding global access.)  KEEP_APART = false lets them merge: the
// class-64 split phase's default products (not its k-half variant), whose
// merged pairs have the right offsets (rows 80 dwords apart, 16- and 32-row
// strides in st64 units), 0.5% faster per FP32 step.
template <typename T, bool KEEP_APART = true>
__device__ __forceinline__ void lds_put(T* p, int idx, T v) {
  p[idx] = v;
  if constexpr (KEEP_APART && std::is_same<T, float>::value) asm volatile("" ::: "memory");
}

// Four floats to bf16 bits, round to nearest even, in two hardware
// conversions of a pair each (v_cvt_pk_bf16_f32): measured 3.5% faster per
// FP32 contour step than rounding each value with integer operations and
// packing the results, with identical bits.
__device__ inline MfmaOp<float>::s4 bf16x4(const float (&x)[4]) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef __bf16 b2 __attribute__((ext_vector_type(2)));
  const b2 lo = __builtin_convertvector((f2{x[0], x[1]}), b2);
  const b2 hi = __builtin_convertvector((f2{x[2], x[3]}), b2);
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(MfmaOp<float>::s4,
                            (u2{__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)}));
}

// Far-first order of the KS k groups of an operator tile: the groups sorted
// by decreasing distance from the tile's diagonal, which sits at group
// position 4 jb + 1.5 for row block jb (two pointers from the ends inwards),
// so that the small entries far from the diagonal are summed first.  The
// orders of all NJB row blocks are built at compile time.
template <int KS, int NJB>
struct FarFirstTable {
  int v[NJB][KS];
  constexpr FarFirstTable() : v{} {
    for (int j = 0; j < NJB; ++j) {
      const float gc = 4.0f * j + 1.5f;
      int lo = 0, hi = KS - 1;
      for (int n = 0; n < KS; ++n) {
        const float dl = lo > gc ? lo - gc : gc - lo;
        const float dh = hi > gc ? hi - gc : gc - hi;
        if (dl >= dh) v[j][n] = lo++;
        else v[j][n] = hi--;
      }
    }
  }
};

// The order of row block jb (0 <= jb < NJB): selects among the compile-time
// orders, without branches or a table in memory.  jb is the same in every
// lane of the wave (every caller derives it from the wave index).  With
// SCALAR the selection runs on the scalar unit and the order lives in scalar
// registers: the FP64 kernels take it (the dense class-64 multi-term phase
// stopped spilling, 43 -> 32 us per step); FP32 kernels measured faster with
// the selection in vector registers (the same phase 27.5 against 30.4 us).
template <int KS, int NJB, bool SCALAR = false>
__device__ __forceinline__ void far_first_order(int jb, int (&ord)[KS]) {
  constexpr FarFirstTable<KS, NJB> t{};
  const int j0 = SCALAR ? __builtin_amdgcn_readfirstlane(jb) : jb;
#pragma unroll
  for (int n = 0; n < KS; ++n) {
    int o = t.v[0][n];
#pragma unroll
    for (int j = 1; j < NJB; ++j) o = j0 == j ? t.v[j][n] : o;
    ord[n] = o;
  }
}

template <int MP>
struct MfmaShape {
  static constexpr int NT = MP <= 64 ? 256 : 512;  // threads: one wave per row tile at 128
  static constexpr int TL = MP <= 32 ? 64 : 32;  // lines per workgroup
  static constexpr int LDT = TL + 16;            // LDS pitch: 4 B-rows hit distinct banks
  static constexpr int JT = MP / 16;             // 16-row output tiles
  static constexpr int TT = TL / 16;             // 16-line output tiles
  static constexpr int KS = MP / 4;              // MFMA k steps
  static constexpr int WPJ = (NT / 64) / JT >= 1 ? (NT / 64) / JT : 1;  // waves per row tile
  static constexpr int TPW = TT / WPJ;           // line tiles per wave
  static_assert(TT % WPJ == 0, "line tiles must split over the waves of a row tile");
  static_assert(JT * WPJ <= NT / 64, "every row tile needs a wave");
};

// One workgroup's tile `tidx` of a pass on one chain; `tile` is LDS of at
// least MP * MfmaShape<MP>::LDT elements.  Shared by the single-field kernel
// and the batched kernel.  With CEN (centrosymmetric operator, M == MP,
// MP % 32 == 0) the lines are split into u/v halves and `mat` is the
// half-size operator At_c (M/2 x M) described at SlabShape, paired as
// cen_partner(., shift) says.  The k groups of every output tile are summed
// far-first (see SlabShape).
template <typename T, int AXIS, int MP, bool CEN = false, bool NTM = true>
__device__ __forceinline__ void mfma_pass_tile(const PassChain<T>& ch, const T* __restrict__ mat,
                                               const unsigned short* __restrict__ mat_lo, int X,
                                               int Y, int Z, int tidx, T* tile,
                                               bool shift = false) {
  using S = MfmaShape<MP>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int NT = S::NT;
  constexpr int TL = S::TL, LDT = S::LDT, JT = S::JT, WPJ = S::WPJ, TPW = S::TPW;
  constexpr int KS = CEN ? MP / 8 : MP / 4;
  constexpr int H = MP / 2;
  constexpr int ROWS = CEN ? H : MP;          // staged rows (pairs for CEN)
  constexpr int KT = (ROWS * TL + NT - 1) / NT;  // slots per thread
  static_assert(!CEN || MP % 32 == 0, "centrosymmetric split needs M % 32 == 0");
  constexpr bool SPLIT = std::is_same<T, float>::value;

  const int M = AXIS == 0 ? X : (AXIS == 1 ? Y : Z);
  const int YZ = Y * Z;
  const int nlines = X * YZ / M;
  const int l0 = tidx * TL;
  const int nl = nlines - l0 < TL ? nlines - l0 : TL;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const bool yplane = AXIS == 1 && Z % TL == 0;
  const int x0 = yplane ? l0 / Z : 0;
  const int z0 = yplane ? l0 - x0 * Z : 0;

  // global index of element (row i, line ll)
  auto gidx = [&](int i, int ll) -> int {
    if constexpr (AXIS == 2) {
      return (l0 + ll) * Z + i;
    } else if constexpr (AXIS == 1) {
      if (yplane) return x0 * YZ + i * Z + z0 + ll;
      const int l = l0 + ll, x = l / Z;
      return x * YZ + i * Z + (l - x * Z);
    } else {
      return i * YZ + l0 + ll;
    }
  };

  // ---- slot k of this thread: row i (< ROWS), line ll ----
  int si[KT], sl[KT];
  bool ok[KT];
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int f = tid + NT * k;
    int i, ll;
    if constexpr (AXIS == 2) {
      ll = f / ROWS;
      i = f - ll * ROWS;
    } else {
      i = f / TL;
      ll = f - i * TL;
    }
    si[k] = i;
    sl[k] = ll;
    ok[k] = f < ROWS * TL && i < (CEN ? H : M) && ll < nl;
  }

  // operands of this wave: row tile jt, k groups far-first
  const int jt = w % JT, sub = w / JT;
  const bool active = w < JT * WPJ;
  const int jrow = jt * 16 + (lane & 15);
  const int jb = CEN ? (jt % (JT / 2)) : jt;
  const int kbase = CEN ? (jt < JT / 2 ? 0 : H) : 0;
  int ord[KS];
  far_first_order<KS, CEN ? JT / 2 : JT, sizeof(T) == 8>(jb, ord);
  const int ldm = CEN ? MP : M;
  T a[KS];
  short alo[SPLIT ? KS : 1];
  int brow[KS];
#pragma unroll
  for (int n = 0; n < KS; ++n) {
    const int i = 4 * ord[n] + (lane >> 4);
    brow[n] = kbase + i;
    const bool okm = active && (CEN || (i < M && jrow < M));
    a[n] = okm ? mat[i * ldm + jrow] : T(0);
    if constexpr (SPLIT) alo[n] = (okm && mat_lo) ? (short)mat_lo[i * ldm + jrow] : (short)0;
  }

  // ---- stage (issue every global load first).  With NTM the lines are read
  // and written non-temporally, so that the E values and the operators stay
  // in L2: measured 7% faster FP64 (3% FP32) on a batch of 256 instances,
  // whose contour steps are these passes; a lone diffusion() (NTM off) measured
  // 4% slower with it ----
  if constexpr (!CEN) {
    T v[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] = ok[k] ? field_load<NTM>(&ch.in[gidx(si[k], sl[k])]) : T(0);
    if (ch.pre) {
      T e[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) e[k] = ok[k] ? ch.pre[gidx(si[k], sl[k])] : T(0);
#pragma unroll
      for (int k = 0; k < KT; ++k) v[k] *= e[k];
    }
#pragma unroll
    for (int k = 0; k < KT; ++k)
      if (tid + NT * k < ROWS * TL) tile[si[k] * LDT + sl[k]] = v[k];
  } else {
    T x1[KT], x2[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      x1[k] = ok[k] ? field_load<NTM>(&ch.in[gidx(si[k], sl[k])]) : T(0);
      x2[k] = ok[k] ? field_load<NTM>(&ch.in[gidx(cen_partner<MP>(si[k], shift), sl[k])]) : T(0);
    }
    if (ch.pre) {
      T e1[KT], e2[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        e1[k] = ok[k] ? ch.pre[gidx(si[k], sl[k])] : T(0);
        e2[k] = ok[k] ? ch.pre[gidx(cen_partner<MP>(si[k], shift), sl[k])] : T(0);
      }
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        x1[k] *= e1[k];
        x2[k] *= e2[k];
      }
    }
#pragma unroll
    for (int k = 0; k < KT; ++k)
      if (tid + NT * k < ROWS * TL) {
        tile[si[k] * LDT + sl[k]] = x1[k] + x2[k];
        tile[(H + si[k]) * LDT + sl[k]] = x1[k] - x2[k];
      }
  }
  __syncthreads();

  // ---- out[j][t] = sum_k A[j][k] * tile[k][t] ----
  V4 acc[TPW];
  const bool use_lo = SPLIT && mat_lo != nullptr;
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    acc[u] = V4{T(0), T(0), T(0), T(0)};
    if (active) {
      const int col = (sub * TPW + u) * 16 + (lane & 15);
      T b[KS];
#pragma unroll
      for (int n = 0; n < KS; ++n) b[n] = tile[brow[n] * LDT + col];
      // low parts first (see slab_product)
      if constexpr (SPLIT) {
        if (use_lo) {
#pragma unroll
          for (int c4 = 0; c4 < (KS + 3) / 4; ++c4) {
            typename MfmaOp<float>::s4 av;
            float fb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int n = 4 * c4 + i;
              av[i] = n < KS ? alo[n] : (short)0;
              fb[i] = n < KS ? b[n] : 0.f;
            }
            acc[u] = MfmaOp<float>::mma_bf16(av, bf16x4(fb), acc[u]);
          }
        }
      }
#pragma unroll
      for (int n = 0; n < KS; ++n) acc[u] = MfmaOp<T>::mma(a[n], b[n], acc[u]);
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int col = (sub * TPW + u) * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lds_put<T>(tile, (jt * 16 + MfmaOp<T>::row(lane, r)) * LDT + col, acc[u][r]);
    }
  }
  __syncthreads();

  // ---- coalesced store, with the optional post multiply ----
  if constexpr (!CEN) {
    T v[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] = ok[k] ? tile[si[k] * LDT + sl[k]] : T(0);
    if (ch.post) {
      T e[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) e[k] = ok[k] ? ch.post[gidx(si[k], sl[k])] : T(0);
#pragma unroll
      for (int k = 0; k < KT; ++k) v[k] *= e[k];
    }
#pragma unroll
    for (int k = 0; k < KT; ++k)
      if (ok[k]) field_store<NTM>(v[k], &ch.out[gidx(si[k], sl[k])]);
  } else {
    T o1[KT], o2[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const T U = ok[k] ? tile[si[k] * LDT + sl[k]] : T(0);
      const T V = ok[k] ? tile[(H + si[k]) * LDT + sl[k]] : T(0);
      o1[k] = U + V;
      o2[k] = U - V;
    }
    if (ch.post) {
      T e1[KT], e2[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        e1[k] = ok[k] ? ch.post[gidx(si[k], sl[k])] : T(0);
        e2[k] = ok[k] ? ch.post[gidx(cen_partner<MP>(si[k], shift), sl[k])] : T(0);
      }
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        o1[k] *= e1[k];
        o2[k] *= e2[k];
      }
    }
#pragma unroll
    for (int k = 0; k < KT; ++k)
      if (ok[k]) {
        field_store<NTM>(o1[k], &ch.out[gidx(si[k], sl[k])]);
        field_store<NTM>(o2[k], &ch.out[gidx(cen_partner<MP>(si[k], shift), sl[k])]);
      }
  }
}

template <typename T, int AXIS, int MP, bool CEN, bool NTM>
__global__ __launch_bounds__(MfmaShape<MP>::NT) void axis_pass_mfma_kernel(PassParams<T> p) {
  __shared__ T tile[MP * MfmaShape<MP>::LDT];
  mfma_pass_tile<T, AXIS, MP, CEN, NTM>(p.c[blockIdx.y], p.mat, p.mat_lo, p.X, p.Y, p.Z,
                                        blockIdx.x, tile, p.shift != 0);
}

// Whether a matrix-core pass reads and writes its lines non-temporally: in a
// propagation (stream), and always for FP32, whose lone diffusion() also
// measured faster with it (FP64: 4% slower).  Separate kernels for the two
// cases: one kernel holding both bodies measured up to 20% slower (FP32).
template <typename T>
inline bool pass_nt(const PassParams<T>& p) {
  return p.stream || sizeof(T) == 4;
}

// Padded length class for an axis of length M (0 if unsupported).
inline int pass_class(int M) {
  if (M <= 0) return 0;
  if (M <= 16) return 16;
  if (M <= 32) return 32;
  if (M <= 48) return 48;
  if (M <= 64) return 64;
  if (M <= 128) return 128;
  return 0;
}

// Which product the passes use: matrix cores (default) or the FMA kernel.
enum class PassImpl { Auto, Mfma, Valu, Radix };
PassImpl pass_impl();

// ---------------------------------------------------------------------------
// Fused z + y passes on x-slabs.  One workgroup (512 threads, 8 waves) owns
// the Y x Z plane of one x index: it stages the plane in LDS transposed
// (z-major), applies Dz along z and Dy along y with the matrix cores, and
// stores the plane back.  With the x pass that follows, a contour step takes
// two launches instead of three.  Requires Y and Z in the same length class
// MP <= 64.
//
// CEN (centrosymmetric operators, D[M-1-j][M-1-k] = D[j][k]: PmmmDCT and P1
// with M = MP a multiple of 32): each line x is split into u = x_j + x_{M-1-j}
// and v = x_j - x_{M-1-j} (j < M/2); then U = Ps u and V = Pa v with the
// half-size blocks Ps = (D[j][k] + D[j][M-1-k])/2, Pa = (D[j][k] -
// D[j][M-1-k])/2, and out_j = U_j + V_j, out_{M-1-j} = U_j - V_j.  This halves
// the products.  The half-size operator is stored as At_c[k][j] with
// k < M/2, j < M: columns j < M/2 hold Ps^T, columns M/2 + j hold Pa^T.  For
// 3m the partner of j is j + M/2 instead of M-1-j (cen_partner).
//
// Every 16-row output tile sums its k groups (4 rows each) from the group
// farthest from the tile's diagonal to the nearest, so the large near-diagonal
// terms are added last and the running sum is rounded fewer times at full
// magnitude (this matters for FP32 accumulation).
// ---------------------------------------------------------------------------
template <int MP, bool CEN>
struct SlabShape {
  static constexpr int NT = 512;
  static constexpr int P = MP + 16;              // LDS pitch: conflict-free B reads
  static constexpr int JT = MP / 16;             // 16-row output tiles
  static constexpr int KS = CEN ? MP / 8 : MP / 4;  // MFMA k steps per output tile
  // Wave w owns row tile w % JT (its operator registers) and the column
  // tiles s, s + WPJ, ... with s = w / JT < WPJ.
  static constexpr int WPJ = (NT / 64) / JT;       // waves per row tile
  static constexpr int TPW = (JT + WPJ - 1) / WPJ;  // column tiles per wave
  static_assert(!CEN || MP % 32 == 0, "centrosymmetric split needs M % 32 == 0");
  // column tile u of wave w, or -1
  static __device__ __forceinline__ int col_tile(int w, int u) {
    const int s = w / JT, ct = s + WPJ * u;
    return (s < WPJ && ct < JT) ? ct : -1;
  }
};

// Operand registers of one wave for one product: A (and its bf16 low part
// for FP32) of row tile jt in the far-first group order, and the LDS row of
// each group.
template <typename T, int MP, bool CEN>
struct SlabOperands {
  static constexpr int KS = SlabShape<MP, CEN>::KS;
  T a[KS];
  short alo[std::is_same<T, float>::value ? KS : 1];
  int brow[KS];  // LDS row of the B operand of this lane for each step
};

// `pitch` (> 0) overrides the row length of the stored operator, for a copy
// held in LDS at a padded pitch.
template <typename T, int MP, bool CEN>
__device__ __forceinline__ void slab_load_operands(const T* __restrict__ mat,
                                                   const unsigned short* __restrict__ lo, int M,
                                                   int jt, int lane,
                                                   SlabOperands<T, MP, CEN>& op, int pitch = 0) {
  using S = SlabShape<MP, CEN>;
  constexpr int KS = S::KS, JT = S::JT;
  constexpr bool SPLIT = std::is_same<T, float>::value;
  const int half = JT / 2;
  // row block of this tile inside its k space (the whole M, or one half)
  const int jb = CEN ? (jt % half) : jt;
  const int kbase = CEN ? (jt < half ? 0 : MP / 2) : 0;
  int ord[KS];
  far_first_order<KS, CEN ? JT / 2 : JT, sizeof(T) == 8>(jb, ord);
  const int jrow = jt * 16 + (lane & 15);
  const int ldm = pitch > 0 ? pitch : (CEN ? MP : M);  // row length of the stored operator
#pragma unroll
  for (int n = 0; n < KS; ++n) {
    const int i = 4 * ord[n] + (lane >> 4);  // k index inside the k space
    op.brow[n] = kbase + i;
    const bool ok = CEN || (i < M && jrow < M);
    op.a[n] = ok ? mat[i * ldm + jrow] : T(0);
    if constexpr (SPLIT) op.alo[n] = (ok && lo) ? (short)lo[i * ldm + jrow] : (short)0;
  }
}

// One product of the slab for the tiles of wave w: out[j][c] = sum_k A[j][k]
// src[k][c], src in lds[k * P + c]; the results stay in registers.
template <typename T, int MP, bool CEN>
__device__ __forceinline__ void slab_product(const SlabOperands<T, MP, CEN>& op, bool use_lo,
                                             const T* lds, int w, int lane,
                                             typename MfmaOp<T>::v4 (&acc)[SlabShape<MP, CEN>::TPW]) {
  using S = SlabShape<MP, CEN>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int P = S::P, KS = S::KS, TPW = S::TPW;
  constexpr bool SPLIT = std::is_same<T, float>::value;
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    acc[u] = V4{T(0), T(0), T(0), T(0)};
    const int ct = S::col_tile(w, u);
    if (ct < 0) continue;
    const int col = ct * 16 + (lane & 15);
    T b[KS];
#pragma unroll
    for (int n = 0; n < KS; ++n) b[n] = lds[op.brow[n] * P + col];
    // the operator's low parts first, while the accumulator is small (added
    // last, a correction below half an ulp of the sum would be rounded away)
    if constexpr (SPLIT) {
      if (use_lo) {
#pragma unroll
        for (int c4 = 0; c4 < (KS + 3) / 4; ++c4) {
          typename MfmaOp<float>::s4 av;
          float fb[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int n = 4 * c4 + i;
            av[i] = n < KS ? op.alo[n] : (short)0;
            fb[i] = n < KS ? b[n] : 0.f;
          }
          acc[u] = MfmaOp<float>::mma_bf16(av, bf16x4(fb), acc[u]);
        }
      }
    }
#pragma unroll
    for (int n = 0; n < KS; ++n) acc[u] = MfmaOp<T>::mma(op.a[n], b[n], acc[u]);
  }
}

template <typename T>
struct SlabParams {
  PassChain<T> c[2];
  const T* mat_z;  // dense At (M x M), or At_c (M/2 x M) when centrosymmetric
  const T* mat_y;
  const unsigned short* lo_z;
  const unsigned short* lo_y;
  int X, Y, Z;
  int shift;       // centrosymmetric pairing (see cen_partner)
  int stream;      // 1: non-temporal field traffic (field_load)
};

template <typename T, int MP, bool CEN, bool NTM>
__device__ __forceinline__ void slab_zy_body(const SlabParams<T>& p, T* lds) {
  using S = SlabShape<MP, CEN>;
  constexpr int NT = S::NT, P = S::P, JT = S::JT, TPW = S::TPW;
  constexpr int H = MP / 2;
  using V4 = typename MfmaOp<T>::v4;

  const PassChain<T> ch = p.c[blockIdx.y];
  const int Y = p.Y, Z = p.Z;
  const long base = (long)blockIdx.x * Y * Z;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  constexpr bool SPLIT = std::is_same<T, float>::value;
  const bool shift = p.shift != 0;

  // ---- operands of both products, loaded with the plane ----
  const int jt = w % JT;  // every tile of this wave shares its row tile
  SlabOperands<T, MP, CEN> oz, oy;
  slab_load_operands<T, MP, CEN>(p.mat_z, p.lo_z, Z, jt, lane, oz);
  slab_load_operands<T, MP, CEN>(p.mat_y, p.lo_y, Y, jt, lane, oy);

  if constexpr (!CEN) {
    // ---- stage the plane transposed: lds[z * P + y], zero padding ----
    constexpr int KE = (MP * MP + NT - 1) / NT;
    T v[KE];
    int gy[KE], gz[KE];
    bool ok[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      gy[k] = f / MP;
      gz[k] = f - (f / MP) * MP;
      ok[k] = f < MP * MP && gy[k] < Y && gz[k] < Z;
      v[k] = ok[k] ? field_load<NTM>(&ch.in[base + (long)gy[k] * Z + gz[k]]) : T(0);
    }
    if (ch.pre) {
      T e[KE];
#pragma unroll
      for (int k = 0; k < KE; ++k) e[k] = ok[k] ? ch.pre[base + (long)gy[k] * Z + gz[k]] : T(0);
#pragma unroll
      for (int k = 0; k < KE; ++k) v[k] *= e[k];
    }
#pragma unroll
    for (int k = 0; k < KE; ++k)
      if (tid + NT * k < MP * MP) lds[gz[k] * P + gy[k]] = v[k];
  } else {
    // ---- stage z pairs: lds[z'][y] = u, lds[H + z'][y] = v (Y = Z = MP) ----
    constexpr int KE = MP * H / NT;
    long g1[KE], g2[KE];
    T x1[KE], x2[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int y = f / H, z = f - (f / H) * H;
      g1[k] = base + (long)y * MP + z;
      g2[k] = base + (long)y * MP + cen_partner<MP>(z, shift);
      x1[k] = field_load<NTM>(&ch.in[g1[k]]);
      x2[k] = field_load<NTM>(&ch.in[g2[k]]);
    }
    if (ch.pre) {
      T e1[KE], e2[KE];
#pragma unroll
      for (int k = 0; k < KE; ++k) {
        e1[k] = ch.pre[g1[k]];
        e2[k] = ch.pre[g2[k]];
      }
#pragma unroll
      for (int k = 0; k < KE; ++k) {
        x1[k] *= e1[k];
        x2[k] *= e2[k];
      }
    }
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int y = f / H, z = f - (f / H) * H;
      lds[z * P + y] = x1[k] + x2[k];
      lds[(H + z) * P + y] = x1[k] - x2[k];
    }
  }
  __syncthreads();

  const bool lo_z = SPLIT && p.lo_z != nullptr, lo_y = SPLIT && p.lo_y != nullptr;
  // ---- z product: rows j (Dz output, or U/V), columns y ----
  V4 acc[TPW];
  slab_product<T, MP, CEN>(oz, lo_z, lds, w, lane, acc);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int ct = S::col_tile(w, u);
    if (ct >= 0) {
      const int col = ct * 16 + (lane & 15);  // y
#pragma unroll
      for (int r = 0; r < 4; ++r) lds[col * P + jt * 16 + MfmaOp<T>::row(lane, r)] = acc[u][r];
    }
  }
  __syncthreads();

  if constexpr (CEN) {
    // lds[y][j] holds U (j < H) and V (j >= H) along z; rebuild the z values
    // and form the y pairs: lds[y'][z] = u, lds[H + y'][z] = v (z and the
    // row y as in memory, the partners at cen_partner).
    constexpr int KB = (H * H + NT - 1) / NT;  // 2 x 2 blocks per thread
    T nv[KB][4];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int f = tid + NT * k;
      if (f >= H * H) continue;
      const int y = f / H, z = f - (f / H) * H;
      const int yb = cen_partner<MP>(y, shift);
      const T U1 = lds[y * P + z], V1 = lds[y * P + H + z];
      const T U2 = lds[yb * P + z], V2 = lds[yb * P + H + z];
      const T t11 = U1 + V1, t12 = U1 - V1;  // (y, z), (y, zb)
      const T t21 = U2 + V2, t22 = U2 - V2;  // (yb, z), (yb, zb)
      nv[k][0] = t11 + t21;  // u at (y, z)
      nv[k][1] = t12 + t22;  // u at (y, zb)
      nv[k][2] = t11 - t21;  // v at (y, z)
      nv[k][3] = t12 - t22;  // v at (y, zb)
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int f = tid + NT * k;
      if (f >= H * H) continue;
      const int y = f / H, z = f - (f / H) * H, zb = cen_partner<MP>(z, shift);
      lds[y * P + z] = nv[k][0];
      lds[y * P + zb] = nv[k][1];
      lds[(H + y) * P + z] = nv[k][2];
      lds[(H + y) * P + zb] = nv[k][3];
    }
    __syncthreads();
  }

  // ---- y product: rows j' (Dy output, or U/V), columns z ----
  slab_product<T, MP, CEN>(oy, lo_y, lds, w, lane, acc);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int ct = S::col_tile(w, u);
    if (ct >= 0) {
      const int col = ct * 16 + (lane & 15);  // z
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lds_put<T>(lds, (jt * 16 + MfmaOp<T>::row(lane, r)) * P + col, acc[u][r]);
    }
  }
  __syncthreads();

  if constexpr (!CEN) {
    constexpr int KE = (MP * MP + NT - 1) / NT;
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int y = f / MP, z = f - (f / MP) * MP;
      if (f < MP * MP && y < Y && z < Z) {
        const long g = base + (long)y * Z + z;
        T v = lds[y * P + z];
        if (ch.post) v *= ch.post[g];
        field_store<NTM>(v, &ch.out[g]);
      }
    }
  } else {
    constexpr int KE = MP * H / NT;
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int y = f / MP, z = f - (f / MP) * MP;  // y < H
      const T U = lds[y * P + z], V = lds[(H + y) * P + z];
      const long g1 = base + (long)y * MP + z;
      const long g2 = base + (long)cen_partner<MP>(y, shift) * MP + z;
      T o1 = U + V, o2 = U - V;
      if (ch.post) {
        o1 *= ch.post[g1];
        o2 *= ch.post[g2];
      }
      field_store<NTM>(o1, &ch.out[g1]);
      field_store<NTM>(o2, &ch.out[g2]);
    }
  }
}

template <typename T, int MP, bool CEN, bool NTM>
__global__ __launch_bounds__(512) void slab_zy_kernel(SlabParams<T> p) {
  __shared__ T lds[MP * SlabShape<MP, CEN>::P];
  slab_zy_body<T, MP, CEN, NTM>(p, lds);
}

// Non-temporal field traffic in the slab when the caller streams fields (a
// propagation): measured 2-8% faster per step for FP64 and 5-7% for the FP32
// half-size products; dense FP32 products measured 11% slower with it (class
// 48) and keep ordinary accesses, as does a lone diffusion().
template <typename T>
inline bool slab_nt(const SlabParams<T>& p, bool cen) {
  return p.stream && (cen || sizeof(T) == 8);
}

bool slab_enabled();

// Whether a contour step fuses its z and y passes on x-slabs (Y and Z share a
// length class <= 64).
inline bool slab_applies(int Y, int Z) {
  const int cy = pass_class(Y), cz = pass_class(Z);
  return slab_enabled() && cy == cz && cy != 0 && cy <= 64;
}

// Launches the fused z + y pass when Y and Z share a class <= 64; returns
// hipErrorNotSupported otherwise (the caller then runs the two passes).
// `cen` selects the centrosymmetric kernel; the caller passes the half-size
// operators in that case and guarantees Y == Z == MP with MP % 32 == 0.
template <typename T, int MP, bool CEN>
void launch_slab_mp(const SlabParams<T>& p, dim3 grid, bool nt, hipStream_t s) {
  if constexpr (CEN || sizeof(T) == 8) {
    if (nt) {
      hipLaunchKernelGGL((slab_zy_kernel<T, MP, CEN, true>), grid, dim3(512), 0, s, p);
      return;
    }
  }
  hipLaunchKernelGGL((slab_zy_kernel<T, MP, CEN, false>), grid, dim3(512), 0, s, p);
}

template <typename T>
hipError_t launch_slab_zy(const SlabParams<T>& p, int nchains, bool cen, hipStream_t s) {
  const int cy = pass_class(p.Y), cz = pass_class(p.Z);
  if (!slab_enabled() || cy != cz || cy > 64 || cy == 0 || nchains < 1 || nchains > 2)
    return hipErrorNotSupported;
  dim3 grid((unsigned)p.X, (unsigned)nchains, 1);
  const bool nt = slab_nt(p, cen);
  if (cen) {
    if (p.Y != cy || p.Z != cz || cy % 32 != 0) return hipErrorInvalidValue;
    if (cy == 32) launch_slab_mp<T, 32, true>(p, grid, nt, s);
    else launch_slab_mp<T, 64, true>(p, grid, nt, s);
    return hipGetLastError();
  }
  switch (cy) {
    case 16: launch_slab_mp<T, 16, false>(p, grid, nt, s); break;
    case 32: launch_slab_mp<T, 32, false>(p, grid, nt, s); break;
    case 48: launch_slab_mp<T, 48, false>(p, grid, nt, s); break;
    default: launch_slab_mp<T, 64, false>(p, grid, nt, s); break;
  }
  return hipGetLastError();
}

template <typename T, int AXIS, int MP, bool CEN>
void launch_mfma_pass(const PassParams<T>& p, dim3 grid, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    if (!pass_nt(p)) {
      hipLaunchKernelGGL((axis_pass_mfma_kernel<T, AXIS, MP, CEN, false>), grid,
                         dim3(MfmaShape<MP>::NT), 0, s, p);
      return;
    }
  }
  hipLaunchKernelGGL((axis_pass_mfma_kernel<T, AXIS, MP, CEN, true>), grid,
                     dim3(MfmaShape<MP>::NT), 0, s, p);
}

template <typename T, int AXIS, int MP>
hipError_t launch_axis_pass_mp(const PassParams<T>& p, int nchains, hipStream_t s) {
  const int M = AXIS == 0 ? p.X : (AXIS == 1 ? p.Y : p.Z);
  const long nlines = (long)p.X * p.Y * p.Z / M;
  if constexpr (MP <= 128) {
    if (p.cen || pass_impl() != PassImpl::Valu) {
      using S = MfmaShape<MP>;
      const long nblk = (nlines + S::TL - 1) / S::TL;
      dim3 grid((unsigned)nblk, (unsigned)nchains, 1);
      if (p.cen) {
        if constexpr (MP % 32 == 0) {
          if (M != MP) return hipErrorInvalidValue;
          launch_mfma_pass<T, AXIS, MP, true>(p, grid, s);
          return hipGetLastError();
        }
        return hipErrorInvalidValue;
      }
      launch_mfma_pass<T, AXIS, MP, false>(p, grid, s);
      return hipGetLastError();
    }
  }
  if (p.cen) return hipErrorInvalidValue;
  using S = PassShape<MP>;
  const long nblk = (nlines + S::TL - 1) / S::TL;
  dim3 grid((unsigned)nblk, (unsigned)nchains, 1);
  hipLaunchKernelGGL((axis_pass_kernel<T, AXIS, MP>), grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

template <typename T, int AXIS>
hipError_t launch_axis_pass_ax(const PassParams<T>& p, int nchains, hipStream_t s) {
  const int M = AXIS == 0 ? p.X : (AXIS == 1 ? p.Y : p.Z);
  switch (pass_class(M)) {
    case 16: return launch_axis_pass_mp<T, AXIS, 16>(p, nchains, s);
    case 32: return launch_axis_pass_mp<T, AXIS, 32>(p, nchains, s);
    case 48: return launch_axis_pass_mp<T, AXIS, 48>(p, nchains, s);
    case 64: return launch_axis_pass_mp<T, AXIS, 64>(p, nchains, s);
    case 128: return launch_axis_pass_mp<T, AXIS, 128>(p, nchains, s);
    default: return hipErrorInvalidValue;
  }
}

// Launch one pass along `axis` for `nchains` (1 or 2) chains.
template <typename T>
hipError_t launch_axis_pass(int axis, const PassParams<T>& p, int nchains, hipStream_t s) {
  if (nchains < 1 || nchains > 2) return hipErrorInvalidValue;
  if (axis == 0) return launch_axis_pass_ax<T, 0>(p, nchains, s);
  if (axis == 1) return launch_axis_pass_ax<T, 1>(p, nchains, s);
  if (axis == 2) return launch_axis_pass_ax<T, 2>(p, nchains, s);
  return hipErrorInvalidValue;
}

}  // namespace crys
