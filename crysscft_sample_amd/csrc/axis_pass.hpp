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
};

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

// float -> bf16 bits, round to nearest even (finite inputs).
__device__ inline short f32_to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

template <int MP>
struct MfmaShape {
  static constexpr int TL = MP <= 32 ? 64 : 32;  // lines per workgroup
  static constexpr int LDT = TL + 16;            // LDS pitch: 4 B-rows hit distinct banks
  static constexpr int JT = MP / 16;             // 16-row output tiles
  static constexpr int TT = TL / 16;             // 16-line output tiles
  static constexpr int KS = MP / 4;              // MFMA k steps
  static constexpr int WPJ = JT >= 4 ? 1 : 4 / JT;  // waves sharing one row tile
  static constexpr int TPW = TT / WPJ;           // line tiles per wave
  static_assert(TT % WPJ == 0, "line tiles must split over the waves of a row tile");
};

// One workgroup's tile `tidx` of a pass on one chain; `tile` is LDS of at
// least MP * MfmaShape<MP>::LDT elements.  Shared by the single-field kernel
// and the batched kernel.
template <typename T, int AXIS, int MP>
__device__ __forceinline__ void mfma_pass_tile(const PassChain<T>& ch, const T* __restrict__ mat,
                                               const unsigned short* __restrict__ mat_lo, int X,
                                               int Y, int Z, int tidx, T* tile) {
  using S = MfmaShape<MP>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int NT = 256;
  constexpr int TL = S::TL, LDT = S::LDT, JT = S::JT, KS = S::KS, WPJ = S::WPJ, TPW = S::TPW;
  constexpr int KT = MP * TL / NT;
  static_assert(MP * TL % NT == 0, "tile must split evenly over the workgroup");

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

  // operands of this wave: rows jt*16.. of D, all k steps
  const int jt = w % JT, sub = w / JT;
  const bool active = w < JT * WPJ;
  const int jrow = jt * 16 + (lane & 15);
  T a[KS];
#pragma unroll
  for (int kk = 0; kk < KS; ++kk) {
    const int i = 4 * kk + (lane >> 4);
    a[kk] = (active && i < M && jrow < M) ? mat[i * M + jrow] : T(0);
  }
  // Split precision (FP32): D = D_hi + D_lo with D_lo in bf16.  The bf16
  // product of a 16-row chunk c pairs k' = 4(l/16)+i with row 16c+4i+l/16,
  // the row this lane already holds from f32 step kk = 4c+i, so it reuses the
  // same A addresses and the same B values.
  constexpr bool SPLIT = std::is_same<T, float>::value;
  short alo[SPLIT ? KS : 1];
  const bool use_lo = SPLIT && mat_lo != nullptr;
  if constexpr (SPLIT) {
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
      const int i = 4 * kk + (lane >> 4);
      alo[kk] = (use_lo && active && i < M && jrow < M) ? (short)mat_lo[i * M + jrow] : (short)0;
    }
  }
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
#pragma unroll
  for (int k = 0; k < KT; ++k) tile[lo[k]] = v[k];
  __syncthreads();

  V4 acc[TPW];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    acc[u] = V4{T(0), T(0), T(0), T(0)};
    if (active) {
      const int col = (sub * TPW + u) * 16 + (lane & 15);
#pragma unroll
      for (int c4 = 0; c4 < KS / 4; ++c4) {
        T b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int kk = 4 * c4 + i;
          b[i] = tile[(4 * kk + (lane >> 4)) * LDT + col];
          acc[u] = MfmaOp<T>::mma(a[kk], b[i], acc[u]);
        }
        if constexpr (SPLIT) {
          if (use_lo) {
            typename MfmaOp<float>::s4 av = {alo[4 * c4], alo[4 * c4 + 1], alo[4 * c4 + 2],
                                             alo[4 * c4 + 3]};
            typename MfmaOp<float>::s4 bv = {f32_to_bf16(b[0]), f32_to_bf16(b[1]),
                                             f32_to_bf16(b[2]), f32_to_bf16(b[3])};
            acc[u] = MfmaOp<float>::mma_bf16(av, bv, acc[u]);
          }
        }
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int col = (sub * TPW + u) * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) tile[(jt * 16 + MfmaOp<T>::row(lane, r)) * LDT + col] = acc[u][r];
    }
  }
  __syncthreads();

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

template <typename T, int AXIS, int MP>
__global__ __launch_bounds__(256) void axis_pass_mfma_kernel(PassParams<T> p) {
  __shared__ T tile[MP * MfmaShape<MP>::LDT];
  mfma_pass_tile<T, AXIS, MP>(p.c[blockIdx.y], p.mat, p.mat_lo, p.X, p.Y, p.Z, blockIdx.x, tile);
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
enum class PassImpl { Mfma, Valu };
PassImpl pass_impl();

// ---------------------------------------------------------------------------
// Fused z + y passes on x-slabs.  One workgroup (512 threads, 8 waves) owns
// the Y x Z plane of one x index: it stages the plane in LDS transposed
// (z-major), applies Dz along z and Dy along y with the matrix cores, and
// stores the plane back.  With the x pass that follows, a contour step takes
// two launches instead of three.  Requires Y and Z in the same length class
// MP <= 64.
// ---------------------------------------------------------------------------
template <int MP>
struct SlabShape {
  static constexpr int NT = 512;
  static constexpr int P = MP + 4;                 // LDS row pitch (elements)
  static constexpr int JT = MP / 16;               // 16-row output tiles
  static constexpr int CT = MP / 16;               // 16-column output tiles
  static constexpr int KS = MP / 4;                // MFMA k steps
  static constexpr int NTILE = JT * CT;            // output tiles per GEMM
  static constexpr int TPW = (NTILE + 7) / 8;      // tiles per wave
};

// One GEMM of the slab: out[c][j] = sum_k A[j][k] * src[k][c] for the 16 x 16
// tiles of wave w, with src stored as lds[k * P + c]; results are returned in
// registers (acc) for the caller to write once every wave has read src.
template <typename T, int MP>
__device__ __forceinline__ void slab_gemm(const T* __restrict__ mat,
                                          const unsigned short* __restrict__ mat_lo, int M,
                                          const T* lds, int w, int lane,
                                          typename MfmaOp<T>::v4 (&acc)[SlabShape<MP>::TPW]) {
  using S = SlabShape<MP>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int P = S::P, JT = S::JT, KS = S::KS, NTILE = S::NTILE, TPW = S::TPW;
  constexpr bool SPLIT = std::is_same<T, float>::value;
  const bool use_lo = SPLIT && mat_lo != nullptr;
  int cur_jt = -1;
  T a[KS];
  short alo[SPLIT ? KS : 1];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    acc[u] = V4{T(0), T(0), T(0), T(0)};
    const int q = w + 8 * u;  // tile index: jt = q % JT, ct = q / JT
    if (q >= NTILE) continue;
    const int jt = q % JT, ct = q / JT;
    if (jt != cur_jt) {
      cur_jt = jt;
      const int jrow = jt * 16 + (lane & 15);
#pragma unroll
      for (int kk = 0; kk < KS; ++kk) {
        const int i = 4 * kk + (lane >> 4);
        a[kk] = (i < M && jrow < M) ? mat[i * M + jrow] : T(0);
        if constexpr (SPLIT)
          alo[kk] = (use_lo && i < M && jrow < M) ? (short)mat_lo[i * M + jrow] : (short)0;
      }
    }
    const int col = ct * 16 + (lane & 15);
#pragma unroll
    for (int c4 = 0; c4 < KS / 4; ++c4) {
      T b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = 4 * c4 + i;
        b[i] = lds[(4 * kk + (lane >> 4)) * P + col];
        acc[u] = MfmaOp<T>::mma(a[kk], b[i], acc[u]);
      }
      if constexpr (SPLIT) {
        if (use_lo) {
          typename MfmaOp<float>::s4 av = {alo[4 * c4], alo[4 * c4 + 1], alo[4 * c4 + 2],
                                           alo[4 * c4 + 3]};
          typename MfmaOp<float>::s4 bv = {f32_to_bf16(b[0]), f32_to_bf16(b[1]),
                                           f32_to_bf16(b[2]), f32_to_bf16(b[3])};
          acc[u] = MfmaOp<float>::mma_bf16(av, bv, acc[u]);
        }
      }
    }
  }
}

template <typename T>
struct SlabParams {
  PassChain<T> c[2];
  const T* mat_z;
  const T* mat_y;
  const unsigned short* lo_z;
  const unsigned short* lo_y;
  int X, Y, Z;
};

template <typename T, int MP>
__global__ __launch_bounds__(512) void slab_zy_kernel(SlabParams<T> p) {
  using S = SlabShape<MP>;
  constexpr int NT = S::NT, P = S::P, JT = S::JT, TPW = S::TPW, NTILE = S::NTILE;
  constexpr int KE = (MP * MP + NT - 1) / NT;  // plane elements per thread
  __shared__ T lds[MP * P];
  using V4 = typename MfmaOp<T>::v4;

  const PassChain<T> ch = p.c[blockIdx.y];
  const int Y = p.Y, Z = p.Z;
  const long base = (long)blockIdx.x * Y * Z;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;

  // ---- stage the plane transposed: lds[z * P + y] (zero padding) ----
  T v[KE];
  int gy[KE], gz[KE];
  bool ok[KE];
#pragma unroll
  for (int k = 0; k < KE; ++k) {
    const int f = tid + NT * k;
    const int y = f / MP, z = f - (f / MP) * MP;
    gy[k] = y;
    gz[k] = z;
    ok[k] = f < MP * MP && y < Y && z < Z;
    v[k] = ok[k] ? ch.in[base + (long)y * Z + z] : T(0);
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
  __syncthreads();

  // ---- z pass: t[y][j] = sum_z Dz[j][z] * in[y][z], src lds[z][y] ----
  V4 acc[TPW];
  slab_gemm<T, MP>(p.mat_z, p.lo_z, Z, lds, w, lane, acc);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int q = w + 8 * u;
    if (q < NTILE) {
      const int jt = q % JT, ct = q / JT;
      const int col = ct * 16 + (lane & 15);  // y
#pragma unroll
      for (int r = 0; r < 4; ++r) lds[col * P + jt * 16 + MfmaOp<T>::row(lane, r)] = acc[u][r];
    }
  }
  __syncthreads();

  // ---- y pass: o[j][z] = sum_y Dy[j][y] * t[y][z], src lds[y][z] ----
  slab_gemm<T, MP>(p.mat_y, p.lo_y, Y, lds, w, lane, acc);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int q = w + 8 * u;
    if (q < NTILE) {
      const int jt = q % JT, ct = q / JT;
      const int col = ct * 16 + (lane & 15);  // z
#pragma unroll
      for (int r = 0; r < 4; ++r) lds[(jt * 16 + MfmaOp<T>::row(lane, r)) * P + col] = acc[u][r];
    }
  }
  __syncthreads();

  // ---- store the plane (row-major y, z), optional post multiply ----
#pragma unroll
  for (int k = 0; k < KE; ++k) v[k] = lds[gy[k] * P + gz[k]];
  if (ch.post) {
    T e[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) e[k] = ok[k] ? ch.post[base + (long)gy[k] * Z + gz[k]] : T(0);
#pragma unroll
    for (int k = 0; k < KE; ++k) v[k] *= e[k];
  }
#pragma unroll
  for (int k = 0; k < KE; ++k)
    if (ok[k]) ch.out[base + (long)gy[k] * Z + gz[k]] = v[k];
}

bool slab_enabled();

// Launches the fused z + y pass when Y and Z share a class <= 64; returns
// hipErrorNotSupported otherwise (the caller then runs the two passes).
template <typename T>
hipError_t launch_slab_zy(const SlabParams<T>& p, int nchains, hipStream_t s) {
  const int cy = pass_class(p.Y), cz = pass_class(p.Z);
  if (!slab_enabled() || cy != cz || cy > 64 || cy == 0 || nchains < 1 || nchains > 2)
    return hipErrorNotSupported;
  dim3 grid((unsigned)p.X, (unsigned)nchains, 1);
  switch (cy) {
    case 16: hipLaunchKernelGGL((slab_zy_kernel<T, 16>), grid, dim3(512), 0, s, p); break;
    case 32: hipLaunchKernelGGL((slab_zy_kernel<T, 32>), grid, dim3(512), 0, s, p); break;
    case 48: hipLaunchKernelGGL((slab_zy_kernel<T, 48>), grid, dim3(512), 0, s, p); break;
    default: hipLaunchKernelGGL((slab_zy_kernel<T, 64>), grid, dim3(512), 0, s, p); break;
  }
  return hipGetLastError();
}

template <typename T, int AXIS, int MP>
hipError_t launch_axis_pass_mp(const PassParams<T>& p, int nchains, hipStream_t s) {
  const int M = AXIS == 0 ? p.X : (AXIS == 1 ? p.Y : p.Z);
  const long nlines = (long)p.X * p.Y * p.Z / M;
  if constexpr (MP <= 64) {
    if (pass_impl() == PassImpl::Mfma) {
      using S = MfmaShape<MP>;
      const long nblk = (nlines + S::TL - 1) / S::TL;
      dim3 grid((unsigned)nblk, (unsigned)nchains, 1);
      hipLaunchKernelGGL((axis_pass_mfma_kernel<T, AXIS, MP>), grid, dim3(256), 0, s, p);
      return hipGetLastError();
    }
  }
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
