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
  const T* mat;  // At, M x M, At[i*M + j]
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
  constexpr int RJ = S::RJ, RT = S::RT, TL = S::TL, LDT = S::LDT;
  __shared__ T tile[MP * LDT];
  __shared__ T mat[S::MAT_LDS ? MP * MP : 1];

  const PassChain<T> ch = p.c[blockIdx.y];
  const int X = p.X, Y = p.Y, Z = p.Z;
  const int M = AXIS == 0 ? X : (AXIS == 1 ? Y : Z);
  const long YZ = (long)Y * Z;
  const long nlines = (long)X * YZ / M;
  const long l0 = (long)blockIdx.x * TL;
  const int nl = (int)((nlines - l0) < TL ? (nlines - l0) : TL);  // valid lines in this tile
  const int tid = threadIdx.x;

  if constexpr (S::MAT_LDS) {
    for (int f = tid; f < MP * MP; f += 256) {
      const int i = f / MP, j = f - i * MP;
      mat[f] = (i < M && j < M) ? p.mat[i * M + j] : T(0);
    }
  }

  // ---- stage the M x TL tile (zero for lines past the end) ----
  if constexpr (AXIS == 2) {
    const long base = l0 * Z;
    for (int f = tid; f < TL * M; f += 256) {
      const int ll = f / M, i = f - ll * M;
      T v = T(0);
      if (ll < nl) {
        v = ch.in[base + f];
        if (ch.pre) v *= ch.pre[base + f];
      }
      tile[i * LDT + ll] = v;
    }
  } else {
    for (int f = tid; f < TL * M; f += 256) {
      const int i = f / TL, ll = f - i * TL;
      T v = T(0);
      if (ll < nl) {
        const long l = l0 + ll;
        long g;
        if constexpr (AXIS == 1) g = (l / Z) * YZ + (long)i * Z + (l % Z);
        else g = (long)i * YZ + l;
        v = ch.in[g];
        if (ch.pre) v *= ch.pre[g];
      }
      tile[i * LDT + ll] = v;
    }
  }
  __syncthreads();

  // ---- out[j][t] = sum_i At[i][j] * tile[i][t] ----
  const int tx = tid & 15, ty = tid >> 4;
  T acc[RJ][RT];
#pragma unroll
  for (int r = 0; r < RJ; ++r)
#pragma unroll
    for (int c = 0; c < RT; ++c) acc[r][c] = T(0);

#pragma unroll 4
  for (int i = 0; i < M; ++i) {
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
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RJ; ++r)
#pragma unroll
    for (int c = 0; c < RT; ++c) tile[(ty * RJ + r) * LDT + tx * RT + c] = acc[r][c];
  __syncthreads();

  // ---- coalesced store, with the optional post multiply ----
  if constexpr (AXIS == 2) {
    const long base = l0 * Z;
    for (int f = tid; f < TL * M; f += 256) {
      const int ll = f / M, i = f - ll * M;
      if (ll < nl) {
        T v = tile[i * LDT + ll];
        if (ch.post) v *= ch.post[base + f];
        ch.out[base + f] = v;
      }
    }
  } else {
    for (int f = tid; f < TL * M; f += 256) {
      const int i = f / TL, ll = f - i * TL;
      if (ll < nl) {
        const long l = l0 + ll;
        long g;
        if constexpr (AXIS == 1) g = (l / Z) * YZ + (long)i * Z + (l % Z);
        else g = (long)i * YZ + l;
        T v = tile[i * LDT + ll];
        if (ch.post) v *= ch.post[g];
        ch.out[g] = v;
      }
    }
  }
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

template <typename T, int AXIS, int MP>
hipError_t launch_axis_pass_mp(const PassParams<T>& p, int nchains, hipStream_t s) {
  using S = PassShape<MP>;
  const int M = AXIS == 0 ? p.X : (AXIS == 1 ? p.Y : p.Z);
  const long nlines = (long)p.X * p.Y * p.Z / M;
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
