// Radix FFT / DCT axis pass on gfx950 (LDS-resident, two lines per complex
// transform).
//
// The same per-axis operator as the matrix passes (operators.hpp), applied
// with fast transforms instead of a dense product:
//
//  - P1 (solver_p1.h:174-202): out = IFFT_N( K'(k) FFT_N(in) ), K' = K / N
//    with k folded as k < (N+1)/2 ? k : N-k.  K' is real and even, so two
//    real lines a, b go through one complex FFT of z = a + i b and come back
//    as Re / Im of the inverse.
//  - PmmmDCT (solver_Pmmm.h:194-215): out = REDFT01( K' REDFT10(in) ),
//    K' = K / (2M), computed with M-point complex FFTs (Makhoul): the line x
//    is reordered to v_j = x_{2j}, v_{M-1-j} = x_{2j+1} (j < M/2), V =
//    FFT_M(v) gives REDFT10 as Y_k = 2 Re(c_k V_k), c_k = exp(-i pi k / 2M),
//    and REDFT01 of Y' = K' Y is the reordered real part of the inverse FFT
//    of W_k = conj(c_k) (Y'_k - i Y'_{M-k}).  With P = 2 c_k V_k =
//    Y_k - i Y_{M-k} one step gives W_k = conj(c_k) (K'_k Re P + i K'_{M-k} Im P)
//    from the pair (V_k, V_{M-k}); for two packed lines V_a, V_b are split
//    from Z = FFT(v_a + i v_b) by Hermitian symmetry.
//  - 3m with zero translational parts (solver_3m.h:281-324): the even
//    sub-lattice e of a mirror-symmetric line (f(N-1-j) = f(j), N = 2M) is
//    exactly the Makhoul order of the first half of f (v_j = f_{2j} = e_j,
//    v_{M-1-j} = f_{2j+1} = f_{N-2-2j} = e_{M-1-j}), so the 3m operator is
//    the PmmmDCT one without the reordering, with the same K'_k, k < M (the
//    Nyquist mode of such a line vanishes).
//
// Each workgroup owns TL lines of one chain.  The tile is loaded coalesced
// (with the optional pre multiply by E) straight into the transform layout:
// packed line c holds line 2c in its real plane and line 2c+1 in its
// imaginary plane, each point at its transform position (the Makhoul order
// for PmmmDCT).  Every stage then works in place on that one LDS buffer:
//
//   forward  A: job (c, t), t < R2: DFT_R1 over the points t + R2 m, times
//               w^(t k1), back to positions t + R2 k1;
//   forward  B: residue s < R1: DFT_R2 over the positions s R2 + t; the
//               output Z_(s + R1 k2) stays at position s R2 + k2;
//   spectral:   the pairs (k, M-k);
//   inverse  A: residue s: DFT_R2 over the positions s R2 + k2, times
//               w^(j2 s), back to positions s R2 + j2;
//   inverse  B: job (c, j2), j2 < R2: DFT_R1 over the positions s R2 + j2;
//               output j1 goes to position j1 R2 + j2 = j, the natural order.
//
// Forward B, the spectral step and inverse A can run as one stage in
// registers: the partner of Z_k (residue s = k mod R1) is Z_(M-k), residue
// R1 - s, so a job that holds residues s and R1 - s (0 and R1 / 2 together)
// has every pair it needs.  The slab and phase kernels below always do this;
// the pass kernel does it for FP32 at length 128 only, because elsewhere its
// R1 / 2 jobs per line leave too many threads idle (measured slower).
//
// Position p sits at p + p / R2 in its plane, so the R1 or R2 lanes of one
// job group and the lanes of neighbouring jobs fall on distinct banks in
// every stage.  The store reads the tile back from the transform layout
// (with the optional post multiply).  Tiles are disjoint, so a pass may run
// in place (out == in).
#pragma once

#include "axis_pass.hpp"
#include "plane_phase.hpp"

namespace crys {

enum RadixKind : int { kRadixP1 = 0, kRadixDct = 1, kRadixM3 = 2 };

// Per-axis table (element type T) of a transform of length M:
//   [0, 2M)   cos(2 pi j / M), sin(2 pi j / M) for j < M (interleaved)
//   [2M, 4M)  cos(pi k / 2M), sin(pi k / 2M) for k < M (DCT kinds; P1: unused)
//   [4M, 5M)  K'_k for k < M
template <int M>
struct RadixShape {
  static constexpr int R1 = M <= 16 ? 4 : (M <= 64 ? 8 : 16);  // points per forward-A job
  static constexpr int R2 = M / R1;                             // points per forward-B job
  // lines per workgroup: 1024 points per tile (small tiles, so that a 64^3
  // pass has 512 workgroups); 4096 at M = 128, whose x / y tiles then load
  // rows of 32 contiguous values
  static constexpr int TL = M == 128 ? 32 : 1024 / M;
  static constexpr int NC = TL / 2;                             // packed complex lines
  static constexpr int NT = NC * (R1 > R2 ? R1 : R2);           // threads
  static constexpr int PP = M + R1 + 1;                         // plane pitch (odd)
  static constexpr int TAB = 5 * M;
  static constexpr int TAB_LDS = 3 * M;  // FFT twiddles and K' (the c_k stay in global memory)
  static_assert(M >= 8 && (M & (M - 1)) == 0 && M <= 128, "power-of-two length 8..128");
  static_assert(R1 <= 16 && R2 <= 16 && R2 >= 2, "register radices");
  static __device__ __forceinline__ int phys(int p) { return p + p / R2; }
};

template <typename T>
struct RadixParams {
  PassChain<T> c[2];
  const T* tab;
  int X, Y, Z;
};

// cos(2 pi p / 16) and sin(2 pi p / 16)
__host__ __device__ constexpr double cos16(int p) {
  constexpr double a = 0.92387953251128675613, b = 0.70710678118654752440,
                   c = 0.38268343236508977173;
  p &= 15;
  return p == 0 ? 1.0 : p == 1 ? a : p == 2 ? b : p == 3 ? c : p == 4 ? 0.0 : p == 5 ? -c
       : p == 6 ? -b : p == 7 ? -a : p == 8 ? -1.0 : p == 9 ? -a : p == 10 ? -b
       : p == 11 ? -c : p == 12 ? 0.0 : p == 13 ? c : p == 14 ? b : a;
}
__host__ __device__ constexpr double sin16(int p) { return cos16(p - 4); }

__host__ __device__ constexpr int bitrev(int i, int R) {
  int r = 0;
  for (int b = 1; b < R; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// In-register DFT of R <= 16 points: X_k = sum_m x_m exp(SIGN 2 pi i m k / R)
// (radix-2, decimation in time; every twiddle is a compile-time constant).
template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft_reg(T (&re)[R], T (&im)[R]) {
  T xr[R], xi[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    xr[bitrev(i, R)] = re[i];
    xi[bitrev(i, R)] = im[i];
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int s0 = 0; s0 < R; s0 += len) {
#pragma unroll
      for (int j = 0; j < len / 2; ++j) {
        const int p = j * (16 / len);  // twiddle angle in units of 2 pi / 16
        const int a = s0 + j, b = a + len / 2;
        T br, bi;
        if (p == 0) {
          br = xr[b];
          bi = xi[b];
        } else if (p == 4) {  // exp(SIGN i pi / 2) = SIGN i
          br = SIGN > 0 ? -xi[b] : xi[b];
          bi = SIGN > 0 ? xr[b] : -xr[b];
        } else {
          const T wc = (T)cos16(p), ws = (T)(SIGN * sin16(p));
          br = xr[b] * wc - xi[b] * ws;
          bi = xr[b] * ws + xi[b] * wc;
        }
        xr[b] = xr[a] - br;
        xi[b] = xi[a] - bi;
        xr[a] = xr[a] + br;
        xi[a] = xi[a] + bi;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    re[i] = xr[i];
    im[i] = xi[i];
  }
}

// One in-place stage over every packed line: job (c, g), g < G, loads the R
// points at positions BASE_MUL g + STEP r (r < R) of both planes, applies
// DFT_R with sign SIGN, multiplies output r by w_M^(SIGN g r) when `twiddle`
// is set, and stores output r back to position BASE_MUL g + STEP r.
template <typename T, int M, int R, int G, int SIGN, int STEP, int BASE_MUL>
__device__ __forceinline__ void radix_stage(T* lds, const T* tw, bool twiddle) {
  using S = RadixShape<M>;
  for (int job = threadIdx.x; job < S::NC * G; job += S::NT) {
    const int c = job / G, g = job - c * G;
    T* const pr = lds + c * S::PP;
    T* const pi = lds + (S::NC + c) * S::PP;
    T vr[R], vi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = S::phys(BASE_MUL * g + STEP * r);
      vr[r] = pr[q];
      vi[r] = pi[q];
    }
    dft_reg<T, R, SIGN>(vr, vi);
    if (twiddle) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int e = g * r;  // < M
        const T wc = tw[2 * e], ws = SIGN > 0 ? tw[2 * e + 1] : -tw[2 * e + 1];
        const T x = vr[r] * wc - vi[r] * ws;
        vi[r] = vr[r] * ws + vi[r] * wc;
        vr[r] = x;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = S::phys(BASE_MUL * g + STEP * r);
      pr[q] = vr[r];
      pi[q] = vi[r];
    }
  }
}

// Transform position of reduced index i (PmmmDCT: the Makhoul order
// v_p = x_{2p}, v_{M-1-p} = x_{2p+1}); identity for P1 and 3m.
template <int KIND, int M>
__device__ __forceinline__ int radix_pos(int i) {
  if constexpr (KIND == kRadixDct) return (i & 1) ? M - 1 - (i >> 1) : (i >> 1);
  else return i;
}

// The spectral pair (Z_k, Z_{M-k}) -> (W_k, W_{M-k}) of the DCT kinds, in
// place (see the pass kernel's spectral step).
template <typename T, int M>
__device__ __forceinline__ void dct_spec_pair(T& zr, T& zi, T& mr, T& mi, int k, const T* kp,
                                              const T* cq) {
  const T yr = mr, yi = -mi;                    // conj(Z_{M-k})
  const T cr = cq[2 * k], ci = -cq[2 * k + 1];  // c_k = exp(-i pi k / 2M)
  const T K0 = kp[k], K1 = kp[(M - k) & (M - 1)];
  const T sr = zr + yr, si = zi + yi, dr = zr - yr, di = zi - yi;
  const T par = cr * sr - ci * si, pai = cr * si + ci * sr;
  const T qr = cr * dr - ci * di, qi = cr * di + ci * dr;
  const T pbr = qi, pbi = -qr;
  const T ar = K0 * par, ai = K1 * pai, br = K0 * pbr, bi = K1 * pbi;
  const T war = cr * ar + ci * ai, wai = cr * ai - ci * ar;
  const T wbr = cr * br + ci * bi, wbi = cr * bi - ci * br;
  zr = war - wbi;
  zi = wai + wbr;
  mr = war + wbi;
  mi = wbr - wai;
}

// Forward stage B, the spectral step and inverse stage A' of the pass kernel
// in one pass over the registers (see slab_fused_bsa below for the pairing of
// residues s and R1 - s): job (c, u), u < R1 / 2, on the planes of packed
// line c.
template <typename T, int M, int KIND>
__device__ __forceinline__ void radix_fused_bsa(T* buf, const T* tw, const T* kp, const T* cq) {
  using S = RadixShape<M>;
  constexpr int R1 = S::R1, R2 = S::R2, NC = S::NC, PP = S::PP, U = R1 / 2;
  for (int job = threadIdx.x; job < NC * U; job += S::NT) {
    const int c = job / U, u = job - c * U;
    const int sa = u, sb = u == 0 ? R1 / 2 : R1 - u;
    T* const pr = buf + c * PP;
    T* const pi = buf + (NC + c) * PP;
    T ar[R2], ai[R2], br[R2], bi[R2];
#pragma unroll
    for (int t = 0; t < R2; ++t) {
      ar[t] = pr[S::phys(sa * R2 + t)];
      ai[t] = pi[S::phys(sa * R2 + t)];
      br[t] = pr[S::phys(sb * R2 + t)];
      bi[t] = pi[S::phys(sb * R2 + t)];
    }
    dft_reg<T, R2, -1>(ar, ai);  // Z_(sa + R1 k2)
    dft_reg<T, R2, -1>(br, bi);  // Z_(sb + R1 k2)
    if constexpr (KIND == kRadixP1) {
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        const T ka = kp[sa + R1 * k2], kb = kp[sb + R1 * k2];
        ar[k2] *= ka;
        ai[k2] *= ka;
        br[k2] *= kb;
        bi[k2] *= kb;
      }
    } else if (u == 0) {
#pragma unroll
      for (int k2 = 0; k2 <= R2 / 2; k2 += R2 / 2) {
        T mr = ar[k2], mi = ai[k2];
        dct_spec_pair<T, M>(ar[k2], ai[k2], mr, mi, R1 * k2, kp, cq);
      }
#pragma unroll
      for (int k2 = 1; k2 < R2 / 2; ++k2)
        dct_spec_pair<T, M>(ar[k2], ai[k2], ar[R2 - k2], ai[R2 - k2], R1 * k2, kp, cq);
#pragma unroll
      for (int k2 = 0; k2 < R2 / 2; ++k2)
        dct_spec_pair<T, M>(br[k2], bi[k2], br[R2 - 1 - k2], bi[R2 - 1 - k2], R1 / 2 + R1 * k2, kp,
                            cq);
    } else {
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2)
        dct_spec_pair<T, M>(ar[k2], ai[k2], br[R2 - 1 - k2], bi[R2 - 1 - k2], sa + R1 * k2, kp,
                            cq);
    }
    dft_reg<T, R2, 1>(ar, ai);
    dft_reg<T, R2, 1>(br, bi);
#pragma unroll
    for (int j2 = 1; j2 < R2; ++j2) {
      const int ea = j2 * sa, eb = j2 * sb;  // < M
      const T ca = tw[2 * ea], sna = tw[2 * ea + 1], cb = tw[2 * eb], snb = tw[2 * eb + 1];
      const T xa = ar[j2] * ca - ai[j2] * sna;
      ai[j2] = ar[j2] * sna + ai[j2] * ca;
      ar[j2] = xa;
      const T xb = br[j2] * cb - bi[j2] * snb;
      bi[j2] = br[j2] * snb + bi[j2] * cb;
      br[j2] = xb;
    }
#pragma unroll
    for (int t = 0; t < R2; ++t) {
      pr[S::phys(sa * R2 + t)] = ar[t];
      pi[S::phys(sa * R2 + t)] = ai[t];
      pr[S::phys(sb * R2 + t)] = br[t];
      pi[S::phys(sb * R2 + t)] = bi[t];
    }
  }
}

// Global field traffic of the radix passes and the radix slab: each value is
// read and written once per launch.  FP32 uses non-temporal loads and stores,
// so that E and the tables stay in L2 (measured 7-11% faster per FP32 step at
// 128^3); FP64 keeps ordinary accesses (non-temporal ones measured 1-2.5%
// slower on three of four workloads).
template <typename T>
__device__ __forceinline__ T radix_line_load(const T* p) {
  if constexpr (sizeof(T) == 4) return __builtin_nontemporal_load(p);
  else return *p;
}
template <typename T>
__device__ __forceinline__ void radix_line_store(T v, T* p) {
  if constexpr (sizeof(T) == 4) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Waves per SIMD the compiler must allow: 4 (four 256-thread workgroups per
// CU) where that fits in 128 VGPRs without spilling; the FP64 length-128
// kernel needs about 150 and measured slower when held to 128 (spills).
template <typename T, int M>
constexpr int radix_waves_per_eu() {
  return (sizeof(T) == 8 && M == 128) ? 1 : 4;
}

template <typename T, int AXIS, int M, int KIND>
__global__ __launch_bounds__(RadixShape<M>::NT)
__attribute__((amdgpu_waves_per_eu(radix_waves_per_eu<T, M>()))) void radix_pass_kernel(
    RadixParams<T> p) {
  using S = RadixShape<M>;
  constexpr int NT = S::NT, TL = S::TL, NC = S::NC, PP = S::PP, R1 = S::R1, R2 = S::R2;
  constexpr int KT = (M * TL + NT - 1) / NT;  // tile elements per thread
  __shared__ T buf[2 * NC * PP];              // planes re (c < NC) | im
  __shared__ T tab[S::TAB_LDS];  // FFT twiddles (2M), then K' (M)

  const PassChain<T> ch = p.c[blockIdx.y];
  const int X = p.X, Y = p.Y, Z = p.Z;
  const int YZ = Y * Z;
  const int nlines = X * YZ / M;
  const int l0 = blockIdx.x * TL;
  const int nl = nlines - l0 < TL ? nlines - l0 : TL;
  const int tid = threadIdx.x;
  const bool yplane = AXIS == 1 && Z % TL == 0;
  const int x0 = yplane ? l0 / Z : 0;
  const int z0 = yplane ? l0 - x0 * Z : 0;

  // ---- element (row i, line ll) of slot k: global index, LDS index, and
  // whether it exists; recomputed for the store instead of held in registers
  auto slot = [&](int k, int& g, int& l) -> bool {
    const int f = tid + NT * k;
    int i, ll;
    if constexpr (AXIS == 2) {
      ll = f / M;
      i = f - ll * M;
    } else {
      i = f / TL;
      ll = f - i * TL;
    }
    l = ((ll & 1) * NC + (ll >> 1)) * PP + S::phys(radix_pos<KIND, M>(i));
    if constexpr (AXIS == 2) {
      g = (l0 + ll) * Z + i;
    } else if constexpr (AXIS == 1) {
      if (yplane) {
        g = x0 * YZ + i * Z + z0 + ll;
      } else {
        const int l1 = l0 + ll, x = l1 / Z;
        g = x * YZ + i * Z + (l1 - x * Z);
      }
    } else {
      g = i * YZ + l0 + ll;
    }
    return f < M * TL && ll < nl;
  };

  // ---- global loads first (tile, E, table), then LDS ----
  {
    int gi[KT], lo[KT];
    bool ok[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      ok[k] = slot(k, gi[k], lo[k]);
      if (!ok[k]) gi[k] = 0;
    }
    T v[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] = ok[k] ? radix_line_load(&ch.in[gi[k]]) : T(0);
    if (ch.pre) {
      T e[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) e[k] = ok[k] ? ch.pre[gi[k]] : T(0);
#pragma unroll
      for (int k = 0; k < KT; ++k) v[k] *= e[k];
    }
    for (int f = tid; f < S::TAB_LDS; f += NT) tab[f] = p.tab[f < 2 * M ? f : f + 2 * M];
#pragma unroll
    for (int k = 0; k < KT; ++k)
      if (tid + NT * k < M * TL) buf[lo[k]] = v[k];
  }
  __syncthreads();

  const T* const tw = tab;
  const T* const cq = p.tab + 2 * M;
  const T* const kp = tab + 2 * M;

  // ---- forward FFT of z = line 2c + i line 2c+1 ----
  radix_stage<T, M, R1, R2, -1, R2, 1>(buf, tw, true);   // A: points t + R2 m
  __syncthreads();
  if constexpr (sizeof(T) == 4 && M == 128) {
    // B, spectral and A' in registers: measured faster only here (FP32 at
    // length 128); elsewhere its R1 / 2 jobs per line leave too many threads
    // idle
    radix_fused_bsa<T, M, KIND>(buf, tw, kp, cq);
    __syncthreads();
  } else {
    radix_stage<T, M, R2, R1, -1, 1, R2>(buf, tw, false);  // B: points s R2 + t
    __syncthreads();

    // ---- spectral step; Z_k sits at position (k % R1) R2 + k / R1 ----
    auto zpos = [](int k) { return S::phys((k % R1) * R2 + k / R1); };
    if constexpr (KIND == kRadixP1) {
      for (int f = tid; f < NC * M; f += NT) {
        const int c = f / M, k = f - c * M;
        const int q = zpos(k);
        buf[c * PP + q] *= kp[k];
        buf[(NC + c) * PP + q] *= kp[k];
      }
    } else {
      constexpr int HP = M / 2 + 1;  // pairs (k, M-k), k = 0 .. M/2
      for (int f = tid; f < NC * HP; f += NT) {
        const int c = f / HP, k = f - c * HP;
        const int km = (M - k) & (M - 1);
        T* const pr = buf + c * PP;
        T* const pi = buf + (NC + c) * PP;
        const int q0 = zpos(k), q1 = zpos(km);
        const T zr = pr[q0], zi = pi[q0];
        const T yr = pr[q1], yi = -pi[q1];           // conj(Z_{M-k})
        const T cr = cq[2 * k], ci = -cq[2 * k + 1];  // c_k = exp(-i pi k / 2M)
        const T K0 = kp[k], K1 = kp[km];
        // line a: P = c_k (Z_k + conj Z_{M-k}); line b: P = -i c_k (Z_k - conj Z_{M-k})
        const T sr = zr + yr, si = zi + yi, dr = zr - yr, di = zi - yi;
        const T par = cr * sr - ci * si, pai = cr * si + ci * sr;
        const T qr = cr * dr - ci * di, qi = cr * di + ci * dr;
        const T pbr = qi, pbi = -qr;
        // W = conj(c_k) (K'_k Re P + i K'_{M-k} Im P)
        const T ar = K0 * par, ai = K1 * pai, br = K0 * pbr, bi = K1 * pbi;
        const T war = cr * ar + ci * ai, wai = cr * ai - ci * ar;
        const T wbr = cr * br + ci * bi, wbi = cr * bi - ci * br;
        // packed: W_k = Wa + i Wb, W_{M-k} = conj(Wa) + i conj(Wb)
        pr[q0] = war - wbi;
        pi[q0] = wai + wbr;
        if (km != k) {
          pr[q1] = war + wbi;
          pi[q1] = wbr - wai;
        }
      }
    }
    __syncthreads();

    // ---- inverse FFT back to the natural order; Re / Im are the two lines ----
    radix_stage<T, M, R2, R1, 1, 1, R2>(buf, tw, true);    // A': positions s R2 + k2
    __syncthreads();
  }
  radix_stage<T, M, R1, R2, 1, R2, 1>(buf, tw, false);   // B': positions s R2 + j2
  __syncthreads();

  // ---- coalesced store, with the optional post multiply ----
  int gi[KT], lo[KT];
  bool ok[KT];
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    ok[k] = slot(k, gi[k], lo[k]);
    if (!ok[k]) gi[k] = 0;
  }
  T v[KT];
#pragma unroll
  for (int k = 0; k < KT; ++k) v[k] = (tid + NT * k < M * TL) ? buf[lo[k]] : T(0);
  if (ch.post) {
    T e[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) e[k] = ok[k] ? ch.post[gi[k]] : T(0);
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] *= e[k];
  }
#pragma unroll
  for (int k = 0; k < KT; ++k)
    if (ok[k]) radix_line_store(v[k], &ch.out[gi[k]]);
}

// ---------------------------------------------------------------------------
// Fused z + y radix slab at length 128: one workgroup holds the whole x-plane
// (Y = Z = 128) in LDS as a real array pl[row][col], row = transform position
// of y, column = padded transform position of z.  The z transforms pack two
// rows as the real and imaginary parts of one complex line, the y transforms
// two columns; every stage works in place on that one array, with the stage
// order and position bookkeeping of the pass kernel above.  A contour step is
// then this kernel (in -> scr, pre E) and one radix x pass (scr -> out, post
// E): 6 n b per chain instead of the 8 n b of three passes.
// ---------------------------------------------------------------------------
template <typename T>
struct RadixSlabParams {
  PassChain<T> c[2];
  const T* tab_z;  // radix_table of the z axis
  const T* tab_y;  // radix_table of the y axis
  int X, Y, Z;
};

template <typename T>
struct RadixSlabShape {
  static constexpr int M = 128, R1 = 16, R2 = 8;
  // one workgroup per CU (the plane fills the LDS): 1024 threads for FP32,
  // 512 for FP64, whose 16-point stages need more than 128 VGPRs
  static constexpr int NT = sizeof(T) == 8 ? 512 : 1024;
  static constexpr int NCL = M / 2;           // packed lines per direction
  static constexpr int RP = 148;              // row pitch: >= M + M / R2, = 4 mod 16
  static_assert(RP >= M + M / R2 && RP % 16 == 4, "row pitch");
  static __device__ __forceinline__ int col(int q) { return q + q / R2; }
};

// In-place stage over the packed lines of one direction (see radix_stage):
// job (c, g), g < G; DFT_R over the positions BASE_MUL g + STEP r; twiddle
// w_M^(SIGN g r) when `twiddle`.  Z_DIR: line c = rows 2c, 2c+1 (positions
// along the row, jobs of one line on neighbouring lanes); else line c =
// columns 2c, 2c+1 (positions down the column, lines on neighbouring lanes).
template <typename T, int NT, bool Z_DIR, int R, int G, int SIGN, int STEP, int BASE_MUL>
__device__ __forceinline__ void slab_stage(T* pl, const T* tw, bool twiddle) {
  using S = RadixSlabShape<T>;
  constexpr int M = S::M, NCL = S::NCL, RP = S::RP;
  for (int job = threadIdx.x; job < NCL * G; job += NT) {
    const int c = Z_DIR ? job / G : job % NCL;
    const int g = Z_DIR ? job - c * G : job / NCL;
    auto at = [&](int pos, int part) {
      return Z_DIR ? (2 * c + part) * RP + S::col(pos) : pos * RP + S::col(2 * c + part);
    };
    T vr[R], vi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = BASE_MUL * g + STEP * r;
      vr[r] = pl[at(q, 0)];
      vi[r] = pl[at(q, 1)];
    }
    dft_reg<T, R, SIGN>(vr, vi);
    if (twiddle) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int e = g * r;  // < M
        const T wc = tw[2 * e], ws = SIGN > 0 ? tw[2 * e + 1] : -tw[2 * e + 1];
        const T x = vr[r] * wc - vi[r] * ws;
        vi[r] = vr[r] * ws + vi[r] * wc;
        vr[r] = x;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = BASE_MUL * g + STEP * r;
      pl[at(q, 0)] = vr[r];
      pl[at(q, 1)] = vi[r];
    }
  }
  (void)M;
}

// Forward stage B, the spectral step and inverse stage A' in one pass over
// the registers (slab and phase kernels, M = 128): forward B of residue s
// leaves Z_(s + R1 k2) at the positions s R2 + k2 that inverse A' of residue
// s reads back, and the spectral partner of residue s is residue R1 - s.  Job
// (c, u), u < R1 / 2, holds residues {u, R1 - u} ({0, R1 / 2} for u = 0), so
// every pair is in one job's registers and the two LDS round trips and
// barriers between the three stages go away.
template <typename T, int NT, bool Z_DIR, int KIND>
__device__ __forceinline__ void slab_fused_bsa(T* pl, const T* tw, const T* kp, const T* cq) {
  using S = RadixSlabShape<T>;
  constexpr int M = S::M, R1 = S::R1, R2 = S::R2, NCL = S::NCL, RP = S::RP, U = R1 / 2;
  for (int job = threadIdx.x; job < NCL * U; job += NT) {
    const int c = Z_DIR ? job / U : job % NCL;
    const int u = Z_DIR ? job - c * U : job / NCL;
    const int sa = u, sb = u == 0 ? R1 / 2 : R1 - u;
    auto at = [&](int pos, int part) {
      return Z_DIR ? (2 * c + part) * RP + S::col(pos) : pos * RP + S::col(2 * c + part);
    };
    T ar[R2], ai[R2], br[R2], bi[R2];
#pragma unroll
    for (int t = 0; t < R2; ++t) {
      ar[t] = pl[at(sa * R2 + t, 0)];
      ai[t] = pl[at(sa * R2 + t, 1)];
      br[t] = pl[at(sb * R2 + t, 0)];
      bi[t] = pl[at(sb * R2 + t, 1)];
    }
    dft_reg<T, R2, -1>(ar, ai);  // Z_(sa + R1 k2)
    dft_reg<T, R2, -1>(br, bi);  // Z_(sb + R1 k2)
    if constexpr (KIND == kRadixP1) {
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        const T ka = kp[sa + R1 * k2], kb = kp[sb + R1 * k2];
        ar[k2] *= ka;
        ai[k2] *= ka;
        br[k2] *= kb;
        bi[k2] *= kb;
      }
    } else if (u == 0) {
      // residue 0: k = 0 and M / 2 pair with themselves, k2 with R2 - k2
#pragma unroll
      for (int k2 = 0; k2 <= R2 / 2; k2 += R2 / 2) {
        T mr = ar[k2], mi = ai[k2];
        dct_spec_pair<T, M>(ar[k2], ai[k2], mr, mi, R1 * k2, kp, cq);
      }
#pragma unroll
      for (int k2 = 1; k2 < R2 / 2; ++k2)
        dct_spec_pair<T, M>(ar[k2], ai[k2], ar[R2 - k2], ai[R2 - k2], R1 * k2, kp, cq);
      // residue R1 / 2: k2 with R2 - 1 - k2
#pragma unroll
      for (int k2 = 0; k2 < R2 / 2; ++k2)
        dct_spec_pair<T, M>(br[k2], bi[k2], br[R2 - 1 - k2], bi[R2 - 1 - k2], R1 / 2 + R1 * k2, kp,
                            cq);
    } else {
      // residue sa with residue sb = R1 - sa: k2 with R2 - 1 - k2
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2)
        dct_spec_pair<T, M>(ar[k2], ai[k2], br[R2 - 1 - k2], bi[R2 - 1 - k2], sa + R1 * k2, kp,
                            cq);
    }
    // inverse A': DFT over k2, twiddle w^(j2 s), back to positions s R2 + j2
    dft_reg<T, R2, 1>(ar, ai);
    dft_reg<T, R2, 1>(br, bi);
#pragma unroll
    for (int j2 = 1; j2 < R2; ++j2) {
      const int ea = j2 * sa, eb = j2 * sb;  // < M
      const T ca = tw[2 * ea], sna = tw[2 * ea + 1], cb = tw[2 * eb], snb = tw[2 * eb + 1];
      const T xa = ar[j2] * ca - ai[j2] * sna;
      ai[j2] = ar[j2] * sna + ai[j2] * ca;
      ar[j2] = xa;
      const T xb = br[j2] * cb - bi[j2] * snb;
      bi[j2] = br[j2] * snb + bi[j2] * cb;
      br[j2] = xb;
    }
#pragma unroll
    for (int t = 0; t < R2; ++t) {
      pl[at(sa * R2 + t, 0)] = ar[t];
      pl[at(sa * R2 + t, 1)] = ai[t];
      pl[at(sb * R2 + t, 0)] = br[t];
      pl[at(sb * R2 + t, 1)] = bi[t];
    }
  }
}

// Forward FFT, spectral step (K' of this direction; cq the DCT twiddles in
// global memory) and inverse FFT of every packed line of one direction.
// NT: the threads of the calling workgroup; `before_last()` runs before the
// last stage (a caller's loads that should be in flight during it).
template <typename T, int NT, bool Z_DIR, int KIND, typename F>
__device__ __forceinline__ void slab_transform(T* pl, const T* tw, const T* kp, const T* cq,
                                               F before_last) {
  using S = RadixSlabShape<T>;
  constexpr int R1 = S::R1, R2 = S::R2;
  slab_stage<T, NT, Z_DIR, R1, R2, -1, R2, 1>(pl, tw, true);   // A: points t + R2 m
  __syncthreads();
  slab_fused_bsa<T, NT, Z_DIR, KIND>(pl, tw, kp, cq);          // B, spectral, A'
  __syncthreads();
  before_last();
  slab_stage<T, NT, Z_DIR, R1, R2, 1, R2, 1>(pl, tw, false);   // B': positions s R2 + j2
  __syncthreads();
}

template <typename T, int NT, bool Z_DIR, int KIND>
__device__ __forceinline__ void slab_transform(T* pl, const T* tw, const T* kp, const T* cq) {
  slab_transform<T, NT, Z_DIR, KIND>(pl, tw, kp, cq, [] {});
}

template <typename T, int KIND>
__global__ __launch_bounds__(RadixSlabShape<T>::NT) void radix_slab_kernel(RadixSlabParams<T> p) {
  using S = RadixSlabShape<T>;
  constexpr int M = S::M, NT = S::NT, RP = S::RP;
  constexpr int KT = M * M / NT;  // plane elements per thread
  constexpr int KB = 8;           // loaded in batches of 8 (registers)
  constexpr int NBAT = KT / KB;
  static_assert(NBAT * KB == KT, "batches");
  __shared__ T pl[M * RP];
  __shared__ T tab[4 * M];  // FFT twiddles (2M), K' of z (M), K' of y (M)

  const PassChain<T> ch = p.c[blockIdx.y];
  const long base = (long)blockIdx.x * M * M;
  const int tid = threadIdx.x;
  auto lds_at = [](int f) {
    const int y = f / M, z = f - y * M;
    return radix_pos<KIND, M>(y) * RP + S::col(radix_pos<KIND, M>(z));
  };

  for (int f = tid; f < 4 * M; f += NT)
    tab[f] = f < 2 * M ? p.tab_z[f] : (f < 3 * M ? p.tab_z[f + 2 * M] : p.tab_y[f + M]);
#pragma unroll
  for (int b = 0; b < NBAT; ++b) {
    T v[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) v[k] = radix_line_load(&ch.in[base + tid + NT * (b * KB + k)]);
    if (ch.pre) {
      T e[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) e[k] = ch.pre[base + tid + NT * (b * KB + k)];
#pragma unroll
      for (int k = 0; k < KB; ++k) v[k] *= e[k];
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) pl[lds_at(tid + NT * (b * KB + k))] = v[k];
  }
  __syncthreads();

  slab_transform<T, NT, true, KIND>(pl, tab, tab + 2 * M, p.tab_z + 2 * M);   // z
  slab_transform<T, NT, false, KIND>(pl, tab, tab + 3 * M, p.tab_y + 2 * M);  // y

#pragma unroll
  for (int b = 0; b < NBAT; ++b) {
    T v[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) v[k] = pl[lds_at(tid + NT * (b * KB + k))];
    if (ch.post) {
      T e[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) e[k] = ch.post[base + tid + NT * (b * KB + k)];
#pragma unroll
      for (int k = 0; k < KB; ++k) v[k] *= e[k];
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) radix_line_store(v[k], &ch.out[base + tid + NT * (b * KB + k)]);
  }
}

// Whether the fused z + y radix slab applies: both axes of length 128.
inline bool radix_slab_applies(int Y, int Z) { return Y == 128 && Z == 128; }

// One z + y radix slab (in -> out per chain, optional pre / post E) for 1 or
// 2 chains; tab_z and tab_y are the radix tables of the z and y axes.
template <typename T>
hipError_t launch_radix_slab(int kind, const RadixSlabParams<T>& p, int nchains, hipStream_t s) {
  if (nchains < 1 || nchains > 2 || !radix_slab_applies(p.Y, p.Z)) return hipErrorInvalidValue;
  dim3 grid((unsigned)p.X, (unsigned)nchains, 1);
  const dim3 block(RadixSlabShape<T>::NT);
  if (kind == kRadixP1)
    hipLaunchKernelGGL((radix_slab_kernel<T, kRadixP1>), grid, block, 0, s, p);
  else if (kind == kRadixDct)
    hipLaunchKernelGGL((radix_slab_kernel<T, kRadixDct>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((radix_slab_kernel<T, kRadixM3>), grid, block, 0, s, p);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// One launch per contour step at length 128 (all three axes): the plane phase
// of plane_phase.hpp with radix transforms.  A workgroup holds one plane of
// one chain (rows = the row axis, y for x-slabs and x for y-slabs, columns =
// z) in LDS as in the slab kernel above, finishes the previous step along the
// row axis, multiplies by E, writes the history entry, multiplies by E of the
// next step, applies the z and row-axis transforms of the next step and
// writes the scratch plane for the next phase: 4 n b per chain and step.
// 512 threads; each thread keeps the E values of its 32 plane elements in
// registers from the start, so the E stage waits for no load.
// ---------------------------------------------------------------------------
template <typename T, int KIND>
__global__ __launch_bounds__(512) void radix_phase_kernel(PhaseParams<T> p) {
  using S = RadixSlabShape<T>;
  constexpr int M = S::M, NT = 512, RP = S::RP;
  constexpr int KT = M * M / NT;  // plane elements per thread
  __shared__ T pl[M * RP];
  __shared__ T tab[4 * M];  // FFT twiddles (2M), K' of z (M), K' of the row axis (M)

  const PhaseChain<T> ch = p.c[blockIdx.y];
  const int RS = (int)p.row_stride;  // < 2^31 / 128 for the grids this kernel runs
  const long base = (long)blockIdx.x * p.plane_stride;
  const int tid = threadIdx.x;
  const bool fin = ch.hist != nullptr, next = ch.out != nullptr;
  // 32-bit offsets from per-plane base pointers (one register per address)
  const T* const in_b = ch.in + base;
  const T* const efin_b = fin ? ch.e_fin + base : nullptr;
  const T* const enext_b = next ? ch.e_next + base : nullptr;
  T* const hist_b = fin ? ch.hist + base : nullptr;
  T* const out_b = next ? ch.out + base : nullptr;
  auto gidx = [&](int f) {
    const int r = f / M, z = f - r * M;
    return r * RS + z;
  };
  auto lds_at = [](int f) {
    const int r = f / M, z = f - r * M;
    return radix_pos<KIND, M>(r) * RP + S::col(radix_pos<KIND, M>(z));
  };

  for (int f = tid; f < 4 * M; f += NT)
    tab[f] = f < 2 * M ? p.tab_c[f] : (f < 3 * M ? p.tab_c[f + 2 * M] : p.tab_r[f + M]);
  // the plane and E (the finished step's; the next step's in a start phase),
  // both in the coalesced element order f = tid + NT k.  A start phase reads
  // E behind the plane (its staging multiplies by it).  A finishing phase
  // reads the finished step's E after the staging barrier, so that the plane
  // does not share the memory system with it; FP64 only before the last
  // stage of the finishing transform: held through the whole transform its
  // 64 registers would spill.
  constexpr bool E_EARLY = sizeof(T) == 4;
  T e[KT];
  auto load_e = [&](const T* ea) {
#pragma unroll
    for (int k = 0; k < KT; ++k) e[k] = ea[gidx(tid + NT * k)];
  };
  {
    T v[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      // FP32: the plane is read once, non-temporally (measured 3% faster per
      // step at 128^3; FP64 unchanged within noise)
      if constexpr (sizeof(T) == 4) v[k] = __builtin_nontemporal_load(&in_b[gidx(tid + NT * k)]);
      else v[k] = in_b[gidx(tid + NT * k)];
    }
    if (!fin) load_e(enext_b);
    // (FP32: two loops, so that a finishing phase's staging does not wait
    // for E)
    if (fin && E_EARLY) {
#pragma unroll
      for (int k = 0; k < KT; ++k) pl[lds_at(tid + NT * k)] = v[k];
    } else {
#pragma unroll
      for (int k = 0; k < KT; ++k) pl[lds_at(tid + NT * k)] = fin ? v[k] : v[k] * e[k];
    }
  }
  __syncthreads();
  if (fin && E_EARLY) load_e(efin_b);

  if (fin) {
    // finish the previous step along the row axis
    slab_transform<T, NT, false, KIND>(pl, tab, tab + 3 * M, p.tab_r + 2 * M, [&] {
      if (!E_EARLY) load_e(efin_b);
    });
    // q_{s+1} = E_s (.) -> history (non-temporal); times E_{s+1}.  The
    // registers of E_s then hold q_{s+1}, and the history stores are issued
    // after the loop: a store in flight keeps its data registers, and a load
    // of the block junction's E between two stores made each wait for the
    // stores before it (32 round trips per thread)
    const bool junction = next && ch.e_next != ch.e_fin;
    if (junction) {
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int f = tid + NT * k;
        const int a = lds_at(f);
        const T q = pl[a] * e[k];
        pl[a] = q * enext_b[gidx(f)];
        e[k] = q;
      }
    } else {
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int a = lds_at(tid + NT * k);
        const T q = pl[a] * e[k];
        if (next) pl[a] = q * e[k];
        e[k] = q;
      }
    }
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int f = tid + NT * k;
      __builtin_nontemporal_store(e[k], &hist_b[gidx(f)]);
    }
    if (!next) return;
    __syncthreads();
  }
  // the next step: z, then the row axis
  slab_transform<T, NT, true, KIND>(pl, tab, tab + 2 * M, p.tab_c + 2 * M);
  slab_transform<T, NT, false, KIND>(pl, tab, tab + 3 * M, p.tab_r + 2 * M);
  // the half-finished plane for the next phase (non-temporal: read from
  // other XCDs)
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int f = tid + NT * k;
    __builtin_nontemporal_store(pl[lds_at(f)], &out_b[gidx(f)]);
  }
}

template <typename T>
hipError_t launch_radix_phase(const PhaseParams<T>& p, int nplanes, int nchains, hipStream_t s) {
  if (nchains < 1 || nchains > 2 || p.MR != 128 || p.MC != 128 || !p.tab_r || !p.tab_c ||
      p.row_stride * 127 + 128 > 0x7fffffffL)
    return hipErrorInvalidValue;
  dim3 grid((unsigned)nplanes, (unsigned)nchains, 1);
  if (p.rkind == kRadixP1)
    hipLaunchKernelGGL((radix_phase_kernel<T, kRadixP1>), grid, dim3(512), 0, s, p);
  else if (p.rkind == kRadixDct)
    hipLaunchKernelGGL((radix_phase_kernel<T, kRadixDct>), grid, dim3(512), 0, s, p);
  else
    hipLaunchKernelGGL((radix_phase_kernel<T, kRadixM3>), grid, dim3(512), 0, s, p);
  return hipGetLastError();
}

// Whether an axis of transform length M can take the radix pass.
inline bool radix_length_ok(int M) { return M >= 8 && M <= 128 && (M & (M - 1)) == 0; }

template <typename T, int AXIS, int KIND>
hipError_t launch_radix_pass_k(const RadixParams<T>& p, int M, int nchains, hipStream_t s) {
  const long n = (long)p.X * p.Y * p.Z;
  auto go = [&](auto mtag) -> hipError_t {
    constexpr int MM = decltype(mtag)::value;
    using S = RadixShape<MM>;
    const long nlines = n / MM;
    dim3 grid((unsigned)((nlines + S::TL - 1) / S::TL), (unsigned)nchains, 1);
    hipLaunchKernelGGL((radix_pass_kernel<T, AXIS, MM, KIND>), grid, dim3(S::NT), 0, s, p);
    return hipGetLastError();
  };
  switch (M) {
    case 8: return go(std::integral_constant<int, 8>{});
    case 16: return go(std::integral_constant<int, 16>{});
    case 32: return go(std::integral_constant<int, 32>{});
    case 64: return go(std::integral_constant<int, 64>{});
    case 128: return go(std::integral_constant<int, 128>{});
    default: return hipErrorInvalidValue;
  }
}

// One radix pass along `axis` (transform length = the axis length) for 1 or
// 2 chains.
template <typename T>
hipError_t launch_radix_pass(int axis, int kind, const RadixParams<T>& p, int nchains,
                             hipStream_t s) {
  if (nchains < 1 || nchains > 2) return hipErrorInvalidValue;
  const int M = axis == 0 ? p.X : (axis == 1 ? p.Y : p.Z);
  if (!radix_length_ok(M)) return hipErrorInvalidValue;
#define CRYS_RADIX_AX(A)                                                              \
  if (axis == A) {                                                                    \
    if (kind == kRadixP1) return launch_radix_pass_k<T, A, kRadixP1>(p, M, nchains, s); \
    if (kind == kRadixDct) return launch_radix_pass_k<T, A, kRadixDct>(p, M, nchains, s); \
    return launch_radix_pass_k<T, A, kRadixM3>(p, M, nchains, s);                      \
  }
  CRYS_RADIX_AX(0)
  CRYS_RADIX_AX(1)
  CRYS_RADIX_AX(2)
#undef CRYS_RADIX_AX
  return hipErrorInvalidValue;
}

}  // namespace crys
