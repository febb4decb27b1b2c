// One launch per contour step: alternating x-slab / y-slab phases.
//
// A contour step is q_{s+1} = E_s Dx Dy Dz E_s q_s.  A workgroup that holds a
// whole plane of the natural (x, y, z) layout can apply the two axes of that
// plane: an x-slab (fixed x, rows y, contiguous z) applies Dy and Dz, a y-slab
// (fixed y, rows x, contiguous z) applies Dx and Dz.  Chaining the steps as
//
//   phase 0 (x-slab):  Dz Dy E_0 q_0                    (start step 0)
//   phase 1 (y-slab):  Dx -> q_1 = E_0 (.) ; Dz Dx E_1 q_1
//   phase 2 (x-slab):  Dy -> q_2 = E_1 (.) ; Dz Dy E_2 q_2
//   ...
//   phase Ns:          finish step Ns-1 -> q_Ns
//
// gives every step one launch: each phase first applies its row-axis
// operator to finish the previous step, multiplies by that step's E, writes
// the history entry, multiplies by the next step's E and applies the row and
// column (z) operators of the next step.  The half-finished field lives in
// two scratch fields per chain that the phases alternate between (phase k
// reads what phase k-1 wrote; the split kernel below has two workgroups per
// plane, so a phase never rewrites its input in place).  Per chain and step this moves 4 n
// values (scratch in, E, history out, scratch out) instead of the 6 n of the
// two-launch step.  Every array stays in the natural layout, so the density
// and stress kernels read the histories unchanged.
//
// In LDS the plane is kept row-major (rows = row axis, contiguous columns =
// z).  Row-axis products read it as the B operand (4 rows x 16 contiguous
// columns per MFMA step, pitch PB); the z product reads it as the A operand
// (16 rows x 4 contiguous columns, pitch PA): no access is transposed.
// With CEN (centrosymmetric operators, all three axes of length MP, a
// multiple of 32) both products use the half-size u/v split described at
// SlabShape, with the mirror or the half-period pairing (cen_partner); the
// splits and recombinations are fused into the staging, the E/history stage
// and the final store.
#pragma once

#include "axis_pass.hpp"

namespace crys {

template <typename T>
struct PhaseChain {
  const T* in;      // plane source: q_0 (start phase) or the scratch field
  T* out;           // scratch field for the next phase, or null (finish only)
  T* hist;          // history entry written after the finishing product, or null (start)
  const T* e_fin;   // E of the step being finished (with hist)
  const T* e_next;  // E of the step being started (with out)
};

template <typename T>
struct PhaseParams {
  PhaseChain<T> c[2];
  const T* mat_r;  // row-axis operator: dense At (M x M) or At_c (M/2 x M)
  const T* mat_c;  // z operator
  const T* tab_r;  // radix tables of the row axis and z (length-128 radix phase), or null
  const T* tab_c;
  int rkind;       // RadixKind of the radix phase
  int shift;       // centrosymmetric pairing (cen_partner in axis_pass.hpp)
  const unsigned short* lo_r;  // bf16 low parts (FP32), or null
  const unsigned short* lo_c;
  int MR, MC;           // row axis length, z length
  int same_rc;          // the row operator equals the z operator (one copy loaded)
  // FP32 class-64 half-size operators in register order (load_operands_reg),
  // or null
  const T* reg_r;
  const T* reg_c;
  const unsigned short* reg_lo_r;
  const unsigned short* reg_lo_c;
  long plane_stride;    // elements between planes
  long row_stride;      // elements between rows of a plane
};

template <typename T, int MP, bool CEN, int NT_ = 512>
struct PhaseShape {
  static constexpr int NT = NT_;
  static constexpr int JT = MP / 16;
  static constexpr int H = MP / 2;
  static constexpr int KS = CEN ? MP / 8 : MP / 4;
  static constexpr int PB = MP + 16;                              // row-product pitch
  static constexpr int PA = MP + (sizeof(T) == 8 ? 2 : 4);        // z-product pitch
  // Wave w owns 16-row tile w % JT of the operator and the tiles s, s + WPJ,
  // ... of the other dimension, s = w / JT < WPJ.
  static constexpr int WPJ = (NT / 64) / JT >= 1 ? (NT / 64) / JT : 1;
  static constexpr int TPW = (JT + WPJ - 1) / WPJ;
  static __device__ __forceinline__ int tile(int w, int u) {
    const int s = w / JT, t = s + WPJ * u;
    return (s < WPJ && t < JT) ? t : -1;
  }
};

// Row-axis product: out[j][c] = sum_k A[j][k] src[k][c], src at pitch PB.
// Wave w owns row tile w % JT (operator registers) and the column tiles
// tile(w, u).  All LDS operands of the wave are read first; the MFMAs of its
// tiles are then issued interleaved (independent accumulators back to back).
template <typename T, int MP, bool CEN, int NT>
__device__ __forceinline__ void phase_row_product(
    const SlabOperands<T, MP, CEN>& op, bool use_lo, const T* lds, int w, int lane,
    typename MfmaOp<T>::v4 (&acc)[PhaseShape<T, MP, CEN, NT>::TPW]) {
  using S = PhaseShape<T, MP, CEN, NT>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int PB = S::PB, KS = S::KS, TPW = S::TPW;
  constexpr bool SPLIT = std::is_same<T, float>::value;
  T b[TPW][KS];
  bool on[TPW];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    acc[u] = V4{T(0), T(0), T(0), T(0)};
    const int ct = S::tile(w, u);
    on[u] = ct >= 0;
    const int col = (on[u] ? ct : 0) * 16 + (lane & 15);
#pragma unroll
    for (int n = 0; n < KS; ++n) b[u][n] = on[u] ? lds[op.brow[n] * PB + col] : T(0);
  }
  // the operator's low parts first, while the accumulators are small (added
  // last, a correction below half an ulp of the sum would be rounded away)
  if constexpr (SPLIT) {
    if (use_lo) {
#pragma unroll
      for (int c4 = 0; c4 < (KS + 3) / 4; ++c4) {
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
          if (!on[u]) continue;
          typename MfmaOp<float>::s4 av;
          float fb[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int n = 4 * c4 + i;
            av[i] = n < KS ? op.alo[n] : (short)0;
            fb[i] = n < KS ? b[u][n] : 0.f;
          }
          acc[u] = MfmaOp<float>::mma_bf16(av, bf16x4(fb), acc[u]);
        }
      }
    }
  }
#pragma unroll
  for (int n = 0; n < KS; ++n) {
#pragma unroll
    for (int u = 0; u < TPW; ++u)
      if (on[u]) acc[u] = MfmaOp<T>::mma(op.a[n], b[u][n], acc[u]);
  }
}

// z product: out[r][j] = sum_k src[r][k] Dz[j][k], src at pitch PA (the A
// operand), the operator in registers as the B operand.  Wave w owns output
// column tile w % JT (its operator registers) and the row tiles tile(w, u).
template <typename T, int MP, bool CEN, int NT>
__device__ __forceinline__ void phase_col_product(
    const SlabOperands<T, MP, CEN>& op, bool use_lo, const T* lds, int w, int lane,
    typename MfmaOp<T>::v4 (&acc)[PhaseShape<T, MP, CEN, NT>::TPW]) {
  using S = PhaseShape<T, MP, CEN, NT>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int PA = S::PA, KS = S::KS, TPW = S::TPW;
  constexpr bool SPLIT = std::is_same<T, float>::value;
  T a[TPW][KS];
  bool on[TPW];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    acc[u] = V4{T(0), T(0), T(0), T(0)};
    const int rt = S::tile(w, u);
    on[u] = rt >= 0;
    const int row = (on[u] ? rt : 0) * 16 + (lane & 15);
#pragma unroll
    for (int n = 0; n < KS; ++n) a[u][n] = on[u] ? lds[row * PA + op.brow[n]] : T(0);
  }
  // low parts first (see phase_row_product)
  if constexpr (SPLIT) {
    if (use_lo) {
#pragma unroll
      for (int c4 = 0; c4 < (KS + 3) / 4; ++c4) {
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
          if (!on[u]) continue;
          typename MfmaOp<float>::s4 bv;
          float fa[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int n = 4 * c4 + i;
            fa[i] = n < KS ? a[u][n] : 0.f;
            bv[i] = n < KS ? op.alo[n] : (short)0;
          }
          acc[u] = MfmaOp<float>::mma_bf16(bf16x4(fa), bv, acc[u]);
        }
      }
    }
  }
#pragma unroll
  for (int n = 0; n < KS; ++n) {
#pragma unroll
    for (int u = 0; u < TPW; ++u)
      if (on[u]) acc[u] = MfmaOp<T>::mma(a[u][n], op.a[n], acc[u]);
  }
}

// No work after the staging barrier (the launch-per-plane kernels).
struct NoStagingHook {
  __device__ void operator()() const {}
};

// One phase of one chain on one plane (base: its first element; RS: its row
// stride), for the single-field and the batched kernels.  `lds` holds at
// least MP * PhaseShape::PB elements.
//
// With RAW the plane is not read from ch.in: a persistent kernel has already
// moved it into LDS as compact rows (`raw`, element (r, c) at r * MC + c, see
// dma_plane), and `after_staging` runs once the plane has been staged, when
// `raw` may be overwritten (the persistent kernel starts the next plane's copy
// there).
template <typename T, int MP, bool CEN, int NT_, bool RAW = false,
          typename Hook = NoStagingHook>
__device__ __forceinline__ void plane_phase_body(const PhaseChain<T>& ch, const T* mat_r,
                                                 const T* mat_c,
                                                 const unsigned short* mat_r_lo,
                                                 const unsigned short* mat_c_lo, int MR, int MC,
                                                 long base, long RS, T* lds,
                                                 bool shift = false, const T* raw = nullptr,
                                                 const Hook& after_staging = Hook{},
                                                 int tid_in = -1) {
  using S = PhaseShape<T, MP, CEN, NT_>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int NT = S::NT, JT = S::JT, H = S::H, PB = S::PB, PA = S::PA, TPW = S::TPW;
  constexpr bool SPLIT = std::is_same<T, float>::value;

  // a persistent caller passes the thread index through an opaque copy, so
  // that nothing derived from it is hoisted out of its loop over planes and
  // held in registers across the planes
  const int tid = tid_in >= 0 ? tid_in : (int)threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bool fin = ch.hist != nullptr, next = ch.out != nullptr;
  const bool lo_r = SPLIT && mat_r_lo != nullptr, lo_c = SPLIT && mat_c_lo != nullptr;

  // operators: the row tile w % JT of the row operator, the output column
  // tile w % JT of the z operator
  SlabOperands<T, MP, CEN> orow, ocol;
  slab_load_operands<T, MP, CEN>(mat_r, mat_r_lo, MR, w % JT, lane, orow);
  slab_load_operands<T, MP, CEN>(mat_c, mat_c_lo, MC, w % JT, lane, ocol);

  // ---- every global load up front: the plane and the E values that the
  // E/history stage needs (same element per task), one memory latency ----
  // CEN task k: rows r < H and its partner (cen_partner) of column c; else
  // task k: (r, c)
  constexpr int TASKS = CEN ? H * MP : MP * MP;
  constexpr int KE = (TASKS + NT - 1) / NT;
  constexpr int NV = CEN ? 2 : 1;  // elements per task
  long g[KE][NV];
  bool ok[KE];
  T x[KE][NV], ef[KE][NV];
#pragma unroll
  for (int k = 0; k < KE; ++k) {
    const int f = tid + NT * k;
    const int r = f / MP, c = f - r * MP;
    ok[k] = f < TASKS && (CEN || (r < MR && c < MC));
    g[k][0] = base + r * RS + c;
    if constexpr (CEN) g[k][1] = base + cen_partner<MP>(r, shift) * RS + c;
  }
  // A start phase stages the plane times E of the next step and reads that E
  // behind the plane.  A finishing phase stages the plane alone and reads the
  // finished step's E after the staging barrier, while the finishing product
  // runs, so that the plane does not share the memory system with it
  // (measured 1% per FP64 step at 32^3 reduced, 5.9% on the two-workgroup
  // class-64 phase).  The other E of a block junction is read in the
  // E/history stage.
  auto load_e = [&](const T* e) {
#pragma unroll
    for (int k = 0; k < KE; ++k)
#pragma unroll
      for (int v = 0; v < NV; ++v) ef[k][v] = ok[k] ? e[g[k][v]] : T(0);
  };
  if constexpr (RAW) {
    // (the plane from its LDS copy)
    if (!fin) load_e(ch.e_next);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      x[k][0] = ok[k] ? raw[r * MC + c] : T(0);
      if constexpr (CEN) x[k][1] = ok[k] ? raw[cen_partner<MP>(r, shift) * MC + c] : T(0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < KE; ++k)
#pragma unroll
      for (int v = 0; v < NV; ++v) x[k][v] = ok[k] ? ch.in[g[k][v]] : T(0);
    if (!fin) load_e(ch.e_next);
  }

  // ---- stage the plane (start: times E of the next step) ----
  // (two loops, so that a finishing phase's staging does not wait for E)
  auto stage = [&](auto with_e) {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      if constexpr (CEN) {
        // rows r < H: u = x_r + x_{p(r)}; rows H + r: v = x_r - x_{p(r)}
        if (f >= TASKS) continue;
        const T a = decltype(with_e)::value ? x[k][0] * ef[k][0] : x[k][0];
        const T b = decltype(with_e)::value ? x[k][1] * ef[k][1] : x[k][1];
        lds[r * PB + c] = a + b;
        lds[(H + r) * PB + c] = a - b;
      } else {
        if (f < MP * MP)
          lds[r * PB + c] = decltype(with_e)::value ? x[k][0] * ef[k][0] : x[k][0];
      }
    }
  };
  if (fin) stage(std::false_type{});
  else stage(std::true_type{});
  __syncthreads();
  // (before a persistent caller starts its next plane's copy, which the
  // E/history stage then need not wait for)
  if (fin) load_e(ch.e_fin);
  if constexpr (RAW) after_staging();

  V4 acc[TPW];
  const int jt = w % JT;
  // the history entry (finishing phases), stored at the end of the phase;
  // non-temporal: read again only by the density kernel
  T hv[KE][NV];
  auto store_history = [&]() {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      if (f >= TASKS || (!CEN && !ok[k])) continue;
#pragma unroll
      for (int v = 0; v < NV; ++v) __builtin_nontemporal_store(hv[k][v], &ch.hist[g[k][v]]);
    }
  };
  if (fin) {
    // ---- finish the previous step along the row axis ----
    phase_row_product<T, MP, CEN, NT>(orow, lo_r, lds, w, lane, acc);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int ct = S::tile(w, u);
      if (ct >= 0) {
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lds_put<T>(lds, (jt * 16 + MfmaOp<T>::row(lane, r)) * PB + col, acc[u][r]);
      }
    }
    __syncthreads();

    // ---- q_{s+1} = E_s (.) -> history; times E_{s+1} for the next step ----
    // (the history values stay in registers and are stored at the end of the
    // phase: a store in flight keeps its data registers, and the compiler
    // makes the first instruction that reuses them wait for the store)
    const bool junction = next && ch.e_next != ch.e_fin;
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      if constexpr (CEN) {
        if (f >= TASKS) continue;
        const T U = lds[r * PB + c], V = lds[(H + r) * PB + c];
        T o1 = (U + V) * ef[k][0], o2 = (U - V) * ef[k][1];
        hv[k][0] = o1;
        hv[k][1] = o2;
        if (next) {
          o1 *= junction ? ch.e_next[g[k][0]] : ef[k][0];
          o2 *= junction ? ch.e_next[g[k][1]] : ef[k][1];
          lds[r * PB + c] = o1 + o2;
          lds[(H + r) * PB + c] = o1 - o2;
        }
      } else if (f < MP * MP) {
        T o = T(0);
        if (ok[k]) {
          o = lds[r * PB + c] * ef[k][0];
          hv[k][0] = o;
          if (next) o *= junction ? ch.e_next[g[k][0]] : ef[k][0];
        }
        if (next) lds[r * PB + c] = o;
      }
    }
    if (!next) {
      store_history();
      return;
    }
    __syncthreads();
  }

  // ---- start the next step: row-axis product ----
  phase_row_product<T, MP, CEN, NT>(orow, lo_r, lds, w, lane, acc);
  __syncthreads();
  if constexpr (CEN) {
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int ct = S::tile(w, u);
      if (ct >= 0) {
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lds_put<T>(lds, (jt * 16 + MfmaOp<T>::row(lane, r)) * PB + col, acc[u][r]);
      }
    }
    __syncthreads();
    // rows U (j < H) / V (H + j) -> values w at rows j, M-1-j of the paired
    // basis; then the z pairs of each row (column c and its partner cp in
    // memory order): A operand [row][u (c < H) | v (H + c)] at pitch PA
    constexpr int KB = (H * H + NT - 1) / NT;
    T nv[KB][4];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int f = tid + NT * k;
      if (f >= H * H) continue;
      const int j = f / H, c = f - (f / H) * H, cp = cen_partner<MP>(c, shift);
      const T P1 = lds[j * PB + c], P2 = lds[(H + j) * PB + c];
      const T P3 = lds[j * PB + cp], P4 = lds[(H + j) * PB + cp];
      const T w11 = P1 + P2, w21 = P1 - P2;  // (j, c), (M-1-j, c)
      const T w12 = P3 + P4, w22 = P3 - P4;  // (j, cp), (M-1-j, cp)
      nv[k][0] = w11 + w12;  // u (j, c)
      nv[k][1] = w11 - w12;  // v (j, c)
      nv[k][2] = w21 + w22;  // u (M-1-j, c)
      nv[k][3] = w21 - w22;  // v (M-1-j, c)
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int f = tid + NT * k;
      if (f >= H * H) continue;
      const int j = f / H, c = f - (f / H) * H;
      lds[j * PA + c] = nv[k][0];
      lds[j * PA + H + c] = nv[k][1];
      lds[(MP - 1 - j) * PA + c] = nv[k][2];
      lds[(MP - 1 - j) * PA + H + c] = nv[k][3];
    }
  } else {
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int ct = S::tile(w, u);
      if (ct >= 0) {
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lds_put<T>(lds, (jt * 16 + MfmaOp<T>::row(lane, r)) * PA + col, acc[u][r]);
      }
    }
  }
  __syncthreads();

  // ---- z product: rows r, output columns j (U | V for CEN) ----
  phase_col_product<T, MP, CEN, NT>(ocol, lo_c, lds, w, lane, acc);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int rt = S::tile(w, u);
    if (rt >= 0) {
      const int col = jt * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lds_put<T>(lds, (rt * 16 + MfmaOp<T>::row(lane, r)) * PB + col, acc[u][r]);
    }
  }
  __syncthreads();

  // ---- store the scratch plane: non-temporal for FP64 (the next phase reads
  // it from other XCDs, so it gains nothing from this XCD's L2); ordinary
  // stores for FP32, measured 6% faster per step at 32^3 reduced and 1.5%
  // at class 48 ----
  if constexpr (CEN) {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      if (f >= TASKS) continue;
      const int r = f / H, j = f - (f / H) * H;
      const T U = lds[r * PB + j], V = lds[r * PB + H + j];
      const long g0 = base + cen_index<MP>(r, shift) * RS;
      const long g3 = g0 + cen_partner<MP>(j, shift);
      if constexpr (SPLIT) {
        ch.out[g0 + j] = U + V;
        ch.out[g3] = U - V;
      } else {
        __builtin_nontemporal_store(U + V, &ch.out[g0 + j]);
        __builtin_nontemporal_store(U - V, &ch.out[g3]);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      if (f < TASKS && r < MR && c < MC) {
        if constexpr (SPLIT) ch.out[base + r * RS + c] = lds[r * PB + c];
        else __builtin_nontemporal_store(lds[r * PB + c], &ch.out[base + r * RS + c]);
      }
    }
  }
  if (fin) store_history();
}

template <typename T, int MP, bool CEN, int NT_ = 512>
__global__ __launch_bounds__(NT_) void plane_phase_kernel(PhaseParams<T> p) {
  __shared__ T lds[MP * PhaseShape<T, MP, CEN, NT_>::PB];
  plane_phase_body<T, MP, CEN, NT_>(p.c[blockIdx.y], p.mat_r, p.mat_c, p.lo_r, p.lo_c, p.MR, p.MC,
                                    (long)blockIdx.x * p.plane_stride, p.row_stride, lds,
                                    p.shift != 0);
}

// One 16 x 16 output tile with the plane as the B operand: out[j][c] =
// sum_k A[j][k] src[k][c], src rows k at `pitch`, column col = tile base +
// lane % 16.  The operator tile is in `op` (rows j of the wave's row tile).
template <typename T, int MP, bool CEN>
__device__ __forceinline__ typename MfmaOp<T>::v4 mfma_tile_b(const SlabOperands<T, MP, CEN>& op,
                                                              bool use_lo, const T* lds, int pitch,
                                                              int col) {
  using V4 = typename MfmaOp<T>::v4;
  constexpr int KS = SlabShape<MP, CEN>::KS;
  constexpr bool SPLIT = std::is_same<T, float>::value;
  T b[KS];
#pragma unroll
  for (int n = 0; n < KS; ++n) b[n] = lds[op.brow[n] * pitch + col];
  V4 acc = V4{T(0), T(0), T(0), T(0)};
  // low parts first (see phase_row_product)
  if constexpr (SPLIT) {
    if (use_lo) {
#pragma unroll
      for (int c4 = 0; c4 < KS / 4; ++c4) {
        typename MfmaOp<float>::s4 av;
        float fb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          av[i] = op.alo[4 * c4 + i];
          fb[i] = b[4 * c4 + i];
        }
        acc = MfmaOp<float>::mma_bf16(av, bf16x4(fb), acc);
      }
    }
  }
#pragma unroll
  for (int n = 0; n < KS; ++n) acc = MfmaOp<T>::mma(op.a[n], b[n], acc);
  return acc;
}

// One 16 x 16 output tile with the plane as the A operand: out[r][j] =
// sum_k src[r][k] D[j][k], src row r = tile base + lane % 16 at `pitch`.
template <typename T, int MP, bool CEN>
__device__ __forceinline__ typename MfmaOp<T>::v4 mfma_tile_a(const SlabOperands<T, MP, CEN>& op,
                                                              bool use_lo, const T* lds, int pitch,
                                                              int row) {
  using V4 = typename MfmaOp<T>::v4;
  constexpr int KS = SlabShape<MP, CEN>::KS;
  constexpr bool SPLIT = std::is_same<T, float>::value;
  T a[KS];
#pragma unroll
  for (int n = 0; n < KS; ++n) a[n] = lds[row * pitch + op.brow[n]];
  V4 acc = V4{T(0), T(0), T(0), T(0)};
  // low parts first (see phase_row_product)
  if constexpr (SPLIT) {
    if (use_lo) {
#pragma unroll
      for (int c4 = 0; c4 < KS / 4; ++c4) {
        typename MfmaOp<float>::s4 bv;
        float fa[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          fa[i] = a[4 * c4 + i];
          bv[i] = op.alo[4 * c4 + i];
        }
        acc = MfmaOp<float>::mma_bf16(bf16x4(fa), bv, acc);
      }
    }
  }
#pragma unroll
  for (int n = 0; n < KS; ++n) acc = MfmaOp<T>::mma(a[n], op.a[n], acc);
  return acc;
}

// Operator registers for half of the k steps of one 16-row tile of a
// centrosymmetric class-MP operator (far-first order as slab_load_operands):
// steps 4 kh .. 4 kh + 3 of the MP / 8.
template <typename T, int MP>
struct KHalfOperands {
  T a[4];
  short alo[std::is_same<T, float>::value ? 4 : 1];
  int brow[4];
};

template <typename T, int MP>
__device__ __forceinline__ void load_operands_khalf(const T* __restrict__ mat,
                                                    const unsigned short* __restrict__ lo, int jt,
                                                    int lane, int kh, KHalfOperands<T, MP>& op,
                                                    int pitch = MP) {
  constexpr int KS = MP / 8, JT = MP / 16;
  static_assert(KS == 8, "k halves of 4 steps");
  constexpr bool SPLIT = std::is_same<T, float>::value;
  const int half = JT / 2;
  const int jb = jt % half;
  const int kbase = jt < half ? 0 : MP / 2;
  int ord[KS];
  far_first_order<KS, JT / 2, sizeof(T) == 8>(jb, ord);
  const int jrow = jt * 16 + (lane & 15);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int o = kh ? ord[4 + i] : ord[i];
    const int k = 4 * o + (lane >> 4);
    op.brow[i] = kbase + k;
    op.a[i] = mat[k * pitch + jrow];
    if constexpr (SPLIT) op.alo[i] = lo ? (short)lo[k * pitch + jrow] : (short)0;
  }
}

// Operator registers of row tile jt of a half-size FP32 operator of length 64
// from its register-order copy (upload_register_order in crys_solver.hip):
// lane l finds its 8 values, in the far-first order of slab_load_operands, at
// reg[(jt * 64 + l) * 8 ..] and their bf16 low parts at the same index of
// lo: three 16-byte loads instead of 16 scattered ones, the same registers.
__device__ __forceinline__ void load_operands_reg(const float* __restrict__ reg,
                                                  const unsigned short* __restrict__ lo, int jt,
                                                  int lane, SlabOperands<float, 64, true>& op) {
  int ord[8];
  far_first_order<8, 2>(jt & 1, ord);
  const int kbase = jt < 2 ? 0 : 32;
#pragma unroll
  for (int n = 0; n < 8; ++n) op.brow[n] = kbase + 4 * ord[n] + (lane >> 4);
  const size_t at = ((size_t)jt * 64 + lane) * 8;
  const float4 v0 = *(const float4*)(reg + at), v1 = *(const float4*)(reg + at + 4);
  op.a[0] = v0.x;
  op.a[1] = v0.y;
  op.a[2] = v0.z;
  op.a[3] = v0.w;
  op.a[4] = v1.x;
  op.a[5] = v1.y;
  op.a[6] = v1.z;
  op.a[7] = v1.w;
  const uint4 l4 = *(const uint4*)(lo + at);
  const unsigned lw[4] = {l4.x, l4.y, l4.z, l4.w};
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    op.alo[2 * n] = (short)(lw[n] & 0xffffu);
    op.alo[2 * n + 1] = (short)(lw[n] >> 16);
  }
}

// Partial tile over one k half, the plane as the B operand (see mfma_tile_b).
template <typename T, int MP>
__device__ __forceinline__ typename MfmaOp<T>::v4 mfma_tile_b_khalf(
    const KHalfOperands<T, MP>& op, bool use_lo, const T* lds, int pitch, int col) {
  using V4 = typename MfmaOp<T>::v4;
  T b[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) b[i] = lds[op.brow[i] * pitch + col];
  V4 acc = V4{T(0), T(0), T(0), T(0)};
  // low parts first (see phase_row_product)
  if constexpr (std::is_same<T, float>::value) {
    if (use_lo) {
      typename MfmaOp<float>::s4 av;
      float fb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        av[i] = op.alo[i];
        fb[i] = b[i];
      }
      acc = MfmaOp<float>::mma_bf16(av, bf16x4(fb), acc);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) acc = MfmaOp<T>::mma(op.a[i], b[i], acc);
  return acc;
}

// Partial tile over one k half, the plane as the A operand (see mfma_tile_a).
template <typename T, int MP>
__device__ __forceinline__ typename MfmaOp<T>::v4 mfma_tile_a_khalf(
    const KHalfOperands<T, MP>& op, bool use_lo, const T* lds, int pitch, int row) {
  using V4 = typename MfmaOp<T>::v4;
  T a[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = lds[row * pitch + op.brow[i]];
  V4 acc = V4{T(0), T(0), T(0), T(0)};
  // low parts first (see phase_row_product)
  if constexpr (std::is_same<T, float>::value) {
    if (use_lo) {
      typename MfmaOp<float>::s4 bv;
      float fa[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        fa[i] = a[i];
        bv[i] = op.alo[i];
      }
      acc = MfmaOp<float>::mma_bf16(bf16x4(fa), bv, acc);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) acc = MfmaOp<T>::mma(a[i], op.a[i], acc);
  return acc;
}

// The class-64 centrosymmetric phase on two workgroups per plane, so that
// 2 chains x 64 planes fill 256 CUs.  Both workgroups of a plane load it
// whole and apply the finishing row product whole (the next step's row
// product needs every row of its result); each then writes the history rows,
// applies the next row product and the z product, and stores only for its
// half h of the plane: the rows j and M-1-j (positions of the paired basis,
// see cen_index) with j in [16 h, 16 h + 16), i.e.
// the U/V row tiles h and 2 + h and the output row tiles h and 3 - h.  Per
// CU this is 2 product units instead of 3.  The partners of a plane have
// workgroup ids 8 apart, which places them on one XCD (shared L2).
// 1024 threads (4 waves per SIMD): the finishing product is 16 tiles, one
// per wave; the other two are 8 tiles each, split over two waves by k half
// (waves w and w + 8).  The second half's partial tile goes to LDS rows this
// workgroup does not own (row tile t xor 1, i.e. rows j xor 16) and is added
// where the tile is read.
//
// The next row product runs on the 8 waves whose finishing row tile is h or
// 2 + h, over the whole k range with the finishing product's operator
// registers, and the z product on waves 0..7 over the whole k range: no
// second row-operator load and no partial tiles to merge.  FP64 also runs the
// finishing product on waves 0..7 with two column tiles each, sharing the
// operator registers (FIN8).  Wave w runs on SIMD w % 4, so the finishing row
// tile of wave w is (w + w / 4) % 4 (PERM): the waves that go on to the next
// row product then fall two (one with FIN8) on each SIMD, where w % 4 would put
// them all on two SIMDs.  Measured per step at 64^3 reduced against the k-half
// split of the two products over all 16 waves (KHALF, kept for A/B timing):
// FP64 9.97 -> 9.70 us (9.17 with PERM), FP32 8.22 -> 7.74 us (7.40 with
// PERM); FIN8 measured 3-4% slower for FP32, 0.5% faster for FP64.
template <typename T, bool KHALF, bool FIN>
__device__ __forceinline__ void plane_phase_half_body(const PhaseParams<T>& p) {
  constexpr bool REUSE = !KHALF, ZFULL = !KHALF;
  constexpr bool FIN8 = !KHALF && std::is_same<T, double>::value;
  constexpr bool PERM = !KHALF;
  constexpr int MP = 64, H = 32, NT = 1024, JT = 4;
  using S = PhaseShape<T, MP, true, NT>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int PB = S::PB, PA = S::PA, TPW = S::TPW;
  static_assert(TPW == 1, "one finishing tile per wave");
  constexpr bool SPLIT = std::is_same<T, float>::value;
  // two plane buffers: every product reads one and writes the other, so
  // only the barrier that publishes a result separates two stages
  __shared__ T lds2[2 * MP * PB];
  T* const buf0 = lds2;
  T* const buf1 = lds2 + MP * PB;
  // FP64: the row and z operators (half-size, H x MP) are loaded once per
  // workgroup into LDS (pitch OP: the four k rows of one read fall on
  // alternate bank halves), and each product reads its operator registers
  // from there: the waves' registers repeat each operator tile 2 to 4 times,
  // 128 KB per workgroup from global memory otherwise, which delays the
  // plane.  FP32 operators are half the bytes and load straight into
  // registers (measured faster for FP32, slower for FP64).
  constexpr bool OPS_LDS = !SPLIT;
  constexpr int OP = MP + 16;
  __shared__ T opl[OPS_LDS ? 2 * H * OP : 1];

  const int id = blockIdx.x;
  const int h = (id >> 3) & 1;
  const int plane = ((id >> 4) << 3) | (id & 7);
  const PhaseChain<T> ch = p.c[blockIdx.y];
  const long RS = p.row_stride;
  const long base = (long)plane * p.plane_stride;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // row tile of wave w's finishing product (and of its next-row tiles): w % 4,
  // or with PERM (w + w / 4) % 4, which spreads the waves of one row tile over
  // the four SIMDs (wave w on SIMD w % 4)
  const int rho = PERM ? ((w + (w >> 2)) & 3) : (w % JT);
  // FIN: the phase finishes a step (every phase but the start phase); both
  // chains are at the same phase
  constexpr bool fin = FIN;
  const bool next = ch.out != nullptr;
  const bool lo_r = SPLIT && p.lo_r != nullptr, lo_c = SPLIT && p.lo_c != nullptr;
  const bool shift = p.shift != 0;

  // operators: finishing row product (row tile w % 4, all k); next row
  // product (row tile h or 2 + h, k half kh); z product (output column tile
  // item / 2, k half kh)
  const int item = w & 7, kh = w >> 3;
  const int r2t = (item & 1) ? 2 + h : h;
  // z product tile of waves 0..7 (ZFULL): the output column tile is the
  // wave's finishing row tile rho, and the row tile h for waves 0..3, 3 - h
  // for waves 4..7, so that where the z operator equals the row operator the
  // finishing product's registers serve the z product too
  const int ct2 = ZFULL ? rho : item >> 1;
  const int rt3 = ZFULL ? (w < 4 ? h : 3 - h) : ((item & 1) ? 3 - h : h);
  SlabOperands<T, MP, true> o1, o3f;  // o3f: whole z operator tile (ZFULL)
  KHalfOperands<T, MP> o2, o3;

  // ---- global loads up front: the operators, the plane and its E values ----
  constexpr int KO = H * MP / NT;
  static_assert(KO * NT == H * MP, "operator copy splits evenly");
  // (the z operator is not loaded where it equals the row operator: equal
  // axis and cell lengths, as on cubic grids; measured 0.37 us per FP64 step
  // for the 32 KB of both operators)
  const bool same_rc = p.same_rc != 0;
  T om[2][OPS_LDS ? KO : 1];
  if constexpr (OPS_LDS) {
#pragma unroll
    for (int k = 0; k < KO; ++k) om[0][k] = p.mat_r[tid + NT * k];
    if (!same_rc) {
#pragma unroll
      for (int k = 0; k < KO; ++k) om[1][k] = p.mat_c[tid + NT * k];
    }
  } else {
    // (register-order copies where they exist: the same values, fewer loads)
    if constexpr (SPLIT) {
      if (p.reg_r && p.reg_lo_r) load_operands_reg(p.reg_r, p.reg_lo_r, rho, lane, o1);
      else slab_load_operands<T, MP, true>(p.mat_r, p.lo_r, MP, rho, lane, o1);
    } else {
      slab_load_operands<T, MP, true>(p.mat_r, p.lo_r, MP, rho, lane, o1);
    }
    if constexpr (!REUSE) load_operands_khalf<T, MP>(p.mat_r, p.lo_r, r2t, lane, kh, o2);
    if constexpr (ZFULL) {
      if (w < 8 && !same_rc) {
        if constexpr (SPLIT) {
          if (p.reg_c && p.reg_lo_c) load_operands_reg(p.reg_c, p.reg_lo_c, ct2, lane, o3f);
          else slab_load_operands<T, MP, true>(p.mat_c, p.lo_c, MP, ct2, lane, o3f);
        } else {
          slab_load_operands<T, MP, true>(p.mat_c, p.lo_c, MP, ct2, lane, o3f);
        }
      }
    } else {
      load_operands_khalf<T, MP>(p.mat_c, p.lo_c, ct2, lane, kh, o3);
    }
  }
  constexpr int KE = H * MP / NT;
  long g1[KE], g2[KE];
  T x1[KE], x2[KE], ef1[KE], ef2[KE];
#pragma unroll
  for (int k = 0; k < KE; ++k) {
    const int f = tid + NT * k;
    const int r = f / MP, c = f - r * MP;
    g1[k] = base + r * RS + c;
    g2[k] = base + cen_partner<MP>(r, shift) * RS + c;
  }
  // The plane is read once, non-temporally, so that it does not displace the
  // E values (read again by every step) from this XCD's L2: measured 0.7% to
  // 2.2% faster per FP64 step at 64^3 reduced (two machines), FP32 within 0.5%.
  // A start phase stages the plane times E of the next step and reads that E
  // behind the plane; a finishing phase reads its E after the staging
  // barrier, while the finishing product runs, so that the plane's loads do
  // not share the memory system with it.  The other E of a block junction is
  // read in the E/history stage.
#pragma unroll
  for (int k = 0; k < KE; ++k) {
    x1[k] = __builtin_nontemporal_load(&ch.in[g1[k]]);
    x2[k] = __builtin_nontemporal_load(&ch.in[g2[k]]);
  }
  if constexpr (!FIN) {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      ef1[k] = ch.e_next[g1[k]];
      ef2[k] = ch.e_next[g2[k]];
    }
  }

  // ---- operators to LDS; stage u/v rows (start: times E of the next step) ----
  if constexpr (OPS_LDS) {
#pragma unroll
    for (int k = 0; k < KO; ++k) {
      const int f = tid + NT * k, r = f / MP, c = f - r * MP;
      opl[r * OP + c] = om[0][k];
      if (!same_rc) opl[H * OP + r * OP + c] = om[1][k];
    }
  }
  const T* const opl_z = same_rc ? opl : opl + H * OP;
  if (fin) {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      buf0[r * PB + c] = x1[k] + x2[k];
      buf0[(H + r) * PB + c] = x1[k] - x2[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      const T a = x1[k] * ef1[k], b = x2[k] * ef2[k];
      buf0[r * PB + c] = a + b;
      buf0[(H + r) * PB + c] = a - b;
    }
  }
  __syncthreads();
  if constexpr (FIN) {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      ef1[k] = ch.e_fin[g1[k]];
      ef2[k] = ch.e_fin[g2[k]];
    }
  }

  // the history entry of this half (finishing phases), stored at the end of
  // the phase; non-temporal: the entries are read again only by the density
  // kernel after the whole propagation
  const bool junction = fin && next && ch.e_next != ch.e_fin;
  T hv1[KE], hv2[KE];
  auto store_history = [&]() {
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int r = (tid + NT * k) / MP;
      if ((r >> 4) == h) {
        __builtin_nontemporal_store(hv1[k], &ch.hist[g1[k]]);
        __builtin_nontemporal_store(hv2[k], &ch.hist[g2[k]]);
      }
    }
  };

  if (fin) {
    if constexpr (OPS_LDS) {
      if (!FIN8 || w < 8) slab_load_operands<T, MP, true>(opl, nullptr, MP, rho, lane, o1, OP);
    }
    // ---- finishing row product, whole plane: one tile per wave, or (FIN8)
    // two column tiles per wave on waves 0..7 sharing the operator registers ----
    if constexpr (FIN8) {
      using S8 = PhaseShape<T, MP, true, 512>;
      if (w < 8) {
        V4 acc8[S8::TPW];
        phase_row_product<T, MP, true, 512>(o1, lo_r, buf0, w, lane, acc8);
#pragma unroll
        for (int u = 0; u < S8::TPW; ++u) {
          const int col = S8::tile(w, u) * 16 + (lane & 15);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            lds_put<T, false>(buf1, (rho * 16 + MfmaOp<T>::row(lane, r)) * PB + col, acc8[u][r]);
        }
      }
    } else {
      V4 acc[TPW];
      phase_row_product<T, MP, true, NT>(o1, lo_r, buf0, w, lane, acc);
      const int ct = S::tile(w, 0);
      const int col = ct * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lds_put<T, false>(buf1, (rho * 16 + MfmaOp<T>::row(lane, r)) * PB + col, acc[0][r]);
    }
    __syncthreads();
    // ---- E, history rows of this half, E of the next step ----
    // (FP32 issues the history stores here, and they complete while the next
    // products run: 1% faster per step than at the end of the phase.  A store
    // in flight keeps its data registers, and the compiler makes the first
    // instruction that reuses them wait for the store to complete, so the
    // history values are held to the end of the phase, where nothing has
    // reused their registers.  FP64 stores them at the end of the phase,
    // measured 1% faster than here)
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      const T U = buf1[r * PB + c], V = buf1[(H + r) * PB + c];
      T o1v = (U + V) * ef1[k], o2v = (U - V) * ef2[k];
      hv1[k] = o1v;
      hv2[k] = o2v;
      if (next) {
        o1v *= junction ? ch.e_next[g1[k]] : ef1[k];
        o2v *= junction ? ch.e_next[g2[k]] : ef2[k];
        buf0[r * PB + c] = o1v + o2v;
        buf0[(H + r) * PB + c] = o1v - o2v;
      }
    }
    if (SPLIT || !next) store_history();
    if (!next) return;
    __syncthreads();
  }

  // ---- next row product: U/V row tiles h and 2 + h ----
  if constexpr (REUSE && FIN8) {
    // the waves 0..7 holding row tile h or 2 + h: two column tiles each
    using S8 = PhaseShape<T, MP, true, 512>;
    const int rt = rho;
    if (w < 8 && (rt == h || rt == 2 + h)) {
      if constexpr (OPS_LDS) {
        if (!fin) slab_load_operands<T, MP, true>(opl, nullptr, MP, rt, lane, o1, OP);
      }
      V4 acc8[S8::TPW];
      phase_row_product<T, MP, true, 512>(o1, lo_r, buf0, w, lane, acc8);
#pragma unroll
      for (int u = 0; u < S8::TPW; ++u) {
        const int col = S8::tile(w, u) * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lds_put<T, false>(buf1, (rt * 16 + MfmaOp<T>::row(lane, r)) * PB + col, acc8[u][r]);
      }
    }
  } else if constexpr (REUSE) {
    const int rt = rho;
    if (rt == h || rt == 2 + h) {
      if constexpr (OPS_LDS) {
        if (!fin) slab_load_operands<T, MP, true>(opl, nullptr, MP, rt, lane, o1, OP);
      }
      V4 acc[TPW];
      phase_row_product<T, MP, true, NT>(o1, lo_r, buf0, w, lane, acc);
      const int col = S::tile(w, 0) * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lds_put<T, false>(buf1, (rt * 16 + MfmaOp<T>::row(lane, r)) * PB + col, acc[0][r]);
    }
  } else {
    // k halves over all 16 waves
    if constexpr (OPS_LDS) load_operands_khalf<T, MP>(opl, nullptr, r2t, lane, kh, o2, OP);
    const V4 acc = mfma_tile_b_khalf<T, MP>(o2, lo_r, buf0, PB, ct2 * 16 + (lane & 15));
    const int rt = kh ? (r2t ^ 1) : r2t;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      lds_put<T>(buf1, (rt * 16 + MfmaOp<T>::row(lane, r)) * PB + ct2 * 16 + (lane & 15), acc[r]);
  }
  __syncthreads();
  // ---- rows j, M-1-j of this half (j = 16 h + f / 32) -> z pairs at PA
  // (column c and its partner cp in memory order) ----
  const bool mid = tid < 16 * H;
  if (mid) {
    const int j = 16 * h + tid / H, c = tid - (tid / H) * H, jo = j ^ 16;
    const int cp = cen_partner<MP>(c, shift);
    T nv[4];
    // (the k-half variant sums its two partial tiles, rows j and j ^ 16)
    const T P1 = REUSE ? buf1[j * PB + c] : buf1[j * PB + c] + buf1[jo * PB + c];
    const T P2 = REUSE ? buf1[(H + j) * PB + c] : buf1[(H + j) * PB + c] + buf1[(H + jo) * PB + c];
    const T P3 = REUSE ? buf1[j * PB + cp] : buf1[j * PB + cp] + buf1[jo * PB + cp];
    const T P4 = REUSE ? buf1[(H + j) * PB + cp]
                       : buf1[(H + j) * PB + cp] + buf1[(H + jo) * PB + cp];
    const T w11 = P1 + P2, w21 = P1 - P2;
    const T w12 = P3 + P4, w22 = P3 - P4;
    nv[0] = w11 + w12;
    nv[1] = w11 - w12;
    nv[2] = w21 + w22;
    nv[3] = w21 - w22;
    buf0[j * PA + c] = nv[0];
    buf0[j * PA + H + c] = nv[1];
    buf0[(MP - 1 - j) * PA + c] = nv[2];
    buf0[(MP - 1 - j) * PA + H + c] = nv[3];
  }
  __syncthreads();
  // ---- z product: output row tiles h and 3 - h ----
  if constexpr (ZFULL) {
    // 8 waves over the whole k range
    // (with equal operators o1 serves as the z operator tile: FP32 loads it
    // at the start, FP64 for the finishing product of waves 0..7)
    const bool zr = same_rc && (SPLIT || (FIN8 && fin));
    if (w < 8) {
      if constexpr (OPS_LDS) {
        if (!zr) slab_load_operands<T, MP, true>(opl_z, nullptr, MP, ct2, lane, o3f, OP);
      }
      const int arow = rt3 * 16 + (lane & 15);
      const V4 acc = zr ? mfma_tile_a<T, MP, true>(o1, lo_c, buf0, PA, arow)
                        : mfma_tile_a<T, MP, true>(o3f, lo_c, buf0, PA, arow);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lds_put<T, false>(buf1, (rt3 * 16 + MfmaOp<T>::row(lane, r)) * PB + ct2 * 16 + (lane & 15),
                   acc[r]);
    }
  } else {
    // k halves over all 16 waves
    if constexpr (OPS_LDS) load_operands_khalf<T, MP>(opl_z, nullptr, ct2, lane, kh, o3, OP);
    const V4 acc = mfma_tile_a_khalf<T, MP>(o3, lo_c, buf0, PA, rt3 * 16 + (lane & 15));
    const int rt = kh ? (rt3 ^ 1) : rt3;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      lds_put<T>(buf1, (rt * 16 + MfmaOp<T>::row(lane, r)) * PB + ct2 * 16 + (lane & 15), acc[r]);
  }
  __syncthreads();
  // ---- store this half's rows: non-temporal for FP64 as in
  // plane_phase_kernel; ordinary FP32 stores measured about 1% faster per
  // step, ordinary FP64 stores 2% slower ----
  {
    const int rr = tid / H, j = tid - (tid / H) * H;
    const int row = rr < 16 ? 16 * h + rr : 48 - 16 * h + (rr - 16);
    const int ro = row ^ 16;
    // (the k-half variant sums its two partial tiles, rows row and row ^ 16)
    const T U = ZFULL ? buf1[row * PB + j] : buf1[row * PB + j] + buf1[ro * PB + j];
    const T V = ZFULL ? buf1[row * PB + H + j] : buf1[row * PB + H + j] + buf1[ro * PB + H + j];
    const long g0 = base + cen_index<MP>(row, shift) * RS;
    const long g3 = g0 + cen_partner<MP>(j, shift);
    if constexpr (SPLIT) {
      ch.out[g0 + j] = U + V;
      ch.out[g3] = U - V;
    } else {
      __builtin_nontemporal_store(U + V, &ch.out[g0 + j]);
      __builtin_nontemporal_store(U - V, &ch.out[g3]);
    }
  }
  if constexpr (FIN) {
    if constexpr (SPLIT) {
      // (the history values of the stores in flight)
#pragma unroll
      for (int k = 0; k < KE; ++k) asm volatile("" ::"v"(hv1[k]), "v"(hv2[k]));
    } else {
      store_history();
    }
  }
}

template <typename T, bool FIN>
__global__ __launch_bounds__(1024) void plane_phase_half_kernel(PhaseParams<T> p) {
  plane_phase_half_body<T, false, FIN>(p);
}

// The same phase with the next row and z products split by k half over all
// 16 waves (CRYS_PHASE_REUSE=0, for A/B timing).
template <typename T, bool FIN>
__global__ __launch_bounds__(1024) void plane_phase_half_khalf_kernel(PhaseParams<T> p) {
  plane_phase_half_body<T, true, FIN>(p);
}

// ---------------------------------------------------------------------------
// One launch per contour step for a 3m solver with translational parts,
// whose diffusion is a sum of separable terms X_t (x) Y_t (x) Z_t (the folded
// operator of m3_glide.hpp).  The phases alternate as above, one scratch
// plane per term: an x-slab phase starts step s with Z_t Y_t (E q_s) for
// every term t and writes nt planes; the y-slab phase after it finishes the
// step with sum_t X_t (plane t), accumulated in registers, applies E, writes
// the history entry, and starts step s + 1 with Z_t X_t (E q_{s+1}) per term;
// and so on.  3 nt products per phase instead of 3, nt planes read and nt
// written, one launch per step.  Dense operators (the terms are not
// centrosymmetric); bufA holds the staged plane for all terms of the start,
// bufB each term's products.
constexpr int kMaxGlideTerms = 8;

template <typename T>
struct PhaseTermsParams {
  PhaseChain<T> c[2];  // in: term 0's plane (k > 0) or q_0 (start phase); out: term 0's plane
  long tstride;        // elements between the term planes of in (k > 0) and of out
  int nt;
  const T* mat_r[kMaxGlideTerms];  // row-axis factor of each term (X_t or Y_t)
  const T* mat_c[kMaxGlideTerms];  // z factor of each term
  const unsigned short* lo_r[kMaxGlideTerms];  // FP32: bf16 low parts, or null
  const unsigned short* lo_c[kMaxGlideTerms];
  int MR, MC;
  long plane_stride, row_stride;
};

// HALF: two workgroups per plane (workgroup ids 8 apart, which places them on
// one XCD), as the class-64 split phase: each finishes the step for the whole
// plane (the next row product needs every row) and starts the next one for
// its half h of the rows only, one output tile per wave: 2 nt products per
// workgroup instead of 3 nt, on twice the workgroups.
template <typename T, int MP, int NT_, bool HALF = false>
__global__ __launch_bounds__(NT_) void plane_phase_terms_kernel(PhaseTermsParams<T> p) {
  using S = PhaseShape<T, MP, false, NT_>;
  using V4 = typename MfmaOp<T>::v4;
  constexpr int NT = S::NT, JT = S::JT, PB = S::PB, PA = S::PA, TPW = S::TPW;
  constexpr bool SPLIT = std::is_same<T, float>::value;
  constexpr int KE = (MP * MP + NT - 1) / NT;
  static_assert(!HALF || NT / 64 == (JT / 2) * JT, "one start tile per wave");
  __shared__ T bufA[MP * PB];
  __shared__ T bufB[MP * PB];
  const PhaseChain<T> ch = p.c[blockIdx.y];
  const int id = blockIdx.x;
  const int h = HALF ? (id >> 3) & 1 : 0;
  const int plane = HALF ? (((id >> 4) << 3) | (id & 7)) : id;
  const long base = (long)plane * p.plane_stride, RS = p.row_stride;
  const int MR = p.MR, MC = p.MC, nt = p.nt;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, jt = w % JT;
  const bool fin = ch.hist != nullptr, next = ch.out != nullptr;

  long g[KE];
  bool ok[KE];
  T ef[KE];
  const T* e_a = fin ? ch.e_fin : ch.e_next;
#pragma unroll
  for (int k = 0; k < KE; ++k) {
    const int f = tid + NT * k;
    const int r = f / MP, c = f - r * MP;
    ok[k] = f < MP * MP && r < MR && c < MC;
    g[k] = base + r * RS + c;
  }
#pragma unroll
  for (int k = 0; k < KE; ++k) ef[k] = ok[k] ? e_a[g[k]] : T(0);
  // one plane into bufA (zero outside MR x MC), times E for the start phase
  auto stage = [&](const T* src, bool times_e) {
    T x[KE];
#pragma unroll
    for (int k = 0; k < KE; ++k) x[k] = ok[k] ? src[g[k]] : T(0);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      if (f < MP * MP) bufA[r * PB + c] = times_e ? x[k] * ef[k] : x[k];
    }
  };

  V4 acc[TPW];
  if (fin) {
    // ---- finish step s: sum over the terms of the row factor times its plane ----
    // (two terms' planes are read at once, into bufA and bufB, so that the
    // second one's memory latency overlaps the first one's)
    V4 sum[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) sum[u] = V4{T(0), T(0), T(0), T(0)};
    auto add_term = [&](int t, const T* buf) {
      SlabOperands<T, MP, false> orow;
      slab_load_operands<T, MP, false>(p.mat_r[t], p.lo_r[t], MR, jt, lane, orow);
      phase_row_product<T, MP, false, NT>(orow, SPLIT && p.lo_r[t] != nullptr, buf, w, lane,
                                          acc);
#pragma unroll
      for (int u = 0; u < TPW; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) sum[u][r] += acc[u][r];
    };
    for (int t = 0; t < nt; t += 2) {
      const bool two = t + 1 < nt;
      const T* src0 = ch.in + t * p.tstride;
      const T* src1 = ch.in + (two ? t + 1 : t) * p.tstride;
      T x0[KE], x1[KE];
#pragma unroll
      for (int k = 0; k < KE; ++k) x0[k] = ok[k] ? src0[g[k]] : T(0);
      if (two) {
#pragma unroll
        for (int k = 0; k < KE; ++k) x1[k] = ok[k] ? src1[g[k]] : T(0);
      }
#pragma unroll
      for (int k = 0; k < KE; ++k) {
        const int f = tid + NT * k;
        const int r = f / MP, c = f - r * MP;
        if (f < MP * MP) {
          bufA[r * PB + c] = x0[k];
          if (two) bufB[r * PB + c] = x1[k];
        }
      }
      __syncthreads();
      add_term(t, bufA);
      if (two) add_term(t + 1, bufB);
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int ct = S::tile(w, u);
      if (ct >= 0) {
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lds_put<T>(bufA, (jt * 16 + MfmaOp<T>::row(lane, r)) * PB + col, sum[u][r]);
      }
    }
    __syncthreads();
    // ---- q_{s+1} = E_s (.) -> history; times E_{s+1} for the next step ----
    // (the registers of E_s then hold q_{s+1}, stored after the loop: a load
    // of the block junction's E between two history stores made each wait
    // for the stores before it)
    const bool junction = next && ch.e_next != ch.e_fin;
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      if (f >= MP * MP) continue;
      T o = T(0);
      if (ok[k]) {
        o = bufA[r * PB + c] * ef[k];
        const T en = junction ? ch.e_next[g[k]] : ef[k];
        ef[k] = o;
        if (next) o *= en;
      }
      if (next) bufA[r * PB + c] = o;
    }
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      if (f < MP * MP && ok[k] && (!HALF || (f / MP >= MP / 2) == (h != 0)))
        __builtin_nontemporal_store(ef[k], &ch.hist[g[k]]);
    }
    if (!next) return;
    __syncthreads();
  } else {
    stage(ch.in, true);
    __syncthreads();
  }
  if constexpr (HALF) {
    // ---- start step s + 1 for rows [MP/2 h, MP/2 (h + 1)): per term, wave w
    // computes row tile rt and column tile ct of the row factor, then of the
    // z factor ----
    const int rt = (JT / 2) * h + w % (JT / 2), ct = w / (JT / 2);
    for (int t = 0; t < nt; ++t) {
      SlabOperands<T, MP, false> orow, ocol;
      slab_load_operands<T, MP, false>(p.mat_r[t], p.lo_r[t], MR, rt, lane, orow);
      slab_load_operands<T, MP, false>(p.mat_c[t], p.lo_c[t], MC, ct, lane, ocol);
      const V4 a = mfma_tile_b<T, MP, false>(orow, SPLIT && p.lo_r[t] != nullptr, bufA, PB,
                                             ct * 16 + (lane & 15));
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lds_put<T>(bufB, (rt * 16 + MfmaOp<T>::row(lane, r)) * PA + ct * 16 + (lane & 15), a[r]);
      __syncthreads();
      const V4 z = mfma_tile_a<T, MP, false>(ocol, SPLIT && p.lo_c[t] != nullptr, bufB, PA,
                                             rt * 16 + (lane & 15));
      // the tile straight from registers: 16 contiguous elements per row
      T* out = ch.out + t * p.tstride;
      const int c = ct * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + MfmaOp<T>::row(lane, r);
        if (row < MR && c < MC) {
          const long gi = base + row * RS + c;
          if constexpr (SPLIT) out[gi] = z[r];
          else __builtin_nontemporal_store(z[r], &out[gi]);
        }
      }
      __syncthreads();  // bufB is rewritten by the next term
    }
    return;
  }
  // ---- start step s + 1: per term, the row factor then the z factor ----
  for (int t = 0; t < nt; ++t) {
    SlabOperands<T, MP, false> orow, ocol;
    slab_load_operands<T, MP, false>(p.mat_r[t], p.lo_r[t], MR, jt, lane, orow);
    slab_load_operands<T, MP, false>(p.mat_c[t], p.lo_c[t], MC, jt, lane, ocol);
    phase_row_product<T, MP, false, NT>(orow, SPLIT && p.lo_r[t] != nullptr, bufA, w, lane,
                                        acc);
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int ct = S::tile(w, u);
      if (ct >= 0) {
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lds_put<T>(bufB, (jt * 16 + MfmaOp<T>::row(lane, r)) * PA + col, acc[u][r]);
      }
    }
    __syncthreads();
    phase_col_product<T, MP, false, NT>(ocol, SPLIT && p.lo_c[t] != nullptr, bufB, w, lane, acc);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int rt = S::tile(w, u);
      if (rt >= 0) {
        const int col = jt * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lds_put<T>(bufB, (rt * 16 + MfmaOp<T>::row(lane, r)) * PB + col, acc[u][r]);
      }
    }
    __syncthreads();
    T* out = ch.out + t * p.tstride;
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int f = tid + NT * k;
      const int r = f / MP, c = f - r * MP;
      if (ok[k]) {
        if constexpr (SPLIT) out[g[k]] = bufB[r * PB + c];
        else __builtin_nontemporal_store(bufB[r * PB + c], &out[g[k]]);
      }
    }
    __syncthreads();  // bufB is rewritten by the next term
  }
}

bool phase_half_enabled();

// Launches one multi-term phase for 1 or 2 chains (MR, MC <= mp).
template <typename T>
hipError_t launch_phase_terms(const PhaseTermsParams<T>& p, int nplanes, int nchains, int mp,
                              hipStream_t s) {
  if (nchains < 1 || nchains > 2 || nplanes < 1 || p.nt < 1 || p.nt > kMaxGlideTerms ||
      p.MR > mp || p.MC > mp)
    return hipErrorInvalidValue;
  dim3 grid((unsigned)nplanes, (unsigned)nchains, 1);
  switch (mp) {
    case 16: hipLaunchKernelGGL((plane_phase_terms_kernel<T, 16, 1024>), grid, dim3(1024), 0, s, p); break;
    case 32: hipLaunchKernelGGL((plane_phase_terms_kernel<T, 32, 1024>), grid, dim3(1024), 0, s, p); break;
    case 48: hipLaunchKernelGGL((plane_phase_terms_kernel<T, 48, 512>), grid, dim3(512), 0, s, p); break;
    case 64:
      // two workgroups per plane where the planes pair up on the XCDs
      if (nplanes % 8 == 0 && phase_half_enabled()) {
        dim3 g2((unsigned)(2 * nplanes), (unsigned)nchains, 1);
        hipLaunchKernelGGL((plane_phase_terms_kernel<T, 64, 512, true>), g2, dim3(512), 0, s, p);
      } else {
        hipLaunchKernelGGL((plane_phase_terms_kernel<T, 64, 512>), grid, dim3(512), 0, s, p);
      }
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

bool phase_reuse_enabled();

// Whether the phase path applies: all three axes in one length class <= 64.
inline int phase_class(int X, int Y, int Z) {
  const int c = pass_class(X);
  if (c == 0 || c > 64 || pass_class(Y) != c || pass_class(Z) != c) return 0;
  return c;
}

bool phase_enabled();
bool centro_enabled();

// Launches one phase for 1 or 2 chains.  `cen` selects the centrosymmetric
// kernel (then the caller passes the half-size operators and MR == MC == mp,
// a multiple of 32).  The caller fills p.c[], the operators and
// p.MR/MC/strides; `nplanes` is the number of slabs.  Workgroups have 1024
// threads (16 waves, one 16 x 16 tile each per product at class 64), except
// the dense class-48/64 kernels, whose operator registers fit 512 threads.
template <typename T>
hipError_t launch_phase(const PhaseParams<T>& p, int nplanes, int nchains, int mp, bool cen,
                        hipStream_t s) {
  if (nchains < 1 || nchains > 2 || nplanes < 1) return hipErrorInvalidValue;
  dim3 grid((unsigned)nplanes, (unsigned)nchains, 1);
  if (cen) {
    if (p.MR != mp || p.MC != mp) return hipErrorInvalidValue;
    // two workgroups per plane: all 256 CUs for 2 chains x 64 planes
    if (mp == 64 && nplanes % 8 == 0 &&
        phase_half_enabled()) {
      dim3 g2((unsigned)(2 * nplanes), (unsigned)nchains, 1);
      // (both chains finish a step in every phase but the start phase)
      const bool fin = p.c[0].hist != nullptr;
      if (nchains == 2 && (p.c[1].hist != nullptr) != fin) return hipErrorInvalidValue;
      if (phase_reuse_enabled()) {
        if (fin) hipLaunchKernelGGL((plane_phase_half_kernel<T, true>), g2, dim3(1024), 0, s, p);
        else hipLaunchKernelGGL((plane_phase_half_kernel<T, false>), g2, dim3(1024), 0, s, p);
      } else {
        if (fin)
          hipLaunchKernelGGL((plane_phase_half_khalf_kernel<T, true>), g2, dim3(1024), 0, s, p);
        else
          hipLaunchKernelGGL((plane_phase_half_khalf_kernel<T, false>), g2, dim3(1024), 0, s, p);
      }
      return hipGetLastError();
    }
    if (mp == 32)
      hipLaunchKernelGGL((plane_phase_kernel<T, 32, true, 1024>), grid, dim3(1024), 0, s, p);
    else if (mp == 64)
      hipLaunchKernelGGL((plane_phase_kernel<T, 64, true, 1024>), grid, dim3(1024), 0, s, p);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
  switch (mp) {
    case 16: hipLaunchKernelGGL((plane_phase_kernel<T, 16, false, 1024>), grid, dim3(1024), 0, s, p); break;
    case 32: hipLaunchKernelGGL((plane_phase_kernel<T, 32, false, 1024>), grid, dim3(1024), 0, s, p); break;
    case 48: hipLaunchKernelGGL((plane_phase_kernel<T, 48, false, 512>), grid, dim3(512), 0, s, p); break;
    case 64: hipLaunchKernelGGL((plane_phase_kernel<T, 64, false, 512>), grid, dim3(512), 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace crys
