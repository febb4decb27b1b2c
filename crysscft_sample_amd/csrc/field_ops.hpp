// Pointwise and reduction kernels of the SCFT iteration
// (crysSCFT_sample.cpp:294-371, :398-424, :476-508) on the device.
//
// Reductions are deterministic: a fixed grid, a fixed per-thread order and a
// fixed LDS tree, then a single-workgroup pass over the per-block partials.
#pragma once

#include <hip/hip_runtime.h>

namespace crys {

constexpr int kRedThreads = 256;
constexpr int kMaxRedBlocks = 512;

inline int red_blocks(size_t n) {
  size_t b = (n + kRedThreads - 1) / kRedThreads;
  return (int)(b < (size_t)kMaxRedBlocks ? (b ? b : 1) : kMaxRedBlocks);
}

template <typename T>
__global__ void fill_kernel(T* p, T v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T>
__global__ void copy_kernel(T* dst, const T* src, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

template <typename D, typename S>
__global__ void convert_kernel(D* dst, const S* src, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (D)src[i];
}

// calc_ksi (:341-346) and array_exp (:315-322) for both blocks.
template <typename T>
__global__ void prep_fields_kernel(const double* wA, const double* wB, double* ksi, T* EA, T* EB,
                                   double xN, double ds, size_t n) {
  const double index = -ds / 2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double a = wA[i], b = wB[i];
    ksi[i] = 0.5 * (a + b - xN);
    EA[i] = (T)(1. * exp(index * a));
    EB[i] = (T)(1. * exp(index * b));
  }
}

// Block reductions over kRedThreads threads: a butterfly inside each wave,
// then the four wave results through LDS, two barriers in all (the LDS tree
// over 256 values took nine).  Every thread gets the same value, and the
// order of the additions is fixed, so results repeat bitwise.
static_assert(kRedThreads == 256, "four waves per reduction block");

__device__ inline double block_sum(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

__device__ inline double block_max(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o);
    v = u > v ? u : v;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double a = sh[0] > sh[1] ? sh[0] : sh[1], b = sh[2] > sh[3] ? sh[2] : sh[3];
  __syncthreads();
  return a > b ? a : b;
}

// Partial sums of q(s=0) of the backward chain for Q = <q_A^dagger(0)>
// (array_average, :324-331, called at :494).
template <typename T>
__global__ __launch_bounds__(kRedThreads) void qsum_partial_kernel(const T* q0, double* part,
                                                                   size_t n) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kRedThreads)
    acc += (double)q0[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// scal[0] = Q = sum(part)/n, scal[1] = fA/(NsA Q), scal[2] = fB/(NsB Q).
static __global__ __launch_bounds__(kRedThreads) void q_finalize_kernel(const double* part, int nparts,
                                                                 double* scal, size_t n, double fA,
                                                                 double fB, size_t NsA,
                                                                 size_t NsB) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kRedThreads) acc += part[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    const double Q = s / n;
    scal[0] = Q;
    scal[1] = fA / (NsA * Q);
    scal[2] = fB / (NsB * Q);
  }
}

// integrate_phi's contour sums (:348-360) over the slots s0 .. s0+cnt-1 of
// one checkpoint segment: qf and qb hold those slots contiguously.  The
// trapezoid weights are those of the whole contour (slot 0, NsA and Ns weigh
// 1/2 in the blocks that end there).  first != 0 starts the sums, else they
// are added to accA / accB; every slot is summed in exactly one segment.
template <typename T>
__global__ __launch_bounds__(kRedThreads) void phi_accumulate_kernel(
    const T* qf, const T* qb, size_t n, size_t s0, size_t cnt, size_t NsA, size_t Ns,
    double* accA, double* accB, int first) {
  for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kRedThreads) {
    double a = 0.0, b = 0.0;
    for (size_t k = 0; k < cnt; ++k) {
      const size_t t = s0 + k;
      // each slot is read once (non-temporal: 1% faster checkpointed FP64)
      const double p = (double)__builtin_nontemporal_load(&qf[k * n + i]) *
                       (double)__builtin_nontemporal_load(&qb[k * n + i]);
      if (t <= NsA) a += (t == 0 || t == NsA) ? 0.5 * p : p;
      if (t >= NsA) b += (t == NsA || t == Ns) ? 0.5 * p : p;
    }
    accA[i] = first ? a : accA[i] + a;
    accB[i] = first ? b : accB[i] + b;
  }
}

// The fields of one SCFT instance that the density/energy/mixing work reads
// and writes (see phi_energy_mix_points).
template <typename T>
struct PhiFields {
  const T* qf;
  const T* qb;
  double *phiA, *phiB, *wA, *wB, *ksi;
  T *EA, *EB;
  size_t n, NsA, NsB;
  double ca, cb, xN, acc_mix, index;
};

// integrate_phi (:348-360) for both blocks from the joined histories
// (forward slots 0..NsA are q_A, NsA..NsA+NsB are q_B; backward slots the
// same for q^dagger), then print_energy's sums (:404-418) with the
// pre-update fields, calc_new_fields (:362-371), and the next iteration's
// calc_ksi / array_exp (:478-480), for the points of thread t0 of nthr.
// With FROM_SUMS, phiA / phiB hold the contour sums of phi_accumulate_kernel
// on entry and qf / qb are not read.
template <typename T, bool FROM_SUMS>
__device__ __forceinline__ void phi_energy_mix_points(const PhiFields<T>& F, size_t t0, size_t nthr,
                                                      double& s_e1, double& s_enth,
                                                      double& s_fk, double& m_inc) {
  const size_t n = F.n;
  // the field work of point i from its two contour sums
  auto finish = [&](size_t i, double sA, double sB) {
    const double pA = F.ca * (FROM_SUMS ? F.phiA[i] : sA);
    const double pB = F.cb * (FROM_SUMS ? F.phiB[i] : sB);
    const double a = F.wA[i], b = F.wB[i], x = F.ksi[i];
    s_e1 += -a * pA - b * pB;
    s_enth += F.xN * pA * pB;
    const double t = 1 - pA - pB;
    s_fk += -x * t;
    m_inc = fabs(t) > m_inc ? fabs(t) : m_inc;
    F.phiA[i] = pA;
    F.phiB[i] = pB;
    const double na = a + F.acc_mix * (F.xN * pB + x - a);
    const double nb = b + F.acc_mix * (F.xN * pA + x - b);
    F.wA[i] = na;
    F.wB[i] = nb;
    F.ksi[i] = 0.5 * (na + nb - F.xN);
    F.EA[i] = (T)(1. * exp(F.index * na));
    F.EB[i] = (T)(1. * exp(F.index * nb));
  };
  // sum_s w_s q[s] q+[s] over slots s0..s1 (trapezoid: the end slots weigh
  // 1/2) for V consecutive points, one 16-byte load per slot and chain; R
  // slots per round with independent partial sums, so that 2 R loads are in
  // flight per thread.  Each history entry is read once here, non-temporally:
  // the FP64 iteration at 128^3 measured 1.7% faster than with ordinary
  // loads.  16-byte loads of V points measured faster than one point per
  // thread (SCFT iteration at 128^3: FP64 0.6%, FP32 1.2%).
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vT __attribute__((ext_vector_type(V)));
  const T* const qf = F.qf;
  const T* const qb = F.qb;
  auto contour_sums = [&](size_t i0, size_t s0, size_t s1, double (&out)[V]) {
    constexpr int R = 4;
    double p[R][V];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int v = 0; v < V; ++v) p[r][v] = 0.0;
    size_t s = s0 + 1;
    for (; s + R <= s1; s += R) {
      vT a[R], b[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        a[r] = __builtin_nontemporal_load((const vT*)&qf[(s + r) * n + i0]);
        b[r] = __builtin_nontemporal_load((const vT*)&qb[(s + r) * n + i0]);
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v) p[r][v] += (double)a[r][v] * (double)b[r][v];
    }
    for (; s < s1; ++s) {
      const vT a = *(const vT*)&qf[s * n + i0], b = *(const vT*)&qb[s * n + i0];
#pragma unroll
      for (int v = 0; v < V; ++v) p[0][v] += (double)a[v] * (double)b[v];
    }
    const vT a0 = *(const vT*)&qf[s0 * n + i0], b0 = *(const vT*)&qb[s0 * n + i0];
    const vT a1 = *(const vT*)&qf[s1 * n + i0], b1 = *(const vT*)&qb[s1 * n + i0];
#pragma unroll
    for (int v = 0; v < V; ++v)
      out[v] = ((p[0][v] + p[1][v]) + (p[2][v] + p[3][v])) +
               0.5 * ((double)a0[v] * (double)b0[v] + (double)a1[v] * (double)b1[v]);
  };
  const size_t nv = n - n % V;  // (the grid rules make n a multiple of 8)
  for (size_t i0 = t0 * V; i0 < nv; i0 += nthr * V) {
    double sA[V], sB[V];
    if constexpr (!FROM_SUMS) {
      contour_sums(i0, 0, F.NsA, sA);
      contour_sums(i0, F.NsA, F.NsA + F.NsB, sB);
    }
#pragma unroll
    for (int v = 0; v < V; ++v) finish(i0 + v, FROM_SUMS ? 0.0 : sA[v], FROM_SUMS ? 0.0 : sB[v]);
  }
  // the last n % V points one by one
  for (size_t i = nv + t0; i < n; i += nthr) {
    double sA = 0.0, sB = 0.0;
    if constexpr (!FROM_SUMS) {
      auto sum1 = [&](size_t s0, size_t s1) {
        double t = 0.5 * ((double)qf[s0 * n + i] * (double)qb[s0 * n + i] +
                          (double)qf[s1 * n + i] * (double)qb[s1 * n + i]);
        for (size_t s = s0 + 1; s < s1; ++s) t += (double)qf[s * n + i] * (double)qb[s * n + i];
        return t;
      };
      sA = sum1(0, F.NsA);
      sB = sum1(F.NsA, F.NsA + F.NsB);
    }
    finish(i, sA, sB);
  }
}

// phi_energy_mix_points over one instance, then the block's partial sums of
// the energy terms and the incompressibility maximum.
template <typename T, bool FROM_SUMS = false>
__global__ __launch_bounds__(kRedThreads) void phi_energy_mix_kernel(
    const T* qf, const T* qb, const double* scal, double* phiA, double* phiB, double* wA,
    double* wB, double* ksi, T* EA, T* EB, double* part, size_t n, size_t NsA, size_t NsB,
    double xN, double acc_mix, double ds) {
  __shared__ double sh[kRedThreads];
  const PhiFields<T> F{qf, qb, phiA, phiB, wA, wB, ksi, EA, EB, n, NsA, NsB,
                       scal[1], scal[2], xN, acc_mix, -ds / 2};
  double s_e1 = 0.0, s_enth = 0.0, s_fk = 0.0, m_inc = 0.0;
  phi_energy_mix_points<T, FROM_SUMS>(F, blockIdx.x * (size_t)kRedThreads + threadIdx.x,
                                      (size_t)gridDim.x * kRedThreads, s_e1, s_enth, s_fk, m_inc);
  const double r0 = block_sum(s_e1, sh);
  const double r1 = block_sum(s_enth, sh);
  const double r2 = block_sum(s_fk, sh);
  const double r3 = block_max(m_inc, sh);
  if (threadIdx.x == 0) {
    part[0 * kMaxRedBlocks + blockIdx.x] = r0;
    part[1 * kMaxRedBlocks + blockIdx.x] = r1;
    part[2 * kMaxRedBlocks + blockIdx.x] = r2;
    part[3 * kMaxRedBlocks + blockIdx.x] = r3;
  }
}

// Sums the partials and writes one trace row at row *counter (mod cap).
static __global__ __launch_bounds__(kRedThreads) void energy_finalize_kernel(const double* part,
                                                                      int nparts,
                                                                      const double* scal,
                                                                      double* trace, int* counter,
                                                                      int cap, size_t n) {
  __shared__ double sh[kRedThreads];
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int i = threadIdx.x; i < nparts; i += kRedThreads) {
    a0 += part[0 * kMaxRedBlocks + i];
    a1 += part[1 * kMaxRedBlocks + i];
    a2 += part[2 * kMaxRedBlocks + i];
    const double m = part[3 * kMaxRedBlocks + i];
    a3 = m > a3 ? m : a3;
  }
  const double e1 = block_sum(a0, sh);
  const double enth = block_sum(a1, sh);
  const double fk = block_sum(a2, sh);
  const double mi = block_max(a3, sh);
  if (threadIdx.x == 0) {
    const double Q = scal[0];
    const double e0 = -log(Q);
    const double E1 = e1 / n, EN = enth / n, FK = fk / n;
    const int row = *counter % cap;
    double* r = trace + (size_t)row * 7;
    r[0] = e0 + E1 + EN + FK;
    r[1] = EN;
    r[2] = e0;
    r[3] = E1;
    r[4] = FK;
    r[5] = mi;
    r[6] = Q;
    *counter = *counter + 1;
  }
}

// Partial sums of u * E * v (the stress inner products).
template <typename T>
__global__ __launch_bounds__(kRedThreads) void dot3_partial_kernel(const T* u, const T* E,
                                                                   const T* v, double* part,
                                                                   size_t n) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kRedThreads)
    acc += (double)u[i] * ((double)E[i] * (double)v[i]);
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Gradient of the field-theoretic free energy
//   H[wA, wB] = (1/n) sum_r [ (wA-wB)^2 / (4 xN) - (wA+wB)/2 + xN/4 ] - ln Q
// from the densities of the same fields: d(-ln Q)/dwA(r) = cA phiA(r) / n with
// cA = ds NsA / fA (and B alike).  Partials of the field part go to part.
static __global__ __launch_bounds__(kRedThreads) void hgrad_kernel(const double* wA, const double* wB,
                                                            const double* phiA,
                                                            const double* phiB, double* gA,
                                                            double* gB, double* part, size_t n,
                                                            double xN, double cA, double cB) {
  __shared__ double sh[kRedThreads];
  const double inv_n = 1.0 / (double)n;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kRedThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kRedThreads) {
    const double a = wA[i], b = wB[i], d = a - b;
    acc += d * d / (4 * xN) - (a + b) / 2 + xN / 4;
    gA[i] = (d / (2 * xN) - 0.5 + cA * phiA[i]) * inv_n;
    gB[i] = (-d / (2 * xN) - 0.5 + cB * phiB[i]) * inv_n;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[k] = sum of part[k * nparts + b] over b, one workgroup per k.
static __global__ __launch_bounds__(kRedThreads) void sum_partials_kernel(const double* part, int nparts,
                                                                   double* out) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  const double* p = part + (size_t)blockIdx.x * nparts;
  for (int i = threadIdx.x; i < nparts; i += kRedThreads) acc += p[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

static __global__ void set_int_kernel(int* p, int v) { *p = v; }

}  // namespace crys
