// Many independent SCFT instances advanced together (include/crys_solver.h,
// crys_batch_*).
//
// Each instance is one diblock_scft_simple_iterate (crysSCFT_sample.cpp:426-
// 531) with its own grid, chemistry, cell and contour.  When every instance
// has its three axes in one length class (16/32/48/64), a contour step of
// the whole batch is one launch of batch_phase_kernel: the alternating
// x-slab / y-slab phases of plane_phase.hpp, with blockIdx.x indexing a work
// list of (instance, chain, plane) items, provided that list holds at most
// 1792 items (small batches).  Otherwise a step is three launches
// of batch_pass_kernel, one per axis, over (instance, chain, tile) items.
// Either list is ordered by decreasing contour length so that the instances
// still running at step t are a prefix of it, and each item's length class
// selects the matrix-core body at run time.  One SCFT iteration of the whole
// batch is captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "axis_pass.hpp"
#include "crys_solver.h"
#include "field_ops.hpp"
#include "operators.hpp"
#include "plane_phase.hpp"

namespace crys {
int set_error(int code, const char* msg);
}

namespace {

constexpr int kCap = 256;  // trace rows kept on the device between copies
// Above this many (instance, chain, plane) items per phase the batch runs
// the per-axis passes instead of the one-launch phases; CRYS_BATCH_PHASE_ITEMS
// overrides it (for A/B timing).
size_t phase_batch_items() {
  static const size_t lim = [] {
    const char* e = getenv("CRYS_BATCH_PHASE_ITEMS");
    return e ? (size_t)strtoull(e, nullptr, 10) : (size_t)1792;
  }();
  return lim;
}

// CRYS_BATCH_PERSISTENT=0 runs the batch phases with one workgroup per item
// instead of the persistent grid (for A/B timing).
bool batch_persistent_enabled() {
  static const bool on = [] {
    const char* e = getenv("CRYS_BATCH_PERSISTENT");
    return !(e && strcmp(e, "0") == 0);
  }();
  return on;
}

int errf(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return crys::set_error(code, buf);
}

#define BHIP(expr)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return errf(e_ == hipErrorOutOfMemory ? CRYS_ERR_OUT_OF_MEMORY : CRYS_ERR_HIP,       \
                  "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define BCHK(expr)                \
  do {                            \
    int r_ = (expr);              \
    if (r_ != CRYS_OK) return r_; \
  } while (0)

}  // namespace

namespace crys {

// Device view of one instance.  Pointers are element pointers into the
// batch's shared allocations.
template <typename T>
struct BatchInst {
  int X, Y, Z, cls[3];
  int NsA, NsB, Ns, nblk;  // nblk: reduction workgroups of this instance
  long n;
  const T* mat[3];
  const unsigned short* mat_lo[3];  // FP32: bf16 low parts of the operators
  const T* mat_c[3];                // half-size centrosymmetric operators (cen)
  const unsigned short* mat_c_lo[3];
  int pcls, cen;                    // phase length class (0: none), centrosymmetric
  int shift;                        // 3m: half-period pairing (cen_partner)
  T *qf, *qb, *scr0, *scr1, *EA, *EB;
  T *scr0b, *scr1b;                 // second scratch field per chain (phases)
  double *wA, *wB, *ksi, *phiA, *phiB;
  double fA, fB, xN, acc, ds;
};

// LDS elements for the largest class (MP = 64 or 48/32/16).
constexpr int kBatchLds = 64 * MfmaShape<64>::LDT > 32 * MfmaShape<32>::LDT
                              ? 64 * MfmaShape<64>::LDT
                              : 32 * MfmaShape<32>::LDT;

template <typename T, int AXIS>
__global__ __launch_bounds__(256) void batch_pass_kernel(const BatchInst<T>* __restrict__ insts,
                                                         const int2* __restrict__ work, int t) {
  __shared__ T lds[kBatchLds];
  const int2 w = work[blockIdx.x];
  const BatchInst<T>& I = insts[w.x];
  const int chain = w.y >> 16, tidx = w.y & 0xffff;
  const long n = I.n;
  T* q = chain == 0 ? I.qf : I.qb;
  T* scr = chain == 0 ? I.scr0 : I.scr1;
  const int s_in = chain == 0 ? t : I.Ns - t;
  const int s_out = chain == 0 ? t + 1 : I.Ns - t - 1;
  const T* E = chain == 0 ? (t < I.NsA ? I.EA : I.EB) : (s_out >= I.NsA ? I.EB : I.EA);
  PassChain<T> ch;
  if constexpr (AXIS == 2) ch = {q + s_in * n, scr, E, nullptr};
  else if constexpr (AXIS == 1) ch = {scr, scr, nullptr, nullptr};
  else ch = {scr, q + s_out * n, nullptr, E};
  if (I.mat_c[AXIS]) {  // half-size product (the axis length is its class, 32 or 64)
    const T* mc = I.mat_c[AXIS];
    const unsigned short* lc = I.mat_c_lo[AXIS];
    const bool sh = I.shift != 0;
    if (I.cls[AXIS] == 32)
      mfma_pass_tile<T, AXIS, 32, true>(ch, mc, lc, I.X, I.Y, I.Z, tidx, lds, sh);
    else
      mfma_pass_tile<T, AXIS, 64, true>(ch, mc, lc, I.X, I.Y, I.Z, tidx, lds, sh);
    return;
  }
  const T* mat = I.mat[AXIS];
  const unsigned short* lo = I.mat_lo[AXIS];
  switch (I.cls[AXIS]) {
    case 16: mfma_pass_tile<T, AXIS, 16>(ch, mat, lo, I.X, I.Y, I.Z, tidx, lds); break;
    case 32: mfma_pass_tile<T, AXIS, 32>(ch, mat, lo, I.X, I.Y, I.Z, tidx, lds); break;
    case 48: mfma_pass_tile<T, AXIS, 48>(ch, mat, lo, I.X, I.Y, I.Z, tidx, lds); break;
    default: mfma_pass_tile<T, AXIS, 64>(ch, mat, lo, I.X, I.Y, I.Z, tidx, lds); break;
  }
}

// Phase k of the batch (see plane_phase.hpp): work item (instance, chain
// << 16 | plane); `yslab` selects y-slab phases (odd k).  Forward chain q_f[t]
// -> q_f[t+1] with E_A for t < NsA; backward chain q_b[Ns-t] -> q_b[Ns-t-1]
// with E_B while Ns-t-1 >= NsA; phase k writes scratch field k % 2.
//
// NT = 1024 threads hold every body but the dense class-64 one within 128
// VGPRs (16 waves per CU); a batch with dense class-64 planes (reduced
// lengths 49..63) runs the NT = 512 kernel, which holds all bodies.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void batch_phase_kernel(const BatchInst<T>* __restrict__ insts,
                                                         const int2* __restrict__ work, int k,
                                                         int yslab) {
  __shared__ T lds[64 * PhaseShape<T, 64, false, NT>::PB];
  const int2 w = work[blockIdx.x];
  const BatchInst<T>& I = insts[w.x];
  const int chain = w.y >> 16, plane = w.y & 0xffff;
  const long n = I.n;
  const int Ns = I.Ns, NsA = I.NsA;
  T* q = chain == 0 ? I.qf : I.qb;
  T* half[2] = {chain == 0 ? I.scr0 : I.scr1, chain == 0 ? I.scr0b : I.scr1b};
  auto E = [&](int t) -> const T* {
    return chain == 0 ? (t < NsA ? I.EA : I.EB) : (Ns - t - 1 >= NsA ? I.EB : I.EA);
  };
  PhaseChain<T> ch{};
  ch.in = k == 0 ? (chain == 0 ? q : q + (long)Ns * n) : half[(k + 1) % 2];
  if (k >= 1) {
    ch.hist = chain == 0 ? q + (long)k * n : q + (long)(Ns - k) * n;
    ch.e_fin = E(k - 1);
  }
  if (k < Ns) {
    ch.out = half[k % 2];
    ch.e_next = E(k);
  }
  const int ra = yslab ? 0 : 1;
  const int MR = yslab ? I.X : I.Y, MC = I.Z;
  const long base = yslab ? (long)plane * I.Z : (long)plane * I.Y * I.Z;
  const long RS = yslab ? (long)I.Y * I.Z : (long)I.Z;
  if (I.cen) {
    const T *mr = I.mat_c[ra], *mc = I.mat_c[2];
    const unsigned short *lr = I.mat_c_lo[ra], *lc = I.mat_c_lo[2];
    const bool sh = I.shift != 0;
    if (I.pcls == 32)
      plane_phase_body<T, 32, true, NT>(ch, mr, mc, lr, lc, MR, MC, base, RS, lds, sh);
    else
      plane_phase_body<T, 64, true, NT>(ch, mr, mc, lr, lc, MR, MC, base, RS, lds, sh);
    return;
  }
  const T *mr = I.mat[ra], *mc = I.mat[2];
  const unsigned short *lr = I.mat_lo[ra], *lc = I.mat_lo[2];
  switch (I.pcls) {
    case 16: plane_phase_body<T, 16, false, NT>(ch, mr, mc, lr, lc, MR, MC, base, RS, lds); break;
    case 32: plane_phase_body<T, 32, false, NT>(ch, mr, mc, lr, lc, MR, MC, base, RS, lds); break;
    case 48: plane_phase_body<T, 48, false, NT>(ch, mr, mc, lr, lc, MR, MC, base, RS, lds); break;
    default:
      // the host runs the NT = 512 kernel whenever a dense class-64 plane exists
      if constexpr (NT == 512)
        plane_phase_body<T, 64, false, NT>(ch, mr, mc, lr, lc, MR, MC, base, RS, lds);
      break;
  }
}

// Copies one plane (rows of MC elements at stride RS from in + base) into
// `raw` as compact rows (element (r, c) at r * MC + c) with gfx950's direct
// global-to-LDS loads: 16 bytes per lane, each wave's load one contiguous
// 1 KB block of LDS, no registers.  Lanes past the plane's end reload its
// first element into LDS slots past the plane (raw holds kRawElems, a whole
// number of blocks).  The copy is complete once every wave has waited for its
// vector memory counter and the workgroup has passed a barrier.
constexpr int kRawElems = 64 * 64;

template <typename T, int NT>
__device__ __forceinline__ void dma_plane(const T* in, long base, long RS, int MR, int MC,
                                          T* raw) {
  constexpr int PER = 16 / (int)sizeof(T);
  constexpr int BLK = 64 * PER;
  static_assert(kRawElems % BLK == 0, "raw holds whole blocks");
  const int total = MR * MC;
  const int blocks = (total + BLK - 1) / BLK;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int j = w; j < blocks; j += NT / 64) {
    int e = j * BLK + lane * PER;
    if (e >= total) e = 0;
    const int r = e / MC, c = e - r * MC;
    __builtin_amdgcn_global_load_lds(in + base + r * RS + c, raw + j * BLK, 16, 0, 0);
  }
}

// Geometry of phase item `item` (instance, chain << 16 | plane) of phase k.
template <typename T>
struct PhaseItem {
  const BatchInst<T>* I;
  PhaseChain<T> ch;
  int ra, MR, MC;
  long base, RS;
};

template <typename T>
__device__ __forceinline__ PhaseItem<T> phase_item(const BatchInst<T>* __restrict__ insts,
                                                   const int2* __restrict__ work, int item, int k,
                                                   int yslab) {
  PhaseItem<T> it;
  const int2 w = work[item];
  const BatchInst<T>& I = insts[w.x];
  it.I = &I;
  const int chain = w.y >> 16, plane = w.y & 0xffff;
  const long n = I.n;
  const int Ns = I.Ns, NsA = I.NsA;
  T* q = chain == 0 ? I.qf : I.qb;
  T* half[2] = {chain == 0 ? I.scr0 : I.scr1, chain == 0 ? I.scr0b : I.scr1b};
  auto E = [&](int t) -> const T* {
    return chain == 0 ? (t < NsA ? I.EA : I.EB) : (Ns - t - 1 >= NsA ? I.EB : I.EA);
  };
  it.ch = PhaseChain<T>{};
  it.ch.in = k == 0 ? (chain == 0 ? q : q + (long)Ns * n) : half[(k + 1) % 2];
  if (k >= 1) {
    it.ch.hist = chain == 0 ? q + (long)k * n : q + (long)(Ns - k) * n;
    it.ch.e_fin = E(k - 1);
  }
  if (k < Ns) {
    it.ch.out = half[k % 2];
    it.ch.e_next = E(k);
  }
  it.ra = yslab ? 0 : 1;
  it.MR = yslab ? I.X : I.Y;
  it.MC = I.Z;
  it.base = yslab ? (long)plane * I.Z : (long)plane * I.Y * I.Z;
  it.RS = yslab ? (long)I.Y * I.Z : (long)I.Z;
  return it;
}

// Phase k of the batch on a persistent grid: workgroup g takes the items
// g, g + G, g + 2 G, ... of the phase's work list (G = gridDim.x, a few per
// CU).  Each item's plane is copied into LDS (dma_plane) while the previous
// item is being computed, so that the plane's load latency, exposed once per
// item when every item is a workgroup of its own, is hidden behind the
// products.  The items of one phase are independent, so no workgroup waits
// for another.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void batch_phase_persistent_kernel(
    const BatchInst<T>* __restrict__ insts, const int2* __restrict__ work, int cnt, int k,
    int yslab) {
  __shared__ T lds[64 * PhaseShape<T, 64, false, NT>::PB];
  __shared__ __align__(16) T raw[kRawElems];
  int item = blockIdx.x;
  if (item >= cnt) return;
  {
    const PhaseItem<T> f = phase_item<T>(insts, work, item, k, yslab);
    dma_plane<T, NT>(f.ch.in, f.base, f.RS, f.MR, f.MC, raw);
  }
  for (; item < cnt; item += gridDim.x) {
    // the plane of this item has landed in raw (every wave's copy), and the
    // previous item is done with lds
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const PhaseItem<T> it = phase_item<T>(insts, work, item, k, yslab);
    const BatchInst<T>& I = *it.I;
    const int nx = item + (int)gridDim.x;
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));  // opaque: recomputed per plane (see plane_phase_body)
    auto prefetch = [&]() {
      if (nx < cnt) {
        const PhaseItem<T> nt = phase_item<T>(insts, work, nx, k, yslab);
        dma_plane<T, NT>(nt.ch.in, nt.base, nt.RS, nt.MR, nt.MC, raw);
      }
    };
    if (I.cen) {
      const T *mr = I.mat_c[it.ra], *mc = I.mat_c[2];
      const unsigned short *lr = I.mat_c_lo[it.ra], *lc = I.mat_c_lo[2];
      const bool sh = I.shift != 0;
      if (I.pcls == 32)
        plane_phase_body<T, 32, true, NT, true>(it.ch, mr, mc, lr, lc, it.MR, it.MC, it.base,
                                                it.RS, lds, sh, raw, prefetch, tid);
      else
        plane_phase_body<T, 64, true, NT, true>(it.ch, mr, mc, lr, lc, it.MR, it.MC, it.base,
                                                it.RS, lds, sh, raw, prefetch, tid);
      continue;
    }
    const T *mr = I.mat[it.ra], *mc = I.mat[2];
    const unsigned short *lr = I.mat_lo[it.ra], *lc = I.mat_lo[2];
    switch (I.pcls) {
      case 16:
        plane_phase_body<T, 16, false, NT, true>(it.ch, mr, mc, lr, lc, it.MR, it.MC, it.base,
                                                 it.RS, lds, false, raw, prefetch, tid);
        break;
      case 32:
        plane_phase_body<T, 32, false, NT, true>(it.ch, mr, mc, lr, lc, it.MR, it.MC, it.base,
                                                 it.RS, lds, false, raw, prefetch, tid);
        break;
      case 48:
        plane_phase_body<T, 48, false, NT, true>(it.ch, mr, mc, lr, lc, it.MR, it.MC, it.base,
                                                 it.RS, lds, false, raw, prefetch, tid);
        break;
      default:
        // the host runs the NT = 512 kernel whenever a dense class-64 plane exists
        if constexpr (NT == 512)
          plane_phase_body<T, 64, false, NT, true>(it.ch, mr, mc, lr, lc, it.MR, it.MC, it.base,
                                                   it.RS, lds, false, raw, prefetch, tid);
        else
          prefetch();
        break;
    }
  }
}

// Q partials: work item (instance, block).
template <typename T>
__global__ __launch_bounds__(kRedThreads) void batch_qsum_kernel(const BatchInst<T>* insts,
                                                                 const int2* work, double* part) {
  __shared__ double sh[kRedThreads];
  const int2 w = work[blockIdx.x];
  const BatchInst<T>& I = insts[w.x];
  double acc = 0.0;
  for (long i = (long)w.y * kRedThreads + threadIdx.x; i < I.n; i += (long)I.nblk * kRedThreads)
    acc += (double)I.qb[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[(size_t)w.x * kMaxRedBlocks + w.y] = s;
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void batch_qfinal_kernel(const BatchInst<T>* insts,
                                                                   const double* part,
                                                                   double* scal) {
  __shared__ double sh[kRedThreads];
  const BatchInst<T>& I = insts[blockIdx.x];
  double acc = 0.0;
  for (int i = threadIdx.x; i < I.nblk; i += kRedThreads)
    acc += part[(size_t)blockIdx.x * kMaxRedBlocks + i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    const double Q = s / I.n;
    double* sc = scal + 4 * (size_t)blockIdx.x;
    sc[0] = Q;
    sc[1] = I.fA / (I.NsA * Q);
    sc[2] = I.fB / (I.NsB * Q);
  }
}

// integrate_phi, print_energy sums, calc_new_fields and the next calc_ksi /
// array_exp of one instance (crysSCFT_sample.cpp:348-371, :398-424, :478-480).
template <typename T>
__global__ __launch_bounds__(kRedThreads) void batch_phi_kernel(const BatchInst<T>* insts,
                                                                const int2* work,
                                                                const double* scal,
                                                                double* part) {
  __shared__ double sh[kRedThreads];
  const int2 w = work[blockIdx.x];
  const BatchInst<T>& I = insts[w.x];
  const PhiFields<T> F{I.qf, I.qb, I.phiA, I.phiB, I.wA, I.wB, I.ksi, I.EA, I.EB,
                       (size_t)I.n, (size_t)I.NsA, (size_t)I.NsB, scal[4 * (size_t)w.x + 1],
                       scal[4 * (size_t)w.x + 2], I.xN, I.acc, -I.ds / 2};
  double s_e1 = 0.0, s_enth = 0.0, s_fk = 0.0, m_inc = 0.0;
  // (the same per-point work as the single-instance kernel, over this
  // instance's blocks)
  phi_energy_mix_points<T, false>(F, (size_t)w.y * kRedThreads + threadIdx.x,
                                  (size_t)I.nblk * kRedThreads, s_e1, s_enth, s_fk, m_inc);
  const double r0 = block_sum(s_e1, sh);
  const double r1 = block_sum(s_enth, sh);
  const double r2 = block_sum(s_fk, sh);
  const double r3 = block_max(m_inc, sh);
  if (threadIdx.x == 0) {
    double* p = part + (size_t)w.x * 4 * kMaxRedBlocks;
    p[0 * kMaxRedBlocks + w.y] = r0;
    p[1 * kMaxRedBlocks + w.y] = r1;
    p[2 * kMaxRedBlocks + w.y] = r2;
    p[3 * kMaxRedBlocks + w.y] = r3;
  }
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void batch_efinal_kernel(const BatchInst<T>* insts,
                                                                   const double* part,
                                                                   const double* scal,
                                                                   double* trace,
                                                                   const int* counter) {
  __shared__ double sh[kRedThreads];
  const BatchInst<T>& I = insts[blockIdx.x];
  const double* p = part + (size_t)blockIdx.x * 4 * kMaxRedBlocks;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int i = threadIdx.x; i < I.nblk; i += kRedThreads) {
    a0 += p[0 * kMaxRedBlocks + i];
    a1 += p[1 * kMaxRedBlocks + i];
    a2 += p[2 * kMaxRedBlocks + i];
    a3 = p[3 * kMaxRedBlocks + i] > a3 ? p[3 * kMaxRedBlocks + i] : a3;
  }
  const double e1 = block_sum(a0, sh);
  const double enth = block_sum(a1, sh);
  const double fk = block_sum(a2, sh);
  const double mi = block_max(a3, sh);
  if (threadIdx.x == 0) {
    const double Q = scal[4 * (size_t)blockIdx.x];
    const double e0 = -log(Q);
    const double n = (double)I.n;
    const double E1 = e1 / n, EN = enth / n, FK = fk / n;
    double* r = trace + ((size_t)blockIdx.x * kCap + (*counter % kCap)) * 7;
    r[0] = e0 + E1 + EN + FK;
    r[1] = EN;
    r[2] = e0;
    r[3] = E1;
    r[4] = FK;
    r[5] = mi;
    r[6] = Q;
  }
}

__global__ void batch_count_kernel(int* counter, int reset) { *counter = reset ? 0 : *counter + 1; }

}  // namespace crys

struct crys_batch {
  int precision = CRYS_FP64;
  size_t count = 0;
  std::vector<crys_instance_t> desc;
  std::vector<size_t> n, NsA, NsB, off_f, off_q;  // per instance sizes and offsets
  std::vector<int> cls;                           // 3 per instance
  std::vector<size_t> off_m;                      // matrix offsets, 3 per instance
  std::vector<size_t> off_mc;                     // half-size matrix offsets (cen instances)
  bool phase_ok = false;                          // every instance can run the phases
  bool phase_dense64 = false;                     // a dense class-64 plane: 512-thread phases
  bool phases = false;                            // one launch per step (batch_phase_kernel)
  // phases on a persistent grid of this many workgroups
  // (batch_phase_persistent_kernel), or 0: one workgroup per item
  int phase_grid = 0;
  int n_cu = 0;                                   // compute units of the batch's device
  int2* pwork[2] = {nullptr, nullptr};            // phase items: x-slab, y-slab phases
  std::vector<int> pprefix[2];                    // phase items active at phase k
  hipStream_t stream = nullptr;
  void *dev = nullptr;  // one allocation for every buffer
  void* insts = nullptr;
  int2 *work[3] = {nullptr, nullptr, nullptr}, *rwork = nullptr;
  std::vector<int> prefix[3];  // work items active at step t, per axis
  int n_rwork = 0, Ns_max = 0;
  double *part_q = nullptr, *part_e = nullptr, *scal = nullptr, *trace = nullptr;
  int* counter = nullptr;
  std::vector<void*> allocs;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<char> fields_set;
  // typed base pointers
  void *qf = nullptr, *qb = nullptr, *scr = nullptr, *EA = nullptr, *EB = nullptr, *mats = nullptr;
  unsigned short* mats_lo = nullptr;  // FP32: bf16 low parts of the operators
  void* mats_c = nullptr;             // half-size centrosymmetric operators
  unsigned short* mats_c_lo = nullptr;
  double *wA = nullptr, *wB = nullptr, *ksi = nullptr, *phiA = nullptr, *phiB = nullptr;
};

namespace {

int lines_tiles(int mp, size_t nlines) {
  const int TL = mp <= 32 ? 64 : 32;
  return (int)((nlines + TL - 1) / TL);
}

void batch_free(crys_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void* p : b->allocs)
    if (p) (void)hipFree(p);
  b->allocs.clear();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->stream = nullptr;
}

int balloc(crys_batch* b, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess)
    return errf(e == hipErrorOutOfMemory ? CRYS_ERR_OUT_OF_MEMORY : CRYS_ERR_HIP,
                "batch allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  b->allocs.push_back(*p);
  return CRYS_OK;
}

template <typename T>
int batch_setup(crys_batch* b) {
  const size_t C = b->count;
  size_t tot_f = 0, tot_q = 0, tot_m = 0, tot_mc = 0;
  // the one-launch phases need every instance's three axes in one class;
  // centrosymmetric products where all three lengths equal a class of 32/64
  // per axis: the half-size product applies where the axis length equals its
  // class of 32 or 64 (PmmmDCT and P1 operators are centrosymmetric, 3m
  // operators have the half-period pairing)
  std::vector<int> pcls(C, 0), cen(C, 0), cax(3 * C, 0);
  bool all = crys::phase_enabled();
  for (size_t i = 0; i < C; ++i) {
    const crys_instance_t& d = b->desc[i];
    const int red = d.kind == CRYS_P1 ? 1 : 2;
    const int c = b->cls[3 * i];
    if (b->cls[3 * i + 1] == c && b->cls[3 * i + 2] == c) pcls[i] = c;
    else all = false;
    bool ce = pcls[i] != 0;
    for (int a = 0; a < 3; ++a) {
      const int ca = b->cls[3 * i + a];
      cax[3 * i + a] = crys::centro_enabled() && ca % 32 == 0 &&
                       (int)(d.grid_log[a] / red) == ca;
      ce = ce && cax[3 * i + a];
    }
    cen[i] = ce ? 1 : 0;
  }
  // Measured on MI355X (PmmmDCT instances of 16^3..64^3, contour steps of one
  // iteration, passes over phases, FP64 / FP32): 1.34 / 1.35 at 656 items,
  // 1.08 / 1.08 at 1264, 0.98 / 0.89 at 2240, 0.90 / 0.75 at 4752 and
  // 0.81 / 0.70 at 18400: the passes' lighter workgroups (and their
  // half-size products on every axis of length 32 or 64) win once the batch
  // fills the chip.
  size_t items = 0;
  for (size_t i = 0; i < C; ++i) {
    const crys_instance_t& d = b->desc[i];
    const size_t red = d.kind == CRYS_P1 ? 1 : 2;
    items += 2 * std::max(d.grid_log[0], d.grid_log[1]) / red;
  }
  b->phase_ok = all;
  b->phase_dense64 = false;
  for (size_t i = 0; i < C; ++i) b->phase_dense64 = b->phase_dense64 || (pcls[i] == 64 && !cen[i]);
  b->phases = all && items <= phase_batch_items();
  if (!all) std::fill(cen.begin(), cen.end(), 0);
  b->off_mc.assign(3 * C, 0);
  for (size_t i = 0; i < C; ++i) {
    b->off_f[i] = tot_f;
    b->off_q[i] = tot_q;
    tot_f += b->n[i];
    tot_q += (b->NsA[i] + b->NsB[i] + 1) * b->n[i];
    const crys_instance_t& d = b->desc[i];
    for (int a = 0; a < 3; ++a) {
      b->off_m[3 * i + a] = tot_m;
      const size_t M = d.kind == CRYS_P1 ? d.grid_log[a] : d.grid_log[a] / 2;
      tot_m += M * M;
      if (cax[3 * i + a]) {
        b->off_mc[3 * i + a] = tot_mc;
        tot_mc += M * M / 2;
      }
    }
  }
  BCHK(balloc(b, &b->qf, tot_q * sizeof(T)));
  BCHK(balloc(b, &b->qb, tot_q * sizeof(T)));
  // one scratch field per chain; the phases alternate between two
  BCHK(balloc(b, &b->scr, (b->phase_ok ? 4 : 2) * tot_f * sizeof(T)));
  BCHK(balloc(b, &b->EA, tot_f * sizeof(T)));
  BCHK(balloc(b, &b->EB, tot_f * sizeof(T)));
  BCHK(balloc(b, &b->mats, tot_m * sizeof(T)));
  double** fd[5] = {&b->wA, &b->wB, &b->ksi, &b->phiA, &b->phiB};
  for (double** p : fd) BCHK(balloc(b, (void**)p, tot_f * 8));
  BCHK(balloc(b, (void**)&b->part_q, C * crys::kMaxRedBlocks * 8));
  BCHK(balloc(b, (void**)&b->part_e, C * 4 * crys::kMaxRedBlocks * 8));
  BCHK(balloc(b, (void**)&b->scal, C * 4 * 8));
  BCHK(balloc(b, (void**)&b->trace, C * kCap * 7 * 8));
  BCHK(balloc(b, (void**)&b->counter, sizeof(int)));

  // operators (and, for the half-size forms, their double values)
  std::vector<T> mats(tot_m);
  std::vector<double> mats_d(tot_mc ? tot_m : 0);
  std::vector<unsigned short> lo(std::is_same<T, float>::value ? tot_m : 0);
  for (size_t i = 0; i < C; ++i) {
    const crys_instance_t& d = b->desc[i];
    for (int a = 0; a < 3; ++a) {
      std::vector<double> At;
      const int Nl = (int)d.grid_log[a];
      // PmmmDCT keeps the reference's cubic cell unless asked not to
      const double L = (d.kind == CRYS_PMMM_DCT && !(d.flags & CRYS_FLAG_PMMM_ORTHORHOMBIC))
                           ? d.cell[0]
                           : d.cell[a];
      if (d.kind == CRYS_PMMM_DCT) At = crys::pmmm_axis_operator(Nl / 2, L, d.ds);
      else if (d.kind == CRYS_P1) At = crys::p1_axis_operator(Nl, L, d.ds);
      else At = crys::m3_axis_operator(Nl, L, d.ds);
      std::copy(At.begin(), At.end(), mats.begin() + b->off_m[3 * i + a]);
      if (!mats_d.empty()) std::copy(At.begin(), At.end(), mats_d.begin() + b->off_m[3 * i + a]);
      if (std::is_same<T, float>::value)
        for (size_t k = 0; k < At.size(); ++k)
          lo[b->off_m[3 * i + a] + k] =
              crys::bf16_bits((float)(At[k] - (double)(float)At[k]));
    }
  }
  BHIP(hipMemcpy(b->mats, mats.data(), tot_m * sizeof(T), hipMemcpyHostToDevice));
  if (std::is_same<T, float>::value) {
    BCHK(balloc(b, (void**)&b->mats_lo, tot_m * sizeof(unsigned short)));
    BHIP(hipMemcpy(b->mats_lo, lo.data(), tot_m * sizeof(unsigned short), hipMemcpyHostToDevice));
  }
  if (tot_mc) {  // half-size forms (centro_half) of the centrosymmetric axes
    std::vector<T> mc(tot_mc);
    std::vector<unsigned short> mclo(std::is_same<T, float>::value ? tot_mc : 0);
    for (size_t i = 0; i < C; ++i) {
      const crys_instance_t& d = b->desc[i];
      for (int a = 0; a < 3; ++a) {
        if (!cax[3 * i + a]) continue;
        const int M = (int)(d.kind == CRYS_P1 ? d.grid_log[a] : d.grid_log[a] / 2);
        const size_t o = b->off_m[3 * i + a];
        std::vector<double> At(M * (size_t)M);
        for (size_t k = 0; k < At.size(); ++k) At[k] = (double)mats_d[o + k];
        const std::vector<double> Cm = crys::centro_half(At, M, d.kind == CRYS_3M);
        std::copy(Cm.begin(), Cm.end(), mc.begin() + b->off_mc[3 * i + a]);
        if (std::is_same<T, float>::value)
          for (size_t k = 0; k < Cm.size(); ++k)
            mclo[b->off_mc[3 * i + a] + k] =
                crys::bf16_bits((float)(Cm[k] - (double)(float)Cm[k]));
      }
    }
    BCHK(balloc(b, &b->mats_c, tot_mc * sizeof(T)));
    BHIP(hipMemcpy(b->mats_c, mc.data(), tot_mc * sizeof(T), hipMemcpyHostToDevice));
    if (std::is_same<T, float>::value) {
      BCHK(balloc(b, (void**)&b->mats_c_lo, tot_mc * sizeof(unsigned short)));
      BHIP(hipMemcpy(b->mats_c_lo, mclo.data(), tot_mc * sizeof(unsigned short),
                     hipMemcpyHostToDevice));
    }
  }

  // instance table
  std::vector<crys::BatchInst<T>> tab(C);
  T* qf = (T*)b->qf;
  T* qb = (T*)b->qb;
  T* scr = (T*)b->scr;
  for (size_t i = 0; i < C; ++i) {
    const crys_instance_t& d = b->desc[i];
    crys::BatchInst<T>& I = tab[i];
    const int red = d.kind == CRYS_P1 ? 1 : 2;
    I.X = (int)(d.grid_log[0] / red);
    I.Y = (int)(d.grid_log[1] / red);
    I.Z = (int)(d.grid_log[2] / red);
    for (int a = 0; a < 3; ++a) {
      I.cls[a] = b->cls[3 * i + a];
      I.mat[a] = (const T*)b->mats + b->off_m[3 * i + a];
      I.mat_lo[a] = b->mats_lo ? b->mats_lo + b->off_m[3 * i + a] : nullptr;
      I.mat_c[a] = cax[3 * i + a] ? (const T*)b->mats_c + b->off_mc[3 * i + a] : nullptr;
      I.mat_c_lo[a] =
          (cax[3 * i + a] && b->mats_c_lo) ? b->mats_c_lo + b->off_mc[3 * i + a] : nullptr;
    }
    I.pcls = pcls[i];
    I.cen = cen[i];
    I.shift = d.kind == CRYS_3M ? 1 : 0;
    I.NsA = (int)b->NsA[i];
    I.NsB = (int)b->NsB[i];
    I.Ns = I.NsA + I.NsB;
    I.n = (long)b->n[i];
    I.nblk = crys::red_blocks(b->n[i]);
    I.qf = qf + b->off_q[i];
    I.qb = qb + b->off_q[i];
    I.scr0 = scr + b->off_f[i];
    I.scr1 = scr + (b->off_f[i] + tot_f);
    I.scr0b = b->phase_ok ? scr + (b->off_f[i] + 2 * tot_f) : nullptr;
    I.scr1b = b->phase_ok ? scr + (b->off_f[i] + 3 * tot_f) : nullptr;
    I.EA = (T*)b->EA + b->off_f[i];
    I.EB = (T*)b->EB + b->off_f[i];
    I.wA = b->wA + b->off_f[i];
    I.wB = b->wB + b->off_f[i];
    I.ksi = b->ksi + b->off_f[i];
    I.phiA = b->phiA + b->off_f[i];
    I.phiB = b->phiB + b->off_f[i];
    I.fA = d.fA;
    I.fB = 1 - d.fA;
    I.xN = d.xN;
    I.acc = d.acceptance;
    I.ds = d.ds;
  }
  BCHK(balloc(b, &b->insts, C * sizeof(crys::BatchInst<T>)));
  BHIP(hipMemcpy(b->insts, tab.data(), C * sizeof(crys::BatchInst<T>), hipMemcpyHostToDevice));

  // work lists, instances by decreasing contour length
  std::vector<size_t> order(C);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    return b->NsA[x] + b->NsB[x] > b->NsA[y] + b->NsB[y];
  });
  b->Ns_max = (int)(b->NsA[order[0]] + b->NsB[order[0]]);
  for (int a = 0; a < 3; ++a) {
    std::vector<int2> wl;
    std::vector<int> end_of;  // cumulative count after each sorted instance
    for (size_t k = 0; k < C; ++k) {
      const size_t i = order[k];
      const crys::BatchInst<T>& I = tab[i];
      const int M = a == 0 ? I.X : (a == 1 ? I.Y : I.Z);
      const int nt = lines_tiles(I.cls[a], (size_t)I.n / M);
      for (int ch = 0; ch < 2; ++ch)
        for (int t = 0; t < nt; ++t) wl.push_back(int2{(int)i, (ch << 16) | t});
      end_of.push_back((int)wl.size());
    }
    b->prefix[a].assign(b->Ns_max, 0);
    for (int t = 0; t < b->Ns_max; ++t) {
      int cnt = 0;
      for (size_t k = 0; k < C; ++k)
        if ((int)(b->NsA[order[k]] + b->NsB[order[k]]) > t) cnt = end_of[k];
      b->prefix[a][t] = cnt;
    }
    BCHK(balloc(b, (void**)&b->work[a], wl.size() * sizeof(int2)));
    BHIP(hipMemcpy(b->work[a], wl.data(), wl.size() * sizeof(int2), hipMemcpyHostToDevice));
  }
  if (b->phase_ok) {
    for (int ys = 0; ys < 2; ++ys) {
      std::vector<int2> wl;
      std::vector<int> end_of;
      for (size_t k = 0; k < C; ++k) {
        const size_t i = order[k];
        const int planes = ys ? tab[i].Y : tab[i].X;
        for (int ch = 0; ch < 2; ++ch)
          for (int pl = 0; pl < planes; ++pl) wl.push_back(int2{(int)i, (ch << 16) | pl});
        end_of.push_back((int)wl.size());
      }
      // phase k (0..Ns) of an instance with Ns contour steps
      b->pprefix[ys].assign(b->Ns_max + 1, 0);
      for (int k = 0; k <= b->Ns_max; ++k) {
        int cnt = 0;
        for (size_t j = 0; j < C; ++j)
          if ((int)(b->NsA[order[j]] + b->NsB[order[j]]) >= k) cnt = end_of[j];
        b->pprefix[ys][k] = cnt;
      }
      BCHK(balloc(b, (void**)&b->pwork[ys], wl.size() * sizeof(int2)));
      BHIP(hipMemcpy(b->pwork[ys], wl.data(), wl.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
  }
  std::vector<int2> rw;
  for (size_t i = 0; i < C; ++i)
    for (int k = 0; k < tab[i].nblk; ++k) rw.push_back(int2{(int)i, k});
  b->n_rwork = (int)rw.size();
  BCHK(balloc(b, (void**)&b->rwork, rw.size() * sizeof(int2)));
  BHIP(hipMemcpy(b->rwork, rw.data(), rw.size() * sizeof(int2), hipMemcpyHostToDevice));

  // fixed ends of the chains: q_A(0) = 1 and q_B^dagger(NsB) = 1
  for (size_t i = 0; i < C; ++i) {
    const size_t n = b->n[i], Ns = b->NsA[i] + b->NsB[i];
    hipLaunchKernelGGL((crys::fill_kernel<T>), dim3(256), dim3(256), 0, b->stream, tab[i].qf, (T)1,
                       n);
    hipLaunchKernelGGL((crys::fill_kernel<T>), dim3(256), dim3(256), 0, b->stream,
                       tab[i].qb + Ns * n, (T)1, n);
  }
  BHIP(hipGetLastError());
  BHIP(hipStreamSynchronize(b->stream));
  return CRYS_OK;
}

// The contour steps of one iteration of every instance: the phases (one
// launch per step) or the per-axis passes (three).
template <typename T>
int batch_record_steps(crys_batch* b, hipStream_t s, bool phases) {
  const crys::BatchInst<T>* I = (const crys::BatchInst<T>*)b->insts;
  for (int k = 0; phases && k <= b->Ns_max; ++k) {
    const int ys = k % 2;
    const int cnt = b->pprefix[ys][k];
    if (cnt == 0) continue;
    if (b->phase_grid > 0) {
      const unsigned g = (unsigned)std::min(cnt, b->phase_grid);
      if (b->phase_dense64)
        hipLaunchKernelGGL((crys::batch_phase_persistent_kernel<T, 512>), dim3(g), dim3(512), 0,
                           s, I, (const int2*)b->pwork[ys], cnt, k, ys);
      else
        hipLaunchKernelGGL((crys::batch_phase_persistent_kernel<T, 1024>), dim3(g), dim3(1024),
                           0, s, I, (const int2*)b->pwork[ys], cnt, k, ys);
    } else if (b->phase_dense64) {
      hipLaunchKernelGGL((crys::batch_phase_kernel<T, 512>), dim3(cnt), dim3(512), 0, s, I,
                         (const int2*)b->pwork[ys], k, ys);
    } else {
      hipLaunchKernelGGL((crys::batch_phase_kernel<T, 1024>), dim3(cnt), dim3(1024), 0, s, I,
                         (const int2*)b->pwork[ys], k, ys);
    }
    BHIP(hipGetLastError());
  }
  for (int t = 0; !phases && t < b->Ns_max; ++t) {
    const int ax_order[3] = {2, 1, 0};
    for (int a : ax_order) {
      const int cnt = b->prefix[a][t];
      if (cnt == 0) continue;
      if (a == 2)
        hipLaunchKernelGGL((crys::batch_pass_kernel<T, 2>), dim3(cnt), dim3(256), 0, s, I,
                           (const int2*)b->work[2], t);
      else if (a == 1)
        hipLaunchKernelGGL((crys::batch_pass_kernel<T, 1>), dim3(cnt), dim3(256), 0, s, I,
                           (const int2*)b->work[1], t);
      else
        hipLaunchKernelGGL((crys::batch_pass_kernel<T, 0>), dim3(cnt), dim3(256), 0, s, I,
                           (const int2*)b->work[0], t);
      BHIP(hipGetLastError());
    }
  }
  return CRYS_OK;
}

template <typename T>
int batch_record(crys_batch* b, hipStream_t s) {
  const crys::BatchInst<T>* I = (const crys::BatchInst<T>*)b->insts;
  BCHK(batch_record_steps<T>(b, s, b->phases));
  const unsigned C = (unsigned)b->count;
  hipLaunchKernelGGL((crys::batch_qsum_kernel<T>), dim3(b->n_rwork), dim3(crys::kRedThreads), 0,
                     s, I, (const int2*)b->rwork, b->part_q);
  hipLaunchKernelGGL((crys::batch_qfinal_kernel<T>), dim3(C), dim3(crys::kRedThreads), 0, s, I,
                     (const double*)b->part_q, b->scal);
  hipLaunchKernelGGL((crys::batch_phi_kernel<T>), dim3(b->n_rwork), dim3(crys::kRedThreads), 0, s,
                     I, (const int2*)b->rwork, (const double*)b->scal, b->part_e);
  hipLaunchKernelGGL((crys::batch_efinal_kernel<T>), dim3(C), dim3(crys::kRedThreads), 0, s, I,
                     (const double*)b->part_e, (const double*)b->scal, b->trace,
                     (const int*)b->counter);
  hipLaunchKernelGGL(crys::batch_count_kernel, dim3(1), dim3(1), 0, s, b->counter, 0);
  BHIP(hipGetLastError());
  return CRYS_OK;
}

void batch_drop_graph(crys_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Mean milliseconds of the contour steps of one iteration on one path, over
// reps back-to-back recordings after one warm-up.  The steps only write the
// histories and the scratch fields, which the next iteration rewrites, so
// timing them leaves the batch's state as it was.
int batch_time_steps(crys_batch* b, bool phases, int reps, double* ms) {
  hipEvent_t e0, e1;
  BHIP(hipEventCreate(&e0));
  BHIP(hipEventCreate(&e1));
  auto rec = [&]() {
    return b->precision == CRYS_FP32 ? batch_record_steps<float>(b, b->stream, phases)
                                     : batch_record_steps<double>(b, b->stream, phases);
  };
  int r = rec();
  if (r == CRYS_OK) r = hipEventRecord(e0, b->stream) == hipSuccess ? CRYS_OK : CRYS_ERR_HIP;
  for (int i = 0; i < reps && r == CRYS_OK; ++i) r = rec();
  if (r == CRYS_OK) r = hipEventRecord(e1, b->stream) == hipSuccess ? CRYS_OK : CRYS_ERR_HIP;
  float t = 0.f;
  if (r == CRYS_OK && hipEventSynchronize(e1) == hipSuccess &&
      hipEventElapsedTime(&t, e0, e1) == hipSuccess)
    *ms = (double)t / reps;
  else if (r == CRYS_OK)
    r = errf(CRYS_ERR_HIP, "event timing failed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return r;
}

int batch_graph(crys_batch* b) {
  if (b->exec) return CRYS_OK;
  BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int r = b->precision == CRYS_FP32 ? batch_record<float>(b, b->stream)
                                    : batch_record<double>(b, b->stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (r != CRYS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return errf(CRYS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  b->graph = g;
  BHIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return CRYS_OK;
}

}  // namespace

extern "C" {

int crys_batch_create(const crys_instance_t* desc, size_t count, int precision,
                      crys_batch_t** out) {
  if (!out || !desc || count == 0) return errf(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  *out = nullptr;
  if (precision != CRYS_FP64 && precision != CRYS_FP32)
    return errf(CRYS_ERR_INVALID_ARGUMENT, "unknown precision %d", precision);
  if (count > 65535) return errf(CRYS_ERR_UNSUPPORTED, "at most 65535 instances");
  crys_batch* b = new (std::nothrow) crys_batch();
  if (!b) return errf(CRYS_ERR_OUT_OF_MEMORY, "host allocation failed");
  b->precision = precision;
  b->count = count;
  b->desc.assign(desc, desc + count);
  b->n.resize(count);
  b->NsA.resize(count);
  b->NsB.resize(count);
  b->off_f.resize(count);
  b->off_q.resize(count);
  b->cls.resize(3 * count);
  b->off_m.resize(3 * count);
  b->fields_set.assign(count, 0);
  for (size_t i = 0; i < count; ++i) {
    const crys_instance_t& d = desc[i];
    const size_t* g = d.grid_log;
    if (d.kind != CRYS_P1 && d.kind != CRYS_3M && d.kind != CRYS_PMMM_DCT) {
      delete b;
      return errf(CRYS_ERR_INVALID_ARGUMENT, "instance %zu: unknown kind", i);
    }
    bool bad = g[0] == 0 || g[1] == 0 || g[2] == 0;
    if (d.kind == CRYS_P1) bad = bad || g[2] % 8 != 0;
    else bad = bad || g[0] % 2 || g[1] % 2 || g[2] % 2 || (g[2] / 2) % 8 != 0;
    if (bad) {
      delete b;
      return errf(CRYS_ERR_BAD_GRID, "instance %zu: grid %zu x %zu x %zu rejected", i, g[0], g[1],
                  g[2]);
    }
    if (!(d.ds > 0) || !(d.fA > 0 && d.fA < 1) || !(d.cell[0] > 0 && d.cell[1] > 0 && d.cell[2] > 0)) {
      delete b;
      return errf(CRYS_ERR_INVALID_ARGUMENT, "instance %zu: bad ds, fA or cell", i);
    }
    b->NsA[i] = (size_t)(d.fA / d.ds);        // crysSCFT_sample.cpp:443
    b->NsB[i] = (size_t)((1 - d.fA) / d.ds);  // :444
    if (b->NsA[i] == 0 || b->NsB[i] == 0) {
      delete b;
      return errf(CRYS_ERR_INVALID_ARGUMENT, "instance %zu: a block has no contour step", i);
    }
    const int red = d.kind == CRYS_P1 ? 1 : 2;
    b->n[i] = (g[0] / red) * (g[1] / red) * (g[2] / red);
    for (int a = 0; a < 3; ++a) {
      const int c = crys::pass_class((int)(g[a] / red));
      if (c == 0 || c > 64) {
        delete b;
        return errf(CRYS_ERR_UNSUPPORTED, "instance %zu: batched axes are limited to 64 points", i);
      }
      b->cls[3 * i + a] = c;
    }
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return errf(CRYS_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
  }
  b->phase_grid = batch_persistent_enabled() ? std::max(b->n_cu, 1) : 0;
  int r = precision == CRYS_FP32 ? batch_setup<float>(b) : batch_setup<double>(b);
  if (r != CRYS_OK) {
    batch_free(b);
    delete b;
    return r;
  }
  *out = b;
  return CRYS_OK;
}

int crys_batch_destroy(crys_batch_t* b) {
  if (!b) return CRYS_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  delete b;
  return CRYS_OK;
}

int crys_batch_set_fields(crys_batch_t* b, size_t i, const double* wA, const double* wB) {
  if (!b || !wA || !wB || i >= b->count) return errf(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  const size_t n = b->n[i], o = b->off_f[i];
  hipStream_t s = b->stream;
  BHIP(hipMemcpyAsync(b->wA + o, wA, n * 8, hipMemcpyHostToDevice, s));
  BHIP(hipMemcpyAsync(b->wB + o, wB, n * 8, hipMemcpyHostToDevice, s));
  const crys_instance_t& d = b->desc[i];
  if (b->precision == CRYS_FP32)
    hipLaunchKernelGGL((crys::prep_fields_kernel<float>), dim3(256), dim3(256), 0, s,
                       (const double*)b->wA + o, (const double*)b->wB + o, b->ksi + o,
                       (float*)b->EA + o, (float*)b->EB + o, d.xN, d.ds, n);
  else
    hipLaunchKernelGGL((crys::prep_fields_kernel<double>), dim3(256), dim3(256), 0, s,
                       (const double*)b->wA + o, (const double*)b->wB + o, b->ksi + o,
                       (double*)b->EA + o, (double*)b->EB + o, d.xN, d.ds, n);
  BHIP(hipGetLastError());
  BHIP(hipStreamSynchronize(s));
  b->fields_set[i] = 1;
  return CRYS_OK;
}

int crys_batch_run(crys_batch_t* b, size_t nstep, double* trace) {
  if (!b) return errf(CRYS_ERR_INVALID_ARGUMENT, "null handle");
  for (size_t i = 0; i < b->count; ++i)
    if (!b->fields_set[i]) return errf(CRYS_ERR_INVALID_ARGUMENT, "instance %zu has no fields", i);
  BCHK(batch_graph(b));
  hipStream_t s = b->stream;
  std::vector<double> rows;
  bool diverged = false;
  size_t done = 0;
  while (done < nstep) {
    const size_t chunk = std::min(nstep - done, (size_t)kCap);
    hipLaunchKernelGGL(crys::batch_count_kernel, dim3(1), dim3(1), 0, s, b->counter, 1);
    BHIP(hipGetLastError());
    for (size_t k = 0; k < chunk; ++k) BHIP(hipGraphLaunch(b->exec, s));
    if (trace) {
      rows.resize(b->count * kCap * 7);
      BHIP(hipMemcpyAsync(rows.data(), b->trace, rows.size() * 8, hipMemcpyDeviceToHost, s));
    }
    BHIP(hipStreamSynchronize(s));
    if (trace)
      for (size_t i = 0; i < b->count; ++i)
        for (size_t k = 0; k < chunk; ++k) {
          const double* r = &rows[(i * kCap + k) * 7];
          memcpy(trace + (i * nstep + done + k) * 7, r, 7 * 8);
          if (!std::isfinite(r[0]) || !std::isfinite(r[6])) diverged = true;
        }
    done += chunk;
  }
  if (diverged) return errf(CRYS_ERR_DIVERGED, "non-finite free energy or Q in the trace");
  return CRYS_OK;
}

int crys_batch_set_path(crys_batch_t* b, int path) {
  if (!b) return errf(CRYS_ERR_INVALID_ARGUMENT, "null handle");
  if (path < CRYS_BATCH_PATH_AUTO || path > CRYS_BATCH_PATH_PASSES)
    return errf(CRYS_ERR_INVALID_ARGUMENT, "unknown path %d", path);
  if (path == CRYS_BATCH_PATH_PHASES && !b->phase_ok)
    return errf(CRYS_ERR_UNSUPPORTED,
                "the phases need every instance's three axes in one length class");
  bool ph = path == CRYS_BATCH_PATH_PHASES;
  if (path == CRYS_BATCH_PATH_AUTO) {
    size_t items = 0;
    for (const crys_instance_t& d : b->desc)
      items += 2 * std::max(d.grid_log[0], d.grid_log[1]) / (d.kind == CRYS_P1 ? 1 : 2);
    ph = b->phase_ok && items <= phase_batch_items();
  }
  if (ph != b->phases) {
    BHIP(hipStreamSynchronize(b->stream));
    batch_drop_graph(b);
    b->phases = ph;
  }
  return CRYS_OK;
}

int crys_batch_get_path(const crys_batch_t* b, int* path) {
  if (!b || !path) return errf(CRYS_ERR_INVALID_ARGUMENT, "null argument");
  *path = b->phases ? CRYS_BATCH_PATH_PHASES : CRYS_BATCH_PATH_PASSES;
  return CRYS_OK;
}

int crys_batch_autotune(crys_batch_t* b, int reps, int* path, double ms[2]) {
  if (!b || reps <= 0) return errf(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  double t_ph = -1.0, t_pa = -1.0;
  if (b->phase_ok) {
    // the phases on one workgroup per item and on persistent grids of 1, 2
    // and 4 workgroups per CU; the fastest grid is kept
    const int keep = b->phase_grid;
    const int grids[4] = {0, b->n_cu, 2 * b->n_cu, 4 * b->n_cu};
    int best = keep;
    for (int gi = 0; gi < 4; ++gi) {
      if (gi > 0 && grids[gi] <= 0) continue;
      b->phase_grid = grids[gi];
      double t = -1.0;
      const int r = batch_time_steps(b, true, reps, &t);
      if (r != CRYS_OK) {
        b->phase_grid = keep;
        return r;
      }
      if (t_ph < 0 || t < t_ph) {
        t_ph = t;
        best = grids[gi];
      }
    }
    if (best != keep) {
      BHIP(hipStreamSynchronize(b->stream));
      batch_drop_graph(b);
    }
    b->phase_grid = best;
  }
  BCHK(batch_time_steps(b, false, reps, &t_pa));
  const bool ph = b->phase_ok && t_ph < t_pa;
  BCHK(crys_batch_set_path(b, ph ? CRYS_BATCH_PATH_PHASES : CRYS_BATCH_PATH_PASSES));
  if (path) *path = ph ? CRYS_BATCH_PATH_PHASES : CRYS_BATCH_PATH_PASSES;
  if (ms) {
    ms[0] = t_ph;
    ms[1] = t_pa;
  }
  return CRYS_OK;
}

int crys_batch_set_phase_grid(crys_batch_t* b, int workgroups) {
  if (!b || workgroups < 0) return errf(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  if (workgroups != b->phase_grid) {
    BHIP(hipStreamSynchronize(b->stream));
    batch_drop_graph(b);
    b->phase_grid = workgroups;
  }
  return CRYS_OK;
}

int crys_batch_get_phase_grid(const crys_batch_t* b, int* workgroups) {
  if (!b || !workgroups) return errf(CRYS_ERR_INVALID_ARGUMENT, "null argument");
  *workgroups = b->phase_grid;
  return CRYS_OK;
}

int crys_batch_get_fields(crys_batch_t* b, size_t i, double* phiA, double* phiB, double* wA,
                          double* wB) {
  if (!b || i >= b->count) return errf(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  const size_t n = b->n[i], o = b->off_f[i];
  hipStream_t s = b->stream;
  if (phiA) BHIP(hipMemcpyAsync(phiA, b->phiA + o, n * 8, hipMemcpyDeviceToHost, s));
  if (phiB) BHIP(hipMemcpyAsync(phiB, b->phiB + o, n * 8, hipMemcpyDeviceToHost, s));
  if (wA) BHIP(hipMemcpyAsync(wA, b->wA + o, n * 8, hipMemcpyDeviceToHost, s));
  if (wB) BHIP(hipMemcpyAsync(wB, b->wB + o, n * 8, hipMemcpyDeviceToHost, s));
  BHIP(hipStreamSynchronize(s));
  return CRYS_OK;
}

int crys_batch_get_info(const crys_batch_t* b, size_t i, size_t* n_phy, size_t* NsA,
                        size_t* NsB) {
  if (!b || i >= b->count) return errf(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  if (n_phy) *n_phy = b->n[i];
  if (NsA) *NsA = b->NsA[i];
  if (NsB) *NsB = b->NsB[i];
  return CRYS_OK;
}

}  // extern "C"
