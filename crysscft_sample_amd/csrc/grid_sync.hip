// Measurement of the synchronisation between contour steps.
//
// Every phase of the one-launch contour step (plane_phase.hpp) ends with all
// workgroups' half-finished planes written and starts by reading planes that
// workgroups on the other XCDs wrote: an x-slab / y-slab transpose through
// the chip.  That boundary is a kernel launch today.  The alternative is a
// persistent kernel with a grid barrier between phases.  crys_time_grid_sync
// times both on the same exchange: each of W workgroups of 1024 threads (one
// per CU, like the phase kernel) writes 64 KB and then reads the 64 KB that a
// workgroup 37 ids away (another XCD) wrote, for R rounds, and counts the
// values that were not yet visible.
//
// Variants (enum in crys_solver.h):
//   LAUNCH      one round per kernel launch (the launch is the barrier)
//   EMPTY       empty launches of the same shape (the launch floor)
//   FENCED      persistent; one counter; agent-scope release / acquire fences
//               around the arrival (L2 write-back and invalidate)
//   TWO_LEVEL   persistent; per-XCD counters (workgroup id mod 8), the last
//               arrival of each group arrives at a root counter and releases
//               its group; agent-scope fences as FENCED
//   BYPASS      persistent; one counter; the exchanged data is stored and
//               loaded with agent-scope accesses (write-through, L2 bypass),
//               so no L2 fence is needed
//   BARRIER     persistent; FENCED without the data (barrier alone)
//   XCD         persistent; the workgroups of one XCD (read from the XCC_ID
//               hardware register) synchronise among themselves on one
//               counter per XCD, with no L2 fence: the partner is 5 places on
//               in the same XCD, and its data is read with loads that miss
//               the CU's L1 and hit the XCD's L2, where every CU's stores
//               land (the vector L1 is write-through)
// Every spin is bounded; a workgroup that times out sets a flag that ends
// every workgroup at its next round, and the count of time-outs is returned.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "crys_solver.h"

namespace crys {
int set_error(int code, const char* msg);
}

namespace {

constexpr int kSyncThreads = 1024;
constexpr int kSyncPerWg = 8192;       // doubles exchanged per workgroup and round (64 KB)
constexpr unsigned kSpinMax = 1u << 20;  // bounded spin (about 30 ms with s_sleep)
constexpr int kSyncMaxWg = 256;
constexpr size_t kSyncLds = 100 * 1024;  // one workgroup per CU, like the phase kernel

struct SyncArgs {
  double* buf;    // 2 * kSyncMaxWg * kSyncPerWg (double buffer by round parity)
  unsigned* cnt;  // [0] flat counter; [1..8] per-XCD leaves; [9] root; [16..23] group release;
                  // [32..39] per-XCD tickets (XCD)
  unsigned* err;  // [0] values not yet visible, [1] spin time-outs
  int nwg, rounds, variant, stride;
};

__device__ inline void bounded_spin(unsigned* p, unsigned target, unsigned* err) {
  unsigned n = 0;
  while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++n > kSpinMax) {
      __hip_atomic_fetch_add(&err[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
  }
}

__device__ inline double exch_value(int round, int wg) { return (double)(round * 1000 + wg); }

// XCD of the calling wave: the XCC_ID hardware register, bits 3:0.
__device__ inline int xcd_of_wave() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20); }

// One double stored by another workgroup of the same XCD: the buffer load
// carries sc0, so it misses this CU's L1 and reads the XCD's shared L2.
__device__ inline double load_from_l2(__amdgpu_buffer_rsrc_t r, int byte_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 1));
}

__global__ __launch_bounds__(kSyncThreads) void sync_persistent_kernel(SyncArgs a) {
  extern __shared__ double sync_lds[];
  __shared__ int stop, where;
  const int g = blockIdx.x, tid = threadIdx.x;
  sync_lds[tid] = 0.0;
  const bool data = a.variant != CRYS_SYNC_BARRIER;
  const bool bypass = a.variant == CRYS_SYNC_BYPASS;
  const bool xcd = a.variant == CRYS_SYNC_XCD;
  const int per = a.nwg / 8;
  // XCD: this workgroup's XCD and its place among the workgroups there
  if (tid == 0) {
    stop = 0;
    where = 0;
    if (xcd) {
      const int x = xcd_of_wave();
      const unsigned t = x < 8 ? __hip_atomic_fetch_add(&a.cnt[32 + x], 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                               : 0u;
      if (x >= 8 || t >= (unsigned)per) {
        __hip_atomic_fetch_add(&a.err[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        stop = 1;
      }
      where = x * per + (int)t;
    }
  }
  __syncthreads();
  if (stop) return;
  const int me = xcd ? where : g;
  const int group = me / per, slot = me - group * per;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      a.buf, (short)0, (int)(2 * kSyncMaxWg * kSyncPerWg * sizeof(double)), 0x00020000);
  for (int r = 0; r < a.rounds; ++r) {
    const size_t moff = ((size_t)(r & 1) * a.nwg + me) * kSyncPerWg;
    double* mine = a.buf + moff;
    if (data) {
      const double v = exch_value(r, me);
      for (int i = tid; i < kSyncPerWg; i += kSyncThreads) {
        if (bypass) {
          __hip_atomic_store(&mine[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (xcd) {
          // an ordinary store (kept in the L2), written as a buffer store so
          // that it is not merged with the non-temporal store below
          __builtin_amdgcn_raw_buffer_store_b64(
              __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), rsrc,
              (int)((moff + i) * sizeof(double)), 0, 0);
        } else {
          __builtin_nontemporal_store(v, &mine[i]);
        }
      }
    }
    // every store of this wave has reached the L2 (vmcnt covers stores)
    if (xcd) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
      if (xcd) {
        __hip_atomic_fetch_add(&a.cnt[1 + group], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bounded_spin(&a.cnt[1 + group], (unsigned)(r + 1) * per, a.err);
      } else if (a.variant == CRYS_SYNC_TWO_LEVEL) {
        const int x = g % 8, per = a.nwg / 8;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned old = __hip_atomic_fetch_add(&a.cnt[1 + x], 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == (unsigned)(r + 1) * per) {
          __hip_atomic_fetch_add(&a.cnt[9], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          bounded_spin(&a.cnt[9], (unsigned)(r + 1) * 8, a.err);
          __hip_atomic_store(&a.cnt[16 + x], (unsigned)(r + 1), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        } else {
          bounded_spin(&a.cnt[16 + x], (unsigned)(r + 1), a.err);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      } else if (bypass) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __hip_atomic_fetch_add(&a.cnt[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bounded_spin(&a.cnt[0], (unsigned)(r + 1) * a.nwg, a.err);
      } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&a.cnt[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bounded_spin(&a.cnt[0], (unsigned)(r + 1) * a.nwg, a.err);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      stop = __hip_atomic_load(&a.err[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    }
    __syncthreads();
    if (stop) return;
    if (data) {
      const int src = xcd ? group * per + (slot + 5) % per : (g + a.stride) % a.nwg;
      const size_t off = ((size_t)(r & 1) * a.nwg + src) * kSyncPerWg;
      const double* theirs = a.buf + off;
      const double want = exch_value(r, src);
      unsigned bad = 0;
      for (int i = tid; i < kSyncPerWg; i += kSyncThreads) {
        const double got =
            xcd ? load_from_l2(rsrc, (int)((off + i) * sizeof(double)))
                : bypass ? __hip_atomic_load(&theirs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : theirs[i];
        bad += got != want;
      }
      if (bad) __hip_atomic_fetch_add(&a.err[0], bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // buffer r & 1 is rewritten in round r + 2, after the barrier of round
    // r + 1, which every reader of round r has passed
  }
}

// One round per launch: read what the previous launch wrote, write this round.
__global__ __launch_bounds__(kSyncThreads) void sync_round_kernel(SyncArgs a, int r) {
  extern __shared__ double sync_lds[];
  const int g = blockIdx.x, tid = threadIdx.x;
  sync_lds[tid] = 0.0;
  if (r > 0) {
    const int src = (g + a.stride) % a.nwg;
    const double* theirs = a.buf + ((size_t)((r - 1) & 1) * a.nwg + src) * kSyncPerWg;
    const double want = exch_value(r - 1, src);
    unsigned bad = 0;
    for (int i = tid; i < kSyncPerWg; i += kSyncThreads) bad += theirs[i] != want;
    if (bad) __hip_atomic_fetch_add(&a.err[0], bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  double* mine = a.buf + ((size_t)(r & 1) * a.nwg + g) * kSyncPerWg;
  const double v = exch_value(r, g);
  for (int i = tid; i < kSyncPerWg; i += kSyncThreads) __builtin_nontemporal_store(v, &mine[i]);
}

__global__ __launch_bounds__(kSyncThreads) void sync_empty_kernel(int r) {
  extern __shared__ double sync_lds[];
  sync_lds[threadIdx.x] = (double)r;
}

int sync_fail(int code, const char* what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof(buf), "crys_time_grid_sync: %s: %s", what, hipGetErrorString(e));
  return crys::set_error(code, buf);
}

}  // namespace

extern "C" int crys_time_grid_sync(int variant, int workgroups, int rounds, double* us_per_round,
                                   unsigned* not_visible, unsigned* timeouts) {
  if (variant < CRYS_SYNC_LAUNCH || variant > CRYS_SYNC_XCD || rounds < 1 || workgroups < 8 ||
      workgroups > kSyncMaxWg || workgroups % 8 != 0 || !us_per_round)
    return crys::set_error(CRYS_ERR_INVALID_ARGUMENT,
                           "crys_time_grid_sync: variant, 8..256 workgroups (multiple of 8), "
                           "rounds >= 1");
  const bool persistent = variant >= CRYS_SYNC_FENCED;
  hipError_t e = hipSuccess;
  double* buf = nullptr;
  unsigned *cnt = nullptr, *err = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float best = 0.f;
  unsigned h_err[2] = {0, 0};
  int rc = CRYS_OK;
  if (persistent) {
    int dev = 0, coop = 0, cus = 0, per_cu = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
    if (e == hipSuccess)
      e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)sync_persistent_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSyncLds);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sync_persistent_kernel,
                                                       kSyncThreads, kSyncLds);
    if (e != hipSuccess) return sync_fail(CRYS_ERR_HIP, "device query", e);
    if (!coop || workgroups > per_cu * cus)
      return crys::set_error(CRYS_ERR_UNSUPPORTED,
                             "crys_time_grid_sync: the workgroups cannot all be resident");
  } else {
    const void* k = variant == CRYS_SYNC_LAUNCH ? (const void*)sync_round_kernel
                                                : (const void*)sync_empty_kernel;
    e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSyncLds);
    if (e != hipSuccess) return sync_fail(CRYS_ERR_HIP, "kernel attribute", e);
  }
  e = hipMalloc((void**)&buf, (size_t)2 * kSyncMaxWg * kSyncPerWg * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&cnt, 64 * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc((void**)&err, 2 * sizeof(unsigned));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  // a warm-up pass, then the timed pass
  for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {
    e = hipMemsetAsync(cnt, 0, 64 * sizeof(unsigned), s);
    if (e == hipSuccess) e = hipMemsetAsync(err, 0, 2 * sizeof(unsigned), s);
    SyncArgs a{buf, cnt, err, workgroups, rounds, variant, 37};
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    if (persistent) {
      void* args[] = {&a};
      if (e == hipSuccess)
        e = hipLaunchCooperativeKernel((const void*)sync_persistent_kernel, dim3(workgroups),
                                       dim3(kSyncThreads), args, kSyncLds, s);
    } else {
      for (int r = 0; r < rounds && e == hipSuccess; ++r) {
        if (variant == CRYS_SYNC_LAUNCH)
          hipLaunchKernelGGL(sync_round_kernel, dim3(workgroups), dim3(kSyncThreads), kSyncLds, s,
                             a, r);
        else
          hipLaunchKernelGGL(sync_empty_kernel, dim3(workgroups), dim3(kSyncThreads), kSyncLds, s,
                             r);
        e = hipGetLastError();
      }
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&best, e0, e1);
    if (e == hipSuccess) e = hipMemcpy(h_err, err, sizeof(h_err), hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) rc = sync_fail(CRYS_ERR_HIP, "timing", e);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (buf) (void)hipFree(buf);
  if (cnt) (void)hipFree(cnt);
  if (err) (void)hipFree(err);
  if (rc != CRYS_OK) return rc;
  *us_per_round = (double)best * 1e3 / rounds;
  if (not_visible) *not_visible = h_err[0];
  if (timeouts) *timeouts = h_err[1];
  return CRYS_OK;
}
