// libcrys_solver: the C ABI of include/crys_solver.h on gfx950.
//
// Host side of the drop-in for CrysPSSolverBase (solver_base.h:8-51) and its
// three solvers, plus the device-resident SCFT iteration of
// diblock_scft_simple_iterate (crysSCFT_sample.cpp:426-531).
#include "crys_solver.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <algorithm>
#include <array>
#include <new>
#include <string>
#include <vector>

#include "axis_pass.hpp"
#include "field_ops.hpp"
#include "m3_glide.hpp"
#include "operators.hpp"
#include "plane_phase.hpp"
#include "radix_pass.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(e_ == hipErrorOutOfMemory ? CRYS_ERR_OUT_OF_MEMORY : CRYS_ERR_HIP,      \
                  "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define CHK(expr)                 \
  do {                            \
    int r_ = (expr);              \
    if (r_ != CRYS_OK) return r_; \
  } while (0)

constexpr int kTraceCap = 1024;
// Iterations per graph launch in crys_scft_run: fewer graph boundaries, 0.2%
// more SCFT iterations per second at 128^3 (FP64 589.1 -> 590.3, FP32
// 801.5 -> 803.3 iter/s; 16 measured the same as 4).
constexpr int kGraphIters = 4;

unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  return (unsigned)(b ? b : 1);
}

}  // namespace

namespace crys {
int set_error(int code, const char* msg) { return fail(code, "%s", msg); }

// CRYS_PHASE=0 records the SCFT iteration with two (or three) launches per
// contour step instead of the one-launch plane phases (for A/B timing).
bool phase_enabled() {
  static const bool on = [] {
    const char* e = getenv("CRYS_PHASE");
    return !(e && strcmp(e, "0") == 0);
  }();
  return on;
}

// CRYS_PHASE_HALF=0 runs the class-64 centrosymmetric phase on one workgroup
// per plane instead of two (for A/B timing).
bool phase_half_enabled() {
  static const bool on = [] {
    const char* e = getenv("CRYS_PHASE_HALF");
    return !(e && strcmp(e, "0") == 0);
  }();
  return on;
}

// CRYS_PHASE_REUSE=0 runs the class-64 split phase with its next row and z
// products split by k half over all 16 waves (for A/B timing; see
// plane_phase_half_body).
bool phase_reuse_enabled() {
  static const bool on = [] {
    const char* e = getenv("CRYS_PHASE_REUSE");
    return !(e && strcmp(e, "0") == 0);
  }();
  return on;
}

// CRYS_CENTRO=0 uses the dense operators instead of the centrosymmetric
// half-size products (for A/B timing).
bool centro_enabled() {
  static const bool on = [] {
    const char* e = getenv("CRYS_CENTRO");
    return !(e && strcmp(e, "0") == 0);
  }();
  return on;
}

// CRYS_SLAB=0 runs the z and y passes as two launches (for A/B timing).
bool slab_enabled() {
  static const bool on = [] {
    const char* e = getenv("CRYS_SLAB");
    return !(e && strcmp(e, "0") == 0);
  }();
  return on;
}

// Per-axis pass kernel.  By default an axis of length 128 takes the radix
// FFT/DCT pass (radix_pass.hpp), measured faster there than the matrix-core
// product in FP64 and FP32, and every other axis the matrix-core pass (or the
// fused z+y slab).  For A/B timing CRYS_AXIS_PASS=mfma forces the matrix-core
// passes, CRYS_AXIS_PASS=radix the radix pass on every power-of-two axis and
// CRYS_AXIS_PASS=valu the FMA pass kernel.
PassImpl pass_impl() {
  static const PassImpl impl = [] {
    const char* e = getenv("CRYS_AXIS_PASS");
    if (e && strcmp(e, "valu") == 0) return PassImpl::Valu;
    if (e && strcmp(e, "radix") == 0) return PassImpl::Radix;
    if (e && strcmp(e, "mfma") == 0) return PassImpl::Mfma;
    return PassImpl::Auto;
  }();
  return impl;
}
}  // namespace crys

// One cached operator set: the three per-axis matrices for one ds value.
struct OpSet {
  double ds;
  void* dmat[3];     // device, element type of the solver's precision
  void* dmat_lo[3];  // FP32: bf16 bits of D - (float)D (split-precision product)
  void* dmat_c[3];   // half-size centrosymmetric form (M % 32 == 0; 3m: half-period pairing) or null
  void* dmat_c_lo[3];
  // FP32, half-size operators of length 64: the same values (and bf16 low
  // parts) in the order of the split phase's operator registers (crys::
  // load_operands_reg), or null
  void* dmat_reg[3] = {};
  void* dmat_reg_lo[3] = {};
  void* dmat_dL[3];  // dD/dL per axis, built on first use (cell stress)
  // whether the operator of axis a equals the z operator bit for bit (equal
  // lengths and cell lengths): a phase then loads one operator for both
  bool same_as_z[3] = {false, false, false};
  void* drad[3];     // radix-pass tables (radix_pass.hpp) when the radix passes are selected
  // 3m with translational parts: the separable terms of the folded operator
  // (crys::m3_glide_terms), per term and axis; their L derivatives on first use
  std::vector<std::array<void*, 3>> gmat, glo, gdL;
};

struct crys_solver {
  int kind = 0, precision = 0, flags = 0;
  size_t nlog[3] = {0, 0, 0}, nphy[3] = {0, 0, 0};
  size_t n_log = 0, n_phy = 0;
  double cell[6] = {4, 4, 4, 0, 0, 0};  // solver_base.h:43
  double ds_current = 0.0;              // solver_base.h:46
  double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  hipStream_t stream = nullptr;
  int device = 0;           // the HIP device current at creation
  double* iomat = nullptr;  // device, n_phy doubles
  void* work = nullptr;     // device, n_phy elements of the precision
  void* scratch = nullptr;  // device, n_phy elements of the precision
  std::vector<OpSet> ops;
  int cur = -1;
  // bumped whenever the operator set in use changes (new cell, new or other
  // cached ds): a captured iteration graph records the operators' addresses
  unsigned long ops_gen = 0;
  // 3m with non-zero translational parts: diffusion() is the sum of the
  // separable terms of OpSet::gmat on the reduced grid; glide_acc and
  // glide_tmp (n_phy elements of the precision) hold the partial sums
  bool glide = false;
  crys::M3Ops m3{};
  void* glide_acc = nullptr;
  void* glide_tmp = nullptr;
};

struct crys_scft {
  crys_solver* sv = nullptr;  // the solver the caller passed (not owned)
  double fA = 0, fB = 0, ds = 0, xN = 0, acc = 0, L[3] = {0, 0, 0};
  size_t NsA = 0, NsB = 0, Ns = 0, n = 0;
  double *wA = nullptr, *wB = nullptr, *ksi = nullptr, *phiA = nullptr, *phiB = nullptr;
  void *EA = nullptr, *EB = nullptr, *qf = nullptr, *qb = nullptr, *scr[2] = {nullptr, nullptr};
  double *qpart = nullptr, *epart = nullptr, *scal = nullptr, *trace = nullptr;
  double *gA = nullptr, *gB = nullptr, *spart = nullptr, *sums = nullptr;  // gradient, lazily
  int* counter = nullptr;
  // Contour storage, allocated on first use.  every == 0: qf and qb hold all
  // Ns+1 slots of each chain.  every > 0: ck holds the forward chain at the
  // slots 0, every, 2 every, ... (nck of them), and qf / qb hold one segment
  // of every+1 slots that is recomputed from its checkpoint.
  size_t every_req = 0, budget = 0;  // as requested (0: automatic, 0: free memory)
  size_t every = 0, nck = 0;
  void* ck = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // kGraphIters iterations in one graph (run() launches it while that many
  // remain, then the one-iteration graph)
  hipGraph_t graph_n = nullptr;
  hipGraphExec_t exec_n = nullptr;
  unsigned long graph_gen = 0;  // sv->ops_gen when the graphs were captured
  bool fields_set = false;
};

namespace {

size_t elem_size(int precision) { return precision == CRYS_FP32 ? sizeof(float) : sizeof(double); }

void free_ops(crys_solver* h) {
  for (auto& o : h->ops)
    for (int a = 0; a < 3; ++a) {
      if (o.dmat[a]) (void)hipFree(o.dmat[a]);
      if (o.dmat_lo[a]) (void)hipFree(o.dmat_lo[a]);
      if (o.dmat_c[a]) (void)hipFree(o.dmat_c[a]);
      if (o.dmat_c_lo[a]) (void)hipFree(o.dmat_c_lo[a]);
      if (o.dmat_reg[a]) (void)hipFree(o.dmat_reg[a]);
      if (o.dmat_reg_lo[a]) (void)hipFree(o.dmat_reg_lo[a]);
      if (o.dmat_dL[a]) (void)hipFree(o.dmat_dL[a]);
      if (o.drad[a]) (void)hipFree(o.drad[a]);
    }
  for (auto& o : h->ops)
    for (auto* v : {&o.gmat, &o.glo, &o.gdL})
      for (auto& t : *v)
        for (void* m : t)
          if (m) (void)hipFree(m);
  h->ops.clear();
  h->cur = -1;
  ++h->ops_gen;
}


// Uploads one operator in the solver's precision (and, for FP32, the bf16
// bits of its low part to *lo when lo is not null).
int upload_matrix(const crys_solver* h, const std::vector<double>& At, void** dst, void** lo) {
  const size_t cnt = At.size();
  HIPCHK(hipMalloc(dst, cnt * elem_size(h->precision)));
  if (h->precision == CRYS_FP32) {
    std::vector<float> f(At.begin(), At.end());
    HIPCHK(hipMemcpy(*dst, f.data(), cnt * sizeof(float), hipMemcpyHostToDevice));
    if (lo) {
      std::vector<unsigned short> b(cnt);
      for (size_t i = 0; i < cnt; ++i) b[i] = crys::bf16_bits((float)(At[i] - (double)f[i]));
      HIPCHK(hipMalloc(lo, cnt * sizeof(unsigned short)));
      HIPCHK(hipMemcpy(*lo, b.data(), cnt * sizeof(unsigned short), hipMemcpyHostToDevice));
    }
  } else {
    HIPCHK(hipMemcpy(*dst, At.data(), cnt * sizeof(double), hipMemcpyHostToDevice));
  }
  return CRYS_OK;
}

// Builds and uploads the three per-axis operators (or their L derivatives).
// For FP32 solvers `lo` (if not null) receives the bf16 bits of
// D - (float)D, the low part of the split-precision product.  When `cen` is
// not null it receives the half-size centrosymmetric forms (axes of length 32,
// 64 or 128; null elsewhere), paired by the half-period translation for 3m
// (centro_half).
// The register-order copy of a half-size FP32 operator of length 64 (At_c,
// 32 x 64): lane l of row tile jt holds, for k step n, the value that
// slab_load_operands reads for it, At_c[i * 64 + jt * 16 + l % 16] with
// i = 4 ord[n] + l / 16 (ord: the far-first order of row block jt % 2), at
// index (jt * 64 + l) * 8 + n; its bf16 low part at the same index.
int upload_register_order(const std::vector<double>& C, void** dst, void** dst_lo) {
  constexpr crys::FarFirstTable<8, 2> ff{};
  std::vector<float> a(4 * 64 * 8);
  std::vector<unsigned short> l(a.size());
  for (int jt = 0; jt < 4; ++jt)
    for (int lane = 0; lane < 64; ++lane)
      for (int n = 0; n < 8; ++n) {
        const int i = 4 * ff.v[jt % 2][n] + (lane >> 4);
        const double v = C[(size_t)i * 64 + jt * 16 + (lane & 15)];
        const float hi = (float)v;
        const size_t at = ((size_t)jt * 64 + lane) * 8 + n;
        a[at] = hi;
        l[at] = crys::bf16_bits((float)(v - (double)hi));
      }
  HIPCHK(hipMalloc(dst, a.size() * sizeof(float)));
  HIPCHK(hipMemcpy(*dst, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(dst_lo, l.size() * sizeof(unsigned short)));
  HIPCHK(hipMemcpy(*dst_lo, l.data(), l.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
  return CRYS_OK;
}

int upload_axis_ops(crys_solver* h, double ds, bool dL, void* dst[3], void* lo[3] = nullptr,
                    void* cen[3] = nullptr, void* cen_lo[3] = nullptr,
                    bool same_as_z[3] = nullptr, void* reg[3] = nullptr,
                    void* reg_lo[3] = nullptr) {
  std::vector<double> ops[3];
  for (int a = 2; a >= 0; --a) {
    std::vector<double>& At = ops[a];
    const int M = (int)h->nphy[a];
    if (h->kind == CRYS_PMMM_DCT) At = crys::pmmm_axis_operator(M, h->cell[a], ds, dL);
    else if (h->kind == CRYS_P1) At = crys::p1_axis_operator((int)h->nlog[a], h->cell[a], ds, dL);
    else At = crys::m3_axis_operator((int)h->nlog[a], h->cell[a], ds, dL);
    if (same_as_z) same_as_z[a] = At == ops[2];
    CHK(upload_matrix(h, At, &dst[a], lo ? &lo[a] : nullptr));
    if (cen && M % 32 == 0 && M <= 128 && M != 96 && crys::centro_enabled()) {
      std::vector<double> C = crys::centro_half(At, M, h->kind == CRYS_3M);
      CHK(upload_matrix(h, C, &cen[a], cen_lo ? &cen_lo[a] : nullptr));
      if (reg && reg_lo && M == 64 && h->precision == CRYS_FP32)
        CHK(upload_register_order(C, &reg[a], &reg_lo[a]));
    }
  }
  return CRYS_OK;
}

// Radix-pass kind of the solver (radix_pass.hpp).
int radix_kind(const crys_solver* h) {
  return h->kind == CRYS_P1 ? crys::kRadixP1 : (h->kind == CRYS_PMMM_DCT ? crys::kRadixDct
                                                                         : crys::kRadixM3);
}

// Whether axis `a` of this solver's pass-based contour steps runs the radix
// pass (see pass_impl).
bool radix_axis(const crys_solver* h, int a) {
  const int M = (int)h->nphy[a];
  switch (crys::pass_impl()) {
    case crys::PassImpl::Radix: return crys::radix_length_ok(M);
    case crys::PassImpl::Auto: return M == 128;
    default: return false;
  }
}

// The separable terms of a 3m solver with translational parts
// (crys::m3_glide_terms) on the device: values (and, for FP32, their bf16 low
// parts) or, with dL, their L derivatives.
int upload_glide_terms(crys_solver* h, double ds, bool dL, OpSet* o) {
  int N[3];
  for (int a = 0; a < 3; ++a) N[a] = (int)h->nlog[a];
  const std::vector<crys::M3Term> terms =
      crys::m3_glide_terms(h->m3.mask, h->m3.shift, N, h->cell, ds, dL);
  auto& dst = dL ? o->gdL : o->gmat;
  dst.assign(terms.size(), {nullptr, nullptr, nullptr});
  if (!dL) o->glo.assign(terms.size(), {nullptr, nullptr, nullptr});
  for (size_t t = 0; t < terms.size(); ++t)
    for (int a = 0; a < 3; ++a)
      CHK(upload_matrix(h, terms[t].At[a], &dst[t][a], dL ? nullptr : &o->glo[t][a]));
  return CRYS_OK;
}

int build_ops(crys_solver* h, double ds, OpSet* out) {
  out->ds = ds;
  for (int a = 0; a < 3; ++a)
    out->dmat[a] = out->dmat_lo[a] = out->dmat_c[a] = out->dmat_c_lo[a] = out->dmat_reg[a] =
        out->dmat_reg_lo[a] = out->dmat_dL[a] = out->drad[a] = nullptr;
  if (h->glide) return upload_glide_terms(h, ds, false, out);
  CHK(upload_axis_ops(h, ds, false, out->dmat, out->dmat_lo, out->dmat_c, out->dmat_c_lo,
                      out->same_as_z, out->dmat_reg, out->dmat_reg_lo));
  for (int a = 0; a < 3; ++a) {
    if (!radix_axis(h, a)) continue;
    const std::vector<double> t = crys::radix_table(radix_kind(h), (int)h->nphy[a], h->cell[a], ds);
    CHK(upload_matrix(h, t, &out->drad[a], nullptr));
  }
  return CRYS_OK;
}

// One contour step E D E (or a bare D with E == nullptr) for 1 or 2 chains:
// z pass in -> scr (pre E), y pass in place on scr, x pass scr -> out (post E).
//
// `stream` marks a step of a propagation, whose fields are read once by the
// next launch: the matrix-core passes and slab then read and write them
// non-temporally.  A lone diffusion() passes false (its field stays in L2).
// `only` >= 0 issues only that launch of the step (0-based, in issue order);
// *nlaunch (if not null) receives the number of launches of a whole step.
template <typename T>
int launch_step(const crys_solver* h, const OpSet& op, int nch, const T* const in[2],
                T* const out[2], T* const scr[2], const T* const E[2], hipStream_t s,
                bool stream, int only = -1, int* nlaunch = nullptr) {
  const int X = (int)h->nphy[0], Y = (int)h->nphy[1], Z = (int)h->nphy[2];
  crys::PassParams<T> p{};
  p.X = X;
  p.Y = Y;
  p.Z = Z;
  p.shift = h->kind == CRYS_3M;
  p.stream = stream;
  crys::RadixParams<T> rp{};
  rp.X = X;
  rp.Y = Y;
  rp.Z = Z;
  // one pass along `ax` with the chains' buffers `cc`: the radix pass when
  // the axis has a radix table, else the matrix-core pass
  auto pass = [&](int ax, const crys::PassChain<T>* cc) -> int {
    if (op.drad[ax]) {
      for (int c = 0; c < nch; ++c) rp.c[c] = cc[c];
      rp.tab = (const T*)op.drad[ax];
      HIPCHK(crys::launch_radix_pass<T>(ax, radix_kind(h), rp, nch, s));
      return CRYS_OK;
    }
    for (int c = 0; c < nch; ++c) p.c[c] = cc[c];
    p.cen = op.dmat_c[ax] != nullptr;
    p.mat = (const T*)(p.cen ? op.dmat_c[ax] : op.dmat[ax]);
    p.mat_lo = (const unsigned short*)(p.cen ? op.dmat_c_lo[ax] : op.dmat_lo[ax]);
    HIPCHK(crys::launch_axis_pass<T>(ax, p, nch, s));
    return CRYS_OK;
  };
  crys::PassChain<T> cc[2];
  int k = 0;
  if (op.drad[1] && op.drad[2] && crys::radix_slab_applies(Y, Z) && crys::slab_enabled()) {
    // z and y radix transforms fused on x-slabs (both axes of length 128)
    crys::RadixSlabParams<T> rs{};
    rs.X = X;
    rs.Y = Y;
    rs.Z = Z;
    rs.tab_z = (const T*)op.drad[2];
    rs.tab_y = (const T*)op.drad[1];
    for (int c = 0; c < nch; ++c) rs.c[c] = {in[c], scr[c], E[c], nullptr};
    if (only < 0 || only == k) HIPCHK(crys::launch_radix_slab<T>(radix_kind(h), rs, nch, s));
    ++k;
  } else if (crys::slab_applies(Y, Z) && !op.drad[1] && !op.drad[2]) {
    // z and y fused on x-slabs when the two axes share a length class; the
    // half-size centrosymmetric operators when both axes have them
    crys::SlabParams<T> sp{};
    sp.X = X;
    sp.Y = Y;
    sp.Z = Z;
    sp.shift = h->kind == CRYS_3M;
    sp.stream = stream;
    const bool cen_zy = op.dmat_c[1] && op.dmat_c[2] && Y == Z;
    for (int c = 0; c < nch; ++c) sp.c[c] = {in[c], scr[c], E[c], nullptr};
    sp.mat_z = (const T*)(cen_zy ? op.dmat_c[2] : op.dmat[2]);
    sp.mat_y = (const T*)(cen_zy ? op.dmat_c[1] : op.dmat[1]);
    sp.lo_z = (const unsigned short*)(cen_zy ? op.dmat_c_lo[2] : op.dmat_lo[2]);
    sp.lo_y = (const unsigned short*)(cen_zy ? op.dmat_c_lo[1] : op.dmat_lo[1]);
    if (only < 0 || only == k) HIPCHK(crys::launch_slab_zy<T>(sp, nch, cen_zy, s));
    ++k;
  } else {
    // z pass in -> scr (pre E), y pass in place on scr
    for (int ax = 2; ax >= 1; --ax, ++k) {
      if (only >= 0 && only != k) continue;
      for (int c = 0; c < nch; ++c)
        cc[c] = ax == 2 ? crys::PassChain<T>{in[c], scr[c], E[c], nullptr}
                        : crys::PassChain<T>{scr[c], scr[c], nullptr, nullptr};
      CHK(pass(ax, cc));
    }
  }
  // x pass scr -> out (post E)
  if (only < 0 || only == k) {
    for (int c = 0; c < nch; ++c) cc[c] = {scr[c], out[c], nullptr, E[c]};
    CHK(pass(0, cc));
  }
  ++k;
  if (nlaunch) *nlaunch = k;
  return CRYS_OK;
}

// One contour step of a 3m solver with translational parts for 1 or 2
// chains: out = E_post (.) sum_t X_t (x) Y_t (x) Z_t (E_pre (.) in) over the
// separable terms of OpSet::gmat.  Term t is a z pass (in -> tmp, times
// E_pre), a y pass in place and an x pass; a single term writes out directly,
// else the first writes acc, each further one tmp2, summed into acc and, for
// the last term, into out (times E_post).  in may equal out: every term reads
// in before out is written.  E_pre, E_post or their entries may be null.
// dL_axis >= 0 applies the L derivative of every term's factor on that axis
// (the stress).
template <typename T>
int launch_glide_step(const crys_solver* h, const OpSet& op, int nch, const T* const in[2],
                      T* const out[2], T* const tmp[2], T* const acc[2], T* const tmp2[2],
                      const T* const E_pre[2], const T* const E_post[2], hipStream_t s,
                      bool stream, int dL_axis = -1) {
  crys::PassParams<T> p{};
  p.X = (int)h->nphy[0];
  p.Y = (int)h->nphy[1];
  p.Z = (int)h->nphy[2];
  p.stream = stream;
  const size_t nt = op.gmat.size();
  const size_t n = h->n_phy;
  if (nt == 0 || (dL_axis >= 0 && op.gdL.size() != nt))
    return fail(CRYS_ERR_HIP, "3m operator terms are missing");
  for (size_t t = 0; t < nt; ++t) {
    const bool last = t + 1 == nt;
    for (int k = 0; k < 3; ++k) {
      const int ax = 2 - k;  // z, y, x
      for (int c = 0; c < nch; ++c) {
        T* xo = nt == 1 ? out[c] : (t == 0 ? acc[c] : tmp2[c]);
        const T* ep = E_pre ? E_pre[c] : nullptr;
        const T* eq = (nt == 1 && E_post) ? E_post[c] : nullptr;
        p.c[c] = k == 0   ? crys::PassChain<T>{in[c], tmp[c], ep, nullptr}
                 : k == 1 ? crys::PassChain<T>{tmp[c], tmp[c], nullptr, nullptr}
                          : crys::PassChain<T>{tmp[c], xo, nullptr, eq};
      }
      const bool d = ax == dL_axis;
      p.mat = (const T*)(d ? op.gdL[t][ax] : op.gmat[t][ax]);
      p.mat_lo = d ? nullptr : (const unsigned short*)op.glo[t][ax];
      p.cen = 0;
      HIPCHK(crys::launch_axis_pass<T>(ax, p, nch, s));
    }
    if (t >= 1) {
      crys::GlideSum<T> g{};
      for (int c = 0; c < nch; ++c)
        g.c[c] = {last ? out[c] : acc[c], acc[c], tmp2[c],
                  (last && E_post) ? E_post[c] : nullptr};
      hipLaunchKernelGGL((crys::glide_sum_kernel<T>), dim3(grid_for(n), (unsigned)nch), dim3(256),
                         0, s, g, n);
      HIPCHK(hipGetLastError());
    }
  }
  return CRYS_OK;
}

// Length class of the one-launch phase path for this solver (0: not used):
// the matrix-core plane phases when all three axes share a length class up
// to 64, the radix plane phase when all three axes have length 128 and take
// the radix pass.
int phase_mp(const crys_solver* h) {
  if (!crys::phase_enabled()) return 0;
  if (h->glide)  // the multi-term phases: one length class up to 64 (dense products)
    return crys::phase_class((int)h->nphy[0], (int)h->nphy[1], (int)h->nphy[2]);
  if (h->nphy[0] == 128 && h->nphy[1] == 128 && h->nphy[2] == 128 && radix_axis(h, 0) &&
      radix_axis(h, 1) && radix_axis(h, 2))
    return 128;
  return crys::phase_class((int)h->nphy[0], (int)h->nphy[1], (int)h->nphy[2]);
}

// Phase k (0..Ns) of both chains of an SCFT propagation (see plane_phase.hpp):
// phase 0 starts step 0 from q_0, phase k finishes step k-1 (history entry k)
// and starts step k, phase Ns only finishes.  Even phases work on x-slabs
// (row axis y), odd phases on y-slabs (row axis x).  Forward chain: q_f[t]
// -> q_f[t+1] with E_A for t < NsA; backward chain: q_b[Ns-t] -> q_b[Ns-t-1]
// with E_B while Ns-t-1 >= NsA.
//
// The chains may also cover a part of the contour: qf / qb then hold slots
// t0 .. t0+Ns (a checkpoint segment), NsA stays the global junction slot, and
// nch = 1 advances the forward chain alone.
template <typename T>
int launch_phase_step(const crys_solver* h, const OpSet& op, size_t k, size_t Ns, size_t NsA,
                      T* qf, T* qb, T* const scr[2], const T* EA, const T* EB, hipStream_t s,
                      int nch = 2, size_t t0 = 0) {
  const int X = (int)h->nphy[0], Y = (int)h->nphy[1], Z = (int)h->nphy[2];
  const size_t n = (size_t)X * Y * Z;
  const int mp = phase_mp(h);
  const bool cen = op.dmat_c[0] && op.dmat_c[1] && op.dmat_c[2] && X == mp && Y == mp && Z == mp;
  const bool yslab = (k % 2) == 1;
  crys::PhaseParams<T> p{};
  p.shift = h->kind == CRYS_3M;
  const int ra = yslab ? 0 : 1;  // row axis
  p.mat_r = (const T*)(cen ? op.dmat_c[ra] : op.dmat[ra]);
  p.mat_c = (const T*)(cen ? op.dmat_c[2] : op.dmat[2]);
  p.lo_r = (const unsigned short*)(cen ? op.dmat_c_lo[ra] : op.dmat_lo[ra]);
  p.lo_c = (const unsigned short*)(cen ? op.dmat_c_lo[2] : op.dmat_lo[2]);
  p.same_rc = op.same_as_z[ra] ? 1 : 0;
  if (cen && op.dmat_reg[ra] && op.dmat_reg[2]) {
    p.reg_r = (const T*)op.dmat_reg[ra];
    p.reg_c = (const T*)op.dmat_reg[2];
    p.reg_lo_r = (const unsigned short*)op.dmat_reg_lo[ra];
    p.reg_lo_c = (const unsigned short*)op.dmat_reg_lo[2];
  }
  p.MR = yslab ? X : Y;
  p.MC = Z;
  p.plane_stride = yslab ? (long)Z : (long)Y * Z;
  p.row_stride = yslab ? (long)Y * Z : (long)Z;
  const int nplanes = yslab ? Y : X;
  auto Ef = [&](size_t t) { return t0 + t < NsA ? EA : EB; };
  auto Eb = [&](size_t t) { return t0 + Ns - t - 1 >= NsA ? EB : EA; };
  crys::PhaseChain<T>& f = p.c[0];
  crys::PhaseChain<T>& b = p.c[1];
  // phase k writes scratch half k % 2 and reads the half phase k-1 wrote; a
  // 3m solver with translational parts keeps one plane per term in each half
  const size_t nt = h->glide ? op.gmat.size() : 1;
  const size_t wr = (k % 2) * nt * n, rd = ((k + 1) % 2) * nt * n;
  f.in = k == 0 ? qf : scr[0] + rd;
  b.in = k == 0 ? qb + Ns * n : scr[1] + rd;
  if (k >= 1) {
    f.hist = qf + k * n;
    f.e_fin = Ef(k - 1);
    b.hist = qb + (Ns - k) * n;
    b.e_fin = Eb(k - 1);
  }
  if (k < Ns) {
    f.out = scr[0] + wr;
    f.e_next = Ef(k);
    b.out = scr[1] + wr;
    b.e_next = Eb(k);
  }
  if (h->glide) {
    crys::PhaseTermsParams<T> tp{};
    tp.c[0] = p.c[0];
    tp.c[1] = p.c[1];
    tp.tstride = (long)n;
    tp.nt = (int)nt;
    for (size_t t = 0; t < nt && t < (size_t)crys::kMaxGlideTerms; ++t) {
      tp.mat_r[t] = (const T*)op.gmat[t][ra];
      tp.mat_c[t] = (const T*)op.gmat[t][2];
      tp.lo_r[t] = (const unsigned short*)op.glo[t][ra];
      tp.lo_c[t] = (const unsigned short*)op.glo[t][2];
    }
    tp.MR = p.MR;
    tp.MC = p.MC;
    tp.plane_stride = p.plane_stride;
    tp.row_stride = p.row_stride;
    HIPCHK(crys::launch_phase_terms<T>(tp, nplanes, nch, mp, s));
    return CRYS_OK;
  }
  if (mp == 128) {
    p.tab_r = (const T*)op.drad[ra];
    p.tab_c = (const T*)op.drad[2];
    p.rkind = radix_kind(h);
    HIPCHK(crys::launch_radix_phase<T>(p, nplanes, nch, s));
    return CRYS_OK;
  }
  HIPCHK(crys::launch_phase<T>(p, nplanes, nch, mp, cen, s));
  return CRYS_OK;
}

int check_handle(const crys_solver* h) {
  if (!h) return fail(CRYS_ERR_INVALID_ARGUMENT, "null solver handle");
  return CRYS_OK;
}

int diffusion_impl(crys_solver* h) {
  if (h->cur < 0) return fail(CRYS_ERR_NO_CONTOUR_STEP, "diffusion() before set_contour_step()");
  const OpSet& op = h->ops[h->cur];
  const size_t n = h->n_phy;
  if (h->precision == CRYS_FP64) {
    double* io = h->iomat;
    const double* in[2] = {io, nullptr};
    double* out[2] = {io, nullptr};
    double* scr[2] = {io, nullptr};
    const double* E[2] = {nullptr, nullptr};
    if (h->glide) {
      double* tmp[2] = {(double*)h->scratch, nullptr};
      double* acc[2] = {(double*)h->glide_acc, nullptr};
      double* tmp2[2] = {(double*)h->glide_tmp, nullptr};
      return launch_glide_step<double>(h, op, 1, in, out, tmp, acc, tmp2, nullptr, nullptr,
                                       h->stream, false);
    }
    return launch_step<double>(h, op, 1, in, out, scr, E, h->stream, false);
  }
  float* w = (float*)h->work;
  hipLaunchKernelGGL((crys::convert_kernel<float, double>), dim3(grid_for(n)), dim3(256), 0,
                     h->stream, w, (const double*)h->iomat, n);
  HIPCHK(hipGetLastError());
  const float* in[2] = {w, nullptr};
  float* out[2] = {w, nullptr};
  float* scr[2] = {w, nullptr};
  const float* E[2] = {nullptr, nullptr};
  if (h->glide) {
    float* tmp[2] = {(float*)h->scratch, nullptr};
    float* acc[2] = {(float*)h->glide_acc, nullptr};
    float* tmp2[2] = {(float*)h->glide_tmp, nullptr};
    CHK(launch_glide_step<float>(h, op, 1, in, out, tmp, acc, tmp2, nullptr, nullptr, h->stream,
                                 false));
  } else {
    CHK(launch_step<float>(h, op, 1, in, out, scr, E, h->stream, false));
  }
  hipLaunchKernelGGL((crys::convert_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0,
                     h->stream, h->iomat, (const float*)w, n);
  HIPCHK(hipGetLastError());
  return CRYS_OK;
}

}  // namespace

namespace {

// solve_diffusion_equation on device buffers; `scr` (n doubles) stages the
// half-finished step between the passes.
int propagate_impl(crys_solver_t* h, const double* E, const double* init, size_t Ns,
                   int direction, double* q, double* scr) {
  CHK(check_handle(h));
  if (!E || !init || !q) return fail(CRYS_ERR_INVALID_ARGUMENT, "null device buffer");
  if (direction != 1 && direction != -1)
    return fail(CRYS_ERR_INVALID_ARGUMENT, "Wrong direction");  // crysSCFT_sample.cpp:381
  if (h->precision != CRYS_FP64)
    return fail(CRYS_ERR_UNSUPPORTED, "crys_propagate takes FP64 histories; use an FP64 solver");
  const size_t n = h->n_phy;
  if (h->cur < 0) return fail(CRYS_ERR_NO_CONTOUR_STEP, "propagate before set_contour_step()");
  const OpSet& op = h->ops[h->cur];
  size_t s = direction == 1 ? 0 : Ns;
  HIPCHK(hipMemcpyAsync(q + s * n, init, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  for (size_t it = 0; it < Ns; ++it) {
    const size_t sn = direction == 1 ? s + 1 : s - 1;
    const double* in[2] = {q + s * n, nullptr};
    double* out[2] = {q + sn * n, nullptr};
    double* sc[2] = {scr, nullptr};
    const double* Ee[2] = {E, nullptr};
    if (h->glide) {
      double* acc[2] = {(double*)h->glide_acc, nullptr};
      double* tmp2[2] = {(double*)h->glide_tmp, nullptr};
      CHK(launch_glide_step<double>(h, op, 1, in, out, sc, acc, tmp2, Ee, Ee, h->stream, true));
    } else {
      CHK(launch_step<double>(h, op, 1, in, out, sc, Ee, h->stream, true));
    }
    s = sn;
  }
  HIPCHK(hipMemcpyAsync(h->iomat, q + s * n, n * sizeof(double), hipMemcpyDeviceToDevice,
                        h->stream));
  return CRYS_OK;
}

}  // namespace

// Staging field of crys_propagate_noalloc: the half-finished contour step
// between the axis passes, one field of the largest physical grid a solver
// accepts (every axis at most 128 points: 2^21 doubles, 16 MiB).  Each GPU
// the library runs on holds its own copy until the process ends; every call
// on one GPU shares it.  Host code uses its device address only (per device,
// from hipGetSymbolAddress), never the name.
constexpr size_t kNoAllocMaxPoints = (size_t)1 << 21;
static __device__ __align__(256) double scratch[kNoAllocMaxPoints];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on device `dev`, looked up on the first
// call for that device with that device current, then kept.
int noalloc_scratch(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(CRYS_ERR_UNSUPPORTED, "device %d is beyond the %d the library tracks", dev,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(CRYS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return CRYS_OK;
}

}  // namespace

extern "C" {

const char* crys_last_error(void) { return g_err.c_str(); }

const char* crys_version(void) { return "crysscft_sample_amd 0.1 (gfx950)"; }

int crys_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int crys_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return CRYS_OK;
}

int crys_solver_create(int kind, const size_t grid_log[3], const double translational_part[9],
                       int precision, int flags, crys_solver_t** out) {
  if (!out || !grid_log) return fail(CRYS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (kind != CRYS_P1 && kind != CRYS_3M && kind != CRYS_PMMM_DCT)
    return fail(CRYS_ERR_INVALID_ARGUMENT, "unknown solver kind %d", kind);
  if (precision != CRYS_FP64 && precision != CRYS_FP32)
    return fail(CRYS_ERR_INVALID_ARGUMENT, "unknown precision %d", precision);
  const size_t Nx = grid_log[0], Ny = grid_log[1], Nz = grid_log[2];
  if (kind == CRYS_P1) {
    // solver_p1.h:101-108
    if (Nx == 0 || Ny == 0 || Nz == 0)
      return fail(CRYS_ERR_BAD_GRID, "ERROR: Bad Grid Size For Group Triclinic P1");
    if (Nz % 8 != 0)
      return fail(CRYS_ERR_BAD_GRID,
                  "ERROR: Bad Grid Size For Group Triclinic P1, alignment cannot be guaranteed");
  } else {
    // solver_3m.h:81-92, solver_Pmmm.h:71-82
    if (Nx % 2 || Ny % 2 || Nz % 2)
      return fail(CRYS_ERR_BAD_GRID, "ERROR: Grid Size Must Be Even For %s",
                  kind == CRYS_3M ? "Generalized 3m symmetry" : "Group Pmmm");
    if (Nx == 0 || Ny == 0 || Nz == 0)
      return fail(CRYS_ERR_BAD_GRID, "ERROR: Bad Grid Size For %s",
                  kind == CRYS_3M ? "Generalized 3m symmetry" : "Group Pmmm");
    if ((Nz / 2) % 8 != 0)
      return fail(CRYS_ERR_BAD_GRID, "ERROR: Bad Grid Size For %s, Alignment Cannot Be Guaranteed",
                  kind == CRYS_3M ? "Generalized 3m symmetry" : "Group Pmmm");
  }
  crys_solver* h = new (std::nothrow) crys_solver();
  if (!h) return fail(CRYS_ERR_OUT_OF_MEMORY, "host allocation failed");
  h->kind = kind;
  h->precision = precision;
  h->flags = flags;
  for (int a = 0; a < 3; ++a) {
    h->nlog[a] = grid_log[a];
    h->nphy[a] = kind == CRYS_P1 ? grid_log[a] : grid_log[a] / 2;
  }
  if (kind == CRYS_3M && translational_part)
    for (int i = 0; i < 9; ++i) h->g[i] = translational_part[i];
  h->n_log = Nx * Ny * Nz;
  h->n_phy = h->nphy[0] * h->nphy[1] * h->nphy[2];
  if (h->n_phy >= (size_t)1 << 31) {
    delete h;
    return fail(CRYS_ERR_UNSUPPORTED, "physical grid of %zu points exceeds 2^31", h->n_phy);
  }
  for (int a = 0; a < 3; ++a)
    if (crys::pass_class((int)h->nphy[a]) == 0) {
      delete h;
      return fail(CRYS_ERR_UNSUPPORTED, "physical grid length %zu exceeds 128 on axis %d",
                  grid_log[a], a);
    }
  const bool glide = kind == CRYS_3M && crys::m3_nonzero(h->g);
  if (glide) {
    const int r = crys::m3_build_ops(h->g, h->nlog, &h->m3);
    if (r != 0) {
      delete h;
      return fail(CRYS_ERR_UNSUPPORTED,
                  r == 1 ? "Solver3m: a translational part is not a whole number of grid points"
                         : "Solver3m: the symmetry operations do not map the 2x2y2z sub-lattice "
                           "onto the grid");
    }
  }
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&h->iomat, h->n_phy * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->work, h->n_phy * elem_size(precision));
  if (e == hipSuccess) e = hipMalloc(&h->scratch, h->n_phy * elem_size(precision));
  h->glide = glide;
  if (glide && e == hipSuccess) e = hipMalloc(&h->glide_acc, h->n_phy * elem_size(precision));
  if (glide && e == hipSuccess) e = hipMalloc(&h->glide_tmp, h->n_phy * elem_size(precision));
  if (e == hipSuccess) e = hipMemset(h->iomat, 0, h->n_phy * sizeof(double));
  if (e != hipSuccess) {
    crys_solver_destroy(h);
    return fail(e == hipErrorOutOfMemory ? CRYS_ERR_OUT_OF_MEMORY : CRYS_ERR_HIP,
                "solver allocation failed: %s", hipGetErrorString(e));
  }
  *out = h;
  return CRYS_OK;
}

int crys_solver_destroy(crys_solver_t* h) {
  if (!h) return CRYS_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_ops(h);
  if (h->glide_acc) (void)hipFree(h->glide_acc);
  if (h->glide_tmp) (void)hipFree(h->glide_tmp);
  if (h->iomat) (void)hipFree(h->iomat);
  if (h->work) (void)hipFree(h->work);
  if (h->scratch) (void)hipFree(h->scratch);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CRYS_OK;
}

int crys_set_cell_para(crys_solver_t* h, const double c[6]) {
  CHK(check_handle(h));
  if (!c) return fail(CRYS_ERR_INVALID_ARGUMENT, "null cell_para");
  if (h->kind != CRYS_P1) {
    // solver_3m.h:117, solver_Pmmm.h:100
    if (c[0] == h->cell[0] && c[1] == h->cell[1] && c[2] == h->cell[2] && c[3] == 0 && c[4] == 0 &&
        c[5] == 0)
      return CRYS_OK;
  }
  if (!(c[0] > 0 && c[1] > 0 && c[2] > 0))
    return fail(CRYS_ERR_INVALID_ARGUMENT, "cell lengths must be positive");
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->kind == CRYS_PMMM_DCT && !(h->flags & CRYS_FLAG_PMMM_ORTHORHOMBIC)) {
    const double nc[6] = {c[0], c[0], c[0], 0, 0, 0};  // solver_Pmmm.h:101
    memcpy(h->cell, nc, sizeof(nc));
  } else {
    const double nc[6] = {c[0], c[1], c[2], 0, 0, 0};
    memcpy(h->cell, nc, sizeof(nc));
  }
  free_ops(h);
  h->ds_current = 0.0;
  return CRYS_OK;
}

int crys_set_contour_step(crys_solver_t* h, double ds) {
  CHK(check_handle(h));
  if (!(ds > 0) || !std::isfinite(ds)) return fail(CRYS_ERR_INVALID_ARGUMENT, "ds must be > 0");
  if (!h->ops.empty() && ds == h->ds_current && h->cur >= 0) return CRYS_OK;
  h->ds_current = ds;
  for (size_t i = 0; i < h->ops.size(); ++i)
    if (h->ops[i].ds == ds) {
      if (h->cur != (int)i) ++h->ops_gen;
      h->cur = (int)i;
      return CRYS_OK;
    }
  OpSet o;
  int r = build_ops(h, ds, &o);
  if (r != CRYS_OK) {
    for (int a = 0; a < 3; ++a) {
      if (o.dmat[a]) (void)hipFree(o.dmat[a]);
      if (o.dmat_lo[a]) (void)hipFree(o.dmat_lo[a]);
      if (o.dmat_c[a]) (void)hipFree(o.dmat_c[a]);
      if (o.dmat_c_lo[a]) (void)hipFree(o.dmat_c_lo[a]);
      if (o.dmat_reg[a]) (void)hipFree(o.dmat_reg[a]);
      if (o.dmat_reg_lo[a]) (void)hipFree(o.dmat_reg_lo[a]);
      if (o.drad[a]) (void)hipFree(o.drad[a]);
    }
    for (auto* v : {&o.gmat, &o.glo})
      for (auto& t : *v)
        for (void* m : t)
          if (m) (void)hipFree(m);
    return r;
  }
  h->ops.push_back(o);
  h->cur = (int)h->ops.size() - 1;
  ++h->ops_gen;
  return CRYS_OK;
}

int crys_diffusion(crys_solver_t* h) {
  CHK(check_handle(h));
  return diffusion_impl(h);
}

int crys_diffusion_time(crys_solver_t* h, int reps, double* ms_per_call) {
  CHK(check_handle(h));
  if (reps <= 0 || !ms_per_call) return fail(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  CHK(diffusion_impl(h));  // warm
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int r = CRYS_OK;
  (void)hipEventRecord(e0, h->stream);
  for (int i = 0; i < reps && r == CRYS_OK; ++i) r = diffusion_impl(h);
  (void)hipEventRecord(e1, h->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (r == CRYS_OK) *ms_per_call = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return r;
}

int crys_get_cell_para(const crys_solver_t* h, double out[6]) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  memcpy(out, h->cell, sizeof(h->cell));
  return CRYS_OK;
}
int crys_get_grid_size_log(const crys_solver_t* h, size_t out[3]) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  for (int a = 0; a < 3; ++a) out[a] = h->nlog[a];
  return CRYS_OK;
}
int crys_get_grid_size_phy(const crys_solver_t* h, size_t out[3]) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  for (int a = 0; a < 3; ++a) out[a] = h->nphy[a];
  return CRYS_OK;
}
int crys_get_space_size_log(const crys_solver_t* h, size_t* out) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  *out = h->n_log;
  return CRYS_OK;
}
int crys_get_space_size_phy(const crys_solver_t* h, size_t* out) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  *out = h->n_phy;
  return CRYS_OK;
}
int crys_get_ds(const crys_solver_t* h, double* out) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  *out = h->ds_current;
  return CRYS_OK;
}
int crys_get_iomat_for_diffusion(const crys_solver_t* h, double** out) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  *out = h->iomat;
  return CRYS_OK;
}
int crys_get_kind(const crys_solver_t* h, int* out) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  *out = h->kind;
  return CRYS_OK;
}
int crys_get_precision(const crys_solver_t* h, int* out) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  *out = h->precision;
  return CRYS_OK;
}

int crys_iomat_upload(crys_solver_t* h, const double* host) {
  CHK(check_handle(h));
  if (!host) return fail(CRYS_ERR_INVALID_ARGUMENT, "null host buffer");
  HIPCHK(hipMemcpyAsync(h->iomat, host, h->n_phy * sizeof(double), hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return CRYS_OK;
}

int crys_iomat_download(crys_solver_t* h, double* host) {
  CHK(check_handle(h));
  if (!host) return fail(CRYS_ERR_INVALID_ARGUMENT, "null host buffer");
  HIPCHK(hipMemcpyAsync(host, h->iomat, h->n_phy * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return CRYS_OK;
}

int crys_solver_synchronize(crys_solver_t* h) {
  CHK(check_handle(h));
  HIPCHK(hipStreamSynchronize(h->stream));
  return CRYS_OK;
}

int crys_solver_get_stream(const crys_solver_t* h, void** out) {
  CHK(check_handle(h));
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  *out = (void*)h->stream;
  return CRYS_OK;
}

int crys_propagate(crys_solver_t* h, const double* E, const double* init, size_t Ns,
                   int direction, double* q) {
  CHK(check_handle(h));
  return propagate_impl(h, E, init, Ns, direction, q, (double*)h->scratch);
}

int crys_propagate_noalloc(crys_solver_t* h, const double* E, const double* init, size_t Ns,
                           int direction, double* q) {
  CHK(check_handle(h));
  if (h->n_phy > kNoAllocMaxPoints)
    return fail(CRYS_ERR_UNSUPPORTED,
                "crys_propagate_noalloc: %zu physical points exceed its static staging field of "
                "%zu points", h->n_phy, kNoAllocMaxPoints);
  double* scr = nullptr;
  CHK(noalloc_scratch(h->device, &scr));
  return propagate_impl(h, E, init, Ns, direction, q, scr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// SCFT iteration
// ---------------------------------------------------------------------------
namespace {

void drop_graphs(crys_scft* h) {
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->exec_n) (void)hipGraphExecDestroy(h->exec_n);
  if (h->graph_n) (void)hipGraphDestroy(h->graph_n);
  h->exec = nullptr;
  h->graph = nullptr;
  h->exec_n = nullptr;
  h->graph_n = nullptr;
}

void scft_free(crys_scft* h) {
  drop_graphs(h);
  void* bufs[] = {h->wA,    h->wB,   h->ksi,   h->phiA,    h->phiB,   h->EA,
                  h->EB,    h->qf,   h->qb,    h->scr[0],  h->scr[1], h->qpart,
                  h->epart, h->scal, h->trace, h->counter, h->gA,     h->gB,
                  h->spart, h->sums,  h->ck};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
}

// Advances the chains over c contour steps of a segment whose slot 0 is the
// global slot t0: q_f[k] -> q_f[k+1] in qf and q_b[c-k] -> q_b[c-k-1] in qb,
// from qf[0] and qb[c] (nch = 1: the forward chain alone).  The whole
// contour is the segment t0 = 0, c = Ns.
template <typename T>
int propagate_segment(crys_scft* h, hipStream_t s, int nch, T* qf, T* qb, size_t c, size_t t0) {
  const crys_solver* sv = h->sv;
  const OpSet& op = sv->ops[sv->cur];
  const size_t n = h->n, NsA = h->NsA;
  T* scr[2] = {(T*)h->scr[0], (T*)h->scr[1]};
  const T* EA = (const T*)h->EA;
  const T* EB = (const T*)h->EB;
  if (phase_mp(sv)) {
    for (size_t k = 0; k <= c; ++k)
      CHK(launch_phase_step<T>(sv, op, k, c, NsA, qf, qb, scr, EA, EB, s, nch, t0));
  } else {
    // a 3m solver with translational parts: scr holds 3 n per chain
    T* acc[2] = {scr[0] + n, scr[1] + n};
    T* tmp2[2] = {scr[0] + 2 * n, scr[1] + 2 * n};
    for (size_t t = 0; t < c; ++t) {
      const size_t bo = c - t - 1;
      const T* in[2] = {qf + t * n, qb + (c - t) * n};
      T* out[2] = {qf + (t + 1) * n, qb + bo * n};
      const T* E[2] = {t0 + t < NsA ? EA : EB, t0 + bo >= NsA ? EB : EA};
      if (sv->glide)
        CHK(launch_glide_step<T>(sv, op, nch, in, out, scr, acc, tmp2, E, E, s, true));
      else
        CHK(launch_step<T>(sv, op, nch, in, out, scr, E, s, true));
    }
  }
  return CRYS_OK;
}

// d Q / d L_a for the contour steps t_begin .. t_end-1: for every step t and
// axis a, < E_t q^dagger_{t+1}, (dD_a (x) D_b (x) D_c) E_t q_t > with the
// derivative operator on axis a.  qf and qb hold the slots from slot0 on.
// Partials go to h->spart[(3 t + a) * nb + block].
template <typename T>
int scft_record_stress(crys_scft* h, hipStream_t s, const T* qf, const T* qb, size_t t_begin,
                       size_t t_end, size_t slot0) {
  const crys_solver* sv = h->sv;
  const OpSet& op = sv->ops[sv->cur];
  const size_t n = h->n, NsA = h->NsA;
  T* w = (T*)h->scr[0];
  const int nb = crys::red_blocks(n);
  crys::PassParams<T> p{};
  p.X = (int)sv->nphy[0];
  p.Y = (int)sv->nphy[1];
  p.Z = (int)sv->nphy[2];
  for (size_t t = t_begin; t < t_end; ++t) {
    const T* E = (const T*)(t < NsA ? h->EA : h->EB);
    const T* q = qf + (t - slot0) * n;
    const T* qd = qb + (t + 1 - slot0) * n;
    if (sv->glide) {
      // the derivative of every separable term on axis a, summed (scr[0]
      // holds 3 n: the term in flight, the sum, the next term)
      T* tmp[2] = {w, nullptr};
      T* acc[2] = {w + n, nullptr};
      T* tmp2[2] = {w + 2 * n, nullptr};
      const T* in[2] = {q, nullptr};
      const T* Ep[2] = {E, nullptr};
      for (int a = 0; a < 3; ++a) {
        CHK(launch_glide_step<T>(sv, op, 1, in, acc, tmp, acc, tmp2, Ep, nullptr, s, false, a));
        hipLaunchKernelGGL((crys::dot3_partial_kernel<T>), dim3(nb), dim3(crys::kRedThreads), 0,
                           s, (const T*)(w + n), E, qd, h->spart + (3 * t + a) * nb, n);
        HIPCHK(hipGetLastError());
      }
      continue;
    }
    for (int a = 0; a < 3; ++a) {
      // z pass (pre E), then y, then x; the axis a pass uses dD/dL_a
      const int order[3] = {2, 1, 0};
      for (int k = 0; k < 3; ++k) {
        const int ax = order[k];
        p.c[0] = {k == 0 ? q : w, w, k == 0 ? E : nullptr, nullptr};
        p.mat = (const T*)(ax == a ? op.dmat_dL[ax] : op.dmat[ax]);
        p.mat_lo = ax == a ? nullptr : (const unsigned short*)op.dmat_lo[ax];
        HIPCHK(crys::launch_axis_pass<T>(ax, p, 1, s));
      }
      hipLaunchKernelGGL((crys::dot3_partial_kernel<T>), dim3(nb), dim3(crys::kRedThreads), 0, s,
                         (const T*)w, E, qd, h->spart + (3 * t + a) * nb, n);
      HIPCHK(hipGetLastError());
    }
  }
  return CRYS_OK;
}

// The propagation and contour sums of one iteration with checkpoints every
// C = h->every slots (sqrt(Ns) checkpointing): a forward sweep keeps q_f at
// the slots j C; then, last segment first, each segment [s0, s1] recomputes
// q_f from its checkpoint while the backward chain advances over it, and
// phi_accumulate_kernel adds the segment's slots to the contour sums (held in
// phiA / phiB).  One propagation of the forward chain more than the plain
// iteration; storage nck + 2 (C + 1) fields instead of 2 (Ns + 1).
template <typename T>
int record_checkpointed(crys_scft* h, hipStream_t s, bool stress) {
  const size_t n = h->n, Ns = h->Ns, C = h->every, J = h->nck;
  T* F = (T*)h->qf;
  T* B = (T*)h->qb;
  T* ck = (T*)h->ck;
  const unsigned g = grid_for(n);
  auto copy = [&](T* dst, const T* src) -> int {
    hipLaunchKernelGGL((crys::copy_kernel<T>), dim3(g), dim3(256), 0, s, dst, src, n);
    HIPCHK(hipGetLastError());
    return CRYS_OK;
  };
  for (size_t j = 0; j + 1 < J; ++j) {
    CHK(copy(F, ck + j * n));
    CHK(propagate_segment<T>(h, s, 1, F, B, C, j * C));
    CHK(copy(ck + (j + 1) * n, F + C * n));
  }
  for (size_t j = J; j-- > 0;) {
    const size_t s0 = j * C, s1 = std::min(s0 + C, Ns), c = s1 - s0;
    const bool last = j + 1 == J;
    CHK(copy(F, ck + j * n));
    if (last) {
      hipLaunchKernelGGL((crys::fill_kernel<T>), dim3(g), dim3(256), 0, s, B + c * n, (T)1, n);
      HIPCHK(hipGetLastError());
    } else {
      CHK(copy(B + c * n, B));  // q_b at slot s1, left in slot 0 by the segment after
    }
    CHK(propagate_segment<T>(h, s, 2, F, B, c, s0));
    hipLaunchKernelGGL((crys::phi_accumulate_kernel<T>), dim3(crys::red_blocks(n)),
                       dim3(crys::kRedThreads), 0, s, (const T*)F, (const T*)B, n, s0,
                       last ? c + 1 : c, h->NsA, Ns, h->phiA, h->phiB, last ? 1 : 0);
    HIPCHK(hipGetLastError());
    if (stress) CHK(scft_record_stress<T>(h, s, F, B, s0, s1, s0));
  }
  return CRYS_OK;
}

// Records one full SCFT iteration (crysSCFT_sample.cpp:476-508) on stream s.
// The forward chain is q_A then q_B (joined at slot NsA, :487), the backward
// chain q_B^dagger then q_A^dagger (joined at slot NsA, :490); both advance
// in the same launches.  With stress, the per-step stress partials of the
// same propagation are recorded too.
template <typename T>
int scft_record_iteration(crys_scft* h, hipStream_t s, double acc_mix, bool stress = false) {
  const size_t n = h->n, Ns = h->Ns;
  T* qf = (T*)h->qf;
  T* qb = (T*)h->qb;
  if (h->every) {
    CHK(record_checkpointed<T>(h, s, stress));
  } else {
    CHK(propagate_segment<T>(h, s, 2, qf, qb, Ns, 0));
    if (stress) CHK(scft_record_stress<T>(h, s, qf, qb, 0, Ns, 0));
  }
  const int nb = crys::red_blocks(n);
  hipLaunchKernelGGL((crys::qsum_partial_kernel<T>), dim3(nb), dim3(crys::kRedThreads), 0, s,
                     (const T*)qb, h->qpart, n);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(crys::q_finalize_kernel, dim3(1), dim3(crys::kRedThreads), 0, s,
                     (const double*)h->qpart, nb, h->scal, n, h->fA, h->fB, h->NsA, h->NsB);
  HIPCHK(hipGetLastError());
  if (h->every)
    hipLaunchKernelGGL((crys::phi_energy_mix_kernel<T, true>), dim3(nb), dim3(crys::kRedThreads),
                       0, s, (const T*)nullptr, (const T*)nullptr, (const double*)h->scal,
                       h->phiA, h->phiB, h->wA, h->wB, h->ksi, (T*)h->EA, (T*)h->EB, h->epart, n,
                       h->NsA, h->NsB, h->xN, acc_mix, h->ds);
  else
    hipLaunchKernelGGL((crys::phi_energy_mix_kernel<T>), dim3(nb), dim3(crys::kRedThreads), 0, s,
                       (const T*)qf, (const T*)qb, (const double*)h->scal, h->phiA, h->phiB,
                       h->wA, h->wB, h->ksi, (T*)h->EA, (T*)h->EB, h->epart, n, h->NsA, h->NsB,
                       h->xN, acc_mix, h->ds);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(crys::energy_finalize_kernel, dim3(1), dim3(crys::kRedThreads), 0, s,
                     (const double*)h->epart, nb, (const double*)h->scal, h->trace, h->counter,
                     kTraceCap, n);
  HIPCHK(hipGetLastError());
  return CRYS_OK;
}

// Bytes of contour storage for checkpoint interval `every` (0: full histories).
size_t contour_bytes(size_t n, size_t Ns, size_t every, size_t es) {
  if (every == 0) return 2 * (Ns + 1) * n * es;
  const size_t nck = (Ns + every - 1) / every;
  return (nck + 2 * (every + 1)) * n * es;
}

// Allocates the contour storage on first use: full histories when they fit
// the budget (the requested one, else 90% of the free device memory at this
// call), else checkpoints every ceil(sqrt(Ns)) slots; a requested interval
// is used as given.
int scft_storage(crys_scft* h) {
  if (h->qf) return CRYS_OK;
  const size_t n = h->n, Ns = h->Ns, es = elem_size(h->sv->precision);
  size_t budget = h->budget;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    budget = fr / 10 * 9;
  }
  size_t every = h->every_req;
  if (every == 0 && contour_bytes(n, Ns, 0, es) > budget)
    every = (size_t)std::ceil(std::sqrt((double)Ns));
  if (every >= Ns) every = 0;  // one segment: the full histories
  const size_t bytes = contour_bytes(n, Ns, every, es);
  if (h->budget && bytes > h->budget)
    return fail(CRYS_ERR_OUT_OF_MEMORY,
                "contour storage needs %zu bytes (checkpoint interval %zu), budget %zu", bytes,
                every, h->budget);
  const size_t slots = every ? every + 1 : Ns + 1;
  const size_t nck = every ? (Ns + every - 1) / every : 0;
  hipError_t e = hipMalloc(&h->qf, slots * n * es);
  if (e == hipSuccess) e = hipMalloc(&h->qb, slots * n * es);
  if (e == hipSuccess && nck) e = hipMalloc(&h->ck, nck * n * es);
  if (e != hipSuccess) {
    for (void** p : {&h->qf, &h->qb, &h->ck}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    return fail(e == hipErrorOutOfMemory ? CRYS_ERR_OUT_OF_MEMORY : CRYS_ERR_HIP,
                "contour storage (%zu bytes): %s", bytes, hipGetErrorString(e));
  }
  h->every = every;
  h->nck = nck;
  // the start values: q_f[0] = 1 and q_b[Ns] = 1 (crysSCFT_sample.cpp:483, :486);
  // with checkpoints, q_f[0] is checkpoint 0 and q_b[Ns] is set per iteration
  hipStream_t s = h->sv->stream;
  const unsigned g = grid_for(n);
  if (es == sizeof(float)) {
    hipLaunchKernelGGL((crys::fill_kernel<float>), dim3(g), dim3(256), 0, s,
                       (float*)(every ? h->ck : h->qf), 1.0f, n);
    if (!every)
      hipLaunchKernelGGL((crys::fill_kernel<float>), dim3(g), dim3(256), 0, s,
                         (float*)h->qb + Ns * n, 1.0f, n);
  } else {
    hipLaunchKernelGGL((crys::fill_kernel<double>), dim3(g), dim3(256), 0, s,
                       (double*)(every ? h->ck : h->qf), 1.0, n);
    if (!every)
      hipLaunchKernelGGL((crys::fill_kernel<double>), dim3(g), dim3(256), 0, s,
                         (double*)h->qb + Ns * n, 1.0, n);
  }
  HIPCHK(hipGetLastError());
  return CRYS_OK;
}

void scft_free_storage(crys_scft* h) {
  for (void** p : {&h->qf, &h->qb, &h->ck}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  h->every = 0;
  h->nck = 0;
}

int scft_build_graph(crys_scft* h) {
  crys_solver* sv = h->sv;
  drop_graphs(h);
  CHK(scft_storage(h));
  // one iteration, then kGraphIters iterations back to back
  for (int iters : {1, kGraphIters}) {
    HIPCHK(hipStreamBeginCapture(sv->stream, hipStreamCaptureModeThreadLocal));
    int r = CRYS_OK;
    for (int i = 0; i < iters && r == CRYS_OK; ++i)
      r = sv->precision == CRYS_FP32 ? scft_record_iteration<float>(h, sv->stream, h->acc)
                                     : scft_record_iteration<double>(h, sv->stream, h->acc);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(sv->stream, &g);
    if (r != CRYS_OK) {
      if (g) (void)hipGraphDestroy(g);
      drop_graphs(h);
      return r;
    }
    if (e != hipSuccess) {
      drop_graphs(h);
      return fail(CRYS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    hipGraph_t& slot = iters == 1 ? h->graph : h->graph_n;
    hipGraphExec_t& ex = iters == 1 ? h->exec : h->exec_n;
    slot = g;
    const hipError_t ie = hipGraphInstantiate(&ex, slot, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      ex = nullptr;
      drop_graphs(h);
      return fail(CRYS_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
    }
  }
  h->graph_gen = sv->ops_gen;
  return CRYS_OK;
}

// The operators of the handle's own contour step.  The caller may have set
// another ds on the solver, or a new cell (which drops its operators), since
// crys_scft_create: E, NsA and NsB belong to the handle's ds, so the solver is
// set back to it, on its current cell.
int scft_ensure_ops(crys_scft* h) {
  const crys_solver* sv = h->sv;
  if (sv->cur >= 0 && sv->ops[sv->cur].ds == h->ds) return CRYS_OK;
  return crys_set_contour_step(h->sv, h->ds);
}

template <typename T>
void scft_prep(crys_scft* h) {
  hipLaunchKernelGGL((crys::prep_fields_kernel<T>), dim3(grid_for(h->n)), dim3(256), 0,
                     h->sv->stream, (const double*)h->wA, (const double*)h->wB, h->ksi, (T*)h->EA,
                     (T*)h->EB, h->xN, h->ds, h->n);
}

// Device fields in the solver's physical layout: copied into the handle's
// fields, then calc_ksi / array_exp.  Asynchronous on the solver's stream.
int scft_set_fields_device(crys_scft* h, const double* wA, const double* wB) {
  hipStream_t s = h->sv->stream;
  HIPCHK(hipMemcpyAsync(h->wA, wA, h->n * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(h->wB, wB, h->n * 8, hipMemcpyDeviceToDevice, s));
  if (h->sv->precision == CRYS_FP32) scft_prep<float>(h);
  else scft_prep<double>(h);
  HIPCHK(hipGetLastError());
  h->fields_set = true;
  return CRYS_OK;
}

// crys_scft_evaluate and crys_scft_evaluate_device: the fields and the field
// gradients are host arrays, or (device) device arrays in the solver's
// physical layout.
int scft_evaluate_impl(crys_scft* h, bool device, const double* wA, const double* wB, double* H,
                       double* dH_dwA, double* dH_dwB, double dH_dL[3], double* trace_row) {
  if (!h) return fail(CRYS_ERR_INVALID_ARGUMENT, "null handle");
  if ((wA == nullptr) != (wB == nullptr))
    return fail(CRYS_ERR_INVALID_ARGUMENT, "pass both fields or neither");
  if (wA) CHK(device ? scft_set_fields_device(h, wA, wB) : crys_scft_set_fields(h, wA, wB));
  if (!h->fields_set) return fail(CRYS_ERR_INVALID_ARGUMENT, "no fields set");
  crys_solver* sv = h->sv;
  CHK(scft_ensure_ops(h));
  OpSet& op = sv->ops[sv->cur];
  hipStream_t s = sv->stream;
  const size_t n = h->n, Ns = h->Ns;
  const int nb = crys::red_blocks(n);
  const bool stress = dH_dL != nullptr;
  CHK(scft_storage(h));
  if (!h->gA) HIPCHK(hipMalloc((void**)&h->gA, n * 8));
  if (!h->gB) HIPCHK(hipMalloc((void**)&h->gB, n * 8));
  if (!h->sums) HIPCHK(hipMalloc((void**)&h->sums, (1 + 3 * Ns) * 8));
  if (stress) {
    if (!h->spart) HIPCHK(hipMalloc((void**)&h->spart, 3 * Ns * nb * 8));
    if (sv->glide) {
      if (op.gdL.empty()) CHK(upload_glide_terms(sv, op.ds, true, &op));
    } else if (!op.dmat_dL[0]) {
      CHK(upload_axis_ops(sv, op.ds, true, op.dmat_dL));
    }
  }
  hipLaunchKernelGGL(crys::set_int_kernel, dim3(1), dim3(1), 0, s, h->counter, 0);
  HIPCHK(hipGetLastError());
  const bool f32 = sv->precision == CRYS_FP32;
  CHK(f32 ? scft_record_iteration<float>(h, s, 0.0, stress)
          : scft_record_iteration<double>(h, s, 0.0, stress));
  const double cA = h->ds * (double)h->NsA / h->fA, cB = h->ds * (double)h->NsB / h->fB;
  hipLaunchKernelGGL(crys::hgrad_kernel, dim3(nb), dim3(crys::kRedThreads), 0, s,
                     (const double*)h->wA, (const double*)h->wB, (const double*)h->phiA,
                     (const double*)h->phiB, h->gA, h->gB, h->qpart, n, h->xN, cA, cB);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(crys::sum_partials_kernel, dim3(1), dim3(crys::kRedThreads), 0, s,
                     (const double*)h->qpart, nb, h->sums);
  HIPCHK(hipGetLastError());
  if (stress) {
    hipLaunchKernelGGL(crys::sum_partials_kernel, dim3((unsigned)(3 * Ns)), dim3(crys::kRedThreads),
                       0, s, (const double*)h->spart, nb, h->sums + 1);
    HIPCHK(hipGetLastError());
  }
  std::vector<double> sums(1 + 3 * Ns);
  double row[CRYS_TRACE_COLS];
  HIPCHK(hipMemcpyAsync(row, h->trace, sizeof(row), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sums.data(), h->sums, (stress ? 1 + 3 * Ns : 1) * 8,
                        hipMemcpyDeviceToHost, s));
  if (device) {
    if (dH_dwA) HIPCHK(hipMemcpyAsync(dH_dwA, h->gA, n * 8, hipMemcpyDeviceToDevice, s));
    if (dH_dwB) HIPCHK(hipMemcpyAsync(dH_dwB, h->gB, n * 8, hipMemcpyDeviceToDevice, s));
  } else {
    if (dH_dwA) HIPCHK(hipMemcpyAsync(dH_dwA, h->gA, n * 8, hipMemcpyDeviceToHost, s));
    if (dH_dwB) HIPCHK(hipMemcpyAsync(dH_dwB, h->gB, n * 8, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  const double Q = row[CRYS_TRACE_Q];
  if (trace_row) memcpy(trace_row, row, sizeof(row));
  if (H) *H = sums[0] / (double)n - log(Q);
  if (stress) {
    for (int a = 0; a < 3; ++a) {
      double acc = 0.0;
      for (size_t t = 0; t < Ns; ++t) acc += sums[1 + 3 * t + a];
      dH_dL[a] = -acc / ((double)n * Q);
    }
  }
  if (!std::isfinite(Q) || !(Q > 0)) return fail(CRYS_ERR_DIVERGED, "non-finite or zero Q");
  return CRYS_OK;
}

}  // namespace

extern "C" {

int crys_scft_create(crys_solver_t* solver, double fA, double ds, double xN, double Lx, double Ly,
                     double Lz, double acceptance, crys_scft_t** out) {
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  CHK(check_handle(solver));
  if (!(ds > 0) || !std::isfinite(ds)) return fail(CRYS_ERR_INVALID_ARGUMENT, "ds must be > 0");
  if (!(fA > 0 && fA < 1)) return fail(CRYS_ERR_INVALID_ARGUMENT, "fA must lie in (0, 1)");
  const double fB = 1 - fA;
  const size_t NsA = (size_t)(fA / ds);  // crysSCFT_sample.cpp:443
  const size_t NsB = (size_t)(fB / ds);  // :444
  if (NsA == 0 || NsB == 0)
    return fail(CRYS_ERR_INVALID_ARGUMENT, "NsA = %zu, NsB = %zu: each block needs >= 1 step", NsA,
                NsB);
  const double cell[6] = {Lx, Ly, Lz, 0, 0, 0};
  CHK(crys_set_cell_para(solver, cell));   // :466
  CHK(crys_set_contour_step(solver, ds));  // :467
  CHK(crys_diffusion(solver));             // :471 warm-up
  crys_scft* h = new (std::nothrow) crys_scft();
  if (!h) return fail(CRYS_ERR_OUT_OF_MEMORY, "host allocation failed");
  h->sv = solver;
  h->fA = fA;
  h->fB = fB;
  h->ds = ds;
  h->xN = xN;
  h->acc = acceptance;
  h->L[0] = Lx;
  h->L[1] = Ly;
  h->L[2] = Lz;
  h->NsA = NsA;
  h->NsB = NsB;
  h->Ns = NsA + NsB;
  h->n = h->sv->n_phy;
  const size_t n = h->n, es = elem_size(solver->precision);
  hipError_t e = hipSuccess;
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  A((void**)&h->wA, n * 8);
  A((void**)&h->wB, n * 8);
  A((void**)&h->ksi, n * 8);
  A((void**)&h->phiA, n * 8);
  A((void**)&h->phiB, n * 8);
  A(&h->EA, n * es);
  A(&h->EB, n * es);
  // two fields per chain: the phases alternate between them (a phase may
  // run on two workgroups per plane, so it cannot rewrite its input in place);
  // for a 3m solver with translational parts two per term (its phases keep
  // one plane per term) and at least three (launch_glide_step: the term in
  // flight, the partial sum and the next term)
  const size_t nterms = solver->glide ? solver->ops[solver->cur].gmat.size() : 1;
  const size_t nscr = solver->glide ? std::max<size_t>(3, 2 * nterms) : 2;
  A(&h->scr[0], nscr * n * es);
  A(&h->scr[1], nscr * n * es);
  A((void**)&h->qpart, crys::kMaxRedBlocks * 8);
  A((void**)&h->epart, 4 * crys::kMaxRedBlocks * 8);
  A((void**)&h->scal, 4 * 8);
  A((void**)&h->trace, (size_t)kTraceCap * CRYS_TRACE_COLS * 8);
  A((void**)&h->counter, sizeof(int));
  if (e != hipSuccess) {
    scft_free(h);
    delete h;
    return fail(e == hipErrorOutOfMemory ? CRYS_ERR_OUT_OF_MEMORY : CRYS_ERR_HIP,
                "SCFT allocation failed: %s", hipGetErrorString(e));
  }
  e = hipStreamSynchronize(h->sv->stream);
  if (e != hipSuccess) {
    scft_free(h);
    delete h;
    return fail(CRYS_ERR_HIP, "SCFT setup failed: %s", hipGetErrorString(e));
  }
  *out = h;
  return CRYS_OK;
}

int crys_scft_destroy(crys_scft_t* h) {
  if (!h) return CRYS_OK;
  if (h->sv && h->sv->stream) (void)hipStreamSynchronize(h->sv->stream);
  scft_free(h);
  delete h;
  return CRYS_OK;
}

int crys_scft_set_checkpointing(crys_scft_t* h, size_t every, size_t budget_bytes) {
  if (!h) return fail(CRYS_ERR_INVALID_ARGUMENT, "null handle");
  HIPCHK(hipStreamSynchronize(h->sv->stream));
  // the iteration graph records the storage addresses: rebuild it on next use
  drop_graphs(h);
  scft_free_storage(h);
  const size_t old_every = h->every_req, old_budget = h->budget;
  h->every_req = every;
  h->budget = budget_bytes;
  const int r = scft_storage(h);
  if (r != CRYS_OK) {
    // keep the handle usable with the previous setting
    std::string keep = g_err;
    h->every_req = old_every;
    h->budget = old_budget;
    (void)scft_storage(h);
    g_err = keep;
    return r;
  }
  HIPCHK(hipStreamSynchronize(h->sv->stream));
  return CRYS_OK;
}

int crys_scft_get_checkpointing(crys_scft_t* h, size_t* every, size_t* storage_bytes) {
  if (!h) return fail(CRYS_ERR_INVALID_ARGUMENT, "null handle");
  CHK(scft_storage(h));
  if (every) *every = h->every;
  if (storage_bytes)
    *storage_bytes = contour_bytes(h->n, h->Ns, h->every, elem_size(h->sv->precision));
  return CRYS_OK;
}

int crys_scft_set_fields(crys_scft_t* h, const double* wA, const double* wB) {
  if (!h || !wA || !wB) return fail(CRYS_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t s = h->sv->stream;
  HIPCHK(hipMemcpyAsync(h->wA, wA, h->n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(h->wB, wB, h->n * 8, hipMemcpyHostToDevice, s));
  if (h->sv->precision == CRYS_FP32) scft_prep<float>(h);
  else scft_prep<double>(h);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  h->fields_set = true;
  return CRYS_OK;
}

int crys_scft_run(crys_scft_t* h, size_t nstep, double* trace_host) {
  if (!h) return fail(CRYS_ERR_INVALID_ARGUMENT, "null handle");
  if (!h->fields_set) return fail(CRYS_ERR_INVALID_ARGUMENT, "crys_scft_set_fields not called");
  crys_solver* sv = h->sv;
  CHK(scft_ensure_ops(h));
  if (!h->exec || h->graph_gen != sv->ops_gen) CHK(scft_build_graph(h));
  hipStream_t s = sv->stream;
  bool diverged = false;
  size_t done = 0;
  // without a trace buffer the rows still come back, for the divergence check
  std::vector<double> rows;
  if (!trace_host) rows.resize((size_t)kTraceCap * CRYS_TRACE_COLS);
  while (done < nstep) {
    const size_t chunk = nstep - done < (size_t)kTraceCap ? nstep - done : (size_t)kTraceCap;
    hipLaunchKernelGGL(crys::set_int_kernel, dim3(1), dim3(1), 0, s, h->counter, 0);
    HIPCHK(hipGetLastError());
    size_t i = 0;
    if (h->exec_n)
      for (; i + kGraphIters <= chunk; i += kGraphIters) HIPCHK(hipGraphLaunch(h->exec_n, s));
    for (; i < chunk; ++i) HIPCHK(hipGraphLaunch(h->exec, s));
    double* dst = trace_host ? trace_host + done * CRYS_TRACE_COLS : rows.data();
    HIPCHK(hipMemcpyAsync(dst, h->trace, chunk * CRYS_TRACE_COLS * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i < chunk; ++i) {
      const double* r = dst + i * CRYS_TRACE_COLS;
      if (!std::isfinite(r[CRYS_TRACE_F]) || !std::isfinite(r[CRYS_TRACE_Q])) diverged = true;
    }
    done += chunk;
  }
  if (diverged) return fail(CRYS_ERR_DIVERGED, "non-finite free energy or Q in the trace");
  return CRYS_OK;
}

int crys_scft_get_fields(crys_scft_t* h, double* phiA, double* phiB, double* wA, double* wB) {
  if (!h) return fail(CRYS_ERR_INVALID_ARGUMENT, "null handle");
  hipStream_t s = h->sv->stream;
  const size_t b = h->n * 8;
  if (phiA) HIPCHK(hipMemcpyAsync(phiA, h->phiA, b, hipMemcpyDeviceToHost, s));
  if (phiB) HIPCHK(hipMemcpyAsync(phiB, h->phiB, b, hipMemcpyDeviceToHost, s));
  if (wA) HIPCHK(hipMemcpyAsync(wA, h->wA, b, hipMemcpyDeviceToHost, s));
  if (wB) HIPCHK(hipMemcpyAsync(wB, h->wB, b, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return CRYS_OK;
}

int crys_scft_get_contour(const crys_scft_t* h, size_t* NsA, size_t* NsB) {
  if (!h) return fail(CRYS_ERR_INVALID_ARGUMENT, "null handle");
  if (NsA) *NsA = h->NsA;
  if (NsB) *NsB = h->NsB;
  return CRYS_OK;
}

int crys_diblock_scft_simple_iterate(crys_solver_t* solver, double fA, double ds, double xN,
                                     double Lx, double Ly, double Lz, double acceptance,
                                     size_t nstep, const double* wA_init, const double* wB_init,
                                     double* trace, double* phiA, double* phiB, double* wA_out,
                                     double* wB_out) {
  crys_scft* h = nullptr;
  CHK(crys_scft_create(solver, fA, ds, xN, Lx, Ly, Lz, acceptance, &h));
  int r = crys_scft_set_fields(h, wA_init, wB_init);
  if (r == CRYS_OK) r = crys_scft_run(h, nstep, trace);
  if (r == CRYS_OK || r == CRYS_ERR_DIVERGED) {
    int r2 = crys_scft_get_fields(h, phiA, phiB, wA_out, wB_out);
    if (r == CRYS_OK) r = r2;
  }
  std::string keep = g_err;
  crys_scft_destroy(h);
  g_err = keep;
  return r;
}

int crys_scft_evaluate(crys_scft_t* h, const double* wA, const double* wB, double* H,
                       double* dH_dwA, double* dH_dwB, double dH_dL[3], double* trace_row) {
  return scft_evaluate_impl(h, false, wA, wB, H, dH_dwA, dH_dwB, dH_dL, trace_row);
}

int crys_scft_evaluate_device(crys_scft_t* h, const double* wA, const double* wB, double* H,
                              double* dH_dwA, double* dH_dwB, double dH_dL[3],
                              double* trace_row) {
  return scft_evaluate_impl(h, true, wA, wB, H, dH_dwA, dH_dwB, dH_dL, trace_row);
}

int crys_axis_operator(int kind, size_t n_logical, double L, double ds, double* out) {
  if (!out) return fail(CRYS_ERR_INVALID_ARGUMENT, "null output");
  if (!(ds > 0) || !(L > 0)) return fail(CRYS_ERR_INVALID_ARGUMENT, "L and ds must be > 0");
  if (n_logical == 0 || n_logical > 256)
    return fail(CRYS_ERR_INVALID_ARGUMENT, "n_logical out of range");
  std::vector<double> At;
  if (kind == CRYS_P1) At = crys::p1_axis_operator((int)n_logical, L, ds);
  else if (kind == CRYS_3M || kind == CRYS_PMMM_DCT) {
    if (n_logical % 2) return fail(CRYS_ERR_BAD_GRID, "n_logical must be even");
    At = kind == CRYS_3M ? crys::m3_axis_operator((int)n_logical, L, ds)
                         : crys::pmmm_axis_operator((int)n_logical / 2, L, ds);
  } else
    return fail(CRYS_ERR_INVALID_ARGUMENT, "unknown solver kind %d", kind);
  memcpy(out, At.data(), At.size() * sizeof(double));
  return CRYS_OK;
}

int crys_scft_time_propagation(crys_scft_t* h, int reps, double* ms_per_step,
                                int* launches_per_step) {
  if (!h || reps <= 0) return fail(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  crys_solver* sv = h->sv;
  CHK(scft_ensure_ops(h));
  CHK(scft_storage(h));
  hipStream_t s = sv->stream;
  // the propagation of both chains exactly as the iteration records it (one
  // segment of `every` steps when checkpointing), captured in a graph
  const size_t c = h->every ? h->every : h->Ns;
  const bool f32 = sv->precision == CRYS_FP32;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int r = f32 ? propagate_segment<float>(h, s, 2, (float*)h->qf, (float*)h->qb, c, 0)
                    : propagate_segment<double>(h, s, 2, (double*)h->qf, (double*)h->qb, c, 0);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (r != CRYS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return fail(CRYS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  size_t nodes = 0;
  hipGraphExec_t x = nullptr;
  e = hipGraphGetNodes(g, nullptr, &nodes);
  if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipGraphLaunch(x, s);  // warm
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int i = 0; i < reps && e == hipSuccess; ++i) e = hipGraphLaunch(x, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (x) (void)hipGraphExecDestroy(x);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(CRYS_ERR_HIP, "propagation timing: %s", hipGetErrorString(e));
  if (ms_per_step) *ms_per_step = (double)ms / ((double)reps * (double)c);
  if (launches_per_step) *launches_per_step = (int)((nodes + c / 2) / c);
  return CRYS_OK;
}

int crys_scft_time_step(crys_scft_t* h, int reps, double* ms_per_step, double* ms_per_launch,
                        int* nkernels) {
  if (!h || reps <= 0) return fail(CRYS_ERR_INVALID_ARGUMENT, "bad arguments");
  crys_solver* sv = h->sv;
  CHK(scft_ensure_ops(h));
  const OpSet& op = sv->ops[sv->cur];
  hipStream_t s = sv->stream;
  CHK(scft_storage(h));
  // the slots held in qf / qb: the whole contour, or one checkpoint segment
  const size_t n = h->n, Ns = h->every ? h->every : h->Ns;
  int nl = 0;
  const bool phase = phase_mp(sv) != 0 && Ns >= 2;
  // one contour step of both chains (q_0 -> q_1 forward, q+_Ns -> q+_Ns-1
  // backward), or only its launch `only`; on the phase path the step is one
  // full phase (phase 1: finishes step 0 and starts step 1)
  auto step = [&](int only) -> int {
    if (phase) {
      nl = 1;
      if (sv->precision == CRYS_FP32) {
        float* sc[2] = {(float*)h->scr[0], (float*)h->scr[1]};
        return launch_phase_step<float>(sv, op, 1, Ns, h->NsA, (float*)h->qf, (float*)h->qb, sc,
                                        (const float*)h->EA, (const float*)h->EB, s);
      }
      double* sc[2] = {(double*)h->scr[0], (double*)h->scr[1]};
      return launch_phase_step<double>(sv, op, 1, Ns, h->NsA, (double*)h->qf, (double*)h->qb, sc,
                                       (const double*)h->EA, (const double*)h->EB, s);
    }
    if (sv->glide) {
      // the whole step of the separable terms, timed as one entry
      nl = 1;
      if (sv->precision == CRYS_FP32)
        return propagate_segment<float>(h, s, 2, (float*)h->qf, (float*)h->qb, 1, 0);
      return propagate_segment<double>(h, s, 2, (double*)h->qf, (double*)h->qb, 1, 0);
    }
    if (sv->precision == CRYS_FP32) {
      float* qf = (float*)h->qf;
      float* qb = (float*)h->qb;
      const float* in[2] = {qf, qb + Ns * n};
      float* out[2] = {qf + n, qb + (Ns - 1) * n};
      float* sc[2] = {(float*)h->scr[0], (float*)h->scr[1]};
      const float* E[2] = {(const float*)h->EA, (const float*)h->EB};
      return launch_step<float>(sv, op, 2, in, out, sc, E, s, true, only, &nl);
    }
    double* qf = (double*)h->qf;
    double* qb = (double*)h->qb;
    const double* in[2] = {qf, qb + Ns * n};
    double* out[2] = {qf + n, qb + (Ns - 1) * n};
    double* sc[2] = {(double*)h->scr[0], (double*)h->scr[1]};
    const double* E[2] = {(const double*)h->EA, (const double*)h->EB};
    return launch_step<double>(sv, op, 2, in, out, sc, E, s, true, only, &nl);
  };
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto timed = [&](int only, double* ms_out) -> int {
    int r = step(only);  // warm
    if (r != CRYS_OK) return r;
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < reps && r == CRYS_OK; ++i) r = step(only);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_out = ms / reps;
    return r;
  };
  double whole = 0.0;
  int r = timed(-1, &whole);
  if (r == CRYS_OK && ms_per_step) *ms_per_step = whole;
  for (int k = 0; r == CRYS_OK && k < nl && k < 3; ++k) {
    double t = 0.0;
    r = timed(k, &t);
    if (r == CRYS_OK && ms_per_launch) ms_per_launch[k] = t;
  }
  if (r == CRYS_OK && nkernels) *nkernels = nl;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return r;
}

}  // extern "C"
