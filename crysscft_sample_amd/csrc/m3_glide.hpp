// 3m with non-zero translational parts (CrysPSSolver3m's translational_part,
// solver_3m.h:17, :176; the symmetry operations of generate_m3_operation,
// solver_utilities.h:55-66).
//
// The reference folds the full-grid multiplier into 16 coefficient arrays
// with a phase per symmetry operation and works on the half grid in Fourier
// space.  Here the same linear map is applied on the reduced grid in real
// space: operation o maps the even sub-lattice point j of an axis to its
// image on the logical grid, and the P1 operator restricted to
// (sub-lattice row, image column) gives one factor per axis and operation;
// the sum over the 8 operations of the tensor products of their factors, with
// terms that share two factors merged (crys::m3_glide_terms in
// operators.cpp), is the diffusion of the reduced field.  The 8 operations
// must map the even sub-lattice one to one onto the grid (checked at solver
// creation).
//
// Operation o mirrors the axes in mask[o] (bit 0 = x, 1 = y, 2 = z) and
// translates by t_o (fractions of the cell).  The translations are those of
// the reference's phase rows (-z: g0..g2, -y: g3..g5, -x: g6..g8 and their
// products); in grid points an image coordinate is
//   j' = 2 j + s        (axis not mirrored)
//   j' = -1 - 2 j + s   (axis mirrored, the mirror plane at -1/2)
// modulo N, with s = t N.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

namespace crys {

struct M3Ops {
  int mask[8];
  int shift[8][3];
};

// Builds the operation table for translational part g on logical grid N.
// Returns 0, or 1 if a translation is not a whole number of grid points, or 2
// if the operations do not map the even sub-lattice onto the whole grid.
inline int m3_build_ops(const double g[9], const size_t N[3], M3Ops* m) {
  const double t[8][3] = {
      {0, 0, 0},
      {g[0], g[1], g[2]},
      {g[3], g[4], g[5]},
      {g[0] + g[3], -g[1] + g[4], g[2] + g[5]},
      {g[6], g[7], g[8]},
      {g[6] - g[0], g[1] + g[7], g[2] + g[8]},
      {-g[3] + g[6], g[4] + g[7], g[5] + g[8]},
      {-g[0] - g[3] + g[6], -g[1] + g[4] + g[7], g[2] + g[5] + g[8]},
  };
  const int mask[8] = {0, 4, 2, 6, 1, 5, 3, 7};
  int seen = 0;
  for (int o = 0; o < 8; ++o) {
    m->mask[o] = mask[o];
    int parity = 0;
    for (int a = 0; a < 3; ++a) {
      const double s = t[o][a] * (double)N[a];
      const double r = std::nearbyint(s);
      if (std::fabs(s - r) > 1e-9) return 1;
      const long n = (long)N[a];
      long si = ((long)r % n + n) % n;
      m->shift[o][a] = (int)si;
      const long off = (mask[o] >> a) & 1 ? si - 1 : si;
      parity |= (int)(((off % 2) + 2) % 2) << a;
    }
    if (seen & (1 << parity)) return 2;
    seen |= 1 << parity;
  }
  return 0;
}

inline bool m3_nonzero(const double g[9]) {
  for (int i = 0; i < 9; ++i)
    if (g[i] != 0.0) return true;
  return false;
}

// Sum of two terms of a contour step, per chain (blockIdx.y): out = (a + b)
// times e where e is not null.
template <typename T>
struct GlideSum {
  struct Chain {
    T* out;
    const T* a;
    const T* b;
    const T* e;
  } c[2];
};

template <typename T>
__global__ void glide_sum_kernel(GlideSum<T> g, size_t n) {
  const typename GlideSum<T>::Chain ch = g.c[blockIdx.y];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const T v = ch.a[i] + ch.b[i];
    ch.out[i] = ch.e ? v * ch.e[i] : v;
  }
}

}  // namespace crys
