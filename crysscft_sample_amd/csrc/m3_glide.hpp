// 3m with non-zero translational parts (CrysPSSolver3m's translational_part,
// solver_3m.h:17, :176; the symmetry operations of generate_m3_operation,
// solver_utilities.h:55-66).
//
// The reference folds the full-grid multiplier into 16 coefficient arrays
// with a phase per symmetry operation.  Here the same linear map is applied
// as: expand the reduced field (the even sub-lattice f[2x,2y,2z]) to the full
// logical grid with the 8 operations of the group, apply the P1 operators
// there, and take the even sub-lattice back.  This is exact: the diffusion
// operator commutes with the operations, and the 8 operations map the even
// sub-lattice one to one onto the grid (checked at solver creation).
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

// image of reduced coordinate j on an axis of logical length n
__host__ __device__ inline long m3_image(long j, int mirrored, int shift, long n) {
  const long v = mirrored ? -1 - 2 * j + shift : 2 * j + shift;
  return ((v % n) + n) % n;
}

// full[op(2 r)] = red[r] for the 8 operations; X, Y, Z the reduced lengths.
static __global__ void m3_expand_kernel(const double* __restrict__ red, double* __restrict__ full,
                                        M3Ops m, int X, int Y, int Z) {
  const size_t n = (size_t)X * Y * Z;
  const long NX = 2 * X, NY = 2 * Y, NZ = 2 * Z;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const long x = (long)(i / ((size_t)Y * Z)), y = (long)((i / Z) % Y), z = (long)(i % Z);
    const double v = red[i];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const long fx = m3_image(x, m.mask[o] & 1, m.shift[o][0], NX);
      const long fy = m3_image(y, (m.mask[o] >> 1) & 1, m.shift[o][1], NY);
      const long fz = m3_image(z, (m.mask[o] >> 2) & 1, m.shift[o][2], NZ);
      full[((size_t)fx * NY + fy) * NZ + fz] = v;
    }
  }
}

// red[r] = scale * full[2 r]
static __global__ void m3_extract_kernel(const double* __restrict__ full, double* __restrict__ red,
                                         int X, int Y, int Z, double scale) {
  const size_t n = (size_t)X * Y * Z;
  const size_t NY = 2 * (size_t)Y, NZ = 2 * (size_t)Z;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t x = i / ((size_t)Y * Z), y = (i / Z) % Y, z = i % Z;
    red[i] = scale * full[((2 * x) * NY + 2 * y) * NZ + 2 * z];
  }
}

// out[i] = a[i] * b[i]
static __global__ void mul_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                  double* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = a[i] * b[i];
}

// Host-side expansion / extraction of full-grid fields (the SCFT of a 3m
// solver with translational parts runs on the expanded fields).
inline void m3_expand_host(const double* red, double* full, const M3Ops& m, const size_t nred[3]) {
  const long X = (long)nred[0], Y = (long)nred[1], Z = (long)nred[2];
  for (long x = 0; x < X; ++x)
    for (long y = 0; y < Y; ++y)
      for (long z = 0; z < Z; ++z) {
        const double v = red[(x * Y + y) * Z + z];
        for (int o = 0; o < 8; ++o) {
          const long fx = m3_image(x, m.mask[o] & 1, m.shift[o][0], 2 * X);
          const long fy = m3_image(y, (m.mask[o] >> 1) & 1, m.shift[o][1], 2 * Y);
          const long fz = m3_image(z, (m.mask[o] >> 2) & 1, m.shift[o][2], 2 * Z);
          full[(fx * 2 * Y + fy) * 2 * Z + fz] = v;
        }
      }
}

inline void m3_extract_host(const double* full, double* red, const size_t nred[3], double scale) {
  const size_t X = nred[0], Y = nred[1], Z = nred[2];
  for (size_t x = 0; x < X; ++x)
    for (size_t y = 0; y < Y; ++y)
      for (size_t z = 0; z < Z; ++z)
        red[(x * Y + y) * Z + z] = scale * full[((2 * x) * 2 * Y + 2 * y) * 2 * Z + 2 * z];
}

}  // namespace crys
