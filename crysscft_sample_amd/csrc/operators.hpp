// Per-axis diffusion operators of the three solvers, built on the host.
//
// In every solver the Fourier multiplier K = exp(-ds (kx^2 + ky^2 + kz^2))
// factorises over the axes, and so do the transforms (FFTW r2c/c2r for P1,
// REDFT10/REDFT01 for Pmmm).  The whole triplet "forward transform, xK,
// inverse transform, normalise" is therefore the tensor product
// Dx (x) Dy (x) Dz of three small real matrices, one per axis.  This file
// builds those matrices in extended precision; the device applies them with
// one pass per axis.
//
//  - PmmmDCT (solver_Pmmm.h:125-215, K from :159-186):
//      D = C3 diag(K) C2 / (2M), C2[k][j] = 2 cos(pi k (j+1/2)/M) (REDFT10),
//      C3[j][k] = c_k cos(pi k (j+1/2)/M), c_0 = 1, c_k = 2 (REDFT01).
//  - P1 (solver_p1.h:83-202, K from :142-172):
//      D[j][j'] = (1/N) sum_k K(k) cos(2 pi k (j-j')/N), k folded as
//      k < (N+1)/2 ? k : N-k.
//  - 3m with zero translational parts (solver_3m.h:143-324): the even
//      sub-lattice e[j] = f[2j] of a field with f(N-1-j) = f(j) fixes f on
//      the odd sites (f[2j+1] = e[M-1-j]), so D3[j][j'] = Dp1[2j][2j'] +
//      Dp1[2j][2M-1-2j'] with the P1 operator of the logical length N = 2M.
#pragma once

#include <vector>

namespace crys {

// Each builder returns At with At[i * M + j] = D[j][i], i.e. out_j =
// sum_i At[i*M + j] * in_i (the layout the device pass streams).  With
// dL = true it returns dD/dL instead: K_k is replaced by
// dK_k/dL = K_k * 2 ds (2 pi k)^2 / L^3 (the cell stress of the operator).
std::vector<double> pmmm_axis_operator(int M, double L, double ds, bool dL = false);
std::vector<double> p1_axis_operator(int N, double L, double ds, bool dL = false);
std::vector<double> m3_axis_operator(int N_logical, double L, double ds, bool dL = false);

// Half-size form of a centrosymmetric operator (D[M-1-j][M-1-k] = D[j][k],
// true for PmmmDCT and P1): from At (M x M, At[k*M + j] = D[j][k]) returns
// At_c (M/2 x M) with At_c[k*M + j] = Ps[j][k] for j < M/2 and
// At_c[k*M + M/2 + j] = Pa[j][k], Ps/Pa = (D[j][k] +/- D[j][M-1-k]) / 2.
//
// With `shift` the pairing is the half-period translation k <-> k + M/2
// instead (D[j+M/2][k+M/2] = D[j][k], indices mod M): the 3m operator with
// zero translational parts (M even).  On the even sub-lattice the mirror
// f(N-1-j) = f(j) composed with the translation by N/2 maps e[m] to
// e[m + M/2], and D3 commutes with it.  In the basis that keeps k < M/2 and
// reverses the upper half (position M-1-i holds index M/2 + i) D3 is
// centrosymmetric; Ps/Pa = (D[j][k] +/- D[j][M/2+k]) / 2 are its blocks, and
// the kernels read the partner of k at M/2 + k.
std::vector<double> centro_half(const std::vector<double>& At, int M, bool shift = false);

// Constant table of the radix pass (radix_pass.hpp) for one axis.  kind 0:
// P1, transform length M = N (the logical length), K' = K(k folded) / N;
// kind 1: PmmmDCT, M = the reduced length, K' = K(k) / (2M); kind 2: 3m with
// zero translational parts, M = N / 2, K' = K(k) / (2M), k < M.  Layout:
// cos, sin of 2 pi j / M (j < M, interleaved), cos, sin of pi k / 2M
// (k < M, interleaved), then K'_k (k < M): 5 M values.
std::vector<double> radix_table(int kind, int M, double L, double ds);

// 3m with translational parts (CrysPSSolver3m's translational_part,
// solver_3m.h:17, folded into its coefficients at solver_3m.h:172-255 with the
// operations of generate_m3_operation, solver_utilities.h:55-66).  Operation o
// maps reduced coordinate j of axis a to the full-grid image
// 2 j + s_oa (not mirrored) or -1 - 2 j + s_oa (mirrored, bit a of mask[o]),
// mod N_a.  The diffusion on the even sub-lattice is then
//   out = sum_o B_ox (x) B_oy (x) B_oz,  B_oa[i][j] = Dp1_a[2 i][image_oa(j)],
// with the P1 operator of each logical axis.  Terms whose factors agree on
// two axes merge by adding the third; the result is a short list of
// separable terms (one for zero or mirror-plane-only translations, two for a
// single glide).  Each factor is returned in the At layout of the passes;
// with dL the factors are the L derivatives of the same terms (same list).
struct M3Term {
  std::vector<double> At[3];
};
std::vector<M3Term> m3_glide_terms(const int mask[8], const int shift[8][3], const int N[3],
                                   const double L[3], double ds, bool dL = false);

// bf16 bits of x, round to nearest even (finite x).
unsigned short bf16_bits(float x);

}  // namespace crys
