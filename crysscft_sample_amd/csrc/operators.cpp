// Host construction of the per-axis diffusion operators (see operators.hpp).
// Sums run in long double with exact integer range reduction of the cosine
// arguments, so each entry is the double nearest the exact operator to within
// a few ulp.
#include "operators.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <utility>

namespace crys {

namespace {

constexpr long double kPiL = 3.141592653589793238462643383279502884L;

// cos(2 pi m / d) for m in [0, d)
std::vector<long double> cos_table(long d) {
  std::vector<long double> t(d);
  for (long m = 0; m < d; ++m) t[m] = cosl(2.0L * kPiL * (long double)m / (long double)d);
  return t;
}

// exp(-ds (2 pi k / L)^2), the per-axis factor of K, or its L derivative.
long double kfactor(long k, double L, double ds, bool dL) {
  long double kk = 2.0L * kPiL * (long double)k / (long double)L;
  long double K = expl(-(long double)ds * kk * kk);
  return dL ? K * 2.0L * (long double)ds * kk * kk / (long double)L : K;
}

std::vector<long double> p1_dense(int N, double L, double ds, bool dL) {
  std::vector<long double> K(N);
  for (int k = 0; k < N; ++k) K[k] = kfactor(k < (N + 1) / 2 ? k : N - k, L, ds, dL);
  std::vector<long double> ct = cos_table(N);
  // circulant: D[j][j'] = c(j - j') with c(d) = (1/N) sum_k K(k) cos(2 pi k d / N)
  std::vector<long double> c(N);
  for (int d = 0; d < N; ++d) {
    long double s = 0;
    for (int k = 0; k < N; ++k) s += K[k] * ct[((long)k * d) % N];
    c[d] = s / (long double)N;
  }
  std::vector<long double> D((size_t)N * N);
  for (int j = 0; j < N; ++j)
    for (int jp = 0; jp < N; ++jp) D[(size_t)j * N + jp] = c[((j - jp) % N + N) % N];
  return D;
}

}  // namespace

std::vector<double> pmmm_axis_operator(int M, double L, double ds, bool dL) {
  // cos(pi k (2j+1) / (2M)) = cos(2 pi m / (4M)), m = k (2j+1) mod 4M
  std::vector<long double> ct = cos_table(4L * M);
  std::vector<long double> K(M);
  for (int k = 0; k < M; ++k) K[k] = kfactor(k, L, ds, dL);
  std::vector<double> At((size_t)M * M);
  for (int j = 0; j < M; ++j)
    for (int jp = 0; jp < M; ++jp) {
      long double s = 0;
      for (int k = 0; k < M; ++k) {
        long double ck = k == 0 ? 1.0L : 2.0L;
        long double a = ct[((long)k * (2 * j + 1)) % (4L * M)];
        long double b = ct[((long)k * (2 * jp + 1)) % (4L * M)];
        s += ck * a * K[k] * 2.0L * b;
      }
      // D[j][jp]; D is symmetric, stored transposed all the same
      At[(size_t)jp * M + j] = (double)(s / (2.0L * M));
    }
  return At;
}

std::vector<double> p1_axis_operator(int N, double L, double ds, bool dL) {
  std::vector<long double> D = p1_dense(N, L, ds, dL);
  std::vector<double> At((size_t)N * N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) At[(size_t)i * N + j] = (double)D[(size_t)j * N + i];
  return At;
}

std::vector<double> m3_axis_operator(int N, double L, double ds, bool dL) {
  const int M = N / 2;
  std::vector<long double> D = p1_dense(N, L, ds, dL);
  std::vector<double> At((size_t)M * M);
  for (int j = 0; j < M; ++j)
    for (int i = 0; i < M; ++i) {
      long double v = D[(size_t)(2 * j) * N + 2 * i] + D[(size_t)(2 * j) * N + (2 * M - 1 - 2 * i)];
      At[(size_t)i * M + j] = (double)v;
    }
  return At;
}

std::vector<M3Term> m3_glide_terms(const int mask[8], const int shift[8][3], const int N[3],
                                   const double L[3], double ds, bool dL) {
  // a factor is a list of base matrices (mirrored, shift), summed
  using Base = std::pair<int, int>;
  using Factor = std::vector<Base>;
  std::vector<std::array<Factor, 3>> terms;
  for (int o = 0; o < 8; ++o) {
    std::array<Factor, 3> t;
    for (int a = 0; a < 3; ++a) t[a] = {Base{(mask[o] >> a) & 1, shift[o][a]}};
    terms.push_back(t);
  }
  // merge two terms that agree on two axes, adding the third; repeat
  for (bool merged = true; merged;) {
    merged = false;
    for (size_t i = 0; i < terms.size() && !merged; ++i)
      for (size_t j = i + 1; j < terms.size() && !merged; ++j) {
        int same = 0, diff = -1;
        for (int a = 0; a < 3; ++a) {
          if (terms[i][a] == terms[j][a]) ++same;
          else diff = a;
        }
        if (same < 2) continue;
        if (diff < 0) diff = 0;  // identical terms: twice one factor
        Factor f = terms[i][diff];
        f.insert(f.end(), terms[j][diff].begin(), terms[j][diff].end());
        std::sort(f.begin(), f.end());
        terms[i][diff] = f;
        terms.erase(terms.begin() + (long)j);
        merged = true;
      }
  }
  std::vector<M3Term> out(terms.size());
  for (int a = 0; a < 3; ++a) {
    const int Na = N[a], M = Na / 2;
    const std::vector<long double> D = p1_dense(Na, L[a], ds, dL);
    for (size_t t = 0; t < terms.size(); ++t) {
      std::vector<double>& At = out[t].At[a];
      At.assign((size_t)M * M, 0.0);
      for (int j = 0; j < M; ++j)      // output (row of B)
        for (int i = 0; i < M; ++i) {  // input
          long double v = 0.0L;
          for (const Base& b : terms[t][a]) {
            const long img = b.first ? -1 - 2L * i + b.second : 2L * i + b.second;
            const long col = ((img % Na) + Na) % Na;
            v += D[(size_t)(2 * j) * Na + col];
          }
          At[(size_t)i * M + j] = (double)v;
        }
    }
  }
  return out;
}

std::vector<double> centro_half(const std::vector<double>& At, int M, bool shift) {
  const int H = M / 2;
  std::vector<double> c((size_t)H * M);
  for (int k = 0; k < H; ++k)
    for (int j = 0; j < H; ++j) {
      const int kp = shift ? H + k : M - 1 - k;  // partner of k
      const double a = At[(size_t)k * M + j], b = At[(size_t)kp * M + j];
      c[(size_t)k * M + j] = 0.5 * (a + b);
      c[(size_t)k * M + H + j] = 0.5 * (a - b);
    }
  return c;
}

std::vector<double> radix_table(int kind, int M, double L, double ds) {
  std::vector<double> t(5 * (size_t)M);
  for (int j = 0; j < M; ++j) {
    // the angles 2 pi j / M and pi j / (2M) in long double
    const long double a = 2.0L * kPiL * (long double)j / (long double)M;
    const long double b = kPiL * (long double)j / (2.0L * (long double)M);
    t[2 * j] = (double)cosl(a);
    t[2 * j + 1] = (double)sinl(a);
    t[2 * M + 2 * j] = (double)cosl(b);
    t[2 * M + 2 * j + 1] = (double)sinl(b);
    if (kind == 0)
      t[4 * M + j] = (double)(kfactor(j < (M + 1) / 2 ? j : M - j, L, ds, false) / (long double)M);
    else
      t[4 * M + j] = (double)(kfactor(j, L, ds, false) / (2.0L * (long double)M));
  }
  return t;
}

unsigned short bf16_bits(float x) {
  unsigned u;
  std::memcpy(&u, &x, sizeof(u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

}  // namespace crys
