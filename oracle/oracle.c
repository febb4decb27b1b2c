/*
 * CPU FP64 oracle: a plain-C restatement of the reference's propagator path.
 *
 * TEST INFRASTRUCTURE ONLY -- only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this.  The HIP library never uses it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The transforms are written from FFTW's published definitions (REDFT10,
 * REDFT01, r2c, c2r) because FFTW3 itself is not available in this image.
 * Pinning of each part is listed in oracle.h.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* solver_base.h:4-6 */
static const double kPi = 3.1415926535897932384626;

/* ------------------------------------------------------------------------ */
/* Complex FFT plans (own code): radix-2 for powers of two, direct DFT else. */
/* ------------------------------------------------------------------------ */
typedef struct {
  size_t n;
  int pow2;
  double* wr;   /* cos(2 pi k / n), k < n */
  double* wi;   /* -sin(2 pi k / n): forward twiddle e^{-2 pi i k / n} */
  size_t* rev;  /* bit reversal (pow2 only) */
  double* hr;   /* cos(pi k / (2n)), k < n  (Makhoul DCT twiddles) */
  double* hi;   /* sin(pi k / (2n)) */
} fft_plan;

#define MAX_PLANS 64
static fft_plan g_plans[MAX_PLANS];
static int g_nplans = 0;
static int g_threads = 1;

void orc_set_num_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_num_threads(void) { return g_threads; }

/* cos/sin of 2*pi*m/d with exact integer range reduction. */
static void unit_root(size_t m, size_t d, double* c, double* s) {
  m %= d;
  double a = 2.0 * kPi * (double)m / (double)d;
  *c = cos(a);
  *s = sin(a);
}

static const fft_plan* get_plan(size_t n) {
  for (int i = 0; i < g_nplans; ++i)
    if (g_plans[i].n == n) return &g_plans[i];
  if (g_nplans == MAX_PLANS) {
    fprintf(stderr, "oracle: plan cache full\n");
    abort();
  }
  fft_plan* p = &g_plans[g_nplans];
  p->n = n;
  p->pow2 = (n & (n - 1)) == 0;
  p->wr = (double*)malloc(sizeof(double) * n);
  p->wi = (double*)malloc(sizeof(double) * n);
  p->hr = (double*)malloc(sizeof(double) * n);
  p->hi = (double*)malloc(sizeof(double) * n);
  p->rev = NULL;
  for (size_t k = 0; k < n; ++k) {
    double c, s;
    unit_root(k, n, &c, &s);
    p->wr[k] = c;
    p->wi[k] = -s;
    unit_root(k, 4 * n, &c, &s); /* angle pi k / (2n) */
    p->hr[k] = c;
    p->hi[k] = s;
  }
  if (p->pow2 && n > 1) {
    size_t bits = 0;
    while (((size_t)1 << bits) < n) ++bits;
    p->rev = (size_t*)malloc(sizeof(size_t) * n);
    for (size_t i = 0; i < n; ++i) {
      size_t r = 0;
      for (size_t b = 0; b < bits; ++b)
        if (i & ((size_t)1 << b)) r |= (size_t)1 << (bits - 1 - b);
      p->rev[i] = r;
    }
  }
  ++g_nplans;
  return p;
}

/* In-place complex DFT of re/im (length p->n). sign = -1 forward, +1 backward.
 * tmp must hold 2n doubles (used by the direct path). */
static void cfft(const fft_plan* p, double* re, double* im, int sign, double* tmp) {
  const size_t n = p->n;
  if (n <= 1) return;
  const double sg = sign < 0 ? 1.0 : -1.0; /* multiplies wi */
  if (p->pow2) {
    for (size_t i = 0; i < n; ++i) {
      size_t r = p->rev[i];
      if (r > i) {
        double t = re[i]; re[i] = re[r]; re[r] = t;
        t = im[i]; im[i] = im[r]; im[r] = t;
      }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
      size_t half = len >> 1, step = n / len;
      for (size_t i = 0; i < n; i += len) {
        for (size_t j = 0; j < half; ++j) {
          double wr = p->wr[j * step], wi = sg * p->wi[j * step];
          double ur = re[i + j], ui = im[i + j];
          double xr = re[i + j + half], xi = im[i + j + half];
          double vr = xr * wr - xi * wi, vi = xr * wi + xi * wr;
          re[i + j] = ur + vr;
          im[i + j] = ui + vi;
          re[i + j + half] = ur - vr;
          im[i + j + half] = ui - vi;
        }
      }
    }
    return;
  }
  double* tr = tmp;
  double* ti = tmp + n;
  for (size_t k = 0; k < n; ++k) {
    double sr = 0.0, si = 0.0;
    for (size_t j = 0; j < n; ++j) {
      size_t m = (j * k) % n;
      double wr = p->wr[m], wi = sg * p->wi[m];
      sr += re[j] * wr - im[j] * wi;
      si += re[j] * wi + im[j] * wr;
    }
    tr[k] = sr;
    ti[k] = si;
  }
  memcpy(re, tr, sizeof(double) * n);
  memcpy(im, ti, sizeof(double) * n);
}

/* REDFT10 of one line (Makhoul): Y_k = 2 Re(e^{-i pi k/(2n)} DFT(v)_k),
 * v_j = x_{2j}, v_{n-1-j} = x_{2j+1}.  Work buffer: 4n doubles. */
static void dct2_line(const fft_plan* p, const double* x, double* y, double* w) {
  const size_t n = p->n;
  double* re = w;
  double* im = w + n;
  double* tmp = w + 2 * n;
  for (size_t j = 0; 2 * j < n; ++j) re[j] = x[2 * j];
  for (size_t j = 0; 2 * j + 1 < n; ++j) re[n - 1 - j] = x[2 * j + 1];
  for (size_t j = 0; j < n; ++j) im[j] = 0.0;
  cfft(p, re, im, -1, tmp);
  for (size_t k = 0; k < n; ++k) y[k] = 2.0 * (p->hr[k] * re[k] + p->hi[k] * im[k]);
}

/* REDFT01 of one line (inverse Makhoul):
 * Z_k = (X_k - i X_{n-k}) e^{+i pi k/(2n)}, z = sum_k Z_k e^{+2 pi i k m / n},
 * y_{2j} = Re z_j, y_{2j+1} = Re z_{n-1-j}.  Work buffer: 4n doubles. */
static void dct3_line(const fft_plan* p, const double* x, double* y, double* w) {
  const size_t n = p->n;
  double* re = w;
  double* im = w + n;
  double* tmp = w + 2 * n;
  for (size_t k = 0; k < n; ++k) {
    double a = x[k];
    double b = k == 0 ? 0.0 : -x[n - k];
    re[k] = a * p->hr[k] - b * p->hi[k];
    im[k] = a * p->hi[k] + b * p->hr[k];
  }
  cfft(p, re, im, +1, tmp);
  for (size_t j = 0; 2 * j < n; ++j) y[2 * j] = re[j];
  for (size_t j = 0; 2 * j + 1 < n; ++j) y[2 * j + 1] = re[n - 1 - j];
}

/* Apply a real 1-D line transform along one axis of an nx*ny*nz array. */
typedef void (*line_fn)(const fft_plan*, const double*, double*, double*);

static void apply_axis_real(line_fn fn, double* a, size_t nx, size_t ny, size_t nz, int axis) {
  size_t dims[3] = {nx, ny, nz};
  size_t len = dims[axis];
  size_t stride = axis == 2 ? 1 : (axis == 1 ? nz : ny * nz);
  size_t nlines = nx * ny * nz / len;
  const fft_plan* p = get_plan(len);
  long long L = (long long)nlines;
#pragma omp parallel num_threads(g_threads)
  {
    double* buf = (double*)malloc(sizeof(double) * 6 * len);
    double* xin = buf + 4 * len;
    double* yout = buf + 5 * len;
#pragma omp for schedule(static)
    for (long long l = 0; l < L; ++l) {
      size_t base;
      if (axis == 2) base = (size_t)l * nz;
      else if (axis == 1) base = ((size_t)l / nz) * ny * nz + ((size_t)l % nz);
      else base = (size_t)l;
      for (size_t i = 0; i < len; ++i) xin[i] = a[base + i * stride];
      fn(p, xin, yout, buf);
      for (size_t i = 0; i < len; ++i) a[base + i * stride] = yout[i];
    }
    free(buf);
  }
}

void orc_dct2_3d(const double* in, double* out, size_t nx, size_t ny, size_t nz) {
  if (out != in) memcpy(out, in, sizeof(double) * nx * ny * nz);
  for (int ax = 2; ax >= 0; --ax) apply_axis_real(dct2_line, out, nx, ny, nz, ax);
}

void orc_dct3_3d(const double* in, double* out, size_t nx, size_t ny, size_t nz) {
  if (out != in) memcpy(out, in, sizeof(double) * nx * ny * nz);
  for (int ax = 2; ax >= 0; --ax) apply_axis_real(dct3_line, out, nx, ny, nz, ax);
}

/* Complex transform along axis 0 or 1 of an nx*ny*nzc complex array. */
static void apply_axis_complex(double* a, size_t nx, size_t ny, size_t nzc, int axis, int sign) {
  size_t len = axis == 0 ? nx : ny;
  size_t stride = axis == 0 ? ny * nzc : nzc;
  size_t nlines = nx * ny * nzc / len;
  const fft_plan* p = get_plan(len);
  long long L = (long long)nlines;
#pragma omp parallel num_threads(g_threads)
  {
    double* buf = (double*)malloc(sizeof(double) * 4 * len);
#pragma omp for schedule(static)
    for (long long l = 0; l < L; ++l) {
      size_t base;
      if (axis == 1) base = ((size_t)l / nzc) * ny * nzc + ((size_t)l % nzc);
      else base = (size_t)l;
      double* re = buf;
      double* im = buf + len;
      for (size_t i = 0; i < len; ++i) {
        re[i] = a[2 * (base + i * stride)];
        im[i] = a[2 * (base + i * stride) + 1];
      }
      cfft(p, re, im, sign, buf + 2 * len);
      for (size_t i = 0; i < len; ++i) {
        a[2 * (base + i * stride)] = re[i];
        a[2 * (base + i * stride) + 1] = im[i];
      }
    }
    free(buf);
  }
}

/* FFTW r2c (forward, unnormalised): out[x][y][z], z <= nz/2. */
void orc_rfft_3d(const double* in, double* out, size_t nx, size_t ny, size_t nz) {
  const size_t nzc = nz / 2 + 1;
  const fft_plan* p = get_plan(nz);
  long long L = (long long)(nx * ny);
#pragma omp parallel num_threads(g_threads)
  {
    double* buf = (double*)malloc(sizeof(double) * 4 * nz);
#pragma omp for schedule(static)
    for (long long l = 0; l < L; ++l) {
      double* re = buf;
      double* im = buf + nz;
      for (size_t z = 0; z < nz; ++z) {
        re[z] = in[(size_t)l * nz + z];
        im[z] = 0.0;
      }
      cfft(p, re, im, -1, buf + 2 * nz);
      for (size_t z = 0; z < nzc; ++z) {
        out[2 * ((size_t)l * nzc + z)] = re[z];
        out[2 * ((size_t)l * nzc + z) + 1] = im[z];
      }
    }
    free(buf);
  }
  apply_axis_complex(out, nx, ny, nzc, 1, -1);
  apply_axis_complex(out, nx, ny, nzc, 0, -1);
}

/* FFTW c2r (backward, unnormalised).  The z line is Hermitian-extended from
 * its nz/2+1 stored entries; the imaginary parts of the self-conjugate
 * entries do not contribute, as in FFTW. */
void orc_irfft_3d(const double* in, double* out, size_t nx, size_t ny, size_t nz) {
  const size_t nzc = nz / 2 + 1;
  double* t = (double*)malloc(sizeof(double) * 2 * nx * ny * nzc);
  memcpy(t, in, sizeof(double) * 2 * nx * ny * nzc);
  apply_axis_complex(t, nx, ny, nzc, 0, +1);
  apply_axis_complex(t, nx, ny, nzc, 1, +1);
  const fft_plan* p = get_plan(nz);
  long long L = (long long)(nx * ny);
#pragma omp parallel num_threads(g_threads)
  {
    double* buf = (double*)malloc(sizeof(double) * 4 * nz);
#pragma omp for schedule(static)
    for (long long l = 0; l < L; ++l) {
      double* re = buf;
      double* im = buf + nz;
      const double* h = t + 2 * (size_t)l * nzc;
      for (size_t z = 0; z < nz; ++z) {
        if (z < nzc) {
          re[z] = h[2 * z];
          im[z] = h[2 * z + 1];
        } else {
          re[z] = h[2 * (nz - z)];
          im[z] = -h[2 * (nz - z) + 1];
        }
      }
      /* self-conjugate entries contribute their real part only */
      im[0] = 0.0;
      if (nz % 2 == 0) im[nz / 2] = 0.0;
      cfft(p, re, im, +1, buf + 2 * nz);
      for (size_t z = 0; z < nz; ++z) out[(size_t)l * nz + z] = re[z];
    }
    free(buf);
  }
  free(t);
}

/* ------------------------------------------------------------------------ */
/* Solvers                                                                   */
/* ------------------------------------------------------------------------ */
#define MAX_K 16

typedef struct {
  double ds;
  double* k;    /* P1: nx*ny*(nz/2+1); Pmmm: n_phy */
  double* kre;  /* 3m: 8 * n_phy */
  double* kim;  /* 3m: 8 * n_phy */
} kentry;

struct orc_solver {
  int kind;
  size_t nlog[3], nphy[3];
  size_t n_log, n_phy;
  double cell[6];
  double ds_current;
  double g[9];
  double* iomat;
  double* step1;  /* P1: complex half array; Pmmm: n_phy; 3m: complex n_phy */
  double* step2;
  double* r_re;   /* 3m twiddles: 8 * n_phy */
  double* r_im;
  kentry kc[MAX_K];
  int nk;
  kentry* cur;
};

static void k_destroy(orc_solver* s) {
  for (int i = 0; i < s->nk; ++i) {
    free(s->kc[i].k);
    free(s->kc[i].kre);
    free(s->kc[i].kim);
  }
  memset(s->kc, 0, sizeof(s->kc));
  s->nk = 0;
  s->cur = NULL;
}

static void set_err(char* err, size_t errlen, const char* msg) {
  if (err && errlen) {
    strncpy(err, msg, errlen - 1);
    err[errlen - 1] = 0;
  }
}

/* 3m twiddles: R_mat_generator_full (solver_3m.h:172-182) with the symmetry
 * operation phases of generate_m3_operation (solver_utilities.h:55-66) and
 * twiddle_factor (solver_utilities.h:27-53) over the half-size grid. */
static void m3_twiddles(orc_solver* s) {
  const double* g = s->g;
  const double sx = (double)s->nlog[0], sy = (double)s->nlog[1], sz = (double)s->nlog[2];
  const double tp = -2.0 * kPi;
  double op[7][3] = {
      {tp * (g[0]), tp * (g[1]), tp * (g[2] - 1. / sz)},
      {tp * (g[3]), tp * (g[4] - 1. / sy), tp * (g[5])},
      {tp * (g[0] + g[3]), tp * (-g[1] + g[4] - 1. / sy), tp * (g[2] + g[5] - 1. / sz)},
      {tp * (g[6] - 1. / sx), tp * (g[7]), tp * (g[8])},
      {tp * (g[6] - g[0] - 1. / sx), tp * (g[1] + g[7]), tp * (g[2] + g[8] - 1. / sz)},
      {tp * (-g[3] + g[6] - 1. / sx), tp * (g[4] + g[7] - 1. / sy), tp * (g[5] + g[8])},
      {tp * (-g[0] - g[3] + g[6] - 1. / sx), tp * (-g[1] + g[4] + g[7] - 1. / sy),
       tp * (g[2] + g[5] + g[8] - 1. / sz)},
  };
  const size_t X = s->nphy[0], Y = s->nphy[1], Z = s->nphy[2], n = s->n_phy;
  for (size_t i = 0; i < n; ++i) {
    s->r_re[i] = 1.0;
    s->r_im[i] = 0.0;
  }
  for (int p = 1; p < 8; ++p) {
    double* dc = s->r_re + (size_t)p * n;
    double* dsn = s->r_im + (size_t)p * n;
    for (size_t x = 0; x < X; ++x)
      for (size_t y = 0; y < Y; ++y)
        for (size_t z = 0; z < Z; ++z) {
          double angle = x * op[p - 1][0] + y * op[p - 1][1] + z * op[p - 1][2];
          dc[(x * Y + y) * Z + z] = cos(angle);
          dsn[(x * Y + y) * Z + z] = sin(angle);
        }
  }
}

orc_solver* orc_solver_create(int kind, const size_t grid_log[3], const double g[9], char* err,
                              size_t errlen) {
  size_t Nx = grid_log[0], Ny = grid_log[1], Nz = grid_log[2];
  if (kind == ORC_P1) {
    /* solver_p1.h:95-125 */
    if (Nx == 0 || Ny == 0 || Nz == 0) {
      set_err(err, errlen, "ERROR: Bad Grid Size For Group Triclinic P1");
      return NULL;
    }
    if (Nz % 8 != 0) {
      set_err(err, errlen, "ERROR: Bad Grid Size For Group Triclinic P1, alignment cannot be guaranteed");
      return NULL;
    }
  } else if (kind == ORC_M3 || kind == ORC_PMMM) {
    /* solver_3m.h:73-113, solver_Pmmm.h:66-96 */
    if (Nx % 2 || Ny % 2 || Nz % 2) {
      set_err(err, errlen, "ERROR: Grid Size Must Be Even");
      return NULL;
    }
    if (Nx == 0 || Ny == 0 || Nz == 0) {
      set_err(err, errlen, "ERROR: Bad Grid Size");
      return NULL;
    }
    if ((Nz / 2) % 8 != 0) {
      set_err(err, errlen, "ERROR: Bad Grid Size, Alignment Cannot Be Guaranteed");
      return NULL;
    }
  } else {
    set_err(err, errlen, "ERROR: unknown solver kind");
    return NULL;
  }
  orc_solver* s = (orc_solver*)calloc(1, sizeof(orc_solver));
  s->kind = kind;
  for (int a = 0; a < 3; ++a) {
    s->nlog[a] = grid_log[a];
    s->nphy[a] = kind == ORC_P1 ? grid_log[a] : grid_log[a] / 2;
  }
  s->n_log = Nx * Ny * Nz;
  s->n_phy = s->nphy[0] * s->nphy[1] * s->nphy[2];
  /* solver_base.h:43 default cell; :46 default ds */
  double c0[6] = {4, 4, 4, 0, 0, 0};
  memcpy(s->cell, c0, sizeof(c0));
  s->ds_current = 0.0;
  for (int i = 0; i < 9; ++i) s->g[i] = g ? g[i] : 0.0;
  s->iomat = (double*)calloc(s->n_phy, sizeof(double));
  if (kind == ORC_P1) {
    s->step1 = (double*)calloc(2 * Nx * Ny * (Nz / 2 + 1), sizeof(double));
  } else if (kind == ORC_PMMM) {
    s->step1 = (double*)calloc(s->n_phy, sizeof(double));
    s->step2 = (double*)calloc(s->n_phy, sizeof(double));
  } else {
    s->step1 = (double*)calloc(2 * s->n_phy, sizeof(double));
    s->step2 = (double*)calloc(2 * s->n_phy, sizeof(double));
    s->r_re = (double*)calloc(8 * s->n_phy, sizeof(double));
    s->r_im = (double*)calloc(8 * s->n_phy, sizeof(double));
    m3_twiddles(s);
  }
  return s;
}

void orc_solver_destroy(orc_solver* s) {
  if (!s) return;
  k_destroy(s);
  free(s->iomat);
  free(s->step1);
  free(s->step2);
  free(s->r_re);
  free(s->r_im);
  free(s);
}

void orc_set_cell_para(orc_solver* s, const double c[6]) {
  if (s->kind == ORC_P1) {
    /* solver_p1.h:64-69: always resets */
    double nc[6] = {c[0], c[1], c[2], 0, 0, 0};
    memcpy(s->cell, nc, sizeof(nc));
  } else {
    /* solver_3m.h:115-121, solver_Pmmm.h:98-104: early return when equal */
    if (c[0] == s->cell[0] && c[1] == s->cell[1] && c[2] == s->cell[2] && c[3] == 0 && c[4] == 0 &&
        c[5] == 0)
      return;
    if (s->kind == ORC_PMMM) {
      /* solver_Pmmm.h:101: cubic only, every length is cell_para[0] */
      double nc[6] = {c[0], c[0], c[0], 0, 0, 0};
      memcpy(s->cell, nc, sizeof(nc));
    } else {
      double nc[6] = {c[0], c[1], c[2], 0, 0, 0};
      memcpy(s->cell, nc, sizeof(nc));
    }
  }
  k_destroy(s);
  s->ds_current = 0.0;
}

/* solver_p1.h:142-172 */
static void k_gen_p1(orc_solver* s, double* K, double dz) {
  size_t Nx = s->nlog[0], Ny = s->nlog[1], Nz = s->nlog[2];
  double Lx = s->cell[0], Ly = s->cell[1], Lz = s->cell[2];
  double* kx = (double*)malloc(sizeof(double) * Nx);
  double* ky = (double*)malloc(sizeof(double) * Ny);
  double* kz = (double*)malloc(sizeof(double) * (Nz / 2 + 1));
  size_t i;
  for (i = 0; i < (Nx + 1) / 2; i++) kx[i] = 2 * kPi * i * 1.0 / Lx;
  for (i = (Nx + 1) / 2; i < Nx; i++) kx[i] = 2 * kPi * (Nx - i) * 1.0 / Lx;
  for (i = 0; i < Nx; i++) kx[i] *= kx[i];
  for (i = 0; i < (Ny + 1) / 2; i++) ky[i] = 2 * kPi * i * 1.0 / Ly;
  for (i = (Ny + 1) / 2; i < Ny; i++) ky[i] = 2 * kPi * (Ny - i) * 1.0 / Ly;
  for (i = 0; i < Ny; i++) ky[i] *= ky[i];
  for (i = 0; i < Nz / 2 + 1; i++) kz[i] = 2 * kPi * i * 1.0 / Lz;
  for (i = 0; i < Nz / 2 + 1; i++) kz[i] *= kz[i];
  size_t nzc = Nz / 2 + 1;
  for (size_t x = 0; x < Nx; ++x)
    for (size_t y = 0; y < Ny; ++y)
      for (size_t m = 0; m < nzc; ++m) K[(x * Ny + y) * nzc + m] = exp(-dz * (kx[x] + ky[y] + kz[m]));
  free(kx);
  free(ky);
  free(kz);
}

/* solver_Pmmm.h:159-186 */
static void k_gen_pmmm(orc_solver* s, double* K, double dz, double L) {
  size_t Nx = s->nlog[0], Ny = s->nlog[1], Nz = s->nlog[2];
  double* kx = (double*)malloc(sizeof(double) * (Nx / 2));
  double* ky = (double*)malloc(sizeof(double) * (Ny / 2));
  double* kz = (double*)malloc(sizeof(double) * (Nz / 2));
  size_t i, j, k, ijk;
  for (i = 0; i < Nx / 2; i++) kx[i] = 2 * kPi * i * 1.0 / L;
  for (i = 0; i < Nx / 2; i++) kx[i] *= kx[i];
  for (j = 0; j < Ny / 2; j++) ky[j] = 2 * kPi * j * 1.0 / L;
  for (j = 0; j < Ny / 2; j++) ky[j] *= ky[j];
  for (k = 0; k < Nz / 2; k++) kz[k] = 2 * kPi * k * 1.0 / L;
  for (k = 0; k < Nz / 2; k++) kz[k] *= kz[k];
  ijk = 0;
  for (i = 0; i < Nx / 2; ++i)
    for (j = 0; j < Ny / 2; ++j)
      for (k = 0; k < Nz / 2; ++k) {
        K[ijk] = exp(-(kx[i] + ky[j] + kz[k]) * dz);
        ++ijk;
      }
  free(kx);
  free(ky);
  free(kz);
}

/* z extent of the 3m coefficient arrays ("cut half", solver_utilities.h:80,
 * :112, :172; equals Nz_qua of solver_3m.h:287). */
static size_t m3_nz_cut(size_t Nz2) { return ((Nz2 / 2 + 1) / 8 + 1) * 8; }

/* solver_3m.h:184-255 with mat_split / add_sub_by_seq / mul_add_sub_by_seq
 * (solver_utilities.h:74-208). */
static void k_gen_3m(orc_solver* s, double* kre, double* kim, double dz) {
  const size_t Nx = s->nlog[0], Ny = s->nlog[1], Nz = s->nlog[2];
  const size_t X = Nx / 2, Y = Ny / 2, Z = Nz / 2, n = s->n_phy;
  const double Lx = s->cell[0], Ly = s->cell[1], Lz = s->cell[2];
  double* kx = (double*)malloc(sizeof(double) * Nx);
  double* ky = (double*)malloc(sizeof(double) * Ny);
  double* kz = (double*)malloc(sizeof(double) * Nz);
  double fx = 2. * kPi / Lx;
  fx *= fx;
  double fy = 2. * kPi / Ly;
  fy *= fy;
  double fz = 2. * kPi / Lz;
  fz *= fz;
  for (size_t i = 0; i < Nx; ++i) {
    size_t t = i > Nx / 2 ? Nx - i : i;
    kx[i] = t * t * fx;
  }
  for (size_t i = 0; i < Ny; ++i) {
    size_t t = i > Ny / 2 ? Ny - i : i;
    ky[i] = t * t * fy;
  }
  for (size_t i = 0; i < Nz; ++i) {
    size_t t = i > Nz / 2 ? Nz - i : i;
    kz[i] = t * t * fz;
  }
  double* full = (double*)malloc(sizeof(double) * s->n_log);
  double factor = 1. / s->n_log;
  for (size_t x = 0; x < Nx; ++x)
    for (size_t y = 0; y < Ny; ++y)
      for (size_t z = 0; z < Nz; ++z)
        full[(x * Ny + y) * Nz + z] = exp(-(kx[x] + ky[y] + kz[z]) * dz) * factor;
  free(kx);
  free(ky);
  free(kz);

  const size_t zc = m3_nz_cut(Z);
  double* sp = (double*)calloc(8 * n, sizeof(double));
  /* mat_split: part p takes the octant offset (px*X, py*Y, pz*Z) */
  for (int p = 0; p < 8; ++p) {
    size_t ox = (size_t)((p & 4) >> 2) * X, oy = (size_t)((p & 2) >> 1) * Y, oz = (size_t)(p & 1) * Z;
    for (size_t x = 0; x < X; ++x)
      for (size_t y = 0; y < Y; ++y)
        for (size_t z = 0; z < zc; ++z)
          sp[p * n + (x * Y + y) * Z + z] = full[((x + ox) * Ny + y + oy) * Nz + z + oz];
  }
  free(full);
  /* three add/sub stages over the pairs that differ in the x, y, z bit */
  static const int seq[3][4][2] = {{{0, 4}, {1, 5}, {2, 6}, {3, 7}},
                                   {{0, 2}, {1, 3}, {4, 6}, {5, 7}},
                                   {{0, 1}, {2, 3}, {4, 5}, {6, 7}}};
  for (int st = 0; st < 3; ++st)
    for (int q = 0; q < 4; ++q) {
      double* d0 = sp + seq[st][q][0] * n;
      double* d1 = sp + seq[st][q][1] * n;
      for (size_t xy = 0; xy < X * Y; ++xy)
        for (size_t z = 0; z < zc; ++z) {
          double a = d0[xy * Z + z], b = d1[xy * Z + z];
          d0[xy * Z + z] = a + b;
          d1[xy * Z + z] = a - b;
        }
    }
  /* multiply by the twiddles, then butterfly the pairs (0,7),(6,1),(2,5),(4,3) */
  static const int pairs[4][2] = {{0, 7}, {6, 1}, {2, 5}, {4, 3}};
  for (int q = 0; q < 4; ++q) {
    int a = pairs[q][0], b = pairs[q][1];
    for (size_t xy = 0; xy < X * Y; ++xy)
      for (size_t z = 0; z < zc; ++z) {
        size_t i = xy * Z + z;
        double ra = sp[a * n + i] * s->r_re[a * n + i];
        double ia = sp[a * n + i] * s->r_im[a * n + i];
        double rb = sp[b * n + i] * s->r_re[b * n + i];
        double ib = sp[b * n + i] * s->r_im[b * n + i];
        kre[a * n + i] = ra + rb;
        kre[b * n + i] = ra - rb;
        kim[a * n + i] = ia + ib;
        kim[b * n + i] = ia - ib;
      }
  }
  free(sp);
}

void orc_set_contour_step(orc_solver* s, double ds) {
  /* solver_p1.h:71-81, solver_3m.h:123-141, solver_Pmmm.h:106-123 */
  if (s->nk > 0 && ds == s->ds_current) return;
  s->ds_current = ds;
  for (int i = 0; i < s->nk; ++i)
    if (s->kc[i].ds == ds) {
      s->cur = &s->kc[i];
      return;
    }
  if (s->nk == MAX_K) {
    fprintf(stderr, "oracle: K cache full\n");
    abort();
  }
  kentry* e = &s->kc[s->nk++];
  e->ds = ds;
  if (s->kind == ORC_P1) {
    e->k = (double*)calloc(s->nlog[0] * s->nlog[1] * (s->nlog[2] / 2 + 1), sizeof(double));
    k_gen_p1(s, e->k, ds);
  } else if (s->kind == ORC_PMMM) {
    e->k = (double*)calloc(s->n_phy, sizeof(double));
    k_gen_pmmm(s, e->k, ds, s->cell[0]);
  } else {
    e->kre = (double*)calloc(8 * s->n_phy, sizeof(double));
    e->kim = (double*)calloc(8 * s->n_phy, sizeof(double));
    k_gen_3m(s, e->kre, e->kim, ds);
  }
  s->cur = e;
}

static void diffusion_p1(orc_solver* s) {
  /* solver_p1.h:83-93, :184-202 */
  size_t Nx = s->nlog[0], Ny = s->nlog[1], Nz = s->nlog[2], nzc = Nz / 2 + 1;
  orc_rfft_3d(s->iomat, s->step1, Nx, Ny, Nz);
  const double* K = s->cur->k;
  for (size_t i = 0; i < Nx * Ny * nzc; ++i) {
    s->step1[2 * i] *= K[i];
    s->step1[2 * i + 1] *= K[i];
  }
  orc_irfft_3d(s->step1, s->iomat, Nx, Ny, Nz);
  double c = 1. / s->n_log;
  for (size_t i = 0; i < s->n_phy; ++i) s->iomat[i] *= c;
}

static void diffusion_pmmm(orc_solver* s) {
  /* solver_Pmmm.h:125-143, :194-215 */
  size_t X = s->nphy[0], Y = s->nphy[1], Z = s->nphy[2];
  orc_dct2_3d(s->iomat, s->step1, X, Y, Z);
  const double* K = s->cur->k;
  for (size_t i = 0; i < s->n_phy; ++i) s->step2[i] = s->step1[i] * K[i];
  orc_dct3_3d(s->step2, s->iomat, X, Y, Z);
  double f = 1. / s->n_log;
  for (size_t i = 0; i < s->n_phy; ++i) s->iomat[i] *= f;
}

static void diffusion_3m(orc_solver* s) {
  /* solver_3m.h:143-146, :281-324 */
  const size_t X = s->nphy[0], Y = s->nphy[1], Z = s->nphy[2], n = s->n_phy;
  const size_t zh = Z / 2 + 1;
  const size_t zq = m3_nz_cut(Z);
  /* guru r2c into rows of stride Z complex entries, Z/2+1 valid */
  double* half = (double*)malloc(sizeof(double) * 2 * X * Y * zh);
  orc_rfft_3d(s->iomat, half, X, Y, Z);
  memset(s->step1, 0, sizeof(double) * 2 * n);
  for (size_t xy = 0; xy < X * Y; ++xy)
    for (size_t z = 0; z < zh; ++z) {
      s->step1[2 * (xy * Z + z)] = half[2 * (xy * zh + z)];
      s->step1[2 * (xy * Z + z) + 1] = half[2 * (xy * zh + z) + 1];
    }
  const double* kre = s->cur->kre;
  const double* kim = s->cur->kim;
  const double* st = s->step1;
  for (size_t x = 0; x < X; ++x)
    for (size_t y = 0; y < Y; ++y) {
      size_t b = (x * Y + y) * Z;
      const double* s000 = st + 2 * b;
      const double* s110 = st + 2 * ((((X - x) % X) * Y + (Y - y) % Y) * Z);
      const double* s010 = st + 2 * ((x * Y + (Y - y) % Y) * Z);
      const double* s100 = st + 2 * ((((X - x) % X) * Y + y) * Z);
      double* d = s->step2 + 2 * b;
      for (size_t z = 0; z < zq; ++z) {
        size_t i = b + z;
        d[2 * z] = s000[2 * z] * kre[0 * n + i] - s000[2 * z + 1] * kim[7 * n + i] +
                   s110[2 * z] * kre[6 * n + i] - s110[2 * z + 1] * kim[1 * n + i] +
                   s010[2 * z] * kre[2 * n + i] - s010[2 * z + 1] * kim[5 * n + i] +
                   s100[2 * z] * kre[4 * n + i] - s100[2 * z + 1] * kim[3 * n + i];
        d[2 * z + 1] = s000[2 * z] * kim[0 * n + i] + s000[2 * z + 1] * kre[7 * n + i] +
                       s110[2 * z] * kim[6 * n + i] + s110[2 * z + 1] * kre[1 * n + i] +
                       s010[2 * z] * kim[2 * n + i] + s010[2 * z + 1] * kre[5 * n + i] +
                       s100[2 * z] * kim[4 * n + i] + s100[2 * z + 1] * kre[3 * n + i];
      }
    }
  /* guru c2r reading the first Z/2+1 entries of each row */
  for (size_t xy = 0; xy < X * Y; ++xy)
    for (size_t z = 0; z < zh; ++z) {
      half[2 * (xy * zh + z)] = s->step2[2 * (xy * Z + z)];
      half[2 * (xy * zh + z) + 1] = s->step2[2 * (xy * Z + z) + 1];
    }
  orc_irfft_3d(half, s->iomat, X, Y, Z);
  free(half);
}

void orc_diffusion(orc_solver* s) {
  if (!s->cur) {
    fprintf(stderr, "oracle: diffusion before set_contour_step\n");
    abort();
  }
  if (s->kind == ORC_P1) diffusion_p1(s);
  else if (s->kind == ORC_PMMM) diffusion_pmmm(s);
  else diffusion_3m(s);
}

double* orc_get_iomat(orc_solver* s) { return s->iomat; }
size_t orc_get_space_size_phy(const orc_solver* s) { return s->n_phy; }
void orc_get_grid_size_phy(const orc_solver* s, size_t out[3]) {
  for (int a = 0; a < 3; ++a) out[a] = s->nphy[a];
}
void orc_get_cell_para(const orc_solver* s, double out[6]) { memcpy(out, s->cell, sizeof(s->cell)); }
double orc_get_ds(const orc_solver* s) { return s->ds_current; }

/* ------------------------------------------------------------------------ */
/* Driver (crysSCFT_sample.cpp:294-531)                                      */
/* ------------------------------------------------------------------------ */
void orc_propagate(orc_solver* s, double* q, const double* init, const double* E, size_t Ns,
                   int direction) {
  /* crysSCFT_sample.cpp:373-396 */
  const size_t n = s->n_phy;
  double* io = s->iomat;
  double* cur = direction == 1 ? q : q + n * Ns;
  memcpy(io, init, sizeof(double) * n);
  memcpy(cur, init, sizeof(double) * n);
  for (size_t it = 0; it < Ns; ++it) {
    for (size_t i = 0; i < n; ++i) io[i] *= E[i];
    orc_diffusion(s);
    for (size_t i = 0; i < n; ++i) io[i] *= E[i];
    cur = direction == 1 ? cur + n : cur - n;
    memcpy(cur, io, sizeof(double) * n);
  }
}

/* crysSCFT_sample.cpp:348-360 */
static void integrate_phi(double* phi, const double* q, const double* qc, double f, double Q,
                          size_t size, size_t Ns) {
  for (size_t s = 0; s <= Ns; ++s) {
    double factor = s == 0 ? 0.5 : (s == Ns ? 0.5 : 1);
    factor *= f / (Ns * Q);
    for (size_t i = 0; i < size; ++i) phi[i] += factor * q[s * size + i] * qc[s * size + i];
  }
}

int orc_scft_simple_iterate(orc_solver* s, double fA, double ds, double xN, double Lx, double Ly,
                            double Lz, double acceptance, size_t Nstep, double* wA, double* wB,
                            double* phiA, double* phiB, double* trace) {
  if (!s || !wA || !wB || !phiA || !phiB || !(ds > 0)) return -1;
  const size_t n = s->n_phy;
  const double fB = 1 - fA;
  const size_t NsA = (size_t)(fA / ds); /* :443 truncation */
  const size_t NsB = (size_t)(fB / ds); /* :444 */
  double* wAds = (double*)malloc(sizeof(double) * n);
  double* wBds = (double*)malloc(sizeof(double) * n);
  double* ksi = (double*)malloc(sizeof(double) * n);
  double* init = (double*)malloc(sizeof(double) * n);
  double* qA = (double*)malloc(sizeof(double) * n * (NsA + 1));
  double* qcA = (double*)malloc(sizeof(double) * n * (NsA + 1));
  double* qB = (double*)malloc(sizeof(double) * n * (NsB + 1));
  double* qcB = (double*)malloc(sizeof(double) * n * (NsB + 1));
  double cell[6] = {Lx, Ly, Lz, 0, 0, 0};
  orc_set_cell_para(s, cell);
  orc_set_contour_step(s, ds);
  orc_diffusion(s); /* :471 warm-up on whatever iomat holds */
  for (size_t itr = 0; itr < Nstep; ++itr) {
    for (size_t i = 0; i < n; ++i) ksi[i] = 0.5 * (wA[i] + wB[i] - xN); /* :341-346 */
    for (size_t i = 0; i < n; ++i) wAds[i] = 1. * exp(-ds / 2 * wA[i]);  /* :315-322 */
    for (size_t i = 0; i < n; ++i) wBds[i] = 1. * exp(-ds / 2 * wB[i]);
    for (size_t i = 0; i < n; ++i) init[i] = 1;
    orc_propagate(s, qA, init, wAds, NsA, 1);
    orc_propagate(s, qcB, init, wBds, NsB, -1);
    memcpy(init, qA + NsA * n, sizeof(double) * n);
    orc_propagate(s, qB, init, wBds, NsB, 1);
    memcpy(init, qcB, sizeof(double) * n);
    orc_propagate(s, qcA, init, wAds, NsA, -1);
    double sum = 0.;
    for (size_t i = 0; i < n; ++i) sum += qcA[i];
    double Q = sum / n; /* :324-331, :494 */
    for (size_t i = 0; i < n; ++i) phiA[i] = 0;
    for (size_t i = 0; i < n; ++i) phiB[i] = 0;
    integrate_phi(phiA, qA, qcA, fA, Q, n, NsA);
    integrate_phi(phiB, qB, qcB, fB, Q, n, NsB);
    /* print_energy (:398-424), recorded every iteration */
    double e0 = -log(Q), e1 = 0, enth = 0, fk = 0;
    double maxinc = fabs(1 - phiA[0] - phiB[0]);
    for (size_t i = 0; i < n; ++i) {
      e1 += -wA[i] * phiA[i] - wB[i] * phiB[i];
      enth += xN * phiA[i] * phiB[i];
      double t = 1 - phiA[i] - phiB[i];
      fk += -ksi[i] * t;
      maxinc = fabs(t) > maxinc ? fabs(t) : maxinc;
    }
    e1 /= n;
    enth /= n;
    fk /= n;
    if (trace) {
      double* r = trace + itr * ORC_TRACE_COLS;
      r[0] = e0 + e1 + enth + fk;
      r[1] = enth;
      r[2] = e0;
      r[3] = e1;
      r[4] = fk;
      r[5] = maxinc;
      r[6] = Q;
    }
    /* calc_new_fields (:362-371) */
    for (size_t i = 0; i < n; ++i) {
      wA[i] += acceptance * (xN * phiB[i] + ksi[i] - wA[i]);
      wB[i] += acceptance * (xN * phiA[i] + ksi[i] - wB[i]);
    }
  }
  free(wAds);
  free(wBds);
  free(ksi);
  free(init);
  free(qA);
  free(qcA);
  free(qB);
  free(qcB);
  return 0;
}
