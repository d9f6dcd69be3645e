/*
 * CPU oracle for the dense FP64 GEMM hot path. TEST INFRASTRUCTURE ONLY: see
 * oracle.h. Built with -ffp-contract=off so that every `a * b + c` written here is
 * rounded twice unless fma() is called explicitly.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GAMMA 0x9E3779B97F4A7C15ULL

/* ---------------- generators ---------------- */

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * GAMMA); }

void orc_fill_uniform(double *x, size_t n, uint64_t seed) {
  for (size_t i = 0; i < n; ++i) {
    uint64_t z = orc_splitmix64_at(seed, i);
    x[i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
  }
}

void orc_fill_wide(double *x, size_t n, uint64_t seed) {
  for (size_t i = 0; i < n; ++i) {
    uint64_t z = orc_splitmix64_at(seed, i);
    int e = (int)(((z >> 32) * 41ULL) >> 32) - 20;
    x[i] = ldexp((z & 1) ? -1.0 : 1.0, e);
  }
}

void orc_fill_ramp(double *x, size_t n) {
  for (size_t i = 0; i < n; ++i) x[i] = (double)i;
}

/* ---------------- reference arithmetic ---------------- */

/* src/iterative.c:8-13 */
void orc_iterative(const double *A, const double *B, double *C, int n) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double a = A[(size_t)i * n + j];
      const double *brow = B + (size_t)j * n;
      double *crow = C + (size_t)i * n;
      for (int k = 0; k < n; ++k) {
        double p = a * brow[k];
        crow[k] = crow[k] + p;
      }
    }
}

/* src/phpc_gemm.cu:25-55 (per-element arithmetic of gemm_kernel) */
void orc_kernel_emul(const double *A, int lda, const double *B, int ldb, double *C, int ldc,
                     int m, int k, int n) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m; ++i) {
    double *acc = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    const double *arow = A + (size_t)i * lda;
    /* k outer so that the row of B streams; each acc[j] still sees ascending kk. */
    for (int kk = 0; kk < k; ++kk) {
      const double a = arow[kk];
      const double *brow = B + (size_t)kk * ldb;
      for (int j = 0; j < n; ++j) acc[j] = fma(a, brow[j], acc[j]);
    }
    double *crow = C + (size_t)i * ldc;
    for (int j = 0; j < n; ++j) crow[j] = crow[j] + acc[j];
    free(acc);
  }
}

/* src/phpc_summa.c:9-22 */
int orc_find_lcm(int a, int b) {
  int x = a, y = b;
  while (y != 0) {
    int r = x % y;
    x = y;
    y = r;
  }
  return a / x * b;
}

/* MPI_Dims_create(P, 2), src/main.c:38-45 */
void orc_dims_create(int nranks, int dims[2]) {
  if (nranks <= 1) {
    dims[0] = dims[1] = 1;
    return;
  }
  int best = 1;
  for (int d = 1; (long)d * d <= nranks; ++d)
    if (nranks % d == 0) best = d;
  dims[0] = nranks / best;
  dims[1] = best;
}

int orc_block_size(int n, int parts, int i) { return n / parts + (i < n % parts); }

int orc_block_start(int n, int parts, int i) {
  int q = n / parts, r = n % parts;
  return i * q + (i < r ? i : r);
}

/* src/phpc_summa.c:36-95 on an emulated p x q grid */
void orc_summa_emul(const double *A, const double *B, double *C, int N, int p, int q, int panel) {
  int w = panel > 0 ? panel : N / orc_find_lcm(p, q);
  if (w <= 0) return;
  int steps = (N + w - 1) / w;
  for (int t = 0; t < steps; ++t) {
    int k0 = t * w;
    int kw = (k0 + w <= N) ? w : N - k0;
    for (int ri = 0; ri < p; ++ri)
      for (int cj = 0; cj < q; ++cj) {
        int r0 = orc_block_start(N, p, ri), mr = orc_block_size(N, p, ri);
        int c0 = orc_block_start(N, q, cj), nc = orc_block_size(N, q, cj);
        orc_kernel_emul(A + (size_t)r0 * N + k0, N, B + (size_t)k0 * N + c0, N,
                        C + (size_t)r0 * N + c0, N, mr, kw, nc);
      }
  }
}

/*
 * The reference's SUMMA arithmetic (src/phpc_summa.c:63-95 over src/phpc_gemm.cu:32,50-55)
 * for a list of entries of C = C0 + A B, k x k panels of width `panel` (<= 0: one panel,
 * which is one local-GEMM call over the whole k): out[e] = ((c0 + s_0) + s_1) + ... with
 * s_t = fma chain from zero over panel t in ascending k. Every rank of a SUMMA grid walks
 * the same global k panels, so the value of an entry does not depend on the grid. This is
 * orc_summa_emul / orc_kernel_emul evaluated only where asked, for sizes where the full
 * emulation would take too long; C0 may be NULL (zero).
 */
void orc_summa_entries(const double *A, long lda, const double *B, long ldb, const double *C0, long ldc0,
                       const int *rows, const int *cols, long count, int k, int panel, double *out) {
  const int w = panel > 0 ? panel : (k > 0 ? k : 1);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < count; ++e) {
    const double *arow = A + (size_t)rows[e] * lda;
    const double *bcol = B + cols[e];
    double c = C0 ? C0[(size_t)rows[e] * ldc0 + cols[e]] : 0.0;
    for (int k0 = 0; k0 < k; k0 += w) {
      const int k1 = k0 + w < k ? k0 + w : k;
      double acc = 0.0;
      for (int kk = k0; kk < k1; ++kk) acc = fma(arow[kk], bcol[(size_t)kk * ldb], acc);
      c = c + acc;
    }
    out[e] = c;
  }
}

/* ---------------- double-double reference ---------------- */

static inline void two_sum(double a, double b, double *s, double *e) {
  double x = a + b;
  double z = x - a;
  *e = (a - (x - z)) + (b - z);
  *s = x;
}

static inline void dd_add_prod(double *hi, double *lo, double a, double b) {
  double p = a * b;
  double pe = fma(a, b, -p);
  double s, se;
  two_sum(*hi, p, &s, &se);
  *hi = s;
  *lo += se + pe;
}

static double dd_dot(const double *a, size_t inca, const double *b, size_t incb, int k, double c0) {
  double hi = c0, lo = 0.0;
  for (int kk = 0; kk < k; ++kk) dd_add_prod(&hi, &lo, a[(size_t)kk * inca], b[(size_t)kk * incb]);
  return hi + lo;
}

void orc_gemm_dd(const double *A, int lda, const double *B, int ldb, const double *C0, int ldc0,
                 double *Cref, int ldcr, int m, int k, int n) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < m; ++i) {
    double *hi = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *lo = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int j = 0; j < n; ++j) hi[j] = C0 ? C0[(size_t)i * ldc0 + j] : 0.0;
    const double *arow = A + (size_t)i * lda;
    for (int kk = 0; kk < k; ++kk) {
      const double a = arow[kk];
      const double *brow = B + (size_t)kk * ldb;
      for (int j = 0; j < n; ++j) dd_add_prod(&hi[j], &lo[j], a, brow[j]);
    }
    for (int j = 0; j < n; ++j) Cref[(size_t)i * ldcr + j] = hi[j] + lo[j];
    free(hi);
    free(lo);
  }
}

void orc_gemm_dd_entries(const double *A, int lda, const double *B, int ldb, const double *C0,
                         int ldc0, const int *rows, const int *cols, double *out, size_t count,
                         int k) {
#pragma omp parallel for schedule(dynamic, 16)
  for (long e = 0; e < (long)count; ++e) {
    int i = rows[e], j = cols[e];
    double c0 = C0 ? C0[(size_t)i * ldc0 + j] : 0.0;
    out[e] = dd_dot(A + (size_t)i * lda, 1, B + j, (size_t)ldb, k, c0);
  }
}

double orc_dot_q(const double *a, int inca, const double *b, int incb, int k) {
  __float128 s = 0;
  for (int kk = 0; kk < k; ++kk) s += (__float128)a[(size_t)kk * inca] * (__float128)b[(size_t)kk * incb];
  return (double)s;
}

double orc_absdot(const double *A, int lda, const double *B, int ldb, int i, int j, int k) {
  double s = 0;
  for (int kk = 0; kk < k; ++kk) s += fabs(A[(size_t)i * lda + kk]) * fabs(B[(size_t)kk * ldb + j]);
  return s;
}

void orc_err_full(const double *C, int ldc, const double *Cref, int ldcr, const double *A, int lda,
                  const double *B, int ldb, const double *C0, int ldc0, int m, int k, int n,
                  double *err_cw, double *err_nw) {
  double cw = 0, maxd = 0, maxr = 0;
#pragma omp parallel for reduction(max : cw, maxd, maxr) schedule(dynamic, 4)
  for (int i = 0; i < m; ++i) {
    double *ab = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int kk = 0; kk < k; ++kk) {
      double a = fabs(A[(size_t)i * lda + kk]);
      const double *brow = B + (size_t)kk * ldb;
      for (int j = 0; j < n; ++j) ab[j] += a * fabs(brow[j]);
    }
    for (int j = 0; j < n; ++j) {
      double d = fabs(C[(size_t)i * ldc + j] - Cref[(size_t)i * ldcr + j]);
      double den = ab[j] + (C0 ? fabs(C0[(size_t)i * ldc0 + j]) : 0.0);
      double r = d == 0.0 ? 0.0 : (den > 0 ? d / den : INFINITY);
      if (r > cw) cw = r;
      if (d > maxd) maxd = d;
      double ar = fabs(Cref[(size_t)i * ldcr + j]);
      if (ar > maxr) maxr = ar;
    }
    free(ab);
  }
  if (err_cw) *err_cw = cw;
  if (err_nw) *err_nw = maxr > 0 ? maxd / maxr : maxd;
}

/* ---------------- closed forms and residuals ---------------- */

double orc_ramp_closed(int N, int r, int c) {
  typedef unsigned __int128 u128;
  u128 n = (u128)N;
  u128 s1 = n * (n - 1) / 2;
  u128 s2 = (n - 1) * n * (2 * n - 1) / 6;
  u128 v = (u128)r * n * n * s1 + (u128)r * (u128)c * n * n + n * s2 + (u128)c * s1;
  return (double)v;
}

double orc_freivalds(const double *A, const double *B, const double *C, int n, uint64_t seed) {
  double *x = (double *)malloc(sizeof(double) * (size_t)n);
  double *bx = (double *)malloc(sizeof(double) * (size_t)n);
  double *absbx = (double *)malloc(sizeof(double) * (size_t)n);
  orc_fill_uniform(x, (size_t)n, seed);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    bx[i] = dd_dot(B + (size_t)i * n, 1, x, 1, n, 0.0);
    double s = 0;
    for (int j = 0; j < n; ++j) s += fabs(B[(size_t)i * n + j]) * fabs(x[j]);
    absbx[i] = s;
  }
  double worst = 0;
#pragma omp parallel for reduction(max : worst) schedule(static)
  for (int i = 0; i < n; ++i) {
    double cx = dd_dot(C + (size_t)i * n, 1, x, 1, n, 0.0);
    double abx = dd_dot(A + (size_t)i * n, 1, bx, 1, n, 0.0);
    double den = 0;
    for (int j = 0; j < n; ++j) den += fabs(A[(size_t)i * n + j]) * absbx[j];
    double r = den > 0 ? fabs(cx - abx) / den : fabs(cx - abx);
    if (r > worst) worst = r;
  }
  free(x);
  free(bx);
  free(absbx);
  return worst;
}

/* ---------------- split-precision restatement ---------------- */

int orc_exp_of_max(double mx) {
  int e = 0;
  if (mx > 0.0) (void)frexp(mx, &e);
  return e;
}

void orc_split_digits_w(double x, int e, int s, int bits, int *d) {
  double v = ldexp(x, -e);
  for (int t = 0; t < s; ++t) {
    v = ldexp(v, bits);
    double q = trunc(v);
    v -= q;
    d[t] = (int)q;
  }
}

void orc_split_digits(double x, int e, int s, int *d) { orc_split_digits_w(x, e, s, 7, d); }

void orc_split_gemm(const double *A, int lda, const double *B, int ldb, double *C, int ldc, int m, int k, int n,
                    int s) {
  orc_split_gemm_w(A, lda, B, ldb, C, ldc, m, k, n, s, 7, 131072);
}

void orc_split_gemm_w(const double *A, int lda, const double *B, int ldb, double *C, int ldc, int m, int k, int n,
                      int s, int bits, long max_terms) {
  if (m <= 0 || n <= 0 || k <= 0 || s <= 0) return;
  const long kp = ((long)k + 127) / 128 * 128; /* the device's plane width */
  int *eA = (int *)malloc(sizeof(int) * (size_t)m);
  int *eB = (int *)malloc(sizeof(int) * (size_t)n);
  signed char *a8 = (signed char *)calloc((size_t)m * (size_t)s * (size_t)k, 1); /* [i][t][l] */
  signed char *b8 = (signed char *)calloc((size_t)n * (size_t)s * (size_t)k, 1); /* [j][u][l] */
  for (int i = 0; i < m; ++i) {
    double mx = 0;
    for (int l = 0; l < k; ++l) mx = fmax(mx, fabs(A[(size_t)i * lda + l]));
    eA[i] = orc_exp_of_max(mx);
  }
  for (int j = 0; j < n; ++j) {
    double mx = 0;
    for (int l = 0; l < k; ++l) mx = fmax(mx, fabs(B[(size_t)l * ldb + j]));
    eB[j] = orc_exp_of_max(mx);
  }
  int d[64];
  for (int i = 0; i < m; ++i)
    for (int l = 0; l < k; ++l) {
      orc_split_digits_w(A[(size_t)i * lda + l], eA[i], s, bits, d);
      for (int t = 0; t < s; ++t) a8[((size_t)i * s + t) * k + l] = (signed char)d[t];
    }
  for (int j = 0; j < n; ++j)
    for (int l = 0; l < k; ++l) {
      orc_split_digits_w(B[(size_t)l * ldb + j], eB[j], s, bits, d);
      for (int u = 0; u < s; ++u) b8[((size_t)j * s + u) * k + l] = (signed char)d[u];
    }
  /* device piece length: whole padded slices per piece, at most max_terms products, else
   * whole 128-byte slabs */
  long piece = max_terms / kp * kp;
  if (piece <= 0) piece = (max_terms / 128) * 128;
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      double acc = 0.0;
      int first = 1;
      for (int g = s + 1; g >= 2; --g) {
        /* level g: A slices t = 1..g-1 against B slices u = g-t, concatenated along k
         * with stride kp (zero padding between slices), cut into pieces of `piece` */
        const long klen = (long)(g - 1) * kp;
        for (long off = 0; off < klen; off += piece) {
          const long end = off + piece < klen ? off + piece : klen;
          long long sum = 0;
          for (long q = off; q < end; ++q) {
            const int t = (int)(q / kp) + 1;
            const long l = q % kp;
            if (l >= k) continue;
            const int u = g - t;
            sum += (long long)a8[((size_t)i * s + (t - 1)) * k + l] * (long long)b8[((size_t)j * s + (u - 1)) * k + l];
          }
          double v = ldexp((double)sum, -bits * g);
          acc = first ? v : acc + v;
          first = 0;
        }
      }
      C[(size_t)i * ldc + j] = C[(size_t)i * ldc + j] + ldexp(acc, eA[i] + eB[j]);
    }
  free(eA);
  free(eB);
  free(a8);
  free(b8);
}

/* ---------------- CRT restatement ---------------- */

static const int orc_moduli[20] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                   223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

static int orc_smod(long long x, int m) {
  long long r = x % m;
  if (r < 0) r += m;
  if (r > (m - 1) / 2) r -= m;
  return (int)r;
}

static int orc_modinv(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((long long)a * x % m == 1) return x;
  return 0;
}

void orc_crt_digits(const int *res, int L, int *v) {
  for (int t = 0; t < L; ++t) {
    const int mt = orc_moduli[t];
    long long prod = 1, y = 0;
    for (int s2 = 0; s2 < t; ++s2) {
      y += (long long)v[s2] * prod; /* prod = (m_1..m_{s2}) mod m_t */
      prod = prod * orc_moduli[s2] % mt;
    }
    const int inv = t == 0 ? 1 : orc_modinv((int)prod, mt);
    long long d = (res[t] - y) % mt;
    if (d < 0) d += mt;
    v[t] = orc_smod(d * inv % mt, mt);
  }
}

static int orc_res_of_double(double x, int m) {
  /* x is an integer-valued double, |x| < 2^127 */
  __int128 xi = (__int128)x;
  __int128 r = xi % m;
  if (r < 0) r += m;
  int ri = (int)r;
  if (ri > (m - 1) / 2) ri -= m;
  return ri;
}

void orc_crt_gemm(const double *A, int lda, const double *B, int ldb, double *C, int ldc, int m, int k, int n,
                  int p, int L) {
  if (m <= 0 || n <= 0 || k <= 0 || L <= 0) return;
  int *eA = (int *)malloc(sizeof(int) * (size_t)m);
  int *eB = (int *)malloc(sizeof(int) * (size_t)n);
  signed char *ra = (signed char *)malloc((size_t)m * L * k);
  signed char *rb = (signed char *)malloc((size_t)n * L * k);
  for (int i = 0; i < m; ++i) {
    double mx = 0;
    for (int l = 0; l < k; ++l) mx = fmax(mx, fabs(A[(size_t)i * lda + l]));
    eA[i] = orc_exp_of_max(mx);
    for (int l = 0; l < k; ++l) {
      const double x = rint(ldexp(A[(size_t)i * lda + l], p - eA[i]));
      for (int t = 0; t < L; ++t) {
        int r = orc_res_of_double(x, orc_moduli[t]);
        if (orc_moduli[t] == 256 && r == 128) r = -128;
        ra[((size_t)i * L + t) * k + l] = (signed char)r;
      }
    }
  }
  for (int j = 0; j < n; ++j) {
    double mx = 0;
    for (int l = 0; l < k; ++l) mx = fmax(mx, fabs(B[(size_t)l * ldb + j]));
    eB[j] = orc_exp_of_max(mx);
    for (int l = 0; l < k; ++l) {
      const double x = rint(ldexp(B[(size_t)l * ldb + j], p - eB[j]));
      for (int t = 0; t < L; ++t) {
        int r = orc_res_of_double(x, orc_moduli[t]);
        if (orc_moduli[t] == 256 && r == 128) r = -128;
        rb[((size_t)j * L + t) * k + l] = (signed char)r;
      }
    }
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      int res[20], v[20];
      for (int t = 0; t < L; ++t) {
        long long sum = 0;
        for (int l = 0; l < k; ++l)
          sum += (long long)ra[((size_t)i * L + t) * k + l] * (long long)rb[((size_t)j * L + t) * k + l];
        res[t] = orc_smod(sum, orc_moduli[t]);
      }
      orc_crt_digits(res, L, v);
      double hi = (double)v[L - 1], lo = 0.0;
      for (int t = L - 2; t >= 0; --t) {
        const double mt = (double)orc_moduli[t];
        const double ph = hi * mt;
        const double pe = fma(hi, mt, -ph);
        const double pl = lo * mt + pe;
        const double sh = ph + (double)v[t];
        const double bb = sh - ph;
        const double se = (ph - (sh - bb)) + ((double)v[t] - bb);
        const double tl = pl + se;
        hi = sh + tl;
        lo = tl - (hi - sh);
      }
      const double cp = hi + lo;
      C[(size_t)i * ldc + j] = C[(size_t)i * ldc + j] + ldexp(cp, eA[i] + eB[j] - 2 * p);
    }
  free(eA);
  free(eB);
  free(ra);
  free(rb);
}
