/*
 * pulsar_oracle.c — FP64 CPU restatement of the reference GRAPE inner loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Structure follows the reference function by function (citations relative to
 * the Pulsar.jl tree): Hamiltonian assembly, eigen-exp propagators, forward and
 * backward chains stored whole, the per-slice gradient contraction, the MR
 * state-pair and Lindblad kernels, and the ensemble mean. All arithmetic is
 * FP64 complex, column-major.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "pulsar_oracle.h"

typedef double complex cx;

static inline cx ld(const double* p, size_t i) { return p[2 * i] + I * p[2 * i + 1]; }
static inline void st(double* p, size_t i, cx v) {
  p[2 * i] = creal(v);
  p[2 * i + 1] = cimag(v);
}

static void mat_load(int n, const double* src, cx* dst) {
  for (size_t i = 0; i < (size_t)n * n; i++) dst[i] = ld(src, i);
}
static void mat_store(int n, const cx* src, double* dst) {
  for (size_t i = 0; i < (size_t)n * n; i++) st(dst, i, src[i]);
}
static void mat_eye(int n, cx* A) {
  memset(A, 0, sizeof(cx) * n * n);
  for (int i = 0; i < n; i++) A[i + i * n] = 1.0;
}

/* Products of n >= 32 outside a parallel region share their columns among
 * tl_columns threads: or_eval sets it for the calling thread when a problem
 * has fewer members than threads (a member evaluated alone, its chains); every
 * column is computed in the same order either way. */
static _Thread_local int tl_columns = 1;
#ifdef _OPENMP
#define PAR_COLUMNS \
  _Pragma("omp parallel for schedule(static) num_threads(tl_columns) if (tl_columns > 1 && n >= 32 && !omp_in_parallel())")
#else
#define PAR_COLUMNS
#endif

/* C = A * B (column-major n x n) */
static void mm_nn(int n, const cx* A, const cx* B, cx* C) {
  PAR_COLUMNS
  for (int j = 0; j < n; j++) {
    cx* c = C + (size_t)j * n;
    for (int i = 0; i < n; i++) c[i] = 0;
    for (int k = 0; k < n; k++) {
      cx b = B[k + (size_t)j * n];
      const cx* a = A + (size_t)k * n;
      for (int i = 0; i < n; i++) c[i] += a[i] * b;
    }
  }
}
/* C = A * B^dag */
static void mm_nc(int n, const cx* A, const cx* B, cx* C) {
  PAR_COLUMNS
  for (int j = 0; j < n; j++) {
    cx* c = C + (size_t)j * n;
    for (int i = 0; i < n; i++) c[i] = 0;
    for (int k = 0; k < n; k++) {
      cx b = conj(B[j + (size_t)k * n]);
      const cx* a = A + (size_t)k * n;
      for (int i = 0; i < n; i++) c[i] += a[i] * b;
    }
  }
}
/* C = A^dag * B */
static void mm_cn(int n, const cx* A, const cx* B, cx* C) {
  PAR_COLUMNS
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      cx s = 0;
      for (int k = 0; k < n; k++) s += conj(A[k + (size_t)i * n]) * B[k + (size_t)j * n];
      C[i + (size_t)j * n] = s;
    }
}
/* y = A x */
static void mv(int n, const cx* A, const cx* x, cx* y) {
  for (int i = 0; i < n; i++) y[i] = 0;
  for (int k = 0; k < n; k++) {
    cx b = x[k];
    for (int i = 0; i < n; i++) y[i] += A[i + (size_t)k * n] * b;
  }
}
/* y = A^dag x */
static void mv_c(int n, const cx* A, const cx* x, cx* y) {
  for (int i = 0; i < n; i++) {
    cx s = 0;
    for (int k = 0; k < n; k++) s += conj(A[k + (size_t)i * n]) * x[k];
    y[i] = s;
  }
}
/* dot(a, b) = sum conj(a) b  (Julia LinearAlgebra.dot) */
static cx vdot(int n, const cx* a, const cx* b) {
  cx s = 0;
  for (int i = 0; i < n; i++) s += conj(a[i]) * b[i];
  return s;
}
static double norm1(int n, const cx* A) {
  double best = 0;
  for (int j = 0; j < n; j++) {
    double s = 0;
    for (int i = 0; i < n; i++) s += cabs(A[i + (size_t)j * n]);
    if (s > best) best = s;
  }
  return best;
}

/* ------------------------------------------------------------------------ */
/* Hermitian eigensolver: cyclic complex Jacobi (Golub & Van Loan 8.5). The   */
/* upper triangle is read, as Julia's Hermitian(H) does (Propagators.jl:80).  */
/* ------------------------------------------------------------------------ */
static int eigh_cx(int d, const cx* Ain, double* w, cx* V) {
  cx* A = (cx*)malloc(sizeof(cx) * d * d);
  if (!A) return PR_ERR_NOMEM;
  for (int j = 0; j < d; j++)
    for (int i = 0; i <= j; i++) {
      cx v = Ain[i + (size_t)j * d];
      if (i == j) v = creal(v);
      A[i + (size_t)j * d] = v;
      A[j + (size_t)i * d] = conj(v);
    }
  mat_eye(d, V);
  double total = 0;
  for (size_t i = 0; i < (size_t)d * d; i++) total += creal(A[i] * conj(A[i]));
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0;
    for (int q = 1; q < d; q++)
      for (int p = 0; p < q; p++) off += creal(A[p + (size_t)q * d] * conj(A[p + (size_t)q * d]));
    if (off <= 1e-34 * total || off == 0.0) break;
    for (int q = 1; q < d; q++)
      for (int p = 0; p < q; p++) {
        cx z = A[p + (size_t)q * d];
        double az = cabs(z);
        if (az == 0.0) continue;
        double a = creal(A[p + (size_t)p * d]), b = creal(A[q + (size_t)q * d]);
        if (sweep > 3 && az < 1e-18 * (fabs(a) + fabs(b))) {
          A[p + (size_t)q * d] = 0;
          A[q + (size_t)p * d] = 0;
          continue;
        }
        cx ph = z / az; /* e^{i phi} */
        double theta = (b - a) / (2.0 * az);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        cx cph = conj(ph);
        /* J = E R: J_pp = c, J_pq = s, J_qp = -s e^{-i phi}, J_qq = c e^{-i phi} */
        cx Jpp = c, Jpq = s, Jqp = -s * cph, Jqq = c * cph;
        for (int i = 0; i < d; i++) { /* A <- A J (columns p, q) */
          cx aip = A[i + (size_t)p * d], aiq = A[i + (size_t)q * d];
          A[i + (size_t)p * d] = aip * Jpp + aiq * Jqp;
          A[i + (size_t)q * d] = aip * Jpq + aiq * Jqq;
          cx vip = V[i + (size_t)p * d], viq = V[i + (size_t)q * d];
          V[i + (size_t)p * d] = vip * Jpp + viq * Jqp;
          V[i + (size_t)q * d] = vip * Jpq + viq * Jqq;
        }
        for (int j = 0; j < d; j++) { /* A <- J^dag A (rows p, q) */
          cx apj = A[p + (size_t)j * d], aqj = A[q + (size_t)j * d];
          A[p + (size_t)j * d] = conj(Jpp) * apj + conj(Jqp) * aqj;
          A[q + (size_t)j * d] = conj(Jpq) * apj + conj(Jqq) * aqj;
        }
        A[p + (size_t)q * d] = 0;
        A[q + (size_t)p * d] = 0;
        A[p + (size_t)p * d] = a - t * az;
        A[q + (size_t)q * d] = b + t * az;
      }
  }
  for (int i = 0; i < d; i++) w[i] = creal(A[i + (size_t)i * d]);
  free(A);
  return PR_OK;
}

int or_eigh(int d, const double* A, double* w, double* V) {
  cx* a = (cx*)malloc(sizeof(cx) * d * d);
  cx* v = (cx*)malloc(sizeof(cx) * d * d);
  if (!a || !v) {
    free(a);
    free(v);
    return PR_ERR_NOMEM;
  }
  mat_load(d, A, a);
  int rc = eigh_cx(d, a, w, v);
  mat_store(d, v, V);
  free(a);
  free(v);
  return rc;
}

/* _expm_neg_i_into! (src/Computation/Propagators.jl:659-674):
 * U = (V diag(cis(-lambda dt))) V^dag. */
static int expm_neg_i_cx(int d, const cx* H, double dt, cx* U, cx* V, cx* T, double* w) {
  if (dt == 0.0) { /* compute_propagator short-circuit, Propagators.jl:72-74 */
    mat_eye(d, U);
    return PR_OK;
  }
  int rc = eigh_cx(d, H, w, V);
  if (rc) return rc;
  for (int j = 0; j < d; j++) {
    cx c = cexp(-I * w[j] * dt);
    for (int i = 0; i < d; i++) T[i + (size_t)j * d] = V[i + (size_t)j * d] * c;
  }
  mm_nc(d, T, V, U);
  return PR_OK;
}

int or_expm_neg_i(int d, const double* H, double dt, double* U) {
  cx* buf = (cx*)malloc(sizeof(cx) * 4 * d * d);
  double* w = (double*)malloc(sizeof(double) * d);
  if (!buf || !w) {
    free(buf);
    free(w);
    return PR_ERR_NOMEM;
  }
  cx *h = buf, *u = buf + d * d, *v = buf + 2 * d * d, *t = buf + 3 * d * d;
  mat_load(d, H, h);
  int rc = expm_neg_i_cx(d, h, dt, u, v, t, w);
  mat_store(d, u, U);
  free(buf);
  free(w);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* General exp: scaling and squaring with Pade approximants (Higham 2005,     */
/* "The scaling and squaring method for the matrix exponential revisited"),  */
/* the algorithm of Julia's LinearAlgebra.exp used by GRAPELindblad.jl:152 and */
/* PropagatorRegistry.jl:148-163 (without the balancing step).                */
/* ------------------------------------------------------------------------ */
static int lu_solve(int n, cx* A, cx* B) { /* solve A X = B in place (B <- X), n RHS */
  int* piv = (int*)malloc(sizeof(int) * n);
  if (!piv) return PR_ERR_NOMEM;
  for (int k = 0; k < n; k++) {
    int pmax = k;
    double vmax = cabs(A[k + (size_t)k * n]);
    for (int i = k + 1; i < n; i++)
      if (cabs(A[i + (size_t)k * n]) > vmax) {
        vmax = cabs(A[i + (size_t)k * n]);
        pmax = i;
      }
    piv[k] = pmax;
    if (vmax == 0.0) {
      free(piv);
      return PR_ERR_ARG;
    }
    if (pmax != k)
      for (int j = 0; j < n; j++) {
        cx t = A[k + (size_t)j * n];
        A[k + (size_t)j * n] = A[pmax + (size_t)j * n];
        A[pmax + (size_t)j * n] = t;
        t = B[k + (size_t)j * n];
        B[k + (size_t)j * n] = B[pmax + (size_t)j * n];
        B[pmax + (size_t)j * n] = t;
      }
    cx inv = 1.0 / A[k + (size_t)k * n];
    for (int i = k + 1; i < n; i++) {
      cx f = A[i + (size_t)k * n] * inv;
      A[i + (size_t)k * n] = f;
      for (int j = k + 1; j < n; j++) A[i + (size_t)j * n] -= f * A[k + (size_t)j * n];
      for (int j = 0; j < n; j++) B[i + (size_t)j * n] -= f * B[k + (size_t)j * n];
    }
  }
  for (int j = 0; j < n; j++)
    for (int i = n - 1; i >= 0; i--) {
      cx s = B[i + (size_t)j * n];
      for (int k = i + 1; k < n; k++) s -= A[i + (size_t)k * n] * B[k + (size_t)j * n];
      B[i + (size_t)j * n] = s / A[i + (size_t)i * n];
    }
  free(piv);
  return PR_OK;
}

static int expm_cx(int n, const cx* Ain, cx* E) {
  static const double th[5] = {1.495585217958292e-2, 2.539398330063230e-1, 9.504178996162932e-1,
                               2.097847961257068e0, 5.371920351148152e0};
  static const double b3[4] = {120., 60., 12., 1.};
  static const double b5[6] = {30240., 15120., 3360., 420., 30., 1.};
  static const double b7[8] = {17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.};
  static const double b9[10] = {17643225600., 8821612800., 2075673600., 302702400., 30270240.,
                                2162160.,     110880.,     3960.,       90.,        1.};
  static const double b13[14] = {64764752532480000., 32382376266240000., 7771770303897600.,
                                 1187353796428800.,  129060195264000.,   10559470521600.,
                                 670442572800.,      33522128640.,       1323241920.,
                                 40840800.,          960960.,            16380.,
                                 182.,               1.};
  size_t nn = (size_t)n * n;
  cx* w = (cx*)malloc(sizeof(cx) * nn * 9);
  if (!w) return PR_ERR_NOMEM;
  cx *A = w, *A2 = w + nn, *A4 = w + 2 * nn, *A6 = w + 3 * nn, *U = w + 4 * nn, *V = w + 5 * nn,
     *T = w + 6 * nn, *T2 = w + 7 * nn, *A8 = w + 8 * nn;
  memcpy(A, Ain, sizeof(cx) * nn);
  double nrm = norm1(n, A);
  int s = 0;
  const double* b = NULL;
  int m = 13;
  if (nrm <= th[0]) { b = b3; m = 3; }
  else if (nrm <= th[1]) { b = b5; m = 5; }
  else if (nrm <= th[2]) { b = b7; m = 7; }
  else if (nrm <= th[3]) { b = b9; m = 9; }
  else {
    b = b13;
    m = 13;
    if (nrm > th[4]) s = (int)ceil(log2(nrm / th[4]));
    if (s < 0) s = 0;
    double sc = ldexp(1.0, -s);
    for (size_t i = 0; i < nn; i++) A[i] *= sc;
  }
  mm_nn(n, A, A, A2);
  if (m <= 9) {
    /* U = A * sum_{odd k} b_k A^{k-1}, V = sum_{even k} b_k A^k */
    cx* P[5] = {NULL, A2, A4, A6, A8};
    int npow = (m - 1) / 2; /* highest even power index: A^{2 npow} */
    if (npow >= 2) mm_nn(n, A2, A2, A4);
    if (npow >= 3) mm_nn(n, A4, A2, A6);
    if (npow >= 4) mm_nn(n, A6, A2, A8);
    for (size_t i = 0; i < nn; i++) {
      T[i] = 0;
      V[i] = 0;
    }
    for (int i = 0; i < n; i++) {
      T[i + (size_t)i * n] = b[1];
      V[i + (size_t)i * n] = b[0];
    }
    for (int q = 1; q <= npow; q++)
      for (size_t i = 0; i < nn; i++) {
        T[i] += b[2 * q + 1] * P[q][i];
        V[i] += b[2 * q] * P[q][i];
      }
    mm_nn(n, A, T, U);
  } else {
    mm_nn(n, A2, A2, A4);
    mm_nn(n, A4, A2, A6);
    for (size_t i = 0; i < nn; i++) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
    mm_nn(n, A6, T, T2);
    for (size_t i = 0; i < nn; i++) T2[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
    for (int i = 0; i < n; i++) T2[i + (size_t)i * n] += b[1];
    mm_nn(n, A, T2, U);
    for (size_t i = 0; i < nn; i++) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
    mm_nn(n, A6, T, V);
    for (size_t i = 0; i < nn; i++) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
    for (int i = 0; i < n; i++) V[i + (size_t)i * n] += b[0];
  }
  /* (V - U) R = (V + U) */
  for (size_t i = 0; i < nn; i++) {
    T[i] = V[i] - U[i];
    E[i] = V[i] + U[i];
  }
  int rc = lu_solve(n, T, E);
  for (int k = 0; k < s && rc == PR_OK; k++) {
    mm_nn(n, E, E, T);
    memcpy(E, T, sizeof(cx) * nn);
  }
  free(w);
  return rc;
}

int or_expm(int n, const double* A, double* E) {
  cx* a = (cx*)malloc(sizeof(cx) * 2 * n * n);
  if (!a) return PR_ERR_NOMEM;
  mat_load(n, A, a);
  int rc = expm_cx(n, a, a + n * n);
  mat_store(n, a + n * n, E);
  free(a);
  return rc;
}

/* Frechet derivative of exp: upper-right block of exp([[X, E], [0, X]]). */
static int frechet_cx(int n, const cx* X, const cx* Ein, cx* L) {
  int m = 2 * n;
  cx* Z = (cx*)calloc((size_t)m * m * 2, sizeof(cx));
  if (!Z) return PR_ERR_NOMEM;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      Z[i + (size_t)j * m] = X[i + (size_t)j * n];
      Z[(i + n) + (size_t)(j + n) * m] = X[i + (size_t)j * n];
      Z[i + (size_t)(j + n) * m] = Ein[i + (size_t)j * n];
    }
  cx* EZ = Z + (size_t)m * m;
  int rc = expm_cx(m, Z, EZ);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) L[i + (size_t)j * n] = EZ[i + (size_t)(j + n) * m];
  free(Z);
  return rc;
}

/* Frechet derivative for a Hermitian generator X = -i dt H via eigen
 * divided differences: L = V (Gamma o (V^dag E V)) V^dag,
 * Gamma_ab = (e^{x_a} - e^{x_b}) / (x_a - x_b), x = -i dt lambda. */
static int frechet_herm_cx(int d, const cx* H, double dt, const cx* Ein, cx* L, cx* V, cx* T1,
                           cx* T2, double* w) {
  int rc = eigh_cx(d, H, w, V);
  if (rc) return rc;
  mm_cn(d, V, Ein, T1); /* V^dag E */
  mm_nn(d, T1, V, T2);  /* V^dag E V */
  for (int b = 0; b < d; b++)
    for (int a = 0; a < d; a++) {
      cx xa = -I * dt * w[a], xb = -I * dt * w[b];
      cx dl = xa - xb, g;
      if (cabs(dl) < 1e-4) {
        cx ser = 1.0 + dl / 2.0 + dl * dl / 6.0 + dl * dl * dl / 24.0 + dl * dl * dl * dl / 120.0;
        g = cexp(xb) * ser;
      } else {
        g = (cexp(xa) - cexp(xb)) / dl;
      }
      T2[a + (size_t)b * d] *= g;
    }
  mm_nn(d, V, T2, T1);
  mm_nc(d, T1, V, L);
  return PR_OK;
}

int or_expm_frechet(int n, const double* X, const double* E, double* L) {
  cx* a = (cx*)malloc(sizeof(cx) * 3 * n * n);
  if (!a) return PR_ERR_NOMEM;
  mat_load(n, X, a);
  mat_load(n, E, a + n * n);
  int rc = frechet_cx(n, a, a + n * n, a + 2 * n * n);
  mat_store(n, a + 2 * n * n, L);
  free(a);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Per-member evaluation                                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  cx *H, *Hs, *Us, *P, *Q, *T1, *T2, *T3, *V, *A0, *psi, *lam, *tmpv, *ps0, *pst;
  double* w;
  int inner; /* threads over the slices of member_props (1: none) */
} work_t;

static void work_free(work_t* W) {
  free(W->H);
  free(W->Hs);
  free(W->Us);
  free(W->P);
  free(W->Q);
  free(W->T1);
  free(W->T2);
  free(W->T3);
  free(W->V);
  free(W->A0);
  free(W->psi);
  free(W->lam);
  free(W->tmpv);
  free(W->ps0);
  free(W->pst);
  free(W->w);
}

static int work_alloc(work_t* W, const pr_problem* p) {
  size_t d = p->dim, dd = d * d, nt = p->n_t, np = p->n_pairs > 0 ? p->n_pairs : 1;
  memset(W, 0, sizeof(*W));
  W->inner = 1;
  W->H = (cx*)malloc(sizeof(cx) * dd);
  W->Hs = (cx*)malloc(sizeof(cx) * dd * p->n_ctrl);
  W->Us = (cx*)malloc(sizeof(cx) * dd * nt);
  W->P = (cx*)malloc(sizeof(cx) * dd * (nt + 1));
  W->Q = (cx*)malloc(sizeof(cx) * dd * (nt + 1));
  W->T1 = (cx*)malloc(sizeof(cx) * dd);
  W->T2 = (cx*)malloc(sizeof(cx) * dd);
  W->T3 = (cx*)malloc(sizeof(cx) * dd);
  W->V = (cx*)malloc(sizeof(cx) * dd);
  W->A0 = (cx*)malloc(sizeof(cx) * dd);
  W->psi = (cx*)malloc(sizeof(cx) * d * (nt + 1) * np);
  W->lam = (cx*)malloc(sizeof(cx) * d * (nt + 1) * np);
  W->tmpv = (cx*)malloc(sizeof(cx) * d);
  W->ps0 = (cx*)malloc(sizeof(cx) * d * np);
  W->pst = (cx*)malloc(sizeof(cx) * d * np);
  W->w = (double*)malloc(sizeof(double) * d);
  if (!W->H || !W->Hs || !W->Us || !W->P || !W->Q || !W->T1 || !W->T2 || !W->T3 || !W->V ||
      !W->A0 || !W->psi || !W->lam || !W->tmpv || !W->ps0 || !W->pst || !W->w) {
    work_free(W);
    return PR_ERR_NOMEM;
  }
  return PR_OK;
}

static inline double slice_dt(const pr_problem* p, int k) {
  return p->dt_slices ? p->dt_slices[k] : p->dt;
}
static inline double cscale(const pr_problem* p, int m, int j) {
  return p->ctrl_scale ? p->ctrl_scale[(size_t)m * p->n_ctrl + j] : 1.0;
}
static const double* member_H0(const pr_problem* p, int m) {
  size_t dd = (size_t)p->dim * p->dim;
  return p->H0_per_member ? p->H0 + 2 * dd * m : p->H0;
}

/* Step propagators of member m for every slice, stored in W->Us.
 * Closed: build_total_hamiltonian! (Propagators.jl:520-544) with the member's
 * controls pre-scaled as _scale_controls does (Platforms/Common.jl:64-69),
 * then _expm_neg_i_into! per slice (Propagators.jl:166-184, 659-674).
 * Lindblad: L_buf = L_drift + (pwr*w) L_ctrl, Phi = exp(L_buf .* dt)
 * (GRAPELindblad.jl:146-153). */
static int member_props(const pr_problem* p, int m, const double* u, work_t* W) {
  int d = p->dim, nc = p->n_ctrl;
  size_t dd = (size_t)d * d;
  const double* h0 = member_H0(p, m);
  for (int j = 0; j < nc; j++) {
    double s = cscale(p, m, j);
    for (size_t i = 0; i < dd; i++) W->Hs[j * dd + i] = s * ld(p->Hc, j * dd + i);
  }
  /* The slices are independent: with W->inner > 1 (fewer members than
   * threads) they are shared among threads, each with scratch of its own
   * (thread 0 uses W's). */
  int err = PR_OK;
#ifdef _OPENMP
#pragma omp parallel num_threads(W->inner) if (W->inner > 1)
#endif
  {
    cx *H = W->H, *V = W->V, *T1 = W->T1;
    double* w = W->w;
    int own = 0;
#ifdef _OPENMP
    if (omp_get_thread_num() > 0) {
      H = (cx*)malloc(sizeof(cx) * dd);
      V = (cx*)malloc(sizeof(cx) * dd);
      T1 = (cx*)malloc(sizeof(cx) * dd);
      w = (double*)malloc(sizeof(double) * d);
      own = 1;
    }
#endif
    int lerr = (!H || !V || !T1 || !w) ? PR_ERR_NOMEM : PR_OK;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int k = 0; k < p->n_t; k++) {
      if (lerr) continue;
      double dtk = slice_dt(p, k);
      cx* Uk = W->Us + k * dd;
      for (size_t i = 0; i < dd; i++) H[i] = ld(h0, i);
      int rc;
      if (p->dynamics == PR_CLOSED) {
        for (int j = 0; j < nc; j++) {
          double a = u[j + (size_t)k * nc];
          const cx* hs = W->Hs + j * dd;
          for (size_t i = 0; i < dd; i++) H[i] += a * hs[i];
        }
        rc = expm_neg_i_cx(d, H, dtk, Uk, V, T1, w);
      } else {
        for (int j = 0; j < nc; j++) {
          double coef = cscale(p, m, j) * u[j + (size_t)k * nc];
          for (size_t i = 0; i < dd; i++) H[i] += coef * ld(p->Hc, j * dd + i);
        }
        for (size_t i = 0; i < dd; i++) T1[i] = H[i] * dtk;
        rc = expm_cx(d, T1, Uk);
      }
      if (rc) lerr = rc;
    }
    if (lerr) {
#ifdef _OPENMP
#pragma omp critical(member_props_err)
#endif
      err = lerr;
    }
    if (own) {
      free(H);
      free(V);
      free(T1);
      free(w);
    }
  }
  return err;
}

/* Exact-mode generator of slice k: closed dynamics rebuild H_k into W->H
 * exactly as member_props assembles it; Lindblad dynamics build
 * X_k = L_k .* dt_k into W->A0. */
static void member_generator(const pr_problem* p, int m, const double* u, int k, work_t* W) {
  int d = p->dim, nc = p->n_ctrl;
  size_t dd = (size_t)d * d;
  const double* h0 = member_H0(p, m);
  double dtk = slice_dt(p, k);
  if (p->dynamics == PR_CLOSED) {
    for (size_t i = 0; i < dd; i++) W->H[i] = ld(h0, i);
    for (int j = 0; j < nc; j++) {
      double a = u[j + (size_t)k * nc];
      const cx* hs = W->Hs + j * dd;
      for (size_t i = 0; i < dd; i++) W->H[i] += a * hs[i];
    }
    return;
  }
  for (size_t i = 0; i < dd; i++) W->A0[i] = ld(h0, i);
  for (int j = 0; j < nc; j++) {
    double coef = cscale(p, m, j) * u[j + (size_t)k * nc];
    for (size_t i = 0; i < dd; i++) W->A0[i] += coef * ld(p->Hc, j * dd + i);
  }
  for (size_t i = 0; i < dd; i++) W->A0[i] *= dtk;
}

/* Gate objective, closed system. compute_grape_gradient
 * (src/Physics/Gradients.jl:80-127): forward P (Propagators.jl:277-301),
 * backward Q (Propagators.jl:358-382), then _gate_gradient_into!
 * (Gradients.jl:254-291): Phi = Tr(Ut^dag U_total)/d,
 * A_k = (P_k Ut^dag) Q_k^dag, s = sum_{q,p} A_k[p,q] H_j[q,p],
 * G[j,k] = 2 Re(conj(Phi) (-i dt s / d)). The reference's loop builds
 * Q[k] = U_k^dag Q[k+1] from propagators[k] (Propagators.jl:375-380), so
 * Q_k^dag = U_n ... U_k includes slice k itself (its docstring's
 * U[k+1]^dag ... U[n]^dag does not): here R_k = U_n ... U_{k+1} and
 * Q_k^dag = R_k U_k. F is compute_fidelity's |Tr(Ut^dag U)/d|^2
 * (Objectives.jl:239-246). Exact mode uses the Frechet derivative
 * L(X_k, P_k Ut^dag R_k), the exact adjoint. PR_GATE_PROPAGATORS is
 * _grape_gradient_from_propagators (MRPhysics.jl:312-336): the trace without
 * Ut^dag and divided by d once more. */
static int gate_member(const pr_problem* p, int m, const double* u, int grad_mode, double* Fm,
                       double* Gm, work_t* W) {
  int d = p->dim, nc = p->n_ctrl, nt = p->n_t;
  size_t dd = (size_t)d * d;
  cx* Ut = W->T3;
  mat_load(d, p->U_target, Ut);
  /* forward propagators: P[0] = I, P[k+1] = U_k P[k] */
  mat_eye(d, W->P);
  for (int k = 0; k < nt; k++) mm_nn(d, W->Us + k * dd, W->P + k * dd, W->P + (k + 1) * dd);
  const cx* Utot = W->P + (size_t)nt * dd;
  cx tr = 0;
  for (int i = 0; i < d; i++)
    for (int r = 0; r < d; r++) tr += conj(Ut[r + (size_t)i * d]) * Utot[r + (size_t)i * d];
  cx Phi = tr / (double)d;
  *Fm = creal(Phi * conj(Phi));
  if (!Gm) return PR_OK;
  /* R_k^dag (0-based slice k): R[nt-1] = I, R[k] = U_{k+1}^dag R[k+1] */
  mat_eye(d, W->Q + (size_t)(nt - 1) * dd);
  for (int k = nt - 2; k >= 0; k--) mm_cn(d, W->Us + (k + 1) * dd, W->Q + (k + 1) * dd, W->Q + k * dd);
  const int helper = p->objective == PR_GATE_PROPAGATORS;
  for (int k = 0; k < nt; k++) {
    double dtk = slice_dt(p, k);
    if (helper) {
      mm_nc(d, W->P + k * dd, W->Q + k * dd, W->T2);  /* P_k R_k */
    } else {
      mm_nc(d, W->P + k * dd, Ut, W->T1);      /* tmp = P_k Ut^dag */
      mm_nc(d, W->T1, W->Q + k * dd, W->T2);   /* tmp R_k */
    }
    cx* A = W->T2;
    if (grad_mode != PR_GRAD_EXACT) {
      mm_nn(d, W->T2, W->Us + k * dd, W->T1);  /* A_k = P_k Ut^dag R_k U_k */
      memcpy(W->T2, W->T1, sizeof(cx) * dd);
    }
    if (grad_mode == PR_GRAD_EXACT) {
      member_generator(p, m, u, k, W);
      int rc;
      if (p->dynamics == PR_CLOSED) {
        rc = frechet_herm_cx(d, W->H, dtk, W->T2, W->T1, W->V, W->Q + (size_t)nt * dd, W->A0, W->w);
      } else {
        rc = frechet_cx(d, W->A0, W->T2, W->T1);
      }
      if (rc) return rc;
      A = W->T1;
    }
    for (int j = 0; j < nc; j++) {
      const cx* Hj = W->Hs + j * dd;
      cx s = 0;
      for (int q = 0; q < d; q++)
        for (int pp = 0; pp < d; pp++) s += A[pp + (size_t)q * d] * Hj[q + (size_t)pp * d];
      cx inner = -I * dtk * s / (double)d;
      if (helper) inner /= (double)d;
      Gm[j + (size_t)k * nc] = 2.0 * creal(conj(Phi) * inner);
    }
  }
  return PR_OK;
}

/* State-pair objectives: _grape_cpu (src/Application/MR/GRAPEState.jl:186-231)
 * for closed dynamics, _grape_lindblad_cpu (GRAPELindblad.jl:146-196) for
 * Lindblad. psi forward, lambda backward (lambda[n+1] = target,
 * lambda[n] = U_n^dag lambda[n+1]); inner = <lambda[n+1]| A_j |psi[n]> with
 * the member's scaled operator; prefactor fidelity_grad_prefactor
 * (Objectives.jl:352-355): :square 2 dt Im(conj(z) inner), :real dt Im(inner)
 * (closed) or lindblad_grad_prefactor (Lindblad.jl:220-232): :square
 * 2 dt Re(conj(z) inner), :real dt Re(inner). F_m sums the pairs
 * (GRAPEState.jl:210-211). */
static int pairs_member(const pr_problem* p, int m, const double* u, int grad_mode, double* Fm,
                        double* Gm, work_t* W) {
  int d = p->dim, nc = p->n_ctrl, nt = p->n_t, np = p->n_pairs;
  size_t dd = (size_t)d * d;
  int closed = p->dynamics == PR_CLOSED;
  int square = p->objective == PR_PAIRS_SQUARE || p->objective == PR_STATE || p->objective == PR_STATE_PROPAGATORS;
  /* :modulus (closed): F = |z| (Objectives.jl:222), prefactor dt Im(inner / |z|),
   * 0 for |z| < 1e-14 (Objectives.jl:356-361) */
  int modulus = p->objective == PR_PAIRS_MODULUS;
  /* PR_STATE_PROPAGATORS: the state branch of _grape_gradient_from_propagators
   * (MRPhysics.jl:337-351), 2 Im(conj(z) <lambda_k| -i dt A |psi_k>), lambda_k
   * with slice k included */
  int helper = p->objective == PR_STATE_PROPAGATORS;
  /* PR_STATE: _state_gradient_into! (Gradients.jl:357-399), lambda_k = Q_k psi_t
   * with Q_k including U_k (Propagators.jl:375-380): <lambda[k]|A|psi[k]>;
   * the MR kernels pair lambda[k+1] with psi[k] (GRAPEState.jl:218-229) */
  int after = p->objective == PR_STATE || helper;
  double F = 0;
  if (Gm) memset(Gm, 0, sizeof(double) * nc * nt);
  cx* zs = (cx*)malloc(sizeof(cx) * np);
  if (!zs) return PR_ERR_NOMEM;
  for (int s = 0; s < np; s++) {
    cx* psi = W->psi + (size_t)s * d * (nt + 1);
    cx* lam = W->lam + (size_t)s * d * (nt + 1);
    for (int i = 0; i < d; i++) {
      psi[i] = ld(p->psi_init, (size_t)s * d + i);
      lam[(size_t)nt * d + i] = ld(p->psi_targ, (size_t)s * d + i);
    }
    for (int k = 0; k < nt; k++) mv(d, W->Us + k * dd, psi + (size_t)k * d, psi + (size_t)(k + 1) * d);
    cx z = vdot(d, lam + (size_t)nt * d, psi + (size_t)nt * d);
    zs[s] = z;
    F += square ? creal(conj(z) * z) : modulus ? cabs(z) : creal(z);
    if (!Gm) continue;
    for (int k = nt - 1; k >= 0; k--) mv_c(d, W->Us + k * dd, lam + (size_t)(k + 1) * d, lam + (size_t)k * d);
    if (grad_mode == PR_GRAD_EXACT) continue;
    for (int k = 0; k < nt; k++) {
      double dtk = slice_dt(p, k);
      const cx* lv = lam + (size_t)(k + 1 - after) * d;
      const cx* pv = psi + (size_t)k * d;
      for (int j = 0; j < nc; j++) {
        mv(d, W->Hs + j * dd, pv, W->tmpv);
        cx inner = vdot(d, lv, W->tmpv);
        double g;
        if (helper) g = 2.0 * cimag(conj(z) * (-I * dtk) * inner);
        else if (modulus) g = cabs(z) < 1e-14 ? 0.0 : dtk * cimag(inner / cabs(z));
        else if (closed) g = square ? 2.0 * dtk * cimag(conj(z) * inner) : dtk * cimag(inner);
        else g = square ? 2.0 * dtk * creal(conj(z) * inner) : dtk * creal(inner);
        Gm[j + (size_t)k * nc] += g;
      }
    }
  }
  if (Gm && grad_mode == PR_GRAD_EXACT) {
    /* exact: M_k = sum_s conj(w_s) psi_k,s lam_{k+1,s}^dag, inner_j = Tr(A_j L(X_k, M_k)) */
    for (int k = 0; k < nt; k++) {
      double dtk = slice_dt(p, k);
      cx* M = W->T2;
      memset(M, 0, sizeof(cx) * dd);
      for (int s = 0; s < np; s++) {
        cx wgt = square ? 2.0 * conj(zs[s]) : 1.0;
        /* :modulus: the exact derivative of |z| */
        if (modulus) wgt = cabs(zs[s]) < 1e-14 ? 0.0 : conj(zs[s]) / cabs(zs[s]);
        const cx* pv = W->psi + (size_t)s * d * (nt + 1) + (size_t)k * d;
        const cx* lv = W->lam + (size_t)s * d * (nt + 1) + (size_t)(k + 1) * d;
        for (int b = 0; b < d; b++)
          for (int a = 0; a < d; a++) M[a + (size_t)b * d] += wgt * pv[a] * conj(lv[b]);
      }
      member_generator(p, m, u, k, W);
      int rc;
      if (closed) {
        rc = frechet_herm_cx(d, W->H, dtk, M, W->T1, W->V, W->Q, W->A0, W->w);
      } else {
        rc = frechet_cx(d, W->A0, M, W->T1);
      }
      if (rc) {
        free(zs);
        return rc;
      }
      cx c = closed ? -I * dtk : dtk;
      for (int j = 0; j < nc; j++) {
        const cx* Hj = W->Hs + j * dd;
        cx tr = 0;
        for (int q = 0; q < d; q++)
          for (int a = 0; a < d; a++) tr += W->T1[a + (size_t)q * d] * Hj[q + (size_t)a * d];
        Gm[j + (size_t)k * nc] = creal(c * tr);
      }
    }
  }
  free(zs);
  *Fm = F;
  return PR_OK;
}

static int is_gate(const pr_problem* p) { return p->objective == PR_GATE || p->objective == PR_GATE_PROPAGATORS; }

static int member_eval(const pr_problem* p, int m, const double* u, int grad_mode, int separate_fg,
                       double* Fm, double* Gm, work_t* W) {
  int rc;
  if (p->dynamics != PR_CLOSED) {
    /* Lindblad Hs_j = s_j L_j for the gradient inner products */
    size_t dd = (size_t)p->dim * p->dim;
    for (int j = 0; j < p->n_ctrl; j++) {
      double s = cscale(p, m, j);
      for (size_t i = 0; i < dd; i++) W->Hs[j * dd + i] = s * ld(p->Hc, j * dd + i);
    }
  }
  if (separate_fg) {
    /* the fidelity closure: its own propagators (EnsembleObjective.jl:299) */
    rc = member_props(p, m, u, W);
    if (rc) return rc;
    double Fdummy;
    rc = is_gate(p) ? gate_member(p, m, u, grad_mode, Fm, NULL, W)
                                 : pairs_member(p, m, u, grad_mode, Fm, NULL, W);
    if (rc || !Gm) return rc;
    rc = member_props(p, m, u, W);
    if (rc) return rc;
    return is_gate(p) ? gate_member(p, m, u, grad_mode, &Fdummy, Gm, W)
                                   : pairs_member(p, m, u, grad_mode, &Fdummy, Gm, W);
  }
  rc = member_props(p, m, u, W);
  if (rc) return rc;
  return is_gate(p) ? gate_member(p, m, u, grad_mode, Fm, Gm, W)
                                 : pairs_member(p, m, u, grad_mode, Fm, Gm, W);
}

static int check_problem(const pr_problem* p) {
  if (!p || p->dim < 1 || p->n_ctrl < 1 || p->n_t < 1 || p->n_members < 1 || !p->H0 || !p->Hc)
    return PR_ERR_ARG;
  if (p->dynamics != PR_CLOSED && p->dynamics != PR_LINDBLAD) return PR_ERR_ARG;
  if (is_gate(p)) {
    if (!p->U_target || p->dynamics != PR_CLOSED) return PR_ERR_ARG;
  } else if (p->objective == PR_PAIRS_SQUARE || p->objective == PR_PAIRS_REAL ||
             ((p->objective == PR_STATE || p->objective == PR_STATE_PROPAGATORS || p->objective == PR_PAIRS_MODULUS) &&
              p->dynamics == PR_CLOSED)) {
    if (p->n_pairs < 1 || !p->psi_init || !p->psi_targ) return PR_ERR_ARG;
  } else {
    return PR_ERR_ARG;
  }
  return PR_OK;
}

/* Ensemble mean: ensemble_value_and_grad (EnsembleObjective.jl:272-307) with
 * _agg_value/_agg_grad! (:169-217) for the gate objective (w_m = 1/M); the MR
 * kernels' sum(fid_buf)/N_ens (GRAPEState.jl:239-244) for pairs
 * (w_m = 1/(M n_pairs)). Members run in parallel, the reduction is sequential
 * in member order. */
int or_eval(const pr_problem* p, const double* controls, int grad_mode, int separate_fg,
            int n_threads, double* F, double* G, double* Fm_out, double* Gm_out) {
  int rc = check_problem(p);
  if (rc) return rc;
  int M = p->n_members, nc = p->n_ctrl, nt = p->n_t;
  size_t ng = (size_t)nc * nt;
  double* Fm = (double*)malloc(sizeof(double) * M);
  double* Gm = G ? (double*)malloc(sizeof(double) * ng * M) : NULL;
  if (!Fm || (G && !Gm)) {
    free(Fm);
    free(Gm);
    return PR_ERR_NOMEM;
  }
  int err = PR_OK;
#ifdef _OPENMP
  /* threads over members; with fewer members than threads, the members run
   * one after another and their slices' propagators in parallel */
  int nth = n_threads > 0 ? n_threads : omp_get_max_threads();
  int outer = M < nth ? 1 : nth, inner = M < nth ? nth : 1;
  const int columns_before = tl_columns;
  tl_columns = inner;
#pragma omp parallel num_threads(outer)
#else
  int inner = 1;
#endif
  {
    work_t W;
    int lrc = work_alloc(&W, p);
    W.inner = inner;
    if (lrc) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      err = lrc;
    }
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int m = 0; m < M; m++) {
      if (lrc) continue;
      int r = member_eval(p, m, controls, grad_mode, separate_fg, &Fm[m], Gm ? Gm + ng * m : NULL, &W);
      if (r) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
        err = r;
      }
    }
    if (!lrc) work_free(&W);
  }
#ifdef _OPENMP
  tl_columns = columns_before;
#endif
  if (err == PR_OK) {
    double Fs = 0;
    for (int m = 0; m < M; m++) {
      double w = p->weight ? p->weight[m]
                           : (is_gate(p) ? 1.0 / M : 1.0 / ((double)M * p->n_pairs));
      Fs += w * Fm[m];
      if (G) {
        if (m == 0)
          for (size_t i = 0; i < ng; i++) G[i] = 0;
        for (size_t i = 0; i < ng; i++) G[i] += w * Gm[ng * m + i];
      }
    }
    *F = Fs;
    if (Fm_out) memcpy(Fm_out, Fm, sizeof(double) * M);
    if (Gm_out && Gm) memcpy(Gm_out, Gm, sizeof(double) * ng * M);
  }
  free(Fm);
  free(Gm);
  return err;
}

int or_total_propagator(const pr_problem* p, int member, const double* controls, double* U) {
  if (!p || member < 0 || member >= p->n_members) return PR_ERR_ARG;
  work_t W;
  int rc = work_alloc(&W, p);
  if (rc) return rc;
  rc = member_props(p, member, controls, &W);
  if (rc == PR_OK) {
    int d = p->dim;
    size_t dd = (size_t)d * d;
    /* compute_total_propagator (Propagators.jl:212-238) */
    mat_eye(d, W.P);
    for (int k = 0; k < p->n_t; k++) {
      mm_nn(d, W.Us + k * dd, W.P, W.T1);
      memcpy(W.P, W.T1, sizeof(cx) * dd);
    }
    mat_store(d, W.P, U);
  }
  work_free(&W);
  return rc;
}
