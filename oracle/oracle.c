/* CPU FP64 oracle of the batched Rodas3 chemistry step. TEST INFRASTRUCTURE ONLY (the parity checker and the CPU
 * baseline); the product never links it.
 *
 * A scalar restatement of the reference solver's arithmetic, one cell at a time:
 *   rates ........ maxichem/kernels/common.cuh:135-242 (law, collision_limit, gibbs_rt, rate_pair, cell_rates)
 *   screen ....... common.cuh:244-279, call.cuh:35-50
 *   rhs, Jacobian  common.cuh:281-402 (reaction_rate, S v, F = (n_tot / nd) S v, u = S v / nd - Jn N / n_tot)
 *   LU, solves ... call.cuh:81-158 and common.cuh:404-566: M0 = I / (gamma h) - Jn factored without row exchanges
 *                  in the elimination order; W^-1 by Sherman-Morrison; the four Rodas3 stages with explicit fma
 *   step control . common.cuh:87-133, 568-599 (step_factor, accept_try, next_try, try_error, write_state)
 *   guard ........ guard.cuh:8-108 (shares_drifted, element_multipliers, check_cell)
 *   recovery ..... recover.cuh:3-310 (dense_lu with threshold pivoting, dense_solve, form_dense_w, retries)
 * The reference's sparse LU skips structural zeros; this oracle factors the same matrix densely in the same order,
 * which changes only the order of no sums and adds only exact zeros. Sums the reference takes across lanes
 * (n_tot, Sigma x, the error norm) are taken here in elimination order.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* The settings of maxichem_amd/_constants.py (a test compares each with oracle_constant). */
#define T_MIN 300.0
#define T_MAX 3000.0
#define X_SUM_TOL 1e-10
#define REC_STEPS 10
#define H_MIN 1e-12
#define CALL_END (1.0 - 1e-14)
#define STEP_SAFETY 0.9
#define ERR_FLOOR 1e-10
#define FAC_NONFINITE 0.1
#define FACACC 0.2
#define FACREJ 0.2
#define FACMAX 20.0
#define FAC_AFTER_REJECT 0.5
#define SECMIN 1e-4
#define SECANT_MIN_ORDER 0.05
#define SECANT_MAX_ORDER 3.0
#define SECANT_TARGET 0.5
#define PIVTAU 0.1
#define ELEMENT_PIVOT_MIN 1e-10
#define REC_RETRIES 3
#define REC_RETRY_SPLIT 0.5
#define GMAX 1e-3
#define KB 1.380649e-16
#define P0 1.0e6
#define NASA_SPLIT_T 1000.0
#define K_COLL 1e-9
#define FL_FALLOFF 1
#define FL_SECOND_K0 2
#define FL_THREE_BODY 4
#define FL_TROE 8
#define FL_EFF 16
#define FL_VM11 32
#define FL_PHOTO 64
#define FL_NF_SHIFT 8
#define FL_NR_SHIFT 10
#define REFUSED_COLD 1
#define REFUSED_BAD_INPUT 2
#define REFUSED_HOT 3
#define GUARD_NONFINITE 4
#define GUARD_UNFINISHED 8
#define GUARD_DRIFT 16
/* Rodas3 (Sandu et al. 1997), gamma = 1/2; the Troe factor's coefficients (Venot et al. 2012, eq. C.2-C.5). */
#define GAMMA 0.5
#define RODAS_C21 4.0
#define RODAS_C43 (8.0 / 3.0)
#define TROE_C0 0.4
#define TROE_C1 0.67
#define TROE_N0 0.75
#define TROE_N1 1.27
#define TROE_D 0.14

double oracle_constant(const char* name) {
  static const struct { const char* n; double v; } table[] = {
      {"T_MIN", T_MIN}, {"T_MAX", T_MAX}, {"X_SUM_TOL", X_SUM_TOL}, {"REC_STEPS", REC_STEPS}, {"H_MIN", H_MIN},
      {"CALL_END", CALL_END}, {"STEP_SAFETY", STEP_SAFETY}, {"ERR_FLOOR", ERR_FLOOR},
      {"FAC_NONFINITE", FAC_NONFINITE}, {"FACACC", FACACC}, {"FACREJ", FACREJ}, {"FACMAX", FACMAX},
      {"FAC_AFTER_REJECT", FAC_AFTER_REJECT}, {"SECMIN", SECMIN}, {"SECANT_MIN_ORDER", SECANT_MIN_ORDER},
      {"SECANT_MAX_ORDER", SECANT_MAX_ORDER}, {"SECANT_TARGET", SECANT_TARGET}, {"PIVTAU", PIVTAU},
      {"ELEMENT_PIVOT_MIN", ELEMENT_PIVOT_MIN}, {"REC_RETRIES", REC_RETRIES}, {"REC_RETRY_SPLIT", REC_RETRY_SPLIT},
      {"GMAX", GMAX}, {"KB", KB}, {"P0", P0}, {"NASA_SPLIT_T", NASA_SPLIT_T}, {"K_COLL", K_COLL},
      {"FL_FALLOFF", FL_FALLOFF}, {"FL_SECOND_K0", FL_SECOND_K0}, {"FL_THREE_BODY", FL_THREE_BODY},
      {"FL_TROE", FL_TROE}, {"FL_EFF", FL_EFF}, {"FL_VM11", FL_VM11}, {"FL_PHOTO", FL_PHOTO},
      {"FL_NF_SHIFT", FL_NF_SHIFT}, {"FL_NR_SHIFT", FL_NR_SHIFT}};
  for (size_t i = 0; i < sizeof(table) / sizeof(table[0]); i++)
    if (!strcmp(table[i].n, name)) return table[i].v;
  return NAN;
}

/* ---- rates (common.cuh:135-242) ---- */

/* Rate coefficient of a law row q at T (ln T, 1 / T) and n_tot nd, M not folded in (common.cuh:136-161). */
static double law(const double* q, int fl, double T, double lT, double iT, double nd) {
  double k = exp(q[0] + q[1] * lT - q[2] * iT);
  if (fl & FL_SECOND_K0) k += exp(q[3] + q[4] * lT - q[5] * iT);
  if (fl & FL_FALLOFF) {
    const double pr = k * nd / exp(q[6] + q[7] * lT - q[8] * iT);
    k = k > 0.0 ? k / (1.0 + pr) : 0.0;
    if ((fl & FL_TROE) && k > 0.0) {
      const double fc = log10((1.0 - q[9]) * exp(-T / q[10]) + q[9] * exp(-T / q[11]) + exp(-q[12] / T));
      const double x = (fl & FL_VM11) ? log10(pr) : log10(pr) - TROE_C0 - TROE_C1 * fc;
      /* at Pr = 0 Troe's ratio tends to -1 / d; the VM11 one to -inf, so F to 1 */
      const double y = (fl & FL_VM11) ? x / (TROE_N0 - TROE_N1 * fc)
                                      : (isfinite(x) ? x / (TROE_N0 - TROE_N1 * fc - TROE_D * x) : -1.0 / TROE_D);
      k *= pow(10.0, fc / (1.0 + y * y));
    }
  }
  if (fl & FL_EFF) k *= q[13];
  return k;
}

/* A direction's bimolecular-equivalent rate (common.cuh:165-167): M-folded k times nd^(m - 2). */
static double bimolecular(int m, int three_body, double k, double nd) {
  return m >= 2 ? k * (m == 3 ? nd : 1.0) : (three_body ? k / nd : 0.0);
}

/* G / RT of every species at T (common.cuh:183-203). */
static void gibbs(const oracle_net* net, double T, double* g) {
  const double lT = log(T), iT = 1.0 / T, T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  const double* cf = net->nasa + (T < NASA_SPLIT_T ? 0 : net->ns * 9);
  for (int i = 0; i < net->ns; i++) {
    const double* a = cf + 9 * i;
    const double h = -a[0] * iT * iT + a[1] * lT * iT + a[2] + a[3] * T / 2 + a[4] * T2 / 3 + a[5] * T3 / 4 +
                     a[6] * T4 / 5 + a[7] * iT;
    const double s = -a[0] * iT * iT / 2 - a[1] * iT + a[2] * lT + a[3] * T + a[4] * T2 / 2 + a[5] * T3 / 3 +
                     a[6] * T4 / 4 + a[8];
    g[i] = h - s;
  }
}

static int n_forward(int fl) { return (fl >> FL_NF_SHIFT) & 3; }
static int n_reverse(int fl) { return (fl >> FL_NR_SHIFT) & 3; }

/* kf, kr of every reaction (common.cuh:208-242): k from the law; three-body both ways times nd; the reverse from
 * exp(dG/RT) (k_B T / P0)^dnu; then the collision limit (common.cuh:171-180). Photolysis: kf = J, kr = 0. */
/* lm: the cell's log rate multipliers [nr] or NULL: both coefficients of a reaction times exp(lm) after the
 * collision limit, a photolysis branch's kf = J exp(lm) (ad_common.cuh:136-168) */
static void rates_lm(const oracle_net* net, double T, double P, const double* J, const double* lm, double* kf,
                     double* kr) {
  const int ns = net->ns;
  const double nd = P / (KB * T), lT = log(T), iT = 1.0 / T, lKT = log(KB * T / P0);
  double* g = (double*)malloc(sizeof(double) * ns);
  gibbs(net, T, g);
  for (int r = 0; r < net->nr; r++) {
    const int fl = net->rf[r];
    if (fl & FL_PHOTO) {
      kf[r] = J ? J[r - net->nrt] : 0.0;
      kr[r] = 0.0;
      if (lm) kf[r] *= exp(lm[r]);
      continue;
    }
    const int* x = net->rx + 6 * r;
    const double k = law(net->rc + 14 * r, fl, T, lT, iT, nd);
    double dg = 0.0;
    for (int m = 3; m < 6; m++) dg += x[m] >= 0 ? g[x[m]] : 0.0;
    for (int m = 0; m < 3; m++) dg -= x[m] >= 0 ? g[x[m]] : 0.0;
    const int nf = n_forward(fl), nrv = n_reverse(fl), tb = (fl & FL_THREE_BODY) != 0;
    const double c3 = tb ? nd : 1.0;
    double f = k * c3, b = (k * exp(dg + (double)(nrv - nf) * lKT)) * c3;
    const double q = fmax(bimolecular(nf, tb, f, nd), bimolecular(nrv, tb, b, nd));
    if (q > K_COLL) {
      const double c = K_COLL / q;
      f *= c;
      b *= c;
    }
    if (lm) {
      const double e = exp(lm[r]);
      f *= e;
      b *= e;
    }
    kf[r] = f;
    kr[r] = b;
  }
  free(g);
}

void oracle_rates(const oracle_net* net, double T, double P, const double* J, double* kf, double* kr) {
  rates_lm(net, T, P, J, NULL, kf, kr);
}

/* ---- the cell's work space ---- */

/* Each reaction's species with a nonzero net coefficient S[p, r] = products - reactants (reactions whose net
 * column is zero, such as A + B -> A + B, change nothing and take no term, as in the reference's streams). */
typedef struct {
  int nsp;
  int sp[6];
  double cf[6];
} net_column;

typedef struct {
  const oracle_net* net;
  const net_column* col;
  int ns, nr;
  double nd;
  double *kf, *kr, *n, *v, *sv, *A, *dinv, *N, *Y, *F, *u, *K1, *K2, *K3, *K4, *un, *rhsb;
  int* pr;
} cell_t;

static net_column* net_columns(const oracle_net* net) {
  net_column* col = (net_column*)calloc((size_t)net->nr, sizeof(net_column));
  for (int r = 0; r < net->nr; r++) {
    const int* x = net->rx + 6 * r;
    for (int m = 0; m < 6; m++) {
      if (x[m] < 0) continue;
      const double d = m < 3 ? -1.0 : 1.0;
      int e = 0;
      while (e < col[r].nsp && col[r].sp[e] != x[m]) e++;
      if (e == col[r].nsp) {
        col[r].sp[e] = x[m];
        col[r].cf[e] = 0.0;
        col[r].nsp++;
      }
      col[r].cf[e] += d;
    }
    int k = 0; /* drop the species whose net coefficient is zero */
    for (int e = 0; e < col[r].nsp; e++)
      if (col[r].cf[e] != 0.0) {
        col[r].sp[k] = col[r].sp[e];
        col[r].cf[k++] = col[r].cf[e];
      }
    col[r].nsp = k;
  }
  return col;
}

static void cell_alloc(cell_t* c, const oracle_net* net, const net_column* col) {
  const int ns = net->ns, nr = net->nr;
  c->net = net;
  c->col = col;
  c->ns = ns;
  c->nr = nr;
  c->kf = (double*)malloc(sizeof(double) * nr);
  c->kr = (double*)malloc(sizeof(double) * nr);
  c->v = (double*)malloc(sizeof(double) * nr);
  c->n = (double*)malloc(sizeof(double) * ns);
  c->sv = (double*)malloc(sizeof(double) * ns);
  c->A = (double*)malloc(sizeof(double) * ns * ns);
  double** vecs[] = {&c->dinv, &c->N, &c->Y, &c->F, &c->u, &c->K1, &c->K2, &c->K3, &c->K4, &c->un, &c->rhsb};
  for (size_t i = 0; i < sizeof(vecs) / sizeof(vecs[0]); i++) *vecs[i] = (double*)malloc(sizeof(double) * ns);
  c->pr = (int*)malloc(sizeof(int) * ns);
}

static void cell_free(cell_t* c) {
  double* all[] = {c->kf, c->kr, c->v, c->n, c->sv, c->A, c->dinv, c->N, c->Y, c->F, c->u, c->K1, c->K2, c->K3,
                   c->K4, c->un, c->rhsb};
  for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); i++) free(all[i]);
  free(c->pr);
}

/* Rates of progress and S v at the state Y (elimination order; common.cuh:282-353): n = nd Y / n_tot; F =
 * (n_tot / nd) S v; u0 = S v / nd when u0 is not NULL; returns n_tot. */
static double rhs(cell_t* c, const double* Y, double* F, double* u0) {
  const oracle_net* net = c->net;
  const int ns = c->ns, *perm = net->perm;
  double ntot = 0.0;
  for (int i = 0; i < ns; i++) ntot += Y[i];
  const double xf = c->nd / ntot, ixf = ntot / c->nd;
  for (int i = 0; i < ns; i++) c->n[perm[i]] = Y[i] * xf; /* n in network order */
  for (int i = 0; i < ns; i++) c->sv[i] = 0.0;
  for (int r = 0; r < c->nr; r++) {
    const int* x = net->rx + 6 * r;
    double a = c->kf[r], b = c->kr[r];
    for (int m = 0; m < 3; m++) a *= x[m] >= 0 ? c->n[x[m]] : 1.0;
    for (int m = 3; m < 6; m++) b *= x[m] >= 0 ? c->n[x[m]] : 1.0;
    c->v[r] = a - b;
    const net_column* col = c->col + r;
    for (int e = 0; e < col->nsp; e++) c->sv[col->sp[e]] += col->cf[e] * c->v[r];
  }
  for (int i = 0; i < ns; i++) {
    const double s = c->sv[perm[i]];
    F[i] = s * ixf;
    if (u0) u0[i] = s / c->nd;
  }
  return ntot;
}

/* -Jn = -d(S v)/dn at the concentrations c->n, dense [ns][ns] in elimination order, into A (common.cuh:355-392). */
static void minus_jn(cell_t* c, const int* inv) {
  const oracle_net* net = c->net;
  const int ns = c->ns;
  memset(c->A, 0, sizeof(double) * ns * ns);
  for (int r = 0; r < c->nr; r++) {
    const int* x = net->rx + 6 * r;
    for (int side = 0; side < 2; side++) {
      const double k = side ? -c->kr[r] : c->kf[r];
      const int* sp = x + 3 * side;
      for (int m = 0; m < 3; m++) {
        if (sp[m] < 0) continue;
        double d = k; /* d v / d n_q through this occurrence of q */
        for (int o = 0; o < 3; o++)
          if (o != m && sp[o] >= 0) d *= c->n[sp[o]];
        const int q = inv[sp[m]];
        const net_column* col = c->col + r;
        for (int e = 0; e < col->nsp; e++) c->A[inv[col->sp[e]] * ns + q] -= col->cf[e] * d;
      }
    }
  }
}

/* The call's factorization of M0 = A in place: no row exchanges, elimination order (call.cuh:81-112). */
static void lu_plain(int n, double* A, double* dinv) {
  for (int k = 0; k < n; k++) {
    const double dk = 1.0 / A[k * n + k];
    dinv[k] = dk;
    for (int i = k + 1; i < n; i++) {
      if (A[i * n + k] == 0.0) continue;
      const double l = A[i * n + k] * dk;
      A[i * n + k] = l;
      for (int j = k + 1; j < n; j++) A[i * n + j] -= l * A[k * n + j];
    }
  }
}

/* y <- U^-1 L^-1 y with lu_plain's factors (common.cuh:435-522). */
static void solve_plain(int n, const double* A, const double* dinv, double* y) {
  for (int k = 0; k < n; k++)
    for (int i = k + 1; i < n; i++) y[i] -= A[i * n + k] * y[k];
  for (int k = n - 1; k >= 0; k--) {
    const double zk = y[k] * dinv[k];
    for (int i = 0; i < k; i++) y[i] -= A[i * n + k] * zk;
    y[k] = zk;
  }
}

void oracle_threshold_lu(int n, double* A, int* pr, double* dinv, int m, double* y) {
  /* recover.cuh:7-54: exchange only where |A_kk| < PIVTAU * column maximum (ties to the lower row; NaN never taken) */
  for (int i = 0; i < n; i++) pr[i] = i;
  for (int k = 0; k < n; k++) {
    double best = -1.0;
    int bi = n;
    for (int i = k; i < n; i++) {
      const double a = fabs(A[i * n + k]);
      if (a > best) {
        best = a;
        bi = i;
      }
    }
    if (bi != n && bi != k && !(fabs(A[k * n + k]) >= PIVTAU * best)) {
      for (int j = 0; j < n; j++) {
        const double t = A[k * n + j];
        A[k * n + j] = A[bi * n + j];
        A[bi * n + j] = t;
      }
      const int t = pr[k];
      pr[k] = pr[bi];
      pr[bi] = t;
    }
    const double dk = 1.0 / A[k * n + k];
    dinv[k] = dk;
    for (int i = k + 1; i < n; i++) {
      const double l = A[i * n + k] * dk;
      A[i * n + k] = l;
      for (int j = k + 1; j < n; j++) A[i * n + j] -= l * A[k * n + j];
    }
  }
  /* dense_solve (recover.cuh:57-82) */
  double* a = (double*)malloc(sizeof(double) * n);
  for (int r = 0; r < m; r++) {
    double* yr = y + (size_t)r * n;
    for (int i = 0; i < n; i++) a[i] = yr[pr[i]];
    for (int k = 0; k < n - 1; k++)
      for (int i = k + 1; i < n; i++) a[i] -= A[i * n + k] * a[k];
    for (int k = n - 1; k > 0; k--) {
      const double zk = a[k] * dinv[k];
      for (int i = 0; i < k; i++) a[i] -= A[i * n + k] * zk;
    }
    for (int i = 0; i < n; i++) yr[i] = a[i] * dinv[i];
  }
  free(a);
}

/* ---- step control (common.cuh:87-133) ---- */

typedef struct {
  double h, t, nok, pen, pha;
  int prej;
} step_t;

static double step_factor(double en, int ok, int prej, double pen, double pha, double ha) {
  const int efin = isfinite(en);
  double fac = efin ? STEP_SAFETY / cbrt(fmax(en, ERR_FLOOR)) : FAC_NONFINITE;
  fac = fmin(fmax(fac, ok ? FACACC : FACREJ), FACMAX);
  if (!ok && efin && prej && en > 1.0 && pen > 1.0 && pha > ha) { /* secant: err ~ h^pe through the two tries */
    const double pe = log(pen / en) / log(pha / ha);
    fac = pe > SECANT_MIN_ORDER ? fmax(pow(SECANT_TARGET / en, 1.0 / fmin(pe, SECANT_MAX_ORDER)), SECMIN) : SECMIN;
  }
  if (!ok) fac = fmin(fac, FAC_AFTER_REJECT);
  return fac;
}

/* A try's error and its new state (common.cuh:570-589): RMS of K4 / (atol + rtol max(|N|, |y|)), and while no step
 * of the call is accepted the larger of RMS and max-norm. */
static double try_error(int ns, const double* N, const double* K1, const double* K3, const double* K4, double* Y,
                        int first, double rtol, double atol, int* ok) {
  double e2 = 0.0, em = 0.0;
  int nn = 1;
  for (int i = 0; i < ns; i++) {
    const double y = N[i] + 2.0 * K1[i] + K3[i] + K4[i];
    Y[i] = y;
    const double q = K4[i] / (atol + rtol * fmax(fabs(N[i]), fabs(y)));
    e2 += q * q;
    em = fmax(em, fabs(q));
    nn = nn && (y >= -atol);
  }
  const double er = sqrt(e2 / ns);
  const double en = (first && isfinite(er)) ? fmax(er, em) : er;
  *ok = (en <= 1.0) && nn && isfinite(en);
  return en;
}

/* ---- one pass of Rodas3 tries from the state N: the call kernel (dense = 0) or the recovery (dense = 1) ---- */

/* Runs the try loop; returns the step state. N (elimination order) is advanced in place; *tries counts. */
/* Test hooks: forced (not NULL) replays exactly nforced tries of the given sizes, each accepted whatever its error
 * (the frozen primal of a tangent); record (not NULL) receives the accepted tries' sizes, at most cap of them, and
 * *nrecord their number. */
typedef struct {
  const double* forced;
  int nforced;
  double* record;
  int cap;
  int* nrecord;
} try_hooks;

static step_t try_loop_hooked(cell_t* c, const int* inv, double h0, double dt, int max_steps, double rtol,
                              double atol, int dense, double* tries, const try_hooks* hk) {
  const int ns = c->ns;
  step_t s = {h0, 0.0, 0.0, 0.0, 0.0, 0};
  const int forced = hk && hk->forced;
  int it = (dt > 0.0 || forced) ? 0 : max_steps;
  double* W = c->A;
  while (forced ? it < hk->nforced : (it < max_steps && s.t < dt * CALL_END)) {
    const double ha = forced ? hk->forced[it] : fmin(s.h, dt - s.t), igh = 1.0 / (GAMMA * ha), ih = 1.0 / ha;
    /* a retry at the same state recomputes F, u and -Jn: the same inputs give the same values */
    const double ntot = rhs(c, c->N, c->F, c->u);
    minus_jn(c, inv);
    for (int i = 0; i < ns; i++) { /* u = w = S v / nd - Jn N / n_tot (common.cuh:394-402) */
      double acc = 0.0;
      for (int j = 0; j < ns; j++) acc += W[i * ns + j] * c->N[j];
      c->u[i] = c->u[i] + acc / ntot;
    }
    double c0 = 0.0;
    if (!dense) {
      for (int i = 0; i < ns; i++) W[i * ns + i] += igh; /* M0 = I / (gamma h) - Jn */
      lu_plain(ns, W, c->dinv);
      memcpy(c->un, c->u, sizeof(double) * ns); /* Sherman-Morrison: c0 = 1 / (1 - 1^T M0^-1 w) */
      solve_plain(ns, W, c->dinv, c->un);
      double su = 0.0;
      for (int i = 0; i < ns; i++) su += c->un[i];
      c0 = 1.0 / (1.0 - su);
    } else {
      for (int i = 0; i < ns; i++) /* W = I / (gamma h) - Jn - w 1^T (recover.cuh:125-130) */
        for (int j = 0; j < ns; j++) W[i * ns + j] = (i == j ? W[i * ns + j] + igh : W[i * ns + j]) - c->u[i];
      oracle_threshold_lu(ns, W, c->pr, c->dinv, 0, NULL);
    }
    double* Ks[4] = {c->K1, c->K2, c->K3, c->K4};
    for (int st = 1; st <= 4; st++) {
      double* K = Ks[st - 1];
      if (st > 2) { /* F at Y3 = N + 2 K1 and Y4 = N + 2 K1 + K3 */
        for (int i = 0; i < ns; i++) {
          double y = c->N[i] + 2.0 * c->K1[i];
          if (st == 4) y = y + c->K3[i];
          c->Y[i] = y;
        }
        rhs(c, c->Y, c->F, NULL);
      }
      for (int i = 0; i < ns; i++) { /* the stage right-hand sides (common.cuh:560-566) */
        if (st == 1)
          K[i] = c->F[i];
        else if (st == 2)
          K[i] = fma(RODAS_C21 * ih, c->K1[i], c->F[i]);
        else if (st == 3)
          K[i] = fma(ih, c->K1[i] - c->K2[i], c->F[i]);
        else
          K[i] = fma(ih, fma(-RODAS_C43, c->K3[i], c->K1[i] - c->K2[i]), c->F[i]);
      }
      if (!dense) {
        solve_plain(ns, W, c->dinv, K);
        double sk = 0.0;
        for (int i = 0; i < ns; i++) sk += K[i];
        const double sc = sk * c0;
        for (int i = 0; i < ns; i++) K[i] = K[i] + sc * c->un[i];
      } else {
        /* dense_solve on a vector already permuted by pr is the threshold LU's solve */
        double* a = c->rhsb;
        for (int i = 0; i < ns; i++) a[i] = K[c->pr[i]];
        for (int k = 0; k < ns - 1; k++)
          for (int i = k + 1; i < ns; i++) a[i] -= W[i * ns + k] * a[k];
        for (int k = ns - 1; k > 0; k--) {
          const double zk = a[k] * c->dinv[k];
          for (int i = 0; i < k; i++) a[i] -= W[i * ns + k] * zk;
        }
        for (int i = 0; i < ns; i++) K[i] = a[i] * c->dinv[i];
      }
    }
    int ok = 0;
    const double en = try_error(ns, c->N, c->K1, c->K3, c->K4, c->Y, s.t == 0.0, rtol, atol, &ok);
    if (forced) ok = 1;
    if (ok) { /* accept_try (common.cuh:111-118) */
      for (int i = 0; i < ns; i++) c->N[i] = fmax(c->Y[i], 0.0);
      s.t += ha;
      s.nok += 1.0;
      if (hk && hk->record && *hk->nrecord < hk->cap) hk->record[(*hk->nrecord)++] = ha;
    }
    const double fac = step_factor(en, ok, s.prej, s.pen, s.pha, ha);
    s.h = ha * fac; /* next_try (common.cuh:128-133) */
    s.prej = !ok;
    s.pen = en;
    s.pha = ha;
    it++;
    *tries += 1.0;
  }
  return s;
}

static step_t try_loop(cell_t* c, const int* inv, double h0, double dt, int max_steps, double rtol, double atol,
                       int dense, double* tries) {
  return try_loop_hooked(c, inv, h0, dt, max_steps, rtol, atol, dense, tries, NULL);
}

/* x = N / sum N in network order (common.cuh:592-599). */
static void write_state(const cell_t* c, double* x) {
  double s = 0.0;
  for (int i = 0; i < c->ns; i++) s += c->N[i];
  for (int i = 0; i < c->ns; i++) x[c->net->perm[i]] = c->N[i] / s;
}

/* ---- the guard (guard.cuh:13-108) ---- */

int oracle_check_cell(const oracle_net* net, const double* x0, double* xout, int done, double nok, double atol) {
  const int ns = net->ns, nel = net->nel;
  double e0[16], e1[16], lam[16], m[256], m0[16], y[16];
  double etot = 0.0;
  for (int i = 0; i < ns * nel; i++) etot += net->atoms[i];
  for (int e = 0; e < nel; e++) e0[e] = e1[e] = 0.0;
  int xfin = 1;
  for (int i = 0; i < ns; i++) {
    xfin = xfin && isfinite(xout[i]);
    for (int e = 0; e < nel; e++) {
      e0[e] = fma(net->atoms[i * nel + e], x0[i], e0[e]);
      e1[e] = fma(net->atoms[i * nel + e], xout[i], e1[e]);
    }
  }
  double s0 = 0.0, s1 = 0.0;
  for (int e = 0; e < nel; e++) {
    s0 += e0[e];
    s1 += e1[e];
  }
  int drift = 0; /* shares_drifted */
  for (int e = 0; e < nel; e++)
    drift = drift || fabs(e1[e] / s1 - e0[e] / s0) > net->drift * e0[e] / s0 + 2.0 * atol * nok * etot / s0;
  const int code = (xfin ? 0 : GUARD_NONFINITE) | (done ? 0 : GUARD_UNFINISHED) | (drift ? GUARD_DRIFT : 0);
  if (code) return code;
  /* element_multipliers: (A^T X A) lambda = e0 s1 / s0 - e1 by elimination, skipping vanished pivots */
  const double r = s1 / s0;
  for (int e = 0; e < nel * nel; e++) m[e] = 0.0;
  for (int i = 0; i < ns; i++) {
    const double* ai = net->atoms + i * nel;
    for (int e = 0; e < nel; e++) {
      if (ai[e] == 0.0) continue;
      for (int f = e; f < nel; f++)
        if (ai[f] != 0.0) m[e * nel + f] = fma(ai[e] * ai[f], xout[i], m[e * nel + f]);
    }
  }
  for (int e = 0; e < nel; e++) {
    m0[e] = m[e * nel + e];
    y[e] = fma(e0[e], r, -e1[e]);
  }
  for (int k = 0; k < nel; k++) {
    if (!(m[k * nel + k] > ELEMENT_PIVOT_MIN * m0[k])) continue;
    for (int j = k + 1; j < nel; j++) {
      const double l = m[k * nel + j] / m[k * nel + k];
      for (int f = j; f < nel; f++) m[j * nel + f] = fma(-l, m[k * nel + f], m[j * nel + f]);
      y[j] = fma(-l, y[k], y[j]);
    }
  }
  for (int k = nel - 1; k >= 0; k--) {
    lam[k] = 0.0;
    if (!(m[k * nel + k] > ELEMENT_PIVOT_MIN * m0[k])) continue;
    double b = y[k];
    for (int f = k + 1; f < nel; f++) b = fma(-m[k * nel + f], lam[f], b);
    lam[k] = b / m[k * nel + k];
  }
  double sum = 0.0;
  for (int i = 0; i < ns; i++) {
    const double* ai = net->atoms + i * nel;
    double g = 1.0;
    for (int e = 0; e < nel; e++)
      if (ai[e] != 0.0) g = fma(lam[e], ai[e], g);
    g = fmin(fmax(g, 1.0 - GMAX), 1.0 + GMAX);
    sum = fma(xout[i], g, sum);
    xout[i] *= g;
  }
  if (!(isfinite(sum) && sum > 0.0)) return GUARD_NONFINITE;
  const double is = 1.0 / sum;
  for (int i = 0; i < ns; i++) xout[i] *= is;
  return 0;
}

/* ---- one cell: screen, call, guard, recovery ---- */

static void one_cell(cell_t* c, const int* inv, int b, const double* x0, double T, double P, double dt, double hin,
                     const double* J, const double* lm, double rtol, double atol, int max_steps, double* xout,
                     double* hout, int* done, double* nok, int* status, int* flag, double* tries, long long* count) {
  const oracle_net* net = c->net;
  const int ns = c->ns, nj = net->nj;
  const double nd = P / (KB * T);
  c->nd = nd;
  for (int i = 0; i < ns; i++) c->N[i] = x0[net->perm[i]];
  /* the screen (common.cuh:244-279): a NaN enters the sum for a negative or non-finite x or J */
  double p = 0.0, pm = 0.0;
  for (int i = 0; i < ns; i++) {
    p += c->N[i] >= 0.0 ? c->N[i] : NAN;
    pm = fmax(pm, c->N[i]);
  }
  for (int q = 0; q < nj && J; q++) p += J[q] >= 0.0 ? 0.0 * J[q] : NAN;
  for (int r = 0; r < net->nr && lm; r++) p += 0.0 * lm[r]; /* a non-finite multiplier: bad input (ad_common.cuh:38-43) */
  const int fine = fabs(p - 1.0) <= X_SUM_TOL && pm > 0.0 && isfinite(T) && isfinite(nd) && nd > 0.0 &&
                   isfinite(dt) && dt >= 0.0;
  const int st = T < T_MIN ? REFUSED_COLD : (T > T_MAX ? REFUSED_HOT : (fine ? 0 : REFUSED_BAD_INPUT));
  *status = st;
  if (st == 0) rates_lm(net, T, P, J, lm, c->kf, c->kr);
  const double h0 = isfinite(hin) ? hin : dt;
  if (!isfinite(hin)) count[OC_NONFINITE_H]++;
  step_t s;
  if (st == 0) {
    s = try_loop(c, inv, fmax(fmin(h0, dt), H_MIN), dt, max_steps, rtol, atol, 0, tries);
  } else {
    s.h = fmax(fmin(h0, dt), H_MIN);
    s.t = 0.0;
    s.nok = 0.0;
  }
  write_state(c, xout);
  *hout = s.h;
  *done = (st == 0 && s.t >= dt * CALL_END) ? 1 : 0;
  *nok = s.nok;
  *flag = 0;
  if (st) {
    count[st == REFUSED_COLD ? OC_COLD : (st == REFUSED_HOT ? OC_HOT : OC_BAD)]++;
    return;
  }
  const int code = oracle_check_cell(net, x0, xout, *done, *nok, atol);
  if (code & GUARD_NONFINITE) count[OC_NONFINITE]++;
  if (code & GUARD_UNFINISHED) count[OC_UNFINISHED]++;
  if (code & GUARD_DRIFT) count[OC_DRIFT]++;
  if (!code) return;
  count[OC_RECOMPUTED]++;
  /* the recovery (recover.cuh:251-308): from the inputs, dense W with threshold pivoting, REC_STEPS times the tries,
   * and up to REC_RETRIES more passes, with a smaller first try, for a result that only drifted */
  const double mh = fmin(h0, dt);
  for (int k = 0;; k++) {
    for (int i = 0; i < ns; i++) c->N[i] = x0[net->perm[i]];
    const double start = k == 0 ? mh : fmin(mh, dt * pow(REC_RETRY_SPLIT, (double)k));
    s = try_loop(c, inv, fmax(start, H_MIN), dt, REC_STEPS * max_steps, rtol, atol, 1, tries);
    write_state(c, xout);
    const int dn = s.t >= dt * CALL_END ? 1 : 0;
    const int cc = oracle_check_cell(net, x0, xout, dn, s.nok, atol);
    if (cc == GUARD_DRIFT && k < REC_RETRIES) continue;
    *hout = s.h;
    *done = cc ? 0 : dn;
    *nok = s.nok;
    *flag = cc ? 2 : 1;
    count[cc ? OC_STILL_FAILED : OC_RECOVERED]++;
    (void)b;
    return;
  }
}

/* oracle_call with per-cell log rate multipliers lm [B][nr] (NULL: none), the tangent kernels' primal */
int oracle_call_lm(const oracle_net* net, int B, const double* x0, const double* T, const double* P, const double* dt,
                   const double* hin, const double* J, const double* lm, double rtol, double atol, int max_steps,
                   double* xout, double* hout, int* done, double* nok, int* status, int* flag, double* tries,
                   long long* count, int threads) {
  const int ns = net->ns, nj = net->nj;
  int* inv = (int*)malloc(sizeof(int) * ns);
  for (int k = 0; k < ns; k++) inv[net->perm[k]] = k;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  long long total[OC_SLOTS] = {0};
  net_column* col = net_columns(net);
#pragma omp parallel
  {
    cell_t c;
    cell_alloc(&c, net, col);
    long long mine[OC_SLOTS] = {0};
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < B; b++) {
      tries[b] = 0.0;
      one_cell(&c, inv, b, x0 + (size_t)b * ns, T[b], P[b], dt[b], hin[b], J ? J + (size_t)b * nj : NULL,
               lm ? lm + (size_t)b * net->nr : NULL, rtol, atol, max_steps, xout + (size_t)b * ns, hout + b, done + b,
               nok + b, status + b, flag + b, tries + b, mine);
    }
#pragma omp critical
    for (int i = 0; i < OC_SLOTS; i++) total[i] += mine[i];
    cell_free(&c);
  }
  for (int i = 0; i < OC_SLOTS; i++) count[i] += total[i];
  free(col);
  free(inv);
  return 0;
}

int oracle_call(const oracle_net* net, int B, const double* x0, const double* T, const double* P, const double* dt,
                const double* hin, const double* J, double rtol, double atol, int max_steps, double* xout,
                double* hout, int* done, double* nok, int* status, int* flag, double* tries, long long* count,
                int threads) {
  return oracle_call_lm(net, B, x0, T, P, dt, hin, J, NULL, rtol, atol, max_steps, xout, hout, done, nok, status,
                        flag, tries, count, threads);
}

/* The step-size law alone (step_factor above), for the tests that pin it against known answers. */
double oracle_step_factor(double en, int ok, int prej, double pen, double pha, double ha) {
  return step_factor(en, ok, prej, pen, pha, ha);
}

/* The tries of one cell without the guard and the recovery: the call kernel's part of one_cell (dense 0), or one pass
 * of the recovery (dense 1: the dense W with threshold pivoting), from x0 (network order) at h0 over dt, at most
 * max_steps tries. Outputs the next step size, the time reached, the accepted steps and the tries. */
/* A cell set up for one test run: rates at (T, P, J), N = x0 in elimination order. */
static void test_cell(cell_t* c, const oracle_net* net, net_column* col, int* inv, const double* x0, double T,
                      double P, const double* J) {
  const int ns = net->ns;
  cell_alloc(c, net, col);
  c->nd = P / (KB * T);
  oracle_rates(net, T, P, J, c->kf, c->kr);
  for (int k = 0; k < ns; k++) inv[net->perm[k]] = k;
  for (int i = 0; i < ns; i++) c->N[i] = x0[net->perm[i]];
}

int oracle_accepted_steps(const oracle_net* net, const double* x0, double T, double P, const double* J, double dt,
                          double h0, double rtol, double atol, int max_steps, double* steps, int cap) {
  net_column* col = net_columns(net);
  cell_t c;
  int* inv = (int*)calloc((size_t)net->ns, sizeof(int));
  test_cell(&c, net, col, inv, x0, T, P, J);
  int n = 0;
  double tries = 0.0;
  const try_hooks hk = {NULL, 0, steps, cap, &n};
  (void)try_loop_hooked(&c, inv, fmax(fmin(h0, dt), H_MIN), dt, max_steps, rtol, atol, 0, &tries, &hk);
  free(inv);
  free(col);
  cell_free(&c);
  return n;
}

int oracle_replay(const oracle_net* net, const double* x0, double T, double P, const double* J, const double* steps,
                  int n, double atol, double* xout) {
  net_column* col = net_columns(net);
  cell_t c;
  int* inv = (int*)calloc((size_t)net->ns, sizeof(int));
  test_cell(&c, net, col, inv, x0, T, P, J);
  double tries = 0.0;
  const try_hooks hk = {steps, n, NULL, 0, NULL};
  const step_t s = try_loop_hooked(&c, inv, steps[0], 1.0, n, 1.0, atol, 0, &tries, &hk);
  write_state(&c, xout);
  const int code = oracle_check_cell(net, x0, xout, 1, s.nok, atol);
  free(inv);
  free(col);
  cell_free(&c);
  return code;
}

void oracle_try_loop(const oracle_net* net, const double* x0, double T, double P, const double* J, double dt,
                     double h0, double rtol, double atol, int max_steps, int dense, double* h, double* t, double* nok,
                     double* tries) {
  const int ns = net->ns;
  net_column* col = net_columns(net);
  cell_t c;
  cell_alloc(&c, net, col);
  c.nd = P / (KB * T);
  oracle_rates(net, T, P, J, c.kf, c.kr);
  int* inv = (int*)calloc((size_t)ns, sizeof(int));
  for (int k = 0; k < ns; k++) inv[net->perm[k]] = k;
  for (int i = 0; i < ns; i++) c.N[i] = x0[net->perm[i]];
  *tries = 0.0;
  const step_t s = try_loop(&c, inv, fmax(fmin(h0, dt), H_MIN), dt, max_steps, rtol, atol, dense, tries);
  *h = s.h;
  *t = s.t;
  *nok = s.nok;
  free(inv);
  free(col);
  cell_free(&c);
}

/* F(N) and the full Jacobian dF/dN = Jn + w 1^T at N (network order): the oracle's own pieces, for the tests that
 * check them against finite differences and dense solves. */
void oracle_rhs_jacobian(const oracle_net* net, double T, double P, const double* J, const double* N, double* F,
                         double* JN) {
  const int ns = net->ns;
  net_column* col = net_columns(net);
  cell_t c;
  cell_alloc(&c, net, col);
  c.nd = P / (KB * T);
  oracle_rates(net, T, P, J, c.kf, c.kr);
  int* inv = (int*)calloc((size_t)ns, sizeof(int));
  for (int k = 0; k < ns; k++) inv[net->perm[k]] = k;
  for (int i = 0; i < ns; i++) c.N[i] = N[net->perm[i]];
  const double ntot = rhs(&c, c.N, c.F, c.u);
  minus_jn(&c, inv);
  for (int i = 0; i < ns; i++) {
    double acc = 0.0;
    for (int j = 0; j < ns; j++) acc += c.A[i * ns + j] * c.N[j];
    c.u[i] = c.u[i] + acc / ntot;
  }
  for (int i = 0; i < ns; i++) {
    F[net->perm[i]] = c.F[i];
    for (int j = 0; j < ns; j++) JN[net->perm[i] * ns + net->perm[j]] = -c.A[i * ns + j] + c.u[i];
  }
  free(inv);
  free(col);
  cell_free(&c);
}
