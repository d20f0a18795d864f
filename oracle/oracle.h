/* The CPU FP64 oracle of the batched Rodas3 chemistry step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline. Nothing in the product (maxichem_amd/) links or
 * calls it. It restates, in plain scalar C, the arithmetic of the reference's call, guard and recover kernels
 * (maxichem/kernels/common.cuh, call.cuh, guard.cuh, recover.cuh), one cell at a time, OpenMP over cells.
 *
 * Every array is in the network's species order except where stated; the elimination order perm only decides the
 * order of the unpivoted LU and of the sums that the reference takes in kernel order.
 */
#ifndef MAXICHEM_ORACLE_H
#define MAXICHEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int ns, nr, nrt, nj, nel;
  const int* perm;      /* [ns] network species index of elimination row k */
  const int* rx;        /* [nr][6] reactants then products, network species indices, -1 pads */
  const double* rc;     /* [nr][14] ln A0, B0, C0, ln A1, B1, C1, ln Ainf, Binf, Cinf, Troe a, T3, T1, T2, eff */
  const int* rf;        /* [nr] law flags (maxichem_amd/_constants.py FL_*) */
  const double* nasa;   /* [2][ns][9] the NASA-9 ranges below and above NASA_SPLIT_T */
  const double* atoms;  /* [ns][nel] */
  double drift;         /* the guard's element-drift allowance */
} oracle_net;

/* Counters in the device count array's order: queue, the ten COUNTS, recovery queue. */
enum { OC_QUEUE, OC_BAD, OC_COLD, OC_HOT, OC_NONFINITE, OC_UNFINISHED, OC_DRIFT, OC_RECOMPUTED, OC_RECOVERED,
       OC_STILL_FAILED, OC_NONFINITE_H, OC_RECOVERY_QUEUE, OC_SLOTS };

/* One call over B cells: the call, the guard and the recovery. J is [B][nj] or NULL. Outputs: xout [B][ns], hout,
 * done, nok, status, flag (0/1/2), tries (every Rodas3 try the cell took, call and recovery) and count[OC_SLOTS]
 * (added to). threads <= 0 takes OpenMP's default. Returns 0. */
int oracle_call(const oracle_net* net, int B, const double* x0, const double* T, const double* P, const double* dt,
                const double* hin, const double* J, double rtol, double atol, int max_steps, double* xout,
                double* hout, int* done, double* nok, int* status, int* flag, double* tries, long long* count,
                int threads);

/* oracle_call with the cells' log rate multipliers lm [B][nr] (NULL: none): both coefficients of a reaction times
 * exp(lm) after the collision limit, a photolysis branch's kf = J exp(lm); a non-finite multiplier makes its cell bad
 * input. */
int oracle_call_lm(const oracle_net* net, int B, const double* x0, const double* T, const double* P, const double* dt,
                   const double* hin, const double* J, const double* lm, double rtol, double atol, int max_steps,
                   double* xout, double* hout, int* done, double* nok, int* status, int* flag, double* tries,
                   long long* count, int threads);

/* Rate coefficients of one cell (T, P, J) before any call: kf, kr [nr] (network reaction order). */
void oracle_rates(const oracle_net* net, double T, double P, const double* J, double* kf, double* kr);

/* The right-hand side F(N) [ns] and the dense Jacobian dF/dN [ns][ns] (network order) at state N. */
void oracle_rhs_jacobian(const oracle_net* net, double T, double P, const double* J, const double* N, double* F,
                         double* JN);

/* The recovery's dense LU with threshold pivoting, in place on A [n][n] (row-major): unit L below the diagonal, U on
 * and above; pr[i] = the original row now in row i; dinv[i] = 1 / U_ii; then y <- A^-1 y for m right-hand sides
 * y [m][n]. */
void oracle_threshold_lu(int n, double* A, int* pr, double* dinv, int m, double* y);

/* The step-size law: the factor on the try's step ha for the next try, from the try's error en, whether it was
 * accepted, and the previous try's rejection, error and step (step_factor of common.cuh:87-100). */
double oracle_step_factor(double en, int ok, int prej, double pen, double pha, double ha);

/* One cell's tries without the guard and the recovery, as the call kernel takes them (dense 0) or as one recovery
 * pass does (dense 1): the next step size h, the time reached t, the accepted steps nok and the tries. x0 in network
 * order, J [nj] or NULL. */
void oracle_try_loop(const oracle_net* net, const double* x0, double T, double P, const double* J, double dt,
                     double h0, double rtol, double atol, int max_steps, int dense, double* h, double* t, double* nok,
                     double* tries);

/* The frozen primal of a tangent (test hooks). oracle_accepted_steps: one cell's call tries as oracle_try_loop takes
 * them (dense 0), recording the accepted tries' sizes into steps (at most cap); returns how many were accepted.
 * oracle_replay: from x0 at (T, P, J), exactly n tries of the sizes steps[], each accepted whatever its error, then
 * x = N / sum N and the guard's check and correction (done, nok n) into xout (network order); returns the guard's
 * bits. */
int oracle_accepted_steps(const oracle_net* net, const double* x0, double T, double P, const double* J, double dt,
                          double h0, double rtol, double atol, int max_steps, double* steps, int cap);
int oracle_replay(const oracle_net* net, const double* x0, double T, double P, const double* J, const double* steps,
                  int n, double atol, double* xout);

/* The guard's test and share correction of one result (check_cell): returns the GUARD_* bits, 0 when it passes, in
 * which case xout is corrected in place. */
int oracle_check_cell(const oracle_net* net, const double* x0, double* xout, int done, double nok, double atol);

#ifdef __cplusplus
}
#endif
#endif
