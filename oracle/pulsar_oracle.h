/*
 * pulsar_oracle.h — FP64 CPU restatement of the reference GRAPE inner loop.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the device path
 * (libpulsar_rocm.so): only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. It is never part of the product path.
 *
 * Every function restates one reference function; the citation is in the
 * comment above its definition in pulsar_oracle.c. Problems are described with
 * the same pr_problem struct as the device library (include/pulsar_rocm.h) so
 * that both are fed identical inputs.
 *
 * Pinning (what independent check holds each function):
 *   or_eigh / or_expm_neg_i  scipy.linalg.expm fixtures; sigma_z / Rabi closed
 *                            forms; unitarity; det = exp(-i dt Tr H)
 *   or_expm                  scipy.linalg.expm fixtures (non-normal 8x8, 64x64)
 *   or_expm_frechet          central finite differences of or_expm; the
 *                            closed form for commuting X, E (L = E exp(X))
 *   or_eval gate, reference  n_t = 1 closed form; FD convergence O(dt^2);
 *                            scipy-built fixtures for n_t <= 5
 *   or_eval gate, exact      central FD to 1e-8 relative
 *   or_eval pairs / Lindblad FD convergence; zero-rate Lindblad equals the
 *                            closed system on vec(rho); trace preservation
 *   ensemble weighting       equals the weighted sum of single-member evals
 */
#ifndef PULSAR_ORACLE_H
#define PULSAR_ORACLE_H

#include "../include/pulsar_rocm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Hermitian eigen-decomposition (cyclic complex Jacobi). A: d x d complex
 * column-major (upper triangle read, as Julia's Hermitian(H)). */
int or_eigh(int d, const double* A, double* w, double* V);

/* exp(-i H dt) via eigen-decomposition (Propagators.jl:659-674). */
int or_expm_neg_i(int d, const double* H, double dt, double* U);

/* exp(A) for a general complex matrix: scaling and squaring with Pade
 * approximants of degree 3..13 (Higham 2005), as LinearAlgebra.exp. */
int or_expm(int n, const double* A, double* E);

/* Frechet derivative L(X, E) of exp at X in direction E (block-triangular
 * exp [[X, E], [0, X]]). */
int or_expm_frechet(int n, const double* X, const double* E, double* L);

/* Ensemble F and G for one problem. grad_mode: PR_GRAD_REFERENCE or
 * PR_GRAD_EXACT. separate_fg = 1 recomputes the exps for F and for G as the
 * reference's per-member closures do (EnsembleObjective.jl:297-301).
 * n_threads <= 0: OpenMP default. G may be NULL (F only). Fm (optional,
 * [n_members]) receives per-member fidelities; Gm (optional,
 * [n_members][n_ctrl*n_t]) per-member gradients. */
int or_eval(const pr_problem* p, const double* controls, int grad_mode, int separate_fg,
            int n_threads, double* F, double* G, double* Fm, double* Gm);

/* Per-member U_total (closed: product of eigen-exps; Lindblad: of Pade exps). */
int or_total_propagator(const pr_problem* p, int member, const double* controls, double* U);

#ifdef __cplusplus
}
#endif

#endif
