/*
 * rp_hip.h -- C ABI of the MI355X ring-polymer hot path (libpimdtunnel.so).
 *
 * Plain C: pointers, sizes and status codes only.  Host buffers use the
 * reference's chain layout x(n, ndim, natom) in Fortran column-major order, i.e.
 * x[dof*n + bead] with dof = atom*ndim + dim (0-based), so a Fortran caller passes
 * its arrays unchanged through iso_c_binding.  Every function returns RP_OK (0) or
 * an error code; rp_last_error() gives the message.  The reference has no error
 * codes: it prints and calls `stop` (verletmodule.f90:533-536, 577-580), which the
 * Fortran binding reproduces on a non-zero status.
 *
 * Threading: one host thread per handle; handles are not re-entrant.  Each handle
 * owns one HIP stream on the device chosen by rp_init().
 */
#ifndef RP_HIP_H
#define RP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RP_OK = 0,
  RP_ERR_ARG = 1,      /* bad argument (size, pointer, unsupported shape) */
  RP_ERR_HIP = 2,      /* HIP runtime failure, or no device */
  RP_ERR_NAN = 3,      /* NaN in the momenta -- the reference's `stop` */
  RP_ERR_NOISE = 4,    /* injected noise stream exhausted */
  RP_ERR_STATE = 5,    /* call out of order */
  RP_ERR_CONVERGE = 6  /* optimiser / eigen-solve did not converge */
};

/* ------------------------------------------------------------------ runtime */
int rp_init(int device);                 /* select the device; idempotent */
int rp_finalize(void);
int rp_last_error(char *buf, int len);   /* copies the last message of this thread */
int rp_version(void);                    /* ABI version, 1 */

/* ------------------------------------------------------------------ PES plugin
 * Device twin of one `module mcmod_mass` (mcmod_waterdimer.f90:1-105).  The kinds
 * and parameter layouts are listed in pimd-tunneling_amd/csrc/rp_pes.h. */
typedef struct rp_pes rp_pes;
enum { RP_PES_QUARTIC = 1, RP_PES_COUPLED2D = 2, RP_PES_EVB = 3, RP_PES_WATER = 4,
       RP_PES_HARMONIC = 5 };

int rp_pes_create(int kind, int ndim, int natom, const double *dp, int ndp, const int *ip,
                  int nip, rp_pes **out);
int rp_pes_destroy(rp_pes *pes);
/* A potential that lives on the host: any mcmod_mass plugin with no device twin, e.g.
 * mcmod_waterdimer.f90:12-103, whose V / Vprime / potforce / Vdoubleprime the reference
 * calls per bead (verletmodule.f90:573, instantonmod.f90:105, 184).  The library calls
 * these back with whole chains: x and grad (nbead, ndim, natom) column-major (the layout of
 * every host buffer here), energy[nbead] as the plugin returns it (its own V - V0);
 * hess[bead][ndof][ndof] with hess[b][i][k] = d grad_k / d x_i.  A callback returns 0, or
 * non-zero to stop the calling routine with RP_ERR_ARG.  hessian may be NULL: central
 * differences of potforce with step 1e-4, as the reference's plugins difference Vprime
 * (mcmod_waterdimer.f90:83-103).  The plain path (rp_chain: each step's kick calls
 * potforce once for the chain's beads) and the ring routines (rp_ring) run with it; the
 * batched engine, whose fused kernel evaluates the potential on the device, refuses it. */
typedef int (*rp_host_potforce_fn)(void *user, int nbead, const double *x, double *grad,
                                   double *energy);
typedef int (*rp_host_hessian_fn)(void *user, int nbead, const double *x, double *hess);
enum { RP_PES_HOST = 6 };
int rp_pes_create_host(int ndim, int natom, rp_host_potforce_fn potforce,
                       rp_host_hessian_fn hessian, void *user, rp_pes **out);
/* module variable V0 (mcmod_waterdimer.f90:3), set by the driver to V(well1)
 * (pimd_par.f90:166, rpi_ser.f90:95) */
int rp_pes_set_v0(rp_pes *pes, double v0);
/* potforce / V / Vprime for nbead beads at once (mcmod_waterdimer.f90:19-80):
 * x and grad are (nbead, ndim, natom) column-major; energy[nbead] (V - V0); grad may be
 * NULL (V only), energy may be NULL (Vprime only). */
int rp_pes_potforce(rp_pes *pes, int nbead, const double *x, double *grad, double *energy);
/* rp_pes_potforce without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it).  Same arguments and results as rp_pes_potforce for a
 * device potential; the staging of x, grad and energy lives in one static device array of
 * 2^20 doubles per GPU, so nbead * (2 * ndof + 1) must stay within 2^20 (RP_ERR_ARG
 * otherwise; 19,065 beads at ndof = 27).  Calls that overlap on one GPU share that array
 * and race.  A potential that lives on the host gets RP_ERR_ARG. */
int rp_pes_potforce_noalloc(rp_pes *pes, int nbead, const double *x, double *grad,
                            double *energy);
/* Vdoubleprime (mcmod_waterdimer.f90:83-103) for nbead beads: hess[bead][ndof][ndof] with
 * hess[b][i][k] = d grad_k / d x_i, from forward-over-reverse (exact, no finite steps). */
int rp_pes_hessian(rp_pes *pes, int nbead, const double *x, double *hess);
/* Adjoint sensitivities of the bead energies to potential parameters (WATER: dV/dqH,
 * dV/dDe); out[bead][nparam].  Returns RP_ERR_ARG for kinds without parameters. */
int rp_pes_param_grad(rp_pes *pes, int nbead, const double *x, double *out, int *nparam);

/* ------------------------------------------------------------------ plain path
 * One trajectory's `verletint` state (verletmodule.f90:10-26) on the device.  Each
 * call is one launch sequence on the chain's stream: the drop-in for callers that
 * step one trajectory at a time, as pimd_par.f90:321-381 does. */
typedef struct rp_chain rp_chain;

/* init_nm (verletmodule.f90:306-338) + alloc of the Langevin table c1/c2 (:380-386):
 * lam, beadmass, transmatrix (truncated pi, instantonmod.f90:4) and beadvec from the
 * fixed ends a, b (ndof each). */
int rp_nm_setup(rp_pes *pes, int n, int ndim, int natom, const double *mass, double betan,
                double tau, double dt, double gamma, int cayley, const double *a,
                const double *b, rp_chain **out);
int rp_chain_destroy(rp_chain *c);
int rp_chain_set_ends(rp_chain *c, const double *a, const double *b); /* re-run init_nm's beadvec */
int rp_chain_set_state(rp_chain *c, const double *x, const double *p);
int rp_chain_get_state(rp_chain *c, double *x, double *p);
/* copies of the init_nm arrays: T[n*n] row-major, lam[n], beadmass[n*natom] ([mode][atom]),
 * beadvec[ndof*n]; any pointer may be NULL */
int rp_chain_get_nm(rp_chain *c, double *T, double *lam, double *beadmass, double *beadvec);
/* Philox stream of this trajectory and its draw cursor (the VSL stream position,
 * verletmodule.f90:363-364); setting the stream also rewinds its Poisson cursor */
int rp_chain_set_rng(rp_chain *c, uint64_t seed, uint64_t stream, int64_t cursor);
int rp_chain_get_cursor(rp_chain *c, int64_t *cursor);
/* position of the stream's Poisson intervals (the Andersen kicks, :202, 230); successive
 * rp_propagate_nm calls continue it, as the reference's VSL Poisson stream continues */
int rp_chain_set_kick_cursor(rp_chain *c, int64_t kick_cursor);
int rp_chain_get_kick_cursor(rp_chain *c, int64_t *kick_cursor);
/* nmtransform_forward (direction 0) / nmtransform_backward (1) (verletmodule.f90:254-286)
 * applied to ncols n-vectors in[col*n + bead]; shift_dof[col] >= 0 subtracts/adds
 * beadvec(:, dof) as the reference does for positions, -1 for momenta. */
int rp_nm_transform(rp_chain *c, int direction, int ncols, const int *shift_dof,
                    const double *in, double *out);
int rp_step_v(rp_chain *c, double h);                    /* step_v        :565-585 */
int rp_step_nm(rp_chain *c, double h);                   /* step_nm       :494-561 */
/* step_langevin :634-667; noise = the n*ndof draws pk((dofi-1)*n+k), or NULL to draw them
 * from the chain's Philox stream */
int rp_step_langevin(rp_chain *c, const double *noise);
int rp_time_step_pile(rp_chain *c, const double *noise); /* time_step_pile :423-435 */
int rp_time_step_nm(rp_chain *c);                        /* time_step_nm   :291-302 */
/* Andersen resample (:211-228) / init_path momenta (:102-115): noise = n*ndof draws in
 * (dim outer, atom inner, n per call) order, or NULL for the Philox stream */
int rp_andersen_resample(rp_chain *c, const double *noise);
/* path spline for init_path and the dHdrlimit re-initialisation: lampath[npath],
 * path[ndof*npath] ((npath, ndim, natom) column-major, as read_path builds it) */
int rp_chain_set_path(rp_chain *c, int npath, const double *lampath, const double *path);
/* init_path (verletmodule.f90:32-48, 102-115); noise as rp_andersen_resample */
int rp_init_path(rp_chain *c, double xi, const double *noise);
/* propagate_pimd_pile (verletmodule.f90:372-416): nmc steps, TI after imin steps,
 * dHdrlimit < 0 disables the re-initialisation.  noise: nnoise draws consumed in order by the
 * O-steps (n*ndof per step) and the re-initialisations (n*ndof each), or NULL for the
 * Philox stream.  dhdr = running sum / (nmc - imin), 0 when nmc = imin. */
int rp_propagate_pile(rp_chain *c, int64_t nmc, int64_t imin, double dhdrlimit,
                      const double *dbdl, double xi, const double *noise, int64_t nnoise,
                      double *dhdr, int64_t *nreinit);
/* propagate_pimd_nm (verletmodule.f90:190-250), Andersen: kicks = Poisson interval draws
 * or NULL (Philox, mean noutput); noise as rp_andersen_resample per resample, or NULL. */
int rp_propagate_nm(rp_chain *c, int64_t nmc, int64_t imin, int64_t noutput,
                    const double *dbdl, const double *noise, int64_t nnoise,
                    const int32_t *kicks, int64_t nkicks, double *dhdr);

/* ------------------------------------------------------------------ batched engine
 * Many trajectories (TI nodes x replicas x bead counts) advanced together by one
 * persistent fused force + thermostat kernel per bead-count group.  Semantics per
 * polymer are those of propagate_pimd_pile / propagate_pimd_nm above. */
typedef struct rp_batch rp_batch;
typedef struct {
  int n;            /* beads (<= 256) */
  double betan;     /* beta/(n+1) for PIMD-TI (pimd_par.f90:94) */
  double xi;        /* TI node: start of init_path and of the re-initialisation */
  const double *a;  /* fixed start (ndof) */
  const double *b;  /* fixed end b(xi) (ndof) */
  const double *dbdl; /* db/dxi (ndof) */
  uint64_t stream;  /* Philox stream id of this trajectory */
} rp_polymer_desc;

enum { RP_THERMO_ANDERSEN = 1, RP_THERMO_LANGEVIN = 2 }; /* pimd_par.f90:103-107 */

int rp_batch_create(rp_pes *pes, const double *mass, int npoly, const rp_polymer_desc *desc,
                    double tau, double dt, double gamma, int thermostat, int64_t noutput,
                    double dhdrlimit, int cayley, uint64_t seed, rp_batch **out);
int rp_batch_destroy(rp_batch *b);
int rp_batch_set_path(rp_batch *b, int npath, const double *lampath, const double *path);
/* init_path for every polymer (verletmodule.f90:32-48, 102-115) from its Philox stream,
 * or from injected draws: noise = each polymer's n*ndof draws in user order, one after
 * the other */
int rp_batch_init_path(rp_batch *b, const double *noise);
int rp_batch_set_state(rp_batch *b, int poly, const double *x, const double *p);
int rp_batch_get_state(rp_batch *b, int poly, double *x, double *p);
/* Andersen parity runs: injected Poisson intervals, nkicks_per_poly per polymer
 * (kicks[poly*nkicks_per_poly + k]); NULL returns to the Philox Poisson draws */
int rp_batch_set_kicks(rp_batch *b, const int32_t *kicks, int64_t nkicks_per_poly);
/* advance every polymer nsteps steps; TI accumulates on steps after imin (steps are
 * counted from the first rp_batch_run).  noise: NULL (Philox) or injected draws for
 * parity runs, nnoise_per_poly per polymer (noise[poly*nnoise_per_poly + k]) consumed in
 * the reference's order by O-steps, resamples and re-initialisations. */
int rp_batch_run(rp_batch *b, int64_t nsteps, int64_t imin, const double *noise,
                 int64_t nnoise_per_poly);
/* per polymer: running TI sum, accumulated step count, re-initialisations */
int rp_batch_ti_readout(rp_batch *b, double *dhdr_sum, int64_t *count, int64_t *nreinit);
/* per polymer: Philox draw cursor (the stream position of its next normal) */
int rp_batch_get_cursor(rp_batch *b, int64_t *cursor);
/* per polymer: Philox cursor of the Andersen Poisson intervals */
int rp_batch_get_kick_cursor(rp_batch *b, int64_t *kick_cursor);
/* resume from a restart file (pimd_par.f90:355-370): set the normal (cursor) and Poisson
 * (kick_cursor) stream positions per polymer; either may be NULL to keep it.  A new
 * kick_cursor discards the Poisson interval in progress, as propagate_pimd_nm draws a fresh
 * one on entry (verletmodule.f90:201). */
int rp_batch_set_cursor(rp_batch *b, const int64_t *cursor, const int64_t *kick_cursor);
/* families of equal (n, betan), one fused-kernel launch each: n[f], count[f] */
int rp_batch_families(rp_batch *b, int *nfam, int *n, int *count);
/* the fused-kernel instance family `fam` runs: waves per workgroup, MFMA units per wave
 * (a unit is 32 normal modes x 16 dofs of x and p) and whether its transforms are folded
 * into two half-size products (even n) */
int rp_batch_family_shape(rp_batch *b, int fam, int *waves, int *units_per_wave, int *folded);
/* launch shape: steps per launch (>= 1; 0 keeps the current value) and waves per
 * polymer workgroup (2, 4 or 8 for every family; 0 keeps each family's value; 8 applies to
 * families of at most 16 MFMA units, 2 to families of at most 2 units with at most 2 dof
 * tiles, the others take 4).  rp_batch_autotune times, on the live batch (state saved and
 * restored), 2, 4 and 8 waves per family from each family's own launch timings, then the
 * steps-per-launch candidates; it keeps the fastest and reports the steps and the most
 * frequent wave count. */
int rp_batch_set_launch(rp_batch *b, int steps_per_launch, int waves);
/* rp_batch_run replays each family's block of steps_per_launch steps as a captured HIP
 * graph, between that family's timing events, when it draws from the Philox streams
 * (enable = 1, the default); enable = 0 launches the kernels directly.  Injected-noise runs
 * always launch directly.  The results are the same either way. */
int rp_batch_set_graphs(rp_batch *b, int enable);
int rp_batch_autotune(rp_batch *b, int trial_steps, int *best_steps, int *best_waves);
/* hipEvent-timed average duration of one fused-kernel launch per family in the last
 * rp_batch_run: avg_ms[nfam]; launches = launches per family */
int rp_batch_last_kernel_ms(rp_batch *b, double *avg_ms, int64_t *launches);

/* ------------------------------------------------------------------ instantons
 * A batch of open fixed-end chains of uneven n (instantonmod.f90:17-217, 679-827). */
typedef struct rp_ring rp_ring;
int rp_ring_create(rp_pes *pes, const double *mass, int nchain, const int *n,
                   const double *betan, const double *a, const double *b, int fixedends,
                   rp_ring **out);
int rp_ring_destroy(rp_ring *r);
/* x: the chains concatenated, each (n_c, ndim, natom) column-major */
int rp_ring_set_x(rp_ring *r, const double *x);
int rp_ring_get_x(rp_ring *r, double *x);
/* UM / UMprime / UMforceenergy (:17-151): energy[nchain] and, if grad != NULL, the
 * adjoint gradient in the layout of x */
int rp_ring_um(rp_ring *r, double *energy, double *grad);
/* Hessian-vector product of U_M (forward-over-reverse through the adjoint) */
int rp_ring_hvp(rp_ring *r, const double *v, double *hv);
/* With no counterpart in the reference (the reference holds its chains in host arrays):
 * rp_ring_set_x, rp_ring_um and rp_ring_hvp for callers whose chains already sit in the
 * device memory of the ring's GPU (torch tensors under torch.autograd): x, grad, v and hv
 * are device pointers in the same layouts, copied device to device; energy stays a host
 * array.  The caller's writes to x and v must be complete before the call (synchronise the
 * stream that made them); the outputs are complete on return. */
int rp_ring_set_x_device(rp_ring *r, const double *x);
int rp_ring_um_device(rp_ring *r, double *energy, double *grad);
int rp_ring_hvp_device(rp_ring *r, const double *v, double *hv);
/* UMhessian (:155-217) of chain `chain`: the reference's answer(ndof+1, n*ndof), LAPACK
 * lower band storage in column-major order (AB(1+i-j, j) at band[j*(ndof+1) + i-j], 0-based),
 * bead blocks from ndof forward-over-reverse sweeps; singlewell reuses bead 1's block
 * (:183-185) */
int rp_ring_hessian_band(rp_ring *r, int chain, int singlewell, double *band);
/* log-determinant of every chain's mass-weighted ring Hessian by FP64 block LU, with
 * the reference's coupling pattern (:203), reduced as detJ's callers reduce the eigenvalues
 * (ascending, instantonmod.f90:782-827): logdet_pos[c] = sum of ln(eta) over eta > 0,
 * skipping eta_1 (the lowest) when skip_lowest (lndetJ, rpi_ser.f90:355-365), else over
 * every eta > 0 (lndetJ0, :227-236); nneg[c] = number of negative eigenvalues (the inertia);
 * eta1[c] = the lowest eigenvalue.  With no negative eigenvalue it is held on a bracket of
 * inertia counts and converged by inverse iteration shifted below it, so a cluster at the
 * bottom of the spectrum cannot trade it for a neighbour; negative eigenvalues beyond the
 * one inverse iteration finds are located by bisection on the inertia, so any spectrum is
 * reduced. */
int rp_ring_logdet(rp_ring *r, int singlewell, int skip_lowest, double *logdet_pos,
                   int *nneg, double *eta1);
/* the k lowest eigenvalues of every chain's mass-weighted ring Hessian (the band of
 * rp_ring_hessian_band), ascending, eigs[chain*k + j], each by bisection on the inertia of
 * the block LDL^T of H - sigma I to a few ulps: the head of the ascending spectrum DSBEVD
 * returns in detJ (instantonmod.f90:782-827) */
int rp_ring_lowest_eigs(rp_ring *r, int singlewell, int k, double *eigs);
/* An option of rp_instanton with no counterpart in the reference: L-BFGS with the
 * initial inverse Hessian gamma P^-1 instead of gamma I, where P is the spring part of the
 * U_M Hessian per dof column, (m/betan^2) tridiag(-1, 2, -1) + sigma I (one-sided end rows
 * for free ends), applied by a tridiagonal solve, and gamma = s.y / y.P^-1.y from the
 * newest correction pair.  The springs set the stiff end of the spectrum (the iteration
 * count of the plain method grows with N and beta); sigma stands in for the potential's
 * curvature.  Same stopping tests, same minimum.  sigma <= 0 switches it off (the default:
 * the reference's method). */
int rp_ring_set_preconditioner(rp_ring *r, double sigma);
/* instanton (:679-777): L-BFGS-B (m <= 32 corrections, unbounded as nbd=0) on every chain
 * at once, resident on the device (iterates, gradients and correction pairs stay in HBM;
 * gradients from the adjoint).  x_inout as rp_ring_set_x; action[c] = U_M at the optimum;
 * iters[c]; status[c] = 0 converged, 1 abnormal line search, 2 iteration cap (maxiter, or
 * 200000 evaluations when maxiter = 0).  The optimum is left as the ring's x. */
int rp_instanton(rp_ring *r, double pgtol, double factr, int m, int maxiter, double *x_inout,
                 double *action, int *iters, int *status);
/* adjoint sensitivity dU_M/dtheta of each chain to the potential parameters
 * (rp_pes_param_grad summed over beads): out[chain][nparam] */
int rp_ring_param_grad(rp_ring *r, double *out, int *nparam);

#ifdef __cplusplus
}
#endif
#endif
