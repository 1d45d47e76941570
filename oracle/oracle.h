/*
 * oracle.h -- CPU restatement of the pimd-tunneling ring-polymer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * Fortran layer, the Python host package) may link, load or call this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker / the timed CPU baseline.
 *
 * Every routine restates one reference routine literally (same loop order,
 * same formulas, same quirks) in plain C99, FP64, single-threaded.  The
 * reference file:line it follows is given next to each declaration.
 *
 * Array layout is the reference's: a chain x(n, ndim, natom) in Fortran
 * column-major order, i.e. x[dof*n + bead] with dof = atom*ndim + dim
 * (0-based).  One bead's coordinates are x(ndim, natom) = x[dof].
 */
#ifndef PIMD_ORACLE_H
#define PIMD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* instantonmod.f90:4 -- the truncated literal the reference uses everywhere */
#define OR_PI 3.14159265358979

/* ------------------------------------------------------------------ PES */
/* Potential kinds.  None of the config potentials exists in the reference; they are
 * written against the mcmod_mass API (mcmod_waterdimer.f90:12-103): V returns
 * V_raw(x) - V0, Vprime returns +dV/dx, potforce both, Vdoubleprime the Hessian
 * hess(ndim,natom,ndim,natom) with hess(i,j,k,l) = d grad(k,l) / d x(i,j). */
enum {
  OR_PES_QUARTIC = 1,   /* sum_d h((x_d/x0)^2-1)^2          (form of mcmod_1d.f90:14-23) */
  OR_PES_COUPLED2D = 2, /* V0(x^2/a^2-1)^2 + m w^2 (y+c x^2)^2 / 2                     */
  OR_PES_EVB = 3,       /* two-state EVB of harmonic+cubic wells                        */
  OR_PES_WATER = 4,     /* flexible water cluster: Morse OH + bend + LJ OO + charges    */
  OR_PES_HARMONIC = 5   /* sum_d k_d (x_d - c_d)^2 / 2 (closed-form tests)             */
};

typedef struct {
  int kind, ndim, natom, ndof;
  double v0;            /* mcmod_mass V0, subtracted from every energy */
  const double *dp;     /* kind-specific real parameters (see oracle.c) */
  int ndp;
  const int *ip;        /* kind-specific integer parameters */
  int nip;
} or_pes;

double or_pes_V(const or_pes *pes, const double *x);
void or_pes_Vprime(const or_pes *pes, const double *x, double *grad);
double or_pes_potforce(const or_pes *pes, const double *x, double *grad);
void or_pes_Vdoubleprime(const or_pes *pes, const double *x, double *hess);
/* d V / d theta for the potential parameters an adjoint sensitivity is asked for:
 * EVB: none; WATER: out[0] = dV/dqH (qO = -2 qH), out[1] = dV/dDe. Returns count. */
int or_pes_param_grad(const or_pes *pes, const double *x, double *out);

/* ------------------------------------------------------------------ noise */
/* A sequential source of N(0,1) draws, as MKL VSL's stream is used by the reference
 * (one stream, consumed in call order, verletmodule.f90:55,106,215,641).  Either an
 * injected array (parity mode) or counter-based Philox4x32-10 (the engine's own RNG). */
typedef struct {
  const double *normals;   /* injected normals, or NULL for Philox */
  int64_t nnormals;
  int64_t cursor;          /* next draw index */
  const int32_t *kicks;    /* injected Poisson draws (Andersen), or NULL */
  int64_t nkicks, kcursor;
  uint64_t seed, stream;   /* Philox key / stream id */
  int error;               /* set when an injected array runs out */
} or_rng;

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_philox_normal(uint64_t seed, uint64_t stream, int64_t index);
int32_t or_philox_poisson(uint64_t seed, uint64_t stream, int64_t index, double mean);
void or_rng_normals(or_rng *r, int64_t count, double stdev, double *out);
int32_t or_rng_poisson(or_rng *r, double mean);

/* ------------------------------------------------------------------ normal modes */
typedef struct {
  int n, ndim, natom, ndof;
  double betan, tau, dt, gamma;
  int cayley;
  const double *mass;      /* natom */
  double *lam;             /* n                       verletmodule.f90:314 */
  double *beadmass;        /* [i*natom + j] = beadmass(j,i)   :315 */
  double *T;               /* [i*n + l] = transmatrix(i,l)    :319 */
  double *beadvec;         /* [dof*n + i] = beadvec(i,dof)    :330-333 */
  double *c1, *c2;         /* [i*natom + j] = c1(j,i)          :380-386 */
  double *work;            /* scratch, 4*n*ndof */
} or_nm;

/* init_nm (verletmodule.f90:306-338) + the c1/c2 table (:380-386).  Allocates. */
int or_nm_init(or_nm *nm, int n, int ndim, int natom, const double *mass, double betan,
               double tau, double dt, double gamma, int cayley, const double *a,
               const double *b);
void or_nm_free(or_nm *nm);
/* nmtransform_forward / backward on one dof's n-vector (verletmodule.f90:254-286);
 * bead < 0 means "no shift" (the reference's bead = 0). */
/* swap a BLAS dsymv_ (Fortran ABI, 32-bit integers) into the transforms; NULL = the loop */
void or_use_blas_dsymv(void *f);
void or_nm_forward(const or_nm *nm, const double *x, double *q, int dof);
void or_nm_backward(const or_nm *nm, const double *q, double *x, int dof);

/* Integrator pieces.  Return 0, or 3 when the reference would stop on a NaN. */
int or_step_v(const or_nm *nm, const or_pes *pes, double h, double *x, double *p);   /* :565-585 */
int or_step_nm(const or_nm *nm, double h, double *x, double *p);                      /* :494-561 */
int or_step_langevin(const or_nm *nm, or_rng *rng, double *p);                        /* :634-667 */
int or_time_step_pile(const or_nm *nm, const or_pes *pes, or_rng *rng, double *x, double *p); /* :423-435 */
int or_time_step_nm(const or_nm *nm, const or_pes *pes, double *x, double *p);        /* :291-302 */
/* Andersen momentum resample (verletmodule.f90:211-228). */
void or_andersen_resample(const or_nm *nm, or_rng *rng, double *p);

/* ------------------------------------------------------------------ path / TI */
typedef struct {
  int npath, ndof;
  const double *lampath;   /* npath */
  const double *path;      /* [dof*npath + k] = path(k, dim, atom) */
  double *y2;              /* [dof*npath + k] spline second derivatives */
} or_path;

void or_spline(const double *x, const double *y, int n, double yp1, double ypn, double *y2); /* instantonmod.f90:397-442 */
double or_splint(const double *xa, const double *ya, const double *y2a, int n, double x);    /* :499-528 */
double or_splin_grad(const double *xa, const double *ya, const double *y2a, int n, double x);/* :532-561 */
int or_locate(const double *xx, int n, double x);                                             /* :565-603 */
int or_path_init(or_path *pth, int npath, int ndof, const double *lampath, const double *path);
void or_path_free(or_path *pth);

/* init_path (verletmodule.f90:32-48, 102-115; readhess branch not restated). */
void or_init_path(const or_nm *nm, const or_path *pth, or_rng *rng, double xi, double *x, double *p);

typedef struct {
  int64_t nmc, imin, restartnmc;
  double dhdrlimit;          /* < 0: no limit (pimd_par.f90:78 default -1) */
  const double *dbdl;        /* ndof, dbdl(dim, atom) */
  double xi;
  const or_path *path;       /* for the dHdrlimit re-initialisation */
  int64_t nreinit;           /* out: how many times the path was re-initialised */
  double dhdr_sum;           /* in (restart) / out: running sum */
} or_ti;

/* propagate_pimd_pile (verletmodule.f90:372-416): returns dHdr mean in *dhdr. */
int or_propagate_pile(const or_nm *nm, const or_pes *pes, or_rng *rng, or_ti *ti, double *x,
                      double *p, double *dhdr);
/* propagate_pimd_nm (verletmodule.f90:190-250), Andersen thermostat with Noutput mean. */
int or_propagate_andersen(const or_nm *nm, const or_pes *pes, or_rng *rng, or_ti *ti,
                          int64_t noutput, double *x, double *p, double *dhdr);

/* gauleg (verletmodule.f90:124-160) */
void or_gauleg(double x1, double x2, double *x, double *w, int n);

/* ------------------------------------------------------------------ instanton */
typedef struct {
  int n, ndim, natom, ndof;
  double betan;
  const double *mass;   /* natom */
  int fixedends;
} or_chain;

double or_UM(const or_chain *c, const or_pes *pes, const double *x, const double *a, const double *b); /* instantonmod.f90:17-46 */
void or_UMprime(const or_chain *c, const or_pes *pes, const double *x, const double *a, const double *b, double *g); /* :50-89 */
double or_UMforceenergy(const or_chain *c, const or_pes *pes, const double *x, const double *a, const double *b, double *g); /* :93-151 */
/* UMhessian (:155-217): band[(row)*(totdof) + col] with row in 0..ndof, LAPACK lower band
 * storage AB(1+i-j, j) -> band[(i-j)*totdof + j]. */
void or_UMhessian(const or_chain *c, const or_pes *pes, const double *x, int singlewell, double *band);

/* ------------------------------------------------------------------ CPU baseline */
/* Runs ntraj independent PILE trajectories of nsteps each (Philox noise, seed+traj as the
 * stream), one trajectory per thread as the reference runs one per MPI rank.  Returns the
 * wall seconds; dhdr_out[ntraj] receives each trajectory's TI mean. */
double or_bench_pile(const or_pes *pes, int n, int ndim, int natom, const double *mass,
                     double betan, double tau, double dt, double gamma, const double *a,
                     const double *b, const double *dbdl, const double *x0, int ntraj,
                     int64_t nsteps, int nthreads, uint64_t seed, double *dhdr_out);

#ifdef __cplusplus
}
#endif
#endif
