// rp_ring.h -- internal interface between the ring kernels and the instanton driver.
#pragma once
#include <vector>

#include "rp_internal.h"

namespace rp {
int ring_eval(rp_ring* r, bool want_grad);            // energies (+ adjoint gradients) on device
int ring_upload_x(rp_ring* r, const double* x);
int ring_download(rp_ring* r, double* e, double* g);  // synchronises the ring's stream
int ring_tot(const rp_ring* r);
int ring_nchain(const rp_ring* r);
int ring_D(const rp_ring* r);
const std::vector<int>& ring_n(const rp_ring* r);
const std::vector<int64_t>& ring_off(const rp_ring* r);
int ring_lowest_eigen(rp_ring* r, const std::vector<int>& ng, double* eta);
// device views used by the instanton search: the evaluation point x, its adjoint gradient
// g and energies e (written by ring_eval), chain offsets and bead counts, the stream
struct RingBufs {
  double* x;
  double* g;
  const double* e;
  const int64_t* off;
  const int* n;
  int D;
  hipStream_t s;
};
RingBufs ring_bufs(rp_ring* r);
// the instanton search's preconditioner (rp_ring_set_preconditioner): Thomas coefficients
// {1/denominator, alpha} per (chain, atom, bead) from tab + off[chain], or tab = nullptr (off)
struct RingPc {
  const double* tab;
  const int64_t* off;
  int ndim;
};
RingPc ring_pc(rp_ring* r);
}  // namespace rp
