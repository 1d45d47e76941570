// rp_batch.hip -- host side of the batched engine (rp_batch_* in rp_hip.h).
//
// Polymers are grouped into families of equal (n, betan): one family shares the
// transform T and the per-mode table, and is advanced by one launch of the fused
// step kernel (rp_batch_kernel.h) with one workgroup per polymer.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>

#include "rp_batch.h"

using namespace rp;

namespace {
struct Family {
  int N = 0, NPAD = 0;
  double betan = 0;
  int first = 0, count = 0;  // polymers [first, first+count) in batch order
  int waves = 4;             // waves per workgroup of this family's step kernel (4 or 8)
  DevBuf<double> d_T, d_Tf, d_modes;
};
}  // namespace

struct rp_batch {
  rp_pes* pes = nullptr;
  int ndim = 0, natom = 0, D = 0, DT = 0, DC = 0;
  int npoly = 0;
  double tau = 0, dt = 0, gamma = 0, limit = -1;
  int thermostat = RP_THERMO_LANGEVIN, cayley = 0;
  int64_t noutput = 1;
  uint64_t seed = 0;
  std::vector<int> order;     // batch position -> user index
  std::vector<int> pos;       // user index -> batch position
  std::vector<int> N;         // per batch position
  std::vector<int64_t> soff;  // per batch position, state offset (doubles)
  std::vector<Family> fam;
  hipStream_t s = nullptr;
  DevBuf<double> d_state, d_pf64, d_pdata, d_massd, d_noise, d_pathd;
  DevBuf<int64_t> d_pi64;
  DevBuf<uint64_t> d_stream;
  DevBuf<int32_t> d_kicks;
  DevBuf<int> d_flags;
  int npath = 0;
  int64_t kick_cap = 0;
  int64_t steps_done = 0;
  int steps_per_launch = 64;
  std::vector<hipEvent_t> fev;  // per family: start, end
  std::vector<double> fam_ms;   // per family: summed launch time of the last run
  int64_t last_launches = 0;
  // each family's launch of a step block (the last family's followed by the step counter
  // update) captured as a HIP graph and replayed between its timing events; a graph is
  // rebuilt when anything it captured changes
  int use_graph = 1;
  struct GraphKey {
    int fam = -1;
    int steps = -1;
    int64_t imin = -1, kick_cap = -1;
    const void* kicks = nullptr;
    const void* pathd = nullptr;
    int npath = -1;
    std::vector<int> waves;
    bool operator==(const GraphKey& o) const {
      return fam == o.fam && steps == o.steps && imin == o.imin && kick_cap == o.kick_cap &&
             kicks == o.kicks &&
             pathd == o.pathd && npath == o.npath && waves == o.waves;
    }
  };
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;  // recent family step blocks
  DevBuf<int64_t> d_step0;
};

static void graph_drop(rp_batch* b) {
  for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.second);
  b->graphs.clear();
}

// Folded halves of T for even N (H = N/2, HPAD = H rounded up to 16): row r of the odd
// half is mode 2r+1 and row r of the even half mode 2r+2 (1-based), both over beads
// 1..H.  Because sin(i (N+1-l) pi/(N+1)) = (-1)^(i+1) sin(i l pi/(N+1)), T x splits into
// odd modes = Todd (x_l + x_{N+1-l}) and even modes = Teven (x_l - x_{N+1-l}), half the
// multiply-adds; the identity holds to 3.4e-13 relative at N = 256 with the reference's
// truncated pi.
// Stored four times, each HPAD x HPAD and zero-padded: [odd^T | even^T | odd | even],
// the transposes for the forward transform and the halves themselves for the backward one,
// so both read rows of 16 consecutive entries.
static void family_folded(const Family& f, const std::vector<double>& T, std::vector<double>& Tf) {
  const int N = f.N, NP = f.NPAD, H = N / 2, HP = (H + 15) / 16 * 16;
  const size_t Q = (size_t)HP * HP;
  Tf.assign(4 * Q, 0.0);
  for (int r = 0; r < H; ++r)
    for (int k = 0; k < H; ++k) {
      const double to = T[(size_t)(2 * r) * NP + k], te = T[(size_t)(2 * r + 1) * NP + k];
      Tf[(size_t)k * HP + r] = to;
      Tf[Q + (size_t)k * HP + r] = te;
      Tf[2 * Q + (size_t)r * HP + k] = to;
      Tf[3 * Q + (size_t)r * HP + k] = te;
    }
}

static void family_tables(const rp_batch* b, const Family& f, std::vector<double>& T,
                          std::vector<double>& modes) {
  const int N = f.N, NP = f.NPAD;
  T.assign((size_t)NP * NP, 0.0);
  modes.assign((size_t)NP * kMODE, 0.0);
  for (int i = 1; i <= N; ++i)
    for (int l = i; l <= N; ++l) {
      double t = std::sin((double)(i * l) * kPiRef / (double)(N + 1)) * std::sqrt(2.0 / (double)(N + 1));
      T[(size_t)(i - 1) * NP + (l - 1)] = t;
      T[(size_t)(l - 1) * NP + (i - 1)] = t;
    }
  for (int i = 0; i < NP; ++i) {
    double* md = modes.data() + (size_t)i * kMODE;
    if (i < N) {
      const int ii = i + 1;
      const double lam = 2.0 * std::sin((double)ii * kPiRef / (double)(2 * N + 2)) / f.betan;
      const double lb = lam * f.betan;
      const double c = std::sqrt(2.0 / (double)(N + 1)) / (lb * lb);
      md[MD_MU] = lam * lam;
      md[MD_IMU] = 1.0 / (lam * lam);
      md[MD_U] = std::sin((double)ii * kPiRef / (double)(N + 1)) * c;
      md[MD_W] = std::sin((double)(N * ii) * kPiRef / (double)(N + 1)) * c;
      md[MD_LAM] = lam;
    } else {  // padded modes: state stays zero; keep every coefficient finite
      md[MD_MU] = 1.0;
      md[MD_IMU] = 1.0;
    }
  }
}

static int batch_flags(rp_batch* b) {
  int f[2] = {0, 0};
  RP_HIP(hipMemcpyAsync(f, b->d_flags.p, sizeof(f), hipMemcpyDeviceToHost, b->s));
  RP_HIP(hipStreamSynchronize(b->s));
  if (f[0] || f[1]) {
    int z[2] = {0, 0};
    RP_HIP(hipMemcpy(b->d_flags.p, z, sizeof(z), hipMemcpyHostToDevice));
    if (f[0]) return set_error(RP_ERR_NAN, "NaN in the momenta of a batched polymer");
    return set_error(RP_ERR_NOISE, "injected noise or Poisson draws exhausted in the batch");
  }
  return RP_OK;
}

bool rp::batch_supports(const rp_pes* pes, int n) {
  if (!pes || n < 2 || n > 256) return false;
  const int D = pes->ndof;
  bool kind_ok = false;
  switch (pes->kind) {
    case PES_EVB: kind_ok = D <= 32; break;
    case PES_WATER: kind_ok = pes->natom == 3 || pes->natom == 6 || pes->natom == 18; break;
    case PES_QUARTIC: kind_ok = D <= 32; break;
    case PES_HARMONIC: kind_ok = D <= 16; break;
    case PES_COUPLED2D: kind_ok = true; break;
    default: kind_ok = false;  // host potentials and unknown kinds
  }
  const int npad = (n + 31) / 32 * 32;
  return kind_ok && batch_lds_bytes(npad, batch_dtiles(pes), batch_pes_lds(pes, n)) <= 160 * 1024;
}

// Waves per workgroup by default: 8 (two per SIMD, which hides the latency of the
// elementwise phases) from 8 to 16 MFMA units, 4 otherwise.  16 waves (four per SIMD, one
// unit each) measured slower at every N (8-step launch of 5632 polymers, ms, 8 / 16 waves:
// N = 128 4.39 / 8.75, 192 7.50 / 11.20, 256 11.31 / 13.59).  Measured on the 27-dof EVB
// (ms per 8-step launch of 5632 polymers, 4 / 8 waves): 2 units 1.37 / 2.15, 4 units
// 1.69 / 2.68, 6 units 3.43 / 3.85 (two 4-wave workgroups share a CU), 8 units 5.30 / 4.37,
// 12 units 8.88 / 7.51, 16 units 13.4 / 11.3.  Four 2-wave workgroups per CU for 2 units (no
// idle wave) measured 1.45 ms against 1.36 for two 4-wave ones.
static int default_waves(int DT, int NPAD) {
  const int units = (NPAD >> 5) * DT;
  if (units <= 2 && DT <= 2) return 2;
  return (units >= 8 && units <= 16) ? 8 : 4;
}

static BatchArgs base_args(rp_batch* b, const Family& f, double h) {
  BatchArgs a;
  memset(&a, 0, sizeof(a));
  a.pes = b->pes->dev();
  a.waves = f.waves;
  a.N = f.N;
  a.NPAD = f.NPAD;
  a.D = b->D;
  a.ndim = b->ndim;
  a.natom = b->natom;
  a.LD = 2 * b->DC + 1;
  a.first = f.first;
  a.dt = b->dt;
  a.limit = b->limit;
  a.stdev0 = std::sqrt(1.0 / f.betan);
  a.tau = b->tau;
  const double tau = b->tau;
  if (b->cayley) {
    const double w2 = h * h / (tau * tau);
    a.c = (4.0 - w2) / (4.0 + w2);
    a.sq = 4.0 * h / (tau * tau * (4.0 + w2));
    a.sp = 4.0 * h / (4.0 + w2);
  } else {
    a.c = std::cos(h / tau);
    a.sq = std::sin(h / tau) / tau;
    a.sp = tau * std::sin(h / tau);
  }
  const double c1 = std::exp(-b->gamma * b->dt / tau);
  const double c2 = std::sqrt(1.0 - c1 * c1);
  a.c1sq = c1 * c1;
  a.kappa = tau * c2 * std::sqrt((1.0 + c1 * c1) / f.betan);
  a.noutput = b->noutput;
  a.seed = b->seed;
  a.state = b->d_state.p;
  a.pi64 = b->d_pi64.p;
  a.pf64 = b->d_pf64.p;
  a.stream = b->d_stream.p;
  a.T = f.d_T.p;
  a.Tf = f.d_Tf.p;
  a.fold = (f.N % 2 == 0) && f.d_Tf.p != nullptr;
  a.modes = f.d_modes.p;
  a.pdata = b->d_pdata.p;
  a.massd = b->d_massd.p;
  a.npath = b->npath;
  a.pathd = b->d_pathd.p;
  a.flags = b->d_flags.p;
  return a;
}

static int launch_family(rp_batch* b, const BatchArgs& a, int npoly) {
  switch (b->pes->kind) {
    case PES_EVB: return launch_batch_evb(b->pes, a, b->thermostat, npoly, b->s);
    case PES_WATER: return launch_batch_water(b->pes, a, b->thermostat, npoly, b->s);
    default: return launch_batch_simple(b->pes, a, b->thermostat, npoly, b->s);
  }
}

extern "C" int rp_batch_create(rp_pes* pes, const double* mass, int npoly,
                               const rp_polymer_desc* desc, double tau, double dt, double gamma,
                               int thermostat, int64_t noutput, double dhdrlimit, int cayley,
                               uint64_t seed, rp_batch** out) {
  RP_REQUIRE(out, "rp_batch_create: out is NULL");
  *out = nullptr;
  RP_REQUIRE(pes && mass && desc && npoly > 0, "rp_batch_create: bad arguments");
  RP_REQUIRE(thermostat == RP_THERMO_LANGEVIN || thermostat == RP_THERMO_ANDERSEN,
             "rp_batch_create: thermostat must be 1 (Andersen) or 2 (Langevin)");
  RP_REQUIRE(tau > 0 && dt > 0, "rp_batch_create: tau and dt must be positive");
  RP_REQUIRE(thermostat != RP_THERMO_ANDERSEN || noutput > 0, "rp_batch_create: noutput <= 0");
  RP_REQUIRE(pes->kind != PES_HOST,
             "rp_batch_create: the fused kernel evaluates the potential on the device; a "
             "potential that lives on the host runs through the plain path (rp_nm_setup)");
  int rc = ensure_init();
  if (rc) return rc;
  rp_batch* b = new rp_batch();
  b->pes = pes;
  b->ndim = pes->ndim;
  b->natom = pes->natom;
  b->D = pes->ndof;
  b->DT = batch_dtiles(pes);
  b->DC = 16 * b->DT;
  b->npoly = npoly;
  b->tau = tau;
  b->dt = dt;
  b->gamma = gamma;
  b->limit = dhdrlimit;
  b->thermostat = thermostat;
  b->cayley = cayley;
  b->noutput = noutput;
  b->seed = seed;
  auto fail = [&](int code) {
    rp_batch_destroy(b);
    return code;
  };
  for (int p = 0; p < npoly; ++p) {
    const int n = desc[p].n;
    if (n < 2 || n > 256) return fail(set_error(RP_ERR_ARG, "polymer %d: n=%d outside 2..256", p, n));
    if (!(desc[p].betan > 0) || !desc[p].a || !desc[p].b || !desc[p].dbdl)
      return fail(set_error(RP_ERR_ARG, "polymer %d: bad descriptor", p));
    const int npad = (n + 31) / 32 * 32;
    if (batch_lds_bytes(npad, b->DT, batch_pes_lds(pes, n)) > 160 * 1024)
      return fail(set_error(RP_ERR_ARG, "polymer %d: n=%d with %d dof needs more LDS than a CU has",
                            p, n, b->D));
  }
  // families of equal (n, betan), each contiguous in batch order
  std::map<std::pair<int, double>, std::vector<int>> groups;
  for (int p = 0; p < npoly; ++p) groups[{desc[p].n, desc[p].betan}].push_back(p);
  for (auto& kv : groups) {
    Family f;
    f.N = kv.first.first;
    f.NPAD = (f.N + 31) / 32 * 32;
    f.betan = kv.first.second;
    f.first = (int)b->order.size();
    f.count = (int)kv.second.size();
    f.waves = default_waves(b->DT, f.NPAD);
    for (int p : kv.second) b->order.push_back(p);
    b->fam.push_back(std::move(f));
  }
  b->pos.assign(npoly, 0);
  for (int k = 0; k < npoly; ++k) b->pos[b->order[k]] = k;
  const int D = b->D, DC = b->DC;
  std::vector<int64_t> pi64((size_t)npoly * kPI64, 0);
  std::vector<double> pf64((size_t)npoly * kPF64, 0.0), pdata((size_t)npoly * 3 * DC, 0.0);
  std::vector<uint64_t> streams(npoly);
  int64_t off = 0;
  b->N.resize(npoly);
  b->soff.resize(npoly);
  for (int k = 0; k < npoly; ++k) {
    const rp_polymer_desc& d = desc[b->order[k]];
    b->N[k] = d.n;
    b->soff[k] = off;
    pi64[(size_t)k * kPI64 + PI_STATE_OFF] = off;
    pi64[(size_t)k * kPI64 + PI_RKICK] = -1;  // first Poisson interval not drawn yet
    off += 2 * (int64_t)D * d.n;
    pf64[(size_t)k * kPF64 + PF_XI] = d.xi;
    for (int j = 0; j < D; ++j) {
      pdata[(size_t)k * 3 * DC + j] = d.a[j];
      pdata[(size_t)k * 3 * DC + DC + j] = d.b[j];
      pdata[(size_t)k * 3 * DC + 2 * DC + j] = d.dbdl[j];
    }
    streams[k] = d.stream;
  }
  std::vector<double> massd(DC, 1.0);
  for (int a = 0; a < b->natom; ++a)
    for (int j = 0; j < b->ndim; ++j) massd[a * b->ndim + j] = mass[a];
  int zf[2] = {0, 0};
  if ((rc = [&]() -> int {
        RP_HIP(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
        b->fev.assign(2 * b->fam.size(), nullptr);
        for (auto& e : b->fev) RP_HIP(hipEventCreate(&e));
        int r;
        if ((r = b->d_state.alloc(off))) return r;
        RP_HIP(hipMemsetAsync(b->d_state.p, 0, off * sizeof(double), b->s));
        if ((r = b->d_pi64.alloc(pi64.size())) || (r = b->d_pi64.upload(pi64.data(), pi64.size(), b->s))) return r;
        if ((r = b->d_pf64.alloc(pf64.size())) || (r = b->d_pf64.upload(pf64.data(), pf64.size(), b->s))) return r;
        if ((r = b->d_pdata.alloc(pdata.size())) || (r = b->d_pdata.upload(pdata.data(), pdata.size(), b->s))) return r;
        if ((r = b->d_stream.alloc(npoly)) || (r = b->d_stream.upload(streams.data(), npoly, b->s))) return r;
        if ((r = b->d_massd.alloc(DC)) || (r = b->d_massd.upload(massd.data(), DC, b->s))) return r;
        if ((r = b->d_flags.alloc(2)) || (r = b->d_flags.upload(zf, 2, b->s))) return r;
        for (Family& f : b->fam) {
          std::vector<double> T, modes;
          family_tables(b, f, T, modes);
          if ((r = f.d_T.alloc(T.size())) || (r = f.d_T.upload(T.data(), T.size(), b->s))) return r;
          if (f.N % 2 == 0) {
            std::vector<double> Tf;
            family_folded(f, T, Tf);
            if ((r = f.d_Tf.alloc(Tf.size())) || (r = f.d_Tf.upload(Tf.data(), Tf.size(), b->s))) return r;
          }
          if ((r = f.d_modes.alloc(modes.size())) || (r = f.d_modes.upload(modes.data(), modes.size(), b->s)))
            return r;
        }
        RP_HIP(hipStreamSynchronize(b->s));
        return RP_OK;
      }()))
    return fail(rc);
  *out = b;
  return RP_OK;
}

extern "C" int rp_batch_destroy(rp_batch* b) {
  if (!b) return RP_OK;
  if (b->s) (void)hipStreamSynchronize(b->s);
  graph_drop(b);
  b->d_step0.free();
  for (Family& f : b->fam) {
    f.d_T.free();
    f.d_Tf.free();
    f.d_modes.free();
  }
  for (DevBuf<double>* d : {&b->d_state, &b->d_pf64, &b->d_pdata, &b->d_massd, &b->d_noise, &b->d_pathd})
    d->free();
  b->d_pi64.free();
  b->d_stream.free();
  b->d_kicks.free();
  b->d_flags.free();
  for (auto& e : b->fev)
    if (e) (void)hipEventDestroy(e);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
  return RP_OK;
}

extern "C" int rp_batch_set_path(rp_batch* b, int npath, const double* lampath, const double* path) {
  RP_REQUIRE(b && lampath && path && npath >= 2, "rp_batch_set_path: bad arguments");
  std::vector<double> y2;
  path_spline_host(npath, b->D, lampath, path, y2);
  std::vector<double> pd((size_t)npath * (1 + 2 * b->D));
  memcpy(pd.data(), lampath, npath * sizeof(double));
  memcpy(pd.data() + npath, path, (size_t)npath * b->D * sizeof(double));
  memcpy(pd.data() + (size_t)npath * (1 + b->D), y2.data(), (size_t)npath * b->D * sizeof(double));
  int rc;
  if ((rc = b->d_pathd.alloc(pd.size())) || (rc = b->d_pathd.upload(pd.data(), pd.size(), b->s))) return rc;
  b->npath = npath;
  RP_HIP(hipStreamSynchronize(b->s));
  return RP_OK;
}

// Injected draws: polymer p (user order) owns noise[p*cap, (p+1)*cap), indexed from 0 for
// this call; the Philox cursors are saved before and restored after, so injected calls
// leave the engine's own streams where they were.
static int noise_begin(rp_batch* b, const double* noise, int64_t cap, std::vector<int64_t>& saved) {
  int rc;
  const size_t tot = (size_t)b->npoly * cap;
  if (b->d_noise.n < tot && (rc = b->d_noise.alloc(tot))) return rc;
  if ((rc = b->d_noise.upload(noise, tot, b->s))) return rc;
  saved.resize((size_t)b->npoly * kPI64);
  if ((rc = b->d_pi64.download(saved.data(), saved.size(), b->s))) return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  std::vector<int64_t> pi = saved;
  for (int k = 0; k < b->npoly; ++k) {
    pi[(size_t)k * kPI64 + PI_NOISE_OFF] = (int64_t)b->order[k] * cap;
    pi[(size_t)k * kPI64 + PI_CURSOR] = 0;
  }
  return b->d_pi64.upload(pi.data(), pi.size(), b->s);
}

static int noise_end(rp_batch* b, const std::vector<int64_t>& saved) {
  std::vector<int64_t> pi(saved.size());
  int rc;
  if ((rc = b->d_pi64.download(pi.data(), pi.size(), b->s))) return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  for (int k = 0; k < b->npoly; ++k)
    pi[(size_t)k * kPI64 + PI_CURSOR] = saved[(size_t)k * kPI64 + PI_CURSOR];
  return b->d_pi64.upload(pi.data(), pi.size(), b->s);
}

// init_path (verletmodule.f90:32-119) of every polymer of a family, one workgroup each:
// NM momenta sqrt(m'/betan) z with the draws in (dim outer, atom inner) blocks of N
// (:102-106), p = T v (:107-115), and the beads spread along the path spline,
// x_k = splint(path, (k-1) xi/(N-1)) (:40-48).  It runs once per task, so a plain dense
// product serves; the step kernel's launches then all have the same shape.
__global__ void __launch_bounds__(256) batch_init_kernel(BatchArgs A) {
  extern __shared__ double V[];
  const int poly = A.first + blockIdx.x, tid = threadIdx.x;
  const int N = A.N, NPAD = A.NPAD, D = A.D, ndim = A.ndim, natom = A.natom;
  int64_t* pi = A.pi64 + (size_t)poly * kPI64;
  double* st = A.state + pi[PI_STATE_OFF];
  const int64_t noff = pi[PI_NOISE_OFF], cur = pi[PI_CURSOR], need = (int64_t)N * D;
  const uint64_t strm = A.stream[poly];
  if (A.noise && cur + need > A.noise_cap) {  // the same for every thread of the block
    if (tid == 0) atomicOr(&A.flags[1], 1);
    return;
  }
  for (int e = tid; e < N * D; e += blockDim.x) {
    const int c = e / N, i = e - c * N;
    const int atom = c / ndim, dim = c - atom * ndim;
    const int64_t idx = cur + (int64_t)(dim * natom + atom) * N + i;
    const double z = A.noise ? A.noise[noff + idx] : philox_normal(A.seed, strm, idx);
    V[e] = (A.stdev0 * z) * (sqrt(A.massd[c]) * A.modes[i * kMODE + MD_LAM] * A.tau);
  }
  __syncthreads();
  const double xi = A.pf64[(size_t)poly * kPF64 + PF_XI];
  const double* lampath = A.pathd;
  for (int e = tid; e < N * D; e += blockDim.x) {
    const int c = e / N, i = e - c * N;
    double acc = 0.0;
    for (int k = 0; k < N; ++k) acc += A.T[(size_t)k * NPAD + i] * V[c * N + k];
    st[(size_t)D * N + e] = acc;
    st[e] = splint_dev(lampath, lampath + A.npath + (size_t)c * A.npath,
                       lampath + A.npath * (1 + D) + (size_t)c * A.npath, A.npath,
                       (double)i * xi / (double)(N - 1));
  }
  if (tid == 0) pi[PI_CURSOR] = cur + need;
}

static int run_init(rp_batch* b, const double* dn, int64_t cap) {
  for (size_t fi = 0; fi < b->fam.size(); ++fi) {
    Family& f = b->fam[fi];
    BatchArgs a = base_args(b, f, 0.5 * b->dt);
    a.nsteps = 0;
    a.noise = dn;
    a.noise_cap = cap;
    const size_t lds = (size_t)f.N * b->D * sizeof(double);
    RP_REQUIRE(lds <= 160 * 1024, "rp_batch_init_path: N*ndof = %d is too large",
               f.N * b->D);
    RP_HIP(hipFuncSetAttribute((const void*)batch_init_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    batch_init_kernel<<<f.count, 256, lds, b->s>>>(a);
    RP_HIP(hipGetLastError());
  }
  RP_HIP(hipStreamSynchronize(b->s));
  return RP_OK;
}

__global__ void batch_advance_kernel(int64_t* step0, int steps) {
  if (threadIdx.x == 0) *step0 += steps;
}

static rp_batch::GraphKey graph_key(const rp_batch* b, size_t fi, int steps, int64_t imin) {
  rp_batch::GraphKey k;
  k.fam = (int)fi;
  k.steps = steps;
  k.imin = imin;
  k.kick_cap = b->kick_cap;
  k.kicks = b->d_kicks.p;
  k.pathd = b->d_pathd.p;
  k.npath = b->npath;
  for (const Family& f : b->fam) k.waves.push_back(f.waves);
  return k;
}

// Capture family fi's launch of one step block (and, for the last family, the device step
// counter advance).  The stream leaves capture mode on every path.
static int graph_build(rp_batch* b, size_t fi, int steps, int64_t imin, hipGraphExec_t* out) {
  int rc;
  if (!b->d_step0.p && (rc = b->d_step0.alloc(1))) return rc;
  Family& f = b->fam[fi];
  BatchArgs a = base_args(b, f, 0.5 * b->dt);
  a.nsteps = steps;
  a.step0 = 0;
  a.step0p = b->d_step0.p;
  a.imin = imin;
  a.kicks = b->d_kicks.p;
  a.kick_cap = b->kick_cap;
  RP_HIP(hipStreamBeginCapture(b->s, hipStreamCaptureModeRelaxed));
  rc = launch_family(b, a, f.count);
  if (rc == RP_OK && fi + 1 == b->fam.size()) {
    batch_advance_kernel<<<1, 64, 0, b->s>>>(b->d_step0.p, steps);
    if (hipGetLastError() != hipSuccess)
      rc = set_error(RP_ERR_HIP, "batch graph: capture of the step counter failed");
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->s, &g);
  if (rc == RP_OK && e != hipSuccess)
    rc = set_error(RP_ERR_HIP, "batch graph: capture failed (%s)", hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  if (rc == RP_OK) {
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess)
      rc = set_error(RP_ERR_HIP, "batch graph: instantiate failed (%s)", hipGetErrorString(ei));
  }
  if (g) (void)hipGraphDestroy(g);
  if (rc != RP_OK) return rc;
  if (b->graphs.size() >= 64) {
    (void)hipGraphExecDestroy(b->graphs.front().second);
    b->graphs.erase(b->graphs.begin());
  }
  b->graphs.emplace_back(graph_key(b, fi, steps, imin), exec);
  *out = exec;
  return RP_OK;
}

// One launch per family of `steps` steps, each bracketed by its own events; per-family
// durations are accumulated into fam_ms.  With the engine's own streams (no injected draws)
// the block is a replay of its captured graph.
static int run_block(rp_batch* b, int steps, int64_t imin, const double* dn, int64_t cap,
                     std::vector<double>& fam_ms) {
  int rc;
  if (b->use_graph && !dn) {
    std::vector<hipGraphExec_t> exec(b->fam.size(), nullptr);
    for (size_t fi = 0; fi < b->fam.size(); ++fi) {
      const rp_batch::GraphKey key = graph_key(b, fi, steps, imin);
      for (auto& g : b->graphs)
        if (g.first == key) exec[fi] = g.second;
      if (!exec[fi] && (rc = graph_build(b, fi, steps, imin, &exec[fi]))) return rc;
    }
    RP_HIP(hipMemcpyAsync(b->d_step0.p, &b->steps_done, sizeof(int64_t), hipMemcpyHostToDevice, b->s));
    for (size_t fi = 0; fi < b->fam.size(); ++fi) {
      RP_HIP(hipEventRecord(b->fev[2 * fi], b->s));
      RP_HIP(hipGraphLaunch(exec[fi], b->s));
      RP_HIP(hipEventRecord(b->fev[2 * fi + 1], b->s));
    }
  } else {
    for (size_t fi = 0; fi < b->fam.size(); ++fi) {
    Family& f = b->fam[fi];
    BatchArgs a = base_args(b, f, 0.5 * b->dt);
    a.nsteps = steps;
    a.step0 = b->steps_done;
    a.imin = imin;
    a.noise = dn;
    a.noise_cap = cap;
    a.kicks = b->d_kicks.p;
    a.kick_cap = b->kick_cap;
    RP_HIP(hipEventRecord(b->fev[2 * fi], b->s));
    if ((rc = launch_family(b, a, f.count))) return rc;
    RP_HIP(hipEventRecord(b->fev[2 * fi + 1], b->s));
    }
  }
  RP_HIP(hipEventSynchronize(b->fev[2 * b->fam.size() - 1]));
  fam_ms.resize(b->fam.size(), 0.0);
  for (size_t fi = 0; fi < b->fam.size(); ++fi) {
    float ms = 0.f;
    RP_HIP(hipEventElapsedTime(&ms, b->fev[2 * fi], b->fev[2 * fi + 1]));
    fam_ms[fi] += ms;
  }
  return RP_OK;
}

extern "C" int rp_batch_init_path(rp_batch* b, const double* noise) {
  RP_REQUIRE(b, "rp_batch_init_path: NULL batch");
  RP_REQUIRE(b->npath >= 2, "rp_batch_init_path: no path set (rp_batch_set_path)");
  int rc;
  std::vector<int64_t> saved;
  int64_t cap = 0;
  if (noise) {
    // polymer p's n*ndof draws follow each other in user order; repack at a uniform stride
    for (int k = 0; k < b->npoly; ++k) cap = std::max<int64_t>(cap, (int64_t)b->N[k] * b->D);
    std::vector<double> packed((size_t)b->npoly * cap, 0.0);
    int64_t src = 0;
    for (int p = 0; p < b->npoly; ++p) {
      const int64_t cnt = (int64_t)b->N[b->pos[p]] * b->D;
      memcpy(packed.data() + (size_t)p * cap, noise + src, cnt * sizeof(double));
      src += cnt;
    }
    if ((rc = noise_begin(b, packed.data(), cap, saved))) return rc;
  }
  if ((rc = run_init(b, noise ? b->d_noise.p : nullptr, cap))) return rc;
  if (noise && (rc = noise_end(b, saved))) return rc;
  return batch_flags(b);
}

extern "C" int rp_batch_set_state(rp_batch* b, int poly, const double* x, const double* p) {
  RP_REQUIRE(b && x && p && poly >= 0 && poly < b->npoly, "rp_batch_set_state: bad arguments");
  const int k = b->pos[poly];
  const size_t nd = (size_t)b->N[k] * b->D;
  RP_HIP(hipMemcpyAsync(b->d_state.p + b->soff[k], x, nd * sizeof(double), hipMemcpyHostToDevice, b->s));
  RP_HIP(hipMemcpyAsync(b->d_state.p + b->soff[k] + nd, p, nd * sizeof(double), hipMemcpyHostToDevice, b->s));
  RP_HIP(hipStreamSynchronize(b->s));
  return RP_OK;
}

extern "C" int rp_batch_get_state(rp_batch* b, int poly, double* x, double* p) {
  RP_REQUIRE(b && poly >= 0 && poly < b->npoly, "rp_batch_get_state: bad arguments");
  const int k = b->pos[poly];
  const size_t nd = (size_t)b->N[k] * b->D;
  if (x) RP_HIP(hipMemcpyAsync(x, b->d_state.p + b->soff[k], nd * sizeof(double), hipMemcpyDeviceToHost, b->s));
  if (p) RP_HIP(hipMemcpyAsync(p, b->d_state.p + b->soff[k] + nd, nd * sizeof(double), hipMemcpyDeviceToHost, b->s));
  RP_HIP(hipStreamSynchronize(b->s));
  return RP_OK;
}

extern "C" int rp_batch_set_kicks(rp_batch* b, const int32_t* kicks, int64_t nkicks_per_poly) {
  RP_REQUIRE(b && (kicks == nullptr || nkicks_per_poly > 0), "rp_batch_set_kicks: bad arguments");
  int rc;
  if (!kicks) {
    b->d_kicks.free();
    b->kick_cap = 0;
    return RP_OK;
  }
  const size_t tot = (size_t)b->npoly * nkicks_per_poly;
  if ((rc = b->d_kicks.alloc(tot)) || (rc = b->d_kicks.upload(kicks, tot, b->s))) return rc;
  std::vector<int64_t> pi((size_t)b->npoly * kPI64);
  if ((rc = b->d_pi64.download(pi.data(), pi.size(), b->s))) return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  for (int k = 0; k < b->npoly; ++k) {
    pi[(size_t)k * kPI64 + PI_KICK_OFF] = (int64_t)b->order[k] * nkicks_per_poly;
    pi[(size_t)k * kPI64 + PI_KCURSOR] = 0;
    pi[(size_t)k * kPI64 + PI_RKICK] = -1;
    pi[(size_t)k * kPI64 + PI_ACOUNT] = 0;
  }
  b->kick_cap = nkicks_per_poly;
  if ((rc = b->d_pi64.upload(pi.data(), pi.size(), b->s))) return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  return RP_OK;
}

extern "C" int rp_batch_run(rp_batch* b, int64_t nsteps, int64_t imin, const double* noise,
                            int64_t nnoise_per_poly) {
  RP_REQUIRE(b && nsteps >= 0 && imin >= 0, "rp_batch_run: bad arguments");
  RP_REQUIRE(!noise || nnoise_per_poly > 0, "rp_batch_run: injected noise needs nnoise_per_poly > 0");
  RP_REQUIRE(!(b->thermostat == RP_THERMO_LANGEVIN && b->limit >= 0.0) || b->npath >= 2,
             "rp_batch_run: dHdrlimit >= 0 re-initialises from the path: call rp_batch_set_path");
  int rc;
  std::vector<int64_t> saved;
  if (noise && (rc = noise_begin(b, noise, nnoise_per_poly, saved))) return rc;
  std::vector<double> fam_ms(b->fam.size(), 0.0);
  int64_t launches = 0, left = nsteps;
  while (left > 0) {
    const int steps = (int)std::min<int64_t>(left, b->steps_per_launch);
    if ((rc = run_block(b, steps, imin, noise ? b->d_noise.p : nullptr, nnoise_per_poly,
                        fam_ms)))
      return rc;
    ++launches;
    b->steps_done += steps;
    left -= steps;
  }
  b->fam_ms = fam_ms;
  b->last_launches = launches;
  if (noise && (rc = noise_end(b, saved))) return rc;
  return batch_flags(b);
}

extern "C" int rp_batch_ti_readout(rp_batch* b, double* dhdr_sum, int64_t* count,
                                   int64_t* nreinit) {
  RP_REQUIRE(b, "rp_batch_ti_readout: NULL batch");
  std::vector<int64_t> pi((size_t)b->npoly * kPI64);
  std::vector<double> pf((size_t)b->npoly * kPF64);
  int rc;
  if ((rc = b->d_pi64.download(pi.data(), pi.size(), b->s)) ||
      (rc = b->d_pf64.download(pf.data(), pf.size(), b->s)))
    return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  for (int p = 0; p < b->npoly; ++p) {
    const int k = b->pos[p];
    if (dhdr_sum) dhdr_sum[p] = pf[(size_t)k * kPF64 + PF_TISUM];
    if (count) count[p] = pi[(size_t)k * kPI64 + PI_TICNT];
    if (nreinit) nreinit[p] = pi[(size_t)k * kPI64 + PI_NREINIT];
  }
  return RP_OK;
}

extern "C" int rp_batch_get_cursor(rp_batch* b, int64_t* cursor) {
  RP_REQUIRE(b && cursor, "rp_batch_get_cursor: bad arguments");
  std::vector<int64_t> pi((size_t)b->npoly * kPI64);
  int rc;
  if ((rc = b->d_pi64.download(pi.data(), pi.size(), b->s))) return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  for (int p = 0; p < b->npoly; ++p) cursor[p] = pi[(size_t)b->pos[p] * kPI64 + PI_CURSOR];
  return RP_OK;
}

extern "C" int rp_batch_get_kick_cursor(rp_batch* b, int64_t* kick_cursor) {
  RP_REQUIRE(b && kick_cursor, "rp_batch_get_kick_cursor: bad arguments");
  std::vector<int64_t> pi((size_t)b->npoly * kPI64);
  int rc;
  if ((rc = b->d_pi64.download(pi.data(), pi.size(), b->s))) return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  for (int p = 0; p < b->npoly; ++p)
    kick_cursor[p] = pi[(size_t)b->pos[p] * kPI64 + PI_KCURSOR];
  return RP_OK;
}

// Resume: place every polymer's normal and Poisson draws at saved stream positions.  A new
// Poisson cursor also discards the interval in progress, so the next step draws a fresh one,
// as propagate_pimd_nm does on entry.
extern "C" int rp_batch_set_cursor(rp_batch* b, const int64_t* cursor, const int64_t* kick_cursor) {
  RP_REQUIRE(b, "rp_batch_set_cursor: NULL batch");
  for (int p = 0; p < b->npoly; ++p)
    RP_REQUIRE((!cursor || cursor[p] >= 0) && (!kick_cursor || kick_cursor[p] >= 0),
               "rp_batch_set_cursor: negative cursor");
  std::vector<int64_t> pi((size_t)b->npoly * kPI64);
  int rc;
  if ((rc = b->d_pi64.download(pi.data(), pi.size(), b->s))) return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  for (int p = 0; p < b->npoly; ++p) {
    int64_t* q = pi.data() + (size_t)b->pos[p] * kPI64;
    if (cursor) q[PI_CURSOR] = cursor[p];
    if (kick_cursor) {
      q[PI_KCURSOR] = kick_cursor[p];
      q[PI_RKICK] = -1;
      q[PI_ACOUNT] = 0;
    }
  }
  if ((rc = b->d_pi64.upload(pi.data(), pi.size(), b->s))) return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  return RP_OK;
}

extern "C" int rp_batch_set_graphs(rp_batch* b, int enable) {
  RP_REQUIRE(b, "rp_batch_set_graphs: NULL batch");
  b->use_graph = enable ? 1 : 0;
  if (!enable) graph_drop(b);
  return RP_OK;
}

// 8 waves per workgroup needs at most 16 MFMA units (2 per wave)
static bool eight_waves_fit(const rp_batch* b, const Family& f) {
  return (f.NPAD >> 5) * b->DT <= 16;
}

// 2 waves run a family only when each wave holds one of at most two units (W is a multiple
// of DT); otherwise such a request falls back to 4
static bool two_waves_fit(const rp_batch* b, const Family& f) {
  return b->DT <= 2 && (f.NPAD >> 5) * b->DT <= 2;
}

static int fitted_waves(const rp_batch* b, const Family& f, int w) {
  if (w == 8) return eight_waves_fit(b, f) ? 8 : 4;
  if (w == 2) return two_waves_fit(b, f) ? 2 : 4;
  return 4;
}

extern "C" int rp_batch_set_launch(rp_batch* b, int steps_per_launch, int waves) {
  RP_REQUIRE(b && steps_per_launch >= 0, "rp_batch_set_launch: bad arguments");
  RP_REQUIRE(waves == 0 || waves == 2 || waves == 4 || waves == 8,
             "rp_batch_set_launch: waves per workgroup must be 2, 4 or 8");
  if (steps_per_launch > 0) b->steps_per_launch = steps_per_launch;
  if (waves)
    for (Family& f : b->fam) f.waves = fitted_waves(b, f, waves);
  return RP_OK;
}

extern "C" int rp_batch_autotune(rp_batch* b, int trial_steps, int* best_steps, int* best_waves) {
  RP_REQUIRE(b && trial_steps > 0, "rp_batch_autotune: bad arguments");
  int rc;
  // save the whole mutable state, time each candidate, restore
  const size_t ns = b->d_state.n;
  std::vector<double> st(ns), pf((size_t)b->npoly * kPF64);
  std::vector<int64_t> pi((size_t)b->npoly * kPI64);
  if ((rc = b->d_state.download(st.data(), ns, b->s)) || (rc = b->d_pf64.download(pf.data(), pf.size(), b->s)) ||
      (rc = b->d_pi64.download(pi.data(), pi.size(), b->s)))
    return rc;
  RP_HIP(hipStreamSynchronize(b->s));
  auto restore = [&]() -> int {
    int r;
    if ((r = b->d_state.upload(st.data(), ns, b->s)) || (r = b->d_pf64.upload(pf.data(), pf.size(), b->s)) ||
        (r = b->d_pi64.upload(pi.data(), pi.size(), b->s)))
      return r;
    RP_HIP(hipStreamSynchronize(b->s));
    return RP_OK;
  };
  auto trial = [&](int cs, std::vector<double>& fam_ms, double& sec) -> int {
    fam_ms.assign(b->fam.size(), 0.0);
    auto t0 = std::chrono::steady_clock::now();
    int64_t left = trial_steps;
    while (left > 0) {
      const int steps = (int)std::min<int64_t>(left, cs);
      int r;
      if ((r = run_block(b, steps, INT64_MAX / 2, nullptr, 0, fam_ms))) return r;
      left -= steps;
    }
    sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return restore();
  };
  const int64_t steps0 = b->steps_done;
  // 1. waves per family, from each family's own hipEvent-timed launches
  const size_t nf = b->fam.size();
  std::vector<double> best_ms(nf, 1e300), fam_ms;
  std::vector<int> best_w(nf, 4);
  double sec = 0.0;
  for (int w : {2, 4, 8}) {
    for (Family& f : b->fam) f.waves = fitted_waves(b, f, w);
    if ((rc = trial(std::min(b->steps_per_launch, trial_steps), fam_ms, sec))) return rc;
    for (size_t fi = 0; fi < nf; ++fi)
      if (b->fam[fi].waves == w && fam_ms[fi] < best_ms[fi]) {
        best_ms[fi] = fam_ms[fi];
        best_w[fi] = w;
      }
  }
  int nw[9] = {0};
  for (size_t fi = 0; fi < nf; ++fi) {
    b->fam[fi].waves = best_w[fi];
    ++nw[best_w[fi]];
  }
  // 2. steps per launch
  const int cands[] = {8, 32, 128};
  double best = 1e300;
  int bs = b->steps_per_launch;
  for (int cs : cands) {
    if (cs > trial_steps) continue;
    if ((rc = trial(cs, fam_ms, sec))) return rc;
    if (sec < best) {
      best = sec;
      bs = cs;
    }
  }
  b->steps_done = steps0;
  b->steps_per_launch = bs;
  if (best_steps) *best_steps = bs;
  if (best_waves)  // the most frequent choice (ties to the wider workgroup)
    *best_waves = (nw[8] >= nw[4] && nw[8] >= nw[2]) ? 8 : (nw[4] >= nw[2] ? 4 : 2);
  return batch_flags(b);
}

extern "C" int rp_batch_last_kernel_ms(rp_batch* b, double* avg_ms, int64_t* launches) {
  RP_REQUIRE(b && avg_ms, "rp_batch_last_kernel_ms: bad arguments");
  // per family: average duration of one launch of the fused kernel in the last run
  for (size_t fi = 0; fi < b->fam.size(); ++fi)
    avg_ms[fi] = b->last_launches ? b->fam_ms[fi] / b->last_launches : 0.0;
  if (launches) *launches = b->last_launches;
  return RP_OK;
}

extern "C" int rp_batch_family_shape(rp_batch* b, int fam, int* waves, int* units_per_wave,
                                     int* folded) {
  RP_REQUIRE(b && fam >= 0 && fam < (int)b->fam.size(), "rp_batch_family_shape: bad arguments");
  const Family& f = b->fam[fam];
  // the kernel instance launch_batch_thermo picks: units of 32 modes x 16 dofs (folded: 16
  // rows of each half x 16 dofs; the same count), dealt over the waves of a workgroup
  const int units = (f.NPAD >> 5) * b->DT;
  if (waves) *waves = f.waves;
  if (units_per_wave) *units_per_wave = (units + f.waves - 1) / f.waves;
  if (folded) *folded = (f.N % 2 == 0) && f.d_Tf.p != nullptr && b->pes->kind != PES_WATER;
  return RP_OK;
}

extern "C" int rp_batch_families(rp_batch* b, int* nfam, int* n, int* count) {
  RP_REQUIRE(b && nfam, "rp_batch_families: bad arguments");
  *nfam = (int)b->fam.size();
  for (size_t fi = 0; fi < b->fam.size(); ++fi) {
    if (n) n[fi] = b->fam[fi].N;
    if (count) count[fi] = b->fam[fi].count;
  }
  return RP_OK;
}
