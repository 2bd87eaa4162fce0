// Host side of the C ABI (include/mose_gpu.h): contexts, cases, the step sequencer and the launches.
//
// The sequencer is Explicit_Step (src/lib/numerics/time/explicit/Mod_Explicit.f90:52-397) for a time-accurate
// step without chemistry: dt of the state at the start of the step, then n_RK stages.  The stage buffers are
// rotated instead of copied (field_copy_P_flat_to_PO_flat, MOSE_Field_m.f90:321-330, is not needed): stage 1
// reads the step base P0 and writes PA, stage 2 reads PA and writes PB, the last stage reads its input and
// writes the new state over P0 (a cell-local read-then-write of the base).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mose_gpu.h"
#include "mose_internal.h"
#include "mose_chem.hpp"
#include "mose_kernels.hpp"
#include "mose_radau.hpp"
#include "mose_stage.hpp"

using namespace mose;

namespace {

struct Case {
  mose_scheme sc{};
  int nb = 0, nprim = 0;
  std::vector<mose_blk> blk;
  std::vector<mose_bc_record> rec;
  std::vector<mose_farfield> ff;
  std::vector<signed char> btyp;   // record type per boundary-plane position
  int64_t NC = 0, NF1 = 0, NF2 = 0, NF3 = 0, NFI[3] = {0, 0, 0}, NBP = 0, ninterior = 0;
  // device memory (one slab)
  void* slab = nullptr;
  size_t slab_bytes = 0;
  double *P0 = nullptr, *PA = nullptr, *PB = nullptr;
  double *dl = nullptr, *m = nullptr, *vol = nullptr, *yn = nullptr, *fm1 = nullptr, *fm2 = nullptr, *fm3 = nullptr;
  double *Fc[3] = {nullptr, nullptr, nullptr}, *Fd[3] = {nullptr, nullptr, nullptr};
  double* binc = nullptr;
  // implicit residual smoothing (scheme irs = 1): the residual and the two Jacobi buffers (flow variables)
  double *dR = nullptr, *drs1 = nullptr, *drs2 = nullptr;
  mose_blk* dblk = nullptr;
  mose_bc_record* drec = nullptr;
  mose_farfield* dff = nullptr;
  signed char* dbtyp = nullptr;
  double* dt = nullptr;          // [1] current dt
  double* dthist = nullptr;      // [hist_cap]
  int hist_cap = 0;
  unsigned long long* dtmin = nullptr;
  unsigned int* bad = nullptr;   // [2]
  unsigned long long* first_bad = nullptr;
  double time = 0.0, last_dt = 0.0;
  int64_t steps = 0;
  bool valid = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // fused path
  int path = MOSE_PATH_FUSED;
  int TI = 32, TJ = 4, TK = 2;
  Tile* dtiles = nullptr;
  int ntiles = 0, tiles_cap = 0;
  unsigned long long* dtnext = nullptr;
  bool dt_valid = false;
  // residual norm of the last step (per-workgroup partials of the last stage) and the state hash slots
  double* rpart = nullptr;
  int64_t rpart_cap = 0, nparts_last = 0;
  double* rnorm = nullptr;
  unsigned long long* hslot = nullptr;
  hipStream_t stream = nullptr;   // the case's own stream: several cases' steps overlap on the device
  // pinned staging of the step report (dt history, bad-cell counters, first bad cell)
  double* stage_hist = nullptr;
  unsigned* stage_bad = nullptr;
  unsigned long long* stage_fb = nullptr;
  void* stage_host = nullptr;
  int pending = 0;
  bool launched = false;          // a mose_case_launch awaits its mose_case_poll
  // one fused RK step captured as a hipGraph (kernels only), replayed per step; off with mose_tuning.no_graph, or
  // for every case of the process with MOSE_NO_GRAPH=1 in the environment (read at case creation): a profiler that
  // intercepts graph launches then sees plain kernel launches
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int* histpos = nullptr;     // device: next slot of the dt history
  // chemistry of Strang-split steps (mose_case_chem_attach): n_strang = 2, strang = 1/2 with a mechanism attached
  int n_strang = 1;
  double strang = 1.0;
  int chem_id = -1;               // the mechanism's chemistry object
  int chem_method = 0;
  ChemArgs chem{};                // tolerances, constants, rows, outputs (state and dt set per launch)
  long long* chem_rows = nullptr; // device: the arena row of each interior cell (one allocation with the outputs)
  unsigned long long* chem_ctr = nullptr;   // device: [0] queue of a launch, [8..13] statistics of a step call
  unsigned long long* stage_chem = nullptr; // pinned: the statistics of the last step call
  // timing
  bool timing = false;
  std::vector<hipEvent_t> tev;
  double ms_stage_total = 0.0;
  int64_t stage_launches = 0;
};

}  // namespace

struct mose_ctx {
  int device = 0;
  int max_lds = 65536;       // dynamic LDS a workgroup may take on this device
  hipStream_t stream = nullptr;
  int nsc = 0;
  bool viscous_tables = false;
  double *dcp = nullptr, *dh = nullptr, *dmi = nullptr, *dk = nullptr;
  mose_dtab htab{};
  mose_dtab* dtab = nullptr;
  std::vector<std::unique_ptr<Case>> cases;
  std::string err;
};

#define HIPCHK(ctx, call)                                                             \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                 \
      return MOSE_ERR_HIP;                                                            \
    }                                                                                 \
  } while (0)

// The scratch of mose_chem_advance_cells, the entry point that allocates nothing: the chemistry kernels' work queue
// and batch statistics ([0] the next cell, [1..6] ChemArgs::stats).  Each GPU the library runs on holds its own copy
// of these 64 bytes until the process ends.  On the host the array's name is not a device address: mose_ctx_create
// looks up the address once per device (hipGetSymbolAddress, with that device current) into g_scratch, and every use
// goes through g_scratch.  Calls of mose_chem_advance_cells that overlap on one GPU share it and race.
constexpr int kScratchWords = 8;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];
static std::mutex g_scratch_mu;
static unsigned long long* g_scratch[kMaxDevices] = {};

__global__ void k_zero_u64(unsigned long long* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

static int fail(mose_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

extern "C" int mose_device_count(int* n) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  if (n) *n = c;
  return MOSE_OK;
}

extern "C" int mose_ctx_info(mose_ctx* ctx, int64_t* total, int64_t* freeb, int* cus, int* lds) {
  if (!ctx) return MOSE_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  size_t f = 0, t = 0;
  HIPCHK(ctx, hipMemGetInfo(&f, &t));
  int n = 0;
  HIPCHK(ctx, hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, ctx->device));
  if (total) *total = (int64_t)t;
  if (freeb) *freeb = (int64_t)f;
  if (cus) *cus = n;
  if (lds) *lds = ctx->max_lds;
  return MOSE_OK;
}

extern "C" const char* mose_last_error(const mose_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int mose_ctx_create(int device, const mose_tables* t, mose_ctx** out) {
  if (!out || !t || !t->wm || !t->cp || !t->h || t->nsc < 1 || t->nsc > MOSE_MAX_SPECIES) return MOSE_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MOSE_ERR_NODEV;
  if (device < 0 || device >= ndev) return MOSE_ERR_ARG;
  auto ctx = std::make_unique<mose_ctx>();
  ctx->device = device;
  ctx->nsc = t->nsc;
  HIPCHK(ctx.get(), hipSetDevice(device));
  if (device < kMaxDevices) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
      unsigned long long* p = nullptr;
      HIPCHK(ctx.get(), hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
      if (!p) return fail(ctx.get(), MOSE_ERR_HIP, "hipGetSymbolAddress(scratch) gave no address");
      g_scratch[device] = p;
    }
  }
  HIPCHK(ctx.get(), hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
      ctx->max_lds = v;
  }
  size_t n = (size_t)t->nsc * MOSE_NTAB;
  HIPCHK(ctx.get(), hipMalloc(&ctx->dcp, 4 * n * sizeof(double)));
  ctx->dh = ctx->dcp + n;
  ctx->dmi = ctx->dh + n;
  ctx->dk = ctx->dmi + n;
  HIPCHK(ctx.get(), hipMemcpy(ctx->dcp, t->cp, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(ctx.get(), hipMemcpy(ctx->dh, t->h, n * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> zero(n, 0.0);
  HIPCHK(ctx.get(), hipMemcpy(ctx->dmi, t->mu ? t->mu : zero.data(), n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(ctx.get(), hipMemcpy(ctx->dk, t->k ? t->k : zero.data(), n * sizeof(double), hipMemcpyHostToDevice));
  ctx->viscous_tables = (t->mu != nullptr && t->k != nullptr);
  mose_dtab& T = ctx->htab;
  T.cp = ctx->dcp; T.h = ctx->dh; T.mi = ctx->dmi; T.k = ctx->dk;
  // Ri_tab = Runiv / wm_tab and the Wilke pair constants (Load_ThermoTransport.f90:51-67)
  for (int s = 0; s < t->nsc; ++s) { T.wm[s] = t->wm[s]; T.Ri[s] = MOSE_RUNIV / t->wm[s]; }
  for (int sj = 0; sj < t->nsc; ++sj)
    for (int si = 0; si < t->nsc; ++si) {
      double ratio = t->wm[si] / t->wm[sj];
      T.mm025[si][sj] = std::pow(ratio, -0.25);
      T.is8[si][sj] = 1.0 / std::sqrt(8.0 * (1.0 + ratio));
    }
  HIPCHK(ctx.get(), hipMalloc(&ctx->dtab, sizeof(mose_dtab)));
  HIPCHK(ctx.get(), hipMemcpy(ctx->dtab, &T, sizeof(mose_dtab), hipMemcpyHostToDevice));
  *out = ctx.release();
  return MOSE_OK;
}

static void case_free(Case* c) {
  if (!c) return;
  if (c->stream) { (void)hipStreamSynchronize(c->stream); (void)hipStreamDestroy(c->stream); }
  c->stream = nullptr;
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  if (c->stage_host) (void)hipHostFree(c->stage_host);
  c->stage_host = nullptr;
  if (c->slab) (void)hipFree(c->slab);
  if (c->rpart) (void)hipFree(c->rpart);
  c->rpart = nullptr;
  c->rpart_cap = 0;
  if (c->dtiles) (void)hipFree(c->dtiles);
  c->dtiles = nullptr;
  if (c->chem_rows) (void)hipFree(c->chem_rows);
  c->chem_rows = nullptr;
  if (c->stage_chem) (void)hipHostFree(c->stage_chem);
  c->stage_chem = nullptr;
  for (auto& e : c->tev)
    if (e) (void)hipEventDestroy(e);
  c->tev.clear();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  c->slab = nullptr;
  c->valid = false;
}

extern "C" int mose_ctx_destroy(mose_ctx* ctx) {
  if (!ctx) return MOSE_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& c : ctx->cases) case_free(c.get());
  if (ctx->dcp) (void)hipFree(ctx->dcp);
  if (ctx->dtab) (void)hipFree(ctx->dtab);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MOSE_OK;
}

// block table, arena sizes and validation of the scheme / records
static int plan_case(const mose_scheme* sc, int nb, const int* dims, int nrec, const mose_bc_record* rec, int nff,
                     const mose_farfield* ff, Case& c, std::string& why) {
  if (!sc || nb <= 0 || !dims) { why = "bad arguments"; return MOSE_ERR_ARG; }
  if (sc->nsc < 1 || sc->nsc > MOSE_MAX_SPECIES) { why = "nsc out of range"; return MOSE_ERR_ARG; }
  if (sc->nrans < 0 || sc->nrans > 1) { why = "nrans must be 0 or 1 (Spalart-Allmaras)"; return MOSE_ERR_ARG; }
  if (sc->nrans > 0 && !sc->viscous) { why = "Spalart-Allmaras needs the viscous fluxes"; return MOSE_ERR_ARG; }
  if (sc->limiter < 1 || sc->limiter > 6) { why = "limiter id must be 1..6"; return MOSE_ERR_ARG; }
  if (!(sc->riemann == MOSE_RS_HLLC || sc->riemann == MOSE_RS_HLLE || sc->riemann == MOSE_RS_SLAU ||
        sc->riemann == MOSE_RS_AUSMP || sc->riemann == MOSE_RS_LMROE)) {
    why = "riemann solver must be HLLC (1), HLLE (2), SLAU (3), AUSM+ (4) or LMRoe (12)";
    return MOSE_ERR_ARG;
  }
  if (sc->n_rk < 1 || sc->n_rk > 3) { why = "n_rk must be 1..3"; return MOSE_ERR_ARG; }
  if (sc->ndir < 1 || sc->ndir > 3) { why = "ndir must be 1..3"; return MOSE_ERR_ARG; }
  if (sc->integration != MOSE_INT_CONS && sc->integration != MOSE_INT_PRIM) {
    why = "integration must be MOSE_INT_CONS (0) or MOSE_INT_PRIM (1)";
    return MOSE_ERR_ARG;
  }
  if (sc->irs != 0 && sc->irs != 1) { why = "irs must be 0 or 1"; return MOSE_ERR_ARG; }
  if (sc->irs && (sc->irs_njacobi < 1 || sc->irs_njacobi > 64 || !(sc->irs_beta >= 0.0) || !std::isfinite(sc->irs_beta))) {
    why = "implicit residual smoothing needs 1..64 Jacobi sweeps and a finite irs_beta >= 0";
    return MOSE_ERR_ARG;
  }
  c.sc = *sc;
  c.nb = nb;
  c.nprim = sc->nsc + 4 + sc->nrans;
  c.blk.resize(nb);
  int64_t NC = 0, NF1 = 0, NF2 = 0, NF3 = 0, NFI[3] = {0, 0, 0}, NBP = 0, nin = 0;
  for (int b = 0; b < nb; ++b) {
    int n1 = dims[3 * b], n2 = dims[3 * b + 1], n3 = dims[3 * b + 2];
    if (n1 < 1 || n2 < 1 || n3 < 1) { why = "block dims must be >= 1"; return MOSE_ERR_ARG; }
    mose_blk& B = c.blk[b];
    B.n1 = n1; B.n2 = n2; B.n3 = n3; B.pi = n1 + 2 * MOSE_GC; B.pj = n2 + 2 * MOSE_GC; B.pad = 0;
    B.boff = NC; B.f1off = NF1; B.f2off = NF2; B.f3off = NF3;
    B.ifoff[0] = NFI[0]; B.ifoff[1] = NFI[1]; B.ifoff[2] = NFI[2];
    int64_t pl[6] = {(int64_t)n2 * n3, (int64_t)n2 * n3, (int64_t)n1 * n3, (int64_t)n1 * n3, (int64_t)n1 * n2, (int64_t)n1 * n2};
    for (int f = 0; f < 6; ++f) { B.bpoff[f] = NBP; NBP += pl[f]; }
    B.cell0 = nin;
    NC += (int64_t)B.pi * B.pj * (n3 + 2 * MOSE_GC);
    NF1 += (int64_t)(n1 + 1) * n2 * n3;
    NF2 += (int64_t)n1 * (n2 + 1) * n3;
    NF3 += (int64_t)n1 * n2 * (n3 + 1);
    NFI[0] += (int64_t)(n1 - 1) * n2 * n3;
    NFI[1] += (int64_t)n1 * (n2 - 1) * n3;
    NFI[2] += (int64_t)n1 * n2 * (n3 - 1);
    nin += (int64_t)n1 * n2 * n3;
  }
  c.NC = NC; c.NF1 = NF1; c.NF2 = NF2; c.NF3 = NF3; c.NBP = NBP; c.ninterior = nin;
  for (int d = 0; d < 3; ++d) c.NFI[d] = std::max<int64_t>(NFI[d], 1);
  // records: every boundary cell-face exactly once, grouped by block then face
  std::vector<uint8_t> seen((size_t)NBP, 0);
  int lastb = -1, lastf = 0;
  for (int e = 0; e < nrec; ++e) {
    const mose_bc_record& r = rec[e];
    if (r.b < 0 || r.b >= nb || r.f < 1 || r.f > 6) { why = "record " + std::to_string(e) + ": bad block or face"; return MOSE_ERR_ARG; }
    if (r.b < lastb || (r.b == lastb && r.f < lastf)) { why = "records must be grouped by block, then face 1..6"; return MOSE_ERR_ARG; }
    lastb = r.b; lastf = r.f;
    const mose_blk& B = c.blk[r.b];
    int n[3] = {B.n1, B.n2, B.n3};
    int dir = (r.f + 1) / 2 - 1;
    int idx[3] = {r.i, r.j, r.k};
    for (int a = 0; a < 3; ++a)
      if (idx[a] < 1 || idx[a] > n[a]) { why = "record " + std::to_string(e) + ": cell outside the block"; return MOSE_ERR_ARG; }
    int want = (r.f % 2 == 1) ? 1 : n[dir];
    if (idx[dir] != want) { why = "record " + std::to_string(e) + ": cell is not on face " + std::to_string(r.f); return MOSE_ERR_ARG; }
    switch (r.typ) {
      case MOSE_BC_CONNECTION: {
        if (r.bs < 0 || r.bs >= nb || r.fs < 1 || r.fs > 6) { why = "record " + std::to_string(e) + ": bad donor"; return MOSE_ERR_ARG; }
        const mose_blk& S = c.blk[r.bs];
        int ns_[3] = {S.n1, S.n2, S.n3}, sidx[3] = {r.is, r.js, r.ks};
        for (int a = 0; a < 3; ++a)
          if (sidx[a] < 1 || sidx[a] > ns_[a]) { why = "record " + std::to_string(e) + ": donor outside its block"; return MOSE_ERR_ARG; }
        int sdir = (r.fs + 1) / 2 - 1;
        if (ns_[sdir] < MOSE_GC) { why = "record " + std::to_string(e) + ": donor block thinner than the halo"; return MOSE_ERR_ARG; }
        // transverse donor neighbours must stay inside the padded donor block
        int d1[3] = {0, 0, 0}, d2[3] = {0, 0, 0}, t1 = 0, t2 = 0;
        for (int a = 0; a < 3; ++a) if (a != sdir) { if (t1 == 0) { d1[a] = r.d11; d2[a] = r.d21; t1 = 1; } else { d1[a] = r.d12; d2[a] = r.d22; t2 = 1; } }
        (void)t2;
        for (int a = 0; a < 3; ++a) {
          int lo = sidx[a] - std::abs(d1[a]) - std::abs(d2[a]), hi = sidx[a] + std::abs(d1[a]) + std::abs(d2[a]);
          if (lo < 1 - MOSE_GC || hi > ns_[a] + MOSE_GC) { why = "record " + std::to_string(e) + ": bad transverse map"; return MOSE_ERR_ARG; }
        }
        break;
      }
      case MOSE_BC_NONE:         // no flux; zero-gradient ghosts
      case MOSE_BC_SYMMETRY:     // mirror ghost of the cell; the record reads the cell and its neighbour along the normal
        break;
      case MOSE_BC_EXTRAPOLATION:
        if (n[dir] == 2) { why = "extrapolation faces need a block extent of 1 or >= 3 along the face normal"; return MOSE_ERR_ARG; }
        break;
      case MOSE_BC_WALL_ADIABATIC:
      case MOSE_BC_WALL_ISOTHERMAL:
        // the cubic wall ghost reads three interior layers; the record stencil reads the interior neighbour
        if (n[dir] < 3) { why = "record " + std::to_string(e) + ": a wall needs a block extent >= 3 along its normal"; return MOSE_ERR_ARG; }
        if (r.typ == MOSE_BC_WALL_ISOTHERMAL && !(r.tw > 1.0 && r.tw < 9990.0)) { why = "record " + std::to_string(e) + ": wall temperature outside (1, 9990) K"; return MOSE_ERR_ARG; }
        break;
      case MOSE_BC_FARFIELD: {
        if (n[dir] < 2) { why = "record " + std::to_string(e) + ": a far-field face needs a block extent >= 2 along its normal"; return MOSE_ERR_ARG; }
        if (r.ff < 0 || r.ff >= nff || !ff) { why = "record " + std::to_string(e) + ": far-field row out of range"; return MOSE_ERR_ARG; }
        const mose_farfield& q = ff[r.ff];
        if (q.mach == 1.0) { why = "far-field row " + std::to_string(r.ff) + ": Mach exactly 1 has no closure"; return MOSE_ERR_ARG; }
        double tn = r.tvec[0] * r.tvec[0] + r.tvec[1] * r.tvec[1] + r.tvec[2] * r.tvec[2];
        if (q.mach >= 0.0 && q.mach < 1.0 && !(tn > 0.5)) { why = "record " + std::to_string(e) + ": far-field face tangent not set (mose_records_tangent)"; return MOSE_ERR_ARG; }
        break;
      }
      default:
        why = "record " + std::to_string(e) + ": boundary type " + std::to_string(r.typ) + " is not supported yet";
        return MOSE_ERR_ARG;
    }
    int64_t pos = B.bpoff[r.f - 1] + (r.f <= 2 ? (r.j - 1) + (int64_t)B.n2 * (r.k - 1)
                                     : r.f <= 4 ? (r.i - 1) + (int64_t)B.n1 * (r.k - 1)
                                                : (r.i - 1) + (int64_t)B.n1 * (r.j - 1));
    if (seen[pos]) { why = "record " + std::to_string(e) + ": boundary face listed twice"; return MOSE_ERR_ARG; }
    seen[pos] = 1;
    if ((int64_t)c.btyp.size() < NBP) c.btyp.assign((size_t)NBP, 0);
    c.btyp[pos] = (signed char)r.typ;
  }
  for (int64_t q = 0; q < NBP; ++q)
    if (!seen[q]) { why = "a boundary cell-face has no record (every block face needs one record per cell)"; return MOSE_ERR_ARG; }
  c.rec.assign(rec, rec + nrec);
  if (nff > 0) c.ff.assign(ff, ff + nff);
  if (c.btyp.empty()) c.btyp.assign((size_t)std::max<int64_t>(NBP, 1), 0);
  return MOSE_OK;
}

static size_t case_bytes(const Case& c) {
  size_t d = sizeof(double);
  size_t b = 0;
  b += 3 * (size_t)c.NC * c.nprim * d;                     // P0 PA PB
  b += (size_t)c.NC * (3 + 9 + 1 + 1) * d;                 // dl m vol yn
  b += (size_t)(c.NF1 + c.NF2 + c.NF3) * 4 * d;            // fm1..3
  for (int q = 0; q < 3; ++q) b += 2 * (size_t)c.NFI[q] * c.nprim * d;   // face buffers
  b += 2 * (size_t)std::max<int64_t>(c.NBP, 1) * c.nprim * d;           // boundary increments
  if (c.sc.irs) b += (size_t)c.NC * (c.nprim + 2 * (c.sc.nsc + 4)) * d + 3 * 256;   // IRS residual and sweeps
  b += c.blk.size() * sizeof(mose_blk) + c.rec.size() * sizeof(mose_bc_record) + 8192 + 64 * 1024;
  b += (std::max<size_t>(c.ff.size(), 1) * sizeof(mose_farfield) + 255) / 256 * 256;
  b += ((size_t)std::max<int64_t>(c.NBP, 1) + 255) / 256 * 256;
  b += 64 * sizeof(double) + 256 + 512;
  return b;
}

extern "C" int mose_case_bytes(const mose_scheme* sc, int nb, const int* dims, int nrec, int64_t* bytes) {
  Case c;
  std::string why;
  if (!bytes) return MOSE_ERR_ARG;
  // sizing does not need the records themselves
  if (!sc || nb <= 0 || !dims) return MOSE_ERR_ARG;
  c.sc = *sc;
  c.nprim = sc->nsc + 4 + sc->nrans;
  int64_t NC = 0, NF = 0, NFI[3] = {0, 0, 0}, NBP = 0;
  for (int b = 0; b < nb; ++b) {
    int64_t n1 = dims[3 * b], n2 = dims[3 * b + 1], n3 = dims[3 * b + 2];
    NC += (n1 + 4) * (n2 + 4) * (n3 + 4);
    NF += (n1 + 1) * n2 * n3 + n1 * (n2 + 1) * n3 + n1 * n2 * (n3 + 1);
    NFI[0] += (n1 - 1) * n2 * n3; NFI[1] += n1 * (n2 - 1) * n3; NFI[2] += n1 * n2 * (n3 - 1);
    NBP += 2 * (n2 * n3 + n1 * n3 + n1 * n2);
  }
  c.NC = NC; c.NF1 = NF; c.NF2 = 0; c.NF3 = 0; c.NBP = NBP;
  for (int d = 0; d < 3; ++d) c.NFI[d] = std::max<int64_t>(NFI[d], 1);
  c.blk.resize(nb);
  c.rec.resize(std::max(nrec, 0));
  *bytes = (int64_t)case_bytes(c);
  return MOSE_OK;
}

static int alloc_case(mose_ctx* ctx, Case& c) {
  size_t d = sizeof(double);
  c.hist_cap = 4096;
  size_t total = case_bytes(c) + (size_t)c.hist_cap * d;
  HIPCHK(ctx, hipMalloc(&c.slab, total));
  c.slab_bytes = total;
  char* p = (char*)c.slab;
  auto take = [&](size_t nbytes) {
    char* q = p;
    p += (nbytes + 255) & ~(size_t)255;
    return q;
  };
  size_t st = (size_t)c.NC * c.nprim * d;
  c.P0 = (double*)take(st);
  c.PA = (double*)take(st);
  c.PB = (double*)take(st);
  c.dl = (double*)take(3 * (size_t)c.NC * d);
  c.m = (double*)take(9 * (size_t)c.NC * d);
  c.vol = (double*)take((size_t)c.NC * d);
  c.yn = (double*)take((size_t)c.NC * d);
  c.fm1 = (double*)take(4 * (size_t)c.NF1 * d);
  c.fm2 = (double*)take(4 * (size_t)c.NF2 * d);
  c.fm3 = (double*)take(4 * (size_t)c.NF3 * d);
  for (int q = 0; q < 3; ++q) {
    c.Fc[q] = (double*)take((size_t)c.NFI[q] * c.nprim * d);
    c.Fd[q] = (double*)take((size_t)c.NFI[q] * c.nprim * d);
  }
  c.binc = (double*)take(2 * (size_t)std::max<int64_t>(c.NBP, 1) * c.nprim * d);
  if (c.sc.irs) {
    c.dR = (double*)take((size_t)c.NC * c.nprim * d);
    c.drs1 = (double*)take((size_t)c.NC * (c.sc.nsc + 4) * d);
    c.drs2 = (double*)take((size_t)c.NC * (c.sc.nsc + 4) * d);
  }
  c.dblk = (mose_blk*)take(c.blk.size() * sizeof(mose_blk));
  c.drec = (mose_bc_record*)take(std::max<size_t>(c.rec.size(), 1) * sizeof(mose_bc_record));
  c.dff = (mose_farfield*)take(std::max<size_t>(c.ff.size(), 1) * sizeof(mose_farfield));
  c.dbtyp = (signed char*)take((size_t)std::max<int64_t>(c.NBP, 1));
  c.dt = (double*)take(64);
  c.dthist = (double*)take((size_t)c.hist_cap * d);
  c.dtmin = (unsigned long long*)take(64);
  c.bad = (unsigned int*)take(64);
  c.first_bad = (unsigned long long*)take(64);
  c.dtnext = (unsigned long long*)take(64);
  c.rnorm = (double*)take(64 * sizeof(double));
  c.histpos = (int*)take(64);
  c.hslot = (unsigned long long*)take(64);
  if ((size_t)(p - (char*)c.slab) > total) return fail(ctx, MOSE_ERR_NOMEM, "internal: slab overflow");
  for (auto& e : c.ev) HIPCHK(ctx, hipEventCreate(&e));
  HIPCHK(ctx, hipHostMalloc(&c.stage_host, (size_t)c.hist_cap * d + 64, hipHostMallocDefault));
  c.stage_hist = (double*)c.stage_host;
  c.stage_bad = (unsigned*)((char*)c.stage_host + (size_t)c.hist_cap * d);
  c.stage_fb = (unsigned long long*)((char*)c.stage_host + (size_t)c.hist_cap * d + 16);
  return MOSE_OK;
}

// workgroups of the per-block update kernels (the plain path's partials of the residual norm)
static int64_t plain_parts(const Case& c) {
  int64_t n = 0;
  for (auto& B : c.blk) n += ((int64_t)B.n1 * B.n2 * B.n3 + 255) / 256;
  return n;
}

static int ensure_rpart(mose_ctx* ctx, Case& c, int64_t nparts) {
  if (nparts <= c.rpart_cap) return MOSE_OK;
  if (c.rpart) HIPCHK(ctx, hipFree(c.rpart));
  c.rpart = nullptr;
  c.rpart_cap = 0;
  HIPCHK(ctx, hipMalloc(&c.rpart, (size_t)nparts * (c.nprim - c.sc.nsc + 1) * sizeof(double)));
  c.rpart_cap = nparts;
  return MOSE_OK;
}

// The case's step graph is destroyed only once its stream has finished every replay queued on it.
static void drop_graph(Case& c) {
  if ((c.gexec || c.graph) && c.stream) (void)hipStreamSynchronize(c.stream);
  if (c.gexec) (void)hipGraphExecDestroy(c.gexec);
  if (c.graph) (void)hipGraphDestroy(c.graph);
  c.gexec = nullptr;
  c.graph = nullptr;
}

static Case* get_case(mose_ctx* ctx, int id) {
  if (!ctx || id < 0 || id >= (int)ctx->cases.size()) return nullptr;
  Case* c = ctx->cases[id].get();
  return (c && c->valid) ? c : nullptr;
}

// The residual kernels are instantiated for (nsc, nrans) = (1, 0), (1, 1), (4, 0) and (9, 0): the species and RANS
// counts are compile-time counts, as in the reference (one binary per count, src/lib/CMakeLists.txt:64-71).  Four
// species are the count of the reference's multi-species cases (test/1D/diff1, diff2), nine that of the hydrogen-air
// mechanism of the reacting cases.  F<NS, NR>(...) of the case's counts:
#define MOSE_DISPATCH(sc, F, ...)                                                                     \
  ((sc).nsc == 9 ? F<9, 0>(__VA_ARGS__) : (sc).nsc == 4 ? F<4, 0>(__VA_ARGS__)                          \
                                        : (sc).nrans == 0 ? F<1, 0>(__VA_ARGS__) : F<1, 1>(__VA_ARGS__))

static size_t stage_lds_(int nprim, bool visc, int TI, int TJ, int TK);
// static LDS of the stage kernel beside its dynamic part (the workgroup reductions of dt and of the residual norm)
static constexpr size_t kStaticLds = 2048;
// the first tile of the preference list that fits the device's LDS and does not straddle a one-cell direction
static void choose_default_tile(mose_ctx* ctx, Case& c) {
  static const int pref[][3] = {{8, 8, 4}, {16, 4, 4}, {32, 4, 2}, {16, 8, 2}, {32, 8, 1}, {64, 4, 1}, {16, 16, 1},
                                {128, 2, 1}, {256, 1, 1}};
  int mx[3] = {0, 0, 0};
  for (auto& B : c.blk) { mx[0] = std::max(mx[0], B.n1); mx[1] = std::max(mx[1], B.n2); mx[2] = std::max(mx[2], B.n3); }
  for (auto& t : pref) {
    if ((t[1] > 1 && mx[1] == 1) || (t[2] > 1 && mx[2] == 1)) continue;
    if (stage_lds_(c.nprim, c.sc.viscous != 0, t[0], t[1], t[2]) + kStaticLds > (size_t)ctx->max_lds) continue;
    c.TI = t[0]; c.TJ = t[1]; c.TK = t[2];
    return;
  }
}

extern "C" int mose_case_create(mose_ctx* ctx, const mose_scheme* sc, int nb, const int* dims, const double* dl,
                                const double* m, const double* vol, const double* yn, const double* fm1,
                                const double* fm2, const double* fm3, int nrec, const mose_bc_record* rec, int nff,
                                const mose_farfield* ff, int* case_id) {
  if (!ctx || !case_id || !dl || !m || !vol || !fm1 || !fm2 || !fm3 || (nrec > 0 && !rec) || nff < 0) return MOSE_ERR_ARG;
  auto c = std::make_unique<Case>();
  std::string why;
  int rc = plan_case(sc, nb, dims, nrec, rec, nff, ff, *c, why);
  if (rc != MOSE_OK) return fail(ctx, rc, why);
  if (sc->nsc != ctx->nsc) return fail(ctx, MOSE_ERR_ARG, "the scheme's nsc differs from the context tables");
  if (sc->viscous && !ctx->viscous_tables) return fail(ctx, MOSE_ERR_ARG, "viscous case needs mu and k tables");
  if (!((sc->nsc == 1 && sc->nrans == 0) || (sc->nsc == 1 && sc->nrans == 1) || (sc->nsc == 4 && sc->nrans == 0) ||
        (sc->nsc == 9 && sc->nrans == 0)))
    return fail(ctx, MOSE_ERR_ARG, "the library instantiates the residual kernels for nsc = 1 (nrans 0 or 1), nsc = 4 "
                                   "and nsc = 9 (nrans 0)");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  rc = alloc_case(ctx, *c);
  if (rc != MOSE_OK) { case_free(c.get()); return rc; }
  size_t d = sizeof(double);
  HIPCHK(ctx, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));   // default priority
  hipStream_t s = c->stream;
  HIPCHK(ctx, hipMemcpyAsync(c->dl, dl, 3 * (size_t)c->NC * d, hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(c->m, m, 9 * (size_t)c->NC * d, hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(c->vol, vol, (size_t)c->NC * d, hipMemcpyHostToDevice, s));
  if (yn) HIPCHK(ctx, hipMemcpyAsync(c->yn, yn, (size_t)c->NC * d, hipMemcpyHostToDevice, s));
  else HIPCHK(ctx, hipMemsetAsync(c->yn, 0, (size_t)c->NC * d, s));
  HIPCHK(ctx, hipMemcpyAsync(c->fm1, fm1, 4 * (size_t)c->NF1 * d, hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(c->fm2, fm2, 4 * (size_t)c->NF2 * d, hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(c->fm3, fm3, 4 * (size_t)c->NF3 * d, hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(c->dblk, c->blk.data(), c->blk.size() * sizeof(mose_blk), hipMemcpyHostToDevice, s));
  if (!c->rec.empty())
    HIPCHK(ctx, hipMemcpyAsync(c->drec, c->rec.data(), c->rec.size() * sizeof(mose_bc_record), hipMemcpyHostToDevice, s));
  if (!c->ff.empty())
    HIPCHK(ctx, hipMemcpyAsync(c->dff, c->ff.data(), c->ff.size() * sizeof(mose_farfield), hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(c->dbtyp, c->btyp.data(), c->btyp.size(), hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemsetAsync(c->P0, 0, 3 * (size_t)c->NC * c->nprim * d, s));
  // the boundary planes start at zero: the record kernel leaves the entries of no-flux (MOSE 0) faces untouched
  HIPCHK(ctx, hipMemsetAsync(c->binc, 0, 2 * (size_t)std::max<int64_t>(c->NBP, 1) * c->nprim * d, s));
  if (c->sc.irs) {
    HIPCHK(ctx, hipMemsetAsync(c->dR, 0, (size_t)c->NC * c->nprim * d, s));
    HIPCHK(ctx, hipMemsetAsync(c->drs1, 0, (size_t)c->NC * (c->sc.nsc + 4) * d, s));
    HIPCHK(ctx, hipMemsetAsync(c->drs2, 0, (size_t)c->NC * (c->sc.nsc + 4) * d, s));
    c->path = MOSE_PATH_PLAIN;   // the smoothing needs the whole residual: the per-block launches
  }
  HIPCHK(ctx, hipStreamSynchronize(s));
  rc = ensure_rpart(ctx, *c, plain_parts(*c));
  if (rc != MOSE_OK) { case_free(c.get()); return rc; }
  c->valid = true;
  {
    const char* e = std::getenv("MOSE_NO_GRAPH");
    if (e && e[0] == '1') c->use_graph = false;
  }
  choose_default_tile(ctx, *c);
  ctx->cases.push_back(std::move(c));
  *case_id = (int)ctx->cases.size() - 1;
  return MOSE_OK;
}

extern "C" int mose_case_destroy(mose_ctx* ctx, int id) {
  Case* c = get_case(ctx, id);
  if (!c) return MOSE_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(c->stream));
  case_free(c);
  return MOSE_OK;
}

extern "C" int mose_state_upload(mose_ctx* ctx, int id, const double* P) {
  Case* c = get_case(ctx, id);
  if (!c || !P) return MOSE_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  size_t n = (size_t)c->NC * c->nprim * sizeof(double);
  HIPCHK(ctx, hipMemcpyAsync(c->P0, P, n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(ctx, hipStreamSynchronize(c->stream));
  c->dt_valid = false;
  return MOSE_OK;
}

extern "C" int mose_state_download(mose_ctx* ctx, int id, double* P) {
  Case* c = get_case(ctx, id);
  if (!c || !P) return MOSE_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  size_t n = (size_t)c->NC * c->nprim * sizeof(double);
  HIPCHK(ctx, hipMemcpyAsync(P, c->P0, n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(ctx, hipStreamSynchronize(c->stream));
  return MOSE_OK;
}

// ----------------------------------------------------------------------------------------------------------------
// launches
// ----------------------------------------------------------------------------------------------------------------
static CaseDev make_dev(mose_ctx* ctx, Case& c, const double* Pin, double* Pout) {
  CaseDev D{};
  D.Pin = Pin; D.Pout = Pout; D.PO = c.P0; D.Pg = const_cast<double*>(Pin);
  D.NC = c.NC; D.NF1 = c.NF1; D.NF2 = c.NF2; D.NF3 = c.NF3;
  D.dl = c.dl; D.m = c.m; D.vol = c.vol; D.yn = c.yn; D.fm1 = c.fm1; D.fm2 = c.fm2; D.fm3 = c.fm3;
  for (int q = 0; q < 3; ++q) { D.Fc[q] = c.Fc[q]; D.Fd[q] = c.Fd[q]; D.NFI[q] = c.NFI[q]; }
  D.binc = c.binc; D.NBP = std::max<int64_t>(c.NBP, 1);
  D.blk = c.dblk; D.rec = c.drec; D.nrec = (int)c.rec.size(); D.ff = c.dff; D.btyp = c.dbtyp;
  D.lim = c.sc.limiter; D.rs = c.sc.riemann; D.visc = c.sc.viscous; D.guard = c.sc.thermo_guard; D.mco = c.sc.mco;
  D.pimp = c.sc.point_implicit;
  D.vc.sc = c.sc.sc; D.vc.sct = c.sc.sct; D.vc.prt = c.sc.prt; D.vc.prl = c.sc.prandtl;
  D.tab = ctx->dtab;
  D.dt = c.dt; D.dtmin = c.dtmin; D.bad = c.bad; D.first_bad = c.first_bad; D.strang = c.strang;
  D.cfl = c.sc.cfl; D.vnn = c.sc.vnn; D.ndir = c.sc.ndir; D.prim = c.sc.integration == MOSE_INT_PRIM;
  return D;
}

static inline unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <int NS, int NR>
static void launch_dt(Case& c, const CaseDev& D, hipStream_t s, int step_index) {
  k_dt_init<<<1, 64, 0, s>>>(c.dtmin);
  for (int b = 0; b < c.nb; ++b) {
    int64_t n = (int64_t)c.blk[b].n1 * c.blk[b].n2 * c.blk[b].n3;
    k_dt<NS, NR><<<nblocks(n, 256), 256, 0, s>>>(D, b);
  }
  k_dt_final<<<1, 64, 0, s>>>(c.dtmin, c.dt, step_index >= 0 ? c.dthist : nullptr, step_index);
}

template <int NS, int NR>
static void launch_flux(Case& c, const CaseDev& D, hipStream_t s) {
  int nrec = (int)c.rec.size();
  if (nrec > 0) {
    k_ghost<NS, NR><<<nblocks(nrec, 256), 256, 0, s>>>(D);
    k_bcrec<NS, NR><<<nblocks(nrec, 128), 128, 0, s>>>(D);
  }
  for (int b = 0; b < c.nb; ++b) {
    const mose_blk& B = c.blk[b];
    int64_t n0 = (int64_t)(B.n1 - 1) * B.n2 * B.n3, n1_ = (int64_t)B.n1 * (B.n2 - 1) * B.n3;
    int64_t n2_ = (int64_t)B.n1 * B.n2 * (B.n3 - 1);
    if (n0 > 0) k_face<NS, NR, 0><<<nblocks(n0, 128), 128, 0, s>>>(D, b);
    if (n1_ > 0) k_face<NS, NR, 1><<<nblocks(n1_, 128), 128, 0, s>>>(D, b);
    if (n2_ > 0) k_face<NS, NR, 2><<<nblocks(n2_, 128), 128, 0, s>>>(D, b);
  }
}

// Implicit residual smoothing (IRS_apply, Lib_IRS.f90:53-96) between the fluxes and the update of a plain-path stage:
// the residual of every interior cell into dR (pass 2, boundary increments, SA source), its scaling for the
// conservative update, then per computational direction the residual's ghost layer (block faces, then the connection
// records' donors) and the Jacobi sweeps of each block (drs1 <- dR over the block, sweeps drs1 -> drs2 -> drs1,
// dR <- drs1 on the interior).  The update then reads dR.
template <int NS, int NR>
static void launch_irs(Case& c, const CaseDev& D, hipStream_t s) {
  const int nf = c.sc.nsc + 4;
  for (int b = 0; b < c.nb; ++b) {
    const int64_t ncell = (int64_t)c.blk[b].n1 * c.blk[b].n2 * c.blk[b].n3;
    k_residual_out<NS, NR><<<nblocks(ncell, 256), 256, 0, s>>>(D, b, c.dR);
    if (c.sc.integration != MOSE_INT_PRIM) k_irs_scale<NS, NR><<<nblocks(ncell, 256), 256, 0, s>>>(D, b, c.dR);
  }
  const int nrec = (int)c.rec.size();
  for (int d = 0; d < c.sc.ndir; ++d) {
    for (int b = 0; b < c.nb; ++b) {
      const mose_blk& B = c.blk[b];
      const int64_t nbp = 2 * ((int64_t)B.n2 * B.n3 + (int64_t)B.n1 * B.n3 + (int64_t)B.n1 * B.n2);
      k_irs_ghost_faces<<<nblocks(nbp, 256), 256, 0, s>>>(c.dblk, b, c.dR, c.NC, nf);
    }
    if (nrec > 0) k_irs_ghost_conn<<<nblocks(nrec, 256), 256, 0, s>>>(c.dblk, c.drec, nrec, c.dR, c.NC, nf);
    for (int b = 0; b < c.nb; ++b) {
      const mose_blk& B = c.blk[b];
      const int64_t ncell = (int64_t)B.n1 * B.n2 * B.n3;
      const int64_t nflat = (int64_t)B.pi * B.pj * (B.n3 + 2 * MOSE_GC);
      const int64_t off = d == 0 ? 1 : (d == 1 ? (int64_t)B.pi : (int64_t)B.pi * B.pj);
      k_irs_copy_block<<<nblocks(nflat, 256), 256, 0, s>>>(B, c.dR, c.drs1, c.NC, nf);
      for (int jac = 0; jac < c.sc.irs_njacobi; ++jac) {
        k_irs_sweep<<<nblocks(ncell, 256), 256, 0, s>>>(B, c.dR, c.drs1, c.drs2, c.NC, nf, off, c.sc.irs_beta);
        k_irs_copy_interior<<<nblocks(ncell, 256), 256, 0, s>>>(B, c.drs2, c.drs1, c.NC, nf);
      }
      k_irs_copy_interior<<<nblocks(ncell, 256), 256, 0, s>>>(B, c.drs1, c.dR, c.NC, nf);
    }
  }
}

// ----------------------------------------------------------------------------------------------------------------
// fused path: tiles, launch, autotune
// ----------------------------------------------------------------------------------------------------------------
static int tile_slots(int TI, int TJ, int TK) { return (TI + 1) * TJ * TK + TI * (TJ + 1) * TK + TI * TJ * (TK + 1); }

static int hs_slots(int TI, int TJ, int TK) { return (TI + 2) * TJ * TK + TI * (TJ + 2) * TK + TI * TJ * (TK + 2); }

// LDS doubles of region 0 (half-slopes, then the viscous cell cache: rho, 1/rho, T and the primitives of the tile
// and its one-cell halo) and the bytes of the whole tile
static int stage_r0(const Case& c, int TI, int TJ, int TK) {
  int r0 = hs_slots(TI, TJ, TK) * c.nprim;
  if (c.sc.viscous) r0 = std::max(r0, (3 + c.nprim) * (TI + 2) * (TJ + 2) * (TK + 2));
  return (r0 + 31) & ~31;
}
static size_t stage_lds(const Case& c, int TI, int TJ, int TK) {
  if (c.path == MOSE_PATH_FAST)
    return fast_lds_doubles(c.sc.nsc, c.sc.nrans, c.nprim, c.sc.viscous != 0, TI, TJ, TK) * sizeof(double);
  return ((size_t)stage_r0(c, TI, TJ, TK) + (size_t)tile_slots(TI, TJ, TK) * c.nprim) * sizeof(double);
}
static size_t stage_lds_(int nprim, bool visc, int TI, int TJ, int TK) {
  Case c;
  c.nprim = nprim;
  c.sc.viscous = visc ? 1 : 0;
  return stage_lds(c, TI, TJ, TK);
}

static int set_tiles(mose_ctx* ctx, Case& c, int TI, int TJ, int TK) {
  if (TI < 1 || TJ < 1 || TK < 1 || (TI * TJ * TK != 128 && TI * TJ * TK != 256 && TI * TJ * TK != 512))
    return fail(ctx, MOSE_ERR_ARG, "a tile holds 128, 256 or 512 cells");
  if (stage_lds(c, TI, TJ, TK) + kStaticLds > (size_t)ctx->max_lds)
    return fail(ctx, MOSE_ERR_ARG, "tile " + std::to_string(TI) + "x" + std::to_string(TJ) + "x" + std::to_string(TK) +
                                       " needs " + std::to_string(stage_lds(c, TI, TJ, TK)) + " B of LDS, the device allows " +
                                       std::to_string(ctx->max_lds));
  // tiles block by block, in super-tiles of 4 x 4 x 4 tiles (i fastest inside a super-tile and between them):
  // neighbouring tiles run close together in time and share the L2 of their XCD (1.4 % faster than plain i-fastest
  // order on the 256^3 Taylor-Green case)
  std::vector<Tile> t;
  for (int b = 0; b < c.nb; ++b) {
    const mose_blk& B = c.blk[b];
    const int na = (B.n1 + TI - 1) / TI, nbj = (B.n2 + TJ - 1) / TJ, nc = (B.n3 + TK - 1) / TK;
    for (int sc = 0; sc < nc; sc += 4)
      for (int sb = 0; sb < nbj; sb += 4)
        for (int sa = 0; sa < na; sa += 4)
          for (int cc = sc; cc < std::min(sc + 4, nc); ++cc)
            for (int bb = sb; bb < std::min(sb + 4, nbj); ++bb)
              for (int a = sa; a < std::min(sa + 4, na); ++a) t.push_back(Tile{b, 1 + a * TI, 1 + bb * TJ, 1 + cc * TK});
  }
  if ((int)t.size() > c.tiles_cap) {
    if (c.dtiles) HIPCHK(ctx, hipFree(c.dtiles));
    c.dtiles = nullptr;
    HIPCHK(ctx, hipMalloc(&c.dtiles, t.size() * sizeof(Tile)));
    c.tiles_cap = (int)t.size();
  }
  HIPCHK(ctx, hipMemcpyAsync(c.dtiles, t.data(), t.size() * sizeof(Tile), hipMemcpyHostToDevice, c.stream));
  HIPCHK(ctx, hipStreamSynchronize(c.stream));
  c.ntiles = (int)t.size();
  c.TI = TI; c.TJ = TJ; c.TK = TK;
  drop_graph(c);
  return ensure_rpart(ctx, c, c.ntiles);
}

// Every tile of 256 or 512 cells runs as one workgroup of 512 threads: two per cell of a 256-cell tile, one per cell
// of a 512-cell tile.  The kernel takes 127 VGPRs, so two 256-cell tiles (76 KB of LDS each at nprim 5, viscous) run
// on a CU at 4 waves per SIMD; with one thread per cell they ran at 2, and the stage took 16 % longer (7.22 against
// 5.96 ms on the 256^3 Taylor-Green case).  A 128-cell tile runs as 256 threads (two per cell): with RANS (nprim 6,
// 149 VGPRs) three of its 48 KB workgroups share a CU, where a 256-cell tile's 92 KB leaves one.  The two preferred
// tile shapes run kernels with the shape fixed at compile time; the others read it at run time (launch_stage_tiles).
template <int NS, int NR>
static void launch_stage(Case& c, const CaseDev& D, hipStream_t s, double rkc, bool last) {
  StageArgs A;
  A.tiles = c.dtiles; A.ntiles = c.ntiles; A.TI = c.TI; A.TJ = c.TJ; A.TK = c.TK;
  A.nslot = tile_slots(c.TI, c.TJ, c.TK);
  A.nhs = hs_slots(c.TI, c.TJ, c.TK);
  A.ncache = (c.TI + 2) * (c.TJ + 2) * (c.TK + 2);
  A.r0 = stage_r0(c, c.TI, c.TJ, c.TK);
  A.last = last ? 1 : 0; A.rkc = rkc; A.dtnext = c.dtnext; A.rpart = c.rpart;
  size_t lds = stage_lds(c, c.TI, c.TJ, c.TK);
  unsigned grid = (unsigned)(((c.ntiles + 7) / 8) * 8);
  if (c.path == MOSE_PATH_FAST) launch_stage_fast<NS, NR>(D, A, c.sc.riemann, c.sc.viscous != 0, grid, lds, s);
  else launch_stage_tiles<NS, NR>(D, A, c.sc.riemann, c.sc.viscous != 0, grid, lds, s);
}

// records the step's dt in the history at the device-side position (the same kernel serves eager and graph steps)
__global__ void k_step_begin(const double* dt, double* hist, int* pos, int cap) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int t = *pos;
    if (t < cap) hist[t] = *dt;
    *pos = t + 1;
  }
}


static int case_chem_launch(mose_ctx* ctx, Case& c, hipStream_t s);

template <int NS, int NR>
static int run_steps(mose_ctx* ctx, Case& c, int nsteps) {
  hipStream_t s = c.stream;
  static const double rk[3][3] = {{1.0, 0.0, 0.0}, {1.0, 0.5, 0.0}, {1.0, 0.25, 2.0 / 3.0}};   // Lib_RK.f90:6-9
  int nrk = c.sc.n_rk;
  bool fused = (c.path == MOSE_PATH_FUSED || c.path == MOSE_PATH_FAST);
  if (fused && c.ntiles == 0) {
    int rc = set_tiles(ctx, c, c.TI, c.TJ, c.TK);
    if (rc != MOSE_OK) return rc;
  }
  if (fused && stage_lds(c, c.TI, c.TJ, c.TK) + kStaticLds > (size_t)ctx->max_lds) return fail(ctx, MOSE_ERR_ARG, "tile does not fit");
  HIPCHK(ctx, hipMemsetAsync(c.bad, 0, 2 * sizeof(unsigned), s));
  HIPCHK(ctx, hipMemsetAsync(c.first_bad, 0xff, sizeof(unsigned long long), s));
  const int nstr = c.n_strang;
  size_t nev = c.timing ? (size_t)2 * nsteps * nrk * nstr : 0;
  while (c.tev.size() < nev) {
    hipEvent_t e;
    HIPCHK(ctx, hipEventCreate(&e));
    c.tev.push_back(e);
  }
  HIPCHK(ctx, hipMemsetAsync(c.histpos, 0, sizeof(int), s));
  if (nstr == 2) k_zero_u64<<<1, 64, 0, s>>>(c.chem_ctr + 8, 8);   // the chemistry statistics of this call
  size_t iev = 0;
  for (int n = 0; n < nsteps; ++n) {
    CaseDev D0 = make_dev(ctx, c, c.P0, c.PA);
    if (!fused || !c.dt_valid) launch_dt<NS, NR>(c, D0, s, -1);
    // the fused step with dt carried over from the last stage is the same launch sequence every step: capture it
    // once and replay it (no memset node in the graph; every node is a kernel)
    const bool graph = fused && c.dt_valid && !c.timing && c.use_graph;
    if (graph && c.gexec) {
      HIPCHK(ctx, hipGraphLaunch(c.gexec, s));
      c.nparts_last = c.ntiles;
      continue;
    }
    if (graph) HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    k_step_begin<<<1, 64, 0, s>>>(c.dt, c.dthist, c.histpos, c.hist_cap);
    // the strang loop of Explicit_Step (Mod_Explicit.f90:155-294): one substage without chemistry; with it, half a
    // transport step (strangcoeff 1/2 in the update), Chemistry_Newstate_blk over the whole dt on the resident state,
    // the other half.  Each substage starts from the state it finds (Copy_State: P0 is the base of its stages).
    for (int istr = 1; istr <= nstr; ++istr) {
      for (int irk = 1; irk <= nrk; ++irk) {
        double* in = irk == 1 ? c.P0 : (irk == 2 ? c.PA : c.PB);
        double* out = irk == nrk ? c.P0 : (irk == 1 ? c.PA : c.PB);
        CaseDev D = make_dev(ctx, c, in, out);
        double rkc = rk[nrk - 1][irk - 1];
        const bool last = irk == nrk && istr == nstr;   // the step's last stage: next dt and the norm partials
        if (fused) {
          int nrec = (int)c.rec.size();
          if (nrec > 0) {
            k_ghost<NS, NR><<<nblocks(nrec, 256), 256, 0, s>>>(D);
            k_bcrec<NS, NR><<<nblocks(nrec, 128), 128, 0, s>>>(D);
          }
          if (last) k_dt_init<<<1, 64, 0, s>>>(c.dtnext);
          if (c.timing) HIPCHK(ctx, hipEventRecord(c.tev[iev++], s));
          launch_stage<NS, NR>(c, D, s, rkc, last);
          if (last) c.nparts_last = c.ntiles;
          if (c.timing) HIPCHK(ctx, hipEventRecord(c.tev[iev++], s));
          if (last) k_dt_final<<<1, 64, 0, s>>>(c.dtnext, c.dt, nullptr, 0);
        } else {
          launch_flux<NS, NR>(c, D, s);
          if (c.sc.irs) launch_irs<NS, NR>(c, D, s);
          if (c.timing) HIPCHK(ctx, hipEventRecord(c.tev[iev++], s));
          int64_t part0 = 0;
          for (int b = 0; b < c.nb; ++b) {
            int64_t ncell = (int64_t)c.blk[b].n1 * c.blk[b].n2 * c.blk[b].n3;
            k_update<NS, NR><<<nblocks(ncell, 256), 256, 0, s>>>(D, b, rkc, last ? c.rpart : nullptr, part0,
                                                                 c.sc.irs ? c.dR : nullptr);
            part0 += nblocks(ncell, 256);
          }
          if (last) c.nparts_last = part0;
          if (c.timing) HIPCHK(ctx, hipEventRecord(c.tev[iev++], s));
        }
      }
      if (istr == 1 && nstr == 2) {
        int rc = case_chem_launch(ctx, c, s);
        if (rc != MOSE_OK) return rc;
      }
    }
    if (graph) {
      HIPCHK(ctx, hipStreamEndCapture(s, &c.graph));
      HIPCHK(ctx, hipGraphInstantiate(&c.gexec, c.graph, nullptr, nullptr, 0));
      HIPCHK(ctx, hipGraphLaunch(c.gexec, s));
    }
    c.dt_valid = fused;
  }
  HIPCHK(ctx, hipGetLastError());
  if (c.timing && iev > 0) {
    HIPCHK(ctx, hipStreamSynchronize(s));
    double tot = 0.0;
    for (size_t q = 0; q + 1 < iev; q += 2) {
      float ms = 0.f;
      HIPCHK(ctx, hipEventElapsedTime(&ms, c.tev[q], c.tev[q + 1]));
      tot += ms;
    }
    c.ms_stage_total = tot;
    c.stage_launches = (int64_t)(iev / 2);
  }
  return MOSE_OK;
}

template <int NS, int NR>
static int autotune(mose_ctx* ctx, Case& c, mose_tuning* chosen) {
  // candidates of 128, 256 and 512 cells; each runs the ghost, record and stage kernels of stage 1 (into the scratch
  // buffer PA, the state P0 untouched) a few times, the fastest is kept
  static const int cand[][3] = {{32, 4, 2}, {64, 2, 2}, {16, 4, 4}, {16, 8, 2}, {8, 8, 4}, {64, 4, 1}, {32, 8, 1},
                                {128, 2, 1}, {256, 1, 1}, {8, 8, 8}, {16, 8, 4}, {16, 4, 8}, {32, 4, 4}, {16, 16, 2},
                                {32, 16, 1}, {64, 8, 1}, {8, 4, 4}, {16, 4, 2}, {16, 8, 1}};
  hipStream_t s = c.stream;
  hipEvent_t e0, e1;
  HIPCHK(ctx, hipEventCreate(&e0));
  HIPCHK(ctx, hipEventCreate(&e1));
  int dims_max[3] = {0, 0, 0};
  for (auto& B : c.blk) { dims_max[0] = std::max(dims_max[0], B.n1); dims_max[1] = std::max(dims_max[1], B.n2); dims_max[2] = std::max(dims_max[2], B.n3); }
  CaseDev D0 = make_dev(ctx, c, c.P0, c.PA);
  launch_dt<NS, NR>(c, D0, s, -1);
  double best = 1e30;
  int bi = 0;
  for (int q = 0; q < (int)(sizeof cand / sizeof cand[0]); ++q) {
    int TI = cand[q][0], TJ = cand[q][1], TK = cand[q][2];
    if (stage_lds(c, TI, TJ, TK) + kStaticLds > (size_t)ctx->max_lds) continue;
    if ((TJ > 1 && dims_max[1] == 1) || (TK > 1 && dims_max[2] == 1)) continue;   // no tile across a 1-cell direction
    if (set_tiles(ctx, c, TI, TJ, TK) != MOSE_OK) continue;
    CaseDev D = make_dev(ctx, c, c.P0, c.PA);
    int nrec = (int)c.rec.size();
    if (nrec > 0) {
      k_ghost<NS, NR><<<nblocks(nrec, 256), 256, 0, s>>>(D);
      k_bcrec<NS, NR><<<nblocks(nrec, 128), 128, 0, s>>>(D);
    }
    launch_stage<NS, NR>(c, D, s, 1.0, false);   // warm
    HIPCHK(ctx, hipEventRecord(e0, s));
    for (int r = 0; r < 3; ++r) launch_stage<NS, NR>(c, D, s, 1.0, false);
    HIPCHK(ctx, hipEventRecord(e1, s));
    HIPCHK(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) { best = ms; bi = q; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(ctx, hipGetLastError());
  int rc = set_tiles(ctx, c, cand[bi][0], cand[bi][1], cand[bi][2]);
  if (rc != MOSE_OK) return rc;
  c.dt_valid = false;
  if (chosen) { chosen->ti = c.TI; chosen->tj = c.TJ; chosen->tk = c.TK; }
  return MOSE_OK;
}

extern "C" int mose_case_tune(mose_ctx* ctx, int id, const mose_tuning* t, mose_tuning* chosen) {
  Case* c = get_case(ctx, id);
  if (!c || !t) return MOSE_ERR_ARG;
  if (t->path != MOSE_PATH_FUSED && t->path != MOSE_PATH_PLAIN && t->path != MOSE_PATH_FAST)
    return fail(ctx, MOSE_ERR_ARG, "unknown path");
  if (c->launched || c->pending) return fail(ctx, MOSE_ERR_ARG, "mose_case_tune: a launch of the case awaits its poll");
  if (t->path == MOSE_PATH_FAST && c->sc.riemann == MOSE_RS_SLAU)
    return fail(ctx, MOSE_ERR_ARG, "the performance path's kernels take HLLC, HLLE, AUSM+ and LMRoe (SLAU: the parity "
                                   "or plain path)");
  if (c->sc.irs && t->path != MOSE_PATH_PLAIN)
    return fail(ctx, MOSE_ERR_ARG, "implicit residual smoothing runs on the plain path (MOSE_PATH_PLAIN)");
  if (t->path == MOSE_PATH_FAST && c->sc.nsc == 9 && c->sc.integration == MOSE_INT_PRIM)
    return fail(ctx, MOSE_ERR_ARG, "the performance path's nine-species kernels take the conservative update only");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  c->path = t->path;
  c->timing = t->timing != 0;
  {
    const char* e = std::getenv("MOSE_NO_GRAPH");
    c->use_graph = t->no_graph == 0 && !(e && e[0] == '1');
  }
  c->dt_valid = false;
  drop_graph(*c);
  if (chosen) *chosen = *t;
  if (c->path == MOSE_PATH_FUSED || c->path == MOSE_PATH_FAST) {
    if (t->ti == 0 && t->tj == 0 && t->tk == 0)
      return MOSE_DISPATCH(c->sc, autotune, ctx, *c, chosen);
    return set_tiles(ctx, *c, t->ti, t->tj, t->tk);
  }
  return MOSE_OK;
}

// Step n whole steps of a case on its own stream and queue the copy of the step report into the case's pinned
// staging; nothing waits.  case_collect waits for the case's stream and fills the report.
static int case_enqueue(mose_ctx* ctx, Case* c, int nsteps) {
  int rc;
  rc = MOSE_DISPATCH(c->sc, run_steps, ctx, *c, nsteps);
  if (rc != MOSE_OK) return rc;
  size_t nh = (size_t)std::min(nsteps, c->hist_cap);
  if (nh > 0)
    HIPCHK(ctx, hipMemcpyAsync(c->stage_hist, c->dthist, nh * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(ctx, hipMemcpyAsync(c->stage_bad, c->bad, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(ctx, hipMemcpyAsync(c->stage_fb, c->first_bad, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  if (c->n_strang == 2)
    HIPCHK(ctx, hipMemcpyAsync(c->stage_chem, c->chem_ctr + 8, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               c->stream));
  c->pending = nsteps;
  return MOSE_OK;
}

static int case_collect(mose_ctx* ctx, Case* c, mose_report* rep) {
  int nsteps = c->pending;
  c->pending = 0;
  HIPCHK(ctx, hipStreamSynchronize(c->stream));
  size_t nh = (size_t)std::min(nsteps, c->hist_cap);
  for (size_t q = 0; q < nh; ++q) c->time += c->stage_hist[q];
  if (nh > 0) c->last_dt = c->stage_hist[nh - 1];
  c->steps += nsteps;
  unsigned bad[2] = {c->stage_bad[0], c->stage_bad[1]};
  unsigned long long fb = *c->stage_fb;
  int status = MOSE_OK;
  const bool chem_failed = c->n_strang == 2 && nsteps > 0 && c->stage_chem[3] > 0;
  if (chem_failed) status = MOSE_ERR_ODE;
  if (bad[1] > 0) status = MOSE_ERR_THERMO;
  if (bad[0] > 0) status = MOSE_ERR_STATE;
  if (rep) {
    rep->dt = c->last_dt;
    rep->time = c->time;
    rep->steps = c->steps;
    rep->bad_cells = (int)(bad[0] + bad[1]);
    rep->first_bad_block = fb == ~0ull ? -1 : (int)(fb >> 48);
    rep->first_bad_k = fb == ~0ull ? -1 : (int)((fb >> 32) & 0xffff);
    rep->first_bad_j = fb == ~0ull ? -1 : (int)((fb >> 16) & 0xffff);
    rep->first_bad_i = fb == ~0ull ? -1 : (int)(fb & 0xffff);
    rep->status = status;
  }
  if (status == MOSE_ERR_ODE) {
    const unsigned long long* st = c->stage_chem;
    ctx->err = std::to_string(st[3]) + " chemistry integrations failed (idid < 0) in the last " +
               std::to_string(nsteps) + " steps, the first at interior cell " + std::to_string(~0ull - st[5]);
  } else if (status != MOSE_OK) {
    char msg[256];
    snprintf(msg, sizeof msg, "%s in %u cells, first at block %d (i,j,k) = (%d,%d,%d)",
             status == MOSE_ERR_STATE ? "p < 0 or non-finite state" : "T outside (1, 9990) K", bad[0] + bad[1],
             fb == ~0ull ? -1 : (int)(fb >> 48), (int)(fb & 0xffff), (int)((fb >> 16) & 0xffff), (int)((fb >> 32) & 0xffff));
    ctx->err = msg;
  }
  return status;
}

extern "C" int mose_case_step(mose_ctx* ctx, int id, int nsteps, mose_report* rep) {
  Case* c = get_case(ctx, id);
  if (!c || nsteps < 0) return MOSE_ERR_ARG;
  if (c->launched) return fail(ctx, MOSE_ERR_ARG, "mose_case_step: a mose_case_launch of this case awaits its poll");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = case_enqueue(ctx, c, nsteps);
  if (rc != MOSE_OK) return rc;
  return case_collect(ctx, c, rep);
}

// Asynchronous stepping for a serving loop: queue now, collect when the case's event has fired.
extern "C" int mose_case_launch(mose_ctx* ctx, int id, int nsteps) {
  Case* c = get_case(ctx, id);
  if (!c || nsteps < 0) return MOSE_ERR_ARG;
  if (c->pending > 0) return fail(ctx, MOSE_ERR_ARG, "mose_case_launch: a launch of this case is still in flight");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = case_enqueue(ctx, c, nsteps);
  if (rc != MOSE_OK) return rc;
  HIPCHK(ctx, hipEventRecord(c->ev[0], c->stream));
  c->launched = true;
  return MOSE_OK;
}

extern "C" int mose_case_poll(mose_ctx* ctx, int id, int* done, mose_report* rep) {
  Case* c = get_case(ctx, id);
  if (!c || !done) return MOSE_ERR_ARG;
  if (!c->launched) return fail(ctx, MOSE_ERR_ARG, "mose_case_poll: no launch of this case is in flight");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipError_t q = hipEventQuery(c->ev[0]);
  if (q == hipErrorNotReady) {
    *done = 0;
    return MOSE_OK;
  }
  if (q != hipSuccess) return fail(ctx, MOSE_ERR_HIP, std::string("mose_case_poll: ") + hipGetErrorString(q));
  *done = 1;
  c->launched = false;
  return case_collect(ctx, c, rep);
}

// All cases are queued first, each on its own stream, then collected: the steps of different cases run
// concurrently on the device (small cases fill the CUs that one case alone leaves idle).
extern "C" int mose_cases_step(mose_ctx* ctx, const int* ids, int n, int nsteps, mose_report* reps) {
  if (!ctx || !ids || n < 0 || nsteps < 0) return MOSE_ERR_ARG;
  for (int q = 0; q < n; ++q) {
    Case* c = get_case(ctx, ids[q]);
    if (!c) return fail(ctx, MOSE_ERR_ARG, "mose_cases_step: bad case id");
    if (c->launched) return fail(ctx, MOSE_ERR_ARG, "mose_cases_step: a mose_case_launch of a case awaits its poll");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int worst = MOSE_OK;
  std::vector<char> queued(n, 0);
  for (int q = 0; q < n; ++q) {
    int rc = case_enqueue(ctx, get_case(ctx, ids[q]), nsteps);
    if (rc != MOSE_OK) { if (worst == MOSE_OK) worst = rc; continue; }
    queued[q] = 1;
  }
  for (int q = 0; q < n; ++q) {
    if (!queued[q]) continue;
    int rc = case_collect(ctx, get_case(ctx, ids[q]), reps ? &reps[q] : nullptr);
    if (rc != MOSE_OK && worst == MOSE_OK) worst = rc;
  }
  return worst;
}

template <int NS, int NR>
static int residual_impl(mose_ctx* ctx, Case& c, double* R) {
  hipStream_t s = c.stream;
  CaseDev D = make_dev(ctx, c, c.P0, c.PA);
  launch_flux<NS, NR>(c, D, s);
  double* dR = c.PB;   // scratch
  HIPCHK(ctx, hipMemsetAsync(dR, 0, (size_t)c.NC * c.nprim * sizeof(double), s));
  for (int b = 0; b < c.nb; ++b) {
    int64_t ncell = (int64_t)c.blk[b].n1 * c.blk[b].n2 * c.blk[b].n3;
    k_residual_out<NS, NR><<<nblocks(ncell, 256), 256, 0, s>>>(D, b, dR);
  }
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(R, dR, (size_t)c.NC * c.nprim * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  return MOSE_OK;
}

extern "C" int mose_case_residual(mose_ctx* ctx, int id, double* R) {
  Case* c = get_case(ctx, id);
  if (!c || !R) return MOSE_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return MOSE_DISPATCH(c->sc, residual_impl, ctx, *c, R);
}

extern "C" int mose_case_dt(mose_ctx* ctx, int id, double* dt) {
  Case* c = get_case(ctx, id);
  if (!c || !dt) return MOSE_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  CaseDev D = make_dev(ctx, *c, c->P0, c->PA);
  MOSE_DISPATCH(c->sc, launch_dt, *c, D, c->stream, -1);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(dt, c->dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(ctx, hipStreamSynchronize(c->stream));
  return MOSE_OK;
}

extern "C" int mose_case_residual_norm(mose_ctx* ctx, int id, double* out) {
  Case* c = get_case(ctx, id);
  if (!c || !out) return MOSE_ERR_ARG;
  if (c->steps == 0 || c->nparts_last == 0) return fail(ctx, MOSE_ERR_ARG, "mose_case_residual_norm: no step taken yet");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int nres = c->nprim - c->sc.nsc + 1;
  k_rnorm_final<<<1, 64, 0, c->stream>>>(c->rpart, c->nparts_last, nres, c->rnorm);
  HIPCHK(ctx, hipGetLastError());
  double h[64];
  HIPCHK(ctx, hipMemcpyAsync(h, c->rnorm, nres * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(ctx, hipStreamSynchronize(c->stream));
  for (int q = 0; q < nres; ++q) out[q] = std::sqrt(h[q]);   // Mod_Explicit.f90:358
  return MOSE_OK;
}

template <int NS, int NR>
static void launch_hash(Case& c, hipStream_t s) {
  for (int b = 0; b < c.nb; ++b) {
    const mose_blk& B = c.blk[b];
    int64_t n = (int64_t)B.n1 * B.n2 * B.n3;
    k_hash<NS, NR><<<nblocks(n, 256), 256, 0, s>>>(c.P0, c.NC, B, c.hslot);
  }
}

extern "C" int mose_case_state_hash(mose_ctx* ctx, int id, uint64_t* lo, uint64_t* hi) {
  Case* c = get_case(ctx, id);
  if (!c || !lo || !hi) return MOSE_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipMemsetAsync(c->hslot, 0, 2 * sizeof(unsigned long long), c->stream));
  MOSE_DISPATCH(c->sc, launch_hash, *c, c->stream);
  HIPCHK(ctx, hipGetLastError());
  unsigned long long h[2];
  HIPCHK(ctx, hipMemcpyAsync(h, c->hslot, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(ctx, hipStreamSynchronize(c->stream));
  *lo = h[0];
  *hi = h[1];
  return MOSE_OK;
}

extern "C" int mose_case_timing(mose_ctx* ctx, int id, double* avg, double* total, int64_t* launches) {
  Case* c = get_case(ctx, id);
  if (!c) return MOSE_ERR_ARG;
  if (avg) *avg = c->stage_launches > 0 ? c->ms_stage_total / (double)c->stage_launches : 0.0;
  if (total) *total = c->ms_stage_total;
  if (launches) *launches = c->stage_launches;
  return MOSE_OK;
}

// ----------------------------------------------------------------------------------------------------------------
// stream-copy probe
// ----------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_copy(const double2* __restrict__ a, double2* __restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

extern "C" int mose_bandwidth_probe(mose_ctx* ctx, int64_t bytes, int reps, double* gbs) {
  if (!ctx || bytes < 4096 || reps < 1 || !gbs) return MOSE_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int64_t n = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  HIPCHK(ctx, hipMalloc(&a, (size_t)n * 16));
  if (hipMalloc(&b, (size_t)n * 16) != hipSuccess) { (void)hipFree(a); return fail(ctx, MOSE_ERR_NOMEM, "probe buffers"); }
  hipStream_t s = ctx->stream;
  (void)hipMemsetAsync(a, 0, (size_t)n * 16, s);
  hipEvent_t e0, e1;
  HIPCHK(ctx, hipEventCreate(&e0));
  HIPCHK(ctx, hipEventCreate(&e1));
  unsigned grid = 256 * 64;
  k_copy<<<grid, 256, 0, s>>>(a, b, n);
  HIPCHK(ctx, hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) k_copy<<<grid, 256, 0, s>>>(a, b, n);
  HIPCHK(ctx, hipEventRecord(e1, s));
  HIPCHK(ctx, hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
  *gbs = 2.0 * (double)n * 16.0 * reps / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return MOSE_OK;
}

// ----------------------------------------------------------------------------------------------------------------
// stiff chemistry: batches of independent cells (Chemistry_Newstate_blk with the host RODAS of the general path)
// ----------------------------------------------------------------------------------------------------------------
__global__ void k_pow(const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = mose_pow(x[i], y[i]);
}

extern "C" int mose_pow_batch(mose_ctx* ctx, int64_t n, const double* x, const double* y, double* out) {
  if (!ctx || n < 0 || (n > 0 && (!x || !y || !out))) return MOSE_ERR_ARG;
  if (n == 0) return MOSE_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  double* d = nullptr;
  const size_t bytes = (size_t)n * sizeof(double);
  if (hipMalloc(&d, 3 * bytes) != hipSuccess) return fail(ctx, MOSE_ERR_NOMEM, "pow buffers");
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemcpyAsync(d, x, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d + n, y, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    const int64_t blocks = (n + 255) / 256;
    k_pow<<<(unsigned)(blocks < 65536 ? blocks : 65536), 256, 0, s>>>(d, d + n, d + 2 * n, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, d + 2 * n, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ctx, MOSE_ERR_HIP, hipGetErrorString(e));
  return MOSE_OK;
}

namespace {
constexpr int kChemTeamsPerWg = TeamShape<9>::per_wg;   // the kernels are instantiated for 9 species
struct ChemObj {
  ChemDev M{};
  void* tabs = nullptr;          // device copies of the mechanism tables (one allocation)
  double* state = nullptr;       // [NZ][cap]
  double* hr = nullptr;
  int* idid = nullptr;
  int* nstep = nullptr;
  unsigned long long* ctr = nullptr;   // [0] queue, [1..6] stats
  int64_t cap = 0, n = 0;
  int grid = 0, cus = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool valid = false;
};
std::vector<std::unique_ptr<ChemObj>> g_chem;   // per process; ids index this list (contexts share it)
}  // namespace

// RADAU5's ITOL = 1 tolerance transform and FNEWT (hairerRadau5.f:430-505; set_radau5_dev_tols,
// Lib_Radau5_dev.f90:71-107) and RADCOR's method constants (hairerRadau5.f:713-728), once per call on the host:
// k = [c1, c2, c1 - 1, c2 - 1, c1 - c2, dd1, dd2, dd3, 1/u1, alph/cno, beta/cno, FNEWT]
static void radau5_host_setup(int n, const double* rt, const double* at, double* rtolt, double* atolt, double* k) {
  const double expm = 2.0 / 3.0, uround = 1.1e-19;
  for (int i = 0; i < n; ++i) {
    const double quot = at[i] / rt[i];
    rtolt[i] = 0.1 * std::pow(rt[i], expm);
    atolt[i] = rtolt[i] * quot;
  }
  const double tolst = rtolt[0];
  const double sq6 = std::sqrt(6.0);
  const double c1 = (4.0 - sq6) / 10.0, c2 = (4.0 + sq6) / 10.0;
  const double u1 = (6.0 + std::pow(81.0, 1.0 / 3.0) - std::pow(9.0, 1.0 / 3.0)) / 30.0;
  const double alph = (12.0 - std::pow(81.0, 1.0 / 3.0) + std::pow(9.0, 1.0 / 3.0)) / 60.0;
  const double beta = (std::pow(81.0, 1.0 / 3.0) + std::pow(9.0, 1.0 / 3.0)) * std::sqrt(3.0) / 60.0;
  const double cno = alph * alph + beta * beta;
  k[0] = c1; k[1] = c2; k[2] = c1 - 1.0; k[3] = c2 - 1.0; k[4] = c1 - c2;
  k[5] = -(13.0 + 7.0 * sq6) / 3.0; k[6] = (-13.0 + 7.0 * sq6) / 3.0; k[7] = -1.0 / 3.0;
  k[8] = 1.0 / u1; k[9] = alph / cno; k[10] = beta / cno;
  k[11] = std::fmax(10 * uround / tolst, std::fmin(0.03, std::pow(tolst, 0.5)));
}

static ChemObj* get_chem(int id) {
  if (id < 0 || id >= (int)g_chem.size() || !g_chem[id] || !g_chem[id]->valid) return nullptr;
  return g_chem[id].get();
}

static void chem_free(ChemObj* o) {
  if (!o) return;
  if (o->tabs) (void)hipFree(o->tabs);
  if (o->state) (void)hipFree(o->state);
  for (auto& e : o->ev)
    if (e) (void)hipEventDestroy(e);
  o->tabs = nullptr; o->state = nullptr;
  o->valid = false;
}

extern "C" int mose_chem_create(mose_ctx* ctx, const mose_mechanism* m, int64_t ncell_max, int* chem_id) {
  if (!ctx || !m || !chem_id || ncell_max < 1) return MOSE_ERR_ARG;
  if (ncell_max > MOSE_CHEM_CELLS_MAX) return fail(ctx, MOSE_ERR_ARG, "mose_chem_create: at most 2^31 - 1 cells");
  if (m->ns != 9) return fail(ctx, MOSE_ERR_ARG, "the library instantiates the chemistry kernel for 9 species");
  if (m->tmax <= m->tmin || m->n_arrh < 0 || m->n_troe < 0 || m->n_lind < 0 || !m->wm || !m->cp || !m->h ||
      !m->ni1 || !m->ni2 || !m->eps || (m->n_arrh > 0 && (!m->kf || !m->kb)) || (m->n_troe > 0 && !m->troe) ||
      (m->n_lind > 0 && !m->lind))
    return fail(ctx, MOSE_ERR_ARG, "mose_chem_create: incomplete mechanism");
  if (m->n_arrh + m->n_troe + m->n_lind > NRMAX)
    return fail(ctx, MOSE_ERR_ARG, "mose_chem_create: the chemistry kernel takes at most 32 reactions");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto o = std::make_unique<ChemObj>();
  const int ns = m->ns, rows = m->tmax - m->tmin + 1, nr = m->n_arrh + m->n_troe + m->n_lind;
  size_t n_thermo = (size_t)ns * rows, n_st = (size_t)nr * (ns + 1), n_kf = (size_t)m->n_arrh * rows;
  size_t n_troe = (size_t)m->n_troe * 4 * rows, n_lind = (size_t)m->n_lind * 3 * rows;
  size_t total = 2 * n_thermo + 3 * n_st + 2 * n_kf + n_troe + n_lind + 8;
  HIPCHK(ctx, hipMalloc(&o->tabs, total * sizeof(double)));
  double* p = (double*)o->tabs;
  auto put = [&](const double* src, size_t cnt) -> double* {
    double* dst = p;
    if (cnt) (void)hipMemcpy(dst, src, cnt * sizeof(double), hipMemcpyHostToDevice);
    p += cnt;
    return dst;
  };
  ChemDev& M = o->M;
  M.ns = ns; M.na = m->n_arrh; M.nt = m->n_troe; M.nl = m->n_lind; M.tmin = m->tmin; M.rows = rows;
  for (int s = 0; s < ns; ++s) { M.wm[s] = m->wm[s]; M.Ri[s] = MOSE_RUNIV / m->wm[s]; }
  M.cp = put(m->cp, n_thermo); M.h = put(m->h, n_thermo);
  M.ni1 = put(m->ni1, n_st); M.ni2 = put(m->ni2, n_st); M.eps = put(m->eps, n_st);
  M.kf = put(m->kf, n_kf); M.kb = put(m->kb, n_kf);
  M.troe = put(m->troe, n_troe); M.lind = put(m->lind, n_lind);
  HIPCHK(ctx, hipGetLastError());
  int cus = 0;
  HIPCHK(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  // persistent grid of 256-lane workgroups (24 teams of 10 lanes, one cell per team at a time): enough to give
  // every SIMD several waves
  o->grid = (int)std::min<int64_t>((ncell_max + kChemTeamsPerWg - 1) / kChemTeamsPerWg, (int64_t)cus * 8);
  o->cus = cus;
  o->cap = ncell_max;
  const int nz = ns + 1;
  size_t sbytes = (size_t)nz * ncell_max * sizeof(double) + (size_t)ncell_max * (sizeof(double) + 2 * sizeof(int)) + 1024;
  HIPCHK(ctx, hipMalloc(&o->state, sbytes));
  o->hr = o->state + (size_t)nz * ncell_max;
  o->idid = (int*)(o->hr + ncell_max);
  o->nstep = o->idid + ncell_max;
  o->ctr = (unsigned long long*)(((uintptr_t)(o->nstep + ncell_max) + 255) & ~(uintptr_t)255);
  for (auto& e : o->ev) HIPCHK(ctx, hipEventCreate(&e));
  o->valid = true;
  g_chem.push_back(std::move(o));
  *chem_id = (int)g_chem.size() - 1;
  return MOSE_OK;
}

extern "C" int mose_chem_destroy(mose_ctx* ctx, int id) {
  ChemObj* o = get_chem(id);
  if (!ctx || !o) return MOSE_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  chem_free(o);
  return MOSE_OK;
}

extern "C" int mose_chem_upload(mose_ctx* ctx, int id, int64_t ncell, const double* state) {
  ChemObj* o = get_chem(id);
  if (!ctx || !o || !state || ncell < 1 || ncell > o->cap) return fail(ctx, MOSE_ERR_ARG, "mose_chem_upload: bad arguments");
  const int nz = o->M.ns + 1;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // [nz][ncell] host rows into [nz][cap] device rows
  for (int q = 0; q < nz; ++q)
    HIPCHK(ctx, hipMemcpyAsync(o->state + (size_t)q * o->cap, state + (size_t)q * ncell, ncell * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  o->n = ncell;
  return MOSE_OK;
}

extern "C" int mose_chem_download(mose_ctx* ctx, int id, double* state, double* hr, int32_t* idid, int32_t* nstep) {
  ChemObj* o = get_chem(id);
  if (!ctx || !o || !state) return MOSE_ERR_ARG;
  const int nz = o->M.ns + 1;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  for (int q = 0; q < nz; ++q)
    HIPCHK(ctx, hipMemcpyAsync(state + (size_t)q * o->n, o->state + (size_t)q * o->cap, o->n * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
  if (hr) HIPCHK(ctx, hipMemcpyAsync(hr, o->hr, o->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (idid) HIPCHK(ctx, hipMemcpyAsync(idid, o->idid, o->n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (nstep) HIPCHK(ctx, hipMemcpyAsync(nstep, o->nstep, o->n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MOSE_OK;
}

// The checks of an advance's options; on success A holds dt, the tolerances (RADAU5: transformed) and the constants.
static int chem_opts(mose_ctx* ctx, const ChemObj* o, double dt, const mose_ode_opts* opts, ChemArgs& A,
                     const char* who) {
  if (!opts || !(dt > 0.0)) return fail(ctx, MOSE_ERR_ARG, std::string(who) + ": bad arguments");
  if (opts->method != MOSE_ODE_RODAS && opts->method != MOSE_ODE_RADAU5)
    return fail(ctx, MOSE_ERR_ARG, std::string(who) + ": method must be MOSE_ODE_RODAS or MOSE_ODE_RADAU5");
  const int nz = o->M.ns + 1;
  for (int q = 0; q < nz; ++q)
    if (!(opts->rtol[q] > 0.0) || !(opts->atol[q] > 0.0))
      return fail(ctx, MOSE_ERR_ARG, std::string(who) + ": tolerances must be > 0");
  A.dt = dt;
  for (int q = 0; q < nz; ++q) { A.rtol[q] = opts->rtol[q]; A.atol[q] = opts->atol[q]; }
  if (opts->method == MOSE_ODE_RADAU5) radau5_host_setup(nz, opts->rtol, opts->atol, A.rtol, A.atol, A.rk);
  return MOSE_OK;
}

// One integration of A.ncell cells with the counters at ctr ([0] queue, [1..6] stats, cleared here by a kernel):
// the persistent launch, then the statistics into the report.  Allocates nothing.
static int chem_run(mose_ctx* ctx, ChemObj* o, ChemArgs& A, unsigned long long* ctr, int method,
                    mose_chem_report* rep) {
  hipStream_t s = ctx->stream;
  A.queue = ctr; A.stats = ctr + 1;
  k_zero_u64<<<1, 64, 0, s>>>(ctr, kScratchWords);
  int grid = (int)std::min<int64_t>((int64_t)o->cus * 8, (A.ncell + kChemTeamsPerWg - 1) / kChemTeamsPerWg);
  HIPCHK(ctx, hipEventRecord(o->ev[0], s));
  if (method == MOSE_ODE_RADAU5) k_radau<9><<<grid, 256, 0, s>>>(o->M, A);
  else k_chem<9><<<grid, 256, 0, s>>>(o->M, A);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(o->ev[1], s));
  unsigned long long st[7];
  HIPCHK(ctx, hipMemcpyAsync(st, ctr, sizeof st, hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  int64_t first_failed = st[6] ? (int64_t)(~0ull - st[6]) : -1;
  int32_t first_idid = 0;
  if (first_failed >= 0) {
    HIPCHK(ctx, hipMemcpyAsync(&first_idid, A.idid + first_failed, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
  }
  float ms = 0.f;
  HIPCHK(ctx, hipEventElapsedTime(&ms, o->ev[0], o->ev[1]));
  if (rep) {
    rep->cells = A.ncell;
    rep->steps_total = (int64_t)st[1];
    rep->rejected_total = (int64_t)st[2];
    rep->rhs_total = (int64_t)st[3];
    rep->failed = (int64_t)st[4];
    rep->steps_max = (int32_t)st[5];
    rep->first_failed_idid = first_idid;
    rep->first_failed_cell = first_failed;
    rep->ms = ms;
  }
  if (st[4] > 0)
    return fail(ctx, MOSE_ERR_ODE, std::to_string(st[4]) + " cells failed (idid < 0), the first is cell " +
                                       std::to_string(first_failed) + " with idid " + std::to_string(first_idid));
  return MOSE_OK;
}

extern "C" int mose_chem_advance(mose_ctx* ctx, int id, double dt, const mose_ode_opts* opts, mose_chem_report* rep) {
  ChemObj* o = get_chem(id);
  if (!ctx || !o) return fail(ctx, MOSE_ERR_ARG, "mose_chem_advance: bad arguments");
  if (o->n < 1) return fail(ctx, MOSE_ERR_ARG, "mose_chem_advance: no cells uploaded");
  ChemArgs A{};
  int rc = chem_opts(ctx, o, dt, opts, A, "mose_chem_advance");
  if (rc != MOSE_OK) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  A.state = o->state; A.hr = o->hr; A.idid = o->idid; A.nstep = o->nstep;
  A.ncell = o->n; A.stride = o->cap; A.prow = o->M.ns;
  return chem_run(ctx, o, A, o->ctr, opts->method, rep);
}

static bool is_device_ptr(const void* p, int device) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice && a.device == device;
}

extern "C" int mose_chem_advance_cells(mose_ctx* ctx, int id, int64_t ncell, int64_t stride, double* state,
                                       double* hr, int32_t* idid, int32_t* nstep, double dt,
                                       const mose_ode_opts* opts, mose_chem_report* rep) {
  ChemObj* o = get_chem(id);
  if (!ctx || !o) return fail(ctx, MOSE_ERR_ARG, "mose_chem_advance_cells: bad arguments");
  if (ncell < 1 || ncell > MOSE_CHEM_CELLS_MAX)
    return fail(ctx, MOSE_ERR_ARG, "mose_chem_advance_cells: ncell must be in [1, 2^31 - 1], got " +
                                       std::to_string(ncell));
  if (stride < ncell) return fail(ctx, MOSE_ERR_ARG, "mose_chem_advance_cells: stride < ncell");
  if (ctx->device >= kMaxDevices) return fail(ctx, MOSE_ERR_ARG, "mose_chem_advance_cells: device index too large");
  ChemArgs A{};
  int rc = chem_opts(ctx, o, dt, opts, A, "mose_chem_advance_cells");
  if (rc != MOSE_OK) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!is_device_ptr(state, ctx->device) || !is_device_ptr(hr, ctx->device) || !is_device_ptr(idid, ctx->device) ||
      !is_device_ptr(nstep, ctx->device))
    return fail(ctx, MOSE_ERR_ARG, "mose_chem_advance_cells: state, hr, idid and nstep must be device memory of the "
                                   "context's GPU");
  unsigned long long* ctr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ctr = g_scratch[ctx->device];
  }
  if (!ctr) return fail(ctx, MOSE_ERR_HIP, "mose_chem_advance_cells: no scratch address for this device");
  A.state = state; A.hr = hr; A.idid = idid; A.nstep = nstep;
  A.ncell = ncell; A.stride = stride; A.prow = o->M.ns;
  return chem_run(ctx, o, A, ctr, opts->method, rep);
}

// ----------------------------------------------------------------------------------------------------------------
// chemistry on a case's resident state: Chemistry_Newstate_blk (src/lib/physics/Lib_Chemistry.f90:33-76) on the
// interior cells of every block, in place on the var-last arena, and the Strang-split step around it
// ----------------------------------------------------------------------------------------------------------------
static int case_chem_launch(mose_ctx* ctx, Case& c, hipStream_t s) {
  ChemObj* o = get_chem(c.chem_id);
  if (!o) return fail(ctx, MOSE_ERR_ARG, "the case's chemistry object was destroyed");
  ChemArgs A = c.chem;
  A.state = c.P0; A.stride = c.NC; A.dtp = c.dt;
  A.queue = c.chem_ctr; A.stats = c.chem_ctr + 8;
  k_zero_u64<<<1, 64, 0, s>>>(c.chem_ctr, 1);
  const int grid = (int)std::min<int64_t>((int64_t)o->cus * 8, (A.ncell + kChemTeamsPerWg - 1) / kChemTeamsPerWg);
  if (c.chem_method == MOSE_ODE_RADAU5) k_radau<9><<<grid, 256, 0, s>>>(o->M, A);
  else k_chem<9><<<grid, 256, 0, s>>>(o->M, A);
  HIPCHK(ctx, hipGetLastError());
  return MOSE_OK;
}

extern "C" int mose_case_chem_attach(mose_ctx* ctx, int id, int chem_id, const mose_ode_opts* opts) {
  Case* c = get_case(ctx, id);
  if (!c) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_attach: bad case id");
  if (c->launched || c->pending) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_attach: a launch of the case awaits its poll");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(c->stream));
  if (chem_id < 0) {
    drop_graph(*c);
    c->n_strang = 1; c->strang = 1.0; c->chem_id = -1;
    return MOSE_OK;
  }
  ChemObj* o = get_chem(chem_id);
  if (!o) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_attach: no such chemistry object");
  if (c->sc.nsc != o->M.ns || c->sc.nrans != 0)
    return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_attach: the case's species must be the mechanism's (and no RANS)");
  if (c->ninterior > MOSE_CHEM_CELLS_MAX) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_attach: too many cells");
  ChemArgs A{};
  int rc = chem_opts(ctx, o, 1.0, opts, A, "mose_case_chem_attach");
  if (rc != MOSE_OK) return rc;
  const int64_t n = c->ninterior;
  if (!c->chem_rows) {
    // one allocation: rows [n] (int64), hr [n], idid [n], nstep [n], counters [16]
    const size_t bytes = (size_t)n * (8 + 8 + 4 + 4) + 256 + 16 * sizeof(unsigned long long);
    HIPCHK(ctx, hipMalloc(&c->chem_rows, bytes));
    HIPCHK(ctx, hipHostMalloc(&c->stage_chem, 8 * sizeof(unsigned long long), hipHostMallocDefault));
    std::vector<long long> rows;
    rows.reserve((size_t)n);
    for (const mose_blk& B : c->blk)
      for (int k = 1; k <= B.n3; ++k)
        for (int j = 1; j <= B.n2; ++j)
          for (int i = 1; i <= B.n1; ++i)
            rows.push_back(B.boff + (i + MOSE_GC - 1) + (long long)(j + MOSE_GC - 1) * B.pi +
                           (long long)(k + MOSE_GC - 1) * B.pi * B.pj);
    HIPCHK(ctx, hipMemcpy(c->chem_rows, rows.data(), rows.size() * sizeof(long long), hipMemcpyHostToDevice));
  }
  double* hr = (double*)(c->chem_rows + n);
  int32_t* idid = (int32_t*)(hr + n);
  int32_t* nstep = idid + n;
  c->chem_ctr = (unsigned long long*)(((uintptr_t)(nstep + n) + 255) & ~(uintptr_t)255);
  A.rows = c->chem_rows; A.prow = c->sc.nsc + 3; A.hr = hr; A.idid = idid; A.nstep = nstep; A.ncell = n;
  c->chem = A;
  c->chem_method = opts->method;
  c->chem_id = chem_id;
  c->n_strang = 2;
  c->strang = 0.5;
  drop_graph(*c);
  return MOSE_OK;
}

static void chem_report_fill(const Case& c, const unsigned long long* st, int32_t first_idid, mose_chem_report* rep) {
  rep->cells = c.ninterior;
  rep->steps_total = (int64_t)st[0];
  rep->rejected_total = (int64_t)st[1];
  rep->rhs_total = (int64_t)st[2];
  rep->failed = (int64_t)st[3];
  rep->steps_max = (int32_t)st[4];
  rep->first_failed_cell = st[5] ? (int64_t)(~0ull - st[5]) : -1;
  rep->first_failed_idid = first_idid;
  rep->ms = 0.0;
}

extern "C" int mose_case_chem_report(mose_ctx* ctx, int id, mose_chem_report* rep) {
  Case* c = get_case(ctx, id);
  if (!c || !rep) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_report: bad arguments");
  if (c->n_strang != 2) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_report: no chemistry attached");
  if (c->launched || c->pending) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_report: a launch awaits its poll");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int32_t first_idid = 0;
  const unsigned long long* st = c->stage_chem;
  if (st[5]) {
    HIPCHK(ctx, hipMemcpy(&first_idid, c->chem.idid + (~0ull - st[5]), sizeof(int32_t), hipMemcpyDeviceToHost));
  }
  chem_report_fill(*c, st, first_idid, rep);
  return MOSE_OK;
}

extern "C" int mose_case_chem_advance(mose_ctx* ctx, int id, double dt, mose_chem_report* rep) {
  Case* c = get_case(ctx, id);
  if (!c || !(dt > 0.0)) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_advance: bad arguments");
  if (c->n_strang != 2) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_advance: no chemistry attached");
  if (c->launched || c->pending) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_advance: a launch awaits its poll");
  ChemObj* o = get_chem(c->chem_id);
  if (!o) return fail(ctx, MOSE_ERR_ARG, "mose_case_chem_advance: the case's chemistry object was destroyed");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = c->stream;
  ChemArgs A = c->chem;
  A.state = c->P0; A.stride = c->NC; A.dtp = nullptr; A.dt = dt;
  A.queue = c->chem_ctr; A.stats = c->chem_ctr + 8;
  k_zero_u64<<<1, 64, 0, s>>>(c->chem_ctr, 16);
  const int grid = (int)std::min<int64_t>((int64_t)o->cus * 8, (A.ncell + kChemTeamsPerWg - 1) / kChemTeamsPerWg);
  HIPCHK(ctx, hipEventRecord(c->ev[1], s));
  if (c->chem_method == MOSE_ODE_RADAU5) k_radau<9><<<grid, 256, 0, s>>>(o->M, A);
  else k_chem<9><<<grid, 256, 0, s>>>(o->M, A);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(c->ev[2], s));
  HIPCHK(ctx, hipMemcpyAsync(c->stage_chem, c->chem_ctr + 8, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  const unsigned long long* st = c->stage_chem;
  int32_t first_idid = 0;
  if (st[5]) HIPCHK(ctx, hipMemcpy(&first_idid, A.idid + (~0ull - st[5]), sizeof(int32_t), hipMemcpyDeviceToHost));
  float ms = 0.f;
  HIPCHK(ctx, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
  if (rep) {
    chem_report_fill(*c, st, first_idid, rep);
    rep->ms = ms;
  }
  if (st[3] > 0)
    return fail(ctx, MOSE_ERR_ODE, std::to_string(st[3]) + " cells failed (idid < 0), the first is interior cell " +
                                       std::to_string(~0ull - st[5]) + " with idid " + std::to_string(first_idid));
  return MOSE_OK;
}
