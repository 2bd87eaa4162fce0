/*
 * mose_gpu.h -- C ABI of the MI355X-native time-step residual path of MOSE-GPU.
 *
 * The reference drives this path through Fortran module procedures called by Explicit_Step
 * (src/lib/numerics/time/explicit/Mod_Explicit.f90:155-294) on module-global device arrays
 * (src/lib/gpu/MOSE_Field_m.f90:42-72).  This header is the drop-in boundary that a Fortran host
 * binds with ISO_C_BINDING (mose-gpu_amd/fortran/mose_gpu_c.f90); every entry point cites the
 * reference procedure(s) it replaces.
 *
 * Conventions
 *   - every function returns an int status: MOSE_OK (0) or a negative MOSE_ERR_*; the process is
 *     never aborted.  mose_last_error() gives the message of the last failure of the context.
 *   - the context owns all device memory; host arrays are borrowed for the duration of a call.
 *   - one host thread per context; streams and graphs are internal.
 *   - FP64 throughout.  Indices in records are the reference's 1-based (i,j,k), ghosts 1-2..n+2.
 *
 * Host layouts (the reference's flat arrays, MOSE_Field_m.f90:42-57 and Lib_Metrics.f90:107-133)
 *   arena rows   NC = sum over blocks of (n1+4)(n2+4)(n3+4); block b starts at row boff[b] (blocks in
 *                order); row(b,i,j,k) = boff + (i+1) + (j+1)*(n1+4) + (k+1)*(n1+4)*(n2+4)
 *   state        P[s*NC + row], s = species 0..nsc-1, u, v, w, p, then the RANS variables (var-last)
 *   cell geometry dl[d*NC+row] (cell lengths), m[(a+3b)*NC+row] (metric dxi_a/dx_b), vol[row], yn[row]
 *   face geometry fm1[q*NF1 + f1] with q = nx, ny, nz, area and f1 = f1off[b] + i + (n1+1)*((j-1)+n2*(k-1)),
 *                i = 0..n1 (fm1(4,0:n1,n2,n3) of the reference); fm2 (j = 0..n2), fm3 (k = 0..n3) alike
 *   nodes        per block (n1+1)(n2+1)(n3+1) points, x,y,z interleaved, i fastest (the mesh file order)
 *   tables       cp, h, mu, k per species on T = 0..10000 K (10001 rows, FLINT's Tmin = 0 layout,
 *                lib/FLINT/src/lib/Load_ThermoTransport.f90:80-96): tab[s*10001 + T]
 */
#ifndef MOSE_GPU_H
#define MOSE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOSE_GC 2
#define MOSE_NTAB 10001
#define MOSE_MAX_SPECIES 16

enum {
  MOSE_OK = 0,
  MOSE_ERR_ARG = -1,       /* invalid argument / unsupported option */
  MOSE_ERR_HIP = -2,       /* a HIP runtime call failed */
  MOSE_ERR_NOMEM = -3,     /* device memory exhausted (or above the case's HBM budget) */
  MOSE_ERR_MESH = -4,      /* degenerate cell (det = 0 or volume <= 0) */
  MOSE_ERR_STATE = -5,     /* p < 0 or non-finite state after an update (State_Check_blk) */
  MOSE_ERR_THERMO = -6,    /* T outside (1, 9990) K (the thermo guard) */
  MOSE_ERR_NODEV = -7,     /* no GPU */
  MOSE_ERR_ODE = -8        /* a chemistry integration returned idid < 0 */
};

/* boundary record types (internal ids of Mod_BC.f90; MOSE types: 0/100 -> 0, 201/1xx -> 1, 300 -> 3, 400 -> 11,
 * 301 -> 5, 302 -> 6, 407 -> 4).  Type 0 is a face with no boundary condition (IO_BC_Open.f90:149-156): no flux is
 * computed there, and its ghost cells take the zero-gradient fill (Lib_Extrapolation.f90:590-607).  Type 3 is an
 * Euler symmetry plane (IO_BC_Open.f90:191-193): BC_Symmetry_Eul and, when viscous, BC_Symmetry_Visc
 * (Lib_BC.f90:640-861), behind the mirror ghost of Ghost_Symmetry (Lib_Extrapolation.f90:558-587). */
enum {
  MOSE_BC_NONE = 0,
  MOSE_BC_CONNECTION = 1,
  MOSE_BC_SYMMETRY = 3,
  MOSE_BC_FARFIELD = 4,
  MOSE_BC_WALL_ADIABATIC = 5,
  MOSE_BC_WALL_ISOTHERMAL = 6,
  MOSE_BC_EXTRAPOLATION = 11
};

/* limiter ids (Lib_Limiters.f90:21-61) and Riemann solver ids (Lib_Riemann.f90:23-107) */
enum { MOSE_LIM_MINMOD = 1, MOSE_LIM_VANLEER = 2, MOSE_LIM_VANALBADA = 3, MOSE_LIM_MC = 4, MOSE_LIM_SUPERBEE = 5, MOSE_LIM_FIRST = 6 };
enum { MOSE_RS_HLLC = 1, MOSE_RS_HLLE = 2, MOSE_RS_SLAU = 3, MOSE_RS_AUSMP = 4, MOSE_RS_LMROE = 12 };
/* integration variables of the state update (Assign_Integration_Variables, Lib_Newstate.f90:43-60): conservative
 * (Newstate_Conservative_blk) or primitive (Newstate_Primitive_blk, Lib_Newstate.f90:340-531, with Theta = (drho/dp)_T;
 * the preconditioned variant, integration-variables = prec, is not built) */
enum { MOSE_INT_CONS = 0, MOSE_INT_PRIM = 1 };

/* one boundary cell-face: bc_bi/bc_bj/bc_bk/bc_bf/bc_typ of Mod_BC.f90, the connection donor of
 * Lib_Extrapolation.f90:1023-1135 (conn_is/js/ks/fs, conn_d11..d22) and the wall / far-field data */
typedef struct {
  int32_t b, i, j, k, f, typ;      /* block (0-based), cell, face 1..6, type */
  int32_t bs, is, js, ks, fs;      /* connection donor block (0-based), cell and face */
  int32_t d11, d12, d21, d22;      /* transverse axis map (identity: 1,0,0,1) */
  int32_t periodic;                /* MOSE 201: ghost nodes translated by one period (metrics only) */
  int32_t ff;                      /* type 4: row of the case's far-field table (mose_farfield) */
  int32_t pad;
  double tw;                       /* isothermal wall temperature (bc_tw) */
  double tvec[3];                  /* type 4: unit tangent of the face, the node diagonal (bc_tvec, Mod_BC.f90:228-240);
                                      mose_metrics_build fills it from the nodes */
} mose_bc_record;

/* The data of a far-field / inflow-outflow face (type 4; MOSE_open 407 -> legacy 4, IO_BC_Open.f90:302-311):
 * bc_mach, bc_t0, bc_p0, bc_pamb, bc_alpha, bc_beta, bc_relfac, bc_ci (mass fractions), bc_rans (nu~ per unit
 * density) of Mod_BC.f90:212-224.  mach >= 0: stagnation inlet (p0, T0) / back-pressure outlet (pamb). */
#define MOSE_MAX_SPECIES 16
typedef struct {
  double mach, t0, p0, pamb, alpha, beta, rel_fac, reserved;
  double ci[MOSE_MAX_SPECIES];
  double rans[2];
} mose_farfield;

typedef struct {
  int32_t nsc;          /* species (nsc >= 1) */
  int32_t nrans;        /* 0, or 1 for Spalart-Allmaras */
  int32_t limiter;      /* MOSE_LIM_* */
  int32_t riemann;      /* MOSE_RS_* */
  int32_t viscous;      /* 0 euler, 1 navier-stokes */
  int32_t n_rk;         /* 1..3 (SSP-RK3 = 3) */
  int32_t ndir;         /* 1, 2 or 3 computational directions */
  int32_t point_implicit;
  int32_t thermo_guard; /* 1: T outside (1, 9990) K is an error (MOSE_ERR_THERMO) */
  int32_t integration;  /* MOSE_INT_CONS (default) or MOSE_INT_PRIM */
  int32_t irs;          /* 1: implicit residual smoothing before the update (IRS_apply, Lib_IRS.f90:53-96); plain path */
  int32_t irs_njacobi; /* its Jacobi sweeps per direction (the reference's 2) */
  int32_t reserved[4];
  double mco;           /* riemann-options-Mco (LMRoe) */
  double cfl, vnn;
  double prandtl;       /* > 0: k = mu cp / Pr */
  double sc, sct, prt;  /* Schmidt, turbulent Schmidt and Prandtl numbers */
  double irs_beta;      /* the smoothing coefficient (irs-beta) */
  double reserved_d[7];
} mose_scheme;

typedef struct {
  int32_t nsc;
  int32_t pad;
  const double* wm;     /* [nsc] molecular weights kg/kmol */
  const double* cp;     /* [nsc*MOSE_NTAB] */
  const double* h;
  const double* mu;     /* NULL when inviscid */
  const double* k;
} mose_tables;

typedef struct {
  double dt;            /* the global time step of the last step */
  double time;          /* accumulated time of the case */
  int64_t steps;        /* steps done by the case */
  int32_t bad_cells;    /* cells failing the state check / thermo guard in the last call */
  int32_t first_bad_block, first_bad_i, first_bad_j, first_bad_k;
  int32_t status;
} mose_report;

typedef struct mose_ctx mose_ctx;

/* Context: device selection and the FLINT tables (flint_acc_upload_thermo, Lib_Radau5_dev.f90:118-159;
 * cuf_init_tables, Mod_Convective.f90:59-76). */
int mose_ctx_create(int device, const mose_tables* tables, mose_ctx** out);
int mose_ctx_destroy(mose_ctx* ctx);
const char* mose_last_error(const mose_ctx* ctx);
int mose_device_count(int* n);
/* Device facts the host sizes work by: HBM bytes (total, free), compute units, LDS bytes a workgroup may take.
 * No reference counterpart (the serving loop admits cases by HBM budget). */
int mose_ctx_info(mose_ctx* ctx, int64_t* hbm_total, int64_t* hbm_free, int* cu_count, int* lds_per_block);

/* Geometry of a multi-block mesh from its nodes (Setup_Metrics, Lib_Metrics.f90:15-136).  Host only.
 * yn may be NULL; ndir as Check_Mesh_Type (Lib_Metrics.f90:139-196). */
int mose_metrics_build(int nb, const int* dims, const double* nodes, int nrec, const mose_bc_record* rec, int ndir,
                       double* dl, double* m, double* vol, double* yn, double* fm1, double* fm2, double* fm3);

/* The wall distance yn[NC] of every interior cell (Compute_Yn, Lib_Metrics.f90:513-744): the distance from the cell
 * centre (the mean of its eight nodes) to the nearest centroid of a wall face (records of types 5 and 6; huge when
 * there is none), then the first ghost layer of every connection face from its donor cell (Yn_Connection,
 * Lib_Metrics.f90:747-763).  Other rows are left as they are.  Host only. */
int mose_wall_distance(int nb, const int* dims, const double* nodes, int nrec, const mose_bc_record* rec, double* yn);

/* The unit tangent of every far-field record (type 4) from the nodes (Setup_BC_Device, Mod_BC.f90:228-240).
 * Host only; writes rec[e].tvec. */
int mose_records_tangent(int nb, const int* dims, const double* nodes, int nrec, mose_bc_record* rec);

/* A case: blocks, geometry and boundary records made device resident (Setup_Device_Residency,
 * Mod_Explicit.f90:404-482; Setup_BC_Device).  Records must be grouped by block, then by face 1..6
 * (the launch order of BC_Fluxes, Mod_BC.f90:1042-1062).  Returns the case id in *case_id. */
int mose_case_create(mose_ctx* ctx, const mose_scheme* scheme, int nb, const int* dims, const double* dl,
                     const double* m, const double* vol, const double* yn, const double* fm1, const double* fm2,
                     const double* fm3, int nrec, const mose_bc_record* rec, int nff, const mose_farfield* ff,
                     int* case_id);
int mose_case_destroy(mose_ctx* ctx, int case_id);
/* bytes of device memory a case of these blocks takes (the serving loop admits by this figure) */
int mose_case_bytes(const mose_scheme* scheme, int nb, const int* dims, int nrec, int64_t* bytes);

/* State in and out (pack_P_to_flat / unpack_flat_to_P, MOSE_Field_m.f90:171-194,334-369). */
int mose_state_upload(mose_ctx* ctx, int case_id, const double* P);
int mose_state_download(mose_ctx* ctx, int case_id, double* P);

/* nsteps explicit steps of one case (Explicit_Step, Mod_Explicit.f90:52-397, time accurate without
 * chemistry): per step the global dt (Compute_dt, Mod_dt.f90:15-243), PO <- P, then n_RK stages of
 * ghosts + halo (Fill_Ghost_Cell, Lib_Extrapolation.f90:1370-1524), fluxes (Fluxes, Mod_Fluxes.f90:10;
 * BC_Fluxes, Mod_BC.f90:646) and the conservative update (Newstate_Conservative_blk, Lib_Newstate.f90:257).
 * The steps run on the case's own stream; the call returns when they are done, with the report. */
int mose_case_step(mose_ctx* ctx, int case_id, int nsteps, mose_report* report);

/* The same on several cases at once: the steps of every case are queued on its own stream before any is waited
 * on, so the cases' kernels overlap on the device; each case keeps its own dt.  (Within one case, all of its
 * blocks already share one launch per stage, the block pack of MOSE_Pack_m.f90:70-251.) */
int mose_cases_step(mose_ctx* ctx, const int* case_ids, int ncases, int nsteps, mose_report* reports);

/* Asynchronous stepping for a serving loop.  mose_case_launch queues nsteps steps of the case on its stream and
 * returns at once; mose_case_poll returns at once too, with done = 0 while they run and done = 1 (and the report of
 * those steps, as mose_case_step gives it) when they have finished.  One launch per case may be in flight; poll it
 * to done before any other call on the case.  No reference counterpart (the serving layer). */
int mose_case_launch(mose_ctx* ctx, int case_id, int nsteps);
int mose_case_poll(mose_ctx* ctx, int case_id, int* done, mose_report* report);

/* Residual of the current state (ghosts, fluxes and boundary fluxes of one stage, no update): R[s*NC+row]
 * on the interior rows.  For checks against Fluxes + BC_Fluxes. */
int mose_case_residual(mose_ctx* ctx, int case_id, double* R);

/* Global dt of the current state (Compute_dt + the min reduction). */
int mose_case_dt(mose_ctx* ctx, int case_id, double* dt);

/* The L2 norm of the last step's change, per variable (Compute_Residual_blk, Mod_Diagnostic.f90:63-106, then the
 * sqrt of Mod_Explicit.f90:358): out[q] = sqrt(sum over the interior cells of d_q^2), d_0 = the sum of the species
 * changes, d_1.. = the changes of u, v, w, p and the RANS variables; nprim - nsc + 1 values.  The last stage of
 * every step leaves per-workgroup partial sums; this call adds them in a fixed order, so it is deterministic. */
int mose_case_residual_norm(mose_ctx* ctx, int case_id, double* out);

/* field_state_hash (MOSE_Field_m.f90:254-284): h_lo = the sum of the low 32 bits and h_hi = the sum of the high
 * 32 bits of the IEEE patterns of every interior value (all variables), as wrapping 64-bit integers.
 * Order-independent: equal states give equal hashes on every launch order. */
int mose_case_state_hash(mose_ctx* ctx, int case_id, uint64_t* h_lo, uint64_t* h_hi);

/* Execution path and tile shape of a case.  path: MOSE_PATH_FUSED (default) runs one fused stage kernel per RK
 * stage over the tiles of every block; MOSE_PATH_PLAIN runs the per-block face kernels + cell update (the
 * per-block launch structure of the reference, kept as the comparison path); MOSE_PATH_FAST is the fused path in
 * its performance formulation: the same discretisation with each tile's primitives staged in LDS once, fused
 * multiply-adds, Newton-refined reciprocals and a shorter form of the Roe (LMRoe) flux; it differs from the other
 * two by rounding only (relative 1e-12 on the parity cases, except where a change of rounding alone moves a case
 * further: Noh87's cold gas moves by ~2e-11 under contraction), not bit for bit.
 * A tile holds 128, 256 or 512 cells
 * (ti * tj * tk); ti = tj = tk = 0 autotunes the tile (each candidate timed on the case itself, the state left
 * unchanged).  timing = 1 records HIP events around
 * every stage kernel of the following step calls (mose_case_timing).  No reference counterpart (launch tuning). */
enum { MOSE_PATH_FUSED = 0, MOSE_PATH_PLAIN = 1, MOSE_PATH_FAST = 2 };
typedef struct {
  int32_t path;
  int32_t ti, tj, tk;
  int32_t timing;
  int32_t no_graph;     /* 1: launch every fused step eagerly instead of replaying the captured step graph */
  int32_t reserved[2];
} mose_tuning;
int mose_case_tune(mose_ctx* ctx, int case_id, const mose_tuning* tuning, mose_tuning* chosen);

/* Device time of the stage kernels of the last timed step call: average per launch, total, launch count. */
int mose_case_timing(mose_ctx* ctx, int case_id, double* ms_stage_avg, double* ms_stage_total, int64_t* launches);

/* Stream-copy probe: copies `bytes` (read `bytes` + write `bytes`) `reps` times with the library's copy kernel and
 * returns the bandwidth in GB/s (read + written bytes over the kernel time).  Calibrates the HBM traffic counters
 * against a kernel of known traffic and measures the box's achievable bandwidth.  No reference counterpart. */
int mose_bandwidth_probe(mose_ctx* ctx, int64_t bytes, int reps, double* gbs);

/* ---------------------------------------------------------------------------------------------------------
 * Stiff chemistry (FLINT): per-cell constant-volume ODE integration over dt of independent cells
 * (Chemistry_Newstate_blk, src/lib/physics/Lib_Chemistry.f90:33-76) with RODAS (hairerRodas.f ROSCOR, METH = 1,
 * the host integrator wrap_rodas, OSlo.f90:485-527) or RADAU5 (hairerRadau5.f RADCOR, wrap_radau5,
 * OSlo.f90:349-412).
 * --------------------------------------------------------------------------------------------------------- */

/* A reaction mechanism in FLINT's table form (read_chemistry, Load_Chemistry.f90:7-164; the general loop of
 * Lib_Chemistry_wdot.f90:163-288).  Every rate table has one row per Kelvin on [tmin, tmax] (row T at
 * index T - tmin), like the thermo tables.  Per reaction r of each class, ni1/ni2/eps hold ns + 1 values:
 * the reactant and product stoichiometric coefficients and third-body efficiencies of every species, and in
 * slot ns the third-body flag (ni1[ns] > 0: the rate is multiplied by sum_s eps_s c_s).  Classes are stored
 * one after the other (Arrhenius, Troe, Lindemann). */
typedef struct {
  int32_t ns;             /* species, <= MOSE_MAX_SPECIES */
  int32_t n_arrh, n_troe, n_lind;
  int32_t tmin, tmax;     /* table rows tmin..tmax (FLINT: 0..10000) */
  const double* wm;       /* [ns] kg/kmol */
  const double* cp;       /* [ns][rows] J/(kg K) */
  const double* h;        /* [ns][rows] J/kg, formation enthalpy included */
  const double* ni1;      /* [n_arrh + n_troe + n_lind][ns + 1] */
  const double* ni2;      /* same shape */
  const double* eps;      /* same shape */
  const double* kf;       /* [n_arrh][rows] forward rate constants (SI: kmol, m^3, s) */
  const double* kb;       /* [n_arrh][rows] reverse rate constants */
  const double* troe;     /* [n_troe][4][rows]: k_inf, k_0, K_c, F_cent */
  const double* lind;     /* [n_lind][3][rows]: k_inf, k_0, K_c */
} mose_mechanism;

/* RODAS: hairerRodas.f ROSCOR (METH = 1) as wrap_rodas runs it (OSlo.f90:485-527).
 * RADAU5: hairerRadau5.f RADCOR as wrap_radau5 runs it (OSlo.f90:349-412: IJAC = 0, ITOL = 1), in the statement
 * order of the device radau5_dev (Lib_Radau5_dev.f90:199-652). */
enum { MOSE_ODE_RODAS = 1, MOSE_ODE_RADAU5 = 2 };
typedef struct {
  int32_t method;         /* MOSE_ODE_RODAS or MOSE_ODE_RADAU5 */
  int32_t reserved[3];
  double rtol[MOSE_MAX_SPECIES + 1];   /* per ODE component (species..., T): raw tolerances, as the wrappers take
                                          them (RADAU5 applies its ITOL = 1 transform itself) */
  double atol[MOSE_MAX_SPECIES + 1];
} mose_ode_opts;

typedef struct {
  int64_t cells;
  int64_t failed;         /* cells with idid < 0 */
  int64_t steps_total, rejected_total, rhs_total;
  int32_t steps_max, first_failed_idid;   /* the idid of first_failed_cell (0 if none failed) */
  int64_t first_failed_cell;   /* the lowest-index cell that ended with idid < 0, -1 if none */
  double ms;              /* device time of the integration kernel */
} mose_chem_report;

/* A batch of independent chemistry cells resident on the device: state [ns + 1][ncell] (species densities,
 * then pressure), the layout of the flow arena's chemistry columns. */
int mose_chem_create(mose_ctx* ctx, const mose_mechanism* mech, int64_t ncell_max, int* chem_id);
int mose_chem_destroy(mose_ctx* ctx, int chem_id);
int mose_chem_upload(mose_ctx* ctx, int chem_id, int64_t ncell, const double* state);
int mose_chem_download(mose_ctx* ctx, int chem_id, double* state, double* hr, int32_t* idid, int32_t* nstep);
/* Every cell: T = p / (rho R), RODAS or RADAU5 over [0, dt] on (rho_1..rho_ns, T), species floored at 1e-20,
 * p = T rho R (Chemistry_Newstate_blk); hr = |T_new - T_old| / T_old.  One persistent launch: each team of ns + 1
 * lanes takes the next cell from an atomic work queue when its cell is done, so cells of uneven stiffness share the
 * device evenly.  The controllers' real powers (ERR**0.25, FACCON**0.8, QNEWT**p) and the Troe factor's 10**x are
 * glibc's pow bit for bit (mose_pow_batch), so the step sequences are those of the reference's libm-linked Fortran. */
int mose_chem_advance(mose_ctx* ctx, int chem_id, double dt, const mose_ode_opts* opts, mose_chem_report* report);

/* The largest cell count of one chemistry batch: the work queue hands out 32-bit cell indices. */
#define MOSE_CHEM_CELLS_MAX 2147483647LL

/* mose_chem_advance on cells that the caller holds in device memory, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  state = [ns + 1][stride] doubles (species densities, then pressure;
 * cell c at column c, 1 <= ncell <= stride), hr = [ncell] doubles, idid and nstep = [ncell] int32, all device
 * memory of the context's GPU (host pointers are refused); chem_id supplies the mechanism (mose_chem_create) and
 * its batch buffer is not touched.  Same kernels, same results as mose_chem_advance on the same cells.
 * The work queue and the statistics live in a 64-byte file-scope device array of the library, one per GPU, held
 * until the process ends and cleared by each call: calls that overlap on one GPU race.  ncell above
 * MOSE_CHEM_CELLS_MAX is refused (MOSE_ERR_ARG).  The call does not call hipMalloc, hipFree or any allocating
 * routine; it waits for its own launch (the report's counters). */
int mose_chem_advance_cells(mose_ctx* ctx, int chem_id, int64_t ncell, int64_t stride, double* state, double* hr,
                            int32_t* idid, int32_t* nstep, double dt, const mose_ode_opts* opts,
                            mose_chem_report* report);

/* Chemistry on a flow case (nsc = the mechanism's ns, no RANS): its steps become Strang-split steps of
 * Explicit_Step (Mod_Explicit.f90:155-294, set_StrangSplitting Lib_Strang.f90:10-22): half a transport step
 * (strangcoeff 1/2), Chemistry_Newstate_blk (Lib_Chemistry.f90:33-76) over the step's whole dt on the case's resident
 * state, the other half.  The mechanism is chem_id's (mose_chem_create), the integrator and tolerances opts'.
 * chem_id < 0 detaches: plain transport steps again.  Cells whose integration ends with idid < 0 make the step call
 * return MOSE_ERR_ODE (mose_case_chem_report gives the counts). */
int mose_case_chem_attach(mose_ctx* ctx, int case_id, int chem_id, const mose_ode_opts* opts);
/* Chemistry_Newstate_blk alone, in place on the interior cells of the case's resident state, over dt (host value):
 * the reference's per-block call (Lib_Chemistry.f90:54-74) on every block, one persistent launch. */
int mose_case_chem_advance(mose_ctx* ctx, int case_id, double dt, mose_chem_report* report);
/* the chemistry counts of the last mose_case_step / mose_cases_step / poll of a case with chemistry attached */
int mose_case_chem_report(mose_ctx* ctx, int case_id, mose_chem_report* report);

/* out[i] = x[i]**y[i] on the device with the chemistry kernels' pow (host arrays of n values): the bits of the host
 * libm pow that the reference's Fortran `**` calls (glibc 2.35 on x86-64 with FMA; hairerRodas.f:805,816,
 * hairerRadau5.f:918,942,969, Lib_Chemistry_falloff.f90:78).  Exposed so that a host can check the device's
 * powers against its own libm. */
int mose_pow_batch(mose_ctx* ctx, int64_t n, const double* x, const double* y, double* out);

#ifdef __cplusplus
}
#endif
#endif
