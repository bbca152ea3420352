/*
 * orctm_nh.h -- C ABI of the MI355X (gfx950) nonhydrostatic pressure Poisson solver.
 *
 * Drop-in for ORCTM's PETSc path (ORCTM/src_orctm/mo_petsc.f90, module MO_PETSC):
 *
 *   reference call                                   this ABI
 *   -----------------------------------------------  -------------------------------------------
 *   call petsc_start      (mo_petsc.f90:41-128,       orctm_nh_create()
 *                          called orctm.f90:970)
 *   compute_matrix        (mo_petsc.f90:543-782)      done once inside orctm_nh_create()
 *   call petsc_solve_3D   (mo_petsc.f90:211-233,      orctm_nh_solve()
 *                          called ocvtot_nh.f90:150)
 *   ComputeRHS            (mo_petsc.f90:237-397)      done inside orctm_nh_solve()/orctm_nh_set_rhs()
 *   call petsc_end        (mo_petsc.f90:131-143)      orctm_nh_destroy()
 *   MatMult (PETSc AIJ, used inside KSPSolve)        orctm_nh_apply()
 *
 * All arrays crossing the boundary are the caller's Fortran arrays, column-major,
 * including the one-cell halo ring of ORCTM's local arrays (mo_commo1.f90:105-146).
 * Every entry point returns an int status (ORCTM_NH_OK == 0) and never aborts.
 * No HIP or torch types appear in the signatures; one handle = one system on one GPU.
 */
#ifndef ORCTM_NH_H
#define ORCTM_NH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORCTM_NH_ABI_VERSION 1

/* Status codes. The reference ignores PETSc's ierr and never queries the converged
 * reason (mo_petsc.f90:219, no KSPGetConvergedReason); here every outcome is reported. */
enum orctm_nh_status {
  ORCTM_NH_OK = 0,              /* converged to rtol */
  ORCTM_NH_NOT_CONVERGED = 1,   /* maxit / max_outer reached; pnh holds the last iterate */
  ORCTM_NH_BREAKDOWN = 2,       /* Krylov breakdown (p.Ap <= 0 or r.z <= 0 or non-finite scalar) */
  ORCTM_NH_ERR_NONFINITE = 3,   /* NaN/Inf in the right-hand side; pnh left untouched */
  ORCTM_NH_ERR_INVALID = 4,     /* invalid geometry or arguments */
  ORCTM_NH_ERR_NOMEM = 5,       /* device allocation failed or admission refused */
  ORCTM_NH_ERR_HIP = 6,         /* HIP runtime error */
  ORCTM_NH_ERR_NODEVICE = 7,    /* no usable gfx950 device */
  ORCTM_NH_PENDING = 8          /* serving: the request has not finished yet */
};

/* Precision modes. */
enum orctm_nh_precision {
  ORCTM_NH_FP64 = 0,   /* FP64 operator, FP64 PCG */
  ORCTM_NH_MIXED = 1   /* FP32 operator copy + FP32 PCG (FP32 reductions) inside FP64 iterative refinement */
};

/* Operator inputs. Field meanings follow the reference exactly:
 *   ie, je, ke  : IE, JE, KE of mo_param1.f90 -- halo'd local sizes. The Poisson unknowns are the
 *                 interior (IE-2) x (JE-2) x KE cells, the DMDA of mo_petsc.f90:48-50,75-95.
 *   i_start     : lower bound of u-point arrays, DLXU(I_start:IE,JE) (mo_commo1.f90:105; 0 on the
 *                 west-most rank, mo_parallel.f90:197-201).
 *   j_start     : lower bound of v-point arrays, DLYV(IE,J_start:JE) (mo_commo1.f90:106).
 *   dlxp, dlyp  : DLXP(IE,JE), DLYP(IE,JE)  grid spacings at p-points  [m]
 *   dlxu        : DLXU(I_start:IE,JE)        x spacing at u-points      [m]
 *   dlyv        : DLYV(IE,J_start:JE)        y spacing at v-points      [m]
 *   ddpo        : DDPO(IE,JE,KE)  p-cell thickness incl. partial bottom cells, 0 below (boden.f90:186-190)
 *   weto        : WETO(IE,JE,KE)  land/sea mask, 1 = wet (boden.f90:185)
 *   dzw         : DZW(KE)         layer thicknesses (beleg.f90:87-105)
 *   cyclic_x/y  : CYCLICWE / CYCLICNS (mo_param1.f90:62-77; periodic DMDA, mo_petsc.f90:75-86)
 *   open_w/e/n/s: MO_OBCS West/East/North/South (mo_obcs.f90:9-10); j grows southward, north = J-1.
 * The library copies what it needs during orctm_nh_create(); the pointers are not kept. */
typedef struct orctm_nh_grid {
  int32_t ie, je, ke;
  int32_t i_start, j_start;
  const double* dlxp;
  const double* dlyp;
  const double* dlxu;
  const double* dlyv;
  const double* ddpo;
  const double* weto;
  const double* dzw;
  int32_t cyclic_x, cyclic_y;
  int32_t open_w, open_e, open_n, open_s;
} orctm_nh_grid_t;

/* Solver options. orctm_nh_default_options() fills the defaults listed here. */
typedef struct orctm_nh_options {
  int32_t precision;    /* ORCTM_NH_MIXED (default) or ORCTM_NH_FP64 */
  int32_t maxit;        /* cap on total inner Krylov iterations per solve (default 10000, the KSP default) */
  int32_t max_outer;    /* cap on FP64 refinement steps per solve (default 40) */
  int32_t warm_start;   /* 1 (default): after the first solve, start from the previous solution,
                           as KSPSetInitialGuessNonzero(ksp,PETSC_TRUE) does (mo_petsc.f90:232) */
  double rtol;          /* outer stop: ||b - A p||_2 <= rtol ||b||_2, FP64, reference operator (default 1e-10) */
  double rtol_inner;    /* inner relative reduction per refinement step (default 1e-5 MIXED, 0 = as needed FP64) */
  int32_t check_every;  /* inner iterations per device-side chunk between host polls (default 16) */
  int32_t device;       /* HIP device ordinal (default 0) */
  int32_t profile;      /* 1: time the inner kernels with HIP events (orctm_nh_profile_read); default 0 */
  int32_t tile_rows;    /* rows (j) per tile of the inner-iteration kernels: 0 (default) = tuned per size
                           class at create / batch create; 4, 8 or 12 = fixed. Results are the same bits
                           for every choice (block sums are formed over fixed 4-row groups). Tuned picks
                           persist per device (arch/CU count), precision and size class in a JSON file:
                           $ORCTM_NH_TUNE_FILE (set but empty: none), else
                           $XDG_CACHE_HOME/orctm_nh/tile_tuning.json, else
                           $HOME/.cache/orctm_nh/tile_tuning.json; later processes skip the timing. */
  int32_t bottom_solve; /* 1: ORCTM's CG2D4NONHY option in orctm_nh_divergence: the 2-D bottom-pressure solve
                           and the corrections of DIVG by PNH_bot before the right-hand side is formed
                           (ocvtot_nh.f90:64-147, mo_petsc.f90:147-208, 252-374, 401-538). Default 0, as
                           the shipped Makefile builds (-DZZCG2D4NONHY). */
  int32_t zw_table;     /* where the iteration kernels read ZW/ZP (stats.zw_table reports the choice):
                           0 (default) = from the per-level table and per-column window when they
                           reproduce every column (checked at create) and the solve launches cover at
                           least 2^20 cells (one system, or a batch in total), else the per-cell planes;
                           1 = the table whenever it reproduces every column; -1 = always the planes.
                           Results are the same bits for every choice. */
  int32_t buffer_addressing; /* how the matvec kernel addresses memory (stats.buffer_addressing reports the
                           choice): 0 (default) = through buffer descriptors, with masked lanes and
                           wavefronts turned away by the hardware's range check instead of branches,
                           when every inner array of the solve launches holds at most 2^31 bytes (a
                           system of at most 2^29 padded cells in FP32, 2^28 in FP64), else through
                           64-bit pointers; -1 = always 64-bit pointers. Same bits either way. */
  int32_t reserved[1];
} orctm_nh_options_t;

/* Per-solve statistics. */
typedef struct orctm_nh_stats {
  int32_t status;       /* enum orctm_nh_status of this solve */
  int32_t outer_iters;  /* FP64 refinement steps taken */
  int32_t inner_iters;  /* total inner PCG iterations */
  int32_t tile_rows;    /* tile height the inner-iteration kernels ran with for this system */
  double rel_res;       /* final ||b - A p||_2 / ||b||_2 in FP64 with the reference operator */
  double bnorm;         /* ||b||_2 */
  double solve_ms;      /* device time of the solve (HIP events), excluding host<->device copies */
  double total_ms;      /* host wall time of the call, including copies */
  int32_t zw_table;     /* 1: the iteration kernels read ZW/ZP from the wet column's per-level table and
                           each column's 3-level window around its bottom (checked bit for bit at create);
                           0: from the per-cell planes */
  int32_t buffer_addressing; /* 1: the matvec kernel addressed memory through buffer descriptors */
  double reserved[3];
} orctm_nh_stats_t;

typedef struct orctm_nh orctm_nh_t;

/* Library and device information. */
int orctm_nh_abi_version(void);
int orctm_nh_device_count(int32_t* n);
void orctm_nh_default_options(orctm_nh_options_t* opts);
const char* orctm_nh_status_string(int status);
/* Text of the last HIP failure seen by the calling thread (empty if none). */
const char* orctm_nh_last_error(void);

/* Replaces petsc_start (mo_petsc.f90:41-128) + compute_matrix (mo_petsc.f90:543-782):
 * validates the grid, uploads the metric/mask arrays and assembles the operator on the device,
 * once, as KSPSetUp assembles it once. opts may be NULL (defaults). */
int orctm_nh_create(const orctm_nh_grid_t* grid, const orctm_nh_options_t* opts, orctm_nh_t** handle);

/* Replaces petsc_solve_3D (mo_petsc.f90:211-233) with its ComputeRHS callback (mo_petsc.f90:376):
 * reads DIVG(IE,JE,KE) (interior x WETO is the right-hand side), solves A p = b and writes the
 * interior of PNH(IE,JE,KE) only (mo_petsc.f90:223-227); the halo of pnh is not touched.
 * Initial guess: the handle's previous solution when warm_start is set (KSPSetInitialGuessNonzero),
 * zero on the first call. stats may be NULL. */
int orctm_nh_solve(orctm_nh_t* h, const double* divg, double* pnh, orctm_nh_stats_t* stats);

/* Replaces petsc_end (mo_petsc.f90:131-143). Frees all device memory of the handle. */
int orctm_nh_destroy(orctm_nh_t* h);

/* Split form of orctm_nh_solve for callers that keep data resident on the device
 * (serving loops, benchmarks):  set_rhs -> solve_resident -> get_solution. */
int orctm_nh_set_rhs(orctm_nh_t* h, const double* divg);            /* ComputeRHS, mo_petsc.f90:376 */
int orctm_nh_set_guess(orctm_nh_t* h, const double* pnh);           /* PNH interior as initial guess */
int orctm_nh_solve_resident(orctm_nh_t* h, orctm_nh_stats_t* stats); /* KSPSolve on resident b, p */
int orctm_nh_get_solution(orctm_nh_t* h, double* pnh);              /* mo_petsc.f90:223-227 */

/* y = A x with the reference operator in FP64 (the AIJ MatMult that KSPSolve applies to the matrix
 * of compute_matrix). x, y: interior vectors, (IE-2)*(JE-2)*KE doubles in DMDA natural order
 * (i fastest, then j, then k) -- the order of res_array in mo_petsc.f90:222-227. */
int orctm_nh_apply(orctm_nh_t* h, const double* x, double* y);

/* orctm_nh_apply without any device allocation: no counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The same y = A x, the same bits. x and y are staged in a fixed device array
 * of the library (128 MiB on each GPU it runs on, held until the process ends), so a system may have
 * at most 8388608 padded cells ((IE-2) rounded up to 32, times (JE-2), times KE); larger ones return
 * ORCTM_NH_ERR_INVALID before any device work. Every call on one GPU shares that array: calls that
 * overlap on one GPU (from several threads) race and must be serialised by the caller. */
int orctm_nh_apply_noalloc(orctm_nh_t* h, const double* x, double* y);

/* Reads back the assembled reference operator: 7 planes of (IE-2)*(JE-2)*KE doubles, in the order
 * centre, west(i-1), east(i+1), north(j-1), south(j+1), up(k-1), down(k+1); absent couplings are 0.
 * These are the MatSetValuesStencil values of mo_petsc.f90:623-765. */
int orctm_nh_get_operator(orctm_nh_t* h, double* coef7);

/* Size of the interior: nx = IE-2, ny = JE-2, nz = KE. */
int orctm_nh_get_dims(orctm_nh_t* h, int32_t* nx, int32_t* ny, int32_t* nz);

/* Bytes of device memory the handle holds (its own state plus the operator it uses). */
int orctm_nh_device_bytes(orctm_nh_t* h, int64_t* bytes);

/* Bytes of device memory this handle owns alone (its state, not the operator it may share). */
int orctm_nh_state_bytes(orctm_nh_t* h, int64_t* bytes);

/* A new handle on the same grid that shares src's assembled operator (read only, freed with the last
 * handle using it) and owns its own right-hand side (a copy of src's), solution (zero) and work
 * vectors: about 40 B/cell (MIXED) instead of ~140. For many states on one grid: ensemble members,
 * serving requests. No counterpart in the reference (one KSP per run, mo_petsc.f90:24-26). */
int orctm_nh_clone(orctm_nh_t* src, orctm_nh_t** out);

/* Synthetic time-step forcing on the resident right-hand side: b <- b * (1 + amplitude * xi),
 * xi ~ N(0,1) per cell from a counter-based stream keyed by (seed, cell); land stays 0. Used by the
 * serving workload (each request's RHS perturbed by 1 % per step). No counterpart in the reference. */
int orctm_nh_perturb_rhs(orctm_nh_t* h, double amplitude, uint64_t seed);

/* Device time of the two inner-iteration kernels, accumulated over the launches that did work
 * (HIP events recorded on the solver's own stream around each kernel; with timing on the iterations
 * are launched directly, not replayed from the captured graphs). Needs options.profile = 1 (or
 * orctm_nh_set_profile). No counterpart in the reference (PETSc's -log_view is never enabled). */
typedef struct orctm_nh_profile {
  double matvec_ms;         /* fused search-direction update + 7-point matvec + p.q */
  double update_ms;         /* x/r update + vertical-line solve + r.z, r.r */
  int64_t matvec_launches;
  int64_t update_launches;
  int64_t cells;            /* interior cells covered by one launch (all systems of the context) */
  int64_t reserved0;
  double reserved[2];
} orctm_nh_profile_t;
int orctm_nh_profile_read(orctm_nh_t* h, orctm_nh_profile_t* prof, int32_t reset);
/* Turns kernel timing on or off between solves (rebuilds the handle's solver context; a batch
 * created afterwards profiles if any member does). With timing on, the inner kernels are launched
 * directly with HIP events around them instead of from the captured graph. */
int orctm_nh_set_profile(orctm_nh_t* h, int32_t on);

/* Streaming bandwidth of the device: a plain 16-B-per-lane copy of `bytes` repeated `reps` times,
 * timed with HIP events; *gbps = bytes read + written per second / 1e9. The copy is also the
 * known-size reference for converting HBM byte counters. Allocates two buffers of `bytes`.
 * No counterpart in the reference. */
int orctm_nh_copy_probe(int32_t device, int64_t bytes, int32_t reps, double* gbps);

/* ------------------------------------------------------------------------------------------
 * The steps of OCVTOT_NH on either side of the solve (ORCTM/src_orctm/ocvtot_nh.f90), on the
 * device, for one process holding the whole domain. All arrays are the caller's Fortran arrays with
 * halos (mo_commo1.f90:105-146); sizes, I_start/J_start and the cyclic/open flags are those of the
 * grid the handle was created with. Expressions are evaluated left to right as the Fortran source
 * writes them (FP64). */
typedef struct orctm_nh_vgrid {
  const double* dlxp;    /* DLXP(IE,JE) */
  const double* dlyp;    /* DLYP(IE,JE) */
  const double* ddpo;    /* DDPO(IE,JE,KE) */
  const double* weto;    /* WETO(IE,JE,KE) */
  const double* amsuo;   /* AMSUO(I_start:IE,JE,KE) */
  const double* dduo;    /* DDUO(I_start:IE,JE,KE) */
  const double* dtdxuo;  /* DTDXUO(I_start:IE,JE) = DT/DLXU (beleg.f90:66) */
  const double* dlyu;    /* DLYU(I_start:IE,JE) */
  const double* amsue;   /* AMSUE(IE,J_start:JE,KE) */
  const double* ddue;    /* DDUE(IE,J_start:JE,KE) */
  const double* dpye;    /* DPYE(IE,J_start:JE) = DT/DLYV (beleg.f90:74) */
  const double* dlxv;    /* DLXV(IE,J_start:JE) */
  const double* dtdxpo;  /* DTDXPO(IE,JE) = DT/DLXP (beleg.f90:48) */
  const double* dtdyo;   /* DTDYO(IE,JE) = DT/DLYP (beleg.f90:49) */
  double dt;             /* DT */
  double dti;            /* DTI = 1/DT */
} orctm_nh_vgrid_t;
/* Uploads the velocity-side metrics once (they are constant over a run, like compute_matrix's). */
int orctm_nh_attach_vgrid(orctm_nh_t* h, const orctm_nh_vgrid_t* vg);
/* ocvtot_nh.f90:11-62 + the right-hand side of petsc_solve_3D: GU = (UOO-UOO_1)/DT, GV, GW likewise,
 * DIVG = DIVGU + DIVGV + DIVGW on the interior cells, then b = DIVG*WETO resident on the device
 * (ComputeRHS, mo_petsc.f90:376), as orctm_nh_set_rhs(DIVG) would. UOO, UOO_1 (I_start:IE,JE,KE);
 * VOE, VOE_1 (IE,J_start:JE,KE); WNO, WNO_1 (IE,JE,KEP); ZO (IE,JE). If divg is not NULL its
 * interior receives DIVG (the halo, which only the exchange of :43-57 would fill, is left as is). */
int orctm_nh_divergence(orctm_nh_t* h, const double* uoo, const double* uoo_1, const double* voe,
                        const double* voe_1, const double* wno, const double* wno_1, const double* zo,
                        double* divg);
/* ocvtot_nh.f90:151-230 after a solve: PNH gets the resident solution in its interior, then
 * bounds_exch(PNH) (one process: cyclic copies only, mo_parallel.f90:551-668) and the open-boundary
 * copies (:154-186) fill its halo (other halo cells keep the caller's values); UOO and VOE receive the
 * pressure correction (:189-201) and bounds_exch; WNO is recomputed from continuity (:207-228) and
 * exchanged. pnh, uoo, voe are read and written, wno (IE,JE,KEP) is written. The copies UOO_1 = UOO,
 * VOE_1 = VOE, WNO_1 = WNO, WO = WNO (:232-249) stay with the caller. */
int orctm_nh_project(orctm_nh_t* h, double* pnh, double* uoo, double* voe, double* wno);
/* With options.bottom_solve: the state of the last 2-D bottom solve (petsc_solve_2D, mo_petsc.f90:147-166,
 * then bounds_exch(PNH_bot), ocvtot_nh.f90:146). pnh_bot (IE,JE), halo included (zero on a closed side,
 * as the source leaves it); divg_bot_surf (IE,JE), interior written, halo untouched (the source's halo
 * values are never read); iterations and rel_res (||b - A2 p||_2 / ||b||_2) of that solve, and its device
 * time from the bottom values to the exchange of PNH_bot (HIP events). Any pointer may be NULL. The 2-D
 * solve's status is also what orctm_nh_divergence returns when it is not OK. */
int orctm_nh_get_bottom(orctm_nh_t* h, double* pnh_bot, double* divg_bot_surf, int32_t* iterations,
                        double* rel_res, double* device_ms);
/* With options.bottom_solve: ||K - K^T||_F / ||K||_F of the 2-D operator the bottom solve's CG runs on,
 * K = -S A2 over the inner rows of the box (S = DLXP*DLYP; A2 from compute_matrix_2d, mo_petsc.f90:401-538),
 * measured at create. It is 0 when DLXP is constant along j and DLYP along i; on other grids (spherical,
 * curvilinear) K is not symmetric: CG with the true-residual check every 8 iterations still converges on
 * the grids tested (defects up to 5e-3), and the solve reports ORCTM_NH_NOT_CONVERGED if it does not.
 * No counterpart in the reference (its cg2d_ KSP is PETSc's default GMRES). */
int orctm_nh_bottom_symmetry(orctm_nh_t* h, double* defect);

/* ------------------------------------------------------------------------------------------
 * The barotropic free-surface solve of the hydrostatic core: ITPREP (ORCTM/src_orctm/itprep.f90:15-38,
 * called once from orctm.f90:872) and TRONEU (troneu.f90:48-131, called every step from orctm.f90:1414),
 * for one process holding the whole domain. Arrays are the caller's Fortran arrays with halos
 * (mo_commo1.f90:105-146): p-points (IE,JE), u-points (I_start:IE,JE), v-points (IE,J_start:JE). */
typedef struct orctm_nh_sor_grid {
  int32_t ie, je, i_start, j_start;
  int32_t cyclic_x, cyclic_y;  /* ICYCLI_X, ICYCLI_Y */
  int32_t iterations;          /* red-black sweeps per solve (250 in troneu.f90:103) */
  int32_t device;              /* HIP device ordinal */
  int32_t half_sweeps;         /* half-sweeps per blocked device launch: 0 = the fastest of 4, 6, 7, 8, 12, timed at
                                  creation; or one of them. Results are bit-identical for every choice */
  int32_t reserved;            /* 0 */
  const double* dlxp;          /* DLXP(IE,JE) */
  const double* dlyp;          /* DLYP(IE,JE) */
  const double* dtdxpo;        /* DTDXPO(IE,JE) = DT/DLXP */
  const double* dtdyo;         /* DTDYO(IE,JE) = DT/DLYP */
  const double* weto;          /* WETO(IE,JE,KE): the first layer is used */
  const double* dlxu;          /* DLXU(I_start:IE,JE) */
  const double* dlyu;          /* DLYU(I_start:IE,JE) */
  const double* deuto;         /* DEUTO(I_start:IE,JE), u-point depth (boden.f90:116-122) */
  const double* dlxv;          /* DLXV(IE,J_start:JE) */
  const double* dlyv;          /* DLYV(IE,J_start:JE) */
  const double* deute;         /* DEUTE(IE,J_start:JE), v-point depth (boden.f90:124-130) */
  double g, conn, cono, stabn; /* G, CONN, CONO = 1-CONN, STABN (orctm.f90:657-661) */
  double dt;                   /* DT */
  double sorpar;               /* SORPAR, the over-relaxation factor (trotest.f90), in (0, 2) */
} orctm_nh_sor_grid_t;
typedef struct orctm_nh_sor orctm_nh_sor_t;
/* Uploads the metrics and forms XX, UF, VF, FF on the device (itprep.f90:15-38). */
int orctm_nh_sor_create(const orctm_nh_sor_grid_t* grid, orctm_nh_sor_t** sor);
/* troneu.f90:48-131: B1O from U1O, UZO (I_start:IE,JE) and V1E, VZE (IE,J_start:JE) on J = J_start+1..JE,
 * I = I_start+1..IE (b1o is read, for the entries the source does not recompute, and written); Z1O
 * starts at 0 with its boundary rows and columns at B1O/XX, then `iterations` red-black SOR sweeps,
 * each half-sweep followed by the one-process bounds_exch (cyclic copies). *device_ms (may be NULL)
 * receives the device time from B1O to the last sweep. */
int orctm_nh_troneu(orctm_nh_sor_t* sor, const double* u1o, const double* uzo, const double* v1e,
                    const double* vze, double* b1o, double* z1o, double* device_ms);
/* The half-sweeps per blocked launch in use and the kernel launches of one call's sweeps. */
int orctm_nh_sor_info(orctm_nh_sor_t* sor, int32_t* half_sweeps, int32_t* launches);
int orctm_nh_sor_destroy(orctm_nh_sor_t* sor);

/* ------------------------------------------------------------------------------------------
 * Building blocks of the inner solve, exposed for verification. No counterpart in the reference
 * (its Krylov internals live inside PETSc). "inner" is the scaled operator C = S (-A),
 * S = DLXP*DLYP/ZP per cell, held in FP32 (MIXED) or FP64 (FP64); x, y, r, z are interior vectors
 * in natural order of that precision (float or double). */
int orctm_nh_inner_apply(orctm_nh_t* h, const void* x, void* y);
/* z = M^{-1} r with M the vertical-line (per-column tridiagonal) preconditioner of C;
 * *rho receives r.z as accumulated by the device kernel (may be NULL). */
int orctm_nh_inner_precond(orctm_nh_t* h, const void* r, void* z, double* rho);

/* ------------------------------------------------------------------------------------------
 * Batched solves: many independent systems (ensemble members, nests) of uneven sizes advanced
 * together, one launch per solver phase for the whole batch, with per-system convergence masks.
 * The handles keep ownership of their device data; a batch only references them. */
typedef struct orctm_nh_batch orctm_nh_batch_t;
int orctm_nh_batch_create(int32_t n, orctm_nh_t* const* systems, orctm_nh_batch_t** batch);
/* Solves every member on its resident b (orctm_nh_set_rhs) and guess; stats has n entries (may be NULL).
 * Returns ORCTM_NH_OK when every member converged, else the first non-OK member status. */
int orctm_nh_batch_solve(orctm_nh_batch_t* batch, orctm_nh_stats_t* stats);
int orctm_nh_batch_destroy(orctm_nh_batch_t* batch);
int orctm_nh_batch_profile_read(orctm_nh_batch_t* batch, orctm_nh_profile_t* prof, int32_t reset);

/* ------------------------------------------------------------------------------------------
 * Serving: a request loop with continuous batching. No counterpart in the reference (ORCTM solves
 * once per model time step on its own ranks). A request is a grid -- given as a "model" handle whose
 * operator and resident right-hand side (orctm_nh_set_rhs) define it -- and nsteps consecutive time
 * steps: between steps the right-hand side is perturbed on the device (the stream of
 * orctm_nh_perturb_rhs, amplitude `perturb`, seed + step) and the solve warm-starts from the pressure
 * kept on the device.
 * The server runs `lanes` worker threads, each with its own stream, so that the work of several
 * admitted requests overlaps on the device. A free lane admits the oldest queued request that fits
 * the device-memory budget and the slot count (lane 0 takes only requests below 2^24 cells when there
 * are two lanes or more, so that small requests are never all queued behind large ones), together
 * with the queued requests of the same size class (floor(log2(cells))) and precision that also fit,
 * up to max_cohort requests or 2^25 cells: a cohort. Each admitted request holds a clone of its model
 * (the operator is shared). A cohort's requests advance together, one batched solve per time step
 * (one launch per solver phase, per-request convergence masks); a request retires at the end of its
 * last step, and the lane admits new requests as soon as its cohort is done, while other lanes keep
 * running. A request larger than the whole budget is refused with ORCTM_NH_ERR_NOMEM. The model handle
 * must stay alive and unchanged until its requests have been admitted. Each request's result is
 * bit-identical alone and in any mix. */
typedef struct orctm_nh_server orctm_nh_server_t;

typedef struct orctm_nh_server_options {
  int64_t hbm_budget;   /* bytes of device memory for admitted requests' state (<= 0: 90 % of what is free,
                           counting the library's idle pooled memory as free) */
  int32_t max_active;   /* admitted requests at most (default 64) */
  int32_t device;       /* HIP device ordinal (default 0) */
  int32_t lanes;        /* worker lanes (streams) running cohorts concurrently (default 4, at most 16) */
  int32_t max_cohort;   /* requests per cohort at most (default 16) */
  int32_t cohort_graphs; /* 1: a cohort replays its inner iterations from captured hipGraphs; 0 (default):
                            it launches them directly. A burst of 48 mixed requests on 4 lanes served
                            49.1 solves/s direct and 49.5 from graphs on MI355X, inside run-to-run noise;
                            direct launches skip the capture, which holds the lock other lanes' admissions
                            take */
  int32_t reserved[5];
} orctm_nh_server_options_t;

typedef struct orctm_nh_request_stats {
  int32_t status;       /* first non-OK step status, else ORCTM_NH_OK */
  int32_t steps_done;
  int32_t inner_iters;  /* summed over the steps */
  int32_t outer_iters;
  double rel_res;       /* of the last step */
  double t_submit_ms;   /* server clock: milliseconds since orctm_nh_server_create */
  double t_admit_ms;
  double t_done_ms;
  double reserved[4];
} orctm_nh_request_stats_t;

void orctm_nh_server_default_options(orctm_nh_server_options_t* opts);
int orctm_nh_server_create(const orctm_nh_server_options_t* opts, orctm_nh_server_t** server);
/* Queues a request; pnh_out (may be NULL): Fortran (IE,JE,KE) array that receives the interior of
 * the last step's solution before the request completes. */
int orctm_nh_server_submit(orctm_nh_server_t* server, orctm_nh_t* model, int32_t nsteps, double perturb,
                           uint64_t seed, double* pnh_out, int64_t* ticket);
/* ORCTM_NH_OK when finished (stats filled), ORCTM_NH_PENDING otherwise. */
int orctm_nh_server_poll(orctm_nh_server_t* server, int64_t ticket, orctm_nh_request_stats_t* stats);
int orctm_nh_server_wait(orctm_nh_server_t* server, int64_t ticket, orctm_nh_request_stats_t* stats);
int orctm_nh_server_drain(orctm_nh_server_t* server);
/* rounds: batched solves run so far (one per cohort and time step); bytes_in_use: state of the admitted
 * requests; active: admitted requests not yet finished; queued: requests waiting for admission. */
int orctm_nh_server_info(orctm_nh_server_t* server, int64_t* rounds, int64_t* bytes_in_use, int32_t* active,
                         int32_t* queued);
/* Finishes every queued and admitted request, then stops the server thread. */
int orctm_nh_server_destroy(orctm_nh_server_t* server);

#ifdef __cplusplus
}
#endif

#endif /* ORCTM_NH_H */
