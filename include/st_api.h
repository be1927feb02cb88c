/*
 * st_api.h -- C-ABI of the MI355X per-RK-stage pipeline for the periodic
 * single-fluid shock-turbulence solver.
 *
 * Each entry point replaces one subroutine of the reference Fortran path
 * (paths relative to the reference repository); the Fortran module
 * st_mod.f90 binds these with iso_c_binding.
 *
 * Arrays are Fortran-ordered F(My,Mz,Mx) (y unit stride), several fields
 * stacked last: Q(My,Mz,Mx,5) = [rho*u, rho*v, rho*w, rho, rho*E]
 * (src/main.f:28, src/header:7-10).  Host pointers unless a name ends in
 * _dev.  All calls return 0 on success, a negative code on failure, and
 * never abort; st_last_error() gives the message.  Calls on one context
 * run in order on that context's stream; host-pointer calls return after
 * their results are on the host.
 */
#ifndef ST_API_H
#define ST_API_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct st_ctx st_ctx;

/* Grid and run parameters: the runtime form of the compile-time header
 * (src/header:5-10) and the input deck (src/define_periodic_sf.f:20-80). */
typedef struct {
    int nx, ny, nz;
    double period_x, period_y, period_z;
    double x_start, y_start, z_start;
    /* 0: periodic in x (src/makefile_periodic:11-22).  1: the non-periodic-x
     * link set (src/makefile_nonperiodic1d:10-17): x spans period_x with both
     * end points (src/define_nonperiodic1d_sf.f:42), x derivatives and
     * filters close with boundary rows (src/derivCDopt6_nonperiodic1d.f,
     * src/filters_nonperiodic1d.f), the Gaussian mirrors at the ends, the
     * artificial transport uses constant coefficients
     * (src/hyper_routines.f:66-214) and the hyper dt limit applies from the
     * first step (src/lddrk4_nonperiodic_sf.f:105). */
    int nonperiodic_x;
} st_grid;

typedef struct {
    double Re, Pr, gamma, vis_exp, bulk_ratio;
    int hyper_param;       /* 1: artificial transport on (src/hyper_rhs.f:50-52) */
    double cfl_cal;
    int d2_mode;           /* 0: consistent FDL3DI D2; 1: coefficients as shipped */
    int rk_scheme;         /* 0: CC5 then LDDRK6; 1: LDDRK5 then LDDRK6; 2: CC5 only */
} st_params;

/* Setup_Grid + Setup_Derivatives + Init_Filters + Initialize_MPI at one rank
 * (src/grid.f:15, src/derivCDopt6_periodic.f:9, src/filters_periodic.f:9,
 * src/initialize_mpi_periodic.f90:1). Allocates all device buffers. */
int st_ctx_create(const st_grid *grid, const st_params *params, st_ctx **out);
void st_ctx_destroy(st_ctx *ctx);
const char *st_last_error(const st_ctx *ctx);
/* Build provenance (no counterpart in the reference): "src-sha256:" and the
 * first 16 hex digits of the SHA-256 of the library's sources (each file's
 * path relative to the package, a newline, then its contents, in sorted
 * path order), fixed at build time. */
const char *st_build_id(void);
int st_sync(st_ctx *ctx);

/* State in / out (the /RKWORK/ Q array, src/main.f:31). */
int st_state_upload(st_ctx *ctx, const double *Q);
int st_state_download(st_ctx *ctx, double *Q);
/* Device address of the resident state Q (5 fields); for callers that keep
 * data on the GPU. */
int st_state_dev(st_ctx *ctx, double **Q_dev);

/* DFDX/DFDY/DFDZ(F, dF, iadd, cadd)  src/derivCDopt6_periodic.f:164,400,631.
 * dir 0,1,2 = x,y,z; iadd 0: dF = cadd*dF/ds, 1: dF += cadd*dF/ds. */
int st_d1(st_ctx *ctx, int dir, const double *F, double *dF, int iadd, double cadd);
/* D2FDX2/D2FDY2/D2FDZ2(F, dF, d2F, iadd)  :281,515,747; iadd=1 also sets dF. */
int st_d2(st_ctx *ctx, int dir, const double *F, double *dF, double *d2F, int iadd);
/* Compact_Filter(F, F0, N)  src/filters_periodic.f:90 (X, then Y, then Z). */
int st_filter(st_ctx *ctx, const double *F, double *F0, int nf);
/* Compact_Filter_X/Y/Z  src/filters_periodic.f:115,254,393. */
int st_filter_dir(st_ctx *ctx, int dir, const double *F, double *F0);
/* Gaussian_Filter(F, Fg, N)  src/filters_periodic.f:540. */
int st_gauss(st_ctx *ctx, const double *F, double *Fg, int nf);
/* Gaussian_Filter_X/Y/Z  src/filters_periodic.f:565,680,793. */
int st_gauss_dir(st_ctx *ctx, int dir, const double *F, double *Fg);

/* RHS(Q, Qrhs)  src/rhsv_periodic_sf.f:9.  Q == NULL uses the resident
 * state; Q is not modified.  Leaves the 12 first derivatives (/GRAD/) on
 * the device for st_hyper_rhs. Qrhs == NULL keeps the result on the device. */
int st_rhs(st_ctx *ctx, const double *Q, double *Qrhs);
/* Hyper_RHS(Q, Qrhs)  src/hyper_rhs.f:12: adds the artificial-transport
 * terms to the right-hand side of the preceding st_rhs (same Q). */
int st_hyper_rhs(st_ctx *ctx, double *Qrhs);
/* Compute_hyper(Q)  src/hyper_dynamic.f:173-318 (non-periodic x: the static
 * coefficients of src/hyper_routines.f:66-214), as Program Main calls it
 * before the first time step (src/main.f:100): beta*, mu*, kappa* (/HYPER/)
 * from the state and the velocity gradients /GRAD/ as the last st_rhs or
 * st_compute_grad left them (zero on a new context, as the reference's
 * /GRAD/ is at start-up).  Q == NULL uses the resident state; otherwise Q
 * is uploaded as the resident state first (as st_rhs does).  The next
 * st_tstep's dt limits read these coefficients.  st_slot_load does the
 * same for its slot with zero gradients. */
int st_compute_hyper(st_ctx *ctx, const double *Q);
/* Compute_Grad  src/grad.f:13-115 (called by Output, src/out.f:718), on the
 * resident state: overwrites the nine velocity gradients of /GRAD/ and
 * returns MaxDil = max |div u| and Enstrophy = (1/2) <|curl u|^2>, the log
 * columns 6 and 7 of src/main.f:110-116.  Either pointer may be NULL. */
int st_compute_grad(st_ctx *ctx, double *maxdil, double *enstrophy);
/* Compute_hyper(Q) coefficients after st_hyper_rhs or st_compute_hyper:
 * beta*, mu*, kappa* (src/hyper_dynamic.f:173, /HYPER/). */
int st_get_hyper(st_ctx *ctx, double *H3);
int st_set_hyper(st_ctx *ctx, const double *H3);
/* /GRAD/ after st_rhs: dux_dx duy_dx duz_dx dux_dy duy_dy duz_dy dux_dz
 * duy_dz duz_dz dTdx dTdy dTdz (src/rhsv_periodic_sf.f:55-58). */
int st_get_grad(st_ctx *ctx, double *G12);

/* Fix_Pressure(Q)  src/lddrk4_periodic_sf.f:436 (host array in place). */
int st_fix_pressure(st_ctx *ctx, double *Q);
/* Compute_CFL  src/lddrk4_periodic_sf.f:7, on the resident state. */
int st_cfl(st_ctx *ctx, double *cfl_stab, double *cfl_resv);
/* Compute_CFL that allocates nothing.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Q_dev is
 * a device array Q(My,Mz,Mx,5) of the given grid; out_dev (device, 2
 * doubles) receives cfl_stab and cfl_resv, the same bits as st_cfl on that
 * state.  The work is enqueued on `stream` (a hipStream_t; NULL: the null
 * stream) and not waited for.  No context: the per-block partials live in
 * a fixed 256 KiB device array of the library, one per GPU, so grids of at
 * most 33,554,432 points are accepted (larger ones return -2 before any
 * launch); calls that overlap on one GPU race on that array.  Errors:
 * st_last_error(NULL). */
int st_cfl_noalloc(const st_grid *grid, double gamma, const double *Q_dev, double *out_dev, void *stream);
/* Compute_Tstep  src/lddrk4_periodic_sf.f:74 (hyper limits when step > 1). */
int st_tstep(st_ctx *ctx, double cfl_resv, int step, double *dt);
/* One stage of LDDRK4 (src/lddrk4_periodic_sf.f:344-356) on the resident
 * state: RHS, Hyper_RHS, QW = a QW + dt dQ, Q += b QW, filter, Fix_Pressure.
 * reset_qw = 1 zeroes QW first (start of each stage loop, :339, :376). */
int st_rk_stage(st_ctx *ctx, double alpha, double beta, double dt, int reset_qw);
/* LDDRK4(time, step)  src/lddrk4_periodic_sf.f:311 with a given dt. */
int st_step(st_ctx *ctx, double dt, double *time, int *step);
/* nsteps of march mode 1 (src/main.f:129-136): Compute_CFL, Compute_Tstep,
 * LDDRK4 per step.  dt_out (may be NULL) receives each step's dt. */
int st_march(st_ctx *ctx, int nsteps, double *time, int *step, double *dt_out);
/* Compute_Mom, Compute_KE, Compute_Mass, Compute_Energy, Compute_Entropy
 * (src/lddrk4_periodic_sf.f:153-302): MomX MomY MomZ KE Mass Energy Entropy. */
int st_diagnostics(st_ctx *ctx, double *out7);

/* Per-pass GPU timing with HIP events (0 off, 1 on).  st_profile_read
 * returns, for up to nmax entries (pass kind x direction), the launch count,
 * total milliseconds and the fields read+written per launch; names
 * ("kind.direction") via st_pass_name. */
int st_profile(st_ctx *ctx, int on);
int st_profile_read(st_ctx *ctx, int nmax, int *count, double *total_ms, double *fields_rw);
const char *st_pass_name(int kind);

/* ---- batches of same-size instances (no counterpart in the reference,
 * which runs one problem per job).  For ensemble and serving callers: a
 * batch holds nslots instances of one grid and parameter set, stacked field
 * by field on the device, and advances every loaded slot with one launch
 * per pass.  Each slot keeps its own state, Re, dt, time and step counter;
 * dt is computed on the device (Compute_CFL + Compute_Tstep per slot), so a
 * step needs no host round trip and is replayed from a captured HIP graph.
 * A slot gives bit-identical results alone and among others.  The
 * single-instance entry points above act on slot 0. */
int st_batch_create(const st_grid *grid, const st_params *params, int nslots, st_ctx **out);
int st_batch_slots(const st_ctx *ctx);
/* Load a slot's state and run settings and mark it active: it advances on
 * each st_batch_march step while step < step0 + nsteps (the driver's step
 * counter starts at step0, src/main.f:129-136); Re <= 0 means inviscid. */
int st_slot_load(st_ctx *ctx, int slot, const double *Q, double Re, double time0, int step0, int nsteps);
int st_slot_store(st_ctx *ctx, int slot, double *Q);
int st_slot_release(st_ctx *ctx, int slot);
/* Enqueue nsteps batch steps on the context's stream and return (st_sync or
 * st_batch_status waits).  Finished and released slots do no work. */
int st_batch_march(st_ctx *ctx, int nsteps);
/* Per-slot step counter, time, last dt and active flag (arrays of nslots;
 * any may be NULL); waits for the stream. */
int st_batch_status(st_ctx *ctx, int *step, double *time, double *dt_last, int *active);
/* 1 when all work enqueued on the context's stream has finished, 0 while
 * it runs (does not wait). */
int st_query(st_ctx *ctx);

/* ---- groups of batches of different sizes (no counterpart in the
 * reference).  For ensemble and serving callers with mixed grid sizes: a
 * group steps up to four batches, one per size of the family 48^3, 64^3,
 * 96^3, 128^3 (periodic, no sponge, one hyper_param and rk_scheme), with
 * ONE launch per pass covering every member's running slots, instead of
 * one launch per pass per batch.  Every pass of a member then takes that
 * size's M/2-point layout (st_pass_config_set narrow = 1, nt = 256 gives a
 * single batch the same layouts and so bit-identical results).  The
 * members stay usable as batches (slot load, store, release, status act
 * on them); st_group_march orders itself after the members' pending work
 * and their later work after its steps.  A step is captured in a HIP graph
 * once per tuple of active-slot counts.  Destroy the group before its
 * members.  Errors: st_last_error(NULL). */
typedef struct st_group st_group;
int st_group_create(st_ctx *const *members, int nmembers, st_group **out);
int st_group_march(st_group *group, int nsteps);
int st_group_sync(st_group *group);
void st_group_destroy(st_group *group);

/* Autotuner (no counterpart in the reference): time every pass of an RK
 * stage over all slots under each lane layout (M or M/2 points per lane) and
 * tile width (256, 512, 1024 threads) and keep the fastest per pass and
 * direction; reps stages per candidate.  The state is restored.
 * st_pass_config reads the choice for kind_dir = kind*3 + direction (the
 * index of st_pass_name); st_pass_config_set sets it (narrow 0 or 1, nt
 * 256, 512 or 1024), e.g. to give another context of the same grid the
 * same lane layouts, and so the same rounding.  A tile width that a pass
 * cannot use falls back to the next smaller one at launch. */
int st_autotune(st_ctx *ctx, int reps);
int st_pass_config(st_ctx *ctx, int kind_dir, int *narrow, int *nt);
int st_pass_config_set(st_ctx *ctx, int kind_dir, int narrow, int nt);

/* ---- on-disk formats (host only, no context) ----
 * Write_Restart / Read_Restart  src/out.f:493-692: a 28-byte header (the
 * extents Ny, Nz, Nx, 5 as int32, the time as float64, the step as int32)
 * then Q(Ny,Nz,Nx,5) as float64 in Fortran order, native byte order (one
 * rank: the whole grid).  st_restart_read fails (-4) on a file of other
 * extents.  The reference names the files <restart_path>/<prefix>.restart.<n>
 * (:528-532), reading n = 0 at start-up (src/main.f:69-71) and writing
 * n = 1, 2, ... (src/out.f:722-725). */
int st_restart_write(const char *path, int nx, int ny, int nz, const double *Q, double time, int step);
int st_restart_read(const char *path, int nx, int ny, int nz, double *Q, double *time, int *step);

/* The input deck of Define_Parameters (src/define_periodic_sf.f:20-80), in
 * its record order: a comment line per section, list-directed values (the
 * first token of each line, d exponents allowed), paths as A20, and a
 * blank record between sections.  Grid extents are not in the deck (they
 * are compile-time in the reference, src/header:5-10). */
typedef struct {
    char vis_path[24], restart_path[24], stats_path[24], probe_path[24], source_path[24], prefix[24];
    int grid_x, grid_y;
    double x_start, y_start, z_start, period_x, period_y, period_z;
    double Re, Pr, Sc, gamma, vis_exp, bulk_ratio, T_ratio;
    int hyper_param, part_param;
    int march_mode;
    double Tstep, cfl_cal;
    int Nsteps;
    double run_time;
    int comp_stats;
    int out_grid, out_init, out_incr_x, out_incr_y, out_incr_z, out_restart, out_vis, out_probe, out_stats,
        out_source;
} st_deck;
int st_deck_read(const char *path, st_deck *deck);

/* Sponge_Forcing  src/sponge1d.f:7-95 in every stage after Hyper_RHS (the
 * hook order of src/lddrk4_nonperiodic_mf.f:462-466): dQ -= sigma(x)
 * (Q - Qref) with sigma = amp (d/len)^power over the first and last
 * ceil(frac nx/2) x points (the reference: amp 5, power 2, frac 0.4).  The
 * reference holds one hard-coded Qref (:51-57); here the left layer relaxes
 * to ref_left and the right to ref_right (conserved variables).  amp = 0
 * turns it off. */
int st_sponge(st_ctx *ctx, double amp, double power, double frac, const double ref_left[5],
              const double ref_right[5]);
/* Planar Mach-`mach` shock at x = frac*period_x (tanh profile of the given
 * width) on the non-periodic x grid, Rankine-Hugoniot states on either
 * side, upstream isotropic turbulence (k0, urms; urms = 0: none).  states
 * (may be NULL) receives the upstream and downstream conserved states (10). */
int st_ic_shock(int nx, int ny, int nz, const double period[3], double mach, double frac, double width,
                double k0, double urms, double gamma, unsigned long long seed, double *Q, double *states);

/* Seeded isotropic-turbulence initial condition (new; the reference restarts
 * from a file, src/main.f:69-71).  Q(My,Mz,Mx,5) on the host. */
int st_ic_isotropic(int nx, int ny, int nz, const double period[3], double k0, double urms,
                    double mt, double gamma, unsigned long long seed, double *Q);

#ifdef __cplusplus
}
#endif
#endif
