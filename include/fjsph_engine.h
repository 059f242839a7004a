/*
 * fjsph_engine.h -- C-ABI of the MI355X WCSPH step engine.
 *
 * The engine replaces the per-step particle operators of FJSPH
 * (src/Neighbours.h, src/Shifting.h, src/Geometry.h, src/Resid.h,
 * src/Newmark_Beta.h, src/Runge_Kutta.h, src/Integration.h) with hand-written
 * gfx950 kernels.  A host driver owns the settings and calls these entry points
 * in the order the reference's Integrator does; every call is synchronous on
 * the context's HIP stream unless stated otherwise.
 *
 * Conventions
 *  - Plain C types only; no exceptions cross this boundary.  Every entry point
 *    returns an fjsph_status; fjsph_last_error(ctx) holds the message of the
 *    last failure on that context.  The reference prints and calls exit(-1)
 *    on fatal input errors (src/IO.cpp:708-712); here the call fails instead.
 *  - Host arrays are in HOST ORDER: the order of the upload, boundary
 *    particles first ([0, n_bound)), then fluid ([n_bound, n_total)), as
 *    Init_Particles lays them out (src/Init.cpp:288-484).  The engine keeps
 *    particles sorted by cell on the device and maps back on download.
 *  - Vectors are interleaved per particle: x[i*dim + k].
 *  - One context is used by one host thread at a time.
 */
#ifndef FJSPH_ENGINE_H
#define FJSPH_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FJSPH_ABI_VERSION 7

typedef enum fjsph_status {
    FJSPH_OK = 0,
    FJSPH_ERR_ARG = 1,         /* bad argument (null pointer, size, enum) */
    FJSPH_ERR_HIP = 2,         /* a HIP runtime call failed */
    FJSPH_ERR_OOM = 3,         /* device allocation failed */
    FJSPH_ERR_STATE = 4,       /* call out of order (e.g. no state uploaded) */
    FJSPH_ERR_UNSUPPORTED = 5, /* option not built into this engine */
    FJSPH_ERR_NO_DEVICE = 6    /* no HIP device visible */
} fjsph_status;

/* Particle type flag b (src/VarDefs.h:92-102). */
enum {
    FJSPH_BOUND = 0,
    FJSPH_PISTON = 1,
    FJSPH_BUFFER = 2,
    FJSPH_BACK = 3,
    FJSPH_PIPE = 4,
    FJSPH_FREE = 5,
    FJSPH_OUTLET = 6,
    FJSPH_LOST = 7
};

/* Arithmetic of the pair passes. */
enum {
    FJSPH_PREC_FP64 = 0,  /* every pair term in FP64 (parity path)               */
    FJSPH_PREC_MIXED = 1  /* FP32 pair terms, FP64 accumulation, FP64 x/t/residual */
};

enum { FJSPH_NEWMARK_BETA = 0, FJSPH_RUNGE_KUTTA = 1 };     /* src/VarDefs.h:129-133 */
enum { FJSPH_DBC = 0, FJSPH_PRESSURE_G = 1, FJSPH_GHOST = 2 }; /* src/VarDefs.h:135-140 */
enum { FJSPH_COLE_EOS = 0, FJSPH_ISO_EOS = 1 };            /* src/VarDefs.h:142-146 */
enum { FJSPH_NO_AERO = 0, FJSPH_GISSLER = 1, FJSPH_INDUCED_PRESSURE = 2, FJSPH_SKIN_FRIC = 3 };
/* Aerodynamic velocity source (src/VarDefs.h:168-173: constVel, meshInfl,
 * VLMInfl).  FJSPH_SURROGATE has no counterpart in the reference: the mesh
 * field replaced by a small neural network (fjsph_set_surrogate). */
enum { FJSPH_CONST_VEL = 0, FJSPH_MESH_INFL = 1, FJSPH_VLM_INFL = 2, FJSPH_SURROGATE = 3 };

/*
 * Settings: the fields of FLUID / AERO / INTEG_SETT / SIM that the hot path
 * reads (src/Var.h:63-92, 160-237, 240-310, 337-394), under the para-file keys
 * of GetInput (src/IO.cpp:349-453).  fjsph_settings_default() fills the
 * reference's defaults; derived constants (Set_Values, src/IO.cpp:26-128) are
 * computed by the engine and can be read back with fjsph_derive().
 */
typedef struct fjsph_settings {
    double particle_step;       /* "SPH initial spacing"                       */
    double H_fac;               /* "SPH smoothing length factor"    (2.0)      */
    double rho_rest;            /* "Reference dispersed density"    (1000)     */
    double speed_sound;         /* "SPH speed of sound"             (300)      */
    double gam;                 /* Tait gamma                       (7)        */
    double mu;                  /* "Reference dispersed viscosity"  (8.94e-4)  */
    double sig;                 /* "Reference surface tension"      (0.0708)   */
    double visc_alpha;          /* "SPH artificial viscosity factor"(0.1)      */
    double dsph_delta;          /* "SPH delta coefficient"          (0.1)      */
    double press_back;          /* "SPH background pressure"        (0)        */
    double press_pipe;          /* "SPH starting pressure"          (0)        */
    double rho_var;             /* "SPH maximum absolute density variation (%)" (50) */
    double rho_max;             /* "SPH maximum density" (1500 = unset)        */
    double rho_min;             /* "SPH minimum density" (500 = unset)         */
    double rho_max_iter;        /* "SPH density variation to reduce timestep (%)" (1) */
    double grav[3];             /* "SPH gravity vector"                        */
    double cfl;                 /* "SPH CFL condition"              (1.0)      */
    double cfl_min;             /* "SPH minimum CFL"                (0.1)      */
    double cfl_max;             /* "SPH maximum CFL"                (2.0)      */
    double cfl_step;            /* "SPH unstable CFL step"          (0.05)     */
    double subits_factor;       /* "SPH stable CFL count iteration factor" (0.333) */
    double min_residual;        /* "SPH solver minimum residual"    (-7)       */
    double delta_t_max;         /* "SPH maximum timestep"           (1.0)      */
    double delta_t_min;         /* "SPH minimum timestep"           (0)        */
    double frame_time_interval; /* "SPH frame time interval"                   */
    double max_shift_vel;       /* "SPH maximum shifting velocity"  (9999999)  */
    double v_inf[3];            /* "SPH freestream velocity"                   */
    double p_ref;               /* "Reference pressure"             (101353)   */
    double rho_g;               /* "Reference density"              (1.29251)  */
    double mu_g;                /* "Sutherland reference viscosity" (1.716e-5) */
    double temp_g;              /* "Reference temperature"          (298)      */
    double R_g;                 /* specific gas constant            (287)      */
    double gamma_g;             /* "Gas constant gamma"             (1.403)    */
    double lam_cutoff;          /* "SPH aerodynamic cutoff value"   (0.75)     */
    double i_interp_fac;        /* "SPH aerodynamic interpolation factor" (0.5) */
    double Cf, Ck, Cd, Cb;      /* TAB constants (1/3, 8, 5, 0.5)              */
    int32_t dim;                /* SIMDIM: 2 or 3                              */
    int32_t precision;          /* FJSPH_PREC_*                                */
    int32_t integrator;         /* "SPH integration solver": FJSPH_NEWMARK_BETA / FJSPH_RUNGE_KUTTA */
    int32_t eos;                /* "SPH equation of state (0=Cole/1=Isothermal)" */
    int32_t aero_case;          /* "SPH aerodynamic case": FJSPH_NO_AERO ...   */
    int32_t aero_source;        /* FJSPH_CONST_VEL / FJSPH_MESH_INFL           */
    int32_t use_lam;            /* "SPH interpolation factor (0=ncount/1=lambda)" (1) */
    int32_t use_TAB_def;        /* "SPH use TAB deformation (0/1)"  (0)        */
    int32_t ale;                /* ALE particle shifting (the reference's -DALE build) */
    int32_t max_subits;         /* "SPH Newmark-Beta iteration limit" (20)     */
    int32_t n_stable_limit;     /* "SPH stable CFL count limit"     (10)       */
    int32_t n_unstable_limit;   /* "SPH unstable CFL count limit"   (3)        */
} fjsph_settings;

/* Constants derived by Set_Values (src/IO.cpp:26-128). */
typedef struct fjsph_derived {
    double H, H_sq, sr, W_correc, W_dx, B, dsph_cont, nu;
    double sim_mass, bnd_mass, rho_max, rho_min, dx;
    double n_full, i_n_full, interp_fac, A_plate, A_sphere, L;
    double sos, i_sos_sq, td, omega, tmax, Cdef, ycoef;
    double lam_st;  /* pairwise surface-tension normaliser (src/Resid.cpp:437-439) */
} fjsph_derived;

/* A boundary or fluid block: the bound_block fields the step reads
 * (src/Var.h:820-881). Blocks are given in host order and must tile
 * [0, n_total) with boundary blocks first. */
typedef struct fjsph_block {
    int64_t first;        /* index.first  */
    int64_t last;         /* index.second (exclusive) */
    int32_t is_fluid;     /* 0 = boundary block, 1 = fluid block */
    int32_t bound_solver; /* FJSPH_DBC / FJSPH_PRESSURE_G / FJSPH_GHOST */
    int32_t no_slip;      /* "Wall is no slip (0/1)" */
    int32_t n_times;      /* 0: constant velocity vel; else see fjsph_set_block_motion */
    double vel[3];        /* vels[0] */
} fjsph_block;

/* Planes of a fluid block (bound_block, src/Var.h:858-866): a particle
 * crosses plane (norm, c) when x . norm > c.  delconst = FJSPH_NO_PLANE
 * (default_val, src/VarDefs.h:195) switches deletion off; the aero plane
 * is applied as given (the reference's unset value is norm = (default_val,
 * default_val, default_val), c = default_val). */
#define FJSPH_NO_PLANE 9999999.0
typedef struct fjsph_block_planes {
    int32_t inlet;                /* block_type == inletZone */
    int32_t fixed_vel_or_dynamic; /* inlet buffer: 0 fixed velocity, 1 dynamic */
    double delete_norm[3], delconst;
    double insert_norm[3], insconst;
    double aero_norm[3], aeroconst;
} fjsph_block_planes;

/* Per-step report of Integrator::integrate (src/Integration.cpp:233-303). */
typedef struct fjsph_step_info {
    double rms_error;   /* log-residual of the last sub-iteration        */
    double dt;          /* timestep used                                  */
    double cfl;         /* CFL after the adaptation                       */
    double time;        /* simulation time after the step                 */
    double max_rho_pc;  /* maxRho_pc                                      */
    double maxf;        /* max |acc|                                      */
    double maxAf;       /* max |Af|                                       */
    double npd;         /* numerical particle density of the step         */
    int32_t iterations; /* sub-iteration counter `iteration`              */
    int32_t rollbacks;  /* rejected sub-iteration sequences (dt halvings) */
} fjsph_step_info;

/* Field identifiers for fjsph_download_field (names follow SPHPart,
 * src/Var.h:600-641).  Vector fields have dim components, L has dim*dim
 * (row-major), scalars one. */
enum {
    FJSPH_F_X = 0, FJSPH_F_V, FJSPH_F_ACC, FJSPH_F_AF, FJSPH_F_AVISC,
    FJSPH_F_RHO, FJSPH_F_P, FJSPH_F_RRHO, FJSPH_F_M, FJSPH_F_DELTAD,
    FJSPH_F_L, FJSPH_F_GRADRHO, FJSPH_F_NORM, FJSPH_F_LAM, FJSPH_F_LAM_NB,
    FJSPH_F_COLOURG, FJSPH_F_COLOUR, FJSPH_F_KERNSUM, FJSPH_F_SURF,
    FJSPH_F_CURVE, FJSPH_F_NORM_CURVE, FJSPH_F_WOCCL, FJSPH_F_PDIST,
    FJSPH_F_CELLV, FJSPH_F_CELLP, FJSPH_F_CELLRHO, FJSPH_F_CELLID,
    FJSPH_F_VPERT, FJSPH_F_SURFZONE, FJSPH_F_B,
    FJSPH_F_PID, /* SPHPart::part_id: host index at upload, then svar.part_id++ per inserted particle */
    FJSPH_F_NFAILED,  /* SPHPart::ipt_n_failed: consecutive mesh lookups that placed the particle in no cell */
    FJSPH_F_INTERNAL, /* SPHPart::internal: a lookup found it across a wall (marker -1) boundary face */
    FJSPH_F_COUNT
};
enum { FJSPH_STATE_N = 0, FJSPH_STATE_NP1 = 1 }; /* pn / pnp1 */

/* ---- settings ------------------------------------------------------------ */

/* Fill the reference defaults (src/Var.h:63-310) for dimension dim. */
void fjsph_settings_default(fjsph_settings* s, int32_t dim);

/* Set_Values (src/IO.cpp:26-128) plus get_n_full (src/Geometry.cpp:282-308). */
fjsph_status fjsph_derive(const fjsph_settings* s, fjsph_derived* out);

/* Parse a para file in the reference format ("key: value # comment",
 * GetInput src/IO.cpp:305-723, Get_Number/Get_Vector src/IOFunctions.h:32-230)
 * into s (which should hold defaults first). Unknown keys are ignored, as in
 * the reference. */
fjsph_status fjsph_settings_parse_file(fjsph_settings* s, const char* path);

/* ---- context ------------------------------------------------------------- */

typedef struct fjsph_ctx fjsph_ctx;

int32_t fjsph_abi_version(void);

/* Create an engine on HIP device `device`. Replaces the driver-side SIM/pn/
 * pnp1/outlist locals of main (src/FJSPH.cpp:49-58). */
fjsph_status fjsph_create(const fjsph_settings* s, int32_t device, fjsph_ctx** out);
void fjsph_destroy(fjsph_ctx* ctx);
const char* fjsph_last_error(const fjsph_ctx* ctx);

/* Upload the initial state (pn = pnp1 = state, src/Init.cpp:493) in host
 * order. x, v: n_total*dim; rho, p, m: n_total; b: particle types. For the
 * constant-velocity aero source, cellV/cellP/cellRho are initialised as in
 * src/FJSPH.cpp:115-126. Time is reset to 0. */
fjsph_status fjsph_upload_state(fjsph_ctx* ctx, int64_t n_total, int64_t n_bound,
                                const double* x, const double* v, const double* rho,
                                const double* p, const double* m, const int32_t* b);

/* Boundary/fluid blocks (LIMITS, src/Var.h:883). */
fjsph_status fjsph_set_blocks(fjsph_ctx* ctx, int32_t n_blocks, const fjsph_block* blocks);
/* Time-velocity table of a moving boundary block (bound_block::times/vels). */
fjsph_status fjsph_set_block_motion(fjsph_ctx* ctx, int32_t block, int32_t n_times,
                                    const double* times, const double* vels);

/* Planes of block `block` (bound_block::delete_norm/delconst,
 * insert_norm/insconst, aero_norm/aeroconst, block_type,
 * fixed_vel_or_dynamic).  Check_Pipe_Outlet (src/Containment.cpp:822-890)
 * turns PIPE particles past the aero plane FREE after every pre-pass;
 * update_data (src/Integration.cpp:109-226) deletes fluid particles past
 * an active deletion plane. */
fjsph_status fjsph_set_block_planes(fjsph_ctx* ctx, int32_t block, const fjsph_block_planes* planes);
/* Inlet columns of an inlet block (bound_block::back and ::buffer, built
 * by the inlet shape, src/shapes/inlet.cpp): back[c] is the host index of
 * column c's BACK particle, buffer[c*n_buf + j] its j-th BUFFER particle
 * behind it.  update_buffer_region (src/shapes/inlet.cpp:578-640) then
 * runs in every step and the buffer follows the inlet mode in Do_NB_Iter
 * (src/Newmark_Beta.cpp:243-292).  Needs the Newmark-Beta integrator. */
fjsph_status fjsph_set_inlet(fjsph_ctx* ctx, int32_t block, int32_t n_cols, int32_t n_buf, const int64_t* back,
                             const int64_t* buffer);
/* SPH maximum particle count (svar.max_points, src/Var.h:369): no new inlet
 * particles past it. */
fjsph_status fjsph_set_max_points(fjsph_ctx* ctx, int64_t max_points);
fjsph_status fjsph_get_max_points(fjsph_ctx* ctx, int64_t* max_points);
/* Current particle counts (inlets add particles, deletion removes them). */
fjsph_status fjsph_get_counts(fjsph_ctx* ctx, int64_t* n_total, int64_t* n_bound);
/* Current host range of a block (index.first, index.second). */
fjsph_status fjsph_get_block_range(fjsph_ctx* ctx, int32_t block, int64_t* first, int64_t* last);

/* Download one field of pn or pnp1 in host order (see FJSPH_F_*). */
fjsph_status fjsph_download_field(fjsph_ctx* ctx, int32_t which, int32_t field, double* out);
/* Overwrite one field of pn or pnp1 from host order (test and restart use). */
fjsph_status fjsph_upload_field(fjsph_ctx* ctx, int32_t which, int32_t field, const double* in);

/* ---- per-pass entry points (same units as the reference's functions) ------ */

/* update_neighbours (src/Neighbours.cpp:7-31) on pnp1: the set of j with
 * |x_i - x_j|^2 < 4H^2 in FP64, self included. */
fjsph_status fjsph_build_neighbours(fjsph_ctx* ctx);
/* Neighbour counts (self included) in host order. */
fjsph_status fjsph_get_neighbour_counts(fjsph_ctx* ctx, int32_t* counts);
/* CSR dump of the lists in host order: offsets[n_total+1], idx[offsets[n]];
 * each list sorted ascending, self included. */
fjsph_status fjsph_get_neighbours(fjsph_ctx* ctx, int64_t* offsets, int64_t* idx, int64_t idx_capacity);

/* dSPH_PreStep (src/Shifting.cpp:12-123) over all particles; returns npd. */
fjsph_status fjsph_dsph_prestep(fjsph_ctx* ctx, double* npd);
/* get_aero_velocity, constant-velocity source (src/Resid.cpp:569-611). */
fjsph_status fjsph_aero_velocity(fjsph_ctx* ctx);
/* Detect_Surface (src/Geometry.cpp:14-280). */
fjsph_status fjsph_detect_surface(fjsph_ctx* ctx);
/* dissipation_terms (src/Shifting.cpp:126-186). */
fjsph_status fjsph_dissipation_terms(fjsph_ctx* ctx);
/* particle_shift (src/Shifting.cpp:189-290); requires settings.ale. */
fjsph_status fjsph_particle_shift(fjsph_ctx* ctx);
/* Boundary velocity, no-slip and boundary solver of Do_NB_Iter
 * (src/Newmark_Beta.cpp:69-132, src/Resid.cpp:21-186). */
fjsph_status fjsph_boundary_update(fjsph_ctx* ctx);
/* get_acc_and_Rrho (src/Resid.cpp:426-469) on pnp1 with the given npd. */
fjsph_status fjsph_acc_rrho(fjsph_ctx* ctx, double npd);
/* Newmark_Beta::Do_NB_Iter (src/Newmark_Beta.cpp:54-301) followed by the
 * residual of Check_Error (src/Newmark_Beta.cpp:16-25): returns
 * log10(sqrt(sum|x'-x_prev|^2 / n_fluid)). */
fjsph_status fjsph_nb_iter(fjsph_ctx* ctx, double npd, double* log_error);
/* Integrator::find_timestep (src/Integration.cpp:370-443) on pnp1. */
fjsph_status fjsph_find_timestep(fjsph_ctx* ctx, double* dt);

/* Integrator::integrate_no_update (src/Integration.cpp:27-107): the start-up
 * pass the driver makes before its frame loop (src/FJSPH.cpp:183). */
fjsph_status fjsph_integrate_no_update(fjsph_ctx* ctx, fjsph_step_info* info);
/* Integrator::integrate (src/Integration.cpp:233-303): one full step,
 * including update_data's pn <- pnp1 and the adaptive CFL. */
fjsph_status fjsph_integrate_step(fjsph_ctx* ctx, fjsph_step_info* info);
/* Advance the frame clock: last_frame_time += frame_time_interval,
 * current_frame + 1 (src/FJSPH.cpp:322-323 and the frame loop). */
fjsph_status fjsph_advance_frame(fjsph_ctx* ctx);
/* Simulation clock and timestep (INTEG_SETT current_time / delta_t). */
fjsph_status fjsph_get_time(fjsph_ctx* ctx, double* time, double* delta_t, double* cfl);
fjsph_status fjsph_set_time(fjsph_ctx* ctx, double time, double delta_t, double cfl);

/* The integrator state a run carries from step to step and from frame to
 * frame (INTEG_SETT current_time, last_frame_time, delta_t, cfl,
 * current_frame, n_stable, n_unstable, src/Var.h:63-92; svar.part_id, the
 * next particle id, src/Var.h:365): what Write_HDF5 / Read_HDF5 carry
 * across a restart (src/H5IO.cpp:915-1318). */
typedef struct fjsph_clock {
    double current_time, last_frame_time, delta_t, cfl;
    int64_t current_frame, n_stable, n_unstable, next_part_id;
} fjsph_clock;
fjsph_status fjsph_get_clock(fjsph_ctx* ctx, fjsph_clock* out);
fjsph_status fjsph_set_clock(fjsph_ctx* ctx, const fjsph_clock* in);
/* The settings the context was created with. */
fjsph_status fjsph_get_settings(fjsph_ctx* ctx, fjsph_settings* out);
/* Current blocks (LIMITS: ranges follow insertions and deletions) and
 * their planes. */
fjsph_status fjsph_get_blocks(fjsph_ctx* ctx, int32_t* n_blocks, fjsph_block* blocks, int32_t capacity);
fjsph_status fjsph_get_block_planes(fjsph_ctx* ctx, int32_t block, fjsph_block_planes* out);
/* Time-velocity table of a block (bound_block::times, ::vels; 3 values
 * per time, as fjsph_set_block_motion takes them); times/vels may be null
 * to query n_times. */
fjsph_status fjsph_get_block_motion(fjsph_ctx* ctx, int32_t block, int32_t* n_times, double* times, double* vels);
/* Current inlet columns of block `block` in host indices, the layout of
 * fjsph_set_inlet (bound_block::back, ::buffer).  A column that advanced
 * without a new particle (max_points reached) lists its deepest buffer
 * particle in every place past its last one, as the reference's list does
 * (src/shapes/inlet.cpp:604-615).  n_cols = 0 when the block has no
 * inlet; back and buffer may be null to query the sizes. */
fjsph_status fjsph_get_inlet(fjsph_ctx* ctx, int32_t block, int32_t* n_cols, int32_t* n_buf, int64_t* back,
                             int64_t* buffer);

/* ---- tile blocks and their autotuner --------------------------------------- */

/* Independent simulations in one context: group[i] (0..n_groups-1, every
 * group non-empty) assigns host particle i to a simulation.  Replaces the
 * reference's one process per simulation (src/FJSPH.cpp): each group keeps
 * its own integrator state -- time step (find_timestep over its own
 * particles), CFL control and counters, clock and frame end
 * (src/Integration.cpp:233-303, 370-443), Newmark-Beta residual and
 * iteration count (a group whose residual has converged keeps its state
 * while the others iterate; a diverged group rolls back alone with its own
 * time step halved, src/Newmark_Beta.cpp:10-52) and npd
 * (src/Shifting.cpp:16-68, 121) -- and one fjsph_integrate_step advances
 * every group by one step.  All pair passes run as one launch over every
 * group.  Groups must not interact: particles of different groups farther
 * apart than 2H (the engine bins each group apart, so nearer particles of
 * two groups are simply not neighbours).  Requires the Newmark-Beta
 * integrator, no boundary particles, no inlets or deletion planes; the
 * per-pass entry points fjsph_acc_rrho, fjsph_nb_iter and
 * fjsph_find_timestep are refused; fjsph_set_time sets every group.
 * n_groups = 0 returns to one simulation.  fjsph_integrate_step's info
 * then describes the batch: largest residual, iteration count, density
 * variation and accelerations, smallest time step, CFL number and clock,
 * summed rollbacks, mean npd. */
fjsph_status fjsph_set_groups(fjsph_ctx* ctx, int32_t n_groups, const int32_t* group);
/* The step information of groups [first, first + count) after the last
 * step (out: count records). */
fjsph_status fjsph_get_group_info(fjsph_ctx* ctx, int32_t first, int32_t count, fjsph_step_info* out);

/* Targets per tile block of the pair passes (the cells of one block row
 * whose particles share one candidate union in LDS); 0 restores the
 * default, 256.  No counterpart in the reference.  Results change only in
 * the order of summation.  A block spans at most 256 cells whatever the
 * setting, and a block whose candidate union overflows the 16-bit list
 * entries (a clustered region) makes the build halve the block length
 * and try again, down to one cell. */
fjsph_status fjsph_set_block_targets(fjsph_ctx* ctx, int32_t targets);
fjsph_status fjsph_get_block_targets(fjsph_ctx* ctx, int32_t* targets);
/* The first list build of a step moves the particles into cell order
 * (a permutation of every per-particle array) once every `steps` steps;
 * the builds in between index the particles through the cell order.
 * Default 4 (the first step after an upload or this call reorders).  No
 * counterpart in the reference; results change only in the order of
 * summation. */
fjsph_status fjsph_set_reorder_interval(fjsph_ctx* ctx, int32_t steps);
/* Autotuner of the tile blocks: for each candidate, one step's pair work
 * (two list builds with their pre-passes, the force pass once per Newmark
 * iteration of the last step) timed over `reps` repetitions on the current
 * state; ms[c] receives the time per repetition, *best the fastest
 * candidate, which is then set.  The particle state is restored bit for bit
 * afterwards.  Allocates a device snapshot of the per-particle arrays. */
fjsph_status fjsph_autotune_blocks(fjsph_ctx* ctx, int32_t n_cand, const int32_t* targets, int32_t reps, double* ms,
                                   int32_t* best);

/* ---- implicit particle tracking (IPT) ------------------------------------ */

/* IPT_SETT (src/Var.h:313-331): with using_ipt on and the mesh aero source,
 * fluid particles removed by a deletion plane are tracked through the gas
 * mesh as point droplets (IPT::Integrate, src/IPT.cpp:870-1106) in
 * update_data (src/Integration.cpp:150-168), on the GPU, one thread per
 * particle; the tracks' ends are kept as records. */
typedef struct fjsph_ipt_settings {
    int32_t using_ipt;  /* "Transition to IPT" (0)                         */
    int32_t eq_order;   /* "Velocity equation order": 1 or 2 (2)           */
    double max_x;       /* "Maximum x trajectory coordinate" (9999999)     */
    double relax;       /* IPT_SETT::relax (0.6)                           */
    int32_t n_relax;    /* IPT_SETT::n_relax (5)                           */
    int32_t max_steps;  /* cell-to-cell steps per track before it stops; no
                           counterpart in the reference (1000000)          */
} fjsph_ipt_settings;

enum { FJSPH_IPT_SUCCESS = 1, FJSPH_IPT_FAILED = 2, FJSPH_IPT_STOPPED = 3 };

/* The end of a track (the last entry Terminate_Particle keeps,
 * src/IPT.cpp:735-818): status success (left the mesh, or past max_x),
 * failed (no face crossed, |v| > 1e3, or a step longer than the mesh's
 * maximum length) or stopped (step cap); the cell (negative: the boundary
 * marker of the face it left through) and the last face crossed. */
typedef struct fjsph_ipt_record {
    int64_t part_id;
    int32_t status, steps, cell, face;
    double t0, t;
    double x0[3], x[3], v[3];
} fjsph_ipt_record;

void fjsph_ipt_settings_default(fjsph_ipt_settings* s);
fjsph_status fjsph_set_ipt(fjsph_ctx* ctx, const fjsph_ipt_settings* s);
/* Track n particles from the given positions, velocities and cells of the
 * current mesh (the cells' gas values start the tracks), at time t0; the
 * records go to out (n of them).  Diameter and area are Set_Values' ipt_diam
 * and ipt_area (src/IO.cpp:126-127). */
fjsph_status fjsph_ipt_track(fjsph_ctx* ctx, int64_t n, const double* x, const double* v, const int32_t* cell,
                             const int64_t* part_id, double t0, fjsph_ipt_record* out);
/* Records of the particles tracked by update_data since the last clear. */
fjsph_status fjsph_ipt_count(fjsph_ctx* ctx, int64_t* n_records, int64_t* n_success, int64_t* n_failed);
fjsph_status fjsph_ipt_records(fjsph_ctx* ctx, int64_t first, int64_t count, fjsph_ipt_record* out);
fjsph_status fjsph_ipt_clear(fjsph_ctx* ctx);

/* ---- aerodynamic mesh (aero_source = FJSPH_MESH_INFL) --------------------- */

/* The MESH of src/Var.h:396-451 as flat arrays (host order, 0-based):
 *   verts[nverts*dim];
 *   face_off[nfaces+1], face_verts[face_off[nfaces]]: the vertices of each
 *     face (3D: polygons, tested as src/Geometry.cpp:485-575 does, through
 *     vertices 0, 1, 2 and the last; 2D: edges of two vertices);
 *   face_lr[2*nfaces]: left cell, right cell or a negative boundary marker
 *     (leftright, -1 inner wall, -2 outer boundary);
 *   cell_off[ncells+1], cell_faces[cell_off[ncells]]: the faces of each cell;
 *   vert_vel[nverts*dim], vert_p[nverts], vert_rho[nverts]: the gas solution
 *     at the vertices, averaged to the cells as Average_Point_Data_to_Cell
 *     (src/CDFIO.cpp:112-155) does.
 * Replaces the mesh readers of src/FJSPH.cpp:66-90 and the cell KD-tree.
 * The lookup (FindCell, src/Containment.cpp:706-820) runs on every
 * get_aero_velocity: FREE particles with lam_nb <= lam_cutoff get the
 * values of the cell that contains them (the previous cell first, then the
 * cells near their position); particles found in no cell keep their values
 * and are counted (deletion is not built). */
fjsph_status fjsph_set_mesh(fjsph_ctx* ctx, int64_t nverts, const double* verts, int64_t nfaces,
                            const int64_t* face_off, const int64_t* face_verts, const int64_t* face_lr,
                            int64_t ncells, const int64_t* cell_off, const int64_t* cell_faces,
                            const double* vert_vel, const double* vert_p, const double* vert_rho);
/* Cell centres and values as averaged: centre, vel (ncells*dim), p, rho (ncells). */
fjsph_status fjsph_get_mesh_cells(fjsph_ctx* ctx, double* centre, double* vel, double* p, double* rho);
/* Aero particles found in no cell by the last lookup. */
fjsph_status fjsph_mesh_lost(fjsph_ctx* ctx, int64_t* lost);

/* What FindCell did with the particles it placed in no cell
 * (src/Containment.cpp:734-805, src/Resid.cpp:481-519), counted since the
 * state was uploaded: a ray from the particle to each of the 500 nearest
 * cell centres is tested against those cells' boundary faces; a crossing of
 * a far-field face (right cell -2) deletes the particle, a wall face (-1)
 * sets FJSPH_F_INTERNAL; with no net crossing FJSPH_F_NFAILED counts up and
 * the particle is deleted once it is past 10.  Deleted particles leave both
 * states, the neighbour lists are built again and dSPH_PreStep runs again,
 * inside the same lookup (Integration.cpp:47-48, 82-83). */
typedef struct fjsph_lost_info {
    int64_t lost;     /* aero particles in no cell at the last lookup */
    int64_t deleted;  /* particles the lookups deleted */
    int64_t internal; /* lookups that found a wall crossing */
    int64_t failed;   /* lookups that found no net crossing */
    int64_t first_cell;         /* particles leaving the pipe in no cell (FirstCell's boundary branch) */
    int64_t first_cell_deleted; /* of those, deleted across a far-field face */
} fjsph_lost_info;
fjsph_status fjsph_mesh_lost_info(fjsph_ctx* ctx, fjsph_lost_info* info);
/* The mesh field at n host points (n*dim): the containing cell (-1 if none)
 * and its (vel[dim], p, rho) into vals (n*(dim+2), zeros if none).  No
 * counterpart in the reference; used to label the surrogate's training
 * samples. */
fjsph_status fjsph_mesh_eval(fjsph_ctx* ctx, int64_t n, const double* points, int32_t* cells, double* vals);

/* ---- surrogate of the mesh field (aero_source = FJSPH_SURROGATE) ----------- */

/* No counterpart in the reference: replaces the mesh lookup by an MLP of
 * n_hidden (1..8) layers of width 64 with GELU (erf form), evaluated with
 * BF16 MFMA and FP32 accumulation.  Input: the position normalised as
 * (x - in_shift) / in_scale (dim values); output: (vel[dim], p, rho) =
 * y * out_scale + out_shift (dim + 2 values).  weights: FP32, each layer
 * [out][in] row-major, in order W_1 [64][dim], W_2 .. W_n [64][64],
 * W_out [dim+2][64]; biases: b_1 .. b_n [64], b_out [dim+2].  Applied to the
 * same particles as the mesh lookup (FREE, lam_nb <= lam_cutoff). */
fjsph_status fjsph_set_surrogate(fjsph_ctx* ctx, int32_t n_hidden, const float* weights, const float* biases,
                                 const double* in_shift, const double* in_scale, const double* out_shift,
                                 const double* out_scale);
/* Evaluate the surrogate at n host points (n*dim) into out (n*(dim+2)). */
fjsph_status fjsph_surrogate_eval(fjsph_ctx* ctx, int64_t n, const double* points, double* out);
/* Train the surrogate on the GPU with the engine's own kernels (BF16 MFMA
 * forward and backward passes with FP32 accumulation, Adam on FP32 master
 * parameters, cosine learning-rate schedule from lr to 0) on n samples:
 * points (n*dim) and values (n*(dim+2), the (vel, p, rho) to fit),
 * normalised with the given shifts and scales.  Each of `steps` steps
 * draws `batch` samples (a multiple of 64) by a hash of seed, step and
 * position, and minimises their mean squared error.  weights/biases (the
 * layout of fjsph_set_surrogate): initial values in, trained values out;
 * loss (steps values, may be NULL): the mean squared error of each step's
 * batch before its update.  No counterpart in the reference. */
fjsph_status fjsph_surrogate_train(fjsph_ctx* ctx, int64_t n, const double* points, const double* values,
                                   int32_t n_hidden, int32_t steps, int32_t batch, double lr, uint64_t seed,
                                   const double* in_shift, const double* in_scale, const double* out_shift,
                                   const double* out_scale, float* weights, float* biases, double* loss);
/* Gradients of the mean squared error over one batch of `batch` (a
 * multiple of 64) normalised samples xn (batch*dim) with targets tn
 * (batch*(dim+2)) for the given parameters, by the training kernels: into
 * dweights/dbiases (the layout of fjsph_set_surrogate), the loss into
 * *loss.  For checks of the training kernels.  No counterpart in the
 * reference. */
fjsph_status fjsph_surrogate_grad(fjsph_ctx* ctx, int32_t n_hidden, int32_t batch, const float* xn, const float* tn,
                                  const float* weights, const float* biases, float* dweights, float* dbiases,
                                  double* loss);

/* Largest batch of fjsph_surrogate_grad_noalloc, and the size of the
 * device scratch array it works in (one per GPU, held until the process
 * ends). */
#define FJSPH_NOALLOC_MAX_BATCH 4096
#define FJSPH_NOALLOC_SCRATCH_BYTES (24u << 20)

/* fjsph_surrogate_grad without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving or
 * training loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same results as fjsph_surrogate_grad, but every pointer is
 * DEVICE memory: xn (batch*dim), tn (batch*(dim+2)), weights, biases in;
 * dweights, dbiases and *loss (may be NULL) out.  The activations, the
 * per-wave gradient partials and the packed weight tiles live in a
 * file-scope device array of FJSPH_NOALLOC_SCRATCH_BYTES; batch must be a
 * multiple of 64 and at most FJSPH_NOALLOC_MAX_BATCH (FJSPH_ERR_ARG
 * otherwise), n_hidden 1..8.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several contexts or threads) race.
 * Synchronous on the context stream, like the other calls. */
fjsph_status fjsph_surrogate_grad_noalloc(fjsph_ctx* ctx, int32_t n_hidden, int32_t batch, const float* xn,
                                          const float* tn, const float* weights, const float* biases,
                                          float* dweights, float* dbiases, double* loss);

/* ---- measurement ---------------------------------------------------------- */

/* Pass ids of the engine's own timers (HIP events on the context stream).
 * FJSPH_T_FORCE is the pair kernel of get_acc_and_Rrho alone;
 * FJSPH_T_AERO the gas-field lookup of get_aero_velocity (mesh FindCell or
 * the surrogate; the constant source is part of FJSPH_T_DSPH);
 * FJSPH_T_UPDATE the per-particle tail (aerodynamic force, artificial
 * viscosity, gravity) and the Newmark-Beta update. */
enum {
    FJSPH_T_NEIGHBOUR = 0, FJSPH_T_DSPH, FJSPH_T_SURFACE, FJSPH_T_DISSIPATION,
    FJSPH_T_SHIFT, FJSPH_T_BOUNDARY, FJSPH_T_FORCE, FJSPH_T_REDUCE,
    FJSPH_T_UPDATE, FJSPH_T_AERO, FJSPH_T_COUNT
};
/* Enable (1) / disable (0) per-pass timing; enabling clears the totals. */
fjsph_status fjsph_enable_timers(fjsph_ctx* ctx, int32_t on);
/* Total milliseconds and launch counts per pass since timers were enabled. */
fjsph_status fjsph_get_timers(fjsph_ctx* ctx, double* ms, int64_t* launches);
/* Number of particles processed per launch of a pass id, and the
 * algorithmic bytes per particle the engine moves in that pass. */
fjsph_status fjsph_pass_bytes(fjsph_ctx* ctx, int32_t pass, double* bytes_per_particle);
/* Current neighbour statistics: total pairs (self excluded) and max count. */
fjsph_status fjsph_neighbour_stats(fjsph_ctx* ctx, int64_t* pairs, int32_t* max_count);

#ifdef __cplusplus
}
#endif

#endif /* FJSPH_ENGINE_H */
