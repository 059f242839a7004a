/*
 * Restart files of the FJSPH step engine, in the reference's layout: the
 * `<prefix>_particles.h5` file that Write_HDF5 writes after every frame and
 * Read_HDF5 reads back (src/H5IO.cpp:915-1318).
 *
 * Layout (HDF5): root attributes (Write_HDF5_Attributes,
 * src/H5IO.cpp:915-1025); one group per block in LIMITS order, "Boundary k"
 * for the boundary blocks and "Fluid k" for the fluid ones, each holding
 * 1-D datasets of its particles (Write_Zone_Data, src/H5IO.cpp:395-523),
 * "Back index" / "Buffer index" for an inlet block (Write_Inlet_Data,
 * src/H5IO.cpp:525-541) and the block attributes (Write_Zone_Attributes,
 * src/H5IO.cpp:1027-1088).  Names, HDF5 types and the order of writing are
 * the reference's, so the files are interchangeable.
 *
 * Two levels: the *_host calls move a host image of the state (no GPU
 * needed); fjsph_restart_write / fjsph_restart_read move an engine
 * context's state through that image.  Built as libfjsph_restart.so, which
 * links libfjsph_engine.so and libhdf5.
 */
#ifndef FJSPH_RESTART_H
#define FJSPH_RESTART_H

#include "fjsph_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FJSPH_RESTART_ABI_VERSION 1
/* Buffer particles per inlet column in the file: Write_Inlet_Data and
 * Read_Inlet_Data fix it at 4 (src/H5IO.cpp:530-540, 903-910). */
#define FJSPH_RESTART_NBUF 4

int32_t fjsph_restart_abi_version(void);

/*
 * What the file holds beside the engine state: the driver's I/O names and
 * settings (IN_OUT_SETT, IPT_SETT and a few SIM/AERO fields, src/Var.h).
 * Null or empty strings are not written, as Write_String_Attribute skips
 * empty ones.  fjsph_restart_meta_default fills the reference's defaults.
 */
typedef struct fjsph_restart_meta {
    const char* input_file;     /* "Input para filename"                    */
    const char* fluid_file;     /* "Input fluid definition filename"        */
    const char* bound_file;     /* "Input boundary definition filename"     */
    const char* tau_mesh;       /* "Primary grid face filename"             */
    const char* tau_bmap;       /* "Boundary mapping filename"              */
    const char* tau_sol;        /* "Restart-data prefix"                    */
    const char* foam_dir;       /* "OpenFOAM input directory"               */
    const char* foam_sol;       /* "OpenFOAM solution directory"            */
    const char* vlm_file;       /* "VLM definition filename"                */
    const char* output_prefix;  /* "Output files prefix"                    */
    const char* restart_prefix; /* "SPH restart prefix"                     */
    const char* output_names;   /* "Variable list"                          */
    const char* solver_name;    /* "SPH integration solver"; null: from settings */
    const char* aero_case_name; /* "SPH aerodynamic case"; null: from settings   */
    const char* const* block_names; /* "Block name" per block, or null      */
    const int32_t* block_types;     /* "Block type" per block, or null: an
                                       inlet block 6 (inletZone), others 0  */
    const int32_t* particle_order;  /* "Lattice or HCPL packing", or null   */
    double scale;               /* "Grid scale" (1)                          */
    double angle_alpha;         /* "Solution angle of attack" (0)            */
    double hydro_height;        /* "Hydrostatic height" (-1)                 */
    double bound_step_factor;   /* "SPH boundary spacing factor" (1)         */
    double restart_tol;         /* "SPH restart fit tolerance" (0.001)       */
    double contangb;            /* "SPH surface tension contact angle" (150) */
    double v_ref;               /* "Reference velocity" (0)                  */
    double M_ref;               /* "Reference Mach number" (-1)              */
    double max_x_sph;           /* "SPH tracking conversion x coordinate"    */
    double max_x;               /* "Maximum x trajectory coordinate"         */
    double offset_vec[3];       /* "SPH global offset coordinate"            */
    int32_t offset_axis;        /* "Dimension offset vector"                 */
    int32_t single_file;        /* "Single file for output" (1)              */
    int32_t out_encoding;       /* "SPH output encoding" (1)                 */
    int32_t max_frames;         /* "SPH frame count"                         */
    int32_t init_hydro_pressure;/* "Init hydrostatic pressure" (0)           */
    int32_t use_dx;             /* "SPH SP diameter definition" (0)          */
    int32_t bound_solver;       /* "SPH boundary solver" (0)                 */
    int32_t using_ipt, ipt_eq_order, part_out, streak_out, cells_out;
    int32_t foam_is_binary, foam_label_size, foam_scalar_size, foam_buoyant_sim, foam_is_incomp;
} fjsph_restart_meta;

void fjsph_restart_meta_default(fjsph_restart_meta* m, int32_t dim);

/*
 * The root attributes Read_HDF5 reads (src/H5IO.cpp:1230-1247), and the
 * counts a reader needs to size its arrays.  dim is 3 when the first
 * group has a "Position coordinate z" dataset, else 2.  clock.last_frame_time
 * is clock.current_time, as Read_HDF5 sets it; clock.delta_t is 0 (the
 * file has none: the first step finds its own).
 */
typedef struct fjsph_restart_header {
    int32_t dim;
    int32_t n_bound_blocks, n_fluid_blocks;
    int64_t bound_points, fluid_points; /* from the groups' lengths */
    fjsph_clock clock;
    double particle_step;       /* "SPH initial spacing"      */
    double frame_time_interval; /* "SPH frame time interval"  */
    double scale;               /* "Grid scale"               */
    int32_t single_file;        /* "Single file for output"   */
} fjsph_restart_header;

fjsph_status fjsph_restart_read_header(const char* path, fjsph_restart_header* out);

/*
 * Host image of the state a restart file carries.  Arrays are in host
 * order: n particles, blocks tiling [0, n) with the boundary blocks first.
 * Vectors are interleaved (x[i*dim + k]).  Inlets: n_cols[k] columns for
 * block k (0 for a block without one), back[k] (n_cols[k] host indices)
 * and buffer[k] (n_cols[k] * FJSPH_RESTART_NBUF).  Motion: n_times[k] and,
 * when nonzero, times[k] (n_times[k]) and vels[k] (3 per time).
 * For *_host reads, the caller allocates every array from the header's
 * counts (and a first call with the arrays null fills n_blocks, the
 * blocks, n_cols and n_times).
 */
typedef struct fjsph_restart_host {
    int32_t dim;
    int64_t n, n_bound;
    fjsph_settings settings;
    fjsph_clock clock;
    int64_t max_points;
    int32_t n_blocks;
    fjsph_block* blocks;
    fjsph_block_planes* planes;
    int32_t* n_times;
    double** times;
    double** vels;
    int32_t* n_cols;
    int64_t** back;
    int64_t** buffer;
    double *x, *v, *acc, *p, *rho, *rrho, *m;
    int32_t* b;
    int64_t* pid;
    int64_t* cell_id;
    double *cell_v, *cell_p, *cell_rho;
} fjsph_restart_host;

/* Write the file (an existing file is replaced, as Write_HDF5 does). */
fjsph_status fjsph_restart_write_host(const char* path, const fjsph_restart_host* st, const fjsph_restart_meta* meta);
/* Read the file into st.  Fields of st->settings that the file holds
 * (particle_step, frame_time_interval) are set; the rest are left as given.
 * Blocks' first/last come from the groups' lengths; their other fields,
 * the planes and the motion tables come from the file's block attributes. */
fjsph_status fjsph_restart_read_host(const char* path, fjsph_restart_host* st);

/* Engine level.  Write: the state of ctx at the end of a step (pnp1) with
 * its blocks, planes, motion tables, inlets and clock. */
fjsph_status fjsph_restart_write(fjsph_ctx* ctx, const char* path, const fjsph_restart_meta* meta);
/* Read: what Init_Particles_Restart + Read_HDF5 do (src/FJSPH.cpp:130-135):
 * the particles (pn = pnp1), block ranges, inlet columns and clock of the
 * file into ctx, which was created with the settings the run uses (its
 * particle_step should be the header's).  blocks/planes give the block
 * definitions (the bmap files of the reference): solver, no-slip, velocity
 * and planes; null takes them from the file's block attributes.  The mesh,
 * surrogate and max_points are set by the caller as for a fresh start. */
fjsph_status fjsph_restart_read(fjsph_ctx* ctx, const char* path, int32_t n_blocks, const fjsph_block* blocks,
                                const fjsph_block_planes* planes);
/* Text of the last restart error of this thread. */
const char* fjsph_restart_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FJSPH_RESTART_H */
