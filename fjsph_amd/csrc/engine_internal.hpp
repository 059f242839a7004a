/*
 * Internal types of the MI355X WCSPH engine: device state layout, constants
 * passed to kernels, and the kernel launch wrappers (kernels.hip).
 *
 * Device layout (all arrays in cell-sorted slot order):
 *   - one DevState per reference state (pn, pnp1): every SPHPart field the
 *     step reads, FP64, vectors interleaved per particle (x[s*D + k]);
 *   - x, v, rho, p of pnp1 exist twice (cur / next) so the fused
 *     force + Newmark kernel reads iterate k and writes k+1 without races;
 *   - neighbour lists as rows of four 16-bit union indices per target,
 *     row-major by row and column-major by the target's position in cell
 *     order (see NbrList).
 */
#ifndef FJSPH_ENGINE_INTERNAL_HPP
#define FJSPH_ENGINE_INTERNAL_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fjsph_engine.h"

namespace fj {

/* Constants of one launch (derived once per context). */
struct Consts {
    double H, H_sq, sr, Wc, W_dx, B, gam, rho0, c0_sq, press_back, nu, visc_alpha, dsph_cont;
    double rho_min, rho_max, dx, particle_step, lam_cutoff, interp_fac, i_n_full;
    double A_plate, A_sphere, L, sos, gamma_g, mu_g, rho_g, ycoef, Cb, max_shift_vel, lam_st, sig;
    double grav[3];
    double v_inf[3];
    int32_t eos, aero_case, aero_source, use_lam, use_TAB_def, ale;
    int64_t n_bound;
};

struct DevState {
    double *x, *v, *acc, *Af, *aVisc, *Rrho, *rho, *p, *m, *deltaD, *L, *gradRho, *norm;
    double *lam, *lam_nb, *colourG, *colour, *kernsum, *curve, *norm_curve, *woccl, *pDist;
    double *cellV, *cellP, *cellRho, *vPert;
    int32_t *surf, *cellID, *surfzone, *b, *pid, *nfail, *internal;
};

/* Neighbour lists in rows of 4 entries: row r of the target at position p
 * of the target order (TileGrid::vorder) is lu4[r*npad + p], four 16-bit
 * entries k*TILE + t (slot t of LDS tile k of the candidate union of the
 * target's tile block), ascending, all in one tile; unused entries hold
 * NO_ENTRY.  The targets of one wave
 * hold the same number of rows per tile (padding rows are all NO_ENTRY).
 * count[i] is the row count of slot i, nnb[i] its neighbours (self
 * excluded).  The frozen distance of src/Resid.cpp:288-289 is not stored:
 * it is recomputed from the build positions Aux::xb. */
struct NbrList {
    uint64_t* lu4;
    int32_t* count;
    int32_t* nnb;
    int64_t npad;
    int32_t cap;  /* entries per slot (4 per row) */
    int32_t f32;  /* mixed precision: FP32 pair terms */
};

/* Candidate-union tile shared by the list build and every pair pass
 * (a multiple of 4, the group size in which union slots are dealt to
 * the tiles). */
#ifndef FJ_TILE
#define FJ_TILE 496
#endif
constexpr int TILE = FJ_TILE;
static_assert(TILE % 4 == 0, "tile slots come in groups of 4");
constexpr uint32_t NO_ENTRY = 0xFFFF;
/* Longest tile block in cells: every coordinate of a block's candidate
 * union then lies within 1000 H of its first target, the bound the FP32
 * prefilter of the list build relies on. */
constexpr int32_t kMaxBlockCells = 256;

/* Tile decomposition of the last neighbour build: block b covers cells
 * [q*C, q*C + C) of cell row b / nchunk, q = b % nchunk. */
struct TileGrid {
    int32_t dims[3];
    int32_t C;
    int32_t nchunk;
    int64_t nblocks;
    const int32_t* start; /* cell_start, ncell + 1 */
    const int32_t* order; /* slots in cell order (the candidate unions) */
    const int32_t* cell;  /* cell of every slot */
    int32_t* vorder;      /* target order of the lists (order[] at the build) */
    int32_t* pos;         /* position of every slot in vorder[] */
    /* independent simulations (null otherwise): the group of every cell
     * layer along the last axis (-1 between groups), whether each group is
     * still iterating, and each group's numerical particle density */
    const int32_t* lgrp;
    const int32_t* gactive;
    const double* gnpd;
    const int32_t* gbuild; /* list build: only the blocks of these groups (null: all) */
};

/* Independent simulations in one context, for the per-slot Newmark update:
 * the group of every slot, whether the group is still iterating (a group
 * that is not keeps its state), each group's time step, and a per-slot
 * array that takes the squared position change for the per-group
 * residuals.  grp == nullptr: one simulation. */
struct DevGroups {
    const int32_t* grp;
    const int32_t* active;
    const double* dt;
    double* d2;
};

/* Per-particle quantities derived from the current rho, p, m, b (and vPert),
 * refreshed by launch_aux before every pass that reads them. */
struct Aux {
    double* vol;     /* m / rho            */
    double* P;       /* p / rho^2          */
    double* invrho;  /* 1 / rho            */
    float4* fq;      /* {m (negative for BOUND), p/rho^2, 1/rho, m/rho} in FP32 */
    float4* fale;    /* {vPert, rho} in FP32 (ALE)                          */
    float4* fs1;     /* {gradRho, lambda} in FP32 (mixed), by dSPH_PreStep   */
    float4* fs2;     /* {m/rho, sqrt(B gamma / rho), m (negative for PISTON
                        and BOUND), 0} in FP32 (mixed)                      */
    double* nhat;    /* normalised surface normal (after Detect_Surface)     */
    double* lnb;     /* dSPH_PreStep: correction matrix over non-boundary
                        neighbours (D x D per slot), for its eigenvalue     */
    double* xb;      /* positions at the last neighbour build (D per slot)   */
    int4* spend;     /* Detect_Surface: the first four neighbours whose FP32
                        point-T test was undecided                          */
    int32_t* sund;   /* the count (sund[0]) and slots (sund[1..]) of the
                        targets with an undecided point-T test, appended by
                        k_surface1_m                                        */
};

/* Per-block boundary settings visible to kernels. */
struct DevBlocks {
    int32_t* solver;   /* per block */
    int32_t* no_slip;  /* per block */
    double* vel;       /* per block, 3 each (current velocity) */
    int32_t n;
};

/* Aerodynamic mesh on the device (vertices with 3 components in 2D too,
 * z = 0), CSR faces and cells, cell values, and the voxel candidate grid:
 * voxel v = (z*gdims[1] + y)*gdims[0] + x lists gcells[gstart[v] ..
 * gstart[v+1]), the cells whose bounding box overlaps it, ascending. */
struct DevMesh {
    int32_t ncells;
    const double* verts;
    const int32_t *face_off, *face_v;
    const int32_t *cell_off, *cell_f;
    const double *cVel, *cP, *cRho;
    double glo[3];
    double ginv;
    int32_t gdims[3];
    const int32_t *gstart, *gcells;
    const int32_t* big; /* cells too large for the grid, scanned by every lookup */
    int32_t nbig;
    const int32_t* face_lr;  /* left, right cell per face (right < 0 on the boundary) */
    const double* cCentre;   /* 3 per cell (z = 0 in 2D) */
    /* cell centres binned on a uniform grid (nearest-centre queries) */
    double clo[3];
    double cinv;
    int32_t cdims[3];
    const int32_t *cstart, *ccells;
};

/* IPT constants (IPT_SETT, src/Var.h:313-331; Set_Values ipt_diam,
 * src/IO.cpp:126-127; MESH::maxlength, src/CDFIO.cpp:867-929, 1117-1220). */
struct IptParams {
    double grav[3];
    double mu_g, rho_l, relax, max_x, max_len, diam, area;
    int32_t eq_order, n_relax, max_subits, max_steps;
};

/* Start of an IPT track (IPTPart from an SPH particle, src/Var.h:734-762). */
struct IptStart {
    double x[3], v[3], cellV[3];
    double cellRho, t;
    int64_t part_id;
    int32_t cell, pad;
};

/* MLP surrogate of the mesh field (surrogate.hip): BF16 weights already in
 * the MFMA A-operand layout, FP32 biases, input/output scaling. */
struct Surrogate {
    const void* w;   /* bf16 tiles: layer 1 (4 tiles), hidden layers (8 each), output (2) */
    const float* b;  /* n_hidden x 64 */
    const float* bo; /* 16 (D + 2 used) */
    int32_t n_hidden;
    double in_shift[3], in_scale_inv[3];
    double out_shift[5], out_scale[5];
};

/* Next Newmark iterate (k_nb_update writes, the engine swaps it in). */
struct NextState {
    double *x, *v, *rho, *p;
};

/* Per-block planes on the device (inlet.hip): 4 doubles each (normal, c)
 * for deletion, insertion and the aero plane; inlet mode and buffer depth. */
struct DevPlanes {
    const double *del, *ins, *aero;
    const int32_t *mode, *nbuf;
};

/* Segmented list of the aero particles (surrogate.hip): segment length and
 * the stride of its counters (one 128-byte line each). */
constexpr int64_t AERO_SEG = 65536;
constexpr int64_t AERO_CSTRIDE = 32;

/* Device buffers of one surrogate training step (surrogate_train.hip). */
struct MlpTrain {
    const float* X;      /* samples: normalised inputs, 4 per sample (D used) */
    const float* T;      /* normalised targets, 8 per sample (D + 2 used)     */
    const int32_t* idx;  /* batch: sample of each position                    */
    const void* wf;      /* forward weight tiles (BF16, A layout)             */
    const void* wt;      /* backward tiles of the transposes                  */
    const float* b;      /* hidden biases, n_hidden x 64                      */
    const float* bo;     /* output biases (D + 2)                             */
    float* Z;            /* pre-activations, n_hidden x batch x 64            */
    __bf16* H;           /* activations (H_0 = input), (n_hidden + 1) x batch x 64 */
    float* dWp;          /* per-wave weight-gradient partials, nw each        */
    float* dbp;          /* per-wave bias-gradient partials, nb each          */
    float* lossp;        /* per-wave sums of squared errors                   */
    int64_t nw, nb;
    int32_t n_hidden, batch;
};

/* Scalars kept on the device between kernels. */
struct DevScalars {
    double npd;       /* numerical particle density of the last dSPH pass */
    double residual;  /* sum |x' - x_prev|^2 of the last Newmark iterate */
    double red[8];    /* find_timestep reduction results */
    double bbox[6];   /* lo[3], hi[3] */
    int32_t max_count;
    int32_t overflow; /* largest candidate union beyond 0xFFFF / TILE tiles (0 if none) */
    int32_t mesh_lost; /* aero particles in no mesh cell at the last lookup */
    int32_t lost_del;  /* of those, deleted by the boundary test */
    int32_t lost_int;  /* of those, found across a wall face */
    int32_t lost_fail; /* of those, with no net boundary crossing */
    int32_t lost_err;  /* 1: a nearest-centre search failed (too many equal distances);
                          2: FirstCell found no boundary crossing */
    int32_t fresh_n;   /* particles leaving the pipe listed for FirstCell */
    int32_t fc_count;  /* of those, in no cell (FirstCell's boundary branch) */
    int32_t fc_del;    /* of those, across a far-field face */
};

/* ---- launch wrappers (kernels.hip); every one enqueues on `st` ---------- */

void launch_bbox(int dim, const double* x, int64_t n, double* partial, int nparts, double* out, hipStream_t st);
void launch_cells(int dim, const double* x, int64_t n, const double* lo, double inv_cs, const int32_t* dims,
                  int32_t* cell, int32_t* cell_count, hipStream_t st, const int32_t* grp = nullptr,
                  const double* glo = nullptr, const int32_t* gl0 = nullptr);
/* Independent simulations: per-group bounding boxes (6 per group), sums
 * over segments, and find_timestep's reduction per group (8 per group). */
void launch_group_bbox(int dim, const double* x, const int64_t* gstart, int32_t ngroups, double* out, hipStream_t st);
void launch_seg_sum(const double* vals, const int64_t* seg, int32_t nseg, const double* scale, double* out,
                    hipStream_t st);
void launch_group_dt(int dim, const Consts& c, DevState s, const int32_t* part_id, const int64_t* gstart,
                     int32_t ngroups, double* out, hipStream_t st);
void launch_scan(const int32_t* count, int64_t ncell, int32_t* start, int32_t* block_sums, hipStream_t st);
void launch_scatter(const int32_t* cell, int64_t n, const int32_t* start, int32_t* fill, int32_t* order,
                    hipStream_t st);
void launch_sort_cells(const int32_t* start, int64_t ncell, int32_t* order, hipStream_t st);
void launch_build_lists(int dim, const double* x, int64_t n, TileGrid tg, double sr, NbrList nl, DevScalars* sc,
                        hipStream_t st);
void launch_export_lists(int dim, TileGrid tg, NbrList nl, int64_t n, int32_t* out, hipStream_t st);
void launch_aux(int dim, DevState s, Aux a, int64_t n, int ale, int mixed, double bgam, hipStream_t st);
void launch_gather_d(const double* src, double* dst, const int32_t* perm, int64_t n, int width, hipStream_t st);
void launch_gather_i(const int32_t* src, int32_t* dst, const int32_t* perm, int64_t n, hipStream_t st);
void launch_iota(int32_t* a, int64_t n, hipStream_t st);

/* Pair passes.  r_from_x = 1 when the positions are those of the last
 * neighbour build (r is then formed from x); otherwise from Aux::xb. */
void launch_dsph(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                 int64_t n, int r_from_x, int with_aero_vel, double* partial, DevScalars* sc, hipStream_t st);
void launch_aero_velocity(int dim, const Consts& c, DevState s, const int32_t* part_id, NbrList nl, int64_t n,
                          hipStream_t st);
/* mode 0: Detect_Surface loop 1; 1: + dissipation_terms; 2: dissipation only */
void launch_surface1(int dim, int prec, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg,
                     NbrList nl, int64_t n, int r_from_x, double* norms_tmp, int mode, hipStream_t st);
void launch_surface2(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                     int64_t n, int r_from_x, const double* norms_tmp, hipStream_t st);
void launch_shift(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                  int64_t n, int r_from_x, hipStream_t st);
/* Detect_Surface loops 2-3 and particle_shift in one pass (mixed precision, ALE) */
void launch_surface2_shift(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg,
                           NbrList nl, int64_t n, int r_from_x, const double* norms_tmp, hipStream_t st);
void launch_boundary(int dim, const Consts& c, DevState s, Aux a, const int32_t* blk, DevBlocks bl, int32_t block,
                     TileGrid tg, NbrList nl, int64_t n, int64_t nb, int32_t* slots, int32_t* near_inlet,
                     hipStream_t st);
/* Force pass on `s`; when next_x is non-null the Newmark update of
 * Do_NB_Iter is fused and writes next x/v/rho/p for every slot, plus one
 * residual partial sum per tile block. */
void launch_force(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                  int r_from_x, const DevScalars* sc, hipStream_t st);
/* next_x == nullptr: only the acceleration tail, no Newmark update. */
void launch_nb_update(int dim, const Consts& c, DevState s, DevState pn, double* next_x, double* next_v,
                      double* next_rho, double* next_p, const int32_t* part_id, const int32_t* blk, DevBlocks bl,
                      const int32_t* near_inlet, NbrList nl, int64_t n, double dt, double* partial, hipStream_t st,
                      DevGroups gr = DevGroups{});
/* Surrogate on the aero particles of s (mode 0) or on points (mode 1). */
void launch_surrogate(int dim, const Consts& c, DevState s, const int32_t* part_id, Surrogate sg, int64_t n, int mode,
                      const double* pts, double* out, hipStream_t st);
void launch_aero_list(int dim, const Consts& c, DevState s, const int32_t* part_id, int64_t n, int32_t* list,
                      int32_t* count, hipStream_t st);
void launch_surrogate_list(int dim, DevState s, Surrogate sg, int64_t n, const int32_t* list, const int32_t* count,
                           hipStream_t st);
void launch_pipe_outlet(int dim, DevState s, const int32_t* part_id, const int32_t* blk, DevPlanes pl, int64_t n,
                        int64_t n_bound, hipStream_t st, int32_t* fresh = nullptr, double lam_cutoff = 0.0,
                        DevScalars* sc = nullptr);
/* FirstCell (src/Containment.cpp:423-573) for the particles leaving the pipe
 * that the lookup placed in no cell: the nearest cells again, then the
 * boundary branch (reflection at walls, deletion mark at the far field). */
void launch_first_cell(int dim, DevState s, DevMesh m, const int32_t* fresh, int32_t nfresh, int32_t* fc_del,
                       DevScalars* sc, hipStream_t st);
void launch_inlet_backslot(const int32_t* col, const int32_t* depth, int64_t n, int32_t* back_slot, hipStream_t st);
void launch_rk_buffer(int dim, const Consts& c, DevState s, DevState pn, DevState s1, DevState s2, DevState s3,
                      const int32_t* col, const int32_t* depth, const int32_t* blk, const int32_t* back_slot,
                      DevPlanes pl, int64_t n, double dt, double dx, int final_stage, hipStream_t st);
void launch_inlet_buffer(int dim, const Consts& c, DevState s, DevState pn, NextState nx, const int32_t* col,
                         const int32_t* depth, const int32_t* blk, const int32_t* back_slot, DevPlanes pl, int64_t n,
                         double dt, double dx, double* partial, hipStream_t st);
void launch_inlet_back(int dim, DevState s, int32_t* col, int32_t* depth, const int32_t* blk, DevPlanes pl, int64_t n,
                       int32_t* adv, int32_t* last_slot, hipStream_t st);
void launch_inlet_shift(DevState s, int32_t* col, int32_t* depth, const int32_t* blk, DevPlanes pl, int64_t n,
                        const int32_t* adv, hipStream_t st);
void launch_inlet_insert(int dim, DevState s, int32_t* blk, int32_t* col, int32_t* depth, const int32_t* col_block,
                         const int32_t* rank, const int32_t* last_slot, DevPlanes pl, int32_t ncols, int64_t n_old,
                         double dx, int64_t id_base, hipStream_t st);
void launch_shift_ids(int32_t* part_id, int64_t n, int32_t from, int32_t delta, hipStream_t st);
void launch_new_ids(int32_t* part_id, const int32_t* col_block, const int32_t* rank, int32_t ncols, int32_t k,
                    int64_t n_old, int32_t rank_base, int32_t first_new, hipStream_t st);
void launch_delete_flag(int dim, DevState s, const int32_t* part_id, const int32_t* blk, DevPlanes pl, int64_t n,
                        int64_t n_bound, int32_t* keep, int32_t* hflag, int32_t* blkcount, hipStream_t st);
void launch_compact_perm(const int32_t* keep, const int32_t* start, int64_t n, int32_t* perm, hipStream_t st);
void launch_compact_ids(int32_t* part_id, const int32_t* hstart, int64_t n, hipStream_t st);
void launch_mlp_batch(uint64_t seed, int64_t step, int64_t n, int32_t batch, int32_t* idx, hipStream_t st);
void launch_mlp_train(int dim, const MlpTrain& m, hipStream_t st);
void launch_mlp_reduce(const float* part, int64_t nwaves, int64_t stride, int64_t n, float* grad, hipStream_t st);
void launch_mlp_adam(float* p, const float* g, float* m1, float* m2, int64_t n, float lr, float b1, float b2,
                     float eps, float c1, float c2, hipStream_t st);
int64_t mlp_tiles_fwd(int n_hidden);
int64_t mlp_tiles_bwd(int n_hidden);
void launch_mlp_pack(int dim, const float* w, int32_t n_hidden, void* wf, void* wt, hipStream_t st);
/* Scratch of the entry points that allocate nothing (noalloc.hip): a
 * file-scope device array per GPU, its address taken once per device by
 * noalloc_init (with that device current) and returned by noalloc_scratch. */
constexpr size_t kNoallocScratchBytes = FJSPH_NOALLOC_SCRATCH_BYTES;
hipError_t noalloc_init(int device);
void* noalloc_scratch(int device);
void launch_mlp_stage(const float* xn, const float* tn, int32_t batch, int32_t dim, float* X, float* T, int32_t* idx,
                      hipStream_t st);
void launch_mlp_loss(const float* sse, double denom, double* loss, hipStream_t st);
/* The mesh field at points: containing cell (-1 if none) and (vel, p, rho). */
void launch_mesh_eval(int dim, DevMesh m, int64_t n, const double* pts, int32_t* cell, double* vals, hipStream_t st);
void launch_ipt_track(int dim, const DevMesh& m, const IptParams& p, int64_t n, const IptStart* in,
                      fjsph_ipt_record* out, hipStream_t st);
void launch_ipt_gather(int dim, DevState s, const int32_t* slots, int64_t n, double t, IptStart* out, hipStream_t st);
/* FindCell on the aero particles of s (mesh.hip). */
/* FindCell over the aero particles: the previous cell, then the voxel
 * candidates.  Particles in no cell are counted in sc->mesh_lost and, when
 * lost_list is given (room for n), listed there. */
void launch_find_cell(int dim, const Consts& c, DevState s, const int32_t* part_id, DevMesh m, int64_t n,
                      DevScalars* sc, int32_t* lost_list, hipStream_t st);
/* FindCell's boundary test for the nlost listed particles, one block each:
 * lost_del[k] = 1 when particle k is to be deleted; sc->lost_* counted. */
void launch_lost_classify(int dim, DevState s, DevMesh m, const int32_t* lost_list, int32_t nlost, int32_t* lost_del,
                          DevScalars* sc, hipStream_t st);
/* keep (per slot) and hflag (per host id) for a deletion of the listed
 * particles with lost_del set; blkcount[k] counts them per block. */
void launch_lost_marks(const int32_t* lost_list, const int32_t* lost_del, int32_t nlost, const int32_t* part_id,
                       const int32_t* blk, int64_t n, int32_t* keep, int32_t* hflag, int32_t* blkcount,
                       hipStream_t st);
/* Runge-Kutta 4 (pairs.hip): stage update of s from pn (final_stage: the
 * weighted combination with stages s1..s3), boundary densities, and the
 * squared position change partials of Check_RK_Error. */
void launch_rk_update(int dim, const Consts& c, DevState s, DevState pn, DevState s1, DevState s2, DevState s3,
                      const int32_t* part_id, int64_t n, double dt, int final_stage, hipStream_t st);
void launch_rk_bnd_rho(int dim, const Consts& c, DevState s, DevState pn, DevState s1, DevState s2, DevState s3,
                       const int32_t* blk, DevBlocks bl, const int32_t* near_inlet, int64_t n, double dt,
                       int final_stage, hipStream_t st);
void launch_sqdiff(int dim, const double* xa, const double* xb, const int32_t* part_id, int64_t n_bound, int64_t n,
                   double* partial, hipStream_t st);
void launch_sum_partials(const double* partial, int nparts, int width, double* out, hipStream_t st);
void launch_sum_partials_scaled(const double* partial, int nparts, int width, double scale, double* out,
                                hipStream_t st);
void launch_timestep_reduce(int dim, const Consts& c, DevState s, const int32_t* part_id, int64_t n,
                            double* partial, int nparts, double* out, hipStream_t st);
void launch_fix_rho(DevState s, const int32_t* part_id, int64_t nb, int64_t n, double rho0, hipStream_t st);

/* Number of blocks used for per-block partial reductions. */
int reduction_parts(int64_t n);

}  // namespace fj

#endif
