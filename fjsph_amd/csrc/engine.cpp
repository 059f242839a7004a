/*
 * C-ABI of the WCSPH step engine (include/fjsph_engine.h): device memory,
 * host<->device transfer, neighbour build with cell-order reordering, and
 * the host control flow of Integrator / Newmark_Beta.
 *
 * Host control mirrors src/Integration.cpp:27-303 and
 * src/Newmark_Beta.cpp:10-331 line for line in its decisions; every
 * particle operation runs in kernels.hip.  The host reads back one scalar
 * per Newmark sub-iteration (the residual), the bounding box and the list
 * length per neighbour build, and the find_timestep reduction per step.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "engine_internal.hpp"

using namespace fj;

namespace {

constexpr int kNumFields = FJSPH_F_COUNT;

struct FieldDesc {
    int width_kind; /* 0 scalar, 1 vector (dim), 2 matrix (dim*dim) */
    bool is_int;
};

FieldDesc field_desc(int f)
{
    switch (f) {
    case FJSPH_F_X: case FJSPH_F_V: case FJSPH_F_ACC: case FJSPH_F_AF: case FJSPH_F_AVISC:
    case FJSPH_F_GRADRHO: case FJSPH_F_NORM: case FJSPH_F_CELLV: case FJSPH_F_VPERT:
        return {1, false};
    case FJSPH_F_L:
        return {2, false};
    case FJSPH_F_SURF: case FJSPH_F_CELLID: case FJSPH_F_SURFZONE: case FJSPH_F_B: case FJSPH_F_PID:
    case FJSPH_F_NFAILED: case FJSPH_F_INTERNAL:
        return {0, true};
    default:
        return {0, false};
    }
}

void** field_slot(DevState& s, int f)
{
    switch (f) {
    case FJSPH_F_X: return (void**)&s.x;
    case FJSPH_F_V: return (void**)&s.v;
    case FJSPH_F_ACC: return (void**)&s.acc;
    case FJSPH_F_AF: return (void**)&s.Af;
    case FJSPH_F_AVISC: return (void**)&s.aVisc;
    case FJSPH_F_RHO: return (void**)&s.rho;
    case FJSPH_F_P: return (void**)&s.p;
    case FJSPH_F_RRHO: return (void**)&s.Rrho;
    case FJSPH_F_M: return (void**)&s.m;
    case FJSPH_F_DELTAD: return (void**)&s.deltaD;
    case FJSPH_F_L: return (void**)&s.L;
    case FJSPH_F_GRADRHO: return (void**)&s.gradRho;
    case FJSPH_F_NORM: return (void**)&s.norm;
    case FJSPH_F_LAM: return (void**)&s.lam;
    case FJSPH_F_LAM_NB: return (void**)&s.lam_nb;
    case FJSPH_F_COLOURG: return (void**)&s.colourG;
    case FJSPH_F_COLOUR: return (void**)&s.colour;
    case FJSPH_F_KERNSUM: return (void**)&s.kernsum;
    case FJSPH_F_SURF: return (void**)&s.surf;
    case FJSPH_F_CURVE: return (void**)&s.curve;
    case FJSPH_F_NORM_CURVE: return (void**)&s.norm_curve;
    case FJSPH_F_WOCCL: return (void**)&s.woccl;
    case FJSPH_F_PDIST: return (void**)&s.pDist;
    case FJSPH_F_CELLV: return (void**)&s.cellV;
    case FJSPH_F_CELLP: return (void**)&s.cellP;
    case FJSPH_F_CELLRHO: return (void**)&s.cellRho;
    case FJSPH_F_CELLID: return (void**)&s.cellID;
    case FJSPH_F_VPERT: return (void**)&s.vPert;
    case FJSPH_F_SURFZONE: return (void**)&s.surfzone;
    case FJSPH_F_B: return (void**)&s.b;
    case FJSPH_F_PID: return (void**)&s.pid;
    case FJSPH_F_NFAILED: return (void**)&s.nfail;
    case FJSPH_F_INTERNAL: return (void**)&s.internal;
    default: return nullptr;
    }
}

struct HipError {
    hipError_t e;
    const char* what;
};

struct HostBlock {
    fjsph_block b;
    std::vector<double> times, vels;
};

struct TimerRec {
    int pass;
    hipEvent_t a, b;
};

}  // namespace

struct fjsph_ctx {
    fjsph_settings s{};
    fjsph_derived d{};
    Consts c{};
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;

    int64_t n = 0, nb = 0;
    bool have_state = false, have_nbrs = false;

    DevState pn{}, pnp1{};
    DevState rk1{}, rk2{}, rk3{}; /* Runge-Kutta stage states, allocated on first use */
    bool rk_alloc = false;

    /* aerodynamic mesh: host copies of the cell data, device arrays */
    bool have_mesh = false;
    int64_t mesh_ncells = 0;
    std::vector<double> mcentre, mvel, mp, mrho;
    double mesh_max_len = 0.0; /* MESH::maxlength */
    int32_t block_targets = 0; /* targets per tile block (0: 256) */
    int32_t reorder_every = 4, steps_to_reorder = 0; /* cell-order permutation interval, steps left */
    /* implicit particle tracking */
    fjsph_ipt_settings ipt{};
    std::vector<fjsph_ipt_record> ipt_records;
    int64_t ipt_success = 0, ipt_failed = 0;
    IptStart* ipt_in = nullptr;
    fjsph_ipt_record* ipt_out = nullptr;
    int32_t* ipt_slots = nullptr;
    int64_t ipt_cap = 0;
    DevMesh dm{};
    std::vector<void*> mesh_bufs;
    /* particles a mesh lookup placed in no cell */
    int32_t *lost_list = nullptr, *lost_del = nullptr;
    int64_t lost_cap = 0;
    int64_t lost_deleted = 0, lost_internal = 0, lost_failed = 0;
    /* particles leaving the pipe: FirstCell's list, deletion marks, counters */
    int32_t *fresh = nullptr, *fc_del = nullptr;
    int64_t fresh_cap = 0, first_cell = 0, first_cell_deleted = 0;
    bool tuning = false; /* the tile-block autotuner is timing passes: no deletions */
    /* surrogate */
    int32_t* aero_list = nullptr;   /* segmented list of the aero particles of the last lookup */
    int32_t* aero_counts = nullptr; /* its per-segment lengths */
    int64_t aero_list_cap = 0;      /* segments allocated */
    bool have_surrogate = false;
    Surrogate sg{};
    std::vector<void*> sg_bufs;
    double *x_nxt = nullptr, *v_nxt = nullptr, *rho_nxt = nullptr, *p_nxt = nullptr;
    int32_t *part_id = nullptr, *blk = nullptr, *near_inlet = nullptr;
    double* norms_tmp = nullptr;
    double* partial = nullptr;
    int nparts = 0;
    double* gather_tmp[3] = {nullptr, nullptr, nullptr}; /* spare field buffers: scalar, vector, matrix */
    int32_t* gather_tmp_i = nullptr;

    /* cell grid */
    int32_t *cell = nullptr, *cell_count = nullptr, *cell_start = nullptr, *cell_fill = nullptr;
    int32_t *order = nullptr, *vorder = nullptr, *pos = nullptr, *tile_sums = nullptr;
    int64_t ncell_cap = 0, ncell = 0;
    int32_t dims[3] = {1, 1, 1};

    NbrList nl{};
    Aux aux{};
    TileGrid tg{};
    int64_t partial_cap = 0;
    int32_t* export_buf = nullptr;
    bool x_at_build = false;   /* pnp1.x unchanged since the last build     */
    bool f64_tail = false;     /* this step's remaining force passes in FP64 */
    int32_t* bnd_slots = nullptr; /* one boundary block's slots (count first), per pass */
    DevScalars* sc = nullptr;
    DevScalars* sc_host = nullptr; /* pinned */
    DevBlocks bl{};
    std::vector<HostBlock> blocks;
    std::vector<int32_t> slot_host; /* host copy of part_id when needed */

    /* integrator state (INTEG_SETT + Integrator members) */
    double current_time = 0.0, last_frame_time = 0.0, delta_t = 2e-10, cfl = 1.0, safe_dt = 0.0;
    unsigned n_stable = 0, n_unstable = 0, iteration = 0;
    int64_t current_frame = 0, next_part_id = 0;
    int rollbacks = 0;
    double maxf = 0, maxAf = 0, maxRho_pc = 0, maxRhoi = 0, maxdrho = 0, minST = 0, maxU = 0, maxShift = 0;
    double last_npd = 1.0;

    /* independent simulations in one context (fjsph_set_groups; 0: one).
     * Slots are group-major: group g holds slots [gstart_h[g], gstart_h[g+1]).
     * Each group keeps the integrator state the reference keeps per
     * simulation (src/Integration.cpp:233-303, src/Newmark_Beta.cpp:10-52). */
    struct Group {
        double time = 0.0, last_frame_time = 0.0, dt = 2e-10, cfl = 1.0, safe_dt = 0.0;
        unsigned n_stable = 0, n_unstable = 0, iteration = 0;
        int rollbacks = 0;
        bool active = true;
        bool f64_tail = false; /* this step's remaining force passes in FP64 */
        double xmax = 0.0;     /* max |coordinate| at the last build */
        double logbase = 0.0, rms = 0.0, npd = 1.0;
        double maxf = 0, maxAf = 0, maxRho_pc = 0, maxRhoi = 0, maxdrho = 0, minST = 0, maxU = 0, maxShift = 0;
    };
    int32_t ngroups = 0;
    std::vector<Group> grs;
    std::vector<int64_t> gstart_h; /* slot ranges, ngroups + 1 */
    std::vector<int64_t> gblk_h;   /* tile-block ranges of the last build, ngroups + 1 */
    std::vector<int32_t> gl0_h;    /* first cell layer and layer count, 2 per group */
    int32_t* grp = nullptr;        /* group of every slot */
    int64_t *gstart = nullptr, *gblk = nullptr;
    double *gbox = nullptr, *glo = nullptr, *gdt = nullptr, *gnpd = nullptr, *gscale = nullptr;
    double *gred = nullptr, *gres = nullptr, *gd2 = nullptr;
    int32_t *gl0 = nullptr, *gactive = nullptr, *gbuild = nullptr, *lgrp = nullptr;
    /* force-pass masks of the groups iterating in FP32 pair terms and of
     * those whose step has gone to FP64 ones */
    int32_t *gact_m = nullptr, *gact_d = nullptr;
    int64_t lgrp_cap = 0, gd2_cap = 0;

    /* timers */
    bool timing = false;
    std::vector<TimerRec> pending;
    std::vector<hipEvent_t> event_pool;
    double t_ms[FJSPH_T_COUNT] = {0};
    int64_t t_launch[FJSPH_T_COUNT] = {0};

    int64_t pairs = 0;
    int32_t max_count = 0;

    /* inlets, deletion and the pipe exit; per-slot arrays hold `cap` slots */
    int64_t cap = 0;
    int64_t max_points = 9999999; /* svar.max_points (src/Var.h:369) */
    std::vector<fjsph_block_planes> planes; /* per block */
    struct InletCols {
        int32_t block, col_base, ncols, nbuf;
    };
    std::vector<InletCols> inlets;
    std::vector<int32_t> starved; /* per column: advances since its last new particle */
    int32_t ncols = 0;
    bool planes_on = false, planes_dirty = true;
    int32_t *col = nullptr, *depth = nullptr;                       /* per slot */
    int32_t *c_adv = nullptr, *c_back = nullptr, *c_last = nullptr; /* per column */
    int32_t *c_rank = nullptr, *c_block = nullptr;
    double *pl_del = nullptr, *pl_ins = nullptr, *pl_aero = nullptr;
    int32_t *pl_mode = nullptr, *pl_nbuf = nullptr;
    DevPlanes dpl{};
    int32_t *keep = nullptr, *hflag = nullptr, *kstart = nullptr, *hstart = nullptr; /* deletion scratch */
    int32_t *perm = nullptr, *del_tiles = nullptr, *blkcount = nullptr;
    size_t blkcount_cap = 0;
};

namespace {

void set_err(fjsph_ctx* ctx, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
}

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) throw HipError{e_, #call};                                    \
    } while (0)

template <typename T>
void dalloc(T*& p, size_t count)
{
    if (p) { (void)hipFree(p); p = nullptr; }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        throw HipError{e, "hipMalloc"};
    }
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int width_of(const fjsph_ctx* ctx, int f)
{
    const FieldDesc fd = field_desc(f);
    return fd.width_kind == 0 ? 1 : (fd.width_kind == 1 ? ctx->s.dim : ctx->s.dim * ctx->s.dim);
}

void alloc_state(fjsph_ctx* ctx, DevState& s)
{
    for (int f = 0; f < kNumFields; ++f) {
        void** slot = field_slot(s, f);
        const size_t cnt = size_t(std::max(ctx->cap, ctx->n)) * width_of(ctx, f);
        if (field_desc(f).is_int) dalloc(*(int32_t**)slot, cnt);
        else dalloc(*(double**)slot, cnt);
        HIPCHK(hipMemsetAsync(*slot, 0, cnt * (field_desc(f).is_int ? 4 : 8), ctx->st));
    }
}

void free_state(DevState& s)
{
    for (int f = 0; f < kNumFields; ++f) {
        void** slot = field_slot(s, f);
        if (*slot) (void)hipFree(*slot);
        *slot = nullptr;
    }
}

void free_all(fjsph_ctx* ctx)
{
    free_state(ctx->pn);
    free_state(ctx->pnp1);
    free_state(ctx->rk1);
    free_state(ctx->rk2);
    free_state(ctx->rk3);
    ctx->rk_alloc = false;
    for (void* p : ctx->mesh_bufs) (void)hipFree(p);
    ctx->mesh_bufs.clear();
    ctx->have_mesh = false;
    dfree(ctx->ipt_in); dfree(ctx->ipt_out); dfree(ctx->ipt_slots);
    ctx->ipt_cap = 0;
    dfree(ctx->lost_list); dfree(ctx->lost_del);
    ctx->lost_cap = 0;
    dfree(ctx->fresh); dfree(ctx->fc_del);
    ctx->fresh_cap = 0;
    ctx->first_cell = ctx->first_cell_deleted = 0;
    ctx->lost_deleted = ctx->lost_internal = ctx->lost_failed = 0;
    for (void* p : ctx->sg_bufs) (void)hipFree(p);
    ctx->sg_bufs.clear();
    ctx->have_surrogate = false;
    dfree(ctx->x_nxt); dfree(ctx->v_nxt); dfree(ctx->rho_nxt); dfree(ctx->p_nxt);
    dfree(ctx->part_id); dfree(ctx->blk); dfree(ctx->near_inlet); dfree(ctx->norms_tmp);
    dfree(ctx->partial); dfree(ctx->gather_tmp[0]); dfree(ctx->gather_tmp[1]); dfree(ctx->gather_tmp[2]);
    dfree(ctx->gather_tmp_i);
    dfree(ctx->cell); dfree(ctx->cell_count); dfree(ctx->cell_start); dfree(ctx->cell_fill);
    dfree(ctx->order); dfree(ctx->vorder); dfree(ctx->pos); dfree(ctx->tile_sums);
    dfree(ctx->nl.lu4); dfree(ctx->nl.count); dfree(ctx->nl.nnb); dfree(ctx->export_buf);
    dfree(ctx->aux.vol); dfree(ctx->aux.P); dfree(ctx->aux.invrho); dfree(ctx->aux.fq); dfree(ctx->aux.fale);
    dfree(ctx->aux.fs1); dfree(ctx->aux.fs2);
    dfree(ctx->aux.nhat); dfree(ctx->aux.lnb); dfree(ctx->aux.xb); dfree(ctx->aux.spend); dfree(ctx->aux.sund); dfree(ctx->bnd_slots); dfree(ctx->aero_list);
    dfree(ctx->aero_counts);
    ctx->aero_list_cap = 0;
    dfree(ctx->bl.solver); dfree(ctx->bl.no_slip); dfree(ctx->bl.vel);
    dfree(ctx->col); dfree(ctx->depth); dfree(ctx->keep); dfree(ctx->hflag); dfree(ctx->kstart);
    dfree(ctx->hstart); dfree(ctx->perm); dfree(ctx->del_tiles); dfree(ctx->blkcount);
    ctx->blkcount_cap = 0;
    dfree(ctx->c_adv); dfree(ctx->c_back); dfree(ctx->c_last); dfree(ctx->c_rank); dfree(ctx->c_block);
    dfree(ctx->pl_del); dfree(ctx->pl_ins); dfree(ctx->pl_aero); dfree(ctx->pl_mode); dfree(ctx->pl_nbuf);
    dfree(ctx->grp); dfree(ctx->gstart); dfree(ctx->gblk); dfree(ctx->gbox); dfree(ctx->glo); dfree(ctx->gdt);
    dfree(ctx->gnpd); dfree(ctx->gscale); dfree(ctx->gred); dfree(ctx->gres); dfree(ctx->gd2); dfree(ctx->gl0);
    dfree(ctx->gactive); dfree(ctx->gbuild); dfree(ctx->lgrp);
    dfree(ctx->gact_m); dfree(ctx->gact_d);
    ctx->lgrp_cap = ctx->gd2_cap = 0;
    ctx->ngroups = 0;
    ctx->grs.clear();
    ctx->inlets.clear();
    ctx->starved.clear();
    ctx->ncols = 0;
    ctx->planes_dirty = true;
    ctx->ncell_cap = 0;
    ctx->nl.cap = 0;
    ctx->cap = 0;
}

/* New allocation of `count` elements keeping the first `keep` of the old. */
template <typename T>
void dgrow(T*& p, size_t count, size_t keep)
{
    T* q = nullptr;
    hipError_t e = hipMalloc((void**)&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) throw HipError{e, "hipMalloc"};
    if (p && keep) {
        e = hipMemcpy(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice);
        if (e != hipSuccess) throw HipError{e, "hipMemcpy"};
    }
    if (p) (void)hipFree(p);
    p = q;
}

/* (Re)allocate every per-slot array for `cap` slots, keeping the first
 * `keep` entries (the particles so far); new slots of the states are zero,
 * of col/depth -1.  The neighbour lists are dropped (rebuilt at the next
 * build). */
void alloc_slots(fjsph_ctx* ctx, int64_t cap, int64_t keep)
{
    const size_t D = size_t(ctx->s.dim), c = size_t(cap), k = size_t(keep);
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (DevState* st : {&ctx->pn, &ctx->pnp1}) {
        for (int f = 0; f < kNumFields; ++f) {
            void** slot = field_slot(*st, f);
            const size_t w = size_t(width_of(ctx, f)), es = field_desc(f).is_int ? 4 : 8;
            char* q = (char*)*slot;
            dgrow(q, c * w * es, k * w * es);
            HIPCHK(hipMemset(q + k * w * es, 0, (c - k) * w * es));
            *slot = q;
        }
    }
    if (ctx->rk_alloc) {
        free_state(ctx->rk1);
        free_state(ctx->rk2);
        free_state(ctx->rk3);
        ctx->rk_alloc = false;
    }
    dgrow(ctx->x_nxt, c * D, k * D);
    dgrow(ctx->v_nxt, c * D, k * D);
    dgrow(ctx->rho_nxt, c, k);
    dgrow(ctx->p_nxt, c, k);
    dgrow(ctx->part_id, c, k);
    dgrow(ctx->blk, c, k);
    dgrow(ctx->near_inlet, c, k);
    HIPCHK(hipMemset(ctx->near_inlet + k, 0, (c - k) * 4));
    dgrow(ctx->col, c, k);
    dgrow(ctx->depth, c, k);
    HIPCHK(hipMemset(ctx->col + k, 0xFF, (c - k) * 4));
    HIPCHK(hipMemset(ctx->depth + k, 0xFF, (c - k) * 4));
    dgrow(ctx->norms_tmp, c * D, 0);
    dgrow(ctx->gather_tmp[0], c, 0);
    dgrow(ctx->gather_tmp[1], c * D, 0);
    dgrow(ctx->gather_tmp[2], c * D * D, 0);
    dgrow(ctx->gather_tmp_i, c, 0);
    if (ctx->grp) dgrow(ctx->grp, c, k);
    if (ctx->gd2) {
        dgrow(ctx->gd2, c, 0);
        ctx->gd2_cap = int64_t(c);
    }
    dgrow(ctx->cell, c, k);
    dgrow(ctx->order, c, k);
    dgrow(ctx->pos, c, k);
    dgrow(ctx->vorder, c, k);
    dgrow(ctx->aux.vol, c, 0);
    dgrow(ctx->aux.P, c, 0);
    dgrow(ctx->aux.invrho, c, 0);
    dgrow(ctx->aux.fq, c, 0);
    dgrow(ctx->aux.fale, c, 0);
    dgrow(ctx->aux.fs1, c, 0);
    dgrow(ctx->aux.fs2, c, 0);
    HIPCHK(hipMemset(ctx->aux.fs1, 0, c * sizeof(float4)));
    HIPCHK(hipMemset(ctx->aux.fs2, 0, c * sizeof(float4)));
    dgrow(ctx->aux.nhat, c * D, k * D);
    HIPCHK(hipMemset(ctx->aux.nhat + k * D, 0, (c - k) * D * 8));
    dgrow(ctx->aux.lnb, c * D * D, 0);
    dgrow(ctx->aux.xb, c * D, 0);
    dgrow(ctx->aux.spend, c, 0);
    dgrow(ctx->aux.sund, c + 1, 0);
    dgrow(ctx->bnd_slots, c + 1, 0);
    dgrow(ctx->keep, c, 0);
    dgrow(ctx->hflag, c, 0);
    dgrow(ctx->kstart, c + 1, 0);
    dgrow(ctx->hstart, c + 1, 0);
    dgrow(ctx->perm, c, 0);
    dgrow(ctx->del_tiles, c / 1024 + 2, 0);
    ctx->nl.npad = (cap + 63) / 64 * 64;
    dgrow(ctx->nl.count, size_t(ctx->nl.npad), 0);
    dgrow(ctx->nl.nnb, size_t(ctx->nl.npad), 0);
    dfree(ctx->nl.lu4);
    ctx->nl.cap = 0;
    const int64_t parts = int64_t(reduction_parts(cap)) * 8;
    if (parts > ctx->partial_cap) {
        dalloc(ctx->partial, size_t(parts));
        ctx->partial_cap = parts;
    }
    ctx->cap = cap;
    ctx->have_nbrs = false;
}

void fill_consts(fjsph_ctx* ctx)
{
    const fjsph_settings& s = ctx->s;
    const fjsph_derived& d = ctx->d;
    Consts& c = ctx->c;
    c.H = d.H; c.H_sq = d.H_sq; c.sr = d.sr; c.Wc = d.W_correc; c.W_dx = d.W_dx; c.B = d.B; c.gam = s.gam;
    c.rho0 = s.rho_rest; c.c0_sq = s.speed_sound * s.speed_sound; c.press_back = s.press_back; c.nu = d.nu;
    c.visc_alpha = s.visc_alpha; c.dsph_cont = d.dsph_cont; c.rho_min = d.rho_min; c.rho_max = d.rho_max;
    c.dx = d.dx; c.particle_step = s.particle_step; c.lam_cutoff = s.lam_cutoff; c.interp_fac = d.interp_fac;
    c.i_n_full = d.i_n_full; c.A_plate = d.A_plate; c.A_sphere = d.A_sphere; c.L = d.L; c.sos = d.sos;
    c.gamma_g = s.gamma_g; c.mu_g = s.mu_g; c.rho_g = s.rho_g; c.ycoef = d.ycoef; c.Cb = s.Cb;
    c.max_shift_vel = s.max_shift_vel; c.lam_st = d.lam_st; c.sig = s.sig;
    for (int k = 0; k < 3; ++k) { c.grav[k] = s.grav[k]; c.v_inf[k] = s.v_inf[k]; }
    c.eos = s.eos; c.aero_case = s.aero_case; c.aero_source = s.aero_source; c.use_lam = s.use_lam;
    c.use_TAB_def = s.use_TAB_def; c.ale = s.ale;
    c.n_bound = ctx->nb;
}

/* ---------------------------------------------------------------- timers */

hipEvent_t take_event(fjsph_ctx* ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
}

struct PassTimer {
    fjsph_ctx* ctx;
    int pass;
    hipEvent_t a = nullptr;
    PassTimer(fjsph_ctx* c, int p) : ctx(c), pass(p)
    {
        if (ctx->timing) {
            a = take_event(ctx);
            HIPCHK(hipEventRecord(a, ctx->st));
        }
    }
    ~PassTimer() noexcept(false)
    {
        if (ctx->timing && a) {
            hipEvent_t b = take_event(ctx);
            HIPCHK(hipEventRecord(b, ctx->st));
            ctx->pending.push_back({pass, a, b});
        }
    }
};

void drain_timers(fjsph_ctx* ctx)
{
    if (ctx->pending.empty()) return;
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (auto& r : ctx->pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        ctx->t_ms[r.pass] += ms;
        ctx->t_launch[r.pass] += 1;
        ctx->event_pool.push_back(r.a);
        ctx->event_pool.push_back(r.b);
    }
    ctx->pending.clear();
}

/* ----------------------------------------------------------- state moves */

void sync(fjsph_ctx* ctx) { HIPCHK(hipStreamSynchronize(ctx->st)); }

void copy_state(fjsph_ctx* ctx, DevState& dst, DevState& src)
{
    for (int f = 0; f < kNumFields; ++f) {
        void* d = *field_slot(dst, f);
        void* s = *field_slot(src, f);
        const size_t bytes = size_t(ctx->n) * width_of(ctx, f) * (field_desc(f).is_int ? 4 : 8);
        HIPCHK(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, ctx->st));
    }
}

/* Permute every per-slot array by perm (new slot s <- old slot perm[s]),
 * for the first `count` new slots (ctx->n when negative). */
void permute_all(fjsph_ctx* ctx, const int32_t* perm, int64_t count = -1)
{
    const int64_t n = count < 0 ? ctx->n : count;
    for (DevState* st : {&ctx->pn, &ctx->pnp1}) {
        for (int f = 0; f < kNumFields; ++f) {
            void** slot = field_slot(*st, f);
            if (field_desc(f).is_int) {
                launch_gather_i((int32_t*)*slot, ctx->gather_tmp_i, perm, n, ctx->st);
                std::swap(*(int32_t**)slot, ctx->gather_tmp_i);
            } else {
                const int w = width_of(ctx, f);
                /* the spare buffer of the field's width (same capacity)
                 * takes the gathered values and becomes the field */
                double*& spare = ctx->gather_tmp[field_desc(f).width_kind];
                launch_gather_d((double*)*slot, spare, perm, n, w, ctx->st);
                std::swap(*(double**)slot, spare);
            }
        }
    }
    for (int32_t** arr : {&ctx->part_id, &ctx->blk, &ctx->cell, &ctx->col, &ctx->depth, &ctx->grp}) {
        if (!*arr) continue;
        launch_gather_i(*arr, ctx->gather_tmp_i, perm, n, ctx->st);
        std::swap(*arr, ctx->gather_tmp_i);
    }
}

/* copy_state for the slots [first, first + count) only */
void copy_state_range(fjsph_ctx* ctx, DevState& dst, DevState& src, int64_t first, int64_t count)
{
    for (int f = 0; f < kNumFields; ++f) {
        const int w = width_of(ctx, f);
        const size_t es = field_desc(f).is_int ? 4 : 8;
        char* d = (char*)*field_slot(dst, f) + size_t(first) * w * es;
        const char* s = (const char*)*field_slot(src, f) + size_t(first) * w * es;
        HIPCHK(hipMemcpyAsync(d, s, size_t(count) * w * es, hipMemcpyDeviceToDevice, ctx->st));
    }
}

bool grouped(const fjsph_ctx* ctx) { return ctx->ngroups > 0; }

/* update_neighbours (src/Neighbours.cpp:7-31) on pnp1.  When `reorder` is
 * set the particles are first moved into cell order (a permutation only;
 * the neighbour sets and every result are unchanged). */
/* Cell layout of independent simulations: each group is binned from its
 * own lower corner, and the groups are stacked along the last axis in
 * group order with one empty cell layer between consecutive groups, so
 * the cell order is group-major, no candidate union crosses a group and
 * every tile block lies in one group.  Sets dims; returns the cell count. */
double group_layout(fjsph_ctx* ctx, double cs)
{
    const int D = ctx->s.dim, G = ctx->ngroups;
    launch_group_bbox(D, ctx->pnp1.x, ctx->gstart, G, ctx->gbox, ctx->st);
    std::vector<double> box(size_t(G) * 6);
    HIPCHK(hipMemcpyAsync(box.data(), ctx->gbox, sizeof(double) * box.size(), hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    std::vector<double> glo(size_t(G) * 3, 0.0);
    ctx->gl0_h.assign(size_t(G) * 2, 0);
    for (int k = 0; k < 3; ++k) ctx->dims[k] = 1;
    int64_t layer = 0;
    for (int g = 0; g < G; ++g) {
        ctx->grs[size_t(g)].xmax = 0.0;
        for (int k = 0; k < D; ++k) {
            const double lo = box[g * 6 + k], hi = box[g * 6 + 3 + k];
            if (!std::isfinite(lo) || !std::isfinite(hi)) throw std::string("non-finite particle position");
            ctx->grs[size_t(g)].xmax = std::max(ctx->grs[size_t(g)].xmax, std::max(std::fabs(lo), std::fabs(hi)));
            glo[g * 3 + k] = lo;
            const double span = (hi - lo) / cs;
            if (span > 1e9) throw std::string("a group's particle cloud is too large for the cell grid");
            const int32_t nk = int32_t(std::floor(span)) + 1;
            if (k < D - 1) ctx->dims[k] = std::max(ctx->dims[k], nk);
            else {
                ctx->gl0_h[g * 2] = int32_t(layer);
                ctx->gl0_h[g * 2 + 1] = nk;
                layer += nk + (g + 1 < G ? 1 : 0);
            }
        }
    }
    if (layer > 0x7fffffff) throw std::string("too many cell layers");
    ctx->dims[D - 1] = int32_t(layer);
    if (layer > ctx->lgrp_cap) {
        ctx->lgrp_cap = layer + layer / 4 + 64;
        dalloc(ctx->lgrp, size_t(ctx->lgrp_cap));
    }
    std::vector<int32_t> lg(size_t(layer), -1);
    for (int g = 0; g < G; ++g)
        for (int32_t l = 0; l < ctx->gl0_h[g * 2 + 1]; ++l) lg[size_t(ctx->gl0_h[g * 2] + l)] = g;
    HIPCHK(hipMemcpyAsync(ctx->glo, glo.data(), sizeof(double) * glo.size(), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipMemcpyAsync(ctx->gl0, ctx->gl0_h.data(), sizeof(int32_t) * ctx->gl0_h.size(), hipMemcpyHostToDevice,
                          ctx->st));
    HIPCHK(hipMemcpyAsync(ctx->lgrp, lg.data(), sizeof(int32_t) * lg.size(), hipMemcpyHostToDevice, ctx->st));
    sync(ctx);
    double ncell_d = 1.0;
    for (int k = 0; k < 3; ++k) ncell_d *= ctx->dims[k];
    return ncell_d;
}

/* Tile-block ranges of the groups for the tile grid of the last build. */
void group_blocks(fjsph_ctx* ctx)
{
    const int D = ctx->s.dim, G = ctx->ngroups;
    const int64_t per_layer = (D == 3 ? int64_t(ctx->tg.dims[1]) : 1) * ctx->tg.nchunk;
    ctx->gblk_h.assign(size_t(G) + 1, 0);
    for (int g = 0; g < G; ++g) ctx->gblk_h[g] = int64_t(ctx->gl0_h[g * 2]) * per_layer;
    ctx->gblk_h[G] = int64_t(ctx->gl0_h[(G - 1) * 2] + ctx->gl0_h[(G - 1) * 2 + 1]) * per_layer;
    /* the gap layer after group g belongs to no group: its blocks are empty
     * and write zero partials, so the ranges may include them */
    for (int g = 0; g + 1 < G; ++g) ctx->gblk_h[g + 1] = std::max(ctx->gblk_h[g + 1], ctx->gblk_h[g]);
    HIPCHK(hipMemcpyAsync(ctx->gblk, ctx->gblk_h.data(), sizeof(int64_t) * ctx->gblk_h.size(), hipMemcpyHostToDevice,
                          ctx->st));
}

void build_neighbours(fjsph_ctx* ctx, bool reorder)
{
    PassTimer t(ctx, FJSPH_T_NEIGHBOUR);
    const int D = ctx->s.dim;
    const int64_t n = ctx->n;
    const double cs = std::sqrt(ctx->d.sr) * (1.0 + 1e-7);
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double ncell_d = 1.0;
    if (grouped(ctx)) {
        ncell_d = group_layout(ctx, cs);
    } else {
        launch_bbox(D, ctx->pnp1.x, n, ctx->partial, ctx->nparts, ctx->sc->bbox, ctx->st);
        HIPCHK(hipMemcpyAsync(ctx->sc_host->bbox, ctx->sc->bbox, sizeof(double) * 6, hipMemcpyDeviceToHost,
                              ctx->st));
        sync(ctx);
        for (int k = 0; k < D; ++k) { lo[k] = ctx->sc_host->bbox[k]; hi[k] = ctx->sc_host->bbox[D + k]; }
        for (int k = 0; k < D; ++k)
            if (!std::isfinite(lo[k]) || !std::isfinite(hi[k]))
                throw std::string("non-finite particle position");
        for (int k = 0; k < 3; ++k) {
            ctx->dims[k] = 1;
            if (k < D) {
                const double span = (hi[k] - lo[k]) / cs;
                ctx->dims[k] = int32_t(std::floor(span)) + 1;
            }
            ncell_d *= ctx->dims[k];
        }
    }
    const double cell_limit = std::max(4.0 * double(n), 1.0e6) + 64.0;
    if (ncell_d > cell_limit || ncell_d > 2.0e9)
        throw std::string("particle cloud too sparse for the dense cell grid (") + std::to_string(ncell_d) + " cells)";
    const int64_t ncell = int64_t(ncell_d);
    if (ncell > ctx->ncell_cap) {
        const int64_t cap = ncell + ncell / 4 + 64;
        dalloc(ctx->cell_count, cap);
        dalloc(ctx->cell_fill, cap);
        dalloc(ctx->cell_start, cap + 1);
        dalloc(ctx->tile_sums, cap / 2048 + 2);
        ctx->ncell_cap = cap;
    }
    ctx->ncell = ncell;
    const double inv_cs = 1.0 / cs;
    HIPCHK(hipMemsetAsync(ctx->cell_count, 0, sizeof(int32_t) * ncell, ctx->st));
    HIPCHK(hipMemsetAsync(ctx->cell_fill, 0, sizeof(int32_t) * ncell, ctx->st));
    launch_cells(D, ctx->pnp1.x, n, lo, inv_cs, ctx->dims, ctx->cell, ctx->cell_count, ctx->st, ctx->grp, ctx->glo,
                 ctx->gl0);
    launch_scan(ctx->cell_count, ncell, ctx->cell_start, ctx->tile_sums, ctx->st);
    launch_scatter(ctx->cell, n, ctx->cell_start, ctx->cell_fill, ctx->order, ctx->st);
    launch_sort_cells(ctx->cell_start, ncell, ctx->order, ctx->st);
    if (reorder) {
        permute_all(ctx, ctx->order);
        launch_iota(ctx->order, n, ctx->st);
        if (ctx->ncols > 0) launch_inlet_backslot(ctx->col, ctx->depth, n, ctx->c_back, ctx->st);
    }
    HIPCHK(hipMemcpyAsync(ctx->aux.xb, ctx->pnp1.x, sizeof(double) * n * D, hipMemcpyDeviceToDevice, ctx->st));
    /* tile grid: chunks of C cells sized for about one workgroup of targets */
    auto set_tiles = [&](int32_t C) {
        TileGrid& tg = ctx->tg;
        for (int k = 0; k < 3; ++k) tg.dims[k] = ctx->dims[k];
        tg.C = C;
        tg.nchunk = (ctx->dims[0] + C - 1) / C;
        tg.nblocks = int64_t(ctx->dims[1]) * ctx->dims[2] * tg.nchunk;
        tg.start = ctx->cell_start;
        tg.order = ctx->order;
        tg.vorder = ctx->vorder;
        tg.pos = ctx->pos;
        tg.cell = ctx->cell;
        tg.lgrp = grouped(ctx) ? ctx->lgrp : nullptr;
        tg.gactive = grouped(ctx) ? ctx->gactive : nullptr;
        tg.gnpd = grouped(ctx) ? ctx->gnpd : nullptr;
        if (tg.nblocks > 0x7fffffff) throw std::string("tile grid too large");
        if (tg.nblocks > ctx->partial_cap) {
            ctx->partial_cap = tg.nblocks + tg.nblocks / 4 + 64;
            dalloc(ctx->partial, size_t(std::max<int64_t>(ctx->partial_cap, int64_t(ctx->nparts) * 8)));
        }
    };
    int32_t C;
    {
        const double occ = std::pow(cs / ctx->s.particle_step, D);
        const double targets = ctx->block_targets > 0 ? double(ctx->block_targets) : 256.0;
        C = int32_t(std::lround(targets / std::max(occ, 1.0)));
        /* at most kMaxBlockCells cells per block: the list build's FP32
         * prefilter band (k_build_t) holds for blocks up to that length */
        C = std::max(1, std::min({C, ctx->dims[0], kMaxBlockCells}));
    }
    set_tiles(C);
    for (int attempt = 0; attempt < 40; ++attempt) {
        HIPCHK(hipMemsetAsync(&ctx->sc->max_count, 0, 2 * sizeof(int32_t), ctx->st));
        launch_build_lists(D, ctx->pnp1.x, n, ctx->tg, ctx->d.sr, ctx->nl, ctx->sc, ctx->st);
        HIPCHK(hipMemcpyAsync(&ctx->sc_host->max_count, &ctx->sc->max_count, 2 * sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->st));
        sync(ctx);
        if (ctx->sc_host->overflow) {
            /* a clustered region: shorter blocks have smaller unions */
            if (ctx->tg.C > 1) {
                set_tiles(std::max(1, ctx->tg.C / 2));
                continue;
            }
            throw std::string("neighbour candidate union of a one-cell tile block exceeds ") +
                std::to_string(NO_ENTRY / TILE) + " LDS tiles (union of " + std::to_string(ctx->sc_host->overflow) +
                "); particles are too clustered for the tile grid";
        }
        const int32_t mc = ctx->sc_host->max_count;
        ctx->max_count = mc;
        if (mc <= ctx->nl.cap) break;
        /* capacity in entries, a multiple of the 4-entry batch */
        const int32_t cap = ((mc + mc / 8 + 8) + 7) / 8 * 8;
        dalloc(ctx->nl.lu4, size_t(cap / 4) * ctx->nl.npad);
        ctx->nl.cap = cap;
    }
    if (ctx->max_count > ctx->nl.cap) throw std::string("neighbour list capacity could not be established");
    if (grouped(ctx)) group_blocks(ctx);
    ctx->have_nbrs = true;
    ctx->x_at_build = true;
}

/* ------------------------------------------------------------ the passes */

void require(fjsph_ctx* ctx, bool nbrs)
{
    if (!ctx->have_state) throw std::string("no particle state uploaded");
    if (nbrs && !ctx->have_nbrs) throw std::string("neighbour lists not built");
}

void refresh_aux(fjsph_ctx* ctx)
{
    launch_aux(ctx->s.dim, ctx->pnp1, ctx->aux, ctx->n, ctx->s.ale, ctx->s.precision == FJSPH_PREC_MIXED, ctx->c.B * ctx->c.gam,
               ctx->st);
}

int r_from_x(const fjsph_ctx* ctx) { return ctx->x_at_build ? 1 : 0; }

void compact_deleted(fjsph_ctx* ctx, const std::vector<int32_t>& cnt, int64_t ndel);

/* Per-block deletion counters, zeroed. */
void zero_blkcount(fjsph_ctx* ctx)
{
    const size_t need = std::max<size_t>(ctx->blocks.size(), 1);
    if (ctx->blkcount_cap < need) {
        dalloc(ctx->blkcount, need);
        ctx->blkcount_cap = need;
    }
    HIPCHK(hipMemsetAsync(ctx->blkcount, 0, sizeof(int32_t) * need, ctx->st));
}

/* get_aero_velocity for the mesh source (src/Resid.cpp:479-523): FindCell.
 * With `classify`, the particles in no cell get FindCell's boundary test
 * and those it condemns are deleted from both states (returns true; the
 * caller builds the neighbour lists again).  Without it (the lookup after
 * Check_Pipe_Outlet, whose FirstCell leaves those counters alone) they are
 * only counted. */
bool mesh_lookup(fjsph_ctx* ctx, bool classify)
{
    if (!ctx->have_mesh) throw std::string("aero_source is the mesh but no mesh was set (fjsph_set_mesh)");
    PassTimer t(ctx, FJSPH_T_AERO);
    const int64_t n = ctx->n;
    if (classify && ctx->lost_cap < n) {
        const int64_t cap = n + n / 8 + 64;
        dalloc(ctx->lost_list, size_t(cap));
        dalloc(ctx->lost_del, size_t(cap));
        ctx->lost_cap = cap;
    }
    HIPCHK(hipMemsetAsync(&ctx->sc->mesh_lost, 0, 5 * sizeof(int32_t), ctx->st));
    launch_find_cell(ctx->s.dim, ctx->c, ctx->pnp1, ctx->part_id, ctx->dm, n, ctx->sc,
                     classify ? ctx->lost_list : nullptr, ctx->st);
    if (!classify) return false;
    HIPCHK(hipMemcpyAsync(&ctx->sc_host->mesh_lost, &ctx->sc->mesh_lost, sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->st));
    sync(ctx);
    const int32_t nlost = ctx->sc_host->mesh_lost;
    if (nlost == 0) return false;
    launch_lost_classify(ctx->s.dim, ctx->pnp1, ctx->dm, ctx->lost_list, nlost, ctx->lost_del, ctx->sc, ctx->st);
    HIPCHK(hipMemcpyAsync(&ctx->sc_host->mesh_lost, &ctx->sc->mesh_lost, 5 * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->st));
    sync(ctx);
    const DevScalars& h = *ctx->sc_host;
    if (h.lost_err)
        throw std::string("mesh lookup: the nearest-centre search of a particle in no cell did not converge "
                          "(thousands of cell centres at one distance)");
    ctx->lost_internal += h.lost_int;
    ctx->lost_failed += h.lost_fail;
    if (h.lost_del == 0) return false;
    const int64_t ndel = h.lost_del;
    const size_t nbk = ctx->blocks.size();
    zero_blkcount(ctx);
    launch_lost_marks(ctx->lost_list, ctx->lost_del, nlost, ctx->part_id, ctx->blk, n, ctx->keep, ctx->hflag,
                      ctx->blkcount, ctx->st);
    std::vector<int32_t> cnt(std::max<size_t>(nbk, 1));
    HIPCHK(hipMemcpyAsync(cnt.data(), ctx->blkcount, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    int64_t inblocks = 0;
    for (size_t k = 0; k < nbk; ++k) inblocks += cnt[k];
    if (nbk > 0 && inblocks != ndel) throw std::string("mesh lookup: a deleted particle belongs to no block");
    compact_deleted(ctx, cnt, ndel);
    ctx->lost_deleted += ndel;
    ctx->have_nbrs = false;
    return true;
}

/* get_aero_velocity for the surrogate source: the MLP in place of FindCell,
 * over the compacted list of the aero particles. */
void surrogate_lookup(fjsph_ctx* ctx)
{
    if (!ctx->have_surrogate) throw std::string("aero_source is the surrogate but none was set (fjsph_set_surrogate)");
    PassTimer t(ctx, FJSPH_T_AERO);
    const int64_t nseg = (ctx->n + AERO_SEG - 1) / AERO_SEG;
    if (ctx->aero_list_cap < nseg) {
        ctx->aero_list_cap = nseg + nseg / 8 + 1;
        dalloc(ctx->aero_list, size_t(ctx->aero_list_cap * AERO_SEG));
        dalloc(ctx->aero_counts, size_t(ctx->aero_list_cap * AERO_CSTRIDE));
    }
    HIPCHK(hipMemsetAsync(ctx->aero_counts, 0, sizeof(int32_t) * nseg * AERO_CSTRIDE, ctx->st));
    launch_aero_list(ctx->s.dim, ctx->c, ctx->pnp1, ctx->part_id, ctx->n, ctx->aero_list, ctx->aero_counts, ctx->st);
    launch_surrogate_list(ctx->s.dim, ctx->pnp1, ctx->sg, ctx->n, ctx->aero_list, ctx->aero_counts, ctx->st);
}

void dsph(fjsph_ctx* ctx, bool with_aero);

/* get_aero_velocity with the mesh: FindCell, and after deletions the
 * neighbour lists and dSPH_PreStep again (src/Resid.cpp:483-519). */
void aero_mesh(fjsph_ctx* ctx)
{
    if (!mesh_lookup(ctx, !ctx->tuning)) return;
    build_neighbours(ctx, false);
    dsph(ctx, false);
}

void dsph(fjsph_ctx* ctx, bool with_aero)
{
    {
        PassTimer t(ctx, FJSPH_T_DSPH);
        refresh_aux(ctx);
        launch_dsph(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl, ctx->n, r_from_x(ctx),
                    with_aero ? 1 : 0, ctx->partial, ctx->sc, ctx->st);
        /* each group's npd: its blocks' partials over its particle count */
        if (grouped(ctx)) launch_seg_sum(ctx->partial, ctx->gblk, ctx->ngroups, ctx->gscale, ctx->gnpd, ctx->st);
    }
    if (with_aero && ctx->s.aero_source == FJSPH_MESH_INFL) aero_mesh(ctx);
    if (with_aero && ctx->s.aero_source == FJSPH_SURROGATE) surrogate_lookup(ctx);
}

/* Detect_Surface loop 1 and dissipation_terms share one staged pass
 * (measured faster than two separate passes at 4M particles). */
#ifndef FJSPH_SURFACE_FUSED
#define FJSPH_SURFACE_FUSED 1
#endif
/* Detect_Surface loops 2-3 and particle_shift share one staged pass in
 * mixed precision (k_surf2_shift_m). */
#ifndef FJSPH_SHIFT_FUSED
#define FJSPH_SHIFT_FUSED 1
#endif

/* with_shift: particle_shift runs in the same pass as Detect_Surface loops
 * 2-3 where that fused form exists (mixed precision). */
void surface_and_dissipation(fjsph_ctx* ctx, bool with_diss, bool with_shift)
{
    PassTimer t(ctx, FJSPH_T_SURFACE);
    if (FJSPH_SURFACE_FUSED || !with_diss) {
        launch_surface1(ctx->s.dim, ctx->s.precision, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl,
                        ctx->n, r_from_x(ctx), ctx->norms_tmp, with_diss ? 1 : 0, ctx->st);
    } else {
        launch_surface1(ctx->s.dim, ctx->s.precision, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl,
                        ctx->n, r_from_x(ctx), ctx->norms_tmp, 0, ctx->st);
        launch_surface1(ctx->s.dim, ctx->s.precision, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl,
                        ctx->n, r_from_x(ctx), ctx->norms_tmp, 2, ctx->st);
    }
    if (with_shift) {
        launch_surface2_shift(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl, ctx->n,
                              r_from_x(ctx), ctx->norms_tmp, ctx->st);
        return;
    }
    launch_surface2(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl, ctx->n, r_from_x(ctx),
                    ctx->norms_tmp, ctx->st);
}

void shift(fjsph_ctx* ctx)
{
    PassTimer t(ctx, FJSPH_T_SHIFT);
    launch_shift(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl, ctx->n, r_from_x(ctx),
                 ctx->st);
}

/* The reference's unset planes (bound_block constructor, src/Var.h:827-831):
 * no deletion, insertion and aero planes at default_val. */
fjsph_block_planes default_planes()
{
    fjsph_block_planes q{};
    for (int a = 0; a < 3; ++a) {
        q.delete_norm[a] = FJSPH_NO_PLANE;
        q.insert_norm[a] = FJSPH_NO_PLANE;
        q.aero_norm[a] = FJSPH_NO_PLANE;
    }
    q.delconst = q.insconst = q.aeroconst = FJSPH_NO_PLANE;
    return q;
}

/* Per-block planes to the device (4 doubles each: normal, constant). */
void upload_planes(fjsph_ctx* ctx)
{
    if (!ctx->planes_dirty) return;
    const size_t nbk = std::max<size_t>(ctx->blocks.size(), 1);
    if (ctx->planes.size() < ctx->blocks.size()) ctx->planes.resize(ctx->blocks.size(), default_planes());
    std::vector<double> del(nbk * 4, 0.0), ins(nbk * 4, 0.0), aero(nbk * 4, 0.0);
    std::vector<int32_t> mode(nbk, 0), nbuf(nbk, 0);
    for (size_t k = 0; k < ctx->blocks.size(); ++k) {
        const fjsph_block_planes& q = ctx->planes[k];
        for (int a = 0; a < 3; ++a) {
            del[k * 4 + a] = q.delete_norm[a];
            ins[k * 4 + a] = q.insert_norm[a];
            aero[k * 4 + a] = q.aero_norm[a];
        }
        del[k * 4 + 3] = q.delconst;
        ins[k * 4 + 3] = q.insconst;
        aero[k * 4 + 3] = q.aeroconst;
        mode[k] = q.fixed_vel_or_dynamic;
    }
    for (const auto& in : ctx->inlets) nbuf[size_t(in.block)] = in.nbuf;
    dalloc(ctx->pl_del, del.size());
    dalloc(ctx->pl_ins, ins.size());
    dalloc(ctx->pl_aero, aero.size());
    dalloc(ctx->pl_mode, mode.size());
    dalloc(ctx->pl_nbuf, nbuf.size());
    HIPCHK(hipMemcpy(ctx->pl_del, del.data(), del.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->pl_ins, ins.data(), ins.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->pl_aero, aero.data(), aero.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->pl_mode, mode.data(), mode.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->pl_nbuf, nbuf.data(), nbuf.size() * 4, hipMemcpyHostToDevice));
    ctx->dpl = DevPlanes{ctx->pl_del, ctx->pl_ins, ctx->pl_aero, ctx->pl_mode, ctx->pl_nbuf};
    ctx->planes_dirty = false;
}

/* Check_Pipe_Outlet (src/Containment.cpp:822-890): PIPE particles past
 * their block's aero plane become FREE; with the mesh source the new FREE
 * particles get their cell (FirstCell; the lookup of the others finds
 * their cached cell again).  The reference's deletion of a particle that
 * FirstCell places in no cell is not made: it is counted by the lookup. */
void pipe_outlet(fjsph_ctx* ctx)
{
    if (!ctx->planes_on) return;
    upload_planes(ctx);
    const bool mesh = ctx->s.aero_source == FJSPH_MESH_INFL && ctx->have_mesh;
    const bool first = mesh && !ctx->tuning;
    if (first) {
        if (ctx->fresh_cap < ctx->n) {
            const int64_t cap = ctx->n + ctx->n / 8 + 64;
            dalloc(ctx->fresh, size_t(cap));
            dalloc(ctx->fc_del, size_t(cap));
            ctx->fresh_cap = cap;
        }
        HIPCHK(hipMemsetAsync(&ctx->sc->fresh_n, 0, 3 * sizeof(int32_t), ctx->st));
    }
    launch_pipe_outlet(ctx->s.dim, ctx->pnp1, ctx->part_id, ctx->blk, ctx->dpl, ctx->n, ctx->nb, ctx->st,
                       first ? ctx->fresh : nullptr, ctx->s.lam_cutoff, ctx->sc);
    if (!mesh) return;
    mesh_lookup(ctx, false);
    if (!first) return;
    /* FirstCell for the particles leaving the pipe that the lookup placed in
     * no cell (src/Containment.cpp:466-573) */
    HIPCHK(hipMemcpyAsync(&ctx->sc_host->fresh_n, &ctx->sc->fresh_n, sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->st));
    sync(ctx);
    const int32_t nfresh = ctx->sc_host->fresh_n;
    if (nfresh == 0) return;
    HIPCHK(hipMemsetAsync(&ctx->sc->lost_err, 0, sizeof(int32_t), ctx->st));
    launch_first_cell(ctx->s.dim, ctx->pnp1, ctx->dm, ctx->fresh, nfresh, ctx->fc_del, ctx->sc, ctx->st);
    HIPCHK(hipMemcpyAsync(&ctx->sc_host->lost_err, &ctx->sc->lost_err, 4 * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->st));
    sync(ctx);
    const DevScalars& h = *ctx->sc_host;
    if (h.lost_err & 1)
        throw std::string("FirstCell: the nearest-centre search of a particle leaving the pipe did not converge");
    if (h.lost_err & 2)
        throw std::string("FirstCell: first containing cell for a particle leaving the pipe not found (no boundary "
                          "crossing; the reference stops here)");
    ctx->first_cell += h.fc_count;
    if (h.fc_del == 0) return;
    /* the reference erases these particles from pnp1 alone and keeps its
     * neighbour lists; here they leave both states and the lists and
     * dSPH_PreStep are made again, as after FindCell's deletions */
    const size_t nbk = ctx->blocks.size();
    zero_blkcount(ctx);
    launch_lost_marks(ctx->fresh, ctx->fc_del, nfresh, ctx->part_id, ctx->blk, ctx->n, ctx->keep, ctx->hflag,
                      ctx->blkcount, ctx->st);
    std::vector<int32_t> cnt(std::max<size_t>(nbk, 1));
    HIPCHK(hipMemcpyAsync(cnt.data(), ctx->blkcount, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    compact_deleted(ctx, cnt, h.fc_del);
    ctx->first_cell_deleted += h.fc_del;
    ctx->lost_deleted += h.fc_del;
    ctx->have_nbrs = false;
    build_neighbours(ctx, false);
    dsph(ctx, false);
}

/* dSPH_PreStep, get_aero_velocity, Detect_Surface, dissipation_terms,
 * particle_shift, Check_Pipe_Outlet (src/Integration.cpp:45-61 and
 * 80-98). */
void prepasses(fjsph_ctx* ctx)
{
    dsph(ctx, true);
    const bool fused_shift = FJSPH_SHIFT_FUSED && ctx->s.ale && ctx->nl.f32;
    surface_and_dissipation(ctx, true, fused_shift);
    if (ctx->s.ale && !fused_shift) shift(ctx);
    pipe_outlet(ctx);
}

void upload_block_velocities(fjsph_ctx* ctx, bool rk)
{
    bool moving = false;
    for (auto& hb : ctx->blocks) moving |= hb.b.n_times != 0;
    if (!moving) return;
    std::vector<double> vel(ctx->blocks.size() * 3, 0.0);
    for (size_t k = 0; k < ctx->blocks.size(); ++k) {
        const HostBlock& hb = ctx->blocks[k];
        if (hb.b.n_times == 0) {
            for (int a = 0; a < 3; ++a) vel[k * 3 + a] = hb.b.vel[a];
            continue;
        }
        for (int t = 0; t < hb.b.n_times; ++t) {
            const bool cond = rk ? (hb.times[t] > ctx->current_time) : (ctx->current_time > hb.times[t]);
            if (cond) for (int a = 0; a < 3; ++a) vel[k * 3 + a] = hb.vels[t * 3 + a];
        }
    }
    HIPCHK(hipMemcpyAsync(ctx->bl.vel, vel.data(), sizeof(double) * vel.size(), hipMemcpyHostToDevice, ctx->st));
    sync(ctx);
}

void boundary_update(fjsph_ctx* ctx)
{
    if (ctx->nb == 0) return;
    PassTimer t(ctx, FJSPH_T_BOUNDARY);
    upload_block_velocities(ctx, false);
    refresh_aux(ctx);
    for (size_t k = 0; k < ctx->blocks.size(); ++k) {
        if (ctx->blocks[k].b.is_fluid) continue;
        launch_boundary(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->blk, ctx->bl, int32_t(k), ctx->tg, ctx->nl,
                        ctx->n, ctx->nb, ctx->bnd_slots, ctx->near_inlet, ctx->st);
    }
}

void force(fjsph_ctx* ctx, bool update)
{
    refresh_aux(ctx);
    {
        PassTimer t(ctx, FJSPH_T_FORCE);
        NbrList nl64 = ctx->nl;
        nl64.f32 = 0;
        bool tail = false;
        if (grouped(ctx))
            for (const fjsph_ctx::Group& g : ctx->grs) tail |= g.active && g.f64_tail;
        if (!ctx->nl.f32 || (!grouped(ctx) && !ctx->f64_tail)) {
            launch_force(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl, r_from_x(ctx),
                         ctx->sc, ctx->st);
        } else if (!grouped(ctx)) {
            launch_force(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, nl64, r_from_x(ctx), ctx->sc,
                         ctx->st);
        } else if (!tail) {
            launch_force(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl, r_from_x(ctx),
                         ctx->sc, ctx->st);
        } else {
            /* groups whose step went to FP64 pair terms and the others, each
             * launch skipping the other's blocks */
            TileGrid tm = ctx->tg, td = ctx->tg;
            tm.gactive = ctx->gact_m;
            td.gactive = ctx->gact_d;
            launch_force(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, tm, ctx->nl, r_from_x(ctx), ctx->sc,
                         ctx->st);
            launch_force(ctx->s.dim, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, td, nl64, r_from_x(ctx), ctx->sc,
                         ctx->st);
        }
    }
    PassTimer t(ctx, FJSPH_T_UPDATE);
    const DevGroups gr = grouped(ctx) ? DevGroups{ctx->grp, ctx->gactive, ctx->gdt, ctx->gd2} : DevGroups{};
    launch_nb_update(ctx->s.dim, ctx->c, ctx->pnp1, ctx->pn, update ? ctx->x_nxt : nullptr, ctx->v_nxt,
                     ctx->rho_nxt, ctx->p_nxt, ctx->part_id, ctx->blk, ctx->bl, ctx->near_inlet, ctx->nl, ctx->n,
                     ctx->delta_t, ctx->partial, ctx->st, gr);
    if (update && ctx->ncols > 0) {
        /* the inlet buffer of Do_NB_Iter, after the fluid update */
        upload_planes(ctx);
        launch_inlet_buffer(ctx->s.dim, ctx->c, ctx->pnp1, ctx->pn,
                            NextState{ctx->x_nxt, ctx->v_nxt, ctx->rho_nxt, ctx->p_nxt}, ctx->col, ctx->depth,
                            ctx->blk, ctx->c_back, ctx->dpl, ctx->n, ctx->delta_t, ctx->d.dx, ctx->partial, ctx->st);
    }
}

/* Do_NB_Iter + the residual of Check_Error; returns log10 rms. */
double nb_iter(fjsph_ctx* ctx)
{
    boundary_update(ctx);
    force(ctx, true);
    std::swap(ctx->pnp1.x, ctx->x_nxt);
    std::swap(ctx->pnp1.v, ctx->v_nxt);
    std::swap(ctx->pnp1.rho, ctx->rho_nxt);
    std::swap(ctx->pnp1.p, ctx->p_nxt);
    ctx->x_at_build = false;
    {
        PassTimer t(ctx, FJSPH_T_REDUCE);
        /* one residual partial per TPB slots (k_nb_update) */
        launch_sum_partials(ctx->partial, reduction_parts(ctx->n), 1, &ctx->sc->residual, ctx->st);
        HIPCHK(hipMemcpyAsync(&ctx->sc_host->residual, &ctx->sc->residual, sizeof(double), hipMemcpyDeviceToHost,
                              ctx->st));
    }
    sync(ctx);
    const double nf = double(ctx->n - ctx->nb);
    return std::log10(std::sqrt(ctx->sc_host->residual / nf));
}

/* The time step of Integrator::find_timestep (src/Integration.cpp:370-443)
 * from the seven reduced maxima/minima r, for one simulation's state st
 * (the context or a group: maxima, CFL number, safe step). */
template <typename S>
double timestep_of(const fjsph_ctx* ctx, S& st, const double* r, double time, double last_frame_time)
{
    const fjsph_settings& s = ctx->s;
    const fjsph_derived& d = ctx->d;
    st.maxf = r[0];
    st.maxAf = r[1];
    st.maxdrho = r[2];
    st.maxRhoi = r[3];
    st.maxU = r[4];
    st.maxShift = r[5];
    st.minST = r[6];
    st.maxRho_pc = 100 * st.maxRhoi / s.rho_rest;
    double mn = 0.25 * std::sqrt(d.H / st.maxf);
    mn = std::min(mn, 2 * d.H / st.maxU);
    if (s.mu != 0.0) mn = std::min(mn, 0.125 * d.H_sq * s.rho_rest / s.mu);
    mn = std::min(mn, 0.067 * st.minST);
    mn = std::min(mn, 0.5 * std::sqrt(d.H / st.maxdrho));
    mn = std::min(mn, 1.5 * d.H / s.speed_sound);
    st.safe_dt = 0.75 * mn;
    double dt = st.cfl * st.safe_dt;
    if (dt < s.delta_t_min) dt = s.delta_t_min;
    else if (dt > s.delta_t_max) dt = s.delta_t_max;
    if (dt > last_frame_time + s.frame_time_interval - time)
        dt = last_frame_time + s.frame_time_interval - time + s.delta_t_min;
    return dt;
}

/* Integrator::find_timestep (src/Integration.cpp:370-443). */
double find_timestep(fjsph_ctx* ctx)
{
    {
        PassTimer t(ctx, FJSPH_T_REDUCE);
        launch_timestep_reduce(ctx->s.dim, ctx->c, ctx->pnp1, ctx->part_id, ctx->n, ctx->partial, ctx->nparts,
                               ctx->sc->red, ctx->st);
        HIPCHK(hipMemcpyAsync(ctx->sc_host->red, ctx->sc->red, sizeof(double) * 7, hipMemcpyDeviceToHost, ctx->st));
    }
    sync(ctx);
    return timestep_of(ctx, *ctx, ctx->sc_host->red, ctx->current_time, ctx->last_frame_time);
}

/* find_timestep of every group: the reduction over each group's particles,
 * then each group's own formula, CFL number and frame end; the time steps
 * go to the device for the Newmark update. */
void find_timestep_groups(fjsph_ctx* ctx)
{
    const int G = ctx->ngroups;
    {
        PassTimer t(ctx, FJSPH_T_REDUCE);
        launch_group_dt(ctx->s.dim, ctx->c, ctx->pnp1, ctx->part_id, ctx->gstart, G, ctx->gred, ctx->st);
    }
    std::vector<double> r(size_t(G) * 8);
    HIPCHK(hipMemcpyAsync(r.data(), ctx->gred, sizeof(double) * r.size(), hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    std::vector<double> dts(static_cast<size_t>(G));
    for (int g = 0; g < G; ++g) {
        fjsph_ctx::Group& gr = ctx->grs[g];
        gr.dt = timestep_of(ctx, gr, &r[size_t(g) * 8], gr.time, gr.last_frame_time);
        dts[g] = gr.dt;
    }
    HIPCHK(hipMemcpyAsync(ctx->gdt, dts.data(), sizeof(double) * G, hipMemcpyHostToDevice, ctx->st));
    sync(ctx);
}

void rollback(fjsph_ctx* ctx)
{
    ctx->f64_tail = false;
    copy_state(ctx, ctx->pnp1, ctx->pn);
    build_neighbours(ctx, false);
    ctx->delta_t = 0.5 * ctx->delta_t;
    ctx->rollbacks++;
}

/* log10 of 100 FP64 ulps of coordinates up to xmax.  Once the rms position
 * change of a sub-iteration is below it, the iterates differ in the last
 * bits of the positions, where FP32 pair terms change from one iterate to
 * the next by more than the FP64 ones they stand for: from there on the
 * mixed path runs the step's force passes in FP64, which settle on a fixed
 * point as the reference's do (a column at rest: 13 sub-iterations to an
 * exact fixed point in FP64, the iteration limit in FP32 pair terms, whose
 * rms change stays at ~1e-18 m).  Steps that converge above this level,
 * the usual case, never reach it. */
double rounding_level(double xmax) { return xmax > 0.0 ? std::log10(std::ldexp(xmax, -52)) + 2.0 : -HUGE_VAL; }

double build_xmax(const fjsph_ctx* ctx)
{
    double m = 0.0;
    for (int k = 0; k < 2 * ctx->s.dim; ++k) m = std::max(m, std::fabs(ctx->sc_host->bbox[k]));
    return m;
}

/* Newmark_Beta::Newmark_Beta + Check_Error (src/Newmark_Beta.cpp:10-52, 303-331). */
double newmark_loop(fjsph_ctx* ctx, double& logbase)
{
    double rms = 0.0;
    while (rms > ctx->s.min_residual) {
        const double le = nb_iter(ctx);
        if (ctx->iteration == 0) logbase = le;
        rms = le - logbase;
        if (ctx->nl.f32 && rms > ctx->s.min_residual && le < rounding_level(build_xmax(ctx))) ctx->f64_tail = true;
        if (ctx->iteration > unsigned(ctx->s.max_subits)) {
            if (rms > 0.0) {
                rollback(ctx);
                ctx->iteration = 0;
                rms = 0.0;
                continue;
            }
            break;
        }
        ctx->iteration++;
    }
    ctx->f64_tail = false;
    return rms;
}

/* ---- independent simulations (fjsph_set_groups) */

bool reorder_due(fjsph_ctx* ctx);

void upload_active(fjsph_ctx* ctx)
{
    const size_t G = ctx->grs.size();
    std::vector<int32_t> a(3 * G);
    for (size_t g = 0; g < G; ++g) {
        const fjsph_ctx::Group& gr = ctx->grs[g];
        a[g] = gr.active ? 1 : 0;
        a[G + g] = (gr.active && !gr.f64_tail) ? 1 : 0;
        a[2 * G + g] = (gr.active && gr.f64_tail) ? 1 : 0;
    }
    HIPCHK(hipMemcpyAsync(ctx->gactive, a.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipMemcpyAsync(ctx->gact_m, a.data() + G, sizeof(int32_t) * G, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipMemcpyAsync(ctx->gact_d, a.data() + 2 * G, sizeof(int32_t) * G, hipMemcpyHostToDevice, ctx->st));
    sync(ctx);
}

void upload_group_dt(fjsph_ctx* ctx)
{
    std::vector<double> d(ctx->grs.size());
    for (size_t g = 0; g < d.size(); ++g) d[g] = ctx->grs[g].dt;
    HIPCHK(hipMemcpyAsync(ctx->gdt, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice, ctx->st));
    sync(ctx);
}

/* Do_NB_Iter + Check_Error's residual for every group still iterating
 * (the others keep their state): le[g] = log10 rms of group g. */
void nb_iter_groups(fjsph_ctx* ctx, std::vector<double>& le)
{
    boundary_update(ctx);
    force(ctx, true);
    std::swap(ctx->pnp1.x, ctx->x_nxt);
    std::swap(ctx->pnp1.v, ctx->v_nxt);
    std::swap(ctx->pnp1.rho, ctx->rho_nxt);
    std::swap(ctx->pnp1.p, ctx->p_nxt);
    ctx->x_at_build = false;
    const int G = ctx->ngroups;
    {
        PassTimer t(ctx, FJSPH_T_REDUCE);
        launch_seg_sum(ctx->gd2, ctx->gstart, G, nullptr, ctx->gres, ctx->st);
    }
    std::vector<double> res(static_cast<size_t>(G));
    HIPCHK(hipMemcpyAsync(res.data(), ctx->gres, sizeof(double) * G, hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    le.resize(size_t(G));
    for (int g = 0; g < G; ++g) {
        const double nf = double(ctx->gstart_h[g + 1] - ctx->gstart_h[g]);
        le[g] = std::log10(std::sqrt(res[g] / nf));
    }
}

/* The lists of the groups in `mask` built again at their current positions
 * (after a rollback); the other groups keep their cells, lists and build
 * positions, as separate simulations would.  Every particle is binned from
 * its build position, which gives the other groups the cells of their last
 * build, so their candidate unions are unchanged.  Falls back to a full
 * build if the rebuilt lists need more room or shorter tile blocks. */
void rebuild_groups(fjsph_ctx* ctx, const std::vector<int32_t>& mask)
{
    const int D = ctx->s.dim;
    const int64_t n = ctx->n;
    {
        PassTimer t(ctx, FJSPH_T_NEIGHBOUR);
        for (int g = 0; g < ctx->ngroups; ++g)
            if (mask[g]) {
                const int64_t a = ctx->gstart_h[g], e = ctx->gstart_h[g + 1];
                HIPCHK(hipMemcpyAsync(ctx->aux.xb + a * D, ctx->pnp1.x + a * D, sizeof(double) * (e - a) * D,
                                      hipMemcpyDeviceToDevice, ctx->st));
            }
        const double cs = std::sqrt(ctx->d.sr) * (1.0 + 1e-7);
        const double lo[3] = {0, 0, 0};
        HIPCHK(hipMemsetAsync(ctx->cell_count, 0, sizeof(int32_t) * ctx->ncell, ctx->st));
        HIPCHK(hipMemsetAsync(ctx->cell_fill, 0, sizeof(int32_t) * ctx->ncell, ctx->st));
        launch_cells(D, ctx->aux.xb, n, lo, 1.0 / cs, ctx->dims, ctx->cell, ctx->cell_count, ctx->st, ctx->grp,
                     ctx->glo, ctx->gl0);
        launch_scan(ctx->cell_count, ctx->ncell, ctx->cell_start, ctx->tile_sums, ctx->st);
        launch_scatter(ctx->cell, n, ctx->cell_start, ctx->cell_fill, ctx->order, ctx->st);
        launch_sort_cells(ctx->cell_start, ctx->ncell, ctx->order, ctx->st);
        HIPCHK(hipMemcpyAsync(ctx->gbuild, mask.data(), sizeof(int32_t) * mask.size(), hipMemcpyHostToDevice,
                              ctx->st));
        HIPCHK(hipMemsetAsync(&ctx->sc->max_count, 0, 2 * sizeof(int32_t), ctx->st));
        TileGrid tg = ctx->tg;
        tg.gbuild = ctx->gbuild;
        launch_build_lists(D, ctx->aux.xb, n, tg, ctx->d.sr, ctx->nl, ctx->sc, ctx->st);
        HIPCHK(hipMemcpyAsync(&ctx->sc_host->max_count, &ctx->sc->max_count, 2 * sizeof(int32_t),
                              hipMemcpyDeviceToHost, ctx->st));
        sync(ctx);
    }
    if (ctx->sc_host->overflow || ctx->sc_host->max_count > ctx->nl.cap) {
        /* every group's lists again, at the current positions */
        build_neighbours(ctx, false);
    } else {
        ctx->max_count = std::max(ctx->max_count, ctx->sc_host->max_count);
    }
    ctx->x_at_build = false;
}

/* Newmark_Beta::Newmark_Beta + Check_Error (src/Newmark_Beta.cpp:10-52,
 * 303-331) for every group at once: each group iterates until its own
 * residual converges (then it keeps its state and its blocks are skipped),
 * rolls back on its own with its own time step halved. */
void newmark_loop_groups(fjsph_ctx* ctx)
{
    const fjsph_settings& s = ctx->s;
    const int G = ctx->ngroups;
    for (fjsph_ctx::Group& g : ctx->grs) {
        g.rms = 0.0;
        g.active = true;
        g.f64_tail = false;
    }
    upload_active(ctx);
    std::vector<double> le;
    std::vector<int32_t> roll(size_t(G), 0);
    for (;;) {
        bool any = false;
        for (const fjsph_ctx::Group& g : ctx->grs) any |= g.active;
        if (!any) break;
        nb_iter_groups(ctx, le);
        bool rolled = false;
        std::fill(roll.begin(), roll.end(), 0);
        for (int k = 0; k < G; ++k) {
            fjsph_ctx::Group& g = ctx->grs[k];
            if (!g.active) continue;
            if (g.iteration == 0) g.logbase = le[k];
            g.rms = le[k] - g.logbase;
            if (ctx->nl.f32 && g.rms > s.min_residual && le[k] < rounding_level(g.xmax)) g.f64_tail = true;
            if (g.iteration > unsigned(s.max_subits)) {
                if (g.rms > 0.0) {
                    roll[k] = 1;
                    rolled = true;
                    g.f64_tail = false;
                    g.dt *= 0.5;
                    g.rollbacks++;
                    g.iteration = 0;
                    g.rms = 0.0;
                    continue;
                }
                g.active = false;
                continue;
            }
            g.iteration++;
            if (!(g.rms > s.min_residual)) g.active = false;
        }
        if (rolled) {
            for (int k = 0; k < G; ++k)
                if (roll[k])
                    copy_state_range(ctx, ctx->pnp1, ctx->pn, ctx->gstart_h[k], ctx->gstart_h[k + 1] - ctx->gstart_h[k]);
            upload_group_dt(ctx);
            rebuild_groups(ctx, roll);
        }
        upload_active(ctx);
    }
    for (fjsph_ctx::Group& g : ctx->grs) g.f64_tail = false;
}

/* integrate_no_update (src/Integration.cpp:27-107) for every group. */
double integrate_no_update_groups(fjsph_ctx* ctx)
{
    const int G = ctx->ngroups;
    for (fjsph_ctx::Group& g : ctx->grs) {
        g.iteration = 0;
        g.rollbacks = 0;
        g.active = true;
    }
    upload_active(ctx);
    find_timestep_groups(ctx);
    build_neighbours(ctx, reorder_due(ctx));
    prepasses(ctx);
    std::vector<double> le;
    nb_iter_groups(ctx, le);
    for (int k = 0; k < G; ++k) {
        ctx->grs[k].logbase = le[k];
        ctx->grs[k].iteration++;
    }
    build_neighbours(ctx, false);
    prepasses(ctx);
    newmark_loop_groups(ctx);
    std::vector<double> npd(static_cast<size_t>(G));
    HIPCHK(hipMemcpyAsync(npd.data(), ctx->gnpd, sizeof(double) * G, hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    double worst = -1e300;
    unsigned it = 0;
    int rb = 0;
    for (int k = 0; k < G; ++k) {
        fjsph_ctx::Group& g = ctx->grs[k];
        g.npd = npd[k];
        worst = std::max(worst, g.rms);
        it = std::max(it, g.iteration);
        rb += g.rollbacks;
    }
    ctx->iteration = it;
    ctx->rollbacks = rb;
    drain_timers(ctx);
    return worst;
}

/* ---- Runge-Kutta 4 (src/Runge_Kutta.cpp:10-515, src/Integration.cpp:305-365).
 * Stage states are full copies of the previous stage, as in the reference;
 * every stage's force pass uses the distances frozen at the last build. */
void rk_stage(fjsph_ctx* ctx, DevState& dst, DevState& prev, double dt, bool final_stage)
{
    const int D = ctx->s.dim;
    const int mixed = ctx->s.precision == FJSPH_PREC_MIXED;
    if (&dst != &prev) copy_state(ctx, dst, prev);
    if (ctx->nb > 0) {
        PassTimer t(ctx, FJSPH_T_BOUNDARY);
        upload_block_velocities(ctx, true);
        launch_aux(D, dst, ctx->aux, ctx->n, ctx->s.ale, mixed, ctx->c.B * ctx->c.gam, ctx->st);
        for (size_t k = 0; k < ctx->blocks.size(); ++k) {
            if (ctx->blocks[k].b.is_fluid) continue;
            launch_boundary(D, ctx->c, dst, ctx->aux, ctx->blk, ctx->bl, int32_t(k), ctx->tg, ctx->nl, ctx->n,
                            ctx->nb, ctx->bnd_slots, ctx->near_inlet, ctx->st);
        }
        launch_rk_bnd_rho(D, ctx->c, dst, ctx->pn, ctx->rk1, ctx->rk2, ctx->rk3, ctx->blk, ctx->bl, ctx->near_inlet,
                          ctx->n, dt, final_stage ? 1 : 0, ctx->st);
    }
    launch_aux(D, dst, ctx->aux, ctx->n, ctx->s.ale, mixed, ctx->c.B * ctx->c.gam, ctx->st);
    {
        PassTimer t(ctx, FJSPH_T_FORCE);
        launch_force(D, ctx->c, dst, ctx->aux, ctx->part_id, ctx->tg, ctx->nl, 0, ctx->sc, ctx->st);
    }
    PassTimer t(ctx, FJSPH_T_UPDATE);
    launch_nb_update(D, ctx->c, dst, ctx->pn, nullptr, nullptr, nullptr, nullptr, ctx->part_id, ctx->blk, ctx->bl,
                     ctx->near_inlet, ctx->nl, ctx->n, dt, ctx->partial, ctx->st);
    launch_rk_update(D, ctx->c, dst, ctx->pn, ctx->rk1, ctx->rk2, ctx->rk3, ctx->part_id, ctx->n, dt,
                     final_stage ? 1 : 0, ctx->st);
    if (ctx->ncols > 0) {
        upload_planes(ctx);
        launch_rk_buffer(D, ctx->c, dst, ctx->pn, ctx->rk1, ctx->rk2, ctx->rk3, ctx->col, ctx->depth, ctx->blk,
                         ctx->c_back, ctx->dpl, ctx->n, dt, ctx->d.dx, final_stage ? 1 : 0, ctx->st);
    }
}

/* Check_RK_Error (src/Runge_Kutta.cpp:10-26): log10 rms of b.x - a.x over
 * the fluid, minus logbase. */
double rk_error(fjsph_ctx* ctx, const DevState& a, const DevState& b, double logbase)
{
    {
        PassTimer t(ctx, FJSPH_T_REDUCE);
        launch_sqdiff(ctx->s.dim, a.x, b.x, ctx->part_id, ctx->nb, ctx->n, ctx->partial, ctx->st);
        launch_sum_partials(ctx->partial, reduction_parts(ctx->n), 1, &ctx->sc->residual, ctx->st);
        HIPCHK(hipMemcpyAsync(&ctx->sc_host->residual, &ctx->sc->residual, sizeof(double), hipMemcpyDeviceToHost,
                              ctx->st));
    }
    sync(ctx);
    const double nf = double(ctx->n - ctx->nb);
    return std::log10(std::sqrt(ctx->sc_host->residual / nf)) - logbase;
}

/* Whether this step's first list build moves the particles into cell
 * order. */
bool reorder_due(fjsph_ctx* ctx)
{
    if (ctx->steps_to_reorder > 0) {
        --ctx->steps_to_reorder;
        return false;
    }
    ctx->steps_to_reorder = ctx->reorder_every - 1;
    return true;
}

/* integrate_no_update with the Runge-Kutta solver: Get_First_RK on the
 * state of the first build, then Runge_Kutta4 after the second. */
double integrate_no_update_rk(fjsph_ctx* ctx)
{
    if (!ctx->rk_alloc) {
        alloc_state(ctx, ctx->rk1);
        alloc_state(ctx, ctx->rk2);
        alloc_state(ctx, ctx->rk3);
        ctx->rk_alloc = true;
    }
    ctx->iteration = 0;
    ctx->rollbacks = 0;
    ctx->delta_t = find_timestep(ctx);
    build_neighbours(ctx, reorder_due(ctx));
    prepasses(ctx);
    rk_stage(ctx, ctx->pnp1, ctx->pn, 0.5 * ctx->delta_t, false);
    ctx->x_at_build = false;
    const double logbase = rk_error(ctx, ctx->pn, ctx->pnp1, 0.0);
    build_neighbours(ctx, false);
    prepasses(ctx);
    copy_state(ctx, ctx->rk1, ctx->pnp1);
    rk_stage(ctx, ctx->rk2, ctx->rk1, 0.5 * ctx->delta_t, false);
    rk_stage(ctx, ctx->rk3, ctx->rk2, ctx->delta_t, false);
    rk_stage(ctx, ctx->pnp1, ctx->rk3, ctx->delta_t, true);
    ctx->x_at_build = false;
    const double rms = rk_error(ctx, ctx->rk3, ctx->pnp1, logbase);
    HIPCHK(hipMemcpyAsync(&ctx->sc_host->npd, &ctx->sc->npd, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    ctx->last_npd = ctx->sc_host->npd;
    drain_timers(ctx);
    return rms;
}

/* integrate_no_update (src/Integration.cpp:27-107). */
double integrate_no_update(fjsph_ctx* ctx)
{
    if (grouped(ctx)) return integrate_no_update_groups(ctx);
    if (ctx->s.integrator == FJSPH_RUNGE_KUTTA) return integrate_no_update_rk(ctx);
    if (ctx->s.integrator != FJSPH_NEWMARK_BETA) throw std::string("unknown integrator");
    ctx->iteration = 0;
    ctx->rollbacks = 0;
    ctx->f64_tail = false;
    double logbase = 0.0;
    ctx->delta_t = find_timestep(ctx);
    build_neighbours(ctx, reorder_due(ctx));
    prepasses(ctx);
    const double le0 = nb_iter(ctx);
    logbase = le0; /* Check_Error at iteration 0 */
    ctx->iteration++;
    build_neighbours(ctx, false);
    prepasses(ctx);
    const double rms = newmark_loop(ctx, logbase);
    HIPCHK(hipMemcpyAsync(&ctx->sc_host->npd, &ctx->sc->npd, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    ctx->last_npd = ctx->sc_host->npd;
    drain_timers(ctx);
    return rms;
}

void set_count(fjsph_ctx* ctx, int64_t n)
{
    ctx->n = n;
    ctx->nparts = reduction_parts(n);
    ctx->have_nbrs = false;
}

/* update_buffer_region (src/shapes/inlet.cpp:578-640): back particles past
 * the insertion plane become PIPE, their columns move up, and each such
 * column gets a new BUFFER particle behind its last one while the count is
 * under max_points.  New particles go to the end of their block in host
 * order, in column order (the reference's insert at index.second). */
void inlet_update(fjsph_ctx* ctx)
{
    const int D = ctx->s.dim;
    const int32_t nc = ctx->ncols;
    upload_planes(ctx);
    HIPCHK(hipMemsetAsync(ctx->c_last, 0xFF, sizeof(int32_t) * nc, ctx->st));
    HIPCHK(hipMemsetAsync(ctx->c_adv, 0, sizeof(int32_t) * nc, ctx->st));
    launch_inlet_back(D, ctx->pnp1, ctx->col, ctx->depth, ctx->blk, ctx->dpl, ctx->n, ctx->c_adv, ctx->c_last, ctx->st);
    launch_inlet_shift(ctx->pnp1, ctx->col, ctx->depth, ctx->blk, ctx->dpl, ctx->n, ctx->c_adv, ctx->st);
    std::vector<int32_t> adv(nc), last(nc), rank(nc, -1);
    HIPCHK(hipMemcpyAsync(adv.data(), ctx->c_adv, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(last.data(), ctx->c_last, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    const int64_t n_old = ctx->n;
    int32_t nadd = 0;
    std::vector<int32_t> per_block(ctx->blocks.size(), 0);
    for (const auto& in : ctx->inlets) /* inlets are kept in block order */
        for (int32_t c = in.col_base; c < in.col_base + in.ncols; ++c) {
            if (!adv[c]) continue;
            if (last[c] >= 0 && n_old + nadd < ctx->max_points) {
                rank[c] = nadd++;
                per_block[size_t(in.block)]++;
                ctx->starved[size_t(c)] = 0;
            } else if (++ctx->starved[size_t(c)] >= in.nbuf) {
                /* the reference's buffer list would now hold the back
                 * particle itself: a state this engine does not mirror */
                throw std::string("inlet column ") + std::to_string(c - in.col_base) + " of block " +
                    std::to_string(in.block) + " advanced " + std::to_string(in.nbuf) +
                    " times without a new particle (max_points reached): its buffer is used up";
            }
        }
    if (nadd > 0) {
        if (n_old + nadd > ctx->cap) alloc_slots(ctx, n_old + nadd + std::max<int64_t>(n_old / 8, 1024), n_old);
        /* new slots start from zero (SPHPart(x, pj, BUFFER, id)) */
        for (int f = 0; f < kNumFields; ++f) {
            const size_t w = size_t(width_of(ctx, f)), es = field_desc(f).is_int ? 4 : 8;
            char* base = (char*)*field_slot(ctx->pnp1, f);
            HIPCHK(hipMemsetAsync(base + size_t(n_old) * w * es, 0, size_t(nadd) * w * es, ctx->st));
        }
        HIPCHK(hipMemsetAsync(ctx->near_inlet + n_old, 0, sizeof(int32_t) * nadd, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->c_rank, rank.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice, ctx->st));
        /* particle ids svar.part_id++ in insertion order (src/shapes/inlet.cpp:614-632) */
        if (ctx->next_part_id + nadd >= (int64_t(1) << 31)) throw std::string("particle ids past 2^31");
        launch_inlet_insert(D, ctx->pnp1, ctx->blk, ctx->col, ctx->depth, ctx->c_block, ctx->c_rank, ctx->c_last,
                            ctx->dpl, nc, n_old, ctx->d.dx, ctx->next_part_id, ctx->st);
        ctx->next_part_id += nadd;
        /* host ids: block by block in host order */
        int32_t rank_base = 0;
        for (size_t k = 0; k < ctx->blocks.size(); ++k) {
            const int32_t a = per_block[k];
            if (a == 0) continue;
            const int32_t at = int32_t(ctx->blocks[k].b.last);
            launch_shift_ids(ctx->part_id, n_old + rank_base, at, a, ctx->st);
            launch_new_ids(ctx->part_id, ctx->c_block, ctx->c_rank, nc, int32_t(k), n_old, rank_base, at, ctx->st);
            ctx->blocks[k].b.last += a;
            for (size_t j = k + 1; j < ctx->blocks.size(); ++j) {
                ctx->blocks[j].b.first += a;
                ctx->blocks[j].b.last += a;
            }
            rank_base += a;
        }
        set_count(ctx, n_old + nadd);
    }
    launch_inlet_backslot(ctx->col, ctx->depth, ctx->n, ctx->c_back, ctx->st);
}

/* Deletion plane of update_data (src/Integration.cpp:126-205): fluid
 * particles past their block's active plane are removed, the rest keep
 * their order (slots and host ids are compacted). */
/* MESH::maxlength as the TAU readers set it (Place_Edges, Place_Faces,
 * src/CDFIO.cpp:880-929, 1130-1220): the longest edge of the 2D edges, or
 * of the triangles and the longer diagonal of the other faces in 3D, times
 * 4 (2D) or 5 (3D).  Computed for every mesh here; the reference sets it
 * only when it reads a TAU mesh. */
double mesh_max_length(int D, int64_t nfaces, const int64_t* face_off, const int64_t* face_verts, const double* verts)
{
    auto dist = [&](int64_t a, int64_t b) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) {
            const double d = verts[a * D + k] - verts[b * D + k];
            s += d * d;
        }
        return std::sqrt(s);
    };
    double mx = 0.0;
    for (int64_t f = 0; f < nfaces; ++f) {
        const int64_t* fv = face_verts + face_off[f];
        const int64_t nv = face_off[f + 1] - face_off[f];
        double e = 0.0;
        if (D == 2) e = dist(fv[0], fv[1]);
        else if (nv == 3) e = std::max(dist(fv[0], fv[1]), std::max(dist(fv[0], fv[2]), dist(fv[1], fv[2])));
        else if (nv >= 4) e = std::max(dist(fv[0], fv[2]), dist(fv[1], fv[3]));
        mx = std::max(mx, e);
    }
    return mx * (D == 2 ? 4.0 : 5.0);
}

IptParams ipt_params(const fjsph_ctx* ctx)
{
    const fjsph_settings& s = ctx->s;
    IptParams p{};
    for (int a = 0; a < 3; ++a) p.grav[a] = s.grav[a];
    p.mu_g = s.mu_g;
    p.rho_l = s.rho_rest;
    p.relax = ctx->ipt.relax;
    p.max_x = ctx->ipt.max_x;
    p.max_len = ctx->mesh_max_len;
    /* Set_Values (src/IO.cpp:126-127) */
    p.diam = std::pow((6.0 * ctx->d.sim_mass) / (3.14159265358979323846 * s.rho_rest), 1.0 / 3.0);
    p.area = 3.14159265358979323846 * p.diam * p.diam / 4.0;
    p.eq_order = ctx->ipt.eq_order;
    p.n_relax = ctx->ipt.n_relax;
    p.max_subits = s.max_subits;
    p.max_steps = ctx->ipt.max_steps;
    return p;
}

void ipt_reserve(fjsph_ctx* ctx, int64_t n)
{
    if (n <= ctx->ipt_cap) return;
    const int64_t c = std::max<int64_t>(n, 2 * ctx->ipt_cap);
    dfree(ctx->ipt_in);
    dfree(ctx->ipt_out);
    dfree(ctx->ipt_slots);
    dalloc(ctx->ipt_in, size_t(c));
    dalloc(ctx->ipt_out, size_t(c));
    dalloc(ctx->ipt_slots, size_t(c));
    ctx->ipt_cap = c;
}

/* update_data's hand-off (src/Integration.cpp:150-168): the particles the
 * deletion planes remove, in ascending host index (the reference's
 * std::set), become IPT particles at the current time and are tracked. */
void ipt_handoff(fjsph_ctx* ctx, int64_t ndel)
{
    const int64_t n = ctx->n;
    std::vector<int32_t> keep(static_cast<size_t>(n)), host(static_cast<size_t>(n));
    HIPCHK(hipMemcpyAsync(keep.data(), ctx->keep, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(host.data(), ctx->part_id, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    std::vector<std::pair<int32_t, int32_t>> del;
    del.reserve(size_t(ndel));
    for (int64_t sl = 0; sl < n; ++sl)
        if (!keep[size_t(sl)]) del.emplace_back(host[size_t(sl)], int32_t(sl));
    std::sort(del.begin(), del.end());
    std::vector<int32_t> slots(del.size());
    for (size_t k = 0; k < del.size(); ++k) slots[k] = del[k].second;
    const int64_t m = int64_t(slots.size());
    if (m == 0) return;
    ipt_reserve(ctx, m);
    HIPCHK(hipMemcpyAsync(ctx->ipt_slots, slots.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, ctx->st));
    launch_ipt_gather(ctx->s.dim, ctx->pnp1, ctx->ipt_slots, m, ctx->current_time, ctx->ipt_in, ctx->st);
    launch_ipt_track(ctx->s.dim, ctx->dm, ipt_params(ctx), m, ctx->ipt_in, ctx->ipt_out, ctx->st);
    const size_t at = ctx->ipt_records.size();
    ctx->ipt_records.resize(at + size_t(m));
    HIPCHK(hipMemcpyAsync(ctx->ipt_records.data() + at, ctx->ipt_out, sizeof(fjsph_ipt_record) * m,
                          hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    for (size_t k = at; k < ctx->ipt_records.size(); ++k) {
        if (ctx->ipt_records[k].status == FJSPH_IPT_SUCCESS) ctx->ipt_success++;
        else if (ctx->ipt_records[k].status == FJSPH_IPT_FAILED) ctx->ipt_failed++;
    }
}

void delete_update(fjsph_ctx* ctx)
{
    const int D = ctx->s.dim;
    const int64_t n = ctx->n;
    const size_t nbk = ctx->blocks.size();
    upload_planes(ctx);
    zero_blkcount(ctx);
    launch_delete_flag(D, ctx->pnp1, ctx->part_id, ctx->blk, ctx->dpl, n, ctx->nb, ctx->keep, ctx->hflag,
                       ctx->blkcount, ctx->st);
    std::vector<int32_t> cnt(std::max<size_t>(nbk, 1));
    HIPCHK(hipMemcpyAsync(cnt.data(), ctx->blkcount, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost, ctx->st));
    sync(ctx);
    int64_t ndel = 0;
    for (size_t k = 0; k < nbk; ++k) ndel += cnt[k];
    if (ndel == 0) return;
    if (ctx->ipt.using_ipt && ctx->s.aero_source == FJSPH_MESH_INFL && ctx->have_mesh) ipt_handoff(ctx, ndel);
    compact_deleted(ctx, cnt, ndel);
}

/* Removes the slots with keep[slot] == 0 (hflag[host id] == 1), cnt[k] of
 * them in block k: both states are compacted in order, host ids close up,
 * block ranges shift. */
void compact_deleted(fjsph_ctx* ctx, const std::vector<int32_t>& cnt, int64_t ndel)
{
    if (grouped(ctx))
        throw std::string("independent simulations (groups): a particle deletion would change the groups' slot "
                          "ranges, which is not supported");
    const int64_t n = ctx->n;
    launch_scan(ctx->keep, n, ctx->kstart, ctx->del_tiles, ctx->st);
    launch_scan(ctx->hflag, n, ctx->hstart, ctx->del_tiles, ctx->st);
    launch_compact_perm(ctx->keep, ctx->kstart, n, ctx->perm, ctx->st);
    launch_compact_ids(ctx->part_id, ctx->hstart, n, ctx->st);
    permute_all(ctx, ctx->perm, n - ndel);
    int64_t shift = 0;
    for (size_t k = 0; k < ctx->blocks.size(); ++k) {
        ctx->blocks[k].b.first -= shift;
        shift += cnt[k];
        ctx->blocks[k].b.last -= shift;
    }
    set_count(ctx, n - ndel);
    if (ctx->ncols > 0) launch_inlet_backslot(ctx->col, ctx->depth, ctx->n, ctx->c_back, ctx->st);
}

/* update_data (src/Integration.cpp:109-226): density floor, inlets,
 * deletion, pn = pnp1. */
void update_data(fjsph_ctx* ctx)
{
    launch_fix_rho(ctx->pnp1, ctx->part_id, ctx->nb, ctx->n, ctx->s.rho_rest, ctx->st);
    if (ctx->ncols > 0) inlet_update(ctx);
    if (ctx->planes_on) delete_update(ctx);
    copy_state(ctx, ctx->pn, ctx->pnp1);
}

/* The CFL control of integrate (src/Integration.cpp:270-300) for one
 * simulation's state st (the context or a group). */
template <typename S>
void cfl_control(const fjsph_settings& s, S& st, double step_error)
{
    if (step_error > s.min_residual || st.maxRho_pc > s.rho_max_iter) {
        if (step_error > 0.6 * s.min_residual) {
            st.cfl = std::max(s.cfl_min, st.cfl - s.cfl_step);
            st.n_unstable = 0;
        } else if (st.n_unstable > unsigned(s.n_unstable_limit)) {
            st.cfl = std::max(s.cfl_min, st.cfl - s.cfl_step);
            st.n_unstable = 0;
        } else
            st.n_unstable++;
    } else
        st.n_unstable = 0;
    if (st.iteration < s.subits_factor * s.max_subits && st.n_unstable == 0) {
        if (st.n_stable > unsigned(s.n_stable_limit)) {
            st.cfl = std::min(s.cfl_max, st.cfl + s.cfl_step);
            st.n_stable = 0;
        } else
            st.n_stable++;
    } else
        st.n_stable = 0;
}

double integrate_groups(fjsph_ctx* ctx);

/* integrate (src/Integration.cpp:233-303). */
double integrate(fjsph_ctx* ctx)
{
    if (grouped(ctx)) return integrate_groups(ctx);
    const double step_error = integrate_no_update(ctx);
    update_data(ctx);
    ctx->current_time += ctx->delta_t;
    cfl_control(ctx->s, *ctx, step_error);
    sync(ctx);
    return step_error;
}

/* integrate for independent simulations: every group takes one step with
 * its own time step, CFL control and clock. */
double integrate_groups(fjsph_ctx* ctx)
{
    const double worst = integrate_no_update(ctx);
    update_data(ctx);
    for (fjsph_ctx::Group& g : ctx->grs) {
        g.time += g.dt;
        cfl_control(ctx->s, g, g.rms);
    }
    sync(ctx);
    return worst;
}

void fill_info(fjsph_ctx* ctx, double rms, fjsph_step_info* info)
{
    if (!info) return;
    if (grouped(ctx)) {
        /* the batch: the largest residual, iteration count, density
         * variation and accelerations, the smallest time step, CFL number
         * and clock, the sum of the rollbacks, the mean npd (each group's
         * own values: fjsph_get_group_info) */
        *info = fjsph_step_info{};
        info->rms_error = rms;
        info->dt = info->cfl = info->time = 1e300;
        double npd = 0.0;
        for (const fjsph_ctx::Group& g : ctx->grs) {
            info->dt = std::min(info->dt, g.dt);
            info->cfl = std::min(info->cfl, g.cfl);
            info->time = std::min(info->time, g.time);
            info->max_rho_pc = std::max(info->max_rho_pc, g.maxRho_pc);
            info->maxf = std::max(info->maxf, g.maxf);
            info->maxAf = std::max(info->maxAf, g.maxAf);
            info->iterations = std::max(info->iterations, int32_t(g.iteration));
            info->rollbacks += g.rollbacks;
            npd += g.npd;
        }
        info->npd = npd / double(ctx->grs.size());
        return;
    }
    info->rms_error = rms;
    info->dt = ctx->delta_t;
    info->cfl = ctx->cfl;
    info->time = ctx->current_time;
    info->max_rho_pc = ctx->maxRho_pc;
    info->maxf = ctx->maxf;
    info->maxAf = ctx->maxAf;
    info->npd = ctx->last_npd;
    info->iterations = int32_t(ctx->iteration);
    info->rollbacks = ctx->rollbacks;
}

template <typename F>
fjsph_status guarded(fjsph_ctx* ctx, F&& f)
{
    if (!ctx) return FJSPH_ERR_ARG;
    try {
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) throw HipError{e, "hipSetDevice"};
        f();
        return FJSPH_OK;
    } catch (const HipError& he) {
        set_err(ctx, "%s failed: %s", he.what, hipGetErrorString(he.e));
        return he.e == hipErrorOutOfMemory ? FJSPH_ERR_OOM : FJSPH_ERR_HIP;
    } catch (const std::string& msg) {
        ctx->err = msg;
        return FJSPH_ERR_STATE;
    } catch (const std::exception& ex) {
        ctx->err = ex.what();
        return FJSPH_ERR_STATE;
    }
}

std::vector<int32_t>& slot_to_host(fjsph_ctx* ctx)
{
    ctx->slot_host.resize(size_t(ctx->n));
    HIPCHK(hipMemcpyAsync(ctx->slot_host.data(), ctx->part_id, sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost,
                          ctx->st));
    sync(ctx);
    /* the host arrays are indexed by it: an inconsistent map must not write
     * out of their bounds */
    for (int32_t h : ctx->slot_host)
        if (h < 0 || h >= ctx->n) throw std::string("internal: slot to host map out of range");
    return ctx->slot_host;
}

}  // namespace

/* =========================================================== C entry points */

extern "C" {

int32_t fjsph_abi_version(void) { return FJSPH_ABI_VERSION; }

fjsph_status fjsph_create(const fjsph_settings* s, int32_t device, fjsph_ctx** out)
{
    if (!s || !out) return FJSPH_ERR_ARG;
    *out = nullptr;
    fjsph_derived d;
    if (fjsph_derive(s, &d) != FJSPH_OK) return FJSPH_ERR_ARG;
    if (s->precision != FJSPH_PREC_FP64 && s->precision != FJSPH_PREC_MIXED) return FJSPH_ERR_ARG;
    if (s->aero_case != FJSPH_NO_AERO && s->aero_case != FJSPH_GISSLER && s->aero_case != FJSPH_INDUCED_PRESSURE &&
        s->aero_case != FJSPH_SKIN_FRIC)
        return FJSPH_ERR_ARG;
    if (s->aero_source != FJSPH_CONST_VEL && s->aero_source != FJSPH_MESH_INFL && s->aero_source != FJSPH_SURROGATE)
        return FJSPH_ERR_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FJSPH_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return FJSPH_ERR_ARG;
    fjsph_ctx* ctx = new fjsph_ctx();
    fjsph_ipt_settings_default(&ctx->ipt);
    ctx->s = *s;
    ctx->d = d;
    ctx->device = device;
    ctx->cfl = s->cfl;
    fjsph_status rc = guarded(ctx, [&] {
        HIPCHK(hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking));
        HIPCHK(hipMalloc((void**)&ctx->sc, sizeof(DevScalars)));
        HIPCHK(hipMemsetAsync(ctx->sc, 0, sizeof(DevScalars), ctx->st));
        HIPCHK(hipHostMalloc((void**)&ctx->sc_host, sizeof(DevScalars), hipHostMallocDefault));
        /* this device's address of the scratch array of the entry points
         * that allocate nothing (once per device, the device current) */
        HIPCHK(noalloc_init(device));
        fill_consts(ctx);
    });
    if (rc != FJSPH_OK) {
        fjsph_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return FJSPH_OK;
}

void fjsph_destroy(fjsph_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    free_all(ctx);
    for (auto& r : ctx->pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->sc) (void)hipFree(ctx->sc);
    if (ctx->sc_host) (void)hipHostFree(ctx->sc_host);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
}

const char* fjsph_last_error(const fjsph_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

fjsph_status fjsph_upload_state(fjsph_ctx* ctx, int64_t n_total, int64_t n_bound, const double* x, const double* v,
                                const double* rho, const double* p, const double* m, const int32_t* b)
{
    if (!ctx) return FJSPH_ERR_ARG;
    if (n_total <= 0 || n_bound < 0 || n_bound >= n_total || n_total >= (int64_t(1) << 31) || !x || !v || !rho ||
        !p || !m || !b) {
        ctx->err = "upload_state: bad sizes or null arrays (need 0 <= n_bound < n_total < 2^31)";
        return FJSPH_ERR_ARG;
    }
    return guarded(ctx, [&] {
        free_all(ctx);
        ctx->partial_cap = 0;
        ctx->n = n_total;
        ctx->nb = n_bound;
        const int D = ctx->s.dim;
        alloc_slots(ctx, n_total, 0);
        ctx->nparts = reduction_parts(n_total);
        ctx->nl.f32 = ctx->s.precision == FJSPH_PREC_MIXED ? 1 : 0;
        sync(ctx); /* the blocking copies below run on the null stream */
        /* host-side staging of the initial state (src/Init.cpp:288-493,
         * src/FJSPH.cpp:115-126) */
        std::vector<int32_t> ids(n_total), blk(n_total, -1);
        for (int64_t i = 0; i < n_total; ++i) ids[i] = int32_t(i);
        for (size_t k = 0; k < ctx->blocks.size(); ++k)
            for (int64_t i = std::max<int64_t>(0, ctx->blocks[k].b.first);
                 i < std::min<int64_t>(n_total, ctx->blocks[k].b.last); ++i)
                blk[i] = int32_t(k);
        std::vector<double> cellV(size_t(n_total) * D), cellP(n_total, ctx->s.p_ref), cellRho(n_total, ctx->s.rho_g);
        for (int64_t i = 0; i < n_total; ++i)
            for (int k = 0; k < D; ++k) cellV[i * D + k] = ctx->s.v_inf[k];
        std::vector<int32_t> cellID(n_total, -3);
        for (DevState* st : {&ctx->pn, &ctx->pnp1}) {
            HIPCHK(hipMemcpy(st->x, x, sizeof(double) * n_total * D, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(st->v, v, sizeof(double) * n_total * D, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(st->rho, rho, sizeof(double) * n_total, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(st->p, p, sizeof(double) * n_total, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(st->m, m, sizeof(double) * n_total, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(st->b, b, sizeof(int32_t) * n_total, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(st->cellID, cellID.data(), sizeof(int32_t) * n_total, hipMemcpyHostToDevice));
            if (ctx->s.aero_source != FJSPH_MESH_INFL) {
                HIPCHK(hipMemcpy(st->cellV, cellV.data(), sizeof(double) * n_total * D, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(st->cellP, cellP.data(), sizeof(double) * n_total, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(st->cellRho, cellRho.data(), sizeof(double) * n_total, hipMemcpyHostToDevice));
            }
        }
        HIPCHK(hipMemcpy(ctx->part_id, ids.data(), sizeof(int32_t) * n_total, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->pn.pid, ids.data(), sizeof(int32_t) * n_total, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->pnp1.pid, ids.data(), sizeof(int32_t) * n_total, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->blk, blk.data(), sizeof(int32_t) * n_total, hipMemcpyHostToDevice));
        ctx->c.n_bound = n_bound;
        ctx->have_state = true;
        ctx->have_nbrs = false;
        ctx->current_time = 0.0;
        ctx->last_frame_time = 0.0;
        ctx->cfl = ctx->s.cfl;
        ctx->delta_t = ctx->s.delta_t_min > 0 ? ctx->s.delta_t_min : 2e-10;
        ctx->n_stable = ctx->n_unstable = 0;
        ctx->current_frame = 0;
        ctx->steps_to_reorder = 0;
        ctx->next_part_id = n_total; /* svar.part_id after Init_Particles (src/Init.cpp:478) */
        /* blocks may have been set before the state: push them */
        if (!ctx->blocks.empty()) {
            const int nbk = int(ctx->blocks.size());
            std::vector<int32_t> solver(nbk), ns(nbk);
            std::vector<double> vel(size_t(nbk) * 3);
            for (int k = 0; k < nbk; ++k) {
                solver[k] = ctx->blocks[k].b.bound_solver;
                ns[k] = ctx->blocks[k].b.no_slip;
                for (int a = 0; a < 3; ++a) vel[k * 3 + a] = ctx->blocks[k].b.vel[a];
            }
            dalloc(ctx->bl.solver, nbk);
            dalloc(ctx->bl.no_slip, nbk);
            dalloc(ctx->bl.vel, size_t(nbk) * 3);
            HIPCHK(hipMemcpy(ctx->bl.solver, solver.data(), sizeof(int32_t) * nbk, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(ctx->bl.no_slip, ns.data(), sizeof(int32_t) * nbk, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(ctx->bl.vel, vel.data(), sizeof(double) * nbk * 3, hipMemcpyHostToDevice));
            ctx->bl.n = nbk;
        }
        sync(ctx);
    });
}

fjsph_status fjsph_set_blocks(fjsph_ctx* ctx, int32_t n_blocks, const fjsph_block* blocks)
{
    if (!ctx || n_blocks < 0 || (n_blocks > 0 && !blocks)) return FJSPH_ERR_ARG;
    for (int k = 0; k < n_blocks; ++k) {
        const fjsph_block& b = blocks[k];
        if (b.first < 0 || b.last < b.first || (b.bound_solver < 0 || b.bound_solver > 2)) {
            ctx->err = "set_blocks: bad block range or solver";
            return FJSPH_ERR_ARG;
        }
    }
    ctx->blocks.clear();
    ctx->planes.assign(size_t(n_blocks), default_planes());
    ctx->planes_on = false;
    ctx->planes_dirty = true;
    for (int k = 0; k < n_blocks; ++k) {
        HostBlock hb;
        hb.b = blocks[k];
        hb.b.n_times = 0;
        ctx->blocks.push_back(hb);
    }
    if (!ctx->have_state) return FJSPH_OK;
    /* re-derive per-slot block ids in the current slot order */
    return guarded(ctx, [&] {
        std::vector<int32_t>& sh = slot_to_host(ctx);
        std::vector<int32_t> blk(size_t(ctx->n), -1);
        for (int64_t s = 0; s < ctx->n; ++s) {
            const int64_t h = sh[s];
            for (int k = 0; k < n_blocks; ++k)
                if (h >= blocks[k].first && h < blocks[k].last) { blk[s] = k; break; }
        }
        HIPCHK(hipMemcpy(ctx->blk, blk.data(), sizeof(int32_t) * ctx->n, hipMemcpyHostToDevice));
        const int nbk = n_blocks;
        std::vector<int32_t> solver(std::max(nbk, 1)), ns(std::max(nbk, 1));
        std::vector<double> vel(size_t(std::max(nbk, 1)) * 3, 0.0);
        for (int k = 0; k < nbk; ++k) {
            solver[k] = blocks[k].bound_solver;
            ns[k] = blocks[k].no_slip;
            for (int a = 0; a < 3; ++a) vel[k * 3 + a] = blocks[k].vel[a];
        }
        dalloc(ctx->bl.solver, solver.size());
        dalloc(ctx->bl.no_slip, ns.size());
        dalloc(ctx->bl.vel, vel.size());
        HIPCHK(hipMemcpy(ctx->bl.solver, solver.data(), sizeof(int32_t) * solver.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->bl.no_slip, ns.data(), sizeof(int32_t) * ns.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->bl.vel, vel.data(), sizeof(double) * vel.size(), hipMemcpyHostToDevice));
        ctx->bl.n = nbk;
    });
}

fjsph_status fjsph_set_block_planes(fjsph_ctx* ctx, int32_t block, const fjsph_block_planes* planes)
{
    if (!ctx || !planes || block < 0 || block >= int32_t(ctx->blocks.size())) return FJSPH_ERR_ARG;
    if (planes->fixed_vel_or_dynamic < 0 || planes->fixed_vel_or_dynamic > 1) return FJSPH_ERR_ARG;
    if (grouped(ctx)) {
        ctx->err = "set_block_planes: not supported with independent simulations (groups)";
        return FJSPH_ERR_STATE;
    }
    if (ctx->planes.size() < ctx->blocks.size()) ctx->planes.resize(ctx->blocks.size(), default_planes());
    ctx->planes[size_t(block)] = *planes;
    ctx->planes_on = true;
    ctx->planes_dirty = true;
    return FJSPH_OK;
}

fjsph_status fjsph_set_inlet(fjsph_ctx* ctx, int32_t block, int32_t n_cols, int32_t n_buf, const int64_t* back,
                             const int64_t* buffer)
{
    if (!ctx || block < 0 || block >= int32_t(ctx->blocks.size()) || n_cols <= 0 || n_buf <= 0 || !back || !buffer)
        return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        require(ctx, false);
        if (grouped(ctx)) throw std::string("set_inlet: not supported with independent simulations (groups)");
        const fjsph_block& bk = ctx->blocks[size_t(block)].b;
        if (!bk.is_fluid) throw std::string("set_inlet: not a fluid block");
        for (const auto& in : ctx->inlets)
            if (in.block == block) throw std::string("set_inlet: block already has an inlet");
        for (int64_t k = 0; k < int64_t(n_cols) * (1 + n_buf); ++k) {
            const int64_t h = k < n_cols ? back[k] : buffer[k - n_cols];
            if (h < bk.first || h >= bk.last) throw std::string("set_inlet: particle outside the block");
        }
        /* host id -> slot */
        std::vector<int32_t>& sh = slot_to_host(ctx);
        std::vector<int32_t> slot_of(size_t(ctx->n));
        for (int64_t sl = 0; sl < ctx->n; ++sl) slot_of[size_t(sh[size_t(sl)])] = int32_t(sl);
        std::vector<int32_t> col(size_t(ctx->n)), depth(size_t(ctx->n));
        HIPCHK(hipMemcpy(col.data(), ctx->col, sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(depth.data(), ctx->depth, sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost));
        const int32_t base = ctx->ncols;
        for (int32_t c = 0; c < n_cols; ++c) {
            const int32_t sb = slot_of[size_t(back[c])];
            col[size_t(sb)] = base + c;
            depth[size_t(sb)] = 0;
            for (int32_t j = 0; j < n_buf; ++j) {
                const int32_t sl = slot_of[size_t(buffer[size_t(c) * n_buf + j])];
                col[size_t(sl)] = base + c;
                depth[size_t(sl)] = j + 1;
            }
        }
        HIPCHK(hipMemcpy(ctx->col, col.data(), sizeof(int32_t) * ctx->n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->depth, depth.data(), sizeof(int32_t) * ctx->n, hipMemcpyHostToDevice));
        ctx->inlets.push_back({block, base, n_cols, n_buf});
        std::sort(ctx->inlets.begin(), ctx->inlets.end(),
                  [](const fjsph_ctx::InletCols& a, const fjsph_ctx::InletCols& b) { return a.block < b.block; });
        ctx->ncols = base + n_cols;
        ctx->starved.resize(size_t(ctx->ncols), 0);
        std::vector<int32_t> cblock(size_t(ctx->ncols));
        for (const auto& in : ctx->inlets)
            for (int32_t c = in.col_base; c < in.col_base + in.ncols; ++c) cblock[size_t(c)] = in.block;
        for (int32_t** a : {&ctx->c_adv, &ctx->c_back, &ctx->c_last, &ctx->c_rank, &ctx->c_block})
            dalloc(*a, size_t(ctx->ncols));
        HIPCHK(hipMemcpy(ctx->c_block, cblock.data(), sizeof(int32_t) * ctx->ncols, hipMemcpyHostToDevice));
        if (ctx->planes.size() < ctx->blocks.size()) ctx->planes.resize(ctx->blocks.size(), default_planes());
        ctx->planes[size_t(block)].inlet = 1;
        ctx->planes_on = true;
        ctx->planes_dirty = true;
        launch_inlet_backslot(ctx->col, ctx->depth, ctx->n, ctx->c_back, ctx->st);
        sync(ctx);
    });
}

fjsph_status fjsph_set_max_points(fjsph_ctx* ctx, int64_t max_points)
{
    if (!ctx || max_points <= 0) return FJSPH_ERR_ARG;
    ctx->max_points = max_points;
    return FJSPH_OK;
}

fjsph_status fjsph_get_max_points(fjsph_ctx* ctx, int64_t* max_points)
{
    if (!ctx || !max_points) return FJSPH_ERR_ARG;
    *max_points = ctx->max_points;
    return FJSPH_OK;
}

fjsph_status fjsph_get_counts(fjsph_ctx* ctx, int64_t* n_total, int64_t* n_bound)
{
    if (!ctx) return FJSPH_ERR_ARG;
    if (n_total) *n_total = ctx->n;
    if (n_bound) *n_bound = ctx->nb;
    return FJSPH_OK;
}

fjsph_status fjsph_get_block_range(fjsph_ctx* ctx, int32_t block, int64_t* first, int64_t* last)
{
    if (!ctx || block < 0 || block >= int32_t(ctx->blocks.size())) return FJSPH_ERR_ARG;
    if (first) *first = ctx->blocks[size_t(block)].b.first;
    if (last) *last = ctx->blocks[size_t(block)].b.last;
    return FJSPH_OK;
}

fjsph_status fjsph_get_blocks(fjsph_ctx* ctx, int32_t* n_blocks, fjsph_block* blocks, int32_t capacity)
{
    if (!ctx || !n_blocks) return FJSPH_ERR_ARG;
    *n_blocks = int32_t(ctx->blocks.size());
    if (!blocks) return FJSPH_OK;
    if (capacity < *n_blocks) return FJSPH_ERR_ARG;
    for (size_t k = 0; k < ctx->blocks.size(); ++k) blocks[k] = ctx->blocks[k].b;
    return FJSPH_OK;
}

fjsph_status fjsph_get_block_motion(fjsph_ctx* ctx, int32_t block, int32_t* n_times, double* times, double* vels)
{
    if (!ctx || !n_times || block < 0 || block >= int32_t(ctx->blocks.size())) return FJSPH_ERR_ARG;
    const HostBlock& hb = ctx->blocks[size_t(block)];
    *n_times = int32_t(hb.times.size());
    if (times) std::copy(hb.times.begin(), hb.times.end(), times);
    if (vels) std::copy(hb.vels.begin(), hb.vels.end(), vels);
    return FJSPH_OK;
}

fjsph_status fjsph_get_block_planes(fjsph_ctx* ctx, int32_t block, fjsph_block_planes* out)
{
    if (!ctx || !out || block < 0 || block >= int32_t(ctx->blocks.size())) return FJSPH_ERR_ARG;
    *out = size_t(block) < ctx->planes.size() ? ctx->planes[size_t(block)] : default_planes();
    return FJSPH_OK;
}

fjsph_status fjsph_get_inlet(fjsph_ctx* ctx, int32_t block, int32_t* n_cols, int32_t* n_buf, int64_t* back,
                             int64_t* buffer)
{
    if (!ctx || !n_cols || !n_buf || block < 0 || block >= int32_t(ctx->blocks.size())) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        *n_cols = 0;
        *n_buf = 0;
        const fjsph_ctx::InletCols* in = nullptr;
        for (const auto& c : ctx->inlets)
            if (c.block == block) in = &c;
        if (!in) return;
        *n_cols = in->ncols;
        *n_buf = in->nbuf;
        if (!back || !buffer) return;
        const int64_t n = ctx->n;
        const std::vector<int32_t>& sh = slot_to_host(ctx);
        std::vector<int32_t> col(static_cast<size_t>(n)), depth(static_cast<size_t>(n));
        HIPCHK(hipMemcpy(col.data(), ctx->col, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(depth.data(), ctx->depth, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        const int32_t nb = in->nbuf;
        std::vector<int64_t> bk(size_t(in->ncols), -1), bf(size_t(in->ncols) * nb, -1);
        for (int64_t sl = 0; sl < n; ++sl) {
            const int32_t c = col[size_t(sl)] - in->col_base;
            if (c < 0 || c >= in->ncols) continue;
            const int32_t d = depth[size_t(sl)];
            if (d == 0) bk[size_t(c)] = sh[size_t(sl)];
            else if (d > 0 && d <= nb) bf[size_t(c) * nb + (d - 1)] = sh[size_t(sl)];
        }
        for (int32_t c = 0; c < in->ncols; ++c) {
            const int64_t deepest = bf[size_t(c) * nb + (nb - 1)];
            if (bk[size_t(c)] < 0 || deepest < 0) throw std::string("get_inlet: a column lost its back or last particle");
            /* places vacated by advances without a new particle repeat the
             * deepest particle, as the reference's shifted list does */
            for (int32_t j = 0; j < nb; ++j)
                if (bf[size_t(c) * nb + j] < 0) bf[size_t(c) * nb + j] = deepest;
            back[c] = bk[size_t(c)];
            for (int32_t j = 0; j < nb; ++j) buffer[size_t(c) * nb + j] = bf[size_t(c) * nb + j];
        }
    });
}

void fjsph_ipt_settings_default(fjsph_ipt_settings* s)
{
    if (!s) return;
    s->using_ipt = 0;
    s->eq_order = 2;        /* IPT_SETT::ipt_eq_order (src/Var.h:316) */
    s->max_x = 9999999.0;   /* IPT_SETT::max_x (src/Var.h:318) */
    s->relax = 0.6;         /* IPT_SETT::relax (src/Var.h:323) */
    s->n_relax = 5;         /* IPT_SETT::n_relax (src/Var.h:324) */
    s->max_steps = 1000000;
}

fjsph_status fjsph_set_ipt(fjsph_ctx* ctx, const fjsph_ipt_settings* s)
{
    if (!ctx || !s || (s->eq_order != 1 && s->eq_order != 2) || s->max_steps <= 0 || s->n_relax < 0)
        return FJSPH_ERR_ARG;
    ctx->ipt = *s;
    return FJSPH_OK;
}

fjsph_status fjsph_ipt_track(fjsph_ctx* ctx, int64_t n, const double* x, const double* v, const int32_t* cell,
                             const int64_t* part_id, double t0, fjsph_ipt_record* out)
{
    if (!ctx || n < 0 || (n > 0 && (!x || !v || !cell || !out))) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        if (!ctx->have_mesh) throw std::string("ipt_track: no mesh set");
        if (n == 0) return;
        const int D = ctx->s.dim;
        std::vector<IptStart> st(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i) {
            IptStart& q = st[size_t(i)];
            std::memset(&q, 0, sizeof(q));
            for (int a = 0; a < D; ++a) {
                q.x[a] = x[i * D + a];
                q.v[a] = v[i * D + a];
            }
            q.cell = cell[i];
            if (q.cell >= 0 && q.cell < ctx->mesh_ncells) {
                for (int a = 0; a < D; ++a) q.cellV[a] = ctx->mvel[size_t(q.cell) * D + a];
                q.cellRho = ctx->mrho[size_t(q.cell)];
            }
            q.t = t0;
            q.part_id = part_id ? part_id[i] : i;
        }
        ipt_reserve(ctx, n);
        HIPCHK(hipMemcpyAsync(ctx->ipt_in, st.data(), sizeof(IptStart) * n, hipMemcpyHostToDevice, ctx->st));
        launch_ipt_track(D, ctx->dm, ipt_params(ctx), n, ctx->ipt_in, ctx->ipt_out, ctx->st);
        HIPCHK(hipMemcpyAsync(out, ctx->ipt_out, sizeof(fjsph_ipt_record) * n, hipMemcpyDeviceToHost, ctx->st));
        sync(ctx);
    });
}

fjsph_status fjsph_ipt_count(fjsph_ctx* ctx, int64_t* n_records, int64_t* n_success, int64_t* n_failed)
{
    if (!ctx) return FJSPH_ERR_ARG;
    if (n_records) *n_records = int64_t(ctx->ipt_records.size());
    if (n_success) *n_success = ctx->ipt_success;
    if (n_failed) *n_failed = ctx->ipt_failed;
    return FJSPH_OK;
}

fjsph_status fjsph_ipt_records(fjsph_ctx* ctx, int64_t first, int64_t count, fjsph_ipt_record* out)
{
    if (!ctx || first < 0 || count < 0 || first + count > int64_t(ctx->ipt_records.size()) || (count > 0 && !out))
        return FJSPH_ERR_ARG;
    std::copy(ctx->ipt_records.begin() + first, ctx->ipt_records.begin() + first + count, out);
    return FJSPH_OK;
}

fjsph_status fjsph_ipt_clear(fjsph_ctx* ctx)
{
    if (!ctx) return FJSPH_ERR_ARG;
    ctx->ipt_records.clear();
    ctx->ipt_success = ctx->ipt_failed = 0;
    return FJSPH_OK;
}

fjsph_status fjsph_set_groups(fjsph_ctx* ctx, int32_t n_groups, const int32_t* group)
{
    if (!ctx || n_groups < 0 || (n_groups > 0 && !group)) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        require(ctx, false);
        if (n_groups == 0) {
            ctx->ngroups = 0;
            ctx->grs.clear();
            dfree(ctx->grp);
            dfree(ctx->gd2);
            ctx->have_nbrs = false;
            return;
        }
        if (ctx->nb != 0)
            throw std::string("set_groups: boundary particles are not supported (each group is a body of fluid)");
        if (ctx->ncols > 0 || ctx->planes_on)
            throw std::string("set_groups: inlets and deletion planes are not supported with groups");
        if (ctx->s.integrator != FJSPH_NEWMARK_BETA)
            throw std::string("set_groups: independent simulations use the Newmark-Beta integrator");
        const int64_t n = ctx->n;
        const std::vector<int32_t> hostof = slot_to_host(ctx);
        std::vector<int32_t> gs(static_cast<size_t>(n));
        std::vector<int64_t> cnt(size_t(n_groups), 0);
        for (int64_t i = 0; i < n; ++i) {
            const int32_t g = group[hostof[i]];
            if (g < 0 || g >= n_groups) throw std::string("set_groups: group index out of range");
            gs[i] = g;
            cnt[g]++;
        }
        for (int32_t g = 0; g < n_groups; ++g)
            if (cnt[g] == 0) throw std::string("set_groups: every group needs at least one particle");
        /* group-major slots: a stable order by group */
        std::vector<int32_t> perm(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i) perm[i] = int32_t(i);
        std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return gs[a] < gs[b]; });
        bool identity = true;
        for (int64_t i = 0; i < n && identity; ++i) identity = perm[i] == i;
        if (!identity) {
            HIPCHK(hipMemcpy(ctx->perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
            permute_all(ctx, ctx->perm);
            std::vector<int32_t> sorted(static_cast<size_t>(n));
            for (int64_t i = 0; i < n; ++i) sorted[i] = gs[perm[i]];
            gs.swap(sorted);
        }
        const size_t G = size_t(n_groups);
        dalloc(ctx->grp, size_t(ctx->cap));
        dalloc(ctx->gd2, size_t(ctx->cap));
        ctx->gd2_cap = ctx->cap;
        dalloc(ctx->gstart, G + 1);
        dalloc(ctx->gblk, G + 1);
        dalloc(ctx->gbox, G * 6);
        dalloc(ctx->glo, G * 3);
        dalloc(ctx->gl0, G * 2);
        dalloc(ctx->gdt, G);
        dalloc(ctx->gnpd, G);
        dalloc(ctx->gscale, G);
        dalloc(ctx->gred, G * 8);
        dalloc(ctx->gres, G);
        dalloc(ctx->gactive, G);
        dalloc(ctx->gact_m, G);
        dalloc(ctx->gact_d, G);
        dalloc(ctx->gbuild, G);
        ctx->gstart_h.assign(G + 1, 0);
        std::vector<double> scale(G);
        for (size_t g = 0; g < G; ++g) {
            ctx->gstart_h[g + 1] = ctx->gstart_h[g] + cnt[g];
            scale[g] = 1.0 / double(cnt[g]);
        }
        sync(ctx);
        HIPCHK(hipMemcpy(ctx->grp, gs.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->gstart, ctx->gstart_h.data(), sizeof(int64_t) * (G + 1), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->gscale, scale.data(), sizeof(double) * G, hipMemcpyHostToDevice));
        fjsph_ctx::Group g0;
        g0.time = ctx->current_time;
        g0.last_frame_time = ctx->last_frame_time;
        g0.dt = ctx->delta_t;
        g0.cfl = ctx->cfl;
        g0.n_stable = ctx->n_stable;
        g0.n_unstable = ctx->n_unstable;
        ctx->grs.assign(G, g0);
        ctx->ngroups = n_groups;
        upload_active(ctx);
        upload_group_dt(ctx);
        ctx->have_nbrs = false;
    });
}

fjsph_status fjsph_get_group_info(fjsph_ctx* ctx, int32_t first, int32_t count, fjsph_step_info* out)
{
    if (!ctx || first < 0 || count < 0 || (count > 0 && !out)) return FJSPH_ERR_ARG;
    if (int64_t(first) + count > ctx->ngroups) {
        ctx->err = "get_group_info: group range out of bounds";
        return FJSPH_ERR_ARG;
    }
    for (int32_t k = 0; k < count; ++k) {
        const fjsph_ctx::Group& g = ctx->grs[size_t(first + k)];
        fjsph_step_info& o = out[k];
        o = fjsph_step_info{};
        o.rms_error = g.rms;
        o.dt = g.dt;
        o.cfl = g.cfl;
        o.time = g.time;
        o.max_rho_pc = g.maxRho_pc;
        o.maxf = g.maxf;
        o.maxAf = g.maxAf;
        o.npd = g.npd;
        o.iterations = int32_t(g.iteration);
        o.rollbacks = g.rollbacks;
    }
    return FJSPH_OK;
}

fjsph_status fjsph_set_block_targets(fjsph_ctx* ctx, int32_t targets)
{
    if (!ctx || targets < 0 || targets > 65536) return FJSPH_ERR_ARG;
    ctx->block_targets = targets;
    ctx->have_nbrs = false;
    return FJSPH_OK;
}

fjsph_status fjsph_set_reorder_interval(fjsph_ctx* ctx, int32_t steps)
{
    if (!ctx || steps < 1) return FJSPH_ERR_ARG;
    ctx->reorder_every = steps;
    ctx->steps_to_reorder = 0;
    return FJSPH_OK;
}

fjsph_status fjsph_get_block_targets(fjsph_ctx* ctx, int32_t* targets)
{
    if (!ctx || !targets) return FJSPH_ERR_ARG;
    *targets = ctx->block_targets > 0 ? ctx->block_targets : 256;
    return FJSPH_OK;
}

fjsph_status fjsph_autotune_blocks(fjsph_ctx* ctx, int32_t n_cand, const int32_t* targets, int32_t reps, double* ms,
                                   int32_t* best)
{
    if (!ctx || n_cand <= 0 || !targets || reps <= 0 || !ms || !best) return FJSPH_ERR_ARG;
    for (int32_t c = 0; c < n_cand; ++c)
        if (targets[c] <= 0 || targets[c] > 65536) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        require(ctx, false);
        const int64_t n = ctx->n;
        /* every per-slot array, so that the state comes back bit for bit.
         * The list builds permute the arrays by swapping buffers, so the
         * pointers are collected again, in the same order, to restore. */
        auto collect = [&] {
            std::vector<std::pair<void*, size_t>> v;
            for (DevState* st : {&ctx->pn, &ctx->pnp1})
                for (int f = 0; f < kNumFields; ++f)
                    v.emplace_back(*field_slot(*st, f),
                                   size_t(n) * size_t(width_of(ctx, f)) * (field_desc(f).is_int ? 4 : 8));
            for (int32_t* a : {ctx->part_id, ctx->blk, ctx->cell, ctx->col, ctx->depth, ctx->grp})
                if (a) v.emplace_back(a, size_t(n) * 4);
            return v;
        };
        std::vector<std::pair<void*, size_t>> arrays = collect();
        size_t total = 0;
        for (auto& a : arrays) total += (a.second + 255) / 256 * 256;
        char* snap = nullptr;
        HIPCHK(hipMalloc((void**)&snap, total));
        struct Free {
            char* p;
            ~Free() { if (p) (void)hipFree(p); }
        } guard{snap};
        size_t off = 0;
        for (auto& a : arrays) {
            HIPCHK(hipMemcpyAsync(snap + off, a.first, a.second, hipMemcpyDeviceToDevice, ctx->st));
            off += (a.second + 255) / 256 * 256;
        }
        const bool timing = ctx->timing;
        drain_timers(ctx);
        ctx->timing = false;
        ctx->tuning = true;
        const int32_t saved = ctx->block_targets;
        /* one step's pair work: two list builds with their pre-passes, and
         * the force pass once per Newmark iteration of the last step */
        const int iters = ctx->iteration > 0 ? int(ctx->iteration) : 10;
        if (grouped(ctx)) {
            for (fjsph_ctx::Group& g : ctx->grs) g.active = true;
            upload_active(ctx);
        }
        auto step_work = [&] {
            build_neighbours(ctx, true);
            prepasses(ctx);
            build_neighbours(ctx, false);
            prepasses(ctx);
            for (int i = 0; i < iters; ++i) force(ctx, false);
        };
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        double best_ms = 0.0;
        int32_t best_k = -1;
        try {
            for (int32_t c = 0; c < n_cand; ++c) {
                ctx->block_targets = targets[c];
                ctx->have_nbrs = false;
                step_work(); /* warm-up: list capacity for this tiling */
                HIPCHK(hipEventRecord(e0, ctx->st));
                for (int32_t r = 0; r < reps; ++r) step_work();
                HIPCHK(hipEventRecord(e1, ctx->st));
                HIPCHK(hipEventSynchronize(e1));
                float el = 0.0f;
                HIPCHK(hipEventElapsedTime(&el, e0, e1));
                ms[c] = double(el) / reps;
                if (best_k < 0 || ms[c] < best_ms) {
                    best_ms = ms[c];
                    best_k = c;
                }
            }
        } catch (...) {
            ctx->block_targets = saved;
            ctx->timing = timing;
            ctx->tuning = false;
            ctx->have_nbrs = false;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            arrays = collect();
            off = 0;
            for (auto& a : arrays) {
                (void)hipMemcpyAsync(a.first, snap + off, a.second, hipMemcpyDeviceToDevice, ctx->st);
                off += (a.second + 255) / 256 * 256;
            }
            /* the back slot of every inlet column follows the restored order */
            if (ctx->ncols > 0) launch_inlet_backslot(ctx->col, ctx->depth, n, ctx->c_back, ctx->st);
            (void)hipStreamSynchronize(ctx->st);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        arrays = collect();
        off = 0;
        for (auto& a : arrays) {
            HIPCHK(hipMemcpyAsync(a.first, snap + off, a.second, hipMemcpyDeviceToDevice, ctx->st));
            off += (a.second + 255) / 256 * 256;
        }
        /* the tuning builds moved the inlet columns' back slots with their
         * permutations: find them again in the restored order */
        if (ctx->ncols > 0) launch_inlet_backslot(ctx->col, ctx->depth, n, ctx->c_back, ctx->st);
        sync(ctx);
        ctx->timing = timing;
        ctx->tuning = false;
        ctx->block_targets = targets[best_k];
        ctx->have_nbrs = false;
        ctx->x_at_build = false;
        *best = targets[best_k];
    });
}

fjsph_status fjsph_get_clock(fjsph_ctx* ctx, fjsph_clock* out)
{
    if (!ctx || !out) return FJSPH_ERR_ARG;
    out->current_time = ctx->current_time;
    out->last_frame_time = ctx->last_frame_time;
    out->delta_t = ctx->delta_t;
    out->cfl = ctx->cfl;
    out->current_frame = ctx->current_frame;
    out->n_stable = ctx->n_stable;
    out->n_unstable = ctx->n_unstable;
    out->next_part_id = ctx->next_part_id;
    return FJSPH_OK;
}

fjsph_status fjsph_set_clock(fjsph_ctx* ctx, const fjsph_clock* in)
{
    if (!ctx || !in || in->n_stable < 0 || in->n_unstable < 0 || in->current_frame < 0 || in->next_part_id < 0 ||
        in->next_part_id >= (int64_t(1) << 31))
        return FJSPH_ERR_ARG;
    ctx->current_time = in->current_time;
    ctx->last_frame_time = in->last_frame_time;
    ctx->delta_t = in->delta_t;
    ctx->cfl = in->cfl;
    ctx->current_frame = in->current_frame;
    ctx->n_stable = unsigned(in->n_stable);
    ctx->n_unstable = unsigned(in->n_unstable);
    ctx->next_part_id = in->next_part_id;
    return FJSPH_OK;
}

fjsph_status fjsph_get_settings(fjsph_ctx* ctx, fjsph_settings* out)
{
    if (!ctx || !out) return FJSPH_ERR_ARG;
    *out = ctx->s;
    return FJSPH_OK;
}

fjsph_status fjsph_set_block_motion(fjsph_ctx* ctx, int32_t block, int32_t n_times, const double* times,
                                    const double* vels)
{
    if (!ctx || block < 0 || block >= int32_t(ctx->blocks.size()) || n_times < 0 || (n_times > 0 && (!times || !vels)))
        return FJSPH_ERR_ARG;
    HostBlock& hb = ctx->blocks[block];
    hb.b.n_times = n_times;
    hb.times.assign(times, times + n_times);
    hb.vels.assign(vels, vels + size_t(n_times) * 3);
    return FJSPH_OK;
}

fjsph_status fjsph_download_field(fjsph_ctx* ctx, int32_t which, int32_t field, double* out)
{
    if (!ctx || !out || field < 0 || field >= FJSPH_F_COUNT || (which != 0 && which != 1)) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        require(ctx, false);
        DevState& st = which == FJSPH_STATE_N ? ctx->pn : ctx->pnp1;
        const int w = width_of(ctx, field);
        const size_t cnt = size_t(ctx->n) * w;
        std::vector<int32_t>& sh = slot_to_host(ctx);
        if (field_desc(field).is_int) {
            std::vector<int32_t> buf(cnt);
            HIPCHK(hipMemcpy(buf.data(), *field_slot(st, field), cnt * 4, hipMemcpyDeviceToHost));
            for (int64_t s = 0; s < ctx->n; ++s) out[sh[s]] = double(buf[s]);
        } else {
            std::vector<double> buf(cnt);
            HIPCHK(hipMemcpy(buf.data(), *field_slot(st, field), cnt * 8, hipMemcpyDeviceToHost));
            for (int64_t s = 0; s < ctx->n; ++s)
                for (int k = 0; k < w; ++k) out[size_t(sh[s]) * w + k] = buf[size_t(s) * w + k];
        }
    });
}

fjsph_status fjsph_upload_field(fjsph_ctx* ctx, int32_t which, int32_t field, const double* in)
{
    if (!ctx || !in || field < 0 || field >= FJSPH_F_COUNT || (which != 0 && which != 1)) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        require(ctx, false);
        DevState& st = which == FJSPH_STATE_N ? ctx->pn : ctx->pnp1;
        if (which == FJSPH_STATE_NP1 && field == FJSPH_F_X) ctx->x_at_build = false;
        const int w = width_of(ctx, field);
        const size_t cnt = size_t(ctx->n) * w;
        std::vector<int32_t>& sh = slot_to_host(ctx);
        if (field_desc(field).is_int) {
            std::vector<int32_t> buf(cnt);
            for (int64_t s = 0; s < ctx->n; ++s) buf[s] = int32_t(in[sh[s]]);
            HIPCHK(hipMemcpy(*field_slot(st, field), buf.data(), cnt * 4, hipMemcpyHostToDevice));
        } else {
            std::vector<double> buf(cnt);
            for (int64_t s = 0; s < ctx->n; ++s)
                for (int k = 0; k < w; ++k) buf[size_t(s) * w + k] = in[size_t(sh[s]) * w + k];
            HIPCHK(hipMemcpy(*field_slot(st, field), buf.data(), cnt * 8, hipMemcpyHostToDevice));
        }
    });
}

fjsph_status fjsph_build_neighbours(fjsph_ctx* ctx)
{
    return guarded(ctx, [&] {
        require(ctx, false);
        build_neighbours(ctx, true);
        drain_timers(ctx);
    });
}

fjsph_status fjsph_get_neighbour_counts(fjsph_ctx* ctx, int32_t* counts)
{
    if (!counts) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        require(ctx, true);
        sync(ctx);
        std::vector<int32_t> c(static_cast<size_t>(ctx->n));
        HIPCHK(hipMemcpy(c.data(), ctx->nl.nnb, sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost));
        std::vector<int32_t>& sh = slot_to_host(ctx);
        for (int64_t s = 0; s < ctx->n; ++s) counts[sh[s]] = c[s] + 1;
    });
}

fjsph_status fjsph_get_neighbours(fjsph_ctx* ctx, int64_t* offsets, int64_t* idx, int64_t idx_capacity)
{
    if (!offsets || !idx) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        require(ctx, true);
        sync(ctx);
        const int64_t n = ctx->n;
        std::vector<int32_t> cnt(static_cast<size_t>(n));
        HIPCHK(hipMemcpy(cnt.data(), ctx->nl.nnb, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        std::vector<int32_t> ids(size_t(ctx->nl.cap) * ctx->nl.npad);
        if (!ids.empty()) {
            dalloc(ctx->export_buf, ids.size());
            launch_export_lists(ctx->s.dim, ctx->tg, ctx->nl, n, ctx->export_buf, ctx->st);
            sync(ctx);
            HIPCHK(hipMemcpy(ids.data(), ctx->export_buf, sizeof(int32_t) * ids.size(), hipMemcpyDeviceToHost));
            dfree(ctx->export_buf);
        }
        std::vector<int32_t>& sh = slot_to_host(ctx);
        std::vector<int64_t> hcount(static_cast<size_t>(n));
        for (int64_t s = 0; s < n; ++s) hcount[sh[s]] = cnt[s] + 1;
        offsets[0] = 0;
        for (int64_t h = 0; h < n; ++h) offsets[h + 1] = offsets[h] + hcount[h];
        if (offsets[n] > idx_capacity) throw std::string("get_neighbours: idx_capacity too small");
        for (int64_t s = 0; s < n; ++s) {
            const int64_t h = sh[s];
            int64_t* o = idx + offsets[h];
            o[0] = h;
            for (int k = 0; k < cnt[s]; ++k) o[k + 1] = sh[ids[size_t(k) * ctx->nl.npad + s]];
            std::sort(o, o + cnt[s] + 1);
        }
    });
}

fjsph_status fjsph_dsph_prestep(fjsph_ctx* ctx, double* npd)
{
    return guarded(ctx, [&] {
        require(ctx, true);
        dsph(ctx, false);
        HIPCHK(hipMemcpyAsync(&ctx->sc_host->npd, &ctx->sc->npd, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
        sync(ctx);
        if (npd) *npd = ctx->sc_host->npd;
        drain_timers(ctx);
    });
}

fjsph_status fjsph_aero_velocity(fjsph_ctx* ctx)
{
    return guarded(ctx, [&] {
        require(ctx, true);
        if (ctx->s.aero_source == FJSPH_MESH_INFL) aero_mesh(ctx);
        else if (ctx->s.aero_source == FJSPH_SURROGATE) surrogate_lookup(ctx);
        else launch_aero_velocity(ctx->s.dim, ctx->c, ctx->pnp1, ctx->part_id, ctx->nl, ctx->n, ctx->st);
        sync(ctx);
    });
}

fjsph_status fjsph_set_mesh(fjsph_ctx* ctx, int64_t nverts, const double* verts, int64_t nfaces,
                            const int64_t* face_off, const int64_t* face_verts, const int64_t* face_lr,
                            int64_t ncells, const int64_t* cell_off, const int64_t* cell_faces,
                            const double* vert_vel, const double* vert_p, const double* vert_rho)
{
    if (!ctx || nverts <= 0 || nfaces <= 0 || ncells <= 0 || !verts || !face_off || !face_verts || !face_lr ||
        !cell_off || !cell_faces || !vert_vel || !vert_p || !vert_rho)
        return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        const int D = ctx->s.dim;
        const int64_t nfv = face_off[nfaces], ncf = cell_off[ncells];
        if (face_off[0] != 0 || cell_off[0] != 0 || nfv > INT32_MAX || ncf > INT32_MAX || nverts > INT32_MAX ||
            ncells > INT32_MAX)
            throw std::string("set_mesh: offsets must start at 0 and sizes fit 32-bit indices");
        for (int64_t f = 0; f < nfaces; ++f) {
            const int64_t nv = face_off[f + 1] - face_off[f];
            if (nv < D || (D == 2 && nv != 2)) throw std::string("set_mesh: a face has too few vertices");
        }
        for (int64_t k = 0; k < nfv; ++k)
            if (face_verts[k] < 0 || face_verts[k] >= nverts) throw std::string("set_mesh: face vertex out of range");
        for (int64_t k = 0; k < ncf; ++k)
            if (cell_faces[k] < 0 || cell_faces[k] >= nfaces) throw std::string("set_mesh: cell face out of range");

        /* Average_Point_Data_to_Cell (src/CDFIO.cpp:112-155): compensated sums
         * over the cell's unique vertices in ascending order */
        ctx->mcentre.assign(size_t(ncells) * D, 0.0);
        ctx->mvel.assign(size_t(ncells) * D, 0.0);
        ctx->mp.assign(size_t(ncells), 0.0);
        ctx->mrho.assign(size_t(ncells), 0.0);
        std::vector<double> cbox(size_t(ncells) * 6);
        std::vector<int64_t> uv;
        auto kahan = [&](auto&& val) {
            double sum = 0.0, cc = 0.0;
            for (int64_t v : uv) {
                const double y = val(v) - cc;
                const double t = sum + y;
                cc = (t - sum) - y;
                sum = t;
            }
            return sum;
        };
        for (int64_t c = 0; c < ncells; ++c) {
            uv.clear();
            for (int64_t k = cell_off[c]; k < cell_off[c + 1]; ++k) {
                const int64_t f = cell_faces[k];
                for (int64_t q = face_off[f]; q < face_off[f + 1]; ++q) uv.push_back(face_verts[q]);
            }
            std::sort(uv.begin(), uv.end());
            uv.erase(std::unique(uv.begin(), uv.end()), uv.end());
            const double nv = double(uv.size());
            for (int a = 0; a < D; ++a) {
                ctx->mcentre[c * D + a] = kahan([&](int64_t v) { return verts[v * D + a]; }) / nv;
                ctx->mvel[c * D + a] = kahan([&](int64_t v) { return vert_vel[v * D + a]; }) / nv;
            }
            ctx->mp[c] = kahan([&](int64_t v) { return vert_p[v]; }) / nv;
            ctx->mrho[c] = kahan([&](int64_t v) { return vert_rho[v]; }) / nv;
            for (int a = 0; a < 3; ++a) {
                double lo = 0.0, hi = 0.0;
                if (a < D) {
                    lo = hi = verts[uv[0] * D + a];
                    for (int64_t v : uv) { lo = std::min(lo, verts[v * D + a]); hi = std::max(hi, verts[v * D + a]); }
                }
                cbox[c * 6 + a] = lo;
                cbox[c * 6 + 3 + a] = hi;
            }
        }

        /* voxel candidate grid, about 4 cells per voxel, over the cells of
         * ordinary size: a cell more than 64 times the median extent (a
         * degenerate cell, say one with a vertex at a file's fill value) goes
         * to an overflow list that every lookup also scans, so that it
         * neither stretches the grid nor fills millions of voxels */
        std::vector<double> ext(size_t(ncells), 0.0);
        for (int64_t c = 0; c < ncells; ++c)
            for (int a = 0; a < D; ++a) ext[size_t(c)] = std::max(ext[size_t(c)], cbox[c * 6 + 3 + a] - cbox[c * 6 + a]);
        std::vector<double> sorted_ext(ext);
        std::nth_element(sorted_ext.begin(), sorted_ext.begin() + ncells / 2, sorted_ext.end());
        const double ext_cap = 64.0 * std::max(sorted_ext[size_t(ncells / 2)], 1e-300);
        std::vector<char> big(size_t(ncells), 0);
        std::vector<int32_t> big_cells;
        for (int64_t c = 0; c < ncells; ++c)
            if (!(ext[size_t(c)] <= ext_cap)) {
                big[size_t(c)] = 1;
                big_cells.push_back(int32_t(c));
            }
        if (int64_t(big_cells.size()) == ncells) throw std::string("set_mesh: no cell of ordinary size");
        double glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        bool first_cell = true;
        for (int64_t c = 0; c < ncells; ++c) {
            if (big[size_t(c)]) continue;
            for (int a = 0; a < D; ++a) {
                glo[a] = first_cell ? cbox[c * 6 + a] : std::min(glo[a], cbox[c * 6 + a]);
                ghi[a] = first_cell ? cbox[c * 6 + 3 + a] : std::max(ghi[a], cbox[c * 6 + 3 + a]);
            }
            first_cell = false;
        }
        double vol = 1.0;
        for (int a = 0; a < D; ++a) vol *= std::max(ghi[a] - glo[a], 1e-300);
        double h = std::pow(vol / std::max(1.0, double(ncells) / 4.0), 1.0 / D);
        int32_t gd[3] = {1, 1, 1};
        for (;;) {
            double tot = 1.0;
            for (int a = 0; a < D; ++a) {
                gd[a] = int32_t(std::min(1.0e6, std::floor((ghi[a] - glo[a]) / h) + 1.0));
                tot *= gd[a];
            }
            if (tot <= 6.4e7) break;
            h *= 1.25;
        }
        const double ginv = 1.0 / h;
        const int64_t nvox = int64_t(gd[0]) * gd[1] * gd[2];
        auto vrange = [&](int64_t c, int a, int32_t& g0, int32_t& g1) {
            g0 = 0;
            g1 = 0;
            if (a >= D) return;
            g0 = int32_t(std::max(0.0, std::floor((cbox[c * 6 + a] - glo[a]) * ginv)));
            g1 = int32_t(std::min(double(gd[a] - 1), std::floor((cbox[c * 6 + 3 + a] - glo[a]) * ginv)));
        };
        std::vector<int64_t> gcount(size_t(nvox) + 1, 0);
        std::vector<int32_t> gcells;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<int64_t> fill;
            if (pass == 1) {
                fill.assign(gcount.begin(), gcount.end() - 1);
                gcells.assign(size_t(gcount[nvox]), 0);
            }
            for (int64_t c = 0; c < ncells; ++c) {
                if (big[size_t(c)]) continue;
                int32_t r[3][2];
                for (int a = 0; a < 3; ++a) vrange(c, a, r[a][0], r[a][1]);
                for (int32_t z = r[2][0]; z <= r[2][1]; ++z)
                    for (int32_t y = r[1][0]; y <= r[1][1]; ++y)
                        for (int32_t x = r[0][0]; x <= r[0][1]; ++x) {
                            const int64_t v = (int64_t(z) * gd[1] + y) * gd[0] + x;
                            if (pass == 0) ++gcount[v + 1];
                            else gcells[fill[v]++] = int32_t(c);
                        }
            }
            if (pass == 0) {
                for (int64_t v = 0; v < nvox; ++v) gcount[v + 1] += gcount[v];
                if (gcount[nvox] > INT32_MAX) throw std::string("set_mesh: voxel lists too long");
                continue;
            }
            /* upload */
            for (void* p : ctx->mesh_bufs) (void)hipFree(p);
            ctx->mesh_bufs.clear();
            auto up = [&](const void* src, size_t bytes) {
                void* d = nullptr;
                HIPCHK(hipMalloc(&d, std::max<size_t>(bytes, 4)));
                ctx->mesh_bufs.push_back(d);
                if (bytes) HIPCHK(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
                return d;
            };
            auto to32 = [](const int64_t* a, int64_t n) { return std::vector<int32_t>(a, a + n); };
            std::vector<double> v3(size_t(nverts) * 3, 0.0);
            for (int64_t v = 0; v < nverts; ++v)
                for (int a = 0; a < D; ++a) v3[v * 3 + a] = verts[v * D + a];
            std::vector<int32_t> fo = to32(face_off, nfaces + 1), fv = to32(face_verts, nfv);
            std::vector<int32_t> co = to32(cell_off, ncells + 1), cf = to32(cell_faces, ncf);
            std::vector<int32_t> gs(size_t(nvox) + 1);
            for (int64_t v = 0; v <= nvox; ++v) gs[v] = int32_t(gcount[v]);
            sync(ctx);
            DevMesh& m = ctx->dm;
            m.ncells = int32_t(ncells);
            m.verts = (const double*)up(v3.data(), v3.size() * 8);
            m.face_off = (const int32_t*)up(fo.data(), fo.size() * 4);
            m.face_v = (const int32_t*)up(fv.data(), fv.size() * 4);
            m.cell_off = (const int32_t*)up(co.data(), co.size() * 4);
            m.cell_f = (const int32_t*)up(cf.data(), cf.size() * 4);
            m.cVel = (const double*)up(ctx->mvel.data(), ctx->mvel.size() * 8);
            m.cP = (const double*)up(ctx->mp.data(), ctx->mp.size() * 8);
            m.cRho = (const double*)up(ctx->mrho.data(), ctx->mrho.size() * 8);
            m.gstart = (const int32_t*)up(gs.data(), gs.size() * 4);
            m.gcells = (const int32_t*)up(gcells.data(), gcells.size() * 4);
            m.big = (const int32_t*)up(big_cells.data(), big_cells.size() * 4);
            m.nbig = int32_t(big_cells.size());
            std::vector<int32_t> lr32(size_t(nfaces) * 2);
            for (int64_t k = 0; k < nfaces * 2; ++k) lr32[size_t(k)] = int32_t(face_lr[k]);
            m.face_lr = (const int32_t*)up(lr32.data(), lr32.size() * 4);
            std::vector<double> c3(size_t(ncells) * 3, 0.0);
            for (int64_t c = 0; c < ncells; ++c)
                for (int a = 0; a < D; ++a) c3[size_t(c) * 3 + a] = ctx->mcentre[size_t(c) * D + a];
            m.cCentre = (const double*)up(c3.data(), c3.size() * 8);
            for (int a = 0; a < 3; ++a) { m.glo[a] = glo[a]; m.gdims[a] = gd[a]; }
            m.ginv = ginv;
            /* cell centres binned on a grid of about 4 centres per voxel
             * (nearest-centre queries of the particles in no cell) */
            double clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
            for (int a = 0; a < D; ++a) {
                clo[a] = chi[a] = c3[size_t(a)];
                for (int64_t c = 0; c < ncells; ++c) {
                    clo[a] = std::min(clo[a], c3[size_t(c) * 3 + a]);
                    chi[a] = std::max(chi[a], c3[size_t(c) * 3 + a]);
                }
            }
            double span = 0.0, vol = 1.0;
            for (int a = 0; a < D; ++a) span = std::max(span, chi[a] - clo[a]);
            if (!(span > 0.0)) span = 1.0;
            for (int a = 0; a < D; ++a) vol *= std::max(chi[a] - clo[a], 1e-6 * span);
            double ch = std::pow(vol * 4.0 / double(ncells), 1.0 / D);
            int64_t cd[3] = {1, 1, 1}, ncv = 1;
            for (int it = 0; it < 64; ++it) {
                ncv = 1;
                for (int a = 0; a < D; ++a) {
                    cd[a] = int64_t(std::floor((chi[a] - clo[a]) / ch)) + 1;
                    ncv *= cd[a];
                }
                if (ncv <= 2 * ncells + 64) break;
                ch *= 1.25;
            }
            const double cinv = 1.0 / ch;
            std::vector<int32_t> cst(size_t(ncv) + 1, 0), ccl(static_cast<size_t>(ncells)), cvox(static_cast<size_t>(ncells));
            for (int64_t c = 0; c < ncells; ++c) {
                int64_t vx = 0;
                for (int a = D - 1; a >= 0; --a) {
                    const int64_t g = std::min<int64_t>(cd[a] - 1, int64_t((c3[size_t(c) * 3 + a] - clo[a]) * cinv));
                    vx = vx * cd[a] + std::max<int64_t>(g, 0);
                }
                cvox[size_t(c)] = int32_t(vx);
                cst[size_t(vx) + 1]++;
            }
            for (int64_t v = 0; v < ncv; ++v) cst[size_t(v) + 1] += cst[size_t(v)];
            {
                std::vector<int32_t> fill(cst.begin(), cst.end() - 1);
                for (int64_t c = 0; c < ncells; ++c) ccl[size_t(fill[size_t(cvox[size_t(c)])]++)] = int32_t(c);
            }
            m.cstart = (const int32_t*)up(cst.data(), cst.size() * 4);
            m.ccells = (const int32_t*)up(ccl.data(), ccl.size() * 4);
            for (int a = 0; a < 3; ++a) { m.clo[a] = clo[a]; m.cdims[a] = int32_t(cd[a]); }
            m.cinv = cinv;
        }
        ctx->mesh_ncells = ncells;
        ctx->mesh_max_len = mesh_max_length(D, nfaces, face_off, face_verts, verts);
        ctx->have_mesh = true;
    });
}

fjsph_status fjsph_set_surrogate(fjsph_ctx* ctx, int32_t n_hidden, const float* weights, const float* biases,
                                 const double* in_shift, const double* in_scale, const double* out_shift,
                                 const double* out_scale)
{
    if (!ctx || n_hidden < 1 || n_hidden > 8 || !weights || !biases || !in_shift || !in_scale || !out_shift ||
        !out_scale)
        return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        const int D = ctx->s.dim, NO = D + 2;
        for (int k = 0; k < D; ++k)
            if (!(in_scale[k] != 0.0)) throw std::string("set_surrogate: zero input scale");
        /* FP32 -> BF16, round to nearest even */
        auto bf16 = [](float f) {
            uint32_t u;
            std::memcpy(&u, &f, 4);
            if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);
            u += 0x7fffu + ((u >> 16) & 1u);
            return uint16_t(u >> 16);
        };
        /* A-operand tiles of a layer: W[out][in] padded to MT*16 x KC*32;
         * lane l of tile (mt, kc) holds W[mt*16 + l%16][kc*32 + 8*(l/16) + e] */
        std::vector<uint16_t> tiles;
        const float* wp = weights;
        auto add_layer = [&](int nout, int nin, int MT, int KC) {
            for (int mt = 0; mt < MT; ++mt)
                for (int kc = 0; kc < KC; ++kc)
                    for (int l = 0; l < 64; ++l)
                        for (int e = 0; e < 8; ++e) {
                            const int r = mt * 16 + l % 16, k = kc * 32 + 8 * (l / 16) + e;
                            tiles.push_back(r < nout && k < nin ? bf16(wp[size_t(r) * nin + k]) : uint16_t(0));
                        }
            wp += size_t(nout) * nin;
        };
        add_layer(64, D, 4, 1);
        for (int l = 1; l < n_hidden; ++l) add_layer(64, 64, 4, 2);
        add_layer(NO, 64, 1, 2);
        std::vector<float> bh(biases, biases + size_t(n_hidden) * 64);
        std::vector<float> bo(16, 0.0f);
        for (int k = 0; k < NO; ++k) bo[k] = biases[size_t(n_hidden) * 64 + k];
        for (void* p : ctx->sg_bufs) (void)hipFree(p);
        ctx->sg_bufs.clear();
        auto up = [&](const void* src, size_t bytes) {
            void* d = nullptr;
            HIPCHK(hipMalloc(&d, bytes));
            ctx->sg_bufs.push_back(d);
            HIPCHK(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
            return d;
        };
        sync(ctx);
        Surrogate& sg = ctx->sg;
        sg.w = up(tiles.data(), tiles.size() * 2);
        sg.b = (const float*)up(bh.data(), bh.size() * 4);
        sg.bo = (const float*)up(bo.data(), bo.size() * 4);
        sg.n_hidden = n_hidden;
        for (int k = 0; k < 3; ++k) {
            sg.in_shift[k] = k < D ? in_shift[k] : 0.0;
            sg.in_scale_inv[k] = k < D ? 1.0 / in_scale[k] : 0.0;
        }
        for (int k = 0; k < 5; ++k) {
            sg.out_shift[k] = k < NO ? out_shift[k] : 0.0;
            sg.out_scale[k] = k < NO ? out_scale[k] : 0.0;
        }
        ctx->have_surrogate = true;
    });
}

extern "C++" {
namespace {

/* Device buffers of surrogate training, freed on scope exit. */
struct TrainBufs {
    std::vector<void*> ptrs;
    template <typename T>
    T* alloc(size_t count)
    {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        ptrs.push_back(p);
        return (T*)p;
    }
    ~TrainBufs()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

int64_t mlp_nw(int D, int L) { return int64_t(64) * D + int64_t(L - 1) * 64 * 64 + int64_t(D + 2) * 64; }
int64_t mlp_nb(int D, int L) { return int64_t(L) * 64 + D + 2; }

/* One forward/backward pass over the batch idx (on the device) with the
 * parameters p (weights then biases): gradients into g, the batch's sum of
 * squared errors into *sse (device). */
void mlp_pass(fjsph_ctx* ctx, MlpTrain& m, float* p, float* g, float* sse, void* wf, void* wt)
{
    const int D = ctx->s.dim, L = m.n_hidden;
    const int64_t nwaves = m.batch / 64;
    launch_mlp_pack(D, p, L, wf, wt, ctx->st);
    m.wf = wf;
    m.wt = wt;
    m.b = p + m.nw;
    m.bo = p + m.nw + int64_t(L) * 64;
    launch_mlp_train(D, m, ctx->st);
    launch_mlp_reduce(m.dWp, nwaves, m.nw, m.nw, g, ctx->st);
    launch_mlp_reduce(m.dbp, nwaves, m.nb, m.nb, g + m.nw, ctx->st);
    launch_mlp_reduce(m.lossp, nwaves, 1, 1, sse, ctx->st);
}

/* Scratch of a pass over a batch of B samples. */
void mlp_scratch(TrainBufs& tb, MlpTrain& m, int D, int L, int32_t B)
{
    m.n_hidden = L;
    m.batch = B;
    m.nw = mlp_nw(D, L);
    m.nb = mlp_nb(D, L);
    const int64_t nwaves = B / 64;
    m.Z = tb.alloc<float>(size_t(L) * B * 64);
    m.H = tb.alloc<__bf16>(size_t(L + 1) * B * 64);
    m.dWp = tb.alloc<float>(size_t(nwaves) * m.nw);
    m.dbp = tb.alloc<float>(size_t(nwaves) * m.nb);
    m.lossp = tb.alloc<float>(size_t(nwaves));
}

}  // namespace
}  // extern "C++"

fjsph_status fjsph_surrogate_grad(fjsph_ctx* ctx, int32_t n_hidden, int32_t batch, const float* xn, const float* tn,
                                  const float* weights, const float* biases, float* dweights, float* dbiases,
                                  double* loss)
{
    if (!ctx || n_hidden < 1 || n_hidden > 8 || batch <= 0 || batch % 64 != 0 || !xn || !tn || !weights ||
        !biases || !dweights || !dbiases)
        return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        const int D = ctx->s.dim, NO = D + 2, L = n_hidden;
        TrainBufs tb;
        MlpTrain m{};
        mlp_scratch(tb, m, D, L, batch);
        std::vector<float> X(size_t(batch) * 4, 0.f), T(size_t(batch) * 8, 0.f);
        std::vector<int32_t> idx(batch);
        for (int32_t j = 0; j < batch; ++j) {
            for (int a = 0; a < D; ++a) X[size_t(j) * 4 + a] = xn[size_t(j) * D + a];
            for (int a = 0; a < NO; ++a) T[size_t(j) * 8 + a] = tn[size_t(j) * NO + a];
            idx[j] = j;
        }
        float* dX = tb.alloc<float>(X.size());
        float* dT = tb.alloc<float>(T.size());
        int32_t* didx = tb.alloc<int32_t>(idx.size());
        float* p = tb.alloc<float>(size_t(m.nw + m.nb));
        float* g = tb.alloc<float>(size_t(m.nw + m.nb));
        float* sse = tb.alloc<float>(1);
        void* wf = tb.alloc<uint16_t>(size_t(mlp_tiles_fwd(L)) * 64 * 8);
        void* wt = tb.alloc<uint16_t>(size_t(mlp_tiles_bwd(L)) * 64 * 8);
        sync(ctx);
        HIPCHK(hipMemcpy(dX, X.data(), X.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dT, T.data(), T.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(didx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(p, weights, size_t(m.nw) * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(p + m.nw, biases, size_t(m.nb) * 4, hipMemcpyHostToDevice));
        m.X = dX;
        m.T = dT;
        m.idx = didx;
        mlp_pass(ctx, m, p, g, sse, wf, wt);
        sync(ctx);
        HIPCHK(hipMemcpy(dweights, g, size_t(m.nw) * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(dbiases, g + m.nw, size_t(m.nb) * 4, hipMemcpyDeviceToHost));
        float s = 0.f;
        HIPCHK(hipMemcpy(&s, sse, 4, hipMemcpyDeviceToHost));
        if (loss) *loss = double(s) / (double(batch) * NO);
    });
}

fjsph_status fjsph_surrogate_grad_noalloc(fjsph_ctx* ctx, int32_t n_hidden, int32_t batch, const float* xn,
                                          const float* tn, const float* weights, const float* biases,
                                          float* dweights, float* dbiases, double* loss)
{
    if (!ctx || n_hidden < 1 || n_hidden > 8 || batch <= 0 || batch % 64 != 0 || !xn || !tn || !weights ||
        !biases || !dweights || !dbiases)
        return FJSPH_ERR_ARG;
    if (batch > FJSPH_NOALLOC_MAX_BATCH) {
        set_err(ctx, "surrogate_grad_noalloc: batch %d exceeds the maximum %d of the scratch array", int(batch),
                int(FJSPH_NOALLOC_MAX_BATCH));
        return FJSPH_ERR_ARG;
    }
    return guarded(ctx, [&] {
        unsigned char* base = (unsigned char*)noalloc_scratch(ctx->device);
        if (!base) throw std::string("surrogate_grad_noalloc: the scratch array of this device was not set up");
        const int D = ctx->s.dim, NO = D + 2, L = n_hidden;
        const int64_t nwaves = batch / 64;
        MlpTrain m{};
        m.n_hidden = L;
        m.batch = batch;
        m.nw = mlp_nw(D, L);
        m.nb = mlp_nb(D, L);
        /* carve the scratch array (256-byte aligned pieces) */
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            unsigned char* p = base + off;
            off += (bytes + 255) / 256 * 256;
            return (void*)p;
        };
        float* X = (float*)carve(size_t(batch) * 4 * sizeof(float));
        float* T = (float*)carve(size_t(batch) * 8 * sizeof(float));
        int32_t* idx = (int32_t*)carve(size_t(batch) * sizeof(int32_t));
        m.Z = (float*)carve(size_t(L) * batch * 64 * sizeof(float));
        m.H = (__bf16*)carve(size_t(L + 1) * batch * 64 * sizeof(__bf16));
        m.dWp = (float*)carve(size_t(nwaves) * m.nw * sizeof(float));
        m.dbp = (float*)carve(size_t(nwaves) * m.nb * sizeof(float));
        m.lossp = (float*)carve(size_t(nwaves) * sizeof(float));
        float* sse = (float*)carve(sizeof(float));
        void* wf = carve(size_t(mlp_tiles_fwd(L)) * 64 * 8 * sizeof(uint16_t));
        void* wt = carve(size_t(mlp_tiles_bwd(L)) * 64 * 8 * sizeof(uint16_t));
        if (off > kNoallocScratchBytes) throw std::string("surrogate_grad_noalloc: scratch array too small");
        /* every scratch word the kernels read is written first in this call */
        launch_mlp_stage(xn, tn, batch, D, X, T, idx, ctx->st);
        launch_mlp_pack(D, weights, L, wf, wt, ctx->st);
        m.X = X;
        m.T = T;
        m.idx = idx;
        m.wf = wf;
        m.wt = wt;
        m.b = biases;
        m.bo = biases + int64_t(L) * 64;
        launch_mlp_train(D, m, ctx->st);
        launch_mlp_reduce(m.dWp, nwaves, m.nw, m.nw, dweights, ctx->st);
        launch_mlp_reduce(m.dbp, nwaves, m.nb, m.nb, dbiases, ctx->st);
        if (loss) {
            launch_mlp_reduce(m.lossp, nwaves, 1, 1, sse, ctx->st);
            launch_mlp_loss(sse, double(batch) * NO, loss, ctx->st);
        }
        sync(ctx);
    });
}

fjsph_status fjsph_surrogate_train(fjsph_ctx* ctx, int64_t n, const double* points, const double* values,
                                   int32_t n_hidden, int32_t steps, int32_t batch, double lr, uint64_t seed,
                                   const double* in_shift, const double* in_scale, const double* out_shift,
                                   const double* out_scale, float* weights, float* biases, double* loss)
{
    if (!ctx || n <= 0 || n > 0x7fffffff || !points || !values || n_hidden < 1 || n_hidden > 8 || steps < 0 ||
        batch <= 0 || batch % 64 != 0 || !(lr > 0.0) || !in_shift || !in_scale || !out_shift || !out_scale ||
        !weights || !biases)
        return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        const int D = ctx->s.dim, NO = D + 2, L = n_hidden;
        for (int a = 0; a < D; ++a)
            if (!(in_scale[a] != 0.0)) throw std::string("surrogate_train: zero input scale");
        for (int a = 0; a < NO; ++a)
            if (!(out_scale[a] != 0.0)) throw std::string("surrogate_train: zero output scale");
        TrainBufs tb;
        MlpTrain m{};
        mlp_scratch(tb, m, D, L, batch);
        std::vector<float> X(size_t(n) * 4, 0.f), T(size_t(n) * 8, 0.f);
        for (int64_t j = 0; j < n; ++j) {
            for (int a = 0; a < D; ++a) X[size_t(j) * 4 + a] = float((points[j * D + a] - in_shift[a]) / in_scale[a]);
            for (int a = 0; a < NO; ++a)
                T[size_t(j) * 8 + a] = float((values[j * NO + a] - out_shift[a]) / out_scale[a]);
        }
        const int64_t np = m.nw + m.nb;
        float* dX = tb.alloc<float>(X.size());
        float* dT = tb.alloc<float>(T.size());
        int32_t* didx = tb.alloc<int32_t>(size_t(batch));
        float* p = tb.alloc<float>(size_t(np));
        float* g = tb.alloc<float>(size_t(np));
        float* m1 = tb.alloc<float>(size_t(np));
        float* m2 = tb.alloc<float>(size_t(np));
        float* sse = tb.alloc<float>(size_t(std::max(steps, 1)));
        void* wf = tb.alloc<uint16_t>(size_t(mlp_tiles_fwd(L)) * 64 * 8);
        void* wt = tb.alloc<uint16_t>(size_t(mlp_tiles_bwd(L)) * 64 * 8);
        sync(ctx);
        HIPCHK(hipMemcpy(dX, X.data(), X.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dT, T.data(), T.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(p, weights, size_t(m.nw) * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(p + m.nw, biases, size_t(m.nb) * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemsetAsync(m1, 0, size_t(np) * 4, ctx->st));
        HIPCHK(hipMemsetAsync(m2, 0, size_t(np) * 4, ctx->st));
        m.X = dX;
        m.T = dT;
        m.idx = didx;
        /* Adam (beta 0.9, 0.999, eps 1e-8) with a cosine learning-rate
         * schedule from lr to 0 over the steps */
        const double b1 = 0.9, b2 = 0.999;
        for (int32_t t = 1; t <= steps; ++t) {
            launch_mlp_batch(seed, t, n, batch, didx, ctx->st);
            mlp_pass(ctx, m, p, g, sse + (t - 1), wf, wt);
            const double lr_t = lr * 0.5 * (1.0 + std::cos(3.14159265358979323846 * double(t - 1) / double(steps)));
            launch_mlp_adam(p, g, m1, m2, np, float(lr_t), float(b1), float(b2), 1e-8f,
                            float(1.0 - std::pow(b1, t)), float(1.0 - std::pow(b2, t)), ctx->st);
        }
        sync(ctx);
        HIPCHK(hipMemcpy(weights, p, size_t(m.nw) * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(biases, p + m.nw, size_t(m.nb) * 4, hipMemcpyDeviceToHost));
        if (loss && steps > 0) {
            std::vector<float> l(steps);
            HIPCHK(hipMemcpy(l.data(), sse, size_t(steps) * 4, hipMemcpyDeviceToHost));
            for (int32_t t = 0; t < steps; ++t) loss[t] = double(l[t]) / (double(batch) * NO);
        }
    });
}

fjsph_status fjsph_surrogate_eval(fjsph_ctx* ctx, int64_t n, const double* points, double* out)
{
    if (!ctx || n < 0 || (n > 0 && (!points || !out))) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        if (!ctx->have_surrogate) throw std::string("no surrogate set");
        if (n == 0) return;
        const int D = ctx->s.dim;
        double *dp = nullptr, *dout = nullptr;
        HIPCHK(hipMalloc((void**)&dp, sizeof(double) * n * D));
        HIPCHK(hipMalloc((void**)&dout, sizeof(double) * n * (D + 2)));
        sync(ctx);
        HIPCHK(hipMemcpy(dp, points, sizeof(double) * n * D, hipMemcpyHostToDevice));
        launch_surrogate(D, ctx->c, DevState{}, nullptr, ctx->sg, n, 1, dp, dout, ctx->st);
        sync(ctx);
        HIPCHK(hipMemcpy(out, dout, sizeof(double) * n * (D + 2), hipMemcpyDeviceToHost));
        (void)hipFree(dp);
        (void)hipFree(dout);
    });
}

fjsph_status fjsph_get_mesh_cells(fjsph_ctx* ctx, double* centre, double* vel, double* p, double* rho)
{
    if (!ctx) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        if (!ctx->have_mesh) throw std::string("no mesh set");
        if (centre) std::copy(ctx->mcentre.begin(), ctx->mcentre.end(), centre);
        if (vel) std::copy(ctx->mvel.begin(), ctx->mvel.end(), vel);
        if (p) std::copy(ctx->mp.begin(), ctx->mp.end(), p);
        if (rho) std::copy(ctx->mrho.begin(), ctx->mrho.end(), rho);
    });
}

fjsph_status fjsph_mesh_eval(fjsph_ctx* ctx, int64_t n, const double* points, int32_t* cells, double* vals)
{
    if (!ctx || n < 0 || (n > 0 && (!points || !cells || !vals))) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        if (!ctx->have_mesh) throw std::string("no mesh set");
        if (n == 0) return;
        const int D = ctx->s.dim;
        double *dp = nullptr, *dv = nullptr;
        int32_t* dc = nullptr;
        HIPCHK(hipMalloc((void**)&dp, sizeof(double) * n * D));
        HIPCHK(hipMalloc((void**)&dv, sizeof(double) * n * (D + 2)));
        HIPCHK(hipMalloc((void**)&dc, sizeof(int32_t) * n));
        sync(ctx);
        HIPCHK(hipMemcpy(dp, points, sizeof(double) * n * D, hipMemcpyHostToDevice));
        launch_mesh_eval(D, ctx->dm, n, dp, dc, dv, ctx->st);
        sync(ctx);
        HIPCHK(hipMemcpy(cells, dc, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(vals, dv, sizeof(double) * n * (D + 2), hipMemcpyDeviceToHost));
        (void)hipFree(dp);
        (void)hipFree(dv);
        (void)hipFree(dc);
    });
}

fjsph_status fjsph_mesh_lost(fjsph_ctx* ctx, int64_t* lost)
{
    if (!ctx || !lost) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        HIPCHK(hipMemcpyAsync(&ctx->sc_host->mesh_lost, &ctx->sc->mesh_lost, sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->st));
        sync(ctx);
        *lost = ctx->sc_host->mesh_lost;
    });
}

fjsph_status fjsph_mesh_lost_info(fjsph_ctx* ctx, fjsph_lost_info* info)
{
    if (!ctx || !info) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        HIPCHK(hipMemcpyAsync(&ctx->sc_host->mesh_lost, &ctx->sc->mesh_lost, sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->st));
        sync(ctx);
        info->lost = ctx->sc_host->mesh_lost;
        info->deleted = ctx->lost_deleted;
        info->internal = ctx->lost_internal;
        info->failed = ctx->lost_failed;
        info->first_cell = ctx->first_cell;
        info->first_cell_deleted = ctx->first_cell_deleted;
    });
}

fjsph_status fjsph_detect_surface(fjsph_ctx* ctx)
{
    return guarded(ctx, [&] {
        require(ctx, true);
        refresh_aux(ctx);
        surface_and_dissipation(ctx, false, false);
        sync(ctx);
        drain_timers(ctx);
    });
}

fjsph_status fjsph_dissipation_terms(fjsph_ctx* ctx)
{
    return guarded(ctx, [&] {
        require(ctx, true);
        PassTimer t(ctx, FJSPH_T_DISSIPATION);
        refresh_aux(ctx);
        launch_surface1(ctx->s.dim, ctx->s.precision, ctx->c, ctx->pnp1, ctx->aux, ctx->part_id, ctx->tg, ctx->nl,
                        ctx->n, r_from_x(ctx), ctx->norms_tmp, 2, ctx->st);
        sync(ctx);
    });
}

fjsph_status fjsph_particle_shift(fjsph_ctx* ctx)
{
    return guarded(ctx, [&] {
        require(ctx, true);
        if (!ctx->s.ale) throw std::string("particle_shift needs settings.ale = 1");
        refresh_aux(ctx);
        shift(ctx);
        sync(ctx);
        drain_timers(ctx);
    });
}

fjsph_status fjsph_boundary_update(fjsph_ctx* ctx)
{
    return guarded(ctx, [&] {
        require(ctx, true);
        boundary_update(ctx);
        sync(ctx);
        drain_timers(ctx);
    });
}

fjsph_status fjsph_acc_rrho(fjsph_ctx* ctx, double npd)
{
    return guarded(ctx, [&] {
        if (grouped(ctx))
            throw std::string("this per-pass entry point acts on one simulation; with groups, step them with "
                              "fjsph_integrate_step");
        require(ctx, true);
        HIPCHK(hipMemcpyAsync(&ctx->sc->npd, &npd, sizeof(double), hipMemcpyHostToDevice, ctx->st));
        force(ctx, false);
        sync(ctx);
        drain_timers(ctx);
    });
}

fjsph_status fjsph_nb_iter(fjsph_ctx* ctx, double npd, double* log_error)
{
    return guarded(ctx, [&] {
        if (grouped(ctx))
            throw std::string("this per-pass entry point acts on one simulation; with groups, step them with "
                              "fjsph_integrate_step");
        require(ctx, true);
        HIPCHK(hipMemcpyAsync(&ctx->sc->npd, &npd, sizeof(double), hipMemcpyHostToDevice, ctx->st));
        const double le = nb_iter(ctx);
        if (log_error) *log_error = le;
        drain_timers(ctx);
    });
}

fjsph_status fjsph_find_timestep(fjsph_ctx* ctx, double* dt)
{
    return guarded(ctx, [&] {
        if (grouped(ctx))
            throw std::string("this per-pass entry point acts on one simulation; with groups, step them with "
                              "fjsph_integrate_step");
        require(ctx, false);
        const double v = find_timestep(ctx);
        if (dt) *dt = v;
        drain_timers(ctx);
    });
}

fjsph_status fjsph_integrate_no_update(fjsph_ctx* ctx, fjsph_step_info* info)
{
    return guarded(ctx, [&] {
        require(ctx, false);
        const double rms = integrate_no_update(ctx);
        fill_info(ctx, rms, info);
    });
}

fjsph_status fjsph_integrate_step(fjsph_ctx* ctx, fjsph_step_info* info)
{
    return guarded(ctx, [&] {
        require(ctx, false);
        const double rms = integrate(ctx);
        fill_info(ctx, rms, info);
        drain_timers(ctx);
    });
}

fjsph_status fjsph_advance_frame(fjsph_ctx* ctx)
{
    if (!ctx) return FJSPH_ERR_ARG;
    ctx->last_frame_time += ctx->s.frame_time_interval;
    ctx->current_frame++;
    for (fjsph_ctx::Group& g : ctx->grs) g.last_frame_time += ctx->s.frame_time_interval;
    return FJSPH_OK;
}

fjsph_status fjsph_get_time(fjsph_ctx* ctx, double* time, double* delta_t, double* cfl)
{
    if (!ctx) return FJSPH_ERR_ARG;
    if (time) *time = ctx->current_time;
    if (delta_t) *delta_t = ctx->delta_t;
    if (cfl) *cfl = ctx->cfl;
    return FJSPH_OK;
}

fjsph_status fjsph_set_time(fjsph_ctx* ctx, double time, double delta_t, double cfl)
{
    if (!ctx) return FJSPH_ERR_ARG;
    ctx->current_time = time;
    ctx->delta_t = delta_t;
    ctx->cfl = cfl;
    for (fjsph_ctx::Group& g : ctx->grs) {
        g.time = time;
        g.dt = delta_t;
        g.cfl = cfl;
    }
    return FJSPH_OK;
}

fjsph_status fjsph_enable_timers(fjsph_ctx* ctx, int32_t on)
{
    if (!ctx) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        drain_timers(ctx);
        ctx->timing = on != 0;
        if (ctx->timing)
            for (int k = 0; k < FJSPH_T_COUNT; ++k) { ctx->t_ms[k] = 0.0; ctx->t_launch[k] = 0; }
    });
}

fjsph_status fjsph_get_timers(fjsph_ctx* ctx, double* ms, int64_t* launches)
{
    if (!ctx || !ms || !launches) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        drain_timers(ctx);
        for (int k = 0; k < FJSPH_T_COUNT; ++k) { ms[k] = ctx->t_ms[k]; launches[k] = ctx->t_launch[k]; }
    });
}

fjsph_status fjsph_pass_bytes(fjsph_ctx* ctx, int32_t pass, double* bytes_per_particle)
{
    if (!ctx || !bytes_per_particle || pass < 0 || pass >= FJSPH_T_COUNT) return FJSPH_ERR_ARG;
    const double D = ctx->s.dim;
    const double ale = ctx->s.ale ? 1.0 : 0.0;
    double b = 0.0;
    switch (pass) {
    case FJSPH_T_FORCE:
        /* k_force_t / k_force_m, once per slot: reads order (4), x, v (2*8D), the FP32
         * record {m, p/rho^2, 1/rho, m/rho} (16), rho, m, p/rho^2, 1/rho,
         * deltaD (40), part_id, b (8), ALE vPert (8D) + surfzone (4);
         * writes acc (8D), Rrho (8). */
        b = 4 + 16 * D + 16 + 40 + 8 + 8 * D + 8 + ale * (8 * D + 4);
        break;
    case FJSPH_T_UPDATE:
        /* k_nb_update: reads part_id, b, cellID, neighbour count (16),
         * x, v, acc, aVisc, cellV (5*8D), cellRho, cellP, m, lam_nb, woccl
         * (40), pn x, v, acc (3*8D), Rrho, rho, p, pn rho, pn Rrho (40),
         * ALE vPert (8D); writes Af, acc (2*8D), next x, v (2*8D), next
         * rho, p (16). */
        b = 16 + 40 * D + 40 + 24 * D + 40 + ale * 8 * D + 32 * D + 16;
        break;
    case FJSPH_T_DSPH:
        /* reads x (8D), rho, m, b, part_id, count; writes L (8D^2), gradRho,
         * norm (16D), lam, lam_nb, colourG, kernsum, colour (40), cellV (8D),
         * cellID (4) */
        b = 8 * D + 16 + 12 + 8 * D * D + 16 * D + 40 + 8 * D + 4;
        break;
    case FJSPH_T_SURFACE:
        b = 2 * (8 * D + 8 * D * D + 8 * D + 48) + 8 * D * 3 + 24;
        /* particle_shift's 48 B when it runs inside loops 2-3 */
        if (FJSPH_SHIFT_FUSED && ctx->s.ale && ctx->s.precision == FJSPH_PREC_MIXED) b += 48;
        break;
    case FJSPH_T_NEIGHBOUR:
        b = 8 * D + 4 * 4 + 4;
        break;
    default:
        b = 48;
    }
    *bytes_per_particle = b;
    return FJSPH_OK;
}

fjsph_status fjsph_neighbour_stats(fjsph_ctx* ctx, int64_t* pairs, int32_t* max_count)
{
    if (!ctx) return FJSPH_ERR_ARG;
    return guarded(ctx, [&] {
        require(ctx, true);
        sync(ctx);
        std::vector<int32_t> c(static_cast<size_t>(ctx->n));
        HIPCHK(hipMemcpy(c.data(), ctx->nl.nnb, sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost));
        int64_t t = 0;
        int32_t mx = 0;
        for (int32_t v : c) { t += v; mx = std::max(mx, v); }
        if (pairs) *pairs = t;
        if (max_count) *max_count = mx;
    });
}

}  // extern "C"
