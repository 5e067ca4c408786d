/*
 * epi_amd.h -- C ABI of the MI355X implementation of the epi5_MT per-step
 * hot path (apqw/epi).  Plain pointers and sizes only.
 *
 * One context holds one or more tissues (a batch) bound to one HIP stream;
 * every call of the reference call surface advances all tissues of the
 * context in the same kernel launches.  A context is not thread-safe;
 * different contexts may be used from different threads.  All entry points
 * return 0 (EPI_OK) or an EPI_ERR_* code; epi_last_error() gives text.
 * Host buffers belong to the caller, device memory to the library.
 *
 * Arrays of cells are in cell-index order; fields are (nx+1)*(ny+1)*(nz+1)
 * doubles in the reference's [x][y][z] order (z fastest, Field.h:9-53).
 */
#ifndef EPI_AMD_H
#define EPI_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    EPI_OK = 0,
    EPI_ERR_CAPACITY = 1,            /* bucket >= 200, list >= 400, cells > capacity */
    EPI_ERR_STRONG_INTERACTION = 2,  /* |coef| > 1000 (cell_interaction.cpp:51) or membrane > 100 */
    EPI_ERR_NO_DERMIS = 3,           /* FIX without a membrane contact (cell_state_renew.cpp:309) */
    EPI_ERR_OUT_OF_GRID = 4,         /* cell outside the bucket or voxel grid */
    EPI_ERR_REMOVE_ALIAS = 5,        /* remove_exec would delete a moved cell (cellmanager.cpp:338-352) */
    EPI_ERR_RNG_EXHAUSTED = 6,       /* replay stream shorter than the draws needed */
    EPI_ERR_INVALID = 7,             /* bad argument or inconsistent input */
    EPI_ERR_HIP = 8,                 /* HIP runtime failure */
    EPI_ERR_PAIR_REMOVED = 9         /* a removed cell still had a division partner */
};

/* cell states (define.h:36-47) */
enum { EPI_ALIVE = 0, EPI_DEAD = 1, EPI_DISA = 2, EPI_UNUSED = 3, EPI_FIX = 4,
       EPI_BLANK = 5, EPI_DER = 6, EPI_MUSUME = 7, EPI_AIR = 8, EPI_MEMB = 9 };

typedef struct epi_ctx epi_ctx;

/* Runtime form of define.h:72-113 (NMX, NMY_tri, NX/NY/NZ, LX/LY/LZ,
 * MAX_CELL_NUM).  row_stride caps a neighbour list (64 to 400, the
 * reference's MAX_CONNECT_CELL_NUM; smaller values are raised to 64);
 * bucket_cap caps a bucket (<= 200, N3). 0 = default. */
typedef struct epi_geometry {
    int32_t nmx, nmy;
    int32_t nx, ny, nz;
    double lx, ly, lz;
    int32_t max_cells;
    int32_t row_stride;
    int32_t bucket_cap;
} epi_geometry;

/* Input rows of the deck (cellmanager.cpp:243-302): MEMB rows first, then
 * DER, then the rest; pair[i] is the partner row or -1. */
typedef struct epi_cell_input {
    int32_t n, nmemb, nder;
    const uint8_t* state; const int32_t* fix_origin;
    const double *x, *y, *z, *radius, *ca2p, *ca2p_avg, *ip3, *ex_inert;
    const double *agek, *ageb, *ex_fat, *in_fat, *spr_nat_len;
    const int32_t* rest_div_times; const uint8_t* is_malignant; const int32_t* pair;
} epi_cell_input;

/* Full per-cell state, download and kernel-level upload.  *_next are the
 * second buffers of the reference's dual_double / SwapData.  Lists are CSR
 * (conn_ptr[n+1], conn, gj per entry).  Null pointers are skipped. */
typedef struct epi_cells {
    int32_t n, nmemb, nder, sw;
    uint8_t* state; int32_t* fix_origin;
    double *x, *y, *z, *xn, *yn, *zn;
    double* radius;
    double *ca2p, *ca2p_next, *ip3, *ip3_next;
    double *ca2p_avg, *ex_inert, *agek, *ageb, *ex_fat, *in_fat, *spr_nat_len, *diff_u;
    int32_t *rest_div_times, *pair, *dermis, *lat;
    uint8_t* is_malignant;
    int64_t* uid;
    int32_t* conn_ptr; int32_t* conn; double* gj; int32_t conn_cap;
    int32_t* memb_adj;
    /* cell_stat per cell (define.h:198-202): k_aging_start, k_cornified,
     * k_disap timesteps, 3 per cell, -1 = never set (download only) */
    int64_t* cst;
} epi_cells;

/* Step driver of proc.cpp:131-173 per tissue, plus a forced Ca cadence:
 * ca_every > 0 also runs calc_ca2p when (step+1) % ca_every == 0. */
typedef struct epi_driver {
    int64_t step;
    int32_t force_cornif;
    int32_t num_sc;
    int32_t ca_every;
    int32_t clean_every;   /* accepted for parity with the oracle; stale gj
                              entries are never kept here, so it is a no-op */
} epi_driver;

typedef struct epi_status {
    int32_t n, err, sw, ndiv, nrm, maxlen;
    int64_t rng_pos, next_uid;
    double zzmax;
    int64_t ca_solves;      /* calc_ca2p calls run by epi_step */
    int64_t steps;          /* steps run by epi_step */
    int64_t cell_steps;     /* sum over those steps of the live cell count */
    int64_t grid_points_b;  /* sum of nx*ny*iz_bound over ext_stim sweeps */
    int64_t grid_points_ca; /* sum of nx*ny*iz_bound over ATP sub-steps */
} epi_status;

/* ---- lifetime ---------------------------------------------------------- */
/* new context with ntissues tissues on the current HIP device */
int epi_create(int32_t ntissues, const epi_geometry* geometries, epi_ctx** out);
void epi_destroy(epi_ctx* ctx);
const char* epi_last_error(const epi_ctx* ctx);
int32_t epi_num_tissues(const epi_ctx* ctx);

/* ---- runtime parameters ----------------------------------------------------
 * The physics constants of the path as runtime values: the key list of
 * epi_intg's CalcParams (epi_intg/calc/CalcParams.cpp:51-162) with the values
 * of epi5_MT as defaults (define.h:80-230, cell_interaction.cpp:21-45,
 * cell_state_renew.cpp:17-119, ca2p.cpp:14-144, ext_stim.cpp:9-20) and the
 * derived quantities of CalcParams::calculate_param (:16-45).  Unsigned and
 * bool keys are int32 (a bool reads as 1 when its text reads as the integer
 * 1, else 0; an unsigned key refuses a negative value).  A context starts
 * with the defaults, its geometry's grid and box filled in.
 * Keys that fix the shape of the build rather than a value must keep the
 * values the context was created with, and epi_set_params refuses a different
 * one with EPI_ERR_INVALID: LX, LY, LZ, NX, NY, NZ (the geometry),
 * MAX_CONNECT_CELL_NUM (400), N2 (400), N3 (200), AREA_GRID_ORIGINAL (2.0,
 * the bucket edge), USE_TRI_MEMB, NEW_BEND_POT and STOCHASTIC (1), SYSTEM
 * (0).  MALIG_NUM, R_memb, MEMB_ADHE_RANGE, ATP_init and ext_stim_init
 * describe an initial state, which the caller builds; NUM_ITR, CUT,
 * force_cornification and outputdir are the driver's.  ca2p_init,
 * ex_inert_init, IP3_init and gj_init are read by epi_value_init (and
 * ca2p_init by the renewal laws, gj_init by new list entries), agki_max,
 * fac and div_max by epi_load_cells.  Every other key is a value the kernels
 * read, per tissue, from the next call on; epi_step also takes DT_Cell,
 * T_TURNOVER, NUM_SC_INIT and SW_THRESH from it, and a Ca solve runs
 * (int)(Ca_avg_time / DT_Ca) sub-steps.  DT_Cell and DT_Ca must be positive
 * and Ca_avg_time / DT_Ca at least 1. */
typedef struct epi_params {
    int32_t MAX_CONNECT_CELL_NUM;
    double DT_Cell, DT_Ca;
    int32_t MALIG_NUM;
    double LX, LY, LZ;
    int32_t NX, NY, NZ;
    double R_max, R_der, R_memb, COMPRESS_FACTOR;
    int32_t MEMB_ADHE_RANGE;
    double THRESH_SP, LJ_THRESH, THRESH_DEAD;
    int32_t NUM_ITR;
    double Ca_avg_time, ca2p_init, ex_inert_init, IP3_init, gj_init, ATP_init, ext_stim_init, FAC_MAP;
    int32_t SW_THRESH;
    double T_TURNOVER;
    int32_t NUM_SC_INIT, USE_TRI_MEMB;
    double agki_max, fac;
    int32_t div_max;
    double Kpp, dp, kbc, Hb, Cout, kg, gamma, mu0, mu1, kmu, para_b, para_bb, para_k1, k_flux, beta_zero, CA_OUT;
    double thgra, delta_th, thpri, kpa, delta_K, Kpri, delta_I, iage_kitei, para_k2, H0, wd, ca2p_du;
    double STIM11, Kaa, Da, AIR_STIM, AREA_GRID_ORIGINAL;
    int32_t N2, N3;
    double eps_m, DER_DER_CONST, K_TOTAL, K_DESMOSOME_RATIO, Kspring, Kspring_d, KBEND, delta_R_coef, para_ljp2,
        Kspring_division;
    double stoch_div_time_ratio, S2, accel_div, eps_kb, alpha_b, S0, eps_ks, accel_diff, alpha_k, eps_kk, S1;
    double ubar, delta_lipid, delta_sig_r1, lipid_rel, lipid, delta_L_coef, ADHE_CONST, DISA_conn_num_thresh;
    double eps_L, unpair_dist_coef, kb, DUR_ALIVE, DUR_DEAD, DB;
    int32_t SYSTEM, CUT, NEW_BEND_POT, STOCHASTIC, force_cornification;
    char outputdir[256];
} epi_params;
/* epi5_MT's values (the grid and box of a given geometry are filled in by
 * epi_get_params) */
void epi_params_default(epi_params* p);
/* Params::load (epi_intg/Params.cpp:91-96) over parse_paramtext
 * (epi_intg/parser.cpp:30-58): "key=value" lines; a line that starts with
 * '#' or '//', or holds no '=', is skipped; spaces around the key and the
 * value are dropped; the first line of a key wins.  Every key of the struct
 * (or its alias: MALIGNANT, LOOP, Ca2P_avg_time, TURNOVER) that appears sets
 * its field; other keys are ignored.  With require_all a key the file lacks is
 * an error, as in the reference; without it the field keeps its value.  A
 * value that does not read as its type is an error.  err (err_len bytes)
 * gets the message. */
int epi_params_load(const char* path, int32_t require_all, epi_params* p, char* err, int32_t err_len);
/* Params::generate_paramfile (epi_intg/Params.cpp:98-110): one "key=value"
 * line per key in the reference's order */
int epi_params_save(const char* path, const epi_params* p);
/* the parameters of one tissue of a context, read from the next call on;
 * refused (EPI_ERR_INVALID) inside a mechanics or reaction-diffusion window */
int epi_set_params(epi_ctx* ctx, int32_t tissue, const epi_params* p);
int epi_get_params(epi_ctx* ctx, int32_t tissue, epi_params* p);

/* ---- initial state: init_gen (epi_intg/calc/init_gen.cpp:23-106) ---------
 * Host code (no device, no context).  epi_sheet_size gives the membrane
 * sheet's columns and rows of a parameter record (MEMB_NUM_X, MEMB_NUM_Y of
 * CalcParams::calculate_param, CalcParams.cpp:20-26): a geometry's nmx, nmy.
 * epi_init_gen writes the cells of init_gen(nfix, der) into caller arrays of
 * cap entries: the sheet (triangular when USE_TRI_MEMB), der layers of
 * dermis, nfix FIX cells on a ring (fix_origin 0..nfix-1), each FIX then
 * lowered by 0.001 until a membrane cell is within 1.1 (radius + R_der) of
 * it (relaxation, :23-46).  Rows come MEMB, DER, FIX, as epi_load_cells
 * wants them; every other cell field is 0.  out->n, nmemb, nder, nmx, nmy
 * are set even when cap is short (EPI_ERR_CAPACITY).  EPI_ERR_INVALID, with
 * err's message, when a FIX sinks below Z_LEVEL - R_max as the reference
 * throws. */
typedef struct epi_gen_cells {
    int32_t n, nmemb, nder, nmx, nmy;     /* out */
    uint8_t* state; int32_t* fix_origin;  /* caller arrays of cap entries */
    double *x, *y, *z, *radius;
} epi_gen_cells;
int epi_sheet_size(const epi_params* p, int32_t* nmx, int32_t* nmy);
int epi_init_gen(const epi_params* p, int32_t nfix, int32_t der, int32_t cap, epi_gen_cells* out, char* err,
                 int32_t err_len);

/* ---- initialisation: cman_init (cell_init.cpp:132-142) ------------------ */
/* CellManager::_load_from_file + _memb_init (cellmanager.cpp:198-323, 66-113) */
int epi_load_cells(epi_ctx* ctx, int32_t tissue, const epi_cell_input* in);
/* _cman_value_init (cell_init.cpp:18-40) */
int epi_value_init(epi_ctx* ctx, int32_t tissue);
/* load_uvp / load_w (cell_init.cpp:47-109); gj in CSR order of the current lists */
int epi_set_cell_values(epi_ctx* ctx, int32_t tissue, const double* ca2p, const double* ip3,
                        const double* ex_inert, const double* ca2p_avg, const double* gj);
/* Field::read_binary into ATP/ext_stim buffers (proc.cpp:107-110) */
int epi_set_fields(epi_ctx* ctx, int32_t tissue, const double* atp, const double* atp_next,
                   const double* b, const double* b_next);
/* replay stream of genrand_real (Random_gen.cpp:14-54), consumed in the
 * reference's serial order; the library keeps its own device copy */
int epi_set_rng(epi_ctx* ctx, int32_t tissue, const double* draws, int64_t n);
/* std::mt19937_64(seed) through uniform_real_distribution<double>(0,1) */
int epi_rng_mt19937_64(uint64_t seed, double* out, int64_t n);

/* ---- the reference call surface (all tissues of the context) ----------- */
int epi_cell_interaction(epi_ctx* ctx);            /* cell_interaction.h:5, cell_interaction.cpp:648 */
int epi_pos_copy(epi_ctx* ctx);                    /* cellmanager.h:117, cellmanager.cpp:19 */
int epi_cell_state_renew(epi_ctx* ctx);            /* cell_state_renew.h:3, cell_state_renew.cpp:400 */
int epi_cell_pos_periodic_fix(epi_ctx* ctx);       /* cellmanager.h:111, cellmanager.cpp:395 */
int epi_connect_cell(epi_ctx* ctx);                /* cell_connection.h:3, cell_connection.cpp:191 */
int epi_calc_zzmax(epi_ctx* ctx, double* zzmax);   /* proc.cpp:34; zzmax[ntissues] */
int epi_setup_map_lat(epi_ctx* ctx);               /* map.h:5-7, map.cpp:24 */
int epi_calc_ext_stim(epi_ctx* ctx, const double* zzmax); /* ext_stim.h:5, ext_stim.cpp:18 */
int epi_initialize_sc(epi_ctx* ctx, const double* zzmax); /* cell_state_renew.h:4, :435 */
int epi_calc_ca2p(epi_ctx* ctx, const double* zzmax);     /* ca2p.h:7, ca2p.cpp:380 */
/* calc_ca2p with n_sub sub-steps in place of Ca_ITR (the averages are still
 * divided by Ca_ITR).  No counterpart in the reference: timing and tuning
 * samples only. */
int epi_calc_ca2p_n(epi_ctx* ctx, const double* zzmax, int32_t n_sub);
int epi_should_calc_ca(epi_ctx* ctx, int32_t* flags);     /* cellmanager.h:97 */
int epi_ca_calc_condition_reset(epi_ctx* ctx);            /* cellmanager.h:101 */
/* CellManager::current_timestep (cellmanager.h:105, set per step at
 * proc.cpp:132): the timestep the next state renewal records in cell_stat;
 * each renewal (and each step of epi_step) advances it by one, and epi_step
 * sets it from its drivers' step counters */
int epi_set_timestep(epi_ctx* ctx, int32_t tissue, int64_t step);
/* the cst_store of CellManager::append_stat_data (cellmanager.cpp:338-352,
 * 559-567): the cell_stat records (k_aging_start, k_cornified, k_disap; -1 =
 * never set) of the DEAD cells removed since the last call, in removal
 * order, 3 int64 per record; the store is then emptied.  EPI_ERR_CAPACITY
 * (store kept) when more than max_records are held; *n_records gets the
 * count either way.  The store holds up to max_cells records. */
int epi_take_cell_stats(epi_ctx* ctx, int32_t tissue, int64_t* records, int32_t max_records, int32_t* n_records);
/* one sub-step of calc_ca2p's loop (ca2p.cpp:392-399) with a given iz_bound,
 * for kernel-level checks; maps from the last setup_map_lat */
int epi_ca_substep(epi_ctx* ctx, int32_t iz_bound);

/* ---- the driver ---------------------------------------------------------- */
/* nsteps iterations of proc.cpp:153-172 for every tissue; drivers[ntissues]
 * hold each tissue's loop state and are updated */
int epi_step(epi_ctx* ctx, int64_t nsteps, epi_driver* drivers);
int epi_synchronize(epi_ctx* ctx);

/* ---- state exchange ------------------------------------------------------ */
int epi_get_status(epi_ctx* ctx, int32_t tissue, epi_status* out);
/* fills the arrays of *out (n entries; conn/gj up to conn_cap) */
int epi_get_cells(epi_ctx* ctx, int32_t tissue, epi_cells* out);
int epi_get_fields(epi_ctx* ctx, int32_t tissue, double* atp, double* atp_next, double* b,
                   double* b_next, int32_t* cmap1, uint8_t* cmap2);
/* ---- asynchronous snapshots ---------------------------------------------
 * The per-cell record CellManager::output writes (cellmanager.cpp:451-538),
 * taken the way proc.cpp:139-141 takes one every CUT steps, but without
 * stopping the step loop.  epi_snapshot_begin queues, behind the work queued
 * so far (call it between epi_step calls), one kernel that packs tissue's
 * record into the context's snapshot buffer `slot` (0 <= slot <
 * EPI_SNAPSHOT_SLOTS) and its download into pinned host memory on a copy
 * stream of its own; it returns at once, and later epi_step calls run beside
 * the download.  The first call on a slot, or one for a tissue of more cells
 * than the slot has held, allocates the slot's device and pinned buffers.
 * A call on a slot whose previous snapshot was not waited for waits for it.
 * epi_snapshot_wait blocks until the slot's download has landed and points
 * *out at it: the pointers stay valid until the next epi_snapshot_begin on
 * the slot or epi_destroy.  touch is the ALIVE cell's "has a DEAD neighbour"
 * flag of the output record. */
#define EPI_SNAPSHOT_SLOTS 2
typedef struct epi_snapshot {
    int32_t tissue, n;
    const uint8_t *state, *touch;
    const int32_t *rest_div_times, *pair, *fix_origin;
    const double *radius, *ageb, *agek, *ca2p, *x, *y, *z, *ca2p_avg, *ex_fat, *in_fat, *spr_nat_len;
} epi_snapshot;
int epi_snapshot_begin(epi_ctx* ctx, int32_t tissue, int32_t slot);
int epi_snapshot_wait(epi_ctx* ctx, int32_t slot, epi_snapshot* out);

/* kernel-level inputs: positions (both buffers), state/agek, lists, maps */
int epi_set_positions(epi_ctx* ctx, int32_t tissue, const double* x, const double* y, const double* z);
int epi_set_state(epi_ctx* ctx, int32_t tissue, const uint8_t* state, const double* agek);
int epi_set_lists(epi_ctx* ctx, int32_t tissue, const int32_t* conn_ptr, const int32_t* conn,
                  const double* gj);
int epi_set_maps(epi_ctx* ctx, int32_t tissue, const int32_t* cmap1, const uint8_t* cmap2,
                 const int32_t* lat);

/* ---- kernel timing -------------------------------------------------------
 * No counterpart in the reference (its only timer is proc.cpp:128-137).
 * Bit k of mask brackets launches of kernel k with HIP events recorded by
 * the dispatch on the context's stream; totals accumulate until
 * epi_profile_reset. */
int epi_profile_enable(epi_ctx* ctx, uint64_t mask);
/* period 1 (the default): every launch of an enabled kernel carries events.
 * period > 1: runs of `run` consecutive launches (of any kernel) carry events,
 * one run every `period` launches, and the enabled kernels' launches in a run
 * are counted except the first launch of the run.  A dispatch-recorded event
 * pair brackets its kernel only when the launch before it also carries
 * events; sampling keeps that while timing a small share of launches. */
int epi_profile_sample(epi_ctx* ctx, int32_t period, int32_t run);
int epi_profile_read(epi_ctx* ctx, int32_t kernel, const char** name, double* total_ms, int64_t* count);
/* every launch of an enabled kernel since the last reset, timed or not */
int epi_profile_launches(epi_ctx* ctx, int32_t kernel, int64_t* launches);
int epi_profile_reset(epi_ctx* ctx);
int32_t epi_profile_kernel_count(void);

/* ---- frozen mechanics window and its adjoint ------------------------------
 * No counterpart in the reference.  A window is cell_interaction + pos_copy
 * (cell_interaction.cpp:648-685, cellmanager.cpp:19-26) run n_steps times
 * from a snapshot of every tissue's positions, with lists, states, pairs and
 * dermis frozen and no periodic fix, and with the parameters below as
 * runtime values (the reference's values give the plain step).  The
 * Ca-dependent parts of the mechanics (the agek switch between K_DESMOSOME
 * and K_TOTAL, fed by Ca through the ageing rate) are held at the window's
 * start.  epi_mech_end restores the snapshot. */
typedef struct epi_mech_params {
    double eps;       /* eps_m of the LJ laws and the wall (cell_interaction.cpp:21), 0.01 */
    double k_adh;     /* K_TOTAL of adhesion; K_DESMOSOME = 0.01 K_TOTAL (:24-25), 3.0 */
    double k_spring;  /* Kspring of the FIX and FIX-paired MUSUME anchors (:26), 25.0 */
    double k_bend;    /* KBEND of the membrane bending (:28), 0.5 */
} epi_mech_params;
typedef epi_mech_params epi_mech_grad;
int epi_mech_begin(epi_ctx* ctx);
int epi_mech_forward(epi_ctx* ctx, const epi_mech_params* params, int32_t n_steps);
/* positions after the last forward: x[n], y[n], z[n] */
int epi_mech_get_positions(epi_ctx* ctx, int32_t tissue, double* xyz);
/* loss[t] = mean over cells of |x_T - target|^2 (minimum image in x, y),
 * targets[t] = x[n], y[n], z[n]; recompute from checkpoints every
 * checkpoint_every steps */
int epi_mech_adjoint(epi_ctx* ctx, const epi_mech_params* params, int32_t n_steps, int32_t checkpoint_every,
                     const double* const* targets, double* loss, epi_mech_grad* grads);
int epi_mech_end(epi_ctx* ctx);

/* ---- frozen reaction-diffusion window and its adjoint --------------------
 * No counterpart in the reference (it has no gradients).  A window is the
 * sub-step of calc_ca2p (ca2p.cpp:392-399) run n_sub times from a snapshot
 * of every tissue, with positions, lists, states and voxel maps frozen, and
 * with the parameters below as runtime values.  With du = 0.01, dp = 0.1,
 * kpa = NULL (calc_Kpa) and gj0 = 0.99 on fresh edges a window is
 * calc_ca2p's loop.  Division and removal are outside a window (they are not
 * differentiable).  epi_rd_end restores the snapshot. */
typedef struct epi_rd_params {
    double du;          /* ca2p_du, the Ca2+ gap-junction coefficient (ca2p.cpp:183) */
    double dp;          /* dp, the IP3 coefficient (ca2p.cpp:17; supra_calc and dead_IP3_calc) */
    double gj0;         /* starting conductance of every directed edge (gj_init) */
    const double* kpa;  /* per-cell IP3 production (calc_Kpa, ca2p.cpp:101-110); NULL = from agek */
} epi_rd_params;
typedef struct epi_rd_grad {
    double du, dp, gj0;
    double* kpa;        /* n doubles filled when non-NULL */
} epi_rd_grad;
/* freeze the current state of every tissue; zzmax[t] sets iz_bound as in calc_ca2p */
int epi_rd_begin(epi_ctx* ctx, const double* zzmax);
/* n_sub sub-steps from the snapshot with params[t] per tissue */
int epi_rd_forward(epi_ctx* ctx, const epi_rd_params* params, int32_t n_sub);
/* ca2p of every cell after the last forward (n doubles) */
int epi_rd_get_ca2p(epi_ctx* ctx, int32_t tissue, double* out);
/* loss[t] = mean over AFM cells of (ca2p after n_sub - targets[t])^2 and
 * its gradient; recompute from checkpoints every checkpoint_every sub-steps */
int epi_rd_adjoint(epi_ctx* ctx, const epi_rd_params* params, int32_t n_sub, int32_t checkpoint_every,
                   const double* const* targets, double* loss, epi_rd_grad* grads);
int epi_rd_end(epi_ctx* ctx);

/* ---- calc_ext_stim on caller-held grids, allocating nothing ---------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): calc_ext_stim (ext_stim.cpp:18-69) over grids the caller
 * holds, with no context, through the same kernels as epi_calc_ext_stim,
 * with epi5_MT's parameters (epi_params_default).
 * The sweep's grid descriptors, per-grid status and tile work list live in
 * one static device array per GPU of EPI_NOALLOC_SCRATCH_BYTES, held until
 * the process ends, so a call takes at most EPI_NOALLOC_MAX_GRIDS grids.
 * Calls that overlap on one GPU (on different streams, or from different
 * threads) race on that array.  The call allocates nothing (no hipMalloc,
 * hipFree or allocating library routine): it queues four kernels on `stream`
 * (a hipStream_t; NULL = the null stream) and returns.
 *
 * Layout: grid point (j, k, l), 0 <= j <= nx, 0 <= k <= ny, 0 <= l <= nz, is
 * element l*pl + k*xp + j of b_in, b_out, cmap1 and cmap2 (x fastest, xp >=
 * nx + 1, pl >= xp*(ny + 1)); planes x = nx and y = ny hold the periodic
 * copies of x = 0 and y = 0, as in Field<T, NX+1, NY+1, NZ+1>.  b_out is the
 * reference's ext_stim.second(): the caller swaps the two afterwards
 * (ext_stim.cpp:68). */
typedef struct epi_bfield_view {
    int32_t nx, ny, nz, xp;
    int64_t pl;
    double lz;              /* dz = lz / nz */
    double zzmax;           /* calc_zzmax: iz_bound = (int)((zzmax + 2.8) * (nz / lz)) */
    const double* b_in;     /* device: B before the sweep */
    double* b_out;          /* device: B after it, on [0,nx) x [0,ny) x [0,iz_bound) and the copies */
    const int32_t* cmap1;   /* device: cell index per point, -1 = none (setup_map_lat) */
    const uint8_t* cmap2;   /* device: 0 / 1 per point */
    const uint8_t* state;   /* device: per cell */
    const double* agek;     /* device: per cell */
} epi_bfield_view;
enum { EPI_NOALLOC_MAX_GRIDS = 1024 };
/* bytes of the static scratch array of epi_calc_ext_stim_noalloc per GPU */
int64_t epi_noalloc_scratch_bytes(void);
/* views: device array of ngrids views (read by the first kernel); err: device
 * array of ngrids codes, EPI_OK or EPI_ERR_OUT_OF_GRID (iz_bound outside
 * [0, nz]: that grid is left as it was).  Returns EPI_ERR_INVALID for ngrids
 * outside [1, EPI_NOALLOC_MAX_GRIDS] or null pointers, EPI_ERR_HIP when the
 * scratch address or a launch fails. */
int epi_calc_ext_stim_noalloc(const epi_bfield_view* views, int32_t ngrids, int32_t* err, void* stream);

/* ---- launch shapes -------------------------------------------------------
 * No counterpart in the reference.  Launch-shape knobs of the grid kernels,
 * chosen per workload by epi_amd/autotune.py; they change timing only, never
 * results.  The stencils walk a list of tiles over every tissue of a launch;
 * a tile is 64 *_ROWS consecutive points of a tissue's x-y plane (the block)
 * by *_ZB z planes.  Values: *_ZB in {1, 2, 4, 8, 16} (z planes per tile),
 * *_ROWS in {1, 2, 4} (64-lane rows per block), CA_GROUP in {4, 8, 16} (lanes per
 * cell in the Ca cell kernel), XCD in {0, 1} (1: the Ca cell and ATP kernels
 * number their workgroups so that each of the 8 XCDs runs one contiguous run
 * of cells / grid rows), RD_GROUP in {8, 16} (lanes per cell in the reverse
 * cell kernel of reaction-diffusion windows), CA_FUSE in {0, 1, 2} (1: a Ca
 * solve launches the ATP update of a sub-step and the cells of the next one
 * together, the cells computing their ATP corners themselves; 2: only for
 * solves of at most 131072 cells in total), MAP_LANES in
 * {1, 4, 16} (lanes per cell in the voxel-map stamp, splitting its voxels),
 * CA_GROUP_SMALL in {4, 8, 16} (lanes per cell in the solves CA_FUSE 2
 * fuses; those solves' ATP tiles are one plane deep when ATP_ZB is above 2), STEP_GRAPH in {0, 1} (0, the default: epi_step launches the
 * renewal and neighbour-build kernels of a step one by one; 1: it replays
 * them, side stream included, and every whole Ca solve (set-up, sub-steps,
 * averages) from HIP graphs captured once per launch key; whether a step
 * solves is still decided on the host from the step's status read-back),
 * B_XCD in {0, 1} (1: the B stencil's tile list is cut into one contiguous
 * run per XCD), SWEEP_WAVE in {0, 1} (lanes per cell of the neighbour
 * sweep: 0, 32 lanes of five buckets each; 1, one wave of two buckets per
 * lane, for small tissues), SIDE_FORK in {0, 1} (where a step's voxel maps
 * and B sweep fork onto the side stream: 0, after the bucket insertion, beside
 * the neighbour sweep; 1, after the whole build, beside the next step's
 * forces).  EPI_ERR_INVALID for other keys or values. */
enum {
    EPI_TUNE_B_ZB = 0, EPI_TUNE_B_ROWS = 1, EPI_TUNE_ATP_ZB = 2, EPI_TUNE_ATP_ROWS = 3,
    EPI_TUNE_CA_GROUP = 4, EPI_TUNE_XCD = 5, EPI_TUNE_RD_GROUP = 6, EPI_TUNE_CA_FUSE = 7,
    EPI_TUNE_MAP_LANES = 8, EPI_TUNE_CA_GROUP_SMALL = 9, EPI_TUNE_STEP_GRAPH = 10,
    EPI_TUNE_B_XCD = 11, EPI_TUNE_SWEEP_WAVE = 12, EPI_TUNE_SIDE_FORK = 13,
    EPI_TUNE_COUNT = 14
};
int epi_set_tuning(epi_ctx* ctx, int32_t key, int32_t value);
int epi_get_tuning(epi_ctx* ctx, int32_t key, int32_t* value);

#ifdef __cplusplus
}
#endif
#endif
