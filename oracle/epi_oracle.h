/*
 * epi_oracle.h -- CPU restatement of the epi5_MT per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X implementation in epi_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product path never calls
 * it and has no CPU fallback.
 *
 * It restates epi5_MT (apqw/epi) serially, in ascending cell index order,
 * which is the canonical order of every parallel loop of the reference
 * (the reference's own TBB loops are racy; the serial order is what its
 * single-thread execution produces).  Geometry is runtime; the physical
 * constants are the reference's compile-time constants.
 *
 * Pinning: the reference cannot be built here (it needs Intel TBB), so every
 * function is pinned by closed forms, invariants, brute force on small inputs
 * and the shipped input deck; see tests/test_oracle_*.py.
 */
#ifndef EPI_ORACLE_H
#define EPI_ORACLE_H
#include <stdint.h>

#include "../include/epi_amd.h"   /* epi_params, the runtime parameter record */

#ifdef __cplusplus
extern "C" {
#endif

/* cell states, epi5_MT/define.h:36-47 */
enum { EO_ALIVE = 0, EO_DEAD = 1, EO_DISA = 2, EO_UNUSED = 3, EO_FIX = 4,
       EO_BLANK = 5, EO_DER = 6, EO_MUSUME = 7, EO_AIR = 8, EO_MEMB = 9 };

/* status codes (same numbering as include/epi_amd.h) */
enum { EO_OK = 0, EO_ERR_CAPACITY = 1, EO_ERR_STRONG_INTERACTION = 2,
       EO_ERR_NO_DERMIS = 3, EO_ERR_OUT_OF_GRID = 4, EO_ERR_REMOVE_ALIAS = 5,
       EO_ERR_RNG_EXHAUSTED = 6, EO_ERR_INVALID = 7, EO_ERR_PAIR_REMOVED = 9 };

typedef struct eo_geometry {
    int32_t nmx, nmy;        /* membrane columns, rows (nmemb = nmx*nmy) */
    int32_t nx, ny, nz;      /* grid divisions (define.h:104-106) */
    double lx, ly, lz;       /* domain (define.h:97-99) */
    int32_t max_cells;       /* capacity (define.h:83) */
} eo_geometry;

typedef struct eo_tissue eo_tissue;

/* a new tissue starts with epi5_MT's parameters and its geometry's grid and box */
eo_tissue* eo_create(const eo_geometry* g);
/* the runtime parameters: epi5_MT's values (define.h:80-230 and the physics
 * files) and a tissue's record, with the derived quantities of
 * CalcParams::calculate_param (epi_intg/calc/CalcParams.cpp:16-45) */
void eo_params_default(epi_params* p);
void eo_set_params(eo_tissue* t, const epi_params* p);
void eo_get_params(const eo_tissue* t, epi_params* p);
/* init_gen (epi_intg/calc/init_gen.cpp:23-106): the cells into o's arrays
 * of cap entries; EO_ERR_CAPACITY when they do not fit (o->n set) */
int eo_init_gen(const epi_params* p, int nfix, int der, int cap, epi_gen_cells* o);
void eo_destroy(eo_tissue* t);
int eo_error(const eo_tissue* t);
void eo_clear_error(eo_tissue* t);

/* CellManager::create (cellmanager.cpp:366-390); returns the new index or -1 */
int eo_create_cell(eo_tissue* t, int state, int fix_origin, double x, double y,
                   double z, double radius, double ca2p, double ca2p_avg,
                   double ip3, double ex_inert, double agek, double ageb,
                   double ex_fat, double in_fat, double spr_nat_len,
                   int rest_div_times, int is_malignant);
/* pair linking done by _load_from_file (cellmanager.cpp:304-313) */
int eo_link_pair(eo_tissue* t, int a, int b);
/* sets nmemb/nder (cellmanager.cpp:317-322) and runs _memb_init (:66-113) */
int eo_finish_load(eo_tissue* t, int nmemb, int nder);
/* _cman_value_init (cell_init.cpp:18-40) */
void eo_value_init(eo_tissue* t);

/* replayed draws for genrand_real (Random_gen.cpp:14-54) */
void eo_set_rng(eo_tissue* t, const double* draws, int64_t n);
int64_t eo_rng_cursor(const eo_tissue* t);
/* std::mt19937_64(seed) + uniform_real_distribution<double>(0,1) */
void eo_mt19937_64_uniform(uint64_t seed, double* out, int64_t n);
void eo_mt19937_64_raw(uint64_t seed, uint64_t* out, int64_t n);

/* the hot path, one function per reference call */
void eo_cell_interaction(eo_tissue* t);        /* cell_interaction.cpp:648 */
void eo_pos_copy(eo_tissue* t);                /* cellmanager.cpp:19 */
void eo_cell_state_renew(eo_tissue* t);        /* cell_state_renew.cpp:400 */
void eo_cell_pos_periodic_fix(eo_tissue* t);   /* cellmanager.cpp:395 */
void eo_connect_cell(eo_tissue* t);            /* cell_connection.cpp:191 */
double eo_calc_zzmax(eo_tissue* t);            /* proc.cpp:34 */
void eo_setup_map_lat(eo_tissue* t);           /* map.cpp:24 */
void eo_calc_ext_stim(eo_tissue* t, double zzmax); /* ext_stim.cpp:18 */
void eo_initialize_sc(eo_tissue* t, double zzmax); /* cell_state_renew.cpp:435 */
void eo_calc_ca2p(eo_tissue* t, double zzmax);     /* ca2p.cpp:380 */
int eo_should_calc_ca(const eo_tissue* t);     /* cellmanager.h:97 */
void eo_ca_calc_condition_reset(eo_tissue* t); /* cellmanager.h:101 */
void eo_clean_up(eo_tissue* t);                /* cellmanager.cpp:543 */
/* cell_stat accounting: CellManager::current_timestep (cellmanager.h:105;
 * eo_step sets it from the driver), per-cell records (k_aging_start,
 * k_cornified, k_disap; -1 = never set) and append_stat_data's store
 * (cellmanager.cpp:338-352, 559-567) */
void eo_set_timestep(eo_tissue* t, int64_t step);
int64_t eo_timestep(const eo_tissue* t);
void eo_get_cst(const eo_tissue* t, int64_t* out);
int eo_take_cell_stats(eo_tissue* t, int64_t* out, int max_records, int* n_records);

/* Sub-steps of calc_ca2p exposed for kernel-level tests. */
void eo_ca_substep(eo_tissue* t, int iz_bound); /* ca2p.cpp:392-399 body */
/* calc_ca2p with n_sub sub-steps instead of Ca_ITR (timing samples) */
void eo_calc_ca2p_n(eo_tissue* t, double zzmax, int n_sub);

/* One iteration of the proc loop body (proc.cpp:153-172), plus an optional
 * forced Ca cadence: ca_every > 0 also fires calc_ca2p when
 * (step+1) % ca_every == 0.  Returns the number of Ca solves run (0/1). */
typedef struct eo_driver {
    int64_t step;          /* i of proc.cpp:131 */
    int32_t force_cornif;  /* flg_forced_sc */
    int32_t num_sc;
    int32_t ca_every;
    int32_t clean_every;   /* clean_up cadence (proc.cpp:141); 0 = never */
} eo_driver;
int eo_step(eo_tissue* t, eo_driver* d);

/* ---- state exchange (logical order: *_cur is the reference's first()) ---- */
typedef struct eo_cells {
    int32_t n, nmemb, nder, sw;
    uint8_t* state; int32_t* fix_origin;
    double *x, *y, *z, *xn, *yn, *zn;     /* dual_double d1 and d2 */
    double* radius;
    double *ca2p, *ca2p_next, *ip3, *ip3_next;
    double *ca2p_avg, *ex_inert, *agek, *ageb, *ex_fat, *in_fat, *spr_nat_len, *diff_u;
    int32_t *rest_div_times, *pair, *dermis, *lat;  /* lat: 3 per cell */
    uint8_t* is_malignant;
    int64_t* uid;
    int32_t* conn_ptr;   /* n+1 */
    int32_t* conn;       /* conn_ptr[n] entries */
    double* gj;          /* gj value of each list entry */
    int32_t conn_cap;    /* capacity of conn/gj on input */
    int32_t* memb_adj;   /* nmemb*6 */
} eo_cells;
int eo_get_cells(const eo_tissue* t, eo_cells* out);
/* overwrite per-cell values (the load_uvp/load_w resume path,
 * cell_init.cpp:47-109): ca2p/ip3 both buffers, ex_inert, ca2p_avg.
 * gj entries for current list entries are set from gj[] when non-null. */
int eo_set_cell_values(eo_tissue* t, const double* ca2p, const double* ip3,
                       const double* ex_inert, const double* ca2p_avg,
                       const double* gj);

/* fields, layout [x][y][z] with z fastest, (nx+1)*(ny+1)*(nz+1) each */
int eo_get_fields(const eo_tissue* t, double* atp, double* atp_next,
                  double* b, double* b_next, int32_t* cmap1, uint8_t* cmap2);
int eo_set_fields(eo_tissue* t, const double* atp, const double* atp_next,
                  const double* b, const double* b_next);
int eo_set_maps(eo_tissue* t, const int32_t* cmap1, const uint8_t* cmap2,
                const int32_t* lat);
/* per-cell positions only (both buffers), for kernel-level tests */
int eo_set_positions(eo_tissue* t, const double* x, const double* y, const double* z);
int eo_set_state(eo_tissue* t, const uint8_t* state, const double* agek);
/* replace neighbour lists and gj (CSR) without running connect_cell */
int eo_set_lists(eo_tissue* t, const int32_t* conn_ptr, const int32_t* conn,
                 const double* gj);

/* ---- frozen mechanics window (cell_interaction + pos_copy with lists,
 * states, pairs and dermis frozen, no periodic fix) and its adjoint ------- */
typedef struct eo_mech_params {
    double eps;       /* eps_m of the LJ laws and the wall (cell_interaction.cpp:21), 0.01 */
    double k_adh;     /* K_TOTAL of adhesion, K_DESMOSOME = 0.01 K_TOTAL (:24-25), 3.0 */
    double k_spring;  /* Kspring of the FIX and FIX-paired MUSUME anchors (:26), 25.0 */
    double k_bend;    /* KBEND of the membrane bending (:28), 0.5 */
} eo_mech_params;
typedef eo_mech_params eo_mech_grad;
/* n_steps steps from the current positions (left unchanged); positions
 * after the window into xyz_out (x[n], y[n], z[n]) */
int eo_mech_forward(eo_tissue* t, const eo_mech_params* p, int n_steps, double* xyz_out);
/* loss = mean over cells of |x_T - target|^2 (minimum image in x, y) and
 * its gradient with respect to the parameters */
int eo_mech_adjoint(eo_tissue* t, const eo_mech_params* p, int n_steps, const double* target_xyz,
                    double* loss, eo_mech_grad* grad);

/* ---- frozen reaction-diffusion window (the calc_ca2p sub-step with
 * positions, lists, states and maps frozen) and its adjoint ---------------- */
typedef struct eo_rd_params {
    double du;          /* ca2p_du (ca2p.cpp:183), 0.01 in the reference */
    double dp;          /* dp (ca2p.cpp:17), 0.1 */
    double gj0;         /* starting gj of every directed edge, gj_init 0.99 */
    const double* kpa;  /* per-cell IP3 production (calc_Kpa); NULL = calc_Kpa(agek) */
} eo_rd_params;
typedef struct eo_rd_grad { double du, dp, gj0; double* kpa; /* n, may be NULL */ } eo_rd_grad;
/* n_sub sub-steps from the tissue's current state (left unchanged); the
 * final ca2p of every cell into ca2p_out (n) */
int eo_rd_forward(eo_tissue* t, double zzmax, const eo_rd_params* p, int n_sub, double* ca2p_out);
/* loss = mean over AFM cells of (ca2p after n_sub - target)^2 and its
 * gradient with respect to the parameters */
int eo_rd_adjoint(eo_tissue* t, double zzmax, const eo_rd_params* p, int n_sub, const double* target,
                  double* loss, eo_rd_grad* grad);

/* helpers used by the pinning tests */
double eo_p_diff_x(const eo_tissue* t, double a, double b);
double eo_coef(eo_tissue* t, int law, int c1, int c2);
void eo_tri_bend(const eo_tissue* t, int m, double* out3);

#ifdef __cplusplus
}
#endif
#endif
