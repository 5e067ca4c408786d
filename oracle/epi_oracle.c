/*
 * epi_oracle.c -- serial CPU restatement of the epi5_MT per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see epi_oracle.h).  Every function names the
 * reference lines it follows.  Arithmetic is written with the same operand
 * order as the reference so that an IEEE double build without FP contraction
 * (gcc -O2 -ffp-contract=off, no -march) reproduces the reference's results;
 * tanh/sin/cos/sqrt come from libm like the reference.
 *
 * Representation: cells are addressed by index, like CellManager's slots.
 * Pointer identity (gj keys, Cell::pair) is represented by a per-cell uid
 * that follows the cell through remove_exec's migration.
 */
#include "epi_oracle.h"
#include "../include/epi_amd.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* EO_OMP builds the multithreaded form used as the CPU throughput baseline:
 * the loops that epi5_MT runs as tbb::parallel_for / parallel_for_each
 * (cellmanager.h:28-35, cell_connection.cpp:41,93,135,179, cell_interaction.cpp:
 * 503-515,652-656, cell_state_renew.cpp:402,422, map.cpp:35-42,156,
 * ext_stim.cpp:26,63, ca2p.cpp:158,186,241,311) run over OpenMP threads, with
 * the reference's atomics: CAS-added position updates (atomics.h:42-65),
 * fetch-add list pushes, a locked cell creation and replay-stream cursor.
 * Like the reference under TBB, its list order, accumulation order and
 * daughter numbering then depend on the thread schedule, so only the
 * default (serial) build is the parity checker. */
#ifdef EO_OMP
#include <omp.h>
#define EO_PAR _Pragma("omp parallel for schedule(static)")
#define EO_PAR_DYN _Pragma("omp parallel for schedule(dynamic, 64)")
#define EO_PAR2 _Pragma("omp parallel for collapse(2) schedule(static)")
#define EO_CRITICAL _Pragma("omp critical(eo_create)")
#else
#define EO_PAR
#define EO_PAR_DYN
#define EO_PAR2
#define EO_CRITICAL
#endif

/* d += v: a compare-and-swap loop when threads share d (atomic_double's
 * operator+=, atomics.h:42-65), plain addition in the serial build */
static inline void add_d(double* d, double v) {
#ifdef EO_OMP
    uint64_t old = __atomic_load_n((uint64_t*)d, __ATOMIC_RELAXED), nw;
    do {
        double x;
        memcpy(&x, &old, 8);
        x += v;
        memcpy(&nw, &x, 8);
    } while (!__atomic_compare_exchange_n((uint64_t*)d, &old, nw, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
#else
    *d += v;
#endif
}

/* ---- the parameter record ------------------------------------------------
 * These functions name the record's fields, so they sit before the names
 * below that stand for them. */
typedef struct eo_prm {
    epi_params in;
    double P_MEMB, agki_max_fix, leak_from_storage, delta_R, delta_L, iz_margin;
    int Ca_ITR;
} eo_prm;

/* CalcParams::calculate_param (CalcParams.cpp:16-45) */
static void derive(eo_prm* P, const epi_params* in) {
    P->in = *in;
    P->P_MEMB = 1.0 / in->COMPRESS_FACTOR;
    P->agki_max_fix = in->fac * in->agki_max;
    P->leak_from_storage = in->CA_OUT * in->beta_zero;
    P->delta_R = in->delta_R_coef * in->R_der;
    P->delta_L = in->delta_L_coef * in->R_max;
    P->iz_margin = in->FAC_MAP * in->R_max;
    P->Ca_ITR = (int)(in->Ca_avg_time / in->DT_Ca);
}

/* epi5_MT's constants: define.h:80-230, cell.h:23-32, ca2p.cpp:16-144,
 * cell_connection.cpp:16-30, cell_interaction.cpp:21-45,
 * cell_state_renew.cpp:17-119, ext_stim.cpp:9-20 */
void eo_params_default(epi_params* p) {
    memset(p, 0, sizeof *p);
    p->MAX_CONNECT_CELL_NUM = 400; p->DT_Cell = 0.01; p->DT_Ca = 0.01; p->MALIG_NUM = 0;
    p->LX = 50.0; p->LY = 50.0; p->LZ = 100.0; p->NX = 100; p->NY = 100; p->NZ = 200;
    p->R_max = 1.4; p->R_der = 1.4; p->R_memb = 1.0; p->COMPRESS_FACTOR = 4; p->MEMB_ADHE_RANGE = 0;
    p->THRESH_SP = 3.0; p->LJ_THRESH = 1.2; p->THRESH_DEAD = 22.0; p->NUM_ITR = 4000000;
    p->Ca_avg_time = 10.0; p->ca2p_init = 0.122; p->ex_inert_init = 0.97; p->IP3_init = 0; p->gj_init = 0.99;
    p->ATP_init = 0; p->ext_stim_init = 0; p->FAC_MAP = 2.0; p->SW_THRESH = 20; p->T_TURNOVER = 6000.0;
    p->NUM_SC_INIT = 1; p->USE_TRI_MEMB = 1;
    p->agki_max = 6.0; p->fac = 1; p->div_max = 15;
    p->Kpp = 0.3; p->dp = 0.1; p->kbc = 0.4 * 1.2; p->Hb = 0.01; p->Cout = 1.0; p->kg = 0.1; p->gamma = 2.0;
    p->mu0 = 0.567; p->mu1 = 0.1; p->kmu = 0.05; p->para_b = 0.11; p->para_bb = 0.89; p->para_k1 = 0.7;
    p->k_flux = 8.1; p->beta_zero = 0.02; p->CA_OUT = 1.0; p->thgra = 0.2; p->delta_th = 1.0; p->thpri = 1.0;
    p->kpa = 4.0; p->delta_K = 1.0; p->Kpri = 6.0; p->delta_I = 1.5; p->iage_kitei = 0; p->para_k2 = 0.7;
    p->H0 = 0.5; p->wd = 0.1; p->ca2p_du = 0.01; p->STIM11 = 0.002; p->Kaa = 0.5; p->Da = 1.0; p->AIR_STIM = 0.1;
    p->AREA_GRID_ORIGINAL = 2.0; p->N2 = 400; p->N3 = 200;
    p->eps_m = 0.01; p->DER_DER_CONST = 0.2; p->K_TOTAL = 3.0; p->K_DESMOSOME_RATIO = 0.01; p->Kspring = 25.0;
    p->Kspring_d = 5.0; p->KBEND = 0.5; p->delta_R_coef = 0.4; p->para_ljp2 = 0.005; p->Kspring_division = 5.0;
    p->stoch_div_time_ratio = 0.25; p->S2 = 0.1; p->accel_div = 1.0; p->eps_kb = 0.12; p->alpha_b = 5.0;
    p->S0 = 0.1 * 0.2; p->eps_ks = 0.10 * 0.5; p->accel_diff = 1.0; p->alpha_k = 2.0; p->eps_kk = 0.10 * 0.5;
    p->S1 = 0.1; p->ubar = 0.45; p->delta_lipid = 0.1; p->delta_sig_r1 = 0.1; p->lipid_rel = 0.05 * 4.0;
    p->lipid = 0.6; p->delta_L_coef = 0.01; p->ADHE_CONST = 31.3; p->DISA_conn_num_thresh = 11;
    p->eps_L = 0.14; p->unpair_dist_coef = 0.9;
    p->kb = 0.025; p->DUR_ALIVE = 0.5; p->DUR_DEAD = 2.0; p->DB = 0.0009;
    p->SYSTEM = 0; p->CUT = 1000; p->NEW_BEND_POT = 1; p->STOCHASTIC = 1; p->force_cornification = 1;
    strcpy(p->outputdir, "output");
}

/* init_gen (epi_intg/calc/init_gen.cpp:48-106) with relaxation (:23-46):
 * the sheet, der dermis layers and a ring of nfix FIX cells, each FIX
 * lowered by 0.001 until within 1.1 (radius + R_der) of a membrane cell.
 * MEMB_NUM_X / MEMB_NUM_Y from calculate_param (CalcParams.cpp:20-26). */
static double gen_pdiff(double a, double b, double L) {
    double d = a - b;
    if (d > 0.5 * L) return d - L;
    if (d <= -0.5 * L) return d + L;
    return d;
}
int eo_init_gen(const epi_params* p, int nfix, int der, int cap, epi_gen_cells* o) {
    const double P_MEMB_ = 1.0 / p->COMPRESS_FACTOR, pratio = 2.0 * P_MEMB_;
    const double ytri = 2.0 / sqrt(3.0);
    unsigned mx = (unsigned)(p->LX / (p->R_memb * pratio));
    unsigned my = (unsigned)(p->LY / (p->R_memb * pratio));
    if (p->USE_TRI_MEMB) my = ((uint32_t)(my * ytri)) & 0xfffffffeu;
    const double zl = der * 2.0 * p->R_der;
    const unsigned ndx = (unsigned)(p->LX / (2.0 * p->R_der)), ndy = (unsigned)(p->LY / (2.0 * p->R_der));
    const int n = (int)(mx * my + (unsigned)der * ndy * ndx + (unsigned)nfix);
    o->nmx = (int)mx; o->nmy = (int)my; o->nmemb = (int)(mx * my); o->nder = (int)((unsigned)der * ndy * ndx);
    o->n = n;
    if (n > cap) return EO_ERR_CAPACITY;
    int c = 0;
    for (unsigned k = 0; k < my; k++)
        for (unsigned j = 0; j < mx; j++, c++) {
            const double shift = (p->USE_TRI_MEMB && k % 2 == 1) ? p->R_memb * pratio * 0.5 : 0.0;
            o->state[c] = EO_MEMB; o->fix_origin[c] = 0; o->radius[c] = p->R_memb;
            o->x[c] = 0.5 * p->R_memb + j * p->R_memb * pratio + shift;
            o->y[c] = p->USE_TRI_MEMB ? 0.5 * p->R_memb + k * p->R_memb * pratio / ytri
                                      : 0.5 * p->R_memb + k * p->R_memb * pratio;
            o->z[c] = zl + p->R_memb;
        }
    for (int l = 0; l < der; l++)
        for (unsigned k = 0; k < ndy; k++)
            for (unsigned j = 0; j < ndx; j++, c++) {
                o->state[c] = EO_DER; o->fix_origin[c] = 0; o->radius[c] = p->R_der;
                o->x[c] = p->R_der + 2.0 * j * p->R_der;
                o->y[c] = p->R_der + 2.0 * k * p->R_der;
                o->z[c] = p->R_der + 2.0 * l * p->R_der;
            }
    const int nb = c;
    for (int l = 0; l < nfix; l++, c++) {
        o->state[c] = EO_FIX; o->fix_origin[c] = l; o->radius[c] = p->R_max;
        o->x[c] = p->LX / 2.0 + (p->LX / 4.0) * cos(2.0 * 3.14159265358979323846 * l / nfix);
        o->y[c] = p->LY / 2.0 + (p->LY / 4.0) * sin(2.0 * 3.14159265358979323846 * l / nfix);
        o->z[c] = p->R_max + p->R_memb + zl + p->R_memb;
    }
    for (int f = nb; f < n; f++) {
        double dmin = 10000.0;
        while (1) {
            for (int l = 0; l < nb && o->state[l] == EO_MEMB; l++) {
                const double dx = gen_pdiff(o->x[f], o->x[l], p->LX), dy = gen_pdiff(o->y[f], o->y[l], p->LY);
                const double dz = o->z[f] - o->z[l];
                const double d = sqrt(dx * dx + dy * dy + dz * dz);
                if (d < dmin) dmin = d;
            }
            if (dmin < 1.1 * (o->radius[f] + p->R_der)) break;
            if (o->z[f] < zl - p->R_max) return EO_ERR_INVALID;
            o->z[f] = o->z[f] - 0.001;
        }
    }
    return EO_OK;
}

/* the mechanics coefficients of a record (cell_interaction.cpp:21-28) */
static void mp_from(const eo_prm* P, eo_mech_params* mp) {
    mp->eps = P->in.eps_m; mp->k_adh = P->in.K_TOTAL; mp->k_spring = P->in.Kspring; mp->k_bend = P->in.KBEND;
}
static unsigned prm_sw_thresh(const eo_prm* P) { return (unsigned)P->in.SW_THRESH; }

/* ---- parameters --------------------------------------------------------
 * The physics constants are runtime values, one record per tissue (the
 * epi_params struct of include/epi_amd.h, keys of epi_intg's CalcParams),
 * with the derived quantities of CalcParams::calculate_param
 * (epi_intg/calc/CalcParams.cpp:16-45).  Every exported function points g_P
 * at its tissue's record first (use()); the names below read it. */
static const eo_prm* g_P;
#define DT_Cell (g_P->in.DT_Cell)
#define DT_Ca (g_P->in.DT_Ca)
#define R_max (g_P->in.R_max)
#define R_der (g_P->in.R_der)
#define THRESH_SP (g_P->in.THRESH_SP)
#define LJ_THRESH (g_P->in.LJ_THRESH)
#define THRESH_DEAD (g_P->in.THRESH_DEAD)
#define ca2p_init (g_P->in.ca2p_init)
#define gj_init (g_P->in.gj_init)
#define FAC_MAP (g_P->in.FAC_MAP)
#define SW_THRESH (g_P->in.SW_THRESH)
#define T_TURNOVER (g_P->in.T_TURNOVER)
#define NUM_SC_INIT ((int)g_P->in.NUM_SC_INIT)
#define P_MEMB (g_P->P_MEMB)
#define DER_DER_CONST (g_P->in.DER_DER_CONST)
#define K_DESMOSOME_RATIO (g_P->in.K_DESMOSOME_RATIO)
#define K_DESMOSOME (g_mp.k_adh * K_DESMOSOME_RATIO)
#define Kspring_d (g_P->in.Kspring_d)
#define delta_R (g_P->delta_R)
#define para_ljp2 (g_P->in.para_ljp2)
#define Kspring_division (g_P->in.Kspring_division)
#define agki_max (g_P->in.agki_max)
#define agki_max_fix (g_P->agki_max_fix)
#define div_max ((int)g_P->in.div_max)
/* the shape of the build */
#define stoch_corr_coef 1.0           /* define.h:195 */
#define MEMB_ADJ_CONN_NUM 6           /* define.h:147-153 (TRI_MEMB) */
#define MAX_CONNECT_CELL_NUM 400      /* define.h:80, cell_connection.cpp:30 */
#define N3 200                        /* cell_connection.cpp:28 */
#define AREA_GRID_ORIGINAL 2.0        /* cell_connection.cpp:16 */
#define AREA_GRID (AREA_GRID_ORIGINAL + 1e-7) /* cell_connection.cpp:19 */

#define POW6(x) ((x) * (x) * (x) * (x) * (x) * (x))
#define POW3(x) ((x) * (x) * (x))
#define DIST_SQ(x, y, z) ((x) * (x) + (y) * (y) + (z) * (z))

/* force laws (cell_interaction.cpp:68-76) */
enum { LAW_NONE = 0, LAW_OTHER, LAW_DER_DER, LAW_FIXMU, LAW_FIX_MEMB,
       LAW_MU_MEMB, LAW_ALAIRDE, LAW_MEMB_MEMB, LAW_PAIR };

/* eps_m, K_TOTAL, Kspring and KBEND as the force laws use them: the
 * tissue's parameters, or a mechanics window's while one runs (g_window) */
static eo_mech_params g_mp;
static int g_window;
#define eps_m (g_mp.eps)
#define K_TOTAL (g_mp.k_adh)
#define Kspring (g_mp.k_spring)
#define KBEND (g_mp.k_bend)

typedef struct { int64_t uid; double v; uint8_t checked; } gj_ent;
typedef struct { gj_ent* e; int n, cap; } gj_map;

struct eo_tissue {
    eo_geometry g;
    double dx, dy, dz, inv_dx, inv_dy, inv_dz;
    int ANX, ANY, ANZ, NX, NY, NZ;
    int n, nmemb, nder, cap;
    int err;
    /* per cell */
    uint8_t* state; int32_t* fix_origin;
    double *x1, *y1, *z1, *x2, *y2, *z2;
    double* radius; double* div_age_thresh;
    double *ca2p_avg, *ex_inert, *agek, *ageb, *ex_fat, *in_fat, *spr_nat_len,
        *diff_u, *spring_force_to_memb;
    int32_t *rest_div_times, *pair, *dermis, *lat;
    uint8_t* is_malignant;
    int64_t* uid; int64_t next_uid;
    int32_t* conn; int32_t* nconn;  /* MAX_CONNECT_CELL_NUM per cell */
    gj_map* gj;
    int32_t* memb_adj;              /* memb[6] of Memb_data (cell.h:58) */
    double* ca_buf[2]; double* ip3_buf[2]; int ca_first;  /* SwapData */
    /* grids [x][y][z], z fastest */
    size_t gsz;
    double* atp[2]; int atp_first;
    double* bst[2]; int b_first;
    int32_t* cmap1; uint8_t* cmap2;
    /* bucket grid, cell_connection.cpp:194-195 */
    int32_t* aindx; int32_t* area;
    /* remove queue */
    int32_t* rq; int nrq;
    uint8_t sw;                     /* std::atomic<uint_fast8_t> (cellmanager.h:56) */
    /* cell_stat (define.h:198-202) per cell: k_aging_start, k_cornified,
     * k_disap timesteps, -1 = never set (the reference leaves the fields of a
     * new Cell uninitialised); cst_store (cellmanager.h:106) of removed DEAD
     * cells, nstat records; current_timestep (cellmanager.h:105) */
    int64_t* cst; int64_t* stat; int nstat;
    int64_t timestep;
    eo_prm P;                       /* runtime parameters */
    const double* rng; int64_t rng_n, rng_pos;
    /* periodic index tables, utils.cpp:32-72 */
    int *px_prev, *px_next, *py_prev, *py_next, *pz_prev, *pz_next;
};

/* the record the parameter names read, and the mechanics coefficients of
 * the tissue unless a window runs */
static void use(const eo_tissue* t) {
    g_P = &t->P;
    if (!g_window) mp_from(&t->P, &g_mp);
}

void eo_set_params(eo_tissue* t, const epi_params* p) { derive(&t->P, p); }
void eo_get_params(const eo_tissue* t, epi_params* p) { *p = t->P.in; }

#define IDX3(t, j, k, l) (((size_t)(j) * ((t)->NY + 1) + (size_t)(k)) * ((t)->NZ + 1) + (size_t)(l))

static void set_err(eo_tissue* t, int e) { if (t->err == EO_OK) t->err = e; }

/* ---- geometry helpers: utils.h:25-53 ---------------------------------- */
static inline double p_diff_x(const eo_tissue* t, double x1, double x2) {
    const double LX = t->g.lx;
    const double diff = x1 - x2;
    if (diff > 0.5 * LX) return diff - LX;
    if (diff <= -0.5 * LX) return diff + LX;
    return diff;
}
static inline double p_diff_y(const eo_tissue* t, double y1, double y2) {
    const double LY = t->g.ly;
    const double diff = y1 - y2;
    if (diff > 0.5 * LY) return diff - LY;
    if (diff <= -0.5 * LY) return diff + LY;
    return diff;
}
static inline double p_dist_sq(const eo_tissue* t, double x1, double y1, double z1,
                               double x2, double y2, double z2) {
    const double diffx = p_diff_x(t, x1, x2);
    const double diffy = p_diff_y(t, y1, y2);
    const double diffz = z1 - z2;
    return diffx * diffx + diffy * diffy + diffz * diffz;
}
static inline double min0(double a) { return a > 0 ? a : 0; }
/* cell.h:287-290 (d1 positions) */
static inline double cell_dist_sq(const eo_tissue* t, int a, int b) {
    return p_dist_sq(t, t->x1[a], t->y1[a], t->z1[a], t->x1[b], t->y1[b], t->z1[b]);
}

/* ---- construction ------------------------------------------------------ */
static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s); return p; }

eo_tissue* eo_create(const eo_geometry* g) {
    eo_tissue* t = (eo_tissue*)xcalloc(1, sizeof(eo_tissue));
    t->g = *g;
    {
        epi_params d;
        eo_params_default(&d);
        d.LX = g->lx; d.LY = g->ly; d.LZ = g->lz; d.NX = g->nx; d.NY = g->ny; d.NZ = g->nz;
        derive(&t->P, &d);
    }
    t->NX = g->nx; t->NY = g->ny; t->NZ = g->nz;
    /* define.h:111-113 */
    t->dx = g->lx / g->nx; t->inv_dx = g->nx / g->lx;
    t->dy = g->ly / g->ny; t->inv_dy = g->ny / g->ly;
    t->dz = g->lz / g->nz; t->inv_dz = g->nz / g->lz;
    /* cell_connection.cpp:22-26 */
    t->ANX = (int)((double)g->lx / AREA_GRID_ORIGINAL + 0.5);
    t->ANY = (int)((double)g->ly / AREA_GRID_ORIGINAL + 0.5);
    t->ANZ = (int)((double)g->lz / AREA_GRID_ORIGINAL);
    int cap = g->max_cells; t->cap = cap;
    t->state = xcalloc(cap, 1); t->fix_origin = xcalloc(cap, 4);
    t->x1 = xcalloc(cap, 8); t->y1 = xcalloc(cap, 8); t->z1 = xcalloc(cap, 8);
    t->x2 = xcalloc(cap, 8); t->y2 = xcalloc(cap, 8); t->z2 = xcalloc(cap, 8);
    t->radius = xcalloc(cap, 8); t->div_age_thresh = xcalloc(cap, 8);
    t->ca2p_avg = xcalloc(cap, 8); t->ex_inert = xcalloc(cap, 8);
    t->agek = xcalloc(cap, 8); t->ageb = xcalloc(cap, 8);
    t->ex_fat = xcalloc(cap, 8); t->in_fat = xcalloc(cap, 8);
    t->spr_nat_len = xcalloc(cap, 8); t->diff_u = xcalloc(cap, 8);
    t->spring_force_to_memb = xcalloc(cap, 8);
    t->rest_div_times = xcalloc(cap, 4); t->pair = xcalloc(cap, 4);
    t->dermis = xcalloc(cap, 4); t->lat = xcalloc((size_t)cap * 3, 4);
    t->is_malignant = xcalloc(cap, 1); t->uid = xcalloc(cap, 8);
    t->conn = xcalloc((size_t)cap * MAX_CONNECT_CELL_NUM, 4);
    t->nconn = xcalloc(cap, 4);
    t->gj = xcalloc(cap, sizeof(gj_map));
    t->memb_adj = xcalloc((size_t)g->nmx * g->nmy * 6, 4);
    for (int i = 0; i < 2; i++) { t->ca_buf[i] = xcalloc(cap, 8); t->ip3_buf[i] = xcalloc(cap, 8); }
    t->gsz = (size_t)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    for (int i = 0; i < 2; i++) { t->atp[i] = xcalloc(t->gsz, 8); t->bst[i] = xcalloc(t->gsz, 8); }
    t->cmap1 = xcalloc(t->gsz, 4); t->cmap2 = xcalloc(t->gsz, 1);
    for (size_t i = 0; i < t->gsz; i++) t->cmap1[i] = -1;
    t->aindx = xcalloc((size_t)t->ANX * t->ANY * t->ANZ, 4);
    t->area = xcalloc((size_t)t->ANX * t->ANY * t->ANZ * N3, 4);
    t->rq = xcalloc(cap, 4);
    t->cst = xcalloc((size_t)cap * 3, 8); t->stat = xcalloc((size_t)cap * 3, 8);
    /* utils.cpp:32-72 */
    int NX = t->NX, NY = t->NY, NZ = t->NZ;
    t->px_prev = xcalloc(NX, sizeof(int)); t->px_next = xcalloc(NX, sizeof(int));
    t->py_prev = xcalloc(NY, sizeof(int)); t->py_next = xcalloc(NY, sizeof(int));
    t->pz_prev = xcalloc(NZ, sizeof(int)); t->pz_next = xcalloc(NZ, sizeof(int));
    for (int k = 0; k < NX; k++) {
        int prev_x = k - 1, next_x = k + 1;
        if (prev_x < 0) prev_x += NX;
        if (next_x >= NX) next_x -= NX;
        t->px_prev[k] = prev_x; t->px_next[k] = next_x;
    }
    for (int k = 0; k < NY; k++) {
        int prev_y = k - 1, next_y = k + 1;
        if (prev_y < 0) prev_y += NY;
        if (next_y >= NY) next_y -= NY;
        t->py_prev[k] = prev_y; t->py_next[k] = next_y;
    }
    for (int l = 0; l < NZ; l++) {
        int prev_z = l - 1, next_z = l + 1;
        if (prev_z == -1) prev_z = 1;
        if (next_z == NZ) next_z = NZ - 1;
        t->pz_prev[l] = prev_z; t->pz_next[l] = next_z;
    }
    return t;
}

void eo_destroy(eo_tissue* t) {
    if (!t) return;
    void* ptrs[] = {t->state, t->fix_origin, t->x1, t->y1, t->z1, t->x2, t->y2, t->z2,
                    t->radius, t->div_age_thresh, t->ca2p_avg, t->ex_inert, t->agek,
                    t->ageb, t->ex_fat, t->in_fat, t->spr_nat_len, t->diff_u,
                    t->spring_force_to_memb, t->rest_div_times, t->pair, t->dermis,
                    t->lat, t->is_malignant, t->uid, t->conn, t->nconn, t->memb_adj,
                    t->ca_buf[0], t->ca_buf[1], t->ip3_buf[0], t->ip3_buf[1],
                    t->atp[0], t->atp[1], t->bst[0], t->bst[1], t->cmap1, t->cmap2,
                    t->aindx, t->area, t->rq, t->px_prev, t->px_next, t->py_prev,
                    t->py_next, t->pz_prev, t->pz_next, t->cst, t->stat};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    for (int i = 0; i < t->cap; i++) free(t->gj[i].e);
    free(t->gj);
    free(t);
}

int eo_error(const eo_tissue* t) { return t->err; }
void eo_clear_error(eo_tissue* t) { t->err = EO_OK; }

/* cell.cpp:11-20 */
static double get_div_age_thresh(int state) {
    return state == EO_FIX ? agki_max_fix : state == EO_MUSUME ? agki_max : 0;
}

/* CellManager::create, cellmanager.cpp:366-390; Cell ctor cell.cpp:58-84 */
int eo_create_cell(eo_tissue* t, int state, int fix_origin, double x, double y,
                   double z, double radius, double ca2p, double ca2p_avg,
                   double ip3, double ex_inert, double agek, double ageb,
                   double ex_fat, double in_fat, double spr_nat_len,
                   int rest_div_times, int is_malignant) {
    use(t);
    if (t->n >= t->cap) { set_err(t, EO_ERR_CAPACITY); return -1; }
    int i = t->n++;
    t->state[i] = (uint8_t)state;
    t->fix_origin[i] = (state == EO_DER || state == EO_MEMB) ? -1 : fix_origin;
    t->x1[i] = t->x2[i] = x; t->y1[i] = t->y2[i] = y; t->z1[i] = t->z2[i] = z;
    t->ca2p_avg[i] = ca2p_avg; t->ex_inert[i] = ex_inert; t->agek[i] = agek;
    t->ageb[i] = ageb; t->ex_fat[i] = ex_fat; t->in_fat[i] = in_fat;
    t->spr_nat_len[i] = spr_nat_len; t->radius[i] = radius;
    t->div_age_thresh[i] = get_div_age_thresh(state);
    t->rest_div_times[i] = state == EO_FIX ? div_max : rest_div_times;
    t->is_malignant[i] = (uint8_t)is_malignant;
    t->diff_u[i] = 0; t->spring_force_to_memb[i] = 0;
    t->pair[i] = -1; t->dermis[i] = -1; t->nconn[i] = 0;
    t->lat[3 * i] = t->lat[3 * i + 1] = t->lat[3 * i + 2] = 0;
    t->gj[i].n = 0;
    t->uid[i] = t->next_uid++;
    t->cst[3 * i] = t->cst[3 * i + 1] = t->cst[3 * i + 2] = -1;
    t->ca_buf[0][i] = t->ca_buf[1][i] = ca2p;      /* ca2p._set */
    t->ip3_buf[0][i] = t->ip3_buf[1][i] = ip3;     /* IP3._set */
    return i;
}

int eo_link_pair(eo_tissue* t, int a, int b) {
    use(t);
    if (a < 0 || b < 0 || a >= t->n || b >= t->n) return EO_ERR_INVALID;
    t->pair[a] = b; t->pair[b] = a;
    return EO_OK;
}

/* _memb_init, cellmanager.cpp:66-113 (TRI_MEMB branch) */
int eo_finish_load(eo_tissue* t, int nmemb, int nder) {
    use(t);
    const int NMX = t->g.nmx;
    if (nmemb != t->g.nmx * t->g.nmy || (t->g.nmy % 2) != 0) return EO_ERR_INVALID;
    t->nmemb = nmemb; t->nder = nder;
    int32_t* m = t->memb_adj;
    for (int j = 0; j < nmemb; j++) {
        int jj = j % NMX;
        m[6 * j + 3] = (jj == 0) ? j + NMX - 1 : j - 1;
        m[6 * j + 0] = (jj == NMX - 1) ? j - (NMX - 1) : j + 1;
    }
    for (int j = 0; j < nmemb; j++) {
        int kk = j / NMX;
        int top = (j - NMX + nmemb) % nmemb;
        int bot = (j + NMX) % nmemb;
        if (kk % 2 == 0) {
            m[6 * j + 1] = top;
            m[6 * j + 2] = m[6 * top + 3];
            m[6 * j + 4] = m[6 * bot + 3];
            m[6 * j + 5] = bot;
        } else {
            m[6 * j + 1] = m[6 * top + 0];
            m[6 * j + 2] = top;
            m[6 * j + 4] = bot;
            m[6 * j + 5] = m[6 * bot + 0];
        }
        for (int i = 0; i < 6; i++) t->conn[(size_t)j * MAX_CONNECT_CELL_NUM + i] = m[6 * j + i];
        t->nconn[j] = 6;
    }
    return EO_OK;
}

/* ---- gj map (std::unordered_map<const Cell*, gj_value>) --------------- */
static gj_ent* gj_find(gj_map* m, int64_t uid) {
    for (int i = 0; i < m->n; i++) if (m->e[i].uid == uid) return &m->e[i];
    return NULL;
}
/* operator[]: default-constructs gj_value (= gj_init, unchecked) if absent */
static gj_ent* gj_get(gj_map* m, int64_t uid) {
    gj_ent* e = gj_find(m, uid);
    if (e) return e;
    if (m->n == m->cap) {
        m->cap = m->cap ? 2 * m->cap : 16;
        m->e = (gj_ent*)realloc(m->e, (size_t)m->cap * sizeof(gj_ent));
    }
    e = &m->e[m->n++];
    e->uid = uid; e->v = gj_init; e->checked = 0;
    return e;
}

/* _cman_value_init, cell_init.cpp:18-40 */
void eo_value_init(eo_tissue* t) {
    use(t);
    for (int i = 0; i < t->n; i++) {
        t->ca_buf[0][i] = t->ca_buf[1][i] = ca2p_init;
        t->ca2p_avg[i] = t->ca_buf[t->ca_first][i];
        t->ex_inert[i] = g_P->in.ex_inert_init;
        t->ip3_buf[0][i] = t->ip3_buf[1][i] = g_P->in.IP3_init;
        for (int k = 0; k < t->nconn[i]; k++) {
            int64_t u = t->uid[t->conn[(size_t)i * MAX_CONNECT_CELL_NUM + k]];
            if (!gj_find(&t->gj[i], u)) { gj_ent* e = gj_get(&t->gj[i], u); e->v = gj_init; }
        }
    }
    for (int i = t->nmemb + t->nder; i < t->n; i++) {
        if (t->state[i] == EO_DEAD) {
            t->ca2p_avg[i] = 0;
            t->ca_buf[0][i] = t->ca_buf[1][i] = 0.0;
        }
    }
}

/* ---- RNG --------------------------------------------------------------- */
void eo_set_rng(eo_tissue* t, const double* draws, int64_t n) {
    t->rng = draws; t->rng_n = n; t->rng_pos = 0;
}
int64_t eo_rng_cursor(const eo_tissue* t) { return t->rng_pos; }
static double genrand_real(eo_tissue* t) {
    if (t->rng_pos >= t->rng_n) { set_err(t, EO_ERR_RNG_EXHAUSTED); return 0.5; }
    return t->rng[t->rng_pos++];
}

/* std::mt19937_64, published parameters (Matsumoto & Nishimura 2004;
 * C++ [rand.predef]) */
typedef struct { uint64_t mt[312]; int mti; } mt64;
static void mt64_seed(mt64* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < 312; i++)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->mti = 312;
}
static uint64_t mt64_next(mt64* s) {
    static const uint64_t mag01[2] = {0ULL, 0xB5026F5AA96619E9ULL};
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    if (s->mti >= 312) {
        int i;
        for (i = 0; i < 312 - 156; i++) {
            uint64_t x = (s->mt[i] & UM) | (s->mt[i + 1] & LM);
            s->mt[i] = s->mt[i + 156] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        for (; i < 311; i++) {
            uint64_t x = (s->mt[i] & UM) | (s->mt[i + 1] & LM);
            s->mt[i] = s->mt[i + (156 - 312)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        uint64_t x = (s->mt[311] & UM) | (s->mt[0] & LM);
        s->mt[311] = s->mt[155] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        s->mti = 0;
    }
    uint64_t x = s->mt[s->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}
void eo_mt19937_64_raw(uint64_t seed, uint64_t* out, int64_t n) {
    mt64 s; mt64_seed(&s, seed);
    for (int64_t i = 0; i < n; i++) out[i] = mt64_next(&s);
}
/* libstdc++ uniform_real_distribution<double>(0,1) over a 64-bit engine:
 * generate_canonical<double,53> takes one engine word, x/2^64 (clamped below
 * 1), then (b-a)*u + a. */
void eo_mt19937_64_uniform(uint64_t seed, double* out, int64_t n) {
    mt64 s; mt64_seed(&s, seed);
    const double two64 = 18446744073709551616.0;
    for (int64_t i = 0; i < n; i++) {
        double sum = 0.0;
        sum += (double)mt64_next(&s) * 1.0;
        double u = sum / two64;
        if (u >= 1.0) u = nextafter(1.0, 0.0);
        out[i] = u * (1.0 - 0.0) + 0.0;
    }
}

/* ---- connect_cell: cell_connection.cpp:38-203 -------------------------- */
static inline int32_t* conn_row(eo_tissue* t, int c) { return &t->conn[(size_t)c * MAX_CONNECT_CELL_NUM]; }

static void conn_push(eo_tissue* t, int c, int o) {
#ifdef EO_OMP
    /* Lockfree_push_stack::push_back_with_index (atomics.h:95-149) */
    const int k = __atomic_fetch_add(&t->nconn[c], 1, __ATOMIC_RELAXED);
    if (k >= MAX_CONNECT_CELL_NUM) { set_err(t, EO_ERR_CAPACITY); __atomic_fetch_sub(&t->nconn[c], 1, __ATOMIC_RELAXED); return; }
    conn_row(t, c)[k] = o;
    if (k + 1 >= MAX_CONNECT_CELL_NUM) set_err(t, EO_ERR_CAPACITY);
#else
    if (t->nconn[c] >= MAX_CONNECT_CELL_NUM) { set_err(t, EO_ERR_CAPACITY); return; }
    conn_row(t, c)[t->nconn[c]++] = o;
    /* the reference asserts size < N2 after a push (cell_connection.cpp:122);
     * this restatement treats a full list as a capacity error for both ends */
    if (t->nconn[c] >= MAX_CONNECT_CELL_NUM) set_err(t, EO_ERR_CAPACITY);
#endif
}

/* grid_init, cell_connection.cpp:38-62 */
static void grid_init(eo_tissue* t) {
    const double LX = t->g.lx, LY = t->g.ly;
    memset(t->aindx, 0, sizeof(int32_t) * (size_t)t->ANX * t->ANY * t->ANZ);
    EO_PAR
    for (int c = 0; c < t->n; c++) {
        long aix = (long)((0.5 * LX - p_diff_x(t, 0.5 * LX, t->x1[c])) / AREA_GRID);
        long aiy = (long)((0.5 * LY - p_diff_y(t, 0.5 * LY, t->y1[c])) / AREA_GRID);
        long aiz = (long)((min0(t->z1[c])) / AREA_GRID);
        if (aix >= t->ANX || aiy >= t->ANY || aiz >= t->ANZ || aix < 0 || aiy < 0 || aiz < 0) {
            set_err(t, EO_ERR_OUT_OF_GRID);
            t->nconn[c] = t->state[c] == EO_MEMB ? MEMB_ADJ_CONN_NUM : 0;
            continue;
        }
        size_t b = ((size_t)aix * t->ANY + aiy) * t->ANZ + aiz;
#ifdef EO_OMP
        const int slot = __atomic_fetch_add(&t->aindx[b], 1, __ATOMIC_RELAXED);
        if (slot >= N3 - 1) { set_err(t, EO_ERR_CAPACITY); __atomic_fetch_sub(&t->aindx[b], 1, __ATOMIC_RELAXED); }
        else t->area[b * N3 + slot] = c;
#else
        t->area[b * N3 + t->aindx[b]++] = c;
        if (t->aindx[b] >= N3) { set_err(t, EO_ERR_CAPACITY); t->aindx[b] = N3 - 1; }
#endif
        t->nconn[c] = t->state[c] == EO_MEMB ? MEMB_ADJ_CONN_NUM : 0;
    }
}

/* connect_proc, cell_connection.cpp:70-129 */
static void connect_proc(eo_tissue* t) {
    const int ii = 2;
    const int ANX = t->ANX, ANY = t->ANY, ANZ = t->ANZ;
    EO_PAR_DYN
    for (int c = t->nmemb; c < t->n; c++) {
        const long anx = (long)(t->x1[c] / AREA_GRID);
        const long any = (long)(t->y1[c] / AREA_GRID);
        const long anz = (long)(t->z1[c] / AREA_GRID);
        if (anx >= ANX || any >= ANY || anz >= ANZ || anx < 0 || any < 0 || anz < 0) {
            set_err(t, EO_ERR_OUT_OF_GRID);
            continue;
        }
        const long xend = anx + ii + ANX, yend = any + ii + ANY, zend = anz + ii + ANZ;
        for (long j = anx - ii + ANX; j <= xend; j++) {
            const long aix = j % ANX;
            for (long k = any - ii + ANY; k <= yend; k++) {
                const long aiy = k % ANY;
                for (long l = anz - ii + ANZ; l <= zend; l++) {
                    const long aiz = l % ANZ;
                    size_t b = ((size_t)aix * ANY + aiy) * ANZ + aiz;
                    const int sz = t->aindx[b];
                    for (int m = 0; m < sz; ++m) {
                        const int o = t->area[b * N3 + m];
                        if (c <= o) continue;
                        const double diffx = p_diff_x(t, t->x1[c], t->x1[o]);
                        const double diffy = p_diff_y(t, t->y1[c], t->y1[o]);
                        const double diffz = t->z1[c] - t->z1[o];
                        const double rad_sum = t->radius[c] + t->radius[o];
                        if (DIST_SQ(diffx, diffy, diffz) <= LJ_THRESH * LJ_THRESH * rad_sum * rad_sum) {
                            conn_push(t, c, o);
                            conn_push(t, o, c);
                        }
                    }
                }
            }
        }
    }
}

/* gj_refresh, cell_connection.cpp:133-152 */
static void gj_refresh(eo_tissue* t) {
    EO_PAR_DYN
    for (int c = 0; c < t->n; c++) {
        gj_map* m = &t->gj[c];
        for (int i = 0; i < m->n; i++) m->e[i].checked = 0;
        int32_t* row = conn_row(t, c);
        for (int k = 0; k < t->nconn[c]; k++) gj_get(m, t->uid[row[k]])->checked = 1;
        for (int i = 0; i < m->n; i++) if (!m->e[i].checked) m->e[i].v = gj_init;
    }
}

/* find_dermis / set_dermis, cell_connection.cpp:158-185 */
static void set_dermis(eo_tissue* t) {
    EO_PAR
    for (int c = t->nmemb + t->nder; c < t->n; c++) {
        if (t->state[c] == EO_FIX || t->state[c] == EO_MUSUME) {
            double d1Sq = t->g.lx * t->g.lx;
            double distanceSq = 0;
            int dermis = -1;
            int32_t* row = conn_row(t, c);
            for (int k = 0; k < t->nconn[c]; k++) {
                int o = row[k];
                if (t->state[o] == EO_MEMB) {
                    distanceSq = cell_dist_sq(t, c, o);
                    if (distanceSq < d1Sq) { d1Sq = distanceSq; dermis = o; }
                }
            }
            t->dermis[c] = dermis;
        }
    }
}

void eo_connect_cell(eo_tissue* t) {
    use(t);
    grid_init(t);
    connect_proc(t);
    gj_refresh(t);
    set_dermis(t);
}

/* ---- cell_interaction: cell_interaction.cpp:48-685 --------------------- */
static inline int is_near(const eo_tissue* t, int c1, int c2) {
    return cell_dist_sq(t, c1, c2) < (t->radius[c1] + t->radius[c2]) * (t->radius[c1] + t->radius[c2]);
}
static inline double ljmain(const eo_tissue* t, int c1, int c2) {
    const double distSq = cell_dist_sq(t, c1, c2);
    const double rad_sum = t->radius[c1] + t->radius[c2];
    const double LJ6 = POW6(rad_sum) / POW3(distSq);
    return 4.0 * eps_m * LJ6 * (LJ6 - 1) / distSq;
}
static inline double adhesion(const eo_tissue* t, int c1, int c2, const double spring_const) {
    const double distlj = sqrt(cell_dist_sq(t, c1, c2));
    const double rad_sum = t->radius[c1] + t->radius[c2];
    const double LJ2_m1 = (distlj / rad_sum) - 1;
    return (LJ2_m1 + 1 > LJ_THRESH ? 0.0
            : -(spring_const / distlj) * LJ2_m1 *
              (1 - LJ2_m1 * LJ2_m1 / ((LJ_THRESH - 1.0) * (LJ_THRESH - 1.0))));
}

static double coef_law(eo_tissue* t, int law, int c1, int c2) {
    switch (law) {
    case LAW_OTHER: /* :277-279 */
        return ljmain(t, c1, c2);
    case LAW_DER_DER: { /* :129-139 with :77-101 */
        const double dist = sqrt(cell_dist_sq(t, c1, c2));
        if (dist + delta_R < t->radius[c1] + t->radius[c2]) {
            const double dist2 = dist + delta_R;
            const double rad_sum = t->radius[c1] + t->radius[c2];
            const double LJ6 = POW6(rad_sum) / POW6(dist2);
            return 4.0 * eps_m * LJ6 * (LJ6 - 1) / (dist * dist2) + para_ljp2;
        } else if (is_near(t, c1, c2)) {
            return para_ljp2;
        } else {
            const double dist2 = dist + delta_R;
            const double rad_sum = t->radius[c1] + t->radius[c2];
            const double LJ6 = POW6(rad_sum) / POW6(dist2);
            return 4.0 * eps_m * LJ6 * (LJ6 - 1) / (dist * dist) + para_ljp2;
        }
    }
    case LAW_ALAIRDE: /* :141-152 */
        if (is_near(t, c1, c2)) return ljmain(t, c1, c2);
        else {
            double spf = K_DESMOSOME;
            if (t->agek[c1] > THRESH_SP && t->agek[c2] > THRESH_SP) spf = K_TOTAL;
            return adhesion(t, c1, c2, spf);
        }
    case LAW_FIXMU: /* :154-162 */
        if (is_near(t, c1, c2)) return ljmain(t, c1, c2);
        return adhesion(t, c1, c2, K_DESMOSOME);
    case LAW_MU_MEMB: /* :164-172 */
        if (is_near(t, c1, c2)) return ljmain(t, c1, c2);
        return adhesion(t, c1, c2, t->spring_force_to_memb[c1]);
    case LAW_FIX_MEMB: /* :174-183 */
        if (is_near(t, c1, c2)) return ljmain(t, c1, c2);
        else {
            const double distlj = sqrt(cell_dist_sq(t, c1, c2));
            const double LJ2 = distlj / (t->radius[c1] + t->radius[c2]);
            return -(Kspring / distlj) * (LJ2 - 1.0);
        }
    case LAW_MEMB_MEMB: { /* :185-234 */
        const double rad_sum = t->radius[c1] + t->radius[c2];
        const double rad_sum_sq = rad_sum * rad_sum;
        double cr_dist_sq;
        const double distSq = cell_dist_sq(t, c1, c2);
        if (distSq < (cr_dist_sq = rad_sum_sq * P_MEMB * P_MEMB)) {
            const double LJ6 = POW3(cr_dist_sq) / POW3(distSq);
            double tmp = 4.0 * eps_m * LJ6 * (LJ6 - 1) / distSq;
            if (fabs(tmp) > 100) set_err(t, EO_ERR_STRONG_INTERACTION);
            return tmp;
        } else if (distSq < rad_sum_sq) {
            const double distlj = sqrt(distSq);
            const double cr_dist = rad_sum * P_MEMB;
            double tmp = -(DER_DER_CONST / distlj) * (distlj / cr_dist - 1.0);
            if (fabs(tmp) > 100) set_err(t, EO_ERR_STRONG_INTERACTION);
            return tmp;
        } else {
            const double distlj = sqrt(distSq);
            const double lambda_dist = (1.0 + P_MEMB) * rad_sum;
            const double cr_dist = rad_sum * P_MEMB;
            const double LJ6 = POW6(cr_dist) / POW6(lambda_dist - distlj);
            double tmp = -(DER_DER_CONST / rad_sum) * ((1.0 - P_MEMB) / P_MEMB)
                         - 4 * eps_m * (LJ6 * (LJ6 - 1.0)) / ((lambda_dist - distlj) * distlj);
            if (fabs(tmp) > 100) set_err(t, EO_ERR_STRONG_INTERACTION);
            return tmp;
        }
    }
    case LAW_PAIR: { /* :281-285 */
        const double dist = sqrt(cell_dist_sq(t, c1, c2));
        const double force = Kspring_division * (dist - t->spr_nat_len[c1]);
        return -force / dist;
    }
    default:
        return 0.0;
    }
}

double eo_coef(eo_tissue* t, int law, int c1, int c2) {
    use(t);
    return coef_law(t, law, c1, c2);
}

/* cell_interaction_apply, cell_interaction.cpp:48-66 */
static int g_tape_on;
static void tape_push(int law, int c1, int c2);
static void apply(eo_tissue* t, int law, int c1, int c2) {
    if (g_tape_on) tape_push(law, c1, c2);
    const double coef = coef_law(t, law, c1, c2);
    if (fabs(coef) > 1000) set_err(t, EO_ERR_STRONG_INTERACTION);
    const double dumx = DT_Cell * coef * p_diff_x(t, t->x1[c1], t->x1[c2]);
    const double dumy = DT_Cell * coef * p_diff_y(t, t->y1[c1], t->y1[c2]);
    const double dumz = DT_Cell * coef * (t->z1[c1] - t->z1[c2]);
    add_d(&t->x2[c1], dumx); add_d(&t->y2[c1], dumy); add_d(&t->z2[c1], dumz);
    add_d(&t->x2[c2], -dumx); add_d(&t->y2[c2], -dumy); add_d(&t->z2[c2], -dumz);
}

/* wall_interaction, cell_interaction.cpp:287-294 */
static void wall_interaction(eo_tissue* t, int c) {
    const double distlj = 2.0 * t->z1[c];
    const double LJ6 = POW6(t->radius[c]) / POW6(t->z1[c]);
    const double ljm = 4.0 * eps_m * LJ6 * (LJ6 - 1.0) / (distlj * distlj);
    add_d(&t->z2[c], DT_Cell * ljm * 2.0 * t->z1[c]);
}

/* vec3 (vec3.h:5-55) */
typedef struct { double x, y, z; } v3;
static inline v3 v3_make(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 v3_cross(v3 a, v3 b) {
    return v3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double v3_norm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline v3 v3_normalize(v3* a) {      /* mutates and returns *this */
    double nm = v3_norm(*a);
    a->x /= nm; a->y /= nm; a->z /= nm;
    return *a;
}
static inline v3 v3_scale(double c, v3 v) { return v3_make(c * v.x, c * v.y, c * v.z); }
static inline v3 v3_div(v3 v, double c) { return v3_make(v.x / c, v.y / c, v.z / c); }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vpsub(const eo_tissue* t, v3 a, v3 b) {
    return v3_make(p_diff_x(t, a.x, b.x), p_diff_y(t, a.y, b.y), a.z - b.z);
}
static inline v3 ctov(const eo_tissue* t, int c) { return v3_make(t->x1[c], t->y1[c], t->z1[c]); }

/* _memb_bend_diff_factor (NEW_BEND_POT), cell_interaction.cpp:349-360 */
static inline double bend_factor(double cdot) {
    if (cdot >= 0) return (1.0 - cdot);
    return 1.0 / ((1.0 + cdot) * (1.0 + cdot));
}
/* tri_rho, :463-465 */
static inline v3 tri_rho(v3 n1, v3 b1) { return v3_sub(b1, v3_scale(v3_dot(n1, b1), n1)); }

/* tri_bend_1, :468-474 -- norm1_pre is normalised in place before its norm
 * is taken as the divisor */
static v3 tri_bend_1(const eo_tissue* t, int r1, int a1, int b1, int c1) {
    v3 r1v = ctov(t, r1), a1v = ctov(t, a1), b1v = ctov(t, b1), c1v = ctov(t, c1);
    v3 norm1_pre = v3_cross(vpsub(t, r1v, b1v), vpsub(t, a1v, b1v));
    v3 norm1 = v3_normalize(&norm1_pre);
    v3 n2tmp = v3_cross(vpsub(t, a1v, b1v), vpsub(t, c1v, b1v));
    v3 norm2 = v3_normalize(&n2tmp);
    double dot1 = v3_dot(norm1, norm2);
    return v3_scale(bend_factor(dot1),
                    v3_div(v3_cross(vpsub(t, a1v, b1v), tri_rho(norm1, norm2)), v3_norm(norm1_pre)));
}
/* tri_bend_2, :476-486 */
static v3 tri_bend_2(const eo_tissue* t, int r1, int a1, int b1, int c1) {
    v3 r1v = ctov(t, r1), a1v = ctov(t, a1), b1v = ctov(t, b1), c1v = ctov(t, c1);
    v3 norm1_pre = v3_cross(vpsub(t, r1v, c1v), vpsub(t, b1v, c1v));
    v3 norm2_pre = v3_cross(vpsub(t, r1v, b1v), vpsub(t, a1v, b1v));
    v3 norm1 = v3_normalize(&norm1_pre), norm2 = v3_normalize(&norm2_pre);
    double dot1 = v3_dot(norm1, norm2);
    return v3_scale(bend_factor(dot1),
                    v3_add(v3_div(v3_cross(vpsub(t, b1v, c1v), tri_rho(norm1, norm2)), v3_norm(norm1_pre)),
                           v3_div(v3_cross(vpsub(t, a1v, b1v), tri_rho(norm2, norm1)), v3_norm(norm2_pre))));
}
/* tri_memb_bend, :488-499 (returns dr) */
static v3 tri_memb_dr(const eo_tissue* t, int m) {
    const int32_t* md = &t->memb_adj[6 * m];
    v3 dr = v3_make(0, 0, 0);
    for (int i = 0; i < 6; i++) {
        int nb = md[(i + 1) % 6];
        dr = v3_add(dr, tri_bend_1(t, m, md[i], nb, t->memb_adj[6 * nb + i]));
    }
    for (int i = 0; i < 6; i++)
        dr = v3_add(dr, tri_bend_2(t, m, md[i], md[(i + 1) % 6], md[(i + 2) % 6]));
    return dr;
}
void eo_tri_bend(const eo_tissue* t, int m, double* out3) {
    use(t);
    v3 dr = tri_memb_dr(t, m);
    out3[0] = dr.x; out3[1] = dr.y; out3[2] = dr.z;
}

static inline int collide_with_wall(const eo_tissue* t, int c) { return t->z1[c] < t->radius[c]; }

void eo_cell_interaction(eo_tissue* t) {
    use(t);
    /* memb_bend, :501-519 */
    EO_PAR
    for (int m = 0; m < t->nmemb; m++) {
        v3 dr = tri_memb_dr(t, m);
        t->x2[m] += DT_Cell * KBEND * dr.x;
        t->y2[m] += DT_Cell * KBEND * dr.y;
        t->z2[m] += DT_Cell * KBEND * dr.z;
    }
    /* _MEMB_interaction, :529-547 */
    EO_PAR
    for (int m = 0; m < t->nmemb; m++) {
        if (collide_with_wall(t, m)) wall_interaction(t, m);
        int32_t* row = conn_row(t, m);
        for (int i = 0; i < MEMB_ADJ_CONN_NUM; i++) {
            int o = row[i];
            if (t->state[o] == EO_MEMB && m > o) apply(t, LAW_MEMB_MEMB, m, o);
        }
    }
    /* non-membrane loop, :656-679 */
    EO_PAR_DYN
    for (int c = t->nmemb; c < t->n; c++) {
        int32_t* row = conn_row(t, c);
        const int nc = t->nconn[c];
        switch (t->state[c]) {
        case EO_DER: /* _DER_interaction :549-561 */
            if (collide_with_wall(t, c)) wall_interaction(t, c);
            for (int k = 0; k < nc; k++) {
                int o = row[k];
                if (t->state[o] == EO_DER && c > o) apply(t, LAW_DER_DER, c, o);
                else apply(t, LAW_OTHER, c, o);
            }
            break;
        case EO_FIX: /* _FIX_interaction :581-604 */
            for (int k = 0; k < nc; k++) {
                int o = row[k];
                switch (t->state[o]) {
                case EO_FIX:
                    if (t->pair[c] != o && c > o) apply(t, LAW_FIXMU, c, o);
                    break;
                case EO_MUSUME:
                    if (t->pair[c] != o) apply(t, LAW_FIXMU, c, o);
                    break;
                case EO_MEMB:
                    /* dermis()->memb_exist<0>(conn) == (conn == dermis) */
                    if (t->dermis[c] == o) apply(t, LAW_FIX_MEMB, c, o);
                    else apply(t, LAW_OTHER, c, o);
                    break;
                default: break;
                }
            }
            break;
        case EO_MUSUME: { /* _MUSUME_interaction :607-634 */
            t->spring_force_to_memb[c] = 0;
            int pr = t->pair[c];
            if (pr >= 0 && t->state[pr] == EO_FIX) t->spring_force_to_memb[c] = Kspring;
            else if (t->rest_div_times[c] > 0) t->spring_force_to_memb[c] = Kspring_d;
            for (int k = 0; k < nc; k++) {
                int o = row[k];
                if (t->state[o] == EO_MEMB) {
                    if (t->dermis[c] == o) apply(t, LAW_MU_MEMB, c, o);
                    else apply(t, LAW_OTHER, c, o);
                } else if (t->state[o] == EO_MUSUME && t->pair[c] != o && c > o) {
                    apply(t, LAW_FIXMU, c, o);
                }
            }
            break;
        }
        case EO_ALIVE: case EO_AIR: case EO_DEAD: /* _AL_AIR_DE_interaction :563-579 */
            for (int k = 0; k < nc; k++) {
                int o = row[k];
                switch (t->state[o]) {
                case EO_ALIVE: case EO_AIR: case EO_DEAD:
                    if (c > o) apply(t, LAW_ALAIRDE, c, o);
                    break;
                case EO_DER: break;
                default: apply(t, LAW_ALAIRDE, c, o); break;
                }
            }
            break;
        default: break;
        }
        /* _pair_interaction :637-641 */
        if (t->pair[c] >= 0 && c > t->pair[c]) apply(t, LAW_PAIR, c, t->pair[c]);
    }
}

/* pos_copy, cellmanager.cpp:19-26 */
void eo_pos_copy(eo_tissue* t) {
    EO_PAR
    for (int c = 0; c < t->n; c++) { t->x1[c] = t->x2[c]; t->y1[c] = t->y2[c]; t->z1[c] = t->z2[c]; }
}

/* ---- cell_state_renew: cell_state_renew.cpp:17-443 --------------------- */
#define stoch_div_time_ratio (g_P->in.stoch_div_time_ratio)
#define S2 (g_P->in.S2)
#define eps_kb (g_P->in.eps_kb)
#define accel_div (g_P->in.accel_div)
#define alpha_b (g_P->in.alpha_b)
#define S0 (g_P->in.S0)
#define eps_ks (g_P->in.eps_ks)
#define accel_diff (g_P->in.accel_diff)
#define alpha_k (g_P->in.alpha_k)
#define eps_kk (g_P->in.eps_kk)
#define S1 (g_P->in.S1)
#define ubar (g_P->in.ubar)
#define delta_lipid (g_P->in.delta_lipid)
#define delta_sig_r1 (g_P->in.delta_sig_r1)
#define lipid_rel (g_P->in.lipid_rel)
#define lipid_c (g_P->in.lipid)

static inline double weighted_eps_kb(const eo_tissue* t, int c) { return t->is_malignant[c] ? accel_div * eps_kb : eps_kb; }
static inline double ageb_const(const eo_tissue* t, int c) {
    return weighted_eps_kb(t, c) * (S2 + alpha_b * min0(t->ca2p_avg[c] - ca2p_init));
}
static inline double weighted_eps_ks(const eo_tissue* t, int c) { return t->is_malignant[c] ? accel_diff * eps_ks : eps_ks; }
static inline double agek_ALIVE_const(const eo_tissue* t, int c) {
    return weighted_eps_ks(t, c) * (S0 + alpha_k * min0(t->ca2p_avg[c] - ca2p_init));
}
static inline double agek_DEAD_AIR_const(void) { return eps_kk * S1; }
static inline double k_lipid_release(const eo_tissue* t, int c) {
    return 0.25 * lipid_rel * (1 + tanh((t->ca2p_avg[c] - ubar) / delta_sig_r1)) *
           (1 + tanh((t->agek[c] - THRESH_SP) / delta_lipid));
}
static inline double k_lipid(const eo_tissue* t, int c) {
    return 0.25 * lipid_c * (1 + tanh((ubar - t->ca2p_avg[c]) / delta_sig_r1)) *
           (1 + tanh((t->agek[c] - THRESH_SP) / delta_lipid));
}

/* stochastic_div_test / is_divide_ready, :133-155 */
static int is_divide_ready(eo_tissue* t, int c) {
    if (t->pair[c] < 0 && (t->ageb[c] >= t->div_age_thresh[c] * (1.0 - stoch_div_time_ratio))) {
        return genrand_real(t) * (t->div_age_thresh[c] * stoch_div_time_ratio) <=
               stoch_corr_coef * DT_Cell * weighted_eps_kb(t, c) * S2;
    }
    return 0;
}

/* is_divide_ready, then cell_divide when it holds; 1 if the cell divided.
 * The draw and the creation touch the shared replay cursor and cell count,
 * so the multithreaded build takes them under one lock, entered only by
 * cells old enough to draw */
static void cell_divide(eo_tissue* t, int div);
static int try_divide(eo_tissue* t, int c) {
#ifdef EO_OMP
    if (!(t->pair[c] < 0 && (t->ageb[c] >= t->div_age_thresh[c] * (1.0 - stoch_div_time_ratio)))) return 0;
#endif
    int divided = 0;
    EO_CRITICAL
    {
        if (is_divide_ready(t, c)) { cell_divide(t, c); divided = 1; }
    }
    return divided;
}

/* calc_cell_uvec + div_direction, :168-217 */
static void div_direction(eo_tissue* t, int me, int dermis, double* outx, double* outy, double* outz) {
    double nvx = p_diff_x(t, t->x1[me], t->x1[dermis]);
    double nvy = p_diff_y(t, t->y1[me], t->y1[dermis]);
    double nvz = t->z1[me] - t->z1[dermis];
    double norm = sqrt(DIST_SQ(nvx, nvy, nvz));
    double nx = nvx / norm, ny = nvy / norm, nz = nvz / norm;
    double ox, oy, oz, sum;
    do {
        double rand_theta = M_PI * genrand_real(t);
        double cr1 = cos(rand_theta);
        double sr1 = sin(rand_theta);
        double rand_phi = 2 * M_PI * genrand_real(t);
        double cr2 = cos(rand_phi);
        double sr2 = sin(rand_phi);
        double rvx = sr1 * cr2;
        double rvy = sr1 * sr2;
        double rvz = cr1;
        ox = ny * rvz - nz * rvy;
        oy = nz * rvx - nx * rvz;
        oz = nx * rvy - ny * rvx;
        if (t->err == EO_ERR_RNG_EXHAUSTED) { ox = 1; oy = 0; oz = 0; }
    } while ((sum = DIST_SQ(ox, oy, oz)) < 1.0e-14);
    sum = sqrt(sum);
    *outx = ox / sum; *outy = oy / sum; *outz = oz / sum;
}

/* cell_divide, :227-270 */
static void cell_divide(eo_tissue* t, int div) {
    const double delta_L = g_P->delta_L;   /* :231, delta_L_coef * R_max (CalcParams.cpp:41) */
    int d = eo_create_cell(t, EO_MUSUME, t->fix_origin[div], t->x1[div], t->y1[div], t->z1[div],
                           t->radius[div], t->ca_buf[t->ca_first][div], t->ca_buf[t->ca_first][div],
                           t->ip3_buf[t->ca_first][div], t->ex_inert[div], 0, 0, 0, 0, delta_L,
                           t->rest_div_times[div], t->is_malignant[div]);
    if (d < 0) return;
    t->pair[div] = d;
    t->pair[d] = div;
    t->spr_nat_len[div] = delta_L;
    t->ageb[div] = 0;
    if (t->state[div] == EO_MUSUME) { t->rest_div_times[div]--; t->rest_div_times[d]--; }
    if (t->dermis[div] < 0) { set_err(t, EO_ERR_NO_DERMIS); return; }
    double divx, divy, divz;
    div_direction(t, div, t->dermis[div], &divx, &divy, &divz);
    double nx = t->x1[div] + divx * 0.5 * delta_L;
    double ny = t->y1[div] + divy * 0.5 * delta_L;
    double nz = t->z1[div] + divz * 0.5 * delta_L;
    t->x1[div] = t->x2[div] = nx; t->y1[div] = t->y2[div] = ny; t->z1[div] = t->z2[div] = nz;
    double dx = t->x1[d] - divx * 0.5 * delta_L;
    double dy = t->y1[d] - divy * 0.5 * delta_L;
    double dz = t->z1[d] - divz * 0.5 * delta_L;
    t->x1[d] = t->x2[d] = dx; t->y1[d] = t->y2[d] = dy; t->z1[d] = t->z2[d] = dz;
}

/* Cell::migrate + slot move, cellmanager.cpp:347-349, cell.cpp:42-48 */
static void move_cell(eo_tissue* t, int dst, int src) {
#define MV(a) t->a[dst] = t->a[src]
    MV(state); MV(fix_origin); MV(x1); MV(y1); MV(z1); MV(x2); MV(y2); MV(z2);
    MV(radius); MV(div_age_thresh); MV(ca2p_avg); MV(ex_inert); MV(agek); MV(ageb);
    MV(ex_fat); MV(in_fat); MV(spr_nat_len); MV(diff_u); MV(spring_force_to_memb);
    MV(rest_div_times); MV(pair); MV(dermis); MV(is_malignant); MV(uid); MV(nconn);
#undef MV
    for (int k = 0; k < 3; k++) t->lat[3 * dst + k] = t->lat[3 * src + k];
    for (int k = 0; k < 3; k++) t->cst[3 * dst + k] = t->cst[3 * src + k];
    memcpy(conn_row(t, dst), conn_row(t, src), sizeof(int32_t) * (size_t)t->nconn[src]);
    for (int b = 0; b < 2; b++) { t->ca_buf[b][dst] = t->ca_buf[b][src]; t->ip3_buf[b][dst] = t->ip3_buf[b][src]; }
    gj_map tmp = t->gj[dst]; t->gj[dst] = t->gj[src]; t->gj[src] = tmp; t->gj[src].n = 0;
    if (t->pair[dst] >= 0) t->pair[t->pair[dst]] = dst;
}

/* CellManager::remove_exec, cellmanager.cpp:338-352.  A queued slot at or
 * beyond the current end means the tail that moved was itself queued: the
 * reference then deletes a live pointer (undefined); reported as
 * EO_ERR_REMOVE_ALIAS. */
static void remove_exec(eo_tissue* t) {
    for (int i = 0; i < t->nrq; i++) {
        int r = t->rq[i];
        if (r >= t->n) { set_err(t, EO_ERR_REMOVE_ALIAS); continue; }
        if (t->pair[r] >= 0) set_err(t, EO_ERR_PAIR_REMOVED);
        if (t->state[r] == EO_DEAD) {   /* :343-346: the record goes to cst_store */
            if (t->nstat >= t->cap) set_err(t, EO_ERR_CAPACITY);
            else {
                int64_t* o = &t->stat[3 * (size_t)t->nstat++];
                o[0] = t->cst[3 * r]; o[1] = t->cst[3 * r + 1]; o[2] = t->timestep;
            }
        }
        int tail = --t->n;
        if (tail != r) move_cell(t, r, tail);
        else t->gj[r].n = 0;
    }
    t->nrq = 0;
}

/* pair_disperse, :364-385 */
static void pair_disperse(eo_tissue* t, int c) {
    const double eps_L = g_P->in.eps_L, unpair_dist_coef = g_P->in.unpair_dist_coef;
    int p = t->pair[c];
    double rad_sum = t->radius[c] + t->radius[p];
    double unpair_th = unpair_dist_coef * rad_sum;
    double distSq = 0;
    if (t->spr_nat_len[c] < 2.0 * t->radius[c]) {
        t->spr_nat_len[c] += DT_Cell * eps_L;
        t->spr_nat_len[p] = t->spr_nat_len[c];
    } else if ((distSq = cell_dist_sq(t, c, p)) > unpair_th * unpair_th) {
        t->spr_nat_len[c] = 0;
        t->spr_nat_len[p] = 0;
        t->pair[p] = -1;
        t->pair[c] = -1;
    }
}

void eo_cell_state_renew(eo_tissue* t) {
    use(t);
    const int start = t->nmemb + t->nder, end = t->n;  /* bounds captured at call */
    EO_PAR_DYN
    for (int c = start; c < end; c++) {
        switch (t->state[c]) {
        case EO_ALIVE: /* :347-358 */
            if (t->agek[c] >= THRESH_DEAD) {
                t->state[c] = EO_DEAD;   /* cornificate, cellmanager.cpp:48-53 */
                t->cst[3 * c + 1] = t->timestep;
#ifdef EO_OMP
                __atomic_fetch_add(&t->sw, 1, __ATOMIC_RELAXED);
#else
                t->sw = (uint8_t)(t->sw + 1);
#endif
            } else {
                const double tmp = k_lipid_release(t, c) * t->in_fat[c];
                t->in_fat[c] += DT_Cell * (k_lipid(t, c) * (1.0 - t->in_fat[c]) - tmp);
                t->ex_fat[c] += DT_Cell * tmp;
                t->agek[c] += DT_Cell * agek_ALIVE_const(t, c);
            }
            break;
        case EO_DEAD: case EO_AIR: /* :330-341 */
            if (t->agek[c] >= g_P->in.ADHE_CONST && (double)t->nconn[c] <= g_P->in.DISA_conn_num_thresh) {
#ifdef EO_OMP
                t->rq[__atomic_fetch_add(&t->nrq, 1, __ATOMIC_RELAXED)] = c;
#else
                t->rq[t->nrq++] = c;
#endif
            } else t->agek[c] += DT_Cell * agek_DEAD_AIR_const();
            break;
        case EO_FIX: /* :308-324 */
            if (t->dermis[c] < 0) { set_err(t, EO_ERR_NO_DERMIS); break; }
            if (try_divide(t, c) == 0) t->ageb[c] += DT_Cell * ageb_const(t, c);
            break;
        case EO_MUSUME: /* :282-302 (SYSTEM == WHOLE) */
            if (t->dermis[c] < 0 && t->pair[c] < 0) {
                t->state[c] = EO_ALIVE;
                t->cst[3 * c] = t->timestep;   /* :287 */
                break;
            }
            if (!(t->rest_div_times[c] > 0 && try_divide(t, c))) t->ageb[c] += DT_Cell * ageb_const(t, c);
            break;
        default: break;
        }
    }
    remove_exec(t);
    EO_PAR
    for (int c = t->nmemb + t->nder; c < t->n; c++)
        if (t->pair[c] >= 0 && c > t->pair[c]) pair_disperse(t, c);
}

/* initialize_sc, :435-443 */
void eo_initialize_sc(eo_tissue* t, double zzmax) {
    use(t);
    for (int c = t->nmemb + t->nder; c < t->n; c++) {
        if (t->state[c] == EO_ALIVE && zzmax - t->z1[c] < 8 * R_max) {
            t->agek[c] = t->agek[c] > THRESH_DEAD ? t->agek[c] : THRESH_DEAD;
            t->state[c] = EO_AIR;
        }
    }
}

/* cell_pos_periodic_fix, cellmanager.cpp:395-414 */
void eo_cell_pos_periodic_fix(eo_tissue* t) {
    use(t);
    const double LX = t->g.lx, LY = t->g.ly;
    EO_PAR
    for (int c = 0; c < t->n; c++) {
        if (t->x2[c] > LX) { double v = t->x2[c] - LX; t->x1[c] = t->x2[c] = v; }
        else if (t->x2[c] < 0) { double v = t->x2[c] + LX; t->x1[c] = t->x2[c] = v; }
        if (t->y2[c] > LY) { double v = t->y2[c] - LY; t->y1[c] = t->y2[c] = v; }
        else if (t->y2[c] < 0) { double v = t->y2[c] + LY; t->y1[c] = t->y2[c] = v; }
    }
}

/* calc_zzmax, proc.cpp:34-43 */
double eo_calc_zzmax(eo_tissue* t) {
    use(t);
    double zmax = 0;
    for (int c = t->nmemb + t->nder; c < t->n; c++) {
        int st = t->state[c];
        if (st == EO_DEAD || st == EO_ALIVE || st == EO_MUSUME || st == EO_FIX)
            if (zmax < t->z1[c]) zmax = t->z1[c];
    }
    return zmax;
}

/* ---- setup_map_lat: map.cpp:18-167 ------------------------------------- */
/* cmap1 voxel v = cell c; threads race on it as in the reference, so the
 * multithreaded build stores atomically (the last writer wins either way) */
#ifdef EO_OMP
#define STAMP1(t, v, c) __atomic_store_n(&(t)->cmap1[v], (c), __ATOMIC_RELAXED)
#else
#define STAMP1(t, v, c) ((t)->cmap1[v] = (c))
#endif
static int lat_wrap(eo_tissue* t, long v, int N) {
    /* precalc_lat_* tables hold i%N for i in [0, 3N) (utils.cpp:18-30) */
    long i = v + N;
    if (i < 0 || i >= 3L * N) { set_err(t, EO_ERR_OUT_OF_GRID); i = ((i % N) + N) % N; }
    return (int)(i % N);
}

/* the stamps of one cell (map.cpp:42-154); dySq, ipy, dzSq are scratch of
 * 4 iry + 1, 4 iry + 1 and 4 irz + 1 entries */
static void stamp_cell(eo_tissue* t, int c, int irx, int iry, int irz, double* dySq, int* ipy, double* dzSq) {
    const int NX = t->NX, NY = t->NY, NZ = t->NZ;
    const int* clat = &t->lat[3 * c];
    const int st = t->state[c];
    const int afm = st == EO_ALIVE || st == EO_FIX || st == EO_MUSUME;
    const int _irx = afm ? 2 * irx : irx, _iry = afm ? 2 * iry : iry, _irz = afm ? 2 * irz : irz;
    const double normal_radSq = t->radius[c] * t->radius[c];
    const double crad = FAC_MAP * t->radius[c];
    const double cradSq = crad * crad;
    const int ymax = clat[1] + _iry, ymin = clat[1] - _iry;
    for (int l = ymin, yc = 0; l <= ymax; l++) {
        const double my = l * t->dy;
        ipy[yc] = lat_wrap(t, l, NY);
        const double diffy = p_diff_y(t, my, t->y1[c]);
        dySq[yc] = diffy * diffy;
        yc++;
    }
    const int z_b = clat[2] - _irz, z_u = clat[2] + _irz;
    const int zmin = z_b >= 1 ? z_b : 1;
    const int zmax = z_u < NZ ? z_u : NZ - 1;
    for (int m = zmin, zc = 0; m <= zmax; m++) {
        const double mz = m * t->dz;
        const double diffz = mz - t->z1[c];
        dzSq[zc] = diffz * diffz;
        zc++;
    }
    const int xmax = clat[0] + _irx, xmin = clat[0] - _irx;
    for (int k = xmin; k <= xmax; k++) {
        const double mx = k * t->dx;
        const int ipx = lat_wrap(t, k, NX);
        const double diffx = p_diff_x(t, mx, t->x1[c]);
        const double diffxSq = diffx * diffx;
        for (int yc = 0; yc < _iry * 2 + 1; yc++) {
            const double tmps = diffxSq + dySq[yc];
            for (int zc = 0, ipz = zmin; ipz <= zmax; ipz++, zc++) {
                const double distSq = tmps + dzSq[zc];
                size_t v = IDX3(t, ipx, ipy[yc], ipz);
                if (!afm) {
                    if (distSq < normal_radSq) STAMP1(t, v, c);
                } else if (distSq < cradSq) {
                    t->cmap2[v] = 1;
                    if (distSq < normal_radSq) STAMP1(t, v, c);
                }
            }
        }
    }
}

void eo_setup_map_lat(eo_tissue* t) {
    use(t);
    const double Rad = 1.4;
    const int NX = t->NX, NY = t->NY, NZ = t->NZ;
    const int irx = (int)(Rad * NX / t->g.lx);
    const int iry = (int)(Rad * NY / t->g.ly);
    const int irz = (int)(Rad * NZ / t->g.lz);
    EO_PAR
    for (size_t i = 0; i < t->gsz; i++) { t->cmap1[i] = -1; t->cmap2[i] = 0; }
    EO_PAR
    for (int c = 0; c < t->n; c++) {  /* set_lattice :18-22 */
        t->lat[3 * c + 0] = lat_wrap(t, (long)(int)(t->x1[c] * t->inv_dx), NX);
        t->lat[3 * c + 1] = lat_wrap(t, (long)(int)(t->y1[c] * t->inv_dy), NY);
        t->lat[3 * c + 2] = lat_wrap(t, (long)(int)(t->z1[c] * t->inv_dz), NZ);
    }
#ifdef EO_OMP
#pragma omp parallel
#endif
    {
        double* dySq = (double*)malloc(sizeof(double) * (4 * iry + 1));
        int* ipy = (int*)malloc(sizeof(int) * (4 * iry + 1));
        double* dzSq = (double*)malloc(sizeof(double) * (4 * irz + 1));
#ifdef EO_OMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int c = 0; c < t->n; c++) stamp_cell(t, c, irx, iry, irz, dySq, ipy, dzSq);
        free(dySq); free(ipy); free(dzSq);
    }
    EO_PAR
    for (int l = 0; l < NZ; l++) {  /* :156-166 */
        for (int j = 0; j <= NX; j++) {
            t->cmap2[IDX3(t, j, NY, l)] = t->cmap2[IDX3(t, j, 0, l)];
            t->cmap1[IDX3(t, j, NY, l)] = t->cmap1[IDX3(t, j, 0, l)];
        }
        for (int k = 0; k <= NY; k++) {
            t->cmap2[IDX3(t, NX, k, l)] = t->cmap2[IDX3(t, 0, k, l)];
            t->cmap1[IDX3(t, NX, k, l)] = t->cmap1[IDX3(t, 0, k, l)];
        }
    }
}

/* ---- calc_ext_stim: ext_stim.cpp:9-69 ---------------------------------- */
static inline double fB(double age, double B, int cornif) {
    const double kb = g_P->in.kb, DUR_ALIVE = g_P->in.DUR_ALIVE, DUR_DEAD = g_P->in.DUR_DEAD;
    return (cornif && age > THRESH_DEAD - DUR_ALIVE && age <= THRESH_DEAD + DUR_DEAD ? 1 : 0) - kb * B;
}

void eo_calc_ext_stim(eo_tissue* t, double zzmax) {
    use(t);
    const double DB = g_P->in.DB;
    const int NX = t->NX, NY = t->NY, NZ = t->NZ;
    const int iz_bound = (int)((zzmax + FAC_MAP * R_max) * t->inv_dz);
    if (iz_bound > NZ || iz_bound < 0) { set_err(t, EO_ERR_OUT_OF_GRID); return; }
    const double* carr = t->bst[t->b_first];
    double* narr = t->bst[1 - t->b_first];
    const double inv_dz = t->inv_dz;
    EO_PAR2
    for (int j = 0; j < NX; j++) {
        for (int k = 0; k < NY; k++) {
            const int prev_x = t->px_prev[j], next_x = t->px_next[j];
            const int prev_y = t->py_prev[k], next_y = t->py_next[k];
            for (int l = 0; l < iz_bound; l++) {
                const int prev_z = t->pz_prev[l], next_z = t->pz_next[l];
                double dum_age = 0;
                int flg_cornified = 0;
                int cc = t->cmap1[IDX3(t, j, k, l)];
                if (cc >= 0 && (t->state[cc] == EO_DEAD || t->state[cc] == EO_ALIVE)) {
                    dum_age = t->agek[cc];
                    flg_cornified = 1;
                }
                const double cext = carr[IDX3(t, j, k, l)];
                narr[IDX3(t, j, k, l)] = cext +
                    DT_Ca * (DB *
                             ((double)t->cmap2[IDX3(t, prev_x, k, l)] * (carr[IDX3(t, prev_x, k, l)] - cext)
                              + (double)t->cmap2[IDX3(t, j, prev_y, l)] * (carr[IDX3(t, j, prev_y, l)] - cext)
                              + (double)t->cmap2[IDX3(t, j, k, prev_z)] * (carr[IDX3(t, j, k, prev_z)] - cext)
                              + (double)t->cmap2[IDX3(t, j, k, next_z)] * (carr[IDX3(t, j, k, next_z)] - cext)
                              + (double)t->cmap2[IDX3(t, j, next_y, l)] * (carr[IDX3(t, j, next_y, l)] - cext)
                              + (double)t->cmap2[IDX3(t, next_x, k, l)] * (carr[IDX3(t, next_x, k, l)] - cext))
                             * inv_dz * inv_dz
                             + fB(dum_age, cext, flg_cornified));
            }
        }
    }
    EO_PAR
    for (int l = 0; l < iz_bound; l++) {
        for (int j = 0; j < NX; j++) narr[IDX3(t, j, NY, l)] = narr[IDX3(t, j, 0, l)];
        for (int k = 0; k <= NY; k++) narr[IDX3(t, NX, k, l)] = narr[IDX3(t, 0, k, l)];
    }
    t->b_first = 1 - t->b_first;
}

/* ---- calc_ca2p: ca2p.cpp:14-402 ---------------------------------------- */
#define Kpp (g_P->in.Kpp)
#define dp_c (g_P->in.dp)

static inline double ca2p_by_ext_stim(double ext_stim) {
    const double kbc = g_P->in.kbc, Hb = g_P->in.Hb, Cout = g_P->in.Cout;
    return kbc * ext_stim * ext_stim * Cout / (Hb + ext_stim * ext_stim);
}
static inline double ca2p_into_storage(double ca2p_in_cell) {
    const double kg = g_P->in.kg, gam = g_P->in.gamma;
    return gam * ca2p_in_cell / (kg + ca2p_in_cell);
}
static inline double ER_domain1_active(double ip3) {
    const double mu0 = g_P->in.mu0, mu1 = g_P->in.mu1, kmu = g_P->in.kmu;
    return (mu0 + mu1 * ip3 / (ip3 + kmu));
}
static inline double ER_domain2_active(double ca2p) {
    const double para_b = g_P->in.para_b, para_bb = g_P->in.para_bb, para_k1 = g_P->in.para_k1;
    return (para_b + para_bb * ca2p / (para_k1 + ca2p));
}
static inline double ca2p_reaction_factor(double u, double ER_domain3_deactive, double p, double B) {
    const double k_flux = g_P->in.k_flux;
    const double leak_from_storage = g_P->leak_from_storage;
    double ca2p_from_storage = k_flux * ER_domain1_active(p) * ER_domain2_active(u) * ER_domain3_deactive;
    return ca2p_from_storage - ca2p_into_storage(u) + leak_from_storage + ca2p_by_ext_stim(B);
}
static inline double calc_th(int is_alive, double agek) {
    const double thgra = g_P->in.thgra, delta_th = g_P->in.delta_th, thpri = g_P->in.thpri;
    return is_alive ? thgra + ((thpri - thgra) * 0.5) * (1.0 + tanh((THRESH_SP - agek) / delta_th)) : thpri;
}
static inline double calc_Kpa(int is_alive, double agek) {
    const double kpa = g_P->in.kpa, Kgra = kpa, delta_K = g_P->in.delta_K, Kpri = g_P->in.Kpri;
    return is_alive ? Kgra + ((Kpri - Kgra) * 0.5) * (1.0 + tanh((THRESH_SP - agek) / delta_K)) : Kpri;
}
static inline double calc_IAG(int is_alive, double agek) {
    const double delta_I = g_P->in.delta_I, iage_kitei = g_P->in.iage_kitei;
    return is_alive ? 0.5 * (1.0 + tanh((agek - THRESH_SP) / delta_I)) : iage_kitei;
}
static inline double ex_inert_diff(double ca2p, double current_ex_inert, double _th) {
    const double para_k2 = g_P->in.para_k2;
    return ((para_k2 * para_k2 / (para_k2 * para_k2 + ca2p * ca2p) - current_ex_inert) / _th);
}
static inline double IP3_default_diff(double _Kpa, double a_avg, double current_IP3) {
    const double H0 = g_P->in.H0;
    return (_Kpa * a_avg / (H0 + a_avg) - Kpp * current_IP3);
}
static inline double fw(double diff, double w) {
    const double wd = g_P->in.wd;
    return (1. - w) + (-1. + tanh((wd - diff) / 0.1)) / 2.;
}
static inline double grid_avg8(const eo_tissue* t, const double* g, int ix, int iy, int iz) {
    return 0.125 * (g[IDX3(t, ix, iy, iz)] + g[IDX3(t, ix + 1, iy, iz)] + g[IDX3(t, ix, iy + 1, iz)]
                    + g[IDX3(t, ix, iy, iz + 1)] + g[IDX3(t, ix + 1, iy + 1, iz)] + g[IDX3(t, ix + 1, iy, iz + 1)]
                    + g[IDX3(t, ix, iy + 1, iz + 1)] + g[IDX3(t, ix + 1, iy + 1, iz + 1)]);
}
static inline double fa(double diffu, double A) {
    const double STIM11 = g_P->in.STIM11, Kaa = g_P->in.Kaa;
    return STIM11 * min0(diffu) - A * Kaa;
}
static inline int is_afm(int st) { return st == EO_ALIVE || st == EO_FIX || st == EO_MUSUME; }

/* per-voxel cell_diffu_map / air_stim_flg (init_ca2p_map :238-278) as cell
 * index of an AFM owner (-1 = dummy) and air flag; the state does not change
 * during calc_ca2p, so a voxel's pointer target is fixed for the solve */
typedef struct { int32_t* diffu_cell; uint8_t* air; int iz_bound; } ca_maps;

static void init_ca2p_map(eo_tissue* t, ca_maps* m) {
    EO_PAR
    for (size_t i = 0; i < t->gsz; i++) { m->diffu_cell[i] = -1; m->air[i] = 0; }
    EO_PAR
    for (int j = 0; j < t->NX; j++)
        for (int k = 0; k < t->NY; k++)
            for (int l = 0; l < m->iz_bound; l++) {
                size_t v = IDX3(t, j, k, l);
                int c = t->cmap1[v];
                int tmp = -1; uint8_t asf = 0;
                if (c >= 0) {
                    if (is_afm(t->state[c])) tmp = c;
                    for (int cc = 0; cc < t->nconn[c]; ++cc)
                        if (t->state[conn_row(t, c)[cc]] == EO_AIR) { asf = 1; break; }
                }
                m->diffu_cell[v] = tmp; m->air[v] = asf;
            }
}

/* dead_IP3_calc :156-173 */
static void dead_IP3_calc(eo_tissue* t) {
    const double* ip3 = t->ip3_buf[t->ca_first];
    double* ip3n = t->ip3_buf[1 - t->ca_first];
    EO_PAR
    for (int c = t->nmemb + t->nder; c < t->n; c++) {
        if (t->state[c] == EO_DEAD) {
            int count = 0;
            double tmp = 0;
            int32_t* row = conn_row(t, c);
            for (int k = 0; k < t->nconn[c]; k++) {
                int o = row[k];
                if (t->state[o] == EO_ALIVE) { tmp += ip3[o]; count++; }
            }
            tmp = DT_Ca * (dp_c * (tmp - count * ip3[c]) - Kpp * ip3[c]);
            ip3n[c] = ip3[c] + tmp;
        }
    }
}

/* supra_calc :182-222 */
static void supra_calc(eo_tissue* t) {
    const double ca2p_du = g_P->in.ca2p_du;
    const double* ca = t->ca_buf[t->ca_first];
    double* can = t->ca_buf[1 - t->ca_first];
    const double* ip3 = t->ip3_buf[t->ca_first];
    double* ip3n = t->ip3_buf[1 - t->ca_first];
    const double* atp = t->atp[t->atp_first];
    const double* B = t->bst[t->b_first];
    EO_PAR_DYN
    for (int c = t->nmemb + t->nder; c < t->n; c++) {
        const int st = t->state[c];
        if (!is_afm(st)) continue;
        const int ix = t->lat[3 * c], iy = t->lat[3 * c + 1], iz = t->lat[3 * c + 2];
        const int is_alive = st == EO_ALIVE;
        const double _th = calc_th(is_alive, t->agek[c]);
        const double _Kpa = calc_Kpa(is_alive, t->agek[c]);
        const double IAGv = calc_IAG(is_alive, t->agek[c]);
        double tmp_diffu = 0, tmp_IP3 = 0;
        int32_t* row = conn_row(t, c);
        for (int k = 0; k < t->nconn[c]; k++) {
            int o = row[k];
            int so = t->state[o];
            if (so == EO_ALIVE || so == EO_DEAD || so == EO_FIX || so == EO_MUSUME) {
                gj_ent* e = gj_find(&t->gj[c], t->uid[o]);
                if (!e) { set_err(t, EO_ERR_INVALID); continue; }   /* gj.at() throws */
                tmp_diffu += e->v * (ca[o] - ca[c]);
                tmp_IP3 += e->v * (ip3[o] - ip3[c]);
                if (so == EO_ALIVE) e->v += DT_Ca * fw(fabs(ca[o] - ca[c]), e->v);
            }
        }
        t->diff_u[c] = ca2p_reaction_factor(ca[c], t->ex_inert[c], ip3[c], grid_avg8(t, B, ix, iy, iz))
                       + ca2p_du * IAGv * tmp_diffu;
        ip3n[c] = ip3[c] + DT_Ca * (IP3_default_diff(_Kpa, grid_avg8(t, atp, ix, iy, iz), ip3[c]) + dp_c * IAGv * tmp_IP3);
        const double cs = ca[c] + DT_Ca * t->diff_u[c];
        can[c] = cs;
        t->ca2p_avg[c] += cs;
        t->ex_inert[c] += DT_Ca * ex_inert_diff(ca[c], t->ex_inert[c], _th);
    }
}

/* ATP_refresh :302-338 */
static void ATP_refresh(eo_tissue* t, const ca_maps* m) {
    const double Da = g_P->in.Da, AIR_STIM = g_P->in.AIR_STIM;
    const int NX = t->NX, NY = t->NY, iz_bound = m->iz_bound;
    const double* carr = t->atp[t->atp_first];
    double* narr = t->atp[1 - t->atp_first];
    const uint8_t* cm2 = t->cmap2;
    const double inv_dx = t->inv_dx;
    EO_PAR2
    for (int j = 0; j < NX; j++) {
        for (int k = 0; k < NY; k++) {
            const int prev_x = t->px_prev[j], next_x = t->px_next[j];
            const int prev_y = t->py_prev[k], next_y = t->py_next[k];
            for (int l = 0; l < iz_bound; l++) {
                const int prev_z = t->pz_prev[l], next_z = t->pz_next[l];
                const size_t v = IDX3(t, j, k, l);
                const double catp = carr[v];
                const int dc = m->diffu_cell[v];
                const double diffu = dc >= 0 ? t->diff_u[dc] : 0.0;
                narr[v] = catp + DT_Ca * (Da * ((double)cm2[IDX3(t, prev_x, k, l)] * (carr[IDX3(t, prev_x, k, l)] - catp)
                                                + (double)cm2[IDX3(t, next_x, k, l)] * (carr[IDX3(t, next_x, k, l)] - catp)
                                                + (double)cm2[IDX3(t, j, prev_y, l)] * (carr[IDX3(t, j, prev_y, l)] - catp)
                                                + (double)cm2[IDX3(t, j, next_y, l)] * (carr[IDX3(t, j, next_y, l)] - catp)
                                                + (double)cm2[IDX3(t, j, k, prev_z)] * (carr[IDX3(t, j, k, prev_z)] - catp)
                                                + (double)cm2[IDX3(t, j, k, next_z)] * (carr[IDX3(t, j, k, next_z)] - catp))
                                          * inv_dx * inv_dx
                                          + fa(diffu, catp) + (double)m->air[v] * AIR_STIM);
            }
        }
    }
    EO_PAR
    for (int l = 0; l <= iz_bound; l++) {
        for (int j = 0; j <= NX; j++) narr[IDX3(t, j, NY, l)] = narr[IDX3(t, j, 0, l)];
        for (int k = 0; k <= NY; k++) narr[IDX3(t, NX, k, l)] = narr[IDX3(t, 0, k, l)];
    }
}

/* update_values :343-350 */
static void update_values(eo_tissue* t) {
    t->ca_first = 1 - t->ca_first;   /* ca2p_swap and IP3_swap swap together */
    for (int c = t->nmemb + t->nder; c < t->n; c++) t->diff_u[c] = 0;
    t->atp_first = 1 - t->atp_first;
}

static int ca_iz_bound(const eo_tissue* t, double zzmax) {
    return (int)((zzmax + FAC_MAP * R_max) / t->dz);  /* ca2p.cpp:386 */
}

static ca_maps* g_maps_cache = NULL;  /* not used: maps are per call */

void eo_ca_substep(eo_tissue* t, int iz_bound) {
    use(t);
    if (iz_bound > t->NZ || iz_bound < 0) { set_err(t, EO_ERR_OUT_OF_GRID); return; }
    ca_maps m;
    m.iz_bound = iz_bound;
    m.diffu_cell = (int32_t*)malloc(sizeof(int32_t) * t->gsz);
    m.air = (uint8_t*)malloc(t->gsz);
    init_ca2p_map(t, &m);
    dead_IP3_calc(t);
    supra_calc(t);
    ATP_refresh(t, &m);
    update_values(t);
    free(m.diffu_cell); free(m.air);
}

static void calc_ca2p_n(eo_tissue* t, double zzmax, unsigned n_sub);

void eo_calc_ca2p(eo_tissue* t, double zzmax) {
    use(t);
    calc_ca2p_n(t, zzmax, (unsigned)g_P->Ca_ITR);
}

/* calc_ca2p with a given number of sub-steps (timing samples only) */
void eo_calc_ca2p_n(eo_tissue* t, double zzmax, int n_sub) {
    use(t);
    calc_ca2p_n(t, zzmax, (unsigned)n_sub);
}

static void calc_ca2p_n(eo_tissue* t, double zzmax, unsigned n_sub) {
    (void)g_maps_cache;
    const unsigned Ca_ITR = (unsigned)g_P->Ca_ITR;
    /* init_ca2p_avg :20-27 */
    for (int c = t->nmemb + t->nder; c < t->n; c++) if (is_afm(t->state[c])) t->ca2p_avg[c] = 0;
    ca_maps m;
    m.iz_bound = ca_iz_bound(t, zzmax);
    if (m.iz_bound > t->NZ || m.iz_bound < 0) { set_err(t, EO_ERR_OUT_OF_GRID); return; }
    m.diffu_cell = (int32_t*)malloc(sizeof(int32_t) * t->gsz);
    m.air = (uint8_t*)malloc(t->gsz);
    init_ca2p_map(t, &m);
    for (unsigned cstp = 0; cstp < n_sub; cstp++) {
        dead_IP3_calc(t);
        supra_calc(t);
        ATP_refresh(t, &m);
        update_values(t);
    }
    /* set_cell_ca2p :355-367 */
    for (int c = t->nmemb + t->nder; c < t->n; c++) {
        if (is_afm(t->state[c])) t->ca2p_avg[c] /= Ca_ITR;
        else { t->ca2p_avg[c] = 0; t->ca_buf[0][c] = t->ca_buf[1][c] = 0; }
    }
    free(m.diffu_cell); free(m.air);
}

int eo_should_calc_ca(const eo_tissue* t) { return t->sw >= prm_sw_thresh(&t->P); }
void eo_ca_calc_condition_reset(eo_tissue* t) { t->sw = 0; }

/* CellManager::clean_up, cellmanager.cpp:543-557 */
void eo_clean_up(eo_tissue* t) {
    use(t);
    for (int c = 0; c < t->n; c++) {
        gj_map* m = &t->gj[c];
        int w = 0;
        for (int i = 0; i < m->n; i++) if (m->e[i].checked) m->e[w++] = m->e[i];
        m->n = w;
    }
}

/* proc.cpp:131-173 loop body; cell_dynamics :22-32 */
int eo_step(eo_tissue* t, eo_driver* d) {
    use(t);
    t->timestep = d->step;   /* cman.current_timestep = i (proc.cpp:132) */
    if (d->clean_every > 0 && d->step % d->clean_every == 0 && d->step > 0) eo_clean_up(t);
    eo_cell_interaction(t);
    eo_pos_copy(t);
    eo_cell_state_renew(t);
    eo_cell_pos_periodic_fix(t);
    eo_connect_cell(t);
    double zzmax = eo_calc_zzmax(t);
    eo_setup_map_lat(t);
    eo_calc_ext_stim(t, zzmax);
    if ((double)d->step * DT_Cell > T_TURNOVER && d->force_cornif) {
        d->force_cornif = 0;
        eo_initialize_sc(t, zzmax);
        d->num_sc = NUM_SC_INIT;
    }
    int ran = 0;
    int forced = d->ca_every > 0 && ((d->step + 1) % d->ca_every) == 0;
    if (eo_should_calc_ca(t) || d->num_sc > 0 || forced) {
        eo_calc_ca2p(t, zzmax);
        if (d->num_sc > 0) d->num_sc--;
        eo_ca_calc_condition_reset(t);
        ran = 1;
    }
    d->step++;
    return ran;
}

/* ---- cell_stat ----------------------------------------------------------- */
void eo_set_timestep(eo_tissue* t, int64_t step) { t->timestep = step; }
int64_t eo_timestep(const eo_tissue* t) { return t->timestep; }
/* the per-cell records, 3 per cell */
void eo_get_cst(const eo_tissue* t, int64_t* out) { memcpy(out, t->cst, sizeof(int64_t) * 3 * (size_t)t->n); }
/* append_stat_data (cellmanager.cpp:559-567): the records of the DEAD cells
 * removed since the last call, in removal order, then the store is cleared */
int eo_take_cell_stats(eo_tissue* t, int64_t* out, int max_records, int* n_records) {
    if (t->nstat > max_records) return EO_ERR_CAPACITY;
    memcpy(out, t->stat, sizeof(int64_t) * 3 * (size_t)t->nstat);
    *n_records = t->nstat;
    t->nstat = 0;
    return EO_OK;
}

/* ---- state exchange ---------------------------------------------------- */
int eo_get_cells(const eo_tissue* t, eo_cells* o) {
    o->n = t->n; o->nmemb = t->nmemb; o->nder = t->nder; o->sw = t->sw;
    int nnz = 0;
    for (int c = 0; c < t->n; c++) nnz += t->nconn[c];
    if (o->conn_ptr) {
        int acc = 0;
        for (int c = 0; c < t->n; c++) { o->conn_ptr[c] = acc; acc += t->nconn[c]; }
        o->conn_ptr[t->n] = acc;
    }
    if (o->conn && nnz > o->conn_cap) return EO_ERR_CAPACITY;
    for (int c = 0, p = 0; c < t->n; c++) {
        const int32_t* row = &t->conn[(size_t)c * MAX_CONNECT_CELL_NUM];
        for (int k = 0; k < t->nconn[c]; k++, p++) {
            if (o->conn) o->conn[p] = row[k];
            if (o->gj) {
                gj_ent* e = gj_find(&t->gj[c], t->uid[row[k]]);
                o->gj[p] = e ? e->v : -1.0;
            }
        }
    }
#define CP(dst, src, ty) if (o->dst) memcpy(o->dst, t->src, sizeof(ty) * (size_t)t->n)
    CP(state, state, uint8_t); CP(fix_origin, fix_origin, int32_t);
    CP(x, x1, double); CP(y, y1, double); CP(z, z1, double);
    CP(xn, x2, double); CP(yn, y2, double); CP(zn, z2, double);
    CP(radius, radius, double);
    CP(ca2p_avg, ca2p_avg, double); CP(ex_inert, ex_inert, double); CP(agek, agek, double);
    CP(ageb, ageb, double); CP(ex_fat, ex_fat, double); CP(in_fat, in_fat, double);
    CP(spr_nat_len, spr_nat_len, double); CP(diff_u, diff_u, double);
    CP(rest_div_times, rest_div_times, int32_t); CP(pair, pair, int32_t);
    CP(dermis, dermis, int32_t); CP(is_malignant, is_malignant, uint8_t); CP(uid, uid, int64_t);
#undef CP
    if (o->lat) memcpy(o->lat, t->lat, sizeof(int32_t) * 3 * (size_t)t->n);
    if (o->ca2p) memcpy(o->ca2p, t->ca_buf[t->ca_first], 8 * (size_t)t->n);
    if (o->ca2p_next) memcpy(o->ca2p_next, t->ca_buf[1 - t->ca_first], 8 * (size_t)t->n);
    if (o->ip3) memcpy(o->ip3, t->ip3_buf[t->ca_first], 8 * (size_t)t->n);
    if (o->ip3_next) memcpy(o->ip3_next, t->ip3_buf[1 - t->ca_first], 8 * (size_t)t->n);
    if (o->memb_adj) memcpy(o->memb_adj, t->memb_adj, sizeof(int32_t) * 6 * (size_t)t->nmemb);
    return EO_OK;
}

int eo_set_cell_values(eo_tissue* t, const double* ca2p, const double* ip3,
                       const double* ex_inert, const double* ca2p_avg, const double* gj) {
    use(t);
    for (int c = 0; c < t->n; c++) {
        if (ca2p) t->ca_buf[0][c] = t->ca_buf[1][c] = ca2p[c];
        if (ip3) t->ip3_buf[0][c] = t->ip3_buf[1][c] = ip3[c];
        if (ex_inert) t->ex_inert[c] = ex_inert[c];
        if (ca2p_avg) t->ca2p_avg[c] = ca2p_avg[c];
    }
    if (gj) {
        int p = 0;
        for (int c = 0; c < t->n; c++)
            for (int k = 0; k < t->nconn[c]; k++, p++)
                gj_get(&t->gj[c], t->uid[conn_row(t, c)[k]])->v = gj[p];
    }
    return EO_OK;
}

int eo_get_fields(const eo_tissue* t, double* atp, double* atp_next, double* b,
                  double* b_next, int32_t* cmap1, uint8_t* cmap2) {
    size_t s = t->gsz;
    if (atp) memcpy(atp, t->atp[t->atp_first], 8 * s);
    if (atp_next) memcpy(atp_next, t->atp[1 - t->atp_first], 8 * s);
    if (b) memcpy(b, t->bst[t->b_first], 8 * s);
    if (b_next) memcpy(b_next, t->bst[1 - t->b_first], 8 * s);
    if (cmap1) memcpy(cmap1, t->cmap1, 4 * s);
    if (cmap2) memcpy(cmap2, t->cmap2, s);
    return EO_OK;
}

int eo_set_fields(eo_tissue* t, const double* atp, const double* atp_next,
                  const double* b, const double* b_next) {
    use(t);
    size_t s = t->gsz;
    if (atp) memcpy(t->atp[t->atp_first], atp, 8 * s);
    if (atp_next) memcpy(t->atp[1 - t->atp_first], atp_next, 8 * s);
    if (b) memcpy(t->bst[t->b_first], b, 8 * s);
    if (b_next) memcpy(t->bst[1 - t->b_first], b_next, 8 * s);
    return EO_OK;
}

int eo_set_maps(eo_tissue* t, const int32_t* cmap1, const uint8_t* cmap2, const int32_t* lat) {
    use(t);
    if (cmap1) memcpy(t->cmap1, cmap1, 4 * t->gsz);
    if (cmap2) memcpy(t->cmap2, cmap2, t->gsz);
    if (lat) memcpy(t->lat, lat, sizeof(int32_t) * 3 * (size_t)t->n);
    return EO_OK;
}

int eo_set_positions(eo_tissue* t, const double* x, const double* y, const double* z) {
    use(t);
    for (int c = 0; c < t->n; c++) {
        t->x1[c] = t->x2[c] = x[c]; t->y1[c] = t->y2[c] = y[c]; t->z1[c] = t->z2[c] = z[c];
    }
    return EO_OK;
}

int eo_set_state(eo_tissue* t, const uint8_t* state, const double* agek) {
    use(t);
    for (int c = 0; c < t->n; c++) {
        if (state) t->state[c] = state[c];
        if (agek) t->agek[c] = agek[c];
    }
    return EO_OK;
}

int eo_set_lists(eo_tissue* t, const int32_t* conn_ptr, const int32_t* conn, const double* gj) {
    use(t);
    for (int c = 0; c < t->n; c++) {
        int len = conn_ptr[c + 1] - conn_ptr[c];
        if (len > MAX_CONNECT_CELL_NUM) return EO_ERR_CAPACITY;
        t->nconn[c] = len;
        memcpy(conn_row(t, c), conn + conn_ptr[c], sizeof(int32_t) * (size_t)len);
        t->gj[c].n = 0;
        for (int k = 0; k < len; k++) {
            gj_ent* e = gj_get(&t->gj[c], t->uid[conn[conn_ptr[c] + k]]);
            e->v = gj ? gj[conn_ptr[c] + k] : gj_init;
            e->checked = 1;
        }
    }
    return EO_OK;
}

double eo_p_diff_x(const eo_tissue* t, double a, double b) { return p_diff_x(t, a, b); }

/* ==== frozen reaction-diffusion window and its adjoint ====================
 * The sub-step of calc_ca2p (ca2p.cpp:392-399: dead_IP3_calc, supra_calc,
 * ATP_refresh, update_values) run n_sub times with positions, lists, states
 * and voxel maps frozen at the tissue's current state, and with four
 * parameter groups made runtime values:
 *   du   the Ca2+ gap-junction coefficient ca2p_du (0.01, ca2p.cpp:183)
 *   dp   the IP3 coefficient dp (0.1, ca2p.cpp:17), in supra_calc and
 *        dead_IP3_calc
 *   gj0  the starting conductance of every directed edge (gj_init = 0.99)
 *   kpa  per-cell IP3 production rate (calc_Kpa, ca2p.cpp:101-110)
 * With du = 0.01, dp = 0.1, kpa = calc_Kpa and gj0 = 0.99 on fresh edges a
 * window is calc_ca2p's loop.  The adjoint is the reverse pass of the same
 * arithmetic for the loss L = mean over AFM cells of (ca2p_T - target)^2.
 * Values the window never writes (ca2p of DEAD cells, grid planes from
 * iz_bound up, both alternating between the two buffers) are constants. */

typedef struct {
    int n, start, NX, NY, NZ, izb;
    size_t gsz;
    int32_t* ptr; int32_t* col; uint8_t* kind;    /* CSR lists, 0 / 1 coupled / 3 coupled ALIVE */
    double *th, *iag, *kpa0;
    int32_t* owner; uint8_t* air;                 /* voxel -> AFM cell, air flag */
    double *u0[2], *p0[2], *h0, *A0[2];
    const double* B;
} rd_win;

static void rd_free(rd_win* w) {
    free(w->ptr); free(w->col); free(w->kind); free(w->th); free(w->iag); free(w->kpa0);
    free(w->owner); free(w->air);
    for (int b = 0; b < 2; b++) { free(w->u0[b]); free(w->p0[b]); free(w->A0[b]); }
    free(w->h0);
}

static int rd_setup(eo_tissue* t, double zzmax, rd_win* w) {
    memset(w, 0, sizeof(*w));
    w->n = t->n; w->start = t->nmemb + t->nder;
    w->NX = t->NX; w->NY = t->NY; w->NZ = t->NZ;
    w->izb = ca_iz_bound(t, zzmax);
    if (w->izb > t->NZ || w->izb < 0) return EO_ERR_OUT_OF_GRID;
    w->gsz = t->gsz;
    const int n = t->n;
    w->ptr = xcalloc(n + 1, 4);
    for (int c = 0; c < n; c++) w->ptr[c + 1] = w->ptr[c] + t->nconn[c];
    w->col = xcalloc(w->ptr[n], 4); w->kind = xcalloc(w->ptr[n], 1);
    for (int c = 0; c < n; c++) {
        int32_t* row = conn_row(t, c);
        for (int e = 0; e < t->nconn[c]; e++) {
            int o = row[e], so = t->state[o];
            w->col[w->ptr[c] + e] = o;
            w->kind[w->ptr[c] + e] = so == EO_ALIVE ? 3 : (so == EO_DEAD || so == EO_FIX || so == EO_MUSUME) ? 1 : 0;
        }
    }
    w->th = xcalloc(n, 8); w->iag = xcalloc(n, 8); w->kpa0 = xcalloc(n, 8);
    for (int c = w->start; c < n; c++) {
        if (!is_afm(t->state[c])) continue;
        int al = t->state[c] == EO_ALIVE;
        w->th[c] = calc_th(al, t->agek[c]); w->iag[c] = calc_IAG(al, t->agek[c]); w->kpa0[c] = calc_Kpa(al, t->agek[c]);
    }
    ca_maps m;
    m.iz_bound = w->izb;
    w->owner = m.diffu_cell = (int32_t*)malloc(sizeof(int32_t) * t->gsz);
    w->air = m.air = (uint8_t*)malloc(t->gsz);
    init_ca2p_map(t, &m);
    for (int b = 0; b < 2; b++) {
        w->u0[b] = xcalloc(n, 8); w->p0[b] = xcalloc(n, 8); w->A0[b] = xcalloc(t->gsz, 8);
        memcpy(w->u0[b], t->ca_buf[b ? 1 - t->ca_first : t->ca_first], 8 * (size_t)n);
        memcpy(w->p0[b], t->ip3_buf[b ? 1 - t->ca_first : t->ca_first], 8 * (size_t)n);
        memcpy(w->A0[b], t->atp[b ? 1 - t->atp_first : t->atp_first], 8 * t->gsz);
    }
    w->h0 = xcalloc(n, 8);
    memcpy(w->h0, t->ex_inert, 8 * (size_t)n);
    w->B = t->bst[t->b_first];
    return EO_OK;
}

typedef struct { double *u[2], *p[2], *h, *g, *A[2], *du; int cur; } rd_state;

static void rd_state_init(const rd_win* w, rd_state* s, double gj0) {
    for (int b = 0; b < 2; b++) {
        s->u[b] = xcalloc(w->n, 8); s->p[b] = xcalloc(w->n, 8); s->A[b] = xcalloc(w->gsz, 8);
        memcpy(s->u[b], w->u0[b], 8 * (size_t)w->n);
        memcpy(s->p[b], w->p0[b], 8 * (size_t)w->n);
        memcpy(s->A[b], w->A0[b], 8 * w->gsz);
    }
    s->h = xcalloc(w->n, 8); memcpy(s->h, w->h0, 8 * (size_t)w->n);
    s->g = xcalloc(w->ptr[w->n], 8);
    for (int e = 0; e < w->ptr[w->n]; e++) s->g[e] = gj0;
    s->du = xcalloc(w->n, 8);
    s->cur = 0;
}
static void rd_state_free(rd_state* s) {
    for (int b = 0; b < 2; b++) { free(s->u[b]); free(s->p[b]); free(s->A[b]); }
    free(s->h); free(s->g); free(s->du);
}

static inline double rd_kpa(const rd_win* w, const eo_rd_params* P, int c) { return P->kpa ? P->kpa[c] : w->kpa0[c]; }

/* one sub-step in the order of ca2p.cpp:392-399 */
static void rd_substep(eo_tissue* t, const rd_win* w, const eo_rd_params* P, rd_state* s) {
    const double* u = s->u[s->cur]; double* un = s->u[1 - s->cur];
    const double* p = s->p[s->cur]; double* pn = s->p[1 - s->cur];
    const double* A = s->A[s->cur]; double* An = s->A[1 - s->cur];
    for (int c = w->start; c < w->n; c++) {          /* dead_IP3_calc */
        if (t->state[c] != EO_DEAD) continue;
        int count = 0; double tmp = 0;
        for (int e = w->ptr[c]; e < w->ptr[c + 1]; e++)
            if (w->kind[e] == 3) { tmp += p[w->col[e]]; count++; }
        tmp = DT_Ca * (P->dp * (tmp - count * p[c]) - Kpp * p[c]);
        pn[c] = p[c] + tmp;
    }
    for (int c = 0; c < w->n; c++) s->du[c] = 0;
    for (int c = w->start; c < w->n; c++) {          /* supra_calc */
        if (!is_afm(t->state[c])) continue;
        const int ix = t->lat[3 * c], iy = t->lat[3 * c + 1], iz = t->lat[3 * c + 2];
        double su = 0, sp = 0;
        for (int e = w->ptr[c]; e < w->ptr[c + 1]; e++) {
            if (!w->kind[e]) continue;
            const int o = w->col[e];
            su += s->g[e] * (u[o] - u[c]);
            sp += s->g[e] * (p[o] - p[c]);
            if (w->kind[e] == 3) s->g[e] += DT_Ca * fw(fabs(u[o] - u[c]), s->g[e]);
        }
        s->du[c] = ca2p_reaction_factor(u[c], s->h[c], p[c], grid_avg8(t, w->B, ix, iy, iz)) + P->du * w->iag[c] * su;
        pn[c] = p[c] + DT_Ca * (IP3_default_diff(rd_kpa(w, P, c), grid_avg8(t, A, ix, iy, iz), p[c]) + P->dp * w->iag[c] * sp);
        un[c] = u[c] + DT_Ca * s->du[c];
        s->h[c] += DT_Ca * ex_inert_diff(u[c], s->h[c], w->th[c]);
    }
    /* ATP_refresh */
    const double Da = g_P->in.Da, AIR_STIM = g_P->in.AIR_STIM;
    const uint8_t* cm2 = t->cmap2;
    const double inv_dx = t->inv_dx;
    for (int j = 0; j < w->NX; j++) {
        const int px = t->px_prev[j], nx = t->px_next[j];
        for (int k = 0; k < w->NY; k++) {
            const int py = t->py_prev[k], ny = t->py_next[k];
            for (int l = 0; l < w->izb; l++) {
                const int pz = t->pz_prev[l], nz = t->pz_next[l];
                const size_t v = IDX3(t, j, k, l);
                const double a = A[v];
                const int dc = w->owner[v];
                const double diffu = dc >= 0 ? s->du[dc] : 0.0;
                An[v] = a + DT_Ca * (Da * ((double)cm2[IDX3(t, px, k, l)] * (A[IDX3(t, px, k, l)] - a)
                                           + (double)cm2[IDX3(t, nx, k, l)] * (A[IDX3(t, nx, k, l)] - a)
                                           + (double)cm2[IDX3(t, j, py, l)] * (A[IDX3(t, j, py, l)] - a)
                                           + (double)cm2[IDX3(t, j, ny, l)] * (A[IDX3(t, j, ny, l)] - a)
                                           + (double)cm2[IDX3(t, j, k, pz)] * (A[IDX3(t, j, k, pz)] - a)
                                           + (double)cm2[IDX3(t, j, k, nz)] * (A[IDX3(t, j, k, nz)] - a))
                                     * inv_dx * inv_dx + fa(diffu, a) + (double)w->air[v] * AIR_STIM);
            }
        }
    }
    for (int l = 0; l <= w->izb; l++) {
        for (int j = 0; j <= w->NX; j++) An[IDX3(t, j, w->NY, l)] = An[IDX3(t, j, 0, l)];
        for (int k = 0; k <= w->NY; k++) An[IDX3(t, w->NX, k, l)] = An[IDX3(t, 0, k, l)];
    }
    s->cur = 1 - s->cur;
}

int eo_rd_forward(eo_tissue* t, double zzmax, const eo_rd_params* P, int n_sub, double* ca2p_out) {
    use(t);
    rd_win w; rd_state s;
    int rc = rd_setup(t, zzmax, &w);
    if (rc) { rd_free(&w); return rc; }
    rd_state_init(&w, &s, P->gj0);
    for (int i = 0; i < n_sub; i++) rd_substep(t, &w, P, &s);
    if (ca2p_out) memcpy(ca2p_out, s.u[s.cur], 8 * (size_t)w.n);
    rd_state_free(&s); rd_free(&w);
    return EO_OK;
}

/* map a grid corner read by grid_avg8 to the variable it reads: the copy
 * planes x = NX and y = NY hold the values of x = 0 and y = 0 written by the
 * last ATP_refresh; planes from iz_bound up are constants (-1) */
static inline long rd_corner_var(const eo_tissue* t, const rd_win* w, int j, int k, int l) {
    if (l >= w->izb) return -1;
    if (j == w->NX) j = 0;
    if (k == w->NY) k = 0;
    return (long)IDX3(t, j, k, l);
}

int eo_rd_adjoint(eo_tissue* t, double zzmax, const eo_rd_params* P, int n_sub, const double* target,
                  double* loss, eo_rd_grad* G) {
    use(t);
    rd_win w; rd_state s;
    int rc = rd_setup(t, zzmax, &w);
    if (rc) { rd_free(&w); return rc; }
    const int n = w.n, E = w.ptr[n];
    const size_t gsz = w.gsz;
    rd_state_init(&w, &s, P->gj0);
    /* forward, keeping every state */
    double** Us = malloc(sizeof(double*) * (n_sub + 1));
    double** Ps = malloc(sizeof(double*) * (n_sub + 1));
    double** Hs = malloc(sizeof(double*) * (n_sub + 1));
    double** Gs = malloc(sizeof(double*) * (n_sub + 1));
    double** As = malloc(sizeof(double*) * (n_sub + 1));
    for (int i = 0; i <= n_sub; i++) {
        Us[i] = xcalloc(n, 8); Ps[i] = xcalloc(n, 8); Hs[i] = xcalloc(n, 8); Gs[i] = xcalloc(E ? E : 1, 8);
        As[i] = xcalloc(gsz, 8);
        memcpy(Us[i], s.u[s.cur], 8 * (size_t)n); memcpy(Ps[i], s.p[s.cur], 8 * (size_t)n);
        memcpy(Hs[i], s.h, 8 * (size_t)n); memcpy(Gs[i], s.g, 8 * (size_t)E); memcpy(As[i], s.A[s.cur], 8 * gsz);
        if (i < n_sub) rd_substep(t, &w, P, &s);
    }
    /* loss and its seed */
    int nafm = 0;
    for (int c = w.start; c < n; c++) nafm += is_afm(t->state[c]);
    double L = 0;
    double *lu = xcalloc(n, 8), *lp = xcalloc(n, 8), *lh = xcalloc(n, 8), *lg = xcalloc(E ? E : 1, 8),
           *lA = xcalloc(gsz, 8);
    double *nu = xcalloc(n, 8), *np_ = xcalloc(n, 8), *nh = xcalloc(n, 8), *ng = xcalloc(E ? E : 1, 8),
           *nA = xcalloc(gsz, 8), *accA = xcalloc(n, 8);
    for (int c = w.start; c < n; c++)
        if (is_afm(t->state[c])) {
            const double d = Us[n_sub][c] - target[c];
            L += d * d;
            lu[c] = 2.0 * d / nafm;
        }
    *loss = L / nafm;
    G->du = G->dp = G->gj0 = 0;
    if (G->kpa) for (int c = 0; c < n; c++) G->kpa[c] = 0;
    const double Da = g_P->in.Da, Kaa = g_P->in.Kaa, STIM11 = g_P->in.STIM11, H0 = g_P->in.H0;
    const double kflux = g_P->in.k_flux, mu0 = g_P->in.mu0, mu1 = g_P->in.mu1, kmu = g_P->in.kmu,
                 pb = g_P->in.para_b, pbb = g_P->in.para_bb, pk1 = g_P->in.para_k1;
    const double kg = g_P->in.kg, gam = g_P->in.gamma, k2 = g_P->in.para_k2, wd = g_P->in.wd;
    const double cD = DT_Ca * Da * t->inv_dx * t->inv_dx;
    const uint8_t* cm2 = t->cmap2;
    for (int i = n_sub - 1; i >= 0; i--) {
        const double *u = Us[i], *p = Ps[i], *h = Hs[i], *g = Gs[i], *A = As[i];
        memset(nu, 0, 8 * (size_t)n); memset(np_, 0, 8 * (size_t)n); memset(nh, 0, 8 * (size_t)n);
        memset(ng, 0, 8 * (size_t)E); memset(nA, 0, 8 * gsz); memset(accA, 0, 8 * (size_t)n);
        /* voxel adjoints gathered per owner cell (for the fa source term) */
        for (int j = 0; j < w.NX; j++)
            for (int k = 0; k < w.NY; k++)
                for (int l = 0; l < w.izb; l++) {
                    const size_t v = IDX3(t, j, k, l);
                    if (w.owner[v] >= 0) accA[w.owner[v]] += lA[v];
                }
        for (int c = w.start; c < n; c++) {
            const int st = t->state[c];
            if (st == EO_DEAD) {
                int cnt = 0; double sum = 0;
                for (int e = w.ptr[c]; e < w.ptr[c + 1]; e++)
                    if (w.kind[e] == 3) { sum += p[w.col[e]]; cnt++; }
                np_[c] += lp[c] * (1.0 - DT_Ca * P->dp * cnt - DT_Ca * Kpp);
                for (int e = w.ptr[c]; e < w.ptr[c + 1]; e++)
                    if (w.kind[e] == 3) np_[w.col[e]] += lp[c] * DT_Ca * P->dp;
                G->dp += lp[c] * DT_Ca * (sum - cnt * p[c]);
                continue;
            }
            if (!is_afm(st)) continue;
            const int ix = t->lat[3 * c], iy = t->lat[3 * c + 1], iz = t->lat[3 * c + 2];
            double su = 0, sp = 0, gsum = 0;
            for (int e = w.ptr[c]; e < w.ptr[c + 1]; e++) {
                if (!w.kind[e]) continue;
                const int o = w.col[e];
                su += g[e] * (u[o] - u[c]); sp += g[e] * (p[o] - p[c]); gsum += g[e];
            }
            const double Bavg = grid_avg8(t, w.B, ix, iy, iz), Aavg = grid_avg8(t, A, ix, iy, iz);
            const double iag = w.iag[c], kpa = rd_kpa(&w, P, c);
            const double diffu = ca2p_reaction_factor(u[c], h[c], p[c], Bavg) + P->du * iag * su;
            const double e1 = mu0 + mu1 * p[c] / (p[c] + kmu), e2 = pb + pbb * u[c] / (pk1 + u[c]);
            const double dRdu = kflux * e1 * (pbb * pk1 / ((pk1 + u[c]) * (pk1 + u[c]))) * h[c]
                                - gam * kg / ((kg + u[c]) * (kg + u[c]));
            const double dRdh = kflux * e1 * e2;
            const double dRdp = kflux * (mu1 * kmu / ((p[c] + kmu) * (p[c] + kmu))) * e2 * h[c];
            const double abar = DT_Ca * lu[c] + (diffu > 0 ? accA[c] * DT_Ca * STIM11 : 0.0);
            const double q = k2 * k2 + u[c] * u[c];
            nu[c] += lu[c] + abar * (dRdu - P->du * iag * gsum) + lh[c] * DT_Ca * (-2.0 * u[c] * k2 * k2 / (q * q)) / w.th[c];
            np_[c] += lp[c] * (1.0 - DT_Ca * Kpp - DT_Ca * P->dp * iag * gsum) + abar * dRdp;
            nh[c] += lh[c] * (1.0 - DT_Ca / w.th[c]) + abar * dRdh;
            for (int e = w.ptr[c]; e < w.ptr[c + 1]; e++) {
                if (!w.kind[e]) continue;
                const int o = w.col[e];
                double to_u = abar * P->du * iag * g[e];
                double lge = lg[e];
                if (w.kind[e] == 3) {
                    const double d = u[o] - u[c];
                    const double th_ = tanh((wd - fabs(d)) / 0.1);
                    const double fwd = -5.0 * (1.0 - th_ * th_);
                    const double sg = d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
                    to_u += lg[e] * DT_Ca * fwd * sg;
                    nu[c] -= lg[e] * DT_Ca * fwd * sg;
                    lge = lg[e] * (1.0 - DT_Ca);
                }
                ng[e] = lge + abar * P->du * iag * (u[o] - u[c]) + lp[c] * DT_Ca * P->dp * iag * (p[o] - p[c]);
                if (is_afm(t->state[o])) nu[o] += to_u;
                np_[o] += lp[c] * DT_Ca * P->dp * iag * g[e];
            }
            /* grid_avg8 of ATP in the IP3 production term */
            const double wA = lp[c] * DT_Ca * kpa * H0 / ((H0 + Aavg) * (H0 + Aavg)) * 0.125;
            for (int q8 = 0; q8 < 8; q8++) {
                const long vv = rd_corner_var(t, &w, ix + (q8 & 1), iy + ((q8 >> 1) & 1), iz + ((q8 >> 2) & 1));
                if (vv >= 0) nA[vv] += wA;
            }
            G->du += abar * iag * su;
            G->dp += lp[c] * DT_Ca * iag * sp;
            if (G->kpa) G->kpa[c] += lp[c] * DT_Ca * Aavg / (H0 + Aavg);
        }
        /* ATP_refresh: every updated point w reads its six neighbours */
        for (int j = 0; j < w.NX; j++) {
            const int px = t->px_prev[j], nx = t->px_next[j];
            for (int k = 0; k < w.NY; k++) {
                const int py = t->py_prev[k], ny = t->py_next[k];
                for (int l = 0; l < w.izb; l++) {
                    const size_t v = IDX3(t, j, k, l);
                    const int pz = t->pz_prev[l], nz = t->pz_next[l];
                    const size_t nb[6] = {IDX3(t, px, k, l), IDX3(t, nx, k, l), IDX3(t, j, py, l),
                                          IDX3(t, j, ny, l), IDX3(t, j, k, pz), IDX3(t, j, k, nz)};
                    const int nbl[6] = {l, l, l, l, pz, nz};
                    double msum = 0;
                    for (int q6 = 0; q6 < 6; q6++) msum += (double)cm2[nb[q6]];
                    nA[v] += lA[v] * (1.0 - cD * msum - DT_Ca * Kaa);
                    for (int q6 = 0; q6 < 6; q6++)
                        if (nbl[q6] < w.izb) nA[nb[q6]] += lA[v] * cD * (double)cm2[nb[q6]];
                }
            }
        }
        double* tmp;
        tmp = lu; lu = nu; nu = tmp;
        tmp = lp; lp = np_; np_ = tmp;
        tmp = lh; lh = nh; nh = tmp;
        tmp = lg; lg = ng; ng = tmp;
        tmp = lA; lA = nA; nA = tmp;
    }
    for (int c = w.start; c < n; c++)
        if (is_afm(t->state[c]))
            for (int e = w.ptr[c]; e < w.ptr[c + 1]; e++)
                if (w.kind[e]) G->gj0 += lg[e];
    for (int i = 0; i <= n_sub; i++) { free(Us[i]); free(Ps[i]); free(Hs[i]); free(Gs[i]); free(As[i]); }
    free(Us); free(Ps); free(Hs); free(Gs); free(As);
    free(lu); free(lp); free(lh); free(lg); free(lA); free(nu); free(np_); free(nh); free(ng); free(nA); free(accA);
    rd_state_free(&s); rd_free(&w);
    return EO_OK;
}

/* ==== frozen mechanics window and its adjoint =============================
 * cell_interaction (cell_interaction.cpp:648-685) + pos_copy repeated
 * n_steps times with lists, states, pairs and dermis frozen and without the
 * periodic fix (a smooth function of the positions), the parameters eps_m,
 * K_TOTAL (K_DESMOSOME = 0.01 K_TOTAL), Kspring and KBEND at runtime.  Which
 * interactions apply depends only on the frozen data, so each step applies
 * the same list of (law, c1, c2); the reverse pass walks it backwards with
 * the derivative of the active branch of each law. */

typedef struct { int law, c1, c2; } mech_app;
static mech_app* g_tape = NULL;
static int g_tape_n = 0, g_tape_cap = 0;

static void tape_push(int law, int c1, int c2) {
    if (g_tape_n == g_tape_cap) {
        g_tape_cap = g_tape_cap ? 2 * g_tape_cap : 1024;
        g_tape = realloc(g_tape, sizeof(mech_app) * g_tape_cap);
    }
    g_tape[g_tape_n].law = law; g_tape[g_tape_n].c1 = c1; g_tape[g_tape_n].c2 = c2;
    g_tape_n++;
}

/* coef of a law at the current positions and its partial derivatives with
 * respect to the squared distance s and to eps_m, K_TOTAL and Kspring */
typedef struct { double coef, ds, deps, dadh, dspr; } law_d;

static law_d ljmain_d(double s, double R, double eps) {
    law_d r = {0, 0, 0, 0, 0};
    const double L = POW6(R) / POW3(s);
    r.coef = 4.0 * eps * L * (L - 1) / s;
    r.ds = 4.0 * eps * (4.0 * L - 7.0 * L * L) / (s * s);
    r.deps = 4.0 * L * (L - 1) / s;
    return r;
}
/* adhesion with spring constant K: h(d) = -(K/d) x (1 - x^2/0.04), x = d/R - 1 */
static law_d adhesion_d(double s, double R, double K) {
    law_d r = {0, 0, 0, 0, 0};
    const double d = sqrt(s);
    const double x = d / R - 1;
    if (x + 1 > LJ_THRESH) return r;
    const double c2 = (LJ_THRESH - 1.0) * (LJ_THRESH - 1.0);
    const double g = x * (1 - x * x / c2);
    r.coef = -(K / d) * g;
    const double dg = (1 - 3 * x * x / c2) / R;
    const double dd = K * g / (d * d) - (K / d) * dg;
    r.ds = dd / (2 * d);
    return r;
}

static law_d coef_law_d(eo_tissue* t, int law, int c1, int c2) {
    const double s = cell_dist_sq(t, c1, c2);
    const double R = t->radius[c1] + t->radius[c2];
    const double eps = g_mp.eps;
    law_d r = {0, 0, 0, 0, 0};
    switch (law) {
    case LAW_OTHER:
        return ljmain_d(s, R, eps);
    case LAW_DER_DER: {
        const double d = sqrt(s), D = d + delta_R;
        const double L = POW6(R) / POW6(D);
        const double dL = -6.0 * L / D;
        double g, dgd;
        if (D < R) {
            g = L * (L - 1) / (d * D);
            dgd = (2 * L - 1) * dL / (d * D) - L * (L - 1) * (D + d) / ((d * D) * (d * D));
        } else if (s < R * R) {
            return r;   /* para_ljp2 */
        } else {
            g = L * (L - 1) / (d * d);
            dgd = (2 * L - 1) * dL / (d * d) - 2 * L * (L - 1) / (d * d * d);
        }
        r.coef = 4.0 * eps * g + para_ljp2;
        r.ds = 4.0 * eps * dgd / (2 * d);
        r.deps = 4.0 * g;
        return r;
    }
    case LAW_ALAIRDE:
        if (s < R * R) return ljmain_d(s, R, eps);
        else {
            const int strong = t->agek[c1] > THRESH_SP && t->agek[c2] > THRESH_SP;
            const double K = strong ? g_mp.k_adh : g_mp.k_adh * K_DESMOSOME_RATIO;
            r = adhesion_d(s, R, K);
            r.dadh = r.coef / g_mp.k_adh;
            return r;
        }
    case LAW_FIXMU:
        if (s < R * R) return ljmain_d(s, R, eps);
        r = adhesion_d(s, R, g_mp.k_adh * K_DESMOSOME_RATIO);
        r.dadh = r.coef / g_mp.k_adh;
        return r;
    case LAW_MU_MEMB:
        if (s < R * R) return ljmain_d(s, R, eps);
        r = adhesion_d(s, R, t->spring_force_to_memb[c1]);
        {
            const int pr = t->pair[c1];
            if (pr >= 0 && t->state[pr] == EO_FIX && t->spring_force_to_memb[c1] != 0)
                r.dspr = r.coef / g_mp.k_spring;
        }
        return r;
    case LAW_FIX_MEMB:
        if (s < R * R) return ljmain_d(s, R, eps);
        else {
            const double d = sqrt(s);
            r.coef = -(g_mp.k_spring / d) * (d / R - 1.0);
            r.ds = -g_mp.k_spring / (2 * d * d * d);
            r.dspr = -(1.0 / d) * (d / R - 1.0);
            return r;
        }
    case LAW_MEMB_MEMB: {
        const double cr = R * P_MEMB, cr2 = cr * cr;
        if (s < cr2) {
            const double L = POW3(cr2) / POW3(s);
            r.coef = 4.0 * eps * L * (L - 1) / s;
            r.ds = 4.0 * eps * (4.0 * L - 7.0 * L * L) / (s * s);
            r.deps = 4.0 * L * (L - 1) / s;
        } else if (s < R * R) {
            const double d = sqrt(s);
            r.coef = -(DER_DER_CONST / d) * (d / cr - 1.0);
            r.ds = -DER_DER_CONST / (2 * d * d * d);
        } else {
            const double d = sqrt(s);
            const double lam = (1.0 + P_MEMB) * R;
            const double m = lam - d;
            const double L = POW6(cr) / POW6(m);
            const double dL = 6.0 * L / m;
            const double q = L * (L - 1.0) / (m * d);
            const double dq = (2 * L - 1) * dL / (m * d) + L * (L - 1.0) * (d - m) / ((m * d) * (m * d));
            r.coef = -(DER_DER_CONST / R) * ((1.0 - P_MEMB) / P_MEMB) - 4 * eps * q;
            r.ds = -4 * eps * dq / (2 * d);
            r.deps = -4 * q;
        }
        return r;
    }
    case LAW_PAIR: {
        const double d = sqrt(s);
        r.coef = -(Kspring_division * (d - t->spr_nat_len[c1])) / d;
        r.ds = -Kspring_division * t->spr_nat_len[c1] / (2 * d * d * d);
        return r;
    }
    default:
        return r;
    }
}

/* reverse of tri_bend_1(r, a, b, c) for the output adjoint o: the result is
 * f(n1.n2) (v x rho(n1, n2)) / |n1|, n1 = unit(u x v), n2 = unit(v x w),
 * u = r-b, v = a-b, w = c-b (minimum image); the divisor |n1| is 1 up to
 * rounding and is held constant */
static v3 v3_unit_bar(v3 N, v3 n_adj) {   /* adjoint of n = N/|N| */
    const double nn = v3_norm(N);
    const v3 n = v3_div(N, nn);
    return v3_div(v3_sub(n_adj, v3_scale(v3_dot(n_adj, n), n)), nn);
}
static inline double bend_factor_d(double c) {
    if (c >= 0) return -1.0;
    return -2.0 / ((1.0 + c) * (1.0 + c) * (1.0 + c));
}
static void bend_pair_bar(v3 n1, v3 n2, v3 v, double scale, v3 o_adj, v3* v_adj, v3* n1_adj, v3* n2_adj, double* d_adj) {
    /* y = scale * f(d) * (v x rho), rho = n2 - d n1 (d given separately) */
    const double d = v3_dot(n1, n2);
    const v3 rho = v3_sub(n2, v3_scale(d, n1));
    const v3 oy = v3_scale(scale * bend_factor(d), o_adj);
    *d_adj += scale * bend_factor_d(d) * v3_dot(o_adj, v3_cross(v, rho));
    *v_adj = v3_add(*v_adj, v3_cross(rho, oy));
    const v3 r_adj = v3_cross(oy, v);
    *n2_adj = v3_add(*n2_adj, r_adj);
    *d_adj += -v3_dot(r_adj, n1);
    *n1_adj = v3_add(*n1_adj, v3_scale(-d, r_adj));
}
static void tri_bend_1_bar(const eo_tissue* t, int r1, int a1, int b1, int c1, v3 o_adj, v3* bar) {
    /* bar[0..3] accumulate the adjoints of r1, a1, b1, c1 */
    v3 r1v = ctov(t, r1), a1v = ctov(t, a1), b1v = ctov(t, b1), c1v = ctov(t, c1);
    const v3 u = vpsub(t, r1v, b1v), v = vpsub(t, a1v, b1v), w = vpsub(t, c1v, b1v);
    v3 N1 = v3_cross(u, v), N2 = v3_cross(v, w);
    v3 t1 = N1, t2 = N2;
    const v3 n1 = v3_normalize(&t1), n2 = v3_normalize(&t2);
    const double div = v3_norm(t1);
    v3 v_adj = {0, 0, 0}, n1_adj = {0, 0, 0}, n2_adj = {0, 0, 0};
    double d_adj = 0;
    bend_pair_bar(n1, n2, v, 1.0 / div, o_adj, &v_adj, &n1_adj, &n2_adj, &d_adj);
    n1_adj = v3_add(n1_adj, v3_scale(d_adj, n2));
    n2_adj = v3_add(n2_adj, v3_scale(d_adj, n1));
    const v3 N1_adj = v3_unit_bar(N1, n1_adj), N2_adj = v3_unit_bar(N2, n2_adj);
    const v3 u_adj = v3_cross(v, N1_adj);
    v_adj = v3_add(v_adj, v3_add(v3_cross(N1_adj, u), v3_cross(w, N2_adj)));
    const v3 w_adj = v3_cross(N2_adj, v);
    bar[0] = v3_add(bar[0], u_adj);
    bar[1] = v3_add(bar[1], v_adj);
    bar[3] = v3_add(bar[3], w_adj);
    bar[2] = v3_sub(bar[2], v3_add(u_adj, v3_add(v_adj, w_adj)));
}
/* reverse of tri_bend_2(r, a, b, c): n1 = unit((r-c) x (b-c)),
 * n2 = unit((r-b) x (a-b)), result f(n1.n2) [(b-c) x rho(n1,n2)/|n1| +
 * (a-b) x rho(n2,n1)/|n2|] */
static void tri_bend_2_bar(const eo_tissue* t, int r1, int a1, int b1, int c1, v3 o_adj, v3* bar) {
    v3 r1v = ctov(t, r1), a1v = ctov(t, a1), b1v = ctov(t, b1), c1v = ctov(t, c1);
    const v3 p = vpsub(t, r1v, c1v), q = vpsub(t, b1v, c1v);       /* N1 = p x q */
    const v3 u = vpsub(t, r1v, b1v), v = vpsub(t, a1v, b1v);       /* N2 = u x v */
    v3 N1 = v3_cross(p, q), N2 = v3_cross(u, v);
    v3 t1 = N1, t2 = N2;
    const v3 n1 = v3_normalize(&t1), n2 = v3_normalize(&t2);
    const double div1 = v3_norm(t1), div2 = v3_norm(t2);
    /* f(d) is shared: d = n1.n2; the two terms are f(d) q x rho(n1,n2)/div1
     * and f(d) v x rho(n2,n1)/div2; bend_pair_bar handles f(d) (x) rho */
    v3 q_adj = {0, 0, 0}, v_adj = {0, 0, 0}, n1_adj = {0, 0, 0}, n2_adj = {0, 0, 0};
    double d_adj = 0;
    bend_pair_bar(n1, n2, q, 1.0 / div1, o_adj, &q_adj, &n1_adj, &n2_adj, &d_adj);
    bend_pair_bar(n2, n1, v, 1.0 / div2, o_adj, &v_adj, &n2_adj, &n1_adj, &d_adj);
    n1_adj = v3_add(n1_adj, v3_scale(d_adj, n2));
    n2_adj = v3_add(n2_adj, v3_scale(d_adj, n1));
    const v3 N1_adj = v3_unit_bar(N1, n1_adj), N2_adj = v3_unit_bar(N2, n2_adj);
    const v3 p_adj = v3_cross(q, N1_adj);
    q_adj = v3_add(q_adj, v3_cross(N1_adj, p));
    const v3 u_adj = v3_cross(v, N2_adj);
    v_adj = v3_add(v_adj, v3_cross(N2_adj, u));
    /* p = r-c, q = b-c, u = r-b, v = a-b */
    bar[0] = v3_add(bar[0], v3_add(p_adj, u_adj));
    bar[1] = v3_add(bar[1], v_adj);
    bar[2] = v3_add(bar[2], v3_sub(q_adj, v3_add(u_adj, v_adj)));
    bar[3] = v3_sub(bar[3], v3_add(p_adj, q_adj));
}

static int mech_window(eo_tissue* t, const eo_mech_params* p, int n_steps, double** keep) {
    use(t);
    const eo_mech_params saved = g_mp;
    g_mp = *p;
    g_window = 1;
    for (int s = 0; s < n_steps; s++) {
        if (keep) {
            memcpy(keep[s], t->x1, 8 * (size_t)t->n);
            memcpy(keep[s] + t->n, t->y1, 8 * (size_t)t->n);
            memcpy(keep[s] + 2 * (size_t)t->n, t->z1, 8 * (size_t)t->n);
        }
        eo_cell_interaction(t);
        eo_pos_copy(t);
    }
    g_mp = saved;
    g_window = 0;
    return t->err;
}

int eo_mech_forward(eo_tissue* t, const eo_mech_params* p, int n_steps, double* xyz_out) {
    const size_t n = (size_t)t->n;
    double* save = malloc(8 * 6 * n);
    memcpy(save, t->x1, 8 * n); memcpy(save + n, t->y1, 8 * n); memcpy(save + 2 * n, t->z1, 8 * n);
    memcpy(save + 3 * n, t->x2, 8 * n); memcpy(save + 4 * n, t->y2, 8 * n); memcpy(save + 5 * n, t->z2, 8 * n);
    int rc = mech_window(t, p, n_steps, NULL);
    if (xyz_out) { memcpy(xyz_out, t->x1, 8 * n); memcpy(xyz_out + n, t->y1, 8 * n); memcpy(xyz_out + 2 * n, t->z1, 8 * n); }
    memcpy(t->x1, save, 8 * n); memcpy(t->y1, save + n, 8 * n); memcpy(t->z1, save + 2 * n, 8 * n);
    memcpy(t->x2, save + 3 * n, 8 * n); memcpy(t->y2, save + 4 * n, 8 * n); memcpy(t->z2, save + 5 * n, 8 * n);
    free(save);
    return rc;
}

int eo_mech_adjoint(eo_tissue* t, const eo_mech_params* p, int n_steps, const double* target, double* loss,
                    eo_mech_grad* G) {
    const int n = t->n;
    double* save = malloc(8 * 6 * (size_t)n);
    memcpy(save, t->x1, 8 * (size_t)n); memcpy(save + n, t->y1, 8 * (size_t)n); memcpy(save + 2 * n, t->z1, 8 * (size_t)n);
    memcpy(save + 3 * n, t->x2, 8 * (size_t)n); memcpy(save + 4 * n, t->y2, 8 * (size_t)n); memcpy(save + 5 * n, t->z2, 8 * (size_t)n);
    double** keep = malloc(sizeof(double*) * (n_steps + 1));
    for (int s = 0; s <= n_steps; s++) keep[s] = malloc(8 * 3 * (size_t)n);
    /* the interactions of one step (the same every step) */
    g_tape_n = 0;
    g_tape_on = 1;
    use(t);
    const eo_mech_params saved = g_mp;
    g_mp = *p;
    g_window = 1;
    int rc = mech_window(t, p, n_steps, keep);
    g_tape_on = 0;
    memcpy(keep[n_steps], t->x1, 8 * (size_t)n); memcpy(keep[n_steps] + n, t->y1, 8 * (size_t)n);
    memcpy(keep[n_steps] + 2 * n, t->z1, 8 * (size_t)n);
    const int napp = g_tape_n / (n_steps > 0 ? n_steps : 1);
    /* loss and seed */
    double* lam = calloc(3 * (size_t)n, 8);
    double L = 0;
    for (int c = 0; c < n; c++) {
        const double dx = p_diff_x(t, keep[n_steps][c], target[c]);
        const double dy = p_diff_y(t, keep[n_steps][n + c], target[n + c]);
        const double dz = keep[n_steps][2 * n + c] - target[2 * n + c];
        L += dx * dx + dy * dy + dz * dz;
        lam[c] = 2 * dx / n; lam[n + c] = 2 * dy / n; lam[2 * n + c] = 2 * dz / n;
    }
    *loss = L / n;
    G->eps = G->k_adh = G->k_spring = G->k_bend = 0;
    double* nl = calloc(3 * (size_t)n, 8);
    v3* b_adj = calloc((size_t)n, sizeof(v3));
    for (int s = n_steps - 1; s >= 0; s--) {
        memcpy(t->x1, keep[s], 8 * (size_t)n); memcpy(t->y1, keep[s] + n, 8 * (size_t)n);
        memcpy(t->z1, keep[s] + 2 * (size_t)n, 8 * (size_t)n);
        memcpy(nl, lam, 8 * 3 * (size_t)n);   /* the identity part of x' = x + ... */
        /* pairs */
        for (int k = 0; k < napp; k++) {
            const mech_app a = g_tape[k];
            if (a.law == LAW_MU_MEMB) {   /* spring_force_to_memb as set in the step */
                int pr = t->pair[a.c1];
                t->spring_force_to_memb[a.c1] = (pr >= 0 && t->state[pr] == EO_FIX) ? Kspring
                                                : (t->rest_div_times[a.c1] > 0 ? Kspring_d : 0);
            }
            const law_d ld = coef_law_d(t, a.law, a.c1, a.c2);
            const double rx = p_diff_x(t, t->x1[a.c1], t->x1[a.c2]);
            const double ry = p_diff_y(t, t->y1[a.c1], t->y1[a.c2]);
            const double rz = t->z1[a.c1] - t->z1[a.c2];
            const double ax = lam[a.c1] - lam[a.c2], ay = lam[n + a.c1] - lam[n + a.c2],
                         az = lam[2 * n + a.c1] - lam[2 * n + a.c2];
            const double ar = ax * rx + ay * ry + az * rz;
            const double gx = DT_Cell * (ld.coef * ax + 2 * ld.ds * ar * rx);
            const double gy = DT_Cell * (ld.coef * ay + 2 * ld.ds * ar * ry);
            const double gz = DT_Cell * (ld.coef * az + 2 * ld.ds * ar * rz);
            nl[a.c1] += gx; nl[n + a.c1] += gy; nl[2 * n + a.c1] += gz;
            nl[a.c2] -= gx; nl[n + a.c2] -= gy; nl[2 * n + a.c2] -= gz;
            G->eps += DT_Cell * ld.deps * ar;
            G->k_adh += DT_Cell * ld.dadh * ar;
            G->k_spring += DT_Cell * ld.dspr * ar;
        }
        /* wall: z' = z + DT 2 eps L(L-1)/z, L = r^6/z^6, for z < r */
        for (int c = 0; c < n; c++) {
            const int st = t->state[c];
            if (!(c < t->nmemb || st == EO_DER)) continue;
            if (!collide_with_wall(t, c)) continue;
            const double z = t->z1[c];
            const double Lw = POW6(t->radius[c]) / POW6(z);
            nl[2 * n + c] += lam[2 * n + c] * DT_Cell * 2.0 * g_mp.eps * (7.0 * Lw - 13.0 * Lw * Lw) / (z * z);
            G->eps += lam[2 * n + c] * DT_Cell * 2.0 * Lw * (Lw - 1.0) / z;
        }
        /* bending of every membrane particle */
        for (int m = 0; m < n; m++) b_adj[m] = v3_make(0, 0, 0);
        for (int m = 0; m < t->nmemb; m++) {
            const v3 lm = v3_make(lam[m], lam[n + m], lam[2 * n + m]);
            const v3 dr = tri_memb_dr(t, m);
            G->k_bend += DT_Cell * v3_dot(lm, dr);
            const v3 ob = v3_scale(DT_Cell * g_mp.k_bend, lm);
            const int32_t* md = &t->memb_adj[6 * m];
            for (int i = 0; i < 6; i++) {
                const int nb = md[(i + 1) % 6];
                const int idx[4] = {m, md[i], nb, t->memb_adj[6 * nb + i]};
                v3 bar[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                tri_bend_1_bar(t, idx[0], idx[1], idx[2], idx[3], ob, bar);
                for (int q = 0; q < 4; q++) b_adj[idx[q]] = v3_add(b_adj[idx[q]], bar[q]);
            }
            for (int i = 0; i < 6; i++) {
                const int idx[4] = {m, md[i], md[(i + 1) % 6], md[(i + 2) % 6]};
                v3 bar[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                tri_bend_2_bar(t, idx[0], idx[1], idx[2], idx[3], ob, bar);
                for (int q = 0; q < 4; q++) b_adj[idx[q]] = v3_add(b_adj[idx[q]], bar[q]);
            }
        }
        for (int m = 0; m < n; m++) { nl[m] += b_adj[m].x; nl[n + m] += b_adj[m].y; nl[2 * n + m] += b_adj[m].z; }
        double* tmp = lam; lam = nl; nl = tmp;
    }
    g_mp = saved;
    g_window = 0;
    memcpy(t->x1, save, 8 * (size_t)n); memcpy(t->y1, save + n, 8 * (size_t)n); memcpy(t->z1, save + 2 * n, 8 * (size_t)n);
    memcpy(t->x2, save + 3 * n, 8 * (size_t)n); memcpy(t->y2, save + 4 * n, 8 * (size_t)n);
    memcpy(t->z2, save + 5 * n, 8 * (size_t)n);
    for (int s = 0; s <= n_steps; s++) free(keep[s]);
    free(keep); free(save); free(lam); free(nl); free(b_adj);
    return rc;
}
