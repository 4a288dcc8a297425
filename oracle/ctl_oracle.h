/*
 * ctl_oracle.h -- CPU oracle for the tau-leaping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and by nothing else.  The product (cutauleaping_amd/) never links it.
 *
 * It restates, in plain C, the per-simulation algorithm of cuTauLeaping
 * (2phase_n-tau-leaping.cu, kernel.cu) in two precisions:
 *   ORC_REPLAY  FP32 propensities / tau selection, FP64 clock: the exact op
 *               order of the gfx950 kernel, used for bit-exact trajectory
 *               parity with the GPU.
 *   ORC_REFSEM  the reference's own types (double a_j, a0, tau; float mu,
 *               sigma, g; float v=2 binomial term), used as the statistical
 *               comparator and as the CPU baseline.
 * Exact SSA is the force_ssa mode of either precision (100-step direct
 * method bursts, 2phase_n-tau-leaping.cu:2280-2454).
 *
 * Group algorithm (job.group = G): the gfx950 kernel for larger networks,
 * where G lanes cooperate on one simulation.  Reaction j and species s
 * belong to lane j mod G and s mod G.  Same per-reaction propensities and
 * per-species tau candidates as REPLAY; a0 and the selection sums are the
 * lane partial sums (ascending reactions) combined by a Hillis-Steele
 * inclusive scan over the G lanes (a selection picks among the lanes with a
 * positive partial); the small-lambda process of each lane uses the
 * lane's Philox stream (counter word 3 = 1 + lane, first gap drawn at the
 * start of the attempt) and every Poisson count (lambda >= 1) the
 * reaction's own stream (word 3 = 2^31 | j); the update and its negativity test run per species over the species-major
 * fear entries (the same rejections as the reaction-major order).
 *
 * Random numbers: Philox4x32-10 (Salmon et al., SC'11) keyed by the seed,
 * counter = {block, event, global simulation index, 0}; cuRAND's XORWOW
 * and curand_poisson (the reference's closed RNG) are replaced by the
 * samplers below, so parity with the reference is distributional.
 */
#ifndef CTL_ORACLE_H
#define CTL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_REPLAY = 0, ORC_REFSEM = 1 };

typedef struct {
    int S, R;
    const int32_t* left;    /* R*S row-major [r][s]   (left_side)  */
    const int32_t* right;   /* R*S row-major [r][s]   (right_side) */
    const int32_t* locked;  /* S flags, non-zero = feed-locked (M_feed); may be NULL */
} orc_net;

typedef struct {
    int n;                          /* simulations */
    const uint32_t* x0; int x0_rows;/* [row][s]; 1 row: shared; >= n rows: one per
                                       simulation; else n/rows simulations per row */
    const double* c; int c_rows;    /* [row][r]; same rule */
    const double* ts; int K;        /* sample times, t_max = ts[K-1] */
    const int32_t* cs; int SS;      /* sampled species */
    uint32_t seed_lo, seed_hi;      /* Philox key */
    uint32_t sim_base;              /* global index of simulation 0 (CLI offset) */
    int nc;                         /* criticality threshold (10) */
    float eps;                      /* tau-leaping error control (0.03) */
    int ssa_steps;                  /* SSA burst length (100) */
    float ssa_factor;               /* switch to SSA when tau' < ssa_factor / a0 (10) */
    int force_ssa;
    int precision;                  /* ORC_REPLAY | ORC_REFSEM */
    int threads;                    /* OpenMP threads, 0 = runtime default */
    int nfs;                        /* per-simulation feeds (read_feeds, T.cu:459-567): */
    const int32_t* fs;              /*   feed species; each is reset to the simulation's */
    const uint32_t* fv; int fv_rows;/*   value [row][nfs] after every leap and SSA step */
    int group;                      /* 0: one lane per simulation (REPLAY / REFSEM);
                                       G in {2,4,...,64}: the group algorithm with G lanes
                                       per simulation (FP32, REPLAY precision) */
} orc_job;

typedef struct {
    uint64_t firings;     /* sum of leap counts + SSA events with time <= t_max */
    uint64_t leaps;       /* accepted leaps */
    uint64_t rejections;  /* tau-halving retries */
    uint64_t switches;    /* leap attempts that fell back to SSA */
    uint64_t ssa_steps;   /* SSA steps executed (including overshoot) */
    uint64_t poisson;     /* individual Poisson draws made */
    uint64_t gaps;        /* exponential gaps of the small-lambda Poisson process (REPLAY) */
    double   t_end;
    uint32_t events;      /* Philox events consumed */
    uint32_t absorbed;
} orc_stat;

/* Compiled model (restatement of read_stoichiometry, 2phase_n-tau-leaping.cu:122-457). */
typedef struct {
    int S, R;
    int nL, nV, nF;
    int *Lptr, *Ls, *Lv;        /* left flat, reaction-major (T.cu:219-232) */
    int *Vptr, *Vr, *Vv;        /* rav: var flat, species-major (T.cu:308-320) */
    int *Fptr, *Fs, *Fv;        /* fear flat, reaction-major (T.cu:323-335) */
    int *FSptr, *FSr, *FSv;     /* the same entries species-major, ascending reaction */
    unsigned char *hor, *hort;  /* HOR / HORtype (T.cu:402-432) */
} orc_model;

int  orc_compile(const orc_net* net, orc_model* m);
void orc_model_free(orc_model* m);

/* Runs every simulation of the job.  storage is [sim][i][sk] (uint32);
 * stats (n entries) and xfinal ([sim][s] final state) may be NULL. */
int orc_run(const orc_net* net, const orc_job* job, uint32_t* storage, orc_stat* stats,
            int32_t* xfinal);

/* Sampler entry points for the tests. */
void  orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_logf(float x);
float orc_expf(float x);
float orc_log1m(float v);
/* count draws from one event stream (key, sim, ev), consumed sequentially. */
void orc_poisson_stream(uint32_t seed_lo, uint32_t seed_hi, uint32_t sim, uint32_t ev,
                        float lam, int count, int32_t* out);
void orc_uniform_stream(uint32_t seed_lo, uint32_t seed_hi, uint32_t sim, uint32_t ev,
                        int open_closed, int count, float* out);

#ifdef __cplusplus
}
#endif
#endif
