/*
 * ctl.h -- C ABI of the MI355X-native tau-leaping library (libctl.so).
 *
 * cuTauLeaping has no library or FFI interface: its boundary is the process
 * (kernel.cu:74-436, argv of kernel.cu:78-119) plus the host functions and
 * kernels that main() calls.  Each entry point below replaces one of those
 * call sites; the citation names the reference code it stands for.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative status; the codes
 *    mirror the reference's exit() codes (CTL_E_*).  ctl_last_error()
 *    returns a thread-local message for the last failure.  No C++
 *    exception crosses this ABI.
 *  - The library owns ctl_model / ctl_sim handles; the caller owns every
 *    output buffer it passes in.
 *  - A handle is not thread-safe.  Different handles may run concurrently;
 *    each ctl_sim owns one HIP stream on its device.
 *  - threads / blocks are semantic (simulation count = threads * blocks,
 *    input-row selection, fitness grouping), not the launch geometry; the
 *    library chooses the geometry itself.
 */
#ifndef CTL_H
#define CTL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (kernel.cu / 2phase_n-tau-leaping.cu exit codes). */
#define CTL_OK            0
#define CTL_E_ARGS       (-1)   /* bad arguments or malformed input files (T.cu:58-71, 211-214, 244-246) */
#define CTL_E_DEVICE     (-2)   /* HIP runtime failure or no usable GPU */
#define CTL_E_NOMEM      (-3)   /* device allocation failed (K.cu:231-234) */
#define CTL_E_ROWS       (-4)   /* MX_0 has fewer rows than offset + simulations (T.cu:596-599) */
#define CTL_E_WIDTH      (-5)   /* MX_0 row width differs from the species count (T.cu:628-631) */
#define CTL_E_NOFILE     (-6)   /* MX_0 missing (T.cu:643-646) */
#define CTL_E_TTS        (-10)  /* tts_vector does not cover threads*blocks (T.cu:1136-1140) */
#define CTL_E_STATE      (-11)  /* call out of order (e.g. samples before run) */
#define CTL_E_LIMIT      (-12)  /* model exceeds a documented size limit */

typedef struct ctl_model ctl_model;
typedef struct ctl_sim ctl_sim;

/* A model given as arrays (used by generators and tests; the folder loader
 * builds the same thing). */
typedef struct {
    int32_t S, R;
    const int32_t* left;       /* R*S row-major [r][s], >= 0 (left_side) */
    const int32_t* right;      /* R*S row-major [r][s], >= 0 (right_side) */
    const int32_t* feed;       /* S flags, non-zero = feed-locked (M_feed); NULL = none */
    const uint32_t* x0;        /* x0_rows * S initial counts (MX_0 rows) */
    int64_t x0_rows;           /* 1: shared by every simulation; >= n: one row per
                                  simulation (the reference); otherwise n/x0_rows
                                  consecutive simulations share a row (replicas) */
    const double* c;           /* c_rows * R rate constants (c_matrix rows) */
    int64_t c_rows;            /* same rule as x0_rows */
    const double* t_vector;    /* K sample times, t_max = t_vector[K-1] */
    int32_t K;
    const int32_t* cs_vector;  /* SS sampled species; NULL = all species */
    int32_t SS;
} ctl_model_desc;

typedef struct {
    int32_t  threads;          /* argv[2]; simulations = threads * blocks */
    int32_t  blocks;           /* argv[3] */
    int32_t  gpu;              /* argv[4]: HIP device ordinal */
    int64_t  offset;           /* argv[5]: first input row; also the Philox stream base */
    int32_t  force_ssa;        /* argv[9] != "0" */
    int32_t  nc;               /* critical threshold n_c (T.cuh:75, default 10) */
    float    eps;              /* error control epsilon (T.cuh:76, default 0.03) */
    int32_t  ssa_steps;        /* SSA burst length (K.cu:83, default 100) */
    float    ssa_factor;       /* switch to SSA when tau' < ssa_factor / a0 (T.cu:1935, default 10) */
    uint64_t seed;             /* Philox key (the reference used time(NULL), K.cu:176) */
    int32_t  verbose;
    /* launch geometry (0 = chosen by the library) */
    int32_t  block_size;       /* threads per workgroup (64..256, multiple of 64; 0: library default;
                                  < 0 with one lane per simulation: default size, state kept in HBM
                                  even when it would fit in LDS) */
    int32_t  slots;            /* resident simulation slots */
    int32_t  iters_per_launch; /* leap/SSA iterations per slot per launch */
    int32_t  ssa_chunk;        /* SSA steps per iteration of a slot (0: 10 for one lane per
                                  simulation, 8 for a regrouping group kernel) */
    int32_t  lanes_per_sim;    /* 1: one lane runs one simulation; 2..64 (power of two): that many
                                  lanes of a wavefront cooperate on one simulation whose state
                                  stays in LDS (the group algorithm); 0: chosen by the library */
    int32_t  compact_interval; /* every this many iterations a workgroup of several wavefronts
                                  regroups its simulations so that leap-mode and SSA-mode ones run
                                  in different wavefronts, SSA bursts then running ssa_chunk steps
                                  per iteration; trajectories do not depend on it.  0: library
                                  default (8 for 16 or 32 lanes per simulation outside
                                  force-SSA mode, else never), < 0: never */
} ctl_opts;

typedef struct {
    uint64_t firings;          /* reactions fired with event time <= t_max */
    uint64_t leaps;            /* accepted tau-leaps */
    uint64_t rejections;       /* tau-halving retries */
    uint64_t switches;         /* leap attempts that fell back to SSA */
    uint64_t ssa_steps;        /* SSA steps executed (incl. the burst overshoot) */
    uint64_t poisson;          /* individual Poisson variates drawn (lambda >= 1) */
    uint64_t gaps;             /* exponential gaps of the small-lambda Poisson process */
    uint64_t simulations;
    uint64_t absorbed;         /* simulations that reached a0 = 0 */
    double   kernel_ms;        /* sum of simulation-kernel durations (HIP events) */
    double   wall_ms;          /* wall time of ctl_sim_run */
    int32_t  launches;         /* simulation-kernel launches */
    int32_t  slots;            /* resident slots used */
    int32_t  block_size;
    int32_t  lds_state;        /* 1 when per-lane state lives in LDS */
    int32_t  lanes_per_sim;    /* lanes per simulation used */
} ctl_stats;

typedef struct {
    uint64_t firings;          /* per simulation, as in ctl_stats */
    uint64_t ssa_steps;
    uint64_t poisson;
    uint64_t gaps;
    double   t_end;            /* final clock (t_max + 0.1 after absorption, T.cu:1693) */
    uint32_t leaps, rejections, switches;
    uint32_t events;           /* Philox events consumed */
    uint32_t absorbed;         /* 1 when a0 reached 0 */
    uint32_t reserved;
} ctl_sim_stat;

typedef struct {
    int32_t S, R, K, SS;
    int32_t nL, nV, nF;        /* non-zeros of the left, var and fear matrices */
    int64_t x0_rows, c_rows;
    int64_t row_offset;        /* input row of x0 row 0 (folder loads) */
    double  t_max;
} ctl_model_info;

/* ---- options ---------------------------------------------------------- */
/* Reference defaults: nc = 10, eps = 0.03 (T.cuh:75-76), 100 SSA steps
 * (K.cu:83), SSA when tau' < 10 / a0 (T.cu:1935). */
void ctl_opts_default(ctl_opts* o);

/* ---- model ------------------------------------------------------------- */
/* Parses an input folder: first_pass + read_stoichiometry (T.cu:90-457),
 * read_quantities_offset (T.cu:570-649), read_parameters_offset
 * (T.cu:717-782), read_instants (T.cu:932-1053).  Rows [offset, offset+n)
 * of MX_0 and c_matrix are loaded (n <= 0: every row from offset on). */
int ctl_model_load_folder(const char* dir, int64_t offset, int64_t n, int verbose, ctl_model** out);

/* Same model from arrays (copies everything). */
int ctl_model_from_arrays(const ctl_model_desc* d, ctl_model** out);

int ctl_model_info_get(const ctl_model* m, ctl_model_info* out);

/* Compiled structure (read_stoichiometry T.cu:219-335, 402-432): HOR and
 * HORtype per species, and the left / var (species-major) / fear flats as
 * (reaction, species, coefficient) triples.  Any pointer may be NULL. */
int ctl_model_structure(const ctl_model* m, uint8_t* hor, uint8_t* hortype,
                        int32_t* left_flat, int32_t* rav_flat, int32_t* fear_flat);

/* Fitness inputs ts_rep, p_rep, ts_numtgt, tts_vector, ts_matrix
 * (read_time_series, T.cu:1056-1261). */
int ctl_model_load_fitness(ctl_model* m, const char* dir, int32_t threads, int32_t blocks, int verbose);

/* Per-simulation feeds, read_feeds (T.cu:459-567): the species with a
 * non-zero value in the last line of MX_feed are reset, after every accepted
 * leap (T.cu:2072-2079), to the value in the simulation's row (rows
 * [offset, offset+n); n <= 0: all from offset).  SSA steps apply no feed, as
 * in the reference's USE_FEAR_SSA build (T.cu:38, 2240-2244, 2420-2424).  The
 * reference compiles feeds out by default (DISABLE_FEED, T.cu:34): a model
 * has none unless one of these is called.  A missing MX_feed prints the
 * reference's warning and leaves the model without feeds (T.cu:473-477).
 * Rows follow the x0_rows rule. */
int ctl_model_load_feeds(ctl_model* m, const char* dir, int64_t offset, int64_t n, int verbose);
int ctl_model_set_feeds(ctl_model* m, const int32_t* species, int32_t nfs, const uint32_t* values, int64_t rows);

void ctl_model_destroy(ctl_model* m);

/* ---- simulation -------------------------------------------------------- */
/* Device allocation and upload: allocate_space (T.cu:2899-2972),
 * set_symbols (T.cu:865-930), setup_random + PrepareVariables
 * (T.cu:2974-2988, 1269-1297). */
int ctl_sim_create(const ctl_model* m, const ctl_opts* o, ctl_sim** out);

/* Runs every simulation to t_max: the host loop of kernel.cu:242-345 with
 * PreCalculations / EventualSSASteps / LightweightSSA / ParallelVoting
 * (T.cu:1568-2454, 2990-3002).  Blocking.  May be called again: each call
 * restarts from the initial state with the same streams. */
int ctl_sim_run(ctl_sim* s);

/* Sample storage (dev_perthread_storage, T.cu:2966).  layout 0: the
 * reference's [sample][sampled species][simulation]; layout 1:
 * [simulation][sample][sampled species].  count must be n*K*SS. */
int ctl_sim_samples(const ctl_sim* s, uint32_t* out, size_t count, int layout);

/* Device pointer of the [simulation][sample][sampled species] storage (for
 * callers that keep the samples on the GPU). */
int ctl_sim_samples_device(const ctl_sim* s, void** dev_ptr, size_t* count);

int ctl_sim_stats(const ctl_sim* s, ctl_stats* out);
int ctl_sim_sim_stats(const ctl_sim* s, ctl_sim_stat* out, size_t n);

/* Final states and clocks (dev_perthread_X / dev_perthread_T). */
int ctl_sim_final_state(const ctl_sim* s, int32_t* x /* n*S or NULL */, double* t /* n or NULL */);

/* WriteDynamics2 / PrintDynamics2 (T.cu:3431-3507): one text file
 * <path_prefix>_<gid> per simulation, or everything on stdout. */
int ctl_write_dynamics(const ctl_sim* s, const char* path_prefix, int to_stdout);

/* calculateFitness (T.cu:3324-3408): one double per (group, thread),
 * n / p_rep values.  Needs ctl_model_load_fitness. */
int ctl_fitness(const ctl_sim* s, double* out, size_t count);

/* Scalable output (no counterpart in the reference, whose only output is
 * one text file per simulation, T.cu:3431-3469).  Per-group moments of the
 * samples, reduced on the device: simulations [g*group, (g+1)*group) form
 * group g; mean and population standard deviation per (group, sample,
 * sampled species), [g][i][sk] doubles, count = (n/group)*K*SS.  n must be a
 * multiple of group.  Either output pointer may be NULL. */
int ctl_sim_moments(const ctl_sim* s, int32_t group, double* mean, double* sd, size_t count);

/* Allocation-free variant of ctl_sim_moments, same arguments and the same
 * results bit for bit.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc /
 * hipFree: the partial sums and the staging of the two reduction passes live
 * in a 32 MiB static device array, one per GPU, set up by the first
 * ctl_sim_create on that GPU.  Fails with CTL_E_LIMIT when the moments need
 * more (about 16 * (n/group) * K * SS * (ceil(group/256), at most 64, + 1)
 * bytes).  Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (from several threads) race and must be serialised by the caller. */
int ctl_sim_moments_noalloc(const ctl_sim* s, int32_t group, double* mean, double* sd, size_t count);

/* Binary dump of the samples: a 32-byte header (magic "CTLS", version 1,
 * n, K, SS as uint32, 12 reserved bytes) then n*K*SS uint32 in
 * [simulation][sample][sampled species] order. */
int ctl_write_samples_binary(const ctl_sim* s, const char* path);

void ctl_sim_destroy(ctl_sim* s);

/* ---- batched models ------------------------------------------------------ */
/* Many independent models of uneven size and replica count in one run of
 * persistent kernels: the models that share a number of lanes per simulation
 * form a class with its own kernel and stream, and the classes run
 * concurrently.  No counterpart in the reference, which runs one model per
 * process (a batch replaces N invocations of kernel.cu's main with the same
 * argv except the input folder).  Model i runs replicas[i] simulations; its
 * Philox stream indices are offset + base_i + replica, base_i being the sum
 * of the replica counts of the models before it (ctl_batch_stream_base).
 * threads / blocks of the options are ignored; every other field applies to
 * all models (block_size to the one-lane class only).  The models must
 * outlive nothing: they are copied.  ctl_stats.kernel_ms of a batch is the
 * elapsed time of its launch sequence (the classes overlap). */
typedef struct ctl_batch ctl_batch;
int ctl_batch_create(const ctl_model* const* models, const int64_t* replicas, int32_t nmodels,
                     const ctl_opts* o, ctl_batch** out);
/* ctl_batch_create with lanes per simulation chosen per model: lanes[i] > 0
 * runs model i with that many lanes per simulation (its kernel class),
 * lanes[i] == 0 leaves it to o->lanes_per_sim / the library's size rule.
 * No counterpart in the reference (no batching there); autotune.tune_batch
 * produces the array. */
int ctl_batch_create_lanes(const ctl_model* const* models, const int64_t* replicas, const int32_t* lanes,
                           int32_t nmodels, const ctl_opts* o, ctl_batch** out);
int ctl_batch_run(ctl_batch* b);
/* Samples of one model, [replica][sample][sampled species]. */
int ctl_batch_samples(const ctl_batch* b, int32_t model, uint32_t* out, size_t count);
int ctl_batch_sim_stats(const ctl_batch* b, int32_t model, ctl_sim_stat* out, size_t count);
int ctl_batch_stats(const ctl_batch* b, ctl_stats* out);
int ctl_batch_stream_base(const ctl_batch* b, int32_t model, int64_t* base);
/* Lanes per simulation model `model` runs with: lanes_per_sim of the options
 * when it is non-zero, otherwise the library's choice for that model (one
 * lane for small working sets, else about 16 reactions per lane, widened
 * until a wavefront's share of LDS fits).  Models of one value form a kernel
 * class; ctl_stats.lanes_per_sim of a batch is that value when every model
 * shares it and 0 when the classes differ. */
int ctl_batch_model_lanes(const ctl_batch* b, int32_t model, int32_t* lanes);
void ctl_batch_destroy(ctl_batch* b);

/* ---- surrogate MLP ------------------------------------------------------- */
/* A multilayer perceptron widths[0] -> ... -> widths[nlayers] with ReLU on the
 * hidden layers and a linear output, trained and evaluated by hand-written
 * BF16-MFMA kernels with FP32 accumulation.  No counterpart in the reference:
 * it is the learned surrogate that screens parameter sweeps (log rate
 * constants -> per-species mean and standard deviation trajectories); the
 * simulator above generates its data and judges its error.
 * Parameters are one flat FP32 vector: W_1 [w0][w1] row-major, b_1 [w1],
 * W_2, b_2, ...  Weights start He-uniform from seed, biases at zero. */
typedef struct ctl_mlp ctl_mlp;
int ctl_mlp_create(const int32_t* widths, int32_t nlayers, int32_t gpu, uint64_t seed, ctl_mlp** out);
int ctl_mlp_num_params(const ctl_mlp* n, int64_t* count);
/* set != 0: upload count FP32 parameters; set == 0: download them. */
int ctl_mlp_params(ctl_mlp* n, float* host, size_t count, int32_t set);
/* y [rows][wL] = f(x [rows][w0]); host buffers. */
int ctl_mlp_forward(ctl_mlp* n, const float* x, int64_t rows, float* y);
/* Same on device buffers, enqueued on the network's stream and waited for;
 * ms (optional) receives the device time of the forward kernels. */
int ctl_mlp_forward_device(ctl_mlp* n, const float* x_dev, int64_t rows, float* y_dev, float* ms);
/* Loss (mean squared error over rows*wL) and its FP32 gradient with respect
 * to every parameter (grad may be NULL) for one batch of rows <= 65536. */
int ctl_mlp_gradient(ctl_mlp* n, const float* x, const float* t, int32_t rows, float* grad, size_t count,
                     float* loss);
/* Training set kept in device memory: x [rows][w0], t [rows][wL]. */
int ctl_mlp_set_data(ctl_mlp* n, const float* x, const float* t, int64_t rows);
/* Mini-batch Adam (beta1 0.9, beta2 0.999, eps 1e-8) over shuffled batches;
 * losses[e] (optional) receives the mean batch loss of epoch e. */
int ctl_mlp_train(ctl_mlp* n, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* losses);
void ctl_mlp_destroy(ctl_mlp* n);

/* ---- misc -------------------------------------------------------------- */
const char* ctl_last_error(void);
const char* ctl_version(void);
/* Number of HIP devices visible (0 when none); never fails loudly. */
int ctl_device_count(void);
/* Profiling build only (libctl_phases.so, compiled with CTL_PHASE_CLOCK=1):
 * out[0 .. CTL_PHASES) receives the shader-clock ticks the group kernel's
 * wavefronts spent in each phase since the last reset, in the order
 * propensities, tau', switch / tau'' / pick, counts and gaps, Poisson queue,
 * update, sampling, SSA burst opening, SSA select, SSA update, SSA
 * dependents, before a leap, before SSA steps, loop (fetch / finish),
 * regrouping, parking.  reset != 0 zeroes the totals after reading.  The
 * normal library returns CTL_E_STATE.  No counterpart in the reference. */
#define CTL_PHASES 16
int ctl_phase_clock(int32_t gpu, uint64_t* out, int32_t count, int32_t reset);

#ifdef __cplusplus
}
#endif
#endif /* CTL_H */
