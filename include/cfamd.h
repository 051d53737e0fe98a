/*
 * cfamd.h - C ABI of the MI355X Monte Carlo pricing and AAD-risk engine.
 *
 * Plain C: POD descriptors, pointers and sizes, int status codes.  Every
 * function returns CF_OK (0) on success or a negative CF_ERR_* code, and
 * cf_last_error() then holds the message (thread local).  The C++ wrappers in
 * include/compfinance_amd/ rethrow it as std::runtime_error, as the reference
 * does (mcBase.h:273, :320; main.h:90-93, :125-128, :202-205, :377-383).
 *
 * Two levels:
 *   1. descriptor level - replaces the simulation templates
 *        mcParallelSimul     mcBase.h:314-400
 *        mcParallelSimulAAD  mcBase.h:564-753
 *      for a model and a product flattened into cf_model / cf_product;
 *   2. store level - replaces the id-keyed registry and entry points the
 *      reference exposes to its own front end (store.h:38-205, main.h:43-411,
 *      exported to Excel as xPutDupire, xPutBlackScholes, xPutEuropean,
 *      xPutBarrier, xPutEuropeans, xValue, xAADrisk, xAADriskAggregate,
 *      xlExport.cpp:84-386, :546-772).
 * All functions may be called from several host threads.
 */
#ifndef CFAMD_H
#define CFAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define CF_OK 0
#define CF_ERR_ARG (-1)          /* invalid argument, incompatible model/product */
#define CF_ERR_NOT_FOUND (-2)    /* unknown store id or payoff label */
#define CF_ERR_UNSUPPORTED (-3)  /* model/product/RNG outside the GPU engine */
#define CF_ERR_DEVICE (-4)       /* HIP runtime error */
#define CF_ERR_CAPACITY (-5)     /* request does not fit the device limits */

/* ------------------------------------------------------------ descriptors */
enum cf_model_kind {
    CF_MODEL_BLACK_SCHOLES = 0,  /* BlackScholes<T>  mcMdlBS.h:21-351     */
    CF_MODEL_DUPIRE = 1          /* Dupire<T>        mcMdlDupire.h:28-281 */
};

/* Model parameters as the reference constructors take them
 * (BlackScholes(spot, vol, spotMeasure, rate, div) mcMdlBS.h:72-93;
 *  Dupire(spot, spots, times, vols, maxDt) mcMdlDupire.h:76-111). */
typedef struct cf_model {
    int kind;
    double spot;
    /* Black-Scholes */
    double vol;
    double rate;
    double div;
    int spot_measure;        /* 0 = risk-neutral, 1 = spot measure */
    /* Dupire local-vol grid, vols[i * n_times + j] = sigma(spots[i], times[j]) */
    int n_spots;
    int n_times;
    const double* spots;
    const double* times;
    const double* vols;
    double max_dt;
} cf_model;

enum cf_product_kind {
    CF_PRODUCT_EUROPEAN = 0,   /* European<T>  mcPrd.h:29-126  */
    CF_PRODUCT_BARRIER = 1,    /* UOC<T>       mcPrd.h:128-289 */
    CF_PRODUCT_EUROPEANS = 2,  /* Europeans<T> mcPrd.h:291-400 */
    CF_PRODUCT_ASIAN = 3,      /* arithmetic-average option; no reference class,
                                  written to the Product<T> interface mcBase.h:122-152 */
    CF_PRODUCT_BOOK = 4        /* several products on the same paths, payoffs
                                  concatenated; no reference class (Product<T>) */
};

/* Product parameters as the reference store functions take them
 * (putEuropean store.h:128-142, putBarrier :144-163, putEuropeans :184-205). */
typedef struct cf_product {
    int kind;
    double strike;           /* European, barrier */
    double exercise;         /* European exercise date */
    double settlement;       /* European settlement date (<= 0: exercise) */
    double barrier;          /* up-and-out barrier */
    double maturity;         /* barrier maturity */
    double monitor_freq;     /* barrier monitoring period */
    double smooth;           /* barrier smoothing, fraction of spot (<= 0: 1e-8) */
    int call_put;            /* barrier: 0 = call, 1 = put */
    int n_options;           /* Europeans: (maturity, strike) pairs */
    const double* maturities;
    const double* strikes;
    /* Asian: strike, maturity, call_put and fixings at fixing_freq, 2 fixing_freq, ...
     * up to the maturity (accumulated as the barrier's monitoring dates, the last
     * one at maturity), each moved to the nearest date of the grid 0, snap_freq,
     * 2 snap_freq, ..., maturity when snap_freq > 0 (ties to the earlier date). */
    double fixing_freq;
    double snap_freq;
    /* Book: the parts (European with settlement = exercise, barrier, Europeans,
     * Asian); at most 16 barriers and 4 distinct fixing schedules. */
    int n_parts;
    const struct cf_product* parts;
} cf_product;

enum cf_precision { CF_FP64 = 0, CF_FP32 = 1 };
enum cf_aad_mode { CF_AAD_CHECKPOINT = 0, CF_AAD_RECOMPUTE = 1 };

/* NumericalParam (main.h:33-40) plus the engine's own options. */
typedef struct cf_num {
    int parallel;            /* accepted for compatibility; the GPU is always parallel */
    int use_sobol;           /* 1: Sobol; 0: mrg32k3a with seed1, seed2 (main.h:51-52) */
    int num_path;            /* 1 .. 2^31 - 1 (NumericalParam::numPath is an int,
                              * main.h:36); a run of more paths is sharded by
                              * first_path into calls of at most that many, and
                              * first_path + num_path stays below 2^32 - 1 (the
                              * Sobol index range) */
    int seed1;               /* mrg32k3a seeds, converted to unsigned as the reference does */
    int seed2;
    int precision;           /* CF_FP64 or CF_FP32 */
    int aad_mode;            /* CF_AAD_CHECKPOINT or CF_AAD_RECOMPUTE */
    int first_path;          /* first Sobol path index; 0 in the reference */
} cf_num;

/* ------------------------------------------------------------ descriptor level */

/* Number of payoffs (Product::payoffLabels().size(), mcBase.h:139) and of model
 * parameters (Model::numParams(), mcBase.h:197-200); negative on error. */
int cf_num_payoffs(const cf_product* prd);
int cf_num_params(const cf_model* mdl);

/* Labels, formatted as the reference does (mcPrd.h:66-80, :197-207, :325-335;
 * mcMdlBS.h:87-90; mcMdlDupire.h:95-107).  Writes a NUL-terminated string. */
int cf_payoff_label(const cf_product* prd, int k, char* buf, size_t cap);
int cf_param_label(const cf_model* mdl, int k, char* buf, size_t cap);

/* Product timeline (Product::timeline(), mcBase.h:131) and the model's
 * simulation timeline with product-date flags (Dupire::allocate
 * mcMdlDupire.h:165-193, BlackScholes::allocate mcMdlBS.h:149-195).
 * Return the length; at most cap entries are written. */
int cf_product_timeline(const cf_product* prd, double* times, int cap);
int cf_sim_timeline(const cf_model* mdl, const cf_product* prd, double* times, int* common, int cap);

/* mcParallelSimul (mcBase.h:314-400).  Simulates num->num_path paths with the
 * Sobol sequence, path p using point p+1 (sobol.h:119-151).
 *   means[n_pay]            average payoffs (main.h:69-74)
 *   std_errs[n_pay]         Monte Carlo standard errors (nullable)
 *   per_path[n_path*n_pay]  the reference's result matrix, row = path (nullable) */
int cf_mc_simul(const cf_model* mdl, const cf_product* prd, const cf_num* num,
                double* means, double* std_errs, double* per_path);

/* mcParallelSimulAAD (mcBase.h:564-753) with a linear aggregator
 * agg = sum_k weights[k] * payoff_k, which covers the reference's two
 * aggregators (one-hot, main.h:135; notionals, main.h:210-213).
 *   payoff_means[n_pay], agg_mean[1]
 *   risks[n_param]          d E[agg] / d parameter, parameter order of the model
 *   per_path_payoffs[n_path*n_pay], per_path_agg[n_path]   (nullable) */
int cf_mc_simul_aad(const cf_model* mdl, const cf_product* prd, const cf_num* num,
                    const double* weights, double* payoff_means, double* agg_mean,
                    double* risks, double* per_path_payoffs, double* per_path_agg);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cf_mc_simul_aad with the same arguments and results,
 * with no device or pinned allocation and no freeing.  Every device buffer
 * of the call comes from a module array of CF_NOALLOC_SCRATCH_MIB MiB per
 * device, held until the process ends; a call that needs more returns
 * CF_ERR_CAPACITY (use fewer paths, the recompute adjoint, or
 * cf_mc_simul_aad).  The array keeps the last call's contents; the engine's
 * lock runs the calls of one process on one device one after the other
 * (calls that overlapped on one device would race on the array). */
#define CF_NOALLOC_SCRATCH_MIB 256
int cf_mc_simul_aad_noalloc(const cf_model* mdl, const cf_product* prd, const cf_num* num,
                            const double* weights, double* payoff_means, double* agg_mean,
                            double* risks, double* per_path_payoffs, double* per_path_agg);

/* mcParallelSimulAADMulti (mcBase.h:858-989): the risk of every payoff
 * separately, risks[n_param][n_pay] = d E[payoff_k] / d parameter_i, on the
 * same paths; one forward sweep (checkpoint adjoint, when its kept states fit
 * the free device memory) and one reverse sweep per payoff, else one adjoint
 * call per payoff.  Returns n_param. */
int cf_mc_simul_aad_multi(const cf_model* mdl, const cf_product* prd, const cf_num* num,
                          double* payoff_means, double* risks);

/* Device memory one AAD call on this product would take (checkpoint buffers
 * dominate); used for admission.  Bytes, or negative on error. */
int64_t cf_aad_footprint(const cf_model* mdl, const cf_product* prd, const cf_num* num);

/* ------------------------------------------------------------ store level */

/* store.h:38-54 */
int cf_put_black_scholes(double spot, double vol, int q_spot, double rate, double div, const char* id);
/* store.h:56-73; vols spot major, n_spots x n_times */
int cf_put_dupire(double spot, const double* spots, int n_spots, const double* times, int n_times,
                  const double* vols, double max_dt, const char* id);
/* store.h:128-142 */
int cf_put_european(double strike, double exercise_date, double settlement_date, const char* id);
/* store.h:144-163 */
int cf_put_barrier(double strike, double barrier, double maturity, double monitor_freq,
                   double smooth, int call_put, const char* id);
/* store.h:184-205 */
int cf_put_europeans(const double* maturities, const double* strikes, int n, const char* id);

/* Arithmetic-average option (CF_PRODUCT_ASIAN; no reference counterpart). */
int cf_put_asian(double strike, double maturity, double fixing_freq, double snap_freq, int call_put, const char* id);
/* Book of stored products on shared paths (CF_PRODUCT_BOOK; no reference
 * counterpart); part_ids name products already in the store. */
int cf_put_book(const char* const* part_ids, int n, const char* id);

/* main.h:80-96: payoff values in label order; returns n_pay, writes at most cap. */
int cf_value(const char* model_id, const char* product_id, const cf_num* num, double* values, int cap);

/* main.h:99-173.  risk_payoff: label, or NULL / "" for the first payoff.
 * Returns n_param; writes payoff_values[min(n_pay, cap_pay)], *risk_value,
 * risks[min(n_param, cap_risk)]. */
int cf_aad_risk_one(const char* model_id, const char* product_id, const cf_num* num,
                    const char* risk_payoff, double* payoff_values, int cap_pay,
                    double* risk_value, double* risks, int cap_risk);

/* main.h:176-254: notionals keyed by payoff label. */
int cf_aad_risk_aggregate(const char* model_id, const char* product_id, const char* const* labels,
                          const double* notionals, int n, const cf_num* num,
                          double* payoff_values, int cap_pay, double* risk_value,
                          double* risks, int cap_risk);

/* main.h:269-312 (AADriskMulti): payoff values and risks[n_param][n_pay],
 * writes at most cap_pay values and cap_risk risks.  Returns n_param. */
int cf_aad_risk_multi(const char* model_id, const char* product_id, const cf_num* num,
                      double* payoff_values, int cap_pay, double* risks, int cap_risk);

/* main.h:364-411: value, delta and the n_spots x n_times vega matrix. */
int cf_dupire_aad_risk(const char* model_id, const char* product_id, const char* const* labels,
                       const double* notionals, int n, const cf_num* num,
                       double* value, double* delta, double* vega, int cap_vega);

/* Remove every stored model and product. */
void cf_store_clear(void);

/* ------------------------------------------------------------ local-vol calibration (FP64, host) */

/* Merton jump-diffusion implied-vol surface (MertonIVS, ivs.h:150-179;
 * merton, blackScholesIvol analytics.h:23-91). */
typedef struct cf_merton_ivs {
    double spot;
    double vol;
    double jmp_intens;
    double jmp_average;
    double jmp_std;
} cf_merton_ivs;

/* dupireCalib (main.h:414-438 -> mcMdlDupire.h:337-388): spots = fillData(incl
 * spots, max_ds, 0.01), times = fillData(incl times, max_dt, ONE_HOUR, {max_dt}),
 * local vols by Dupire's formula on the surface's call prices (ivs.h:119-138),
 * flat outside 2.5 ATM standard deviations.  Writes *n_spots, *n_times and,
 * when n_spots x n_times <= cap, spots, times and lvols (spot major). */
int cf_dupire_calib(const cf_merton_ivs* ivs, const double* incl_spots, int n_incl_spots, double max_ds,
                    const double* incl_times, int n_incl_times, double max_dt, double* spots, double* times,
                    double* lvols, int* n_spots, int* n_times, int cap);

/* dupireCalib with a risk view (mcMdlDupire.h:337-388, ivs.h:26-138): the
 * local vols of the Merton surface plus implied-vol spreads on strikes x mats
 * (spreads[n_strikes * n_mats], strike major; interp2D<true> between them,
 * interp.h:66-110).  Outputs as cf_dupire_calib. */
int cf_dupire_calib_view(const cf_merton_ivs* ivs, const double* incl_spots, int n_incl_spots, double max_ds,
                         const double* incl_times, int n_incl_times, double max_dt, const double* strikes,
                         int n_strikes, const double* mats, int n_mats, const double* spreads, double* spots,
                         double* times, double* lvols, int* n_spots, int* n_times, int cap);

/* The chain step of dupireSuperbucket (main.h:536-566): risks to the risk
 * view of implied-vol spreads on strikes x mats (vega, strike major) from the
 * local-vol risks `microbucket` (n_spots x n_times of the calibration above,
 * spot major), through the calibration's derivative. */
int cf_superbucket_vega(const cf_merton_ivs* ivs, const double* incl_spots, int n_incl_spots, double max_ds,
                        const double* incl_times, int n_incl_times, double max_dt_vol, const double* strikes,
                        int n_strikes, const double* mats, int n_mats, const double* microbucket, int n_micro,
                        double* vega);

/* dupireSuperbucket (main.h:453-569): calibrates, stores the model as
 * "superbucket", runs dupireAADRisk on the GPU for the stored product and
 * notionals, and chains the local-vol risks through the calibration to the
 * risk view of implied-vol spreads on strikes x mats.  Writes value, delta and
 * vega[n_strikes * n_mats] (strike major). */
int cf_dupire_superbucket(const cf_merton_ivs* ivs, double max_dt, const char* product_id,
                          const char* const* labels, const double* notionals, int n,
                          const double* incl_spots, int n_incl_spots, double max_ds,
                          const double* incl_times, int n_incl_times, double max_dt_vol,
                          const double* strikes, int n_strikes, const double* mats, int n_mats,
                          const cf_num* num, double* value, double* delta, double* vega);

/* dupireSuperbucketBump (main.h:575-696): the superbucket by bump and
 * reprice on the GPU (FP64 or FP32 as num asks): value, delta by a 1e-8 spot
 * bump, and vega[n_strikes * n_mats] (strike major), each risk-view spread
 * bumped by 1e-5 with the local vols recalibrated.  The product is the stored
 * product_id; labels/notionals as cf_dupire_superbucket. */
int cf_dupire_superbucket_bump(const cf_merton_ivs* ivs, double max_dt, const char* product_id,
                               const char* const* labels, const double* notionals, int n,
                               const double* incl_spots, int n_incl_spots, double max_ds,
                               const double* incl_times, int n_incl_times, double max_dt_vol,
                               const double* strikes, int n_strikes, const double* mats, int n_mats,
                               const cf_num* num, double* value, double* delta, double* vega);

/* ------------------------------------------------------------ request server */

/* Continuous batching of pricing and risk requests on one GPU; no counterpart
 * in the reference (its entry points are synchronous, AAD.cpp:22-26).  Each
 * request is one mcParallelSimul (mcBase.h:314-400) or, with aad,
 * mcParallelSimulAAD (mcBase.h:564-753) call with a linear aggregator.
 * Requests are admitted in arrival order while one of `slots` HIP streams is
 * free and the request's device block (tables, checkpoints, partial sums)
 * fits in the server's device arena of hbm_budget bytes, allocated once at
 * creation; a checkpoint AAD request that does not fit runs the recompute
 * variant when that fits and auto_recompute is set.  Requests arriving while
 * others run start at once and their kernels overlap on the device. */
typedef struct cf_server cf_server;

typedef struct cf_server_opts {
    int slots;               /* requests on the device at once, 1..16 */
    int64_t hbm_budget;      /* device bytes of the arena */
    int auto_recompute;      /* 1: switch AAD to the recompute variant to fit */
    int graphs;              /* 1: a slot that launches the same sequence a
                              * third time (same kernels, parameters,
                              * addresses and sizes) captures it in a HIP graph
                              * and replays the graph for later repeats;
                              * 0: direct launches only */
    int max_pack;            /* queued requests of one path kernel that a slot
                              * runs in one packed launch, 1..8 (0: 8; 1: no
                              * packing).  Only requests too small to fill the
                              * device alone are packed. */
} cf_server_opts;

typedef struct cf_request {
    cf_model model;          /* copied at submission, arrays included */
    cf_product product;
    cf_num num;
    int aad;                 /* 0: prices; 1: prices, aggregate and risks */
    const double* weights;   /* aad: aggregation weights [n_pay]; NULL: first payoff */
} cf_request;

typedef struct cf_result {
    /* in: caller buffers and their capacities (nullable) */
    double* means;
    double* std_errs;
    int cap_pay;
    double* risks;
    int cap_risk;
    /* out */
    int status;              /* CF_OK or the request's error code */
    char error[256];
    int n_pay, n_param;
    double agg_mean;
    double submit_ms;        /* submission time since the server's creation */
    double queue_ms;         /* submission to launch */
    double latency_ms;       /* submission to results on the host */
    double kernel_ms;        /* path kernel (HIP events); with graphs, the
                              * request's whole device sequence */
    int64_t device_bytes;    /* arena bytes held while on the device */
    int recomputed;          /* 1 if admitted as the recompute variant */
    int packed;              /* requests in the launch that ran it (1: alone) */
    double plan_ms;          /* submission: validation and planning (part of queue_ms) */
    double launch_ms;        /* host time queueing its copies and kernels */
    double device_ms;        /* its launch's whole device sequence (HIP events) */
} cf_result;

typedef struct cf_server_stats {
    int64_t submitted, completed, failed;
    int64_t queued, running;
    int64_t peak_running;         /* most requests on the device at once */
    int64_t admitted_busy;        /* requests started while others ran */
    int64_t memory_waits;         /* admission passes that found no room */
    int64_t recompute_switches;   /* AAD requests switched to recompute */
    int64_t arena_bytes, peak_arena_bytes;
    int64_t graph_captures;       /* launch sequences captured and instantiated
                                   * (and launched) */
    int64_t graph_replays;        /* launches that replayed a captured graph */
    int64_t packed_launches;      /* launches that ran two or more requests */
    int64_t packed_requests;      /* requests run in those launches */
} cf_server_stats;

/* opts NULL: 4 slots, 16 GiB, auto_recompute on, graphs on, packs of up to 8.
 * Uses the current device. */
int cf_server_create(const cf_server_opts* opts, cf_server** out);
/* Validates and plans the request on the calling thread, then queues it. */
int cf_server_submit(cf_server* srv, const cf_request* req, int64_t* id);
/* 1 when the request's results are ready, 0 when not, negative on error. */
int cf_server_poll(cf_server* srv, int64_t id);
/* Blocks until the request is done, writes its results and releases it;
 * returns its status.  One wait per request: a second wait on the same id,
 * while the first blocks or after it, returns an error. */
int cf_server_wait(cf_server* srv, int64_t id, cf_result* out);
int cf_server_get_stats(cf_server* srv, cf_server_stats* out);
/* Finishes every queued request, then releases the server. */
int cf_server_destroy(cf_server* srv);

/* ------------------------------------------------------------ runtime */

/* Thread-local message of the last failing call. */
const char* cf_last_error(void);

/* Timings of the last simulation call on this thread, from HIP events on the
 * stream the kernels run on. */
typedef struct cf_kernel_stats {
    double main_kernel_ms;   /* path kernel (forward, or forward + adjoint) */
    double reduce_ms;        /* cross-block reductions */
    double total_ms;         /* whole call on the device, uploads included */
    int grid_blocks;
    int block_threads;
    int lds_bytes;
    int64_t path_steps;      /* paths x simulation steps */
    double setup_ms;         /* host time from the call to the kernel launch */
    double host_ms;          /* host time of the whole call */
    int lean_layout;         /* 1 when the LEAN shared-memory layout was used */
    double enqueue_ms;       /* host time queueing the call's copies and kernels */
    double wait_ms;          /* host time waiting for the device after queueing */
} cf_kernel_stats;
int cf_last_kernel_stats(cf_kernel_stats* out);

/* Wait for all work the engine queued on the current device (the simulation
 * calls already return only when their results are on the host). */
int cf_device_synchronize(void);

/* Sobol points of paths [first_path, first_path + n_paths) in dim (1..1101)
 * dimensions, drawn on the device exactly as the path kernels draw them:
 * path p uses Sobol index p + 1, the state after skipTo(p) and next()
 * (sobol.h:77-151).  words: [n_paths][dim] uint32 state words (nullable);
 * normals: [n_paths][dim] Gaussians in the precision's real type, the
 * reference's nextG (sobol.h:110-116, invNormalCdf gaussians.h:47-87) in
 * FP64, the kernels' table form in FP32 (nullable).  layout: 0 the kernels'
 * FULL form (low words staged per lane, [dimension][lane]), 1 LEAN (formed
 * from six direction numbers), 2 FULL interleaved by groups of four
 * dimensions and read four at a time (the FP32 adjoint kernels' form), -1
 * FULL when its tables fit the block's shared memory. */
int cf_sobol_points(int64_t first_path, int64_t n_paths, int dim, int precision, int layout, uint32_t* words,
                    void* normals);

/* Library version string. */
const char* cf_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CFAMD_H */
