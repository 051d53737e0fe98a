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
    CF_PRODUCT_EUROPEANS = 2   /* Europeans<T> mcPrd.h:291-400 */
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
} cf_product;

enum cf_precision { CF_FP64 = 0, CF_FP32 = 1 };
enum cf_aad_mode { CF_AAD_CHECKPOINT = 0, CF_AAD_RECOMPUTE = 1 };

/* NumericalParam (main.h:33-40) plus the engine's own options. */
typedef struct cf_num {
    int parallel;            /* accepted for compatibility; the GPU is always parallel */
    int use_sobol;           /* must be 1: the engine implements Sobol only */
    int num_path;
    int seed1;               /* unused with Sobol (main.h:51-52) */
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

/* main.h:364-411: value, delta and the n_spots x n_times vega matrix. */
int cf_dupire_aad_risk(const char* model_id, const char* product_id, const char* const* labels,
                       const double* notionals, int n, const cf_num* num,
                       double* value, double* delta, double* vega, int cap_vega);

/* Remove every stored model and product. */
void cf_store_clear(void);

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
} cf_kernel_stats;
int cf_last_kernel_stats(cf_kernel_stats* out);

/* Wait for all work the engine queued on the current device (the simulation
 * calls already return only when their results are on the host). */
int cf_device_synchronize(void);

/* Diagnostics: copies `bytes` bytes between two device buffers with a plain
 * streaming kernel, `reps` times.  Used to calibrate hardware byte counters
 * (a known traffic of 2 x bytes per launch).  No counterpart in the reference. */
int cf_diag_copy(size_t bytes, int reps);

/* Library version string. */
const char* cf_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CFAMD_H */
