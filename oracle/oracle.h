/*
 * oracle.h -- CPU restatement of the reference's connected-components
 * contraction path.  TEST INFRASTRUCTURE ONLY: it is linked or called by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, to
 * check results and as the CPU baseline; the product (libkacomp) never calls it.
 *
 * Pinned by (tests/test_oracle.py):
 *   or_canonical_labels   closed form (min id per component) against brute force
 *                         BFS on every graph with <= 6 vertices, invariants, and the
 *                         reference fixture example/simple.graph -> [0,0,0,3,3,3]
 *   or_exp_deviate        against std::minstd_rand0 + std::exponential_distribution
 *                         (the calls of exponential_contraction.h:467-471) built with
 *                         and without fused multiply-add contraction
 *   or_lp_fixpoint        against or_lp_closed_form (multi-source Dijkstra over the
 *                         same deviates), which is the fixed point's closed form
 *   or_splitmix64_mix     the published SplitMix64 output sequences (seeds 0 and
 *                         1234567), bijectivity on a sample, avalanche
 *                         (tests/test_oracle_streams.py, as the next three)
 *   or_hashed_uniform /   Kolmogorov-Smirnov against U(0,1) and against the
 *   or_hashed_deviate     continuous exponential law; the truncated tail law
 *                         P(dev >= k) = exp(-beta k) (exponential_contraction.h:
 *                         467-471); no correlation across ids and iterations
 *   or_uniform_deviate_100  chi-squared against the uniform law on 0..99 of
 *                         local_contraction.h:173-176
 *   or_run (LOCAL) keys   level 0 against the closed form of two synchronous
 *                         passes: the minimum key over the closed 2-hop
 *                         neighbourhood, from the sparse product (A+I)^2
 *   or_run (EXP/LOCAL)    labels against or_canonical_labels; per-level sizes and
 *                         keys against independent recomputation in the tests
 *   or_run (LABELS/SHORTCUT) labels against or_canonical_labels
 *   or_mlp_grad (FP64)    central finite differences of its own loss
 *   or_mlp_* (bf16 points) the same network with the device's bf16 rounding;
 *                         pinned through the FP64 form (rounding a network whose
 *                         weights are already bf16 values changes nothing else)
 *   or_level_features     hand-computed features of small graphs
 *   or_replicate          the split graph is symmetric and simple, every proxy
 *                         hangs off its hub, hubs keep only their proxies, and
 *                         merging every proxy into its hub gives back the input's
 *                         edges exactly (tests/test_oracle_streams.py)
 */
#ifndef KC_ORACLE_H_
#define KC_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Canonical labels: minimum vertex id of each component (union-find). */
void or_canonical_labels(uint64_t n, const uint64_t *off, const uint32_t *tgt, uint32_t *labels);

/* Deviate of RunContractionLP restated step by step (fma selects the fused
 * canonical sum). */
uint32_t or_exp_deviate(uint32_t gid, int32_t seed, uint32_t iteration, uint64_t rng_offset, double beta, int32_t fma);
/* The same draw through the standard library objects the reference uses,
 * compiled with contraction off / with FMA contraction. */
uint64_t or_std_deviate_nofma(uint32_t gid, int32_t seed, uint32_t iteration, uint64_t rng_offset, double beta);
uint64_t or_std_deviate_fma(uint32_t gid, int32_t seed, uint32_t iteration, uint64_t rng_offset, double beta);
int32_t or_std_fma_compiled(void);

/* LP sweep loop (exponential_contraction.h:518-586) from key0 to its fixed
 * point, vertex order, in place.  Returns the number of sweeps (exchange
 * rounds, the last one without change). */
uint64_t or_lp_fixpoint(uint64_t n, const uint64_t *off, const uint32_t *tgt, const uint64_t *key0, uint64_t hop,
                        uint64_t *key_out);
/* Closed form of the same fixed point: key(v) = min_u (key0(u) + d(u,v)*hop). */
void or_lp_closed_form(uint64_t n, const uint64_t *off, const uint32_t *tgt, const uint64_t *key0, uint64_t hop,
                       uint64_t *key_out);

/* Full runs.  variant: 0 exp, 1 local, 2 shortcut, 3 labels (kc_variant). */
typedef struct or_result or_result;
/* stream: 0 reference draw, 1 hashed draw (the library's KC_STREAM_*).
 * pre_mode: 0 none, 1 local, 2 single, 3 double (the library's KC_PRE_*): the
 * components inside pre_blocks blocks of consecutive level indices contracted
 * first (levels with distance-0 keys, sweeps 0, radius 0). */
/* replicate: level 0 with its hubs (degree >= hd_threshold * sqrt(n0)) split
 * into proxies (the library's kc_options.replicate; exp only). */
or_result *or_run(uint64_t n, const uint64_t *off, const uint32_t *tgt, int32_t variant, int32_t seed, double beta,
                  int32_t fma, int32_t stream, uint64_t seq_limit, int32_t keep_keys, int32_t pre_mode,
                  uint64_t pre_blocks, int32_t replicate, double hd_threshold);
/* hubs split and proxies added by replicate */
void or_replication(const or_result *r, uint64_t out[2]);
/* Level 0 of the input with its hubs split: returns its vertex count n0 + R;
 * with off_out non-null also writes its CSR (n0 + R + 1 offsets; the caller
 * sizes tgt_out by off[n] + 2 * R) and the id of every vertex (a proxy has its
 * hub's). */
uint64_t or_replicate(uint64_t n, const uint64_t *off, const uint32_t *tgt, int32_t seed, double hd_threshold,
                      uint64_t *off_out, uint32_t *tgt_out, uint32_t *gid_out);
void or_free(or_result *r);
void or_labels(const or_result *r, uint32_t *labels);
uint32_t or_levels(const or_result *r);
/* info: n_l, m_l (undirected), sweeps_l, max radius_l */
void or_level_info(const or_result *r, uint32_t level, uint64_t info[4]);
void or_level_keys(const or_result *r, uint32_t level, uint64_t *keys, uint32_t *gid);
uint64_t or_total_sweeps(const or_result *r);
double or_seconds(const or_result *r);

/* Hashed-stream deviate: SplitMix64 uniform, then -log(1-u)/beta with std::log. */
uint32_t or_hashed_deviate(uint32_t gid, int32_t seed, uint32_t iteration, double beta);
/* Uniform deviate in [0,99] used by local contraction (SplitMix64). */
uint32_t or_uniform_deviate_100(uint32_t gid, int32_t seed, uint32_t iteration);
/* SplitMix64's output function, shared by both streams above. */
uint64_t or_splitmix64_mix(uint64_t z);
/* The 53-bit uniform in [0,1) behind or_hashed_deviate. */
double or_hashed_uniform(uint32_t gid, int32_t seed, uint32_t iteration);

/* Learned centre proposals (no reference counterpart).  w: KC_ORACLE_MLP_PARAMS
 * parameters in the library's public order (W1[6][64], b1, W2[64][64], b2, w3, b3). */
#define KC_ORACLE_MLP_PARAMS 4673
void or_mlp_forward(const double *w, const float *X, uint64_t n, int32_t bf16_points, double *scores);
void or_mlp_grad(const double *w, const float *X, const float *y, uint64_t n, int32_t bf16_points, double *grads,
                 double *loss);
void or_level_features(uint64_t n, const uint64_t *off, const uint32_t *tgt, uint32_t round, float *X);

#ifdef __cplusplus
}
#endif

#endif
