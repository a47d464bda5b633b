/*
 * dqmc_hip.h -- C ABI of the MI355X (gfx950) DQMC local-sweep library.
 *
 * Drop-in boundary for MonteCarlo.jl's DQMC hot path.  The reference keeps one chain's
 * Green's-function stack in `DQMCStack` (src/flavors/DQMC/stack.jl:1-74) and runs the
 * sweep through update(::LocalSweep, mc, model, field) (updates/local_updates.jl:82).
 * A `HIPStack <: AbstractDQMCStack` on the Julia side owns one `dqmc_chain_t` and forwards
 * each stack method to the entry point named below (ccall; plain pointers and sizes).
 *
 * Conventions
 *   - Matrices are column-major FP64 with an explicit leading dimension `ld` (>= n).
 *   - A block-diagonal (two-flavour, MagneticHirschField) matrix is exchanged as an
 *     n x (flavors*n) array, block b in columns b*n .. b*n+n-1.
 *   - The Hirsch field is an n x slices Int8 array of +-1 (fields.jl:329-330).
 *   - Slice numbers are 1-based as in the reference; pivots are 0-based.
 *   - Every entry point returns 0 on success and a negative DQMC_E* code on failure;
 *     dqmc_last_error() describes the failure.  Numerical events (negative or tied
 *     acceptance probabilities, propagation errors) never fail a call: they are counted
 *     in dqmc_stats_t, as the reference records them in DQMCAnalysis.
 *   - A chain handle is used by one host thread at a time; its work runs on its own
 *     HIP stream on the device chosen at creation.
 */
#ifndef DQMC_HIP_H
#define DQMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DQMC_OK 0
#define DQMC_E_INVALID (-1)
#define DQMC_E_HIP (-2)
#define DQMC_E_STATE (-3)
#define DQMC_E_ALLOC (-4)
#define DQMC_E_DEVICE (-5) /* a kernel gave up a bounded wait or met a broken invariant:
                              the chain's G is not valid (count in dqmc_stats_t.device_errors) */

#define DQMC_E_PRECISION (-6) /* mixed precision: a stabilisation found max|G_wrapped - G_scratch|
                                 above dqmc_config_t.mixed_tolerance (count and largest value
                                 in dqmc_stats_t; G was replaced by the FP64 recompute)     */

#define DQMC_PRECISION_FP64 0  /* everything in FP64 */
#define DQMC_PRECISION_MIXED 1 /* FP32 wraps and rank-1 updates between stabilisations; FP64
                                  UDT, recompute and detratio.  Every stabilisation checks
                                  max|G_wrapped - G_scratch| against mixed_tolerance.        */

typedef struct dqmc_chain_s* dqmc_chain_t;

typedef struct {
  int n_sites;      /* N = length(lattice(model)); the block size                          */
  int flavors;      /* 1 = DensityHirschField, 2 = MagneticHirschField (fields.jl:378,427)  */
  int slices;       /* M = DQMCParameters.slices (parameters.jl:117)                        */
  int safe_mult;    /* DQMCParameters.safe_mult                                             */
  double delta_tau; /* DQMCParameters.delta_tau                                             */
  double alpha;     /* Hirsch coupling field.alpha (fields.jl:370-376, 419-425)             */
  int precision;    /* DQMC_PRECISION_*                                                     */
  int check_propagation_error; /* DQMCParameters.check_propagation_error                   */
  int delay_depth;  /* sites per local-update window: <= 16 -> 16, <= 32 -> 32, else 64;
                       0 = automatic (32)                                              */
  int device;       /* HIP device ordinal                                                   */
  double mixed_tolerance; /* DQMC_PRECISION_MIXED: bound on max|G_wrapped - G_scratch| at each
                             stabilisation; 0 = 1e-6.  Ignored in FP64.                    */
} dqmc_config_t;

typedef struct {
  int64_t accepted;             /* accepted local flips (sum of sweep_spatial returns)       */
  int64_t proposed;             /* proposals                                                 */
  int64_t uniforms_used;        /* uniforms consumed (one per proposal with p <= 1)          */
  int64_t negative_probability; /* proposals with p < 0 (local_updates.jl:41-47)             */
  int64_t ties;                 /* |u - p| < 1e-12, or 1 < p < 1 + 1e-12                     */
  int64_t stream_exhausted;     /* supplied uniform stream ran out (then u = 1, reject)      */
  int64_t propagation_error_count; /* stabilisations with max|dG| > 1e-7 (stack.jl:644-654) */
  int64_t stabilizations;       /* stabilisations checked                                    */
  int64_t sweeps;               /* completed sweeps                                          */
  int64_t delayed_flushes;      /* rank-k flushes of the delayed update                      */
  double propagation_error_max; /* max over stabilisations of max|G_wrapped - G_scratch|    */
  double propagation_error_last;
  int64_t device_errors;        /* kernels that gave up a bounded wait or met a broken
                                   invariant (each makes the call that sees it return
                                   DQMC_E_DEVICE)                                          */
  int64_t precision_violations; /* mixed precision: stabilisations above mixed_tolerance    */
} dqmc_stats_t;

/* DQMCStack(field, model, checkerboard=false) + initialize_stack (stack.jl:113-206).
 * Allocates every device buffer of the chain (footprint: dqmc_footprint_bytes).        */
int dqmc_create(const dqmc_config_t* cfg, dqmc_chain_t* out);
int dqmc_destroy(dqmc_chain_t h);
const char* dqmc_last_error(dqmc_chain_t h); /* h may be NULL: last creation error */
int dqmc_footprint_bytes(const dqmc_config_t* cfg, int64_t* bytes);

/* init_hopping_matrices (stack.jl:209-249): exp(-+dtau T/2), exp(-+dtau T) of the N x N
 * hopping matrix (hopping_matrix, HubbardModel.jl:112-124), computed on the device.    */
int dqmc_set_hopping(dqmc_chain_t h, const double* T, int ld);
/* Same, with the four exponentials supplied (e.g. the reference's own fallback_exp).   */
int dqmc_set_hopping_exp(dqmc_chain_t h, const double* exp_mhalf, const double* exp_phalf,
                         const double* exp_m, const double* exp_p, int ld);
int dqmc_get_hopping_exp(dqmc_chain_t h, double* exp_mhalf, double* exp_phalf, double* exp_m,
                         double* exp_p, int ld);

/* field.conf (fields.jl:329-334), N x M Int8 +-1. */
int dqmc_set_field(dqmc_chain_t h, const int8_t* conf, int ld);
int dqmc_get_field(dqmc_chain_t h, int8_t* conf, int ld);
/* The field in the reference's compressed form: compress(f) = BitArray(f.conf .== 1),
 * decompress!(f, c) = (f.conf .= Int8.(2c .- 1)) (fields.jl:331-334), as Julia stores a
 * BitArray: nchunks = ceil(n_sites * slices / 64) uint64 chunks, bit k % 64 of chunk k / 64
 * for column-major entry k, least significant bit first, high bits of the last chunk zero
 * (set_field_packed refuses others).  Packed and unpacked on the device; set, then
 * build_stack, resumes a chain from a stored field.                                      */
int dqmc_get_field_packed(dqmc_chain_t h, uint64_t* chunks, int64_t nchunks);
int dqmc_set_field_packed(dqmc_chain_t h, const uint64_t* chunks, int64_t nchunks);

/* Randomness of the Metropolis test `p > 1 || rand() < p` (local_updates.jl:53): a
 * uniform is consumed only when p <= 1.  Either a supplied stream (copied to the
 * device) or the counter stream u_k = SplitMix64(seed, k) >> 11 * 2^-53 from `position`. */
int dqmc_set_uniforms(dqmc_chain_t h, const double* u, int64_t count);
int dqmc_set_seed(dqmc_chain_t h, uint64_t seed, uint64_t position);
/* Optional per-proposal log (accept flag and p) for parity checks. */
int dqmc_set_decision_log(dqmc_chain_t h, int64_t capacity);
int dqmc_get_decision_log(dqmc_chain_t h, uint8_t* accepted, double* p, int64_t capacity,
                          int64_t* length);

/* reverse_build_stack + propagate, as initialize_run does (DQMC.jl:178-179).           */
int dqmc_build_stack(dqmc_chain_t h);
/* build_stack (stack.jl:257-281) + propagate.                                          */
int dqmc_build_stack_forward(dqmc_chain_t h);
/* reverse_build_stack (stack.jl:284-308) alone.                                        */
int dqmc_reverse_build_stack(dqmc_chain_t h);
/* propagate (stack.jl:605-730).                                                        */
int dqmc_propagate(dqmc_chain_t h);
/* sweep_spatial (local_updates.jl:23-60) at the current slice; *accepted may be NULL.  */
int dqmc_sweep_spatial(dqmc_chain_t h, int64_t* accepted);
/* nsweeps x local_sweep (local_updates.jl:7-14): 2M x (sweep_spatial; propagate).      */
int dqmc_sweep(dqmc_chain_t h, int nsweeps, int64_t* accepted);

/* Batched form: `count` distinct chains sweep `nsweeps` times each, each exactly as
 * dqmc_sweep would alone (bit for bit).  accepted[i] (optional) receives chain i's accepted
 * flips.  FP64 chains on one device run in lockstep slice visits on one stream: every launch
 * covers all chains (window kernels grouped by flavours and window width, the propagate
 * kernels of chains at the same kind of step together), with per-chain item tables in device
 * memory; chains with fewer slices finish their sweeps earlier.  Other mixes (mixed precision,
 * several devices, kernel timing on) run one host thread and stream per chain.  No reference
 * counterpart: the reference runs one chain per Julia process (run!, DQMC.jl:252-394); this
 * is the many-chain form of update(::LocalSweep) (local_updates.jl:82) for independent
 * chains. */
int dqmc_batch_sweep(dqmc_chain_t* chains, int count, int nsweeps, int64_t* accepted);

/* Serving form: a request loop over independent chains with continuous batching.  No
 * reference counterpart as a library call: the reference serves one chain per Julia process
 * through run! (DQMC.jl:144-394); a request is what one such process runs (create the stack,
 * build it, nsweeps x update(::LocalSweep), read back field, G and statistics).
 *   - Requests are admitted in arrival order while fewer than max_active chains run and the
 *     sum of their footprints (dqmc_footprint_bytes) fits memory_budget; requests arriving
 *     while others run are admitted as slots and memory free up, and run concurrently, each
 *     on its own host thread and HIP stream on `device`.
 *   - A result is the same as the chain run alone (its kernels touch only its own buffers).
 *   - The server's entry points may be called from several host threads;
 *     dqmc_server_last_error describes the last failure of the calling thread.
 *   - dqmc_server_destroy marks queued requests failed, waits for running ones and for
 *     threads inside dqmc_server_wait to return, then frees the server.
 *   - request_state: 0 queued, 1 running, 2 done, 3 failed.                              */
typedef struct dqmc_server_s* dqmc_server_t;
typedef struct {
  int device;             /* HIP device ordinal of every chain                              */
  int max_active;         /* chains running at once (>= 1)                                  */
  int64_t memory_budget;  /* bytes of device memory the running chains may hold together    */
} dqmc_server_config_t;
typedef struct {
  dqmc_config_t chain;    /* shape of the chain (chain.device is replaced by the server's)  */
  const double* hopping;  /* n x n hopping matrix T (copied at submit)                       */
  int ld_hopping;
  const int8_t* field;    /* n x slices initial +-1 field (copied at submit)                 */
  int ld_field;
  uint64_t seed;          /* counter-stream seed, position 0 (dqmc_set_seed)                 */
  int nsweeps;            /* sweeps after the stack build                                    */
} dqmc_request_t;
int dqmc_server_create(const dqmc_server_config_t* cfg, dqmc_server_t* out);
int dqmc_server_submit(dqmc_server_t s, const dqmc_request_t* request, int64_t* id);
int dqmc_server_poll(dqmc_server_t s, int64_t id, int* request_state);
int dqmc_server_wait(dqmc_server_t s, int64_t id);
/* Copies the final field (n x slices), G (n x flavors*n, physical = 0) and statistics of a
 * finished request and releases it; latency[0] = submit -> done, latency[1] = time queued
 * (seconds).  Any output pointer may be NULL.                                             */
int dqmc_server_result(dqmc_server_t s, int64_t id, int8_t* field, int ld_field, double* G, int ld_G,
                       int64_t* accepted, dqmc_stats_t* stats, double* latency);
int dqmc_server_status(dqmc_server_t s, int* queued, int* active, int* peak_active,
                       int64_t* used_bytes);
/* Drops queued requests, waits for running ones, frees the server. */
int dqmc_server_destroy(dqmc_server_t s);
const char* dqmc_server_last_error(dqmc_server_t s);

/* Equal-time observables on the device from the physical G (greens.jl:114-125):
 * occupation (measurements/constructors/occupation.jl:44-55), kinetic/total energy
 * (measurements/constructors/energy.jl:119-166), interaction energy with the reference's U
 * (models/HubbardModel.jl:160-182), double occupancy (1 - G_up_ii)(1 - G_dn_ii).
 * out[0] density per site, [1] kinetic, [2] interaction, [3] total energy, [4] mean double
 * occupancy, [5, 5 + flavors n) occupation per flavour block, then n values of double
 * occupancy per site; nout >= 5 + (flavors + 1) n.  Needs dqmc_set_hopping (T itself). */
int dqmc_measure_equal_time(dqmc_chain_t h, double U, double* out, int nout);

/* mc.stack.greens (physical = 0) or greens(mc) = e^{dtau T/2} G e^{-dtau T/2}
 * (physical = 1, greens.jl:87-125); n x (flavors*n).                                   */
int dqmc_get_greens(dqmc_chain_t h, double* G, int ld, int physical);
int dqmc_set_greens(dqmc_chain_t h, const double* G, int ld);
/* calculate_greens(mc, slice, output) from scratch (stack.jl:525-583); does not touch
 * the chain's current G.                                                               */
int dqmc_calculate_greens(dqmc_chain_t h, int slice, double* G, int ld);

/* Time-displaced Green's function G(k <- l) = <c_i(k dtau) c_j^dagger(l dtau)>, 0 <= k, l <=
 * slices: greens(mc, k, l) (physical = 1: e^{dtau T/2} G e^{-dtau T/2}, greens.jl:114-125)
 * or calculate_greens(mc, k, l) (physical = 0), unequal_time_stack.jl:297-338, from the
 * chain's unequal-time stack (forward, backward and inverse UDT slots per safe_mult range,
 * :1-42, 124-278, built lazily on first use for the current field and kept until the field
 * or the hopping changes); k >= l by calculate_greens_full1! (:535-617), k < l by
 * calculate_greens_full2! (:620-697).  n x (flavors*n).  Leaves the chain's G, stack and
 * state untouched.                                                                     */
int dqmc_unequal_time_greens(dqmc_chain_t h, int k, int l, double* G, int ld, int physical);

/* CombinedGreensIterator(mc; start, stop, recalculate) iterated to its end, which is also
 * TimeIntegral's iteration (measurements/greens_iterators.jl:83-407): for l = start..stop
 * the physical G(0, l), G(l, 0), G(l, l), written to `out` as 3 (stop - start + 1)
 * consecutive n x (flavors*n) matrices (leading dimension ld) in that order per l.  G(0, 0)
 * is the chain's own G when its current slice is 1 (FP64 form), else full1!(0, 0); every
 * `recalculate` slices (the reference's default is 2 safe_mult) the triple is recomputed
 * from the unequal-time stack, in between advanced by one slice product per step and
 * re-decomposed every safe_mult slices.  out_on_device = 1: `out` is device memory on the
 * chain's GPU and ld must be n.  Leaves the chain's G, stack and state untouched.       */
int dqmc_combined_greens(dqmc_chain_t h, int start, int stop, int recalculate, double* out, int ld,
                         int out_on_device);

/* charge_density_susceptibility(mc, model) and spin_density_susceptibility(mc, model, dir)
 * measured once for the current field (measurements/constructors/charge_density.jl:47,
 * spin_density.jl:54): the TimeIntegral iteration over l = 0..slices (trapezoid weights dtau,
 * halved at both ends; measurements/generic.jl:337-370) with G00 = the chain's current
 * physical G, summed by EachSitePairByDistance (generic.jl:434-460) with the observable's
 * full kernel (reduced = 0) or reduced kernel (reduced = 1): full/reduced_cdc_kernel
 * (charge_density.jl:62-170) over the flavour pairs, full/reduced_sdc_{x,y,z}_kernel
 * (spin_density.jl:68-330); one flavour block is the reference's DiagonallyRepeatingMatrix
 * form (with total_flavors), two blocks its BlockDiagonal form.  The sum is divided by the
 * number of sites (finish!, generic.jl:574-586).  dir2srctrg: the lattice's
 * Bravais_dir2srctrg (lattice_cache.jl:226-242), 0-based, n_bravais rows (row = shift) of
 * n_bravais targets; sites are numbered Bravais cell + n_bravais * basis.  out: n_bravais *
 * n_basis^2 values, index dir + n_bravais (b1 + n_basis b2).  Allocates its accumulator per
 * call.                                                                                  */
#define DQMC_CHARGE_DENSITY 0
#define DQMC_SPIN_DENSITY_X 1
#define DQMC_SPIN_DENSITY_Y 2
#define DQMC_SPIN_DENSITY_Z 3
int dqmc_density_susceptibility(dqmc_chain_t h, int observable, int recalculate, const int32_t* dir2srctrg,
                                int n_bravais, int n_basis, int total_flavors, int reduced, double* out);

/* global_update(mc, model, field) with `proposal` (n x slices, +-1) as the proposed conf
 * (global_updates.jl:191-207; propose_global_from_conf :136-169, inv_det :64-121,
 * calculate_inv_greens_udt :25-47, accept_global! :171-188): p = |exp(-dE) detratio| from
 * the D vectors of the current G and of the proposal's G from scratch; accepted iff p > 1 or
 * u < p with u the chain's next uniform (consumed only when p <= 1); an accepted proposal
 * becomes the field and the stack is rebuilt (reverse_build_stack + propagate), a rejected
 * one leaves field and G as they were.  Needs current_slice == 1 and direction == 1 right
 * after a sweep or a stack build.  *accepted and *p (either may be NULL) get the outcome. */
int dqmc_global_update(dqmc_chain_t h, const int8_t* proposal, int ld, int* accepted, double* p);

int dqmc_get_state(dqmc_chain_t h, int* current_slice, int* direction, int* current_range);
int dqmc_get_stats(dqmc_chain_t h, dqmc_stats_t* stats);
/* Bound of the window kernel's internal wait (polls of an LDS progress counter, default
 * 2^22, about 0.1 s): a wave that reaches it leaves its loop, the launch still ends, and
 * the event is a device error.  No reference counterpart (the reference's update is
 * sequential host code).                                                               */
int dqmc_set_wait_limit(dqmc_chain_t h, int64_t polls);
/* A sweep's launch sequence is the same every sweep (the decisions are made on the device),
 * so FP64 chains capture one sweep as a HIP graph on first use and replay it while the host
 * state it baked in (propagate state, buffer roles, seed / uniform / log buffers, wait bound)
 * is unchanged, recapturing otherwise.  enable = 0 launches kernel by kernel.  Results are
 * the same either way.  No reference counterpart (the reference's sweep is host code). */
int dqmc_set_sweep_graph(dqmc_chain_t h, int enable);
/* FP64 chains with windows of <= 32 sites run each slice's local updates
 * (sweep_spatial!, local_updates.jl:23-60) as one launch: the windows back to back in one
 * workgroup, each window's update of G applied beside the next window's decisions, ordered by
 * bounded device-side waits (dqmc_set_wait_limit).  The launch's 1 + min(b (n/64)^2, 128)
 * workgroups must be resident together (one per CU): chains swept concurrently on other
 * streams of the same GPU should use enable = 0 or dqmc_batch_sweep, otherwise a wait can
 * run to its bound (reported as a device error).  enable = 0 launches once per window.
 * Results are bit-identical either way.  No reference counterpart (launch scheduling). */
int dqmc_set_slice_launch(dqmc_chain_t h, int enable);
/* max|G_wrapped - G_scratch| at every stabilisation since the last call (mixed precision:
 * FP32-propagated G against the FP64 recompute).                                       */
int dqmc_get_stabilization_log(dqmc_chain_t h, double* out, int64_t capacity, int64_t* length);
int dqmc_synchronize(dqmc_chain_t h);

/* Kernel timing (the reference's @bm / TimerOutputs timers, helpers.jl:83-131): when
 * enabled, every launch of the chain is bracketed by HIP events on the chain's stream and
 * the device time is accumulated per kernel kind.  Enabling resets the totals.          */
#define DQMC_KERNEL_GEMM_F64 0 /* slice products, wraps, calculate_greens products */
#define DQMC_KERNEL_GEMM_F32 1 /* mixed-precision wraps                            */
#define DQMC_KERNEL_UDT 2      /* UDT: column norms, Q initialisation, D/T output   */
#define DQMC_KERNEL_RDIVP 3    /* triangular solve of calculate_greens             */
#define DQMC_KERNEL_LOCAL 4    /* Metropolis decisions of one window of sites      */
#define DQMC_KERNEL_LOCAL_APPLY 5 /* rank-m application of a window's flips to G    */
#define DQMC_KERNEL_UDT_PANEL 6  /* UDT: pivot choice + Householder panel (per panel) */
#define DQMC_KERNEL_UDT_UPDATE 7 /* UDT: block-reflector trailing update / Q build   */
#define DQMC_KERNEL_KINDS 8
int dqmc_set_kernel_timing(dqmc_chain_t h, int enable);
int dqmc_get_kernel_times(dqmc_chain_t h, double* total_ms, int64_t* launches, int kinds);

/* Device linear algebra on host arrays (unit parity of the kernels).
 *   udt_AVX_pivot! (UDT.jl:216-334), rdivp! (real.jl:198-226),
 *   calculate_greens_AVX! (stack.jl:442-496), vmul! (real.jl:7-102).                   */
int dqmc_la_udt(int n, const double* X, int apply_pivot, double* U, double* D, double* T,
                int32_t* pivot);
int dqmc_la_rdivp(int n, double* A, const double* T, const int32_t* pivot);
int dqmc_la_calc_greens_udt(int n, const double* Ul, const double* Dl, const double* Tl,
                            const double* Ur, const double* Dr, const double* Tr, double* G);
int dqmc_la_gemm(int n, int trans_a, int trans_b, const double* A, const double* B, double* C,
                 int precision);

/* The UDT of dqmc_la_udt on device memory, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  X, U, T (n x n, column-major,
 * ld = n), D (n) and pivot (n) are device pointers on the current device; U must not alias
 * X or T, nor T alias X.  device_errors (device pointer, may be NULL) is incremented by
 * broken pivot selections.  Runs asynchronously on `stream` (a hipStream_t; NULL = the null
 * stream), with its working storage in a static device array of the library sized for
 * n <= DQMC_NOALLOC_MAX_N (about 8.4 MB per GPU, held until the process ends); larger n is
 * refused with DQMC_E_INVALID.  Calls that overlap on one GPU share that array and race.  */
#define DQMC_NOALLOC_MAX_N 1024
int dqmc_la_udt_noalloc(int n, const double* X, int apply_pivot, double* U, double* D, double* T,
                        int32_t* pivot, uint64_t* device_errors, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DQMC_HIP_H */
