/*
 * rt_hip.h — C ABI of the MI355X (gfx950) flip-MCMC library `librt_hip.so`.
 *
 * This is the boundary the per-model C++ classes (DominoTiler, LozengeTiler,
 * RectTriangleTiler, TriangleDimerTiler; randomtilings_amd/src/) drive, and what a
 * ctypes/cgo/JNI binding would bind.  Signatures use plain pointers and sizes only.
 * Every entry point returns RT_OK (0) or a negative RT_ERR_* code; rt_last_error()
 * returns a thread-local message for the last failure on the calling thread.
 *
 * Ownership: host arrays passed in are borrowed for the duration of the call.  The
 * library owns all device memory and one HIP stream per context.  Calls on one context
 * are serialized by the caller (as with the reference's one in-order queue per tiler);
 * separate contexts may be used concurrently.
 *
 * Tiling layouts are the reference's: an N x N row-major int32 vertex-state array
 * (domino: DominoTiler.h:17-37, lozenge: LozengeTiler.h:8-28, rect-triangle:
 * RectTriangleTiler.h:16-27, triangle-dimer: TriangleDimerTiler.h:13-69).
 */
#ifndef RT_HIP_H_
#define RT_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RT_OK 0
#define RT_ERR_ARG (-1)     /* invalid argument or shape */
#define RT_ERR_PARAMS (-2)  /* TinyMT parameter table missing or shorter than the chain needs */
#define RT_ERR_HIP (-3)     /* HIP runtime error */
#define RT_ERR_IO (-4)      /* parameter file missing or malformed */
#define RT_ERR_DEVICE (-5)  /* device-side error flag (e.g. a bounded grid-barrier wait expired) */
#define RT_ERR_UNTILEABLE (-6) /* domain has no tiling (max/min construction failed) */
#define RT_ERR_NOT_COALESCED (-7) /* coupling from the past did not coalesce within the bound */

/* models */
#define RT_DOMINO 0
#define RT_LOZENGE 1
#define RT_RECTTRI 2
#define RT_TRIDIMER 3
#define RT_DOMINO_PAIR 4 /* rt_chain_create only: the two coupled domino chains of coupling from
                            the past (rt_domino_cftp) as one chain.  Each int of the N x N tiling
                            holds the chain started from the maximal tiling in bits 0-7 and the
                            one started from the minimal tiling in bits 8-15; one random number
                            per site and step drives both, exactly as two Walk calls with the
                            same seeds (Examples/AztecDiamondCftpEx.cpp:49-60).  Rules:
                            RT_DOMINO_CFTP, RT_DOMINO_WEIGHTED.  No reference counterpart. */

/* domino rules (the reference selects them by kernel file name) */
#define RT_DOMINO_LAZY 0     /* dominokernel.cl: toggle 3<->12 iff rd < .8            */
#define RT_DOMINO_CFTP 1     /* dominokernelCFTP.cl: rd < .5 -> 12, else -> 3           */
#define RT_DOMINO_WEIGHTED 2 /* heat-bath, P(12) per (class, row parity); rparams = p12[4] */

/* lozenge rules */
#define RT_LOZENGE_UNIFORM 0 /* lozengekernel.cl: rd < .4 add, .4 <= rd < .8 remove     */
#define RT_LOZENGE_QREF 1    /* qlozengekernel.cl rule with a runtime q; rparams = {q}   */
#define RT_LOZENGE_QHEAT 2   /* heat-bath: add with prob q/(1+q), remove otherwise; {q}  */

/* seeding modes for rt_chain_seed */
#define RT_SEED_WALK 0  /* single-seed Walk: schedule RNG seeded with `seed`; TinyMT seed as the
                           reference's Walk(t, steps, seed) derives it per model            */
#define RT_SEED_EPOCH 1 /* epoch of Walk(t, vector steps, vector seeds): DominoTiler.cpp:79-82 */

/* coupling-from-the-past epoch schedules (rt_domino_cftp) */
#define RT_CFTP_REFERENCE 0 /* Examples/AztecDiamondCftpEx.cpp:42-68: start with one epoch of
                               first_steps; each restart prepends an epoch of twice the last
                               epoch's length (so 2*first_steps each time)                   */
#define RT_CFTP_DOUBLING 1  /* each restart prepends an epoch as long as all current epochs
                               together, doubling the total (Propp-Wilson)                   */

/* execution paths */
#define RT_PATH_AUTO 0       /* persistent temporal-blocked kernel where available        */
#define RT_PATH_PERSISTENT 1 /* persistent LDS-tiled kernel, generators resident in VGPRs */
#define RT_PATH_STEPWISE 2   /* one Rotate + Update launch pair per step (reference shape) */
#define RT_PATH_BLOCKED 3    /* LDS-tiled kernel, one launch per block of h steps: no
                                co-residency needed (large grids, shared devices)        */

typedef struct rt_ctx rt_ctx;
typedef struct rt_chain rt_chain;
typedef struct rt_batch rt_batch;
typedef struct rt_unet rt_unet;
typedef struct rt_observer rt_observer;

/* Thread-local description of the last error on this thread. */
const char* rt_last_error(void);

/* Library version string. */
const char* rt_version(void);

/* Number of HIP devices (0 when none or the runtime is unavailable). */
int rt_device_count(void);

/* Marketing name of HIP device `device` into buf (NUL-terminated, truncated to cap). */
int rt_device_name(int device, char* buf, int32_t cap);

/* Create a context on HIP device `device` (XTiler(ctx, queue, devices, kernelPath, err),
 * e.g. DominoTiler.h:61-67, minus the runtime OpenCL compile: kernels are AOT gfx950). */
int rt_create(int device, rt_ctx** out);
int rt_destroy(rt_ctx* ctx);

/* TinyMT parameter table: n entries of (mat1, mat2, tmat) interleaved in `mat3`.
 * Replaces get_params_buffer (src/TinyMT/file_reader.cpp:22-38). */
int rt_load_params(rt_ctx* ctx, const uint32_t* mat3, int64_t n);

/* Read the first n parameter sets of a tinymtdc-format file (lines starting with '#'
 * skipped; mat1/mat2/tmat are the hex fields after the 3rd/4th/5th comma).
 * Replaces XTiler::LoadTinyMT(path, size) (DominoTiler.cpp:109-111,
 * file_reader.cpp:51-70,139-166). */
int rt_load_params_file(rt_ctx* ctx, const char* path, int64_t n);

/* Host only (no device): parse the first n parameter sets of a tinymtdc-format file into
 * mat3 (3n words: mat1, mat2, tmat per entry), as tinymt::file_reader::get reads them
 * (src/TinyMT/file_reader.cpp:51-70,114-166): lines starting with '#' are skipped; an empty
 * line or the end of the file where an entry is expected, a line with fewer than five commas,
 * or a field strtoul reports out of range is RT_ERR_IO (the reference throws
 * std::runtime_error).  rt_load_params_file is this followed by rt_load_params. */
int rt_parse_params_file(const char* path, int64_t n, uint32_t* mat3);

/* Number of parameter sets currently loaded (0 if none). */
int64_t rt_params_count(rt_ctx* ctx);

/* Default execution path and tile configuration for chains created on this context
 * afterwards (no reference counterpart: the reference has one kernel shape). */
int rt_ctx_set_path(rt_ctx* ctx, int path, int32_t tile_config);

/* The integer thresholds on k = u>>9 that replace the reference's float decisions for
 * (model, rule, rparams): 24 words of domino/lozenge transition intervals
 * [class][row parity] (A->B lo, A->B hi, B->A lo, B->A hi) followed by the
 * rectangle-triangle rule: 36 words of thresholds (gate, 3 band limits, 5x4 cycle and 12
 * toggle acceptance limits), then its lookup tables (32 hash keys, 32 indices, 5x4 cycle
 * and 12 toggle targets).  Returns the word count (156), writing them when cap is at least
 * that (out may be NULL to ask for the count).  Host-only; used to check the derivation
 * against the reference predicates without a GPU. */
int32_t rt_rule_thresholds(int model, int rule, const double* rparams, int32_t nrparams, uint32_t* out,
                           int32_t cap);

/* Record HIP event `slot` (0..15) on the context's stream, and read the GPU time between
 * two recorded slots (synchronizes on `b`).  Used to time a region of queued work on the
 * stream the kernels actually run on. */
int rt_ctx_mark(rt_ctx* ctx, int32_t slot);
int rt_ctx_elapsed_ms(rt_ctx* ctx, int32_t a, int32_t b, double* ms);

/* Wait for all work queued on the context's stream. */
int rt_ctx_sync(rt_ctx* ctx);

/* ---- whole-walk entry points (what XTiler::Walk calls) ----------------------------- */

/* DominoTiler::Walk(tiling&, long steps, long seed) — DominoTiler.cpp:9-55.
 * rparams: NULL, or p12[4] for RT_DOMINO_WEIGHTED. */
int rt_domino_walk(rt_ctx* ctx, int rule, const double* rparams, int32_t* tiling, int32_t N, int64_t steps,
                   int64_t seed);

/* DominoTiler::Walk(tiling&, vector<long> steps, vector<long> seeds) — DominoTiler.cpp:58-107. */
int rt_domino_walk_epochs(rt_ctx* ctx, int rule, const double* rparams, int32_t* tiling, int32_t N,
                          const int64_t* steps, const int64_t* seeds, int32_t K);

/* LozengeTiler::Walk(tiling&, int steps, long seed) — LozengeTiler.cpp:18-88.
 * q is used by RT_LOZENGE_QREF / RT_LOZENGE_QHEAT (qlozengekernel.cl:3 hard-codes 0.99). */
int rt_lozenge_walk(rt_ctx* ctx, int rule, double q, int32_t* tiling, int32_t N, int64_t steps, int64_t seed);

/* rt_lozenge_walk that allocates nothing: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, semantics and result as rt_lozenge_walk; the device
 * memory it needs (two state copies, two generator copies, exchange flags, error word, colour
 * schedule) lies in a file-scope array of 64 MiB that each device the library runs on holds
 * until the process ends, its address taken once per device by rt_create.  Limit: the layout
 * must fit that array (about 12.7 * N^2 + 1.6 MiB bytes: N <= 2271, e.g. the hexagon
 * a = b = c = 1000 at N = 2007); a larger N is RT_ERR_ARG.  Tile skipping is off.  Calls on
 * one device share the array (each clears what must start at zero): calls that overlap on one
 * device race.  The parameter table must have been loaded beforehand (rt_load_params). */
int rt_lozenge_walk_noalloc(rt_ctx* ctx, int rule, double q, int32_t* tiling, int32_t N, int64_t steps,
                            int64_t seed);

/* RectTriangleTiler::Walk(tiling&, int steps, long seed) — RectTriangleTiler.cpp:17-92.
 * weights16: NULL for the reference's compile-time weights (recttrianglekernel.cl:3-13,
 * all 1.0), or the 16-entry per-digit weight table. */
int rt_recttri_walk(rt_ctx* ctx, const float* weights16, int32_t* tiling, int32_t N, int64_t steps, int64_t seed);

/* TriangleDimerTiler::Walk(tiling&, int steps, long seed) — TriangleDimerTiler.cpp:86-483.
 * The N x N lozenge-state tiling (N even, states 0..15); triangle and butterfly states are
 * derived on the device as TilingToTriangleFlips / TilingToButterflyFlips do.  Chains of
 * model RT_TRIDIMER: RT_PATH_PERSISTENT = every step in one single-workgroup launch (move
 * lattice N/2 <= 64), RT_PATH_STEPWISE = tiled launches (tile_config selects the tile, and
 * configurations 3-5 run the tiles of 0-2 two steps per launch with a halo of 12 instead of
 * 6), RT_PATH_AUTO = the former when it fits, else the tiled launches of the configuration
 * rt_chain_autotune chose for the lattice (configuration 0 without tuning). */
int rt_tridimer_walk(rt_ctx* ctx, int32_t* tiling, int32_t N, int64_t steps, int64_t seed);

/* Coupling from the past for the monotone domino rules (RT_DOMINO_CFTP, RT_DOMINO_WEIGHTED),
 * as Examples/AztecDiamondCftpEx.cpp:42-68 runs it over DominoTiler::Walk(t, steps, seeds):
 * epoch seeds come from std::mt19937(seed); every restart replays all epochs, oldest first,
 * on copies of tmax and tmin (the maximal and minimal tilings of the domain) and stops when
 * the two agree everywhere.  Both chains run as one coupled device chain (one random number
 * per site and step drives both) and the comparison is a device reduction.  On success
 * `out` holds the sample, *total_steps the steps of the final replay and *restarts the number
 * of restarts (either may be NULL).  RT_ERR_NOT_COALESCED after max_restarts restarts.
 * Unless rt_ctx_set_path fixed a path, the first call for a side times the tile
 * configurations on a scratch copy (as rt_chain_autotune does) and later calls reuse the
 * choice; the sample does not depend on it. */
int rt_domino_cftp(rt_ctx* ctx, int rule, const double* rparams, const int32_t* tmax, const int32_t* tmin,
                   int32_t N, int64_t seed, int mode, int64_t first_steps, int32_t max_restarts, int32_t* out,
                   int64_t* total_steps, int32_t* restarts);

/* rt_domino_cftp with a progress callback (NULL: none), called after every replay's coupling
 * test with the replay's index (0 = the first), its length in steps and whether it coalesced
 * (1 or 0), and within a replay after each epoch but the last with the steps done so far and
 * coalesced = -1 (the epoch is waited for first); for callers of long runs (order-1000
 * diamonds take minutes).  No reference counterpart. */
typedef void (*rt_cftp_progress)(void* user, int32_t restart, int64_t replay_steps, int32_t coalesced);
int rt_domino_cftp_ex(rt_ctx* ctx, int rule, const double* rparams, const int32_t* tmax, const int32_t* tmin,
                      int32_t N, int64_t seed, int mode, int64_t first_steps, int32_t max_restarts, int32_t* out,
                      int64_t* total_steps, int32_t* restarts, rt_cftp_progress progress, void* user);

/* ---- height functions ----------------------------------------------------------------- */

/* MaxTiling (dir = +1) / MinTiling (dir = -1) on the device — DominoTiler.cpp:214-400,
 * LozengeTiler.cpp:140-318.  domain: RT_DOMINO an M x M array of squares, RT_LOZENGE an
 * M x 2M array of triangles; tiling (and heights, may be NULL): (M+1) x (M+1).  The boundary
 * walk runs on the host, the closure of the boundary heights over the domain's directed unit
 * edges on the GPU, and the tiling is read off the heights on the GPU.  Returns
 * RT_ERR_UNTILEABLE with an all-zero tiling when the region has no tiling (the reference
 * prints a message and returns a partial height vector). */
int rt_extremal_tiling(rt_ctx* ctx, int model, const int32_t* domain, int32_t M, int dir, int32_t* tiling,
                       int32_t* heights);

/* Closure of a height field over the domain's directed unit edges, in place: dir = +1 gives
 * the largest field below `heights` with h(v) <= h(u) + 1 along every "+1" edge u -> v
 * (dir = -1 the smallest above it along the "-1" edges).  INT32_MAX marks unset vertices.
 * Vertices with a nonzero `fixed` entry (NULL: none) may not move: RT_ERR_UNTILEABLE if they
 * would.  Projects a predicted height field (rounded to the lattice's residue classes) onto
 * admissible ones.  *launches (may be NULL) receives the relaxation launches used.  No
 * counterpart in the reference (its construction runs on the host). */
int rt_height_closure(rt_ctx* ctx, int model, const int32_t* domain, int32_t M, int dir, int32_t* heights,
                      const uint8_t* fixed, int32_t* launches);

/* ---- device-resident chains (benchmarks, CFTP, batching) --------------------------- */

/* Upload a tiling and split it into colour classes on the device.  rparams as in the
 * walk entry points (domino: p12[4] or NULL; lozenge: {q}; rect-triangle: 16 weights). */
int rt_chain_create(rt_ctx* ctx, int model, int rule, const double* rparams, int32_t nrparams,
                    const int32_t* tiling, int32_t N, rt_chain** out);
int rt_chain_destroy(rt_chain* ch);

/* Select the execution path and, for RT_PATH_PERSISTENT, a tile configuration index
 * (-1 = the autotuned/default choice). */
int rt_chain_set_path(rt_chain* ch, int path, int32_t tile_config);

/* Autotune the persistent tile configuration for this chain's shape: every configuration
 * whose tile grid fits on the device runs `probe_steps` steps on a scratch clone of the
 * chain (the chain's own state and generators are untouched) and the fastest is selected
 * for this chain and remembered for later chains of the same model, side and parameter
 * mode.  Writes the chosen index to *best (may be NULL). */
int rt_chain_autotune(rt_chain* ch, int64_t probe_steps, int32_t* best);

/* (Re)seed: initialise the TinyMT generators on the device and the host colour-schedule
 * RNG exactly as the reference's Walk does for `mode` (RT_SEED_WALK / RT_SEED_EPOCH). */
int rt_chain_seed(rt_chain* ch, int mode, int64_t seed);

/* Enqueue `steps` chain steps on the context stream (asynchronous). */
int rt_chain_run(rt_chain* ch, int64_t steps);

/* Wait for the chain's work; returns RT_ERR_DEVICE if a kernel raised its error flag. */
int rt_chain_sync(rt_chain* ch);

/* GPU milliseconds between the first and last kernel of the most recent rt_chain_run,
 * measured with HIP events on the chain's stream (synchronizes). */
int rt_chain_elapsed_ms(rt_chain* ch, double* ms);

/* Download and merge the colour classes into an N x N tiling (synchronizes). */
int rt_chain_read(rt_chain* ch, int32_t* tiling);

/* Flip attempts per step (Rotate work-items, padding included, as in the reference). */
int64_t rt_chain_attempts_per_step(rt_chain* ch);

/* Bytes of packed tiling state on the device (all colour classes). */
int64_t rt_chain_state_bytes(rt_chain* ch);

/* Kernel launches issued by the most recent rt_chain_run, and the tile configuration it
 * used (-1 for the stepwise path). */
int64_t rt_chain_last_launches(rt_chain* ch);
int32_t rt_chain_last_config(rt_chain* ch);

/* Resident (persistent) grids are launched cooperatively, so the runtime either guarantees
 * that all their workgroups are resident at once or refuses the launch; a refused run
 * continues on the blocked path (same result).  This reports how many launches of this chain
 * were refused and the hipError_t of the last refusal (0 if none).  Diagnostic; no
 * counterpart in the reference. */
int rt_chain_coop_refusals(rt_chain* ch, int64_t* count, int32_t* last_error);

/* Skipped tiles (lozenge chains, packed sweep).  The reference rotates every position of the
 * N x N/3 class arrays (lozengekernel.cl:11-12); positions outside the domain hold code 0,
 * which never flips, and a cell whose neighbourhood is all zero is never rewritten by the
 * update.  When the uploaded tiling is a fixed point of the update (every valid tiling is),
 * tiles whose cells, halo and a two-cell margin are all zero never change, and the packed
 * persistent sweep does not run them: the tiling after any number of steps is bit-identical.
 * The sweep checks while it runs that no change reaches a skipped tile (rt_chain_sync then
 * returns RT_ERR_DEVICE).  No counterpart in the reference.
 * rt_chain_set_skip: enable (default) or disable skipping for this chain.
 * rt_chain_last_grid: tiles of the most recent persistent run's grid and how many ran.
 * rt_lozenge_skip_map: host only (no device): for an N x N lozenge tiling and a tile of
 * tr x tc positions (rows x class columns) with halo h, one byte per tile, row-major, 1 =
 * runs (map may be NULL; every tile runs when the tiling is not a fixed point of the update),
 * the number that run, and whether the tiling is a fixed point. */
int rt_chain_set_skip(rt_chain* ch, int enable);
int rt_chain_last_grid(rt_chain* ch, int32_t* tiles, int32_t* running);
int rt_lozenge_skip_map(const int32_t* tiling, int32_t N, int32_t tr, int32_t tc, int32_t h, uint8_t* map,
                        int32_t* running, int32_t* consistent);

/* ---- observables on the device (no counterpart in the reference) ---------------------------
 * Height-function statistics of domino chains computed where the chains live, without reading
 * the tilings back.  The heights are DominoTiler::TilingToHeightfunc's (src/Domino/
 * DominoTiler.cpp:423-461: 0 at the first boundary vertex in row-major order, +-1 along
 * uncrossed edges) for the domain's tilings.  An observer holds, for one domain, its M x M
 * squares (`domain`), its minimal height function `hmin` ((M+1) x (M+1), INT32_MAX outside,
 * e.g. from rt_extremal_tiling: only its boundary heights are used, which every tiling shares),
 * the vertices the statistics cover (`mask`, NULL: every vertex of the domain) and an optional
 * reference field `ref256` (256 x a height field, e.g. a long-run mean, rounded; NULL: none).
 * rt_observe computes, for each of n domino chains of side M+1 on the observer's context, in
 * one launch (a segmented row scan per row, wave64 reductions): out[4k+0] = masked vertices,
 * out[4k+1] = sum of their heights, out[4k+2] = sum of squares, out[4k+3] = sum of
 * |256 h - ref256| (0 without a reference), all exact integers; with heights != NULL also
 * every chain's height function (n x (M+1)^2, INT32_MAX outside).  Synchronizes. */
int rt_observer_create(rt_ctx* ctx, const int32_t* domain, int32_t M, const int32_t* hmin, const uint8_t* mask,
                       const int32_t* ref256, rt_observer** out);
int rt_observer_destroy(rt_observer* obs);
int rt_observe(rt_observer* obs, rt_chain* const* chains, int32_t n, int64_t* out, int32_t* heights);

/* ---- batches of independent chains (no counterpart in the reference: its callers loop
 * over tilers) ---------------------------------------------------------------------------
 * A batch runs many seeded chains of one context together: the domino, lozenge and
 * rectangle-triangle chains of each model advance in lockstep, one blocked launch per block of
 * h steps covering the tiles of every chain of the model (a chain-index table maps blocks to
 * chains); triangle-dimer chains advance together, one tiled launch per step on 32 x 32 tiles,
 * or on the configuration set on the largest of them (rt_chain_set_path; 3-5: two steps per
 * launch).  With use_graph, one round of 32 steps of every model (its launches at their offsets
 * from a device step counter, then one counter update) is captured as a HIP graph and replayed;
 * the graph is kept by the batch and replayed by later runs as long as its chains, their tile
 * configurations and the parameter table are unchanged (rt_batch_last_graphs is then 0).
 * After rt_batch_run(b, k) every chain is exactly where rt_chain_run(ch, k) would have left it.
 * Chains stay owned by the caller and must outlive the batch. */
int rt_batch_create(rt_ctx* ctx, rt_batch** out);
int rt_batch_destroy(rt_batch* b);
int rt_batch_add(rt_batch* b, rt_chain* ch);
int rt_batch_run(rt_batch* b, int64_t steps, int use_graph);
/* Waits for the batch's runs; RT_ERR_DEVICE if a batched launch raised the device error flag
 * since the last rt_batch_sync (the flag is then cleared). */
int rt_batch_sync(rt_batch* b);
/* Measure every tile configuration that applies to each model of the batch (probe_steps steps
 * of that model's chains alone, after a warm-up) and keep the fastest for later runs of the
 * batch; chosen (may be NULL) receives the choice per model RT_DOMINO .. RT_TRIDIMER (-1: none
 * in the batch).  Every chain is put back exactly as it was (state, generators and host
 * schedule RNGs), so the trajectories do not depend on the probes.  No reference counterpart. */
int rt_batch_autotune(rt_batch* b, int64_t probe_steps, int32_t* chosen);
/* Set the tile configuration later runs of the batch use for `model` (RT_DOMINO .. RT_TRIDIMER;
 * idx -1 returns to the automatic choice), e.g. one an earlier rt_batch_autotune chose for the
 * same chains.  RT_ERR_ARG for an index out of range; a run fails with RT_ERR_ARG when the
 * configuration is a packed one and a chain's rule cannot use it.  No reference counterpart. */
int rt_batch_set_config(rt_batch* b, int model, int32_t idx);
/* GPU milliseconds of the most recent rt_batch_run (HIP events on the context stream). */
int rt_batch_elapsed_ms(rt_batch* b, double* ms);
/* Kernel launches (those of graph replays included) and graphs instantiated by the last run
 * (0 when it replayed the graph kept from an earlier run). */
int64_t rt_batch_last_launches(rt_batch* b);
int32_t rt_batch_last_graphs(rt_batch* b);

/* ---- learned warm start (no counterpart in the reference) --------------------------------
 * A U-Net trained and run on the GPU by this library's kernels: 3 x 3 convolutions as bf16
 * matrix-core products (v_mfma_f32_32x32x8_bf16, fp32 accumulation), ReLU, 2 x 2 average
 * pooling, nearest upsampling, additive skips; `levels` levels of widths base * 2^l, in_ch
 * input channels, one output channel.  Images are NHWC fp32 host arrays (B x H x W x in_ch
 * in, B x H x W out), H and W divisible by 2^(levels-1).  Parameters form one flat fp32
 * vector in layer order (encoder levels 0..L-1 two convs each, decoder levels L-2..0 an
 * up-conv and a conv each, the output conv; each conv W [9 cin x cout] with
 * W[(tap*cin + ci)*cout + co], tap = 3*dy + dx, then its bias). */
int rt_unet_create(rt_ctx* ctx, int levels, int base, int in_ch, int64_t seed, rt_unet** out);
int rt_unet_destroy(rt_unet* net);
int64_t rt_unet_param_count(rt_unet* net);
int rt_unet_get_params(rt_unet* net, float* out);
int rt_unet_set_params(rt_unet* net, const float* in);
int rt_unet_forward(rt_unet* net, const float* x, int B, int H, int W, float* y);
/* Masked mean squared error over pixels with a nonzero mask, and (grad != NULL) its gradient
 * with respect to every parameter. */
int rt_unet_loss_grad(rt_unet* net, const float* x, const float* target, const float* mask, int B, int H, int W,
                      float* loss, float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on a batch; *loss before the step. */
int rt_unet_train_step(rt_unet* net, const float* x, const float* target, const float* mask, int B, int H, int W,
                       float lr, float* loss);
/* C = A B through the bf16 matrix-core kernel (A: M x K, B: K x N row-major fp32, rounded to
 * bf16; C fp32): the product on its own, for tests. */
int rt_nn_gemm(rt_ctx* ctx, int M, int N, int K, const float* A, const float* B, float* C);
/* The same with the operands stored as `layout` says (0: A [M x K] and B [K x N] row-major;
 * 1: A stored as [K x M]; 2: B stored as [N x K]). */
int rt_nn_gemm_ex(rt_ctx* ctx, int M, int N, int K, int layout, const float* A, const float* B, float* C);
/* Device-resident timing of that product for one operand layout (0: A [M x K] and B [K x N]
 * row-major, a convolution's forward product; 1: A stored as [K x M], the weight gradient;
 * 2: B stored as [N x K], the input gradient): *ms = milliseconds per product over reps. */
int rt_nn_gemm_time(rt_ctx* ctx, int M, int N, int K, int layout, int reps, double* ms);

/* Workgroups a persistent launch of tile configuration idx of `model` may use (at most one per
 * CU; uniform_params: the single-parameter-set kernel), 0 when the configuration never runs
 * persistently (its kernel needs scratch memory), -1 for a bad index.  Larger tile grids run
 * one launch per block of steps.  Diagnostic; no counterpart in the reference. */
int32_t rt_persist_capacity(int model, int32_t idx, int device, int uniform_params);

/* Persistent tile configurations compiled for `model` (no device needed): count, and the
 * rows, columns, halo width and threads per workgroup of configuration idx. */
int32_t rt_tile_config_count(int model);
int rt_tile_config(int model, int32_t idx, int32_t* rows, int32_t* cols, int32_t* halo, int32_t* threads);
/* 1 when tile configuration idx of the model runs the specialised sweep: for lozenge and
 * domino chains the packed sweep (four positions per 32-bit word; dominoes when the rule
 * splits k at one threshold per class and row parity or is the lazy toggle), for
 * rectangle-triangle chains the row-wave sweep (extended tile rows of 64 positions, one
 * wavefront per row); 0 otherwise.  No counterpart in the reference. */
int32_t rt_tile_config_packed(int model, int32_t idx);

/* Number of persistent tile configurations compiled for this chain's model. */
int32_t rt_chain_tile_configs(rt_chain* ch);

/* Describe tile configuration `idx`: rows, cols, halo, threads per workgroup. */
int rt_chain_tile_config_info(rt_chain* ch, int32_t idx, int32_t* rows, int32_t* cols, int32_t* halo,
                              int32_t* threads);

#ifdef __cplusplus
}
#endif

#endif /* RT_HIP_H_ */
