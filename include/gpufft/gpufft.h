/*
 * gpufft.h -- C ABI of the MI355X (gfx950) 3D R2C/C2R FFT engine.
 *
 * This is the drop-in boundary that replaces the reference's hot path:
 *   INIT_GPU_FFT<T1,T2>(Nx,Ny,Nz,procs,rank,comm)
 *        GPU_FFT/src/GPU_FFT/GPU_FFT.h:114-115, GPU_FFT_init.cu:41-93
 *   GPU_FFT_R2C<T1,T2>(T1*)   GPU_FFT.h:117-118, FFT_R2C.cu:41-70
 *   GPU_FFT_C2R<T1,T2>(T2*)   GPU_FFT.h:120-121, FFT_C2R.cu:41-70
 * The reference exposes only C++ templates over process-global state
 * (GPU_FFT.cu:43-87); this ABI is handle-based, stream-ordered and returns
 * status codes.  include/GPU_FFT/GPU_FFT.h rebuilds the reference templates on
 * top of it.
 *
 * Data layout (identical to the reference at p = 1, Readme.md:47-52):
 *   real input   (Nx/p, Ny, 2K) reals, row-major, K = Nz/2 + 1 (integer
 *                division), the last 2K - Nz reals of every row are padding;
 *   complex out  (Nx, Ny/p, K) complex (interleaved re, im), kz fastest,
 *                in the same bytes (in place).
 * Forward uses exp(-2 pi i jk/N); neither direction normalises, so
 * C2R(R2C(f)) = Nx*Ny*Nz*f.  C2R overwrites its input; its pad slots are
 * unspecified.  C2R ignores the imaginary parts of the kz = 0 and (even Nz)
 * kz = Nz/2 bins after the x and y passes (numpy irfftn convention).
 *
 * All entry points are thread-safe for distinct plans.  Execution is
 * asynchronous on the given HIP stream (NULL = the null stream), exactly like
 * the reference, which returns with the last FFT still in flight
 * (FFT_R2C.cu:69, caller syncs at GPU_FFT_TEST_SCRIPT/test.cu:146).
 */
#ifndef GPUFFT_GPUFFT_H_
#define GPUFFT_GPUFFT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFFT_VERSION 100 /* 1.0.0 */

/* Compatible with hipStream_t (which is `struct ihipStream_t*`). */
typedef struct ihipStream_t* gfft_stream_t;

typedef struct gfft_plan_s* gfft_plan;

typedef enum {
  GFFT_F32 = 0, /* float  / float2   (reference T1_f / T2_f, transitions.h:77-78) */
  GFFT_F64 = 1  /* double / double2  (reference T1_d / T2_d, transitions.h:79-80) */
} gfft_prec;

typedef enum {
  GFFT_OK = 0,
  GFFT_ERR_INVALID_ARG = 1,       /* NULL handle/pointer, data not aligned to a complex element (8 B FP32, 16 B FP64), bad flag */
  GFFT_ERR_INVALID_DIMS = 2,      /* a dimension < 1, p does not divide Nx and Ny, sizes overflow */
  GFFT_ERR_UNSUPPORTED_SIZE = 3,  /* an edge has a prime factor > 13 or exceeds the LDS-resident limit */
  GFFT_ERR_UNSUPPORTED_PROCS = 4, /* full transform requested with procs > 1 (needs an exchange) */
  GFFT_ERR_INVALID_RANK = 5,      /* rank outside [0, procs) */
  GFFT_ERR_OUT_OF_MEMORY = 6,     /* device allocation failed */
  GFFT_ERR_HIP = 7,               /* a HIP runtime call or kernel launch failed */
  GFFT_ERR_DEVICE_MISMATCH = 8,   /* executed with another current device than the plan's */
  GFFT_ERR_BUFFER_TOO_SMALL = 9,  /* text query buffer too small */
  GFFT_ERR_OVER_BUDGET = 10,      /* serving: one request needs more HBM than the whole budget */
  GFFT_ERR_TIMEOUT = 11,          /* serving: gfft_server_wait ran out of time; batch: a persistent wait hit its bound */
  GFFT_ERR_ABORTED = 12,          /* batch: a persistent launch gave up its waits on gfft_batch_set_abort */
  GFFT_ERR_SCRATCH_TOO_SMALL = 13 /* batch: gfft_batch_exec_noalloc's plan exceeds the static counter array */
} gfft_status;

/* Version number (GFFT_VERSION of the built library). */
int gfft_version(void);

/* Static description of a status code. */
const char* gfft_strerror(int status);

/* Plan a transform of an Nx x Ny x Nz grid on the current HIP device.
 * Replaces INIT_GPU_FFT (GPU_FFT_init.cu:41-93): validates the geometry,
 * picks a kernel per axis, uploads exact twiddle tables (computed in extended
 * precision on the host, rounded once).  Unlike the reference it allocates
 * no work buffer: every pass is in place.  procs > 1 plans the local share of
 * a slab decomposition (see gfft_exec_r2c_local).  Where the process-wide
 * tune cache holds choices for this geometry on this device type (an earlier
 * gfft_plan_tune, or gfft_tune_cache_import), the plan starts with them;
 * otherwise with the default variants.  Serving and batched plans do the
 * same. */
int gfft_plan_create(gfft_plan* plan, int64_t nx, int64_t ny, int64_t nz, int procs, int rank, int prec);

/* Release a plan and its device tables (the reference never frees). */
int gfft_plan_destroy(gfft_plan plan);

/* In-place forward 3D R2C (replaces GPU_FFT_R2C, FFT_R2C.cu:41-70). */
int gfft_exec_r2c(gfft_plan plan, void* data, gfft_stream_t stream);

/* In-place inverse 3D C2R, unnormalised (replaces GPU_FFT_C2R,
 * FFT_C2R.cu:41-70). */
int gfft_exec_c2r(gfft_plan plan, void* data, gfft_stream_t stream);

/* The local halves of a slab-decomposed transform (any procs >= 1), for a
 * caller that performs the inter-rank exchange itself (MPI, RCCL, or copies).
 * Exchange buffers hold p chunks of (Ny/p, Nx/p, K) complex values, chunk s
 * for (send) or from (receive) rank s -- the reference's layout around its
 * MPI block (FFT_R2C.cu:51-63, FFT_C2R.cu:51-63).  All buffers are
 * gfft_plan_data_bytes long and must not alias.
 *   R2C = gfft_exec_r2c_local -> all-to-all of chunks -> gfft_exec_r2c_finish
 *   C2R = gfft_exec_c2r_start -> all-to-all of chunks -> gfft_exec_c2r_local */

/* z R2C + y C2C on the local x-slab (in place in `data`, which is then
 * scratch), with the slab transpose fused into the y pass's stores into
 * `sendbuf`.  Replaces FFT_R2C.cu:45-48 (hipFFT R2C + __transpose_slab__). */
int gfft_exec_r2c_local(gfft_plan plan, void* data, void* sendbuf, gfft_stream_t stream);

/* x C2C reading the received chunks (the chunk transpose fused into its
 * loads), writing the (Nx, Ny/p, K) y-slab to `out`.  Replaces
 * FFT_R2C.cu:66-69 (__chunk_transpose__ + hipFFT C2C forward). */
int gfft_exec_r2c_finish(gfft_plan plan, const void* recvbuf, void* out, gfft_stream_t stream);

/* x C2C inverse of the y-slab `data`, packing into p chunks in `sendbuf`.
 * Replaces FFT_C2R.cu:45-48 (hipFFT C2C inverse + __chunk_transpose_inverse__). */
int gfft_exec_c2r_start(gfft_plan plan, const void* data, void* sendbuf, gfft_stream_t stream);

/* y C2C inverse reading the received slab-transposed layout, then z C2R in
 * place, leaving the real x-slab in `data`.  Replaces FFT_C2R.cu:66-69
 * (__transpose_slab__ + hipFFT C2R). */
int gfft_exec_c2r_local(gfft_plan plan, const void* recvbuf, void* data, gfft_stream_t stream);

/* Run one axis pass of the plan in place (axis 0 = z rows, 1 = y, 2 = x;
 * inverse 0 = the R2C direction, 1 = the C2R direction), for per-pass
 * measurement and for callers that interleave their own work between passes.
 * R2C = passes (z,0), (y,0), (x,0); C2R = (x,1), (y,1), (z,1).  Axis 3 is the
 * fused z + y plane pass (R2C: z then y; C2R: y then z), valid only while the
 * plan uses it (describe: "zy_fused"); then R2C = (3,0), (x,0) and C2R =
 * (x,1), (3,1).  A length-1 strided axis is a no-op.  No counterpart in the
 * reference (its passes are hipFFT plans, GPU_FFT_init.cu:86-91). */
int gfft_exec_axis(gfft_plan plan, int axis, int inverse, void* data, gfft_stream_t stream);

/* Autotuner: time every compiled variant of every axis pass (tile width,
 * points per thread, rows per workgroup, and the run-time-schedule kernel)
 * on `scratch` (>= gfft_plan_data_bytes bytes, contents destroyed), `reps`
 * forward+inverse pairs each, and keep the fastest.  Choices are cached per
 * geometry for the life of the process.  Synchronises `stream`.  The
 * reference delegates this choice to hipFFT's planner (GPU_FFT_init.cu:86-91). */
int gfft_plan_tune(gfft_plan plan, void* scratch, gfft_stream_t stream, int reps);

/* Select variant `variant` (0 = default; gfft_plan_describe lists the
 * count as "candidates") for axis 0 = z, 1 = y, 2 = x. */
int gfft_plan_set_variant(gfft_plan plan, int axis, int variant);

/* Cache blocking of the z and y passes: with planes > 0 (and < Nx/p) the
 * R2C runs z then y on each block of `planes` consecutive x-planes before
 * the x pass, and the C2R runs the x pass, then y then z per block, so a
 * block's intermediate can be served from the memory-side cache.  0 (the
 * default) runs whole-array passes.  Same results either way. */
int gfft_plan_set_chunking(gfft_plan plan, int64_t planes);

/* Tile order of the passes: 0 = every pass walks its tiles first to last;
 * 1 (the default) = alternate, so that each pass of a round trip starts on
 * the tiles the previous pass wrote last (R2C: z forward, y reversed, x
 * forward; C2R: x reversed, y forward, z reversed), which are still in the
 * memory-side cache.  Bitwise the same results either way. */
int gfft_plan_set_walk(gfft_plan plan, int mode);

/* Per-pass tile order: bit (3 * direction + axis) set = that pass walks its
 * tiles last to first (direction 0 = R2C, 1 = C2R; axis 0 = z, 1 = y,
 * 2 = x), mask in [0, 63].  gfft_plan_set_walk(plan, 1) is mask 42 and
 * mode 0 is mask 0.  Bitwise the same results for every mask.  No reference
 * counterpart (tile order is internal to its library calls). */
int gfft_plan_set_walk_mask(gfft_plan plan, int mask);

/* Order of the two strided passes of full transforms: bit 0 set = the R2C
 * runs x before y (z, x, y), bit 1 set = the C2R runs y before x (y, x, z);
 * 0 (the default) is the reference's order (FFT_R2C.cu:45-69 z+y then x,
 * FFT_C2R.cu:45-69 x then y+z).  The axes are separable, so every order
 * computes the same transform; results agree to rounding.  The local
 * halves of a slab exchange keep the reference's order. */
int gfft_plan_set_pass_order(gfft_plan plan, int order);

/* Persistent autotuning: the process-wide cache that gfft_plan_tune fills
 * and consults, as text, one line per tuned geometry:
 *   "<gfx arch>/<CUs>|<pass geometry> <chosen kernel description>"
 * Export (buf = NULL queries the size into *needed, NUL included; a short
 * buffer gives GFFT_ERR_BUFFER_TOO_SMALL), import (adds or replaces entries;
 * '#' lines are comments; a malformed line rejects the whole text with
 * GFFT_ERR_INVALID_ARG), clear.  Entries of another device type or of a
 * kernel the library does not have are ignored at lookup (the plan is then
 * tuned again).  No counterpart in the reference (it has no tuning). */
int gfft_tune_cache_export(char* buf, size_t len, size_t* needed);
int gfft_tune_cache_import(const char* text);
int gfft_tune_cache_clear(void);

/* Fused z + y kernels for grids whose x-planes fit in on-chip memory (small
 * Ny x Nz, see gfft_plan_describe's "zy_plane_kernel"): each x-plane makes one
 * HBM round trip for both axes instead of two.  1 (the default where such a
 * kernel is compiled) = use them, 0 = separate passes; gfft_plan_tune keeps
 * the faster.  Results agree to rounding either way.  Returns
 * GFFT_ERR_INVALID_ARG for 1 where no fused kernel exists.  No reference
 * counterpart (the reference always runs separate library passes). */
int gfft_plan_set_fused(gfft_plan plan, int on);

/* Output scale factors fused into the last pass of each direction (default
 * 1.0 and 1.0: unnormalised, the reference's semantics).  c2r_scale =
 * 1/(Nx*Ny*Nz) gives the normalised inverse that the reference harness
 * obtains with a separate kernel (GPU_FFT_TEST_SCRIPT/test.cu:53-65,149) at
 * no extra memory traffic.  Not thread-safe against concurrent execs of the
 * same plan. */
int gfft_plan_set_scale(gfft_plan plan, double r2c_scale, double c2r_scale);

/* Device bytes held by the plan (twiddle tables), for admission control. */
int gfft_plan_device_bytes(gfft_plan plan, size_t* bytes);

/* Bytes of the caller's in-place array for this plan's local share:
 * (Nx/p) * Ny * K complex values. */
int gfft_plan_data_bytes(gfft_plan plan, size_t* bytes);

/* Text (JSON) description of the passes the plan will launch. */
int gfft_plan_describe(gfft_plan plan, char* buf, size_t len);

/* Host-only planning of a geometry (no device work): same validation and
 * kernel choice as gfft_plan_create, written as JSON into buf.  Returns the
 * status gfft_plan_create would return for a geometry problem. */
int gfft_query_plan(int64_t nx, int64_t ny, int64_t nz, int procs, int rank, int prec, char* buf, size_t len);

/* ---------------------------------------------------------------------------
 * Batched mode: many independent in-place transforms of uneven sizes run as
 * one unit of work, eagerly or as one hipGraph (no counterpart in the
 * reference, which plans and runs one grid per process).  FP32; all edges of
 * a transform are one size of {32, 48, 64, 96, 128, 160, 192, 256} (Nx and
 * Ny may also be 1).  Each transform uses the layout of gfft_exec_r2c on its
 * own array.
 *
 * Two schedulers.  GFFT_BATCH_PERSISTENT (default): one persistent launch
 * runs every axis pass of every transform, one workgroup per slot of a plan
 * made at creation: small transforms are run whole by one workgroup each
 * (several per workgroup, packed longest first), larger ones by a team of
 * workgroups whose members own interleaved tile shares of each pass (the
 * largest by a team spanning XCDs, see gfft_batch_set_span); a workgroup
 * with nothing left helps its XCD's teams.  A
 * team member waits only for the shares of its own transform's pass (per
 * share state words, no grid-wide barrier) and takes over shares whose owner
 * has not started, so completion never depends on how many workgroups are
 * resident.  Every wait is bounded (gfft_batch_set_wait_limit) and gives up
 * when the host sets the abort word (gfft_batch_set_abort); the launch then
 * drains and reports through gfft_batch_status.  The last workgroup to leave
 * resets the counters, so launches and graph replays need no clearing
 * between them.  Executions of one batch must not overlap (they share the
 * counters).
 * GFFT_BATCH_LEVELS: one launch per (size, axis pass) covering the tiles of
 * every transform of that size, the sizes' chains on internal streams.
 * ------------------------------------------------------------------------- */
typedef struct gfft_batch_s* gfft_batch;

#define GFFT_BATCH_PERSISTENT 0
#define GFFT_BATCH_LEVELS 1

/* GFFT_OK if (nx, ny, nz, prec) is supported by the batched kernels. */
int gfft_batch_query(int64_t nx, int64_t ny, int64_t nz, int prec);

/* dims: count triples (nx, ny, nz).  Allocates descriptors, tile tables,
 * twiddles and internal streams. */
int gfft_batch_create(gfft_batch* batch, int count, const int64_t* dims, int prec);

/* Device pointers of the count in-place arrays (host array of pointers).
 * Synchronises the device first (earlier launches read the old pointers). */
int gfft_batch_set_data(gfft_batch batch, void* const* ptrs);

/* Per-transform output scales (NULL = all 1.0), as gfft_plan_set_scale.
 * Synchronises the device first. */
int gfft_batch_set_scales(gfft_batch batch, const double* r2c_scale, const double* c2r_scale);

/* mode 0 = R2C of every transform, 1 = C2R, 2 = R2C then C2R, on `stream`
 * (persistent: one launch; levels: the sizes' launches run on internal
 * streams forked from and joined back into `stream`).  Work queued on
 * `stream` later starts after the batch. */
int gfft_batch_exec(gfft_batch batch, int mode, gfft_stream_t stream);

/* No counterpart in the reference; for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid it).  The persistent scheduler's execution of `mode` (whatever
 * the batch's current scheduler), with its counter words in a static array
 * of 65536 words that the library holds on each GPU until the process ends
 * (device address taken once per device with hipGetSymbolAddress; zeroed by a
 * fill kernel on the first call on a device).  Calls no allocating routine.
 * A plan needing more than 65524 share words returns
 * GFFT_ERR_SCRATCH_TOO_SMALL.  Every call on one GPU shares the array:
 * calls that overlap on one GPU (other streams, other batches) race.
 * Failures are reported by gfft_batch_status like those of gfft_batch_exec. */
int gfft_batch_exec_noalloc(gfft_batch batch, int mode, gfft_stream_t stream);

/* Same work as one hipGraph launch on `stream` (built on first use per
 * mode: the persistent kernel as its one node, or one branch of level
 * launches per size).  The graph reads pointers and scales from device
 * memory, so later gfft_batch_set_data / _set_scales calls take effect in
 * replays. */
int gfft_batch_exec_graph(gfft_batch batch, int mode, gfft_stream_t stream);

/* Autotuner of the batched kernels.  Tunes both schedulers, then keeps the
 * one whose tuned execution of `mode` is faster (gfft_batch_scheduler tells
 * which; gfft_batch_tuned_ms gives both times).  Persistent: every variant
 * set (one compiled variant per size: radix order and points per thread,
 * which sets the pencils per 256-thread tile) and team factor in {1, 2, 4},
 * each timed as `reps` eager launches.  Levels: every size's variants and
 * the per-transform-plan choice, each size's level chain timed alone.  Runs
 * on the arrays set by gfft_batch_set_data (contents destroyed).
 * Synchronises `stream`; later graph launches are rebuilt. */
int gfft_batch_tune(gfft_batch batch, int mode, int reps, gfft_stream_t stream);

/* Level scheduler: select variant `variant` (0 = default) of the size class
 * with edge `edge`; query the current choice.  set synchronises the device. */
int gfft_batch_set_variant(gfft_batch batch, int64_t edge, int variant);
int gfft_batch_variant(gfft_batch batch, int64_t edge, int* variant);

int gfft_batch_data_bytes(gfft_batch batch, int index, size_t* bytes);

/* Tiles one execution of `mode` processes, and (if launches is not NULL)
 * the number of kernel launches it takes. */
int gfft_batch_tiles(gfft_batch batch, int mode, uint32_t* tiles, uint32_t* launches);

/* Scheduler: GFFT_BATCH_PERSISTENT (default) or GFFT_BATCH_LEVELS. */
int gfft_batch_set_scheduler(gfft_batch batch, int scheduler);
int gfft_batch_scheduler(gfft_batch batch, int* scheduler);

/* Persistent scheduler: variant set (0..1) and team factor (1..64: a
 * transform gets a team of workgroups in one XCD when its work exceeds that
 * many slot shares, a slot share being the batch's work divided by the
 * slots; 0: global-level plan, every workgroup takes shares of every
 * transform's axis pass, one hand-over across all XCDs per pass, the
 * smallest transforms run whole from a queue meanwhile); query them and (if
 * variants is not NULL) the set's variant per size class, 8 ints for the
 * sizes 32 ... 256.  set rebuilds the plans and synchronises the device.
 * gfft_batch_tune tries team factors 0, 1, 2 and 4 with both sets, and
 * factors 1 and 2 with teams kept to one XCD (span 1). */
int gfft_batch_set_persist(gfft_batch batch, int set, int team_factor);
int gfft_batch_persist(gfft_batch batch, int* set, int* team_factor, int* variants);

/* Team plans: the most XCDs one team may span (1..8, default 2).  A
 * transform whose share of workgroups exceeds one XCD's slots gets a team
 * split over that many XCDs, whose members hand each axis pass over through
 * an L2 write-back per XCD part.  set rebuilds the plans and synchronises
 * the device. */
int gfft_batch_set_span(gfft_batch batch, int parts);
int gfft_batch_span(gfft_batch batch, int* parts);

/* The persistent plan of `mode`: slots (workgroups of the launch), teams
 * (transforms run by two or more workgroups) and their members (global-level
 * plans: no teams, every slot a member). */
int gfft_batch_plan_info(gfft_batch batch, int mode, uint32_t* slots, uint32_t* teams, uint32_t* members);

/* ms per eager execution of the tuned mode measured by the last
 * gfft_batch_tune for each scheduler (-1 before any tune). */
int gfft_batch_tuned_ms(gfft_batch batch, float* persistent_ms, float* levels_ms);

/* Bound of one wait of the persistent kernel, seconds of device wall clock
 * (default 2).  A wait that reaches it marks the launch failed. */
int gfft_batch_set_wait_limit(gfft_batch batch, double seconds);

/* Nonzero: every wait of the batch's persistent launches, running or later,
 * gives up at once (the launch drains without computing the waiting work and
 * reports GFFT_ERR_ABORTED); 0 clears the request.  Host memory only, no
 * synchronisation, callable while a launch runs. */
int gfft_batch_set_abort(gfft_batch batch, int abort);

/* Synchronises the device, then reports and clears the sticky outcome of
 * the persistent launches since the last call: GFFT_OK, GFFT_ERR_TIMEOUT (a
 * wait reached its bound) or GFFT_ERR_ABORTED.  Their results are undefined
 * unless it is GFFT_OK. */
int gfft_batch_status(gfft_batch batch, int* status);

int gfft_batch_destroy(gfft_batch batch);

/* ---------------------------------------------------------------------------
 * Serving runtime (no counterpart in the reference): R2C / C2R requests of
 * mixed sizes and precisions arrive one by one and are executed as they are
 * admitted.  Plans and twiddle tables stay resident per (Nx, Ny, Nz, prec).
 * Admission is in arrival order (one FIFO for requests at or under the
 * small-lane threshold, one for larger ones) under an HBM byte budget: a
 * FIFO's head starts once its array fits beside everything in flight, a
 * stream it may use has fewer than the stream depth of unfinished requests,
 * and, if host-staged, the server's arena has a free block for it.  Admitted
 * requests are spread over `nstreams` internal streams; one waiter thread
 * per stream waits for its oldest request's completion event (timed polling
 * or a blocking synchronise) and stamps the completion.  Every entry point
 * is thread-safe.
 * ------------------------------------------------------------------------- */
typedef struct gfft_server_s* gfft_server;

#define GFFT_SERVER_STAGING 1 /* create: allocate a budget-sized arena for host-staged requests */
#define GFFT_SERVER_SHARED_QUEUES 2 /* create: plain non-blocking streams from the runtime's queue pool */
#define GFFT_REQ_DEVICE 1     /* request: data is device memory, transformed in place */
#define GFFT_REQ_PRODUCER 2   /* request: start only after the work queued on `producer` before submit */

typedef struct {
  int64_t nx, ny, nz;
  int prec;       /* gfft_prec */
  int direction;  /* 0 = R2C, 1 = C2R */
  int flags;      /* GFFT_REQ_DEVICE, or 0 = host-staged through the arena */
  int reserved;   /* 0 */
  void* data;     /* device: the in-place array; host-staged: host input (pinned for overlap) */
  void* host_out; /* host-staged: host output, may equal data; ignored for device requests */
  double scale;   /* output scale fused into the last pass; 0 means 1 (unnormalised) */
  double arrival; /* arrival time on gfft_server_clock(); < 0 = at submission */
  uint64_t tag;   /* returned unchanged in the result */
  gfft_stream_t producer; /* with GFFT_REQ_PRODUCER: the stream that writes `data` (an event is recorded
                           * on it at submission and the server's stream waits for that event; for the
                           * null stream with the default blocking server streams the ordering is
                           * implicit and no event is recorded) */
} gfft_request;

typedef struct {
  uint64_t id, tag;
  int status;                     /* GFFT_OK, or why this request failed */
  int stream;                     /* internal stream it ran on, -1 if it never ran */
  double arrival, admitted, done; /* gfft_server_clock() seconds */
  double gpu_ms;                  /* transform kernels only (device events) */
  double staged_ms;               /* copy in + transform + copy out (device events); = gpu_ms for device */
  size_t bytes;                   /* array bytes counted against the budget */
  double started;                 /* when the device began the request's first operation:
                                   * done - staged_ms (started - admitted = time in the stream's queue) */
} gfft_result;

typedef struct {
  uint64_t submitted, completed, rejected, waiting, running, plans;
  size_t budget, in_flight_bytes, peak_in_flight_bytes, arena_bytes, arena_used, plan_bytes;
} gfft_server_stats_t;

/* Monotonic clock (seconds) used for arrival / admission / completion. */
double gfft_server_clock(void);

/* Bytes of a request's in-place array, Nx*Ny*2K reals (host-only). */
int gfft_request_bytes(int64_t nx, int64_t ny, int64_t nz, int prec, size_t* bytes);

/* budget_bytes bounds the arrays of admitted, unfinished requests; flags:
 * GFFT_SERVER_STAGING, GFFT_SERVER_SHARED_QUEUES.  Uses the current HIP
 * device.  By default each of the nstreams streams is created with a CU mask
 * of every CU, which gives it a hardware queue of its own (the streams then
 * synchronise with the legacy null stream, like any blocking stream); with
 * GFFT_SERVER_SHARED_QUEUES they are plain non-blocking streams, which the
 * runtime may map several to one hardware queue.  nstreams = 3 (two streams
 * for large requests and the small-request lane) measured the lowest p99
 * latency on a mixed-size Poisson load at 80 % of capacity; more streams
 * share the GPU among more large requests and stretch each one's service. */
int gfft_server_create(gfft_server* srv, size_t budget_bytes, int nstreams, int flags);

/* Reserve the last stream for requests of at most max_bytes (array bytes):
 * larger requests never queue there, so a small request is never stuck
 * behind a large one on every stream.  Default with nstreams >= 2: 32 MiB;
 * 0 turns the reservation off (every request goes to the stream with the
 * least outstanding bytes).  Applies to requests submitted afterwards. */
int gfft_server_set_small_lane(gfft_server srv, size_t max_bytes);

/* At most `depth` admitted, unfinished requests per stream (default 1: a
 * request is bound to a stream once that stream is free); further requests
 * wait in the server's FIFO, where their time shows as admitted - arrival.
 * 0 = no limit.  Applies to admissions from now on. */
int gfft_server_set_stream_depth(gfft_server srv, int depth);

/* How a stream's waiter waits for its oldest request: sleeping
 * `microseconds` between event queries (default 20), or with 0 blocking in
 * hipEventSynchronize (no polling; wake-ups come later, so the streams may
 * idle between requests). */
int gfft_server_set_poll_interval(gfft_server srv, int microseconds);

/* Build and keep the plan of a geometry now (otherwise built on first use,
 * which allocates and synchronises). */
int gfft_server_prepare(gfft_server srv, int64_t nx, int64_t ny, int64_t nz, int prec);

/* Queue a request.  Invalid geometry, sizes without a plan
 * (GFFT_ERR_UNSUPPORTED_SIZE), GFFT_ERR_OVER_BUDGET (the array alone exceeds
 * the budget, or for a host-staged request the arena) and bad pointers are
 * rejected here, nothing queued.  A device array must be complete when the
 * server's stream reaches it: either its producer's work is finished before
 * submit, or the request names the producer stream (GFFT_REQ_PRODUCER).
 * Requests at or under the small-lane threshold and larger ones are admitted
 * from two FIFOs, each in arrival order, so a small request never waits
 * behind a large one. */
int gfft_server_submit(gfft_server srv, const gfft_request* req, uint64_t* id);

/* Non-blocking: move up to max finished results (completion order) to out. */
int gfft_server_poll(gfft_server srv, gfft_result* out, int max, int* count);

/* Block until nothing is waiting or running; GFFT_ERR_TIMEOUT after timeout_s
 * seconds (<= 0: no limit). */
int gfft_server_wait(gfft_server srv, double timeout_s);

int gfft_server_stats(gfft_server srv, gfft_server_stats_t* st);

/* Waits for everything queued, then frees the server. */
int gfft_server_destroy(gfft_server srv);

#ifdef __cplusplus
}
#endif

#endif /* GPUFFT_GPUFFT_H_ */
