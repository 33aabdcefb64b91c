/* spectrapack_hip.h — C ABI of the MI355X spectral placement step.
 *
 * Plain C: POD structs, pointers and sizes, int status codes; no exceptions
 * cross this boundary.  Status codes map one-to-one onto the reference's
 * failure code strings (sp_hip_status_code returns the string).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   engine/pack_compute/include/spectrapack/compute/correlation.hpp:17-117
 *     CorrelationSpec, CorrelationLimits, CorrelationStats, NumericReport,
 *     CorrelationFailure/Result, FftExecution, estimate_correlation_cpu,
 *     correlate_binary_cpu, correlate_proximity_cpu.
 *   engine/pack_solver/src/gap_fill.cpp:363-509 (per-rotation free masks and
 *     the lowest (z, y, x) choice) — the batched placement step.
 *   engine/pack_geometry/include/spectrapack/geometry/conservative_fields.hpp:100-164
 *     voxelize_object / voxelize_container — the rasteriser.
 *   engine/pack_solver/src/orientation_cube.hpp:10-80, orientations.cpp:44-172 —
 *     orientation catalogs.
 *   engine/pack_solver/src/field_proximity.hpp:32-132, spectral_pipeline.hpp:26-51,
 *     spectral_pipeline.cpp:1740-1857, spectral.cpp:1485-1590 — proximity field,
 *     ranked candidate pages and the greedy loop over them.
 */
#ifndef SPECTRAPACK_HIP_H
#define SPECTRAPACK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (correlation.hpp:52-56 CorrelationFailure::code values) */
enum sp_status {
    SP_OK = 0,
    SP_CORRELATION_MEMORY_LIMIT = 1,
    SP_CORRELATION_SHAPE_OVERFLOW = 2,
    SP_CORRELATION_INDEX_OVERFLOW = 3,
    SP_CORRELATION_INVALID_FIELD = 4,
    SP_CORRELATION_INVALID_ENVIRONMENT = 5,
    SP_CORRELATION_INVALID_KERNEL = 6,
    SP_CORRELATION_NUMERIC_RANGE = 7,
    SP_CORRELATION_PADDED_CELL_LIMIT = 8,
    SP_CORRELATION_NUMERIC = 9,
    SP_CORRELATION_DIRECT_TERM_LIMIT = 10,
    SP_CORRELATION_ALLOCATION = 11,
    SP_CORRELATION_OPERATION = 12,
    SP_CORRELATION_FLOATING_ENVIRONMENT = 13,
    SP_CORRELATION_EXECUTION_CLOSED = 14,
    SP_CPU_THREAD_COUNT_UNSUPPORTED = 15,
    SP_FFT_THREAD_START_FAILURE = 16,
    SP_OPERATION_CANCELLED = 17,
    SP_DEADLINE_EXCEEDED = 18,
    /* Codes of the GPU build itself (no reference counterpart). */
    SP_DEVICE_ERROR = 32,     /* a HIP runtime call failed; see sp_hip_last_error */
    SP_INVALID_ARGUMENT = 33, /* null pointer, bad handle or unsupported option */
    SP_FIELD_INPUT = 34,      /* conservative_fields.cpp FIELD_INPUT */
    SP_FIELD_INDEX_OVERFLOW = 35,
    SP_FFT_LENGTH_UNSUPPORTED = 36, /* a padded length above the kernels' maximum */
    SP_CAPACITY = 37,         /* caller buffer too small; required size reported */
    SP_SCRATCH_LIMIT = 38,    /* the allocation-free correlation needs more than its static scratch */
    SP_FIELD_COPY_ID = 39,    /* conservative_fields.cpp FIELD_COPY_ID: no such (removable) copy */
    SP_FIELD_COUNT_UNDERFLOW = 40 /* conservative_fields.cpp FIELD_COUNT_UNDERFLOW */
};

/* Returns the reference's code string for a status ("CORRELATION_NUMERIC", ...). */
const char* sp_hip_status_code(int status);

/* ---- context: correlation.hpp:78-98 FftExecution (one coordinator owns it and
 * serialises calls).  Owns a HIP stream, device buffers and FFT tables. */
typedef struct sp_hip_ctx sp_hip_ctx;
int sp_hip_ctx_create(int device, uint64_t bytes_cap, sp_hip_ctx** out);
void sp_hip_ctx_destroy(sp_hip_ctx* ctx);
const char* sp_hip_last_error(const sp_hip_ctx* ctx);

/* correlation.hpp:22-27 (X-fastest storage). */
typedef struct {
    uint32_t environment_shape[3];
    uint32_t kernel_shape[3];
    int64_t environment_first[3];
    int64_t kernel_first[3];
} sp_correlation_spec;

/* correlation.hpp:29-36, plus the bytes of the caller's result objects that
 * the reference counts in its admission (correlation.cpp:67-72: two
 * CorrelationOutcome, one vector<double>, one OperationControl; the C++
 * mirror fills it from its own types, C callers state their own). */
typedef struct {
    uint64_t max_working_bytes;
    uint64_t reserved_bytes;
    uint64_t max_padded_cells;
    uint64_t max_direct_terms;
    uint64_t metadata_bytes;
} sp_correlation_limits;

/* correlation.hpp:38-44 */
typedef struct {
    uint64_t padded_cells;
    uint64_t working_bytes_peak;
    uint64_t direct_terms;
} sp_correlation_stats;

/* correlation.hpp:46-50, plus the precision the GPU used and its bound. */
typedef struct {
    double max_integer_residual;
    double mass_residual;
    double max_probe_error;
    double error_bound;  /* a-priori rounding bound of the chosen precision */
    int32_t precision;   /* 32 or 64 */
    uint32_t fft_shape[3];
} sp_numeric_report;

/* correlation.hpp:100-103, plus the device bytes of this build. */
typedef struct {
    uint32_t padded_shape[3];
    uint64_t padded_cells;
    uint64_t working_bytes; /* the reference's admission figure (see sp_hip_estimate) */
    uint64_t device_bytes;  /* device memory one FP64 correlation of this spec takes here;
                               0 when a padded axis exceeds the device FFT's maximum */
} sp_correlation_estimate;

enum sp_precision { SP_PRECISION_AUTO = 0, SP_PRECISION_FP32 = 32, SP_PRECISION_FP64 = 64 };

/* correlation.cpp:596-641 estimate_correlation_cpu: padded shape P = E+K-1 and
 * working_bytes = 40 prod(P) + fft_threads (64 * 16 max(P) + 64 KiB)
 *                 + metadata_bytes + 64
 * (two complex<double> buffers and the double output per padded cell, the
 * pinned pocketfft workspace bound per FFT thread, the caller's metadata
 * objects, the probe index array; the execution owner reserves nothing here).
 * fft_threads outside 1..8: CPU_THREAD_COUNT_UNSUPPORTED.  A correlation
 * admits reserved_bytes + working_bytes (one thread) against
 * max_working_bytes and reports that sum as working_bytes_peak, as the
 * reference does. */
int sp_hip_estimate(const sp_correlation_spec* spec, uint32_t fft_threads, uint64_t metadata_bytes,
                    sp_correlation_estimate* out);

/* correlation.hpp:108-111 correlate_binary_cpu: full-linear correlation of a
 * uint8 environment {0,1} with a uint8 kernel {0,1}.  out_values receives
 * prod(E+K-1) integer-valued doubles, X-fastest; out_first/out_shape the
 * translation_first = E.first - K.first - (K-1) and shape E+K-1.  Host buffers. */
int sp_hip_correlate_binary(sp_hip_ctx* ctx, const sp_correlation_spec* spec, const uint8_t* environment,
                            uint64_t environment_count, const uint8_t* kernel, uint64_t kernel_count,
                            const sp_correlation_limits* limits, int precision, double* out_values,
                            uint64_t out_capacity, int64_t out_first[3], uint32_t out_shape[3],
                            sp_numeric_report* report, sp_correlation_stats* stats);

/* No counterpart in the reference: sp_hip_correlate_binary for callers that
 * cannot allocate device memory (code recorded into a HIP graph, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, checks and
 * results; every device buffer of the call, twiddle tables included, lives in
 * one static device array of SP_NOALLOC_SCRATCH_BYTES per GPU (held from the
 * library's first use on that GPU until the process ends), and the call
 * makes no device allocation.  A call that needs more returns
 * SP_SCRATCH_LIMIT (64 MiB fits, e.g., an FP32 correlation on a 128^3 FFT
 * grid).  Calls on one GPU share its array: the library serialises the calls
 * of one process, and work that overlaps it on that GPU from outside these
 * calls races. */
#define SP_NOALLOC_SCRATCH_BYTES (64ull << 20)
int sp_hip_correlate_binary_noalloc(sp_hip_ctx* ctx, const sp_correlation_spec* spec, const uint8_t* environment,
                                    uint64_t environment_count, const uint8_t* kernel, uint64_t kernel_count,
                                    const sp_correlation_limits* limits, int precision, double* out_values,
                                    uint64_t out_capacity, int64_t out_first[3], uint32_t out_shape[3],
                                    sp_numeric_report* report, sp_correlation_stats* stats);

/* correlation.hpp:113-117 correlate_proximity_cpu: environment values in [0,1]
 * (double), FP64 on the device (the reference pins these to 1e-10 relative). */
int sp_hip_correlate_proximity(sp_hip_ctx* ctx, const sp_correlation_spec* spec, const double* environment,
                               uint64_t environment_count, const uint8_t* kernel, uint64_t kernel_count,
                               const sp_correlation_limits* limits, double* out_values, uint64_t out_capacity,
                               int64_t out_first[3], uint32_t out_shape[3], sp_numeric_report* report,
                               sp_correlation_stats* stats);

/* ---- batched placement step (gap_fill.cpp:363-509) -------------------------
 * A session keeps the environment occupancy (uint8, X-fastest) resident in HBM
 * together with O oriented kernels.  One step correlates every orientation in
 * one graph launch, forms each feasible mask (rounded overlap == 0 over legal
 * translations), picks the lowest kernel-window corner in (z, y, x) order with
 * the earliest orientation winning ties, and (optionally) stamps that kernel. */
typedef struct sp_hip_session sp_hip_session;

typedef struct {
    uint32_t environment_shape[3];
    int64_t environment_first[3];
    uint32_t orientations;
    int32_t precision;      /* sp_precision */
    uint32_t capture_graph; /* 1: replay each step as one hipGraph */
} sp_session_desc;

typedef struct {
    int32_t orientation;     /* -1 when no orientation has a free translation */
    uint32_t run_flags;      /* numeric_flags of every earlier step since the result was last read (OR) */
    int64_t translation[3];  /* t: kernel cell i sits at world cell t + kernel_first + i */
    int64_t key[3];          /* (t + kernel_first) as x, y, z */
    double max_residual;     /* max |v - round(v)| over legal translations, all orientations */
    uint32_t numeric_flags;  /* bit0: residual >= 0.25, bit1: range check failed, bit2: the chosen
                                corner's direct recheck found an overlap.  A step with any flag
                                stamps nothing. */
    int32_t precision;
    double error_bound;      /* a-priori bound of the step's precision */
} sp_step_result;

int sp_hip_session_create(sp_hip_ctx* ctx, const sp_session_desc* desc, sp_hip_session** out);
void sp_hip_session_destroy(sp_hip_session* s);
/* Host occupancy grid (uint8 {0,1}, prod(environment_shape) cells). */
int sp_hip_session_set_environment(sp_hip_session* s, const uint8_t* cells, uint64_t count);
int sp_hip_session_get_environment(sp_hip_session* s, uint8_t* cells, uint64_t count);
/* Kernel o: uint8 {0,1} cells of shape kernel_shape (X-fastest) whose cell 0
 * is world cell kernel_first; count must equal prod(kernel_shape)
 * (CORRELATION_INVALID_FIELD otherwise, as a shape/cells mismatch is in
 * correlation.cpp:310-313).  All orientations must be set before a step. */
int sp_hip_session_set_kernel(sp_hip_session* s, uint32_t orientation, const uint8_t* cells, uint64_t count,
                              const uint32_t kernel_shape[3], const int64_t kernel_first[3]);
/* One step; stamp != 0 ORs the chosen kernel into the occupancy on the device.
 * CORRELATION_NUMERIC when this step or an earlier unread one failed a check
 * (out is filled; the failing step stamped nothing). */
int sp_hip_session_step(sp_hip_session* s, int stamp, sp_step_result* out);
/* gap_fill.cpp:512-523 verdicts outside the graph: stamp kernel `orientation` at
 * `translation` (an accepted proposal), or keep that translation blocked for
 * that orientation in every later step (a rejected proposal). */
int sp_hip_session_stamp(sp_hip_session* s, uint32_t orientation, const int64_t translation[3]);
int sp_hip_session_exclude(sp_hip_session* s, uint32_t orientation, const int64_t translation[3]);
/* Placed copies (conservative_fields.hpp:127-164, conservative_fields.cpp:1666-1887
 * BlockedField add / remove / placed_count).  The occupancy given to
 * set_environment is the fixed mask.  Every copy the session stamps (a step's
 * choice, or sp_hip_session_stamp) gets the next copy id (0, 1, ... since the
 * environment was set) and adds one to the placed count of each cell of its
 * stamp footprint (its kernel, dilated by the pair clearance).  A cell is
 * blocked while the mask blocks it or its count is nonzero.  remove_copy
 * subtracts the copy's counts and frees the cells no copy and no mask block
 * any longer; the next step sees the new occupancy (incremental sessions
 * subtract the freed cells' correlation from their fields).  FIELD_COPY_ID:
 * the id was never stamped, is removed already, lies beyond the first 2^20
 * copies, or was stamped before the kernels or the pair clearance changed.
 * copies lists every stamped copy in id order (*count = copies; CAPACITY
 * when capacity is smaller); placed_counts reads the per-cell counts. */
typedef struct {
    uint64_t copy_id;
    int64_t translation[3];
    uint32_t orientation;
    uint32_t removed;
} sp_placed_copy;
int sp_hip_session_copies(sp_hip_session* s, sp_placed_copy* out, uint64_t capacity, uint64_t* count);
int sp_hip_session_remove_copy(sp_hip_session* s, uint64_t copy_id);
int sp_hip_session_placed_counts(sp_hip_session* s, uint32_t* counts, uint64_t count);
/* Enqueue `steps` steps without reading results back (for throughput timing);
 * sp_hip_session_last_result then reads the last one, and returns
 * CORRELATION_NUMERIC when any step since the previous read failed a check
 * (run_flags / numeric_flags say which). */
int sp_hip_session_run(sp_hip_session* s, uint32_t steps, int stamp);
int sp_hip_session_last_result(sp_hip_session* s, sp_step_result* out);
/* Legal overlap counts of orientation o from the last step (int32, X-fastest,
 * shape environment_shape - kernel_shape + 1); requires keep_fields. */
int sp_hip_session_keep_fields(sp_hip_session* s, int keep);
int sp_hip_session_overlap_field(sp_hip_session* s, uint32_t orientation, int32_t* out, uint64_t count,
                                 uint32_t legal_shape[3]);
/* Runs one step without the graph, timing each kernel with HIP events:
 * kernel_ms[i] and names (32 chars each) for up to capacity kernels; *count
 * receives the number of kernels. */
int sp_hip_session_profile_step(sp_hip_session* s, int stamp, float* kernel_ms, char* names, uint32_t capacity,
                                uint32_t* count);
/* 1 (default): register-resident FFT passes for 128-, 256- and 512-point axes;
 * 0: the generic LDS Stockham passes for every axis (same results).  A change
 * after an incremental session's first step fails with INVALID_ARGUMENT. */
int sp_hip_session_set_fast_path(sp_hip_session* s, int on);
/* 1 (default): on FP32 sessions over a 256^3 FFT grid whose steps keep no
 * per-value output (no overlap fields, no raw values, the inverse y and z
 * tiles not pinned), the inverse y and inverse z passes run as one pass per
 * legal x column that keeps the column's spectrum on chip (the product pass
 * then writes its output column-major); 0: the two passes run separately.
 * Same results.  get_fused_passes reports whether the next step fuses them.
 * No counterpart in the reference (a scheduling choice of this build). */
int sp_hip_session_set_fused_passes(sp_hip_session* s, int on);
int sp_hip_session_get_fused_passes(sp_hip_session* s, int32_t* active);
/* Unrounded legal values (double) of the last step, for measuring the FP32
 * deviation from FP64 on the same occupancy; requires keep_raw. */
int sp_hip_session_keep_raw(sp_hip_session* s, int keep);
int sp_hip_session_raw_values(sp_hip_session* s, uint32_t orientation, double* out, uint64_t count);
/* Autotuner of the step's three main passes.  Register-FFT passes (256- and
 * 512-point axes) choose (radix, tile width); generic passes (other lengths)
 * choose the tile width with radix 0.  pass 0:
 * kernel x transform + product + inverse x; 1: inverse y; 2: C2R z + argmin.
 * Each candidate (radix, tile width) whose LDS fits the device is timed on
 * the session's own buffers before the first step (events, 3 launches) and
 * the fastest is kept, cached per shape for the context.  set_autotune(0)
 * uses the fixed defaults; set_tile pins one pass (radix 0 and width 0
 * unpin it); get_tile reports the tile in use and its measured ms per launch
 * (0: not measured, -1: taken from the cache); tile_candidates lists the
 * fitting (radix, width) pairs.  Every tile gives identical results. */
int sp_hip_session_set_autotune(sp_hip_session* s, int on);
int sp_hip_session_set_tile(sp_hip_session* s, int pass, int radix, int width);
int sp_hip_session_get_tile(sp_hip_session* s, int pass, int32_t* radix, int32_t* width, float* ms);
int sp_hip_session_tile_candidates(sp_hip_session* s, int pass, int32_t* pairs, uint32_t capacity, uint32_t* count);
/* Pair clearance between placed copies (conservative_fields.cpp:189-204,
 * 398-434, 1015-1122): a stamped copy also blocks every cell within the
 * clearance stencil of its cells (offsets within halo = ceil(c / pitch) + 1
 * cells that validation_kernel.cpp:3492-3522 does not rule out).  0 (the
 * default) stamps the kernel's own cells.  Takes effect at the next step. */
int sp_hip_session_set_pair_clearance(sp_hip_session* s, double clearance_mm, double pitch_mm);
/* Incremental steps (gap_fill.cpp:426-451,524-584: blocked offsets from
 * small correlations).  The first step computes every orientation's overlap
 * field in full; afterwards the fields (int32, exact) are updated by the
 * correlation of each step's newly set cells with every kernel, computed as
 * a circular correlation on an FFT grid of at least 2 Kmax - 1 cells per
 * axis, and each orientation's
 * lowest feasible corner is rescanned forward only when it became blocked.
 * Fields, choices and occupancy are identical to full steps.  Used when the
 * window's FFT grid is at most half the session's; otherwise steps stay
 * full.  Not combinable with set_active or proposal steps. */
int sp_hip_session_set_incremental(sp_hip_session* s, int on);
/* 1 (default): every orientation's kernel spectrum (forward transform up to
 * its y pass; it does not depend on the occupancy) is computed once after the
 * kernels are set and stays resident in HBM; steps start from it.  0: each
 * step recomputes it.  Results are identical. */
int sp_hip_session_set_resident_kernels(sp_hip_session* s, int on);
/* Search restriction (no reference counterpart; used by the learned
 * proposal).  The next steps consider only the listed orientations (count 0:
 * every orientation) and, when region_mask is not null, only kernel-window
 * corners u (relative to environment_first) whose region
 * (u * regions / environment_shape per axis, X-fastest, regions[a] in 1..64)
 * has a nonzero mask byte.  count 0 and a null mask restore the plain search.
 * Orientations whose kernel does not fit are dropped from the list. */
int sp_hip_session_set_active(sp_hip_session* s, const uint32_t* orientations, uint32_t count,
                              const uint8_t* region_mask, const uint32_t regions[3]);
/* Lowest feasible key of every orientation in the last step, packed as
 * (uz << 48) | (uy << 32) | (ux << 16) | orientation with u the kernel-window
 * corner relative to environment_first; all ones when the orientation had no
 * feasible translation (or was not searched).  count >= orientations. */
int sp_hip_session_orientation_minima(sp_hip_session* s, uint64_t* keys, uint32_t count);
/* Milliseconds per launch of one pass of the step ("mul_x", "yz_argmin",
 * "inv_c2c_y" or "c2r_z_argmin"), launched back to back: `launches` launches
 * (1..1000) recorded in one graph and replayed between two events after an
 * unstamped step has filled the pass's inputs.  Each launch rewrites the same
 * output from the same inputs, so results and occupancy are unchanged.  No
 * counterpart in the reference (measurement of this build). */
int sp_hip_session_time_pass(sp_hip_session* s, const char* pass, uint32_t launches, float* ms);
/* Grid of the last column-major product pass (fused path): its blocks and its
 * row tiles; fewer blocks than tiles means each block took an even share of
 * row tiles x active orientations.  0, 0 before one ran.  No counterpart in
 * the reference (a scheduling detail of this build, for tests). */
int sp_hip_session_product_grid(sp_hip_session* s, uint32_t* blocks, uint64_t* tiles);
/* Stream the session launches on (hipStream_t), for events around its work. */
void* sp_hip_session_stream(sp_hip_session* s);
/* Device bytes held by the session. */
uint64_t sp_hip_session_device_bytes(const sp_hip_session* s);

/* ---- ranked candidate pages ------------------------------------------------
 * The pack_solver greedy path's ranking (spectral_pipeline.hpp:26-51
 * RankedCandidate / CandidatePageMode / CandidatePageQuery;
 * spectral_pipeline.cpp:1343 build_proximity, :1563 correlate_proximity_cpu,
 * :1740-1857 the page), batched over orientations on one session:
 *   proximity field  exp(-sqrt(d2)/2) of the occupancy (field_proximity.hpp:32-132),
 *   proximity        its FP64 correlation with each kernel at every legal corner,
 *   score            .65 * top - .35 * proximity / kernel ones, where
 *                    top = (origin_z + pitch * t_z + object_max_z - container_min_z)
 *                          / (container_max_z - container_min_z),
 *   page             the 32 best candidates after the cursor by (score,
 *                    translation z, y, x); ordinary pages keep only zero binary
 *                    overlap and recheck each entry directly against the
 *                    occupancy; low-overlap pages order by overlap first.
 * Rejected translations (sp_hip_session_exclude) never appear in a page. */
typedef struct sp_ranking_desc {
    double lattice_origin_z_mm;  /* lattice.origin_mm[2] */
    double pitch_mm;             /* lattice.pitch_mm */
    double container_min_z_mm;   /* container bounds min z */
    double container_max_z_mm;   /* container bounds max z (height must be > 0) */
} sp_ranking_desc;

typedef struct sp_ranked_candidate {
    int64_t translation[3];
    double score;
    double binary_overlap;        /* exact count; 0 in ordinary pages */
    uint32_t orientation;
    uint32_t direct_zero_overlap; /* ordinary pages: the direct recheck found no overlap */
} sp_ranked_candidate;

#define SP_PAGE_SIZE 32
#define SP_PAGE_ORDINARY 0
#define SP_PAGE_LOW_OVERLAP 1

/* Enables ranking: the constants above and each orientation's oriented-bounds
 * max z in mm (geometry::oriented_bounds(...).bounds_mm.max[2]), count =
 * orientations.  Environments of at most 1024 cells per axis.  Keeps the
 * binary overlap fields. */
int sp_hip_session_set_ranking(sp_hip_session* s, const sp_ranking_desc* desc, const double* object_max_z_mm,
                               uint32_t count);
/* Binary overlaps, proximity field and proximity correlation of the current
 * occupancy (ranked_pages runs this by itself when the occupancy changed). */
int sp_hip_session_rank(sp_hip_session* s);
/* One page per listed orientation: pages[i * SP_PAGE_SIZE + j], j <
 * page_counts[i], in rank order; cursors[i] is used when has_cursor[i] is
 * nonzero (both may be null).  An orientation whose kernel does not fit the
 * window has an empty page. */
int sp_hip_session_ranked_pages(sp_hip_session* s, int mode, const uint32_t* orientations, uint32_t count,
                                const sp_ranked_candidate* cursors, const uint8_t* has_cursor,
                                sp_ranked_candidate* pages, uint32_t* page_counts);
/* spectral.cpp:1485-1590, the ordinary greedy loop over ranked pages: fresh
 * pages of every orientation for the current occupancy; the best page head by
 * (score, orientation, translation z, y, x) in turn; a consumed full page
 * refilled after its cursor; the first head whose direct recheck found no
 * overlap and that `validate` accepts (the caller's exact check, nonzero =
 * accept; every head accepted when null) is stamped, and the search starts
 * over.  Stops when no page has a candidate or min(max_copies, capacity)
 * copies are placed; *placed = copies written to placements. */
typedef int (*sp_validate_fn)(void* user, uint32_t orientation, const int64_t translation[3]);
typedef struct sp_placement {
    int64_t translation[3];
    uint32_t orientation;
    uint32_t reserved;
    double score;
} sp_placement;
int sp_hip_session_ranked_fill(sp_hip_session* s, uint32_t max_copies, sp_validate_fn validate, void* user,
                               sp_placement* placements, uint32_t capacity, uint32_t* placed);
/* The proximity field of the last rank: values (may be null) and exact squared
 * distances (may be null; 0x7fffffff where no cell is occupied), X-fastest. */
int sp_hip_session_proximity_field(sp_hip_session* s, double* values, uint32_t* squared, uint64_t count);
/* The proximity correlation of the last rank at orientation o's legal corners
 * (X-fastest, shape environment_shape - kernel_shape + 1). */
int sp_hip_session_proximity_values(sp_hip_session* s, uint32_t o, double* values, uint64_t count);

/* ---- conservative rasteriser (conservative_fields.hpp:100-164) -------------
 * Mesh: vertex_count x 3 doubles (local frame), triangle_count x 3 indices.
 *
 * voxelize_object (conservative_fields.cpp:919-1003) for `count` canonical
 * quaternions (x, y, z, w) at the lattice origin (which must be 0,0,0):
 * per orientation the trimmed window (firsts/shapes, 3 per orientation) and
 * {0,1} cells (X-fastest) concatenated at offsets[k].  When cells is null or
 * capacity < required, returns SP_CAPACITY with every window and *required
 * filled. */
int sp_hip_voxelize_objects(sp_hip_ctx* ctx, const double* vertices, uint32_t vertex_count, const uint32_t* triangles,
                            uint32_t triangle_count, const double origin[3], double pitch, const double* quaternions,
                            uint32_t count, int64_t* firsts, uint32_t* shapes, uint64_t* offsets, uint8_t* cells,
                            uint64_t capacity, uint64_t* required);
/* voxelize_container, analytic box (conservative_fields.cpp:1146-1198):
 * 1 = blocked, 0 = certified inside the box shrunk by the clearance. */
int sp_hip_voxelize_box_container(sp_hip_ctx* ctx, const int64_t first[3], const uint32_t shape[3], const double origin[3],
                                  double pitch, const double dims[3], double clearance, uint8_t* cells);
/* voxelize_container, closed mesh (conservative_fields.cpp:1200-1275): only
 * cells certified inside the solid are free (0), and with a wall clearance
 * every cell within the clearance stencil of a cell that is not certified
 * inside is blocked too (raster and fill over the window grown by the halo,
 * cf:189-204 and 436-504). */
int sp_hip_voxelize_mesh_container(sp_hip_ctx* ctx, const double* vertices, uint32_t vertex_count,
                                   const uint32_t* triangles, uint32_t triangle_count, const int64_t first[3],
                                   const uint32_t shape[3], const double origin[3], double pitch, double clearance,
                                   uint8_t* cells);

/* ---- learned proposal (no reference counterpart) ---------------------------
 * A BF16 3D CNN (gfx950 MFMA, FP32 accumulate) ranks the orientations of a
 * step and the 16^3 regions of the environment; the exact search then runs on
 * the top orientations in the top regions only, and the full search is the
 * fallback when that finds nothing.  Inputs: the occupancy downsampled to 64^3
 * (uint8 fraction q = round(255 occupied / cells) per bin, bin b of an axis of
 * extent e covering cells [b e / 64, (b + 1) e / 64)) and a 32-float part
 * descriptor.  Outputs: 96 orientation logits and 4096 region logits
 * (X-fastest 16^3).  Parameters are one flat FP32 array in this order:
 *   w1[32][27] b1[32]  w2[32][27][32] b2[32]  w3[32][27][32] b3[32]
 *   w4[32][27][32] b4[32]  wf[32][32]  wh[32] bh  wa[64][96] ba[64]
 *   wo[96][64] bo[96]
 * (conv weights [out][tap][in], tap = 9 dz + 3 dy + dx, source voxel
 * stride * v + d - 1; wa reads [mean of layer 4, layer 4 pooled by the
 * softmax of the region logits, descriptor]). */
typedef struct sp_hip_proposal sp_hip_proposal;

uint32_t sp_hip_proposal_parameter_count(void);
/* max_batch: largest training or forward batch (1..1024). */
int sp_hip_proposal_create(sp_hip_ctx* ctx, uint32_t max_batch, sp_hip_proposal** out);
void sp_hip_proposal_destroy(sp_hip_proposal* p);
/* Setting parameters also resets the Adam state. */
int sp_hip_proposal_set_parameters(sp_hip_proposal* p, const float* params, uint64_t count);
int sp_hip_proposal_get_parameters(sp_hip_proposal* p, float* params, uint64_t count);
/* Device bank of 64^3 uint8 samples for training: reserve slots, fill them
 * from the host or from a session's current occupancy, read one back. */
int sp_hip_proposal_reserve(sp_hip_proposal* p, uint32_t samples);
int sp_hip_proposal_load_samples(sp_hip_proposal* p, const uint8_t* samples, uint32_t first, uint32_t count);
int sp_hip_proposal_record(sp_hip_proposal* p, sp_hip_session* s, uint32_t slot);
int sp_hip_proposal_read_sample(sp_hip_proposal* p, uint32_t slot, uint8_t* sample);
/* Forward over bank samples: logits [batch][96] and heat [batch][4096]
 * (either may be null). */
int sp_hip_proposal_forward(sp_hip_proposal* p, const uint32_t* slots, uint32_t batch, const float* descriptors,
                            float* logits, float* heat);
/* Loss and gradients over bank samples.  Orientation loss: cross entropy of
 * the softmax over orientations with a nonzero mask byte against
 * orientation_target [batch][96]; region loss: cross entropy of the softmax
 * over the 4096 heat logits against region[b] (none when < 0), times
 * heat_weight; both averaged over the batch.  losses receives the two means;
 * gradients (nullable) the flat FP32 gradient; update != 0 applies one Adam
 * step (beta 0.9 / 0.999, eps 1e-8) with learning_rate. */
int sp_hip_proposal_train(sp_hip_proposal* p, const uint32_t* slots, uint32_t batch, const float* descriptors,
                          const float* orientation_target, const uint8_t* orientation_mask, const int32_t* region,
                          float heat_weight, float learning_rate, int update, float losses[2], float* gradients);
/* The scene's orientation prior for sp_hip_proposal_step: orientations are
 * ranked by logit[o] + weight * prior[o] (prior[o] = 0 for o >= count, and
 * for every o until this is called).  count <= 96; values finite. */
int sp_hip_proposal_set_prior(sp_hip_proposal* p, const float* prior, uint32_t count, float weight);
/* One learned placement step on a session with at most 96 orientations and
 * at least 64 cells per axis: the top_orientations fitting orientations by
 * score (logit plus the prior term; ties: lower index) searched exactly in
 * the top_regions regions by heat logit (4096: every region, no region
 * mask), all in one graph.  When nothing there is feasible and fallback
 * != 0, the full search runs (*used_fallback = 1).  out->orientation is -1
 * when no step was found. */
int sp_hip_proposal_step(sp_hip_proposal* p, sp_hip_session* s, const float descriptor[32], uint32_t top_orientations,
                         uint32_t top_regions, int stamp, int fallback, sp_step_result* out, int* used_fallback);
/* One lazy learned step.  The proposal keeps every orientation's last
 * lowest feasible corner; occupancy only grows, so a corner still free is
 * still that orientation's minimum, and a blocked one bounds the new minimum
 * from below.  The orientations whose corner the occupancy blocked and whose
 * bound lies below the lowest current minimum are ranked by score (logit
 * plus the prior term; ties: lower index), the top_orientations of them are
 * searched exactly (every region), and the step places the lowest current
 * minimum: the plain search's choice whenever no more than
 * top_orientations orientations needed a search.  The full search runs
 * instead (*used_full = 1), making every minimum current, on the first lazy
 * step on a session, after sp_hip_proposal_lazy_reset, after any change to
 * the session other than stamping that alters which positions are feasible
 * (sp_hip_session_set_environment, a kernel, an excluded translation, a
 * removed copy, the pair clearance), and when the known minima hold nothing
 * feasible.  *searched: the orientations searched (nullable outputs). */
int sp_hip_proposal_lazy_step(sp_hip_proposal* p, sp_hip_session* s, const float descriptor[32],
                              uint32_t top_orientations, int stamp, sp_step_result* out, int* used_full,
                              uint32_t* searched);
/* Forgets the kept minima: the next lazy step runs the full search. */
int sp_hip_proposal_lazy_reset(sp_hip_proposal* p);
/* Self-test of the MFMA tile layout: c[32][32] = a[32][16] * b[16][32], BF16
 * bit patterns in, FP32 out. */
int sp_hip_proposal_mfma_probe(sp_hip_proposal* p, const uint16_t* a, const uint16_t* b, float* c);

/* A-priori FP32 rounding bound for a correlation of an environment of mass
 * env_mass with a kernel of mass kernel_mass on an FFT grid of fft_cells cells:
 * lambda * 2u * sqrt(3 log2 N) * rms, with rms^2 <= min(mB, mK) mB mK / N and
 * lambda = sqrt(2 (ln 2N + 40 ln 2)).  FP64 is chosen when > 0.4. */
double sp_hip_fp32_error_bound(uint64_t fft_cells, uint64_t env_mass, uint64_t kernel_mass);

#ifdef __cplusplus
}
#endif

#endif
