/*
 * neonufft_amd.h -- C ABI of the MI355X (gfx950) NeoNUFFT GPU plans.
 *
 * Every entry point is a plain-C mirror of one member of the reference's GPU
 * plan classes neonufft::gpu::Plan<T,DIM> / neonufft::gpu::PlanT3<T,DIM>
 * (include/neonufft/gpu/plan.hpp of the reference; the C++ classes of this
 * library, include/neonufft/gpu/plan.hpp, are built on top of the same code).
 * A binding (ctypes, cffi, JNI, cgo) needs nothing but this header.
 *
 * Conventions:
 *  - precision: NNF_F32 (T = float, complex = 2 x float interleaved) or
 *    NNF_F64 (T = double). dim: 1, 2 or 3. Type-3 plans also take NNF_MIXED
 *    (extension, no reference counterpart): double coordinates and complex128
 *    inputs / outputs, FP64 fold and phase arguments, fp32 spreading, FFT and
 *    interpolation.
 *  - Every data pointer is a DEVICE pointer. Coordinate arrays are passed as
 *    an array of `dim` pointers to contiguous T arrays (struct of arrays),
 *    as in the reference's std::array<const T*, DIM>.
 *  - Mode arrays are column-major (x fastest) with element strides.
 *  - `stream` is a hipStream_t (NULL = the null stream). Work is enqueued on it.
 *    Plan creation and set_points synchronise the stream; transforms do not.
 *  - Return value: a NeonufftError code (include/neonufft/errors.h, same values
 *    as the reference). On failure nnf_last_error() returns a message for the
 *    calling thread.
 *  - A plan must not be used from two host threads at once (it owns its grids).
 */
#ifndef NEONUFFT_AMD_H
#define NEONUFFT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NNF_F32 0
#define NNF_F64 1
#define NNF_MIXED 2

/* Mirrors neonufft::Options (include/neonufft/plan.hpp:19-29 of the reference),
 * with the C++ bools as ints. order: 0 = NEONUFFT_MODE_ORDER_FFT,
 * 1 = NEONUFFT_MODE_ORDER_CMCL (default). */
typedef struct nnf_options {
  double tol;
  double upsampfac;
  double recenter_threshold;
  int num_threads;
  int sort_input;
  int sort_output;
  int kernel_approximation;
  int order;
  int kernel_type;
} nnf_options;

typedef struct nnf_plan_s* nnf_plan;
typedef struct nnf_plan_t3_s* nnf_plan_t3;
typedef struct nnf_batch_s* nnf_batch;

/* Fills the reference defaults: tol 1e-3, upsampfac 2, recenter 0.1, CMCL. */
void nnf_options_default(nnf_options* opt);

/* Message of the last failing call on this thread ("" if none). */
const char* nnf_last_error(void);

/* Library version string. */
const char* nnf_version(void);

/* ---- Type 1 / type 2 plan --------------------------------------------- */

/* gpu::Plan<T,DIM>::Plan(opt, sign, num_nu, loc, modes, stream)
 * (include/neonufft/gpu/plan.hpp:30-32). Folds, rescales and bins the points
 * (as set_points) and synchronises the stream. */
int nnf_plan_create(nnf_plan* plan, int precision, int dim, const nnf_options* opt, int sign,
                    int64_t num_nu, const void* const* loc, const int64_t* modes, void* stream);

/* gpu::Plan<T,DIM>::set_points (include/neonufft/gpu/plan.hpp:42). */
int nnf_plan_set_points(nnf_plan plan, int64_t num_nu, const void* const* loc);

/* No counterpart in the reference: nnf_plan_set_points without any device
 * allocation, for callers that cannot allocate (serving loops that must not
 * call hipMalloc between requests, allocation-free sections of a program).
 * The temporaries of the fold and the counting sorts live in a fixed array
 * inside the library, 64 MiB per GPU (held from the library's load to the
 * process's end), which limits a call to about 16 million points; the
 * results go to the memory that earlier nnf_plan_create / nnf_plan_set_points
 * calls left in the plan. Fails with NEONUFFT_INPUT_ERROR, before any work,
 * when num_nu exceeds the largest point count the plan has held or the
 * temporaries exceed the array; fails the same way, after the sort, when the
 * new points need more subproblems or spreading tiles than the plan holds
 * (clustered points after uniform ones), and the plan then refuses
 * transforms until a set_points succeeds. All plans on one GPU share its
 * array: calls are serialised per device and return with the plan's stream
 * synchronised, so the array is free again when a call returns. */
int nnf_plan_set_points_noalloc(nnf_plan plan, int64_t num_nu, const void* const* loc);

/* gpu::Plan<T,DIM>::transform_type_1 (include/neonufft/gpu/plan.hpp:34-35):
 * in = num_nu complex strengths, out = modes written with out_strides. */
int nnf_plan_transform_type_1(nnf_plan plan, const void* in, void* out,
                              const int64_t* out_strides);

/* gpu::Plan<T,DIM>::transform_type_2 (include/neonufft/gpu/plan.hpp:37-39):
 * in = modes read with in_strides, out = num_nu complex values. */
int nnf_plan_transform_type_2(nnf_plan plan, const void* in, const int64_t* in_strides,
                              void* out);

/* Destroys the plan (the reference's pImpl deleter, src/neonufft/gpu/plan.cpp:17-21).
 * Does not synchronise: the caller must not destroy a plan with work in flight. */
int nnf_plan_destroy(nnf_plan plan);

/* ---- Type 3 plan --------------------------------------------------------- */

/* gpu::PlanT3<T,DIM>::grid_memory_size (include/neonufft/gpu/plan.hpp:53-57).
 * Bounds are given as doubles (exact for float plans). */
int nnf_plan_t3_grid_memory_size(int precision, int dim, const nnf_options* opt,
                                 const double* input_min, const double* input_max,
                                 const double* output_min, const double* output_max,
                                 uint64_t* size);

/* gpu::PlanT3<T,DIM>::PlanT3(opt, sign, num_in, input_points, num_out,
 * output_points, stream, batch_size) (include/neonufft/gpu/plan.hpp:59-62). */
int nnf_plan_t3_create(nnf_plan_t3* plan, int precision, int dim, const nnf_options* opt,
                       int sign, int64_t num_in, const void* const* input_points, int64_t num_out,
                       const void* const* output_points, void* stream, int64_t batch_size);

/* gpu::PlanT3<T,DIM>::PlanT3(opt, sign, input_min, input_max, output_min,
 * output_max, stream, batch_size) (include/neonufft/gpu/plan.hpp:64-67). */
int nnf_plan_t3_create_bounds(nnf_plan_t3* plan, int precision, int dim, const nnf_options* opt,
                              int sign, const double* input_min, const double* input_max,
                              const double* output_min, const double* output_max, void* stream,
                              int64_t batch_size);

/* gpu::PlanT3<T,DIM>::set_input_points (include/neonufft/gpu/plan.hpp:69-70). */
int nnf_plan_t3_set_input_points(nnf_plan_t3 plan, int64_t num_in, const void* const* points);

/* gpu::PlanT3<T,DIM>::set_output_points (include/neonufft/gpu/plan.hpp:72-73). */
int nnf_plan_t3_set_output_points(nnf_plan_t3 plan, int64_t num_out, const void* const* points);

/* gpu::PlanT3<T,DIM>::add_input (include/neonufft/gpu/plan.hpp:75); bdist <= 0 means num_in. */
int nnf_plan_t3_add_input(nnf_plan_t3 plan, const void* in, int64_t bdist);

/* gpu::PlanT3<T,DIM>::transform (include/neonufft/gpu/plan.hpp:77); bdist <= 0 means num_out. */
int nnf_plan_t3_transform(nnf_plan_t3 plan, void* out, int64_t bdist);

/* gpu::PlanT3<T,DIM>::reset (include/neonufft/gpu/plan.hpp:79). */
int nnf_plan_t3_reset(nnf_plan_t3 plan);

int nnf_plan_t3_destroy(nnf_plan_t3 plan);

/* ---- Tuning extension (no reference counterpart) -------------------------
 * The subgrid tile (bin) edge per dimension and the maximum number of points
 * one workgroup handles. Zero entries keep the built-in defaults.
 * spread_mode selects the type-1 spreading of 2D plans:
 *   NNF_SPREAD_TILES: per-subproblem tiles + tile reduction;
 *   NNF_SPREAD_BANDS: one workgroup per band of whole fine-grid rows writes
 *     its cells once; when n0 is a power of two >= 32, type 1 also runs the
 *     x-direction FFT of the band in LDS and writes only the retained modes;
 *   NNF_SPREAD_BANDS_GRID: row bands that always write the fine grid;
 *   NNF_SPREAD_AUTO: NNF_SPREAD_BANDS whenever the grid suits the row view
 *     (n0 + ns/2 and n1 at most 65535), else tiles.
 * band_rows: fine-grid rows per band (0 = default). Other dimensions always
 * use tiles. */
#define NNF_SPREAD_AUTO 0
#define NNF_SPREAD_TILES 1
#define NNF_SPREAD_BANDS 2
#define NNF_SPREAD_BANDS_GRID 3
typedef struct nnf_tuning {
  int32_t bin_size[3];
  int32_t max_points;
  int32_t spread_mode;
  int32_t band_rows;
  int32_t reserved[2];
} nnf_tuning;

/* nnf_plan_create with explicit tuning (NULL = defaults). */
int nnf_plan_create_tuned(nnf_plan* plan, int precision, int dim, const nnf_options* opt,
                          int sign, int64_t num_nu, const void* const* loc, const int64_t* modes,
                          void* stream, const nnf_tuning* tuning);

/* ---- Autotuner (no reference counterpart) ---------------------------------
 * Times the spreading (types bit 0) and / or interpolation (bit 1) kernels of
 * a plan with these arguments for a list of subgrid tile shapes and
 * per-workgroup point caps, on the given device points, and returns the
 * fastest in `best` (pass it to nnf_plan_create_tuned) with its time per
 * call in `best_ms` (sum over the selected kernels). 2D plans that spread
 * with row bands time their type 1 for a list of band heights instead of
 * the tile shapes (best->band_rows). `reps` timed calls per
 * candidate. With use_cache non-zero, a result is reused for the same
 * precision, dim, kernel width, fine grid, floor(log2 num_nu), types and
 * point-distribution signature (largest over mean occupied-bin count and share
 * of empty bins on the default tiles, so clustered and uniform points of one
 * size are tuned separately), from this process or from the cache file
 * written by earlier processes: NEONUFFT_AMD_TUNE_CACHE names it ("" = no
 * file), else $XDG_CACHE_HOME or $HOME/.cache /neonufft_amd/
 * autotune-<version>-<gfx arch>.txt. best->reserved[0] tells where the
 * answer came from: 0 = timed by this call, 1 = this process's cache,
 * 2 = the cache file. */
int nnf_autotune(int precision, int dim, const nnf_options* opt, int sign, int64_t num_nu,
                 const void* const* loc, const int64_t* modes, void* stream, int types, int reps,
                 int use_cache, nnf_tuning* best, double* best_ms);
int nnf_autotune_clear_cache(void);

/* ---- Stage-level extension (no reference counterpart) --------------------
 * The reference runs these stages only inside transform_type_{1,2}
 * (src/neonufft/gpu/plan_impl.hpp:91-136). They are exposed for stage-level
 * parity tests and for timing the spreading / interpolation kernels alone. */

typedef struct nnf_plan_info {
  int precision;
  int dim;
  int n_spread;          /* kernel width ns */
  double es_beta;        /* ES kernel shape */
  int64_t modes[3];      /* N_d (type 1/2) or spread-grid size (type 3) */
  int64_t grid[3];       /* fine (FFT) grid n_d */
  int64_t spread_grid[3];/* grid the points are spread onto (type 3: n_d, else = grid) */
  int64_t bin_size[3];   /* subgrid tile (bin) edge per dim */
  int64_t num_bins;
  int64_t num_subproblems;
  int64_t num_points;
  int64_t num_out_points;/* type 3 targets */
  int horner_degree;     /* fp32 ES polynomial degree (0 = exact exp/sqrt) */
  int quad_nodes;        /* Gauss-Legendre nodes on (0,1) for phihat */
  int spread_mode;       /* type-1 spreading: NNF_SPREAD_TILES, _BANDS (fused x-FFT) or _BANDS_GRID */
  int band_rows;         /* fine-grid rows per band (row bands), else 0 */
} nnf_plan_info;

int nnf_plan_get_info(nnf_plan plan, nnf_plan_info* info);
int nnf_plan_t3_get_info(nnf_plan_t3 plan, nnf_plan_info* info);

/* Spread `in` onto the fine grid (type-1 stage 1): the grid is overwritten,
 * or added to when `accumulate` is non-zero. Equals nnf_plan_spread_tiles
 * followed by nnf_plan_reduce_tiles. */
int nnf_plan_spread(nnf_plan plan, const void* in, int accumulate);
/* The spreading kernel alone: every subproblem's subgrid tile into the plan's
 * tile buffer. */
int nnf_plan_spread_tiles(nnf_plan plan, const void* in);
/* The tile reduction alone: fine grid (+)= sum of the covering tiles. */
int nnf_plan_reduce_tiles(nnf_plan plan, int accumulate);
/* Interpolate the current fine grid at the plan's points into `out`
 * (type-2 stage 3). */
int nnf_plan_interp(nnf_plan plan, void* out);
/* Device pointer of the plan's fine grid (n_0 x n_1 x n_2 complex, x fastest). */
int nnf_plan_fine_grid(nnf_plan plan, void** grid);
/* Device pointer of the deconvolution factors fs_d[0..n_d/2] (T). */
int nnf_plan_fseries(nnf_plan plan, int d, const void** fs);

/* ---- Batched uneven transforms (no reference counterpart) ----------------
 * `count` independent type 1 / type 2 transforms of one precision, dimension
 * and sign, each with its own points and mode counts, issued together: one
 * zero fill, one spreading / interpolation launch over all of their
 * subproblems, one batched FFT per distinct mode shape. Replaces a loop over
 * `count` plans of the reference's gpu::Plan (include/neonufft/gpu/plan.hpp),
 * with the same result per transform.
 *   num_nu[t]            points of transform t
 *   loc[t * dim + d]     device coordinates of transform t, dimension d
 *   modes[t * dim + d]   N_d of transform t
 *   use_graph            non-zero: capture each transform type once into a
 *                        HIP graph on first use and replay it afterwards.
 * Mode arrays of every transform are contiguous, x fastest (strides
 * 1, N_0, N_0 N_1). The pointer tables of transform_type_{1,2} are host
 * arrays of `count` device pointers. */
int nnf_batch_create(nnf_batch* batch, int precision, int dim, const nnf_options* opt, int sign,
                     int64_t count, const int64_t* num_nu, const void* const* loc,
                     const int64_t* modes, void* stream, int use_graph);
/* nnf_batch_create with explicit tuning (NULL = defaults; bin_size,
 * max_points, spread_mode and band_rows apply to every transform). */
int nnf_batch_create_tuned(nnf_batch* batch, int precision, int dim, const nnf_options* opt,
                           int sign, int64_t count, const int64_t* num_nu,
                           const void* const* loc, const int64_t* modes, void* stream,
                           int use_graph, const nnf_tuning* tuning);
int nnf_batch_set_points(nnf_batch batch, const int64_t* num_nu, const void* const* loc);
int nnf_batch_transform_type_1(nnf_batch batch, const void* const* in, void* const* out);
int nnf_batch_transform_type_2(nnf_batch batch, const void* const* in, void* const* out);

typedef struct nnf_batch_info {
  int precision;
  int dim;
  int n_spread;
  int64_t count;
  int64_t num_points;       /* sum over transforms */
  int64_t num_subproblems;  /* one launch covers all of them */
  int64_t num_shapes;       /* distinct mode shapes = batched FFT plans */
  int64_t grid_cells;       /* fine-grid cells over all transforms */
  int graph;                /* 1 once a captured graph is being replayed */
  int spread_mode;          /* type-1 spreading: NNF_SPREAD_TILES, _BANDS or _BANDS_GRID */
} nnf_batch_info;

int nnf_batch_get_info(nnf_batch batch, nnf_batch_info* info);
int nnf_batch_destroy(nnf_batch batch);

#ifdef __cplusplus
}
#endif

#endif /* NEONUFFT_AMD_H */
