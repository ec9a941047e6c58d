/*
 * libmiccai_hip.so - C ABI of the gfx950 HIP implementation of the
 * tubular-segmentation path (Frangi vesselness -> direction adjustment ->
 * gradient vector flow -> anisotropic-TV continuous max-flow).
 *
 * Conventions
 *  - Volumes are C-order float32 (D0, D1, D2), axis 2 fastest; dims[] holds
 *    (D0, D1, D2) and every entry point requires D0*D1*D2 < 2^31.
 *  - "AoS x k": k floats per voxel, interleaved. "SoA x k": k separate fields
 *    of N floats, field f at ptr + f*N.
 *  - Pointers named *_dev are device pointers (e.g. torch.Tensor.data_ptr()),
 *    owned by the caller. Pointers named *_host are host memory.
 *  - stream is a hipStream_t passed as void* (NULL = the null stream).
 *    Every entry point only enqueues work unless its comment says it blocks.
 *  - Return: MH_OK (0); < 0 invalid argument or HIP error (message in
 *    mh_last_error); > 0 a numerical early stop (NaN criterion).
 *
 * Reference paths below are relative to the reference repository's src/.
 */
#ifndef MICCAI_HIP_H
#define MICCAI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_OK 0
#define MH_ERR_ARG -1
#define MH_ERR_HIP -2
#define MH_STOPPED_NAN 1
#define MH_DEFER 2 /* the device form does not apply here: nothing was changed */

typedef struct mh_ctx mh_ctx;

/* ---------------------------------------------------------------- context */

/* Replaces ClProgramLoader.__init__ (helpers/cl_helpers.py:93-114): binds a
 * device and allocates the context's small reduction/control block. */
int mh_create(int device, mh_ctx **out);
int mh_destroy(mh_ctx *ctx);
/* Last error message of this context (never NULL). */
const char *mh_last_error(mh_ctx *ctx);
/* ABI version (major*100 + minor). */
int mh_version(void);

/* ------------------------------------------------------- elementwise/util */

/* Replaces helpers.cl to_zeros / to_value (clcode/helpers.cl:13-34). */
int mh_fill(mh_ctx *ctx, float *x_dev, int64_t n, float value, void *stream);

/* Replaces Reducer.reduce with sum_abs.cl / max.cl (helpers/cl_helpers.py:
 * 232-298, clcode/sum_abs.cl:25-58, clcode/max.cl:25-58). The result is written
 * to out_dev (1 float, device). Max ignores NaN (fmax); an all-NaN input gives
 * NaN. */
int mh_reduce_sum_abs(mh_ctx *ctx, const float *x_dev, int64_t n, float *out_dev, void *stream);
int mh_reduce_max(mh_ctx *ctx, const float *x_dev, int64_t n, float *out_dev, void *stream);

/* ------------------------------------------------------------------ Frangi */

/* One separable 3-pass filter (axis 0, then 1, then 2) with mirror boundary
 * of period 2(n-1) and true-convolution tap order. Replaces
 * GaussianFilter3D.execute + convolve_naive.cl axis0/1/2 (processing/
 * qfrangi_gpu.py:91-155, clcode/convolve_naive.cl:25-166). taps*_host hold
 * 2*r+1 floats; result in out_dev; swap_dev is N floats of scratch. Every dim
 * used with r > 0 must be >= 2. */
int mh_gauss_filter(mh_ctx *ctx, const float *in_dev, float *out_dev, float *swap_dev,
                    const int64_t dims[3], const float *taps0_host, int r0,
                    const float *taps1_host, int r1, const float *taps2_host, int r2,
                    void *stream);

/* All six scale-normalised Gaussian-derivative Hessian entries at one scale,
 * written AoS x6 [H00, H11, H22, H01, H02, H12]. Replaces Hessian.execute
 * (processing/qfrangi_gpu.py:208-271). taps_host is [axis][order] packed:
 * for axis a (0..2) and derivative order o (0..2) the 2*radius[a]+1 taps start
 * at taps_host + (3*a + o) * (2*max_radius + 1), max_radius = max(radius).
 * work_dev must hold mh_hessian_work_floats(dims) floats. */
int64_t mh_hessian_work_floats(const int64_t dims[3]);
int mh_gauss_hessian(mh_ctx *ctx, const float *img_dev, const int64_t dims[3],
                     const float *taps_host, const int radius[3], float *hess6_dev,
                     float *work_dev, void *stream);

/* 3x3 symmetric Jacobi eigen-decomposition per voxel, in place allowed.
 * in: AoS x6 Hessian; out: AoS x6 [l1, l2, l3, phi, theta, psi] sorted by
 * ascending |l|, frame squeezed (eigenvectors in columns). NaN x6 if not
 * converged within maxiter sweeps. If norm_max_dev != NULL, the maximum over
 * voxels of ||lambda||_2 (NaN skipped) is max-accumulated into *norm_max_dev
 * (initialise it to 0). Replaces Eigendecomposition.execute + qeigs.cl eigs
 * and the calculateNorm/redMax part of VesselnessUpdate (processing/
 * qfrangi_gpu.py:274-332, 429-433; clcode/qeigs.cl:146-256;
 * clcode/frangi.cl:107-115). */
int mh_eigs(mh_ctx *ctx, const float *hess6_dev, float *eig6_dev, int64_t n, int maxiter,
            float *norm_max_dev, void *stream);

/* max over voxels of ||lambda||_2 of an AoS x6 eigen buffer (NaN skipped;
 * 0 if every voxel is NaN) into *norm_max_dev. Replaces calculateNorm + the
 * redMax reduction (clcode/frangi.cl:107-115, clcode/max.cl:25-58). */
int mh_eig_norm_max(mh_ctx *ctx, const float *eig6_dev, int64_t n, float *norm_max_dev,
                    void *stream);

/* Frangi response of one scale and running maximum with strict '>' carrying
 * the winner's squeezed frame. c = c_value if c_value >= 0, else
 * c = 0.5 * (*norm_max_dev). Replaces updateVesselness (clcode/frangi.cl:46-92,
 * processing/qfrangi_gpu.py:383-439). */
int mh_update_vesselness(mh_ctx *ctx, float *val_dev, float *quat3_dev, const float *eig6_dev,
                         int64_t n, float alpha, float beta, float c_value,
                         const float *norm_max_dev, int bright_objects, void *stream);

/* -------------------------------------------------------------- directions */

/* Ridge mask: resp > threshold and resp > max of 8 trilinear samples in the
 * cross-section plane. threshold_f must be the largest float <= the double
 * threshold (so that the float comparison equals the double one). Replaces
 * RawRidgeDetector.execute + detectRidges (processing/ridge_detection.py:
 * 12-77, clcode/detect_ridges.cl:34-102). */
int mh_detect_ridges(mh_ctx *ctx, float *ridge_dev, const float *resp_dev,
                     const float *qeig3_dev, const int64_t dims[3], float threshold_f,
                     void *stream);

/* v1 = column 0 of the squeezed frame, AoS x3, in place allowed. Replaces
 * extractMainDirections (clcode/adjust_directions.cl:29-43). */
int mh_main_dirs(mh_ctx *ctx, float *dirs3_dev, const float *qeig3_dev, int64_t n, void *stream);

/* For every voxel, nearest ridge point by scaled squared distance (lowest
 * index wins ties), then flip the voxel's direction if its dot product with
 * that ridge direction is negative. coords_dev: uint16 R x 3, rdirs_dev:
 * float R x 3. With work_dev (mh_nonridge_work_bytes(dims) bytes), an
 * isotropic voxel size and many ridge points, the nearest point comes from an
 * exact separable distance transform (same result as the scan); otherwise
 * every ridge point is scanned. Replaces adjustNonRidgeSigns
 * (clcode/adjust_directions.cl:182-246, processing/adjust_directions.py:419-446). */
int64_t mh_nonridge_work_bytes(const int64_t dims[3]);
int mh_nonridge_signs(mh_ctx *ctx, float *dirs3_dev, const int64_t dims[3],
                      const uint16_t *coords_dev, const float *rdirs_dev, int64_t nridge,
                      const float scaling_host[3], void *work_dev, int64_t work_bytes,
                      void *stream);

/* Direction -> orthonormal frame (v, v2, v x v2) -> squeezed quaternion,
 * AoS x3, in place allowed; a zero vector yields the identity frame.
 * Replaces TransformationCompleter.complete + completeTransformations
 * (processing/adjust_directions.py:486-504, clcode/adjust_directions.cl:68-100). */
int mh_complete_frames(mh_ctx *ctx, float *out3_dev, const float *dirs3_dev, int64_t n,
                       void *stream);

/* Host (CPU) hysteresis: out = candidate voxels 26-connected through
 * candidates to a seed voxel (seed = sure && cand). Replaces
 * hysteresis_threshold (processing/ridge_detection.py:80-131). Blocks. */
int mh_hysteresis_host(const uint8_t *cand_host, const uint8_t *sure_host, uint8_t *out_host,
                       const int64_t dims[3]);

/* The ridge's hysteresis and its 26-connected segments on the device:
 * union-find over the candidate voxels (cand != 0) and their 26 neighbours.
 * With use_upper, a segment is kept when one of its voxels has resp >= upper
 * (the flood fill from the seeds of hysteresis_threshold,
 * processing/ridge_detection.py:80-131, with seeds = candidate and response
 * >= the upper threshold, adjust_directions.py:258-288); without, every
 * segment is kept. ridge (may alias cand) receives the kept voxels as 1.f,
 * labels (int32, N) the segment of every kept voxel as its smallest voxel
 * index (so segments in increasing label order are scipy.ndimage.label's
 * numbering, as the cleaning at adjust_directions.py:290-339 uses) and -1
 * elsewhere. work: mh_ridge_segments_work_bytes(dims) bytes. */
int64_t mh_ridge_segments_work_bytes(const int64_t dims[3]);
int mh_ridge_segments(mh_ctx *ctx, const float *cand_dev, const float *resp_dev, float upper,
                      int use_upper, const int64_t dims[3], float *ridge_dev, int *labels_dev,
                      void *work_dev, int64_t work_bytes, void *stream);

/* The direction adjustment after the ridge's hysteresis, on the device:
 * dirs_dev holds every voxel's main direction (AoS x3, mh_main_dirs) and
 * receives the adjusted directions; ridge_dev and labels_dev are
 * mh_ridge_segments' outputs. The ridge's minimum spanning forest (the
 * 26-neighbour pairs by squared length, then point indices), the signs along
 * it, with clean the ceil(3 ln(e + S)) segments of longest summed signed
 * direction kept, the minimum spanning tree between the kept segments, the
 * signs from the smallest kept point along it, and every voxel's sign from
 * its nearest kept ridge point. kept_dev (N floats) receives the kept ridge
 * (1.f). stats_host (8 int64): ridge points, segments, kept segments, kept
 * points, zero-product forest edges, microseconds in the cleaning's full-tree
 * walk (0 when the bounds decide), microseconds in the call, forest rounds. Replaces
 * AdvancedAdjuster's ridge steps (processing/adjust_directions.py:213-256,
 * :290-339, :365-446) for an isotropic voxel size that is a power of two;
 * the results equal the host construction (mh_ridge_forest_host,
 * mh_forest_pieces_host, mh_ridge_tree_from_forest_host, mh_tree_signs_host,
 * mh_nonridge_signs) bit for bit. Returns MH_DEFER, with dirs_dev unchanged,
 * where that cannot be guaranteed or the host's own rule differs: another
 * voxel size, 2^30 voxels or more or an axis longer than 16384 (30-bit point
 * indices), at most 256 ridge or kept points (the dense tree), a NaN
 * direction product, segment lengths the bounds do not separate, more than
 * 65536 kept points outside the largest kept segment, more than 4096 kept
 * segments. Blocks until done. */
int mh_adjust_ridge_device(mh_ctx *ctx, float *dirs_dev, const float *ridge_dev,
                           const int *labels_dev, const int64_t dims[3],
                           const double voxel_size_host[3], int clean, float *kept_dev,
                           int64_t *stats_host, void *stream);

/* Host: a minimum spanning tree of the complete Euclidean graph over n points
 * (pts_host: n x 3 doubles, e.g. voxel-size scaled ridge coordinates), as
 * n - 1 index pairs in edges_host (2(n - 1) int64). r0 is the longest
 * neighbour distance (the scaled 26-neighbourhood diagonal): Kruskal on the
 * pairs within r0, then Boruvka rounds by k-d tree search. Ties in length go
 * to the lexicographically smallest index pair. Replaces the pdist +
 * squareform + minimum_spanning_tree of _adjust_ridge_directions
 * (processing/adjust_directions.py:365-379) without its O(n^2) matrix. */
int mh_spanning_tree_host(const double *pts_host, int64_t n, double r0, int64_t *edges_host);
/* Host: the same tree for n distinct voxels (coords_host: n x 3 int64 voxel
 * indices inside dims) at scaled positions coords * voxel_size; the short
 * edges come from a dense neighbour lookup instead of a k-d tree query. */
int mh_ridge_tree_host(const int64_t *coords_host, int64_t n, const int64_t dims[3],
                       const double voxel_size[3], int64_t *edges_host);
/* Host: the first phase of mh_ridge_tree_host alone, the spanning forest of
 * the neighbour pairs shorter than twice the smallest voxel size; returns the
 * number of edges written (< 0 on error). The direction adjustment's cleaning
 * scores need no more than this (processing/adjust_directions.py:290-339). */
int64_t mh_ridge_forest_host(const int64_t *coords_host, int64_t n, const int64_t dims[3],
                             const double voxel_size[3], int64_t *edges_host);
/* Host: mh_ridge_tree_host with its first phase given (replaces the same
 * pdist + minimum_spanning_tree, processing/adjust_directions.py:365-379, for
 * the cleaned ridge): forest_host (nf x 2) is mh_ridge_forest_host's output
 * for these points, or that output for a superset restricted to a union of
 * its trees and renumbered in order, as after the cleaning keeps whole
 * segments (adjust_directions.py:290-339). edges_host receives exactly
 * mh_ridge_tree_host's edges. */
int mh_ridge_tree_from_forest_host(const int64_t *coords_host, int64_t n, const int64_t dims[3],
                                   const double voxel_size[3], const int64_t *forest_host,
                                   int64_t nf, int64_t *edges_host);
/* Host: the pieces of a forest between its zero-product edges and signed
 * directions within each piece (every non-zero edge joins points whose
 * signed directions have a non-negative product; a piece's smallest point
 * keeps its direction), by a union-find carrying sign parities. pieces_host:
 * n labels 0 .. P-1; counts_host[0]: zero products, counts_host[1]: NaN
 * products (then the tree walk itself is needed). Feeds the cleaning scores
 * of _do_clean_ridge (processing/adjust_directions.py:290-339), which only
 * need the walk's signs up to one flip per piece; without zero or NaN
 * products the signs are mh_tree_signs_host's from root 0 bit for bit.
 * n < 2^31. */
int mh_forest_pieces_host(const float *dirs_host, int64_t n, const int64_t *edges_host,
                          int64_t nedges, float *out_host, int32_t *pieces_host,
                          int64_t counts_host[2]);
/* Host: directions (n x 3 floats) with signs propagated along a tree from
 * root: a child keeps its direction when its dot product with its adjusted
 * parent is >= 0, else flips (NaN flips). Trees of a forest not reached from
 * root start at their smallest index, which keeps its direction. Replaces
 * the traversal of _adjust_ridge_directions (processing/adjust_directions.py:384-417),
 * computed without walking the points: parity sets within the pieces between
 * zero and NaN products, a walk over the pieces for their anchors. n < 2^31. */
int mh_tree_signs_host(const float *dirs_host, int64_t n, const int64_t *edges_host, int64_t nedges,
                       int64_t root, float *out_host);

/* --------------------------------------------------------------------- GVF */

typedef struct {
    float mu;      /* regularisation weight */
    float dt;      /* time step, gradient_vector_flow.py:58-65 */
    float rss[3];  /* 1 / voxel_size^2 per axis */
} mh_gvf_params;

/* r = |v|^2, v SoA x3. Replaces the response kernel (clcode/gvf.cl:246-253). */
int mh_gvf_response(mh_ctx *ctx, float *r_dev, const float *v3_dev, int64_t n, void *stream);

/* `iters` explicit Euler steps g <- g - dt (r (g - v) - mu lap(g)) with
 * mirrored neighbours at the faces (gvf.cl:44-173), ping-ponging between
 * g3_dev and swap3_dev (both SoA x3); the result always ends in g3_dev.
 * If mag_dev != NULL, *mag_dev receives sum over voxels of |delta| of the
 * last step (the caller divides by N); work_dev (mh_gvf_work_floats(dims)
 * floats) then holds the partial sums. Every dim must be >= 2. Replaces one
 * chunk of GvfGpu._execute (processing/gradient_vector_flow.py:179-232). */
int64_t mh_gvf_work_floats(const int64_t dims[3]);
int mh_gvf(mh_ctx *ctx, float *g3_dev, float *swap3_dev, const float *v3_dev, const float *r_dev,
           const int64_t dims[3], const mh_gvf_params *params, int iters, float *mag_dev,
           float *work_dev, void *stream);
/* How mh_gvf runs its steps on this context (no counterpart in the
 * reference): 0 (default) two steps per launch, blocked along axis 0, for
 * volumes of 2^20 voxels or more; 1 one step per launch; 2 two per launch
 * whenever the volume allows. Results do not depend on it, bit for bit (the
 * magnitude sums of the checks may differ in their summation order). */
int mh_gvf_set_sweep(mh_ctx *ctx, int mode);

/* Normalise non-zero vectors of g (SoA x3) to unit length
 * (clcode/gvf.cl:215-232). */
int mh_gvf_normalize(mh_ctx *ctx, float *g3_dev, int64_t n, void *stream);

/* ---------------------------------------------------------------- max-flow */

typedef struct {
    float rcc;        /* 1 / (cc * min_s^2) */
    float rs[3];      /* 1 / s_k (voxel size s) */
    float ss[3];      /* steps * min_s^2 / s_k */
    float sqev[3];    /* sqrt of the constant eigenvalues (const_evals) */
    int const_evals;  /* 1: eigs AoS x3 squeezed; 0: AoS x6 [sqrt(l) x3, squeezed x3] */
    int num_iter;
    int update;       /* criterion every `update` iterations, at i % update == 0 */
    float error_bound;
} mh_cmf_params;

typedef struct {
    int iterations;   /* update_u applications performed */
    int stopped;      /* 0: ran num_iter; 1: NaN criterion; 2: criterion < error_bound */
    int ncrit;        /* number of criterion values written */
    float last_crit;
} mh_cmf_result;

/* Device workspace needed by mh_cmf_solve for a volume of n voxels. */
int64_t mh_cmf_workspace_bytes(int64_t n);

/* One ATV continuous max-flow solve. u_dev holds u0 on entry and the result
 * on exit. labels_dev (uint8, 1 = foreground seed, 2 = background seed) may be
 * NULL for the unsupervised cut. crit_host (optional, ceil(num_iter/update)
 * floats) receives every criterion sum|err| / (N * rcc). Blocks until the
 * solve has finished. Replaces CmfCutQAnisotropicGpu._execute with update_p,
 * update_u, update_tmp and the sum_abs criterion (processing/
 * continuous_cut_anisotropic_gpu.py:198-344, clcode-anisotropic/
 * update_qp.cl:150-316). */
int mh_cmf_solve(mh_ctx *ctx, float *u_dev, const float *eigs_dev, const float *alpha_dev,
                 const float *cs_dev, const float *ct_dev, const uint8_t *labels_dev,
                 const int64_t dims[3], const mh_cmf_params *params, void *work_dev,
                 int64_t work_bytes, mh_cmf_result *result_host, float *crit_host,
                 void *stream);

/* mh_cmf_solve without a workspace argument and without any device
 * allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The workspace lives in a static device
 * array of 64 MiB per GPU, held from the first mh_create on that GPU until the
 * process ends, so volumes are limited to MH_CMF_NOALLOC_MAX_VOXELS voxels
 * (larger ones are rejected with MH_ERR_ARG). Same arguments, results (bit
 * for bit) and blocking as mh_cmf_solve; the context's device must be
 * current. The method is the one mh_cmf_solve tuned for these dims on this
 * context, else the untuned default; it never tunes and never captures a
 * graph. Every call on a GPU shares its array: calls that overlap on one GPU
 * (other streams or threads) race and must be serialised by the caller. */
#define MH_CMF_NOALLOC_MAX_VOXELS (1 << 20)
int mh_cmf_solve_noalloc(mh_ctx *ctx, float *u_dev, const float *eigs_dev, const float *alpha_dev,
                         const float *cs_dev, const float *ct_dev, const uint8_t *labels_dev,
                         const int64_t dims[3], const mh_cmf_params *params,
                         mh_cmf_result *result_host, float *crit_host, void *stream);

/* ------------------------------------------------------ batched max-flow */

/* Many independent unsupervised constant-eigenvalue solves on volumes that
 * share the frame field, the terminal capacities and |grad I|^2. Solve s uses
 * alpha = w[s] * exp(-gms / decay[s]^2) (phantom_experiment.py:359), computed
 * on the fly. Solve s owns u_dev + s*N (u0 on entry, result on exit).
 * No counterpart in the reference, which runs these solves one by one
 * (phantom_experiment.py:350-368); each solve equals mh_cmf_solve on the same
 * inputs within FP32 rounding. Blocks. */
int64_t mh_cmf_batched_workspace_bytes(int64_t n, int nsolves);
int mh_cmf_solve_batched(mh_ctx *ctx, float *u_dev, int nsolves, const float *w_host,
                         const float *decay_host, const float *eigs_dev, const float *gms_dev,
                         const float *cs_dev, const float *ct_dev, const int64_t dims[3],
                         const mh_cmf_params *params, void *work_dev, int64_t work_bytes,
                         mh_cmf_result *results_host, void *stream);
/* Schedule of mh_cmf_solve_batched on this context (no counterpart in the
 * reference): `group` solves (1..8) share one thread per voxel, and a group
 * advances `visit` iterations before the next group takes its turn (visit >=
 * num_iter runs the groups one after another). 0 selects the default: groups
 * of 4 run one after another, which keeps a 128^3 group's state within the
 * last-level cache; when a multi-level sweep runs the solves, groups of 2,
 * which the 1024-thread tile shapes advance two solves per thread (group 1:
 * one per thread). Results do not depend on the schedule, bit for bit. */
int mh_cmf_batched_set_schedule(mh_ctx *ctx, int group, int visit);
/* Multi-level sweep of mh_cmf_solve_batched's groups (no counterpart in the
 * reference). 0 (default): when the groups run one after another and the
 * single-solve tuner picks a sweep for these dims (it is run on a probe copy
 * of solve 0 if it has not been for these dims), the solves run through a
 * sweep in pairs (two per thread) where the volume allows: the first group's
 * passes time the single-solve choice and each pair-capable tile shape, and
 * the fastest runs the rest (kept per dims, mh_cmf_batched_sweep_choice);
 * otherwise groups of 4 use the per-iteration kernels, which share the frame
 * and capacities in registers. 1: never sweep. 2: sweep the groups whenever
 * eligible, with the tile shape forced by mh_cmf_set_sweep(mode 2) or the
 * tuned one. Results do not depend on it, bit for bit. */
int mh_cmf_batched_set_sweep(mh_ctx *ctx, int mode);
/* Multi-level sweep of mh_cmf_solve on this context (no counterpart in the
 * reference): unsupervised constant-eigenvalue solves may advance two, three
 * or four iterations per pass over the state (temporal blocking along axis
 * 0; tile shapes 1-5 two levels, 6-7 three, 8-9 four). mode 0
 * (default) tunes: the second mh_cmf_solve of a volume shape on this context
 * (>= 96 iterations; the first runs the untuned default, so shapes met once
 * pay nothing for timing) times the plain kernels and the sweep tile shapes
 * on its own iterations and the context keeps the fastest for that shape;
 * mh_cmf_solve_batched tunes on a probe at once; mode 1 never sweeps; mode 2 always
 * sweeps when the solve is eligible, with tile (1-based index of the built-in
 * tile shapes, 0 the first) and zchunk (planes per work-group, 0: the plane
 * chunk a cost model of the launch favours for that tile shape).
 * Results do not depend on any of them, bit for bit. */
int mh_cmf_set_sweep(mh_ctx *ctx, int mode, int tile, int zchunk);
/* Graph replay in mh_cmf_solve and the resident solves (no counterpart in the
 * reference): when the per-iteration kernels run, one criterion cycle (update
 * iterations starting at a criterion iteration) is captured as a HIP graph
 * and replayed for the following cycles: on 1, off 0 (the default: on gfx950
 * launched kernels were measured as fast or faster). The graph holds kernel
 * launches only and runs on a stream of the context's own, ordered after and
 * before the caller's stream. Results do not depend on it, bit for bit. */
int mh_cmf_set_graphs(mh_ctx *ctx, int on);
/* The method the tuner chose on this context for volumes of these dims:
 * returns 1 and sets tile (0: the per-iteration kernels, else the 1-based tile
 * shape) and zchunk, or returns 0 when nothing was tuned for them. */
int mh_cmf_sweep_choice(mh_ctx *ctx, const int64_t dims[3], int *tile, int *zchunk);
/* The method mh_cmf_solve_batched's pair tuner chose on this context for
 * volumes of these dims (no counterpart in the reference): returns 1 and sets
 * tile (1-based tile shape) and zchunk, or returns 0 when no pairs were timed
 * for them (the batched solver then uses the single-solve choice). */
int mh_cmf_batched_sweep_choice(mh_ctx *ctx, const int64_t dims[3], int *tile, int *zchunk);

/* Packed ragged batch: volume v has dims vdims_host[3v..3v+2] and its fields
 * start at element offset voff_host[v] in every packed array (u, eigs*3,
 * alpha, cs, ct). No counterpart in the reference. Blocks. */
int64_t mh_cmf_ragged_workspace_bytes(int64_t total_voxels, int nvol);
int mh_cmf_solve_ragged(mh_ctx *ctx, float *u_dev, const float *eigs_dev, const float *alpha_dev,
                        const float *cs_dev, const float *ct_dev, int nvol,
                        const int64_t *vdims_host, const int64_t *voff_host,
                        const mh_cmf_params *params, void *work_dev, int64_t work_bytes,
                        mh_cmf_result *results_host, void *stream);

/* ---------------------------------------------- fields around the max-flow */

/* (cs, ct) = rq * (max(r, 0), max(-r, 0)), r = log((N(x;1,s^2) + eps) /
 * (N(x;0,s^2) + eps)), rq = 1 / log((p_hat + eps) / eps); FP32. Replaces the
 * host weights of phantom_experiment.py:280-301. */
int mh_terminal_capacities(mh_ctx *ctx, const float *img_dev, int64_t n, float noise_level,
                           float eps, float *cs_dev, float *ct_dev, void *stream);
/* |grad I|^2 with central differences inside, one-sided at the faces
 * (numpy.gradient; phantom_experiment.py:345). */
int mh_grad_mag_sq(mh_ctx *ctx, const float *img_dev, const int64_t dims[3], float *gms_dev,
                   void *stream);
/* alpha = w exp(-gms / decay^2) (phantom_experiment.py:359). */
int mh_nonterminal_weight(mh_ctx *ctx, const float *gms_dev, int64_t n, float weighting,
                          float decay, float *alpha_dev, void *stream);
/* Min-max normalisation to [0, 1] (0 everywhere for a constant input);
 * work_dev holds mh_normalize_work_floats(n) floats (helpers/misc.py:63-114,
 * without percentiles). */
int64_t mh_normalize_work_floats(int64_t n);
int mh_minmax_normalize(mh_ctx *ctx, const float *x_dev, int64_t n, float *out_dev,
                        float *work_dev, void *stream);
/* v3 (SoA x3) = dirs3 (AoS x3) * resp: the GVF input of
 * phantom_experiment.py:335. */
int mh_weighted_dirs(mh_ctx *ctx, const float *dirs3_dev, const float *resp_dev, int64_t n,
                     float *v3_dev, void *stream);
/* SoA x3 -> AoS x3 (the GVF result layout, gradient_vector_flow.py:241-253). */
int mh_soa_to_aos3(mh_ctx *ctx, const float *soa_dev, int64_t n, float *aos_dev, void *stream);
/* Synthetic helix phantom for load generation: tube of radius r_tube around a
 * helix of radius r_helix and pitch `pitch` along axis 0, plus N(0, sigma^2)
 * noise from a counter-based hash of (seed, voxel). truth_dev (u8) and
 * tangent_dev (AoS x3: unit tangent of the closest centre-line point, in axis
 * order) may be NULL. No counterpart in the reference (its tvtk voxeliser,
 * helpers/helix_phantom.py, draws a different, toric helix). */
int mh_helix_phantom(mh_ctx *ctx, const int64_t dims[3], float r_tube, float r_helix,
                     float pitch, float sigma, uint64_t seed, float *img_dev, uint8_t *truth_dev,
                     float *tangent_dev, void *stream);

/* ------------------------------------------- resident jobs (continuous batching) */

/* u0 = img / (max(img) + FLT_MIN) in FP32, a NaN in img giving a NaN maximum
 * (replaces the host initialisation at continuous_cut_anisotropic_gpu.py:282-283
 * for device-resident inputs). work: mh_initial_u_work_floats(n) floats. */
int64_t mh_initial_u_work_floats(int64_t n);
int mh_initial_u(mh_ctx *ctx, const float *img_dev, int64_t n, float *u_dev, float *work_dev,
                 void *stream);

/* Dice counts of nsolves labellings u (nsolves x n, u >= 0.5 is foreground)
 * against one mask truth (n bytes, nonzero = foreground): counts_dev[3 s + k]
 * = |S|, |S & T|, |T| for solve s, exact. The host forms
 * 2 |S & T| / (|S| + |T|) as fast_dice_for does (phantom_experiment.py:36-47)
 * without copying the labellings. work: mh_dice_work_bytes(n, nsolves). */
int64_t mh_dice_work_bytes(int64_t n, int nsolves);
int mh_dice_counts(mh_ctx *ctx, const float *u_dev, int nsolves, int64_t n, const uint8_t *truth_dev,
                   void *work_dev, int64_t work_bytes, unsigned long long *counts_dev, void *stream);

/* A plain solve advanced in steps (no counterpart in the reference, whose
 * solves run start to finish, continuous_cut_anisotropic_gpu.py:276-343): the
 * serving loop keeps large cuts resident this way and advances each by a
 * chunk of iterations per turn, with the same methods as mh_cmf_solve
 * (per-iteration kernels or multi-level sweep), so the result is mh_cmf_solve's
 * bit for bit. Unsupervised. work_dev: mh_cmf_workspace_bytes(n) bytes,
 * 256-byte aligned; every pointer is a device pointer the caller owns. */
typedef struct {
    int64_t dims[3];
    float *u_dev;            /* u0 before init, the labelling after finish */
    const float *eigs_dev;   /* squeezed frames AoS x3 (const_evals) or AoS x6 */
    const float *alpha_dev;
    const float *cs_dev;
    const float *ct_dev;
    void *work_dev;
    mh_cmf_params params;
    int iter;                /* iterations enqueued so far (set by the calls) */
    int cur;                 /* buffer set holding the state (set by the calls) */
    int swept;               /* a multi-level sweep has run (set by the calls) */
    int method;              /* -2 not chosen yet, -1 per-iteration kernels, else tile shape */
    int zchunk;
} mh_cmf_resident;

/* Enqueue the initialisation; sets iter = 0 and the method to "not chosen". */
int mh_cmf_resident_init(mh_ctx *ctx, mh_cmf_resident *r, void *stream);
/* Enqueue up to `iters` more iterations (never past num_iter). The method is
 * chosen at the first step: tuned for these dims on this context, tuned now
 * on this solve's iterations when allow_tune (blocks while timing), or the
 * default (the three-level sweep with tile shape 7 for volumes of at least
 * 800,000 voxels, where it measured faster, else the per-iteration kernels). */
int mh_cmf_resident_step(mh_ctx *ctx, mh_cmf_resident *r, int iters, int allow_tune, void *stream);
/* Blocks; *stopped = 0 running, 1 NaN criterion, 2 converged. */
int mh_cmf_resident_status(mh_ctx *ctx, const mh_cmf_resident *r, int *stopped, void *stream);
/* Blocks; puts the final u in u_dev and reports like mh_cmf_solve. */
int mh_cmf_resident_finish(mh_ctx *ctx, mh_cmf_resident *r, mh_cmf_result *res, float *crit_host,
                           void *stream);

/* Dice of the largest face-connected component of u >= threshold against the
 * mask truth (the experiment's dice_for / select_largest, phantom_experiment.py:
 * 50-103, with scipy's default structure): out_dev[0] = |L| (-1 if there is no
 * foreground), [1] = |L & T|, [2] = |T|, [3] = the smallest voxel index of L.
 * Ties of the largest size go to the component whose first voxel comes first
 * in raster order, as there. labels_dev (int32 per voxel, optional) receives
 * every foreground voxel's component as its smallest voxel index, -1 for
 * background. work: mh_largest_component_work_bytes(dims). */
int64_t mh_largest_component_work_bytes(const int64_t dims[3]);
int mh_largest_component_dice(mh_ctx *ctx, const float *u_dev, float threshold,
                              const uint8_t *truth_dev, const int64_t dims[3], int *labels_dev,
                              void *work_dev, int64_t work_bytes, long long *out_dev, void *stream);

/* One unsupervised constant-eigenvalue solve whose state stays on the device
 * between calls, so jobs of different sizes can join and leave a running set
 * at iteration-chunk boundaries. All pointers are device pointers owned by
 * the caller; state_dev holds mh_cmf_job_state_bytes(n, params) bytes. No
 * counterpart in the reference (each of its solves runs alone, start to
 * finish, continuous_cut_anisotropic_gpu.py:276-343); a job run in chunks
 * computes the same iterations as mh_cmf_solve. */
typedef struct {
    int64_t dims[3];
    float *u_dev;            /* u0 before mh_cmf_jobs_init, the labelling after */
    const float *eigs_dev;   /* squeezed frames AoS x3 */
    const float *alpha_dev;
    const float *cs_dev;
    const float *ct_dev;
    void *state_dev;
    mh_cmf_params params;    /* const_evals must be 1 */
    int iter;                /* iterations done so far (updated by mh_cmf_jobs_step) */
    int stopped;             /* 0 running, 1 NaN criterion, 2 converged (updated) */
} mh_cmf_job;

int64_t mh_cmf_job_state_bytes(int64_t n, const mh_cmf_params *params);
/* Initialise the state of njobs jobs (frames expanded, hp = 0, pt = min(cs, ct),
 * tmp = -rcc u0); sets iter = 0. Enqueues only. */
int mh_cmf_jobs_init(mh_ctx *ctx, mh_cmf_job *jobs_host, int njobs, void *stream);
/* Advance every job by min(iters, num_iter - iter) iterations (a stopped job
 * does nothing) in one launch sequence covering all jobs; updates iter and
 * stopped. Blocks. At most 1024 jobs per call. */
int mh_cmf_jobs_step(mh_ctx *ctx, mh_cmf_job *jobs_host, int njobs, int iters, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MICCAI_HIP_H */
