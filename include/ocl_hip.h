/*
 * ocl_hip.h -- C ABI of the MI355X (gfx950) batch drop-cutter / push-cutter.
 *
 * This is the drop-in boundary behind OpenCAMLib's C++ operation classes:
 * the host classes in opencamlib_amd/host (ocl::BatchDropCutter,
 * ocl::BatchPushCutter, ocl::STLSurf, the MillingCutter family) call these
 * entry points, and so does the Python binding.  Plain pointers and sizes
 * only; no C++ or torch types.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative OCL_E_* code on
 *     failure; ocl_hip_last_error() returns the message (thread-local).
 *   - Host-pointer entry points copy in and out and synchronise the stream.
 *     *_device entry points take device pointers and are asynchronous on the
 *     given stream (NULL = the default stream).
 *   - A mesh handle owns its device memory.  Handles may be used from several
 *     threads; calls on one handle are serialised by the caller (the reference
 *     Operation objects are not re-entrant either).
 *   - Triangle indices returned are positions in the input triangle array.
 */
#ifndef OCL_HIP_H
#define OCL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCL_OK 0
#define OCL_E_ARG (-1)
#define OCL_E_HIP (-2)
#define OCL_E_NOMEM (-3)
#define OCL_E_UNSUPPORTED (-4)

/* cutter kinds (src/cutters: cylcutter, ballcutter, bullcutter, conecutter, compositecutter) */
#define OCL_CUTTER_CYL 0
#define OCL_CUTTER_BALL 1
#define OCL_CUTTER_BULL 2
#define OCL_CUTTER_CONE 3
#define OCL_CUTTER_COMPOSITE 4

/* precision modes of the drop/push kernels */
#define OCL_PREC_EXACT64 0 /* reference-order FP64 on every pair */
#define OCL_PREC_FP64 1    /* FP64 selection kernel + FP64 exact refine */
#define OCL_PREC_FP32 2    /* FP32 selection kernel only (no refine) */
#define OCL_PREC_MIXED 3   /* FP32 selection + FP64 refine + FP64 escalation */
/* push only, OR-ed into precision: evaluate only (fiber, triangle) pairs whose
 * search boxes overlap, instead of the reference's KD-search superset */
#define OCL_PUSH_EXACT_FILTER 0x100

/* One simple cutter: CylCutter(d, L), BallCutter(d, L), BullCutter(d, r, L),
 * ConeCutter(d, angle, L) (cylcutter.cpp:39, ballcutter.cpp:38,
 * bullcutter.cpp:38, conecutter.cpp:39). */
typedef struct {
    int32_t type;
    double diameter;
    double length;
    double corner_radius; /* Bull */
    double angle;         /* Cone half-angle, radians */
} ocl_cutter_simple;

/* Cutter descriptor.  For OCL_CUTTER_COMPOSITE the sub-cutters with their
 * radius limit, height limit and z offset follow CompositeCutter::addCutter
 * (compositecutter.cpp:43-53); `length` is the composite length. */
typedef struct {
    ocl_cutter_simple c;
    int32_t nsub;
    ocl_cutter_simple sub[4];
    double radius_limit[4];
    double height_limit[4];
    double zoffset[4];
} ocl_cutter_desc;

typedef struct ocl_mesh ocl_mesh;

/* Counters of one call (stats may be NULL). */
typedef struct {
    uint64_t points;        /* CL points (drop) or fibers (push) */
    uint64_t pairs;         /* (point, triangle) pairs passing the overlap test */
    uint64_t escalated;     /* points the first selection pass could not decide */
    uint64_t escalated_exact; /* of those, points finished by the exact FP64 pass */
    uint64_t brent_iters;   /* Bull edge Brent iterations (FP64) */
    uint64_t intervals;     /* push: output intervals */
    double kernel_ms;       /* device time of the compute kernels */
    double select_ms;       /* device time of the first selection kernel (the
                               whole compute for exact64 and composites) */
    uint64_t wave_evals;    /* wavefront evaluations of the per-pair contact
                               code in the first selection kernel (pairs /
                               (64 wave_evals) = lane utilisation) */
    uint64_t skipped;       /* push: candidate pairs left out because they
                               cannot change their fiber (not in pairs) */
} ocl_stats;

/* Mesh: ntri triangles, 9 doubles each (p0, p1, p2).  bucket = the KD-tree
 * bucket size of Operation::setBucketSize (default 1), which fixes the
 * reference's triangle evaluation order (kdtree.hpp:141-210).
 * Replaces BatchDropCutter::setSTL (batchdropcutter.cpp:54-63) and
 * BatchPushCutter::setSTL (batchpushcutter.cpp:55-70). */
int ocl_hip_mesh_create(const double* tri_xyz, size_t ntri, unsigned bucket, ocl_mesh** out);
int ocl_hip_mesh_destroy(ocl_mesh* mesh);

/* Host only (no device work): STLReader::read_from_file
 * (stlreader.cpp:57-150), binary or text STL, float32 coordinates promoted
 * to double, 9 doubles per triangle in file order.  Sets *ntri_out to the
 * triangle count and fills tri_xyz_out when cap >= that count (call with
 * NULL first to size the array).  OCL_E_ARG if the file cannot be read. */
int ocl_hip_stl_read(const char* path, double* tri_xyz_out, size_t cap, size_t* ntri_out);

/* An STL file straight to a device mesh (STLReader + setSTL without the
 * STLSurf triangle list): ocl_hip_stl_read then ocl_hip_mesh_create. */
int ocl_hip_mesh_create_stl(const char* path, unsigned bucket, ocl_mesh** out);
size_t ocl_hip_mesh_size(const ocl_mesh* mesh);

/* Host only (no device work): the reference KD-tree's leaf order for a
 * triangle array, order_out[rank] = triangle index.  dimset 0 = XY (drop,
 * KDTree::setXYDimensions), 1 = YZ (X fibers, setYZDimensions), 2 = XZ
 * (Y fibers, setXZDimensions); kdtree.hpp:117-268.  It is the order in which
 * KDTree::search returns overlapping triangles (kdtree.hpp:273-304), used for
 * getTrianglesUnderCutter / getOverlapTriangles (batchdropcutter_py.hpp:46-55,
 * batchpushcutter_py.hpp:59-84).  accept_out (optional, 4 doubles per rank):
 * the search reaches rank k for a query box with, per search axis
 * a0 = first axis, a1 = second, max_a0 >= A0, min_a0 <= B0, max_a1 >= A1,
 * min_a1 <= B1, stored (A0, B0, A1, B1). */
int ocl_hip_kd_order(const double* tri_xyz, size_t ntri, int dimset, unsigned bucket, uint32_t* order_out,
                     double* accept_out);

/* Batch drop-cutter: replaces BatchDropCutter::run / dropCutter5
 * (batchdropcutter.cpp:187-231).  cl_xyz: n x (x, y, z0).  Outputs (any may
 * be NULL except z_out): lifted z, CC point xyz, CCType, index of the contact
 * triangle (UINT32_MAX if none), contact feature code. */
int ocl_hip_drop_batch(ocl_mesh* mesh, const ocl_cutter_desc* cutter, const double* cl_xyz,
                       size_t n, int precision, double* z_out, double* cc_xyz_out,
                       uint8_t* cc_type_out, uint32_t* tri_out, uint8_t* feat_out,
                       ocl_stats* stats, void* stream);

/* Same as ocl_hip_drop_batch with every array in device memory (asynchronous). */
int ocl_hip_drop_device(ocl_mesh* mesh, const ocl_cutter_desc* cutter, const double* d_cl_xyz,
                        size_t n, int precision, double* d_z_out, double* d_cc_xyz_out,
                        uint8_t* d_cc_type_out, uint32_t* d_tri_out, uint8_t* d_feat_out,
                        ocl_stats* stats, void* stream);

/* Allocation-free variant of ocl_hip_drop_device (same arguments without
 * stats, same results bit for bit).  No counterpart in the reference: it is
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture, tight serving loops and harnesses that forbid hipMalloc.
 * The per-call workspace (keys, sort scratch, per-point selection records,
 * queues, counters) is one fixed array per GPU in the library, about 52 MB,
 * held until the process ends; the call makes no hipMalloc, hipFree or
 * other allocating call and clears what must start at zero with a kernel.
 * Limits: n <= 524288 CL points per call (OCL_E_ARG beyond; split larger
 * sets into several calls on one stream).  The selection records of the
 * mesh for this cutter and precision are mesh state: an earlier
 * ocl_hip_drop_device call on the mesh with this cutter and precision
 * builds them (OCL_E_ARG otherwise; EXACT64 and composite cutters need
 * none).  Every call on one GPU shares the array: calls that overlap on one
 * GPU (other streams, other threads) race; calls in order on one stream do
 * not. */
int ocl_hip_drop_device_noalloc(ocl_mesh* mesh, const ocl_cutter_desc* cutter, const double* d_cl_xyz, size_t n,
                                int precision, double* d_z_out, double* d_cc_xyz_out, uint8_t* d_cc_type_out,
                                uint32_t* d_tri_out, uint8_t* d_feat_out, void* stream);

/* Backward of the drop-cutter.  No counterpart in the reference: it serves the
 * differentiable path (the torch.autograd.Function of the Python binding and
 * slope-driven path refinement).  From the active contact a forward call on
 * the same mesh and cutter recorded (its cc_xyz, cc_type, tri and feat
 * outputs, device arrays), gives per point dz/dx, dz/dy, dz/dr and the
 * derivative of z with respect to the contact triangle's vertices.  Device
 * pointers, asynchronous on stream.
 *   d_grad_z   upstream dL/dz per point (NULL: all ones)
 *   d_dz       optional, n x 3: dz/dx, dz/dy, dz/dr per point (dz/dr is NaN
 *              for composite cutters, which have no single radius)
 *   d_grad_tri optional, ntri x 9 in input triangle order: ADDS
 *              sum_i grad_z[i] dz_i/d(p0, p1, p2 xyz); zero it first
 *   d_grad_r   optional, one double: ADDS sum_i grad_z[i] dz_i/dr
 * precision: OCL_PREC_FP32 / OCL_PREC_MIXED compute facet and vertex contacts
 * in FP32 (FP64 where the profile is ill-conditioned) and edge contacts in
 * FP64; OCL_PREC_FP64 / OCL_PREC_EXACT64 compute everything in FP64.
 * Every output is deterministic, bit for bit: d_grad_r is summed per block,
 * then by one fixed-order reduction; d_grad_tri sums each triangle's terms in
 * point order (a stable sort of the points by contact triangle, then
 * segmented sums), with no atomics.  The partial sums and the sort's buffers
 * live in the mesh handle: device calls on one mesh must run on one stream,
 * or be ordered with events, like the reference's non-re-entrant Operation. */
int ocl_hip_drop_backward(ocl_mesh* mesh, const ocl_cutter_desc* cutter, const double* d_cl_xyz, size_t n,
                          const double* d_cc_xyz, const uint8_t* d_cc_type, const uint32_t* d_tri,
                          const uint8_t* d_feat, const double* d_grad_z, int precision, double* d_dz,
                          double* d_grad_tri, double* d_grad_r, void* stream);

/* Forward plus slopes with host buffers (no counterpart in the reference; the
 * input of slope-driven path refinement, which replaces the flatness test of
 * AdaptivePathDropCutter::adaptive_sample, adaptivepathdropcutter.cpp:72-94):
 * z_out (n), dz_dxy_out (n x 2: dz/dx, dz/dy of each point's active contact,
 * 0 where nothing was touched) and optionally the CC point and type. */
int ocl_hip_drop_slopes(ocl_mesh* mesh, const ocl_cutter_desc* cutter, const double* cl_xyz, size_t n,
                        int precision, double* z_out, double* dz_dxy_out, double* cc_xyz_out,
                        uint8_t* cc_type_out, ocl_stats* stats, void* stream);

/* ---------------------------------------------------------------- batches
 * Many drop-cutter jobs of uneven mesh size, cutter and CL count in one
 * submission.  No counterpart in the reference, where each job is its own
 * BatchDropCutter::run (batchdropcutter.cpp:187-231); every job's results
 * are those ocl_hip_drop_device gives for it alone.  The batch runs
 *   - one key pass and one deterministic radix sort over all jobs' points,
 *   - the selection passes of all jobs in ONE persistent launch whose
 *     wavefronts pull 64-point chunks from a device work queue (largest jobs
 *     first), or one launch per job with per_job_launch (the baseline),
 *   - each job's resolve and exact passes,
 * with no host synchronisation and no memset, so with use_graph the whole
 * sequence is captured once in a hipGraph and every ocl_hip_batch_run
 * replays it (re-reading the jobs' CL arrays).  Composite cutters and
 * OCL_PREC_EXACT64 take the exact pass.  Each job has its own workspace and
 * keeps the selection records of its cutter, so a mesh may appear in several
 * jobs (with different cutters) and other calls on a mesh, drop calls with
 * other cutters or more points and other batches included, do not change
 * what a batch replays.  A mesh must outlive the batches that use it:
 * ocl_hip_mesh_destroy fails with OCL_E_ARG while one is alive. */
typedef struct {
    ocl_mesh* mesh;
    ocl_cutter_desc cutter;
    const double* d_cl_xyz; /* device, n x 3 */
    size_t n;
    double* d_z_out;        /* device outputs as ocl_hip_drop_device; all but z may be NULL */
    double* d_cc_xyz_out;
    uint8_t* d_cc_type_out;
    uint32_t* d_tri_out;
    uint8_t* d_feat_out;
} ocl_drop_job;

/* Tunables of a batch (zero = default). */
typedef struct {
    int32_t hilbert_bits;   /* CL binning: Hilbert curve over 2^bits x 2^bits cells per job domain (4..10, 9) */
    int32_t blocks_per_cu;  /* persistent selection grid: 256-thread blocks per CU (4) */
    int32_t use_graph;      /* capture in a hipGraph at create time and replay it in run */
    int32_t per_job_launch; /* one selection launch per job instead of the work queue */
} ocl_tune;

typedef struct ocl_batch ocl_batch;

/* Builds the batch: validates the jobs, builds and caches each mesh's
 * selection records for its cutter (synchronous), uploads the job tables
 * and, with use_graph, captures the graph. */
int ocl_hip_batch_create(const ocl_drop_job* jobs, size_t njobs, int precision, const ocl_tune* tune,
                         ocl_batch** out);
/* Runs (or replays) the batch on stream; with stats, waits for it and fills
 * points, pairs, escalated, wave_evals and kernel_ms (select_ms: the
 * selection launch, eager runs only). */
int ocl_hip_batch_run(ocl_batch* batch, ocl_stats* stats, void* stream);
/* Counters of one job of the last ocl_hip_batch_run (synchronous): points,
 * pairs evaluated by the selection, escalated, wave_evals.  No counterpart
 * in the reference (a diagnostic, as BatchDropCutter::getCalls). */
int ocl_hip_batch_job_stats(ocl_batch* batch, size_t job, ocl_stats* stats);
int ocl_hip_batch_destroy(ocl_batch* batch);
/* Autotuner: times every combination of hilbert_bits {7, 8, 9} and
 * blocks_per_cu {2, 4, 8} on the jobs themselves (graph replays, best of
 * `trials` after a warm-up run) and returns the fastest in *best. */
int ocl_hip_batch_autotune(const ocl_drop_job* jobs, size_t njobs, int precision, int trials, ocl_tune* best,
                           double* best_ms);

/* Batch push-cutter: replaces BatchPushCutter::run / pushCutter3
 * (batchpushcutter.cpp:136-188) for one fiber direction (dir 0 = X fibers,
 * setXDirection; 1 = Y fibers, setYDirection; batchpushcutter.cpp:55-66).
 * fib: nf x (p1.xyz, p2.xyz), horizontal, along the chosen axis.  Computes
 * every fiber's interval list (the union of its per-triangle intervals,
 * Fiber::addInterval, fiber.cpp:58-93) on the device and returns
 * offsets_out[0..nf] (fiber i owns rows offsets_out[i] .. offsets_out[i+1]-1)
 * and the row total; ocl_hip_push_fetch then copies the rows out.  The
 * search box is across +- R, z in [f.z, f.z + L]; the computation is FP64 in
 * every precision mode.
 * Pairs: every triangle the reference's KD search returns for the fiber
 * (kdtree.hpp:273-304, a superset of the box overlaps, all of them pushed as
 * in batchpushcutter.cpp:170-181); with OCL_PUSH_EXACT_FILTER only the
 * triangles whose boxes overlap the search box.  OCL_PREC_EXACT64 pushes
 * every such pair; the other modes leave out a pair whose interval bound
 * (the triangle's extent along the fiber widened by the cutter's profile
 * width) lies strictly inside an interval the fiber already holds, or whose
 * triangle lies below the fiber: Fiber::addInterval would ignore it
 * (fiber.cpp:58-66), so the results are the same (stats->skipped counts
 * them; the reference notes the idea at batchpushcutter.cpp:177). */
int ocl_hip_push_batch(ocl_mesh* mesh, const ocl_cutter_desc* cutter, int dir, const double* fib_p1p2, size_t nf,
                       int precision, uint64_t* offsets_out, uint64_t* total_out, ocl_stats* stats, void* stream);

/* Rows of the last ocl_hip_push_batch on this mesh, in fiber order, each
 * fiber's intervals sorted by lower end: lohi = (lower, upper) fiber
 * parameters t of p1 + t (p2 - p1); cc = lower CC xyz, lower CCType,
 * upper CC xyz, upper CCType (8 doubles per row). */
int ocl_hip_push_fetch(ocl_mesh* mesh, double* lohi, double* cc, void* stream);

/* X and Y fibers in one call: replaces the two BatchPushCutter runs of
 * Waterline::run (waterline.cpp:78-87, which notes they "could/should run
 * in parallel").  Both directions' fibers form one device queue in one
 * persistent launch, so the second direction fills the GPU while the first
 * one's last fibers finish.  xfib: nx X fibers, yfib: ny Y fibers (as
 * ocl_hip_push_batch with dir 0 and 1); the results per direction are those
 * of ocl_hip_push_batch for that direction alone: x_offsets_out[0..nx],
 * y_offsets_out[0..ny], totals_out[0..1] = row totals (X, Y); rows via
 * ocl_hip_push_fetch_dir.  stats: both directions together. */
int ocl_hip_push_xy(ocl_mesh* mesh, const ocl_cutter_desc* cutter, const double* xfib, size_t nx,
                    const double* yfib, size_t ny, int precision, uint64_t* x_offsets_out,
                    uint64_t* y_offsets_out, uint64_t* totals_out, ocl_stats* stats, void* stream);

/* Rows of the last push in direction dir (0 = X, 1 = Y; by ocl_hip_push_batch
 * or ocl_hip_push_xy), as ocl_hip_push_fetch.  No reference counterpart:
 * the fibers of BatchPushCutter::getFibers (batchpushcutter.hpp:65) for one
 * direction. */
int ocl_hip_push_fetch_dir(ocl_mesh* mesh, int dir, double* lohi, double* cc, void* stream);

/* Thread-local message of the last failure. */
const char* ocl_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
