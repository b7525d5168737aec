/*
 * ocl_oracle.h -- FP64 CPU restatement of OpenCAMLib's batch drop-cutter and
 * batch push-cutter, used ONLY as test infrastructure: the parity checker of
 * the HIP path and the timed CPU baseline.  Nothing in the product library
 * links or calls this code.
 *
 * Every routine follows the reference operation order so that results match
 * the reference's FP64 arithmetic (x86-64 SSE2, no FMA contraction).  The
 * per-function comments cite the reference file:line restated.
 *
 * Pinning (what checks this restatement independently of its own formulas):
 *   - closed forms: plane facet contacts, vertex contacts, Cyl/Ball edge
 *     contacts against lines (tests/test_oracle_closed_form.py);
 *   - brute force: drop z of every cutter, composites included, against a
 *     triangle equals the supremum of p.z - height(|p - cl|_xy) over densely
 *     sampled triangle points, refined by golden-section search
 *     (tests/test_oracle_bruteforce.py, tests/test_oracle_composite.py);
 *     where a reference rule discards a contact the supremum keeps, the
 *     oracle's diagnostic entry points flag the point (OC_Q_* below) and
 *     the check is that the result is only lower there;
 *   - push vs drop duality, composites included: push interval endpoints
 *     equal the fiber positions where the drop height crosses the fiber z;
 *   - ConeCutter with l > 0: the rim facet contact is the l = 0 one moved
 *     down by l (conecutter.cpp:92); the hyperbola edge contact
 *     (conecutter.cpp:122) is the tangency point of the cone of height
 *     `length` over the same radius (tests/test_oracle_composite.py);
 *   - KD tree leaf order with ties in spread and the Fiber::addInterval CC
 *     attribution against hand-derived cases (tests/test_oracle_hand.py);
 *     KD search vs exhaustive overlap scan (candidate sets equal);
 *   - interval union vs an independent sort-and-merge union.
 */
#ifndef OCL_ORACLE_H
#define OCL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* CCType numbering of src/geo/ccpoint.hpp:33-39 */
enum oc_cctype {
    OC_NONE = 0, OC_VERTEX, OC_VERTEX_CYL, OC_EDGE, OC_EDGE_HORIZ, OC_EDGE_SHAFT,
    OC_EDGE_HORIZ_CYL, OC_EDGE_HORIZ_TOR, OC_EDGE_BALL, OC_EDGE_POS, OC_EDGE_NEG,
    OC_EDGE_CYL, OC_EDGE_CONE, OC_EDGE_CONE_BASE, OC_FACET, OC_FACET_TIP,
    OC_FACET_CYL, OC_ERROR
};

enum oc_cutter_kind { OC_CYL = 0, OC_BALL = 1, OC_BULL = 2, OC_CONE = 3, OC_COMPOSITE = 4 };

#define OC_MAX_SUB 4

typedef struct oc_cutter {
    int kind;
    double diameter, radius, length;
    double xy_normal_length, normal_length, center_height;
    double radius1, radius2; /* bull */
    double angle;            /* cone half-angle */
    /* composite */
    int nsub;
    struct oc_cutter* sub[OC_MAX_SUB];
    double radiusvec[OC_MAX_SUB], heightvec[OC_MAX_SUB], zoffset[OC_MAX_SUB];
} oc_cutter;

typedef struct oc_mesh oc_mesh;

/* cutters (constructors of cylcutter.cpp:39-46, ballcutter.cpp:38-45,
   bullcutter.cpp:38-47, conecutter.cpp:39-49) */
void oc_cutter_cyl(oc_cutter* c, double d, double l);
void oc_cutter_ball(oc_cutter* c, double d, double l);
void oc_cutter_bull(oc_cutter* c, double d, double r, double l);
void oc_cutter_cone(oc_cutter* c, double d, double angle, double l);
void oc_cutter_composite(oc_cutter* c);
/* CompositeCutter::addCutter (compositecutter.cpp:43-53); sub is borrowed */
void oc_composite_add(oc_cutter* c, oc_cutter* sub, double r, double h, double zoff);

/* mesh from 9 doubles per triangle; triangle normal/bbox as triangle.cpp:39-75 */
oc_mesh* oc_mesh_create(const double* xyz9, long ntri);
void oc_mesh_free(oc_mesh* m);
long oc_mesh_size(const oc_mesh* m);

/* KD-tree leaf order (rank) of kdtree.hpp:117-268.  dimset 0 = XY (drop),
   1 = YZ (X-fibers), 2 = XZ (Y-fibers).  rank_out[tri] = position. */
int oc_kd_rank(const oc_mesh* m, int dimset, unsigned bucket, int* rank_out);

/* KD search (kdtree.hpp:123-135, 273-304) for a drop query; returns the number
   of triangles written (in search order) or -needed if cap is too small. */
long oc_kd_search_drop(oc_mesh* m, const oc_cutter* c, double x, double y, double z,
                       unsigned bucket, int* out, long cap);

/* KD search of one push fiber (batchpushcutter.cpp:160-170), dir 0 = X
   fiber (YZ tree), 1 = Y fiber (XZ tree); same return convention. */
long oc_kd_search_push(oc_mesh* m, const oc_cutter* c, int dir, const double* fib6, unsigned bucket, int* out,
                       long cap);

/* Batch drop-cutter (batchdropcutter.cpp:187-231).
   search_mode 0: reference KD search then overlaps/below filter;
   search_mode 1: exhaustive overlap scan sorted by KD rank (same result);
   search_mode 2: every triangle in input order, like dropCutter1 (no filter).
   Outputs: z, cc xyz, cc type, index of the triangle and feature code of the
   last lift (-1 if none).  ncalls = dropCutter() calls made. */
int oc_batch_drop(oc_mesh* m, const oc_cutter* c, const double* cl_xyz, long n,
                  int search_mode, int nthreads, unsigned bucket,
                  double* z_out, double* cc_out, int* cctype_out,
                  int* tri_out, int* feat_out, long* ncalls_out);

/* single-triangle MillingCutter::dropCutter (millingcutter.cpp:386-409) on a
   CL point starting at z_io; returns 1 if lifted. */
int oc_drop_triangle(const oc_cutter* c, const double* tri9, double x, double y,
                     double* z_io, double* cc_out, int* cctype_out, int* feat_out);

/* One edge of MillingCutter::edgeDrop for a simple cutter (the radius test
   and singleEdgeDrop, millingcutter.cpp:84-91, 121-136); p1p2 = 6 doubles;
   z_io = running height in and out; returns 1 if lifted (cc_out = CC). */
int oc_edge_drop(const oc_cutter* c, const double* p1p2, double x, double y, double* z_io, double* cc_out);

/* Parity diagnostics (no reference counterpart).  The _diag variants give
   the same results and also OR, per CL point, the bits below wherever one of
   the reference's rules discarded a contact that the geometric supremum of
   p.z - h(|p - cl|_xy) over the triangles would keep.  Tests compare with
   that supremum at every point without a bit. */
#define OC_Q_COMP_EDGE 1   /* composite: the sub-cutter's highest edge contact
                              failed the radius window and a single edge had
                              an in-window contact above the running height
                              (compositecutter.cpp:175-194) */
#define OC_Q_CONE_TIP 2    /* composite: a cone sub-cutter's tip lift skipped
                              its rim contact (conecutter.cpp:100-102), the
                              tip failed the window and the rim would have
                              lifted in the window */
#define OC_Q_FACET_FIRST 4 /* the facet lifted, so the vertex and edge tests
                              were skipped (millingcutter.cpp:398-405), and
                              one of them would have lifted higher */
#define OC_Q_COMP_BELOW 8  /* composite push: a contact below the first
                              height window was kept for the last sub-cutter,
                              the fall-back of height_to_index
                              (compositecutter.cpp:71-83, 196-270) */
int oc_push_triangle_diag(const oc_cutter* c, const double* fib6, const double* tri9, double* out10, int* quirk_out);
int oc_drop_triangle_diag(const oc_cutter* c, const double* tri9, double x, double y, double* z_io, double* cc_out,
                          int* cctype_out, int* feat_out, int* quirk_out);
int oc_batch_drop_diag(oc_mesh* m, const oc_cutter* c, const double* cl_xyz, long n, int search_mode, int nthreads,
                       unsigned bucket, double* z_out, double* cc_out, int* cctype_out, int* tri_out, int* feat_out,
                       long* ncalls_out, int* quirk_out);

/* Analytic gradient of the drop height at a CL point from its active contact
   (feature code and CC point of the forward pass).  Not a reference
   function: the FP64 check of the device backward kernel.  out (12) = dz/dx,
   dz/dy, dz/dr (NaN for composites), dz/d(p0, p1, p2 xyz).  Returns 1 when
   there is no contact. */
int oc_drop_grad(const oc_cutter* c, const double* tri9, int feat, int cctype, double x, double y,
                 const double* cc3, double* out);

/* oc_drop_grad over a batch (OpenMP); tri = input triangle index per point
   (-1: no contact, zero gradient), out12 = 12 doubles per point. */
void oc_batch_drop_grad(const oc_mesh* m, const oc_cutter* c, const double* cl_xyz, long n, const int* tri,
                        const int* feat, const int* cctype, const double* cc, int nthreads, double* out12);

/* Batch push-cutter (batchpushcutter.cpp:136-188).  dir 0 = X fibers,
   1 = Y fibers.  fib = 6 doubles per fiber (p1, p2).
   search_mode 0: reference KD superset (no filter);
   search_mode 2: KD search then exact bbox overlap in the two search dims.
   Output: offsets[nf+1]; for interval k: lohi[2k], lohi[2k+1];
   cc[8k..8k+7] = lower cc xyz, type, upper cc xyz, type.
   Returns total intervals, or -(needed) when cap is too small. */
long oc_batch_push(oc_mesh* m, const oc_cutter* c, int dir, const double* fib, long nf,
                   int search_mode, int nthreads, unsigned bucket,
                   long cap, long* offsets, double* lohi, double* cc, long* ncalls_out);

/* single (fiber, triangle) MillingCutter::pushCutter (millingcutter.cpp:377-382).
   Returns 1 if the interval is non-empty. out = lower, upper, lower cc xyz,
   type, upper cc xyz, type. */
int oc_push_triangle(const oc_cutter* c, const double* fib6, const double* tri9, double* out10);

/* Fiber::addInterval (fiber.cpp:58-93) applied to a list of per-pair
   intervals in order (10 doubles each, as oc_push_triangle).  Returns count. */
long oc_fiber_union(const double* ints10, long n, long cap, double* out10);

#ifdef __cplusplus
}
#endif
#endif
