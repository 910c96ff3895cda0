/*
 * atac: MI355X (gfx950) point-vs-SDF contact for RedMax-style simulation.
 *
 * Plain C ABI of libatac_contact.so. Host buffers are plain pointers and sizes;
 * "device" buffers are HIP device pointers allocated by the caller; streams are
 * hipStream_t passed as void* (NULL = the default stream). Every entry point
 * returns ATAC_OK (0) or a negative status; atac_last_error() returns the message
 * of the last failure on the calling thread. Handles are not thread-safe.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   simulation/redmax/Body/BodySDFObj.{h,cpp}          the SDF body (grid + queries)
 *   simulation/redmax/CollisionDetection/CollisionDetection.cpp:124-147
 *   simulation/redmax/Force/ForceGeneralSDFContact.{h,cpp}
 *   simulation/redmax/Mesh.cpp:143-198                 SDFMesh distance queries
 */
#ifndef ATAC_H
#define ATAC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATAC_OK 0
#define ATAC_ERR_INVALID (-1)   /* bad argument or size */
#define ATAC_ERR_HIP (-2)       /* HIP runtime error */
#define ATAC_ERR_NOMEM (-3)     /* allocation failed */
#define ATAC_ERR_STATE (-4)     /* call not valid in the current state */

typedef struct atac_grid_s* atac_grid;     /* FP32 SDF grid resident in HBM */
typedef struct atac_points_s* atac_points; /* Morton-sorted SoA FP32 contact points */
typedef struct atac_batch_s* atac_batch;   /* instance table, tile plan, workspace */
typedef struct atac_mesh_s* atac_mesh;     /* triangle mesh + BVH + pseudonormals (BVH contact variant) */

/* One force object: ForceGeneralSDFContact(contact_body, SDF_body, kn, kt, mu, damping)
 * with its _scale, the contact body's _contact_scale, the SDF body's _col_th and the
 * two bodies' _E_0i / _phi (ForceGeneralSDFContact.h:14-31, Body.h:48-61,
 * BodySDFObj.h:13-15). E is 4x4 row-major [R p; 0 1]; phi = [w; v] in the body frame. */
typedef struct {
    int32_t grid;     /* index into the grid array given to atac_batch_create */
    int32_t points;   /* index into the point-set array */
    int32_t segment;  /* segment of that point set */
    int32_t reserved;
    double E1[16];    /* contact body pose */
    double phi1[6];
    double E2[16];    /* SDF body pose */
    double phi2[6];
    double col_th;    /* BodySDFObj::_col_th */
    double kn, kt, mu, damping;
    double scale;         /* ForceGeneralSDFContact::_scale (set_contact_scale) */
    double contact_scale; /* contact body's _contact_scale */
} atac_instance;

typedef struct {
    int32_t tile_points;   /* points per tile (0: 4096); a multiple of 256, at most 8192 */
    int32_t detect_block;  /* threads per detection workgroup: 64, 128, 256, 512, 1024 (0: 256) */
    int32_t blocks_per_cu; /* detection: 0 = one workgroup per tile (default); n > 0 = n persistent
                              workgroups per CU (at most the occupancy limit) walking the tiles */
    int32_t order;         /* tile queues: 0 default (1); 1 one queue in instance order;
                              2 one queue grouped by grid; 3 one queue per XCD, grouped by grid;
                              4 one queue per XCD by the tile's place in its instance's Morton-sorted
                              points (the same part of a grid on one XCD, every grid on every XCD) */
    int32_t points_per_lane; /* points a detection lane takes per step (64 apart): 1, 2, 4 (0: 4) */
    int32_t max_active;    /* bound on the active contacts of one evaluation (the record pool, 48 B each;
                              96 B with ATAC_EVAL_PRECISE): 0 = one per point up to 4M points, a quarter of
                              the points beyond; > 0 that many; < 0 one per point (needed by
                              ATAC_EVAL_ALL_ACTIVE). An evaluation with more contacts gives NaN outputs and
                              n_active -1, and atac_batch_workspace reports the overflow */
    int32_t group_cull;    /* 0 default (1); 1: before detection, drop the groups of 256 consecutive sorted
                              points whose bounding box maps only to cells where no point can be active
                              (the decisions are unchanged); -1: detect every point */
} atac_tuning;

/* flags of atac_batch_eval / atac_eval_plain */
#define ATAC_EVAL_FORCE 0u       /* fm only: ForceGeneralSDFContact::computeForce */
#define ATAC_EVAL_DERIV 1u       /* + Km, Dm: computeForceWithDerivative */
#define ATAC_EVAL_P6 2u          /* + contact-scale column of dfm_dp */
#define ATAC_EVAL_CHECKPOINT 4u  /* accepted for compatibility: every evaluation keeps its active-contact
                                    records (d, grad phi, Hessian) for atac_batch_vjp */
#define ATAC_EVAL_PRECISE 8u     /* FP64 contact arithmetic (detection stays FP32) */
#define ATAC_EVAL_FUSED_ADJOINT 32u /* atac_batch_graph_capture with g_fm: the fused pass of
                                       atac_batch_eval_vjp instead of evaluation + adjoint */
#define ATAC_EVAL_ALL_ACTIVE 16u /* every point is a contact, as for the reference's contact-list
                                    overloads (ForceGeneralSDFContact.cpp:147, :221, :449) */

/* Outputs of one evaluation, device pointers, any may be NULL. Per instance:
 * fm 12 = [tau1 f1 tau2 f2] (each body's frame, angular first), Km/Dm 144 row-major
 * (d fm / d [q1 q2], d fm / d [phi1 phi2], pose perturbation E <- E exp([w; v])),
 * dfm_dp6 12 (d fm / d contact_scale as the reference assembles it), n_active 1. */
typedef struct {
    double* fm;
    double* Km;
    double* Dm;
    double* dfm_dp6;
    int32_t* n_active;
} atac_outputs;

/* Gradients from atac_batch_vjp, device pointers, any may be NULL. Per instance:
 * g_q1/g_q2 6 (right-multiplied pose twists), g_phi1/g_phi2 6, g_kn 1, g_kt 1
 * (with respect to the force's kn and kt parameters). */
typedef struct {
    double* g_q1;
    double* g_q2;
    double* g_phi1;
    double* g_phi2;
    double* g_kn;
    double* g_kt;
} atac_gradients;

const char* atac_last_error(void);
int atac_version(void);

/* Selects the HIP device used by subsequent calls of this thread and reports its
 * compute-unit count (may be NULL). */
int atac_device_init(int32_t device, int32_t* n_cu);

/* BodySDFObj's grid (_SDF, _min_box, _max_box, _dx/_dy/_dz; BodySDFObj.h:13-16,
 * BodySDFObj.cpp:77-83): ni*nj*nk FP32 values in a[i + ni*(j + nj*k)] order. */
int atac_grid_create(const float* values, int32_t ni, int32_t nj, int32_t nk, const double min_box[3],
                     const double max_box[3], const double dxyz[3], atac_grid* out);
int atac_grid_destroy(atac_grid g);

/* Body::_contact_points (Body.h:37,129; BodyMeshObj.cpp:184-238): n points, FP64 xyz
 * interleaved, in the contact body's frame. seg_offsets (n_segments + 1 entries,
 * NULL = one segment) splits them into independent sets. Each segment is Morton-
 * sorted by the cell of edge sort_cell it falls in (<= 0: automatic) and stored
 * as FP32 SoA. */
int atac_points_create(const double* xyz, int64_t n, const int64_t* seg_offsets, int32_t n_segments,
                       double sort_cell, atac_points* out);
int atac_points_destroy(atac_points p);
/* For each stored (sorted) point of a segment, its original index within the segment. */
int atac_points_permutation(atac_points p, int32_t segment, int64_t* sorted_to_original, int64_t capacity);

/* A set of force objects evaluated together in one persistent launch. */
int atac_batch_create(const atac_grid* grids, int32_t n_grids, const atac_points* points, int32_t n_point_sets,
                      const atac_instance* inst, int32_t n_inst, const atac_tuning* tuning, atac_batch* out);
int atac_batch_destroy(atac_batch b);
/* New poses / parameters for the same (grid, points, segment) structure. The
 * adjoint records of the last evaluation are dropped (atac_batch_vjp then returns
 * ATAC_ERR_STATE until the next evaluation). stream NULL: a synchronous upload
 * ordered after the work already queued on the batch's own stream. */
int atac_batch_set_instances(atac_batch b, const atac_instance* inst, int32_t n_inst, void* stream);
/* Re-plans tiles and launch shape (used by the autotuner). */
int atac_batch_retune(atac_batch b, const atac_tuning* tuning);
int atac_batch_info(atac_batch b, int64_t* n_points, int32_t* n_tiles, int32_t* persistent_blocks,
                    atac_tuning* tuning);

/* ForceGeneralSDFContact::computeForce / computeForceWithDerivative for every instance
 * (ForceGeneralSDFContact.cpp:46-144 with collision_detection_general_SDF,
 * CollisionDetection.cpp:124-147): detection over all points, then the contact
 * wrench (and Km, Dm, p6 column) over the active ones. Stream-ordered. */
int atac_batch_eval(atac_batch b, uint32_t flags, const atac_outputs* out, void* stream);

/* Adjoint of the last evaluation of the batch (at its precision), from the
 * active-contact records it kept: g_fm (device, 12 per instance) -> gradients.
 * Stream-ordered after that evaluation. */
int atac_batch_vjp(atac_batch b, const double* g_fm, const atac_gradients* out, void* stream);

/* atac_batch_eval then atac_batch_vjp for a cotangent g_fm (n_inst x 12, device) known when the
 * contacts are evaluated (a backward pass that re-evaluates a step's contact forces,
 * ForceGeneralSDFContact::computeForce, ForceGeneralSDFContact.cpp:46-144, with the adjoint
 * of its wrench): one pass over the active contacts computes both, without writing the
 * records. The same outputs as the two calls (wrench only: flags ATAC_EVAL_PRECISE and
 * ATAC_EVAL_ALL_ACTIVE; no Km/Dm or p6). No records are kept: atac_batch_vjp then returns
 * ATAC_ERR_STATE until the next atac_batch_eval. Stream-ordered. No counterpart call in the
 * reference, which forms Km and Dm instead. */
int atac_batch_eval_vjp(atac_batch b, uint32_t flags, const atac_outputs* out, const double* g_fm,
                        const atac_gradients* grads, void* stream);

/* Captures [reset, detect, slow, contact, reduce] (and the adjoint when g_fm != NULL;
 * with ATAC_EVAL_FUSED_ADJOINT, the fused pass of atac_batch_eval_vjp instead)
 * into a HIP graph bound to these outputs; atac_batch_graph_launch replays it.
 * Capturing runs nothing: atac_batch_vjp returns ATAC_ERR_STATE until the graph
 * (or an eager evaluation) has been launched. */
int atac_batch_graph_capture(atac_batch b, uint32_t flags, const atac_outputs* out, const double* g_fm,
                             const atac_gradients* grads);
int atac_batch_graph_launch(atac_batch b, void* stream);

/* The record pool of the batch and the last evaluation's use of it (synchronous):
 * capacity (records), used (contacts the last evaluation found; may exceed the
 * capacity) and overflowed (1: its outputs are NaN, its adjoint refused; plan the
 * batch again with a larger max_active). Any output may be NULL. */
int atac_batch_workspace(atac_batch b, int64_t* capacity, int64_t* used, int32_t* overflowed);

/* The points the last evaluation's detection read (live_points: those of the groups the
 * culling pass kept; all of them without culling) and the batch's points (synchronous).
 * No counterpart in the reference: a diagnostic of this library's group culling. */
int atac_batch_live_points(atac_batch b, int64_t* live_points, int64_t* points);

/* Per-stage kernel timing with HIP events recorded on the launch stream around
 * each kernel of atac_batch_eval / atac_batch_vjp (not inside captured graphs).
 * Reads and clears what was recorded since the last call (ms_sum, launches: 6
 * entries each for detect, slow, contact, reduce, vjp, vjp-reduce; may be NULL), then
 * enables (enable != 0) or disables recording. Synchronous when reading. */
int atac_batch_timing(atac_batch b, int32_t enable, double* ms_sum, int32_t* launches);

/* Active contacts of one instance from the last evaluation (host out): original
 * point indices (within the instance's segment), ascending. Synchronous. */
int atac_batch_active_ids(atac_batch b, int32_t inst, int64_t* ids, int64_t capacity, int64_t* count);

/* The plain path with the outputs of an evaluation followed by its adjoint: the
 * wrench and n_active (out; Km, Dm, dfm_dp6 are not used) and, for the cotangent
 * g_fm (device, 12 per instance), the gradients of atac_batch_vjp, with one launch
 * per instance, one lane per point and FP64 atomics. The per-instance baseline of
 * the batched evaluation + adjoint. */
int atac_eval_plain_vjp(atac_batch b, uint32_t flags, const double* g_fm, const atac_outputs* out,
                        const atac_gradients* grads, void* stream);

/* Distinct grid values read by the detection stencils of all instances (summed
 * over grids). Synchronous diagnostic used for the algorithmic byte count. */
int atac_batch_touched_values(atac_batch b, int64_t* count);

/* The plain per-instance path: one launch per instance, one lane per point, the
 * contact math inline and FP64 atomics; same outputs as atac_batch_eval. */
int atac_eval_plain(atac_batch b, uint32_t flags, const atac_outputs* out, void* stream);

/* BodySDFObj::distance (BodySDFObj.cpp:207-213) for n world points (device, n x 3):
 * query_SDF(R2^T (xw - p2)) + col_th, FLT_MAX + col_th outside the box. With E2 = I
 * and col_th = 0 this is SDFMesh::distance (Mesh.cpp:143-198). */
int atac_sdf_distance(atac_grid g, const double E2[16], double col_th, const double* xw, int64_t n, double* d,
                      void* stream);

/* BodySDFObj::collision (BodySDFObj.cpp:239-377) for n world points and velocities
 * (device, n x 3 each): values (n x 11: d, n[3], ddot, tdot[3], xi2[3]) and, when
 * derivs != NULL, the 14 derivative blocks (n x 135: dd_dxw 3, dd_dq2 6, dn_dxw 9,
 * dn_dq2 18, dddot_dxw 3, dddot_dxwdot 3, dddot_dq2 6, dddot_dphi2 6, dtdot_dxw 9,
 * dtdot_dxwdot 9, dtdot_dq2 18, dtdot_dphi2 18, dxi2_dxw 9, dxi2_dq2 18, row-major). */
int atac_sdf_collision(atac_grid g, const double E2[16], const double phi2[6], double col_th, uint32_t flags,
                       const double* xw, const double* xw_dot, int64_t n, double* values, double* derivs,
                       void* stream);

/* Synchronous host-buffer variants for callers that hold no device memory (the C++
 * Force/Body classes, ctypes without torch): they run on the batch's own stream (or
 * the default stream) with library-owned device staging and return when the host
 * buffers are filled. Same meaning and layout as the device-pointer entry points. */
int atac_batch_eval_host(atac_batch b, uint32_t flags, double* fm, double* Km, double* Dm, double* dfm_dp6,
                         int32_t* n_active);
/* atac_batch_set_instances + atac_batch_eval_host in one call for host loops that
 * evaluate the same force objects at changing poses (Simulation::newton's force
 * evaluations, Simulation.cpp:1081-1117; the reference calls computeForce* per force
 * object each time): the instance upload and the evaluation, whose reduction writes the
 * results straight into pinned host memory, are replayed as one HIP graph with one wait
 * (captured on first use per flag combination and again after the batch's buffers
 * change). Same instance rules, meaning and layout as those two; not for use while the
 * batch's own stream is being captured. */
int atac_batch_eval_instances_host(atac_batch b, const atac_instance* inst, int32_t n_inst, uint32_t flags, double* fm,
                                   double* Km, double* Dm, double* dfm_dp6, int32_t* n_active);
int atac_batch_vjp_host(atac_batch b, const double* g_fm, double* g_q1, double* g_q2, double* g_phi1, double* g_phi2,
                        double* g_kn, double* g_kt);
int atac_sdf_distance_host(atac_grid g, const double E2[16], double col_th, const double* xw, int64_t n, double* d);

/* Signed distance grid of a closed triangle mesh: the level-set construction the
 * reference's SDF bodies are built with (sdfgen::make_level_set3,
 * simulation/externals/sdfgen/makelevelset3.cpp:163-230, called from
 * BodySDFObj::precompute_SDF, Body/BodySDFObj.cpp:38-77, and SDFMesh::precompute_SDF,
 * Mesh.cpp:54-95): exact distances within exact_band cells of each triangle, fast
 * sweeping elsewhere, inside (negative) where the +x ray crossings so far are odd;
 * the same values node for node. x: 3*nv FP32 vertex coordinates; tri: 3*nf vertex
 * indices (< nv); origin: node (0,0,0); phi (host out): ni*nj*nk FP32, x fastest.
 * Runs on the current device; synchronous. */
int atac_sdf_from_mesh(const float* x, int64_t nv, const uint32_t* tri, int64_t nf, const float origin[3], float dx,
                       float dy, float dz, int32_t ni, int32_t nj, int32_t nk, int32_t exact_band, float* phi);

/* The grid precompute_SDF places around a mesh (BodySDFObj.cpp:38-77 with
 * min_size = res; Mesh.cpp:54-95 with min_size = 20): V is 3*nv FP64 vertex
 * coordinates; outputs the FP32 min_box (the grid origin) and max_box (the
 * bounding box padded by one cell), the spacing d (FP64) and the node counts n. */
int atac_sdf_mesh_grid(const double* V, int64_t nv, double dx, int32_t min_size, float min_box[3], float max_box[3],
                       double d[3], int32_t n[3]);
int atac_sdf_collision_host(atac_grid g, const double E2[16], const double phi2[6], double col_th, uint32_t flags,
                            const double* xw, const double* xw_dot, int64_t n, double* values, double* derivs);

/* ---- BVH contact variant ----
 * The triangle mesh of a BVH body in its body frame, prepared for exact signed
 * distance queries: a BVH, face normals, uniform edge normals and angle-weighted
 * vertex normals, as BodyBVHObj sets up libigl's AABB tree
 * (Body/BVH/BVHEngine.h:33-55, Body/BVH/AABBTree.h:56-60; BodyMeshObj.cpp:179-182 for
 * the bounding box). V: 3*nv FP64 coordinates; F: 3*nf vertex indices (< nv). */
int atac_mesh_create(const double* V, int64_t nv, const int64_t* F, int64_t nf, atac_mesh* out);
int atac_mesh_destroy(atac_mesh m);

/* BodyBVHObj::distance_normal_parallel (Body/BodyBVHObj.cpp:322-356): for each world
 * point xw[i] (3*n FP64) with the body at E2, d = signed distance + col_th inside the
 * vertices' bounding box (libigl signed_distance_pseudonormal: closest point over
 * all triangles, sign from the pseudonormal of the closest feature) and 10 outside;
 * normal (3*n) = that pseudonormal (0 outside); face (n) = the closest triangle
 * (-1 outside). normal and face may be NULL. Device pointers, asynchronous on stream. */
int atac_mesh_distance_normal(atac_mesh m, const double E2[16], double col_th, const double* xw, int64_t n, double* d,
                              double* normal, int32_t* face, void* stream);
/* The same with host pointers (synchronous). */
int atac_mesh_distance_normal_host(atac_mesh m, const double E2[16], double col_th, const double* xw, int64_t n,
                                   double* d, double* normal, int32_t* face);

/* ForceGeneralBVHContact::computeForce / computeForceWithDerivative
 * (Force/ForceGeneralBVHContact.cpp:46-219 and its derivative overloads): the
 * contacts of the contact body's points xi (3*n FP64, body frame) against the mesh
 * (collision_detection_general_BVH, CollisionDetection/CollisionDetection.cpp:150-185:
 * inside the bounding box and d = signed distance + col_th < 0), the wrench fm[12]
 * (contact body then BVH body, each in its own frame), with ATAC_EVAL_P6 the
 * contact-scale column dfm_dp6[12], with ATAC_EVAL_DERIV Km[144] and Dm[144]
 * (row-major, w.r.t. [q1; q2] and [phi1; phi2]); n_active (may be NULL) = contacts.
 * inst->grid is ignored. FP64 arithmetic, a fixed summation order. Host pointers. */
int atac_bvh_force_host(atac_mesh m, const atac_instance* inst, const double* xi, int64_t n, uint32_t flags,
                        double* fm, double* Km, double* Dm, double* dfm_dp6, int64_t* n_active);

/* No counterpart in the reference. atac_bvh_force_host for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): it calls no allocating routine. The partial
 * rows of its force workgroups live in a fixed scratch array of each device
 * (1.3 MB, held until the process ends), so n is at most
 * ATAC_BVH_NOALLOC_MAX_POINTS; larger n is refused with ATAC_ERR_INVALID.
 * Every call on a device shares that device's array: calls that overlap on one
 * GPU (other streams, other threads) race and must be serialised by the caller.
 * xi: device, 3*n FP64 (contact body frame); out: device, 312 doubles
 * (fm 12 | dfm_dp6 12 | Km 144 | Dm 144, row-major; the blocks flags do not ask
 * for hold zeros); n_active: device, 1 (may be NULL). inst is a host pointer.
 * Asynchronous on stream, same results bit for bit as atac_bvh_force_host. */
#define ATAC_BVH_NOALLOC_MAX_POINTS 1048576
int atac_bvh_force_noalloc(atac_mesh m, const atac_instance* inst, const double* xi, int64_t n, uint32_t flags,
                           double* out, int64_t* n_active, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ATAC_H */
