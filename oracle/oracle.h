/*
 * oracle.h -- CPU restatement of the reference's TSDF fusion path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library never links it.
 *
 * Everything here follows cu-Reconstruction's source (paths relative to its
 * root) under the build contract listed per function: the reference's own
 * operation order in IEEE float32, no FMA contraction (built with
 * -ffp-contract=off), correctly rounded division.  Three reference quirks are
 * replaced, identically in the oracle and the GPU path:
 *   - sdf uses the voxel's camera-space z, not its world z
 *     (tsdf_volume.cu:172-173 compute voxel_cam and then use voxel_ct.z);
 *     identical to the reference at the identity pose;
 *   - a voxel must have camera z > 0 (camera_util.cu:9-10,46-47 map z == 0
 *     to pixel (0,0) and project points behind the camera mirrored);
 *   - the weight is capped, W' = min(W+1, max_weight), with
 *     D' = (D*W + tsdf)/(W+1) (tsdf_volume.cu:179-185 ignore max_weight; the
 *     cap follows octree.cu:125-127);
 *   - marching cubes uses the layer stride gx*gy (marching_cubes.cu:211
 *     uses gx+gy, correct only for 2x2xN grids);
 *   - the trilinear sampler weights neighbour (1,1,1) by xyz and (0,1,1) by
 *     (1-x)yz (tsdf_volume.cu:100-101 swap the two);
 *   - the z-buffer render takes fx, fy, cx, cy explicitly
 *     (cuda_point_cloud.cu:248-251 read cx, cy from K[2], K[5], which are 0
 *     in the column-major K the point extraction reads).
 */
#ifndef RECON_ORACLE_H
#define RECON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_volume {
    int32_t grid[3];        /* gx, gy, gz */
    float voxel_size[3];
    float offset[3];
    float truncation;
    float max_weight;
    float max_depth;
    float fx, fy, cx, cy;
    int32_t weight_type;    /* 0: weight 1; 1: sigma(1)^2 / sigma(d)^2 */
    float noise[3];         /* sigma(d) = noise[0] + noise[1] * (d - noise[2])^2 */
} oracle_volume;

/* Pose {qw,qx,qy,qz,tx,ty,tz} narrowed to float and expanded to a row-major
 * 3x3 rotation + translation in float32 (tsdf_volume.cu:271-272,
 * se3.cuh:42-57,124-140). */
void oracle_pose_matrix_f32(const double pose[7], float R[9], float t[3]);

/* Voxel centre in float32, ((v + 0.5f) * vs) + off, each step rounded
 * (tsdf_volume.cu:117-119). */
void oracle_voxel_centre_f32(const oracle_volume* vol, int32_t x, int32_t y, int32_t z, float c[3]);

/* D = truncation, W = 0 (tsdf_volume.cu:227-254). */
void oracle_tsdf_init(const oracle_volume* vol, float* D, float* W);

/* One frame of cu_tsdf_integrate (tsdf_volume.cu:144-190) under the contract
 * above.  depth: rows x cols floats with `pitch` floats per row.  When
 * `update_mask` is non-NULL it receives 1 for every voxel updated by this
 * frame, 0 otherwise.  Returns the number of updated voxels.  OpenMP over
 * (y,z) rows with `nthreads` threads (<= 0: library default), which is the
 * reference kernel's own decomposition (tsdf_volume.cu:151-163). */
int64_t oracle_tsdf_integrate_f32(const oracle_volume* vol, float* D, float* W,
                                  const float* depth, int32_t rows, int32_t cols,
                                  int64_t pitch, const double pose[7],
                                  uint8_t* update_mask, int nthreads);

/* The same frame with every branch decision (pixel index, depth validity,
 * the sdf >= -truncation test) taken from the float32 path, and the values
 * (camera z, sdf, clamp, running average) computed in float64 from the
 * float-narrowed pose.  *decision_flips (may be NULL) receives the number of
 * voxels whose decision an independent float64 evaluation would change.
 * Returns the number of updated voxels. */
int64_t oracle_tsdf_integrate_f64(const oracle_volume* vol, double* D, double* W,
                                  const float* depth, int32_t rows, int32_t cols,
                                  int64_t pitch, const double pose[7],
                                  int64_t* decision_flips, int nthreads);

/* Marching cubes, class pass (marching_cubes.cu:76-84,111-130).  Fills
 * class_hist[256] (may be NULL) over all (gx-1)(gy-1)(gz-1) cubes and returns
 * the triangle count. */
int64_t oracle_mc_count(const float* field, const int32_t grid[3],
                        int64_t* class_hist, int nthreads);

/* Marching cubes emission (marching_cubes.cu:23-74,86-189,191-236): triangle
 * soup in cube order z, y, x, slot.  Writes at most `capacity` triangles:
 * verts 9 floats, tris 3 int32 (3t,3t+1,3t+2), edge_ids 3 int32 per triangle
 * (the table edge of each vertex; may be NULL), cube_ids (linear cube index
 * x + (gx-1)*(y + (gy-1)*z); may be NULL).  Returns the total count. */
int64_t oracle_mc_extract(const float* field, const int32_t grid[3],
                          const float voxel_size[3], const float offset[3],
                          float* verts, int32_t* tris, int32_t* edge_ids,
                          int64_t* cube_ids, int64_t capacity, int nthreads);

/* point_cloud_extract (cuda_point_cloud.cu:22-70): float64 back-projection,
 * float32 transform.  Writes every pixel: invalid (d <= 0) pixels get
 * (0,0,0) and valid = 0.  colors_out may be NULL; rgb may be NULL (white). */
void oracle_depth_to_points(const float* depth, int32_t rows, int32_t cols, int64_t pitch,
                            double fx, double fy, double cx, double cy,
                            const double transform[7],
                            const uint8_t* rgb, int64_t rgb_pitch,
                            float* xyz, uint8_t* valid, uint8_t* colors_out);

/* The colour integrate: oracle_tsdf_integrate_f32 plus, for voxels updated
 * with -trunc <= sdf <= trunc, c = ((c*wc) + rgb)/(wc+1) per channel and
 * wc = min(wc+1, max_weight) in C (4 floats per voxel).  No reference
 * counterpart on the dense path (octree.cu:118-137 averages colour per node). */
int64_t oracle_tsdf_integrate_color_f32(const oracle_volume* vol, float* D, float* W, float* C,
                                        const float* depth, int32_t rows, int32_t cols,
                                        int64_t pitch, const uint8_t* rgb, int64_t rgb_pitch,
                                        const double pose[7], int nthreads);
/* Colour at float32 points from C: the device sampler's support and
 * coefficients, neighbours weighted by coef * wc, rounded to 0..255. */
void oracle_sample_colors_f32(const oracle_volume* vol, const float* C, const float* points,
                              int64_t npoints, uint8_t* out);

/* Sparse (brick) volume: the dense lattice and update over 8x8x8-voxel
 * bricks allocated along each frame's truncation band (sparse_oracle.c). */
typedef struct oracle_sparse oracle_sparse;
oracle_sparse* oracle_sparse_create(const oracle_volume* vol);
void oracle_sparse_destroy(oracle_sparse* s);
void oracle_sparse_set_bounds(oracle_sparse* s, const int32_t lo[3], const int32_t hi[3]);
void oracle_sparse_allocate_box(oracle_sparse* s, const int32_t lo[3], const int32_t hi[3]);
int64_t oracle_sparse_integrate(oracle_sparse* s, const float* depth, int32_t rows, int32_t cols,
                                int64_t pitch, const double pose[7]);
int64_t oracle_sparse_num_bricks(const oracle_sparse* s);
void oracle_sparse_export(const oracle_sparse* s, const int32_t origin[3], const int32_t size[3],
                          float* D, float* W);
int64_t oracle_sparse_bricks(const oracle_sparse* s, int32_t* coords, float* D, float* W,
                             int64_t capacity);

/* Welded marching cubes (no reference counterpart; the reference emits a
 * soup).  One vertex per grid edge whose end values differ in sign, numbered
 * by (owning lower-end voxel, axis x, y, z); the vertex runs from the negative
 * end (equal to the soup's vertex unless the other end is exactly +-0);
 * normals are the lerped central-difference gradient, normalised.  Triangles
 * are the soup's, in its order, as int32 vertex indices.  Writes up to the
 * capacities; returns the triangle count and *num_vertices. */
int64_t oracle_mc_weld(const float* field, const int32_t grid[3], const float voxel_size[3],
                       const float offset[3], float* verts, float* normals, int64_t vertex_capacity,
                       int32_t* tris, int64_t triangle_capacity, int64_t* num_vertices);

/* project_to_depth / project_to_color_depth + filter_image_if
 * (cuda_point_cloud.cu:229-353, 89-99): points in index order, world to
 * camera with the float32 pose matrix, pixel = trunc((X*fx)/Z + cx) in
 * float32, Z == 0 skipped, a pixel keeps the smallest Z (the first of equal
 * ones) and that point's colour (255,255,255 when colors is NULL); unwritten
 * pixels get depth 0 and colour 0.  valid (may be NULL) skips points whose
 * byte is 0.  depth is rows x pitch floats; color_out (may be NULL) is rows x
 * color_pitch bytes of rgb triples. */
void oracle_render_points(const float* xyz, const uint8_t* valid, const uint8_t* colors,
                          int64_t num_points, const double pose[7], float fx, float fy, float cx,
                          float cy, int32_t rows, int32_t cols, int64_t pitch, float* depth,
                          uint8_t* color_out, int64_t color_pitch);

/* point_cloud_from_depth (point_cloud_utility.cpp:76-115): all-float64 path,
 * keeps min < d < max, transform applied in float64 (the caller's Sophus SE3
 * as a rotation from the normalised quaternion).  Writes up to `capacity`
 * points (3 doubles each) and their pixel indices; returns the count. */
int64_t oracle_point_cloud_from_depth_f64(const float* depth, int32_t rows, int32_t cols,
                                          int64_t pitch, double fx, double fy, double cx,
                                          double cy, const double transform[7],
                                          double depth_min, double depth_max,
                                          double* xyz, int64_t* pixel_index,
                                          int64_t capacity);

/* ---- differentiable integrate (adjoint_oracle.c), float64 ---------------- */
/* Rt: per frame a world->camera row-major rotation (9 doubles) then the
 * translation (3 doubles).  Depth is sampled bilinearly: depth[j][i] sits at
 * pixel coordinates (i, j); a voxel needs z > 0, its 2x2 taps inside the
 * image and all four valid.  Otherwise as oracle_tsdf_integrate_f32. */
void oracle_diff_forward_f64(const oracle_volume* vol, double* D, double* W, const float* depth,
                             int32_t rows, int32_t cols, const double* Rt, int32_t nframes,
                             int nthreads);
/* The same forward in float32 in the GPU kernel's operation order; rt: 12
 * floats per frame [R row-major | t].  D, W start as oracle_tsdf_init. */
void oracle_diff_forward_f32(const oracle_volume* vol, float* D, float* W, const float* depth,
                             int32_t rows, int32_t cols, const float* rt, int32_t nframes,
                             int nthreads);
/* Trilinear support of a world point (tsdf_volume.cu:46-104). */
int oracle_trilinear_support(const oracle_volume* vol, const double pw[3], int64_t ids[8],
                             double coef[8]);
/* L = (1/K) sum_k w_k r_k^2, r_k = trilinear TSDF at point k; lambda (may be
 * NULL) <- dL/dD. */
double oracle_surface_loss_f64(const oracle_volume* vol, const double* D, const double* points,
                               const double* weights, int64_t npoints, double* lambda);
/* dL/dxi (6 per frame, xi = (rho, omega), left perturbation) and dL/ddepth
 * (may be NULL) from lambda = dL/dD of the final volume, by per-voxel
 * recomputation of the frame decisions. */
void oracle_diff_adjoint_f64(const oracle_volume* vol, const float* depth, int32_t rows,
                             int32_t cols, const double* Rt, int32_t nframes, const double* lambda,
                             double* grad_pose, double* grad_depth);
/* Per-frame (model-to-frame) loss of pose refinement: pair k = (frame
 * pairs[2k], point pairs[2k+1]) has residual r = d(s) - z, the frame's
 * bilinear depth at the projection s of world point p = points[3*point]
 * minus the point's camera z (0 when the sample is invalid: z <= 0, the 2x2
 * footprint leaves the image or a tap is not a valid depth).  With tau > 0
 * a residual of |r| >= tau counts tau^2 and has no gradient (a truncated
 * quadratic).  Returns L = sum_k w_k min(r_k^2, tau^2) / npairs (w =
 * weights, NULL for all 1) and adds dL/dxi (6 per frame, left perturbation,
 * the adjoint's convention) into grad_pose (may be NULL); residuals (may be
 * NULL) receives r. */
double oracle_frame_loss_f64(const oracle_volume* vol, const float* depth, int32_t rows,
                             int32_t cols, const double* Rt, int32_t nframes, const double* points,
                             const int32_t* pairs, const double* weights, int64_t npairs,
                             double tau, double* grad_pose, double* residuals);
/* Gauss-Newton system of that term, (nframes, 6, 6) added to H and
 * (nframes, 6) to grad (may be NULL): the point-to-plane linearisation with
 * world normals (one per point), or the bilinear depth slopes when normals is
 * NULL. */
void oracle_frame_normal_matrix_f64(const oracle_volume* vol, const float* depth, int32_t rows,
                                    int32_t cols, const double* Rt, int32_t nframes,
                                    const double* points, const double* normals,
                                    const int32_t* pairs, const double* weights, int64_t npairs,
                                    double tau, double* H, double* grad);
/* Per-point weights: 1 when no decision of the point's support voxels (frame
 * validity, footprint, the -trunc skip, the +trunc clamp) can change under a
 * perturbation of size <= h of one pose coordinate or one depth tap, and the
 * samples cross no pixel line with a bilinear slope jump above `kink`. */
void oracle_stable_points(const oracle_volume* vol, const float* depth, int32_t rows,
                          int32_t cols, const double* Rt, int32_t nframes, const double* points,
                          int64_t npoints, double h, double kink, double* weights);

#ifdef __cplusplus
}
#endif

#endif /* RECON_ORACLE_H */
