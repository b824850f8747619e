/*
 * recon.h -- C-ABI of the MI355X (gfx950) TSDF fusion path.
 *
 * Plain C: pointers, sizes and status codes only.  Every device pointer is a
 * HIP device allocation; every `stream` argument is a hipStream_t passed as
 * an opaque pointer (NULL = the default stream).  Calls marked "async" only
 * enqueue work on the stream; calls marked "sync" wait for the stream before
 * returning.  A handle is not thread-safe; distinct handles are independent.
 *
 * Reference interface replaced (cu-Reconstruction, paths relative to its root):
 *   TSDFVolumeConfig              modules/include/cuda/tsdf_volume.cuh:17-29
 *   TSDFVolume ctor / dtor        modules/src/cuda/tsdf_volume.cu:194-254
 *   TSDFVolume::integrate         modules/include/cuda/tsdf_volume.cuh:49-51,
 *                                 modules/src/cuda/tsdf_volume.cu:257-290
 *   accessors                     modules/include/cuda/tsdf_volume.cuh:57-92
 *   generate_triangular_surface   modules/include/cuda/marching_cubes.cuh:13-15,
 *                                 modules/src/cuda/marching_cubes.cu:191-236
 *   cudaPointCloud::extract_points modules/include/cuda/cuda_point_cloud.cuh:55-58,
 *                                 modules/src/cuda/cuda_point_cloud.cu:179-226
 * The reference reports errors as bool + a log line (tsdf_volume.cu:260-268);
 * here each failure class has its own negative status and recon_last_error()
 * returns the message.
 */
#ifndef RECON_RECON_H
#define RECON_RECON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RECON_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
    RECON_OK = 0,
    RECON_ERR_NULL = -1,      /* a required pointer argument is NULL            */
    RECON_ERR_SIZE = -2,      /* zero / negative / inconsistent image or buffer size */
    RECON_ERR_CONFIG = -3,    /* invalid volume configuration                   */
    RECON_ERR_OOM = -4,       /* device allocation failed                       */
    RECON_ERR_HIP = -5,       /* a HIP runtime call or kernel launch failed     */
    RECON_ERR_CAPACITY = -6,  /* output buffer too small (count still reported) */
    RECON_ERR_RANGE = -7,     /* sizes overflow the kernels' 32-bit indexing    */
    RECON_ERR_DEVICE = -8     /* no usable gfx950 device / wrong device         */
};

typedef void* recon_stream_t; /* hipStream_t */

#if defined(__GNUC__)
#define RECON_API __attribute__((visibility("default")))
#else
#define RECON_API
#endif

/* ---- volume configuration ---------------------------------------------- */
/* Mirrors TSDFVolumeConfig (tsdf_volume.cuh:17-29).  The last three fields are
 * additions: the reference hard-codes truncation = 1.1*|voxel_size|
 * (tsdf_volume.cu:199), max depth = 24 (tsdf_volume.cu:169) and never applies
 * max_weight (tsdf_volume.cu:179-185).  A value <= 0 in `truncation` or
 * `max_depth` selects the reference's constant. */
typedef struct recon_tsdf_config {
    int32_t voxel_grid_size[3];   /* gx, gy, gz (voxels)                        */
    float physical_size[3];       /* metres; voxel_size = physical / grid       */
    float global_offset[3];       /* world position of the grid's min corner    */
    float max_weight;             /* weight cap, >= 1                            */
    float focal_x, focal_y;       /* pinhole intrinsics of the depth camera     */
    float cx, cy;
    uint32_t camera_width;        /* informational, as in the reference         */
    uint32_t camera_height;
    float max_cell_size;          /* unused by the reference; kept for layout   */
    float truncation;             /* metres; <= 0 -> 1.1*|voxel_size|           */
    float max_depth;              /* metres; <= 0 -> 24                         */
    /* WEIGHT_TYPE (tsdf_volume.cuh:12-15, declared and unused there):
     * RECON_WEIGHT_BY_DEPTH (0) weighs every observation 1, the reference's
     * update; RECON_WEIGHT_BY_VARIANCE (1) weighs a depth sample d by
     * sigma(1 m)^2 / sigma(d)^2 with the axial noise model
     * sigma(d) = a + b*(d - c)^2 (Nguyen, Izadi, Lovell 2012), so
     * D' = (D*W + w*tsdf) / (W + w), W' = min(W + w, max_weight). */
    int32_t weight_type;
    float noise_model[3];         /* a, b, c; all 0 -> 0.0012, 0.0019, 0.4     */
} recon_tsdf_config;

#define RECON_WEIGHT_BY_DEPTH 0
#define RECON_WEIGHT_BY_VARIANCE 1

/* Read-only description of a created volume. */
typedef struct recon_tsdf_info {
    int32_t grid[3];
    float voxel_size[3];
    float global_offset[3];
    float truncation;
    float max_weight;
    float max_depth;
    float intrinsics[4];          /* fx, fy, cx, cy */
    int64_t num_voxels;
    int32_t weight_type;          /* as configured */
    float noise_model[3];         /* a, b, c in use */
} recon_tsdf_info;

/* Pitched single-channel float32 image in device memory, the shape of
 * cv::cuda::PtrStepSzf (data, step in bytes, rows, cols). */
typedef struct recon_depth_view {
    const float* data;
    size_t step;                  /* bytes between rows, >= 4*cols, multiple of 4 */
    int32_t rows;
    int32_t cols;
} recon_depth_view;

typedef struct recon_tsdf recon_tsdf; /* opaque handle */

/* ---- library ------------------------------------------------------------ */
/* Thread-local message of the last failing call on this thread ("" if none). */
RECON_API const char* recon_last_error(void);
RECON_API int recon_abi_version(void);
/* sync.  Fills a volume description without allocating: validates `cfg`. */
RECON_API int recon_tsdf_describe(const recon_tsdf_config* cfg, recon_tsdf_info* info);

/* ---- lifecycle (TSDFVolume ctor/dtor, tsdf_volume.cu:194-254) ---------- */
/* sync.  Allocates D and W (float32[gz][gy][gx], linear x + gx*(y + gy*z)) on
 * the current device and sets D = truncation, W = 0 (tsdf_volume.cu:227-254).
 * No voxel-centre table is allocated: centres are computed analytically. */
RECON_API int recon_tsdf_create(const recon_tsdf_config* cfg, recon_tsdf** out);
RECON_API void recon_tsdf_destroy(recon_tsdf* h);
/* async.  D = truncation, W = 0. */
RECON_API int recon_tsdf_reset(recon_tsdf* h, recon_stream_t stream);
RECON_API int recon_tsdf_get_info(const recon_tsdf* h, recon_tsdf_info* info);

/* ---- accessors (tsdf_volume.cuh:57-92) --------------------------------- */
/* async.  The reference's per-voxel transformation records
 * (TSDFVolume::transformation_voxels_data, tsdf_volume.cuh:33-38,84-90,
 * filled by tsdf_volume.cu:106-127), which this library does not store:
 * 28-byte packed {float3 translation = ((v+0.5)*vs)+off; float4
 * quat_rotation = (x 0, y 0, z 0, w 1)} for every voxel into `out` (device,
 * 28*N bytes). */
RECON_API int recon_tsdf_transformation_voxels(recon_tsdf* h, void* out, recon_stream_t stream);
RECON_API float* recon_tsdf_distances(recon_tsdf* h);   /* device pointer, N floats */
RECON_API float* recon_tsdf_weights(recon_tsdf* h);     /* device pointer, N floats */
/* sync.  Copies D and/or W (either may be NULL) to host memory of N floats. */
RECON_API int recon_tsdf_download(recon_tsdf* h, float* host_d, float* host_w, recon_stream_t stream);
/* sync.  Copies D and/or W from host memory of N floats into the volume. */
RECON_API int recon_tsdf_upload(recon_tsdf* h, const float* host_d, const float* host_w, recon_stream_t stream);

/* ---- integrate (TSDFVolume::integrate, tsdf_volume.cu:144-190,257-290) -- */
/* async.  One depth frame (metres, float32, device memory) seen from
 * camera_pose = {qw,qx,qy,qz,tx,ty,tz}, world->camera, narrowed to float32
 * as the reference does (tsdf_volume.cu:271-272).  `rgb` (pitched uchar3, the
 * depth's size) may be NULL; it is ignored, as in the reference, unless
 * colour is enabled on the volume (recon_tsdf_enable_color): then a voxel
 * updated inside the truncation band (-trunc <= sdf <= trunc) averages the
 * rgb pixel of its depth sample into its colour, c' = ((c*wc) + rgb)/(wc+1)
 * per channel, wc' = min(wc+1, max_weight). */
RECON_API int recon_tsdf_integrate(recon_tsdf* h, const recon_depth_view* depth,
                         const uint8_t* rgb, size_t rgb_step,
                         const double camera_pose[7], recon_stream_t stream);
/* async.  Same, for `num_frames` frames applied in order in one pass over
 * the volume: depth frame f starts at depth0 + f*frame_stride_bytes, all
 * frames share rows/cols/step, poses are num_frames*7 doubles.  Results are
 * identical to num_frames calls of recon_tsdf_integrate.  The depth maps must
 * stay valid until the stream reaches this work. */
RECON_API int recon_tsdf_integrate_frames(recon_tsdf* h, const recon_depth_view* depth0,
                                size_t frame_stride_bytes, int32_t num_frames,
                                const double* camera_poses, recon_stream_t stream);
/* When `counter` (device uint64) is non-NULL, later integrate calls add the
 * number of voxel updates they perform to it (diagnostic; slower path).
 * Pass NULL to switch counting off. */
RECON_API int recon_tsdf_set_update_counter(recon_tsdf* h, unsigned long long* counter);
/* The same for the number of distinct voxels each sweep over the volume
 * updates at least once: one sweep per recon_tsdf_integrate call, one per
 * 16 frames of a recon_tsdf_integrate_frames call (the voxels whose D and W
 * a sweep reads and writes).  No counterpart in the reference; colour
 * integrates do not count. */
RECON_API int recon_tsdf_set_touch_counter(recon_tsdf* h, unsigned long long* counter);

/* ---- marching cubes (generate_triangular_surface, marching_cubes.cu) --- */
/* The mesh is the reference's triangle soup (marching_cubes.cu:159-189):
 * triangle t owns vertices 3t, 3t+1, 3t+2 and triangles are ordered by cube
 * z (outer), y, x, then table slot.  verts: 9 floats per triangle (x,y,z of
 * its three vertices); tris: 3 int32 per triangle (the vertex indices).
 * sync.  Total triangle count of the current volume. */
RECON_API int recon_tsdf_count_triangles(recon_tsdf* h, int64_t* num_triangles, recon_stream_t stream);
/* sync.  Writes min(count, capacity) triangles; *num_triangles = count.
 * Returns RECON_ERR_CAPACITY when count > capacity.  `tris` may be NULL. */
RECON_API int recon_tsdf_extract_mesh(recon_tsdf* h, float* verts, int32_t* tris,
                            int64_t capacity_triangles, int64_t* num_triangles,
                            recon_stream_t stream);
/* Same two calls over any dense float32[gz][gy][gx] field in device memory,
 * with corner positions (v + 0.5) * voxel_size + offset. */
RECON_API int recon_mc_count(const float* field, const int32_t grid[3], const float voxel_size[3],
                   const float offset[3], int64_t* num_triangles, recon_stream_t stream);
RECON_API int recon_mc_extract(const float* field, const int32_t grid[3], const float voxel_size[3],
                     const float offset[3], float* verts, int32_t* tris,
                     int64_t capacity_triangles, int64_t* num_triangles,
                     recon_stream_t stream);
/* No counterpart in the reference: recon_mc_extract for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and status
 * codes as recon_mc_extract, but it calls no hipMalloc / hipFree: its
 * scratch (2 B per 64-cube chunk of each of the (gy-1)*(gz-1) cube rows,
 * plus 20 B per 16 cube rows) lives in one static 48 MiB device array per
 * GPU, held until the process ends; a 1024^3 grid fits, and grids that need
 * more fail with RECON_ERR_RANGE.  Runs on the current device.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other host
 * threads, or other streams with work in flight) race.  sync. */
RECON_API int recon_mc_extract_noalloc(const float* field, const int32_t grid[3],
                                       const float voxel_size[3], const float offset[3],
                                       float* verts, int32_t* tris, int64_t capacity_triangles,
                                       int64_t* num_triangles, recon_stream_t stream);
/* async.  hist_device (256 uint64, device memory) <- number of cubes of each
 * class, bit i of a class set iff corner i is < 0 (marching_cubes.cu:76-84).
 * The reference has no such call; it exposes the classification that the
 * extraction is built on. */
RECON_API int recon_mc_class_histogram(const float* field, const int32_t grid[3],
                             unsigned long long* hist_device, recon_stream_t stream);

/* ---- colour ----------------------------------------------------------------- */
/* No counterpart on the reference's dense path (its integrate ignores colour;
 * its octree keeps a colour average per node, octree.cu:118-137).  enable != 0
 * allocates (once) a colour volume of 4 floats per voxel (r, g, b, weight),
 * zeroed, which later recon_tsdf_integrate calls with an rgb image update and
 * recon_tsdf_reset clears; enable == 0 frees it.  async (enable), sync
 * (disable). */
RECON_API int recon_tsdf_enable_color(recon_tsdf* h, int enable, recon_stream_t stream);
RECON_API float* recon_tsdf_colors(recon_tsdf* h);  /* device pointer, 4N floats, or NULL */
/* sync.  Copies the colour volume (4N floats) to host memory. */
RECON_API int recon_tsdf_download_colors(recon_tsdf* h, float* host, recon_stream_t stream);
/* async.  Colour at `num_points` world points (3 floats each, device): the
 * trilinear support of the TSDF sampler, each neighbour weighted by its
 * coefficient times its colour weight, c = sum(k*w*c) / sum(k*w) in
 * neighbour order, rounded to 0..255 (0,0,0 without colour or support);
 * 3 bytes per point into colors_out (device). */
RECON_API int recon_tsdf_sample_colors(recon_tsdf* h, const float* points, int64_t num_points,
                                       uint8_t* colors_out, recon_stream_t stream);

/* ---- sparse volume --------------------------------------------------------- */
/* No counterpart on the reference's path (a dense grid, tsdf_volume.cu:194-254;
 * octree.cu is an unfinished sparse prototype).  The dense volume's voxel
 * lattice -- voxel v (int32 triple, |v| < 2^23) centred at ((v+0.5)*vs)+off,
 * vs = physical_size / voxel_grid_size of `cfg` -- stored in 8x8x8-voxel
 * bricks, brick b = floor(v/8), found through a hash table; at most
 * `brick_capacity` (1..2^24) bricks, 4 KiB each.  Unallocated voxels read as
 * D = truncation, W = 0.  The grid size of `cfg` sets the voxel size only:
 * the sparse volume is unbounded. */
typedef struct recon_sparse recon_sparse;
RECON_API int recon_sparse_create(const recon_tsdf_config* cfg, int64_t brick_capacity,
                                  recon_sparse** out);               /* sync */
RECON_API void recon_sparse_destroy(recon_sparse* h);
/* async.  Frees every brick. */
RECON_API int recon_sparse_reset(recon_sparse* h, recon_stream_t stream);
/* async.  One depth frame: allocates every brick that meets, one voxel wide,
 * the world box of some valid pixel's frustum slab between d - trunc and
 * d + trunc, then applies the dense update rule of recon_tsdf_integrate to
 * every voxel of every allocated brick.  Voxels the dense integrate updates
 * inside the band always lie in allocated bricks; with every brick already
 * allocated the result equals the dense integrate bit for bit. */
RECON_API int recon_sparse_integrate(recon_sparse* h, const recon_depth_view* depth,
                                     const double camera_pose[7], recon_stream_t stream);
/* Host only.  Later integrate calls allocate only bricks that meet the
 * voxel box [lo, hi) (NULL, NULL: unbounded, the default). */
RECON_API int recon_sparse_set_bounds(recon_sparse* h, const int32_t lo[3], const int32_t hi[3]);
/* async.  Allocates every brick that meets the voxel box [lo, hi). */
RECON_API int recon_sparse_allocate_box(recon_sparse* h, const int32_t lo[3], const int32_t hi[3],
                                        recon_stream_t stream);
/* sync.  Allocated bricks and overflow flags (bit 0: the pool filled and
 * bricks were dropped, bit 1: a pixel slab spanned more than 8 bricks on an
 * axis and was cut, bit 2: a slab left the |v| < 2^23 range, bit 3: a brick
 * was dropped after 128 hash probes). */
RECON_API int recon_sparse_stats(recon_sparse* h, int64_t* num_bricks, uint32_t* overflow,
                                 recon_stream_t stream);
/* async.  Dense box of voxels [origin, origin + size) into device arrays
 * (x fastest; D and/or W may be NULL) -- e.g. for marching cubes. */
RECON_API int recon_sparse_export(recon_sparse* h, const int32_t origin[3], const int32_t size[3],
                                  float* D, float* W, recon_stream_t stream);
/* sync.  Brick coordinates (3 int32 each) and contents (512 floats each,
 * voxel (x, y, z) of a brick at 64z + 8y + x) to host memory, in pool order
 * (the allocation order, which may vary between runs). */
RECON_API int recon_sparse_download_bricks(recon_sparse* h, int32_t* coords, float* D, float* W,
                                           int64_t capacity, int64_t* num_bricks,
                                           recon_stream_t stream);

/* ---- welded mesh ---------------------------------------------------------- */
/* sync.  No counterpart in the reference, whose extraction emits a triangle
 * soup (marching_cubes.cu:171-179 never shares a vertex): the same surface
 * with one vertex per grid edge whose end values differ in sign (v < 0
 * against not), about a third of the soup's vertices.  Vertices are numbered
 * by (owning lower-end voxel x + gx*(y + gy*z), axis x, y, z); each runs from
 * the negative end N to the other end M, p = sN + r*(sM - sN) with
 * r = (-dN)/(dM - dN), the soup's formula (bit-identical to the soup's
 * vertex except where dM is exactly +-0).  Normals (may be NULL) are the
 * field's gradient -- central differences over 2*vs inside the grid,
 * one-sided over vs on its border -- lerped from N to M with r and
 * normalised, (0,0,0) where it vanishes.  Triangles (may be NULL) are the
 * soup's, in its order, as int32 vertex indices.  Always returns both
 * counts; with every output NULL it only counts.  Fails with
 * RECON_ERR_CAPACITY, writing nothing, when a requested output does not fit,
 * and with RECON_ERR_RANGE above 2^31-1 vertices. */
RECON_API int recon_mc_extract_welded(const float* field, const int32_t grid[3],
                                      const float voxel_size[3], const float offset[3],
                                      float* vertices, float* normals, int64_t vertex_capacity,
                                      int32_t* triangles, int64_t triangle_capacity,
                                      int64_t* num_vertices, int64_t* num_triangles,
                                      recon_stream_t stream);
/* sync.  The same on a volume's TSDF (the handle's scratch is reused). */
RECON_API int recon_tsdf_extract_welded_mesh(recon_tsdf* h, float* vertices, float* normals,
                                             int64_t vertex_capacity, int32_t* triangles,
                                             int64_t triangle_capacity, int64_t* num_vertices,
                                             int64_t* num_triangles, recon_stream_t stream);

/* ---- depth -> point cloud (point_cloud_extract, cuda_point_cloud.cu:22-70) */
/* async.  Pixel (x,y) with depth d>0: p = ((x-cx)*d/fx, (y-cy)*d/fy, d)
 * computed in float64 and rounded to float32, then p_world = T*p in float32
 * with transform = {qw,qx,qy,qz,tx,ty,tz} (camera->world, as the reference
 * uses it, cuda_point_cloud.cu:60).  Output index y*cols + x:
 * xyz[3*i..3*i+2], valid[i] = 1; pixels with d <= 0 write (0,0,0) and
 * valid = 0 (the reference leaves them unwritten).  `colors_out` (3 bytes
 * per point, may be NULL) receives the rgb pixel, or 255,255,255 when `rgb`
 * is NULL, or 0,0,0 for invalid pixels. */
RECON_API int recon_depth_to_points(const recon_depth_view* depth, double fx, double fy,
                          double cx, double cy, const double transform[7],
                          const uint8_t* rgb, size_t rgb_step,
                          float* xyz, uint8_t* valid, uint8_t* colors_out,
                          recon_stream_t stream);

/* ---- point cloud -> depth / colour z-buffer render ------------------------ */
/* async.  Replaces cudaPointCloud::project_to_depth / project_to_color_depth
 * with filter_image_if (cuda_point_cloud.cu:229-353, 89-99), which splat on
 * the host.  Point i (xyz[3i..3i+2], world) is moved to the camera with
 * camera_pose = {qw,qx,qy,qz,tx,ty,tz} (world->camera, float32 as integrate
 * narrows it) and lands on pixel (trunc((X*fx)/Z + cx), trunc((Y*fy)/Z + cy))
 * in float32 with the intrinsics narrowed to float; points with Z == 0, with
 * valid[i] == 0 (valid may be NULL) or off the image are skipped.  A pixel
 * keeps the smallest Z, the lowest index among equal Z, and that point's
 * colours[3i..3i+2] (255,255,255 when colours is NULL); other pixels get depth
 * 0 and colour 0.  Results do not depend on the order of execution.  depth is
 * rows x depth_step bytes of float; color_out (may be NULL) rows x color_step
 * bytes of rgb triples; zbuffer is device scratch of rows*cols 8-byte words.
 * The reference reads cx, cy from K[2], K[5], 0 in the column-major K it
 * fills; here they are explicit.  num_points < 2^32. */
RECON_API int recon_points_render(const float* xyz, const uint8_t* valid, const uint8_t* colors,
                                  int64_t num_points, const double camera_pose[7], double fx,
                                  double fy, double cx, double cy, float* depth,
                                  size_t depth_step, int32_t rows, int32_t cols,
                                  uint8_t* color_out, size_t color_step,
                                  unsigned long long* zbuffer, recon_stream_t stream);

/* ---- differentiable integrate -------------------------------------------- */
/* No counterpart in the reference.  An extension of cu_tsdf_integrate
 * (tsdf_volume.cu:144-190) whose depth lookup is bilinear, so that the
 * volume is differentiable in the camera poses and the depth maps:
 * depth[j][i] sits at pixel coordinates (i, j); a voxel samples a frame when
 * its camera z > 0, its continuous pixel coordinates s = ((x*f)/z) + c lie in
 * [0, size-1) on both axes and the four taps are valid (0 < d <= max_depth).
 * The update is otherwise the plain one (skip below -trunc, clamp at +trunc,
 * weight cap).  Poses are DEVICE float32 world->camera matrices, 12 per frame
 * (row-major R, then t), so a caller can refine them on the device.  Calls
 * grow per-handle scratch on first use (hipMalloc); a graph capture should
 * follow one uncaptured call of the same sizes. */

/* host.  poses (num_poses*7 doubles {qw,qx,qy,qz,tx,ty,tz}) -> rt (12 floats
 * each) narrowed as recon_tsdf_integrate narrows them (tsdf_volume.cu:271-272,
 * se3.cuh:42-57). */
RECON_API int recon_pose_to_matrix(const double* poses, int32_t num_poses, float* rt);
/* async.  Integrates num_frames frames (depth frame f at depth0 + f *
 * frame_stride_bytes, at least 2x2) in order with bilinear sampling. */
RECON_API int recon_tsdf_diff_integrate(recon_tsdf* h, const recon_depth_view* depth0,
                                        size_t frame_stride_bytes, int32_t num_frames,
                                        const float* poses_rt, recon_stream_t stream);
/* async.  L = (1/K) sum_k w_k r_k^2, r_k the trilinear TSDF at world point k
 * (interpolate_trilinearly, tsdf_volume.cu:46-104, with neighbour (1,1,1)
 * weighted x*y*z and (0,1,1) (1-x)*y*z, which the reference swaps; a point
 * whose 2x2x2 support leaves the grid has r = 0 and no gradient).  Device
 * arrays: points 3K floats, weights K floats (NULL: all 1), residuals K
 * floats (may be NULL), *loss one double.  lambda_ids / lambda (8K each, both
 * or neither) receive dL/dD as (voxel index, value) pairs, 8 per point in
 * the sampler's neighbour order, index -1 for a point without support. */
RECON_API int recon_tsdf_surface_loss(recon_tsdf* h, const float* points, const float* weights,
                                      int64_t num_points, float* residuals, double* loss,
                                      int64_t* lambda_ids, float* lambda, recon_stream_t stream);
/* async.  Given dL/dD of the volume after recon_tsdf_diff_integrate of the
 * same frames and poses (the pairs of recon_tsdf_surface_loss, or any list;
 * repeated indices add; indices outside the volume are ignored):
 * grad_pose (device, num_frames*6 doubles) <- dL/dxi_f, xi = (rho, omega)
 * the left perturbation T_f <- Exp(xi) T_f (a camera point moves by
 * rho + omega x p); grad_depth (device, may be NULL; the depth frames'
 * layout) <- dL/ddepth.  The +trunc clamp and the skip below -trunc are
 * taken as constant (their subgradient 0).  Pose sums are reduced in a fixed
 * order; depth sums are exact int64 fixed-point sums (each contribution
 * rounded to 2^-41 of the largest |dL/dD|, then added in any order): both
 * are run-to-run identical. */
RECON_API int recon_tsdf_diff_adjoint(recon_tsdf* h, const recon_depth_view* depth0,
                                      size_t frame_stride_bytes, int32_t num_frames,
                                      const float* poses_rt, const int64_t* lambda_ids,
                                      const float* lambda, int64_t num_lambda, double* grad_pose,
                                      float* grad_depth, recon_stream_t stream);

/* ---- batched pose refinement --------------------------------------------- */
/* No counterpart in the reference.  Many independent fragments (volumes of
 * their own size, frames and image size), each with its surface points,
 * fused with recon_tsdf_diff_integrate's bilinear update and scored with
 * recon_tsdf_surface_loss, over one arena; the poses of all frames are
 * refined on the device.  All device memory is allocated at creation. */
typedef struct recon_diff_batch recon_diff_batch;

typedef struct recon_diff_volume {
    recon_tsdf_config config;  /* volume and the intrinsics of its frames */
    int32_t num_frames;        /* 1..64 */
    int32_t rows, cols;        /* depth image size (>= 2x2) */
    size_t depth_step;         /* bytes per depth row */
    size_t frame_stride;       /* bytes between frames */
    const float* depth;        /* device: frame f at depth + f * frame_stride */
    const float* points;       /* device: 3 floats per surface point (world) */
    const float* weights;      /* device: one per point, or NULL (all 1) */
    int64_t num_points;
    /* Model-to-frame term, optional (num_frame_pairs 0): pair k = (frame
     * frame_pairs[2k], point frame_pairs[2k+1]) scores that frame's depth
     * against that point, r = d(s) - z: d(s) the bilinear depth at the
     * point's projection s (the diff_integrate sampling rule), z its camera
     * z; r = 0 when the sample is invalid.  The volume's loss adds
     * frame_weight * sum_k min(r_k^2, frame_trunc^2) / num_frame_pairs
     * (frame_trunc <= 0: not truncated; a truncated pair has no gradient),
     * which observes every frame's own pose (the fused TSDF averages the
     * frames).  Device int32 pairs, grouped by frame in increasing frame
     * order. */
    const int32_t* frame_pairs;
    int64_t num_frame_pairs;
    double frame_weight;
    double frame_trunc;
    /* Optional (NULL): device, 3 floats per point, the points' unit surface
     * normals (world).  Used only by recon_diff_batch_refine_gn, whose
     * Gauss-Newton system then takes the point-to-plane linearisation of
     * each pair's residual instead of the bilinear depth slopes. */
    const float* normals;
} recon_diff_volume;

typedef struct recon_adam_params {
    double lr_translation;     /* step on rho (metres) at the first step */
    double lr_rotation;        /* step on omega (radians) at the first step */
    double beta1, beta2, eps;  /* Adam moments */
    double decay;              /* both steps are multiplied by decay per step */
} recon_adam_params;

/* sync.  The depth maps and points must stay valid and unchanged while the
 * batch lives (their depth bounds are computed here, once). */
RECON_API int recon_diff_batch_create(const recon_diff_volume* volumes, int32_t num_volumes,
                                      recon_diff_batch** out);
RECON_API void recon_diff_batch_destroy(recon_diff_batch* b);
RECON_API int recon_diff_batch_volume_data(recon_diff_batch* b, int32_t i, float** distances,
                                           float** weights, recon_tsdf_info* info);
/* sync.  Copies volume i's D and/or W (either may be NULL) to host memory. */
RECON_API int recon_diff_batch_download(recon_diff_batch* b, int32_t i, float* host_d,
                                        float* host_w, recon_stream_t stream);
/* async.  Resets every volume and integrates all of its frames seen from
 * poses_rt (device, 12 floats per frame, volumes in order); loss (device,
 * one double per volume) <- each volume's surface loss. */
RECON_API int recon_diff_batch_forward(recon_diff_batch* b, const float* poses_rt, double* loss,
                                       recon_stream_t stream);
/* async.  grad_pose (device, 6 doubles per frame) <- d(loss_v)/dxi of each
 * frame's volume for the last recon_diff_batch_forward (same poses), xi the
 * left perturbation of recon_tsdf_diff_adjoint.  Run-to-run identical. */
RECON_API int recon_diff_batch_backward(recon_diff_batch* b, double* grad_pose,
                                        recon_stream_t stream);
/* sync.  `steps` steps of forward, backward and Adam on every frame's xi,
 * T <- Exp(-step) T with R re-orthonormalised, poses_rt updated in place;
 * one graph launch per step.  loss_history (device, steps * num_volumes
 * doubles, may be NULL) <- each step's losses before its update. */
RECON_API int recon_diff_batch_refine(recon_diff_batch* b, float* poses_rt, int32_t steps,
                                      const recon_adam_params* opt, double* loss_history,
                                      recon_stream_t stream);

typedef struct recon_gn_params {
    double step;         /* fraction of the Gauss-Newton step taken, (0, 1] */
    double damping;      /* Levenberg-Marquardt: solve (H + damping diag(H)) x = g */
    double trunc_ratio;  /* (0, 1]: at step k every volume's frame_trunc is scaled
                          * by trunc_ratio^min(1, k / max(1, steps / 2 - 1))
                          * (integer division); 1: fixed */
} recon_gn_params;

/* sync.  `steps` damped Gauss-Newton steps on every frame's xi: forward,
 * backward, then per frame x = M^-1 g, M = H + damping diag(H) +
 * 1e-12 max(diag H) I factored by Cholesky (a coordinate whose pivot is at
 * most 1e-9 max(diag M), one the pairs do not fix, takes no step).  H and
 * the model-to-frame part of g are the Gauss-Newton system of that term,
 * frame_weight * sum 2 J J^T / num_frame_pairs and frame_weight *
 * sum 2 r J / num_frame_pairs over the pairs with |r| < trunc, J = (j, p x j)
 * the residual's derivative with respect to xi linearised as
 * j = -(z / (n . p)) n from the point's normal n (camera frame) when the
 * volume has normals (point-to-plane; pairs seen at |n . p| < 0.1 |n| |p| add
 * nothing), else j = dr/dp from the bilinear depth slopes (the exact
 * gradient); the surface loss adds its exact gradient to g.
 * T <- Exp(-step x) T in float64, R re-orthonormalised, poses_rt updated in
 * place; a frame without pairs (H = 0) keeps its pose.  Every volume needs
 * frame pairs.  loss_history as recon_diff_batch_refine (each step's loss at
 * its own truncation).  No counterpart in the reference. */
RECON_API int recon_diff_batch_refine_gn(recon_diff_batch* b, float* poses_rt, int32_t steps,
                                         const recon_gn_params* opt, double* loss_history,
                                         recon_stream_t stream);

/* ---- batched volumes ----------------------------------------------------- */
/* No counterpart in the reference, which builds one TSDFVolume per scene and
 * blocks on every frame (tsdf_surface_reconstruction.cpp:125-138).  A batch
 * holds many independent volumes (sizes, frame counts and image resolutions
 * may all differ); one recon_batch_run replays a hipGraph that resets every
 * volume, integrates all of each volume's frames in order and meshes every
 * volume.  Results equal recon_tsdf_reset + recon_tsdf_integrate per frame +
 * recon_tsdf_extract_mesh on each volume. */
typedef struct recon_batch recon_batch;

typedef struct recon_batch_volume {
    recon_tsdf_config config;     /* volume and the camera intrinsics of its frames */
    int32_t num_frames;           /* 0..64 */
    int32_t rows, cols;           /* depth image size of every frame */
    size_t depth_step;            /* bytes per depth row */
    size_t frame_stride;          /* bytes between consecutive frames */
    const float* depth;           /* device; frame f at depth + f*frame_stride; must stay
                                     valid while the batch exists */
    const double* poses;          /* host, num_frames * 7 {qw,qx,qy,qz,tx,ty,tz}; copied */
} recon_batch_volume;

/* sync.  Plans the batch (device tables, one arena for all D and W, work
 * lists) and captures its graph.  mesh_capacity_triangles = 0 builds it
 * without meshing.  tune_cache_path (may be NULL) is a JSON file written by
 * recon_batch_autotune; volumes whose (grid, image size) class it lists use
 * the tile shape found there. */
RECON_API int recon_batch_create(const recon_batch_volume* volumes, int32_t num_volumes,
                                 int64_t mesh_capacity_triangles, const char* tune_cache_path,
                                 recon_batch** out);
RECON_API void recon_batch_destroy(recon_batch* b);
/* async.  One graph launch: reset, integrate and mesh every volume. */
RECON_API int recon_batch_run(recon_batch* b, recon_stream_t stream);
/* Device pointers of volume i's D and W (float32[gz][gy][gx]) and its description. */
RECON_API int recon_batch_volume_data(recon_batch* b, int32_t i, float** distances, float** weights,
                                      recon_tsdf_info* info);
/* sync.  ranges[2i], ranges[2i+1] <- first triangle and triangle count of
 * volume i in the shared vertex buffer *verts (9 floats per triangle, the
 * soup order of recon_tsdf_extract_mesh).  RECON_ERR_CAPACITY when the
 * meshes exceeded the capacity given at creation (the counts are still set). */
RECON_API int recon_batch_meshes(recon_batch* b, int64_t* ranges, float** verts,
                                 recon_stream_t stream);
/* sync.  Copies volume i's D and/or W (either may be NULL) to host memory. */
RECON_API int recon_batch_download(recon_batch* b, int32_t i, float* host_d, float* host_w,
                                   recon_stream_t stream);
/* sync.  Copies volume i's mesh (9 floats per triangle) to host memory;
 * *num_triangles <- its triangle count; RECON_ERR_CAPACITY if it exceeds
 * capacity_triangles (nothing is copied then). */
RECON_API int recon_batch_download_mesh(recon_batch* b, int32_t i, float* host_verts,
                                        int64_t capacity_triangles, int64_t* num_triangles,
                                        recon_stream_t stream);
/* sync.  Tunes the tile shape of each (grid, image size) class of `volumes`:
 * starting from the variants in cache_path (or the default), each class in
 * turn tries every variant, timing the whole batch's integrate graph
 * (`iterations` runs), and keeps a variant that makes it >= 2 % faster.  The
 * choices are merged into the JSON file cache_path. */
RECON_API int recon_batch_autotune(const recon_batch_volume* volumes, int32_t num_volumes,
                                   const char* cache_path, int32_t iterations,
                                   recon_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* RECON_RECON_H */
