/*
 * scene.h -- synthetic depth frames for the TSDF fusion path.
 *
 * Ray-casts z-depth maps (depth = camera z of the first hit, the quantity
 * pixel_to_camera expects, cu-Reconstruction modules/src/cuda/camera_util.cu:
 * 21-31) of scenes made of spheres, axis-aligned boxes and horizontal planes,
 * then adds seeded Gaussian noise and dropout.  All randomness is
 * counter-based (splitmix64 of seed, frame, pixel), so every frame is
 * bit-identical for the same arguments whatever the thread count.
 * Host-only C++; the workload generator of the tests and the benchmark.
 */
#ifndef RECON_SCENE_H
#define RECON_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { RECON_PRIM_SPHERE = 0, RECON_PRIM_BOX = 1, RECON_PRIM_PLANE_Z = 2 };

typedef struct recon_scene_prim {
    int32_t kind;
    int32_t pad_;
    double p[6]; /* sphere: cx,cy,cz,r; box: min xyz, max xyz; plane: z */
} recon_scene_prim;

typedef struct recon_scene_noise {
    double sigma0;    /* depth noise sigma = sigma0 + sigma_z2 * z^2 (metres) */
    double sigma_z2;
    double dropout;   /* probability that a pixel is set to 0 */
    uint64_t seed;
} recon_scene_noise;

/* Renders `num_frames` frames; frame f uses pose poses_cw[7f..7f+6]
 * ({qw,qx,qy,qz,tx,ty,tz}, world->camera) and noise stream index
 * frame_ids[f] (NULL: f).  out: num_frames * rows * cols floats, metres,
 * 0 for misses and dropped pixels.  Pixel (u,v) looks along camera direction
 * ((u - cx)/fx, (v - cy)/fy, 1).  noise may be NULL (exact depth). */
int recon_scene_render(const recon_scene_prim* prims, int32_t num_prims,
                       const double* poses_cw, const uint64_t* frame_ids, int32_t num_frames,
                       double fx, double fy, double cx, double cy, int32_t cols, int32_t rows,
                       const recon_scene_noise* noise, float* out, int32_t nthreads);

/* Exact signed distance of world point p to the union of the primitives
 * (negative inside), for checking extracted meshes. */
double recon_scene_sdf(const recon_scene_prim* prims, int32_t num_prims, const double p[3]);

#ifdef __cplusplus
}
#endif

#endif /* RECON_SCENE_H */
