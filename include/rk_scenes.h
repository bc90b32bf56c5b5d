/*
 * rk_scenes.h — seeded synthetic scenes and rays (host code, no GPU).
 * Inputs for tests and benchmarks; every generator is deterministic for a
 * given seed (SplitMix64, counter-based, so generation order does not matter).
 * Call each generator once with NULL outputs to get the counts.
 */
#ifndef RK_SCENES_H
#define RK_SCENES_H

#include "rk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Icosahedron subdivided 'levels' times by edge midpoints projected to the unit
 * sphere: 20*4^levels faces, 10*4^levels+2 vertices (w = 0). */
rk_status rk_scene_icosphere(uint32_t levels, rk_float4* vertices, rk_uint4* faces, uint32_t* vertex_count,
                             uint32_t* face_count);

/* grid x grid vertices over [0,1]^2 with z = amplitude * fBm normalised to [0,1]
 * (6 octaves of hashed value noise, base frequency 4, lacunarity 2, gain 0.5),
 * two faces per cell: (v00,v10,v11), (v00,v11,v01). */
rk_status rk_scene_heightfield(uint32_t grid, uint64_t seed, float amplitude, rk_float4* vertices, rk_uint4* faces,
                               uint32_t* vertex_count, uint32_t* face_count);

/* n independent triangles: centroid c ~ U[lo,hi]^3, edges e1, e2 isotropic with
 * lognormal(mu, sigma) length; v0 = c - (e1+e2)/3, v1 = v0+e1, v2 = v0+e2. */
rk_status rk_scene_soup(uint32_t n, uint64_t seed, float mu, float sigma, float lo, float hi, rk_float4* vertices,
                        rk_uint4* faces, uint32_t* vertex_count, uint32_t* face_count);

/* The mesh hitbox as CUDAModel accumulates it (CUDAModel.h:178-183): the
 * merge of every vertex as a float4 with w = 0, starting from an empty box. */
rk_status rk_scene_hitbox(const rk_float4* vertices, uint32_t vertex_count, rk_aabox* hitbox);

/* The camera of CUDACamera (CUDACamera.cpp:56-88, y up) after setCenter(eye) and
 * lookAt(target), and the per-frame constants of CUDA_calcPrimaryRays
 * (CUDA_calcPrimaryRays.cu:43-55) for a width x height surface.
 * out[0] = ray origin, out[1] = ray dir at pixel (0,0), out[2] = step per x,
 * out[3] = step per y. fov in radians, ratio = width/height as the caller sets it. */
rk_status rk_camera_lookat(rk_float4 eye, rk_float4 target, float fov, float ratio, int width, int height,
                           rk_float4 out[4]);
/* The same constants from an explicit up vector (for z-up scenes). */
rk_status rk_camera_lookat_up(rk_float4 eye, rk_float4 target, rk_float4 up, float fov, float ratio, int width,
                              int height, rk_float4 out[4]);
/* Host evaluation of the primary rays: dir = (ray_dir - x*ray_x) - y*ray_y. */
rk_status rk_primary_rays_host(const rk_float4 cam[4], int width, int height, rk_float4* origins,
                               rk_float4* directions);

/* n rays, origins ~ U[lo,hi]^3, directions uniform on the sphere (w = 0). */
rk_status rk_rays_random(uint32_t n, uint64_t seed, float lo, float hi, rk_float4* origins, rk_float4* directions);

#ifdef __cplusplus
}
#endif

#endif /* RK_SCENES_H */
