/*
 * fclx_synth.h — seeded synthetic workloads for the mesh hot path: perturbed
 * geodesic icospheres, random-displacement heightfield terrains, and query
 * poses.  Host-only; no counterpart in the reference (its tests load OBJ
 * files and draw poses from an unseeded rand(), test_fcl_utility.h:304-373).
 *
 * Random numbers: splitmix64 seeds a xoshiro256** stream; uniforms are
 * (x >> 11) * 2^-53; normals come from Box-Muller.  A rotation is the matrix
 * of a normalised 4-vector of standard normals (a uniform random rotation).
 */
#ifndef FCLX_SYNTH_H
#define FCLX_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Frequency-k geodesic icosahedron: 10k^2+2 vertices, 20k^2 triangles. */
void fclx_synth_icosphere_size(int32_t k, int32_t* num_vertices, int32_t* num_tris);
/* Unit sphere with each vertex radius scaled by 1 + U(-noise, noise). */
int fclx_synth_icosphere(int32_t k, double noise, uint64_t seed, double* vertices,
                         int32_t* tris);

/* g x g grid over [-1,1]^2, heights U(-h, h) with h = height_frac * 2. */
void fclx_synth_heightfield_size(int32_t g, int32_t* num_vertices, int32_t* num_tris);
int fclx_synth_heightfield(int32_t g, double height_frac, uint64_t seed, double* vertices,
                           int32_t* tris);

/* Scale vertices about the origin so the farthest one is at distance r. */
void fclx_synth_scale_to_radius(double* vertices, int32_t num_vertices, double r);

/* Poses (12 doubles each).  tf1 is the identity; tf2 has a uniform random
 * rotation and a translation:
 *   kind 0: uniform in [-a, a]^3                       (p0 = a)
 *   kind 1: s * u, u uniform on the unit sphere, s ~ U[p0, p1]. */
int fclx_synth_poses(int32_t n, int32_t kind, double p0, double p1, uint64_t seed,
                     double* tf1, double* tf2);

/* Uniform doubles in [0,1) from the stream seeded with seed. */
void fclx_synth_uniform(int64_t n, uint64_t seed, double* out);

#ifdef __cplusplus
}
#endif

#endif
