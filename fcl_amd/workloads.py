"""Seeded synthetic workloads for the mesh hot path.

- mesh_collide_pair(): two perturbed icospheres (k = 11, 2,420 triangles
  each, radii 1 +- 15 %), poses with a uniform random rotation and a
  translation uniform in [-2.2, 2.2]^3; CollisionRequest(1) -- the plain
  mesh-vs-mesh collide case.
- mesh_distance_pair(): icospheres with k = 32 (20,480 triangles), BVHModel
  <RSS>, translation |t| ~ U[2.1, 4] along a uniform direction.
- uneven_batch(): 64 mesh pairs, each mesh a perturbed icosphere or a
  heightfield terrain (p = 1/2), triangle counts log-uniform in
  [500, 200,000] (realised as 20k^2 or 2(g-1)^2), every mesh scaled to unit
  bounding radius; 1,000,000 queries split over the pairs by a Dirichlet(1)
  draw (normalised Exp(1), largest-remainder rounding); translation uniform
  in [-h, h]^3 with h = 1.1 * (R_A + R_B) = 2.2; exhaustive contacts
  (num_max_contacts = 100,000).
- primitive_batch(): 4,000,000 primitive pairs, every shape drawn uniformly
  from {Box, Sphere, Capsule, Cylinder, Convex}; box sides, radii and
  lengths log-uniform in [0.05, 1]; a convex is 32 points from N(0, I)
  scaled by a log-uniform factor in [0.05, 1] (its GJK support -- a scan of
  the vertex list -- equals that of the points' hull, since points strictly
  inside the hull never attain the maximum); both poses per pair with a
  uniform random rotation (normalised Gaussian quaternion) and a translation
  uniform in [-1, 1]^3; collide with enable_contact, GST_INDEP, FP64.
"""
import numpy as np

from . import CONVEX, SHAPE_DTYPE, heightfield, icosphere, poses, scale_to_radius


def synth_mesh(kind, n_tris, seed):
    """kind 0: perturbed icosphere, 1: heightfield; about n_tris triangles,
    scaled to unit bounding radius."""
    if kind == 0:
        k = max(1, int(round(np.sqrt(n_tris / 20.0))))
        V, T = icosphere(k, 0.15, seed)
    else:
        g = max(2, int(round(np.sqrt(n_tris / 2.0))) + 1)
        V, T = heightfield(g, 0.05, seed)
    return scale_to_radius(V, 1.0), T


def mesh_collide_pair(n_queries=10000):
    meshes = [icosphere(11, 0.15, 0xA0), icosphere(11, 0.15, 0xB0)]
    tf1, tf2 = poses(n_queries, 0, 2.2, 0.0, 0xC0)
    qm = np.tile(np.array([0, 1], np.int32), (n_queries, 1))
    return {"meshes": meshes, "query_meshes": qm, "tf1": tf1, "tf2": tf2, "cap": 1}


def mesh_distance_pair(n_queries=10000):
    meshes = [icosphere(32, 0.15, 0xD0), icosphere(32, 0.15, 0xD1)]
    tf1, tf2 = poses(n_queries, 1, 2.1, 4.0, 0xD2)
    qm = np.tile(np.array([0, 1], np.int32), (n_queries, 1))
    return {"meshes": meshes, "query_meshes": qm, "tf1": tf1, "tf2": tf2}


def dirichlet_split(total, n, rng):
    w = rng.exponential(1.0, n)
    w /= w.sum()
    raw = w * total
    counts = np.floor(raw).astype(np.int64)
    rest = total - counts.sum()
    order = np.argsort(-(raw - counts), kind="stable")
    counts[order[:rest]] += 1
    return counts


def uneven_batch(n_queries=1_000_000, n_pairs=64, seed=0x5EED, lo=500, hi=200_000,
                 mesh_builder=None):
    """Returns dict with 'meshes' [(V,T)], 'pairs' (n_pairs,2) mesh indices,
    'query_meshes' (n,2), 'tf1', 'tf2', 'cap', 'pair_counts'."""
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, 2, 2 * n_pairs)
    tris = np.exp(rng.uniform(np.log(lo), np.log(hi), 2 * n_pairs))
    seeds = rng.integers(1, 2**62, 2 * n_pairs)
    build = mesh_builder or (lambda i: synth_mesh(int(kinds[i]), float(tris[i]), int(seeds[i])))
    meshes = [build(i) for i in range(2 * n_pairs)]
    pairs = np.arange(2 * n_pairs, dtype=np.int32).reshape(n_pairs, 2)
    counts = dirichlet_split(n_queries, n_pairs, rng)
    qm = np.repeat(pairs, counts, axis=0).astype(np.int32)
    tf1, tf2 = poses(n_queries, 0, 2.2, 0.0, int(rng.integers(1, 2**62)))
    return {"meshes": meshes, "pairs": pairs, "query_meshes": qm, "tf1": tf1, "tf2": tf2,
            "cap": 100000, "pair_counts": counts, "kinds": kinds, "tris_target": tris}


def random_poses(n, rng, half_extent=1.0):
    """(n, 12) poses: rotation from a normalised Gaussian quaternion
    (row-major), translation uniform in [-half_extent, half_extent]^3."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    tf = np.empty((n, 12))
    tf[:, 0] = 1 - 2 * (y * y + z * z)
    tf[:, 1] = 2 * (x * y - w * z)
    tf[:, 2] = 2 * (x * z + w * y)
    tf[:, 3] = 2 * (x * y + w * z)
    tf[:, 4] = 1 - 2 * (x * x + z * z)
    tf[:, 5] = 2 * (y * z - w * x)
    tf[:, 6] = 2 * (x * z - w * y)
    tf[:, 7] = 2 * (y * z + w * x)
    tf[:, 8] = 1 - 2 * (x * x + y * y)
    tf[:, 9:] = rng.uniform(-half_extent, half_extent, (n, 3))
    return tf


def primitive_batch(n_pairs=4_000_000, seed=0x5A9E, lo=0.05, hi=1.0, convex_points=32):
    """Returns dict with 'shapes' (2n fclx_shape records; pair q uses shapes
    2q and 2q+1), 'vertices' (V,3) convex pool, 'pair_shapes' (n,2), 'tf1',
    'tf2', 'cap' (1) and 'enable_contact' (True)."""
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    types = rng.integers(0, 5, n).astype(np.int32)
    dims = np.exp(rng.uniform(np.log(lo), np.log(hi), (n, 3)))
    rec = np.zeros(n, SHAPE_DTYPE)
    rec["type"] = types
    # unused dimensions are zero: sphere keeps the radius, capsule and
    # cylinder (radius, lz), a convex none (its scale is baked into the points)
    dims[types == 1, 1:] = 0
    dims[(types == 2) | (types == 3), 2] = 0
    conv = np.flatnonzero(types == CONVEX)
    scale = dims[conv, 0].copy()
    dims[conv] = 0
    rec["dims"] = dims
    rec["num_vertices"][conv] = convex_points
    rec["vertex_offset"][conv] = np.arange(len(conv), dtype=np.int64) * convex_points
    pts = rng.standard_normal((len(conv), convex_points, 3))
    pts *= scale[:, None, None]
    tf1 = random_poses(n_pairs, rng)
    tf2 = random_poses(n_pairs, rng)
    return {"shapes": rec, "vertices": pts.reshape(-1, 3), "tf1": tf1, "tf2": tf2,
            "pair_shapes": np.arange(n, dtype=np.int32).reshape(n_pairs, 2), "cap": 1,
            "enable_contact": True}
