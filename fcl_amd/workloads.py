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


# ------------------------------------------------------------------ serving
def box_mesh(center, size, yaw):
    """A box as 8 vertices and 12 triangles (outward), turned by yaw about z."""
    c, s = np.cos(yaw), np.sin(yaw)
    sx, sy, sz = np.asarray(size, float) / 2
    corners = np.array([[x, y, z] for x in (-sx, sx) for y in (-sy, sy) for z in (-sz, sz)])
    R = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    V = corners @ R.T + np.asarray(center, float)
    T = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                  [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.int32)
    return V, T


def serving_scene(grid=400, n_boxes=15133, extent=4.0, seed=0x5CE7E):
    """One mesh: a grid x grid heightfield over [-extent, extent]^2 (heights
    U(-0.04, 0.04) * extent) and n_boxes boxes (sides U[0.05, 0.25], on the
    ground plane, random yaw).  grid 400 and 15,133 boxes give 318,402 +
    181,596 = 499,998 triangles."""
    rng = np.random.default_rng(seed)
    V, T = heightfield(grid, 0.02, int(rng.integers(1, 2**62)))
    V = V * extent
    Vs, Ts, off = [V], [T], len(V)
    sizes = rng.uniform(0.05, 0.25, (n_boxes, 3))
    xy = rng.uniform(-0.95 * extent, 0.95 * extent, (n_boxes, 2))
    yaw = rng.uniform(0, 2 * np.pi, n_boxes)
    for i in range(n_boxes):
        bv, bt = box_mesh((xy[i, 0], xy[i, 1], sizes[i, 2] / 2), sizes[i], yaw[i])
        Vs.append(bv)
        Ts.append(bt + off)
        off += 8
    return np.ascontiguousarray(np.concatenate(Vs)), np.ascontiguousarray(np.concatenate(Ts))


def capsule_mesh(radius, length, n_tris, noise=0.1, seed=0):
    """A capsule along z (segment `length`, radius `radius`) tessellated into
    about n_tris triangles: rings along the stadium profile between two pole
    vertices; every vertex's distance from the axis is scaled by
    1 + U(-noise, noise)."""
    rng = np.random.default_rng(seed)
    m = max(8, int(round(np.sqrt(n_tris))))          # vertices around
    rings = max(3, int(round(n_tris / (2.0 * m))))   # interior rings
    half_arc = np.pi * radius / 2
    total = 2 * half_arc + length
    s = np.arange(1, rings + 1) / (rings + 1) * total
    rho = np.empty(rings)
    z = np.empty(rings)
    for i, si in enumerate(s):
        if si < half_arc:
            a = si / radius
            rho[i], z[i] = radius * np.sin(a), -length / 2 - radius * np.cos(a)
        elif si <= half_arc + length:
            rho[i], z[i] = radius, -length / 2 + (si - half_arc)
        else:
            a = (si - half_arc - length) / radius
            rho[i], z[i] = radius * np.cos(a), length / 2 + radius * np.sin(a)
    th = np.arange(m) * (2 * np.pi / m)
    scale = 1 + rng.uniform(-noise, noise, (rings, m))
    X = (rho[:, None] * scale) * np.cos(th)[None, :]
    Y = (rho[:, None] * scale) * np.sin(th)[None, :]
    Z = np.repeat(z[:, None], m, 1)
    ring_v = np.stack([X, Y, Z], -1).reshape(-1, 3)
    V = np.concatenate([[[0, 0, -length / 2 - radius]], ring_v, [[0, 0, length / 2 + radius]]])
    top = len(V) - 1
    tris = []
    j = np.arange(m)
    jn = (j + 1) % m
    tris.append(np.stack([np.zeros(m, int), 1 + jn, 1 + j], 1))       # bottom fan
    for r in range(rings - 1):
        a0, a1 = 1 + r * m, 1 + (r + 1) * m
        tris.append(np.stack([a0 + j, a0 + jn, a1 + jn], 1))
        tris.append(np.stack([a0 + j, a1 + jn, a1 + j], 1))
    last = 1 + (rings - 1) * m
    tris.append(np.stack([last + j, last + jn, np.full(m, top)], 1))  # top fan
    return np.ascontiguousarray(V, np.float64), np.ascontiguousarray(np.concatenate(tris), np.int32)


def serving_robot(n_links=20, lo=1000, hi=50000, seed=0x40B07):
    """20 capsule-like link meshes, triangle counts log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    links = []
    for i in range(n_links):
        nt = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        links.append(capsule_mesh(rng.uniform(0.03, 0.08), rng.uniform(0.15, 0.4), nt, 0.1,
                                  int(rng.integers(1, 2**62))))
    return links


def _small_rotations(rng, n, sigma):
    """n rotation matrices about uniform random axes by angles N(0, sigma)."""
    ax = rng.standard_normal((n, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ang = rng.normal(0, sigma, n)
    K = np.zeros((n, 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -ax[:, 2], ax[:, 1], -ax[:, 0]
    K[:, 1, 0], K[:, 2, 0], K[:, 2, 1] = ax[:, 2], -ax[:, 1], ax[:, 0]
    s, c = np.sin(ang)[:, None, None], np.cos(ang)[:, None, None]
    return np.eye(3)[None] + s * K + (1 - c) * (K @ K)


def serving_request(n_links, n_waypoints, rng, extent=4.0, sigma_t=0.01, sigma_r=0.02):
    """Link poses (n_waypoints, n_links, 12), waypoint-major: each link
    starts near a common base (x, y in [-0.85, 0.85] * extent, z in
    [0.05, 0.45]) with a uniform random orientation and moves by a random
    walk: translation steps N(0, sigma_t) per axis, rotation steps about a
    random axis by N(0, sigma_r) rad."""
    base = np.array([*rng.uniform(-0.85 * extent, 0.85 * extent, 2), rng.uniform(0.05, 0.45)])
    t = base + rng.uniform(-0.4, 0.4, (n_links, 3))
    t[:, 2] = np.clip(t[:, 2], -0.05, None)
    q = rng.standard_normal((n_links, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
                 1).reshape(n_links, 3, 3)
    out = np.empty((n_waypoints, n_links, 12))
    steps_t = rng.normal(0, sigma_t, (n_waypoints, n_links, 3))
    for k in range(n_waypoints):
        if k:
            t = t + steps_t[k]
            R = _small_rotations(rng, n_links, sigma_r) @ R
        out[k, :, :9] = R.reshape(n_links, 9)
        out[k, :, 9:] = t
    return out


def serving_requests(n, n_links=20, lo=64, hi=4096, seed=0x5E4E):
    """n requests: waypoint counts log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    lens = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(int)
    return [np.ascontiguousarray(serving_request(n_links, int(L), rng)) for L in lens]
