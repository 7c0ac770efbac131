"""Shared test helpers: OBJ loading, pose generators in the style of FCL's
tests, and independent brute-force geometry used to pin the oracle."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_obj(name):
    """Vertices and fan-triangulated faces of an OBJ file, like
    test::loadOBJFile (test/test_fcl_utility.h:193-278)."""
    V, T = [], []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if tok[0] == "v":
                V.append([float(tok[1]), float(tok[2]), float(tok[3])])
            elif tok[0] == "f":
                idx = [int(t.split("/")[0]) - 1 for t in tok[1:]]
                for t in range(len(idx) - 2):
                    T.append([idx[0], idx[t + 1], idx[t + 2]])
    return np.array(V, np.float64), np.array(T, np.int32)


def euler_to_matrix(a, b, c):
    """test_fcl_utility.h:312-326 (eulerToMatrix)."""
    c1, c2, c3 = np.cos(a), np.cos(b), np.cos(c)
    s1, s2, s3 = np.sin(a), np.sin(b), np.sin(c)
    return np.array([[c1 * c2, -c2 * s1, s2],
                     [c3 * s1 + c1 * s2 * s3, c1 * c3 - s1 * s2 * s3, -c2 * s3],
                     [s1 * s3 - c1 * c3 * s2, c3 * s1 * s2 + c1 * s3, c2 * c3]])


def random_transforms(n, extents, seed):
    """Poses like test::generateRandomTransforms (test_fcl_utility.h:348-373)
    from an explicit seed: translation uniform in the extents box, Euler
    angles uniform in [0, 2pi)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 12))
    for i in range(n):
        x = rng.uniform(extents[0], extents[3])
        y = rng.uniform(extents[1], extents[4])
        z = rng.uniform(extents[2], extents[5])
        a, b, c = rng.uniform(0, 2 * np.pi, 3)
        out[i, :9] = euler_to_matrix(a, b, c).reshape(9)
        out[i, 9:] = [x, y, z]
    return out


def identity_pose():
    t = np.zeros(12)
    t[[0, 4, 8]] = 1
    return t


def pose_apply(tf, P):
    R = np.asarray(tf[:9]).reshape(3, 3)
    return P @ R.T + np.asarray(tf[9:])


# --------------------------------------------------------------- brute force
def seg_seg_dist(p, a, q, b):
    """Distance between segments p+s*a and q+t*b (s,t in [0,1]) by checking
    the interior critical point and all four clamped edges."""
    best = np.inf
    cands = []
    A = np.array([[a @ a, -(a @ b)], [-(a @ b), b @ b]])
    rhs = np.array([(q - p) @ a, -((q - p) @ b)])
    if abs(np.linalg.det(A)) > 1e-300:
        s, t = np.linalg.solve(A, rhs)
        if 0 <= s <= 1 and 0 <= t <= 1:
            cands.append((s, t))
    for s in (0.0, 1.0):
        pt = p + s * a
        t = 0.0 if b @ b == 0 else np.clip((pt - q) @ b / (b @ b), 0, 1)
        cands.append((s, t))
    for t in (0.0, 1.0):
        qt = q + t * b
        s = 0.0 if a @ a == 0 else np.clip((qt - p) @ a / (a @ a), 0, 1)
        cands.append((s, t))
    for s, t in cands:
        best = min(best, np.linalg.norm(p + s * a - (q + t * b)))
    return best


def point_tri_dist(x, A, B, C):
    """Distance from x to triangle ABC (closed) via projection + edges."""
    n = np.cross(B - A, C - A)
    nn = n @ n
    if nn > 0:
        d = (x - A) @ n / nn
        y = x - d * n
        # barycentric inside test
        c0 = np.cross(B - A, y - A) @ n
        c1 = np.cross(C - B, y - B) @ n
        c2 = np.cross(A - C, y - C) @ n
        if c0 >= 0 and c1 >= 0 and c2 >= 0:
            return abs(d) * np.sqrt(nn)
    return min(seg_seg_dist(x, np.zeros(3), A, B - A), seg_seg_dist(x, np.zeros(3), B, C - B),
               seg_seg_dist(x, np.zeros(3), C, A - C))


def seg_crosses_tri(p, q, A, B, C, eps):
    """Segment pq crosses the interior of triangle ABC with every orientation
    value at least eps away from zero (an unambiguous crossing)."""
    def orient(a, b, c, d):
        return np.dot(np.cross(b - a, c - a), d - a)
    s1, s2 = orient(A, B, C, p), orient(A, B, C, q)
    if not (s1 * s2 < 0 and abs(s1) > eps and abs(s2) > eps):
        return False
    o = [orient(p, q, A, B), orient(p, q, B, C), orient(p, q, C, A)]
    if min(abs(v) for v in o) <= eps:
        return False
    return all(v > 0 for v in o) or all(v < 0 for v in o)


def _edge_plane_points(A, B):
    """Points where the edges of triangle A cross the plane of triangle B."""
    n = np.cross(B[1] - B[0], B[2] - B[0])
    out = []
    for i in range(3):
        p, q = A[i], A[(i + 1) % 3]
        sp, sq = (p - B[0]) @ n, (q - B[0]) @ n
        if sp * sq < 0:
            out.append(p + (q - p) * (sp / (sp - sq)))
    return out


def tri_tri_dist(S, T):
    """Brute-force distance between closed triangles: min over the 9 edge
    pairs, the 6 vertex-face pairs, and the edge-plane crossing points of
    each triangle against the other (0 when an edge pierces a face)."""
    best = np.inf
    for i in range(3):
        for j in range(3):
            best = min(best, seg_seg_dist(S[i], S[(i + 1) % 3] - S[i], T[j],
                                          T[(j + 1) % 3] - T[j]))
    for i in range(3):
        best = min(best, point_tri_dist(S[i], *T), point_tri_dist(T[i], *S))
    for x in _edge_plane_points(S, T):
        best = min(best, point_tri_dist(x, *T))
    for x in _edge_plane_points(T, S):
        best = min(best, point_tri_dist(x, *S))
    return best


def rect_corners(axis, to, l):
    """The four corners of the RSS rectangle To + axis[:, :2] @ [s*l0, t*l1]."""
    ax = np.asarray(axis).reshape(3, 3)
    c = []
    for s, t in ((0, 0), (1, 0), (1, 1), (0, 1)):
        c.append(np.asarray(to) + ax[:, 0] * s * l[0] + ax[:, 1] * t * l[1])
    return c


def rect_rect_dist(c1, c2):
    """Brute-force distance between two planar rectangles given by corners:
    they are unions of two triangles each."""
    t1 = [np.array([c1[0], c1[1], c1[2]]), np.array([c1[0], c1[2], c1[3]])]
    t2 = [np.array([c2[0], c2[1], c2[2]]), np.array([c2[0], c2[2], c2[3]])]
    return min(tri_tri_dist(a, b) for a in t1 for b in t2)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def random_transform_pairs(n, extents, delta_trans, delta_rot, seed):
    """Pose pairs like test::generateRandomTransforms with deltas
    (test_fcl_utility.h:375-417) from an explicit seed: the first pose as
    random_transforms draws it, the second with translation offsets uniform
    in [-delta_trans, delta_trans] and Euler angles offset uniformly in
    [-delta_rot, delta_rot]."""
    rng = np.random.default_rng(seed)
    a_out = np.zeros((n, 12))
    b_out = np.zeros((n, 12))
    for i in range(n):
        t = np.array([rng.uniform(extents[k], extents[k + 3]) for k in range(3)])
        ang = rng.uniform(0, 2 * np.pi, 3)
        dt = rng.uniform(-np.asarray(delta_trans, float), np.asarray(delta_trans, float))
        da = rng.uniform(-delta_rot, delta_rot, 3)
        a_out[i, :9] = euler_to_matrix(*ang).reshape(9)
        a_out[i, 9:] = t
        b_out[i, :9] = euler_to_matrix(*(ang + da)).reshape(9)
        b_out[i, 9:] = t + dt
    return a_out, b_out
