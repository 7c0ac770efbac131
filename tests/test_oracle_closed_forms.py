"""Pins the oracle's sphere-cylinder and sphere-box closed forms
(sphereCylinderIntersect, sphereBoxIntersect, reached through the
GST_INDEP solver's shapeIntersect) against the known answers of the
reference's own primitive tests:

- test/narrowphase/detail/primitive_shape_algorithm/test_sphere_cylinder.cpp
  :222-660 (GetUniformConfigurations, GetNonUniformConfigurations), run as
  its CollisionAcrossVaryingWorldFrames, CollisionWithSphereRotations and
  CollisionIncompatibleScales tests (:833-856) do: each configuration in the
  cylinder frame, posed in seven world frames (:781-823), with the sphere
  unrotated and in five orientations (:829-852); expected collision flag,
  penetration depth, normal (from the sphere into the cylinder) and
  contact position at the tolerance 22 eps (:66-69).
- test_sphere_box.cpp:207-487 the same way (tolerance 18 eps, :58-61).

The configurations are restated here as data (positions, radii, expected
depth / normal / position); the oracle under test is oracle/fclo_shape.c.
"""
import numpy as np
import pytest

import fclo

EPS = np.finfo(np.float64).eps


def unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def angle_axis(angle, axis):
    """Rotation matrix of an angle about a unit axis (Rodrigues)."""
    k = unit(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def pose(R, t):
    out = np.zeros(12)
    out[:9] = np.asarray(R).reshape(9)
    out[9:] = t
    return out


def world_frames():
    """The seven frames of QueryWithVaryingWorldFrames: identity, a
    translation, then that translation with a 90-degree turn about each
    axis, an arbitrary rotation and a near-identity one."""
    t = np.array([1.3, 2.7, 6.5])
    frames = [(np.eye(3), np.zeros(3)), (np.eye(3), t)]
    for a in range(3):
        frames.append((angle_axis(np.pi / 2, np.eye(3)[a]), t))
    frames.append((angle_axis(np.pi / 3, [1, 2, 3]), t))
    frames.append((angle_axis(np.sqrt(EPS), [1, 0, 0]), t))
    return frames


def sphere_rotations():
    rots = [np.eye(3)]
    rots += [angle_axis(np.pi / 2, np.eye(3)[a]) for a in range(3)]
    rots += [angle_axis(np.pi / 3, [1, 2, 3]), angle_axis(np.sqrt(EPS), [1, 0, 0])]
    return rots


BARREL = [unit(v) for v in ([1, 0, 0], [0, 1, 0], [1, 0.5, 0], [-1, 0.5, 0], [-1, -0.5, 0],
                            [1, -0.5, 0])]
Z = np.array([0.0, 0, 1])


def cylinder_uniform():
    """(name, r_c, h_c, r_s, p_CS, colliding, depth, normal, pos)."""
    h_c, r_c, r_s = 0.6, 1.2, 0.7
    hh = h_c / 2
    out = []
    p = np.array([r_c * 0.25, r_c * 0.25, hh + r_s * 1.1])
    out.append(("separated, +z face", r_c, h_c, r_s, p, False, None, None, None))
    n = unit([-1, -1, 0])
    p = np.array([0, 0, hh * 0.5]) - n * (r_s + r_c + 0.1)
    out.append(("separated, barrel", r_c, h_c, r_s, p, False, None, None, None))
    n = unit([-1, -1, -1])
    cs = np.array([0, 0, hh]) + unit([-n[0], -n[1], 0]) * r_c
    out.append(("separated, edge", r_c, h_c, r_s, cs - n * (r_s + 0.1), False, None, None, None))
    d = min(r_c, r_s) * 0.25
    for sign in (-1.0, 1.0):
        n = -sign * Z
        cs = np.array([r_c * 0.25, r_c * 0.25, hh * sign])
        out.append(("outside centre, cap %+d" % sign, r_c, h_c, r_s, cs - n * (r_s - d), True,
                    d, n, cs + n * (d / 2)))
    for ncs in BARREL:
        cs = np.array([0, 0, hh * 0.1]) + ncs * r_c
        out.append(("outside centre, barrel", r_c, h_c, r_s, cs + ncs * (r_s - d), True, d,
                    -ncs, cs - ncs * (d / 2)))
    for sign in (-1.0, 1.0):
        for nxy in BARREL:
            cs = np.array([0, 0, sign * hh]) + nxy * r_c
            ncs = unit(cs)
            out.append(("outside centre, edge", r_c, h_c, r_s, cs + ncs * (r_s - d), True, d,
                        -ncs, cs - ncs * (d / 2)))
    for sign in (-1.0, 1.0):
        n = -sign * Z
        cs = np.array([r_c * 0.25, r_c * 0.25, hh * sign])
        dep = 0.1 + r_s
        out.append(("inside centre, cap %+d" % sign, r_c, h_c, r_s, cs + n * 0.1, True, dep, n,
                    cs + n * (dep / 2)))
    for ncs in BARREL:
        cs = np.array([0, 0, hh * 0.1]) + ncs * r_c
        dep = r_s + 0.1
        out.append(("inside centre, barrel", r_c, h_c, r_s, cs - ncs * 0.1, True, dep, -ncs,
                    cs - ncs * (dep / 2)))
    tiny = 22 * EPS / 2
    for sign in (-1.0, 1.0):
        n = -sign * Z
        cs = np.array([r_c * 0.25, r_c * 0.25, hh * sign])
        out.append(("eps outside, cap", r_c, h_c, r_s, cs - n * tiny, True, r_s, n,
                    cs + n * (r_s / 2)))
    for ncs in BARREL:
        cs = np.array([0, 0, hh * 0.1]) + ncs * r_c
        out.append(("eps outside, barrel", r_c, h_c, r_s, cs - ncs * tiny, True, r_s, -ncs,
                    cs - ncs * (r_s / 2)))
    for sign in (-1.0, 1.0):
        for nxy in BARREL:
            cs = np.array([0, 0, sign * hh]) + nxy * r_c
            ncs = unit(cs)
            n = -sign * Z  # epsilon outside the edge counts as inside: the face wins
            out.append(("eps outside, edge", r_c, h_c, r_s, cs + ncs * tiny, True, r_s, n,
                        cs + n * (r_s / 2)))
    big_r = h_c * 2
    dep = r_s + h_c / 2
    out.append(("origin, face nearest", big_r, h_c, r_s, np.zeros(3), True, dep, -Z,
                np.array([0, 0, h_c / 2 - dep / 2])))
    big_h = r_c * 4
    dep = r_s + r_c
    out.append(("origin, barrel nearest", r_c, big_h, r_s, np.zeros(3), True, dep,
                np.array([-1.0, 0, 0]), np.array([r_c - dep / 2, 0, 0])))
    return out


def cylinder_nonuniform():
    out = []
    r_c, h_c, r_s = 9.0, 0.1, 0.025
    d = r_s / 2
    cs = unit([1, 2, 0]) * (r_c - r_s) + Z * (h_c / 2)
    out.append(("large disk, small sphere, face", r_c, h_c, r_s, cs + Z * (r_s - d), True, d,
                -Z, cs - Z * (d / 2)))
    out.append(("large disk, small sphere, face, apart", r_c, h_c, r_s, cs + Z * (r_s + d),
                False, None, None, None))
    ncs = unit([1, 2, 0])
    cs = ncs * r_c + Z * (r_s * 0.1)
    out.append(("large disk, small sphere, barrel", r_c, h_c, r_s, cs + ncs * (r_s - d), True,
                d, -ncs, cs - ncs * (d / 2)))
    out.append(("large disk, small sphere, barrel, apart", r_c, h_c, r_s, cs + ncs * (r_s + d),
                False, None, None, None))
    r_c, h_c, r_s = 0.025, 0.1, 9.0
    d = r_c / 2
    cs = unit([1, 2, 0]) * (r_c * 0.5) + Z * (h_c / 2)
    out.append(("large sphere, small disk, face", r_c, h_c, r_s, cs + Z * (r_s - d), True, d,
                -Z, cs - Z * (d / 2)))
    out.append(("large sphere, small disk, face, apart", r_c, h_c, r_s, cs + Z * (r_s + d),
                False, None, None, None))
    ncs = unit([1, 2, 0])
    cs = ncs * r_c + Z * (h_c * 0.1)
    out.append(("large sphere, small disk, barrel", r_c, h_c, r_s, cs + ncs * (r_s - d), True,
                d, -ncs, cs - ncs * (d / 2)))
    out.append(("large sphere, small disk, barrel, apart", r_c, h_c, r_s,
                cs + ncs * (r_s + d), False, None, None, None))
    return out


def _eval(sphere, other, X_W, R_S, cfg, eps):
    name, p_S, colliding, depth, normal, pos = cfg
    Rw, tw = X_W
    X_WO = pose(Rw, tw)
    X_WS = pose(Rw @ R_S, Rw @ p_S + tw)
    hit, _ = fclo.shape_intersect(sphere, X_WS, other, X_WO, contacts=False)
    assert hit == colliding, name
    hit, got = fclo.shape_intersect(sphere, X_WS, other, X_WO)
    assert hit == colliding, name
    if not colliding:
        assert got == [], name
        return
    assert len(got) == 1, name
    n, p, d = got[0]
    assert abs(d - depth) <= eps, (name, d, depth)
    assert np.max(np.abs(n - Rw @ normal)) <= eps, (name, n, Rw @ normal)
    assert np.max(np.abs(p - (Rw @ pos + tw))) <= eps, (name, p, Rw @ pos + tw)


def _cyl_case(c):
    name, r_c, h_c, r_s, p, col, d, n, pos = c
    return (fclo.make_shape(fclo.SPHERE, (r_s,)), fclo.make_shape(fclo.CYLINDER, (r_c, h_c)),
            (name, np.asarray(p, float), col, d, n, pos))


@pytest.mark.parametrize("rotated", [False, True])
def test_sphere_cylinder_known_answers(rotated):
    eps = 22 * EPS
    configs = [_cyl_case(c) for c in cylinder_uniform()]
    rots = sphere_rotations() if rotated else [np.eye(3)]
    n = 0
    for R_S in rots:
        for X_W in world_frames():
            for sph, cyl, cfg in configs:
                _eval(sph, cyl, X_W, R_S, cfg, eps)
                n += 1
    assert n == len(rots) * 7 * 53


def test_sphere_cylinder_incompatible_scales():
    eps = 22 * EPS
    for X_W in world_frames():
        for c in cylinder_nonuniform():
            sph, cyl, cfg = _cyl_case(c)
            _eval(sph, cyl, X_W, np.eye(3), cfg, eps)


def box_uniform():
    """(name, half_size, r, p_BS, colliding, depth, normal, pos)."""
    w, d, h, r = 0.6, 1.2, 3.6, 0.7
    hs = np.array([w / 2, d / 2, h / 2])
    out = []
    p = np.array([hs[0] * 0.5, hs[1] * 0.5, hs[2] + r * 1.1])
    out.append(("separated, +z face", hs, r, p, False, None, None, None))
    out.append(("separated, corner", hs, r, hs + np.array([r, r, r]) * 1.25, False, None, None,
                None))
    td = r * 0.5
    p = hs + np.array([0, 0, r - td])
    out.append(("outside centre, +z face", hs, r, p, True, td, -Z,
                np.array([p[0], p[1], (h - td) / 2])))
    n = unit([-1, -2, -3])
    out.append(("outside centre, corner", hs, r, hs - n * (r - td), True, td, n,
                hs + n * (td * 0.5)))
    inset = hs.min() * 0.5
    for axis in range(3):
        for sign in (-1, 1):
            dr = sign * np.eye(3)[axis]
            out.append(("inside centre, face %d%+d" % (axis, sign), hs, r,
                        dr * (inset - hs[axis]), True, inset + r, dr,
                        dr * ((r + inset) / 2 - hs[axis])))
    p = np.array([0.0, 0.0, hs[2]])
    out.append(("centre on +z face", hs, r, p, True, r, -Z, np.array([p[0], p[1], p[2] - r / 2])))
    out.append(("centre on corner", hs, r, hs.copy(), True, r, np.array([-1.0, 0, 0]),
                np.array([hs[0] - r / 2, hs[1], hs[2]])))
    X, Y = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    for half, n, p in (((10, 10, 10), -X, (2.5, 0, 0)), ((10, 15, 10), -X, (2.5, 0, 0)),
                       ((10, 12, 14), -X, (2.5, 0, 0)), ((15, 10, 10), -Y, (0, 2.5, 0)),
                       ((15, 10, 14), -Y, (0, 2.5, 0)), ((15, 12, 10), -Z, (0, 0, 2.5))):
        out.append(("coincident centres %s" % (half,), np.array(half, float), 5.0, np.zeros(3),
                    True, 15.0, n, np.array(p, float)))
    return out


def box_nonuniform():
    out = []
    hs = np.array([15.0, 1, 1])
    r = 0.01
    p = np.array([hs[0] * 0.95, 0, hs[2] + r * 0.5])
    dep = r - (p[2] - hs[2])
    out.append(("skinny box, small sphere", hs, r, p, True, dep, -Z,
                np.array([p[0], p[1], hs[2] - dep / 2])))
    out.append(("skinny box, small sphere, apart", hs, r,
                np.array([hs[0] * 0.95, 0, hs[2] + r + r * 0.1]), False, None, None, None))
    hs = np.array([0.1, 0.15, 0.2])
    r = 10.0
    n = unit([-1, -2, -3])
    td = hs.min() * 0.5
    out.append(("large sphere, tiny box", hs, r, hs - n * (r - td), True, td, n,
                hs + n * (td * 0.5)))
    out.append(("large sphere, tiny box, apart", hs, r, hs - n * (r + hs.min() * 0.1), False,
                None, None, None))
    return out


def _box_case(c):
    name, hs, r, p, col, d, n, pos = c
    return (fclo.make_shape(fclo.SPHERE, (r,)), fclo.make_shape(fclo.BOX, tuple(2 * hs)),
            (name, np.asarray(p, float), col, d, n, pos))


@pytest.mark.parametrize("rotated", [False, True])
def test_sphere_box_known_answers(rotated):
    eps = 18 * EPS
    configs = [_box_case(c) for c in box_uniform()]
    rots = sphere_rotations() if rotated else [np.eye(3)]
    for R_S in rots:
        for X_W in world_frames():
            for sph, box, cfg in configs:
                _eval(sph, box, X_W, R_S, cfg, eps)


def test_sphere_box_incompatible_scales():
    eps = 18 * EPS
    for X_W in world_frames():
        for c in box_nonuniform():
            sph, box, cfg = _box_case(c)
            _eval(sph, box, X_W, np.eye(3), cfg, eps)


def test_a_planted_mistake_fails():
    """The check is sharp: the same configurations with the expected
    normal's sign flipped, or the depth off by 1e-12, are rejected."""
    sph, cyl, cfg = _cyl_case(cylinder_uniform()[3])
    name, p, col, d, n, pos = cfg
    with pytest.raises(AssertionError):
        _eval(sph, cyl, world_frames()[5], np.eye(3), (name, p, col, d, -n, pos), 22 * EPS)
    with pytest.raises(AssertionError):
        _eval(sph, cyl, world_frames()[5], np.eye(3), (name, p, col, d + 1e-12, n, pos), 22 * EPS)
