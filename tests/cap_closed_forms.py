"""Closed forms and limits of the complex absorbing potential's corner and edge blends.

The reference kernel (sources/cuda_kernels/absorbing_potential.cu:58-278) blends the
maximum distance, exponent and outer potential of the axes a voxel lies outside of with
mix(q, u, v) = q_z v + (1 - v)(q_y u + q_x (1 - u)) (common.cu:394-397) at atan angles.
Two facts follow from that definition without restating any branch:

* Symmetric parameters (one outer potential p, one exponent e and one layer depth, so
  that every axis's max distance is the same c): every blend whose three entries are
  p, e or c collapses to them, and the voxel holds the 1-D ramp
  p * (max(|d| - offset, 0) / c)^e of its distance |d| from the box (over the outside
  axes).  The exceptions are exactly the entries the kernel writes as 0 or as the box
  half-size: on the (0, -1, +-1) edges (:167-176, :187-196) the max distance is
  mix((box_half_x, c, c), 0, v) = c v + (1 - v) box_half_x; on the (0, +1, -1) edge
  (:269-278) so is it, and the exponent and potential are mix((0, e, e), 0, v) = e v and
  p v, with v = atan(|d_z| / |d_y|) * 2 / pi.
* Limits: a blend with u -> 0 or 1, v -> 0 or 1 takes the entry of one axis, so a voxel
  just outside a box plane (distance eps) in an edge region meets the 1-D ramp of the
  face it borders, and a corner voxel meets the edge it borders -- wherever the kernel's
  fixed mix() arguments make the limit that axis (listed in ``edge_face_limits``).

These functions compute the expected values from the geometry alone (FP64), so that the
oracle (and, through it, the GPU kernel) is checked against them, not against itself.
"""

import numpy as np

f32 = np.float32


def grid_coordinates(shape, dx):
    """Voxel coordinates r = dr*i - 0.5*N*dr in float32 (absorbing_potential.cu:52)."""
    return [f32(dx[d]) * np.arange(shape[d], dtype=f32) - f32(0.5) * f32(shape[d]) * f32(dx[d]) for d in range(3)]


def box_half(shape, dx):
    return np.array([0.5 * shape[d] * dx[d] for d in range(3)])


def side_distances(shape, dx, bottom, top):
    """Per axis: side (-1 below the box, +1 above, 0 strictly inside, 9 on a plane) and
    |d| (distance from the plane on the outside side, 0 inside), float64 from float32 d."""
    r = grid_coordinates(shape, dx)
    sides, dist = [], []
    for d in range(3):
        dB = (r[d] - f32(bottom[d])).astype(np.float64)
        dT = (r[d] - f32(top[d])).astype(np.float64)
        s = np.where(dB < 0, -1, np.where(dT > 0, 1, np.where((dB > 0) & (dT < 0), 0, 9)))
        sides.append(s)
        dist.append(np.where(s == -1, -dB, np.where(s == 1, dT, 0.0)))
    return sides, dist


def ramp(dist, offset, maxd, e, p):
    with np.errstate(divide="ignore", invalid="ignore"):
        return p * np.power(np.maximum(dist - offset, 0.0) / maxd, e)


def symmetric_closed_form(shape, dx, bottom, top, offset, p, e):
    """Expected CAP for one outer potential p and exponent e on every axis and side, and a
    box whose layer depth box_half - |corner| is the same on every axis and side."""
    bh = box_half(shape, dx)
    depth = [bh[d] - abs(bottom[d]) for d in range(3)] + [bh[d] - abs(top[d]) for d in range(3)]
    assert np.ptp(depth) < 1e-9, "symmetric closed form needs one layer depth"
    c = depth[0] - offset
    sides, dist = side_distances(shape, dx, bottom, top)
    S = np.stack(np.meshgrid(*sides, indexing="ij"))
    D = np.stack(np.meshgrid(*dist, indexing="ij"))
    on_plane = np.any(S == 9, axis=0)
    outside = (S != 0) & (S != 9)
    rho = np.sqrt(np.sum(np.where(outside, D, 0.0) ** 2, axis=0))
    out = ramp(rho, offset, c, e, p)
    # the edges with x inside and y below, and x inside, y above, z below
    v = np.arctan2(D[2], D[1]) * 2.0 / np.pi
    maxd_q = c * v + (1.0 - v) * bh[0]
    yz = (S[0] == 0) & outside[1] & outside[2]
    q1 = yz & (S[1] == -1)
    q2 = yz & (S[1] == 1) & (S[2] == -1)
    out = np.where(q1, ramp(rho, offset, maxd_q, e, p), out)
    out = np.where(q2, ramp(rho, offset, maxd_q, e * v, p * v), out)
    out = np.where(on_plane | ~np.any(outside, axis=0), 0.0, out)
    return out


def per_side_closed_form(shape, dx, bottom, top, offset, p_neg, p_pos, e_neg, e_pos):
    """Expected CAP where a voxel's outside axes all lie on one side s, for parameters that
    are one value per side (p_s, e_s, layer depth per side the same on every axis).
    Returns (expected, mask of the voxels it covers)."""
    bh = box_half(shape, dx)
    dneg = [bh[d] - abs(bottom[d]) for d in range(3)]
    dpos = [bh[d] - abs(top[d]) for d in range(3)]
    assert np.ptp(dneg) < 1e-9 and np.ptp(dpos) < 1e-9
    sides, dist = side_distances(shape, dx, bottom, top)
    S = np.stack(np.meshgrid(*sides, indexing="ij"))
    D = np.stack(np.meshgrid(*dist, indexing="ij"))
    outside = (S != 0) & (S != 9)
    nout = outside.sum(axis=0)
    allneg = np.all(np.where(outside, S == -1, True), axis=0) & (nout > 0) & ~np.any(S == 9, axis=0)
    allpos = np.all(np.where(outside, S == 1, True), axis=0) & (nout > 0) & ~np.any(S == 9, axis=0)
    rho = np.sqrt(np.sum(np.where(outside, D, 0.0) ** 2, axis=0))
    out = np.where(allneg, ramp(rho, offset, dneg[0] - offset, e_neg, p_neg),
                   ramp(rho, offset, dpos[0] - offset, e_pos, p_pos))
    v = np.arctan2(D[2], D[1]) * 2.0 / np.pi
    # the (0, -1, -1) edge reads the box half-size of x for its max distance at u = 0
    q = allneg & (S[0] == 0) & outside[1] & outside[2]
    out = np.where(q, ramp(rho, offset, (dneg[0] - offset) * v + (1.0 - v) * bh[0], e_neg, p_neg), out)
    return out, allneg | allpos


def face_ramp(shape, dx, bottom, top, offset, pot_pos, pot_neg, exp_pos, exp_neg, axis, side, dist):
    """1-D ramp of one face (absorbing_potential.cu:279-303) at distance dist."""
    bh = box_half(shape, dx)
    corner = bottom[axis] if side < 0 else top[axis]
    maxd = bh[axis] - (abs(corner) + offset)
    e = exp_neg[axis] if side < 0 else exp_pos[axis]
    p = pot_neg[axis] if side < 0 else pot_pos[axis]
    return ramp(dist, offset, maxd, e, p)


def eps_planes(shape, dx, depth_index, eps):
    """Box planes eps beyond voxel planes: voxel depth_index lies eps below the bottom
    plane and voxel N-1-depth_index eps above the top plane on every axis."""
    r = grid_coordinates(shape, dx)
    bottom = tuple(float(r[d][depth_index]) + eps for d in range(3))
    top = tuple(float(r[d][shape[d] - 1 - depth_index]) - eps for d in range(3))
    return bottom, top


def edge_face_limits(cap, shape, dx, bottom, top, offset, pot_pos, pot_neg, exp_pos, exp_neg, depth_index):
    """Worst relative deviation of every edge voxel just outside a box plane from the limit the
    kernel's mix() arguments give there (eps-geometry of ``eps_planes``):

    * "u" edges (x and y outside) and pillar edges (x and z outside) meet both faces;
    * the (0, +1, +1) edge meets both faces;
    * the (0, -1, +-1) and (0, +1, -1) edges meet the z face as |d_y| -> 0;  as |d_z| -> 0
      the (0, -1, +-1) edges tend to p_y (|d_y| - offset)^e_y / box_half_x^e_y and the
      (0, +1, -1) edge to 0.

    Returns {label: deviation} with deviations relative to the face's outer potential."""
    bh = box_half(shape, dx)
    n = shape
    k = depth_index
    eps_idx = {-1: [k, k, k], 1: [n[0] - 1 - k, n[1] - 1 - k, n[2] - 1 - k]}
    deep = {-1: [list(range(0, k)) for _ in range(3)],
            1: [list(range(n[d] - k, n[d])) for d in range(3)]}
    mid = [n[d] // 2 for d in range(3)]
    sides, dist = side_distances(shape, dx, bottom, top)
    params = (shape, dx, bottom, top, offset, pot_pos, pot_neg, exp_pos, exp_neg)
    devs = {}

    def check(label, axes_sides, eps_axis, limit):
        # voxels: eps_axis at its eps index, the other outside axis deep, the inside axis mid
        (a1, s1), (a2, s2) = axes_sides
        far_axis, far_side = (a2, s2) if eps_axis == a1 else (a1, s1)
        eps_side = s1 if eps_axis == a1 else s2
        worst = 0.0
        for j in deep[far_side][far_axis]:
            idx = list(mid)
            idx[eps_axis] = eps_idx[eps_side][eps_axis]
            idx[far_axis] = j
            assert sides[eps_axis][idx[eps_axis]] == eps_side and sides[far_axis][j] == far_side
            got = float(cap[tuple(idx)])
            want = limit(far_axis, far_side, dist[far_axis][j])
            scale = abs((pot_neg if far_side < 0 else pot_pos)[far_axis])
            worst = max(worst, abs(got - want) / scale)
        devs[label] = worst

    def meets_face(far_axis, far_side, d):
        return face_ramp(*params, far_axis, far_side, d)

    for sx in (-1, 1):
        for sy in (-1, 1):
            for eps_axis in (0, 1):
                check(f"u-edge ({sx},{sy},0) eps on {'xy'[eps_axis]}", ((0, sx), (1, sy)), eps_axis, meets_face)
        for sz in (-1, 1):
            for eps_axis in (0, 2):
                check(f"pillar ({sx},0,{sz}) eps on {'x z'[eps_axis]}", ((0, sx), (2, sz)), eps_axis, meets_face)
    for sy, sz in ((1, 1), (-1, 1), (-1, -1), (1, -1)):
        check(f"yz-edge (0,{sy},{sz}) eps on y", ((1, sy), (2, sz)), 1, meets_face)
    check("yz-edge (0,1,1) eps on z", ((1, 1), (2, 1)), 2, meets_face)
    for sz in (-1, 1):
        def quirk(far_axis, far_side, d):
            return ramp(d, offset, bh[0], exp_neg[1], pot_neg[1])
        check(f"yz-edge (0,-1,{sz}) eps on z", ((1, -1), (2, sz)), 2, quirk)
    check("yz-edge (0,1,-1) eps on z", ((1, 1), (2, -1)), 2, lambda *a: 0.0)
    return devs


def corner_edge_continuity(cap, shape, depth_index):
    """Worst relative difference between corner voxels just outside a z or y box plane and the
    edge voxel with the same outside coordinates (z or y moved inside): as |d_z| -> 0 a
    corner's blend tends to the "u" edge's, as |d_y| -> 0 to the pillar edge's."""
    n = shape
    k = depth_index
    eps_idx = {-1: [k] * 3, 1: [n[d] - 1 - k for d in range(3)]}
    mid = [n[d] // 2 for d in range(3)]
    devs = {}
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                worst = 0.0
                xs = [0, 1] if sx < 0 else [n[0] - 1, n[0] - 2]
                ys = [0, 1] if sy < 0 else [n[1] - 1, n[1] - 2]
                zs = [0, 1] if sz < 0 else [n[2] - 1, n[2] - 2]
                for i in xs:
                    for j in ys:
                        corner = cap[i, j, eps_idx[sz][2]]
                        edge = cap[i, j, mid[2]]
                        worst = max(worst, abs(float(corner) - float(edge)) / max(abs(float(edge)), 1e-30))
                    for kk in zs:
                        corner = cap[i, eps_idx[sy][1], kk]
                        edge = cap[i, mid[1], kk]
                        worst = max(worst, abs(float(corner) - float(edge)) / max(abs(float(edge)), 1e-30))
                devs[f"corner ({sx},{sy},{sz})"] = worst
    return devs
