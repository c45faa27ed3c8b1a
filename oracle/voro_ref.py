"""CPU restatement of the Voronoi-side steps (test infrastructure: the
checker for include/vc_voro.h, never part of the product path).

Each function follows the reference line by line in plain Python / numpy
float32 arithmetic:
  tag_points        VoroInfo::tagVert (src/voroinfo.cpp:447-454):
                    std::round of each coordinate, then in-bounds and
                    occupied (include/spaceinfo.h:53-58, :93-105)
  measures          computeFacesMeasure / computeEdgesMeasure /
                    computeVertexMeasure, LAMBDA (src/voroinfo.cpp:1432-1574)
  extract_inside    extractInsideWithMeasure (src/voroinfo.cpp:624-729)
  merge_close       mergeCloseVts (src/voroinfo.cpp:1064-1384), new edges
                    in ascending order (the reference iterates an
                    unordered_set)
  radii             the radius loop of computeInfoRelatedtoSites and the
                    TetGen loader (src/voroinfo.cpp:290-325, :404-441): faces
                    in order, each vertex takes its distance to the face's
                    first site, the last assignment stays
  part_inside       getPartInside (src/voroinfo.cpp:573-620)
  voronoi_from_ft   the discrete Voronoi diagram of the feature transform
                    (the construction that replaces computeVD's TetGen step,
                    src/highlevelalgo.cpp:487-529): a face per pair of
                    6-adjacent inside voxels with different nearest
                    background voxels
Pinned by hand-derived known answers in tests/test_voro.py; voronoi_from_ft
also against a brute-force nearest-background search and a per-voxel loop.
"""
import collections

import numpy as np


def round_half_away(x):
    """std::round on float32 values, exactly."""
    x = np.asarray(x, np.float32)
    t = np.trunc(x)
    frac = (x - t).astype(np.float32)
    return (t + np.where(np.abs(frac) >= np.float32(0.5), np.sign(x), 0)).astype(np.float64)


def tag_points(occ, pts):
    occ = np.asarray(occ)
    r = round_half_away(np.asarray(pts, np.float32).reshape(-1, 3))
    ok = np.all(np.isfinite(r), axis=1) & np.all(np.abs(r) < 1e9, axis=1)
    ri = np.where(ok[:, None], r, 0).astype(np.int64)
    ok &= np.all((ri >= 0) & (ri < np.array(occ.shape)), axis=1)
    out = np.zeros(len(r), np.uint8)
    idx = np.nonzero(ok)[0]
    out[idx] = occ[ri[idx, 0], ri[idx, 1], ri[idx, 2]] != 0
    return out


def site_dist(s, a, b):
    d = (s[b] - s[a]).astype(np.float32)
    q = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1]))
    q = np.float32(q + np.float32(d[2] * d[2]))
    return np.float32(np.sqrt(q))


def measures(nv, edges, faces_vts, face_sites, sites, face_valid):
    """(lam_v, lam_e, lam_f) with the adjacency of a finalized complex:
    edges of a vertex in edge order, faces of an edge in face order (faces
    whose polygon has the edge as a side)."""
    sites = np.asarray(sites, np.float32).reshape(-1, 3)
    nf, ne = len(faces_vts), len(edges)
    lam_f = np.array([site_dist(sites, *face_sites[f]) for f in range(nf)], np.float32)
    key = {}
    for e, (a, b) in enumerate(edges):
        key.setdefault((min(a, b), max(a, b)), e)
    ef = [[] for _ in range(ne)]
    for f, poly in enumerate(faces_vts):
        for j in range(len(poly)):
            a, b = poly[j], poly[(j + 1) % len(poly)]
            ef[key[(min(a, b), max(a, b))]].append(f)
    ve = [[] for _ in range(nv)]
    for e, (a, b) in enumerate(edges):
        ve[a].append(e)
        ve[b].append(e)
    lam_e = np.zeros(ne, np.float32)
    for e in range(ne):
        for f in ef[e]:
            if face_valid[f]:
                lam_e[e] = max(lam_e[e], lam_f[f])
    lam_v = np.zeros(nv, np.float32)
    for v in range(nv):
        for e in ve[v]:
            for f in ef[e]:
                if face_valid[f]:
                    lam_v[v] = max(lam_v[v], lam_f[f])
    return lam_v, lam_e, lam_f, ve, ef


def zigzag(p):
    tris, i, j, inc = [], len(p) - 1, 0, True
    while i - 1 > j:
        if inc:
            tris.append((p[i], p[j], p[j + 1]))
            j += 1
        else:
            tris.append((p[i], p[j], p[i - 1]))
            i -= 1
        inc = not inc
    return tris


def extract_inside(nv, edges, faces_vts, v_valid, lam_e, lam_f):
    vid = [v for v in range(nv) if v_valid[v]]
    new_id = {v: k for k, v in enumerate(vid)}
    e_out, em, seen = [], [], set()
    for e, (a, b) in enumerate(edges):
        if v_valid[a] and v_valid[b]:
            e_out.append((a, b))
            em.append(lam_e[e])
            seen.add((a, b))
    tris, from_fi, fm = [], [], []
    for f, poly in enumerate(faces_vts):
        if not all(v_valid[v] for v in poly):
            continue
        for t in zigzag(poly):
            fm.append(lam_f[f])
            tris.append(t)
            from_fi.append(f)
            for s in range(3):
                a, b = t[s], t[(s + 1) % 3]
                k = (min(a, b), max(a, b))
                if k not in seen:
                    seen.add(k)
                    e_out.append(k)
                    em.append(lam_f[f])
    e_out = [(new_id[a], new_id[b]) for a, b in e_out]
    tris = [tuple(new_id[v] for v in t) for t in tris]
    return vid, e_out, np.array(em, np.float32), tris, from_fi, np.array(fm, np.float32)


def merge_close(vts, edges, faces_vts, v_valid, eps, only_inside):
    vts = np.asarray(vts, np.float32).reshape(-1, 3)
    nv = len(vts)
    ve = [[] for _ in range(nv)]
    for e, (a, b) in enumerate(edges):
        ve[a].append(e)
        ve[b].append(e)
    sq_r = np.float32(np.float32(eps) * np.float32(eps))

    def d2(a, b):
        d = (vts[b] - vts[a]).astype(np.float32)
        q = np.float32(d[0] * d[0])
        q = np.float32(q + np.float32(d[1] * d[1]))
        return np.float32(q + np.float32(d[2] * d[2]))

    visited = [False] * nv
    old_to_new = [-1] * nv
    new_vts = []
    for vi in range(nv):
        if visited[vi] or (only_inside and not v_valid[vi]):
            continue
        new_vts.append(vi)
        q = collections.deque([vi])
        visited[vi] = True
        old_to_new[vi] = len(new_vts) - 1
        while q:
            cur = q.popleft()
            for e in ve[cur]:
                a, b = edges[e]
                ni = b if a == cur else a
                if not visited[ni] and d2(vi, ni) <= sq_r:
                    q.append(ni)
                    visited[ni] = True
                    old_to_new[ni] = len(new_vts) - 1
    eset, keep = set(), []
    for (a, b) in edges:
        if only_inside and not (v_valid[a] and v_valid[b]):
            keep.append(0)
            continue
        u, w = old_to_new[a], old_to_new[b]
        if u == w:
            keep.append(0)
            continue
        eset.add((min(u, w), max(u, w)))
        keep.append(1)
    new_faces, remove = [], []
    for poly in faces_vts:
        if only_inside and not all(v_valid[v] for v in poly):
            remove.append(1)
            continue
        fv = []
        for v in poly:
            c = old_to_new[v]
            if c not in fv:
                fv.append(c)
        if len(fv) < 2:
            remove.append(1)
        elif len(fv) == 2:
            eset.add((min(fv), max(fv)))
            remove.append(1)
        else:
            new_faces.append(fv)
            remove.append(0)
    return old_to_new, new_vts, new_faces, remove, sorted(eset), keep


def radii(vts, faces_vts, face_sites, sites):
    vts = np.asarray(vts, np.float32).reshape(-1, 3)
    sites = np.asarray(sites, np.float32).reshape(-1, 3)
    r = np.zeros(len(vts), np.float32)
    for f, poly in enumerate(faces_vts):
        s = sites[face_sites[f][0]]
        for v in poly:
            d = (vts[v] - s).astype(np.float32)
            q = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1]))
            q = np.float32(q + np.float32(d[2] * d[2]))
            r[v] = np.sqrt(q)
    return r


def part_inside(nv, faces_vts, v_valid):
    vi = [v for v in range(nv) if v_valid[v]]
    new = {v: i for i, v in enumerate(vi)}
    tris, es = [], set()
    for poly in faces_vts:
        for t in zigzag(list(poly)):
            if all(v_valid[v] for v in t):
                c = [new[v] for v in t]
                tris.append(tuple(c))
                for k in range(3):
                    a, b = c[k], c[(k + 1) % 3]
                    es.add((min(a, b), max(a, b)))
    return vi, sorted(es), tris


def voronoi_from_ft(occ, ft):
    """occ: (nx, ny, nz) uint8; ft: the same shape, linear index of each
    voxel's nearest background voxel (lowest index among the nearest).
    Returns (vts (V, 3) float32, faces (F, 4) int32, face_sites (F, 2)
    int32, sites (S, 3) float32) in the order of vc_voro_from_ft: faces by
    lower voxel (linear index) then axis; corners (0,0), (1,0), (1,1), (0,1)
    over the axes a+1, a+2; vertices in corner-lattice order; sites in
    linear order."""
    occ = np.asarray(occ) != 0
    ft = np.asarray(ft, np.int64)
    nx, ny, nz = occ.shape
    shape = np.array(occ.shape)
    rec = []  # (p linear, axis, corner lattice ids (4), site p, site q)
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, shape[a] - 1)
        hi[a] = slice(1, shape[a])
        m = occ[tuple(lo)] & occ[tuple(hi)] & (ft[tuple(lo)] != ft[tuple(hi)])
        p = np.argwhere(m)  # voxel coordinates of the lower voxel, C order
        if not len(p):
            continue
        b, c = (a + 1) % 3, (a + 2) % 3
        corners = []
        for db, dc in ((0, 0), (1, 0), (1, 1), (0, 1)):
            q = p.copy()
            q[:, a] += 1
            q[:, b] += db
            q[:, c] += dc
            corners.append((q[:, 0] * (ny + 1) + q[:, 1]) * (nz + 1) + q[:, 2])
        lin = (p[:, 0] * ny + p[:, 1]) * nz + p[:, 2]
        qn = p.copy()
        qn[:, a] += 1
        lin_q = (qn[:, 0] * ny + qn[:, 1]) * nz + qn[:, 2]
        rec.append((lin * 3 + a, np.stack(corners, 1),
                    ft.reshape(-1)[lin], ft.reshape(-1)[lin_q]))
    if not rec:
        return (np.zeros((0, 3), np.float32), np.zeros((0, 4), np.int32),
                np.zeros((0, 2), np.int32), np.zeros((0, 3), np.float32))
    key = np.concatenate([r[0] for r in rec])
    order = np.argsort(key, kind="stable")
    corners = np.concatenate([r[1] for r in rec])[order]
    s_p = np.concatenate([r[2] for r in rec])[order]
    s_q = np.concatenate([r[3] for r in rec])[order]
    lat, faces = np.unique(corners.reshape(-1), return_inverse=True)
    faces = faces.reshape(-1, 4).astype(np.int32)
    k = lat % (nz + 1)
    r = lat // (nz + 1)
    vts = np.stack([r // (ny + 1), r % (ny + 1), k], 1).astype(np.float32) - np.float32(0.5)
    used, sid = np.unique(np.concatenate([s_p, s_q]), return_inverse=True)
    face_sites = np.stack([sid[:len(s_p)], sid[len(s_p):]], 1).astype(np.int32)
    sites = np.stack([used // (ny * nz), (used // nz) % ny, used % nz], 1).astype(np.float32)
    return vts, faces, face_sites, sites
