"""CPU tests of the cell-complex collapse (cellcomplex::finalize +
CellComplexThinning::prune restated, src/cellcomplex.cpp:364-491,
src/ccthin.cpp:201-771): hand-derived known answers, closed forms, the chi
invariant, and the product's host implementation (vc_cc_*) against the
oracle on random complexes."""
import numpy as np
import pytest


def _faces_erep(foff, fe, f):
    return list(fe[foff[f]:foff[f + 1]])


def test_finalize_known_answer_voro_fixture(oracle):
    """The hand-built complex of the reference's test_voro_merge.cpp:10-47:
    duplicate given edges collapse, face sides reuse the given ids."""
    import voxel_ma_amd as vm
    edges = [(0, 1), (1, 2), (2, 3), (0, 3), (2, 4), (1, 4), (1, 5), (4, 5), (5, 7), (6, 7),
             (4, 6), (0, 3), (3, 8), (0, 8), (0, 8), (8, 9), (0, 9)]
    faces = [[0, 1, 2, 3], [1, 2, 4], [1, 4, 5], [4, 5, 7, 6], [0, 3, 8], [0, 8, 9]]
    exp_edges = [(0, 1), (1, 2), (2, 3), (0, 3), (2, 4), (1, 4), (1, 5), (4, 5), (5, 7), (6, 7),
                 (4, 6), (3, 8), (0, 8), (8, 9), (0, 9)]
    exp_fe = [0, 1, 2, 3, 1, 4, 5, 5, 7, 6, 7, 8, 9, 10, 3, 11, 12, 12, 13, 14]
    for impl in (oracle, vm):
        e, foff, fe = impl.cc_finalize(10, edges, faces)
        assert [tuple(x) for x in e] == exp_edges
        assert list(fe) == exp_fe


def test_prune_known_answer_triangle(oracle):
    """One triangle, everything below threshold (hand-derived FIFO trace):
    FE(f0,e0) removes the face, then EV(e2,v0) and EV(e1,v1): vertex 2 stays."""
    import voxel_ma_amd as vm
    for impl in (oracle, vm):
        e, foff, fe = impl.cc_finalize(3, [], [[0, 1, 2]])
        assert [tuple(x) for x in e] == [(0, 1), (1, 2), (0, 2)]
        rc, rv, re, rf, ref = impl.cc_prune(3, e, foff, fe, np.zeros(3), np.zeros(1), 1.0, 1.0)
        assert rc == 0
        assert list(rv) == [1, 1, 0] and list(re) == [1, 1, 1] and list(rf) == [1]


def test_prune_known_answer_thresholded_square(oracle):
    """Two triangles, only face 0 below the face threshold: face 0, edges
    (0,1), (1,2) and vertex 1 go; face 1 stays with its three edges."""
    import voxel_ma_amd as vm
    for impl in (oracle, vm):
        e, foff, fe = impl.cc_finalize(4, [], [[0, 1, 2], [0, 2, 3]])
        assert [tuple(x) for x in e] == [(0, 1), (1, 2), (0, 2), (2, 3), (0, 3)]
        rc, rv, re, rf, ref = impl.cc_prune(4, e, foff, fe, np.zeros(5), [0.5, 2.0], 1.0, 1.0)
        assert rc == 0
        assert list(rv) == [0, 1, 0, 0]
        assert list(re) == [1, 1, 0, 0, 0]
        assert list(rf) == [1, 0]
        assert list(ref) == [0, 0, 1, 1, 1]


def test_simple_triangulate_zigzag(oracle):
    # (last, 0, 1), (last, 1, last-1), (last-1, 1, 2), ... (src/geomalgo.cpp:17-34)
    assert [tuple(t) for t in oracle.simple_triangulate([10, 11, 12, 13, 14])] == \
        [(14, 10, 11), (14, 11, 13), (13, 11, 12)]
    assert [tuple(t) for t in oracle.simple_triangulate([5, 6, 7])] == [(7, 5, 6)]


def _grid(m, n):
    faces = []
    for i in range(m):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, i * (n + 1) + j + 1, (i + 1) * (n + 1) + j + 1, \
                (i + 1) * (n + 1) + j
            faces += [[a, b, c], [a, c, d]]
    return (m + 1) * (n + 1), faces


def _chi(nv, e, foff, rv, re, rf):
    return int((rv == 0).sum()) - int((re == 0).sum()) + int((rf == 0).sum())


def test_disk_collapses_to_a_vertex(oracle):
    nv, faces = _grid(4, 5)
    e, foff, fe = oracle.cc_finalize(nv, [], faces)
    rc, rv, re, rf, ref = oracle.cc_prune(nv, e, foff, fe, np.zeros(len(e)), np.zeros(len(faces)),
                                          1.0, 1.0)
    assert rc == 0 and rf.all() and re.all() and (rv == 0).sum() == 1


def test_annulus_collapses_to_a_cycle(oracle):
    # ring of 8 x 2 quads, each split in two triangles
    n, w = 8, 2
    faces = []
    for i in range(n):
        for j in range(w):
            a = j * n + i
            b = j * n + (i + 1) % n
            c = (j + 1) * n + (i + 1) % n
            d = (j + 1) * n + i
            faces += [[a, b, c], [a, c, d]]
    nv = (w + 1) * n
    e, foff, fe = oracle.cc_finalize(nv, [], faces)
    rc, rv, re, rf, ref = oracle.cc_prune(nv, e, foff, fe, np.zeros(len(e)), np.zeros(len(faces)),
                                          1.0, 1.0)
    assert rc == 0 and rf.all()
    left_e = np.nonzero(re == 0)[0]
    left_v = np.nonzero(rv == 0)[0]
    assert len(left_e) == len(left_v) > 2        # a single cycle of isolated edges
    assert all(ref[i] == 0 for i in left_e)


def test_closed_surface_is_untouched(oracle):
    faces = [[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]]
    e, foff, fe = oracle.cc_finalize(6, [], faces)
    rc, rv, re, rf, ref = oracle.cc_prune(6, e, foff, fe, np.zeros(len(e)), np.zeros(8), 1.0, 1.0)
    assert rc == 0 and not rv.any() and not re.any() and not rf.any()


def _random_complex(rng):
    m, n = int(rng.integers(1, 6)), int(rng.integers(1, 6))
    nv, faces = _grid(m, n)
    keep = rng.random(len(faces)) < rng.uniform(0.4, 1.0)
    faces = [f for f, k in zip(faces, keep) if k]
    # merge some triangle pairs into quads (polygon faces of mixed sizes)
    out = []
    i = 0
    while i < len(faces):
        if i + 1 < len(faces) and rng.random() < 0.3 and faces[i][0] == faces[i + 1][0] and \
                faces[i][2] == faces[i + 1][1]:
            a, b, c = faces[i]
            out.append([a, b, c, faces[i + 1][2]])
            i += 2
        else:
            out.append(faces[i])
            i += 1
    extra = [(int(a), int(b)) for a, b in rng.integers(0, nv, size=(int(rng.integers(0, 6)), 2))
             if a != b]
    return nv, extra, out


@pytest.mark.parametrize("seed", range(4))
def test_product_matches_oracle_on_random_complexes(oracle, seed):
    import voxel_ma_amd as vm
    rng = np.random.default_rng(100 + seed)
    for _ in range(60):
        nv, extra, faces = _random_complex(rng)
        eo, fo, feo = oracle.cc_finalize(nv, extra, faces)
        ep, fp, fep = vm.cc_finalize(nv, extra, faces)
        assert np.array_equal(eo, ep) and np.array_equal(fo, fp) and np.array_equal(feo, fep)
        em = rng.random(len(eo)).astype(np.float32)
        fm = rng.random(len(faces)).astype(np.float32)
        f_t, l_t = float(rng.uniform(0, 1.2)), float(rng.uniform(0, 1.2))
        ro = oracle.cc_prune(nv, eo, fo, feo, em, fm, f_t, l_t)
        rp = vm.cc_prune(nv, ep, fp, fep, em, fm, f_t, l_t)
        assert ro[0] == rp[0] == 0
        for a, b in zip(ro[1:], rp[1:]):
            assert np.array_equal(a, b)
        # every collapse removes a simple pair: chi of what is left is unchanged
        zeros = (np.zeros(nv, np.uint8), np.zeros(len(eo), np.uint8), np.zeros(len(faces), np.uint8))
        assert _chi(nv, eo, fo, *ro[1:4]) == _chi(nv, eo, fo, *zeros)
