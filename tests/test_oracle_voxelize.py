"""Pins of the voxeliser oracle (oracle/voxelize_oracle.c) by closed forms and
brute force: straight streamlines with hand-computed voxel crossings,
orientations and norms; dense sampling of random smooth streamlines; the
centroid mode; blur replicas along offset lines; the EC peak selection; the
ISO voxels and the TDI map (trk2dictionary_c.cpp:289-869,
trk2dictionary.pyx:50-59)."""
import numpy as np
import pytest

from commit_amd import synth
from oracle import voxelize as vo

DIRS = synth.hemisphere_directions(500)
H = synth.hash_table(DIRS)
H_X = int(H[90 * 181 + 0])       # colatitude 90, longitude 0: the x axis
H_Y = int(H[90 * 181 + 90])      # the y axis
H_Z = int(H[0])                  # colatitude 0: the z axis
H_MX = int(H[90 * 181 + 180])    # longitude 180: -x with y = 0


def lin(x, y, z, dim):
    return x + dim[0] * (y + dim[1] * z)


def run(fibs, dim, pix=(1.0, 1.0, 1.0), **kw):
    pts = np.concatenate([np.asarray(f, np.float32).reshape(-1, 3) for f in fibs]) if fibs else \
        np.zeros((0, 3), np.float32)
    ptr = np.zeros(len(fibs) + 1, np.int64)
    ptr[1:] = np.cumsum([len(f) for f in fibs])
    r = vo.trk2dictionary(pts, ptr, dim, pix, H, n_threads=3, **kw)
    assert r["status"] == 0
    return r


def test_hash_axes_are_distinct():
    assert len({H_X, H_Y, H_Z, H_MX}) >= 3
    # the table returns the nearest direction to each axis
    for idx, u in ((H_X, (1, 0, 0)), (H_Y, (0, 1, 0)), (H_Z, (0, 0, 1))):
        assert abs(DIRS[idx] @ np.array(u, float)) >= np.max(np.abs(DIRS @ np.array(u, float))) - 1e-12


def test_straight_streamline_along_x():
    dim = (6, 4, 4)
    r = run([[[0.3, 1.5, 1.5], [2.5, 1.5, 1.5], [4.7, 1.5, 1.5]]], dim)
    np.testing.assert_array_equal(r["IC_v"], [lin(x, 1, 1, dim) for x in range(5)])
    np.testing.assert_array_equal(r["IC_o"], [H_X] * 5)
    np.testing.assert_array_equal(r["IC_f"], 0)
    np.testing.assert_allclose(r["IC_len"], [0.7, 1, 1, 1, 0.7], atol=1e-6)
    np.testing.assert_allclose(r["TRK_len"], [4.4], rtol=1e-6)
    np.testing.assert_allclose(r["TRK_lenTot"], [4.4], rtol=1e-6)
    np.testing.assert_allclose(r["TRK_norm"], [np.sqrt(2 * 0.49 + 3)], rtol=1e-6)
    np.testing.assert_array_equal(r["TRK_kept"], [1])
    np.testing.assert_array_equal(r["ISO_v"], r["IC_v"])


def test_straight_streamlines_along_y_and_z_anisotropic():
    dim, pix = (3, 5, 7), (2.0, 1.5, 0.5)
    # along -y (flipped into y >= 0), then along +z
    r = run([[[3.0, 6.9, 1.2], [3.0, 3.0, 1.2], [3.0, 0.2, 1.2]],
             [[1.0, 1.0, 0.1], [1.0, 1.0, 1.6], [1.0, 1.0, 3.4]]], dim, pix)
    f0 = r["IC_f"] == 0
    # y from 6.9 down to 0.2 crosses voxel boundaries 6.0, 4.5, 3.0, 1.5
    np.testing.assert_array_equal(r["IC_v"][f0], [lin(1, y, 2, dim) for y in range(5)])
    np.testing.assert_array_equal(r["IC_o"][f0], H_Y)
    np.testing.assert_allclose(r["IC_len"][f0], [1.3, 1.5, 1.5, 1.5, 0.9], atol=1e-5)
    f1 = r["IC_f"] == 1
    np.testing.assert_array_equal(r["IC_v"][f1], [lin(0, 0, z, dim) for z in range(7)])
    np.testing.assert_array_equal(r["IC_o"][f1], H_Z)
    np.testing.assert_allclose(r["IC_len"][f1], [0.4, 0.5, 0.5, 0.5, 0.5, 0.5, 0.4], atol=1e-5)
    np.testing.assert_allclose(r["TRK_len"], [6.7, 3.3], rtol=1e-6)


def test_segments_in_map_order_and_merged():
    """Key order (o, z, y, x) across orientations; a voxel revisited with the
    same orientation merges into one segment."""
    dim = (4, 4, 4)
    # along +x in row (y=0, z=0), then up +y, then back along -x in row y=2
    r = run([[[0.5, 0.5, 0.5], [2.5, 0.5, 0.5], [2.5, 2.5, 0.5], [0.5, 2.5, 0.5]]], dim)
    o, v = r["IC_o"].astype(np.int64), r["IC_v"].astype(np.int64)
    key = o * 64 + v
    assert np.all(np.diff(key) > 0)
    # the x-runs (and the -x run, flipped to +x) share one orientation
    xs = v[o == H_X]
    np.testing.assert_array_equal(np.sort(xs), sorted([lin(0, 0, 0, dim), lin(1, 0, 0, dim), lin(2, 0, 0, dim),
                                                       lin(0, 2, 0, dim), lin(1, 2, 0, dim), lin(2, 2, 0, dim)]))
    # voxel (2,0,0) holds 0.5 along x then 0.5 along y
    sel = (v == lin(2, 0, 0, dim))
    assert sel.sum() == 2
    np.testing.assert_allclose(np.sort(r["IC_len"][sel]), [0.5, 0.5], atol=1e-6)
    np.testing.assert_allclose(r["TRK_len"], [6.0], rtol=1e-6)


def _steps(f, shift=(0, 0, 0), pix=(1, 1, 1)):
    sx = np.float32(shift[0]) * np.float32(pix[0])
    sy = np.float32(shift[1]) * np.float32(pix[1])
    sz = np.float32(shift[2]) * np.float32(pix[2])
    return (f + np.array([sx, sy, sz], np.float32)).astype(np.float64)


def _dense(fibs, dim, pix, K=4000, shift=(0, 0, 0)):
    """Per-voxel length of each streamline by dense sampling of its polyline."""
    out = []
    for f in fibs:
        S = _steps(np.asarray(f, np.float32), shift, pix)
        acc = {}
        for i in range(1, len(S)):
            a, b = S[i - 1], S[i]
            t = (np.arange(K) + 0.5) / K
            mid = a[None] + (b - a)[None] * t[:, None]
            vox = np.floor(mid / np.array(pix)).astype(np.int64)
            ok = np.all((vox >= 0) & (vox < np.array(dim)), axis=1)
            seg = np.linalg.norm(b - a) / K
            for vv in np.unique(lin(vox[ok, 0], vox[ok, 1], vox[ok, 2], dim)):
                acc[int(vv)] = acc.get(int(vv), 0.0) + seg * np.count_nonzero(
                    lin(vox[ok, 0], vox[ok, 1], vox[ok, 2], dim) == vv)
        out.append(acc)
    return out


def test_random_streamlines_against_dense_sampling():
    dim, pix = (12, 10, 8), (1.0, 1.25, 0.8)
    pts, ptr = synth.streamlines(5, 40, dim, 1.0, mode=0, step_vox=0.7)
    # scale into the anisotropic box
    pts = (pts * np.array([pix[0], pix[1], pix[2]], np.float32)).astype(np.float32)
    fibs = [pts[ptr[i]:ptr[i + 1]] for i in range(len(ptr) - 1)]
    r = vo.trk2dictionary(pts, ptr, dim, pix, H, min_seg_len=0.0, n_threads=2)
    assert r["status"] == 0
    dense = _dense(fibs, dim, pix)
    kept = np.nonzero(r["TRK_kept"])[0]
    assert kept.size == len(fibs)
    for kf, f in enumerate(kept):
        sel = r["IC_f"] == kf
        got = {}
        for v, ln in zip(r["IC_v"][sel], r["IC_len"][sel]):
            got[int(v)] = got.get(int(v), 0.0) + float(ln)
        exp = dense[f]
        keys = set(got) | set(exp)
        for k in keys:
            assert abs(got.get(k, 0.0) - exp.get(k, 0.0)) < 2e-3, (f, k, got.get(k), exp.get(k))
        # FiberLen: the polyline length inside the grid
        np.testing.assert_allclose(r["TRK_len"][kf], sum(exp.values()), rtol=1e-4)
        # TRK_norm from the per-voxel sums
        np.testing.assert_allclose(r["TRK_norm"][kf], np.sqrt(sum(x * x for x in got.values())), rtol=1e-6)
    # every orientation is the hashed direction of its step: within the
    # table's 1-degree resolution of the nearest direction
    for kf, f in enumerate(kept):
        S = _steps(fibs[f], (0, 0, 0), pix)
        d = np.diff(S, axis=0)
        d /= np.linalg.norm(d, axis=1)[:, None]
        best = np.max(np.abs(d @ DIRS.T), axis=1)
        for o in np.unique(r["IC_o"][r["IC_f"] == kf]):
            assert np.any(np.abs(d @ DIRS[o]) >= best - 0.03)


def test_centroid_mode():
    dim, pix = (10, 9, 8), (1.0, 1.0, 1.0)
    pts, ptr = synth.streamlines(9, 15, dim, 1.0, mode=0, step_vox=1.3)
    fibs = [pts[ptr[i]:ptr[i + 1]] for i in range(len(ptr) - 1)]
    r = vo.trk2dictionary(pts, ptr, dim, pix, H, do_intersect=False, min_seg_len=0.0, n_threads=2)
    for kf, f in enumerate(np.nonzero(r["TRK_kept"])[0]):
        S = _steps(fibs[f])
        mid = np.floor(0.5 * (S[1:] + S[:-1])).astype(np.int64)
        ok = np.all((mid >= 0) & (mid < np.array(dim)), axis=1)
        ln = np.linalg.norm(np.diff(S, axis=0), axis=1)
        exp = np.bincount(lin(mid[ok, 0], mid[ok, 1], mid[ok, 2], dim), weights=ln[ok], minlength=np.prod(dim))
        sel = r["IC_f"] == kf
        got = np.bincount(r["IC_v"][sel], weights=r["IC_len"][sel].astype(np.float64), minlength=np.prod(dim))
        np.testing.assert_allclose(got, exp, atol=1e-5)


def test_mask_grid_bounds_and_shift():
    dim = (8, 3, 3)
    f = [[-2.2, 1.5, 1.5], [3.0, 1.5, 1.5], [10.6, 1.5, 1.5]]
    full = run([f], dim)
    np.testing.assert_array_equal(full["IC_v"], [lin(x, 1, 1, dim) for x in range(8)])
    np.testing.assert_allclose(full["TRK_len"], [8.0], rtol=1e-6)     # only what lies in the grid
    mask = np.ones(dim, np.float32)
    mask[2:5, 1, 1] = 0
    m = run([f], dim, mask=mask)
    np.testing.assert_array_equal(m["IC_v"], [lin(x, 1, 1, dim) for x in (0, 1, 5, 6, 7)])
    np.testing.assert_allclose(m["TRK_len"], [5.0], rtol=1e-6)
    # fiber_shift in voxel units moves every point
    s = run([f], dim, fiber_shift=(0.0, 1.0, -1.0))
    np.testing.assert_array_equal(s["IC_v"], [lin(x, 2, 0, dim) for x in range(8)])


def test_min_seg_len_drops_short_pieces():
    dim = (4, 2, 2)
    f = [[0.8, 0.5, 0.5], [2.0, 0.5, 0.5], [3.3, 0.5, 0.5]]
    r = run([f], dim, min_seg_len=0.25)
    # piece 0.8 -> 1.0 (0.2) is dropped; 3.0 -> 3.3 (0.3) kept
    np.testing.assert_array_equal(r["IC_v"], [1, 2, 3])
    np.testing.assert_allclose(r["IC_len"], [1.0, 1.0, 0.3], atol=1e-6)
    np.testing.assert_allclose(r["TRK_len"], [2.3], rtol=1e-6)


def test_length_filters_renumbering_and_short_streamlines():
    dim = (30, 3, 3)
    mk = lambda L: [[0.5, 1.5, 1.5], [0.5 + L / 2, 1.5, 1.5], [0.5 + L, 1.5, 1.5]]
    fibs = [mk(5.0), mk(15.0), [[0.5, 1.5, 1.5], [9.5, 1.5, 1.5]], mk(25.0), mk(12.0), [[1, 1, 1]]]
    r = run(fibs, dim, min_fiber_len=10.0, max_fiber_len=20.0)
    np.testing.assert_array_equal(r["TRK_kept"], [0, 1, 0, 0, 1, 0])
    np.testing.assert_allclose(r["TRK_len"], [15.0, 12.0], rtol=1e-6)
    assert set(np.unique(r["IC_f"])) == {0, 1}
    assert np.all(np.diff(r["IC_f"].astype(np.int64)) >= 0)
    np.testing.assert_allclose(np.bincount(r["IC_f"], weights=r["IC_len"]), [15.0, 12.0], rtol=1e-5)
    # a two-point streamline records nothing (fiberForwardModel returns for pts <= 2)
    r2 = run([[[0.5, 1.5, 1.5], [9.5, 1.5, 1.5]]], dim)
    assert r2["IC_v"].size == 0 and r2["TRK_kept"][0] == 0


def test_blur_replicas_follow_offset_lines():
    """Replica k starts at R_k * (n rotated by angle_k about the first step's
    direction), n = normalize(d.y - d.z, d.z - d.x, d.x - d.y)
    (trk2dictionary_c.cpp:597-641): for d = +x, n = (0, -1, 1)/sqrt(2), and a
    rotation by +90 degrees about x gives (0, -1, -1)/sqrt(2)."""
    dim = (12, 10, 10)
    f = [[0.5, 5.0, 5.0], [6.0, 5.0, 5.0], [11.5, 5.0, 5.0]]
    R = 0.6
    blur = (np.array([0.0, R, R, R]), np.array([0.0, 0.0, np.pi / 2, 1.0]), np.array([1.0, 0.5, 0.25, 5e-4]))
    r = run([f], dim, blur=blur)
    o = R / np.sqrt(2)
    rows = {(5, 5): 1.0, (int(np.floor(5 - o)), int(np.floor(5 + o))): 0.5,
            (int(np.floor(5 - o)), int(np.floor(5 - o))): 0.25}
    exp = np.zeros(np.prod(dim))
    for (y, z), w in rows.items():
        for x in range(12):
            # x runs over [0.5, 11.5]: half voxels at both ends
            exp[lin(x, y, z, dim)] += w * (0.5 if x in (0, 11) else 1.0)
    got = np.bincount(r["IC_v"], weights=r["IC_len"].astype(np.float64), minlength=np.prod(dim))
    np.testing.assert_allclose(got, exp, atol=1e-5)
    np.testing.assert_allclose(r["TRK_len"], [11.0], rtol=1e-6)
    np.testing.assert_allclose(r["TRK_lenTot"], [11.0 * 1.75], rtol=1e-6)
    # blur_apply_to = 0: the original streamline only
    r0 = run([f], dim, blur=blur, blur_apply_to=np.array([0], np.uint8))
    np.testing.assert_allclose(r0["TRK_lenTot"], [11.0], rtol=1e-6)
    np.testing.assert_array_equal(r0["IC_v"], [lin(x, 5, 5, dim) for x in range(12)])


def test_ec_peaks_threshold_flip_affine():
    dim = (4, 3, 3)
    f = [[0.5, 1.5, 1.5], [1.9, 1.5, 1.5], [3.5, 1.5, 1.5]]      # traverses (x, 1, 1)
    peaks = np.zeros(dim + (9,), np.float32)
    peaks[1, 1, 1, 0:3] = (1, 0, 0)
    peaks[1, 1, 1, 3:6] = (0, 0.05, 0)        # below vf_THR * max: dropped
    peaks[1, 1, 1, 6:9] = (0, -0.5, 0)        # kept, flipped into y >= 0
    peaks[2, 1, 1, 0:3] = (0, 0, 2)
    peaks[0, 0, 0, 0:3] = (1, 0, 0)           # voxel without IC segments: ignored
    r = run([f], dim, peaks=peaks, vf_THR=0.1)
    np.testing.assert_array_equal(r["EC_v"], [lin(1, 1, 1, dim), lin(1, 1, 1, dim), lin(2, 1, 1, dim)])
    np.testing.assert_array_equal(r["EC_o"], [H_X, H_Y, H_Z])
    r = run([f], dim, peaks=peaks, vf_THR=0.1, flip_peaks=(True, False, False))
    np.testing.assert_array_equal(r["EC_o"], [H_MX, H_Y, H_Z])
    rot = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float64)    # x -> y
    r = run([f], dim, peaks=peaks, vf_THR=0.1, peaks_affine=rot)
    np.testing.assert_array_equal(r["EC_o"], [H_Y, H_MX, H_Z])
    r = run([f], dim, peaks=peaks, vf_THR=0.04)
    assert r["EC_v"].size == 4


def test_iso_voxels_and_tdi():
    dim = (5, 4, 3)
    fibs = [[[0.5, 0.5, 0.5], [2.5, 0.6, 0.5], [4.5, 0.7, 0.5]], [[0.5, 3.5, 2.5], [2.5, 2.5, 2.5], [4.5, 1.5, 2.5]],
            [[2.2, 0.2, 0.3], [2.3, 3.8, 0.4], [2.4, 3.9, 2.6]]]
    r = run(fibs, dim)
    tdi = np.bincount(r["IC_v"], weights=r["IC_len"].astype(np.float64), minlength=np.prod(dim))
    np.testing.assert_allclose(r["tdi"], tdi, rtol=1e-6)
    np.testing.assert_array_equal(r["ISO_v"], np.nonzero(tdi > 0)[0])
    iso = np.ones(dim, np.float32)
    iso[0, 0, 0] = 0
    mask = np.ones(dim, np.float32)
    mask[4, 0, 0] = 0
    r2 = run(fibs, dim, iso_mask=iso, mask=mask)
    exp = [v for v in np.nonzero(np.bincount(r2["IC_v"], minlength=np.prod(dim)))[0] if v != lin(0, 0, 0, dim)]
    np.testing.assert_array_equal(r2["ISO_v"], exp)
    assert lin(4, 0, 0, dim) not in set(r2["IC_v"].tolist())


def test_threads_do_not_change_results():
    dim = (16, 16, 12)
    pts, ptr = synth.streamlines(3, 300, dim, 1.0, mode=0, step_vox=0.6)
    a = vo.trk2dictionary(pts, ptr, dim, (1, 1, 1), H, n_threads=1)
    b = vo.trk2dictionary(pts, ptr, dim, (1, 1, 1), H, n_threads=7)
    for k in a:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]))
