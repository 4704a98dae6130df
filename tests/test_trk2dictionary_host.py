"""Host side of the GPU voxeliser's flow (commit_amd/trk2dictionary.py): the
TRK and TCK readers and writers, the TCK-to-voxmm matrix
(trk2dictionary.pyx:406-411, trk2dictionary_c.cpp:896-911) and the blur
replicas (trk2dictionary.pyx:210-239).  No GPU needed."""
import numpy as np
import pytest

from commit_amd import synth
from commit_amd import trk2dictionary as t2d


def test_trk_round_trip(tmp_path):
    pts, ptr = synth.streamlines(1, 60, (10, 10, 10), 1.0, mode=0, step_vox=0.7)
    path = str(tmp_path / "a.trk")
    t2d.write_trk(path, pts, ptr, (10, 11, 12), (1.0, 1.5, 2.0))
    f, p, h = t2d.read_trk(path)
    np.testing.assert_array_equal(f, pts)
    np.testing.assert_array_equal(p, ptr)
    assert list(h["dimensions"]) == [10, 11, 12] and h["nb_streamlines"] == 60
    np.testing.assert_array_equal(h["voxel_sizes"], np.float32([1.0, 1.5, 2.0]))


def test_trk_with_scalars_and_properties(tmp_path):
    """Per-point scalars and per-streamline properties are skipped (read_fiberTRK, :873-893)."""
    path = str(tmp_path / "s.trk")
    hdr = bytearray(1000)
    hdr[0:6] = b"TRACK\0"
    np.frombuffer(hdr, "<i2", 3, 6)[:] = (4, 4, 4)
    np.frombuffer(hdr, "<f4", 3, 12)[:] = (1, 1, 1)
    np.frombuffer(hdr, "<i2", 1, 36)[0] = 2
    np.frombuffer(hdr, "<i2", 1, 238)[0] = 1
    np.frombuffer(hdr, "<i4", 1, 988)[0] = 2
    np.frombuffer(hdr, "<i4", 1, 996)[0] = 1000
    body = b""
    fibs = [np.float32([[0, 0, 0], [1, 1, 1], [2, 2, 2]]), np.float32([[3, 3, 3], [2, 1, 0]])]
    for fb in fibs:
        body += np.int32(len(fb)).tobytes()
        for q in fb:
            body += q.tobytes() + np.float32([9, 9]).tobytes()
        body += np.float32([7]).tobytes()
    with open(path, "wb") as fh:
        fh.write(bytes(hdr) + body)
    f, p, _ = t2d.read_trk(path)
    np.testing.assert_array_equal(f, np.concatenate(fibs))
    np.testing.assert_array_equal(p, [0, 3, 5])


def test_trk_rejects_max_fib_len(tmp_path):
    pts = np.zeros((10000, 3), np.float32)
    path = str(tmp_path / "l.trk")
    t2d.write_trk(path, pts, np.array([0, 10000]), (4, 4, 4), (1, 1, 1))
    with pytest.raises(ValueError, match="10000"):
        t2d.read_trk(path)


def test_tck_round_trip_and_voxmm(tmp_path):
    pts, ptr = synth.streamlines(2, 40, (10, 10, 10), 1.0, mode=0, step_vox=0.7)
    pix = (1.0, 2.0, 0.5)
    # world = A (voxmm / pix) with A = diag(pix) plus a translation
    aff = np.array([[1.0, 0, 0, -10.0], [0, 2.0, 0, 4.0], [0, 0, 0.5, 2.5], [0, 0, 0, 1]])
    world = (pts.astype(np.float64) + aff[:3, 3]).astype(np.float32)
    path = str(tmp_path / "a.tck")
    t2d.write_tck(path, world, ptr)
    M = t2d.tck_to_voxmm(aff, pix)
    np.testing.assert_allclose(M.reshape(4, 4)[:3], np.hstack([np.eye(3), -aff[:3, 3:]]), atol=1e-6)
    f, p, h = t2d.read_tck(path, M)
    np.testing.assert_array_equal(p, ptr)
    np.testing.assert_allclose(f, pts, atol=1e-5)
    # the conversion is ((p0 m0 + p1 m1) + p2 m2) + m3 in float32
    r = world[:, 0] * M[0] + world[:, 1] * M[1]
    r = (r + world[:, 2] * M[2]) + M[3]
    np.testing.assert_array_equal(f[:, 0], r.astype(np.float32))


def test_blur_replicas_closed_forms():
    rho, ang, w, sigma = t2d.blur_replicas(0.0, 0.0)
    np.testing.assert_array_equal(rho, [0.0]) and np.testing.assert_array_equal(w, [1.0])
    rho, ang, w, sigma = t2d.blur_replicas(blur_core_extent=0.5, blur_gauss_extent=0.5, blur_spacing=0.25,
                                           blur_gauss_min=0.1)
    # lattice points of spacing 0.25 in the disc of radius 1: i^2 + j^2 <= 16
    assert rho.size == sum(1 for i in range(-4, 5) for j in range(-4, 5) if i * i + j * j <= 16)
    assert rho[0] == 0.0 and ang[0] == 0.0 and w[0] == 1.0
    np.testing.assert_array_equal(w[rho <= 0.5], 1.0)
    # the Gaussian reaches blur_gauss_min at the outer edge
    edge = np.isclose(rho, 1.0)
    assert edge.sum() == 4
    np.testing.assert_allclose(w[edge], 0.1, rtol=1e-6)
    assert np.all(np.diff(w[np.argsort(rho)]) <= 1e-15)
    # core only: every replica weighs 1
    rho, ang, w, _ = t2d.blur_replicas(blur_core_extent=0.5, blur_spacing=0.25)
    assert rho.size == 13 and np.all(w == 1.0)
    with pytest.raises(ValueError):
        t2d.blur_replicas(blur_core_extent=-1)
