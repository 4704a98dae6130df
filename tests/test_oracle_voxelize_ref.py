"""The voxeliser oracle against the reference's own trk2dictionary()
(trk2dictionary_c.cpp, compiled from the reference tree into oracle/_ref by
oracle/Makefile): every dictionary file bitwise equal on the same streamlines,
with and without intersection, a mask, an ISO mask, peaks with flips and an
affine, blur replicas, a shift, anisotropic voxels, length limits and several
threads.  Skipped where the reference build is absent."""
import numpy as np
import pytest

from commit_amd import synth
from oracle import voxelize as vo

pytestmark = pytest.mark.skipif(not vo.reference_available(), reason="reference trk2dictionary not built")

H = synth.hash_table(synth.hemisphere_directions(500))
KEYS = ("IC_f", "IC_v", "IC_o", "IC_len", "TRK_norm", "TRK_len", "TRK_lenTot", "TRK_kept", "EC_v", "EC_o", "ISO_v")


def check(tmp_path, pts, ptr, dim, pix, **kw):
    ref = vo.reference_trk2dictionary(str(tmp_path), pts, ptr, dim, pix, H, **kw)
    kw.pop("n_threads", None)
    orc = vo.trk2dictionary(pts, ptr, dim, pix, H, n_threads=3, **kw)
    assert orc["status"] == 0
    for k in KEYS:
        a, b = ref[k], orc[k]
        assert a.dtype == b.dtype and a.shape == b.shape, (k, a.shape, b.shape)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (k, np.nonzero(a != b)[0][:5])
    np.testing.assert_array_equal(ref["has_ic"], orc["tdi"] > 0)
    return ref


@pytest.fixture(scope="module")
def streamlines():
    dim = (24, 22, 16)
    pts, ptr = synth.streamlines(77, 1500, dim, 1.0, mode=0, step_vox=0.45)
    return pts, ptr, dim


@pytest.mark.parametrize("intersect", [True, False])
def test_plain(tmp_path, streamlines, intersect):
    pts, ptr, dim = streamlines
    r = check(tmp_path, pts, ptr, dim, (1.0, 1.0, 1.0), do_intersect=intersect)
    assert r["IC_v"].size > 10_000


def test_mask_iso_limits_shift_anisotropic_threads(tmp_path, streamlines):
    pts, ptr, dim = streamlines
    pix = (1.25, 1.0, 0.8)
    p2 = (pts * np.array(pix, np.float32)).astype(np.float32)
    rng = np.random.default_rng(4)
    mask = (rng.random(dim) < 0.85).astype(np.float32)
    iso = (rng.random(dim) < 0.6).astype(np.float32)
    r = check(tmp_path, p2, ptr, dim, pix, mask=mask, iso_mask=iso, min_fiber_len=6.0, max_fiber_len=25.0,
              min_seg_len=0.02, fiber_shift=(0.5, -0.25, 0.125), n_threads=4)
    assert 0 < r["TRK_kept"].sum() < ptr.size - 1


def test_peaks(tmp_path, streamlines):
    pts, ptr, dim = streamlines
    rng = np.random.default_rng(6)
    peaks = rng.standard_normal(dim + (9,)).astype(np.float32)
    peaks[..., 6:9] *= 0.08
    th = 0.4
    rot = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    r = check(tmp_path, pts, ptr, dim, (1.0, 1.0, 1.0), peaks=peaks, vf_THR=0.1, flip_peaks=(False, True, True),
              peaks_affine=rot)
    assert r["EC_v"].size > 1000


def test_blur(tmp_path, streamlines):
    pts, ptr, dim = streamlines
    n = 300
    from commit_amd import trk2dictionary as t2d
    rho, ang, wts, _ = t2d.blur_replicas(blur_core_extent=0.5, blur_gauss_extent=0.75, blur_spacing=0.25)
    apply = (np.arange(n) % 3 != 0).astype(np.uint8)
    check(tmp_path, pts[:ptr[n]], ptr[:n + 1], dim, (1.0, 1.0, 1.0), blur=(rho, ang, wts), blur_apply_to=apply,
          n_threads=2)


def test_edge_streamlines(tmp_path):
    dim = (6, 5, 4)
    fibs = [
        [[1.0, 1.0, 1.0]],
        [[0.5, 0.5, 0.5], [3.5, 0.5, 0.5]],
        [[1.0, 1.0, 1.0], [3.0, 1.0, 1.0], [5.0, 1.0, 1.0]],
        [[0.0, 0.0, 0.0], [2.0, 2.0, 2.0], [4.0, 4.0, 3.0]],
        [[2.5, 2.5, 2.5], [2.5, 2.5, 2.5], [2.6, 2.5, 2.5]],
        [[-5.0, 2.5, 2.5], [-3.0, 2.5, 2.5], [-1.0, 2.5, 2.5]],
        [[5.9, 4.9, 3.9], [0.1, 0.1, 0.1], [5.9, 0.1, 3.9]],
    ]
    pts = np.concatenate([np.asarray(f, np.float32) for f in fibs])
    ptr = np.concatenate(([0], np.cumsum([len(f) for f in fibs]))).astype(np.int64)
    for i, kw in enumerate(({}, {"do_intersect": False}, {"min_seg_len": 0.0})):
        check(tmp_path / str(i), pts, ptr, dim, (1.0, 1.0, 1.0), **kw)
