"""Device voxeliser (commit_hip_trk2dictionary) against the oracle
(oracle/voxelize_oracle.c, trk2dictionary_c.cpp:111-869): every output
bit-for-bit, on synthetic streamlines with and without voxel intersection, a
mask, length limits, a coordinate shift, anisotropic voxels, blur replicas,
EC peaks and an ISO mask, edge cases, the full-brain tractogram, device-
resident inputs, handle reuse and the file-based flow."""
import os

import numpy as np
import pytest

from commit_amd import synth
from commit_amd import trk2dictionary as t2d
from oracle import voxelize as vo

pytestmark = pytest.mark.gpu

H = synth.hash_table(synth.hemisphere_directions(500))
KEYS = ("IC_f", "IC_v", "IC_o", "IC_len", "TRK_norm", "TRK_len", "TRK_lenTot", "TRK_kept", "EC_v", "EC_o",
        "ISO_v", "tdi")


def same(g, o, what=""):
    for k in KEYS:
        a, b = np.asarray(g[k]), np.asarray(o[k])
        assert a.dtype == b.dtype and a.shape == b.shape, (what, k, a.shape, b.shape)
        bad = np.nonzero(a.view(np.uint8).reshape(a.size, -1) != b.view(np.uint8).reshape(b.size, -1))[0] \
            if a.size else []
        assert len(bad) == 0, (what, k, len(bad), a[bad[:5]] if len(bad) else None, b[bad[:5]] if len(bad) else None)


def both(pts, ptr, dim, pix, **kw):
    print(f"voxelise {len(ptr) - 1} streamlines, grid {dim}, {sorted(kw)}", flush=True)
    g = t2d.voxelize(pts, ptr, dim, pix, hash_table=H, **kw)
    o = vo.trk2dictionary(pts, ptr, dim, pix, H, **kw)
    assert o["status"] == 0
    return g, o


@pytest.fixture(scope="module")
def small():
    dim = (32, 32, 20)
    pts, ptr = synth.streamlines(1001, 10_000, dim, 1.0, mode=0, step_vox=0.4)
    return pts, ptr, dim


@pytest.mark.parametrize("intersect", [True, False])
def test_small_tractogram(small, intersect):
    pts, ptr, dim = small
    g, o = both(pts, ptr, dim, (1.0, 1.0, 1.0), do_intersect=intersect)
    assert g["IC_v"].size > 100_000
    same(g, o)


def test_mask_limits_shift_anisotropic(small):
    pts, ptr, dim = small
    pix = (1.0, 1.25, 0.8)
    p2 = (pts * np.array(pix, np.float32)).astype(np.float32)
    mask = synth.ellipsoid_mask(dim, (14.0, 13.0, 8.0)).astype(np.float32)
    g, o = both(p2, ptr, dim, pix, mask=mask, min_fiber_len=8.0, max_fiber_len=30.0, min_seg_len=0.05,
                fiber_shift=(0.25, -0.5, 0.125))
    assert 0 < g["TRK_kept"].sum() < ptr.size - 1
    same(g, o)


def test_blur_replicas(small):
    pts, ptr, dim = small
    rho, ang, wts, _ = t2d.blur_replicas(blur_core_extent=0.5, blur_gauss_extent=0.5, blur_spacing=0.25)
    assert rho.size > 10
    rng = np.random.default_rng(3)
    n = 2000
    sub_ptr = ptr[:n + 1]
    sub = pts[:sub_ptr[-1]]
    apply = (rng.random(n) < 0.7).astype(np.uint8)
    g, o = both(sub, sub_ptr, dim, (1.0, 1.0, 1.0), blur=(rho, ang, wts), blur_apply_to=apply)
    same(g, o)


@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("weight", [1.0, 0.75])
def test_key_and_value_widths(wide, weight):
    """The device keeps a piece's key in 32 bits when (largest orientation + 1)
    x voxels fits, and its value w * len as a float when the one replica has
    weight 1; each of the four forms against the oracle.  Orientations above
    32000 on a 163,840-voxel grid need 33-bit keys."""
    dim = (64, 64, 40)
    pts, ptr = synth.streamlines(1007, 3000, dim, 1.0, mode=0, step_vox=0.45)
    h = (H.astype(np.int32) + (32000 if wide else 0)).astype(np.int16)
    kbits = int((int(h.max()) + 1) * int(np.prod(dim)) - 1).bit_length()
    assert (kbits > 32) == wide
    kw = {} if weight == 1.0 else {"blur": (np.zeros(1), np.zeros(1), np.array([weight]))}
    print(f"voxelise {len(ptr) - 1} streamlines, grid {dim}, {kbits}-bit keys, weight {weight}", flush=True)
    g = t2d.voxelize(pts, ptr, dim, (1.0, 1.0, 1.0), hash_table=h, **kw)
    o = vo.trk2dictionary(pts, ptr, dim, (1.0, 1.0, 1.0), h, **kw)
    assert o["status"] == 0 and g["IC_v"].size > 10_000
    assert (g["IC_o"].min() >= 32000) == wide
    same(g, o, (wide, weight))


def test_regions_outgrown():
    """Single-replica tractograms are walked once into regions of 4 slots per
    point plus 16; streamlines with steps of tens of voxels outgrow them, and
    the tractogram goes through the counting walk instead: same outputs."""
    dim = (64, 64, 64)
    pts, ptr = synth.streamlines(1011, 500, dim, 1.0, mode=0, step_vox=0.5)
    rng = np.random.default_rng(12)
    longs = [rng.uniform(0.5, 63.5, size=(5, 3)).astype(np.float32) for _ in range(20)]
    pts = np.concatenate([pts] + longs)
    ptr = np.concatenate([ptr, ptr[-1] + 5 * np.arange(1, 21)]).astype(np.int64)
    g, o = both(pts, ptr, dim, (1.0, 1.0, 1.0), max_fiber_len=1000.0)
    assert g["IC_v"].size > 10_000
    same(g, o)


def _helix(npts, dim, turns, seed):
    """A helix of npts points through the grid, steps of about 0.4 voxel."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 1.0, npts)
    c = np.asarray(dim, np.float64) / 2
    r = min(dim[0], dim[1]) * 0.35
    ph = rng.uniform(0, 2 * np.pi)
    pts = np.stack([c[0] + r * np.cos(2 * np.pi * turns * t + ph), c[1] + r * np.sin(2 * np.pi * turns * t + ph),
                    1.0 + (dim[2] - 2.0) * t], axis=1)
    return pts.astype(np.float32)


@pytest.mark.parametrize("longest,turns", [(900, 3.0), (2600, 9.0)])
def test_long_streamlines(longest, turns):
    """Streamlines of hundreds to thousands of pieces: up to 512 pieces the
    fused sort-merge kernel sorts in registers, up to 2,048 in LDS, and a
    tractogram with a longer streamline takes the segmented sorts instead."""
    dim = (64, 64, 48)
    fibs = [_helix(n, dim, tn, k) for k, (n, tn) in enumerate([(longest, turns), (300, 1.0), (60, 0.2)])]
    pts0, ptr0 = synth.streamlines(1013, 200, dim, 1.0, mode=0, step_vox=0.5)
    pts = np.concatenate([pts0] + fibs)
    ptr = np.concatenate([ptr0, ptr0[-1] + np.cumsum([len(f) for f in fibs])]).astype(np.int64)
    g, o = both(pts, ptr, dim, (1.0, 1.0, 1.0), max_fiber_len=5000.0)
    kept_f = np.nonzero(g["TRK_kept"])[0]
    per = np.bincount(g["IC_f"], minlength=kept_f.size)
    assert per.max() > (2048 if longest > 2000 else 512) // 2
    same(g, o, longest)


def test_peaks_iso_flips_affine(small):
    pts, ptr, dim = small
    rng = np.random.default_rng(5)
    peaks = rng.standard_normal(dim + (9,)).astype(np.float32)
    peaks[rng.random(dim) < 0.2] = 0
    peaks[..., 3:6] *= 0.05
    iso = (rng.random(dim) < 0.7).astype(np.float32)
    mask = (rng.random(dim) < 0.9).astype(np.float32)
    th = 0.3
    rot = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    for flips in ((False, False, False), (True, False, True)):
        g, o = both(pts, ptr, dim, (1.0, 1.0, 1.0), peaks=peaks, vf_THR=0.1, flip_peaks=flips,
                    peaks_affine=rot, iso_mask=iso, mask=mask)
        assert g["EC_v"].size > 1000 and g["ISO_v"].size > 1000
        same(g, o, flips)


def test_edge_cases():
    dim = (6, 5, 4)
    fibs = [
        [[1.0, 1.0, 1.0]],                                          # one point
        [[0.5, 0.5, 0.5], [3.5, 0.5, 0.5]],                         # two points: nothing recorded
        [[1.0, 1.0, 1.0], [3.0, 1.0, 1.0], [5.0, 1.0, 1.0]],        # along voxel edges (integer coordinates)
        [[0.0, 0.0, 0.0], [2.0, 2.0, 2.0], [4.0, 4.0, 3.0]],        # through voxel corners
        [[2.5, 2.5, 2.5], [2.5, 2.5, 2.5], [2.6, 2.5, 2.5]],        # a repeated point
        [[-5.0, 2.5, 2.5], [-3.0, 2.5, 2.5], [-1.0, 2.5, 2.5]],     # outside the grid
        [[5.9, 4.9, 3.9], [0.1, 0.1, 0.1], [5.9, 0.1, 3.9]],        # long diagonal steps
    ]
    pts = np.concatenate([np.asarray(f, np.float32) for f in fibs])
    ptr = np.concatenate(([0], np.cumsum([len(f) for f in fibs]))).astype(np.int64)
    for kw in ({}, {"do_intersect": False}, {"min_seg_len": 0.0}):
        g, o = both(pts, ptr, dim, (1.0, 1.0, 1.0), **kw)
        same(g, o, kw)
    # empty tractogram and a tractogram with nothing kept
    g, o = both(np.zeros((0, 3), np.float32), np.zeros(1, np.int64), dim, (1.0, 1.0, 1.0))
    same(g, o)
    g, o = both(pts, ptr, dim, (1.0, 1.0, 1.0), min_fiber_len=1000.0, max_fiber_len=2000.0)
    assert g["IC_v"].size == 0 and g["ISO_v"].size == 0
    same(g, o)


def test_rejects_too_long_streamline_and_bad_offsets():
    dim = (4, 4, 4)
    pts = np.tile(np.array([[0.5, 0.5, 0.5]], np.float32), (10000, 1))
    with pytest.raises(t2d._lib.CommitHipError, match="10000 points"):
        t2d.voxelize(pts, np.array([0, 10000], np.int64), dim, (1, 1, 1), hash_table=H)
    with pytest.raises(t2d._lib.CommitHipError, match="fib_ptr"):
        t2d.voxelize(pts[:10], np.array([0, 4, 11], np.int64), dim, (1, 1, 1), hash_table=H)


def test_handle_reuse_and_device_inputs(small):
    import torch
    pts, ptr, dim = small
    v = t2d.Voxeliser(0)
    try:
        v.run(pts, ptr, dim, (1.0, 1.0, 1.0), H)
        a = v.fetch()
        n = 500
        v.run(pts[:ptr[n]], ptr[:n + 1], (20, 20, 20), (1.0, 1.0, 1.0), H)
        b = v.fetch()
        ob = vo.trk2dictionary(pts[:ptr[n]], ptr[:n + 1], (20, 20, 20), (1.0, 1.0, 1.0), H)
        same(b, ob, "second, smaller run")
        tp = torch.from_numpy(pts).cuda()
        tq = torch.from_numpy(ptr).cuda()
        v.run(tp, tq, dim, (1.0, 1.0, 1.0), H)
        c = v.fetch()
        same(c, a, "device inputs")
    finally:
        v.close()


def test_fullsize_brain_tractogram():
    cfg = synth.WORKLOADS["brain"]
    dim, pix = cfg["dim"], (cfg["pixdim"],) * 3
    mask = synth.ellipsoid_mask(dim, cfg["mask"][1]).astype(np.float32)
    print("generating 1M streamlines", flush=True)
    pts, ptr = synth.streamlines(1002, cfg["nF"], dim, cfg["pixdim"], mask=mask, mode=1, step_vox=0.5)
    g, o = both(pts, ptr, dim, pix, mask=mask, min_fiber_len=20.0, max_fiber_len=250.0)
    print(f"{g['IC_v'].size} segments, {g['TRK_kept'].sum()} kept", flush=True)
    assert g["IC_v"].size > 50_000_000
    same(g, o)


def test_file_flow_trk_and_tck(tmp_path, small):
    from commit_amd import nifti
    pts, ptr, dim = small
    n = 3000
    pts, ptr = pts[:ptr[n]], ptr[:n + 1]
    pix = (1.0, 1.0, 1.0)
    trk = str(tmp_path / "t.trk")
    t2d.write_trk(trk, pts, ptr, dim, pix)
    mask = synth.ellipsoid_mask(dim, (14.0, 13.0, 8.0)).astype(np.float32)
    nifti.save(nifti.Nifti1Image(mask, np.diag([1.0, 1.0, 1.0, 1.0])), str(tmp_path / "mask.nii.gz"))
    out = str(tmp_path / "COMMIT")
    t2d.run(trk, path_out=out, filename_mask=str(tmp_path / "mask.nii.gz"), hash_table=H, min_fiber_len=5.0)
    o = vo.trk2dictionary(pts, ptr, dim, pix, H, mask=mask, min_fiber_len=5.0)
    for key, dt in (("IC_f", "<u4"), ("IC_v", "<u4"), ("IC_o", "<u2"), ("IC_len", "<f4"), ("TRK_norm", "<f4"),
                    ("TRK_len", "<f4"), ("TRK_lenTot", "<f4"), ("TRK_kept", "u1"), ("ISO_v", "<u4")):
        got = np.fromfile(os.path.join(out, f"dictionary_{key}.dict"), dtype=dt)
        np.testing.assert_array_equal(got, o[key].astype(dt), err_msg=key)
    tdi = nifti.load(os.path.join(out, "dictionary_tdi.nii.gz")).data
    np.testing.assert_array_equal(np.asarray(tdi).ravel(order="F"), o["tdi"])
    # the same streamlines as a .tck in world coordinates, with a reference image
    aff = np.array([[1.0, 0, 0, -10.0], [0, 1.0, 0, 4.0], [0, 0, 1.0, 2.5], [0, 0, 0, 1]])
    world = (pts.astype(np.float64) @ aff[:3, :3].T + aff[:3, 3]).astype(np.float32)
    tck = str(tmp_path / "t.tck")
    t2d.write_tck(tck, world, ptr)
    nifti.save(nifti.Nifti1Image(mask, aff, pix), str(tmp_path / "ref.nii.gz"))
    out2 = str(tmp_path / "COMMIT2")
    t2d.run(tck, path_out=out2, TCK_ref_image=str(tmp_path / "ref.nii.gz"), hash_table=H)
    fib2, ptr2, _ = t2d.read_tck(tck, t2d.tck_to_voxmm(aff, pix))
    o2 = vo.trk2dictionary(fib2, ptr2, dim, pix, H)
    got = np.fromfile(os.path.join(out2, "dictionary_IC_len.dict"), dtype="<f4")
    np.testing.assert_array_equal(got, o2["IC_len"])
    # and Evaluation reads the dictionary the GPU wrote
    from commit_amd import core
    raw = {k: np.fromfile(os.path.join(out, f"dictionary_{k}.dict"), dtype=dt) for k, dt in
           (("IC_f", "<u4"), ("IC_v", "<u4"), ("IC_o", "<u2"), ("IC_len", "<f4"), ("TRK_norm", "<f4"))}
    assert raw["IC_f"].max() + 1 == raw["TRK_norm"].size
    assert hasattr(core, "Evaluation")
