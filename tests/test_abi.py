"""C-ABI and host-side logic (CPU only: no kernel is launched here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from robartsvtk_amd import abi, configs

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(rvk_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = abi.lib()
    names = header_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_function_has_a_binding():
    names = set(header_functions())
    assert names <= set(abi.exported_symbols()), sorted(names - set(abi.exported_symbols()))


def test_version_and_status_strings():
    lib = abi.lib()
    assert lib.rvk_version() >= 100
    assert lib.rvk_status_string(0) == b"ok"
    assert lib.rvk_status_string(2) == b"no device"
    assert lib.rvk_status_string(99) == b"unknown status"


def test_no_device_is_reported_not_crashed():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    assert abi.device_count() == 0
    h = C.c_void_p()
    rc = abi.lib().rvk_ctx_create(0, C.byref(h))
    assert rc == 2 and not h.value
    assert b"no HIP device" in abi.lib().rvk_last_error_string()


def test_null_arguments_are_rejected():
    lib = abi.lib()
    h = C.c_void_p()
    dims = (C.c_int32 * 3)(8, 8, 8)
    sp = (C.c_double * 3)(1, 1, 1)
    assert lib.rvk_volume_create(None, dims, sp, 0, 1, C.byref(h)) == 1
    assert lib.rvk_render_drr(None, None, 0, None, None, None, None) == 1
    assert lib.rvk_tf1d_create(None, 4, None, None, None, None, None, None, C.byref(h)) == 1


# ---------------------------------------------------------------------------
# geometry (ComputeMatrices restatement)
def _cam():
    return abi.camera((300.0, -200.0, 150.0), (64.0, 60.0, 70.0), (0.0, 0.0, 1.0), 30.0, (100.0, 700.0))


def test_view_matrices_are_inverse_and_project_focal_point_to_center():
    cam = _cam()
    m, mi = abi.compute_matrices(cam, 1.5, origin=(2.0, 3.0, 4.0), spacing=(0.8, 0.9, 1.1))
    np.testing.assert_allclose(m @ mi, np.eye(4), atol=1e-9)
    # voxel coordinates of the focal point project to NDC (0, 0)
    f = (np.array(cam.focal_point[:]) - np.array([2.0, 3.0, 4.0])) / np.array([0.8, 0.9, 1.1])
    ndc = mi @ np.append(f, 1.0)
    ndc = ndc[:3] / ndc[3]
    assert abs(ndc[0]) < 1e-9 and abs(ndc[1]) < 1e-9 and 0.0 < ndc[2] < 1.0
    # near and far planes map to depth 0 and 1
    d = np.array(cam.focal_point[:]) - np.array(cam.position[:])
    d /= np.linalg.norm(d)
    for dist, depth in ((100.0, 0.0), (700.0, 1.0)):
        w = (np.array(cam.position[:]) + dist * d - np.array([2.0, 3.0, 4.0])) / np.array([0.8, 0.9, 1.1])
        q = mi @ np.append(w, 1.0)
        assert q[2] / q[3] == pytest.approx(depth, abs=1e-9)


def test_prebake_maps_pixels_to_ndc():
    cam = _cam()
    W, H = 64, 48
    view = abi.compute_view(cam, W, H)
    m, _ = abi.compute_matrices(cam, W / H)
    mb = abi.view_matrix(view).astype(np.float64)
    for (x, y) in [(0, 0), (17, 30), (63, 47)]:
        vx, vy = 1.0 - x / W, y / H
        a = mb @ np.array([vx, vy, 0.5, 1.0])
        b = m @ np.array([2.0 * x / W - 1.0, 2.0 * y / H - 1.0, 0.5, 1.0])
        np.testing.assert_allclose(a[:3] / a[3], b[:3] / b[3], rtol=2e-5, atol=2e-4)


def test_pose_matrix_derivatives_match_finite_differences():
    pose = np.array([0.1, -0.2, 0.15, 3.0, -4.0, 5.0])
    c = (10.0, 20.0, 30.0)
    m, dm = abi.pose_matrix(pose, c)
    R = m[:3, :3]
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(m @ np.array([*c, 1.0]), np.array([*c, 1.0]) + np.append(pose[3:], 0), atol=1e-12)
    h = 1e-6
    for i in range(6):
        e = np.zeros(6)
        e[i] = h
        fd = (abi.pose_matrix(pose + e, c)[0] - abi.pose_matrix(pose - e, c)[0]) / (2 * h)
        np.testing.assert_allclose(dm[i], fd, atol=1e-8)


def test_posed_view_derivatives_match_finite_differences():
    cam = configs.drr_camera(np.array([100.0, 110.0, 90.0]), 0.4, 64)
    pose = np.array([0.05, -0.03, 0.02, 2.0, -1.0, 3.0])
    c = (100.0, 110.0, 90.0)
    _, dm = abi.posed_view(cam, pose, c, 64, 64, spacing=(0.8, 0.8, 0.8), with_derivatives=True)

    def baked(p):
        u, _ = abi.pose_matrix(p, c)
        m, _ = abi.compute_matrices(cam, 1.0, spacing=(0.8, 0.8, 0.8), user_matrix=u)
        return np.array([[m[r, 0] * -2.0, m[r, 1] * 2.0, m[r, 2], m[r, 3] + m[r, 0] - m[r, 1]] for r in range(4)])

    h = 1e-5
    for i in range(6):
        e = np.zeros(6)
        e[i] = h
        fd = (baked(pose + e) - baked(pose - e)) / (2 * h)
        np.testing.assert_allclose(dm[i].reshape(4, 4), fd, rtol=1e-4, atol=1e-6 * np.abs(fd).max())


# ---------------------------------------------------------------------------
# synthetic inputs
def test_synthetic_volumes_are_deterministic_and_plausible():
    a = abi.synth_ct_phantom((48, 40, 44), (2.0, 2.0, 2.0), 11)
    b = abi.synth_ct_phantom((48, 40, 44), (2.0, 2.0, 2.0), 11)
    assert np.array_equal(a, b)
    assert a[0, 0, 0] == -1000 and a[-1, -1, -1] == -1000
    tissue = a[(a > -200) & (a < 300)]
    assert abs(tissue.mean() - 40) < 3 and abs(tissue.std() - 20) < 3
    assert ((a >= 400) & (a <= 1500)).any()
    c = abi.synth_ct_phantom((48, 40, 44), (2.0, 2.0, 2.0), 12)
    assert not np.array_equal(a, c)
    blobs = abi.synth_blobs((32, 32, 32), 5)
    assert blobs.dtype == np.uint8 and blobs.max() > 100
    p = abi.synth_blobs_perlin_u16((24, 24, 24), 3, base_period=12.0)
    assert p.dtype == np.uint16 and p.min() == 0 and p.max() == 4095
    u = abi.synth_uniform(1, 1000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.05
