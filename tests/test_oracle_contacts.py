"""Pins of the oracle's platform-box and grid-SDF contact finders
(ora_box_contacts, ora_sdf_contacts): closed forms in the box's own frame
(axis-aligned and rotated, the rotation convention against scipy), a
sampled plane against the plane finder (trilinear interpolation is exact
for it), a sampled sphere against its closed form."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from multilayersolver2024_amd.integrator import (PlaneContactFinder, PlatformContactDetector,
                                                 SdfContactFinder, eul2rotm_xyz)


def test_euler_xyz_matches_scipy():
    deg = np.array([12.0, -35.0, 70.0])
    R = eul2rotm_xyz(np.radians(deg))
    assert np.allclose(R, Rotation.from_euler("XYZ", deg, degrees=True).as_matrix(), atol=1e-15)


@pytest.mark.parametrize("degrees", [(0, 0, 0), (20, -10, 35)])
def test_box_contacts_in_the_box_frame(degrees, oracle_lib):
    rng = np.random.default_rng(4)
    box = PlatformContactDetector(degrees, [1.0, 0.6, 0.3], [0.2, -0.1, 0.4])
    loc = rng.uniform(-0.6, 0.6, (400, 3)) * np.array([1.0, 0.6, 0.3])  # half inside
    x = box.center + loc @ box.axes  # world = center + sum_k local_k axis_k
    c = O.box_contacts(x.reshape(-1), box.center, box.axes, box.halfExtent, 1e-4)
    inside = np.all(np.abs(loc) <= box.halfExtent, axis=1)
    assert np.array_equal(c["vertex"], np.nonzero(inside)[0])
    dist = box.halfExtent - np.abs(loc[inside])
    k = np.argmin(dist, axis=1)
    phi = -dist[np.arange(len(k)), k]
    nrm = np.sign(loc[inside][np.arange(len(k)), k])[:, None] * box.axes[k]
    assert np.allclose(c["normal"].reshape(-1, 3), nrm, atol=1e-12)
    assert np.allclose(c["point"].reshape(-1, 3), x[inside] - nrm * phi[:, None], atol=1e-12)
    # the collider point lies on the face of the minimum
    face = (c["point"].reshape(-1, 3) - box.center) @ box.axes.T
    assert np.allclose(np.abs(face[np.arange(len(k)), k]), box.halfExtent[k], atol=1e-12)
    assert 50 < len(k) < 350


def _grid(fn, lo, hi, h):
    ax = [np.arange(lo[a], hi[a] + h / 2, h) for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return fn(np.stack([X, Y, Z], axis=-1)), np.array(lo, dtype=np.float64)


def test_sdf_of_a_plane_is_the_plane_finder(oracle_lib):
    n = np.array([0.3, -0.2, 0.9])
    n = n / np.linalg.norm(n)
    pos = np.array([0.1, 0.0, 0.2])
    vals, org = _grid(lambda X: X @ n - n @ pos, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 0.05)
    sdf = SdfContactFinder(vals, org, 0.05)
    rng = np.random.default_rng(5)
    x = rng.uniform(-0.45, 0.45, (500, 3))
    x = x[np.abs(x @ n - n @ pos) > 1e-9]
    a = O.sdf_contacts(x.reshape(-1), sdf.values, sdf.dims, sdf.origin, sdf.spacing, 1e-4)
    b = O.plane_contacts(3, x.reshape(-1), n, pos, 1e-4, strict=False)
    assert np.array_equal(a["vertex"], b["vertex"]) and a["K"] > 100
    assert np.allclose(a["normal"], b["normal"], atol=1e-12)
    assert np.allclose(a["point"], b["point"], atol=1e-12)


def test_sdf_of_a_sphere_approximates_its_closed_form(oracle_lib):
    c0, R = np.array([0.05, -0.02, 0.01]), 0.3
    h = 0.01
    vals, org = _grid(lambda X: np.linalg.norm(X - c0, axis=-1) - R, (-0.4, -0.4, -0.4),
                      (0.4, 0.4, 0.4), h)
    sdf = SdfContactFinder(vals, org, h)
    rng = np.random.default_rng(6)
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.2, 0.35, 400)
    x = c0 + d * r[:, None]
    a = O.sdf_contacts(x.reshape(-1), sdf.values, sdf.dims, sdf.origin, sdf.spacing, 1e-4)
    inside = r < R - 2e-3
    outside = r > R + 2e-3
    got = np.zeros(400, bool)
    got[a["vertex"]] = True
    assert np.all(got[inside]) and not np.any(got[outside])
    nr = a["normal"].reshape(-1, 3)
    # the interpolant's gradient is first-order accurate: error ~ h / R,
    # halving with the spacing
    err = np.max(np.linalg.norm(nr - d[a["vertex"]], axis=1))
    assert err < 1.5 * h / R
    vals2, org2 = _grid(lambda X: np.linalg.norm(X - c0, axis=-1) - R, (-0.4, -0.4, -0.4),
                        (0.4, 0.4, 0.4), 2 * h)
    a2 = O.sdf_contacts(x.reshape(-1), np.ascontiguousarray(vals2.ravel(order="F")),
                        vals2.shape, org2, 2 * h, 1e-4)
    common, i1, i2 = np.intersect1d(a["vertex"], a2["vertex"], return_indices=True)
    err2 = np.max(np.linalg.norm(a2["normal"].reshape(-1, 3)[i2] - d[common], axis=1))
    assert 1.4 < err2 / err < 3.0
    phi_true = r[a["vertex"]] - R
    pts = a["point"].reshape(-1, 3)
    assert np.max(np.abs(np.linalg.norm(pts - c0, axis=1) - R)) < 2e-3
    assert np.max(np.abs((x[a["vertex"]] - pts) @ np.eye(3) - nr * phi_true[:, None])) < 2e-3


def test_finders_concatenate_in_order(oracle_lib):
    plane = PlaneContactFinder([0, 0, 1], [0, 0, 0])
    box = PlatformContactDetector([0, 0, 0], [0.4, 0.4, 0.4], [0.5, 0.0, 0.3])
    x = np.array([[0.0, 0.0, -0.1], [0.5, 0.0, 0.35], [1.0, 0.0, -0.2], [0.45, 0.1, 0.2]])
    c = O.finder_contacts(x.reshape(-1), [plane, box])
    assert list(c["vertex"]) == [0, 2, 1, 3]
