"""CPU checks of fix mesh/surface ... generate (dem_mesh_generate, the host
restatement of InputMeshTri::meshgenerator) against geometry fixed by the
mathematics: every node on the generated surface, the triangle count each
face bit adds, and the total area from closed forms (a polygonal cylinder of
n segments has side area n * 2R sin(pi/n) * H and cap area
n * R^2 sin(2 pi/n) / 2); box faces point outward. Bad arguments are refused."""
import math

import numpy as np
import pytest

import pydem


def area(t):
    return 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)


def test_plane():
    t = pydem.mesh_generate("plane", [0.3])
    assert t.shape == (2, 3, 3)
    assert np.all(t[..., 2] == 0.0) and np.abs(t[..., :2]).max() == 0.15
    assert area(t).sum() == pytest.approx(0.09, rel=1e-14)


@pytest.mark.parametrize("kind,dims", [("box", [0.2, 0.1, 0.4]), ("cube", [0.25])])
def test_box_faces(kind, dims):
    sx, sy, sz = (dims * 3)[:3] if kind == "cube" else dims
    half = np.array([sx, sy, sz]) / 2
    t = pydem.mesh_generate(kind, dims, mask=63)
    assert t.shape == (12, 3, 3)
    assert np.allclose(np.abs(t), half, rtol=0, atol=0)  # every node a corner
    assert area(t).sum() == pytest.approx(2 * (sx * sy + sy * sz + sz * sx), rel=1e-14)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    assert np.all(np.einsum("ij,ij->i", n, t.mean(axis=1)) > 0)  # outward
    # each face bit adds the two triangles of its face
    faces = {1: (2, 1), 2: (2, -1), 4: (1, -1), 8: (1, 1), 16: (0, -1), 32: (0, 1)}
    for bit, (ax, sgn) in faces.items():
        f = pydem.mesh_generate(kind, dims, mask=bit)
        assert f.shape == (2, 3, 3)
        assert np.all(f[..., ax] == sgn * half[ax])


@pytest.mark.parametrize("nseg", [3, 16, 64])
def test_cylinder(nseg):
    R, H = 0.15, 0.1
    t = pydem.mesh_generate("cylinder", [R, H], nsegments=nseg, mask=7)
    assert t.shape == (4 * nseg, 3, 3)
    rad = np.hypot(t[..., 0], t[..., 1])
    on_axis = rad < 1e-15
    assert np.allclose(rad[~on_axis], R, rtol=1e-15)
    assert np.all(np.isin(t[..., 2], [-H / 2, H / 2]))
    side = 2 * R * math.sin(math.pi / nseg) * H * nseg
    caps = 2 * nseg * 0.5 * R * R * math.sin(2 * math.pi / nseg)
    assert area(t).sum() == pytest.approx(side + caps, rel=1e-12)
    assert pydem.mesh_generate("cylinder", [R, H], nsegments=nseg, mask=4).shape[0] == 2 * nseg
    assert pydem.mesh_generate("cylinder", [R, H], nsegments=nseg, mask=1).shape[0] == nseg


def test_pipe_and_disk():
    ri, ro, H, n = 0.05, 0.08, 0.2, 24
    t = pydem.mesh_generate("pipe", [ri, ro, H], nsegments=n, mask=63)
    assert t.shape == ((1 + 2 + 1 + 2 + 2 + 2) * n, 3, 3)
    s = math.sin(math.pi / n)
    ring = n * 0.5 * math.sin(2 * math.pi / n)
    expect = (2 * ri * s * H * n + 2 * ro * s * H * n      # sides
              + 2 * ri * ri * ring                          # inner caps (full disks)
              + 2 * (ro * ro - ri * ri) * ring)             # annuli
    assert area(t).sum() == pytest.approx(expect, rel=1e-12)
    d = pydem.mesh_generate("disk", [0.1], nsegments=n)
    assert d.shape == (n, 3, 3) and np.all(d[..., 2] == 0.0)
    assert area(d).sum() == pytest.approx(0.01 * ring, rel=1e-12)


@pytest.mark.parametrize("args", [("sphere", [0.1], 0, 0), ("box", [0.1, 0.0, 0.1], 0, 63),
                                  ("cylinder", [0.1, 0.1], 0, 7), ("box", [0.1, 0.1, 0.1], 0, 0),
                                  ("pipe", [0.1, 0.05, 0.1], 8, 63)])
def test_bad_arguments_are_refused(args):
    with pytest.raises(pydem.DemError):
        pydem.mesh_generate(*args)
