"""Gaussian PLY files: the property layout of GaussianIO.save_as_ply
(GaussianIO.py:9-45, misc_utils.py:41-43) and exact round trips."""
import numpy as np
import pytest
import torch

from word2world_amd import ply
from word2world_amd.scene import load_scene, make_scene, save_scene
from word2world_amd.serve import ScenePool, save_result_ply
from word2world_amd.train import GaussianStore

# the header plyfile writes for a degree-1 scene of two Gaussians
HEADER_DEG1 = """ply
format binary_little_endian 1.0
element vertex 2
property float x
property float y
property float z
property float sh_coefficients_dc_0
property float sh_coefficients_dc_1
property float sh_coefficients_dc_2
property float sh_coefficients_ac_0
property float sh_coefficients_ac_1
property float sh_coefficients_ac_2
property float sh_coefficients_ac_3
property float sh_coefficients_ac_4
property float sh_coefficients_ac_5
property float sh_coefficients_ac_6
property float sh_coefficients_ac_7
property float sh_coefficients_ac_8
property float opacity
property float svec_0
property float svec_1
property float svec_2
property float quaternion_0
property float quaternion_1
property float quaternion_2
property float quaternion_3
end_header
"""


def _labelled(n=2, K=3):
    """Parameters whose values name their position."""
    idx = np.arange(n, dtype=np.float32)[:, None]
    means = 1000 * idx + np.array([1, 2, 3], np.float32)
    sh_dc = (1000 * idx + np.array([11, 12, 13], np.float32))[:, None, :]
    k = np.arange(K, dtype=np.float32)[None, :, None]
    c = np.arange(3, dtype=np.float32)[None, None, :]
    sh_ac = 1000 * idx[:, :, None] + 100 * c + k + 0.5          # [n, K, 3]
    opacity = 1000 * idx[:, 0] + 21
    svec = 1000 * idx + np.array([31, 32, 33], np.float32)
    quat = 1000 * idx + np.array([41, 42, 43, 44], np.float32)
    return means, sh_dc, sh_ac.astype(np.float32), opacity, svec, quat


def test_header_and_record_layout(tmp_path):
    p = str(tmp_path / "g.ply")
    ply.save_ply(p, *_labelled())
    raw = open(p, "rb").read()
    assert raw.startswith(HEADER_DEG1.encode())
    body = np.frombuffer(raw[len(HEADER_DEG1):], "<f4").reshape(2, 23)
    # AC columns are channel-major: column c*K + k holds coefficient k of channel c
    want0 = [1, 2, 3, 11, 12, 13, 0.5, 1.5, 2.5, 100.5, 101.5, 102.5, 200.5, 201.5, 202.5,
             21, 31, 32, 33, 41, 42, 43, 44]
    np.testing.assert_array_equal(body[0], np.array(want0, np.float32))
    np.testing.assert_array_equal(body[1], np.array(want0, np.float32) + 1000)


def test_round_trip_bit_exact(tmp_path):
    rng = np.random.default_rng(3)
    n, K = 257, 15
    arrs = [rng.standard_normal(s).astype(np.float32) for s in
            [(n, 3), (n, 1, 3), (n, K, 3), (n,), (n, 3), (n, 4)]]
    arrs[0][0] = [np.nan, np.inf, -0.0]
    arrs[2][1, 2, 0] = np.float32(1e-45)  # denormal
    p = str(tmp_path / "sub" / "r.ply")  # directory is created
    ply.save_ply(p, *arrs)
    d = ply.load_ply(p)
    for name, a in zip(["means", "sh_dc", "sh_ac", "opacity", "svec", "quaternion"], arrs):
        assert d[name].dtype == np.float32 and d[name].shape == a.shape
        assert d[name].tobytes() == a.tobytes(), name


def test_degree0_and_empty(tmp_path):
    for n in (0, 5):
        a = _labelled(n, 0)
        p = str(tmp_path / f"d0_{n}.ply")
        ply.save_ply(p, *a)
        d = ply.load_ply(p)
        assert d["sh_ac"].shape == (n, 0, 3)
        np.testing.assert_array_equal(d["means"], a[0])
        assert "sh_coefficients_ac_0" not in open(p, "rb").read(2000).decode("latin1")


def _write(path, header, body: bytes):
    with open(path, "wb") as f:
        f.write(header.encode())
        f.write(body)


def test_ascii_and_big_endian(tmp_path):
    props = ["x", "y", "z", "sh_coefficients_dc_0", "sh_coefficients_dc_1", "sh_coefficients_dc_2",
             "sh_coefficients_ac_0", "sh_coefficients_ac_1", "sh_coefficients_ac_2", "opacity",
             "svec_0", "svec_1", "svec_2", "quaternion_0", "quaternion_1", "quaternion_2",
             "quaternion_3"]
    rows = np.arange(34, dtype=np.float32).reshape(2, 17) * 0.25
    head = lambda fmt: "ply\nformat %s 1.0\ncomment made by hand\nelement vertex 2\n%send_header\n" % (
        fmt, "".join(f"property float {q}\n" for q in props))
    pa = str(tmp_path / "a.ply")
    _write(pa, head("ascii"), "\n".join(" ".join(repr(float(v)) for v in r) for r in rows).encode() + b"\n")
    pb = str(tmp_path / "b.ply")
    _write(pb, head("binary_big_endian"), rows.astype(">f4").tobytes())
    for p in (pa, pb):
        d = ply.load_ply(p)
        np.testing.assert_array_equal(d["means"], rows[:, 0:3])
        np.testing.assert_array_equal(d["sh_dc"][:, 0], rows[:, 3:6])
        # one coefficient per channel: channel c is column 6 + c
        np.testing.assert_array_equal(d["sh_ac"][:, 0], rows[:, 6:9])
        np.testing.assert_array_equal(d["opacity"], rows[:, 9])
        np.testing.assert_array_equal(d["svec"], rows[:, 10:13])
        np.testing.assert_array_equal(d["quaternion"], rows[:, 13:17])


def test_malformed(tmp_path):
    p = str(tmp_path / "bad.ply")
    _write(p, "ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\nend_header\n",
           b"\0" * 4)
    with pytest.raises(ValueError, match="lacks"):
        ply.load_ply(p)
    _write(p, "ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty double x\nend_header\n",
           b"\0" * 8)
    with pytest.raises(ValueError, match="expected float"):
        ply.load_ply(p)
    _write(p, "ply\nformat binary_little_endian 1.0\nelement face 1\nproperty list uchar int vi\n"
              "element vertex 1\nproperty float x\nend_header\n", b"\0" * 9)
    with pytest.raises(ValueError, match="before 'vertex'"):
        ply.load_ply(p)
    good = str(tmp_path / "good.ply")
    ply.save_ply(good, *_labelled())
    raw = open(good, "rb").read()
    _write(p, "", raw[:-4])
    with pytest.raises(ValueError, match="shorter"):
        ply.load_ply(p)
    _write(p, "PLY\n", b"")
    with pytest.raises(ValueError):
        ply.load_ply(p)
    with pytest.raises(ValueError, match="inconsistent"):
        a = list(_labelled())
        a[4] = a[4][:1]
        ply.save_ply(p, *a)


def test_scene_store_and_pool(tmp_path):
    sc = make_scene(300, 2, seed=5)
    p = str(tmp_path / "scene.ply")
    save_scene(p, sc)
    back = load_scene(p)
    assert back.sh_degree == 2
    for k, v in sc.params().items():
        assert torch.equal(back.params()[k], v), k
    with pytest.raises(ValueError, match="needs"):
        load_scene(p, sh_degree=3)
    assert load_scene(p, sh_degree=1).sh_degree == 1

    st = GaussianStore(sc.params(), 2, "cpu")
    ps = str(tmp_path / "store.ply")
    st.save_ply(ps)
    st2 = GaussianStore.from_ply(ps, "cpu")
    assert st2.sh_degree == 2 and st2.n == 300 and torch.equal(st2.flat, st.flat)

    pool = ScenePool.from_ply([p, ps], [p], "cpu")
    assert len(pool.scenes) == 2 and torch.equal(pool.params[1]["means"], sc.means)

    # an optimisation result in the server's arena layout (sh rows of M coefficients)
    M = 9
    sh = torch.cat((sc.sh_dc, sc.sh_ac), 1)
    res = {"means": sc.means.reshape(-1), "sh": sh.reshape(-1), "opacity": sc.opacity.reshape(-1),
           "svec": sc.svec.reshape(-1), "quaternion": sc.quaternion.reshape(-1)}
    pr = str(tmp_path / "result.ply")
    save_result_ply(pr, res, M)
    assert open(pr, "rb").read() == open(p, "rb").read()
