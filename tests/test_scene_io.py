"""Scene files (RustGS io/scene_io.rs): the oracle (oracle/scene_io.py)
pinned by the reference's own tests (io/scene_io/tests.rs) and by Rust's
float formatting, then the library (rgs_save_splats, rgs_scene_load, host
code behind the C ABI; no device) against the oracle:

* PLY and .splat files written by the library are byte-identical to the
  oracle's, for SH degrees 0-4 and values that stress the float text (NaN,
  infinities, -0, subnormals, f32::MAX, long fractions);
* files read by the library give the oracle's arrays bit for bit: ASCII
  and binary PLY, properties in any order and with extras, interleaved or
  channel-major f_rest, probability or raw-logit opacity, CRLF and blank
  lines, a second element, and every error the reference raises;
* a multi-threaded write and read (8192 rows) and save -> load round trips
  that return the same bits.
"""
import math
import struct

import numpy as np
import pytest

from oracle import scene_io as OS
from rustscan_amd import scene
from rustscan_amd.scene_io import SceneIoError, SplatMetadata, load_splats, save_splats

F = np.float32


def rgb_to_sh_dc(v):
    return (F(v) - F(0.5)) / F(0.2820948)


# ---------------------------------------------------------------- oracle pins


def test_rust_float_display():
    """Rust's `{}` for f32 (shortest round-trip digits, never an exponent)."""
    cases = [(0.0, "0"), (-0.0, "-0"), (1.0, "1"), (0.1, "0.1"), (1e-7, "0.0000001"), (0.25, "0.25"),
             (1e30, "1000000000000000000000000000000"),
             (3.4028234663852886e38, "340282350000000000000000000000000000000"),
             (-2.3025851, "-2.3025851"), (16777216.0, "16777216"), (float("nan"), "NaN"), (float("inf"), "inf"),
             (float("-inf"), "-inf"), (1e-45, "0." + "0" * 44 + "1")]
    for v, want in cases:
        assert OS.rust_f32(v) == want, (v, OS.rust_f32(v))
    rng = np.random.default_rng(1)
    for v in rng.normal(size=2000).astype(F) * F(10.0) ** rng.integers(-8, 8, 2000).astype(F):
        assert F(float(OS.rust_f32(v))) == v  # round trip


def _host(pos, ls, rot, op, sh):
    return [np.asarray(x, F) for x in (pos, ls, rot, op, sh)]


def test_oracle_reference_ply_roundtrip():
    """scene_io/tests.rs test_save_splats_ply_roundtrip and
    test_save_splats_dispatches_ply_by_extension."""
    pos, ls, rot, op, sh = _host([[0, 0, 1]], [[math.log(F(0.1))] * 3], [[1, 0, 0, 0]], [0.0],
                                 [[rgb_to_sh_dc(0.2), rgb_to_sh_dc(0.3), rgb_to_sh_dc(0.4)]])
    text = OS.ply_text(pos, ls, rot, op, sh, 0, iterations=12, final_loss=0.25)
    assert text.startswith(b"ply\n")
    out, meta = OS.load_ply(text)
    assert len(out["opacity_logits"]) == 1 and meta["iterations"] == 12 and meta["sh_degree"] == 0
    assert abs(out["opacity_logits"][0] - op[0]) < 1e-6


def test_oracle_reference_ply_roundtrip_with_sh():
    """scene_io/tests.rs test_save_splats_ply_roundtrip_with_sh_coeffs."""
    sh = [rgb_to_sh_dc(0.2), rgb_to_sh_dc(0.3), rgb_to_sh_dc(0.4)] + [F(i) * F(0.01) for i in range(45)]
    pos, ls, rot, op, sh = _host([[1, 2, 3]], [[math.log(F(0.2)), math.log(F(0.3)), math.log(F(0.4))]],
                                 [[1, 0, 0, 0]], [0.25], [sh])
    out, meta = OS.load_ply(OS.ply_text(pos, ls, rot, op, sh, 3, iterations=7, final_loss=0.125))
    assert meta["sh_degree"] == 3
    assert abs(out["opacity_logits"][0] - F(0.25)) < 1e-6
    assert np.abs(out["sh_coeffs"][0] - sh[0]).max() < 1e-6


def test_oracle_reference_splat_roundtrip():
    """scene_io/tests.rs test_save_splats_splat_roundtrip."""
    pos = [[0, 0, 1], [1, 2, 3]]
    ls = [[math.log(F(x)) for x in (0.1, 0.2, 0.3)], [math.log(F(x)) for x in (0.4, 0.5, 0.6)]]
    rot = [[1, 0, 0, 0], [0.707, 0, 0.707, 0]]
    op = [0.0, 0.42]
    sh = [[rgb_to_sh_dc(x) for x in (0.2, 0.3, 0.4)], [rgb_to_sh_dc(x) for x in (0.8, 0.7, 0.6)]]
    data = OS.splat_bytes(*_host(pos, ls, rot, op, sh), 0)
    assert len(data) == 64
    out, meta = OS.load_splat(data)
    assert len(out["opacity_logits"]) == 2 and meta["gaussian_count"] == 2 and meta["sh_degree"] == 0
    assert list(out["positions"][1]) == [1.0, 2.0, 3.0]
    scale0 = [OS.expf(x) for x in out["log_scales"][0]]
    assert max(abs(a - b) for a, b in zip(scale0, (0.1, 0.2, 0.3))) < 1e-6
    opacity0 = F(1.0) / (F(1.0) + OS.expf(-out["opacity_logits"][0]))
    assert abs(opacity0 - 0.5019608) < 1e-6
    rgb1 = out["sh_coeffs"][1][0] * F(0.2820948) + F(0.5)
    assert abs(rgb1 - 0.8) <= 1.0 / 255.0 + 1e-6


PROPS_3DGS = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(45)]
              + ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])
PROPS_BRUSH = (["x", "y", "z", "scale_0", "scale_1", "scale_2", "opacity", "rot_0", "rot_1", "rot_2", "rot_3",
                "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(45)])


def ply_file(comments, props, rows, fmt="ascii", newline="\n", types=None, extra_header=()):
    h = ["ply", f"format {'ascii 1.0' if fmt == 'ascii' else 'binary_little_endian 1.0'}"]
    h += [f"comment {c}" for c in comments]
    h.append(f"element vertex {len(rows)}")
    h += [f"property {(types or {}).get(p, 'float')} {p}" for p in props]
    h += list(extra_header)
    h.append("end_header")
    head = newline.join(h) + newline
    if fmt == "ascii":
        body = "".join(" ".join(str(v) for v in r) + newline for r in rows)
        return (head + body).encode()
    return head.encode() + b"".join(struct.pack(f"<{len(r)}f", *r) for r in rows)


def _external_raw_logit_rows():
    row = [0, 0, 1, 0, 0, 0, 0.1, 0.2, 0.3] + [0] * 45 + [2.0, -2.3025851, -2.3025851, -2.3025851, 1, 0, 0, 0]
    return [row]


def test_oracle_reference_external_files():
    """scene_io/tests.rs test_load_external_raw_logit_opacity,
    test_load_external_channel_major_sh_coeffs, test_load_binary_brush_export_ply."""
    f1 = ply_file(["exported elsewhere", "iterations 1", "final_loss 0.0", "sh_degree 0"], PROPS_3DGS,
                  _external_raw_logit_rows())
    out, _ = OS.load_ply(f1)
    assert abs(out["opacity_logits"][0] - 2.0) < 1e-6
    row = [0, 0, 1, 0, 0, 0, 0.1, 0.2, 0.3] + list(range(1, 10)) + [0] * 36 + [2.0, -2.3025851, -2.3025851,
                                                                                 -2.3025851, 1, 0, 0, 0]
    out, meta = OS.load_ply(ply_file(["Exported from Brush", "SH degree: 1"], PROPS_3DGS, [row]))
    assert meta["sh_degree"] == 1
    assert list(out["sh_coeffs"][0]) == [F(x) for x in (0.1, 0.2, 0.3, 1, 4, 7, 2, 5, 8, 3, 6, 9)]
    brow = [1.0, 2.0, 3.0, -math.log(10.0), -1.609438, -1.2039728, 2.0, 1.0, 0.0, 0.0, 0.0, 0.1, 0.2, 0.3] + \
        [float(v) for v in range(1, 46)]
    out, meta = OS.load_ply(ply_file(["Exported from Brush", "SH degree: 3", "SplatRenderMode: default"],
                                     PROPS_BRUSH, [brow], fmt="binary"))
    assert meta["sh_degree"] == 3 and list(out["positions"][0]) == [1.0, 2.0, 3.0]
    assert abs(out["opacity_logits"][0] - 2.0) < 1e-6
    want = [F(0.1), F(0.2), F(0.3)]
    for k in range(15):
        want += [F(k + 1), F(k + 16), F(k + 31)]
    assert list(out["sh_coeffs"][0]) == want


# ---------------------------------------------------------------- library vs oracle


def test_library_reference_external_files(tmp_path):
    """The reference's external-file tests through the library."""
    got, _ = _check_load(tmp_path / "a.ply", ply_file(["exported elsewhere", "iterations 1", "final_loss 0.0",
                                                        "sh_degree 0"], PROPS_3DGS, _external_raw_logit_rows()))
    assert abs(got.raw_opac[0] - 2.0) < 1e-6
    row = [0, 0, 1, 0, 0, 0, 0.1, 0.2, 0.3] + list(range(1, 10)) + [0] * 36 + [2.0, -2.3025851, -2.3025851,
                                                                                 -2.3025851, 1, 0, 0, 0]
    got, meta = _check_load(tmp_path / "b.ply", ply_file(["Exported from Brush", "SH degree: 1"], PROPS_3DGS, [row]))
    assert meta.sh_degree == 1
    assert list(got.sh.reshape(-1)) == [F(x) for x in (0.1, 0.2, 0.3, 1, 4, 7, 2, 5, 8, 3, 6, 9)]
    brow = [1.0, 2.0, 3.0, -math.log(10.0), -1.609438, -1.2039728, 2.0, 1.0, 0.0, 0.0, 0.0, 0.1, 0.2, 0.3] + \
        [float(v) for v in range(1, 46)]
    got, meta = _check_load(tmp_path / "c.ply", ply_file(["Exported from Brush", "SH degree: 3"], PROPS_BRUSH,
                                                          [brow], fmt="binary"))
    assert meta.sh_degree == 3 and list(got.transforms[0, :3]) == [1.0, 2.0, 3.0]


def special_scene(n, deg, seed):
    sc = scene.gaussian_scene(n, seed=seed, sh_degree=deg, rest_sigma=0.3)
    tf, sh, op = sc.transforms.copy(), sc.sh.copy(), sc.raw_opac.copy()
    specials = [np.nan, np.inf, -np.inf, -0.0, 0.0, 1e-45, 3e-39, 3.4028235e38, -1e30, 1.0 / 3.0, 123456.789,
                1e-7, 16777217.0]
    rng = np.random.default_rng(seed)
    for i, v in enumerate(specials):
        r = i % n
        tf[r, rng.integers(0, 10)] = v
        sh[(r + 1) % n, rng.integers(0, sh.shape[1]), rng.integers(0, 3)] = v
        op[(r + 2) % n] = v
    return scene.Splats(tf, sh, op, deg)


def _arrays(sp):
    tf = np.asarray(sp.transforms, F)
    return tf[:, 0:3], tf[:, 7:10], tf[:, 3:7], np.asarray(sp.raw_opac, F), np.asarray(sp.sh, F).reshape(sp.n, -1)


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, F).view(np.uint32), np.asarray(b, F).view(np.uint32))


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_ply_bytes_equal_oracle(tmp_path, deg):
    sp = special_scene(97, deg, 10 + deg)
    p = tmp_path / "s.ply"
    save_splats(p, sp, SplatMetadata(iterations=1234, final_loss=0.0625, gaussian_count=97, sh_degree=deg))
    pos, ls, rot, op, sh = _arrays(sp)
    want = OS.ply_text(pos, ls, rot, op, sh, deg, iterations=1234, final_loss=0.0625)
    assert p.read_bytes() == want


@pytest.mark.parametrize("deg", [0, 3])
def test_splat_bytes_equal_oracle(tmp_path, deg):
    sp = special_scene(61, deg, 20 + deg)
    p = tmp_path / "s.SPLAT"  # the extension is matched case-insensitively
    save_splats(p, sp)
    pos, ls, rot, op, sh = _arrays(sp)
    assert p.read_bytes() == OS.splat_bytes(pos, ls, rot, op, sh, deg)


def _check_load(path, data):
    path.write_bytes(data)
    got, meta = load_splats(path)
    ref, rmeta = OS.load_ply(data) if path.suffix.lower() == ".ply" else OS.load_splat(data)
    assert meta.sh_degree == rmeta["sh_degree"] and meta.iterations == rmeta["iterations"]
    assert meta.gaussian_count == rmeta["gaussian_count"]
    assert _same_bits(meta.final_loss, rmeta["final_loss"])
    pos, ls, rot, op, sh = _arrays(got)
    assert _same_bits(pos, ref["positions"]) and _same_bits(ls, ref["log_scales"])
    assert _same_bits(rot, ref["rotations"]) and _same_bits(op, ref["opacity_logits"])
    assert _same_bits(sh, ref["sh_coeffs"])
    return got, meta


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_save_load_round_trip_is_exact(tmp_path, deg):
    sp = special_scene(53, deg, 30 + deg)
    p = tmp_path / "r.ply"
    save_splats(p, sp, SplatMetadata(5, 0.5, 53, deg))
    got, meta = _check_load(p, p.read_bytes())
    assert meta == SplatMetadata(5, 0.5, 53, deg)
    for a, b in zip(_arrays(got), _arrays(sp)):
        assert _same_bits(a, b)


def test_degree_four_is_written_cut_and_refused_on_load(tmp_path):
    sp = special_scene(5, 4, 40)
    p = tmp_path / "d4.ply"
    save_splats(p, sp)
    with pytest.raises(SceneIoError, match="unsupported sh_degree 4"):
        load_splats(p)


def test_load_variants_equal_oracle(tmp_path):
    rng = np.random.default_rng(7)
    n = 23
    rows3 = rng.normal(size=(n, len(PROPS_3DGS))).astype(F).tolist()
    # shuffled property order with extra properties, interleaved layout, probability opacity
    props = PROPS_3DGS + ["red", "f_rest_x"]
    perm = rng.permutation(len(props))
    sprops = [props[i] for i in perm]
    rows = [[(r + [0.5, 1.0])[i] for i in perm] for r in rows3]
    for r in rows:
        r[sprops.index("opacity")] = float(rng.uniform(0, 1))
    comments = ["sh_coeff_layout interleaved", "opacity_representation probability", "iterations +17",
                "final_loss 1e-3"]
    _check_load(tmp_path / "a.ply", ply_file(comments, sprops, rows))
    _check_load(tmp_path / "b.ply", ply_file(comments, sprops, rows, fmt="binary"))
    # CRLF, blank lines, a second element, degree inferred from 24 rest properties (degree 2)
    props2 = PROPS_3DGS[:9] + [f"f_rest_{i}" for i in range(24)] + PROPS_3DGS[54:]
    rows2 = rng.normal(size=(n, len(props2))).astype(F).tolist()
    data = ply_file([], props2, rows2, newline="\r\n", extra_header=["element face 0", "property list uchar int i"])
    data = data.replace(b"end_header\r\n", b"end_header\r\n\r\n   \r\n")
    _check_load(tmp_path / "c.ply", data)
    # degree 1 declared, 45 rest values present: only the first 9 kept
    _check_load(tmp_path / "d.ply", ply_file(["sh_degree 1"], PROPS_3DGS, rows3))
    # .splat
    _check_load(tmp_path / "e.splat", bytes(rng.integers(0, 256, 32 * 9, dtype=np.uint8)))


BAD = [
    (b"plx\nformat ascii 1.0\nend_header\n", "missing PLY magic header"),
    (b"ply\nformat binary_big_endian 1.0\nend_header\n", "unsupported PLY format binary_big_endian 1.0"),
    (b"ply\nformat ascii 1.0\nelement vertex 0\nproperty list uchar int f\nend_header\n",
     "list vertex properties are not supported"),
    (b"ply\nformat ascii 1.0\nelement vertex 0\nproperty float\nend_header\n", "invalid property declaration"),
    (b"ply\nformat ascii 1.0\nelement vertex 0\nproperty double x\nend_header\n",
     "unsupported vertex property type double"),
    (b"ply\nformat ascii 1.0\nelement vertex 0\nproperty float x\n", "missing end_header"),
]


@pytest.mark.parametrize("data,msg", BAD)
def test_header_errors(tmp_path, data, msg):
    p = tmp_path / "bad.ply"
    p.write_bytes(data)
    with pytest.raises(SceneIoError, match=msg):
        load_splats(p)
    with pytest.raises(OS.PlyError, match=msg):
        OS.load_ply(data)


def test_body_errors(tmp_path):
    good = _external_raw_logit_rows()[0]
    cases = [
        (ply_file([], PROPS_3DGS[:-1], [good[:-1]]), "missing required PLY property rot_3"),
        (ply_file([], PROPS_3DGS, [good[:-1]]), "expected 62 values per vertex in splat PLY, got 61"),
        (ply_file([], PROPS_3DGS, [good]).replace(b" 2.0 ", b" 2.0x "), "invalid float literal"),
        (ply_file([], PROPS_3DGS, [good]).replace(b"element vertex 1", b"element vertex 2"),
         "vertex count mismatch: header 2, parsed 1"),
        (ply_file([], PROPS_3DGS, [good], fmt="binary")[:-4], "binary PLY truncated"),
        (ply_file([], PROPS_3DGS[:9] + [f"f_rest_{i}" for i in range(4)] + PROPS_3DGS[54:],
                  [good[:13] + good[54:]]), "invalid f_rest property count 4"),
        (ply_file([], PROPS_3DGS[:9] + [f"f_rest_{i}" for i in range(6)] + PROPS_3DGS[54:],
                  [good[:15] + good[54:]]), "could not infer SH degree from 6"),
    ]
    for data, msg in cases:
        p = tmp_path / "bad.ply"
        p.write_bytes(data)
        with pytest.raises(SceneIoError, match=msg):
            load_splats(p)
        with pytest.raises(OS.PlyError, match=msg):
            OS.load_ply(data)


def test_extension_and_splat_length_errors(tmp_path):
    sp = special_scene(3, 0, 50)
    with pytest.raises(SceneIoError, match="unsupported splat output extension '.txt'"):
        save_splats(tmp_path / "s.txt", sp)
    with pytest.raises(SceneIoError, match="missing splat output extension"):
        save_splats(tmp_path / "noext", sp)
    (tmp_path / "s.splat").write_bytes(b"\0" * 33)
    with pytest.raises(SceneIoError, match="invalid .splat byte length 33"):
        load_splats(tmp_path / "s.splat")
    with pytest.raises(SceneIoError, match="failed to read scene"):
        load_splats(tmp_path / "missing.ply")


def test_threaded_write_and_read(tmp_path):
    sp = scene.gaussian_scene(8192, seed=60, sh_degree=3)
    p = tmp_path / "big.ply"
    save_splats(p, sp, SplatMetadata(3, 0.75, 8192, 3))
    pos, ls, rot, op, sh = _arrays(sp)
    data = p.read_bytes()
    assert data == OS.ply_text(pos, ls, rot, op, sh, 3, iterations=3, final_loss=0.75)
    got, meta = load_splats(p)
    for a, b in zip(_arrays(got), _arrays(sp)):
        assert _same_bits(a, b)
