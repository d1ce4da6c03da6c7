"""The scene file readers (host/io.cpp) on inputs the golden fixtures do not exercise: JSON escapes, unicode, deep
nesting, number forms and error offsets (through the camera file); OBJ corner forms v, v/t, v//n, v/t/n, negative
indices, polygon fans, usemtl with unknown names, indented MTL keys; ASCII PLY with an extra element. The expected
values follow the file formats (README.md "Input formats" of the reference, src/io/json.cpp, src/io/mesh_loader.cpp)."""
import json
import os

import numpy as np
import pytest

import cwp

IDENT = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 5, 0, 0, 0, 1]
C2W = ",".join(str(x) for x in IDENT)


def _setup_with_camera(tmp_path, text, golden_dir, extra=()):
    cam = tmp_path / "cam.json"
    cam.write_text(text)
    return cwp.Setup(["--mesh", os.path.join(golden_dir, "boxes.obj"), "--camera", str(cam), "--width", "64",
                      "--height", "48"] + list(extra))


def test_json_escapes_unicode_nesting_and_numbers(tmp_path, golden_dir):
    doc = ('{"note": "tab\\t quote\\" slash\\/ back\\\\ \\u00e9\\u4e2d", "deep": [[[[[[1]]]]]], "e": -1.5E+1, '
           '"arr": [], "obj": {}, "t": true, "f": false, "n": null,\n\r\t "fx": 5, "width": 1, "height": 1, '
           '"hfov_deg": 3.0e1, '
           '"c2w_blender_zup": [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 5], [0, 0, 0, 1]]}')
    json.loads(doc)   # the document is valid JSON
    s = _setup_with_camera(tmp_path, doc, golden_dir)
    p = s.params()
    # fx belongs to another resolution, so the file's hfov (30 degrees, written 3.0e1) sets the focal length
    assert p.fx == pytest.approx(32 / np.tan(np.radians(15.0)), rel=1e-6)
    assert p.pos[:] == pytest.approx([0, 0, 5])


def test_json_matrix_found_below_other_keys_and_frames(tmp_path, golden_dir):
    frames = [IDENT, [1, 0, 0, 2, 0, 1, 0, 0, 0, 0, 1, 5, 0, 0, 0, 1]]
    doc = json.dumps({"meta": {"hfov_deg": 40}, "anim": {"c2w_blender_zup": frames}})
    cam = tmp_path / "c.json"
    cam.write_text(doc)
    s = cwp.Setup(["--mesh", os.path.join(golden_dir, "boxes.obj"), "--camera", str(cam), "--width", "64", "--height", "48"])
    assert s.params().pos[:] == pytest.approx([0, 0, 5])


@pytest.mark.parametrize("text,where", [('{"a": [1, 2,, 3]}', "offset 12"), ('{"a" 1}', "offset 5"),
                                        ('{"a": tru}', "offset 6"), ('{"a": "x}', "unterminated"),
                                        ('{"a": 1} x', "trailing"), ('[1, 2', "unexpected end")])
def test_json_errors_name_the_offset(tmp_path, golden_dir, text, where):
    with pytest.raises(cwp.CwpError, match=where):
        _setup_with_camera(tmp_path, text, golden_dir)


def test_json_nesting_limit(tmp_path, golden_dir):
    with pytest.raises(cwp.CwpError, match="nesting too deep"):
        _setup_with_camera(tmp_path, "[" * 300 + "]" * 300, golden_dir)


def test_obj_corner_forms_fans_and_materials(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl red\n  Kd 0.9 0.1 0.1\nnewmtl blue\nKd 0.1 0.1 0.9\n")
    (tmp_path / "m.obj").write_text(
        "mtllib m.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0\n"
        "vt 0 0\nvn 0 0 1\nvn 0 0 -1\n"
        "usemtl blue\n"
        "f 1/1/1 2/1/1 3/1/1 4/1/1 5/1/1\n"      # pentagon: fanned into 3 triangles
        "usemtl nosuch\n"
        "f -5//2 -4//2 -3//2\n"                   # negative indices, v//n
        "usemtl red\n"
        "f 1/1 3/1 4/1\n"                         # v/t: no normal
        "\tf 2 3 4\n")                            # indented plain corners
    s = cwp.Setup(["--mesh", str(tmp_path / "m.obj"), "--hfov", "40", "--c2w", C2W])
    tri = s.array(cwp.ARR_TRIANGLES)
    np.testing.assert_array_equal(tri, [[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 1, 2], [0, 2, 3], [1, 2, 3]])
    np.testing.assert_array_equal(s.array(cwp.ARR_MATERIAL), [2, 2, 2, 0, 1, 1])
    kd = s.array(cwp.ARR_ALBEDO)
    np.testing.assert_allclose(kd[1, :3], [0.9, 0.1, 0.1], rtol=1e-6)   # indented Kd is read
    np.testing.assert_allclose(kd[2, :3], [0.1, 0.1, 0.9], rtol=1e-6)


def test_obj_bad_index_is_rejected(tmp_path):
    (tmp_path / "b.obj").write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 2 9\n")
    with pytest.raises(cwp.CwpError, match="failed to load"):
        cwp.Setup(["--mesh", str(tmp_path / "b.obj"), "--hfov", "40", "--c2w", C2W])


def test_ply_ascii_with_extra_element(tmp_path):
    (tmp_path / "a.ply").write_text(
        "ply\nformat ascii 1.0\ncomment x\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n"
        "element face 2\nproperty list uchar int vertex_indices\nproperty int material_index\n"
        "element edge 1\nproperty int vertex1\nproperty int vertex2\nend_header\n"
        "0 0 0\n1 0 0\n1 1 0\n0 1 0\n3 0 1 2 1\n3 0 2 3 2\n0 1\n")
    s = cwp.Setup(["--mesh", str(tmp_path / "a.ply"), "--hfov", "40", "--c2w", C2W])
    np.testing.assert_array_equal(s.array(cwp.ARR_TRIANGLES), [[0, 1, 2], [0, 2, 3]])
    np.testing.assert_array_equal(s.array(cwp.ARR_MATERIAL), [1, 2])
