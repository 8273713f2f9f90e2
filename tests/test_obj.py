"""OBJ ingest (srg_obj_parse / srg_obj_load), restating Obj::Model::Load
(Viewer/Obj.cpp:374-567). Host code: these run without a GPU. Each rule is
checked on hand-written files whose expected meshes follow from the rule
itself; the cube grid of the 1280x720 workload goes through the writer and
the reader and must come back triangle for triangle, bit for bit."""
import numpy as np
import pytest

import softrast_amd as sr
from softrast_amd import scenes

QUAD = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"


def _pos(mesh):
    return mesh.vertices["pos"]


def test_quad_splits_into_0_1_2_and_0_2_3():
    m = sr.obj_parse(QUAD + "f 1 2 3 4\n").meshes[0]
    assert m.indices.dtype == np.uint16 and m.indices.tolist() == [0, 1, 2, 0, 2, 3]
    assert _pos(m).tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]


def test_only_four_corners_of_a_polygon_are_read():
    m = sr.obj_parse(QUAD + "v 0.5 2 0\nf 1 2 3 4 5\n").meshes[0]
    assert m.indices.tolist() == [0, 1, 2, 0, 2, 3] and len(m.vertices) == 4


def test_zero_index_ends_the_face():
    m = sr.obj_parse(QUAD + "f 1 2 3 0 4\n").meshes[0]
    assert m.indices.tolist() == [0, 1, 2]


def test_faces_of_one_or_two_corners_stay_in_the_index_list():
    # no triangle is formed and nothing is dropped: the corners' vertices are
    # appended as read, so later triangles start mid-triple
    m = sr.obj_parse(QUAD + "f 1 2\nf 3 4 1\nf 2\n").meshes[0]
    assert m.indices.tolist() == [0, 1, 2, 3, 0, 1]
    assert _pos(m).tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]


def test_index_forms_and_missing_attribute_lists():
    text = QUAD + "vt 0.25 0.75\nvt 0.5 0.5\nvn 0 0 1\nvn 0 1 0\n"
    m = sr.obj_parse(text + "f 1/2/1 2//2 3/1\n").meshes[0]
    v = m.vertices
    assert v["uv"].tolist() == [[0.5, 0.5], [0.25, 0.75], [0.25, 0.75]]  # "2//2" has uv index 0 -> element 0
    assert v["nrm"].tolist() == [[0, 0, 1], [0, 1, 0], [0, 0, 1]]       # "3/1" has normal index 0 -> element 0
    bare = sr.obj_parse(QUAD + "f 1 2 3\n").meshes[0]                    # no vt / vn lines at all: zeros
    assert not bare.vertices["uv"].any() and not bare.vertices["nrm"].any()


def test_negative_indices_count_from_the_last_element_read():
    m = sr.obj_parse("v 0 0 0\nv 1 0 0\nv 1 1 0\nf -3 -2 -1\nv 5 5 5\nf -4 -1 -3\n").meshes[0]
    assert _pos(m).tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [5, 5, 5]]
    assert m.indices.tolist() == [0, 1, 2, 0, 3, 1]


def test_one_vertex_per_distinct_triple():
    text = QUAD + "vt 0 0\nvt 1 1\n"
    m = sr.obj_parse(text + "f 1/1 2/1 3/1\nf 1/1 3/1 4/1\nf 1/2 2/1 3/1\n").meshes[0]
    # (1,1) (2,1) (3,1) (4,1) (1,2): five vertices, repeats reuse the first
    assert len(m.vertices) == 5
    assert m.indices.tolist() == [0, 1, 2, 0, 2, 3, 4, 1, 2]


def test_groups_split_meshes_and_share_attribute_lists():
    m = sr.obj_parse(QUAD + "g a\nf 1 2 3\ng b\nf 3 4 1\ng c\n")
    assert len(m.meshes) == 2
    assert m.meshes[1].indices.tolist() == [0, 1, 2]
    assert _pos(m.meshes[1]).tolist() == [[1, 1, 0], [0, 1, 0], [0, 0, 0]]


def test_flip_winding_and_flip_uvs():
    text = QUAD + "vt 0.25 0.125\nf 1/1 2/1 3/1 4/1\n"
    m = sr.obj_parse(text, sr.OBJ_FLIP_WINDING | sr.OBJ_FLIP_UVS).meshes[0]
    assert m.indices.tolist() == [0, 2, 1, 0, 3, 2]
    assert m.vertices["uv"][0].tolist() == [0.25, 0.875]


@pytest.mark.parametrize("nverts,dtype", [(65535, np.uint16), (65536, np.uint32)])
def test_index_type_follows_vertex_count(nverts, dtype):
    lines = [f"v {k} 0 0" for k in range(nverts)]
    lines += [f"f {k + 1} {k + 2} {k + 3}" for k in range(0, nverts - 2, 3)]
    lines.append(f"f {nverts} 1 2")
    m = sr.obj_parse("\n".join(lines) + "\n").meshes[0]
    assert len(m.vertices) == nverts and m.indices.dtype == dtype
    assert m.indices[-3:].tolist() == [nverts - 1, 0, 1]


def test_lines_are_read_in_fgets_pieces():
    """A line longer than 2,047 bytes arrives in pieces, as fgets(buf, 2048)
    reads it: the blanks after a vertex are a separate, empty line."""
    m = sr.obj_parse("v 1 2 3" + " " * 3000 + "\nv 4 5 6\nv 7 8 9\nf 1 2 3\n").meshes[0]
    assert _pos(m).tolist() == [[1, 2, 3], [4, 5, 6], [7, 8, 9]]


def test_crlf_and_indented_lines():
    m = sr.obj_parse("  v 0 0 0\r\n\tv 1 0 0\r\nv 0 1 0 \r\n  f 1 2 3  \r\n").meshes[0]
    assert m.indices.tolist() == [0, 1, 2] and _pos(m)[2].tolist() == [0, 1, 0]


@pytest.mark.parametrize("text", ["v 1 2\nf 1 1 1\n", "vt 1\n", "vn 0 1\n", QUAD + "f 1 2 9\n",
                                  QUAD + "vt 0 0\nf 1/3 2/1 3/1\n", QUAD + "vn 0 0 1\nf 1//2 2//1 3//1\n",
                                  QUAD + "f -9 1 2\n"])
def test_malformed_files_are_errors(text):
    with pytest.raises(sr.SrgError):
        sr.obj_parse(text)


def test_materials_and_files(tmp_path):
    (tmp_path / "mats.mtl").write_text("# comment\nnewmtl first\n  Kd 1 1 1\nnewmtl second\n  map_Kd tex/b.png\n")
    (tmp_path / "m.obj").write_text("mtllib mats.mtl\n" + QUAD + "usemtl second\nf 1 2 3\ng\nusemtl first\n"
                                    "f 1 3 4\nusemtl missing\ng\nf 2 3 4\n")
    m = sr.obj_load(tmp_path / "m.obj", sr.OBJ_FLIP_WINDING)
    assert m.materials == [("first", ""), ("second", str(tmp_path) + "/tex/b.png")]
    assert [x.material for x in m.meshes] == [1, 0, 0]  # an unknown name keeps the current material
    assert m.meshes[0].indices.tolist() == [0, 2, 1]
    with pytest.raises(sr.SrgError):
        sr.obj_load(tmp_path / "absent.obj")


def test_cube_grid_round_trips_through_obj():
    """The 1280x720 cube grid, written by scenes.obj_text and read back with
    FlipWinding: the same triangles in the same order with the same vertex
    bytes, u16 indices (24,576 vertices), one vertex per distinct triple."""
    v, i = scenes.cube_grid_mesh()
    m = sr.obj_parse(scenes.obj_text(v, i), sr.OBJ_FLIP_WINDING).meshes[0]
    assert m.indices.dtype == np.uint16 and len(m.vertices) == len(v) == 24576
    assert v[i.reshape(-1, 3)].tobytes() == m.vertices[m.indices.reshape(-1, 3)].tobytes()
