"""The reference's text outputs (carpios_amd/output.py): Output_Scalar
(lib/structure/s_io_file.hpp:42-72), the residual file
(s_solver.hpp:193-203), the step line (s_event.hpp:374-384) and the output
file name (s_event.hpp:236-239).  The reference tree holds no file its
structured code wrote, so the expected text below is spelled out from the
fmt specs of those lines (fmt's "{:10.6e}" is printf's "%10.6e"), and the
header is read back the way TextFile::parse_config reads it
(lib/io/io_define.hpp:157-170)."""
import io

import numpy as np

from carpios_amd import output as out


def test_output_scalar_text_and_order(tmp_path):
    cx = np.array([0.1, 0.3, 0.5])
    cy = np.array([0.25, 0.75])
    gl = 2
    f = np.full((3 + 2 * gl, 2 + 2 * gl), 99.0)  # ghosts never written
    f[gl:gl + 3, gl:gl + 2] = [[1.0, -2.0], [0.0, 1e-300], [123456.789, -0.5]]
    name = tmp_path / out.scalar_name("phi", 12, 0.03125)
    assert name.name == "phi_12_3.1250e-02"
    out.output_scalar(name, cx, cy, f)
    text = name.read_text().split("\n")
    assert text[:4] == ["## Size : 6", "## Dim : 2", "## NX :  3", "## NY :  2"]
    # Ijk order: i fastest, then j; z = 0 in 2-D
    assert text[4] == "1.000000e-01, 2.500000e-01, 0.000000e+00, 1.000000e+00"
    assert text[5] == "3.000000e-01, 2.500000e-01, 0.000000e+00, 0.000000e+00"
    assert text[6] == "5.000000e-01, 2.500000e-01, 0.000000e+00, 1.234568e+05"
    assert text[7] == "1.000000e-01, 7.500000e-01, 0.000000e+00, -2.000000e+00"
    assert text[8] == "3.000000e-01, 7.500000e-01, 0.000000e+00, 1.000000e-300"
    assert text[9] == "5.000000e-01, 7.500000e-01, 0.000000e+00, -5.000000e-01"
    assert text[10] == "" and len(text) == 11  # every line ends with "\n"
    cfg, data = out.read_scalar(name)
    assert cfg == {"Size": "6", "Dim": "2", "NX": "3", "NY": "2"}
    assert data.shape == (6, 4)
    assert np.allclose(data[:, 3], f[gl:gl + 3, gl:gl + 2].T.ravel(), rtol=1e-6)


def test_output_scalar_accepts_interior_fields(tmp_path):
    v = np.arange(6.0).reshape(3, 2)
    a = out.output_scalar(tmp_path / "a", [0, 1, 2], [0, 1], v)
    b = out.output_scalar(tmp_path / "b", [0, 1, 2], [0, 1], np.pad(v, 1), gl=1)
    assert (tmp_path / "a").read_text() == (tmp_path / "b").read_text()


def test_residual_file_appends_with_an_empty_line():
    f = io.StringIO()
    f.write("previous solve\n")
    out.output_residual(f, [1.0, 0.5, 2.5e-11])
    assert f.getvalue() == ("previous solve\n"
                            "         0,1.0000000e+00\n"
                            "         1,5.0000000e-01\n"
                            "         2,2.5000000e-11\n"
                            "\n")


def test_step_line():
    assert out.step_line(42, 0.0125, 0.5, 1.25) == \
        "step:       42 t: 1.25000e-02 cpu:  0.50000 wall:  1.25000\n"


def test_reader_on_a_file_the_reference_holds():
    # tests/golden/textfile_pointchain_shui.txt is the data file
    # example/geometry/ear_clipping/shui of the reference tree, written in
    # the same TextFile layout as Output_Scalar ("## key : value" header
    # lines, then comma-separated numbers): the header reads as
    # TextFile::parse_config reads it and every data line parses
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "textfile_pointchain_shui.txt")
    cfg, data = out.read_scalar(path)
    # the file's "Size" header (67) is as its writer left it; it holds 81
    # points, and neither reader cross-checks the two
    assert cfg == {"Dim": "2", "Size": "67", "Type": "PointChain", "Closed": "1"}
    assert data.shape == (81, 2)
    assert tuple(data[0]) == (128.20655244, 815.0344063)
