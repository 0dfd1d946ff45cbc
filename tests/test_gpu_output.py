"""The reference's text output written from device-resident fields: fields
fetched from the device (NS after its first step, where the divergence is
bitwise the oracle's; the Advection_ driver, bitwise the oracle's sweep)
give Output_Scalar files byte for byte equal to the oracle's."""
import numpy as np
import pytest

import carpios_amd as ca
import oracle as O
from carpios_amd import output as out
from test_gpu_advection import pair as adv_pair
from test_gpu_ns import pair as ns_pair

pytestmark = pytest.mark.gpu


def centres(g):
    gl = g.gl
    return g.ax.c[gl:gl + g.nx], g.ay.c[gl:gl + g.ny]


def test_ns_divergence_file_from_the_device(ctx, tmp_path):
    g, ora, gpu = ns_pair(ctx, 33, 29)
    ora.step()
    gpu.step(1)
    cx, cy = centres(g)
    a = out.output_scalar(tmp_path / out.scalar_name("div", 1, gpu.time()[0]), cx, cy,
                          gpu.get_field("div"))
    b = out.output_scalar(tmp_path / "oracle_div", cx, cy, ora.f["div"])
    assert open(a).read() == open(b).read()
    cfg, data = out.read_scalar(a)
    assert cfg["NX"] == "33" and cfg["NY"] == "29" and data.shape == (33 * 29, 4)
    gpu.close()


def test_advection_phi_file_from_the_device(ctx, tmp_path):
    g, ora, dev = adv_pair(ctx, 37, 53, "VanLeer", seed=11)
    ora.step(4)
    dev.step(4)
    cx, cy = centres(g)
    a = out.output_scalar(tmp_path / "phi_dev", cx, cy, dev.get_field("phi"))
    b = out.output_scalar(tmp_path / "phi_ora", cx, cy, ora.f["phi"])
    assert open(a).read() == open(b).read()
    dev.close()
