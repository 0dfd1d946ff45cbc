"""The reference's text outputs for device-resident fields.

Host-side formatting only: a run keeps its fields in HBM and fetches them
(NS.get_field, Poisson.get_phi) on output events; these functions then write
the same files the reference writes, so existing post-processing reads them
unchanged.

  output_scalar   Output_Scalar        (lib/structure/s_io_file.hpp:42-72)
  scalar_name     EventOutputScalar_'s file name (s_event.hpp:236-239)
  output_residual Solver_::output_residual (s_solver.hpp:193-203)
  step_line       EventOutputTime_'s line (s_event.hpp:374-384)
  read_config     TextFile::parse_config's "## key : value" header lines
                  (lib/io/io_define.hpp:157-170)

Formats follow the reference's fmt specs, which Python's format mini-language
shares ("{:10.6e}", "{:10d}", "{:>8.5f}").
"""
import numpy as np


def _interior(field, nx, ny, gl):
    f = np.asarray(field, dtype=np.float64)
    if f.shape == (nx, ny):
        return f
    if f.shape == (nx + 2 * gl, ny + 2 * gl):
        return f[gl:gl + nx, gl:gl + ny]
    raise ValueError(f"field of shape {f.shape} is neither ({nx}, {ny}) nor with {gl} ghost layers")


def output_scalar(filename, cx, cy, field, gl=None):
    """Output_Scalar<2>: a header of "## key : value" lines, then one line per
    cell in the reference's Ijk order (i fastest): the three centre
    coordinates (z = 0 in 2-D) and the value, each "{:10.6e}", separated by
    ", ".  cx, cy are the interior cell centres; field is (nx, ny) or with
    gl ghost layers on each side (j contiguous, as the device stores it)."""
    cx = np.asarray(cx, dtype=np.float64)
    cy = np.asarray(cy, dtype=np.float64)
    nx, ny = len(cx), len(cy)
    if gl is None:
        gl = (np.shape(field)[0] - nx) // 2
    v = _interior(field, nx, ny, gl)
    lines = [f"## Size : {nx * ny:d}", "## Dim : 2", f"## NX :  {nx:d}", f"## NY :  {ny:d}"]
    for j in range(ny):
        yj = format(cy[j], "10.6e")
        for i in range(nx):
            lines.append(f"{cx[i]:10.6e}, {yj}, {0.0:10.6e}, {v[i, j]:10.6e}")
    with open(filename, "w") as f:
        f.write("\n".join(lines) + "\n")
    return filename


def scalar_name(fname, step, t):
    """EventOutputScalar_'s file name: "{:s}_{:d}_{:.4e}"."""
    return f"{fname}_{step:d}_{t:.4e}"


def output_residual(f, residuals):
    """Solver_::output_residual: appends "{:10d},{:10.7e}" per iteration and
    an empty line to an open text file (the reference seeks to its end)."""
    f.seek(0, 2)
    for n, v in enumerate(residuals):
        f.write(f"{n:10d},{v:10.7e}\n")
    f.write("\n")


def step_line(step, t, cpu, wall):
    """EventOutputTime_'s line: step, time, and the last interval's cpu and
    wall seconds."""
    return f"step: {step:>8d} t: {t:>10.5e} cpu: {cpu:>8.5f} wall: {wall:>8.5f}\n"


def read_config(filename):
    """The "## key : value" header lines of a text output, as
    TextFile::parse_config reads them (whitespace tokens; token 1 is the key,
    token 3 the value)."""
    out = {}
    with open(filename) as f:
        for line in f:
            tok = line.split()
            if tok and tok[0] == "##" and len(tok) >= 4:
                out[tok[1]] = tok[3]
    return out


def read_scalar(filename):
    """Reads an Output_Scalar file back: (header dict, (n, 4) array of x, y,
    z, value)."""
    cfg = read_config(filename)
    rows = []
    with open(filename) as f:
        for line in f:
            if line.startswith("##") or not line.strip():
                continue
            rows.append([float(t) for t in line.split(",")])
    return cfg, np.array(rows)
