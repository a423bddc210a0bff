"""Output formats of the reference's post-processing, written from device samples.

* gnuplot rows (output/gnuplot.hpp:39-55 with output_manager.hpp:23's fixed_format(10, 18)):
  one line per grid point, x outermost, each number right-aligned in 18 characters with 10
  decimals;
* VTK ImageData in ASCII (output/vtk.hpp:41-120), the format output_manager<3> writes;
* coefficient dumps "i j value" / "i j k value..." with the stream's default formatting
  (implicit.hpp:121-127), and their reader (tools/error.cpp:16-30, :88-105).
"""
import numpy as np

from . import linspace  # noqa: F401  (re-exported for grid construction)


def _num(v):
    return f"{v:18.10f}"


def gnuplot_2d(xs, ys, vals):
    """Text of gnuplot_printer<2>: vals indexed [y][x] (as Simulation.sample returns)."""
    vals = np.asarray(vals)
    out = []
    for i, x in enumerate(xs):
        for j, y in enumerate(ys):
            out.append(_num(x) + _num(y) + _num(vals[j, i]) + "\n")
    return "".join(out)


def vtk_3d(*fields):
    """Text of output::vtk for one or more fields indexed [z][y][x] (x fastest in the file)."""
    fields = [np.asarray(f) for f in fields]
    nz, ny, nx = fields[0].shape
    extent = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    head = ('<?xml version="1.0"?>\n'
            '<VTKFile type="ImageData" version="0.1">\n'
            f'  <ImageData WholeExtent="{extent}" origin="0 0 0" spacing="1 1 1">\n'
            f'    <Piece Extent="{extent}">\n'
            '      <PointData Scalars="Result">\n'
            f'        <DataArray Name="Result"  type="Float32" format="ascii" '
            f'NumberOfComponents="{len(fields)}">\n')
    flat = [f.ravel() for f in fields]
    rows = ["".join(_num(f[i]) for f in flat) + "\n" for i in range(flat[0].size)]
    tail = ("        </DataArray>\n"
            "      </PointData>\n"
            "    </Piece>\n"
            "  </ImageData>\n"
            "</VTKFile>\n")
    return head + "".join(rows) + tail


def write_coefficients(path, u, shape):
    """implicit.hpp:121-127: one line "i j value" (2D) or "i j k value" (3D) per coefficient,
    first index outermost, values in the stream's default format (6 significant digits)."""
    u = np.asarray(u, dtype=np.float64).reshape(tuple(shape)[::-1])
    with open(path, "w") as fh:
        for idx in np.ndindex(*shape):
            fh.write(" ".join(str(i) for i in idx) + f" {u[idx[::-1]]:g}\n")


def read_coefficients(path, shape):
    """tools/error.cpp:16-30 (read_solution): lines "i j value" (or "i j k value"), entries
    not listed stay 0; returns the coefficients first-index-fastest."""
    shape = tuple(shape)
    u = np.zeros(shape[::-1])
    d = len(shape)
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) < d + 1:
                continue
            idx = tuple(int(v) for v in parts[:d])
            u[idx[::-1]] = float(parts[d])
    return u.ravel()
