"""Small I/O helpers (jax_fem/utils.py): ASCII VTU read/write of nodal solutions.

``read_vtu`` reads the ASCII VTU files the reference's FEniCSx fixtures use
(tests/benchmarks/*/fenicsx/sol_p0_000000.vtu) without meshio; ``save_sol``
writes a nodal solution in the same format (utils.py:13-57 writes float32).
"""
import xml.etree.ElementTree as ET

import numpy as onp


def read_vtu(path):
    """(points (P,3) f64, cells (C,n) i64, point_data dict) from an ASCII VTU file."""
    root = ET.parse(path).getroot()
    piece = root.find('.//Piece')
    npts = int(piece.get('NumberOfPoints'))
    ncells = int(piece.get('NumberOfCells'))

    def arr(elem, dtype):
        return onp.array(elem.text.split(), dtype=dtype)

    pts = arr(piece.find('Points/DataArray'), onp.float64).reshape(npts, -1)
    cell_arrays = {d.get('Name'): d for d in piece.find('Cells').findall('DataArray')}
    conn = arr(cell_arrays['connectivity'], onp.int64)
    offsets = arr(cell_arrays['offsets'], onp.int64)
    nper = offsets[0]
    if not onp.all(onp.diff(onp.concatenate([[0], offsets])) == nper):
        raise ValueError("mixed cell sizes are not supported")
    cells = conn.reshape(ncells, nper)
    pdata = {}
    pd = piece.find('PointData')
    if pd is not None:
        for d in pd.findall('DataArray'):
            ncomp = int(d.get('NumberOfComponents', '1'))
            a = arr(d, onp.float64)
            pdata[d.get('Name')] = a.reshape(npts, ncomp) if ncomp > 1 else a
    return pts, cells, pdata


_VTK_TYPES = {8: 12, 27: 29}


def save_sol(fe, sol, sol_file, cell_infos=None, point_infos=None):
    """Write points, cells and the nodal solution to an ASCII VTU file (float32 data)."""
    import os
    sol = onp.asarray(sol.detach().cpu().numpy() if hasattr(sol, 'detach') else sol)
    os.makedirs(os.path.dirname(sol_file) or '.', exist_ok=True)
    pts, cells = onp.asarray(fe.points), onp.asarray(fe.cells)
    n = cells.shape[1]

    def da(name, a, ncomp, typ='Float32'):
        body = ' '.join(repr(float(x)) if typ.startswith('Float') else str(int(x)) for x in onp.ravel(a))
        return f'<DataArray type="{typ}" Name="{name}" NumberOfComponents="{ncomp}" format="ascii">{body}</DataArray>'

    sol2 = sol.reshape(len(pts), -1).astype(onp.float32)
    parts = ['<?xml version="1.0"?>', '<VTKFile type="UnstructuredGrid" version="1.0">', '<UnstructuredGrid>',
             f'<Piece NumberOfPoints="{len(pts)}" NumberOfCells="{len(cells)}">',
             '<Points>' + da('Points', pts, 3, 'Float64') + '</Points>',
             '<Cells>' + da('connectivity', cells, 1, 'Int64') +
             da('offsets', onp.arange(1, len(cells) + 1) * n, 1, 'Int64') +
             da('types', onp.full(len(cells), _VTK_TYPES.get(n, 12)), 1, 'UInt8') + '</Cells>',
             '<PointData>' + da('sol', sol2, sol2.shape[1]) + ''.join(
                 da(k, onp.asarray(v, onp.float32), onp.asarray(v).reshape(len(pts), -1).shape[1])
                 for k, v in (point_infos or [])) + '</PointData>',
             '<CellData>' + ''.join(
                 da(k, onp.asarray(v, onp.float32), onp.asarray(v).reshape(len(cells), -1).shape[1])
                 for k, v in (cell_infos or [])) + '</CellData>',
             '</Piece>', '</UnstructuredGrid>', '</VTKFile>']
    with open(sol_file, 'w') as f:
        f.write('\n'.join(parts))
