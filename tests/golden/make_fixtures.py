"""Convert the reference's FEniCSx fixture files into compact .npz test data.

Reads (once, in the build container) the ASCII VTU fixtures and .npy values that
jax-fem's own benchmark tests compare against:
  tests/benchmarks/{linear_poisson,linear_elasticity_cube,linear_elasticity_cylinder,
                    hyperelasticity,plasticity}/fenicsx/sol_p0_000000.vtu
  tests/benchmarks/linear_elasticity_cylinder/fenicsx/surface_area.npy
  tests/benchmarks/hyperelasticity/fenicsx/traction.npy
  tests/benchmarks/plasticity/fenicsx/{disps,avg_stresses}.npy
and writes tests/golden/fenicsx_<case>.npz with points (P,3), cells (C,8) (VTK
order), sol (P,vec) and the scalar goldens.  Only data is stored.
Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
OUT = os.path.dirname(os.path.abspath(__file__))
CASES = ['linear_poisson', 'linear_elasticity_cube', 'linear_elasticity_cylinder', 'hyperelasticity', 'plasticity']


def read(path):
    piece = ET.parse(path).getroot().find('.//Piece')
    n = int(piece.get('NumberOfPoints'))
    nc = int(piece.get('NumberOfCells'))
    arr = lambda e, t: np.array(e.text.split(), dtype=t)
    pts = arr(piece.find('Points/DataArray'), np.float64).reshape(n, 3)
    cd = {d.get('Name'): d for d in piece.find('Cells').findall('DataArray')}
    cells = arr(cd['connectivity'], np.int64).reshape(nc, -1)
    sol = None
    for d in piece.find('PointData').findall('DataArray'):
        if d.get('Name') == 'sol':
            k = int(d.get('NumberOfComponents', '1'))
            sol = arr(d, np.float64).reshape(n, k)
    return pts, cells, sol


for case in CASES:
    d = os.path.join(REF, 'tests', 'benchmarks', case, 'fenicsx')
    pts, cells, sol = read(os.path.join(d, 'sol_p0_000000.vtu'))
    extra = {}
    for f in ('surface_area', 'traction', 'disps', 'avg_stresses'):
        fp = os.path.join(d, f + '.npy')
        if os.path.exists(fp):
            extra[f] = np.load(fp, allow_pickle=False)
    np.savez_compressed(os.path.join(OUT, f'fenicsx_{case}.npz'), points=pts, cells=cells, sol=sol, **extra)
    print(case, pts.shape, cells.shape, sol.shape, {k: np.asarray(v).tolist() for k, v in extra.items()})
