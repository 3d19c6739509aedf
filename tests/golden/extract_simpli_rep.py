"""Write the reference's golden regression vectors as small .npy fixtures.

The reference holds one golden file for the simulation path,
``tests/simulation/simpli_rep.h5`` (written by ``tests/simulation/simpli_rep.py``
and checked by ``tests/simulation/simpli_test.py:360-383`` with
``h5diff --relative=0.00001``), plus its input fiber model ``tests/cube.dat``.
h5py is not installed here, so the datasets are exported as raw little-endian
binaries with the HDF5 command line tool ``h5dump -b LE`` and reshaped with the
shapes ``h5dump -H`` reports.

Run once, in a container that has the reference mounted:

    python tests/golden/extract_simpli_rep.py

Outputs (all data, no code): ``tests/golden/simpli_rep/*.npy`` and
``tests/golden/cube.dat`` (a verbatim copy of the reference's fiber data file).
"""

import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = '/root/reference'
H5 = os.path.join(REF, 'tests', 'simulation', 'simpli_rep.h5')
H5DUMP = shutil.which('h5dump') or '/opt/conda/bin/h5dump'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'simpli_rep')

DATASETS = {
    '/tissue/tissue': ('tissue.npy', np.uint16),
    '/tissue/optical_axis': ('optical_axis.npy', np.float32),
    '/tissue/tissue_properties': ('tissue_properties.npy', np.float64),
}
for t in range(5):
    DATASETS[f'/simulation/data/{t}'] = (f'data_{t}.npy', np.float64)
    for m in ('transmittance', 'direction', 'retardation'):
        DATASETS[f'/analysis/epa/{t}/{m}'] = (f'epa_{t}_{m}.npy', np.float64)
    DATASETS[f'/simulation/optic/{t}'] = (f'optic_{t}.npy', np.int64)
for m in ('direction', 'inclination', 'trel'):
    DATASETS[f'/analysis/rofl/{m}'] = (f'rofl_{m}.npy', np.float64)


def shape_of(dset):
    hdr = subprocess.run([H5DUMP, '-H', '-d', dset, H5],
                         check=True,
                         capture_output=True,
                         text=True).stdout
    m = re.search(r'DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)', hdr)
    return tuple(int(s) for s in m.group(1).split(','))


def main():
    if not os.path.exists(H5):
        sys.exit(f'{H5} not found')
    os.makedirs(OUT, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for dset, (name, dtype) in DATASETS.items():
            shape = shape_of(dset)
            raw = os.path.join(tmp, 'raw.bin')
            subprocess.run([H5DUMP, '-d', dset, '-b', 'LE', '-o', raw, H5],
                           check=True,
                           capture_output=True)
            arr = np.fromfile(raw, dtype=dtype).reshape(shape)
            np.save(os.path.join(OUT, name), arr)
            print(dset, arr.shape, arr.dtype)
    shutil.copyfile(os.path.join(REF, 'tests', 'cube.dat'),
                    os.path.join(os.path.dirname(OUT), 'cube.dat'))


if __name__ == '__main__':
    main()
