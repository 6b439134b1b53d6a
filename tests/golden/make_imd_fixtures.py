"""Write the IMD golden-curve fixtures used by tests/test_oracle_golden.py.

Source: the data files the reference's multilayer tests compare against
(optika/materials/_tests/_data/{Si,SiO2,SiO2_100A,SiC_Cr}.txt, read by
optika/materials/_tests/test_multilayers.py:178-287).  Each is an IMD
export: wavelength (Angstrom) and normal-incidence transmissivity (Si,
SiO2 stacks) or reflectivity (SiC/Cr).  The rows are subsampled to keep the
fixtures small; values are copied unchanged.

Run from the repository root:  python tests/golden/make_imd_fixtures.py
"""

import pathlib

import numpy as np

SRC = pathlib.Path("/root/reference/optika/materials/_tests/_data")
DST = pathlib.Path(__file__).resolve().parent

FILES = {"Si": 20, "SiO2": 20, "SiO2_100A": 20, "SiC_Cr": 2}

for name, step in FILES.items():
    rows = []
    with open(SRC / f"{name}.txt") as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith(";"):
                continue
            rows.append([float(v) for v in s.split()[:2]])
    data = np.asarray(rows)[::step]
    np.savetxt(DST / f"imd_{name}.csv", data, delimiter=",", fmt="%.10g",
               header=f"wavelength_angstrom,efficiency  (every {step}th row of {name}.txt)")
    print(name, data.shape)
