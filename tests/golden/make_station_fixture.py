"""Writes tests/golden/stations.npz from data files the reference holds:

* the DMS latitude / longitude strings of the 43 project stations, read from
  the reference's station workbook
  (LecturaEstacionesSismologicasProyecto/Localizacion_estacionesSismologicas.xlsx,
  columns Latitud_1 / Longitud_1) with the package's own .xlsx reader;
* the grid positions the reference computed from them (its
  npyFiles/rec_pos_{x,y,z}.npy, dx = 1 km), read with numpy.load (no pickles).

    python tests/golden/make_station_fixture.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from elastic_fm_vp_vs_rho_amd import stations  # noqa: E402

REF = "/root/reference/Elastic3DGPU_MT_VP_VS_RHO_MPI/LecturaEstacionesSismologicasProyecto"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stations.npz")


def main():
    rows = stations.read_xlsx(os.path.join(REF, "Localizacion_estacionesSismologicas.xlsx"))
    head = [str(h).strip() if h is not None else "" for h in rows[0]]
    i, j = head.index("Latitud_1"), head.index("Longitud_1")
    lat = [str(r[i]) for r in rows[1:] if len(r) > j and r[i]]
    lon = [str(r[j]) for r in rows[1:] if len(r) > j and r[j]]
    ref = {c: np.load(os.path.join(REF, "npyFiles", f"rec_pos_{c}.npy")) for c in "xyz"}
    np.savez_compressed(OUT, lat_dms=np.array(lat), lon_dms=np.array(lon),
                        rec_x=ref["x"], rec_y=ref["y"], rec_z=ref["z"], dx=1e3)
    print(f"wrote {OUT}: {len(lat)} stations")


if __name__ == "__main__":
    main()
