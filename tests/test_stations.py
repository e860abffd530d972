"""Real-station geometry ingestion (stations.py) against the reference's own
data: the DMS coordinates of its station workbook turn into exactly the grid
positions the reference saved (npyFiles/rec_pos_{x,y,z}.npy, EST:113-157), and
the .xlsx reader reads a workbook built here cell for cell."""
import io
import os
import zipfile

import numpy as np
import pytest

from elastic_fm_vp_vs_rho_amd import driver as D
from elastic_fm_vp_vs_rho_amd import stations as St

FIX = os.path.join(os.path.dirname(__file__), "golden", "stations.npz")


def test_reference_station_positions():
    f = np.load(FIX)
    lat = [St.parse_dms(s) for s in f["lat_dms"]]
    lon = [St.parse_dms(s) for s in f["lon_dms"]]
    x, y, z = St.grid_positions(lat, lon, float(f["dx"]))
    np.testing.assert_array_equal(x, f["rec_x"])
    np.testing.assert_array_equal(y, f["rec_y"])
    np.testing.assert_array_equal(z, f["rec_z"])
    assert x.dtype == f["rec_x"].dtype and z.dtype == f["rec_z"].dtype


def test_parse_dms():
    assert St.parse_dms('6° 15\' 45.650" N') == pytest.approx(6 + 15 / 60 + 45.65 / 3600)
    assert St.parse_dms('73° 8\' 45.540" W') == pytest.approx(-(73 + 8 / 60 + 45.54 / 3600))
    assert St.parse_dms("1 0 36 S") == pytest.approx(-1.01)
    with pytest.raises(ValueError):
        St.parse_dms("6 15 N")


def test_projection_scale():
    """One degree of latitude is D·π/360 km; one of longitude that times cos(6°)."""
    x, y, _ = St.grid_positions([7.0], [-73.0], 1.0)      # 1 m cells
    assert y[0] == int(12756.320 * np.pi / 360 * 1e3)
    assert x[0] == int(12756.320 * np.pi * np.cos(np.radians(6.0)) / 360 * 1e3)


def test_loader_rescales_and_places_at_depth_2():
    # M:2866-2878 divides by 2e3/dx: halves at dx = 1 km, quarters at 0.5 km (truncating)
    rx, ry, rz = St.load_positions(np.array([94, 155, 3]), np.array([29, 51, 1]), dx=1e3)
    np.testing.assert_array_equal(rx, [47, 77, 1])
    np.testing.assert_array_equal(ry, [14, 25, 0])
    rx, _, _ = St.load_positions(np.array([94, 155, 3]), np.array([29, 51, 1]), dx=500.0)
    np.testing.assert_array_equal(rx, [23, 38, 0])
    assert set(rz) == {2}


def _xlsx(rows):
    """A minimal workbook: row 1 as shared strings, the rest inline strings / numbers."""
    ss = sorted({c for c in rows[0]})
    sheet = ['<worksheet xmlns="http://schemas.openxmlformats.org/spreadsheetml/2006/main"><sheetData>']
    for i, row in enumerate(rows, 1):
        sheet.append(f'<row r="{i}">')
        for j, v in enumerate(row):
            ref = chr(65 + j) + str(i)
            if i == 1:
                sheet.append(f'<c r="{ref}" t="s"><v>{ss.index(v)}</v></c>')
            elif isinstance(v, str):
                sheet.append(f'<c r="{ref}" t="inlineStr"><is><t>{v}</t></is></c>')
            else:
                sheet.append(f'<c r="{ref}"><v>{v}</v></c>')
        sheet.append("</row>")
    sheet.append("</sheetData></worksheet>")
    sst = ['<sst xmlns="http://schemas.openxmlformats.org/spreadsheetml/2006/main">']
    sst += [f"<si><t>{s}</t></si>" for s in ss] + ["</sst>"]
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w") as z:
        z.writestr("xl/sharedStrings.xml", "".join(sst))
        z.writestr("xl/worksheets/sheet1.xml", "".join(sheet))
    return buf.getvalue()


def test_spreadsheet_to_geometry(tmp_path):
    f = np.load(FIX)
    rows = [["Estacion", "Latitud_1", "Longitud_1", "Distancia"]]
    for k in range(5):
        rows.append([f"ST{k}", str(f["lat_dms"][k]), str(f["lon_dms"][k]), 10.5 * k])
    p = tmp_path / "st.xlsx"
    p.write_bytes(_xlsx(rows))
    back = St.read_xlsx(str(p))
    assert back[0] == rows[0] and back[3][3] == 21.0
    x, y, z = St.stations_from_spreadsheet(str(p), 1e3)
    np.testing.assert_array_equal(x, f["rec_x"][:5])
    np.testing.assert_array_equal(y, f["rec_y"][:5])
    # into the driver's Geometry, stations shared by the earthquakes (M:2882-2890)
    g = D.Geometry()
    eq = D.Earthquake(50, 60, 30)
    for s in zip(x, y, z):
        eq.addStation(D.Station(*s))
    g.addSismo(eq)
    gx, gy, gz = g.getStationsEarthquake(0)
    np.testing.assert_array_equal(gx, x)
