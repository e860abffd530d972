"""Real-station geometry ingestion (host-side setup, run once per survey).

Reference: LecturaEstacionesSismologicasProyecto/lecturaEstacionesSismologicasProyecto.py
(EST:113-157) turns the stations' latitude/longitude in degrees-minutes-seconds
into grid indices with an equirectangular projection about (6° N, 74° W):

    x = int((lon - lon_ref)·D·π·cos(lat_ref)/360 / (dx/1 km))
    y = int((lat - lat_ref)·D·π/360 / (dx/1 km)),        z = 1.0

with D = 12756.320 km (the Earth's equatorial diameter) and int() truncating
towards zero (numpy astype(int)), and saves rec_pos_{x,y,z}.npy.  The driver's
commented loader (mainPropElastic3D_MT.py M:2866-2878) reads such files back as
(positions / (2e3/dx)).astype(int) with the receivers at z = 2.  The
coordinates come from the project's station spreadsheet
(Localizacion_estacionesSismologicas.xlsx, columns Latitud_1 / Longitud_1);
read_xlsx reads such a workbook with the standard library alone (no openpyxl).
"""
from __future__ import annotations

import math
import re
import zipfile
import xml.etree.ElementTree as ET

import numpy as np

EARTH_DIAMETER_KM = 12756.320
REF_LAT, REF_LON = 6.0, -74.0
_NS = {"m": "http://schemas.openxmlformats.org/spreadsheetml/2006/main"}


def parse_dms(text):
    """Degrees from a degrees-minutes-seconds string such as 6° 15' 45.650" N
    (N and E positive, S and W negative)."""
    nums = re.findall(r"\d+(?:\.\d+)?", text)
    if len(nums) < 3:
        raise ValueError(f"not a degrees-minutes-seconds coordinate: {text!r}")
    deg = float(nums[0]) + float(nums[1]) / 60.0 + float(nums[2]) / 3600.0
    hemi = text.strip()[-1:].upper()
    if hemi not in "NSEW":
        raise ValueError(f"coordinate without a hemisphere letter: {text!r}")
    return -deg if hemi in "SW" else deg


def grid_positions(lat, lon, dx, ref_lat=REF_LAT, ref_lon=REF_LON, earth_diameter_km=EARTH_DIAMETER_KM):
    """Grid indices (x, y) int64 and z = 1.0 (float64) of stations at latitudes /
    longitudes in degrees, for a cell size dx in metres (EST:139-150)."""
    lat = np.asarray(lat, np.float64)
    lon = np.asarray(lon, np.float64)
    km = dx / 1e3
    x = np.trunc((lon - ref_lon) * earth_diameter_km * math.pi * math.cos(math.radians(ref_lat)) / 360.0 / km)
    y = np.trunc((lat - ref_lat) * earth_diameter_km * math.pi / 360.0 / km)
    return x.astype(np.int64), y.astype(np.int64), np.ones(lat.shape, np.float64)


def load_positions(x, y, dx, file_dx=2e3, z=2):
    """The driver's loader (M:2866-2878): station indices saved on a grid of
    spacing `file_dx` metres, rescaled to spacing dx ((v / (file_dx/dx)).astype(int)),
    all at depth index z.  x, y: arrays or .npy paths."""
    ax = np.load(x) if isinstance(x, str) else np.asarray(x)
    ay = np.load(y) if isinstance(y, str) else np.asarray(y)
    f = file_dx / dx
    rx = (ax / f).astype(int)
    ry = (ay / f).astype(int)
    return rx, ry, np.full(rx.shape, int(z))


def read_xlsx(path, sheet=1):
    """Rows of one worksheet of an .xlsx workbook as lists of cell values
    (str, float or None), shared strings resolved; columns by letter order."""
    with zipfile.ZipFile(path) as z:
        strings = []
        if "xl/sharedStrings.xml" in z.namelist():
            root = ET.fromstring(z.read("xl/sharedStrings.xml"))
            for si in root.findall("m:si", _NS):
                strings.append("".join(t.text or "" for t in si.iter(f"{{{_NS['m']}}}t")))
        root = ET.fromstring(z.read(f"xl/worksheets/sheet{sheet}.xml"))
    rows = []
    for row in root.findall(".//m:row", _NS):
        cells = {}
        for c in row.findall("m:c", _NS):
            col = re.match(r"[A-Z]+", c.get("r")).group(0)
            idx = 0
            for ch in col:
                idx = idx * 26 + (ord(ch) - 64)
            t, v = c.get("t"), c.find("m:v", _NS)
            if t == "s" and v is not None:
                val = strings[int(v.text)]
            elif t == "inlineStr":
                val = "".join(x.text or "" for x in c.iter(f"{{{_NS['m']}}}t"))
            elif v is not None:
                try:
                    val = float(v.text)
                except ValueError:
                    val = v.text
            else:
                val = None
            cells[idx - 1] = val
        n = max(cells) + 1 if cells else 0
        rows.append([cells.get(i) for i in range(n)])
    return rows


def stations_from_spreadsheet(path, dx, lat_col="Latitud_1", lon_col="Longitud_1", **kw):
    """(x, y, z) grid positions of the stations listed in a workbook whose
    header row names a DMS latitude and longitude column (EST:113-157)."""
    rows = read_xlsx(path)
    head = [str(h).strip() if h is not None else "" for h in rows[0]]
    i, j = head.index(lat_col), head.index(lon_col)
    lat, lon = [], []
    for r in rows[1:]:
        if len(r) > max(i, j) and r[i] and r[j]:
            lat.append(parse_dms(str(r[i])))
            lon.append(parse_dms(str(r[j])))
    return grid_positions(lat, lon, dx, **kw)
