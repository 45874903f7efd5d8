"""Readers for the reference's project input files (the hydrology part of the formats).

``read_project(dir, name, lc_table)`` reads ``<name>.mesh``, ``.att``, ``.soil``, ``.riv``,
``.calib`` and ``.para`` plus the land-cover table and returns a ``Mesh`` whose parameters
are derived as the reference's InitSoil / InitLc / InitRiver derive them:

* soil (src/init_soil.c:24-66): every table value times its calibration multiplier;
  smcmin and smcmax are each scaled by POROSITY before porosity = smcmax - smcmin;
  kmacv = KMACSATV * KMACV_RO * kinfv(table), kmach = KMACSATH * KMACH_RO * ksath(table);
  dmac is capped at the soil depth by ``Mesh.init``;
* land cover (src/init_lc.c:22-37): rzd = DROOT * table DROOT, rough = ROUGH * table ROUGH;
* rivers (src/init_river.c:62-89): depth = RIV_DPTH * shape depth, coeff = RIV_WDTH * shape
  CWID, order = shape OINT, rough = ROUGH_RIV * material ROUGH, ksath = KRIVH * KH,
  ksatv = KRIVV * KV, bedthick = BEDTHCK * BEDTHCK, cwr = CWR.

File layouts follow the reference readers (src/read_mesh.c, src/read_att.c,
src/read_soil.c, src/read_river.c, src/read_calib.c, src/read_lc.c, src/read_para.c).
Only what the right-hand side uses is read.
"""
import calendar
import os
import time

import numpy as np

from .mesh import DIFF_WAVE, Mesh


def _lines(path):
    out = []
    with open(path) as f:
        for raw in f:
            s = raw.split("#", 1)[0].strip()
            if s:
                out.append(s.split())
    return out


def _keyword(rows, key, cast=float):
    for r in rows:
        if r[0] == key:
            return cast(r[1])
    raise KeyError(key)


def read_mesh(path):
    rows = _lines(path)
    ne = int(rows[0][1])
    body = rows[2:2 + ne]
    tri = np.array([[int(v) for v in r[1:4]] for r in body], np.int32)
    nabr = np.array([[int(v) for v in r[4:7]] for r in body], np.int32)
    k = 2 + ne
    nn = int(rows[k][1])
    nodes = np.array([[float(v) for v in r[1:5]] for r in rows[k + 2:k + 2 + nn]])
    return tri, nabr, nodes


def read_att(path, ne):
    rows = _lines(path)[1:1 + ne]
    a = np.array([[int(v) for v in r[:10]] for r in rows], np.int64)
    return dict(soil=a[:, 1], lc=a[:, 3], bc=a[:, 7:10].astype(np.int32))


def read_soil(path):
    rows = _lines(path)
    n = int(rows[0][1])
    t = np.array([[float(v) for v in r[1:16]] for r in rows[2:2 + n]])
    cols = ["silt", "clay", "om", "bd", "kinf", "ksatv", "ksath", "maxsmc", "minsmc", "alpha", "beta",
            "machf", "macvf", "dmac", "qtz"]
    tab = {c: t[:, i] for i, c in enumerate(cols)}
    rest = rows[2 + n:]
    tab["dinf"] = _keyword(rest, "DINF")
    tab["kmacv_ro"] = _keyword(rest, "KMACV_RO")
    tab["kmach_ro"] = _keyword(rest, "KMACH_RO")
    if np.any(t[:, 4:11] < 0.0):
        raise ValueError("pedotransfer-function soil entries (-999) are not supported")
    return tab


def read_lc(path):
    rows = _lines(path)
    n = int(rows[0][1])
    t = np.array([[float(v) for v in r[1:16]] for r in rows[2:2 + n]])
    return dict(rzd=t[:, 1], rough=t[:, 14])


def read_calib(path):
    rows = _lines(path)
    keys = ["KSATH", "KSATV", "KINF", "KMACSATH", "KMACSATV", "DINF", "DROOT", "DMAC", "POROSITY", "ALPHA",
            "BETA", "MACVF", "MACHF", "ROUGH", "ROUGH_RIV", "KRIVH", "KRIVV", "BEDTHCK", "RIV_DPTH", "RIV_WDTH"]
    return {k: _keyword(rows, k) for k in keys}


def read_riv(path):
    rows = _lines(path)
    nr = int(rows[0][1])
    body = np.array([[int(v) for v in r[:10]] for r in rows[2:2 + nr]], np.int64)
    k = 2 + nr
    nshp = int(rows[k][1])
    shp = np.array([[float(v) for v in r[1:4]] for r in rows[k + 2:k + 2 + nshp]])
    k += 2 + nshp
    nmat = int(rows[k][1])
    mat = np.array([[float(v) for v in r[1:6]] for r in rows[k + 2:k + 2 + nmat]])
    k += 2 + nmat
    nbc = int(rows[k][1])
    series = None
    if nbc > 0:
        # RIV_TS blocks up to RES or the end (src/read_river.c:118-175)
        lines = _data_lines(path)
        at = next(i for i, ln in enumerate(lines) if ln.split()[0] == "BC")
        blocks, _ = _ts_blocks(lines, at + 1, "RIV_TS", 1, stop=("RES",))
        if len(blocks) != nbc:
            raise ValueError("river boundary series count does not match BC")
        _equal_length(blocks, "river boundary series")
        series = (np.stack([b[1] for b in blocks]), np.stack([b[2][:, 0] for b in blocks]))
    return dict(bc_series=series, fromnode=body[:, 1], tonode=body[:, 2], down=body[:, 3], left=body[:, 4], right=body[:, 5],
                shape=body[:, 6], matl=body[:, 7], bc=body[:, 8], shp=shp, mat=mat)


def read_para(path):
    rows = _lines(path)
    return dict(surf_mode=_keyword(rows, "SURF_MODE", int), riv_mode=_keyword(rows, "RIV_MODE", int),
                stepsize=_keyword(rows, "MODEL_STEPSIZE", float), reltol=_keyword(rows, "RELTOL"),
                abstol=_keyword(rows, "ABSTOL"), initstep=_keyword(rows, "INIT_SOLVER_STEP"),
                lsm_step=_keyword(rows, "LSM_STEP", float))


def str_time(s):
    """"YYYY-MM-DD HH:MM" (UTC) to integral seconds since the epoch (StrTime,
    src/time_func.c:30-58, which calls timegm)."""
    return float(calendar.timegm(time.strptime(s, "%Y-%m-%d %H:%M")))


def _ts_blocks(lines, i, key, nvar, stop=()):
    """Time-series blocks from lines[i:]: each opens with ``<key> <index> ...`` and two
    header lines, then one record per line, ``YYYY-MM-DD HH:MM v1 ... vn``, up to the next
    block, a line starting with a ``stop`` keyword, or the end.  Returns (blocks, next i)
    with blocks a list of (header fields, ftime[nrec], data[nrec, nvar])."""
    blocks = []
    while i < len(lines) and not lines[i].split()[0] in stop:
        head = lines[i].split()
        if head[0] != key:
            raise ValueError(f"expected {key} at '{lines[i]}'")
        if int(head[1]) != len(blocks) + 1:
            raise ValueError(f"{key} blocks out of order")
        i += 3
        ft, rows = [], []
        while i < len(lines) and lines[i].split()[0] not in (key,) + tuple(stop):
            tok = lines[i].split()
            ft.append(str_time(tok[0] + " " + tok[1]))
            rows.append([float(v) for v in tok[2:2 + nvar]])
            i += 1
        blocks.append((head, np.array(ft), np.array(rows).reshape(len(ft), nvar)))
    return blocks, i


def _data_lines(path):
    with open(path) as f:
        return [ln.rstrip("\n") for ln in f if ln.strip() and not ln.lstrip().startswith("#")]


def read_ts_blocks(path, key, nvar):
    """Every ``key`` block of a forcing file (ReadForc, src/read_forc.c:3-80; ReadLai,
    src/read_lai.c; ReadBc, src/read_bc.c)."""
    return _ts_blocks(_data_lines(path), 0, key, nvar)[0]


def _equal_length(blocks, what):
    if len({len(b[1]) for b in blocks}) > 1:
        raise ValueError(f"{what} with different record counts are not supported")


def read_bc(path):
    """Boundary-condition series of a ``.bc`` file (BC_TS blocks, src/read_bc.c):
    (ftime[nts, nrec], value[nts, nrec])."""
    blocks = read_ts_blocks(path, "BC_TS", 1)
    _equal_length(blocks, "boundary series")
    return np.stack([b[1] for b in blocks]), np.stack([b[2][:, 0] for b in blocks])


def read_meteo(path):
    """Station records of a ``.meteo`` file: (ftime[nstation, nrec], data[nstation, nrec, 5],
    zlvl_wind[nstation]) with PRCP, SFCTMP, RH, SFCSPD, SOLAR, the variables the simple
    land-surface step reads (src/forcing.c:340-349).  Every station must hold the same
    number of records (the device layout is one [nstation, nrec] table)."""
    blocks = read_ts_blocks(path, "METEO_TS", 5)
    _equal_length(blocks, "stations")
    ftime = np.stack([b[1] for b in blocks])
    data = np.stack([b[2] for b in blocks])
    zlvl = np.array([float(b[0][3]) for b in blocks])
    return ftime, data, zlvl


def read_lai(path):
    """LAI forcing of a ``.lai`` file: (ftime[nts, nrec], lai[nts, nrec])."""
    blocks = read_ts_blocks(path, "LAI_TS", 1)
    _equal_length(blocks, "LAI series")
    return np.stack([b[1] for b in blocks]), np.stack([b[2][:, 0] for b in blocks])


def read_project(directory, name, lc_table):
    p = lambda ext: os.path.join(directory, f"{name}.{ext}")
    tri, nabr, nodes = read_mesh(p("mesh"))
    ne = tri.shape[0]
    att = read_att(p("att"), ne)
    soil = read_soil(p("soil"))
    cal = read_calib(p("calib"))
    lc = read_lc(lc_table)
    riv = read_riv(p("riv"))
    para = read_para(p("para"))
    si = att["soil"] - 1
    li = att["lc"] - 1
    # InitSoil (src/init_soil.c:24-66)
    smcmin = cal["POROSITY"] * soil["minsmc"][si]
    smcmax = cal["POROSITY"] * soil["maxsmc"][si]
    s = dict(ksath=cal["KSATH"] * soil["ksath"][si], ksatv=cal["KSATV"] * soil["ksatv"][si],
             kinfv=cal["KINF"] * soil["kinf"][si], smcmin=smcmin, smcmax=smcmax,
             alpha=cal["ALPHA"] * soil["alpha"][si], beta=cal["BETA"] * soil["beta"][si],
             dinf=np.full(ne, cal["DINF"] * soil["dinf"]), dmac=cal["DMAC"] * soil["dmac"][si],
             areafh=cal["MACHF"] * soil["machf"][si], areafv=cal["MACVF"] * soil["macvf"][si],
             kmacv=cal["KMACSATV"] * soil["kmacv_ro"] * soil["kinf"][si],
             kmach=cal["KMACSATH"] * soil["kmach_ro"] * soil["ksath"][si],
             # InitLc (src/init_lc.c:22-37)
             rough=cal["ROUGH"] * lc["rough"][li], rzd=cal["DROOT"] * lc["rzd"][li])
    # InitRiver (src/init_river.c:62-89)
    sh = riv["shape"] - 1
    mt = riv["matl"] - 1
    rivers = dict(fromnode=riv["fromnode"].astype(np.int32), tonode=riv["tonode"].astype(np.int32),
                  down=riv["down"].astype(np.int32), leftele=riv["left"].astype(np.int32),
                  rightele=riv["right"].astype(np.int32), intrpl_ord=riv["shp"][sh, 1].astype(np.int32),
                  depth=cal["RIV_DPTH"] * riv["shp"][sh, 0], coeff=cal["RIV_WDTH"] * riv["shp"][sh, 2],
                  bc_type=riv["bc"].astype(np.int32), rough=cal["ROUGH_RIV"] * riv["mat"][mt, 0],
                  cwr=riv["mat"][mt, 1], ksath=cal["KRIVH"] * riv["mat"][mt, 2],
                  ksatv=cal["KRIVV"] * riv["mat"][mt, 3], bedthick=cal["BEDTHCK"] * riv["mat"][mt, 4])
    m = Mesh(nodes[:, 0:2], nodes[:, 2], nodes[:, 3], tri, nabr, rivers, s,
             surf_mode=para["surf_mode"], riv_mode=para["riv_mode"], stepsize=para["stepsize"],
             bc_type=att["bc"])
    m.para = para
    return m.init()
