"""The reference's parameter files (.par.json / .par.toml) read into the
right-hand side's parameters, so a run configured for the reference can be
replayed here.

readParamFile (BSSN_GR/src/grUtils.cpp:37-52) picks TOML for a ".toml" suffix
(readParamTOMLFile, BSSN_GR/src/parameters.cpp:183-560) and JSON otherwise
(readParamJSONFile, grUtils.cpp:54-468). Only the keys that reach the
right-hand side, the constraint diagnostic and the time step are read; the
rest of the file (IO, AMR, initial data, solver control) belongs to parts of
the reference that are not on this path and is returned untouched in `raw`.

Keys read, with the reference's behaviour:
  BSSN_LAMBDA       four integers (JSON: BSSN_LAMBDA_1..4, cast to unsigned;
                    TOML: an array)                                 required
  BSSN_LAMBDA_F     two floats (JSON: BSSN_LAMBDA_F0/F1; TOML: array) required
  CHI_FLOOR, KO_DISS_SIGMA                                          required
  RIT_ETA_CENTRAL, RIT_ETA_OUTER, RIT_ETA_WIDTH                     optional
  BSSN_KO_SIGMA_SCALE_BY_CONFORMAL (enables CAKO) and
  BSSN_KO_SIGMA_SCALE_BY_CONFORMAL_POST_MERGER_ONLY (if true, CAKO stays off
  before the merger: grUtils.cpp:186-191)                           optional
  BSSN_EPSILON_CAKO_GAUGE, BSSN_EPSILON_CAKO_OTHER                  optional
  BSSN_BH1_CONSTRAINT_R, BSSN_BH2_CONSTRAINT_R (default 5.0)        optional
  BSSN_BH1, BSSN_BH2 tables (X, Y, Z: the norm's excision centres)  optional here
  BSSN_CFL_FACTOR, BSSN_RK_TIME_BEGIN, BSSN_RK_TIME_END             optional here
  BSSN_RK_TYPE (0 RK3, 1 RK4, 2 RK45; grUtils.cpp:100)              optional here
  BSSN_ETA_R0, BSSN_ETA_POWER (JSON: BSSN_ETA_POWER_1/_2; TOML: array;
  the eta function's constants, grUtils.cpp:205-207)                 optional
  BSSN_RK45_TIME_STEP_SIZE, BSSN_RK45_DESIRED_TOL (parameters.cpp:
  220-223: the first step size and the tolerance of dgr_rk45_step)   optional here
  BSSN_XI (JSON: BSSN_XI_0.._2, TOML: an array; cast to unsigned like
  grUtils.cpp:161-163, parameters.cpp:315-317: the Rochester gauge's
  coefficients)                                                     optional
A required key that is missing raises KeyError, where the reference's JSON
reader throws on the null it would convert.
"""
from __future__ import annotations

import dataclasses
import json
import os

from .api import BssnParams, default_params


@dataclasses.dataclass
class ParamFile:
    params: BssnParams                      # for dgr_bssn_rhs / dgr_rk4_step(s)
    bh: tuple | None                        # ((x1, y1, z1), (x2, y2, z2)) or None
    constraint_r: tuple                     # (BSSN_BH1_CONSTRAINT_R, BSSN_BH2_CONSTRAINT_R)
    cfl: float | None                       # BSSN_CFL_FACTOR (dt = cfl * h_min in the reference)
    time: tuple | None                      # (BSSN_RK_TIME_BEGIN, BSSN_RK_TIME_END)
    raw: dict                               # the whole file as parsed
    rk_type: int | None = None              # BSSN_RK_TYPE: 0 RK3 (dgr_rk3_step), 1 RK4, 2 RK45 (dgr_rk45_step)
    rk45: tuple | None = None               # (BSSN_RK45_TIME_STEP_SIZE, BSSN_RK45_DESIRED_TOL)


def _load(path: str) -> tuple[dict, bool]:
    if path.endswith(".toml"):
        try:
            import tomllib as toml  # Python >= 3.11
        except ImportError:
            import tomli as toml
        with open(path, "rb") as fh:
            return toml.load(fh), True
    with open(path) as fh:
        return json.load(fh), False


def _required(d: dict, key: str):
    if key not in d:
        raise KeyError(f"parameter file lacks the required key {key}")
    return d[key]


def read_param_file(path: str | os.PathLike) -> ParamFile:
    path = os.fspath(path)
    d, is_toml = _load(path)
    p = default_params()
    lam = _required(d, "BSSN_LAMBDA")
    lamf = _required(d, "BSSN_LAMBDA_F")
    if is_toml:
        lam4 = [int(lam[i]) for i in range(4)]
        lamf2 = [float(lamf[0]), float(lamf[1])]
    else:
        lam4 = [int(_required(lam, f"BSSN_LAMBDA_{i + 1}")) for i in range(4)]
        lamf2 = [float(_required(lamf, f"BSSN_LAMBDA_F{i}")) for i in range(2)]
    for i in range(4):
        p.lambda_[i] = float(lam4[i])
    p.lambda_f[0], p.lambda_f[1] = lamf2
    p.chi_floor = float(_required(d, "CHI_FLOOR"))
    p.ko_sigma = float(_required(d, "KO_DISS_SIGMA"))
    for key, field in (("RIT_ETA_CENTRAL", "eta_central"), ("RIT_ETA_OUTER", "eta_outer"),
                       ("RIT_ETA_WIDTH", "eta_width"), ("BSSN_EPSILON_CAKO_GAUGE", "eps_cako_gauge"),
                       ("BSSN_EPSILON_CAKO_OTHER", "eps_cako_other")):
        if key in d:
            setattr(p, field, float(d[key]))
    scale = bool(d.get("BSSN_KO_SIGMA_SCALE_BY_CONFORMAL", False))
    if bool(d.get("BSSN_KO_SIGMA_SCALE_BY_CONFORMAL_POST_MERGER_ONLY", False)):
        scale = False
    p.cako_enabled = 1 if scale else 0
    # the eta function's constants (grUtils.cpp:205-207, parameters.cpp:329-331);
    # whether it is used is a build option there (BSSN_ETA_FUNCTION), here
    # params.eta_func, left to the caller
    if "BSSN_ETA_R0" in d:
        p.eta_r0 = float(d["BSSN_ETA_R0"])
    if "BSSN_ETA_POWER" in d:
        pw = d["BSSN_ETA_POWER"]
        if isinstance(pw, dict):
            pw = (pw["BSSN_ETA_POWER_1"], pw["BSSN_ETA_POWER_2"])
        p.eta_power[0], p.eta_power[1] = float(pw[0]), float(pw[1])
    # the Rochester gauge's coefficients; whether that gauge (USE_ROCHESTER_GAUGE)
    # or the slow-start lapse (BSSN_ENABLE_SSL_IN_RHS) is used is a build option
    # there, here params.rochester / params.ssl, left to the caller
    if "BSSN_XI" in d:
        xi = d["BSSN_XI"]
        if isinstance(xi, dict):
            xi = [xi[f"BSSN_XI_{i}"] for i in range(3)]
        for i in range(3):
            p.xi[i] = float(int(xi[i]) & 0xFFFFFFFF)

    bh = None
    if "BSSN_BH1" in d and "BSSN_BH2" in d:
        bh = tuple(tuple(float(d[b][c]) for c in ("X", "Y", "Z")) for b in ("BSSN_BH1", "BSSN_BH2"))
    r = (float(d.get("BSSN_BH1_CONSTRAINT_R", 5.0)), float(d.get("BSSN_BH2_CONSTRAINT_R", 5.0)))
    cfl = float(d["BSSN_CFL_FACTOR"]) if "BSSN_CFL_FACTOR" in d else None
    time = None
    if "BSSN_RK_TIME_BEGIN" in d and "BSSN_RK_TIME_END" in d:
        time = (float(d["BSSN_RK_TIME_BEGIN"]), float(d["BSSN_RK_TIME_END"]))
    rk = int(d["BSSN_RK_TYPE"]) if "BSSN_RK_TYPE" in d else None
    rk45 = None
    if "BSSN_RK45_TIME_STEP_SIZE" in d and "BSSN_RK45_DESIRED_TOL" in d:
        rk45 = (float(d["BSSN_RK45_TIME_STEP_SIZE"]), float(d["BSSN_RK45_DESIRED_TOL"]))
    return ParamFile(params=p, bh=bh, constraint_r=r, cfl=cfl, time=time, raw=d, rk_type=rk, rk45=rk45)
