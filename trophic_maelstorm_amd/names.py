"""The reference's output folders and file names, built the way its drivers
build them with std::stringstream (a double at precision p prints as
printf("%.<p>g"); the stream default is p = 6).

The reference puts a random number in some names (the thread's jID); here
jID is an argument, so a run can name its files deterministically.

Citations are to Rietkerk_FPE/rietkerk_bjork_FKE.cpp and
Rietkerk_FPE/rietkerk_bjork_basic.h in the reference's simulations/ directory.
"""
from __future__ import annotations

import math

# rietkerk_bjork_basic.h:154-159
FRAME_PREFIX = "/FRAME_P_c_DP_G_"
PRELIM_PREFIX = "/PRELIM_AGGRAND_P_c_ID_"
REPLICATE_PREFIX = "/PRELIM_TSERIES_P_c_DP_G_"


def ostream(x, prec: int = 6) -> str:
    """x as `stream << setprecision(prec) << x` prints it (ints as ints)."""
    if isinstance(x, int) and not isinstance(x, bool):
        return str(x)
    return "%.*g" % (prec, float(x))


def _c_round(x: float) -> float:
    """C's round(): halves away from zero."""
    r = math.floor(abs(x))
    if abs(x) - r >= 0.5:
        r += 1
    return math.copysign(r, x)


def folders(prefix: str, n_basal: int = 1, mG: float = 0.0, mP: float = 0.0) -> dict:
    """frame_folder, prelim_folder and stat_prefix after set_Prefix (FKE:143-165)."""
    mG = int(mG) if mG == int(mG) else _c_round(mG * 100) / 100
    mP = int(mP) if mP == int(mP) else _c_round(mP * 100) / 100
    sp = f"{n_basal}Sp"
    if mG > 0 and mP > 0:
        sp += f"_mG_{ostream(mG)}_mP_{ostream(mP)}"
    elif mG > 0 and mP == 0:
        sp += f"_mG_{ostream(mG)}"
    return {
        "frame_folder": f"../Data/Rietkerk/Frames/Stochastic/{sp}/{prefix}_",
        "prelim_folder": f"../Data/Rietkerk/Prelims/Stochastic/{sp}/{prefix}_",
        "stat_prefix": f"../Data/Rietkerk/Stochastic/1stOrderCC_Rietkerk_{prefix}_P_c_G_",
    }


def _run_dir(base: str, a_st, a_end, dP, Gstar, Vstar, geq_tail: str = "", veq_tail: str = "",
             plain_tail: str = "") -> str:
    d = f"{base}{ostream(a_st)}-{ostream(a_end)}_dP_{ostream(dP)}"
    if Gstar != -1:
        return d + "_Geq_" + ostream(Gstar, 5) + geq_tail
    if Vstar != -1:
        return d + "_Veq_" + ostream(Vstar, 5) + veq_tail
    return d + plain_tail


def frame_path(frame_folder, a_st, a_end, dP, g, t, dt, a, D1, jID, dx, r, Gstar=-1, Vstar=-1) -> str:
    """A FRAME file (save_framefileswrapper, FKE:2890-2912; save_frame, FKE:2831-2834)."""
    d = _run_dir(frame_folder, a_st, a_end, dP, Gstar, Vstar)
    name = (f"{FRAME_PREFIX}{g}_T_{ostream(t)}_dt_{ostream(dt, 3)}_a_{ostream(a, 4)}_D1_{ostream(D1, 3)}"
            f"_jID_{ostream(jID)}_dx_{ostream(dx, 2)}_R_{r}.csv")
    return d + name


def timeseries_path(prelim_folder, a_st, a_end, dP, g, t, dt, a, jID, r, Gstar=-1, Vstar=-1) -> str:
    """A replicate's PRELIM time series (save_prelimfileswrapper, FKE:3051-3075)."""
    d = _run_dir(prelim_folder, a_st, a_end, dP, Gstar, Vstar, "/TimeSeries", "/TimeSeries", "/TimeSeries")
    name = f"{REPLICATE_PREFIX}{g}_T_{ostream(t)}_dt_{ostream(dt, 3)}_a_{ostream(a, 5)}_jID_{ostream(jID)}_R_{r}.csv"
    return d + name


def aggregate_prelim_path(prelim_folder, a_st, a_end, dP, jID, g, t_max, dt, a, D2, r, Gstar=-1, Vstar=-1) -> str:
    """The running replicate statistics after replicate r (FKE:4606-4630); the
    reference adds /TimeSeries only in the Veq case."""
    d = _run_dir(prelim_folder, a_st, a_end, dP, Gstar, Vstar, "", "/TimeSeries", "")
    name = (f"{PRELIM_PREFIX}{ostream(jID)}_DP_G_{g}_T_{ostream(t_max)}_dt_{ostream(dt, 3)}_a_{ostream(a, 5)}"
            f"_D2_{ostream(D2, 4)}_R_{r}.csv")
    return d + name


def sweep_path(stat_prefix, g, t_max, dt, D1, a_start, a_end, dx, r, Gstar=-1, Vstar=-1) -> str:
    """The sweep's final statistics file (FKE:4843-4863)."""
    s = (f"{stat_prefix}{g}_T_{ostream(t_max)}_dt_{ostream(dt, 3)}_D1_{ostream(D1, 3)}_a1_{ostream(a_start)}"
         f"_a2_{ostream(a_end)}_dx_{ostream(dx, 2)}")
    if Gstar != -1:
        s += "_Geq_" + ostream(Gstar, 5)
    elif Vstar != -1:
        s += "_Veq_" + ostream(Vstar, 5)
    return s + f"_R_{r}.csv"
