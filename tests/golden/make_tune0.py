"""Write tests/golden/tune0_params.json: the model and numerics values the
preset workloads use, read from the reference's data files (TUNE0 model card,
PDG mass table, eikonal numerics, ALICE cut card). Data only."""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def json5(path):
    txt = open(path).read()
    txt = re.sub(r"//[^\n]*", "", txt)           # line comments
    txt = re.sub(r",(\s*[}\]])", r"\1", txt)     # trailing commas
    return json.loads(txt)


def pdg_line(pdg):
    for line in open(os.path.join(REF, "modeldata", "mass_width_2018.mcd")):
        if line.startswith("*"):
            continue
        f = line.split()
        if f and f[0] == str(pdg):
            # after the integer id columns: mass, +err, -err, width, +err, -err, name
            vals = [x for x in f[1:] if ("." in x or "E" in x.upper()) and
                    re.match(r"^[-+]?\d*\.?\d*(E[-+]?\d+)?$", x, re.I)]
            return float(vals[0]), (float(vals[3]) if len(vals) > 3 else 0.0)
    raise KeyError(pdg)


def main():
    g = json5(os.path.join(REF, "modeldata", "TUNE0", "GENERAL.json"))
    num = json5(os.path.join(REF, "modeldata", "NUMERICS.json"))
    alice = json5(os.path.join(REF, "tests", "processes", "ALICECUTS.json"))
    mpi, wpi = pdg_line(211)
    mp, _ = pdg_line(2212)
    mrho, wrho = pdg_line(113)
    out = {
        "PARAM_SOFT": g["PARAM_SOFT"], "PARAM_REGGE": g["PARAM_REGGE"], "PARAM_PI": g["PARAM_PI"],
        "PARAM_RHO": g["PARAM_RHO"], "PARAM_TENSORPOM": g["PARAM_TENSORPOM"],
        "NUMERICS_EIKONAL": num["NUMERICS_EIKONAL"],
        "pdg": {"pi+": [mpi, wpi], "p": [mp, 0.0], "rho0": [mrho, wrho]},
        "alice": {"ENERGY": alice["PROCESSPARAM"]["ENERGY"], "GENCUTS_C_Rap": alice["GENCUTS"]["<C>"]["Rap"],
                  "FIDCUTS": alice["FIDCUTS"]},
    }
    path = os.path.join(HERE, "tune0_params.json")
    json.dump(out, open(path, "w"), indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
