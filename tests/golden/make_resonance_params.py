"""Write tests/golden/resonance_params.json: the resonance inputs of the
PP[RES+CON] workload, read from the reference's data files (JSON with comments):
the "RES" list of tests/processes/ALICECUTS.json, modeldata/TUNE0/RES_<name>.json,
the pi+ pi- entries of modeldata/TUNE0/BRANCHING.json and PARAM_SPIN / PARAM_REGGE
of modeldata/TUNE0/GENERAL.json. Run once where the reference tree is present:
    python tests/golden/make_resonance_params.py /root/reference
"""
import json
import os
import re
import sys


def load(path):
    t = open(path).read()
    t = re.sub(r"//.*", "", t)
    t = re.sub(r",(\s*[}\]])", r"\1", t)
    return json.loads(t)


def main(ref):
    card = load(os.path.join(ref, "tests/processes/ALICECUTS.json"))
    names = card["PROCESSPARAM"]["RES"]
    br = load(os.path.join(ref, "modeldata/TUNE0/BRANCHING.json"))
    gen = load(os.path.join(ref, "modeldata/TUNE0/GENERAL.json"))
    out = {"process": card["PROCESSPARAM"]["PROCESS"],
           "spin": gen["PARAM_SPIN"], "omega": gen["PARAM_REGGE"]["omega"], "resonances": {}}
    for name in names:
        r = load(os.path.join(ref, f"modeldata/TUNE0/RES_{name}.json"))["PARAM_RES"]
        e = [v for v in br[str(r["PDG"])].values()
             if isinstance(v, dict) and sorted(v.get("PDG", [])) == [-211, 211]][0]
        out["resonances"][name] = {
            "pdg": r["PDG"], "M": r["M"], "W": r["W"], "J": r["J"], "P": r["P"], "BW": r["BW"],
            "g_A": r["g_A"], "g_phi": r["g_phi"], "g_FF": r["g_FF"],
            "rho_real": r["rho_real"], "rho_imag": r["rho_imag"],
            "BR": e["BR"], "alpha_ls": e["alpha_ls"]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "resonance_params.json")
    json.dump(out, open(dst, "w"), indent=1, sort_keys=True)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
