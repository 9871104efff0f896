"""Extracts the inputs and recorded outputs of Chroma's RNG-free regression test prec_wilson-twisted
(tests/chroma/hadron/propagator/prec_wilson-twisted.{ini,out,metric}.xml) into prec_wilson_twisted.json.
Run in a container where the reference tree is mounted at /root/reference."""
import json
import os
import xml.etree.ElementTree as ET

REF = "/root/reference/tests/chroma/hadron/propagator"
ini = ET.parse(os.path.join(REF, "prec_wilson-twisted.ini.xml")).getroot()
out = ET.parse(os.path.join(REF, "prec_wilson-twisted.out.xml")).getroot()
met = ET.parse(os.path.join(REF, "prec_wilson-twisted.metric.xml")).getroot()

prop = ini.find(".//elem[Name='PROPAGATOR']/Param")
fa = prop.find("FermionAction")
bc = fa.find("FermionBC")
inv = prop.find("InvertParam")
g = {
    "source": "tests/chroma/hadron/propagator/prec_wilson-twisted.{ini,out,metric}.xml",
    "nrow": [int(v) for v in ini.find("Param/nrow").text.split()],
    "cfg_type": ini.find("Cfg/cfg_type").text.strip(),
    "FermAct": fa.find("FermAct").text.strip(),
    "kappa": float(fa.find("Kappa").text),
    "boundary": [int(v) for v in bc.find("boundary").text.split()],
    "phases_by_pi": [float(v) for v in bc.find("phases_by_pi").text.split()],
    "phases_dir": [int(v) for v in bc.find("phases_dir").text.split()],
    "invType": inv.find("invType").text.strip(),
    "RsdCG": float(inv.find("RsdCG").text),
    "MaxCG": int(inv.find("MaxCG").text),
    "t_srce": [int(v) for v in ini.find(".//elem[Name='MAKE_SOURCE']/Param/Source/t_srce").text.split()],
    "n_count": int(out.find(".//QuarkProp4/Qprop/n_count").text),
    "n_counts": [int(e.text) for e in out.findall(".//QuarkProp4/Qprop/n_count")],
    "ncg_had": int(out.find(".//Relaxation_Iterations/ncg_had").text),
    "prop_corr": [float(v) for v in out.find(".//Prop_correlator/prop_corr").text.split()],
    "tolerance": float(met.find("assertion").get("tolerance")),
}
json.dump(g, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "prec_wilson_twisted.json"), "w"), indent=1)
print(g)
