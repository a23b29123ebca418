"""Writes the committed golden fixtures from the reference checkout.

Run once in a container that has the reference at /root/reference:
    python tests/golden/make_fixtures.py [/root/reference]

Outputs (all data, no code):
  tests/golden/data/*.ini|.acc|.pot   -- the reference's data sets (CC0,
      data/LICENSE); Box44.ini is stored gzip-compressed.
  tests/golden/published.json         -- values the reference's own runs
      published in results/: total energies of TableII, Figure4 and Figure6
      (Box44) with their pair counts and pair-list capacities, and the
      4-digit force errors of Figure1 and Figure2; and
      the PME comparator's runs of Figure4 ("Figure4_pme": energy, beta,
      pair count, list capacity; the reference's pme CLI ignores --nu, so every run has
      the order "nu" = 6 it reports, whatever "nu_arg" asked for).
"""
import gzip
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SMALL = ["CsClN2", "CsClN16", "CsClN128", "NaClN8", "NaClN64", "NaClN512", "changaN8",
         "changaN64", "changaN512", "spceN300", "spceN600", "spceN900", "spceN2250"]


def main():
    out = os.path.join(HERE, "data")
    os.makedirs(out, exist_ok=True)
    for name in SMALL:
        for ext in (".ini", ".acc", ".pot"):
            src = os.path.join(REF, "data", name + ext)
            if os.path.exists(src):
                shutil.copyfile(src, os.path.join(out, name + ext))
    with open(os.path.join(REF, "data", "Box44.ini"), "rb") as f, \
            gzip.GzipFile(os.path.join(out, "Box44.ini.gz"), "wb", mtime=0) as g:
        g.write(f.read())
    shutil.copyfile(os.path.join(REF, "data", "Box44.pot"), os.path.join(out, "Box44.pot"))
    shutil.copyfile(os.path.join(REF, "data", "LICENSE"), os.path.join(out, "LICENSE"))

    pub = {}
    keep = ["NumberOfLevels", "nu", "cutoff", "TopLevelMx", "utotal", "nlist_interaction_count",
            "nlist_len", "poterror", "time_energy", "time_nlist"]
    for fig in ("TableII", "Figure4", "Figure6"):
        with open(os.path.join(REF, "results", fig + ".json")) as f:
            d = json.load(f)
        pub[fig] = [{"L": t.get("L", t["msmout"]["NumberOfLevels"]), "nu": t["nu"], "M": t["M"],
                     "a0": t["a0"], **{k: t["msmout"][k] for k in keep if k in t["msmout"]}}
                    for t in d["trial"]]
        if fig == "Figure4":
            pk = ["nu", "cutoff", "TopLevelMx", "beta", "tol_dir", "utotal",
                  "nlist_interaction_count", "nlist_len", "poterror", "deltaF/Fref",
                  "time_energy", "time_nlist"]
            pub["Figure4_pme"] = [{"M": t["M"], "a0": t["a0"], "nu_arg": t["nu"],
                                   **{k: t["pmeout"][k] for k in pk}} for t in d["trial"]]
    rows = []
    with open(os.path.join(REF, "results", "Figure1.txt")) as f:
        f.readline()
        for line in f:
            v = line.split()
            if len(v) == 5:
                rows.append({"N": int(v[0]), "M": int(v[1]), "a0": float(v[2]), "nu": int(v[3]),
                             "deltaF/Fref": float(v[4])})
    pub["Figure1"] = rows
    rows = []
    with open(os.path.join(REF, "results", "Figure2.txt")) as f:
        for line in f:
            v = line.split()
            if len(v) == 10 and v[0].startswith("spce"):
                rows.append({"data": v[0], "N": int(v[1]), "A": int(v[2]), "M": int(v[4]),
                             "a0": float(v[5]), "nu": int(v[6]), "deltaF/Fref": float(v[7]),
                             "deltaFest/Fref": float(v[8])})
            if line.startswith("Fudge"):
                break
    pub["Figure2"] = rows
    with open(os.path.join(HERE, "published.json"), "w") as f:
        json.dump(pub, f, indent=1)


if __name__ == "__main__":
    main()
