"""Writes reference_out_peptides.json from the SEQUEST-style result files of
the reference's own tests (tests/spectra/*.out): for every listed match the
printed (M+H)+ of the peptide and its sequence, and for every file the
printed precursor (M+H)+ and the charge in the file name.  Data only; run
from the repository root with the reference checked out at /root/reference.
"""
import glob
import json
import os
import re

SRC = "/root/reference/tests/spectra"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_out_peptides.json")

ROW = re.compile(r"\s*(\d+)\.\s+([0-9.]+)\s+([0-9.]+)\s+([0-9.]+)\s+.*\s(\S)\.([A-Z]+)\.(\S)\s*$")
HDR = re.compile(r"\(M\+H\)\+ Mass\s*::\s*([0-9.]+)\s*~\s*([0-9.]+)")


def main():
    files = []
    for path in sorted(glob.glob(os.path.join(SRC, "*.out"))):
        name = os.path.basename(path)
        m = re.search(r"\.(\d)\.out$", name)
        entry = dict(file=name, charge=int(m.group(1)) if m else None, precursor_mh=None,
                     tolerance=None, matches=[])
        for line in open(path):
            h = HDR.search(line)
            if h:
                entry["precursor_mh"], entry["tolerance"] = h.group(1), h.group(2)
            r = ROW.match(line)
            if r:
                entry["matches"].append(dict(rank=int(r.group(1)), mh=r.group(2),
                                             deltcn=r.group(3), xcorr=r.group(4),
                                             before=r.group(5), peptide=r.group(6),
                                             after=r.group(7)))
        files.append(entry)
    with open(OUT, "w") as f:
        json.dump(files, f, indent=1)
    print(OUT, sum(len(e["matches"]) for e in files), "matches in", len(files), "files")


if __name__ == "__main__":
    main()
