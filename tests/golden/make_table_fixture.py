"""Writes reference_out_tables.json from the reference's own result files
(tests/spectra/*.out): for every file its three printed blocks as they are
(configuration, match table, protein details) and the cells each line is
made of (a cell starts where its column's dashes start; trailing spaces are
layout).  Data only; run from the repository root with the reference
checked out at /root/reference.
"""
import glob
import json
import os

SRC = "/root/reference/tests/spectra"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_out_tables.json")


def blocks(lines):
    out, cur = [], []
    for l in lines:
        if l.strip() == "":
            if cur:
                out.append(cur)
            cur = []
        else:
            cur.append(l)
    if cur:
        out.append(cur)
    return out


def main():
    files = []
    for path in sorted(glob.glob(os.path.join(SRC, "*.out"))):
        lines = open(path).read().split("\n")
        cfg, tab, det = blocks(lines)[:3]
        dash = tab[1]
        starts = [i for i, ch in enumerate(dash) if ch == "-" and (i == 0 or dash[i - 1] == " ")]
        cells = []
        for l in tab:
            bounds = starts + [len(l)]
            cells.append("\t".join(l[bounds[c]:bounds[c + 1]].rstrip() for c in range(len(starts))))
        k = cfg[0].index("::")
        ccells = ["\t".join([l[:k].rstrip(), "::", l[k + 3:].rstrip()]) for l in cfg]
        files.append(dict(file=os.path.basename(path), config_lines=cfg, config_cells=ccells,
                          table_lines=tab, table_cells=cells, detail_lines=det,
                          detail_headers=[l[5:].rstrip() for l in det]))
    with open(OUT, "w") as f:
        json.dump(files, f, indent=1)
    print(OUT, len(files), "files")


if __name__ == "__main__":
    main()
